// Parameter blocks and launchers of the PLS-LDA cross-validation kernels
// (pls_kernels.hip): the per-fold cache (standardisation and Gram matrix) and
// the per-evaluation fused fold kernel (class cross-products, power-refresh
// SIMPLS, LDA fit, validation argmax). All arithmetic is FP64.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace kodama_hip {

constexpr int kPlsThreads = 256;      // general PLS kernels
constexpr int kPlsEvalThreads = 64;   // one wave per fold in the cross-validation kernels
constexpr int kPlsMaxComponents = 64;
constexpr int kPlsMaxDims = 256;
constexpr int kPlsMaxSLds = 8192;     // doubles of S = p x classes kept in LDS

// Rows of every fold, concatenated: fold f owns [offsets[f], offsets[f+1]).
struct PlsFoldRows {
  const int* rows = nullptr;
  const int* offsets = nullptr;  // [folds + 1]
};

// Fold cache (mean, scale, column sums of the standardised training rows,
// Gram matrix G = Xs'Xs of every fold's training rows) from moments: each
// fold's own rows, shifted by one row of their run (the run's first row, so
// that the moments hold no large offsets), give the moment matrix of
// [x - shift | 1] on the FP64 matrix cores, in row chunks; a fold's training
// moments are the sum of its run's other folds' moments, from which mean,
// scale (the scale accumulator starts at 1), column sums and G follow.
struct PlsCacheParams {
  const float* x = nullptr;      // [n x p] row-major
  int p = 0;
  int folds = 0;
  int center = 1;
  int scale = 1;
  PlsFoldRows val;               // each fold's own rows
  const int* shift_row = nullptr;    // [folds] the row that shifts the fold's run
  const int* run_first = nullptr;    // [folds] as PlsEvalParams
  const int* run_end = nullptr;
  int chunks = 0;                    // row chunks of all folds, fold by fold
  const int* chunk_fold = nullptr;   // [chunks]
  const int* chunk_begin = nullptr;  // [chunks] positions in the concatenated own rows
  const int* chunk_end = nullptr;
  const int* chunk_off = nullptr;    // [folds + 1] first chunk of each fold
  double* partial = nullptr;         // [chunks x z16 x z16], z16 = p + 1 rounded up to 16
  double* moments = nullptr;         // [folds x (p + 1) x (p + 1)]
  double* mean = nullptr;        // [folds x p]
  double* sd = nullptr;          // [folds x p]
  double* xsum = nullptr;        // [folds x p] column sums of the standardised training rows
  double* gram = nullptr;        // [folds x p x p]
};
constexpr int kPlsMomentRows = 256;  // rows per moment chunk

// Class sums of each fold's validation rows (ascending row order, FP64):
// fsum[f][c][j] = sum of x[row][j] over the fold's rows of code c, fcnt[f][c]
// their number. A fold's training sums are its run's other folds' sums.
struct PlsFoldSumsParams {
  const float* x = nullptr;
  int p = 0;
  int folds = 0;
  int max_classes = 0;
  const int* codes = nullptr;        // [rows] label code per row
  PlsFoldRows val;
  double* fsum = nullptr;            // [folds x max_classes x p]
  int* fcnt = nullptr;               // [folds x max_classes]
};

struct PlsEvalParams {
  const float* x = nullptr;
  int p = 0;
  int folds = 0;
  int requested = 0;             // components evaluated (fixed or max)
  int da = 0;                    // 1: PLS-DA head (least squares on the scores), 0: LDA
  int rank_cap = 0;              // min(max_components, p) before the per-fold n - 1 cap
  int max_classes = 0;           // label codes (0 .. max_classes - 1)
  const double* w0 = nullptr;    // [p] first power-iteration start (normal draws)
  const double* mean = nullptr;  // fold cache
  const double* sd = nullptr;
  const double* xsum = nullptr;
  const double* gram = nullptr;
  // The folds of fold f's run are [run_first[f], run_end[f]); the training
  // rows of f are the rows of the run's other folds.
  const int* run_first = nullptr;    // [folds]
  const int* run_end = nullptr;      // [folds]
  const double* fsum = nullptr;      // class sums of every fold's rows (PlsFoldSumsParams)
  const int* fcnt = nullptr;
  PlsFoldRows val;
  // Per fold: the label codes present among its training rows, ascending.
  int* seg_code = nullptr;           // [folds x max_classes]
  int* seg_count = nullptr;          // [folds]
  // Per-fold scratch in global memory; the large one-wave layout also keeps
  // the loadings, G W, class score sums and LDA rows here instead of LDS.
  double* weights = nullptr;         // [folds x kPlsMaxComponents x p], component-major [a * p + j]
  double* loadings = nullptr;        // [folds x kPlsMaxComponents x p]
  double* gw = nullptr;              // [folds x kPlsMaxComponents x p] G W
  double* zsum = nullptr;            // [folds x max_classes x p] standardised class sums (present classes)
  double* mu = nullptr;              // [folds x max_classes x kPlsMaxComponents]
  double* linear = nullptr;          // [folds x max_classes x kPlsMaxComponents] (DA: coefficients)
  double* constant = nullptr;        // [folds x max_classes] (DA: class frequencies)
  // Outputs. components[f] = -1 reports a failed LDA factorisation.
  int* pred_code = nullptr;          // [n] written at validation rows
  int* components = nullptr;         // [folds]
  int* rank = nullptr;               // [folds] fitted SIMPLS components (0 = degenerate fit)
  // Validation scores [val total x kPlsMaxComponents] and discriminants
  // [val total x max_classes] (rows in validation-list order).
  double* val_scores = nullptr;
  double* discriminants = nullptr;
};

// General PLS (fit_general_pls, plscv.cpp:2765-2878) on the device, FP64.
// Column statistics of a matrix m [n x cols] in row chunks: pass 0 sums the
// columns, pass 1 the squared deviations from `mean`; partial [chunks x cols],
// reduced in chunk order by launch_pls_colstats_reduce into mean (pass 0:
// sum / n when centring, else 0) or sd (pass 1: sqrt((1 + sum) / max(1, n-1)),
// non-finite or zero -> 1; the scale accumulator starts at 1 as in
// train_center_scale_float).
struct PlsColStatsParams {
  const float* m = nullptr;
  int n = 0;
  int cols = 0;
  int chunks = 0;
  int chunk_rows = 0;
  int pass = 0;
  int center = 1;
  double* mean = nullptr;      // [cols]
  double* sd = nullptr;        // [cols] (pass 1)
  double* sum = nullptr;       // [cols] (pass 0, may be null) the column sums
  double* partial = nullptr;   // [chunks x cols]
};

// Cross-products of the standardised rows: Z = Xs' [Xs | Yc] on the matrix
// cores (v_mfma_f64_16x16x4f64), one row chunk per workgroup, partial
// [chunks x p16 x z16] (p16, z16 = p and p + q rounded up to 16), then reduced
// in chunk order into gram [p x p] and cross [q x p] (row r = X' y_r).
struct PlsCrossParams {
  const float* x = nullptr;    // [n x p]
  const float* y = nullptr;    // [n x q]
  int n = 0, p = 0, q = 0;
  int chunks = 0, chunk_rows = 0;  // chunk_rows a multiple of 16
  const double* x_mean = nullptr;
  const double* x_sd = nullptr;
  const double* y_mean = nullptr;
  double* partial = nullptr;
  double* gram = nullptr;      // [p x p]
  double* cross = nullptr;     // [q x p]
};

// One wave: power-refresh SIMPLS of the cross-product (q rows in LDS) with
// the Gram matrix, then the response weights Q = S0' W and the coefficients
// W Q'.
struct PlsFitParams {
  int p = 0, q = 0;
  int rank_cap = 0;            // max(1, min(requested, p, n - 1))
  const double* w0 = nullptr;  // [p]
  const double* gram = nullptr;
  const double* cross = nullptr;   // [q x p], undeflated
  double* weights = nullptr;       // [p x kPlsMaxComponents]
  double* loadings = nullptr;      // [p x kPlsMaxComponents]
  double* yweights = nullptr;      // [q x kPlsMaxComponents]
  double* coefficients = nullptr;  // [p x q]
  int* fitted = nullptr;           // [1] rank (0: degenerate)
  double* wscratch = nullptr;      // [2 x kPlsMaxComponents x p] component-major weights and loadings
};

// Scores T = Xs W [n x A] and fitted values ybar + T Q' [n x q], one thread
// per row, written as float.
struct PlsScoreParams {
  const float* x = nullptr;
  int n = 0, p = 0, q = 0, A = 0;
  const double* x_mean = nullptr;
  const double* x_sd = nullptr;
  const double* y_mean = nullptr;
  const double* weights = nullptr;   // [p x kPlsMaxComponents]
  const double* yweights = nullptr;  // [q x kPlsMaxComponents]
  double* tbuf = nullptr;            // [n x A] scratch: the scores in FP64 (unstaged kernel only)
  float* scores = nullptr;           // [n x A]
  float* fitted = nullptr;           // [n x q]
};

hipError_t launch_pls_colstats(const PlsColStatsParams& prm, hipStream_t s);
hipError_t launch_pls_colstats_reduce(const PlsColStatsParams& prm, int scale, hipStream_t s);
hipError_t launch_pls_cross(const PlsCrossParams& prm, hipStream_t s);
size_t pls_fit_lds_bytes(int p, int q, int rc, bool staged);
hipError_t launch_pls_fit(const PlsFitParams& prm, hipStream_t s);
// True when the scores take the LDS-staged kernel (up to 16 components),
// which needs no tbuf.
bool pls_scores_staged(int p, int q, int A);
hipError_t launch_pls_scores(const PlsScoreParams& prm, hipStream_t s);

// LDS of one fold's wave: the compact layout keeps the loadings, G W, the
// class score sums and the LDA rows in LDS (and G when p <= 64), the wide
// layout keeps them in the global scratch. pls_eval_layout picks the compact
// one when it fits 64 KiB.
size_t pls_eval_lds_bytes(int p, int max_classes, int rank_cap, bool compact);
bool pls_eval_compact(int p, int max_classes, int rank_cap);
hipError_t launch_pls_fold_sums(const PlsFoldSumsParams& prm, hipStream_t s);
hipError_t launch_pls_cache(const PlsCacheParams& prm, hipStream_t s);
hipError_t launch_pls_eval(const PlsEvalParams& prm, hipStream_t s);

}  // namespace kodama_hip
