// Parameter blocks and launchers of the PLS-LDA cross-validation kernels
// (pls_kernels.hip): the per-fold cache (standardisation and Gram matrix) and
// the per-evaluation fused fold kernel (class cross-products, power-refresh
// SIMPLS, LDA fit, validation argmax). All arithmetic is FP64.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace kodama_hip {

constexpr int kPlsThreads = 256;      // one workgroup per fold
constexpr int kPlsMaxComponents = 64;
constexpr int kPlsMaxDims = 256;
constexpr int kPlsMaxSLds = 8192;     // doubles of S = p x classes kept in LDS

// Rows of every fold, concatenated: fold f owns [offsets[f], offsets[f+1]).
struct PlsFoldRows {
  const int* rows = nullptr;
  const int* offsets = nullptr;  // [folds + 1]
};

struct PlsCacheParams {
  const float* x = nullptr;      // [n x p] row-major
  int p = 0;
  int folds = 0;
  int center = 1;
  int scale = 1;
  PlsFoldRows train;             // any order; sums run in this order
  double* mean = nullptr;        // [folds x p]
  double* sd = nullptr;          // [folds x p]
  double* xsum = nullptr;        // [folds x p] column sums of the standardised training rows
  double* gram = nullptr;        // [folds x p x p]
};

struct PlsEvalParams {
  const float* x = nullptr;
  int p = 0;
  int folds = 0;
  int requested = 0;             // components evaluated (fixed or max)
  int rank_cap = 0;              // min(max_components, p) before the per-fold n - 1 cap
  const double* w0 = nullptr;    // [p] first power-iteration start (normal draws)
  const double* mean = nullptr;  // fold cache
  const double* sd = nullptr;
  const double* xsum = nullptr;
  const double* gram = nullptr;
  // Training rows grouped by class code (ascending code, rows ascending inside
  // a class), as pls_group_kernel writes them: fold f has seg_count[f]
  // classes, class s is seg_code[f * max_classes + s] and owns train.rows[i]
  // for seg_start[f * (max_classes + 1) + s] <= i < seg_start[... + s + 1]
  // (indices into the concatenated training rows).
  PlsFoldRows train;
  const int* seg_code = nullptr;     // [folds x max_classes]
  const int* seg_start = nullptr;    // [folds x (max_classes + 1)]
  const int* seg_count = nullptr;    // [folds]
  PlsFoldRows val;
  int max_classes = 0;               // label codes (0 .. max_classes - 1)
  // Scratch (device, per fold).
  double* weights = nullptr;         // [folds x p x kPlsMaxComponents], [j * kA + a]
  double* loadings = nullptr;        // [folds x p x kPlsMaxComponents]
  double* s0 = nullptr;              // [folds x max_classes x p] S before deflation
  double* gw = nullptr;              // [folds x p x kPlsMaxComponents] G W
  double* linear = nullptr;          // [folds x max_classes x kPlsMaxComponents]
  double* constant = nullptr;        // [folds x max_classes]
  // Outputs. components[f] = -1 reports a failed LDA factorisation.
  int* pred_code = nullptr;          // [n] written at validation rows
  int* components = nullptr;         // [folds]
  int* rank = nullptr;               // [folds] fitted SIMPLS components (0 = degenerate fit)
  // Validation scores [val total x kPlsMaxComponents] and discriminants
  // [val total x max_classes] (rows in validation-list order).
  double* val_scores = nullptr;
  double* discriminants = nullptr;
};

// Groups each fold's training rows (ascending, `train`) by the class code of
// each row (codes[row]) into `sorted`, stably, and writes the segment tables.
struct PlsGroupParams {
  PlsFoldRows train;
  const int* codes = nullptr;        // [rows] label code per row
  int folds = 0;
  int max_classes = 0;
  int* sorted = nullptr;             // [train total]
  int* seg_code = nullptr;           // [folds x max_classes]
  int* seg_start = nullptr;          // [folds x (max_classes + 1)]
  int* seg_count = nullptr;          // [folds]
};

size_t pls_eval_lds_bytes(int p, int max_classes);
hipError_t launch_pls_group(const PlsGroupParams& prm, hipStream_t s);
hipError_t launch_pls_cache(const PlsCacheParams& prm, hipStream_t s);
hipError_t launch_pls_eval(const PlsEvalParams& prm, hipStream_t s);

}  // namespace kodama_hip
