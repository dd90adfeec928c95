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
  // Training rows grouped by class code (ascending code, rows ascending within
  // a class): fold f's classes are seg_code[seg_offsets[f] .. seg_offsets[f+1])
  // and class s of that range owns train.rows[seg_start[s] .. seg_start[s+1])
  // (seg_start indexes the concatenated training rows; one extra entry per fold).
  PlsFoldRows train;
  const int* seg_code = nullptr;     // [total classes]
  const int* seg_offsets = nullptr;  // [folds + 1]
  const int* seg_start = nullptr;    // [total classes + folds]
  PlsFoldRows val;
  int max_classes = 0;               // largest class count of a fold
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

size_t pls_eval_lds_bytes(int p, int max_classes);
hipError_t launch_pls_cache(const PlsCacheParams& prm, hipStream_t s);
hipError_t launch_pls_eval(const PlsEvalParams& prm, hipStream_t s);

}  // namespace kodama_hip
