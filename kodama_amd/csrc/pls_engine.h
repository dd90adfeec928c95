// Host side of the device PLS-LDA cross-validation (pls_engine.cpp): the data
// stays on the device, the fold cache is rebuilt only when the folds change,
// and each evaluation uploads the training rows grouped by class.
#pragma once

#include <vector>

#include "engine.h"

namespace kodama_hip {

// Per-fold detail of one evaluation (validation rows in ascending order).
struct PlsFoldDetail {
  int rank = 0;                     // fitted SIMPLS components (0: degenerate)
  int components = 0;
  std::vector<double> weights;      // [p x rank]
  std::vector<double> val_scores;   // [val rows x components]
  std::vector<double> discriminants;  // [val rows x classes of the fold]
  IVec classes;                     // label codes of the fold's training classes
  IVec val_rows;
};

class PlsCv {
 public:
  // x [n x p] host rows; `classes` = number of label codes (0..classes-1).
  // With run_offsets, rows [run_offsets[r], run_offsets[r+1]) form run r: a
  // fold's ids belong to one run and it trains on that run's other rows only
  // (empty: one run of all rows).
  PlsCv(int device, const float* x, int n, int p, const kodama_hip_core_options& options, int classes,
        const IVec& run_offsets = IVec());
  ~PlsCv();
  PlsCv(const PlsCv&) = delete;
  PlsCv& operator=(const PlsCv&) = delete;

  // One F-fold evaluation of the labelling `codes` under `folds` (same
  // indexing). Writes the predicted code of every row; *selected = the
  // component count PLSLDACV reports (min over folds, capped by
  // max_components). Fold statistics are recomputed when `folds` differs from
  // the previous call's.
  void evaluate(const IVec& codes, const IVec& folds, int* pred_codes, int* selected);
  // The same on pinned host buffers of n entries: the caller writes the
  // codes into staged_codes() and reads the predictions from staged_pred()
  // after the call (no copies through pageable memory).
  void evaluate_staged(const IVec& folds, int* selected);
  int* staged_codes();
  const int* staged_pred() const;
  // Builds the fold cache for `folds` now (otherwise the first evaluation
  // with new folds does).
  void prepare(const IVec& folds);
  // Head of the fold fits: false (default) LDA (PLSLDACV), true the
  // least-squares PLS-DA head (PLSDACV).
  void set_da(bool da);
  // Detail of fold `fold` of the last evaluation.
  PlsFoldDetail detail(int fold) const;
  // Average duration of the fused evaluation kernel (HIP events).
  double eval_kernel_ms_average() const;

 private:
  struct Impl;
  Impl* impl_ = nullptr;
};

// General PLS fit (fit_general_pls, plscv.cpp:2765-2878) on the device in
// FP64: column statistics, the cross-products Xs'[Xs | Yc] on the matrix
// cores, power-refresh SIMPLS, response weights, coefficients, scores and
// fitted values. Outputs as the reference's PLSFitResult (float, row-major).
struct PlsGeneralResult {
  int components = 0;
  std::vector<float> weights, response_loadings, scores, coefficients, fitted, x_center, x_scale, y_center;
  double cross_ms = 0.0;   // the MFMA cross-product kernel (HIP events)
  double total_ms = 0.0;   // every kernel of the fit
};
PlsGeneralResult pls_general_fit(int device, const float* x, const float* y, int n, int p, int q,
                                 const kodama_hip_core_options& options);

// std::mt19937(1) + std::normal_distribution<float>(0, 1), the first power
// iteration start of every SIMPLS fit (plscv.cpp:844-846), widened to double.
std::vector<double> pls_start_vector(int p);

}  // namespace kodama_hip
