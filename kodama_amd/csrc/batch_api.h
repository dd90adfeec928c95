// The M-run kNN evolution behind kodama_hip_core_knn_batch, shared with the
// learned-proposal entry points (capi_mlp.cpp).
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/kodama_hip.h"
#include "mlp.h"

namespace kodama_hip {

struct BatchRequest {
  const float* x;
  int n, p, runs;
  const int32_t* run_offsets;
  const int32_t* landmark_rows;
  const int32_t* initial_labels;
  const int32_t* fold_ids;
  const kodama_hip_core_options* options;
  // Optional, same indexing as landmark_rows: constraint groups and fixed
  // flags of the plain form (CoreKNN's constrain / fixed).
  const int32_t* constrain = nullptr;
  const int32_t* fixed = nullptr;
  // Optional caller tables [total x table_k] with run-local neighbour ids
  // (-1 / -inf padded) in place of the per-fold exact tables (CoreKNNGraph);
  // fold_ids must then be the folds the tables were built with.
  const int32_t* table_nbr = nullptr;
  const float* table_score = nullptr;
  int table_k = 0;
};

// mode 0: plain (host proposals with the reference streams; with `trace`, the
// accept/reject samples are recorded); 1: device-resident; 2: device-resident
// with the learned proposal (mlp_params host [kMlpParams], uniform fallback
// probability `fallback`). Outputs as kodama_hip_core_knn_batch.
void run_knn_batch(const BatchRequest& q, int mode, const float* mlp_params, float fallback, TraceSink* trace,
                   int32_t* clbest, double* accbest, double* vect_acc, double* timings);

// The M-run CorePLSLDA (kodama_hip_core_plslda_batch) with the request's
// optional constraint groups and fixed rows; caller tables are not used.
void run_plslda_batch(const BatchRequest& q, int32_t* clbest, double* accbest, double* vect_acc,
                      double* timings);

// PLSLDAPredict for `fits` independent fits: fit r trains on the data rows
// train_rows[train_off[r] .. train_off[r+1]) with `labels` (same indexing) and
// predicts the data rows test_rows[test_off[r] .. test_off[r+1]) into pred
// (same indexing). Options: max / fixed components, center, scale, device.
void plslda_predict_batch(const float* x, int n, int p, int fits, const std::vector<int>& train_off,
                          const int* train_rows, const int* labels, const std::vector<int>& test_off,
                          const int* test_rows, const kodama_hip_core_options& options, int* pred);

}  // namespace kodama_hip
