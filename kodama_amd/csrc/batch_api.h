// The M-run kNN evolution behind kodama_hip_core_knn_batch, shared with the
// learned-proposal entry points (capi_mlp.cpp).
#pragma once

#include <cstdint>

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
};

// mode 0: plain (host proposals with the reference streams; with `trace`, the
// accept/reject samples are recorded); 1: device-resident; 2: device-resident
// with the learned proposal (mlp_params host [kMlpParams], uniform fallback
// probability `fallback`). Outputs as kodama_hip_core_knn_batch.
void run_knn_batch(const BatchRequest& q, int mode, const float* mlp_params, float fallback, TraceSink* trace,
                   int32_t* clbest, double* accbest, double* vect_acc, double* timings);

}  // namespace kodama_hip
