// Kernels of the KODAMA.matrix orchestration (matrix_kernels.hip): the Lloyd
// steps of the reference k-means in its own float order, and the projection of
// non-landmark rows through the global graph.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace kodama_hip {

// out[r] = dot(x_r, x_r) in the reference's float order (see matrix_kernels.hip).
hipError_t launch_sse_norms(const float* x, int rows, int p, float* out, hipStream_t s);

// Nearest centroid of every row (smallest (|x|^2 + |c|^2) - 2 x.c, first
// cluster on ties); rows whose assignment changes are counted in *changed.
// With runs > 1, `runs` problems of n rows and k clusters each, laid out one
// after another (x, x_norm, centroids, c_norm, assign, changed strided by
// problem), solved by one launch; active (may be null) skips problems whose
// entry is 0.
hipError_t launch_kmeans_assign(const float* x, int n, int p, const float* x_norm, const float* centroids,
                                const float* c_norm, int k, int* assign, int* changed, hipStream_t s,
                                int runs = 1, const int* active = nullptr);

// Per cluster: member rows summed in ascending row order, then scaled by
// 1 / count (clusters without members stay zero); sizes as float counts.
// Batched as launch_kmeans_assign.
hipError_t launch_kmeans_update(const float* x, int n, int p, const int* assign, int k, float* centroids,
                                float* sizes, hipStream_t s, int runs = 1, const int* active = nullptr);

struct ProjectParams {
  const int* graph = nullptr;        // [n x neighbors] 0-based ids, -1 = none
  int n = 0;
  int neighbors = 0;
  const int* landmarks = nullptr;    // per run, ascending data rows, concatenated
  const int* run_offsets = nullptr;  // [runs + 1]
  const int* clbest = nullptr;       // per landmark, same indexing as landmarks
  int runs = 0;
  int votes = 0;                     // most landmark neighbours that vote (1..64)
  int max_landmarks = 0;             // largest run's landmark count
  int* local = nullptr;              // workspace [runs x n]: landmark position of each row, or -1
  int* res = nullptr;                // [runs x n]
};

// res[r][i]: the landmark's own label, else the majority label (smallest on
// ties) of the first `votes` landmarks in row i's graph neighbours, else the
// run's first landmark label. Fills prm.local first (row -> position in the
// run's landmark list), so each lookup is one load.
hipError_t launch_project(const ProjectParams& prm, hipStream_t s);

}  // namespace kodama_hip
