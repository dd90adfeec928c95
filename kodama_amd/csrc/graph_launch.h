// Launchers of the graph-side kernels (graph_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace kodama_hip {

// Size of the static device scratch of the allocation-free dissimilarity.
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;
// Device address and size of that scratch on `device`, which must be the
// current device (resolved with hipGetSymbolAddress on first use per device).
hipError_t noalloc_scratch(int device, int** ptr, size_t* bytes);

// In place: internal self-kNN distances and 0-based ids -> public distances and
// 1-based ids (missing entries: -1 / +inf). metric as KODAMA_HIP_* .
hipError_t launch_graph_publish(int* idx, float* dist, size_t total, int metric, hipStream_t s);
// res [runs x samples] -> labels [samples x stride] (stride >= runs, zero padded).
hipError_t launch_labels_transpose(const int* res, int runs, int samples, int stride, int* out, hipStream_t s);
// Dynamic LDS of the dissimilarity kernel; the launcher refuses more than 64 KiB.
size_t dissimilarity_lds_bytes(int neighbors);
// In place rewrite of a graph [samples x neighbors] from sample-major labels of
// `stride` (a multiple of 4) int32 per sample.
hipError_t launch_dissimilarity(int* indices, float* distances, int samples, int neighbors, const int* labels,
                                int stride, bool input_one_based, bool output_one_based, hipStream_t s);

}  // namespace kodama_hip
