// Kernel parameter blocks and host launchers shared by the HIP translation
// units and the C++ host engine of libkodama_hip.so.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace kodama_hip {

constexpr int kMaxK = 64;       // neighbours kept per row by the register top-k
constexpr int kMaxScanK = 128;  // query-vs-train scan with at most 128 dimensions
constexpr int kMaxDims = 256;   // feature dimensions of the register query path

// Exact kNN. MODE 0 (query vs train): query/train row-major, out_index/out_dist
// [query_rows x k], internal distances; with slice_rows > 0 block row y scans
// train rows [y * slice_rows, (y + 1) * slice_rows) into the lists
// [y][query_rows x k] (global train indices), which launch_knn_merge_slices
// combines. MODE 1 (per-fold tables of a batch of
// runs): rows[] are the data rows of each run, concatenated with run_offsets;
// a row's candidates are the rows of its run in another fold; output goes to the
// transposed padded vote layout [k x pad_total] as padded row ids and scores.
struct KnnTopkParams {
  const float* train = nullptr;
  const float* query = nullptr;
  int train_rows = 0;
  int query_rows = 0;
  int dims = 0;
  int k = 0;
  const int* exclude = nullptr;      // MODE 0, per query, may be null
  int slice_rows = 0;                // MODE 0, train rows per block row (0: all)
  const int* rows = nullptr;         // MODE 1
  const int* folds = nullptr;        // MODE 1, per concatenated row
  const int* run_offsets = nullptr;  // MODE 1, [runs + 1]
  const int* pad_offsets = nullptr;  // MODE 1, [runs + 1]
  int pad_total = 0;                 // MODE 1
  // MODE 1, optional: only the run-local rows qlist[qlist_offsets[r] ..
  // qlist_offsets[r + 1]) of each run r are queried (null: every row). With
  // slice_rows > 0 as well, block z scans the run's rows [z * slice_rows,
  // (z + 1) * slice_rows) and writes its list of listed query i to
  // [z][list_total x k] (run-local ids, distances; -1 / +inf), for a merge.
  const int* qlist = nullptr;
  const int* qlist_offsets = nullptr;  // [runs + 1]
  int list_total = 0;
  int* out_index = nullptr;
  float* out_dist = nullptr;         // MODE 0 distances, MODE 1 scores (-distance)
};

// Batched vote over the padded layout (each run's rows start at a multiple of
// 256, so every block belongs to one run).
struct VoteParams {
  const int* neighbors = nullptr;    // [k x pad_total] padded row ids or -1
  const float* scores = nullptr;     // [k x pad_total]
  const int* labels = nullptr;       // [pad_total] label codes
  const int* block_run = nullptr;    // [blocks]
  const int* run_offsets = nullptr;  // [runs + 1] unpadded offsets (row counts)
  const int* pad_offsets = nullptr;  // [runs + 1]
  const int* fallback_code = nullptr;  // [runs] code for rows without neighbours, may be null
  int pad_total = 0;
  int k = 0;
  int classes = 0;                   // codes per run (array stride)
  int max_rows = 0;                  // largest run (LDS label staging when it fits)
  int* predictions = nullptr;        // [pad_total]
  int* counts = nullptr;             // [runs x classes] label counts, null = skip reductions
  int* firsts = nullptr;             // [runs x classes] first local row of each label, may be null
  int* matches = nullptr;            // [runs] rows with prediction == label
  int runs = 0;                      // stride of the parity double buffers
  const int* cycle = nullptr;        // device cycle counter: reductions use buffer (cycle & 1)
  // Device form: labels of absorbed runs go through remap [runs x classes]
  // (-1 = unchanged) when decision[run] & remap_flag; the row's own remapped
  // label is written to relabel (the labels buffer itself).
  const int* remap = nullptr;
  const int* decision = nullptr;
  int remap_flag = 0;
  int* relabel = nullptr;
};

hipError_t launch_knn_topk(const KnnTopkParams& prm, int mode, bool euclid, int max_queries, int runs,
                           hipStream_t stream, int slices = 1);
// The k smallest by (distance, index) of `slices` lists [slices][query_rows x k]
// (each ascending, -1 = empty) -> [query_rows x k], -1 / +inf past the end.
hipError_t launch_knn_merge_slices(const int* idx, const float* dist, int slices, int query_rows, int k,
                                   int* out_index, float* out_dist, hipStream_t stream);
hipError_t launch_vote(const VoteParams& prm, int blocks, hipStream_t stream);
hipError_t launch_vote_reset(int* counts, int* firsts, int* matches, int runs, int classes,
                             hipStream_t stream);

}  // namespace kodama_hip
