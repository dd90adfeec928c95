// Split-BF16 MFMA exact kNN (knn_mfma.hip) and its host driver (knn_search.cpp).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <vector>

namespace kodama_hip {

constexpr int kMfmaMaxK = 128;        // k served by the MFMA path
// Candidates kept per (query, lane column): 16 x this per query.
inline int mfma_list_len(int k) { return k <= 32 ? 8 : 16; }
constexpr int kMfmaMinTrain = 4096;   // smaller train sets use the register scan
constexpr float kMfmaErrorScale = 8e-5f;  // E = scale * (|q|^2 + max |t|^2)

struct KnnMfmaParams {
  const float* train = nullptr;          // [train_rows x dims] f32 (exact re-score)
  const float* query = nullptr;          // [query_rows x dims] f32
  const unsigned short* train_hi = nullptr;  // [train_rows x 64] bf16 bits
  const unsigned short* train_lo = nullptr;
  const unsigned short* query_hi = nullptr;  // [query_rows x 64]
  const unsigned short* query_lo = nullptr;
  const float* train_norm2 = nullptr;    // [train_rows]
  const float* query_norm2 = nullptr;    // [query_rows]
  float train_norm2_max = 0.0f;
  float error_scale = kMfmaErrorScale;
  int train_rows = 0;
  int query_rows = 0;
  int dims = 0;
  int k = 0;
  const int* exclude = nullptr;          // [query_rows] train index to skip, may be null
  // Per-fold tables of a batch of runs (launch_knn_mfma_tables): query = train
  // = the concatenated rows of all runs; each row's candidates are the rows of
  // its run in another fold. Indices stay concatenated.
  const int* run_offsets = nullptr;      // [runs + 1]
  const int* folds = nullptr;            // [rows]
  float* cand_value = nullptr;           // [query_rows x 16 x mfma_list_len(k)]
  int* cand_index = nullptr;
  int* flags = nullptr;                  // [query_rows] 1 = needs the exact scan
  int* out_index = nullptr;              // [query_rows x k]
  float* out_dist = nullptr;             // internal distances
  // Schedule: candidates kept per (query, lane column) (0: mfma_list_len(k);
  // 8 serves k <= 32, 16 any k), and waves of 16 queries per workgroup (4 or
  // 8). Every schedule gives the same neighbours.
  int list = 0;
  int waves = 8;
  // Candidate lists as packed 32-bit keys (set by the launchers): the
  // order-preserving key of the candidate value with its low `index_bits`
  // bits replaced by the train row's index within its lane column (row / 16),
  // in cand_index; or, for larger train sets, values quantised to the same
  // bits (their lower limits in cand_value) with rows; -1 keeps separate
  // exact values and indices. `runs` is the batch's run count for the
  // rerank's run lookup.
  int index_bits = -1;
  int runs = 0;
};
// Largest column index width of the packed lists: the key keeps 21 of the
// value's 32 order bits, a relative quantisation of at most 2^-12.
constexpr int kMfmaMaxIndexBits = 11;
// Index bits of the per-segment keys of larger train sets (8-entry lists):
// segments of 8192 rows, values kept to 23 order bits (2^-14 relative).
constexpr int kMfmaSegmentBits = 9;
// Packed lists on (1, the default, where the train rows fit) or off (0).
void set_mfma_packed_lists(int on);
int mfma_packed_lists();
inline int mfma_list(const KnnMfmaParams& p) { return p.list != 0 ? p.list : mfma_list_len(p.k); }

hipError_t launch_split_rows(const float* x, int rows, int p, unsigned short* hi, unsigned short* lo, float* norm2,
                             hipStream_t s);
hipError_t launch_knn_mfma(const KnnMfmaParams& prm, bool euclid, hipStream_t s);
hipError_t launch_knn_mfma_tables(const KnnMfmaParams& prm, bool euclid, int runs, int max_rows, hipStream_t s);
hipError_t launch_publish_tables(const int* idx, const float* dist, const int* flags, const int* run_offsets,
                                 int runs, const int* pad_offsets, int pad_total, int rows, int k, int* out_index,
                                 float* out_score, hipStream_t s);
hipError_t launch_publish_listed(const int* idx, const float* dist, const int* qlist, const int* qrun,
                                 const int* pad_offsets, int pad_total, int count, int k, int* out_index,
                                 float* out_score, hipStream_t s);
// rows[i] of src [* x p] -> dst [count x p]; and the inverse scatter of [count x k] results.
hipError_t launch_gather_rows(const float* src, const int* rows, int count, int p, float* dst, hipStream_t s);
hipError_t launch_scatter_results(const int* src_idx, const float* src_dist, const int* rows, int count, int k,
                                  int* dst_idx, float* dst_dist, hipStream_t s);

// Exact kNN of device-resident rows (internal distances, (distance, index)
// order, -1 / +inf beyond the available rows), bit-identical to the register
// scan: the MFMA path when dims <= 64, k <= 128 and train_rows >= 4096,
// otherwise the register scan. Queries the 8-entry lists (k <= 32) cannot
// certify take, when there are 4096 or more, a second MFMA pass with 16-entry
// lists; the rest (those it flags again, fewer, or of 16-entry lists) take
// the register scan. `train_norm2_max`
// bounds the squared norms of the train rows (prepared rows for cosine).
void device_exact_knn(const float* d_train, int train_rows, const float* d_query, int query_rows, int dims, int k,
                      int metric, const int* d_exclude, int* d_idx, float* d_dist, float train_norm2_max,
                      hipStream_t s, int* mfma_fallbacks = nullptr);

// Per-fold kNN tables of a batch of runs (precompute_knn_cv_cuda,
// core.cpp:694-782, for all runs at once) into the transposed padded vote
// layout [k x pad_total] (padded row ids, scores = minus the internal
// distance; -1 / -inf past the available rows). Two paths give identical
// tables: 1 the register scan, 2 split-BF16 MFMA candidates with the exact
// rerank (p <= 64, k <= 32; flagged rows rescanned). path 0 follows the
// process-wide setting (set_fold_table_path), whose 0 tunes: on the first
// table of a shape bucket (p, k class, metric, log2 of the largest run and of
// all rows, folds) both paths run and timed, and the faster serves that bucket
// from then on. Returns the path used.
// Schedule of a per-fold table build: path 1 scan, 2 MFMA (with `list`
// candidates per lane column and `waves` waves per workgroup); 0 = tuned.
struct TableSchedule {
  int path = 0;
  int list = 0;
  int waves = 0;
};
void set_fold_table_schedule(const TableSchedule& t);
TableSchedule last_fold_table_schedule();

struct FoldTableJob {
  const float* d_x = nullptr;       // prepared rows [n x p]
  int p = 0, k = 0, metric = 0;
  const int* d_rows = nullptr;      // concatenated data rows of the runs
  const int* d_folds = nullptr;     // fold of each concatenated row
  std::vector<int> run_offsets;     // host copy, [runs + 1]
  const int* d_run_offsets = nullptr;
  const int* d_pad_offsets = nullptr;
  int pad_total = 0, max_rows = 0, folds = 0;
  float norm2_max = 0.0f;           // bound of the rows' squared norms
  int path = 0;
  int* d_nbr = nullptr;
  float* d_score = nullptr;
  int* fallbacks = nullptr;         // rows the MFMA path rescanned (may be null)
};
int device_fold_tables(const FoldTableJob& job, hipStream_t s);
void set_fold_table_path(int path);
int fold_table_path();
int last_fold_table_path();

}  // namespace kodama_hip
