// Host driver of the exact kNN on the device: the split-BF16 MFMA candidates
// with the exact rerank where they apply, the register scan otherwise and for
// the queries the rerank flags.
#include <algorithm>
#include <vector>

#include "../../include/kodama_hip.h"
#include "device.h"
#include "knn_mfma.h"
#include "launch.h"

namespace kodama_hip {

namespace {

constexpr int kScanBlock = 256;          // queries per register-scan block
constexpr int kScanTargetBlocks = 512;   // two blocks per CU
constexpr int kMaxSlices = 64;           // lists one merge wave combines
constexpr int kSliceAlign = 64;          // LDS tile rows of the scan

}  // namespace

void device_exact_knn(const float* d_train, int train_rows, const float* d_query, int query_rows, int dims, int k,
                      int metric, const int* d_exclude, int* d_idx, float* d_dist, float train_norm2_max,
                      hipStream_t s, int* mfma_fallbacks) {
  const bool euclid = metric == KODAMA_HIP_EUCLIDEAN;
  if (mfma_fallbacks) *mfma_fallbacks = -1;  // -1: the MFMA path was not used
  // Register scan. Few queries (one block per 256) would leave most CUs idle,
  // so the train rows are cut into slices scanned by separate blocks, whose
  // lists are merged after: about 512 blocks in all, at most 64 slices.
  auto scan = [&](const float* q, int rows, const int* ex, int* idx, float* dist) {
    KnnTopkParams prm;
    prm.train = d_train;
    prm.query = q;
    prm.train_rows = train_rows;
    prm.query_rows = rows;
    prm.dims = dims;
    prm.k = k;
    prm.exclude = ex;
    prm.out_index = idx;
    prm.out_dist = dist;
    const int gx = (rows + kScanBlock - 1) / kScanBlock;
    int slices = std::min(kMaxSlices, std::max(1, (kScanTargetBlocks + gx - 1) / gx));
    int slice_rows = (train_rows + slices - 1) / slices;
    slice_rows = std::max(kSliceAlign, (slice_rows + kSliceAlign - 1) / kSliceAlign * kSliceAlign);
    slices = (train_rows + slice_rows - 1) / slice_rows;
    if (slices <= 1) {
      check(launch_knn_topk(prm, 0, euclid, rows, 1, s), "knn_topk launch");
      return;
    }
    DeviceBuffer<int> part_idx(static_cast<size_t>(slices) * rows * k);
    DeviceBuffer<float> part_dist(part_idx.size());
    prm.slice_rows = slice_rows;
    prm.out_index = part_idx.get();
    prm.out_dist = part_dist.get();
    check(launch_knn_topk(prm, 0, euclid, rows, slices, s), "knn_topk launch");
    check(launch_knn_merge_slices(part_idx.get(), part_dist.get(), slices, rows, k, idx, dist, s),
          "knn_merge_slices launch");
    check(hipStreamSynchronize(s), "knn scan sync");  // the partial lists are freed on return
  };
  if (dims > 64 || k > kMfmaMaxK || train_rows < kMfmaMinTrain || query_rows == 0) {
    scan(d_query, query_rows, d_exclude, d_idx, d_dist);
    return;
  }
  const bool same = d_query == d_train;
  DeviceBuffer<unsigned short> t_hi(static_cast<size_t>(train_rows) * 64), t_lo(t_hi.size());
  DeviceBuffer<float> t_n2(static_cast<size_t>(train_rows));
  check(launch_split_rows(d_train, train_rows, dims, t_hi.get(), t_lo.get(), t_n2.get(), s), "split_rows launch");
  DeviceBuffer<unsigned short> q_hi, q_lo;
  DeviceBuffer<float> q_n2;
  if (!same) {
    q_hi.resize(static_cast<size_t>(query_rows) * 64);
    q_lo.resize(q_hi.size());
    q_n2.resize(static_cast<size_t>(query_rows));
    check(launch_split_rows(d_query, query_rows, dims, q_hi.get(), q_lo.get(), q_n2.get(), s), "split_rows launch");
  }
  const size_t cand = static_cast<size_t>(query_rows) * 16 * kMfmaListLen;
  DeviceBuffer<float> c_val(cand);
  DeviceBuffer<int> c_idx(cand), flags(static_cast<size_t>(query_rows));
  KnnMfmaParams m;
  m.train = d_train;
  m.query = d_query;
  m.train_hi = t_hi.get();
  m.train_lo = t_lo.get();
  m.query_hi = same ? t_hi.get() : q_hi.get();
  m.query_lo = same ? t_lo.get() : q_lo.get();
  m.train_norm2 = t_n2.get();
  m.query_norm2 = same ? t_n2.get() : q_n2.get();
  m.train_norm2_max = train_norm2_max;
  m.train_rows = train_rows;
  m.query_rows = query_rows;
  m.dims = dims;
  m.k = k;
  m.exclude = d_exclude;
  m.cand_value = c_val.get();
  m.cand_index = c_idx.get();
  m.flags = flags.get();
  m.out_index = d_idx;
  m.out_dist = d_dist;
  check(launch_knn_mfma(m, euclid, s), "knn_mfma launch");
  std::vector<int> h_flags(static_cast<size_t>(query_rows));
  flags.download(h_flags.data(), h_flags.size(), s);
  check(hipStreamSynchronize(s), "knn_mfma sync");
  std::vector<int> redo;
  for (int q = 0; q < query_rows; ++q) if (h_flags[q]) redo.push_back(q);
  if (mfma_fallbacks) *mfma_fallbacks = static_cast<int>(redo.size());
  if (redo.empty()) return;
  const int nr = static_cast<int>(redo.size());
  DeviceBuffer<int> d_rows(static_cast<size_t>(nr));
  d_rows.upload(redo.data(), redo.size(), s);
  DeviceBuffer<float> q_sub(static_cast<size_t>(nr) * dims), r_dist(static_cast<size_t>(nr) * k);
  DeviceBuffer<int> r_idx(static_cast<size_t>(nr) * k), ex_sub;
  check(launch_gather_rows(d_query, d_rows.get(), nr, dims, q_sub.get(), s), "gather_rows launch");
  if (d_exclude) {
    std::vector<int> h_ex(static_cast<size_t>(query_rows)), sub(redo.size());
    // Exclusions of the flagged queries (host copy: the list is short).
    check(hipMemcpyAsync(h_ex.data(), d_exclude, h_ex.size() * sizeof(int), hipMemcpyDeviceToHost, s), "download");
    check(hipStreamSynchronize(s), "exclusion sync");
    for (int i = 0; i < nr; ++i) sub[i] = h_ex[redo[i]];
    ex_sub.resize(sub.size());
    ex_sub.upload(sub.data(), sub.size(), s);
  }
  scan(q_sub.get(), nr, d_exclude ? ex_sub.get() : nullptr, r_idx.get(), r_dist.get());
  check(launch_scatter_results(r_idx.get(), r_dist.get(), d_rows.get(), nr, k, d_idx, d_dist, s),
        "scatter_results launch");
  check(hipStreamSynchronize(s), "fallback sync");
}

}  // namespace kodama_hip
