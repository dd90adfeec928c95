// Host driver of the exact kNN on the device: the split-BF16 MFMA candidates
// with the exact rerank where they apply, the register scan otherwise and for
// the queries the rerank flags.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
#include <map>
#include <mutex>
#include <tuple>
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
constexpr int kTableTargetBlocks = 2048; // sliced rescans of flagged table rows
constexpr int kMfmaSecondPassMin = 4096; // flagged queries for a 16-entry MFMA pass before the scan

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
  const size_t cand = static_cast<size_t>(query_rows) * 16 * mfma_list_len(k);
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
  m.waves = 4;  // two 4-wave blocks per CU overlap each other's staging barriers
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
  const int* d_ex_sub = d_exclude ? ex_sub.get() : nullptr;
  if (mfma_list(m) == 8 && nr >= kMfmaSecondPassMin) {
    // Second pass over the flagged queries with 16 candidates per lane column
    // (most certify there); only queries flagged again take the scan. Its
    // workgroups stream every train row, so it pays once the flagged queries
    // would keep the sliced scan busy for longer.
    DeviceBuffer<unsigned short> s_hi(static_cast<size_t>(nr) * 64), s_lo(s_hi.size());
    DeviceBuffer<float> s_n2(static_cast<size_t>(nr));
    check(launch_split_rows(q_sub.get(), nr, dims, s_hi.get(), s_lo.get(), s_n2.get(), s), "split_rows launch");
    const size_t cand2 = static_cast<size_t>(nr) * 16 * 16;
    DeviceBuffer<float> c2_val(cand2);
    DeviceBuffer<int> c2_idx(cand2), flags2(static_cast<size_t>(nr));
    KnnMfmaParams m2 = m;
    m2.query = q_sub.get();
    m2.query_hi = s_hi.get();
    m2.query_lo = s_lo.get();
    m2.query_norm2 = s_n2.get();
    m2.query_rows = nr;
    m2.exclude = d_ex_sub;
    m2.cand_value = c2_val.get();
    m2.cand_index = c2_idx.get();
    m2.flags = flags2.get();
    m2.out_index = r_idx.get();
    m2.out_dist = r_dist.get();
    m2.list = 16;
    check(launch_knn_mfma(m2, euclid, s), "knn_mfma launch");
    std::vector<int> h_flags2(static_cast<size_t>(nr));
    flags2.download(h_flags2.data(), h_flags2.size(), s);
    check(hipStreamSynchronize(s), "knn_mfma sync");
    std::vector<int> redo2;
    for (int i = 0; i < nr; ++i) if (h_flags2[i]) redo2.push_back(i);
    if (mfma_fallbacks) *mfma_fallbacks = static_cast<int>(redo2.size());
    if (!redo2.empty()) {
      const int n2 = static_cast<int>(redo2.size());
      DeviceBuffer<int> d_rows2(static_cast<size_t>(n2));
      d_rows2.upload(redo2.data(), redo2.size(), s);
      DeviceBuffer<float> q_sub2(static_cast<size_t>(n2) * dims), r2_dist(static_cast<size_t>(n2) * k);
      DeviceBuffer<int> r2_idx(static_cast<size_t>(n2) * k), ex_sub2;
      check(launch_gather_rows(q_sub.get(), d_rows2.get(), n2, dims, q_sub2.get(), s), "gather_rows launch");
      if (d_exclude) {
        std::vector<int> h_ex(static_cast<size_t>(query_rows)), sub2(redo2.size());
        check(hipMemcpyAsync(h_ex.data(), d_exclude, h_ex.size() * sizeof(int), hipMemcpyDeviceToHost, s),
              "download");
        check(hipStreamSynchronize(s), "exclusion sync");
        for (int i = 0; i < n2; ++i) sub2[i] = h_ex[redo[redo2[i]]];
        ex_sub2.resize(sub2.size());
        ex_sub2.upload(sub2.data(), sub2.size(), s);
      }
      scan(q_sub2.get(), n2, d_exclude ? ex_sub2.get() : nullptr, r2_idx.get(), r2_dist.get());
      check(launch_scatter_results(r2_idx.get(), r2_dist.get(), d_rows2.get(), n2, k, r_idx.get(), r_dist.get(), s),
            "scatter_results launch");
    }
  } else {
    scan(q_sub.get(), nr, d_ex_sub, r_idx.get(), r_dist.get());
  }
  check(launch_scatter_results(r_idx.get(), r_dist.get(), d_rows.get(), nr, k, d_idx, d_dist, s),
        "scatter_results launch");
  check(hipStreamSynchronize(s), "fallback sync");
}

namespace {

// Tuning cache of the per-fold tables: shape bucket -> the fastest schedule.
struct TableShape {
  int p, kmax, euclid, rows_log2, total_log2, folds;
  bool operator<(const TableShape& o) const {
    return std::tie(p, kmax, euclid, rows_log2, total_log2, folds) <
           std::tie(o.p, o.kmax, o.euclid, o.rows_log2, o.total_log2, o.folds);
  }
};
// Shapes the tuner does not time: the largest run and the landmark rows in
// all runs at or above these take the MFMA path.
constexpr int kMfmaSureRows = 2048;
constexpr long long kMfmaSureTotal = 1ll << 16;
std::mutex g_tune_mutex;
std::map<TableShape, TableSchedule> g_tuned;
std::atomic<int> g_table_path{0}, g_table_list{0}, g_table_waves{0};
thread_local TableSchedule g_last_table;

int ceil_log2(int v) {
  int b = 0;
  while ((1 << b) < v) ++b;
  return b;
}

double seconds_of(const std::function<void()>& f, hipStream_t s) {
  check(hipStreamSynchronize(s), "tuning sync");
  const auto t0 = std::chrono::steady_clock::now();
  f();
  check(hipStreamSynchronize(s), "tuning sync");
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

void set_fold_table_path(int path) { set_fold_table_schedule(TableSchedule{path, 0, 0}); }
int fold_table_path() { return g_table_path.load(); }
int last_fold_table_path() { return g_last_table.path; }
void set_fold_table_schedule(const TableSchedule& t) {
  g_table_path.store(t.path);
  g_table_list.store(t.list);
  g_table_waves.store(t.waves);
}
TableSchedule last_fold_table_schedule() { return g_last_table; }

int device_fold_tables(const FoldTableJob& job, hipStream_t s) {
  const bool euclid = job.metric == KODAMA_HIP_EUCLIDEAN;
  if (job.fallbacks) *job.fallbacks = -1;  // -1: the register scan served every row
  const int runs = static_cast<int>(job.run_offsets.size()) - 1;
  const int total = job.run_offsets[runs];
  const bool mfma_ok = job.p <= 64 && job.k <= kMfmaMaxK;
  int path = job.path != 0 ? job.path : g_table_path.load();
  if (path == 2 && !mfma_ok) path = 1;

  // Register scan of every row of every run, or of listed rows only.
  auto scan = [&](const int* d_qlist, const int* d_qlist_offsets, int max_queries) {
    KnnTopkParams prm;
    prm.train = job.d_x;
    prm.dims = job.p;
    prm.k = job.k;
    prm.rows = job.d_rows;
    prm.folds = job.d_folds;
    prm.run_offsets = job.d_run_offsets;
    prm.pad_offsets = job.d_pad_offsets;
    prm.pad_total = job.pad_total;
    prm.qlist = d_qlist;
    prm.qlist_offsets = d_qlist_offsets;
    prm.out_index = job.d_nbr;
    prm.out_dist = job.d_score;
    check(launch_knn_topk(prm, 1, euclid, max_queries, runs, s), "knn_topk launch");
  };
  // Split-BF16 MFMA candidates (with `list` candidates per lane column and
  // `waves` waves of queries per workgroup), exact rerank, the flagged rows
  // rescanned.
  auto mfma = [&](int list, int waves) {
    DeviceBuffer<float> xc(static_cast<size_t>(total) * job.p);
    check(launch_gather_rows(job.d_x, job.d_rows, total, job.p, xc.get(), s), "gather_rows launch");
    DeviceBuffer<unsigned short> hi(static_cast<size_t>(total) * 64), lo(hi.size());
    DeviceBuffer<float> n2(static_cast<size_t>(total));
    check(launch_split_rows(xc.get(), total, job.p, hi.get(), lo.get(), n2.get(), s), "split_rows launch");
    const size_t cand = static_cast<size_t>(total) * 16 * list;
    DeviceBuffer<float> c_val(cand), r_dist(static_cast<size_t>(total) * job.k);
    DeviceBuffer<int> c_idx(cand), flags(static_cast<size_t>(total)), r_idx(r_dist.size());
    KnnMfmaParams m;
    m.train = xc.get();
    m.query = xc.get();
    m.train_hi = m.query_hi = hi.get();
    m.train_lo = m.query_lo = lo.get();
    m.train_norm2 = m.query_norm2 = n2.get();
    m.train_norm2_max = job.norm2_max;
    m.train_rows = m.query_rows = total;
    m.dims = job.p;
    m.k = job.k;
    m.run_offsets = job.d_run_offsets;
    m.folds = job.d_folds;
    m.cand_value = c_val.get();
    m.cand_index = c_idx.get();
    m.flags = flags.get();
    m.out_index = r_idx.get();
    m.out_dist = r_dist.get();
    m.list = list;
    m.waves = waves;
    check(launch_knn_mfma_tables(m, euclid, runs, job.max_rows, s), "knn_mfma_tables launch");
    check(launch_publish_tables(r_idx.get(), r_dist.get(), flags.get(), job.d_run_offsets, runs, job.d_pad_offsets,
                                job.pad_total, total, job.k, job.d_nbr, job.d_score, s),
          "publish_tables launch");
    std::vector<int> h_flags(static_cast<size_t>(total));
    flags.download(h_flags.data(), h_flags.size(), s);
    check(hipStreamSynchronize(s), "fold tables sync");
    std::vector<int> qlist, qrun, qoff(static_cast<size_t>(runs) + 1, 0);
    int max_q = 0;
    for (int r = 0; r < runs; ++r) {
      for (int i = job.run_offsets[r]; i < job.run_offsets[r + 1]; ++i) {
        if (h_flags[i]) {
          qlist.push_back(i - job.run_offsets[r]);
          qrun.push_back(r);
        }
      }
      qoff[r + 1] = static_cast<int>(qlist.size());
      max_q = std::max(max_q, qoff[r + 1] - qoff[r]);
    }
    const int nflag = static_cast<int>(qlist.size());
    if (job.fallbacks) *job.fallbacks = nflag;
    if (nflag == 0) return;
    DeviceBuffer<int> d_ql(qlist.size()), d_qo(qoff.size()), d_qr(qrun.size());
    d_ql.upload(qlist.data(), qlist.size(), s);
    d_qo.upload(qoff.data(), qoff.size(), s);
    d_qr.upload(qrun.data(), qrun.size(), s);
    // A few flagged rows per run would leave one block per run scanning the
    // whole run; their train rows are cut into slices (about 2048 blocks in
    // all, at most 64 slices) whose lists are merged.
    const int gx = (max_q + kScanBlock - 1) / kScanBlock;
    int slices = std::min(kMaxSlices, std::max(1, (kTableTargetBlocks + gx * runs - 1) / (gx * runs)));
    int slice_rows = (job.max_rows + slices - 1) / slices;
    slice_rows = std::max(kSliceAlign, (slice_rows + kSliceAlign - 1) / kSliceAlign * kSliceAlign);
    slices = (job.max_rows + slice_rows - 1) / slice_rows;
    if (slices <= 1) {
      scan(d_ql.get(), d_qo.get(), max_q);
    } else {
      DeviceBuffer<int> part_idx(static_cast<size_t>(slices) * nflag * job.k), m_idx(static_cast<size_t>(nflag) * job.k);
      DeviceBuffer<float> part_dist(part_idx.size()), m_dist(m_idx.size());
      KnnTopkParams prm;
      prm.train = job.d_x;
      prm.dims = job.p;
      prm.k = job.k;
      prm.rows = job.d_rows;
      prm.folds = job.d_folds;
      prm.run_offsets = job.d_run_offsets;
      prm.pad_offsets = job.d_pad_offsets;
      prm.pad_total = job.pad_total;
      prm.qlist = d_ql.get();
      prm.qlist_offsets = d_qo.get();
      prm.list_total = nflag;
      prm.slice_rows = slice_rows;
      prm.out_index = part_idx.get();
      prm.out_dist = part_dist.get();
      check(launch_knn_topk(prm, 1, euclid, max_q, runs, s, slices), "knn_topk launch");
      check(launch_knn_merge_slices(part_idx.get(), part_dist.get(), slices, nflag, job.k, m_idx.get(), m_dist.get(), s),
            "knn_merge_slices launch");
      check(launch_publish_listed(m_idx.get(), m_dist.get(), d_ql.get(), d_qr.get(), job.d_pad_offsets,
                                  job.pad_total, nflag, job.k, job.d_nbr, job.d_score, s),
            "publish_listed launch");
    }
    check(hipStreamSynchronize(s), "fold tables sync");
  };

  // MFMA schedules: lists of 8 (k <= 32) and 16 candidates, 4 and 8 waves
  // per workgroup; a forced list or wave count narrows them.
  std::vector<TableSchedule> schedules;
  for (int list : {8, 16}) {
    if (list == 8 && job.k > 32) continue;
    if (g_table_list.load() != 0 && list != g_table_list.load()) continue;
    for (int waves : {8, 4}) {
      if (g_table_waves.load() != 0 && waves != g_table_waves.load()) continue;
      schedules.push_back(TableSchedule{2, list, waves});
    }
  }
  // A forced list that the job's k excludes leaves the scan.
  const bool mfma_usable = mfma_ok && !schedules.empty();
  TableSchedule choice{path == 2 && !mfma_usable ? 1 : path, 0, 0};
  if (path == 0) {
    if (!mfma_usable) {
      choice = TableSchedule{1, 0, 0};
    } else {
      // k enters by its power-of-two bucket (8, 16, 32, 64, 128): the
      // schedules open to a k differ across 32 (lists of 8 need k <= 32).
      const TableShape key{job.p, 1 << std::max(3, ceil_log2(job.k)), euclid ? 1 : 0, ceil_log2(job.max_rows),
                           ceil_log2(total), job.folds};
      bool cached = false;
      {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
          choice = it->second;
          cached = true;
        }
      }
      if (!cached) {
        // First use of this shape bucket: time every candidate schedule (they
        // give identical tables) and keep the fastest. Runs this large sit
        // well inside the region where the MFMA path measured several times
        // faster than the scan (7.1 against 29.7 ms for 100 runs of
        // 1000-5000 rows, p = 50, k = 30): the scan is not timed there.
        const bool sure = job.max_rows >= kMfmaSureRows && total >= kMfmaSureTotal;
        double best = 0.0;
        if (!sure) {
          best = seconds_of([&] { scan(nullptr, nullptr, job.max_rows); }, s);
          choice = TableSchedule{1, 0, 0};
        }
        for (const TableSchedule& t : schedules) {
          const double sec = seconds_of([&] { mfma(t.list, t.waves); }, s);
          if (choice.path == 0 || sec < best) {
            best = sec;
            choice = t;
          }
        }
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tuned[key] = choice;
        g_last_table = choice;
        return choice.path;
      }
    }
  } else if (choice.path == 2) {
    choice = schedules.front();
  }
  if (choice.path == 2) {
    mfma(choice.list, choice.waves);
  } else {
    choice = TableSchedule{1, 0, 0};
    scan(nullptr, nullptr, job.max_rows);
  }
  g_last_table = choice;
  return choice.path;
}

}  // namespace kodama_hip
