// C ABI of libkodama_hip.so, orchestration half: KODAMA.matrix for the kNN
// classifier (run_kodama_matrix, kodama_matrix.cpp:3324-4051, and
// run_iteration, 2870-3322) on one device. The shared k-means atlas, the
// per-run k-means starting labels and the global graph run on the device; the
// landmark quotas draw on the host from the reference's streams; the M Core
// runs go through the batched engine (plain or device-resident form); the other
// rows are projected through the global graph on the device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <numeric>
#include <random>
#include <string>

#include "../../include/kodama_hip.h"
#include "capi_common.h"
#include "engine.h"
#include "graph_launch.h"
#include "knn_mfma.h"
#include "launch.h"
#include "matrix.h"
#include "matrix_launch.h"

using namespace kodama_hip;

namespace {

void require(bool ok, const char* msg) {
  if (!ok) throw InvalidArgument(msg);
}

struct MatrixSizes {
  int landmarks = 0;   // L per run
  int splitting = 0;   // k-means classes (atlas and starting labels)
  int coarse_k = 0;    // atlas classes
  int neighbors = 0;   // global graph width
};

// Option resolution of run_kodama_matrix (kodama_matrix.cpp:3340-3346,
// 3463-3467, 3662-3667) and run_iteration (2889-2895).
MatrixSizes resolve(int n, const kodama_hip_matrix_options& o) {
  MatrixSizes z;
  int L = o.landmarks > 0 ? o.landmarks : 10000;
  if (L >= n) L = static_cast<int>(std::ceil(static_cast<double>(n) * 0.75));
  z.landmarks = std::max(2, std::min(L, n - 1));
  z.splitting = o.splitting > 0 ? o.splitting : (n < 40000 ? 100 : 300);
  z.coarse_k = std::max(2, std::min(z.splitting, n));
  const int requested = o.graph_neighbors > 0 ? o.graph_neighbors : 100;
  z.neighbors = std::max(1, static_cast<int>(std::floor(std::min(
    {static_cast<double>(z.landmarks), static_cast<double>(n) * 0.75 - 1.0, static_cast<double>(requested)}))));
  return z;
}

double seconds_since(std::chrono::steady_clock::time_point& t) {
  const auto now = std::chrono::steady_clock::now();
  const double s = std::chrono::duration<double>(now - t).count();
  t = now;
  return s;
}

}  // namespace

extern "C" {

int kodama_hip_matrix_sizes(int32_t n, const kodama_hip_matrix_options* options, int32_t* landmarks,
                            int32_t* graph_neighbors) {
  return guarded([&] {
    require(options != nullptr && landmarks != nullptr && graph_neighbors != nullptr, "null pointer.");
    require(n >= 3, "KODAMAMatrix requires at least 3 rows.");
    const MatrixSizes z = resolve(n, *options);
    *landmarks = z.landmarks;
    *graph_neighbors = z.neighbors;
  });
}

}  // extern "C"

namespace {

// A caller-supplied neighbour graph (KODAMAMatrixFromGraphData).
struct ExternalGraph {
  const int32_t* indices = nullptr;
  const float* distances = nullptr;
  int neighbors = 0;
  int index_base = -1;  // 0, 1, or -1: one-based when every id is in 1..n
};

// normalize_external_graph (kodama_matrix.cpp:1292-1330): the first `width`
// valid neighbours of each row (in range, not the row itself, finite
// distance), 0-based, distances clamped at 0; -1 / +inf past them.
void normalize_graph(const ExternalGraph& g, int n, int width, IVec& idx, std::vector<float>& dist) {
  require(g.indices != nullptr && g.distances != nullptr, "null graph.");
  require(g.neighbors > 0, "NeighborGraph.neighbors must be positive.");
  int lo = std::numeric_limits<int>::max(), hi = std::numeric_limits<int>::min();
  const size_t total = static_cast<size_t>(n) * g.neighbors;
  for (size_t e = 0; e < total; ++e) {
    const int id = g.indices[e];
    if (id < 0) continue;
    lo = std::min(lo, id);
    hi = std::max(hi, id);
  }
  const bool one_based = g.index_base == 1 || (g.index_base < 0 && lo >= 1 && hi <= n);
  idx.assign(static_cast<size_t>(n) * width, -1);
  dist.assign(idx.size(), std::numeric_limits<float>::infinity());
  for (int i = 0; i < n; ++i) {
    int col = 0;
    for (int j = 0; j < g.neighbors && col < width; ++j) {
      const size_t src = static_cast<size_t>(i) * g.neighbors + j;
      int id = g.indices[src];
      if (one_based && id > 0) --id;
      if (id < 0 || id >= n || id == i) continue;
      const float d = g.distances[src];
      if (!std::isfinite(d)) continue;
      idx[static_cast<size_t>(i) * width + col] = id;
      dist[static_cast<size_t>(i) * width + col] = std::max(0.0f, d);
      ++col;
    }
  }
}

void matrix_impl(const float* x, int32_t n, int32_t p, const kodama_hip_matrix_options* options,
                 const ExternalGraph* ext, int32_t* res, double* acc, double* v, int32_t* landmark_rows,
                 int32_t* graph_indices_one_based, float* graph_distances, double* timings) {
  {
    require(x != nullptr && options != nullptr && res != nullptr && acc != nullptr, "null pointer.");
    const kodama_hip_matrix_options& o = *options;
    require(n >= 3, "KODAMAMatrix requires at least 3 rows.");
    require(p >= 1 && p <= kMaxDims, "KODAMAMatrix on the HIP backend supports 1..256 feature columns.");
    require(o.runs >= 1, "KODAMAMatrixOptions::runs must be positive.");
    require(o.cycles >= 0, "KODAMAMatrixOptions::cycles must be non-negative.");
    require(o.folds >= 2, "KODAMAMatrixOptions::folds must be at least two.");
    require(o.metric >= KODAMA_HIP_COSINE && o.metric <= KODAMA_HIP_EUCLIDEAN, "unknown metric.");
    require(o.k >= 1 && o.k <= kMaxK, "knn.k must be in 1..64 on the HIP backend.");
    require(o.form >= 0 && o.form <= 2, "form must be 0 (plain), 1 (device-resident) or 2 (learned).");
    require(o.form != 2 || o.mlp_params != nullptr, "the learned form needs network parameters.");
    require(o.form != 2 || (o.mlp_fallback >= 0.0f && o.mlp_fallback <= 1.0f), "fallback must be a probability.");
    const MatrixSizes z = resolve(n, o);
    require(z.neighbors <= (p <= 128 ? kMaxScanK : kMaxK),
            "graph neighbours above 128 (above 64 beyond 128 feature columns) are not supported by the HIP kNN kernel.");
    const int L = z.landmarks;
    select_device(o.device);
    Stream st(o.device);
    const hipStream_t s = st.get();
    auto clock = std::chrono::steady_clock::now();
    double t[6] = {0, 0, 0, 0, 0, 0};

    // Shared landmark atlas: k-means of every row with mt19937_64(seed + 1)
    // (kodama_matrix.cpp:3662-3726), its strata indexed by class.
    DeviceBuffer<float> d_x(static_cast<size_t>(n) * p);
    d_x.upload(x, d_x.size(), s);
    std::mt19937_64 shared(o.seed + 1u);
    const IVec atlas = device_kmeans(x, d_x.get(), n, p, z.coarse_k, shared, 10, s);
    IVec strata_off, strata_rows;
    index_strata(atlas, z.coarse_k, strata_off, strata_rows);
    t[0] = seconds_since(clock);

    // Global graph: exact self-kNN of the prepared rows, self excluded; or
    // the caller's graph, normalised.
    const int K = ext ? std::max(1, std::min(z.neighbors, ext->neighbors)) : z.neighbors;
    const size_t edges = static_cast<size_t>(n) * K;
    DeviceBuffer<float> d_gdist(edges);
    DeviceBuffer<int> d_gidx(edges);
    IVec ext_idx;
    std::vector<float> ext_dist;
    if (ext) {
      normalize_graph(*ext, n, K, ext_idx, ext_dist);
      d_gidx.upload(ext_idx.data(), edges, s);
      d_gdist.upload(ext_dist.data(), edges, s);
    } else {
      const std::vector<float> xs = prepare_rows(x, n, p, o.metric, 1);
      DeviceBuffer<float> d_xs(xs.size());
      DeviceBuffer<int> d_self(static_cast<size_t>(n));
      IVec self(static_cast<size_t>(n));
      std::iota(self.begin(), self.end(), 0);
      d_xs.upload(xs.data(), xs.size(), s);
      d_self.upload(self.data(), self.size(), s);
      device_exact_knn(d_xs.get(), n, d_xs.get(), n, p, z.neighbors, o.metric, d_self.get(), d_gidx.get(),
                       d_gdist.get(), max_row_norm2(xs.data(), n, p), s, &last_knn_fallbacks());
    }
    check(hipStreamSynchronize(s), "graph sync");
    t[1] = seconds_since(clock);

    // Per run r = 1..M: mt19937_64(seed + r) skips one draw (the atlas is
    // shared), quota-samples the landmarks and draws the k-means starting
    // labels of the landmark rows (kodama_matrix.cpp:2897, 2937-2949, 3088-3103).
    // The quota samples first (host threads, one stream per run), then the
    // starting-label k-means of all runs in lockstep on the device.
    IVec offsets(static_cast<size_t>(o.runs) + 1, 0), land_all(static_cast<size_t>(o.runs) * L), init_all;
    init_all.reserve(land_all.size());
    {
      std::vector<std::mt19937_64> rngs(static_cast<size_t>(o.runs));
#pragma omp parallel for schedule(dynamic)
      for (int r = 0; r < o.runs; ++r) {
        rngs[r].seed(o.seed + static_cast<std::uint64_t>(r + 1));
        (void)rngs[r]();
        const IVec land = quota_sample_landmarks(strata_off, strata_rows, L, rngs[r]);
        std::copy(land.begin(), land.end(), land_all.begin() + static_cast<size_t>(r) * L);
      }
      for (int r = 1; r <= o.runs; ++r) offsets[r] = offsets[r - 1] + L;
      DeviceBuffer<int> d_land(land_all.size());
      DeviceBuffer<float> d_xl(land_all.size() * static_cast<size_t>(p));
      d_land.upload(land_all.data(), land_all.size(), s);
      check(launch_gather_rows(d_x.get(), d_land.get(), static_cast<int>(land_all.size()), p, d_xl.get(), s),
            "gather_rows launch");
      const int init_k = std::max(2, std::min(z.splitting, L));
      const std::vector<IVec> xw = device_kmeans_batch(x, land_all.data(), d_xl.get(), o.runs, L, p, init_k, rngs, 10, s);
      for (const IVec& w : xw) init_all.insert(init_all.end(), w.begin(), w.end());
    }
    t[2] = seconds_since(clock);

    // Core of every run on its landmarks (run_iteration's CoreOptions,
    // kodama_matrix.cpp:3108-3139): run r (0-based) draws from seed + 1 + r and
    // folds from cv_seed + r, i.e. seed + run id for both.
    kodama_hip_core_options co{};
    co.cycles = o.cycles;
    co.classifier = 1;
    co.seed = o.seed + 1u;
    co.evolutionary_search = 1;
    co.guarded_diversity = 1;
    co.many_to_one_absorption = 1;
    co.adaptive_proposal_size = 1;
    co.prediction_guidance = co.policy_adaptive_proposal_size = co.transition_proposal = 1;
    co.stochastic_acceptance = co.diversity_multiplier = co.pls_transition_coarsening = 1;
    co.pls_fragmentation_penalty = co.reject_single_class = 1;
    co.folds = o.folds;
    co.stratified = 0;
    co.cv_seed = o.seed + 1u;
    co.k = o.k;
    co.metric = o.metric;
    co.device = o.device;
    co.n_threads = o.n_threads;
    const int total = offsets[o.runs];
    IVec clbest(static_cast<size_t>(total));
    std::vector<double> vect(static_cast<size_t>(o.runs) * std::max(1, o.cycles));
    double core_t[4] = {0, 0, 0, 0};
    const int rc =
      o.form == 2 ? kodama_hip_core_knn_batch_learned(x, n, p, o.runs, offsets.data(), land_all.data(),
                                                      init_all.data(), nullptr, &co, o.mlp_params, o.mlp_fallback,
                                                      clbest.data(), acc, vect.data(), core_t)
                  : kodama_hip_core_knn_batch(x, n, p, o.runs, offsets.data(), land_all.data(), init_all.data(),
                                              nullptr, &co, o.form, clbest.data(), acc, vect.data(), core_t);
    if (rc != KODAMA_HIP_OK) {
      const std::string msg = std::string("KODAMAMatrix core: ") + kodama_hip_last_error();
      if (rc == KODAMA_HIP_EINVAL) throw InvalidArgument(msg);
      throw DeviceError(msg);
    }
    t[3] = core_t[0];
    t[4] = core_t[1];
    seconds_since(clock);

    // Projection of the other rows through the global graph (2979-3000 of
    // run_iteration's kNN branch), all runs in one launch.
    DeviceBuffer<int> d_land(land_all.size()), d_off(offsets.size()), d_cl(clbest.size()),
      d_res(static_cast<size_t>(o.runs) * n);
    d_land.upload(land_all.data(), land_all.size(), s);
    d_off.upload(offsets.data(), offsets.size(), s);
    d_cl.upload(clbest.data(), clbest.size(), s);
    ProjectParams pp;
    pp.graph = d_gidx.get();
    pp.n = n;
    pp.neighbors = K;
    pp.landmarks = d_land.get();
    pp.run_offsets = d_off.get();
    pp.clbest = d_cl.get();
    pp.runs = o.runs;
    pp.votes = std::max(1, o.k);
    pp.max_landmarks = L;
    DeviceBuffer<int> d_local(static_cast<size_t>(o.runs) * n);
    pp.local = d_local.get();
    pp.res = d_res.get();
    check(launch_project(pp, s), "project launch");
    d_res.download(res, d_res.size(), s);
    if (graph_indices_one_based != nullptr && graph_distances != nullptr) {
      if (ext) {
        // The caller's distances are public already; ids become 1-based.
        for (size_t e = 0; e < edges; ++e) {
          graph_indices_one_based[e] = ext_idx[e] >= 0 ? ext_idx[e] + 1 : ext_idx[e];
          graph_distances[e] = ext_dist[e];
        }
      } else {
        check(launch_graph_publish(d_gidx.get(), d_gdist.get(), edges, o.metric, s), "graph_publish launch");
        d_gidx.download(graph_indices_one_based, edges, s);
        d_gdist.download(graph_distances, edges, s);
      }
    }
    check(hipStreamSynchronize(s), "projection sync");
    t[5] = seconds_since(clock);

    if (v != nullptr) {
      for (int r = 0; r < o.runs; ++r) {
        std::copy_n(vect.data() + static_cast<size_t>(r) * o.cycles, o.cycles, v + static_cast<size_t>(r) * o.cycles);
      }
    }
    if (landmark_rows != nullptr) std::copy(land_all.begin(), land_all.end(), landmark_rows);
    if (timings != nullptr) std::copy_n(t, 6, timings);
  }
}

}  // namespace

extern "C" {

int kodama_hip_kodama_matrix(const float* x, int32_t n, int32_t p, const kodama_hip_matrix_options* options,
                             int32_t* res, double* acc, double* v, int32_t* landmark_rows,
                             int32_t* graph_indices_one_based, float* graph_distances, double* timings) {
  return guarded([&] {
    matrix_impl(x, n, p, options, nullptr, res, acc, v, landmark_rows, graph_indices_one_based, graph_distances,
                timings);
  });
}

// KODAMAMatrixFromGraphData_CUDA (kodama_matrix.cpp:4676-4690).
int kodama_hip_kodama_matrix_from_graph(const float* x, int32_t n, int32_t p, const int32_t* graph_indices,
                                        const float* graph_distances_in, int32_t graph_neighbors,
                                        int32_t index_base, const kodama_hip_matrix_options* options, int32_t* res,
                                        double* acc, double* v, int32_t* landmark_rows,
                                        int32_t* graph_indices_one_based, float* graph_distances,
                                        double* timings) {
  return guarded([&] {
    require(n >= 2, "KODAMAMatrixFromGraphData requires at least two rows.");
    require(index_base >= -1 && index_base <= 1, "index_base must be 0, 1 or -1 (auto).");
    require(graph_indices != nullptr && graph_distances_in != nullptr, "null graph.");
    require(graph_neighbors > 0, "NeighborGraph.neighbors must be positive.");
    ExternalGraph g;
    g.indices = graph_indices;
    g.distances = graph_distances_in;
    g.neighbors = graph_neighbors;
    g.index_base = index_base;
    matrix_impl(x, n, p, options, &g, res, acc, v, landmark_rows, graph_indices_one_based, graph_distances,
                timings);
  });
}

}  // extern "C"
