// C ABI of libkodama_hip.so, orchestration half: KODAMA.matrix for the kNN
// and PLS-LDA classifiers (run_kodama_matrix, kodama_matrix.cpp:3324-4051, and
// run_iteration, 2870-3322) on one device, with the caller's starting labels,
// constraint groups and fixed rows, from the data or from a caller graph. The
// shared k-means atlas, the per-run k-means starting labels and the global
// graph run on the device; the landmark quotas draw on the host from the
// reference's streams; the M Core runs go through the batched engines (kNN:
// plain or device-resident form; PLS-LDA: plain); the other rows are projected
// through the global graph on the device (kNN) or by batched PLSLDAPredict.
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
#include "batch_api.h"
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

// Per-run tables of CoreKNNGraph (core.cpp:624-692) on the run's landmark
// subgraph (subset_graph_to_rows, kodama_matrix.cpp:1332-1366): the global
// neighbours of each landmark row that are landmarks of the run, as run-local
// ids in graph order (self dropped); the first k of them outside the row's
// fold vote with score -d^2 (Euclidean) or 1 - d. nbr / score [total x k].
void subgraph_tables(const IVec& gidx, const std::vector<float>& gdist, int n, int width, const IVec& offsets,
                     const IVec& land_all, const IVec& fold_all, int k, int metric, IVec& nbr,
                     std::vector<float>& score) {
  const int runs = static_cast<int>(offsets.size()) - 1;
  const size_t total = static_cast<size_t>(offsets[runs]);
  nbr.assign(total * k, -1);
  score.assign(total * k, -std::numeric_limits<float>::infinity());
#pragma omp parallel
  {
    IVec local(static_cast<size_t>(n), -1);
#pragma omp for schedule(dynamic, 1)
    for (int r = 0; r < runs; ++r) {
      const int b = offsets[r], L = offsets[r + 1] - offsets[r];
      for (int i = 0; i < L; ++i) local[land_all[b + i]] = i;
      for (int i = 0; i < L; ++i) {
        const size_t row = static_cast<size_t>(land_all[b + i]) * width;
        int col = 0;
        for (int j = 0; j < width && col < k; ++j) {
          const int g = gidx[row + j];
          if (g < 0 || g >= n) continue;
          const int nb = local[g];
          if (nb < 0 || nb == i || fold_all[b + nb] == fold_all[b + i]) continue;
          const float d = gdist[row + j];
          if (!(std::isfinite(d) && d >= 0.0f)) continue;
          nbr[static_cast<size_t>(b + i) * k + col] = nb;
          score[static_cast<size_t>(b + i) * k + col] = metric == KODAMA_HIP_EUCLIDEAN ? -(d * d) : 1.0f - d;
          ++col;
        }
      }
      for (int i = 0; i < L; ++i) local[land_all[b + i]] = -1;
    }
  }
}

void matrix_impl(const float* x, int32_t n, int32_t p, const kodama_hip_matrix_options* options,
                 const ExternalGraph* ext, const int32_t* starting_labels, const int32_t* constrain,
                 const int32_t* fixed, int32_t* res, double* acc, double* v, int32_t* landmark_rows,
                 int32_t* graph_indices_one_based, float* graph_distances, double* timings) {
  require(x != nullptr && options != nullptr && res != nullptr && acc != nullptr, "null pointer.");
  const kodama_hip_matrix_options& o = *options;
  require(n >= 3, "KODAMAMatrix requires at least 3 rows.");
  require(p >= 1 && p <= kMaxDims, "KODAMAMatrix on the HIP backend supports 1..256 feature columns.");
  require(o.runs >= 1, "KODAMAMatrixOptions::runs must be positive.");
  require(o.cycles >= 0, "KODAMAMatrixOptions::cycles must be non-negative.");
  require(o.folds >= 2, "KODAMAMatrixOptions::folds must be at least two.");
  require(o.classifier == 0 || o.classifier == 1, "Unsupported KODAMA.matrix classifier.");
  require(o.metric >= KODAMA_HIP_COSINE && o.metric <= KODAMA_HIP_EUCLIDEAN, "unknown metric.");
  const bool knn = o.classifier == 1;
  require(!knn || (o.k >= 1 && o.k <= kMaxK), "knn.k must be in 1..64 on the HIP backend.");
  require(o.form >= 0 && o.form <= 2, "form must be 0 (plain), 1 (device-resident) or 2 (learned).");
  require(o.form == 0 || knn, "the device-resident forms run the kNN classifier only.");
  require(o.form == 0 || starting_labels == nullptr,
          "starting labels turn the evolutionary search off, which only the plain form runs.");
  require(o.form != 2 || o.mlp_params != nullptr, "the learned form needs network parameters.");
  require(o.form != 2 || (o.mlp_fallback >= 0.0f && o.mlp_fallback <= 1.0f), "fallback must be a probability.");
  const MatrixSizes z = resolve(n, o);
  const bool want_graph = graph_indices_one_based != nullptr && graph_distances != nullptr;
  // The global graph serves the kNN classifier (the Core from a caller graph,
  // the projection) and the graph outputs.
  const bool build_graph = ext == nullptr && (knn || want_graph);
  require(!build_graph || z.neighbors <= (p <= 128 ? kMaxScanK : kMaxK),
          "graph neighbours above 128 (above 64 beyond 128 feature columns) are not supported by the HIP kNN kernel.");
  const int L = z.landmarks;
  const int comps = std::max(1, std::min(o.components, std::min(n, p)));
  const IVec cons_all = normalize_constrain(constrain, n);
  const bool identity = is_identity_constrain(cons_all);
  IVec fixed_all(static_cast<size_t>(n), 0);
  bool any_fixed = false;
  if (fixed) {
    for (int i = 0; i < n; ++i) {
      fixed_all[i] = fixed[i] != 0 ? 1 : 0;
      any_fixed = any_fixed || fixed_all[i] != 0;
    }
  }
  require(o.form == 0 || (identity && !any_fixed),
          "the device-resident forms support neither constraint groups nor fixed rows.");
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
  DeviceBuffer<float> d_gdist;
  DeviceBuffer<int> d_gidx;
  IVec ext_idx;
  std::vector<float> ext_dist;
  if (ext) {
    normalize_graph(*ext, n, K, ext_idx, ext_dist);
    if (knn) {
      d_gidx.resize(edges);
      d_gidx.upload(ext_idx.data(), edges, s);
    }
  } else if (build_graph) {
    d_gidx.resize(edges);
    d_gdist.resize(edges);
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
  // shared), quota-samples the landmarks and, without starting labels, draws
  // the k-means starting labels of the landmark rows (kodama_matrix.cpp:2897,
  // 2937-2949, 3071-3103). The quota samples first (host threads, one stream
  // per run), then the starting-label k-means of all runs in lockstep on the
  // device. Constraint groups of the landmarks (factor_subset) take their
  // majority starting label.
  IVec offsets(static_cast<size_t>(o.runs) + 1, 0), land_all(static_cast<size_t>(o.runs) * L),
    init_all(land_all.size()), cons_land, fixed_land;
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
    if (starting_labels) {
      for (size_t i = 0; i < land_all.size(); ++i) init_all[i] = starting_labels[land_all[i]];
    } else {
      DeviceBuffer<int> d_land(land_all.size());
      DeviceBuffer<float> d_xl(land_all.size() * static_cast<size_t>(p));
      d_land.upload(land_all.data(), land_all.size(), s);
      check(launch_gather_rows(d_x.get(), d_land.get(), static_cast<int>(land_all.size()), p, d_xl.get(), s),
            "gather_rows launch");
      const int init_k = std::max(2, std::min(z.splitting, L));
      const std::vector<IVec> xw =
        device_kmeans_batch(x, land_all.data(), d_xl.get(), o.runs, L, p, init_k, rngs, 10, s);
      for (int r = 0; r < o.runs; ++r) std::copy(xw[r].begin(), xw[r].end(), init_all.begin() + offsets[r]);
    }
    if (!identity) {
      cons_land.resize(land_all.size());
#pragma omp parallel for schedule(dynamic)
      for (int r = 0; r < o.runs; ++r) {
        const IVec land(land_all.begin() + offsets[r], land_all.begin() + offsets[r + 1]);
        const IVec g = factor_subset(cons_all, land);
        const IVec w = constrained_majority(IVec(init_all.begin() + offsets[r], init_all.begin() + offsets[r + 1]), g);
        std::copy(g.begin(), g.end(), cons_land.begin() + offsets[r]);
        std::copy(w.begin(), w.end(), init_all.begin() + offsets[r]);
      }
    }
    if (any_fixed) {
      fixed_land.resize(land_all.size());
      for (size_t i = 0; i < land_all.size(); ++i) fixed_land[i] = fixed_all[land_all[i]];
    }
  }
  t[2] = seconds_since(clock);

  // Core of every run on its landmarks (run_iteration's CoreOptions,
  // kodama_matrix.cpp:3108-3139): run r (0-based) draws from seed + 1 + r and
  // folds from cv_seed + r, i.e. seed + run id for both; greedy search with
  // starting labels; PLS-LDA with `components` fixed components and class
  // coarsening.
  kodama_hip_core_options co{};
  co.cycles = o.cycles;
  co.classifier = o.classifier;
  co.seed = o.seed + 1u;
  co.evolutionary_search = starting_labels == nullptr ? 1 : 0;
  co.guarded_diversity = 1;
  co.many_to_one_absorption = 1;
  co.auto_class_coarsening = knn ? 0 : 1;
  co.adaptive_proposal_size = 1;
  co.prediction_guidance = co.policy_adaptive_proposal_size = co.transition_proposal = 1;
  co.stochastic_acceptance = co.diversity_multiplier = co.pls_transition_coarsening = 1;
  co.pls_fragmentation_penalty = co.reject_single_class = 1;
  co.folds = o.folds;
  co.stratified = 0;
  co.cv_seed = o.seed + 1u;
  co.k = std::max(1, o.k);
  co.metric = o.metric;
  co.max_components = comps;
  co.fixed_components = comps;
  co.center = 1;
  co.scale = 1;
  co.pls_fold_cache = 1;
  co.device = o.device;
  co.n_threads = o.n_threads;
  const int total = offsets[o.runs];
  IVec clbest(static_cast<size_t>(total));
  std::vector<double> vect(static_cast<size_t>(o.runs) * std::max(1, o.cycles));
  double core_t[5] = {0, 0, 0, 0, 0};
  BatchRequest q{x, n, p, o.runs, offsets.data(), land_all.data(), init_all.data(), nullptr, &co};
  q.constrain = identity ? nullptr : cons_land.data();
  q.fixed = any_fixed ? fixed_land.data() : nullptr;
  IVec fold_all, tab_nbr;
  std::vector<float> tab_score;
  if (knn && ext) {
    // CoreKNNGraph on each run's landmark subgraph, with the folds
    // make_folds draws for the run.
    fold_all.resize(static_cast<size_t>(total));
#pragma omp parallel for schedule(dynamic)
    for (int r = 0; r < o.runs; ++r) {
      const IVec lab(init_all.begin() + offsets[r], init_all.begin() + offsets[r + 1]);
      const IVec g = identity ? IVec() : IVec(cons_land.begin() + offsets[r], cons_land.begin() + offsets[r + 1]);
      const IVec f = make_folds(lab, g, o.folds, false, co.cv_seed + static_cast<std::uint64_t>(r));
      std::copy(f.begin(), f.end(), fold_all.begin() + offsets[r]);
    }
    const int kt = std::max(1, std::min(co.k, K));
    subgraph_tables(ext_idx, ext_dist, n, K, offsets, land_all, fold_all, kt, o.metric, tab_nbr, tab_score);
    q.fold_ids = fold_all.data();
    q.table_nbr = tab_nbr.data();
    q.table_score = tab_score.data();
    q.table_k = kt;
  }
  if (knn) {
    run_knn_batch(q, o.form, o.form == 2 ? o.mlp_params : nullptr, o.mlp_fallback, nullptr, clbest.data(), acc,
                  vect.data(), core_t);
  } else {
    run_plslda_batch(q, clbest.data(), acc, vect.data(), core_t);
  }
  t[3] = core_t[0];
  t[4] = core_t[1];
  seconds_since(clock);

  // Projection of the other rows: through the global graph (kNN: the
  // majority of the first k landmarks among a row's neighbours, 2979-3000;
  // all runs in one launch), or PLSLDAPredict of a fit on the run's landmarks
  // (3262-3276).
  if (knn) {
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
    check(hipStreamSynchronize(s), "projection sync");
  } else {
    IVec test_off(static_cast<size_t>(o.runs) + 1, 0), test_rows;
    test_rows.reserve(static_cast<size_t>(o.runs) * (n - L));
    IVec mark(static_cast<size_t>(n));
    for (int r = 0; r < o.runs; ++r) {
      std::fill(mark.begin(), mark.end(), 0);
      int32_t* out = res + static_cast<size_t>(r) * n;
      for (int i = offsets[r]; i < offsets[r + 1]; ++i) {
        mark[land_all[i]] = 1;
        out[land_all[i]] = clbest[i];
      }
      for (int i = 0; i < n; ++i) if (!mark[i]) test_rows.push_back(i);
      test_off[r + 1] = static_cast<int>(test_rows.size());
    }
    IVec pred(test_rows.size());
    plslda_predict_batch(x, n, p, o.runs, offsets, land_all.data(), clbest.data(), test_off, test_rows.data(), co,
                         pred.data());
    for (int r = 0; r < o.runs; ++r) {
      for (int i = test_off[r]; i < test_off[r + 1]; ++i) res[static_cast<size_t>(r) * n + test_rows[i]] = pred[i];
    }
  }
  if (!identity) {
#pragma omp parallel for schedule(dynamic)
    for (int r = 0; r < o.runs; ++r) {
      int32_t* out = res + static_cast<size_t>(r) * n;
      const IVec cm = constrained_majority(IVec(out, out + n), cons_all);
      std::copy(cm.begin(), cm.end(), out);
    }
  }
  if (want_graph) {
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
      check(hipStreamSynchronize(s), "graph download sync");
    }
  }
  t[5] = seconds_since(clock);

  if (v != nullptr) {
    for (int r = 0; r < o.runs; ++r) {
      std::copy_n(vect.data() + static_cast<size_t>(r) * o.cycles, o.cycles, v + static_cast<size_t>(r) * o.cycles);
    }
  }
  if (landmark_rows != nullptr) std::copy(land_all.begin(), land_all.end(), landmark_rows);
  if (timings != nullptr) std::copy_n(t, 6, timings);
}

}  // namespace

extern "C" {

int kodama_hip_kodama_matrix(const float* x, int32_t n, int32_t p, const kodama_hip_matrix_options* options,
                             const int32_t* starting_labels, const int32_t* constrain, const int32_t* fixed,
                             int32_t* res, double* acc, double* v, int32_t* landmark_rows,
                             int32_t* graph_indices_one_based, float* graph_distances, double* timings) {
  return guarded([&] {
    matrix_impl(x, n, p, options, nullptr, starting_labels, constrain, fixed, res, acc, v, landmark_rows,
                graph_indices_one_based, graph_distances, timings);
  });
}

// KODAMAMatrixFromGraphData_CUDA (kodama_matrix.cpp:4676-4690).
int kodama_hip_kodama_matrix_from_graph(const float* x, int32_t n, int32_t p, const int32_t* graph_indices,
                                        const float* graph_distances_in, int32_t graph_neighbors,
                                        int32_t index_base, const kodama_hip_matrix_options* options,
                                        const int32_t* starting_labels, const int32_t* constrain,
                                        const int32_t* fixed, int32_t* res, double* acc, double* v,
                                        int32_t* landmark_rows, int32_t* graph_indices_one_based,
                                        float* graph_distances, double* timings) {
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
    matrix_impl(x, n, p, options, &g, starting_labels, constrain, fixed, res, acc, v, landmark_rows,
                graph_indices_one_based, graph_distances, timings);
  });
}

}  // extern "C"
