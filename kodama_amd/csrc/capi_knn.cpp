// C ABI of libkodama_hip.so, kNN half: the exact-kNN and vote-graph device seam,
// KNNCV, CoreKNN, CoreKNNGraph and the batched M-run evolution.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <numeric>
#include <string>

#include "../../include/kodama_hip.h"
#include "capi_common.h"
#include "engine.h"
#include "batch_api.h"
#include "evolve_device.h"
#include "evolve_launch.h"
#include "knn_mfma.h"
#include "launch.h"
#include "mlp.h"

using namespace kodama_hip;

namespace {

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void require(bool ok, const char* msg) {
  if (!ok) throw InvalidArgument(msg);
}

void validate_matrix(const float* x, int n, int p, const int* labels) {
  require(x != nullptr, "MatrixView data pointer is null.");
  require(n > 0 && p > 0, "Input matrix must have at least one row and one column.");
  require(labels != nullptr, "labels size must match number of rows.");
}

IVec vec_or_empty(const int* p, int n) { return p ? IVec(p, p + n) : IVec(); }

}  // namespace

struct kodama_hip_vote_graph {
  std::unique_ptr<KnnBatch> batch;
  int samples = 0;
};

namespace kodama_hip {

thread_local int g_last_evolution_workgroups = 0;
thread_local int g_last_evolution_lds_flows = 0;

void run_knn_batch(const BatchRequest& q, int mode, const float* mlp_params, float fallback, TraceSink* trace,
                   int32_t* clbest, double* accbest, double* vect_acc, double* timings) {
  const float* x = q.x;
  const int n = q.n, p = q.p, runs = q.runs;
  const kodama_hip_core_options* options = q.options;
  const bool tables = q.table_nbr != nullptr;
  require(tables || (x != nullptr && n > 0 && p > 0), "invalid data matrix.");
  require(runs >= 1, "runs must be positive.");
  require(q.run_offsets != nullptr && q.initial_labels != nullptr, "null input.");
  require(tables || q.landmark_rows != nullptr, "null input.");
  require(!tables || (q.table_score != nullptr && q.fold_ids != nullptr && q.table_k >= 1),
          "caller tables need scores, fold ids and a width.");
  require(options != nullptr && clbest != nullptr && accbest != nullptr, "null options or output.");
  require(options->cycles >= 0, "Core cycles must be non-negative.");
  require(mode != 2 || mlp_params != nullptr, "the learned form needs network parameters.");
  require(mode == 0 || (q.constrain == nullptr && q.fixed == nullptr),
          "the device-resident evolution supports neither constraint groups nor fixed rows.");
  select_device(options->device);
  const auto t0 = std::chrono::steady_clock::now();
  const IVec offsets(q.run_offsets, q.run_offsets + runs + 1);
  require(offsets[0] == 0, "run_offsets[0] must be 0.");
  for (int r = 0; r < runs; ++r) require(offsets[r + 1] > offsets[r] + 1, "every run needs at least two rows.");
  const int total = offsets[runs];
  IVec rows;
  if (!tables) {
    rows.assign(q.landmark_rows, q.landmark_rows + total);
    for (int v : rows) require(v >= 0 && v < n, "landmark row outside the data matrix.");
  }
  std::vector<IVec> folds(static_cast<size_t>(runs)), values(static_cast<size_t>(runs)),
    codes(static_cast<size_t>(runs)), cons(static_cast<size_t>(runs)), fixed(static_cast<size_t>(runs));
  IVec fold_concat(static_cast<size_t>(total));
  int classes = 1;
  for (int r = 0; r < runs; ++r) {
    const int L = offsets[r + 1] - offsets[r];
    const int* lab = q.initial_labels + offsets[r];
    if (q.constrain) cons[r].assign(q.constrain + offsets[r], q.constrain + offsets[r + 1]);
    if (q.fixed) fixed[r].assign(q.fixed + offsets[r], q.fixed + offsets[r + 1]);
    folds[r] = q.fold_ids ? IVec(q.fold_ids + offsets[r], q.fold_ids + offsets[r + 1])
                          : make_folds(IVec(lab, lab + L), cons[r], options->folds, options->stratified != 0,
                                       options->cv_seed + static_cast<std::uint64_t>(r));
    {
      // Two folds or more: every row then has training rows, so the exact
      // search gives it at least one neighbour.
      const auto mm = std::minmax_element(folds[r].begin(), folds[r].end());
      require(*mm.first != *mm.second, "every run needs at least two distinct folds.");
    }
    encode_labels(lab, L, values[r], codes[r]);
    classes = std::max(classes, static_cast<int>(values[r].size()));
    std::copy(folds[r].begin(), folds[r].end(), fold_concat.begin() + offsets[r]);
  }
  std::unique_ptr<KnnBatch> owned;
  if (tables) {
    owned = std::make_unique<KnnBatch>(options->device, q.table_nbr, q.table_score, offsets, q.table_k, classes);
  } else {
    const std::vector<float> xs = prepare_rows(x, n, p, options->metric, 0);
    owned = std::make_unique<KnnBatch>(options->device, xs.data(), n, p, offsets, rows, fold_concat, options->k,
                                       options->metric, classes);
  }
  KnnBatch& batch = *owned;
  // The device-resident forms' label-independent set-up (state arena, 16-bit
  // table, cleared flow table, packed network) belongs to the precompute.
  std::unique_ptr<DeviceEvolution> dev;
  if (mode != 0) dev = std::make_unique<DeviceEvolution>(batch, values, *options, mode == 2 ? mlp_params : nullptr);
  const double t_pre = seconds_since(t0);
  const auto t1 = std::chrono::steady_clock::now();
  double vote_seconds = 0.0;
  if (mode == 0) {
    std::vector<RunEvolution> ev;
    ev.reserve(static_cast<size_t>(runs));
    for (int r = 0; r < runs; ++r) {
      CoreFlags f = CoreFlags::from(*options);
      f.knn = true;
      f.seed = options->seed + static_cast<std::uint64_t>(r);
      ev.emplace_back(values[r], codes[r], cons[r], fixed[r], f);
    }
    if (trace) evolve_plain_traced(ev, batch, folds, options->cycles, options->n_threads, *trace);
    else evolve_plain(ev, batch, folds, options->cycles, options->n_threads);
    vote_seconds = batch.vote_ms_total() * 1e-3 / std::max(1, batch.vote_calls());
    for (int r = 0; r < runs; ++r) {
      const IVec& b = ev[r].best_codes();
      for (size_t i = 0; i < b.size(); ++i) clbest[offsets[r] + i] = values[r][b[i]];
      accbest[r] = ev[r].accbest();
      if (vect_acc) std::copy(ev[r].vect_acc().begin(), ev[r].vect_acc().end(), vect_acc + static_cast<size_t>(r) * options->cycles);
    }
  }
  double t_evo = 0.0;
  if (mode == 0) {
    t_evo = seconds_since(t1);
  } else {
    // The evolution proper runs with its inputs in device memory: the
    // starting labels are uploaded before its clock starts and the results
    // downloaded after it stops; the wall time with both copies is reported
    // apart.
    dev->start(codes);
    const auto tr = std::chrono::steady_clock::now();
    if (mode == 2) dev->run_learned(options->cycles, fallback);
    else dev->run(options->cycles);
    dev->wait();
    t_evo = seconds_since(tr);
    g_last_evolution_workgroups = dev->workgroups();
    g_last_evolution_lds_flows = dev->lds_flow_table() ? 1 : 0;
    dev->export_results(clbest, accbest, vect_acc);
  }
  const double t_wall = seconds_since(t1);
  if (dev) vote_seconds = dev->kernel_seconds();
  if (timings) {
    timings[0] = t_pre;
    timings[1] = t_evo;
    timings[2] = static_cast<double>(runs) * options->cycles;
    timings[3] = vote_seconds;
    timings[4] = t_wall;
  }
}

}  // namespace kodama_hip

extern "C" {

int kodama_hip_exact_knn_search(const float* train, int32_t train_rows, const float* query, int32_t query_rows,
                                int32_t dimensions, int32_t k, int32_t metric, int32_t device,
                                const int32_t* query_train_indices, int32_t* out_indices, float* out_distances) {
  return guarded([&] {
    require(train != nullptr && query != nullptr, "null matrix.");
    require(train_rows >= 1 && query_rows >= 0 && dimensions >= 1 && k >= 1,
            "invalid exact kNN dimensions or neighbor count.");
    require(k <= (dimensions <= 128 ? kMaxScanK : kMaxK),
            "k above 128 (above 64 beyond 128 feature columns) is not supported by the HIP kNN kernel.");
    require(dimensions <= kMaxDims, "the HIP kNN kernel supports 1..256 feature columns.");
    select_device(device);
    Stream st(device);
    const hipStream_t s = st.get();
    DeviceBuffer<float> d_train(static_cast<size_t>(train_rows) * dimensions);
    DeviceBuffer<float> d_query(static_cast<size_t>(query_rows) * dimensions);
    DeviceBuffer<int> d_ex(query_train_indices ? static_cast<size_t>(query_rows) : 0);
    DeviceBuffer<int> d_idx(static_cast<size_t>(query_rows) * k);
    DeviceBuffer<float> d_dist(static_cast<size_t>(query_rows) * k);
    d_train.upload(train, d_train.size(), s);
    d_query.upload(query, d_query.size(), s);
    if (query_train_indices) d_ex.upload(query_train_indices, d_ex.size(), s);
    device_exact_knn(d_train.get(), train_rows, d_query.get(), query_rows, dimensions, k, metric,
                     query_train_indices ? d_ex.get() : nullptr, d_idx.get(), d_dist.get(),
                     max_row_norm2(train, train_rows, dimensions), s, &last_knn_fallbacks());
    d_idx.download(out_indices, d_idx.size(), s);
    d_dist.download(out_distances, d_dist.size(), s);
    st.sync();
  });
}

int kodama_hip_build_knn_vote_graph(const int32_t* neighbor_rows, const float* scores, int32_t samples,
                                    int32_t neighbors, int32_t device, kodama_hip_vote_graph** out_graph) {
  return guarded([&] {
    require(out_graph != nullptr, "null output handle.");
    require(neighbor_rows != nullptr && scores != nullptr, "null graph arrays.");
    require(samples >= 1 && neighbors >= 1, "invalid vote graph shape.");
    select_device(device);
    auto g = std::make_unique<kodama_hip_vote_graph>();
    g->batch = std::make_unique<KnnBatch>(device, neighbor_rows, scores, IVec{0, samples}, neighbors, 1);
    g->samples = samples;
    *out_graph = g.release();
  });
}

int kodama_hip_knn_vote_predict(kodama_hip_vote_graph* graph, const int32_t* label_codes, int32_t fallback_code,
                                int32_t* prediction_codes) {
  return guarded([&] {
    require(graph != nullptr && graph->batch != nullptr, "invalid vote graph.");
    require(label_codes != nullptr && prediction_codes != nullptr, "null label arrays.");
    KnnBatch& b = *graph->batch;
    std::copy(label_codes, label_codes + graph->samples, b.labels_host());
    b.vote_all(false);
    const int* pred = b.predictions_host();
    for (int i = 0; i < graph->samples; ++i) prediction_codes[i] = pred[i] >= 0 ? pred[i] : fallback_code;
  });
}

void kodama_hip_release_knn_vote_graph(kodama_hip_vote_graph* graph) { delete graph; }

// KNNCV_CUDA (knncv.cpp:422-488): every fold's exact neighbours come from one
// fold-masked launch; the vote is the same (count, score sum, label) rule; a
// row with no neighbour keeps the label of its fold's first training row.
int kodama_hip_knncv(const float* x, int32_t n, int32_t p, const int32_t* labels, const int32_t* constrain,
                     const kodama_hip_core_options* options, int32_t* predicted_labels,
                     int32_t* fold_assignments, double* global_accuracy) {
  return guarded([&] {
    validate_matrix(x, n, p, labels);
    require(options != nullptr, "null options.");
    require(options->k >= 1, "KNNOptions::k must be positive.");
    select_device(options->device);
    const IVec lab(labels, labels + n);
    const IVec folds = make_folds(lab, vec_or_empty(constrain, n), options->folds, options->stratified != 0,
                                  options->cv_seed);
    IVec values, codes;
    encode_labels(labels, n, values, codes);
    const std::vector<float> xs = prepare_rows(x, n, p, options->metric, 1);
    IVec rows(static_cast<size_t>(n));
    std::iota(rows.begin(), rows.end(), 0);
    for (int f : folds) require(f >= 0, "invalid fold id.");
    {
      std::map<int, int> fold_rows;
      for (int f : folds) ++fold_rows[f];
      for (const auto& kv : fold_rows) {
        if (kv.second == n) throw std::runtime_error("KNN fold has no training samples.");
      }
    }
    KnnBatch batch(options->device, xs.data(), n, p, IVec{0, n}, rows, folds, options->k, options->metric,
                   static_cast<int>(values.size()));
    std::copy(codes.begin(), codes.end(), batch.labels_host());
    batch.vote_all(false);
    const int* pred = batch.predictions_host();
    std::map<int, int> first_train;
    size_t ok = 0;
    for (int i = 0; i < n; ++i) {
      int lab_i;
      if (pred[i] >= 0) {
        lab_i = values[pred[i]];
      } else {
        auto it = first_train.find(folds[i]);
        if (it == first_train.end()) {
          int t = 0;
          while (t < n && folds[t] == folds[i]) ++t;
          it = first_train.emplace(folds[i], t).first;
        }
        lab_i = labels[it->second];
      }
      predicted_labels[i] = lab_i;
      ok += lab_i == labels[i] ? 1u : 0u;
    }
    if (fold_assignments) std::copy(folds.begin(), folds.end(), fold_assignments);
    if (global_accuracy) *global_accuracy = static_cast<double>(ok) / static_cast<double>(n);
  });
}

// CoreKNN_CUDA (core.cpp:1782-1812).
int kodama_hip_core_knn(const float* x, int32_t n, int32_t p, const int32_t* labels, const int32_t* constrain,
                        const int32_t* fixed, const kodama_hip_core_options* options,
                        kodama_hip_core_result* result) {
  return guarded([&] {
    validate_matrix(x, n, p, labels);
    require(options != nullptr && result != nullptr, "null options or result.");
    require(options->k >= 1, "KNNOptions::k must be positive.");
    require(options->cycles >= 0, "Core cycles must be non-negative.");
    select_device(options->device);
    CoreFlags f = CoreFlags::from(*options);
    f.knn = true;
    const IVec lab(labels, labels + n);
    const IVec cons = vec_or_empty(constrain, n);
    const IVec folds = make_folds(lab, cons, options->folds, options->stratified != 0, options->cv_seed);
    {
      std::map<int, int> fold_rows;
      for (int fo : folds) ++fold_rows[fo];
      for (const auto& kv : fold_rows) {
        if (kv.second == n) throw std::runtime_error("KNN fold has no training samples.");
      }
    }
    IVec values, codes;
    encode_labels(labels, n, values, codes);
    const std::vector<float> xs = prepare_rows(x, n, p, options->metric, 0);
    IVec rows(static_cast<size_t>(n));
    std::iota(rows.begin(), rows.end(), 0);
    KnnBatch batch(options->device, xs.data(), n, p, IVec{0, n}, rows, folds, options->k, options->metric,
                   static_cast<int>(values.size()));
    std::vector<RunEvolution> runs;
    runs.emplace_back(values, codes, cons, vec_or_empty(fixed, n), f);
    evolve_plain(runs, batch, std::vector<IVec>{folds}, options->cycles, 1);
    runs[0].export_result(result);
    if (result->fold_assignments) std::copy(folds.begin(), folds.end(), result->fold_assignments);
  });
}

// CoreKNNGraph_CUDA (core.cpp:1684-1700 -> precompute_knn_cv_graph 624-692):
// for each row, the first k graph neighbours outside its fold, score -d^2
// (Euclidean) or 1 - d.
int kodama_hip_core_knn_graph(const int32_t* graph_indices, const float* graph_distances, int32_t samples,
                              int32_t neighbors, int32_t index_base, const int32_t* labels,
                              const int32_t* constrain, const int32_t* fixed,
                              const kodama_hip_core_options* options, kodama_hip_core_result* result) {
  return guarded([&] {
    require(options != nullptr && result != nullptr, "null options or result.");
    require(samples >= 2, "CoreKNNGraph requires at least two samples.");
    require(neighbors > 0, "NeighborGraph.neighbors must be positive.");
    require(graph_indices != nullptr && graph_distances != nullptr && labels != nullptr, "null input.");
    require(options->k >= 1, "KNNOptions::k must be positive.");
    select_device(options->device);
    const size_t total = static_cast<size_t>(samples) * neighbors;
    // normalize_graph_indices (core.cpp:586-622).
    int lo = std::numeric_limits<int>::max(), hi = std::numeric_limits<int>::min();
    for (size_t i = 0; i < total; ++i) {
      if (graph_indices[i] < 0) continue;
      lo = std::min(lo, graph_indices[i]);
      hi = std::max(hi, graph_indices[i]);
    }
    const bool one_based = index_base == 2 || (index_base == 0 && lo >= 1 && hi <= samples);
    CoreFlags f = CoreFlags::from(*options);
    f.knn = true;
    const IVec lab(labels, labels + samples);
    const IVec cons = vec_or_empty(constrain, samples);
    const IVec folds = make_folds(lab, cons, options->folds, options->stratified != 0, options->cv_seed);
    const int k = std::max(1, std::min(options->k, neighbors));
    require(k <= kMaxK, "k above 64 is not supported by the HIP vote kernel.");
    std::vector<int> nbr(static_cast<size_t>(samples) * k, -1);
    std::vector<float> score(static_cast<size_t>(samples) * k, -std::numeric_limits<float>::infinity());
    for (int row = 0; row < samples; ++row) {
      int col = 0;
      for (int j = 0; j < neighbors && col < k; ++j) {
        int id = graph_indices[static_cast<size_t>(row) * neighbors + j];
        if (one_based && id > 0) --id;
        if (id < 0 || id >= samples || id == row) continue;
        if (folds[id] == folds[row]) continue;
        const float d0 = graph_distances[static_cast<size_t>(row) * neighbors + j];
        const float dist = std::isfinite(d0) && d0 >= 0.0f ? d0 : std::numeric_limits<float>::infinity();
        if (!std::isfinite(dist)) continue;
        const float d = std::max(0.0f, dist);
        nbr[static_cast<size_t>(row) * k + col] = id;
        score[static_cast<size_t>(row) * k + col] =
          options->metric == KODAMA_HIP_EUCLIDEAN ? -(d * d) : 1.0f - d;
        ++col;
      }
    }
    IVec values, codes;
    encode_labels(labels, samples, values, codes);
    KnnBatch batch(options->device, nbr.data(), score.data(), IVec{0, samples}, k, static_cast<int>(values.size()));
    std::vector<RunEvolution> runs;
    runs.emplace_back(values, codes, cons, vec_or_empty(fixed, samples), f);
    evolve_plain(runs, batch, std::vector<IVec>{folds}, options->cycles, 1);
    runs[0].export_result(result);
    if (result->fold_assignments) std::copy(folds.begin(), folds.end(), result->fold_assignments);
  });
}

int kodama_hip_core_knn_batch(const float* x, int32_t n, int32_t p, int32_t runs, const int32_t* run_offsets,
                              const int32_t* landmark_rows, const int32_t* initial_labels, const int32_t* fold_ids,
                              const kodama_hip_core_options* options, int32_t mode, int32_t* clbest,
                              double* accbest, double* vect_acc, double* timings) {
  return guarded([&] {
    require(mode == 0 || mode == 1, "mode must be 0 (plain) or 1 (device-resident).");
    BatchRequest q{x, n, p, runs, run_offsets, landmark_rows, initial_labels, fold_ids, options};
    run_knn_batch(q, mode, nullptr, 0.0f, nullptr, clbest, accbest, vect_acc, timings);
  });
}

// precompute_knn_cv_cuda (core.cpp:694-782) for a batch of runs.
int kodama_hip_knn_fold_tables(const float* x, int32_t n, int32_t p, int32_t runs, const int32_t* run_offsets,
                               const int32_t* rows, const int32_t* fold_ids, int32_t k, int32_t metric,
                               int32_t path, int32_t device, int32_t* neighbors, float* scores,
                               int32_t* path_used) {
  return guarded([&] {
    require(x != nullptr && n > 0 && p > 0, "invalid data matrix.");
    require(runs >= 1, "runs must be positive.");
    require(run_offsets != nullptr && rows != nullptr && fold_ids != nullptr, "null input.");
    require(neighbors != nullptr && scores != nullptr, "null output.");
    require(path >= 0 && path <= 2, "path must be 0 (tuned), 1 (scan) or 2 (MFMA).");
    select_device(device);
    const IVec offsets(run_offsets, run_offsets + runs + 1);
    require(offsets[0] == 0, "run_offsets[0] must be 0.");
    for (int r = 0; r < runs; ++r) require(offsets[r + 1] > offsets[r], "every run needs at least one row.");
    const int total = offsets[runs];
    const IVec rw(rows, rows + total), folds(fold_ids, fold_ids + total);
    for (int v : rw) require(v >= 0 && v < n, "run row outside the data matrix.");
    for (int f : folds) require(f >= 0, "invalid fold id.");
    const std::vector<float> xs = prepare_rows(x, n, p, metric, 0);
    KnnBatch batch(device, xs.data(), n, p, offsets, rw, folds, k, metric, 1, path);
    for (int r = 0; r < runs; ++r) {
      batch.download_tables(r, neighbors + static_cast<size_t>(offsets[r]) * k,
                            scores + static_cast<size_t>(offsets[r]) * k);
    }
    if (path_used) *path_used = last_fold_table_path();
  });
}

int kodama_hip_set_fold_table_path(int32_t path) {
  return guarded([&] {
    require(path >= 0 && path <= 2, "path must be 0 (tuned), 1 (scan) or 2 (MFMA).");
    set_fold_table_path(path);
  });
}

int kodama_hip_last_fold_table_path(void) { return last_fold_table_path(); }

int kodama_hip_set_fold_table_schedule(int32_t path, int32_t list, int32_t waves) {
  return guarded([&] {
    require(path >= 0 && path <= 2, "path must be 0 (tuned), 1 (scan) or 2 (MFMA).");
    require(list == 0 || list == 8 || list == 16, "list must be 0 (tuned), 8 or 16.");
    require(waves == 0 || waves == 4 || waves == 8, "waves must be 0 (tuned), 4 or 8.");
    set_fold_table_schedule(TableSchedule{path, list, waves});
  });
}

int kodama_hip_set_mfma_packed_lists(int32_t on) {
  return guarded([&] {
    require(on == 0 || on == 1, "on must be 0 or 1.");
    set_mfma_packed_lists(on);
  });
}

int kodama_hip_last_fold_table_schedule(int32_t* path, int32_t* list, int32_t* waves) {
  return guarded([&] {
    require(path != nullptr && list != nullptr && waves != nullptr, "null pointer.");
    const TableSchedule t = last_fold_table_schedule();
    *path = t.path;
    *list = t.list;
    *waves = t.waves;
  });
}

int kodama_hip_set_evolution_split(int32_t max_group) {
  return guarded([&] {
    require(max_group >= 0 && max_group <= kEvoMaxGroup, "max_group must be 0 (automatic) to 8.");
    set_evolution_split(max_group);
  });
}

int kodama_hip_last_evolution_workgroups(void) { return g_last_evolution_workgroups; }

int kodama_hip_set_cooperative_launch(int32_t cooperative) {
  return guarded([&] {
    require(cooperative == 0 || cooperative == 1, "cooperative must be 0 or 1.");
    set_cooperative_launch(cooperative != 0);
  });
}

int kodama_hip_set_lds_flow_table(int32_t enabled) {
  return guarded([&] {
    require(enabled == 0 || enabled == 1, "enabled must be 0 or 1.");
    set_lds_flow_table(enabled != 0);
  });
}

int kodama_hip_last_evolution_lds_flows(void) { return g_last_evolution_lds_flows; }

}  // extern "C"
