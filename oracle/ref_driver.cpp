// C-ABI driver over the reference's own vendored CPU core (kodama-cpp), used by
// the test suite to write golden vectors and by bench.py as the timed CPU
// baseline. TEST INFRASTRUCTURE ONLY: it is compiled together with the
// reference sources into oracle/_ref/libkodama_ref.so by oracle/Makefile.
//
// Every function converts flat arrays to the reference API types
// (kodama.hpp:114-143 MatrixView, 294-349 CoreOptions/CoreResult) and calls the
// public *_CPU entry point named in its comment. The kNN classifiers are run with
// hnsw_ef_search = 4096 and n_threads = 1 by default, which makes the reference
// HNSW search exhaustive (ef >= training rows) and its build deterministic.

#include "kodama/kodama.hpp"
#include "common.hpp"
#include "native_knn.hpp"

#include "kodama_oracle.h"

#include <algorithm>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  } catch (...) {
    g_error = "unknown exception";
    return 1;
  }
}

kodama::DistanceMetric metric_of(int m) {
  if (m == 0) return kodama::DistanceMetric::Cosine;
  if (m == 1) return kodama::DistanceMetric::InnerProduct;
  return kodama::DistanceMetric::Euclidean;
}

std::vector<int> vec_or_empty(const int32_t* p, int n) {
  if (p == nullptr) return {};
  return std::vector<int>(p, p + n);
}

kodama::CoreOptions core_options(const oracle_core_options* o) {
  kodama::CoreOptions c;
  c.cycles = o->cycles;
  c.classifier = o->classifier == 1 ? kodama::CoreClassifier::KNN : kodama::CoreClassifier::PLS_LDA;
  c.seed = o->seed;
  c.shake = o->shake != 0;
  c.evolutionary_search = o->evolutionary_search != 0;
  c.guarded_diversity = o->guarded_diversity != 0;
  c.many_to_one_absorption = o->many_to_one_absorption != 0;
  c.auto_class_coarsening = o->auto_class_coarsening != 0;
  c.adaptive_proposal_size = o->adaptive_proposal_size != 0;
  c.evolution.prediction_guidance = o->prediction_guidance != 0;
  c.evolution.adaptive_proposal_size = o->policy_adaptive_proposal_size != 0;
  c.evolution.transition_proposal = o->transition_proposal != 0;
  c.evolution.stochastic_acceptance = o->stochastic_acceptance != 0;
  c.evolution.diversity_multiplier = o->diversity_multiplier != 0;
  c.evolution.pls_transition_coarsening = o->pls_transition_coarsening != 0;
  c.evolution.pls_fragmentation_penalty = o->pls_fragmentation_penalty != 0;
  c.evolution.reject_single_class = o->reject_single_class != 0;
  c.knn.k = o->k;
  c.knn.metric = metric_of(o->metric);
  c.knn.backend = kodama::Backend::CPU;
  c.knn.cv.folds = o->folds;
  c.knn.cv.stratified = o->stratified != 0;
  c.knn.cv.seed = o->cv_seed;
  c.knn.hnsw_ef_search = o->hnsw_ef_search > 0 ? o->hnsw_ef_search : 4096;
  c.knn.n_threads = o->n_threads > 0 ? o->n_threads : 1;
  c.pls.cv.folds = o->folds;
  c.pls.cv.stratified = o->stratified != 0;
  c.pls.cv.seed = o->cv_seed;
  c.pls.max_components = o->max_components;
  c.pls.fixed_components = o->fixed_components;
  c.pls.center = o->center != 0;
  c.pls.scale = o->scale != 0;
  c.pls.n_threads = o->n_threads > 0 ? o->n_threads : 1;
  c.pls.backend = kodama::Backend::CPU;
  return c;
}

void export_core(const kodama::CoreResult& r, int n, int cycles, oracle_core_result* out) {
  std::memcpy(out->clbest, r.clbest.data(), sizeof(int32_t) * n);
  std::memcpy(out->clbest_dirty, r.clbest_dirty.data(), sizeof(int32_t) * n);
  std::memcpy(out->cvpredbest, r.cvpredbest.data(), sizeof(int32_t) * n);
  if (out->fold_assignments != nullptr && r.fold_assignments.size() == static_cast<size_t>(n)) {
    std::memcpy(out->fold_assignments, r.fold_assignments.data(), sizeof(int32_t) * n);
  }
  for (int c = 0; c < cycles; ++c) {
    out->vect_acc[c] = r.vect_acc[c];
    out->vect_score[c] = r.vect_score[c];
    out->proposal_size[c] = r.proposal_size[c];
    out->active_classes[c] = r.active_classes[c];
    out->accepted[c] = r.accepted[c];
    out->improving_acceptance[c] = r.improving_acceptance[c];
    out->temperature_acceptance[c] = r.temperature_acceptance[c];
  }
  const int counters[ORACLE_N_COUNTERS] = {
    r.cycles_completed, r.proposals_evaluated, r.best_state_updates,
    r.current_state_accepts, r.stochastic_state_attempts, r.stochastic_state_accepts,
    r.current_state_rejections, r.coarsening_moves, r.absorption_moves,
    r.transition_attempted, r.transition_accepted, r.many_to_one_attempted,
    r.many_to_one_accepted, r.pls_coarsening_attempted, r.pls_coarsening_accepted,
    r.cv_evaluations, r.success ? 1 : 0};
  std::memcpy(out->counters, counters, sizeof(counters));
  out->accbest = r.accbest;
  out->scorebest = r.scorebest;
}

}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_error.c_str(); }

// kodama::detail::make_folds (common.cpp:108-214)
int ref_make_folds(const int32_t* labels, int32_t n, const int32_t* constrain,
                   int32_t folds, int32_t stratified, uint64_t seed, int32_t* out) {
  return guarded([&] {
    kodama::FoldOptions f;
    f.folds = folds;
    f.stratified = stratified != 0;
    f.seed = seed;
    const std::vector<int> r = kodama::detail::make_folds(
      std::vector<int>(labels, labels + n), vec_or_empty(constrain, n), f);
    std::memcpy(out, r.data(), sizeof(int32_t) * n);
  });
}

// kodama::detail::native_hnsw_search (native_knn.cpp:1097-1151), single thread.
int ref_hnsw_search(const float* train, int32_t nt, const float* query, int32_t nq,
                    int32_t p, int32_t k, int32_t metric, int32_t ef_search,
                    int32_t* out_idx, float* out_dist) {
  return guarded([&] {
    kodama::detail::NativeHNSWParameters hp{32, 200, ef_search};
    const kodama::detail::NativeKNNResult r = kodama::detail::native_hnsw_search(
      std::vector<float>(train, train + static_cast<size_t>(nt) * p), nt,
      std::vector<float>(query, query + static_cast<size_t>(nq) * p), nq, p, k,
      metric_of(metric), hp, 1);
    const size_t items = static_cast<size_t>(nq) * r.neighbors;
    for (size_t i = 0; i < static_cast<size_t>(nq) * k; ++i) {
      out_idx[i] = -1;
      out_dist[i] = 0.0f;
    }
    for (int q = 0; q < nq; ++q) {
      for (int j = 0; j < r.neighbors; ++j) {
        out_idx[static_cast<size_t>(q) * k + j] = r.indices[static_cast<size_t>(q) * r.neighbors + j];
        out_dist[static_cast<size_t>(q) * k + j] = r.distances[static_cast<size_t>(q) * r.neighbors + j];
      }
    }
    (void)items;
  });
}

// kodama::CoreKNN_CPU (core.cpp:1648-1669)
int ref_core_knn(const float* x, int32_t n, int32_t p, const int32_t* labels,
                 const int32_t* constrain, const int32_t* fixed,
                 const oracle_core_options* opt, oracle_core_result* out) {
  return guarded([&] {
    const kodama::MatrixView xv(x, static_cast<size_t>(n), static_cast<size_t>(p));
    const kodama::CoreResult r = kodama::CoreKNN_CPU(
      xv, std::vector<int>(labels, labels + n), vec_or_empty(constrain, n),
      vec_or_empty(fixed, n), core_options(opt));
    export_core(r, n, opt->cycles, out);
  });
}

// kodama::CoreKNNGraph_CPU (core.cpp:1671-1682); graph indices are 0-based.
int ref_core_knn_graph(const int32_t* idx, const float* dist, int32_t samples,
                       int32_t neighbors, const int32_t* labels, const int32_t* constrain,
                       const int32_t* fixed, const oracle_core_options* opt,
                       oracle_core_result* out) {
  return guarded([&] {
    kodama::NeighborGraph g;
    g.neighbors = neighbors;
    g.index_base = kodama::GraphIndexBase::Zero;
    g.indices.assign(idx, idx + static_cast<size_t>(samples) * neighbors);
    g.distances.assign(dist, dist + static_cast<size_t>(samples) * neighbors);
    const kodama::CoreResult r = kodama::CoreKNNGraph_CPU(
      g, samples, std::vector<int>(labels, labels + samples),
      vec_or_empty(constrain, samples), vec_or_empty(fixed, samples), core_options(opt));
    export_core(r, samples, opt->cycles, out);
  });
}

// kodama::KNNCV_CPU (knncv.cpp:366-419)
int ref_knncv(const float* x, int32_t n, int32_t p, const int32_t* labels,
              const int32_t* constrain, const oracle_core_options* opt,
              int32_t* pred, int32_t* folds_out) {
  return guarded([&] {
    const kodama::MatrixView xv(x, static_cast<size_t>(n), static_cast<size_t>(p));
    kodama::KNNOptions k = core_options(opt).knn;
    const kodama::KNNCVResult r = kodama::KNNCV_CPU(
      xv, std::vector<int>(labels, labels + n), vec_or_empty(constrain, n), k);
    std::memcpy(pred, r.predicted_labels.data(), sizeof(int32_t) * n);
    std::memcpy(folds_out, r.fold_assignments.data(), sizeof(int32_t) * n);
  });
}

// kodama::PLSLDACV_CPU (plscv.cpp:2930-2932)
int ref_plsldacv(const float* x, int32_t n, int32_t p, const int32_t* labels,
                 const int32_t* constrain, const oracle_core_options* opt,
                 int32_t* pred, int32_t* folds_out, int32_t* selected_components) {
  return guarded([&] {
    const kodama::MatrixView xv(x, static_cast<size_t>(n), static_cast<size_t>(p));
    kodama::PLSOptions po = core_options(opt).pls;
    po.data_epoch = 0;
    const kodama::PLSCVResult r = kodama::PLSLDACV_CPU(
      xv, std::vector<int>(labels, labels + n), vec_or_empty(constrain, n), po);
    std::memcpy(pred, r.predicted_labels.data(), sizeof(int32_t) * n);
    std::memcpy(folds_out, r.fold_assignments.data(), sizeof(int32_t) * n);
    *selected_components = r.selected_components;
  });
}

// kodama::PLSDACV_CPU (plscv.cpp:2926-2928)
int ref_plsdacv(const float* x, int32_t n, int32_t p, const int32_t* labels,
                const int32_t* constrain, const oracle_core_options* opt,
                int32_t* pred, int32_t* folds_out, int32_t* selected_components) {
  return guarded([&] {
    const kodama::MatrixView xv(x, static_cast<size_t>(n), static_cast<size_t>(p));
    kodama::PLSOptions po = core_options(opt).pls;
    po.data_epoch = 0;
    const kodama::PLSCVResult r = kodama::PLSDACV_CPU(
      xv, std::vector<int>(labels, labels + n), vec_or_empty(constrain, n), po);
    std::memcpy(pred, r.predicted_labels.data(), sizeof(int32_t) * n);
    std::memcpy(folds_out, r.fold_assignments.data(), sizeof(int32_t) * n);
    *selected_components = r.selected_components;
  });
}

// kodama::PLS_CPU (plscv.cpp:2899-2901); outputs widened to double.
int ref_pls(const float* x, int32_t n, int32_t p, const float* y, int32_t q,
            const oracle_core_options* opt, double* weights, double* response_loadings,
            double* scores, double* coefficients, double* fitted, double* x_center,
            double* x_scale, double* y_center, int32_t* components) {
  return guarded([&] {
    const kodama::MatrixView xv(x, static_cast<size_t>(n), static_cast<size_t>(p));
    const kodama::MatrixView yv(y, static_cast<size_t>(n), static_cast<size_t>(q));
    const kodama::PLSFitResult r = kodama::PLS_CPU(xv, yv, core_options(opt).pls);
    *components = r.components;
    std::copy(r.weights.begin(), r.weights.end(), weights);
    std::copy(r.response_loadings.begin(), r.response_loadings.end(), response_loadings);
    std::copy(r.scores.begin(), r.scores.end(), scores);
    std::copy(r.coefficients.begin(), r.coefficients.end(), coefficients);
    std::copy(r.fitted.begin(), r.fitted.end(), fitted);
    std::copy(r.x_center.begin(), r.x_center.end(), x_center);
    std::copy(r.x_scale.begin(), r.x_scale.end(), x_scale);
    std::copy(r.y_center.begin(), r.y_center.end(), y_center);
  });
}

// kodama::PLSLDAPredict_CPU (plscv.cpp:2976-3005)
int ref_plslda_predict(const float* train, int32_t n_train, const int32_t* labels, const float* test,
                       int32_t n_test, int32_t p, const oracle_core_options* opt, int32_t* pred) {
  return guarded([&] {
    const kodama::MatrixView tv(train, static_cast<size_t>(n_train), static_cast<size_t>(p));
    const kodama::MatrixView sv(test, static_cast<size_t>(n_test), static_cast<size_t>(p));
    const std::vector<int> out =
      kodama::PLSLDAPredict_CPU(tv, std::vector<int>(labels, labels + n_train), sv, core_options(opt).pls);
    std::copy(out.begin(), out.end(), pred);
  });
}

// kodama::CorePLSLDA_CPU (core.cpp:1626-1646)
int ref_core_plslda(const float* x, int32_t n, int32_t p, const int32_t* labels,
                    const int32_t* constrain, const int32_t* fixed,
                    const oracle_core_options* opt, oracle_core_result* out) {
  return guarded([&] {
    const kodama::MatrixView xv(x, static_cast<size_t>(n), static_cast<size_t>(p));
    const kodama::CoreResult r = kodama::CorePLSLDA_CPU(
      xv, std::vector<int>(labels, labels + n), vec_or_empty(constrain, n),
      vec_or_empty(fixed, n), core_options(opt));
    export_core(r, n, opt->cycles, out);
  });
}

// kodama::KODAMADissimilarityInPlace (kodama_matrix.cpp:4383-4463); 1-based graph.
int ref_dissimilarity(int32_t* indices, float* distances, int32_t samples,
                      int32_t neighbors, const int32_t* res, int32_t runs) {
  return guarded([&] {
    kodama::NeighborGraph g;
    g.neighbors = neighbors;
    g.index_base = kodama::GraphIndexBase::One;
    g.indices.assign(indices, indices + static_cast<size_t>(samples) * neighbors);
    g.distances.assign(distances, distances + static_cast<size_t>(samples) * neighbors);
    kodama::KODAMADissimilarityInPlace(
      g, std::vector<int>(res, res + static_cast<size_t>(runs) * samples), runs, samples,
      kodama::Backend::CPU, 1, 0);
    std::memcpy(indices, g.indices.data(), sizeof(int32_t) * g.indices.size());
    std::memcpy(distances, g.distances.data(), sizeof(float) * g.distances.size());
  });
}

// kodama::KODAMAKNNGraph_CPU (graph_cluster.cpp:774-788): HNSW, 1-based output.
int ref_knn_graph(const float* x, int32_t n, int32_t p, int32_t k, int32_t metric,
                  int32_t n_threads, int32_t* idx_one_based, float* dist) {
  return guarded([&] {
    const kodama::MatrixView xv(x, static_cast<size_t>(n), static_cast<size_t>(p));
    kodama::GraphClusterOptions o;
    o.k = k;
    o.metric = metric_of(metric);
    o.n_threads = n_threads;
    const kodama::NeighborGraph g = kodama::KODAMAKNNGraph_CPU(xv, o);
    for (size_t i = 0; i < static_cast<size_t>(n) * k; ++i) {
      idx_one_based[i] = -1;
      dist[i] = 0.0f;
    }
    const int w = g.neighbors;
    const bool one = g.index_base == kodama::GraphIndexBase::One;
    for (int r = 0; r < n; ++r) {
      for (int j = 0; j < w && j < k; ++j) {
        const int id = g.indices[static_cast<size_t>(r) * w + j];
        idx_one_based[static_cast<size_t>(r) * k + j] = id < 0 ? -1 : (one ? id : id + 1);
        dist[static_cast<size_t>(r) * k + j] = g.distances[static_cast<size_t>(r) * w + j];
      }
    }
  });
}

// kodama::KODAMAMatrix_CPU / KODAMAMatrixFromGraphData_CPU
// (kodama_matrix.cpp:4465-4475, 4656-4670): one thread, no PCA visual
// initialisation, no graph dissimilarity; the run diagnostics' hashes of the
// landmark rows, starting labels and fold ids go to diag [runs x 3].
int ref_kodama_matrix(const float* x, int32_t n, int32_t p, const oracle_matrix_options* o,
                      const int32_t* graph_indices, const float* graph_distances, int32_t graph_neighbors,
                      int32_t index_base, const int32_t* starting_labels, const int32_t* constrain,
                      const int32_t* fixed, int32_t* res, double* acc, double* v, uint64_t* diag) {
  return guarded([&] {
    const kodama::MatrixView xv(x, static_cast<size_t>(n), static_cast<size_t>(p));
    kodama::KODAMAMatrixOptions m;
    m.runs = o->runs;
    m.cycles = o->cycles;
    m.components = o->components;
    m.landmarks = o->landmarks;
    m.splitting = o->splitting;
    m.graph_neighbors = o->graph_neighbors;
    m.folds = o->folds;
    m.n_threads = 1;
    m.seed = o->seed;
    m.metric = metric_of(o->metric);
    m.backend = kodama::Backend::CPU;
    m.classifier = o->classifier == 1 ? kodama::CoreClassifier::KNN : kodama::CoreClassifier::PLS_LDA;
    m.apply_kodama_dissimilarity = false;
    m.compute_visual_init = false;
    m.materialize_graph = false;
    m.knn.k = o->k;
    m.knn.n_threads = 1;
    const std::vector<int> w = vec_or_empty(starting_labels, n), c = vec_or_empty(constrain, n),
                           f = vec_or_empty(fixed, n);
    kodama::KODAMAMatrixResult r;
    if (graph_indices != nullptr) {
      kodama::NeighborGraph g;
      g.neighbors = graph_neighbors;
      g.index_base = index_base == 0 ? kodama::GraphIndexBase::Zero
                   : index_base == 1 ? kodama::GraphIndexBase::One : kodama::GraphIndexBase::Auto;
      g.indices.assign(graph_indices, graph_indices + static_cast<size_t>(n) * graph_neighbors);
      g.distances.assign(graph_distances, graph_distances + static_cast<size_t>(n) * graph_neighbors);
      r = kodama::KODAMAMatrixFromGraphData_CPU(xv, g, w, c, f, m);
    } else {
      r = kodama::KODAMAMatrix_CPU(xv, w, c, f, m);
    }
    std::memcpy(res, r.res.data(), sizeof(int32_t) * r.res.size());
    std::memcpy(acc, r.acc.data(), sizeof(double) * r.acc.size());
    if (v) std::memcpy(v, r.v.data(), sizeof(double) * r.v.size());
    if (diag) {
      for (int i = 0; i < o->runs; ++i) {
        diag[static_cast<size_t>(i) * 3 + 0] = r.run_diagnostics[i].landmark_rows_hash;
        diag[static_cast<size_t>(i) * 3 + 1] = r.run_diagnostics[i].initial_labels_hash;
        diag[static_cast<size_t>(i) * 3 + 2] = r.run_diagnostics[i].fold_assignments_hash;
      }
    }
  });
}

}  // extern "C"
