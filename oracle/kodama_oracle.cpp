// CPU oracle for the KODAMA label-evolution hot path -- TEST INFRASTRUCTURE ONLY.
//
// A C++ restatement of the reference algorithms, used as the checker of the HIP
// path. It is C++ rather than C because exact trajectory replay depends on the
// libstdc++ random streams the reference draws from (std::mt19937_64,
// std::uniform_int_distribution, std::shuffle, std::discrete_distribution,
// std::uniform_real_distribution, std::normal_distribution); calling the same
// library with the same sequence of requests reproduces them.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
// library. The product path never does.
//
// Pinning (what checks each function; tests/test_oracle.py and
// tests/test_oracle_matrix.py, fixtures in tests/golden/ written by the
// compiled reference through oracle/_ref):
//   oracle_make_folds        reference make_folds fixtures (24 cases) + fold-size invariants
//   oracle_exact_knn         brute force with integer coordinates and ties; reference exhaustive HNSW fixtures
//   oracle_knn_vote          hand-built ties (count, score, label) + empty-row fallback; double score sums
//   oracle_objective         closed form acc*sqrt(1-sum p^2), single-class -inf
//   oracle_core_knn          CoreKNN_CPU trajectory fixtures (Core API and KODAMA.matrix flags); live reference
//   oracle_knncv             KNNCV_CPU fixtures (cosine, Euclidean)
//   oracle_knn_graph         live KODAMAKNNGraph_CPU (three metrics, 3 / 6 / 40 columns); brute force with ties
//   oracle_dissimilarity     KODAMADissimilarityInPlace fixtures + brute-force co-assignment counts; live reference
//   oracle_plsldacv (fp32)   PLSLDACV_CPU fixtures, bit-exact; live reference
//   oracle_plsldacv (fp64)   two-class closed form (rank-1 S), G-orthonormal weights, one-component LDA
//   oracle_core_plslda       CorePLSLDA_CPU trajectory fixtures (fp32)
//   oracle_plsdacv           PLSDACV_CPU fixtures, bit-exact (fp32; fp64 the same predictions); live reference
//   oracle_pls               PLS_CPU fixtures, bit-exact (fp32), fp64 within 1e-4; closed forms (rank one,
//                            full-rank least squares)
//   oracle_kmeans_labels,    the reference's own kmeans_labels / quota_sample_landmarks fixtures and live calls
//   oracle_quota_sample
//   oracle_normalize_graph   the reference's normalize_external_graph, live, every index base; a hand-worked case
//   oracle_kodama_matrix,    KODAMAMatrix_CPU (PLS-LDA) and KODAMAMatrixFromGraphData_CPU (kNN) fixtures:
//   oracle_kodama_matrix_    res, acc, v and the run diagnostics' hashes (landmarks, starting labels, folds)
//   from_graph               bit-exact, with starting labels, constraint groups and fixed rows; live
//                            reference on fresh seeds; the exact-graph projection against a numpy brute force
//   PLSLDAPredict (fp32/64)  PLSLDAPredict_CPU fixtures (true and random labels, 2 / 5 / 8 components)

#include "kodama_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <exception>
#include <limits>
#include <map>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

using IVec = std::vector<int>;

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

IVec to_vec(const int32_t* p, int n) { return p ? IVec(p, p + n) : IVec(); }

// ---------------------------------------------------------------------------
// Folds. Follows common.cpp:108-214 (detail::make_folds): one mt19937_64 seeded
// with the fold seed; unconstrained rows are either shuffled and dealt i % F, or
// (stratified) shuffled per label and dealt to the least-loaded fold. Constrained
// rows move as groups (majority label per group).
// ---------------------------------------------------------------------------
IVec make_folds(const IVec& labels, const IVec& constrain, int n_folds, bool stratified,
                std::uint64_t seed) {
  if (n_folds < 2) throw std::invalid_argument("Number of folds must be at least 2.");
  const int n = static_cast<int>(labels.size());
  std::mt19937_64 rng(seed);
  IVec out(static_cast<size_t>(n), 0);
  if (constrain.empty()) {
    if (stratified) {
      std::map<int, IVec> rows_of_label;
      for (int i = 0; i < n; ++i) rows_of_label[labels[i]].push_back(i);
      IVec load(static_cast<size_t>(n_folds), 0);
      for (auto& entry : rows_of_label) {
        IVec& rows = entry.second;
        std::shuffle(rows.begin(), rows.end(), rng);
        std::fill(load.begin(), load.end(), 0);
        for (int r : rows) {
          const int f = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
          out[r] = f;
          ++load[f];
        }
      }
    } else {
      IVec perm(static_cast<size_t>(n));
      std::iota(perm.begin(), perm.end(), 0);
      std::shuffle(perm.begin(), perm.end(), rng);
      for (size_t i = 0; i < perm.size(); ++i) out[perm[i]] = static_cast<int>(i % n_folds);
    }
    return out;
  }

  std::map<int, IVec> members;
  for (int i = 0; i < n; ++i) members[constrain[i]].push_back(i);
  struct Group { int key; int label; int size; };
  std::vector<Group> groups;
  for (const auto& kv : members) {
    std::map<int, int> tally;
    for (int r : kv.second) ++tally[labels[r]];
    int best = tally.begin()->first, best_n = tally.begin()->second;
    for (const auto& t : tally) {
      if (t.second > best_n) { best = t.first; best_n = t.second; }
    }
    groups.push_back({kv.first, best, static_cast<int>(kv.second.size())});
  }
  IVec group_fold(groups.size(), 0);
  if (stratified) {
    std::map<int, std::vector<size_t>> groups_of_label;
    for (size_t g = 0; g < groups.size(); ++g) groups_of_label[groups[g].label].push_back(g);
    IVec load(static_cast<size_t>(n_folds), 0);
    for (auto& entry : groups_of_label) {
      auto& ids = entry.second;
      std::shuffle(ids.begin(), ids.end(), rng);
      std::fill(load.begin(), load.end(), 0);
      for (size_t g : ids) {
        const int f = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
        group_fold[g] = f;
        load[f] += groups[g].size;
      }
    }
  } else {
    std::vector<size_t> perm(groups.size());
    std::iota(perm.begin(), perm.end(), 0);
    std::shuffle(perm.begin(), perm.end(), rng);
    for (size_t i = 0; i < perm.size(); ++i) group_fold[perm[i]] = static_cast<int>(i % n_folds);
  }
  size_t g = 0;
  for (const auto& kv : members) {
    for (int r : kv.second) out[r] = group_fold[g];
    ++g;
  }
  return out;
}

IVec sorted_unique(IVec v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return v;
}

// ---------------------------------------------------------------------------
// Distances. The reference's CPU kNN (native_knn.cpp:41-95, taken on any x86-64
// with AVX2+FMA) accumulates 32 lane partials with fused multiply-adds over blocks
// of 32 dimensions, reduces them lane-wise (s0+s1)+(s2+s3), then 8 -> 4 lanes,
// then (l0+l1)+(l2+l3), and adds the remaining dimensions one at a time with a
// fused multiply-add (GCC contracts `r += d*d` inside the avx2,fma function).
// Squared L2 for Euclidean; minus the dot product for cosine / inner product.
// ---------------------------------------------------------------------------
float ref_order_reduce(const float* a, const float* b, int p, bool euclid) {
  float acc[32];
  for (int s = 0; s < 32; ++s) acc[s] = 0.0f;
  int d = 0;
  for (; d + 31 < p; d += 32) {
    for (int s = 0; s < 32; ++s) {
      if (euclid) {
        const float t = a[d + s] - b[d + s];
        acc[s] = std::fma(t, t, acc[s]);
      } else {
        acc[s] = std::fma(a[d + s], b[d + s], acc[s]);
      }
    }
  }
  float v8[8];
  for (int l = 0; l < 8; ++l) v8[l] = (acc[l] + acc[8 + l]) + (acc[16 + l] + acc[24 + l]);
  float v4[4];
  for (int l = 0; l < 4; ++l) v4[l] = v8[l] + v8[4 + l];
  float r = (v4[0] + v4[1]) + (v4[2] + v4[3]);
  for (; d < p; ++d) {
    if (euclid) {
      const float t = a[d] - b[d];
      r = std::fma(t, t, r);
    } else {
      r = std::fma(a[d], b[d], r);
    }
  }
  return euclid ? r : -r;
}

float distance_of(const float* a, const float* b, int p, int metric) {
  return ref_order_reduce(a, b, p, metric == 2);
}

// Row preparation for cosine search. variant 0 = core.cpp:159-193
// (make_search_matrix: long double sum of squares, double scale); variant 1 =
// native_knn.cpp:1005-1035 (prepare_native_matrix: float sum, float scale).
std::vector<float> prepare_rows(const float* x, int n, int p, int metric, int variant) {
  std::vector<float> out(x, x + static_cast<size_t>(n) * p);
  if (metric != 0) return out;
  for (int r = 0; r < n; ++r) {
    const float* row = x + static_cast<size_t>(r) * p;
    float* dst = out.data() + static_cast<size_t>(r) * p;
    if (variant == 0) {
      long double ss = 0.0L;
      for (int j = 0; j < p; ++j) ss += static_cast<long double>(row[j]) * static_cast<long double>(row[j]);
      const double nrm = std::sqrt(static_cast<double>(ss));
      const double scale = nrm > 0.0 && std::isfinite(nrm) ? 1.0 / nrm : 1.0;
      for (int j = 0; j < p; ++j) dst[j] = static_cast<float>(static_cast<double>(row[j]) * scale);
    } else {
      float ss = 0.0f;
      for (int j = 0; j < p; ++j) ss += row[j] * row[j];
      const float scale = ss > 0.0f && std::isfinite(ss) ? 1.0f / std::sqrt(ss) : 1.0f;
      for (int j = 0; j < p; ++j) dst[j] = row[j] * scale;
    }
  }
  return out;
}

// Exact k nearest neighbours, ascending (distance, index); exclude[q] (if given)
// removes one training index per query. Width-k output, -1 / +inf padded beyond
// min(k, available), as the reference search reports (native_knn.cpp:218-266).
void exact_knn(const float* train, int nt, const float* query, int nq, int p, int k,
               int metric, const int32_t* exclude, int32_t* out_idx, float* out_dist) {
  // Queries are independent: one thread per query block, each with its own
  // candidate list, so the result does not depend on the thread count.
#pragma omp parallel if (static_cast<long long>(nq) * nt * p > (1ll << 24))
  {
    std::vector<std::pair<float, int>> cand;
    cand.reserve(static_cast<size_t>(nt));
#pragma omp for schedule(dynamic, 8)
    for (int q = 0; q < nq; ++q) {
      cand.clear();
      const int ex = exclude ? exclude[q] : -1;
      for (int t = 0; t < nt; ++t) {
        if (t == ex) continue;
        cand.emplace_back(distance_of(query + static_cast<size_t>(q) * p,
                                      train + static_cast<size_t>(t) * p, p, metric), t);
      }
      const int kk = std::min<int>(k, static_cast<int>(cand.size()));
      std::partial_sort(cand.begin(), cand.begin() + kk, cand.end(),
                        [](const std::pair<float, int>& a, const std::pair<float, int>& b) {
                          return a.first < b.first || (a.first == b.first && a.second < b.second);
                        });
      for (int j = 0; j < k; ++j) {
        const size_t o = static_cast<size_t>(q) * k + j;
        out_idx[o] = j < kk ? cand[j].second : -1;
        out_dist[o] = j < kk ? cand[j].first : std::numeric_limits<float>::infinity();
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Per-fold kNN tables (core.cpp:463-584 with an exact search in place of HNSW,
// flattened as core.cpp:311-341): row i of a fold's validation set gets the
// min(k, |train|) nearest training rows of that fold, as landmark-row ids, with
// score = -internal distance (native_knn.cpp:1153-1155).
// ---------------------------------------------------------------------------
struct KnnTables {
  IVec folds;
  int width = 0;
  IVec nbr;                 // [n * width]
  std::vector<float> score; // [n * width]
};

KnnTables precompute_tables(const float* x, int n, int p, const IVec& folds, int k, int metric) {
  if (k < 1) throw std::invalid_argument("KNNOptions::k must be positive.");
  KnnTables t;
  t.folds = folds;
  t.width = k;
  t.nbr.assign(static_cast<size_t>(n) * k, -1);
  t.score.assign(static_cast<size_t>(n) * k, -std::numeric_limits<float>::infinity());
  const std::vector<float> xs = prepare_rows(x, n, p, metric, 0);
  for (int f : sorted_unique(folds)) {
    IVec val, tr;
    for (int i = 0; i < n; ++i) (folds[i] == f ? val : tr).push_back(i);
    if (tr.empty()) throw std::runtime_error("KNN fold has no training samples.");
    const int kf = std::min<int>(k, static_cast<int>(tr.size()));
    std::vector<float> trx(tr.size() * p), vx(val.size() * p);
    for (size_t i = 0; i < tr.size(); ++i) std::memcpy(&trx[i * p], &xs[static_cast<size_t>(tr[i]) * p], sizeof(float) * p);
    for (size_t i = 0; i < val.size(); ++i) std::memcpy(&vx[i * p], &xs[static_cast<size_t>(val[i]) * p], sizeof(float) * p);
    std::vector<int32_t> idx(val.size() * kf);
    std::vector<float> dist(val.size() * kf);
    exact_knn(trx.data(), static_cast<int>(tr.size()), vx.data(), static_cast<int>(val.size()), p, kf,
              metric, nullptr, idx.data(), dist.data());
    for (size_t q = 0; q < val.size(); ++q) {
      for (int j = 0; j < kf; ++j) {
        const int id = idx[q * kf + j];
        const size_t o = static_cast<size_t>(val[q]) * k + j;
        if (id >= 0) t.nbr[o] = tr[id];
        t.score[o] = -dist[q * kf + j];
      }
    }
  }
  return t;
}

// Majority label with ties to the smallest label (core.cpp:292-303, 810-824).
int majority_smallest(const std::map<int, int>& counts) {
  int best = counts.empty() ? 0 : counts.begin()->first;
  int best_n = -1;
  for (const auto& kv : counts) {
    if (kv.second > best_n) { best = kv.first; best_n = kv.second; }
  }
  return best;
}

// The kNN CV vote (core.cpp:784-914 with the fallbacks of 258-309): per validation
// row, labels of its neighbours are tallied; the winner has the most votes, then
// the largest sum of scores (double, neighbour-rank order), then the smallest
// label. Rows with no neighbour take the majority label of their training fold.
IVec vote_predict(const KnnTables& t, const IVec& labels) {
  const int n = static_cast<int>(labels.size());
  IVec pred(static_cast<size_t>(n), 0);
  std::map<int, int> all;
  for (int l : labels) ++all[l];
  const int global_fallback = majority_smallest(all);
  std::vector<int> empty_rows;
  for (int i = 0; i < n; ++i) {
    std::map<int, std::pair<int, double>> tally;
    for (int j = 0; j < t.width; ++j) {
      const int nb = t.nbr[static_cast<size_t>(i) * t.width + j];
      if (nb < 0) continue;
      auto& e = tally[labels[nb]];
      e.first += 1;
      e.second += static_cast<double>(t.score[static_cast<size_t>(i) * t.width + j]);
    }
    if (tally.empty()) {
      pred[i] = global_fallback;
      empty_rows.push_back(i);
      continue;
    }
    int best = 0, best_n = -1;
    double best_s = -std::numeric_limits<double>::infinity();
    for (const auto& kv : tally) {
      const int c = kv.second.first;
      const double s = kv.second.second;
      if (c > best_n || (c == best_n && s > best_s) || (c == best_n && s == best_s && kv.first < best)) {
        best = kv.first; best_n = c; best_s = s;
      }
    }
    pred[i] = best;
  }
  if (!empty_rows.empty()) {
    for (int f : sorted_unique(t.folds)) {
      std::map<int, int> train_counts;
      for (const auto& kv : all) train_counts[kv.first] = kv.second;
      for (int i = 0; i < n; ++i) if (t.folds[i] == f) --train_counts[labels[i]];
      const int fb = majority_smallest(train_counts);
      for (int r : empty_rows) if (t.folds[r] == f) pred[r] = fb;
    }
  }
  return pred;
}

// ---------------------------------------------------------------------------
// Objective (core.cpp:916-972). Class frequencies are visited in the iteration
// order of an unordered_map filled in sample order after reserve(n), as the
// reference does, so the double sum of p^2 rounds identically.
// ---------------------------------------------------------------------------
struct Flags {
  int cycles;
  bool knn;
  std::uint64_t seed;
  bool shake, evolutionary, guarded, absorption, coarsening, adaptive;
  bool guidance, policy_adaptive, transition, stochastic, diversity, pls_coarsening, fragmentation, reject_single;
};

Flags flags_of(const oracle_core_options* o) {
  Flags f;
  f.cycles = o->cycles;
  f.knn = o->classifier == 1;
  f.seed = o->seed;
  f.shake = o->shake; f.evolutionary = o->evolutionary_search; f.guarded = o->guarded_diversity;
  f.absorption = o->many_to_one_absorption; f.coarsening = o->auto_class_coarsening;
  f.adaptive = o->adaptive_proposal_size;
  f.guidance = o->prediction_guidance; f.policy_adaptive = o->policy_adaptive_proposal_size;
  f.transition = o->transition_proposal; f.stochastic = o->stochastic_acceptance;
  f.diversity = o->diversity_multiplier; f.pls_coarsening = o->pls_transition_coarsening;
  f.fragmentation = o->pls_fragmentation_penalty; f.reject_single = o->reject_single_class;
  return f;
}

double objective(const IVec& labels, double acc, const Flags& f) {
  const bool diversity = f.guarded && f.diversity;
  const bool fragmentation = f.coarsening && !f.knn && f.fragmentation;
  const bool reject_single = f.guarded && f.reject_single;
  if (!diversity && !fragmentation && !reject_single) return acc;
  if (labels.empty()) return acc;
  std::unordered_map<int, int> counts;
  counts.reserve(labels.size());
  for (int l : labels) counts[l]++;
  const int k = static_cast<int>(counts.size());
  if (k < 1) return acc;
  if (reject_single && k == 1) return -std::numeric_limits<double>::infinity();
  const double n = static_cast<double>(labels.size());
  double score = acc;
  if (diversity) {
    double same = 0.0;
    for (const auto& kv : counts) {
      const double q = static_cast<double>(kv.second) / n;
      same += q * q;
    }
    score *= std::sqrt(std::max(0.0, 1.0 - same));
  }
  if (fragmentation && k > 1) {
    double h = 0.0;
    for (const auto& kv : counts) {
      const double q = static_cast<double>(kv.second) / n;
      if (q > 0.0) h -= q * std::log(q);
    }
    const double hmax = n > 1.0 ? std::log(n) : 0.0;
    const double code_cost = hmax > 0.0 ? std::clamp(h / hmax, 0.0, 1.0) : 0.0;
    const double keff = std::exp(h);
    const double frag = std::max(0.0, std::log(static_cast<double>(k) / std::max(1.0, keff)));
    const double frag_cost = frag / (1.0 + frag);
    score -= (1.0 - acc) * std::max(code_cost, frag_cost);
  }
  return score;
}

double accuracy(const IVec& truth, const IVec& pred) {
  if (truth.empty()) return 0.0;
  size_t ok = 0;
  for (size_t i = 0; i < truth.size(); ++i) ok += truth[i] == pred[i] ? 1u : 0u;
  return static_cast<double>(ok) / static_cast<double>(truth.size());
}

// ---------------------------------------------------------------------------
// Transition proposals (core.cpp:974-1226).
// ---------------------------------------------------------------------------
struct TransitionStats {
  std::map<int, int> sizes;
  std::map<int, int> movable;
  std::map<int, std::map<int, int>> flows;  // label -> (previous prediction -> rows)
};

TransitionStats transition_stats(const IVec& labels, const IVec& prev_pred, const IVec& fixed) {
  TransitionStats s;
  for (size_t i = 0; i < labels.size(); ++i) {
    ++s.sizes[labels[i]];
    ++s.flows[labels[i]][prev_pred[i]];
    if (fixed.empty() || fixed[i] != 1) ++s.movable[labels[i]];
  }
  return s;
}

// core.cpp:995-1108: merge unstable small classes into the class most of their
// members were predicted as, with a budget set by the class fragmentation.
bool coarsen_classes(IVec& labels, const IVec& fixed, const TransitionStats& st, std::mt19937_64& rng) {
  const int classes = static_cast<int>(st.sizes.size());
  if (classes <= 2) return false;
  const double n = static_cast<double>(labels.size());
  double h = 0.0;
  for (const auto& kv : st.sizes) {
    const double q = static_cast<double>(kv.second) / n;
    if (q > 0.0) h -= q * std::log(q);
  }
  const double keff = std::exp(h);
  const double frag = std::max(0.0, std::log(static_cast<double>(classes) / std::max(1.0, keff)));
  double flow_entropy = 0.0;
  struct Cand { int source; int dest; double score; };
  std::vector<Cand> cands;
  for (const auto& kv : st.sizes) {
    const int src = kv.first;
    const int src_n = kv.second;
    const auto mv = st.movable.find(src);
    const int movable = mv == st.movable.end() ? 0 : mv->second;
    if (movable <= 0) continue;
    int dst = src, dst_n = src_n, best = 0;
    double row_h = 0.0;
    const auto& row = st.flows.at(src);
    for (const auto& e : row) {
      const double q = static_cast<double>(e.second) / static_cast<double>(src_n);
      if (q > 0.0) row_h -= q * std::log(q);
      const auto cs = st.sizes.find(e.first);
      if (e.first != src && cs != st.sizes.end() &&
          (e.second > best || (e.second == best && cs->second > dst_n) ||
           (e.second == best && cs->second == dst_n && e.first < dst))) {
        dst = e.first; dst_n = cs->second; best = e.second;
      }
    }
    flow_entropy += (static_cast<double>(src_n) / n) * row_h;
    if (dst == src || best == 0) continue;
    const double stay = static_cast<double>(row.count(src) ? row.at(src) : 0) / static_cast<double>(src_n);
    cands.push_back({src, dst, (1.0 / std::sqrt(static_cast<double>(movable))) * (1.0 - stay)});
  }
  if (cands.empty()) return false;
  const double temp = frag / (frag + flow_entropy + 1.0);
  const double redundant = std::max(0.0, static_cast<double>(classes) - keff);
  int budget = static_cast<int>(std::ceil(temp * redundant));
  if (budget < 1 && frag > 0.0) budget = 1;
  budget = std::min(budget, classes - 2);
  if (budget <= 0) return false;
  std::shuffle(cands.begin(), cands.end(), rng);
  std::stable_sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.score > b.score; });
  std::unordered_map<int, int> remap;
  remap.reserve(st.sizes.size());
  for (const auto& kv : st.sizes) remap[kv.first] = kv.first;
  int merged = 0;
  for (const Cand& c : cands) {
    if (merged >= budget) break;
    if (remap[c.source] != c.source) continue;
    int d = c.dest;
    while (remap[d] != d) d = remap[d];
    if (d == c.source) continue;
    remap[c.source] = d;
    ++merged;
  }
  if (merged == 0) return false;
  bool changed = false;
  for (size_t i = 0; i < labels.size(); ++i) {
    if (!fixed.empty() && fixed[i] == 1) continue;
    int d = labels[i];
    while (remap[d] != d) d = remap[d];
    if (d != labels[i]) { labels[i] = d; changed = true; }
  }
  return changed;
}

// core.cpp:1110-1226: each class votes for the class it flows into most beyond
// chance; one target (weighted by its surplus) absorbs all its sources.
bool absorb_many_to_one(IVec& labels, const IVec& fixed, const TransitionStats& st, std::mt19937_64& rng) {
  const int classes = static_cast<int>(st.sizes.size());
  if (classes <= 2) return false;
  const double n = static_cast<double>(labels.size());
  struct Src { int source; double surplus; };
  struct Target { int target = 0; double weight = 0.0; std::vector<Src> sources; };
  std::map<int, Target> by_target;
  for (const auto& kv : st.sizes) {
    const int src = kv.first;
    const int src_n = kv.second;
    const auto mv = st.movable.find(src);
    if (mv == st.movable.end() || mv->second <= 0) continue;
    const auto& row = st.flows.at(src);
    const int stay = row.count(src) ? row.at(src) : 0;
    int tgt = src, best = 0, tgt_n = src_n;
    double surplus_best = 0.0;
    for (const auto& e : row) {
      if (e.first == src) continue;
      const auto cs = st.sizes.find(e.first);
      if (cs == st.sizes.end()) continue;
      const double expected = static_cast<double>(src_n) * static_cast<double>(cs->second) / std::max(1.0, n);
      const double surplus = static_cast<double>(e.second) - expected;
      if (surplus > surplus_best || (surplus == surplus_best && e.second > best) ||
          (surplus == surplus_best && e.second == best && cs->second > tgt_n) ||
          (surplus == surplus_best && e.second == best && cs->second == tgt_n && e.first < tgt)) {
        tgt = e.first; tgt_n = cs->second; best = e.second; surplus_best = surplus;
      }
    }
    if (tgt == src || surplus_best <= 0.0 || best <= stay) continue;
    Target& t = by_target[tgt];
    t.target = tgt;
    t.weight += surplus_best;
    t.sources.push_back({src, surplus_best});
  }
  if (by_target.empty()) return false;
  std::vector<Target> multi, single;
  for (auto& kv : by_target) {
    std::sort(kv.second.sources.begin(), kv.second.sources.end(), [](const Src& a, const Src& b) {
      if (a.surplus != b.surplus) return a.surplus > b.surplus;
      return a.source < b.source;
    });
    (kv.second.sources.size() > 1 ? multi : single).push_back(std::move(kv.second));
  }
  std::vector<Target>& moves = multi.empty() ? single : multi;
  if (moves.empty()) return false;
  std::vector<double> w;
  for (const Target& t : moves) w.push_back(std::max(t.weight, std::numeric_limits<double>::min()));
  std::discrete_distribution<int> pick(w.begin(), w.end());
  const Target& chosen = moves[static_cast<size_t>(pick(rng))];
  std::unordered_map<int, int> remap;
  for (const Src& s : chosen.sources) if (s.source != chosen.target) remap[s.source] = chosen.target;
  if (remap.empty()) return false;
  IVec proposed = labels;
  bool changed = false;
  for (size_t i = 0; i < proposed.size(); ++i) {
    if (!fixed.empty() && fixed[i] == 1) continue;
    const auto it = remap.find(proposed[i]);
    if (it == remap.end()) continue;
    proposed[i] = it->second;
    changed = true;
  }
  if (!changed) return false;
  const int after = static_cast<int>(sorted_unique(proposed).size());
  if (after <= 1 || after >= classes) return false;
  labels.swap(proposed);
  return true;
}

// core.cpp:1228-1254: how many groups to relabel this cycle (smoothstep-annealed
// upper bound when adaptive).
int proposal_count(int groups, int cycle, int cycles, bool adaptive, std::mt19937_64& rng) {
  if (groups <= 1) return std::max(0, groups);
  if (!adaptive || cycles <= 1) {
    std::uniform_int_distribution<int> d(1, groups);
    return d(rng);
  }
  const double t = std::clamp(static_cast<double>(cycle + 1) / static_cast<double>(cycles + 1), 0.0, 1.0);
  const double smooth = t * t * (3.0 - 2.0 * t);
  const double temp = std::clamp(1.0 - smooth, 0.0, 1.0);
  const int hi = std::max(1, 1 + static_cast<int>(std::floor(static_cast<double>(groups - 1) * temp)));
  std::uniform_int_distribution<int> d(1, hi);
  return d(rng);
}

// ---------------------------------------------------------------------------
// Label evolution (core.cpp:1268-1549, maximize_core).
// ---------------------------------------------------------------------------
template <class Predict>
void evolve(int n, const IVec& initial, const IVec& constrain, const IVec& fixed_in, const Flags& f,
            Predict&& predict, oracle_core_result* out) {
  if (f.cycles < 0) throw std::invalid_argument("Core cycles must be non-negative.");
  const int T = f.cycles;
  IVec clbest = initial, clbest_dirty = initial;
  std::vector<double> vacc(T, -1.0), vscore(T, -1.0);
  IVec psize(T, 0), active(T, 0);
  std::vector<uint8_t> acc_flag(T, 0), impr(T, 0), temp_flag(T, 0);
  int counters[ORACLE_N_COUNTERS] = {0};
  enum { kCompleted, kProposals, kBest, kCurAcc, kStochAtt, kStochAcc, kCurRej, kCoarse, kAbsorb,
         kTransAtt, kTransAcc, kM2OAtt, kM2OAcc, kPlsAtt, kPlsAcc, kCvEval, kSuccess };
  const IVec fixed = fixed_in.empty() ? IVec(static_cast<size_t>(n), 0) : fixed_in;

  // Groups: one per row without constraints, else one per distinct constraint
  // value in ascending order, members in row order.
  const bool singletons = constrain.empty();
  std::vector<IVec> members;
  IVec groups;
  if (singletons) {
    groups.resize(n);
    std::iota(groups.begin(), groups.end(), 0);
  } else {
    std::map<int, IVec> by_key;
    for (int i = 0; i < n; ++i) by_key[constrain[i]].push_back(i);
    for (auto& kv : by_key) {
      groups.push_back(static_cast<int>(members.size()));
      members.push_back(kv.second);
    }
  }

  std::mt19937_64 rng(f.seed);
  IVec predbest = predict(clbest);
  counters[kCvEval] = 1;
  double accbest = f.shake ? 0.0 : accuracy(clbest, predbest);
  double scorebest = f.shake ? -std::numeric_limits<double>::infinity() : objective(clbest, accbest, f);
  IVec cur = clbest, curpred = predbest;
  double cur_acc = accbest, cur_score = scorebest;
  bool success = false;

  for (int cycle = 0; cycle < T; ++cycle) {
    IVec cl = f.evolutionary ? cur : clbest;
    const IVec cl_dirty = cl;
    const IVec& guide = f.evolutionary ? curpred : predbest;
    if (!groups.empty()) {
      const int take = proposal_count(static_cast<int>(groups.size()), cycle, T,
                                      f.adaptive && f.policy_adaptive, rng);
      psize[cycle] = take;
      IVec order = groups;
      std::shuffle(order.begin(), order.end(), rng);
      order.resize(static_cast<size_t>(take));
      IVec emp_labels, emp_counts;
      if (!f.guidance) {
        std::map<int, int> tally;
        for (int l : cl) ++tally[l];
        for (const auto& kv : tally) { emp_labels.push_back(kv.first); emp_counts.push_back(kv.second); }
      }
      for (int g : order) {
        IVec elig;
        if (singletons) {
          if (fixed[g] != 1) elig.push_back(g);
        } else {
          for (int r : members[g]) if (fixed[r] != 1) elig.push_back(r);
        }
        if (elig.empty()) continue;
        IVec cand, cnt;
        if (f.guidance) {
          std::map<int, int> tally;
          for (int r : elig) ++tally[guide[r]];
          for (const auto& kv : tally) { cand.push_back(kv.first); cnt.push_back(kv.second); }
        } else {
          cand = emp_labels;
          cnt = emp_counts;
        }
        if (cand.empty()) continue;
        std::discrete_distribution<int> pick(cnt.begin(), cnt.end());
        const int repl = cand[static_cast<size_t>(pick(rng))];
        for (int r : elig) cl[r] = repl;
      }
    }

    bool coarsened = false, absorbed = false;
    const bool try_coarsen = f.transition && f.coarsening && f.pls_coarsening;
    const bool try_absorb = f.transition && f.absorption;
    TransitionStats st;
    if (try_coarsen || try_absorb) {
      ++counters[kTransAtt];
      st = transition_stats(cl, guide, fixed);
    }
    if (try_coarsen) {
      ++counters[kPlsAtt];
      coarsened = coarsen_classes(cl, fixed, st, rng);
      if (coarsened) ++counters[kCoarse];
    }
    if (try_absorb) {
      ++counters[kM2OAtt];
      if (coarsened) {
        const TransitionStats st2 = transition_stats(cl, guide, fixed);
        absorbed = absorb_many_to_one(cl, fixed, st2, rng);
      } else {
        absorbed = absorb_many_to_one(cl, fixed, st, rng);
      }
      if (absorbed) ++counters[kAbsorb];
    }

    const IVec pred = predict(cl);
    ++counters[kCvEval];
    ++counters[kProposals];
    const double acc = accuracy(cl, pred);
    const double score = objective(cl, acc, f);
    const bool improves = score > scorebest;
    if (improves) {
      predbest = pred;
      clbest = cl;
      clbest_dirty = cl_dirty;
      accbest = acc;
      scorebest = score;
      ++counters[kBest];
    }
    bool accepted = improves, improving = improves, stochastic = false;
    if (f.evolutionary) {
      const double cooling = 1.0 - static_cast<double>(cycle + 1) / static_cast<double>(std::max(1, T));
      const double temperature = std::max(1.0e-9, 0.10 * std::max(0.0, 1.0 - cur_acc) * cooling);
      bool take = score >= cur_score;
      improving = take;
      if (!take && f.stochastic && temperature > 1.0e-9) {
        ++counters[kStochAtt];
        std::uniform_real_distribution<double> u(0.0, 1.0);
        take = u(rng) < std::exp((score - cur_score) / temperature);
        stochastic = take;
      }
      if (take) {
        cur = cl;
        curpred = pred;
        cur_acc = acc;
        cur_score = score;
        ++counters[kCurAcc];
        if (stochastic) ++counters[kStochAcc];
      } else {
        ++counters[kCurRej];
      }
      accepted = take;
    }
    acc_flag[cycle] = accepted ? 1 : 0;
    impr[cycle] = accepted && improving ? 1 : 0;
    temp_flag[cycle] = stochastic ? 1 : 0;
    if (accepted && (coarsened || absorbed)) {
      ++counters[kTransAcc];
      if (coarsened) ++counters[kPlsAcc];
      if (absorbed) ++counters[kM2OAcc];
    }
    active[cycle] = static_cast<int>(sorted_unique(f.evolutionary ? cur : clbest).size());
    vacc[cycle] = accbest;
    vscore[cycle] = scorebest;
    counters[kCompleted] = cycle + 1;
    if (acc == 1.0 && std::isfinite(score)) success = true;
  }
  counters[kSuccess] = success ? 1 : 0;

  std::memcpy(out->clbest, clbest.data(), sizeof(int32_t) * n);
  std::memcpy(out->clbest_dirty, clbest_dirty.data(), sizeof(int32_t) * n);
  std::memcpy(out->cvpredbest, predbest.data(), sizeof(int32_t) * n);
  for (int c = 0; c < T; ++c) {
    out->vect_acc[c] = vacc[c];
    out->vect_score[c] = vscore[c];
    out->proposal_size[c] = psize[c];
    out->active_classes[c] = active[c];
    out->accepted[c] = acc_flag[c];
    out->improving_acceptance[c] = impr[c];
    out->temperature_acceptance[c] = temp_flag[c];
  }
  std::memcpy(out->counters, counters, sizeof(counters));
  out->accbest = accbest;
  out->scorebest = scorebest;
}

// ---------------------------------------------------------------------------
// PLS-LDA (plscv.cpp). Templated on the working precision T: T = float mirrors
// the reference's float32 arithmetic operation for operation (its accumulators
// that are double stay double), T = double is the FP64 oracle.
// ---------------------------------------------------------------------------
template <class T>
struct Mat {
  int r = 0, c = 0;
  std::vector<T> v;
  Mat() = default;
  Mat(int rr, int cc) : r(rr), c(cc), v(static_cast<size_t>(rr) * cc, T(0)) {}
  T& operator()(int i, int j) { return v[static_cast<size_t>(i) * c + j]; }
  T operator()(int i, int j) const { return v[static_cast<size_t>(i) * c + j]; }
};

// Per-training-set centring and scaling (plscv.cpp:669-698). The scale
// accumulator starts at 1, as in the reference.
template <class T>
void center_scale(const float* x, int p, const IVec& rows, bool center, bool scale,
                  std::vector<T>& mean, std::vector<T>& sd) {
  mean.assign(p, T(0));
  sd.assign(p, T(1));
  if (center) {
    for (int r : rows) for (int j = 0; j < p; ++j) mean[j] += static_cast<T>(x[static_cast<size_t>(r) * p + j]);
    const T inv = rows.empty() ? T(0) : T(1) / static_cast<T>(rows.size());
    for (T& m : mean) m *= inv;
  }
  if (scale) {
    for (int r : rows) for (int j = 0; j < p; ++j) {
      const T d = static_cast<T>(x[static_cast<size_t>(r) * p + j]) - mean[j];
      sd[j] += d * d;
    }
    for (int j = 0; j < p; ++j) {
      const T s = std::sqrt(sd[j] / std::max(T(1), static_cast<T>(rows.size() - 1)));
      sd[j] = s > T(0) && std::isfinite(s) ? s : T(1);
    }
  }
}

template <class T>
Mat<T> standardize(const float* x, int p, const IVec& rows, const std::vector<T>& mean, const std::vector<T>& sd) {
  Mat<T> out(static_cast<int>(rows.size()), p);
  for (int i = 0; i < out.r; ++i)
    for (int j = 0; j < p; ++j)
      out(i, j) = (static_cast<T>(x[static_cast<size_t>(rows[i]) * p + j]) - mean[j]) / sd[j];
  return out;
}

template <class T>
Mat<T> gram_of(const Mat<T>& x) {
  Mat<T> g(x.c, x.c);
  for (int a = 0; a < x.c; ++a)
    for (int b = 0; b <= a; ++b) {
      T v = T(0);
      for (int i = 0; i < x.r; ++i) v += x(i, a) * x(i, b);
      g(a, b) = v;
      g(b, a) = v;
    }
  return g;
}

// dot / mat-vec helpers with double accumulation (plscv.cpp:774-814).
template <class T>
T dotd(const std::vector<T>& a, const std::vector<T>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += static_cast<double>(a[i]) * static_cast<double>(b[i]);
  return static_cast<T>(s);
}
template <class T>
T norm2(const std::vector<T>& a) { return std::sqrt(std::max(T(0), dotd(a, a))); }
template <class T>
T power_norm(const std::vector<T>& a) {
  double s = 0.0;
  for (T v : a) s += static_cast<double>(v) * static_cast<double>(v);
  if (!std::isfinite(s) || s < 0.0) return std::numeric_limits<T>::infinity();
  if (std::is_same<T, float>::value) {
    if (s <= static_cast<double>(std::numeric_limits<float>::max())) return std::sqrt(static_cast<T>(s));
    return static_cast<T>(std::sqrt(s));
  }
  return static_cast<T>(std::sqrt(s));
}
template <class T>
std::vector<T> matvec(const Mat<T>& a, const std::vector<T>& x) {
  std::vector<T> o(a.r, T(0));
  for (int i = 0; i < a.r; ++i) {
    double s = 0.0;
    for (int j = 0; j < a.c; ++j) s += static_cast<double>(a(i, j)) * static_cast<double>(x[j]);
    o[i] = static_cast<T>(s);
  }
  return o;
}
template <class T>
std::vector<T> tmatvec(const Mat<T>& a, const std::vector<T>& x) {
  std::vector<T> o(a.c, T(0));
  for (int j = 0; j < a.c; ++j) {
    double s = 0.0;
    for (int i = 0; i < a.r; ++i) s += static_cast<double>(a(i, j)) * static_cast<double>(x[i]);
    o[j] = static_cast<T>(s);
  }
  return o;
}

// S = X'(Y - ybar) through per-class column sums (plscv.cpp:700-746).
template <class T>
Mat<T> label_crossprod(const Mat<T>& x, const IVec& y, const IVec& classes) {
  const int p = x.c, m = static_cast<int>(classes.size());
  std::vector<T> s(static_cast<size_t>(p) * m, T(0)), xsum(p, T(0)), counts(m, T(0));
  IVec code(y.size());
  for (size_t i = 0; i < y.size(); ++i) {
    code[i] = static_cast<int>(std::lower_bound(classes.begin(), classes.end(), y[i]) - classes.begin());
    counts[code[i]] += T(1);
  }
  for (int i = 0; i < x.r; ++i)
    for (int j = 0; j < p; ++j) {
      const T v = x(i, j);
      xsum[j] += v;
      s[j + static_cast<size_t>(code[i]) * p] += v;
    }
  const T inv = x.r > 0 ? T(1) / static_cast<T>(x.r) : T(0);
  for (int c = 0; c < m; ++c) {
    const T ym = counts[c] * inv;
    for (int j = 0; j < p; ++j) s[j + static_cast<size_t>(c) * p] -= ym * xsum[j];
  }
  Mat<T> out(p, m);
  for (int c = 0; c < m; ++c) for (int j = 0; j < p; ++j) out(j, c) = s[j + static_cast<size_t>(c) * p];
  return out;
}

struct Degenerate : std::runtime_error { Degenerate() : std::runtime_error("degenerate PLS fit") {} };

// Power-refresh SIMPLS (plscv.cpp:816-955): per component, the previous weight
// (or N(0,1) from mt19937(1) for the first) is refreshed by two S S' power steps,
// normalised, scaled to unit score norm, and the loading is Gram-Schmidt
// orthogonalised against earlier loadings before S is deflated.
template <class T>
Mat<T> simpls_power(const Mat<T>& x, Mat<T> s, int max_components, const Mat<T>* gram, Mat<T>* qout = nullptr) {
  const int rank_cap = std::min({max_components, x.c, std::max(1, x.r - 1)});
  const Mat<T> s0 = s;
  Mat<T> w(x.c, rank_cap), vb(x.c, rank_cap), q(s.c, rank_cap);
  std::vector<T> prev;
  int fitted = 0;
  auto refresh = [](const Mat<T>& cp, std::vector<T>& cand) {
    for (int it = 0; it < 2; ++it) {
      const std::vector<T> r = tmatvec(cp, cand);
      cand = matvec(cp, r);
    }
  };
  for (int a = 0; a < rank_cap; ++a) {
    std::vector<T> wt(x.c, T(0));
    if (!prev.empty()) {
      wt = prev;
    } else {
      std::mt19937 gen(1u);
      std::normal_distribution<float> nd(0.0f, 1.0f);
      for (T& v : wt) v = static_cast<T>(nd(gen));
    }
    refresh(s, wt);
    T nrm = power_norm(wt);
    if (prev.empty() && (!std::isfinite(nrm) || nrm <= T(1.0e-10f))) {
      int strongest = -1;
      double strongest_sq = 0.0;
      for (int c = 0; c < s.c; ++c) {
        double sq = 0.0;
        for (int j = 0; j < s.r; ++j) sq += static_cast<double>(s(j, c)) * static_cast<double>(s(j, c));
        if (std::isfinite(sq) && sq > strongest_sq) { strongest_sq = sq; strongest = c; }
      }
      if (strongest >= 0 && strongest_sq > 1.0e-20) {
        for (int j = 0; j < s.r; ++j) wt[j] = s(j, strongest);
        refresh(s, wt);
        nrm = power_norm(wt);
      }
    }
    if (!std::isfinite(nrm) || nrm <= T(1.0e-10f)) break;
    for (T& v : wt) v /= nrm;
    std::vector<T> load;
    if (gram != nullptr) {
      load = matvec(*gram, wt);
      const T tt = dotd(wt, load);
      if (!std::isfinite(tt) || tt <= T(1.0e-20f)) break;
      const T inv = T(1) / std::sqrt(tt);
      for (T& v : wt) v *= inv;
      for (T& v : load) v *= inv;
    } else {
      std::vector<T> t(x.r, T(0));
      for (int i = 0; i < x.r; ++i) {
        T v = T(0);
        for (int j = 0; j < x.c; ++j) v += x(i, j) * wt[j];
        t[i] = v;
      }
      const T tn = norm2(t);
      if (!std::isfinite(tn) || tn <= T(1.0e-10f)) break;
      const T inv = T(1) / tn;
      for (T& v : t) v *= inv;
      for (T& v : wt) v *= inv;
      load = tmatvec(x, t);
    }
    // Response weights q = S0' w of the scaled weight (plscv.cpp:905).
    const std::vector<T> qa = tmatvec(s0, wt);
    for (int c = 0; c < s.c; ++c) q(c, a) = qa[c];
    prev = wt;
    for (int b = 0; b < a; ++b) {
      T proj = T(0);
      for (int j = 0; j < x.c; ++j) proj += vb(j, b) * load[j];
      for (int j = 0; j < x.c; ++j) load[j] -= proj * vb(j, b);
    }
    const T ln = norm2(load);
    if (!std::isfinite(ln) || ln <= T(1.0e-10f)) break;
    for (T& v : load) v /= ln;
    for (int j = 0; j < x.c; ++j) { w(j, a) = wt[j]; vb(j, a) = load[j]; }
    std::vector<T> vs(s.c, T(0));
    for (int c = 0; c < s.c; ++c) {
      T v = T(0);
      for (int j = 0; j < s.r; ++j) v += load[j] * s(j, c);
      vs[c] = v;
    }
    for (int j = 0; j < s.r; ++j) for (int c = 0; c < s.c; ++c) s(j, c) -= load[j] * vs[c];
    fitted = a + 1;
  }
  if (fitted < 1) throw Degenerate();
  if (qout) {
    *qout = Mat<T>(s.c, fitted);
    for (int a = 0; a < fitted; ++a) for (int c = 0; c < s.c; ++c) (*qout)(c, a) = q(c, a);
  }
  if (fitted == rank_cap) return w;
  Mat<T> out(x.c, fitted);
  for (int a = 0; a < fitted; ++a) for (int j = 0; j < x.c; ++j) out(j, a) = w(j, a);
  return out;
}

// Gauss-Jordan elimination with partial pivoting on a + 1e-9 I, columns whose
// pivot is below 1e-12 skipped (solve_linear_float, plscv.cpp:1346-1377).
template <class T>
Mat<T> solve_linear(Mat<T> a, Mat<T> b) {
  const int n = a.r, m = b.c;
  for (int i = 0; i < n; ++i) a(i, i) += static_cast<T>(1.0e-9f);
  for (int col = 0; col < n; ++col) {
    int piv = col;
    T best = std::abs(a(col, col));
    for (int r = col + 1; r < n; ++r) {
      const T v = std::abs(a(r, col));
      if (v > best) { best = v; piv = r; }
    }
    if (best < static_cast<T>(1.0e-12f)) continue;
    if (piv != col) {
      for (int j = 0; j < n; ++j) std::swap(a(col, j), a(piv, j));
      for (int j = 0; j < m; ++j) std::swap(b(col, j), b(piv, j));
    }
    const T div = a(col, col);
    for (int j = col; j < n; ++j) a(col, j) /= div;
    for (int j = 0; j < m; ++j) b(col, j) /= div;
    for (int r = 0; r < n; ++r) {
      if (r == col) continue;
      const T f = a(r, col);
      if (f == T(0)) continue;
      for (int j = col; j < n; ++j) a(r, j) -= f * a(col, j);
      for (int j = 0; j < m; ++j) b(r, j) -= f * b(col, j);
    }
  }
  return b;
}

// Scores X W[:, :A] with a row-sequential accumulation (transform_pls_scores_float,
// plscv.cpp:995-1008).
template <class T>
Mat<T> transform_scores(const Mat<T>& x, const Mat<T>& w, int A) {
  Mat<T> out(x.r, A);
  for (int i = 0; i < x.r; ++i)
    for (int j = 0; j < x.c; ++j) {
      const T v = x(i, j);
      for (int a = 0; a < A; ++a) out(i, a) += v * w(j, a);
    }
  return out;
}

// PLS-DA prediction (plscv.cpp:1210-1296): class frequencies ybar; B solves
// (T'T) B = T'Y - (T'1) ybar' over the first A training scores (per-class
// score sums in the working type, the rest accumulated in double); the
// prediction is the first class of largest ybar_c + t'B[:, c].
template <class T>
IVec predict_pls_da(const Mat<T>& tt, const IVec& y, const Mat<T>& tv, const IVec& classes, int A) {
  const int C = static_cast<int>(classes.size());
  IVec code(y.size());
  std::vector<T> ymean(C, T(0));
  for (size_t i = 0; i < y.size(); ++i) {
    code[i] = static_cast<int>(std::lower_bound(classes.begin(), classes.end(), y[i]) - classes.begin());
    ymean[code[i]] += T(1);
  }
  const T inv = y.empty() ? T(0) : T(1) / static_cast<T>(y.size());
  for (T& m : ymean) m *= inv;
  Mat<T> lhs(A, A), rhs(A, C);
  std::vector<double> sums(A, 0.0), acc(static_cast<size_t>(A) * A, 0.0);
  for (int r = 0; r < tt.r; ++r) {
    const int c = code[r];
    for (int a = 0; a < A; ++a) {
      const double v = static_cast<double>(tt(r, a));
      rhs(a, c) += static_cast<T>(v);
      sums[a] += v;
      for (int b = a; b < A; ++b) acc[static_cast<size_t>(a) * A + b] += v * static_cast<double>(tt(r, b));
    }
  }
  for (int a = 0; a < A; ++a)
    for (int b = a; b < A; ++b) {
      const T v = static_cast<T>(acc[static_cast<size_t>(a) * A + b]);
      lhs(a, b) = v;
      lhs(b, a) = v;
    }
  for (int a = 0; a < A; ++a)
    for (int c = 0; c < C; ++c) rhs(a, c) -= static_cast<T>(sums[a] * static_cast<double>(ymean[c]));
  const Mat<T> coef = solve_linear(lhs, rhs);
  IVec pred(static_cast<size_t>(tv.r), classes.front());
  for (int i = 0; i < tv.r; ++i) {
    int best = 0;
    T best_s = -std::numeric_limits<T>::infinity();
    for (int c = 0; c < C; ++c) {
      T sc = ymean[c];
      for (int a = 0; a < A; ++a) sc += tv(i, a) * coef(a, c);
      if (sc > best_s) { best_s = sc; best = c; }
    }
    pred[i] = classes[best];
  }
  return pred;
}

// Cholesky solve (plscv.cpp:1379-1412).
template <class T>
bool chol_solve(const Mat<T>& m, const Mat<T>& rhs, Mat<T>& sol) {
  const int n = m.r;
  Mat<T> lo(n, n);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c <= r; ++c) {
      T v = m(r, c);
      for (int k = 0; k < c; ++k) v -= lo(r, k) * lo(c, k);
      if (r == c) {
        if (!std::isfinite(v) || v <= T(0)) return false;
        lo(r, c) = std::sqrt(v);
      } else {
        const T d = lo(c, c);
        if (!std::isfinite(d) || d <= T(0)) return false;
        lo(r, c) = v / d;
      }
    }
  sol = rhs;
  for (int col = 0; col < rhs.c; ++col) {
    for (int r = 0; r < n; ++r) {
      T v = sol(r, col);
      for (int k = 0; k < r; ++k) v -= lo(r, k) * sol(k, col);
      sol(r, col) = v / lo(r, r);
    }
    for (int r = n - 1; r >= 0; --r) {
      T v = sol(r, col);
      for (int k = r + 1; k < n; ++k) v -= lo(k, r) * sol(k, col);
      sol(r, col) = v / lo(r, r);
    }
  }
  return true;
}

template <class T>
struct LdaFit { Mat<T> linear; std::vector<T> constant; };

// LDA from score statistics with the ridge grid (plscv.cpp:1419-1487).
template <class T>
LdaFit<T> lda_fit(Mat<T> sums, const IVec& counts, const Mat<T>& tt, int n_train) {
  const int C = sums.r, A = sums.c;
  for (int c = 0; c < C; ++c) {
    const T inv = counts[c] > 0 ? T(1) / static_cast<T>(counts[c]) : T(0);
    for (int a = 0; a < A; ++a) sums(c, a) *= inv;
  }
  Mat<T> pooled(A, A);
  const T df = static_cast<T>(std::max(1, n_train - C));
  T trace = T(0);
  for (int r = 0; r < A; ++r) {
    for (int c2 = 0; c2 <= r; ++c2) {
      T between = T(0);
      for (int c = 0; c < C; ++c) between += static_cast<T>(counts[c]) * sums(c, r) * sums(c, c2);
      const T cov = (tt(r, c2) - between) / df;
      pooled(r, c2) = cov;
      pooled(c2, r) = cov;
    }
    trace += pooled(r, r);
  }
  Mat<T> rhs(A, C);
  for (int c = 0; c < C; ++c) for (int a = 0; a < A; ++a) rhs(a, c) = sums(c, a);
  const T rs = std::isfinite(trace) && trace > T(0) ? trace / static_cast<T>(std::max(1, A)) : T(1);
  const float grid[] = {1e-8f, 1e-6f, 1e-5f, 1e-4f, 1e-3f, 1e-2f};
  Mat<T> sol;
  bool ok = false;
  for (float g : grid) {
    Mat<T> cov = pooled;
    const T lam = static_cast<T>(g) * rs;
    for (int a = 0; a < A; ++a) cov(a, a) += lam;
    if (chol_solve(cov, rhs, sol)) { ok = true; break; }
  }
  if (!ok) throw std::runtime_error("PLS-LDA Cholesky factorization failed.");
  LdaFit<T> m{Mat<T>(C, A), std::vector<T>(C, T(0))};
  for (int c = 0; c < C; ++c) {
    T dm = T(0);
    for (int a = 0; a < A; ++a) {
      m.linear(c, a) = sol(a, c);
      dm += sums(c, a) * m.linear(c, a);
    }
    const T prior = std::max(static_cast<T>(counts[c]) / static_cast<T>(std::max(1, n_train)),
                             static_cast<T>(std::numeric_limits<float>::min()));
    m.constant[c] = T(-0.5) * dm + std::log(prior);
  }
  return m;
}

template <class T>
struct FoldOut {
  IVec pred;                 // per validation row
  int components = 1;
  Mat<T> weights;            // p x rank (empty on fallback)
  Mat<T> val_scores;         // n_val x A
  Mat<T> discriminant;       // n_val x C
  IVec classes;
};

// One fold of PLS-LDA CV (plscv.cpp:2340-2438 with the streamed predictor
// 1544-1608). `gram` selects the Gram-matrix loading path of the fold cache.
template <class T>
FoldOut<T> pls_fold(const float* x, int p, const IVec& labels, const IVec& train, const IVec& val,
                    const oracle_core_options* o, bool use_gram, bool da = false) {
  FoldOut<T> out;
  std::vector<T> mean, sd;
  center_scale<T>(x, p, train, o->center != 0, o->scale != 0, mean, sd);
  const Mat<T> xt = standardize<T>(x, p, train, mean, sd);
  const Mat<T> xv = standardize<T>(x, p, val, mean, sd);
  IVec y(train.size());
  for (size_t i = 0; i < train.size(); ++i) y[i] = labels[train[i]];
  const IVec classes = sorted_unique(y);
  out.classes = classes;
  if (classes.size() <= 1) {
    const int lab = classes.empty() ? (labels.empty() ? 0 : labels.front()) : classes.front();
    out.pred.assign(val.size(), lab);
    out.components = 1;
    return out;
  }
  const int rank_cap = std::max(1, std::min({o->max_components, p, std::max(1, xt.r - 1)}));
  Mat<T> g;
  const bool gram_path = use_gram && xt.c <= xt.r;
  if (gram_path) g = gram_of(xt);
  Mat<T> w;
  try {
    w = simpls_power<T>(xt, label_crossprod(xt, y, classes), rank_cap, gram_path ? &g : nullptr);
  } catch (const Degenerate&) {
    std::map<int, int> tally;
    for (int l : y) ++tally[l];
    int lab = classes.front(), best = -1;
    for (const auto& kv : tally) if (kv.second > best) { lab = kv.first; best = kv.second; }
    out.pred.assign(val.size(), lab);
    out.components = 1;
    return out;
  }
  const int requested = o->fixed_components > 0 ? o->fixed_components : o->max_components;
  const int A = std::min(requested, w.c);
  out.components = A;
  out.weights = w;
  if (da) {
    // PLS-DA head (plscv.cpp:2430-2442): scores of every fitted component,
    // the regression over the first A.
    const Mat<T> tt = transform_scores(xt, w, w.c);
    const Mat<T> tv = transform_scores(xv, w, w.c);
    out.val_scores = tv;
    out.pred = predict_pls_da(tt, y, tv, classes, A);
    return out;
  }
  const int C = static_cast<int>(classes.size());
  IVec code(y.size()), counts(C, 0);
  for (size_t i = 0; i < y.size(); ++i) {
    code[i] = static_cast<int>(std::lower_bound(classes.begin(), classes.end(), y[i]) - classes.begin());
    ++counts[code[i]];
  }
  Mat<T> sums(C, A), tt(A, A);
  std::vector<T> row(A, T(0));
  for (int i = 0; i < xt.r; ++i) {
    std::fill(row.begin(), row.end(), T(0));
    for (int j = 0; j < p; ++j) {
      const T v = xt(i, j);
      for (int a = 0; a < A; ++a) row[a] += v * w(j, a);
    }
    for (int a = 0; a < A; ++a) sums(code[i], a) += row[a];
    for (int r = 0; r < A; ++r) for (int c2 = 0; c2 <= r; ++c2) tt(r, c2) += row[r] * row[c2];
  }
  for (int r = 0; r < A; ++r) for (int c2 = 0; c2 < r; ++c2) tt(c2, r) = tt(r, c2);
  const LdaFit<T> m = lda_fit(sums, counts, tt, xt.r);
  out.val_scores = Mat<T>(xv.r, A);
  out.discriminant = Mat<T>(xv.r, C);
  out.pred.assign(val.size(), classes.front());
  for (int i = 0; i < xv.r; ++i) {
    std::fill(row.begin(), row.end(), T(0));
    for (int j = 0; j < p; ++j) {
      const T v = xv(i, j);
      for (int a = 0; a < A; ++a) row[a] += v * w(j, a);
    }
    for (int a = 0; a < A; ++a) out.val_scores(i, a) = row[a];
    int best = 0;
    T best_s = -std::numeric_limits<T>::infinity();
    for (int c = 0; c < C; ++c) {
      T sc = m.constant[c];
      for (int a = 0; a < A; ++a) sc += row[a] * m.linear(c, a);
      out.discriminant(i, c) = sc;
      if (sc > best_s) { best_s = sc; best = c; }
    }
    out.pred[i] = classes[best];
  }
  return out;
}

// F-fold PLS-LDA CV (plscv.cpp:2293-2507). With the fold cache (non-stratified
// folds and a data epoch, as CorePLSLDA runs it) folds come from the first call
// and the Gram loading path is used; otherwise folds are drawn from the labels.
template <class T>
IVec plsldacv(const float* x, int n, int p, const IVec& labels, const IVec& constrain,
              const oracle_core_options* o, const IVec* cached_folds, IVec* folds_out, int* selected,
              bool da = false) {
  if (o->max_components < 1) throw std::invalid_argument("PLSOptions::max_components must be positive.");
  const bool cache = o->stratified == 0 && o->pls_fold_cache != 0;
  const IVec folds = (cache && cached_folds) ? *cached_folds
                                               : make_folds(labels, constrain, o->folds, o->stratified != 0, o->cv_seed);
  if (folds_out) *folds_out = folds;
  IVec pred(static_cast<size_t>(n), labels.empty() ? 0 : labels.front());
  int evaluated = o->fixed_components > 0 ? o->fixed_components : o->max_components;
  for (int f : sorted_unique(folds)) {
    IVec val, tr;
    for (int i = 0; i < n; ++i) (folds[i] == f ? val : tr).push_back(i);
    const FoldOut<T> fo = pls_fold<T>(x, p, labels, tr, val, o, cache, da);
    for (size_t i = 0; i < val.size(); ++i) pred[val[i]] = fo.pred[i];
    evaluated = std::min(evaluated, fo.components);
  }
  if (selected) *selected = std::min(evaluated, o->max_components);
  return pred;
}

// General PLS (fit_general_pls, plscv.cpp:2765-2878): X centred / scaled as
// the training rows of a fold, Y centred by its float column means, S = X'Y
// summed over rows in order, power-refresh SIMPLS without the Gram path,
// scores T = X W, coefficients W Q' and fitted values ybar + T Q'.
template <class T>
struct PlsFit {
  Mat<T> w, q, scores, coef, fitted;
  std::vector<T> xc, xs, yc;
};

template <class T>
PlsFit<T> pls_general(const float* x, const float* y, int n, int p, int q, const oracle_core_options* o) {
  if (x == nullptr || y == nullptr || n < 2 || p < 1 || q < 1) {
    throw std::invalid_argument("PLS requires non-empty X and Y matrices with at least two rows.");
  }
  const int requested = o->fixed_components > 0 ? o->fixed_components : o->max_components;
  if (requested < 1) throw std::invalid_argument("PLS component count must be positive.");
  IVec rows(static_cast<size_t>(n));
  std::iota(rows.begin(), rows.end(), 0);
  PlsFit<T> f;
  center_scale<T>(x, p, rows, o->center != 0, o->scale != 0, f.xc, f.xs);
  const Mat<T> xp = standardize<T>(x, p, rows, f.xc, f.xs);
  Mat<T> yp(n, q);
  f.yc.assign(q, T(0));
  for (int i = 0; i < n; ++i)
    for (int r = 0; r < q; ++r) {
      const T v = static_cast<T>(y[static_cast<size_t>(i) * q + r]);
      if (!std::isfinite(v)) throw std::invalid_argument("PLS Y must contain only finite values.");
      f.yc[r] += v;
      yp(i, r) = v;
    }
  if (o->center) {
    const T inv = T(1) / static_cast<T>(n);
    for (T& v : f.yc) v *= inv;
    for (int i = 0; i < n; ++i) for (int r = 0; r < q; ++r) yp(i, r) -= f.yc[r];
  }
  const int comps = std::max(1, std::min({requested, p, std::max(1, n - 1)}));
  Mat<T> cross(p, q);
  for (int j = 0; j < p; ++j)
    for (int r = 0; r < q; ++r) {
      T v = T(0);
      for (int i = 0; i < n; ++i) v += xp(i, j) * yp(i, r);
      cross(j, r) = v;
    }
  f.w = simpls_power<T>(xp, cross, comps, nullptr, &f.q);
  const int A = f.w.c;
  f.scores = transform_scores(xp, f.w, A);
  f.coef = Mat<T>(p, q);
  for (int j = 0; j < p; ++j)
    for (int r = 0; r < q; ++r) {
      T v = T(0);
      for (int a = 0; a < A; ++a) v += f.w(j, a) * f.q(r, a);
      f.coef(j, r) = v;
    }
  f.fitted = Mat<T>(n, q);
  for (int i = 0; i < n; ++i)
    for (int r = 0; r < q; ++r) {
      T v = f.yc[r];
      for (int a = 0; a < A; ++a) v += f.scores(i, a) * f.q(r, a);
      f.fitted(i, r) = v;
    }
  return f;
}

// ---------------------------------------------------------------------------
// KODAMA dissimilarity (kodama_matrix.cpp:2685-2776): per edge, the fraction of
// runs that put both ends in the same (non-zero) class rescales (1 + d); rows are
// re-sorted by (distance, neighbour id).
// ---------------------------------------------------------------------------
void dissimilarity(int32_t* idx, float* dist, int samples, int K, const int32_t* res, int runs,
                   bool in_one, bool out_one) {
  std::vector<std::pair<float, int>> row(K);
  for (int i = 0; i < samples; ++i) {
    for (int j = 0; j < K; ++j) {
      const size_t o = static_cast<size_t>(i) * K + j;
      const int stored = idx[o];
      const int nb = stored >= 0 && in_one ? stored - 1 : stored;
      float d = dist[o];
      if (nb < 0 || nb >= samples || !std::isfinite(d)) {
        row[j] = {std::numeric_limits<float>::infinity(), stored};
        continue;
      }
      int same = 0, valid = 0;
      for (int r = 0; r < runs; ++r) {
        const int a = res[static_cast<size_t>(r) * samples + i];
        const int b = res[static_cast<size_t>(r) * samples + nb];
        const int v = a != 0 && b != 0;
        valid += v;
        same += v && a == b;
      }
      if (same == 0 || valid == 0) {
        d = std::numeric_limits<float>::infinity();
      } else {
        const double agree = static_cast<double>(same) / static_cast<double>(valid);
        d = static_cast<float>((1.0 + static_cast<double>(d)) / (agree * agree));
      }
      row[j] = {d, out_one ? nb + 1 : nb};
    }
    std::stable_sort(row.begin(), row.end(), [](const std::pair<float, int>& a, const std::pair<float, int>& b) {
      if (a.first != b.first) return a.first < b.first;
      return a.second < b.second;
    });
    for (int j = 0; j < K; ++j) {
      const size_t o = static_cast<size_t>(i) * K + j;
      dist[o] = row[j].first;
      idx[o] = row[j].second;
    }
  }
}

// ---------------------------------------------------------------------------
// KODAMA.matrix orchestration (kodama_matrix.cpp): the shared k-means landmark
// atlas, per-run quota landmarks and k-means starting labels, the Core on the
// landmarks and the projection of the other rows through the global graph.
// ---------------------------------------------------------------------------

// dot_float32 as the reference's portable x86-64 build computes it
// (kodama_matrix.cpp:91-124; SSE without FMA): 16 running partials over blocks
// of 16 (product rounded, then sum rounded), lanes (s[l]+s[4+l])+(s[8+l]+s[12+l]),
// then ((l0+l1)+l2)+l3, then the remaining products one at a time.
float sse_dot(const float* a, const float* b, int p) {
  float s[16];
  for (int j = 0; j < 16; ++j) s[j] = 0.0f;
  int i = 0;
  for (; i + 15 < p; i += 16) {
    for (int j = 0; j < 16; ++j) {
      const float prod = a[i + j] * b[i + j];
      s[j] = s[j] + prod;
    }
  }
  float lane[4];
  for (int l = 0; l < 4; ++l) lane[l] = (s[l] + s[4 + l]) + (s[8 + l] + s[12 + l]);
  float sum = ((lane[0] + lane[1]) + lane[2]) + lane[3];
  for (; i < p; ++i) {
    const float prod = a[i] * b[i];
    sum = sum + prod;
  }
  return sum;
}

// FaissCompatibleRandom (kodama_matrix.cpp:535-549): mt19937 of the low 32 bits.
struct FaissRng {
  std::mt19937 g;
  explicit FaissRng(std::uint64_t seed) : g(static_cast<std::uint32_t>(seed)) {}
  int integer(int m) { return static_cast<int>(g() % static_cast<std::uint32_t>(m)); }
  float uniform() { return static_cast<float>(g()) / static_cast<float>(std::mt19937::max()); }
};

// faiss_compatible_permutation (kodama_matrix.cpp:551-564).
IVec faiss_permutation(int n, std::uint64_t seed) {
  IVec perm(static_cast<size_t>(n));
  std::iota(perm.begin(), perm.end(), 0);
  FaissRng r(seed);
  for (int i = 0; i + 1 < n; ++i) std::swap(perm[i], perm[i + r.integer(n - i)]);
  return perm;
}

// Float Fenwick tree of the empty-cluster donor weights (kodama_matrix.cpp:566-603).
struct Fenwick {
  std::vector<float> tree, vals;
  explicit Fenwick(const std::vector<float>& v) : tree(v.size() + 1, 0.0f), vals(v.size(), 0.0f) {
    for (size_t i = 0; i < v.size(); ++i) set(i, v[i]);
  }
  void set(size_t i, float v) {
    const float delta = v - vals[i];
    vals[i] = v;
    for (++i; i < tree.size(); i += i & (~i + 1)) tree[i] += delta;
  }
  float total() const {
    float r = 0.0f;
    for (size_t i = vals.size(); i > 0; i -= i & (~i + 1)) r += tree[i];
    return r;
  }
  int select(float target) const {
    size_t i = 0, step = 1;
    float prefix = 0.0f;
    while ((step << 1) < tree.size()) step <<= 1;
    for (; step > 0; step >>= 1) {
      const size_t nx = i + step;
      if (nx < tree.size() && prefix + tree[nx] <= target) {
        i = nx;
        prefix += tree[nx];
      }
    }
    return i < vals.size() ? static_cast<int>(i) : -1;
  }
};

// split_empty_kmeans_clusters (kodama_matrix.cpp:605-662).
void split_empty(int rows, int p, int k, std::vector<float>& sizes, std::vector<float>& cent) {
  if (rows <= k) return;
  const float eps = 1.0f / 1024.0f;
  FaissRng r(1234u);
  std::vector<float> w(static_cast<size_t>(k));
  for (int c = 0; c < k; ++c) w[c] = std::max(0.0f, sizes[c] - 1.0f);
  Fenwick tree(w);
  for (int e = 0; e < k; ++e) {
    if (sizes[e] != 0.0f) continue;
    const float total = tree.total();
    int d = total > 0.0f ? tree.select(r.uniform() * total) : -1;
    if (d < 0) {
      d = 0;
      for (int c = 1; c < k; ++c) if (sizes[c] > sizes[d]) d = c;
    }
    float* ce = cent.data() + static_cast<size_t>(e) * p;
    float* cd = cent.data() + static_cast<size_t>(d) * p;
    std::copy_n(cd, p, ce);
    for (int j = 0; j < p; ++j) {
      if ((j & 1) == 0) {
        ce[j] *= 1.0f + eps;
        cd[j] *= 1.0f - eps;
      } else {
        ce[j] *= 1.0f - eps;
        cd[j] *= 1.0f + eps;
      }
    }
    sizes[e] = sizes[d] / 2.0f;
    sizes[d] -= sizes[e];
    tree.set(static_cast<size_t>(e), std::max(0.0f, sizes[e] - 1.0f));
    tree.set(static_cast<size_t>(d), std::max(0.0f, sizes[d] - 1.0f));
  }
}

// kmeans_labels, CPU path (kodama_matrix.cpp:664-799): 1-based labels.
IVec kmeans_labels(const float* x, int n, int p, int k, std::mt19937_64& rng, int max_iter) {
  k = std::max(1, std::min(k, n));
  const std::uint64_t kseed = rng() & 0x7fffffffULL;
  const IVec order = faiss_permutation(n, kseed + 1u);
  std::vector<float> cent(static_cast<size_t>(k) * p), cn(static_cast<size_t>(k)), pn(static_cast<size_t>(n)),
    sizes(static_cast<size_t>(k), 0.0f);
  for (int c = 0; c < k; ++c) std::copy_n(x + static_cast<size_t>(order[c]) * p, p, cent.data() + static_cast<size_t>(c) * p);
  for (int i = 0; i < n; ++i) pn[i] = sse_dot(x + static_cast<size_t>(i) * p, x + static_cast<size_t>(i) * p, p);
  IVec asg(static_cast<size_t>(n), -1);
  auto assign = [&]() {
    for (int c = 0; c < k; ++c) {
      const float* cc = cent.data() + static_cast<size_t>(c) * p;
      cn[c] = sse_dot(cc, cc, p);
    }
    int changed = 0;
    for (int i = 0; i < n; ++i) {
      int best = 0;
      float bd = std::numeric_limits<float>::infinity();
      const float* xi = x + static_cast<size_t>(i) * p;
      for (int c = 0; c < k; ++c) {
        const float dot = sse_dot(xi, cent.data() + static_cast<size_t>(c) * p, p);
        const float sum = pn[i] + cn[c];
        const float twice = 2.0f * dot;
        const float d = sum - twice;
        if (d < bd || (d == bd && c < best)) {
          bd = d;
          best = c;
        }
      }
      if (asg[i] != best) {
        asg[i] = best;
        ++changed;
      }
    }
    return changed;
  };
  for (int it = 0; it < std::max(1, max_iter); ++it) {
    const int changed = assign();
    std::fill(cent.begin(), cent.end(), 0.0f);
    std::fill(sizes.begin(), sizes.end(), 0.0f);
    for (int i = 0; i < n; ++i) {
      const int c = asg[i];
      sizes[c] += 1.0f;
      float* cc = cent.data() + static_cast<size_t>(c) * p;
      const float* xi = x + static_cast<size_t>(i) * p;
      for (int j = 0; j < p; ++j) cc[j] += xi[j];
    }
    for (int c = 0; c < k; ++c) {
      if (sizes[c] == 0.0f) continue;
      const float scale = 1.0f / sizes[c];
      float* cc = cent.data() + static_cast<size_t>(c) * p;
      for (int j = 0; j < p; ++j) cc[j] *= scale;
    }
    split_empty(n, p, k, sizes, cent);
    if (changed == 0) break;
  }
  assign();
  IVec labels(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) labels[i] = asg[i] + 1;
  return labels;
}

// index_strata (kodama_matrix.cpp:800-824): rows of each stratum, ascending.
void index_strata(const IVec& labels, int strata, IVec& offsets, IVec& rows) {
  offsets.assign(static_cast<size_t>(strata) + 1, 0);
  rows.assign(labels.size(), 0);
  for (int l : labels) {
    if (l < 1 || l > strata) throw std::invalid_argument("stratum label out of range.");
    ++offsets[l];
  }
  for (int s = 1; s <= strata; ++s) offsets[s] += offsets[s - 1];
  IVec cur = offsets;
  for (size_t i = 0; i < labels.size(); ++i) rows[cur[labels[i] - 1]++] = static_cast<int>(i);
}

// quota_sample_landmarks with shuffled strata (kodama_matrix.cpp:856-944):
// floor quotas, residual slots by a systematic draw over the shuffled strata,
// std::sample inside each stratum; sorted rows.
IVec quota_sample(const IVec& offsets, const IVec& rows, int target, std::mt19937_64& rng) {
  const int samples = static_cast<int>(rows.size());
  target = std::max(1, std::min(target, samples));
  const int strata = static_cast<int>(offsets.size()) - 1;
  IVec quota(static_cast<size_t>(strata), 0), order;
  std::vector<long double> frac(static_cast<size_t>(strata), 0.0L);
  int assigned = 0;
  for (int s = 0; s < strata; ++s) {
    const int count = offsets[s + 1] - offsets[s];
    if (count <= 0) continue;
    order.push_back(s);
    const long double expected =
      static_cast<long double>(target) * static_cast<long double>(count) / static_cast<long double>(samples);
    const int base = std::min(count, static_cast<int>(std::floor(expected)));
    quota[s] = base;
    frac[s] = expected - static_cast<long double>(base);
    assigned += base;
  }
  std::shuffle(order.begin(), order.end(), rng);
  const int residual = target - assigned;
  if (residual > 0) {
    const long double offset = std::generate_canonical<long double, 64>(rng);
    long double cum = 0.0L;
    int selected = 0;
    std::vector<char> up(static_cast<size_t>(strata), 0);
    for (int s : order) {
      cum += frac[s];
      const long double next = offset + static_cast<long double>(selected);
      if (selected < residual && next < cum) {
        ++quota[s];
        up[s] = 1;
        ++selected;
      }
    }
    while (selected < residual) {
      int best = -1;
      long double bf = -1.0L;
      for (int s : order) {
        if (up[s]) continue;
        if (frac[s] > bf) {
          bf = frac[s];
          best = s;
        }
      }
      if (best < 0) throw std::runtime_error("landmark quota allocation failed.");
      ++quota[best];
      up[best] = 1;
      ++selected;
    }
  }
  IVec out;
  out.reserve(static_cast<size_t>(target));
  for (int s = 0; s < strata; ++s) {
    if (quota[s] <= 0) continue;
    std::sample(rows.cbegin() + offsets[s], rows.cbegin() + offsets[s + 1], std::back_inserter(out), quota[s], rng);
  }
  std::sort(out.begin(), out.end());
  return out;
}

// majority_label (kodama_matrix.cpp:169-181): most votes, then smallest label.
int majority_label(const IVec& v) {
  std::map<int, int> counts;
  for (int x : v) counts[x] += 1;
  int best = counts.begin()->first, bc = counts.begin()->second;
  for (const auto& kv : counts) {
    if (kv.second > bc || (kv.second == bc && kv.first < best)) {
      best = kv.first;
      bc = kv.second;
    }
  }
  return best;
}

void copy_fold_out(const IVec& folds, int32_t* out) {
  if (out) std::memcpy(out, folds.data(), sizeof(int32_t) * folds.size());
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

int oracle_make_folds(const int32_t* labels, int32_t n, const int32_t* constrain, int32_t folds,
                      int32_t stratified, uint64_t seed, int32_t* out) {
  return guarded([&] {
    const IVec f = make_folds(IVec(labels, labels + n), to_vec(constrain, n), folds, stratified != 0, seed);
    std::memcpy(out, f.data(), sizeof(int32_t) * n);
  });
}

float oracle_distance(const float* a, const float* b, int32_t p, int32_t metric) {
  return distance_of(a, b, p, metric);
}

void oracle_prepare_rows(const float* x, int32_t n, int32_t p, int32_t metric, int32_t variant, float* out) {
  const std::vector<float> r = prepare_rows(x, n, p, metric, variant);
  std::memcpy(out, r.data(), sizeof(float) * r.size());
}

int oracle_exact_knn(const float* train, int32_t nt, const float* query, int32_t nq, int32_t p,
                     int32_t k, int32_t metric, const int32_t* exclude, int32_t* out_idx, float* out_dist) {
  return guarded([&] { exact_knn(train, nt, query, nq, p, k, metric, exclude, out_idx, out_dist); });
}

int oracle_precompute_knn_cv(const float* x, int32_t n, int32_t p, const int32_t* folds, int32_t k,
                             int32_t metric, int32_t* nbr, float* score) {
  return guarded([&] {
    const KnnTables t = precompute_tables(x, n, p, IVec(folds, folds + n), k, metric);
    std::memcpy(nbr, t.nbr.data(), sizeof(int32_t) * t.nbr.size());
    std::memcpy(score, t.score.data(), sizeof(float) * t.score.size());
  });
}

int oracle_knn_vote(const int32_t* nbr, const float* score, int32_t n, int32_t width,
                    const int32_t* folds, const int32_t* labels, int32_t* pred) {
  return guarded([&] {
    KnnTables t;
    t.folds.assign(folds, folds + n);
    t.width = width;
    t.nbr.assign(nbr, nbr + static_cast<size_t>(n) * width);
    t.score.assign(score, score + static_cast<size_t>(n) * width);
    const IVec r = vote_predict(t, IVec(labels, labels + n));
    std::memcpy(pred, r.data(), sizeof(int32_t) * n);
  });
}

double oracle_objective(const int32_t* labels, int32_t n, double acc, const oracle_core_options* opt) {
  return objective(IVec(labels, labels + n), acc, flags_of(opt));
}

int oracle_core_knn(const float* x, int32_t n, int32_t p, const int32_t* labels, const int32_t* constrain,
                    const int32_t* fixed, const oracle_core_options* opt, oracle_core_result* out) {
  return guarded([&] {
    const IVec init(labels, labels + n);
    const IVec cons = to_vec(constrain, n);
    const IVec folds = make_folds(init, cons, opt->folds, opt->stratified != 0, opt->cv_seed);
    const KnnTables t = precompute_tables(x, n, p, folds, opt->k, opt->metric);
    Flags f = flags_of(opt);
    f.knn = true;
    evolve(n, init, cons, to_vec(fixed, n), f, [&](const IVec& l) { return vote_predict(t, l); }, out);
    copy_fold_out(folds, out->fold_assignments);
  });
}

int oracle_core_knn_tables(const int32_t* nbr, const float* score, int32_t n, int32_t width,
                           const int32_t* folds, const int32_t* labels, const int32_t* constrain,
                           const int32_t* fixed, const oracle_core_options* opt, oracle_core_result* out) {
  return guarded([&] {
    KnnTables t;
    t.folds.assign(folds, folds + n);
    t.width = width;
    t.nbr.assign(nbr, nbr + static_cast<size_t>(n) * width);
    t.score.assign(score, score + static_cast<size_t>(n) * width);
    Flags f = flags_of(opt);
    f.knn = true;
    evolve(n, IVec(labels, labels + n), to_vec(constrain, n), to_vec(fixed, n), f,
           [&](const IVec& l) { return vote_predict(t, l); }, out);
    copy_fold_out(t.folds, out->fold_assignments);
  });
}

// KNNCV (knncv.cpp:366-419) with an exact search: per fold, the k nearest
// training rows vote with (count, sum of scores, smallest label); rows with no
// neighbour keep the label of the fold's first training row.
int oracle_knncv(const float* x, int32_t n, int32_t p, const int32_t* labels, const int32_t* constrain,
                 const oracle_core_options* opt, int32_t* pred, int32_t* folds_out) {
  return guarded([&] {
    if (opt->k < 1) throw std::invalid_argument("KNNOptions::k must be positive.");
    const IVec lab(labels, labels + n);
    const IVec folds = make_folds(lab, to_vec(constrain, n), opt->folds, opt->stratified != 0, opt->cv_seed);
    const std::vector<float> xs = prepare_rows(x, n, p, opt->metric, 1);
    IVec out(static_cast<size_t>(n), 0);
    for (int f : sorted_unique(folds)) {
      IVec val, tr;
      for (int i = 0; i < n; ++i) (folds[i] == f ? val : tr).push_back(i);
      if (tr.empty()) throw std::runtime_error("KNN fold has no training samples.");
      const int kf = std::min<int>(opt->k, static_cast<int>(tr.size()));
      std::vector<float> trx(tr.size() * p), vx(val.size() * p);
      for (size_t i = 0; i < tr.size(); ++i) std::memcpy(&trx[i * p], &xs[static_cast<size_t>(tr[i]) * p], sizeof(float) * p);
      for (size_t i = 0; i < val.size(); ++i) std::memcpy(&vx[i * p], &xs[static_cast<size_t>(val[i]) * p], sizeof(float) * p);
      std::vector<int32_t> idx(val.size() * kf);
      std::vector<float> dist(val.size() * kf);
      exact_knn(trx.data(), static_cast<int>(tr.size()), vx.data(), static_cast<int>(val.size()), p, kf,
                opt->metric, nullptr, idx.data(), dist.data());
      for (size_t q = 0; q < val.size(); ++q) {
        std::map<int, std::pair<int, double>> tally;
        for (int j = 0; j < kf; ++j) {
          const int id = idx[q * kf + j];
          if (id < 0) continue;
          auto& e = tally[lab[tr[id]]];
          e.first += 1;
          e.second += static_cast<double>(-dist[q * kf + j]);
        }
        int best = lab[tr.front()];
        if (!tally.empty()) {
          best = tally.begin()->first;
          int bn = tally.begin()->second.first;
          double bs = tally.begin()->second.second;
          for (const auto& kv : tally) {
            if (kv.second.first > bn || (kv.second.first == bn && kv.second.second > bs) ||
                (kv.second.first == bn && kv.second.second == bs && kv.first < best)) {
              best = kv.first; bn = kv.second.first; bs = kv.second.second;
            }
          }
        }
        out[val[q]] = best;
      }
    }
    std::memcpy(pred, out.data(), sizeof(int32_t) * n);
    copy_fold_out(folds, folds_out);
  });
}

// Exact self kNN graph (graph_cluster.cpp:790-837 contract): k = min(k, n-1),
// self excluded, 1-based ids, public distance sqrt(max(0, d2)) for Euclidean,
// max(0, 1 + internal) for cosine, 1 + internal for inner product
// (native_knn.cpp:1157-1164).
int oracle_knn_graph(const float* x, int32_t n, int32_t p, int32_t k, int32_t metric,
                     int32_t* idx_one_based, float* dist) {
  return guarded([&] {
    if (n < 2) throw std::invalid_argument("KODAMAKNNGraph requires at least two rows.");
    const int kk = std::max(1, std::min(k, n - 1));
    const std::vector<float> xs = prepare_rows(x, n, p, metric, 1);
    IVec ex(n);
    std::iota(ex.begin(), ex.end(), 0);
    std::vector<int32_t> idx(static_cast<size_t>(n) * kk);
    std::vector<float> d(static_cast<size_t>(n) * kk);
    exact_knn(xs.data(), n, xs.data(), n, p, kk, metric, ex.data(), idx.data(), d.data());
    for (size_t i = 0; i < idx.size(); ++i) {
      idx_one_based[i] = idx[i] >= 0 ? idx[i] + 1 : -1;
      const float v = d[i];
      dist[i] = metric == 2 ? std::sqrt(std::max(0.0f, v)) : (metric == 0 ? std::max(0.0f, 1.0f + v) : 1.0f + v);
    }
  });
}

int oracle_dissimilarity(int32_t* indices, float* distances, int32_t samples, int32_t neighbors,
                         const int32_t* res, int32_t runs, int32_t input_one_based, int32_t output_one_based) {
  return guarded([&] {
    dissimilarity(indices, distances, samples, neighbors, res, runs, input_one_based != 0, output_one_based != 0);
  });
}

int oracle_plsldacv(const float* x, int32_t n, int32_t p, const int32_t* labels, const int32_t* constrain,
                    const oracle_core_options* opt, int32_t precision, int32_t* pred, int32_t* folds_out,
                    int32_t* selected_components) {
  return guarded([&] {
    const IVec lab(labels, labels + n);
    IVec folds;
    int sel = 1;
    const IVec r = precision == 32
      ? plsldacv<float>(x, n, p, lab, to_vec(constrain, n), opt, nullptr, &folds, &sel)
      : plsldacv<double>(x, n, p, lab, to_vec(constrain, n), opt, nullptr, &folds, &sel);
    std::memcpy(pred, r.data(), sizeof(int32_t) * n);
    copy_fold_out(folds, folds_out);
    if (selected_components) *selected_components = sel;
  });
}

// PLSDACV (plscv.cpp:2926-2928): the PLS-LDA CV with the PLS-DA head.
int oracle_plsdacv(const float* x, int32_t n, int32_t p, const int32_t* labels, const int32_t* constrain,
                   const oracle_core_options* opt, int32_t precision, int32_t* pred, int32_t* folds_out,
                   int32_t* selected_components) {
  return guarded([&] {
    const IVec lab(labels, labels + n);
    IVec folds;
    int sel = 1;
    const IVec r = precision == 32
      ? plsldacv<float>(x, n, p, lab, to_vec(constrain, n), opt, nullptr, &folds, &sel, true)
      : plsldacv<double>(x, n, p, lab, to_vec(constrain, n), opt, nullptr, &folds, &sel, true);
    std::memcpy(pred, r.data(), sizeof(int32_t) * n);
    copy_fold_out(folds, folds_out);
    if (selected_components) *selected_components = sel;
  });
}

// PLS_CPU (plscv.cpp:2899-2901). Outputs in double: weights [p x A], response
// loadings [q x A], scores [n x A], coefficients [p x q], fitted [n x q],
// centres / scales; *components = A (buffers sized for the requested count).
int oracle_pls(const float* x, int32_t n, int32_t p, const float* y, int32_t q, const oracle_core_options* opt,
               int32_t precision, double* weights, double* response_loadings, double* scores,
               double* coefficients, double* fitted, double* x_center, double* x_scale, double* y_center,
               int32_t* components) {
  return guarded([&] {
    auto emit = [&](const auto& f) {
      const int A = f.w.c;
      *components = A;
      for (int j = 0; j < p; ++j) for (int a = 0; a < A; ++a) weights[static_cast<size_t>(j) * A + a] = f.w(j, a);
      for (int r = 0; r < q; ++r) for (int a = 0; a < A; ++a) response_loadings[static_cast<size_t>(r) * A + a] = f.q(r, a);
      for (int i = 0; i < n; ++i) for (int a = 0; a < A; ++a) scores[static_cast<size_t>(i) * A + a] = f.scores(i, a);
      for (size_t e = 0; e < f.coef.v.size(); ++e) coefficients[e] = f.coef.v[e];
      for (size_t e = 0; e < f.fitted.v.size(); ++e) fitted[e] = f.fitted.v[e];
      for (int j = 0; j < p; ++j) { x_center[j] = f.xc[j]; x_scale[j] = f.xs[j]; }
      for (int r = 0; r < q; ++r) y_center[r] = f.yc[r];
    };
    if (precision == 32) emit(pls_general<float>(x, y, n, p, q, opt));
    else emit(pls_general<double>(x, y, n, p, q, opt));
  });
}

// Per-fold detail of one PLS-LDA fold for tolerance checks: weights [p x rank]
// (row-major, first `rank` columns), validation scores [n_val x A], discriminant
// scores [n_val x C] and the class list. Buffers must hold p*max_components,
// n*max_components, n*n_classes_max (n) and n entries.
int oracle_pls_fold_detail(const float* x, int32_t n, int32_t p, const int32_t* labels, const int32_t* folds,
                           int32_t fold, const oracle_core_options* opt, int32_t precision, double* weights,
                           int32_t* rank, double* val_scores, double* discriminants, int32_t* classes,
                           int32_t* n_classes, int32_t* n_val) {
  return guarded([&] {
    IVec val, tr;
    for (int i = 0; i < n; ++i) (folds[i] == fold ? val : tr).push_back(i);
    const IVec lab(labels, labels + n);
    auto emit = [&](const auto& fo) {
      *rank = fo.weights.c;
      for (int j = 0; j < fo.weights.r; ++j)
        for (int a = 0; a < fo.weights.c; ++a)
          weights[static_cast<size_t>(j) * fo.weights.c + a] = static_cast<double>(fo.weights(j, a));
      for (size_t i = 0; i < fo.val_scores.v.size(); ++i) val_scores[i] = static_cast<double>(fo.val_scores.v[i]);
      for (size_t i = 0; i < fo.discriminant.v.size(); ++i) discriminants[i] = static_cast<double>(fo.discriminant.v[i]);
      for (size_t i = 0; i < fo.classes.size(); ++i) classes[i] = fo.classes[i];
      *n_classes = static_cast<int>(fo.classes.size());
      *n_val = static_cast<int>(val.size());
    };
    const bool cache = opt->stratified == 0 && opt->pls_fold_cache != 0;
    if (precision == 32) emit(pls_fold<float>(x, p, lab, tr, val, opt, cache));
    else emit(pls_fold<double>(x, p, lab, tr, val, opt, cache));
  });
}

int oracle_core_plslda(const float* x, int32_t n, int32_t p, const int32_t* labels, const int32_t* constrain,
                       const int32_t* fixed, const oracle_core_options* opt, int32_t precision,
                       oracle_core_result* out) {
  return guarded([&] {
    const IVec init(labels, labels + n);
    const IVec cons = to_vec(constrain, n);
    oracle_core_options o = *opt;
    o.pls_fold_cache = 1;
    // The fold cache is keyed on data, not labels: folds come from the first call.
    const IVec first_folds = make_folds(init, cons, o.folds, o.stratified != 0, o.cv_seed);
    Flags f = flags_of(&o);
    f.knn = false;
    evolve(n, init, cons, to_vec(fixed, n), f, [&](const IVec& l) {
      return precision == 32 ? plsldacv<float>(x, n, p, l, cons, &o, &first_folds, nullptr, nullptr)
                             : plsldacv<double>(x, n, p, l, cons, &o, &first_folds, nullptr, nullptr);
    }, out);
    copy_fold_out(first_folds, out->fold_assignments);
  });
}


int oracle_kmeans_labels(const float* x, int32_t n, int32_t p, int32_t k, uint64_t rng_seed, int32_t discard,
                         int32_t max_iter, int32_t* labels) {
  return guarded([&] {
    std::mt19937_64 rng(rng_seed);
    for (int i = 0; i < discard; ++i) (void)rng();
    const IVec l = kmeans_labels(x, n, p, k, rng, max_iter);
    std::memcpy(labels, l.data(), sizeof(int32_t) * l.size());
  });
}

int oracle_quota_sample(const int32_t* strata_labels, int32_t n, int32_t strata, int32_t target, uint64_t rng_seed,
                        int32_t discard, int32_t* rows_out, int32_t* count_out) {
  return guarded([&] {
    IVec off, rows;
    index_strata(IVec(strata_labels, strata_labels + n), strata, off, rows);
    std::mt19937_64 rng(rng_seed);
    for (int i = 0; i < discard; ++i) (void)rng();
    const IVec r = quota_sample(off, rows, target, rng);
    std::memcpy(rows_out, r.data(), sizeof(int32_t) * r.size());
    *count_out = static_cast<int32_t>(r.size());
  });
}

}  // extern "C"

namespace {

// normalize_external_graph (kodama_matrix.cpp:1292-1330): a caller's graph in
// 0-based ids, `width` wide. One-based when index_base is 1, or when it is -1
// (auto) and every non-negative id lies in 1..n; per row the first `width`
// entries that name another row in range with a finite distance, that
// distance clamped at 0; -1 / +inf past them.
void normalize_graph(const int32_t* gi, const float* gd, int n, int given, int index_base, int width,
                     std::vector<int32_t>& idx, std::vector<float>& dist) {
  if (n < 2) throw std::invalid_argument("NeighborGraph samples must be at least 2.");
  if (given <= 0 || width <= 0) throw std::invalid_argument("NeighborGraph.neighbors must be positive.");
  bool any = false, in_one_based_range = true;
  for (size_t e = 0; e < static_cast<size_t>(n) * given; ++e) {
    if (gi[e] < 0) continue;
    any = true;
    if (gi[e] < 1 || gi[e] > n) in_one_based_range = false;
  }
  // No non-negative id at all: the reference's min/max stay at their sentinels
  // (INT_MAX >= 1 but INT_MIN <= n), so auto detection reads it as one-based.
  const bool one_based = index_base == 1 || (index_base < 0 && (any ? in_one_based_range : true));
  idx.assign(static_cast<size_t>(n) * width, -1);
  dist.assign(idx.size(), std::numeric_limits<float>::infinity());
  for (int i = 0; i < n; ++i) {
    int filled = 0;
    for (int j = 0; j < given && filled < width; ++j) {
      int id = gi[static_cast<size_t>(i) * given + j];
      const float d = gd[static_cast<size_t>(i) * given + j];
      if (one_based && id > 0) id -= 1;
      const bool keep = id >= 0 && id < n && id != i && std::isfinite(d);
      if (!keep) continue;
      idx[static_cast<size_t>(i) * width + filled] = id;
      dist[static_cast<size_t>(i) * width + filled] = d < 0.0f ? 0.0f : d;
      ++filled;
    }
  }
}

// normalize_constrain / normalize_fixed / is_identity_constrain
// (kodama_matrix.cpp:183-216): constraint values renumbered 1.. by first
// occurrence (none: every row its own group), fixed flags as 0/1.
IVec normalize_constrain(const int32_t* c, int n) {
  IVec out(static_cast<size_t>(n));
  if (c == nullptr) {
    std::iota(out.begin(), out.end(), 1);
    return out;
  }
  std::map<int, int> ids;
  for (int i = 0; i < n; ++i) {
    auto it = ids.find(c[i]);
    if (it == ids.end()) it = ids.emplace(c[i], static_cast<int>(ids.size()) + 1).first;
    out[i] = it->second;
  }
  return out;
}

bool is_identity(const IVec& c) {
  for (size_t i = 0; i < c.size(); ++i) if (c[i] != static_cast<int>(i + 1)) return false;
  return true;
}

// factor_subset (kodama_matrix.cpp:988-1000): values of `rows` renumbered 1..
// by first occurrence.
IVec factor_subset(const IVec& values, const IVec& rows) {
  std::map<int, int> ids;
  IVec out(rows.size());
  for (size_t i = 0; i < rows.size(); ++i) {
    const int v = values[rows[i]];
    auto it = ids.find(v);
    if (it == ids.end()) it = ids.emplace(v, static_cast<int>(ids.size()) + 1).first;
    out[i] = it->second;
  }
  return out;
}

// constrained_majority (kodama_matrix.cpp:1003-1011): every member of a
// constraint group takes the group's majority label (ties: smaller label).
IVec constrained_majority(const IVec& labels, const IVec& groups) {
  std::map<int, IVec> by;
  for (size_t i = 0; i < labels.size(); ++i) by[groups[i]].push_back(labels[i]);
  std::map<int, int> lab;
  for (const auto& kv : by) lab[kv.first] = majority_label(kv.second);
  IVec out(labels.size());
  for (size_t i = 0; i < labels.size(); ++i) out[i] = lab[groups[i]];
  return out;
}

// PLSLDAPredict_CPU (plscv.cpp:2976-3005): one fit on every training row (no
// Gram loading path), validation scores of the test rows; a fit without a
// component is the reference's DegeneratePLSFit.
template <class T>
IVec plslda_predict(const float* train, int nt, const IVec& labels, const float* test, int nq, int p,
                    const oracle_core_options* o) {
  if (nt == 0 || nq == 0 || p == 0) return IVec();
  const IVec classes = sorted_unique(labels);
  if (classes.size() == 1) return IVec(static_cast<size_t>(nq), classes[0]);
  std::vector<float> x(static_cast<size_t>(nt + nq) * p);
  std::copy_n(train, static_cast<size_t>(nt) * p, x.begin());
  std::copy_n(test, static_cast<size_t>(nq) * p, x.begin() + static_cast<size_t>(nt) * p);
  IVec lab(static_cast<size_t>(nt + nq), labels[0]), tr(static_cast<size_t>(nt)), te(static_cast<size_t>(nq));
  std::copy(labels.begin(), labels.end(), lab.begin());
  std::iota(tr.begin(), tr.end(), 0);
  std::iota(te.begin(), te.end(), nt);
  const FoldOut<T> fo = pls_fold<T>(x.data(), p, lab, tr, te, o, false);
  if (fo.weights.c == 0) throw std::runtime_error("PLSLDAPredict: degenerate PLS fit");
  return fo.pred;
}

// CoreKNNGraph's tables (core.cpp:624-692) of a run's landmark subgraph
// (subset_graph_to_rows, kodama_matrix.cpp:1332-1366): per landmark row the
// global neighbours that are landmarks, as run-local ids in graph order; the
// first k of them outside the row's fold vote, score -d^2 (Euclidean) or 1 - d.
KnnTables subgraph_tables(const int32_t* gi, const float* gd, int n, int width, const IVec& land,
                          const IVec& folds, int k_opt, int metric) {
  const int L = static_cast<int>(land.size());
  IVec local(static_cast<size_t>(n), -1);
  for (int i = 0; i < L; ++i) local[land[i]] = i;
  KnnTables t;
  t.folds = folds;
  t.width = std::max(1, std::min(k_opt, width));
  t.nbr.assign(static_cast<size_t>(L) * t.width, -1);
  t.score.assign(t.nbr.size(), -std::numeric_limits<float>::infinity());
  for (int i = 0; i < L; ++i) {
    int col = 0;
    for (int j = 0; j < width && col < t.width; ++j) {
      const size_t src = static_cast<size_t>(land[i]) * width + j;
      const int g = gi[src];
      if (g < 0 || g >= n) continue;
      const int nb = local[g];
      if (nb < 0 || nb == i || folds[nb] == folds[i]) continue;
      const float d0 = gd[src];
      if (!(std::isfinite(d0) && d0 >= 0.0f)) continue;
      t.nbr[static_cast<size_t>(i) * t.width + col] = nb;
      t.score[static_cast<size_t>(i) * t.width + col] = metric == 2 ? -(d0 * d0) : 1.0f - d0;
      ++col;
    }
  }
  return t;
}

// FNV-1a over the little-endian bytes of 32-bit values (the hash of the
// reference's run diagnostics, kodama_matrix.cpp:1689-1699).
uint64_t fnv1a(const IVec& v) {
  uint64_t h = 1469598103934665603ull;
  for (int x : v) {
    const uint32_t b = static_cast<uint32_t>(x);
    for (int s = 0; s < 32; s += 8) {
      h ^= static_cast<uint8_t>((b >> s) & 0xffu);
      h *= 1099511628211ull;
    }
  }
  return h;
}

// KODAMA.matrix without spatial data (kodama_matrix.cpp:3324-4051 with
// run_iteration 2870-3322), kNN or PLS-LDA classifier: the shared k-means atlas
// (mt19937_64(seed + 1)); per run r = 1..M an mt19937_64(seed + r) that skips
// one draw, quota-samples the landmarks and, without starting labels, draws the
// k-means starting labels (with them: the landmarks' starting labels, and the
// search is greedy); constraint groups take their majority label before the
// Core and after the projection. Core: kNN with per-fold exact neighbours, or
// with the caller's graph subset to the landmarks (CoreKNNGraph); PLS-LDA with
// `components` fixed components. Projection: kNN majority of the first k
// landmarks among a row's graph neighbours (clbest[0] when none), or
// PLSLDAPredict. diag (may be NULL) [runs x 3]: FNV-1a hashes of the landmark
// rows, the starting labels and the fold ids, as the reference's run
// diagnostics.
int matrix_impl(const float* x, int32_t n, int32_t p, const oracle_matrix_options* o, const int32_t* ext_idx,
                const float* ext_dist, int32_t ext_width, int32_t ext_base, const int32_t* starting_labels,
                const int32_t* constrain_in, const int32_t* fixed_in, int32_t* res, double* acc, double* v,
                int32_t* landmark_rows, int32_t* landmarks_used, uint64_t* diag) {
  return guarded([&] {
    if (n < 3) throw std::invalid_argument("KODAMAMatrix requires at least 3 rows.");
    if (o->runs < 1) throw std::invalid_argument("KODAMAMatrixOptions::runs must be positive.");
    if (o->cycles < 0) throw std::invalid_argument("KODAMAMatrixOptions::cycles must be non-negative.");
    if (o->folds < 2) throw std::invalid_argument("KODAMAMatrixOptions::folds must be at least two.");
    if (o->classifier != 0 && o->classifier != 1) throw std::invalid_argument("Unsupported KODAMA.matrix classifier.");
    const bool knn = o->classifier == 1;
    int L = o->landmarks > 0 ? o->landmarks : 10000;
    if (L >= n) L = static_cast<int>(std::ceil(static_cast<double>(n) * 0.75));
    L = std::max(2, std::min(L, n - 1));
    const int split = o->splitting > 0 ? o->splitting : (n < 40000 ? 100 : 300);
    const int coarse_k = std::max(2, std::min(split, n));
    const int requested = o->graph_neighbors > 0 ? o->graph_neighbors : 100;
    const int resolved = std::max(1, static_cast<int>(std::floor(std::min(
      {static_cast<double>(L), static_cast<double>(n) * 0.75 - 1.0, static_cast<double>(requested)}))));
    const int neighbors = ext_idx != nullptr ? std::max(1, std::min(ext_width, resolved)) : resolved;
    const int comps = std::max(1, std::min(o->components, std::min(n, p)));
    const IVec cons_all = normalize_constrain(constrain_in, n);
    const bool identity = is_identity(cons_all);
    IVec fixed_all(static_cast<size_t>(n), 0);
    if (fixed_in) for (int i = 0; i < n; ++i) fixed_all[i] = fixed_in[i] != 0 ? 1 : 0;
    std::mt19937_64 shared(o->seed + 1u);
    const IVec atlas = kmeans_labels(x, n, p, coarse_k, shared, 10);
    IVec off, rows;
    index_strata(atlas, coarse_k, off, rows);
    // The global graph serves the kNN classifier only (the caller's graph in
    // KODAMAMatrixFromGraphData, else the exact self-kNN graph).
    std::vector<int32_t> gi;
    std::vector<float> gd;
    if (ext_idx != nullptr) {
      normalize_graph(ext_idx, ext_dist, n, ext_width, ext_base, neighbors, gi, gd);
    } else if (knn) {
      const std::vector<float> xs = prepare_rows(x, n, p, o->metric, 1);
      gi.resize(static_cast<size_t>(n) * neighbors);
      gd.resize(gi.size());
      IVec self(static_cast<size_t>(n));
      std::iota(self.begin(), self.end(), 0);
      exact_knn(xs.data(), n, xs.data(), n, p, neighbors, o->metric, self.data(), gi.data(), gd.data());
    }
    *landmarks_used = L;
    for (int r = 1; r <= o->runs; ++r) {
      std::mt19937_64 rng(o->seed + static_cast<std::uint64_t>(r));
      (void)rng();
      const IVec land = quota_sample(off, rows, L, rng);
      std::vector<float> xl(static_cast<size_t>(L) * p);
      for (int i = 0; i < L; ++i) std::copy_n(x + static_cast<size_t>(land[i]) * p, p, xl.data() + static_cast<size_t>(i) * p);
      IVec cons(static_cast<size_t>(L));
      if (identity) std::iota(cons.begin(), cons.end(), 1);
      else cons = factor_subset(cons_all, land);
      IVec fx(static_cast<size_t>(L));
      for (int i = 0; i < L; ++i) fx[i] = fixed_all[land[i]];
      IVec xw(static_cast<size_t>(L));
      if (starting_labels != nullptr) {
        for (int i = 0; i < L; ++i) xw[i] = starting_labels[land[i]];
      } else {
        xw = kmeans_labels(xl.data(), L, p, std::max(2, std::min(split, L)), rng, 10);
      }
      if (!identity) xw = constrained_majority(xw, cons);
      oracle_core_options co{};
      co.cycles = o->cycles;
      co.classifier = o->classifier;
      co.seed = o->seed + static_cast<std::uint64_t>(r);
      co.evolutionary_search = starting_labels == nullptr ? 1 : 0;
      co.guarded_diversity = 1;
      co.many_to_one_absorption = 1;
      co.auto_class_coarsening = knn ? 0 : 1;
      co.adaptive_proposal_size = 1;
      co.prediction_guidance = co.policy_adaptive_proposal_size = co.transition_proposal = 1;
      co.stochastic_acceptance = co.diversity_multiplier = co.pls_transition_coarsening = 1;
      co.pls_fragmentation_penalty = co.reject_single_class = 1;
      co.folds = o->folds;
      co.stratified = 0;
      co.cv_seed = o->seed + static_cast<std::uint64_t>(r);
      co.k = std::max(1, o->k);
      co.metric = o->metric;
      co.max_components = comps;
      co.fixed_components = comps;
      co.center = 1;
      co.scale = 1;
      co.pls_fold_cache = 1;
      const IVec folds = make_folds(xw, cons, co.folds, false, co.cv_seed);
      Flags f = flags_of(&co);
      f.knn = knn;
      const int T = std::max(0, o->cycles);
      IVec clbest(static_cast<size_t>(L)), dirty(static_cast<size_t>(L)), cvp(static_cast<size_t>(L)),
        psize(static_cast<size_t>(T) + 1), active(static_cast<size_t>(T) + 1);
      std::vector<double> vacc(static_cast<size_t>(T) + 1), vscore(static_cast<size_t>(T) + 1);
      std::vector<uint8_t> a1(static_cast<size_t>(T) + 1), a2(static_cast<size_t>(T) + 1), a3(static_cast<size_t>(T) + 1);
      int32_t counters[ORACLE_N_COUNTERS];
      oracle_core_result cr{};
      cr.clbest = clbest.data();
      cr.clbest_dirty = dirty.data();
      cr.cvpredbest = cvp.data();
      cr.vect_acc = vacc.data();
      cr.vect_score = vscore.data();
      cr.proposal_size = psize.data();
      cr.active_classes = active.data();
      cr.accepted = a1.data();
      cr.improving_acceptance = a2.data();
      cr.temperature_acceptance = a3.data();
      cr.counters = counters;
      if (knn) {
        const KnnTables t = ext_idx != nullptr
          ? subgraph_tables(gi.data(), gd.data(), n, neighbors, land, folds, co.k, co.metric)
          : precompute_tables(xl.data(), L, p, folds, co.k, co.metric);
        evolve(L, xw, cons, fx, f, [&](const IVec& l) { return vote_predict(t, l); }, &cr);
      } else {
        const bool f32 = o->precision == 32;
        evolve(L, xw, cons, fx, f, [&](const IVec& l) {
          return f32 ? plsldacv<float>(xl.data(), L, p, l, cons, &co, &folds, nullptr, nullptr)
                     : plsldacv<double>(xl.data(), L, p, l, cons, &co, &folds, nullptr, nullptr);
        }, &cr);
      }
      int32_t* out = res + static_cast<size_t>(r - 1) * n;
      IVec local(static_cast<size_t>(n), -1);
      for (int i = 0; i < L; ++i) local[land[i]] = i;
      IVec tpoints;
      for (int i = 0; i < n; ++i) {
        if (local[i] >= 0) out[i] = clbest[local[i]];
        else tpoints.push_back(i);
      }
      if (knn) {
        for (int i : tpoints) {
          IVec votes;
          for (int j = 0; j < neighbors && static_cast<int>(votes.size()) < co.k; ++j) {
            const int g = gi[static_cast<size_t>(i) * neighbors + j];
            if (g < 0 || g >= n) continue;
            if (local[g] >= 0) votes.push_back(clbest[local[g]]);
          }
          out[i] = votes.empty() ? clbest[0] : majority_label(votes);
        }
      } else if (!tpoints.empty()) {
        std::vector<float> xt(tpoints.size() * static_cast<size_t>(p));
        for (size_t i = 0; i < tpoints.size(); ++i)
          std::copy_n(x + static_cast<size_t>(tpoints[i]) * p, p, xt.data() + i * p);
        const int nt = static_cast<int>(tpoints.size());
        const IVec pr = o->precision == 32 ? plslda_predict<float>(xl.data(), L, clbest, xt.data(), nt, p, &co)
                                           : plslda_predict<double>(xl.data(), L, clbest, xt.data(), nt, p, &co);
        for (int i = 0; i < nt; ++i) out[tpoints[i]] = pr[i];
      }
      if (!identity) {
        const IVec all(out, out + n);
        const IVec cm = constrained_majority(all, cons_all);
        std::copy(cm.begin(), cm.end(), out);
      }
      acc[r - 1] = cr.accbest;
      for (int c = 0; c < T; ++c) v[static_cast<size_t>(r - 1) * T + c] = vacc[c];
      if (landmark_rows) std::memcpy(landmark_rows + static_cast<size_t>(r - 1) * L, land.data(), sizeof(int32_t) * L);
      if (diag) {
        diag[static_cast<size_t>(r - 1) * 3 + 0] = fnv1a(land);
        diag[static_cast<size_t>(r - 1) * 3 + 1] = fnv1a(xw);
        diag[static_cast<size_t>(r - 1) * 3 + 2] = fnv1a(folds);
      }
    }
  });
}

}  // namespace

extern "C" {

int oracle_kodama_matrix(const float* x, int32_t n, int32_t p, const oracle_matrix_options* o,
                         const int32_t* starting_labels, const int32_t* constrain, const int32_t* fixed,
                         int32_t* res, double* acc, double* v, int32_t* landmark_rows, int32_t* landmarks_used,
                         uint64_t* diag) {
  return matrix_impl(x, n, p, o, nullptr, nullptr, 0, -1, starting_labels, constrain, fixed, res, acc, v,
                     landmark_rows, landmarks_used, diag);
}

// KODAMAMatrixFromGraphData (kodama_matrix.cpp:4656-4670, 3526-3556): the
// caller's graph, normalised to min(its width, the resolved graph width),
// replaces the exact global graph, in the Core step as in the projection.
int oracle_kodama_matrix_from_graph(const float* x, int32_t n, int32_t p, const oracle_matrix_options* o,
                                    const int32_t* graph_indices, const float* graph_distances,
                                    int32_t graph_neighbors, int32_t index_base, const int32_t* starting_labels,
                                    const int32_t* constrain, const int32_t* fixed, int32_t* res, double* acc,
                                    double* v, int32_t* landmark_rows, int32_t* landmarks_used, uint64_t* diag) {
  if (graph_indices == nullptr || graph_distances == nullptr || graph_neighbors <= 0) {
    return guarded([] { throw std::invalid_argument("null or empty graph."); });
  }
  return matrix_impl(x, n, p, o, graph_indices, graph_distances, graph_neighbors, index_base, starting_labels,
                     constrain, fixed, res, acc, v, landmark_rows, landmarks_used, diag);
}

int oracle_normalize_graph(const int32_t* graph_indices, const float* graph_distances, int32_t n,
                           int32_t graph_neighbors, int32_t index_base, int32_t width, int32_t* indices_out,
                           float* distances_out) {
  return guarded([&] {
    std::vector<int32_t> gi;
    std::vector<float> gd;
    normalize_graph(graph_indices, graph_distances, n, graph_neighbors, index_base, width, gi, gd);
    std::copy(gi.begin(), gi.end(), indices_out);
    std::copy(gd.begin(), gd.end(), distances_out);
  });
}

}  // extern "C"
