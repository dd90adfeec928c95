// Host engine of the HIP backend (see engine.h).
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <thread>
#include <unordered_map>

#include <omp.h>

#include "capi_common.h"
#include "knn_mfma.h"
#include "launch.h"

namespace kodama_hip {

// ---------------------------------------------------------------------------
// Folds (common.cpp:108-214). The same generator, the same std::shuffle calls
// and the same dealing rules as the reference, so fold ids are identical.
// ---------------------------------------------------------------------------
IVec make_folds(const IVec& labels, const IVec& constrain, int folds, bool stratified, std::uint64_t seed) {
  if (folds < 2) throw InvalidArgument("Number of folds must be at least 2.");
  const int n = static_cast<int>(labels.size());
  std::mt19937_64 rng(seed);
  IVec out(static_cast<size_t>(n), 0);
  auto least_loaded = [](const IVec& load) {
    return static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
  };
  if (constrain.empty()) {
    if (!stratified) {
      IVec perm(static_cast<size_t>(n));
      std::iota(perm.begin(), perm.end(), 0);
      std::shuffle(perm.begin(), perm.end(), rng);
      for (int i = 0; i < n; ++i) out[perm[i]] = i % folds;
      return out;
    }
    std::map<int, IVec> by_label;
    for (int i = 0; i < n; ++i) by_label[labels[i]].push_back(i);
    IVec load(static_cast<size_t>(folds));
    for (auto& kv : by_label) {
      std::shuffle(kv.second.begin(), kv.second.end(), rng);
      std::fill(load.begin(), load.end(), 0);
      for (int r : kv.second) {
        const int f = least_loaded(load);
        out[r] = f;
        ++load[f];
      }
    }
    return out;
  }
  // Constrained rows move as groups, each group carrying its majority label.
  std::map<int, IVec> members;
  for (int i = 0; i < n; ++i) members[constrain[i]].push_back(i);
  IVec group_label, group_size;
  for (const auto& kv : members) {
    std::map<int, int> tally;
    for (int r : kv.second) ++tally[labels[r]];
    auto best = tally.begin();
    for (auto it = tally.begin(); it != tally.end(); ++it) if (it->second > best->second) best = it;
    group_label.push_back(best->first);
    group_size.push_back(static_cast<int>(kv.second.size()));
  }
  const size_t G = group_label.size();
  IVec group_fold(G, 0);
  if (stratified) {
    std::map<int, std::vector<size_t>> by_label;
    for (size_t g = 0; g < G; ++g) by_label[group_label[g]].push_back(g);
    IVec load(static_cast<size_t>(folds));
    for (auto& kv : by_label) {
      std::shuffle(kv.second.begin(), kv.second.end(), rng);
      std::fill(load.begin(), load.end(), 0);
      for (size_t g : kv.second) {
        const int f = least_loaded(load);
        group_fold[g] = f;
        load[f] += group_size[g];
      }
    }
  } else {
    std::vector<size_t> perm(G);
    std::iota(perm.begin(), perm.end(), 0);
    std::shuffle(perm.begin(), perm.end(), rng);
    for (size_t i = 0; i < G; ++i) group_fold[perm[i]] = static_cast<int>(i % folds);
  }
  size_t g = 0;
  for (const auto& kv : members) {
    for (int r : kv.second) out[r] = group_fold[g];
    ++g;
  }
  return out;
}

std::vector<float> prepare_rows(const float* x, int n, int p, int metric, int variant) {
  std::vector<float> out(x, x + static_cast<size_t>(n) * p);
  if (metric != KODAMA_HIP_COSINE) return out;
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

CoreFlags CoreFlags::from(const kodama_hip_core_options& o) {
  CoreFlags f;
  f.cycles = o.cycles;
  f.knn = o.classifier == 1;
  f.seed = o.seed;
  f.shake = o.shake != 0;
  f.evolutionary = o.evolutionary_search != 0;
  f.guarded = o.guarded_diversity != 0;
  f.absorption = o.many_to_one_absorption != 0;
  f.coarsening = o.auto_class_coarsening != 0;
  f.adaptive = o.adaptive_proposal_size != 0;
  f.guidance = o.prediction_guidance != 0;
  f.policy_adaptive = o.policy_adaptive_proposal_size != 0;
  f.transition = o.transition_proposal != 0;
  f.stochastic = o.stochastic_acceptance != 0;
  f.diversity = o.diversity_multiplier != 0;
  f.pls_coarsening = o.pls_transition_coarsening != 0;
  f.fragmentation = o.pls_fragmentation_penalty != 0;
  f.reject_single = o.reject_single_class != 0;
  return f;
}

void encode_labels(const int* labels, int n, IVec& values, IVec& codes) {
  values.assign(labels, labels + n);
  std::sort(values.begin(), values.end());
  values.erase(std::unique(values.begin(), values.end()), values.end());
  codes.resize(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    codes[i] = static_cast<int>(std::lower_bound(values.begin(), values.end(), labels[i]) - values.begin());
  }
}

double objective_from_counts(const int* counts, const int* firsts, const IVec& label_values, int n,
                             double acc, const CoreFlags& f) {
  const bool diversity = f.guarded && f.diversity;
  const bool fragmentation = f.coarsening && !f.knn && f.fragmentation;
  const bool reject_single = f.guarded && f.reject_single;
  if (!diversity && !fragmentation && !reject_single) return acc;
  if (n == 0) return acc;
  const int C = static_cast<int>(label_values.size());
  IVec present;
  for (int c = 0; c < C; ++c) if (counts[c] > 0) present.push_back(c);
  std::sort(present.begin(), present.end(), [&](int a, int b) { return firsts[a] < firsts[b]; });
  std::unordered_map<int, int> tally;
  tally.reserve(static_cast<size_t>(n));
  for (int c : present) tally.emplace(label_values[c], counts[c]);
  const int k = static_cast<int>(tally.size());
  if (k < 1) return acc;
  if (reject_single && k == 1) return -std::numeric_limits<double>::infinity();
  const double nn = static_cast<double>(n);
  double score = acc;
  if (diversity) {
    double same = 0.0;
    for (const auto& kv : tally) {
      const double q = static_cast<double>(kv.second) / nn;
      same += q * q;
    }
    score *= std::sqrt(std::max(0.0, 1.0 - same));
  }
  if (fragmentation && k > 1) {
    double h = 0.0;
    for (const auto& kv : tally) {
      const double q = static_cast<double>(kv.second) / nn;
      if (q > 0.0) h -= q * std::log(q);
    }
    const double hmax = nn > 1.0 ? std::log(nn) : 0.0;
    const double code_cost = hmax > 0.0 ? std::clamp(h / hmax, 0.0, 1.0) : 0.0;
    const double keff = std::exp(h);
    const double frag = std::max(0.0, std::log(static_cast<double>(k) / std::max(1.0, keff)));
    score -= (1.0 - acc) * std::max(code_cost, frag / (1.0 + frag));
  }
  return score;
}

namespace {

// sample_proposal_group_count (core.cpp:1228-1254).
int proposal_count(int groups, int cycle, int cycles, bool adaptive, std::mt19937_64& rng) {
  if (groups <= 1) return std::max(0, groups);
  if (!adaptive || cycles <= 1) {
    std::uniform_int_distribution<int> pick(1, groups);
    return pick(rng);
  }
  const double t = std::clamp(static_cast<double>(cycle + 1) / static_cast<double>(cycles + 1), 0.0, 1.0);
  const double smooth = t * t * (3.0 - 2.0 * t);
  const double temperature = std::clamp(1.0 - smooth, 0.0, 1.0);
  const int hi = std::max(1, 1 + static_cast<int>(std::floor(static_cast<double>(groups - 1) * temperature)));
  std::uniform_int_distribution<int> pick(1, hi);
  return pick(rng);
}

enum Counter {
  kCompleted, kProposals, kBest, kCurAcc, kStochAtt, kStochAcc, kCurRej, kCoarse, kAbsorb,
  kTransAtt, kTransAcc, kM2OAtt, kM2OAcc, kPlsAtt, kPlsAcc, kCvEval, kSuccess
};

int count_present(const IVec& counts) {
  int k = 0;
  for (int v : counts) k += v > 0 ? 1 : 0;
  return k;
}

}  // namespace

RunEvolution::RunEvolution(IVec label_values, IVec initial_codes, const IVec& constrain, IVec fixed,
                           const CoreFlags& f)
    : f_(f), n_(static_cast<int>(initial_codes.size())), label_values_(std::move(label_values)),
      fixed_(std::move(fixed)), rng_(f.seed) {
  if (f_.cycles < 0) throw InvalidArgument("Core cycles must be non-negative.");
  if (fixed_.empty()) fixed_.assign(static_cast<size_t>(n_), 0);
  if (static_cast<int>(fixed_.size()) != n_) throw InvalidArgument("fixed size must be zero or match number of rows.");
  singletons_ = constrain.empty();
  if (singletons_) {
    groups_ = n_;
  } else {
    std::map<int, IVec> by_key;
    for (int i = 0; i < n_; ++i) by_key[constrain[i]].push_back(i);
    for (auto& kv : by_key) members_.push_back(std::move(kv.second));
    groups_ = static_cast<int>(members_.size());
  }
  order_.resize(static_cast<size_t>(groups_));
  best_ = std::move(initial_codes);
  best_dirty_ = best_;
  const int T = f_.cycles;
  vacc_.assign(T, -1.0);
  vscore_.assign(T, -1.0);
  psize_.assign(T, 0);
  active_.assign(T, 0);
  accepted_.assign(T, 0);
  improving_.assign(T, 0);
  temperature_.assign(T, 0);
  const int C = classes();
  sizes_.assign(C, 0);
  movable_.assign(C, 0);
  tmp_.assign(C, 0);
  bucket_start_.assign(C + 1, 0);
}

void RunEvolution::begin(const int* pred, const Tally& t) {
  const int C = classes();
  bestpred_.assign(pred, pred + n_);
  counters_[kCvEval] = 1;
  best_counts_.assign(t.counts, t.counts + C);
  accbest_ = f_.shake ? 0.0 : (n_ ? static_cast<double>(t.matches) / static_cast<double>(n_) : 0.0);
  scorebest_ = f_.shake ? -std::numeric_limits<double>::infinity()
                        : objective_from_counts(t.counts, t.firsts, label_values_, n_, accbest_, f_);
  cur_ = best_;
  curpred_ = bestpred_;
  cur_counts_ = best_counts_;
  cur_acc_ = accbest_;
  cur_score_ = scorebest_;
}

// Transition counts label -> previous prediction, grouped by label in ascending
// order with destinations ascending: the iteration order of the reference's
// std::map<int, std::map<int, int>> (core.cpp:974-993).
void RunEvolution::build_flows(const IVec& cl, const IVec& guide) {
  const int C = classes();
  std::fill(sizes_.begin(), sizes_.end(), 0);
  std::fill(movable_.begin(), movable_.end(), 0);
  for (int i = 0; i < n_; ++i) {
    ++sizes_[cl[i]];
    if (fixed_[i] != 1) ++movable_[cl[i]];
  }
  // Bucket rows by label (counting sort), then tally predictions per bucket.
  bucket_start_[0] = 0;
  for (int c = 0; c < C; ++c) bucket_start_[c + 1] = bucket_start_[c] + sizes_[c];
  bucket_.resize(static_cast<size_t>(n_));
  {
    IVec fill(bucket_start_.begin(), bucket_start_.end() - 1);
    for (int i = 0; i < n_; ++i) bucket_[fill[cl[i]]++] = i;
  }
  present_.clear();
  flow_start_.clear();
  flow_dest_.clear();
  flow_count_.clear();
  for (int c = 0; c < C; ++c) {
    if (sizes_[c] == 0) continue;
    present_.push_back(c);
    flow_start_.push_back(static_cast<int>(flow_dest_.size()));
    touched_.clear();
    for (int b = bucket_start_[c]; b < bucket_start_[c + 1]; ++b) {
      const int d = guide[bucket_[b]];
      if (tmp_[d]++ == 0) touched_.push_back(d);
    }
    std::sort(touched_.begin(), touched_.end());
    for (int d : touched_) {
      flow_dest_.push_back(d);
      flow_count_.push_back(tmp_[d]);
      tmp_[d] = 0;
    }
  }
  flow_start_.push_back(static_cast<int>(flow_dest_.size()));
}

// propose_auto_class_coarsening (core.cpp:995-1108).
bool RunEvolution::coarsen(IVec& cl) {
  const int classes_now = static_cast<int>(present_.size());
  if (classes_now <= 2) return false;
  const double n = static_cast<double>(n_);
  double h = 0.0;
  for (int c : present_) {
    const double q = static_cast<double>(sizes_[c]) / n;
    if (q > 0.0) h -= q * std::log(q);
  }
  const double keff = std::exp(h);
  const double frag = std::max(0.0, std::log(static_cast<double>(classes_now) / std::max(1.0, keff)));
  double flow_entropy = 0.0;
  struct Cand { int source; int dest; double score; };
  std::vector<Cand> cands;
  for (size_t s = 0; s < present_.size(); ++s) {
    const int src = present_[s];
    const int src_n = sizes_[src];
    const int movable = movable_[src];
    if (movable <= 0) continue;
    int dst = src, dst_n = src_n, best = 0, stay_n = 0;
    double row_h = 0.0;
    for (int e = flow_start_[s]; e < flow_start_[s + 1]; ++e) {
      const int d = flow_dest_[e];
      const int cnt = flow_count_[e];
      if (d == src) stay_n = cnt;
      const double q = static_cast<double>(cnt) / static_cast<double>(src_n);
      if (q > 0.0) row_h -= q * std::log(q);
      const int dn = sizes_[d];
      if (d != src && dn > 0 &&
          (cnt > best || (cnt == best && dn > dst_n) || (cnt == best && dn == dst_n && d < dst))) {
        dst = d; dst_n = dn; best = cnt;
      }
    }
    flow_entropy += (static_cast<double>(src_n) / n) * row_h;
    if (dst == src || best == 0) continue;
    const double stay = static_cast<double>(stay_n) / static_cast<double>(src_n);
    cands.push_back({src, dst, (1.0 / std::sqrt(static_cast<double>(movable))) * (1.0 - stay)});
  }
  if (cands.empty()) return false;
  const double temp = frag / (frag + flow_entropy + 1.0);
  const double redundant = std::max(0.0, static_cast<double>(classes_now) - keff);
  int budget = static_cast<int>(std::ceil(temp * redundant));
  if (budget < 1 && frag > 0.0) budget = 1;
  budget = std::min(budget, classes_now - 2);
  if (budget <= 0) return false;
  std::shuffle(cands.begin(), cands.end(), rng_);
  std::stable_sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.score > b.score; });
  IVec remap(static_cast<size_t>(classes()));
  std::iota(remap.begin(), remap.end(), 0);
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
  for (int i = 0; i < n_; ++i) {
    if (fixed_[i] == 1) continue;
    int d = cl[i];
    while (remap[d] != d) d = remap[d];
    if (d != cl[i]) { cl[i] = d; changed = true; }
  }
  return changed;
}

// propose_many_to_one_absorption (core.cpp:1110-1226).
bool RunEvolution::absorb(IVec& cl) {
  const int classes_now = static_cast<int>(present_.size());
  if (classes_now <= 2) return false;
  const double n = static_cast<double>(n_);
  struct Src { int source; double surplus; };
  struct Target { int target; double weight; std::vector<Src> sources; };
  std::map<int, Target> by_target;
  for (size_t s = 0; s < present_.size(); ++s) {
    const int src = present_[s];
    const int src_n = sizes_[src];
    if (movable_[src] <= 0) continue;
    int stay = 0;
    for (int e = flow_start_[s]; e < flow_start_[s + 1]; ++e) if (flow_dest_[e] == src) stay = flow_count_[e];
    int tgt = src, best = 0, tgt_n = src_n;
    double surplus_best = 0.0;
    for (int e = flow_start_[s]; e < flow_start_[s + 1]; ++e) {
      const int d = flow_dest_[e];
      if (d == src) continue;
      const int dn = sizes_[d];
      if (dn == 0) continue;
      const int cnt = flow_count_[e];
      const double expected = static_cast<double>(src_n) * static_cast<double>(dn) / std::max(1.0, n);
      const double surplus = static_cast<double>(cnt) - expected;
      if (surplus > surplus_best || (surplus == surplus_best && cnt > best) ||
          (surplus == surplus_best && cnt == best && dn > tgt_n) ||
          (surplus == surplus_best && cnt == best && dn == tgt_n && d < tgt)) {
        tgt = d; tgt_n = dn; best = cnt; surplus_best = surplus;
      }
    }
    if (tgt == src || surplus_best <= 0.0 || best <= stay) continue;
    auto it = by_target.find(tgt);
    if (it == by_target.end()) it = by_target.emplace(tgt, Target{tgt, 0.0, {}}).first;
    it->second.weight += surplus_best;
    it->second.sources.push_back({src, surplus_best});
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
  std::vector<double> weights;
  weights.reserve(moves.size());
  for (const Target& t : moves) weights.push_back(std::max(t.weight, std::numeric_limits<double>::min()));
  std::discrete_distribution<int> pick(weights.begin(), weights.end());
  const Target& chosen = moves[static_cast<size_t>(pick(rng_))];
  // Every source has movable rows, so the remap changes the labelling; a
  // source class disappears when none of its rows is fixed.
  IVec remap(static_cast<size_t>(classes()), -1);
  int vanished = 0;
  bool any = false;
  for (const Src& s : chosen.sources) {
    if (s.source == chosen.target) continue;
    remap[s.source] = chosen.target;
    any = true;
    if (sizes_[s.source] == movable_[s.source]) ++vanished;
  }
  if (!any) return false;
  const int after = classes_now - vanished;
  if (after <= 1 || after >= classes_now) return false;
  for (int i = 0; i < n_; ++i) {
    if (fixed_[i] == 1) continue;
    const int r = remap[cl[i]];
    if (r >= 0) cl[i] = r;
  }
  return true;
}

const IVec& RunEvolution::propose(int cycle) {
  const IVec& base = f_.evolutionary ? cur_ : best_;
  const IVec& guide = f_.evolutionary ? curpred_ : bestpred_;
  cl_ = base;
  cl_dirty_ = base;
  if (groups_ > 0) {
    const int take = proposal_count(groups_, cycle, f_.cycles, f_.adaptive && f_.policy_adaptive, rng_);
    psize_[cycle] = take;
    std::iota(order_.begin(), order_.end(), 0);
    std::shuffle(order_.begin(), order_.end(), rng_);
    IVec emp_codes, emp_counts;
    if (!f_.guidance) {
      IVec tally(static_cast<size_t>(classes()), 0);
      for (int c : cl_) ++tally[c];
      for (int c = 0; c < classes(); ++c) if (tally[c] > 0) { emp_codes.push_back(c); emp_counts.push_back(tally[c]); }
    }
    IVec elig, cand, cnt;
    for (int t = 0; t < take; ++t) {
      const int g = order_[t];
      if (singletons_ && f_.guidance) {
        // One candidate label: the distribution is degenerate and draws nothing.
        if (fixed_[g] != 1) cl_[g] = guide[g];
        continue;
      }
      elig.clear();
      if (singletons_) {
        if (fixed_[g] != 1) elig.push_back(g);
      } else {
        for (int r : members_[g]) if (fixed_[r] != 1) elig.push_back(r);
      }
      if (elig.empty()) continue;
      if (f_.guidance) {
        cand.clear();
        for (int r : elig) cand.push_back(guide[r]);
        std::sort(cand.begin(), cand.end());
        cnt.clear();
        size_t w = 0;
        for (size_t i = 0; i < cand.size();) {
          size_t j = i + 1;
          while (j < cand.size() && cand[j] == cand[i]) ++j;
          cnt.push_back(static_cast<int>(j - i));
          cand[w++] = cand[i];
          i = j;
        }
        cand.resize(w);
        std::discrete_distribution<int> pick(cnt.begin(), cnt.end());
        const int repl = cand[static_cast<size_t>(pick(rng_))];
        for (int r : elig) cl_[r] = repl;
      } else {
        if (emp_codes.empty()) continue;
        std::discrete_distribution<int> pick(emp_counts.begin(), emp_counts.end());
        const int repl = emp_codes[static_cast<size_t>(pick(rng_))];
        for (int r : elig) cl_[r] = repl;
      }
    }
  }
  coarsened_ = false;
  absorbed_ = false;
  const bool try_coarsen = f_.transition && f_.coarsening && f_.pls_coarsening;
  const bool try_absorb = f_.transition && f_.absorption;
  if (try_coarsen || try_absorb) {
    ++counters_[kTransAtt];
    build_flows(cl_, guide);
  }
  if (try_coarsen) {
    ++counters_[kPlsAtt];
    coarsened_ = coarsen(cl_);
    if (coarsened_) ++counters_[kCoarse];
  }
  if (try_absorb) {
    ++counters_[kM2OAtt];
    if (coarsened_) build_flows(cl_, guide);
    absorbed_ = absorb(cl_);
    if (absorbed_) ++counters_[kAbsorb];
  }
  return cl_;
}

void RunEvolution::finish(int cycle, const int* pred, const Tally& t) {
  const int C = classes();
  ++counters_[kCvEval];
  ++counters_[kProposals];
  const double acc = n_ ? static_cast<double>(t.matches) / static_cast<double>(n_) : 0.0;
  const double score = objective_from_counts(t.counts, t.firsts, label_values_, n_, acc, f_);
  const bool improves = score > scorebest_;
  if (improves) {
    bestpred_.assign(pred, pred + n_);
    best_ = cl_;
    best_dirty_ = cl_dirty_;
    best_counts_.assign(t.counts, t.counts + C);
    accbest_ = acc;
    scorebest_ = score;
    ++counters_[kBest];
  }
  bool accepted = improves, improving = improves, stochastic = false;
  if (f_.evolutionary) {
    const double cooling = 1.0 - static_cast<double>(cycle + 1) / static_cast<double>(std::max(1, f_.cycles));
    const double temperature = std::max(1.0e-9, 0.10 * std::max(0.0, 1.0 - cur_acc_) * cooling);
    bool take = score >= cur_score_;
    improving = take;
    if (!take && f_.stochastic && temperature > 1.0e-9) {
      ++counters_[kStochAtt];
      std::uniform_real_distribution<double> u(0.0, 1.0);
      take = u(rng_) < std::exp((score - cur_score_) / temperature);
      stochastic = take;
    }
    if (take) {
      cur_.swap(cl_);
      curpred_.assign(pred, pred + n_);
      cur_counts_.assign(t.counts, t.counts + C);
      cur_acc_ = acc;
      cur_score_ = score;
      ++counters_[kCurAcc];
      if (stochastic) ++counters_[kStochAcc];
    } else {
      ++counters_[kCurRej];
    }
    accepted = take;
  }
  accepted_[cycle] = accepted ? 1 : 0;
  improving_[cycle] = accepted && improving ? 1 : 0;
  temperature_[cycle] = stochastic ? 1 : 0;
  if (accepted && (coarsened_ || absorbed_)) {
    ++counters_[kTransAcc];
    if (coarsened_) ++counters_[kPlsAcc];
    if (absorbed_) ++counters_[kM2OAcc];
  }
  active_[cycle] = count_present(f_.evolutionary ? cur_counts_ : best_counts_);
  vacc_[cycle] = accbest_;
  vscore_[cycle] = scorebest_;
  counters_[kCompleted] = cycle + 1;
  if (acc == 1.0 && std::isfinite(score)) success_ = true;
}

void RunEvolution::export_result(kodama_hip_core_result* out) const {
  for (int i = 0; i < n_; ++i) {
    out->clbest[i] = label_values_[best_[i]];
    out->clbest_dirty[i] = label_values_[best_dirty_[i]];
    out->cvpredbest[i] = label_values_[bestpred_[i]];
  }
  for (int c = 0; c < f_.cycles; ++c) {
    out->vect_acc[c] = vacc_[c];
    out->vect_score[c] = vscore_[c];
    out->proposal_size[c] = psize_[c];
    out->active_classes[c] = active_[c];
    out->accepted[c] = accepted_[c];
    out->improving_acceptance[c] = improving_[c];
    out->temperature_acceptance[c] = temperature_[c];
  }
  int counters[KODAMA_HIP_N_COUNTERS];
  std::memcpy(counters, counters_, sizeof(counters));
  counters[kSuccess] = success_ ? 1 : 0;
  std::memcpy(out->counters, counters, sizeof(counters));
  out->accbest = accbest_;
  out->scorebest = scorebest_;
}

// ---------------------------------------------------------------------------
// Device kNN tables and batched vote.
// ---------------------------------------------------------------------------
void KnnBatch::allocate(const IVec& run_offsets) {
  const hipStream_t s = stream_.get();
  pad_offsets_.assign(static_cast<size_t>(runs_) + 1, 0);
  IVec block_run;
  for (int r = 0; r < runs_; ++r) {
    const int L = run_offsets[r + 1] - run_offsets[r];
    if (L < 1) throw InvalidArgument("every run needs at least one row.");
    max_rows_ = std::max(max_rows_, L);
    const int padded = (L + 255) / 256 * 256;
    pad_offsets_[r + 1] = pad_offsets_[r] + padded;
    for (int b = 0; b < padded / 256; ++b) block_run.push_back(r);
  }
  pad_total_ = pad_offsets_[runs_];
  blocks_ = static_cast<int>(block_run.size());
  d_run_off_.resize(run_offsets.size());
  d_run_off_.upload(run_offsets.data(), run_offsets.size(), s);
  d_pad_off_.resize(pad_offsets_.size());
  d_pad_off_.upload(pad_offsets_.data(), pad_offsets_.size(), s);
  d_block_run_.resize(block_run.size());
  d_block_run_.upload(block_run.data(), block_run.size(), s);
  d_nbr_.resize(static_cast<size_t>(k_) * pad_total_);
  d_score_.resize(static_cast<size_t>(k_) * pad_total_);
  d_labels_.resize(static_cast<size_t>(pad_total_));
  d_pred_.resize(static_cast<size_t>(pad_total_));
  d_counts_.resize(static_cast<size_t>(runs_) * classes_);
  d_firsts_.resize(static_cast<size_t>(runs_) * classes_);
  d_matches_.resize(static_cast<size_t>(runs_));
  h_labels_.resize(static_cast<size_t>(pad_total_));
  h_pred_.resize(static_cast<size_t>(pad_total_));
  h_counts_.resize(static_cast<size_t>(runs_) * classes_);
  h_firsts_.resize(static_cast<size_t>(runs_) * classes_);
  h_matches_.resize(static_cast<size_t>(runs_));
  // Device-evolution staging: 16-bit labels per padded row plus per-run words
  // (keys, widths, state, the elapsed ticks of up to eight workgroups), and
  // the accuracy history of up to 128 cycles.
  h_stage_up_.resize(2 * static_cast<size_t>(pad_total_) + 16 * static_cast<size_t>(runs_) + 1024);
  h_stage_down_.resize(2 * static_cast<size_t>(pad_total_) + (32 + 64 + 8 * 128) * static_cast<size_t>(runs_) + 1024);
  std::fill(h_labels_.get(), h_labels_.get() + pad_total_, 0);
  std::fill(h_pred_.get(), h_pred_.get() + pad_total_, 0);
}

void KnnBatch::find_empty_rows() {
  // Filled slots come first, so a row whose first slot is empty has no neighbour.
  std::vector<int> first_col(static_cast<size_t>(pad_total_));
  d_nbr_.download(first_col.data(), static_cast<size_t>(pad_total_), stream_.get());
  stream_.sync();
  empty_rows_.assign(static_cast<size_t>(runs_), IVec());
  for (int r = 0; r < runs_; ++r) {
    const int L = run_offsets_[r + 1] - run_offsets_[r];
    for (int i = 0; i < L; ++i) if (first_col[pad_offsets_[r] + i] < 0) empty_rows_[r].push_back(i);
  }
}

KnnBatch::KnnBatch(int device, const int* neighbors, const float* scores, const IVec& run_offsets, int k,
                   int classes)
    : stream_(device), runs_(static_cast<int>(run_offsets.size()) - 1), k_(k), classes_(std::max(1, classes)),
      run_offsets_(run_offsets) {
  if (k < 1 || k > kMaxK) throw InvalidArgument("neighbours per row must be in 1..64.");
  if (runs_ < 1) throw InvalidArgument("at least one run is required.");
  allocate(run_offsets_);
  // Transpose into the padded vote layout with padded row ids, filled slots
  // first so that empty rows are detectable from column 0.
  std::vector<int> tn(static_cast<size_t>(k) * pad_total_, -1);
  std::vector<float> ts(static_cast<size_t>(k) * pad_total_, -INFINITY);
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < runs_; ++r) {
    const int L = run_offsets_[r + 1] - run_offsets_[r];
    for (int i = 0; i < L; ++i) {
      const size_t src = static_cast<size_t>(run_offsets_[r] + i) * k;
      const size_t row = static_cast<size_t>(pad_offsets_[r] + i);
      int w = 0;
      for (int j = 0; j < k; ++j) {
        const int nb = neighbors[src + j];
        if (nb < 0 || nb >= L) continue;
        tn[static_cast<size_t>(w) * pad_total_ + row] = pad_offsets_[r] + nb;
        ts[static_cast<size_t>(w) * pad_total_ + row] = scores[src + j];
        ++w;
      }
    }
  }
  d_nbr_.upload(tn.data(), tn.size(), stream_.get());
  d_score_.upload(ts.data(), ts.size(), stream_.get());
  find_empty_rows();
}

KnnBatch::KnnBatch(int device, const float* x, int n, int p, const IVec& run_offsets, const IVec& rows,
                   const IVec& folds, int k, int metric, int classes, int table_path)
    : stream_(device), runs_(static_cast<int>(run_offsets.size()) - 1), k_(k), classes_(classes),
      run_offsets_(run_offsets) {
  if (k < 1) throw InvalidArgument("KNNOptions::k must be positive.");
  if (k > kMaxK) throw InvalidArgument("k above 64 is not supported by the HIP kNN kernel.");
  if (p < 1 || p > kMaxDims) throw InvalidArgument("the HIP kNN kernel supports 1..256 feature columns.");
  if (classes < 1) throw InvalidArgument("at least one class is required.");
  if (runs_ < 1) throw InvalidArgument("at least one run is required.");
  const hipStream_t s = stream_.get();
  allocate(run_offsets);
  int max_rows = 0;
  for (int r = 0; r < runs_; ++r) max_rows = std::max(max_rows, run_offsets[r + 1] - run_offsets[r]);

  DeviceBuffer<float> d_x(static_cast<size_t>(n) * p);
  DeviceBuffer<int> d_rows(rows.size()), d_folds(folds.size());
  d_x.upload(x, static_cast<size_t>(n) * p, s);
  d_rows.upload(rows.data(), rows.size(), s);
  d_folds.upload(folds.data(), folds.size(), s);

  FoldTableJob job;
  job.d_x = d_x.get();
  job.p = p;
  job.k = k;
  job.metric = metric;
  job.d_rows = d_rows.get();
  job.d_folds = d_folds.get();
  job.run_offsets = run_offsets;
  job.d_run_offsets = d_run_off_.get();
  job.d_pad_offsets = d_pad_off_.get();
  job.pad_total = pad_total_;
  job.max_rows = max_rows;
  for (int f : folds) job.folds = std::max(job.folds, f + 1);
  job.norm2_max = max_row_norm2(x, n, p);
  job.path = table_path;
  job.fallbacks = &last_knn_fallbacks();
  job.d_nbr = d_nbr_.get();
  job.d_score = d_score_.get();
  device_fold_tables(job, s);
  find_empty_rows();
}

KnnBatch::~KnnBatch() {
  if (graph_exec_ != nullptr) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_ != nullptr) (void)hipGraphDestroy(graph_);
}

void KnnBatch::enqueue_vote(bool reduce, bool timed) {
  const hipStream_t s = stream_.get();
  d_labels_.upload(h_labels_.get(), static_cast<size_t>(pad_total_), s);
  if (reduce) check(launch_vote_reset(d_counts_.get(), d_firsts_.get(), d_matches_.get(), runs_, classes_, s), "vote reset");
  if (timed) timer_.start(s);
  VoteParams prm;
  prm.neighbors = d_nbr_.get();
  prm.scores = d_score_.get();
  prm.labels = d_labels_.get();
  prm.block_run = d_block_run_.get();
  prm.run_offsets = d_run_off_.get();
  prm.pad_offsets = d_pad_off_.get();
  prm.pad_total = pad_total_;
  prm.k = k_;
  prm.max_rows = max_rows_;
  prm.classes = classes_;
  prm.predictions = d_pred_.get();
  if (reduce) {
    prm.counts = d_counts_.get();
    prm.firsts = d_firsts_.get();
    prm.matches = d_matches_.get();
  }
  check(launch_vote(prm, blocks_, s), "vote launch");
  if (timed) timer_.stop(s);
  d_pred_.download(h_pred_.get(), static_cast<size_t>(pad_total_), s);
  if (reduce) {
    d_counts_.download(h_counts_.get(), static_cast<size_t>(runs_) * classes_, s);
    d_firsts_.download(h_firsts_.get(), static_cast<size_t>(runs_) * classes_, s);
    d_matches_.download(h_matches_.get(), static_cast<size_t>(runs_), s);
  }
}

void KnnBatch::vote_all(bool reduce) {
  const hipStream_t s = stream_.get();
  if (graph_on_ && reduce && graph_exec_ != nullptr) {
    check(hipGraphLaunch(graph_exec_, s), "hipGraphLaunch");
    stream_.sync();
    return;
  }
  enqueue_vote(reduce, true);
  stream_.sync();
  vote_ms_ += timer_.ms();
  ++vote_calls_;
  if (graph_on_ && reduce) {
    // Capture the same sequence (without the timer) for the later calls.
    check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    enqueue_vote(true, false);
    check(hipStreamEndCapture(s, &graph_), "hipStreamEndCapture");
    check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
  }
}

void KnnBatch::download_tables(int r, int* nbr, float* score) {
  const hipStream_t s = stream_.get();
  const int L = run_offsets_[r + 1] - run_offsets_[r];
  std::vector<int> col(static_cast<size_t>(L));
  std::vector<float> scol(static_cast<size_t>(L));
  for (int j = 0; j < k_; ++j) {
    const size_t off = static_cast<size_t>(j) * pad_total_ + pad_offsets_[r];
    d_nbr_.download(col.data(), static_cast<size_t>(L), s, off);
    d_score_.download(scol.data(), static_cast<size_t>(L), s, off);
    stream_.sync();
    for (int i = 0; i < L; ++i) {
      nbr[static_cast<size_t>(i) * k_ + j] = col[i] >= 0 ? col[i] - pad_offsets_[r] : -1;
      score[static_cast<size_t>(i) * k_ + j] = scol[i];
    }
  }
}

int resolve_threads(int requested, int runs) {
  int t = requested > 0 ? requested : static_cast<int>(std::thread::hardware_concurrency());
  if (t < 1) t = 1;
  return std::max(1, std::min(t, runs));
}

namespace {

// Training-fold majority for rows without neighbours (core.cpp:274-309): the
// label with most training rows in the fold, ties to the smaller label.
void apply_fold_fallbacks(const IVec& empty_rows, const IVec& folds, const int* labels, int classes,
                          int* pred, int* matches) {
  if (empty_rows.empty()) return;
  const int n = static_cast<int>(folds.size());
  IVec total(static_cast<size_t>(classes), 0);
  for (int i = 0; i < n; ++i) ++total[labels[i]];
  std::map<int, int> fold_fallback;
  for (int r : empty_rows) {
    const int f = folds[r];
    auto it = fold_fallback.find(f);
    if (it == fold_fallback.end()) {
      IVec tr = total;
      for (int i = 0; i < n; ++i) if (folds[i] == f) --tr[labels[i]];
      int best = 0;
      for (int c = 1; c < classes; ++c) if (tr[c] > tr[best]) best = c;
      it = fold_fallback.emplace(f, best).first;
    }
    const int prev = pred[r];
    pred[r] = it->second;
    if (prev != labels[r] && pred[r] == labels[r]) ++*matches;
    if (prev == labels[r] && pred[r] != labels[r]) --*matches;
  }
}

}  // namespace

void evolve_plain(std::vector<RunEvolution>& runs, KnnBatch& batch, const std::vector<IVec>& folds,
                  int cycles, int n_threads, const CycleHook& hook) {
  const int M = static_cast<int>(runs.size());
  const int C = batch.classes();
  const int threads = resolve_threads(n_threads, M);
  int* h_labels = batch.labels_host();

  auto finish_run = [&](int r, int cycle, const int* labels) {
    int* pred = const_cast<int*>(batch.predictions_host()) + batch.pad_offset(r);
    int matches = batch.matches_host()[r];
    apply_fold_fallbacks(batch.empty_rows(r), folds[r], labels, runs[r].classes(), pred, &matches);
    Tally t;
    t.matches = matches;
    t.counts = batch.counts_host() + static_cast<size_t>(r) * C;
    t.firsts = batch.firsts_host() + static_cast<size_t>(r) * C;
    if (cycle < 0) runs[r].begin(pred, t);
    else runs[r].finish(cycle, pred, t);
  };

  // Initial evaluation of the starting labels (eager, timed; it also
  // captures the per-cycle graph that the cycles replay).
  for (int r = 0; r < M; ++r) {
    const IVec& init = runs[r].initial();
    std::copy(init.begin(), init.end(), h_labels + batch.pad_offset(r));
  }
  batch.use_graph(cycles > 0);
  batch.vote_all();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (int r = 0; r < M; ++r) finish_run(r, -1, h_labels + batch.pad_offset(r));

  for (int cycle = 0; cycle < cycles; ++cycle) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int r = 0; r < M; ++r) {
      const IVec& cl = runs[r].propose(cycle);
      std::copy(cl.begin(), cl.end(), h_labels + batch.pad_offset(r));
    }
    batch.vote_all();
    if (hook) hook(cycle, false);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int r = 0; r < M; ++r) finish_run(r, cycle, h_labels + batch.pad_offset(r));
    if (hook) hook(cycle, true);
  }
}

}  // namespace kodama_hip
