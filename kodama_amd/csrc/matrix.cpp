// KODAMA.matrix building blocks on the host: the reference's random streams
// (libstdc++ mt19937 / mt19937_64, std::shuffle, std::generate_canonical,
// std::sample, called in the reference's order so the draws coincide) and the
// k-means driver whose Lloyd steps run on the device (matrix_kernels.hip).
#include "matrix.h"

#include <algorithm>
#include <map>
#include <cmath>
#include <iterator>
#include <limits>
#include <numeric>
#include <stdexcept>

#include "device.h"
#include "matrix_launch.h"

namespace kodama_hip {

namespace {

// mt19937 seeded with the low 32 bits: integers by modulo, floats as
// draw / 2^32-1 in float (kodama_matrix.cpp:535-549).
class FaissStream {
 public:
  explicit FaissStream(std::uint64_t seed) : g_(static_cast<std::uint32_t>(seed)) {}
  int below(int m) { return static_cast<int>(g_() % static_cast<std::uint32_t>(m)); }
  float unit() { return static_cast<float>(g_()) / static_cast<float>(std::mt19937::max()); }

 private:
  std::mt19937 g_;
};

// Fisher-Yates from the front (kodama_matrix.cpp:551-564).
IVec seed_order(int n, std::uint64_t seed) {
  IVec order(static_cast<size_t>(n));
  std::iota(order.begin(), order.end(), 0);
  FaissStream r(seed);
  for (int i = 0; i + 1 < n; ++i) std::swap(order[i], order[i + r.below(n - i)]);
  return order;
}

// Float prefix-sum tree over the donor weights, updated in place
// (kodama_matrix.cpp:566-603): the same float additions in the same order.
class DonorTree {
 public:
  explicit DonorTree(const std::vector<float>& w) : t_(w.size() + 1, 0.0f), w_(w.size(), 0.0f) {
    for (size_t i = 0; i < w.size(); ++i) set(i, w[i]);
  }
  void set(size_t i, float v) {
    const float delta = v - w_[i];
    w_[i] = v;
    for (size_t j = i + 1; j < t_.size(); j += j & (~j + 1)) t_[j] += delta;
  }
  float total() const {
    float acc = 0.0f;
    for (size_t j = w_.size(); j > 0; j -= j & (~j + 1)) acc += t_[j];
    return acc;
  }
  // Largest prefix whose float sum stays <= target; -1 past the end.
  int find(float target) const {
    size_t pos = 0, step = 1;
    float prefix = 0.0f;
    while ((step << 1) < t_.size()) step <<= 1;
    for (; step > 0; step >>= 1) {
      const size_t nx = pos + step;
      if (nx < t_.size() && prefix + t_[nx] <= target) {
        pos = nx;
        prefix += t_[nx];
      }
    }
    return pos < w_.size() ? static_cast<int>(pos) : -1;
  }

 private:
  std::vector<float> t_, w_;
};

// Each empty cluster takes half of a donor drawn by weight max(0, size - 1)
// (the largest cluster if every weight is zero); the two centroids are pushed
// apart by a relative 1/1024 with alternating signs (kodama_matrix.cpp:605-662).
void split_empty_clusters(int rows, int p, int k, std::vector<float>& sizes, std::vector<float>& cent) {
  if (rows <= k) return;
  const float eps = 1.0f / 1024.0f;
  FaissStream r(1234u);
  std::vector<float> w(static_cast<size_t>(k));
  for (int c = 0; c < k; ++c) w[c] = std::max(0.0f, sizes[c] - 1.0f);
  DonorTree tree(w);
  for (int e = 0; e < k; ++e) {
    if (sizes[e] != 0.0f) continue;
    const float total = tree.total();
    int donor = total > 0.0f ? tree.find(r.unit() * total) : -1;
    if (donor < 0) donor = static_cast<int>(std::max_element(sizes.begin(), sizes.end()) - sizes.begin());
    float* ce = cent.data() + static_cast<size_t>(e) * p;
    float* cd = cent.data() + static_cast<size_t>(donor) * p;
    std::copy_n(cd, p, ce);
    for (int j = 0; j < p; ++j) {
      const bool even = (j & 1) == 0;
      ce[j] *= even ? 1.0f + eps : 1.0f - eps;
      cd[j] *= even ? 1.0f - eps : 1.0f + eps;
    }
    sizes[e] = sizes[donor] / 2.0f;
    sizes[donor] -= sizes[e];
    tree.set(static_cast<size_t>(e), std::max(0.0f, sizes[e] - 1.0f));
    tree.set(static_cast<size_t>(donor), std::max(0.0f, sizes[donor] - 1.0f));
  }
}

}  // namespace

IVec device_kmeans(const float* h_x, const float* d_x, int n, int p, int k, std::mt19937_64& rng, int max_iter,
                   hipStream_t s) {
  k = std::max(1, std::min(k, n));
  const std::uint64_t kseed = rng() & 0x7fffffffULL;
  const IVec order = seed_order(n, kseed + 1u);
  std::vector<float> cent(static_cast<size_t>(k) * p), sizes(static_cast<size_t>(k));
  for (int c = 0; c < k; ++c) std::copy_n(h_x + static_cast<size_t>(order[c]) * p, p, cent.data() + static_cast<size_t>(c) * p);
  DeviceBuffer<float> d_xn(static_cast<size_t>(n)), d_cent(cent.size()), d_cn(static_cast<size_t>(k)),
    d_sizes(static_cast<size_t>(k));
  DeviceBuffer<int> d_asg(static_cast<size_t>(n)), d_changed(1);
  check(launch_sse_norms(d_x, n, p, d_xn.get(), s), "sse_norms launch");
  check(hipMemsetAsync(d_asg.get(), 0xFF, sizeof(int) * d_asg.size(), s), "memset");  // -1: unassigned
  auto assign = [&]() {
    d_cent.upload(cent.data(), cent.size(), s);
    check(launch_sse_norms(d_cent.get(), k, p, d_cn.get(), s), "sse_norms launch");
    check(hipMemsetAsync(d_changed.get(), 0, sizeof(int), s), "memset");
    check(launch_kmeans_assign(d_x, n, p, d_xn.get(), d_cent.get(), d_cn.get(), k, d_asg.get(), d_changed.get(), s),
          "kmeans_assign launch");
    int changed = 0;
    d_changed.download(&changed, 1, s);
    check(hipStreamSynchronize(s), "kmeans sync");
    return changed;
  };
  for (int it = 0; it < std::max(1, max_iter); ++it) {
    const int changed = assign();
    check(launch_kmeans_update(d_x, n, p, d_asg.get(), k, d_cent.get(), d_sizes.get(), s), "kmeans_update launch");
    d_cent.download(cent.data(), cent.size(), s);
    d_sizes.download(sizes.data(), sizes.size(), s);
    check(hipStreamSynchronize(s), "kmeans sync");
    split_empty_clusters(n, p, k, sizes, cent);
    if (changed == 0) break;
  }
  assign();
  IVec labels(static_cast<size_t>(n));
  d_asg.download(labels.data(), labels.size(), s);
  check(hipStreamSynchronize(s), "kmeans sync");
  for (int& l : labels) ++l;
  return labels;
}

std::vector<IVec> device_kmeans_batch(const float* h_x, const int* rows, const float* d_rows, int runs, int n, int p,
                                      int k, std::vector<std::mt19937_64>& rngs, int max_iter, hipStream_t s) {
  k = std::max(1, std::min(k, n));
  const size_t kp = static_cast<size_t>(k) * p;
  std::vector<float> cent(static_cast<size_t>(runs) * kp), sizes(static_cast<size_t>(runs) * k);
  for (int r = 0; r < runs; ++r) {
    const std::uint64_t kseed = rngs[r]() & 0x7fffffffULL;
    const IVec order = seed_order(n, kseed + 1u);
    for (int c = 0; c < k; ++c) {
      const int data_row = rows[static_cast<size_t>(r) * n + order[c]];
      std::copy_n(h_x + static_cast<size_t>(data_row) * p, p, cent.data() + r * kp + static_cast<size_t>(c) * p);
    }
  }
  DeviceBuffer<float> d_xn(static_cast<size_t>(runs) * n), d_cent(cent.size()), d_cn(static_cast<size_t>(runs) * k),
    d_sizes(sizes.size());
  DeviceBuffer<int> d_asg(static_cast<size_t>(runs) * n), d_changed(static_cast<size_t>(runs)),
    d_active(static_cast<size_t>(runs));
  check(launch_sse_norms(d_rows, runs * n, p, d_xn.get(), s), "sse_norms launch");
  check(hipMemsetAsync(d_asg.get(), 0xFF, sizeof(int) * d_asg.size(), s), "memset");  // -1: unassigned
  // Phase of each problem: Lloyd rounds (assign, update, split), then one
  // final assignment with the last centroids, then done.
  enum Phase { kLoop, kFinal, kDone };
  std::vector<int> phase(static_cast<size_t>(runs), kLoop), rounds(static_cast<size_t>(runs), 0),
    act(static_cast<size_t>(runs)), changed(static_cast<size_t>(runs));
  for (;;) {
    bool any = false;
    for (int r = 0; r < runs; ++r) {
      act[r] = phase[r] != kDone ? 1 : 0;
      any = any || act[r];
    }
    if (!any) break;
    d_cent.upload(cent.data(), cent.size(), s);
    check(launch_sse_norms(d_cent.get(), runs * k, p, d_cn.get(), s), "sse_norms launch");
    check(hipMemsetAsync(d_changed.get(), 0, sizeof(int) * runs, s), "memset");
    d_active.upload(act.data(), act.size(), s);
    check(launch_kmeans_assign(d_rows, n, p, d_xn.get(), d_cent.get(), d_cn.get(), k, d_asg.get(), d_changed.get(), s,
                               runs, d_active.get()),
          "kmeans_assign launch");
    d_changed.download(changed.data(), changed.size(), s);
    check(hipStreamSynchronize(s), "kmeans sync");  // act and changed are reused below
    bool update = false;
    for (int r = 0; r < runs; ++r) {
      act[r] = phase[r] == kLoop ? 1 : 0;
      update = update || act[r];
    }
    if (update) {
      d_active.upload(act.data(), act.size(), s);
      check(launch_kmeans_update(d_rows, n, p, d_asg.get(), k, d_cent.get(), d_sizes.get(), s, runs, d_active.get()),
            "kmeans_update launch");
      d_cent.download(cent.data(), cent.size(), s);
      d_sizes.download(sizes.data(), sizes.size(), s);
      check(hipStreamSynchronize(s), "kmeans sync");
#pragma omp parallel for schedule(dynamic)
      for (int r = 0; r < runs; ++r) {
        if (!act[r]) continue;
        std::vector<float> sz(sizes.begin() + static_cast<size_t>(r) * k, sizes.begin() + static_cast<size_t>(r + 1) * k);
        std::vector<float> ce(cent.begin() + r * kp, cent.begin() + (r + 1) * kp);
        split_empty_clusters(n, p, k, sz, ce);
        std::copy(ce.begin(), ce.end(), cent.begin() + r * kp);
      }
    }
    for (int r = 0; r < runs; ++r) {
      if (phase[r] == kFinal) {
        phase[r] = kDone;
      } else if (phase[r] == kLoop) {
        ++rounds[r];
        if (changed[r] == 0 || rounds[r] >= std::max(1, max_iter)) phase[r] = kFinal;
      }
    }
  }
  std::vector<int> flat(static_cast<size_t>(runs) * n);
  d_asg.download(flat.data(), flat.size(), s);
  check(hipStreamSynchronize(s), "kmeans sync");
  std::vector<IVec> labels(static_cast<size_t>(runs));
  for (int r = 0; r < runs; ++r) {
    labels[r].assign(flat.begin() + static_cast<size_t>(r) * n, flat.begin() + static_cast<size_t>(r + 1) * n);
    for (int& l : labels[r]) ++l;
  }
  return labels;
}

void index_strata(const IVec& labels, int strata, IVec& offsets, IVec& rows) {
  if (strata < 1) throw InvalidArgument("landmark strata count must be positive.");
  offsets.assign(static_cast<size_t>(strata) + 1, 0);
  for (int l : labels) {
    if (l < 1 || l > strata) throw InvalidArgument("landmark stratum label is out of range.");
    ++offsets[l];
  }
  std::partial_sum(offsets.begin(), offsets.end(), offsets.begin());
  rows.assign(labels.size(), 0);
  IVec cursor(offsets.begin(), offsets.end() - 1);
  for (size_t i = 0; i < labels.size(); ++i) rows[cursor[labels[i] - 1]++] = static_cast<int>(i);
}

IVec quota_sample_landmarks(const IVec& offsets, const IVec& rows, int target, std::mt19937_64& rng) {
  const int total = static_cast<int>(rows.size());
  target = std::max(1, std::min(target, total));
  const int strata = static_cast<int>(offsets.size()) - 1;
  // Proportional quotas: floor of target * count / total (long double), the
  // fractions kept for the residual slots.
  IVec quota(static_cast<size_t>(strata), 0), live;
  std::vector<long double> rest(static_cast<size_t>(strata), 0.0L);
  int given = 0;
  for (int st = 0; st < strata; ++st) {
    const int count = offsets[st + 1] - offsets[st];
    if (count <= 0) continue;
    live.push_back(st);
    const long double share =
      static_cast<long double>(target) * static_cast<long double>(count) / static_cast<long double>(total);
    quota[st] = std::min(count, static_cast<int>(std::floor(share)));
    rest[st] = share - static_cast<long double>(quota[st]);
    given += quota[st];
  }
  std::shuffle(live.begin(), live.end(), rng);
  const int residual = target - given;
  if (residual > 0) {
    // Systematic draw: one uniform offset, then every unit step of the running
    // fraction sum over the shuffled strata takes one more row.
    const long double start = std::generate_canonical<long double, 64>(rng);
    long double running = 0.0L;
    int taken = 0;
    std::vector<char> bumped(static_cast<size_t>(strata), 0);
    for (int st : live) {
      running += rest[st];
      if (taken < residual && start + static_cast<long double>(taken) < running) {
        ++quota[st];
        bumped[st] = 1;
        ++taken;
      }
    }
    // A last boundary lost to rounding goes to the largest remaining fraction.
    while (taken < residual) {
      int pick = -1;
      long double best = -1.0L;
      for (int st : live) {
        if (!bumped[st] && rest[st] > best) {
          best = rest[st];
          pick = st;
        }
      }
      if (pick < 0) throw DeviceError("landmark quota allocation could not be completed.");
      ++quota[pick];
      bumped[pick] = 1;
      ++taken;
    }
  }
  IVec out;
  out.reserve(static_cast<size_t>(target));
  for (int st = 0; st < strata; ++st) {
    if (quota[st] <= 0) continue;
    std::sample(rows.cbegin() + offsets[st], rows.cbegin() + offsets[st + 1], std::back_inserter(out), quota[st],
                rng);
  }
  if (static_cast<int>(out.size()) != target) throw DeviceError("landmark quota allocation gave the wrong size.");
  std::sort(out.begin(), out.end());
  return out;
}

IVec normalize_constrain(const int32_t* constrain, int n) {
  IVec out(static_cast<size_t>(n));
  if (constrain == nullptr) {
    std::iota(out.begin(), out.end(), 1);
    return out;
  }
  std::map<int, int> ids;
  for (int i = 0; i < n; ++i) out[i] = ids.emplace(constrain[i], static_cast<int>(ids.size()) + 1).first->second;
  return out;
}

bool is_identity_constrain(const IVec& groups) {
  for (size_t i = 0; i < groups.size(); ++i) {
    if (groups[i] != static_cast<int>(i + 1)) return false;
  }
  return true;
}

IVec factor_subset(const IVec& values, const IVec& rows) {
  std::map<int, int> ids;
  IVec out(rows.size());
  for (size_t i = 0; i < rows.size(); ++i) {
    out[i] = ids.emplace(values[rows[i]], static_cast<int>(ids.size()) + 1).first->second;
  }
  return out;
}

int majority_label(const IVec& labels) {
  std::map<int, int> counts;
  for (int v : labels) ++counts[v];
  int best = counts.begin()->first, best_n = counts.begin()->second;
  for (const auto& kv : counts) {
    if (kv.second > best_n) {  // ascending keys: a tie keeps the smaller label
      best = kv.first;
      best_n = kv.second;
    }
  }
  return best;
}

IVec constrained_majority(const IVec& labels, const IVec& groups) {
  std::map<int, IVec> members;
  for (size_t i = 0; i < labels.size(); ++i) members[groups[i]].push_back(labels[i]);
  std::map<int, int> label_of;
  for (const auto& kv : members) label_of[kv.first] = majority_label(kv.second);
  IVec out(labels.size());
  for (size_t i = 0; i < labels.size(); ++i) out[i] = label_of[groups[i]];
  return out;
}

}  // namespace kodama_hip
