// Learned proposal, host side: the row features (the twin of the device
// kernel), accept/reject traces of the plain evolution, and the SGD driver.
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>

#include "device.h"

namespace kodama_hip {

void host_row_features(const int* codes, int k, int own, float* feat, int* ranked) {
  // Distinct voted codes with their counts, ordered (count desc, code asc).
  std::vector<int> c;
  c.reserve(static_cast<size_t>(k));
  for (int j = 0; j < k; ++j) if (codes[j] >= 0) c.push_back(codes[j]);
  std::sort(c.begin(), c.end());
  std::vector<std::pair<int, int>> voted;  // (count, code)
  int own_count = 0;
  for (size_t i = 0; i < c.size();) {
    size_t j = i + 1;
    while (j < c.size() && c[j] == c[i]) ++j;
    voted.emplace_back(static_cast<int>(j - i), c[i]);
    if (c[i] == own) own_count = static_cast<int>(j - i);
    i = j;
  }
  std::sort(voted.begin(), voted.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
    return a.first != b.first ? a.first > b.first : a.second < b.second;
  });
  std::fill(feat, feat + kMlpIn, 0.0f);
  int own_rank = kMlpRanks;
  for (int j = 0; j < kMlpRanks; ++j) {
    const bool real = j < static_cast<int>(voted.size());
    feat[j] = real ? static_cast<float>(voted[j].first) / static_cast<float>(k) : 0.0f;
    ranked[j] = real ? voted[j].second : -1;
    if (real && voted[j].second == own && own_rank == kMlpRanks) own_rank = j;
  }
  const int top = voted.empty() ? 0 : voted[0].first;
  feat[kMlpRanks + own_rank] = 1.0f;
  feat[61] = static_cast<float>(top - own_count) / static_cast<float>(k);
  feat[62] = 1.0f;
}

void evolve_plain_traced(std::vector<RunEvolution>& runs, KnnBatch& batch, const std::vector<IVec>& folds,
                         int cycles, int n_threads, TraceSink& sink) {
  const int M = static_cast<int>(runs.size());
  const int k = batch.k();
  const int pad = batch.pad_total();
  // Host copy of the neighbour tables ([k][pad_total] padded ids) in one download.
  std::vector<int> nbr(static_cast<size_t>(k) * pad);
  batch.d_neighbors().download(nbr.data(), nbr.size(), batch.stream());
  check(hipStreamSynchronize(batch.stream()), "trace table sync");
  // Samples of the current cycle per run, resolved after acceptance.
  std::vector<std::vector<float>> feat(static_cast<size_t>(M));
  std::vector<std::vector<int>> rank(static_cast<size_t>(M));
  std::vector<int> codes(static_cast<size_t>(k)), ranked(kMlpRanks);
  float f[kMlpIn];
  // Outcome samples wait for the final labelling: (run, row, current code,
  // ranked codes) beside their features.
  struct Pending {
    int run, row, cur;
    int ranked[kMlpRanks];
  };
  std::vector<Pending> pending;
  std::vector<float> pending_feat;
  auto record_outcomes = [&]() {
    for (int r = 0; r < M; ++r) {
      RunEvolution& ev = runs[r];
      const IVec& cur = ev.current_labels();
      const IVec& pred = ev.current_predictions();
      const int base = batch.pad_offset(r);
      for (int i = 0; i < ev.rows(); ++i) {
        if (static_cast<int>(pending.size()) >= sink.max_samples) return;
        if (ev.fixed_row(i) || pred[i] == cur[i]) continue;
        for (int j = 0; j < k; ++j) {
          const int id = nbr[static_cast<size_t>(j) * pad + base + i];
          codes[j] = id >= 0 ? cur[id - base] : -1;
        }
        Pending pd;
        pd.run = r;
        pd.row = i;
        pd.cur = cur[i];
        host_row_features(codes.data(), k, cur[i], f, pd.ranked);
        pending.push_back(pd);
        pending_feat.insert(pending_feat.end(), f, f + kMlpIn);
      }
    }
  };
  auto hook = [&](int cycle, bool after_finish) {
    if (sink.outcome) {
      if (!after_finish) record_outcomes();
      return;
    }
    for (int r = 0; r < M; ++r) {
      RunEvolution& ev = runs[r];
      if (!after_finish) {
        feat[r].clear();
        rank[r].clear();
        if (!ev.singleton_groups()) continue;
        const IVec& cur = ev.current_labels();
        const IVec& pred = ev.current_predictions();
        const IVec& order = ev.sampled_groups();
        const int base = batch.pad_offset(r);
        const int take = ev.proposal_size(cycle);
        for (int t = 0; t < take; ++t) {
          const int i = order[t];
          if (ev.fixed_row(i) || pred[i] == cur[i]) continue;
          for (int j = 0; j < k; ++j) {
            const int id = nbr[static_cast<size_t>(j) * pad + base + i];
            codes[j] = id >= 0 ? cur[id - base] : -1;
          }
          host_row_features(codes.data(), k, cur[i], f, ranked.data());
          int rk = -1;
          for (int j = 0; j < kMlpRanks; ++j) if (ranked[j] == pred[i]) rk = j;
          if (rk < 0) continue;  // a fallback prediction outside the votes
          feat[r].insert(feat[r].end(), f, f + kMlpIn);
          rank[r].push_back(rk);
        }
      } else {
        const bool acc = ev.accepted(cycle);
        for (size_t s = 0; s < rank[r].size(); ++s) {
          if (static_cast<int>(sink.target.size()) >= sink.max_samples) return;
          sink.feat.insert(sink.feat.end(), feat[r].begin() + s * kMlpIn, feat[r].begin() + (s + 1) * kMlpIn);
          sink.target.push_back(acc ? rank[r][s] : kMlpKeep);
        }
      }
    }
  };
  evolve_plain(runs, batch, folds, cycles, n_threads, hook);
  if (sink.outcome) {
    for (size_t e = 0; e < pending.size(); ++e) {
      const Pending& pd = pending[e];
      const int fin = runs[pd.run].best_codes()[pd.row];
      int target = -1;
      if (fin == pd.cur) {
        target = kMlpKeep;
      } else {
        for (int j = 0; j < kMlpRanks; ++j) if (pd.ranked[j] == fin) target = j;
      }
      if (target < 0) continue;
      sink.feat.insert(sink.feat.end(), pending_feat.begin() + e * kMlpIn, pending_feat.begin() + (e + 1) * kMlpIn);
      sink.target.push_back(target);
    }
  }
}

void mlp_init(std::uint64_t seed, float* params) {
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> g(0.0f, 1.0f);
  std::fill(params, params + kMlpParams, 0.0f);
  auto fill = [&](int off, int fan_in, int count) {
    const float sd = std::sqrt(2.0f / static_cast<float>(fan_in));
    for (int e = 0; e < count; ++e) params[off + e] = sd * g(rng);
  };
  fill(kMlpW1, kMlpIn, kMlpIn * kMlpHidden);
  fill(kMlpW2, kMlpHidden, kMlpHidden * kMlpHidden);
  fill(kMlpW3, kMlpHidden, kMlpHidden * kMlpOut);
}

void mlp_train(int device, const float* feat, const int* target, int samples, int epochs, int batch, float lr,
               float momentum, std::uint64_t seed, float* params, double* loss_per_epoch) {
  if (samples < 1 || epochs < 0 || batch < 1) throw InvalidArgument("MLP training needs samples and a batch size.");
  for (int s = 0; s < samples; ++s) {
    if (target[s] < 0 || target[s] > kMlpKeep) throw InvalidArgument("MLP target outside 0..30.");
  }
  Stream st(device);
  const hipStream_t s = st.get();
  DeviceBuffer<float> d_feat(static_cast<size_t>(samples) * kMlpIn), d_params(kMlpParams), d_vel(kMlpParams),
    d_partial(static_cast<size_t>(kMlpGradWaves) * (kMlpParams + 1));
  DeviceBuffer<int> d_target(static_cast<size_t>(samples)), d_order(static_cast<size_t>(samples));
  const int batches = (samples + batch - 1) / batch;
  DeviceBuffer<float> d_loss(static_cast<size_t>(std::max(1, batches)));
  d_feat.upload(feat, d_feat.size(), s);
  d_target.upload(target, d_target.size(), s);
  d_params.upload(params, kMlpParams, s);
  check(hipMemsetAsync(d_vel.get(), 0, sizeof(float) * kMlpParams, s), "memset");
  std::vector<int> order(static_cast<size_t>(samples));
  std::vector<float> loss(static_cast<size_t>(batches));
  for (int e = 0; e < epochs; ++e) {
    std::iota(order.begin(), order.end(), 0);
    std::mt19937_64 rng(seed + static_cast<std::uint64_t>(e));
    std::shuffle(order.begin(), order.end(), rng);
    d_order.upload(order.data(), order.size(), s);
    for (int b = 0; b < batches; ++b) {
      const int begin = b * batch;
      const int count = std::min(batch, samples - begin);
      check(launch_mlp_grad(d_params.get(), d_feat.get(), d_target.get(), d_order.get() + begin, count,
                            d_partial.get(), s),
            "mlp_grad launch");
      check(launch_mlp_sgd(d_params.get(), d_vel.get(), d_partial.get(), count, lr, momentum, d_loss.get() + b, s),
            "mlp_sgd launch");
    }
    d_loss.download(loss.data(), loss.size(), s);
    st.sync();
    if (loss_per_epoch) {
      double sum = 0.0;
      for (float v : loss) sum += v;
      loss_per_epoch[e] = sum / batches;
    }
  }
  d_params.download(params, kMlpParams, s);
  st.sync();
}

}  // namespace kodama_hip
