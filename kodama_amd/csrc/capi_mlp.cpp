// C ABI of libkodama_hip.so, learned-proposal half: the network's features,
// forward pass and gradients (for checks against a CPU restatement), trace
// collection from the plain evolution, device training, and the batched
// device-resident evolution with the learned proposal.
#include <algorithm>
#include <numeric>

#include "../../include/kodama_hip.h"
#include "batch_api.h"
#include "capi_common.h"
#include "device.h"
#include "mlp.h"
#include "mlp_launch.h"

using namespace kodama_hip;

namespace {

void require(bool ok, const char* msg) {
  if (!ok) throw InvalidArgument(msg);
}

}  // namespace

extern "C" {

int kodama_hip_mlp_init(uint64_t seed, float* params) {
  return guarded([&] {
    require(params != nullptr, "null pointer.");
    mlp_init(seed, params);
  });
}

int kodama_hip_mlp_features(const int32_t* neighbors, const int32_t* labels, int32_t rows, int32_t k,
                            int32_t device, float* features, int32_t* ranked) {
  return guarded([&] {
    require(neighbors != nullptr && labels != nullptr && features != nullptr && ranked != nullptr, "null pointer.");
    require(rows >= 1 && k >= 1 && k <= 64, "rows must be positive and k in 1..64.");
    for (int i = 0; i < rows; ++i) require(labels[i] >= 0 && labels[i] < 0xFFFF, "label codes must be in 0..65534.");
    select_device(device);
    Stream st(device);
    const hipStream_t s = st.get();
    DeviceBuffer<int> d_nbr(static_cast<size_t>(rows) * k), d_lab(static_cast<size_t>(rows));
    DeviceBuffer<float> d_feat(static_cast<size_t>(rows) * kMlpIn);
    DeviceBuffer<unsigned short> d_rank(static_cast<size_t>(rows) * kMlpRanks);
    d_nbr.upload(neighbors, d_nbr.size(), s);
    d_lab.upload(labels, d_lab.size(), s);
    check(launch_mlp_features(d_nbr.get(), d_lab.get(), rows, k, d_feat.get(), d_rank.get(), s), "mlp_features");
    std::vector<unsigned short> r16(d_rank.size());
    d_feat.download(features, d_feat.size(), s);
    d_rank.download(r16.data(), r16.size(), s);
    st.sync();
    for (size_t i = 0; i < r16.size(); ++i) ranked[i] = r16[i] == 0xFFFF ? -1 : r16[i];
  });
}

int kodama_hip_mlp_forward(const float* params, const float* features, int32_t rows, int32_t device, float* probs) {
  return guarded([&] {
    require(params != nullptr && features != nullptr && probs != nullptr, "null pointer.");
    require(rows >= 1, "rows must be positive.");
    select_device(device);
    Stream st(device);
    const hipStream_t s = st.get();
    DeviceBuffer<float> d_par(kMlpParams), d_feat(static_cast<size_t>(rows) * kMlpIn),
      d_prob(static_cast<size_t>(rows) * kMlpOut);
    DeviceBuffer<uint32_t> d_packed(kMlpTiled);
    d_par.upload(params, kMlpParams, s);
    d_feat.upload(features, d_feat.size(), s);
    check(launch_mlp_pack(d_par.get(), d_packed.get(), s), "mlp_pack");
    check(launch_mlp_forward(d_packed.get(), d_feat.get(), rows, d_prob.get(), s), "mlp_forward");
    d_prob.download(probs, d_prob.size(), s);
    st.sync();
  });
}

int kodama_hip_mlp_gradients(const float* params, const float* features, const int32_t* targets, int32_t rows,
                             int32_t device, float* grads, float* loss) {
  return guarded([&] {
    require(params != nullptr && features != nullptr && targets != nullptr && grads != nullptr, "null pointer.");
    require(rows >= 1, "rows must be positive.");
    for (int i = 0; i < rows; ++i) require(targets[i] >= 0 && targets[i] <= kMlpKeep, "targets must be in 0..30.");
    select_device(device);
    Stream st(device);
    const hipStream_t s = st.get();
    DeviceBuffer<float> d_par(kMlpParams), d_vel(kMlpParams), d_feat(static_cast<size_t>(rows) * kMlpIn),
      d_part(static_cast<size_t>(kMlpGradWaves) * (kMlpParams + 1)), d_loss(1);
    DeviceBuffer<int> d_tgt(static_cast<size_t>(rows)), d_idx(static_cast<size_t>(rows));
    std::vector<int> idx(static_cast<size_t>(rows));
    std::iota(idx.begin(), idx.end(), 0);
    d_par.upload(params, kMlpParams, s);
    d_feat.upload(features, d_feat.size(), s);
    d_tgt.upload(targets, d_tgt.size(), s);
    d_idx.upload(idx.data(), idx.size(), s);
    check(hipMemsetAsync(d_vel.get(), 0, sizeof(float) * kMlpParams, s), "memset");
    check(launch_mlp_grad(d_par.get(), d_feat.get(), d_tgt.get(), d_idx.get(), rows, d_part.get(), s), "mlp_grad");
    // A zero learning rate and momentum leave the parameters and put the mean
    // gradient in the velocity buffer.
    check(launch_mlp_sgd(d_par.get(), d_vel.get(), d_part.get(), rows, 0.0f, 0.0f, d_loss.get(), s), "mlp_sgd");
    d_vel.download(grads, kMlpParams, s);
    if (loss) d_loss.download(loss, 1, s);
    st.sync();
  });
}

}  // extern "C"

namespace {

void collect(const float* x, int32_t n, int32_t p, int32_t runs, const int32_t* run_offsets,
             const int32_t* landmark_rows, const int32_t* initial_labels, const int32_t* fold_ids,
             const kodama_hip_core_options* options, int32_t max_samples, bool outcome, float* features,
             int32_t* targets, int32_t* samples, double* accbest) {
  require(features != nullptr && targets != nullptr && samples != nullptr, "null pointer.");
  require(max_samples >= 1, "max_samples must be positive.");
  BatchRequest q{x, n, p, runs, run_offsets, landmark_rows, initial_labels, fold_ids, options};
  TraceSink sink;
  sink.max_samples = max_samples;
  sink.outcome = outcome;
  const int total = run_offsets ? run_offsets[runs] : 0;
  std::vector<int32_t> clbest(static_cast<size_t>(std::max(total, 1)));
  std::vector<double> acc(static_cast<size_t>(std::max(runs, 1)));
  run_knn_batch(q, 0, nullptr, 0.0f, &sink, clbest.data(), acc.data(), nullptr, nullptr);
  const int count = static_cast<int>(sink.target.size());
  std::copy(sink.feat.begin(), sink.feat.end(), features);
  std::copy(sink.target.begin(), sink.target.end(), targets);
  *samples = count;
  if (accbest) std::copy_n(acc.data(), runs, accbest);
}

}  // namespace

extern "C" {

int kodama_hip_mlp_collect_traces(const float* x, int32_t n, int32_t p, int32_t runs, const int32_t* run_offsets,
                                  const int32_t* landmark_rows, const int32_t* initial_labels,
                                  const int32_t* fold_ids, const kodama_hip_core_options* options,
                                  int32_t max_samples, float* features, int32_t* targets, int32_t* samples,
                                  double* accbest) {
  return guarded([&] {
    collect(x, n, p, runs, run_offsets, landmark_rows, initial_labels, fold_ids, options, max_samples, false,
            features, targets, samples, accbest);
  });
}

int kodama_hip_mlp_collect_outcomes(const float* x, int32_t n, int32_t p, int32_t runs,
                                    const int32_t* run_offsets, const int32_t* landmark_rows,
                                    const int32_t* initial_labels, const int32_t* fold_ids,
                                    const kodama_hip_core_options* options, int32_t max_samples, float* features,
                                    int32_t* targets, int32_t* samples, double* accbest) {
  return guarded([&] {
    collect(x, n, p, runs, run_offsets, landmark_rows, initial_labels, fold_ids, options, max_samples, true,
            features, targets, samples, accbest);
  });
}

int kodama_hip_mlp_train(const float* features, const int32_t* targets, int32_t samples, int32_t epochs,
                         int32_t batch, float learning_rate, float momentum, uint64_t seed, int32_t device,
                         float* params, double* loss_per_epoch) {
  return guarded([&] {
    require(features != nullptr && targets != nullptr && params != nullptr, "null pointer.");
    require(samples >= 1 && epochs >= 0 && batch >= 1, "samples and batch must be positive.");
    select_device(device);
    mlp_train(device, features, targets, samples, epochs, batch, learning_rate, momentum, seed, params,
              loss_per_epoch);
  });
}

int kodama_hip_core_knn_batch_learned(const float* x, int32_t n, int32_t p, int32_t runs,
                                      const int32_t* run_offsets, const int32_t* landmark_rows,
                                      const int32_t* initial_labels, const int32_t* fold_ids,
                                      const kodama_hip_core_options* options, const float* params, float fallback,
                                      int32_t* clbest, double* accbest, double* vect_acc, double* timings) {
  return guarded([&] {
    require(params != nullptr, "null network parameters.");
    require(fallback >= 0.0f && fallback <= 1.0f, "fallback must be a probability.");
    BatchRequest q{x, n, p, runs, run_offsets, landmark_rows, initial_labels, fold_ids, options};
    run_knn_batch(q, 2, params, fallback, nullptr, clbest, accbest, vect_acc, timings);
  });
}

}  // extern "C"
