// Host side of the learned proposal (mlp.cpp): trace collection from the plain
// evolution, the training driver over the device kernels, and the host twin
// of the row features.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

#include "engine.h"
#include "mlp_launch.h"

namespace kodama_hip {

// Features of one row (the layout of mlp_launch.h) from its k neighbour codes
// (-1 = missing) and its own code; ranked[30] gets the voted codes by rank
// (-1 past them). Identical to the device kernel's result.
void host_row_features(const int* codes, int k, int own, float* feat, int* ranked);

// Accept/reject samples of the plain evolution: for every row a cycle's
// proposal relabels (a sampled singleton whose prediction differs from its
// label), its features under the labels the proposal started from, and the
// target: the rank of its new label when the proposal became the current
// state, keep (30) otherwise.
// With outcome targets every row whose label differs from its prediction is
// a sample at every cycle (features under the current labels), and its target
// is the rank of the label the row holds in the run's final best labelling
// (keep when that is its current label; rows whose final label is not among
// the voted ranks are not recorded): the network then learns which rows end
// up relabelled, and to which label, rather than the proposal-level accept
// flag that every row of a proposal shares.
struct TraceSink {
  std::vector<float> feat;    // [samples x 64]
  std::vector<int> target;    // [samples]
  int max_samples = 0;        // stop recording beyond this many
  bool outcome = false;       // outcome targets instead of accept / reject
};

// The plain batched evolution (evolve_plain) with trace recording; runs are
// recorded in run order within each cycle, so the samples are deterministic.
void evolve_plain_traced(std::vector<RunEvolution>& runs, KnnBatch& batch, const std::vector<IVec>& folds,
                         int cycles, int n_threads, TraceSink& sink);

// Minibatch SGD with momentum on the device: `epochs` passes over the samples
// in an order drawn from mt19937_64(seed + epoch); params [kMlpParams] fp32
// in/out; loss_per_epoch [epochs] (may be null) = mean minibatch loss.
void mlp_train(int device, const float* feat, const int* target, int samples, int epochs, int batch, float lr,
               float momentum, std::uint64_t seed, float* params, double* loss_per_epoch);

// He-normal weights (std::normal_distribution over mt19937_64(seed)), zero biases.
void mlp_init(std::uint64_t seed, float* params);

}  // namespace kodama_hip
