// The learned proposal of the device-resident evolution (mlp_kernels.hip): a
// 64-64-64-32 MLP over per-row neighbour-vote features that scores which rows
// to relabel and to which of their voted labels.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace kodama_hip {

// Features of a row (64 floats): [0, 30) vote counts of its 30 most voted
// labels (count desc, code asc) over k; [30, 61) one-hot of its own label's
// rank (60: not among them); 61 the margin (top count - own count) / k;
// 62 a constant 1; 63 zero. Outputs (32 logits): j < 30 relabel to the label
// of rank j, 30 keep, 31 unused; rank j is valid when it exists and is not
// the row's own label.
constexpr int kMlpIn = 64;
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 32;
constexpr int kMlpRanks = 30;
constexpr int kMlpKeep = 30;
// Parameter block (fp32 master): W1 [in x hidden], b1, W2 [hidden x hidden],
// b2, W3 [hidden x out], b3; W is row-major [input][output].
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpIn * kMlpHidden;
constexpr int kMlpW2 = kMlpB1 + kMlpHidden;
constexpr int kMlpB2 = kMlpW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpW3 = kMlpB2 + kMlpHidden;
constexpr int kMlpB3 = kMlpW3 + kMlpHidden * kMlpOut;
constexpr int kMlpParams = kMlpB3 + kMlpOut;  // 10400
constexpr int kMlpGradWaves = 64;             // waves per training step (fixed reduction order)

// Features of the rows of one table: nbr [rows x k] local ids (-1 missing),
// labels [rows] codes. feat [rows x 64], ranked [rows x 30] codes (0xFFFF past
// the voted labels).
hipError_t launch_mlp_features(const int* nbr, const int* labels, int rows, int k, float* feat,
                               unsigned short* ranked, hipStream_t s);

// Parameters -> the per-lane packed form the scoring and forward kernels keep
// in registers: [3][32][64] bf16 pairs (rows 2q and 2q + 1 of column j of W1,
// W2, W3), then b1, b2, b3 as fp32 bits.
constexpr int kMlpPacked = 3 * 32 * 64 + 2 * kMlpHidden + kMlpOut;
hipError_t launch_mlp_pack(const float* params, uint32_t* packed, hipStream_t s);

// Forward pass of `rows` feature rows with bf16-rounded weights and inputs and
// FP32 accumulation: probabilities [rows x 32] (masked entries 0).
hipError_t launch_mlp_forward(const uint32_t* packed, const float* feat, int rows, float* probs, hipStream_t s);

// One minibatch: per-wave partial gradients of the mean cross-entropy of the
// samples idx[0..count) (targets in 0..30) -> partial [kMlpGradWaves x
// (kMlpParams + 1)] (the last entry the loss sum).
hipError_t launch_mlp_grad(const float* params, const float* feat, const int* targets, const int* idx, int count,
                           float* partial, hipStream_t s);
// Sum of the partials in wave order; grad = sum / count; velocity = momentum *
// velocity + grad; params -= lr * velocity. loss_out (may be null) = mean loss.
hipError_t launch_mlp_sgd(float* params, float* velocity, const float* partial, int count, float lr, float momentum,
                          float* loss_out, hipStream_t s);

// Per cycle of the learned device evolution: the persistent kernel lists the
// rows whose label differs from their prediction (cand_rows, *cand_count);
// launch_mlp_candidates computes their features (one thread each) and zeroes
// *next_count, the counter the next persistent launch appends to;
// launch_mlp_score runs the network (one wave per row): score = 1 - p(keep),
// target = the label of the most probable valid rank (score 0 when no rank is
// valid).
struct MlpScoreParams {
  const uint32_t* packed = nullptr;           // launch_mlp_pack of the parameters
  const unsigned short* neighbors = nullptr;  // [k x pad_total] run-local ids (n = none)
  const unsigned short* labels = nullptr;     // [pad_total] current codes
  const int* run_offsets = nullptr;
  const int* pad_offsets = nullptr;
  const int* block_run = nullptr;             // [pad_total / 256]
  int pad_total = 0;
  int k = 0;
  const int* cand_rows = nullptr;             // [pad_total] padded rows of the candidates
  const int* cand_count = nullptr;
  int* next_count = nullptr;
  float* feat = nullptr;                      // [pad_total x 64] candidate features
  unsigned short* cand_ranked = nullptr;      // [pad_total x 30]
  float* score = nullptr;                     // [pad_total]
  unsigned short* target = nullptr;           // [pad_total]
};
hipError_t launch_mlp_candidates(const MlpScoreParams& p, hipStream_t s);
hipError_t launch_mlp_score(const MlpScoreParams& p, hipStream_t s);

}  // namespace kodama_hip
