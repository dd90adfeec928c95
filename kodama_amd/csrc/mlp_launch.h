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

// Parameters -> the tiled form of the matrix-core forward pass: the A
// operands of v_mfma_f32_16x16x32_bf16 per lane, 8 bf16 (one uint4) each --
// layer 1 [4 unit tiles][2 K blocks] (ops 0..7), layer 2 [4][2] (8..15), layer
// 3 [2 output tiles][2] (16..19), as [op][64 lanes] -- then b1, b2, b3 as fp32.
// Lane l of op (t, b) holds unit 16t + l % 16 against inputs 32b + 8g + e
// (layer 1) or 32b + 4g + e, 32b + 16 + 4g + e - 4 (layers 2 and 3), e = 0..7,
// g = l / 16.
constexpr int kMlpTileOps = 20;
constexpr int kMlpTiled = kMlpTileOps * 64 * 4 + 2 * kMlpHidden + kMlpOut;  // 32-bit words
hipError_t launch_mlp_pack(const float* params, uint32_t* tiled, hipStream_t s);

// Forward pass of `rows` feature rows, 16 rows per matrix-core tile, with
// bf16 weights and inputs and FP32 sums: probabilities [rows x 32] (masked
// entries 0).
hipError_t launch_mlp_forward(const uint32_t* tiled, const float* feat, int rows, float* probs, hipStream_t s);

// One minibatch: per-wave partial gradients of the mean cross-entropy of the
// samples idx[0..count) (targets in 0..30) -> partial [kMlpGradWaves x
// (kMlpParams + 1)] (the last entry the loss sum).
hipError_t launch_mlp_grad(const float* params, const float* feat, const int* targets, const int* idx, int count,
                           float* partial, hipStream_t s);
// Sum of the partials in wave order; grad = sum / count; velocity = momentum *
// velocity + grad; params -= lr * velocity. loss_out (may be null) = mean loss.
hipError_t launch_mlp_sgd(float* params, float* velocity, const float* partial, int count, float lr, float momentum,
                          float* loss_out, hipStream_t s);

}  // namespace kodama_hip
