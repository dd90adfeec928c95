// Learned proposal network (gfx950): the table features, the forward pass,
// and training (backward pass and momentum SGD).
//
// The features and the forward pass used at inference (16-row tiles on the
// matrix cores) live in mlp_device.h, shared with the learned variant of the
// persistent evolution kernel. The MLP is 64 -> 64 (ReLU) -> 64 (ReLU) -> 32
// with bf16-rounded weights and inputs (weights kept as an fp32 master copy)
// and FP32 accumulation; logits of invalid outputs are masked before the
// softmax.
//
// Training runs one wave per sample: lane l owns hidden unit l (and output
// l < 32), the sample's vectors sit in the wave's LDS slice, the fp32 weight
// matrices in the workgroup's LDS (41.6 KB), and lane l keeps row l of every
// gradient matrix in registers, so a wave accumulates its samples without
// atomics; the per-wave partials are summed in a fixed order, so training is
// deterministic for a given sample order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "mlp_device.h"
#include "mlp_launch.h"

namespace kodama_hip {

namespace {

constexpr int kWave = mlp::kWave;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

struct MlpLds {
  float w1[kMlpIn * kMlpHidden];
  float b1[kMlpHidden];
  float w2[kMlpHidden * kMlpHidden];
  float b2[kMlpHidden];
  float w3[kMlpHidden * kMlpOut];
  float b3[kMlpOut];
};
static_assert(sizeof(MlpLds) == sizeof(float) * kMlpParams, "MLP parameter layout");

// Weight matrices rounded to bf16, biases as they are.
__device__ void load_weights(const float* __restrict__ params, MlpLds& w) {
  float* dst = reinterpret_cast<float*>(&w);
  for (int e = threadIdx.x; e < kMlpParams; e += blockDim.x) {
    const bool bias = (e >= kMlpB1 && e < kMlpW2) || (e >= kMlpB2 && e < kMlpW3) || e >= kMlpB3;
    dst[e] = bias ? params[e] : mlp::bf16_round(params[e]);
  }
}

// Forward pass of the sample whose features sit in xs (wave LDS slice); lanes
// < 32 return their probability (0 when masked). h1s / h2s keep the hidden
// activations for the backward pass.
__device__ float forward_wave(const MlpLds& w, const float* __restrict__ feat, float* xs, float* h1s, float* h2s,
                              int lane, float* z_out) {
  xs[lane] = mlp::bf16_round(feat[lane]);
  mlp::wave_sync();
  float a = w.b1[lane];
  for (int i = 0; i < kMlpIn; ++i) a = fmaf(xs[i], w.w1[i * kMlpHidden + lane], a);
  h1s[lane] = fmaxf(a, 0.0f);
  mlp::wave_sync();
  a = w.b2[lane];
  for (int i = 0; i < kMlpHidden; ++i) a = fmaf(h1s[i], w.w2[i * kMlpHidden + lane], a);
  h2s[lane] = fmaxf(a, 0.0f);
  mlp::wave_sync();
  const int o = lane & (kMlpOut - 1);
  float z = w.b3[o];
  for (int i = 0; i < kMlpHidden; ++i) z = fmaf(h2s[i], w.w3[i * kMlpOut + o], z);
  const bool valid = lane < kMlpOut && mlp::output_valid(feat, lane);
  float m = valid ? z : -INFINITY;
#pragma unroll
  for (int off = kWave / 2; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  const float e = valid ? expf(z - m) : 0.0f;
  float sum = e;
#pragma unroll
  for (int off = kWave / 2; off >= 1; off >>= 1) sum += __shfl_xor(sum, off);
  if (z_out) *z_out = z;
  return e / sum;
}

}  // namespace

// Parameters -> the tiled layout of mlp::tile_forward (launch_mlp_pack).
__global__ void __launch_bounds__(kBlock) mlp_pack_kernel(const float* __restrict__ params,
                                                          uint32_t* __restrict__ tiled) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  constexpr int kOpWords = kMlpTileOps * kWave * 4;
  if (e >= kMlpTiled) return;
  if (e < kOpWords) {
    const int op = e / (kWave * 4), lane = (e / 4) % kWave, word = e % 4;
    const int layer = op / 8, t = (op % 8) / 2, b = op % 2;
    const int unit = 16 * t + (lane & 15), g = lane >> 4;
    uint32_t bits[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int el = 2 * word + h;
      float v;
      if (layer == 0) {
        v = params[kMlpW1 + (32 * b + 8 * g + el) * kMlpHidden + unit];
      } else {
        const int in = 32 * b + (el < 4 ? 4 * g + el : 16 + 4 * g + el - 4);
        v = layer == 1 ? params[kMlpW2 + in * kMlpHidden + unit] : params[kMlpW3 + in * kMlpOut + unit];
      }
      bits[h] = mlp::bf16_bits(v);
    }
    tiled[e] = bits[0] | bits[1] << 16;
  } else {
    const int b = e - kOpWords;
    const float v = b < kMlpHidden ? params[kMlpB1 + b]
                                   : (b < 2 * kMlpHidden ? params[kMlpB2 + b - kMlpHidden]
                                                         : params[kMlpB3 + b - 2 * kMlpHidden]);
    tiled[e] = __float_as_uint(v);
  }
}

// ---- features of caller tables (tests, traces) ---------------------------------

// One lane per row: the row vote of mlp_device.h over N >= k neighbours.
template <int N>
__global__ void __launch_bounds__(kBlock) mlp_table_features_kernel(const int* __restrict__ nbr,
                                                                     const int* __restrict__ labels, int rows, int k,
                                                                     float* __restrict__ feat,
                                                                     unsigned short* __restrict__ ranked) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= rows) return;
  uint32_t key[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const int id = j < k ? nbr[static_cast<size_t>(r) * k + j] : -1;
    key[j] = id >= 0 && id < rows ? static_cast<uint32_t>(labels[id]) : mlp::kNone;
  }
  const mlp::RowVote v = mlp::row_vote<N>(key, static_cast<uint32_t>(labels[r]));
  const auto frac = [k](int c) { return __fdiv_rn(static_cast<float>(c), static_cast<float>(k)); };
  float* f = feat + static_cast<size_t>(r) * kMlpIn;
#pragma unroll
  for (int j = 0; j < kMlpIn; ++j) f[j] = mlp::feature<N>(key, v, j, frac);
#pragma unroll
  for (int q = 0; q < kMlpRanks; ++q) {
    ranked[static_cast<size_t>(r) * kMlpRanks + q] =
      static_cast<unsigned short>(q < N && q < v.distinct ? key[q < N ? q : 0] & 0xFFFFu : mlp::kNone);
  }
}

// One tile of 16 rows per wave; the operands come straight from global memory.
__global__ void __launch_bounds__(kBlock) mlp_forward_kernel(const uint32_t* __restrict__ tiled,
                                                              const float* __restrict__ feat, int rows,
                                                              float* __restrict__ probs) {
  const int lane = threadIdx.x & (kWave - 1);
  const int g = lane >> 4;
  const int tiles = (rows + mlp::kTileRows - 1) / mlp::kTileRows;
  const int waves = gridDim.x * kWavesPerBlock;
  for (int tl = blockIdx.x * kWavesPerBlock + static_cast<int>(threadIdx.x) / kWave; tl < tiles; tl += waves) {
    const int row = tl * mlp::kTileRows + (lane & 15);
    uint4 x[2] = {make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)};
    // Rows past the end: zero inputs and only the keep output (never written).
    uint32_t valid = g == 3 ? 1u << 6 : 0u;
    if (row < rows) {
      const float* f = feat + static_cast<size_t>(row) * kMlpIn;
      uint32_t v[2][8];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[b][e] = mlp::bf16_bits(f[32 * b + 8 * g + e]);
        x[b] = make_uint4(v[b][0] | v[b][1] << 16, v[b][2] | v[b][3] << 16, v[b][4] | v[b][5] << 16,
                          v[b][6] | v[b][7] << 16);
      }
      valid = mlp::tile_valid(f, lane);
    }
    float prob[8];
    mlp::tile_forward(tiled, x, valid, lane, prob);
    if (row < rows) {
#pragma unroll
      for (int e = 0; e < 8; ++e) probs[static_cast<size_t>(row) * kMlpOut + mlp::tile_unit(e >> 2, g, e & 3)] = prob[e];
    }
  }
}

__global__ void __launch_bounds__(kBlock) mlp_grad_kernel(const float* __restrict__ params,
                                                           const float* __restrict__ feat,
                                                           const int* __restrict__ targets,
                                                           const int* __restrict__ idx, int count,
                                                           float* __restrict__ partial) {
  __shared__ MlpLds w;
  __shared__ float vec[kWavesPerBlock][6][kWave];
  load_weights(params, w);
  __syncthreads();
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int gw = blockIdx.x * kWavesPerBlock + wave;
  float* xs = vec[wave][0];
  float* h1s = vec[wave][1];
  float* h2s = vec[wave][2];
  float* dzs = vec[wave][3];
  float* dh2s = vec[wave][4];
  float* dh1s = vec[wave][5];
  float g1[kMlpHidden], g2[kMlpHidden], g3[kMlpOut];
#pragma unroll
  for (int j = 0; j < kMlpHidden; ++j) g1[j] = g2[j] = 0.0f;
#pragma unroll
  for (int j = 0; j < kMlpOut; ++j) g3[j] = 0.0f;
  float gb1 = 0.0f, gb2 = 0.0f, gb3 = 0.0f, loss = 0.0f;
  for (int s = gw; s < count; s += kMlpGradWaves) {
    const int sample = idx[s];
    const float* f = feat + static_cast<size_t>(sample) * kMlpIn;
    const int t = targets[sample];
    const float p = forward_wave(w, f, xs, h1s, h2s, lane, nullptr);
    const float pt = __shfl(p, t);
    if (lane == 0) loss += -logf(fmaxf(pt, 1e-30f));
    const float dz = lane < kMlpOut ? p - (lane == t ? 1.0f : 0.0f) : 0.0f;
    if (lane < kMlpOut) dzs[lane] = dz;
    mlp::wave_sync();
    float a = 0.0f;
    for (int j = 0; j < kMlpOut; ++j) a = fmaf(w.w3[lane * kMlpOut + j], dzs[j], a);
    const float h2 = h2s[lane];
    const float dh2 = h2 > 0.0f ? a : 0.0f;
    dh2s[lane] = dh2;
    mlp::wave_sync();
    a = 0.0f;
    for (int j = 0; j < kMlpHidden; ++j) a = fmaf(w.w2[lane * kMlpHidden + j], dh2s[j], a);
    const float h1 = h1s[lane];
    const float dh1 = h1 > 0.0f ? a : 0.0f;
    dh1s[lane] = dh1;
    mlp::wave_sync();
    const float x = xs[lane];
#pragma unroll
    for (int j = 0; j < kMlpHidden; ++j) {
      g1[j] = fmaf(x, dh1s[j], g1[j]);
      g2[j] = fmaf(h1, dh2s[j], g2[j]);
    }
#pragma unroll
    for (int j = 0; j < kMlpOut; ++j) g3[j] = fmaf(h2, dzs[j], g3[j]);
    gb1 += dh1;
    gb2 += dh2;
    gb3 += dz;
    mlp::wave_sync();
  }
  float* out = partial + static_cast<size_t>(gw) * (kMlpParams + 1);
#pragma unroll
  for (int j = 0; j < kMlpHidden; ++j) {
    out[kMlpW1 + lane * kMlpHidden + j] = g1[j];
    out[kMlpW2 + lane * kMlpHidden + j] = g2[j];
  }
#pragma unroll
  for (int j = 0; j < kMlpOut; ++j) out[kMlpW3 + lane * kMlpOut + j] = g3[j];
  out[kMlpB1 + lane] = gb1;
  out[kMlpB2 + lane] = gb2;
  if (lane < kMlpOut) out[kMlpB3 + lane] = gb3;
  if (lane == 0) out[kMlpParams] = loss;
}

__global__ void __launch_bounds__(kBlock) mlp_sgd_kernel(float* __restrict__ params, float* __restrict__ velocity,
                                                          const float* __restrict__ partial, int count, float lr,
                                                          float momentum, float* __restrict__ loss_out) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e > kMlpParams) return;
  float g = 0.0f;
  for (int w = 0; w < kMlpGradWaves; ++w) g += partial[static_cast<size_t>(w) * (kMlpParams + 1) + e];
  g /= static_cast<float>(count);
  if (e == kMlpParams) {
    if (loss_out) *loss_out = g;
    return;
  }
  const float v = fmaf(momentum, velocity[e], g);
  velocity[e] = v;
  params[e] = fmaf(-lr, v, params[e]);
}

// ---- launchers ----------------------------------------------------------------

hipError_t launch_mlp_features(const int* nbr, const int* labels, int rows, int k, float* feat,
                               unsigned short* ranked, hipStream_t s) {
  if (k < 1 || k > kWave) return hipErrorInvalidValue;
  if (rows == 0) return hipSuccess;
  const dim3 grid((rows + kBlock - 1) / kBlock);
  if (k <= 8) {
    hipLaunchKernelGGL(mlp_table_features_kernel<8>, grid, dim3(kBlock), 0, s, nbr, labels, rows, k, feat, ranked);
  } else if (k <= 16) {
    hipLaunchKernelGGL(mlp_table_features_kernel<16>, grid, dim3(kBlock), 0, s, nbr, labels, rows, k, feat, ranked);
  } else if (k <= 32) {
    hipLaunchKernelGGL(mlp_table_features_kernel<32>, grid, dim3(kBlock), 0, s, nbr, labels, rows, k, feat, ranked);
  } else {
    hipLaunchKernelGGL(mlp_table_features_kernel<64>, grid, dim3(kBlock), 0, s, nbr, labels, rows, k, feat, ranked);
  }
  return hipGetLastError();
}

hipError_t launch_mlp_pack(const float* params, uint32_t* tiled, hipStream_t s) {
  hipLaunchKernelGGL(mlp_pack_kernel, dim3((kMlpTiled + kBlock - 1) / kBlock), dim3(kBlock), 0, s, params, tiled);
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const uint32_t* tiled, const float* feat, int rows, float* probs, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  const int tiles = (rows + mlp::kTileRows - 1) / mlp::kTileRows;
  const int blocks = (tiles + kWavesPerBlock - 1) / kWavesPerBlock;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks < 1024 ? blocks : 1024), dim3(kBlock), 0, s, tiled, feat, rows,
                     probs);
  return hipGetLastError();
}

hipError_t launch_mlp_grad(const float* params, const float* feat, const int* targets, const int* idx, int count,
                           float* partial, hipStream_t s) {
  if (count < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(kMlpGradWaves / kWavesPerBlock), dim3(kBlock), 0, s, params, feat,
                     targets, idx, count, partial);
  return hipGetLastError();
}

hipError_t launch_mlp_sgd(float* params, float* velocity, const float* partial, int count, float lr, float momentum,
                          float* loss_out, hipStream_t s) {
  if (count < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mlp_sgd_kernel, dim3((kMlpParams + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, params,
                     velocity, partial, count, lr, momentum, loss_out);
  return hipGetLastError();
}

}  // namespace kodama_hip
