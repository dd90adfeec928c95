// Learned proposal network of the device-resident evolution (gfx950).
//
// Features (mlp_launch.h): a row's k neighbour labels are sorted (bitonic
// network in registers), run lengths give each voted label's count, and a
// second network orders (count desc, code asc). The MLP is 64 -> 64 (ReLU) ->
// 64 (ReLU) -> 32 with bf16-rounded weights and inputs (weights kept as an
// fp32 master copy) and FP32 accumulation; logits of invalid outputs are
// masked before the softmax.
//
// One wave evaluates one sample: lane l owns hidden unit l (and output l <
// 32), the sample's vectors sit in the wave's LDS slice and every weight
// matrix in the workgroup's LDS (41.6 KB). For training, lane l also owns row
// l of every gradient matrix in registers, so a wave accumulates its samples
// without atomics and the per-wave partials are summed in a fixed order:
// training is deterministic for a given sample order. Batches of a few
// thousand rows per cycle leave MFMA tiles nothing to amortise; the network
// is latency-bound, and the wave-per-sample form keeps it to a few
// microseconds per cycle.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "mlp_launch.h"

namespace kodama_hip {

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kNone = 0xFFFFu;

__device__ __forceinline__ float bf16_round(float x) {
  // Round to nearest even (finite inputs).
  const uint32_t b = __float_as_uint(x);
  return __uint_as_float((b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u);
}

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
    dst[e] = bias ? params[e] : bf16_round(params[e]);
  }
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <int N>
__device__ __forceinline__ void sort_u32(uint32_t (&a)[N]) {
#pragma unroll
  for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int j = i ^ stride;
        if (j > i) {
          const uint32_t lo = min(a[i], a[j]), hi = max(a[i], a[j]);
          const bool up = (i & size) == 0;
          a[i] = up ? lo : hi;
          a[j] = up ? hi : lo;
        }
      }
    }
  }
}

// Features of one row from its neighbour codes (kNone = missing) and its own
// code; writes feat[64] and ranked[30].
template <int KMAX>
__device__ void row_features(uint32_t (&code)[KMAX], int k, uint32_t own, float* __restrict__ feat,
                             unsigned short* __restrict__ ranked) {
  sort_u32<KMAX>(code);
  int own_count = 0;
#pragma unroll
  for (int j = 0; j < KMAX; ++j) own_count += code[j] == own ? 1 : 0;
  // Run length from each position to the end of its run, then one key per
  // distinct voted label: (255 - count) << 16 | code.
  uint32_t len[KMAX];
  len[KMAX - 1] = 1;
#pragma unroll
  for (int j = KMAX - 2; j >= 0; --j) len[j] = code[j + 1] == code[j] ? len[j + 1] + 1 : 1;
  uint32_t key[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    const bool start = j == 0 || code[j] != code[j - 1];
    key[j] = start && code[j] != kNone ? ((255u - len[j]) << 16) | code[j] : 0xFFFFFFFFu;
  }
  sort_u32<KMAX>(key);
  int own_rank = kMlpRanks;
  int top = 0;
#pragma unroll
  for (int j = 0; j < kMlpRanks; ++j) {
    const uint32_t kj = j < KMAX ? key[j] : 0xFFFFFFFFu;  // j is a constant after unrolling
    const bool real = kj != 0xFFFFFFFFu;
    const int cnt = real ? static_cast<int>(255u - (kj >> 16)) : 0;
    const uint32_t c = real ? (kj & 0xFFFFu) : kNone;
    if (j == 0) top = cnt;
    feat[j] = real ? __fdiv_rn(static_cast<float>(cnt), static_cast<float>(k)) : 0.0f;
    ranked[j] = static_cast<unsigned short>(c);
    if (real && c == own && own_rank == kMlpRanks) own_rank = j;
  }
  for (int j = kMlpRanks; j < kMlpIn; ++j) feat[j] = 0.0f;
  feat[kMlpRanks + own_rank] = 1.0f;
  feat[61] = __fdiv_rn(static_cast<float>(top - own_count), static_cast<float>(k));
  feat[62] = 1.0f;
}

// Output validity: keep, or an existing rank that is not the row's own label.
__device__ __forceinline__ bool output_valid(const float* __restrict__ feat, int j) {
  if (j == kMlpKeep) return true;
  if (j >= kMlpRanks) return false;
  return feat[j] > 0.0f && feat[kMlpRanks + j] == 0.0f;
}

// Forward pass of the sample whose features sit in xs (wave LDS slice); lanes
// < 32 return their probability (0 when masked). h1s / h2s keep the hidden
// activations for the backward pass.
__device__ float forward_wave(const MlpLds& w, const float* __restrict__ feat, float* xs, float* h1s, float* h2s,
                              int lane, float* z_out) {
  xs[lane] = bf16_round(feat[lane]);
  wave_sync();
  float a = w.b1[lane];
  for (int i = 0; i < kMlpIn; ++i) a = fmaf(xs[i], w.w1[i * kMlpHidden + lane], a);
  h1s[lane] = fmaxf(a, 0.0f);
  wave_sync();
  a = w.b2[lane];
  for (int i = 0; i < kMlpHidden; ++i) a = fmaf(h1s[i], w.w2[i * kMlpHidden + lane], a);
  h2s[lane] = fmaxf(a, 0.0f);
  wave_sync();
  const int o = lane & (kMlpOut - 1);
  float z = w.b3[o];
  for (int i = 0; i < kMlpHidden; ++i) z = fmaf(h2s[i], w.w3[i * kMlpOut + o], z);
  const bool valid = lane < kMlpOut && output_valid(feat, lane);
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

// The forward pass with the weights held in registers (the scoring and test
// kernels): lane j keeps column j of each matrix as bf16 pairs (rows 2p and
// 2p + 1 in one register), so each layer is 64 LDS-broadcast multiply-adds
// per lane with no weight traffic. Same operation order as forward_wave.
struct PackedWeights {
  uint32_t w1[kMlpIn / 2], w2[kMlpHidden / 2], w3[kMlpHidden / 2];
  float b1, b2, b3;

  __device__ void load(const uint32_t* __restrict__ packed, int lane) {
#pragma unroll
    for (int q = 0; q < kMlpIn / 2; ++q) w1[q] = packed[(0 * 32 + q) * kWave + lane];
#pragma unroll
    for (int q = 0; q < kMlpHidden / 2; ++q) w2[q] = packed[(1 * 32 + q) * kWave + lane];
#pragma unroll
    for (int q = 0; q < kMlpHidden / 2; ++q) w3[q] = packed[(2 * 32 + q) * kWave + lane];
    b1 = __uint_as_float(packed[3 * 32 * kWave + lane]);
    b2 = __uint_as_float(packed[3 * 32 * kWave + kMlpHidden + lane]);
    b3 = __uint_as_float(packed[3 * 32 * kWave + 2 * kMlpHidden + (lane & (kMlpOut - 1))]);
  }
};

__device__ __forceinline__ float lo_bf16(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float hi_bf16(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }

__device__ float forward_packed(const PackedWeights& w, const float* __restrict__ feat, float* xs, float* h1s,
                                float* h2s, int lane) {
  xs[lane] = bf16_round(feat[lane]);
  wave_sync();
  float a = w.b1;
#pragma unroll
  for (int q = 0; q < kMlpIn / 2; ++q) {
    a = fmaf(xs[2 * q], lo_bf16(w.w1[q]), a);
    a = fmaf(xs[2 * q + 1], hi_bf16(w.w1[q]), a);
  }
  h1s[lane] = fmaxf(a, 0.0f);
  wave_sync();
  a = w.b2;
#pragma unroll
  for (int q = 0; q < kMlpHidden / 2; ++q) {
    a = fmaf(h1s[2 * q], lo_bf16(w.w2[q]), a);
    a = fmaf(h1s[2 * q + 1], hi_bf16(w.w2[q]), a);
  }
  h2s[lane] = fmaxf(a, 0.0f);
  wave_sync();
  float z = w.b3;
#pragma unroll
  for (int q = 0; q < kMlpHidden / 2; ++q) {
    z = fmaf(h2s[2 * q], lo_bf16(w.w3[q]), z);
    z = fmaf(h2s[2 * q + 1], hi_bf16(w.w3[q]), z);
  }
  const bool valid = lane < kMlpOut && output_valid(feat, lane);
  float m = valid ? z : -INFINITY;
#pragma unroll
  for (int off = kWave / 2; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  const float e = valid ? expf(z - m) : 0.0f;
  float sum = e;
#pragma unroll
  for (int off = kWave / 2; off >= 1; off >>= 1) sum += __shfl_xor(sum, off);
  return e / sum;
}

}  // namespace

// Parameters -> the packed layout of PackedWeights: [3][32][64] bf16 pairs
// (matrix, row pair, column), then b1, b2, b3 as fp32 bits.
__global__ void __launch_bounds__(kBlock) mlp_pack_kernel(const float* __restrict__ params,
                                                          uint32_t* __restrict__ packed) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  constexpr int kPairs = 3 * 32 * kWave;
  if (e >= kMlpPacked) return;
  if (e < kPairs) {
    const int m = e / (32 * kWave), q = (e / kWave) % 32, j = e % kWave;
    float lo = 0.0f, hi = 0.0f;
    if (m == 0) {
      lo = params[kMlpW1 + (2 * q) * kMlpHidden + j];
      hi = params[kMlpW1 + (2 * q + 1) * kMlpHidden + j];
    } else if (m == 1) {
      lo = params[kMlpW2 + (2 * q) * kMlpHidden + j];
      hi = params[kMlpW2 + (2 * q + 1) * kMlpHidden + j];
    } else if (j < kMlpOut) {
      lo = params[kMlpW3 + (2 * q) * kMlpOut + j];
      hi = params[kMlpW3 + (2 * q + 1) * kMlpOut + j];
    }
    packed[e] = (__float_as_uint(bf16_round(lo)) >> 16) | (__float_as_uint(bf16_round(hi)) & 0xFFFF0000u);
  } else {
    const int b = e - kPairs;
    const float v = b < kMlpHidden ? params[kMlpB1 + b]
                                   : (b < 2 * kMlpHidden ? params[kMlpB2 + b - kMlpHidden]
                                                         : params[kMlpB3 + b - 2 * kMlpHidden]);
    packed[e] = __float_as_uint(v);
  }
}

// ---- features of caller tables (tests, traces) ---------------------------------

template <int KMAX>
__global__ void __launch_bounds__(kBlock) mlp_table_features_kernel(const int* __restrict__ nbr,
                                                                     const int* __restrict__ labels, int rows, int k,
                                                                     float* __restrict__ feat,
                                                                     unsigned short* __restrict__ ranked) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= rows) return;
  uint32_t code[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    const int id = j < k ? nbr[static_cast<size_t>(r) * k + j] : -1;
    code[j] = id >= 0 && id < rows ? static_cast<uint32_t>(labels[id]) : kNone;
  }
  row_features<KMAX>(code, k, static_cast<uint32_t>(labels[r]), feat + static_cast<size_t>(r) * kMlpIn,
                     ranked + static_cast<size_t>(r) * kMlpRanks);
}

__global__ void __launch_bounds__(kBlock) mlp_forward_kernel(const uint32_t* __restrict__ packed,
                                                              const float* __restrict__ feat, int rows,
                                                              float* __restrict__ probs) {
  __shared__ float vec[kWavesPerBlock][3][kWave];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  PackedWeights w;
  w.load(packed, lane);
  const int waves = gridDim.x * kWavesPerBlock;
  for (int s = blockIdx.x * kWavesPerBlock + wave; s < rows; s += waves) {
    const float* f = feat + static_cast<size_t>(s) * kMlpIn;
    const float p = forward_packed(w, f, vec[wave][0], vec[wave][1], vec[wave][2], lane);
    if (lane < kMlpOut) probs[static_cast<size_t>(s) * kMlpOut + lane] = p;
    wave_sync();
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
    wave_sync();
    float a = 0.0f;
    for (int j = 0; j < kMlpOut; ++j) a = fmaf(w.w3[lane * kMlpOut + j], dzs[j], a);
    const float h2 = h2s[lane];
    const float dh2 = h2 > 0.0f ? a : 0.0f;
    dh2s[lane] = dh2;
    wave_sync();
    a = 0.0f;
    for (int j = 0; j < kMlpHidden; ++j) a = fmaf(w.w2[lane * kMlpHidden + j], dh2s[j], a);
    const float h1 = h1s[lane];
    const float dh1 = h1 > 0.0f ? a : 0.0f;
    dh1s[lane] = dh1;
    wave_sync();
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
    wave_sync();
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

// ---- the learned proposal inside the evolution -----------------------------------

// Features of the listed candidates, one thread each (their gathers overlap
// across the many threads in flight); also zeroes the counter the next
// persistent launch appends to.
template <int KMAX>
__global__ void __launch_bounds__(kBlock) mlp_candidates_kernel(MlpScoreParams p) {
  const int count = *p.cand_count;
  if (blockIdx.x == 0 && threadIdx.x == 0) *p.next_count = 0;
  const int stride = gridDim.x * kBlock;
  for (int c = blockIdx.x * kBlock + threadIdx.x; c < count; c += stride) {
    const int row = p.cand_rows[c];
    const int run = p.block_run[row / kBlock];
    const int base = p.pad_offsets[run];
    const int n = p.run_offsets[run + 1] - p.run_offsets[run];
    uint32_t code[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      const int id = j < p.k ? static_cast<int>(p.neighbors[static_cast<size_t>(j) * p.pad_total + row]) : n;
      code[j] = id < n ? static_cast<uint32_t>(p.labels[base + id]) : kNone;
    }
    row_features<KMAX>(code, p.k, p.labels[row], p.feat + static_cast<size_t>(c) * kMlpIn,
                       p.cand_ranked + static_cast<size_t>(c) * kMlpRanks);
  }
}

// The network over the candidates, one wave per candidate, weights in
// registers: score = 1 - p(keep), target = the label of the most probable
// valid rank (score 0 when no rank is valid).
__global__ void __launch_bounds__(kBlock) mlp_score_kernel(MlpScoreParams p) {
  __shared__ float vec[kWavesPerBlock][3][kWave];
  const int count = *p.cand_count;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (static_cast<int>(blockIdx.x) * kWavesPerBlock + wave >= count) return;  // whole wave idle; no block barrier
  PackedWeights w;
  w.load(p.packed, lane);
  const int waves = gridDim.x * kWavesPerBlock;
  for (int c = blockIdx.x * kWavesPerBlock + wave; c < count; c += waves) {
    const float* feat = p.feat + static_cast<size_t>(c) * kMlpIn;
    const float prob = forward_packed(w, feat, vec[wave][0], vec[wave][1], vec[wave][2], lane);
    // Most probable valid relabel rank (smallest rank on ties).
    const bool rank_ok = lane < kMlpRanks && output_valid(feat, lane);
    uint64_t key = rank_ok ? (static_cast<uint64_t>(__float_as_uint(prob)) << 32) |
                               static_cast<uint32_t>(kMlpRanks - lane)
                           : 0ull;
#pragma unroll
    for (int off = kWave / 2; off >= 1; off >>= 1) {
      const uint64_t other = __shfl_xor(key, off);
      key = other > key ? other : key;
    }
    const float keep = __shfl(prob, kMlpKeep);
    if (lane == 0) {
      const int row = p.cand_rows[c];
      if (key != 0ull) {
        const int best = kMlpRanks - static_cast<int>(key & 0xFFFFFFFFull);
        p.score[row] = 1.0f - keep;
        p.target[row] = p.cand_ranked[static_cast<size_t>(c) * kMlpRanks + best];
      } else {
        p.score[row] = 0.0f;
      }
    }
    wave_sync();
  }
}

// ---- launchers ----------------------------------------------------------------

hipError_t launch_mlp_features(const int* nbr, const int* labels, int rows, int k, float* feat,
                               unsigned short* ranked, hipStream_t s) {
  if (k < 1 || k > 64) return hipErrorInvalidValue;
  if (rows == 0) return hipSuccess;
  const dim3 grid((rows + kBlock - 1) / kBlock), block(kBlock);
  if (k <= 8) hipLaunchKernelGGL((mlp_table_features_kernel<8>), grid, block, 0, s, nbr, labels, rows, k, feat, ranked);
  else if (k <= 16) hipLaunchKernelGGL((mlp_table_features_kernel<16>), grid, block, 0, s, nbr, labels, rows, k, feat, ranked);
  else if (k <= 32) hipLaunchKernelGGL((mlp_table_features_kernel<32>), grid, block, 0, s, nbr, labels, rows, k, feat, ranked);
  else hipLaunchKernelGGL((mlp_table_features_kernel<64>), grid, block, 0, s, nbr, labels, rows, k, feat, ranked);
  return hipGetLastError();
}

hipError_t launch_mlp_pack(const float* params, uint32_t* packed, hipStream_t s) {
  hipLaunchKernelGGL(mlp_pack_kernel, dim3((kMlpPacked + kBlock - 1) / kBlock), dim3(kBlock), 0, s, params, packed);
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const uint32_t* packed, const float* feat, int rows, float* probs, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  const int blocks = (rows + kWavesPerBlock - 1) / kWavesPerBlock;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks < 1024 ? blocks : 1024), dim3(kBlock), 0, s, packed, feat, rows,
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

hipError_t launch_mlp_candidates(const MlpScoreParams& p, hipStream_t s) {
  if (p.k < 1 || p.k > 64 || p.pad_total % kBlock != 0) return hipErrorInvalidValue;
  const dim3 grid(512), block(kBlock);
  if (p.k <= 8) hipLaunchKernelGGL((mlp_candidates_kernel<8>), grid, block, 0, s, p);
  else if (p.k <= 16) hipLaunchKernelGGL((mlp_candidates_kernel<16>), grid, block, 0, s, p);
  else if (p.k <= 32) hipLaunchKernelGGL((mlp_candidates_kernel<32>), grid, block, 0, s, p);
  else hipLaunchKernelGGL((mlp_candidates_kernel<64>), grid, block, 0, s, p);
  return hipGetLastError();
}

hipError_t launch_mlp_score(const MlpScoreParams& p, hipStream_t s) {
  if (p.k < 1 || p.k > kWave || p.pad_total % kBlock != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mlp_score_kernel, dim3(1024), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

}  // namespace kodama_hip
