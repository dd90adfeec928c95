// Device code of the learned proposal shared by mlp_kernels.hip (tests,
// training, the table features) and the learned variant of the persistent
// evolution kernel (evolve_kernels.hip): the neighbour vote of a row computed
// by one lane (from which its features follow), and the forward pass of a
// tile of 16 rows on the matrix cores.
//
// Tile forward pass: every layer is D^T = W^T X^T on v_mfma_f32_16x16x32_bf16
// with the weights as the A operand (lane l: unit 16t + l % 16 of unit tile t)
// and the tile's rows as the B operand (lane l: row l % 16), so lane l ends a
// layer holding units 16t + 4 (l / 16) + r (r = 0..3) of row l % 16. The next
// layer takes exactly these values as its B operand: its K index is permuted
// so that lane group g = l / 16 of K block b holds units 32b + 4g + r and
// 32b + 16 + 4g + r, and the packed weights use the same permutation -- no
// exchange of activations between layers. Inputs and weights are bf16 (the
// network's definition); hidden activations (FP32) enter as three bf16 parts
// hi + mid + lo, which leaves a residual below 2^-24 of each activation, and
// the bf16 x bf16 products are exact in FP32, so the tile computes the FP32
// network up to the order of its FP32 sums.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "mlp_launch.h"

namespace kodama_hip {
namespace mlp {

constexpr int kWave = 64;
constexpr uint32_t kNone = 0xFFFFu;  // code of a missing neighbour
constexpr int kTileRows = 16;        // rows of one tile forward pass

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf16_round(float x) {
  // Round to nearest even (finite inputs).
  const uint32_t b = __float_as_uint(x);
  return __uint_as_float((b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u);
}
__device__ __forceinline__ uint32_t bf16_bits(float x) { return __float_as_uint(bf16_round(x)) >> 16; }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Ascending bitonic sort of N (a power of two) values held in registers. A
// scheduling barrier after each group of four compare-exchanges keeps the
// sort in place (N + 8 live values); unconstrained interleaving of a stage's
// exchanges doubles the registers.
template <int N>
__device__ __forceinline__ void bitonic_sort(uint32_t (&a)[N]) {
#pragma unroll
  for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      int done = 0;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int j = i ^ stride;
        if (j > i) {
          const uint32_t x = a[i], y = a[j];
          const uint32_t lo = min(x, y), hi = max(x, y);
          a[i] = (i & size) == 0 ? lo : hi;
          a[j] = (i & size) == 0 ? hi : lo;
          if (++done % 4 == 0) __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
  }
}

// The neighbour vote of one row, computed by one lane: a[] holds the codes of
// its neighbours (kNone when missing; N >= k entries). Sorting the codes gives
// each voted code's count as a run length; the run ends become keys
// (255 - count) << 16 | code, in place, and sorting the keys ranks the voted
// codes by count desc, code asc. On return a[r] is the key of rank r
// (0xFFFFFFFF past the voted codes).
struct RowVote {
  int distinct;   // voted codes
  int own_rank;   // rank of the row's own label if voted and < 30, else 30
  int own_count;  // votes for the row's own label
  int top;        // votes of rank 0 (0 when nothing is voted)
};
template <int N>
__device__ __forceinline__ RowVote row_vote(uint32_t (&a)[N], uint32_t own) {
  bitonic_sort<N>(a);
  RowVote v{0, kMlpRanks, 0, 0};
  int start = 0;
  uint32_t prev = 0xFFFFFFFFu, cur = a[0];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const uint32_t next = j + 1 < N ? a[j + 1] : 0xFFFFFFFFu;
    if (cur != prev) start = j;
    const int len = j - start + 1;
    const bool voted = cur != next && cur != kNone;
    a[j] = voted ? (static_cast<uint32_t>(255 - len) << 16) | cur : 0xFFFFFFFFu;
    v.distinct += voted ? 1 : 0;
    v.own_count = voted && cur == own ? len : v.own_count;
    prev = cur;
    cur = next;
  }
  bitonic_sort<N>(a);
  v.top = a[0] != 0xFFFFFFFFu ? 255 - static_cast<int>(a[0] >> 16) : 0;
#pragma unroll
  for (int r = 0; r < (N < kMlpRanks ? N : kMlpRanks); ++r) {
    if (a[r] != 0xFFFFFFFFu && (a[r] & 0xFFFFu) == own) v.own_rank = r;
  }
  return v;
}

// Feature j of a row from its vote (see mlp_launch.h), through frac(c) =
// c / k rounded to FP32 (c = 0..64).
template <int N, class Frac>
__device__ __forceinline__ float feature(const uint32_t (&key)[N], const RowVote& v, int j, Frac frac) {
  if (j < kMlpRanks) return j < N && j < v.distinct ? frac(255 - static_cast<int>(key[j < N ? j : 0] >> 16)) : 0.0f;
  if (j <= 2 * kMlpRanks) return j - kMlpRanks == v.own_rank ? 1.0f : 0.0f;
  if (j == 61) return frac(v.top - v.own_count);
  return j == 62 ? 1.0f : 0.0f;
}

// Output validity: keep, or an existing rank that is not the row's own label.
__device__ __forceinline__ bool output_valid(const float* feat, int j) {
  if (j == kMlpKeep) return true;
  if (j >= kMlpRanks) return false;
  return feat[j] > 0.0f && feat[kMlpRanks + j] == 0.0f;
}
// The same on a feature row held as bf16 bits (features are >= 0).
__device__ __forceinline__ bool output_valid(const unsigned short* feat, int j) {
  if (j == kMlpKeep) return true;
  if (j >= kMlpRanks) return false;
  return feat[j] != 0 && feat[kMlpRanks + j] == 0;
}

// Unit (or output) held by lane group g in element r of tile t.
__device__ __forceinline__ int tile_unit(int t, int g, int r) { return 16 * t + 4 * g + r; }

// Validity bits of this lane's 8 outputs (bit 4t + r: output tile_unit(t, g, r)).
template <class F>
__device__ __forceinline__ uint32_t tile_valid(const F* row_feat, int lane) {
  const int g = lane >> 4;
  uint32_t bits = 0;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) bits |= output_valid(row_feat, tile_unit(t, g, r)) ? 1u << (4 * t + r) : 0u;
  }
  return bits;
}

__device__ __forceinline__ f32x4 mfma(uint4 a, uint4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}

// Keeps the weight loads of one K block next to their MFMAs: neither the IR
// passes (memory clobber) nor the machine scheduler may hoist the next loads
// above it. Hoisting all of them would hold 80 registers.
__device__ __forceinline__ void layer_barrier() {
  __asm__ volatile("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
}

// Three bf16 parts of the ReLU of 8 activations, packed as B operands.
__device__ __forceinline__ void split3(const f32x4& u, const f32x4& v, uint4& hi, uint4& mid, uint4& lo) {
  uint32_t ph[8], pm[8], pl[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const float a = fmaxf(e < 4 ? u[e] : v[e - 4], 0.0f);
    const float h = bf16_round(a);
    const float r1 = a - h;  // exact
    const float m = bf16_round(r1);
    ph[e] = __float_as_uint(h) >> 16;
    pm[e] = __float_as_uint(m) >> 16;
    pl[e] = bf16_bits(r1 - m);  // r1 - m exact
  }
  hi = make_uint4(ph[0] | ph[1] << 16, ph[2] | ph[3] << 16, ph[4] | ph[5] << 16, ph[6] | ph[7] << 16);
  mid = make_uint4(pm[0] | pm[1] << 16, pm[2] | pm[3] << 16, pm[4] | pm[5] << 16, pm[6] | pm[7] << 16);
  lo = make_uint4(pl[0] | pl[1] << 16, pl[2] | pl[3] << 16, pl[4] | pl[5] << 16, pl[6] | pl[7] << 16);
}

// Forward pass of one tile. ops: the tiled weights of launch_mlp_pack (uint4
// operands [kMlpTileOps][64], then the biases as fp32); x: this lane's B
// operands of the input (row l % 16, features 32b + 8g .. + 7 as bf16, b = 0,
// 1); valid: tile_valid of the lane's row. Returns the lane's 8 probabilities
// (0 where masked; element 4t + r is output tile_unit(t, g, r)).
template <class P>
__device__ __forceinline__ void tile_forward(const P* ops, const uint4 (&x)[2], uint32_t valid, int lane,
                                             float (&prob)[8]) {
  const uint4* w = reinterpret_cast<const uint4*>(ops);
  const float* bias = reinterpret_cast<const float*>(ops) + kMlpTileOps * kWave * 4;
  const int g = lane >> 4;
  // Each layer's bias is loaded just before its MFMAs and layer_barrier
  // keeps the weight loads of a K block next to their MFMAs.
  f32x4 h1[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) h1[t][r] = bias[tile_unit(t, g, r)];
  }
#pragma unroll
  for (int b = 0; b < 2; ++b) {
#pragma unroll
    for (int t = 0; t < 4; ++t) h1[t] = mfma(w[(0 + 2 * t + b) * kWave + lane], x[b], h1[t]);
  }
  layer_barrier();
  f32x4 h2[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) h2[t][r] = bias[kMlpHidden + tile_unit(t, g, r)];
  }
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    uint4 hi, mid, lo;
    split3(h1[2 * b], h1[2 * b + 1], hi, mid, lo);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const uint4 a = w[(8 + 2 * t + b) * kWave + lane];
      h2[t] = mfma(a, lo, h2[t]);
      h2[t] = mfma(a, mid, h2[t]);
      h2[t] = mfma(a, hi, h2[t]);
    }
    layer_barrier();
  }
  f32x4 z[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) z[t][r] = bias[2 * kMlpHidden + tile_unit(t, g, r)];
  }
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    uint4 hi, mid, lo;
    split3(h2[2 * b], h2[2 * b + 1], hi, mid, lo);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const uint4 a = w[(16 + 2 * t + b) * kWave + lane];
      z[t] = mfma(a, lo, z[t]);
      z[t] = mfma(a, mid, z[t]);
      z[t] = mfma(a, hi, z[t]);
    }
    layer_barrier();
  }
  // Masked softmax of row l % 16 over the 4 lanes l % 16 + 16g.
  float m = -INFINITY;
#pragma unroll
  for (int e = 0; e < 8; ++e) m = (valid >> e & 1u) ? fmaxf(m, z[e >> 2][e & 3]) : m;
  m = fmaxf(m, __shfl_xor(m, 16));
  m = fmaxf(m, __shfl_xor(m, 32));
  float sum = 0.0f;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    prob[e] = (valid >> e & 1u) ? expf(z[e >> 2][e & 3] - m) : 0.0f;
    sum += prob[e];
  }
  sum += __shfl_xor(sum, 16);
  sum += __shfl_xor(sum, 32);
#pragma unroll
  for (int e = 0; e < 8; ++e) prob[e] = prob[e] / sum;
}

// Score and target of the lane's row: score = 1 - p(keep); target rank = the
// most probable valid relabel rank (smallest on ties), -1 (score 0) when no
// rank is valid. All lanes call; the 4 lanes of a row agree.
__device__ __forceinline__ void tile_pick(const float (&prob)[8], uint32_t valid, int lane, float* score,
                                          int* rank) {
  const int g = lane >> 4;
  uint64_t key = 0ull;
  float keep = 0.0f;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int j = tile_unit(e >> 2, g, e & 3);
    if (j == kMlpKeep) keep = prob[e];
    if (j < kMlpRanks && (valid >> e & 1u)) {
      const uint64_t kj = (static_cast<uint64_t>(__float_as_uint(prob[e])) << 32) | static_cast<uint32_t>(kMlpRanks - j);
      key = kj > key ? kj : key;
    }
  }
#pragma unroll
  for (int off = 16; off <= 32; off <<= 1) {
    const uint64_t other = __shfl_xor(key, off);
    key = other > key ? other : key;
    keep += __shfl_xor(keep, off);
  }
  *score = key != 0ull ? 1.0f - keep : 0.0f;
  *rank = key != 0ull ? kMlpRanks - static_cast<int>(key & 0xFFFFFFFFull) : -1;
}

}  // namespace mlp
}  // namespace kodama_hip
