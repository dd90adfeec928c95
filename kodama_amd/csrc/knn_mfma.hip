// Exact kNN through split-BF16 MFMA candidates and an exact FP32 rerank.
//
// Every row x is split into bf16 parts hi = bf16(x) and lo = bf16(x - hi)
// (64 dimensions, zero padded). The dot product of a query and a train row is
// approximated by hi.hi + hi.lo + lo.hi, three v_mfma_f32_16x16x32_bf16 chains
// accumulated in FP32 (the bf16 x bf16 products are exact in FP32). What the
// split drops (lo.lo and the residuals below lo) is at most 3.01 * 2^-18
// sum |q_i t_i| and the FP32 accumulation of 192 terms at most 192 * 2^-24 of
// the same sum, so for squared Euclidean distances (and for minus the inner
// product) the candidate value differs from the exact value by at most
// E = 8e-5 (|q|^2 + max |t|^2), about twice the sum of these terms, the FP32
// rounding of the norms and of the final combination, and the reference-order
// rounding of the exact distance.
//
// knn_mfma_candidates_kernel: WAVES (4 or 8) waves of 16 queries per
// workgroup; train rows stream through LDS 64 at a time. MFMA output
// lane l holds queries 4 (l / 16) + r (r = 0..3) against train column l % 16 of
// each 16-row tile, so each lane keeps, per query, a sorted list of its LIST
// (8 or 16) smallest candidate values over the train rows of its column
// class. The schedule (LIST, WAVES) does not change the result: the rerank
// certifies every query against its own lists.
//
// knn_rerank_kernel: one wave per query. A_k = the k-th smallest candidate
// value over the query's 16 lists. Any train row whose exact distance can be
// among the k smallest has a candidate value <= A_k + 2E; if a full list's
// largest value is <= A_k + 2E it may have dropped such a row, and the query
// is flagged for the exact scan. Otherwise every candidate <= A_k + 2E is
// re-scored exactly (the reference's FP32 summation order) and the k smallest
// by (distance, index) are written: the same rows and values as the exact scan.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <atomic>
#include <cstdint>

#include "knn_mfma.h"

namespace kodama_hip {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;
constexpr int kChunk = 64;              // train rows per LDS stage
constexpr int kRowStride = 72;          // bf16 per staged row (64 + 8 pad: spreads LDS banks)

__device__ __forceinline__ unsigned short bf16_bits(float x) {
  // Round to nearest even (finite inputs).
  const uint32_t b = __float_as_uint(x);
  return static_cast<unsigned short>((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float bf16_value(unsigned short h) {
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// Order-preserving map of a float to an unsigned key (finite values).
__device__ __forceinline__ uint32_t order_key(float v) {
  const uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key_value(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Ascending insert into a sorted list of distinct packed keys, branch-free:
// slot s takes the median of (old s - 1, old s, key), which shifts the tail
// up by one from the key's position; a key above the last one (or the empty
// key 0xFFFFFFFF) leaves the list as it is.
// v_med3_u32 (the compiler forms a min and a max from the same expression).
__device__ __forceinline__ uint32_t med3_u32(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

template <int LIST>
__device__ __forceinline__ void key_insert(uint32_t (&kk)[LIST], uint32_t key) {
#pragma unroll
  for (int s = LIST - 1; s > 0; --s) kk[s] = med3_u32(kk[s - 1], kk[s], key);
  kk[0] = min(kk[0], key);
}

// Merge of a segment's packed keys (ascending) into the ascending (value key,
// row) pairs: the key of column row m stands for row `row0` + 16 m. The
// half-cleaner min(pairs[s], keys[LIST - 1 - s]) keeps the LIST smallest as a
// bitonic sequence, which the clean-up stages sort.
template <int LIST>
__device__ __forceinline__ void merge_segment(uint64_t (&k2)[LIST], const uint32_t (&kk)[LIST], uint32_t imask,
                                              uint32_t row0) {
  uint64_t c[LIST];
#pragma unroll
  for (int s = 0; s < LIST; ++s) {
    const uint32_t key = kk[LIST - 1 - s];
    const uint64_t pair = key == 0xFFFFFFFFu ? ~0ull
                                             : static_cast<uint64_t>(key & ~imask) << 32 | (row0 + 16u * (key & imask));
    c[s] = min(k2[s], pair);
  }
#pragma unroll
  for (int stride = LIST / 2; stride > 0; stride >>= 1) {
#pragma unroll
    for (int s = 0; s < LIST; ++s) {
      if ((s & stride) == 0) {
        const uint64_t lo = min(c[s], c[s + stride]), hi = max(c[s], c[s + stride]);
        c[s] = lo;
        c[s + stride] = hi;
      }
    }
  }
#pragma unroll
  for (int s = 0; s < LIST; ++s) k2[s] = c[s];
}

// Ascending list insert; a value equal to the largest kept one is dropped.
template <int LIST>
__device__ __forceinline__ void list_insert(float (&d)[LIST], int (&id)[LIST], float v, int j) {
  if (!(v < d[LIST - 1])) return;
  d[LIST - 1] = v;
  id[LIST - 1] = j;
#pragma unroll
  for (int s = LIST - 1; s > 0; --s) {
    if (d[s] < d[s - 1]) {
      const float td = d[s - 1];
      const int ti = id[s - 1];
      d[s - 1] = d[s];
      id[s - 1] = id[s];
      d[s] = td;
      id[s] = ti;
    }
  }
}

// Reference-order FP32 distance (native_knn.cpp:41-95) for p <= 64, one thread.
template <bool EUCLID>
__device__ __forceinline__ float exact_distance(const float (&q)[64], const float* __restrict__ c, int p) {
  const int nb = p >> 5;
  float acc[32];
#pragma unroll
  for (int s = 0; s < 32; ++s) acc[s] = 0.0f;
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    if (b < nb) {
#pragma unroll
      for (int s = 0; s < 32; ++s) {
        const float cv = c[b * 32 + s];
        if (EUCLID) {
          const float t = __fsub_rn(q[b * 32 + s], cv);
          acc[s] = __fmaf_rn(t, t, acc[s]);
        } else {
          acc[s] = __fmaf_rn(q[b * 32 + s], cv, acc[s]);
        }
      }
    }
  }
  float v4[4];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const float lo = __fadd_rn(__fadd_rn(acc[l], acc[8 + l]), __fadd_rn(acc[16 + l], acc[24 + l]));
    const float hi = __fadd_rn(__fadd_rn(acc[4 + l], acc[12 + l]), __fadd_rn(acc[20 + l], acc[28 + l]));
    v4[l] = __fadd_rn(lo, hi);
  }
  float r = __fadd_rn(__fadd_rn(v4[0], v4[1]), __fadd_rn(v4[2], v4[3]));
  const int tail = nb * 32;
#pragma unroll
  for (int d = 0; d < 64; ++d) {
    if (d >= tail && d < p) {
      const float cv = c[d];
      if (EUCLID) {
        const float t = __fsub_rn(q[d], cv);
        r = __fmaf_rn(t, t, r);
      } else {
        r = __fmaf_rn(q[d], cv, r);
      }
    }
  }
  return EUCLID ? r : -r;
}

}  // namespace

// Rows -> bf16 hi / lo parts [rows x 64] and FP32 squared norms. Eight
// threads per row (a row's eight are in one wave), eight dimensions each: one
// 16-byte store per part, so a wave writes eight whole rows. The norm is the
// sum of the eight partial sums over an xor tree (the same value in all eight
// lanes); its rounding is within the candidate error bound like any order's.
__global__ void split_rows_kernel(const float* __restrict__ x, int rows, int p, unsigned short* __restrict__ hi,
                                  unsigned short* __restrict__ lo, float* __restrict__ norm2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = e >> 3, part = e & 7;
  const bool live = r < rows;
  float s = 0.0f;
  if (live) {
    const float* src = x + static_cast<size_t>(r) * p;
    uint32_t h[4], l[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int d = part * 8 + 2 * t;
      const float v0 = d < p ? src[d] : 0.0f;
      const float v1 = d + 1 < p ? src[d + 1] : 0.0f;
      const unsigned short h0 = bf16_bits(v0), h1 = bf16_bits(v1);
      h[t] = h0 | static_cast<uint32_t>(h1) << 16;
      l[t] = bf16_bits(v0 - bf16_value(h0)) | static_cast<uint32_t>(bf16_bits(v1 - bf16_value(h1))) << 16;
      s = fmaf(v0, v0, s);
      s = fmaf(v1, v1, s);
    }
    const size_t o = static_cast<size_t>(r) * 64 + part * 8;
    *reinterpret_cast<uint4*>(hi + o) = make_uint4(h[0], h[1], h[2], h[3]);
    *reinterpret_cast<uint4*>(lo + o) = make_uint4(l[0], l[1], l[2], l[3]);
  }
  s += __shfl_xor(s, 1);
  s += __shfl_xor(s, 2);
  s += __shfl_xor(s, 4);
  if (live && part == 0) norm2[r] = s;
}

// PACK 0: lists of values and indices. 1: lists of packed keys (the train
// rows fit the index bits). 2: packed keys per segment of 16 << index_bits
// train rows, merged at each segment's end into a sorted list of (value key,
// row) pairs, the value key being the packed key with its index bits cleared:
// within a segment the packed order is this pair order, so the merged list
// keeps the LIST smallest pairs of all rows.
template <bool EUCLID, bool TABLES, int LIST, int WAVES, int PACK>
__global__ void __launch_bounds__(WAVES * 64) knn_mfma_candidates_kernel(KnnMfmaParams prm) {
  constexpr int kCandThreads = WAVES * 64;
  constexpr int kQueriesPerBlock = WAVES * 16;
  __shared__ unsigned short s_hi[kChunk * kRowStride];
  __shared__ unsigned short s_lo[kChunk * kRowStride];
  __shared__ float s_tn[kChunk];
  __shared__ int s_fold[kChunk];
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int grp = lane >> 4;       // k group of the operands, query group of the output
  const int col = lane & 15;       // A row / B column / output column
  // Query and train rows: the two sets, or (TABLES) the rows of run
  // blockIdx.y on both sides, pairs of one fold skipped.
  int q_base = 0, q_rows = prm.query_rows, t_base = 0, t_rows = prm.train_rows;
  if (TABLES) {
    q_base = t_base = prm.run_offsets[blockIdx.y];
    q_rows = t_rows = prm.run_offsets[blockIdx.y + 1] - q_base;
    if (static_cast<int>(blockIdx.x) * kQueriesPerBlock >= q_rows) return;  // whole block idle for this run
  }
  const int q_tile = blockIdx.x * kQueriesPerBlock + wave * 16;

  // A operands: query q_tile + col, dimensions kb * 32 + 8 grp .. + 7.
  uint4 a_hi[2], a_lo[2];
  {
    const int q = q_tile + col;
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      if (q < q_rows) {
        const size_t o = static_cast<size_t>(q_base + q) * 64 + kb * 32 + 8 * grp;
        a_hi[kb] = *reinterpret_cast<const uint4*>(prm.query_hi + o);
        a_lo[kb] = *reinterpret_cast<const uint4*>(prm.query_lo + o);
      } else {
        a_hi[kb] = make_uint4(0, 0, 0, 0);
        a_lo[kb] = make_uint4(0, 0, 0, 0);
      }
    }
  }
  // This lane's output queries: norms, and the excluded train index (or, in
  // TABLES, the query's fold).
  float qn[4];
  int excl[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int q = q_tile + 4 * grp + r;
    qn[r] = q < q_rows ? prm.query_norm2[q_base + q] : 0.0f;
    if (TABLES) {
      excl[r] = q < q_rows ? prm.folds[q_base + q] : -1;
    } else {
      excl[r] = (q < q_rows && prm.exclude != nullptr) ? prm.exclude[q] : -1;
    }
  }
  // Unpacked lists: values and indices; packed: one key per entry (the
  // index arrays are not used); segments: the merged pairs beside them.
  constexpr bool kPacked = PACK != 0;
  float kd[4][kPacked ? 1 : LIST];
  int ki[4][kPacked ? 1 : LIST];
  uint32_t kk[4][kPacked ? LIST : 1];
  uint64_t k2[4][PACK == 2 ? LIST : 1];
  const uint32_t imask = kPacked ? (1u << prm.index_bits) - 1u : 0u;
  const int seg_mask = PACK == 2 ? (16 << prm.index_bits) - 1 : 0;  // segment rows - 1 (a power of two)
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if constexpr (kPacked) {
#pragma unroll
      for (int s = 0; s < LIST; ++s) kk[r][s] = 0xFFFFFFFFu;
      if constexpr (PACK == 2) {
#pragma unroll
        for (int s = 0; s < LIST; ++s) k2[r][s] = ~0ull;
      }
    } else {
#pragma unroll
      for (int s = 0; s < LIST; ++s) {
        kd[r][s] = INFINITY;
        ki[r][s] = -1;
      }
    }
  }

  // Staging of 64 train rows x (8 hi + 8 lo 16-byte parts): the next chunk
  // is loaded into registers while this one is scored, so the loads' latency
  // (L2 for a batch's runs, beyond it for a large train set) is not paid at
  // the barrier.
  constexpr int kStage = kChunk * 16 / kCandThreads;  // parts per thread and chunk
  uint4 pre[kStage];
  float pre_tn = 0.0f;
  int pre_fold = -2;
  auto fetch = [&](int c0) {
#pragma unroll
    for (int i = 0; i < kStage; ++i) {
      const int e = static_cast<int>(threadIdx.x) + i * kCandThreads;
      const int row = e >> 4, part = e & 15;
      const int tr = c0 + row;
      pre[i] = make_uint4(0, 0, 0, 0);
      if (tr < t_rows) {
        const unsigned short* src =
            (part >= 8 ? prm.train_lo : prm.train_hi) + static_cast<size_t>(t_base + tr) * 64 + (part & 7) * 8;
        pre[i] = *reinterpret_cast<const uint4*>(src);
      }
    }
    if (threadIdx.x < kChunk) {
      const int tr = c0 + static_cast<int>(threadIdx.x);
      pre_tn = tr < t_rows ? prm.train_norm2[t_base + tr] : 0.0f;
      if (TABLES) pre_fold = tr < t_rows ? prm.folds[t_base + tr] : -2;
    }
  };
  fetch(0);
  for (int c0 = 0; c0 < t_rows; c0 += kChunk) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kStage; ++i) {
      const int e = static_cast<int>(threadIdx.x) + i * kCandThreads;
      const int row = e >> 4, part = e & 15;
      *reinterpret_cast<uint4*>((part >= 8 ? s_lo : s_hi) + row * kRowStride + (part & 7) * 8) = pre[i];
    }
    if (threadIdx.x < kChunk) {
      s_tn[threadIdx.x] = pre_tn;
      if (TABLES) s_fold[threadIdx.x] = pre_fold;
    }
    __syncthreads();
    if (c0 + kChunk < t_rows) fetch(c0 + kChunk);  // in flight during the scoring below
    f32x4 acc[4];
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) acc[tt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      uint4 b_hi[4], b_lo[4];
#pragma unroll
      for (int tt = 0; tt < 4; ++tt) {
        const int off = (tt * 16 + col) * kRowStride + kb * 32 + 8 * grp;
        b_hi[tt] = *reinterpret_cast<const uint4*>(s_hi + off);
        b_lo[tt] = *reinterpret_cast<const uint4*>(s_lo + off);
      }
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
        acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a_hi[kb]), as_bf16x8(b_hi[tt]), acc[tt], 0, 0, 0);
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
        acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a_hi[kb]), as_bf16x8(b_lo[tt]), acc[tt], 0, 0, 0);
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
        acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a_lo[kb]), as_bf16x8(b_hi[tt]), acc[tt], 0, 0, 0);
    }
    if constexpr (kPacked) {
      // Row j = c0 + 16 tt + col is row (j - seg0) / 16 = (c0 - seg0) / 16 + tt
      // of its lane column within the segment starting at row seg0, the same
      // for every lane. A candidate that is excluded, past the rows or not
      // below +inf (the unpacked lists drop those too) gets the empty key.
      const int seg0 = PACK == 2 ? c0 & ~seg_mask : 0;
#pragma unroll
      for (int tt = 0; tt < 4; ++tt) {
        const int j = c0 + tt * 16 + col;
        const uint32_t m = static_cast<uint32_t>(((c0 - seg0) >> 4) + tt);
        const float tn = s_tn[tt * 16 + col];
        const int skip = TABLES ? s_fold[tt * 16 + col] : j;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float v = EUCLID ? (qn[r] + tn) - 2.0f * acc[tt][r] : -acc[tt][r];
          const bool ok = j < t_rows && skip != excl[r] && v < INFINITY;
          key_insert<LIST>(kk[r], ok ? ((order_key(v) & ~imask) | m) : 0xFFFFFFFFu);
        }
      }
      if constexpr (PACK == 2) {
        // Segment end (a uniform condition): fold the keys into the pairs.
        if (((c0 + kChunk) & seg_mask) == 0 || c0 + kChunk >= t_rows) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            merge_segment<LIST>(k2[r], kk[r], imask, static_cast<uint32_t>(t_base + seg0 + col));
#pragma unroll
            for (int s = 0; s < LIST; ++s) kk[r][s] = 0xFFFFFFFFu;
          }
        }
      }
    } else {
#pragma unroll
      for (int tt = 0; tt < 4; ++tt) {
        const int j = c0 + tt * 16 + col;
        if (j < t_rows) {
          const float tn = s_tn[tt * 16 + col];
          const int skip = TABLES ? s_fold[tt * 16 + col] : j;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            if (skip != excl[r]) {
              const float v = EUCLID ? (qn[r] + tn) - 2.0f * acc[tt][r] : -acc[tt][r];
              list_insert<LIST>(kd[r], ki[r], v, t_base + j);
            }
          }
        }
      }
    }
  }

  // Lists: query q owns [16 columns x LIST] entries, column-major by lane column.
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int q = q_tile + 4 * grp + r;
    if (q < q_rows) {
      const size_t base = (static_cast<size_t>(q_base + q) * 16 + col) * LIST;
      if constexpr (PACK == 1) {
        uint32_t* keys = reinterpret_cast<uint32_t*>(prm.cand_index);
#pragma unroll
        for (int s = 0; s < LIST; ++s) keys[base + s] = kk[r][s];
      } else if constexpr (PACK == 2) {
        // The value's lower limit (the rerank knows the index bits) and the row.
#pragma unroll
        for (int s = 0; s < LIST; ++s) {
          const bool full = k2[r][s] != ~0ull;
          prm.cand_value[base + s] = full ? key_value(static_cast<uint32_t>(k2[r][s] >> 32)) : INFINITY;
          prm.cand_index[base + s] = full ? static_cast<int>(static_cast<uint32_t>(k2[r][s])) : -1;
        }
      } else {
#pragma unroll
        for (int s = 0; s < LIST; ++s) {
          prm.cand_value[base + s] = kd[r][s];
          prm.cand_index[base + s] = ki[r][s];
        }
      }
    }
  }
}

// Ascending bitonic sort of 64 R keys over a wave: element e is register
// e / 64 of lane e % 64.
template <int R>
__device__ __forceinline__ void wave_sort(uint32_t (&v)[R], int lane) {
  constexpr int kN = 64 * R;
#pragma unroll
  for (int size = 2; size <= kN; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= 64) {  // between registers of the same lane
        const int rs = stride / 64;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if ((r & rs) == 0) {
            const bool asc = ((r * 64 + lane) & size) == 0;
            const uint32_t lo = min(v[r], v[r + rs]), hi = max(v[r], v[r + rs]);
            v[r] = asc ? lo : hi;
            v[r + rs] = asc ? hi : lo;
          }
        }
        continue;
      }
      const bool lower = (lane & stride) == 0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint32_t other = __shfl_xor(v[r], stride);
        const bool asc = ((r * 64 + lane) & size) == 0;
        v[r] = lower == asc ? min(v[r], other) : max(v[r], other);
      }
    }
  }
}

// One wave per query; 4 queries per 256-thread workgroup. Each lane holds R =
// LIST / 4 of the query's 16 LIST candidates. A_k (the k-th smallest
// candidate value; ties of value do not move it) comes from a bitonic sort of
// the values over the wave; the exact re-scores of the candidates within the
// bound are ranked by (distance, index) among themselves.
//
// Packed lists: a key stands for a value between key_value(key with its index
// bits cleared) and key_value(key with them set), and its row is 16 x its
// index bits + the lane column of its list, from the start of the query's run.
// The k smallest keys' rows then have exact distances <= hi(k-th key) + E, so
// every row of the true k nearest has a value <= A = hi(k-th key) + 2E; the
// candidates re-scored are those with lo(key) <= A, and a full list whose
// last key has lo <= A may have dropped such a row (the keys of a list are
// distinct, its rows being). Lists merged over segments hold (value lower
// limit, row) with the same limits: the same bound and tests apply.
template <bool EUCLID, int LIST, bool PACKED>
__global__ void __launch_bounds__(256) knn_rerank_kernel(KnnMfmaParams prm) {
  constexpr int kCand = 16 * LIST;
  constexpr int R = kCand / kWave;
  static_assert(R * kWave == kCand, "whole registers of candidates per lane");
  __shared__ float s_ex[4][kCand];
  __shared__ int s_id[4][kCand];
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  // Wave-uniform query (readfirstlane): its row, norm and run come through
  // scalar loads.
  const int q = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.x) * 4 + wave);
  const bool live = q < prm.query_rows;  // whole waves; no block barrier below
  if (!live) return;
  const int k = prm.k;
  const int p = prm.dims;
  const size_t base = static_cast<size_t>(q) * kCand;
  const uint32_t* __restrict__ ckey = reinterpret_cast<const uint32_t*>(prm.cand_index);
  const uint32_t imask = prm.index_bits >= 0 ? (1u << prm.index_bits) - 1u : 0u;
  int t_base = 0;  // packed rows count from the start of the query's run
  if (PACKED && prm.runs > 0) {
    int lo = 0, hi = prm.runs - 1;  // run_offsets[lo] <= q < run_offsets[lo + 1]
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prm.run_offsets[mid] <= q) lo = mid; else hi = mid - 1;
    }
    t_base = prm.run_offsets[lo];
  }
  float val[R];  // the candidate value (packed: its lower limit)
  int idx[R];
  uint32_t key[R];
  int valid = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if constexpr (PACKED) {
      const int slot = r * kWave + lane;
      const uint32_t kv = ckey[base + slot];
      idx[r] = kv != 0xFFFFFFFFu ? t_base + 16 * static_cast<int>(kv & imask) + slot / LIST : -1;
      val[r] = key_value(kv & ~imask);
      key[r] = kv;
    } else {
      val[r] = prm.cand_value[base + r * kWave + lane];
      idx[r] = prm.cand_index[base + r * kWave + lane];
      key[r] = idx[r] >= 0 ? order_key(val[r]) : 0xFFFFFFFFu;
    }
    valid += __popcll(__ballot(idx[r] >= 0));
  }
  wave_sort<R>(key, lane);
  // Rank k - 1 sits in register (k - 1) / 64 of lane (k - 1) % 64.
  uint32_t kth = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) kth = r == (k - 1) / kWave ? key[r] : kth;
  const float ak = key_value(__shfl(kth, (k - 1) % kWave) | imask);
  // Fewer valid candidates than k: every one is kept.
  const float bound = valid >= k ? ak + 2.0f * (prm.error_scale * (prm.query_norm2[q] + prm.train_norm2_max)) : INFINITY;
  // A full list whose largest value is within the bound may have dropped a row.
  bool overflow = false;
  if (lane < 16) {
    const size_t last = base + lane * LIST + LIST - 1;
    if constexpr (PACKED) {
      const uint32_t kv = ckey[last];
      overflow = kv != 0xFFFFFFFFu && key_value(kv & ~imask) <= bound;
    } else {
      overflow = prm.cand_index[last] >= 0 && prm.cand_value[last] <= bound;
    }
  }
  if (__ballot(overflow) != 0ull) {
    if (lane == 0) prm.flags[q] = 1;
    return;
  }
  if (lane == 0) prm.flags[q] = 0;
  // Exact re-score of the candidates within the bound, compacted into the
  // wave's LDS slice.
  float qv[64];
  const float* qrow = prm.query + static_cast<size_t>(q) * p;
#pragma unroll
  for (int d = 0; d < 64; ++d) qv[d] = d < p ? qrow[d] : 0.0f;
  const uint64_t below = (1ull << lane) - 1ull;
  float* ex = s_ex[wave];
  int* id = s_id[wave];
  float e[R];
  bool use[R];
  int nsel = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    use[r] = idx[r] >= 0 && val[r] <= bound;
    e[r] = use[r] ? exact_distance<EUCLID>(qv, prm.train + static_cast<size_t>(idx[r]) * p, p) : INFINITY;
    const uint64_t m = __ballot(use[r]);
    if (use[r]) {
      ex[nsel + __popcll(m & below)] = e[r];
      id[nsel + __popcll(m & below)] = idx[r];
    }
    nsel += __popcll(m);
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  int rank[R];
#pragma unroll
  for (int r = 0; r < R; ++r) rank[r] = 0;
  for (int t = 0; t < nsel; ++t) {
    const float x = ex[t];
    const int j = id[t];
#pragma unroll
    for (int r = 0; r < R; ++r) rank[r] += (x < e[r] || (x == e[r] && j < idx[r])) ? 1 : 0;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (use[r] && rank[r] < k) {
      prm.out_index[static_cast<size_t>(q) * k + rank[r]] = idx[r];
      prm.out_dist[static_cast<size_t>(q) * k + rank[r]] = e[r];
    }
  }
  for (int t = nsel + lane; t < k; t += kWave) {
    prm.out_index[static_cast<size_t>(q) * k + t] = -1;
    prm.out_dist[static_cast<size_t>(q) * k + t] = INFINITY;
  }
}

hipError_t launch_split_rows(const float* x, int rows, int p, unsigned short* hi, unsigned short* lo, float* norm2,
                             hipStream_t s) {
  if (rows == 0) return hipSuccess;
  if (p < 1 || p > 64) return hipErrorInvalidValue;
  const size_t threads = static_cast<size_t>(rows) * 8;
  hipLaunchKernelGGL(split_rows_kernel, dim3(static_cast<unsigned>((threads + 255) / 256)), dim3(256), 0, s, x, rows,
                     p, hi, lo, norm2);
  return hipGetLastError();
}

template <int LIST, int WAVES, int PACK>
static void launch_mfma_schedule(const KnnMfmaParams& prm, bool euclid, bool tables, dim3 grid, hipStream_t s) {
  const dim3 block(WAVES * 64);
  if (tables) {
    if (euclid) hipLaunchKernelGGL((knn_mfma_candidates_kernel<true, true, LIST, WAVES, PACK>), grid, block, 0, s, prm);
    else hipLaunchKernelGGL((knn_mfma_candidates_kernel<false, true, LIST, WAVES, PACK>), grid, block, 0, s, prm);
  } else {
    if (euclid) hipLaunchKernelGGL((knn_mfma_candidates_kernel<true, false, LIST, WAVES, PACK>), grid, block, 0, s, prm);
    else hipLaunchKernelGGL((knn_mfma_candidates_kernel<false, false, LIST, WAVES, PACK>), grid, block, 0, s, prm);
  }
}

template <int LIST, int PACK>
static void launch_mfma_packing(const KnnMfmaParams& prm, bool euclid, bool tables, dim3 grid, hipStream_t s) {
  if (prm.waves == 4) launch_mfma_schedule<LIST, 4, PACK>(prm, euclid, tables, grid, s);
  else launch_mfma_schedule<LIST, 8, PACK>(prm, euclid, tables, grid, s);
  const unsigned rblocks = static_cast<unsigned>((prm.query_rows + 3) / 4);
  if (euclid) hipLaunchKernelGGL((knn_rerank_kernel<true, LIST, PACK == 1>), dim3(rblocks), dim3(256), 0, s, prm);
  else hipLaunchKernelGGL((knn_rerank_kernel<false, LIST, PACK == 1>), dim3(rblocks), dim3(256), 0, s, prm);
}

static std::atomic<int> g_packed_lists{1};

void set_mfma_packed_lists(int on) { g_packed_lists.store(on != 0 ? 1 : 0); }
int mfma_packed_lists() { return g_packed_lists.load(); }

// `train_rows` rows of one search (the largest run for a batch of tables):
// packed lists when their column indices (row / 16) fit kMfmaMaxIndexBits;
// past that, 8-entry lists are packed per segment of kMfmaSegmentBits index
// bits (16 << kMfmaSegmentBits rows) and merged into (value, row) pairs.
template <int LIST>
static void launch_mfma_list(KnnMfmaParams prm, bool euclid, bool tables, int rows, int train_rows, hipStream_t s,
                             unsigned runs) {
  const int qpb = prm.waves * 16;
  const dim3 grid(static_cast<unsigned>((rows + qpb - 1) / qpb), runs);
  const unsigned top = static_cast<unsigned>((train_rows + 15) / 16 - 1);  // largest column index
  int bits = 0;
  while (bits < 32 && (top >> bits) != 0u) ++bits;
  prm.runs = tables ? static_cast<int>(runs) : 0;
  const bool packed = g_packed_lists.load() != 0;
  if (packed && bits <= kMfmaMaxIndexBits) {
    prm.index_bits = bits;
    launch_mfma_packing<LIST, 1>(prm, euclid, tables, grid, s);
    return;
  }
  if constexpr (LIST == 8) {
    if (packed) {
      prm.index_bits = kMfmaSegmentBits;
      launch_mfma_packing<LIST, 2>(prm, euclid, tables, grid, s);
      return;
    }
  }
  prm.index_bits = -1;
  launch_mfma_packing<LIST, 0>(prm, euclid, tables, grid, s);
}

static bool schedule_ok(const KnnMfmaParams& prm) {
  const int list = mfma_list(prm);
  return (prm.waves == 4 || prm.waves == 8) && (list == 16 || (list == 8 && prm.k <= 32));
}

hipError_t launch_knn_mfma(const KnnMfmaParams& prm, bool euclid, hipStream_t s) {
  if (prm.query_rows == 0) return hipSuccess;
  if (prm.dims < 1 || prm.dims > 64 || prm.k < 1 || prm.k > kMfmaMaxK || prm.train_rows < 1 || !schedule_ok(prm)) {
    return hipErrorInvalidValue;
  }
  if (mfma_list(prm) == 8) launch_mfma_list<8>(prm, euclid, false, prm.query_rows, prm.train_rows, s, 1);
  else launch_mfma_list<16>(prm, euclid, false, prm.query_rows, prm.train_rows, s, 1);
  return hipGetLastError();
}

hipError_t launch_knn_mfma_tables(const KnnMfmaParams& prm, bool euclid, int runs, int max_rows, hipStream_t s) {
  if (prm.query_rows == 0 || runs == 0) return hipSuccess;
  if (prm.dims < 1 || prm.dims > 64 || prm.k < 1 || prm.k > kMfmaMaxK || prm.run_offsets == nullptr ||
      prm.folds == nullptr || prm.query_rows != prm.train_rows || max_rows < 1 || !schedule_ok(prm)) {
    return hipErrorInvalidValue;
  }
  if (mfma_list(prm) == 8) launch_mfma_list<8>(prm, euclid, true, max_rows, max_rows, s, static_cast<unsigned>(runs));
  else launch_mfma_list<16>(prm, euclid, true, max_rows, max_rows, s, static_cast<unsigned>(runs));
  return hipGetLastError();
}

}  // namespace kodama_hip

namespace kodama_hip {

__global__ void gather_rows_kernel(const float* __restrict__ src, const int* __restrict__ rows, int count, int p,
                                   float* __restrict__ dst) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(count) * p) return;
  const int r = static_cast<int>(i / p), d = static_cast<int>(i % p);
  dst[i] = src[static_cast<size_t>(rows[r]) * p + d];
}

__global__ void scatter_results_kernel(const int* __restrict__ src_idx, const float* __restrict__ src_dist,
                                       const int* __restrict__ rows, int count, int k, int* __restrict__ dst_idx,
                                       float* __restrict__ dst_dist) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(count) * k) return;
  const int r = static_cast<int>(i / k), j = static_cast<int>(i % k);
  dst_idx[static_cast<size_t>(rows[r]) * k + j] = src_idx[i];
  dst_dist[static_cast<size_t>(rows[r]) * k + j] = src_dist[i];
}

// Reranked per-fold results [rows x k] (concatenated indices, internal
// distances) -> the transposed padded vote layout [k x pad_total]: padded row
// ids and scores (minus the distance); rows the rerank flagged are left to
// the exact scan.
__global__ void publish_tables_kernel(const int* __restrict__ idx, const float* __restrict__ dist,
                                      const int* __restrict__ flags, const int* __restrict__ run_offsets, int runs,
                                      const int* __restrict__ pad_offsets, int pad_total, int rows, int k,
                                      int* __restrict__ out_index, float* __restrict__ out_score) {
  const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<size_t>(rows) * k) return;
  const int i = static_cast<int>(e / k), j = static_cast<int>(e % k);
  if (flags[i]) return;
  int lo = 0, hi = runs - 1;  // run of row i: run_offsets[run] <= i < run_offsets[run + 1]
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (run_offsets[mid] <= i) lo = mid; else hi = mid - 1;
  }
  const int id = idx[e];
  const size_t o = static_cast<size_t>(j) * pad_total + pad_offsets[lo] + (i - run_offsets[lo]);
  out_index[o] = id >= 0 ? pad_offsets[lo] + (id - run_offsets[lo]) : -1;
  out_score[o] = id >= 0 ? -dist[e] : -INFINITY;
}

hipError_t launch_publish_tables(const int* idx, const float* dist, const int* flags, const int* run_offsets,
                                 int runs, const int* pad_offsets, int pad_total, int rows, int k, int* out_index,
                                 float* out_score, hipStream_t s) {
  const size_t total = static_cast<size_t>(rows) * k;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(publish_tables_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s, idx,
                     dist, flags, run_offsets, runs, pad_offsets, pad_total, rows, k, out_index, out_score);
  return hipGetLastError();
}

// Merged lists of listed rows [count x k] (run-local ids, distances) -> the
// padded vote layout.
__global__ void publish_listed_kernel(const int* __restrict__ idx, const float* __restrict__ dist,
                                      const int* __restrict__ qlist, const int* __restrict__ qrun,
                                      const int* __restrict__ pad_offsets, int pad_total, int count, int k,
                                      int* __restrict__ out_index, float* __restrict__ out_score) {
  const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<size_t>(count) * k) return;
  const int f = static_cast<int>(e / k), j = static_cast<int>(e % k);
  const int run = qrun[f];
  const int id = idx[e];
  const size_t o = static_cast<size_t>(j) * pad_total + pad_offsets[run] + qlist[f];
  out_index[o] = id >= 0 ? pad_offsets[run] + id : -1;
  out_score[o] = id >= 0 ? -dist[e] : -INFINITY;
}

hipError_t launch_publish_listed(const int* idx, const float* dist, const int* qlist, const int* qrun,
                                 const int* pad_offsets, int pad_total, int count, int k, int* out_index,
                                 float* out_score, hipStream_t s) {
  const size_t total = static_cast<size_t>(count) * k;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(publish_listed_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s, idx,
                     dist, qlist, qrun, pad_offsets, pad_total, count, k, out_index, out_score);
  return hipGetLastError();
}

hipError_t launch_gather_rows(const float* src, const int* rows, int count, int p, float* dst, hipStream_t s) {
  const size_t total = static_cast<size_t>(count) * p;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_rows_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s, src,
                     rows, count, p, dst);
  return hipGetLastError();
}

hipError_t launch_scatter_results(const int* src_idx, const float* src_dist, const int* rows, int count, int k,
                                  int* dst_idx, float* dst_dist, hipStream_t s) {
  const size_t total = static_cast<size_t>(count) * k;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(scatter_results_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s,
                     src_idx, src_dist, rows, count, k, dst_idx, dst_dist);
  return hipGetLastError();
}

}  // namespace kodama_hip
