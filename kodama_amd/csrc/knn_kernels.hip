// Exact kNN and kNN-vote kernels for gfx950 (CDNA4, 64-wide waves).
//
// knn_topk_kernel: one thread owns one query row and keeps its k best candidates
// in registers; candidate rows stream through LDS in tiles shared by the 256
// queries of the block. Distances are summed in the order of the reference CPU
// kernel (native_knn.cpp:41-95): 32 fused-multiply-add lane partials per block of
// 32 dimensions, the lane tree (s0+s1)+(s2+s3) -> 8 -> 4 -> (l0+l1)+(l2+l3), then
// the remaining dimensions one fused multiply-add at a time. The result is the
// exact float32 value the reference computes, so neighbour sets and scores are
// bit-identical to its exhaustive search. Candidates are visited in increasing
// index order and only a strictly smaller distance displaces a kept one, which
// realises the (distance, index) order of native_knn.cpp:121-123.
//
// vote_kernel: one thread per validation row of every run of a batch. The k
// neighbour labels are tallied in registers; the winner has the most votes, then
// the largest double sum of scores (added in neighbour-rank order), then the
// smallest label code (core.cpp:871-911). The block also reduces, per run, the
// number of rows whose prediction equals the label, the class counts of the
// labels and the first row of each class (the objective's inputs).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cstdint>

#include "launch.h"

namespace kodama_hip {

namespace {

constexpr int kBlock = 256;

// Candidate rows per LDS tile: 16 KiB of candidates for up to 64 dimensions,
// 32 KiB beyond.
template <int PMAX>
constexpr int tile_rows() { return PMAX <= 64 ? 64 : 8192 / PMAX; }

template <int PMAX, bool EUCLID>
__device__ __forceinline__ float ref_order_distance(const float (&q)[PMAX], const float* __restrict__ c,
                                                    int p) {
  // Lane partials of the full 32-dimension blocks.
  const int nb = p >> 5;
  float acc[32];
#pragma unroll
  for (int s = 0; s < 32; ++s) acc[s] = 0.0f;
#pragma unroll
  for (int b = 0; b < PMAX / 32; ++b) {
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
  for (int d = 0; d < PMAX; ++d) {
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

// Insert (d, id) into an ascending register list when it beats the last entry
// or the list still has an empty slot; equal distances keep the earlier index.
template <int KMAX>
__device__ __forceinline__ void topk_insert(float (&kd)[KMAX], int (&ki)[KMAX], int k, float d, int id) {
  const float worst_d = kd[KMAX - 1];
  const int worst_i = ki[KMAX - 1];
  if (!(d < worst_d || worst_i < 0)) return;
  kd[KMAX - 1] = d;
  ki[KMAX - 1] = id;
#pragma unroll
  for (int j = KMAX - 1; j > 0; --j) {
    // Move the new entry up while it is strictly closer or the slot above is empty.
    const bool up = (ki[j - 1] < 0) || (kd[j] < kd[j - 1]);
    if (up) {
      const float td = kd[j - 1];
      const int ti = ki[j - 1];
      kd[j - 1] = kd[j];
      ki[j - 1] = ki[j];
      kd[j] = td;
      ki[j] = ti;
    }
  }
  (void)k;
}

}  // namespace

// MODE 0: queries vs a separate train set, optional per-query excluded train
// index. MODE 1: every run's rows against the rows of the same run whose fold
// differs (the per-fold tables of core.cpp:463-584, all folds at once).
template <int PMAX, int KMAX, int MODE, bool EUCLID>
__global__ void __launch_bounds__(kBlock) knn_topk_kernel(KnnTopkParams prm) {
  constexpr int kTile = tile_rows<PMAX>();
  __shared__ float tile[kTile * PMAX];
  __shared__ int tile_fold[kTile];

  const int run = blockIdx.y;
  int q_begin, q_count, t_begin, t_count;
  if (MODE == 0) {
    q_begin = 0;
    q_count = prm.query_rows;
    t_begin = 0;
    t_count = prm.train_rows;
  } else {
    q_begin = prm.run_offsets[run];
    q_count = prm.run_offsets[run + 1] - q_begin;
    t_begin = q_begin;
    t_count = q_count;
  }
  const int qi = blockIdx.x * kBlock + threadIdx.x;
  if (blockIdx.x * kBlock >= q_count) return;  // whole block idle for this run
  const bool active = qi < q_count;
  const int p = prm.dims;
  const int k = prm.k;

  float q[PMAX];
  int my_fold = -1;
  int excluded = -1;
  if (active) {
    const int qrow = MODE == 0 ? qi : prm.rows[q_begin + qi];
    const float* src = (MODE == 0 ? prm.query : prm.train) + static_cast<size_t>(qrow) * p;
#pragma unroll
    for (int d = 0; d < PMAX; ++d) q[d] = d < p ? src[d] : 0.0f;
    if (MODE == 1) my_fold = prm.folds[q_begin + qi];
    if (MODE == 0 && prm.exclude != nullptr) excluded = prm.exclude[qi];
  } else {
#pragma unroll
    for (int d = 0; d < PMAX; ++d) q[d] = 0.0f;
  }

  float kd[KMAX];
  int ki[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    kd[j] = FLT_MAX;
    ki[j] = -1;
  }

  for (int base = 0; base < t_count; base += kTile) {
    const int nt = min(kTile, t_count - base);
    __syncthreads();
    for (int e = threadIdx.x; e < kTile * p; e += kBlock) {
      const int r = e / p;
      const int d = e - r * p;
      float v = 0.0f;
      if (r < nt) {
        const int trow = MODE == 0 ? base + r : prm.rows[t_begin + base + r];
        v = prm.train[static_cast<size_t>(trow) * p + d];
      }
      tile[r * PMAX + d] = v;
    }
    if (MODE == 1 && threadIdx.x < kTile) {
      tile_fold[threadIdx.x] = threadIdx.x < nt ? prm.folds[t_begin + base + threadIdx.x] : -1;
    }
    __syncthreads();
    if (active) {
      for (int r = 0; r < nt; ++r) {
        const int cand = base + r;
        if (MODE == 1) {
          if (tile_fold[r] == my_fold) continue;
        } else {
          if (cand == excluded) continue;
        }
        const float d = ref_order_distance<PMAX, EUCLID>(q, &tile[r * PMAX], p);
        topk_insert<KMAX>(kd, ki, k, d, cand);
      }
    }
  }

  if (!active) return;
  if (MODE == 0) {
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      if (j < k) {
        const size_t o = static_cast<size_t>(qi) * k + j;
        prm.out_index[o] = ki[j];
        prm.out_dist[o] = ki[j] >= 0 ? kd[j] : INFINITY;
      }
    }
  } else {
    // Transposed, padded vote layout: column j of run r at padded row.
    const int prow = prm.pad_offsets[run] + qi;
    const int pad_base = prm.pad_offsets[run];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      if (j < k) {
        const size_t o = static_cast<size_t>(j) * prm.pad_total + prow;
        prm.out_index[o] = ki[j] >= 0 ? pad_base + ki[j] : -1;
        prm.out_dist[o] = ki[j] >= 0 ? -kd[j] : -INFINITY;
      }
    }
  }
}

template <int KMAX>
__global__ void __launch_bounds__(kBlock) vote_kernel(VoteParams prm) {
  extern __shared__ int smem[];
  int* hist = smem;                    // [classes]
  int* first = smem + prm.classes;     // [classes]
  __shared__ int block_matches;

  const int run = prm.block_run[blockIdx.x];
  const int row = blockIdx.x * kBlock + threadIdx.x;
  const int local = row - prm.pad_offsets[run];
  const int L = prm.run_offsets[run + 1] - prm.run_offsets[run];
  const bool active = local < L;
  const bool reduce = prm.counts != nullptr;

  if (reduce) {
    for (int c = threadIdx.x; c < prm.classes; c += kBlock) {
      hist[c] = 0;
      first[c] = INT_MAX;
    }
    if (threadIdx.x == 0) block_matches = 0;
    __syncthreads();
  }

  int best_code = -1;
  int label = -1;
  if (active) {
    label = prm.labels[row];
    int code[KMAX];
    float sc[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      code[j] = -1;
      sc[j] = 0.0f;
      if (j < prm.k) {
        const size_t o = static_cast<size_t>(j) * prm.pad_total + row;
        const int nb = prm.neighbors[o];
        if (nb >= 0) {
          code[j] = prm.labels[nb];
          sc[j] = prm.scores[o];
        }
      }
    }
    int best_count = -1;
    double best_sum = -INFINITY;
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      bool fresh = code[j] >= 0;
#pragma unroll
      for (int i = 0; i < j; ++i) fresh = fresh && (code[i] != code[j]);
      if (fresh) {
        int cnt = 0;
        double sum = 0.0;
#pragma unroll
        for (int i = j; i < KMAX; ++i) {
          if (code[i] == code[j]) {
            ++cnt;
            sum += static_cast<double>(sc[i]);
          }
        }
        if (cnt > best_count || (cnt == best_count && (sum > best_sum || (sum == best_sum && code[j] < best_code)))) {
          best_code = code[j];
          best_count = cnt;
          best_sum = sum;
        }
      }
    }
    if (best_code < 0) best_code = prm.fallback_code != nullptr ? prm.fallback_code[run] : -1;
    prm.predictions[row] = best_code;
  }

  if (!reduce) return;
  if (active) {
    if (best_code == label) atomicAdd(&block_matches, 1);
    atomicAdd(&hist[label], 1);
    atomicMin(&first[label], local);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < prm.classes; c += kBlock) {
    if (hist[c] != 0) {
      atomicAdd(&prm.counts[static_cast<size_t>(run) * prm.classes + c], hist[c]);
      atomicMin(&prm.firsts[static_cast<size_t>(run) * prm.classes + c], first[c]);
    }
  }
  if (threadIdx.x == 0 && block_matches != 0) atomicAdd(&prm.matches[run], block_matches);
}

__global__ void vote_reset_kernel(int* counts, int* firsts, int* matches, int runs, int classes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = runs * classes;
  if (i < total) {
    counts[i] = 0;
    firsts[i] = INT_MAX;
  }
  if (i < runs) matches[i] = 0;
}

// ---- host launchers --------------------------------------------------------

template <int PMAX, int KMAX>
static hipError_t launch_topk_pk(const KnnTopkParams& prm, int mode, bool euclid, int grid_x, int grid_y,
                                 hipStream_t s) {
  const dim3 grid(grid_x, grid_y), block(kBlock);
  if (mode == 0) {
    if (euclid) hipLaunchKernelGGL((knn_topk_kernel<PMAX, KMAX, 0, true>), grid, block, 0, s, prm);
    else hipLaunchKernelGGL((knn_topk_kernel<PMAX, KMAX, 0, false>), grid, block, 0, s, prm);
  } else {
    if (euclid) hipLaunchKernelGGL((knn_topk_kernel<PMAX, KMAX, 1, true>), grid, block, 0, s, prm);
    else hipLaunchKernelGGL((knn_topk_kernel<PMAX, KMAX, 1, false>), grid, block, 0, s, prm);
  }
  return hipGetLastError();
}

template <int PMAX>
static hipError_t launch_topk_p(const KnnTopkParams& prm, int mode, bool euclid, int gx, int gy, hipStream_t s) {
  if (prm.k <= 8) return launch_topk_pk<PMAX, 8>(prm, mode, euclid, gx, gy, s);
  if (prm.k <= 16) return launch_topk_pk<PMAX, 16>(prm, mode, euclid, gx, gy, s);
  if (prm.k <= 32) return launch_topk_pk<PMAX, 32>(prm, mode, euclid, gx, gy, s);
  return launch_topk_pk<PMAX, 64>(prm, mode, euclid, gx, gy, s);
}

hipError_t launch_knn_topk(const KnnTopkParams& prm, int mode, bool euclid, int max_queries, int runs,
                           hipStream_t stream) {
  if (prm.k < 1 || prm.k > kMaxK || prm.dims < 1 || prm.dims > kMaxDims) return hipErrorInvalidValue;
  const int gx = (max_queries + kBlock - 1) / kBlock;
  if (gx == 0 || runs == 0) return hipSuccess;
  if (prm.dims <= 32) return launch_topk_p<32>(prm, mode, euclid, gx, runs, stream);
  if (prm.dims <= 64) return launch_topk_p<64>(prm, mode, euclid, gx, runs, stream);
  if (prm.dims <= 128) return launch_topk_p<128>(prm, mode, euclid, gx, runs, stream);
  return launch_topk_p<256>(prm, mode, euclid, gx, runs, stream);
}

hipError_t launch_vote(const VoteParams& prm, int blocks, hipStream_t stream) {
  if (prm.k < 1 || prm.k > kMaxK) return hipErrorInvalidValue;
  if (blocks == 0) return hipSuccess;
  const size_t shmem = prm.counts != nullptr ? sizeof(int) * 2 * static_cast<size_t>(prm.classes) : 0;
  if (shmem > 64 * 1024) return hipErrorInvalidValue;
  const dim3 grid(blocks), block(kBlock);
  if (prm.k <= 8) hipLaunchKernelGGL((vote_kernel<8>), grid, block, shmem, stream, prm);
  else if (prm.k <= 16) hipLaunchKernelGGL((vote_kernel<16>), grid, block, shmem, stream, prm);
  else if (prm.k <= 32) hipLaunchKernelGGL((vote_kernel<32>), grid, block, shmem, stream, prm);
  else hipLaunchKernelGGL((vote_kernel<64>), grid, block, shmem, stream, prm);
  return hipGetLastError();
}

hipError_t launch_vote_reset(int* counts, int* firsts, int* matches, int runs, int classes, hipStream_t stream) {
  const int total = runs * classes > runs ? runs * classes : runs;
  const int blocks = (total + kBlock - 1) / kBlock;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(vote_reset_kernel, dim3(blocks), dim3(kBlock), 0, stream, counts, firsts, matches, runs,
                     classes);
  return hipGetLastError();
}

}  // namespace kodama_hip
