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
// vote_kernel: one thread per validation row of every run of a batch; each block
// belongs to one run and stages that run's labels in LDS. The k neighbour labels
// are tallied in registers; the winner has the most votes, then the largest
// double sum of scores (added in neighbour-rank order), then the smallest label
// code (core.cpp:871-911). The block also reduces, per run, the number of rows
// whose prediction equals the label, the class counts of the labels and the
// first row of each class (the objective's inputs).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cstdint>

#include "launch.h"

namespace kodama_hip {

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;

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
// BLOCK: threads (queries) per block; the listed rescans of flagged table rows
// (a few queries per run) take one wave, so idle waves hold no CU slots.
template <int PMAX, int KMAX, int MODE, bool EUCLID, int BLOCK = kBlock>
__global__ void __launch_bounds__(BLOCK) knn_topk_kernel(KnnTopkParams prm) {
  constexpr int kTile = tile_rows<PMAX>();
  __shared__ float tile[kTile * PMAX];
  __shared__ int tile_fold[kTile];

  const int run = blockIdx.y;
  int q_begin, q_count, t_begin, t_count;
  if (MODE == 0) {
    q_begin = 0;
    q_count = prm.query_rows;
    t_begin = prm.slice_rows > 0 ? static_cast<int>(blockIdx.y) * prm.slice_rows : 0;
    t_count = prm.slice_rows > 0 ? min(prm.slice_rows, prm.train_rows - t_begin) : prm.train_rows;
  } else {
    q_begin = prm.run_offsets[run];
    const int L = prm.run_offsets[run + 1] - q_begin;
    q_count = prm.qlist != nullptr ? prm.qlist_offsets[run + 1] - prm.qlist_offsets[run] : L;
    // Run-local offset of this block's train rows (a slice, or all of them).
    t_begin = prm.slice_rows > 0 ? static_cast<int>(blockIdx.z) * prm.slice_rows : 0;
    t_count = prm.slice_rows > 0 ? min(prm.slice_rows, L - t_begin) : L;
  }
  static_assert(kTile <= BLOCK, "one thread per tile row stages the folds");
  const int qi = blockIdx.x * BLOCK + threadIdx.x;
  if (blockIdx.x * BLOCK >= q_count) return;  // whole block idle for this run
  const bool active = qi < q_count;
  const int p = prm.dims;
  const int k = prm.k;

  // MODE 1: the run-local row of this query (all rows, or the listed ones).
  const int lq = MODE == 1 && active && prm.qlist != nullptr ? prm.qlist[prm.qlist_offsets[run] + qi] : qi;
  float q[PMAX];
  int my_fold = -1;
  int excluded = -1;
  if (active) {
    const int qrow = MODE == 0 ? qi : prm.rows[q_begin + lq];
    const float* src = (MODE == 0 ? prm.query : prm.train) + static_cast<size_t>(qrow) * p;
#pragma unroll
    for (int d = 0; d < PMAX; ++d) q[d] = d < p ? src[d] : 0.0f;
    if (MODE == 1) my_fold = prm.folds[q_begin + lq];
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
    for (int e = threadIdx.x; e < kTile * p; e += BLOCK) {
      const int r = e / p;
      const int d = e - r * p;
      float v = 0.0f;
      if (r < nt) {
        const int trow = MODE == 0 ? t_begin + base + r : prm.rows[q_begin + t_begin + base + r];
        v = prm.train[static_cast<size_t>(trow) * p + d];
      }
      tile[r * PMAX + d] = v;
    }
    if (MODE == 1 && threadIdx.x < kTile) {
      tile_fold[threadIdx.x] = threadIdx.x < nt ? prm.folds[q_begin + t_begin + base + threadIdx.x] : -1;
    }
    __syncthreads();
    if (active) {
      for (int r = 0; r < nt; ++r) {
        // Global train index in MODE 0, run-local row in MODE 1.
        const int cand = t_begin + base + r;
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
        const size_t o = (static_cast<size_t>(blockIdx.y) * prm.query_rows + qi) * k + j;
        prm.out_index[o] = ki[j];
        prm.out_dist[o] = ki[j] >= 0 ? kd[j] : INFINITY;
      }
    }
  } else if (prm.slice_rows > 0) {
    // Slice lists of the listed queries, for the merge.
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      if (j < k) {
        const size_t o = (static_cast<size_t>(blockIdx.z) * prm.list_total + prm.qlist_offsets[run] + qi) * k + j;
        prm.out_index[o] = ki[j];
        prm.out_dist[o] = ki[j] >= 0 ? kd[j] : INFINITY;
      }
    }
  } else {
    // Transposed, padded vote layout: column j of run r at padded row.
    const int prow = prm.pad_offsets[run] + lq;
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

// Label codes of a row's neighbours (-1 for empty slots). All neighbour ids are
// loaded first and all label lookups issued next, with no branch on a loaded
// value in between, so each batch of loads is in flight together. With STAGE the
// run's labels (already remapped) sit in LDS at lab_s[local id]; otherwise they
// are gathered from HBM and remapped here. Returns the number of neighbours.
template <int KMAX, bool STAGE>
__device__ __forceinline__ int load_codes(const VoteParams& prm, int row, int pad_base, const int* lab_s,
                                          const int* remap, bool remapping, int (&code)[KMAX]) {
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    code[j] = j < prm.k ? prm.neighbors[static_cast<size_t>(j) * prm.pad_total + row] : -1;
  }
  int gathered[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    if (STAGE) gathered[j] = lab_s[code[j] >= 0 ? code[j] - pad_base : 0];
    else gathered[j] = prm.labels[code[j] >= 0 ? code[j] : row];
  }
  int valid = 0;
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    if (code[j] >= 0) {
      int c = gathered[j];
      if (!STAGE && remapping && remap[c] >= 0) c = remap[c];
      code[j] = c;
      ++valid;
    }
  }
  return valid;
}

// The full vote of one row: most votes, then the larger double sum of scores
// in neighbour-rank order, then the smaller label (core.cpp:871-911).
template <int KMAX>
__device__ int full_tally(const VoteParams& prm, int row, const int (&code)[KMAX], int valid) {
  // Pass 1, integers only, stopping once no unseen label can reach the best.
  int best_code = -1, best_count = 0, seen = 0;
  bool tie = false;
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    if (code[j] >= 0 && best_count <= valid - seen) {
      bool fresh = true;
#pragma unroll
      for (int i = 0; i < j; ++i) fresh = fresh && (code[i] != code[j]);
      if (fresh) {
        int cnt = 0;
#pragma unroll
        for (int i = j; i < KMAX; ++i) cnt += code[i] == code[j] ? 1 : 0;
        seen += cnt;
        if (cnt > best_count) {
          best_count = cnt;
          best_code = code[j];
          tie = false;
        } else if (cnt == best_count) {
          tie = true;
        }
      }
    }
  }
  if (!tie) return best_code;
  // Pass 2, only when several labels share the best count.
  float sc[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) sc[j] = code[j] >= 0 ? prm.scores[static_cast<size_t>(j) * prm.pad_total + row] : 0.0f;
  double best_sum = -INFINITY;
  best_code = -1;
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    if (code[j] >= 0) {
      bool fresh = true;
#pragma unroll
      for (int i = 0; i < j; ++i) fresh = fresh && (code[i] != code[j]);
      if (fresh) {
        int cnt = 0;
#pragma unroll
        for (int i = j; i < KMAX; ++i) cnt += code[i] == code[j] ? 1 : 0;
        if (cnt == best_count) {
          double sum = 0.0;
#pragma unroll
          for (int i = j; i < KMAX; ++i) {
            if (code[i] == code[j]) sum += static_cast<double>(sc[i]);
          }
          if (sum > best_sum || (sum == best_sum && code[j] < best_code)) {
            best_sum = sum;
            best_code = code[j];
          }
        }
      }
    }
  }
  return best_code;
}

template <int KMAX, bool STAGE>
__global__ void __launch_bounds__(kBlock) vote_kernel(VoteParams prm) {
  extern __shared__ int smem[];
  const int C = prm.classes;
  int* hist = smem;              // [C] label counts of this block
  int* first = smem + C;         // [C] first row of each label in this block
  int* remap = smem + 2 * C;     // [C] absorption remap of this run (device form)
  int* queue = smem + 3 * C;     // [kBlock] rows left to the full tally
  int* pred_s = queue + kBlock;  // [kBlock] their predictions
  int* scal = pred_s + kBlock;   // block matches, queue length
  int* lab_s = scal + 2;         // [rows of the run] staged labels (STAGE)

  const int run = prm.block_run[blockIdx.x];
  const int row = blockIdx.x * kBlock + threadIdx.x;
  const int pad_base = prm.pad_offsets[run];
  const int local = row - pad_base;
  const int L = prm.run_offsets[run + 1] - prm.run_offsets[run];
  const bool active = local < L;
  const bool reduce = prm.counts != nullptr;
  const bool remapping = prm.remap != nullptr && (prm.decision[run] & prm.remap_flag) != 0;
  const size_t cb = static_cast<size_t>(run) * C;
  // Reductions are double-buffered by cycle parity when a cycle counter is given.
  const int parity = prm.cycle != nullptr ? (*prm.cycle & 1) : 0;
  const size_t red = static_cast<size_t>(parity) * prm.runs * C;

  for (int c = threadIdx.x; c < C; c += kBlock) {
    hist[c] = 0;
    first[c] = INT_MAX;
    remap[c] = remapping ? prm.remap[cb + c] : -1;
  }
  if (threadIdx.x == 0) {
    scal[0] = 0;
    scal[1] = 0;
  }
  __syncthreads();
  if (STAGE) {
    // The run's labels, remapped. The remap is idempotent (targets map to
    // themselves), so a label another block already rewrote stages the same.
    for (int i = threadIdx.x; i < L; i += kBlock) {
      const int c = prm.labels[pad_base + i];
      lab_s[i] = remapping && remap[c] >= 0 ? remap[c] : c;
    }
    __syncthreads();
  }

  // Fast path: a label holding a strict majority of the votes is the unique
  // winner. Other rows are queued and tallied densely below, so the quadratic
  // tally only runs on full waves of mixed rows.
  int best_code = -1;
  int label = -1;
  bool queued = false;
  if (active) {
    if (STAGE) {
      label = lab_s[local];
      if (remapping && label != prm.labels[row]) prm.relabel[row] = label;
    } else {
      label = prm.labels[row];
      if (remapping && remap[label] >= 0) {
        label = remap[label];
        prm.relabel[row] = label;
      }
    }
    int code[KMAX];
    const int valid = load_codes<KMAX, STAGE>(prm, row, pad_base, lab_s, remap, remapping, code);
    if (valid > 0) {
      int c0 = 0;
#pragma unroll
      for (int j = 0; j < KMAX; ++j) c0 += code[j] == code[0] ? 1 : 0;
      if (2 * c0 > valid) {
        best_code = code[0];
      } else {
        queue[atomicAdd(&scal[1], 1)] = threadIdx.x;
        queued = true;
      }
    }
  }
  __syncthreads();
  const int qn = scal[1];
  for (int q = threadIdx.x; q < qn; q += kBlock) {
    const int t = queue[q];
    const int qrow = blockIdx.x * kBlock + t;
    int code[KMAX];
    const int valid = load_codes<KMAX, STAGE>(prm, qrow, pad_base, lab_s, remap, remapping, code);
    pred_s[t] = full_tally<KMAX>(prm, qrow, code, valid);
  }
  __syncthreads();
  if (active) {
    if (queued) best_code = pred_s[threadIdx.x];
    if (best_code < 0) best_code = prm.fallback_code != nullptr ? prm.fallback_code[run] : -1;
    prm.predictions[row] = best_code;
  }

  if (!reduce) return;
  if (active) {
    if (best_code == label) atomicAdd(&scal[0], 1);
    atomicAdd(&hist[label], 1);
    if (prm.firsts != nullptr) atomicMin(&first[label], local);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += kBlock) {
    if (hist[c] != 0) {
      atomicAdd(&prm.counts[red + cb + c], hist[c]);
      if (prm.firsts != nullptr) atomicMin(&prm.firsts[red + cb + c], first[c]);
    }
  }
  if (threadIdx.x == 0 && scal[0] != 0) {
    atomicAdd(&prm.matches[static_cast<size_t>(parity) * prm.runs + run], scal[0]);
  }
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

// Merge of per-slice top-k lists, one wave per query: lane s holds the head of
// list s; each round the wave takes the smallest (distance, index) key over the
// heads and the owning lane advances. Keys are unique (every train index lies
// in one slice), so exactly one lane wins each round.
__global__ void __launch_bounds__(kBlock) knn_merge_slices_kernel(const int* __restrict__ idx,
                                                                  const float* __restrict__ dist, int slices,
                                                                  int query_rows, int k, int* __restrict__ out_index,
                                                                  float* __restrict__ out_dist) {
  const int q = blockIdx.x * (kBlock / 64) + static_cast<int>(threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (q >= query_rows) return;  // whole waves; no block barrier below
  const size_t stride = static_cast<size_t>(query_rows) * k;
  const size_t row = static_cast<size_t>(q) * k;
  auto key_at = [&](int h) -> uint64_t {
    if (lane >= slices || h >= k) return ~0ull;
    const size_t o = static_cast<size_t>(lane) * stride + row + h;
    const int i = idx[o];
    if (i < 0) return ~0ull;
    // Order-preserving map of the float onto unsigned bits, index below it.
    const uint32_t b = __float_as_uint(dist[o]);
    const uint32_t ord = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return (static_cast<uint64_t>(ord) << 32) | static_cast<uint32_t>(i);
  };
  int head = 0;
  uint64_t mine = key_at(0);
  for (int j = 0; j < k; ++j) {
    uint64_t best = mine;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      const uint64_t other = __shfl_xor(best, m);
      best = other < best ? other : best;
    }
    if (lane == 0) {
      if (best == ~0ull) {
        out_index[row + j] = -1;
        out_dist[row + j] = INFINITY;
      } else {
        const uint32_t ord = static_cast<uint32_t>(best >> 32);
        const uint32_t b = (ord & 0x80000000u) ? (ord & 0x7FFFFFFFu) : ~ord;
        out_index[row + j] = static_cast<int>(static_cast<uint32_t>(best));
        out_dist[row + j] = __uint_as_float(b);
      }
    }
    if (best != ~0ull && mine == best) mine = key_at(++head);
  }
}

// ---- host launchers --------------------------------------------------------

// MODES = 1: query vs train only; 2: both modes.
template <int PMAX, int KMAX, int MODES = 2>
static hipError_t launch_topk_pk(const KnnTopkParams& prm, int mode, bool euclid, int max_queries, int grid_y,
                                 int grid_z, hipStream_t s) {
  if constexpr (MODES == 2) {
    if (mode == 1 && prm.qlist != nullptr) {
      const dim3 grid((max_queries + kWave - 1) / kWave, grid_y, grid_z), block(kWave);
      if (euclid) hipLaunchKernelGGL((knn_topk_kernel<PMAX, KMAX, 1, true, kWave>), grid, block, 0, s, prm);
      else hipLaunchKernelGGL((knn_topk_kernel<PMAX, KMAX, 1, false, kWave>), grid, block, 0, s, prm);
      return hipGetLastError();
    }
  }
  const dim3 grid((max_queries + kBlock - 1) / kBlock, grid_y, grid_z), block(kBlock);
  if (mode == 0) {
    if (euclid) hipLaunchKernelGGL((knn_topk_kernel<PMAX, KMAX, 0, true>), grid, block, 0, s, prm);
    else hipLaunchKernelGGL((knn_topk_kernel<PMAX, KMAX, 0, false>), grid, block, 0, s, prm);
  } else if constexpr (MODES == 2) {
    if (euclid) hipLaunchKernelGGL((knn_topk_kernel<PMAX, KMAX, 1, true>), grid, block, 0, s, prm);
    else hipLaunchKernelGGL((knn_topk_kernel<PMAX, KMAX, 1, false>), grid, block, 0, s, prm);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int PMAX>
static hipError_t launch_topk_p(const KnnTopkParams& prm, int mode, bool euclid, int gx, int gy, int gz,
                                hipStream_t s) {
  // gx: the most queries of one run (the blocks along x follow the block size).
  if (prm.k <= 8) return launch_topk_pk<PMAX, 8>(prm, mode, euclid, gx, gy, gz, s);
  if (prm.k <= 16) return launch_topk_pk<PMAX, 16>(prm, mode, euclid, gx, gy, gz, s);
  if (prm.k <= 32) return launch_topk_pk<PMAX, 32>(prm, mode, euclid, gx, gy, gz, s);
  if (prm.k <= 64) return launch_topk_pk<PMAX, 64>(prm, mode, euclid, gx, gy, gz, s);
  // 65..128 neighbours: query vs train only, up to 128 dimensions (the lists
  // then sit partly in the accumulation registers; beyond they would spill).
  if constexpr (PMAX <= 128) return launch_topk_pk<PMAX, kMaxScanK, 1>(prm, mode, euclid, gx, gy, gz, s);
  return hipErrorInvalidValue;
}

hipError_t launch_knn_topk(const KnnTopkParams& prm, int mode, bool euclid, int max_queries, int runs,
                           hipStream_t stream, int slices) {
  if (prm.k < 1 || prm.dims < 1 || prm.dims > kMaxDims) return hipErrorInvalidValue;
  if (prm.k > (mode == 0 && prm.dims <= 128 ? kMaxScanK : kMaxK)) return hipErrorInvalidValue;
  // Sliced per-fold scans write per-listed-query lists: they need the list.
  if (slices < 1 || (slices > 1 && (mode != 1 || prm.qlist == nullptr || prm.slice_rows <= 0))) {
    return hipErrorInvalidValue;
  }
  if (mode == 1 && prm.slice_rows > 0 && prm.qlist == nullptr) return hipErrorInvalidValue;
  const int gx = max_queries;
  if (gx == 0 || runs == 0) return hipSuccess;
  if (prm.dims <= 32) return launch_topk_p<32>(prm, mode, euclid, gx, runs, slices, stream);
  if (prm.dims <= 64) return launch_topk_p<64>(prm, mode, euclid, gx, runs, slices, stream);
  if (prm.dims <= 128) return launch_topk_p<128>(prm, mode, euclid, gx, runs, slices, stream);
  return launch_topk_p<256>(prm, mode, euclid, gx, runs, slices, stream);
}

template <bool STAGE>
static void launch_vote_k(const VoteParams& prm, dim3 grid, size_t shmem, hipStream_t stream) {
  const dim3 block(kBlock);
  if (prm.k <= 8) hipLaunchKernelGGL((vote_kernel<8, STAGE>), grid, block, shmem, stream, prm);
  else if (prm.k <= 16) hipLaunchKernelGGL((vote_kernel<16, STAGE>), grid, block, shmem, stream, prm);
  else if (prm.k <= 32) hipLaunchKernelGGL((vote_kernel<32, STAGE>), grid, block, shmem, stream, prm);
  else hipLaunchKernelGGL((vote_kernel<64, STAGE>), grid, block, shmem, stream, prm);
}

hipError_t launch_knn_merge_slices(const int* idx, const float* dist, int slices, int query_rows, int k,
                                   int* out_index, float* out_dist, hipStream_t stream) {
  if (slices < 1 || slices > 64 || k < 1 || k > kMaxScanK) return hipErrorInvalidValue;
  if (query_rows == 0) return hipSuccess;
  const int per_block = kBlock / 64;
  hipLaunchKernelGGL(knn_merge_slices_kernel, dim3((query_rows + per_block - 1) / per_block), dim3(kBlock), 0, stream,
                     idx, dist, slices, query_rows, k, out_index, out_dist);
  return hipGetLastError();
}

hipError_t launch_vote(const VoteParams& prm, int blocks, hipStream_t stream) {
  if (prm.k < 1 || prm.k > kMaxK) return hipErrorInvalidValue;
  if (blocks == 0) return hipSuccess;
  const size_t base = sizeof(int) * (3 * static_cast<size_t>(prm.classes) + 2 * kBlock + 2);
  if (base > 64 * 1024) return hipErrorInvalidValue;
  // Stage the run's labels in LDS when they fit (48 KiB of labels at most).
  const size_t staged = base + sizeof(int) * static_cast<size_t>(prm.max_rows);
  if (prm.max_rows > 0 && staged <= 64 * 1024 && prm.max_rows <= 12288) {
    launch_vote_k<true>(prm, dim3(blocks), staged, stream);
  } else {
    launch_vote_k<false>(prm, dim3(blocks), base, stream);
  }
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
