// Device-resident label evolution of a batch of runs (the KODAMA.matrix kNN
// loop of core.cpp:1349-1544 with the per-run flags of kodama_matrix.cpp:
// 3108-3128: prediction-guided singleton relabelling, many-to-one absorption,
// diversity-guarded objective, evolutionary acceptance).
//
// One cycle is a fixed sequence of kernels over all runs, so it is captured once
// in a hipGraph and replayed; the cycle index lives in device memory. Random
// draws come from a counter-based SplitMix64 hash of (run key, cycle, stream,
// index) instead of the host's mt19937_64: the algorithm is the reference's,
// the draws are not.
//
//   evo_clear     zero the per-run flow table, class sizes and vote reductions
//   evo_propose   choose n ~ U{1..hi(cycle)} rows by a keyed permutation and set
//                 them to their current prediction; count label->prediction flows
//   evo_absorb    one workgroup per run: many-to-one absorption decision
//   evo_remap     apply the absorption remap to the proposal
//   vote          (knn_kernels.hip) predictions + per-run reductions
//   evo_accept    one thread per run: objective, best update, acceptance
//   evo_commit    copy accepted / improving proposals into the run state
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>

#include "evolve_launch.h"

namespace kodama_hip {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t draw(uint64_t key, uint64_t cycle, uint64_t stream, uint64_t index) {
  return mix64(key ^ mix64((cycle << 8) ^ stream ^ mix64(index + 0x632BE59BD9B4E019ull)));
}

__device__ __forceinline__ double unit(uint64_t h) { return static_cast<double>(h >> 11) * 0x1.0p-53; }

// Four-round Feistel network on 2*half bits keyed by `key`, cycle-walked into
// [0, n): a pseudo-random permutation of the rows of one run.
__device__ __forceinline__ uint32_t feistel(uint32_t x, int half, uint64_t key) {
  const uint32_t mask = (1u << half) - 1u;
  uint32_t l = x >> half, r = x & mask;
#pragma unroll
  for (int round = 0; round < 4; ++round) {
    const uint32_t f = static_cast<uint32_t>(mix64(key + 0x9E37ull * round + r)) & mask;
    const uint32_t nl = r;
    r = l ^ f;
    l = nl;
  }
  return (l << half) | r;
}

__device__ __forceinline__ uint32_t permute(uint32_t i, uint32_t n, int half, uint64_t key) {
  uint32_t y = feistel(i, half, key);
  while (y >= n) y = feistel(y, half, key);  // terminates: i's orbit returns below n
  return y;
}

// Upper bound of the proposal size (core.cpp:1237-1253, adaptive form).
__device__ __forceinline__ int proposal_hi(int groups, int cycle, int cycles) {
  if (groups <= 1) return groups;
  if (cycles <= 1) return groups;
  double t = static_cast<double>(cycle + 1) / static_cast<double>(cycles + 1);
  t = fmin(fmax(t, 0.0), 1.0);
  const double smooth = t * t * (3.0 - 2.0 * t);
  const double temp = fmin(fmax(1.0 - smooth, 0.0), 1.0);
  return max(1, 1 + static_cast<int>(floor(static_cast<double>(groups - 1) * temp)));
}

}  // namespace

__global__ void evo_clear_kernel(EvoParams p) {
  const size_t tid = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  const size_t flows = static_cast<size_t>(p.runs) * p.classes * p.classes;
  for (size_t i = tid; i < flows; i += stride) p.flows[i] = 0;
  const size_t rc = static_cast<size_t>(p.runs) * p.classes;
  for (size_t i = tid; i < rc; i += stride) {
    p.sizes[i] = 0;
    p.counts[i] = 0;
    p.firsts[i] = INT_MAX;
    p.remap[i] = -1;
  }
  for (size_t i = tid; i < static_cast<size_t>(p.runs); i += stride) p.matches[i] = 0;
}

__global__ void __launch_bounds__(kBlock) evo_propose_kernel(EvoParams p) {
  const int run = p.block_run[blockIdx.x];
  const int row = blockIdx.x * kBlock + threadIdx.x;
  const int base = p.pad_offsets[run];
  const int local = row - base;
  const int n = p.run_offsets[run + 1] - p.run_offsets[run];
  if (local >= n) return;
  const int cycle = *p.cycle;
  const uint64_t key = p.run_key[run];
  const int hi = proposal_hi(n, cycle, p.cycles);
  const int take = 1 + static_cast<int>(unit(draw(key, cycle, 1, 0)) * hi);
  const int half = p.perm_half[run];
  const uint64_t pkey = draw(key, cycle, 2, 0);
  const bool chosen = static_cast<int>(permute(static_cast<uint32_t>(local), static_cast<uint32_t>(n), half, pkey)) <
                      min(take, hi);
  const int guide = p.curpred[row];
  const int label = chosen ? guide : p.cur[row];
  p.prop[row] = label;
  const size_t cb = static_cast<size_t>(run) * p.classes;
  atomicAdd(&p.sizes[cb + label], 1);
  atomicAdd(&p.flows[(cb + label) * p.classes + guide], 1);
}

// Many-to-one absorption (core.cpp:1110-1226) for one run per workgroup: each
// class points at the class it flows into most beyond chance; one target,
// drawn with probability proportional to its total surplus (multi-source
// targets preferred), absorbs all its sources.
__global__ void __launch_bounds__(kBlock) evo_absorb_kernel(EvoParams p) {
  // All LDS of this kernel lives in one double-typed dynamic block starting at
  // offset 0, so the f64 atomics on weight[] are 8-byte aligned.
  extern __shared__ double absorb_lds[];
  const int C = p.classes;
  double* weight = absorb_lds;                                // [C]
  int* sources = reinterpret_cast<int*>(absorb_lds + C);      // [C] sources per target
  int* point = sources + C;                                   // [C] target of each source or -1
  int& present = point[C];
  int& chosen = point[C + 1];
  int& vanished = point[C + 2];

  const int run = blockIdx.x;
  const int n = p.run_offsets[run + 1] - p.run_offsets[run];
  const size_t cb = static_cast<size_t>(run) * C;
  if (threadIdx.x == 0) {
    present = 0;
    chosen = -1;
  }
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    weight[c] = 0.0;
    sources[c] = 0;
    point[c] = -1;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) if (p.sizes[cb + c] > 0) atomicAdd(&present, 1);
  __syncthreads();
  if (present <= 2) {
    if (threadIdx.x == 0) p.decision[run] &= ~kEvoAbsorbed;
    return;
  }
  const double nn = static_cast<double>(n);
  for (int src = threadIdx.x; src < C; src += blockDim.x) {
    const int src_n = p.sizes[cb + src];
    if (src_n <= 0) continue;  // every row is movable in a batch run
    const int* row = p.flows + (cb + src) * C;
    const int stay = row[src];
    int tgt = src, best = 0, tgt_n = src_n;
    double surplus_best = 0.0;
    for (int d = 0; d < C; ++d) {
      const int cnt = row[d];
      if (d == src || cnt == 0) continue;
      const int dn = p.sizes[cb + d];
      if (dn == 0) continue;
      const double surplus = static_cast<double>(cnt) - static_cast<double>(src_n) * static_cast<double>(dn) / fmax(1.0, nn);
      if (surplus > surplus_best || (surplus == surplus_best && cnt > best) ||
          (surplus == surplus_best && cnt == best && dn > tgt_n) ||
          (surplus == surplus_best && cnt == best && dn == tgt_n && d < tgt)) {
        tgt = d; tgt_n = dn; best = cnt; surplus_best = surplus;
      }
    }
    if (tgt == src || surplus_best <= 0.0 || best <= stay) continue;
    point[src] = tgt;
    atomicAdd(&weight[tgt], surplus_best);
    atomicAdd(&sources[tgt], 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // Targets with several sources win over single-source ones (core.cpp:1188-1202).
    bool multi = false;
    for (int t = 0; t < C; ++t) multi = multi || sources[t] > 1;
    double total = 0.0;
    for (int t = 0; t < C; ++t) {
      if (sources[t] == 0 || (multi && sources[t] < 2)) continue;
      total += fmax(weight[t], DBL_MIN);
    }
    const int cycle = *p.cycle;
    if (total > 0.0) {
      const double u = unit(draw(p.run_key[run], cycle, 3, 0)) * total;
      double acc = 0.0;
      int pick = -1;
      for (int t = 0; t < C; ++t) {
        if (sources[t] == 0 || (multi && sources[t] < 2)) continue;
        pick = t;
        acc += fmax(weight[t], DBL_MIN);
        if (u < acc) break;
      }
      chosen = pick;
    }
  }
  __syncthreads();
  const int tgt = chosen;
  if (tgt < 0) {
    if (threadIdx.x == 0) p.decision[run] &= ~kEvoAbsorbed;
    return;
  }
  // All rows are movable: every source vanishes. The result must keep >= 2
  // classes and lose at least one (core.cpp:1219-1223).
  if (threadIdx.x == 0) vanished = 0;
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) if (point[c] == tgt) atomicAdd(&vanished, 1);
  __syncthreads();
  const int after = present - vanished;
  const bool ok = vanished > 0 && after > 1 && after < present;
  if (ok) {
    for (int c = threadIdx.x; c < C; c += blockDim.x) p.remap[cb + c] = point[c] == tgt ? tgt : -1;
  }
  if (threadIdx.x == 0) {
    if (ok) p.decision[run] |= kEvoAbsorbed;
    else p.decision[run] &= ~kEvoAbsorbed;
  }
}

__global__ void __launch_bounds__(kBlock) evo_remap_kernel(EvoParams p) {
  const int run = p.block_run[blockIdx.x];
  const int row = blockIdx.x * kBlock + threadIdx.x;
  const int local = row - p.pad_offsets[run];
  const int n = p.run_offsets[run + 1] - p.run_offsets[run];
  if (local >= n) return;
  if ((p.decision[run] & kEvoAbsorbed) == 0) return;
  const int t = p.remap[static_cast<size_t>(run) * p.classes + p.prop[row]];
  if (t >= 0) p.prop[row] = t;
}

// Objective (core.cpp:916-972 with the diversity guard) and evolutionary
// acceptance (core.cpp:1471-1508); init = 1 evaluates the starting labels.
__global__ void evo_accept_kernel(EvoParams p, int init) {
  const int run = blockIdx.x * blockDim.x + threadIdx.x;
  if (run >= p.runs) return;
  const int n = p.run_offsets[run + 1] - p.run_offsets[run];
  const int C = p.classes;
  const int* cnt = p.counts + static_cast<size_t>(run) * C;
  int k = 0;
  double same = 0.0;
  const double nn = static_cast<double>(n);
  for (int c = 0; c < C; ++c) {
    if (cnt[c] > 0) {
      ++k;
      const double q = static_cast<double>(cnt[c]) / nn;
      same += q * q;
    }
  }
  const double acc = n ? static_cast<double>(p.matches[run]) / nn : 0.0;
  double score = acc;
  if (p.guarded) {
    if (k == 1) score = -INFINITY;
    else score = acc * sqrt(fmax(0.0, 1.0 - same));
  }
  double* st = p.state + static_cast<size_t>(run) * kEvoStateStride;
  const int cycle = *p.cycle;
  if (init) {
    st[0] = acc; st[1] = score; st[2] = acc; st[3] = score;
    p.decision[run] = kEvoAccepted | kEvoImproved;
    return;
  }
  int dec = p.decision[run] & kEvoAbsorbed;
  if (score > st[3]) {
    st[2] = acc;
    st[3] = score;
    dec |= kEvoImproved;
  }
  const double cooling = 1.0 - static_cast<double>(cycle + 1) / static_cast<double>(max(1, p.cycles));
  const double temperature = fmax(1.0e-9, 0.10 * fmax(0.0, 1.0 - st[0]) * cooling);
  bool take = score >= st[1];
  if (!take && temperature > 1.0e-9) {
    const double u = unit(draw(p.run_key[run], cycle, 4, 0));
    take = u < exp((score - st[1]) / temperature);
  }
  if (take) {
    st[0] = acc;
    st[1] = score;
    dec |= kEvoAccepted;
  }
  p.decision[run] = dec;
  if (p.history) p.history[static_cast<size_t>(run) * p.cycles + cycle] = st[2];
}

__global__ void __launch_bounds__(kBlock) evo_commit_kernel(EvoParams p) {
  const int run = p.block_run[blockIdx.x];
  const int row = blockIdx.x * kBlock + threadIdx.x;
  const int local = row - p.pad_offsets[run];
  const int n = p.run_offsets[run + 1] - p.run_offsets[run];
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(p.cycle, 1);
  if (local >= n) return;
  const int dec = p.decision[run];
  if ((dec & (kEvoAccepted | kEvoImproved)) == 0) return;
  const int lab = p.prop[row];
  const int pred = p.proppred[row];
  if (dec & kEvoAccepted) {
    p.cur[row] = lab;
    p.curpred[row] = pred;
  }
  if (dec & kEvoImproved) {
    p.best[row] = lab;
    p.bestpred[row] = pred;
  }
}

// ---- launchers --------------------------------------------------------------

hipError_t launch_evo_clear(const EvoParams& p, hipStream_t s) {
  const size_t work = static_cast<size_t>(p.runs) * p.classes * p.classes;
  const int blocks = static_cast<int>(std::min<size_t>(4096, (work + kBlock - 1) / kBlock + 1));
  hipLaunchKernelGGL(evo_clear_kernel, dim3(blocks), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_evo_propose(const EvoParams& p, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(evo_propose_kernel, dim3(blocks), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_evo_absorb(const EvoParams& p, hipStream_t s) {
  const size_t shmem = static_cast<size_t>(p.classes) * (sizeof(double) + 2 * sizeof(int)) + 3 * sizeof(int);
  if (shmem > 64 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(evo_absorb_kernel, dim3(p.runs), dim3(kBlock), shmem, s, p);
  return hipGetLastError();
}

hipError_t launch_evo_remap(const EvoParams& p, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(evo_remap_kernel, dim3(blocks), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_evo_accept(const EvoParams& p, int init, hipStream_t s) {
  hipLaunchKernelGGL(evo_accept_kernel, dim3((p.runs + 63) / 64), dim3(64), 0, s, p, init);
  return hipGetLastError();
}

hipError_t launch_evo_commit(const EvoParams& p, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(evo_commit_kernel, dim3(blocks), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

}  // namespace kodama_hip
