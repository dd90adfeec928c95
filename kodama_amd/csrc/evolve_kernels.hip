// Device-resident label evolution of a batch of runs: the KODAMA.matrix kNN loop
// (core.cpp:1349-1544 with the per-run flags of kodama_matrix.cpp:3108-3128:
// prediction-guided singleton relabelling, many-to-one absorption,
// diversity-guarded objective, evolutionary acceptance).
//
// One workgroup (1024 threads, 512 for k > 16) owns one run for a whole launch: the run's
// current labels and predictions, the proposal, the class tables and the
// absorption scratch live in LDS (16-bit label codes), so a launch runs any
// number of cycles with only block barriers between phases. The neighbour table
// (16-bit local ids, [k x rows]) streams from HBM; flow counts use a per-run HBM
// table of which only the touched entries are read and reset. Random draws come
// from a counter-based SplitMix64 hash of (run key, cycle, stream) instead of the
// host's mt19937_64: the algorithm is the reference's, the draws are not.
//
// Cycle phases (all within the workgroup):
//   propose  choose n ~ U{1..hi(cycle)} rows by a keyed permutation and set them
//            to their current prediction; class sizes, label->prediction flows,
//            the distinct flow pairs
//   absorb   many-to-one absorption over the distinct pairs (core.cpp:1110-1226)
//   vote     kNN vote of the proposal (core.cpp:784-914), class counts, matches
//   accept   objective (core.cpp:916-972) and acceptance (core.cpp:1471-1508),
//            then commit by swapping the proposal and current buffers
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>

#include "evolve_launch.h"

namespace kodama_hip {

namespace {

constexpr int kThreads = kEvoThreads;
constexpr int kWave = 64;
constexpr int kWaves = kThreads / kWave;
constexpr unsigned short kNoNeighbour = 0xFFFFu;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t draw(uint64_t key, uint64_t cycle, uint64_t stream) {
  return mix64(key ^ mix64((cycle << 8) ^ stream));
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
  if (groups <= 1 || cycles <= 1) return groups;
  double t = static_cast<double>(cycle + 1) / static_cast<double>(cycles + 1);
  t = fmin(fmax(t, 0.0), 1.0);
  const double smooth = t * t * (3.0 - 2.0 * t);
  const double temp = fmin(fmax(1.0 - smooth, 0.0), 1.0);
  return max(1, 1 + static_cast<int>(floor(static_cast<double>(groups - 1) * temp)));
}

// LDS layout of one run-workgroup (8-byte entries first).
struct Lds {
  unsigned long long* weight;   // [C] absorption weight per target, times n (exact)
  double* wl;                   // [C] eligible weight per present-list entry
  double* wave_total;           // [kWaves]
  double* state;                // [4] cur_acc, cur_score, best_acc, best_score
  unsigned long long* best_key; // [C] packed best target per source
  int* sizes;                   // [C] class sizes of the proposal
  int* stay;                    // [C] rows whose prediction equals their label, per label
  int* counts;                  // [C] class counts after absorption (objective)
  int* nsrc;                    // [C] absorbed sources per target
  int* point;                   // [C] target of each source or -1
  int* list;                    // [C] present classes, ascending
  int* remap;                   // [C] absorption remap or -1
  int* pairs;                   // [Lmax] distinct flow entries (label * C + prediction)
  int* queue;                   // [Lmax] rows left to the full tally
  int* scal;                    // [16] scalars
  unsigned short* labels;       // [4 x Lmax] label buffers 0, 1 then prediction buffers 0, 1
  int lmax;
  // Buffer b of the labels / predictions (the current state and the proposal
  // swap on acceptance); plain arithmetic keeps the struct in registers.
  __device__ unsigned short* lab(int b) const { return labels + b * lmax; }
  __device__ unsigned short* pred(int b) const { return labels + (2 + b) * lmax; }
};

enum Scalar { kPresent, kPairs, kMatches, kChosen, kVanished, kPick, kDecision, kCur, kQueue };

__device__ __forceinline__ Lds carve(unsigned char* base, int C, int Lmax) {
  Lds s;
  double* d = reinterpret_cast<double*>(base);
  s.weight = reinterpret_cast<unsigned long long*>(d);
  s.wl = d + C;
  s.wave_total = d + 2 * C;
  s.state = d + 2 * C + kWaves;
  s.best_key = reinterpret_cast<unsigned long long*>(d + 2 * C + kWaves + 4);
  int* i = reinterpret_cast<int*>(s.best_key + C);
  s.sizes = i;
  s.stay = i + C;
  s.counts = i + 2 * C;
  s.nsrc = i + 3 * C;
  s.point = i + 4 * C;
  s.list = i + 5 * C;
  s.remap = i + 6 * C;
  s.pairs = i + 7 * C;
  s.queue = s.pairs + Lmax;
  s.scal = s.queue + Lmax;
  unsigned short* h = reinterpret_cast<unsigned short*>(s.scal + 16);
  s.labels = h;
  s.lmax = Lmax;
  return s;
}

// Codes of row i's neighbours in label buffer `lab` (-1 for empty slots): all
// id loads issued first, then the LDS lookups. Returns the neighbour count.
template <int KMAX>
__device__ __forceinline__ int neighbour_codes(int k, int pad_total, const unsigned short* __restrict__ nbr,
                                               const unsigned short* lab, int i, int (&code)[KMAX]) {
  unsigned short id[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) id[j] = j < k ? nbr[static_cast<size_t>(j) * pad_total + i] : kNoNeighbour;
  int valid = 0;
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    const bool ok = id[j] != kNoNeighbour;
    code[j] = ok ? static_cast<int>(lab[ok ? id[j] : 0]) : -1;
    valid += ok ? 1 : 0;
  }
  return valid;
}

// Full vote of one row: most votes, then the larger double sum of scores in
// neighbour-rank order, then the smaller label (core.cpp:871-911).
template <int KMAX>
__device__ __forceinline__ int full_tally(int pad_total, const float* __restrict__ scores, int i,
                                          const int (&code)[KMAX],
                          int valid) {
  int best_code = -1, best_count = 0, seen = 0;
  bool tie = false;
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    if (code[j] >= 0 && best_count <= valid - seen) {
      bool fresh = true;
#pragma unroll
      for (int t = 0; t < j; ++t) fresh = fresh && (code[t] != code[j]);
      if (fresh) {
        int cnt = 0;
#pragma unroll
        for (int t = j; t < KMAX; ++t) cnt += code[t] == code[j] ? 1 : 0;
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
  // Scores are read on demand: ties are rare and the reads hit the cache, and
  // keeping them out of registers leaves room for the codes.
  double best_sum = -INFINITY;
  best_code = -1;
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    if (code[j] >= 0) {
      bool fresh = true;
#pragma unroll
      for (int t = 0; t < j; ++t) fresh = fresh && (code[t] != code[j]);
      if (fresh) {
        int cnt = 0;
#pragma unroll
        for (int t = j; t < KMAX; ++t) cnt += code[t] == code[j] ? 1 : 0;
        if (cnt == best_count) {
          double sum = 0.0;
#pragma unroll
          for (int t = j; t < KMAX; ++t) {
            if (code[t] == code[j]) sum += static_cast<double>(scores[static_cast<size_t>(t) * pad_total + i]);
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

// Vote of the labels in `lab` into `pred`, with class counts of `lab` and the
// number of rows whose prediction equals their label. Rows without a strict
// majority are queued and tallied densely.
template <int KMAX, int THREADS>
__device__ __forceinline__ void vote_phase(int k, int pad_total, const Lds& s, const unsigned short* __restrict__ nbr,
                           const float* __restrict__ scores, const unsigned short* lab, unsigned short* pred, int n) {
  int matches = 0;
  for (int i = threadIdx.x; i < n; i += THREADS) {
    int code[KMAX];
    const int valid = neighbour_codes<KMAX>(k, pad_total, nbr, lab, i, code);
    const int own = lab[i];
    atomicAdd(&s.counts[own], 1);
    int c0 = 0;
#pragma unroll
    for (int j = 0; j < KMAX; ++j) c0 += code[j] == code[0] ? 1 : 0;
    if (valid > 0 && 2 * c0 > valid) {
      pred[i] = static_cast<unsigned short>(code[0]);
      matches += code[0] == own ? 1 : 0;
    } else {
      s.queue[atomicAdd(&s.scal[kQueue], 1)] = i;
    }
  }
  __syncthreads();
  const int qn = s.scal[kQueue];
  for (int q = threadIdx.x; q < qn; q += THREADS) {
    const int i = s.queue[q];
    int code[KMAX];
    const int valid = neighbour_codes<KMAX>(k, pad_total, nbr, lab, i, code);
    // Every row of a batch run has a neighbour (each run has two folds or more).
    const int best = valid > 0 ? full_tally<KMAX>(pad_total, scores, i, code, valid) : static_cast<int>(lab[i]);
    pred[i] = static_cast<unsigned short>(best);
    matches += best == static_cast<int>(lab[i]) ? 1 : 0;
  }
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1) matches += __shfl_xor(matches, m);
  if ((threadIdx.x & (kWave - 1)) == 0 && matches != 0) atomicAdd(&s.scal[kMatches], matches);
  __syncthreads();
}

// Objective with the diversity guard from the class counts (wave 0).
__device__ __forceinline__ double objective(const Lds& s, bool guarded, int C, int n, double* acc_out) {
  const int lane = threadIdx.x & (kWave - 1);
  int k = 0;
  double same = 0.0;
  const double nn = static_cast<double>(n);
  for (int c = lane; c < C; c += kWave) {
    const int v = s.counts[c];
    if (v > 0) {
      ++k;
      const double q = static_cast<double>(v) / nn;
      same += q * q;
    }
  }
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1) {
    k += __shfl_xor(k, m);
    same += __shfl_xor(same, m);
  }
  const double acc = n ? static_cast<double>(s.scal[kMatches]) / nn : 0.0;
  *acc_out = acc;
  if (!guarded) return acc;
  return k == 1 ? -INFINITY : acc * sqrt(fmax(0.0, 1.0 - same));
}

}  // namespace

template <int KMAX, int THREADS>
__global__ void __launch_bounds__(THREADS) evo_persistent_kernel(EvoParams p) {
  extern __shared__ double lds_raw[];
  const int C = p.classes;
  const Lds s = carve(reinterpret_cast<unsigned char*>(lds_raw), C, p.max_rows);
  const int run = blockIdx.x;
  const int n = p.run_offsets[run + 1] - p.run_offsets[run];
  const int base = p.pad_offsets[run];
  const unsigned short* __restrict__ nbr = p.neighbors + base;
  const float* __restrict__ scores = p.scores + base;
  const uint64_t key = p.run_key[run];
  const int half = p.perm_half[run];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  int* flows = p.flows + static_cast<size_t>(run) * C * C;
  const int k = p.k;
  const int pad_total = p.pad_total;
  const bool guarded = p.guarded != 0;

  // Per-launch set-up: clear the class tables, load the run state.
  for (int c = threadIdx.x; c < C; c += THREADS) {
    s.weight[c] = 0ull;
    s.best_key[c] = 0ull;
    s.sizes[c] = 0;
    s.stay[c] = 0;
    s.counts[c] = 0;
    s.nsrc[c] = 0;
    s.point[c] = -1;
    s.remap[c] = -1;
  }
  if (threadIdx.x < 16) s.scal[threadIdx.x] = 0;
  for (int i = threadIdx.x; i < n; i += THREADS) {
    s.lab(0)[i] = p.cur[base + i];
    s.pred(0)[i] = p.curpred[base + i];
  }
  if (threadIdx.x < 4) s.state[threadIdx.x] = p.state[static_cast<size_t>(run) * kEvoStateStride + threadIdx.x];
  __syncthreads();

  if (p.init) {
    // Evaluate the starting labels: predictions and scores become current and best.
    vote_phase<KMAX, THREADS>(k, pad_total, s, nbr, scores, s.lab(0), s.pred(0), n);
    if (wave == 0) {
      double acc;
      const double score = objective(s, guarded, C, n, &acc);
      if (lane == 0) {
        s.state[0] = acc;
        s.state[1] = score;
        s.state[2] = acc;
        s.state[3] = score;
      }
    }
    for (int i = threadIdx.x; i < n; i += THREADS) {
      p.best[base + i] = s.lab(0)[i];
      p.bestpred[base + i] = s.pred(0)[i];
    }
    __syncthreads();
    for (int c = threadIdx.x; c < C; c += THREADS) s.counts[c] = 0;
    if (threadIdx.x == 0) {
      s.scal[kMatches] = 0;
      s.scal[kQueue] = 0;
    }
    __syncthreads();
  }

  for (int step = 0; step < p.cycle_count; ++step) {
    const int cycle = p.cycle_begin + step;
    const int cur = s.scal[kCur];
    const unsigned short* cur_lab = s.lab(cur);
    const unsigned short* cur_pred = s.pred(cur);
    unsigned short* prop = s.lab(cur ^ 1);
    unsigned short* prop_pred = s.pred(cur ^ 1);

    // -- propose ----------------------------------------------------------------
    {
      const int hi = proposal_hi(n, cycle, p.cycles);
      const int take = min(hi, 1 + static_cast<int>(unit(draw(key, cycle, 1)) * hi));
      const uint64_t pkey = draw(key, cycle, 2);
      for (int i = threadIdx.x; i < n; i += THREADS) {
        const bool chosen = static_cast<int>(permute(static_cast<uint32_t>(i), static_cast<uint32_t>(n), half, pkey)) < take;
        const int guide = cur_pred[i];
        const int label = chosen ? guide : static_cast<int>(cur_lab[i]);
        prop[i] = static_cast<unsigned short>(label);
        atomicAdd(&s.sizes[label], 1);
        if (label == guide) {
          atomicAdd(&s.stay[label], 1);
        } else if (atomicAdd(&flows[label * C + guide], 1) == 0) {
          s.pairs[atomicAdd(&s.scal[kPairs], 1)] = label * C + guide;
        }
      }
    }
    __syncthreads();

    // -- absorb -----------------------------------------------------------------
    if (wave == 0) {
      int running = 0;
      for (int b = 0; b < C; b += kWave) {
        const int c = b + lane;
        const bool here = c < C && s.sizes[c] > 0;
        const uint64_t mask = __ballot(here);
        if (here) s.list[running + __popcll(mask & ((1ull << lane) - 1ull))] = c;
        running += __popcll(mask);
      }
      if (lane == 0) {
        s.scal[kPresent] = running;
        s.scal[kChosen] = -1;
        s.scal[kVanished] = 0;
        s.scal[kPick] = INT_MAX;
        s.scal[kDecision] = 0;
      }
    }
    __syncthreads();
    const int present = s.scal[kPresent];
    const bool eligible = present > 2;
    {
      // Best target per source: atomicMax of a key ordered as the reference's
      // candidate comparison -- surplus (exactly, as cnt*n - size_src*size_dst),
      // then flow, then target size, then the smaller label.
      const int np = s.scal[kPairs];
      for (int q = threadIdx.x; q < np; q += THREADS) {
        const int idx = s.pairs[q];
        // Read and clear in one L2 atomic: the propose-phase atomics of this
        // launch bypass the vector L1, which may hold an older copy.
        const int cnt = atomicExch(&flows[idx], 0);
        if (!eligible) continue;
        const int src = idx / C;
        const int d = idx - src * C;
        const int dn = s.sizes[d];
        if (dn == 0) continue;
        const long long surplus_n = static_cast<long long>(cnt) * n - static_cast<long long>(s.sizes[src]) * dn;
        const unsigned long long k64 = (static_cast<unsigned long long>(surplus_n + (1ll << 27)) << 36) |
                                       (static_cast<unsigned long long>(cnt) << 23) |
                                       (static_cast<unsigned long long>(dn) << 10) |
                                       static_cast<unsigned long long>(1023 - d);
        atomicMax(&s.best_key[src], k64);
      }
    }
    __syncthreads();
    if (eligible) {
      const double nn = fmax(1.0, static_cast<double>(n));
      for (int e = threadIdx.x; e < present; e += THREADS) {
        const int src = s.list[e];
        const unsigned long long k64 = s.best_key[src];
        if (k64 == 0ull) continue;
        const long long surplus_n = static_cast<long long>(k64 >> 36) - (1ll << 27);
        const int cnt = static_cast<int>((k64 >> 23) & 0x1FFFull);
        const int d = 1023 - static_cast<int>(k64 & 0x3FFull);
        if (surplus_n > 0 && cnt > s.stay[src]) {
          s.point[src] = d;
          atomicAdd(&s.weight[d], static_cast<unsigned long long>(surplus_n));
          atomicAdd(&s.nsrc[d], 1);
        }
      }
      __syncthreads();
      // Targets with several sources win over single-source ones
      // (core.cpp:1188-1202); one is drawn in ascending label order with weights
      // max(surplus, DBL_MIN), through a block-wide prefix sum.
      int local_multi = 0;
      for (int e = threadIdx.x; e < present; e += THREADS) local_multi |= s.nsrc[s.list[e]] > 1 ? 1 : 0;
      const bool multi = __syncthreads_or(local_multi) != 0;
      const int per = (present + THREADS - 1) / THREADS;
      const int b0 = min(present, static_cast<int>(threadIdx.x) * per), b1 = min(present, b0 + per);
      double part = 0.0;
      for (int e = b0; e < b1; ++e) {
        const int t = s.list[e];
        const double w =
          s.nsrc[t] > 0 && (!multi || s.nsrc[t] > 1) ? fmax(static_cast<double>(s.weight[t]) / nn, DBL_MIN) : 0.0;
        s.wl[e] = w;
        part += w;
      }
      double incl = part;
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        const double y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
      }
      if (lane == kWave - 1) s.wave_total[wave] = incl;
      __syncthreads();
      double before = 0.0, total = 0.0;
      for (int w = 0; w < THREADS / kWave; ++w) {
        if (w < wave) before += s.wave_total[w];
        total += s.wave_total[w];
      }
      if (total > 0.0) {
        const double u = unit(draw(key, cycle, 3)) * total;
        double acc = before + incl - part;
        for (int e = b0; e < b1; ++e) {
          const double w = s.wl[e];
          // Neighbouring threads' boundaries can differ by a rounding step; the
          // smallest matching entry wins, so the pick is deterministic.
          if (w > 0.0 && u >= acc && u < acc + w) atomicMin(&s.scal[kPick], e);
          acc += w;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
          int e = s.scal[kPick];
          if (e == INT_MAX) {
            // u at or past the last boundary: the last eligible target, as the
            // sequential scan would take.
            for (e = present - 1; e > 0 && s.wl[e] <= 0.0; --e) {
            }
          }
          s.scal[kChosen] = s.list[e];
        }
      }
      __syncthreads();
      const int tgt = s.scal[kChosen];
      if (tgt >= 0) {
        int v = 0;
        for (int e = threadIdx.x; e < present; e += THREADS) v += s.point[s.list[e]] == tgt ? 1 : 0;
        if (v) atomicAdd(&s.scal[kVanished], v);
      }
      __syncthreads();
      // Every row of a batch run is movable, so each source vanishes; the result
      // must keep two classes or more and lose at least one (core.cpp:1219-1223).
      const int vanished = s.scal[kVanished];
      if (tgt >= 0 && vanished > 0 && present - vanished > 1) {
        for (int e = threadIdx.x; e < present; e += THREADS) {
          const int c = s.list[e];
          if (s.point[c] == tgt) s.remap[c] = tgt;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += THREADS) {
          const int r = s.remap[prop[i]];
          if (r >= 0) prop[i] = static_cast<unsigned short>(r);
        }
      }
    }
    // Reset the class tables of this cycle (present classes only).
    for (int e = threadIdx.x; e < present; e += THREADS) {
      const int c = s.list[e];
      s.weight[c] = 0ull;
      s.best_key[c] = 0ull;
      s.sizes[c] = 0;
      s.stay[c] = 0;
      s.nsrc[c] = 0;
      s.point[c] = -1;
      s.remap[c] = -1;
    }
    if (threadIdx.x == 0) s.scal[kPairs] = 0;
    __syncthreads();

    // -- vote -------------------------------------------------------------------
    vote_phase<KMAX, THREADS>(k, pad_total, s, nbr, scores, prop, prop_pred, n);

    // -- accept -----------------------------------------------------------------
    if (wave == 0) {
      double acc;
      const double score = objective(s, guarded, C, n, &acc);
      if (lane == 0) {
        double cur_acc = s.state[0], cur_score = s.state[1];
        int dec = 0;
        if (score > s.state[3]) {
          s.state[2] = acc;
          s.state[3] = score;
          dec |= kEvoImproved;
        }
        const double cooling = 1.0 - static_cast<double>(cycle + 1) / static_cast<double>(max(1, p.cycles));
        const double temperature = fmax(1.0e-9, 0.10 * fmax(0.0, 1.0 - cur_acc) * cooling);
        bool take = score >= cur_score;
        if (!take && temperature > 1.0e-9) take = unit(draw(key, cycle, 4)) < exp((score - cur_score) / temperature);
        if (take) {
          s.state[0] = acc;
          s.state[1] = score;
          dec |= kEvoAccepted;
        }
        s.scal[kDecision] = dec;
        if (p.history) p.history[static_cast<size_t>(run) * p.cycles + cycle] = s.state[2];
      }
    }
    __syncthreads();
    const int dec = s.scal[kDecision];
    if (dec & kEvoImproved) {
      for (int i = threadIdx.x; i < n; i += THREADS) {
        p.best[base + i] = prop[i];
        p.bestpred[base + i] = prop_pred[i];
      }
    }
    // Clear the class counts of this cycle (at most n distinct labels were hit).
    for (int c = threadIdx.x; c < C; c += THREADS) s.counts[c] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
      if (dec & kEvoAccepted) s.scal[kCur] = cur ^ 1;
      s.scal[kMatches] = 0;
      s.scal[kQueue] = 0;
    }
    __syncthreads();
  }

  // Write the run state back for the next launch.
  const int cur = s.scal[kCur];
  for (int i = threadIdx.x; i < n; i += THREADS) {
    p.cur[base + i] = s.lab(cur)[i];
    p.curpred[base + i] = s.pred(cur)[i];
  }
  if (threadIdx.x < 4) p.state[static_cast<size_t>(run) * kEvoStateStride + threadIdx.x] = s.state[threadIdx.x];
}

// Padded rows come in blocks of 256 that each belong to one run (block_run).
__global__ void evo_narrow_kernel(const int* __restrict__ nbr32, const int* __restrict__ pad_offsets,
                                  const int* __restrict__ block_run, unsigned short* __restrict__ nbr16, size_t total,
                                  int pad_total) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int v = nbr32[i];
  const int row = static_cast<int>(i % static_cast<size_t>(pad_total));
  nbr16[i] = v < 0 ? kNoNeighbour : static_cast<unsigned short>(v - pad_offsets[block_run[row >> 8]]);
}

__global__ void evo_zero_kernel(int* a, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// ---- launchers --------------------------------------------------------------

size_t evo_lds_bytes(int classes, int max_rows) {
  return sizeof(double) * (2 * static_cast<size_t>(classes) + kWaves + 4) +
         sizeof(unsigned long long) * static_cast<size_t>(classes) +
         sizeof(int) * (7 * static_cast<size_t>(classes) + 2 * static_cast<size_t>(max_rows) + 16) +
         sizeof(unsigned short) * 4 * static_cast<size_t>(max_rows);
}

// Wider neighbour lists hold more values in registers per thread: their
// kernels run 512 threads per workgroup so that nothing spills.
template <int KMAX>
static hipError_t launch_persistent_k(const EvoParams& p, size_t shmem, hipStream_t s) {
  constexpr int threads = KMAX <= 16 ? kThreads : kThreads / 2;
  const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&evo_persistent_kernel<KMAX, threads>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((evo_persistent_kernel<KMAX, threads>), dim3(p.runs), dim3(threads), shmem, s, p);
  return hipGetLastError();
}

hipError_t launch_evo_persistent(const EvoParams& p, hipStream_t s) {
  const size_t shmem = evo_lds_bytes(p.classes, p.max_rows);
  if (shmem > kEvoMaxLds) return hipErrorInvalidValue;
  if (p.k <= 8) return launch_persistent_k<8>(p, shmem, s);
  if (p.k <= 16) return launch_persistent_k<16>(p, shmem, s);
  if (p.k <= 32) return launch_persistent_k<32>(p, shmem, s);
  return launch_persistent_k<64>(p, shmem, s);
}

hipError_t launch_evo_narrow(const int* nbr32, const int* pad_offsets, const int* row_run, unsigned short* nbr16,
                             size_t total, int pad_total, hipStream_t s) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(evo_narrow_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s, nbr32,
                     pad_offsets, row_run, nbr16, total, pad_total);
  return hipGetLastError();
}

hipError_t launch_evo_zero(int* a, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(evo_zero_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, a, n);
  return hipGetLastError();
}

}  // namespace kodama_hip
