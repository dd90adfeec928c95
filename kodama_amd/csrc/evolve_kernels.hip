// Device-resident label evolution of a batch of runs: the KODAMA.matrix kNN loop
// (core.cpp:1349-1544 with the per-run flags of kodama_matrix.cpp:3108-3128:
// prediction-guided singleton relabelling, many-to-one absorption,
// diversity-guarded objective, evolutionary acceptance).
//
// A workgroup (1024 threads; 512 for k > 32 and the learned variant) holds one
// run for a whole launch: the run's current labels and predictions, the
// proposal, the class tables and the absorption scratch live in LDS (16-bit
// label codes), so a launch runs any number of cycles with only block
// barriers between phases. The neighbour table (16-bit local ids, [k x rows])
// streams from HBM; flow counts use a per-workgroup HBM table of which only
// the touched entries are read and reset. Random draws come from a
// counter-based SplitMix64 hash of (run key, cycle, stream) instead of the
// host's mt19937_64: the algorithm is the reference's, the draws are not. A
// large run may be split over up to eight workgroups (see EvoParams): each
// replays every phase but votes only its slice of rows, and the slices'
// predictions and match counts are exchanged after the vote.
//
// Cycle phases (all within the workgroup):
//   propose  choose n ~ U{1..hi(cycle)} rows by a keyed permutation and set them
//            to their current prediction; class sizes, label->prediction flows,
//            the distinct flow pairs
//   absorb   many-to-one absorption over the distinct pairs (core.cpp:1110-1226)
//   vote     kNN vote of the proposal (core.cpp:784-914) and the number of rows
//            whose prediction equals their label
//   accept   objective (core.cpp:916-972) and acceptance (core.cpp:1471-1508),
//            then commit by swapping the proposal and current buffers
// A launch that starts an evolution first runs an evaluation-only step that
// votes the starting labels in place; both share one copy of the vote code,
// which keeps the kernel inside the instruction cache.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>

#include "evolve_launch.h"
#include "mlp_device.h"
#include "mlp_launch.h"

namespace kodama_hip {

namespace {

constexpr int kThreads = kEvoThreads;
constexpr int kWave = 64;
constexpr int kWaves = kThreads / kWave;
// Workgroup of the learned variant up to 32 neighbours (the largest learned one).
constexpr int kLearnedThreads = kThreads / 2;
constexpr int kLearnedWaves = kLearnedThreads / kWave;
// Label code of the sentinel slot lab[n] of each label buffer: a missing
// neighbour points there (local id n) and matches no label (codes <= 1023).
constexpr uint32_t kSentinel = 0xFFFFu;
static_assert(kSentinel == mlp::kNone, "a missing neighbour has one code in the vote and the features");
constexpr unsigned short kChosenMark = 0xFFFFu;  // above every label code

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t draw(uint64_t key, uint64_t cycle, uint64_t stream) {
  return mix64(key ^ mix64((cycle << 8) ^ stream));
}

// The thread index, read anew at the top of every cycle. Values derived from
// threadIdx.x (row addresses, bounds masks) are loop-invariant across the cycle
// loop; hoisted out of it they outlive the 128-register budget of a
// 1024-thread workgroup and spill to scratch. From this opaque copy they are
// recomputed each cycle instead, for a few instructions.
__device__ __forceinline__ int opaque_tid() {
  int t;
  asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"(static_cast<int>(threadIdx.x)));
  return t;
}

// exp of the acceptance test, out of line: one lane calls it once per cycle,
// and inlined its polynomial constants would be hoisted into registers.
__device__ __noinline__ double accept_exp(double x) { return exp(x); }

__device__ __forceinline__ double unit(uint64_t h) { return static_cast<double>(h >> 11) * 0x1.0p-53; }

__device__ __forceinline__ uint32_t round_fn(uint32_t r, uint32_t key) {
  uint32_t x = (r + key) * 0x9E3779B1u;
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  return x;
}

// Inclusive prefix sums over the 64 lanes of a wave in DPP moves: shifts by
// 1, 2, 4, 8 inside each row of 16 lanes (lanes shifted in from outside the
// row read 0), then lane 15 of each row added to the next row and lane 31 to
// the upper two rows.
template <int CTRL, int ROWS>
__device__ __forceinline__ unsigned dpp_u32(unsigned v) {
  return static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, ROWS, 0xF, true));
}
__device__ __forceinline__ unsigned wave_scan_u32(unsigned v) {
  v += dpp_u32<0x111, 0xF>(v);
  v += dpp_u32<0x112, 0xF>(v);
  v += dpp_u32<0x114, 0xF>(v);
  v += dpp_u32<0x118, 0xF>(v);
  v += dpp_u32<0x142, 0xA>(v);
  v += dpp_u32<0x143, 0xC>(v);
  return v;
}
template <int CTRL, int ROWS>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
  return static_cast<unsigned long long>(dpp_u32<CTRL, ROWS>(static_cast<unsigned>(v >> 32))) << 32 |
         dpp_u32<CTRL, ROWS>(static_cast<unsigned>(v));
}
__device__ __forceinline__ unsigned long long wave_scan_u64(unsigned long long v) {
  v += dpp_u64<0x111, 0xF>(v);
  v += dpp_u64<0x112, 0xF>(v);
  v += dpp_u64<0x114, 0xF>(v);
  v += dpp_u64<0x118, 0xF>(v);
  v += dpp_u64<0x142, 0xA>(v);
  v += dpp_u64<0x143, 0xC>(v);
  return v;
}
// Wave total of a 32-bit value (the last lane of the inclusive sums).
__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
  return static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(wave_scan_u32(v)), 63));
}

// Keyed pseudo-random permutation of [0, n): a four-round Feistel network
// over the mixed-radix domain [0, A * B), A = 2^a with a = floor(bits / 2)
// (bits = ceil(log2 n)) and B = ceil(n / A), a position x being the pair
// (hi, lo) = (x >> a, x mod A). The rounds add keyed functions of one half to
// the other, modulo its radix: lo += F0(hi), hi += F1(lo), lo += F2(hi),
// hi += F3(lo) (an F for hi scaled into [0, B) by a high multiply), each
// undone by a subtraction. It is cycle-walked into [0, n): the row at
// position y < n is the first of network_inverse(y),
// network_inverse(network_inverse(y)), ... below n (the orbit of y under the
// bijection returns below n, so the walk ends). The domain exceeds n by less
// than A, about sqrt(n), so nearly every walk takes one step (a power-of-two
// domain of up to 2n took up to twice as many on average).
struct Permutation {
  uint32_t k0, k1, k2, k3;
  int a;       // bits of lo
  uint32_t b;  // radix of hi

  // Round keys from a 64-bit key: its halves, then the halves of mix64(key).
  __device__ static void round_keys(uint64_t key, uint32_t (&k)[4]) {
    const uint64_t h = mix64(key);
    k[0] = static_cast<uint32_t>(key);
    k[1] = static_cast<uint32_t>(key >> 32);
    k[2] = static_cast<uint32_t>(h);
    k[3] = static_cast<uint32_t>(h >> 32);
  }
  __device__ Permutation(uint32_t r0, uint32_t r1, uint32_t r2, uint32_t r3, int bits, int n)
      : k0(r0), k1(r1), k2(r2), k3(r3), a(bits >> 1),
        b(static_cast<uint32_t>((n + (1 << (bits >> 1)) - 1) >> (bits >> 1))) {}

  __device__ __forceinline__ uint32_t sub_hi(uint32_t hi, uint32_t lo, uint32_t key) const {
    const uint32_t t = __umulhi(round_fn(lo, key), b);  // in [0, b)
    return hi >= t ? hi - t : hi + b - t;
  }

  __device__ __forceinline__ uint32_t network_inverse(uint32_t y) const {
    const uint32_t mask = (1u << a) - 1u;
    uint32_t hi = y >> a, lo = y & mask;
    hi = sub_hi(hi, lo, k3);
    lo = (lo - round_fn(hi, k2)) & mask;
    hi = sub_hi(hi, lo, k1);
    lo = (lo - round_fn(hi, k0)) & mask;
    return (hi << a) | lo;
  }
};

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
  double* wave_total;           // [kWaves]
  double* state;                // [4] cur_acc, cur_score, best_acc, best_score
  unsigned long long* weight;   // [C] absorption weight per target, times n (exact)
  unsigned long long* best_key; // [C] packed best target per source
  int* sizes;                   // [C] class sizes of the proposal (the current state's before relabelling)
  int* stay;                    // [C] rows per label whose prediction differs (the outflow)
  int* counts;                  // [C] class counts of the current state
  int* nsrc;                    // [C] absorbed sources per target
  int* point;                   // [C] target of each source or -1
  int* pairs;                   // [max_rows] distinct flow entries (label * C + prediction)
  int* queue;                   // [max_rows] rows left to the full vote; flow keys while proposing
  int* scal;                    // [16] scalars
  unsigned short* labels;       // [4 x lmax] label buffers 0, 1 then prediction buffers 0, 1
  int lmax;                     // rows of the largest run + the sentinel slot, even
  // LDS flow table (EvoParams::hash_bits > 0): per slot a flow entry (label *
  // C + prediction, kEmptySlot when free) and its count.
  unsigned* hkey;
  int* hcnt;
  // Learned variant only: the network's tiled weights [kMlpTiled] (16-byte
  // aligned); per wave (kLearnedWaves) a tile of 16 rows x 64 features (bf16
  // bits) and 16 rows x 32 slots (ranked codes 0..29, the row's index in slot
  // 31); bf16(c / k) for c = 0..64.
  uint32_t* mlp_w;
  unsigned short* mlp_tile;
  unsigned short* mlp_rank;
  unsigned short* mlp_frac;
  // Buffer b of the labels / predictions (the current state and the proposal
  // swap on acceptance); plain arithmetic keeps the struct in registers.
  __device__ unsigned short* lab(int b) const { return labels + b * lmax; }
  __device__ unsigned short* pred(int b) const { return labels + (2 + b) * lmax; }
};

constexpr unsigned kEmptySlot = 0xFFFFFFFFu;  // above every flow entry (< 1024^2)

// kSumSq: sum of the proposal's squared class sizes; kMulti: some target has
// two sources or more; kVanished: classes absorbed into kChosen (0: none);
// kMoved / kMovedSq: rows and sum of squared sizes of those classes.
// kPending (split runs): bit 0, the other slices' predictions of the current
// state are still to be read (from exchange half kPendingPar, epoch
// kPendingEp); bit 1, the current state is a new best to be written.
// kTake .. kAbsorbHi: the next cycle's draws, made once per workgroup (see
// plan_cycle).
enum Scalar {
  kPresent, kPairs, kMatches, kChosen, kDecision, kCur, kQueue, kCands, kSumSq, kMulti, kVanished, kMoved, kMovedSq,
  kPending, kPendingEp, kPendingPar, kTake, kPerm0, kPerm1, kPerm2, kPerm3, kLearnCycle, kRowKeyLo, kRowKeyHi,
  kAbsorbLo, kAbsorbHi, kAnyTarget, kScalars
};
constexpr int kScalarSlots = 32;
static_assert(kScalars <= kScalarSlots, "scalar slots");

__host__ __device__ constexpr int label_slots(int max_rows) { return (max_rows + 2) & ~1; }

__device__ __forceinline__ Lds carve(unsigned char* base, int C, int max_rows, int* large_rows, int hash_bits) {
  Lds s;
  double* d = reinterpret_cast<double*>(base);
  s.wave_total = d;
  s.state = d + kWaves;
  s.weight = reinterpret_cast<unsigned long long*>(d + kWaves + 4);
  s.best_key = s.weight + C;
  int* i = reinterpret_cast<int*>(s.best_key + C);
  s.sizes = i;
  s.stay = i + C;
  s.counts = i + 2 * C;
  s.nsrc = i + 3 * C;
  s.point = i + 4 * C;
  if (large_rows != nullptr) {
    // Large-run variant: the flow list and the row queue in global scratch.
    s.pairs = large_rows;
    s.queue = large_rows + max_rows;
    s.scal = i + 5 * C;
  } else {
    s.pairs = i + 5 * C;
    s.queue = s.pairs + max_rows;
    s.scal = s.queue + max_rows;
  }
  s.labels = reinterpret_cast<unsigned short*>(s.scal + kScalarSlots);
  s.lmax = label_slots(max_rows);
  const int slots = hash_bits > 0 ? 1 << hash_bits : 0;
  s.hkey = reinterpret_cast<unsigned*>(s.labels + 4 * s.lmax);
  s.hcnt = reinterpret_cast<int*>(s.hkey + slots);
  const size_t used = static_cast<size_t>(reinterpret_cast<unsigned char*>(s.hcnt + slots) - base);
  s.mlp_w = reinterpret_cast<uint32_t*>(base + ((used + 15) & ~static_cast<size_t>(15)));
  s.mlp_tile = reinterpret_cast<unsigned short*>(s.mlp_w + kMlpTiled);
  s.mlp_rank = s.mlp_tile + kLearnedWaves * mlp::kTileRows * kWave;
  s.mlp_frac = s.mlp_rank + kLearnedWaves * mlp::kTileRows * 32;
  return s;
}

// Bitonic sorting network on a register array (every index is a compile-time
// constant after unrolling, so the array stays in registers).
template <int N>
__device__ __forceinline__ void sort_network(uint32_t (&a)[N]) {
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

// Two 16-bit lanes per register: min and max of each lane (v_pk_min_u16 /
// v_pk_max_u16).
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_min16(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_max16(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}

// Bitonic sort of 2M 16-bit keys held two to a register: element i is half
// i / M of register i % M. Every comparator sends the minimum to the lower
// index; each merge of size s opens by comparing i with i ^ (s - 1) (the
// mirror image, which makes the second half of a bitonic merge ascending too)
// and continues with i ^ d for d = s/4 .. 1. Comparators between registers
// act on both halves with one packed min and one packed max; only the
// opening of the last merge pairs the low half of register r with the high
// half of register M - 1 - r.
template <int M>
__device__ __forceinline__ void sort_packed(uint32_t (&v)[M]) {
#pragma unroll
  for (int size = 2; size <= 2 * M; size <<= 1) {
    if (size <= M) {
#pragma unroll
      for (int r = 0; r < M; ++r) {
        const int q = r ^ (size - 1);
        if (q > r) {
          const uint32_t a = v[r], b = v[q];
          v[r] = pk_min16(a, b);
          v[q] = pk_max16(a, b);
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < M / 2; ++r) {
        const int q = M - 1 - r;
        const uint32_t a = v[r], b = __builtin_amdgcn_alignbit(v[q], v[q], 16);  // halves of v[q] swapped
        const uint32_t mn = pk_min16(a, b), mx = pk_max16(a, b);
        v[r] = (mn & 0xFFFFu) | (mx & 0xFFFF0000u);
        v[q] = __builtin_amdgcn_alignbit(mx, mn, 16);  // high half of mn, low half of mx
      }
    }
#pragma unroll
    for (int d = size >> 2; d > 0; d >>= 1) {
#pragma unroll
      for (int r = 0; r < M; ++r) {
        const int q = r ^ d;
        if (q > r) {
          const uint32_t a = v[r], b = v[q];
          v[r] = pk_min16(a, b);
          v[q] = pk_max16(a, b);
        }
      }
    }
  }
}

// Full vote of row i (core.cpp:871-911): most votes, then the larger double sum
// of scores taken in neighbour-rank order, then the smaller label. Neighbours
// become keys (code << kRankBits | rank) sorted ascending, so each label's
// votes form one run with ranks ascending inside it; missing neighbours (the
// sentinel code) sort last. Up to 32 neighbours a key has 5 rank bits and an
// 11-bit code (labels up to 1023, a missing neighbour as 1024 + rank) and two
// keys share a register in the sort; at 64 the keys are 32-bit. Returns -1
// for a row without neighbours. TWO_PASS picks the packed keys' two-pass run
// scan (fewer instructions); the 512-thread learned launch, whose scoring
// holds more registers, keeps the one-pass scan, which does not spill there.
template <int KMAX, bool TWO_PASS>
__device__ __forceinline__ int full_vote(int k, int pad_total, const unsigned short* __restrict__ nbr,
                                         const float* __restrict__ scores, const unsigned short* lab, int i) {
  uint32_t key[KMAX];
  int valid = 0;
  // All of the row's id loads are in flight before the first lookup, and the
  // lookups go out eight at a time; ranks past k re-read rank k - 1 and are
  // masked afterwards, so neither needs a branch.
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    const int jj = j < k ? j : k - 1;
    key[j] = nbr[static_cast<uint32_t>(jj * pad_total + i)];
  }
  __builtin_amdgcn_sched_barrier(0);
  constexpr int kLook = KMAX < 8 ? KMAX : 8;
#pragma unroll
  for (int j0 = 0; j0 < KMAX; j0 += kLook) {
#pragma unroll
    for (int j = j0; j < j0 + kLook; ++j) key[j] = lab[key[j]];
    __builtin_amdgcn_sched_barrier(0);
  }
  constexpr bool kPacked = KMAX <= 32;
  constexpr int kRankBits = kPacked ? 5 : 6;
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    const uint32_t c = j < k ? key[j] : kSentinel;
    valid += c != kSentinel ? 1 : 0;
    // Packed keys: a missing neighbour's code is above every label; for the
    // two-pass scan one of its own (1024 + rank), so missing neighbours never
    // form a run longer than one (the one-pass scan stops at the present
    // ones and takes 2047 for all).
    const uint32_t missing = TWO_PASS ? 1024u + static_cast<uint32_t>(j) : 2047u;
    const uint32_t code = kPacked && c == kSentinel ? missing : c;
    key[j] = (code << kRankBits) | static_cast<uint32_t>(j);
  }
  if constexpr (kPacked) {
    constexpr int M = KMAX / 2;
    uint32_t v[M];
#pragma unroll
    for (int r = 0; r < M; ++r) v[r] = key[r] | key[r + M] << 16;
    sort_packed<M>(v);
#pragma unroll
    for (int r = 0; r < M; ++r) {
      key[r] = v[r] & 0xFFFFu;
      key[r + M] = v[r] >> 16;
    }
  } else {
    sort_network<KMAX>(key);
  }
  // Longest run of one code among the valid entries (the smallest such code)
  // and whether another run is as long, in integer selects: branches per
  // entry would cost the wave an exec-mask update each.
  int best_count = 0, best_code = -1, run = 0;
  bool tie = false;
  if constexpr (kPacked && TWO_PASS) {
    // Two passes: the longest run, then how many runs reach it and the first
    // of them. Missing neighbours sort last as runs of one, so they reach the
    // longest run only when it is 1: then a tie means two present votes.
    uint32_t prev = 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      const uint32_t code = key[j] >> kRankBits;
      run = code == prev ? run + 1 : 1;
      prev = code;
      best_count = max(best_count, run);
    }
    int hits = 0;
    uint32_t first = 0xFFFFFFFFu;
    prev = 0xFFFFFFFFu;
    run = 0;
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      const uint32_t code = key[j] >> kRankBits;
      run = code == prev ? run + 1 : 1;
      prev = code;
      const bool hit = run == best_count;
      hits += hit ? 1 : 0;
      first = min(first, hit ? code : 0xFFFFFFFFu);
    }
    best_code = valid > 0 ? static_cast<int>(first) : -1;
    tie = best_count == 1 ? valid > 1 : hits > 1;
  } else {
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      const uint32_t code = key[j] >> kRankBits;
      run = (j > 0 && code == (key[j - 1] >> kRankBits)) ? run + 1 : 1;
      const bool in = j < valid;
      const bool longer = in && run > best_count;
      tie = !longer && (tie || (in && run == best_count));
      best_code = longer ? static_cast<int>(code) : best_code;
      best_count = longer ? run : best_count;
    }
  }
  if (!tie) return best_code;
  // The scores are read only here, all of a row's loads in flight before the
  // first sum: a wave of 64 rows almost always holds a tie, so the tie path's
  // load latency is paid by the whole wave (missing neighbours read rank 0 and
  // are not summed). Element offsets fit 32 bits (k * pad_total < 2^31,
  // checked at launch), so each load is a scalar base plus one offset register.
  constexpr int kGroup = KMAX;
  double best_sum = -INFINITY, sum = 0.0;
#pragma unroll
  for (int j0 = 0; j0 < KMAX; j0 += kGroup) {
    float score[kGroup];
#pragma unroll
    for (int t = 0; t < kGroup; ++t) {
      const uint32_t rank = j0 + t < valid ? key[j0 + t] & ((1u << kRankBits) - 1u) : 0u;
      score[t] = scores[static_cast<uint32_t>(static_cast<int>(rank) * pad_total + i)];
    }
#pragma unroll
    for (int t = 0; t < kGroup; ++t) {
      const int j = j0 + t;
      const bool fresh = j == 0 || (key[j] >> kRankBits) != (key[j - 1] >> kRankBits);
      run = fresh ? 1 : run + 1;
      sum = (fresh ? 0.0 : sum) + static_cast<double>(score[t]);
      const bool better = j < valid && run == best_count && sum > best_sum;
      best_sum = better ? sum : best_sum;
      best_code = better ? static_cast<int>(key[j] >> kRankBits) : best_code;
    }
  }
  return best_code;
}

// Neighbour ids of rows 2m and 2m+1 (low and high halves), one 32-bit load per
// neighbour rank; run bases and the table stride are multiples of 256 rows.
// Ranks past k get the sentinel slot in both halves, so the lookups that
// follow need no branch.
template <int KMAX>
__device__ __forceinline__ void load_pair(const uint32_t* __restrict__ nbr2, int pad2, int k, int m,
                                          uint32_t sentinel2, uint32_t (&ids)[KMAX]) {
#pragma unroll
  for (int j = 0; j < KMAX; ++j) ids[j] = j < k ? nbr2[static_cast<size_t>(j) * pad2 + m] : sentinel2;
}

// Vote of the labels in `lab` into `pred` and the number of rows whose
// prediction equals their label; with `count_labels` also the class counts of
// `lab`. A row whose nearest neighbour's label holds more than k/2 of the k
// votes is settled at once (the sentinel of a missing neighbour matches no
// label, so this implies a strict majority of the present neighbours); the
// rest are queued for the full vote.
template <int KMAX, int THREADS>
__device__ __forceinline__ void vote_phase(int tid, int k, int pad_total, const Lds& s,
                                           const unsigned short* __restrict__ nbr,
                                           const float* __restrict__ scores, const unsigned short* lab,
                                           unsigned short* pred, int n, bool count_labels, int pair_lo,
                                           int pair_hi) {
  const uint32_t* __restrict__ nbr2 = reinterpret_cast<const uint32_t*>(nbr);
  const int pad2 = pad_total >> 1;
  const int npairs = pair_hi;
  const int lim = min(n, 2 * pair_hi);  // rows past the range are another member's
  const int lane = tid & (kWave - 1);
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t sentinel2 = static_cast<uint32_t>(n) | static_cast<uint32_t>(n) << 16;
  int matches = 0;
  // Uniform trip count: every lane reaches the ballots below. One pair of
  // rows per thread and pass: with the lookups issued in groups, loading two
  // pairs' ids first (as the 8-wave learned launch once did) gains nothing.
  constexpr int kBatch = 1;
  for (int m0 = pair_lo; m0 < npairs; m0 += kBatch * THREADS) {
    uint32_t ids_b[kBatch][KMAX];
#pragma unroll
    for (int b = 0; b < kBatch; ++b) {
      const int m = m0 + b * THREADS + tid;
      if (m < npairs) load_pair<KMAX>(nbr2, pad2, k, m, sentinel2, ids_b[b]);
    }
    // The thread's rows first (no LDS atomic between them, so their lookups
    // overlap), then one queue allocation for all of the wave's undecided rows.
    uint64_t mask[2 * kBatch];
    int queued_total = 0;
#pragma unroll
    for (int hb = 0; hb < 2 * kBatch; ++hb) {
      const int h = hb & 1;
      const uint32_t(&ids)[KMAX] = ids_b[hb >> 1];
      const int i = 2 * (m0 + (hb >> 1) * THREADS + tid) + h;
      bool queued = false;
      if (i < lim) {
        const int own = lab[i];
        if (count_labels) atomicAdd(&s.counts[own], 1);
        // Ranks past k read the sentinel slot, which matches code0 only when
        // code0 is the sentinel itself, and such a row is queued anyway.
        // Lookups go out eight at a time before their compares.
        constexpr int kGroup = KMAX < 8 ? KMAX : 8;
        int code0 = 0, c0 = 0;
#pragma unroll
        for (int j0 = 0; j0 < KMAX; j0 += kGroup) {
          int code[kGroup];
#pragma unroll
          for (int t = 0; t < kGroup; ++t) code[t] = lab[ids[j0 + t] >> (16 * h) & 0xFFFFu];
          if (j0 == 0) code0 = code[0];
#pragma unroll
          for (int t = 0; t < kGroup; ++t) c0 += code[t] == code0 ? 1 : 0;
          __builtin_amdgcn_sched_barrier(0);
        }
        if (code0 != kSentinel && 2 * c0 > k) {
          pred[i] = static_cast<unsigned short>(code0);
          matches += code0 == own ? 1 : 0;
        } else {
          queued = true;
        }
      }
      mask[hb] = __ballot(queued);
      queued_total += __popcll(mask[hb]);
    }
    if (queued_total) {
      int slot = 0;
      if (lane == 0) slot = atomicAdd(&s.scal[kQueue], queued_total);
      slot = __shfl(slot, 0);
#pragma unroll
      for (int hb = 0; hb < 2 * kBatch; ++hb) {
        if ((mask[hb] >> lane) & 1ull) {
          s.queue[slot + __popcll(mask[hb] & below)] = 2 * (m0 + (hb >> 1) * THREADS + tid) + (hb & 1);
        }
        slot += __popcll(mask[hb]);
      }
    }
  }
  __syncthreads();
  const int qn = s.scal[kQueue];
  for (int q = tid; q < qn; q += THREADS) {
    const int i = s.queue[q];
    const int own = lab[i];
    int best = full_vote<KMAX, THREADS == 1024>(k, pad_total, nbr, scores, lab, i);
    // Every row of a batch run has a neighbour (each run has two folds or
    // more); a row without one would keep its label.
    if (best < 0) best = own;
    pred[i] = static_cast<unsigned short>(best);
    matches += best == own ? 1 : 0;
  }
#pragma unroll
  for (int off = kWave / 2; off >= 1; off >>= 1) matches += __shfl_xor(matches, off);
  if (lane == 0 && matches != 0) atomicAdd(&s.scal[kMatches], matches);
  __syncthreads();
}

// Objective with the diversity guard (wave 0), from the number of classes
// and the sum of their squared counts in integers (at most n^2 < 2^28),
// divided by n^2 once. The evaluation of the starting labels counts them in
// s.counts; a proposal's come from its class sizes and the absorbed classes:
// absorbing classes of m rows and squared sizes q into a target of t rows
// adds (t + m)^2 - t^2 - q.
__device__ __forceinline__ double objective(const Lds& s, bool guarded, bool evaluate, int C, int n, int tid,
                                            double* acc_out) {
  const int lane = tid & (kWave - 1);  // from opaque_tid: the count addresses are not kept across the cycle loop
  int k = 0;
  unsigned sq = 0u;
  if (evaluate) {
    for (int c = lane; c < C; c += kWave) {
      const int v = s.counts[c];
      if (v > 0) {
        ++k;
        sq += static_cast<unsigned>(v) * static_cast<unsigned>(v);
      }
    }
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) {
      k += __shfl_xor(k, m);
      sq += __shfl_xor(sq, m);
    }
  } else {
    const int vanished = s.scal[kVanished];
    k = s.scal[kPresent] - vanished;
    sq = static_cast<unsigned>(s.scal[kSumSq]);
    if (vanished > 0) {
      const unsigned t = static_cast<unsigned>(s.sizes[s.scal[kChosen]]);
      const unsigned m = static_cast<unsigned>(s.scal[kMoved]);
      sq = sq - static_cast<unsigned>(s.scal[kMovedSq]) + 2u * t * m + m * m;
    }
  }
  const double nn = static_cast<double>(n);
  const double acc = n ? static_cast<double>(s.scal[kMatches]) / nn : 0.0;
  *acc_out = acc;
  if (!guarded) return acc;
  const double same = n ? static_cast<double>(sq) / (nn * nn) : 0.0;
  return k == 1 ? -INFINITY : acc * sqrt(fmax(0.0, 1.0 - same));
}

}  // namespace

template <int KMAX, int THREADS, bool LEARNED, bool LARGE>
__global__ void __launch_bounds__(THREADS) evo_persistent_kernel(EvoParams p) {
  // Rows per thread and pass of the flow counting, for its register arrays
  // (one pass covers kEvoMaxRows rows at 1024 threads).
  constexpr int kRowsPerThread = (kEvoMaxRows + THREADS - 1) / THREADS;
  // Classes per thread of the per-class phases (label codes are below 1024).
  constexpr int kClassesPerThread = (1024 + THREADS - 1) / THREADS;
  extern __shared__ __align__(16) double lds_raw[];
  const int C = p.classes;
  const int wg = blockIdx.x;
  const int run = p.wg_run ? p.wg_run[wg] : wg;
  const int member = p.wg_run ? p.wg_member[wg] : 0;
  const int group = p.wg_run ? p.run_group[run] : 1;
  const bool lead = member == 0;  // writes the run's results
  // LDS flow table (compile-time off for the large-run layout).
  const int hash_bits = LARGE ? 0 : p.hash_bits;
  const bool hashed = hash_bits > 0;
  const Lds s = carve(reinterpret_cast<unsigned char*>(lds_raw), C, p.max_rows,
                      LARGE ? p.large_scratch + static_cast<size_t>(wg) * 2 * p.max_rows : nullptr, hash_bits);
  const int n = p.run_offsets[run + 1] - p.run_offsets[run];
  // This member's slice of rows for the vote (even start, so whole id pairs).
  const int share = ((n + group - 1) / group + 1) & ~1;
  const int slice_lo = min(n, member * share), slice_hi = min(n, slice_lo + share);
  const int base = p.pad_offsets[run];
  const unsigned short* __restrict__ nbr = p.neighbors + base;
  const float* __restrict__ scores = p.scores + base;
  const uint64_t key = p.run_key[run];
  const int bits = p.perm_bits[run];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  int* flows = p.flows + static_cast<size_t>(wg) * C * C;
  const int k = p.k;
  const int pad_total = p.pad_total;
  const bool guarded = p.guarded != 0;

  // The workgroup's elapsed time: its start subtracted here, its end added
  // before the write-back (one thread, one word per workgroup).
  if (threadIdx.x == 0 && p.wg_time != nullptr) p.wg_time[wg] -= wall_clock64();
  // Per-launch set-up: clear the class tables, load the run state.
  for (int c = threadIdx.x; c < C; c += THREADS) {
    s.weight[c] = 0ull;
    s.best_key[c] = 0ull;
    s.sizes[c] = 0;
    s.stay[c] = 0;
    s.counts[c] = 0;
    s.nsrc[c] = 0;
    s.point[c] = -1;
  }
  if (threadIdx.x < kScalarSlots) s.scal[threadIdx.x] = 0;
  if (hashed) {
    for (int e = threadIdx.x; e < 1 << hash_bits; e += THREADS) {
      s.hkey[e] = kEmptySlot;
      s.hcnt[e] = 0;
    }
  }
  for (int i = threadIdx.x; i < n; i += THREADS) {
    s.lab(0)[i] = p.cur[base + i];
    // Before the first evaluation the predictions start equal to the labels
    // (that evaluation rewrites them).
    s.pred(0)[i] = p.init ? p.cur[base + i] : p.curpred[base + i];
    s.lab(1)[i] = 0;  // no stale chosen marks in the first proposal buffer
  }
  if (threadIdx.x < 2) s.lab(threadIdx.x)[n] = static_cast<unsigned short>(kSentinel);
  if (threadIdx.x < 4) s.state[threadIdx.x] = p.state[static_cast<size_t>(run) * kEvoStateStride + threadIdx.x];
  if (LEARNED) {
    for (int e = threadIdx.x; e < kMlpTiled; e += THREADS) s.mlp_w[e] = p.learn_packed[e];
    if (threadIdx.x <= 64) {
      s.mlp_frac[threadIdx.x] =
        static_cast<unsigned short>(mlp::bf16_bits(__fdiv_rn(static_cast<float>(threadIdx.x), static_cast<float>(k))));
    }
  }
  __syncthreads();

  // Split runs (see EvoParams): a word of the exchange carries its step's
  // epoch (1..65535, never the zero of a cleared word, different for
  // consecutive steps); a read waits for the expected epoch. A wait that runs
  // out of polls (or finds that another one did) flags the error and stops
  // waiting, so every member finishes.
  auto await = [&](unsigned* word, int shift, unsigned ep) {
    unsigned w = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int polls = 0;
    while ((w >> shift) != ep) {
      if (++polls > p.spin_limit || __hip_atomic_load(p.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        __hip_atomic_store(p.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
      w = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return w;
  };
  // The other members' predictions of a step (epoch ep, exchange half par)
  // into `pred`, for this thread's rows i = tid + j * THREADS only (the rows
  // every per-row loop gives the thread, so no barrier is needed before the
  // thread reads them). Eight words in flight at once (the first eight may
  // have been loaded earlier into `first`); a word not yet at the epoch is
  // then awaited on its own.
  constexpr int kAhead = 8;
  auto exchange_half = [&](int par) { return p.xbuf + static_cast<size_t>(par) * p.pad_total + base; };
  auto load_words = [&](unsigned (&w)[kAhead], const unsigned* xb, int i0) {
#pragma unroll
    for (int u = 0; u < kAhead; ++u) {
      const int i = min(i0 + u * THREADS, n - 1);
      w[u] = __hip_atomic_load(&xb[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  };
  auto fill_preds = [&](unsigned short* pred, unsigned ep, int par, int tid, unsigned (&first)[kAhead],
                        bool loaded) {
    const unsigned* xb = exchange_half(par);
    for (int i0 = tid; i0 < n; i0 += kAhead * THREADS) {
      unsigned w[kAhead];
      if (i0 == tid && loaded) {
#pragma unroll
        for (int u = 0; u < kAhead; ++u) w[u] = first[u];
      } else {
        load_words(w, xb, i0);
      }
#pragma unroll
      for (int u = 0; u < kAhead; ++u) {
        const int i = i0 + u * THREADS;
        if (i >= n || (i >= slice_lo && i < slice_hi)) continue;
        if ((w[u] >> 16) != ep) w[u] = await(const_cast<unsigned*>(&xb[i]), 16, ep);
        pred[i] = static_cast<unsigned short>(w[u] & 0xFFFFu);
      }
    }
  };
  // A split run's step ends once the members have the match totals: the
  // predictions of the other slices are read when first needed, by the next
  // step's relabelling (loaded before its marking, stored after) or before
  // the write-back, and only when the step's proposal became the current
  // state (kPending). The best state of an improving step is written then
  // too.
  auto settle = [&](int pending, int tid, unsigned (&first)[kAhead], bool loaded) {
    const int cur = s.scal[kCur];
    fill_preds(s.pred(cur), static_cast<unsigned>(s.scal[kPendingEp]), s.scal[kPendingPar], tid, first, loaded);
    if ((pending & 2) && lead) {
      for (int i = tid; i < n; i += THREADS) {
        p.best[base + i] = s.lab(cur)[i];
        p.bestpred[base + i] = s.pred(cur)[i];
      }
    }
  };

  // A cycle's draws (proposal size, permutation round keys, the learned
  // variant's fallback choice and row key, the absorption draw) are made by
  // one thread of the workgroup and read from LDS: made by every wave, their
  // 64-bit hashing would queue on the compute unit's one scalar unit. A lane
  // of the last wave plans the next cycle during the acceptance test (thread
  // 0 the first one, before the loop); the barriers that follow publish it.
  auto plan_cycle = [&](int cycle) {
    const int hi = proposal_hi(n, cycle, p.cycles);
    s.scal[kTake] = min(hi, 1 + static_cast<int>(unit(draw(key, cycle, 1)) * hi));
    uint32_t rk[4];
    Permutation::round_keys(draw(key, cycle, 2), rk);
    s.scal[kPerm0] = static_cast<int>(rk[0]);
    s.scal[kPerm1] = static_cast<int>(rk[1]);
    s.scal[kPerm2] = static_cast<int>(rk[2]);
    s.scal[kPerm3] = static_cast<int>(rk[3]);
    s.scal[kLearnCycle] = LEARNED && unit(draw(key, cycle, 5)) >= p.learn_fallback ? 1 : 0;
    const uint64_t row_key = LEARNED ? draw(key, cycle, 6) : 0ull;
    s.scal[kRowKeyLo] = static_cast<int>(static_cast<uint32_t>(row_key));
    s.scal[kRowKeyHi] = static_cast<int>(static_cast<uint32_t>(row_key >> 32));
    const uint64_t absorb = draw(key, cycle, 3);
    s.scal[kAbsorbLo] = static_cast<int>(static_cast<uint32_t>(absorb));
    s.scal[kAbsorbHi] = static_cast<int>(static_cast<uint32_t>(absorb >> 32));
  };
  const auto scal_u64 = [&](int lo) {
    return static_cast<uint64_t>(static_cast<uint32_t>(s.scal[lo + 1])) << 32 | static_cast<uint32_t>(s.scal[lo]);
  };
  if (!p.init && threadIdx.x == 0 && p.cycle_count > 0) plan_cycle(p.cycle_begin);
  // The class counts of the current state (a launch that starts with the
  // evaluation counts them in its vote), then the first proposal's sizes.
  if (!p.init) {
    for (int i = threadIdx.x; i < n; i += THREADS) atomicAdd(&s.counts[s.lab(0)[i]], 1);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += THREADS) s.sizes[c] = s.counts[c];

  // Step -1 evaluates the starting labels in place: their predictions and
  // scores become the current and the best state.
  for (int step = p.init ? -1 : 0; step < p.cycle_count; ++step) {
    const bool evaluate = step < 0;
    const int tid = opaque_tid();
    const int cycle = p.cycle_begin + step;
    const int cur = s.scal[kCur];
    const unsigned short* cur_lab = s.lab(cur);
    const unsigned short* cur_pred = s.pred(cur);
    unsigned short* prop = s.lab(evaluate ? cur : cur ^ 1);
    unsigned short* prop_pred = s.pred(evaluate ? cur : cur ^ 1);

    if (!evaluate) {
      // -- propose --------------------------------------------------------------
      {
        const int take = s.scal[kTake];
        // A learned cycle (learned variant, unless the fallback draw picks the
        // uniform proposal) first scores the rows whose label differs from
        // their prediction; the others score 0. The candidates are listed
        // (in the flow-key buffer, free until the flows are counted) and
        // dealt to the waves in tiles of 16; one lane per row computes the
        // row's vote (a register sort of its neighbours' codes), and a
        // matrix-core forward pass scores each tile. The target label
        // goes straight into the proposal buffer and the score into the row
        // queue. Each row is then chosen with probability take * score / (sum
        // of scores), capped at 1; the others get their label back. Otherwise
        // the rows at positions [0, take) of a keyed permutation are marked
        // and take their prediction. The proposal buffer holds only label
        // codes otherwise.
        const bool learned = LEARNED && s.scal[kLearnCycle] != 0;
        // The learned scoring lists its candidates by their predictions.
        const int pending = s.scal[kPending];
        unsigned first[kAhead];
        if (LEARNED && (pending & 1)) settle(pending, tid, first, false);
        if (!LEARNED && (pending & 1)) load_words(first, exchange_half(s.scal[kPendingPar]), tid);
        if (learned) {
          const int ln = tid & (kWave - 1);  // from opaque_tid: not hoisted out of the cycle loop
          float* score = reinterpret_cast<float*>(s.queue);
          int* cands = s.pairs;
          {
            const uint64_t below = (1ull << ln) - 1ull;
            for (int b0 = wave * kWave; b0 < n; b0 += THREADS) {  // uniform in the wave
              const int i = b0 + ln;
              const bool cand = i < n && cur_lab[i] != cur_pred[i];
              if (i < n && !cand) score[i] = 0.0f;
              const uint64_t mask = __ballot(cand);
              if (mask) {
                int at = 0;
                if (ln == 0) at = atomicAdd(&s.scal[kCands], __popcll(mask));
                at = __shfl(at, 0);
                if (cand) cands[at + __popcll(mask & below)] = i;
              }
            }
          }
          __syncthreads();
          const int nc = s.scal[kCands];
          unsigned short* tile = s.mlp_tile + wave * mlp::kTileRows * kWave;
          unsigned short* rk = s.mlp_rank + wave * mlp::kTileRows * 32;
          const int g = ln >> 4, slot = ln & 15;
          // Tiles of 16 candidates are dealt to the waves in turn; lane group
          // g of a wave holds tile t0 + g * waves.
          constexpr int kNw = THREADS / kWave;
          for (int t0 = wave; t0 * mlp::kTileRows < nc; t0 += 4 * kNw) {  // uniform in the wave
            const int c = (t0 + g * kNw) * mlp::kTileRows + slot;
            const bool have = c < nc;
            const int row = have ? cands[c] : 0;
            // Neighbour codes, eight loads in flight at a time.
            uint32_t vkey[KMAX];
#pragma unroll
            for (int j0 = 0; j0 < KMAX; j0 += 8) {
#pragma unroll
              for (int j = j0; j < j0 + 8 && j < KMAX; ++j) {
                const int jj = j < k ? j : k - 1;
                const uint32_t id = nbr[static_cast<uint32_t>(jj * pad_total + row)];
                vkey[j] = have && j < k ? id : static_cast<uint32_t>(n);
              }
              // Slot n (a missing neighbour, a rank past k, a lane past the
              // list) holds the sentinel code, which is mlp::kNone.
#pragma unroll
              for (int j = j0; j < j0 + 8 && j < KMAX; ++j) vkey[j] = cur_lab[vkey[j]];
              __builtin_amdgcn_sched_barrier(0);
            }
            const mlp::RowVote vote = mlp::row_vote<KMAX>(vkey, cur_lab[row]);
            for (int q = 0; q < 4 && (t0 + q * kNw) * mlp::kTileRows < nc; ++q) {  // uniform in the wave
              const int cnt = min(mlp::kTileRows, nc - (t0 + q * kNw) * mlp::kTileRows);
              if (g == q) {
                // This lane's row becomes tile row `slot` (zeros past the list).
                const auto frac = [&](int c) { return __uint_as_float(static_cast<uint32_t>(s.mlp_frac[c]) << 16); };
#pragma unroll
                for (int v4 = 0; v4 < kMlpIn / 8; ++v4) {
                  uint32_t word[4];
#pragma unroll
                  for (int w = 0; w < 4; ++w) {
                    const int j = 8 * v4 + 2 * w;
                    const uint32_t lo = mlp::bf16_bits(mlp::feature<KMAX>(vkey, vote, j, frac));
                    const uint32_t hi = mlp::bf16_bits(mlp::feature<KMAX>(vkey, vote, j + 1, frac));
                    word[w] = have ? lo | hi << 16 : 0u;
                  }
                  *reinterpret_cast<uint4*>(tile + slot * kWave + 8 * v4) = make_uint4(word[0], word[1], word[2], word[3]);
                  __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int r = 0; r < kMlpRanks; ++r) {
                  rk[slot * 32 + r] =
                    static_cast<unsigned short>(r < KMAX && r < vote.distinct ? vkey[r < KMAX ? r : 0] & 0xFFFFu : mlp::kNone);
                }
                rk[slot * 32 + 31] = static_cast<unsigned short>(row);
              }
              mlp::wave_sync();
              const unsigned short* frow = tile + slot * kWave;
              const uint4 x[2] = {*reinterpret_cast<const uint4*>(frow + 8 * g),
                                  *reinterpret_cast<const uint4*>(frow + 32 + 8 * g)};
              const uint32_t valid = mlp::tile_valid(frow, ln);
              float prob[8];
              mlp::tile_forward(s.mlp_w, x, valid, ln, prob);
              float sc;
              int rank;
              mlp::tile_pick(prob, valid, ln, &sc, &rank);
              if (g == 0 && slot < cnt) {
                const int r = rk[slot * 32 + 31];
                score[r] = sc;
                if (rank >= 0) prop[r] = rk[slot * 32 + rank];
              }
              mlp::wave_sync();  // the tile is reused by the next group
            }
          }
          __syncthreads();
          if (tid == 0) s.scal[kCands] = 0;
          float part = 0.0f;
          for (int i = tid; i < n; i += THREADS) part += score[i];
#pragma unroll
          for (int off = kWave / 2; off >= 1; off >>= 1) part += __shfl_xor(part, off);
          if (lane == 0) s.wave_total[wave] = part;
          __syncthreads();
          double total = 0.0;
          for (int w = 0; w < THREADS / kWave; ++w) total += s.wave_total[w];
          const uint64_t row_key = scal_u64(kRowKeyLo);
          for (int i = tid; i < n; i += THREADS) {
            const double u = unit(mix64(row_key + static_cast<uint64_t>(i)));
            const bool chosen = score[i] > 0.0f && u * total < static_cast<double>(take) * score[i];
            if (!chosen) prop[i] = cur_lab[i];
          }
        } else {
          // Each lane walks the preimages of its positions y = tid, tid +
          // THREADS, ... one network step per pass, taking its next position
          // as soon as a walk lands below n: the wave's passes follow the
          // longest lane's total walk, not the longest walk of every position.
          const Permutation perm(static_cast<uint32_t>(s.scal[kPerm0]), static_cast<uint32_t>(s.scal[kPerm1]),
                                 static_cast<uint32_t>(s.scal[kPerm2]), static_cast<uint32_t>(s.scal[kPerm3]), bits,
                                 n);
          uint32_t y = static_cast<uint32_t>(tid), x = y;
          bool live = static_cast<int>(y) < take;
          while (__ballot(live) != 0ull) {
            if (live) {
              x = perm.network_inverse(x);
              if (x < static_cast<uint32_t>(n)) {
                prop[x] = kChosenMark;
                y += THREADS;
                live = static_cast<int>(y) < take;
                x = y;
              }
            }
          }
        }
        __syncthreads();
        if (!LEARNED && (pending & 1)) settle(pending, tid, first, true);
        // The sizes hold the current state's class counts: only a row whose
        // label changes moves one from its old class to its new one.
        for (int i = tid; i < n; i += THREADS) {
          const bool chosen = prop[i] == kChosenMark;
          const int guide = cur_pred[i];
          const int old = cur_lab[i];
          const int label = learned ? static_cast<int>(prop[i]) : (chosen ? guide : old);
          prop[i] = static_cast<unsigned short>(label);
          if (label != old) {
            atomicAdd(&s.sizes[label], 1);
            atomicSub(&s.sizes[old], 1);
          }
          const int fk = label == guide ? -1 : label * C + guide;
          if (hashed) {
            // Flow counts in the LDS table, right here: open addressing with
            // linear probing from a multiplicative hash of the entry; the row
            // that claims a free slot lists the slot. The table has more
            // slots than the run has rows, so a probe always ends.
            if (fk >= 0) {
              const unsigned mask = (1u << hash_bits) - 1u;
              unsigned h = (static_cast<unsigned>(fk) * 0x9E3779B1u) >> (32 - hash_bits);
              while (true) {
                const unsigned old_key = atomicCAS(&s.hkey[h], kEmptySlot, static_cast<unsigned>(fk));
                if (old_key == kEmptySlot || old_key == static_cast<unsigned>(fk)) {
                  atomicAdd(&s.hcnt[h], 1);
                  if (old_key == kEmptySlot) s.pairs[atomicAdd(&s.scal[kPairs], 1)] = static_cast<int>(h);
                  break;
                }
                h = (h + 1u) & mask;
              }
            }
          } else {
            s.queue[i] = fk;
          }
        }
        if (!hashed) {
          // Flow counts in the workgroup's HBM table: all of a thread's
          // returning atomics of a pass are in flight before the first result
          // is used; the first increment of an entry lists it. One pass
          // covers kRowsPerThread * THREADS rows.
          for (int pass0 = 0; pass0 < n; pass0 += kRowsPerThread * THREADS) {
            int fkey[kRowsPerThread], old[kRowsPerThread];
#pragma unroll
            for (int r = 0; r < kRowsPerThread; ++r) {
              const int i = pass0 + tid + r * THREADS;
              fkey[r] = i < n ? s.queue[i] : -1;
              old[r] = fkey[r] >= 0 ? atomicAdd(&flows[fkey[r]], 1) : 1;
            }
#pragma unroll
            for (int r = 0; r < kRowsPerThread; ++r) {
              if (old[r] == 0) s.pairs[atomicAdd(&s.scal[kPairs], 1)] = fkey[r];
            }
          }
        }
      }
      __syncthreads();

      // -- absorb ---------------------------------------------------------------
      // One thread per class (kClassesPerThread consecutive classes per thread
      // in workgroups of fewer than 1024 threads), so no list of the present
      // classes is built: the present count comes from one ballot per wave.
      {
        // The present count and the sum of squared class sizes (for the
        // objective: at most n^2 < 2^28).
        int here = 0;
        unsigned sq = 0u;
#pragma unroll
        for (int j = 0; j < kClassesPerThread; ++j) {
          const int c = tid * kClassesPerThread + j;
          const int v = c < C ? s.sizes[c] : 0;
          here += __popcll(__ballot(v > 0));
          sq += static_cast<unsigned>(v) * static_cast<unsigned>(v);
        }
        sq = wave_sum_u32(sq);
        if (lane == 0 && here != 0) {
          atomicAdd(&s.scal[kPresent], here);
          atomicAdd(reinterpret_cast<unsigned*>(&s.scal[kSumSq]), sq);
        }
        // Best target per source: atomicMax of a key ordered as the reference's
        // candidate comparison -- surplus (exactly, as cnt*n - size_src*size_dst),
        // then flow, then target size, then the smaller label. Keys are built
        // for every flow entry; they are read only when absorption is eligible.
        const int np = s.scal[kPairs];
        for (int q = tid; q < np; q += THREADS) {
          int idx, cnt;
          if (hashed) {
            // The listed slot: its entry and count, cleared for the next cycle.
            const int slot = s.pairs[q];
            idx = static_cast<int>(s.hkey[slot]);
            cnt = s.hcnt[slot];
            s.hkey[slot] = kEmptySlot;
            s.hcnt[slot] = 0;
          } else {
            idx = s.pairs[q];
            // Read and clear in one L2 atomic: the propose-phase atomics of
            // this launch bypass the vector L1, which may hold an older copy.
            cnt = atomicExch(&flows[idx], 0);
          }
          if (LARGE) s.queue[q] = cnt;  // the second pass below needs it again
          const int src = idx / C;
          const int d = idx - src * C;
          atomicAdd(&s.stay[src], cnt);  // the source's outflow
          const int dn = s.sizes[d];
          if (dn == 0) continue;
          const long long surplus_n = static_cast<long long>(cnt) * n - static_cast<long long>(s.sizes[src]) * dn;
          if (LARGE) {
            // Up to 16383 rows: surplus in 29 bits, flow and target size in 14
            // each; the label is settled in a second pass.
            const unsigned long long k64 = (static_cast<unsigned long long>(surplus_n + (1ll << 28)) << 28) |
                                           (static_cast<unsigned long long>(cnt) << 14) |
                                           static_cast<unsigned long long>(dn);
            atomicMax(&s.best_key[src], k64);
          } else {
            const unsigned long long k64 = (static_cast<unsigned long long>(surplus_n + (1ll << 27)) << 36) |
                                           (static_cast<unsigned long long>(cnt) << 23) |
                                           (static_cast<unsigned long long>(dn) << 10) |
                                           static_cast<unsigned long long>(1023 - d);
            atomicMax(&s.best_key[src], k64);
          }
        }
      }
      __syncthreads();
      const int present = s.scal[kPresent];
      const bool eligible = present > 2;
      if (LARGE && eligible) {
        // Second pass: among the targets of a source's best key, the smallest
        // label (kept in `point`, which is -1 until the consumer).
        const int np = s.scal[kPairs];
        for (int q = tid; q < np; q += THREADS) {
          const int idx = s.pairs[q];
          const int cnt = s.queue[q];
          const int src = idx / C;
          const int d = idx - src * C;
          const int dn = s.sizes[d];
          if (dn == 0) continue;
          const long long surplus_n = static_cast<long long>(cnt) * n - static_cast<long long>(s.sizes[src]) * dn;
          const unsigned long long k64 = (static_cast<unsigned long long>(surplus_n + (1ll << 28)) << 28) |
                                         (static_cast<unsigned long long>(cnt) << 14) |
                                         static_cast<unsigned long long>(dn);
          if (k64 == s.best_key[src]) atomicMin(reinterpret_cast<unsigned int*>(&s.point[src]), static_cast<unsigned int>(d));
        }
        __syncthreads();
      }
      int tgt = -1;
      if (eligible) {
        // Each source with a positive surplus that outflows its stayers points
        // at its best target, which gathers the surplus (times n, exact) and
        // counts its sources.
#pragma unroll
        for (int j = 0; j < kClassesPerThread; ++j) {
          const int src = tid * kClassesPerThread + j;
          if (src >= C) continue;
          const unsigned long long k64 = s.best_key[src];
          if (k64 == 0ull) continue;
          long long surplus_n;
          int cnt, d;
          if (LARGE) {
            surplus_n = static_cast<long long>(k64 >> 28) - (1ll << 28);
            cnt = static_cast<int>((k64 >> 14) & 0x3FFFull);
            d = s.point[src];
            s.point[src] = -1;  // the smallest label of the best key, read once
          } else {
            surplus_n = static_cast<long long>(k64 >> 36) - (1ll << 27);
            cnt = static_cast<int>((k64 >> 23) & 0x1FFFull);
            d = 1023 - static_cast<int>(k64 & 0x3FFull);
          }
          // Stayers: the rows whose prediction is their label.
          if (surplus_n > 0 && cnt > s.sizes[src] - s.stay[src]) {
            s.point[src] = d;
            atomicAdd(&s.weight[d], static_cast<unsigned long long>(surplus_n));
            if (atomicAdd(&s.nsrc[d], 1) == 1) s.scal[kMulti] = 1;  // a second source
            s.scal[kAnyTarget] = 1;
          }
        }
        __syncthreads();
        // Targets with several sources win over single-source ones
        // (core.cpp:1188-1202); one is drawn in ascending label order with
        // probability proportional to its surplus. Each thread holds
        // kClassesPerThread consecutive classes; a wave prefix sum (DPP) and
        // the wave totals place them, and the thread whose range holds the
        // draw names the target. The weights are integers, so every sum is
        // exact and the target does not depend on the summation order. No
        // target at all: nothing to draw.
        if (s.scal[kAnyTarget] != 0) {
          const bool multi = s.scal[kMulti] != 0;
          unsigned long long w[kClassesPerThread];
          unsigned long long part = 0ull;
#pragma unroll
          for (int j = 0; j < kClassesPerThread; ++j) {
            const int c = tid * kClassesPerThread + j;
            const int m = c < C ? s.nsrc[c] : 0;
            w[j] = m > 0 && (!multi || m > 1) ? s.weight[c] : 0ull;
            part += w[j];
          }
          const unsigned long long incl = wave_scan_u64(part);
          unsigned long long* wave_sum = reinterpret_cast<unsigned long long*>(s.wave_total);
          const int wv = tid / kWave;  // from opaque_tid: the address is not kept across the cycle loop
          if ((tid & (kWave - 1)) == kWave - 1) wave_sum[wv] = incl;
          __syncthreads();
          unsigned long long before = 0ull, total = 0ull;
#pragma unroll
          for (int v = 0; v < THREADS / kWave; ++v) {
            const unsigned long long t = wave_sum[v];
            before += v < wv ? t : 0ull;
            total += t;
          }
          if (total == 0ull) {
            if (tid == 0) s.scal[kChosen] = -1;
          } else {
            // floor(u * total) for the 64-bit uniform u / 2^64: below total.
            const unsigned long long r = __umul64hi(scal_u64(kAbsorbLo), total);
            unsigned long long acc = before + incl - part;
            if (r >= acc && r < before + incl) {
              int pick = -1;
#pragma unroll
              for (int j = 0; j < kClassesPerThread; ++j) {
                if (pick < 0 && r < acc + w[j]) pick = tid * kClassesPerThread + j;
                acc += w[j];
              }
              s.scal[kChosen] = pick;
            }
          }
          __syncthreads();
          tgt = s.scal[kChosen];
        }
      }
      // Every row of a batch run is movable, so each source of the chosen
      // target vanishes (nsrc of them); the result must keep two classes or
      // more and lose at least one (core.cpp:1219-1223).
      const int vanished = tgt >= 0 ? s.nsrc[tgt] : 0;
      const bool absorb = vanished > 0 && present - vanished > 1;
      if (absorb) {
        for (int i = tid; i < n; i += THREADS) {
          if (s.point[prop[i]] == tgt) prop[i] = static_cast<unsigned short>(tgt);
        }
        // Rows and squared sizes of the absorbed classes, for the objective.
        unsigned moved = 0u, moved_sq = 0u;
#pragma unroll
        for (int j = 0; j < kClassesPerThread; ++j) {
          const int c = tid * kClassesPerThread + j;
          const unsigned v = c < C && s.point[c] == tgt ? static_cast<unsigned>(s.sizes[c]) : 0u;
          moved += v;
          moved_sq += v * v;
        }
        moved = wave_sum_u32(moved);
        moved_sq = wave_sum_u32(moved_sq);
        if (lane == 0) {
          atomicAdd(reinterpret_cast<unsigned*>(&s.scal[kMoved]), moved);
          atomicAdd(reinterpret_cast<unsigned*>(&s.scal[kMovedSq]), moved_sq);
        }
        if (tid == 0) s.scal[kVanished] = vanished;
      }
      __syncthreads();
    }
    // -- vote -------------------------------------------------------------------
    vote_phase<KMAX, THREADS>(tid, k, pad_total, s, nbr, scores, prop, prop_pred, n, evaluate && group == 1,
                              slice_lo >> 1, (slice_hi + 1) >> 1);
    unsigned ep = 0;
    int par = 0;
    if (group > 1) {
      // Publish this slice's predictions and match count in the step's half
      // of the exchange (halves alternate, so a member a step ahead never
      // overwrites words another still reads), then gather the match totals.
      ep = 1u + static_cast<unsigned>((step + 1) % 65535);
      par = (step + 1) & 1;
      unsigned* xb = p.xbuf + static_cast<size_t>(par) * pad_total + base;
      unsigned* xm = p.xmatch + (static_cast<size_t>(par) * p.runs + run) * kEvoMaxGroup;
      for (int i = slice_lo + tid; i < slice_hi; i += THREADS) {
        __hip_atomic_store(&xb[i], (ep << 16) | prop_pred[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (tid == 0) {
        __hip_atomic_store(&xm[member], (ep << 14) | static_cast<unsigned>(s.scal[kMatches]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        int total = 0;
        for (int m = 0; m < group; ++m) {
          total += m == member ? s.scal[kMatches] : static_cast<int>(await(&xm[m], 14, ep) & 0x3FFFu);
        }
        s.scal[kMatches] = total;
      }
      // The evaluation's class counts, over all rows (the sliced vote
      // counted none).
      if (evaluate) {
        for (int i = tid; i < n; i += THREADS) atomicAdd(&s.counts[prop[i]], 1);
      }
      __syncthreads();
    }

    // -- accept -----------------------------------------------------------------
    // The last wave plans the next cycle meanwhile (this cycle's draws are
    // all read by now).
    if (wave == THREADS / kWave - 1 && lane == 0 && step + 1 < p.cycle_count) plan_cycle(cycle + 1);
    if (wave == 0) {
      double acc;
      const double score = objective(s, guarded, evaluate, C, n, tid, &acc);
      if (lane == 0) {
        int dec = 0;
        if (evaluate) {
          s.state[0] = acc;
          s.state[1] = score;
          s.state[2] = acc;
          s.state[3] = score;
          dec = kEvoImproved;
        } else {
          const double cur_acc = s.state[0], cur_score = s.state[1];
          if (score > s.state[3]) {
            s.state[2] = acc;
            s.state[3] = score;
            dec |= kEvoImproved;
          }
          const double cooling = 1.0 - static_cast<double>(cycle + 1) / static_cast<double>(max(1, p.cycles));
          const double temperature = fmax(1.0e-9, 0.10 * fmax(0.0, 1.0 - cur_acc) * cooling);
          bool take = score >= cur_score;
          if (!take && temperature > 1.0e-9) {
            take = unit(draw(key, cycle, 4)) < accept_exp((score - cur_score) / temperature);
          }
          if (take) {
            s.state[0] = acc;
            s.state[1] = score;
            dec |= kEvoAccepted;
          }
          if (p.history && lead) p.history[static_cast<size_t>(run) * p.cycles + cycle] = s.state[2];
        }
        s.scal[kDecision] = dec;
        if (group > 1) {
          s.scal[kPending] = (evaluate || (dec & kEvoAccepted) ? 1 : 0) | (dec & kEvoImproved ? 2 : 0);
          s.scal[kPendingEp] = static_cast<int>(ep);
          s.scal[kPendingPar] = par;
        }
      }
    }
    __syncthreads();
    const int dec = s.scal[kDecision];
    if (group == 1 && (dec & kEvoImproved) && lead) {
      for (int i = tid; i < n; i += THREADS) {
        p.best[base + i] = prop[i];
        p.bestpred[base + i] = prop_pred[i];
      }
    }
    // An accepted proposal's class counts become the current state's: its
    // sizes, with the absorbed classes moved into the target. The next
    // proposal's sizes start from the current counts.
    {
      const bool took = !evaluate && (dec & kEvoAccepted) != 0;
      const int vanished = s.scal[kVanished], tgt = s.scal[kChosen], moved = s.scal[kMoved];
      for (int c = tid; c < C; c += THREADS) {
        if (took) {
          const bool gone = vanished > 0 && s.point[c] == tgt;
          s.counts[c] = gone ? 0 : s.sizes[c] + (vanished > 0 && c == tgt ? moved : 0);
        }
        s.sizes[c] = s.counts[c];
        s.stay[c] = 0;
        s.nsrc[c] = 0;
        s.point[c] = -1;
        s.weight[c] = 0ull;
        s.best_key[c] = 0ull;
      }
    }
    __syncthreads();
    if (tid == 0) {
      if (dec & kEvoAccepted) s.scal[kCur] = cur ^ 1;
      s.scal[kMatches] = 0;
      s.scal[kQueue] = 0;
      s.scal[kDecision] = 0;
      s.scal[kPairs] = 0;
      s.scal[kPresent] = 0;
      s.scal[kSumSq] = 0;
      s.scal[kMulti] = 0;
      s.scal[kVanished] = 0;
      s.scal[kMoved] = 0;
      s.scal[kMovedSq] = 0;
      s.scal[kAnyTarget] = 0;
    }
    __syncthreads();
  }

  if (threadIdx.x == 0 && p.wg_time != nullptr) p.wg_time[wg] += wall_clock64();
  // Write the run state back for the next launch.
  if (!lead) return;
  if (s.scal[kPending] & 1) {
    unsigned first[kAhead];
    settle(s.scal[kPending], static_cast<int>(threadIdx.x), first, false);
  }
  const int cur = s.scal[kCur];
  for (int i = threadIdx.x; i < n; i += THREADS) {
    p.cur[base + i] = s.lab(cur)[i];
    p.curpred[base + i] = s.pred(cur)[i];
  }
  if (threadIdx.x < 4) p.state[static_cast<size_t>(run) * kEvoStateStride + threadIdx.x] = s.state[threadIdx.x];
}

// Padded rows come in blocks of 256 that each belong to one run (block_run).
// A missing neighbour becomes local id n, the run's sentinel slot.
__global__ void evo_narrow_kernel(const int* __restrict__ nbr32, const int* __restrict__ run_offsets,
                                  const int* __restrict__ pad_offsets, const int* __restrict__ block_run,
                                  unsigned short* __restrict__ nbr16, size_t total, int pad_total) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int v = nbr32[i];
  const int run = block_run[static_cast<int>(i % static_cast<size_t>(pad_total)) >> 8];
  const int local = v < 0 ? run_offsets[run + 1] - run_offsets[run] : v - pad_offsets[run];
  nbr16[i] = static_cast<unsigned short>(local);
}

__global__ void evo_zero_kernel(int* a, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// ---- launchers --------------------------------------------------------------

int evo_hash_bits(int max_rows) {
  int bits = 1;
  while ((1 << bits) <= max_rows) ++bits;
  return bits;
}

size_t evo_lds_bytes(int classes, int max_rows, bool learned, bool large, int hash_bits) {
  const size_t rows_lds = large ? 0 : 2 * static_cast<size_t>(max_rows);
  const size_t hash_lds = !large && hash_bits > 0 ? 2 * sizeof(int) * (static_cast<size_t>(1) << hash_bits) : 0;
  const size_t base = sizeof(double) * (kWaves + 4) + sizeof(unsigned long long) * 2 * static_cast<size_t>(classes) +
                      sizeof(int) * (5 * static_cast<size_t>(classes) + rows_lds + kScalarSlots) +
                      sizeof(unsigned short) * 4 * static_cast<size_t>(label_slots(max_rows)) + hash_lds;
  if (!learned) return base;
  return base + 16 + sizeof(uint32_t) * kMlpTiled +
         sizeof(unsigned short) * (kLearnedWaves * mlp::kTileRows * (kWave + 32) + 66);
}

// 1024 threads per workgroup (16 waves hide the gather latency) up to 32
// neighbours; 512 for 64, whose id and key arrays need 256 registers; 768 for
// the learned variant up to 32 (its matrix-core tile pass needs about 140
// registers, over the 128 of 16 waves).
template <int KMAX, bool LEARNED, bool LARGE>
static hipError_t launch_persistent_k(const EvoParams& p, size_t shmem, hipStream_t s) {
  constexpr int threads = KMAX > 32 ? kThreads / 2 : (LEARNED ? kLearnedThreads : kThreads);
  const void* fn = reinterpret_cast<const void*>(&evo_persistent_kernel<KMAX, threads, LEARNED, LARGE>);
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem));
  if (e != hipSuccess) return e;
  if (p.wg_run != nullptr && p.cooperative) {
    // Split runs wait on each other: all workgroups must be resident at once,
    // which a cooperative launch guarantees (or refuses).
    EvoParams arg = p;
    void* args[] = {&arg};
    return hipLaunchCooperativeKernel(fn, dim3(p.workgroups), dim3(threads), args, static_cast<unsigned>(shmem), s);
  }
  // One workgroup per run, or split runs launched as an ordinary grid (sized
  // to what the device holds at once; the bounded waits turn a member that is
  // not resident into an error, not a hang).
  hipLaunchKernelGGL((evo_persistent_kernel<KMAX, threads, LEARNED, LARGE>),
                     dim3(p.wg_run != nullptr ? p.workgroups : p.runs), dim3(threads), shmem, s, p);
  return hipGetLastError();
}

template <bool LEARNED, bool LARGE>
static hipError_t launch_persistent(const EvoParams& p, size_t shmem, hipStream_t s) {
  if (p.k <= 8) return launch_persistent_k<8, LEARNED, LARGE>(p, shmem, s);
  if (p.k <= 16) return launch_persistent_k<16, LEARNED, LARGE>(p, shmem, s);
  if (p.k <= 32) return launch_persistent_k<32, LEARNED, LARGE>(p, shmem, s);
  return launch_persistent_k<64, LEARNED, LARGE>(p, shmem, s);
}

template <int KMAX, bool LEARNED, bool LARGE>
static int resident_per_cu_k(size_t shmem) {
  constexpr int threads = KMAX > 32 ? kThreads / 2 : (LEARNED ? kLearnedThreads : kThreads);
  const void* fn = reinterpret_cast<const void*>(&evo_persistent_kernel<KMAX, threads, LEARNED, LARGE>);
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem)) != hipSuccess) return 0;
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, threads, shmem) != hipSuccess) return 0;
  return blocks;
}

template <bool LEARNED, bool LARGE>
static int resident_per_cu(int k, size_t shmem) {
  if (k <= 8) return resident_per_cu_k<8, LEARNED, LARGE>(shmem);
  if (k <= 16) return resident_per_cu_k<16, LEARNED, LARGE>(shmem);
  if (k <= 32) return resident_per_cu_k<32, LEARNED, LARGE>(shmem);
  return resident_per_cu_k<64, LEARNED, LARGE>(shmem);
}

int evo_resident_per_cu(int k, int classes, int max_rows, bool learned, bool large, int hash_bits) {
  const size_t shmem = evo_lds_bytes(classes, max_rows, learned, large, hash_bits);
  if (large) return learned ? resident_per_cu<true, true>(k, shmem) : resident_per_cu<false, true>(k, shmem);
  return learned ? resident_per_cu<true, false>(k, shmem) : resident_per_cu<false, false>(k, shmem);
}

hipError_t launch_evo_persistent(const EvoParams& p, hipStream_t s) {
  const bool learned = p.learn_packed != nullptr;
  const bool large = p.large_scratch != nullptr;
  const size_t shmem = evo_lds_bytes(p.classes, p.max_rows, learned, large, p.hash_bits);
  if (shmem > kEvoMaxLds || p.max_rows > (large ? kEvoMaxRowsLarge : kEvoMaxRows) || p.classes > 1024 ||
      p.hash_bits < 0 || (p.hash_bits > 0 && (large || (1 << p.hash_bits) <= p.max_rows || p.hash_bits > 15)) ||
      (p.hash_bits == 0 && p.flows == nullptr) ||
      p.k < 1 || p.k > 64 || p.pad_total % 256 != 0 || static_cast<long long>(p.k) * p.pad_total >= (1ll << 31) ||
      (p.wg_run != nullptr && (p.wg_member == nullptr || p.run_group == nullptr || p.xbuf == nullptr ||
                               p.xmatch == nullptr || p.error == nullptr || p.workgroups < p.runs ||
                               p.spin_limit < 1))) {
    return hipErrorInvalidValue;
  }
  if (large) return learned ? launch_persistent<true, true>(p, shmem, s) : launch_persistent<false, true>(p, shmem, s);
  return learned ? launch_persistent<true, false>(p, shmem, s) : launch_persistent<false, false>(p, shmem, s);
}

hipError_t launch_evo_narrow(const int* nbr32, const int* run_offsets, const int* pad_offsets, const int* block_run,
                             unsigned short* nbr16, size_t total, int pad_total, hipStream_t s) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(evo_narrow_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s, nbr32,
                     run_offsets, pad_offsets, block_run, nbr16, total, pad_total);
  return hipGetLastError();
}

hipError_t launch_evo_zero(int* a, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(evo_zero_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, a, n);
  return hipGetLastError();
}

}  // namespace kodama_hip
