// Parameter block and launchers of the persistent device-resident evolution
// kernel (evolve_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace kodama_hip {

constexpr int kEvoThreads = 1024;           // one workgroup per run (512 for k > 32)
constexpr size_t kEvoMaxLds = 160 * 1024;   // LDS of one CDNA4 compute unit
constexpr int kEvoAccepted = 1;
constexpr int kEvoImproved = 2;
constexpr int kEvoStateStride = 4;  // cur_acc, cur_score, best_acc, best_score
constexpr int kEvoMaxRows = 8191;   // rows of one run (13-bit fields of the absorption keys)
// Runs of up to kEvoMaxRowsLarge rows take the large-run variant: the row
// queue and the distinct flow list live in a per-run global scratch instead
// of LDS, and the absorption keys are compared in two passes (surplus, flow,
// target size in 57 bits, then the smallest label).
constexpr int kEvoMaxRowsLarge = 16383;

struct EvoParams {
  int runs = 0;
  int classes = 0;        // label codes per run (array stride), <= 1024
  int cycles = 0;         // total cycles of the evolution (annealing, history stride)
  int cycle_begin = 0;    // first cycle of this launch
  int cycle_count = 0;    // cycles run by this launch
  int init = 0;           // evaluate the starting labels first
  int guarded = 1;        // diversity-guarded objective
  int k = 0;
  int max_rows = 0;       // largest run, <= kEvoMaxRows (kEvoMaxRowsLarge with large_scratch)
  int pad_total = 0;
  const int* run_offsets = nullptr;          // [runs + 1]
  const int* pad_offsets = nullptr;          // [runs + 1]
  const unsigned short* neighbors = nullptr; // [k x pad_total] local ids, run rows = none
  const float* scores = nullptr;             // [k x pad_total]
  const uint64_t* run_key = nullptr;         // [runs]
  const int* perm_bits = nullptr;            // [runs] ceil(log2 rows), at least 1
  unsigned short* cur = nullptr;             // [pad_total] run state between launches
  unsigned short* curpred = nullptr;
  unsigned short* best = nullptr;
  unsigned short* bestpred = nullptr;
  double* state = nullptr;                   // [runs x kEvoStateStride]
  int* flows = nullptr;                      // [workgroups x classes x classes], zero between cycles
  // LDS flow table instead of `flows` (0: `flows`): 2^hash_bits slots, more
  // than max_rows, so a cycle's distinct flow entries (one per moved row at
  // most) never fill it; each slot holds an entry and its count, both cleared
  // when the absorption reads them. LDS-resident layout only.
  int hash_bits = 0;
  double* history = nullptr;                 // [runs x cycles] best accuracy per cycle
  // [workgroups] elapsed wall-clock ticks (100 MHz) summed over launches, for
  // the member assignment of later evolutions of the same batch shape; may
  // be null.
  unsigned long long* wg_time = nullptr;
  // Learned proposal (null: every cycle uniform): the network in the tiled
  // layout of launch_mlp_pack, and the probability that a cycle falls back to
  // the uniform proposal.
  const uint32_t* learn_packed = nullptr;
  float learn_fallback = 0.0f;
  // Large-run variant: [workgroups x 2 x max_rows] ints (distinct flow
  // entries, then the row queue of each workgroup); null for the
  // LDS-resident layout.
  int* large_scratch = nullptr;
  // Split runs (null wg_run: one workgroup per run, workgroup r = run r).
  // Workgroup b works on run wg_run[b] as member wg_member[b] of
  // run_group[run] members. Every member replays the same proposal,
  // absorption and acceptance (identical draws and integer decisions) in its
  // own LDS and votes only its slice of rows; slices are exchanged through
  // xbuf ([2 x pad_total] words, (epoch << 16) | prediction) and the match
  // counts through xmatch ([2 x runs x kEvoMaxGroup], (epoch << 14) |
  // matches), written and read with device-scope relaxed atomics; a reader
  // waits (bounded by spin_limit polls) until a word carries the expected
  // epoch. Steps alternate between the two halves. The match totals are
  // gathered before the acceptance test; the other slices' predictions only
  // when the proposal became the current state, after the next step's
  // marking. A member can overwrite a half only after every member has
  // published the next step's matches, which each does after reading the
  // half. Members need to be resident together: the launch is cooperative.
  // The flow tables and the large-run scratch are per workgroup. *error is
  // set when a wait ran out.
  const int* wg_run = nullptr;
  const int* wg_member = nullptr;
  const int* run_group = nullptr;
  unsigned* xbuf = nullptr;
  unsigned* xmatch = nullptr;
  int* error = nullptr;
  int spin_limit = 0;
  int workgroups = 0;                        // grid size (runs when not split)
  int cooperative = 1;                       // split runs: cooperative launch (else an ordinary one)
};
constexpr int kEvoMaxGroup = 8;  // workgroups per run at most

// LDS of one run-workgroup; `large` leaves the row queue and the flow list
// out; hash_bits > 0 adds the LDS flow table of 2^hash_bits slots.
size_t evo_lds_bytes(int classes, int max_rows, bool learned = false, bool large = false, int hash_bits = 0);
// Bits of the LDS flow table for runs of up to max_rows rows: the smallest
// table with more slots than max_rows.
int evo_hash_bits(int max_rows);
hipError_t launch_evo_persistent(const EvoParams& p, hipStream_t s);
// Workgroups of the persistent kernel's variant for these sizes that one CU
// holds at once (0 when the query fails).
int evo_resident_per_cu(int k, int classes, int max_rows, bool learned, bool large, int hash_bits);
// 32-bit padded neighbour ids of the vote layout -> 16-bit run-local ids; a
// missing neighbour (-1) becomes the run's row count, its sentinel slot.
hipError_t launch_evo_narrow(const int* nbr32, const int* run_offsets, const int* pad_offsets, const int* block_run,
                             unsigned short* nbr16, size_t total, int pad_total, hipStream_t s);
hipError_t launch_evo_zero(int* a, size_t n, hipStream_t s);

}  // namespace kodama_hip
