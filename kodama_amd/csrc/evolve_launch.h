// Parameter block and launchers of the device-resident evolution kernels
// (evolve_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace kodama_hip {

constexpr int kEvoAccepted = 1;
constexpr int kEvoImproved = 2;
constexpr int kEvoAbsorbed = 4;
constexpr int kEvoStateStride = 4;  // cur_acc, cur_score, best_acc, best_score

struct EvoParams {
  int runs = 0;
  int classes = 0;       // label codes per run (array stride)
  int cycles = 0;
  int guarded = 1;       // diversity-guarded objective
  const int* block_run = nullptr;
  const int* run_offsets = nullptr;
  const int* pad_offsets = nullptr;
  int* cycle = nullptr;              // device cycle counter
  const uint64_t* run_key = nullptr; // [runs]
  const int* perm_half = nullptr;    // [runs] Feistel half width
  int* cur = nullptr;
  int* curpred = nullptr;
  int* best = nullptr;
  int* bestpred = nullptr;
  int* prop = nullptr;
  int* proppred = nullptr;
  int* sizes = nullptr;              // [runs x classes]
  int* flows = nullptr;              // [runs x classes x classes]
  int* remap = nullptr;              // [runs x classes]
  int* counts = nullptr;             // vote reductions [runs x classes]
  int* firsts = nullptr;
  int* matches = nullptr;            // [runs]
  double* state = nullptr;           // [runs x kEvoStateStride]
  int* decision = nullptr;           // [runs]
  double* history = nullptr;         // [runs x cycles] best accuracy per cycle
};

hipError_t launch_evo_clear(const EvoParams& p, hipStream_t s);
hipError_t launch_evo_propose(const EvoParams& p, int blocks, hipStream_t s);
hipError_t launch_evo_absorb(const EvoParams& p, hipStream_t s);
hipError_t launch_evo_remap(const EvoParams& p, int blocks, hipStream_t s);
hipError_t launch_evo_accept(const EvoParams& p, int init, hipStream_t s);
hipError_t launch_evo_commit(const EvoParams& p, int blocks, hipStream_t s);

}  // namespace kodama_hip
