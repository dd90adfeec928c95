// Device-resident label evolution of a batch of runs (evolve_device.cpp).
#pragma once

#include <memory>
#include <vector>

#include "engine.h"

namespace kodama_hip {

class DeviceEvolution {
 public:
  // Label-independent set-up of a batch's evolution: the state arena, the
  // 16-bit neighbour table, the cleared flow table and, with `params` (host
  // [kMlpParams], else null), the learned proposal's network in its tiled
  // form. Nothing here depends on the starting labels.
  DeviceEvolution(KnnBatch& batch, const std::vector<IVec>& values, const kodama_hip_core_options& options,
                  const float* params = nullptr);
  ~DeviceEvolution();
  DeviceEvolution(const DeviceEvolution&) = delete;
  DeviceEvolution& operator=(const DeviceEvolution&) = delete;
  // Uploads the starting label codes of every run (same order as `values`);
  // returns once they are on the device.
  void start(const std::vector<IVec>& codes);
  // Enqueues `cycles` more cycles of every run in one persistent launch
  // (the first launch evaluates the starting labels first); with the network
  // of the constructor the learned variant, whose cycles fall back to the
  // uniform proposal with probability `fallback`. Does not block.
  void run(int cycles);
  void run_learned(int cycles, float fallback);
  // Blocks until the enqueued launches are done (results in device memory).
  void wait();
  // Downloads and converts the results; blocks until the launches are done.
  void export_results(int* clbest, double* accbest, double* vect_acc);
  // Time from the start of the first launch to the end of the last (HIP
  // events on the batch stream); valid after export_results.
  double kernel_seconds() const;
  // Workgroups of the launches (the runs, or more when large runs are split).
  int workgroups() const;
  // Whether the launches count flows in the LDS table.
  bool lds_flow_table() const;

  struct Impl;

 private:
  Impl* impl_ = nullptr;
};

// Largest number of workgroups per run in later evolutions: 0 = automatic
// (up to kEvoMaxGroup while the device holds them all at once), 1 = one
// workgroup per run.
void set_evolution_split(int max_group);
int evolution_split();
// Split runs take a cooperative launch (default) or an ordinary one (for
// tools that cannot follow cooperative launches).
void set_cooperative_launch(bool on);
// Later evolutions count flows in an LDS table when it fits beside the run
// state (default) or always in the per-workgroup HBM table.
void set_lds_flow_table(bool on);

}  // namespace kodama_hip
