// Device-resident label evolution of a batch of runs (evolve_device.cpp).
#pragma once

#include <memory>
#include <vector>

#include "engine.h"

namespace kodama_hip {

class DeviceEvolution {
 public:
  DeviceEvolution(KnnBatch& batch, const std::vector<IVec>& values, const std::vector<IVec>& codes,
                  const std::vector<IVec>& folds, const kodama_hip_core_options& options);
  ~DeviceEvolution();
  DeviceEvolution(const DeviceEvolution&) = delete;
  DeviceEvolution& operator=(const DeviceEvolution&) = delete;
  // Runs `cycles` more cycles of every run in one persistent launch; blocks.
  void run(int cycles);
  // Runs `cycles` more cycles with the learned proposal (the persistent
  // kernel's learned variant, one launch): in each learned cycle the network
  // (host params [kMlpParams], copied into the arena) scores the rows whose
  // label differs from their prediction; a cycle falls back to the uniform
  // proposal with probability `fallback`. Blocks.
  void run_learned(int cycles, const float* params, float fallback);
  // Duration of the last launch (HIP events on the batch stream).
  double kernel_seconds() const { return kernel_seconds_; }
  void export_results(int* clbest, double* accbest, double* vect_acc);

 private:
  struct Impl;
  Impl* impl_ = nullptr;
  double kernel_seconds_ = 0.0;
};

}  // namespace kodama_hip
