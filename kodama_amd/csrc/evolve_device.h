// Device-resident label evolution of a batch of runs (declared here, defined in
// evolve_device.cpp).
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
  void run(int cycles);
  double vote_seconds() const { return vote_seconds_; }
  void export_results(int* clbest, double* accbest, double* vect_acc);

 private:
  struct Impl;
  Impl* impl_ = nullptr;
  double vote_seconds_ = 0.0;
};

}  // namespace kodama_hip
