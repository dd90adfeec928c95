#include "evolve_device.h"

namespace kodama_hip {

struct DeviceEvolution::Impl {};

DeviceEvolution::DeviceEvolution(KnnBatch&, const std::vector<IVec>&, const std::vector<IVec>&,
                                 const std::vector<IVec>&, const kodama_hip_core_options&) {
  throw std::runtime_error("device-resident evolution is not available in this build.");
}
DeviceEvolution::~DeviceEvolution() { delete impl_; }
void DeviceEvolution::run(int) {}
void DeviceEvolution::export_results(int*, double*, double*) {}

}  // namespace kodama_hip
