// Shared helpers of the C ABI: exception -> status code mapping with a
// thread-local message, and device selection with an architecture check.
#pragma once

#include <exception>
#include <stdexcept>
#include <string>

#include "../../include/kodama_hip.h"
#include "device.h"

namespace kodama_hip {

struct NoDevice : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void set_last_error(const std::string& msg);

// Makes `device` current after checking that it exists and is a gfx950 part.
void select_device(int device);

// Largest squared row norm (double accumulation, rounded up to float), the
// scale of the MFMA candidate error bound.
float max_row_norm2(const float* x, int rows, int p);

// Queries of the calling thread's last exact kNN search that fell back to the
// register scan (-1: the MFMA path was not used).
int& last_knn_fallbacks();

template <class F>
int guarded(F&& f) {
  try {
    f();
    return KODAMA_HIP_OK;
  } catch (const NoDevice& e) {
    set_last_error(e.what());
    return KODAMA_HIP_ENODEV;
  } catch (const std::invalid_argument& e) {
    set_last_error(e.what());
    return KODAMA_HIP_EINVAL;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return KODAMA_HIP_ERUNTIME;
  } catch (...) {
    set_last_error("unknown exception");
    return KODAMA_HIP_ERUNTIME;
  }
}

}  // namespace kodama_hip
