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
