// Library-level entry points of the C ABI and the helpers of capi_common.h.
#include "capi_common.h"

#include <cstring>
#include <mutex>
#include <vector>

namespace kodama_hip {

namespace {
thread_local std::string g_last_error;
thread_local int g_knn_fallbacks = -1;
std::mutex g_arch_mutex;
std::vector<int> g_arch_ok;  // -1 unknown, 0 not gfx950, 1 gfx950
}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }

int& last_knn_fallbacks() { return g_knn_fallbacks; }

float max_row_norm2(const float* x, int rows, int p) {
  double best = 0.0;
  for (int r = 0; r < rows; ++r) {
    double s = 0.0;
    for (int j = 0; j < p; ++j) {
      const double v = x[static_cast<size_t>(r) * p + j];
      s += v * v;
    }
    best = s > best ? s : best;
  }
  return static_cast<float>(best * (1.0 + 1e-6));
}

void select_device(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    throw NoDevice("no HIP device is visible; the HIP backend requires an MI355X (gfx950).");
  }
  if (device < 0 || device >= count) throw NoDevice("gpu_device is outside the visible HIP devices.");
  {
    std::lock_guard<std::mutex> lock(g_arch_mutex);
    if (g_arch_ok.size() != static_cast<size_t>(count)) g_arch_ok.assign(static_cast<size_t>(count), -1);
    if (g_arch_ok[device] < 0) {
      hipDeviceProp_t prop;
      check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
      g_arch_ok[device] = std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
    }
    if (g_arch_ok[device] != 1) throw NoDevice("the HIP backend is built for gfx950 (MI355X) only.");
  }
  check(hipSetDevice(device), "hipSetDevice");
}

}  // namespace kodama_hip

extern "C" {

const char* kodama_hip_last_error(void) { return kodama_hip::g_last_error.c_str(); }

const char* kodama_hip_version(void) { return "kodama_hip 0.1.0 gfx950"; }

int kodama_hip_last_knn_fallbacks(void) { return kodama_hip::g_knn_fallbacks; }

int kodama_hip_device_count(int32_t* count) {
  return kodama_hip::guarded([&] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    int ok = 0;
    for (int d = 0; d < n; ++d) {
      hipDeviceProp_t prop;
      if (hipGetDeviceProperties(&prop, d) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) ++ok;
    }
    *count = ok;
  });
}

}  // extern "C"
