// Graph-side kernels of the KODAMA matrix: publication of an exact self-kNN
// graph (graph_cluster.cpp:790-837) and the KODAMA dissimilarity rewrite of a
// graph from the run labels (kodama_matrix.cpp:2685-2776).
//
// Dissimilarity layout: the run labels res [runs x samples] are transposed once
// to sample-major rows of `stride` int32 (runs rounded up to a multiple of 4,
// padded with 0 = unassigned, which never counts as valid). One wave owns one
// graph row: per edge each lane loads 4 of the neighbour's labels (one 16-byte
// load, the whole row in one coalesced instruction for up to 256 runs) and the
// co-assignment counts come from ballots. The row is then ordered stably by
// (rewritten distance, output id) by rank counting in LDS.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "graph_launch.h"

// Scratch of the allocation-free dissimilarity entry point: the transposed run
// labels [samples x stride] int32 (stride = runs rounded up to 4), 64 MiB per
// GPU, held until the process ends. Host code never names it directly: the
// per-device address comes from noalloc_scratch below.
static __device__ __align__(256) int scratch[kodama_hip::kNoallocScratchBytes / sizeof(int)];

namespace kodama_hip {

namespace {

constexpr int kWave = 64;
constexpr int kRowsPerBlock = 4;  // waves (graph rows) per workgroup

// Float -> unsigned key with the same order (negative values reversed).
__device__ __forceinline__ uint32_t ordered_bits(float v) {
  const uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

}  // namespace

// Internal distances of a self-kNN search -> 1-based ids and public distances
// (native_knn.cpp:1157-1164): sqrt(max(0, d)) for Euclidean, max(0, 1 + d) for
// cosine, 1 + d for the inner product; missing entries stay -1 / +inf.
__global__ void graph_publish_kernel(int* __restrict__ idx, float* __restrict__ dist, size_t total, int metric) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int id = idx[i];
  if (id < 0) {
    dist[i] = INFINITY;
    return;
  }
  idx[i] = id + 1;
  const float d = dist[i];
  dist[i] = metric == 2 ? sqrtf(fmaxf(0.0f, d)) : (metric == 0 ? fmaxf(0.0f, 1.0f + d) : 1.0f + d);
}

// res [runs x samples] -> labels [samples x stride], zero padded; 32 x 32 tiles.
__global__ void labels_transpose_kernel(const int* __restrict__ res, int runs, int samples, int stride,
                                        int* __restrict__ out) {
  __shared__ int tile[32][33];
  const int s0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  for (int dy = threadIdx.y; dy < 32; dy += blockDim.y) {
    const int r = r0 + dy, s = s0 + threadIdx.x;
    tile[dy][threadIdx.x] = r < runs && s < samples ? res[static_cast<size_t>(r) * samples + s] : 0;
  }
  __syncthreads();
  for (int dy = threadIdx.y; dy < 32; dy += blockDim.y) {
    const int s = s0 + dy, r = r0 + threadIdx.x;
    if (s < samples && r < stride) out[static_cast<size_t>(s) * stride + r] = tile[threadIdx.x][dy];
  }
}

// One wave per graph row. `keys`/`ids` live in dynamic LDS: kRowsPerBlock x
// neighbors entries each.
__global__ void __launch_bounds__(kWave * kRowsPerBlock)
dissimilarity_kernel(int* __restrict__ indices, float* __restrict__ distances, int samples, int neighbors,
                     const int* __restrict__ labels, int stride, int input_one_based, int output_one_based) {
  extern __shared__ unsigned long long dis_lds[];
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int row = blockIdx.x * kRowsPerBlock + wave;
  const bool live = row < samples;  // every wave reaches the barrier below
  unsigned long long* keys = dis_lds + static_cast<size_t>(wave) * neighbors;
  int* ids = reinterpret_cast<int*>(dis_lds + static_cast<size_t>(kRowsPerBlock) * neighbors) +
             static_cast<size_t>(wave) * neighbors;
  const int quads = stride >> 2;
  const int4* lab4 = reinterpret_cast<const int4*>(labels);
  const size_t base = static_cast<size_t>(row) * neighbors;

  // Co-assignment counts: 4 runs per lane per chunk of 256 runs.
  for (int j = 0; live && j < neighbors; ++j) {
    const int stored = indices[base + j];
    const int nb = stored >= 0 && input_one_based ? stored - 1 : stored;
    float d = distances[base + j];
    int out_id = stored;
    if (nb >= 0 && nb < samples && isfinite(d)) {
      int same = 0, valid = 0;
      for (int q0 = 0; q0 < quads; q0 += kWave) {
        const int q = q0 + lane;
        int4 a = make_int4(0, 0, 0, 0), b = make_int4(0, 0, 0, 0);
        if (q < quads) {
          a = lab4[static_cast<size_t>(row) * quads + q];
          b = lab4[static_cast<size_t>(nb) * quads + q];
        }
        const bool v0 = a.x != 0 && b.x != 0, v1 = a.y != 0 && b.y != 0;
        const bool v2 = a.z != 0 && b.z != 0, v3 = a.w != 0 && b.w != 0;
        valid += __popcll(__ballot(v0)) + __popcll(__ballot(v1)) + __popcll(__ballot(v2)) + __popcll(__ballot(v3));
        same += __popcll(__ballot(v0 && a.x == b.x)) + __popcll(__ballot(v1 && a.y == b.y)) +
                __popcll(__ballot(v2 && a.z == b.z)) + __popcll(__ballot(v3 && a.w == b.w));
      }
      if (same == 0 || valid == 0) {
        d = INFINITY;
      } else {
        const double agreement = static_cast<double>(same) / static_cast<double>(valid);
        d = static_cast<float>((1.0 + static_cast<double>(d)) / (agreement * agreement));
      }
      out_id = output_one_based ? nb + 1 : nb;
    } else {
      d = INFINITY;
    }
    if (lane == 0) {
      keys[j] = (static_cast<unsigned long long>(ordered_bits(d)) << 32) |
                static_cast<unsigned long long>(static_cast<uint32_t>(out_id) ^ 0x80000000u);
      ids[j] = out_id;
    }
  }
  __syncthreads();  // orders lane 0's LDS writes before the other lanes' reads
  // Stable order by (distance, id): rank = smaller keys + equal keys earlier.
  for (int e = lane; live && e < neighbors; e += kWave) {
    const unsigned long long ke = keys[e];
    int rank = 0;
    for (int f = 0; f < neighbors; ++f) {
      const unsigned long long kf = keys[f];
      rank += (kf < ke || (kf == ke && f < e)) ? 1 : 0;
    }
    const uint32_t bits = static_cast<uint32_t>(ke >> 32);
    const uint32_t fb = (bits & 0x80000000u) ? (bits & 0x7FFFFFFFu) : ~bits;
    distances[base + rank] = __uint_as_float(fb);
    indices[base + rank] = ids[e];
  }
}

hipError_t launch_graph_publish(int* idx, float* dist, size_t total, int metric, hipStream_t s) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(graph_publish_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s, idx,
                     dist, total, metric);
  return hipGetLastError();
}

hipError_t launch_labels_transpose(const int* res, int runs, int samples, int stride, int* out, hipStream_t s) {
  if (samples == 0 || stride == 0) return hipSuccess;
  const dim3 grid(static_cast<unsigned>((samples + 31) / 32), static_cast<unsigned>((stride + 31) / 32));
  hipLaunchKernelGGL(labels_transpose_kernel, grid, dim3(32, 8), 0, s, res, runs, samples, stride, out);
  return hipGetLastError();
}

hipError_t noalloc_scratch(int device, int** ptr, size_t* bytes) {
  // Resolved once per device, with that device current: every GPU holds its
  // own copy of `scratch` at its own address.
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static int* addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (addr[device] == nullptr) {
    int current = -1;
    hipError_t e = hipGetDevice(&current);
    if (e != hipSuccess) return e;
    if (current != device) return hipErrorInvalidDevice;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = static_cast<int*>(p);
  }
  *ptr = addr[device];
  *bytes = kNoallocScratchBytes;
  return hipSuccess;
}

size_t dissimilarity_lds_bytes(int neighbors) {
  return static_cast<size_t>(kRowsPerBlock) * neighbors * (sizeof(unsigned long long) + sizeof(int));
}

hipError_t launch_dissimilarity(int* indices, float* distances, int samples, int neighbors, const int* labels,
                                int stride, bool input_one_based, bool output_one_based, hipStream_t s) {
  if (samples == 0 || neighbors == 0) return hipSuccess;
  if (stride % 4 != 0 || dissimilarity_lds_bytes(neighbors) > 64 * 1024) return hipErrorInvalidValue;
  const unsigned blocks = static_cast<unsigned>((samples + kRowsPerBlock - 1) / kRowsPerBlock);
  hipLaunchKernelGGL(dissimilarity_kernel, dim3(blocks), dim3(kWave * kRowsPerBlock),
                     dissimilarity_lds_bytes(neighbors), s, indices, distances, samples, neighbors, labels, stride,
                     input_one_based ? 1 : 0, output_one_based ? 1 : 0);
  return hipGetLastError();
}

}  // namespace kodama_hip
