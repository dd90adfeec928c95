// KODAMA.matrix orchestration kernels for gfx950.
//
// k-means (kodama_matrix.cpp:664-799). The reference's portable x86-64 build
// computes every dot product in SSE registers without FMA
// (kodama_matrix.cpp:91-124): 16 running partials over blocks of 16 dimensions
// (each product rounded, then added), the lane sums (s[l] + s[4+l]) +
// (s[8+l] + s[12+l]), then ((l0 + l1) + l2) + l3, then the remaining
// dimensions one product at a time. The kernels spell every operation with
// __fmul_rn / __fadd_rn so nothing is contracted into an FMA, and the
// assignments, centroid sums and norms come out bit-identical to that build.
//
// kmeans_assign_kernel: one thread per row (256 per workgroup) with the row in
// registers; centroids stream through LDS in tiles of 32 KiB, in ascending
// cluster order, so the first cluster wins ties as in the reference loop.
// kmeans_update_kernel: one workgroup per cluster, one thread per feature; the
// members are found 1024 rows at a time with ballots and added in ascending
// row order, the order of the reference's sequential sum.
//
// project_kernel: one thread per (row, run); landmark lookups through a
// per-run map from data row to landmark position (filled by
// local_fill_kernel and local_scatter_kernel); the votes of up to 64 landmark
// neighbours sit in the thread's LDS slice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "matrix_launch.h"

namespace kodama_hip {

namespace {

constexpr int kBlock = 256;
constexpr int kProjectBlock = 128;
constexpr int kMaxVotes = 64;

// Reference-order dot of a register row and an LDS / global row (p <= PMAX).
template <int PMAX>
__device__ __forceinline__ float sse_dot(const float (&q)[PMAX], const float* __restrict__ c, int p) {
  float s[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) s[j] = 0.0f;
  const int nb = p >> 4;
#pragma unroll
  for (int b = 0; b < PMAX / 16; ++b) {
    if (b < nb) {
#pragma unroll
      for (int j = 0; j < 16; ++j) s[j] = __fadd_rn(s[j], __fmul_rn(q[b * 16 + j], c[b * 16 + j]));
    }
  }
  float lane[4];
#pragma unroll
  for (int l = 0; l < 4; ++l) lane[l] = __fadd_rn(__fadd_rn(s[l], s[4 + l]), __fadd_rn(s[8 + l], s[12 + l]));
  float sum = __fadd_rn(__fadd_rn(__fadd_rn(lane[0], lane[1]), lane[2]), lane[3]);
#pragma unroll
  for (int d = 0; d < PMAX; ++d) {
    if (d >= nb * 16 && d < p) sum = __fadd_rn(sum, __fmul_rn(q[d], c[d]));
  }
  return sum;
}

}  // namespace

__global__ void __launch_bounds__(kBlock) sse_norms_kernel(const float* __restrict__ x, int rows, int p,
                                                            float* __restrict__ out) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= rows) return;
  const float* a = x + static_cast<size_t>(r) * p;
  float s[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) s[j] = 0.0f;
  int i = 0;
  for (; i + 15 < p; i += 16) {
#pragma unroll
    for (int j = 0; j < 16; ++j) s[j] = __fadd_rn(s[j], __fmul_rn(a[i + j], a[i + j]));
  }
  float lane[4];
#pragma unroll
  for (int l = 0; l < 4; ++l) lane[l] = __fadd_rn(__fadd_rn(s[l], s[4 + l]), __fadd_rn(s[8 + l], s[12 + l]));
  float sum = __fadd_rn(__fadd_rn(__fadd_rn(lane[0], lane[1]), lane[2]), lane[3]);
  for (; i < p; ++i) sum = __fadd_rn(sum, __fmul_rn(a[i], a[i]));
  out[r] = sum;
}

template <int PMAX>
__global__ void __launch_bounds__(kBlock) kmeans_assign_kernel(const float* __restrict__ x, int n, int p,
                                                                const float* __restrict__ x_norm,
                                                                const float* __restrict__ centroids,
                                                                const float* __restrict__ c_norm, int k,
                                                                int* __restrict__ assign, int* __restrict__ changed,
                                                                const int* __restrict__ active_runs) {
  constexpr int kTile = 8192 / PMAX;  // centroids per 32 KiB LDS tile
  __shared__ float tile[kTile * PMAX];
  __shared__ float tile_norm[kTile];
  // Problem blockIdx.y of a batch of equal-sized ones (rows n, clusters k).
  const int run = blockIdx.y;
  if (active_runs != nullptr && !active_runs[run]) return;  // whole block
  x += static_cast<size_t>(run) * n * p;
  x_norm += static_cast<size_t>(run) * n;
  centroids += static_cast<size_t>(run) * k * p;
  c_norm += static_cast<size_t>(run) * k;
  assign += static_cast<size_t>(run) * n;
  changed += run;
  const int row = blockIdx.x * kBlock + threadIdx.x;
  const bool active = row < n;
  float q[PMAX];
#pragma unroll
  for (int d = 0; d < PMAX; ++d) q[d] = active && d < p ? x[static_cast<size_t>(row) * p + d] : 0.0f;
  const float xn = active ? x_norm[row] : 0.0f;
  int best = 0;
  float best_d = INFINITY;
  for (int c0 = 0; c0 < k; c0 += kTile) {
    const int nt = min(kTile, k - c0);
    __syncthreads();
    for (int e = threadIdx.x; e < kTile * PMAX; e += kBlock) {
      const int r = e / PMAX, d = e - r * PMAX;
      tile[e] = r < nt && d < p ? centroids[static_cast<size_t>(c0 + r) * p + d] : 0.0f;
    }
    for (int e = threadIdx.x; e < kTile; e += kBlock) tile_norm[e] = e < nt ? c_norm[c0 + e] : 0.0f;
    __syncthreads();
    if (active) {
      for (int r = 0; r < nt; ++r) {
        const float dot = sse_dot<PMAX>(q, &tile[r * PMAX], p);
        const float dist = __fsub_rn(__fadd_rn(xn, tile_norm[r]), __fmul_rn(2.0f, dot));
        const int c = c0 + r;
        if (dist < best_d || (dist == best_d && c < best)) {
          best_d = dist;
          best = c;
        }
      }
    }
  }
  if (active && assign[row] != best) {
    assign[row] = best;
    atomicAdd(changed, 1);
  }
}

__global__ void __launch_bounds__(kBlock) kmeans_update_kernel(const float* __restrict__ x, int n, int p,
                                                                const int* __restrict__ assign, int k,
                                                                float* __restrict__ centroids,
                                                                float* __restrict__ sizes,
                                                                const int* __restrict__ active_runs) {
  constexpr int kChunk = 4 * kBlock;  // rows per ballot round
  __shared__ uint64_t masks[kChunk / 64];
  const int run = blockIdx.y;
  if (active_runs != nullptr && !active_runs[run]) return;  // whole block
  x += static_cast<size_t>(run) * n * p;
  assign += static_cast<size_t>(run) * n;
  centroids += static_cast<size_t>(run) * k * p;
  sizes += static_cast<size_t>(run) * k;
  const int c = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int f = threadIdx.x;
  float acc = 0.0f, count = 0.0f;
  for (int base = 0; base < n; base += kChunk) {
    __syncthreads();
#pragma unroll
    for (int part = 0; part < kChunk / kBlock; ++part) {
      const int row = base + part * kBlock + static_cast<int>(threadIdx.x);
      const uint64_t b = __ballot(row < n && assign[row] == c);
      if (lane == 0) masks[part * (kBlock / 64) + wave] = b;
    }
    __syncthreads();
    // Mask w covers rows base + 64 w .. + 63: ascending w and bits give
    // ascending rows.
    for (int w = 0; w < kChunk / 64; ++w) {
      uint64_t b = masks[w];
      while (b != 0ull) {
        const int bit = __ffsll(static_cast<unsigned long long>(b)) - 1;
        b &= b - 1ull;
        const int row = base + w * 64 + bit;
        count = __fadd_rn(count, 1.0f);
        if (f < p) acc = __fadd_rn(acc, x[static_cast<size_t>(row) * p + f]);
      }
    }
  }
  if (f < p) {
    float v = acc;
    if (count != 0.0f) v = __fmul_rn(acc, __fdiv_rn(1.0f, count));
    centroids[static_cast<size_t>(c) * p + f] = v;
  }
  if (threadIdx.x == 0) sizes[c] = count;
}

__global__ void __launch_bounds__(kBlock) local_fill_kernel(int* __restrict__ local, size_t total) {
  for (size_t e = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * kBlock) {
    local[e] = -1;
  }
}

// local[r][landmark row] = its position in run r's landmark list (rows are
// distinct within a run, so no two threads write one entry).
__global__ void __launch_bounds__(kBlock) local_scatter_kernel(ProjectParams prm) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  const int r = blockIdx.y;
  const int begin = prm.run_offsets[r];
  if (j >= prm.run_offsets[r + 1] - begin) return;
  prm.local[static_cast<size_t>(r) * prm.n + prm.landmarks[begin + j]] = j;
}

__global__ void __launch_bounds__(kProjectBlock) project_kernel(ProjectParams prm) {
  __shared__ int votes_lds[kProjectBlock * kMaxVotes];
  const int i = blockIdx.x * kProjectBlock + threadIdx.x;
  const int r = blockIdx.y;
  if (i >= prm.n) return;  // no block barrier below
  const int begin = prm.run_offsets[r];
  const int* lab = prm.clbest + begin;
  const int* local = prm.local + static_cast<size_t>(r) * prm.n;
  auto find = [&](int g) -> int { return local[g]; };
  int out;
  const int self = find(i);
  if (self >= 0) {
    out = lab[self];
  } else {
    // Votes interleaved by thread: slot v of thread t at v * block + t.
    int* votes = votes_lds + threadIdx.x;
    int nv = 0;
    const int* row = prm.graph + static_cast<size_t>(i) * prm.neighbors;
    for (int j = 0; j < prm.neighbors && nv < prm.votes; ++j) {
      const int g = row[j];
      if (g < 0 || g >= prm.n) continue;
      const int pos = find(g);
      if (pos >= 0) votes[(nv++) * kProjectBlock] = lab[pos];
    }
    if (nv == 0) {
      out = lab[0];
    } else {
      int best = 0, best_count = 0;
      for (int a = 0; a < nv; ++a) {
        const int va = votes[a * kProjectBlock];
        int cnt = 0;
        for (int b = 0; b < nv; ++b) cnt += votes[b * kProjectBlock] == va ? 1 : 0;
        if (cnt > best_count || (cnt == best_count && va < best)) {
          best = va;
          best_count = cnt;
        }
      }
      out = best;
    }
  }
  prm.res[static_cast<size_t>(r) * prm.n + i] = out;
}

// ---- launchers --------------------------------------------------------------

hipError_t launch_sse_norms(const float* x, int rows, int p, float* out, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  hipLaunchKernelGGL(sse_norms_kernel, dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0, s, x, rows, p, out);
  return hipGetLastError();
}

template <int PMAX>
static void launch_assign_p(const float* x, int n, int p, const float* xn, const float* cent, const float* cn,
                            int k, int* assign, int* changed, int runs, const int* active, hipStream_t s) {
  hipLaunchKernelGGL((kmeans_assign_kernel<PMAX>), dim3((n + kBlock - 1) / kBlock, runs), dim3(kBlock), 0, s, x, n,
                     p, xn, cent, cn, k, assign, changed, active);
}

hipError_t launch_kmeans_assign(const float* x, int n, int p, const float* x_norm, const float* centroids,
                                const float* c_norm, int k, int* assign, int* changed, hipStream_t s, int runs,
                                const int* active) {
  if (p < 1 || p > 256 || k < 1 || runs < 1) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  if (p <= 32) launch_assign_p<32>(x, n, p, x_norm, centroids, c_norm, k, assign, changed, runs, active, s);
  else if (p <= 64) launch_assign_p<64>(x, n, p, x_norm, centroids, c_norm, k, assign, changed, runs, active, s);
  else if (p <= 128) launch_assign_p<128>(x, n, p, x_norm, centroids, c_norm, k, assign, changed, runs, active, s);
  else launch_assign_p<256>(x, n, p, x_norm, centroids, c_norm, k, assign, changed, runs, active, s);
  return hipGetLastError();
}

hipError_t launch_kmeans_update(const float* x, int n, int p, const int* assign, int k, float* centroids,
                                float* sizes, hipStream_t s, int runs, const int* active) {
  if (p < 1 || p > kBlock || k < 1 || runs < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(kmeans_update_kernel, dim3(k, runs), dim3(kBlock), 0, s, x, n, p, assign, k, centroids, sizes,
                     active);
  return hipGetLastError();
}

hipError_t launch_project(const ProjectParams& prm, hipStream_t s) {
  if (prm.votes < 1 || prm.votes > kMaxVotes || prm.neighbors < 0 || prm.runs < 1 || prm.local == nullptr ||
      prm.max_landmarks < 1) {
    return hipErrorInvalidValue;
  }
  if (prm.n == 0) return hipSuccess;
  const size_t total = static_cast<size_t>(prm.runs) * prm.n;
  const size_t fill_blocks = std::min<size_t>((total + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL(local_fill_kernel, dim3(static_cast<unsigned>(fill_blocks)), dim3(kBlock), 0, s, prm.local,
                     total);
  hipLaunchKernelGGL(local_scatter_kernel, dim3((prm.max_landmarks + kBlock - 1) / kBlock, prm.runs), dim3(kBlock),
                     0, s, prm);
  hipLaunchKernelGGL(project_kernel, dim3((prm.n + kProjectBlock - 1) / kProjectBlock, prm.runs),
                     dim3(kProjectBlock), 0, s, prm);
  return hipGetLastError();
}

}  // namespace kodama_hip
