// PLS-LDA cross-validation on the device (plscv.cpp:2293-2507 with the fold
// cache 593-698, SIMPLS 816-955 and the LDA 1379-1510), all in FP64.
//
// Fold cache (once per set of folds): pls_moments_kernel forms the moment
// matrix of [x - shift | 1] over row chunks of each fold's own rows on the
// FP64 matrix cores (the shift is one row of the fold's run), batched over
// all chunks of all folds; pls_moments_reduce_kernel sums each fold's chunks;
// pls_cache_finalize_kernel takes a fold's training moments as the sum of its
// run's other folds' and derives the training mean and scale (the scale
// accumulator starts at 1, as train_center_scale_float does), the column sums
// of the standardised training rows and their Gram matrix G = Xs'Xs.
//
// Per labelling, two launches:
// pls_fold_sums_kernel (one wave per fold): the class sums of the fold's own
//          rows, sum x[row] per label code, and the class counts. Every row
//          is read once per evaluation; a fold's training rows are the other
//          folds of its run, so their class sums are sums of fold sums.
// pls_eval_kernel (one wave per fold; the PLS-LDA head, or with `da` the
//          least-squares PLS-DA head of PLSDACV, plscv.cpp:1210-1296):
//   S      Z[c] = (training class sum - n_c mean) / sd, the class's
//          standardised column sums, and S[c] = Z[c] - (n_c / n) xsum;
//   SIMPLS power refresh (two S S' products), normalisation through G, one
//          sequential Gram-Schmidt pass against earlier loadings, deflation;
//   LDA    class score means W'Z[c] / n_c and T'T = W'GW (equal to the sums
//          over the training scores), pooled covariance, ridge grid,
//          Cholesky solve;
//   score  validation scores, discriminants, argmax (strict >, first wins),
//          one lane per validation row.
// With one wave per fold the block barriers are a wave's own and a fold's
// arrays stay in LDS (G too for p <= 64) when they fit 64 KiB. Reductions are
// fixed butterflies, so a result never depends on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>

#include "pls_launch.h"

namespace kodama_hip {

namespace {

constexpr int kWave = 64;
constexpr int kT = kPlsThreads;
constexpr int kWaves = kT / kWave;
constexpr int kA = kPlsMaxComponents;
// The reference compares against float literals converted to the working type.
constexpr double kNormFloor = static_cast<double>(1.0e-10f);
constexpr double kTinyFloor = static_cast<double>(1.0e-20f);


}  // namespace

// ---- fold cache ---------------------------------------------------------------

namespace {
typedef double f64x4m __attribute__((ext_vector_type(4)));
constexpr int kMomentTilesPerWave = 8;
constexpr int kMomentTilesPerBlock = kMomentTilesPerWave * kWaves;
}  // namespace

// Moments of one chunk of a fold's own rows (blockIdx.x) for one group of up
// to 32 of the 16 x 16 output tiles (blockIdx.y): Z'Z with Z = [x - shift | 1]
// (zero padded to z16 columns), rows through LDS 16 at a time as doubles, the
// next block's loads in flight during the products (v_mfma_f64_16x16x4f64;
// the operand layout of pls_cross_kernel). Z'Z holds the shifted cross
// moments, the shifted column sums (column p) and the row count (p, p).
template <int PER>
__global__ void __launch_bounds__(kT) pls_moments_kernel(PlsCacheParams prm) {
  extern __shared__ double mst[];  // [16][z16]
  const int p = prm.p, zc = p + 1;
  const int z16 = (zc + 15) / 16 * 16;
  const int tiles_c = z16 / 16, tiles = tiles_c * tiles_c;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int k = blockIdx.x;
  const int f = prm.chunk_fold[k];
  const int b0 = prm.chunk_begin[k], b1 = prm.chunk_end[k];
  const int sr = prm.shift_row[f];
  const int tile0 = blockIdx.y * kMomentTilesPerBlock + wave;
  f64x4m acc[kMomentTilesPerWave];
#pragma unroll
  for (int t = 0; t < kMomentTilesPerWave; ++t) acc[t] = f64x4m{0.0, 0.0, 0.0, 0.0};
  const int kr = lane / 16, kc = lane % 16;
  // Element q of this thread in a block: e = threadIdx.x + q * 256 < 16 z16;
  // indices are clamped (the values are masked when stored).
  auto fetch = [&](int rb, float (&xv)[PER], float (&sv)[PER]) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = min(static_cast<int>(threadIdx.x) + q * kT, 16 * z16 - 1);
      const int rr = e / z16, c = min(e - rr * z16, p - 1);
      const int row = prm.val.rows[min(rb + rr, b1 - 1)];
      xv[q] = prm.x[static_cast<size_t>(row) * p + c];
      sv[q] = prm.x[static_cast<size_t>(sr) * p + c];
    }
  };
  float xv[PER], sv[PER];
  fetch(b0, xv, sv);
  for (int rb = b0; rb < b1; rb += 16) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = static_cast<int>(threadIdx.x) + q * kT;
      if (e < 16 * z16) {
        const int rr = e / z16, c = e - rr * z16;
        const double v = c < p ? static_cast<double>(xv[q]) - static_cast<double>(sv[q]) : (c == p ? 1.0 : 0.0);
        mst[e] = rb + rr < b1 ? v : 0.0;
      }
    }
    __syncthreads();
    if (rb + 16 < b1) fetch(rb + 16, xv, sv);
#pragma unroll
    for (int t = 0; t < kMomentTilesPerWave; ++t) {
      const int tile = tile0 + t * kWaves;
      if (tile < tiles) {
        const int tj = tile / tiles_c, tc = tile - tj * tiles_c;
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4) {
          const double a = mst[(k0 + kr) * z16 + tj * 16 + kc];
          const double b = mst[(k0 + kr) * z16 + tc * 16 + kc];
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  double* out = prm.partial + static_cast<size_t>(k) * z16 * z16;
#pragma unroll
  for (int t = 0; t < kMomentTilesPerWave; ++t) {
    const int tile = tile0 + t * kWaves;
    if (tile >= tiles) break;
    const int tj = tile / tiles_c, tc = tile - tj * tiles_c;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      out[static_cast<size_t>(tj * 16 + 4 * i + kr) * z16 + tc * 16 + kc] = acc[t][i];
    }
  }
}

// Moments of each fold (one workgroup per fold): its chunks' partials in
// chunk order.
__global__ void __launch_bounds__(kT) pls_moments_reduce_kernel(PlsCacheParams prm) {
  const int f = blockIdx.x, p = prm.p, zc = p + 1;
  const int z16 = (zc + 15) / 16 * 16;
  const int k0 = prm.chunk_off[f], k1 = prm.chunk_off[f + 1];
  double* out = prm.moments + static_cast<size_t>(f) * zc * zc;
  for (int e = threadIdx.x; e < zc * zc; e += kT) {
    const int a = e / zc, b = e - a * zc;
    double t = 0.0;
    for (int k = k0; k < k1; ++k) t += prm.partial[(static_cast<size_t>(k) * z16 + a) * z16 + b];
    out[e] = t;
  }
}

// Training statistics of each fold (one workgroup per fold) from the sum T of
// its run's other folds' moments (fixed fold order): with n = T[p][p],
// Y = T[.][p] the shifted column sums, d = mean - shift (center: Y / n,
// otherwise -shift), sum (x - mean)^2 = T[j][j] - 2 d Y + n d^2, the scale
// sqrt((1 + that) / max(1, n - 1)) (non-finite or zero -> 1), the
// standardised column sums (Y - n d) / sd and
// G[a][b] = (T[a][b] - d_a Y_b - d_b Y_a + n d_a d_b) / (sd_a sd_b).
__global__ void __launch_bounds__(kT) pls_cache_finalize_kernel(PlsCacheParams prm) {
  __shared__ double st[3 * kPlsMaxDims + 1];  // d, Y, sd, n
  const int f = blockIdx.x, p = prm.p, zc = p + 1;
  const int g0 = prm.run_first[f], g1 = prm.run_end[f];
  auto total = [&](int a, int b) {
    double t = 0.0;
    for (int g = g0; g < g1; ++g) {
      if (g != f) t += prm.moments[(static_cast<size_t>(g) * zc + a) * zc + b];
    }
    return t;
  };
  double* dv = st;
  double* yv = st + kPlsMaxDims;
  double* sdv = st + 2 * kPlsMaxDims;
  const double n = total(p, p);
  const int sr = prm.shift_row[f];
  for (int j = threadIdx.x; j < p; j += kT) {
    const double y = total(j, p);
    const double c = static_cast<double>(prm.x[static_cast<size_t>(sr) * p + j]);
    const double d = prm.center ? y / n : -c;
    double s = 1.0;
    if (prm.scale) {
      const double ss = total(j, j) - 2.0 * d * y + n * d * d;
      s = sqrt((1.0 + ss) / fmax(1.0, n - 1.0));
      s = s > 0.0 && isfinite(s) ? s : 1.0;
    }
    dv[j] = d;
    yv[j] = y;
    sdv[j] = s;
    prm.mean[static_cast<size_t>(f) * p + j] = prm.center ? c + d : 0.0;
    prm.sd[static_cast<size_t>(f) * p + j] = s;
    prm.xsum[static_cast<size_t>(f) * p + j] = (y - n * d) / s;
  }
  __syncthreads();
  double* gram = prm.gram + static_cast<size_t>(f) * p * p;
  for (int e = threadIdx.x; e < p * p; e += kT) {
    const int a = e / p, b = e - a * p;
    if (b > a) continue;
    const double v = (total(a, b) - dv[a] * yv[b] - dv[b] * yv[a] + n * dv[a] * dv[b]) / (sdv[a] * sdv[b]);
    gram[static_cast<size_t>(a) * p + b] = v;
    gram[static_cast<size_t>(b) * p + a] = v;
  }
}

// ---- evaluation -------------------------------------------------------------

namespace {

constexpr int kE = kPlsEvalThreads;  // one wave per fold

// Wave sums on the DPP crossbar instead of LDS permutes: within each row of
// 16 lanes, pairs (quad_perm [1,0,3,2]), quads ([2,3,0,1]), then rotations
// by 4 and 8 leave the row's sum in every lane of the row; lane 0 of each of
// the four rows is read into a scalar and the four row sums are added in row
// order, so every lane gets the same total and the order is fixed.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<unsigned>(b)), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<unsigned>(b >> 32)), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                                     static_cast<unsigned>(lo)));
}

__device__ __forceinline__ double read_lane(double v, int lane) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned>(b)), lane));
  const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned>(b >> 32)), lane));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// N independent wave sums at once (their cross-lane steps interleave).
template <int N>
__device__ __forceinline__ void lane_totals(double (&v)[N]) {
#pragma unroll
  for (int u = 0; u < N; ++u) v[u] += dpp_move<0xB1>(v[u]);   // quad_perm [1,0,3,2]
#pragma unroll
  for (int u = 0; u < N; ++u) v[u] += dpp_move<0x4E>(v[u]);   // quad_perm [2,3,0,1]
#pragma unroll
  for (int u = 0; u < N; ++u) v[u] += dpp_move<0x124>(v[u]);  // row_ror:4
#pragma unroll
  for (int u = 0; u < N; ++u) v[u] += dpp_move<0x128>(v[u]);  // row_ror:8
#pragma unroll
  for (int u = 0; u < N; ++u) {
    v[u] = ((read_lane(v[u], 0) + read_lane(v[u], 16)) + read_lane(v[u], 32)) + read_lane(v[u], 48);
  }
}

// Sum over the wave; every lane gets the total.
__device__ __forceinline__ double lane_total(double v) {
  double t[1] = {v};
  lane_totals(t);
  return t[0];
}
constexpr int kBatch = 8;  // reductions in flight

// LDS carve of one fold's wave (doubles first, then ints).
struct EvalLds {
  double *S, *wt, *ld, *mean, *sd, *rv, *mat, *tt;
  double *G, *W, *V, *GW, *Z, *mu, *LIN, *CONST;  // null when in global memory
  double *Wp, *LINp;  // compact: weights [8 x 64] and LDA rows [C x 8], zero padded
  int *segc, *cnt, *scal;
};

// Columns of the padded per-column arrays (mean, sd: 64 at least, so that the
// compact validation loop runs over 64 columns without bounds).
__host__ __device__ inline int eval_cols(int p) { return p > 64 ? p : 64; }

__host__ __device__ inline size_t eval_lds_doubles(int p, int C, int rc, bool compact) {
  size_t d = static_cast<size_t>(C) * p + 2 * static_cast<size_t>(p) + 2 * static_cast<size_t>(eval_cols(p)) + C +
             2 * static_cast<size_t>(rc) * rc;
  if (compact) {
    d += static_cast<size_t>(p) * p + 3 * static_cast<size_t>(rc) * p + static_cast<size_t>(C) * p +
         2 * static_cast<size_t>(C) * rc + C + 8 * 64 + 8 * static_cast<size_t>(C);
  }
  return d;
}

__device__ EvalLds eval_carve(double* base, int p, int C, int rc, bool compact) {
  EvalLds s{};
  double* at = base;
  auto take = [&](size_t n) {
    double* r = at;
    at += n;
    return r;
  };
  s.S = take(static_cast<size_t>(C) * p);
  s.wt = take(p);
  s.ld = take(p);
  s.mean = take(eval_cols(p));
  s.sd = take(eval_cols(p));
  s.rv = take(C);
  s.mat = take(static_cast<size_t>(rc) * rc);
  s.tt = take(static_cast<size_t>(rc) * rc);
  if (compact) {
    s.G = take(static_cast<size_t>(p) * p);
    s.W = take(static_cast<size_t>(rc) * p);
    s.V = take(static_cast<size_t>(rc) * p);
    s.GW = take(static_cast<size_t>(rc) * p);
    s.Z = take(static_cast<size_t>(C) * p);
    s.mu = take(static_cast<size_t>(C) * rc);
    s.LIN = take(static_cast<size_t>(C) * rc);
    s.CONST = take(C);
    s.Wp = take(8 * 64);
    s.LINp = take(8 * static_cast<size_t>(C));
  }
  int* ip = reinterpret_cast<int*>(at);
  s.segc = ip;
  s.cnt = ip + C;
  s.scal = ip + 2 * C;
  return s;
}

// Power-refresh SIMPLS (plscv.cpp:816-955) by one wave, lanes over columns:
// S [m x p] (LDS, row c at S[c * p], deflated in place), G [p x p] symmetric
// (read as G[j * p + i] so that the lanes of one j are contiguous), weights and
// loadings component-major (W[a * p + j]). Per component the previous weight
// (w0 for the first) takes two S'S power steps, is normalised (a failed first
// refresh restarts from the strongest row of S), scaled to unit score norm
// through G, its loading G w is Gram-Schmidt orthogonalised once against the
// earlier loadings and normalised, and S is deflated. Returns the fitted rank
// (0: degenerate).
__device__ __forceinline__ int simpls_wave(double* S, int m, int p, int rc, const double* G, const double* w0, double* W, double* V,
                           double* wt, double* ld, double* rv, int* s_int) {
  const int lane = threadIdx.x;
  int fitted = 0;
  auto refresh = [&]() {
    for (int it = 0; it < 2; ++it) {
      for (int c0 = 0; c0 < m; c0 += kBatch) {
        double a[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          a[u] = 0.0;
          if (c0 + u < m) {
            for (int j = lane; j < p; j += kE) a[u] += S[(c0 + u) * p + j] * wt[j];
          }
        }
        lane_totals(a);
        if (lane == 0) {
#pragma unroll
          for (int u = 0; u < kBatch; ++u) {
            if (c0 + u < m) rv[c0 + u] = a[u];
          }
        }
      }
      __syncthreads();
      for (int j = lane; j < p; j += kE) {
        double a = 0.0;
        for (int c = 0; c < m; ++c) a += S[c * p + j] * rv[c];
        wt[j] = a;
      }
      __syncthreads();
    }
  };
  auto norm_of = [&](const double* v) {
    double part = 0.0;
    for (int j = lane; j < p; j += kE) part += v[j] * v[j];
    return lane_total(part);
  };
  for (int a = 0; a < rc; ++a) {
    if (a == 0) {
      for (int j = lane; j < p; j += kE) wt[j] = w0[j];
      __syncthreads();
    }
    refresh();
    double ss = norm_of(wt);
    double nrm = isfinite(ss) && ss >= 0.0 ? sqrt(ss) : INFINITY;
    if (a == 0 && (!isfinite(nrm) || nrm <= kNormFloor)) {
      // The strongest row of S restarts the power iteration.
      if (lane == 0) {
        int strongest = -1;
        double best = 0.0;
        for (int c = 0; c < m; ++c) {
          double sq = 0.0;
          for (int j = 0; j < p; ++j) sq += S[c * p + j] * S[c * p + j];
          if (isfinite(sq) && sq > best) {
            best = sq;
            strongest = c;
          }
        }
        s_int[0] = strongest >= 0 && best > 1.0e-20 ? strongest : -1;
      }
      __syncthreads();
      const int strongest = s_int[0];
      if (strongest >= 0) {
        for (int j = lane; j < p; j += kE) wt[j] = S[strongest * p + j];
        __syncthreads();
        refresh();
        ss = norm_of(wt);
        nrm = isfinite(ss) && ss >= 0.0 ? sqrt(ss) : INFINITY;
      }
    }
    if (!isfinite(nrm) || nrm <= kNormFloor) break;
    for (int j = lane; j < p; j += kE) wt[j] /= nrm;
    __syncthreads();
    for (int i = lane; i < p; i += kE) {
      // Four partial sums (columns j mod 4), added in order.
      double q[4] = {0.0, 0.0, 0.0, 0.0};
      int j = 0;
      for (; j + 4 <= p; j += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] += G[static_cast<size_t>(j + u) * p + i] * wt[j + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (j + u < p) q[u] += G[static_cast<size_t>(j + u) * p + i] * wt[j + u];
      }
      ld[i] = (q[0] + q[1]) + (q[2] + q[3]);
    }
    __syncthreads();
    double part = 0.0;
    for (int j = lane; j < p; j += kE) part += wt[j] * ld[j];
    const double tnorm = lane_total(part);
    if (!isfinite(tnorm) || tnorm <= kTinyFloor) break;
    const double inv = 1.0 / sqrt(tnorm);
    for (int j = lane; j < p; j += kE) {
      wt[j] *= inv;
      ld[j] *= inv;
    }
    __syncthreads();
    // One sequential Gram-Schmidt pass against the earlier loadings.
    for (int b = 0; b < a; ++b) {
      double pp = 0.0;
      const double* vb = V + static_cast<size_t>(b) * p;
      for (int j = lane; j < p; j += kE) pp += vb[j] * ld[j];
      const double proj = lane_total(pp);
      for (int j = lane; j < p; j += kE) ld[j] -= proj * vb[j];
      __syncthreads();
    }
    const double lnn = norm_of(ld);
    const double ln = sqrt(fmax(0.0, lnn));
    if (!isfinite(ln) || ln <= kNormFloor) break;
    for (int j = lane; j < p; j += kE) {
      ld[j] /= ln;
      W[static_cast<size_t>(a) * p + j] = wt[j];
      V[static_cast<size_t>(a) * p + j] = ld[j];
    }
    __syncthreads();
    for (int c0 = 0; c0 < m; c0 += kBatch) {
      double d[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        d[u] = 0.0;
        if (c0 + u < m) {
          for (int j = lane; j < p; j += kE) d[u] += ld[j] * S[(c0 + u) * p + j];
        }
      }
      lane_totals(d);
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        if (c0 + u < m) {
          for (int j = lane; j < p; j += kE) S[(c0 + u) * p + j] -= ld[j] * d[u];
        }
      }
    }
    __syncthreads();
    fitted = a + 1;
  }
  return fitted;
}

}  // namespace

// Class sums of each fold's validation rows: one workgroup per fold of
// `waves` waves, each summing a contiguous share of the rows (ascending) into
// its own LDS partial with lanes over columns, so every partial sum has one
// owner lane and a fixed order; the partials are then added in wave order.
// Per 64 rows the lanes load the row ids and codes, and sixteen rows' columns
// are loaded before they are added.
__global__ void __launch_bounds__(kT) pls_fold_sums_kernel(PlsFoldSumsParams prm, int waves) {
  extern __shared__ double fs_lds[];
  const int f = blockIdx.x, p = prm.p, C = prm.max_classes;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const size_t cp = static_cast<size_t>(C) * p;
  double* acc = fs_lds;                                                          // [waves x C x p]
  int* cnt = reinterpret_cast<int*>(fs_lds + static_cast<size_t>(waves) * cp);  // [waves x C]
  for (size_t e = threadIdx.x; e < waves * cp; e += blockDim.x) acc[e] = 0.0;
  for (int e = threadIdx.x; e < waves * C; e += blockDim.x) cnt[e] = 0;
  __syncthreads();
  const int v0 = prm.val.offsets[f];
  const int nv = prm.val.offsets[f + 1] - v0;
  const int share = (nv + waves - 1) / waves;
  const int r0 = min(nv, wave * share), r1 = min(nv, r0 + share);
  double* my = acc + wave * cp;
  for (int i0 = r0; i0 < r1; i0 += kWave) {
    const int mine = i0 + lane < r1 ? prm.val.rows[v0 + i0 + lane] : -1;
    const int mycode = mine >= 0 ? prm.codes[mine] : 0;
    const int cnt_rows = min(kWave, r1 - i0);
    if (mine >= 0) atomicAdd(&cnt[wave * C + mycode], 1);
    for (int u0 = 0; u0 < cnt_rows; u0 += 16) {  // uniform in the wave
      int row[16], code[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        row[u] = __shfl(mine, u0 + u);
        code[u] = __shfl(mycode, u0 + u);
      }
      for (int j0 = 0; j0 < p; j0 += kWave) {
        const int j = j0 + lane;
        // Unconditional loads: rows past the batch re-read the batch's first
        // row and columns past p column p - 1; neither is added.
        float xv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int r = u0 + u < cnt_rows ? row[u] : row[0];
          xv[u] = prm.x[static_cast<size_t>(r) * p + min(j, p - 1)];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          if (u0 + u < cnt_rows && j < p) my[code[u] * p + j] += static_cast<double>(xv[u]);
        }
      }
    }
  }
  __syncthreads();
  double* out = prm.fsum + static_cast<size_t>(f) * cp;
  for (size_t e = threadIdx.x; e < cp; e += blockDim.x) {
    double t = acc[e];
    for (int w = 1; w < waves; ++w) t += acc[w * cp + e];
    out[e] = t;
  }
  int* oc = prm.fcnt + static_cast<size_t>(f) * C;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    int t = 0;
    for (int w = 0; w < waves; ++w) t += cnt[w * C + c];
    oc[c] = t;
  }
}

// One wave per fold (see the file comment): S from the class sums of the
// run's other folds, SIMPLS, the LDA or PLS-DA head, the validation argmax.
template <bool COMPACT>
__global__ void __launch_bounds__(kE) pls_eval_kernel(PlsEvalParams prm) {
  extern __shared__ double pls_lds[];
  const int f = blockIdx.x;
  const int p = prm.p;
  const int lane = threadIdx.x;
  const int C = prm.max_classes;
  const int rcap = prm.rank_cap;
  const EvalLds L = eval_carve(pls_lds, p, C, rcap, COMPACT);
  const size_t fA = static_cast<size_t>(f) * kA;
  const double* Gg = prm.gram + static_cast<size_t>(f) * p * p;
  // The compact layout's arrays are all LDS (known at compile time, so their
  // accesses are LDS instructions); the wide layout's are global.
  const double* G = COMPACT ? L.G : Gg;
  double* Wg = prm.weights + fA * p;
  double* W = COMPACT ? L.W : Wg;
  double* V = COMPACT ? L.V : prm.loadings + fA * p;
  double* GW = COMPACT ? L.GW : prm.gw + fA * p;
  double* Z = COMPACT ? L.Z : prm.zsum + static_cast<size_t>(f) * C * p;
  double* MU = COMPACT ? L.mu : prm.mu + static_cast<size_t>(f) * C * kA;
  double* LIN = COMPACT ? L.LIN : prm.linear + static_cast<size_t>(f) * C * kA;
  double* CONST = COMPACT ? L.CONST : prm.constant + static_cast<size_t>(f) * C;
  const int v0 = prm.val.offsets[f];
  const int nv = prm.val.offsets[f + 1] - v0;
  const int* vrows = prm.val.rows + v0;
  const double* xsum = prm.xsum + static_cast<size_t>(f) * p;
  double* TV = prm.val_scores + static_cast<size_t>(v0) * kA;       // [i * kA + a]
  double* DISC = prm.discriminants + static_cast<size_t>(v0) * C;    // [i * C + c]

  for (int j = lane; j < eval_cols(p); j += kE) {
    L.mean[j] = j < p ? prm.mean[static_cast<size_t>(f) * p + j] : 0.0;
    L.sd[j] = j < p ? prm.sd[static_cast<size_t>(f) * p + j] : 1.0;
  }
  if (COMPACT) {
    // Eight loads in flight per lane.
    for (int e0 = 0; e0 < p * p; e0 += 8 * kE) {
      double g[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) g[u] = Gg[min(e0 + u * kE + lane, p * p - 1)];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * kE + lane;
        if (e < p * p) L.G[e] = g[u];
      }
    }
  }
  // Training rows per class: the run's other folds.
  const int g0 = prm.run_first[f], g1 = prm.run_end[f];
  for (int c = lane; c < C; c += kE) {
    int t = 0;
    for (int g = g0; g < g1; ++g) t += g != f ? prm.fcnt[static_cast<size_t>(g) * C + c] : 0;
    L.cnt[c] = t;
  }
  __syncthreads();
  if (lane == 0) {
    int s = 0, nt = 0;
    for (int c = 0; c < C; ++c) {
      if (L.cnt[c] > 0) {
        L.segc[s++] = c;
        nt += L.cnt[c];
      }
    }
    L.scal[0] = s;
    L.scal[1] = nt;
    prm.seg_count[f] = s;
  }
  __syncthreads();
  const int cf = L.scal[0];
  const int nt = L.scal[1];
  for (int s = lane; s < cf; s += kE) prm.seg_code[static_cast<size_t>(f) * C + s] = L.segc[s];

  if (cf <= 1) {
    // One class in the training fold: every validation row gets it.
    const int code = cf == 1 ? L.segc[0] : 0;
    for (int i = lane; i < nv; i += kE) prm.pred_code[vrows[i]] = code;
    if (lane == 0) {
      prm.components[f] = 1;
      prm.rank[f] = 0;
    }
    return;
  }

  // -- S = X'(Y - ybar) from the class sums ----------------------------------------
  // Z[c] = sum over the class's training rows of the standardised row
  // ((sum x - n_c mean) / sd), S[c] = Z[c] - (n_c / n) xsum.
  const double inv_nt = 1.0 / static_cast<double>(nt);
  for (int ci = 0; ci < cf; ++ci) {
    const int c = L.segc[ci];
    const int nc = L.cnt[c];
    const double ym = static_cast<double>(nc) * inv_nt;
    for (int j = lane; j < p; j += kE) {
      // The other folds' sums in ascending fold order, four loads in flight
      // (a skipped fold adds 0).
      double z = 0.0;
      for (int gb = g0; gb < g1; gb += 4) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int g = min(gb + u, g1 - 1);
          v[u] = prm.fsum[(static_cast<size_t>(g) * C + c) * p + j];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = gb + u < g1 && gb + u != f ? v[u] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) z += v[u];
      }
      const double zs = (z - static_cast<double>(nc) * L.mean[j]) / L.sd[j];
      Z[ci * p + j] = zs;
      L.S[ci * p + j] = zs - ym * xsum[j];
    }
  }
  __syncthreads();

  // -- power-refresh SIMPLS ----------------------------------------------------
  const int rc = max(1, min(rcap, max(1, nt - 1)));
  const int fitted = simpls_wave(L.S, cf, p, rc, G, prm.w0, W, V, L.wt, L.ld, L.rv, L.scal + 2);
  if (fitted < 1) {
    // Degenerate fit: the training majority (the smaller code on ties).
    if (lane == 0) {
      int best = L.segc[0], best_n = -1;
      for (int c = 0; c < cf; ++c) {
        const int cnt = L.cnt[L.segc[c]];
        if (cnt > best_n) {
          best = L.segc[c];
          best_n = cnt;
        }
      }
      L.scal[2] = best;
      prm.components[f] = 1;
      prm.rank[f] = 0;
    }
    __syncthreads();
    const int code = L.scal[2];
    for (int i = lane; i < nv; i += kE) prm.pred_code[vrows[i]] = code;
    return;
  }
  const int A = min(prm.requested, fitted);
  if (lane == 0) {
    prm.components[f] = A;
    prm.rank[f] = fitted;
  }
  if (COMPACT) {
    for (int e = lane; e < fitted * p; e += kE) Wg[e] = L.W[e];
  }

  // -- class statistics ------------------------------------------------------------
  // Compact layout with up to 8 components: the weights zero-padded to
  // 8 x 64, so that the products below read them without bounds checks.
  const bool padded = COMPACT && A <= 8;
  if (padded) {
    for (int e = lane; e < 8 * 64; e += kE) {
      const int u = e >> 6, j = e & 63;
      L.Wp[e] = u < A && j < p ? W[static_cast<size_t>(u) * p + j] : 0.0;
    }
    __syncthreads();
    for (int i = lane; i < p; i += kE) {
      double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
      for (int j = 0; j < p; ++j) {
        const double g = G[static_cast<size_t>(j) * p + i];
#pragma unroll
        for (int u = 0; u < 8; ++u) s[u] += g * L.Wp[u * 64 + j];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (u < A) GW[static_cast<size_t>(u) * p + i] = s[u];
      }
    }
  }
  // GW = G W (component-major), eight components per pass over G.
  for (int a0 = 0; a0 < (padded ? 0 : A); a0 += 8) {
    for (int i = lane; i < p; i += kE) {
      double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      for (int j = 0; j < p; ++j) {
        const double g = G[static_cast<size_t>(j) * p + i];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (a0 + u < A) s[u] += g * W[static_cast<size_t>(a0 + u) * p + j];
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (a0 + u < A) GW[static_cast<size_t>(a0 + u) * p + i] = s[u];
      }
    }
  }
  // PLS-DA: T'(Y - ybar) = W'S0 with S0 = Z - (n_c / n) xsum; LDA: the class
  // score means W'Z[c] / n_c. Kept in MU as [c * A + a].
  for (int q0 = 0; q0 < cf * A; q0 += kBatch) {
    double v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      v[u] = 0.0;
      const int q = q0 + u;
      if (q < cf * A) {
        const int ci = q / A, a = q - ci * A;
        const double ym = static_cast<double>(L.cnt[L.segc[ci]]) * inv_nt;
        for (int j = lane; j < p; j += kE) {
          const double zj = prm.da ? Z[ci * p + j] - ym * xsum[j] : Z[ci * p + j];
          v[u] += W[static_cast<size_t>(a) * p + j] * zj;
        }
      }
    }
    lane_totals(v);
    if (lane == 0) {
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int q = q0 + u;
        if (q < cf * A) {
          const int cnt = L.cnt[L.segc[q / A]];
          MU[q] = prm.da ? v[u] : (cnt > 0 ? v[u] * (1.0 / static_cast<double>(cnt)) : 0.0);
        }
      }
    }
  }
  __syncthreads();
  // T'T = W' G W.
  for (int q0 = 0; q0 < A * A; q0 += kBatch) {
    double v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      v[u] = 0.0;
      const int q = q0 + u;
      if (q < A * A) {
        const int r = q / A, c2 = q - r * A;
        for (int j = lane; j < p; j += kE) v[u] += W[static_cast<size_t>(r) * p + j] * GW[static_cast<size_t>(c2) * p + j];
      }
    }
    lane_totals(v);
    if (lane == 0) {
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        if (q0 + u < A * A) L.tt[q0 + u] = v[u];
      }
    }
  }
  __syncthreads();

  double* mat = L.mat;
  double* tt = L.tt;
  double* mu = MU;
  if (prm.da) {
    // -- PLS-DA regression (plscv.cpp:1220-1296): Gauss-Jordan elimination with
    // partial pivoting on (T'T + 1e-9 I) B = T'(Y - ybar), columns whose pivot
    // is below 1e-12 skipped; lanes over matrix entries. mat holds T'T, mu the
    // right-hand sides (row a of column c at mu[c * A + a]).
    for (int q = lane; q < A * A; q += kE) {
      const int r = q / A, c2 = q - r * A;
      const int lo = min(r, c2), hi = max(r, c2);
      mat[q] = tt[hi * A + lo] + (r == c2 ? static_cast<double>(1.0e-9f) : 0.0);
    }
    __syncthreads();
    for (int col = 0; col < A; ++col) {
      int piv = col;
      double best = fabs(mat[col * A + col]);
      for (int r = col + 1; r < A; ++r) {
        const double v = fabs(mat[r * A + col]);
        if (v > best) {
          best = v;
          piv = r;
        }
      }
      if (best < static_cast<double>(1.0e-12f)) continue;
      if (piv != col) {
        for (int j = lane; j < A; j += kE) {
          const double t = mat[col * A + j];
          mat[col * A + j] = mat[piv * A + j];
          mat[piv * A + j] = t;
        }
        for (int c = lane; c < cf; c += kE) {
          const double t = mu[c * A + col];
          mu[c * A + col] = mu[c * A + piv];
          mu[c * A + piv] = t;
        }
        __syncthreads();
      }
      const double div = mat[col * A + col];
      __syncthreads();
      for (int j = col + lane; j < A; j += kE) mat[col * A + j] /= div;
      for (int c = lane; c < cf; c += kE) mu[c * A + col] /= div;
      __syncthreads();
      // Row factors read before any row changes (tt is free once copied).
      for (int r = lane; r < A; r += kE) tt[r] = mat[r * A + col];
      __syncthreads();
      for (int q = lane; q < A * (A + cf); q += kE) {
        const int r = q / (A + cf), j = q - r * (A + cf);
        const double fr = tt[r];
        if (r == col || fr == 0.0) continue;
        if (j < A) {
          if (j >= col) mat[r * A + j] -= fr * mat[col * A + j];
        } else {
          const int c = j - A;
          mu[c * A + r] -= fr * mu[c * A + col];
        }
      }
      __syncthreads();
    }
    // Discriminant of class c: ybar_c + t' B[:, c].
    for (int c = lane; c < cf; c += kE) {
      for (int a = 0; a < A; ++a) LIN[static_cast<size_t>(c) * A + a] = mu[c * A + a];
      CONST[c] = static_cast<double>(L.cnt[L.segc[c]]) * inv_nt;
    }
    __syncthreads();
  } else {
    // -- LDA: pooled covariance, ridge grid + Cholesky (lanes over rows /
    // right-hand sides) ---------------------------------------------------------------
    const double df = static_cast<double>(max(1, nt - cf));
    for (int q = lane; q < A * A; q += kE) {
      const int r = q / A, c2 = q - r * A;
      const int lo = min(r, c2), hi = max(r, c2);
      double between = 0.0;
      for (int c = 0; c < cf; ++c) {
        between += static_cast<double>(L.cnt[L.segc[c]]) * mu[c * A + hi] * mu[c * A + lo];
      }
      // The lower triangle is the one computed (symmetric use of T'T).
      mat[q] = (tt[hi * A + lo] - between) / df;
    }
    __syncthreads();
    // The factor is written over tt (lower triangle); the solution of each
    // class column goes to LIN[c * A + a].
    double trace = 0.0;
    for (int a = 0; a < A; ++a) trace += mat[a * A + a];
    const double rs = isfinite(trace) && trace > 0.0 ? trace / static_cast<double>(max(1, A)) : 1.0;
    const float grid[6] = {1e-8f, 1e-6f, 1e-5f, 1e-4f, 1e-3f, 1e-2f};
    bool ok = false;
    double* lo = tt;
    for (int g = 0; g < 6 && !ok; ++g) {
      const double lam = static_cast<double>(grid[g]) * rs;
      bool good = true;
      for (int c = 0; c < A && good; ++c) {
        // Diagonal entry (every lane computes it; lanes agree).
        double v = mat[c * A + c] + lam;
        for (int k = 0; k < c; ++k) v -= lo[c * A + k] * lo[c * A + k];
        if (!isfinite(v) || v <= 0.0) {
          good = false;
          break;
        }
        const double d = sqrt(v);
        __syncthreads();
        if (lane == 0) lo[c * A + c] = d;
        for (int r = c + 1 + lane; r < A; r += kE) {
          double w = mat[r * A + c];
          for (int k = 0; k < c; ++k) w -= lo[r * A + k] * lo[c * A + k];
          lo[r * A + c] = w / d;
        }
        __syncthreads();
      }
      ok = good;
    }
    if (!ok) {
      // The host reports the failed factorisation (components = -1).
      if (lane == 0) prm.components[f] = -1;
      return;
    }
    for (int c = lane; c < cf; c += kE) {
      double* sol = LIN + static_cast<size_t>(c) * A;
      for (int r = 0; r < A; ++r) {
        double v = mu[c * A + r];
        for (int k = 0; k < r; ++k) v -= lo[r * A + k] * sol[k];
        sol[r] = v / lo[r * A + r];
      }
      for (int r = A - 1; r >= 0; --r) {
        double v = sol[r];
        for (int k = r + 1; k < A; ++k) v -= lo[k * A + r] * sol[k];
        sol[r] = v / lo[r * A + r];
      }
      double dm = 0.0;
      for (int a = 0; a < A; ++a) dm += mu[c * A + a] * sol[a];
      const int cnt = L.cnt[L.segc[c]];
      const double prior = fmax(static_cast<double>(cnt) / static_cast<double>(max(1, nt)),
                                static_cast<double>(FLT_MIN));
      CONST[c] = -0.5 * dm + log(prior);
    }
    __syncthreads();
  }

  if (padded) {
    for (int e = lane; e < cf * 8; e += kE) {
      const int c = e >> 3, u = e & 7;
      L.LINp[e] = u < A ? LIN[static_cast<size_t>(c) * A + u] : 0.0;
    }
    __syncthreads();
  }
  // -- validation scores, discriminants, argmax (one lane per row) ----------------
  for (int i = lane; i < nv; i += kE) {
    const size_t row = static_cast<size_t>(vrows[i]) * p;
    double* tv = TV + static_cast<size_t>(i) * kA;
    int best = 0;
    double best_s = -INFINITY;
    if (padded) {
      // p <= 64, 64 columns in chunks of 16 whose loads are issued one chunk
      // ahead (columns past p re-read column p - 1, have mean 0, sd 1 and
      // zero weights, so they add exact zeros); the scores stay in registers
      // for the discriminants.
      double t[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      float cur[16];
#pragma unroll
      for (int v = 0; v < 16; ++v) cur[v] = prm.x[row + min(v, p - 1)];
#pragma unroll 1
      for (int j0 = 0; j0 < 64; j0 += 16) {
        float nxt[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) nxt[v] = prm.x[row + min(j0 + 16 + v, p - 1)];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int j = j0 + v;
          const double z = (static_cast<double>(cur[v]) - L.mean[j]) / L.sd[j];
#pragma unroll
          for (int u = 0; u < 8; ++u) t[u] += z * L.Wp[u * 64 + j];
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) cur[v] = nxt[v];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (u < A) tv[u] = t[u];
      }
      for (int c = 0; c < cf; ++c) {
        double sc = CONST[c];
#pragma unroll
        for (int u = 0; u < 8; ++u) sc += t[u] * L.LINp[c * 8 + u];
        DISC[static_cast<size_t>(i) * C + c] = sc;
        if (sc > best_s) {
          best_s = sc;
          best = c;
        }
      }
    } else {
      for (int a0 = 0; a0 < A; a0 += 8) {
        double t[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        // Sixteen of the row's loads in flight before their use.
        for (int j0 = 0; j0 < p; j0 += 16) {
          float xv[16];
#pragma unroll
          for (int v = 0; v < 16; ++v) xv[v] = prm.x[row + min(j0 + v, p - 1)];
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            const int j = j0 + v;
            if (j < p) {
              const double z = (static_cast<double>(xv[v]) - L.mean[j]) / L.sd[j];
#pragma unroll
              for (int u = 0; u < 8; ++u) {
                if (a0 + u < A) t[u] += z * W[static_cast<size_t>(a0 + u) * p + j];
              }
            }
          }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (a0 + u < A) tv[a0 + u] = t[u];
        }
      }
      for (int c = 0; c < cf; ++c) {
        double sc = CONST[c];
        for (int a = 0; a < A; ++a) sc += tv[a] * LIN[static_cast<size_t>(c) * A + a];
        DISC[static_cast<size_t>(i) * C + c] = sc;
        if (sc > best_s) {
          best_s = sc;
          best = c;
        }
      }
    }
    prm.pred_code[vrows[i]] = L.segc[best];
  }
}

// ---- general PLS ------------------------------------------------------------------

// One chunk of rows per workgroup: the threads form R = 256 / cols row lanes
// (at least one) of `cols` column lanes; row lane l sums rows l, l + R, ...
// of the chunk for its column, four loads in flight, and the R partials are
// added in row-lane order. Consecutive threads read consecutive columns of
// one row.
__global__ void __launch_bounds__(kT) pls_colstats_kernel(PlsColStatsParams prm) {
  __shared__ double part[kT];
  const int cols = prm.cols;
  const int r0 = blockIdx.x * prm.chunk_rows;
  const int r1 = min(prm.n, r0 + prm.chunk_rows);
  for (int c0 = 0; c0 < cols; c0 += kT) {
    const int width = min(kT, cols - c0);
    const int lanes = max(1, kT / width);
    const int rl = threadIdx.x / width, c = c0 + threadIdx.x % width;
    const bool active = rl < lanes && threadIdx.x < lanes * width;
    const double mu = prm.pass == 0 ? 0.0 : prm.mean[min(c, cols - 1)];
    double acc = 0.0;
    if (active) {
      int r = r0 + rl;
      for (; r + 3 * lanes < r1; r += 4 * lanes) {
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = prm.m[static_cast<size_t>(r + u * lanes) * cols + c];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double d = static_cast<double>(v[u]) - mu;
          acc += prm.pass == 0 ? static_cast<double>(v[u]) : d * d;
        }
      }
      for (; r < r1; r += lanes) {
        const double v = static_cast<double>(prm.m[static_cast<size_t>(r) * cols + c]);
        acc += prm.pass == 0 ? v : (v - mu) * (v - mu);
      }
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < width) {
      double t = 0.0;
      for (int l = 0; l < lanes; ++l) t += part[l * width + threadIdx.x];
      prm.partial[static_cast<size_t>(blockIdx.x) * cols + c0 + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

// Chunk partials of one column per workgroup: thread t sums chunks t, t + 256,
// ... in order, then a fixed tree over the threads.
__global__ void __launch_bounds__(kT) pls_colstats_reduce_kernel(PlsColStatsParams prm, int scale) {
  __shared__ double red[kT];
  const int c = blockIdx.x;
  double t = 0.0;
  for (int k = threadIdx.x; k < prm.chunks; k += kT) t += prm.partial[static_cast<size_t>(k) * prm.cols + c];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int w = kT / 2; w >= 1; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  t = red[0];
  if (prm.pass == 0) {
    prm.mean[c] = prm.center ? t / static_cast<double>(prm.n) : 0.0;
    if (prm.sum) prm.sum[c] = t;
  } else {
    double sd = 1.0;
    if (scale) {
      sd = sqrt((1.0 + t) / fmax(1.0, static_cast<double>(prm.n - 1)));
      sd = sd > 0.0 && isfinite(sd) ? sd : 1.0;
    }
    prm.sd[c] = sd;
  }
}

namespace {
typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int kCrossTilesPerWave = 8;
constexpr int kCrossTilesPerBlock = kCrossTilesPerWave * kWaves;
}  // namespace

// Z = Xs' [Xs | Yc] over one chunk of rows (blockIdx.x) for one group of up
// to 32 of the 16 x 16 output tiles (blockIdx.y). Rows go through LDS 16 at a
// time as doubles (standardised X columns, centred Y columns, zero padding);
// each thread holds PER = z16 / 16 elements of a 16-row block, and the next
// block's elements are loaded (without bounds checks: indices are clamped and
// the values masked) while the matrix cores work on the current one. Each
// wave accumulates up to eight tiles with v_mfma_f64_16x16x4f64 (A: lane l
// holds Z'[tile row l % 16][k = l / 16], B: Z[k = l / 16][tile column l % 16];
// D: register i of lane l holds row 4 i + l / 16, column l % 16).
template <int PER>
__global__ void __launch_bounds__(kT) pls_cross_kernel(PlsCrossParams prm) {
  extern __shared__ double zst[];  // [16][z16], then scale [z16] and shift [z16]
  const int p = prm.p, q = prm.q, zc = p + q, n = prm.n;
  const int p16 = (p + 15) / 16 * 16, z16 = (zc + 15) / 16 * 16;
  const int tiles_c = z16 / 16, tiles = (p16 / 16) * tiles_c;
  double* scale = zst + 16 * z16;
  double* shift = scale + z16;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int r0 = blockIdx.x * prm.chunk_rows;
  const int r1 = min(n, r0 + prm.chunk_rows);
  // Tiles of the block dealt round robin: wave w takes tiles w, w + 4, ...
  const int tile0 = blockIdx.y * kCrossTilesPerBlock + wave;
  // Column c of Z is (v - shift[c]) * scale[c]: X columns standardised
  // (1 / sd), Y columns centred, padding columns zero.
  for (int c = threadIdx.x; c < z16; c += kT) {
    scale[c] = c < p ? 1.0 / prm.x_sd[c] : (c < zc ? 1.0 : 0.0);
    shift[c] = c < p ? prm.x_mean[c] : (c < zc ? prm.y_mean[c - p] : 0.0);
  }
  f64x4 acc[kCrossTilesPerWave];
#pragma unroll
  for (int t = 0; t < kCrossTilesPerWave; ++t) acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
  const int kr = lane / 16, kc = lane % 16;
  // Element k of this thread in a block: e = threadIdx.x + k * 256 < 16 z16.
  auto fetch = [&](int rb, float (&xv)[PER], float (&yv)[PER]) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = min(static_cast<int>(threadIdx.x) + k * kT, 16 * z16 - 1);
      const int rr = e / z16, c = e - rr * z16;
      const int r = min(rb + rr, n - 1);
      xv[k] = prm.x[static_cast<size_t>(r) * p + min(c, p - 1)];
      yv[k] = q > 0 ? prm.y[static_cast<size_t>(r) * q + min(max(c - p, 0), q - 1)] : 0.0f;
    }
  };
  float xv[PER], yv[PER];
  fetch(r0, xv, yv);
  __syncthreads();
  for (int rb = r0; rb < r1; rb += 16) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = static_cast<int>(threadIdx.x) + k * kT;
      if (e < 16 * z16) {
        const int rr = e / z16, c = e - rr * z16;
        const double v = c < p ? static_cast<double>(xv[k]) : static_cast<double>(yv[k]);
        zst[e] = rb + rr < r1 ? (v - shift[c]) * scale[c] : 0.0;
      }
    }
    __syncthreads();
    if (rb + 16 < r1) fetch(rb + 16, xv, yv);  // in flight during the products
#pragma unroll
    for (int t = 0; t < kCrossTilesPerWave; ++t) {
      const int tile = tile0 + t * kWaves;
      if (tile < tiles) {
        const int tj = tile / tiles_c, tc = tile - tj * tiles_c;
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4) {
          const double a = zst[(k0 + kr) * z16 + tj * 16 + kc];
          const double b = zst[(k0 + kr) * z16 + tc * 16 + kc];
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  double* out = prm.partial + static_cast<size_t>(blockIdx.x) * p16 * z16;
#pragma unroll
  for (int t = 0; t < kCrossTilesPerWave; ++t) {
    const int tile = tile0 + t * kWaves;
    if (tile >= tiles) break;
    const int tj = tile / tiles_c, tc = tile - tj * tiles_c;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      out[static_cast<size_t>(tj * 16 + 4 * i + kr) * z16 + tc * 16 + kc] = acc[t][i];
    }
  }
}

// Chunk partials of 16 outputs per workgroup: chunk lane l (16 of them) sums
// chunks l, l + 16, ... in order for its output, and the 16 partials are
// added in chunk-lane order. Gram [p x p] and cross [q x p].
constexpr int kReduceOutputs = 16;
__global__ void __launch_bounds__(kT) pls_cross_reduce_kernel(PlsCrossParams prm) {
  __shared__ double part[kT];
  const int p = prm.p, zc = prm.p + prm.q;
  const int p16 = (p + 15) / 16 * 16, z16 = (zc + 15) / 16 * 16;
  const int lanes = kT / kReduceOutputs;
  const int o = threadIdx.x % kReduceOutputs, l = threadIdx.x / kReduceOutputs;
  const int e = blockIdx.x * kReduceOutputs + o;
  const int ec = min(e, p * zc - 1);
  const int j = ec / zc, c = ec - j * zc;
  double t = 0.0;
  for (int k = l; k < prm.chunks; k += lanes) t += prm.partial[(static_cast<size_t>(k) * p16 + j) * z16 + c];
  part[threadIdx.x] = t;
  __syncthreads();
  if (l != 0 || e >= p * zc) return;
  double v = 0.0;
  for (int q = 0; q < lanes; ++q) v += part[q * kReduceOutputs + o];
  if (c < p) prm.gram[static_cast<size_t>(j) * p + c] = v;
  else prm.cross[static_cast<size_t>(c - p) * p + j] = v;
}

// One wave: power-refresh SIMPLS of the cross-product (q rows in LDS) with
// the Gram matrix, then the response weights Q = S0' W and the coefficients
// W Q'. The weights and loadings are fitted component-major in `wscratch` and
// written out as [j * kPlsMaxComponents + a].
// STAGED: the Gram matrix, the weights and the loadings live in LDS (their
// accesses are LDS instructions); otherwise they stay in global memory.
template <bool STAGED>
__global__ void __launch_bounds__(kE) pls_fit_kernel(PlsFitParams prm) {
  extern __shared__ double fit_lds[];
  const int p = prm.p, q = prm.q, lane = threadIdx.x, rc = prm.rank_cap;
  double* S = fit_lds;                              // [q x p]
  double* wt = S + static_cast<size_t>(q) * p;     // [p]
  double* ld = wt + p;                              // [p]
  double* rv = ld + p;                              // [q]
  double* Ql = rv + q;                              // [q x rc] response weights
  double* Wl = Ql + static_cast<size_t>(q) * rc;    // STAGED: [rc x p] weights, component-major
  double* Vl = Wl + static_cast<size_t>(rc) * p;    // STAGED: [rc x p] loadings
  double* Gl = Vl + static_cast<size_t>(rc) * p;    // STAGED: [p x p]
  int* s_int = reinterpret_cast<int*>(STAGED ? Gl + static_cast<size_t>(p) * p : Wl);  // [4]
  for (int e = lane; e < q * p; e += kE) S[e] = prm.cross[e];
  if (STAGED) {
    for (int e = lane; e < p * p; e += kE) Gl[e] = prm.gram[e];
  }
  __syncthreads();
  double* W = STAGED ? Wl : prm.wscratch;                                    // [a * p + j]
  double* V = STAGED ? Vl : prm.wscratch + static_cast<size_t>(kA) * p;
  const double* G = STAGED ? Gl : prm.gram;
  const int fitted = simpls_wave(S, q, p, rc, G, prm.w0, W, V, wt, ld, rv, s_int);
  for (int e = lane; e < fitted * p; e += kE) {
    const int a = e / p, j = e - a * p;
    prm.weights[static_cast<size_t>(j) * kA + a] = W[e];
    prm.loadings[static_cast<size_t>(j) * kA + a] = V[e];
  }
  // Response weights of the scaled weights against the undeflated S
  // (plscv.cpp:905), lanes over columns, eight sums at a time.
  for (int e0 = 0; e0 < q * fitted; e0 += kBatch) {
    double v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      v[u] = 0.0;
      const int e = e0 + u;
      if (e < q * fitted) {
        const int c = e / fitted, a = e - c * fitted;
        for (int j = lane; j < p; j += kE) v[u] += prm.cross[static_cast<size_t>(c) * p + j] * W[static_cast<size_t>(a) * p + j];
      }
    }
    lane_totals(v);
    if (lane == 0) {
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int e = e0 + u;
        if (e < q * fitted) {
          const int c = e / fitted, a = e - c * fitted;
          Ql[c * rc + a] = v[u];
          prm.yweights[static_cast<size_t>(c) * kA + a] = v[u];
        }
      }
    }
  }
  __syncthreads();
  // Coefficients W Q'.
  for (int e = lane; e < p * q; e += kE) {
    const int j = e / q, c = e - j * q;
    double v = 0.0;
    for (int a = 0; a < fitted; ++a) v += W[static_cast<size_t>(a) * p + j] * Ql[c * rc + a];
    prm.coefficients[e] = v;
  }
  if (lane == 0) prm.fitted[0] = fitted;
}

// Scores and fitted values with up to 16 components, one row per thread: the
// weights, response weights, means and scales staged in LDS, zero padded to
// 16 components and a multiple of 16 columns (padding columns have mean 0,
// scale 1 and zero weights, so they add exact zeros), and each row's columns
// loaded sixteen at a time without bounds checks.
constexpr int kScoreA = 16;
__host__ __device__ inline int score_cols(int p) { return (p + 15) / 16 * 16; }
__host__ __device__ inline size_t score_lds_doubles(int p, int q) {
  return static_cast<size_t>(score_cols(p)) * (kScoreA + 2) + static_cast<size_t>(q) * (kScoreA + 1);
}

__global__ void __launch_bounds__(kT) pls_scores_staged_kernel(PlsScoreParams prm) {
  extern __shared__ double sc_lds[];
  const int p = prm.p, q = prm.q, A = prm.A, pc = score_cols(p);
  double* w = sc_lds;                               // [pc x 16]
  double* qw = w + static_cast<size_t>(pc) * kScoreA;  // [q x 16]
  double* mean = qw + static_cast<size_t>(q) * kScoreA;  // [pc]
  double* sd = mean + pc;                           // [pc]
  double* yb = sd + pc;                             // [q]
  for (int e = threadIdx.x; e < pc * kScoreA; e += kT) {
    const int j = e / kScoreA, u = e % kScoreA;
    w[e] = j < p && u < A ? prm.weights[static_cast<size_t>(j) * kA + u] : 0.0;
  }
  for (int e = threadIdx.x; e < q * kScoreA; e += kT) {
    const int c = e / kScoreA, u = e % kScoreA;
    qw[e] = u < A ? prm.yweights[static_cast<size_t>(c) * kA + u] : 0.0;
  }
  // sd holds the inverse scales here.
  for (int j = threadIdx.x; j < pc; j += kT) {
    mean[j] = j < p ? prm.x_mean[j] : 0.0;
    sd[j] = j < p ? 1.0 / prm.x_sd[j] : 1.0;
  }
  for (int c = threadIdx.x; c < q; c += kT) yb[c] = prm.y_mean[c];
  __syncthreads();
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= prm.n) return;
  const size_t row = static_cast<size_t>(i) * p;
  double t[kScoreA];
#pragma unroll
  for (int u = 0; u < kScoreA; ++u) t[u] = 0.0;
  for (int j0 = 0; j0 < pc; j0 += 16) {
    float xv[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) xv[v] = prm.x[row + min(j0 + v, p - 1)];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int j = j0 + v;
      const double z = (static_cast<double>(xv[v]) - mean[j]) * sd[j];
#pragma unroll
      for (int u = 0; u < kScoreA; ++u) t[u] += z * w[j * kScoreA + u];
    }
  }
#pragma unroll
  for (int u = 0; u < kScoreA; ++u) {
    if (u < A) prm.scores[static_cast<size_t>(i) * A + u] = static_cast<float>(t[u]);
  }
  for (int c = 0; c < q; ++c) {
    double v = yb[c];
#pragma unroll
    for (int u = 0; u < kScoreA; ++u) v += t[u] * qw[c * kScoreA + u];
    prm.fitted[static_cast<size_t>(i) * q + c] = static_cast<float>(v);
  }
}

__global__ void pls_scores_kernel(PlsScoreParams prm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= prm.n) return;
  const int p = prm.p, q = prm.q, A = prm.A;
  const float* xr = prm.x + static_cast<size_t>(i) * p;
  double* tr = prm.tbuf + static_cast<size_t>(i) * A;
  // Components in blocks of 16 register accumulators.
  for (int a0 = 0; a0 < A; a0 += 16) {
    double t[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) t[u] = 0.0;
    for (int j = 0; j < p; ++j) {
      const double xs = (static_cast<double>(xr[j]) - prm.x_mean[j]) / prm.x_sd[j];
      const double* w = prm.weights + static_cast<size_t>(j) * kA + a0;
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        if (a0 + u < A) t[u] += xs * w[u];
      }
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      if (a0 + u < A) {
        tr[a0 + u] = t[u];
        prm.scores[static_cast<size_t>(i) * A + a0 + u] = static_cast<float>(t[u]);
      }
    }
  }
  for (int c = 0; c < q; ++c) {
    double v = prm.y_mean[c];
    for (int a = 0; a < A; ++a) v += tr[a] * prm.yweights[static_cast<size_t>(c) * kA + a];
    prm.fitted[static_cast<size_t>(i) * q + c] = static_cast<float>(v);
  }
}

// ---- launchers ----------------------------------------------------------------

hipError_t launch_pls_fold_sums(const PlsFoldSumsParams& prm, hipStream_t s) {
  if (prm.folds == 0) return hipSuccess;
  if (prm.p < 1 || prm.p > kPlsMaxDims || prm.max_classes < 1 ||
      static_cast<size_t>(prm.max_classes) * prm.p > kPlsMaxSLds) {
    return hipErrorInvalidValue;
  }
  // Up to four waves, as many as keep the partials within 64 KiB (one wave
  // at the largest class x column tables).
  const size_t per_wave = sizeof(double) * static_cast<size_t>(prm.max_classes) * prm.p +
                          sizeof(int) * static_cast<size_t>(prm.max_classes);
  const int waves = static_cast<int>(std::max<size_t>(1, std::min<size_t>(4, (64 * 1024) / per_wave)));
  const size_t shmem = per_wave * waves;
  const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&pls_fold_sums_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(pls_fold_sums_kernel, dim3(prm.folds), dim3(kWave * waves), shmem, s, prm, waves);
  return hipGetLastError();
}

size_t pls_eval_lds_bytes(int p, int max_classes, int rank_cap, bool compact) {
  const size_t ints = 2 * static_cast<size_t>(max_classes) + 4;
  return sizeof(double) * eval_lds_doubles(p, max_classes, rank_cap, compact) + sizeof(int) * ints;
}

bool pls_eval_compact(int p, int max_classes, int rank_cap) {
  return p <= 64 && pls_eval_lds_bytes(p, max_classes, rank_cap, true) <= 64 * 1024;
}

hipError_t launch_pls_cache(const PlsCacheParams& prm, hipStream_t s) {
  if (prm.folds == 0) return hipSuccess;
  if (prm.p < 1 || prm.p > kPlsMaxDims || prm.chunks < 1) return hipErrorInvalidValue;
  const int z16 = (prm.p + 1 + 15) / 16 * 16;
  const int tiles = (z16 / 16) * (z16 / 16);
  const size_t shmem = sizeof(double) * 16 * static_cast<size_t>(z16);
  const dim3 grid(static_cast<unsigned>(prm.chunks),
                  static_cast<unsigned>((tiles + kMomentTilesPerBlock - 1) / kMomentTilesPerBlock));
  const int per = z16 / 16;  // elements per thread of a 16-row block
  if (per <= 5) {
    hipLaunchKernelGGL(pls_moments_kernel<5>, grid, dim3(kT), shmem, s, prm);
  } else if (per <= 9) {
    hipLaunchKernelGGL(pls_moments_kernel<9>, grid, dim3(kT), shmem, s, prm);
  } else {
    hipLaunchKernelGGL(pls_moments_kernel<17>, grid, dim3(kT), shmem, s, prm);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(pls_moments_reduce_kernel, dim3(prm.folds), dim3(kT), 0, s, prm);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(pls_cache_finalize_kernel, dim3(prm.folds), dim3(kT), 0, s, prm);
  return hipGetLastError();
}

hipError_t launch_pls_eval(const PlsEvalParams& prm, hipStream_t s) {
  if (prm.folds == 0) return hipSuccess;
  if (prm.p < 1 || prm.p > kPlsMaxDims || prm.rank_cap < 1 || prm.rank_cap > kPlsMaxComponents ||
      prm.rank_cap > prm.p || prm.requested < 1 || prm.max_classes < 1 ||
      static_cast<size_t>(prm.max_classes) * prm.p > kPlsMaxSLds) {
    return hipErrorInvalidValue;
  }
  const bool compact = pls_eval_compact(prm.p, prm.max_classes, prm.rank_cap);
  const size_t shmem = pls_eval_lds_bytes(prm.p, prm.max_classes, prm.rank_cap, compact);
  if (shmem > 160 * 1024) return hipErrorInvalidValue;
  const void* fn = compact ? reinterpret_cast<const void*>(&pls_eval_kernel<true>)
                           : reinterpret_cast<const void*>(&pls_eval_kernel<false>);
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem));
  if (e != hipSuccess) return e;
  if (compact) {
    hipLaunchKernelGGL(pls_eval_kernel<true>, dim3(prm.folds), dim3(kE), shmem, s, prm);
  } else {
    hipLaunchKernelGGL(pls_eval_kernel<false>, dim3(prm.folds), dim3(kE), shmem, s, prm);
  }
  return hipGetLastError();
}

hipError_t launch_pls_colstats(const PlsColStatsParams& prm, hipStream_t s) {
  if (prm.chunks == 0 || prm.cols == 0) return hipSuccess;
  hipLaunchKernelGGL(pls_colstats_kernel, dim3(prm.chunks), dim3(kT), 0, s, prm);
  return hipGetLastError();
}

hipError_t launch_pls_colstats_reduce(const PlsColStatsParams& prm, int scale, hipStream_t s) {
  if (prm.cols == 0) return hipSuccess;
  hipLaunchKernelGGL(pls_colstats_reduce_kernel, dim3(prm.cols), dim3(kT), 0, s, prm, scale);
  return hipGetLastError();
}

hipError_t launch_pls_cross(const PlsCrossParams& prm, hipStream_t s) {
  if (prm.chunks == 0) return hipSuccess;
  if (prm.p < 1 || prm.q < 0 || prm.chunk_rows % 16 != 0 || prm.n < 1) return hipErrorInvalidValue;
  const int p16 = (prm.p + 15) / 16 * 16, z16 = (prm.p + prm.q + 15) / 16 * 16;
  const int tiles = (p16 / 16) * (z16 / 16);
  const size_t shmem = sizeof(double) * 18 * static_cast<size_t>(z16);
  if (shmem > 64 * 1024) return hipErrorInvalidValue;
  const dim3 grid(static_cast<unsigned>(prm.chunks), static_cast<unsigned>((tiles + kCrossTilesPerBlock - 1) / kCrossTilesPerBlock));
  const int per = z16 / 16;  // elements per thread of a 16-row block
  if (per <= 5) {
    hipLaunchKernelGGL(pls_cross_kernel<5>, grid, dim3(kT), shmem, s, prm);
  } else if (per <= 9) {
    hipLaunchKernelGGL(pls_cross_kernel<9>, grid, dim3(kT), shmem, s, prm);
  } else if (per <= 17) {
    hipLaunchKernelGGL(pls_cross_kernel<17>, grid, dim3(kT), shmem, s, prm);
  } else if (per <= 33) {
    hipLaunchKernelGGL(pls_cross_kernel<33>, grid, dim3(kT), shmem, s, prm);
  } else {
    return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int total = prm.p * (prm.p + prm.q);
  hipLaunchKernelGGL(pls_cross_reduce_kernel, dim3((total + kReduceOutputs - 1) / kReduceOutputs), dim3(kT), 0, s,
                     prm);
  return hipGetLastError();
}

size_t pls_fit_lds_bytes(int p, int q, int rc, bool staged) {
  return sizeof(double) * (static_cast<size_t>(q) * p + 2 * static_cast<size_t>(p) + q + static_cast<size_t>(q) * rc +
                           (staged ? 2 * static_cast<size_t>(rc) * p + static_cast<size_t>(p) * p : 0)) +
         4 * sizeof(int);
}

hipError_t launch_pls_fit(const PlsFitParams& prm, hipStream_t s) {
  if (prm.p < 1 || prm.p > kPlsMaxDims || prm.q < 1 || prm.rank_cap < 1 || prm.rank_cap > kPlsMaxComponents ||
      static_cast<size_t>(prm.q) * prm.p > kPlsMaxSLds || prm.wscratch == nullptr) {
    return hipErrorInvalidValue;
  }
  // The Gram matrix, weights and loadings join the LDS state when all fits
  // 150 KiB.
  const bool staged = pls_fit_lds_bytes(prm.p, prm.q, prm.rank_cap, true) <= 150 * 1024;
  const size_t shmem = pls_fit_lds_bytes(prm.p, prm.q, prm.rank_cap, staged);
  if (shmem > 160 * 1024) return hipErrorInvalidValue;
  const void* fn = staged ? reinterpret_cast<const void*>(&pls_fit_kernel<true>)
                          : reinterpret_cast<const void*>(&pls_fit_kernel<false>);
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem));
  if (e != hipSuccess) return e;
  if (staged) {
    hipLaunchKernelGGL(pls_fit_kernel<true>, dim3(1), dim3(kE), shmem, s, prm);
  } else {
    hipLaunchKernelGGL(pls_fit_kernel<false>, dim3(1), dim3(kE), shmem, s, prm);
  }
  return hipGetLastError();
}

bool pls_scores_staged(int p, int q, int A) {
  return A <= kScoreA && sizeof(double) * score_lds_doubles(p, q) <= 48 * 1024;
}

hipError_t launch_pls_scores(const PlsScoreParams& prm, hipStream_t s) {
  if (prm.n == 0) return hipSuccess;
  if (prm.A < 1 || prm.A > kPlsMaxComponents) return hipErrorInvalidValue;
  if (pls_scores_staged(prm.p, prm.q, prm.A)) {
    const size_t shmem = sizeof(double) * score_lds_doubles(prm.p, prm.q);
    hipLaunchKernelGGL(pls_scores_staged_kernel, dim3((prm.n + kT - 1) / kT), dim3(kT), shmem, s, prm);
    return hipGetLastError();
  }
  if (prm.tbuf == nullptr) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pls_scores_kernel, dim3((prm.n + 127) / 128), dim3(128), 0, s, prm);
  return hipGetLastError();
}

}  // namespace kodama_hip
