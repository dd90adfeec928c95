// PLS-LDA cross-validation on the device (plscv.cpp:2293-2507 with the fold
// cache 593-698, SIMPLS 816-955 and the LDA 1379-1510), all in FP64.
//
// pls_cache_kernel (one workgroup per fold, once per data epoch): training mean
// and scale (the scale accumulator starts at 1, as train_center_scale_float
// does), column sums of the standardised training rows and their Gram matrix
// G = Xs'Xs, each summed in training-row order.
//
// pls_eval_kernel (one workgroup per fold, once per labelling; the PLS-LDA head,
// or with `da` the least-squares PLS-DA head of PLSDACV, plscv.cpp:1210-1296):
//   S      S[c] = sum_{rows of class c} xs - (n_c / n) xsum, from the class
//          segments of the training rows (host-sorted by class, rows ascending
//          inside a class), one thread per (class, column) in fixed order;
//   SIMPLS power refresh (two S S' products), normalisation through G, one
//          sequential Gram-Schmidt pass against earlier loadings, deflation;
//   LDA    class score means W'(S0[c] + (n_c/n) xsum) / n_c and T'T = W'GW
//          (equal to the sums over the training scores), pooled covariance,
//          ridge grid, Cholesky solve;
//   score  validation scores, discriminants, argmax (strict >, first wins).
// Reductions are fixed trees (wave butterflies, waves summed in order), so a
// result never depends on scheduling.
#include <hip/hip_runtime.h>

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

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

// Sum of v over the block; every thread gets the total. `red` holds kWaves
// doubles; two barriers make back-to-back calls safe.
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) t += red[w];
  __syncthreads();
  return t;
}

}  // namespace

// Power-refresh SIMPLS (plscv.cpp:816-955) of the cross-product S [m x p]
// (LDS, row c at S[c * p], deflated in place) with the Gram matrix G [p x p]:
// per component the previous weight (w0 for the first) takes two S'S power
// steps, is normalised (a failed first refresh restarts from the strongest
// row of S), scaled to unit score norm through G, its loading G w is
// Gram-Schmidt orthogonalised once against the earlier loadings and
// normalised, and S is deflated. Weights and loadings go to W and V
// ([j * kA + a], global). Returns the fitted rank (0: degenerate). wt, ld
// [p], rv [m] and red [kWaves] are LDS scratch, s_int[0] an LDS int. Every
// thread of the block calls it.
__device__ int simpls_power_fit(double* S, int m, int p, int rc, const double* G, const double* w0, double* W,
                                double* V, double* wt, double* ld, double* rv, double* red, int* s_int) {
  const int tid = threadIdx.x;
  const int cf = m;
  int fitted = 0;
  auto refresh = [&]() {
    for (int it = 0; it < 2; ++it) {
      for (int c = tid; c < cf; c += kT) {
        double a = 0.0;
        for (int j = 0; j < p; ++j) a += S[c * p + j] * wt[j];
        rv[c] = a;
      }
      __syncthreads();
      for (int j = tid; j < p; j += kT) {
        double a = 0.0;
        for (int c = 0; c < cf; ++c) a += S[c * p + j] * rv[c];
        wt[j] = a;
      }
      __syncthreads();
    }
  };
  auto norm_of = [&](const double* v) {
    double part = 0.0;
    for (int j = tid; j < p; j += kT) part += v[j] * v[j];
    return block_sum(part, red);
  };
  for (int a = 0; a < rc; ++a) {
    if (a == 0) {
      for (int j = tid; j < p; j += kT) wt[j] = w0[j];
      __syncthreads();
    }
    refresh();
    double ss = norm_of(wt);
    double nrm = isfinite(ss) && ss >= 0.0 ? sqrt(ss) : INFINITY;
    if (a == 0 && (!isfinite(nrm) || nrm <= kNormFloor)) {
      // The strongest column of S restarts the power iteration.
      if (tid == 0) {
        int strongest = -1;
        double best = 0.0;
        for (int c = 0; c < cf; ++c) {
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
        for (int j = tid; j < p; j += kT) wt[j] = S[strongest * p + j];
        __syncthreads();
        refresh();
        ss = norm_of(wt);
        nrm = isfinite(ss) && ss >= 0.0 ? sqrt(ss) : INFINITY;
      }
    }
    if (!isfinite(nrm) || nrm <= kNormFloor) break;
    for (int j = tid; j < p; j += kT) wt[j] /= nrm;
    __syncthreads();
    for (int i = tid; i < p; i += kT) {
      double s = 0.0;
      for (int j = 0; j < p; ++j) s += G[static_cast<size_t>(i) * p + j] * wt[j];
      ld[i] = s;
    }
    __syncthreads();
    double part = 0.0;
    for (int j = tid; j < p; j += kT) part += wt[j] * ld[j];
    const double tnorm = block_sum(part, red);
    if (!isfinite(tnorm) || tnorm <= kTinyFloor) break;
    const double inv = 1.0 / sqrt(tnorm);
    for (int j = tid; j < p; j += kT) {
      wt[j] *= inv;
      ld[j] *= inv;
    }
    __syncthreads();
    // One sequential Gram-Schmidt pass against the earlier loadings.
    for (int b = 0; b < a; ++b) {
      double pp = 0.0;
      for (int j = tid; j < p; j += kT) pp += V[static_cast<size_t>(j) * kA + b] * ld[j];
      const double proj = block_sum(pp, red);
      for (int j = tid; j < p; j += kT) ld[j] -= proj * V[static_cast<size_t>(j) * kA + b];
      __syncthreads();
    }
    const double lnn = norm_of(ld);
    const double ln = sqrt(fmax(0.0, lnn));
    if (!isfinite(ln) || ln <= kNormFloor) break;
    for (int j = tid; j < p; j += kT) {
      ld[j] /= ln;
      W[static_cast<size_t>(j) * kA + a] = wt[j];
      V[static_cast<size_t>(j) * kA + a] = ld[j];
    }
    __syncthreads();
    for (int c = tid; c < cf; c += kT) {
      double s = 0.0;
      for (int j = 0; j < p; ++j) s += ld[j] * S[c * p + j];
      rv[c] = s;
    }
    __syncthreads();
    for (int q = tid; q < cf * p; q += kT) {
      const int c = q / p, j = q - c * p;
      S[q] -= ld[j] * rv[c];
    }
    __syncthreads();
    fitted = a + 1;
  }
  return fitted;
}

// ---- fold cache ---------------------------------------------------------------

__global__ void __launch_bounds__(kT) pls_cache_kernel(PlsCacheParams prm) {
  constexpr int kTile = 16, kChunk = 32;
  __shared__ double stage[kChunk][2 * kTile + 1];
  const int f = blockIdx.x;
  const int p = prm.p;
  const int t0 = prm.train.offsets[f];
  const int nt = prm.train.offsets[f + 1] - t0;
  const int* rows = prm.train.rows + t0;
  double* mean = prm.mean + static_cast<size_t>(f) * p;
  double* sd = prm.sd + static_cast<size_t>(f) * p;
  double* xsum = prm.xsum + static_cast<size_t>(f) * p;
  double* gram = prm.gram + static_cast<size_t>(f) * p * p;
  for (int j = threadIdx.x; j < p; j += kT) {
    double m = 0.0;
    if (prm.center) {
      for (int i = 0; i < nt; ++i) m += static_cast<double>(prm.x[static_cast<size_t>(rows[i]) * p + j]);
      m *= nt > 0 ? 1.0 / static_cast<double>(nt) : 0.0;
    }
    double s = 1.0;
    if (prm.scale) {
      double acc = 1.0;
      for (int i = 0; i < nt; ++i) {
        const double d = static_cast<double>(prm.x[static_cast<size_t>(rows[i]) * p + j]) - m;
        acc += d * d;
      }
      s = sqrt(acc / fmax(1.0, static_cast<double>(nt - 1)));
      s = s > 0.0 && isfinite(s) ? s : 1.0;
    }
    double xs = 0.0;
    for (int i = 0; i < nt; ++i) xs += (static_cast<double>(prm.x[static_cast<size_t>(rows[i]) * p + j]) - m) / s;
    mean[j] = m;
    sd[j] = s;
    xsum[j] = xs;
  }
  __syncthreads();
  // Gram in 16 x 16 tiles of the lower triangle; rows staged 32 at a time.
  const int tiles = (p + kTile - 1) / kTile;
  const int u = threadIdx.x / kTile, v = threadIdx.x % kTile;
  for (int ta = 0; ta < tiles; ++ta) {
    for (int tb = 0; tb <= ta; ++tb) {
      const int a = ta * kTile + u, b = tb * kTile + v;
      double acc = 0.0;
      for (int c0 = 0; c0 < nt; c0 += kChunk) {
        // Stage the two 16-column blocks of up to 32 standardised rows.
        for (int e = threadIdx.x; e < kChunk * 2 * kTile; e += kT) {
          const int r = e / (2 * kTile), q = e % (2 * kTile);
          const int col = q < kTile ? ta * kTile + q : tb * kTile + (q - kTile);
          double val = 0.0;
          if (c0 + r < nt && col < p) {
            val = (static_cast<double>(prm.x[static_cast<size_t>(rows[c0 + r]) * p + col]) - mean[col]) / sd[col];
          }
          stage[r][q] = val;
        }
        __syncthreads();
        const int len = min(kChunk, nt - c0);
        for (int r = 0; r < len; ++r) acc += stage[r][u] * stage[r][kTile + v];
        __syncthreads();
      }
      if (a < p && b < p && b <= a) {
        gram[static_cast<size_t>(a) * p + b] = acc;
        gram[static_cast<size_t>(b) * p + a] = acc;
      }
    }
  }
}

// ---- evaluation -------------------------------------------------------------

__global__ void __launch_bounds__(kT) pls_eval_kernel(PlsEvalParams prm) {
  extern __shared__ double pls_lds[];
  __shared__ int s_int[4];
  const int f = blockIdx.x;
  const int p = prm.p;
  const int tid = threadIdx.x;
  const int t0 = prm.train.offsets[f];
  const int nt = prm.train.offsets[f + 1] - t0;
  const int cf = prm.seg_count[f];
  const int* seg_code = prm.seg_code + static_cast<size_t>(f) * prm.max_classes;
  const int* seg_start = prm.seg_start + static_cast<size_t>(f) * (prm.max_classes + 1);  // cf + 1 entries
  const int v0 = prm.val.offsets[f];
  const int nv = prm.val.offsets[f + 1] - v0;
  const int* vrows = prm.val.rows + v0;
  const int cmax = prm.max_classes;
  const double* mean = prm.mean + static_cast<size_t>(f) * p;
  const double* sd = prm.sd + static_cast<size_t>(f) * p;
  const double* xsum = prm.xsum + static_cast<size_t>(f) * p;
  const double* G = prm.gram + static_cast<size_t>(f) * p * p;
  double* W = prm.weights + static_cast<size_t>(f) * p * kA;   // [j * kA + a]
  double* V = prm.loadings + static_cast<size_t>(f) * p * kA;
  double* S0 = prm.s0 + static_cast<size_t>(f) * cmax * p;  // undeflated S, [c * p + j]
  double* GW = prm.gw + static_cast<size_t>(f) * p * kA;     // [j * kA + a]
  double* LIN = prm.linear + static_cast<size_t>(f) * cmax * kA;  // [c * kA + a]
  double* CONST = prm.constant + static_cast<size_t>(f) * cmax;
  double* TV = prm.val_scores + static_cast<size_t>(v0) * kA;       // [i * kA + a]
  double* DISC = prm.discriminants + static_cast<size_t>(v0) * cmax;  // [i * cmax + c]

  double* S = pls_lds;                                  // [cmax * p], later LDA rows
  double* wt = S + static_cast<size_t>(cmax) * p;       // [p]
  double* ld = wt + p;                                  // [p]
  double* rv = ld + p;                                  // [cmax]
  double* red = rv + cmax;                              // [kWaves]
  double* mat = red + kWaves;                           // [rc * rc] pooled covariance, then its factor
  double* tt = mat + static_cast<size_t>(prm.rank_cap) * prm.rank_cap;  // [rc * rc] W'GW
  double* mu = tt + static_cast<size_t>(prm.rank_cap) * prm.rank_cap;   // [cmax * rc] class score means

  if (cf <= 1) {
    // One class in the training fold: every validation row gets it.
    const int code = cf == 1 ? seg_code[0] : 0;
    for (int i = tid; i < nv; i += kT) prm.pred_code[vrows[i]] = code;
    if (tid == 0) {
      prm.components[f] = 1;
      prm.rank[f] = 0;
    }
    return;
  }

  // -- S = X'(Y - ybar) -----------------------------------------------------------
  const double inv_nt = 1.0 / static_cast<double>(nt);
  for (int q = tid; q < cf * p; q += kT) {
    const int c = q / p, j = q - c * p;
    const int* r = prm.train.rows;
    double acc = 0.0;
    for (int i = seg_start[c]; i < seg_start[c + 1]; ++i) {
      acc += (static_cast<double>(prm.x[static_cast<size_t>(r[i]) * p + j]) - mean[j]) / sd[j];
    }
    const double ym = static_cast<double>(seg_start[c + 1] - seg_start[c]) * inv_nt;
    const double s = acc - ym * xsum[j];
    S[q] = s;
    S0[q] = s;
  }
  __syncthreads();

  // -- power-refresh SIMPLS ----------------------------------------------------
  const int rc = max(1, min(prm.rank_cap, max(1, nt - 1)));
  const int fitted = simpls_power_fit(S, cf, p, rc, G, prm.w0, W, V, wt, ld, rv, red, s_int);

  if (fitted < 1) {
    // Degenerate fit: the training majority (the smaller code on ties).
    const int code = [&] {
      int best = seg_code[0], best_n = -1;
      for (int c = 0; c < cf; ++c) {
        const int cnt = seg_start[c + 1] - seg_start[c];
        if (cnt > best_n) {
          best = seg_code[c];
          best_n = cnt;
        }
      }
      return best;
    }();
    for (int i = tid; i < nv; i += kT) prm.pred_code[vrows[i]] = code;
    if (tid == 0) {
      prm.components[f] = 1;
      prm.rank[f] = 0;
    }
    return;
  }
  const int A = min(prm.requested, fitted);
  if (tid == 0) {
    prm.components[f] = A;
    prm.rank[f] = fitted;
  }

  // -- class statistics ------------------------------------------------------------
  // GW = G W [p x A] (global scratch, [j * kA + a]).
  for (int q = tid; q < p * A; q += kT) {
    const int i = q / A, a = q - i * A;
    double s = 0.0;
    for (int j = 0; j < p; ++j) s += G[static_cast<size_t>(i) * p + j] * W[static_cast<size_t>(j) * kA + a];
    GW[static_cast<size_t>(i) * kA + a] = s;
  }
  if (prm.da) {
    // PLS-DA: T'(Y - ybar) = W'S0 (class score sums minus ybar times the
    // score totals), kept in mu as [c * A + a].
    for (int q = tid; q < cf * A; q += kT) {
      const int c = q / A, a = q - c * A;
      double s = 0.0;
      for (int j = 0; j < p; ++j) s += W[static_cast<size_t>(j) * kA + a] * S0[c * p + j];
      mu[c * A + a] = s;
    }
  } else {
    // Class score means W'(S0[c] + (n_c / n) xsum) / n_c.
    for (int q = tid; q < cf * A; q += kT) {
      const int c = q / A, a = q - c * A;
      const int cnt = seg_start[c + 1] - seg_start[c];
      const double ym = static_cast<double>(cnt) * inv_nt;
      double s = 0.0;
      for (int j = 0; j < p; ++j) s += W[static_cast<size_t>(j) * kA + a] * (S0[c * p + j] + ym * xsum[j]);
      mu[c * A + a] = cnt > 0 ? s * (1.0 / static_cast<double>(cnt)) : 0.0;
    }
  }
  __syncthreads();
  // T'T = W' G W.
  for (int q = tid; q < A * A; q += kT) {
    const int r = q / A, c2 = q - r * A;
    double s = 0.0;
    for (int j = 0; j < p; ++j) s += W[static_cast<size_t>(j) * kA + r] * GW[static_cast<size_t>(j) * kA + c2];
    tt[q] = s;
  }
  __syncthreads();

  if (prm.da) {
    // -- PLS-DA regression (plscv.cpp:1220-1296): Gauss-Jordan elimination with
    // partial pivoting on (T'T + 1e-9 I) B = T'(Y - ybar), columns whose pivot
    // is below 1e-12 skipped; one wave, lanes over matrix entries. mat holds
    // T'T, mu the right-hand sides (row a of column c at mu[c * A + a]).
    if (tid < kWave) {
      const int lane = tid;
      for (int q = lane; q < A * A; q += kWave) {
        const int r = q / A, c2 = q - r * A;
        const int lo = min(r, c2), hi = max(r, c2);
        mat[q] = tt[hi * A + lo] + (r == c2 ? static_cast<double>(1.0e-9f) : 0.0);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
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
          for (int j = lane; j < A; j += kWave) {
            const double t = mat[col * A + j];
            mat[col * A + j] = mat[piv * A + j];
            mat[piv * A + j] = t;
          }
          for (int c = lane; c < cf; c += kWave) {
            const double t = mu[c * A + col];
            mu[c * A + col] = mu[c * A + piv];
            mu[c * A + piv] = t;
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
        }
        const double div = mat[col * A + col];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (int j = col + lane; j < A; j += kWave) mat[col * A + j] /= div;
        for (int c = lane; c < cf; c += kWave) mu[c * A + col] /= div;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // Row factors read before any row changes (tt is free once copied).
        for (int r = lane; r < A; r += kWave) tt[r] = mat[r * A + col];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (int q = lane; q < A * (A + cf); q += kWave) {
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
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      // Discriminant of class c: ybar_c + t' B[:, c].
      for (int c = lane; c < cf; c += kWave) {
        for (int a = 0; a < A; ++a) LIN[static_cast<size_t>(c) * kA + a] = mu[c * A + a];
        CONST[c] = static_cast<double>(seg_start[c + 1] - seg_start[c]) * inv_nt;
      }
    }
    __syncthreads();
  } else {
    // -- LDA: pooled covariance, ridge grid + Cholesky (one wave; lanes over
    // rows / right-hand sides) ---------------------------------------------------
    const double df = static_cast<double>(max(1, nt - cf));
    for (int q = tid; q < A * A; q += kT) {
      const int r = q / A, c2 = q - r * A;
      const int lo = min(r, c2), hi = max(r, c2);
      double between = 0.0;
      for (int c = 0; c < cf; ++c) {
        between += static_cast<double>(seg_start[c + 1] - seg_start[c]) * mu[c * A + hi] * mu[c * A + lo];
      }
      // The lower triangle is the one computed (symmetric use of T'T).
      mat[q] = (tt[hi * A + lo] - between) / df;
    }
    __syncthreads();
    // The factor is written over tt (lower triangle); the solution of each
    // class column goes to LIN[c * kA + a].
    if (tid < kWave) {
      const int lane = tid;
      double trace = 0.0;
      for (int a = 0; a < A; ++a) trace += mat[a * A + a];
      const double rs = isfinite(trace) && trace > 0.0 ? trace / static_cast<double>(max(1, A)) : 1.0;
      const float grid[6] = {1e-8f, 1e-6f, 1e-5f, 1e-4f, 1e-3f, 1e-2f};
      bool ok = false;
      // tt is free now: it holds the factor while `mat` keeps the covariance.
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
          if (lane == 0) lo[c * A + c] = d;
          for (int r = c + 1 + lane; r < A; r += kWave) {
            double w = mat[r * A + c];
            for (int k = 0; k < c; ++k) w -= lo[r * A + k] * lo[c * A + k];
            lo[r * A + c] = w / d;
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
        }
        ok = good;
      }
      if (lane == 0) s_int[1] = ok ? 1 : 0;
      if (ok) {
        for (int c = lane; c < cf; c += kWave) {
          double* sol = LIN + static_cast<size_t>(c) * kA;
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
          const int cnt = seg_start[c + 1] - seg_start[c];
          const double prior = fmax(static_cast<double>(cnt) / static_cast<double>(max(1, nt)),
                                    static_cast<double>(FLT_MIN));
          CONST[c] = -0.5 * dm + log(prior);
        }
      }
    }
    __syncthreads();
    if (!s_int[1]) {
      // The host reports the failed factorisation (components = -1).
      if (tid == 0) prm.components[f] = -1;
      return;
    }
  }

  // -- validation scores and discriminants ----------------------------------------
  for (int q = tid; q < nv * A; q += kT) {
    const int i = q / A, a = q - i * A;
    const size_t row = static_cast<size_t>(vrows[i]) * p;
    double s = 0.0;
    for (int j = 0; j < p; ++j) {
      s += ((static_cast<double>(prm.x[row + j]) - mean[j]) / sd[j]) * W[static_cast<size_t>(j) * kA + a];
    }
    TV[static_cast<size_t>(i) * kA + a] = s;
  }
  __syncthreads();
  for (int q = tid; q < nv * cf; q += kT) {
    const int i = q / cf, c = q - i * cf;
    double sc = CONST[c];
    for (int a = 0; a < A; ++a) sc += TV[static_cast<size_t>(i) * kA + a] * LIN[static_cast<size_t>(c) * kA + a];
    DISC[static_cast<size_t>(i) * cmax + c] = sc;
  }
  __syncthreads();
  for (int i = tid; i < nv; i += kT) {
    int best = 0;
    double best_s = -INFINITY;
    for (int c = 0; c < cf; ++c) {
      const double sc = DISC[static_cast<size_t>(i) * cmax + c];
      if (sc > best_s) {
        best_s = sc;
        best = c;
      }
    }
    prm.pred_code[vrows[i]] = seg_code[best];
  }
}

// ---- class grouping -----------------------------------------------------------

// One workgroup per fold: class counts, the segment table (classes ascending),
// then one wave scatters the rows in their original order -- per 64-row chunk,
// the lanes of one class take consecutive slots in lane order -- so every class
// segment lists its rows ascending.
__global__ void __launch_bounds__(kT) pls_group_kernel(PlsGroupParams prm) {
  extern __shared__ int grp_lds[];
  const int C = prm.max_classes;
  int* cnt = grp_lds;       // [C]
  int* base = grp_lds + C;  // [C]
  const int f = blockIdx.x;
  const int t0 = prm.train.offsets[f];
  const int nt = prm.train.offsets[f + 1] - t0;
  const int* rows = prm.train.rows + t0;
  for (int c = threadIdx.x; c < C; c += kT) cnt[c] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < nt; i += kT) atomicAdd(&cnt[prm.codes[rows[i]]], 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0, acc = 0;
    int* code = prm.seg_code + static_cast<size_t>(f) * C;
    int* start = prm.seg_start + static_cast<size_t>(f) * (C + 1);
    for (int c = 0; c < C; ++c) {
      if (cnt[c] == 0) continue;
      code[s] = c;
      start[s] = t0 + acc;
      base[c] = t0 + acc;
      acc += cnt[c];
      ++s;
    }
    start[s] = t0 + acc;
    prm.seg_count[f] = s;
  }
  __syncthreads();
  if (threadIdx.x < kWave) {
    const int lane = threadIdx.x;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int c0 = 0; c0 < nt; c0 += kWave) {
      const int i = c0 + lane;
      const int row = i < nt ? rows[i] : -1;
      const int code = i < nt ? prm.codes[row] : -1;
      uint64_t remaining = __ballot(i < nt);
      while (remaining != 0ull) {
        const int leader = __ffsll(static_cast<unsigned long long>(remaining)) - 1;
        const int lc = __shfl(code, leader);
        const uint64_t mask = __ballot(code == lc) & remaining;
        const int b = base[lc];
        if ((mask >> lane) & 1ull) prm.sorted[b + __popcll(mask & below)] = row;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (lane == leader) base[lc] = b + __popcll(mask);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        remaining &= ~mask;
      }
    }
  }
}

// ---- general PLS ------------------------------------------------------------------

__global__ void __launch_bounds__(kT) pls_colstats_kernel(PlsColStatsParams prm) {
  const int r0 = blockIdx.x * prm.chunk_rows;
  const int r1 = min(prm.n, r0 + prm.chunk_rows);
  for (int c = threadIdx.x; c < prm.cols; c += kT) {
    double acc = 0.0;
    if (prm.pass == 0) {
      for (int r = r0; r < r1; ++r) acc += static_cast<double>(prm.m[static_cast<size_t>(r) * prm.cols + c]);
    } else {
      const double mu = prm.mean[c];
      for (int r = r0; r < r1; ++r) {
        const double d = static_cast<double>(prm.m[static_cast<size_t>(r) * prm.cols + c]) - mu;
        acc += d * d;
      }
    }
    prm.partial[static_cast<size_t>(blockIdx.x) * prm.cols + c] = acc;
  }
}

__global__ void pls_colstats_reduce_kernel(PlsColStatsParams prm, int scale) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= prm.cols) return;
  double t = 0.0;
  for (int k = 0; k < prm.chunks; ++k) t += prm.partial[static_cast<size_t>(k) * prm.cols + c];
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
constexpr int kCrossTilesPerWave = 4;
constexpr int kCrossTilesPerBlock = kCrossTilesPerWave * kWaves;
}  // namespace

// Z = Xs' [Xs | Yc] over one chunk of rows (blockIdx.x) for one group of 16 x
// 16 output tiles (blockIdx.y). Rows are staged 16 at a time in LDS as
// doubles (standardised X columns, centred Y columns, zero padding); each
// wave accumulates up to four tiles with v_mfma_f64_16x16x4f64 (A: lane l
// holds Z'[tile row l % 16][k = l / 16], B: Z[k = l / 16][tile column l % 16];
// D: register i of lane l holds row 4 i + l / 16, column l % 16).
__global__ void __launch_bounds__(kT) pls_cross_kernel(PlsCrossParams prm) {
  extern __shared__ double zst[];  // [16][z16]
  const int p = prm.p, q = prm.q, zc = p + q;
  const int p16 = (p + 15) / 16 * 16, z16 = (zc + 15) / 16 * 16;
  const int tiles_c = z16 / 16, tiles = (p16 / 16) * tiles_c;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int r0 = blockIdx.x * prm.chunk_rows;
  const int r1 = min(prm.n, r0 + prm.chunk_rows);
  const int tile0 = blockIdx.y * kCrossTilesPerBlock + wave * kCrossTilesPerWave;
  f64x4 acc[kCrossTilesPerWave];
#pragma unroll
  for (int t = 0; t < kCrossTilesPerWave; ++t) acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
  const int kr = lane / 16, kc = lane % 16;
  for (int rb = r0; rb < r1; rb += 16) {
    for (int e = threadIdx.x; e < 16 * z16; e += kT) {
      const int rr = e / z16, c = e - rr * z16, r = rb + rr;
      double v = 0.0;
      if (r < r1) {
        if (c < p) {
          v = (static_cast<double>(prm.x[static_cast<size_t>(r) * p + c]) - prm.x_mean[c]) / prm.x_sd[c];
        } else if (c < zc) {
          v = static_cast<double>(prm.y[static_cast<size_t>(r) * q + (c - p)]) - prm.y_mean[c - p];
        }
      }
      zst[e] = v;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kCrossTilesPerWave; ++t) {
      const int tile = tile0 + t;
      if (tile >= tiles) break;
      const int tj = tile / tiles_c, tc = tile - tj * tiles_c;
#pragma unroll
      for (int k0 = 0; k0 < 16; k0 += 4) {
        const double a = zst[(k0 + kr) * z16 + tj * 16 + kc];
        const double b = zst[(k0 + kr) * z16 + tc * 16 + kc];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  double* out = prm.partial + static_cast<size_t>(blockIdx.x) * p16 * z16;
#pragma unroll
  for (int t = 0; t < kCrossTilesPerWave; ++t) {
    const int tile = tile0 + t;
    if (tile >= tiles) break;
    const int tj = tile / tiles_c, tc = tile - tj * tiles_c;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      out[static_cast<size_t>(tj * 16 + 4 * i + kr) * z16 + tc * 16 + kc] = acc[t][i];
    }
  }
}

// Chunk partials summed in chunk order: Gram [p x p] and cross [q x p].
__global__ void pls_cross_reduce_kernel(PlsCrossParams prm) {
  const int p = prm.p, zc = prm.p + prm.q;
  const int p16 = (p + 15) / 16 * 16, z16 = (zc + 15) / 16 * 16;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= p * zc) return;
  const int j = e / zc, c = e - j * zc;
  double t = 0.0;
  for (int k = 0; k < prm.chunks; ++k) t += prm.partial[(static_cast<size_t>(k) * p16 + j) * z16 + c];
  if (c < p) prm.gram[static_cast<size_t>(j) * p + c] = t;
  else prm.cross[static_cast<size_t>(c - p) * p + j] = t;
}

__global__ void __launch_bounds__(kT) pls_fit_kernel(PlsFitParams prm) {
  extern __shared__ double fit_lds[];
  __shared__ int s_int[4];
  const int p = prm.p, q = prm.q, tid = threadIdx.x;
  double* S = fit_lds;                            // [q x p]
  double* wt = S + static_cast<size_t>(q) * p;   // [p]
  double* ld = wt + p;                            // [p]
  double* rv = ld + p;                            // [q]
  double* red = rv + q;                           // [kWaves]
  for (int e = tid; e < q * p; e += kT) S[e] = prm.cross[e];
  __syncthreads();
  const int fitted = simpls_power_fit(S, q, p, prm.rank_cap, prm.gram, prm.w0, prm.weights, prm.loadings, wt, ld,
                                      rv, red, s_int);
  // Response weights of the scaled weights against the undeflated S
  // (plscv.cpp:905), then the coefficients W Q'.
  for (int e = tid; e < q * fitted; e += kT) {
    const int c = e / fitted, a = e - c * fitted;
    double v = 0.0;
    for (int j = 0; j < p; ++j) v += prm.cross[static_cast<size_t>(c) * p + j] * prm.weights[static_cast<size_t>(j) * kA + a];
    prm.yweights[static_cast<size_t>(c) * kA + a] = v;
  }
  __syncthreads();
  for (int e = tid; e < p * q; e += kT) {
    const int j = e / q, c = e - j * q;
    double v = 0.0;
    for (int a = 0; a < fitted; ++a) {
      v += prm.weights[static_cast<size_t>(j) * kA + a] * prm.yweights[static_cast<size_t>(c) * kA + a];
    }
    prm.coefficients[e] = v;
  }
  if (tid == 0) prm.fitted[0] = fitted;
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

hipError_t launch_pls_group(const PlsGroupParams& prm, hipStream_t s) {
  if (prm.folds == 0) return hipSuccess;
  if (prm.max_classes < 1) return hipErrorInvalidValue;
  const size_t shmem = 2 * sizeof(int) * static_cast<size_t>(prm.max_classes);
  if (shmem > 64 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pls_group_kernel, dim3(prm.folds), dim3(kT), shmem, s, prm);
  return hipGetLastError();
}

size_t pls_eval_lds_bytes(int p, int max_classes) {
  const int rc = p < kPlsMaxComponents ? p : kPlsMaxComponents;
  return sizeof(double) * (static_cast<size_t>(max_classes) * p + 2 * static_cast<size_t>(p) + max_classes +
                           kWaves + 2 * static_cast<size_t>(rc) * rc + static_cast<size_t>(max_classes) * rc);
}

hipError_t launch_pls_cache(const PlsCacheParams& prm, hipStream_t s) {
  if (prm.folds == 0) return hipSuccess;
  if (prm.p < 1 || prm.p > kPlsMaxDims) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pls_cache_kernel, dim3(prm.folds), dim3(kT), 0, s, prm);
  return hipGetLastError();
}

hipError_t launch_pls_eval(const PlsEvalParams& prm, hipStream_t s) {
  if (prm.folds == 0) return hipSuccess;
  if (prm.p < 1 || prm.p > kPlsMaxDims || prm.rank_cap < 1 || prm.rank_cap > kPlsMaxComponents ||
      prm.rank_cap > prm.p || prm.requested < 1 || static_cast<size_t>(prm.max_classes) * prm.p > kPlsMaxSLds) {
    return hipErrorInvalidValue;
  }
  const size_t shmem = pls_eval_lds_bytes(prm.p, prm.max_classes);
  if (shmem > 160 * 1024) return hipErrorInvalidValue;
  const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&pls_eval_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(pls_eval_kernel, dim3(prm.folds), dim3(kT), shmem, s, prm);
  return hipGetLastError();
}

hipError_t launch_pls_colstats(const PlsColStatsParams& prm, hipStream_t s) {
  if (prm.chunks == 0 || prm.cols == 0) return hipSuccess;
  hipLaunchKernelGGL(pls_colstats_kernel, dim3(prm.chunks), dim3(kT), 0, s, prm);
  return hipGetLastError();
}

hipError_t launch_pls_colstats_reduce(const PlsColStatsParams& prm, int scale, hipStream_t s) {
  if (prm.cols == 0) return hipSuccess;
  hipLaunchKernelGGL(pls_colstats_reduce_kernel, dim3((prm.cols + 127) / 128), dim3(128), 0, s, prm, scale);
  return hipGetLastError();
}

hipError_t launch_pls_cross(const PlsCrossParams& prm, hipStream_t s) {
  if (prm.chunks == 0) return hipSuccess;
  if (prm.p < 1 || prm.q < 0 || prm.chunk_rows % 16 != 0) return hipErrorInvalidValue;
  const int p16 = (prm.p + 15) / 16 * 16, z16 = (prm.p + prm.q + 15) / 16 * 16;
  const int tiles = (p16 / 16) * (z16 / 16);
  const size_t shmem = sizeof(double) * 16 * static_cast<size_t>(z16);
  if (shmem > 64 * 1024) return hipErrorInvalidValue;
  const dim3 grid(static_cast<unsigned>(prm.chunks), static_cast<unsigned>((tiles + kCrossTilesPerBlock - 1) / kCrossTilesPerBlock));
  hipLaunchKernelGGL(pls_cross_kernel, grid, dim3(kT), shmem, s, prm);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int total = prm.p * (prm.p + prm.q);
  hipLaunchKernelGGL(pls_cross_reduce_kernel, dim3((total + 255) / 256), dim3(256), 0, s, prm);
  return hipGetLastError();
}

size_t pls_fit_lds_bytes(int p, int q) {
  return sizeof(double) * (static_cast<size_t>(q) * p + 2 * static_cast<size_t>(p) + q + kWaves);
}

hipError_t launch_pls_fit(const PlsFitParams& prm, hipStream_t s) {
  if (prm.p < 1 || prm.p > kPlsMaxDims || prm.q < 1 || prm.rank_cap < 1 || prm.rank_cap > kPlsMaxComponents ||
      static_cast<size_t>(prm.q) * prm.p > kPlsMaxSLds) {
    return hipErrorInvalidValue;
  }
  const size_t shmem = pls_fit_lds_bytes(prm.p, prm.q);
  const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&pls_fit_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(pls_fit_kernel, dim3(1), dim3(kT), shmem, s, prm);
  return hipGetLastError();
}

hipError_t launch_pls_scores(const PlsScoreParams& prm, hipStream_t s) {
  if (prm.n == 0) return hipSuccess;
  if (prm.A < 1 || prm.A > kPlsMaxComponents) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pls_scores_kernel, dim3((prm.n + 127) / 128), dim3(128), 0, s, prm);
  return hipGetLastError();
}

}  // namespace kodama_hip
