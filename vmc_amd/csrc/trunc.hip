// Zip-up truncation step on gfx950: one workgroup (512 threads, 8 wave64s)
// per sample computes the top-k left singular subspace of the theta matrix
// mat (m x n, m, n <= 128) entirely in FP64:
//
//   G = mat mat^T                       (Gram, register-resident 4x8 blocks)
//   G = Q T Q^T                         (Householder tridiagonalisation, LAPACK
//                                        dsytd2-style reflectors kept in registers)
//   lambda_{m-k..m-1}(T)                (Sturm-count multisection, 16 lanes per
//                                        eigenvalue, 17 sub-intervals per step)
//   z_l = twisted-factorisation vector  (Dhillon/Parlett twisted LDL^T, O(m))
//   U = Q Z, modified Gram-Schmidt      (blocked reflector application)
//   carry = U^T mat / |U^T mat|_F        (log of the norm added to the boundary scale)
//
// This replaces `jnp.linalg.svd(mat)` + keep-k of the reference zip-up
// (src/vmc/peps/common/strategy.py:142-148). The truncated boundary state
// U U^T theta equals the reference's U_k S_k V_k^H whenever sigma_k > sigma_{k+1};
// the gauge of U (signs, rotations inside degenerate blocks) may differ, which
// leaves every contracted amplitude and environment unchanged.
#include "common.hpp"

#include <cfloat>

namespace vmc {
namespace {

constexpr int NT = 512;    // threads per workgroup
constexpr int MMAX = 128;  // max rows/cols of mat
constexpr int KMAX = 32;   // max kept vectors
constexpr int LDY = 129;   // padded column stride of the eigenvector tiles
constexpr int STG = 17;    // padded row stride of the Gram staging tile
constexpr int NCH = 16;    // mat columns staged per Gram pass

__device__ inline double guard_pivot(double d, double pivmin) {
  if (fabs(d) < pivmin) d = (d < 0.0) ? -pivmin : pivmin;
  return d;
}

template <typename T>
__global__ __launch_bounds__(NT) void k_trunc(TruncArgs<T> a) {
  __shared__ double sBig[KMAX * LDY];  // Gram staging | Dm of twisted solves | reflector block
  __shared__ double sY[KMAX * LDY];    // Dp / eigenvectors of T / orthonormal U (column-major)
  __shared__ double sX[MMAX], sV[MMAX], sP[MMAX], sD[MMAX], sE[MMAX], sE2[MMAX], sTau[MMAX];
  __shared__ double sLam[KMAX];
  __shared__ double sRed[3 * 16];

  const int b = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int rb = t >> 4, cb = t & 15;  // register block: rows 4rb..4rb+3, cols 8cb..8cb+7
  const int m = a.m, n = a.n, k = a.k;
  const T *__restrict__ mat = a.mat + (int64_t)b * a.mat_stride;
  T *__restrict__ uo = a.u + (int64_t)b * a.u_stride;
  T *__restrict__ co = a.carry + (int64_t)b * a.carry_stride;

  if (k == m) {
    // Nothing is discarded: U = I is an exact gauge choice, carry = mat.
    double ss = 0.0;
    for (int i = t; i < m * n; i += NT) {
      const double v = (double)mat[i];
      ss = fma(v, v, ss);
    }
    ss = block_sum(ss, sRed);
    const double nrm = sqrt(ss);
    const double sc = nrm > 0.0 ? 1.0 / nrm : 1.0;
    for (int i = t; i < m * k; i += NT) {
      const int r = i / k, c = i - r * k;
      uo[i] = (T)(r == c ? 1.0 : 0.0);
    }
    for (int i = t; i < m * n; i += NT) co[i] = (T)((double)mat[i] * sc);
    if (t == 0) a.logscale[(int64_t)b * a.log_stride] += log(nrm);
    return;
  }

  // ---------------- Gram G = mat mat^T (FP64) --------------------------------
  double g[4][8];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 8; ++c) g[r][c] = 0.0;

  for (int n0 = 0; n0 < n; n0 += NCH) {
    for (int idx = t; idx < MMAX * NCH; idx += NT) {
      const int i = idx / NCH, c = idx - i * NCH;
      double v = 0.0;
      if (i < m && n0 + c < n) v = (double)mat[(int64_t)i * n + n0 + c];
      sBig[i * STG + c] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < NCH; ++kk) {
      double av[4], bv[8];
#pragma unroll
      for (int r = 0; r < 4; ++r) av[r] = sBig[(4 * rb + r) * STG + kk];
#pragma unroll
      for (int c = 0; c < 8; ++c) bv[c] = sBig[(8 * cb + c) * STG + kk];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) g[r][c] = fma(av[r], bv[c], g[r][c]);
    }
    __syncthreads();
  }

  // ---------------- Householder tridiagonalisation ---------------------------
  // The column loop is unrolled by 8 so the register column jo is a
  // compile-time index (a runtime index would spill g to scratch memory).
  for (int jb = 0; 8 * jb < m - 2; ++jb) {
#pragma unroll
    for (int jo = 0; jo < 8; ++jo) {
      const int j = 8 * jb + jo;
      if (j >= m - 2) break;
      if (cb == jb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 4 * rb + r;
          if (i > j && i < m) sX[i] = g[r][jo];
          if (i == j) sD[j] = g[r][jo];
        }
      }
      __syncthreads();
      if (wave == 0) {
        double s = 0.0;
        for (int i = j + 2 + lane; i < m; i += 64) s = fma(sX[i], sX[i], s);
        s = wave_sum(s);
        const double alpha = sX[j + 1];
        double tau = 0.0, beta = alpha, scal = 0.0;
        if (s > 0.0) {
          beta = -copysign(sqrt(fma(alpha, alpha, s)), alpha);
          tau = (beta - alpha) / beta;
          scal = 1.0 / (alpha - beta);
        }
        for (int i = lane; i < MMAX; i += 64) {
          double v = 0.0;
          if (i == j + 1)
            v = 1.0;
          else if (i > j + 1 && i < m)
            v = sX[i] * scal;
          sV[i] = v;
        }
        if (lane == 0) {
          sTau[j] = tau;
          sE[j] = beta;
        }
      }
      __syncthreads();
      const double tau = sTau[j];
      double vc[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) vc[c] = sV[8 * cb + c];
      double acc[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 8; ++c) s = fma(g[r][c], vc[c], s);
        acc[r] = group16_sum(s);
      }
      if (cb == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 4 * rb + r;
          sP[i] = (i > j && i < m) ? tau * acc[r] : 0.0;
        }
      }
      __syncthreads();
      double dp = 0.0;
      for (int i = lane; i < MMAX; i += 64) dp = fma(sP[i], sV[i], dp);
      dp = wave_sum(dp);
      const double hc = 0.5 * tau * dp;
      double vr[4], wr[4], wc[8];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        vr[r] = sV[4 * rb + r];
        wr[r] = fma(-hc, vr[r], sP[4 * rb + r]);
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) wc[c] = fma(-hc, vc[c], sP[8 * cb + c]);
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) g[r][c] = fma(-vr[r], wc[c], fma(-wr[r], vc[c], g[r][c]));
      if (cb == jb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 4 * rb + r;
          if (i > j + 1 && i < m) g[r][jo] = vr[r];
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int i = 4 * rb + r, jj = 8 * cb + c;
      if (i == jj && i >= m - 2 && i < m) sD[i] = g[r][c];
      if (m >= 2 && i == m - 1 && jj == m - 2) sE[m - 2] = g[r][c];
    }
  __syncthreads();

  // ---------------- top-k eigenvalues of T: multisection --------------------
  double glo = INFINITY, ghi = -INFINITY, e2max = 0.0;
  if (t < m) {
    const double el = t > 0 ? fabs(sE[t - 1]) : 0.0;
    const double er = t < m - 1 ? fabs(sE[t]) : 0.0;
    glo = sD[t] - el - er;
    ghi = sD[t] + el + er;
    if (t < m - 1) {
      sE2[t] = sE[t] * sE[t];
      e2max = sE2[t];
    }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    glo = fmin(glo, __shfl_xor(glo, off, 64));
    ghi = fmax(ghi, __shfl_xor(ghi, off, 64));
    e2max = fmax(e2max, __shfl_xor(e2max, off, 64));
  }
  if (lane == 0) {
    sRed[wave] = glo;
    sRed[16 + wave] = ghi;
    sRed[32 + wave] = e2max;
  }
  __syncthreads();
  glo = sRed[0];
  ghi = sRed[16];
  e2max = sRed[32];
  for (int w = 1; w < NT / 64; ++w) {
    glo = fmin(glo, sRed[w]);
    ghi = fmax(ghi, sRed[16 + w]);
    e2max = fmax(e2max, sRed[32 + w]);
  }
  const double pivmin = DBL_MIN * fmax(1.0, e2max);
  const double tnorm = fmax(fabs(glo), fabs(ghi));
  glo -= 2.1 * DBL_EPSILON * tnorm * m + 4.2 * pivmin;
  ghi += 2.1 * DBL_EPSILON * tnorm * m + 4.2 * pivmin;
  const double abstol = DBL_EPSILON * tnorm;

  const int gid = rb;  // 32 groups of 16 lanes
  {
    const bool active = gid < k;
    const int idx = m - k + gid;
    double lo = glo, hi = ghi;
    for (int it = 0; it < 80; ++it) {
      const bool conv =
          !active || (hi - lo) <= fmax(2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)), abstol);
      if (__all(conv)) break;
      if (!conv) {
        const double x = lo + (hi - lo) * (double)(cb + 1) * (1.0 / 17.0);
        double q = guard_pivot(sD[0] - x, pivmin);
        int cnt = q <= 0.0;
        for (int i = 1; i < m; ++i) {
          q = (sD[i] - x) - sE2[i - 1] / q;
          if (fabs(q) < pivmin) q = -pivmin;
          cnt += q <= 0.0;
        }
        double cl = cnt <= idx ? x : -INFINITY;
        double ch = cnt > idx ? x : INFINITY;
        cl = group16_max(cl);
        ch = group16_min(ch);
        lo = fmax(lo, cl);
        hi = fmin(hi, ch);
      }
    }
    if (active && cb == 0) sLam[k - 1 - gid] = 0.5 * (lo + hi);  // descending order
  }
  __syncthreads();

  // ---------------- eigenvectors of T: twisted factorisations ---------------
  if (cb == 0 && gid < k) {
    const double lam = sLam[gid];
    double *Dp = sY + gid * LDY;
    double *Dm = sBig + gid * LDY;
    double dp = sD[0] - lam;
    for (int i = 0; i < m - 1; ++i) {
      dp = guard_pivot(dp, pivmin);
      Dp[i] = dp;
      dp = (sD[i + 1] - lam) - sE2[i] / dp;
    }
    Dp[m - 1] = guard_pivot(dp, pivmin);
    double dm = sD[m - 1] - lam;
    int r = m - 1;
    double best = INFINITY;
    for (int i = m - 1; i >= 1; --i) {
      dm = guard_pivot(dm, pivmin);
      Dm[i] = dm;
      const double gam = Dp[i] + dm - (sD[i] - lam);
      if (fabs(gam) < best) {
        best = fabs(gam);
        r = i;
      }
      dm = (sD[i - 1] - lam) - sE2[i - 1] / dm;
    }
    dm = guard_pivot(dm, pivmin);
    Dm[0] = dm;
    {
      const double gam = Dp[0] + dm - (sD[0] - lam);
      if (fabs(gam) < best) r = 0;
    }
    double nrm2 = 1.0, z = 1.0;
    for (int i = r - 1; i >= 0; --i) {
      z = -(sE[i] / Dp[i]) * z;
      Dp[i] = z;
      nrm2 = fma(z, z, nrm2);
    }
    z = 1.0;
    for (int i = r + 1; i < m; ++i) {
      z = -(sE[i - 1] / Dm[i]) * z;
      Dp[i] = z;
      nrm2 = fma(z, z, nrm2);
    }
    Dp[r] = 1.0;
    const double inv = 1.0 / sqrt(nrm2);
    for (int i = 0; i < MMAX; ++i) Dp[i] = i < m ? Dp[i] * inv : 0.0;
  }
  __syncthreads();

  // ---------------- back-transform U = Q Z ----------------------------------
  double y[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) y[e] = gid < k ? sY[gid * LDY + 8 * cb + e] : 0.0;
  const int nref = m - 2;
  double *sVb = sBig;  // 8 reflectors x MMAX
  __syncthreads();
  for (int blk = (nref - 1) >> 3; blk >= 0 && nref > 0; --blk) {
    if (cb == blk) {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int j = 8 * blk + c;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 4 * rb + r;
          const double v = (i == j + 1) ? 1.0 : ((i > j + 1 && i < m) ? g[r][c] : 0.0);
          sVb[c * MMAX + i] = v;
        }
      }
    }
    __syncthreads();
    if (gid < k) {
      const int jhi = min(8 * blk + 7, nref - 1);
      for (int j = jhi; j >= 8 * blk; --j) {
        const double *vj = sVb + (j & 7) * MMAX + 8 * cb;
        double s = 0.0;
#pragma unroll
        for (int e = 0; e < 8; ++e) s = fma(vj[e], y[e], s);
        s = group16_sum(s);
        const double f = sTau[j] * s;
#pragma unroll
        for (int e = 0; e < 8; ++e) y[e] = fma(-f, vj[e], y[e]);
      }
    }
    __syncthreads();
  }

  // ---------------- modified Gram-Schmidt ------------------------------------
  for (int i = 0; i < k; ++i) {
    if (gid == i) {
      double s = 0.0;
#pragma unroll
      for (int e = 0; e < 8; ++e) s = fma(y[e], y[e], s);
      s = group16_sum(s);
      const double inv = 1.0 / sqrt(s);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        y[e] *= inv;
        sY[i * LDY + 8 * cb + e] = y[e];
      }
    }
    __syncthreads();
    if (gid > i && gid < k) {
      const double *q = sY + i * LDY + 8 * cb;
      double s = 0.0;
#pragma unroll
      for (int e = 0; e < 8; ++e) s = fma(q[e], y[e], s);
      s = group16_sum(s);
#pragma unroll
      for (int e = 0; e < 8; ++e) y[e] = fma(-s, q[e], y[e]);
    }
  }
  __syncthreads();

  // ---------------- outputs: U and normalised carry = U^T mat ---------------
  for (int idx = t; idx < m * k; idx += NT) {
    const int i = idx / k, l = idx - i * k;
    uo[idx] = (T)sY[l * LDY + i];
  }
  const int nn = t & 127, l0 = t >> 7;
  double acc[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) acc[e] = 0.0;
  if (nn < n) {
    for (int i = 0; i < m; ++i) {
      const double mv = (double)mat[(int64_t)i * n + nn];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int l = l0 + 4 * e;
        if (l < k) acc[e] = fma(sY[l * LDY + i], mv, acc[e]);
      }
    }
  }
  double ss = 0.0;
#pragma unroll
  for (int e = 0; e < 8; ++e)
    if (nn < n && l0 + 4 * e < k) ss = fma(acc[e], acc[e], ss);
  ss = block_sum(ss, sRed);
  const double nrm = sqrt(ss);
  const double sc = nrm > 0.0 ? 1.0 / nrm : 1.0;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int l = l0 + 4 * e;
    if (nn < n && l < k) co[(int64_t)l * n + nn] = (T)(acc[e] * sc);
  }
  if (t == 0) a.logscale[(int64_t)b * a.log_stride] += log(nrm);
}

}  // namespace

template <typename T>
void launch_trunc(const TruncArgs<T> &a, int batch, hipStream_t stream) {
  hipLaunchKernelGGL(k_trunc<T>, dim3(batch), dim3(NT), 0, stream, a);
}

template void launch_trunc<float>(const TruncArgs<float> &, int, hipStream_t);
template void launch_trunc<double>(const TruncArgs<double> &, int, hipStream_t);

}  // namespace vmc
