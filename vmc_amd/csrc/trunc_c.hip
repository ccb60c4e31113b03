// Complex zip-up truncation step on gfx950 (complex64 / complex128 plans).
//
// Same algorithm as trunc.hip for a complex theta matrix mat (m x n, m, n <= 64):
//   G = mat mat^H                      (Hermitian Gram, complex FP64, registers)
//   G = Q T Q^H                        (complex Householder tridiagonalisation in the
//                                       LAPACK zhetd2 form: H = I - tau v v^H with a
//                                       real beta, so T is real symmetric; the last
//                                       reflector is a phase that makes e_{m-2} real)
//   top-k eigenpairs of T              (real multisection + twisted factorisation)
//   U = Q Z, complex MGS, carry = U^H mat / |U^H mat|_F
// replacing the complex SVD of the reference zip-up
// (src/vmc/peps/common/strategy.py:142-148). The kept subspace equals the span
// of the reference's U_k; its gauge may differ.
//
// Register layout: 512 threads, thread t owns G[rb + 32 r][cb + 16 c]
// (rb = t / 16, cb = t % 16, r < 2, c < 4).
#include "common.hpp"

#include <cfloat>

namespace vmc {
namespace {

constexpr int NT = 512;
constexpr int MMAX = 64;
constexpr int KMAX = 32;
constexpr int LDY = 65;
constexpr int STG = 17;
constexpr int NCH = 16;

__device__ inline double guard_pivot(double d, double pivmin) {
  if (fabs(d) < pivmin) d = (d < 0.0) ? -pivmin : pivmin;
  return d;
}

__device__ inline double frcp(double q) {
  const double r = __builtin_amdgcn_rcp(q);
  return fma(r, fma(-q, r, 1.0), r);
}

__device__ inline cf64 cmulc(cf64 a, cf64 b) {  // conj(a) * b
  return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}

__device__ inline cf64 group16_csum(cf64 v) { return {group16_sum(v.x), group16_sum(v.y)}; }
__device__ inline cf64 wave_csum(cf64 v) { return {wave_sum(v.x), wave_sum(v.y)}; }

template <typename T, bool DM>  // DM: density-matrix cut (see trunc.hip)
__global__ __launch_bounds__(NT) void k_trunc_c(TruncArgs<T> a) {
  __shared__ cf64 sBig[KMAX * LDY];  // Gram staging | Dm | reflector block | mat rows
  __shared__ cf64 sY[KMAX * LDY];    // Dp (.x) / eigenvectors / orthonormal U, column-major
  __shared__ cf64 sX[2][MMAX], sP[2][MMAX];
  __shared__ double sD[MMAX], sE[MMAX], sE2[MMAX];
  __shared__ cf64 sTau[MMAX];
  __shared__ double sLam[KMAX];
  __shared__ double sRed[3 * 16];

  const int b = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int rb = t >> 4, cb = t & 15;
  const int m = a.m, n = a.n, k = a.k;
  const T *__restrict__ mat = a.mat + (int64_t)b * a.mat_stride;
  T *__restrict__ uo = a.u + (int64_t)b * a.u_stride;
  T *__restrict__ co = a.carry + (int64_t)b * a.carry_stride;

  if (k == m) {
    double ss = 0.0;
    for (int i = t; i < m * n; i += NT) ss += abs2(mat[i]);
    ss = block_sum(ss, sRed);
    const double nrm = sqrt(ss);
    const double sc = nrm > 0.0 ? 1.0 / nrm : 1.0;
    T *__restrict__ eye = DM ? co : uo;  // density-matrix: V = I (see trunc.hip)
    T *__restrict__ body = DM ? uo : co;
    for (int i = t; i < m * k; i += NT) {
      const int r = i / k, c = i - r * k;
      eye[i] = from_acc<T>(cf64(r == c ? 1.0 : 0.0));
    }
    for (int i = t; i < m * n; i += NT) body[i] = from_acc<T>(sc * to_acc(mat[i]));
    if (t == 0) a.logscale[(int64_t)b * a.log_stride] += log(nrm);
    return;
  }

  // ---------------- Gram G = mat mat^H ---------------------------------------
  cf64 g[2][4];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) g[r][c] = cf64(0.0);
  for (int n0 = 0; n0 < n; n0 += NCH) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int idx = t + NT * e;
      const int i = idx >> 4, c = idx & 15;
      sBig[i * STG + c] = (i < m && n0 + c < n) ? to_acc(mat[DM ? (int64_t)(n0 + c) * m + i : (int64_t)i * n + n0 + c]) : cf64(0.0);
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < NCH; ++kk) {
      cf64 av[2], bv[4];
#pragma unroll
      for (int r = 0; r < 2; ++r) av[r] = sBig[(rb + 32 * r) * STG + kk];
#pragma unroll
      for (int c = 0; c < 4; ++c) bv[c] = conj(sBig[(cb + 16 * c) * STG + kk]);
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) g[r][c] = fma(av[r], bv[c], g[r][c]);
    }
    __syncthreads();
  }

  // ---------------- complex Householder tridiagonalisation ------------------
  {
    int buf = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      for (int jo = 0; jo < 16; ++jo) {
        const int j = 16 * c + jo;
        if (j >= m - 1) break;
        cf64 *X = sX[buf];
        cf64 *Pv = sP[buf];
        if (cb == jo) {
#pragma unroll
          for (int r = 0; r < 2; ++r) {
            const int i = rb + 32 * r;
            if (i > j && i < m) X[i] = g[r][c];
            if (i == j) sD[j] = g[r][c].x;
          }
        }
        __syncthreads();
        double s = 0.0;
        for (int i = j + 2 + lane; i < m; i += 64) s += abs2(X[i]);
        s = wave_sum(s);
        const cf64 alpha = X[j + 1];
        cf64 tau(0.0), scal(0.0);
        double beta = alpha.x;
        if (s > 0.0 || alpha.y != 0.0) {
          beta = -copysign(sqrt(abs2(alpha) + s), alpha.x);
          tau = cf64((beta - alpha.x) / beta, -alpha.y / beta);
          const cf64 dnm = cf64(alpha.x - beta, alpha.y);
          const double inv = 1.0 / abs2(dnm);
          scal = cf64(dnm.x * inv, -dnm.y * inv);
        }
        if (t == 0) {
          sTau[j] = tau;
          sE[j] = beta;
        }
        auto vat = [&](int i) -> cf64 {
          return i == j + 1 ? cf64(1.0) : ((i > j + 1 && i < m) ? X[i] * scal : cf64(0.0));
        };
        cf64 vc[4], vr[2];
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) vc[cc] = vat(cb + 16 * cc);
#pragma unroll
        for (int r = 0; r < 2; ++r) vr[r] = vat(rb + 32 * r);
        cf64 acc[2];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          cf64 p(0.0);
#pragma unroll
          for (int cc = 0; cc < 4; ++cc) p = fma(g[r][cc], vc[cc], p);
          acc[r] = group16_csum(p);
        }
        if (cb == 0) {
#pragma unroll
          for (int r = 0; r < 2; ++r) {
            const int i = rb + 32 * r;
            Pv[i] = (i > j && i < m) ? tau * acc[r] : cf64(0.0);
          }
        }
        __syncthreads();
        // x^H v with x = tau G v; lanes 0..63 cover every row
        cf64 dp = cmulc(Pv[lane], vat(lane));
        dp = wave_csum(dp);
        const cf64 hc = 0.5 * (tau * dp);
        cf64 wr[2], wc[4];
#pragma unroll
        for (int r = 0; r < 2; ++r) wr[r] = Pv[rb + 32 * r] - hc * vr[r];
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) wc[cc] = Pv[cb + 16 * cc] - hc * vc[cc];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
          for (int cc = 0; cc < 4; ++cc)
            g[r][cc] = g[r][cc] - (vr[r] * conj(wc[cc]) + wr[r] * conj(vc[cc]));
        if (cb == jo) {
#pragma unroll
          for (int r = 0; r < 2; ++r) {
            const int i = rb + 32 * r;
            if (i > j + 1 && i < m) g[r][c] = vr[r];
          }
        }
        buf ^= 1;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = rb + 32 * r, jj = cb + 16 * c;
      if (i == jj && i == m - 1) sD[i] = g[r][c].x;
    }
  __syncthreads();

  // ---------------- top-k eigenvalues of the real T --------------------------
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
  glo = wave_min(glo);
  ghi = wave_max(ghi);
  e2max = wave_max(e2max);
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
  const int gid = rb;
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
#pragma unroll 4
        for (int i = 1; i < m; ++i) {
          q = fma(-sE2[i - 1], frcp(q), sD[i] - x);
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
    if (active && cb == 0) sLam[k - 1 - gid] = 0.5 * (lo + hi);
  }
  __syncthreads();

  // ---------------- real eigenvectors of T (twisted) -------------------------
  if (cb == 0 && gid < k) {
    const double lam = sLam[gid];
    cf64 *Dp = sY + gid * LDY;
    cf64 *Dm = sBig + gid * LDY;
    double dp = sD[0] - lam;
    for (int i = 0; i < m - 1; ++i) {
      dp = guard_pivot(dp, pivmin);
      Dp[i].x = dp;
      dp = fma(-sE2[i], frcp(dp), sD[i + 1] - lam);
    }
    Dp[m - 1].x = guard_pivot(dp, pivmin);
    double dm = sD[m - 1] - lam;
    int r = m - 1;
    double best = INFINITY;
    for (int i = m - 1; i >= 1; --i) {
      dm = guard_pivot(dm, pivmin);
      Dm[i].x = dm;
      const double gam = Dp[i].x + dm - (sD[i] - lam);
      if (fabs(gam) < best) {
        best = fabs(gam);
        r = i;
      }
      dm = fma(-sE2[i - 1], frcp(dm), sD[i - 1] - lam);
    }
    dm = guard_pivot(dm, pivmin);
    Dm[0].x = dm;
    {
      const double gam = Dp[0].x + dm - (sD[0] - lam);
      if (fabs(gam) < best) r = 0;
    }
    double nrm2 = 1.0, z = 1.0;
    for (int i = r - 1; i >= 0; --i) {
      z = -(sE[i] * frcp(Dp[i].x)) * z;
      Dp[i].x = z;
      nrm2 = fma(z, z, nrm2);
    }
    z = 1.0;
    for (int i = r + 1; i < m; ++i) {
      z = -(sE[i - 1] * frcp(Dm[i].x)) * z;
      Dp[i].x = z;
      nrm2 = fma(z, z, nrm2);
    }
    Dp[r].x = 1.0;
    const double inv = 1.0 / sqrt(nrm2);
    for (int i = 0; i < MMAX; ++i) Dp[i] = cf64(i < m ? Dp[i].x * inv : 0.0);
  }
  __syncthreads();

  // ---------------- back-transform U = Q Z (complex reflectors) --------------
  cf64 y[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) y[e] = gid < k ? sY[gid * LDY + cb + 16 * e] : cf64(0.0);
  const int nref = m - 1;
  cf64 *sVb = sBig;  // 16 reflectors x MMAX
#pragma unroll
  for (int c = 3; c >= 0; --c) {
    if (16 * c >= nref) continue;
    __syncthreads();
    {
      const int j = 16 * c + cb;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int i = rb + 32 * r;
        cf64 v(0.0);
        if (j < nref) v = (i == j + 1) ? cf64(1.0) : ((i > j + 1 && i < m) ? g[r][c] : cf64(0.0));
        sVb[cb * MMAX + i] = v;
      }
    }
    __syncthreads();
    if (gid < k) {
      const int jhi = min(16 * c + 15, nref - 1);
      for (int j = jhi; j >= 16 * c; --j) {
        const cf64 *vj = sVb + (j - 16 * c) * MMAX + cb;
        cf64 s(0.0);
#pragma unroll
        for (int e = 0; e < 4; ++e) s = s + cmulc(vj[16 * e], y[e]);
        s = group16_csum(s);
        const cf64 f = sTau[j] * s;
#pragma unroll
        for (int e = 0; e < 4; ++e) y[e] = y[e] - f * vj[16 * e];
      }
    }
  }

  // ---------------- complex modified Gram-Schmidt -----------------------------
  __syncthreads();
  for (int i = 0; i < k; ++i) {
    if (gid == i) {
      double s = 0.0;
#pragma unroll
      for (int e = 0; e < 4; ++e) s += abs2(y[e]);
      s = group16_sum(s);
      if (s < 0.25) {
        // Repeated eigenvalue (e.g. the exact zeros of a rank-deficient theta):
        // replace by the unit vector with the largest part outside
        // span(q_0..q_{i-1}), orthogonalised twice (see trunc.hip).
        double best = -2.0;
        int bp = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = cb + 16 * e;
          double res = 1.0;
          for (int j = 0; j < i; ++j) res -= abs2(sY[j * LDY + row]);
          if (row < m && res > best) {
            best = res;
            bp = row;
          }
        }
        const double mx = group_max<16>(best);
        const int p = (int)group_min<16>(best == mx ? (double)bp : 1e9);
#pragma unroll
        for (int e = 0; e < 4; ++e) y[e] = cf64(cb + 16 * e == p ? 1.0 : 0.0);
        for (int pass = 0; pass < 2; ++pass)
          for (int j = 0; j < i; ++j) {
            const cf64 *q = sY + j * LDY + cb;
            cf64 d(0.0);
#pragma unroll
            for (int e = 0; e < 4; ++e) d = d + cmulc(q[16 * e], y[e]);
            d = group16_csum(d);
#pragma unroll
            for (int e = 0; e < 4; ++e) y[e] = y[e] - d * q[16 * e];
          }
        s = 0.0;
#pragma unroll
        for (int e = 0; e < 4; ++e) s += abs2(y[e]);
        s = group16_sum(s);
      }
      const double inv = 1.0 / sqrt(s);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        y[e] = inv * y[e];
        sY[i * LDY + cb + 16 * e] = y[e];
      }
    }
    __syncthreads();
    if (gid > i && gid < k) {
      const cf64 *q = sY + i * LDY + cb;
      cf64 s(0.0);
#pragma unroll
      for (int e = 0; e < 4; ++e) s = s + cmulc(q[16 * e], y[e]);
      s = group16_csum(s);
#pragma unroll
      for (int e = 0; e < 4; ++e) y[e] = y[e] - s * q[16 * e];
    }
  }
  __syncthreads();

  // ---------------- outputs: U and carry = U^H mat ----------------------------
  for (int idx = t; idx < m * k; idx += NT) {
    const int i = idx / k, l = idx - i * k;
    if (DM)
      co[(int64_t)l * m + i] = from_acc<T>(sY[l * LDY + i]);
    else
      uo[idx] = from_acc<T>(sY[l * LDY + i]);
  }
  const int nn = t & 63, l0 = t >> 6;
  cf64 acc[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) acc[e] = cf64(0.0);
  for (int i0 = 0; i0 < m; i0 += 16) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int idx = t + NT * e;
      const int ii = idx >> 6, c = idx & 63;
      sBig[idx] = (i0 + ii < m && c < n) ? to_acc(mat[DM ? (int64_t)c * m + i0 + ii : (int64_t)(i0 + ii) * n + c]) : cf64(0.0);
    }
    __syncthreads();
    const int ni = min(16, m - i0);
    for (int ii = 0; ii < ni; ++ii) {
      const cf64 mv = sBig[ii * 64 + nn];
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[e] = acc[e] + cmulc(sY[(l0 + 8 * e) * LDY + i0 + ii], mv);
    }
    __syncthreads();
  }
  double ss = 0.0;
#pragma unroll
  for (int e = 0; e < 4; ++e)
    if (nn < n && l0 + 8 * e < k) ss += abs2(acc[e]);
  ss = block_sum(ss, sRed);
  const double nrm = sqrt(ss);
  const double sc = nrm > 0.0 ? 1.0 / nrm : 1.0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int l = l0 + 8 * e;
    if (nn < n && l < k) {
      if (DM)
        uo[(int64_t)nn * k + l] = from_acc<T>(sc * acc[e]);
      else
        co[(int64_t)l * n + nn] = from_acc<T>(sc * acc[e]);
    }
  }
  if (t == 0) a.logscale[(int64_t)b * a.log_stride] += log(nrm);
}

}  // namespace

template <>
void launch_trunc<cf32>(const TruncArgs<cf32> &a, int batch, hipStream_t stream) {
  if (a.dm)
    hipLaunchKernelGGL((k_trunc_c<cf32, true>), dim3(batch), dim3(NT), 0, stream, a);
  else
    hipLaunchKernelGGL((k_trunc_c<cf32, false>), dim3(batch), dim3(NT), 0, stream, a);
}

template <>
void launch_trunc<cf64>(const TruncArgs<cf64> &a, int batch, hipStream_t stream) {
  if (a.dm)
    hipLaunchKernelGGL((k_trunc_c<cf64, true>), dim3(batch), dim3(NT), 0, stream, a);
  else
    hipLaunchKernelGGL((k_trunc_c<cf64, false>), dim3(batch), dim3(NT), 0, stream, a);
}

}  // namespace vmc
