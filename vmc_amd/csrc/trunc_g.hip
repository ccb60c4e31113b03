// Truncation step for theta matrices beyond the register-resident kernels
// (trunc.hip: real m, n <= 128; trunc_c.hip: complex m, n <= 64; both k <= 32):
// any m, n <= GMAX and k <= GKMAX, real or complex, in FP64. One workgroup of
// GT threads per sample; the Gram matrix and the eigenvector block live in a
// per-sample global scratch (L2 / MALL resident), vectors in LDS or registers:
//
//   G = X X^H                        (LDS-tiled, lower tiles, mirrored)
//   G = Q T Q^H                      (LAPACK zhetrd/dsytrd lower form in panels of
//                                     NB columns: H_j = I - tau_j v_j v_j^H with
//                                     v_j(j+1) = 1 and a real beta, so T is real
//                                     symmetric; reflector j is kept in row j of G
//                                     above the diagonal)
//   lambda_{m-k..m-1}(T)             (Sturm-count multisection)
//   z_l = twisted-factorisation vector
//   U = Q Z, modified Gram-Schmidt   (repeated eigenvalues handled as in trunc.hip)
//   carry = U^H X / |U^H X|_F         (log of the norm added to the boundary scale)
//
// X = mat (m x n) for zip-up; for a density-matrix cut X = theta^T and the
// outputs swap as described at TruncArgs::dm. Same results as the register
// kernels (same subspace; gauge may differ); this one trades their speed for
// size: each Householder step streams the trailing Gram matrix from the cache
// once, and each panel reads and writes it once more.
#include "common.hpp"

#include <cfloat>

namespace vmc {
namespace {

constexpr int GT = 256;     // threads per workgroup
constexpr int GMAX = TRUNC_G_MAX_M;   // max rows of X (LDS vector length)
constexpr int GKMAX = TRUNC_G_MAX_K;  // max kept vectors
constexpr int TL = 32;      // Gram tile
constexpr int NB = TRUNC_G_NB;  // tridiagonalisation panel width
constexpr int RC = 64;      // trailing-update rows staged in LDS per pass
constexpr int PR = GMAX / 64;  // rows per lane of a register-held column
constexpr int CG = 2;       // back-transform columns per wave and pass
constexpr int CL = 8;       // carry rows per thread
static_assert(2 * RC * NB <= 2 * GMAX, "staged panel rows fit over sV and sW");
static_assert(GMAX % 64 == 0, "register-held columns");

__device__ inline double gcj(double v) { return v; }
__device__ inline cf64 gcj(cf64 v) { return {v.x, -v.y}; }
__device__ inline double gre(double v) { return v; }
__device__ inline double gre(cf64 v) { return v.x; }
__device__ inline double gim(double) { return 0.0; }
__device__ inline double gim(cf64 v) { return v.y; }
template <typename S>
__device__ inline S gmk(double re, double im);
template <>
__device__ inline double gmk<double>(double re, double) { return re; }
template <>
__device__ inline cf64 gmk<cf64>(double re, double im) { return {re, im}; }
__device__ inline double gscl(double s, double v) { return s * v; }
__device__ inline cf64 gscl(double s, cf64 v) { return {s * v.x, s * v.y}; }

__device__ inline double gguard(double d, double pivmin) {
  if (fabs(d) < pivmin) d = (d < 0.0) ? -pivmin : pivmin;
  return d;
}

// Block-wide sum (GT threads); red holds >= 2 * GT / 64 doubles.
__device__ inline double gsum(double v, double *red) { return block_sum(v, red); }
__device__ inline cf64 gsum(cf64 v, double *red) { return block_sum(v, red); }

// Sturm multisection for the top-k eigenvalues of T (d, e2 = e^2) with L lanes
// per eigenvalue: every step evaluates L counts and shrinks the interval
// (L + 1)-fold. Writes sLam[0..k) in descending order.
template <int L>
__device__ void top_eigenvalues(const double *sD, const double *sE2, int m, int k, double glo, double ghi,
                                double pivmin, double abstol, double *sLam) {
  const int t = threadIdx.x, eg = t / L, sub = t % L;
  const int nper = GT / L;  // eigenvalues per pass
  for (int base = 0; base < k; base += nper) {
    const int me = base + eg;
    const bool active = me < k;
    const int idx = m - k + me;  // ascending index of this eigenvalue
    double lo = glo, hi = ghi;
    for (int it = 0; it < 400; ++it) {
      const bool conv = !active || (hi - lo) <= fmax(2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)), abstol);
      if (__syncthreads_and(conv)) break;
      if (!conv) {
        const double x = lo + (hi - lo) * (double)(sub + 1) * (1.0 / (L + 1));
        double q = gguard(sD[0] - x, pivmin);
        int cnt = q <= 0.0;
        for (int i = 1; i < m; ++i) {
          q = sD[i] - x - sE2[i - 1] / q;
          if (fabs(q) < pivmin) q = -pivmin;
          cnt += q <= 0.0;
        }
        double cl = cnt <= idx ? x : -INFINITY;
        double ch = cnt > idx ? x : INFINITY;
        cl = group_max<L>(cl);
        ch = group_min<L>(ch);
        lo = fmax(lo, cl);
        hi = fmin(hi, ch);
      }
    }
    if (active && sub == 0) sLam[k - 1 - me] = 0.5 * (lo + hi);
    __syncthreads();
  }
}

template <typename T, typename S, bool DM>
__global__ __launch_bounds__(GT, 4) void k_trunc_g(TruncArgs<T> a) {
  // The Gram tiles are dead once G is formed, and the vectors of the later
  // phases are first written after it, so the two share one buffer: 38 KB of
  // LDS instead of 72 KB (complex) lets more workgroups share a CU.
  constexpr size_t kGramBytes = 2 * sizeof(S) * TL * (TL + 1);
  constexpr size_t kVecBytes = 3 * sizeof(S) * GMAX + 3 * sizeof(double) * GMAX;
  __shared__ __align__(16) unsigned char sRaw[kGramBytes > kVecBytes ? kGramBytes : kVecBytes];
  S(*const sA)[TL + 1] = reinterpret_cast<S(*)[TL + 1]>(sRaw);  // Gram tiles
  S(*const sB)[TL + 1] = sA + TL;
  S *const sTau = reinterpret_cast<S *>(sRaw);  // after the Gram
  double *const sD = reinterpret_cast<double *>(sTau + GMAX);
  double *const sE = sD + GMAX;
  double *const sE2 = sE + GMAX;
  S *const sV = reinterpret_cast<S *>(sE2 + GMAX);
  S *const sW = sV + GMAX;
  // the back-transform stages reflectors over sD .. sW (dead by then)
  constexpr size_t kRefBytes = 3 * sizeof(double) * GMAX + 2 * sizeof(S) * GMAX;
  __shared__ double sLam[GKMAX];
  __shared__ S sPan[2 * NB];  // panel products W^H v, V^H v
  __shared__ double red[2 * GT / 64 + 16];
  __shared__ int sIdx;

  const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m = a.m, n = a.n, k = a.k;
  const T *__restrict__ mat = a.mat + (int64_t)b * a.mat_stride;
  T *__restrict__ uo = a.u + (int64_t)b * a.u_stride;
  T *__restrict__ co = a.carry + (int64_t)b * a.carry_stride;
  S *__restrict__ G = (S *)a.gscr + (int64_t)b * a.gscr_stride;  // m x m, row-major
  S *__restrict__ Z = G + (int64_t)m * m;                         // k columns of length m
  S *__restrict__ Dm = Z + (int64_t)m * k;                         // twisted-solve scratch
  // X(i, c): row i of the factored matrix
  auto X = [&](int i, int c) -> S {
    return (S)to_acc(mat[DM ? (int64_t)c * m + i : (int64_t)i * n + c]);
  };

  if (k == m) {
    // nothing is discarded: U = I, carry = X (density-matrix: new = theta, carry = I)
    double ss = 0.0;
    for (int i = t; i < m * n; i += GT) ss += abs2(mat[i]);
    ss = block_sum(ss, red);
    const double nrm = sqrt(ss);
    const double sc = nrm > 0.0 ? 1.0 / nrm : 1.0;
    T *__restrict__ eye = DM ? co : uo;
    T *__restrict__ body = DM ? uo : co;
    for (int i = t; i < m * k; i += GT) {
      const int r = i / k, c = i - r * k;
      eye[i] = from_acc<T>(gmk<S>(r == c ? 1.0 : 0.0, 0.0));
    }
    for (int i = t; i < m * n; i += GT) body[i] = from_acc<T>(gscl(sc, (S)to_acc(mat[i])));
    if (t == 0) a.logscale[(int64_t)b * a.log_stride] += log(nrm);
    return;
  }

  // ---------------- Gram G = X X^H (lower tiles, mirrored) ------------------
  {
    const int r = t >> 3, c0 = (t & 7) * 4;
    const int nt = (m + TL - 1) / TL;
    for (int ib = 0; ib < nt; ++ib)
      for (int jb = 0; jb <= ib; ++jb) {
        S acc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = gmk<S>(0.0, 0.0);
        for (int k0 = 0; k0 < n; k0 += TL) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int idx = t + GT * e, rr = idx >> 5, cc = idx & 31;
            const int ia = ib * TL + rr, jbr = jb * TL + rr, kk = k0 + cc;
            sA[rr][cc] = (ia < m && kk < n) ? X(ia, kk) : gmk<S>(0.0, 0.0);
            sB[rr][cc] = (jbr < m && kk < n) ? X(jbr, kk) : gmk<S>(0.0, 0.0);
          }
          __syncthreads();
          for (int kk = 0; kk < TL; ++kk) {
            const S av = sA[r][kk];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = acc[q] + av * gcj(sB[c0 + q][kk]);
          }
          __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = ib * TL + r, j = jb * TL + c0 + q;
          if (i < m && j < m && i >= j) {
            G[(int64_t)i * m + j] = acc[q];
            G[(int64_t)j * m + i] = gcj(acc[q]);
          }
        }
      }
  }
  __syncthreads();

  // ---------------- Householder tridiagonalisation (lower, blocked) ---------
  // Panels of NB columns (LAPACK zlatrd/dlatrd form): inside a panel G keeps
  // its panel-start values; column j first takes the earlier reflectors of the
  // panel, A(:, j) -= V W(j, :)^H + W V(j, :)^H, and the product A v is formed
  // from the panel-start matrix and corrected by V (W^H v) + W (V^H v). The
  // trailing matrix takes the panel's rank-2 NB update once, at the panel's end,
  // so each Householder step streams the trailing matrix once (the product)
  // instead of three times (product, update read and write).
  // Panel vectors, column-major m x NB: V(:, q) (zero above row j0 + q + 1, one
  // there) and W(:, q); both live in the eigenvector scratch, unused until later.
  S *__restrict__ Vp = Z;
  S *__restrict__ Wp = Z + (int64_t)m * NB;
  for (int j0 = 0; j0 < m - 1; j0 += NB) {
    const int nbp = min(NB, m - 1 - j0);
    for (int p = 0; p < nbp; ++p) {
      const int j = j0 + p;
      // column j (rows >= j) takes the panel's earlier reflectors; |x|^2 of the
      // part reflected away (rows >= j + 2) is summed on the way
      double xs = 0.0;
      for (int i = j + t; i < m; i += GT) {
        S g = G[(int64_t)i * m + j];
        for (int q = 0; q < p; ++q)
          g = g - Vp[(int64_t)q * m + i] * gcj(Wp[(int64_t)q * m + j]) - Wp[(int64_t)q * m + i] * gcj(Vp[(int64_t)q * m + j]);
        if (p > 0) G[(int64_t)i * m + j] = g;
        if (i >= j + 2) xs += abs2(g);
      }
      xs = block_sum(xs, red);  // its barriers also publish the updated column
      const S alpha = G[(int64_t)(j + 1) * m + j];
      const double ar = gre(alpha), ai = gim(alpha);
      S tau = gmk<S>(0.0, 0.0), scal = gmk<S>(0.0, 0.0);
      double beta = ar;
      if (xs > 0.0 || ai != 0.0) {
        beta = -copysign(sqrt(ar * ar + ai * ai + xs), ar);
        tau = gmk<S>((beta - ar) / beta, -ai / beta);
        // scal = 1 / (alpha - beta)
        const double dr = ar - beta, di = ai, den = dr * dr + di * di;
        scal = gmk<S>(dr / den, -di / den);
      }
      if (t == 0) {
        sE[j] = beta;
        sD[j] = gre(G[(int64_t)j * m + j]);
        sTau[j] = tau;
      }
      const bool reflect = gre(tau) != 0.0 || gim(tau) != 0.0;
      for (int i = j + 1 + t; i < m; i += GT) {
        const S v = i == j + 1 ? gmk<S>(1.0, 0.0) : scal * G[(int64_t)i * m + j];
        sV[i] = v;
        Vp[(int64_t)p * m + i] = v;
        if (i > j + 1) G[(int64_t)j * m + i] = v;  // reflector kept for the back-transform (row j, no later step reads it)
        if (!reflect) Wp[(int64_t)p * m + i] = gmk<S>(0.0, 0.0);
      }
      __syncthreads();
      if (reflect) {
        // sPan[2q] = W(:, q)^H v, sPan[2q + 1] = V(:, q)^H v over rows > j: a wave per product
        for (int d = wave; d < 2 * p; d += GT / 64) {
          const S *col = ((d & 1) ? Vp : Wp) + (int64_t)(d >> 1) * m;
          S s = gmk<S>(0.0, 0.0);
          for (int i = j + 1 + lane; i < m; i += 64) s = s + gcj(col[i]) * sV[i];
          s = gmk<S>(wave_sum(gre(s)), wave_sum(gim(s)));
          if (lane == 0) sPan[d] = s;
        }
        if (p > 0) __syncthreads();
        // x = tau (A0 v - V (W^H v) - W (V^H v)), A0 = the panel-start G[j+1.., j+1..]:
        // x_i = sum_l conj(G[l][i]) v_l (column reads are coalesced across threads)
        for (int i = j + 1 + t; i < m; i += GT) {
          // eight loads in flight per thread, four partial sums
          S s0 = gmk<S>(0.0, 0.0), s1 = s0, s2 = s0, s3 = s0;
          int l = j + 1;
          for (; l + 7 < m; l += 8) {
            const S *gp = G + (int64_t)l * m + i;
            S g[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) g[u] = gp[(int64_t)u * m];
#pragma unroll
            for (int u = 0; u < 8; u += 4) {
              s0 = s0 + gcj(g[u]) * sV[l + u];
              s1 = s1 + gcj(g[u + 1]) * sV[l + u + 1];
              s2 = s2 + gcj(g[u + 2]) * sV[l + u + 2];
              s3 = s3 + gcj(g[u + 3]) * sV[l + u + 3];
            }
          }
          for (; l < m; ++l) s0 = s0 + gcj(G[(int64_t)l * m + i]) * sV[l];
          for (int q = 0; q < p; ++q)
            s1 = s1 - Vp[(int64_t)q * m + i] * sPan[2 * q] - Wp[(int64_t)q * m + i] * sPan[2 * q + 1];
          sW[i] = tau * ((s0 + s1) + (s2 + s3));
        }
        __syncthreads();
        // w = x - (tau / 2)(x^H v) v
        S xv = gmk<S>(0.0, 0.0);
        for (int i = j + 1 + t; i < m; i += GT) xv = xv + gcj(sW[i]) * sV[i];
        xv = gsum(xv, red);
        const S al2 = gscl(-0.5, tau * xv);
        for (int i = j + 1 + t; i < m; i += GT) Wp[(int64_t)p * m + i] = sW[i] + al2 * sV[i];
      }
      __syncthreads();
    }
    // trailing matrix (rows, columns >= j1): A -= V W^H + W V^H over the panel.
    // Rows are staged RC at a time in LDS (over sV, sW: free between steps);
    // a thread owns column l and keeps conj(W(l, :)), conj(V(l, :)) in registers.
    const int j1 = j0 + nbp;
    S(*const sPV)[NB] = reinterpret_cast<S(*)[NB]>(sV);
    S(*const sPW)[NB] = sPV + RC;
    for (int i0 = j1; i0 < m; i0 += RC) {
      for (int e = t; e < RC * NB; e += GT) {
        const int r = e / NB, q = e - r * NB, i = i0 + r;
        const bool ok = i < m && q < nbp;
        sPV[r][q] = ok ? Vp[(int64_t)q * m + i] : gmk<S>(0.0, 0.0);
        sPW[r][q] = ok ? Wp[(int64_t)q * m + i] : gmk<S>(0.0, 0.0);
      }
      __syncthreads();
      const int nr = min(RC, m - i0);
      for (int l = j1 + t; l < m; l += GT) {
        S cw[NB], cv[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) {
          cw[q] = q < nbp ? gcj(Wp[(int64_t)q * m + l]) : gmk<S>(0.0, 0.0);
          cv[q] = q < nbp ? gcj(Vp[(int64_t)q * m + l]) : gmk<S>(0.0, 0.0);
        }
        S *g = G + (int64_t)i0 * m + l;
        for (int r = 0; r < nr; ++r) {
          S g0 = g[(int64_t)r * m];
#pragma unroll
          for (int q = 0; q < NB; ++q) g0 = g0 - sPV[r][q] * cw[q] - sPW[r][q] * cv[q];
          g[(int64_t)r * m] = g0;
        }
      }
      __syncthreads();
    }
  }
  if (t == 0) sD[m - 1] = gre(G[(int64_t)(m - 1) * m + (m - 1)]);
  __syncthreads();

  // ---------------- top-k eigenvalues of T ----------------------------------
  double glo = INFINITY, ghi = -INFINITY, e2max = 0.0;
  for (int i = t; i < m; i += GT) {
    const double el = i > 0 ? fabs(sE[i - 1]) : 0.0;
    const double er = i < m - 1 ? fabs(sE[i]) : 0.0;
    glo = fmin(glo, sD[i] - el - er);
    ghi = fmax(ghi, sD[i] + el + er);
    if (i < m - 1) {
      sE2[i] = sE[i] * sE[i];
      e2max = fmax(e2max, sE2[i]);
    }
  }
  glo = wave_min(glo);
  ghi = wave_max(ghi);
  e2max = wave_max(e2max);
  __syncthreads();
  if (lane == 0) {
    red[wave] = glo;
    red[4 + wave] = ghi;
    red[8 + wave] = e2max;
  }
  __syncthreads();
  glo = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
  ghi = fmax(fmax(red[4], red[5]), fmax(red[6], red[7]));
  e2max = fmax(fmax(red[8], red[9]), fmax(red[10], red[11]));
  __syncthreads();
  const double pivmin = DBL_MIN * fmax(1.0, e2max);
  const double tnorm = fmax(fabs(glo), fabs(ghi));
  glo -= 2.1 * DBL_EPSILON * tnorm * m + 4.2 * pivmin;
  ghi += 2.1 * DBL_EPSILON * tnorm * m + 4.2 * pivmin;
  const double abstol = DBL_EPSILON * tnorm;
  if (k <= 32)
    top_eigenvalues<8>(sD, sE2, m, k, glo, ghi, pivmin, abstol, sLam);
  else if (k <= 64)
    top_eigenvalues<4>(sD, sE2, m, k, glo, ghi, pivmin, abstol, sLam);
  else
    top_eigenvalues<2>(sD, sE2, m, k, glo, ghi, pivmin, abstol, sLam);
  __syncthreads();

  // ---------------- eigenvectors of T: twisted factorisations ---------------
  for (int l = t; l < k; l += GT) {
    const double lam = sLam[l];
    S *Dp = Z + (int64_t)l * m;    // holds D+ then z (real values)
    S *Dmn = Dm + (int64_t)l * m;
    double dp = sD[0] - lam;
    for (int i = 0; i < m - 1; ++i) {
      dp = gguard(dp, pivmin);
      Dp[i] = gmk<S>(dp, 0.0);
      dp = sD[i + 1] - lam - sE2[i] / dp;
    }
    Dp[m - 1] = gmk<S>(gguard(dp, pivmin), 0.0);
    double dm = sD[m - 1] - lam;
    int r = m - 1;
    double best = INFINITY;
    for (int i = m - 1; i >= 1; --i) {
      dm = gguard(dm, pivmin);
      Dmn[i] = gmk<S>(dm, 0.0);
      const double gam = gre(Dp[i]) + dm - (sD[i] - lam);
      if (fabs(gam) < best) {
        best = fabs(gam);
        r = i;
      }
      dm = sD[i - 1] - lam - sE2[i - 1] / dm;
    }
    dm = gguard(dm, pivmin);
    Dmn[0] = gmk<S>(dm, 0.0);
    if (fabs(gre(Dp[0]) + dm - (sD[0] - lam)) < best) r = 0;
    double nrm2 = 1.0, z = 1.0;
    for (int i = r - 1; i >= 0; --i) {
      z = -(sE[i] / gre(Dp[i])) * z;
      Dp[i] = gmk<S>(z, 0.0);
      nrm2 += z * z;
    }
    z = 1.0;
    for (int i = r + 1; i < m; ++i) {
      z = -(sE[i - 1] / gre(Dmn[i])) * z;
      Dp[i] = gmk<S>(z, 0.0);
      nrm2 += z * z;
    }
    Dp[r] = gmk<S>(1.0, 0.0);
    const double inv = 1.0 / sqrt(nrm2);
    for (int i = 0; i < m; ++i) Dp[i] = gscl(inv, Dp[i]);
  }
  __syncthreads();

  // ---------------- back-transform U = H_0 H_1 ... H_{m-2} Z ----------------
  // H_j = I - tau_j v v^H; v = e_{j+1} + G[j, j+2..] (row j above the diagonal).
  // Wave w holds CG columns of Z in registers (row lane + 64 e) per pass, and
  // the reflectors are staged nbr at a time in LDS, last one first, so Z is
  // read and written once and the reflectors once per pass.
  {
    S *const sRef = reinterpret_cast<S *>(sD);
    const int nbr = (int)(kRefBytes / (sizeof(S) * m));
    const int ne = (m + 63) / 64;
    for (int l0 = 0; l0 < k; l0 += (GT / 64) * CG) {
      S z[CG][PR];
#pragma unroll
      for (int c = 0; c < CG; ++c) {
        const int col = l0 + wave * CG + c;
#pragma unroll
        for (int e = 0; e < PR; ++e) {
          const int i = lane + 64 * e;
          z[c][e] = (col < k && i < m) ? Z[(int64_t)col * m + i] : gmk<S>(0.0, 0.0);
        }
      }
      for (int jt = m - 2; jt >= 0; jt -= nbr) {
        const int jb = max(0, jt - nbr + 1);
        __syncthreads();  // the previous block's reads are done
        for (int e = t; e < (jt - jb + 1) * m; e += GT) {
          const int r = e / m, i = e - r * m, j = jb + r;
          sRef[e] = i == j + 1 ? gmk<S>(1.0, 0.0) : (i > j + 1 ? G[(int64_t)j * m + i] : gmk<S>(0.0, 0.0));
        }
        __syncthreads();
        for (int j = jt; j >= jb; --j) {
          const S tj = sTau[j];
          if (gre(tj) == 0.0 && gim(tj) == 0.0) continue;
          const S *v = sRef + (j - jb) * m;
          const int e0 = (j + 1) >> 6;  // v vanishes on rows <= j
          S s[CG];
#pragma unroll
          for (int c = 0; c < CG; ++c) s[c] = gmk<S>(0.0, 0.0);
#pragma unroll
          for (int e = 0; e < PR; ++e) {
            const int i = lane + 64 * e;
            if (e >= e0 && e < ne && i < m) {
              const S vc = gcj(v[i]);
#pragma unroll
              for (int c = 0; c < CG; ++c) s[c] = s[c] + vc * z[c][e];
            }
          }
#pragma unroll
          for (int c = 0; c < CG; ++c) s[c] = tj * gmk<S>(wave_sum(gre(s[c])), wave_sum(gim(s[c])));
#pragma unroll
          for (int e = 0; e < PR; ++e) {
            const int i = lane + 64 * e;
            if (e >= e0 && e < ne && i < m) {
              const S vi = v[i];
#pragma unroll
              for (int c = 0; c < CG; ++c) z[c][e] = z[c][e] - s[c] * vi;
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < CG; ++c) {
        const int col = l0 + wave * CG + c;
#pragma unroll
        for (int e = 0; e < PR; ++e) {
          const int i = lane + 64 * e;
          if (col < k && i < m) Z[(int64_t)col * m + i] = z[c][e];
        }
      }
    }
    __syncthreads();
  }

  // ---------------- modified Gram-Schmidt ------------------------------------
  for (int i = 0; i < k; ++i) {
    S *qi = Z + (int64_t)i * m;
    double s = 0.0;
    for (int r = t; r < m; r += GT) s += abs2(qi[r]);
    s = block_sum(s, red);
    if (s < 0.25) {
      // a repeated eigenvalue repeated an earlier vector: take the unit vector
      // with the largest part outside span(q_0..q_{i-1}), orthogonalised twice
      double bestv = -2.0;
      int bp = 0;
      for (int r = t; r < m; r += GT) {
        double res = 1.0;
        for (int j = 0; j < i; ++j) res -= abs2(Z[(int64_t)j * m + r]);
        if (res > bestv) {
          bestv = res;
          bp = r;
        }
      }
      // block argmax (first index on ties); the barrier also ends block_sum's reads of red
      if (t == 0) sIdx = m;
      __syncthreads();
      const double wm = wave_max(bestv);
      if (lane == 0) red[wave] = wm;
      __syncthreads();
      const double gm = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
      if (bestv == gm) atomicMin(&sIdx, bp);
      __syncthreads();
      const int p = sIdx;
      for (int r = t; r < m; r += GT) qi[r] = gmk<S>(r == p ? 1.0 : 0.0, 0.0);
      __syncthreads();
      for (int pass = 0; pass < 2; ++pass)
        for (int j = 0; j < i; ++j) {
          const S *qj = Z + (int64_t)j * m;
          S d = gmk<S>(0.0, 0.0);
          for (int r = t; r < m; r += GT) d = d + gcj(qj[r]) * qi[r];
          d = gsum(d, red);
          for (int r = t; r < m; r += GT) qi[r] = qi[r] - d * qj[r];
          __syncthreads();
        }
      s = 0.0;
      for (int r = t; r < m; r += GT) s += abs2(qi[r]);
      s = block_sum(s, red);
    }
    const double inv = 1.0 / sqrt(s);
    for (int r = t; r < m; r += GT) qi[r] = gscl(inv, qi[r]);
    __syncthreads();
    // remove q_i from the later vectors (one wave per vector)
    for (int l = i + 1 + wave; l < k; l += GT / 64) {
      S *ql = Z + (int64_t)l * m;
      S d = gmk<S>(0.0, 0.0);
      for (int r = lane; r < m; r += 64) d = d + gcj(qi[r]) * ql[r];
      d = gmk<S>(wave_sum(gre(d)), wave_sum(gim(d)));
      for (int r = lane; r < m; r += 64) ql[r] = ql[r] - d * qi[r];
    }
    __syncthreads();
  }

  // ---------------- outputs: U and normalised carry = U^H X -------------------
  for (int idx = t; idx < m * k; idx += GT) {
    const int i = idx / k, l = idx - i * k;
    const S u = Z[(int64_t)l * m + i];
    if (DM)
      co[(int64_t)l * m + i] = from_acc<T>(u);
    else
      uo[idx] = from_acc<T>(u);
  }
  // carry[l][c] = sum_i conj(U[i][l]) X[i][c]: a norm pass, then the products
  // again for the normalised write (no k x n buffer, one rounding to T). A
  // thread forms CL rows of one column c, so X is read once per CL rows.
  const int kb = (k + CL - 1) / CL;
  auto carry_block = [&](int item, S (&acc)[CL]) {
    const int lb = item / n, c = item - lb * n;
    const S *ul[CL];
#pragma unroll
    for (int q = 0; q < CL; ++q) {
      acc[q] = gmk<S>(0.0, 0.0);
      ul[q] = Z + (int64_t)min(lb * CL + q, k - 1) * m;  // rows past k are computed and dropped
    }
    for (int i = 0; i < m; ++i) {
      const S x = X(i, c);
#pragma unroll
      for (int q = 0; q < CL; ++q) acc[q] = acc[q] + gcj(ul[q][i]) * x;
    }
  };
  double ss = 0.0;
  for (int item = t; item < kb * n; item += GT) {
    S acc[CL];
    carry_block(item, acc);
    const int lb = item / n;
#pragma unroll
    for (int q = 0; q < CL; ++q)
      if (lb * CL + q < k) ss += abs2(acc[q]);
  }
  ss = block_sum(ss, red);
  const double nrm = sqrt(ss);
  const double sc = nrm > 0.0 ? 1.0 / nrm : 1.0;
  for (int item = t; item < kb * n; item += GT) {
    S acc[CL];
    carry_block(item, acc);
    const int lb = item / n, c = item - lb * n;
#pragma unroll
    for (int q = 0; q < CL; ++q) {
      const int l = lb * CL + q;
      if (l >= k) continue;
      if (DM)
        uo[(int64_t)c * k + l] = from_acc<T>(gscl(sc, acc[q]));
      else
        co[(int64_t)l * n + c] = from_acc<T>(gscl(sc, acc[q]));
    }
  }
  if (t == 0) a.logscale[(int64_t)b * a.log_stride] += log(nrm);
}

template <typename T, typename S>
void launch_g(const TruncArgs<T> &a, int batch, hipStream_t stream) {
  if (a.dm)
    hipLaunchKernelGGL((k_trunc_g<T, S, true>), dim3(batch), dim3(GT), 0, stream, a);
  else
    hipLaunchKernelGGL((k_trunc_g<T, S, false>), dim3(batch), dim3(GT), 0, stream, a);
}

}  // namespace

template <>
void launch_trunc_g<float>(const TruncArgs<float> &a, int batch, hipStream_t s) {
  launch_g<float, double>(a, batch, s);
}
template <>
void launch_trunc_g<double>(const TruncArgs<double> &a, int batch, hipStream_t s) {
  launch_g<double, double>(a, batch, s);
}
template <>
void launch_trunc_g<cf32>(const TruncArgs<cf32> &a, int batch, hipStream_t s) {
  launch_g<cf32, cf64>(a, batch, s);
}
template <>
void launch_trunc_g<cf64>(const TruncArgs<cf64> &a, int batch, hipStream_t s) {
  launch_g<cf64, cf64>(a, batch, s);
}

}  // namespace vmc
