// Zip-up truncation step on gfx950: one workgroup (512 threads, 8 wave64s)
// per sample computes the top-k left singular subspace of the theta matrix
// mat (m x n, m, n <= 128) entirely in FP64:
//
//   G = mat mat^T                       (Gram, register-resident)
//   G = Q T Q^T                         (Householder tridiagonalisation, LAPACK
//                                        dsytd2-style reflectors kept in registers)
//   lambda_{m-k..m-1}(T)                (Sturm-count multisection, LPE = 8 or 16
//                                        lanes per eigenvalue, LPE + 1 sub-intervals
//                                        per step; TruncArgs::variant picks LPE;
//                                        counts from the three-term minor recurrence,
//                                        one FMA on the dependency chain)
//   z_l = twisted-factorisation vector  (Dhillon/Parlett twisted LDL^T, O(m))
//   U = Q Z, modified Gram-Schmidt      (reflectors applied 16 at a time)
//   carry = U^T mat / |U^T mat|_F        (log of the norm added to the boundary scale)
//
// This replaces `jnp.linalg.svd(mat)` + keep-k of the reference zip-up
// (src/vmc/peps/common/strategy.py:142-148). The truncated boundary state
// U U^T theta equals the reference's U_k S_k V_k^H whenever sigma_k > sigma_{k+1};
// the gauge of U (signs, rotations inside degenerate blocks) may differ, which
// leaves every contracted amplitude and environment unchanged.
//
// Register layout: thread t owns G[rb + 32 r][cb + 16 c] (rb = t / 16,
// cb = t % 16, r < 4, c < 8), so the 16 lanes of a DPP row share a row block
// and column accesses of the 16 lanes hit consecutive LDS words. Every phase is
// latency bound; cross-lane sums use DPP rows, reciprocals on the recurrence
// critical paths use v_rcp_f64 plus one Newton step.
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

// Minimum waves per SIMD requested from the register allocator: 4 lets two
// 512-thread workgroups share a CU (128 VGPRs each).
#ifndef VMC_TRUNC_MIN_WAVES
#define VMC_TRUNC_MIN_WAVES 4
#endif

// Phase timestamps of workgroup 0 for offline tuning builds (-DVMC_TRUNC_PROBE);
// compiled out otherwise.
#ifdef VMC_TRUNC_PROBE
__device__ unsigned long long g_trunc_probe[16];
#define TRUNC_PROBE(i)                                                              \
  do {                                                                              \
    __syncthreads();                                                                \
    if (blockIdx.x == 0 && threadIdx.x == 0) g_trunc_probe[i] = wall_clock64();     \
  } while (0)
#else
#define TRUNC_PROBE(i) \
  do {                 \
  } while (0)
#endif

__device__ inline double guard_pivot(double d, double pivmin) {
  if (fabs(d) < pivmin) d = (d < 0.0) ? -pivmin : pivmin;
  return d;
}

// 1/q from v_rcp_f64 (about 24 correct bits on gfx950) refined by one Newton
// step (relative error below 2^-48). frcp(0) and frcp(+-inf) are NaN.
__device__ inline double frcp(double q) {
  const double r = __builtin_amdgcn_rcp(q);
  return fma(r, fma(-q, r, 1.0), r);
}

// Modified Gram-Schmidt of the k columns of sY (column-major, stride LDY,
// length len <= MMAX, 512 threads), in place: 16-lane group gid holds column gid, lane cb
// rows cb + 16 e, in y. Columns are taken in order; a column whose residual
// keeps less than a quarter of its norm (a repeated twisted vector, or a zero
// column) is replaced by the unit vector with the largest part outside the
// span of the earlier ones, orthogonalised twice. Inputs must be normalised.
__device__ inline void mgs_columns(double (&y)[8], int gid, int cb, int k, int len, double *sY) {
  __syncthreads();
  for (int i = 0; i < k; ++i) {
    if (gid == i) {
      double s = 0.0;
#pragma unroll
      for (int e = 0; e < 8; ++e) s = fma(y[e], y[e], s);
      s = group16_sum(s);
      if (s < 0.25) {
        // For the trailing zero cluster of a rank-deficient theta the span of
        // the earlier vectors holds every nonzero direction, so the result
        // lies in the null space.
        double best = -2.0;
        int bp = 0;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int row = cb + 16 * e;
          double res = 1.0;
          for (int j = 0; j < i; ++j) {
            const double q = sY[j * LDY + row];
            res = fma(-q, q, res);
          }
          if (row < len && res > best) {
            best = res;
            bp = row;
          }
        }
        const double mx = group_max<16>(best);
        const int p = (int)group_min<16>(best == mx ? (double)bp : 1e9);
#pragma unroll
        for (int e = 0; e < 8; ++e) y[e] = cb + 16 * e == p ? 1.0 : 0.0;
        for (int pass = 0; pass < 2; ++pass)
          for (int j = 0; j < i; ++j) {
            const double *q = sY + j * LDY + cb;
            double d = 0.0;
#pragma unroll
            for (int e = 0; e < 8; ++e) d = fma(q[16 * e], y[e], d);
            d = group16_sum(d);
#pragma unroll
            for (int e = 0; e < 8; ++e) y[e] = fma(-d, q[16 * e], y[e]);
          }
        s = 0.0;
#pragma unroll
        for (int e = 0; e < 8; ++e) s = fma(y[e], y[e], s);
        s = group16_sum(s);
      }
      const double inv = 1.0 / sqrt(s);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        y[e] *= inv;
        sY[i * LDY + cb + 16 * e] = y[e];
      }
    }
    __syncthreads();
    if (gid > i && gid < k) {
      const double *q = sY + i * LDY + cb;
      double s = 0.0;
#pragma unroll
      for (int e = 0; e < 8; ++e) s = fma(q[16 * e], y[e], s);
      s = group16_sum(s);
#pragma unroll
      for (int e = 0; e < 8; ++e) y[e] = fma(-s, q[16 * e], y[e]);
    }
  }
  __syncthreads();
}

typedef double d4 __attribute__((ext_vector_type(4)));

// cacc[mt] (+)= U^T A on v_mfma_f64_16x16x4 for the k <= 32 columns of U in
// sY (length rows) and A = rows x cols (cols <= 128), A(i, c) = mat[c * rows + i]
// when TR, else mat[i * cols + c]. Wave w owns result columns 16 w .. 16 w + 15
// and both 16-row tiles; A is staged 16 rows at a time in stg (16 x 128).
template <typename T, bool TR>
__device__ inline void ut_times(const T *__restrict__ mat, int rows, int cols, const double *sY, double *stg,
                                d4 (&cacc)[2]) {
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int rho = lane & 15, kq = lane >> 4;
  cacc[0] = d4{0.0, 0.0, 0.0, 0.0};
  cacc[1] = d4{0.0, 0.0, 0.0, 0.0};
  for (int i0 = 0; i0 < rows; i0 += 16) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = t + NT * e;
      const int ii = idx >> 7, c = idx & 127;
      stg[idx] = (i0 + ii < rows && c < cols)
                     ? (double)mat[TR ? (int64_t)c * rows + i0 + ii : (int64_t)(i0 + ii) * cols + c]
                     : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k4 = 0; k4 < 4; ++k4) {
      const int ii = 4 * k4 + kq;
      const double bv = stg[ii * 128 + 16 * wave + rho];
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
        cacc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(sY[(16 * mt + rho) * LDY + i0 + ii], bv, cacc[mt], 0, 0, 0);
    }
    __syncthreads();
  }
}

// Second half of a zip-up cut of theta (m x n, k < n < m) through the smaller
// Gram theta^T theta: k_trunc<T, true, LPE> on X = theta^T has written
// u = W / |W|_F with W = theta V_k (m x k, V_k the top-k eigenvectors of
// theta^T theta, so range(W) is the top-k left singular subspace of theta)
// and carry = V_k^T. Here W's columns are orthonormalised into U in place and
// carry = U^T theta / |U^T theta|_F (k x n) is recomputed; the log-scale
// increment log |W|_F of the first kernel equals log |U^T theta|_F (both are
// |S_k|_F) up to rounding.
template <typename T>
__global__ __launch_bounds__(NT, 2) void k_zipt_post(TruncArgs<T> a) {
  __shared__ double sY[KMAX * LDY];
  __shared__ double stg[16 * 128];
  __shared__ double sRed[16];
  const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int gid = t >> 4, cb = t & 15;
  const int m = a.m, n = a.n, k = a.k;
  const T *__restrict__ mat = a.mat + (int64_t)b * a.mat_stride;
  T *__restrict__ uo = a.u + (int64_t)b * a.u_stride;
  T *__restrict__ co = a.carry + (int64_t)b * a.carry_stride;
  for (int idx = t; idx < KMAX * MMAX; idx += NT) {
    const int l = idx / MMAX, i = idx - l * MMAX;
    sY[l * LDY + i] = (l < k && i < m) ? (double)uo[(int64_t)i * k + l] : 0.0;
  }
  __syncthreads();
  // each column normalised by its group; a column below the rounding level of
  // the largest (column 0: rank deficiency) stays zero and mgs_columns completes
  // the basis with a unit vector
  double w[8], s0 = 0.0, s00 = 0.0;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    w[e] = gid < k ? sY[gid * LDY + cb + 16 * e] : 0.0;
    s0 = fma(w[e], w[e], s0);
    const double w0 = sY[cb + 16 * e];
    s00 = fma(w0, w0, s00);
  }
  s0 = group16_sum(s0);
  s00 = group16_sum(s00);
  const double sc0 = s0 > 1e-28 * s00 && s0 > 0.0 ? 1.0 / sqrt(s0) : 0.0;
#pragma unroll
  for (int e = 0; e < 8; ++e) w[e] *= sc0;
  mgs_columns(w, gid, cb, k, m, sY);
  for (int idx = t; idx < m * k; idx += NT) {
    const int i = idx / k, l = idx - i * k;
    uo[idx] = (T)sY[l * LDY + i];
  }
  d4 cacc[2];
  ut_times<T, false>(mat, m, n, sY, stg, cacc);
  const int rho = lane & 15, kq = lane >> 4, nn = 16 * wave + rho;
  double ss = 0.0;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (nn < n && 16 * mt + kq + 4 * r < k) ss = fma(cacc[mt][r], cacc[mt][r], ss);
  ss = block_sum(ss, sRed);
  const double sc = ss > 0.0 ? 1.0 / sqrt(ss) : 1.0;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int l = 16 * mt + kq + 4 * r;
      if (nn < n && l < k) co[(int64_t)l * n + nn] = (T)(cacc[mt][r] * sc);
    }
}

// DM: density-matrix cut (TruncArgs::dm), a compile-time switch so the zip-up
// loads and stores carry no extra branches. LPE: lanes (Sturm counts per step)
// per eigenvalue; 8 issues the multisection from 4 of the 8 waves, 16 from all.
template <typename T, bool DM, int LPE, int NW>
__global__ __launch_bounds__(64 * NW, VMC_TRUNC_MIN_WAVES) void k_trunc(TruncArgs<T> a) {
  // NW waves: the register tile covers MMAX = 16 NW rows (row stride RS = 4 NW,
  // NB = NW column blocks of 16); each 16-lane group carries CPG kept columns
  // through the back-transform and Gram-Schmidt; each wave TPW 16-column carry
  // tiles (n <= 128)
  constexpr int NT = 64 * NW, MMAX = 16 * NW, RS = 4 * NW, NB = NW, LDY = MMAX + 1;
  constexpr int CPG = KMAX / RS, TPW = 8 / NW, EPT = 16 * 128 / NT;
  static_assert(NW == 4 || NW == 8, "workgroup width");
  static_assert(LPE * KMAX / 64 <= NW, "multisection lanes");
  __shared__ double sBig[KMAX * LDY];  // Gram staging | Dm of twisted solves | reflector block | mat rows
  __shared__ double sY[KMAX * LDY];    // Dp / eigenvectors of T / orthonormal U (column-major)
  __shared__ double sX[MMAX], sP[MMAX], sV[MMAX], sAl;
  __shared__ double sD[MMAX], sE[MMAX], sE2[MMAX], sTau[MMAX];
  __shared__ double sLam[KMAX], sInv[KMAX];
  __shared__ double sRed[3 * 16];

  const int b = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int rb = t >> 4, cb = t & 15;
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
    // density-matrix: V = I, so the new tensor is theta itself (normalised)
    // and the carry the k x k identity
    T *__restrict__ eye = DM ? co : uo;
    T *__restrict__ body = DM ? uo : co;
    for (int i = t; i < m * k; i += NT) {
      const int r = i / k, c = i - r * k;
      eye[i] = (T)(r == c ? 1.0 : 0.0);
    }
    for (int i = t; i < m * n; i += NT) body[i] = (T)((double)mat[i] * sc);
    if (t == 0) a.logscale[(int64_t)b * a.log_stride] += log(nrm);
    return;
  }

  TRUNC_PROBE(0);
  // ---------------- Gram G = mat mat^T (FP64 MFMA) ---------------------------
  // v_mfma_f64_16x16x4: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]
  // and receives C[l / 16 + 4 r][l % 16], r < 4. Wave w takes the A rows
  // 4w + (rho & 3) + RS (rho >> 2) and one MFMA tile per register column c, so
  // C lands exactly in g[r][c] = G[rb + RS r][cb + 16 c].
  double g[4][NB];
  {
    typedef double d4 __attribute__((ext_vector_type(4)));
    d4 acc[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) acc[c] = d4{0.0, 0.0, 0.0, 0.0};
    const int rho = lane & 15, kq = lane >> 4;
    const int arow = 4 * wave + (rho & 3) + RS * (rho >> 2);
    // each thread stages 4 elements of the MMAX x 16 chunk; the next chunk's
    // global loads are issued before the current chunk is consumed
    double pre[4];
    auto load = [&](int n0) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int idx = t + NT * e;
        const int i = idx >> 4, c = idx & 15;
        pre[e] = (i < m && n0 + c < n) ? (double)mat[DM ? (int64_t)(n0 + c) * m + i : (int64_t)i * n + n0 + c]
                                       : 0.0;
      }
    };
    load(0);
    for (int n0 = 0; n0 < n; n0 += NCH) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int idx = t + NT * e;
        sBig[(idx >> 4) * STG + (idx & 15)] = pre[e];
      }
      __syncthreads();
      if (n0 + NCH < n) load(n0 + NCH);
#pragma unroll
      for (int k4 = 0; k4 < NCH / 4; ++k4) {
        const int kk = 4 * k4 + kq;
        const double av = sBig[arow * STG + kk];
#pragma unroll
        for (int c = 0; c < NB; ++c)
          acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, sBig[(rho + 16 * c) * STG + kk], acc[c], 0, 0, 0);
      }
      __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < NB; ++c) g[r][c] = acc[c][r];
  }

  TRUNC_PROBE(1);
  // ---------------- Householder tridiagonalisation ---------------------------
  // Column j = 16 c + jo lives in register column c of the lanes with cb == jo;
  // c is unrolled so register indices stay compile-time. Per column j:
  //   the owner lanes publish X' = G[j+2.., j] (zeros elsewhere; rows >= m of
  //   G are zero) and alpha = G[j+1, j];
  //   barrier; every 16-lane group forms beta, tau, scal from |X'|^2 (butterfly
  //   sums: bitwise-identical in all groups) and P = tau G v on its rows as
  //   tau (G[:, j+1] + scal G X'), v = e_{j+1} + scal X'; lanes 0..15 publish v;
  //   barrier; every lane reads v and P at its positions and applies
  //   G -= v w^T + w v^T, w = P - (tau / 2)(v^T P) v, to rows / columns > j,
  //   then column j keeps v below its subdiagonal for the back-transform.
  // X', alpha, P and v are each written between two barriers that bracket all
  // of their reads, so single buffers suffice.
  {
    const int row_hi = 4 * wave + 3;  // wave w owns rows 4w..4w+3 (+RS r)
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      for (int jo = 0; jo < 16; ++jo) {
        const int j = 16 * c + jo;
        if (j >= m - 2) break;
        if (cb == jo) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = rb + RS * r;
            sX[i] = i > j + 1 ? g[r][c] : 0.0;
            if (i == j + 1) sAl = g[r][c];
            if (i == j) sD[j] = g[r][c];
          }
        }
        __syncthreads();
        const double al = sAl;
        double xc[NB];
        double s = 0.0;
#pragma unroll
        for (int cc = 0; cc < NB; ++cc) {
          xc[cc] = sX[cb + 16 * cc];
          s = fma(xc[cc], xc[cc], s);
        }
        s = group16_sum(s);
        double tau = 0.0, beta = al, scal = 0.0;
        if (s > 0.0) {
          beta = -copysign(sqrt(fma(al, al, s)), al);
          tau = (beta - al) * frcp(beta);
          scal = frcp(al - beta);
        }
        if (t == 0) {
          sTau[j] = tau;
          sE[j] = beta;
        }
        // P = tau (G[:, j+1] + scal G X') on rows > j; X' vanishes at columns <= j + 1
        const int j1o = (j + 1) & 15;
        double acc[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          acc[r] = 0.0;
          if (row_hi + RS * r > j) {
            double p = 0.0;
#pragma unroll
            for (int cc = 0; cc < NB; ++cc)
              if (16 * cc + 15 > j + 1) p = fma(g[r][cc], xc[cc], p);
            const double gj1 = jo == 15 ? g[r][c < NB - 1 ? c + 1 : c] : g[r][c];
            acc[r] = group16_sum(fma(scal, p, cb == j1o ? gj1 : 0.0));
          }
        }
        if (cb == 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = rb + RS * r;
            sP[i] = i > j ? tau * acc[r] : 0.0;
          }
        }
        if (t < 16) {
#pragma unroll
          for (int cc = 0; cc < NB; ++cc) {
            const int i = cb + 16 * cc;
            sV[i] = i == j + 1 ? 1.0 : scal * xc[cc];
          }
        }
        __syncthreads();
        double pc[NB], vc[NB], pr[4], vr[4];
#pragma unroll
        for (int cc = 0; cc < NB; ++cc) {
          pc[cc] = sP[cb + 16 * cc];
          vc[cc] = sV[cb + 16 * cc];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          pr[r] = sP[rb + RS * r];
          vr[r] = sV[rb + RS * r];
        }
        double dp = 0.0;
#pragma unroll
        for (int cc = 0; cc < NB; ++cc) dp = fma(vc[cc], pc[cc], dp);
        const double hc = 0.5 * tau * group16_sum(dp);
#pragma unroll
        for (int cc = 0; cc < NB; ++cc) pc[cc] = fma(-hc, vc[cc], pc[cc]);  // w at columns
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (row_hi + RS * r <= j) continue;
          const double wr = fma(-hc, vr[r], pr[r]);
#pragma unroll
          for (int cc = 0; cc < NB; ++cc)
            if (16 * cc + 15 > j) g[r][cc] = fma(-vr[r], pc[cc], fma(-wr, vc[cc], g[r][cc]));
        }
        // column j: v (0 at rows <= j, 1 at j + 1) replaces entries no later step reads
        if (cb == jo) {
#pragma unroll
          for (int r = 0; r < 4; ++r) g[r][c] = vr[r];
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const int i = rb + RS * r, jj = cb + 16 * c;
      if (i == jj && i >= m - 2 && i < m) sD[i] = g[r][c];
      if (m >= 2 && i == m - 1 && jj == m - 2) sE[m - 2] = g[r][c];
    }
  __syncthreads();

  TRUNC_PROBE(2);
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

  // LPE lanes per eigenvalue, each evaluating NPC Sturm counts per step, so an
  // interval shrinks (LPE * NPC + 1)-fold per step. The phase issues only from
  // the first LPE * KMAX / 64 waves; the others (LPE = 8) wait at the barrier,
  // leaving their SIMD slots to the co-resident workgroup.
  {
    constexpr int NPC = 1;  // interleaved counts per lane (measured: 2 and 3 are no faster)
    constexpr double SEC = 1.0 / (LPE * NPC + 1);
    const int eg = t / LPE, sub = t % LPE;
    if (wave < LPE * KMAX / 64) {
      const bool active = eg < k;
      const int idx = m - k + eg;
      double lo = glo, hi = ghi;
      for (int it = 0; it < 200; ++it) {
        const bool conv =
            !active || (hi - lo) <= fmax(2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)), abstol);
        if (__all(conv)) break;
        if (!conv) {
          double x[NPC], p0[NPC], p1[NPC];
          int cnt[NPC], bad[NPC];
#pragma unroll
          for (int j = 0; j < NPC; ++j) {
            x[j] = lo + (hi - lo) * (double)(NPC * sub + j + 1) * SEC;
            p0[j] = 1.0;
            p1[j] = sD[0] - x[j];
            cnt[j] = p1[j] < 0.0;
            bad[j] = p1[j] == 0.0;
          }
          // leading principal minors p_{i+1} = (d_i - x) p_i - e2_{i-1} p_{i-1}:
          // the eigenvalues below x are the sign changes; rescaled by powers of
          // two every 8 steps. An exact zero or a non-finite minor sends the
          // count to the guarded pivot recurrence (LAPACK dlaneg's scheme).
          int i0 = 1;
          for (; i0 + 8 <= m; i0 += 8) {
            double dv[8], ev[8];
#pragma unroll
            for (int ii = 0; ii < 8; ++ii) {
              dv[ii] = sD[i0 + ii];
              ev[ii] = sE2[i0 + ii - 1];
            }
#pragma unroll
            for (int ii = 0; ii < 8; ++ii)
#pragma unroll
              for (int j = 0; j < NPC; ++j) {
                const double pn = fma(dv[ii] - x[j], p1[j], -ev[ii] * p0[j]);
                cnt[j] += (int)((unsigned)(__double2hiint(pn) ^ __double2hiint(p1[j])) >> 31);
                bad[j] |= pn == 0.0;
                p0[j] = p1[j];
                p1[j] = pn;
              }
#pragma unroll
            for (int j = 0; j < NPC; ++j) {
              const int ex = __builtin_amdgcn_frexp_exp(p1[j]);
              p1[j] = __builtin_amdgcn_ldexp(p1[j], -ex);
              p0[j] = __builtin_amdgcn_ldexp(p0[j], -ex);
            }
          }
          for (; i0 < m; ++i0) {
            const double di = sD[i0], ei = sE2[i0 - 1];
#pragma unroll
            for (int j = 0; j < NPC; ++j) {
              const double pn = fma(di - x[j], p1[j], -ei * p0[j]);
              cnt[j] += (int)((unsigned)(__double2hiint(pn) ^ __double2hiint(p1[j])) >> 31);
              bad[j] |= pn == 0.0;
              p0[j] = p1[j];
              p1[j] = pn;
            }
          }
          double cl = -INFINITY, ch = INFINITY;
#pragma unroll
          for (int j = 0; j < NPC; ++j) {
            int c = cnt[j];
            if (bad[j] || !isfinite(p1[j])) {
              double q = guard_pivot(sD[0] - x[j], pivmin);
              c = q <= 0.0;
              for (int i = 1; i < m; ++i) {
                q = fma(-sE2[i - 1], frcp(q), sD[i] - x[j]);
                if (fabs(q) < pivmin) q = -pivmin;
                c += q <= 0.0;
              }
            }
            if (c <= idx) cl = fmax(cl, x[j]);
            if (c > idx) ch = fmin(ch, x[j]);
          }
          cl = group_max<LPE>(cl);
          ch = group_min<LPE>(ch);
          lo = fmax(lo, cl);
          hi = fmin(hi, ch);
        }
      }
      if (active && sub == 0) sLam[k - 1 - eg] = 0.5 * (lo + hi);  // descending order
    }
  }
  __syncthreads();

  TRUNC_PROBE(3);
  // ---------------- eigenvectors of T: twisted factorisations ---------------
  // lane l of wave 0 solves for eigenvector l
  if (t < k) {
    const int gid = t;
    const double lam = sLam[gid];
    double *Dp = sY + gid * LDY;
    double *Dm = sBig + gid * LDY;
    double dp = sD[0] - lam;
    for (int i = 0; i < m - 1; ++i) {
      dp = guard_pivot(dp, pivmin);
      Dp[i] = dp;
      dp = fma(-sE2[i], frcp(dp), sD[i + 1] - lam);
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
      dm = fma(-sE2[i - 1], frcp(dm), sD[i - 1] - lam);
    }
    dm = guard_pivot(dm, pivmin);
    Dm[0] = dm;
    {
      const double gam = Dp[0] + dm - (sD[0] - lam);
      if (fabs(gam) < best) r = 0;
    }
    // z from the twist index outwards, both sides in one loop (no divergence
    // on r); the normalisation is applied where the back-transform loads z
    double nrm2 = 1.0, zl = 1.0, zh = 1.0;
    const int span = max(r, m - 1 - r);
    for (int d = 1; d <= span; ++d) {
      const int il = r - d, ih = r + d;
      if (il >= 0) {
        zl = -(sE[il] * frcp(Dp[il])) * zl;
        Dp[il] = zl;
        nrm2 = fma(zl, zl, nrm2);
      }
      if (ih < m) {
        zh = -(sE[ih - 1] * frcp(Dm[ih])) * zh;
        Dp[ih] = zh;
        nrm2 = fma(zh, zh, nrm2);
      }
    }
    Dp[r] = 1.0;
    sInv[gid] = 1.0 / sqrt(nrm2);
  }
  __syncthreads();

  TRUNC_PROBE(4);
  // ---------------- back-transform U = Q Z ----------------------------------
  // group gid holds columns gid + RS s (s < CPG) of Z; lane cb holds rows cb + 16 e.
  const int gid = rb;
  double y[CPG][NB];
#pragma unroll
  for (int q = 0; q < CPG; ++q) {
    const int col = gid + RS * q;
#pragma unroll
    for (int e = 0; e < NB; ++e)
      y[q][e] = (col < k && cb + 16 * e < m) ? sY[col * LDY + cb + 16 * e] * sInv[col] : 0.0;
  }
  const int nref = m - 2;
  double *sVb = sBig;  // 16 reflectors x MMAX
#pragma unroll
  for (int c = NB - 1; c >= 0; --c) {
    if (16 * c >= nref) continue;
    __syncthreads();
    {
      const int j = 16 * c + cb;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = rb + RS * r;
        double v = 0.0;
        if (j < nref) v = (i == j + 1) ? 1.0 : ((i > j + 1 && i < m) ? g[r][c] : 0.0);
        sVb[cb * MMAX + i] = v;
      }
    }
    __syncthreads();
    if (gid < k) {
      const int jhi = min(16 * c + 15, nref - 1);
      for (int j = jhi; j >= 16 * c; --j) {
        const double *vj = sVb + (j - 16 * c) * MMAX + cb;
        double s[CPG];
#pragma unroll
        for (int q = 0; q < CPG; ++q) {
          s[q] = 0.0;
#pragma unroll
          for (int e = 0; e < NB; ++e) s[q] = fma(vj[16 * e], y[q][e], s[q]);
        }
#pragma unroll
        for (int q = 0; q < CPG; ++q) {
          const double f = sTau[j] * group16_sum(s[q]);
#pragma unroll
          for (int e = 0; e < NB; ++e) y[q][e] = fma(-f, vj[16 * e], y[q][e]);
        }
      }
    }
  }

  TRUNC_PROBE(5);
  // ---------------- modified Gram-Schmidt ------------------------------------
  // column i lives in slot i / RS of group i % RS
  __syncthreads();
  for (int i = 0; i < k; ++i) {
    auto finish = [&](double(&yv)[NB]) {
      double s = 0.0;
#pragma unroll
      for (int e = 0; e < NB; ++e) s = fma(yv[e], yv[e], s);
      s = group16_sum(s);
      if (s < 0.25) {
        // Repeated eigenvalue (a split T, e.g. a cluster of exact zeros of a
        // rank-deficient theta): the twisted vector repeats an earlier one.
        // Take the unit vector with the largest part outside span(q_0..q_{i-1}),
        // orthogonalised twice; for the trailing zero cluster that span holds
        // every nonzero eigenvector, so the result lies in the null space.
        double best = -2.0;
        int bp = 0;
#pragma unroll
        for (int e = 0; e < NB; ++e) {
          const int row = cb + 16 * e;
          double res = 1.0;
          for (int j = 0; j < i; ++j) {
            const double q = sY[j * LDY + row];
            res = fma(-q, q, res);
          }
          if (row < m && res > best) {
            best = res;
            bp = row;
          }
        }
        const double mx = group_max<16>(best);
        const int p = (int)group_min<16>(best == mx ? (double)bp : 1e9);
#pragma unroll
        for (int e = 0; e < NB; ++e) yv[e] = cb + 16 * e == p ? 1.0 : 0.0;
        for (int pass = 0; pass < 2; ++pass)
          for (int j = 0; j < i; ++j) {
            const double *q = sY + j * LDY + cb;
            double d = 0.0;
#pragma unroll
            for (int e = 0; e < NB; ++e) d = fma(q[16 * e], yv[e], d);
            d = group16_sum(d);
#pragma unroll
            for (int e = 0; e < NB; ++e) yv[e] = fma(-d, q[16 * e], yv[e]);
          }
        s = 0.0;
#pragma unroll
        for (int e = 0; e < NB; ++e) s = fma(yv[e], yv[e], s);
        s = group16_sum(s);
      }
      const double inv = 1.0 / sqrt(s);
#pragma unroll
      for (int e = 0; e < NB; ++e) {
        yv[e] *= inv;
        sY[i * LDY + cb + 16 * e] = yv[e];
      }
    };
    if (gid == i % RS) {
      if (CPG == 1 || i < RS)
        finish(y[0]);
      else
        finish(y[CPG - 1]);
    }
    __syncthreads();
    const double *qi = sY + i * LDY + cb;
#pragma unroll
    for (int q = 0; q < CPG; ++q) {
      const int col = gid + RS * q;
      if (col > i && col < k) {
        double s = 0.0;
#pragma unroll
        for (int e = 0; e < NB; ++e) s = fma(qi[16 * e], y[q][e], s);
        s = group16_sum(s);
#pragma unroll
        for (int e = 0; e < NB; ++e) y[q][e] = fma(-s, qi[16 * e], y[q][e]);
      }
    }
  }
  __syncthreads();

  TRUNC_PROBE(6);
  // ---------------- outputs: U and normalised carry = U^T mat ---------------
  for (int idx = t; idx < m * k; idx += NT) {
    const int i = idx / k, l = idx - i * k;
    if (DM)
      co[(int64_t)l * m + i] = (T)sY[l * LDY + i];
    else
      uo[idx] = (T)sY[l * LDY + i];
  }
  // carry tiles on v_mfma_f64_16x16x4: wave w owns the 16-column tiles
  // w + NW tt (tt < TPW) and both 16-row tiles of the k <= 32 kept vectors;
  // A = U^T from sY, B = 16-row chunks of mat staged in LDS.
  typedef double d4 __attribute__((ext_vector_type(4)));
  d4 cacc[TPW][2];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) cacc[tt][0] = cacc[tt][1] = d4{0.0, 0.0, 0.0, 0.0};
  const int rho = lane & 15, kq = lane >> 4;
  for (int i0 = 0; i0 < m; i0 += 16) {
    // stage 16 rows of mat (16 x 128) in LDS
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int idx = t + NT * e;
      const int ii = idx >> 7, c = idx & 127;
      sBig[idx] = (i0 + ii < m && c < n)
                      ? (double)mat[DM ? (int64_t)c * m + i0 + ii : (int64_t)(i0 + ii) * n + c]
                      : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k4 = 0; k4 < 4; ++k4) {
      const int ii = 4 * k4 + kq;
#pragma unroll
      for (int tt = 0; tt < TPW; ++tt) {
        const double bv = sBig[ii * 128 + 16 * (wave + NW * tt) + rho];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
          cacc[tt][mt] =
              __builtin_amdgcn_mfma_f64_16x16x4f64(sY[(16 * mt + rho) * LDY + i0 + ii], bv, cacc[tt][mt], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  double ss = 0.0;
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (16 * (wave + NW * tt) + rho < n && 16 * mt + kq + 4 * r < k) ss = fma(cacc[tt][mt][r], cacc[tt][mt][r], ss);
  ss = block_sum(ss, sRed);
  const double nrm = sqrt(ss);
  const double sc = nrm > 0.0 ? 1.0 / nrm : 1.0;
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    const int nn = 16 * (wave + NW * tt) + rho;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int l = 16 * mt + kq + 4 * r;
        if (nn < n && l < k) {
          if (DM)
            uo[(int64_t)nn * k + l] = (T)(cacc[tt][mt][r] * sc);
          else
            co[(int64_t)l * n + nn] = (T)(cacc[tt][mt][r] * sc);
        }
      }
  }
  if (t == 0) a.logscale[(int64_t)b * a.log_stride] += log(nrm);
  TRUNC_PROBE(7);
}

// ---------------------------------------------------------------------------
// Cuts that discard nothing on the column side: k = n < m (zip-up only). The
// kept subspace is then all of range(mat), and any orthonormal basis of it is
// an exact gauge choice: Householder QR mat = Q R (LAPACK dgeqr2 reflectors,
// H = I - tau v v^T, v(j) = 1, beta = -sign(alpha) |x|), u = Q (m x k, formed
// as in dorg2r) and carry = R / |R|_F = U^T mat normalised. U U^T mat = mat,
// exactly as the reference's U_k S_k V_k^H with every singular value kept
// (src/vmc/peps/common/strategy.py:142-148).
//
// One wave64 per sample, four samples per workgroup. Lane l holds rows
// l + 64 r (r < NR) of all NC columns in registers; column sums are wave
// butterflies, so no LDS and no barriers.
template <typename T, int NR, int NC>
__global__ __launch_bounds__(256) void k_trunc_qr(TruncArgs<T> a, int batch) {
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (b >= batch) return;
  const int m = a.m, n = a.n;
  const T *__restrict__ mat = a.mat + (int64_t)b * a.mat_stride;
  T *__restrict__ uo = a.u + (int64_t)b * a.u_stride;
  T *__restrict__ co = a.carry + (int64_t)b * a.carry_stride;
  double x[NR][NC];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int i = lane + 64 * r;
#pragma unroll
    for (int c = 0; c < NC; ++c) x[r][c] = (i < m && c < n) ? (double)mat[(int64_t)i * n + c] : 0.0;
  }
  double tau[NC];
  // ---- factorisation: column j gets v below the diagonal and R above ----
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    tau[j] = 0.0;
    if (j >= n) continue;
    const double al = readlane_f64(x[j / 64][j], j % 64);
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (lane + 64 * r > j) s = fma(x[r][j], x[r][j], s);
    s = wave_sum(s);
    double beta = al;
    if (s > 0.0) {
      beta = -copysign(sqrt(fma(al, al, s)), al);
      tau[j] = (beta - al) / beta;
      const double scal = 1.0 / (al - beta);
      double v[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int i = lane + 64 * r;
        v[r] = i > j ? x[r][j] * scal : (i == j ? 1.0 : 0.0);
      }
      double w[NC];
#pragma unroll
      for (int c = j + 1; c < NC; ++c) {
        double d = 0.0;
#pragma unroll
        for (int r = 0; r < NR; ++r) d = fma(v[r], x[r][c], d);
        w[c] = wave_sum(d) * tau[j];
      }
#pragma unroll
      for (int c = j + 1; c < NC; ++c)
#pragma unroll
        for (int r = 0; r < NR; ++r) x[r][c] = fma(-w[c], v[r], x[r][c]);
#pragma unroll
      for (int r = 0; r < NR; ++r)
        if (lane + 64 * r > j) x[r][j] = v[r];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (lane + 64 * r == j) x[r][j] = beta;
  }
  // ---- carry = R / |R|_F (k = n rows) ----
  double ss = 0.0;
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (lane + 64 * r <= c) ss = fma(x[r][c], x[r][c], ss);
  ss = wave_sum(ss);
  const double nrm = sqrt(ss);
  const double sc = nrm > 0.0 ? 1.0 / nrm : 1.0;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int i = lane + 64 * r;
    if (i < n)
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (c < n) co[(int64_t)i * n + c] = (T)(i <= c ? x[r][c] * sc : 0.0);
  }
  if (lane == 0) a.logscale[(int64_t)b * a.log_stride] += log(nrm);
  // ---- u = H_0 ... H_{n-1} [I_n; 0] ----
  double q[NR][NC];
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int c = 0; c < NC; ++c) q[r][c] = lane + 64 * r == c ? 1.0 : 0.0;
#pragma unroll
  for (int j = NC - 1; j >= 0; --j) {
    if (j >= n || tau[j] == 0.0) continue;
    double v[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int i = lane + 64 * r;
      v[r] = i > j ? x[r][j] : (i == j ? 1.0 : 0.0);
    }
    double w[NC];
#pragma unroll
    for (int c = j; c < NC; ++c) {
      double d = 0.0;
#pragma unroll
      for (int r = 0; r < NR; ++r) d = fma(v[r], q[r][c], d);
      w[c] = wave_sum(d) * tau[j];
    }
#pragma unroll
    for (int c = j; c < NC; ++c)
#pragma unroll
      for (int r = 0; r < NR; ++r) q[r][c] = fma(-w[c], v[r], q[r][c]);
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int i = lane + 64 * r;
    if (i < m)
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (c < n) uo[(int64_t)i * n + c] = (T)q[r][c];
  }
}

template <typename T, int NR, int NC>
void launch_qr_nc(const TruncArgs<T> &a, int batch, hipStream_t stream) {
  hipLaunchKernelGGL((k_trunc_qr<T, NR, NC>), dim3((batch + 3) / 4), dim3(256), 0, stream, a, batch);
}

template <typename T, int NR>
void launch_qr(const TruncArgs<T> &a, int batch, hipStream_t stream) {
  if (a.n <= 4)
    launch_qr_nc<T, NR, 4>(a, batch, stream);
  else if (a.n <= 8)
    launch_qr_nc<T, NR, 8>(a, batch, stream);
  else if (a.n <= 16)
    launch_qr_nc<T, NR, 16>(a, batch, stream);
  else
    launch_qr_nc<T, 1, 32>(a, batch, stream);  // trunc_qr_shape: n > 16 only with m <= 64
}

// Factored matrices of at most 64 rows take the 4-wave kernel (four
// workgroups per CU); the 16-lane multisection needs all 8 waves.
template <typename T, bool DM>
void launch_lpe(const TruncArgs<T> &a, int batch, hipStream_t stream) {
  if (a.m <= 64)
    hipLaunchKernelGGL((k_trunc<T, DM, 8, 4>), dim3(batch), dim3(256), 0, stream, a);
  else if (a.variant == TRUNC_LPE16)
    hipLaunchKernelGGL((k_trunc<T, DM, 16, 8>), dim3(batch), dim3(512), 0, stream, a);
  else
    hipLaunchKernelGGL((k_trunc<T, DM, 8, 8>), dim3(batch), dim3(512), 0, stream, a);
}

template <typename T>
void launch_real(const TruncArgs<T> &a, int batch, hipStream_t stream) {
  if (trunc_qr_shape(false, a.m, a.n, a.k, a.dm)) {
    if (a.m <= 64)
      launch_qr<T, 1>(a, batch, stream);
    else
      launch_qr<T, 2>(a, batch, stream);
    return;
  }
  if (trunc_zipt_shape(false, a.m, a.n, a.k, a.dm)) {
    // density-matrix form on X = theta^T (its m, n are the factored dims),
    // then orthonormalise W = theta V_k and recompute the carry
    TruncArgs<T> x = a;
    x.m = a.n;
    x.n = a.m;
    x.dm = 1;
    launch_lpe<T, true>(x, batch, stream);
    hipLaunchKernelGGL(k_zipt_post<T>, dim3(batch), dim3(NT), 0, stream, a);
    return;
  }
  if (a.dm)
    launch_lpe<T, true>(a, batch, stream);
  else
    launch_lpe<T, false>(a, batch, stream);
}

}  // namespace

template <>
void launch_trunc<float>(const TruncArgs<float> &a, int batch, hipStream_t stream) {
  launch_real<float>(a, batch, stream);
}

template <>
void launch_trunc<double>(const TruncArgs<double> &a, int batch, hipStream_t stream) {
  launch_real<double>(a, batch, stream);
}

}  // namespace vmc
