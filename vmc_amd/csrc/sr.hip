// Stochastic-reconfiguration solve on gfx950, FP64 (real) or complex128.
//
// Mirrors SRPreconditioner.apply with DirectSolve (Cholesky) in parameter or
// sample space (src/vmc/preconditioners/preconditioners.py:117-163, 268-330)
// over the sliced Jacobian of src/vmc/qgt/jacobian.py:56-93 / qgt.py:173-223:
//
//   O_exp[i, (s, k, j)] = o[i, (s, j)] if p[i, (s, j)] == k else 0   (site-major)
//   O_c = O_exp - mean_i O_exp                                        (centring)
//   ParameterSpace:  (O_c^H O_c / Ns + eps I) x = O_c^H dv
//   SampleSpace:     (O_c O_c^H / Ns + eps I) y = dv,   x = O_c^H y
//   dv = grad_factor * (E - mean E) / Ns
//   real parameters: x <- 2 Re x for a real grad_factor, Re x otherwise
//
// The reference forms the uncentred Gram and subtracts mean terms
// (qgt.py:206-223); centring O first is the same matrix in exact arithmetic.
// Columns are built directly in the site-major order the parameter tree uses,
// which is the reference's SliceOrdering result after _reorder_updates
// (preconditioners.py:61-86) and its SiteOrdering result as is.
//
// Every kernel is a template over the working scalar: double when o, E and the
// grad factor are all real, cf64 (complex128) otherwise. Complex matrices are
// Hermitian; the Cholesky factor satisfies A = L L^H.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vmc_amd.h"
#include "common.hpp"

namespace vmc {
int sr_fail(const std::string &msg);
}

namespace {

using vmc::cf64;
using vmc::vmc_d4;

constexpr int TB = 64;  // GEMM tile
constexpr int NB = 32;  // Cholesky panel

__host__ __device__ inline double cj(double v) { return v; }
__host__ __device__ inline cf64 cj(cf64 v) { return {v.x, -v.y}; }
__host__ __device__ inline double re(double v) { return v; }
__host__ __device__ inline double re(cf64 v) { return v.x; }
__host__ __device__ inline double imv(double) { return 0.0; }
__host__ __device__ inline double imv(cf64 v) { return v.y; }
__host__ __device__ inline double sq(double v) { return v * v; }
__host__ __device__ inline double sq(cf64 v) { return v.x * v.x + v.y * v.y; }
__host__ __device__ inline double scl(double s, double v) { return s * v; }
__host__ __device__ inline cf64 scl(double s, cf64 v) { return {s * v.x, s * v.y}; }
template <typename T>
__host__ __device__ inline T mkT(double r, double i);
template <>
__host__ __device__ inline double mkT<double>(double r, double) { return r; }
template <>
__host__ __device__ inline cf64 mkT<cf64>(double r, double i) { return {r, i}; }

// Read element `idx` of a log-derivative buffer of vmc_dtype `dt`.
template <typename T>
__device__ inline T load_o(const void *o, int dt, int64_t idx) {
  switch (dt) {
    case VMC_F32: return mkT<T>((double)((const float *)o)[idx], 0.0);
    case VMC_F64: return mkT<T>(((const double *)o)[idx], 0.0);
    case VMC_C64: {
      const float2 v = ((const float2 *)o)[idx];
      return mkT<T>((double)v.x, (double)v.y);
    }
    default: {
      const double2 v = ((const double2 *)o)[idx];
      return mkT<T>(v.x, v.y);
    }
  }
}

// Column c of the site-major expanded Jacobian -> (source column q, slice k).
template <typename T>
__global__ void k_expand(const void *o, int o_dtype, const int8_t *p, const int32_t *col_src, const int8_t *col_k,
                         int ns, int64_t np, int64_t npx, T *oc) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (col >= npx) return;
  const int64_t q = col_src[col];
  const int8_t k = col_k[col];
  const int64_t src = (int64_t)i * np + q;
  oc[(int64_t)i * npx + col] = (p[src] == k) ? load_o<T>(o, o_dtype, src) : T(0.0);
}

// Column means, then centring in place.
template <typename T>
__global__ void k_colmean(const T *oc, int ns, int64_t npx, T *mean) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= npx) return;
  T s = 0.0;
  for (int i = 0; i < ns; ++i) s = s + oc[(int64_t)i * npx + col];
  mean[col] = scl(1.0 / ns, s);
}

template <typename T>
__global__ void k_center(T *oc, int ns, int64_t npx, const T *mean) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (col >= npx) return;
  T &v = oc[(int64_t)i * npx + col];
  v = v - mean[col];
}

__device__ inline double wave_red(double s) {
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  return s;
}

// dv = gf * (E - mean E) / Ns (single workgroup of 256). E is real or complex.
template <typename T>
__global__ __launch_bounds__(256) void k_dv(const double *e, int e_cplx, int ns, double gf_re, double gf_im,
                                            T *dv) {
  __shared__ double red[2][4];
  double sr = 0.0, si = 0.0;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    sr += e_cplx ? e[2 * i] : e[i];
    si += e_cplx ? e[2 * i + 1] : 0.0;
  }
  sr = wave_red(sr);
  si = wave_red(si);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = sr;
    red[1][threadIdx.x >> 6] = si;
  }
  __syncthreads();
  const double mr = (red[0][0] + red[0][1] + red[0][2] + red[0][3]) / ns;
  const double mi = (red[1][0] + red[1][1] + red[1][2] + red[1][3]) / ns;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    const double a = ((e_cplx ? e[2 * i] : e[i]) - mr) / ns;
    const double b = ((e_cplx ? e[2 * i + 1] : 0.0) - mi) / ns;
    dv[i] = mkT<T>(gf_re * a - gf_im * b, gf_re * b + gf_im * a);
  }
}

// C[i, j] = alpha * sum_k opA(A[i*sai + k*sak]) * opB(B[j*sbj + k*sbk]) + beta * C[i, j]
// with op = conj when CA / CB (C row-major with leading dimension ldc).
// 64x64 tiles, 256 threads, 4x4 per thread, K staged through LDS in chunks of 16.
template <typename T, bool CA, bool CB>
__global__ __launch_bounds__(256) void k_gemm(int m, int n, int64_t kdim, double alpha, const T *A, int64_t sai,
                                              int64_t sak, const T *B, int64_t sbj, int64_t sbk, double beta, T *C,
                                              int64_t ldc) {
  __shared__ T sA[16][TB + 1];
  __shared__ T sB[16][TB + 1];
  const int t = threadIdx.x;
  const int ti = t >> 4, tj = t & 15;
  const int i0 = blockIdx.y * TB, j0 = blockIdx.x * TB;
  T acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
  for (int64_t k0 = 0; k0 < kdim; k0 += 16) {
    for (int idx = t; idx < 16 * TB; idx += 256) {
      // consecutive lanes walk the operand's contiguous axis
      int kk = idx / TB, r = idx - kk * TB;
      if (sak == 1) {
        r = idx >> 4;
        kk = idx & 15;
      }
      const int64_t k = k0 + kk;
      T a = (i0 + r < m && k < kdim) ? A[(int64_t)(i0 + r) * sai + k * sak] : T(0.0);
      sA[kk][r] = CA ? cj(a) : a;
      int kb = idx / TB, rb = idx - kb * TB;
      if (sbk == 1) {
        rb = idx >> 4;
        kb = idx & 15;
      }
      const int64_t k2 = k0 + kb;
      T b = (j0 + rb < n && k2 < kdim) ? B[(int64_t)(j0 + rb) * sbj + k2 * sbk] : T(0.0);
      sB[kb][rb] = CB ? cj(b) : b;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
      T a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = sA[kk][ti + 16 * u];
        b[u] = sB[kk][tj + 16 * u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + ti + 16 * u, j = j0 + tj + 16 * v;
      if (i < m && j < n) {
        T *c = C + (int64_t)i * ldc + j;
        *c = scl(alpha, acc[u][v]) + (beta != 0.0 ? scl(beta, *c) : T(0.0));
      }
    }
}

template <typename T>
__global__ void k_add_diag(T *A, int n, int64_t lda, double eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[(int64_t)i * lda + i] = A[(int64_t)i * lda + i] + T(eps);
}

// Unblocked Cholesky of the NB x NB diagonal block at (j0, j0), in LDS.
// Hermitian: the diagonal is real, L[r][c] -= L[r][j] conj(L[c][j]).
template <typename T>
__global__ void k_potf2(T *A, int n, int64_t lda, int j0, int *info) {
  __shared__ T L[NB][NB + 1];
  const int nb = min(NB, n - j0);
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int r = idx / NB, c = idx - r * NB;
    L[r][c] = (r < nb && c < nb) ? A[(int64_t)(j0 + r) * lda + j0 + c] : T(0.0);
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    if (threadIdx.x == 0) {
      const double djj = re(L[j][j]);
      if (!(djj > 0.0)) atomicExch(info, j0 + j + 1);
      L[j][j] = T(sqrt(fmax(djj, 1e-300)));
    }
    __syncthreads();
    const double inv = 1.0 / re(L[j][j]);
    for (int r = j + 1 + threadIdx.x; r < nb; r += blockDim.x) L[r][j] = scl(inv, L[r][j]);
    __syncthreads();
    for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
      const int r = idx / nb, c = idx - r * nb;
      if (c > j && r >= c) L[r][c] = L[r][c] - L[r][j] * cj(L[c][j]);
    }
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int r = idx / NB, c = idx - r * NB;
    if (r < nb && c <= r) A[(int64_t)(j0 + r) * lda + j0 + c] = L[r][c];
  }
}

// Panel below the diagonal block: A21 <- A21 L11^-H (one row per thread).
template <typename T>
__global__ void k_trsm_panel(T *A, int n, int64_t lda, int j0) {
  __shared__ T L[NB][NB + 1];
  const int nb = min(NB, n - j0);
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int r = idx / NB, c = idx - r * NB;
    L[r][c] = (r < nb && c <= r) ? A[(int64_t)(j0 + r) * lda + j0 + c] : T(0.0);
  }
  __syncthreads();
  const int row = j0 + nb + blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  T x[NB];
  T *a = A + (int64_t)row * lda + j0;
#pragma unroll
  for (int c = 0; c < NB; ++c) x[c] = c < nb ? a[c] : T(0.0);
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    if (c < nb) {
      T s = x[c];
#pragma unroll
      for (int q = 0; q < c; ++q) s = s - x[q] * cj(L[c][q]);
      x[c] = scl(1.0 / re(L[c][c]), s);
    }
  }
#pragma unroll
  for (int c = 0; c < NB; ++c)
    if (c < nb) a[c] = x[c];
}

// Blocked substitution with L (lower, lda) on many workgroups, b overwritten
// by the solution. The columns are cut into blocks of TS = 64; launch jb
// (ascending for L z = b, descending for L^H y = z) has workgroup 0 apply the
// previous block's update to block jb's rows and then solve block jb's
// diagonal triangle in one wave (lane r holds row r, the pivot is broadcast
// with a shuffle), while every other workgroup applies the previous block's
// update to 64 rows of the part still unsolved. Each launch reads only what
// the launch before it wrote, so the stream order is the only synchronisation.
constexpr int TS = 64;

__device__ inline double shfl_t(double v, int src) { return __shfl(v, src, 64); }
__device__ inline cf64 shfl_t(cf64 v, int src) { return {__shfl(v.x, src, 64), __shfl(v.y, src, 64)}; }
__device__ inline double xor_t(double v, int m) { return __shfl_xor(v, m, 64); }
__device__ inline cf64 xor_t(cf64 v, int m) { return {__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64)}; }

// sum of L[r, c0 + c] x[c] over c < nc for 64 rows r0.. (16 lanes per row,
// 4 adjacent columns per lane); lane 0 of each 16-lane group gets the sum
template <typename T>
__device__ inline void trsv_rows(const T *A, int n, int64_t lda, int r0, int c0, int nc, const T *sx, T *out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sub = lane >> 4, l16 = lane & 15;
  for (int it = 0; it < 4; ++it) {
    const int rl = wave * 16 + it * 4 + sub;
    const int r = r0 + rl;
    T s = 0.0;
    if (r < n) {
      const T *row = A + (int64_t)r * lda + c0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = 4 * l16 + q;
        if (c < nc) s = fma(row[c], sx[c], s);
      }
    }
    for (int m = 8; m >= 1; m >>= 1) s = s + xor_t(s, m);
    if (l16 == 0) out[rl] = s;
  }
}

// sum of conj(L[c0 + q, r]) x[q] over q < nq for 64 columns r0.. (lane r,
// wave w takes q in [16w, 16w + 16), partial sums reduced through LDS)
template <typename T>
__device__ inline void trsv_cols(const T *A, int n, int64_t lda, int r0, int c0, int nq, const T *sx, T *red,
                                 T *out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = r0 + lane;
  T s = 0.0;
  if (r < n)
    for (int q = 16 * wave; q < min(nq, 16 * wave + 16); ++q) s = fma(cj(A[(int64_t)(c0 + q) * lda + r]), sx[q], s);
  red[wave * 64 + lane] = s;
  __syncthreads();
  if (wave == 0) out[lane] = red[lane] + red[64 + lane] + red[128 + lane] + red[192 + lane];
}

template <typename T>
__global__ __launch_bounds__(256) void k_trsv_fwd(const T *A, int n, int64_t lda, T *b, int jb) {
  __shared__ T sx[TS], upd[TS], red[256];
  __shared__ T sL[TS][TS + 1];
  const int t = threadIdx.x;
  const int j0 = jb * TS, nb = min(TS, n - j0);
  const int p0 = j0 - TS;  // previous block (solved), if any
  if (jb > 0 && t < TS) sx[t] = b[p0 + t];
  if (blockIdx.x > 0) {
    const int r0 = j0 + TS + (blockIdx.x - 1) * 64;
    __syncthreads();
    trsv_rows(A, n, lda, r0, p0, TS, sx, upd);
    __syncthreads();
    if (t < 64 && r0 + t < n) b[r0 + t] = b[r0 + t] - upd[t];
    return;
  }
  (void)red;
  for (int idx = t; idx < TS * TS; idx += 256) {
    const int r = idx / TS, c = idx - r * TS;
    sL[r][c] = (r < nb && c <= r) ? A[(int64_t)(j0 + r) * lda + j0 + c] : T(0.0);
  }
  if (jb > 0) {
    __syncthreads();
    trsv_rows(A, n, lda, j0, p0, TS, sx, upd);
  }
  __syncthreads();
  if (t < 64) {
    T v = t < nb ? b[j0 + t] : T(0.0);
    if (jb > 0) v = v - upd[t];
    for (int c = 0; c < nb; ++c) {
      const T xc = scl(1.0 / re(sL[c][c]), shfl_t(v, c));
      if (t > c) v = v - sL[t][c] * xc;
      if (t == c) v = xc;
    }
    if (t < nb) b[j0 + t] = v;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_trsv_bwd(const T *A, int n, int64_t lda, T *b, int jb) {
  __shared__ T sx[TS], upd[TS], red[256];
  __shared__ T sL[TS][TS + 1];
  const int t = threadIdx.x;
  const int j0 = jb * TS, nb = min(TS, n - j0);
  const int p0 = j0 + TS, np = max(0, min(TS, n - p0));  // next block (solved), if any
  if (np > 0 && t < np) sx[t] = b[p0 + t];
  if (blockIdx.x > 0) {
    const int r0 = (blockIdx.x - 1) * 64;
    __syncthreads();
    trsv_cols(A, n, lda, r0, p0, np, sx, red, upd);
    __syncthreads();
    if (t < 64 && r0 + t < j0) b[r0 + t] = b[r0 + t] - upd[t];
    return;
  }
  for (int idx = t; idx < TS * TS; idx += 256) {
    const int r = idx / TS, c = idx - r * TS;
    sL[r][c] = (r < nb && c <= r) ? A[(int64_t)(j0 + r) * lda + j0 + c] : T(0.0);
  }
  if (np > 0) {
    __syncthreads();
    trsv_cols(A, n, lda, j0, p0, np, sx, red, upd);
  }
  __syncthreads();
  if (t < 64) {
    T v = t < nb ? b[j0 + t] : T(0.0);
    if (np > 0) v = v - upd[t];
    for (int c = nb - 1; c >= 0; --c) {
      const T xc = scl(1.0 / re(sL[c][c]), shfl_t(v, c));
      if (t < c) v = v - cj(sL[c][t]) * xc;
      if (t == c) v = xc;
    }
    if (t < nb) b[j0 + t] = v;
  }
}

// L z = b, then L^H y = z (b overwritten by y)
template <typename T>
void trsv_both(const T *A, int n, int64_t lda, T *b, hipStream_t st) {
  const int nblk = (n + TS - 1) / TS;
  for (int jb = 0; jb < nblk; ++jb) {
    const int below = n - (jb + 1) * TS;
    const int g = 1 + (jb > 0 && below > 0 ? (below + 63) / 64 : 0);
    hipLaunchKernelGGL(k_trsv_fwd<T>, dim3(g), dim3(256), 0, st, A, n, lda, b, jb);
  }
  for (int jb = nblk - 1; jb >= 0; --jb) {
    const int g = 1 + (jb < nblk - 1 ? jb : 0);
    hipLaunchKernelGGL(k_trsv_bwd<T>, dim3(g), dim3(256), 0, st, A, n, lda, b, jb);
  }
}

__global__ void k_info(const int *info, double *out) { *out = (double)*info; }

// Final update vector: complex / real copy, or factor * Re x for real parameters.
template <typename T>
__global__ void k_finish(const T *x, int64_t n, int real_out, double factor, void *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (real_out)
    ((double *)out)[i] = factor * re(x[i]);
  else
    ((T *)out)[i] = x[i];
}

// |g - h|^2 / |h|^2 (0 for h = 0), the residual_error metric
template <typename T>
__global__ __launch_bounds__(256) void k_resid_vec(const T *g, const T *h, int64_t n, double *out) {
  __shared__ double red[2][4];
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    a += sq(g[i] - h[i]);
    b += sq(h[i]);
  }
  a = wave_red(a);
  b = wave_red(b);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double na = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    const double nb = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    out[0] = nb > 0.0 ? na / nb : 0.0;
  }
}

// FP64 products on the matrix cores (v_mfma_f64_16x16x4f64), same operand
// convention as k_gemm with real operands: 128 x 128 tiles, 4 waves each
// owning a 64 x 64 quarter (4 x 4 blocks of 16 x 16), K staged through LDS in
// chunks of 16 while the next chunk is loaded into registers. Consecutive
// lanes walk whichever operand axis is contiguous in memory.
// lower: only tiles on or below the diagonal (tile row >= tile column), for
// the Gram matrices and the Cholesky trailing update, which read the lower
// triangle only; mirror = 1 also writes the final C[i, j] to C[j, i], mirror =
// 2 its negative (the imaginary part of a Hermitian result).
// Workgroups are spread over the 8 XCDs so that each XCD works through a
// contiguous run of tiles (shared operand panels stay in its L2).
constexpr int MT = 128, MKC = 16;

__device__ inline int xcd_run(int b, int nb) {
  const int x = b & 7, local = b >> 3, q = nb >> 3, rem = nb & 7;
  return x * q + min(x, rem) + local;
}

template <bool KA, bool KB>
__global__ __launch_bounds__(256) void k_gemm_mfma(int m, int n, int64_t kdim, double alpha, const double *A,
                                                   int64_t sai, int64_t sak, const double *B, int64_t sbj,
                                                   int64_t sbk, double beta, double *C, int64_t ldc, int lower,
                                                   int mirror, int64_t csc) {
  __shared__ double sA[MKC][MT + 2];
  __shared__ double sB[MKC][MT + 2];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int b = xcd_run(blockIdx.x, gridDim.x);
  int ti, tj;
  if (lower) {
    ti = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while (ti * (ti + 1) / 2 > b) --ti;
    while ((ti + 1) * (ti + 2) / 2 <= b) ++ti;
    tj = b - ti * (ti + 1) / 2;
  } else {
    const int ntj = (n + MT - 1) / MT;
    ti = b / ntj;
    tj = b - ti * ntj;
  }
  const int i0 = ti * MT, j0 = tj * MT;
  const int wr0 = (wave >> 1) * 64, wc0 = (wave & 1) * 64;
  constexpr bool ka = KA, kb = KB;  // K is the contiguous axis of A / B
  vmc_d4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = vmc_d4{0.0, 0.0, 0.0, 0.0};
  double pa[8], pb[8];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      const int idx = l * 256 + t;
      const int r = ka ? idx >> 4 : idx & (MT - 1), kk = ka ? idx & 15 : idx >> 7;
      const int64_t k = k0 + kk;
      pa[l] = (i0 + r < m && k < kdim) ? A[(int64_t)(i0 + r) * sai + k * sak] : 0.0;
      const int rb = kb ? idx >> 4 : idx & (MT - 1), kc = kb ? idx & 15 : idx >> 7;
      const int64_t k2 = k0 + kc;
      pb[l] = (j0 + rb < n && k2 < kdim) ? B[(int64_t)(j0 + rb) * sbj + k2 * sbk] : 0.0;
    }
  };
  fetch(0);
  for (int64_t k0 = 0; k0 < kdim; k0 += MKC) {
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      const int idx = l * 256 + t;
      sA[ka ? idx & 15 : idx >> 7][ka ? idx >> 4 : idx & (MT - 1)] = pa[l];
      sB[kb ? idx & 15 : idx >> 7][kb ? idx >> 4 : idx & (MT - 1)] = pb[l];
    }
    __syncthreads();
    if (k0 + MKC < kdim) fetch(k0 + MKC);
#pragma unroll
    for (int ks = 0; ks < MKC; ks += 4) {
      double av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = sA[ks + lk][wr0 + 16 * i + li];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = sB[ks + lk][wc0 + 16 * j + li];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // acc[i][j][r] holds C[i0 + wr0 + 16 i + lk + 4 r, j0 + wc0 + 16 j + li]
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int mm = i0 + wr0 + 16 * i + lk + 4 * r, nn = j0 + wc0 + 16 * j + li;
        if (mm < m && nn < n) {
          double v = alpha * acc[i][j][r];
          if (beta != 0.0) v += beta * C[(int64_t)mm * ldc + (int64_t)nn * csc];
          C[(int64_t)mm * ldc + (int64_t)nn * csc] = v;
          if (mirror && ti != tj) C[(int64_t)nn * ldc + (int64_t)mm * csc] = mirror == 2 ? -v : v;
        }
      }
}

void gemm_mfma(int nb, int m, int n, int64_t k, double alpha, const double *A, int64_t sai, int64_t sak,
               const double *B, int64_t sbj, int64_t sbk, double beta, double *C, int64_t ldc, int lower, int mirror,
               hipStream_t st, int64_t csc = 1) {
  // K "contiguous" also for stride 2: the real or imaginary parts of interleaved complex storage
  auto kern = sak <= 2 ? (sbk <= 2 ? k_gemm_mfma<true, true> : k_gemm_mfma<true, false>)
                       : (sbk <= 2 ? k_gemm_mfma<false, true> : k_gemm_mfma<false, false>);
  hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, st, m, n, k, alpha, A, sai, sak, B, sbj, sbk, beta, C, ldc, lower,
                     mirror, csc);
}

// Complex128 product C = alpha op(A) op(B)^T + beta C as four real MFMA
// products over the interleaved (re, im) storage (strides doubled, imaginary
// parts at offset 1): with op = conj when CA / CB (signs sa, sb),
//   Re C = Re A Re B^T - sa sb Im A Im B^T,  Im C = sb Re A Im B^T + sa Im A Re B^T.
// lower: only the tiles on and below the diagonal of a Hermitian result;
// mirror then fills the upper triangle (the real part symmetric, the
// imaginary part antisymmetric), else the upper triangle is not read.
template <bool CA, bool CB>
void gemm_mfma_c(int nb, int m, int n, int64_t k, double alpha, const cf64 *A, int64_t sai, int64_t sak,
                 const cf64 *B, int64_t sbj, int64_t sbk, double beta, cf64 *C, int64_t ldc, int lower, bool mirror,
                 hipStream_t st) {
  const double sa = CA ? -1.0 : 1.0, sb = CB ? -1.0 : 1.0;
  const double *a = reinterpret_cast<const double *>(A), *b = reinterpret_cast<const double *>(B);
  double *c = reinterpret_cast<double *>(C);
  const int64_t ai = 2 * sai, ak = 2 * sak, bj = 2 * sbj, bk = 2 * sbk, lc = 2 * ldc;
  const int mr = mirror ? 1 : 0, mi = mirror ? 2 : 0;
  gemm_mfma(nb, m, n, k, alpha, a, ai, ak, b, bj, bk, beta, c, lc, lower, mr, st, 2);
  gemm_mfma(nb, m, n, k, -sa * sb * alpha, a + 1, ai, ak, b + 1, bj, bk, 1.0, c, lc, lower, mr, st, 2);
  gemm_mfma(nb, m, n, k, sb * alpha, a, ai, ak, b + 1, bj, bk, beta, c + 1, lc, lower, mi, st, 2);
  gemm_mfma(nb, m, n, k, sa * alpha, a + 1, ai, ak, b, bj, bk, 1.0, c + 1, lc, lower, mi, st, 2);
}

// Real and complex128 products with both output sides at least 64 run on the
// matrix cores (complex128 as four real products); complex64 ones and the
// matrix-vector products stay on k_gemm.
template <typename T, bool CA, bool CB>
void gemm(int m, int n, int64_t k, double alpha, const T *A, int64_t sai, int64_t sak, const T *B, int64_t sbj,
          int64_t sbk, double beta, T *C, int64_t ldc, hipStream_t st) {
  if constexpr (sizeof(T) == 8 || std::is_same<T, cf64>::value) {
    if (m >= 64 && n >= 64) {
      const int nb = ((m + MT - 1) / MT) * ((n + MT - 1) / MT);
      if constexpr (std::is_same<T, cf64>::value)
        gemm_mfma_c<CA, CB>(nb, m, n, k, alpha, A, sai, sak, B, sbj, sbk, beta, C, ldc, 0, false, st);
      else
        gemm_mfma(nb, m, n, k, alpha, A, sai, sak, B, sbj, sbk, beta, C, ldc, 0, 0, st);
      return;
    }
  }
  dim3 grid((n + TB - 1) / TB, (m + TB - 1) / TB);
  hipLaunchKernelGGL((k_gemm<T, CA, CB>), grid, dim3(256), 0, st, m, n, k, alpha, A, sai, sak, B, sbj, sbk, beta, C,
                     ldc);
}

// C = alpha op(A) op(B)^T + beta C for an n x n result that is Hermitian
// (Gram matrices, the Cholesky trailing update): the tiles on and below the
// diagonal only (mirror fills the upper triangle as well); complex128 as four
// real products, complex64 on k_gemm over the whole matrix.
template <typename T, bool CA, bool CB>
void gemm_herm(int n, int64_t k, double alpha, const T *A, int64_t sai, int64_t sak, double beta, T *C, int64_t ldc,
               bool mirror, hipStream_t st) {
  if constexpr (sizeof(T) == 8) {
    if (n >= 64) {
      const int nt = (n + MT - 1) / MT;
      gemm_mfma(nt * (nt + 1) / 2, n, n, k, alpha, A, sai, sak, A, sai, sak, beta, C, ldc, 1, mirror ? 1 : 0, st);
      return;
    }
  } else if constexpr (std::is_same<T, cf64>::value) {
    if (n >= 64) {
      const int nt = (n + MT - 1) / MT;
      gemm_mfma_c<CA, CB>(nt * (nt + 1) / 2, n, n, k, alpha, A, sai, sak, A, sai, sak, beta, C, ldc, 1, mirror, st);
      return;
    }
  }
  gemm<T, CA, CB>(n, n, k, alpha, A, sai, sak, A, sai, sak, beta, C, ldc, st);
}

// In-place lower Cholesky of A (n x n); A[i,j] for j <= i is L afterwards.
// Two levels: panels of NB2 = 256 columns are factored NB = 32 columns at a
// time (potf2, trsm, then an update restricted to the panel's columns), and
// the rest of the matrix gets one rank-256 update per panel on the matrix
// cores, so the trailing matrix is read and written n / 256 times, not n / 32.
template <typename T>
void cholesky(T *A, int n, int64_t lda, int *info, hipStream_t st) {
  constexpr int NB2 = 256;
  for (int p0 = 0; p0 < n; p0 += NB2) {
    const int pend = std::min(n, p0 + NB2);
    for (int j0 = p0; j0 < pend; j0 += NB) {
      const int nb = std::min(NB, n - j0);
      hipLaunchKernelGGL(k_potf2<T>, dim3(1), dim3(256), 0, st, A, n, lda, j0, info);
      const int rest = n - j0 - nb;
      if (rest <= 0) break;
      hipLaunchKernelGGL(k_trsm_panel<T>, dim3((rest + 255) / 256), dim3(256), 0, st, A, n, lda, j0);
      // A[i, j] -= A21[i, :] A21[j, :]^H for rows below the block and the
      // panel's remaining columns j < pend
      const int pc = pend - (j0 + nb);
      T *A21 = A + (int64_t)(j0 + nb) * lda + j0;
      T *A22 = A + (int64_t)(j0 + nb) * lda + j0 + nb;
      if (pc > 0) gemm<T, false, true>(rest, pc, nb, -1.0, A21, lda, 1, A21, lda, 1, 1.0, A22, lda, st);
    }
    const int rest = n - pend;
    if (rest > 0) {
      // trailing update A22 -= A21 A21^H with the whole panel (K = pend - p0)
      T *A21 = A + (int64_t)pend * lda + p0;
      T *A22 = A + (int64_t)pend * lda + pend;
      gemm_herm<T, false, true>(rest, pend - p0, -1.0, A21, lda, 1, 1.0, A22, lda, false, st);
    }
  }
}

// ---- iterative and certified solves of M z = b (M = S + eps or T + eps) ----
// Re(a^H b) over n entries into *out, one workgroup.
template <typename T>
__global__ __launch_bounds__(256) void k_dot_re(const T *a, const T *b, int64_t n, double *out) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc += re(cj(a[i]) * b[i]);
  acc = wave_red(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) *out = red[0] + red[1] + red[2] + red[3];
}

// Conjugate gradients (jax.scipy.sparse.linalg.cg as called by solve_cg,
// solvers.py:16-21: x0 = 0, stop when r^H r <= tol^2 |b|^2 or after maxiter
// steps). Real scalars in s: s[0] = r^H r, s[1] = p^H M p, s[2] = the new
// r^H r, s[3] = tol^2 |b|^2, s[4] = steps taken. Every update is predicated on
// s[0] > s[3], so launches after convergence change nothing and the step
// count equals the reference loop's.
template <typename T>
__global__ void k_cg_init(const T *b, T *x, T *r, T *pv, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = mkT<T>(0.0, 0.0);
  r[i] = b[i];
  pv[i] = b[i];
}

__global__ void k_cg_setup(double *s, double tol2) {
  s[3] = tol2 * s[0];
  s[4] = 0.0;
}

template <typename T>
__global__ void k_cg_step(T *x, T *r, const T *pv, const T *ap, const double *s, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !(s[0] > s[3])) return;
  const double alpha = s[0] / s[1];
  x[i] = x[i] + scl(alpha, pv[i]);
  r[i] = r[i] - scl(alpha, ap[i]);
}

template <typename T>
__global__ void k_cg_dir(const T *r, T *pv, const double *s, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !(s[0] > s[3])) return;
  pv[i] = r[i] + scl(s[2] / s[0], pv[i]);
}

__global__ void k_cg_iters(const double *s, int *info) { *info = (int)s[4]; }

__global__ void k_cg_shift(double *s) {
  if (s[0] > s[3]) {
    s[0] = s[2];
    s[4] += 1.0;
  }
}

// Gershgorin bound max_i sum_j |M_ij| >= lambda_max(M): one workgroup per row
// writes its row sum, then one workgroup takes the maximum.
template <typename T>
__global__ __launch_bounds__(256) void k_row_abs(const T *M, int n, int64_t lda, double *rows) {
  __shared__ double red[4];
  const T *row = M + (int64_t)blockIdx.x * lda;
  double acc = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) acc += sqrt(sq(row[j]));
  acc = wave_red(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) rows[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_max(const double *v, int n, double *out) {
  __shared__ double red[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, v[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

struct SrLayout {
  size_t oc, mean, dv, mat, rhs, y, x, map_src, map_k, info, cg, total;
};

SrLayout sr_layout(int ns, int64_t np, int d, int space, int cplx) {
  SrLayout L{};
  const size_t es = cplx ? 16 : 8;
  const int64_t npx = np * d;
  const int64_t nmat = space == 0 ? npx : ns;
  size_t off = 0;
  auto take = [&](size_t &r, size_t bytes) {
    r = off;
    off = al(off + bytes);
  };
  take(L.oc, (size_t)ns * npx * es);
  take(L.mean, (size_t)npx * es);
  take(L.dv, (size_t)ns * es);
  take(L.mat, (size_t)nmat * nmat * es);
  take(L.rhs, (size_t)nmat * es);
  take(L.y, (size_t)ns * es);
  take(L.x, (size_t)npx * es);
  take(L.map_src, (size_t)npx * 4);
  take(L.map_k, (size_t)npx);
  take(L.info, 8);
  // CG vectors x, r, p, A p and 8 scalars; the SVD bound's row sums fit inside
  take(L.cg, (size_t)4 * nmat * es + 8 * sizeof(double));
  L.total = off;
  return L;
}

struct SrCall {
  int ns, d, n_sites, o_dtype, e_cplx, space, real_params;
  int64_t np;
  const int32_t *pps;
  const int32_t *dsite;  // slices per site (NULL: d everywhere)
  const void *o;
  const int8_t *p;
  const double *eloc;
  double diag_shift, gf_re, gf_im;
  int solver;         // VMC_SOLVE_CHOLESKY, VMC_SOLVE_CG or VMC_SOLVE_SVD
  double solver_tol;  // CG: relative residual tol; SVD: rcond
  int maxiter;        // CG step limit
  void *x_out, *dense_out;
  double *metrics_out;
  void *ws;
  size_t ws_bytes;
  hipStream_t st;
};

int64_t expanded_cols(const SrCall &a) {
  if (!a.dsite) return a.np * a.d;
  int64_t t = 0;
  for (int s = 0; s < a.n_sites; ++s) t += (int64_t)a.dsite[s] * a.pps[s];
  return t;
}

// Solve mat z = b in place of b (mat n x n, Hermitian positive definite once
// shifted) with the call's solver; `cg` is the CG/Gershgorin scratch and info
// receives the Cholesky pivot (Cholesky, SVD) or the CG step count (CG).
template <typename T>
int solve_shifted(const SrCall &a, T *mat, int n, T *b, char *cg, int *info) {
  hipStream_t st = a.st;
  if (a.solver == VMC_SOLVE_CG) {
    // scratch: x, r, p, A p (n entries each), then the CG scalars
    T *x = (T *)cg, *r = x + n, *pv = r + n, *apx = pv + n;
    double *sc = (double *)(apx + n);
    const dim3 g((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(k_cg_init<T>, g, dim3(256), 0, st, b, x, r, pv, (int64_t)n);
    hipLaunchKernelGGL(k_dot_re<T>, dim3(1), dim3(256), 0, st, r, r, (int64_t)n, sc);
    hipLaunchKernelGGL(k_cg_setup, dim3(1), dim3(1), 0, st, sc, a.solver_tol * a.solver_tol);
    for (int it = 0; it < a.maxiter; ++it) {
      gemm<T, false, false>(n, 1, n, 1.0, mat, n, 1, pv, 0, 1, 0.0, apx, 1, st);
      hipLaunchKernelGGL(k_dot_re<T>, dim3(1), dim3(256), 0, st, pv, apx, (int64_t)n, sc + 1);
      hipLaunchKernelGGL(k_cg_step<T>, g, dim3(256), 0, st, x, r, pv, apx, sc, (int64_t)n);
      hipLaunchKernelGGL(k_dot_re<T>, dim3(1), dim3(256), 0, st, r, r, (int64_t)n, sc + 2);
      hipLaunchKernelGGL(k_cg_dir<T>, g, dim3(256), 0, st, r, pv, sc, (int64_t)n);
      hipLaunchKernelGGL(k_cg_shift, dim3(1), dim3(1), 0, st, sc);
      if ((it & 15) == 15 || it == a.maxiter - 1) {
        double hs[5];
        if (hipMemcpyAsync(hs, sc, sizeof(hs), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
          return vmc::sr_fail("CG convergence check failed");
        if (!(hs[0] > hs[3])) break;
      }
    }
    (void)hipMemcpyAsync(b, x, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, st);
    hipLaunchKernelGGL(k_cg_iters, dim3(1), dim3(1), 0, st, sc, info);
    return 0;
  }
  if (a.solver == VMC_SOLVE_SVD) {
    // solve_svd (solvers.py:31-37) keeps singular values above rcond * max(S).
    // mat = S + eps with S Hermitian PSD has every eigenvalue >= eps and
    // <= the Gershgorin bound; when eps > rcond * bound nothing is cut and the
    // pseudo-inverse solve is the Cholesky solve below.
    double *rows = (double *)cg, *bound = rows + n;
    hipLaunchKernelGGL(k_row_abs<T>, dim3(n), dim3(256), 0, st, mat, n, (int64_t)n, rows);
    hipLaunchKernelGGL(k_max, dim3(1), dim3(256), 0, st, rows, n, bound);
    double hb = 0.0;
    if (hipMemcpyAsync(&hb, bound, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return vmc::sr_fail("solve_svd bound failed");
    if (!(a.diag_shift > a.solver_tol * hb))
      return vmc::sr_fail("solve_svd: diag_shift " + std::to_string(a.diag_shift) + " is not above rcond * " +
                          std::to_string(hb) + " (a Gershgorin bound of the largest singular value), so the cutoff "
                          "may remove singular values; that case needs a dense eigensolver, which is not built");
  }
  cholesky<T>(mat, n, n, info, st);
  trsv_both<T>(mat, n, n, b, st);
  return 0;
}

template <typename T>
int sr_run(const SrCall &a) {
  constexpr int cplx = sizeof(T) == 16;
  const int ns = a.ns;
  const int64_t np = a.np, npx = expanded_cols(a);
  // the layout for phys_dim d (>= every site's slice count) bounds this one
  const SrLayout L = sr_layout(ns, np, a.d, a.space, cplx);
  if (a.ws_bytes < L.total) return vmc::sr_fail("SR workspace too small: need " + std::to_string(L.total));
  hipStream_t st = a.st;
  char *w = (char *)a.ws;
  T *oc = (T *)(w + L.oc), *mean = (T *)(w + L.mean), *dv = (T *)(w + L.dv);
  T *mat = (T *)(w + L.mat), *rhs = (T *)(w + L.rhs), *y = (T *)(w + L.y), *x = (T *)(w + L.x);
  int32_t *map_src = (int32_t *)(w + L.map_src);
  int8_t *map_k = (int8_t *)(w + L.map_k);
  int *info = (int *)(w + L.info);

  // site-major column map: site s with n_s params at sliced offset off_s and
  // d_s slices owns expanded columns xoff_s + k*n_s + j  <-  (sliced column
  // off_s + j, slice k), xoff_s = sum of d_t n_t over earlier sites
  std::vector<int32_t> src(npx);
  std::vector<int8_t> kk(npx);
  int64_t off = 0, xoff = 0;
  for (int s = 0; s < a.n_sites; ++s) {
    const int n = a.pps[s];
    const int ds = a.dsite ? a.dsite[s] : a.d;
    if (ds > a.d) return vmc::sr_fail("slices_per_site exceeds phys_dim");
    for (int k = 0; k < ds; ++k)
      for (int j = 0; j < n; ++j) {
        src[xoff + (int64_t)k * n + j] = (int32_t)(off + j);
        kk[xoff + (int64_t)k * n + j] = (int8_t)k;
      }
    off += n;
    xoff += (int64_t)ds * n;
  }
  if (off != np) return vmc::sr_fail("params_per_site does not sum to n_sliced");
  if (hipMemcpyAsync(map_src, src.data(), npx * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(map_k, kk.data(), npx, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemsetAsync(info, 0, sizeof(int), st) != hipSuccess)
    return vmc::sr_fail("hipMemcpyAsync failed");
  // the host maps must outlive the asynchronous copies
  if (hipStreamSynchronize(st) != hipSuccess) return vmc::sr_fail("hipStreamSynchronize failed");

  const dim3 cgrid((unsigned)((npx + 255) / 256), ns);
  const dim3 cols((unsigned)((npx + 255) / 256));
  hipLaunchKernelGGL(k_expand<T>, cgrid, dim3(256), 0, st, a.o, a.o_dtype, a.p, map_src, map_k, ns, np, npx, oc);
  hipLaunchKernelGGL(k_colmean<T>, cols, dim3(256), 0, st, oc, ns, npx, mean);
  hipLaunchKernelGGL(k_center<T>, cgrid, dim3(256), 0, st, oc, ns, npx, mean);
  if (a.dense_out) {
    // QGT.to_dense (qgt.py:173-223): S (space 0) or T (space 1), no shift
    T *dense = (T *)a.dense_out;
    if (a.space == 0)
      gemm_herm<T, true, false>((int)npx, ns, 1.0 / ns, oc, 1, npx, 0.0, dense, npx, true, st);
    else
      gemm_herm<T, false, true>(ns, npx, 1.0 / ns, oc, npx, 1, 0.0, dense, ns, true, st);
    return 0;
  }
  hipLaunchKernelGGL(k_dv<T>, dim3(1), dim3(256), 0, st, a.eloc, a.e_cplx, ns, a.gf_re, a.gf_im, dv);

  if (a.space == 0) {
    const int n = (int)npx;
    // rhs = O_c^H dv ; S = O_c^H O_c / Ns
    gemm<T, true, false>(n, 1, ns, 1.0, oc, 1, npx, dv, 0, 1, 0.0, rhs, 1, st);
    // Cholesky reads the lower triangle only; CG's products and the SVD bound
    // read whole rows, so the other solvers mirror the Gram
    gemm_herm<T, true, false>(n, ns, 1.0 / ns, oc, 1, npx, 0.0, mat, n, a.solver != VMC_SOLVE_CHOLESKY, st);
    hipLaunchKernelGGL(k_add_diag<T>, dim3((n + 255) / 256), dim3(256), 0, st, mat, n, n, a.diag_shift);
    (void)hipMemcpyAsync(x, rhs, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, st);
    if (const int rc = solve_shifted<T>(a, mat, n, x, w + L.cg, info)) return rc;
    if (a.metrics_out) {
      // unshifted S x = O_c^H (O_c x) / Ns through two matrix-vector products
      // (y, then mat, are free here)
      gemm<T, false, false>(ns, 1, n, 1.0, oc, npx, 1, x, 0, 1, 0.0, y, 1, st);
      gemm<T, true, false>(n, 1, ns, 1.0 / ns, oc, 1, npx, y, 0, 1, 0.0, mat, 1, st);
      hipLaunchKernelGGL(k_resid_vec<T>, dim3(1), dim3(256), 0, st, mat, rhs, (int64_t)n, a.metrics_out);
    }
  } else {
    // T = O_c O_c^H / Ns ; (T + eps) y = dv ; x = O_c^H y
    gemm_herm<T, false, true>(ns, npx, 1.0 / ns, oc, npx, 1, 0.0, mat, ns, a.solver != VMC_SOLVE_CHOLESKY, st);
    hipLaunchKernelGGL(k_add_diag<T>, dim3((ns + 255) / 256), dim3(256), 0, st, mat, ns, ns, a.diag_shift);
    (void)hipMemcpyAsync(y, dv, (size_t)ns * sizeof(T), hipMemcpyDeviceToDevice, st);
    if (const int rc = solve_shifted<T>(a, mat, ns, y, w + L.cg, info)) return rc;
    gemm<T, true, false>((int)npx, 1, ns, 1.0, oc, 1, npx, y, 0, 1, 0.0, x, 1, st);
    if (a.metrics_out) {
      // unshifted T y = O_c (O_c^H y) / Ns = O_c x / Ns (rhs is free here)
      gemm<T, false, false>(ns, 1, npx, 1.0 / ns, oc, npx, 1, x, 0, 1, 0.0, rhs, 1, st);
      hipLaunchKernelGGL(k_resid_vec<T>, dim3(1), dim3(256), 0, st, rhs, dv, (int64_t)ns, a.metrics_out);
    }
  }
  // preconditioners.py:321-329: 2 Re x for real parameters under a real grad
  // factor; a complex factor with real parameters keeps Re x (tree_cast).
  const double factor = a.gf_im == 0.0 ? 2.0 : 1.0;
  hipLaunchKernelGGL(k_finish<T>, cols, dim3(256), 0, st, x, npx, a.real_params, factor, a.x_out);
  if (a.metrics_out) hipLaunchKernelGGL(k_info, dim3(1), dim3(1), 0, st, info, a.metrics_out + 1);
  return 0;
}

bool solve_is_complex(int o_dtype, int e_cplx, double gf_im) {
  return o_dtype == VMC_C64 || o_dtype == VMC_C128 || e_cplx || gf_im != 0.0;
}

int sr_dispatch(const SrCall &a) {
  if (!a.o || !a.p || !a.ws || !a.pps) return vmc::sr_fail("null argument");
  if (!a.dense_out && (!a.eloc || !a.x_out)) return vmc::sr_fail("null argument");
  if (a.space != 0 && a.space != 1) return vmc::sr_fail("space must be 0 (parameter) or 1 (sample)");
  if (a.o_dtype < VMC_F32 || a.o_dtype > VMC_C128) return vmc::sr_fail("o_dtype must be a vmc_dtype");
  if (a.ns < 1 || a.np < 1 || a.d < 1) return vmc::sr_fail("empty Jacobian");
  return solve_is_complex(a.o_dtype, a.e_cplx, a.gf_im) ? sr_run<cf64>(a) : sr_run<double>(a);
}


// ------------------------------------------------------------------------- //
// QRSolve (preconditioners.py:166-214): pivoted Householder QR of the centred
// full Jacobian O_c (Ns x n), column-major on the device, then the rank-r
// (minimum-norm) solve of R11 x1 + R12 x2 = Q^H dv.
// ------------------------------------------------------------------------- //

__device__ inline double wave_max_d(double v) {
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// A[c*ns + i] = o[i*n + c] (transpose of the sample-major Jacobian).
template <typename T>
__global__ void k_qr_load(const void *o, int o_dtype, int ns, int64_t n, T *A) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)ns * n) return;
  const int64_t c = idx / ns, i = idx - c * ns;
  A[idx] = load_o<T>(o, o_dtype, i * n + c);
}

// One wave per column: centre, and (norms != null) column norms + identity pivots.
template <typename T>
__global__ __launch_bounds__(256) void k_qr_center(T *A, int ns, int64_t n, double *vn1, double *vn2, int *piv) {
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= n) return;
  T *col = A + c * ns;
  double sr = 0.0, si = 0.0;
  for (int i = lane; i < ns; i += 64) {
    const T v = col[i];
    sr += re(v);
    si += imv(v);
  }
  sr = wave_red(sr);
  si = wave_red(si);
  const T mean = mkT<T>(sr / ns, si / ns);
  double nn = 0.0;
  for (int i = lane; i < ns; i += 64) {
    const T v = col[i] - mean;
    col[i] = v;
    nn += sq(v);
  }
  nn = wave_red(nn);
  if (vn1 && lane == 0) {
    vn1[c] = sqrt(nn);
    vn2[c] = sqrt(nn);
    piv[c] = (int)c;
  }
}

// Pivot of step j: the first column of largest partial norm among j..n-1
// (xLAQP2 / IDAMAX), swapped into column j (single workgroup of 1024).
template <typename T>
__global__ __launch_bounds__(1024) void k_qr_pivot(T *A, int ns, int64_t n, int j, double *vn1, double *vn2,
                                                   int *piv) {
  __shared__ double bv[16];
  __shared__ int64_t bi[16];
  __shared__ int64_t pstar;
  double best = -1.0;
  int64_t bidx = n;
  for (int64_t c = j + threadIdx.x; c < n; c += blockDim.x)
    if (vn1[c] > best) {
      best = vn1[c];
      bidx = c;
    }
  for (int off = 32; off >= 1; off >>= 1) {
    const double ov = __shfl_xor(best, off, 64);
    const int64_t oi = __shfl_xor(bidx, off, 64);
    if (ov > best || (ov == best && oi < bidx)) {
      best = ov;
      bidx = oi;
    }
  }
  if ((threadIdx.x & 63) == 0) {
    bv[threadIdx.x >> 6] = best;
    bi[threadIdx.x >> 6] = bidx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = -1.0;
    int64_t p = n;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
      if (bv[w] > b || (bv[w] == b && bi[w] < p)) {
        b = bv[w];
        p = bi[w];
      }
    pstar = p < n ? p : j;
  }
  __syncthreads();
  const int64_t p = pstar;
  if (p == j) return;
  T *a = A + (int64_t)j * ns, *bcol = A + p * ns;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    const T t = a[i];
    a[i] = bcol[i];
    bcol[i] = t;
  }
  if (threadIdx.x == 0) {
    const int t = piv[j];
    piv[j] = piv[p];
    piv[p] = t;
    vn1[p] = vn1[j];
    vn2[p] = vn2[j];
  }
}

// Householder reflector of column j below row j (xLARFG), single workgroup:
// A[j][j] <- beta, A[j][j+1:] <- v (unit leading entry implied), tau[j].
template <typename T>
__global__ __launch_bounds__(256) void k_qr_house(T *A, int ns, int j, T *tau) {
  __shared__ double red[4];
  T *col = A + (int64_t)j * ns;
  double s = 0.0;
  for (int i = j + 1 + threadIdx.x; i < ns; i += blockDim.x) s += sq(col[i]);
  s = wave_red(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  const double sigma = red[0] + red[1] + red[2] + red[3];
  const T alpha = col[j];
  const double ar = re(alpha), ai = imv(alpha);
  T t, scale;
  double beta;
  if (sigma == 0.0 && ai == 0.0) {
    t = T(0.0);
    scale = T(0.0);
    beta = ar;
  } else {
    const double nrm = sqrt(ar * ar + ai * ai + sigma);
    beta = ar >= 0.0 ? -nrm : nrm;
    t = mkT<T>((beta - ar) / beta, -ai / beta);
    const double er = ar - beta, ei = ai, den = 1.0 / (er * er + ei * ei);
    scale = mkT<T>(er * den, -ei * den);
  }
  __syncthreads();
  for (int i = j + 1 + threadIdx.x; i < ns; i += blockDim.x) col[i] = col[i] * scale;
  if (threadIdx.x == 0) {
    col[j] = T(beta);
    tau[j] = t;
  }
}

// Apply H_j^H to the trailing columns (one wave per column) and downdate the
// partial norms as xLAQP2 does (recomputing them when cancellation is severe).
template <typename T>
__global__ __launch_bounds__(256) void k_qr_apply(T *A, int ns, int64_t n, int j, const T *tau, double *vn1,
                                                  double *vn2) {
  const int64_t c = (int64_t)j + 1 + (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= n) return;
  const T *v = A + (int64_t)j * ns;
  T *col = A + c * ns;
  T w = T(0.0);
  for (int i = j + lane; i < ns; i += 64) w = w + (i == j ? col[i] : cj(v[i]) * col[i]);
  const double wr = wave_red(re(w)), wi = wave_red(imv(w));
  w = mkT<T>(wr, wi);
  const T f = cj(tau[j]) * w;
  for (int i = j + lane; i < ns; i += 64) col[i] = col[i] - (i == j ? f : v[i] * f);
  if (vn1[c] != 0.0) {
    const double rj = sqrt(sq(col[j]));
    double temp = 1.0 - (rj / vn1[c]) * (rj / vn1[c]);
    temp = fmax(temp, 0.0);
    const double ratio = vn1[c] / vn2[c];
    const double temp2 = temp * ratio * ratio;
    if (temp2 <= 1.4901161193847656e-08) {  // sqrt(eps): recompute
      double nn = 0.0;
      for (int i = j + 1 + lane; i < ns; i += 64) nn += sq(col[i]);
      nn = wave_red(nn);
      if (lane == 0) {
        vn1[c] = sqrt(nn);
        vn2[c] = sqrt(nn);
      }
    } else if (lane == 0) {
      vn1[c] *= sqrt(temp);
    }
  }
}

// y <- Q^H y = H_{k-1}^H ... H_0^H y (single workgroup).
template <typename T>
__global__ __launch_bounds__(256) void k_qr_qty(const T *A, int ns, int k, const T *tau, T *y) {
  __shared__ double red[2][4];
  for (int j = 0; j < k; ++j) {
    const T *v = A + (int64_t)j * ns;
    T w = T(0.0);
    for (int i = j + threadIdx.x; i < ns; i += blockDim.x) w = w + (i == j ? y[i] : cj(v[i]) * y[i]);
    const double wr = wave_red(re(w)), wi = wave_red(imv(w));
    if ((threadIdx.x & 63) == 0) {
      red[0][threadIdx.x >> 6] = wr;
      red[1][threadIdx.x >> 6] = wi;
    }
    __syncthreads();
    w = mkT<T>(red[0][0] + red[0][1] + red[0][2] + red[0][3], red[1][0] + red[1][1] + red[1][2] + red[1][3]);
    const T f = cj(tau[j]) * w;
    for (int i = j + threadIdx.x; i < ns; i += blockDim.x) y[i] = y[i] - (i == j ? f : v[i] * f);
    __syncthreads();
  }
}

// Numerical rank from |R_jj| (pivoted QR keeps them non-increasing):
// count of |R_jj| > tol_rel * |R_00| (single thread).
template <typename T>
__global__ void k_qr_rank(const T *A, int ns, int k, double tol_rel, int *rank) {
  const double r0 = sqrt(sq(A[0]));
  int r = 0;
  for (int j = 0; j < k; ++j)
    if (sqrt(sq(A[(int64_t)j * ns + j])) > tol_rel * r0) r = j + 1;
  *rank = r0 > 0.0 ? r : 0;
}

// Upper-triangular solves with R11 = A[0:r, 0:r] (column-major, lda ns):
// every thread solves one right-hand side (column c of B, leading dimension
// ldb) in place by column-oriented back substitution.
template <typename T>
__global__ void k_qr_trsm(const T *A, int ns, int r, T *B, int64_t ldb, int64_t nrhs) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  T *b = B + c * ldb;
  for (int l = r - 1; l >= 0; --l) {
    const T *rl = A + (int64_t)l * ns;
    // b_l / R_ll (complex division through the conjugate)
    const T rll = rl[l];
    const T xl = scl(1.0 / sq(rll), b[l] * cj(rll));
    b[l] = xl;
    for (int i = 0; i < l; ++i) b[i] = b[i] - rl[i] * xl;
  }
}

template <typename T>
__global__ void k_qr_scatter(const T *z, const int *piv, int64_t n, int r, T *x) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  x[piv[j]] = z[j];
}

template <typename T>
__global__ void k_qr_fillz(T *z, int64_t from, int64_t n) {
  const int64_t j = from + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) z[j] = T(0.0);
}

// out = |g - h|^2 / |h|^2 (single workgroup).

struct QrLayout {
  size_t A, vn1, vn2, piv, tau, dv, y, z, lhs, rhs2, t, g, h, info, rank, total;
};

QrLayout qr_layout(int ns, int64_t n, int cplx) {
  QrLayout L{};
  const size_t es = cplx ? 16 : 8;
  const int64_t k = std::min<int64_t>(ns, n);
  const int64_t n2max = n;  // n - rank <= n
  size_t off = 0;
  auto take = [&](size_t &r, size_t bytes) {
    r = off;
    off = al(off + bytes);
  };
  take(L.A, (size_t)ns * n * es);
  take(L.vn1, (size_t)n * 8);
  take(L.vn2, (size_t)n * 8);
  take(L.piv, (size_t)n * 4);
  take(L.tau, (size_t)k * es);
  take(L.dv, (size_t)ns * es);
  take(L.y, (size_t)ns * es);
  take(L.z, (size_t)n * es);
  take(L.lhs, (size_t)n2max * n2max * es);
  take(L.rhs2, (size_t)n2max * es);
  take(L.t, (size_t)ns * es);
  take(L.g, (size_t)n * es);
  take(L.h, (size_t)n * es);
  take(L.info, 8);
  take(L.rank, 8);
  L.total = off;
  return L;
}

template <typename T>
int qr_run(int ns, int64_t n, const void *o, int o_dtype, const double *eloc, int e_cplx, double gf_re,
           double gf_im, double rcond, int min_norm, int real_params, void *x_out, double *metrics, void *ws,
           size_t ws_bytes, hipStream_t st) {
  constexpr int cplx = sizeof(T) == 16;
  const QrLayout L = qr_layout(ns, n, cplx);
  if (ws_bytes < L.total) return vmc::sr_fail("QRSolve workspace too small: need " + std::to_string(L.total));
  char *w = (char *)ws;
  T *A = (T *)(w + L.A), *tau = (T *)(w + L.tau), *dv = (T *)(w + L.dv), *y = (T *)(w + L.y), *z = (T *)(w + L.z);
  T *lhs = (T *)(w + L.lhs), *rhs2 = (T *)(w + L.rhs2), *t = (T *)(w + L.t), *g = (T *)(w + L.g), *h = (T *)(w + L.h);
  double *vn1 = (double *)(w + L.vn1), *vn2 = (double *)(w + L.vn2);
  int *piv = (int *)(w + L.piv), *info = (int *)(w + L.info), *rank_d = (int *)(w + L.rank);
  const int64_t k = std::min<int64_t>(ns, n);
  const int64_t tot = (int64_t)ns * n;
  const unsigned cblocks = (unsigned)((n + 3) / 4);
  hipLaunchKernelGGL(k_qr_load<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, o, o_dtype, ns, n, A);
  hipLaunchKernelGGL(k_qr_center<T>, dim3(cblocks), dim3(256), 0, st, A, ns, n, vn1, vn2, piv);
  hipLaunchKernelGGL(k_dv<T>, dim3(1), dim3(256), 0, st, eloc, e_cplx, ns, gf_re, gf_im, dv);
  for (int64_t j = 0; j < k; ++j) {
    hipLaunchKernelGGL(k_qr_pivot<T>, dim3(1), dim3(1024), 0, st, A, ns, n, (int)j, vn1, vn2, piv);
    hipLaunchKernelGGL(k_qr_house<T>, dim3(1), dim3(256), 0, st, A, ns, (int)j, tau);
    if (j + 1 < n)
      hipLaunchKernelGGL(k_qr_apply<T>, dim3((unsigned)((n - j - 1 + 3) / 4)), dim3(256), 0, st, A, ns, n, (int)j,
                         tau, vn1, vn2);
  }
  (void)hipMemcpyAsync(y, dv, (size_t)ns * sizeof(T), hipMemcpyDeviceToDevice, st);
  hipLaunchKernelGGL(k_qr_qty<T>, dim3(1), dim3(256), 0, st, A, ns, (int)k, tau, y);
  // jnp.linalg.matrix_rank(r) (preconditioners.py:183-186): default relative
  // tolerance max(r.shape) eps with r of shape (min(ns, n), n), i.e. n eps,
  // unless rcond is given; applied to the pivoted |R_jj| against |R_00|
  const double tol_rel = rcond > 0.0 ? rcond : (double)n * 2.220446049250313e-16;
  hipLaunchKernelGGL(k_qr_rank<T>, dim3(1), dim3(1), 0, st, A, ns, (int)k, tol_rel, rank_d);
  int rank = 0;
  if (hipMemcpyAsync(&rank, rank_d, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return vmc::sr_fail("rank readback failed");
  const int64_t n2 = n - rank;
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_qr_fillz<T>, dim3(nb), dim3(256), 0, st, z, 0, n);
  (void)hipMemsetAsync(info, 0, sizeof(int), st);
  if (rank > 0) {
    // a = R11^-1 y1 into z[0:r]
    (void)hipMemcpyAsync(z, y, (size_t)rank * sizeof(T), hipMemcpyDeviceToDevice, st);
    hipLaunchKernelGGL(k_qr_trsm<T>, dim3(1), dim3(1), 0, st, A, ns, rank, z, (int64_t)rank, (int64_t)1);
    if (min_norm && n2 > 0) {
      // B = R11^-1 R12 in place of R12 (columns rank .. n-1 of A)
      T *B = A + (int64_t)rank * ns;
      hipLaunchKernelGGL(k_qr_trsm<T>, dim3((unsigned)((n2 + 63) / 64)), dim3(64), 0, st, A, ns, rank, B,
                         (int64_t)ns, n2);
      // (B^H B + I) x2 = B^H a
      gemm_herm<T, true, false>((int)n2, rank, 1.0, B, ns, 1, 0.0, lhs, n2, false, st);
      hipLaunchKernelGGL(k_add_diag<T>, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, lhs, (int)n2, n2, 1.0);
      gemm<T, true, false>((int)n2, 1, rank, 1.0, B, ns, 1, z, 0, 1, 0.0, rhs2, 1, st);
      cholesky<T>(lhs, (int)n2, n2, info, st);
      trsv_both<T>(lhs, (int)n2, n2, rhs2, st);
      // x1 = a - B x2; z = [x1; x2]
      gemm<T, false, false>(rank, 1, n2, -1.0, B, 1, ns, rhs2, 0, 1, 1.0, z, 1, st);
      (void)hipMemcpyAsync(z + rank, rhs2, (size_t)n2 * sizeof(T), hipMemcpyDeviceToDevice, st);
    }
  }
  T *x = h;  // scatter target (reuses h before the residual)
  hipLaunchKernelGGL(k_qr_scatter<T>, dim3(nb), dim3(256), 0, st, z, piv, n, rank, x);
  (void)hipMemcpyAsync(z, x, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, st);
  // residual |O_c^H O_c x - O_c^H dv|^2 / |O_c^H dv|^2 with O_c rebuilt in A
  hipLaunchKernelGGL(k_qr_load<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, o, o_dtype, ns, n, A);
  hipLaunchKernelGGL(k_qr_center<T>, dim3(cblocks), dim3(256), 0, st, A, ns, n, nullptr, nullptr, nullptr);
  gemm<T, false, false>(ns, 1, n, 1.0, A, 1, ns, z, 0, 1, 0.0, t, 1, st);
  gemm<T, true, false>((int)n, 1, ns, 1.0, A, ns, 1, t, 0, 1, 0.0, g, 1, st);
  gemm<T, true, false>((int)n, 1, ns, 1.0, A, ns, 1, dv, 0, 1, 0.0, h, 1, st);
  if (metrics) {
    hipLaunchKernelGGL(k_resid_vec<T>, dim3(1), dim3(256), 0, st, g, h, n, metrics);
    hipLaunchKernelGGL(k_info, dim3(1), dim3(1), 0, st, rank_d, metrics + 1);
  }
  const double factor = gf_im == 0.0 ? 2.0 : 1.0;
  hipLaunchKernelGGL(k_finish<T>, dim3(nb), dim3(256), 0, st, z, n, real_params, factor, x_out);
  return 0;
}

}  // namespace

extern "C" {

size_t vmc_sr_workspace_size(int n_samples, int64_t n_sliced, int phys_dim, int space, int complex_solve) {
  if (n_samples < 1 || n_sliced < 1 || phys_dim < 1) return 0;
  return sr_layout(n_samples, n_sliced, phys_dim, space, complex_solve != 0).total;
}

int vmc_sr_solve(int n_samples, int64_t n_sliced, int phys_dim, const int32_t *params_per_site, int n_sites,
                 const int32_t *slices_per_site, const void *o, int o_dtype, const int8_t *p, const double *eloc,
                 int eloc_complex, int space, double diag_shift, double grad_factor_re, double grad_factor_im,
                 int real_params, void *x_out, double *metrics_out, void *workspace, size_t workspace_bytes,
                 void *stream) {
  SrCall a{};
  a.dsite = slices_per_site;
  a.ns = n_samples;
  a.np = n_sliced;
  a.d = phys_dim;
  a.pps = params_per_site;
  a.n_sites = n_sites;
  a.o = o;
  a.o_dtype = o_dtype;
  a.p = p;
  a.eloc = eloc;
  a.e_cplx = eloc_complex != 0;
  a.space = space;
  a.diag_shift = diag_shift;
  a.gf_re = grad_factor_re;
  a.gf_im = grad_factor_im;
  a.real_params = real_params != 0;
  a.x_out = x_out;
  a.metrics_out = metrics_out;
  a.ws = workspace;
  a.ws_bytes = workspace_bytes;
  a.st = (hipStream_t)stream;
  a.solver = VMC_SOLVE_CHOLESKY;
  return sr_dispatch(a);
}

int vmc_sr_solve_ex(int n_samples, int64_t n_sliced, int phys_dim, const int32_t *params_per_site, int n_sites,
                    const int32_t *slices_per_site, const void *o, int o_dtype, const int8_t *p, const double *eloc,
                    int eloc_complex, int space, double diag_shift, double grad_factor_re, double grad_factor_im,
                    int real_params, int solver, double solver_tol, int maxiter, void *x_out, double *metrics_out,
                    void *workspace, size_t workspace_bytes, void *stream) {
  if (solver != VMC_SOLVE_CHOLESKY && solver != VMC_SOLVE_CG && solver != VMC_SOLVE_SVD)
    return vmc::sr_fail("solver must be VMC_SOLVE_CHOLESKY, VMC_SOLVE_CG or VMC_SOLVE_SVD");
  if (solver == VMC_SOLVE_CG && (maxiter < 1 || !(solver_tol >= 0.0)))
    return vmc::sr_fail("CG needs maxiter >= 1 and tol >= 0");
  if (solver == VMC_SOLVE_SVD && !(solver_tol >= 0.0)) return vmc::sr_fail("rcond must be non-negative");
  SrCall a{};
  a.dsite = slices_per_site;
  a.ns = n_samples;
  a.np = n_sliced;
  a.d = phys_dim;
  a.pps = params_per_site;
  a.n_sites = n_sites;
  a.o = o;
  a.o_dtype = o_dtype;
  a.p = p;
  a.eloc = eloc;
  a.e_cplx = eloc_complex != 0;
  a.space = space;
  a.diag_shift = diag_shift;
  a.gf_re = grad_factor_re;
  a.gf_im = grad_factor_im;
  a.real_params = real_params != 0;
  a.solver = solver;
  a.solver_tol = solver_tol;
  a.maxiter = maxiter;
  a.x_out = x_out;
  a.metrics_out = metrics_out;
  a.ws = workspace;
  a.ws_bytes = workspace_bytes;
  a.st = (hipStream_t)stream;
  return sr_dispatch(a);
}

int vmc_qgt_dense(int n_samples, int64_t n_sliced, int phys_dim, const int32_t *params_per_site, int n_sites,
                  const int32_t *slices_per_site, const void *o, int o_dtype, const int8_t *p, int space,
                  void *dense_out, void *workspace, size_t workspace_bytes, void *stream) {
  if (!dense_out) return vmc::sr_fail("null argument");
  SrCall a{};
  a.dsite = slices_per_site;
  a.ns = n_samples;
  a.np = n_sliced;
  a.d = phys_dim;
  a.pps = params_per_site;
  a.n_sites = n_sites;
  a.o = o;
  a.o_dtype = o_dtype;
  a.p = p;
  a.space = space;
  a.gf_re = 1.0;
  a.dense_out = dense_out;
  a.ws = workspace;
  a.ws_bytes = workspace_bytes;
  a.st = (hipStream_t)stream;
  return sr_dispatch(a);
}

size_t vmc_sr_qr_workspace_size(int n_samples, int64_t n_params, int complex_solve) {
  if (n_samples < 1 || n_params < 1) return 0;
  return qr_layout(n_samples, n_params, complex_solve != 0).total;
}

int vmc_sr_qr_solve(int n_samples, int64_t n_params, const void *o, int o_dtype, const double *eloc,
                    int eloc_complex, double grad_factor_re, double grad_factor_im, double rcond, int min_norm,
                    int real_params, void *x_out, double *metrics_out, void *workspace, size_t workspace_bytes,
                    void *stream) {
  if (!o || !eloc || !x_out || !workspace) return vmc::sr_fail("null argument");
  if (n_samples < 1 || n_params < 1) return vmc::sr_fail("empty Jacobian");
  if (o_dtype < VMC_F32 || o_dtype > VMC_C128) return vmc::sr_fail("o_dtype must be a vmc_dtype");
  hipStream_t st = (hipStream_t)stream;
  if (solve_is_complex(o_dtype, eloc_complex, grad_factor_im))
    return qr_run<cf64>(n_samples, n_params, o, o_dtype, eloc, eloc_complex, grad_factor_re, grad_factor_im, rcond,
                        min_norm, real_params, x_out, metrics_out, workspace, workspace_bytes, st);
  return qr_run<double>(n_samples, n_params, o, o_dtype, eloc, eloc_complex, grad_factor_re, grad_factor_im, rcond,
                        min_norm, real_params, x_out, metrics_out, workspace, workspace_bytes, st);
}

}  // extern "C"
