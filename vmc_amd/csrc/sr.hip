// Stochastic-reconfiguration solve on gfx950, FP64 throughout.
//
// Mirrors SRPreconditioner.apply with DirectSolve (Cholesky) in parameter or
// sample space (src/vmc/preconditioners/preconditioners.py:117-163, 268-330)
// over the sliced Jacobian of src/vmc/qgt/jacobian.py:56-93 / qgt.py:173-223:
//
//   O_exp[i, (s, k, j)] = o[i, (s, j)] if p[i, (s, j)] == k else 0   (site-major)
//   O_c = O_exp - mean_i O_exp                                        (centring)
//   ParameterSpace:  (O_c^T O_c / Ns + eps I) x = O_c^T dv
//   SampleSpace:     (O_c O_c^T / Ns + eps I) y = dv,   x = O_c^T y
//   dv = grad_factor * (E - mean E) / Ns;  real parameters: x <- 2 Re x
//
// The reference forms the uncentred Gram and subtracts mean terms
// (qgt.py:206-223); centring O first is the same matrix in exact arithmetic.
// Columns are built directly in the site-major order the parameter tree uses,
// which is the reference's SliceOrdering result after _reorder_updates
// (preconditioners.py:61-86) and its SiteOrdering result as is.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vmc_amd.h"

namespace vmc {
int sr_fail(const std::string &msg);
}

namespace {

constexpr int TB = 64;  // GEMM tile
constexpr int NB = 32;  // Cholesky panel

// Column c of the site-major expanded Jacobian -> (source column q, slice k).
__global__ void k_expand(const void *o, int o_f32, const int8_t *p, const int32_t *col_src,
                         const int8_t *col_k, int ns, int64_t np, int64_t npx, double *oc) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (col >= npx) return;
  const int64_t q = col_src[col];
  const int8_t k = col_k[col];
  const int64_t src = (int64_t)i * np + q;
  const double v = o_f32 ? (double)((const float *)o)[src] : ((const double *)o)[src];
  oc[(int64_t)i * npx + col] = (p[src] == k) ? v : 0.0;
}

// Column means, then centring in place.
__global__ void k_colmean(const double *oc, int ns, int64_t npx, double *mean) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= npx) return;
  double s = 0.0;
  for (int i = 0; i < ns; ++i) s += oc[(int64_t)i * npx + col];
  mean[col] = s / ns;
}

__global__ void k_center(double *oc, int ns, int64_t npx, const double *mean) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (col >= npx) return;
  oc[(int64_t)i * npx + col] -= mean[col];
}

// dv = gf * (E - mean E) / Ns (single workgroup).
__global__ void k_dv(const double *e, int ns, double gf, double *dv) {
  __shared__ double red[16];
  double s = 0.0;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) s += e[i];
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double tot = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += red[w];
  const double mean = tot / ns;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) dv[i] = gf * (e[i] - mean) / ns;
}

// C[i, j] = alpha * sum_k A[i*sai + k*sak] * B[j*sbj + k*sbk] + beta * C[i, j]
// (C row-major with leading dimension ldc). 64x64 tiles, 256 threads, 4x4 per
// thread, K staged through LDS in chunks of 16.
__global__ __launch_bounds__(256) void k_dgemm(int m, int n, int64_t kdim, double alpha, const double *A,
                                               int64_t sai, int64_t sak, const double *B, int64_t sbj,
                                               int64_t sbk, double beta, double *C, int64_t ldc) {
  __shared__ double sA[16][TB + 1];
  __shared__ double sB[16][TB + 1];
  const int t = threadIdx.x;
  const int ti = t >> 4, tj = t & 15;
  const int i0 = blockIdx.y * TB, j0 = blockIdx.x * TB;
  double acc[4][4] = {};
  for (int64_t k0 = 0; k0 < kdim; k0 += 16) {
    for (int idx = t; idx < 16 * TB; idx += 256) {
      // consecutive lanes walk the operand's contiguous axis
      int kk = idx / TB, r = idx - kk * TB;
      if (sak == 1) {
        r = idx >> 4;
        kk = idx & 15;
      }
      const int64_t k = k0 + kk;
      sA[kk][r] = (i0 + r < m && k < kdim) ? A[(int64_t)(i0 + r) * sai + k * sak] : 0.0;
      int kb = idx / TB, rb = idx - kb * TB;
      if (sbk == 1) {
        rb = idx >> 4;
        kb = idx & 15;
      }
      const int64_t k2 = k0 + kb;
      sB[kb][rb] = (j0 + rb < n && k2 < kdim) ? B[(int64_t)(j0 + rb) * sbj + k2 * sbk] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
      double a[4], b[4];
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
        double *c = C + (int64_t)i * ldc + j;
        *c = alpha * acc[u][v] + (beta != 0.0 ? beta * *c : 0.0);
      }
    }
}

__global__ void k_add_diag(double *A, int n, int64_t lda, double eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[(int64_t)i * lda + i] += eps;
}

// Unblocked Cholesky of the NB x NB diagonal block at (j0, j0), in LDS.
__global__ void k_potf2(double *A, int n, int64_t lda, int j0, int *info) {
  __shared__ double L[NB][NB + 1];
  const int nb = min(NB, n - j0);
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int r = idx / NB, c = idx - r * NB;
    L[r][c] = (r < nb && c < nb) ? A[(int64_t)(j0 + r) * lda + j0 + c] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    if (threadIdx.x == 0) {
      const double djj = L[j][j];
      if (!(djj > 0.0)) atomicExch(info, j0 + j + 1);
      L[j][j] = sqrt(fmax(djj, 1e-300));
    }
    __syncthreads();
    const double ljj = L[j][j];
    for (int r = j + 1 + threadIdx.x; r < nb; r += blockDim.x) L[r][j] /= ljj;
    __syncthreads();
    for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
      const int r = idx / nb, c = idx - r * nb;
      if (c > j && r >= c) L[r][c] -= L[r][j] * L[c][j];
    }
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int r = idx / NB, c = idx - r * NB;
    if (r < nb && c <= r) A[(int64_t)(j0 + r) * lda + j0 + c] = L[r][c];
  }
}

// Panel below the diagonal block: A21 <- A21 L11^-T (one row per thread).
__global__ void k_trsm_panel(double *A, int n, int64_t lda, int j0) {
  __shared__ double L[NB][NB + 1];
  const int nb = min(NB, n - j0);
  for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    const int r = idx / NB, c = idx - r * NB;
    L[r][c] = (r < nb && c <= r) ? A[(int64_t)(j0 + r) * lda + j0 + c] : 0.0;
  }
  __syncthreads();
  const int row = j0 + nb + blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  double x[NB];
  double *a = A + (int64_t)row * lda + j0;
#pragma unroll
  for (int c = 0; c < NB; ++c) x[c] = c < nb ? a[c] : 0.0;
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    if (c < nb) {
      double s = x[c];
#pragma unroll
      for (int q = 0; q < c; ++q) s -= x[q] * L[c][q];
      x[c] = s / L[c][c];
    }
  }
#pragma unroll
  for (int c = 0; c < NB; ++c)
    if (c < nb) a[c] = x[c];
}

// Blocked forward (L z = b) or backward (L^T y = z) substitution in one
// workgroup; b is overwritten by the solution.
__global__ __launch_bounds__(256) void k_trsv(const double *A, int n, int64_t lda, double *b, int transpose) {
  __shared__ double xs[NB];
  if (!transpose) {
    for (int j0 = 0; j0 < n; j0 += NB) {
      const int nb = min(NB, n - j0);
      if (threadIdx.x == 0) {
        for (int c = 0; c < nb; ++c) {
          double s = b[j0 + c];
          for (int q = 0; q < c; ++q) s -= A[(int64_t)(j0 + c) * lda + j0 + q] * xs[q];
          xs[c] = s / A[(int64_t)(j0 + c) * lda + j0 + c];
          b[j0 + c] = xs[c];
        }
      }
      __syncthreads();
      for (int r = j0 + nb + threadIdx.x; r < n; r += blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < nb; ++q) s += A[(int64_t)r * lda + j0 + q] * xs[q];
        b[r] -= s;
      }
      __syncthreads();
    }
  } else {
    const int nblk = (n + NB - 1) / NB;
    for (int bi = nblk - 1; bi >= 0; --bi) {
      const int j0 = bi * NB;
      const int nb = min(NB, n - j0);
      if (threadIdx.x == 0) {
        for (int c = nb - 1; c >= 0; --c) {
          double s = b[j0 + c];
          for (int q = c + 1; q < nb; ++q) s -= A[(int64_t)(j0 + q) * lda + j0 + c] * xs[q];
          xs[c] = s / A[(int64_t)(j0 + c) * lda + j0 + c];
          b[j0 + c] = xs[c];
        }
      }
      __syncthreads();
      for (int r = threadIdx.x; r < j0; r += blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < nb; ++q) s += A[(int64_t)(j0 + q) * lda + r] * xs[q];
        b[r] -= s;
      }
      __syncthreads();
    }
  }
}

__global__ void k_info(const int *info, double *out) { *out = (double)*info; }

__global__ void k_scale(double *x, int64_t n, double s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= s;
}

// out = |A x - b|^2 / |b|^2 with A (n x n) row-major (single workgroup).
__global__ __launch_bounds__(256) void k_residual(const double *A, int n, int64_t lda, const double *x,
                                                  const double *b, double *out) {
  __shared__ double red[2][8];
  double num = 0.0, den = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < n; ++c) s += A[(int64_t)r * lda + c] * x[c];
    const double d = s - b[r];
    num += d * d;
    den += b[r] * b[r];
  }
  for (int off = 32; off >= 1; off >>= 1) {
    num += __shfl_xor(num, off, 64);
    den += __shfl_xor(den, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = num;
    red[1][threadIdx.x >> 6] = den;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, c = 0.0;
    for (int w = 0; w < 4; ++w) {
      a += red[0][w];
      c += red[1][w];
    }
    *out = c > 0.0 ? a / c : 0.0;
  }
}

static void gemm(int m, int n, int64_t k, double alpha, const double *A, int64_t sai, int64_t sak, const double *B,
                 int64_t sbj, int64_t sbk, double beta, double *C, int64_t ldc, hipStream_t st) {
  dim3 grid((n + TB - 1) / TB, (m + TB - 1) / TB);
  hipLaunchKernelGGL(k_dgemm, grid, dim3(256), 0, st, m, n, k, alpha, A, sai, sak, B, sbj, sbk, beta, C, ldc);
}

// In-place lower Cholesky of A (n x n); A[i,j] for j <= i is L afterwards.
static void cholesky(double *A, int n, int64_t lda, int *info, hipStream_t st) {
  for (int j0 = 0; j0 < n; j0 += NB) {
    const int nb = std::min(NB, n - j0);
    hipLaunchKernelGGL(k_potf2, dim3(1), dim3(256), 0, st, A, n, lda, j0, info);
    const int rest = n - j0 - nb;
    if (rest <= 0) break;
    hipLaunchKernelGGL(k_trsm_panel, dim3((rest + 255) / 256), dim3(256), 0, st, A, n, lda, j0);
    // trailing update A22 -= A21 A21^T
    double *A21 = A + (int64_t)(j0 + nb) * lda + j0;
    double *A22 = A + (int64_t)(j0 + nb) * lda + j0 + nb;
    gemm(rest, rest, nb, -1.0, A21, lda, 1, A21, lda, 1, 1.0, A22, lda, st);
  }
}

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

struct SrLayout {
  size_t oc, mean, dv, mat, rhs, y, map_src, map_k, info, total;
};

static SrLayout sr_layout(int ns, int64_t np, int d, int space) {
  SrLayout L{};
  const int64_t npx = np * d;
  const int64_t nmat = space == 0 ? npx : ns;
  size_t off = 0;
  auto take = [&](size_t &r, size_t bytes) {
    r = off;
    off = al(off + bytes);
  };
  take(L.oc, (size_t)ns * npx * 8);
  take(L.mean, (size_t)npx * 8);
  take(L.dv, (size_t)ns * 8);
  take(L.mat, (size_t)nmat * nmat * 8);
  take(L.rhs, (size_t)nmat * 8);
  take(L.y, (size_t)ns * 8);
  take(L.map_src, (size_t)npx * 4);
  take(L.map_k, (size_t)npx);
  take(L.info, 8);
  L.total = off;
  return L;
}

}  // namespace

extern "C" {

size_t vmc_sr_workspace_size(int n_samples, int64_t n_sliced, int phys_dim, int space) {
  if (n_samples < 1 || n_sliced < 1 || phys_dim < 1) return 0;
  return sr_layout(n_samples, n_sliced, phys_dim, space).total;
}

static int sr_common(int n_samples, int64_t n_sliced, int phys_dim, const int32_t *params_per_site, int n_sites,
                     const void *o, int o_is_f32, const int8_t *p, const double *eloc, int space, double diag_shift,
                     double grad_factor, int real_params, double *x_out, double *metrics_out, double *dense_out,
                     void *workspace, size_t workspace_bytes, void *stream) {
  if (!o || !p || !workspace || !params_per_site) return vmc::sr_fail("null argument");
  if (!dense_out && (!eloc || !x_out)) return vmc::sr_fail("null argument");
  if (space != 0 && space != 1) return vmc::sr_fail("space must be 0 (parameter) or 1 (sample)");
  const int ns = n_samples, d = phys_dim;
  const int64_t np = n_sliced, npx = np * d;
  const SrLayout L = sr_layout(ns, np, d, space);
  if (workspace_bytes < L.total) return vmc::sr_fail("SR workspace too small: need " + std::to_string(L.total));
  hipStream_t st = (hipStream_t)stream;
  char *w = (char *)workspace;
  double *oc = (double *)(w + L.oc), *mean = (double *)(w + L.mean), *dv = (double *)(w + L.dv);
  double *mat = (double *)(w + L.mat), *rhs = (double *)(w + L.rhs), *y = (double *)(w + L.y);
  int32_t *map_src = (int32_t *)(w + L.map_src);
  int8_t *map_k = (int8_t *)(w + L.map_k);
  int *info = (int *)(w + L.info);

  // site-major column map: site s with n_s params at sliced offset off_s owns
  // expanded columns d*off_s + k*n_s + j  <-  (sliced column off_s + j, slice k)
  std::vector<int32_t> src(npx);
  std::vector<int8_t> kk(npx);
  int64_t off = 0;
  for (int s = 0; s < n_sites; ++s) {
    const int n = params_per_site[s];
    for (int k = 0; k < d; ++k)
      for (int j = 0; j < n; ++j) {
        src[d * off + (int64_t)k * n + j] = (int32_t)(off + j);
        kk[d * off + (int64_t)k * n + j] = (int8_t)k;
      }
    off += n;
  }
  if (off != np) return vmc::sr_fail("params_per_site does not sum to n_sliced");
  if (hipMemcpyAsync(map_src, src.data(), npx * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(map_k, kk.data(), npx, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemsetAsync(info, 0, sizeof(int), st) != hipSuccess)
    return vmc::sr_fail("hipMemcpyAsync failed");
  // the host maps must outlive the asynchronous copies
  if (hipStreamSynchronize(st) != hipSuccess) return vmc::sr_fail("hipStreamSynchronize failed");

  const dim3 cgrid((unsigned)((npx + 255) / 256), ns);
  hipLaunchKernelGGL(k_expand, cgrid, dim3(256), 0, st, o, o_is_f32, p, map_src, map_k, ns, np, npx, oc);
  hipLaunchKernelGGL(k_colmean, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, st, oc, ns, npx, mean);
  hipLaunchKernelGGL(k_center, cgrid, dim3(256), 0, st, oc, ns, npx, mean);
  if (dense_out) {
    // QGT.to_dense (qgt.py:173-223): S (space 0) or T (space 1), no shift
    if (space == 0)
      gemm((int)npx, (int)npx, ns, 1.0 / ns, oc, 1, npx, oc, 1, npx, 0.0, dense_out, npx, st);
    else
      gemm(ns, ns, npx, 1.0 / ns, oc, npx, 1, oc, npx, 1, 0.0, dense_out, ns, st);
    return 0;
  }
  hipLaunchKernelGGL(k_dv, dim3(1), dim3(256), 0, st, eloc, ns, grad_factor, dv);

  if (space == 0) {
    const int n = (int)npx;
    // rhs = O_c^T dv ; S = O_c^T O_c / Ns
    gemm(n, 1, ns, 1.0, oc, 1, npx, dv, 0, 1, 0.0, rhs, 1, st);
    gemm(n, n, ns, 1.0 / ns, oc, 1, npx, oc, 1, npx, 0.0, mat, n, st);
    hipLaunchKernelGGL(k_add_diag, dim3((n + 255) / 256), dim3(256), 0, st, mat, n, n, diag_shift);
    double *x = x_out;
    (void)hipMemcpyAsync(x, rhs, (size_t)n * 8, hipMemcpyDeviceToDevice, st);
    cholesky(mat, n, n, info, st);
    hipLaunchKernelGGL(k_trsv, dim3(1), dim3(256), 0, st, mat, n, n, x, 0);
    hipLaunchKernelGGL(k_trsv, dim3(1), dim3(256), 0, st, mat, n, n, x, 1);
    if (metrics_out) {
      gemm(n, n, ns, 1.0 / ns, oc, 1, npx, oc, 1, npx, 0.0, mat, n, st);  // unshifted S
      hipLaunchKernelGGL(k_residual, dim3(1), dim3(256), 0, st, mat, n, n, x, rhs, metrics_out);
    }
  } else {
    // T = O_c O_c^T / Ns ; (T + eps) y = dv ; x = O_c^T y
    gemm(ns, ns, npx, 1.0 / ns, oc, npx, 1, oc, npx, 1, 0.0, mat, ns, st);
    hipLaunchKernelGGL(k_add_diag, dim3((ns + 255) / 256), dim3(256), 0, st, mat, ns, ns, diag_shift);
    (void)hipMemcpyAsync(y, dv, (size_t)ns * 8, hipMemcpyDeviceToDevice, st);
    cholesky(mat, ns, ns, info, st);
    hipLaunchKernelGGL(k_trsv, dim3(1), dim3(256), 0, st, mat, ns, ns, y, 0);
    hipLaunchKernelGGL(k_trsv, dim3(1), dim3(256), 0, st, mat, ns, ns, y, 1);
    gemm((int)npx, 1, ns, 1.0, oc, 1, npx, y, 0, 1, 0.0, x_out, 1, st);
    if (metrics_out) {
      gemm(ns, ns, npx, 1.0 / ns, oc, npx, 1, oc, npx, 1, 0.0, mat, ns, st);
      hipLaunchKernelGGL(k_residual, dim3(1), dim3(256), 0, st, mat, ns, ns, y, dv, metrics_out);
    }
  }
  if (real_params)
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, st, x_out, npx, 2.0);
  if (metrics_out) hipLaunchKernelGGL(k_info, dim3(1), dim3(1), 0, st, info, metrics_out + 1);
  return 0;
}

int vmc_sr_solve(int n_samples, int64_t n_sliced, int phys_dim, const int32_t *params_per_site, int n_sites,
                 const void *o, int o_is_f32, const int8_t *p, const double *eloc, int space, double diag_shift,
                 double grad_factor, int real_params, double *x_out, double *metrics_out, void *workspace,
                 size_t workspace_bytes, void *stream) {
  return sr_common(n_samples, n_sliced, phys_dim, params_per_site, n_sites, o, o_is_f32, p, eloc, space, diag_shift,
                   grad_factor, real_params, x_out, metrics_out, nullptr, workspace, workspace_bytes, stream);
}

int vmc_qgt_dense(int n_samples, int64_t n_sliced, int phys_dim, const int32_t *params_per_site, int n_sites,
                  const void *o, int o_is_f32, const int8_t *p, int space, double *dense_out, void *workspace,
                  size_t workspace_bytes, void *stream) {
  if (!dense_out) return vmc::sr_fail("null argument");
  return sr_common(n_samples, n_sliced, phys_dim, params_per_site, n_sites, o, o_is_f32, p, nullptr, space, 0.0,
                   1.0, 0, nullptr, nullptr, dense_out, workspace, workspace_bytes, stream);
}

}  // extern "C"
