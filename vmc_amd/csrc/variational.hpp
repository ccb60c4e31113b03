// Variational boundary compression on gfx950: QR-initialised one-site ALS
// (src/vmc/peps/common/strategy.py:179-427, QR from
// src/vmc/utils/factorizations.py:9-54).
//
// One row application, per sample (one 256-thread workgroup per sample, one
// launch per site step for the whole batch):
//   init   theta_c = carry . m_c . w_c (k_theta); thin Householder QR of the
//          (left*q) x (Dr*wr) theta matrix; keep Q[:, :k] as the result tensor
//          and carry = Q[:, :k]^H theta = R[:k, :]   (k_var_qr_init; scratch
//          panels beyond the register tile: k_var_qr_init_g)
//   R envs R_{c-1} = conj(res_c) . R_c . w_c . m_c              (k_var_renv)
//   sweeps LR: opt_c = L_c . m_c . w_c . R_c; opt_c = A_c C_c (QR);
//              L_{c+1} = (L_c . m_c . w_c) . conj(A_c)          (k_var_lr)
//          RL: opt_c = L_c . m_c . w_c . R_c; opt_c = C_c B_c (LQ);
//              R_{c-1} = conj(B_c) . R_c . w_c . m_c; at c = 0 the optimal
//              tensor is the result's first tensor             (k_var_rl)
// Contraction orders are the reference's pairwise `optimize` paths. The C
// factors the reference absorbs into the neighbouring tensor are overwritten
// by that neighbour's own optimal tensor before any read, so they are not
// formed here. Any QR of the same matrix differs from LAPACK's by a diagonal
// phase gauge on the new bond, which leaves the compressed state (and the
// column span the next truncated QR keeps) unchanged.
//
// QR of an m x n panel (m <= 128, n <= 32): the panel lives in registers,
// thread t owning column t & 31 and rows (t >> 5) + 8 i, i < 16. Reflectors
// follow LAPACK's xLARFG convention (beta = -sign(Re alpha) |x|), so Q agrees
// with geqrf/orgqr to rounding. Factorisations run in FP64 (complex128 for
// complex plans) whatever the storage type.
#pragma once

#include "common.hpp"

namespace vmc {

constexpr int VQR_R = 16;   // rows per thread of the register panel (VQR_M x VQR_N, common.hpp)

__device__ inline double vcj(double v) { return v; }
__device__ inline cf64 vcj(cf64 v) { return {v.x, -v.y}; }
__device__ inline double vre(double v) { return v; }
__device__ inline double vre(cf64 v) { return v.x; }
__device__ inline double vim(double) { return 0.0; }
__device__ inline double vim(cf64 v) { return v.y; }
__device__ inline double vxor32(double v) { return __shfl_xor(v, 32, 64); }
__device__ inline cf64 vxor32(cf64 v) { return {__shfl_xor(v.x, 32, 64), __shfl_xor(v.y, 32, 64)}; }
template <typename A>
__device__ inline A vmk(double re, double im);
template <>
__device__ inline double vmk<double>(double re, double) { return re; }
template <>
__device__ inline cf64 vmk<cf64>(double re, double im) { return {re, im}; }
__device__ inline float vcjT(float v) { return v; }
__device__ inline double vcjT(double v) { return v; }
__device__ inline cf32 vcjT(cf32 v) { return {v.x, -v.y}; }
__device__ inline cf64 vcjT(cf64 v) { return {v.x, -v.y}; }

template <typename A>
struct QrShared {
  A xs[2][VQR_M];      // column j of the panel (published by its owners)
  A rowj[2][VQR_N];    // row j of the panel
  A red[2][4][VQR_N];  // per-wave partial column dots
  A tau[VQR_N];
};

// Householder QR of the m x n register panel P: afterwards P holds R on and
// above the diagonal and the reflector vectors (unit leading entry implied)
// below it, as xGEQR2 leaves them; tau in sm.tau. kref = min(m, n).
template <typename A>
__device__ inline void vqr_factor(A (&P)[VQR_R], int m, int n, QrShared<A> &sm) {
  const int t = threadIdx.x, c = t & 31, g = t >> 5, lane = t & 63, wave = t >> 6;
  const int kref = min(m, n);
  for (int j = 0; j < kref; ++j) {
    const int b = j & 1;
    if (c == j) {
#pragma unroll
      for (int i = 0; i < VQR_R; ++i) {
        const int r = g + 8 * i;
        if (r >= j && r < m) sm.xs[b][r] = P[i];
      }
    }
    if (g == (j & 7)) {
#pragma unroll
      for (int i = 0; i < VQR_R; ++i)
        if (i == (j >> 3)) sm.rowj[b][c] = P[i];
    }
    __syncthreads();
    A d = A(0.0);
#pragma unroll
    for (int i = 0; i < VQR_R; ++i) {
      const int r = g + 8 * i;
      if (r > j && r < m) d = d + vcj(sm.xs[b][r]) * P[i];
    }
    d = d + vxor32(d);
    if (lane < 32) sm.red[b][wave][c] = d;
    __syncthreads();
    const A alpha = sm.xs[b][j];
    const double sigma = vre(sm.red[b][0][j] + sm.red[b][1][j] + sm.red[b][2][j] + sm.red[b][3][j]);
    const A dc = sm.red[b][0][c] + sm.red[b][1][c] + sm.red[b][2][c] + sm.red[b][3][c];
    A tau, scale;
    double beta;
    if (sigma == 0.0 && vim(alpha) == 0.0) {
      // xLARFG: H = I when x(2:) is zero and alpha is real
      tau = A(0.0);
      scale = A(0.0);
      beta = vre(alpha);
    } else {
      const double ar = vre(alpha), ai = vim(alpha);
      const double nrm = sqrt(ar * ar + ai * ai + sigma);
      beta = ar >= 0.0 ? -nrm : nrm;
      tau = vmk<A>((beta - ar) / beta, -ai / beta);
      // scale = 1 / (alpha - beta)
      const double er = ar - beta, ei = ai;
      const double den = 1.0 / (er * er + ei * ei);
      scale = vmk<A>(er * den, -ei * den);
    }
    const A wc = sm.rowj[b][c] + vcj(scale) * dc;  // v^H P[:, c]
#pragma unroll
    for (int i = 0; i < VQR_R; ++i) {
      const int r = g + 8 * i;
      if (r < m && r >= j) {
        if (c > j && c < n) {
          const A vr = r == j ? A(1.0) : sm.xs[b][r] * scale;
          P[i] = P[i] - vcj(tau) * vr * wc;
        } else if (c == j) {
          P[i] = r == j ? A(beta) : sm.xs[b][r] * scale;
        }
      }
    }
    if (t == 0) sm.tau[j] = tau;
  }
  __syncthreads();
}

// Explicit Q (m x kq, kq <= kref) of the factored panel (xUNG2R): Q holds the
// same register layout as P.
template <typename A>
__device__ inline void vqr_form_q(const A (&P)[VQR_R], A (&Q)[VQR_R], int m, int kq, QrShared<A> &sm) {
  const int t = threadIdx.x, c = t & 31, g = t >> 5, lane = t & 63, wave = t >> 6;
#pragma unroll
  for (int i = 0; i < VQR_R; ++i) Q[i] = A((g + 8 * i) == c && c < kq ? 1.0 : 0.0);
  for (int j = kq - 1; j >= 0; --j) {
    const int b = j & 1;
    if (c == j) {
#pragma unroll
      for (int i = 0; i < VQR_R; ++i) {
        const int r = g + 8 * i;
        if (r > j && r < m) sm.xs[b][r] = P[i];
      }
    }
    __syncthreads();
    const A tau = sm.tau[j];
    // w_c = v^H Q[:, c] for c > j (v_j = 1)
    A d = A(0.0);
#pragma unroll
    for (int i = 0; i < VQR_R; ++i) {
      const int r = g + 8 * i;
      if (r < m && r >= j) d = d + (r == j ? Q[i] : vcj(sm.xs[b][r]) * Q[i]);
    }
    d = d + vxor32(d);
    if (lane < 32) sm.red[b][wave][c] = d;
    __syncthreads();
    const A wc = sm.red[b][0][c] + sm.red[b][1][c] + sm.red[b][2][c] + sm.red[b][3][c];
#pragma unroll
    for (int i = 0; i < VQR_R; ++i) {
      const int r = g + 8 * i;
      if (r >= m) continue;
      const A vr = r == j ? A(1.0) : (r > j ? sm.xs[b][r] : A(0.0));
      if (c > j && c < kq) {
        if (r >= j) Q[i] = Q[i] - tau * vr * wc;
      } else if (c == j) {
        Q[i] = r == j ? A(1.0) - tau : (r > j ? A(0.0) - tau * vr : A(0.0));
      }
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------------------- //
// Panels beyond the register tile (m > 128 rows or n > 32 columns, up to
// VQG_M x VQG_N): the same reflectors as xGEQR2 / xUNG2R on a panel in global
// memory, element (r, c) at P[r * rs + c * cs], applied in blocks of VQB: a
// block reduction forms each reflector, the block's own columns take it at
// once, and every later column takes the whole block from registers.
// ------------------------------------------------------------------------- //

template <typename A>
__device__ inline A vwave_sum(A v);
template <>
__device__ inline double vwave_sum<double>(double v) {
  return wave_sum(v);
}
template <>
__device__ inline cf64 vwave_sum<cf64>(cf64 v) {
  return {wave_sum(v.x), wave_sum(v.y)};
}

// Reflectors per block of the global-panel routines, and rows per lane of a
// register-held panel column (m <= VQG_M).
constexpr int VQB = 8;
constexpr int VQE = VQG_M / 64;

// Entry r of reflector j of a factored panel: 1 at r = j, the stored vector
// below, zero above.
template <typename A>
__device__ inline A vqr_refl(const A *P, int64_t rs, int64_t cs, int j, int r) {
  return r == j ? vmk<A>(1.0, 0.0) : (r > j ? P[(int64_t)r * rs + (int64_t)j * cs] : vmk<A>(0.0, 0.0));
}

// x <- (I - f v v^H) x for one register-held column (rows lane + 64 e), f =
// tau or conj(tau) of reflector j; rows below lo are zero in v and skipped.
template <typename A>
__device__ inline void vqr_apply_col(A (&x)[VQE], const A *P, int64_t rs, int64_t cs, int m, int j, A f) {
  const int lane = threadIdx.x & 63;
  A v[VQE];
  A d = vmk<A>(0.0, 0.0);
#pragma unroll
  for (int e = 0; e < VQE; ++e) {
    const int r = lane + 64 * e;
    v[e] = (r < m) ? vqr_refl<A>(P, rs, cs, j, r) : vmk<A>(0.0, 0.0);
    d = d + vcj(v[e]) * x[e];
  }
  const A g = f * vwave_sum<A>(d);
#pragma unroll
  for (int e = 0; e < VQE; ++e) x[e] = x[e] - g * v[e];
}

// Householder QR of the m x n global panel: R on and above the diagonal,
// reflector vectors (unit leading entry implied) below it, tau in sTau.
// Blocked by VQB reflectors: a block's reflectors are formed from its own
// columns (updating only those), then each later column, held in registers
// by one wave, takes the whole block at once, so it is read and written once
// per block instead of once per reflector.
template <typename A>
__device__ inline void vqr_factor_g(A *P, int64_t rs, int64_t cs, int m, int n, A *sV, A *sTau, double *red) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nw = blockDim.x >> 6;
  const int kref = min(m, n);
  for (int j0 = 0; j0 < kref; j0 += VQB) {
    const int j1 = min(kref, j0 + VQB);
    for (int j = j0; j < j1; ++j) {
      const A alpha = P[(int64_t)j * rs + (int64_t)j * cs];
      double sigma = 0.0;
      for (int r = j + 1 + t; r < m; r += blockDim.x) sigma += abs2(P[(int64_t)r * rs + (int64_t)j * cs]);
      sigma = block_sum(sigma, red);  // its barriers also order every read of alpha before the writes below
      A tau, scale;
      double beta;
      if (sigma == 0.0 && vim(alpha) == 0.0) {
        tau = vmk<A>(0.0, 0.0);
        scale = vmk<A>(0.0, 0.0);
        beta = vre(alpha);
      } else {
        const double ar = vre(alpha), ai = vim(alpha);
        const double nrm = sqrt(ar * ar + ai * ai + sigma);
        beta = ar >= 0.0 ? -nrm : nrm;
        tau = vmk<A>((beta - ar) / beta, -ai / beta);
        const double er = ar - beta, ei = ai;
        const double den = 1.0 / (er * er + ei * ei);
        scale = vmk<A>(er * den, -ei * den);
      }
      for (int r = j + 1 + t; r < m; r += blockDim.x) {
        A &x = P[(int64_t)r * rs + (int64_t)j * cs];
        const A v = x * scale;
        sV[r] = v;
        x = v;
      }
      if (t == 0) {
        sV[j] = vmk<A>(1.0, 0.0);
        sTau[j] = tau;
      }
      __syncthreads();
      if (vre(tau) != 0.0 || vim(tau) != 0.0) {
        // P[:, c] -= conj(tau) v (v^H P[:, c]) for the block's columns c > j
        for (int c = j + 1 + wave; c < j1; c += nw) {
          A d = vmk<A>(0.0, 0.0);
          for (int r = j + lane; r < m; r += 64) d = d + vcj(sV[r]) * P[(int64_t)r * rs + (int64_t)c * cs];
          const A f = vcj(tau) * vwave_sum<A>(d);
          for (int r = j + lane; r < m; r += 64) {
            A &x = P[(int64_t)r * rs + (int64_t)c * cs];
            x = x - f * sV[r];
          }
        }
      }
      if (t == 0) P[(int64_t)j * rs + (int64_t)j * cs] = vmk<A>(beta, 0.0);
      __syncthreads();
    }
    // later columns: H_{j1-1}^H ... H_{j0}^H applied in order, rows >= j0
    for (int c = j1 + wave; c < n; c += nw) {
      A x[VQE];
#pragma unroll
      for (int e = 0; e < VQE; ++e) {
        const int r = lane + 64 * e;
        x[e] = (r >= j0 && r < m) ? P[(int64_t)r * rs + (int64_t)c * cs] : vmk<A>(0.0, 0.0);
      }
      for (int j = j0; j < j1; ++j) {
        const A tj = sTau[j];
        if (vre(tj) != 0.0 || vim(tj) != 0.0) vqr_apply_col<A>(x, P, rs, cs, m, j, vcj(tj));
      }
#pragma unroll
      for (int e = 0; e < VQE; ++e) {
        const int r = lane + 64 * e;
        if (r >= j0 && r < m) P[(int64_t)r * rs + (int64_t)c * cs] = x[e];
      }
    }
    __syncthreads();
  }
}

// Explicit Q (m x kq, column-major with leading dimension m) of a panel
// factored by vqr_factor_g: Q = H_0 H_1 ... H_{kq-1} [I; 0], by blocks of VQB
// reflectors from the last; each column of Q at or right of the block, held in
// registers by one wave, takes the block's reflectors at once.
template <typename A>
__device__ inline void vqr_form_q_g(const A *P, int64_t rs, int64_t cs, A *Q, int m, int kq, const A *sTau) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nw = blockDim.x >> 6;
  for (int64_t i = t; i < (int64_t)m * kq; i += blockDim.x) {
    const int c = (int)(i / m), r = (int)(i - (int64_t)c * m);
    Q[i] = vmk<A>(r == c ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int j1 = kq; j1 > 0; j1 -= VQB) {
    const int j0 = max(0, j1 - VQB);
    for (int c = j0 + wave; c < kq; c += nw) {
      A *qc = Q + (int64_t)c * m;
      A x[VQE];
#pragma unroll
      for (int e = 0; e < VQE; ++e) {
        const int r = lane + 64 * e;
        x[e] = (r >= j0 && r < m) ? qc[r] : vmk<A>(0.0, 0.0);
      }
      // H_j e_c = e_c for j > c: column c starts at reflector min(c, j1 - 1)
      for (int j = min(c, j1 - 1); j >= j0; --j) {
        const A tj = sTau[j];
        if (vre(tj) != 0.0 || vim(tj) != 0.0) vqr_apply_col<A>(x, P, rs, cs, m, j, tj);
      }
#pragma unroll
      for (int e = 0; e < VQE; ++e) {
        const int r = lane + 64 * e;
        if (r >= j0 && r < m) qc[r] = x[e];
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------- //
// Per-site contractions (one sample, whole workgroup)
// ------------------------------------------------------------------------- //

// MPO slice view w[wl, wr, p, q] inside the site tensor slice.
struct WView {
  int64_t sl, sr, sp, sq;
  int wl, wr, P, Q;
};

// opt[a,q,b] = sum L[d,l,a] m[d,p,r] w[l,w,p,q] R[r,w,b] with the reference's
// path [(1,0),(2,0),(1,0)]: Y1 = L.m over d, Y2 = Y1.w over (l,p), opt = Y2.R.
// Leaves Y2[a,r,w,q] in y2 for the left-environment update.
template <typename T>
__device__ inline void var_optimal(const T *L, const T *m, int Dl, int Dr, const T *w, const WView &v, const T *R,
                                   int rl, int rr, T *y1, T *y2, T *opt) {
  const int P = v.P, Q = v.Q, wl = v.wl, wr = v.wr;
  gemm2<T, true>(L, g2(wl * rl, 1), g2(Dl, (int64_t)wl * rl),
           m, g2(Dl, (int64_t)P * Dr), g2(P * Dr, 1),
           y1, g2(wl * rl, (int64_t)P * Dr), g2(P * Dr, 1));
  __syncthreads();
  gemm2<T, true>(y1, g2(rl, (int64_t)P * Dr, Dr, 1), g2(wl, (int64_t)rl * P * Dr, P, Dr),
           w, g2(wl, v.sl, P, v.sp), g2(wr, v.sr, Q, v.sq),
           y2, g2(rl * Dr, (int64_t)wr * Q), g2(wr * Q, 1));
  __syncthreads();
  gemm2<T, true>(y2, g2(rl, (int64_t)Dr * wr * Q, Q, 1), g2(Dr * wr, Q),
           R, g2(Dr * wr, rr), g2(rr, 1),
           opt, g2(rl * Q, rr), g2(rr, 1));
  __syncthreads();
}

// Rout[d,l,a] = sum cres[a,q,b] R[r,w,b] w[l,w,p,q] m[d,p,r] (cres = conj of
// the result tensor, dims (ra, Q, rr)), path [(1,0),(2,0),(1,0)]:
// X1 = R.cres over b, X2 = X1.w over (w,q), Rout = X2.m over (r,p).
template <typename T>
__device__ inline void var_right_update(const T *cres, int ra, int rr, const T *R, const T *w, const WView &v,
                                        const T *m, int Dl, int Dr, T *x1, T *x2, T *Rout) {
  const int P = v.P, Q = v.Q, wl = v.wl, wr = v.wr;
  gemm2<T, true>(R, g2(Dr * wr, rr), g2(rr, 1),
           cres, g2(rr, 1), g2(ra * Q, rr),
           x1, g2(Dr * wr, (int64_t)ra * Q), g2(ra * Q, 1));
  __syncthreads();
  gemm2<T, true>(x1, g2(Dr, (int64_t)wr * ra * Q, ra, Q), g2(wr, (int64_t)ra * Q, Q, 1),
           w, g2(wr, v.sr, Q, v.sq), g2(wl, v.sl, P, v.sp),
           x2, g2(Dr * ra, (int64_t)wl * P), g2(wl * P, 1));
  __syncthreads();
  gemm2<T, true>(x2, g2(ra, (int64_t)wl * P, wl, P), g2(Dr, (int64_t)ra * wl * P, P, 1),
           m, g2(Dr, 1, P, Dr), g2(Dl, (int64_t)P * Dr),
           Rout, g2(ra, 1, wl, ra), g2(Dl, (int64_t)wl * ra));
  __syncthreads();
}

// conj copy of n elements (complex plans only; real plans pass the tensor).
template <typename T>
__device__ inline const T *var_conj(const T *x, int n, T *scratch) {
  if constexpr (!Scalar<T>::complex) {
    return x;
  } else {
    for (int i = threadIdx.x; i < n; i += blockDim.x) scratch[i] = vcjT(x[i]);
    __syncthreads();
    return scratch;
  }
}

// ------------------------------------------------------------------------- //
// Kernels
// ------------------------------------------------------------------------- //

// Truncated QR of the init pass: u = Q[:, :k] of the (m x n) theta matrix,
// carry = Q[:, :k]^H theta normalised, logscale += log |carry|_F
// (strategy.py:222-254; factorizations.py:9-16).
template <typename T>
__global__ __launch_bounds__(256) void k_var_qr_init(TruncArgs<T> a) {
  using A = typename Scalar<T>::acc;
  __shared__ QrShared<A> sm;
  __shared__ A Qs[VQR_M * VQR_N];
  __shared__ double red[32];
  const int b = blockIdx.x;
  const T *mat = a.mat + (int64_t)b * a.mat_stride;
  const int t = threadIdx.x, c = t & 31, g = t >> 5;
  const int m = a.m, n = a.n, k = a.k;
  A P[VQR_R], Qr[VQR_R];
#pragma unroll
  for (int i = 0; i < VQR_R; ++i) {
    const int r = g + 8 * i;
    P[i] = (r < m && c < k) ? to_acc(mat[(int64_t)r * n + c]) : A(0.0);
  }
  vqr_factor<A>(P, m, k, sm);
  vqr_form_q<A>(P, Qr, m, k, sm);
  T *u = a.u + (int64_t)b * a.u_stride;
#pragma unroll
  for (int i = 0; i < VQR_R; ++i) {
    const int r = g + 8 * i;
    if (r < m && c < k) {
      u[(int64_t)r * k + c] = from_acc<T>(Qr[i]);
      Qs[r * VQR_N + c] = Qr[i];
    }
  }
  __syncthreads();
  // carry[j, col] = sum_r conj(Q[r, j]) theta[r, col]; thread -> column t & 127,
  // rows j = (t >> 7) + 2 jj of the carry
  const int col = t & 127, jg = t >> 7;
  A acc[16];
#pragma unroll
  for (int jj = 0; jj < 16; ++jj) acc[jj] = A(0.0);
  if (col < n) {
    for (int r = 0; r < m; ++r) {
      const A x = to_acc(mat[(int64_t)r * n + col]);
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) {
        const int j = jg + 2 * jj;
        if (j < k) acc[jj] = acc[jj] + vcj(Qs[r * VQR_N + j]) * x;
      }
    }
  }
  double s = 0.0;
#pragma unroll
  for (int jj = 0; jj < 16; ++jj) s += abs2(acc[jj]);
  const double nrm2 = block_sum(s, red);
  const double inv = nrm2 > 0.0 ? 1.0 / sqrt(nrm2) : 1.0;
  T *carry = a.carry + (int64_t)b * a.carry_stride;
  if (col < n) {
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) {
      const int j = jg + 2 * jj;
      if (j < k) carry[(int64_t)j * n + col] = from_acc<T>(inv * acc[jj]);
    }
  }
  if (t == 0 && nrm2 > 0.0) a.logscale[(int64_t)b * a.log_stride] += 0.5 * log(nrm2);
}

// The same for panels beyond the register tile (m > VQR_M, k > VQR_N or
// n > 128): the panel and Q live in the per-sample scratch. A kernel of its
// own, so the register kernel keeps its occupancy.
template <typename T>
__global__ __launch_bounds__(256) void k_var_qr_init_g(TruncArgs<T> a) {
  using A = typename Scalar<T>::acc;
  __shared__ double red[32];
  __shared__ A sV[VQG_M], sTau[VQG_N];
  const int b = blockIdx.x;
  const T *mat = a.mat + (int64_t)b * a.mat_stride;
  const int t = threadIdx.x;
  const int m = a.m, n = a.n, k = a.k;
  // generic panel: copy theta[:, :k] column-major to the scratch, factor,
  // form Q, then carry = Q^H theta normalised (norm pass, then write pass)
  A *Pg = (A *)a.gscr + (int64_t)b * a.gscr_stride;
  A *Qg = Pg + (int64_t)m * k;
  for (int64_t i = t; i < (int64_t)m * k; i += blockDim.x) {
    const int cc = (int)(i / m), r = (int)(i - (int64_t)cc * m);
    Pg[i] = to_acc(mat[(int64_t)r * n + cc]);
  }
  __syncthreads();
  vqr_factor_g<A>(Pg, 1, m, m, k, sV, sTau, red);
  vqr_form_q_g<A>(Pg, 1, m, Qg, m, k, sTau);
  T *u = a.u + (int64_t)b * a.u_stride;
  for (int64_t i = t; i < (int64_t)m * k; i += blockDim.x) {
    const int r = (int)(i / k), cc = (int)(i - (int64_t)r * k);
    u[i] = from_acc<T>(Qg[(int64_t)cc * m + r]);
  }
  if constexpr (std::is_same<T, double>::value) {
    // FP64: carry = Q^T theta on MFMA straight into the output, then normalised in place
    T *carry = a.carry + (int64_t)b * a.carry_stride;
    gemm2<T, true>(Qg, g2(k, m), g2(m, 1), mat, g2(m, n), g2(n, 1), carry, g2(k, n), g2(n, 1));
    __syncthreads();
    double s2 = 0.0;
    for (int64_t i = t; i < (int64_t)k * n; i += blockDim.x) s2 += carry[i] * carry[i];
    const double nrm2g = block_sum(s2, red);
    const double invg = nrm2g > 0.0 ? 1.0 / sqrt(nrm2g) : 1.0;
    for (int64_t i = t; i < (int64_t)k * n; i += blockDim.x) carry[i] *= invg;
    if (t == 0 && nrm2g > 0.0) a.logscale[(int64_t)b * a.log_stride] += 0.5 * log(nrm2g);
    return;
  }
  // a thread forms CL carry rows of one column, so theta is read once per CL rows
  constexpr int CL = 8;
  const int kb = (k + CL - 1) / CL;
  auto carry_block = [&](int64_t item, A (&acc)[CL]) {
    const int jb = (int)(item / n), col = (int)(item - (int64_t)jb * n);
    const A *qj[CL];
#pragma unroll
    for (int q = 0; q < CL; ++q) {
      acc[q] = vmk<A>(0.0, 0.0);
      qj[q] = Qg + (int64_t)min(jb * CL + q, k - 1) * m;  // rows past k are computed and dropped
    }
    for (int r = 0; r < m; ++r) {
      const A x = to_acc(mat[(int64_t)r * n + col]);
#pragma unroll
      for (int q = 0; q < CL; ++q) acc[q] = acc[q] + vcj(qj[q][r]) * x;
    }
  };
  double s2 = 0.0;
  for (int64_t i = t; i < (int64_t)kb * n; i += blockDim.x) {
    A acc[CL];
    carry_block(i, acc);
    const int jb = (int)(i / n);
#pragma unroll
    for (int q = 0; q < CL; ++q)
      if (jb * CL + q < k) s2 += abs2(acc[q]);
  }
  const double nrm2g = block_sum(s2, red);
  const double invg = nrm2g > 0.0 ? 1.0 / sqrt(nrm2g) : 1.0;
  T *carry = a.carry + (int64_t)b * a.carry_stride;
  for (int64_t i = t; i < (int64_t)kb * n; i += blockDim.x) {
    A acc[CL];
    carry_block(i, acc);
    const int jb = (int)(i / n), col = (int)(i - (int64_t)jb * n);
#pragma unroll
    for (int q = 0; q < CL; ++q) {
      const int j = jb * CL + q;
      if (j < k) carry[(int64_t)j * n + col] = from_acc<T>(invg * acc[q]);
    }
  }
  if (t == 0 && nrm2g > 0.0) a.logscale[(int64_t)b * a.log_stride] += 0.5 * log(nrm2g);
  return;
}

template <typename T>
struct VarArgs {
  const T *m;            // incoming boundary tensor (Dl, P, Dr)
  int64_t m_stride;
  int Dl, Dr;
  SiteRef<T> site;
  WView w;
  const T *L;            // (Dl, wl, rl)
  const T *R;            // (Dr, wr, rr)
  int64_t env_stride;
  int rl, rr;
  const T *res;          // k_var_renv: result tensor (ra, Q, rr) of this site
  int64_t res_stride;
  int ra;
  T *env_out;            // L_{c+1} (Dr, wr, kk) or R_{c-1} (Dl, wl, kk)
  T *y1, *y2, *opt, *cj;
  int64_t s_stride;
  T *out;                // k_var_rl: B_c (kk, Q, rr), or the normalised first tensor
  int64_t out_stride;
  const double *log_in;  // k_var_rl at c = 0: log_out = log_in + log |opt|_F
  double *log_out;
  int64_t log_in_stride, log_out_stride;
  int first;             // k_var_rl at c = 0
  void *gscr;            // panels beyond the register tile: per-sample FP64 scratch for Q
  int64_t gscr_stride;   // elements of the accumulator type
};

template <typename T>
__global__ __launch_bounds__(256) void k_var_renv(VarArgs<T> a) {
  const int b = blockIdx.x;
  const int64_t so = (int64_t)b * a.s_stride;
  const int n = a.ra * a.w.Q * a.rr;
  const T *cres = var_conj<T>(a.res + (int64_t)b * a.res_stride, n, a.cj + so);
  var_right_update<T>(cres, a.ra, a.rr, a.R + (int64_t)b * a.env_stride, a.site.get(b), a.w,
                      a.m + (int64_t)b * a.m_stride, a.Dl, a.Dr, a.y1 + so, a.y2 + so,
                      a.env_out + (int64_t)b * a.env_stride);
}

template <typename T>
__global__ __launch_bounds__(256) void k_var_lr(VarArgs<T> a) {
  using A = typename Scalar<T>::acc;
  __shared__ QrShared<A> sm;
  const int b = blockIdx.x;
  const int64_t so = (int64_t)b * a.s_stride;
  const T *w = a.site.get(b);
  T *opt = a.opt + so, *y2 = a.y2 + so, *cj = a.cj + so;
  var_optimal<T>(a.L + (int64_t)b * a.env_stride, a.m + (int64_t)b * a.m_stride, a.Dl, a.Dr, w, a.w,
                 a.R + (int64_t)b * a.env_stride, a.rl, a.rr, a.y1 + so, y2, opt);
  const int t = threadIdx.x, c = t & 31, g = t >> 5;
  const int m = a.rl * a.w.Q, n = a.rr, kk = min(m, n);
  if (m > VQR_M || n > VQR_N) {
    __shared__ A sV[VQG_M], sTau[VQG_N];
    __shared__ double redg[32];
    // opt (row-major m x n) is copied column-major into the scratch after Q's
    // block (widened for FP32 storage), so the Householder steps read and
    // write its columns coalesced; Q goes to the scratch
    A *Qg = (A *)a.gscr + (int64_t)b * a.gscr_stride;
    A *Pw = Qg + (int64_t)m * kk;
    for (int64_t i = t; i < (int64_t)m * n; i += blockDim.x) {
      const int cc = (int)(i / m), r = (int)(i - (int64_t)cc * m);
      Pw[i] = to_acc(opt[(int64_t)r * n + cc]);
    }
    __syncthreads();
    vqr_factor_g<A>(Pw, 1, m, m, n, sV, sTau, redg);
    vqr_form_q_g<A>(Pw, 1, m, Qg, m, kk, sTau);
    for (int64_t i = t; i < (int64_t)m * kk; i += blockDim.x) {
      const int r = (int)(i / kk), cc = (int)(i - (int64_t)r * kk);
      cj[i] = from_acc<T>(vcj(Qg[(int64_t)cc * m + r]));
    }
    __syncthreads();
  } else {
    A P[VQR_R], Qr[VQR_R];
#pragma unroll
    for (int i = 0; i < VQR_R; ++i) {
      const int r = g + 8 * i;
      P[i] = (r < m && c < n) ? to_acc(opt[(int64_t)r * n + c]) : A(0.0);
    }
    vqr_factor<A>(P, m, n, sm);
    vqr_form_q<A>(P, Qr, m, kk, sm);
#pragma unroll
    for (int i = 0; i < VQR_R; ++i) {
      const int r = g + 8 * i;
      if (r < m && c < kk) cj[(int64_t)r * kk + c] = from_acc<T>(vcj(Qr[i]));
    }
    __syncthreads();
  }
  // L_{c+1}[r,w,b] = sum_{a,q} Y2[a,r,w,q] conj(A)[a,q,b]  (path step 3 of the L update)
  const int Q = a.w.Q, wr = a.w.wr;
  gemm2<T, true>(y2, g2(a.Dr * wr, Q), g2(a.rl, (int64_t)a.Dr * wr * Q, Q, 1),
           cj, g2(a.rl * Q, kk), g2(kk, 1),
           a.env_out + (int64_t)b * a.env_stride, g2(a.Dr * wr, kk), g2(kk, 1));
}

template <typename T>
__global__ __launch_bounds__(256) void k_var_rl(VarArgs<T> a) {
  using A = typename Scalar<T>::acc;
  __shared__ QrShared<A> sm;
  __shared__ double red[32];
  const int b = blockIdx.x;
  const int64_t so = (int64_t)b * a.s_stride;
  const T *w = a.site.get(b);
  const T *m_in = a.m + (int64_t)b * a.m_stride;
  const T *R = a.R + (int64_t)b * a.env_stride;
  T *opt = a.opt + so, *cj = a.cj + so;
  var_optimal<T>(a.L + (int64_t)b * a.env_stride, m_in, a.Dl, a.Dr, w, a.w, R, a.rl, a.rr, a.y1 + so, a.y2 + so,
                 opt);
  const int Q = a.w.Q;
  T *out = a.out ? a.out + (int64_t)b * a.out_stride : nullptr;
  if (a.first) {
    const int n = a.rl * Q * a.rr;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += abs2(opt[i]);
    const double nrm2 = block_sum(s, red);
    const double inv = nrm2 > 0.0 ? 1.0 / sqrt(nrm2) : 1.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) out[i] = from_acc<T>(inv * to_acc(opt[i]));
    if (threadIdx.x == 0)
      a.log_out[(int64_t)b * a.log_out_stride] =
          a.log_in[(int64_t)b * a.log_in_stride] + (nrm2 > 0.0 ? 0.5 * log(nrm2) : -INFINITY);
    return;
  }
  // LQ of opt as rl x (Q rr): QR of its transpose, B = Q_t^T
  const int t = threadIdx.x, c = t & 31, g = t >> 5;
  const int m = Q * a.rr, n = a.rl, kk = min(m, n);
  // non-final sweeps keep only the environment (a.out is null)
  const bool keep = a.out != nullptr;
  if (m > VQR_M || n > VQR_N) {
    __shared__ A sV[VQG_M], sTau[VQG_N];
    // the transpose of opt is column-major m x n with leading dimension m
    A *Qg = (A *)a.gscr + (int64_t)b * a.gscr_stride;
    A *Pw = (A *)opt;
    if constexpr (sizeof(A) != sizeof(T)) {
      Pw = Qg + (int64_t)m * kk;
      for (int64_t i = t; i < (int64_t)m * n; i += blockDim.x) Pw[i] = to_acc(opt[i]);
      __syncthreads();
    }
    vqr_factor_g<A>(Pw, 1, m, m, n, sV, sTau, red);
    vqr_form_q_g<A>(Pw, 1, m, Qg, m, kk, sTau);
    for (int64_t i = t; i < (int64_t)m * kk; i += blockDim.x) {
      if (keep) out[i] = from_acc<T>(Qg[i]);
      cj[i] = from_acc<T>(vcj(Qg[i]));
    }
    __syncthreads();
  } else {
    A P[VQR_R], Qr[VQR_R];
#pragma unroll
    for (int i = 0; i < VQR_R; ++i) {
      const int r = g + 8 * i;
      P[i] = (r < m && c < n) ? to_acc(opt[(int64_t)c * m + r]) : A(0.0);
    }
    vqr_factor<A>(P, m, n, sm);
    vqr_form_q<A>(P, Qr, m, kk, sm);
#pragma unroll
    for (int i = 0; i < VQR_R; ++i) {
      const int r = g + 8 * i;
      if (r < m && c < kk) {
        if (keep) out[(int64_t)c * m + r] = from_acc<T>(Qr[i]);
        cj[(int64_t)c * m + r] = from_acc<T>(vcj(Qr[i]));
      }
    }
    __syncthreads();
  }
  var_right_update<T>(cj, kk, a.rr, R, w, a.w, m_in, a.Dl, a.Dr, a.y1 + so, a.y2 + so,
                      a.env_out + (int64_t)b * a.env_stride);
}

}  // namespace vmc
