// Shared device helpers of the gfx950 PEPS engine.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

namespace vmc {

// ---------------------------------------------------------------------------
// Scalar types: float, double and the complex pairs cf32 / cf64 (interleaved
// re, im -- the layout of numpy/torch complex64 and complex128).
// ---------------------------------------------------------------------------
template <typename R>
struct cplx {
  R x, y;
  cplx() = default;  // trivial, so cplx can live in __shared__ arrays
  __host__ __device__ cplx(R re) : x(re), y(0) {}
  __host__ __device__ cplx(R re, R im) : x(re), y(im) {}
  __host__ __device__ cplx &operator*=(const cplx &b) {
    const R t = x * b.x - y * b.y;
    y = x * b.y + y * b.x;
    x = t;
    return *this;
  }
  __host__ __device__ cplx &operator+=(const cplx &b) {
    x += b.x;
    y += b.y;
    return *this;
  }
};
using cf32 = cplx<float>;
using cf64 = cplx<double>;
using ::fma;  // keep the real overloads visible next to the complex one below

template <typename R>
__host__ __device__ inline cplx<R> operator+(cplx<R> a, cplx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__host__ __device__ inline cplx<R> operator-(cplx<R> a, cplx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__host__ __device__ inline cplx<R> operator*(cplx<R> a, cplx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__host__ __device__ inline cplx<R> operator*(R s, cplx<R> a) { return {s * a.x, s * a.y}; }
template <typename R>
__host__ __device__ inline cplx<R> conj(cplx<R> a) { return {a.x, -a.y}; }
// a * b + c (holomorphic, no conjugation)
template <typename R>
__host__ __device__ inline cplx<R> fma(cplx<R> a, cplx<R> b, cplx<R> c) {
  return {(R)fma(a.x, b.x, (R)fma(-a.y, b.y, c.x)), (R)fma(a.x, b.y, (R)fma(a.y, b.x, c.y))};
}

template <typename T>
struct Scalar;
template <>
struct Scalar<float> {
  using real = float;
  using acc = double;  // energy / dot accumulator
  static constexpr bool complex = false;
};
template <>
struct Scalar<double> {
  using real = double;
  using acc = double;
  static constexpr bool complex = false;
};
template <>
struct Scalar<cf32> {
  using real = float;
  using acc = cf64;
  static constexpr bool complex = true;
};
template <>
struct Scalar<cf64> {
  using real = double;
  using acc = cf64;
  static constexpr bool complex = true;
};

__host__ __device__ inline double abs2(float v) { return (double)v * v; }
__host__ __device__ inline double abs2(double v) { return v * v; }
__host__ __device__ inline double abs2(cf32 v) { return (double)v.x * v.x + (double)v.y * v.y; }
__host__ __device__ inline double abs2(cf64 v) { return v.x * v.x + v.y * v.y; }
__host__ __device__ inline double to_acc(float v) { return v; }
__host__ __device__ inline double to_acc(double v) { return v; }
__host__ __device__ inline cf64 to_acc(cf32 v) { return {v.x, v.y}; }
__host__ __device__ inline cf64 to_acc(cf64 v) { return v; }
template <typename T>
__host__ __device__ inline T from_acc(double v) { return (T)v; }
template <typename T>
__host__ __device__ inline T from_acc(cf64 v) { return T((typename Scalar<T>::real)v.x, (typename Scalar<T>::real)v.y); }
__host__ __device__ inline double acc_abs(double v) { return fabs(v); }
__host__ __device__ inline double acc_abs(cf64 v) { return sqrt(v.x * v.x + v.y * v.y); }
__host__ __device__ inline double acc_arg(double v) { return v < 0.0 ? M_PI : 0.0; }
__host__ __device__ inline double acc_arg(cf64 v) { return atan2(v.y, v.x); }
// scale * exp(i * phase) in the accumulator type of T (phase 0 / pi for real T)
template <typename T>
__host__ __device__ inline typename Scalar<T>::acc polar_acc(double scale, double phase) {
  if constexpr (Scalar<T>::complex)
    return cf64(scale * cos(phase), scale * sin(phase));
  else
    return phase != 0.0 ? -scale : scale;
}
template <typename T>
__host__ __device__ inline T scale_of(double s) {
  return T((typename Scalar<T>::real)s);
}

// A three-level index i = (i1 * n2 + i2) * n3 + i3 mapped to the element offset
// i1 * s1 + i2 * s2 + i3 * s3. Every pairwise einsum step of the reference's
// contraction orders groups at most three tensor axes per GEMM dimension, so
// M, K and N of each step are described by one G2 per operand.
struct G2 {
  int64_t s1, s2, s3;
  int n1, n2, n3;
};

__host__ __device__ inline G2 g2(int n1, int64_t s1, int n2 = 1, int64_t s2 = 0) {
  G2 g;
  g.s1 = s1;
  g.s2 = s2;
  g.s3 = 0;
  g.n1 = n1;
  g.n2 = n2;
  g.n3 = 1;
  return g;
}

__host__ __device__ inline G2 g3(int n1, int64_t s1, int n2, int64_t s2, int n3, int64_t s3) {
  G2 g;
  g.s1 = s1;
  g.s2 = s2;
  g.s3 = s3;
  g.n1 = n1;
  g.n2 = n2;
  g.n3 = n3;
  return g;
}

__device__ inline int64_t g2off(const G2 &g, int i) {
  const int i12 = i / g.n3;
  const int i3 = i - i12 * g.n3;
  const int i1 = i12 / g.n2;
  const int i2 = i12 - i1 * g.n2;
  return (int64_t)i1 * g.s1 + (int64_t)i2 * g.s2 + (int64_t)i3 * g.s3;
}

// Cross-lane moves through DPP (VALU, no LDS round trip). CTRL is a DPP
// control: 0xB1 / 0x4E are quad_perm xor 1 / xor 2, 0x124 / 0x128 rotate a
// row of 16 lanes by 4 / 8, 0x141 / 0x140 mirror each half row / the row.
template <int CTRL>
__device__ inline double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// Reductions over each row of 16 lanes; every lane of the row gets the result.
// The sum is a butterfly (every step pairs two lanes that hold equal partial
// sums), so all 16 lanes hold bitwise-identical results.
__device__ inline double group16_sum(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  v += dpp_f64<0x140>(v);
  return v;
}

__device__ inline double group16_max(double v) {
  v = fmax(v, dpp_f64<0xB1>(v));
  v = fmax(v, dpp_f64<0x4E>(v));
  v = fmax(v, dpp_f64<0x124>(v));
  v = fmax(v, dpp_f64<0x128>(v));
  return v;
}

__device__ inline double group16_min(double v) {
  v = fmin(v, dpp_f64<0xB1>(v));
  v = fmin(v, dpp_f64<0x4E>(v));
  v = fmin(v, dpp_f64<0x124>(v));
  v = fmin(v, dpp_f64<0x128>(v));
  return v;
}

// Max / min over aligned groups of N lanes (N = 2, 4, 8 or 16) of a row.
template <int N>
__device__ inline double group_max(double v) {
  static_assert(N == 2 || N == 4 || N == 8 || N == 16, "group size");
  v = fmax(v, dpp_f64<0xB1>(v));
  if (N >= 4) v = fmax(v, dpp_f64<0x4E>(v));
  if (N >= 8) v = fmax(v, dpp_f64<0x141>(v));
  if (N >= 16) v = fmax(v, dpp_f64<0x140>(v));
  return v;
}

template <int N>
__device__ inline double group_min(double v) {
  static_assert(N == 2 || N == 4 || N == 8 || N == 16, "group size");
  v = fmin(v, dpp_f64<0xB1>(v));
  if (N >= 4) v = fmin(v, dpp_f64<0x4E>(v));
  if (N >= 8) v = fmin(v, dpp_f64<0x141>(v));
  if (N >= 16) v = fmin(v, dpp_f64<0x140>(v));
  return v;
}

__device__ inline double readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wave (all lanes active); the result is uniform.
__device__ inline double wave_sum(double v) {
  v = group16_sum(v);
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

__device__ inline double wave_max(double v) {
  v = group16_max(v);
  return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)), fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}

__device__ inline double wave_min(double v) {
  v = group16_min(v);
  return fmin(fmin(readlane_f64(v, 0), readlane_f64(v, 16)), fmin(readlane_f64(v, 32), readlane_f64(v, 48)));
}

// Block-wide sum of a double over all threads of the workgroup. `red` holds at
// least blockDim.x / 64 doubles. Contains two barriers.
__device__ inline double block_sum(double v, double *red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int w = 0; w < nw; ++w) s += red[w];
  return s;
}

// Advance a 3-level index offset from index i to i + 1.
__device__ inline void g2next(const G2 &g, int &i2, int &i3, int64_t &off) {
  if (++i3 < g.n3) {
    off += g.s3;
    return;
  }
  i3 = 0;
  off -= (int64_t)(g.n3 - 1) * g.s3;
  if (++i2 < g.n2) {
    off += g.s2;
    return;
  }
  i2 = 0;
  off -= (int64_t)(g.n2 - 1) * g.s2;
  off += g.s1;
}

__device__ inline void g2start(const G2 &g, int i, int &i2, int &i3, int64_t &off) {
  const int i12 = i / g.n3;
  i3 = i - i12 * g.n3;
  const int i1 = i12 / g.n2;
  i2 = i12 - i1 * g.n2;
  off = (int64_t)i1 * g.s1 + (int64_t)i2 * g.s2 + (int64_t)i3 * g.s3;
}

constexpr int GEMM_KC = 16;                       // K chunk staged per step
constexpr int GEMM_SMEM = GEMM_KC * (17 + 257);   // elements of the staging union

// One tile shape of gemm2: TY x TX threads (TY*TX == 256), 4 x 4 outputs per
// thread at rows ty + TY*i and columns tx + TX*j of a (4TY) x (4TX) tile.
template <typename T, int TY, int TX>
__device__ inline void gemm2_tiles(const T *__restrict__ A, const G2 &am, const G2 &ak, const T *__restrict__ B,
                                   const G2 &bk, const G2 &bn, T *__restrict__ C, const G2 &cm, const G2 &cn,
                                   T alpha, T *smem) {
  constexpr int BM = 4 * TY, BN = 4 * TX, KC = GEMM_KC;
  constexpr int LA = BM * KC / 256, LB = BN * KC / 256;
  T *sA = smem;                  // [KC][BM + 1]
  T *sB = smem + KC * (BM + 1);  // [KC][BN + 1]
  const int t = threadIdx.x;
  const int ty = t / TX, tx = t - ty * TX;
  const int M = am.n1 * am.n2 * am.n3, N = bn.n1 * bn.n2 * bn.n3, K = ak.n1 * ak.n2 * ak.n3;
  const int a_mm = t % BM, a_k0 = (t / BM) * LA;
  const int b_nn = t % BN, b_k0 = (t / BN) * LB;
  for (int m0 = 0; m0 < M; m0 += BM) {
    for (int n0 = 0; n0 < N; n0 += BN) {
      T acc[4][4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = T(0);
      const bool a_row_ok = m0 + a_mm < M, b_col_ok = n0 + b_nn < N;
      const int64_t a_moff = a_row_ok ? g2off(am, m0 + a_mm) : 0;
      const int64_t b_noff = b_col_ok ? g2off(bn, n0 + b_nn) : 0;
      for (int k0 = 0; k0 < K; k0 += KC) {
        {
          int i2, i3;
          int64_t off;
          const int kb = k0 + a_k0;
          if (kb < K) g2start(ak, kb, i2, i3, off);
#pragma unroll
          for (int l = 0; l < LA; ++l) {
            const int k = kb + l;
            T v = T(0);
            if (a_row_ok && k < K) {
              v = A[a_moff + off];
              if (k + 1 < K) g2next(ak, i2, i3, off);
            }
            sA[(a_k0 + l) * (BM + 1) + a_mm] = v;
          }
        }
        {
          int i2, i3;
          int64_t off;
          const int kb = k0 + b_k0;
          if (kb < K) g2start(bk, kb, i2, i3, off);
#pragma unroll
          for (int l = 0; l < LB; ++l) {
            const int k = kb + l;
            T v = T(0);
            if (b_col_ok && k < K) {
              v = B[b_noff + off];
              if (k + 1 < K) g2next(bk, i2, i3, off);
            }
            sB[(b_k0 + l) * (BN + 1) + b_nn] = v;
          }
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < KC; ++kk) {
          T av[4], bv[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) av[i] = sA[kk * (BM + 1) + ty + TY * i];
#pragma unroll
          for (int j = 0; j < 4; ++j) bv[j] = sB[kk * (BN + 1) + tx + TX * j];
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = fma(av[i], bv[j], acc[i][j]);
        }
        __syncthreads();
      }
      int64_t coff_n[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int n = n0 + tx + TX * j;
        coff_n[j] = n < N ? g2off(cn, n) : 0;
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = m0 + ty + TY * i;
        if (m >= M) continue;
        const int64_t coff_m = g2off(cm, m);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int n = n0 + tx + TX * j;
          if (n < N) C[coff_m + coff_n[j]] = alpha * acc[i][j];
        }
      }
    }
  }
}


// MFMA 16x16x4 tiles for real operands. Lane l feeds A[l % 16][l / 16] and
// B[l / 16][l % 16] of a 16x4 / 4x16 pair; accumulator register r holds
// D[row(l, r)][l % 16] with row = l/16 + 4r for FP64 and 4(l/16) + r for FP32
// (checked on gfx950 against a host product).
typedef double vmc_d4 __attribute__((ext_vector_type(4)));
typedef float vmc_f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
  using V = vmc_d4;
  __device__ static V op(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int l, int r) { return (l >> 4) + 4 * r; }
};
template <>
struct Mfma16<float> {
  using V = vmc_f4;
  __device__ static V op(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int l, int r) { return 4 * (l >> 4) + r; }
};

// gemm2 tile on MFMA: BM x BN with 4 waves, each wave a WR x WC grid of
// 16x16 blocks (64x64: 2x2, 16x256: 1x4, 256x16: 4x1). Operand staging is the
// VALU path's (LDS [KC][BM+1] / [KC][BN+1] through the G2 descriptors).
template <typename T, int BM, int BN, int WR, int WC>
__device__ inline void gemm2_mfma(const T *__restrict__ A, const G2 &am, const G2 &ak, const T *__restrict__ B,
                                  const G2 &bk, const G2 &bn, T *__restrict__ C, const G2 &cm, const G2 &cn,
                                  T alpha, T *smem) {
  using M = Mfma16<T>;
  constexpr int KC = GEMM_KC;
  constexpr int LA = BM * KC / 256, LB = BN * KC / 256;
  constexpr int WAVES_C = BN / (16 * WC);  // waves along N
  T *sA = smem;
  T *sB = smem + KC * (BM + 1);
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wr0 = (wave / WAVES_C) * 16 * WR, wc0 = (wave % WAVES_C) * 16 * WC;
  const int li = lane & 15, lk = lane >> 4;
  const int M_ = am.n1 * am.n2 * am.n3, N_ = bn.n1 * bn.n2 * bn.n3, K_ = ak.n1 * ak.n2 * ak.n3;
  const int a_mm = t % BM, a_k0 = (t / BM) * LA;
  const int b_nn = t % BN, b_k0 = (t / BN) * LB;
  for (int m0 = 0; m0 < M_; m0 += BM) {
    for (int n0 = 0; n0 < N_; n0 += BN) {
      typename M::V acc[WR][WC];
#pragma unroll
      for (int i = 0; i < WR; ++i)
#pragma unroll
        for (int j = 0; j < WC; ++j) acc[i][j] = typename M::V{0, 0, 0, 0};
      const bool a_row_ok = m0 + a_mm < M_, b_col_ok = n0 + b_nn < N_;
      const int64_t a_moff = a_row_ok ? g2off(am, m0 + a_mm) : 0;
      const int64_t b_noff = b_col_ok ? g2off(bn, n0 + b_nn) : 0;
      // 64 x 64 tiles load the next K chunk into registers while the current
      // one is consumed from LDS, so the global latency overlaps the MFMAs;
      // the 16-wide shapes (16 values per thread) stage it directly, which
      // keeps their callers' registers free
      constexpr bool PF = LA + LB <= 8;
      T pa[LA], pb[LB];
      auto fetch = [&](int k0) {
        {
          int i2 = 0, i3 = 0;
          int64_t off = 0;
          const int kb = k0 + a_k0;
          if (kb < K_) g2start(ak, kb, i2, i3, off);
#pragma unroll
          for (int l = 0; l < LA; ++l) {
            const int k = kb + l;
            T v = T(0);
            if (a_row_ok && k < K_) {
              v = A[a_moff + off];
              if (k + 1 < K_) g2next(ak, i2, i3, off);
            }
            pa[l] = v;
          }
        }
        {
          int i2 = 0, i3 = 0;
          int64_t off = 0;
          const int kb = k0 + b_k0;
          if (kb < K_) g2start(bk, kb, i2, i3, off);
#pragma unroll
          for (int l = 0; l < LB; ++l) {
            const int k = kb + l;
            T v = T(0);
            if (b_col_ok && k < K_) {
              v = B[b_noff + off];
              if (k + 1 < K_) g2next(bk, i2, i3, off);
            }
            pb[l] = v;
          }
        }
      };
      if constexpr (PF) fetch(0);
      for (int k0 = 0; k0 < K_; k0 += KC) {
        if constexpr (!PF) fetch(k0);
#pragma unroll
        for (int l = 0; l < LA; ++l) sA[(a_k0 + l) * (BM + 1) + a_mm] = pa[l];
#pragma unroll
        for (int l = 0; l < LB; ++l) sB[(b_k0 + l) * (BN + 1) + b_nn] = pb[l];
        __syncthreads();
        if constexpr (PF)
          if (k0 + KC < K_) fetch(k0 + KC);
#pragma unroll
        for (int ks = 0; ks < KC; ks += 4) {
          T av[WR], bv[WC];
#pragma unroll
          for (int i = 0; i < WR; ++i) av[i] = sA[(ks + lk) * (BM + 1) + wr0 + 16 * i + li];
#pragma unroll
          for (int j = 0; j < WC; ++j) bv[j] = sB[(ks + lk) * (BN + 1) + wc0 + 16 * j + li];
#pragma unroll
          for (int i = 0; i < WR; ++i)
#pragma unroll
            for (int j = 0; j < WC; ++j) acc[i][j] = M::op(av[i], bv[j], acc[i][j]);
        }
        __syncthreads();
      }
#pragma unroll
      for (int j = 0; j < WC; ++j) {
        const int n = n0 + wc0 + 16 * j + li;
        if (n >= N_) continue;
        const int64_t coff_n = g2off(cn, n);
#pragma unroll
        for (int i = 0; i < WR; ++i)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int m = m0 + wr0 + 16 * i + M::row(lane, r);
            if (m < M_) C[g2off(cm, m) + coff_n] = alpha * acc[i][j][r];
          }
      }
    }
  }
}

// Split-K MFMA GEMM for outputs of at most 16 x 16 (the env-gradient
// contraction: 16 x 16 output, K = 1024 at D = 4): the four waves take
// consecutive 16-wide K slices of each staged 64-wide chunk into one 16x16
// accumulator each, and the four partial tiles are summed through LDS at the
// end. A 16 x 256 tile would pad this product 16-fold.
template <typename T>
__device__ inline void gemm2_mfma_splitk(const T *__restrict__ A, const G2 &am, const G2 &ak,
                                         const T *__restrict__ B, const G2 &bk, const G2 &bn, T *__restrict__ C,
                                         const G2 &cm, const G2 &cn, T alpha, T *smem) {
  using M = Mfma16<T>;
  constexpr int KCH = 64, LD = 17;
  T *sA = smem;             // [KCH][LD]
  T *sB = smem + KCH * LD;  // [KCH][LD]
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int M_ = am.n1 * am.n2 * am.n3, N_ = bn.n1 * bn.n2 * bn.n3, K_ = ak.n1 * ak.n2 * ak.n3;
  const int mm = t & 15, kq = (t >> 4) * 4;  // staging: row / column mm, K offsets kq .. kq + 3
  const bool a_ok = mm < M_, b_ok = mm < N_;
  const int64_t a_moff = a_ok ? g2off(am, mm) : 0;
  const int64_t b_noff = b_ok ? g2off(bn, mm) : 0;
  typename M::V acc = typename M::V{0, 0, 0, 0};
  for (int k0 = 0; k0 < K_; k0 += KCH) {
    {
      int i2 = 0, i3 = 0;
      int64_t off = 0;
      const int kb = k0 + kq;
      if (kb < K_) g2start(ak, kb, i2, i3, off);
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        const int k = kb + l;
        T v = T(0);
        if (a_ok && k < K_) {
          v = A[a_moff + off];
          if (k + 1 < K_) g2next(ak, i2, i3, off);
        }
        sA[(kq + l) * LD + mm] = v;
      }
    }
    {
      int i2 = 0, i3 = 0;
      int64_t off = 0;
      const int kb = k0 + kq;
      if (kb < K_) g2start(bk, kb, i2, i3, off);
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        const int k = kb + l;
        T v = T(0);
        if (b_ok && k < K_) {
          v = B[b_noff + off];
          if (k + 1 < K_) g2next(bk, i2, i3, off);
        }
        sB[(kq + l) * LD + mm] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kk = 16 * wave + 4 * q + lk;
      acc = M::op(sA[kk * LD + li], sB[kk * LD + li], acc);
    }
    __syncthreads();
  }
  // partial tiles of the four waves: red[wave][row][col]
  T *red = smem;
#pragma unroll
  for (int r = 0; r < 4; ++r) red[wave * 256 + M::row(lane, r) * 16 + li] = acc[r];
  __syncthreads();
  {
    const int m = t >> 4, n = t & 15;
    if (m < M_ && n < N_)
      C[g2off(cm, m) + g2off(cn, n)] = alpha * ((red[t] + red[256 + t]) + (red[512 + t] + red[768 + t]));
  }
  __syncthreads();
}

__device__ inline cf64 block_sum(cf64 v, double *red) {
  // red needs 2 * (blockDim.x / 64) doubles
  const double re = block_sum(v.x, red);
  const double im = block_sum(v.y, red + 16);
  return {re, im};
}

// C[m, n] = alpha * sum_k A[m, k] B[k, n] for one sample, executed by the
// whole workgroup (blockDim.x == 256). Operands live in global memory
// (per-sample scratch, L2-resident) and are staged through LDS in K chunks;
// the tile shape follows the output shape (64x64, 16x256 or 256x16) unless
// the caller names one (tile: GEMM_TILE_AUTO, _WIDE 16x256, _TALL 256x16,
// _SQUARE 64x64; every tile computes any shape, in the same K order).
//
// FP32 operands run on MFMA (v_mfma_f32_16x16x4f32). FP64 operands run on
// MFMA (v_mfma_f64_16x16x4f64) only where the caller asks for it (F64_MFMA):
// in the register-heavy environment kernels the accumulators push FP64 to one
// wave per SIMD, which costs more than the MFMA saves. Complex operands stay
// on the VALU path.
// GEMM_TILE_SPLITK (outputs of at most 16 x 16, MFMA types only; other types
// take the automatic tile) splits K over the four waves: its K order differs
// from the other tiles'.
enum GemmTile { GEMM_TILE_AUTO = 0, GEMM_TILE_WIDE = 1, GEMM_TILE_TALL = 2, GEMM_TILE_SQUARE = 3, GEMM_TILE_SPLITK = 4 };

template <typename T, bool F64_MFMA = false>
__device__ inline void gemm2(const T *__restrict__ A, G2 am, G2 ak,
                             const T *__restrict__ B, G2 bk, G2 bn,
                             T *__restrict__ C, G2 cm, G2 cn, T alpha = T(1), int tile = GEMM_TILE_AUTO) {
  __shared__ T smem[GEMM_SMEM];
  if (tile == GEMM_TILE_AUTO) {
    const int M = am.n1 * am.n2 * am.n3;
    const int N = bn.n1 * bn.n2 * bn.n3;
    tile = M <= 16 ? GEMM_TILE_WIDE : N <= 16 ? GEMM_TILE_TALL : GEMM_TILE_SQUARE;
  }
  constexpr bool mfma = std::is_same<T, float>::value || (std::is_same<T, double>::value && F64_MFMA);
  if constexpr (!mfma) {
    if (tile == GEMM_TILE_SPLITK) {
      const int M = am.n1 * am.n2 * am.n3;
      const int N = bn.n1 * bn.n2 * bn.n3;
      tile = M <= 16 ? GEMM_TILE_WIDE : N <= 16 ? GEMM_TILE_TALL : GEMM_TILE_SQUARE;
    }
  }
  if constexpr (mfma) {
    if (tile == GEMM_TILE_SPLITK)
      gemm2_mfma_splitk<T>(A, am, ak, B, bk, bn, C, cm, cn, alpha, smem);
    else if (tile == GEMM_TILE_WIDE)
      gemm2_mfma<T, 16, 256, 1, 4>(A, am, ak, B, bk, bn, C, cm, cn, alpha, smem);
    else if (tile == GEMM_TILE_TALL)
      gemm2_mfma<T, 256, 16, 4, 1>(A, am, ak, B, bk, bn, C, cm, cn, alpha, smem);
    else
      gemm2_mfma<T, 64, 64, 2, 2>(A, am, ak, B, bk, bn, C, cm, cn, alpha, smem);
  } else {
    if (tile == GEMM_TILE_WIDE)
      gemm2_tiles<T, 4, 64>(A, am, ak, B, bk, bn, C, cm, cn, alpha, smem);
    else if (tile == GEMM_TILE_TALL)
      gemm2_tiles<T, 64, 4>(A, am, ak, B, bk, bn, C, cm, cn, alpha, smem);
    else
      gemm2_tiles<T, 16, 16>(A, am, ak, B, bk, bn, C, cm, cn, alpha, smem);
  }
}

// Arguments of one zip-up truncation step (see trunc.hip).
template <typename T>
struct TruncArgs {
  const T *mat;      // (m, n) row-major per sample
  int64_t mat_stride;
  T *u;              // (m, k) row-major per sample -> new boundary tensor (left, q, k)
  int64_t u_stride;
  T *carry;          // (k, n) row-major per sample, normalised
  int64_t carry_stride;
  double *logscale;  // per-sample log of the boundary scale, += log|carry|
  int64_t log_stride;
  int m, n, k;
  // Density-matrix cut (strategy.py:153-176): mat holds theta (n x m row-major)
  // and the kernel factors its transpose X = theta^T (m x n). The top-k left
  // vectors of X are V_k, the eigenvectors of theta^H theta, and X^T-side
  // outputs swap: u <- (U_X^H X)^T = theta V_k normalised (k columns, written
  // (n, k) row-major), carry <- U_X^T = V_k^H ((k, m) row-major). Transposes
  // only, no conjugation: U_X = conj(V_k) for complex theta.
  int dm;
  // Generic kernel only (trunc_g.hip): per-sample FP64 scratch of gscr_stride
  // elements (double for real T, complex double for complex T):
  // m*m + 2*m*max(k, TRUNC_G_NB) (trunc_g_need).
  void *gscr;
  int64_t gscr_stride;
  // Register kernel variant (real plans): TRUNC_LPE8 or TRUNC_LPE16 lanes per
  // eigenvalue in the Sturm multisection. Zero-initialised args take LPE8.
  int variant;
};

// Truncation kernel choices of one cut shape: the register kernel with 8 or
// 16 multisection lanes per eigenvalue (launch_trunc), or the generic kernel
// (launch_trunc_g).
enum TruncVariant { TRUNC_LPE8 = 0, TRUNC_LPE16 = 1, TRUNC_GENERIC = 2 };

template <typename T>
void launch_trunc(const TruncArgs<T> &a, int batch, hipStream_t stream);

// Shapes the register-resident kernels take (trunc.hip: real m, n <= 128;
// trunc_c.hip: complex m, n <= 64; k <= 32 for both); the rest go to
// trunc_g.hip (m <= 512, k <= 128), whose Gram matrix lives in global scratch.
inline bool trunc_fast_shape(bool cplx, int m, int n, int k) {
  const int mx = cplx ? 64 : 128;
  return m <= mx && n <= mx && k <= 32;
}
constexpr int TRUNC_G_MAX_M = 512, TRUNC_G_MAX_K = 128;
// Panel width of the generic kernel's blocked tridiagonalisation: the panel's
// reflectors and w vectors (2 * m * TRUNC_G_NB elements) share the scratch of
// the eigenvector phases (2 * m * k), which start after the reduction.
constexpr int TRUNC_G_NB = 8;
// Generic-kernel scratch elements per sample of an m-row cut keeping k.
inline int64_t trunc_g_need(int m, int k) {
  return (int64_t)m * m + 2 * (int64_t)m * (k > TRUNC_G_NB ? k : TRUNC_G_NB);
}

// Zip-up cuts that keep every column (k = n < m) run as a Householder QR in
// registers (trunc.hip: k_trunc_qr, real m <= 128, n <= 32, m * n <= 64 * 32);
// launch_trunc routes them there, and there is nothing to tune.
inline bool trunc_qr_shape(bool cplx, int m, int n, int k, int dm) {
  return !cplx && !dm && k == n && k < m && m <= 128 && n <= 32 && (m <= 64 || n <= 16);
}

// Zip-up cuts with more rows than columns that still discard (k < n < m) run
// through the smaller Gram theta^T theta (trunc.hip: TM_ZIPT, real shapes of
// the register kernel).
inline bool trunc_zipt_shape(bool cplx, int m, int n, int k, int dm) {
  return !cplx && !dm && k < n && n < m && m <= 128 && k <= 32;
}

// Variational QR / LQ panels: register-resident up to VQR_M x VQR_N
// (variational.hpp), in global scratch up to VQG_M x VQG_N.
constexpr int VQR_M = 128, VQR_N = 32;
constexpr int VQG_M = 512, VQG_N = 128;

template <typename T>
void launch_trunc_g(const TruncArgs<T> &a, int batch, hipStream_t stream);

}  // namespace vmc
