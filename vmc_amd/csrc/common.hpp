// Shared device helpers of the gfx950 PEPS engine.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

namespace vmc {

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

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T>
__device__ inline T group16_sum(T v) {
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) v += __shfl_xor(v, off, 16);
  return v;
}

__device__ inline double group16_max(double v) {
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 16));
  return v;
}

__device__ inline double group16_min(double v) {
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) v = fmin(v, __shfl_xor(v, off, 16));
  return v;
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

// C[m, n] = alpha * sum_k A[m, k] B[k, n] for one sample, executed by the whole
// workgroup. Operands live in global memory (per-sample scratch, L2-resident).
// Outputs are assigned with n fastest so consecutive lanes write consecutive
// columns.
template <typename T>
__device__ inline void gemm2(const T *__restrict__ A, G2 am, G2 ak,
                             const T *__restrict__ B, G2 bk, G2 bn,
                             T *__restrict__ C, G2 cm, G2 cn, T alpha = T(1)) {
  const int M = am.n1 * am.n2 * am.n3;
  const int N = bn.n1 * bn.n2 * bn.n3;
  for (int idx = threadIdx.x; idx < M * N; idx += blockDim.x) {
    const int mi = idx / N;
    const int ni = idx - mi * N;
    const T *a = A + g2off(am, mi);
    const T *b = B + g2off(bn, ni);
    T acc = T(0);
    for (int k1 = 0; k1 < ak.n1; ++k1) {
      const T *a1 = a + (int64_t)k1 * ak.s1;
      const T *b1 = b + (int64_t)k1 * bk.s1;
      for (int k2 = 0; k2 < ak.n2; ++k2) {
        const T *a2 = a1 + (int64_t)k2 * ak.s2;
        const T *b2 = b1 + (int64_t)k2 * bk.s2;
        for (int k3 = 0; k3 < ak.n3; ++k3) acc = fma(a2[(int64_t)k3 * ak.s3], b2[(int64_t)k3 * bk.s3], acc);
      }
    }
    C[g2off(cm, mi) + g2off(cn, ni)] = alpha * acc;
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
};

template <typename T>
void launch_trunc(const TruncArgs<T> &a, int batch, hipStream_t stream);

}  // namespace vmc
