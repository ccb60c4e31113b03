// Blockade PEPS on gfx950 (src/vmc/peps/blockade/*): site (r, c) stores a
// (2 * nc, mu_u, mu_d, mu_l, mu_r) tensor, nc = 2^(incoming bonds), and a
// configuration of occupations n selects slice n * nc + cfg with
// cfg = kL * stride + kU for n = 0 and 0 for n = 1, where kL, kU are the
// occupations of the left and upper neighbours and stride = 2 below the first
// row (blockade/model.py:157-212). Amplitudes, environments and gradients are
// the standard engine's over these slices; this file adds:
//   k_bl_slices      slice of every site + the blockade check (compat.py:43-62)
//   k_bl_mask_sites  the charge masks of D0 != D1 applied once per slice
//                    (model.py:201-212; valid samples fix every mask by slice)
//   k_bl_pair        row pair (r, r+1): Metropolis flips of row r over
//                    two-column windows (model.py:820-994), or the one-site
//                    energies of row r (model.py:371-512)
//   k_bl_row         the last row with one-row windows: flips
//                    (model.py:668-817) or one-site energies (model.py:513-599)
// A flip of (r, c) rewires the slices of (r, c + 1) (its kL) and (r + 1, c)
// (its kU); flips to n = 1 next to an occupied site are refused
// (model.py:220-239). Draws: counter (BL_BASE + r * n_cols + c, sweep, chain, 0).
#pragma once

#include "common.hpp"

namespace vmc {

constexpr uint32_t BL_BASE = 4u << 20;
constexpr int BL_MAX_SITES = 1024;  // occupations staged in LDS per workgroup

struct BlDev {
  int R, C;
  int D0, D1, Dmax;
};

__device__ inline int bl_nc(int r, int c) { return 1 << ((r > 0) + (c > 0)); }

__device__ inline int bl_slice(int r, int c, int n, int kl, int ku) {
  return n * bl_nc(r, c) + (n == 0 ? kl * (r > 0 ? 2 : 1) + ku : 0);
}

static __global__ void k_bl_slices(BlDev g, const int32_t *occ, int batch, int32_t *slices, uint8_t *valid) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const int ns = g.R * g.C;
  const int32_t *o = occ + (int64_t)b * ns;
  int32_t *s = slices + (int64_t)b * ns;
  bool ok = true;
  for (int r = 0; r < g.R; ++r)
    for (int c = 0; c < g.C; ++c) {
      const int n = o[r * g.C + c];
      const int kl = c > 0 ? o[r * g.C + c - 1] : 0;
      const int ku = r > 0 ? o[(r - 1) * g.C + c] : 0;
      s[r * g.C + c] = bl_slice(r, c, n, kl, ku);
      ok = ok && !(n && (kl || ku));
    }
  if (valid) valid[b] = ok;
}

// One workgroup per site: out = in with the bonds of every slice masked by
// their sectors, (kU, n, kL, n) for (up, down, left, right); slice (0, cfg)
// has n = 0, kL = cfg / 2, kU = cfg % 2 (kL = cfg on the first row, kU = cfg in
// the first column), slice (1, 0) has kL = kU = 0.
template <typename T>
__global__ void k_bl_mask_sites(BlDev g, const T *in, T *out) {
  const int site = blockIdx.x, r = site / g.C, c = site % g.C;
  const int Dm = g.Dmax;
  auto dims = [&](int rr, int cc, int &u, int &d, int &l, int &rt) {
    u = rr > 0 ? Dm : 1;
    d = rr < g.R - 1 ? Dm : 1;
    l = cc > 0 ? Dm : 1;
    rt = cc < g.C - 1 ? Dm : 1;
  };
  int64_t off = 0;
  for (int s = 0; s < site; ++s) {
    int u, d, l, rt;
    dims(s / g.C, s % g.C, u, d, l, rt);
    off += (int64_t)2 * bl_nc(s / g.C, s % g.C) * u * d * l * rt;
  }
  int u, d, l, rt;
  dims(r, c, u, d, l, rt);
  const int nc = bl_nc(r, c);
  const int per = u * d * l * rt;
  for (int e = threadIdx.x; e < 2 * nc * per; e += blockDim.x) {
    const int sl = e / per, rem = e - sl * per;
    const int n = sl / nc, cfg = sl - n * nc;
    int kl = 0, ku = 0;
    if (n == 0) {
      if (c == 0)
        ku = cfg;
      else if (r == 0)
        kl = cfg;
      else
        kl = cfg >> 1, ku = cfg & 1;
    }
    const int ir = rem % rt, il = (rem / rt) % l, id = (rem / (rt * l)) % d, iu = rem / (rt * l * d);
    const int dn = n ? g.D1 : g.D0, dl = kl ? g.D1 : g.D0, du = ku ? g.D1 : g.D0;
    const bool keep = iu < du && id < dn && il < dl && ir < dn;
    out[off + e] = keep ? in[off + e] : T(0);
  }
}

template <typename T>
struct BlPairArgs {
  int C, row;
  BlDev g;
  const T *top;
  int64_t top_stride;  // boundary before row r
  const T *bot;
  int64_t bot_stride;  // boundary below row r+1
  const T *sites;
  int32_t *occ;     // (B, R*C), updated by the transition
  int32_t *slices;  // (B, R*C), updated with the occupations
  const ColGeo *geo0, *geo1;
  T *r2;
  int64_t r2_stride, cap2;
  T *l2, *lx, *ly, *lxp, *lyp;  // cap2 each
  int64_t l_stride;
  T *v1, *v2, *v3;
  int64_t v_stride;
  int transition;
  uint64_t seed;
  int sweep;
  // energy
  TermsDev terms;
  const double *top_log;
  int64_t top_log_stride;
  const double *bot_log;
  int64_t bot_log_stride;
  const double *log_amp;
  double *eloc;
};

// out = in advanced through column c of the row pair with slices (s0, s1),
// normalised; returns the log of the removed norm.
template <typename T>
__device__ inline double bl_advance(const BlPairArgs<T> &a, const T *top, const T *bot, int c, int s0, int s1,
                                    const T *in, T *out, T *v1, T *v2, T *v3, double *red) {
  const ColGeo G0 = a.geo0[c], G1 = a.geo1[c];
  const T *w0 = a.sites + G0.site_off + (int64_t)s0 * G0.site_slice;
  const T *w1 = a.sites + G1.site_off + (int64_t)s1 * G1.site_slice;
  left2_v1<T>(G0, G1, in, top + G0.top_off, v1);
  __syncthreads();
  left2_from_v1<T>(G0, G1, v1, w0, w1, bot + G1.bot_off, v2, v3, out);
  double lg = 0.0;
  normalise(out, (int64_t)G0.tr * G0.sr * G1.sr * G1.br, red, &lg);
  return lg;
}

template <typename T>
__device__ inline void bl_copy(const T *x, T *y, int64_t n) {
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) y[i] = x[i];
  __syncthreads();
}

// Metropolis ratio |psi'|^2 / |psi|^2 from (dot, log scale) pairs with the
// common scales removed; contraction.py:88-94's rule for zero amplitudes.
__device__ inline double bl_ratio(double pc, double xc, double pp, double xp) {
  if (pc > 0.0 && pp > 0.0 && isfinite(xc) && isfinite(xp)) return exp(2.0 * (log(pp) - log(pc) + xp - xc));
  if (pc > 0.0 && isfinite(xc)) return 0.0;
  return (pp > 0.0 && isfinite(xp)) ? INFINITY : 0.0;
}

__device__ inline double bl_uniform(uint64_t seed, int r, int c, int C, int sweep, int chain) {
  uint32_t ctr[4] = {BL_BASE + (uint32_t)(r * C + c), (uint32_t)sweep, (uint32_t)chain, 0u};
  philox4x32_10(ctr, (uint32_t)(seed & 0xFFFFFFFFu), (uint32_t)(seed >> 32));
  const uint64_t u53 = ((uint64_t)ctr[0] << 21) | (ctr[1] >> 11);
  return (double)u53 * 0x1.0p-53;
}

// Proposal of a flip of (r, c): the flipped occupation, whether it is allowed,
// and the slices of (r, c), (r, c+1) and (r+1, c) after it (model.py:870-897).
struct BlFlip {
  int n_cur, n_flip, can;
  int s_rc, s_rc1, s_r1c;
};

__device__ inline BlFlip bl_flip(const BlDev &g, const int32_t *o, int r, int c) {
  const int C = g.C;
  BlFlip f;
  f.n_cur = o[r * C + c];
  f.n_flip = 1 - f.n_cur;
  bool blocked = false;
  if (c > 0) blocked |= o[r * C + c - 1] == 1;
  if (c < C - 1) blocked |= o[r * C + c + 1] == 1;
  if (r > 0) blocked |= o[(r - 1) * C + c] == 1;
  if (r < g.R - 1) blocked |= o[(r + 1) * C + c] == 1;
  f.can = f.n_flip == 0 || !blocked;
  const int kl = c > 0 ? o[r * C + c - 1] : 0, ku = r > 0 ? o[(r - 1) * C + c] : 0;
  f.s_rc = bl_slice(r, c, f.n_flip, kl, ku);
  f.s_rc1 = c + 1 < C ? bl_slice(r, c + 1, o[r * C + c + 1], f.n_flip, r > 0 ? o[(r - 1) * C + c + 1] : 0) : 0;
  f.s_r1c = r + 1 < g.R ? bl_slice(r + 1, c, o[(r + 1) * C + c], c > 0 ? o[(r + 1) * C + c - 1] : 0, f.n_flip) : 0;
  return f;
}

template <typename T>
__global__ __launch_bounds__(256) void k_bl_pair(BlPairArgs<T> a) {
  using Acc = typename Scalar<T>::acc;
  __shared__ double red[32];
  __shared__ double r2log[64];
  __shared__ int32_t so[BL_MAX_SITES];  // occupations
  __shared__ int32_t ssl[128];          // slices of rows r and r+1
  __shared__ BlFlip fl;
  __shared__ double draw_u;
  const int b = blockIdx.x;
  const int C = a.C, r = a.row;
  const BlDev &g = a.g;
  const int ns = g.R * C;
  const T *top = a.top + (int64_t)b * a.top_stride;
  const T *bot = a.bot + (int64_t)b * a.bot_stride;
  T *r2 = a.r2 + (int64_t)b * a.r2_stride;
  T *l2 = a.l2 + (int64_t)b * a.l_stride, *lx = a.lx + (int64_t)b * a.l_stride;
  T *ly = a.ly + (int64_t)b * a.l_stride, *lxp = a.lxp + (int64_t)b * a.l_stride;
  T *lyp = a.lyp + (int64_t)b * a.l_stride;
  T *v1 = a.v1 + (int64_t)b * a.v_stride, *v2 = a.v2 + (int64_t)b * a.v_stride, *v3 = a.v3 + (int64_t)b * a.v_stride;
  int32_t *occ = a.occ + (int64_t)b * ns;
  int32_t *sl = a.slices + (int64_t)b * ns;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) so[i] = occ[i];
  for (int i = threadIdx.x; i < 2 * C; i += blockDim.x) ssl[i] = sl[r * C + i];
  __syncthreads();
  int32_t *s0 = ssl, *s1 = ssl + C;

  right2_envs<T>(C, top, bot, a.sites, a.geo0, a.geo1, s0, s1, r2, a.cap2, v1, v2, v3, r2log, red);
  if (threadIdx.x == 0) l2[0] = T(1);
  __syncthreads();
  double llog = 0.0;
  Acc e_row = Acc(0.0);
  double base = 0.0, phase = 0.0;
  if (!a.transition) {
    base = a.top_log[(int64_t)b * a.top_log_stride] + a.bot_log[(int64_t)b * a.bot_log_stride] -
           a.log_amp[2 * (int64_t)b];
    phase = -a.log_amp[2 * (int64_t)b + 1];
  }
  for (int c = 0; c < C; ++c) {
    const int site = r * C + c;
    const int oi = a.transition ? -1 : a.terms.one_idx[site];
    const ColGeo G0 = a.geo0[c], G1 = a.geo1[c];
    const int64_t nl = (int64_t)G0.tr * G0.sr * G1.sr * G1.br;  // L2 after column c
    if (a.transition || oi >= 0) {
      if (threadIdx.x == 0) {
        fl = bl_flip(g, so, r, c);
        if (a.transition) draw_u = bl_uniform(a.seed, r, c, C, a.sweep, b);
      }
      __syncthreads();
      const BlFlip f = fl;
      const bool two = c + 1 < C;
      const T *Rn = r2 + (int64_t)(two ? c + 1 : c) * a.cap2;
      const int64_t nr = two ? (int64_t)a.geo0[c + 1].tr * a.geo0[c + 1].sr * a.geo1[c + 1].sr * a.geo1[c + 1].br
                             : nl;
      double xl = 0.0, xlp = 0.0;
      Acc dprop = Acc(0.0);
      if (f.can) {
        xlp = bl_advance(a, top, bot, c, f.s_rc, f.s_r1c, l2, lxp, v1, v2, v3, red);
        double ylp = 0.0;
        if (two) ylp = bl_advance(a, top, bot, c + 1, f.s_rc1, s1[c + 1], lxp, lyp, v1, v2, v3, red);
        dprop = block_dot(two ? lyp : lxp, Rn, nr, red);
        if (a.transition) {
          xl = bl_advance(a, top, bot, c, s0[c], s1[c], l2, lx, v1, v2, v3, red);
          double yl = 0.0;
          if (two) yl = bl_advance(a, top, bot, c + 1, s0[c + 1], s1[c + 1], lx, ly, v1, v2, v3, red);
          const Acc dcur = block_dot(two ? ly : lx, Rn, nr, red);
          const bool accept = draw_u < fmin(1.0, bl_ratio(acc_abs(dcur), xl + yl, acc_abs(dprop), xlp + ylp));
          if (accept) {
            if (threadIdx.x == 0) {
              so[site] = f.n_flip;
              s0[c] = f.s_rc;
              s1[c] = f.s_r1c;
              if (two) s0[c + 1] = f.s_rc1;
            }
            bl_copy(lxp, l2, nl);
          } else {
            bl_copy(lx, l2, nl);
          }
          __syncthreads();
          continue;
        }
        // sum_ops op[n_flip, n_cur] * amp_flip / amp
        const double sc = exp(base + llog + xlp + ylp + r2log[two ? c + 1 : c]);
        e_row += op_at<Acc>(a.terms, oi + f.n_flip * 2 + f.n_cur) * (dprop * polar_acc<T>(sc, phase));
      } else if (a.transition) {
        // a refused flip has ratio 0: the draw is spent, the state kept
        xl = bl_advance(a, top, bot, c, s0[c], s1[c], l2, l2, v1, v2, v3, red);
        continue;
      }
    }
    llog += bl_advance(a, top, bot, c, s0[c], s1[c], l2, l2, v1, v2, v3, red);
  }
  if (a.transition) {
    for (int i = threadIdx.x; i < C; i += blockDim.x) occ[r * C + i] = so[r * C + i];
    for (int i = threadIdx.x; i < 2 * C; i += blockDim.x) sl[r * C + i] = ssl[i];
  } else if (threadIdx.x == 0) {
    eloc_add(a.eloc, b, e_row);
  }
}

// The last row (one-row windows): right environments, then per site the
// amplitude L . [c] . [c+1] . R[c+1] (or L . [c] . R[c] in the last column)
// with the current and the flipped slices, and the left environment through
// the chosen slice of column c. a.cfg holds the slices and is updated.
template <typename T>
__global__ __launch_bounds__(256) void k_bl_row(RowArgs<T> a, BlDev g, int32_t *occ_all, int transition) {
  using Acc = typename Scalar<T>::acc;
  __shared__ double red[32];
  __shared__ double rlog[64];
  __shared__ int32_t so[BL_MAX_SITES];
  __shared__ BlFlip fl;
  __shared__ double draw_u;
  const int b = blockIdx.x;
  const int C = a.C, r = a.row;
  const int ns = g.R * C;
  const T *top = a.top + (int64_t)b * a.top_stride;
  const T *bot = a.bot + (int64_t)b * a.bot_stride;
  T *renv = a.renv + (int64_t)b * a.renv_stride;
  T *lenv = a.lenv + (int64_t)b * a.lenv_stride;
  T *s1 = a.s1 + (int64_t)b * a.s_stride;
  T *s2 = a.s2 + (int64_t)b * a.s_stride;
  T *s3 = a.s3 + (int64_t)b * a.s_stride;
  T *s4 = a.s4 + (int64_t)b * a.s_stride;
  T *s5 = a.s5 + (int64_t)b * a.s_stride;
  int32_t *cfg = a.cfg + (int64_t)b * a.n_sites + (int64_t)r * C;
  int32_t *occ = occ_all + (int64_t)b * ns;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) so[i] = occ[i];
  __syncthreads();

  right_envs(a, top, bot, renv, s1, s2, cfg, rlog, red);
  if (threadIdx.x == 0) lenv[0] = T(1);
  __syncthreads();
  double llog = 0.0;
  Acc e_row = Acc(0.0);
  double base = 0.0, phase = 0.0;
  if (!transition) {
    base = a.top_log[(int64_t)b * a.top_log_stride] + a.bot_log[(int64_t)b * a.bot_log_stride] -
           a.log_amp[2 * (int64_t)b];
    phase = -a.log_amp[2 * (int64_t)b + 1];
  }
  for (int c = 0; c < C; ++c) {
    const int site = r * C + c;
    const int oi = transition ? -1 : a.terms.one_idx[site];
    const ColGeo G0 = a.geo[c];
    const T *tp0 = top + G0.top_off, *bp0 = bot + G0.bot_off;
    const T *A0 = a.sites + G0.site_off;
    const int64_t n0 = (int64_t)G0.tr * G0.sr * G0.br;
    // Y1 = L . top[c], shared by both slices of column c; s3 = L through the current slice
    gemm2<T>(lenv, g2(G0.sl, G0.bl, G0.bl, 1), g2(G0.tl, (int64_t)G0.sl * G0.bl), tp0, g2(G0.tl, (int64_t)G0.tp * G0.tr),
             g2(G0.tp * G0.tr, 1), s1, g2(G0.sl * G0.bl, (int64_t)G0.tp * G0.tr), g2(G0.tp * G0.tr, 1));
    __syncthreads();
    double x0c = 0.0;
    left_update(G0, A0 + (int64_t)cfg[c] * G0.site_slice, s1, bp0, s2, s3);
    normalise(s3, n0, red, &x0c);
    if (transition || oi >= 0) {
      if (threadIdx.x == 0) {
        fl = bl_flip(g, so, r, c);
        if (transition) draw_u = bl_uniform(a.seed, r, c, C, a.sweep, b);
      }
      __syncthreads();
      const BlFlip f = fl;
      const bool two = c + 1 < C;
      Acc dcur = Acc(0.0), dprop = Acc(0.0);
      double x0p = 0.0, x1c = 0.0, x1p = 0.0;
      if (f.can) {
        left_update(G0, A0 + (int64_t)f.s_rc * G0.site_slice, s1, bp0, s2, s4);
        normalise(s4, n0, red, &x0p);
      }
      if (two) {
        const ColGeo G1 = a.geo[c + 1];
        const T *tp1 = top + G1.top_off, *bp1 = bot + G1.bot_off;
        const T *A1 = a.sites + G1.site_off;
        const int64_t n1 = (int64_t)G1.tr * G1.sr * G1.br;
        const T *Rn = renv + (int64_t)(c + 1) * a.env_cap;
        if (transition) {
          gemm2<T>(s3, g2(G1.sl, G1.bl, G1.bl, 1), g2(G1.tl, (int64_t)G1.sl * G1.bl), tp1,
                   g2(G1.tl, (int64_t)G1.tp * G1.tr), g2(G1.tp * G1.tr, 1), s1, g2(G1.sl * G1.bl, (int64_t)G1.tp * G1.tr),
                   g2(G1.tp * G1.tr, 1));
          __syncthreads();
          left_update(G1, A1 + (int64_t)cfg[c + 1] * G1.site_slice, s1, bp1, s2, s5);
          normalise(s5, n1, red, &x1c);
          dcur = block_dot(s5, Rn, n1, red);
        }
        if (f.can) {
          gemm2<T>(s4, g2(G1.sl, G1.bl, G1.bl, 1), g2(G1.tl, (int64_t)G1.sl * G1.bl), tp1,
                   g2(G1.tl, (int64_t)G1.tp * G1.tr), g2(G1.tp * G1.tr, 1), s1, g2(G1.sl * G1.bl, (int64_t)G1.tp * G1.tr),
                   g2(G1.tp * G1.tr, 1));
          __syncthreads();
          left_update(G1, A1 + (int64_t)f.s_rc1 * G1.site_slice, s1, bp1, s2, s5);
          normalise(s5, n1, red, &x1p);
          dprop = block_dot(s5, Rn, n1, red);
        }
      } else {
        const T *Rc = renv + (int64_t)c * a.env_cap;
        if (transition) dcur = block_dot(s3, Rc, n0, red);
        if (f.can) dprop = block_dot(s4, Rc, n0, red);
      }
      if (transition) {
        const bool accept = f.can && draw_u < fmin(1.0, bl_ratio(acc_abs(dcur), x0c + x1c, acc_abs(dprop), x0p + x1p));
        if (accept) {
          if (threadIdx.x == 0) {
            so[site] = f.n_flip;
            cfg[c] = f.s_rc;
            if (two) cfg[c + 1] = f.s_rc1;
          }
          bl_copy(s4, lenv, n0);
        } else {
          bl_copy(s3, lenv, n0);
        }
        __syncthreads();
        continue;
      }
      if (f.can) {
        const double sc = exp(base + llog + x0p + x1p + rlog[two ? c + 1 : c]);
        e_row += op_at<Acc>(a.terms, oi + f.n_flip * 2 + f.n_cur) * (dprop * polar_acc<T>(sc, phase));
      }
    }
    bl_copy(s3, lenv, n0);
    llog += x0c;
  }
  if (transition) {
    for (int i = threadIdx.x; i < C; i += blockDim.x) occ[r * C + i] = so[r * C + i];
  } else if (threadIdx.x == 0) {
    eloc_add(a.eloc, b, e_row);
  }
}

}  // namespace vmc
