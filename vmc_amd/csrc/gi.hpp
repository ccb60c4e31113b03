// Gauge-invariant PEPS on gfx950 (src/vmc/peps/gi/*): a site's effective
// tensor is the slice sigma * nc + cfg of its stored (d * nc, mu_u, mu_d,
// mu_l, mu_r) tensor, where cfg encodes the independent link charges around
// the site (gi/model.py:308-343). Amplitudes, environments and gradients are
// therefore the standard engine's, indexed by these combined slices; this
// file adds the link-specific pieces:
//   k_gi_slices        combined slice index per site + Gauss-law validity
//                      (gi/compat.py:79-88)
//   k_gi_mask          zero amplitude / log-derivatives of invalid samples
//                      (gi/compat.py:90-99)
//   k_gi_plaquette     one row pair (r, r+1): Metropolis plaquette sweep
//                      (gi/model.py:623-700), vertical link sweep with
//                      matter (gi/model.py:795-878) or plaquette energies
//                      (gi/model.py:546-608)
//   k_gi_hlink         one row: horizontal link sweep with matter
//                      (gi/model.py:703-792)
//   k_gi_link_energy   link-diagonal terms (gi/local_terms.py:13-36)
//   k_gi_site_diag     diagonal terms on matter states (MatterMassTerm,
//                      gi/local_terms.py:39-52)
// Sample layout per chain: sites (R*C), horizontal links (R*(C-1)) then
// vertical links ((R-1)*C), int32 (gi/model.py:121-141).
#pragma once

#include "common.hpp"

namespace vmc {

constexpr int GI_MAX_LINKS = 1024;
constexpr uint32_t GI_PLAQ_BASE = 1u << 20;   // Philox site word of plaquette (r, c): base + r*C + c
constexpr uint32_t GI_HLINK_BASE = 2u << 20;  // horizontal link (r, c): base + r*(C-1) + c
constexpr uint32_t GI_VLINK_BASE = 3u << 20;  // vertical link (r, c): base + r*C + c

struct GiDev {
  int R, C, N, Qx, d;
  const int32_t *charge;  // charge of each matter state (d entries)
  const int32_t *ctab;    // (N, maxdeg) matter states of each charge, ascending, -1 padded
  const int32_t *cdeg;    // (N) number of matter states of each charge
  int maxdeg;
};

__device__ inline int gi_mod(int x, int n) { return ((x % n) + n) % n; }

// One link move with matter (gi/model.py:740-760, 812-832): link += delta,
// the first site's charge shifts by s1 * delta and the second's by -s1 * delta,
// and each site takes a uniformly drawn matter state of its new charge
// (gi/model.py:209-217). Draws: counter word 0 = ctr0, block 0 gives the
// accept uniform and delta = 1 + word % (N - 1), block 1 the two state picks.
struct GiLinkMove {
  double u, ratio_prop;  // accept uniform, proposal ratio deg'/deg
  int delta, sa, sb;     // link shift, new matter states of the two sites
};

__device__ inline double gi_u53(uint32_t hi, uint32_t lo) {
  return (double)((((uint64_t)hi) << 21) | (lo >> 11)) * 0x1.0p-53;
}

__device__ inline GiLinkMove gi_link_move(const GiDev &g, uint64_t seed, uint32_t ctr0, int sweep, int chain, int ma,
                                          int mb, int sign) {
  uint32_t x[4] = {ctr0, (uint32_t)sweep, (uint32_t)chain, 0u};
  uint32_t y[4] = {ctr0, (uint32_t)sweep, (uint32_t)chain, 1u};
  philox4x32_10(x, (uint32_t)(seed & 0xFFFFFFFFu), (uint32_t)(seed >> 32));
  philox4x32_10(y, (uint32_t)(seed & 0xFFFFFFFFu), (uint32_t)(seed >> 32));
  GiLinkMove mv;
  mv.u = gi_u53(x[0], x[1]);
  mv.delta = 1 + (int)(x[2] % (uint32_t)max(g.N - 1, 1));
  const int qa = g.charge[ma], qb = g.charge[mb];
  const int qa2 = gi_mod(qa + sign * mv.delta, g.N), qb2 = gi_mod(qb - sign * mv.delta, g.N);
  const int na = g.cdeg[qa2], nb = g.cdeg[qb2];
  mv.sa = g.ctab[qa2 * g.maxdeg + (int)floor(gi_u53(y[0], y[1]) * na)];
  mv.sb = g.ctab[qb2 * g.maxdeg + (int)floor(gi_u53(y[2], y[3]) * nb)];
  mv.ratio_prop = (double)(na * nb) / (double)(g.cdeg[qa] * g.cdeg[qb]);
  return mv;
}

// Metropolis acceptance from two amplitudes given as (dot, log scale), with
// contraction.py:88-94's rule for zero amplitudes, times the proposal ratio.
__device__ inline double gi_ratio(double pc, double xc, double pp, double xp, double prop) {
  double ratio;
  if (pc > 0.0 && pp > 0.0 && isfinite(xc) && isfinite(xp))
    ratio = exp(2.0 * (log(pp) - log(pc) + xp - xc));
  else if (pc > 0.0 && isfinite(xc))
    ratio = 0.0;
  else
    ratio = (pp > 0.0 && isfinite(xp)) ? INFINITY : 0.0;
  return ratio * prop;
}

__device__ inline int gi_nc(const GiDev &g, int r, int c) {
  const int nl = (r > 0) + (r < g.R - 1) + (c > 0) + (c < g.C - 1);
  int v = 1;
  for (int i = 1; i < nl; ++i) v *= g.N;
  return v;
}

// Config index of site (r, c) from the link array lk (horizontal then
// vertical): the first active link in (right, down, up, left) is fixed by
// Gauss law, the others are base-N digits in (left, up, down, right) order.
__device__ inline int gi_cfg(const GiDev &g, const int32_t *lk, int r, int c) {
  const int C = g.C, nh = g.R * (C - 1);
  const bool al = c > 0, ar = c < C - 1, au = r > 0, ad = r < g.R - 1;
  const int dep = ar ? 0 : ad ? 1 : au ? 2 : al ? 3 : -1;
  int idx = 0;
  if (al && dep != 3) idx = idx * g.N + lk[r * (C - 1) + c - 1];
  if (au && dep != 2) idx = idx * g.N + lk[nh + (r - 1) * C + c];
  if (ad && dep != 1) idx = idx * g.N + lk[nh + r * C + c];
  if (ar && dep != 0) idx = idx * g.N + lk[r * (C - 1) + c];
  return idx;
}

static __global__ void k_gi_slices(GiDev g, const int32_t *samples, int batch, int32_t *slices, uint8_t *valid) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const int R = g.R, C = g.C, ns = R * C, nh = R * (C - 1);
  const int n_sample = ns + nh + (R - 1) * C;
  const int32_t *s = samples + (int64_t)b * n_sample;
  const int32_t *lk = s + ns;
  bool ok = true;
  // values outside [0, d) / [0, N) make the sample invalid (slice 0 keeps
  // every tensor read in bounds; the amplitude is then masked to zero)
  bool in_range = true;
  for (int i = 0; i < ns; ++i) in_range = in_range && s[i] >= 0 && s[i] < g.d;
  for (int i = 0; i < n_sample - ns; ++i) in_range = in_range && lk[i] >= 0 && lk[i] < g.N;
  if (!in_range) {
    for (int i = 0; i < ns; ++i) slices[(int64_t)b * ns + i] = 0;
    if (valid) valid[b] = 0;
    return;
  }
  for (int r = 0; r < R; ++r)
    for (int c = 0; c < C; ++c) {
      const int site = s[r * C + c];
      slices[(int64_t)b * ns + r * C + c] = site * gi_nc(g, r, c) + gi_cfg(g, lk, r, c);
      const int nl = c > 0 ? lk[r * (C - 1) + c - 1] : 0;
      const int nr = c < C - 1 ? lk[r * (C - 1) + c] : 0;
      const int nu = r > 0 ? lk[nh + (r - 1) * C + c] : 0;
      const int nd = r < R - 1 ? lk[nh + r * C + c] : 0;
      const int div = gi_mod(nl + nd - nu - nr, g.N);
      ok = ok && gi_mod(div + g.charge[site], g.N) == gi_mod(g.Qx, g.N);
    }
  if (valid) valid[b] = ok ? 1 : 0;
}

template <typename T>
__global__ void k_gi_mask(const uint8_t *valid, int batch, double *log_amp, T *o, int64_t n_o) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)batch * n_o) return;
  const int64_t b = i / n_o;
  if (valid[b]) return;
  if (o) o[i] = T(0);
  if (i - b * n_o == 0) {
    log_amp[2 * b] = -INFINITY;
    log_amp[2 * b + 1] = 0.0;
  }
}

// Charge masks of a non-uniform degeneracy_per_charge (gi/model.py:52-60,
// 292-305), applied once per slice: one workgroup per site writes its
// (d * nc, u, dn, l, r) tensor with every slice (sigma, cfg) masked by the
// sectors of its own links. The links of a slice are the base-N digits of
// cfg in (left, up, down, right) order, skipping boundary links and the
// dependent one (the first of right, down, up, left present), which Gauss law
// fixes: k_l + k_d - k_u - k_r + charge(sigma) = Qx (mod N). On Gauss-valid
// samples these are the links the per-sample masks use.
template <typename T>
__global__ void k_gi_mask_sites(GiDev g, int Dm, const int32_t *deg, const T *in, T *out) {
  const int site = blockIdx.x, r = site / g.C, c = site % g.C;
  const int R = g.R, C = g.C, N = g.N;
  int64_t off = 0;
  for (int s = 0; s < site; ++s) {
    const int rr = s / C, cc = s % C;
    off += (int64_t)g.d * gi_nc(g, rr, cc) * (rr > 0 ? Dm : 1) * (rr < R - 1 ? Dm : 1) * (cc > 0 ? Dm : 1) *
           (cc < C - 1 ? Dm : 1);
  }
  const bool al = c > 0, ar = c < C - 1, au = r > 0, ad = r < R - 1;
  const int u = au ? Dm : 1, dn = ad ? Dm : 1, l = al ? Dm : 1, rt = ar ? Dm : 1;
  const int dep = ar ? 3 : ad ? 2 : au ? 1 : al ? 0 : -1;  // 0 left, 1 up, 2 down, 3 right
  const int nc = gi_nc(g, r, c), per = u * dn * l * rt;
  const int qx = gi_mod(g.Qx, N);
  for (int e = threadIdx.x; e < g.d * nc * per; e += blockDim.x) {
    const int sl = e / per, rem = e - sl * per;
    const int sigma = sl / nc;
    int rest = sl - sigma * nc;
    int k[4] = {0, 0, 0, 0};
    const bool act[4] = {al, au, ad, ar};
    for (int i = 3; i >= 0; --i)
      if (act[i] && i != dep) {
        k[i] = rest % N;
        rest /= N;
      }
    const int q = g.charge[sigma];
    if (dep == 3) k[3] = gi_mod(k[0] + k[2] - k[1] + q - qx, N);
    if (dep == 2) k[2] = gi_mod(qx - q - k[0] + k[1] + k[3], N);
    if (dep == 1) k[1] = gi_mod(k[0] + k[2] - k[3] + q - qx, N);
    if (dep == 0) k[0] = gi_mod(qx - q - k[2] + k[1] + k[3], N);
    const int ir = rem % rt, il = (rem / rt) % l, id = (rem / (rt * l)) % dn, iu = rem / (rt * l * dn);
    const bool keep = iu < deg[k[1]] && id < deg[k[2]] && il < deg[k[0]] && ir < deg[k[3]];
    out[off + e] = keep ? in[off + e] : T(0);
  }
}

static __global__ void k_gi_link_energy(GiDev g, const int32_t *samples, int batch, const double *ltab, double *eloc,
                                 int cplx_out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const int R = g.R, C = g.C, ns = R * C, nl = R * (C - 1) + (R - 1) * C;
  const int32_t *lk = samples + (int64_t)b * (ns + nl) + ns;
  double e = 0.0;
  for (int l = 0; l < nl; ++l) e += ltab[(int64_t)l * g.N + lk[l]];
  if (cplx_out) {
    eloc[2 * b] = e;
    eloc[2 * b + 1] = 0.0;
  } else {
    eloc[b] = e;
  }
}

// Diagonal terms on matter states (gi/model.py:422-431: diag[sum_k sigma_k d^k]
// over the term's sites, MatterMassTerm among them), added to eloc; the matter
// states are the first R*C entries of each sample.
static __global__ void k_gi_site_diag(TermsDev t, const int32_t *samples, int n_sample, int d, double *eloc, int batch,
                               int cplx_out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const int32_t *s = samples + (int64_t)b * n_sample;
  double e = 0.0, ei = 0.0;
  for (int i = 0; i < t.n_diag; ++i) {
    const int a0 = t.diag_sites[2 * i], a1 = t.diag_sites[2 * i + 1];
    int idx = s[a0];
    if (a1 >= 0) idx = idx * d + s[a1];
    e += t.ops[t.diag_off[i] + idx];
    ei += t.ops_im[t.diag_off[i] + idx];
  }
  if (cplx_out) {
    eloc[2 * b] += e;
    eloc[2 * b + 1] += ei;
  } else {
    eloc[b] += e;
  }
}

template <typename T>
struct GiPlaqArgs {
  int C, row;
  GiDev g;
  const T *top;
  int64_t top_stride;  // boundary before row r (top_off of row r in geo0)
  const T *bot;
  int64_t bot_stride;  // boundary below row r+1 (bot_off of row r+1 in geo1)
  const T *sites;
  int32_t *samples;    // (B, n_sample), links updated by the transition
  int32_t *slices;     // (B, R*C), updated with the links
  const ColGeo *geo0, *geo1;
  T *r2;
  int64_t r2_stride, cap2;
  T *l2, *lx, *ly, *lxp, *lyp;  // cap2 each
  int64_t l_stride;
  T *v1, *v2, *v3;
  int64_t v_stride;
  // transition: plaquette moves, or (vlink) vertical link moves with matter
  int transition, vlink;
  uint64_t seed;
  int sweep;
  // energy
  const double *coeff;  // (R-1)*(C-1) plaquette coefficients (NaN: no term)
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
__device__ inline double gi_advance(const GiPlaqArgs<T> &a, const T *top, const T *bot, int c, int s0, int s1,
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
__device__ inline void gi_copy(const T *x, T *y, int64_t n) {
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) y[i] = x[i];
  __syncthreads();
}

// Plaquette (r, c) flipped by delta: the four links it touches
// (gi/model.py:369-384).
__device__ inline void gi_flip(const GiDev &g, int32_t *lk, int r, int c, int delta) {
  const int C = g.C, nh = g.R * (C - 1);
  lk[r * (C - 1) + c] = gi_mod(lk[r * (C - 1) + c] + delta, g.N);
  lk[(r + 1) * (C - 1) + c] = gi_mod(lk[(r + 1) * (C - 1) + c] - delta, g.N);
  lk[nh + r * C + c] = gi_mod(lk[nh + r * C + c] + delta, g.N);
  lk[nh + r * C + c + 1] = gi_mod(lk[nh + r * C + c + 1] - delta, g.N);
}

template <typename T>
__global__ __launch_bounds__(256) void k_gi_plaquette(GiPlaqArgs<T> a) {
  using Acc = typename Scalar<T>::acc;
  __shared__ double red[32];
  __shared__ double r2log[64];
  __shared__ int32_t lk[GI_MAX_LINKS];
  __shared__ int32_t prop[4];  // proposed slices of (r,c), (r,c+1), (r+1,c), (r+1,c+1)
  __shared__ int32_t ssl[128];  // slices of rows r and r+1
  __shared__ double draw_u, draw_pr;
  __shared__ int draw_delta;
  const int b = blockIdx.x;
  const int C = a.C, r = a.row;
  const GiDev &g = a.g;
  const int ns = g.R * C, nlinks = g.R * (C - 1) + (g.R - 1) * C;
  const T *top = a.top + (int64_t)b * a.top_stride;
  const T *bot = a.bot + (int64_t)b * a.bot_stride;
  T *r2 = a.r2 + (int64_t)b * a.r2_stride;
  T *l2 = a.l2 + (int64_t)b * a.l_stride, *lx = a.lx + (int64_t)b * a.l_stride;
  T *ly = a.ly + (int64_t)b * a.l_stride, *lxp = a.lxp + (int64_t)b * a.l_stride;
  T *lyp = a.lyp + (int64_t)b * a.l_stride;
  T *v1 = a.v1 + (int64_t)b * a.v_stride, *v2 = a.v2 + (int64_t)b * a.v_stride, *v3 = a.v3 + (int64_t)b * a.v_stride;
  int32_t *sample = a.samples + (int64_t)b * (ns + nlinks);
  int32_t *sl = a.slices + (int64_t)b * ns;
  for (int i = threadIdx.x; i < nlinks; i += blockDim.x) lk[i] = sample[ns + i];
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
  if (a.transition && a.vlink) {
    // vertical links (r, c) between rows r and r+1, one column at a time
    for (int c = 0; c < C; ++c) {
      const ColGeo G0 = a.geo0[c], G1 = a.geo1[c];
      const int64_t nl = (int64_t)G0.tr * G0.sr * G1.sr * G1.br;  // L2 after column c = R2[c]
      const T *Rc = r2 + (int64_t)c * a.cap2;
      if (threadIdx.x == 0) {
        const GiLinkMove mv = gi_link_move(g, a.seed, GI_VLINK_BASE + (uint32_t)(r * C + c), a.sweep, b,
                                           sample[r * C + c], sample[(r + 1) * C + c], -1);
        draw_u = mv.u;
        draw_pr = mv.ratio_prop;
        draw_delta = mv.delta;
        prop[1] = mv.sa;
        prop[3] = mv.sb;
        const int li = g.R * (C - 1) + r * C + c, old = lk[li];
        lk[li] = gi_mod(old + mv.delta, g.N);
        prop[0] = mv.sa * gi_nc(g, r, c) + gi_cfg(g, lk, r, c);
        prop[2] = mv.sb * gi_nc(g, r + 1, c) + gi_cfg(g, lk, r + 1, c);
        lk[li] = old;
      }
      __syncthreads();
      const double xl = gi_advance(a, top, bot, c, s0[c], s1[c], l2, lx, v1, v2, v3, red);
      const Acc dcur = block_dot(lx, Rc, nl, red);
      const double xlp = gi_advance(a, top, bot, c, prop[0], prop[2], l2, lxp, v1, v2, v3, red);
      const Acc dprop = block_dot(lxp, Rc, nl, red);
      const bool accept = draw_u < fmin(1.0, gi_ratio(acc_abs(dcur), xl, acc_abs(dprop), xlp, draw_pr));
      if (accept) {
        if (threadIdx.x == 0) {
          const int li = g.R * (C - 1) + r * C + c;
          lk[li] = gi_mod(lk[li] + draw_delta, g.N);
          sample[r * C + c] = prop[1];
          sample[(r + 1) * C + c] = prop[3];
          s0[c] = prop[0];
          s1[c] = prop[2];
        }
        gi_copy(lxp, l2, nl);
      } else {
        gi_copy(lx, l2, nl);
      }
      __syncthreads();
    }
  }
  for (int c = 0; c < (a.transition && a.vlink ? 0 : C - 1); ++c) {
    const ColGeo G0n = a.geo0[c + 1], G1n = a.geo1[c + 1];
    const int64_t nl2 = (int64_t)G0n.tr * G0n.sr * G1n.sr * G1n.br;  // L2 after column c+1 = R2[c+1]
    const T *Rn = r2 + (int64_t)(c + 1) * a.cap2;
    if (a.transition) {
      if (threadIdx.x == 0) {
        uint32_t ctr[4] = {GI_PLAQ_BASE + (uint32_t)(r * C + c), (uint32_t)a.sweep, (uint32_t)b, 0u};
        philox4x32_10(ctr, (uint32_t)(a.seed & 0xFFFFFFFFu), (uint32_t)(a.seed >> 32));
        const uint64_t u53 = ((uint64_t)ctr[0] << 21) | (ctr[1] >> 11);
        draw_u = (double)u53 * 0x1.0p-53;
        draw_delta = 1 + (int)(ctr[2] % (uint32_t)max(g.N - 1, 1));
        gi_flip(g, lk, r, c, draw_delta);
        prop[0] = (sample[r * C + c]) * gi_nc(g, r, c) + gi_cfg(g, lk, r, c);
        prop[1] = (sample[r * C + c + 1]) * gi_nc(g, r, c + 1) + gi_cfg(g, lk, r, c + 1);
        prop[2] = (sample[(r + 1) * C + c]) * gi_nc(g, r + 1, c) + gi_cfg(g, lk, r + 1, c);
        prop[3] = (sample[(r + 1) * C + c + 1]) * gi_nc(g, r + 1, c + 1) + gi_cfg(g, lk, r + 1, c + 1);
        gi_flip(g, lk, r, c, -draw_delta);
      }
      __syncthreads();
      const double xl = gi_advance(a, top, bot, c, s0[c], s1[c], l2, lx, v1, v2, v3, red);
      const double yl = gi_advance(a, top, bot, c + 1, s0[c + 1], s1[c + 1], lx, ly, v1, v2, v3, red);
      const Acc dcur = block_dot(ly, Rn, nl2, red);
      const double xlp = gi_advance(a, top, bot, c, prop[0], prop[2], l2, lxp, v1, v2, v3, red);
      const double ylp = gi_advance(a, top, bot, c + 1, prop[1], prop[3], lxp, lyp, v1, v2, v3, red);
      const Acc dprop = block_dot(lyp, Rn, nl2, red);
      // contraction.py:88-94 on |amp|^2 with the common scale removed
      const double pc = acc_abs(dcur), pp = acc_abs(dprop);
      double ratio;
      if (pc > 0.0 && pp > 0.0 && isfinite(xl + yl) && isfinite(xlp + ylp))
        ratio = exp(2.0 * (log(pp) - log(pc) + (xlp + ylp) - (xl + yl)));
      else if (pc > 0.0 && isfinite(xl + yl))
        ratio = 0.0;
      else
        ratio = (pp > 0.0 && isfinite(xlp + ylp)) ? INFINITY : 0.0;
      const bool accept = draw_u < fmin(1.0, ratio);
      if (accept) {
        if (threadIdx.x == 0) {
          gi_flip(g, lk, r, c, draw_delta);
          s0[c] = prop[0];
          s0[c + 1] = prop[1];
          s1[c] = prop[2];
          s1[c + 1] = prop[3];
        }
        gi_copy(lxp, l2, (int64_t)a.geo0[c].tr * a.geo0[c].sr * a.geo1[c].sr * a.geo1[c].br);
        llog += xlp;
      } else {
        gi_copy(lx, l2, (int64_t)a.geo0[c].tr * a.geo0[c].sr * a.geo1[c].sr * a.geo1[c].br);
        llog += xl;
      }
      __syncthreads();
    } else {
      const double cf = a.coeff[r * (C - 1) + c];
      if (!isnan(cf)) {
        Acc tot = Acc(0.0);
        const int nflip = g.N == 2 ? 1 : 2;  // +1 and -1 coincide for Z2
        for (int f = 0; f < nflip; ++f) {
          const int delta = f == 0 ? 1 : -1;
          if (threadIdx.x == 0) {
            gi_flip(g, lk, r, c, delta);
            prop[0] = (sample[r * C + c]) * gi_nc(g, r, c) + gi_cfg(g, lk, r, c);
            prop[1] = (sample[r * C + c + 1]) * gi_nc(g, r, c + 1) + gi_cfg(g, lk, r, c + 1);
            prop[2] = (sample[(r + 1) * C + c]) * gi_nc(g, r + 1, c) + gi_cfg(g, lk, r + 1, c);
            prop[3] = (sample[(r + 1) * C + c + 1]) * gi_nc(g, r + 1, c + 1) + gi_cfg(g, lk, r + 1, c + 1);
            gi_flip(g, lk, r, c, -delta);
          }
          __syncthreads();
          const double xlp = gi_advance(a, top, bot, c, prop[0], prop[2], l2, lxp, v1, v2, v3, red);
          const double ylp = gi_advance(a, top, bot, c + 1, prop[1], prop[3], lxp, lyp, v1, v2, v3, red);
          const Acc dp = block_dot(lyp, Rn, nl2, red);
          const double sc = exp(base + llog + xlp + ylp + r2log[c + 1]);
          tot += (g.N == 2 ? 2.0 : 1.0) * (dp * polar_acc<T>(sc, phase));
        }
        e_row += cf * tot;
      }
      llog += gi_advance(a, top, bot, c, s0[c], s1[c], l2, l2, v1, v2, v3, red);
    }
  }
  if (a.transition) {
    for (int i = threadIdx.x; i < nlinks; i += blockDim.x) sample[ns + i] = lk[i];
    for (int i = threadIdx.x; i < 2 * C; i += blockDim.x) sl[r * C + i] = ssl[i];
  } else if (threadIdx.x == 0) {
    eloc_add(a.eloc, b, e_row);
  }
}


// Horizontal links of row r with matter moves (gi/model.py:703-792): right
// environments of the row, then per link (r, c) the two-site amplitude
// L . [c] . [c+1] . R[c+1] for the current and the proposed slices (two
// chained left updates and a dot with R[c+1]), Metropolis with the proposal
// ratio, and the left environment through the chosen slice of column c.
// cfg holds the combined slice indices and is updated with the samples.
template <typename T>
__global__ __launch_bounds__(256) void k_gi_hlink(RowArgs<T> a, GiDev g, int32_t *samples) {
  using Acc = typename Scalar<T>::acc;
  __shared__ double red[32];
  __shared__ double rlog[64];
  __shared__ int32_t prop[4];  // proposed slices of (r,c), (r,c+1), then their matter states
  __shared__ double draw_u, draw_pr;
  __shared__ int draw_delta;
  const int b = blockIdx.x;
  const int C = a.C, r = a.row;
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
  const int ns = g.R * C, nlinks = g.R * (C - 1) + (g.R - 1) * C;
  int32_t *sample = samples + (int64_t)b * (ns + nlinks);
  int32_t *lk = sample + ns;

  right_envs(a, top, bot, renv, s1, s2, cfg, rlog, red);
  if (threadIdx.x == 0) lenv[0] = T(1);
  __syncthreads();
  for (int c = 0; c < C - 1; ++c) {
    const ColGeo G0 = a.geo[c], G1 = a.geo[c + 1];
    if (threadIdx.x == 0) {
      const GiLinkMove mv = gi_link_move(g, a.seed, GI_HLINK_BASE + (uint32_t)(r * (C - 1) + c), a.sweep, b,
                                         sample[r * C + c], sample[r * C + c + 1], 1);
      draw_u = mv.u;
      draw_pr = mv.ratio_prop;
      draw_delta = mv.delta;
      prop[2] = mv.sa;
      prop[3] = mv.sb;
      const int li = r * (C - 1) + c, old = lk[li];
      lk[li] = gi_mod(old + mv.delta, g.N);
      prop[0] = mv.sa * gi_nc(g, r, c) + gi_cfg(g, lk, r, c);
      prop[1] = mv.sb * gi_nc(g, r, c + 1) + gi_cfg(g, lk, r, c + 1);
      lk[li] = old;
    }
    __syncthreads();
    const T *tp0 = top + G0.top_off, *bp0 = bot + G0.bot_off;
    const T *tp1 = top + G1.top_off, *bp1 = bot + G1.bot_off;
    const T *A0 = a.sites + G0.site_off, *A1 = a.sites + G1.site_off;
    const int64_t n0 = (int64_t)G0.tr * G0.sr * G0.br, n1 = (int64_t)G1.tr * G1.sr * G1.br;
    const T *Rn = renv + (int64_t)(c + 1) * a.env_cap;
    // Y1 = L . top[c], shared by both slices of column c
    gemm2<T>(lenv, g2(G0.sl, G0.bl, G0.bl, 1), g2(G0.tl, (int64_t)G0.sl * G0.bl), tp0, g2(G0.tl, (int64_t)G0.tp * G0.tr),
             g2(G0.tp * G0.tr, 1), s1, g2(G0.sl * G0.bl, (int64_t)G0.tp * G0.tr), g2(G0.tp * G0.tr, 1));
    __syncthreads();
    double x0c = 0.0, x0p = 0.0, x1c = 0.0, x1p = 0.0;
    left_update(G0, A0 + (int64_t)cfg[c] * G0.site_slice, s1, bp0, s2, s3);
    normalise(s3, n0, red, &x0c);
    left_update(G0, A0 + (int64_t)prop[0] * G0.site_slice, s1, bp0, s2, s4);
    normalise(s4, n0, red, &x0p);
    // two-site amplitudes: column c+1 with the current / proposed slice
    gemm2<T>(s3, g2(G1.sl, G1.bl, G1.bl, 1), g2(G1.tl, (int64_t)G1.sl * G1.bl), tp1, g2(G1.tl, (int64_t)G1.tp * G1.tr),
             g2(G1.tp * G1.tr, 1), s1, g2(G1.sl * G1.bl, (int64_t)G1.tp * G1.tr), g2(G1.tp * G1.tr, 1));
    __syncthreads();
    left_update(G1, A1 + (int64_t)cfg[c + 1] * G1.site_slice, s1, bp1, s2, s5);
    normalise(s5, n1, red, &x1c);
    const Acc dcur = block_dot(s5, Rn, n1, red);
    gemm2<T>(s4, g2(G1.sl, G1.bl, G1.bl, 1), g2(G1.tl, (int64_t)G1.sl * G1.bl), tp1, g2(G1.tl, (int64_t)G1.tp * G1.tr),
             g2(G1.tp * G1.tr, 1), s1, g2(G1.sl * G1.bl, (int64_t)G1.tp * G1.tr), g2(G1.tp * G1.tr, 1));
    __syncthreads();
    left_update(G1, A1 + (int64_t)prop[1] * G1.site_slice, s1, bp1, s2, s5);
    normalise(s5, n1, red, &x1p);
    const Acc dprop = block_dot(s5, Rn, n1, red);
    const bool accept =
        draw_u < fmin(1.0, gi_ratio(acc_abs(dcur), x0c + x1c, acc_abs(dprop), x0p + x1p, draw_pr));
    if (accept) {
      if (threadIdx.x == 0) {
        const int li = r * (C - 1) + c;
        lk[li] = gi_mod(lk[li] + draw_delta, g.N);
        sample[r * C + c] = prop[2];
        sample[r * C + c + 1] = prop[3];
        cfg[c] = prop[0];
        cfg[c + 1] = prop[1];
      }
      gi_copy(s4, lenv, n0);
    } else {
      gi_copy(s3, lenv, n0);
    }
    __syncthreads();
  }
}

}  // namespace vmc
