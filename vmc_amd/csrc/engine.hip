// Batched PEPS amplitude + env-gradient engine for gfx950 and its C ABI.
//
// Execution model: layer-synchronous over the lattice, data-parallel over
// samples. Every kernel launch processes one lattice step (one zip-up column,
// one row of environments) for the whole batch, one workgroup per sample, so a
// launch has `batch` (thousands of) workgroups. Boundary-MPS tensors, carries,
// environments and scratch are per-sample slices of one caller-owned device
// workspace (region-major: region + sample * stride).
//
// Scaling: every boundary MPS carries an FP64 log-scale. Zip-up truncations
// normalise the carry (trunc.hip) and environments are renormalised after each
// step, so FP32 contractions never under- or overflow (|psi| ~ 1e-85 at 10x10).
#include "common.hpp"
#include "../../include/vmc_amd.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace vmc {

static thread_local std::string g_last_error;

static int fail(const std::string &msg) {
  g_last_error = msg;
  return -1;
}

#define VMC_HIP(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) return fail(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Dims3 {
  int l, p, r;
};

struct SiteInfo {
  int u, dn, l, r;   // bond dims of the site tensor (up, down, left, right)
  int64_t off;       // element offset of A[r][c] in the parameter vector
  int64_t slice;     // u * dn * l * r
  int64_t o_off;     // offset of the site's sliced block in o
};

constexpr int MAX_MAT = 128;  // rows/cols of a zip-up theta matrix handled by trunc.hip
constexpr int MAX_KEEP = 32;  // kept singular vectors per truncation

// Geometry of one lattice site inside the row environment pass.
struct ColGeo {
  int64_t top_off, bot_off, site_off, site_slice, o_off;
  int tl, tp, tr;  // top boundary tensor (a, u, b) before this row
  int bl, bp, br;  // bottom boundary tensor (e, v, f) below this row
  int su, sd, sl, sr;
};

}  // namespace vmc

struct vmc_plan {
  int R, C, d, D, chi, strategy, dtype;
  int64_t nparams, nsliced;
  std::vector<vmc::SiteInfo> site;
  std::vector<std::vector<vmc::Dims3>> top;      // [R+1][C]: boundary before row s
  std::vector<std::vector<int64_t>> top_off;     // per-sample element offsets
  int64_t top_size;
  std::vector<std::vector<vmc::Dims3>> bot;      // [R][C]: bottom boundary used by row r
  std::vector<std::vector<int64_t>> bot_off;
  int64_t bot_cap;
  int64_t mat_cap, carry_cap, scratch_cap, env_cap;
  void *geo_dev = nullptr;  // vmc::ColGeo[R * C], uploaded once at plan creation
  // Optional per-kernel-class timing with HIP events (vmc_profile_begin/end).
  bool prof = false;
  std::vector<int> prof_cls;
  std::vector<hipEvent_t> prof_ev;  // start/stop pairs
};

namespace vmc {

// ------------------------------------------------------------------------- //
// Static shape bookkeeping (zip-up bond dimensions depend only on shapes)
// ------------------------------------------------------------------------- //

// Dimensions of the boundary after applying a row MPO with the zip-up rule of
// src/vmc/peps/common/strategy.py:128-150. `from_below` swaps the roles of the
// up and down bonds (src/vmc/peps/common/contraction.py:59-65).
static std::vector<Dims3> zipup_dims(const std::vector<Dims3> &in, const vmc_plan &p, int row,
                                     bool from_below, std::vector<int> *mats_m = nullptr,
                                     std::vector<int> *mats_n = nullptr) {
  std::vector<Dims3> out(p.C);
  int carry_k = -1;
  for (int c = 0; c < p.C; ++c) {
    const SiteInfo &s = p.site[row * p.C + c];
    const int q = from_below ? s.u : s.dn;
    const int left = (c == 0) ? in[c].l * s.l : carry_k;
    const int ncols = in[c].r * s.r;
    if (c == p.C - 1) {
      out[c] = {left, q, ncols};
      break;
    }
    const int k = std::min(p.chi, std::min(left * q, ncols));
    out[c] = {left, q, k};
    carry_k = k;
    if (mats_m) mats_m->push_back(left * q);
    if (mats_n) mats_n->push_back(ncols);
  }
  return out;
}

static int64_t dsize(const Dims3 &d) { return (int64_t)d.l * d.p * d.r; }

static int build_plan(vmc_plan *p) {
  const int R = p->R, C = p->C, D = p->D;
  p->site.resize((size_t)R * C);
  int64_t off = 0, ooff = 0;
  for (int r = 0; r < R; ++r)
    for (int c = 0; c < C; ++c) {
      SiteInfo &s = p->site[r * C + c];
      s.u = r == 0 ? 1 : D;
      s.dn = r == R - 1 ? 1 : D;
      s.l = c == 0 ? 1 : D;
      s.r = c == C - 1 ? 1 : D;
      s.slice = (int64_t)s.u * s.dn * s.l * s.r;
      s.off = off;
      s.o_off = ooff;
      off += s.slice * p->d;
      ooff += s.slice;
    }
  p->nparams = off;
  p->nsliced = ooff;

  std::vector<int> mm, mn;
  // top sweep
  p->top.assign(R + 1, std::vector<Dims3>(C, Dims3{1, 1, 1}));
  for (int r = 0; r < R; ++r) p->top[r + 1] = zipup_dims(p->top[r], *p, r, false, &mm, &mn);
  p->top_off.assign(R + 1, std::vector<int64_t>(C, 0));
  int64_t t = 0;
  for (int s = 0; s <= R; ++s)
    for (int c = 0; c < C; ++c) {
      p->top_off[s][c] = t;
      t += dsize(p->top[s][c]);
    }
  p->top_size = t;
  // bottom sweep: bot[R-1] = ones, bot[r] = apply(bot[r+1], row r+1 from below)
  p->bot.assign(R, std::vector<Dims3>(C, Dims3{1, 1, 1}));
  for (int r = R - 2; r >= 0; --r) p->bot[r] = zipup_dims(p->bot[r + 1], *p, r + 1, true, &mm, &mn);
  p->bot_off.assign(R, std::vector<int64_t>(C, 0));
  p->bot_cap = 0;
  for (int r = 0; r < R; ++r) {
    int64_t b = 0;
    for (int c = 0; c < C; ++c) {
      p->bot_off[r][c] = b;
      b += dsize(p->bot[r][c]);
    }
    p->bot_cap = std::max(p->bot_cap, b);
  }
  // capacities
  int64_t mat_cap = 1, carry_cap = 1, scr = 1, env = 1;
  for (size_t i = 0; i < mm.size(); ++i) {
    if (mm[i] > MAX_MAT || mn[i] > MAX_MAT)
      return fail("theta matrix " + std::to_string(mm[i]) + "x" + std::to_string(mn[i]) +
                  " exceeds the 128x128 truncation kernel");
    mat_cap = std::max<int64_t>(mat_cap, (int64_t)mm[i] * mn[i]);
    carry_cap = std::max<int64_t>(carry_cap, (int64_t)std::min(p->chi, std::min(mm[i], mn[i])) * mn[i]);
    if (std::min(p->chi, std::min(mm[i], mn[i])) > MAX_KEEP)
      return fail("chi above 32 is not supported by the truncation kernel");
  }
  // last-column thetas and intermediates: bound every einsum intermediate by
  // (max boundary bond)^2 * D^2 and environments by bond * D * bond.
  int maxb = 1;
  for (auto &row : p->top)
    for (auto &d3 : row) maxb = std::max({maxb, d3.l, d3.r});
  for (auto &row : p->bot)
    for (auto &d3 : row) maxb = std::max({maxb, d3.l, d3.r});
  scr = (int64_t)maxb * maxb * D * D * std::max(1, p->d);
  env = (int64_t)maxb * maxb * D;
  p->mat_cap = mat_cap;
  p->carry_cap = carry_cap;
  p->scratch_cap = std::max(scr, mat_cap);
  p->env_cap = env;

  std::vector<ColGeo> geo((size_t)R * C);
  for (int r = 0; r < R; ++r)
    for (int c = 0; c < C; ++c) {
      const SiteInfo &s = p->site[r * C + c];
      ColGeo &G = geo[r * C + c];
      G.top_off = p->top_off[r][c];
      G.bot_off = p->bot_off[r][c];
      G.site_off = s.off;
      G.site_slice = s.slice;
      G.o_off = s.o_off;
      G.tl = p->top[r][c].l;
      G.tp = p->top[r][c].p;
      G.tr = p->top[r][c].r;
      G.bl = p->bot[r][c].l;
      G.bp = p->bot[r][c].p;
      G.br = p->bot[r][c].r;
      G.su = s.u;
      G.sd = s.dn;
      G.sl = s.l;
      G.sr = s.r;
    }
  VMC_HIP(hipMalloc(&p->geo_dev, geo.size() * sizeof(ColGeo)));
  VMC_HIP(hipMemcpy(p->geo_dev, geo.data(), geo.size() * sizeof(ColGeo), hipMemcpyHostToDevice));
  return 0;
}

// ------------------------------------------------------------------------- //
// Kernels
// ------------------------------------------------------------------------- //

enum KernelClass { KC_THETA = 0, KC_TRUNC = 1, KC_ROW = 2, KC_AMP = 3, KC_FILL = 4, KC_COUNT = 5 };

// Brackets one launch with HIP events when the plan is profiling.
struct ProfScope {
  vmc_plan *p;
  hipStream_t st;
  ProfScope(vmc_plan *p_, int cls, hipStream_t st_) : p(p_), st(st_) {
    if (!p->prof) return;
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    p->prof_cls.push_back(cls);
    p->prof_ev.push_back(a);
    p->prof_ev.push_back(b);
    (void)hipEventRecord(a, st);
  }
  ~ProfScope() {
    if (p->prof) (void)hipEventRecord(p->prof_ev.back(), st);
  }
};

// x[b * stride + i] = v for i < count (small initialisations inside the call,
// kept as kernels so the call stays capturable in a HIP graph).
template <typename T>
__global__ void k_fill(T *x, int64_t stride, int count, T v, int batch) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)batch * count) return;
  const int64_t b = i / count, j = i - b * count;
  x[b * stride + j] = v;
}

template <typename T>
static void fill(T *x, int64_t stride, int count, T v, int batch, hipStream_t st) {
  const int64_t n = (int64_t)batch * count;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, stride, count, v, batch);
}

// Per-sample site slice: A[r][c][config] with layout (up, down, left, right).
template <typename T>
struct SiteRef {
  const T *base;      // A[r][c] (all physical slices)
  int64_t slice;      // elements per physical slice
  const int32_t *cfg; // configs (batch, n_sites)
  int n_sites, idx;   // site index
  __device__ const T *get(int b) const {
    return base + (int64_t)cfg[(int64_t)b * n_sites + idx] * slice;
  }
};

// theta = carry . m . w  (src/vmc/peps/common/strategy.py:95-115, pairwise
// order [(0,1),(0,1)]: carry.m over Dl first, then .w over (wl, p)).
template <typename T>
struct ThetaArgs {
  const T *carry;    // (k, Dl, wl) or null at column 0 (then k = Dl = wl = 1)
  int64_t carry_stride;
  int k, Dl, wl;
  const T *m;        // incoming boundary tensor (Dl, P, Dr)
  int64_t m_stride;
  int P, Dr;
  SiteRef<T> site;
  int64_t w_sl, w_sr, w_sp, w_sq;  // strides of w[wl, wr, p, q] inside the site slice
  int WR, Q;
  T *t1;             // scratch (k, wl, P, Dr)
  int64_t t1_stride;
  T *out;            // (k*Q, Dr*WR)
  int64_t out_stride;
  const double *log_in;  // column 0: log_out = log_in
  double *log_out;
  int64_t log_in_stride, log_out_stride;
};

template <typename T>
__global__ __launch_bounds__(256) void k_theta(ThetaArgs<T> a) {
  const int b = blockIdx.x;
  const T *m = a.m + (int64_t)b * a.m_stride;
  const T *w = a.site.get(b);
  T *out = a.out + (int64_t)b * a.out_stride;
  const T *t1 = m;
  int k = 1, wl = 1;
  if (a.carry) {
    k = a.k;
    wl = a.wl;
    const T *carry = a.carry + (int64_t)b * a.carry_stride;
    T *t1w = a.t1 + (int64_t)b * a.t1_stride;
    const int PD = a.P * a.Dr;
    gemm2<T>(carry, g2(k, (int64_t)a.Dl * wl, wl, 1), g2(a.Dl, wl),
             m, g2(a.Dl, PD), g2(PD, 1),
             t1w, g2(k, (int64_t)wl * PD, wl, PD), g2(PD, 1));
    __syncthreads();
    t1 = t1w;
  }
  const int PD = a.P * a.Dr;
  gemm2<T>(t1, g2(k, (int64_t)wl * PD, a.Dr, 1), g2(wl, PD, a.P, a.Dr),
           w, g2(wl, a.w_sl, a.P, a.w_sp), g2(a.Q, a.w_sq, a.WR, a.w_sr),
           out, g2(k, (int64_t)a.Q * a.Dr * a.WR, a.Dr, a.WR), g2(a.Q, (int64_t)a.Dr * a.WR, a.WR, 1));
  if (a.log_in && threadIdx.x == 0)
    a.log_out[(int64_t)b * a.log_out_stride] = a.log_in[(int64_t)b * a.log_in_stride];
}

// Amplitude of the final boundary (src/vmc/peps/common/contraction.py:30-35).
template <typename T>
struct AmpArgs {
  const T *bnd;
  int64_t bnd_stride;
  const double *logscale;
  int64_t log_stride;
  double *log_amp;  // (batch, 2)
  int C;
  int64_t off[64];
  int l[64], r[64];
};

template <typename T>
__global__ __launch_bounds__(64) void k_amp(AmpArgs<T> a) {
  __shared__ double st[2][MAX_MAT];
  const int b = blockIdx.x, t = threadIdx.x;
  const T *bnd = a.bnd + (int64_t)b * a.bnd_stride;
  double lsc = a.logscale[(int64_t)b * a.log_stride];
  int cur = 0;
  if (t == 0) st[0][0] = 1.0;
  __syncthreads();
  for (int c = 0; c < a.C; ++c) {
    const T *M = bnd + a.off[c];
    const int L = a.l[c], Rr = a.r[c];
    double mx = 0.0;
    for (int j = t; j < Rr; j += 64) {
      double s = 0.0;
      for (int i = 0; i < L; ++i) s = fma(st[cur][i], (double)M[(int64_t)i * Rr + j], s);
      st[cur ^ 1][j] = s;
      mx = fmax(mx, fabs(s));
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    __syncthreads();
    cur ^= 1;
    if (mx > 0.0) {
      for (int j = t; j < Rr; j += 64) st[cur][j] /= mx;
      lsc += log(mx);
    }
    __syncthreads();
  }
  if (t == 0) {
    const double s = st[cur][0];
    a.log_amp[2 * (int64_t)b] = log(fabs(s)) + lsc;
    a.log_amp[2 * (int64_t)b + 1] = s < 0.0 ? M_PI : 0.0;
  }
}

// Row environment pass: right environments (src/vmc/peps/common/contraction.py:68-86)
// and then the left-to-right env-gradient sweep (src/vmc/peps/common/energy.py:98-120,
// 259-273), one workgroup per sample looping over the columns of the row.
template <typename T>
struct RowArgs {
  int C;
  const T *top;  int64_t top_stride;  // top boundary region base (per-sample stride)
  const T *bot;  int64_t bot_stride;
  const T *sites;                      // parameter vector
  const int32_t *cfg; int n_sites; int row;
  T *renv; int64_t renv_stride; int64_t env_cap;
  double *renv_log; int64_t renv_log_stride;
  T *lenv; int64_t lenv_stride;
  T *s1, *s2; int64_t s_stride;
  const double *top_log; int64_t top_log_stride;
  const double *bot_log; int64_t bot_log_stride;
  const double *log_amp;
  T *o; int64_t o_stride;
  int8_t *p;
  const ColGeo *geo;  // this row's C entries
};

template <typename T>
__device__ inline void normalise(T *x, int64_t n, double *red, double *logacc) {
  double ss = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = (double)x[i];
    ss = fma(v, v, ss);
  }
  ss = block_sum(ss, red);
  const double nrm = sqrt(ss);
  if (nrm > 0.0) {
    const T sc = (T)(1.0 / nrm);
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) x[i] *= sc;
    *logacc += log(nrm);
  } else {
    *logacc += -INFINITY;
  }
  __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(256) void k_row_grad(RowArgs<T> a) {
  __shared__ double red[8];
  __shared__ double rlog[64];
  const int b = blockIdx.x;
  const int C = a.C;
  const T *top = a.top + (int64_t)b * a.top_stride;
  const T *bot = a.bot + (int64_t)b * a.bot_stride;
  T *renv = a.renv + (int64_t)b * a.renv_stride;
  T *lenv = a.lenv + (int64_t)b * a.lenv_stride;
  T *s1 = a.s1 + (int64_t)b * a.s_stride;
  T *s2 = a.s2 + (int64_t)b * a.s_stride;
  const int32_t *cfg = a.cfg + (int64_t)b * a.n_sites + (int64_t)a.row * C;

  // ---- right environments R[c], c = C-1 .. 0 ----
  if (threadIdx.x == 0) {
    renv[(int64_t)(C - 1) * a.env_cap] = T(1);
    rlog[C - 1] = 0.0;
  }
  __syncthreads();
  for (int c = C - 2; c >= 0; --c) {
    const int cc = c + 1;
    const ColGeo G = a.geo[cc];
    const T *tp = top + G.top_off;
    const T *bp = bot + G.bot_off;
    const T *w = a.sites + G.site_off + (int64_t)cfg[cc] * G.site_slice;
    const T *Rin = renv + (int64_t)cc * a.env_cap;
    T *Rout = renv + (int64_t)c * a.env_cap;
    const int ta = G.tl, tu = G.tp, tb = G.tr;
    const int be = G.bl, bv = G.bp, bf = G.br;
    const int wl = G.sl, wr = G.sr, wu = G.su, wd = G.sd;
    const int64_t s_u = (int64_t)wd * wl * wr, s_dn = (int64_t)wl * wr, s_l = wr, s_r = 1;
    // X1[a,u,d,f] = top[a,u,b] R[b,d,f]
    gemm2<T>(tp, g2(ta * tu, tb), g2(tb, 1), Rin, g2(tb, (int64_t)wr * bf), g2(wr * bf, 1),
             s1, g2(ta * tu, (int64_t)wr * bf), g2(wr * bf, 1));
    __syncthreads();
    // X2[c,v,a,f] = mpo[c,d,u,v] X1[a,u,d,f]   (mpo = w[l, r, u, dn])
    gemm2<T>(w, g2(wl, s_l, wd, s_dn), g2(wr, s_r, wu, s_u),
             s1, g2(wr, bf, tu, (int64_t)wr * bf), g2(ta, (int64_t)tu * wr * bf, bf, 1),
             s2, g2(wl, (int64_t)wd * ta * bf, wd, (int64_t)ta * bf), g2(ta, bf, bf, 1));
    __syncthreads();
    // R[a,c,e] = X2[c,v,a,f] bot[e,v,f]
    gemm2<T>(s2, g2(ta, bf, wl, (int64_t)wd * ta * bf), g2(wd, (int64_t)ta * bf, bf, 1),
             bp, g2(bv, bf, bf, 1), g2(be, (int64_t)bv * bf),
             Rout, g2(ta, (int64_t)wl * be, wl, be), g2(be, 1));
    __syncthreads();
    double lg = rlog[cc];
    normalise(Rout, (int64_t)ta * wl * be, red, &lg);
    if (threadIdx.x == 0) rlog[c] = lg;
    __syncthreads();
  }

  // ---- left-to-right env gradients ----
  if (threadIdx.x == 0) lenv[0] = T(1);
  double llog = 0.0;
  __syncthreads();
  const double base = a.top_log[(int64_t)b * a.top_log_stride] + a.bot_log[(int64_t)b * a.bot_log_stride] -
                      a.log_amp[2 * (int64_t)b];
  const double sgn = a.log_amp[2 * (int64_t)b + 1] != 0.0 ? -1.0 : 1.0;
  for (int c = 0; c < C; ++c) {
    const ColGeo G = a.geo[c];
    const T *tp = top + G.top_off;
    const T *bp = bot + G.bot_off;
    const T *w = a.sites + G.site_off + (int64_t)cfg[c] * G.site_slice;
    const T *Rc = renv + (int64_t)c * a.env_cap;
    const int ta = G.tl, tu = G.tp, tb = G.tr;
    const int be = G.bl, bv = G.bp, bf = G.br;
    const int wl = G.sl, wr = G.sr, wu = G.su, wd = G.sd;
    const int64_t s_u = (int64_t)wd * wl * wr, s_dn = (int64_t)wl * wr, s_l = wr, s_r = 1;
    // Y1[c,e,u,b] = L[a,c,e] top[a,u,b]
    gemm2<T>(lenv, g2(wl, be, be, 1), g2(ta, (int64_t)wl * be),
             tp, g2(ta, (int64_t)tu * tb), g2(tu * tb, 1),
             s1, g2(wl * be, (int64_t)tu * tb), g2(tu * tb, 1));
    // Y2[e,v,b,d] = bot[e,v,f] R[b,d,f]
    gemm2<T>(bp, g2(be * bv, bf), g2(bf, 1),
             Rc, g2(bf, 1), g2(tb, (int64_t)wr * bf, wr, bf),
             s2, g2(be * bv, (int64_t)tb * wr), g2(tb * wr, 1));
    __syncthreads();
    if (a.o) {
      // grad[c,u,v,d] = Y1[c,e,u,b] Y2[e,v,b,d], stored as (up, down, left, right) / psi
      const double scale = sgn * exp(base + llog + rlog[c]);
      T *o = a.o + (int64_t)b * a.o_stride + G.o_off;
      gemm2<T>(s1, g2(wl, (int64_t)be * tu * tb, tu, tb), g2(be, (int64_t)tu * tb, tb, 1),
               s2, g2(be, (int64_t)bv * tb * wr, tb, wr), g2(bv, (int64_t)tb * wr, wr, 1),
               o, g2(wl, wr, tu, (int64_t)bv * wl * wr), g2(bv, (int64_t)wl * wr, wr, 1), (T)scale);
      if (a.p) {
        int8_t *pp = a.p + (int64_t)b * a.o_stride + G.o_off;
        const int64_t n = (int64_t)tu * bv * wl * wr;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) pp[i] = (int8_t)cfg[c];
      }
    }
    if (c < C - 1) {
      // Z2[d,v,e,b] = mpo[c,d,u,v] Y1[c,e,u,b]  -> into s2 after Y2 use
      __syncthreads();
      gemm2<T>(w, g2(wr, s_r, wd, s_dn), g2(wl, s_l, wu, s_u),
               s1, g2(wl, (int64_t)be * tu * tb, tu, tb), g2(be, (int64_t)tu * tb, tb, 1),
               s2, g2(wr, (int64_t)wd * be * tb, wd, (int64_t)be * tb), g2(be, tb, tb, 1));
      __syncthreads();
      // L[b,d,f] = bot[e,v,f] Z2[d,v,e,b]
      gemm2<T>(s2, g2(tb, 1, wr, (int64_t)wd * be * tb), g2(be, tb, wd, (int64_t)be * tb),
               bp, g2(be, (int64_t)bv * bf, bv, bf), g2(bf, 1),
               lenv, g2(tb, (int64_t)wr * bf, wr, bf), g2(bf, 1));
      __syncthreads();
      normalise(lenv, (int64_t)tb * wr * bf, red, &llog);
    }
  }
}

// ------------------------------------------------------------------------- //
// Host orchestration
// ------------------------------------------------------------------------- //

struct Region {
  size_t off, bytes;
};

struct Layout {
  Region top, toplog, bot, botlog, carry, mat, s1, s2, s3, renv, renvlog, lenv;
  size_t total;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static Layout make_layout(const vmc_plan *p, int batch) {
  const size_t es = p->dtype == VMC_F64 ? 8 : 4;
  Layout L{};
  size_t off = 0;
  auto take = [&](Region &r, size_t bytes) {
    r.off = off;
    r.bytes = bytes;
    off = align256(off + bytes);
  };
  const size_t B = (size_t)batch;
  take(L.top, B * p->top_size * es);
  take(L.toplog, B * (p->R + 1) * 8);
  take(L.bot, B * 2 * p->bot_cap * es);
  take(L.botlog, B * 2 * 8);
  take(L.carry, B * p->carry_cap * es);
  take(L.mat, B * p->mat_cap * es);
  take(L.s1, B * p->scratch_cap * es);
  take(L.s2, B * p->scratch_cap * es);
  take(L.s3, B * p->scratch_cap * es);
  take(L.renv, B * p->C * p->env_cap * es);
  take(L.renvlog, B * p->C * 8);
  take(L.lenv, B * p->env_cap * es);
  L.total = off;
  return L;
}

// Apply one row MPO to a boundary with zip-up (top sweep: from_below = false).
template <typename T>
static void apply_row(const vmc_plan *p, int row, bool from_below, const T *sites, const int32_t *cfg,
                      const T *in, int64_t in_stride, const std::vector<int64_t> &in_off,
                      const std::vector<Dims3> &in_dims, const double *in_log, int64_t in_log_stride,
                      T *out, int64_t out_stride, const std::vector<int64_t> &out_off,
                      const std::vector<Dims3> &out_dims, double *out_log, int64_t out_log_stride,
                      T *carry, int64_t carry_stride, T *mat, int64_t mat_stride, T *t1,
                      int64_t t1_stride, int batch, hipStream_t st) {
  const int C = p->C;
  int k_prev = 1;
  for (int c = 0; c < C; ++c) {
    const SiteInfo &s = p->site[row * C + c];
    ThetaArgs<T> a{};
    a.carry = c == 0 ? nullptr : carry;
    a.carry_stride = carry_stride;
    a.k = k_prev;
    a.Dl = in_dims[c].l;
    a.wl = s.l;
    a.m = in + in_off[c];
    a.m_stride = in_stride;
    a.P = in_dims[c].p;
    a.Dr = in_dims[c].r;
    a.site = SiteRef<T>{sites + s.off, s.slice, cfg, p->R * C, row * C + c};
    const int64_t s_u = (int64_t)s.dn * s.l * s.r, s_dn = (int64_t)s.l * s.r;
    a.w_sl = s.r;
    a.w_sr = 1;
    a.w_sp = from_below ? s_dn : s_u;
    a.w_sq = from_below ? s_u : s_dn;
    a.WR = s.r;
    a.Q = from_below ? s.u : s.dn;
    a.t1 = t1;
    a.t1_stride = t1_stride;
    const bool last = c == C - 1;
    a.out = last ? out + out_off[c] : mat;
    a.out_stride = last ? out_stride : mat_stride;
    a.log_in = c == 0 ? in_log : nullptr;
    a.log_in_stride = in_log_stride;
    a.log_out = out_log;
    a.log_out_stride = out_log_stride;
    {
      ProfScope ps(const_cast<vmc_plan *>(p), KC_THETA, st);
      hipLaunchKernelGGL(k_theta<T>, dim3(batch), dim3(256), 0, st, a);
    }
    if (last) break;
    TruncArgs<T> tr{};
    tr.mat = mat;
    tr.mat_stride = mat_stride;
    tr.m = (c == 0 ? in_dims[c].l * s.l : k_prev) * a.Q;
    tr.n = in_dims[c].r * s.r;
    tr.k = out_dims[c].r;
    tr.u = out + out_off[c];
    tr.u_stride = out_stride;
    tr.carry = carry;
    tr.carry_stride = carry_stride;
    tr.logscale = out_log;
    tr.log_stride = out_log_stride;
    {
      ProfScope ps(const_cast<vmc_plan *>(p), KC_TRUNC, st);
      launch_trunc<T>(tr, batch, st);
    }
    k_prev = out_dims[c].r;
  }
}

template <typename T>
static int amp_grad_impl(vmc_plan *p, const T *sites, const int32_t *cfg, int batch, double *log_amp,
                         T *o, int8_t *pout, void *ws, size_t ws_bytes, hipStream_t st) {
  const Layout L = make_layout(p, batch);
  if (ws_bytes < L.total) return fail("workspace too small: need " + std::to_string(L.total));
  char *base = (char *)ws;
  T *top = (T *)(base + L.top.off);
  double *toplog = (double *)(base + L.toplog.off);
  T *bot = (T *)(base + L.bot.off);
  double *botlog = (double *)(base + L.botlog.off);
  T *carry = (T *)(base + L.carry.off);
  T *mat = (T *)(base + L.mat.off);
  T *s1 = (T *)(base + L.s1.off);
  T *s2 = (T *)(base + L.s2.off);
  T *renv = (T *)(base + L.renv.off);
  double *renvlog = (double *)(base + L.renvlog.off);
  T *lenv = (T *)(base + L.lenv.off);
  const int R = p->R, C = p->C;
  if (C > 64) return fail("n_cols above 64 is not supported");
  const int64_t tstride = p->top_size;

  // top boundary before row 0: C tensors of shape (1,1,1) equal to one, log 0
  fill<double>(toplog, R + 1, 1, 0.0, batch, st);
  fill<T>(top, tstride, C, T(1), batch, st);
  for (int r = 0; r < R; ++r) {
    std::vector<int64_t> in_off(p->top_off[r]), out_off(p->top_off[r + 1]);
    apply_row<T>(p, r, false, sites, cfg, top, tstride, in_off, p->top[r], toplog + r, R + 1, top, tstride,
                 out_off, p->top[r + 1], toplog + r + 1, R + 1, carry, p->carry_cap, mat, p->mat_cap, s1,
                 p->scratch_cap, batch, st);
  }
  {
    AmpArgs<T> a{};
    a.bnd = top;
    a.bnd_stride = tstride;
    a.logscale = toplog + R;
    a.log_stride = R + 1;
    a.log_amp = log_amp;
    a.C = C;
    for (int c = 0; c < C; ++c) {
      a.off[c] = p->top_off[R][c];
      a.l[c] = p->top[R][c].l;
      a.r[c] = p->top[R][c].r;
    }
    ProfScope ps(p, KC_AMP, st);
    hipLaunchKernelGGL(k_amp<T>, dim3(batch), dim3(64), 0, st, a);
  }
  if (!o && !pout) return 0;

  // bottom sweep with env-gradients, row R-1 .. 0
  int cur = 0;
  fill<double>(botlog, 2, 1, 0.0, batch, st);
  fill<T>(bot, 2 * p->bot_cap, C, T(1), batch, st);
  for (int r = R - 1; r >= 0; --r) {
    RowArgs<T> a{};
    a.C = C;
    a.top = top;
    a.top_stride = tstride;
    a.bot = bot + (int64_t)cur * p->bot_cap;
    a.bot_stride = 2 * p->bot_cap;
    a.sites = sites;
    a.cfg = cfg;
    a.n_sites = R * C;
    a.row = r;
    a.renv = renv;
    a.renv_stride = (int64_t)C * p->env_cap;
    a.env_cap = p->env_cap;
    a.renv_log = renvlog;
    a.renv_log_stride = C;
    a.lenv = lenv;
    a.lenv_stride = p->env_cap;
    a.s1 = s1;
    a.s2 = s2;
    a.s_stride = p->scratch_cap;
    a.top_log = toplog + r;
    a.top_log_stride = R + 1;
    a.bot_log = botlog + cur;
    a.bot_log_stride = 2;
    a.log_amp = log_amp;
    a.o = o;
    a.o_stride = p->nsliced;
    a.p = pout;
    a.geo = (const ColGeo *)p->geo_dev + (int64_t)r * C;
    {
      ProfScope ps(p, KC_ROW, st);
      hipLaunchKernelGGL(k_row_grad<T>, dim3(batch), dim3(256), 0, st, a);
    }
    if (r == 0) break;
    const int nxt = cur ^ 1;
    apply_row<T>(p, r, true, sites, cfg, bot + (int64_t)cur * p->bot_cap, 2 * p->bot_cap, p->bot_off[r],
                 p->bot[r], botlog + cur, 2, bot + (int64_t)nxt * p->bot_cap, 2 * p->bot_cap,
                 p->bot_off[r - 1], p->bot[r - 1], botlog + nxt, 2, carry, p->carry_cap, mat, p->mat_cap,
                 s1, p->scratch_cap, batch, st);
    cur = nxt;
  }
  return 0;
}

}  // namespace vmc

// ------------------------------------------------------------------------- //
// C ABI
// ------------------------------------------------------------------------- //

extern "C" {

int vmc_plan_create(vmc_plan **plan, int n_rows, int n_cols, int phys_dim, int bond_dim, int chi,
                    int strategy, int dtype) {
  if (!plan) return vmc::fail("plan is null");
  if (n_rows < 1 || n_cols < 1 || phys_dim < 1 || bond_dim < 1)
    return vmc::fail("lattice shape, phys_dim and bond_dim must be positive");
  if (n_cols > 64) return vmc::fail("n_cols above 64 is not supported");
  if (dtype != VMC_F32 && dtype != VMC_F64) return vmc::fail("dtype must be VMC_F32 or VMC_F64");
  if (strategy != VMC_ZIPUP && strategy != VMC_NO_TRUNCATION && strategy != VMC_DENSITY_MATRIX)
    return vmc::fail("unsupported strategy");
  if (strategy == VMC_NO_TRUNCATION) chi = 1 << 30;  // keep every singular vector (fails above 32)
  if (chi < 1) return vmc::fail("chi must be positive");
  vmc_plan *p = new vmc_plan();
  p->R = n_rows;
  p->C = n_cols;
  p->d = phys_dim;
  p->D = bond_dim;
  p->chi = chi;
  p->strategy = strategy;
  p->dtype = dtype;
  if (vmc::build_plan(p) != 0) {
    if (p->geo_dev) (void)hipFree(p->geo_dev);
    delete p;
    return -1;
  }
  *plan = p;
  return 0;
}

int vmc_plan_destroy(vmc_plan *plan) {
  if (plan && plan->geo_dev) (void)hipFree(plan->geo_dev);
  delete plan;
  return 0;
}

int64_t vmc_plan_num_params(const vmc_plan *plan) { return plan ? plan->nparams : -1; }
int64_t vmc_plan_num_sliced(const vmc_plan *plan) { return plan ? plan->nsliced : -1; }

size_t vmc_amp_grad_workspace_size(const vmc_plan *plan, int batch) {
  if (!plan || batch < 1) return 0;
  return vmc::make_layout(plan, batch).total;
}

int vmc_amp_grad(vmc_plan *plan, const void *tensors, const int32_t *configs, int batch, double *log_amp,
                 void *o, int8_t *p, void *workspace, size_t workspace_bytes, void *stream) {
  if (!plan || !tensors || !configs || !log_amp || !workspace) return vmc::fail("null argument");
  if (batch < 1) return vmc::fail("batch must be positive");
  hipStream_t st = (hipStream_t)stream;
  if (plan->dtype == VMC_F32)
    return vmc::amp_grad_impl<float>(plan, (const float *)tensors, configs, batch, log_amp, (float *)o, p,
                                     workspace, workspace_bytes, st);
  return vmc::amp_grad_impl<double>(plan, (const double *)tensors, configs, batch, log_amp, (double *)o, p,
                                    workspace, workspace_bytes, st);
}

int vmc_profile_begin(vmc_plan *plan) {
  if (!plan) return vmc::fail("plan is null");
  plan->prof = true;
  plan->prof_cls.clear();
  plan->prof_ev.clear();
  return 0;
}

int vmc_profile_end(vmc_plan *plan, double *ms, int *launches, int n_classes) {
  if (!plan) return vmc::fail("plan is null");
  for (int c = 0; c < n_classes; ++c) {
    ms[c] = 0.0;
    launches[c] = 0;
  }
  int rc = 0;
  for (size_t i = 0; i < plan->prof_cls.size(); ++i) {
    hipEvent_t a = plan->prof_ev[2 * i], b = plan->prof_ev[2 * i + 1];
    if (hipEventSynchronize(b) != hipSuccess) rc = vmc::fail("hipEventSynchronize failed");
    float t = 0.f;
    (void)hipEventElapsedTime(&t, a, b);
    const int c = plan->prof_cls[i];
    if (c < n_classes) {
      ms[c] += t;
      launches[c] += 1;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  plan->prof = false;
  plan->prof_cls.clear();
  plan->prof_ev.clear();
  return rc;
}

const char *vmc_last_error(void) { return vmc::g_last_error.c_str(); }

}  // extern "C"
