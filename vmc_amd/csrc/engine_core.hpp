#pragma once
// Batched PEPS amplitude + env-gradient engine for gfx950: plan bookkeeping,
// kernels and the per-dtype drivers. Included by engine.hip (the C ABI, no
// kernel instantiation) and by engine_dt.hip (compiled once per dtype, each
// object exporting one EngineOps table), so the four dtypes build in parallel.
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
#include <array>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace vmc {

// Records msg as the calling thread's last error and returns -1 (engine.hip).
int fail(const std::string &msg);

#define VMC_HIP(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) return ::vmc::fail(std::string(#call) + ": " + hipGetErrorString(e_)); \
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
constexpr int MAX_KEEP = 32;  // kept vectors per variational QR panel
constexpr int AMP_MAX_BOND = 512;  // bonds of the final boundary contracted by k_amp

// Autotuned choice of one launch shape (vmc_plan_autotune): the kernel variant
// or tile pair and the measured average launch times it was picked on.
enum TuneKind { TK_TRUNC = 1, TK_THETA = 2 };
constexpr int TUNE_KEY = 10;  // kind + 9 shape fields
struct TuneVal {
  int choice;
  int t_choice_ns, t_default_ns;
};
struct Tuner;

// Geometry of one lattice site inside the row environment pass.
struct ColGeo {
  int64_t top_off, bot_off, site_off, site_slice, o_off;
  int nc;          // slices per physical state: physical state p of a site with slice s is the
                   // slice p * nc + s % nc, its operator index s / nc (1 except gauge-invariant)
  int tl, tp, tr;  // top boundary tensor (a, u, b) before this row
  int bl, bp, br;  // bottom boundary tensor (e, v, f) below this row
  int su, sd, sl, sr;
};

}  // namespace vmc

struct vmc_plan {
  int R, C, d, D, chi, strategy, dtype;
  int n_sweeps = 0;       // Variational: ALS sweeps after the QR initialisation
  // Gauge-invariant PEPS (vmc_gi_plan_create): Z_N links, per-site slice counts d * nc
  bool gi = false;
  int gi_N = 0, gi_Qx = 0;
  std::vector<int> nsl;             // slices per site (d for a standard plan)
  int32_t *gi_charge = nullptr;     // device: charge of each matter state
  int32_t *gi_ctab = nullptr;       // device: (N, gi_maxdeg) matter states of each charge
  int32_t *gi_cdeg = nullptr;       // device: matter states per charge
  int gi_maxdeg = 1;
  bool gi_all_charges = false;      // every charge 0..N-1 has a matter state
  bool gi_has_site = false;         // diagonal terms on matter states (vmc_plan_set_terms)
  double *gi_plaq = nullptr;        // device: (R-1)*(C-1) plaquette coefficients, NaN = none
  double *gi_link = nullptr;        // device: (n_links, N) link-diagonal table
  bool gi_has_plaq = false, gi_has_link = false;
  int32_t *gi_deg = nullptr;        // device: degeneracy_per_charge (vmc_gi_plan_set_degeneracy)
  bool gi_masked = false;           // some charge sector is below dmax: slices carry charge masks
  std::vector<char> gi_plaq_row;    // row pair r has plaquette terms
  // Blockade PEPS (vmc_bl_plan_create): occupations select the slices, charge
  // degeneracies D0 / D1 of the sectors k = 0 / 1 (D = max of the two)
  bool bl = false;
  int bl_D0 = 0, bl_D1 = 0;
  int64_t vres_cap = 0;   // Variational: elements of the QR-initialised result per sample
  int64_t nparams, nsliced;
  std::vector<vmc::SiteInfo> site;
  std::vector<std::vector<vmc::Dims3>> top;      // [R+1][C]: boundary before row s
  std::vector<std::vector<int64_t>> top_off;     // per-sample element offsets
  int64_t top_size;
  std::vector<std::vector<vmc::Dims3>> bot;      // [R][C]: bottom boundary used by row r
  std::vector<std::vector<int64_t>> bot_off;
  int64_t bot_cap;
  int64_t mat_cap, carry_cap, scratch_cap, env_cap;
  int64_t gtr_cap = 0;    // FP64 scratch elements per sample of the generic truncation kernel
  void *geo_dev = nullptr;  // vmc::ColGeo[R * C], uploaded once at plan creation
  // Per-chain sampler cache: bottom boundary of every row (row_base[r] + bot_off[r][c]).
  std::vector<int64_t> row_base;
  int64_t cache_elems = 0;
  int64_t cap2 = 0, vcap = 0;  // 2-row environment and vertical-energy scratch per sample
  // Local operator tables (vmc_plan_set_terms)
  void *terms_i = nullptr, *terms_d = nullptr;
  int n_diag = 0;
  int64_t n_ops = 0;      // real parts first, then as many imaginary parts
  bool has_terms = false, has_one = false, has_h = false, has_v = false;
  std::vector<char> v_row;  // row r has vertical terms towards r+1
  // Kernel choices per launch shape, recorded by vmc_plan_autotune or
  // vmc_plan_set_tuning; shapes without an entry take the default kernel.
  std::map<std::array<int, vmc::TUNE_KEY>, vmc::TuneVal> tune;
  int64_t gtr_base = 0;           // gtr_cap of the untuned plan
  vmc::Tuner *tuner = nullptr;    // non-null only inside vmc_plan_autotune
  // Optional per-kernel-class timing with HIP events (vmc_profile_begin/end).
  bool prof = false;
  std::vector<int> prof_cls;
  std::vector<hipEvent_t> prof_ev;  // start/stop pairs
};

namespace vmc {

// ------------------------------------------------------------------------- //
// Static shape bookkeeping (zip-up bond dimensions depend only on shapes)
// ------------------------------------------------------------------------- //

// Vectors kept at a cut of an m x n theta matrix: zip-up keeps
// min(chi, m, n) singular vectors (strategy.py:143-145); density-matrix keeps
// min(chi, n) eigenvectors of the n x n theta^H theta (strategy.py:169-171),
// which can exceed the rank m of theta.
static int keep_k(const vmc_plan &p, int m, int n) {
  return p.strategy == VMC_DENSITY_MATRIX ? std::min(p.chi, n) : std::min(p.chi, std::min(m, n));
}

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
    const int k = keep_k(p, left * q, ncols);
    out[c] = {left, q, k};
    carry_k = k;
    if (mats_m) mats_m->push_back(left * q);
    if (mats_n) mats_n->push_back(ncols);
  }
  return out;
}

static int64_t dsize(const Dims3 &d) { return (int64_t)d.l * d.p * d.r; }

// Bond bookkeeping of one variational row application
// (src/vmc/peps/common/strategy.py:179-427). The QR initialisation has the
// zip-up shapes. A left-to-right pass keeps the bonds (A_c is (rl q) x
// min(rl q, rr)); a right-to-left pass shrinks each bond to
// min(left bond, q * right bond) of its LQ factor, so the result's bonds after
// the first sweep are the exact ranks from the right capped by the
// initialisation.
struct VarTrace {
  std::vector<Dims3> init;
  std::vector<std::vector<int>> lr_rl;  // [sweep][c]: left bond of L_c (and of the optimal tensor)
  std::vector<std::vector<int>> lr_rr;  // [sweep][c]: right bond of R_c in the left-to-right pass
  std::vector<std::vector<int>> rl_rr;  // [sweep][c]: right bond of R_c in the right-to-left pass
  std::vector<Dims3> out;
};

static VarTrace var_trace(const std::vector<Dims3> &in, const vmc_plan &p, int row, bool from_below,
                          std::vector<int> *mats_m = nullptr, std::vector<int> *mats_n = nullptr) {
  VarTrace t;
  t.init = zipup_dims(in, p, row, from_below, mats_m, mats_n);
  const int C = p.C;
  std::vector<int> rb(C), q(C);
  for (int c = 0; c < C; ++c) {
    rb[c] = t.init[c].r;
    q[c] = t.init[c].p;
  }
  for (int s = 0; s < p.n_sweeps; ++s) {
    std::vector<int> lrl(C), lrr(C), rrr(C);
    int rl = t.init[0].l;
    for (int c = 0; c < C; ++c) {
      lrl[c] = rl;
      lrr[c] = rb[c];
      if (c < C - 1) rl = std::min(rl * q[c], rb[c]);
    }
    int rr = rb[C - 1];
    for (int c = C - 1; c >= 0; --c) {
      rrr[c] = rr;
      if (c > 0) {
        rr = std::min(lrl[c], q[c] * rr);
        rb[c - 1] = rr;
      }
    }
    t.lr_rl.push_back(lrl);
    t.lr_rr.push_back(lrr);
    t.rl_rr.push_back(rrr);
  }
  if (p.n_sweeps == 0) {
    t.out = t.init;
  } else {
    t.out.resize(C);
    for (int c = 0; c < C; ++c) t.out[c] = {c == 0 ? t.init[0].l : rb[c - 1], q[c], rb[c]};
  }
  return t;
}

// Boundary dimensions after one row application with the plan's strategy.
static std::vector<Dims3> row_dims(const std::vector<Dims3> &in, const vmc_plan &p, int row, bool from_below,
                                   std::vector<int> *mats_m = nullptr, std::vector<int> *mats_n = nullptr) {
  if (p.strategy == VMC_VARIATIONAL) return var_trace(in, p, row, from_below, mats_m, mats_n).out;
  return zipup_dims(in, p, row, from_below, mats_m, mats_n);
}

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
      off += s.slice * (p->nsl.empty() ? p->d : p->nsl[r * C + c]);
      ooff += s.slice;
    }
  p->nparams = off;
  p->nsliced = ooff;

  std::vector<int> mm, mn;
  // top sweep
  p->top.assign(R + 1, std::vector<Dims3>(C, Dims3{1, 1, 1}));
  for (int r = 0; r < R; ++r) p->top[r + 1] = row_dims(p->top[r], *p, r, false, &mm, &mn);
  p->top_off.assign(R + 1, std::vector<int64_t>(C, 0));
  int64_t t = 0;
  for (int s = 0; s <= R; ++s)
    for (int c = 0; c < C; ++c) {
      p->top_off[s][c] = t;
      t += dsize(p->top[s][c]);
    }
  p->top_size = t;
  for (int c = 0; c < C; ++c)
    if (p->top[R][c].l > AMP_MAX_BOND || p->top[R][c].r > AMP_MAX_BOND)
      return fail("final boundary bond above " + std::to_string(AMP_MAX_BOND));
  // bottom sweep: bot[R-1] = ones, bot[r] = apply(bot[r+1], row r+1 from below)
  p->bot.assign(R, std::vector<Dims3>(C, Dims3{1, 1, 1}));
  for (int r = R - 2; r >= 0; --r) p->bot[r] = row_dims(p->bot[r + 1], *p, r + 1, true, &mm, &mn);
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
  const bool cplx_plan = p->dtype == VMC_C64 || p->dtype == VMC_C128;
  const bool var = p->strategy == VMC_VARIATIONAL;
  // the complex Hermitian-Gram kernel holds 64x64 theta matrices; the
  // variational QR panels take up to 128 rows for every dtype
  for (size_t i = 0; i < mm.size(); ++i) {
    const int kk = keep_k(*p, mm[i], mn[i]);
    mat_cap = std::max<int64_t>(mat_cap, (int64_t)mm[i] * mn[i]);
    carry_cap = std::max<int64_t>(carry_cap, (int64_t)kk * mn[i]);
    if (var) {
      // QR initialisation of theta[:, :k]: register panel (128 x 32, theta width <= 128)
      // or the generic panel in global scratch (512 x 128)
      if (mm[i] <= VQR_M && kk <= VQR_N && mn[i] <= 128) continue;
      if (mm[i] > VQG_M || kk > VQG_N)
        return fail("variational QR initialisation of a " + std::to_string(mm[i]) + "x" + std::to_string(kk) +
                    " panel exceeds the " + std::to_string(VQG_M) + "x" + std::to_string(VQG_N) + " limit");
      p->gtr_cap = std::max<int64_t>(p->gtr_cap, 2 * (int64_t)mm[i] * kk);
      continue;
    }
    // the factored matrix: theta (m x n), or theta^T for a density-matrix cut
    const int fm = p->strategy == VMC_DENSITY_MATRIX ? mn[i] : mm[i];
    const int fn = p->strategy == VMC_DENSITY_MATRIX ? mm[i] : mn[i];
    if (trunc_fast_shape(cplx_plan, fm, fn, kk)) continue;
    if (fm > TRUNC_G_MAX_M || kk > TRUNC_G_MAX_K)
      return fail("truncation of a " + std::to_string(fm) + "x" + std::to_string(fn) + " matrix keeping " +
                  std::to_string(kk) + " vectors exceeds the truncation kernels (rows <= " +
                  std::to_string(TRUNC_G_MAX_M) + ", kept <= " + std::to_string(TRUNC_G_MAX_K) + ")");
    if (kk < fm) p->gtr_cap = std::max<int64_t>(p->gtr_cap, trunc_g_need(fm, kk));
  }
  // last-column thetas and intermediates: bound every einsum intermediate by
  // (max boundary bond)^2 * D^2 and environments by bond * D * bond.
  int maxb = 1;
  for (auto &row : p->top)
    for (auto &d3 : row) maxb = std::max({maxb, d3.l, d3.r});
  for (auto &row : p->bot)
    for (auto &d3 : row) maxb = std::max({maxb, d3.l, d3.r});
  if (var) {
    // QR-initialised results keep the zip-up bonds until the first RL pass
    int64_t vres = 1;
    auto visit = [&](const std::vector<Dims3> &in, int row, bool below) -> int {
      const VarTrace t = var_trace(in, *p, row, below);
      int64_t tot = 0;
      for (int c = 0; c < C; ++c) {
        tot += dsize(t.init[c]);
        maxb = std::max({maxb, t.init[c].l, t.init[c].r});
      }
      vres = std::max(vres, tot);
      // QR (left-to-right) and LQ (right-to-left) panels: m x n with m rows of the
      // factored matrix; beyond the 128 x 32 register tile they run in global
      // scratch (Q, and the FP64 copy of an FP32 panel)
      auto panel = [&](int pm, int pn, const char *what) -> int {
        if (pm <= VQR_M && pn <= VQR_N) return 0;
        if (pm > VQG_M || pn > VQG_N)
          return fail(std::string("variational ") + what + " panel " + std::to_string(pm) + "x" + std::to_string(pn) +
                      " exceeds the " + std::to_string(VQG_M) + "x" + std::to_string(VQG_N) + " limit");
        p->gtr_cap = std::max<int64_t>(p->gtr_cap, (int64_t)pm * std::min(pm, pn) + (int64_t)pm * pn);
        return 0;
      };
      for (size_t s = 0; s < t.lr_rl.size(); ++s)
        for (int c = 0; c < C; ++c) {
          const int q = t.init[c].p;
          if (c < C - 1 && panel(t.lr_rl[s][c] * q, t.lr_rr[s][c], "QR")) return -1;
          if (c > 0 && panel(q * t.rl_rr[s][c], t.lr_rl[s][c], "LQ")) return -1;
        }
      return 0;
    };
    for (int r = 0; r < R; ++r)
      if (visit(p->top[r], r, false)) return -1;
    for (int r = R - 2; r >= 0; --r)
      if (visit(p->bot[r + 1], r + 1, true)) return -1;
    p->vres_cap = vres;
  }
  scr = (int64_t)maxb * maxb * D * D * std::max(1, p->d);
  env = (int64_t)maxb * maxb * D;
  p->gtr_base = p->gtr_cap;
  p->mat_cap = mat_cap;
  p->carry_cap = carry_cap;
  p->scratch_cap = std::max(scr, mat_cap);
  p->env_cap = env;
  p->row_base.assign(R, 0);
  int64_t rb = 0;
  for (int r = 0; r < R; ++r) {
    p->row_base[r] = rb;
    for (int c = 0; c < C; ++c) rb += dsize(p->bot[r][c]);
  }
  p->cache_elems = rb;
  p->cap2 = (int64_t)maxb * maxb * D * D;
  // vertical-energy scratch: the largest pairwise intermediate is
  // d * D^3 * maxb^2 (E3), plus E4 and the d x d amplitudes behind E2.
  p->vcap = (int64_t)std::max(1, p->d) * D * D * D * maxb * maxb + 2 * (int64_t)p->d * D * D * D * D + 64;
  p->v_row.assign(R, 0);

  std::vector<ColGeo> geo((size_t)R * C);
  for (int r = 0; r < R; ++r)
    for (int c = 0; c < C; ++c) {
      const SiteInfo &s = p->site[r * C + c];
      ColGeo &G = geo[r * C + c];
      G.top_off = p->top_off[r][c];
      G.bot_off = p->bot_off[r][c];
      G.site_off = s.off;
      G.site_slice = s.slice;
      G.nc = p->gi || p->bl ? p->nsl[r * C + c] / p->d : 1;
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

}  // namespace vmc

#include "variational.hpp"

namespace vmc {

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
  int tile1, tile2;  // GemmTile of carry.m and of t1.w (GEMM_TILE_AUTO when zero)
};

// At most 128 registers (4 waves per SIMD, four workgroups per CU): the FP64
// instance fits without spills and runs 36.4 -> 34.9 ms per step at 10x10 D=4 chi=32.
template <typename T>
__global__ __launch_bounds__(256, 4) void k_theta(ThetaArgs<T> a) {
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
    gemm2<T, true>(carry, g2(k, (int64_t)a.Dl * wl, wl, 1), g2(a.Dl, wl),
             m, g2(a.Dl, PD), g2(PD, 1),
             t1w, g2(k, (int64_t)wl * PD, wl, PD), g2(PD, 1), T(1), a.tile1);
    __syncthreads();
    t1 = t1w;
  }
  const int PD = a.P * a.Dr;
  gemm2<T, true>(t1, g2(k, (int64_t)wl * PD, a.Dr, 1), g2(wl, PD, a.P, a.Dr),
           w, g2(wl, a.w_sl, a.P, a.w_sp), g2(a.Q, a.w_sq, a.WR, a.w_sr),
           out, g2(k, (int64_t)a.Q * a.Dr * a.WR, a.Dr, a.WR), g2(a.Q, (int64_t)a.Dr * a.WR, a.WR, 1), T(1),
           a.tile2);
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
  using Acc = typename Scalar<T>::acc;
  __shared__ Acc st[2][AMP_MAX_BOND];
  const int b = blockIdx.x, t = threadIdx.x;
  const T *bnd = a.bnd + (int64_t)b * a.bnd_stride;
  double lsc = a.logscale[(int64_t)b * a.log_stride];
  int cur = 0;
  if (t == 0) st[0][0] = Acc(1.0);
  __syncthreads();
  for (int c = 0; c < a.C; ++c) {
    const T *M = bnd + a.off[c];
    const int L = a.l[c], Rr = a.r[c];
    double mx = 0.0;
    for (int j = t; j < Rr; j += 64) {
      Acc s = Acc(0.0);
      for (int i = 0; i < L; ++i) s = fma(st[cur][i], to_acc(M[(int64_t)i * Rr + j]), s);
      st[cur ^ 1][j] = s;
      mx = fmax(mx, acc_abs(s));
    }
    mx = wave_max(mx);
    __syncthreads();
    cur ^= 1;
    if (mx > 0.0) {
      for (int j = t; j < Rr; j += 64) st[cur][j] = (1.0 / mx) * st[cur][j];
      lsc += log(mx);
    }
    __syncthreads();
  }
  if (t == 0) {
    const Acc s = st[cur][0];
    a.log_amp[2 * (int64_t)b] = log(acc_abs(s)) + lsc;
    a.log_amp[2 * (int64_t)b + 1] = acc_arg(s);
  }
}

// Local operator tables on the device (vmc_plan_set_terms). Operators are
// dense row-major <out|op|in> matrices; the energy uses op[:, s] as the
// reference does (src/vmc/peps/common/energy.py:88, 179, 200).
struct TermsDev {
  const int *one_idx;     // [R*C] offset into ops of the summed one-site op, or -1
  const int *h_idx;       // [R*C] bond (r,c)-(r,c+1): summed d^2 x d^2 op, or -1
  const int *v_idx;       // [R*C] bond (r,c)-(r+1,c): summed d^2 x d^2 op, or -1
  const int *diag_sites;  // [n_diag * 2] flat site ids (second = -1 for one site)
  const int *diag_off;    // [n_diag] offset of the diagonal table in ops
  int n_diag;
  const double *ops;
  const double *ops_im;   // imaginary parts, same offsets (zeros for real terms)
};

// Operator entry i as the accumulator type: the complex coefficient on complex
// plans; real plans take only real operators (vmc_plan_set_terms_c refuses others).
template <typename Acc>
__device__ inline Acc op_at(const TermsDev &t, int i);
template <>
__device__ inline double op_at<double>(const TermsDev &t, int i) {
  return t.ops[i];
}
template <>
__device__ inline cf64 op_at<cf64>(const TermsDev &t, int i) {
  return cf64{t.ops[i], t.ops_im[i]};
}

// Row environment pass: right environments (src/vmc/peps/common/contraction.py:68-86)
// and then the left-to-right env-gradient sweep (src/vmc/peps/common/energy.py:98-120,
// 259-273) with the one-site and horizontal local-energy terms of
// _compute_all_env_grads_and_energy (energy.py:123-205); one workgroup per
// sample looping over the columns of the row.
template <typename T>
struct RowArgs {
  int C;
  const T *top;  int64_t top_stride;  // top boundary region base (per-sample stride)
  const T *bot;  int64_t bot_stride;
  const T *sites;                      // parameter vector
  int32_t *cfg; int n_sites; int row;
  T *renv; int64_t renv_stride; int64_t env_cap;
  T *lenv; int64_t lenv_stride;
  T *s1, *s2, *s3, *s4, *s5; int64_t s_stride;
  const double *top_log; int64_t top_log_stride;
  const double *bot_log; int64_t bot_log_stride;
  const double *log_amp;  // (batch, 2); read by the estimate pass, written by none
  T *o; int64_t o_stride;
  int8_t *p;
  const ColGeo *geo;  // this row's C entries
  // energies
  int d;
  bool energy;
  TermsDev terms;
  double *eloc;
  // sampling
  uint64_t seed;
  int sweep;
  double *log_amp_out;  // sampler: log|psi| of the accepted configuration (unused)
};

template <typename T>
__device__ inline void normalise(T *x, int64_t n, double *red, double *logacc) {
  double ss = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) ss += abs2(x[i]);
  ss = block_sum(ss, red);
  const double nrm = sqrt(ss);
  if (nrm > 0.0) {
    const T sc = scale_of<T>(1.0 / nrm);
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) x[i] *= sc;
    *logacc += log(nrm);
  } else {
    *logacc += -INFINITY;
  }
  __syncthreads();
}

// sum_i x_i y_i (no conjugation) accumulated in FP64 / complex FP64.
template <typename T>
__device__ inline typename Scalar<T>::acc block_dot(const T *x, const T *y, int64_t n, double *red) {
  using Acc = typename Scalar<T>::acc;
  Acc s = Acc(0.0);
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s = fma(to_acc(x[i]), to_acc(y[i]), s);
  return block_sum(s, red);
}

// eloc[b] += e (complex eloc: two doubles per sample)
__device__ inline void eloc_add(double *eloc, int b, double e) { eloc[b] += e; }
__device__ inline void eloc_add(double *eloc, int b, cf64 e) {
  eloc[2 * b] += e.x;
  eloc[2 * b + 1] += e.y;
}

// R[c] for c = C-1..0 into renv, with log scales in rlog (shared).
template <typename T, bool MF = false>
__device__ inline void right_envs(const RowArgs<T> &a, const T *top, const T *bot, T *renv, T *s1, T *s2,
                                  const int32_t *cfg, double *rlog, double *red) {
  const int C = a.C;
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
    gemm2<T, MF>(tp, g2(ta * tu, tb), g2(tb, 1), Rin, g2(tb, (int64_t)wr * bf), g2(wr * bf, 1),
             s1, g2(ta * tu, (int64_t)wr * bf), g2(wr * bf, 1));
    __syncthreads();
    // X2[c,v,a,f] = mpo[c,d,u,v] X1[a,u,d,f]   (mpo = w[l, r, u, dn])
    gemm2<T, MF>(w, g2(wl, s_l, wd, s_dn), g2(wr, s_r, wu, s_u),
             s1, g2(wr, bf, tu, (int64_t)wr * bf), g2(ta, (int64_t)tu * wr * bf, bf, 1),
             s2, g2(wl, (int64_t)wd * ta * bf, wd, (int64_t)ta * bf), g2(ta, bf, bf, 1));
    __syncthreads();
    // R[a,c,e] = X2[c,v,a,f] bot[e,v,f]
    gemm2<T, MF>(s2, g2(ta, bf, wl, (int64_t)wd * ta * bf), g2(wd, (int64_t)ta * bf, bf, 1),
             bp, g2(bv, bf, bf, 1), g2(be, (int64_t)bv * bf),
             Rout, g2(ta, (int64_t)wl * be, wl, be), g2(be, 1));
    __syncthreads();
    double lg = rlog[cc];
    normalise(Rout, (int64_t)ta * wl * be, red, &lg);
    if (threadIdx.x == 0) rlog[c] = lg;
    __syncthreads();
  }
}

// Y1[c,e,u,b] = L[a,c,e] top[a,u,b] and Y2[e,v,b,d] = bot[e,v,f] R[b,d,f].
template <typename T, bool MF = false>
__device__ inline void env_halves(const ColGeo &G, const T *lenv, const T *tp, const T *bp, const T *Rc, T *y1,
                                  T *y2) {
  const int ta = G.tl, tu = G.tp, tb = G.tr;
  const int be = G.bl, bv = G.bp, bf = G.br;
  const int wl = G.sl, wr = G.sr;
  gemm2<T, MF>(lenv, g2(wl, be, be, 1), g2(ta, (int64_t)wl * be), tp, g2(ta, (int64_t)tu * tb), g2(tu * tb, 1),
           y1, g2(wl * be, (int64_t)tu * tb), g2(tu * tb, 1));
  gemm2<T, MF>(bp, g2(be * bv, bf), g2(bf, 1), Rc, g2(bf, 1), g2(tb, (int64_t)wr * bf, wr, bf),
           y2, g2(be * bv, (int64_t)tb * wr), g2(tb * wr, 1));
}

// grad[c,u,v,d] = Y1[c,e,u,b] Y2[e,v,b,d], stored as (up, down, left, right) * alpha.
template <typename T, bool MF = false>
__device__ inline void env_grad(const ColGeo &G, const T *y1, const T *y2, T *out, T alpha) {
  const int tu = G.tp, tb = G.tr;
  const int be = G.bl, bv = G.bp;
  const int wl = G.sl, wr = G.sr;
  // a 16 x 16 output at D = 4 with K = be * tb: split K over the waves
  const int tile = wl * tu <= 16 && bv * wr <= 16 ? GEMM_TILE_SPLITK : GEMM_TILE_AUTO;
  gemm2<T, MF>(y1, g2(wl, (int64_t)be * tu * tb, tu, tb), g2(be, (int64_t)tu * tb, tb, 1),
           y2, g2(be, (int64_t)bv * tb * wr, tb, wr), g2(bv, (int64_t)tb * wr, wr, 1),
           out, g2(wl, wr, tu, (int64_t)bv * wl * wr), g2(bv, (int64_t)wl * wr, wr, 1), alpha, tile);
}

// L <- bot . (mpo . Y1): the left-environment update (energy.py:201-205).
template <typename T, bool MF = false>
__device__ inline void left_update(const ColGeo &G, const T *w, const T *y1, const T *bp, T *z2, T *lenv) {
  const int tu = G.tp, tb = G.tr;
  const int be = G.bl, bv = G.bp, bf = G.br;
  const int wl = G.sl, wr = G.sr, wu = G.su, wd = G.sd;
  const int64_t s_u = (int64_t)wd * wl * wr, s_dn = (int64_t)wl * wr, s_l = wr, s_r = 1;
  // Z2[d,v,e,b] = mpo[c,d,u,v] Y1[c,e,u,b]
  gemm2<T, MF>(w, g2(wr, s_r, wd, s_dn), g2(wl, s_l, wu, s_u),
           y1, g2(wl, (int64_t)be * tu * tb, tu, tb), g2(be, (int64_t)tu * tb, tb, 1),
           z2, g2(wr, (int64_t)wd * be * tb, wd, (int64_t)be * tb), g2(be, tb, tb, 1));
  __syncthreads();
  // L[b,d,f] = bot[e,v,f] Z2[d,v,e,b]
  gemm2<T, MF>(z2, g2(tb, 1, wr, (int64_t)wd * be * tb), g2(be, tb, wd, (int64_t)be * tb),
           bp, g2(be, (int64_t)bv * bf, bv, bf), g2(bf, 1),
           lenv, g2(tb, (int64_t)wr * bf, wr, bf), g2(bf, 1));
  __syncthreads();
}

// ENERGY selects the variant with one-site / horizontal local energies; the
// gradient-only variant (amp+grad, GI) carries none of their registers. It runs
// 4 waves per SIMD in FP32 (measured 51.0 -> 48.9 ms per step at 10x10 D=4
// chi=32) and 3 in FP64, where the 128-register bound spills (77.0 -> 85.7 ms).
template <typename T, bool ENERGY>
__global__ __launch_bounds__(256, ENERGY ? 2 : (std::is_same<T, float>::value ? 4 : 3)) void k_row_grad(RowArgs<T> a) {
  using Acc = typename Scalar<T>::acc;
  __shared__ double red[32];
  __shared__ double rlog[64];
  const int b = blockIdx.x;
  const int C = a.C;
  const T *top = a.top + (int64_t)b * a.top_stride;
  const T *bot = a.bot + (int64_t)b * a.bot_stride;
  T *renv = a.renv + (int64_t)b * a.renv_stride;
  T *lenv = a.lenv + (int64_t)b * a.lenv_stride;
  T *s1 = a.s1 + (int64_t)b * a.s_stride;
  T *s2 = a.s2 + (int64_t)b * a.s_stride;
  T *s3 = a.s3 + (int64_t)b * a.s_stride;
  T *s4 = a.s4 + (int64_t)b * a.s_stride;
  T *s5 = a.s5 + (int64_t)b * a.s_stride;
  const int32_t *cfg = a.cfg + (int64_t)b * a.n_sites + (int64_t)a.row * C;
  const int d = a.d;

  right_envs<T, true>(a, top, bot, renv, s1, s2, cfg, rlog, red);

  // ---- left-to-right env gradients and local energies ----
  if (threadIdx.x == 0) lenv[0] = T(1);
  double llog = 0.0;
  Acc e_row = Acc(0.0);
  __syncthreads();
  const double base = a.top_log[(int64_t)b * a.top_log_stride] + a.bot_log[(int64_t)b * a.bot_log_stride] -
                      a.log_amp[2 * (int64_t)b];
  const double phase = -a.log_amp[2 * (int64_t)b + 1];  // 1/psi = exp(-log|psi| - i arg psi)
  for (int c = 0; c < C; ++c) {
    const ColGeo G = a.geo[c];
    const int site = a.row * C + c;
    const T *tp = top + G.top_off;
    const T *bp = bot + G.bot_off;
    const T *A = a.sites + G.site_off;  // all physical slices
    const T *w = A + (int64_t)cfg[c] * G.site_slice;
    const T *Rc = renv + (int64_t)c * a.env_cap;
    env_halves<T, true>(G, lenv, tp, bp, Rc, s1, s2);
    __syncthreads();
    // G / psi into o (or scratch when only energies are wanted)
    T *gdst = a.o ? a.o + (int64_t)b * a.o_stride + G.o_off : s3;
    env_grad<T, true>(G, s1, s2, gdst, from_acc<T>(polar_acc<T>(exp(base + llog + rlog[c]), phase)));
    if (a.p) {
      int8_t *pp = a.p + (int64_t)b * a.o_stride + G.o_off;
      for (int64_t i = threadIdx.x; i < G.site_slice; i += blockDim.x) pp[i] = (int8_t)cfg[c];
    }
    __syncthreads();
    if (ENERGY && a.energy) {
      const int oi = a.terms.one_idx[site];
      const int sg0 = cfg[c] / G.nc, cm0 = cfg[c] - sg0 * G.nc;
      if (oi >= 0) {
        // amps_site[p] / psi = <A[p], G / psi>   (energy.py:175-179)
        for (int q = 0; q < d; ++q) {
          const Acc amp_q = block_dot(A + ((int64_t)q * G.nc + cm0) * G.site_slice, gdst, G.site_slice, red);
          e_row += op_at<Acc>(a.terms, oi + q * d + sg0) * amp_q;
        }
      }
      const int hi = c < C - 1 ? a.terms.h_idx[site] : -1;
      if (hi >= 0) {
        // amps_edge[p, q]: the right environment of column c recomputed with the
        // physical index of site c+1 fixed to q (same pairwise steps as R[c]).
        const ColGeo G1 = a.geo[c + 1];
        const T *tp1 = top + G1.top_off;
        const T *bp1 = bot + G1.bot_off;
        const T *R1 = renv + (int64_t)(c + 1) * a.env_cap;
        const int ta = G1.tl, tu = G1.tp, tb = G1.tr;
        const int be = G1.bl, bv = G1.bp, bf = G1.br;
        const int wl = G1.sl, wr = G1.sr, wu = G1.su, wd = G1.sd;
        const int64_t s_u = (int64_t)wd * wl * wr, s_dn = (int64_t)wl * wr, s_l = wr, s_r = 1;
        gemm2<T, true>(tp1, g2(ta * tu, tb), g2(tb, 1), R1, g2(tb, (int64_t)wr * bf), g2(wr * bf, 1),
                 s4, g2(ta * tu, (int64_t)wr * bf), g2(wr * bf, 1));
        __syncthreads();
        const T alpha1 = from_acc<T>(polar_acc<T>(exp(base + llog + rlog[c + 1]), phase));
        const int sg1 = cfg[c + 1] / G1.nc, cm1 = cfg[c + 1] - sg1 * G1.nc;
        const int s01 = sg0 * d + sg1;
        for (int q = 0; q < d; ++q) {
          const T *w1 = a.sites + G1.site_off + ((int64_t)q * G1.nc + cm1) * G1.site_slice;
          gemm2<T, true>(w1, g2(wl, s_l, wd, s_dn), g2(wr, s_r, wu, s_u),
                   s4, g2(wr, bf, tu, (int64_t)wr * bf), g2(ta, (int64_t)tu * wr * bf, bf, 1),
                   s5, g2(wl, (int64_t)wd * ta * bf, wd, (int64_t)ta * bf), g2(ta, bf, bf, 1));
          __syncthreads();
          gemm2<T, true>(s5, g2(ta, bf, wl, (int64_t)wd * ta * bf), g2(wd, (int64_t)ta * bf, bf, 1),
                   bp1, g2(bv, bf, bf, 1), g2(be, (int64_t)bv * bf),
                   s3, g2(ta, (int64_t)wl * be, wl, be), g2(be, 1));
          __syncthreads();
          // Y2' = bot[c] . R'_q, G' = Y1 . Y2'
          {
            const int tbc = G.tr, bec = G.bl, bvc = G.bp, bfc = G.br, wrc = G.sr;
            gemm2<T, true>(bp, g2(bec * bvc, bfc), g2(bfc, 1), s3, g2(bfc, 1), g2(tbc, (int64_t)wrc * bfc, wrc, bfc),
                     s5, g2(bec * bvc, (int64_t)tbc * wrc), g2(tbc * wrc, 1));
          }
          __syncthreads();
          env_grad<T, true>(G, s1, s5, s2, alpha1);
          __syncthreads();
          for (int pq = 0; pq < d; ++pq) {
            const Acc amp_pq = block_dot(A + ((int64_t)pq * G.nc + cm0) * G.site_slice, s2, G.site_slice, red);
            e_row += op_at<Acc>(a.terms, hi + (pq * d + q) * d * d + s01) * amp_pq;
          }
        }
      }
    }
    if (c < C - 1) {
      left_update<T, true>(G, w, s1, bp, s2, lenv);
      normalise(lenv, (int64_t)G.tr * G.sr * G.br, red, &llog);
    }
  }
  if (ENERGY && a.energy && threadIdx.x == 0) eloc_add(a.eloc, b, e_row);
}

// Metropolis sweep over one row (src/vmc/peps/standard/kernels.py:142-204):
// right environments of the current row, then per site the env-gradient G,
// amp_flip = <A[flip], G>, accept with probability min(1, |amp_flip/amp_cur|^2)
// using a Philox draw, and the left environment with the chosen slice.
__device__ inline void philox4x32_10(uint32_t (&c)[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
  }
}

template <typename T>
__global__ __launch_bounds__(256, 3) void k_row_sample(RowArgs<T> a) {
  __shared__ double red[32];
  __shared__ double rlog[64];
  const int b = blockIdx.x;
  const int C = a.C;
  const T *top = a.top + (int64_t)b * a.top_stride;
  const T *bot = a.bot + (int64_t)b * a.bot_stride;
  T *renv = a.renv + (int64_t)b * a.renv_stride;
  T *lenv = a.lenv + (int64_t)b * a.lenv_stride;
  T *s1 = a.s1 + (int64_t)b * a.s_stride;
  T *s2 = a.s2 + (int64_t)b * a.s_stride;
  T *s3 = a.s3 + (int64_t)b * a.s_stride;
  int32_t *cfg = a.cfg + (int64_t)b * a.n_sites + (int64_t)a.row * C;
  const int d = a.d;

  right_envs(a, top, bot, renv, s1, s2, cfg, rlog, red);
  if (threadIdx.x == 0) lenv[0] = T(1);
  double llog = 0.0;
  __syncthreads();
  const double base = a.top_log[(int64_t)b * a.top_log_stride] + a.bot_log[(int64_t)b * a.bot_log_stride];
  double lc = -INFINITY;  // log |amp_cur|
  for (int c = 0; c < C; ++c) {
    const ColGeo G = a.geo[c];
    const T *tp = top + G.top_off;
    const T *bp = bot + G.bot_off;
    const T *A = a.sites + G.site_off;
    const T *Rc = renv + (int64_t)c * a.env_cap;
    env_halves(G, lenv, tp, bp, Rc, s1, s2);
    __syncthreads();
    env_grad(G, s1, s2, s3, T(1));
    __syncthreads();
    const double S = base + llog + rlog[c];
    const int cur = cfg[c];
    uint32_t x[4] = {(uint32_t)(a.row * C + c), (uint32_t)a.sweep, (uint32_t)b, 0u};
    philox4x32_10(x, (uint32_t)(a.seed & 0xFFFFFFFFu), (uint32_t)(a.seed >> 32));
    const double u = (double)((((uint64_t)x[0]) << 21) | (x[1] >> 11)) * 0x1.0p-53;
    int flip = cur;
    if (d == 2)
      flip = 1 - cur;
    else if (d > 2)
      flip = (cur + 1 + (int)(x[2] % (uint32_t)(d - 1))) % d;
    if (c == 0) {
      const auto v = block_dot(A + (int64_t)cur * G.site_slice, s3, G.site_slice, red);
      lc = log(acc_abs(v)) + S;
    }
    const auto vf = block_dot(A + (int64_t)flip * G.site_slice, s3, G.site_slice, red);
    const double lf = log(acc_abs(vf)) + S;
    double ratio;
    if (lc > -INFINITY)
      ratio = exp(2.0 * (lf - lc));
    else
      ratio = lf > -INFINITY ? INFINITY : 0.0;
    const bool accept = u < fmin(1.0, ratio);
    const int sel = accept ? flip : cur;
    if (accept) lc = lf;
    __syncthreads();
    if (threadIdx.x == 0) cfg[c] = sel;
    if (c < C - 1) {
      left_update(G, A + (int64_t)sel * G.site_slice, s1, bp, s2, lenv);
      normalise(lenv, (int64_t)G.tr * G.sr * G.br, red, &llog);
    }
  }
}

// Diagonal terms (energy.py:150-155): one thread per sample.
static __global__ void k_diag(TermsDev t, const int32_t *cfg, int n_sites, int d, double *eloc, int batch, int cplx_out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const int32_t *s = cfg + (int64_t)b * n_sites;
  double e = 0.0, ei = 0.0;
  for (int i = 0; i < t.n_diag; ++i) {
    const int a0 = t.diag_sites[2 * i], a1 = t.diag_sites[2 * i + 1];
    int idx = s[a0];
    if (a1 >= 0) idx = idx * d + s[a1];
    e += t.ops[t.diag_off[i] + idx];
    ei += t.ops_im[t.diag_off[i] + idx];
  }
  if (cplx_out) {
    eloc[2 * b] = e;
    eloc[2 * b + 1] = ei;
  } else {
    eloc[b] = e;
  }
}

// Vertical bond energies between rows r and r+1
// (_compute_row_pair_vertical_energy, src/vmc/peps/common/energy.py:24-95):
// 2-row right environments (a, l, x, e), then left to right the 2-row left
// environment and, at columns with a vertical term, amps_edge[p, q].
// 2-row right environments of the row pair (r, r + 1) (energy.py:24-44):
// r2[c] (a, l, x, e) for c = C-1 .. 0, normalised, with log-scales in r2log.
template <typename T, bool MF = false>
__device__ inline void right2_envs(int C, const T *top, const T *bot, const T *sites, const ColGeo *geo0,
                                   const ColGeo *geo1, const int32_t *cfg0, const int32_t *cfg1, T *r2,
                                   int64_t cap2, T *v1, T *v2, T *v3, double *r2log, double *red) {
  if (threadIdx.x == 0) {
    r2[(int64_t)(C - 1) * cap2] = T(1);
    r2log[C - 1] = 0.0;
  }
  __syncthreads();
  for (int c = C - 2; c >= 0; --c) {
    const int cc = c + 1;
    const ColGeo G0 = geo0[cc];  // top tensor and row-r site
    const ColGeo G1 = geo1[cc];  // row r+1 site and bottom tensor (bl/bp/br)
    const T *tp = top + G0.top_off;
    const T *bp = bot + G1.bot_off;
    const T *w0 = sites + G0.site_off + (int64_t)cfg0[cc] * G0.site_slice;
    const T *w1 = sites + G1.site_off + (int64_t)cfg1[cc] * G1.site_slice;
    const T *Rin = r2 + (int64_t)cc * cap2;
    T *Rout = r2 + (int64_t)c * cap2;
    const int ta = G0.tl, tu = G0.tp, tb = G0.tr;
    const int be = G1.bl, bw = G1.bp, bf = G1.br;
    const int l0 = G0.sl, r0 = G0.sr, u0 = G0.su, dn0 = G0.sd;
    const int l1 = G1.sl, r1 = G1.sr, u1 = G1.su, dn1 = G1.sd;
    const int64_t su0 = (int64_t)dn0 * l0 * r0, sd0 = (int64_t)l0 * r0, sl0 = r0;
    const int64_t su1 = (int64_t)dn1 * l1 * r1, sd1 = (int64_t)l1 * r1, sl1 = r1;
    const int64_t ryf = (int64_t)r0 * r1 * bf;
    // W1[a,u,r,y,f] = top[a,u,b] R2[b,r,y,f]
    gemm2<T, MF>(tp, g2(ta * tu, tb), g2(tb, 1), Rin, g2(tb, ryf), g2((int)ryf, 1),
             v1, g2(ta * tu, ryf), g2((int)ryf, 1));
    __syncthreads();
    // W2[l,v,a,y,f] = mpo0[l,r,u,v] W1[a,u,r,y,f]
    gemm2<T, MF>(w0, g2(l0, sl0, dn0, sd0), g2(r0, 1, u0, su0),
             v1, g2(r0, (int64_t)r1 * bf, tu, ryf), g3(ta, (int64_t)tu * ryf, r1, bf, bf, 1),
             v2, g2(l0, (int64_t)dn0 * ta * r1 * bf, dn0, (int64_t)ta * r1 * bf), g3(ta, (int64_t)r1 * bf, r1, bf, bf, 1));
    __syncthreads();
    // W3[x,w,l,a,f] = mpo1[x,y,v,w] W2[l,v,a,y,f]
    gemm2<T, MF>(w1, g2(l1, sl1, dn1, sd1), g2(r1, 1, u1, su1),
             v2, g2(r1, bf, dn0, (int64_t)ta * r1 * bf), g3(l0, (int64_t)dn0 * ta * r1 * bf, ta, (int64_t)r1 * bf, bf, 1),
             v3, g2(l1, (int64_t)dn1 * l0 * ta * bf, dn1, (int64_t)l0 * ta * bf), g3(l0, (int64_t)ta * bf, ta, bf, bf, 1));
    __syncthreads();
    // R2[a,l,x,e] = bot[e,w,f] W3[x,w,l,a,f]
    gemm2<T, MF>(v3, g3(ta, bf, l0, (int64_t)ta * bf, l1, (int64_t)dn1 * l0 * ta * bf), g2(dn1, (int64_t)l0 * ta * bf, bf, 1),
             bp, g2(bw, bf, bf, 1), g2(be, (int64_t)bw * bf),
             Rout, g3(ta, (int64_t)l0 * l1 * be, l0, (int64_t)l1 * be, l1, be), g2(be, 1));
    __syncthreads();
    double lg = r2log[cc];
    normalise(Rout, (int64_t)ta * l0 * l1 * be, red, &lg);
    if (threadIdx.x == 0) r2log[c] = lg;
    __syncthreads();
  }
}

// V1[l,x,e,u,b] = L2[a,l,x,e] top[a,u,b] for column c of a row pair.
template <typename T, bool MF = false>
__device__ inline void left2_v1(const ColGeo &G0, const ColGeo &G1, const T *l2, const T *tp, T *v1) {
  const int ta = G0.tl, tu = G0.tp, tb = G0.tr;
  const int64_t lxe = (int64_t)G0.sl * G1.sl * G1.bl, ub = (int64_t)tu * tb;
  gemm2<T, MF>(l2, g2((int)lxe, 1), g2(ta, lxe), tp, g2(ta, ub), g2((int)ub, 1), v1, g2((int)lxe, ub), g2((int)ub, 1));
}

// L2out[b,r,y,f] = V1[l,x,e,u,b] mpo0[l,r,u,v] mpo1[x,y,v,w] bot[e,w,f]
// (energy.py:201-205 order for two rows); l2out may alias the L2 behind v1.
template <typename T, bool MF = false>
__device__ inline void left2_from_v1(const ColGeo &G0, const ColGeo &G1, const T *v1, const T *w0, const T *w1,
                                     const T *bp, T *v2, T *v3, T *l2out) {
  const int tu = G0.tp, tb = G0.tr;
  const int be = G1.bl, bw = G1.bp, bf = G1.br;
  const int l0 = G0.sl, r0 = G0.sr, u0 = G0.su, dn0 = G0.sd;
  const int l1 = G1.sl, r1 = G1.sr, u1 = G1.su, dn1 = G1.sd;
  const int64_t su0 = (int64_t)dn0 * l0 * r0, sd0 = (int64_t)l0 * r0, sl0 = r0;
  const int64_t su1 = (int64_t)dn1 * l1 * r1, sd1 = (int64_t)l1 * r1, sl1 = r1;
  const int64_t ub = (int64_t)tu * tb;
  // V2[r,v,x,e,b] = mpo0[l,r,u,v] V1[l,x,e,u,b]
  const int64_t xeb = (int64_t)l1 * be * tb;
  gemm2<T, MF>(w0, g2(r0, 1, dn0, sd0), g2(l0, sl0, tu, su0),
           v1, g2(l0, (int64_t)l1 * be * ub, tu, tb), g3(l1, (int64_t)be * ub, be, ub, tb, 1),
           v2, g2(r0, (int64_t)dn0 * xeb, dn0, xeb), g3(l1, (int64_t)be * tb, be, tb, tb, 1));
  __syncthreads();
  // V3[y,w,r,e,b] = mpo1[x,y,v,w] V2[r,v,x,e,b]
  const int64_t reb = (int64_t)r0 * be * tb;
  gemm2<T, MF>(w1, g2(r1, 1, dn1, sd1), g2(l1, sl1, u1, su1),
           v2, g2(l1, (int64_t)be * tb, dn0, xeb), g3(r0, (int64_t)dn0 * xeb, be, tb, tb, 1),
           v3, g2(r1, (int64_t)dn1 * reb, dn1, reb), g3(r0, (int64_t)be * tb, be, tb, tb, 1));
  __syncthreads();
  // L2[b,r,y,f] = bot[e,w,f] V3[y,w,r,e,b]
  gemm2<T, MF>(v3, g3(tb, 1, r0, (int64_t)be * tb, r1, (int64_t)dn1 * reb), g2(be, tb, dn1, reb),
           bp, g2(be, (int64_t)bw * bf, bw, bf), g2(bf, 1),
           l2out, g3(tb, (int64_t)r0 * r1 * bf, r0, (int64_t)r1 * bf, r1, bf), g2(bf, 1));
  __syncthreads();
}

template <typename T>
struct VertArgs {
  int C;
  const T *top;  int64_t top_stride;   // top boundary before row r
  const T *bot;  int64_t bot_stride;   // bottom boundary below row r+1
  const T *sites;
  const int32_t *cfg; int n_sites; int row;
  const ColGeo *geo0, *geo1;           // rows r and r+1
  T *r2; int64_t r2_stride; int64_t cap2;
  T *l2; int64_t l2_stride;
  T *v1, *v2, *v3; int64_t v_stride;
  const double *top_log; int64_t top_log_stride;
  const double *bot_log; int64_t bot_log_stride;
  const double *log_amp;
  int d;
  TermsDev terms;
  double *eloc;
};

template <typename T>
__global__ __launch_bounds__(256, 3) void k_row_vertical(VertArgs<T> a) {
  using Acc = typename Scalar<T>::acc;
  __shared__ double red[32];
  __shared__ double r2log[64];
  const int b = blockIdx.x;
  const int C = a.C, d = a.d;
  const T *top = a.top + (int64_t)b * a.top_stride;
  const T *bot = a.bot + (int64_t)b * a.bot_stride;
  T *r2 = a.r2 + (int64_t)b * a.r2_stride;
  T *l2 = a.l2 + (int64_t)b * a.l2_stride;
  T *v1 = a.v1 + (int64_t)b * a.v_stride;
  T *v2 = a.v2 + (int64_t)b * a.v_stride;
  T *v3 = a.v3 + (int64_t)b * a.v_stride;
  const int32_t *cfg0 = a.cfg + (int64_t)b * a.n_sites + (int64_t)a.row * C;
  const int32_t *cfg1 = cfg0 + C;

  // ---- 2-row right environments R2[c] (a, l, x, e) ----
  right2_envs<T, true>(C, top, bot, a.sites, a.geo0, a.geo1, cfg0, cfg1, r2, a.cap2, v1, v2, v3, r2log, red);

  // ---- left to right ----
  if (threadIdx.x == 0) l2[0] = T(1);
  double llog = 0.0;
  Acc e_row = Acc(0.0);
  __syncthreads();
  const double base = a.top_log[(int64_t)b * a.top_log_stride] + a.bot_log[(int64_t)b * a.bot_log_stride] -
                      a.log_amp[2 * (int64_t)b];
  const double phase = -a.log_amp[2 * (int64_t)b + 1];
  for (int c = 0; c < C; ++c) {
    const ColGeo G0 = a.geo0[c];
    const ColGeo G1 = a.geo1[c];
    const T *tp = top + G0.top_off;
    const T *bp = bot + G1.bot_off;
    const T *A0 = a.sites + G0.site_off;
    const T *A1 = a.sites + G1.site_off;
    const T *w0 = A0 + (int64_t)cfg0[c] * G0.site_slice;
    const T *w1 = A1 + (int64_t)cfg1[c] * G1.site_slice;
    // physical state p of the two sites: slices p * nc + s % nc
    const int sg0 = cfg0[c] / G0.nc, sg1 = cfg1[c] / G1.nc;
    const T *A0p = A0 + (int64_t)(cfg0[c] - sg0 * G0.nc) * G0.site_slice;
    const T *A1p = A1 + (int64_t)(cfg1[c] - sg1 * G1.nc) * G1.site_slice;
    const int64_t st0 = (int64_t)G0.nc * G0.site_slice, st1 = (int64_t)G1.nc * G1.site_slice;
    const T *Rc = r2 + (int64_t)c * a.cap2;
    const int ta = G0.tl, tu = G0.tp, tb = G0.tr;
    const int be = G1.bl, bw = G1.bp, bf = G1.br;
    const int l0 = G0.sl, r0 = G0.sr, u0 = G0.su, dn0 = G0.sd;
    const int l1 = G1.sl, r1 = G1.sr, u1 = G1.su, dn1 = G1.sd;
    const int64_t su0 = (int64_t)dn0 * l0 * r0, sd0 = (int64_t)l0 * r0, sl0 = r0;
    const int64_t su1 = (int64_t)dn1 * l1 * r1, sd1 = (int64_t)l1 * r1, sl1 = r1;
    const int64_t lxe = (int64_t)l0 * l1 * be, ub = (int64_t)tu * tb;
    // V1[l,x,e,u,b] = L2[a,l,x,e] top[a,u,b]   (= E1 of amps_edge)
    left2_v1<T, true>(G0, G1, l2, tp, v1);
    __syncthreads();
    const int vi = a.terms.v_idx[a.row * C + c];
    if (vi >= 0) {
      // E2[g,w,b,r,n] = bot[g,w,f] R2[b,r,n,f]
      const int64_t brn = (int64_t)tb * r0 * r1;
      gemm2<T, true>(bp, g2(be * bw, bf), g2(bf, 1), Rc, g2(bf, 1), g3(tb, (int64_t)r0 * r1 * bf, r0, (int64_t)r1 * bf, r1, bf),
               v2, g2(be * bw, brn), g2((int)brn, 1));
      // E3[p,v,r,m,g,b] = A0[p,u,v,l,r] E1[l,m,g,u,b]
      const int64_t mgb = (int64_t)l1 * be * tb;
      gemm2<T, true>(A0p, g3(d, st0, dn0, (int64_t)l0 * r0, r0, 1), g2(u0, su0, l0, sl0),
               v1, g2(tu, tb, l0, (int64_t)l1 * be * ub), g3(l1, (int64_t)be * ub, be, ub, tb, 1),
               v3, g3(d, (int64_t)dn0 * r0 * mgb, dn0, (int64_t)r0 * mgb, r0, mgb), g3(l1, (int64_t)be * tb, be, tb, tb, 1));
      __syncthreads();
      // E4[p,v,w,m,n] = E2[g,w,b,r,n] E3[p,v,r,m,g,b]  (into l2-free scratch: v2 tail)
      T *e4 = v2 + (int64_t)be * bw * brn;
      const int64_t wbrn = (int64_t)bw * brn;
      gemm2<T, true>(v2, g2(bw, brn, r1, 1), g3(be, wbrn, tb, (int64_t)r0 * r1, r0, r1),
               v3, g3(be, tb, tb, 1, r0, mgb), g3(d, (int64_t)dn0 * r0 * mgb, dn0, (int64_t)r0 * mgb, l1, (int64_t)be * tb),
               e4, g2(bw, (int64_t)l1 * r1, r1, 1), g3(d, (int64_t)dn0 * bw * l1 * r1, dn0, (int64_t)bw * l1 * r1, l1, r1));
      __syncthreads();
      // amps[p, q] / psi = E4[p, K] A1[q, K]
      T *amps = e4 + (int64_t)d * G1.site_slice;
      gemm2<T, true>(e4, g2(d, G1.site_slice), g2((int)G1.site_slice, 1), A1p, g2((int)G1.site_slice, 1), g2(d, st1),
               amps, g2(d, d), g2(d, 1), from_acc<T>(polar_acc<T>(exp(base + llog + r2log[c]), phase)));
      __syncthreads();
      const int s01 = sg0 * d + sg1;
      for (int pq = 0; pq < d * d; ++pq) e_row += op_at<Acc>(a.terms, vi + pq * d * d + s01) * to_acc(amps[pq]);
      __syncthreads();
    }
    if (c < C - 1) {
      left2_from_v1<T, true>(G0, G1, v1, w0, w1, bp, v2, v3, l2);
      normalise(l2, (int64_t)tb * r0 * r1 * bf, red, &llog);
    }
  }
  if (threadIdx.x == 0) eloc_add(a.eloc, b, e_row);
}

}  // namespace vmc

#include "gi.hpp"
#include "blockade.hpp"

namespace vmc {

// ------------------------------------------------------------------------- //
// Host orchestration
// ------------------------------------------------------------------------- //

struct Region {
  size_t off, bytes;
};

struct Layout {
  Region top, toplog, bot, botlog, carry, mat, s1, s2, s3, s4, s5, renv, lenv, r2, l2, v1, v2, v3, eloc, vres, vlenv;
  Region gsl, gvalid, gx, gy, gxp, gyp, gtr, bmask;
  size_t total;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t elem_size(int dtype) {
  return dtype == VMC_F32 ? 4 : (dtype == VMC_F64 || dtype == VMC_C64) ? 8 : 16;
}

static bool is_complex(int dtype) { return dtype == VMC_C64 || dtype == VMC_C128; }

static Layout make_layout(const vmc_plan *p, int batch, bool vertical) {
  const size_t es = elem_size(p->dtype);
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
  take(L.s4, B * p->scratch_cap * es);
  take(L.s5, B * p->scratch_cap * es);
  take(L.renv, B * p->C * p->env_cap * es);
  take(L.lenv, B * p->env_cap * es);
  take(L.eloc, B * 16);
  if (p->gtr_cap > 0) take(L.gtr, B * p->gtr_cap * (is_complex(p->dtype) ? 16 : 8));
  if (p->strategy == VMC_VARIATIONAL) {
    take(L.vres, B * p->vres_cap * es);
    take(L.vlenv, B * p->C * p->env_cap * es);
  }
  if ((p->bl && p->bl_D0 != p->bl_D1) || (p->gi && p->gi_masked)) take(L.bmask, (size_t)p->nparams * es);
  if (p->gi || p->bl) {
    take(L.gsl, B * p->R * p->C * 4);
    take(L.gvalid, B);
    if (vertical) {
      take(L.gx, B * p->cap2 * es);
      take(L.gy, B * p->cap2 * es);
      take(L.gxp, B * p->cap2 * es);
      take(L.gyp, B * p->cap2 * es);
    }
  }
  if (vertical) {
    take(L.r2, B * p->C * p->cap2 * es);
    take(L.l2, B * p->cap2 * es);
    take(L.v1, B * p->vcap * es);
    take(L.v2, B * p->vcap * es);
    take(L.v3, B * p->vcap * es);
  }
  L.total = off;
  return L;
}

template <typename T>
struct Bufs {
  T *top, *bot, *carry, *mat, *s1, *s2, *s3, *s4, *s5, *renv, *lenv, *r2, *l2, *v1, *v2, *v3, *vres, *vlenv;
  double *toplog, *botlog, *eloc;
  void *gtr;  // scratch of the generic truncation kernel (FP64 / complex FP64 elements)
};

template <typename T>
static Bufs<T> bufs(const Layout &L, void *ws) {
  char *base = (char *)ws;
  Bufs<T> b{};
  b.top = (T *)(base + L.top.off);
  b.toplog = (double *)(base + L.toplog.off);
  b.bot = (T *)(base + L.bot.off);
  b.botlog = (double *)(base + L.botlog.off);
  b.carry = (T *)(base + L.carry.off);
  b.mat = (T *)(base + L.mat.off);
  b.s1 = (T *)(base + L.s1.off);
  b.s2 = (T *)(base + L.s2.off);
  b.s3 = (T *)(base + L.s3.off);
  b.s4 = (T *)(base + L.s4.off);
  b.s5 = (T *)(base + L.s5.off);
  b.renv = (T *)(base + L.renv.off);
  b.lenv = (T *)(base + L.lenv.off);
  b.eloc = (double *)(base + L.eloc.off);
  b.r2 = (T *)(base + L.r2.off);
  b.l2 = (T *)(base + L.l2.off);
  b.v1 = (T *)(base + L.v1.off);
  b.v2 = (T *)(base + L.v2.off);
  b.v3 = (T *)(base + L.v3.off);
  b.vres = (T *)(base + L.vres.off);
  b.vlenv = (T *)(base + L.vlenv.off);
  b.gtr = base + L.gtr.off;
  return b;
}

// ------------------------------------------------------------------------- //
// Autotuning (vmc_plan_autotune)
// ------------------------------------------------------------------------- //
//
// A tuning pass is an ordinary amp+grad call with plan->tuner set. At the first
// launch of each truncation or theta shape it times every candidate on the live
// operands (one warm-up launch, then `reps` launches between two HIP events),
// records the fastest in plan->tune and launches it for real. Truncation
// candidates write the same outputs (u, carry) as the real launch, which follows
// them on the stream and overwrites them; their log-scale increments go to a
// throwaway buffer. Theta launches are idempotent. Generic-kernel candidates of
// register-sized shapes take their scratch from the tuner, so the tuning call
// runs in the untuned workspace; a generic choice then raises gtr_cap and with
// it the plan's workspace sizes.
struct Tuner {
  int reps = 3;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double *log = nullptr;   // throwaway log scales, one per sample
  void *gscr = nullptr;    // generic-kernel scratch for register-sized shapes
  size_t gscr_bytes = 0;
  int err = 0;             // first failure inside the pass
  std::vector<std::array<int, TUNE_KEY>> seen;  // shapes tuned in this pass
};

// A candidate must beat the default by this factor to replace it, so that
// timing noise does not flip equal kernels.
constexpr float TUNE_MARGIN = 0.99f;

static int tune_ns(float ms) { return (int)std::min(1e9, (double)ms * 1e6); }

static bool tune_first(Tuner *tu, const std::array<int, TUNE_KEY> &key) {
  if (!tu || tu->err) return false;
  if (std::find(tu->seen.begin(), tu->seen.end(), key) != tu->seen.end()) return false;
  tu->seen.push_back(key);
  return true;
}

// Average milliseconds of one launch.
template <typename F>
static float tune_time(Tuner *tu, hipStream_t st, F &&launch) {
  launch();
  (void)hipEventRecord(tu->e0, st);
  for (int i = 0; i < tu->reps; ++i) launch();
  (void)hipEventRecord(tu->e1, st);
  float ms = 0.f;
  if (hipGetLastError() != hipSuccess || hipEventSynchronize(tu->e1) != hipSuccess ||
      hipEventElapsedTime(&ms, tu->e0, tu->e1) != hipSuccess) {
    if (!tu->err) tu->err = fail("autotune: a candidate launch failed");
    return INFINITY;
  }
  return ms / (float)tu->reps;
}

static std::array<int, TUNE_KEY> trunc_key(int m, int n, int k, int dm) {
  return {TK_TRUNC, m, n, k, dm, 0, 0, 0, 0, 0};
}

// Theta shape: the per-sample work of k_theta is fixed by these fields (the
// site strides follow from them and the sweep direction, w_sp > w_sq from above).
template <typename T>
static std::array<int, TUNE_KEY> theta_key(const ThetaArgs<T> &a) {
  const bool c = a.carry != nullptr;
  return {TK_THETA, (int)c, c ? a.k : 1, a.Dl, c ? a.wl : 1, a.P, a.Dr, a.Q, a.WR, (int)(a.w_sp > a.w_sq)};
}

// The truncation kernel that may run a cut of this shape.
static bool trunc_choice_ok(const vmc_plan *p, int m, int n, int k, int dm, int v) {
  const bool cplx = is_complex(p->dtype);
  if (trunc_qr_shape(cplx, m, n, k, dm)) return false;  // the QR kernel, no choice
  const bool fast = trunc_fast_shape(cplx, m, n, k);
  if (v == TRUNC_LPE8) return fast;
  // the 16-lane multisection exists in the 512-thread kernel only, which runs
  // cuts factoring more than 64 rows (n rows on the theta^T route)
  if (v == TRUNC_LPE16) return fast && !cplx && (trunc_zipt_shape(cplx, m, n, k, dm) ? n : m) > 64;
  if (v == TRUNC_GENERIC) return k < m && m <= TRUNC_G_MAX_M && k <= TRUNC_G_MAX_K;
  return false;
}

// Generic-kernel scratch elements per sample of an m-row cut keeping k.
static int64_t gtr_need(int m, int k) { return trunc_g_need(m, k); }

// gtr_cap = untuned need + every register-sized shape tuned to the generic kernel.
static void refresh_gtr(vmc_plan *p) {
  int64_t g = p->gtr_base;
  for (const auto &e : p->tune)
    if (e.first[0] == TK_TRUNC && e.second.choice == TRUNC_GENERIC) g = std::max(g, gtr_need(e.first[1], e.first[3]));
  p->gtr_cap = g;
}

template <typename T>
static void launch_theta(vmc_plan *p, ThetaArgs<T> a, int batch, hipStream_t st) {
  const auto key = theta_key(a);
  Tuner *tu = p->tuner;
  if (tune_first(tu, key)) {
    auto run = [&](int t1, int t2) {
      ThetaArgs<T> b = a;
      b.tile1 = t1;
      b.tile2 = t2;
      return tune_time(tu, st, [&] { hipLaunchKernelGGL(k_theta<T>, dim3(batch), dim3(256), 0, st, b); });
    };
    // coordinate search: the first product's tile (when there is a carry), then the second's
    const float t0 = run(GEMM_TILE_AUTO, GEMM_TILE_AUTO);
    float best = t0;
    int b1 = GEMM_TILE_AUTO, b2 = GEMM_TILE_AUTO;
    if (a.carry)
      for (int t1 = GEMM_TILE_WIDE; t1 <= GEMM_TILE_SQUARE; ++t1) {
        const float t = run(t1, GEMM_TILE_AUTO);
        if (t < TUNE_MARGIN * best) best = t, b1 = t1;
      }
    for (int t2 = GEMM_TILE_WIDE; t2 <= GEMM_TILE_SQUARE; ++t2) {
      const float t = run(b1, t2);
      if (t < TUNE_MARGIN * best) best = t, b2 = t2;
    }
    p->tune[key] = TuneVal{b1 + 4 * b2, tune_ns(best), tune_ns(t0)};
  }
  const auto it = p->tune.find(key);
  if (it != p->tune.end()) {
    a.tile1 = it->second.choice & 3;
    a.tile2 = (it->second.choice >> 2) & 3;
  }
  ProfScope ps(p, KC_THETA, st);
  hipLaunchKernelGGL(k_theta<T>, dim3(batch), dim3(256), 0, st, a);
}

// One zip-up / density-matrix cut with the plan's kernel choice for its shape.
template <typename T>
static void launch_cut(vmc_plan *p, TruncArgs<T> tr, int batch, hipStream_t st) {
  const bool cplx = is_complex(p->dtype);
  const bool fast = trunc_fast_shape(cplx, tr.m, tr.n, tr.k);
  int v = fast || tr.k == tr.m ? TRUNC_LPE8 : TRUNC_GENERIC;
  Tuner *tu = p->tuner;
  const int64_t need = gtr_need(tr.m, tr.k);
  const size_t gbytes = (size_t)batch * need * (cplx ? 16 : 8);
  if (fast && tr.k < tr.m && !trunc_qr_shape(cplx, tr.m, tr.n, tr.k, tr.dm)) {
    const auto key = trunc_key(tr.m, tr.n, tr.k, tr.dm);
    if (tune_first(tu, key)) {
      if (tu->gscr_bytes < gbytes) {
        (void)hipStreamSynchronize(st);
        if (tu->gscr) (void)hipFree(tu->gscr);
        tu->gscr = nullptr;
        tu->gscr_bytes = 0;
        if (hipMalloc(&tu->gscr, gbytes) == hipSuccess)
          tu->gscr_bytes = gbytes;
        else
          (void)hipGetLastError();  // no generic candidate for this shape
      }
      auto run = [&](int cand) {
        TruncArgs<T> c = tr;
        c.logscale = tu->log;
        c.log_stride = 1;
        c.variant = cand == TRUNC_LPE16 ? TRUNC_LPE16 : TRUNC_LPE8;
        c.gscr = tu->gscr;
        c.gscr_stride = need;
        return tune_time(tu, st, [&] {
          if (cand == TRUNC_GENERIC)
            launch_trunc_g<T>(c, batch, st);
          else
            launch_trunc<T>(c, batch, st);
        });
      };
      const float t0 = run(TRUNC_LPE8);
      float best = t0;
      int bv = TRUNC_LPE8;
      for (int cand = TRUNC_LPE16; cand <= TRUNC_GENERIC; ++cand) {
        if (!trunc_choice_ok(p, tr.m, tr.n, tr.k, tr.dm, cand)) continue;
        if (cand == TRUNC_GENERIC && tu->gscr_bytes < gbytes) continue;  // no scratch for it
        const float t = run(cand);
        if (t < TUNE_MARGIN * best) best = t, bv = cand;
      }
      p->tune[key] = TuneVal{bv, tune_ns(best), tune_ns(t0)};
    }
    const auto it = p->tune.find(key);
    if (it != p->tune.end()) v = it->second.choice;
  }
  if (v == TRUNC_GENERIC && fast && p->gtr_cap < need) {
    // only inside a tuning pass (afterwards gtr_cap covers every generic
    // choice): the workspace has no generic room for this shape yet
    if (tu && tu->gscr_bytes >= gbytes) {
      tr.gscr = tu->gscr;
      tr.gscr_stride = need;
    } else {
      v = TRUNC_LPE8;
    }
  }
  tr.variant = v == TRUNC_LPE16 ? TRUNC_LPE16 : TRUNC_LPE8;
  ProfScope ps(p, KC_TRUNC, st);
  if (v == TRUNC_GENERIC)
    launch_trunc_g<T>(tr, batch, st);
  else
    launch_trunc<T>(tr, batch, st);
}

// Apply one row MPO with variational compression (QR-initialised one-site
// ALS, src/vmc/peps/common/strategy.py:179-427); the kernels and their
// contraction orders are in variational.hpp.
template <typename T>
static void apply_row_var(vmc_plan *p, int row, bool from_below, const T *sites, const int32_t *cfg,
                          const T *in, int64_t in_stride, const std::vector<int64_t> &in_off,
                          const std::vector<Dims3> &in_dims, const double *in_log, int64_t in_log_stride,
                          T *out, int64_t out_stride, const std::vector<int64_t> &out_off,
                          const std::vector<Dims3> &out_dims, double *out_log, int64_t out_log_stride,
                          const Bufs<T> &B, int batch, hipStream_t st) {
  (void)out_dims;
  const int C = p->C;
  const VarTrace tr = var_trace(in_dims, *p, row, from_below);
  const bool direct = p->n_sweeps == 0;  // no sweeps: the QR-initialised state is the result
  std::vector<int64_t> roff(C);
  {
    int64_t o = 0;
    for (int c = 0; c < C; ++c) {
      roff[c] = direct ? out_off[c] : o;
      o += dsize(tr.init[c]);
    }
  }
  T *res = direct ? out : B.vres;
  const int64_t res_stride = direct ? out_stride : p->vres_cap;
  auto wview = [&](const SiteInfo &s) {
    const int64_t s_u = (int64_t)s.dn * s.l * s.r, s_dn = (int64_t)s.l * s.r;
    WView v;
    v.sl = s.r;
    v.sr = 1;
    v.sp = from_below ? s_dn : s_u;
    v.sq = from_below ? s_u : s_dn;
    v.wl = s.l;
    v.wr = s.r;
    v.P = from_below ? s.dn : s.u;
    v.Q = from_below ? s.u : s.dn;
    return v;
  };
  // ---- initialisation: zip-up shapes with a truncated QR (strategy.py:222-254)
  int k_prev = 1;
  for (int c = 0; c < C; ++c) {
    const SiteInfo &s = p->site[row * C + c];
    const WView v = wview(s);
    ThetaArgs<T> a{};
    a.carry = c == 0 ? nullptr : B.carry;
    a.carry_stride = p->carry_cap;
    a.k = k_prev;
    a.Dl = in_dims[c].l;
    a.wl = s.l;
    a.m = in + in_off[c];
    a.m_stride = in_stride;
    a.P = in_dims[c].p;
    a.Dr = in_dims[c].r;
    a.site = SiteRef<T>{sites + s.off, s.slice, cfg, p->R * C, row * C + c};
    a.w_sl = v.sl;
    a.w_sr = v.sr;
    a.w_sp = v.sp;
    a.w_sq = v.sq;
    a.WR = s.r;
    a.Q = v.Q;
    a.t1 = B.s1;
    a.t1_stride = p->scratch_cap;
    const bool last = c == C - 1;
    a.out = last ? res + roff[c] : B.mat;
    a.out_stride = last ? res_stride : p->mat_cap;
    a.log_in = c == 0 ? in_log : nullptr;
    a.log_in_stride = in_log_stride;
    a.log_out = out_log;
    a.log_out_stride = out_log_stride;
    launch_theta<T>(p, a, batch, st);
    if (last) break;
    TruncArgs<T> q{};
    q.mat = B.mat;
    q.mat_stride = p->mat_cap;
    q.m = tr.init[c].l * tr.init[c].p;
    q.n = in_dims[c].r * s.r;
    q.k = tr.init[c].r;
    q.u = res + roff[c];
    q.u_stride = res_stride;
    q.carry = B.carry;
    q.carry_stride = p->carry_cap;
    q.logscale = out_log;
    q.log_stride = out_log_stride;
    q.gscr = B.gtr;
    q.gscr_stride = p->gtr_cap;
    {
      ProfScope ps(p, KC_TRUNC, st);
      if (q.m > VQR_M || q.k > VQR_N || q.n > 128)
        hipLaunchKernelGGL(k_var_qr_init_g<T>, dim3(batch), dim3(256), 0, st, q);
      else
        hipLaunchKernelGGL(k_var_qr_init<T>, dim3(batch), dim3(256), 0, st, q);
    }
    k_prev = tr.init[c].r;
  }
  if (direct) return;

  const int64_t ecap = p->env_cap, estride = (int64_t)C * p->env_cap;
  T *Renv = B.renv, *Lenv = B.vlenv;
  auto site_args = [&](int c) {
    const SiteInfo &s = p->site[row * C + c];
    VarArgs<T> a{};
    a.m = in + in_off[c];
    a.m_stride = in_stride;
    a.Dl = in_dims[c].l;
    a.Dr = in_dims[c].r;
    a.site = SiteRef<T>{sites + s.off, s.slice, cfg, p->R * C, row * C + c};
    a.w = wview(s);
    a.env_stride = estride;
    a.y1 = B.s2;
    a.y2 = B.s3;
    a.opt = B.s4;
    a.cj = B.s5;
    a.s_stride = p->scratch_cap;
    a.gscr = B.gtr;
    a.gscr_stride = p->gtr_cap;
    return a;
  };
  // ---- right environments of the initial result (strategy.py:257-289)
  fill<T>(Renv + (C - 1) * ecap, estride, 1, T(1), batch, st);
  for (int c = C - 1; c >= 1; --c) {
    VarArgs<T> a = site_args(c);
    a.R = Renv + c * ecap;
    a.rr = tr.init[c].r;
    a.res = res + roff[c];
    a.res_stride = res_stride;
    a.ra = tr.init[c].l;
    a.env_out = Renv + (c - 1) * ecap;
    ProfScope ps(p, KC_ROW, st);
    hipLaunchKernelGGL(k_var_renv<T>, dim3(batch), dim3(256), 0, st, a);
  }
  for (int sw = 0; sw < p->n_sweeps; ++sw) {
    const bool final_sweep = sw == p->n_sweeps - 1;
    // left-to-right (strategy.py:292-356); the last site's optimal tensor is
    // overwritten by the right-to-left pass before any read, so it is skipped
    fill<T>(Lenv, estride, 1, T(1), batch, st);
    for (int c = 0; c < C - 1; ++c) {
      VarArgs<T> a = site_args(c);
      a.L = Lenv + c * ecap;
      a.rl = tr.lr_rl[sw][c];
      a.R = Renv + c * ecap;
      a.rr = tr.lr_rr[sw][c];
      a.env_out = Lenv + (c + 1) * ecap;
      ProfScope ps(p, KC_TRUNC, st);
      hipLaunchKernelGGL(k_var_lr<T>, dim3(batch), dim3(256), 0, st, a);
    }
    // right-to-left (strategy.py:359-427); only the last sweep's tensors are kept
    fill<T>(Renv + (C - 1) * ecap, estride, 1, T(1), batch, st);
    for (int c = C - 1; c >= 0; --c) {
      if (c == 0 && !final_sweep) break;
      VarArgs<T> a = site_args(c);
      a.L = Lenv + c * ecap;
      a.rl = tr.lr_rl[sw][c];
      a.R = Renv + c * ecap;
      a.rr = tr.rl_rr[sw][c];
      a.env_out = c > 0 ? Renv + (c - 1) * ecap : nullptr;
      a.out = final_sweep ? out + out_off[c] : nullptr;
      a.out_stride = out_stride;
      a.first = c == 0;
      a.log_in = in_log;
      a.log_in_stride = in_log_stride;
      a.log_out = out_log;
      a.log_out_stride = out_log_stride;
      ProfScope ps(p, KC_TRUNC, st);
      hipLaunchKernelGGL(k_var_rl<T>, dim3(batch), dim3(256), 0, st, a);
    }
  }
}

// Apply one row MPO to a boundary with the plan's strategy (top sweep:
// from_below = false); zip-up / density-matrix / no truncation below.
template <typename T>
static void apply_row(vmc_plan *p, int row, bool from_below, const T *sites, const int32_t *cfg,
                      const T *in, int64_t in_stride, const std::vector<int64_t> &in_off,
                      const std::vector<Dims3> &in_dims, const double *in_log, int64_t in_log_stride,
                      T *out, int64_t out_stride, const std::vector<int64_t> &out_off,
                      const std::vector<Dims3> &out_dims, double *out_log, int64_t out_log_stride,
                      const Bufs<T> &B, int batch, hipStream_t st) {
  if (p->strategy == VMC_VARIATIONAL) {
    apply_row_var<T>(p, row, from_below, sites, cfg, in, in_stride, in_off, in_dims, in_log, in_log_stride, out,
                     out_stride, out_off, out_dims, out_log, out_log_stride, B, batch, st);
    return;
  }
  T *carry = B.carry, *mat = B.mat, *t1 = B.s1;
  const int64_t carry_stride = p->carry_cap, mat_stride = p->mat_cap, t1_stride = p->scratch_cap;
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
    launch_theta<T>(p, a, batch, st);
    if (last) break;
    TruncArgs<T> tr{};
    tr.mat = mat;
    tr.mat_stride = mat_stride;
    // density-matrix: the kernel works on theta^T, whose top-k left vectors are
    // the eigenvectors V_k of theta^H theta (strategy.py:166-174)
    const int rows = (c == 0 ? in_dims[c].l * s.l : k_prev) * a.Q, cols = in_dims[c].r * s.r;
    tr.dm = p->strategy == VMC_DENSITY_MATRIX;
    tr.m = tr.dm ? cols : rows;
    tr.n = tr.dm ? rows : cols;
    tr.k = out_dims[c].r;
    tr.u = out + out_off[c];
    tr.u_stride = out_stride;
    tr.carry = carry;
    tr.carry_stride = carry_stride;
    tr.logscale = out_log;
    tr.log_stride = out_log_stride;
    tr.gscr = B.gtr;
    tr.gscr_stride = p->gtr_cap;
    launch_cut<T>(p, tr, batch, st);
    k_prev = out_dims[c].r;
  }
}

static TermsDev terms_dev(const vmc_plan *p) {
  TermsDev t{};
  if (!p->has_terms || !p->terms_i) return t;
  const int RC = p->R * p->C;
  const int *ti = (const int *)p->terms_i;
  t.one_idx = ti;
  t.h_idx = ti + RC;
  t.v_idx = ti + 2 * RC;
  t.diag_sites = ti + 3 * RC;
  t.diag_off = ti + 3 * RC + 2 * p->n_diag;
  t.n_diag = p->n_diag;
  t.ops = (const double *)p->terms_d;
  t.ops_im = t.ops + p->n_ops;
  return t;
}

template <typename T>
static RowArgs<T> row_args(vmc_plan *p, const Bufs<T> &B, const T *sites, int32_t *cfg, int row, const T *bot,
                           int64_t bot_stride, const double *bot_log, int64_t bot_log_stride,
                           const double *log_amp) {
  RowArgs<T> a{};
  a.C = p->C;
  a.top = B.top;
  a.top_stride = p->top_size;
  a.bot = bot;
  a.bot_stride = bot_stride;
  a.sites = sites;
  a.cfg = cfg;
  a.n_sites = p->R * p->C;
  a.row = row;
  a.renv = B.renv;
  a.renv_stride = (int64_t)p->C * p->env_cap;
  a.env_cap = p->env_cap;
  a.lenv = B.lenv;
  a.lenv_stride = p->env_cap;
  a.s1 = B.s1;
  a.s2 = B.s2;
  a.s3 = B.s3;
  a.s4 = B.s4;
  a.s5 = B.s5;
  a.s_stride = p->scratch_cap;
  a.top_log = B.toplog + row;
  a.top_log_stride = p->R + 1;
  a.bot_log = bot_log;
  a.bot_log_stride = bot_log_stride;
  a.log_amp = log_amp;
  a.geo = (const ColGeo *)p->geo_dev + (int64_t)row * p->C;
  a.d = p->d;
  a.terms = terms_dev(p);
  return a;
}

// Top sweep of all rows into the top-boundary slots + final amplitude.
template <typename T>
static void top_sweep(vmc_plan *p, const Bufs<T> &B, const T *sites, const int32_t *cfg, double *log_amp, int batch,
                      hipStream_t st) {
  const int R = p->R, C = p->C;
  fill<double>(B.toplog, R + 1, 1, 0.0, batch, st);
  fill<T>(B.top, p->top_size, C, T(1), batch, st);
  for (int r = 0; r < R; ++r)
    apply_row<T>(p, r, false, sites, cfg, B.top, p->top_size, p->top_off[r], p->top[r], B.toplog + r, R + 1,
                 B.top, p->top_size, p->top_off[r + 1], p->top[r + 1], B.toplog + r + 1, R + 1, B, batch, st);
  AmpArgs<T> a{};
  a.bnd = B.top;
  a.bnd_stride = p->top_size;
  a.logscale = B.toplog + R;
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

template <typename T>
static int amp_grad_impl(vmc_plan *p, const T *sites, const int32_t *cfg, int batch, double *log_amp, T *o,
                         int8_t *pout, void *ws, size_t ws_bytes, hipStream_t st) {
  const Layout L = make_layout(p, batch, false);
  if (ws_bytes < L.total) return fail("workspace too small: need " + std::to_string(L.total));
  const Bufs<T> B = bufs<T>(L, ws);
  const int R = p->R, C = p->C;
  top_sweep<T>(p, B, sites, cfg, log_amp, batch, st);
  if (!o && !pout) return 0;
  // bottom sweep with env-gradients, row R-1 .. 0 (ping-pong bottom slots)
  int cur = 0;
  fill<double>(B.botlog, 2, 1, 0.0, batch, st);
  fill<T>(B.bot, 2 * p->bot_cap, C, T(1), batch, st);
  for (int r = R - 1; r >= 0; --r) {
    RowArgs<T> a = row_args<T>(p, B, sites, const_cast<int32_t *>(cfg), r, B.bot + (int64_t)cur * p->bot_cap,
                               2 * p->bot_cap, B.botlog + cur, 2, log_amp);
    a.o = o;
    a.o_stride = p->nsliced;
    a.p = pout;
    a.energy = false;
    {
      ProfScope ps(p, KC_ROW, st);
      hipLaunchKernelGGL((k_row_grad<T, false>), dim3(batch), dim3(256), 0, st, a);
    }
    if (r == 0) break;
    const int nxt = cur ^ 1;
    apply_row<T>(p, r, true, sites, cfg, B.bot + (int64_t)cur * p->bot_cap, 2 * p->bot_cap, p->bot_off[r],
                 p->bot[r], B.botlog + cur, 2, B.bot + (int64_t)nxt * p->bot_cap, 2 * p->bot_cap,
                 p->bot_off[r - 1], p->bot[r - 1], B.botlog + nxt, 2, B, batch, st);
    cur = nxt;
  }
  return 0;
}

// Cache layout per chain: bottom boundary of row r at row_base[r] (elements of
// T), then R FP64 log-scales after all chains' tensors.
template <typename T>
struct CacheView {
  T *bnd;
  double *log;
  int64_t stride;
};

template <typename T>
static CacheView<T> cache_view(const vmc_plan *p, void *cache, int n) {
  CacheView<T> v;
  v.bnd = (T *)cache;
  v.stride = p->cache_elems;
  v.log = (double *)((char *)cache + align256((size_t)n * p->cache_elems * sizeof(T)));
  return v;
}

// Bottom boundaries of every row for the given configs (init_cache,
// src/vmc/peps/standard/kernels.py:115-129), also used by the estimate pass.
template <typename T>
static void bottom_rows(vmc_plan *p, const Bufs<T> &B, const T *sites, const int32_t *cfg, const CacheView<T> &cv,
                        int r_hi, int r_lo, int batch, hipStream_t st) {
  const int R = p->R;
  for (int r = r_hi; r > r_lo; --r)
    apply_row<T>(p, r, true, sites, cfg, cv.bnd + p->row_base[r], cv.stride, p->bot_off[r], p->bot[r], cv.log + r, R,
                 cv.bnd + p->row_base[r - 1], cv.stride, p->bot_off[r - 1], p->bot[r - 1], cv.log + r - 1, R,
                 B, batch, st);
}

template <typename T>
static int init_cache_impl(vmc_plan *p, const T *sites, const int32_t *cfg, int n, void *cache, void *ws,
                           size_t ws_bytes, hipStream_t st) {
  const Layout L = make_layout(p, n, false);
  if (ws_bytes < L.total) return fail("workspace too small: need " + std::to_string(L.total));
  const Bufs<T> B = bufs<T>(L, ws);
  const CacheView<T> cv = cache_view<T>(p, cache, n);
  const int R = p->R;
  fill<T>(cv.bnd + p->row_base[R - 1], cv.stride, p->C, T(1), n, st);
  fill<double>(cv.log + R - 1, R, 1, 0.0, n, st);
  bottom_rows<T>(p, B, sites, cfg, cv, R - 1, 0, n, st);
  return 0;
}

// One Metropolis sweep + estimate (mc_sweep of src/vmc/core/sampling.py:28-36 over
// transition kernels.py:131-209 and estimate kernels.py:211-245).
template <typename T>
static int sweep_impl(vmc_plan *p, const T *sites, int32_t *cfg, int n, uint64_t seed, int sweep, void *cache,
                      double *log_amp, T *o, int8_t *pout, double *eloc, void *ws, size_t ws_bytes,
                      hipStream_t st, bool do_transition = true, bool do_estimate = true) {
  const Layout L = make_layout(p, n, p->has_v);
  if (ws_bytes < L.total) return fail("workspace too small: need " + std::to_string(L.total));
  const Bufs<T> B = bufs<T>(L, ws);
  const CacheView<T> cv = cache_view<T>(p, cache, n);
  const int R = p->R, C = p->C;
  if (do_transition) {
  // ---- transition: top -> bottom ----
  fill<double>(B.toplog, R + 1, 1, 0.0, n, st);
  fill<T>(B.top, p->top_size, C, T(1), n, st);
  for (int r = 0; r < R; ++r) {
    RowArgs<T> a = row_args<T>(p, B, sites, cfg, r, cv.bnd + p->row_base[r], cv.stride, cv.log + r, R, nullptr);
    a.seed = seed;
    a.sweep = sweep;
    {
      ProfScope ps(p, KC_ROW, st);
      hipLaunchKernelGGL(k_row_sample<T>, dim3(n), dim3(256), 0, st, a);
    }
    apply_row<T>(p, r, false, sites, cfg, B.top, p->top_size, p->top_off[r], p->top[r], B.toplog + r, R + 1, B.top,
                 p->top_size, p->top_off[r + 1], p->top[r + 1], B.toplog + r + 1, R + 1, B, n, st);
  }
  {
    AmpArgs<T> a{};
    a.bnd = B.top;
    a.bnd_stride = p->top_size;
    a.logscale = B.toplog + R;
    a.log_stride = R + 1;
    a.log_amp = log_amp;
    a.C = C;
    for (int c = 0; c < C; ++c) {
      a.off[c] = p->top_off[R][c];
      a.l[c] = p->top[R][c].l;
      a.r[c] = p->top[R][c].r;
    }
    hipLaunchKernelGGL(k_amp<T>, dim3(n), dim3(64), 0, st, a);
  }
  }
  if (!do_estimate) return 0;
  // ---- estimate: bottom -> top ----
  const bool energy = p->has_terms && eloc != nullptr;
  double *el = eloc ? eloc : B.eloc;
  if (energy && p->n_diag > 0)
    hipLaunchKernelGGL(k_diag, dim3((n + 255) / 256), dim3(256), 0, st, terms_dev(p), cfg, R * C, p->d, el, n,
                       (int)is_complex(p->dtype));
  else
    fill<double>(el, 1, 1, 0.0, (is_complex(p->dtype) ? 2 : 1) * n, st);
  fill<T>(cv.bnd + p->row_base[R - 1], cv.stride, C, T(1), n, st);
  fill<double>(cv.log + R - 1, R, 1, 0.0, n, st);
  for (int r = R - 1; r >= 0; --r) {
    RowArgs<T> a = row_args<T>(p, B, sites, cfg, r, cv.bnd + p->row_base[r], cv.stride, cv.log + r, R, log_amp);
    a.o = o;
    a.o_stride = p->nsliced;
    a.p = pout;
    a.energy = energy;
    a.eloc = el;
    {
      ProfScope ps(p, KC_ROW, st);
      if (energy)
        hipLaunchKernelGGL((k_row_grad<T, true>), dim3(n), dim3(256), 0, st, a);
      else
        hipLaunchKernelGGL((k_row_grad<T, false>), dim3(n), dim3(256), 0, st, a);
    }
    if (energy && r < R - 1 && p->v_row[r]) {
      VertArgs<T> v{};
      v.C = C;
      v.top = B.top + p->top_off[r][0] * 0;
      v.top_stride = p->top_size;
      v.bot = cv.bnd + p->row_base[r + 1];
      v.bot_stride = cv.stride;
      v.sites = sites;
      v.cfg = cfg;
      v.n_sites = R * C;
      v.row = r;
      v.geo0 = (const ColGeo *)p->geo_dev + (int64_t)r * C;
      v.geo1 = (const ColGeo *)p->geo_dev + (int64_t)(r + 1) * C;
      v.r2 = B.r2;
      v.r2_stride = (int64_t)C * p->cap2;
      v.cap2 = p->cap2;
      v.l2 = B.l2;
      v.l2_stride = p->cap2;
      v.v1 = B.v1;
      v.v2 = B.v2;
      v.v3 = B.v3;
      v.v_stride = p->vcap;
      v.top_log = B.toplog + r;
      v.top_log_stride = R + 1;
      v.bot_log = cv.log + r + 1;
      v.bot_log_stride = R;
      v.log_amp = log_amp;
      v.d = p->d;
      v.terms = terms_dev(p);
      v.eloc = el;
      ProfScope ps(p, KC_ROW, st);
      hipLaunchKernelGGL(k_row_vertical<T>, dim3(n), dim3(256), 0, st, v);
    }
    if (r > 0) bottom_rows<T>(p, B, sites, cfg, cv, r, r - 1, n, st);
  }
  return 0;
}


// ------------------------------------------------------------------------- //
// Gauge-invariant PEPS (src/vmc/peps/gi/*)
// ------------------------------------------------------------------------- //

static GiDev gi_dev(const vmc_plan *p) {
  GiDev g;
  g.R = p->R;
  g.C = p->C;
  g.N = p->gi_N;
  g.Qx = p->gi_Qx;
  g.d = p->d;
  g.charge = p->gi_charge;
  g.ctab = p->gi_ctab;
  g.cdeg = p->gi_cdeg;
  g.maxdeg = p->gi_maxdeg;
  return g;
}

static void gi_slices(const vmc_plan *p, const int32_t *samples, int n, int32_t *slices, uint8_t *valid,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_gi_slices, dim3((n + 127) / 128), dim3(128), 0, st, gi_dev(p), samples, n, slices, valid);
}

// The parameters the GI kernels read: with a non-uniform degeneracy_per_charge
// a copy in the workspace with every slice's charge masks applied, else the
// caller's tensors.
template <typename T>
static const T *gi_sites(const vmc_plan *p, const T *sites, const Layout &L, void *ws, hipStream_t st) {
  if (!p->gi_masked) return sites;
  T *m = (T *)((char *)ws + L.bmask.off);
  hipLaunchKernelGGL(k_gi_mask_sites<T>, dim3(p->R * p->C), dim3(256), 0, st, gi_dev(p), p->D, p->gi_deg, sites, m);
  return m;
}

template <typename T>
static GiPlaqArgs<T> gi_plaq_args(vmc_plan *p, const Bufs<T> &B, const Layout &L, void *ws, const T *sites,
                                  int32_t *samples, int32_t *gsl, const CacheView<T> &cv, int r) {
  char *base = (char *)ws;
  GiPlaqArgs<T> a{};
  a.C = p->C;
  a.row = r;
  a.g = gi_dev(p);
  a.top = B.top;
  a.top_stride = p->top_size;
  a.bot = cv.bnd + p->row_base[r + 1];
  a.bot_stride = cv.stride;
  a.sites = sites;
  a.samples = samples;
  a.slices = gsl;
  a.geo0 = (const ColGeo *)p->geo_dev + (int64_t)r * p->C;
  a.geo1 = (const ColGeo *)p->geo_dev + (int64_t)(r + 1) * p->C;
  a.r2 = B.r2;
  a.r2_stride = (int64_t)p->C * p->cap2;
  a.cap2 = p->cap2;
  a.l2 = B.l2;
  a.lx = (T *)(base + L.gx.off);
  a.ly = (T *)(base + L.gy.off);
  a.lxp = (T *)(base + L.gxp.off);
  a.lyp = (T *)(base + L.gyp.off);
  a.l_stride = p->cap2;
  a.v1 = B.v1;
  a.v2 = B.v2;
  a.v3 = B.v3;
  a.v_stride = p->vcap;
  a.coeff = p->gi_plaq;
  a.top_log = B.toplog + r;
  a.top_log_stride = p->R + 1;
  a.bot_log = cv.log + r + 1;
  a.bot_log_stride = p->R;
  return a;
}

// One GI sweep: pure-gauge plaquette transition (gi/model.py:901-960) and
// estimate (gi/model.py:392-612 + gi/kernels.py:90-126).
template <typename T>
static int gi_sweep_impl(vmc_plan *p, const T *sites_in, int32_t *samples, int n, uint64_t seed, int sweep, void *cache,
                         double *log_amp, T *o, int8_t *pout, double *eloc, void *ws, size_t ws_bytes,
                         hipStream_t st) {
  const Layout L = make_layout(p, n, true);
  if (ws_bytes < L.total) return fail("workspace too small: need " + std::to_string(L.total));
  const Bufs<T> B = bufs<T>(L, ws);
  int32_t *gsl = (int32_t *)((char *)ws + L.gsl.off);
  const CacheView<T> cv = cache_view<T>(p, cache, n);
  const int R = p->R, C = p->C;
  gi_slices(p, samples, n, gsl, nullptr, st);
  const T *sites = gi_sites<T>(p, sites_in, L, ws, st);
  // ---- transition: plaquette sweep of each row pair, then the row joins the top boundary
  // (the plaquette stage reads the cached bottom boundaries of the previous sweep)
  fill<double>(B.toplog, R + 1, 1, 0.0, n, st);
  fill<T>(B.top, p->top_size, C, T(1), n, st);
  for (int r = 0; r < R; ++r) {
    if (r < R - 1) {
      GiPlaqArgs<T> a = gi_plaq_args<T>(p, B, L, ws, sites, samples, gsl, cv, r);
      a.transition = 1;
      a.seed = seed;
      a.sweep = sweep;
      ProfScope ps(p, KC_ROW, st);
      hipLaunchKernelGGL(k_gi_plaquette<T>, dim3(n), dim3(256), 0, st, a);
    }
    apply_row<T>(p, r, false, sites, gsl, B.top, p->top_size, p->top_off[r], p->top[r], B.toplog + r, R + 1, B.top,
                 p->top_size, p->top_off[r + 1], p->top[r + 1], B.toplog + r + 1, R + 1, B, n, st);
  }
  if (p->d > 1) {
    // with matter: horizontal link sweeps of every row, then vertical link
    // sweeps of every row pair, each over freshly rebuilt bottom boundaries
    // (gi/model.py:962-1036); the top boundary is rebuilt along each stage
    auto rebuild_bottoms = [&]() {
      fill<T>(cv.bnd + p->row_base[R - 1], cv.stride, C, T(1), n, st);
      fill<double>(cv.log + R - 1, R, 1, 0.0, n, st);
      bottom_rows<T>(p, B, sites, gsl, cv, R - 1, 0, n, st);
    };
    auto top_row = [&](int r) {
      apply_row<T>(p, r, false, sites, gsl, B.top, p->top_size, p->top_off[r], p->top[r], B.toplog + r, R + 1,
                   B.top, p->top_size, p->top_off[r + 1], p->top[r + 1], B.toplog + r + 1, R + 1, B, n, st);
    };
    rebuild_bottoms();
    fill<double>(B.toplog, R + 1, 1, 0.0, n, st);
    fill<T>(B.top, p->top_size, C, T(1), n, st);
    for (int r = 0; r < R; ++r) {
      RowArgs<T> a = row_args<T>(p, B, sites, gsl, r, cv.bnd + p->row_base[r], cv.stride, cv.log + r, R, log_amp);
      a.seed = seed;
      a.sweep = sweep;
      {
        ProfScope ps(p, KC_ROW, st);
        hipLaunchKernelGGL(k_gi_hlink<T>, dim3(n), dim3(256), 0, st, a, gi_dev(p), samples);
      }
      top_row(r);
    }
    if (R > 1) {
      rebuild_bottoms();
      fill<double>(B.toplog, R + 1, 1, 0.0, n, st);
      fill<T>(B.top, p->top_size, C, T(1), n, st);
      for (int r = 0; r < R; ++r) {
        if (r < R - 1) {
          GiPlaqArgs<T> a = gi_plaq_args<T>(p, B, L, ws, sites, samples, gsl, cv, r);
          a.transition = 1;
          a.vlink = 1;
          a.seed = seed;
          a.sweep = sweep;
          ProfScope ps(p, KC_ROW, st);
          hipLaunchKernelGGL(k_gi_plaquette<T>, dim3(n), dim3(256), 0, st, a);
        }
        top_row(r);
      }
    }
  }
  {
    AmpArgs<T> a{};
    a.bnd = B.top;
    a.bnd_stride = p->top_size;
    a.logscale = B.toplog + R;
    a.log_stride = R + 1;
    a.log_amp = log_amp;
    a.C = C;
    for (int c = 0; c < C; ++c) {
      a.off[c] = p->top_off[R][c];
      a.l[c] = p->top[R][c].l;
      a.r[c] = p->top[R][c].r;
    }
    hipLaunchKernelGGL(k_amp<T>, dim3(n), dim3(64), 0, st, a);
  }
  // ---- estimate: gradients, link and plaquette energies, new bottom boundaries
  const bool energy = eloc != nullptr;
  double *el = eloc ? eloc : B.eloc;
  const int cplx = is_complex(p->dtype);
  if (energy && p->gi_has_link)
    hipLaunchKernelGGL(k_gi_link_energy, dim3((n + 127) / 128), dim3(128), 0, st, gi_dev(p), samples, n,
                       p->gi_link, el, cplx);
  else
    fill<double>(el, 1, 1, 0.0, (cplx ? 2 : 1) * n, st);
  if (energy && p->gi_has_site)
    hipLaunchKernelGGL(k_gi_site_diag, dim3((n + 127) / 128), dim3(128), 0, st, terms_dev(p), samples,
                       R * C + R * (C - 1) + (R - 1) * C, p->d, el, n, cplx);
  fill<T>(cv.bnd + p->row_base[R - 1], cv.stride, C, T(1), n, st);
  fill<double>(cv.log + R - 1, R, 1, 0.0, n, st);
  // matter operators (one-site / horizontal in the row kernel, vertical in
  // k_row_vertical) act on the physical states p * nc + s % nc of each site
  const bool matter_ops = energy && (p->has_one || p->has_h);
  for (int r = R - 1; r >= 0; --r) {
    RowArgs<T> a = row_args<T>(p, B, sites, gsl, r, cv.bnd + p->row_base[r], cv.stride, cv.log + r, R, log_amp);
    a.o = o;
    a.o_stride = p->nsliced;
    a.p = pout;
    a.energy = matter_ops;
    a.eloc = el;
    {
      ProfScope ps(p, KC_ROW, st);
      if (matter_ops)
        hipLaunchKernelGGL((k_row_grad<T, true>), dim3(n), dim3(256), 0, st, a);
      else
        hipLaunchKernelGGL((k_row_grad<T, false>), dim3(n), dim3(256), 0, st, a);
    }
    if (energy && r < R - 1 && p->has_v && p->v_row[r]) {
      VertArgs<T> v{};
      v.C = C;
      v.top = B.top;
      v.top_stride = p->top_size;
      v.bot = cv.bnd + p->row_base[r + 1];
      v.bot_stride = cv.stride;
      v.sites = sites;
      v.cfg = gsl;
      v.n_sites = R * C;
      v.row = r;
      v.geo0 = (const ColGeo *)p->geo_dev + (int64_t)r * C;
      v.geo1 = (const ColGeo *)p->geo_dev + (int64_t)(r + 1) * C;
      v.r2 = B.r2;
      v.r2_stride = (int64_t)C * p->cap2;
      v.cap2 = p->cap2;
      v.l2 = B.l2;
      v.l2_stride = p->cap2;
      v.v1 = B.v1;
      v.v2 = B.v2;
      v.v3 = B.v3;
      v.v_stride = p->vcap;
      v.top_log = B.toplog + r;
      v.top_log_stride = R + 1;
      v.bot_log = cv.log + r + 1;
      v.bot_log_stride = R;
      v.log_amp = log_amp;
      v.d = p->d;
      v.terms = terms_dev(p);
      v.eloc = el;
      ProfScope ps(p, KC_ROW, st);
      hipLaunchKernelGGL(k_row_vertical<T>, dim3(n), dim3(256), 0, st, v);
    }
    if (energy && r < R - 1 && p->gi_has_plaq && p->gi_plaq_row[r]) {
      GiPlaqArgs<T> q = gi_plaq_args<T>(p, B, L, ws, sites, samples, gsl, cv, r);
      q.transition = 0;
      q.log_amp = log_amp;
      q.eloc = el;
      ProfScope ps(p, KC_ROW, st);
      hipLaunchKernelGGL(k_gi_plaquette<T>, dim3(n), dim3(256), 0, st, q);
    }
    if (r > 0) bottom_rows<T>(p, B, sites, gsl, cv, r, r - 1, n, st);
  }
  return 0;
}

// ------------------------------------------------------------------------- //
// Blockade PEPS (src/vmc/peps/blockade/*)
// ------------------------------------------------------------------------- //

static BlDev bl_dev(const vmc_plan *p) {
  BlDev g;
  g.R = p->R;
  g.C = p->C;
  g.D0 = p->bl_D0;
  g.D1 = p->bl_D1;
  g.Dmax = p->D;
  return g;
}

static void bl_slices(const vmc_plan *p, const int32_t *occ, int n, int32_t *slices, uint8_t *valid, hipStream_t st) {
  hipLaunchKernelGGL(k_bl_slices, dim3((n + 127) / 128), dim3(128), 0, st, bl_dev(p), occ, n, slices, valid);
}

// The parameters the kernels read: with D0 != D1 a copy in the workspace with
// every slice's charge masks applied, else the caller's tensors.
template <typename T>
static const T *bl_sites(const vmc_plan *p, const T *sites, const Layout &L, void *ws, hipStream_t st) {
  if (p->bl_D0 == p->bl_D1) return sites;
  T *m = (T *)((char *)ws + L.bmask.off);
  hipLaunchKernelGGL(k_bl_mask_sites<T>, dim3(p->R * p->C), dim3(256), 0, st, bl_dev(p), sites, m);
  return m;
}

// Slice indices of a gauge-invariant or blockade sample batch into gsl (and the
// sample's validity into gvalid when given); returns the parameters to read:
// the caller's, or a workspace copy with the charge masks of every slice.
template <typename T>
static const T *sliced_inputs(const vmc_plan *p, const T *sites, const int32_t *samples, int n, const Layout &L,
                              void *ws, uint8_t *gvalid, hipStream_t st) {
  int32_t *gsl = (int32_t *)((char *)ws + L.gsl.off);
  if (p->bl) {
    bl_slices(p, samples, n, gsl, gvalid, st);
    return bl_sites<T>(p, sites, L, ws, st);
  }
  gi_slices(p, samples, n, gsl, gvalid, st);
  return gi_sites<T>(p, sites, L, ws, st);
}

template <typename T>
static BlPairArgs<T> bl_pair_args(vmc_plan *p, const Bufs<T> &B, const Layout &L, void *ws, const T *sites,
                                  int32_t *occ, int32_t *gsl, const CacheView<T> &cv, int r) {
  char *base = (char *)ws;
  BlPairArgs<T> a{};
  a.C = p->C;
  a.row = r;
  a.g = bl_dev(p);
  a.top = B.top;
  a.top_stride = p->top_size;
  a.bot = cv.bnd + p->row_base[r + 1];
  a.bot_stride = cv.stride;
  a.sites = sites;
  a.occ = occ;
  a.slices = gsl;
  a.geo0 = (const ColGeo *)p->geo_dev + (int64_t)r * p->C;
  a.geo1 = (const ColGeo *)p->geo_dev + (int64_t)(r + 1) * p->C;
  a.r2 = B.r2;
  a.r2_stride = (int64_t)p->C * p->cap2;
  a.cap2 = p->cap2;
  a.l2 = B.l2;
  a.lx = (T *)(base + L.gx.off);
  a.ly = (T *)(base + L.gy.off);
  a.lxp = (T *)(base + L.gxp.off);
  a.lyp = (T *)(base + L.gyp.off);
  a.l_stride = p->cap2;
  a.v1 = B.v1;
  a.v2 = B.v2;
  a.v3 = B.v3;
  a.v_stride = p->vcap;
  a.terms = terms_dev(p);
  a.top_log = B.toplog + r;
  a.top_log_stride = p->R + 1;
  a.bot_log = cv.log + r + 1;
  a.bot_log_stride = p->R;
  return a;
}

// One blockade sweep: transition (blockade/model.py:620-665: row pairs, then
// the last row) and estimate (model.py:295-617 + kernels.py:50-117).
template <typename T>
static int bl_sweep_impl(vmc_plan *p, const T *sites_in, int32_t *occ, int n, uint64_t seed, int sweep, void *cache,
                         double *log_amp, T *o, int8_t *pout, double *eloc, void *ws, size_t ws_bytes,
                         hipStream_t st) {
  const Layout L = make_layout(p, n, true);
  if (ws_bytes < L.total) return fail("workspace too small: need " + std::to_string(L.total));
  const Bufs<T> B = bufs<T>(L, ws);
  int32_t *gsl = (int32_t *)((char *)ws + L.gsl.off);
  const CacheView<T> cv = cache_view<T>(p, cache, n);
  const int R = p->R, C = p->C;
  const T *sites = sliced_inputs<T>(p, sites_in, occ, n, L, ws, nullptr, st);
  // ---- transition: row pairs over the cached bottom boundaries, each row then
  // joining the top boundary; the last row over a bottom of ones
  fill<double>(B.toplog, R + 1, 1, 0.0, n, st);
  fill<T>(B.top, p->top_size, C, T(1), n, st);
  for (int r = 0; r < R; ++r) {
    if (r < R - 1) {
      BlPairArgs<T> a = bl_pair_args<T>(p, B, L, ws, sites, occ, gsl, cv, r);
      a.transition = 1;
      a.seed = seed;
      a.sweep = sweep;
      ProfScope ps(p, KC_ROW, st);
      hipLaunchKernelGGL(k_bl_pair<T>, dim3(n), dim3(256), 0, st, a);
    } else {
      // the cache row R-1 holds the bottom of ones below the last row
      RowArgs<T> a = row_args<T>(p, B, sites, gsl, r, cv.bnd + p->row_base[r], cv.stride, cv.log + r, R, log_amp);
      a.seed = seed;
      a.sweep = sweep;
      ProfScope ps(p, KC_ROW, st);
      hipLaunchKernelGGL(k_bl_row<T>, dim3(n), dim3(256), 0, st, a, bl_dev(p), occ, 1);
    }
    apply_row<T>(p, r, false, sites, gsl, B.top, p->top_size, p->top_off[r], p->top[r], B.toplog + r, R + 1, B.top,
                 p->top_size, p->top_off[r + 1], p->top[r + 1], B.toplog + r + 1, R + 1, B, n, st);
  }
  {
    AmpArgs<T> a{};
    a.bnd = B.top;
    a.bnd_stride = p->top_size;
    a.logscale = B.toplog + R;
    a.log_stride = R + 1;
    a.log_amp = log_amp;
    a.C = C;
    for (int c = 0; c < C; ++c) {
      a.off[c] = p->top_off[R][c];
      a.l[c] = p->top[R][c].l;
      a.r[c] = p->top[R][c].r;
    }
    hipLaunchKernelGGL(k_amp<T>, dim3(n), dim3(64), 0, st, a);
  }
  // ---- estimate: gradients, diagonal and one-site energies, new bottom boundaries
  const bool energy = eloc != nullptr;
  double *el = eloc ? eloc : B.eloc;
  const int cplx = is_complex(p->dtype);
  fill<double>(el, 1, 1, 0.0, (cplx ? 2 : 1) * n, st);
  if (energy && p->n_diag > 0)
    hipLaunchKernelGGL(k_diag, dim3((n + 127) / 128), dim3(128), 0, st, terms_dev(p), occ, R * C, 2, el, n, cplx);
  fill<T>(cv.bnd + p->row_base[R - 1], cv.stride, C, T(1), n, st);
  fill<double>(cv.log + R - 1, R, 1, 0.0, n, st);
  for (int r = R - 1; r >= 0; --r) {
    RowArgs<T> a = row_args<T>(p, B, sites, gsl, r, cv.bnd + p->row_base[r], cv.stride, cv.log + r, R, log_amp);
    a.o = o;
    a.o_stride = p->nsliced;
    a.p = pout;
    {
      ProfScope ps(p, KC_ROW, st);
      hipLaunchKernelGGL((k_row_grad<T, false>), dim3(n), dim3(256), 0, st, a);
    }
    if (energy && p->has_one) {
      ProfScope ps(p, KC_ROW, st);
      if (r < R - 1) {
        BlPairArgs<T> q = bl_pair_args<T>(p, B, L, ws, sites, occ, gsl, cv, r);
        q.transition = 0;
        q.log_amp = log_amp;
        q.eloc = el;
        hipLaunchKernelGGL(k_bl_pair<T>, dim3(n), dim3(256), 0, st, q);
      } else {
        a.eloc = el;
        hipLaunchKernelGGL(k_bl_row<T>, dim3(n), dim3(256), 0, st, a, bl_dev(p), occ, 0);
      }
    }
    if (r > 0) bottom_rows<T>(p, B, sites, gsl, cv, r, r - 1, n, st);
  }
  return 0;
}

template <typename T>
static int gi_amp_grad_impl(vmc_plan *p, const T *sites_in, const int32_t *samples, int batch, double *log_amp, T *o,
                            int8_t *pout, void *ws, size_t ws_bytes, hipStream_t st) {
  const Layout L = make_layout(p, batch, false);
  if (ws_bytes < L.total) return fail("workspace too small: need " + std::to_string(L.total));
  int32_t *gsl = (int32_t *)((char *)ws + L.gsl.off);
  uint8_t *gvalid = (uint8_t *)ws + L.gvalid.off;
  const T *sites = sliced_inputs<T>(p, sites_in, samples, batch, L, ws, gvalid, st);
  const int rc = amp_grad_impl<T>(p, sites, gsl, batch, log_amp, o, pout, ws, ws_bytes, st);
  if (rc) return rc;
  const int64_t n_o = o ? p->nsliced : 1;
  const int64_t tot = (int64_t)batch * n_o;
  hipLaunchKernelGGL(k_gi_mask<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, gvalid, batch, log_amp, o,
                     n_o);
  return 0;
}

template <typename T>
static int gi_init_cache_impl(vmc_plan *p, const T *sites_in, const int32_t *samples, int n, void *cache, void *ws,
                              size_t ws_bytes, hipStream_t st) {
  const Layout L = make_layout(p, n, false);
  if (ws_bytes < L.total) return fail("workspace too small: need " + std::to_string(L.total));
  int32_t *gsl = (int32_t *)((char *)ws + L.gsl.off);
  const T *sites = sliced_inputs<T>(p, sites_in, samples, n, L, ws, nullptr, st);
  return init_cache_impl<T>(p, sites, gsl, n, cache, ws, ws_bytes, st);
}

// Per-dtype drivers behind the C ABI. engine_dt.hip instantiates one table per
// storage scalar; the C ABI in engine.hip picks the table of the plan dtype.
struct EngineOps {
  int (*amp_grad)(vmc_plan *, const void *, const int32_t *, int, double *, void *, int8_t *, void *, size_t,
                  hipStream_t);
  int (*init_cache)(vmc_plan *, const void *, const int32_t *, int, void *, void *, size_t, hipStream_t);
  int (*sweep)(vmc_plan *, const void *, int32_t *, int, uint64_t, int, void *, double *, void *, int8_t *,
               double *, void *, size_t, hipStream_t, bool, bool);
  int (*top_sweep)(vmc_plan *, const void *, const int32_t *, double *, int, void *, hipStream_t);
  int (*gi_amp_grad)(vmc_plan *, const void *, const int32_t *, int, double *, void *, int8_t *, void *, size_t,
                     hipStream_t);
  int (*gi_init_cache)(vmc_plan *, const void *, const int32_t *, int, void *, void *, size_t, hipStream_t);
  int (*gi_sweep)(vmc_plan *, const void *, int32_t *, int, uint64_t, int, void *, double *, void *, int8_t *,
                  double *, void *, size_t, hipStream_t);
  int (*bl_sweep)(vmc_plan *, const void *, int32_t *, int, uint64_t, int, void *, double *, void *, int8_t *,
                  double *, void *, size_t, hipStream_t);
};
extern EngineOps ops_f32, ops_f64, ops_c64, ops_c128;  // non-const: a const table would also be emitted for the device

}  // namespace vmc
