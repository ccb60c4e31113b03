// C ABI of the PEPS amplitude engine (include/vmc_amd.h). Plan bookkeeping and
// argument checks live here; the dtype-specific drivers are the EngineOps
// tables that engine_dt.hip instantiates once per storage scalar.
#include "engine_core.hpp"

#include <mutex>

namespace vmc {

static thread_local std::string g_last_error;

int fail(const std::string &msg) {
  g_last_error = msg;
  return -1;
}

int sr_fail(const std::string &msg) { return fail(msg); }

// ---- Static scratch of vmc_amp_grad_noalloc -------------------------------
// The workspace of the no-allocation entry point: a file-scope device array of
// VMC_NOALLOC_SCRATCH_BYTES per GPU, held from module load until the process
// ends. In host code `scratch` is only a symbol handle; its device address on
// each GPU comes from hipGetSymbolAddress (once per device, with that device
// current) and only that address is passed to kernels.
static __device__ __align__(256) unsigned char scratch[VMC_NOALLOC_SCRATCH_BYTES];
constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[SCRATCH_MAX_DEVICES];

// Device address of the current device's scratch array.
static int scratch_address(void **out) {
  int dev = -1;
  VMC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return fail("device index beyond the static scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    VMC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail("hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// Driver table of a plan dtype (nullptr for an unknown dtype).
static const EngineOps *ops_of(int dtype) {
  switch (dtype) {
    case VMC_F32:
      return &ops_f32;
    case VMC_F64:
      return &ops_f64;
    case VMC_C64:
      return &ops_c64;
    case VMC_C128:
      return &ops_c128;
  }
  return nullptr;
}


// Operator tables of vmc_plan_set_terms (cplx = false: real values) and
// vmc_plan_set_terms_c (cplx = true: interleaved re, im). Stored on the device
// as all real parts followed by all imaginary parts.
static int set_terms_impl(vmc_plan *plan, int n_diag, const int32_t *diag_sites, const double *diag_vals, int n_one,
                          const int32_t *one_sites, const double *one_ops, int n_h, const int32_t *h_sites,
                          const double *h_ops, int n_v, const int32_t *v_sites, const double *v_ops, bool cplx) {
  if (!plan) return fail("plan is null");
  const int R = plan->R, C = plan->C, RC = R * C, d = plan->d, d2 = d * d, d4 = d2 * d2;
  if (d * d > 64) return fail("phys_dim above 8 is not supported by the energy kernels");
  const int cs = cplx ? 2 : 1;  // doubles per value
  std::vector<double> re, im;
  std::vector<int> one(RC, -1), h(RC, -1), v(RC, -1);
  auto accumulate = [&](std::vector<int> &idx, int site, const double *op, int n) {
    if (idx[site] < 0) {
      idx[site] = (int)re.size();
      for (int i = 0; i < n; ++i) {
        re.push_back(op[cs * i]);
        im.push_back(cplx ? op[cs * i + 1] : 0.0);
      }
    } else {
      for (int i = 0; i < n; ++i) {
        re[idx[site] + i] += op[cs * i];
        if (cplx) im[idx[site] + i] += op[cs * i + 1];
      }
    }
  };
  for (int i = 0; i < n_one; ++i) {
    if (one_sites[i] < 0 || one_sites[i] >= RC) return fail("one-site term outside the lattice");
    accumulate(one, one_sites[i], one_ops + (int64_t)i * d2 * cs, d2);
  }
  for (int i = 0; i < n_h; ++i) {
    const int st = h_sites[i];
    if (st < 0 || st >= RC || st % C == C - 1) return fail("horizontal term outside the lattice");
    accumulate(h, st, h_ops + (int64_t)i * d4 * cs, d4);
  }
  std::vector<char> vrow(R, 0);
  for (int i = 0; i < n_v; ++i) {
    const int st = v_sites[i];
    if (st < 0 || st >= RC - C) return fail("vertical term outside the lattice");
    accumulate(v, st, v_ops + (int64_t)i * d4 * cs, d4);
    vrow[st / C] = 1;
  }
  std::vector<int> dsites(2 * (size_t)std::max(n_diag, 0)), doff(std::max(n_diag, 0));
  for (int i = 0; i < n_diag; ++i) {
    const int a0 = diag_sites[2 * i], a1 = diag_sites[2 * i + 1];
    if (a0 < 0 || a0 >= RC || a1 >= RC) return fail("diagonal term outside the lattice");
    dsites[2 * i] = a0;
    dsites[2 * i + 1] = a1;
    doff[i] = (int)re.size();
    const int n = a1 >= 0 ? d2 : d;
    const double *vals = diag_vals + (int64_t)i * d2 * cs;
    for (int j = 0; j < n; ++j) {
      re.push_back(vals[cs * j]);
      im.push_back(cplx ? vals[cs * j + 1] : 0.0);
    }
  }
  const bool cplx_plan = plan->dtype == VMC_C64 || plan->dtype == VMC_C128;
  if (!cplx_plan)
    for (double x : im)
      if (x != 0.0) return fail("complex operator values need a complex plan (VMC_C64 or VMC_C128)");
  std::vector<int> ints;
  ints.insert(ints.end(), one.begin(), one.end());
  ints.insert(ints.end(), h.begin(), h.end());
  ints.insert(ints.end(), v.begin(), v.end());
  ints.insert(ints.end(), dsites.begin(), dsites.end());
  ints.insert(ints.end(), doff.begin(), doff.end());
  if (re.empty()) {
    re.push_back(0.0);
    im.push_back(0.0);
  }
  std::vector<double> ops(re);
  ops.insert(ops.end(), im.begin(), im.end());
  if (plan->terms_i) (void)hipFree(plan->terms_i);
  if (plan->terms_d) (void)hipFree(plan->terms_d);
  plan->terms_i = plan->terms_d = nullptr;
  VMC_HIP(hipMalloc(&plan->terms_i, ints.size() * sizeof(int)));
  VMC_HIP(hipMemcpy(plan->terms_i, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice));
  VMC_HIP(hipMalloc(&plan->terms_d, ops.size() * sizeof(double)));
  VMC_HIP(hipMemcpy(plan->terms_d, ops.data(), ops.size() * sizeof(double), hipMemcpyHostToDevice));
  plan->n_ops = (int64_t)re.size();
  plan->n_diag = n_diag;
  plan->has_terms = true;
  if (plan->gi) {
    plan->gi_has_site = n_diag > 0;
    plan->has_terms = plan->gi_has_plaq || plan->gi_has_link || plan->gi_has_site || n_one > 0 || n_h > 0 ||
                      n_v > 0;
  }
  plan->has_one = n_one > 0;
  plan->has_h = n_h > 0;
  plan->has_v = n_v > 0;
  plan->v_row = vrow;
  return 0;
}

}  // namespace vmc

// ------------------------------------------------------------------------- //
// C ABI
// ------------------------------------------------------------------------- //

extern "C" {

int vmc_plan_create(vmc_plan **plan, int n_rows, int n_cols, int phys_dim, int bond_dim, int chi,
                    int strategy, int n_sweeps, int dtype) {
  if (!plan) return vmc::fail("plan is null");
  if (n_rows < 1 || n_cols < 1 || phys_dim < 1 || bond_dim < 1)
    return vmc::fail("lattice shape, phys_dim and bond_dim must be positive");
  if (n_cols > 64) return vmc::fail("n_cols above 64 is not supported");
  if (dtype != VMC_F32 && dtype != VMC_F64 && dtype != VMC_C64 && dtype != VMC_C128)
    return vmc::fail("dtype must be VMC_F32, VMC_F64, VMC_C64 or VMC_C128");
  if (strategy != VMC_ZIPUP && strategy != VMC_NO_TRUNCATION && strategy != VMC_DENSITY_MATRIX &&
      strategy != VMC_VARIATIONAL)
    return vmc::fail("unsupported strategy");
  if (strategy == VMC_VARIATIONAL && n_sweeps < 0) return vmc::fail("n_sweeps must be non-negative");
  if (strategy == VMC_NO_TRUNCATION) chi = 1 << 30;  // keep every singular vector (fails above 32)
  if (chi < 1) return vmc::fail("chi must be positive");
  vmc_plan *p = new vmc_plan();
  p->R = n_rows;
  p->C = n_cols;
  p->d = phys_dim;
  p->D = bond_dim;
  p->chi = chi;
  p->strategy = strategy;
  p->n_sweeps = strategy == VMC_VARIATIONAL ? n_sweeps : 0;
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
  if (plan && plan->gi_charge) (void)hipFree(plan->gi_charge);
  if (plan && plan->gi_ctab) (void)hipFree(plan->gi_ctab);
  if (plan && plan->gi_cdeg) (void)hipFree(plan->gi_cdeg);
  if (plan && plan->gi_plaq) (void)hipFree(plan->gi_plaq);
  if (plan && plan->gi_link) (void)hipFree(plan->gi_link);
  if (plan && plan->gi_deg) (void)hipFree(plan->gi_deg);
  if (plan && plan->terms_i) (void)hipFree(plan->terms_i);
  if (plan && plan->terms_d) (void)hipFree(plan->terms_d);
  if (plan && plan->geo_dev) (void)hipFree(plan->geo_dev);
  delete plan;
  return 0;
}

int64_t vmc_plan_num_params(const vmc_plan *plan) { return plan ? plan->nparams : -1; }
int64_t vmc_plan_num_sliced(const vmc_plan *plan) { return plan ? plan->nsliced : -1; }

size_t vmc_amp_grad_workspace_size(const vmc_plan *plan, int batch) {
  if (!plan || batch < 1) return 0;
  return vmc::make_layout(plan, batch, false).total;
}

int vmc_plan_set_terms(vmc_plan *plan, int n_diag, const int32_t *diag_sites, const double *diag_vals,
                       int n_one, const int32_t *one_sites, const double *one_ops, int n_h,
                       const int32_t *h_sites, const double *h_ops, int n_v, const int32_t *v_sites,
                       const double *v_ops) {
  return vmc::set_terms_impl(plan, n_diag, diag_sites, diag_vals, n_one, one_sites, one_ops, n_h, h_sites, h_ops, n_v,
                             v_sites, v_ops, false);
}

int vmc_plan_set_terms_c(vmc_plan *plan, int n_diag, const int32_t *diag_sites, const double *diag_vals,
                         int n_one, const int32_t *one_sites, const double *one_ops, int n_h,
                         const int32_t *h_sites, const double *h_ops, int n_v, const int32_t *v_sites,
                         const double *v_ops) {
  return vmc::set_terms_impl(plan, n_diag, diag_sites, diag_vals, n_one, one_sites, one_ops, n_h, h_sites, h_ops, n_v,
                             v_sites, v_ops, true);
}

size_t vmc_cache_bytes(const vmc_plan *plan, int n_chains) {
  if (!plan || n_chains < 1) return 0;
  const size_t es = vmc::elem_size(plan->dtype);
  return vmc::align256((size_t)n_chains * plan->cache_elems * es) + (size_t)n_chains * plan->R * 8;
}

size_t vmc_sweep_workspace_size(const vmc_plan *plan, int n_chains) {
  if (!plan || n_chains < 1) return 0;
  return vmc::make_layout(plan, n_chains, plan->has_v).total;
}

int vmc_init_cache(vmc_plan *plan, const void *tensors, const int32_t *configs, int n_chains, void *cache,
                   void *workspace, size_t workspace_bytes, void *stream) {
  if (!plan || !tensors || !configs || !cache || !workspace) return vmc::fail("null argument");
  if (n_chains < 1) return vmc::fail("n_chains must be positive");
  hipStream_t st = (hipStream_t)stream;
  return vmc::ops_of(plan->dtype)->init_cache(plan, tensors, configs, n_chains, cache, workspace, workspace_bytes, st);
}

static int sweep_dispatch(vmc_plan *plan, const void *tensors, int32_t *configs, int n_chains, uint64_t seed,
                          int sweep, void *cache, double *log_amp, void *o, int8_t *p, double *eloc, void *workspace,
                          size_t workspace_bytes, void *stream, bool tr, bool es) {
  if (!plan || !tensors || !configs || !cache || !log_amp || !workspace) return vmc::fail("null argument");
  if (plan->gi) return vmc::fail("gauge-invariant plans sample with vmc_gi_mc_sweep");
  if (plan->bl) return vmc::fail("blockade plans sample with vmc_bl_mc_sweep");
  if (n_chains < 1) return vmc::fail("n_chains must be positive");
  if (es && eloc && !plan->has_terms)
    return vmc::fail("local energies requested but no terms set (vmc_plan_set_terms)");
  hipStream_t st = (hipStream_t)stream;
  return vmc::ops_of(plan->dtype)->sweep(plan, tensors, configs, n_chains, seed, sweep, cache, log_amp, o, p, eloc,
                                         workspace, workspace_bytes, st, tr, es);
}

int vmc_mc_sweep(vmc_plan *plan, const void *tensors, int32_t *configs, int n_chains, uint64_t seed, int sweep,
                 void *cache, double *log_amp, void *o, int8_t *p, double *eloc, void *workspace,
                 size_t workspace_bytes, void *stream) {
  return sweep_dispatch(plan, tensors, configs, n_chains, seed, sweep, cache, log_amp, o, p, eloc, workspace,
                        workspace_bytes, stream, true, true);
}

int vmc_local_estimate(vmc_plan *plan, const void *tensors, const int32_t *configs, int n, void *cache,
                       double *log_amp, double *eloc, void *workspace, size_t workspace_bytes, void *stream) {
  if (!plan || !tensors || !configs || !cache || !log_amp || !eloc || !workspace) return vmc::fail("null argument");
  if (!plan->has_terms) return vmc::fail("no terms set (vmc_plan_set_terms)");
  const vmc::Layout L = vmc::make_layout(plan, n, plan->has_v);
  if (workspace_bytes < L.total) return vmc::fail("workspace too small: need " + std::to_string(L.total));
  hipStream_t st = (hipStream_t)stream;
  vmc::ops_of(plan->dtype)->top_sweep(plan, tensors, configs, log_amp, n, workspace, st);
  return sweep_dispatch(plan, tensors, const_cast<int32_t *>(configs), n, 0, 0, cache, log_amp, nullptr, nullptr,
                        eloc, workspace, workspace_bytes, stream, false, true);
}

int vmc_transition(vmc_plan *plan, const void *tensors, int32_t *configs, int n_chains, uint64_t seed, int sweep,
                   const void *cache, double *log_amp, void *workspace, size_t workspace_bytes, void *stream) {
  return sweep_dispatch(plan, tensors, configs, n_chains, seed, sweep, const_cast<void *>(cache), log_amp, nullptr,
                        nullptr, nullptr, workspace, workspace_bytes, stream, true, false);
}

int vmc_estimate(vmc_plan *plan, const void *tensors, const int32_t *configs, int n_chains, void *cache,
                 const double *log_amp, void *o, int8_t *p, double *eloc, void *workspace, size_t workspace_bytes,
                 void *stream) {
  return sweep_dispatch(plan, tensors, const_cast<int32_t *>(configs), n_chains, 0, 0, cache,
                        const_cast<double *>(log_amp), o, p, eloc, workspace, workspace_bytes, stream, false, true);
}

int vmc_amp_grad(vmc_plan *plan, const void *tensors, const int32_t *configs, int batch, double *log_amp,
                 void *o, int8_t *p, void *workspace, size_t workspace_bytes, void *stream) {
  if (!plan || !tensors || !configs || !log_amp || !workspace) return vmc::fail("null argument");
  if (batch < 1) return vmc::fail("batch must be positive");
  hipStream_t st = (hipStream_t)stream;
  return vmc::ops_of(plan->dtype)->amp_grad(plan, tensors, configs, batch, log_amp, o, p, workspace, workspace_bytes,
                                            st);
}

int vmc_amp_grad_noalloc(vmc_plan *plan, const void *tensors, const int32_t *configs, int batch, double *log_amp,
                         void *o, int8_t *p, void *stream) {
  if (!plan || !tensors || !configs || !log_amp) return vmc::fail("null argument");
  if (batch < 1) return vmc::fail("batch must be positive");
  if (plan->prof) return vmc::fail("vmc_amp_grad_noalloc does not run between vmc_profile_begin and _end");
  const size_t need = vmc::make_layout(plan, batch, false).total;
  if (need > (size_t)VMC_NOALLOC_SCRATCH_BYTES)
    return vmc::fail("vmc_amp_grad_noalloc: batch " + std::to_string(batch) + " needs " + std::to_string(need) +
                     " workspace bytes, above the static scratch of " +
                     std::to_string((size_t)VMC_NOALLOC_SCRATCH_BYTES) + " bytes");
  void *ws = nullptr;
  if (vmc::scratch_address(&ws) != 0) return -1;
  return vmc::ops_of(plan->dtype)->amp_grad(plan, tensors, configs, batch, log_amp, o, p, ws,
                                            (size_t)VMC_NOALLOC_SCRATCH_BYTES, (hipStream_t)stream);
}

int vmc_gi_plan_create(vmc_plan **plan, int n_rows, int n_cols, int N, int phys_dim, int dmax, int chi, int strategy,
                       int n_sweeps, int dtype, int Qx, const int32_t *charge_of_site) {
  if (!plan || !charge_of_site) return vmc::fail("null argument");
  if (n_rows < 1 || n_cols < 1 || N < 2 || phys_dim < 1 || dmax < 1)
    return vmc::fail("lattice shape, N >= 2, phys_dim and dmax must be positive");
  if (n_rows * (n_cols - 1) + (n_rows - 1) * n_cols > vmc::GI_MAX_LINKS) return vmc::fail("too many links");
  if (n_cols > 64) return vmc::fail("n_cols above 64 is not supported");
  if (dtype != VMC_F32 && dtype != VMC_F64 && dtype != VMC_C64 && dtype != VMC_C128)
    return vmc::fail("dtype must be VMC_F32, VMC_F64, VMC_C64 or VMC_C128");
  if (strategy != VMC_ZIPUP && strategy != VMC_NO_TRUNCATION && strategy != VMC_DENSITY_MATRIX &&
      strategy != VMC_VARIATIONAL)
    return vmc::fail("unsupported strategy");
  if (strategy == VMC_VARIATIONAL && n_sweeps < 0) return vmc::fail("n_sweeps must be non-negative");
  if (strategy == VMC_NO_TRUNCATION) chi = 1 << 30;
  if (chi < 1) return vmc::fail("chi must be positive");
  vmc_plan *p = new vmc_plan();
  p->R = n_rows;
  p->C = n_cols;
  p->d = phys_dim;
  p->D = dmax;
  p->chi = chi;
  p->strategy = strategy;
  p->n_sweeps = strategy == VMC_VARIATIONAL ? n_sweeps : 0;
  p->dtype = dtype;
  p->gi = true;
  p->gi_N = N;
  p->gi_Qx = Qx;
  p->nsl.resize((size_t)n_rows * n_cols);
  for (int r = 0; r < n_rows; ++r)
    for (int c = 0; c < n_cols; ++c) {
      const int nl = (r > 0) + (r < n_rows - 1) + (c > 0) + (c < n_cols - 1);
      int nc = 1;
      for (int i = 1; i < nl; ++i) nc *= N;
      p->nsl[r * n_cols + c] = phys_dim * nc;
      if (phys_dim * nc > 127) {
        delete p;
        return vmc::fail("phys_dim * N^(links-1) above 127 does not fit the int8 slice indices");
      }
    }
  std::vector<int32_t> ch(charge_of_site, charge_of_site + phys_dim);
  for (auto &q : ch) q = ((q % N) + N) % N;
  // matter states of each charge, ascending (gi/model.py:190-206)
  std::vector<int32_t> deg(N, 0);
  for (int32_t q : ch) ++deg[q];
  int maxdeg = 1;
  for (int32_t v : deg) maxdeg = std::max(maxdeg, (int)v);
  std::vector<int32_t> ctab((size_t)N * maxdeg, -1), fill_n(N, 0);
  for (int s = 0; s < phys_dim; ++s) ctab[(size_t)ch[s] * maxdeg + fill_n[ch[s]]++] = s;
  p->gi_maxdeg = maxdeg;
  p->gi_all_charges = true;
  for (int32_t v : deg) p->gi_all_charges = p->gi_all_charges && v > 0;
  auto cleanup = [&]() {
    if (p->geo_dev) (void)hipFree(p->geo_dev);
    if (p->gi_charge) (void)hipFree(p->gi_charge);
    if (p->gi_ctab) (void)hipFree(p->gi_ctab);
    if (p->gi_cdeg) (void)hipFree(p->gi_cdeg);
    delete p;
  };
  if (hipMalloc(&p->gi_charge, ch.size() * 4) != hipSuccess ||
      hipMemcpy(p->gi_charge, ch.data(), ch.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMalloc(&p->gi_ctab, ctab.size() * 4) != hipSuccess ||
      hipMemcpy(p->gi_ctab, ctab.data(), ctab.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMalloc(&p->gi_cdeg, deg.size() * 4) != hipSuccess ||
      hipMemcpy(p->gi_cdeg, deg.data(), deg.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
    cleanup();
    return vmc::fail("hipMalloc failed");
  }
  if (vmc::build_plan(p) != 0) {
    cleanup();
    return -1;
  }
  *plan = p;
  return 0;
}

int vmc_gi_plan_set_degeneracy(vmc_plan *plan, const int32_t *degeneracy_per_charge) {
  if (!plan || !degeneracy_per_charge) return vmc::fail("null argument");
  if (!plan->gi) return vmc::fail("not a gauge-invariant plan");
  const int N = plan->gi_N;
  int dmax = 0;
  bool masked = false;
  for (int q = 0; q < N; ++q) {
    const int d = degeneracy_per_charge[q];
    if (d < 1 || d > plan->D) return vmc::fail("degeneracy_per_charge entries must lie in [1, dmax]");
    dmax = std::max(dmax, d);
    masked = masked || d < plan->D;
  }
  if (dmax != plan->D) return vmc::fail("the largest degeneracy_per_charge entry must equal the plan's dmax");
  if (!plan->gi_deg) VMC_HIP(hipMalloc(&plan->gi_deg, (size_t)N * 4));
  VMC_HIP(hipMemcpy(plan->gi_deg, degeneracy_per_charge, (size_t)N * 4, hipMemcpyHostToDevice));
  plan->gi_masked = masked;
  return 0;
}

int vmc_gi_set_terms(vmc_plan *plan, const double *plaquette_coeff, const double *link_diag) {
  if (!plan || !plan->gi) return vmc::fail("not a gauge-invariant plan");
  const int R = plan->R, C = plan->C, N = plan->gi_N;
  const int np = (R - 1) * (C - 1), nl = R * (C - 1) + (R - 1) * C;
  if (plan->gi_plaq) (void)hipFree(plan->gi_plaq);
  if (plan->gi_link) (void)hipFree(plan->gi_link);
  plan->gi_plaq = nullptr;
  plan->gi_link = nullptr;
  plan->gi_has_plaq = false;
  plan->gi_has_link = false;
  plan->gi_plaq_row.assign(R, 0);
  std::vector<double> pc(std::max(np, 1), NAN);
  if (plaquette_coeff)
    for (int i = 0; i < np; ++i) {
      pc[i] = plaquette_coeff[i];
      if (!std::isnan(pc[i])) {
        plan->gi_has_plaq = true;
        plan->gi_plaq_row[i / (C - 1)] = 1;
      }
    }
  std::vector<double> lt((size_t)std::max(nl, 1) * N, 0.0);
  if (link_diag) {
    std::copy(link_diag, link_diag + (size_t)nl * N, lt.begin());
    plan->gi_has_link = true;
  }
  VMC_HIP(hipMalloc(&plan->gi_plaq, pc.size() * 8));
  VMC_HIP(hipMemcpy(plan->gi_plaq, pc.data(), pc.size() * 8, hipMemcpyHostToDevice));
  VMC_HIP(hipMalloc(&plan->gi_link, lt.size() * 8));
  VMC_HIP(hipMemcpy(plan->gi_link, lt.data(), lt.size() * 8, hipMemcpyHostToDevice));
  plan->has_terms = plan->gi_has_plaq || plan->gi_has_link || plan->gi_has_site || plan->has_one || plan->has_h ||
                    plan->has_v;
  return 0;
}

int vmc_gi_slices(vmc_plan *plan, const int32_t *samples, int batch, int32_t *slices, uint8_t *valid,
                  void *stream) {
  if (!plan || !plan->gi) return vmc::fail("not a gauge-invariant plan");
  if (!samples || !slices) return vmc::fail("null argument");
  if (batch < 1) return vmc::fail("batch must be positive");
  vmc::gi_slices(plan, samples, batch, slices, valid, (hipStream_t)stream);
  return 0;
}

size_t vmc_gi_workspace_size(const vmc_plan *plan, int batch) {
  if (!plan || !plan->gi || batch < 1) return 0;
  return vmc::make_layout(plan, batch, true).total;
}

int vmc_gi_amp_grad(vmc_plan *plan, const void *tensors, const int32_t *samples, int batch, double *log_amp, void *o,
                    int8_t *p, void *workspace, size_t workspace_bytes, void *stream) {
  if (!plan || !plan->gi) return vmc::fail("not a gauge-invariant plan");
  if (!tensors || !samples || !log_amp || !workspace) return vmc::fail("null argument");
  if (batch < 1) return vmc::fail("batch must be positive");
  return vmc::ops_of(plan->dtype)->gi_amp_grad(plan, tensors, samples, batch, log_amp, o, p, workspace,
                                               workspace_bytes, (hipStream_t)stream);
}

int vmc_gi_init_cache(vmc_plan *plan, const void *tensors, const int32_t *samples, int n_chains, void *cache,
                      void *workspace, size_t workspace_bytes, void *stream) {
  if (!plan || !plan->gi) return vmc::fail("not a gauge-invariant plan");
  if (!tensors || !samples || !cache || !workspace) return vmc::fail("null argument");
  if (n_chains < 1) return vmc::fail("n_chains must be positive");
  return vmc::ops_of(plan->dtype)->gi_init_cache(plan, tensors, samples, n_chains, cache, workspace, workspace_bytes,
                                                 (hipStream_t)stream);
}

int vmc_gi_mc_sweep(vmc_plan *plan, const void *tensors, int32_t *samples, int n_chains, uint64_t seed, int sweep,
                    void *cache, double *log_amp, void *o, int8_t *p, double *eloc, void *workspace,
                    size_t workspace_bytes, void *stream) {
  if (!plan || !plan->gi) return vmc::fail("not a gauge-invariant plan");
  if (plan->d > 1 && !plan->gi_all_charges) return vmc::fail("charge_of_site must include all charges 0..N-1.");
  if (!tensors || !samples || !cache || !log_amp || !workspace) return vmc::fail("null argument");
  if (n_chains < 1) return vmc::fail("n_chains must be positive");
  if (eloc && !plan->has_terms) return vmc::fail("local energies requested but no terms set (vmc_gi_set_terms)");
  return vmc::ops_of(plan->dtype)->gi_sweep(plan, tensors, samples, n_chains, seed, sweep, cache, log_amp, o, p, eloc,
                                            workspace, workspace_bytes, (hipStream_t)stream);
}

int vmc_bl_plan_create(vmc_plan **plan, int n_rows, int n_cols, int D0, int D1, int chi, int strategy, int n_sweeps,
                       int dtype) {
  if (!plan) return vmc::fail("null argument");
  if (n_rows < 1 || n_cols < 1 || D0 < 1 || D1 < 1) return vmc::fail("lattice shape, D0 and D1 must be positive");
  if (n_rows * n_cols > vmc::BL_MAX_SITES) return vmc::fail("more than 1024 sites is not supported");
  if (n_cols > 64) return vmc::fail("n_cols above 64 is not supported");
  if (dtype != VMC_F32 && dtype != VMC_F64 && dtype != VMC_C64 && dtype != VMC_C128)
    return vmc::fail("dtype must be VMC_F32, VMC_F64, VMC_C64 or VMC_C128");
  if (strategy != VMC_ZIPUP && strategy != VMC_NO_TRUNCATION && strategy != VMC_DENSITY_MATRIX &&
      strategy != VMC_VARIATIONAL)
    return vmc::fail("unsupported strategy");
  if (strategy == VMC_VARIATIONAL && n_sweeps < 0) return vmc::fail("n_sweeps must be non-negative");
  if (strategy == VMC_NO_TRUNCATION) chi = 1 << 30;
  if (chi < 1) return vmc::fail("chi must be positive");
  vmc_plan *p = new vmc_plan();
  p->R = n_rows;
  p->C = n_cols;
  p->d = 2;
  p->D = std::max(D0, D1);
  p->chi = chi;
  p->strategy = strategy;
  p->n_sweeps = strategy == VMC_VARIATIONAL ? n_sweeps : 0;
  p->dtype = dtype;
  p->bl = true;
  p->bl_D0 = D0;
  p->bl_D1 = D1;
  p->nsl.resize((size_t)n_rows * n_cols);
  for (int r = 0; r < n_rows; ++r)
    for (int c = 0; c < n_cols; ++c) p->nsl[r * n_cols + c] = 2 << ((r > 0) + (c > 0));
  if (vmc::build_plan(p) != 0) {
    if (p->geo_dev) (void)hipFree(p->geo_dev);
    delete p;
    return -1;
  }
  *plan = p;
  return 0;
}

int vmc_bl_slices(vmc_plan *plan, const int32_t *occupations, int batch, int32_t *slices, uint8_t *valid,
                  void *stream) {
  if (!plan || !plan->bl) return vmc::fail("not a blockade plan");
  if (!occupations || !slices) return vmc::fail("null argument");
  if (batch < 1) return vmc::fail("batch must be positive");
  vmc::bl_slices(plan, occupations, batch, slices, valid, (hipStream_t)stream);
  return 0;
}

size_t vmc_bl_workspace_size(const vmc_plan *plan, int batch) {
  if (!plan || !plan->bl || batch < 1) return 0;
  return vmc::make_layout(plan, batch, true).total;
}

int vmc_bl_amp_grad(vmc_plan *plan, const void *tensors, const int32_t *occupations, int batch, double *log_amp,
                    void *o, int8_t *p, void *workspace, size_t workspace_bytes, void *stream) {
  if (!plan || !plan->bl) return vmc::fail("not a blockade plan");
  if (!tensors || !occupations || !log_amp || !workspace) return vmc::fail("null argument");
  if (batch < 1) return vmc::fail("batch must be positive");
  return vmc::ops_of(plan->dtype)->gi_amp_grad(plan, tensors, occupations, batch, log_amp, o, p, workspace,
                                               workspace_bytes, (hipStream_t)stream);
}

int vmc_bl_init_cache(vmc_plan *plan, const void *tensors, const int32_t *occupations, int n_chains, void *cache,
                      void *workspace, size_t workspace_bytes, void *stream) {
  if (!plan || !plan->bl) return vmc::fail("not a blockade plan");
  if (!tensors || !occupations || !cache || !workspace) return vmc::fail("null argument");
  if (n_chains < 1) return vmc::fail("n_chains must be positive");
  return vmc::ops_of(plan->dtype)->gi_init_cache(plan, tensors, occupations, n_chains, cache, workspace,
                                                 workspace_bytes, (hipStream_t)stream);
}

int vmc_bl_mc_sweep(vmc_plan *plan, const void *tensors, int32_t *occupations, int n_chains, uint64_t seed, int sweep,
                    void *cache, double *log_amp, void *o, int8_t *p, double *eloc, void *workspace,
                    size_t workspace_bytes, void *stream) {
  if (!plan || !plan->bl) return vmc::fail("not a blockade plan");
  if (!tensors || !occupations || !cache || !log_amp || !workspace) return vmc::fail("null argument");
  if (n_chains < 1) return vmc::fail("n_chains must be positive");
  if (eloc && !plan->has_terms) return vmc::fail("local energies requested but no terms set (vmc_plan_set_terms)");
  if (eloc && (plan->has_h || plan->has_v))
    return vmc::fail("blockade local energies take diagonal and one-site terms only");
  return vmc::ops_of(plan->dtype)->bl_sweep(plan, tensors, occupations, n_chains, seed, sweep, cache, log_amp, o, p,
                                            eloc, workspace, workspace_bytes, (hipStream_t)stream);
}

static_assert(VMC_TUNE_TRUNC == vmc::TK_TRUNC && VMC_TUNE_THETA == vmc::TK_THETA, "tuning kinds");
static_assert(VMC_TRUNC_LPE8 == vmc::TRUNC_LPE8 && VMC_TRUNC_LPE16 == vmc::TRUNC_LPE16 &&
                  VMC_TRUNC_GENERIC == vmc::TRUNC_GENERIC, "truncation variants");
static_assert(VMC_TILE_AUTO == vmc::GEMM_TILE_AUTO && VMC_TILE_WIDE == vmc::GEMM_TILE_WIDE &&
                  VMC_TILE_TALL == vmc::GEMM_TILE_TALL && VMC_TILE_SQUARE == vmc::GEMM_TILE_SQUARE, "tiles");
static_assert(VMC_TUNE_ENTRY_INTS == vmc::TUNE_KEY + 3, "tuning entry layout");

int vmc_plan_autotune(vmc_plan *plan, const void *tensors, const int32_t *samples, int batch, double *log_amp,
                      void *o, int8_t *p, void *workspace, size_t workspace_bytes, int reps, void *stream) {
  if (!plan || !tensors || !samples || !log_amp || !workspace) return vmc::fail("null argument");
  if (batch < 1) return vmc::fail("batch must be positive");
  if (reps < 1 || reps > 100) return vmc::fail("reps must be in 1..100");
  if (plan->prof) return vmc::fail("vmc_plan_autotune does not run between vmc_profile_begin and _end");
  hipStream_t st = (hipStream_t)stream;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  VMC_HIP(hipStreamIsCapturing(st, &cap));
  if (cap != hipStreamCaptureStatusNone) return vmc::fail("vmc_plan_autotune cannot run under stream capture");
  vmc::Tuner tu;
  tu.reps = reps;
  const auto saved = plan->tune;
  int rc = 0;
  if (hipEventCreate(&tu.e0) != hipSuccess || hipEventCreate(&tu.e1) != hipSuccess ||
      hipMalloc((void **)&tu.log, (size_t)batch * sizeof(double)) != hipSuccess)
    rc = vmc::fail("vmc_plan_autotune: could not create its events and log-scale buffer");
  if (rc == 0) {
    plan->tuner = &tu;
    const vmc::EngineOps *ops = vmc::ops_of(plan->dtype);
    rc = plan->gi || plan->bl
             ? ops->gi_amp_grad(plan, tensors, samples, batch, log_amp, o, p, workspace, workspace_bytes, st)
                  : ops->amp_grad(plan, tensors, samples, batch, log_amp, o, p, workspace, workspace_bytes, st);
    plan->tuner = nullptr;
    if (hipStreamSynchronize(st) != hipSuccess && rc == 0) rc = vmc::fail("vmc_plan_autotune: stream failed");
    if (rc == 0 && tu.err) rc = tu.err;
  }
  if (tu.log) (void)hipFree(tu.log);
  if (tu.gscr) (void)hipFree(tu.gscr);
  if (tu.e0) (void)hipEventDestroy(tu.e0);
  if (tu.e1) (void)hipEventDestroy(tu.e1);
  if (rc != 0) plan->tune = saved;
  vmc::refresh_gtr(plan);
  return rc;
}

int vmc_plan_tuning(const vmc_plan *plan, int32_t *entries, int capacity) {
  if (!plan) return vmc::fail("plan is null");
  if (capacity > 0 && !entries) return vmc::fail("null argument");
  int i = 0;
  for (const auto &e : plan->tune) {
    if (i < capacity) {
      int32_t *row = entries + (int64_t)i * VMC_TUNE_ENTRY_INTS;
      for (int j = 0; j < vmc::TUNE_KEY; ++j) row[j] = e.first[j];
      row[vmc::TUNE_KEY] = e.second.choice;
      row[vmc::TUNE_KEY + 1] = e.second.t_choice_ns;
      row[vmc::TUNE_KEY + 2] = e.second.t_default_ns;
    }
    ++i;
  }
  return i;
}

int vmc_plan_set_tuning(vmc_plan *plan, const int32_t *entries, int count) {
  if (!plan) return vmc::fail("plan is null");
  if (count < 0 || (count > 0 && !entries)) return vmc::fail("invalid tuning entries");
  std::map<std::array<int, vmc::TUNE_KEY>, vmc::TuneVal> t;
  for (int i = 0; i < count; ++i) {
    const int32_t *row = entries + (int64_t)i * VMC_TUNE_ENTRY_INTS;
    std::array<int, vmc::TUNE_KEY> key;
    for (int j = 0; j < vmc::TUNE_KEY; ++j) key[j] = row[j];
    const int choice = row[vmc::TUNE_KEY];
    if (key[0] == vmc::TK_TRUNC) {
      if (!vmc::trunc_choice_ok(plan, key[1], key[2], key[3], key[4], choice) || key[3] >= key[1] || key[3] < 1)
        return vmc::fail("tuning entry " + std::to_string(i) + ": truncation choice " + std::to_string(choice) +
                         " does not run a " + std::to_string(key[1]) + "x" + std::to_string(key[2]) + " cut keeping " +
                         std::to_string(key[3]));
    } else if (key[0] == vmc::TK_THETA) {
      if (choice < 0 || choice > 15) return vmc::fail("tuning entry " + std::to_string(i) + ": tile pair out of range");
    } else {
      return vmc::fail("tuning entry " + std::to_string(i) + ": unknown kind " + std::to_string(key[0]));
    }
    t[key] = vmc::TuneVal{choice, row[vmc::TUNE_KEY + 1], row[vmc::TUNE_KEY + 2]};
  }
  plan->tune = std::move(t);
  vmc::refresh_gtr(plan);
  return 0;
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
