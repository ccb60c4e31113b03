/*
 * vmc_amd.h -- C ABI of the MI355X (gfx950) PEPS amplitude engine.
 *
 * Plain pointers and sizes only. Device pointers are HIP device memory owned by
 * the caller; every compute call is asynchronous on the given HIP stream and
 * returns 0 on success or a negative status, with a message available from
 * vmc_last_error() (thread-local). No C++ exception crosses this boundary.
 *
 * Tensor layout (the reference's ravel_pytree order of PEPS.tensors):
 *   sites in row-major order, each site tensor A[r][c] of shape
 *   (d, up, down, left, right), C-contiguous, edge bonds of size 1
 *   (reference: src/vmc/peps/standard/model.py:27-35, 57-74).
 * Configurations: int32 occupancies in [0, d), shape (batch, n_rows*n_cols),
 *   row-major (reference: src/vmc/utils/utils.py:55-57).
 * Log-derivatives o: per sample the concatenation over sites (row-major) of the
 *   env-gradient (up, down, left, right) divided by the amplitude -- the
 *   "sliced" layout of src/vmc/peps/standard/kernels.py:54-88.
 * Amplitudes are returned as log psi = log|psi| + i*phase (two doubles per
 *   sample) because at 10x10 lattices |psi| underflows FP32 and nears the FP64
 *   range.
 */
#ifndef VMC_AMD_H
#define VMC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vmc_plan vmc_plan;

/* Storage/arithmetic of tensors, boundaries and contractions (complex types are
 * interleaved re, im like numpy complex64 / complex128). Truncations always run
 * in FP64 (complex FP64 for complex plans). */
enum vmc_dtype { VMC_F32 = 0, VMC_F64 = 1, VMC_C64 = 2, VMC_C128 = 3 };

/* Boundary-MPS strategies (reference: src/vmc/peps/common/strategy.py:44-92).
 * VMC_ZIPUP keeps the top-min(chi, m, n) left singular vectors U of each m x n
 * theta (new tensor U, carry S V^H; strategy.py:128-150). VMC_DENSITY_MATRIX
 * keeps the top-min(chi, n) eigenvectors V of theta^H theta (new tensor
 * theta V, carry V^H; strategy.py:153-176); where rank(theta) < k < n the kept
 * null-space vectors are arbitrary and a later truncating cut depends on them,
 * as in the reference. VMC_NO_TRUNCATION is zip-up with chi large enough that
 * nothing is discarded; VMC_VARIATIONAL is the QR-initialised one-site ALS
 * compression with n_sweeps sweeps (strategy.py:74-92, 179-427). */
enum vmc_strategy { VMC_NO_TRUNCATION = 0, VMC_ZIPUP = 1, VMC_DENSITY_MATRIX = 2, VMC_VARIATIONAL = 3 };

/* Create a plan for an open-boundary n_rows x n_cols PEPS with physical
 * dimension phys_dim, bond dimension bond_dim and boundary bond chi.
 * n_sweeps is Variational's sweep count (ignored by the other strategies).
 * Replaces the static configuration captured by build_mc_kernels
 * (src/vmc/peps/standard/kernels.py:91-113) and PEPS.__init__
 * (src/vmc/peps/standard/model.py:37-74). */
int vmc_plan_create(vmc_plan **plan, int n_rows, int n_cols, int phys_dim,
                    int bond_dim, int chi, int strategy, int n_sweeps, int dtype);
int vmc_plan_destroy(vmc_plan *plan);

/* Number of site-tensor elements (d * sum of params per site) and of sliced
 * log-derivative entries per sample (sum of params per site,
 * src/vmc/utils/smallo.py:23-35). */
int64_t vmc_plan_num_params(const vmc_plan *plan);
int64_t vmc_plan_num_sliced(const vmc_plan *plan);

/* Device workspace bytes needed by vmc_amp_grad for `batch` samples. */
size_t vmc_amp_grad_workspace_size(const vmc_plan *plan, int batch);

/* Amplitude and sliced log-derivative of a batch of configurations:
 * the work of _value_and_grad(model, samples, full_gradient=False)
 * (src/vmc/peps/standard/compat.py:94-133) -- top boundary sweep with cached
 * boundaries (_forward_with_cache, src/vmc/peps/common/contraction.py:38-56)
 * then the bottom env-gradient sweep (_compute_all_gradients,
 * src/vmc/peps/common/energy.py:276-301).
 *   tensors  : device, vmc_plan_num_params elements of the plan dtype
 *   configs  : device int32 (batch, n_sites)
 *   log_amp  : device double (batch, 2) = (log|psi|, phase)
 *   o        : device, (batch, vmc_plan_num_sliced) of the plan dtype, or NULL
 *              for amplitudes only (_value, compat.py:67-78)
 *   p        : device int8 (batch, vmc_plan_num_sliced) active slice index, or NULL
 */
int vmc_amp_grad(vmc_plan *plan, const void *tensors, const int32_t *configs,
                 int batch, double *log_amp, void *o, int8_t *p,
                 void *workspace, size_t workspace_bytes, void *stream);

/* Bytes of the static device scratch behind vmc_amp_grad_noalloc (per GPU). */
#define VMC_NOALLOC_SCRATCH_BYTES (64u << 20)

/* vmc_amp_grad without a workspace argument and without any allocation. It has
 * no counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The workspace is a file-scope device array of
 * VMC_NOALLOC_SCRATCH_BYTES on each GPU the library runs on, held until the
 * process ends; batches whose vmc_amp_grad_workspace_size exceeds it are
 * refused with an error. Every call on one GPU shares that GPU's array, so
 * calls that overlap on one GPU (other streams or host threads) race; serialise
 * them. Outputs and inputs as vmc_amp_grad; the stream must belong to the
 * current device. Not available between vmc_profile_begin and vmc_profile_end. */
int vmc_amp_grad_noalloc(vmc_plan *plan, const void *tensors, const int32_t *configs, int batch, double *log_amp,
                         void *o, int8_t *p, void *stream);

/* Local operator terms for the estimate pass (reference: the term types and
 * bucket_terms of src/vmc/operators/local_terms.py:22-200, consumed by
 * _compute_all_env_grads_and_energy, src/vmc/peps/common/energy.py:123-224).
 * Sites are flat row-major indices. Operators are dense row-major
 * <out|op|in> matrices (d x d one-site, d^2 x d^2 two-site with the first
 * site's index major); the energy applies op[:, s] exactly as the reference
 * does. Diagonal terms: diag_sites holds two site ids per term (second = -1
 * for a one-site diagonal) and diag_vals d^2 doubles per term (d used for one
 * site). Several terms at the same location are summed. Copies the tables to
 * the device (allocates); call it once before sampling. On a gauge-invariant
 * plan the terms act on the matter states with the links held fixed (site
 * index s = matter state; gi/model.py:422-431 diagonal, MatterMassTerm of
 * gi/local_terms.py:39-52, one-site / horizontal / vertical operators of
 * gi/model.py:474-536) and add to the terms of vmc_gi_set_terms. */
int vmc_plan_set_terms(vmc_plan *plan, int n_diag, const int32_t *diag_sites, const double *diag_vals,
                       int n_one, const int32_t *one_sites, const double *one_ops,
                       int n_h, const int32_t *h_sites, const double *h_ops,
                       int n_v, const int32_t *v_sites, const double *v_ops);

/* vmc_plan_set_terms with complex operator values: every value array
 * (diag_vals, one_ops, h_ops, v_ops) holds complex128 entries as interleaved
 * (re, im) doubles, in the same layout and counts. The reference's terms take
 * any dtype (src/vmc/operators/local_terms.py:22-200), e.g. the complex
 * hoppings of the Hofstadter model; the energy still applies op[:, s]
 * (energy.py:88, 179, 200), without conjugation. Values with a nonzero
 * imaginary part need a complex plan (VMC_C64 / VMC_C128). */
int vmc_plan_set_terms_c(vmc_plan *plan, int n_diag, const int32_t *diag_sites, const double *diag_vals,
                         int n_one, const int32_t *one_sites, const double *one_ops,
                         int n_h, const int32_t *h_sites, const double *h_ops,
                         int n_v, const int32_t *v_sites, const double *v_ops);

/* Per-chain sampler cache: the bottom boundary of every row (Cache.bottom_envs,
 * src/vmc/peps/standard/kernels.py:32-35), caller-owned device memory. */
size_t vmc_cache_bytes(const vmc_plan *plan, int n_chains);
size_t vmc_sweep_workspace_size(const vmc_plan *plan, int n_chains);

/* init_cache (src/vmc/peps/standard/kernels.py:115-129): bottom boundaries of
 * each chain's configuration, built bottom-up with the plan's strategy. */
int vmc_init_cache(vmc_plan *plan, const void *tensors, const int32_t *configs, int n_chains,
                   void *cache, void *workspace, size_t workspace_bytes, void *stream);

/* One sweep of every chain: transition (sequential top-to-bottom Metropolis
 * over all sites, src/vmc/peps/standard/kernels.py:131-209) followed by
 * estimate (env-gradients, local energy, new bottom cache; kernels.py:211-245).
 * Draw for site i of sweep `sweep` of chain b: Philox4x32-10 counter
 * (i, sweep, b, 0) under key `seed` (replaces JAX threefry key splitting).
 *   configs : device int32 (n_chains, n_sites), updated in place
 *   cache   : device, vmc_cache_bytes, read (transition) and rewritten (estimate)
 *   log_amp : device double (n_chains, 2): log psi of the new configuration
 *             from the truncated top sweep (Context.amp)
 *   o, p    : sliced log-derivatives and active slice indices, or NULL
 *   eloc    : device double (n_chains) local energies (complex plans: n_chains
 *             (re, im) pairs), or NULL */
int vmc_mc_sweep(vmc_plan *plan, const void *tensors, int32_t *configs, int n_chains, uint64_t seed,
                 int sweep, void *cache, double *log_amp, void *o, int8_t *p, double *eloc,
                 void *workspace, size_t workspace_bytes, void *stream);

/* The two halves of vmc_mc_sweep, for callers that keep the reference's
 * transition / estimate split (kernels.py:131-209 and :211-245). The
 * transition leaves the top boundaries of every row (Context.top_envs) in
 * `workspace`; the next vmc_estimate on the same plan, chains and workspace
 * consumes them together with the transition's log_amp (Context.amp). */
int vmc_transition(vmc_plan *plan, const void *tensors, int32_t *configs, int n_chains, uint64_t seed,
                   int sweep, const void *cache, double *log_amp, void *workspace,
                   size_t workspace_bytes, void *stream);
int vmc_estimate(vmc_plan *plan, const void *tensors, const int32_t *configs, int n_chains, void *cache,
                 const double *log_amp, void *o, int8_t *p, double *eloc, void *workspace,
                 size_t workspace_bytes, void *stream);

/* local_estimate (src/vmc/peps/standard/compat.py:136-194): local energies of
 * given configurations from a top sweep with cached boundaries and the
 * bottom-up energy pass. `cache` (vmc_cache_bytes) receives the bottom
 * boundaries; log_amp receives log psi of the top sweep used as denominator. */
int vmc_local_estimate(vmc_plan *plan, const void *tensors, const int32_t *configs, int n, void *cache,
                       double *log_amp, double *eloc, void *workspace, size_t workspace_bytes, void *stream);

/* SR / TDVP update (SRPreconditioner.apply with DirectSolve = Cholesky,
 * src/vmc/preconditioners/preconditioners.py:117-163, 268-330, over the sliced
 * Jacobian of src/vmc/qgt/jacobian.py:56-93 and the QGT of src/vmc/qgt/qgt.py).
 * FP64 on the device; complex128 when o, eloc or the grad factor is complex.
 *   o, p           : (n_samples, n_sliced) log-derivatives of vmc_dtype o_dtype
 *                    and int8 active slices, as produced by vmc_mc_sweep
 *   params_per_site: n_sites entries summing to n_sliced
 *   slices_per_site: n_sites slice counts (GI-PEPS: phys_dim * nc,
 *                    gi/model.py:1065-1080), each <= phys_dim; NULL means
 *                    phys_dim everywhere
 *   eloc           : (n_samples) local energies, FP64 or (eloc_complex) complex128
 *   space          : 0 ParameterSpace (S = O^H O / Ns, d*n_sliced unknowns)
 *                    1 SampleSpace (T = O O^H / Ns, n_samples unknowns)
 *   grad_factor    : factor of dv, re + i*im (-1 imaginary time, -i real time)
 *   real_params    : x_out is real FP64: 2 Re(x) for a real grad factor,
 *                    Re(x) for a complex one (the reference's tree_cast);
 *                    otherwise x_out has the solve's type (complex128 for a
 *                    complex solve)
 *   x_out          : device (phys_dim * n_sliced), site-major update
 *                    (the flattened parameter pytree order)
 *   metrics_out    : device double[2] = (residual_error, Cholesky info) or NULL
 * vmc_sr_workspace_size: complex_solve = 1 when the solve is complex. */
size_t vmc_sr_workspace_size(int n_samples, int64_t n_sliced, int phys_dim, int space, int complex_solve);
int vmc_sr_solve(int n_samples, int64_t n_sliced, int phys_dim, const int32_t *params_per_site, int n_sites,
                 const int32_t *slices_per_site, const void *o, int o_dtype, const int8_t *p, const double *eloc,
                 int eloc_complex, int space, double diag_shift, double grad_factor_re, double grad_factor_im,
                 int real_params, void *x_out, double *metrics_out, void *workspace, size_t workspace_bytes,
                 void *stream);

/* Linear solvers of DirectSolve (src/vmc/qgt/solvers.py:16-37) for
 * vmc_sr_solve_ex; vmc_sr_solve is VMC_SOLVE_CHOLESKY. */
enum vmc_solver { VMC_SOLVE_CHOLESKY = 0, VMC_SOLVE_CG = 1, VMC_SOLVE_SVD = 2 };

/* vmc_sr_solve with a choice of solver for (M + diag_shift) z = b, M = S or T:
 *   VMC_SOLVE_CHOLESKY : solve_cholesky (solvers.py:24-27);
 *   VMC_SOLVE_CG       : solve_cg (solvers.py:16-21) = conjugate gradients from
 *                        z = 0 until |r| <= solver_tol |b| or maxiter steps
 *                        (the reference defaults: tol 1e-5, maxiter 1000);
 *                        metrics_out[1] = steps taken;
 *   VMC_SOLVE_SVD      : solve_svd (solvers.py:30-37) with rcond = solver_tol.
 *                        M + diag_shift has every eigenvalue >= diag_shift, so
 *                        when diag_shift > rcond * (Gershgorin bound of the
 *                        largest eigenvalue) no singular value is cut and the
 *                        result is the Cholesky solve; otherwise the call fails
 *                        (no dense eigensolver is built).
 * Other arguments and the workspace size as vmc_sr_solve. */
int vmc_sr_solve_ex(int n_samples, int64_t n_sliced, int phys_dim, const int32_t *params_per_site, int n_sites,
                    const int32_t *slices_per_site, const void *o, int o_dtype, const int8_t *p, const double *eloc,
                    int eloc_complex, int space, double diag_shift, double grad_factor_re, double grad_factor_im,
                    int real_params, int solver, double solver_tol, int maxiter, void *x_out, double *metrics_out,
                    void *workspace, size_t workspace_bytes, void *stream);

/* SR update with QRSolve (src/vmc/preconditioners/preconditioners.py:166-214):
 * ParameterSpace over the full Jacobian o (n_samples x n_params, vmc_dtype
 * o_dtype; the reference has no sliced or sample-space QRSolve). Pivoted
 * Householder QR of O_c = o - mean (xGEQP3 pivot rule: largest partial column
 * norm, first on ties), rank r = #{|R_jj| > tol |R_00|} with tol = rcond, or
 * n_params * eps when rcond <= 0 (the default rtol max(r.shape) * eps of the
 * reference's jnp.linalg.matrix_rank(r), r being min(n_samples, n_params) x
 * n_params; the reference thresholds singular values of R, this the pivoted
 * diagonal, which agree unless a singular value lies within the diagonal's
 * bracketing of the tolerance), then
 *   min_norm: x = P [a - B x2; x2], (B^H B + I) x2 = B^H a, a = R11^-1 y1,
 *             B = R11^-1 R12, y = Q^H dv;   else x = P [R11^-1 y1; 0].
 * diag_shift is not used (as in the reference). x_out and real_params as in
 * vmc_sr_solve; metrics_out = (residual |O^H O x - O^H dv|^2 / |O^H dv|^2, r). */
size_t vmc_sr_qr_workspace_size(int n_samples, int64_t n_params, int complex_solve);
int vmc_sr_qr_solve(int n_samples, int64_t n_params, const void *o, int o_dtype, const double *eloc,
                    int eloc_complex, double grad_factor_re, double grad_factor_im, double rcond, int min_norm,
                    int real_params, void *x_out, double *metrics_out, void *workspace, size_t workspace_bytes,
                    void *stream);

/* QGT.to_dense (src/vmc/qgt/qgt.py:173-223) of the centred sliced Jacobian:
 * space 0 -> S (d*n_sliced square, site-major), space 1 -> T (n_samples
 * square, or the expanded width of slices_per_site). dense_out is device
 * FP64, or complex128 for a complex o_dtype;
 * workspace as vmc_sr_workspace_size. */
int vmc_qgt_dense(int n_samples, int64_t n_sliced, int phys_dim, const int32_t *params_per_site, int n_sites,
                  const int32_t *slices_per_site, const void *o, int o_dtype, const int8_t *p, int space,
                  void *dense_out, void *workspace, size_t workspace_bytes, void *stream);

/* ---- Gauge-invariant PEPS (src/vmc/peps/gi/*) ----
 * Z_N lattice gauge theory with links in the sample: per chain, int32 sites
 * (n_rows*n_cols), horizontal links (n_rows*(n_cols-1)) then vertical links
 * ((n_rows-1)*n_cols) (gi/model.py:121-141). Site (r, c) stores a
 * (phys_dim * nc, mu_u, mu_d, mu_l, mu_r) tensor with nc = N^(links-1)
 * and bonds of dmax (gi/model.py:99-118); the sample selects slice
 * sigma * nc + cfg (gi/model.py:308-343, gi/kernels.py:116-118).
 *
 * vmc_gi_plan_create: GIPEPS.__init__ + GIPEPSConfig (gi/model.py:40-118)
 *   with the strategy arguments of vmc_plan_create; charge_of_site has
 *   phys_dim entries. */
int vmc_gi_plan_create(vmc_plan **plan, int n_rows, int n_cols, int N, int phys_dim, int dmax, int chi,
                       int strategy, int n_sweeps, int dtype, int Qx, const int32_t *charge_of_site);
/* GIPEPSConfig.degeneracy_per_charge / mask_per_charge (gi/model.py:52-60)
 * and their use in _assemble_site (gi/model.py:292-305): N entries in
 * [1, dmax] with maximum dmax. A bond index i of a link carrying charge q is
 * kept iff i < degeneracy_per_charge[q]; the library applies these masks to
 * every slice of the caller's tensors (into the workspace) before each call.
 * Without this call every sector has dmax. Allocates once; not during
 * capture. */
int vmc_gi_plan_set_degeneracy(vmc_plan *plan, const int32_t *degeneracy_per_charge);
/* Plaquette terms ((n_rows-1)*(n_cols-1) coefficients by top-left corner,
 * NaN = none; PlaquetteTerm, src/vmc/operators/local_terms.py:127-144 as used
 * by gi/model.py:546-608) and link-diagonal terms (an (n_links, N) table,
 * horizontal links first; LinkDiagonalTerm / build_electric_terms,
 * gi/local_terms.py:13-90). Either pointer may be NULL. */
int vmc_gi_set_terms(vmc_plan *plan, const double *plaquette_coeff, const double *link_diag);
/* Combined slice index of every site and the Gauss-law check
 * (gi/compat.py:79-88); valid may be NULL. */
int vmc_gi_slices(vmc_plan *plan, const int32_t *samples, int batch, int32_t *slices, uint8_t *valid,
                  void *stream);
size_t vmc_gi_workspace_size(const vmc_plan *plan, int batch);
/* gi_apply and its VJP (gi/compat.py:18-99): log psi, sliced o = dpsi/dA / psi
 * and slice indices p; a sample violating Gauss law gets log|psi| = -inf
 * and o = 0. */
int vmc_gi_amp_grad(vmc_plan *plan, const void *tensors, const int32_t *samples, int batch, double *log_amp,
                    void *o, int8_t *p, void *workspace, size_t workspace_bytes, void *stream);
/* init_cache of gi/kernels.py:33-47 (cache as vmc_cache_bytes). */
int vmc_gi_init_cache(vmc_plan *plan, const void *tensors, const int32_t *samples, int n_chains, void *cache,
                      void *workspace, size_t workspace_bytes, void *stream);
/* One transition + estimate of every chain (gi/kernels.py:49-126): the
 * transition of gi/model.py:901-1043, i.e. the plaquette sweep of every row
 * pair (gi/model.py:623-700) and, with matter (phys_dim > 1), the horizontal
 * link sweep of every row and the vertical link sweep of every row pair
 * (gi/model.py:703-878), each over freshly rebuilt bottom boundaries. Philox
 * draws: plaquette (r, c) at counter (2^20 + r*n_cols + c, sweep, chain, 0),
 * horizontal link (r, c) at 2*2^20 + r*(n_cols-1) + c, vertical link (r, c) at
 * 3*2^20 + r*n_cols + c (block 0: accept uniform and delta = 1 + word %
 * (N-1); block 1: the two matter-state picks). Then gradients,
 * link-diagonal + plaquette energies (gi/model.py:392-612) and the new bottom
 * boundaries. samples is updated in place. With matter, every charge
 * 0..N-1 needs a matter state (the reference's ValueError, gi/model.py:92-93). */
int vmc_gi_mc_sweep(vmc_plan *plan, const void *tensors, int32_t *samples, int n_chains, uint64_t seed, int sweep,
                    void *cache, double *log_amp, void *o, int8_t *p, double *eloc, void *workspace,
                    size_t workspace_bytes, void *stream);

/* ---- Blockade PEPS (src/vmc/peps/blockade/*) ----
 * Nearest-neighbour blockade (Rydberg) PEPS: a sample is the int32 occupation
 * (0/1) of every site, row-major. Site (r, c) stores a (2 * nc, mu_u, mu_d,
 * mu_l, mu_r) tensor with nc = 2^((r > 0) + (c > 0)) and bonds of
 * max(D0, D1) (blockade/model.py:100-121); occupations n select slice
 * n * nc + cfg, cfg = kL * stride + kU for n = 0 and 0 for n = 1, with kL / kU
 * the left / upper neighbour's occupation and stride 2 below the first row
 * (model.py:157-199). With D0 != D1 the bonds of each slice are masked by their
 * sectors (kU, n, kL, n) (model.py:201-212) inside the library; o is the
 * gradient of the masked (effective) tensor, as the reference's.
 *
 * vmc_bl_plan_create: BlockadePEPS.__init__ + BlockadePEPSConfig
 *   (model.py:43-121) with the strategy arguments of vmc_plan_create. */
int vmc_bl_plan_create(vmc_plan **plan, int n_rows, int n_cols, int D0, int D1, int chi, int strategy, int n_sweeps,
                       int dtype);
/* Slice index of every site and the blockade check (compat.py:43-50); valid
 * may be NULL. */
int vmc_bl_slices(vmc_plan *plan, const int32_t *occupations, int batch, int32_t *slices, uint8_t *valid,
                  void *stream);
size_t vmc_bl_workspace_size(const vmc_plan *plan, int batch);
/* blockade_apply and its gradient (compat.py:37-62): log psi, sliced
 * o = dpsi/dA / psi and slice indices p (kernels.py:93-103); a sample with two
 * adjacent occupied sites gets log|psi| = -inf and o = 0. Workspace:
 * vmc_amp_grad_workspace_size. */
int vmc_bl_amp_grad(vmc_plan *plan, const void *tensors, const int32_t *occupations, int batch, double *log_amp,
                    void *o, int8_t *p, void *workspace, size_t workspace_bytes, void *stream);
/* init_cache of blockade/kernels.py:34-47 (cache as vmc_cache_bytes). */
int vmc_bl_init_cache(vmc_plan *plan, const void *tensors, const int32_t *occupations, int n_chains, void *cache,
                      void *workspace, size_t workspace_bytes, void *stream);
/* One transition + estimate of every chain (kernels.py:49-117): the row-pair
 * sweeps (model.py:820-994) of rows 0..n_rows-2 over the cached bottom
 * boundaries and the single-row sweep of the last row (model.py:668-817),
 * flips refused next to an occupied site (model.py:220-239); Philox draw of
 * site (r, c) at counter (4 * 2^20 + r * n_cols + c, sweep, chain, 0). Then
 * gradients, diagonal energies and the one-site energies
 * sum op[n_flip, n_cur] psi(flipped) / psi over two-row (last row: one-row)
 * windows (model.py:295-617) from the tables of vmc_plan_set_terms (its
 * horizontal and vertical terms are ignored, as by the reference's estimate),
 * and the new bottom boundaries. occupations is updated in place. */
int vmc_bl_mc_sweep(vmc_plan *plan, const void *tensors, int32_t *occupations, int n_chains, uint64_t seed, int sweep,
                    void *cache, double *log_amp, void *o, int8_t *p, double *eloc, void *workspace,
                    size_t workspace_bytes, void *stream);

/* Autotuning (no reference counterpart: XLA picks the reference's kernels when
 * it compiles _value_and_grad). vmc_plan_autotune runs one vmc_amp_grad (or, for
 * a gauge-invariant plan, vmc_gi_amp_grad) call on the caller's batch and, at
 * the first launch of each distinct shape, times every kernel variant that can
 * run it (a warm-up launch, then `reps` launches between HIP events) and records
 * the fastest in the plan; later calls of every entry point launch the recorded
 * choice. Shapes tuned:
 *   VMC_TUNE_TRUNC  a zip-up or density-matrix cut that the register kernels
 *                   take: key (m, n, k, dm), choice VMC_TRUNC_LPE8 (register
 *                   kernel, 8 Sturm lanes per eigenvalue; the default),
 *                   VMC_TRUNC_LPE16 (16 lanes, real plans) or VMC_TRUNC_GENERIC
 *                   (the kernel of cuts beyond the register limits);
 *   VMC_TUNE_THETA  a theta contraction carry . m . w: key (has_carry, k, Dl,
 *                   wl, P, Dr, Q, WR, from_above), choice tile1 + 4 * tile2 with
 *                   VMC_TILE_* for the two products (AUTO is the output-shape rule).
 * A variant replaces the default only when it is at least 1% faster. Outputs
 * are those of vmc_amp_grad with the chosen kernels. The workspace is the one
 * vmc_amp_grad_workspace_size gave before the call; a VMC_TRUNC_GENERIC choice
 * raises vmc_amp_grad_workspace_size, vmc_sweep_workspace_size and the batch
 * limit of vmc_amp_grad_noalloc afterwards, so query them again. Refused under
 * stream capture and between vmc_profile_begin and _end; synchronises `stream`.
 * On failure the plan keeps its previous choices. */
#define VMC_TUNE_TRUNC 1
#define VMC_TUNE_THETA 2
#define VMC_TRUNC_LPE8 0
#define VMC_TRUNC_LPE16 1
#define VMC_TRUNC_GENERIC 2
#define VMC_TILE_AUTO 0
#define VMC_TILE_WIDE 1   /* 16 x 256 output tile */
#define VMC_TILE_TALL 2   /* 256 x 16 */
#define VMC_TILE_SQUARE 3 /* 64 x 64 */
/* One tuning entry: kind, 9 key fields (unused ones 0), choice, average launch
 * ns of the choice and of the default when tuned (0 if set by the caller). */
#define VMC_TUNE_ENTRY_INTS 13
int vmc_plan_autotune(vmc_plan *plan, const void *tensors, const int32_t *samples, int batch, double *log_amp,
                      void *o, int8_t *p, void *workspace, size_t workspace_bytes, int reps, void *stream);
/* Writes min(count, capacity) entries and returns count (the number recorded). */
int vmc_plan_tuning(const vmc_plan *plan, int32_t *entries, int capacity);
/* Replaces the recorded choices with `count` entries (e.g. saved from
 * vmc_plan_tuning of an identical plan); 0 restores the defaults. An entry whose
 * choice cannot run its shape is refused and the table left unchanged. */
int vmc_plan_set_tuning(vmc_plan *plan, const int32_t *entries, int count);

/* Kernel timing with HIP events, for measurement only (no reference
 * counterpart: the reference has no profiler hooks). Between begin and end,
 * every launch is bracketed by events; end synchronises and returns, per
 * kernel class (0 theta contraction, 1 truncation, 2 row environments and
 * gradients, 3 final amplitude, 4 fills), the summed milliseconds and the
 * launch count. */
int vmc_profile_begin(vmc_plan *plan);
int vmc_profile_end(vmc_plan *plan, double *ms, int *launches, int n_classes);

const char *vmc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* VMC_AMD_H */
