"""Device plans over the C ABI: torch owns the device memory and the stream."""
from __future__ import annotations

import ctypes

import torch

from vmc_amd import _lib
from vmc_amd._lib import LIB, check

_DTYPES = {torch.float32: _lib.VMC_F32, torch.float64: _lib.VMC_F64,
           torch.complex64: _lib.VMC_C64, torch.complex128: _lib.VMC_C128}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class AmpGradPlan:
    """Plan + workspace for batched amplitude / sliced log-derivative evaluation.

    Mirrors the static configuration captured by ``build_mc_kernels``
    (src/vmc/peps/standard/kernels.py:104-113): lattice shape, bond and boundary
    dimensions and contraction strategy.
    """

    def __init__(self, shape, bond_dim, chi, phys_dim=2, strategy=_lib.VMC_ZIPUP,
                 dtype=torch.float32, device="cuda", n_sweeps=2):
        self.shape = (int(shape[0]), int(shape[1]))
        self.bond_dim, self.chi, self.phys_dim = int(bond_dim), int(chi), int(phys_dim)
        self.dtype = dtype
        self.device = torch.device(device)
        handle = ctypes.c_void_p()
        check(LIB.vmc_plan_create(ctypes.byref(handle), self.shape[0], self.shape[1], self.phys_dim,
                                  self.bond_dim, self.chi, int(strategy), int(n_sweeps), _DTYPES[dtype]))
        self._h = handle
        self.num_params = int(LIB.vmc_plan_num_params(handle))
        self.is_complex = dtype.is_complex
        self.energy_dtype = torch.complex128 if dtype.is_complex else torch.float64
        self.num_sliced = int(LIB.vmc_plan_num_sliced(handle))
        self._ws = None
        self._ws_batch = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            LIB.vmc_plan_destroy(h)
            self._h = None

    def workspace(self, batch: int) -> torch.Tensor:
        if self._ws is None or self._ws_batch < batch:
            nbytes = int(LIB.vmc_amp_grad_workspace_size(self._h, int(batch)))
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            self._ws_batch = batch
        return self._ws

    KERNEL_CLASSES = ("theta", "trunc", "row_grad", "amp", "fill")

    # ---------------- autotuning (vmc_plan_autotune) ----------------

    _TUNE_KINDS = {_lib.VMC_TUNE_TRUNC: "trunc", _lib.VMC_TUNE_THETA: "theta"}

    def autotune(self, params, samples, *, reps=3):
        """Times every kernel variant of each truncation and theta shape of an
        amp+grad call on this batch and records the fastest in the plan; later
        calls (amp_grad, sweeps, graphs captured afterwards) launch the recorded
        kernels. Returns the recorded choices (see tuning)."""
        samples = samples.to(device=self.device, dtype=torch.int32).contiguous()
        B = samples.shape[0]
        log_amp = torch.empty((B, 2), dtype=torch.float64, device=self.device)
        o = torch.empty((B, self.num_sliced), dtype=self.dtype, device=self.device)
        ws = self.workspace(B)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_plan_autotune(self._h, _ptr(params), _ptr(samples), int(B), _ptr(log_amp), _ptr(o), None,
                                    _ptr(ws), ctypes.c_size_t(ws.numel()), int(reps), ctypes.c_void_p(stream)))
        self._reset_workspaces()
        return self.tuning()

    def tuning(self):
        """Recorded choices: dicts with kind ("trunc": key (m, n, k, dm), choice
        VMC_TRUNC_*; "theta": key (has_carry, k, Dl, wl, P, Dr, Q, WR, from_above),
        choice tile1 + 4 * tile2 with VMC_TILE_*), the key, the choice and the
        average launch ns of the choice and of the default."""
        n = int(LIB.vmc_plan_tuning(self._h, None, 0))
        if n < 0:
            check(n)
        w = _lib.VMC_TUNE_ENTRY_INTS
        buf = (ctypes.c_int32 * max(1, n * w))()
        LIB.vmc_plan_tuning(self._h, buf, n)
        out = []
        for i in range(n):
            row = list(buf[i * w:(i + 1) * w])
            nkey = 4 if row[0] == _lib.VMC_TUNE_TRUNC else 9
            out.append({"kind": self._TUNE_KINDS.get(row[0], row[0]), "key": tuple(row[1:1 + nkey]),
                        "choice": row[10], "ns": row[11], "default_ns": row[12]})
        return out

    def set_tuning(self, entries):
        """Replaces the recorded choices with entries of the form tuning()
        returns (for instance saved from an identical plan); [] restores the
        default kernels."""
        kinds = {v: k for k, v in self._TUNE_KINDS.items()}
        rows = []
        for e in entries:
            key = list(e["key"]) + [0] * (9 - len(e["key"]))
            rows += [kinds[e["kind"]], *key, int(e["choice"]), int(e.get("ns", 0)), int(e.get("default_ns", 0))]
        buf = (ctypes.c_int32 * max(1, len(rows)))(*rows)
        check(LIB.vmc_plan_set_tuning(self._h, buf, len(entries)))
        self._reset_workspaces()

    def _reset_workspaces(self):
        # a generic-kernel choice changes the workspace sizes
        self._ws = None
        self._ws_batch = 0
        self._sweep_ws = None

    def profile_begin(self):
        check(LIB.vmc_profile_begin(self._h))

    def profile_end(self):
        n = len(self.KERNEL_CLASSES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        check(LIB.vmc_profile_end(self._h, ms, cnt, n))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.KERNEL_CLASSES)}

    # ---------------- sampler (transition / estimate) ----------------

    def set_terms(self, tables):
        """Upload operator tables (see vmc_amd.operators.local_terms.to_tables)."""
        import numpy as np

        def arr(x, dt):
            return np.ascontiguousarray(np.asarray(x, dtype=dt)) if x is not None and len(x) else None

        t = {k: v for k, v in tables.items()}
        keep = []

        def ptr(x):
            if x is None:
                return None
            keep.append(x)
            return ctypes.c_void_p(x.ctypes.data)

        # complex values go to vmc_plan_set_terms_c as interleaved (re, im) doubles
        cplx = any(np.iscomplexobj(np.asarray(t[k])) for k in ("diag_vals", "one_ops", "h_ops", "v_ops")
                   if t.get(k) is not None)

        def vals(x):
            if x is None or not len(x):
                return None
            if cplx:
                return np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).view(np.float64)
            return arr(x, np.float64)

        ds, dv = arr(t["diag_sites"], np.int32), vals(t["diag_vals"])
        os_, oo = arr(t["one_sites"], np.int32), vals(t["one_ops"])
        hs, ho = arr(t["h_sites"], np.int32), vals(t["h_ops"])
        vs, vo = arr(t["v_sites"], np.int32), vals(t["v_ops"])
        fn = LIB.vmc_plan_set_terms_c if cplx else LIB.vmc_plan_set_terms
        check(fn(self._h, len(t["diag_sites"]) // 2 if t["diag_sites"] is not None else 0,
                 ptr(ds), ptr(dv), 0 if os_ is None else len(os_), ptr(os_), ptr(oo),
                 0 if hs is None else len(hs), ptr(hs), ptr(ho),
                 0 if vs is None else len(vs), ptr(vs), ptr(vo)))
        self._has_terms = True

    def new_cache(self, n_chains: int) -> torch.Tensor:
        nbytes = int(LIB.vmc_cache_bytes(self._h, int(n_chains)))
        return torch.empty(nbytes, dtype=torch.uint8, device=self.device)

    def sweep_workspace(self, n_chains: int) -> torch.Tensor:
        ws = getattr(self, "_sweep_ws", None)
        nbytes = int(LIB.vmc_sweep_workspace_size(self._h, int(n_chains)))
        if ws is None or ws.numel() < nbytes:
            ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            self._sweep_ws = ws
        return ws

    def init_cache(self, params, configs, cache):
        configs = configs.to(device=self.device, dtype=torch.int32).contiguous()
        n = configs.shape[0]
        ws = self.sweep_workspace(n)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_init_cache(self._h, _ptr(params), _ptr(configs), int(n), _ptr(cache), _ptr(ws),
                                 ctypes.c_size_t(ws.numel()), ctypes.c_void_p(stream)))
        return cache

    def mc_sweep(self, params, configs, cache, seed, sweep, *, grad=True, want_p=True, energy=True):
        """One transition + estimate of every chain; `configs` (int32, device) is updated in place."""
        assert configs.dtype == torch.int32 and configs.is_contiguous() and configs.device.type == "cuda"
        n = configs.shape[0]
        log_amp = torch.empty((n, 2), dtype=torch.float64, device=self.device)
        o = torch.empty((n, self.num_sliced), dtype=self.dtype, device=self.device) if grad else None
        p = torch.empty((n, self.num_sliced), dtype=torch.int8, device=self.device) if want_p else None
        eloc = torch.empty((n,), dtype=self.energy_dtype, device=self.device) if energy else None
        ws = self.sweep_workspace(n)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_mc_sweep(self._h, _ptr(params), _ptr(configs), int(n), ctypes.c_uint64(int(seed)),
                               int(sweep), _ptr(cache), _ptr(log_amp), _ptr(o), _ptr(p), _ptr(eloc), _ptr(ws),
                               ctypes.c_size_t(ws.numel()), ctypes.c_void_p(stream)))
        return log_amp, o, p, eloc

    def transition(self, params, configs, cache, seed, sweep, ws=None):
        """Metropolis sweep only; the top boundaries stay in the workspace `ws`
        (a fresh one, or this plan's own when None) for the next estimate."""
        n = configs.shape[0]
        log_amp = torch.empty((n, 2), dtype=torch.float64, device=self.device)
        ws = self.sweep_workspace(n) if ws is None else ws
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_transition(self._h, _ptr(params), _ptr(configs), int(n), ctypes.c_uint64(int(seed)),
                                 int(sweep), _ptr(cache), _ptr(log_amp), _ptr(ws), ctypes.c_size_t(ws.numel()),
                                 ctypes.c_void_p(stream)))
        return log_amp

    def new_sweep_workspace(self, n_chains: int) -> torch.Tensor:
        """A workspace owned by the caller (vmc_sweep_workspace_size bytes)."""
        nbytes = int(LIB.vmc_sweep_workspace_size(self._h, int(n_chains)))
        return torch.empty(nbytes, dtype=torch.uint8, device=self.device)

    def estimate(self, params, configs, cache, log_amp, *, grad=True, want_p=True, energy=True, ws=None):
        n = configs.shape[0]
        o = torch.empty((n, self.num_sliced), dtype=self.dtype, device=self.device) if grad else None
        p = torch.empty((n, self.num_sliced), dtype=torch.int8, device=self.device) if want_p else None
        eloc = torch.empty((n,), dtype=self.energy_dtype, device=self.device) if energy else None
        ws = self.sweep_workspace(n) if ws is None else ws
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_estimate(self._h, _ptr(params), _ptr(configs), int(n), _ptr(cache), _ptr(log_amp), _ptr(o),
                               _ptr(p), _ptr(eloc), _ptr(ws), ctypes.c_size_t(ws.numel()),
                               ctypes.c_void_p(stream)))
        return o, p, eloc

    def local_estimate(self, params, configs):
        configs = configs.to(device=self.device, dtype=torch.int32).contiguous()
        n = configs.shape[0]
        cache = self.new_cache(n)
        log_amp = torch.empty((n, 2), dtype=torch.float64, device=self.device)
        eloc = torch.empty((n,), dtype=self.energy_dtype, device=self.device)
        ws = self.sweep_workspace(n)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_local_estimate(self._h, _ptr(params), _ptr(configs), int(n), _ptr(cache), _ptr(log_amp),
                                     _ptr(eloc), _ptr(ws), ctypes.c_size_t(ws.numel()), ctypes.c_void_p(stream)))
        return log_amp, eloc

    def amp_grad(self, params: torch.Tensor, configs: torch.Tensor, *, grad=True, want_p=False,
                 out=None):
        """Returns (log_amp (B,2) float64, o (B, num_sliced) or None, p or None)."""
        assert params.dtype == self.dtype and params.device.type == "cuda" and params.is_contiguous()
        assert params.numel() == self.num_params, (params.numel(), self.num_params)
        configs = configs.to(device=self.device, dtype=torch.int32).contiguous()
        B = configs.shape[0]
        assert configs.shape[1] == self.shape[0] * self.shape[1]
        if out is None:
            log_amp = torch.empty((B, 2), dtype=torch.float64, device=self.device)
            o = torch.empty((B, self.num_sliced), dtype=self.dtype, device=self.device) if grad else None
            p = torch.empty((B, self.num_sliced), dtype=torch.int8, device=self.device) if want_p else None
        else:
            log_amp, o, p = out
        ws = self.workspace(B)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_amp_grad(self._h, _ptr(params), _ptr(configs), int(B), _ptr(log_amp), _ptr(o),
                               _ptr(p), _ptr(ws), ctypes.c_size_t(ws.numel()), ctypes.c_void_p(stream)))
        return log_amp, o, p


class GIPlan(AmpGradPlan):
    """Plan of a gauge-invariant PEPS (src/vmc/peps/gi/model.py:40-118): Z_N
    links, site tensors (d, nc, dmax...) with nc = N^(links - 1), samples
    laid out as sites, horizontal links, vertical links (int32)."""

    def __init__(self, shape, N, phys_dim, dmax, chi, charge_of_site, Qx=0, strategy=_lib.VMC_ZIPUP,
                 dtype=torch.float64, device="cuda", n_sweeps=2, degeneracy_per_charge=None):
        import numpy as np

        self.shape = (int(shape[0]), int(shape[1]))
        self.N, self.Qx = int(N), int(Qx)
        self.bond_dim, self.chi, self.phys_dim = int(dmax), int(chi), int(phys_dim)
        self.dtype = dtype
        self.device = torch.device(device)
        ch = np.ascontiguousarray(np.asarray(charge_of_site, dtype=np.int32))
        handle = ctypes.c_void_p()
        check(LIB.vmc_gi_plan_create(ctypes.byref(handle), self.shape[0], self.shape[1], self.N, self.phys_dim,
                                     self.bond_dim, self.chi, int(strategy), int(n_sweeps), _DTYPES[dtype], self.Qx,
                                     ctypes.c_void_p(ch.ctypes.data)))
        self._h = handle
        self.num_params = int(LIB.vmc_plan_num_params(handle))
        self.num_sliced = int(LIB.vmc_plan_num_sliced(handle))
        self.is_complex = dtype.is_complex
        self.energy_dtype = torch.complex128 if dtype.is_complex else torch.float64
        R, C = self.shape
        self.n_sample = R * C + R * (C - 1) + (R - 1) * C
        self._ws = None
        self._ws_batch = 0
        if degeneracy_per_charge is not None:
            self.set_degeneracy(degeneracy_per_charge)

    def set_degeneracy(self, degeneracy_per_charge):
        """Charge masks of GIPEPSConfig.degeneracy_per_charge (gi/model.py:52-60):
        bond index i of a link with charge q is kept iff i < degeneracy_per_charge[q]."""
        import numpy as np

        deg = np.ascontiguousarray(np.asarray(degeneracy_per_charge, dtype=np.int32))
        assert deg.shape == (self.N,), deg.shape
        check(LIB.vmc_gi_plan_set_degeneracy(self._h, ctypes.c_void_p(deg.ctypes.data)))
        self._ws = None  # the masked copy of the parameters takes workspace room
        self._ws_batch = 0

    def workspace(self, batch: int) -> torch.Tensor:
        if self._ws is None or self._ws_batch < batch:
            nbytes = int(LIB.vmc_gi_workspace_size(self._h, int(batch)))
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            self._ws_batch = batch
        return self._ws

    sweep_workspace = workspace

    def set_gi_terms(self, plaquette_coeff=None, link_diag=None):
        """plaquette_coeff: ((R-1)*(C-1),) with NaN where no term; link_diag:
        (n_links, N) table of link-diagonal values (horizontal links first)."""
        import numpy as np

        keep = []

        def ptr(x):
            if x is None:
                return None
            a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
            keep.append(a)
            return ctypes.c_void_p(a.ctypes.data)

        check(LIB.vmc_gi_set_terms(self._h, ptr(plaquette_coeff), ptr(link_diag)))
        self._has_terms = plaquette_coeff is not None or link_diag is not None

    def slices(self, samples):
        samples = samples.to(device=self.device, dtype=torch.int32).contiguous()
        B = samples.shape[0]
        sl = torch.empty((B, self.shape[0] * self.shape[1]), dtype=torch.int32, device=self.device)
        valid = torch.empty((B,), dtype=torch.uint8, device=self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_gi_slices(self._h, _ptr(samples), int(B), _ptr(sl), _ptr(valid), ctypes.c_void_p(stream)))
        return sl, valid.bool()

    def amp_grad(self, params, samples, *, grad=True, want_p=False, out=None):
        assert params.dtype == self.dtype and params.device.type == "cuda" and params.is_contiguous()
        assert params.numel() == self.num_params, (params.numel(), self.num_params)
        samples = samples.to(device=self.device, dtype=torch.int32).contiguous()
        B = samples.shape[0]
        assert samples.shape[1] == self.n_sample
        if out is None:
            log_amp = torch.empty((B, 2), dtype=torch.float64, device=self.device)
            o = torch.empty((B, self.num_sliced), dtype=self.dtype, device=self.device) if grad else None
            p = torch.empty((B, self.num_sliced), dtype=torch.int8, device=self.device) if want_p else None
        else:
            log_amp, o, p = out
        ws = self.workspace(B)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_gi_amp_grad(self._h, _ptr(params), _ptr(samples), int(B), _ptr(log_amp), _ptr(o), _ptr(p),
                                  _ptr(ws), ctypes.c_size_t(ws.numel()), ctypes.c_void_p(stream)))
        return log_amp, o, p

    def init_cache(self, params, samples, cache):
        samples = samples.to(device=self.device, dtype=torch.int32).contiguous()
        n = samples.shape[0]
        ws = self.workspace(n)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_gi_init_cache(self._h, _ptr(params), _ptr(samples), int(n), _ptr(cache), _ptr(ws),
                                    ctypes.c_size_t(ws.numel()), ctypes.c_void_p(stream)))
        return cache

    def mc_sweep(self, params, samples, cache, seed, sweep, *, grad=True, want_p=True, energy=True):
        """One transition (plaquette sweep; with matter also the horizontal and vertical link
        sweeps) + estimate of every chain; `samples` (int32, device) is updated in place."""
        assert samples.dtype == torch.int32 and samples.is_contiguous() and samples.device.type == "cuda"
        assert samples.dim() == 2 and samples.shape[1] == self.n_sample
        n = samples.shape[0]
        log_amp = torch.empty((n, 2), dtype=torch.float64, device=self.device)
        o = torch.empty((n, self.num_sliced), dtype=self.dtype, device=self.device) if grad else None
        p = torch.empty((n, self.num_sliced), dtype=torch.int8, device=self.device) if want_p else None
        eloc = torch.empty((n,), dtype=self.energy_dtype, device=self.device) if energy else None
        ws = self.workspace(n)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_gi_mc_sweep(self._h, _ptr(params), _ptr(samples), int(n), ctypes.c_uint64(int(seed)),
                                  int(sweep), _ptr(cache), _ptr(log_amp), _ptr(o), _ptr(p), _ptr(eloc), _ptr(ws),
                                  ctypes.c_size_t(ws.numel()), ctypes.c_void_p(stream)))
        return log_amp, o, p, eloc


class BLPlan(AmpGradPlan):
    """Plan of a blockade PEPS (src/vmc/peps/blockade/model.py:43-121): site
    tensors (2, nc, Dmax...) with nc = 2^(incoming bonds), sector degeneracies
    D0 / D1, samples are the int32 occupations of the sites (row-major)."""

    def __init__(self, shape, D0, D1, chi, strategy=_lib.VMC_ZIPUP, dtype=torch.float64, device="cuda", n_sweeps=2):
        self.shape = (int(shape[0]), int(shape[1]))
        self.D0, self.D1 = int(D0), int(D1)
        self.bond_dim, self.chi, self.phys_dim = max(self.D0, self.D1), int(chi), 2
        self.dtype = dtype
        self.device = torch.device(device)
        handle = ctypes.c_void_p()
        check(LIB.vmc_bl_plan_create(ctypes.byref(handle), self.shape[0], self.shape[1], self.D0, self.D1, self.chi,
                                     int(strategy), int(n_sweeps), _DTYPES[dtype]))
        self._h = handle
        self.num_params = int(LIB.vmc_plan_num_params(handle))
        self.num_sliced = int(LIB.vmc_plan_num_sliced(handle))
        self.is_complex = dtype.is_complex
        self.energy_dtype = torch.complex128 if dtype.is_complex else torch.float64
        self.n_sample = self.shape[0] * self.shape[1]
        self._ws = None
        self._ws_batch = 0

    def workspace(self, batch: int) -> torch.Tensor:
        if self._ws is None or self._ws_batch < batch:
            nbytes = int(LIB.vmc_bl_workspace_size(self._h, int(batch)))
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            self._ws_batch = batch
        return self._ws

    sweep_workspace = workspace

    def slices(self, occupations):
        occupations = occupations.to(device=self.device, dtype=torch.int32).contiguous()
        B = occupations.shape[0]
        sl = torch.empty((B, self.n_sample), dtype=torch.int32, device=self.device)
        valid = torch.empty((B,), dtype=torch.uint8, device=self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_bl_slices(self._h, _ptr(occupations), int(B), _ptr(sl), _ptr(valid), ctypes.c_void_p(stream)))
        return sl, valid.bool()

    def amp_grad(self, params, occupations, *, grad=True, want_p=False, out=None):
        assert params.dtype == self.dtype and params.device.type == "cuda" and params.is_contiguous()
        assert params.numel() == self.num_params, (params.numel(), self.num_params)
        occupations = occupations.to(device=self.device, dtype=torch.int32).contiguous()
        B = occupations.shape[0]
        assert occupations.shape[1] == self.n_sample
        if out is None:
            log_amp = torch.empty((B, 2), dtype=torch.float64, device=self.device)
            o = torch.empty((B, self.num_sliced), dtype=self.dtype, device=self.device) if grad else None
            p = torch.empty((B, self.num_sliced), dtype=torch.int8, device=self.device) if want_p else None
        else:
            log_amp, o, p = out
        ws = self.workspace(B)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_bl_amp_grad(self._h, _ptr(params), _ptr(occupations), int(B), _ptr(log_amp), _ptr(o), _ptr(p),
                                  _ptr(ws), ctypes.c_size_t(ws.numel()), ctypes.c_void_p(stream)))
        return log_amp, o, p

    def init_cache(self, params, occupations, cache):
        occupations = occupations.to(device=self.device, dtype=torch.int32).contiguous()
        n = occupations.shape[0]
        ws = self.workspace(n)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_bl_init_cache(self._h, _ptr(params), _ptr(occupations), int(n), _ptr(cache), _ptr(ws),
                                    ctypes.c_size_t(ws.numel()), ctypes.c_void_p(stream)))
        return cache

    def mc_sweep(self, params, occupations, cache, seed, sweep, *, grad=True, want_p=True, energy=True):
        """One transition (row-pair sweeps, then the last row) + estimate of every
        chain; `occupations` (int32, device) is updated in place."""
        assert occupations.dtype == torch.int32 and occupations.is_contiguous()
        assert occupations.device.type == "cuda" and occupations.dim() == 2 and occupations.shape[1] == self.n_sample
        n = occupations.shape[0]
        log_amp = torch.empty((n, 2), dtype=torch.float64, device=self.device)
        o = torch.empty((n, self.num_sliced), dtype=self.dtype, device=self.device) if grad else None
        p = torch.empty((n, self.num_sliced), dtype=torch.int8, device=self.device) if want_p else None
        eloc = torch.empty((n,), dtype=self.energy_dtype, device=self.device) if energy else None
        ws = self.workspace(n)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(LIB.vmc_bl_mc_sweep(self._h, _ptr(params), _ptr(occupations), int(n), ctypes.c_uint64(int(seed)),
                                  int(sweep), _ptr(cache), _ptr(log_amp), _ptr(o), _ptr(p), _ptr(eloc), _ptr(ws),
                                  ctypes.c_size_t(ws.numel()), ctypes.c_void_p(stream)))
        return log_amp, o, p, eloc


class AmpGradGraph:
    """One amp+grad step of a plan captured in a HIP graph (torch.cuda.CUDAGraph
    records the library's launches on the capture stream: vmc_amp_grad allocates,
    copies and synchronises nothing, and its initialisations are fill kernels).

    The graph reads ``params`` and ``configs`` and writes ``log_amp``, ``o`` and
    ``p`` in place: copy new inputs into those buffers, then ``replay()``. The
    captured step is the same launch sequence as ``plan.amp_grad``, so replays
    return the eager results bit for bit.
    """

    def __init__(self, plan: AmpGradPlan, params: torch.Tensor, configs: torch.Tensor, *, grad=True, want_p=False):
        self.plan = plan
        B = configs.shape[0]
        self.params = params.detach().clone().contiguous()
        self.configs = configs.to(device=plan.device, dtype=torch.int32).contiguous().clone()
        self.log_amp = torch.empty((B, 2), dtype=torch.float64, device=plan.device)
        self.o = torch.empty((B, plan.num_sliced), dtype=plan.dtype, device=plan.device) if grad else None
        self.p = torch.empty((B, plan.num_sliced), dtype=torch.int8, device=plan.device) if want_p else None
        out = (self.log_amp, self.o, self.p)
        plan.workspace(B)  # allocate before capture
        side = torch.cuda.Stream(device=plan.device)
        side.wait_stream(torch.cuda.current_stream(plan.device))
        with torch.cuda.stream(side):  # warm-up on the capture stream
            plan.amp_grad(self.params, self.configs, out=out)
        torch.cuda.current_stream(plan.device).wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            plan.amp_grad(self.params, self.configs, out=out)

    def replay(self):
        self.graph.replay()
        return self.log_amp, self.o, self.p


class MixedAmpGradGraph:
    """Plans of different lattice and bond sizes evaluated by one HIP-graph launch.

    Each member (plan, params, samples) gets its own stream; the capture forks
    every stream from the capture stream, records the member's amp+grad step on
    it and joins them back, so one ``replay()`` launches all members as
    independent branches that run concurrently on the GPU. Works for standard
    and gauge-invariant plans (anything with ``amp_grad(params, samples, out=)``).
    Replays reproduce each member's eager result bit for bit.
    """

    def __init__(self, members, *, grad=True):
        self.members = []
        cur = torch.cuda.current_stream()
        self.streams = [torch.cuda.Stream(device=m[0].device) for m in members]
        for (plan, params, samples), st in zip(members, self.streams):
            B = samples.shape[0]
            rec = {
                "plan": plan,
                "params": params.detach().clone().contiguous(),
                "samples": samples.to(device=plan.device, dtype=torch.int32).contiguous().clone(),
                "log_amp": torch.empty((B, 2), dtype=torch.float64, device=plan.device),
                "o": torch.empty((B, plan.num_sliced), dtype=plan.dtype, device=plan.device) if grad else None,
            }
            plan.workspace(B)
            st.wait_stream(cur)
            with torch.cuda.stream(st):  # warm-up on the member's stream
                plan.amp_grad(rec["params"], rec["samples"], out=(rec["log_amp"], rec["o"], None))
            cur.wait_stream(st)
            self.members.append(rec)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            cap = torch.cuda.current_stream()
            for rec, st in zip(self.members, self.streams):
                st.wait_stream(cap)
                with torch.cuda.stream(st):
                    rec["plan"].amp_grad(rec["params"], rec["samples"], out=(rec["log_amp"], rec["o"], None))
            for st in self.streams:
                cap.wait_stream(st)

    def replay(self):
        self.graph.replay()
        return [(rec["log_amp"], rec["o"]) for rec in self.members]
