"""FP64 CPU oracle for the PEPS amplitude engine.

TEST INFRASTRUCTURE ONLY. Nothing in the product path (``vmc_amd``) may import,
call or link anything under ``oracle/``. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker (or the timed CPU baseline), never as the thing measured
or shipped.

The oracle is a NumPy restatement of the reference's hot path, batched over
samples (leading axis ``b``), using the reference's own pairwise einsum orders
and LAPACK factorizations so rounding stays comparable:

* ``oracle.peps``   -- boundary-MPS strategies, forward/backward sweeps,
  env-gradients, local energies and the sequential-sampling sweep
  (src/vmc/peps/common/*.py, src/vmc/peps/standard/*.py, src/vmc/core/sampling.py);
* ``oracle.sr``     -- sliced Jacobian, QGT, SR solves, Euler/RK4
  (src/vmc/qgt/*.py, src/vmc/preconditioners/preconditioners.py,
  src/vmc/drivers/integrators.py);
* ``oracle.exact``  -- independent brute-force checks (full-network einsum,
  dense Hamiltonians) that pin the restatement;
* ``oracle.rng``    -- Philox4x32-10, the counter-based generator shared with the
  GPU sampler so sample sequences can be compared draw for draw;
* ``oracle.flops``  -- shape-replay counter of algorithmic work per sample.

Pinning status of every function is listed in the module headers.
"""
