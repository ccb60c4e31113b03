"""Build the gfx950 shared library ``vmc_amd/libvmc_amd.so`` with hipcc.

Run ``python -m vmc_amd.build`` (or ``__graft_entry__.build()``). Every source is
compiled to its own object in parallel under ``vmc_amd/_build/`` and the objects
are linked into the library in-tree, so it travels with the repository snapshot.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvmc_amd.so")
# (object name, source, extra defines). engine_dt.hip is compiled once per
# storage scalar so the four dtype instantiations of the engine build in
# parallel; engine.hip holds the C ABI and instantiates no kernels.
UNITS = [
    ("engine", "engine.hip", []),
    ("engine_f32", "engine_dt.hip", ["-DVMC_ENGINE_T=float", "-DVMC_ENGINE_OPS=ops_f32"]),
    ("engine_f64", "engine_dt.hip", ["-DVMC_ENGINE_T=double", "-DVMC_ENGINE_OPS=ops_f64"]),
    ("engine_c64", "engine_dt.hip", ["-DVMC_ENGINE_T=cf32", "-DVMC_ENGINE_OPS=ops_c64"]),
    ("engine_c128", "engine_dt.hip", ["-DVMC_ENGINE_T=cf64", "-DVMC_ENGINE_OPS=ops_c128"]),
    ("trunc", "trunc.hip", []),
    ("trunc_c", "trunc_c.hip", []),
    ("trunc_g", "trunc_g.hip", []),
    ("sr", "sr.hip", []),
]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    # Packed FP32 VALU ops stay disabled: hipcc -O3 would otherwise form
    # v_pk_* FP32 instructions from scalar float code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _stale(obj: str, src: str) -> bool:
    """An object is stale when its source or any shared header is newer."""
    if not os.path.exists(obj):
        return True
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    deps.append(os.path.join(HERE, "..", "include", "vmc_amd.h"))
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def _compile(unit, verbose: bool, incremental: bool = False) -> str:
    name, src, defines = unit
    obj = os.path.join(OBJ, name + ".o")
    if incremental and not _stale(obj, src):
        return obj
    cmd = [_hipcc(), *HIPCC_FLAGS, *defines, "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{res.stderr[-8000:]}")
    return obj


def build(verbose: bool = True, jobs: int | None = None, incremental: bool = False) -> str:
    """Compile and link the library. incremental=True (``--incremental``) skips
    objects newer than their source and every shared header; the default
    recompiles everything."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(len(UNITS), os.cpu_count() or 1, 16)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda u: _compile(u, verbose, incremental), UNITS))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "--offload-arch=gfx950", "-fPIC", "-shared", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
