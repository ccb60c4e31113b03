"""The C-ABI library loads and exports every function declared in include/*.h.

No compute call is made (this runs without a GPU).
"""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(vmc_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert {"vmc_plan_create", "vmc_amp_grad", "vmc_last_error"} <= names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "vmc_amd", "libvmc_amd.so"))
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_every_symbol():
    from vmc_amd import _lib

    assert set(_lib.EXPORTED) == _declared()


def test_errors_are_reported_without_a_device():
    from vmc_amd import _lib

    h = ctypes.c_void_p()
    rc = _lib.LIB.vmc_plan_create(ctypes.byref(h), 0, 3, 2, 2, 4, _lib.VMC_ZIPUP, 0, _lib.VMC_F32)
    assert rc != 0
    assert b"positive" in _lib.LIB.vmc_last_error()
    rc = _lib.LIB.vmc_plan_create(ctypes.byref(h), 3, 3, 2, 2, 4, _lib.VMC_VARIATIONAL, -1, _lib.VMC_F64)
    assert rc != 0
    assert b"n_sweeps" in _lib.LIB.vmc_last_error()


def test_header_constants_match_the_binding():
    from vmc_amd import _lib

    defs = {}
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        for m in re.finditer(r"^#define\s+(VMC_\w+)\s+\(?([0-9]+)(?:u\s*<<\s*([0-9]+))?\)?", open(h).read(), flags=re.M):
            defs[m.group(1)] = int(m.group(2)) << int(m.group(3) or 0)
    names = [n for n in defs if n.startswith(("VMC_TUNE", "VMC_TRUNC", "VMC_TILE", "VMC_NOALLOC"))]
    assert len(names) >= 11
    for n in names:
        assert getattr(_lib, n) == defs[n], n


def test_tuning_calls_reject_null_plans():
    from vmc_amd import _lib

    assert _lib.LIB.vmc_plan_tuning(None, None, 0) != 0
    assert b"plan is null" in _lib.LIB.vmc_last_error()
    assert _lib.LIB.vmc_plan_set_tuning(None, None, 0) != 0
    assert _lib.LIB.vmc_plan_autotune(None, None, None, 1, None, None, None, None, 0, 3, None) != 0
    assert b"null argument" in _lib.LIB.vmc_last_error()
