"""NetKet-facing adapters (src/vmc/qgt/netket_compat.py:21-78, vmc_utils.py:34-99)
on the device, against the FP64 oracle.

* build_dense_jac of a PEPS == the oracle's full-layout dpsi/dA / psi, centred;
  a generic torch apply function with a closed-form Jacobian.
* QGTOperator: to_dense == S + diag_shift (S from the oracle's centred
  Jacobian), ``@`` on a parameter pytree == to_dense @ flat, and _solve with cg.
* DenseSR over a variational-state stand-in (samples, parameters, model_state,
  _apply_fun as nk.vqs.MCState exposes them): x == solve(S + eps, g).
"""
import numpy as np
import pytest
import torch

from oracle import peps as P
from oracle import sr as S

pytestmark = pytest.mark.gpu


def _peps(shape=(3, 3), D=2, seed=3):
    from vmc_amd.peps import PEPS, ZipUp

    T = P.random_peps(shape, D, seed=seed)
    return PEPS(shape=shape, bond_dim=D, contraction_strategy=ZipUp(8), tensors=T, dtype=torch.float64), T


def _oracle_jac(T, cfg, shape, D):
    amp, grad, _ = P.value_and_grad(T, cfg, shape, P.ZipUp(8), D, full_gradient=True)
    O = grad / amp[:, None]
    return O - O.mean(axis=0, keepdims=True)


def test_build_dense_jac_peps():
    from vmc_amd.utils import build_dense_jac

    print("[netket] build_dense_jac 3x3", flush=True)
    shape, D = (3, 3), 2
    model, T = _peps(shape, D)
    cfg = P.random_configs(9, 40, seed=4)
    J = build_dense_jac(model, None, None, torch.tensor(cfg).reshape(4, 10, 9)).cpu().numpy()
    ref = _oracle_jac(T, cfg, shape, D)
    assert J.shape == ref.shape
    assert np.max(np.abs(J - ref)) < 1e-10 * np.max(np.abs(ref))
    # params given as the pytree of site tensors
    J2 = build_dense_jac(model.__call__, [[t for t in row] for row in model.tensors], None,
                         torch.tensor(cfg)).cpu().numpy()
    assert np.array_equal(J, J2)


def test_build_dense_jac_generic_apply_fun():
    from vmc_amd.utils import build_dense_jac

    rng = np.random.default_rng(1)
    x = torch.tensor(rng.integers(0, 2, size=(32, 6)).astype(np.float64), device="cuda")
    params = {"w": torch.tensor(rng.standard_normal(6), device="cuda"), "b": torch.tensor([0.3], dtype=torch.float64, device="cuda")}

    def apply_fun(variables, xs):  # log psi = w . x + b^2 x_0
        p = variables["params"]
        return xs @ p["w"] + p["b"][0] ** 2 * xs[:, 0]

    J = build_dense_jac(apply_fun, params, {}, x).cpu().numpy()
    xn = x.cpu().numpy()
    # ravel order: sorted keys (b, then w)
    ref = np.concatenate([(2 * 0.3 * xn[:, 0])[:, None], xn], axis=1)
    ref = ref - ref.mean(axis=0, keepdims=True)
    assert np.max(np.abs(J - ref)) < 1e-12


def test_qgt_operator_and_dense_sr():
    from vmc_amd.qgt import QGT, Jacobian, ParameterSpace
    from vmc_amd.qgt.netket_compat import DenseSR, QGTOperator, cg

    print("[netket] QGTOperator / DenseSR", flush=True)
    shape, D = (3, 3), 2
    model, T = _peps(shape, D)
    cfg = P.random_configs(9, 64, seed=5)
    Oc = _oracle_jac(T, cfg, shape, D)
    Sref = S.qgt_dense_parameter(Oc)
    eps = 1e-2
    op = QGTOperator(_qgt=QGT(Jacobian(torch.tensor(Oc, device="cuda")), ParameterSpace()), diag_shift=eps)
    dense = op.to_dense().cpu().numpy()
    assert np.max(np.abs(dense - (Sref + eps * np.eye(Sref.shape[0])))) < 1e-12 * np.abs(Sref).max()
    rng = np.random.default_rng(6)
    tree = [[torch.tensor(rng.standard_normal(t.shape), device="cuda") for t in row] for row in model.tensors]
    flat = np.concatenate([t.cpu().numpy().ravel() for row in tree for t in row])
    out = op @ tree
    out_flat = np.concatenate([t.cpu().numpy().ravel() for row in out for t in row])
    assert np.max(np.abs(out_flat - dense @ flat)) < 1e-12 * np.abs(dense @ flat).max()
    x, _ = op.solve(lambda A, b, x0=None: cg(A, b, x0=x0, tol=1e-12), tree)
    x_flat = np.concatenate([t.cpu().numpy().ravel() for row in x for t in row])
    ref = np.linalg.solve(Sref + eps * np.eye(Sref.shape[0]), flat)
    assert np.max(np.abs(x_flat - ref)) < 1e-8 * np.abs(ref).max()

    class VState:  # the attributes DenseSR reads from nk.vqs.MCState
        samples = torch.tensor(cfg).reshape(8, 8, 9)
        parameters = [[t for t in row] for row in model.tensors]
        model_state = {}
        _apply_fun = model

    sr = DenseSR(diag_shift=eps, solver=lambda A, b, x0=None: cg(A, b, x0=x0, tol=1e-12))
    g = [[torch.tensor(rng.standard_normal(t.shape), device="cuda") for t in row] for row in model.tensors]
    gf = np.concatenate([t.cpu().numpy().ravel() for row in g for t in row])
    xs = sr(VState(), g)
    xs_flat = np.concatenate([t.cpu().numpy().ravel() for row in xs for t in row])
    ref = np.linalg.solve(Sref + eps * np.eye(Sref.shape[0]), gf)
    assert np.max(np.abs(xs_flat - ref)) < 1e-8 * np.abs(ref).max()
