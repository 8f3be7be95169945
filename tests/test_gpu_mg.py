"""GPU tests of the multigrid-preconditioned batched GMRES (phg_mg_*,
phg_gmres_batch_mg) on the pole-pinned Laplace stiffness (the Poisson solve
of SphereLaplace::solve) and on a Helmholtz operator.

Bars: true FP64 relative residual <= 1e-10 checked with the oracle's SpMV,
forward error against the oracle solve, a level-independent iteration count,
and a column's result independent of the rest of the batch."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import mke_amd as P
from oracle import mesh_ref, oracle as O
from tests._common import oracle_matrix, oracle_mesh, real_sph_harm_field

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _rhs(level, K, seed=11):
    text, om = oracle_mesh(level)
    pr = np.asarray(mesh_ref.load_text(text)["pr"]).reshape(-1, 3)[om.inner]
    rpm, cim, vm = oracle_matrix(level, O.KIND_MASS)
    return np.stack([O.spmv(rpm, cim, vm, real_sph_harm_field(pr, 10, seed + q))
                     for q in range(K)], axis=1)


@pytest.mark.parametrize("level", [3, 5, 7])
def test_poisson_mg_converges_level_independent(level):
    print("mg poisson level=%d" % level, flush=True)
    pm = P.Mesh.load(oracle_mesh(level)[0])
    mg = P.Multigrid(pm, P.LAPLACE)
    assert mg.levels == level - 1
    rp, ci, v = mg.A.csr()
    assert np.array_equal(v, oracle_matrix(level, O.KIND_LAPLACE)[2])
    B = _rhs(level, 2)
    dB = torch.as_tensor(B, device=DEV)
    dX = torch.zeros_like(dB)
    infos = mg.solve_batch(dX, dB, rtol=1e-10)
    X = dX.cpu().numpy()
    for q in range(2):
        assert infos[q]["converged"] and infos[q]["iterations"] <= 30, infos[q]
        r = B[:, q] - O.spmv(rp, ci, v, X[:, q])
        assert np.linalg.norm(r) / np.linalg.norm(B[:, q]) <= 1e-10 * 1.000001


def test_poisson_mg_forward_error_vs_oracle():
    level = 4
    pm = P.Mesh.load(oracle_mesh(level)[0])
    mg = P.Multigrid(pm, P.LAPLACE)
    rp, ci, v = oracle_matrix(level, O.KIND_LAPLACE)
    B = _rhs(level, 1)
    dX = torch.zeros((len(B), 1), dtype=torch.float64, device=DEV)
    mg.solve_batch(dX, torch.as_tensor(B, device=DEV), rtol=1e-12)
    xt, info = O.gmres(rp, ci, v, B[:, 0], rtol=1e-13, max_iters=400000)
    assert info["converged"]
    assert np.linalg.norm(dX.cpu().numpy()[:, 0] - xt) / np.linalg.norm(xt) < 1e-8


def test_mg_batch_columns_independent_and_helmholtz():
    level = 5
    pm = P.Mesh.load(oracle_mesh(level)[0])
    mg = P.Multigrid(pm, P.LAPLACE)
    B = _rhs(level, 4)
    X4 = torch.zeros((len(B), 4), dtype=torch.float64, device=DEV)
    mg.solve_batch(X4, torch.as_tensor(B, device=DEV), rtol=1e-10)
    B3 = _rhs(level, 3, seed=40)
    B3[:, 1] = B[:, 2]
    X3 = torch.zeros((len(B), 3), dtype=torch.float64, device=DEV)
    mg.solve_batch(X3, torch.as_tensor(B3, device=DEV), rtol=1e-10)
    assert torch.equal(X4[:, 2], X3[:, 1])
    # a Helmholtz operator (I - tau Laplace) with its own hierarchy
    cm, cl = P.helmholtz_coeffs(tau=1.0, sigma=0.0, theta=1.0, mu=1e-2)
    mh = P.Multigrid(pm, P.HELMHOLTZ, cm, cl)
    Xh = torch.zeros_like(X4)
    infos = mh.solve_batch(Xh, torch.as_tensor(B, device=DEV), rtol=1e-10)
    assert all(i["converged"] and i["iterations"] <= 20 for i in infos), infos
    rp, ci, v = oracle_matrix(level, O.KIND_HELMHOLTZ, cm, cl)
    Xh = Xh.cpu().numpy()
    for q in range(4):
        r = B[:, q] - O.spmv(rp, ci, v, Xh[:, q])
        assert np.linalg.norm(r) / np.linalg.norm(B[:, q]) <= 1e-10 * 1.000001


@pytest.mark.parametrize("prec", [P.F32])
def test_mg_mixed_precision_helmholtz_and_poisson(prec):
    """FP32 Krylov vectors and V-cycle under FP64 refinement reach the same
    true FP64 residual (checked with the oracle's SpMV) in few steps.  (BF16
    values also converge, but each refinement then gains only ~0.8x: the
    inner operator differs from A by ~4e-3, so it is not a useful form.)"""
    print("mg mixed prec=%d" % prec, flush=True)
    level = 5
    pm = P.Mesh.load(oracle_mesh(level)[0])
    B = _rhs(level, 4)
    cm, cl = P.helmholtz_coeffs(tau=1.0, sigma=0.0, theta=1.0, mu=1e-2)
    for kind, a, b, okind in ((P.HELMHOLTZ, cm, cl, O.KIND_HELMHOLTZ),
                              (P.LAPLACE, 1.0, 0.0, O.KIND_LAPLACE)):
        mg = P.Multigrid(pm, kind, a, b)
        X = torch.zeros((len(B), 4), dtype=torch.float64, device=DEV)
        infos = mg.solve_batch(X, torch.as_tensor(B, device=DEV), rtol=1e-10, precision=prec)
        assert all(i["converged"] and i["iterations"] <= 60 for i in infos), infos
        rp, ci, v = oracle_matrix(level, okind, a, b)
        Xh = X.cpu().numpy()
        for q in range(4):
            r = B[:, q] - O.spmv(rp, ci, v, Xh[:, q])
            assert np.linalg.norm(r) / np.linalg.norm(B[:, q]) <= 1e-10 * 1.000001, (kind, q)
