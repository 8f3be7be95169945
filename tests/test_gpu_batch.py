"""GPU tests of the batched independent-column GMRES (phg_gmres_batch) and
of the GPU SphereJacobian (phg_jacobian_*), through the C ABI.

Bars: each column's true FP64 residual <= rtol (checked on the host with the
oracle's SpMV) and its forward error against a tight oracle solve; a
column's result is bit-identical whatever else is in the batch (same padded
width); GPU calc2 against the oracle restatement (direct mass solves) within
1e-11 of max |J| (the GPU mass solves stop at 1e-13)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import mke_amd as P
from oracle import mesh_ref, oracle as O
from tests._common import oracle_matrix, oracle_mesh, real_sph_harm_field

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
CM, CL = P.helmholtz_coeffs(tau=1.0, sigma=0.0, theta=1.0, mu=1e-2)


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


def _rhs(level, K, seed0=100):
    text, om = oracle_mesh(level)
    pr = np.asarray(mesh_ref.load_text(text)["pr"]).reshape(-1, 3)[om.inner]
    rpm, cim, vm = oracle_matrix(level, O.KIND_MASS)
    cols = [O.spmv(rpm, cim, vm, real_sph_harm_field(pr, 8, seed0 + q)) for q in range(K)]
    return np.stack(cols, axis=1)


def _check_columns(level, B, X, rtol, fwd=None):
    rp, ci, v = oracle_matrix(level, O.KIND_HELMHOLTZ, CM, CL)
    for q in range(B.shape[1]):
        r = B[:, q] - O.spmv(rp, ci, v, X[:, q])
        rel = np.linalg.norm(r) / np.linalg.norm(B[:, q])
        assert rel <= rtol * 1.000001, (q, rel)
        if fwd is not None:
            xt, _ = O.gmres(rp, ci, v, B[:, q], rtol=1e-14)
            assert np.linalg.norm(X[:, q] - xt) / np.linalg.norm(xt) <= fwd, q


@pytest.mark.parametrize("K", [1, 3, 8])
def test_batch_fp64(K):
    print("batch fp64 K=%d" % K, flush=True)
    level = 4
    pm = P.Mesh.load(oracle_mesh(level)[0])
    A = P.Matrix.assemble(pm, P.HELMHOLTZ, CM, CL)
    B = _rhs(level, K)
    dB = _dev(B)
    dX = torch.zeros_like(dB)
    infos = A.solve_batch(dX, dB, rtol=1e-10)
    assert all(i["converged"] for i in infos)
    _check_columns(level, B, dX.cpu().numpy(), 1e-10, fwd=1e-8)


@pytest.mark.parametrize("prec", [P.F32, P.BF16])
def test_batch_mixed(prec):
    print("batch mixed prec=%d" % prec, flush=True)
    level = 4
    pm = P.Mesh.load(oracle_mesh(level)[0])
    A = P.Matrix.assemble(pm, P.HELMHOLTZ, CM, CL)
    B = _rhs(level, 8)
    dB = _dev(B)
    dX = torch.zeros_like(dB)
    infos = A.solve_batch(dX, dB, rtol=1e-10, precision=prec)
    assert all(i["converged"] for i in infos)
    assert all(i["refinements"] >= 1 for i in infos)
    _check_columns(level, B, dX.cpu().numpy(), 1e-10, fwd=1e-8)


def test_batch_column_is_independent_of_the_mix():
    print("batch independence", flush=True)
    level = 4
    pm = P.Mesh.load(oracle_mesh(level)[0])
    A = P.Matrix.assemble(pm, P.HELMHOLTZ, CM, CL)
    B8 = _rhs(level, 8)
    X8 = torch.zeros(B8.shape, dtype=torch.float64, device=DEV)
    A.solve_batch(X8, _dev(B8), rtol=1e-10)
    # the same right-hand side alone at another position of a 5-wide batch
    # (padded to 8 like the first) with other companions
    B5 = _rhs(level, 5, seed0=500)
    B5[:, 4] = B8[:, 2]
    X5 = torch.zeros(B5.shape, dtype=torch.float64, device=DEV)
    A.solve_batch(X5, _dev(B5), rtol=1e-10)
    assert torch.equal(X8[:, 2], X5[:, 4])
    # eager launches give the same bits as the captured graph
    Xe = torch.zeros_like(X8)
    A.solve_batch(Xe, _dev(B8), rtol=1e-10, graph=False)
    assert torch.equal(Xe, X8)


def test_batch_jacobian_mass_cgs1():
    print("batch mass cgs1", flush=True)
    level = 5
    pm = P.Mesh.load(oracle_mesh(level)[0])
    M = P.Matrix.assemble(pm, P.JAC_MASS)
    rp, ci, v = oracle_matrix(level, O.KIND_JAC_MASS)
    B = np.random.default_rng(3).standard_normal((M.n_rows, 8))
    dX = torch.zeros(B.shape, dtype=torch.float64, device=DEV)
    infos = M.solve_batch(dX, _dev(B), rtol=1e-13, ortho=P.CGS1, restart=32)
    X = dX.cpu().numpy()
    for q in range(8):
        rel = np.linalg.norm(B[:, q] - O.spmv(rp, ci, v, X[:, q])) / np.linalg.norm(B[:, q])
        assert rel <= 1e-13 * 1.000001 and infos[q]["iterations"] < 64, (q, rel, infos[q])


def _full_fields(level, npairs, seed=7):
    text, om = oracle_mesh(level)
    pr = np.asarray(mesh_ref.load_text(text)["pr"]).reshape(-1, 3)
    us = np.stack([real_sph_harm_field(pr, 6, seed + 2 * k) for k in range(npairs)])
    vs = np.stack([real_sph_harm_field(pr, 6, seed + 2 * k + 1) for k in range(npairs)])
    return om, us, vs


@pytest.mark.parametrize("level", [3, 5])
def test_jacobian_calc2_vs_oracle(level):
    print("jacobian calc2 level=%d" % level, flush=True)
    from oracle.sjacobian import SphereJacobian as OJ
    om, us, vs = _full_fields(level, 3)
    pm = P.Mesh.load(oracle_mesh(level)[0])
    J = P.SphereJacobian(pm)
    for name, kind in (("idt_", O.KIND_JAC_MASS), ("diff2_cos_", O.KIND_DIFF2_COS)):
        rp, ci, v = getattr(J, name).csr()
        assert np.array_equal(v, oracle_matrix(level, kind)[2]), name
    ans = torch.zeros((3, om.n_inner), dtype=torch.float64, device=DEV)
    infos = J.calc2(ans, _dev(us), _dev(vs))
    assert len(infos) == 24 and all(i["converged"] for i in infos)
    oj = OJ(om)
    got = ans.cpu().numpy()
    for k in range(3):
        ref = oj.calc2(us[k], vs[k])
        assert np.abs(got[k] - ref).max() <= 1e-11 * np.abs(ref).max(), k
    # one pair alone (padded width 8 instead of 32) agrees to the solve tolerance
    one = torch.zeros(om.n_inner, dtype=torch.float64, device=DEV)
    J.calc2(one, _dev(us[1]), _dev(vs[1]))
    assert torch.allclose(one, ans[1], rtol=0, atol=1e-11 * float(ans[1].abs().max()))
    # calc2t = -calc2 (same launch shape: same bits)
    onet = torch.zeros_like(one)
    J.calc2t(onet, _dev(us[1]), _dev(vs[1]))
    assert torch.equal(onet, -one)


def test_batch_autotuned_launch_shapes():
    """The autotuner picks one of the rows-per-thread values; every value
    solves to rtol, and a fixed value gives bitwise-repeatable results."""
    print("batch autotune", flush=True)
    level, K = 5, 8
    pm = P.Mesh.load(oracle_mesh(level)[0])
    A = P.Matrix.assemble(pm, P.HELMHOLTZ, CM, CL)
    n = A.n_rows
    assert P.batch_shape(n, K) == 1          # untuned: 10241 rows, 8 columns -> 321 blocks
    best, ms = A.tune_batch(K, restart=30)
    assert best in (1, 2, 4, 8, 16) and P.batch_shape(n, K) == best
    assert all(t > 0 for t in ms.values()) and ms[best] == min(ms.values())
    B = _rhs(level, K)
    dB = _dev(B)
    try:
        for rpt in (1, 2, 4, 8, 16):
            P.set_batch_shape(n, K, P.F64, rpt)
            X1 = torch.zeros_like(dB)
            assert all(i["converged"] for i in A.solve_batch(X1, dB, rtol=1e-10))
            _check_columns(level, B, X1.cpu().numpy(), 1e-10)
            X2 = torch.zeros_like(dB)
            A.solve_batch(X2, dB, rtol=1e-10)
            assert torch.equal(X1, X2), rpt
    finally:
        P.set_batch_shape(n, K, P.F64, 0)
    assert P.batch_shape(n, K) == 1
    with pytest.raises(P.PhelmError):
        P.set_batch_shape(n, K, P.F64, 3)


def test_batch_fused_first_pass_gives_the_same_bits(tmp_path):
    """The fused CGS2 first pass (update + next dots in one read of the basis)
    makes the same operations in the same order as the two-kernel form
    (PHG_BATCH_FUSED=0, run in a child process): identical solutions."""
    import os
    import subprocess
    import sys
    print("batch fused vs unfused", flush=True)
    level, K = 4, 8
    B = _rhs(level, K)
    np.save(tmp_path / "B.npy", B)
    pm = P.Mesh.load(oracle_mesh(level)[0])
    A = P.Matrix.assemble(pm, P.HELMHOLTZ, CM, CL)
    for prec in (P.F64, P.F32):
        X = torch.zeros(B.shape, dtype=torch.float64, device=DEV)
        A.solve_batch(X, _dev(B), rtol=1e-10, ortho=P.CGS2, precision=prec)
        np.save(tmp_path / ("X%d.npy" % prec), X.cpu().numpy())
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = (
        "import sys, numpy as np, torch\n"
        "sys.path.insert(0, %r)\n"
        "import mke_amd as P\n"
        "from tests._common import oracle_mesh\n"
        "from tests.test_gpu_batch import CM, CL\n"
        "B = np.load(%r)\n"
        "A = P.Matrix.assemble(P.Mesh.load(oracle_mesh(%d)[0]), P.HELMHOLTZ, CM, CL)\n"
        "for prec in (P.F64, P.F32):\n"
        "    X = torch.zeros(B.shape, dtype=torch.float64, device='cuda:0')\n"
        "    A.solve_batch(X, torch.as_tensor(B, device='cuda:0'), rtol=1e-10, ortho=P.CGS2,"
        " precision=prec)\n"
        "    np.save(%r %% prec, X.cpu().numpy())\n"
    ) % (root, str(tmp_path / "B.npy"), level, str(tmp_path / "U%d.npy"))
    env = dict(os.environ, PHG_BATCH_FUSED="0")
    r = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    for prec in (P.F64, P.F32):
        assert np.array_equal(np.load(tmp_path / ("X%d.npy" % prec)),
                              np.load(tmp_path / ("U%d.npy" % prec))), prec


def test_batch_diagonal_right_preconditioning():
    """diag_scale solves (A D^-1) y = b, x = D^-1 y: the same true residual
    bound, in fewer steps for the cos-weighted mass matrix (its diagonal
    spans the cos(latitude) range)."""
    print("batch diag scale", flush=True)
    level = 5
    pm = P.Mesh.load(oracle_mesh(level)[0])
    J = P.SphereJacobian(pm)
    M = J.idt_
    B = _rhs(level, 8)
    dB = _dev(B)
    X0, X1 = torch.zeros_like(dB), torch.zeros_like(dB)
    i0 = M.solve_batch(X0, dB, rtol=1e-13, ortho=P.CGS1, restart=32)
    i1 = M.solve_batch(X1, dB, rtol=1e-13, ortho=P.CGS1, restart=32, diag_scale=True)
    assert all(i["converged"] for i in i0 + i1)
    assert max(i["iterations"] for i in i1) < max(i["iterations"] for i in i0), (i0, i1)
    rp, ci, v = oracle_matrix(level, O.KIND_JAC_MASS)
    X1 = X1.cpu().numpy()
    for q in range(8):
        r = B[:, q] - O.spmv(rp, ci, v, X1[:, q])
        assert np.linalg.norm(r) / np.linalg.norm(B[:, q]) <= 1e-13 * 1.000001, q
