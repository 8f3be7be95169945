"""GPU test of the C++ host classes (include/phelm_gpu.hpp) through the
compiled check program tests/cpp/mirror_check: Solver<double> built by
add() equals the GPU-assembled matrix bit for bit, and Solver::solve,
SphereLaplace::calc1/calc2/solve, SphereNorm and SphereJacobian agree with
the oracle's matrices (scipy sparse direct solves) or, for the Jacobian and
BarVortex, with the Python host classes bit for bit (same library, same
options)."""
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")
sp = pytest.importorskip("scipy.sparse")
spla = pytest.importorskip("scipy.sparse.linalg")

import mke_amd as P
from oracle import oracle as O
from tests._common import oracle_matrix, oracle_mesh

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "mirror_check")
DEV = "cuda:0"
LEVEL = 4


def _csr(t, shape):
    rp, ci, v = t
    return sp.csr_matrix((v, ci, rp), shape=shape)


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/mirror_check is not built (run __graft_entry__.build())")
    out = str(tmp_path_factory.mktemp("mirror"))
    print("mirror_check level %d" % LEVEL, flush=True)
    r = subprocess.run([EXE, str(LEVEL), out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    summ = {}
    for line in open(os.path.join(out, "summary.txt")):
        k, v = line.split()
        summ[k] = float(v)
    load = lambda name: np.fromfile(os.path.join(out, name), dtype=np.float64)  # noqa: E731
    return summ, load


def test_solver_add_and_solve(run):
    summ, load = run
    assert np.array_equal(load("spmv_add.bin"), load("spmv_asm.bin"))
    cm, cl = 1.0, -1e-2
    rp, ci, v = oracle_matrix(LEVEL, O.KIND_HELMHOLTZ, cm, cl)
    b = O.spmv(rp, ci, v, load("x.bin"))
    assert np.array_equal(b, load("spmv_asm.bin"))
    x = load("solve_x.bin")
    rel = np.linalg.norm(b - O.spmv(rp, ci, v, x)) / np.linalg.norm(b)
    assert summ["solve_converged"] == 1 and rel <= 1e-10
    assert np.linalg.norm(x - load("x.bin")) / np.linalg.norm(load("x.bin")) <= 1e-8


def test_sphere_laplace(run):
    summ, load = run
    text, om = oracle_mesh(LEVEL)
    n, V = om.n_inner, len(load("u.bin"))
    inner = np.asarray(om.inner)
    outer = np.setdiff1d(np.arange(V), inner)
    Mm = _csr(oracle_matrix(LEVEL, O.KIND_MASS), (n, n))
    Lm = _csr(oracle_matrix(LEVEL, O.KIND_LAPLACE), (n, n))
    B3 = _csr(om.assemble_boundary(O.KIND_LAPLACE), (n, len(outer)))
    u, v = load("u.bin"), load("v.bin")
    # calc2 = M^-1 (L u_in + B3 u_out)
    want = spla.spsolve(Mm.tocsc(), Lm @ u[inner] + B3 @ u[outer])
    got = load("lap_calc2.bin")
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-11
    c1 = load("lap_calc1.bin")
    assert np.array_equal(c1[inner], got) and np.all(c1[outer] == 0.0)
    # solve: L x = M v_in, zero at the pole
    rhs = Mm @ v[inner]
    xs = load("lap_solve.bin")
    assert summ["lap_solve_converged"] == 1
    assert np.linalg.norm(rhs - Lm @ xs[inner]) / np.linalg.norm(rhs) <= 1e-10
    xd = spla.spsolve(Lm.tocsc(), rhs)
    assert np.linalg.norm(xs[inner] - xd) / np.linalg.norm(xd) <= 1e-7
    assert np.all(xs[outer] == 0.0)


def test_sphere_norm(run):
    summ, load = run
    _, om = oracle_mesh(LEVEL)
    u, v = load("u.bin"), load("v.bin")
    V = len(u)
    N = _csr(om.assemble_full(O.KIND_MASS), (V, V))
    assert abs(summ["norm_u"] - np.sqrt(u @ (N @ u))) <= 1e-12 * summ["norm_u"]
    assert abs(summ["scalar_uv"] - u @ (N @ v)) <= 1e-12 * np.sqrt((u @ (N @ u)) * (v @ (N @ v)))
    d = u - v
    assert abs(summ["dist_uv"] - np.sqrt(d @ (N @ d))) <= 1e-12 * summ["dist_uv"]


def test_sphere_jacobian_and_barvortex_equal_python(run):
    summ, load = run
    m = P.Mesh.sphere(LEVEL)
    du = torch.as_tensor(load("u.bin"), device=DEV)
    dv = torch.as_tensor(load("v.bin"), device=DEV)
    J = P.SphereJacobian(m)
    ans = torch.zeros(m.inner_size, dtype=torch.float64, device=DEV)
    J.calc2(ans, du, dv)
    assert np.array_equal(ans.cpu().numpy(), load("jac_calc2.bin"))
    assert np.array_equal(-load("jac_calc2.bin"), load("jac_calc2t.bin"))
    pr = m.arrays()["pr"]
    lh = torch.as_tensor(2.0 * np.clip(pr[:, 2], -1, 1), device=DEV)
    bv = P.BarVortex(m, lh, tau=600.0 * 7.292e-5, sigma=1.6e-2, mu=8e-5, k1=1.0, k2=1.0,
                     theta=0.5, slots=1)
    u = torch.as_tensor(load("psi.bin"), device=DEV).reshape(1, -1)
    u1 = torch.zeros_like(u)
    bv.calc(u1, u)
    assert summ["bv_converged"] == 1
    assert np.array_equal(u1.cpu().numpy()[0], load("bv_u1.bin"))
    # calc with a pole value of psi (bnd_u = 1, bnd_w = 0)
    ub = torch.zeros_like(u)
    one = torch.ones(1, m.outer_size, dtype=torch.float64, device=DEV)
    bv.calc(ub, u, bnd_u=one)
    assert np.array_equal(ub.cpu().numpy()[0], load("bv_u1_bnd.bin"))
    # the linearised and adjoint steps about psi with that pole value
    for name, f in (("bv_L_bnd.bin", bv.calc_L), ("bv_LT_bnd.bin", bv.calc_LT)):
        out = torch.zeros_like(u)
        f(out, u, u, bnd_u=one)
        assert np.array_equal(out.cpu().numpy()[0], load(name)), name
    assert ub[0, int(m.arrays()["outer"][0])].item() == 1.0


def test_barvortex_callbacks_and_linear_steps_equal_python(run):
    """BarVortex(m, conf) with the rp / coriolis callbacks and zero boundary
    arrays, and L_step / L_1_step / LT_step, equal the Python host class
    given the same mesh functions, bit for bit."""
    summ, load = run
    assert summ["bv_rp_converged"] == 1
    m = P.Mesh.sphere(LEVEL)
    bv = P.BarVortex(m, torch.as_tensor(load("bv_rp_lh.bin"), device=DEV), tau=600.0 * 7.292e-5,
                     sigma=1.6e-2, mu=8e-5, k1=1.0, k2=1.0, theta=0.5, slots=1)
    psi = torch.as_tensor(load("psi.bin"), device=DEV).reshape(1, -1)
    u1 = torch.zeros_like(psi)
    bv.calc(u1, psi, rp=torch.as_tensor(load("bv_rp_f.bin"), device=DEV).reshape(1, -1))
    assert np.array_equal(u1.cpu().numpy()[0], load("bv_rp_u1.bin"))
    du = torch.as_tensor(load("u.bin"), device=DEV).reshape(1, -1)
    dv = torch.as_tensor(load("v.bin"), device=DEV).reshape(1, -1)
    a, b, c = (torch.zeros_like(psi) for _ in range(3))
    bv.L_step(a, du, psi)
    bv.L_1_step(b, a, psi)
    bv.LT_step(c, dv, psi)
    assert np.array_equal(a.cpu().numpy()[0], load("bv_L.bin"))
    assert np.array_equal(b.cpu().numpy()[0], load("bv_L1.bin"))
    assert np.array_equal(c.cpu().numpy()[0], load("bv_LT.bin"))


def test_python_sphere_laplace_and_norm_equal_cpp(run):
    """The Python SphereLaplace / SphereNorm methods make the same library
    calls with the same options as the C++ classes: same bits."""
    summ, load = run
    m = P.Mesh.sphere(LEVEL)
    du = torch.as_tensor(load("u.bin"), device=DEV)
    dv = torch.as_tensor(load("v.bin"), device=DEV)
    L = P.SphereLaplace(m)
    c2 = torch.zeros(m.inner_size, dtype=torch.float64, device=DEV)
    L.calc2(c2, du)
    assert np.array_equal(c2.cpu().numpy(), load("lap_calc2.bin"))
    c1 = torch.zeros(m.n_points, dtype=torch.float64, device=DEV)
    L.calc1(c1, du)
    assert np.array_equal(c1.cpu().numpy(), load("lap_calc1.bin"))
    ls = torch.zeros(m.n_points, dtype=torch.float64, device=DEV)
    info = L.solve(ls, dv)
    assert info["converged"] and np.array_equal(ls.cpu().numpy(), load("lap_solve.bin"))
    N = P.SphereNorm(m)
    assert N.norm(du) == summ["norm_u"]
    assert N.scalar(du, dv) == summ["scalar_uv"]
    assert N.dist(du, dv) == summ["dist_uv"]


def test_baroclin_equals_python(run):
    """The C++ Baroclin (callbacks evaluated on the host at the zone-0 chart
    coordinates) and the Python Baroclin given the same forcing vectors make
    the same library call: same bits."""
    summ, load = run
    m = P.Mesh.sphere(LEVEL)
    bc = P.Baroclin(m, torch.as_tensor(load("bc_lh.bin"), device=DEV), tau=1e-3, sigma=1.6e-2,
                    mu=8e-2, sigma1=1.6e-2, mu1=8e-2, alpha=1.0)
    u1 = torch.as_tensor(load("u.bin"), device=DEV).reshape(1, -1)
    u2 = torch.as_tensor(load("v.bin"), device=DEV).reshape(1, -1)
    o1, o2 = torch.zeros_like(u1), torch.zeros_like(u2)
    info = bc.calc(o1, o2, u1, u2, torch.as_tensor(load("bc_f.bin"), device=DEV),
                   torch.as_tensor(load("bc_g.bin"), device=DEV))[0]
    assert summ["bc_converged"] == 1 and info["converged"]
    assert info["iterations"] == summ["bc_iterations"]
    assert np.array_equal(o1.cpu().numpy()[0], load("bc_u11.bin"))
    assert np.array_equal(o2.cpu().numpy()[0], load("bc_u21.bin"))
    # L_step of (u, v) about (psi, u), L_1_step of that; LT_step changes nothing
    psi = torch.as_tensor(load("psi.bin"), device=DEV).reshape(1, -1)
    l1, l2, k1, k2 = (torch.zeros_like(u1) for _ in range(4))
    bc.L_step(l1, l2, u1, u2, psi, u1)
    bc.L_1_step(k1, k2, l1, l2, psi, u1)
    for t, name in ((l1, "bc_L1.bin"), (l2, "bc_L2.bin"), (k1, "bc_B1.bin"), (k2, "bc_B2.bin")):
        assert np.array_equal(t.cpu().numpy()[0], load(name)), name
    assert np.array_equal(load("bc_LT1.bin"), load("bc_L1.bin"))


def test_baroclin_linear_steps_with_boundary_values_equal_python(run):
    """calc_L of (u, v) about (psi, u) and calc_L_1 of that with a pole value
    of 0.5 through the C++ class and through the Python class: same bits, the
    pole value in the results."""
    _, load = run
    m = P.Mesh.sphere(LEVEL)
    bc = P.Baroclin(m, torch.as_tensor(load("bc_lh.bin"), device=DEV), tau=1e-3, sigma=1.6e-2,
                    mu=8e-2, sigma1=1.6e-2, mu1=8e-2, alpha=1.0)
    u1 = torch.as_tensor(load("u.bin"), device=DEV).reshape(1, -1)
    u2 = torch.as_tensor(load("v.bin"), device=DEV).reshape(1, -1)
    psi = torch.as_tensor(load("psi.bin"), device=DEV).reshape(1, -1)
    bnd = torch.full((1, m.outer_size), 0.5, dtype=torch.float64, device=DEV)
    l1, l2, k1, k2 = (torch.zeros_like(u1) for _ in range(4))
    bc.calc_L(l1, l2, u1, u2, psi, u1, bnd=bnd)
    bc.calc_L_1(k1, k2, l1, l2, psi, u1, bnd=bnd)
    outer = m.arrays()["outer"]
    for t, name in ((l1, "bc_Lb1.bin"), (l2, "bc_Lb2.bin"), (k1, "bc_Bb1.bin"),
                    (k2, "bc_Bb2.bin")):
        assert np.array_equal(t.cpu().numpy()[0], load(name)), name
        assert np.all(load(name)[outer] == 0.5), name


def test_sphere_chafe_cpp_vs_oracle(run):
    """Two SphereChafe::solve steps of the C++ class (A_ solved to 1e-14)
    against the oracle restatement (sparse LU) from the same inputs."""
    import oracle.schafe as S
    from oracle import mesh_ref
    summ, load = run
    a = mesh_ref.load_text(mesh_ref.sphere_mesh_text(LEVEL))
    om = O.Mesh(a)
    xy = np.asarray(a["coords"]).reshape(len(om.p2io), -1, 2)[:, 0, :]
    osc = S.SphereChafe(om, xy, 0.01, 70.0, 1.0)
    U = osc.solve(load("sc_u0.bin"), load("sc_b1.bin"), 0.0)
    g1 = load("sc_u1.bin")
    assert S.norm_dist(om, g1, U) <= 1e-11 * S.norm_dist(om, U, 0 * U)
    U = osc.solve(U, load("sc_b2.bin"), 0.01)
    g2 = load("sc_u2.bin")
    assert S.norm_dist(om, g2, U) <= 1e-11 * S.norm_dist(om, U, 0 * U)
    assert summ["sc_rel_residual"] <= 1e-13
