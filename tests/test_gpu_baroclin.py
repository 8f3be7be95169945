"""GPU Baroclin step (phg_baroclin_*) against the oracle restatement
(oracle/baroclin.py: literal integrate_cb blocks, element-order right part,
direct LU solve of the 4n x 4n system), through the C ABI.

Bars: A_ (4n x 4n CSR) bit-identical to the oracle's; one step of two
members (the reference driver's manufactured pair and a perturbed copy, with
its forcing) within 1e-9 of the oracle step (SphereNorm distance relative to
the state) with the block preconditioner, within 1e-6 without it (the plain
solve's forward error at its 1e-12 residual); a member's result
bit-identical alone and beside another member; the manufactured-solution
tendency error falls with refinement as the oracle's does; the linearised
steps calc_L (within 1e-9 of the oracle's) and calc_L_1 (within 1e-8: up to
10 fixed-point passes, stopping at a distance of 1e-8), member bits
independent of the batch, and L_1(L(u)) = u without the Jacobian terms; the
same bars for the linearised steps with boundary (pole) values."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytest.importorskip("scipy")

import mke_amd as P
from oracle import mesh_ref, oracle as O
from oracle.baroclin import Baroclin as OB, manufactured

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
PRM = dict(sigma=1.6e-2, mu=8e-2, sigma1=1.6e-2, mu1=8e-2, alpha=1.0)
TAU = 1e-3


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


def _setup(level):
    text = mesh_ref.sphere_mesh_text(level)
    a = mesh_ref.load_text(text)
    om = O.Mesh(a)
    xy = np.asarray(a["coords"]).reshape(len(om.p2io), -1, 2)[:, 0, :]
    return P.Mesh.load(text), om, xy


def test_baroclin_matrix_bitwise():
    print("baroclin matrix", flush=True)
    pm, om, xy = _setup(3)
    bc = P.Baroclin(pm, _dev(np.zeros(len(xy))), tau=TAU, slots=1, **PRM)
    rp, ci, v = bc.matrix().csr()
    orp, oci, ov = O.baroclin_matrix(*O.baroclin_blocks(om, TAU, PRM["sigma"], PRM["mu"],
                                                       PRM["sigma1"], PRM["mu1"], PRM["alpha"]))
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and np.array_equal(v, ov)


def _states(xy, om):
    u1, u2, f, g = manufactured(xy, 0.0, **PRM)
    rng = np.random.default_rng(5)
    bump = np.exp(-8.0 * ((xy[:, 0] - 0.3) ** 2 + (np.cos(xy[:, 1]) - 1.0) ** 2))
    bump[om.outer] = 0.0
    U1 = np.stack([u1, u1 + 0.05 * bump])
    U2 = np.stack([u2, u2 - 0.03 * bump * rng.uniform(0.5, 1.0)])
    return U1, U2, f, g


@pytest.mark.parametrize("precondition", [1, 0])
def test_baroclin_step_vs_oracle(precondition):
    print("baroclin step precondition=%d" % precondition, flush=True)
    pm, om, xy = _setup(3)
    U1, U2, f, g = _states(xy, om)
    bc = P.Baroclin(pm, _dev(np.zeros(len(xy))), tau=TAU, slots=2, **PRM)
    D1, D2 = torch.zeros_like(_dev(U1)), torch.zeros_like(_dev(U2))
    infos = bc.calc(D1, D2, _dev(U1), _dev(U2), _dev(f), _dev(g),
                    opts=P.bc_opts(precondition=precondition, max_solver_iters=20000))
    ob = OB(om, tau=TAU, lh=np.zeros(len(xy)), **PRM)
    got1, got2 = D1.cpu().numpy(), D2.cpu().numpy()
    for q in range(2):
        r1, r2, it, nr = ob.calc(U1[q], U2[q], f, g)
        assert infos[q]["converged"] and abs(infos[q]["iterations"] - it) <= 1, (infos[q], it)
        # the 4n solve stops at a relative residual of 1e-12; cond(A_) = 1.2e4 at
        # level 3, and the plain GMRES iterate carries that error into the
        # (small) u-blocks: 1.2e-7 measured; the preconditioned one 1e-9
        tol = 1e-9 if precondition else 1e-6
        assert ob.dist(got1[q], r1) <= tol * ob.bv.norm(r1), q
        assert ob.dist(got2[q], r2) <= tol * ob.bv.norm(r2), q
        assert ob.dist(got1[q], U1[q]) > 1e-4 * ob.bv.norm(U1[q])   # the step moved u1


def test_baroclin_member_independent_of_batch():
    print("baroclin independence", flush=True)
    pm, om, xy = _setup(3)
    U1, U2, f, g = _states(xy, om)
    bc = P.Baroclin(pm, _dev(np.zeros(len(xy))), tau=TAU, slots=2, **PRM)
    D1, D2 = torch.zeros_like(_dev(U1)), torch.zeros_like(_dev(U2))
    bc.calc(D1, D2, _dev(U1), _dev(U2), _dev(f), _dev(g))
    a1 = torch.zeros_like(D1[1:])
    a2 = torch.zeros_like(D2[1:])
    bc.calc(a1, a2, _dev(U1[1:]), _dev(U2[1:]), _dev(f), _dev(g))
    # member 1 alone runs in slot 0 of the same 2-slot engine
    assert torch.equal(a1[0], D1[1]) and torch.equal(a2[0], D2[1])


def test_baroclin_manufactured_tendency_converges():
    print("baroclin manufactured", flush=True)
    errs = []
    for level in (3, 4, 5):
        pm, om, xy = _setup(level)
        bc = P.Baroclin(pm, _dev(np.zeros(len(xy))), tau=TAU, slots=1, **PRM)
        u1, u2, f, g = manufactured(xy, 0.0, **PRM)
        r1, r2, _, _ = manufactured(xy, TAU, **PRM)
        D1, D2 = torch.zeros_like(_dev(u1)), torch.zeros_like(_dev(u2))
        info = bc.calc(D1, D2, _dev(u1), _dev(u2), _dev(f), _dev(g))[0]
        assert info["converged"], info
        N = O.Mesh(mesh_ref.load_text(mesh_ref.sphere_mesh_text(level))).assemble_full(O.KIND_MASS)
        nrm = lambda d: float(np.sqrt(d @ O.spmv(*N, d)))  # noqa: E731
        v1, v2 = D1.cpu().numpy(), D2.cpu().numpy()
        errs.append((nrm((v1 - u1) - (r1 - u1)) / nrm(r1 - u1),
                     nrm((v2 - u2) - (r2 - u2)) / nrm(r2 - u2)))
    # the oracle gives 0.118 / 0.040 / 0.0072 (layer 1) and 0.190 / 0.079 /
    # 0.037 (layer 2) at levels 3 / 4 / 5
    assert errs[1][0] < errs[0][0] / 2.5 and errs[2][0] < errs[1][0] / 2.5, errs
    assert errs[1][1] < errs[0][1] / 1.8 and errs[2][1] < errs[1][1] / 1.8, errs


def _lin_fields(pm, om, seeds):
    pr = pm.arrays()["pr"]
    x, y, z = pr[:, 0], pr[:, 1], pr[:, 2]
    out = []
    for sd in seeds:
        r = np.random.default_rng(sd)
        f = (r.standard_normal() * x * y + r.standard_normal() * (x * x - y * y) * z
             + r.standard_normal() * np.sin(2 * x + y) * z + 0.3 * r.standard_normal() * y)
        out.append(f - f[om.outer[0]])
    return np.stack(out), 2.0 * np.clip(z, -1, 1)


def _rel2(ob, a1, a2, r1, r2):
    return float(np.hypot(ob.dist(a1, r1), ob.dist(a2, r2))
                 / np.hypot(ob.bv.norm(r1), ob.bv.norm(r2)))


def test_baroclin_linear_steps_vs_oracle():
    print("baroclin calc_L / calc_L_1 level=3", flush=True)
    pm, om, xy = _setup(3)
    U1, lh = _lin_fields(pm, om, (1, 2))
    U2, _ = _lin_fields(pm, om, (3, 4))
    Z1, _ = _lin_fields(pm, om, (5, 6))
    Z2, _ = _lin_fields(pm, om, (7, 8))
    tau = 2.5e-4
    bc = P.Baroclin(pm, _dev(lh), tau=tau, slots=2, **PRM)
    ob = OB(om, tau=tau, lh=lh, **PRM)
    L1, L2 = torch.zeros_like(_dev(U1)), torch.zeros_like(_dev(U2))
    for i in bc.calc_L(L1, L2, _dev(U1), _dev(U2), _dev(Z1), _dev(Z2)):
        assert i["converged"] and i["iterations"] == 1 and i["it_jacobian"] > 0, i
    B1, B2 = torch.zeros_like(L1), torch.zeros_like(L2)
    inf_b = bc.calc_L_1(B1, B2, L1, L2, _dev(Z1), _dev(Z2))
    g = [t.cpu().numpy() for t in (L1, L2, B1, B2)]
    for q in range(2):
        r1, r2, _, _ = ob.calc_lin(0, U1[q], U2[q], Z1[q], Z2[q])
        assert _rel2(ob, g[0][q], g[1][q], r1, r2) <= 1e-9, q
        r1, r2, it, nr = ob.calc_lin(1, g[0][q], g[1][q], Z1[q], Z2[q])
        assert abs(inf_b[q]["iterations"] - it) <= 1, (inf_b[q], it)
        assert _rel2(ob, g[2][q], g[3][q], r1, r2) <= 1e-8, q
        assert _rel2(ob, g[2][q], g[3][q], U1[q], U2[q]) <= 3e-4   # inverts to O(tau^2)
    # member 1 alone (slot 0 of the same engine) gives the same bits
    a1, a2 = torch.zeros_like(L1[:1]), torch.zeros_like(L2[:1])
    bc.calc_L(a1, a2, _dev(U1[1:]), _dev(U2[1:]), _dev(Z1[1:]), _dev(Z2[1:]))
    assert torch.equal(a1[0], L1[1]) and torch.equal(a2[0], L2[1])


def test_baroclin_inverse_step_without_jacobian():
    print("baroclin L_1(L(u)) level=3", flush=True)
    pm, om, xy = _setup(3)
    U1, _ = _lin_fields(pm, om, (1,))
    U2, _ = _lin_fields(pm, om, (2,))
    zero = np.zeros((1, pm.n_points))
    bc = P.Baroclin(pm, _dev(zero[0]), tau=1e-2, slots=1, **PRM)
    ob = OB(om, tau=1e-2, lh=zero[0], **PRM)
    L1, L2, B1, B2 = (torch.zeros_like(_dev(U1)) for _ in range(4))
    bc.L_step(L1, L2, _dev(U1), _dev(U2), _dev(zero), _dev(zero))
    bc.L_1_step(B1, B2, L1, L2, _dev(zero), _dev(zero))
    g = [t.cpu().numpy()[0] for t in (L1, B1, B2)]
    assert ob.dist(g[0], U1[0]) > 1e-3 * ob.bv.norm(U1[0])
    assert _rel2(ob, g[1], g[2], U1[0], U2[0]) <= 1e-9


def test_baroclin_boundary_values_vs_oracle():
    """Baroclin::calc with non-zero boundary (pole) values, one array for w1,
    w2, u1, u2 (baroclin.cpp:445-606), two members with different values,
    against the oracle; afterwards a call without values on the same object
    gives the plain step's bits."""
    print("baroclin boundary values", flush=True)
    pm, om, xy = _setup(3)
    U1, U2, f, g = _states(xy, om)
    assert pm.outer_size == len(om.outer) >= 1
    bnd = np.array([[0.02], [-0.03]]) * np.ones((2, pm.outer_size))
    opts = P.bc_opts(precondition=1, max_solver_iters=20000)
    bc = P.Baroclin(pm, _dev(np.zeros(len(xy))), tau=TAU, slots=2, **PRM)
    plain1, plain2 = torch.zeros_like(_dev(U1)), torch.zeros_like(_dev(U2))
    bc.calc(plain1, plain2, _dev(U1), _dev(U2), _dev(f), _dev(g), opts=opts)
    D1, D2 = torch.zeros_like(_dev(U1)), torch.zeros_like(_dev(U2))
    infos = bc.calc(D1, D2, _dev(U1), _dev(U2), _dev(f), _dev(g), opts=opts, bnd=_dev(bnd))
    ob = OB(om, tau=TAU, lh=np.zeros(len(xy)), **PRM)
    got1, got2 = D1.cpu().numpy(), D2.cpu().numpy()
    for q in range(2):
        r1, r2, it, nr = ob.calc(U1[q], U2[q], f, g, bnd=bnd[q])
        assert infos[q]["converged"] and abs(infos[q]["iterations"] - it) <= 1, (infos[q], it)
        assert ob.dist(got1[q], r1) <= 1e-9 * ob.bv.norm(r1), q
        assert ob.dist(got2[q], r2) <= 1e-9 * ob.bv.norm(r2), q
        assert np.all(got1[q][om.outer] == bnd[q]) and np.all(got2[q][om.outer] == bnd[q])
        p1 = plain1.cpu().numpy()[q]
        assert ob.dist(got1[q], p1) > 1e-6 * ob.bv.norm(p1)    # the values matter
    again1, again2 = torch.zeros_like(D1), torch.zeros_like(D2)
    bc.calc(again1, again2, _dev(U1), _dev(U2), _dev(f), _dev(g), opts=opts)
    assert torch.equal(again1, plain1) and torch.equal(again2, plain2)


def test_baroclin_linear_steps_after_boundary_values():
    """calc_L and calc_L_1 after a calc with non-zero boundary values on the
    same object give the bits of a fresh object: the linearised steps run with
    zero boundary values, so the right part's u rows that the boundary step
    wrote must not leak into their solves."""
    print("baroclin calc_L after calc with boundary values", flush=True)
    pm, om, xy = _setup(3)
    U1, U2, f, g = _states(xy, om)
    L1f, lh = _lin_fields(pm, om, (1, 2))
    L2f, _ = _lin_fields(pm, om, (3, 4))
    Z1, _ = _lin_fields(pm, om, (5, 6))
    Z2, _ = _lin_fields(pm, om, (7, 8))
    tau = 2.5e-4
    bnd = np.array([[0.02], [-0.03]]) * np.ones((2, pm.outer_size))
    opts = P.bc_opts(precondition=1, max_solver_iters=20000)
    used = P.Baroclin(pm, _dev(lh), tau=tau, slots=2, **PRM)
    D1, D2 = torch.zeros_like(_dev(U1)), torch.zeros_like(_dev(U2))
    used.calc(D1, D2, _dev(U1), _dev(U2), _dev(f), _dev(g), opts=opts, bnd=_dev(bnd))
    fresh = P.Baroclin(pm, _dev(lh), tau=tau, slots=2, **PRM)
    outs = []
    for bc in (used, fresh):
        L1, L2 = torch.zeros_like(_dev(L1f)), torch.zeros_like(_dev(L2f))
        bc.calc_L(L1, L2, _dev(L1f), _dev(L2f), _dev(Z1), _dev(Z2))
        B1, B2 = torch.zeros_like(L1), torch.zeros_like(L2)
        bc.calc_L_1(B1, B2, L1, L2, _dev(Z1), _dev(Z2))
        outs.append((L1, L2, B1, B2))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_baroclin_linear_steps_with_boundary_values_vs_oracle():
    """calc_L and calc_L_1 with boundary (pole) values (baroclin.cpp:620-986:
    the one bnd array in every Laplacian, the right part of A_ or Ab_ and the
    p2u), two members with different values, Jacobian terms on, against the
    oracle; the results hold the values at the boundary points, and a call
    without values afterwards gives a fresh object's bits."""
    print("baroclin calc_L / calc_L_1 with boundary values level=3", flush=True)
    pm, om, xy = _setup(3)
    U1, lh = _lin_fields(pm, om, (1, 2))
    U2, _ = _lin_fields(pm, om, (3, 4))
    Z1, _ = _lin_fields(pm, om, (5, 6))
    Z2, _ = _lin_fields(pm, om, (7, 8))
    tau = 2.5e-4
    bnd = np.array([[0.3], [-0.45]]) * np.ones((2, pm.outer_size))
    bc = P.Baroclin(pm, _dev(lh), tau=tau, slots=2, **PRM)
    ob = OB(om, tau=tau, lh=lh, **PRM)
    L1, L2 = torch.zeros_like(_dev(U1)), torch.zeros_like(_dev(U2))
    for i in bc.calc_L(L1, L2, _dev(U1), _dev(U2), _dev(Z1), _dev(Z2), bnd=_dev(bnd)):
        assert i["converged"] and i["iterations"] == 1, i
    B1, B2 = torch.zeros_like(L1), torch.zeros_like(L2)
    inf_b = bc.calc_L_1(B1, B2, L1, L2, _dev(Z1), _dev(Z2), bnd=_dev(bnd))
    g = [t.cpu().numpy() for t in (L1, L2, B1, B2)]
    for q in range(2):
        r1, r2, _, _ = ob.calc_lin(0, U1[q], U2[q], Z1[q], Z2[q], bnd=bnd[q])
        assert _rel2(ob, g[0][q], g[1][q], r1, r2) <= 1e-9, q
        p1, p2, _, _ = ob.calc_lin(0, U1[q], U2[q], Z1[q], Z2[q])
        assert _rel2(ob, r1, r2, p1, p2) > 1e-4, q                  # the values matter
        r1, r2, it, nr = ob.calc_lin(1, g[0][q], g[1][q], Z1[q], Z2[q], bnd=bnd[q])
        assert inf_b[q]["converged"] and abs(inf_b[q]["iterations"] - it) <= 1, (inf_b[q], it)
        assert _rel2(ob, g[2][q], g[3][q], r1, r2) <= 1e-8, q
        for t in g:
            assert np.all(t[q][om.outer] == bnd[q]), q
    fresh = P.Baroclin(pm, _dev(lh), tau=tau, slots=2, **PRM)
    outs = []
    for b in (bc, fresh):
        a1, a2 = torch.zeros_like(L1), torch.zeros_like(L2)
        b.calc_L(a1, a2, _dev(U1), _dev(U2), _dev(Z1), _dev(Z2))
        c1, c2 = torch.zeros_like(L1), torch.zeros_like(L2)
        b.calc_L_1(c1, c2, a1, a2, _dev(Z1), _dev(Z2))
        outs.append((a1, a2, c1, c2))
    for a, b in zip(*outs):
        assert torch.equal(a, b)
