"""The workloads of the benchmark's configurations 3-5 at their own sizes,
on the GPU, checked with the oracle:

* level 8 (655,361 unknowns), f of degree <= 64 with l^-2 coefficients:
  mixed FP32 and BF16-stored inner solves under FP64 refinement, FP32 MGS
  in persistent cycles, and FP64 MGS; every solution's true FP64 residual
  (oracle CSR SpMV) <= 1e-10 and its forward error against the oracle's
  GMRES driven to 1e-12;
* level 7 x 64 right-hand sides (64 seeds, degree <= 32): block GMRES with
  MFMA block Gram-Schmidt and independent columns, FP32 Krylov + FP64
  refinement; every column's true residual <= 1e-10, forward error of
  three columns against oracle solves driven to 1e-14;
* the served barotropic ensemble on levels 5/6/7/8: every member's state
  after its requests bit-identical to stepping it alone, request after
  request, with a standalone engine; the level-5 member's first step
  within 1e-9 of the oracle's BarVortex::calc (direct solves)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytest.importorskip("scipy")

import mke_amd as P
from oracle import mesh_ref
from oracle import oracle as O
from tests._common import real_sph_harm_field_lean, rh_state

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TAU = 1e-2


def _system(level):
    mesh = P.Mesh.sphere(level)
    cm, cl = P.helmholtz_coeffs(tau=1.0, sigma=0.0, theta=1.0, mu=TAU)
    A = P.Matrix.assemble(mesh, P.HELMHOLTZ, cm, cl)
    M = P.Matrix.assemble(mesh, P.MASS)
    return mesh, A, M


def _rhs(mesh, M, lmax, seed, decay=0.0):
    a = mesh.arrays()
    f = real_sph_harm_field_lean(a["pr"], lmax, seed, decay=decay, device="cuda")[a["inner"]]
    b = torch.zeros(M.n_rows, dtype=torch.float64, device=DEV)
    M.mult_vector(b, torch.as_tensor(np.ascontiguousarray(f), device=DEV))
    return b


def _true_rel(csr, b, x):
    rp, ci, v = csr
    return np.linalg.norm(b - O.spmv(rp, ci, v, x)) / np.linalg.norm(b)


def test_config3_level8_mixed_precision():
    print("config 3: level 8 mixed precision", flush=True)
    mesh, A, M = _system(8)
    assert A.n_rows == 655361 and A.nnz == 4587509
    b = _rhs(mesh, M, 64, 20251203, decay=2.0)
    csr = A.csr()
    bh = b.cpu().numpy()
    xt, oinfo = O.gmres(*csr, bh, restart=30, rtol=1e-12, max_iters=40000)
    assert oinfo["converged"]
    nt = np.linalg.norm(xt)
    for name, kw in (("mixed f32 cgs2", dict(precision=P.F32, ortho=P.CGS2)),
                     ("mixed bf16 cgs2", dict(precision=P.BF16, ortho=P.CGS2)),
                     ("mixed f32 mgs persistent", dict(precision=P.F32, ortho=P.MGS)),
                     ("fp64 mgs", dict(precision=P.F64, ortho=P.MGS))):
        x = torch.zeros_like(b)
        info = A.solve(x, b, restart=30, rtol=1e-10, max_iters=40000, **kw)
        xh = x.cpu().numpy()
        rel = _true_rel(csr, bh, xh)
        fwd = np.linalg.norm(xh - xt) / nt
        print("  %s: %d steps, %d refinements, true rel %.2e, forward error %.2e"
              % (name, info["iterations"], info["refinements"], rel, fwd), flush=True)
        assert info["converged"] and rel <= 1e-10, (name, info, rel)
        assert fwd <= 1e-8, (name, fwd)


def test_config4_level7_64_right_hand_sides():
    print("config 4: level 7 x 64 right-hand sides", flush=True)
    mesh, A, M = _system(7)
    cols = [_rhs(mesh, M, 32, 20251202 + q) for q in range(64)]
    B = torch.stack(cols, 1).contiguous()
    csr = A.csr()
    Bh = B.cpu().numpy()
    check_cols = (0, 17, 63)
    xt = {q: O.gmres(*csr, Bh[:, q].copy(), restart=30, rtol=1e-14, max_iters=40000)[0]
          for q in check_cols}
    for name, solve in (
            ("block gmres, mfma bcgs1", lambda X: A.solve_block(X, B, rtol=1e-10, ortho=P.CGS1)),
            ("independent columns, f32 cgs1",
             lambda X: A.solve_batch(X, B, rtol=1e-10, precision=P.F32, ortho=P.CGS1))):
        X = torch.zeros_like(B)
        info = solve(X)
        Xh = X.cpu().numpy()
        worst = max(_true_rel(csr, Bh[:, q].copy(), Xh[:, q].copy()) for q in range(64))
        fwd = max(np.linalg.norm(Xh[:, q] - xt[q]) / np.linalg.norm(xt[q]) for q in check_cols)
        print("  %s: worst true rel %.2e, forward error %.2e" % (name, worst, fwd), flush=True)
        assert worst <= 1e-10, (name, worst, info if isinstance(info, dict) else None)
        assert fwd <= 1e-8, (name, fwd)


def test_config5_served_ensemble_levels_5_to_8():
    print("config 5: served ensemble on levels 5-8", flush=True)
    levels, slots, sub = (5, 6, 7, 8), (2, 2, 2, 2), (1, 1, 1, 2)
    srv = P.Server(levels=levels, slots=slots, substeps=sub)
    meshes = {lv: P.Mesh.sphere(lv) for lv in levels}
    pr = {lv: meshes[lv].arrays()["pr"] for lv in levels}
    plan = [5, 5, 6, 7, 8]                       # member -> level
    init = [rh_state(pr[lv], seed=30 + m) for m, lv in enumerate(plan)]
    ids = [srv.add_member(lv, init[m]) for m, lv in enumerate(plan)]
    members = np.array([0, 2, 1, 3, 4, 0, 1, 2])
    arrivals = np.cumsum(np.random.default_rng(11).exponential(1.0 / 400.0, len(members)))
    start, finish, infos, st = srv.run(members, arrivals)
    assert st["completed"] == len(members) and all(i["converged"] for i in infos), infos
    c = srv.cfg.conf
    conf = dict(tau=c.tau, sigma=c.sigma, mu=c.mu, k1=c.k1, k2=c.k2, theta=c.theta)
    first_l5 = None
    for m, lv in enumerate(plan):
        k = sub[levels.index(lv)]
        lh = torch.as_tensor(2.0 * np.clip(pr[lv][:, 2], -1, 1), device=DEV)
        bv = P.BarVortex(meshes[lv], lh, slots=2, **dict(conf, tau=conf["tau"] / k))
        u = torch.as_tensor(init[m], device=DEV).reshape(1, -1)
        for step in range(k * int((members == m).sum())):
            u1 = torch.zeros_like(u)
            bv.calc(u1, u)
            u = u1
            if m == 0 and step == 0:
                first_l5 = u.cpu().numpy()[0].copy()
        got = srv.member_state(ids[m], meshes[lv].n_points)
        assert np.array_equal(got, u.cpu().numpy()[0]), (m, lv)
        bv.close()
    srv.close()
    # the level-5 member's first step against the oracle (direct solves)
    from oracle.barvortex import BarVortex as OB
    om = O.Mesh(mesh_ref.load_text(mesh_ref.sphere_mesh_text(5)))
    lh5 = 2.0 * np.clip(pr[5][:, 2], -1, 1)
    ob = OB(om, lh=lh5, **conf)
    ref, it, rel = ob.calc(init[0], max_iters=100)
    err = ob.norm(first_l5 - ref) / ob.norm(ref)
    print("  level-5 step vs oracle: %.2e (oracle %d passes)" % (err, it), flush=True)
    assert err <= 1e-9, err
