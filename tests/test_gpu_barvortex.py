"""GPU BarVortex::calc (phg_barvortex_*) against the oracle restatement
(oracle/barvortex.py, direct solves) and the member-independence property the
serving loop relies on; the right part conf.rp on the reference driver's
analytic case; the linearised, inverse and adjoint steps (calc_L, calc_L_1,
calc_LT) against the oracle's.

Bars: the GPU step's u1 within 1e-9 (SphereNorm-relative) of the oracle's
in at most 2 more fixed-point passes (it may end earlier by the stall rule, once
the change no longer contracts below 1e-11): the GPU solves stop at
relative residuals 1e-12..1e-13; a member's u1 is bit-identical alone, at
another slot, in a full batch, and when admitted while others are running."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import mke_amd as P
from oracle import mesh_ref
from tests._common import oracle_mesh, real_sph_harm_field

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
OMEGA = 7.292e-5
CONF = dict(tau=600 * OMEGA, sigma=1.6e-2, mu=8e-5, k1=1.0, k2=1.0, theta=0.5)


def _states(level, k):
    """Rossby-Haurwitz wave 4 stream function (nondimensional, pole value 0)
    plus k-1 perturbed copies."""
    text, om = oracle_mesh(level)
    pr = np.asarray(mesh_ref.load_text(text)["pr"]).reshape(-1, 3)
    phi = np.arcsin(np.clip(pr[:, 2], -1, 1))
    lam = np.arctan2(pr[:, 1], pr[:, 0])
    w = 7.848e-6 / OMEGA
    psi = -w * np.sin(phi) + w * np.cos(phi) ** 4 * np.sin(phi) * np.cos(4 * lam)
    out = []
    for q in range(k):
        u = psi + (0 if q == 0 else 2e-3 * real_sph_harm_field(pr, 6, 100 + q))
        out.append(u - u[om.outer[0]])
    return om, pr, np.stack(out), 2.0 * np.sin(phi)


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


@pytest.mark.parametrize("level", [3, 4])
def test_barvortex_step_vs_oracle(level):
    print("barvortex level=%d" % level, flush=True)
    from oracle.barvortex import BarVortex as OB
    om, pr, U, lh = _states(level, 2)
    pm = P.Mesh.load(oracle_mesh(level)[0])
    bv = P.BarVortex(pm, _dev(lh), slots=2, **CONF)
    U1 = torch.zeros_like(_dev(U))
    infos = bv.calc(U1, _dev(U))
    ob = OB(om, lh=lh, **CONF)
    got = U1.cpu().numpy()
    for q in range(2):
        ref, it, rel = ob.calc(U[q], max_iters=100)
        assert infos[q]["converged"], infos[q]
        assert 3 <= infos[q]["iterations"] <= it + 2, (infos[q], it)
        err = ob.norm(got[q] - ref) / ob.norm(ref)
        assert err <= 1e-9, (q, err)
        assert ob.norm(got[q] - U[q]) > 1e-5 * ob.norm(U[q])  # the step moved the state


def test_barvortex_member_independent_of_batch():
    print("barvortex independence", flush=True)
    level = 3
    om, pr, U, lh = _states(level, 4)
    pm = P.Mesh.load(oracle_mesh(level)[0])
    bv = P.BarVortex(pm, _dev(lh), slots=4, **CONF)
    U1 = torch.zeros_like(_dev(U))
    bv.calc(U1, _dev(U))                                   # all four together
    alone = torch.zeros_like(U1[:1])
    bv.calc(alone, _dev(U[2:3]))                          # member 2 alone in slot 0
    assert torch.equal(alone[0], U1[2])
    # continuous batching: member 1 runs two passes alone, member 3 joins
    # at slot 2, member 0 at slot 3 later; results equal the batch run
    out = {q: torch.zeros_like(U1[0]) for q in (0, 1, 3)}
    bv.load(0, _dev(U[1]))
    bv.iterate()
    bv.iterate()
    bv.load(2, _dev(U[3]))
    state, _ = bv.iterate()
    bv.load(3, _dev(U[0]))
    where = {0: 1, 2: 3, 3: 0}
    done = set()
    for _ in range(200):
        state, infos = bv.iterate()
        for slot, member in where.items():
            if state[slot] == 2 and slot not in done:
                bv.store(slot, out[member])
                done.add(slot)
        if len(done) == 3:
            break
    assert len(done) == 3
    for q in (0, 1, 3):
        assert torch.equal(out[q], U1[q]), q


def test_barvortex_mixed_precision_vs_oracle():
    level = 3
    from oracle.barvortex import BarVortex as OB
    om, pr, U, lh = _states(level, 1)
    pm = P.Mesh.load(oracle_mesh(level)[0])
    bv = P.BarVortex(pm, _dev(lh), slots=1, **CONF)
    U1 = torch.zeros_like(_dev(U))
    infos = bv.calc(U1, _dev(U), opts=P.bv_opts(precision=P.F32))
    assert infos[0]["converged"]
    ob = OB(om, lh=lh, **CONF)
    ref, it, rel = ob.calc(U[0], max_iters=100)
    err = ob.norm(U1.cpu().numpy()[0] - ref) / ob.norm(ref)
    assert err <= 1e-9, err


def test_barvortex_forced_step_vs_oracle():
    """The reference driver's an1 / rp1 run (examples/test_barvortex.cpp:489-558)
    at level 4: 10 steps of tau = 1e-3 with the right part, no Coriolis."""
    print("barvortex forced an1/rp1 level=4", flush=True)
    from oracle.barvortex import BarVortex as OB, an1, rp1
    text, om = oracle_mesh(4)
    pm = P.Mesh.load(text)
    conf = dict(tau=1e-3, sigma=1.6e-2, mu=8e-5, k1=1.0, k2=1.0, theta=0.5)
    bv = P.BarVortex(pm, _dev(np.zeros(pm.n_points)), slots=1, **conf)
    ob = OB(om, lh=np.zeros(pm.n_points), **conf)
    xy = pm.zone0_xy()
    x, y = xy[:, 0], xy[:, 1]
    u = an1(x, y, 0.0)
    ug, uo, t = _dev(u).reshape(1, -1), u.copy(), 0.0
    while t < 0.01:
        rp = bv.rp_values(rp1, t)
        assert np.array_equal(rp, rp1(x, y, t, conf["mu"], conf["sigma"]))
        out = torch.zeros_like(ug)
        info = bv.calc(out, ug, rp=_dev(rp).reshape(1, -1))[0]
        assert info["converged"], info
        ug = out
        uo, it, rel = ob.calc(uo, forcing=rp)
        if t == 0.0:        # one step: the single-step bar
            assert ob.norm(ug.cpu().numpy()[0] - uo) <= 1e-9 * ob.norm(uo)
        t += conf["tau"]
    got = ug.cpu().numpy()[0]
    # ten steps, each within 1e-9 of the oracle's from the same start
    assert ob.norm(got - uo) <= 1e-8 * ob.norm(uo)
    e_gpu, e_or = ob.norm(got - an1(x, y, t)), ob.norm(uo - an1(x, y, t))
    assert abs(e_gpu - e_or) <= 1e-6 * e_or, (e_gpu, e_or)


def _lin_fields(pr, om, seeds):
    x, y, z = pr[:, 0], pr[:, 1], pr[:, 2]
    out = []
    for sd in seeds:
        r = np.random.default_rng(sd)
        f = (r.standard_normal() * x * y + r.standard_normal() * (x * x - y * y) * z
             + r.standard_normal() * np.sin(2 * x + y) * z + 0.3 * r.standard_normal() * y)
        out.append(f - f[om.outer[0]])
    return np.stack(out)


def test_barvortex_linear_steps_vs_oracle():
    print("barvortex calc_L / calc_L_1 / calc_LT level=3", flush=True)
    from oracle.barvortex import BarVortex as OB
    level = 3
    om, pr, Z, lh = _states(level, 2)
    pm = P.Mesh.load(oracle_mesh(level)[0])
    bv = P.BarVortex(pm, _dev(lh), slots=2, **CONF)
    ob = OB(om, lh=lh, **CONF)
    U, W, Vv = _lin_fields(pr, om, (1, 2)), _lin_fields(pr, om, (3, 4)), _lin_fields(pr, om, (5, 6))
    opts = P.bv_opts(rtol_poisson=1e-12)
    L = torch.zeros_like(_dev(U))
    for i in bv.calc_L(L, _dev(U), _dev(Z), opts=opts):
        assert i["converged"] and i["it_jacobian"] > 0 and i["it_helmholtz"] > 0, i
    L1 = _dev(W).clone()
    bv.calc_L_1(L1, _dev(U), _dev(Z), opts=opts)
    LT = torch.zeros_like(_dev(Vv))
    bv.calc_LT(LT, _dev(Vv), _dev(Z), opts=opts)
    for q in range(2):
        for got, ref in ((L, ob.calc_L(U[q], Z[q])), (L1, ob.calc_L_1(W[q], U[q], Z[q])),
                         (LT, ob.calc_LT(Vv[q], Z[q]))):
            g = got.cpu().numpy()[q]
            assert ob.norm(g - ref) <= 1e-9 * ob.norm(ref), (q, ob.norm(g - ref) / ob.norm(ref))
    # member 1 alone (slot 0 of the same engine) gives the same bits
    alone = torch.zeros_like(L[:1])
    bv.calc_L(alone, _dev(U[1:]), _dev(Z[1:]), opts=opts)
    assert torch.equal(alone[0], L[1])
    # L_1_step inverts L_step when the Jacobian terms are off
    b0 = P.BarVortex(pm, _dev(lh), slots=1, **dict(CONF, k1=0.0, k2=0.0))
    a, back = torch.zeros_like(L[:1]), torch.zeros_like(L[:1])
    b0.L_step(a, _dev(U[:1]), _dev(Z[:1]), opts=opts)
    b0.L_1_step(back, a, _dev(Z[:1]), opts=opts)
    assert ob.norm(back.cpu().numpy()[0] - U[0]) <= 1e-9 * ob.norm(U[0])


@pytest.mark.parametrize("level", [4, 5])
def test_barvortex_boundary_values_vs_oracle(level):
    """BarVortex::calc with non-zero boundary (pole) values bnd_u, bnd_w
    (barvortex_impl.h:171-310) at levels 4 and 5, two members with different values
    against the oracle; a member without boundary values keeps the bits of
    the plain step, and a member with them the bits it has alone."""
    print("barvortex boundary values level=%d" % level, flush=True)
    from oracle.barvortex import BarVortex as OB
    om, pr, U, lh = _states(level, 2)
    pm = P.Mesh.load(oracle_mesh(level)[0])
    assert pm.outer_size == len(om.outer) == 1
    ob = OB(om, lh=lh, **CONF)
    U = U + np.array([[0.3], [-0.2]])                      # pole values 0.3, -0.2
    bu = np.array([[0.31], [-0.19]])                       # new-time psi at the pole
    bw = np.stack([ob.calc1(U[q])[om.outer] * (1.0 + 0.1 * q) for q in range(2)])
    bv = P.BarVortex(pm, _dev(lh), slots=2, **CONF)
    U1 = torch.zeros_like(_dev(U))
    infos = bv.calc(U1, _dev(U), bnd_u=_dev(bu), bnd_w=_dev(bw))
    got = U1.cpu().numpy()
    for q in range(2):
        ref, it, rel = ob.calc(U[q], max_iters=100, bnd_u=bu[q], bnd_w=bw[q])
        assert infos[q]["converged"], infos[q]
        assert infos[q]["iterations"] <= it + 2, (infos[q], it)
        err = ob.norm(got[q] - ref) / ob.norm(ref)
        assert err <= 1e-9, (q, err)
        assert got[q][om.outer[0]] == bu[q][0]
        plain, _, _ = ob.calc(U[q], max_iters=100)
        assert ob.norm(plain - ref) > 1e-3 * ob.norm(ref)   # the values matter
    # member 0 alone with its values; member 1 without values = the plain step
    one = torch.zeros(1, pm.n_points, dtype=torch.float64, device=DEV)
    bv.calc(one, _dev(U[:1]), bnd_u=_dev(bu[:1]), bnd_w=_dev(bw[:1]))
    assert torch.equal(one[0], U1[0])
    bv.load(0, _dev(U[1]))
    bv.load(1, _dev(U[0]), bnd_u=_dev(bu[0]), bnd_w=_dev(bw[0]))
    while True:
        st = bv.iterate()[0]
        if all(x != 1 for x in st):
            break
    mix = torch.zeros(2, pm.n_points, dtype=torch.float64, device=DEV)
    bv.store(0, mix[0])
    bv.store(1, mix[1])
    plain = torch.zeros(1, pm.n_points, dtype=torch.float64, device=DEV)
    bv.calc(plain, _dev(U[1:]))
    assert torch.equal(mix[0], plain[0])
    assert torch.equal(mix[1], U1[0])


def test_barvortex_linear_steps_with_boundary_values_vs_oracle():
    """calc_L, calc_L_1 and calc_LT with non-zero boundary (pole) values
    (barvortex_impl.h:326-535), two members with different values, against
    the oracle (oracle/barvortex.py); zero values give the plain steps' bits;
    a later plain step on the same engine is unaffected; calc_LT ends with
    set_bnd(v1, bnd_u)."""
    print("barvortex linear steps with boundary values level=3", flush=True)
    from oracle.barvortex import BarVortex as OB
    level = 3
    om, pr, Z, lh = _states(level, 2)
    pm = P.Mesh.load(oracle_mesh(level)[0])
    bv = P.BarVortex(pm, _dev(lh), slots=2, **CONF)
    ob = OB(om, lh=lh, **CONF)
    U = _lin_fields(pr, om, (1, 2)) + np.array([[0.3], [-0.2]])
    W, Vv = _lin_fields(pr, om, (3, 4)), _lin_fields(pr, om, (5, 6))
    bu = np.array([[0.31], [-0.19]])
    bw = np.array([[-0.7], [0.45]])
    opts = P.bv_opts(rtol_poisson=1e-12)
    L = torch.zeros_like(_dev(U))
    for i in bv.calc_L(L, _dev(U), _dev(Z), opts=opts, bnd_u=_dev(bu), bnd_w=_dev(bw)):
        assert i["converged"], i
    L1 = _dev(W).clone()
    bv.calc_L_1(L1, _dev(U), _dev(Z), opts=opts, bnd_u=_dev(bu), bnd_w=_dev(bw))
    LT = torch.zeros_like(_dev(Vv))
    bv.calc_LT(LT, _dev(Vv), _dev(Z), opts=opts, bnd_u=_dev(bu), bnd_w=_dev(bw))
    for q in range(2):
        for name, got, ref in (
                ("L", L, ob.calc_L(U[q], Z[q], bu[q], bw[q])),
                ("L_1", L1, ob.calc_L_1(W[q], U[q], Z[q], bu[q], bw[q])),
                ("LT", LT, ob.calc_LT(Vv[q], Z[q], bu[q], bw[q]))):
            g = got.cpu().numpy()[q]
            err = ob.norm(g - ref) / ob.norm(ref)
            assert err <= 1e-9, (name, q, err)
            if name == "L":   # the values matter
                assert ob.norm(ob.calc_L(U[q], Z[q]) - ref) > 1e-4 * ob.norm(ref)
        assert L.cpu().numpy()[q][om.outer[0]] == bu[q][0]
        assert LT.cpu().numpy()[q][om.outer[0]] == bu[q][0]
    # zero boundary values: the plain steps' bits
    zero = torch.zeros(2, pm.outer_size, dtype=torch.float64, device=DEV)
    Lp, Lz = torch.zeros_like(L), torch.zeros_like(L)
    bv.calc_L(Lp, _dev(U), _dev(Z), opts=opts)
    bv.calc_L(Lz, _dev(U), _dev(Z), opts=opts, bnd_u=zero, bnd_w=zero)
    assert torch.equal(Lp, Lz)
    # a plain step after them on the same engine: the bits of a fresh engine
    S1, S2 = torch.zeros_like(L), torch.zeros_like(L)
    bv.calc(S1, _dev(U))
    fresh = P.BarVortex(pm, _dev(lh), slots=2, **CONF)
    fresh.calc(S2, _dev(U))
    assert torch.equal(S1, S2)
