"""Serving loop (phg_serve_*): requests of members on two mesh levels arrive
in a Poisson stream and are continuously batched; every member's final state
must be bit-identical to stepping that member alone, one request after the
other, with a BarVortex engine of the same slot count (one level runs each
request as two sub-steps of tau / 2)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import mke_amd as P
from tests._common import rh_state

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def test_serve_results_equal_alone_and_in_the_mix():
    print("serve mix", flush=True)
    levels, slots, sub = (3, 4), (2, 2), (1, 2)
    srv = P.Server(levels=levels, slots=slots, substeps=sub)
    assert [srv.slots(lv) for lv in levels] == list(slots)
    meshes = {lv: P.Mesh.sphere(lv) for lv in levels}
    pr = {lv: meshes[lv].arrays()["pr"] for lv in levels}
    plan = [3, 3, 4, 3, 4]                      # member -> level
    init = [rh_state(pr[lv], seed=10 + m) for m, lv in enumerate(plan)]
    ids = [srv.add_member(lv, init[m]) for m, lv in enumerate(plan)]
    rng = np.random.default_rng(7)
    members = rng.integers(0, len(plan), 12)
    arrivals = np.cumsum(rng.exponential(1.0 / 400.0, 12))
    start, finish, infos, st = srv.run(members, arrivals)
    assert st["completed"] == 12 and all(i["converged"] for i in infos)
    assert (finish >= start).all() and (start >= arrivals - 1e-9).all()
    assert st["p50_latency_s"] <= st["p99_latency_s"]
    # the same steps member by member, alone, on standalone engines
    c = srv.cfg.conf
    conf = dict(tau=c.tau, sigma=c.sigma, mu=c.mu, k1=c.k1, k2=c.k2, theta=c.theta)
    for m, lv in enumerate(plan):
        k = sub[levels.index(lv)]              # level 4 runs two steps of tau / 2
        lh = torch.as_tensor(2.0 * np.clip(pr[lv][:, 2], -1, 1), device=DEV)
        bv = P.BarVortex(meshes[lv], lh, slots=2, **dict(conf, tau=conf["tau"] / k))
        u = torch.as_tensor(init[m], device=DEV).reshape(1, -1)
        for _ in range(k * int((members == m).sum())):
            u1 = torch.zeros_like(u)
            bv.calc(u1, u)
            u = u1
        got = srv.member_state(ids[m], meshes[lv].n_points)
        assert np.array_equal(got, u.cpu().numpy()[0]), m
