"""GPU SphereChafe (mke_amd.SphereChafe over the C ABI: GPU-assembled A_ and
bnd_, SphereLaplace::calc2, GMRES) against the oracle restatement
(oracle/schafe.py, sparse LU solves), on the reference driver's analytic
case (examples/test_schafe.cpp:25-34, 61-125: tau = 0.01, mu = 1,
sigma = 70, boundary values of ans at the new time).

Bars: A_ and bnd_ bit-identical to the oracle's; one step within 1e-11 of
the oracle step (SphereNorm distance relative to the state; the A_ solve is
driven to 1e-14, calc2's mass solve to 1e-13); ten steps within 1e-10; the
error against the analytic solution equals the oracle's to 1e-8 of it."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytest.importorskip("scipy")

import mke_amd as P
import oracle.schafe as S
from oracle import mesh_ref
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TAU, MU, SIGMA = 0.01, 1.0, 70.0


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


def _setup(level):
    text = mesh_ref.sphere_mesh_text(level)
    a = mesh_ref.load_text(text)
    om = O.Mesh(a)
    xy = np.asarray(a["coords"]).reshape(len(om.p2io), -1, 2)[:, 0, :]
    return P.Mesh.load(text), om, xy


def test_schafe_matrices_bitwise():
    print("schafe matrices level=3", flush=True)
    pm, om, xy = _setup(3)
    sc = P.SphereChafe(pm, TAU, SIGMA, MU)
    osc = S.SphereChafe(om, xy, TAU, SIGMA, MU)
    for got, want in ((sc.A_.csr(), osc.A), (sc.bnd_.csr(), osc.bndA)):
        for g, w in zip(got, want):
            assert np.array_equal(g, w)


@pytest.mark.parametrize("level", [3, 4])
def test_schafe_steps_vs_oracle(level):
    print("schafe steps level=%d" % level, flush=True)
    pm, om, xy = _setup(level)
    sc = P.SphereChafe(pm, TAU, SIGMA, MU)
    osc = S.SphereChafe(om, xy, TAU, SIGMA, MU)
    U = S.schafe_ans(xy[:, 0], xy[:, 1], 0.0)
    dU = _dev(U)
    for i in range(10):
        B = S.schafe_ans(xy[om.outer, 0], xy[om.outer, 1], TAU * (i + 1))
        info = sc.solve(dU, dU, _dev(B), TAU * i, rtol=1e-14)
        assert info["converged"] or info["rel_residual"] < 1e-13, info
        U = osc.solve(U, B, TAU * i)
        got = dU.cpu().numpy()
        rel = S.norm_dist(om, got, U) / S.norm_dist(om, U, 0 * U)
        assert rel <= (1e-11 if i == 0 else 1e-10), (i, rel)
        assert np.array_equal(got[om.outer], B)
    ans = S.schafe_ans(xy[:, 0], xy[:, 1], TAU * 10)
    e_gpu, e_or = S.norm_dist(om, got, ans), S.norm_dist(om, U, ans)
    assert abs(e_gpu - e_or) <= 1e-8 * e_or, (e_gpu, e_or)
