"""Pins the oracle's BarVortex::calc restatement (oracle/barvortex.py,
examples/barvortex_impl.h:171-310) by what the mathematics fixes, on the CPU:

* Without the Jacobian terms (k1 = k2 = 0) the step is the theta scheme of
  d(lap psi)/dt = -sigma lap psi + mu lap^2 psi.  For psi = (x^2 - y^2) z, a
  degree-3 spherical harmonic (lap psi = -12 psi, zero at the poles), the
  step multiplies psi by (1/tau - (1-theta) s) / (1/tau + theta s) with
  s = sigma + 12 mu.  The discrete factor approaches it at the P1 rate
  O(h^2); theta = 0.75 tells theta from 1 - theta, and a sign error in the
  sigma or mu terms moves the factor by O(1).
* SphereNorm's matrix integrates 1 over the sphere: norm(1)^2 -> 4 pi.
"""
import math

import numpy as np
import pytest

pytest.importorskip("scipy")

from oracle import mesh_ref, oracle as O
from oracle.barvortex import BarVortex


def _setup(level, **kw):
    a = mesh_ref.load_text(mesh_ref.sphere_mesh_text(level))
    M = O.Mesh(a)
    pr = np.asarray(a["pr"]).reshape(-1, 3)
    bv = BarVortex(M, lh=np.zeros(len(pr)), **kw)
    return bv, pr


def _factor(s, tau, theta):
    return (1.0 / tau - (1.0 - theta) * s) / (1.0 / tau + theta * s)


def test_linear_step_is_the_theta_scheme():
    tau, sigma, mu, theta = 1.0, 0.5, 0.05, 0.75
    want = _factor(sigma + 12.0 * mu, tau, theta)
    errs = []
    for level in (3, 4, 5):
        bv, pr = _setup(level, tau=tau, sigma=sigma, mu=mu, k1=0.0, k2=0.0, theta=theta)
        psi = (pr[:, 0] ** 2 - pr[:, 1] ** 2) * pr[:, 2]
        u1, it, rel = bv.calc(psi, max_iters=50, tol=1e-14)
        assert it <= 3 and rel < 1e-14          # no Jacobian: F is fixed after one pass
        got = float(np.dot(u1, psi) / np.dot(psi, psi))
        resid = np.linalg.norm(u1 - got * psi) / np.linalg.norm(u1)
        errs.append(abs(got - want))
        assert resid < 0.05, (level, resid)
    assert errs[-1] < 2e-3, errs
    assert errs[0] / errs[1] > 3.0 and errs[1] / errs[2] > 3.0, errs   # O(h^2)
    # the plausible mistakes land far from the computed factor
    for wrong in (_factor(sigma + 12.0 * mu, tau, 1.0 - theta),
                  _factor(-sigma + 12.0 * mu, tau, theta),
                  _factor(sigma - 12.0 * mu, tau, theta)):
        assert abs(wrong - want) > 50 * errs[-1]


def test_sphere_norm_of_one_is_the_sphere_area():
    errs = []
    for level in (2, 3, 4):
        bv, pr = _setup(level, tau=1.0, sigma=0.0, mu=0.0, k1=0.0, k2=0.0, theta=0.5)
        errs.append(abs(bv.norm(np.ones(len(pr))) ** 2 - 4.0 * math.pi))
    assert errs[-1] < 1e-2 and errs[0] / errs[1] > 3.0 and errs[1] / errs[2] > 3.0, errs
