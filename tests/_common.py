"""Shared helpers of the test suite: seeded spherical-harmonic fields and
cached meshes / oracle matrices."""
import functools
import math

import numpy as np


def real_sph_harm_field(xyz, lmax, seed, decay=0.0):
    """f = sum_{l<=lmax, m} c_lm Y_lm(x, y, z) with real orthonormal
    spherical harmonics evaluated at unit vectors xyz (n x 3); c_lm ~ N(0,1)
    times max(l,1)^-decay, drawn from numpy's default_rng(seed) in (l, m)
    order, m = -l..l."""
    rng = np.random.default_rng(seed)
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    ct = np.clip(z, -1.0, 1.0)
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    phi = np.arctan2(y, x)
    f = np.zeros(len(z))
    # associated Legendre P_l^m(ct) by the standard stable recurrences
    pmm = np.ones_like(ct)
    P = {}
    for m in range(lmax + 1):
        if m > 0:
            pmm = -pmm * (2 * m - 1) * st
        P[(m, m)] = pmm.copy()
        if m + 1 <= lmax:
            P[(m + 1, m)] = ct * (2 * m + 1) * pmm
        for l in range(m + 2, lmax + 1):
            P[(l, m)] = ((2 * l - 1) * ct * P[(l - 1, m)] - (l + m - 1) * P[(l - 2, m)]) / (l - m)
    for l in range(lmax + 1):
        scale = max(l, 1) ** (-decay)
        for m in range(-l, l + 1):
            c = rng.standard_normal() * scale
            am = abs(m)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
            if m == 0:
                Y = norm * P[(l, 0)]
            elif m > 0:
                Y = math.sqrt(2) * norm * P[(l, am)] * np.cos(am * phi)
            else:
                Y = math.sqrt(2) * norm * P[(l, am)] * np.sin(am * phi)
            f += c * Y
    return f


@functools.lru_cache(maxsize=8)
def oracle_mesh(level):
    from oracle import mesh_ref, oracle as O
    text = mesh_ref.sphere_mesh_text(level)
    return text, O.Mesh(mesh_ref.load_text(text))


@functools.lru_cache(maxsize=16)
def oracle_matrix(level, kind, c_mass=1.0, c_lap=0.0):
    _, M = oracle_mesh(level)
    return M.assemble(kind, c_mass, c_lap)


def rh_state(pr, seed=None, noise=1e-3, omega=7.292e-5):
    """Stream function of the Rossby-Haurwitz wave 4 (Williamson et al. 1992,
    case 6: omega = K = 7.848e-6 s^-1) in units of Earth's angular velocity,
    plus, with a seed, a smooth N(0, noise * max|psi|)-scaled perturbation;
    shifted so that the south pole (the Dirichlet point) is 0."""
    z = np.clip(pr[:, 2], -1.0, 1.0)
    phi = np.arcsin(z)
    lam = np.arctan2(pr[:, 1], pr[:, 0])
    w = 7.848e-6 / omega
    psi = -w * np.sin(phi) + w * np.cos(phi) ** 4 * np.sin(phi) * np.cos(4 * lam)
    if seed is not None:
        f = real_sph_harm_field(pr, 8, seed)
        psi = psi + noise * np.abs(psi).max() * f / np.abs(f).max()
    pole = int(np.argmin(z))
    return psi - psi[pole]
