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


def real_sph_harm_field_lean(xyz, lmax, seed, decay=0.0, device=None):
    """The field of real_sph_harm_field (same coefficients, drawn in the same
    (l, m) order) with O(n) memory: the Legendre recurrence runs per m with
    two rolling rows, so degree 64 on 655,362 points needs a few vectors
    instead of one per (l, m).  Sums in m-major order (differs from
    real_sph_harm_field by rounding).  With `device`, evaluates in torch
    FP64 on that device and returns a numpy array."""
    rng = np.random.default_rng(seed)
    coef = {}
    for l in range(lmax + 1):
        scale = max(l, 1) ** (-decay)
        for m in range(-l, l + 1):
            coef[(l, m)] = rng.standard_normal() * scale
    if device is not None:
        import torch
        xp = torch
        t = torch.as_tensor(np.ascontiguousarray(xyz), dtype=torch.float64, device=device)
        x, y, z = t[:, 0], t[:, 1], t[:, 2]
        ct = torch.clamp(z, -1.0, 1.0)
        st = torch.sqrt(torch.clamp(1.0 - ct * ct, min=0.0))
        phi = torch.atan2(y, x)
    else:
        xp = np
        x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
        ct = np.clip(z, -1.0, 1.0)
        st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
        phi = np.arctan2(y, x)
    f = xp.zeros_like(ct)
    pmm = xp.ones_like(ct)
    for m in range(lmax + 1):
        if m > 0:
            pmm = -pmm * (2 * m - 1) * st
        cosm, sinm = xp.cos(m * phi), xp.sin(m * phi)
        p2, p1 = None, pmm
        for l in range(m, lmax + 1):
            if l == m + 1:
                p2, p1 = p1, ct * (2 * m + 1) * pmm
            elif l > m + 1:
                p2, p1 = p1, ((2 * l - 1) * ct * p1 - (l + m - 1) * p2) / (l - m)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))
            if m == 0:
                f = f + coef[(l, 0)] * norm * p1
            else:
                f = f + (coef[(l, m)] * math.sqrt(2) * norm) * p1 * cosm
                f = f + (coef[(l, -m)] * math.sqrt(2) * norm) * p1 * sinm
    return f.cpu().numpy() if device is not None else f


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
