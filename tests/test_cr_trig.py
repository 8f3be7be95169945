"""The correctly rounded cos/sin shared by the device kernels and the oracle
(mke_amd/csrc/cr_trig.h), pinned against mpmath at 200 bits.

The reference evaluates the cos weight of its element integrals with the C
library's cos (src/util.cpp:194-215).  Correct rounding is the one definition
of that value that every correctly rounding library shares; these tests check
it on random arguments, on the slow path alone, through the rounding test
with a looser bound (which sends many arguments down the slow path), on every
distinct argument the level 4-6 assemblies evaluate, and the constants the
header carries.  The spread between assemblies with this cos and with the C
library's cos is measured per kind and level."""
import ctypes as C
import math
import os
import re

import mpmath
import numpy as np
import pytest
from mpmath import mp, mpf
from mpmath.libmp import to_float

from oracle import mesh_ref
from oracle import oracle as O

L = O.lib()
HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "mke_amd", "csrc", "cr_trig.h")
mp.prec = 200


def rn(v):
    """v rounded to the nearest double, ties to even."""
    return to_float(mpf(v)._mpf_, rnd="n")


def _args(rng):
    xs = list(rng.uniform(-3.5, 3.5, 8000)) + list(rng.uniform(-math.pi / 2, math.pi / 2, 8000))
    xs += list(rng.uniform(-1e-3, 1e-3, 500)) + list(2.0 ** rng.uniform(-60, -20, 500))
    specials = [0.0, -0.0, 5e-324, 2.0 ** -1022, 2.0 ** -30, 2.0 ** -26.5, 1e-8, 0.5,
                math.pi / 4, math.pi / 2, -math.pi / 2, 3 * math.pi / 4, math.pi, 3.5, -3.5,
                math.nextafter(math.pi / 2, 0), math.nextafter(math.pi / 2, 4),
                math.nextafter(math.pi / 4, 0), math.nextafter(math.pi / 4, 4)]
    return [float(x) for x in xs] + specials


def test_constants_pinned():
    text = open(HEADER).read()
    mp.prec = 600
    try:
        def parts(v, n):
            out = []
            for _ in range(n):
                d = rn(v)
                out.append(d)
                v -= mpf(d)
            return out
        want = parts(mp.pi / 2, 4)
        for i, w in enumerate(want, 1):
            m = re.search(r"#define CRT_P%d \(?(-?0x[0-9a-fp.+-]+)\)?" % i, text)
            assert m and float.fromhex(m.group(1)) == w, i
        m = re.search(r"#define CRT_TWO_OVER_PI (0x[0-9a-fp.+-]+)", text)
        assert float.fromhex(m.group(1)) == rn(2 / mp.pi)
        for n in range(2, 28):
            c = (-1) ** (n // 2) / mpf(mpmath.factorial(n))
            hi = rn(c)
            lo = rn(c - mpf(hi))
            m = re.search(r"#define CRT_C%d \(?(-?0x[0-9a-fp.+-]+)\)?, \(?(-?0x[0-9a-fp.+-]+|0\.0)\)?"
                          % n, text)
            if m:
                assert float.fromhex(m.group(1)) == hi and float(
                    float.fromhex(m.group(2)) if m.group(2) != "0.0" else 0.0) == lo, n
            else:
                m = re.search(r"#define CRT_D%d \(?(-?0x[0-9a-fp.+-]+)\)?" % n, text)
                assert m and float.fromhex(m.group(1)) == hi, n
        # 256-bit pi/2 of the slow path: floor(pi/2 2^254), limbs low to high
        N = int(mpmath.floor(mp.pi / 2 * mpf(2) ** 254))
        limbs = re.findall(r"pio2\.w\[(\d)\] = (0x[0-9a-f]+)ull;", text)
        assert len(limbs) == 4
        for i, v in limbs:
            assert int(v, 16) == (N >> (64 * int(i))) & ((1 << 64) - 1), i
    finally:
        mp.prec = 200


def test_random_arguments_correctly_rounded():
    rng = np.random.default_rng(11)
    hi, lo = C.c_double(), C.c_double()
    worst = 0.0
    for x in _args(rng):
        c, s = mpmath.cos(mpf(x)), mpmath.sin(mpf(x))
        assert L.or_cos(x) == rn(c), x.hex()
        got = L.or_sin(x)
        assert got == rn(s), x.hex()
        if x == 0.0:
            assert math.copysign(1, got) == math.copysign(1, x)
        for odd, ref in ((0, c), (1, s)):
            if ref == 0:
                continue
            L.or_cr_fast_dd(x, odd, C.byref(hi), C.byref(lo))
            worst = max(worst, float(abs((mpf(hi.value) + mpf(lo.value) - ref) / ref)))
    # the fast path's double-double error, against the 2^-95 the rounding test
    # assumes
    assert worst < 2.0 ** -100, math.log2(worst)


def test_slow_path_and_rounding_test():
    rng = np.random.default_rng(12)
    xs = _args(rng)[:3000] + [float(v) for v in rng.uniform(-3.5, 3.5, 1000)]
    for x in xs:
        for odd in (0, 1):
            ref = rn(mpmath.sin(mpf(x)) if odd else mpmath.cos(mpf(x)))
            assert L.or_cr_slow(x, odd) == ref, (x.hex(), odd)
    # a bound of 2^-60 makes the rounding test fail for ~1 argument in 60:
    # those take the slow path, the rest the fast one; all must round right
    slow_taken = 0
    for x in (float(v) for v in rng.uniform(-1.6, 1.6, 6000)):
        ref = rn(mpmath.cos(mpf(x)))
        assert L.or_cr_cos_bound(x, 2.0 ** -60) == ref, x.hex()
        hi, lo = C.c_double(), C.c_double()
        L.or_cr_fast_dd(x, 0, C.byref(hi), C.byref(lo))
        e = 2 * 2.0 ** -60 * abs(hi.value)
        slow_taken += (hi.value + (lo.value - e)) != (hi.value + (lo.value + e))
    assert slow_taken > 20, slow_taken


KINDS_WITH_COS = (O.KIND_MASS, O.KIND_LAPLACE, O.KIND_JAC_MASS, O.KIND_DIFF1_COS,
                  O.KIND_DIFF2_COS)


def assembly_cos_args(level):
    """Every distinct argument the inner assemblies of the cos-weighted kinds
    pass to cos at this level."""
    om = O.Mesh(mesh_ref.load_text(mesh_ref.sphere_mesh_text(level)))
    cap = 1 << 22
    buf = np.zeros(cap)
    L.or_record_cos_args(buf.ctypes.data_as(C.POINTER(C.c_double)), cap)
    try:
        for kind in KINDS_WITH_COS:
            om.assemble(kind, 1.0, 0.0)
        n = L.or_recorded_cos_args()
    finally:
        L.or_record_cos_args(None, 0)
    assert n > 0, "recording overflowed"
    args = buf[buf.view(np.uint64) != np.uint64(0xFFFFFFFFFFFFFFFF)]
    assert len(args) == n
    return args


@pytest.mark.parametrize("level", [4, 5, 6])
def test_every_assembly_argument(level):
    args = assembly_cos_args(level)
    assert np.abs(args).max() < 1.6
    libm_differs = 0
    for x in args:
        x = float(x)
        ref = rn(mpmath.cos(mpf(x)))
        assert L.or_cos(x) == ref, x.hex()
        libm_differs += math.cos(x) != ref
    print("level %d: %d distinct cos arguments, all correctly rounded; the C library's "
          "cos differs on %d" % (level, len(args), libm_differs))


def _spread(om, kind, c_mass=1.0, c_lap=0.0):
    rp, ci, a = om.assemble(kind, c_mass, c_lap)
    L.or_use_libm_cos(1)
    try:
        _, _, b = om.assemble(kind, c_mass, c_lap)
    finally:
        L.or_use_libm_cos(0)
    diff = np.abs(a - b)
    nz = np.abs(b) > 0
    ent = float((diff[nz] / np.abs(b[nz])).max())
    # relative to the largest entry of the row
    rowmax = np.maximum.reduceat(np.abs(b), rp[:-1])
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    rowrel = float((diff / rowmax[rows]).max())
    return ent, rowrel, int((diff > 0).sum()), len(a)


# measured (this container's C library): max entrywise relative spread, and
# max spread relative to the row's largest entry; the bounds below hold them
# with a margin of ~1.5-2.  The Helmholtz operator of the solve (M - 1e-2 L)
# stays within 1e-12 entrywise at every level; the mass matrix alone cancels
# ~1e7-fold, so where the library's cos is off by one ulp its entries move by
# up to 6e-10 at level 7 (about 1 % of the entries differ at all).
SPREAD_BOUNDS = {
    "mass": {4: (1e-11, 3e-12), 5: (2e-11, 6e-12), 6: (1.5e-10, 3e-11), 7: (1.2e-9, 2e-10)},
    "laplace": {4: (1e-15, 1e-15), 5: (1e-15, 1e-15), 6: (1e-15, 1e-15), 7: (1e-15, 1e-15)},
    "helmholtz": {4: (1e-12, 2e-13), 5: (1e-12, 2e-13), 6: (1e-12, 2e-13), 7: (1e-12, 2e-13)},
}


@pytest.mark.parametrize("level", [4, 5, 6, pytest.param(7, marks=pytest.mark.slow)])
def test_spread_against_c_library_cos(level):
    """The entrywise spread between this build's assembly (correctly rounded
    cos) and the same assembly with the C library's cos.  Both follow the
    reference's arithmetic; they differ only where the library's cos is not
    correctly rounded.  The spread is printed and bounded by what was
    measured."""
    om = O.Mesh(mesh_ref.load_text(mesh_ref.sphere_mesh_text(level)))
    for name, kind, cm, cl in (("mass", O.KIND_MASS, 1.0, 0.0),
                               ("laplace", O.KIND_LAPLACE, 1.0, 0.0),
                               ("helmholtz", O.KIND_HELMHOLTZ, 1.0, -1e-2)):
        ent, rowrel, ndiff, nnz = _spread(om, kind, cm, cl)
        print("level %d %-9s: %d of %d entries differ; max entrywise relative %.2e, "
              "max relative to the row's largest entry %.2e" % (level, name, ndiff, nnz, ent,
                                                                 rowrel))
        bound_ent, bound_row = SPREAD_BOUNDS[name][level]
        assert ent <= bound_ent and rowrel <= bound_row, (name, ent, rowrel)
