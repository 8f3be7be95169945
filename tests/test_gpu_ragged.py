"""Matrices a caller builds (Solver<T>(n) + add, phg_matrix_from_csr), not
from the sphere: ragged rows (empty rows, rows shorter than the widest,
rows wider than the unrolled ELL paths take), rectangular shapes and
matrices with no entries, through the C ABI on the GPU.

Bars: the FP64 SpMV equals the oracle's CSR loop (oracle.spmv) bit for bit
on every path (the unrolled row loop with slots past a row's width
clamped, the loop over wide rows, the streamed kernel for 2^18 rows and
more, at ELL widths 6 and 8); GMRES reaches a true FP64 relative residual
of 1e-10 by the oracle's SpMV, within 1e-8 of the oracle's GMRES driven to
1e-13, and the persistent MGS cycle gives the kernel-per-pass form's bits on
these rows too."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import mke_amd as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _ragged(n_rows, n_cols, maxw, seed, empty_frac=0.1, dominant=False):
    """CSR with row lengths in 0..maxw (about empty_frac of the rows empty;
    repeated column draws merge, so some rows come out shorter), distinct
    ascending columns, N(0,1) values; with dominant (square), every row also
    holds its diagonal, set above the sum of the row's other |values|, and at
    most maxw entries in all."""
    rng = np.random.default_rng(seed)
    w = rng.integers(1, maxw + 1 - (1 if dominant else 0), n_rows)
    w[rng.random(n_rows) < empty_frac] = 0
    rows = np.repeat(np.arange(n_rows, dtype=np.int64), w)
    cols = rng.integers(0, n_cols, rows.size)
    if dominant:
        rows = np.concatenate([rows, np.arange(n_rows)])
        cols = np.concatenate([cols, np.arange(n_rows)])
    key = np.unique(rows * n_cols + cols)
    rows, cols = key // n_cols, key % n_cols
    v = rng.standard_normal(rows.size)
    if dominant:
        diag = rows == cols
        v[diag] = 0.0
        v[diag] = np.bincount(rows, weights=np.abs(v), minlength=n_rows)[rows[diag]] + 1.0
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n_rows))])
    return rp.astype(np.int32), cols.astype(np.int32), v


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


@pytest.mark.parametrize("n,maxw", [(3000, 12), (3000, 8), (3000, 5), (300000, 8), (300000, 6)])
def test_spmv_ragged_bitwise(n, maxw):
    print("ragged spmv n=%d maxw=%d" % (n, maxw), flush=True)
    rp, ci, v = _ragged(n, n, maxw, seed=n + maxw)
    A = P.Matrix.from_csr(n, n, rp, ci, v)
    x = np.random.default_rng(1).standard_normal(n)
    y = torch.full((n,), np.nan, dtype=torch.float64, device=DEV)
    A.mult_vector(y, _dev(x))
    want = O.spmv(rp, ci, v, x)
    assert np.array_equal(y.cpu().numpy(), want)
    A.close()


def test_spmv_rectangular_and_empty():
    print("rectangular and empty spmv", flush=True)
    rp, ci, v = _ragged(1000, 50, 9, seed=5)
    A = P.Matrix.from_csr(1000, 50, rp, ci, v)
    x = np.random.default_rng(2).standard_normal(50)
    y = torch.full((1000,), np.nan, dtype=torch.float64, device=DEV)
    A.mult_vector(y, _dev(x))
    assert np.array_equal(y.cpu().numpy(), O.spmv(rp, ci, v, x))
    A.close()
    # no entries at all: y = 0
    E = P.Matrix.from_csr(7, 7, np.zeros(8, np.int32), np.zeros(0, np.int32), np.zeros(0))
    y = torch.full((7,), np.nan, dtype=torch.float64, device=DEV)
    E.mult_vector(y, _dev(np.ones(7)))
    assert np.array_equal(y.cpu().numpy(), np.zeros(7))
    E.close()


@pytest.mark.parametrize("maxw", [12, 8])
def test_gmres_ragged_rows(maxw):
    print("ragged gmres maxw=%d" % maxw, flush=True)
    n = 20000
    rp, ci, v = _ragged(n, n, maxw, seed=40 + maxw, empty_frac=0.0, dominant=True)
    A = P.Matrix.from_csr(n, n, rp, ci, v)
    b = np.random.default_rng(3).standard_normal(n)
    xt, oinfo = O.gmres(rp, ci, v, b, restart=30, rtol=1e-13)
    assert oinfo["converged"]
    xs = {}
    for name, kw in (("mgs per pass", dict(ortho=P.MGS, persistent=False)),
                     ("mgs persistent", dict(ortho=P.MGS, persistent=True)),
                     ("cgs2", dict(ortho=P.CGS2))):
        x = torch.zeros(n, dtype=torch.float64, device=DEV)
        info = A.solve(x, _dev(b), restart=30, rtol=1e-10, **kw)
        xh = x.cpu().numpy()
        rel = np.linalg.norm(b - O.spmv(rp, ci, v, xh)) / np.linalg.norm(b)
        fwd = np.linalg.norm(xh - xt) / np.linalg.norm(xt)
        print("  %s: %d steps, true rel %.2e, forward error %.2e" %
              (name, info["iterations"], rel, fwd), flush=True)
        assert info["converged"] and rel <= 1e-10, (name, info, rel)
        assert fwd <= 1e-8, (name, fwd)
        xs[name] = x
    assert torch.equal(xs["mgs persistent"], xs["mgs per pass"])
    A.close()
