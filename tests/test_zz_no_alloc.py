"""phg_dot_noalloc, the allocation-free dot product (device result, block
partials in the library's file-scope device array), against phg_dot through
the C ABI: two calls in a row, the second with a smaller input, each within
1e-13 relative of phg_dot (the two add the products in different orders),
bits repeatable call to call; the largest n it serves; and the same results
when the array was left holding a larger call's partials."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import mke_amd as P
from mke_amd.phelm import check

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dot(n, x, y):
    r = C.c_double()
    check(P.lib().phg_dot(n, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
                            C.byref(r), None), "phg_dot")
    return r.value


def _dot_noalloc(n, x, y, out):
    check(P.lib().phg_dot_noalloc(n, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
                                    C.c_void_p(out.data_ptr()), None), "phg_dot_noalloc")
    torch.cuda.synchronize()
    return float(out.cpu().numpy()[0])


def test_dot_noalloc_matches_dot():
    print("phg_dot_noalloc vs phg_dot", flush=True)
    rng = np.random.default_rng(7)
    out = torch.zeros(1, dtype=torch.float64, device=DEV)
    results = []
    for n in (163841, 10241, 163841, 1, 0, 1023, 1025, 4194304):
        x = torch.as_tensor(rng.standard_normal(max(n, 1)), device=DEV)
        y = torch.as_tensor(rng.standard_normal(max(n, 1)), device=DEV)
        ref = _dot(n, x, y)
        got = _dot_noalloc(n, x, y, out)
        again = _dot_noalloc(n, x, y, out)
        exact = float(np.dot(x.cpu().numpy()[:n], y.cpu().numpy()[:n])) if n else 0.0
        scale = float(np.dot(np.abs(x.cpu().numpy()[:n]), np.abs(y.cpu().numpy()[:n]))) if n else 1.0
        assert got == again, n
        assert abs(got - ref) <= 1e-13 * scale and abs(got - exact) <= 1e-13 * scale, (n, got, ref)
        results.append((n, got))
    assert results[1][0] == 10241


def test_dot_noalloc_rejects_sizes_beyond_its_scratch():
    L = P.lib()
    out = torch.zeros(1, dtype=torch.float64, device=DEV)
    x = torch.zeros(8, dtype=torch.float64, device=DEV)
    rc = L.phg_dot_noalloc(4194305, C.c_void_p(x.data_ptr()), C.c_void_p(x.data_ptr()),
                           C.c_void_p(out.data_ptr()), None)
    assert rc == -1 and b"4194304" in L.phg_last_error()
