"""Writes the golden fixtures of this directory (run from the repository root:
``python tests/golden/make_golden.py``).

The reference ships no numeric fixtures for this path (its tests print
errors against analytic fields), so these are the build's own, made with the
oracle and checked in so that later changes of the oracle or of the GPU path
are compared against fixed data:

* sphere_L2.txt          mesh text of ``mke_mesh_sphere --type full --coord
                         local --iter 2`` (oracle restatement)
* matrices_L2.npz        CSR structure and values of every element kind
                         (inner x inner and inner x outer) at level 2
* helmholtz_L3.npz       A = M - 1e-2 L at level 3, b = M f (f: real Y_lm,
                         l <= 8, seed 20251202) and x = A^-1 b by a sparse LU
                         (scipy), a solve independent of GMRES
* jacobian_L3.npz        u, v (real Y_lm fields, l <= 6, seeds 7 and 8) and
                         SphereJacobian::calc2(u, v) of the oracle restatement
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import mesh_ref, oracle as O  # noqa: E402
from tests._common import real_sph_harm_field  # noqa: E402

KINDS = {"mass": O.KIND_MASS, "laplace": O.KIND_LAPLACE, "jac_mass": O.KIND_JAC_MASS,
         "diff1": O.KIND_DIFF1, "diff2": O.KIND_DIFF2, "diff1_cos": O.KIND_DIFF1_COS,
         "diff2_cos": O.KIND_DIFF2_COS}
CM, CL = 1.0, -1e-2   # M - 1e-2 L


def main():
    text2 = mesh_ref.sphere_mesh_text(2)
    with open(os.path.join(HERE, "sphere_L2.txt"), "w") as f:
        f.write(text2)
    M2 = O.Mesh(mesh_ref.load_text(text2))
    out = {}
    rp, ci = M2.csr_structure()
    out["row_ptr"], out["col_idx"] = rp, ci
    brp, bci = M2.boundary_structure()
    out["bnd_row_ptr"], out["bnd_col_idx"] = brp, bci
    for name, kind in list(KINDS.items()) + [("helmholtz", O.KIND_HELMHOLTZ)]:
        cm, cl = (CM, CL) if kind == O.KIND_HELMHOLTZ else (1.0, 0.0)
        out[name] = M2.assemble(kind, cm, cl)[2]
        out["bnd_" + name] = M2.assemble_boundary(kind, cm, cl)[2]
    np.savez_compressed(os.path.join(HERE, "matrices_L2.npz"), **out)

    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    text3 = mesh_ref.sphere_mesh_text(3)
    a3 = mesh_ref.load_text(text3)
    M3 = O.Mesh(a3)
    pr = np.asarray(a3["pr"]).reshape(-1, 3)
    rp, ci, va = M3.assemble(O.KIND_HELMHOLTZ, CM, CL)
    _, _, vm = M3.assemble(O.KIND_MASS)
    f = real_sph_harm_field(pr[M3.inner], 8, 20251202)
    b = O.spmv(rp, ci, vm, f)
    n = M3.n_inner
    x = spla.spsolve(sp.csr_matrix((va, ci, rp), shape=(n, n)).tocsc(), b)
    np.savez_compressed(os.path.join(HERE, "helmholtz_L3.npz"), b=b, x=x, f=f)

    from oracle.sjacobian import SphereJacobian
    u = real_sph_harm_field(pr, 6, 7)
    v = real_sph_harm_field(pr, 6, 8)
    ans = SphereJacobian(M3).calc2(u, v)
    np.savez_compressed(os.path.join(HERE, "jacobian_L3.npz"), u=u, v=v, ans=ans)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
