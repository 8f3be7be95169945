"""CPU restatement of SphereJacobian (src/sjacobian.cpp:136-468).

TEST INFRASTRUCTURE ONLY: the checker for the GPU Jacobian; the product never
imports it.

The nine matrices are the oracle's assemblies of the reference callbacks
(kinds JAC_MASS, DIFF1, DIFF2, DIFF1_COS, DIFF2_COS, inner and boundary).
The reference solves with idt_ through its external solver library (GMRES,
parity unpinned); the checker solves the same systems with a sparse LU
factorisation, i.e. to rounding, which is the answer any converged solve
approximates.  Every vector operation follows the reference's order:
rp = D_rp u_bnd + D u_in (vec_sum(rp, rp, tmp)), pt = solve(rp) * solve(..)
(vec_mult), Ans = pt1 - pt2 (vec_diff), calc2 = ((c1 + c2) + c3) + c4, then
times 0.25 (sjacobian.cpp:348-362).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import oracle as O

_KINDS = {
    "diff1": O.KIND_DIFF1,
    "diff2": O.KIND_DIFF2,
    "diff1_cos": O.KIND_DIFF1_COS,
    "diff2_cos": O.KIND_DIFF2_COS,
}


def _csr(rp, ci, v, n_cols):
    return sp.csr_matrix((v, ci, rp), shape=(len(rp) - 1, n_cols))


class SphereJacobian:
    def __init__(self, mesh):
        """mesh: oracle.Mesh."""
        self.m = mesh
        n, no = mesh.n_inner, len(mesh.outer)
        rp, ci, v = mesh.assemble(O.KIND_JAC_MASS)
        self.idt = _csr(rp, ci, v, n)
        self._lu = spla.splu(self.idt.tocsc())
        self.D = {}
        self.Db = {}
        for name, kind in _KINDS.items():
            self.D[name] = mesh.assemble(kind)
            brp, bci, bv = mesh.assemble_boundary(kind)
            self.Db[name] = (brp, bci, bv)

    def solve(self, rhs):
        return self._lu.solve(rhs)

    def _rp(self, name, u_in, u_bnd):
        """diff*_.mult_vector(tmp, u_in); diff*_rp_.mult_vector(rp, u_bnd);
        vec_sum(rp, rp, tmp)."""
        tmp = O.spmv(*self.D[name], u_in)
        brp, bci, bv = self.Db[name]
        rp = O.spmv(brp, bci, bv, np.ascontiguousarray(u_bnd)) if len(bci) else np.zeros(len(u_in))
        return rp + tmp

    def projections(self, u):
        """idt^-1 (D u_in + D_rp u_bnd) for the four operators: the 4 distinct
        mass solves per input vector of calc2."""
        u = np.asarray(u, dtype=np.float64)
        u_in, u_bnd = u[self.m.inner], u[self.m.outer]
        return {name: self.solve(self._rp(name, u_in, u_bnd)) for name in _KINDS}

    def calc2(self, u, v):
        pu, pv = self.projections(u), self.projections(v)
        c1 = pu["diff2_cos"] * pv["diff1"] - pu["diff1_cos"] * pv["diff2"]   # :136-185
        c2 = pu["diff2"] * pv["diff1_cos"] - pu["diff1"] * pv["diff2_cos"]   # :187-235
        c3 = pu["diff2_cos"] * pv["diff1"] - pu["diff1"] * pv["diff2_cos"]   # :237-285
        c4 = pu["diff2"] * pv["diff1_cos"] - pu["diff1_cos"] * pv["diff2"]   # :287-345
        ans = c1
        ans = ans + c2
        ans = ans + c3
        ans = ans + c4
        return ans * 0.25

    def calc2t(self, u, v):
        return self.calc2(u, v) * -1.0
