"""dump_operator_matrix (utils/dump_operator_matrix.cpp:41-61, options :19-28
and :65-151): writes the matrix of a mesh operator to a file.

    python -m mke_amd.dump_operator [--double|-d] --mesh FILE --operator NAME
                                    [--store CSR|ELL] -o OUT

The reference writes Solver::dump of its external solver library, whose
format is not in its tree; this tool writes the matrix in the Matrix Market
coordinate format (1-based, "real general"), which every sparse package
reads, with the values as float unless --double is given (the reference's
default is float as well).  --store ELL writes the entries in ELL slot order
(entry k of every row, then k + 1, padding skipped) instead of row order;
the matrix is the same.

Operators: "slaplace" (SphereLaplace::laplace_, the stiffness matrix of the
inner points), and, beside the reference's list, "mass" (SphereLaplace::idt_)
and "helmholtz:TAU" (M - TAU L).  The reference's flat-geometry "laplace" is
not built here.  --mesh takes Mesh::load text (mke_mesh_sphere output) or the
binary image that Mesh.save writes.
"""
import argparse
import sys

import numpy as np

from . import phelm as P


def operator_matrix(mesh, name):
    """The device matrix of operator `name` on `mesh`."""
    if name == "slaplace":
        return P.Matrix.assemble(mesh, P.LAPLACE)
    if name == "mass":
        return P.Matrix.assemble(mesh, P.MASS)
    if name.startswith("helmholtz:"):
        tau = float(name.split(":", 1)[1])
        return P.Matrix.assemble(mesh, P.HELMHOLTZ, 1.0, -tau)
    if name == "laplace":
        raise P.PhelmError("dump_operator: the flat-geometry laplace operator is not built")
    raise P.PhelmError("dump_operator: unknown operator name %s" % name)


def write_matrix_market(f, rp, ci, v, n_cols, double=True, store="CSR"):
    """Matrix Market coordinate text of the CSR (rp, ci, v)."""
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    if store == "ELL":   # slot k of every row, then slot k + 1
        slot = np.arange(len(ci)) - np.repeat(rp[:-1], np.diff(rp))
        order = np.lexsort((rows, slot))
        rows, ci, v = rows[order], ci[order], v[order]
    vals = v if double else v.astype(np.float32)
    f.write("%%MatrixMarket matrix coordinate real general\n")
    f.write("%d %d %d\n" % (n, n_cols, len(ci)))
    fmt = "%d %d %.17g\n" if double else "%d %d %.9g\n"
    for r, c, x in zip(rows.tolist(), ci.tolist(), vals.tolist()):
        f.write(fmt % (r + 1, c + 1, x))


def load_mesh(path):
    with open(path, "rb") as fh:
        head = fh.read(8)
    if head == b"PHGMESH1":
        return P.Mesh.load_file(path)
    with open(path, "r") as fh:
        return P.Mesh.load(fh.read())


def main(argv=None):
    ap = argparse.ArgumentParser(prog="dump_operator", description=__doc__.split("\n\n")[0])
    ap.add_argument("--double", "-d", action="store_true", help="write FP64 values")
    ap.add_argument("--mesh", required=True, help="mesh file")
    ap.add_argument("--operator", required=True, help="operator name")
    ap.add_argument("--store", choices=("CSR", "ELL"), default="CSR", help="store format")
    ap.add_argument("-o", dest="out", required=True, help="output file")
    args = ap.parse_args(argv)
    mesh = load_mesh(args.mesh)
    A = operator_matrix(mesh, args.operator)
    rp, ci, v = A.csr()
    with open(args.out, "w") as f:
        write_matrix_market(f, rp, ci, v, A.n_cols, args.double, args.store)
    return 0


if __name__ == "__main__":
    sys.exit(main())
