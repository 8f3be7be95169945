"""The committed fixtures of tests/golden (made by tests/golden/make_golden.py)
against the oracle (CPU) and against the GPU path (gpu marker)."""
import os

import numpy as np
import pytest

from oracle import mesh_ref, oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KINDS = {"mass": O.KIND_MASS, "laplace": O.KIND_LAPLACE, "jac_mass": O.KIND_JAC_MASS,
         "diff1": O.KIND_DIFF1, "diff2": O.KIND_DIFF2, "diff1_cos": O.KIND_DIFF1_COS,
         "diff2_cos": O.KIND_DIFF2_COS, "helmholtz": O.KIND_HELMHOLTZ}
CM, CL = 1.0, -1e-2


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def _mesh_text():
    with open(os.path.join(G, "sphere_L2.txt")) as f:
        return f.read()


def test_oracle_mesh_matches_fixture():
    assert mesh_ref.sphere_mesh_text(2) == _mesh_text()


def test_oracle_matrices_match_fixture():
    g = _load("matrices_L2.npz")
    M = O.Mesh(mesh_ref.load_text(_mesh_text()))
    rp, ci = M.csr_structure()
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(ci, g["col_idx"])
    for name, kind in KINDS.items():
        cm, cl = (CM, CL) if name == "helmholtz" else (1.0, 0.0)
        assert np.array_equal(M.assemble(kind, cm, cl)[2], g[name]), name
        assert np.array_equal(M.assemble_boundary(kind, cm, cl)[2], g["bnd_" + name]), name


def test_oracle_gmres_matches_direct_solve_fixture():
    g = _load("helmholtz_L3.npz")
    M = O.Mesh(mesh_ref.load_text(mesh_ref.sphere_mesh_text(3)))
    rp, ci, v = M.assemble(O.KIND_HELMHOLTZ, CM, CL)
    x, info = O.gmres(rp, ci, v, g["b"], rtol=1e-12)
    assert info["converged"]
    assert np.linalg.norm(x - g["x"]) / np.linalg.norm(g["x"]) < 1e-10


def test_oracle_jacobian_matches_fixture():
    from oracle.sjacobian import SphereJacobian
    g = _load("jacobian_L3.npz")
    M = O.Mesh(mesh_ref.load_text(mesh_ref.sphere_mesh_text(3)))
    ans = SphereJacobian(M).calc2(g["u"], g["v"])
    assert np.abs(ans - g["ans"]).max() <= 1e-13 * np.abs(g["ans"]).max()


@pytest.mark.gpu
def test_gpu_matches_fixtures():
    import torch
    import mke_amd as P
    dev = "cuda:0"
    g = _load("matrices_L2.npz")
    pm = P.Mesh.load(_mesh_text())
    for name, kind in KINDS.items():
        cm, cl = (CM, CL) if name == "helmholtz" else (1.0, 0.0)
        rp, ci, v = P.Matrix.assemble(pm, kind, cm, cl).csr()
        assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(ci, g["col_idx"])
        assert np.array_equal(v, g[name]), name
        assert np.array_equal(P.Matrix.assemble_boundary(pm, kind, cm, cl).csr()[2],
                              g["bnd_" + name]), name
    h = _load("helmholtz_L3.npz")
    pm3 = P.Mesh.sphere(3)
    A = P.Matrix.assemble(pm3, P.HELMHOLTZ, CM, CL)
    b = torch.as_tensor(h["b"], device=dev)
    x = torch.zeros_like(b)
    info = A.solve(x, b, rtol=1e-12)
    assert info["converged"]
    assert np.linalg.norm(x.cpu().numpy() - h["x"]) / np.linalg.norm(h["x"]) < 1e-10
    j = _load("jacobian_L3.npz")
    J = P.SphereJacobian(pm3)
    ans = torch.zeros(len(j["ans"]), dtype=torch.float64, device=dev)
    J.calc2(ans, torch.as_tensor(j["u"], device=dev), torch.as_tensor(j["v"], device=dev))
    assert np.abs(ans.cpu().numpy() - j["ans"]).max() <= 1e-11 * np.abs(j["ans"]).max()
