"""MI355X-native (gfx950) sphere-FEM solve path of Phelm (resetius/mke):
P1 assembly of the spherical mass/Laplace/Helmholtz operators and the GMRES
solve behind the BarVortex implicit step, as a C-ABI HIP library
(libphelm_hip.so, include/phelm_gpu.h) with a thin Python host mirror."""
from .phelm import (BF16, CGS1, CGS2, F32, F64, HELMHOLTZ, LAPLACE, MASS, MGS,  # noqa: F401
                    Helmholtz, Matrix, Mesh, PhelmError, Solver, SphereLaplace,
                    helmholtz_coeffs, lib, sphere_mesh_text)
