"""MI355X-native (gfx950) sphere-FEM solve path of Phelm (resetius/mke):
P1 assembly of the spherical mass/Laplace/Helmholtz and Jacobian operators
and the GMRES solves behind the BarVortex implicit step, as a C-ABI HIP
library (libphelm_hip.so, include/phelm_gpu.h) with a thin Python host
mirror."""
from .phelm import (BF16, CGS1, CGS2, DIFF1, DIFF1_COS, DIFF2, DIFF2_COS,  # noqa: F401
                    F32, F64, HELMHOLTZ, JAC_MASS, LAPLACE, MASS, MGS, Helmholtz, Matrix,
                    BarVortex, Baroclin, Mesh, Multigrid, PhelmError, Server, Solver, SphereJacobian,
                    SphereChafe, SphereLaplace, SphereNorm, batch_shape, bc_opts, bv_opts, set_batch_shape,
                    helmholtz_coeffs, lib, schafe_f, sphere_mesh_text)
