"""MI355X (gfx950) residual / SSP-RK3 / discrete-adjoint path of the
Optimisation2Forme compressible Navier-Stokes solver, behind a C ABI
(include/o2f.h, libo2f.so).  See capi.Solver."""
from .capi import Solver, make_params, read_msh2, hllc_flux_batch, O2FError  # noqa: F401
