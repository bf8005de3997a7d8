"""famrdp_jax on MI355X: drop-in for famrdp_jax's RHS + TVD-RK3 path.

Same module layout and function signatures as the reference package
(famrdp_jax/{core,physics,flux,rhs,time,mesh,solver}); every operation runs
as hand-written FP64 HIP kernels for gfx950 behind the C ABI declared in
include/famrdp_capi.h.  There is no CPU fallback.
"""
__version__ = "0.1.0"
