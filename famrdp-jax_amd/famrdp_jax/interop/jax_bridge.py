"""Eager JAX front end for the HIP path (JAX is optional and absent in some
images: ``available()`` says whether this module can be used).

The reference differentiates ``step`` with ``jax.grad``
(tests/solver/test_grad.py:30-35).  ``make_steps`` returns a function of a
single block's q0 (a JAX array on the GPU) that advances ``nsteps`` steps on
the HIP kernels and carries a ``jax.custom_vjp`` whose backward pass is the
hand-written adjoint with checkpointing.  Arrays cross between JAX and the
device buffers the kernels use through DLPack, without copies to the host.
Inside ``jax.jit`` use an XLA FFI custom call instead (the handler forwards
XLA's stream to famrdp_step / famrdp_step_vjp).
"""
from __future__ import annotations


def available() -> bool:
    try:
        import jax  # noqa: F401
        import jax.dlpack  # noqa: F401
    except ImportError:
        return False
    return True


def _jax():
    if not available():
        raise ImportError("famrdp_jax.interop.jax_bridge needs jax, which is not installed")
    import jax
    import jax.dlpack
    return jax


def make_steps(state, cfg, dt: float, nsteps: int, checkpoint_every: int = 10):
    """q0 (5, ni, nj, nk) float64 JAX array -> q after ``nsteps`` steps, with
    a custom VJP.  ``state`` supplies the metrics; its q is not used."""
    jax = _jax()
    from torch.utils import dlpack as tdl

    from ..adjoint import Adjoint
    from ..solver.stepper import state_to_device
    if len(state.blocks) != 1:
        raise ValueError("make_steps: single-block configurations only")
    adj = Adjoint(state_to_device(state), cfg)
    to_torch = lambda x: tdl.from_dlpack(x).contiguous().clone()
    to_jax = lambda t: jax.dlpack.from_dlpack(t)

    @jax.custom_vjp
    def steps(q0):
        out, _ = adj.forward_checkpointed([to_torch(q0)], dt, nsteps, checkpoint_every)
        return to_jax(out[0])

    def fwd(q0):
        return steps(q0), q0          # residuals must be JAX arrays: keep q0 only

    def bwd(q0, g):
        _, ckpts = adj.forward_checkpointed([to_torch(q0)], dt, nsteps, checkpoint_every)
        lam = [to_torch(g)]
        adj.reverse(lam, dt, nsteps, ckpts, checkpoint_every, None)
        return (to_jax(lam[0]),)

    steps.defvjp(fwd, bwd)
    return steps
