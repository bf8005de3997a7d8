"""Device-resident stepping: the state lives in HBM and advances in place,
one launch sequence (or one hipGraph replay) per step.  Used by the bench
and by callers that run many steps; ``step`` above is the functional API."""
from __future__ import annotations

from .. import _lib as L
from .._plan import Plan, bcs_from_config, cuts_from_topology, plan_for
from ..core.types import Block, Config, State


def state_to_device(state: State) -> State:
    """Copy every array of a State onto the current GPU (float64, contiguous)."""
    from ..core.types import Metrics
    dev = L.device_f64
    return State(blocks=tuple(Block(q=dev(b.q), xyz=dev(b.xyz)) for b in state.blocks),
                 metrics=tuple(Metrics(jac=dev(m.jac), kxyz=dev(m.kxyz), vol=dev(m.vol)) for m in state.metrics),
                 t=state.t, step=state.step)


class Stepper:
    """``mu0`` / ``dt_scale`` (one value per block) give each block its own
    Sutherland mu0 and time step dt * dt_scale (ensembles of independent
    cases in one batched launch); such a stepper owns an uncached plan."""

    def __init__(self, state: State, cfg: Config, *, mu0=None, dt_scale=None):
        self.cfg = cfg
        self.qs = [L.device_f64(b.q).clone() for b in state.blocks]
        self.kxyz = [L.device_f64(m.kxyz).contiguous() for m in state.metrics]
        self.vol = [L.device_f64(m.vol).contiguous() for m in state.metrics]
        self.xyz = [b.xyz for b in state.blocks]
        self.metrics = state.metrics
        self.t = state.t
        self.nstep = state.step
        dims = [tuple(q.shape[1:]) for q in self.qs]
        if mu0 is None and dt_scale is None:
            self.plan = plan_for(dims, cfg.topology, cfg.bc_params, cfg.gas, cfg.scheme.limiter, cfg.ghost)
        else:
            self.plan = Plan(dims, cuts_from_topology(cfg.topology), bcs_from_config(cfg.topology, cfg.bc_params),
                             cfg.gas, cfg.scheme.limiter, cfg.ghost)
            self.plan.set_block_params(mu0, dt_scale)

    def advance(self, dt: float, nsteps: int = 1, *, batched: bool = False, graph: bool = False,
                direct: bool = False):
        if graph:
            self.plan.step(self.qs, self.kxyz, self.vol, dt, batched=batched, graph_steps=nsteps, direct=direct)
        else:
            for _ in range(nsteps):
                self.plan.step(self.qs, self.kxyz, self.vol, dt, batched=batched, direct=direct)
        self.t += dt * nsteps
        self.nstep += nsteps
        if self.cfg.nan_check and nsteps > 0:
            # once per call, after its last step (one reduction over the state
            # and one read of its count)
            self.plan.raise_if_bad(self.qs)

    def autotune(self, dt: float, *, batched: bool = False, reps: int = 3):
        """Pick the fused kernel's tile shape for these blocks (state untouched)."""
        return self.plan.autotune(self.qs, self.kxyz, self.vol, dt, batched=batched, reps=reps)

    def state(self) -> State:
        blocks = tuple(Block(q=q, xyz=x) for q, x in zip(self.qs, self.xyz))
        return State(blocks=blocks, metrics=self.metrics, t=self.t, step=self.nstep)
