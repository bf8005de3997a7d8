"""Gradients through the TVD-RK3 step with hand-written adjoint kernels.

The reference differentiates its eager step with jax.grad
(tests/solver/test_grad.py:10-38), storing every intermediate.  Here the
reverse pass is famrdp_step_vjp (HIP), driven with checkpointing: the forward
keeps the state every ``checkpoint_every`` steps and, for the last segment
(where the reverse pass starts), every step's start and stage states; the
reverse pass recomputes the other segments one at a time (keeping their
step-start and stage states), and each step's reverse rebuilds its halo/BC
state before reversing stages 3, 2, 1, the BCs and the halo copy.

Besides d J / d q0, the farfield states q_inf receive their cotangents
(summed over steps) per (block, face); ``mach_derivative`` turns them into
d J / d M_inf for q_inf(M) = (1, M, 0, 0, p/(gamma-1) + M^2/2).
"""
from __future__ import annotations

import time

from . import _lib as L
from ._plan import Plan, bcs_from_config, cuts_from_topology
from .core.types import State

OBJ_KINETIC_ENERGY = 0
OBJ_WALL_PRESSURE = 1


def torch_empty_like(t):
    return L._torch().empty_like(t)


class Adjoint:
    """Forward and reverse sweeps for one configuration (all blocks resident)."""

    def __init__(self, state: State, cfg, *, batched=False, direct=False, autotune=False):
        torch = L._torch()
        L.require_gpu()
        self.cfg = cfg
        self.batched, self.direct = batched, direct
        self.q0 = [L.device_f64(b.q).clone() for b in state.blocks]
        self.kxyz = [L.device_f64(m.kxyz).contiguous() for m in state.metrics]
        self.vol = [L.device_f64(m.vol).contiguous() for m in state.metrics]
        dims = [tuple(q.shape[1:]) for q in self.q0]
        self.plan = Plan(dims, cuts_from_topology(cfg.topology), bcs_from_config(cfg.topology, cfg.bc_params),
                         cfg.gas, cfg.scheme.limiter, cfg.ghost, workspace=True, adjoint=True)
        self.nb = len(self.q0)
        if autotune:   # stage-kernel variant for these blocks (state untouched)
            self.plan.autotune(self.q0, self.kxyz, self.vol, 0.0, batched=batched)
        self._work = torch.empty(1024, dtype=torch.float64, device="cuda")
        self._J = torch.empty(1, dtype=torch.float64, device="cuda")

    def set_block_params(self, mu0=None, dt_scale=None):
        self.plan.set_block_params(mu0, dt_scale)

    def _step(self, qs, dt):
        self.plan.step(qs, self.kxyz, self.vol, dt, batched=self.batched, direct=self.direct)

    def _step_stages(self, qs, dt):
        """_step that also returns the step's stage-1/2 states (new tensors the
        step writes them into)."""
        q1s = [torch_empty_like(q) for q in qs]
        q2s = [torch_empty_like(q) for q in qs]
        self.plan.step(qs, self.kxyz, self.vol, dt, batched=self.batched, direct=self.direct, stages=(q1s, q2s))
        return q1s, q2s

    def forward(self, dt, nsteps, qs=None):
        qs = [q.clone() for q in (qs or self.q0)]
        for _ in range(nsteps):
            self._step(qs, dt)
        return qs

    def forward_checkpointed(self, qs, dt, nsteps, checkpoint_every, keep_last_segment=False):
        """Advance copies of ``qs``; returns (final states, {step: states})
        with the states kept every ``checkpoint_every`` steps (step 0 included).
        keep_last_segment also keeps the last segment's step-start and stage
        states, returned third as (first step of the segment, starts, stages),
        so that the reverse pass does not recompute that segment."""
        cur = [q.detach().clone() for q in qs]
        checkpoints = {0: [q.clone() for q in cur]}
        last_start = (nsteps - 1) // checkpoint_every * checkpoint_every if nsteps > 0 else 0
        starts, stages = [], []
        for n in range(nsteps):
            keep = keep_last_segment and n >= last_start
            if keep:
                starts.append([q.clone() for q in cur])
                stages.append(self._step_stages(cur, dt))
            else:
                self._step(cur, dt)
            if (n + 1) % checkpoint_every == 0 and n + 1 < nsteps:
                checkpoints[n + 1] = [q.clone() for q in cur]
        if keep_last_segment:
            return cur, checkpoints, (last_start, starts, stages)
        return cur, checkpoints

    def objective(self, qs, kind, blocks=None):
        """J (float) and dJ/dq (list of device tensors) for a build-defined
        objective summed over ``blocks`` (default: all)."""
        torch = L._torch()
        blocks = range(self.nb) if blocks is None else blocks
        J = 0.0
        grads = [torch.zeros_like(q) for q in qs]
        for b in blocks:
            _, ni, nj, nk = qs[b].shape
            L.check(L.lib().famrdp_objective(int(kind), L.ptr(qs[b]), L.ptr(self.kxyz[b]), L.ptr(self.vol[b]), ni, nj,
                                             nk, int(self.cfg.ghost), float(self.cfg.gas.gamma), L.ptr(self._J),
                                             L.ptr(grads[b]), L.ptr(self._work), L.stream()))
            J += float(self._J.item())
        return J, grads

    def reverse(self, lam, dt, nsteps, checkpoints, checkpoint_every, dqinf, *, keep_stages=True, last=None):
        """lam (cotangent of the final state) -> cotangent of q0, in place.
        Each segment is recomputed from its checkpoint keeping every step's
        start state and (keep_stages) its stage-1/2 states, so the reverse of
        a step does not recompute its stages; ``last`` (from
        forward_checkpointed(..., keep_last_segment=True)) supplies the last
        segment's states instead of a recomputation."""
        k = checkpoint_every
        for seg_start in reversed(range(0, nsteps, k)):
            seg_end = min(seg_start + k, nsteps)
            if last is not None and keep_stages and last[0] == seg_start:
                _, starts, stages = last
                last = None
            else:
                cur = [c.clone() for c in checkpoints[seg_start]]
                starts, stages = [], []
                for _ in range(seg_start, seg_end):
                    starts.append([c.clone() for c in cur])
                    if keep_stages:
                        stages.append(self._step_stages(cur, dt))
                    else:
                        self._step(cur, dt)
            for n in reversed(range(seg_end - seg_start)):
                self.plan.step_vjp(starts[n], self.kxyz, self.vol, lam, dt, dqinf, batched=self.batched,
                                   direct=self.direct, stages=stages[n] if keep_stages else None,
                                   consume_start=True)
                starts[n] = None
                if keep_stages:
                    stages[n] = None    # release as the reverse pass goes
        return lam

    def gradient(self, dt, nsteps, kind=OBJ_KINETIC_ENERGY, *, checkpoint_every=10, blocks=None, timing=None):
        """Returns (J, dJ/dq0 per block, dJ/dq_inf as (nblocks, 6, 5) tensor).
        checkpoint_every = k > 0 keeps the state every k steps and recomputes
        one segment at a time in the reverse pass; k <= 0 keeps every step's
        start and stage states from the forward pass (3 states per step in
        device memory) and recomputes nothing."""
        torch = L._torch()
        t0 = time.perf_counter()
        dqinf = torch.zeros(self.nb * 6 * 5, dtype=torch.float64, device="cuda")
        if checkpoint_every <= 0:
            cur = [q.clone() for q in self.q0]
            starts, stages = [], []
            for _ in range(nsteps):
                starts.append([q.clone() for q in cur])
                stages.append(self._step_stages(cur, dt))
            J, lam = self.objective(cur, kind, blocks)
            t1 = time.perf_counter()
            for n in reversed(range(nsteps)):
                self.plan.step_vjp(starts[n], self.kxyz, self.vol, lam, dt, dqinf, batched=self.batched,
                                   direct=self.direct, stages=stages[n], consume_start=True)
                starts[n] = stages[n] = None    # release as the reverse pass goes
        else:
            cur, checkpoints, last = self.forward_checkpointed(self.q0, dt, nsteps, checkpoint_every,
                                                               keep_last_segment=True)
            J, lam = self.objective(cur, kind, blocks)
            t1 = time.perf_counter()
            self.reverse(lam, dt, nsteps, checkpoints, checkpoint_every, dqinf, last=last)
        torch.cuda.synchronize()
        if timing is not None:
            timing["forward_with_checkpoints_s"] = t1 - t0
            timing["total_s"] = time.perf_counter() - t0
        return J, lam, dqinf.reshape(self.nb, 6, 5)


def mach_derivative(dqinf, mach):
    """d J / d M_inf from the q_inf cotangents: q_inf = (1, M, 0, 0, p/(g-1) + M^2/2)."""
    return float(dqinf[..., 1].sum().item() + mach * dqinf[..., 4].sum().item())


def _torch_function():
    torch = L._torch()

    class StepsFunction(torch.autograd.Function):
        """torch.autograd wrapper: q_final = steps(q0) for a single-block
        configuration, differentiated by the hand-written adjoint with
        checkpointing (q0: (5, ni, nj, nk) float64 CUDA tensor)."""

        @staticmethod
        def forward(ctx, q0, adj: Adjoint, dt: float, nsteps: int, checkpoint_every: int = 10):
            cur, ckpts = adj.forward_checkpointed([q0], dt, nsteps, checkpoint_every)
            ctx.adj, ctx.dt, ctx.nsteps, ctx.k, ctx.ckpts = adj, dt, nsteps, checkpoint_every, ckpts
            return cur[0]

        @staticmethod
        def backward(ctx, grad):
            lam = [grad.detach().contiguous().clone()]
            ctx.adj.reverse(lam, ctx.dt, ctx.nsteps, ctx.ckpts, ctx.k, None)
            return lam[0], None, None, None, None

    return StepsFunction


def steps_autograd(q0, adj: Adjoint, dt: float, nsteps: int, checkpoint_every: int = 10):
    """Differentiable ``nsteps`` steps of one block for PyTorch callers."""
    return _torch_function().apply(q0, adj, dt, nsteps, checkpoint_every)
