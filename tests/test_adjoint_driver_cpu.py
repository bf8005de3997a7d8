"""CPU check of the checkpointing driver (famrdp_jax.adjoint.Adjoint:
forward_checkpointed / reverse) with a stand-in plan whose step adds 1 to the
state and whose stage states are tagged, so every reverse call shows which
step-start and stage states it received.  The reverse pass must reverse every
step exactly once, last step first, with that step's own start and stage
states, whether the last segment's states come from the forward pass (no
recompute of that segment) or are recomputed.  No GPU: the HIP reverse sweep
itself is covered by tests/test_gpu_adjoint.py."""
import pytest
import torch

from famrdp_jax.adjoint import Adjoint


class _Plan:
    def __init__(self):
        self.steps = 0
        self.vjps = []
        self._last = None

    def step(self, qs, kxyz, vol, dt, batched=False, direct=False, stages=None):
        for q in qs:
            q += 1.0
        self._last = [q.clone() for q in qs]
        self.steps += 1
        if stages is not None:   # the step writes its stage states into the caller's arrays
            self.stage_states(*stages)

    def stage_states(self, q1s, q2s):
        # stage states of the step that just ran (ending at state n + 1): tags n + 1.25, n + 1.5
        for a, b, q in zip(q1s, q2s, self._last):
            a.copy_(q + 0.25)
            b.copy_(q + 0.5)

    def step_vjp(self, starts, kxyz, vol, lam, dt, dqinf, batched=False, direct=False, stages=None,
                 consume_start=False):
        s = float(starts[0].flatten()[0])
        tags = None if stages is None else (float(stages[0][0].flatten()[0]), float(stages[1][0].flatten()[0]))
        self.vjps.append((s, tags))
        if consume_start:   # the plan may overwrite a consumed start: the driver must not read it again
            for q in starts:
                q += 1000.0


def _adjoint():
    adj = Adjoint.__new__(Adjoint)   # the driver only; no plan or GPU behind it
    adj.plan = _Plan()
    adj.kxyz, adj.vol = [None], [None]
    adj.batched = adj.direct = False
    return adj


@pytest.mark.parametrize("nsteps,k", [(7, 3), (6, 3), (1, 10), (10, 10), (50, 10), (5, 1)])
@pytest.mark.parametrize("keep_last", [True, False])
def test_reverse_order_states_and_recompute_count(nsteps, k, keep_last):
    adj = _adjoint()
    q0 = [torch.zeros(5, 2, 2, 2, dtype=torch.float64)]
    out = adj.forward_checkpointed(q0, 0.1, nsteps, k, keep_last_segment=keep_last)
    final, ckpts = out[0], out[1]
    last = out[2] if keep_last else None
    assert float(final[0].flatten()[0]) == nsteps
    assert sorted(ckpts) == list(range(0, nsteps, k))
    assert all(float(ckpts[s][0].flatten()[0]) == s for s in ckpts)
    forward_steps = adj.plan.steps
    assert forward_steps == nsteps
    if keep_last:
        seg = last[0]
        assert seg == (nsteps - 1) // k * k
        assert [float(s[0].flatten()[0]) for s in last[1]] == list(range(seg, nsteps))
    lam = [torch.zeros_like(q0[0])]
    adj.reverse(lam, 0.1, nsteps, ckpts, k, None, last=last)
    # every step reversed once, last first, with its own start and stage states
    assert adj.plan.vjps == [(float(n), (n + 1.25, n + 1.5)) for n in reversed(range(nsteps))]
    # the kept segment is not recomputed; every other one is, once
    kept = nsteps - ((nsteps - 1) // k * k) if keep_last else 0
    assert adj.plan.steps - forward_steps == nsteps - kept


def test_reverse_without_kept_stages_recomputes_nothing_extra():
    adj = _adjoint()
    q0 = [torch.zeros(5, 2, 2, 2, dtype=torch.float64)]
    final, ckpts = adj.forward_checkpointed(q0, 0.1, 5, 2)
    adj.reverse([torch.zeros_like(q0[0])], 0.1, 5, ckpts, 2, None, keep_stages=False)
    assert adj.plan.vjps == [(float(n), None) for n in reversed(range(5))]
    assert adj.plan.steps == 5 + 5
