"""Checks at the benchmark's own sizes (each well under a minute).

* 256^3 curvilinear TGV: after 99 graph-replayed steps, step 100 on a
  centre sub-box matches one oracle step from the GPU's own step-99 state.
* 64 uneven blocks (19,996,944 interior cells): the batched hipGraph step
  equals per-block plain launches bitwise and stays finite.
* 96^3 wall/farfield gradient over 50 steps: checkpointing every 10 steps
  gives the same bits as keeping every step's states, and is finite.
* 32-case ensemble: the batched step equals per-block launches bitwise.
"""
import numpy as np
import pytest
import torch

from famrdp_jax import cases
from famrdp_jax.core.types import Block, BlockTopology, Metrics, State
from famrdp_jax.solver.cfl import cfl_dt
from famrdp_jax.solver.stepper import Stepper, state_to_device
from oracle import blocks as OB
from _util import rel_linf

pytestmark = pytest.mark.gpu


def _curvilinear_on_gpu(q, xyz_ext, cfg):
    from famrdp_jax.mesh.metric import compute_metrics
    m = compute_metrics(torch.from_numpy(xyz_ext).cuda(), ghost=2)
    sl = (slice(2, -2),) * 3
    mc = Metrics(jac=m.jac[sl].contiguous(), kxyz=m.kxyz[(slice(None), slice(None)) + sl].contiguous(),
                 vol=m.vol[sl].contiguous())
    del m
    qd = torch.from_numpy(q).cuda()
    dt, nbad = cfl_dt(qd, mc, cfg)
    assert nbad == 0
    state = State(blocks=(Block(q=qd, xyz=torch.zeros(1, dtype=torch.float64, device="cuda")),), metrics=(mc,))
    return state, cfg, dt


def test_c2_256_step_100_centre_subbox_vs_oracle():
    print("case: C2 256^3, 99 graph steps, then step 100 vs oracle on a centre sub-box", flush=True)
    state, cfg, dt = _curvilinear_on_gpu(*cases.case_c2(256))
    st = Stepper(state, cfg)
    st.advance(dt, 99, graph=True)
    torch.cuda.synchronize()
    c0, w = 117, 26
    sl = (slice(None),) + (slice(c0, c0 + w),) * 3
    q_before = st.qs[0][sl].cpu().numpy()
    k_sub = state.metrics[0].kxyz[(slice(None), slice(None)) + sl[1:]].cpu().numpy()
    v_sub = state.metrics[0].vol[sl[1:]].cpu().numpy()
    st.advance(dt)
    torch.cuda.synchronize()
    assert torch.isfinite(st.qs[0]).all()
    got = st.qs[0][sl].cpu().numpy()
    # an isolated sub-box: no halo, no BC; cells >= 7 from its faces are
    # outside the crop's influence (3 stages x stencil radius 2)
    ref = OB.step([q_before], [(k_sub, v_sub)], [BlockTopology(0, {}, {})], {}, cases.gas_dict(cfg.gas), dt)[0]
    inner = (slice(None),) + (slice(7, -7),) * 3
    err = rel_linf(got[inner], ref[inner])
    moved = rel_linf(got[inner], q_before[inner])
    print(f"  step 100 sub-box rel Linf {err:.3e} (state change in the step {moved:.3e})", flush=True)
    assert err <= 1e-12


def test_c3_full_64_blocks_batched_graph_equals_per_block():
    print("case: C3 64 blocks (seed 883), 2 steps, batched graph vs per-block plain", flush=True)
    state, cfg = cases.case_c3()
    dev = state_to_device(state)
    del state
    assert sum((b.q.shape[1] - 4) * (b.q.shape[2] - 4) * (b.q.shape[3] - 4) for b in dev.blocks) == 19996944
    dt = min(cfl_dt(b.q, m, cfg)[0] for b, m in zip(dev.blocks, dev.metrics))
    outs = {}
    for name, kw in (("per_block_plain", dict()), ("batched_graph", dict(batched=True, graph=True))):
        st = Stepper(dev, cfg)
        st.autotune(dt, batched=kw.get("batched", False), reps=1)
        print(f"  {name}: tile variant {st.plan.tile()[0]}", flush=True)
        st.advance(dt, 2, **kw)
        torch.cuda.synchronize()
        outs[name] = [q.clone() for q in st.qs]
        del st
    for a, b in zip(outs["batched_graph"], outs["per_block_plain"]):
        assert torch.isfinite(a).all()
        assert torch.equal(a, b)
    moved = max(float((a - q.q).abs().max()) for a, q in zip(outs["batched_graph"], dev.blocks))
    print(f"  bitwise equal, finite; largest change {moved:.3e}", flush=True)
    assert moved > 0.0


def test_c4_96_checkpointed_gradient_equals_all_states():
    from famrdp_jax.adjoint import OBJ_WALL_PRESSURE, Adjoint
    print("case: C4 96^3 wall/farfield, 50 steps, checkpoint 10 vs every state kept", flush=True)
    state, cfg, dt = _curvilinear_on_gpu(*cases.case_c4(96))
    adj = Adjoint(state, cfg, autotune=True)
    J1, g1, q1 = adj.gradient(dt, 50, OBJ_WALL_PRESSURE, checkpoint_every=10)
    J0, g0, q0 = adj.gradient(dt, 50, OBJ_WALL_PRESSURE, checkpoint_every=0)
    assert J1 == J0
    assert torch.isfinite(g1[0]).all() and bool((g1[0] != 0).any())
    assert torch.equal(g1[0], g0[0])
    assert torch.equal(q1, q0)
    print(f"  J {J1:.12e}; gradients bitwise equal, max |dJ/dq0| {float(g1[0].abs().max()):.4e}", flush=True)


def test_c5_32_cases_batched_equals_per_block():
    print("case: C5 32 cases, 3 steps, batched (graph) vs per-block plain", flush=True)
    pairs = cases.case_c5(32)
    state, cfg, mu0 = cases.ensemble(pairs)
    dev = state_to_device(state)
    dts = [cfl_dt(b.q, m, c)[0] for b, m, (_, c) in zip(dev.blocks, dev.metrics, pairs)]
    outs = {}
    for name, kw in (("per_block_plain", dict()), ("batched", dict(batched=True)),
                     ("batched_graph", dict(batched=True, graph=True))):
        st = Stepper(dev, cfg, mu0=mu0, dt_scale=dts)
        st.advance(1.0, 3, **kw)
        torch.cuda.synchronize()
        outs[name] = [q.clone() for q in st.qs]
        del st
    for name in ("batched", "batched_graph"):
        for a, b in zip(outs[name], outs["per_block_plain"]):
            assert torch.isfinite(a).all()
            assert torch.equal(a, b), name
    print("  bitwise equal, finite", flush=True)
