"""Invalid states go NaN the way the reference's do.

The reference propagates NaN through jnp.maximum floors (flux/roe.py:23-24,
33, 39; rhs/viscous.py:36), the minmod limiter's sign/maximum/minimum
(physics/reconstruct/muscl2.py:5-7) and Sutherland's (T/T0)**1.5, which is
NaN for a negative temperature (physics/eos.py:24-25).  A C1 state with a NaN
density cell, a NaN momentum cell and a negative-pressure cell is advanced
one step by every kernel form; the set of non-finite values must equal the
oracle's exactly and the finite values must match it to 1e-12.
"""
import numpy as np
import pytest
import torch

from famrdp_jax import _lib, cases
from famrdp_jax.solver.stepper import Stepper
from oracle import blocks as OB
from oracle import metric as OM
from _util import to_dev

pytestmark = pytest.mark.gpu


def _poisoned(limiter):
    from famrdp_jax.core.types import Block, Config, SchemeChoice, State
    state, cfg = cases.case_c1(n=16, seed=6, p_noise=1e-3)
    q = state.blocks[0].q.copy()
    m = state.metrics[0]
    dt = OM.cfl_dt(q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5)
    q[0, 6, 7, 8] = np.nan                  # density: every primitive NaN
    q[2, 14, 12, 3] = np.nan                # one momentum: v and p NaN (halo wraps it)
    ke = 0.5 * (q[1, 10, 4, 13] ** 2 + q[2, 10, 4, 13] ** 2 + q[3, 10, 4, 13] ** 2) / q[0, 10, 4, 13]
    q[4, 10, 4, 13] = ke - 0.3              # negative pressure: T < 0, Sutherland NaN
    cfg = Config(gas=cfg.gas, scheme=SchemeChoice("muscl2", "roe", 3, limiter), bc_params=cfg.bc_params,
                 topology=cfg.topology, ghost=2, nan_check=False)
    st = State(blocks=(Block(q=q, xyz=state.blocks[0].xyz),), metrics=state.metrics)
    return st, cfg, dt


@pytest.mark.parametrize("limiter", ["minmod", "vanleer"])
def test_nan_and_negative_pressure_masks_match_oracle(limiter):
    print(f"case: C1 16^3 with NaN / negative-pressure cells, {limiter}, 1 step", flush=True)
    state, cfg, dt = _poisoned(limiter)
    m = state.metrics[0]
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        ref = OB.step([state.blocks[0].q], [(m.kxyz, m.vol)], list(cfg.topology), cfg.bc_params,
                      cases.gas_dict(cfg.gas), dt, limiter=limiter)[0]
    bad = ~np.isfinite(ref)
    assert bad.any() and not bad.all()
    print(f"  oracle: {int(bad.any(axis=0).sum())} non-finite cells", flush=True)
    split = [v for v in range(_lib.NUM_TILE_VARIANTS) if v not in _lib.FUSED_VARIANTS][0]
    runs = [("autotuned", dict(), "auto"), ("batched", dict(batched=True), "auto"),
            ("split", dict(), split), ("split_batched", dict(batched=True), split),
            ("direct", dict(direct=True), None), ("direct_batched", dict(direct=True, batched=True), None)]
    scale = [np.abs(ref[v][np.isfinite(ref[v])]).max() for v in range(5)]
    for name, kw, tile in runs:
        st = Stepper(to_dev(state), cfg)
        if tile == "auto":
            st.autotune(dt, batched=kw.get("batched", False), reps=1)
        elif tile is not None:
            st.plan.set_tile(tile)
        st.advance(dt, **kw)
        torch.cuda.synchronize()
        got = st.qs[0].cpu().numpy()
        np.testing.assert_array_equal(np.isnan(got), np.isnan(ref), err_msg=name)
        np.testing.assert_array_equal(np.isfinite(got), np.isfinite(ref), err_msg=name)
        ok = ~bad
        err = max(float(np.abs(got[v][ok[v]] - ref[v][ok[v]]).max()) / scale[v] for v in range(5))
        print(f"  {name}: masks equal, finite cells rel Linf {err:.3e}", flush=True)
        assert err <= 1e-12, name


def test_nan_check_raises_after_the_step():
    """Config.nan_check (core/types.py:187, default True): the functional step
    and Stepper.advance monitor the new state on the GPU and raise
    FloatingPointError for a non-finite or non-positive (rho, p) interior
    cell; a valid state passes, and nan_check=False never checks."""
    from famrdp_jax.core.types import Config
    from famrdp_jax.solver.step import step
    print("case: nan_check on a C1 16^3 state with a NaN cell", flush=True)
    state, cfg, dt = _poisoned("minmod")
    on = Config(gas=cfg.gas, scheme=cfg.scheme, bc_params=cfg.bc_params, topology=cfg.topology, ghost=2,
                nan_check=True)
    with pytest.raises(FloatingPointError):
        step(to_dev(state), on, dt)
    st = Stepper(to_dev(state), on)
    with pytest.raises(FloatingPointError):
        st.advance(dt, 2, graph=True)
    Stepper(to_dev(state), cfg).advance(dt, 1)          # nan_check=False: no check
    clean, ccfg = cases.case_c1(n=16, seed=6)
    ok = Config(gas=ccfg.gas, scheme=ccfg.scheme, bc_params={}, topology=ccfg.topology, ghost=2, nan_check=True)
    out = step(to_dev(clean), ok, dt)
    assert out.step == 1
    Stepper(to_dev(clean), ok).advance(dt, 3, batched=True)
