"""The C + OpenMP restatement of the step (oracle/c_step.py, used for the CPU
baseline on all host cores) against the numpy oracle: same results to a few
units in the last place (bitwise where no libm call differs), independent of
the thread count, and refused for topologies it does not cover."""
import numpy as np
import pytest

from famrdp_jax import cases
from oracle import blocks as OB
from oracle import c_step
from oracle import metric as OM


@pytest.fixture(scope="module")
def built():
    try:
        c_step.build()
    except Exception as e:  # pragma: no cover - gcc/make absent
        pytest.skip(f"C oracle not buildable here: {e!r}")
    return c_step.load()


def _c2(n):
    q, xyz_ext, cfg = cases.case_c2(n)
    jac, k, vol = OM.compute_metrics(xyz_ext)
    m = cases.crop_metrics(jac, k, vol)
    return q, m, cfg


@pytest.mark.parametrize("limiter", ["minmod", "vanleer"])
@pytest.mark.parametrize("case", ["c1", "c2"])
def test_c_step_matches_numpy_oracle(built, case, limiter):
    if case == "c1":
        state, cfg = cases.case_c1(n=12, seed=1)
        q, m = state.blocks[0].q, state.metrics[0]
    else:
        q, m, cfg = _c2(10)
    gas = cases.gas_dict(cfg.gas)
    dt = OM.cfl_dt(q, m.kxyz, m.vol, gas, 0.5)
    ref = OB.step([q], [(m.kxyz, m.vol)], list(cfg.topology), cfg.bc_params, gas, dt, limiter=limiter)[0]
    got = c_step.step(q, m.kxyz, m.vol, cfg.topology[0], gas, dt, limiter=limiter)
    for v in range(5):
        assert np.abs(got[v] - ref[v]).max() <= 1e-15 * np.abs(ref[v]).max()
    one = c_step.step(q, m.kxyz, m.vol, cfg.topology[0], gas, dt, limiter=limiter, threads=1)
    np.testing.assert_array_equal(got, one)


def test_c_step_refuses_non_periodic_blocks(built):
    q, xyz_ext, cfg = cases.case_c4(n=8)
    jac, k, vol = OM.compute_metrics(xyz_ext)
    m = cases.crop_metrics(jac, k, vol)
    with pytest.raises(NotImplementedError):
        c_step.step(q, m.kxyz, m.vol, cfg.topology[0], cases.gas_dict(cfg.gas), 1e-3)
