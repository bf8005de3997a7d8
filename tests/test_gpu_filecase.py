"""A multi-block State built from files (mesh/io_grd.py, io_top.py with the
orientation normalisation, build_state.py), metrics computed on the GPU,
advanced three steps on the GPU against the oracle on the same inputs."""
import numpy as np
import pytest
import torch

from famrdp_jax import cases
from famrdp_jax.mesh.metric import compute_metrics
from famrdp_jax.solver.stepper import Stepper
from oracle import blocks as OB
from oracle import metric as OM
from _filecase import file_case
from _util import rel_linf, to_dev

pytestmark = pytest.mark.gpu


def _gpu_metrics(xyz):
    m = compute_metrics(xyz, ghost=2)     # numpy in, numpy out (computed on the GPU)
    return m.jac, m.kxyz, m.vol


def test_file_built_two_block_case_vs_oracle(tmp_path):
    print("case: mini.inp topology (aligned orientation 0) + two 5x5x3 blocks, GPU metrics, 3 steps", flush=True)
    state, cfg, dt = file_case(str(tmp_path), _gpu_metrics)
    # the GPU metrics are the oracle's to rounding
    ref_state = file_case(str(tmp_path), OM.compute_metrics)[0]
    for a, b in zip(state.metrics, ref_state.metrics):
        np.testing.assert_allclose(a.kxyz, b.kxyz, rtol=0, atol=1e-12 * np.abs(b.kxyz).max())
    qs = [b.q for b in state.blocks]
    mets = [(m.kxyz, m.vol) for m in state.metrics]
    ref = qs
    for _ in range(3):
        ref = OB.step(ref, mets, list(cfg.topology), cfg.bc_params, cases.gas_dict(cfg.gas), dt)
    outs = {}
    for name, kw in (("plain", dict()), ("batched", dict(batched=True)), ("direct", dict(direct=True))):
        st = Stepper(to_dev(state), cfg)
        st.advance(dt, 3, **kw)
        torch.cuda.synchronize()
        outs[name] = [q.cpu().numpy() for q in st.qs]
        err = max(rel_linf(a, b) for a, b in zip(outs[name], ref))
        print(f"  {name}: rel Linf after 3 steps {err:.3e}", flush=True)
        assert err <= 1e-12
    for a, b in zip(outs["batched"], outs["plain"]):
        np.testing.assert_array_equal(a, b)
