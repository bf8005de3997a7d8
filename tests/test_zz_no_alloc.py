"""famrdp_cfl_dt_noalloc (no counterpart in the reference; for callers that
cannot allocate) against famrdp_cfl_dt, through the C ABI: two calls in a row
on one GPU, the second on a smaller block with a NaN cell, each giving the
same dt, spectral radius and bad-cell count as the entry point that takes a
caller workspace, bit for bit (the same kernels over the same partials)."""
import ctypes

import numpy as np
import pytest
import torch

from famrdp_jax import _lib as L, cases
from _util import to_dev

pytestmark = pytest.mark.gpu


def _both(state, cfg):
    dev = to_dev(state)
    q, m = dev.blocks[0].q, dev.metrics[0]
    _, ni, nj, nk = q.shape
    g = L.gas_struct(cfg.gas)
    out_w = torch.full((3,), -1.0, dtype=torch.float64, device="cuda")
    out_n = torch.full((3,), -2.0, dtype=torch.float64, device="cuda")
    work = torch.empty(2048, dtype=torch.float64, device="cuda")
    args = (L.ptr(q), L.ptr(m.kxyz), L.ptr(m.vol), ni, nj, nk, 2, ctypes.byref(g), 0.5)
    L.check(L.lib().famrdp_cfl_dt(*args, L.ptr(out_w), L.ptr(work), L.stream()))
    L.check(L.lib().famrdp_cfl_dt_noalloc(*args, L.ptr(out_n), L.stream()))
    torch.cuda.synchronize()
    return out_w.cpu().numpy(), out_n.cpu().numpy()


def test_cfl_dt_noalloc_matches_workspace_entry_point():
    print("case: famrdp_cfl_dt_noalloc vs famrdp_cfl_dt, C1 24^3 then 12^3 with a NaN cell", flush=True)
    big, cfg = cases.case_c1(n=24, seed=3, p_noise=1e-3)
    w, n = _both(big, cfg)
    print(f"  24^3: dt {n[0]:.17g} lam {n[1]:.17g} bad {n[2]:.0f}", flush=True)
    assert np.array_equal(w, n) and n[0] > 0 and n[2] == 0
    small, cfg2 = cases.case_c1(n=12, seed=4, p_noise=1e-3)
    small.blocks[0].q[0, 6, 7, 8] = np.nan
    w2, n2 = _both(small, cfg2)
    print(f"  12^3: dt {n2[0]:.17g} lam {n2[1]:.17g} bad {n2[2]:.0f}", flush=True)
    assert np.array_equal(w2, n2) and n2[0] > 0 and n2[2] == 1
    assert n2[1] != n[1]       # the second call's result is its own, not the first's
