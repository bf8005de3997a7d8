"""CFL time step and positivity monitor on the GPU.

The reference carries ``Config.cfl`` and ``Config.nan_check``
(famrdp_jax/core/types.py:187-189) but never reads them and has no CFL->dt
map.  This build defines one global dt per run:

    dt = cfl / max_interior [ sum_d (|u . k_d| + c |k_d|) / vol
                              + max(4/3, gamma/Pr) (mu/rho) sum_d |k_d|^2 ]

(k_d = kxyz[d, :]); it follows the reference operator's actual scaling (the
inviscid term divided by vol, the viscous one not).  The same reduction
counts interior cells that are non-finite or have rho <= 0 or p <= 0.
"""
from __future__ import annotations

import ctypes

from .. import _lib as L


def cfl_dt(q, metrics, cfg, *, ghost=None):
    """Returns (dt, n_bad) for one block; n_bad counts non-finite or
    non-positive (rho, p) interior cells."""
    torch = L._torch()
    qd = L.device_f64(q)
    kx = L.device_f64(metrics.kxyz)
    vo = L.device_f64(metrics.vol)
    _, ni, nj, nk = qd.shape
    out = torch.empty(3, dtype=torch.float64, device=qd.device)
    work = torch.empty(2048, dtype=torch.float64, device=qd.device)
    g = L.gas_struct(cfg.gas)
    gh = cfg.ghost if ghost is None else ghost
    L.check(L.lib().famrdp_cfl_dt(L.ptr(qd), L.ptr(kx), L.ptr(vo), ni, nj, nk, gh, ctypes.byref(g), float(cfg.cfl),
                                  L.ptr(out), L.ptr(work), L.stream()))
    res = out.cpu().tolist()
    return res[0], int(res[2])


def check_state(state, cfg):
    """Config.nan_check: raise FloatingPointError if any block has a
    non-finite or non-positive interior cell."""
    for b, (blk, m) in enumerate(zip(state.blocks, state.metrics)):
        _, nbad = cfl_dt(blk.q, m, cfg)
        if nbad:
            raise FloatingPointError(f"block {b}: {nbad} interior cells non-finite or with rho/p <= 0")
