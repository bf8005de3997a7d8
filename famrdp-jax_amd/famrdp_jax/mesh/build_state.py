"""Initial multi-block State from grid and topology files
(famrdp_jax/mesh/build_state.py:13-45)."""
from __future__ import annotations

import numpy as np

from ..core.types import Block, Metrics, State
from .io_grd import read_grd
from .io_top import parse_topology


def build_initial_state(grd_path, inp_path, ghost: int, q_init_fn, *, aligned_orient=None):
    """Coordinates padded by ``ghost`` layers per side by edge replication,
    q = q_init_fn((5, ni, nj, nk)) on the padded shape, and placeholder
    metrics (jac = vol = 1, kxyz = 0) as in the reference; compute real
    metrics with famrdp_jax.mesh.metric.compute_metrics.  Returns (State,
    tuple of BlockTopology).  ``aligned_orient`` is passed to parse_topology
    (0: aligned interfaces get orientation 0, which the halo exchange runs)."""
    topos = tuple(parse_topology(inp_path, aligned_orient=aligned_orient))
    blocks, metrics = [], []
    for xyz in read_grd(grd_path):
        xp = np.pad(xyz, ((0, 0),) + ((ghost, ghost),) * 3, mode="edge")
        shape = xp.shape[1:]
        q = np.asarray(q_init_fn((5,) + shape), dtype=np.float64)
        blocks.append(Block(q=q, xyz=xp))
        metrics.append(Metrics(jac=np.ones(shape), kxyz=np.zeros((3, 3) + shape), vol=np.ones(shape)))
    return State(blocks=tuple(blocks), metrics=tuple(metrics), t=0.0, step=0), topos
