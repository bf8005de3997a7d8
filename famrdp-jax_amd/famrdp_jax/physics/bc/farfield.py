"""Simplified farfield on the GPU (famrdp_jax/physics/bc/farfield.py:7-21)."""
from __future__ import annotations

from ..._plan import BC_FARFIELD, _floats
from ._run import run_face


def apply_farfield(q, face, ghost, q_inf):
    """Every ghost layer of ``face`` set to q_inf."""
    return run_face(q, face, ghost, BC_FARFIELD, q_inf=_floats(q_inf))
