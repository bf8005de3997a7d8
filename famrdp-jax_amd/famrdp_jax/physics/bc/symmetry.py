"""Symmetry plane on the GPU (famrdp_jax/physics/bc/symmetry.py:8-11)."""
from __future__ import annotations

from ..._plan import BC_MIRROR_NORMAL
from ._run import run_face


def apply_symmetry_plane(q, face, ghost):
    return run_face(q, face, ghost, BC_MIRROR_NORMAL)
