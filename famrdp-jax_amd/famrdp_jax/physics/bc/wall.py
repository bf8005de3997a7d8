"""Wall BCs on the GPU (famrdp_jax/physics/bc/wall.py)."""
from __future__ import annotations

from ..._plan import BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL, BC_MIRROR_NORMAL
from ._run import run_face


def apply_wall_noslip_adiabatic(q, face, ghost):
    """wall.py:30-32: mirror with sign [1, -1, -1, -1, 1]."""
    return run_face(q, face, ghost, BC_WALL_ADIABATIC)


def apply_wall_noslip_isothermal(q, face, ghost, T_wall, gamma, R):
    """wall.py:35-58: p_g = p_i, rho_g = p_g/(R T_wall), momentum reversed and
    scaled by rho_g/rho_i, energy rebuilt."""
    return run_face(q, face, ghost, BC_WALL_ISOTHERMAL, T_wall=T_wall, gamma=gamma, R=R)


def apply_wall_slip(q, face, ghost):
    """wall.py:61-64: mirror flipping the face-normal momentum."""
    return run_face(q, face, ghost, BC_MIRROR_NORMAL)
