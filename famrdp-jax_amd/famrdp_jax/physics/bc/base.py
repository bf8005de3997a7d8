"""Face helpers of famrdp_jax/physics/bc/base.py (host-side index logic)."""
from ...core.constants import Face

_AXIS = {Face.IMIN: 1, Face.IMAX: 1, Face.JMIN: 2, Face.JMAX: 2, Face.KMIN: 3, Face.KMAX: 3}


def face_axis(face: Face) -> int:
    return _AXIS[Face(int(face))]


def vel_component(face: Face) -> int:
    """Index of the face-normal momentum in q (1, 2 or 3)."""
    return _AXIS[Face(int(face))]


def ghost_indices(face: Face, ghost: int, ni_total: int):
    """(outer, inner) mirror pairs: IMIN outer [0..g-1] <- inner [2g-1..g];
    IMAX outer [n-1..n-g] <- inner [n-2g..n-g-1] (bc/base.py:21-33)."""
    if int(face) in (Face.IMIN, Face.JMIN, Face.KMIN):
        outer = list(range(ghost))
        inner = [2 * ghost - 1 - g for g in outer]
    else:
        outer = list(range(ni_total - 1, ni_total - 1 - ghost, -1))
        inner = [ni_total - 2 * ghost + (ni_total - 1 - o) for o in outer]
    return outer, inner
