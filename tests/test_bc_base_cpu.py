"""Known pairs of the BC face helpers (famrdp_jax/physics/bc/base.py:15-33):
the mirror pairs of every face at ghost = 1, 2, 3 and the face-axis maps."""
import pytest

from famrdp_jax.core.constants import Face
from famrdp_jax.physics.bc.base import face_axis, ghost_indices, vel_component


def test_face_axis_and_velocity_component():
    assert [face_axis(f) for f in Face] == [1, 1, 2, 2, 3, 3]
    assert [vel_component(f) for f in Face] == [1, 1, 2, 2, 3, 3]
    assert face_axis(int(Face.KMAX)) == 3           # plain ints are accepted too


@pytest.mark.parametrize("face", [Face.IMIN, Face.JMIN, Face.KMIN])
def test_min_faces(face):
    # bc/base.py:23-25 docstring: ghost = 2 -> outer [0, 1], inner [3, 2]
    assert ghost_indices(face, 2, 10) == ([0, 1], [3, 2])
    assert ghost_indices(face, 1, 10) == ([0], [1])
    assert ghost_indices(face, 3, 10) == ([0, 1, 2], [5, 4, 3])


@pytest.mark.parametrize("face", [Face.IMAX, Face.JMAX, Face.KMAX])
def test_max_faces(face):
    # outer n-1, n-2 mirror n-4, n-3 (the layers next to the last interior cell)
    assert ghost_indices(face, 2, 10) == ([9, 8], [6, 7])
    assert ghost_indices(face, 1, 10) == ([9], [8])
    assert ghost_indices(face, 3, 12) == ([11, 10, 9], [6, 7, 8])


def test_pairs_are_mirror_images_about_the_face():
    # every pair sits symmetric about the cell face between the last ghost and
    # the first interior layer: outer + inner = 2 g - 1 (min) or 2 n - 2 g - 1 (max)
    for g in (1, 2, 3, 4):
        for n in (2 * g + 1, 2 * g + 5, 40):
            o, i = ghost_indices(Face.JMIN, g, n)
            assert all(a + b == 2 * g - 1 for a, b in zip(o, i))
            o, i = ghost_indices(Face.JMAX, g, n)
            assert all(a + b == 2 * n - 2 * g - 1 for a, b in zip(o, i))
