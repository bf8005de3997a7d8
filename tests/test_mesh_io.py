"""Grid and topology readers and the initial State (reference
tests/mesh/test_io_grd.py, test_io_top.py, test_build_state.py) on the
reference's own fixtures (tests/golden/mini.grd, mini.inp: two blocks,
(4,3,2) and (3,3,2) grids; 5x5x3 topology with a j-face interface), plus
known values derived by hand from the files."""
import os

import numpy as np
import pytest

from famrdp_jax.core.constants import BCType, Face
from famrdp_jax.mesh.build_state import build_initial_state
from famrdp_jax.mesh.io_grd import read_grd
from famrdp_jax.mesh.io_top import parse_topology

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GRD = os.path.join(GOLD, "mini.grd")
INP = os.path.join(GOLD, "mini.inp")


def test_read_grd_shapes_and_values():
    blocks = read_grd(GRD)
    assert [b.shape for b in blocks] == [(3, 4, 3, 2), (3, 3, 3, 2)]
    assert all(b.dtype == np.float64 and b.flags.c_contiguous for b in blocks)
    for b in blocks:
        ni, nj, nk = b.shape[1:]
        X, Y, Z = np.meshgrid(np.linspace(0, 1, ni), np.linspace(0, 1, nj), np.linspace(0, 1, nk), indexing="ij")
        np.testing.assert_allclose(b[0], X, atol=1e-14)
        np.testing.assert_allclose(b[1], Y, atol=1e-14)
        np.testing.assert_allclose(b[2], Z, atol=1e-14)


def test_read_grd_fortran_order_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    blocks = [rng.normal(size=(3, 5, 4, 3)), rng.normal(size=(3, 2, 6, 7))]
    p = tmp_path / "g.grd"
    with open(p, "wb") as f:
        np.array([len(blocks)], "<i4").tofile(f)
        for b in blocks:
            np.array(b.shape[1:], "<i4").tofile(f)
        for b in blocks:
            for c in range(3):
                b[c].astype("<f8").flatten(order="F").tofile(f)
    for a, b in zip(read_grd(p), blocks):
        np.testing.assert_array_equal(a, b)
    with open(p, "rb") as f:
        head = f.read(200)
    (tmp_path / "t.grd").write_bytes(head)
    with pytest.raises(ValueError):
        read_grd(tmp_path / "t.grd")


def test_parse_topology_known_faces():
    t0, t1 = parse_topology(INP)
    assert (t0.block_id, t1.block_id) == (0, 1)
    for t in (t0, t1):
        assert t.bc_type[Face.KMIN] == BCType.SYMMETRY and t.bc_type[Face.KMAX] == BCType.SYMMETRY
        assert t.bc_type[Face.IMIN] == BCType.WALL and t.bc_type[Face.IMAX] == BCType.FARFIELD
        assert t.bc_type[Face.JMIN] == BCType.CUT1TO1 and t.bc_type[Face.JMAX] == BCType.CUT1TO1
        assert all(t.neighbors[f] is None for f in (Face.IMIN, Face.IMAX, Face.KMIN, Face.KMAX))
        assert all(isinstance(v, BCType) for v in t.bc_type.values()) and len(t.bc_type) == 6
    assert t0.neighbors[Face.JMIN] == (1, Face.JMAX, 1) and t0.neighbors[Face.JMAX] == (1, Face.JMIN, 1)
    assert t1.neighbors[Face.JMIN] == (0, Face.JMAX, 1) and t1.neighbors[Face.JMAX] == (0, Face.JMIN, 1)


def test_parse_topology_face_rules(tmp_path):
    """Degenerate-axis rule with placeholders, j placeholder 0 -> JMAX,
    k placeholder 0 -> KMIN, smallest-extent fallback, later entries win,
    unknown BC codes raise ValueError."""
    txt = "\n".join([
        "1", "1", "8 8 8", "blk", "5",
        "1 8 0 0 1 8 2",        # j degenerate at 0: JMAX (only index 1 is low along j)
        "1 8 1 8 0 0 3",        # k degenerate at 0: KMIN
        "1 2 1 8 1 8 4",        # no degenerate axis; i extent smallest, starts low: IMIN
        "7 8 1 8 1 8 6",        # i extent smallest, starts high: IMAX
        "1 8 1 8 0 0 4",        # KMIN again: later entry wins
    ])
    p = tmp_path / "t.inp"
    p.write_text(txt)
    (t,) = parse_topology(p)
    assert t.bc_type == {Face.JMAX: BCType.WALL, Face.KMIN: BCType.FARFIELD, Face.IMIN: BCType.FARFIELD,
                         Face.IMAX: BCType.OUTFLOW}
    p.write_text(txt.replace("1 8 0 0 1 8 2", "1 8 0 0 1 8 9"))
    with pytest.raises(ValueError):
        parse_topology(p)


def test_build_initial_state():
    state, topos = build_initial_state(GRD, INP, ghost=2, q_init_fn=lambda shape: np.zeros(shape))
    assert len(state.blocks) == 2 and len(topos) == 2
    assert state.blocks[0].xyz.shape == (3, 8, 7, 6) and state.blocks[0].q.shape == (5, 8, 7, 6)
    assert state.blocks[1].q.shape == (5, 7, 7, 6)
    xyz = read_grd(GRD)[0]
    np.testing.assert_array_equal(state.blocks[0].xyz[:, 2:-2, 2:-2, 2:-2], xyz)
    np.testing.assert_array_equal(state.blocks[0].xyz[:, 0, 2:-2, 2:-2], xyz[:, 0])   # edge replication
    m = state.metrics[0]
    assert m.kxyz.shape == (3, 3, 8, 7, 6) and (m.kxyz == 0).all() and (m.vol == 1).all() and (m.jac == 1).all()
    assert state.t == 0.0 and state.step == 0


def test_file_interfaces_carry_orientation_one():
    """Interfaces read from files carry orient = 1, which the halo exchange
    rejects (reference halo.py:86-87); the host side passes it through and
    the plan refuses it."""
    from famrdp_jax._plan import cuts_from_topology
    cuts = cuts_from_topology(parse_topology(INP))
    assert cuts and all(c[4] == 1 for c in cuts)


def test_aligned_orient_normalisation(tmp_path):
    """parse_topology(..., aligned_orient=0): the reference fixture's four
    j-interfaces (same axis, equal tangential extents, identical signed
    tangential ranges) become orientation 0; a reversed tangential range,
    unequal tangential extents or different face axes keep orientation 1."""
    for t in parse_topology(INP, aligned_orient=0):
        assert t.neighbors[Face.JMIN][2] == 0 and t.neighbors[Face.JMAX][2] == 0
        assert t.neighbors[Face.JMIN][:2] == (1 - t.block_id, Face.JMAX)
    txt = open(INP).read()
    lines = txt.split("\n")
    # block A's JMIN pair: reverse the neighbour's k range
    i = lines.index("       1       5        5        5       -3       -1        2")
    rev = list(lines)
    rev[i] = "       1       5        5        5       -1       -3        2"
    p = tmp_path / "rev.inp"
    p.write_text("\n".join(rev))
    a, b = parse_topology(p, aligned_orient=0)
    assert a.neighbors[Face.JMIN][2] == 1 and a.neighbors[Face.JMAX][2] == 0 and b.neighbors[Face.JMIN][2] == 0
    # unequal tangential extents (block B 6 points along i)
    ext = list(lines)
    j = [n for n, ln in enumerate(lines) if ln.strip() == "5      5      3"][1]
    ext[j] = "     6      5      3"
    p.write_text("\n".join(ext))
    assert all(t.neighbors[f][2] == 1 for t in parse_topology(p, aligned_orient=0) for f in (Face.JMIN, Face.JMAX))
    # a j-face glued to an i-face
    ax = list(lines)
    ax[i] = "       5       5        1        5       -3       -1        2"
    p.write_text("\n".join(ax))
    a, _ = parse_topology(p, aligned_orient=0)
    assert a.neighbors[Face.JMIN] == (1, Face.IMAX, 1)
    assert all(c[4] == 1 for c in __import__("famrdp_jax._plan", fromlist=["x"]).cuts_from_topology(
        parse_topology(INP)))
    with pytest.raises(ValueError):
        parse_topology(INP, aligned_orient=2)


def test_file_built_case_steps_on_the_oracle(tmp_path):
    """The file-built two-block case (mini.inp topology, orientation
    normalised) runs through the oracle's halo, BCs and three steps and stays
    finite; without the normalisation the halo rejects it, as the
    reference's does."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from _filecase import file_case
    from famrdp_jax import cases
    from oracle import blocks as OB
    from oracle import metric as OM
    state, cfg, dt = file_case(str(tmp_path), OM.compute_metrics)
    qs = [b.q for b in state.blocks]
    mets = [(m.kxyz, m.vol) for m in state.metrics]
    for _ in range(3):
        qs = OB.step(qs, mets, list(cfg.topology), cfg.bc_params, cases.gas_dict(cfg.gas), dt)
    assert all(np.isfinite(q).all() for q in qs)
    assert max(float(np.abs(a - b.q).max()) for a, b in zip(qs, state.blocks)) > 1e-6
    with pytest.raises(NotImplementedError):
        OB.halo_exchange([b.q for b in state.blocks], parse_topology(INP), 2)
