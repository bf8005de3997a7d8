"""GridGen .inp topology reader (famrdp_jax/mesh/io_top.py:1-196).

Text, blank lines ignored: a header line (ignored), the block count, then per
block its dimensions, a label line, the number of logical BC entries and the
entries ``is ie js je ks ke value``.  A value of -1 (CUT1TO1) is followed by a
second line giving the neighbour's face and, as its value, the neighbour's
1-based block number; other values are BCType codes.  The face of an entry is
the axis whose index range is degenerate (first of i, j, k in that order,
compared on absolute values) at its low end (index 1; along k also 0) or
high end; with no degenerate axis, the axis of smallest extent, at its low
end when the range starts in the lower half.  Later entries for a face
replace earlier ones; faces without an entry stay absent.  Interfaces are
recorded with orientation 1, as the reference does (io_top.py:179), which its
halo exchange rejects (halo.py:86-87).

Orientation normalisation (not in the reference): with ``aligned_orient=0``
an interface whose two faces lie on the same axis, whose blocks have equal
extents along both tangential axes, and whose two index ranges are
identical along both tangential axes (same start, same end, same sign:
the neighbour's cells follow the block's own in the same order) is
recorded with orientation 0 instead -- exactly the copy the halo exchange
performs for orientation 0 (full-extent slabs, no permutation or reversal).
Any other interface keeps orientation 1 and is still rejected by the halo.
The default (``aligned_orient=None``) keeps the reference's behaviour.
"""
from __future__ import annotations

from ..core.constants import BCType, Face
from ..core.types import BlockTopology


def _face_of(rng, dims):
    a = [abs(v) for v in rng]
    lo = (Face.IMIN, Face.JMIN, Face.KMIN)
    hi = (Face.IMAX, Face.JMAX, Face.KMAX)
    for ax in range(3):
        s, e = a[2 * ax], a[2 * ax + 1]
        if s == e:
            # low end: index 1; along k also the placeholder 0 (io_top.py:55-70)
            return lo[ax] if (s == 1 or (ax == 2 and s == 0)) else hi[ax]
    ext = [abs(a[2 * ax + 1] - a[2 * ax]) for ax in range(3)]
    ax = ext.index(min(ext))
    return lo[ax] if a[2 * ax] <= dims[ax] // 2 else hi[ax]


def _aligned(face, row, nbr_face, row2, dims, nbr_dims) -> bool:
    """Same face axis, equal tangential block extents, identical tangential
    index ranges (signed: the same traversal)."""
    ax = int(face) // 2
    if int(nbr_face) // 2 != ax:
        return False
    for t in range(3):
        if t == ax:
            continue
        if dims[t] != nbr_dims[t] or row[2 * t:2 * t + 2] != row2[2 * t:2 * t + 2]:
            return False
    return True


def parse_topology(path, *, aligned_orient=None) -> list:
    """One BlockTopology per block, block ids 0-based in file order.
    ``aligned_orient=0`` records aligned interfaces with orientation 0 (see
    the module docstring); the default keeps orientation 1 everywhere."""
    if aligned_orient not in (None, 0):
        raise ValueError("aligned_orient must be None or 0")
    with open(path) as fh:
        rows = [ln.split() for ln in fh if ln.strip()]
    it = iter(rows[1:])                 # the first line is ignored
    nblocks = int(next(it)[0])
    parsed = []
    for bid in range(nblocks):
        dims = [int(v) for v in next(it)[:3]]
        next(it)                        # label
        nbc = int(next(it)[0])
        entries = []
        for _ in range(nbc):
            row = [int(v) for v in next(it)]
            face = _face_of(row[:6], dims)
            if row[6] == int(BCType.CUT1TO1):
                row2 = [int(v) for v in next(it)]
                entries.append((face, row, row2))
            else:
                entries.append((face, row, None))
        parsed.append((dims, entries))
    topos = []
    for bid, (dims, entries) in enumerate(parsed):
        bc_type, neighbors = {}, {}
        for face, row, row2 in entries:
            if row2 is not None:
                nb = row2[6] - 1
                nbr_face = _face_of(row2[:6], dims)
                orient = 1
                if aligned_orient is not None and 0 <= nb < nblocks and \
                        _aligned(face, row[:6], nbr_face, row2[:6], dims, parsed[nb][0]):
                    orient = aligned_orient
                neighbors[face] = (nb, nbr_face, orient)
                bc_type[face] = BCType.CUT1TO1
            else:
                neighbors[face] = None
                bc_type[face] = BCType(row[6])
        topos.append(BlockTopology(block_id=bid, neighbors=neighbors, bc_type=bc_type))
    return topos
