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


def parse_topology(path) -> list:
    """One BlockTopology per block, block ids 0-based in file order."""
    with open(path) as fh:
        rows = [ln.split() for ln in fh if ln.strip()]
    it = iter(rows[1:])                 # the first line is ignored
    nblocks = int(next(it)[0])
    topos = []
    for bid in range(nblocks):
        dims = [int(v) for v in next(it)[:3]]
        next(it)                        # label
        nbc = int(next(it)[0])
        bc_type, neighbors = {}, {}
        for _ in range(nbc):
            row = [int(v) for v in next(it)]
            face = _face_of(row[:6], dims)
            if row[6] == int(BCType.CUT1TO1):
                row2 = [int(v) for v in next(it)]
                neighbors[face] = (row2[6] - 1, _face_of(row2[:6], dims), 1)
                bc_type[face] = BCType.CUT1TO1
            else:
                neighbors[face] = None
                bc_type[face] = BCType(row[6])
        topos.append(BlockTopology(block_id=bid, neighbors=neighbors, bc_type=bc_type))
    return topos
