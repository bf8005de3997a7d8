"""CPU checks of the drop-in boundary: the built C-ABI library loads and
exports every symbol include/famrdp_capi.h declares (no compute calls), and
the host-side translation of topology/Config objects matches the reference's
visiting order and BC dispatch rules."""
import ctypes
import os
import re

import pytest

from famrdp_jax import _lib
from famrdp_jax._plan import (BC_EXTRAPOLATE, BC_FARFIELD, BC_MIRROR_NORMAL, BC_WALL_ADIABATIC,
                              BC_WALL_ISOTHERMAL, bcs_from_config, cuts_from_topology)
from famrdp_jax.core.constants import BCType, Face
from famrdp_jax.core.types import Block, BlockTopology, Metrics

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "famrdp_capi.h")


def _declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:const char\*|int)\s+(famrdp_\w+)\(", src, re.M)))


def test_header_symbols_exported_and_bound():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # the Python binding table covers exactly the declared API
    assert sorted(_lib.SIGNATURES) == names
    L = _lib.load()
    assert L.famrdp_version().startswith(b"famrdp-mi355x")
    assert L.famrdp_last_error() == b""


def test_variant_count_matches_header():
    import re
    hdr = open(HEADER).read()
    assert int(re.search(r"#define FAMRDP_NUM_TILE_VARIANTS (\d+)", hdr).group(1)) == _lib.NUM_TILE_VARIANTS


def test_plan_create_rejects_bad_block_sizes_before_allocating():
    """famrdp_plan_create validates every block before any HIP call: padded
    sizes below 2*ghost, blocks of 2^31 or more cells (the tile kernels
    index cells in 32 bits) and, in a plan with FAMRDP_PLAN_ADJOINT (2),
    blocks of 2^29 or more (the reverse kernels' 32-bit byte offsets) return
    FAMRDP_E_INVALID with a message."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    L = _lib.load()
    gas = _lib.Gas(1.4, 1.0 / 1.4, 1e-3, 1.0, 0.38, 0.72)
    for dims, flags, msg in (((3, 10, 10), 0, b"2*ghost"), ((2048, 1100, 1000), 0, b"2^31"),
                             ((1024, 1024, 512), 2, b"2^29 or more cells in an adjoint plan")):
        plan = _lib.vp()
        d = (_lib.i32 * 3)(*dims)
        rc = L.famrdp_plan_create(ctypes.byref(plan), 1, d, 0, None, 0, None, ctypes.byref(gas), 0, 2, flags)
        assert rc == _lib.E_INVALID and plan.value is None
        assert msg in L.famrdp_last_error()


def test_step_entry_points_refuse_null_plan():
    """The step entry points check their arguments before any HIP call."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    L = _lib.load()
    assert L.famrdp_step(None, None, None, None, 0.1, 0, None) == _lib.E_INVALID
    assert L.famrdp_step_stages(None, None, None, None, 0.1, 0, None, None, None) == _lib.E_INVALID
    assert b"step_stages" in L.famrdp_last_error()


def test_float64_guard():
    import numpy as np
    with pytest.raises(TypeError):
        Block(q=np.zeros((5, 2, 2, 2), dtype=np.float32), xyz=np.zeros((3, 2, 2, 2)))
    with pytest.raises(TypeError):
        Metrics(jac=np.ones((2, 2, 2)), kxyz=np.zeros((3, 3, 2, 2, 2), dtype=np.float32), vol=np.ones((2, 2, 2)))


def test_cut_order_follows_dict_order():
    t0 = BlockTopology(0, {Face.JMAX: (7, Face.JMIN, 0), Face.IMIN: (0, Face.IMAX, 0), Face.KMIN: None},
                       {Face.JMAX: BCType.CUT1TO1, Face.IMIN: BCType.CUT1TO1, Face.KMIN: BCType.WALL})
    t1 = BlockTopology(7, {Face.JMIN: (0, Face.JMAX, 0)}, {Face.JMIN: BCType.CUT1TO1})
    assert cuts_from_topology((t0, t1)) == [(0, 3, 1, 2, 0), (0, 0, 0, 1, 0), (1, 2, 0, 3, 0)]
    with pytest.raises(KeyError):
        cuts_from_topology((BlockTopology(0, {Face.IMIN: (9, Face.IMAX, 0)}, {Face.IMIN: BCType.CUT1TO1}),))


def test_bc_dispatch_table():
    topo = BlockTopology(0, {}, {Face.KMAX: BCType.FARFIELD, Face.IMIN: BCType.WALL, Face.JMIN: BCType.SYMMETRY,
                                 Face.IMAX: BCType.OUTFLOW, Face.JMAX: BCType.INFLOW, Face.KMIN: BCType.WALL})
    params = {Face.KMIN: {"wall_type": "isothermal", "T_wall": 2.0}, Face.KMAX: {"q_inf": [1, 2, 3, 4, 5]}}
    got = bcs_from_config((topo,), params)
    # Face order, INFLOW / plain OUTFLOW skipped
    assert [(b, f, k) for (b, f, k, _, _) in got] == [(0, 0, BC_WALL_ADIABATIC), (0, 2, BC_MIRROR_NORMAL),
                                                      (0, 4, BC_WALL_ISOTHERMAL), (0, 5, BC_FARFIELD)]
    assert got[2][3] == 2.0 and got[3][4] == (1.0, 2.0, 3.0, 4.0, 5.0)
    got = bcs_from_config((topo,), {Face.IMAX: {"extrapolate": True}})
    assert (0, 1, BC_EXTRAPOLATE) in [(b, f, k) for (b, f, k, _, _) in got]


def test_product_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from famrdp_jax.physics.eos import cons_to_prim
    import numpy as np
    with pytest.raises(RuntimeError):
        cons_to_prim(np.ones((5, 3)), 1.4)


def test_jax_bridge_gated():
    """JAX is optional: the bridge reports whether it can run and refuses
    loudly when it cannot."""
    from famrdp_jax.interop import jax_bridge
    try:
        import jax  # noqa: F401
        have = True
    except ImportError:
        have = False
    assert jax_bridge.available() == have
    if not have:
        with pytest.raises(ImportError):
            jax_bridge.make_steps(None, None, 1.0, 1)
