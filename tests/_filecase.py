"""A multi-block case built from files: the reference's topology fixture
(tests/golden/mini.inp: two 5x5x3 blocks glued j-periodically to each other,
wall / farfield / symmetry elsewhere) with a grid file written here at the
dimensions it declares (mini.grd's own blocks, (4,3,2) and (3,3,2) points,
do not fit that topology's interfaces)."""
import os

import numpy as np

from famrdp_jax import cases
from famrdp_jax.core.constants import Face
from famrdp_jax.core.types import Block, Config, Metrics, State
from famrdp_jax.mesh.build_state import build_initial_state

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INP = os.path.join(GOLD, "mini.inp")
DIMS = (5, 5, 3)
GHOST = 2


def mapping(i, j, k, b):
    """Smooth curvilinear coordinates of point (i, j, k) of block b."""
    x = 0.1 * i + 0.01 * np.sin(0.7 * j + 0.3 * k + b)
    y = 0.1 * (j + 5 * b) + 0.01 * np.sin(0.5 * i - 0.4 * k)
    z = 0.1 * k + 0.01 * np.sin(0.6 * i + 0.2 * j)
    return np.stack([x, y, z])


def write_grd(path, blocks):
    """GridGen binary .grd (mesh/io_grd.py): int32 block count, int32 dims per
    block, then per block x, y, z in Fortran order."""
    with open(path, "wb") as f:
        np.array([len(blocks)], "<i4").tofile(f)
        for b in blocks:
            np.array(b.shape[1:], "<i4").tofile(f)
        for b in blocks:
            for c in range(3):
                b[c].astype("<f8").flatten(order="F").tofile(f)


def _grid(b, lo, hi):
    idx = [np.arange(lo, n + hi, dtype=np.float64) for n in DIMS]
    I, J, K = np.meshgrid(*idx, indexing="ij")
    return mapping(I, J, K, b)


def file_case(tmp_dir, compute_metrics):
    """(State, Config, dt): the State from build_initial_state with aligned
    interfaces normalised to orientation 0, metrics from ``compute_metrics``
    on the coordinates extended by ghost + 2 layers of the same mapping (edge
    replication, the reference's padding, makes the Jacobian singular in the
    ghost layers), cropped to the padded block."""
    grd = os.path.join(tmp_dir, "twoblock.grd")
    write_grd(grd, [_grid(b, 0, 0) for b in range(2)])
    gamma, mach = 1.4, 0.3
    rng = np.random.default_rng(12)

    def q_init(shape):
        pert = lambda: rng.uniform(-1e-3, 1e-3, shape[1:])
        return cases.prim_to_cons_np(1.0 + pert(), mach + pert(), pert(), pert(), (1.0 + pert()) / gamma, gamma)

    state, topos = build_initial_state(grd, INP, GHOST, q_init, aligned_orient=0)
    metrics = []
    for b in range(2):
        jac, kxyz, vol = compute_metrics(_grid(b, -GHOST - 2, GHOST + 2))
        m = cases.crop_metrics(np.asarray(jac), np.asarray(kxyz), np.asarray(vol))
        metrics.append(m)
    state = State(blocks=state.blocks, metrics=tuple(metrics))
    gas = cases.gas_nondim(gamma=gamma, Pr=0.72, Re=500.0, U_ref=mach)
    qinf = cases.freestream(mach, gamma)
    bc_params = {Face.IMIN: {"wall_type": "adiabatic"}, Face.IMAX: {"q_inf": qinf}}
    cfg = Config(gas=gas, scheme=cases.SCHEME, bc_params=bc_params, topology=topos, ghost=GHOST)
    from oracle import metric as OM
    dt = min(OM.cfl_dt(b.q, m.kxyz, m.vol, cases.gas_dict(gas), 0.5) for b, m in zip(state.blocks, metrics))
    return state, cfg, dt
