"""Synthetic workloads for the parity tests and the bench (host-side numpy).

The reference defines no Reynolds number, CFL map, periodic BC or
extrapolation BC, so these builders fix them once (used identically by the
GPU path and the CPU oracle):

* non-dimensional gas: R = 1/gamma (T = 1 at rho = 1, p = 1/gamma), T0 = 1,
  Ts = 110.4/288.15, mu0 = rho_ref * U_ref * L_ref / Re with L_ref = 1;
* periodic = self-neighbour CUT1TO1 on every face, orient 0, Face order;
* curvilinear grid x_i = 2 pi (xi_i + 0.1 sin(2 pi xi_j) sin(2 pi xi_k)),
  (i, j, k) cyclic, xi = (index - ghost) / n, generated with two extra layers
  beyond the ghosts so the 4th-order metric stencil never wraps into ghost
  cells (``crop_metrics`` removes them).

Workloads:
  C1  32^3 Cartesian, periodic, M0.5 along x + Gaussian bump + noise, Re 1000
  C2  256^3 curvilinear TGV, M0.1, Re 1600, Pr 0.71
  C3  64 blocks (4x4x4 lattice, per-axis sizes U[16,96]), seeded gluing,
      farfield / extrapolation outer faces, M2 (seed 883: 19,996,944 cells)
  C4  96^3 curvilinear, isothermal wall at KMIN, farfield at KMAX, M0.8, Re 5000
  C5  32 periodic cases with sizes from {48,64,80}^3, M~U[0.3,0.9],
      Re~logU[500,5000]
"""
from __future__ import annotations

import math

import numpy as np

from .core.constants import BCType, Face
from .core.types import Block, BlockTopology, Config, GasModel, Metrics, SchemeChoice, State

GHOST = 2
TS_NONDIM = 110.4 / 288.15
SCHEME = SchemeChoice(reconstruct="muscl2", flux="roe", rk_order=3, limiter="minmod")
_PAIR = {Face.IMIN: Face.IMAX, Face.IMAX: Face.IMIN, Face.JMIN: Face.JMAX,
         Face.JMAX: Face.JMIN, Face.KMIN: Face.KMAX, Face.KMAX: Face.KMIN}


def gas_nondim(gamma=1.4, Pr=0.72, Re=1000.0, U_ref=1.0, rho_ref=1.0, L_ref=1.0) -> GasModel:
    return GasModel(gamma=gamma, R=1.0 / gamma, mu0=rho_ref * U_ref * L_ref / Re, T0=1.0, Ts=TS_NONDIM,
                    Pr_lam=Pr)


def gas_dict(gas: GasModel) -> dict:
    return dict(gamma=gas.gamma, R=gas.R, mu0=gas.mu0, T0=gas.T0, Ts=gas.Ts, Pr_lam=gas.Pr_lam)


def periodic_topology(block_id=0) -> BlockTopology:
    return BlockTopology(block_id=block_id, neighbors={f: (block_id, _PAIR[f], 0) for f in Face},
                         bc_type={f: BCType.CUT1TO1 for f in Face})


def prim_to_cons_np(rho, u, v, w, p, gamma):
    return np.stack([rho, rho * u, rho * v, rho * w, p / (gamma - 1.0) + rho * (0.5 * (u * u + v * v + w * w))])


def cartesian_metrics(shape) -> Metrics:
    kxyz = np.zeros((3, 3) + tuple(shape))
    for d in range(3):
        kxyz[d, d] = 1.0
    return Metrics(jac=np.ones(shape), kxyz=kxyz, vol=np.ones(shape))


def curvilinear_xyz(n, extra=2, ghost=GHOST, amp=0.1):
    """Coordinates of the padded block plus ``extra`` layers per side."""
    m = n + 2 * ghost + 2 * extra
    xi = (np.arange(m, dtype=np.float64) - ghost - extra) / n
    X1, X2, X3 = np.meshgrid(xi, xi, xi, indexing="ij")
    s = lambda a: np.sin(2.0 * math.pi * a)
    x = 2.0 * math.pi * (X1 + amp * s(X2) * s(X3))
    y = 2.0 * math.pi * (X2 + amp * s(X3) * s(X1))
    z = 2.0 * math.pi * (X3 + amp * s(X1) * s(X2))
    return np.stack([x, y, z])


def crop_metrics(jac, kxyz, vol, extra=2) -> Metrics:
    sl = (slice(extra, -extra),) * 3
    return Metrics(jac=np.ascontiguousarray(jac[sl]), kxyz=np.ascontiguousarray(kxyz[(slice(None), slice(None)) + sl]),
                   vol=np.ascontiguousarray(vol[sl]))


def crop_xyz(xyz_ext, extra=2):
    return np.ascontiguousarray(xyz_ext[(slice(None),) + (slice(extra, -extra),) * 3])


def _bump(shape_int, ghost, center, sigma, rng_shape=None):
    """Gaussian of periodic (minimum-image) distance in cells from ``center``."""
    n = np.asarray(shape_int)
    idx = [np.arange(s + 2 * ghost, dtype=np.float64) - ghost for s in shape_int]
    I, J, K = np.meshgrid(*idx, indexing="ij")
    r2 = np.zeros_like(I)
    for X, c, L in zip((I, J, K), center, n):
        d = np.abs(X - c)
        d = np.minimum(d, L - d)
        r2 = r2 + d * d
    return np.exp(-r2 / (2.0 * sigma * sigma))


def periodic_bump_state(shape_int, *, mach, gamma=1.4, amp=0.05, sigma=4.0, noise=1e-3, seed=0, p_noise=0.0):
    """Density bump plus seeded velocity noise; ``p_noise`` > 0 adds relative
    pressure noise (a uniform pressure puts every pressure slope of the MUSCL
    limiter at roundoff level, on the limiter's kink, where the derivative of
    the step is decided by roundoff)."""
    rng = np.random.default_rng(seed)
    center = rng.uniform(0, 1, 3) * np.asarray(shape_int)
    pad = tuple(s + 2 * GHOST for s in shape_int)
    rho = 1.0 + amp * _bump(shape_int, GHOST, center, sigma)
    u = mach + rng.uniform(-noise, noise, pad)
    v = rng.uniform(-noise, noise, pad)
    w = rng.uniform(-noise, noise, pad)
    p = np.full(pad, 1.0 / gamma)
    if p_noise:
        p = p * (1.0 + rng.uniform(-p_noise, p_noise, pad))
    return prim_to_cons_np(rho, u, v, w, p, gamma)


def _config(gas, topos, bc_params=None, cfl=0.5):
    return Config(gas=gas, scheme=SCHEME, bc_params=bc_params or {}, topology=tuple(topos), ghost=GHOST,
                  nan_check=False, cfl=cfl)


def case_c1(n=32, seed=0, p_noise=0.0):
    """Periodic Cartesian n^3 (C1 at n=32).  Returns (state, cfg)."""
    gamma = 1.4
    gas = gas_nondim(gamma=gamma, Pr=0.72, Re=1000.0, U_ref=0.5)
    q = periodic_bump_state((n, n, n), mach=0.5, gamma=gamma, seed=seed, p_noise=p_noise)
    shape = q.shape[1:]
    state = State(blocks=(Block(q=q, xyz=np.zeros((3,) + shape)),), metrics=(cartesian_metrics(shape),))
    return state, _config(gas, [periodic_topology()])


def tgv_state(xyz, *, mach=0.1, gamma=1.4):
    x, y, z = xyz
    V0, p0, rho0 = mach, 1.0 / gamma, 1.0
    u = V0 * np.sin(x) * np.cos(y) * np.cos(z)
    v = -V0 * np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(x)
    p = p0 + rho0 * V0 * V0 / 16.0 * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    rho = np.full_like(x, rho0)
    return prim_to_cons_np(rho, u, v, w, p, gamma)


def case_c2(n=256):
    """Curvilinear TGV (C2 at n=256).  Returns (q, xyz_ext, cfg); metrics come
    from compute_metrics on xyz_ext followed by crop_metrics."""
    gamma = 1.4
    gas = gas_nondim(gamma=gamma, Pr=0.71, Re=1600.0, U_ref=0.1)
    xyz_ext = curvilinear_xyz(n)
    q = tgv_state(crop_xyz(xyz_ext), mach=0.1, gamma=gamma)
    return q, xyz_ext, _config(gas, [periodic_topology()])


def case_c3(nb_axis=4, lo=16, hi=96, seed=883, p_glue=0.85, mach=2.0, p_noise=0.0):
    """Multi-block lattice (C3 at the defaults: 64 blocks).  Blocks (a,b,c)
    have interior sizes (NI[a], NJ[b], NK[c]) so glued neighbours share their
    tangential extents (the only shape-valid gluing with orient 0).  Returns
    (state, cfg) with unit Cartesian metrics."""
    rng = np.random.default_rng(seed)
    gamma = 1.4
    gas = gas_nondim(gamma=gamma, Pr=0.72, Re=1000.0, U_ref=mach)
    sizes = [rng.integers(lo, hi + 1, nb_axis) for _ in range(3)]
    bid = lambda a, b, c: (a * nb_axis + b) * nb_axis + c
    glue = {}
    for ax in range(3):
        for a in range(nb_axis):
            for b in range(nb_axis):
                for c in range(nb_axis):
                    idx = [a, b, c]
                    if idx[ax] + 1 < nb_axis:
                        nb = list(idx)
                        nb[ax] += 1
                        glue[(bid(*idx), ax)] = bid(*nb) if rng.uniform() < p_glue else None
    topos, blocks, metrics = [], [], []
    qinf = prim_to_cons_np(np.array(1.0), np.array(mach), np.array(0.0), np.array(0.0), np.array(1.0 / gamma), gamma)
    for a in range(nb_axis):
        for b in range(nb_axis):
            for c in range(nb_axis):
                me = bid(a, b, c)
                nbrs, bct = {}, {}
                idx = (a, b, c)
                for f in Face:
                    ax = int(f) // 2
                    lo_side = int(f) % 2 == 0
                    other = None
                    if lo_side and idx[ax] > 0:
                        prev = list(idx); prev[ax] -= 1
                        if glue.get((bid(*prev), ax)) == me:
                            other = bid(*prev)
                    elif not lo_side:
                        other = glue.get((me, ax))
                    if other is not None:
                        nbrs[f] = (other, _PAIR[f], 0)
                        bct[f] = BCType.CUT1TO1
                    else:
                        nbrs[f] = None
                        bct[f] = BCType.FARFIELD if (f == Face.IMIN and a == 0) else BCType.OUTFLOW
                topos.append(BlockTopology(block_id=me, neighbors=nbrs, bc_type=bct))
                shape_int = (int(sizes[0][a]), int(sizes[1][b]), int(sizes[2][c]))
                pad = tuple(s + 2 * GHOST for s in shape_int)
                prng = np.random.default_rng(seed * 1000 + me)
                pert = lambda: prng.uniform(-1e-3, 1e-3, pad)
                p = np.full(pad, 1.0 / gamma)
                if p_noise:
                    p = p * (1.0 + prng.uniform(-p_noise, p_noise, pad))
                q = prim_to_cons_np(1.0 + pert(), mach + pert(), pert(), pert(), p, gamma)
                blocks.append(Block(q=q, xyz=np.zeros((3,) + pad)))
                metrics.append(cartesian_metrics(pad))
    bc_params = {f: {"q_inf": qinf, "extrapolate": True} for f in Face}
    return State(blocks=tuple(blocks), metrics=tuple(metrics)), _config(gas, topos, bc_params)


def case_c4(n=96, mach=0.8, Re=5000.0, seed=4):
    """Wall-bounded curvilinear block (C4 at n=96): KMIN isothermal no-slip
    wall (T_wall = 1 = T_inf), KMAX farfield, periodic in i and j; seeded
    1e-3 perturbations of all five primitives (pressure included, so that the
    gradient of the step is not decided by roundoff at limiter kinks).
    Returns (q, xyz_ext, cfg)."""
    gamma = 1.4
    gas = gas_nondim(gamma=gamma, Pr=0.72, Re=Re, U_ref=mach)
    xyz_ext = curvilinear_xyz(n)
    pad = (n + 2 * GHOST,) * 3
    rng = np.random.default_rng(seed)
    pert = lambda: rng.uniform(-1e-3, 1e-3, pad)
    q = prim_to_cons_np(1.0 + pert(), mach + pert(), pert(), pert(), (1.0 + pert()) / gamma, gamma)
    qinf = freestream(mach, gamma)
    nbrs = {f: (0, _PAIR[f], 0) for f in (Face.IMIN, Face.IMAX, Face.JMIN, Face.JMAX)}
    nbrs[Face.KMIN] = None
    nbrs[Face.KMAX] = None
    bct = {f: BCType.CUT1TO1 for f in (Face.IMIN, Face.IMAX, Face.JMIN, Face.JMAX)}
    bct[Face.KMIN] = BCType.WALL
    bct[Face.KMAX] = BCType.FARFIELD
    bc_params = {Face.KMIN: {"wall_type": "isothermal", "T_wall": 1.0}, Face.KMAX: {"q_inf": qinf}}
    return q, xyz_ext, _config(gas, [BlockTopology(0, nbrs, bct)], bc_params)


def freestream(mach, gamma=1.4):
    return prim_to_cons_np(np.array(1.0), np.array(mach), np.array(0.0), np.array(0.0), np.array(1.0 / gamma), gamma)


def case_c5(ncases=32, seed=5, sizes=(48, 64, 80), p_noise=0.0):
    """Ensemble of independent periodic cases (C5).  Returns a list of
    (state, cfg) pairs, each with its own Mach and Reynolds number."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(ncases):
        shape = tuple(int(s) for s in rng.choice(sizes, 3))
        mach = float(rng.uniform(0.3, 0.9))
        Re = float(math.exp(rng.uniform(math.log(500.0), math.log(5000.0))))
        gas = gas_nondim(gamma=1.4, Pr=0.72, Re=Re, U_ref=mach)
        q = periodic_bump_state(shape, mach=mach, seed=seed * 1000 + c, p_noise=p_noise)
        pad = q.shape[1:]
        st = State(blocks=(Block(q=q, xyz=np.zeros((3,) + pad)),), metrics=(cartesian_metrics(pad),))
        out.append((st, _config(gas, [periodic_topology()])))
    return out


def ensemble(pairs):
    """Independent single-block cases as one multi-block configuration, for a
    batched plan: block c keeps its own topology (renumbered to c), BC
    parameters must agree between cases (they are keyed by face), and the gas
    of case 0 is shared except for mu0, returned per block (Plan
    set_block_params).  Returns (state, cfg, mu0 list)."""
    blocks, metrics, topos, mu0 = [], [], [], []
    gas0 = pairs[0][1].gas
    for c, (st, cfg) in enumerate(pairs):
        if len(st.blocks) != 1:
            raise ValueError("ensemble: every case must be a single block")
        g = cfg.gas
        if (g.gamma, g.R, g.T0, g.Ts, g.Pr_lam) != (gas0.gamma, gas0.R, gas0.T0, gas0.Ts, gas0.Pr_lam):
            raise ValueError("ensemble: cases may differ only in mu0")
        t = cfg.topology[0]
        nbrs = {f: (None if v is None else (c, v[1], v[2])) for f, v in t.neighbors.items()}
        topos.append(BlockTopology(block_id=c, neighbors=nbrs, bc_type=dict(t.bc_type)))
        blocks.append(st.blocks[0])
        metrics.append(st.metrics[0])
        mu0.append(g.mu0)
    cfg0 = pairs[0][1]
    cfg = Config(gas=gas0, scheme=cfg0.scheme, bc_params=cfg0.bc_params, topology=tuple(topos), ghost=cfg0.ghost,
                 nan_check=False, cfl=cfg0.cfl)
    return State(blocks=tuple(blocks), metrics=tuple(metrics)), cfg, mu0


def interior(a, ghost=GHOST):
    return a[(slice(None),) * (a.ndim - 3) + (slice(ghost, -ghost),) * 3]
