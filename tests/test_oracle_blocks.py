"""Pin the oracle's halo exchange, BCs, step and metrics (CPU only).

Known-answer tests restated from the reference's tests/ (cited), the
Fortran test3 Jacobian golden fixture, and brute-force per-cell checks of
the halo's dict-order / pre-exchange semantics.
"""
import os

import numpy as np
import pytest

from famrdp_jax.core.constants import BCType, Face
from famrdp_jax.core.types import BlockTopology
from oracle import blocks as B
from oracle import metric as M
from oracle import numerics as N

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _q_const(ni=8, nj=4, nk=4, rho=1.0, u=2.0, v=3.0, w=4.0, p=1.0, g=1.4):
    q = np.zeros((5, ni, nj, nk))
    q[0], q[1], q[2], q[3] = rho, rho * u, rho * v, rho * w
    q[4] = p / (g - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return q


# ---- halo (reference tests/mesh/test_halo.py:10-46) -------------------------

def test_halo_two_blocks_known_values():
    g = 2
    shape = (5, 8, 6, 4)
    q0 = np.zeros(shape); q0[:, 2:6] = np.arange(1.0, 5.0)[None, :, None, None]
    q1 = np.zeros(shape); q1[:, 2:6] = np.arange(10.0, 14.0)[None, :, None, None]
    t0 = BlockTopology(0, {**{f: None for f in Face}, Face.IMAX: (1, Face.IMIN, 0)}, {Face.IMAX: BCType.CUT1TO1})
    t1 = BlockTopology(1, {**{f: None for f in Face}, Face.IMIN: (0, Face.IMAX, 0)}, {Face.IMIN: BCType.CUT1TO1})
    out = B.halo_exchange([q0, q1], [t0, t1], g)
    assert out[0][0, -2, 0, 0] == 10.0 and out[0][0, -1, 0, 0] == 11.0
    assert out[1][0, 0, 0, 0] == 3.0 and out[1][0, 1, 0, 0] == 4.0


def periodic_topology(block_id=0):
    pairs = {Face.IMIN: Face.IMAX, Face.IMAX: Face.IMIN, Face.JMIN: Face.JMAX,
             Face.JMAX: Face.JMIN, Face.KMIN: Face.KMAX, Face.KMAX: Face.KMIN}
    return BlockTopology(block_id, {f: (block_id, pairs[f], 0) for f in Face},
                         {f: BCType.CUT1TO1 for f in Face})


def test_halo_periodic_pre_exchange_semantics_bruteforce():
    """Distinct values everywhere; every ghost cell must equal the value the
    LAST copy (dict order) covering it reads from the ORIGINAL array."""
    shape = (5, 8, 7, 6)
    q = np.arange(np.prod(shape), dtype=float).reshape(shape)
    topo = periodic_topology()
    out = B.halo_exchange([q], [topo], 2)[0]
    n = shape[1:]
    expect = q.copy()
    for idx in np.ndindex(*n):
        src = None
        for face, (_, sface, _) in topo.neighbors.items():
            ax = B.FACE_AXIS[int(face)] - 1
            lo = int(face) % 2 == 0
            in_ghost = idx[ax] < 2 if lo else idx[ax] >= n[ax] - 2
            if in_ghost:
                s = list(idx)
                s[ax] = idx[ax] + (n[ax] - 4) if lo else idx[ax] - (n[ax] - 4)
                src = tuple(s)         # later faces overwrite earlier ones
        if src is not None:
            expect[(slice(None),) + idx] = q[(slice(None),) + src]
    np.testing.assert_array_equal(out, expect)


def test_halo_rejects_orient():
    t = BlockTopology(0, {Face.IMIN: (0, Face.IMAX, 1)}, {Face.IMIN: BCType.CUT1TO1})
    with pytest.raises(NotImplementedError):
        B.halo_exchange([np.zeros((5, 8, 4, 4))], [t], 2)


# ---- BCs (reference tests/physics/test_bc_*.py) -----------------------------

def test_wall_adiabatic_imin_imax():
    q = np.zeros((5, 8, 4, 4)); q[0], q[1], q[2], q[3], q[4] = 1, 2, 3, 4, 10
    out = B.apply_wall_noslip_adiabatic(q, Face.IMIN, 2)
    for g in range(2):
        np.testing.assert_allclose(out[:4, g], np.array([1, -2, -3, -4.0])[:, None, None] * np.ones((4, 4)))
    out = B.apply_wall_noslip_adiabatic(q, Face.IMAX, 2)
    for g in (6, 7):
        np.testing.assert_allclose(out[:4, g], np.array([1, -2, -3, -4.0])[:, None, None] * np.ones((4, 4)))


def test_wall_isothermal_density():
    q = _q_const()
    out = B.apply_wall_noslip_isothermal(q, Face.IMIN, 2, 300.0, 1.4, 287.058)
    assert out[0, 0, 0, 0] == pytest.approx(1.0 / (287.058 * 300.0), rel=1e-12)
    # momentum reversed and scaled by rho_g/rho_i, energy from p_g and ke_g
    rg = 1.0 / (287.058 * 300.0)
    assert out[1, 0, 0, 0] == pytest.approx(-2.0 * rg, rel=1e-12)
    assert out[4, 1, 0, 0] == pytest.approx(1.0 / 0.4 + 0.5 * rg * (4 + 9 + 16), rel=1e-12)


def test_slip_and_symmetry():
    q = _q_const(p=5.0)
    out = B.apply_wall_slip(q, Face.IMIN, 2)
    assert out[1, 0, 0, 0] == -2.0 and out[2, 0, 0, 0] == 3.0
    out = B.apply_symmetry_plane(q, Face.JMIN, 2)
    assert out[2, 0, 0, 0] == -3.0 and out[1, 0, 0, 0] == 2.0
    out = B.apply_wall_slip(q, Face.KMIN, 2)
    assert out[3, 0, 0, 0] == -4.0 and out[1, 0, 0, 0] == 2.0 and out[2, 0, 0, 0] == 3.0
    out = B.apply_symmetry_plane(q, Face.KMAX, 2)
    assert out[3, 0, 0, 3] == -4.0 and out[1, 0, 0, 3] == 2.0


def test_farfield():
    q = np.ones((5, 8, 4, 4)) * 99.0
    out = B.apply_farfield(q, Face.IMAX, 2, np.array([1.0, 0.5, 0, 0, 2.5]))
    assert out[0, 7, 0, 0] == 1.0 and out[1, 6, 0, 0] == 0.5
    out = B.apply_farfield(q, Face.IMIN, 2, np.array([1.2, 0.3, 0.1, 0, 3.0]))
    assert out[0, 0, 0, 0] == 1.2 and out[1, 1, 0, 0] == 0.3 and out[0, 2, 0, 0] == 99.0
    q = np.ones((5, 8, 4, 4)) * 5.0
    out = B.apply_farfield(q, Face.JMIN, 2, np.array([1.0, 0, 0, 0, 2.5]))
    assert (out[0, 0, 2:, 0] == 5.0).all()


def test_dispatch_known_values():
    gas = dict(gamma=1.4, R=287.058)
    q = _q_const()
    out = B.apply_bc(q, BlockTopology(0, {}, {Face.IMIN: BCType.WALL}), {}, 1.4, 287.058, 2)
    assert out[1, 0, 0, 0] == -2.0 and out[0, 1, 0, 0] == 1.0
    out = B.apply_bc(q, BlockTopology(0, {}, {Face.JMIN: BCType.SYMMETRY}), {}, 1.4, 287.058, 2)
    assert out[2, 0, 0, 0] == -3.0 and out[1, 0, 0, 0] == 2.0
    qi = [1.2, 0.5, 0.1, 0.0, 3.0]
    out = B.apply_bc(q, BlockTopology(0, {}, {Face.IMAX: BCType.FARFIELD}), {Face.IMAX: {"q_inf": qi}}, 1.4, 287.058, 2)
    assert out[0, 7, 0, 0] == 1.2 and out[1, 6, 0, 0] == 0.5
    out = B.apply_bc(q, BlockTopology(0, {}, {Face.IMIN: BCType.CUT1TO1}), {}, 1.4, 287.058, 2)
    np.testing.assert_array_equal(out, q)
    # OUTFLOW without the extension flag stays a no-op (dispatch.py:15-31)
    out = B.apply_bc(q, BlockTopology(0, {}, {Face.IMIN: BCType.OUTFLOW}), {}, 1.4, 287.058, 2)
    np.testing.assert_array_equal(out, q)


def test_bc_sequential_face_order_corners():
    """IMIN is applied before JMIN, so JMIN's mirror reads i-ghost cells that
    IMIN already wrote (in place, dispatch.py:15)."""
    rng = np.random.default_rng(7)
    q = rng.uniform(1, 2, (5, 8, 8, 6))
    topo = BlockTopology(0, {}, {Face.IMIN: BCType.WALL, Face.JMIN: BCType.SYMMETRY})
    out = B.apply_bc(q, topo, {}, 1.4, 287.058, 2)
    s_wall = np.array([1, -1, -1, -1, 1.0]); s_sym = np.array([1, 1, -1, 1, 1.0])
    # corner (i=0, j=0): JMIN mirror of (i=0, j=3), which IMIN set from (i=3, j=3)
    np.testing.assert_array_equal(out[:, 0, 0, 0], q[:, 3, 3, 0] * s_wall * s_sym)


# ---- step (reference tests/solver/test_step.py:7-42) -----------------------

def test_step_uniform_stays_uniform():
    g = 1.4
    q = _q_const(16, 12, 10, rho=1.225, u=50.0, v=0.0, w=0.0, p=101325.0)
    k = np.zeros((3, 3, 16, 12, 10))
    for d in range(3):
        k[d, d] = 1.0
    vol = np.ones((16, 12, 10))
    topo = BlockTopology(0, {f: None for f in Face}, {f: BCType.FARFIELD for f in Face})
    qinf = q[:, 0, 0, 0].copy()
    gas = dict(gamma=1.4, R=287.058, mu0=1.71608e-5, T0=273.15, Ts=110.4, Pr_lam=0.72)
    qs = [q]
    for _ in range(3):
        qs = B.step(qs, [(k, vol)], [topo], {f: {"q_inf": qinf} for f in Face}, gas, 1e-6)
    np.testing.assert_allclose(qs[0][:, 4:-4, 4:-4, 4:-4], q[:, 4:-4, 4:-4, 4:-4], rtol=1e-5, atol=1e-5)


# ---- metric (reference tests/mesh/test_metric.py, tests/validation/test_m11) -

def test_metric_uniform_cube():
    xs, ys, zs = np.linspace(0, 1, 20), np.linspace(0, 2, 20), np.linspace(0, 3, 20)
    xyz = np.stack(np.meshgrid(xs, ys, zs, indexing="ij"))
    jac, k, _ = M.compute_metrics(xyz)
    np.testing.assert_allclose(jac[4:-4, 4:-4, 4:-4], (1 / 19) * (2 / 19) * (3 / 19), rtol=1e-10)
    np.testing.assert_allclose(k[0, 0, 4:-4, 4:-4, 4:-4], 19.0, rtol=1e-9)
    np.testing.assert_allclose(k[1, 0, 4:-4, 4:-4, 4:-4], 0.0, atol=1e-9)


def test_metric_matches_fortran_test3_probe():
    d = np.load(os.path.join(GOLD, "test3_metric_subbox.npz"))
    for b in (1, 2):
        jac, _, _ = M.compute_metrics(d[f"xyz{b}"])
        # stencil inside the sub-box in i and j; k is the full block extent
        np.testing.assert_allclose(jac[2:22, 2:22, :], d[f"jac{b}"][2:22, 2:22, :], atol=1e-13, rtol=0)


def test_cfl_dt_closed_form():
    q = _q_const(8, 8, 8, rho=1.0, u=0.5, v=0.0, w=0.0, p=1 / 1.4)
    k = np.zeros((3, 3, 8, 8, 8))
    for d in range(3):
        k[d, d] = 1.0
    gas = dict(gamma=1.4, R=1 / 1.4, mu0=1e-3, T0=1.0, Ts=0.383, Pr_lam=0.72)
    dt = M.cfl_dt(q, k, np.ones((8, 8, 8)), gas, 0.5)
    lam = (0.5 + 1.0) + 1.0 + 1.0 + (1.4 / 0.72) * 1e-3 * 3
    assert dt == pytest.approx(0.5 / lam, rel=1e-14)


# ---- step composition (reference solver/step.py:32-53, time/rk.py:5-20) ------
#
# The oracle's ``step`` inlines the Shu-Osher stages over every block after one
# halo exchange and one BC pass.  Here it is checked against an independent
# composition of separately pinned pieces: ``halo_exchange`` and ``apply_bc``
# once, then ``numerics.rk3_step`` (pinned by its exact amplification
# polynomial) per block over ``rhs_block`` (pinned by the RHS tests).  This
# fixes the 0.75/0.25 and 1/3, 2/3 coefficients, the halo/BC-once order, and
# the advancement of ghost cells through the three stages.  The planted
# mistakes below show that a wrong composition fails the comparison.

from famrdp_jax import cases as CS  # noqa: E402


def _metrics_c4(xyz_ext):
    jac, kxyz, vol = M.compute_metrics(xyz_ext)
    m = CS.crop_metrics(jac, kxyz, vol)
    return m.kxyz, m.vol


def _composition_cases():
    st, cfg = CS.case_c1(n=12, seed=3, p_noise=1e-3)
    c1 = ([b.q for b in st.blocks], [(m.kxyz, m.vol) for m in st.metrics], cfg)
    q, xyz_ext, cfg4 = CS.case_c4(n=10)
    c4 = ([q], [_metrics_c4(xyz_ext)], cfg4)
    st3, cfg3 = CS.case_c3(nb_axis=2, lo=6, hi=10, seed=11, p_noise=1e-3)
    c3 = ([b.q for b in st3.blocks], [(m.kxyz, m.vol) for m in st3.metrics], cfg3)
    return {"c1_periodic": c1, "c4_wall_farfield": c4, "c3_2x2x2": c3}


_CASES = _composition_cases()


def _dt(qs, metrics, gas):
    return min(M.cfl_dt(q, m[0], m[1], gas, 0.5) for q, m in zip(qs, metrics))


def _bc_all(qs, cfg, gas):
    return [B.apply_bc(q, t, cfg.bc_params, gas["gamma"], gas["R"], 2) for q, t in zip(qs, cfg.topology)]


def _composed(qs, metrics, cfg, gas, dt, limiter="minmod"):
    qs = B.halo_exchange(qs, list(cfg.topology), 2)
    qs = _bc_all(qs, cfg, gas)
    return [N.rk3_step(q, lambda x, m=m: N.rhs_block(x, m[0], m[1], gas, limiter), dt) for q, m in zip(qs, metrics)]


@pytest.mark.parametrize("name", sorted(_CASES))
@pytest.mark.parametrize("limiter", ["minmod", "vanleer"])
def test_step_equals_halo_bc_once_then_rk3(name, limiter):
    qs, metrics, cfg = _CASES[name]
    gas = CS.gas_dict(cfg.gas)
    dt = _dt(qs, metrics, gas)
    got = B.step(qs, metrics, list(cfg.topology), cfg.bc_params, gas, dt, limiter=limiter)
    want = _composed(qs, metrics, cfg, gas, dt, limiter)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    # and the step moved the state by much more than rounding
    assert max(float(np.abs(g - q).max()) for g, q in zip(got, qs)) > 1e-6


def _mutant_refresh_each_stage(qs, metrics, cfg, gas, dt):
    """Planted mistake: halo + BCs refreshed before every stage."""
    topos = list(cfg.topology)
    fix = lambda z: _bc_all(B.halo_exchange(z, topos, 2), cfg, gas)
    L = lambda z: [N.rhs_block(q, m[0], m[1], gas) for q, m in zip(z, metrics)]
    q0 = fix(qs)
    q1 = fix([q + dt * r for q, r in zip(q0, L(q0))])
    q2 = fix([0.75 * q + 0.25 * (a + dt * r) for q, a, r in zip(q0, q1, L(q1))])
    return [(1.0 / 3.0) * q + (2.0 / 3.0) * (b + dt * r) for q, b, r in zip(q0, q2, L(q2))]


def _mutant_swapped_coefficients(qs, metrics, cfg, gas, dt):
    """Planted mistake: 0.25/0.75 instead of 0.75/0.25 in stage 2."""
    qs = _bc_all(B.halo_exchange(qs, list(cfg.topology), 2), cfg, gas)
    L = lambda z: [N.rhs_block(q, m[0], m[1], gas) for q, m in zip(z, metrics)]
    q1 = [q + dt * r for q, r in zip(qs, L(qs))]
    q2 = [0.25 * q + 0.75 * (a + dt * r) for q, a, r in zip(qs, q1, L(q1))]
    return [(1.0 / 3.0) * q + (2.0 / 3.0) * (b + dt * r) for q, b, r in zip(qs, q2, L(q2))]


def _mutant_frozen_ghosts(qs, metrics, cfg, gas, dt):
    """Planted mistake: ghost cells held at their post-BC values through the
    stages (only interior cells advance)."""
    q0 = _bc_all(B.halo_exchange(qs, list(cfg.topology), 2), cfg, gas)
    L = lambda z: [N.rhs_block(q, m[0], m[1], gas) for q, m in zip(z, metrics)]

    def keep(z):
        out = []
        for a, g in zip(z, q0):
            r = g.copy()
            r[:, 2:-2, 2:-2, 2:-2] = a[:, 2:-2, 2:-2, 2:-2]
            out.append(r)
        return out

    q1 = keep([q + dt * r for q, r in zip(q0, L(q0))])
    q2 = keep([0.75 * q + 0.25 * (a + dt * r) for q, a, r in zip(q0, q1, L(q1))])
    return keep([(1.0 / 3.0) * q + (2.0 / 3.0) * (b + dt * r) for q, b, r in zip(q0, q2, L(q2))])


@pytest.mark.parametrize("mutant", [_mutant_refresh_each_stage, _mutant_swapped_coefficients, _mutant_frozen_ghosts])
def test_step_composition_mutants_are_detected(mutant):
    qs, metrics, cfg = _CASES["c1_periodic"]
    gas = CS.gas_dict(cfg.gas)
    dt = _dt(qs, metrics, gas)
    ref = B.step(qs, metrics, list(cfg.topology), cfg.bc_params, gas, dt)
    bad = mutant(qs, metrics, cfg, gas, dt)
    # interior cells only: every mutant changes what the interior receives
    diff = max(float(np.abs(a - b)[:, 2:-2, 2:-2, 2:-2].max()) for a, b in zip(ref, bad))
    assert diff > 1e-10, f"{mutant.__name__} not detected ({diff:.3e})"
