"""Pin the oracle's pointwise numerics and RHS operators (CPU only).

Each check is independent of the formula it pins: the reference's own
known-answer tests (cited), closed forms, invariants, and brute-force loop
evaluations written from the mathematical definition rather than from the
oracle's slicing/rolling code.
"""
import math

import numpy as np
import pytest

from oracle import numerics as N

GAMMA = 1.4
GAS_STD = dict(gamma=1.4, R=287.058, mu0=1.71608e-5, T0=273.15, Ts=110.4, Pr_lam=0.72)


def _cons(rho, u, v, w, p, g=GAMMA):
    return np.array([rho, rho * u, rho * v, rho * w, p / (g - 1.0) + 0.5 * rho * (u * u + v * v + w * w)])


# ---- eos (reference tests/physics/test_eos.py:6-17) ------------------------

def test_eos_roundtrip_and_sound_speed():
    prim = np.array([1.225, 50.0, 10.0, -5.0, 101325.0])[:, None, None, None]
    back = N.cons_to_prim(N.prim_to_cons(prim, GAMMA), GAMMA)
    np.testing.assert_allclose(back, prim, rtol=1e-14)
    np.testing.assert_allclose(N.speed_of_sound(1.225, 101325.0, GAMMA), (1.4 * 101325.0 / 1.225) ** 0.5, rtol=1e-12)


def test_sutherland_closed_form():
    # mu(T0) = mu0 exactly; mu(T) ~ T^1.5 ratio by hand
    assert N.sutherland_viscosity(273.15, 1.7e-5, 273.15, 110.4) == pytest.approx(1.7e-5, rel=1e-15)
    T = 400.0
    expect = 1.7e-5 * math.pow(T / 273.15, 1.5) * (273.15 + 110.4) / (T + 110.4)
    assert N.sutherland_viscosity(T, 1.7e-5, 273.15, 110.4) == pytest.approx(expect, rel=1e-15)


# ---- limiters and MUSCL (reference tests/physics/test_muscl2.py) -----------

def test_minmod_known_values():
    np.testing.assert_allclose(N.minmod(np.array([1.0, -1.0]), np.array([-1.0, 1.0])), 0.0, atol=1e-15)
    assert float(N.minmod(np.array(2.0), np.array(5.0))) == 2.0
    assert float(N.minmod(np.array(-3.0), np.array(-7.0))) == -3.0
    assert float(N.minmod(np.array(0.0), np.array(7.0))) == 0.0


def test_vanleer_harmonic_mean():
    rng = np.random.default_rng(1)
    a = rng.uniform(-2, 2, 200); b = rng.uniform(-2, 2, 200)
    expect = np.array([2 * x * y / (x + y) if x * y > 0 else 0.0 for x, y in zip(a, b)])
    np.testing.assert_allclose(N.vanleer(a, b), expect, rtol=1e-15)


def _scalar_minmod(a, b):
    if a * b <= 0.0:
        return 0.0
    return math.copysign(min(abs(a), abs(b)), a)


def test_muscl_linear_profile_and_bruteforce():
    pv = np.stack([np.linspace(1.0, 10.0, 10)] * 5, axis=0)[:, :, None, None]
    L, R = N.muscl2_reconstruct(pv, axis=1, limiter="minmod")
    np.testing.assert_allclose(L[0, :, 0, 0], R[0, :, 0, 0], rtol=1e-12)
    # brute force: face f between cells f+1 and f+2 along axis 2
    rng = np.random.default_rng(2)
    pv = rng.normal(size=(5, 3, 9, 4))
    L, R = N.muscl2_reconstruct(pv, axis=2, limiter="minmod")
    for v in range(5):
        for i in range(3):
            for k in range(4):
                col = pv[v, i, :, k]
                for f in range(9 - 3):
                    a = f + 1
                    sL = _scalar_minmod(col[a] - col[a - 1], col[a + 1] - col[a])
                    sR = _scalar_minmod(col[a + 1] - col[a], col[a + 2] - col[a + 1])
                    assert L[v, i, f, k] == pytest.approx(col[a] + 0.5 * sL, abs=1e-15)
                    assert R[v, i, f, k] == pytest.approx(col[a + 1] - 0.5 * sR, abs=1e-15)


# ---- Roe flux (reference tests/flux/test_roe.py:5-26) ----------------------

def _phys(U, n, g=GAMMA):
    rho, u, v, w = U[0], U[1] / U[0], U[2] / U[0], U[3] / U[0]
    p = (g - 1) * (U[4] - 0.5 * rho * (u * u + v * v + w * w))
    un = u * n[0] + v * n[1] + w * n[2]
    return np.array([rho * un, rho * u * un + p * n[0], rho * v * un + p * n[1], rho * w * un + p * n[2],
                     (U[4] + p) * un])


@pytest.mark.parametrize("prim,n", [((1.0, 50.0, 0.0, 0.0, 101325.0), (1.0, 0.0, 0.0)),
                                    ((1.0, 30.0, 20.0, 10.0, 50000.0), (0.6, 0.8, 0.0))])
def test_roe_identical_states_is_physical_flux(prim, n):
    U = _cons(*prim)
    F = N.roe_flux(U, U, np.array(n), gamma=GAMMA)
    np.testing.assert_allclose(F, _phys(U, n), rtol=1e-12)


def _unit(v):
    v = np.asarray(v, dtype=float)
    return v / np.linalg.norm(v)


def test_roe_supersonic_upwinding_pins_wave_decomposition():
    """With all Roe eigenvalues of one sign and |lambda| >= eps (no Harten),
    F = 1/2(FL+FR) - 1/2 sum|lambda| alpha R reduces to F(UL) (or F(UR)) iff
    sum lambda alpha R = F(UR) - F(UL).  That identity holds for this
    flux's wave strengths only when rhoL*rhoR = 1 (its dun carries rho_hat),
    so the check runs at rhoL*rhoR = 1; it fails if any wave term is wrong."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        n = _unit(rng.normal(size=3))
        rhoL = rng.uniform(0.5, 2.0); rhoR = 1.0 / rhoL
        c = 1.0
        base = n * rng.uniform(2.5, 4.0) * c + rng.normal(size=3) * 0.1    # supersonic along +n
        UL = _cons(rhoL, *(base + rng.normal(size=3) * 0.02), 1.0 / 1.4 * rng.uniform(0.9, 1.1))
        UR = _cons(rhoR, *(base + rng.normal(size=3) * 0.02), 1.0 / 1.4 * rng.uniform(0.9, 1.1))
        F = N.roe_flux(UL, UR, n, gamma=GAMMA)
        np.testing.assert_allclose(F, _phys(UL, n), rtol=1e-12, atol=1e-12 * np.abs(_phys(UL, n)).max())
        F = N.roe_flux(UL, UR, -n, gamma=GAMMA)   # now supersonic against the normal
        np.testing.assert_allclose(F, _phys(UR, -n), rtol=1e-12, atol=1e-12 * np.abs(_phys(UR, n)).max())


def test_roe_rotational_invariance():
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    for _ in range(20):
        pl = rng.uniform(0.5, 1.5, 5); pr = rng.uniform(0.5, 1.5, 5)
        n = _unit(rng.normal(size=3))
        UL = _cons(pl[0], *(pl[1:4] - 1.0), pl[4]); UR = _cons(pr[0], *(pr[1:4] - 1.0), pr[4])
        F = N.roe_flux(UL, UR, n, gamma=GAMMA)
        rot = lambda U: np.concatenate([[U[0]], Q @ U[1:4], [U[4]]])
        Fr = N.roe_flux(rot(UL), rot(UR), Q @ n, gamma=GAMMA)
        np.testing.assert_allclose(Fr, rot(F), rtol=1e-12, atol=1e-13)


def test_roe_harten_branch_is_continuous():
    """The entropy fix 1/2(lam^2/eps + eps) meets |lam| at |lam| = eps; sweep
    the normal velocity through the sonic point and require no jump."""
    n = np.array([1.0, 0.0, 0.0])
    UR = _cons(1.3, 0.9, 0.0, 0.0, 1.0)
    us = np.linspace(0.3, 1.7, 14001)          # crosses |lambda1| = eps on both sides
    F = np.array([N.roe_flux(_cons(1.0, u, 0.0, 0.0, 1.0 / 1.4), UR, n, gamma=GAMMA) for u in us])[:, [0, 1, 4]]
    jumps = np.abs(np.diff(F, axis=0))
    # a dropped 1/2 in the fix gives a jump ~700x the median step; smooth is < 2x
    assert (jumps.max(axis=0) / np.median(jumps, axis=0)).max() < 5.0


# ---- inviscid RHS (reference tests/rhs/test_invscd.py:7-23) ----------------

def _identity_metrics(shape):
    k = np.zeros((3, 3) + shape)
    for d in range(3):
        k[d, d] = 1.0
    return k, np.ones(shape)


def test_invscd_uniform_flow_zero():
    shape = (16, 12, 10)
    q = np.broadcast_to(_cons(1.225, 50.0, 0.0, 0.0, 101325.0)[:, None, None, None], (5,) + shape).copy()
    k, vol = _identity_metrics(shape)
    rhs = N.compute_invscd_rhs(q, k, vol, gamma=GAMMA)
    np.testing.assert_allclose(rhs[:, 4:-4, 4:-4, 4:-4], 0.0, atol=1e-8)


def _random_state(shape, rng):
    rho = rng.uniform(0.8, 1.2, shape)
    u, v, w = (rng.uniform(-0.3, 0.3, shape) for _ in range(3))
    p = rng.uniform(0.6, 0.8, shape)
    return np.stack([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)])


def _random_metrics(shape, rng):
    k = np.zeros((3, 3) + shape)
    for d in range(3):
        k[d, d] = 1.0
    k = k + 0.2 * rng.normal(size=k.shape)
    vol = rng.uniform(0.5, 1.5, shape)
    return k, vol


def test_invscd_bruteforce_per_cell():
    """Per-cell assembly from the docstring semantics: cells 2..n-3 along each
    axis, faces c-1/2 and c+1/2 with averaged metrics, scalar minmod."""
    rng = np.random.default_rng(5)
    shape = (7, 6, 6)
    q = _random_state(shape, rng)
    k, vol = _random_metrics(shape, rng)
    rhs = N.compute_invscd_rhs(q, k, vol, gamma=GAMMA)
    pv = N.cons_to_prim(q, GAMMA)
    expect = np.zeros_like(q)

    def face_flux(idx, d, a):          # face between cells a and a+1 along d
        def cell(c):
            j = list(idx); j[d] = c
            return tuple(j)
        col = [pv[(slice(None),) + cell(c)] for c in (a - 1, a, a + 1, a + 2)]
        sl = np.array([_scalar_minmod(col[1][v] - col[0][v], col[2][v] - col[1][v]) for v in range(5)])
        sr = np.array([_scalar_minmod(col[2][v] - col[1][v], col[3][v] - col[2][v]) for v in range(5)])
        pL = col[1] + 0.5 * sl; pR = col[2] - 0.5 * sr
        nf = 0.5 * (k[(d, slice(None)) + cell(a)] + k[(d, slice(None)) + cell(a + 1)])
        nm = np.sqrt(np.sum(nf ** 2)) + 1e-300
        F = N.roe_flux(N.prim_to_cons(pL, GAMMA), N.prim_to_cons(pR, GAMMA), nf / nm, gamma=GAMMA)
        return F * nm

    for idx in np.ndindex(*shape):
        for d in range(3):
            c = idx[d]
            if 2 <= c <= shape[d] - 3:
                expect[(slice(None),) + idx] += -(face_flux(idx, d, c) - face_flux(idx, d, c - 1)) / (vol[idx] + 1e-300)
    np.testing.assert_allclose(rhs, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())


# ---- viscous RHS (reference tests/rhs/test_viscous.py:7-20) ----------------

def test_viscous_static_isothermal_zero():
    shape = (16, 12, 10)
    q = np.zeros((5,) + shape); q[0] = 1.225; q[4] = 101325.0 / 0.4
    k, _ = _identity_metrics(shape)
    rhs = N.compute_viscous_rhs(q, k, **GAS_STD)
    np.testing.assert_allclose(rhs[:, 4:-4, 4:-4, 4:-4], 0.0, atol=1e-10)


def test_viscous_linear_shear_closed_form():
    """u = a*y on unit Cartesian metrics, uniform rho and p: tau_xy = mu*a is
    constant so the momentum RHS is 0 and the energy RHS is mu*a^2 (viscous
    heating d(u tau_xy)/dy) away from the roll wrap."""
    shape = (8, 12, 8); a = 0.01
    y = np.arange(shape[1], dtype=float)[None, :, None] * np.ones(shape)
    rho = 1.0; p = 1.0 / 1.4
    u = a * y
    q = np.stack([np.full(shape, rho), rho * u, np.zeros(shape), np.zeros(shape), p / 0.4 + 0.5 * rho * u * u])
    k, _ = _identity_metrics(shape)
    gas = dict(gamma=1.4, R=1 / 1.4, mu0=0.01, T0=1.0, Ts=0.383, Pr_lam=0.72)
    rhs = N.compute_viscous_rhs(q, k, **gas)
    mu = 0.01  # T = p/(rho R) = 1 = T0
    inner = (slice(None), slice(2, -2), slice(2, -2), slice(2, -2))
    np.testing.assert_allclose(rhs[inner][:4], 0.0, atol=1e-15)
    np.testing.assert_allclose(rhs[inner][4], mu * a * a, rtol=1e-10)


def test_viscous_bruteforce_modular_loops():
    rng = np.random.default_rng(6)
    shape = (5, 6, 4)
    q = _random_state(shape, rng)
    k, _ = _random_metrics(shape, rng)
    gas = dict(gamma=1.4, R=1 / 1.4, mu0=0.02, T0=1.0, Ts=0.383, Pr_lam=0.72)
    rhs = N.compute_viscous_rhs(q, k, **gas)
    pv = N.cons_to_prim(q, 1.4)
    T = pv[4] / (pv[0] * gas["R"])
    mu = gas["mu0"] * (T / gas["T0"]) ** 1.5 * (gas["T0"] + gas["Ts"]) / (T + gas["Ts"])
    kap = 1.4 * gas["R"] / 0.4 * mu / gas["Pr_lam"]

    def nb(idx, l, s):
        j = list(idx); j[l] = (j[l] + s) % shape[l]
        return tuple(j)

    def grad(f, idx):
        return np.array([sum(0.5 * (f[nb(idx, l, 1)] - f[nb(idx, l, -1)]) * k[(l, m) + idx] for l in range(3))
                         for m in range(3)])

    G = np.zeros((4, 3) + shape)  # rows: x-,y-,z-momentum, energy; columns j
    for idx in np.ndindex(*shape):
        gu = [grad(pv[1 + i], idx) for i in range(3)]
        gT = grad(T, idx)
        dv = gu[0][0] + gu[1][1] + gu[2][2]
        tau = np.array([[mu[idx] * (gu[i][j] + gu[j][i]) - (2.0 / 3.0) * mu[idx] * dv * (i == j)
                         for j in range(3)] for i in range(3)])
        vel = np.array([pv[1][idx], pv[2][idx], pv[3][idx]])
        G[:3, :, idx[0], idx[1], idx[2]] = tau
        G[3, :, idx[0], idx[1], idx[2]] = vel @ tau + kap[idx] * gT
    expect = np.zeros_like(q)
    for idx in np.ndindex(*shape):
        for r in range(4):
            expect[(r + 1,) + idx] = sum(0.5 * (G[(r, j) + nb(idx, l, 1)] - G[(r, j) + nb(idx, l, -1)]) * k[(l, j) + idx]
                                         for j in range(3) for l in range(3))
    np.testing.assert_allclose(rhs, expect, rtol=1e-11, atol=1e-13 * np.abs(expect).max())


# ---- RK3 (reference tests/time/test_rk.py:6-13, strengthened) --------------

def test_rk3_exact_amplification():
    for dt in (0.1, 0.37, 1.3):
        q1 = N.rk3_step(np.array([1.0]), lambda q: -q, dt)
        assert q1[0] == pytest.approx(1 - dt + dt * dt / 2 - dt ** 3 / 6, rel=1e-14)
    assert abs(N.rk3_step(np.array([1.0]), lambda q: -q, 0.1)[0] - math.exp(-0.1)) < 1e-5
