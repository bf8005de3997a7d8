"""GPU parity of the pointwise / operator entry points against the oracle.

Tolerances: elementwise formulas within a few ulp (FMA contraction on the
GPU, none in numpy), RHS operators within 1e-12 relative L-inf.
"""
import os

import numpy as np
import pytest

from famrdp_jax import cases
from famrdp_jax.core.constants import BCType, Face
from famrdp_jax.core.types import Block, BlockTopology, Config, GasModel, Metrics
from oracle import blocks as OB
from oracle import metric as OM
from oracle import numerics as ON
from _util import rel_linf

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
G = 1.4


def _state(shape, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.7, 1.3, shape)
    u, v, w = (rng.uniform(-0.6, 0.6, shape) for _ in range(3))
    p = rng.uniform(0.5, 0.9, shape)
    return np.stack([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)])


def _metrics(shape, seed):
    rng = np.random.default_rng(seed)
    k = np.zeros((3, 3) + shape)
    for d in range(3):
        k[d, d] = 1.0
    k += 0.25 * rng.normal(size=k.shape)
    return Metrics(jac=np.ones(shape), kxyz=k, vol=rng.uniform(0.5, 1.5, shape))


def test_eos_ops():
    print("case: eos", flush=True)
    from famrdp_jax.physics.eos import cons_to_prim, prim_to_cons, speed_of_sound, sutherland_viscosity
    q = _state((7, 6, 5), 1)
    np.testing.assert_allclose(cons_to_prim(q, G), ON.cons_to_prim(q, G), rtol=1e-14, atol=1e-15)
    pv = ON.cons_to_prim(q, G)
    np.testing.assert_allclose(prim_to_cons(pv, G), ON.prim_to_cons(pv, G), rtol=1e-14)
    T = np.linspace(0.5, 3.0, 101)
    np.testing.assert_allclose(sutherland_viscosity(T, 1e-3, 1.0, 0.38), ON.sutherland_viscosity(T, 1e-3, 1.0, 0.38),
                               rtol=1e-14)
    # reference tests/physics/test_eos.py:6-17
    prim = np.array([1.225, 50.0, 10.0, -5.0, 101325.0])[:, None, None, None]
    np.testing.assert_allclose(cons_to_prim(prim_to_cons(prim, G), G), prim, rtol=1e-14)
    np.testing.assert_allclose(speed_of_sound(np.array([1.225]), np.array([101325.0]), G),
                               (1.4 * 101325.0 / 1.225) ** 0.5, rtol=1e-12)


def test_limiters_and_muscl():
    print("case: limiters/muscl", flush=True)
    from famrdp_jax.physics.reconstruct.muscl2 import minmod, muscl2_reconstruct, vanleer
    rng = np.random.default_rng(2)
    a, b = rng.normal(size=500), rng.normal(size=500)
    a[:10] = 0.0; b[5:15] = a[5:15]
    np.testing.assert_array_equal(minmod(a, b), ON.minmod(a, b))
    np.testing.assert_allclose(vanleer(a, b), ON.vanleer(a, b), rtol=1e-15)
    pv = rng.normal(size=(5, 9, 7, 6))
    for axis in (1, 2, 3):
        for lim in ("minmod", "vanleer"):
            gL, gR = muscl2_reconstruct(pv, axis=axis, limiter=lim)
            rL, rR = ON.muscl2_reconstruct(pv, axis=axis, limiter=lim)
            np.testing.assert_allclose(gL, rL, rtol=1e-14, atol=1e-15)
            np.testing.assert_allclose(gR, rR, rtol=1e-14, atol=1e-15)
    with pytest.raises(KeyError):
        muscl2_reconstruct(pv, axis=1, limiter="superbee")
    # reference tests/physics/test_muscl2.py:14-22
    lin = np.stack([np.linspace(1.0, 10.0, 10)] * 5, axis=0)[:, :, None, None]
    gL, gR = muscl2_reconstruct(lin, axis=1, limiter="minmod")
    np.testing.assert_allclose(gL[0, :, 0, 0], gR[0, :, 0, 0], rtol=1e-12)


def test_roe_flux():
    print("case: roe", flush=True)
    from famrdp_jax.flux.roe import roe_flux
    rng = np.random.default_rng(3)
    UL = _state((4000,), 4); UR = _state((4000,), 5)
    n = rng.normal(size=(3, 4000)); n /= np.linalg.norm(n, axis=0)
    F = roe_flux(UL, UR, n, gamma=G)
    ref = ON.roe_flux(UL, UR, n, gamma=G)
    for v in range(5):
        assert np.abs(F[v] - ref[v]).max() <= 1e-13 * np.abs(ref[v]).max()
    # reference tests/flux/test_roe.py:5-26 (identical states = physical flux)
    rho, u, v, w, p = 1.0, 30.0, 20.0, 10.0, 50000.0
    U = np.array([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)])
    nn = np.array([0.6, 0.8, 0.0])
    un = u * 0.6 + v * 0.8
    expect = np.array([rho * un, rho * u * un + p * 0.6, rho * v * un + p * 0.8, rho * w * un, (U[4] + p) * un])
    np.testing.assert_allclose(roe_flux(U, U, nn, gamma=G), expect, rtol=1e-12)


def test_invscd_rhs():
    print("case: invscd", flush=True)
    from famrdp_jax.rhs.invscd import compute_invscd_rhs
    shape = (12, 10, 9)
    q = _state(shape, 6)
    m = _metrics(shape, 7)
    for lim in ("minmod", "vanleer"):
        got = compute_invscd_rhs(q, m, ghost=2, gamma=G, reconstruct="muscl2", flux="roe", limiter=lim)
        ref = ON.compute_invscd_rhs(q, m.kxyz, m.vol, gamma=G, limiter=lim)
        assert rel_linf(got, ref) <= 1e-12
    # reference tests/rhs/test_invscd.py:7-23
    st, _ = cases.case_c1(n=12)
    qu = np.broadcast_to(np.array([1.225, 1.225 * 50, 0, 0, 101325 / 0.4 + 0.5 * 1.225 * 2500])[:, None, None, None],
                         (5, 16, 12, 10)).copy()
    mu = cases.cartesian_metrics((16, 12, 10))
    got = compute_invscd_rhs(qu, mu, ghost=2, gamma=G, reconstruct="muscl2", flux="roe", limiter="minmod")
    np.testing.assert_allclose(got[:, 4:-4, 4:-4, 4:-4], 0.0, atol=1e-8)


def test_viscous_rhs():
    print("case: viscous", flush=True)
    from famrdp_jax.rhs.viscous import compute_viscous_rhs
    shape = (11, 9, 8)
    q = _state(shape, 8)
    m = _metrics(shape, 9)
    gas = GasModel(gamma=1.4, R=1 / 1.4, mu0=0.02, T0=1.0, Ts=0.383, Pr_lam=0.72)
    got = compute_viscous_rhs(q, m, gas=gas, ghost=2)
    ref = ON.compute_viscous_rhs(q, m.kxyz, **cases.gas_dict(gas))
    assert float(np.abs(got[0]).max()) == 0.0
    for v in range(1, 5):
        assert np.abs(got[v] - ref[v]).max() <= 1e-12 * np.abs(ref[v]).max()


def test_rk3_step_generic():
    print("case: rk3", flush=True)
    import torch
    from famrdp_jax.time.rk import rk3_step
    q0 = np.linspace(0.5, 1.5, 64)
    out = rk3_step(q0, lambda q: -q, 0.37)
    dt = 0.37
    np.testing.assert_allclose(out, q0 * (1 - dt + dt * dt / 2 - dt ** 3 / 6), rtol=1e-14)
    # the reference's keyword name and pytree states (time/rk.py:5-20)
    a, b = np.linspace(0.5, 1.5, 64), torch.linspace(-2.0, 3.0, 10, dtype=torch.float64, device="cuda")
    amp = 1 - dt + dt * dt / 2 - dt ** 3 / 6
    out = rk3_step((a, b), L=lambda q: (-q[0], -2.0 * q[1]), dt=dt)
    assert isinstance(out, tuple) and isinstance(out[0], np.ndarray) and isinstance(out[1], torch.Tensor)
    np.testing.assert_allclose(out[0], a * amp, rtol=1e-14)
    z = -2.0 * dt
    np.testing.assert_allclose(out[1].cpu().numpy(), b.cpu().numpy() * (1 + z + z * z / 2 + z ** 3 / 6), rtol=1e-14)
    out = rk3_step({"u": [a], "v": b}, lambda q: {"u": [-q["u"][0]], "v": -q["v"]}, dt)
    assert set(out) == {"u", "v"} and isinstance(out["u"], list)
    np.testing.assert_allclose(out["u"][0], a * amp, rtol=1e-14)


def _periodic_cfg(shape, gas):
    return Config(gas=gas, scheme=cases.SCHEME, bc_params={}, topology=(cases.periodic_topology(),), ghost=2)


def test_halo_exchange_periodic_and_two_block():
    print("case: halo", flush=True)
    from famrdp_jax.mesh.halo import halo_exchange
    shape = (5, 9, 8, 7)
    q = np.arange(np.prod(shape), dtype=float).reshape(shape)
    topo = cases.periodic_topology()
    got = halo_exchange((Block(q=q, xyz=np.zeros((3,) + shape[1:])),), (topo,), 2)[0].q
    np.testing.assert_array_equal(got, OB.halo_exchange([q], [topo], 2)[0])
    # reference tests/mesh/test_halo.py:10-46
    s2 = (5, 8, 6, 4)
    q0 = np.zeros(s2); q0[:, 2:6] = np.arange(1.0, 5.0)[None, :, None, None]
    q1 = np.zeros(s2); q1[:, 2:6] = np.arange(10.0, 14.0)[None, :, None, None]
    t0 = BlockTopology(0, {**{f: None for f in Face}, Face.IMAX: (1, Face.IMIN, 0)}, {Face.IMAX: BCType.CUT1TO1})
    t1 = BlockTopology(1, {**{f: None for f in Face}, Face.IMIN: (0, Face.IMAX, 0)}, {Face.IMIN: BCType.CUT1TO1})
    z = np.zeros((3,) + s2[1:])
    out = halo_exchange((Block(q=q0, xyz=z), Block(q=q1, xyz=z)), (t0, t1), 2)
    assert out[0].q[0, -2, 0, 0] == 10.0 and out[0].q[0, -1, 0, 0] == 11.0
    assert out[1].q[0, 0, 0, 0] == 3.0 and out[1].q[0, 1, 0, 0] == 4.0
    bad = BlockTopology(0, {Face.IMIN: (0, Face.IMAX, 1)}, {Face.IMIN: BCType.CUT1TO1})
    with pytest.raises(NotImplementedError):
        halo_exchange((Block(q=q, xyz=np.zeros((3,) + shape[1:])),), (bad,), 2)


def test_bcs_against_oracle():
    print("case: bc", flush=True)
    from famrdp_jax.physics.bc.dispatch import apply_bc
    from famrdp_jax.physics.bc.farfield import apply_farfield
    from famrdp_jax.physics.bc.symmetry import apply_symmetry_plane
    from famrdp_jax.physics.bc.wall import apply_wall_noslip_adiabatic, apply_wall_noslip_isothermal, apply_wall_slip
    q = _state((9, 8, 7), 10)
    for f in Face:
        np.testing.assert_array_equal(apply_wall_noslip_adiabatic(q, f, 2), OB.apply_wall_noslip_adiabatic(q, f, 2))
        np.testing.assert_array_equal(apply_wall_slip(q, f, 2), OB.apply_wall_slip(q, f, 2))
        np.testing.assert_array_equal(apply_symmetry_plane(q, f, 2), OB.apply_symmetry_plane(q, f, 2))
        qi = np.array([1.1, 0.2, 0.3, 0.4, 2.7])
        np.testing.assert_array_equal(apply_farfield(q, f, 2, qi), OB.apply_farfield(q, f, 2, qi))
        np.testing.assert_allclose(apply_wall_noslip_isothermal(q, f, 2, 1.3, 1.4, 1 / 1.4),
                                   OB.apply_wall_noslip_isothermal(q, f, 2, 1.3, 1.4, 1 / 1.4), rtol=1e-14)
    # all kinds on one block, sequential Face order with shared edges/corners
    topo = BlockTopology(0, {}, {Face.IMIN: BCType.WALL, Face.IMAX: BCType.FARFIELD, Face.JMIN: BCType.SYMMETRY,
                                 Face.JMAX: BCType.OUTFLOW, Face.KMIN: BCType.WALL, Face.KMAX: BCType.OUTFLOW})
    params = {Face.KMIN: {"wall_type": "isothermal", "T_wall": 1.2}, Face.IMAX: {"q_inf": [1.0, 0.3, 0, 0, 2.0]},
              Face.KMAX: {"extrapolate": True}}
    cfg = Config(gas=GasModel(R=1 / 1.4), scheme=cases.SCHEME, bc_params=params, topology=(topo,), ghost=2)
    got = apply_bc(Block(q=q, xyz=np.zeros((3, 9, 8, 7))), topo, cfg).q
    ref = OB.apply_bc(q, topo, params, 1.4, 1 / 1.4, 2)
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=0)


def test_metrics_golden_and_oracle():
    print("case: metrics", flush=True)
    from famrdp_jax.mesh.metric import compute_metrics
    d = np.load(os.path.join(GOLD, "test3_metric_subbox.npz"))
    for b in (1, 2):
        m = compute_metrics(d[f"xyz{b}"], ghost=0)
        np.testing.assert_allclose(m.jac[2:22, 2:22, :], d[f"jac{b}"][2:22, 2:22, :], atol=1e-13, rtol=0)
    xyz = cases.curvilinear_xyz(12)
    m = compute_metrics(xyz, ghost=2)
    jac, k, vol = OM.compute_metrics(xyz)
    np.testing.assert_allclose(m.jac, jac, rtol=1e-12)
    np.testing.assert_allclose(m.kxyz, k, rtol=1e-12, atol=1e-12 * np.abs(k).max())
    np.testing.assert_allclose(m.vol, vol, rtol=1e-12)


def test_cfl_dt_and_monitor():
    print("case: cfl", flush=True)
    from famrdp_jax.solver.cfl import cfl_dt
    st, cfg = cases.case_c1(n=16, seed=3)
    q, m = st.blocks[0].q, st.metrics[0]
    dt, nbad = cfl_dt(q, m, cfg)
    ref = OM.cfl_dt(q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), cfg.cfl)
    assert dt == pytest.approx(ref, rel=1e-13) and nbad == 0
    q2 = q.copy(); q2[0, 5, 6, 7] = -1.0; q2[4, 8, 8, 8] = np.nan
    _, nbad = cfl_dt(q2, m, cfg)
    assert nbad == 2


def test_fastmath_rcp_sqrt_accuracy():
    """The stage kernels' reciprocal, square root and reciprocal square root
    stay within 2 ulp of the IEEE results (1/x, sqrt(x), 1/sqrt(x)) over the
    magnitudes the solver meets (and the 1e-300 floors)."""
    print("case: fastmath probe", flush=True)
    import ctypes
    import torch
    from famrdp_jax import _lib as L
    rng = np.random.default_rng(12)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 200000), rng.uniform(0.5, 2.0, 200000),
                        np.array([1e-300, 1e-150, 1e-30, 1.0, 2.0, 3.0, 0.1, 1.4 - 1.0])])
    xd = torch.from_numpy(x).cuda()
    r = torch.empty_like(xd); s = torch.empty_like(xd); q = torch.empty_like(xd)
    L.check(L.lib().famrdp_fastmath_probe(xd.data_ptr(), r.data_ptr(), s.data_ptr(), q.data_ptr(), x.size,
                                          L.stream()))
    r = r.cpu().numpy(); s = s.cpu().numpy(); q = q.cpu().numpy()
    ulp = np.spacing(1.0 / x)
    assert np.abs(r - 1.0 / x).max() <= 0 or (np.abs(r - 1.0 / x) / ulp).max() <= 2.0
    ulp_s = np.spacing(np.sqrt(x))
    assert (np.abs(s - np.sqrt(x)) / ulp_s).max() <= 2.0
    rq = 1.0 / np.sqrt(x)
    eq = (np.abs(q - rq) / np.spacing(rq)).max()
    print(f"  rsqrt max error {eq:.2f} ulp", flush=True)
    assert eq <= 2.0
    z = torch.zeros(4, dtype=torch.float64, device="cuda")
    r0 = torch.empty_like(z); s0 = torch.empty_like(z)
    L.check(L.lib().famrdp_fastmath_probe(z.data_ptr(), r0.data_ptr(), s0.data_ptr(), None, 4, L.stream()))
    assert (s0.cpu().numpy() == 0.0).all()
    # the face flux's one-step forms: a few tens of ulp (hardware seeds are good
    # to 2^-24, one Newton step squares that)
    r1 = torch.empty_like(xd); q1 = torch.empty_like(xd)
    L.check(L.lib().famrdp_fastmath_probe1(xd.data_ptr(), r1.data_ptr(), q1.data_ptr(), x.size, L.stream()))
    r1 = r1.cpu().numpy(); q1 = q1.cpu().numpy()
    e1 = (np.abs(r1 - 1.0 / x) / ulp).max()
    eq1 = (np.abs(q1 - rq) / np.spacing(rq)).max()
    print(f"  one-step rcp max error {e1:.2f} ulp, rsqrt {eq1:.2f} ulp", flush=True)
    assert e1 <= 32.0 and eq1 <= 32.0
