"""CPU check of the hand-written reverse sweeps (famrdp_adjoint.h) against
torch autograd through the oracle, which follows JAX's derivative
conventions (tie splitting in maximum/minimum, abs'(0) = +1, where selects).

The sweeps are compiled for the host into a test-only library
(tests/hostcheck/adjoint_hostcheck.hip); the product never loads it.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest
import torch

from oracle import blocks as OB
from oracle import numerics as ON

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "famrdp-jax_amd", "csrc")
G = 1.4
dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def hc(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("hc") / "libhc.so")
    subprocess.run([hipcc, "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + CSRC,
                    os.path.join(HERE, "hostcheck", "adjoint_hostcheck.hip"), "-o", out],
                   check=True, capture_output=True)
    return ctypes.CDLL(out)


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(dp)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _prims(n, rng, mach=0.6):
    rho = rng.uniform(0.6, 1.4, n)
    u, v, w = (rng.uniform(-mach, mach, n) for _ in range(3))
    p = rng.uniform(0.5, 0.9, n)
    return np.stack([rho, u, v, w, p], axis=1)


def test_roe_vjp_matches_autograd(hc):
    rng = np.random.default_rng(21)
    n = 400
    pl, pr = _prims(n, rng), _prims(n, rng)
    pr[:50] = pl[:50] * (1 + 1e-3 * rng.normal(size=(50, 5)))        # near-identical states (Harten branch)
    nrm = rng.normal(size=(n, 3)); nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm[:20] = [0.0, 1.0, 0.0]; pl[:20, 2] = 0.0; pr[:20, 2] = 0.0     # u.n = 0 exactly: abs'(0)
    lF = rng.normal(size=(n, 5))
    lpl = np.zeros((n, 5)); lpr = np.zeros((n, 5))
    hc.hc_roe_vjp(n, P(pl), P(pr), P(nrm), ctypes.c_double(G), P(lF), P(lpl), P(lpr))
    tl = torch.tensor(pl.T, requires_grad=True); tr = torch.tensor(pr.T, requires_grad=True)
    F = ON.roe_flux(ON.prim_to_cons(tl, G), ON.prim_to_cons(tr, G), torch.tensor(nrm.T), gamma=G)
    (F * torch.tensor(lF.T)).sum().backward()
    assert _rel(lpl, tl.grad.numpy().T) < 1e-10
    assert _rel(lpr, tr.grad.numpy().T) < 1e-10


def test_cons_to_prim_vjp(hc):
    rng = np.random.default_rng(22)
    n = 200
    pv = _prims(n, rng)
    q = ON.prim_to_cons(pv.T, G).T.copy()
    lp = rng.normal(size=(n, 5))
    lq = np.zeros((n, 5))
    hc.hc_cons_to_prim_vjp(n, P(q), ctypes.c_double(G), P(lp), P(lq))
    tq = torch.tensor(q.T, requires_grad=True)
    (ON.cons_to_prim(tq, G) * torch.tensor(lp.T)).sum().backward()
    assert _rel(lq, tq.grad.numpy().T) < 1e-12


def test_minmod_fast_equals_reference_minmod(hc):
    """The stage kernels' minmod (median of a, b and 0 in four min/max) has
    the value of muscl2.py:5-7 for every finite or infinite pair; a NaN
    argument is dropped (the other one is returned, two NaNs give 0) -- the
    face routine re-injects the stencil's NaN itself."""
    rng = np.random.default_rng(5)
    a = rng.normal(size=400); b = rng.normal(size=400)
    b[:30] = a[:30]; b[30:60] = -a[30:60]; b[60:80] = 2 * a[60:80]
    a[80:90] = 0.0; b[90:100] = -0.0; a[100:104] = [np.inf, -np.inf, np.inf, 1.0]
    b[100:104] = [2.0, -3.0, -1.0, np.inf]
    out = np.zeros(400)
    hc.hc_minmod_fast(400, P(a), P(b), P(out))
    np.testing.assert_array_equal(out, ON.minmod(a, b))
    nan = np.nan
    a2 = np.array([nan, 2.0, nan, nan, -1.5]); b2 = np.array([3.0, nan, -2.0, nan, nan])
    out2 = np.zeros(5)
    hc.hc_minmod_fast(5, P(a2), P(b2), P(out2))
    np.testing.assert_array_equal(out2, [3.0, 2.0, -2.0, 0.0, -1.5])


@pytest.mark.parametrize("lim", [0, 1])
def test_limiter_vjp_with_ties_and_zeros(hc, lim):
    rng = np.random.default_rng(23 + lim)
    A = rng.normal(size=300); B = rng.normal(size=300)
    B[:20] = A[:20]                # |A| == |B| ties (min splits 1/2)
    B[20:40] = -A[20:40]           # opposite signs
    A[40:50] = 0.0                 # zeros
    ls = rng.normal(size=300)
    lA = np.zeros(300); lB = np.zeros(300)
    hc.hc_limiter_vjp(300, lim, P(A), P(B), P(ls), P(lA), P(lB))
    tA = torch.tensor(A, requires_grad=True); tB = torch.tensor(B, requires_grad=True)
    f = ON.minmod if lim == 0 else ON.vanleer
    (f(tA, tB) * torch.tensor(ls)).sum().backward()
    np.testing.assert_allclose(lA, tA.grad.numpy(), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(lB, tB.grad.numpy(), rtol=1e-13, atol=1e-15)


def _torch_viscous_flux(pv, T, d, k, gas):
    """Pointwise G of rhs/viscous.py:36-108 for autograd (torch)."""
    gamma, R, mu0, T0, Ts, Pr = gas
    mu = ON.sutherland_viscosity(T, mu0, T0, Ts)
    kappa = gamma * R / (gamma - 1.0) * mu / Pr
    gr = [[d[:, 0, s] * k[:, 0 * 3 + m] + d[:, 1, s] * k[:, 1 * 3 + m] + d[:, 2, s] * k[:, 2 * 3 + m]
           for m in range(3)] for s in range(4)]
    div = gr[0][0] + gr[1][1] + gr[2][2]
    tau = [[2 * mu * (gr[0][0] - div / 3.0), mu * (gr[0][1] + gr[1][0]), mu * (gr[0][2] + gr[2][0])],
           [mu * (gr[0][1] + gr[1][0]), 2 * mu * (gr[1][1] - div / 3.0), mu * (gr[1][2] + gr[2][1])],
           [mu * (gr[0][2] + gr[2][0]), mu * (gr[1][2] + gr[2][1]), 2 * mu * (gr[2][2] - div / 3.0)]]
    u, v, w = pv[:, 1], pv[:, 2], pv[:, 3]
    E = [u * tau[0][j] + v * tau[1][j] + w * tau[2][j] - (-kappa * gr[3][j]) for j in range(3)]
    return torch.stack([tau[0][0], tau[0][1], tau[0][2], tau[1][1], tau[1][2], tau[2][2], E[0], E[1], E[2]], dim=1)


def test_viscous_flux_vjp(hc):
    rng = np.random.default_rng(25)
    n = 200
    pv = _prims(n, rng)
    T = rng.uniform(0.7, 1.4, n)
    d = rng.normal(size=(n, 3, 4)) * 0.1
    k = rng.normal(size=(n, 9)) + np.tile(np.eye(3).reshape(9), (n, 1))
    gas = np.array([1.4, 1 / 1.4, 0.02, 1.0, 0.383, 0.72])
    lG = rng.normal(size=(n, 9))
    out = np.zeros((n, 16))
    hc.hc_viscous_vjp(n, P(pv), P(T), P(d), P(k), P(gas), P(lG), P(out))
    tp = torch.tensor(pv, requires_grad=True); tT = torch.tensor(T, requires_grad=True)
    td = torch.tensor(d, requires_grad=True)
    Gt = _torch_viscous_flux(tp, tT, td, torch.tensor(k), gas)
    (Gt * torch.tensor(lG)).sum().backward()
    assert _rel(out[:, 0:3], tp.grad.numpy()[:, 1:4]) < 1e-12
    assert float(np.abs(tp.grad.numpy()[:, [0, 4]]).max()) == 0.0
    assert _rel(out[:, 3], tT.grad.numpy()) < 1e-12
    assert _rel(out[:, 4:], td.grad.numpy().reshape(n, 12)) < 1e-12


def test_isothermal_vjp(hc):
    rng = np.random.default_rng(26)
    n = 100
    q = ON.prim_to_cons(_prims(n, rng).T, G).T.copy()
    lo = rng.normal(size=(n, 5))
    li = np.zeros((n, 5))
    hc.hc_isothermal_vjp(n, P(q), ctypes.c_double(G), ctypes.c_double(1 / 1.4), ctypes.c_double(1.2), P(lo), P(li))
    # oracle: a 1-cell-wide block whose IMIN ghost layer 1 mirrors inner cell 2
    tq = torch.tensor(q.T.reshape(5, 1, 1, n), requires_grad=True)
    block = torch.cat([torch.zeros(5, 2, 1, n, dtype=torch.float64), tq, torch.ones(5, 5, 1, n, dtype=torch.float64)], 1)
    out = OB.apply_wall_noslip_isothermal(block, 0, 2, 1.2, G, 1 / 1.4)
    (out[:, 1, 0, :] * torch.tensor(lo.T)).sum().backward()
    assert _rel(li, tq.grad.numpy().reshape(5, n).T) < 1e-12
