"""Oracle: CPU restatement of famrdp_jax's pointwise and RHS numerics.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg, and only as the checker.  The product
path never imports this module.

Every function restates the reference formula op for op, in the same
evaluation order (numpy evaluates ``a + b + c`` as ``(a + b) + c`` like the
jnp expression it mirrors), so the rounding differs from XLA:CPU only where
XLA contracts or reorders.  Citations are to the reference repository.

The same code runs on numpy arrays (forward oracle) or on CPU float64 torch
tensors (autograd oracle for gradients); see ``_ops``.
"""
from __future__ import annotations

from . import _ops as X


# --------------------------------------------------------------------------
# Equation of state  (famrdp_jax/physics/eos.py:5-25)
# --------------------------------------------------------------------------

def prim_to_cons(prim, gamma):
    """famrdp_jax/physics/eos.py:5-9."""
    rho = prim[0]; u = prim[1]; v = prim[2]; w = prim[3]; p = prim[4]
    ke = 0.5 * (u * u + v * v + w * w)
    rhoE = p / (gamma - 1.0) + rho * ke
    return X.stack([rho, rho * u, rho * v, rho * w, rhoE], axis=0)


def cons_to_prim(q, gamma):
    """famrdp_jax/physics/eos.py:12-17."""
    rho = q[0]
    u = q[1] / rho; v = q[2] / rho; w = q[3] / rho
    ke = 0.5 * (u * u + v * v + w * w)
    p = (gamma - 1.0) * (q[4] - rho * ke)
    return X.stack([rho, u, v, w, p], axis=0)


def speed_of_sound(rho, p, gamma):
    """famrdp_jax/physics/eos.py:20-21."""
    return X.sqrt(gamma * p / rho)


def sutherland_viscosity(T, mu0, T0, Ts):
    """famrdp_jax/physics/eos.py:24-25: mu0 * (T/T0)**1.5 * (T0+Ts)/(T+Ts)."""
    return mu0 * (T / T0) ** 1.5 * (T0 + Ts) / (T + Ts)


# --------------------------------------------------------------------------
# MUSCL-2 reconstruction  (famrdp_jax/physics/reconstruct/muscl2.py)
# --------------------------------------------------------------------------

def minmod(a, b):
    """muscl2.py:5-7."""
    s = X.sign(a) * X.maximum(X.sign(a) * X.sign(b), 0.0)
    return s * X.minimum(X.abs_(a), X.abs_(b))


def vanleer(a, b):
    """muscl2.py:10-12."""
    safe_denom = X.where(X.abs_(a + b) > 1e-30, a + b, 1.0)
    return X.where(a * b > 0.0, 2.0 * a * b / safe_denom, 0.0)


LIMITERS = {"minmod": minmod, "vanleer": vanleer}


def _slc(arr, axis, start, end):
    idx = [slice(None)] * arr.ndim
    idx[axis] = slice(start, end)
    return arr[tuple(idx)]


def muscl2_reconstruct(pv, *, axis, limiter="minmod"):
    """muscl2.py:18-37.  Face j (0..n-4) lies between cells j+1 and j+2."""
    lim = LIMITERS[limiter]
    n = pv.shape[axis]
    pmm = _slc(pv, axis, 0, n - 3)
    pc = _slc(pv, axis, 1, n - 2)
    pp1 = _slc(pv, axis, 2, n - 1)
    pp2 = _slc(pv, axis, 3, n)
    slope_L = lim(pc - pmm, pp1 - pc)
    slope_R = lim(pp1 - pc, pp2 - pp1)
    pv_L = pc + 0.5 * slope_L
    pv_R = pp1 - 0.5 * slope_R
    return pv_L, pv_R


# --------------------------------------------------------------------------
# Roe flux with Harten entropy fix  (famrdp_jax/flux/roe.py)
# --------------------------------------------------------------------------

def physical_flux(U, n, gamma):
    """roe.py:5-18."""
    rho = X.maximum(U[0], 1e-300)
    u = U[1] / rho; v = U[2] / rho; w = U[3] / rho
    ke = 0.5 * (u * u + v * v + w * w)
    p = (gamma - 1.0) * (U[4] - rho * ke)
    un = u * n[0] + v * n[1] + w * n[2]
    H = (U[4] + p) / rho
    return X.stack([
        rho * un,
        rho * u * un + p * n[0],
        rho * v * un + p * n[1],
        rho * w * un + p * n[2],
        rho * H * un,
    ], axis=0)


def roe_flux(UL, UR, nxyz, *, gamma, eps_entropy=0.1):
    """roe.py:21-85 (general-normal Roe flux, Harten fix on lambda1/lambda5)."""
    rhoL = X.maximum(UL[0], 1e-300)
    rhoR = X.maximum(UR[0], 1e-300)
    uL, vL, wL = UL[1] / rhoL, UL[2] / rhoL, UL[3] / rhoL
    uR, vR, wR = UR[1] / rhoR, UR[2] / rhoR, UR[3] / rhoR
    keL = 0.5 * (uL * uL + vL * vL + wL * wL); keR = 0.5 * (uR * uR + vR * vR + wR * wR)
    pL = (gamma - 1.0) * (UL[4] - rhoL * keL); pR = (gamma - 1.0) * (UR[4] - rhoR * keR)
    HL = (UL[4] + pL) / rhoL; HR = (UR[4] + pR) / rhoR

    sqL = X.sqrt(rhoL); sqR = X.sqrt(rhoR)
    ss = X.maximum(sqL + sqR, 1e-150)
    u_h = (sqL * uL + sqR * uR) / ss
    v_h = (sqL * vL + sqR * vR) / ss
    w_h = (sqL * wL + sqR * wR) / ss
    H_h = (sqL * HL + sqR * HR) / ss
    q2 = u_h ** 2 + v_h ** 2 + w_h ** 2
    c2_safe = X.maximum((gamma - 1.0) * (H_h - 0.5 * q2), 1e-30)
    c_h = X.sqrt(c2_safe)
    un_h = u_h * nxyz[0] + v_h * nxyz[1] + w_h * nxyz[2]
    rho_h = sqL * sqR

    eps = eps_entropy * (X.abs_(un_h) + c_h)

    def harten(lam):
        return X.where(X.abs_(lam) >= eps, X.abs_(lam), 0.5 * (lam ** 2 / eps + eps))

    lam1 = un_h - c_h; lam2 = un_h; lam5 = un_h + c_h
    al1 = harten(lam1); al2 = X.abs_(lam2); al5 = harten(lam5)

    dU = UR - UL
    dp = (gamma - 1.0) * (dU[4] - u_h * dU[1] - v_h * dU[2] - w_h * dU[3] + 0.5 * q2 * dU[0])
    dun = (dU[1] * nxyz[0] + dU[2] * nxyz[1] + dU[3] * nxyz[2]) - un_h * dU[0]
    a1 = (dp - rho_h * c_h * dun) / (2.0 * c2_safe)
    a5 = (dp + rho_h * c_h * dun) / (2.0 * c2_safe)
    a23 = dU[0] - dp / c2_safe

    one = X.ones_like(u_h)
    R1 = X.stack([one, u_h - c_h * nxyz[0], v_h - c_h * nxyz[1], w_h - c_h * nxyz[2], H_h - c_h * un_h])
    R5 = X.stack([one, u_h + c_h * nxyz[0], v_h + c_h * nxyz[1], w_h + c_h * nxyz[2], H_h + c_h * un_h])
    R2 = X.stack([one, u_h, v_h, w_h, 0.5 * q2])

    du_t = dU[1] / rho_h - u_h * dU[0] / rho_h
    dv_t = dU[2] / rho_h - v_h * dU[0] / rho_h
    dw_t = dU[3] / rho_h - w_h * dU[0] / rho_h
    dun_full = du_t * nxyz[0] + dv_t * nxyz[1] + dw_t * nxyz[2]
    dvt = X.stack([X.zeros_like(u_h),
                   rho_h * (du_t - dun_full * nxyz[0]),
                   rho_h * (dv_t - dun_full * nxyz[1]),
                   rho_h * (dw_t - dun_full * nxyz[2]),
                   rho_h * (u_h * (du_t - dun_full * nxyz[0]) + v_h * (dv_t - dun_full * nxyz[1])
                            + w_h * (dw_t - dun_full * nxyz[2]))])

    absA_dU = al1 * a1 * R1 + al5 * a5 * R5 + al2 * (a23 * R2 + dvt)

    FL = physical_flux(UL, nxyz, gamma)
    FR = physical_flux(UR, nxyz, gamma)
    return 0.5 * (FL + FR) - 0.5 * absA_dU


# --------------------------------------------------------------------------
# Inviscid RHS  (famrdp_jax/rhs/invscd.py:18-133)
# --------------------------------------------------------------------------

def compute_invscd_rhs(q, kxyz, vol, *, gamma, limiter="minmod"):
    """invscd.py:18-133.

    Per axis d: faces f = 0..n-4 between cells f+1 and f+2, face metric
    0.5*(kxyz[d][f+1] + kxyz[d][f+2]), area weight |n| + 1e-300, Roe flux on
    the unit normal times |n|, and cells c = 2..n-3 along d receive
    -(F[c-1] - F[c-2]) / (vol[c] + 1e-300), over the full extent of the other
    two axes.  ``reconstruct``/``flux``/``ghost`` of the reference are unused
    there and are therefore not parameters here.
    """
    pv = cons_to_prim(q, gamma)
    rhs = X.zeros_like(q)
    for axis in (1, 2, 3):
        n = pv.shape[axis]
        d = axis - 1
        pv_L, pv_R = muscl2_reconstruct(pv, axis=axis, limiter=limiter)
        UL = prim_to_cons(pv_L, gamma)
        UR = prim_to_cons(pv_R, gamma)
        k_dir = kxyz[d]
        k_L = _slc(k_dir, axis, 1, n - 2)
        k_R = _slc(k_dir, axis, 2, n - 1)
        n_face = 0.5 * (k_L + k_R)
        n_mag = X.sqrt(n_face[0] ** 2 + n_face[1] ** 2 + n_face[2] ** 2) + 1e-300
        n_unit = n_face / n_mag[None]
        F_face = roe_flux(UL, UR, n_unit, gamma=gamma)
        F_area = F_face * n_mag[None]
        div = _slc(F_area, axis, 1, None) - _slc(F_area, axis, 0, -1)
        vol_cell = _slc(vol, d, 2, n - 2)
        rhs_dir = -div / (vol_cell[None] + 1e-300)
        idx = [slice(None)] * 4
        idx[axis] = slice(2, n - 2)
        idx = tuple(idx)
        if X.is_torch(rhs):
            rhs = rhs.clone()
        rhs[idx] = rhs[idx] + rhs_dir
    return rhs


# --------------------------------------------------------------------------
# Viscous RHS  (famrdp_jax/rhs/viscous.py:19-119)
# --------------------------------------------------------------------------

def central2(f, axis):
    """viscous.py:12-16: 0.5*(roll(f,-1) - roll(f,+1)), periodic wrap."""
    fp = X.roll(f, -1, axis)
    fm = X.roll(f, 1, axis)
    return 0.5 * (fp - fm)


def compute_viscous_rhs(q, kxyz, *, gamma, R, mu0, T0, Ts, Pr_lam):
    """viscous.py:19-119 (non-conservative chain-rule form, roll wrap)."""
    pv = cons_to_prim(q, gamma)
    rho = pv[0]; u = pv[1]; v = pv[2]; w = pv[3]; p = pv[4]
    T = p / X.maximum(rho * R, 1e-300)
    mu = sutherland_viscosity(T, mu0, T0, Ts)
    cp = gamma * R / (gamma - 1.0)
    kappa = cp * mu / Pr_lam

    def grad_xyz(phi):
        d = [central2(phi, a) for a in (0, 1, 2)]
        # einsum("lijk,lmijk->mijk"): sum over l in order 0,1,2
        return [d[0] * kxyz[0, m] + d[1] * kxyz[1, m] + d[2] * kxyz[2, m] for m in range(3)]

    du = grad_xyz(u); dv = grad_xyz(v); dw = grad_xyz(w); dT = grad_xyz(T)
    div_u = du[0] + dv[1] + dw[2]
    tau = [
        [2 * mu * (du[0] - div_u / 3.0), mu * (du[1] + dv[0]), mu * (du[2] + dw[0])],
        [mu * (du[1] + dv[0]), 2 * mu * (dv[1] - div_u / 3.0), mu * (dv[2] + dw[1])],
        [mu * (du[2] + dw[0]), mu * (dv[2] + dw[1]), 2 * mu * (dw[2] - div_u / 3.0)],
    ]
    heat = [-kappa * dT[m] for m in range(3)]
    vel = [u, v, w]

    def div_physical_scalar(f):
        result = X.zeros_like(f[0])
        for j in range(3):
            d = [central2(f[j], a) for a in (0, 1, 2)]
            result = result + (d[0] * kxyz[0, j] + d[1] * kxyz[1, j] + d[2] * kxyz[2, j])
        return result

    rhs_mom = [div_physical_scalar(tau[i]) for i in range(3)]
    G_energy = [vel[0] * tau[0][j] + vel[1] * tau[1][j] + vel[2] * tau[2][j] - heat[j]
                for j in range(3)]
    rhs_energy = div_physical_scalar(G_energy)
    return X.stack([X.zeros_like(u), rhs_mom[0], rhs_mom[1], rhs_mom[2], rhs_energy], axis=0)


def rhs_block(q, kxyz, vol, gas, limiter="minmod"):
    """solver/step.py:10-15: inviscid + viscous."""
    inv = compute_invscd_rhs(q, kxyz, vol, gamma=gas["gamma"], limiter=limiter)
    vis = compute_viscous_rhs(q, kxyz, **gas)
    return inv + vis


# --------------------------------------------------------------------------
# TVD-RK3  (famrdp_jax/time/rk.py:5-20)
# --------------------------------------------------------------------------

def rk3_step(q, L, dt):
    """time/rk.py:5-20 (Shu-Osher)."""
    q1 = q + dt * L(q)
    q2 = 0.75 * q + 0.25 * (q1 + dt * L(q1))
    return (1.0 / 3.0) * q + (2.0 / 3.0) * (q2 + dt * L(q2))
