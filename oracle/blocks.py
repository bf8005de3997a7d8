"""Oracle: halo exchange, ghost-cell BCs and the TVD-RK3 step.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates famrdp_jax/mesh/halo.py, famrdp_jax/physics/bc/*.py and
famrdp_jax/solver/step.py.  Topology objects are duck-typed: anything with
``block_id``, ``neighbors`` (face -> None | (block_id, face, orient)) and
``bc_type`` (face -> BC code) works, with faces and BC codes compared as
plain integers (the reference's IntEnum values).
"""
from __future__ import annotations

from . import _ops as X
from .numerics import rhs_block

# famrdp_jax/core/constants.py:4-13, 55-61
CUT1TO1, WALL, SYMMETRY, FARFIELD, INFLOW, OUTFLOW, POLE, PATCHED = -1, 2, 3, 4, 5, 6, 7, 8
IMIN, IMAX, JMIN, JMAX, KMIN, KMAX = range(6)
FACES = (IMIN, IMAX, JMIN, JMAX, KMIN, KMAX)
FACE_AXIS = {IMIN: 1, IMAX: 1, JMIN: 2, JMAX: 2, KMIN: 3, KMAX: 3}   # halo.py:12-16, bc/base.py:3-7


def _is_min(face):
    return int(face) in (IMIN, JMIN, KMIN)


def _ghost_slice(face, ghost):
    """halo.py:19-23."""
    return slice(0, ghost) if _is_min(face) else slice(-ghost, None)


def _interior_adj_slice(face, ghost):
    """halo.py:26-33."""
    return slice(ghost, 2 * ghost) if _is_min(face) else slice(-2 * ghost, -ghost)


def _set(arr, idx, val):
    """Functional slab assignment.  numpy arrays are copied once on entry to
    each public function (``_own``) and then written in place; torch tensors
    are cloned per write so autograd never sees a saved view modified."""
    if X.is_torch(arr):
        arr = arr.clone()
    arr[idx] = val
    return arr


def _own(q):
    return q if X.is_torch(q) else X.copy(q)


def halo_exchange(qs, topos, ghost):
    """halo.py:56-92.  Whole-face slab copies in dict order, each reading the
    pre-exchange source array; orient != 0 raises NotImplementedError."""
    id_to_idx = {t.block_id: i for i, t in enumerate(topos)}
    out = [_own(q) for q in qs]
    for i, topo in enumerate(topos):
        for face, info in topo.neighbors.items():
            if info is None:
                continue
            if topo.bc_type.get(face) is None or int(topo.bc_type.get(face)) != CUT1TO1:
                continue
            nb_id, src_face, orient = info
            if orient != 0:
                raise NotImplementedError(f"orient={orient} not supported in stage 1")
            src = qs[id_to_idx[nb_id]]
            idx_t = [slice(None)] * 4
            idx_s = [slice(None)] * 4
            idx_t[FACE_AXIS[int(face)]] = _ghost_slice(face, ghost)
            idx_s[FACE_AXIS[int(src_face)]] = _interior_adj_slice(src_face, ghost)
            out[i] = _set(out[i], tuple(idx_t), src[tuple(idx_s)])
    return out


def ghost_indices(face, ghost, n_total):
    """bc/base.py:21-33: (outer, inner) mirror pairs."""
    if _is_min(face):
        outer = list(range(ghost))
        inner = [2 * ghost - 1 - g for g in outer]
    else:
        outer = list(range(n_total - 1, n_total - 1 - ghost, -1))
        inner = [n_total - 2 * ghost + (n_total - 1 - o) for o in outer]
    return outer, inner


def _plane(ax, i):
    idx = [slice(None)] * 4
    idx[ax] = i
    return tuple(idx)


def mirror_ghost(q, face, ghost, sign_mask):
    """bc/wall.py:7-27."""
    q = _own(q)
    ax = FACE_AXIS[int(face)]
    outer, inner = ghost_indices(face, ghost, q.shape[ax])
    sign = X.asarray_like(sign_mask, q).reshape((5, 1, 1))
    for o, n in zip(outer, inner):
        q = _set(q, _plane(ax, o), q[_plane(ax, n)] * sign)
    return q


def apply_wall_noslip_adiabatic(q, face, ghost):
    """bc/wall.py:30-32."""
    return mirror_ghost(q, face, ghost, [1.0, -1.0, -1.0, -1.0, 1.0])


def apply_wall_noslip_isothermal(q, face, ghost, T_wall, gamma, R):
    """bc/wall.py:35-58."""
    q = _own(q)
    ax = FACE_AXIS[int(face)]
    outer, inner = ghost_indices(face, ghost, q.shape[ax])
    for o, n in zip(outer, inner):
        qi = q[_plane(ax, n)]
        rho_i = qi[0]
        ke_i = 0.5 * (qi[1] ** 2 + qi[2] ** 2 + qi[3] ** 2) / rho_i
        p_i = (gamma - 1.0) * (qi[4] - ke_i)
        p_g = p_i
        rho_g = p_g / (R * T_wall)
        scale = rho_g / rho_i
        rhou_g = -qi[1] * scale
        rhov_g = -qi[2] * scale
        rhow_g = -qi[3] * scale
        ke_g = 0.5 * (rhou_g ** 2 + rhov_g ** 2 + rhow_g ** 2) / rho_g
        rhoE_g = p_g / (gamma - 1.0) + ke_g
        q = _set(q, _plane(ax, o), X.stack([rho_g, rhou_g, rhov_g, rhow_g, rhoE_g], axis=0))
    return q


def _flip_face_momentum(face):
    sign = [1.0] * 5
    sign[FACE_AXIS[int(face)]] = -1.0      # bc/base.py:9-13 (velocity index == face axis)
    return sign


def apply_wall_slip(q, face, ghost):
    """bc/wall.py:61-64."""
    return mirror_ghost(q, face, ghost, _flip_face_momentum(face))


def apply_symmetry_plane(q, face, ghost):
    """bc/symmetry.py:8-11."""
    return mirror_ghost(q, face, ghost, _flip_face_momentum(face))


def apply_farfield(q, face, ghost, q_inf):
    """bc/farfield.py:7-21: both ghost layers = q_inf."""
    q = _own(q)
    ax = FACE_AXIS[int(face)]
    outer, _ = ghost_indices(face, ghost, q.shape[ax])
    for o in outer:
        qi = q[_plane(ax, o)]
        q = _set(q, _plane(ax, o), q_inf.reshape((5, 1, 1)) * X.ones_like(qi))
    return q


def apply_extrapolate(q, face, ghost):
    """Zeroth-order extrapolation: every ghost layer copies the first interior
    layer.  NOT in the reference (OUTFLOW is a silent no-op there,
    bc/dispatch.py:15-31); this is the build's extension, enabled by
    ``bc_params[face]["extrapolate"] = True`` on an OUTFLOW face."""
    q = _own(q)
    ax = FACE_AXIS[int(face)]
    n = q.shape[ax]
    src = ghost if _is_min(face) else n - 1 - ghost
    outer, _ = ghost_indices(face, ghost, n)
    for o in outer:
        q = _set(q, _plane(ax, o), q[_plane(ax, src)])
    return q


def apply_bc(q, topo, bc_params, gamma, R, ghost):
    """bc/dispatch.py:11-32: faces in Face order, sequential, in place."""
    for face in FACES:
        bct = None
        for k, v in topo.bc_type.items():
            if int(k) == face:
                bct = int(v)
        if bct is None or bct == CUT1TO1:
            continue
        bc_p = {}
        for k, v in bc_params.items():
            if int(k) == face:
                bc_p = v
        if bct == WALL:
            if bc_p.get("wall_type", "adiabatic") == "isothermal":
                q = apply_wall_noslip_isothermal(q, face, ghost, bc_p.get("T_wall", 300.0), gamma, R)
            else:
                q = apply_wall_noslip_adiabatic(q, face, ghost)
        elif bct == SYMMETRY:
            q = apply_symmetry_plane(q, face, ghost)
        elif bct == FARFIELD:
            q_inf = bc_p.get("q_inf", [1.0, 0.0, 0.0, 0.0, 2.5])
            if not X.is_torch(q_inf):
                q_inf = X.asarray_like(q_inf, q)
            q = apply_farfield(q, face, ghost, q_inf)
        elif bct == OUTFLOW and bc_p.get("extrapolate", False):
            q = apply_extrapolate(q, face, ghost)
    return q


def step(qs, metrics, topos, bc_params, gas, dt, *, ghost=2, limiter="minmod"):
    """solver/step.py:18-56.

    qs: list of (5, ni, nj, nk) arrays; metrics: list of (kxyz, vol); gas: dict
    with gamma, R, mu0, T0, Ts, Pr_lam.  Halo once, BCs once, then the inlined
    Shu-Osher RK3 over every padded cell.  Returns the new list of q arrays.
    """
    qs = halo_exchange(qs, topos, ghost)
    qs = [apply_bc(q, t, bc_params, gas["gamma"], gas["R"], ghost) for q, t in zip(qs, topos)]

    def L(qt):
        return [rhs_block(q, m[0], m[1], gas, limiter) for q, m in zip(qt, metrics)]

    L0 = L(qs)
    qs1 = [q + dt * r for q, r in zip(qs, L0)]
    L1 = L(qs1)
    qs2 = [0.75 * q + 0.25 * (q1 + dt * r) for q, q1, r in zip(qs, qs1, L1)]
    L2 = L(qs2)
    return [(1.0 / 3.0) * q + (2.0 / 3.0) * (q2 + dt * r) for q, q2, r in zip(qs, qs2, L2)]


def stage_states(qs, metrics, topos, bc_params, gas, dt, *, ghost=2, limiter="minmod"):
    """Same as ``step`` but returns (q0 after halo+BC, q1, q2, q_new) for
    stage-by-stage comparisons."""
    qs = halo_exchange(qs, topos, ghost)
    qs = [apply_bc(q, t, bc_params, gas["gamma"], gas["R"], ghost) for q, t in zip(qs, topos)]

    def L(qt):
        return [rhs_block(q, m[0], m[1], gas, limiter) for q, m in zip(qt, metrics)]

    qs1 = [q + dt * r for q, r in zip(qs, L(qs))]
    qs2 = [0.75 * q + 0.25 * (q1 + dt * r) for q, q1, r in zip(qs, qs1, L(qs1))]
    qn = [(1.0 / 3.0) * q + (2.0 / 3.0) * (q2 + dt * r) for q, q2, r in zip(qs, qs2, L(qs2))]
    return qs, qs1, qs2, qn
