"""Host side of a step plan: translates the reference's topology / Config
objects into the C-ABI plan (CUT1TO1 copies in the reference's visiting
order, BC faces with their parameters) and caches plans by configuration."""
from __future__ import annotations

import ctypes

from . import _lib as L
from .core.constants import BCType, Face

BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL, BC_MIRROR_NORMAL, BC_FARFIELD, BC_EXTRAPOLATE = range(5)
DEFAULT_Q_INF = (1.0, 0.0, 0.0, 0.0, 2.5)      # physics/bc/dispatch.py:30


def _bct(topo, face):
    for k, v in topo.bc_type.items():
        if int(k) == int(face):
            return int(v)
    return None


def _params(bc_params, face):
    for k, v in bc_params.items():
        if int(k) == int(face):
            return v
    return {}


def _floats(x):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return tuple(float(v) for v in x.detach().cpu().reshape(-1).tolist())
    except ImportError:  # pragma: no cover
        pass
    return tuple(float(v) for v in (x.tolist() if hasattr(x, "tolist") else x))


def cuts_from_topology(topos):
    """mesh/halo.py:76-90: blocks in order, faces in neighbors-dict order;
    only faces whose bc_type is CUT1TO1 copy; unknown neighbour ids raise
    KeyError like the reference's id_to_idx lookup."""
    id_to_idx = {t.block_id: i for i, t in enumerate(topos)}
    cuts = []
    for i, topo in enumerate(topos):
        for face, info in topo.neighbors.items():
            if info is None or _bct(topo, face) != int(BCType.CUT1TO1):
                continue
            nb_id, src_face, orient = info
            cuts.append((i, int(face), id_to_idx[nb_id], int(src_face), int(orient)))
    return cuts


def bcs_from_config(topos, bc_params, ghost_unused=None):
    """physics/bc/dispatch.py:11-32 for every block: Face order, CUT1TO1 and
    the no-op kinds (INFLOW, OUTFLOW, POLE, PATCHED) skipped; OUTFLOW with
    ``{"extrapolate": True}`` selects the extrapolation extension."""
    out = []
    for b, topo in enumerate(topos):
        for face in Face:
            bct = _bct(topo, face)
            if bct is None or bct == int(BCType.CUT1TO1):
                continue
            p = _params(bc_params, face)
            if bct == int(BCType.WALL):
                if p.get("wall_type", "adiabatic") == "isothermal":
                    out.append((b, int(face), BC_WALL_ISOTHERMAL, float(p.get("T_wall", 300.0)), DEFAULT_Q_INF))
                else:
                    out.append((b, int(face), BC_WALL_ADIABATIC, 0.0, DEFAULT_Q_INF))
            elif bct == int(BCType.SYMMETRY):
                out.append((b, int(face), BC_MIRROR_NORMAL, 0.0, DEFAULT_Q_INF))
            elif bct == int(BCType.FARFIELD):
                out.append((b, int(face), BC_FARFIELD, 0.0, _floats(p.get("q_inf", DEFAULT_Q_INF))))
            elif bct == int(BCType.OUTFLOW) and p.get("extrapolate", False):
                out.append((b, int(face), BC_EXTRAPOLATE, 0.0, DEFAULT_Q_INF))
    return out


class Plan:
    """Owns one ``famrdp_plan``: block sizes, halo table, BC tables, tiles and
    the per-block stage workspace (allocated once, on the current GPU)."""

    def __init__(self, dims, cuts, bcs, gas, limiter="minmod", ghost=2, workspace=True, adjoint=False):
        L.require_gpu()
        lib = L.load()
        self.nblocks = len(dims)
        self.dims = [tuple(int(v) for v in d) for d in dims]
        flat = (L.i32 * (3 * self.nblocks))(*[v for d in self.dims for v in d])
        ccuts = (L.Cut * max(1, len(cuts)))(*[L.Cut(*c) for c in cuts])
        cbcs = (L.Bc * max(1, len(bcs)))(*[L.Bc(b, f, k, 0, t, (L.f64 * 5)(*qi)) for (b, f, k, t, qi) in bcs])
        g = L.gas_struct(gas)
        h = L.vp()
        L.check(lib.famrdp_plan_create(ctypes.byref(h), self.nblocks, flat, len(cuts), ccuts, len(bcs), cbcs,
                                       ctypes.byref(g), L.limiter_code(limiter), int(ghost),
                                       (0 if workspace else 1) | (2 if adjoint else 0)))
        self._h = h
        self._lib = lib

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.famrdp_plan_destroy(h)
            self._h = None

    @staticmethod
    def _ptrs(ts):
        return (L.vp * len(ts))(*[L.ptr(t) for t in ts])

    def step(self, qs, kxyzs, vols, dt, *, batched=False, graph_steps=0, direct=False, stages=None):
        """One step in place on the device tensors ``qs`` (or ``graph_steps``
        steps replayed from a captured hipGraph).  ``direct`` selects the
        reference-order one-thread-per-cell kernels instead of the tiled ones.
        ``stages`` = (q1s, q2s), per-block device tensors shaped like ``qs``:
        the step writes its stage-1/2 states there (eager steps only)."""
        flags = (1 if batched else 0) | (2 if direct else 0)
        if stages is not None:
            if graph_steps:
                raise ValueError("stages: eager steps only")
            q1s, q2s = stages
            L.check(self._lib.famrdp_step_stages(self._h, self._ptrs(qs), self._ptrs(kxyzs), self._ptrs(vols),
                                                 float(dt), flags, self._ptrs(q1s), self._ptrs(q2s), L.stream()))
        elif graph_steps:
            L.check(self._lib.famrdp_step_graph(self._h, self._ptrs(qs), self._ptrs(kxyzs), self._ptrs(vols),
                                                float(dt), flags, int(graph_steps), L.stream()))
        else:
            L.check(self._lib.famrdp_step(self._h, self._ptrs(qs), self._ptrs(kxyzs), self._ptrs(vols), float(dt),
                                          flags, L.stream()))

    def autotune(self, qs, kxyzs, vols, dt, *, batched=False, reps=3):
        """Time every stage-kernel variant (tile shapes of the fused kernel and
        of the split pair) on these arrays (stages 1 and 2, which write only
        the plan's workspace: the state is not modified) and keep the fastest.
        Returns (chosen variant, tile dims, per-variant ms of stages 1 + 2)."""
        n = L.NUM_TILE_VARIANTS
        ms = (L.f64 * n)()
        L.check(self._lib.famrdp_plan_autotune(self._h, self._ptrs(qs), self._ptrs(kxyzs), self._ptrs(vols), float(dt),
                                               1 if batched else 0, int(reps), ms, L.stream()))
        v, dims = self.tile()
        return v, dims, list(ms)

    def tile(self):
        v = L.i32(0)
        d = (L.i32 * 3)()
        L.check(self._lib.famrdp_plan_get_tile(self._h, ctypes.byref(v), d))
        return v.value, tuple(d)

    def set_tile(self, variant):
        L.check(self._lib.famrdp_plan_set_tile(self._h, int(variant)))

    def timing_start(self):
        L.check(self._lib.famrdp_plan_timing(self._h, 1, None, None))

    def timing_stop(self):
        """Returns (summed stage ms, number of stages timed)."""
        ms, n = L.f64(0.0), L.i64(0)
        L.check(self._lib.famrdp_plan_timing(self._h, 0, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def step_vjp(self, q_start, kxyzs, vols, lams, dt, dqinf=None, *, batched=False, direct=False, stages=None,
                 consume_start=False):
        """In place: lams (cotangent of the step's result) -> cotangent of
        q_start.  ``stages`` = (q1s, q2s) saved by stage_states after the
        forward step from q_start; without it they are recomputed.
        consume_start: q_start may be overwritten (its ghost cells receive the
        halo and BCs) instead of being copied first."""
        flags = (1 if batched else 0) | (2 if direct else 0) | (4 if consume_start else 0)
        if stages is None:
            L.check(self._lib.famrdp_step_vjp(self._h, self._ptrs(q_start), self._ptrs(kxyzs), self._ptrs(vols),
                                              self._ptrs(lams), float(dt), L.ptr(dqinf), flags, L.stream()))
        else:
            q1s, q2s = stages
            L.check(self._lib.famrdp_step_vjp_stages(self._h, self._ptrs(q_start), self._ptrs(q1s), self._ptrs(q2s),
                                                     self._ptrs(kxyzs), self._ptrs(vols), self._ptrs(lams), float(dt),
                                                     L.ptr(dqinf), flags, L.stream()))

    def stage_states(self, q1s, q2s):
        """Copy the stage-1/2 states of the last step into q1s, q2s."""
        L.check(self._lib.famrdp_plan_stage_states(self._h, self._ptrs(q1s), self._ptrs(q2s), L.stream()))

    def set_block_params(self, mu0=None, dt_scale=None):
        arr = lambda x: None if x is None else (L.f64 * self.nblocks)(*[float(v) for v in x])
        L.check(self._lib.famrdp_plan_set_block_params(self._h, arr(mu0), arr(dt_scale)))

    def check_state(self, qs, out):
        """Enqueue the state monitor: out (device tensor, >= 1 double) gets the
        number of bad interior cells over all blocks."""
        L.check(self._lib.famrdp_plan_check_state(self._h, self._ptrs(qs), L.ptr(out), L.stream()))

    def raise_if_bad(self, qs):
        """Config.nan_check: monitor the state and raise FloatingPointError if
        any interior cell is non-finite or has rho <= 0 or p <= 0."""
        torch = L._torch()
        out = torch.empty(1, dtype=torch.float64, device="cuda")
        self.check_state(qs, out)
        nbad = int(out.item())
        if nbad:
            raise FloatingPointError(f"nan_check: {nbad} interior cells are non-finite or have rho <= 0 or p <= 0")

    def halo_exchange(self, qs):
        L.check(self._lib.famrdp_halo_exchange(self._h, self._ptrs(qs), L.stream()))

    def apply_bc(self, qs):
        L.check(self._lib.famrdp_apply_bc(self._h, self._ptrs(qs), L.stream()))


_CACHE: dict = {}


def plan_for(dims, topos, bc_params, gas, limiter, ghost):
    cuts = cuts_from_topology(topos)
    bcs = bcs_from_config(topos, bc_params)
    key = (tuple(tuple(d) for d in dims), tuple(cuts), tuple(bcs),
           (gas.gamma, gas.R, gas.mu0, gas.T0, gas.Ts, gas.Pr_lam), limiter, ghost)
    p = _CACHE.get(key)
    if p is None:
        if len(_CACHE) > 16:
            _CACHE.clear()
        p = Plan(dims, cuts, bcs, gas, limiter, ghost)
        _CACHE[key] = p
    return p
