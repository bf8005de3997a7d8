"""Runs one ghost-cell BC face on the GPU through a halo/BC-only plan."""
from __future__ import annotations

from ... import _lib as L
from ..._plan import Plan, DEFAULT_Q_INF
from ...core.types import GasModel


def run_face(q, face, ghost, kind, *, T_wall=0.0, q_inf=DEFAULT_Q_INF, gamma=1.4, R=287.058):
    qd = L.device_f64(q).clone()
    plan = Plan([tuple(qd.shape[1:])], [], [(0, int(face), kind, float(T_wall), tuple(float(v) for v in q_inf))],
                GasModel(gamma=gamma, R=R), "minmod", ghost, workspace=False)
    plan.apply_bc([qd])
    return L.like_input(qd, q)
