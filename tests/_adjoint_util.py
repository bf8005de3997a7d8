"""Shared pieces of the adjoint tests: the autograd oracle gradient (torch
float64 through the oracle step) and the small wall/farfield case."""
import numpy as np
import torch

from famrdp_jax import cases
from famrdp_jax.adjoint import OBJ_KINETIC_ENERGY
from famrdp_jax.core.types import Block, Config, SchemeChoice, State
from oracle import blocks as OB
from oracle import metric as OM


def _torch_objective(qs, mets, kind, ghost=2, gamma=1.4):
    J = 0.0
    for q, (k, vol) in zip(qs, mets):
        sl = slice(ghost, -ghost)
        if kind == OBJ_KINETIC_ENERGY:
            qi = q[:, sl, sl, sl]
            J = J + (0.5 * (qi[1] ** 2 + qi[2] ** 2 + qi[3] ** 2) / qi[0] * vol[sl, sl, sl]).sum()
        else:
            qi = q[:, sl, sl, ghost]
            ke = 0.5 * (qi[1] ** 2 + qi[2] ** 2 + qi[3] ** 2) / qi[0]
            p = (gamma - 1.0) * (qi[4] - ke)
            A = torch.sqrt(k[2, 0, sl, sl, ghost] ** 2 + k[2, 1, sl, sl, ghost] ** 2 + k[2, 2, sl, sl, ghost] ** 2)
            J = J + (p * A * vol[sl, sl, ghost]).sum()
    return J


def _oracle_grad(state, cfg, dt, nsteps, kind, seed_all=None):
    qs = [torch.tensor(b.q, requires_grad=True) for b in state.blocks]
    mets = [(torch.tensor(m.kxyz), torch.tensor(m.vol)) for m in state.metrics]
    params, qinf = {}, {}
    for f, p in cfg.bc_params.items():
        p = dict(p)
        if "q_inf" in p:
            qinf[int(f)] = torch.tensor(np.asarray(p["q_inf"], dtype=np.float64), requires_grad=True)
            p["q_inf"] = qinf[int(f)]
        params[f] = p
    gas = cases.gas_dict(cfg.gas)
    cur = list(qs)
    for _ in range(nsteps):
        cur = OB.step(cur, mets, list(cfg.topology), params, gas, dt, limiter=cfg.scheme.limiter)
    if seed_all is not None:
        J = sum(c.sum() for c in cur)
    else:
        J = _torch_objective(cur, mets, kind)
    J.backward()
    # faces whose params carry q_inf but are not FARFIELD leave it unused
    return float(J.detach()), [q.grad.numpy() for q in qs], {
        f: (t.grad.numpy() if t.grad is not None else np.zeros(5)) for f, t in qinf.items()}


def _c4_small(n=12, limiter="minmod"):
    q, xyz_ext, cfg = cases.case_c4(n=n)
    jac, k, vol = OM.compute_metrics(xyz_ext)
    m = cases.crop_metrics(jac, k, vol)
    state = State(blocks=(Block(q=q, xyz=cases.crop_xyz(xyz_ext)),), metrics=(m,))
    if limiter != "minmod":
        cfg = Config(gas=cfg.gas, scheme=SchemeChoice("muscl2", "roe", 3, limiter), bc_params=cfg.bc_params,
                     topology=cfg.topology, ghost=2)
    dt = OM.cfl_dt(q, m.kxyz, m.vol, cases.gas_dict(cfg.gas), 0.5)
    return state, cfg, dt
