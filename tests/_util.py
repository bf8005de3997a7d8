"""Helpers shared by the tests."""
import numpy as np


def rel_linf(a, b):
    """Relative L-infinity error per conserved variable (max over variables of
    max|a-b| / max|b|), the metric of the step tolerances."""
    a = np.asarray(a); b = np.asarray(b)
    if a.ndim >= 4 and a.shape[0] == 5:
        return max(float(np.abs(a[v] - b[v]).max() / max(np.abs(b[v]).max(), 1e-300)) for v in range(5))
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def to_dev(state):
    """Copy a numpy State onto cuda:0 (torch tensors)."""
    import torch
    from famrdp_jax.core.types import Block, Metrics, State
    cu = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    return State(blocks=tuple(Block(q=cu(b.q), xyz=cu(b.xyz)) for b in state.blocks),
                 metrics=tuple(Metrics(jac=cu(m.jac), kxyz=cu(m.kxyz), vol=cu(m.vol)) for m in state.metrics),
                 t=state.t, step=state.step)
