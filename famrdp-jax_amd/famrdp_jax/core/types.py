"""Data contract: the same objects as famrdp_jax/core/types.py:26-189.

Array fields hold float64 ``torch`` tensors (device memory on the GPU path)
or numpy arrays (host data handed to the GPU path, which copies them in).
Layout is the reference's: ``q`` (5, ni, nj, nk), ``xyz`` (3, ni, nj, nk),
``kxyz`` (3, 3, ni, nj, nk) with ``kxyz[l, m] = d xi_l / d x_m``, ``jac`` and
``vol`` (ni, nj, nk); C order, k unit stride.

``Block``/``Metrics``/``State`` keep the reference's flatten/unflatten
protocol (leaf order and aux data) and are registered as JAX pytrees when
JAX is importable; the static configuration types are frozen dataclasses.
"""
from __future__ import annotations

from dataclasses import dataclass

from .dtypes import check_array_f64


@dataclass
class Block:
    """q: (5, ni, nj, nk) conservative [rho, rho u, rho v, rho w, rho E];
    xyz: (3, ni, nj, nk) coordinates.  Both float64 (types.py:26-56)."""
    q: object
    xyz: object

    def __post_init__(self):
        check_array_f64({"q": self.q, "xyz": self.xyz})

    def tree_flatten(self):
        return (self.q, self.xyz), {}

    @classmethod
    def tree_unflatten(cls, aux, leaves):
        obj = object.__new__(cls)
        object.__setattr__(obj, "q", leaves[0])
        object.__setattr__(obj, "xyz", leaves[1])
        return obj


@dataclass
class Metrics:
    """jac (ni,nj,nk), kxyz (3,3,ni,nj,nk), vol (ni,nj,nk); float64 (types.py:63-93)."""
    jac: object
    kxyz: object
    vol: object

    def __post_init__(self):
        check_array_f64({"jac": self.jac, "kxyz": self.kxyz, "vol": self.vol})

    def tree_flatten(self):
        return (self.jac, self.kxyz, self.vol), {}

    @classmethod
    def tree_unflatten(cls, aux, leaves):
        obj = object.__new__(cls)
        object.__setattr__(obj, "jac", leaves[0])
        object.__setattr__(obj, "kxyz", leaves[1])
        object.__setattr__(obj, "vol", leaves[2])
        return obj


@dataclass
class State:
    """blocks, metrics (tuples), t (leaf), step (aux) (types.py:100-139)."""
    blocks: tuple
    metrics: tuple
    t: float = 0.0
    step: int = 0

    def tree_flatten(self):
        leaves = (*self.blocks, *self.metrics, self.t)
        aux = {"n_blocks": len(self.blocks), "n_metrics": len(self.metrics), "step": self.step}
        return leaves, aux

    @classmethod
    def tree_unflatten(cls, aux, leaves):
        nb, nm = aux["n_blocks"], aux["n_metrics"]
        obj = object.__new__(cls)
        object.__setattr__(obj, "blocks", tuple(leaves[:nb]))
        object.__setattr__(obj, "metrics", tuple(leaves[nb:nb + nm]))
        object.__setattr__(obj, "t", leaves[nb + nm])
        object.__setattr__(obj, "step", aux["step"])
        return obj


try:  # JAX is optional: register the pytrees where it exists
    import jax as _jax

    for _cls in (Block, Metrics, State):
        _jax.tree_util.register_pytree_node_class(_cls)
except Exception:  # jax absent (this image) -> plain dataclasses
    pass


@dataclass(frozen=True)
class BlockTopology:
    """neighbors: Face -> None | (block_id, Face, orient); bc_type: Face -> BCType."""
    block_id: int
    neighbors: dict
    bc_type: dict


@dataclass(frozen=True)
class GasModel:
    """Ideal gas + Sutherland; defaults as types.py:160-167."""
    gamma: float = 1.4
    R: float = 287.058
    mu0: float = 1.71608e-5
    T0: float = 273.15
    Ts: float = 110.4
    Pr_lam: float = 0.72


@dataclass(frozen=True)
class SchemeChoice:
    reconstruct: str
    flux: str
    rk_order: int
    limiter: str


@dataclass(frozen=True)
class Config:
    gas: GasModel
    scheme: SchemeChoice
    bc_params: dict
    topology: tuple
    ghost: int = 2
    nan_check: bool = True
    debug: bool = False
    cfl: float = 1.0
