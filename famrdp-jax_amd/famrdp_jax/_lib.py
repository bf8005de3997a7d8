"""Loader for the HIP library (``_lib/libfamrdp_hip.so``) and device helpers.

The library is the only compute path: if it is missing, or no GPU is
visible, every operation raises ``RuntimeError`` instead of computing on the
CPU.  Arrays cross the C ABI as raw device pointers of contiguous float64
torch tensors, with torch's current HIP stream.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libfamrdp_hip.so")

_lib = None

c_double_p = ctypes.c_void_p
i32 = ctypes.c_int32
i64 = ctypes.c_int64
f64 = ctypes.c_double
vp = ctypes.c_void_p


class Gas(ctypes.Structure):
    _fields_ = [("gamma", f64), ("R", f64), ("mu0", f64), ("T0", f64), ("Ts", f64), ("Pr_lam", f64)]


class Cut(ctypes.Structure):
    _fields_ = [("tgt_block", i32), ("tgt_face", i32), ("src_block", i32), ("src_face", i32), ("orient", i32)]


class Bc(ctypes.Structure):
    _fields_ = [("block", i32), ("face", i32), ("kind", i32), ("pad_", i32), ("T_wall", f64), ("q_inf", f64 * 5)]


# (name, argtypes) of every symbol in include/famrdp_capi.h
# stage-kernel variants of famrdp_plan_autotune / famrdp_plan_set_tile:
# 0-4 and 10-13 tile shapes of the fused kernel (10-13 and 4 re-stage the
# primitives), 5-9 of the split pair, 14-18 of the i-marching kernel (the
# fused kernel's arithmetic: the same bits)
NUM_TILE_VARIANTS = 19
FUSED_VARIANTS = (0, 1, 2, 3, 4, 10, 11, 12, 13, 14, 15, 16, 17, 18)
MARCH_VARIANTS = (14, 15, 16, 17, 18)

SIGNATURES = {
    "famrdp_last_error": ([], ctypes.c_char_p),
    "famrdp_version": ([], ctypes.c_char_p),
    "famrdp_cons_to_prim": ([vp, vp, i64, f64, vp], i32),
    "famrdp_prim_to_cons": ([vp, vp, i64, f64, vp], i32),
    "famrdp_sutherland": ([vp, vp, i64, f64, f64, f64, vp], i32),
    "famrdp_speed_of_sound": ([vp, vp, vp, i64, f64, vp], i32),
    "famrdp_limiter": ([vp, vp, vp, i64, i32, vp], i32),
    "famrdp_fastmath_probe": ([vp, vp, vp, vp, i64, vp], i32),
    "famrdp_fastmath_probe1": ([vp, vp, vp, i64, vp], i32),
    "famrdp_muscl2": ([vp, vp, vp, i32, i32, i32, i32, i32, i32, vp], i32),
    "famrdp_roe_flux": ([vp, vp, vp, vp, i64, f64, f64, vp], i32),
    "famrdp_invscd_rhs": ([vp, vp, vp, vp, i32, i32, i32, f64, i32, vp], i32),
    "famrdp_viscous_rhs": ([vp, vp, vp, i32, i32, i32, ctypes.POINTER(Gas), vp, vp], i32),
    "famrdp_axpby": ([vp, f64, vp, f64, vp, f64, vp, i64, i32, vp], i32),
    "famrdp_compute_metrics": ([vp, vp, vp, vp, i32, i32, i32, vp], i32),
    "famrdp_plan_create": ([ctypes.POINTER(vp), i32, ctypes.POINTER(i32), i32, ctypes.POINTER(Cut), i32,
                            ctypes.POINTER(Bc), ctypes.POINTER(Gas), i32, i32, i32], i32),
    "famrdp_plan_destroy": ([vp], i32),
    "famrdp_step": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), f64, i32, vp], i32),
    "famrdp_step_graph": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), f64, i32, i32, vp], i32),
    "famrdp_step_stages": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), f64, i32, ctypes.POINTER(vp),
                            ctypes.POINTER(vp), vp], i32),
    "famrdp_plan_autotune": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), f64, i32, i32,
                              ctypes.POINTER(f64), vp], i32),
    "famrdp_plan_set_tile": ([vp, i32], i32),
    "famrdp_plan_get_tile": ([vp, ctypes.POINTER(i32), ctypes.POINTER(i32)], i32),
    "famrdp_plan_timing": ([vp, i32, ctypes.POINTER(f64), ctypes.POINTER(i64)], i32),
    "famrdp_plan_set_block_params": ([vp, ctypes.POINTER(f64), ctypes.POINTER(f64)], i32),
    "famrdp_step_vjp": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), f64, vp,
                         i32, vp], i32),
    "famrdp_step_vjp_stages": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp),
                                ctypes.POINTER(vp), ctypes.POINTER(vp), f64, vp, i32, vp], i32),
    "famrdp_plan_stage_states": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp), vp], i32),
    "famrdp_objective": ([i32, vp, vp, vp, i32, i32, i32, i32, f64, vp, vp, vp, vp], i32),
    "famrdp_halo_exchange": ([vp, ctypes.POINTER(vp), vp], i32),
    "famrdp_apply_bc": ([vp, ctypes.POINTER(vp), vp], i32),
    "famrdp_cfl_dt": ([vp, vp, vp, i32, i32, i32, i32, ctypes.POINTER(Gas), f64, vp, vp, vp], i32),
    "famrdp_cfl_dt_noalloc": ([vp, vp, vp, i32, i32, i32, i32, ctypes.POINTER(Gas), f64, vp, vp], i32),
    "famrdp_plan_check_state": ([vp, ctypes.POINTER(vp), vp, vp], i32),
}

E_INVALID, E_HIP, E_UNSUPPORTED, E_KEY = -1, -2, -3, -4
LIMITERS = {"minmod": 0, "vanleer": 1}


def load(path: str = LIB_PATH):
    """Load the library (no GPU needed) and bind every declared symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"famrdp HIP library not built: {path} is missing (run __graft_entry__.build())")
    # torch first: its bundled HIP runtime (soname libamdhip64.so.7) then also
    # serves this library.  Loaded the other way round, the process would hold
    # two HIP runtimes and torch's kernels would fail to launch.
    import torch  # noqa: F401
    lib = ctypes.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc == 0:
        return
    msg = (_lib.famrdp_last_error() or b"").decode()
    if rc == E_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == E_KEY:
        raise KeyError(msg)
    if rc == E_INVALID:
        raise ValueError(msg)
    raise RuntimeError(f"famrdp HIP error: {msg}")


def lib():
    return load()


# ---------------------------------------------------------------------------
# device helpers
# ---------------------------------------------------------------------------

def _torch():
    import torch
    return torch


def require_gpu():
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("famrdp_jax computes on an MI355X through HIP; no GPU is visible")
    load()


def device_f64(x):
    """Contiguous float64 tensor on the current GPU (copies host data in)."""
    torch = _torch()
    require_gpu()
    if isinstance(x, torch.Tensor):
        if x.dtype != torch.float64:
            raise TypeError(f"expected float64, got {x.dtype}")
        return x.to("cuda").contiguous()
    arr = np.asarray(x)
    if arr.dtype != np.float64:
        arr = arr.astype(np.float64) if arr.dtype.kind in "iu" or arr.ndim == 0 else arr
        if arr.dtype != np.float64:
            raise TypeError(f"expected float64, got {arr.dtype}")
    return torch.from_numpy(np.ascontiguousarray(arr)).to("cuda")


def like_input(t, ref):
    """Return ``t`` (device tensor) in the container type of ``ref``: torch
    tensors stay on the device, anything else comes back as numpy."""
    torch = _torch()
    if isinstance(ref, torch.Tensor):
        return t
    return t.detach().cpu().numpy()


def ptr(t) -> int:
    return t.data_ptr() if t is not None else 0


def stream() -> int:
    return _torch().cuda.current_stream().cuda_stream


def gas_struct(gas) -> Gas:
    return Gas(float(gas.gamma), float(gas.R), float(gas.mu0), float(gas.T0), float(gas.Ts), float(gas.Pr_lam))


def limiter_code(name: str) -> int:
    if name not in LIMITERS:
        raise KeyError(name)
    return LIMITERS[name]
