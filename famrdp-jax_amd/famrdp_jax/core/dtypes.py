"""Float64 guard (famrdp_jax/core/dtypes.py:5-17): every array of the data
contract must be float64; anything else raises TypeError."""


def _is_f64(arr) -> bool:
    dt = getattr(arr, "dtype", None)
    if dt is None:
        return True
    name = str(dt)
    return name in ("float64", "torch.float64")


def assert_float64(arr, *, name: str) -> None:
    if not _is_f64(arr):
        raise TypeError(f"{name!r} must be float64, got {arr.dtype}.")


def _leaves(tree):
    if isinstance(tree, dict):
        for v in tree.values():
            yield from _leaves(v)
    elif isinstance(tree, (list, tuple)):
        for v in tree:
            yield from _leaves(v)
    else:
        yield tree


def check_array_f64(tree) -> None:
    for i, leaf in enumerate(_leaves(tree)):
        if hasattr(leaf, "dtype"):
            assert_float64(leaf, name=f"leaf[{i}]")
