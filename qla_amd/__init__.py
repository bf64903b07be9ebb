"""qla_amd -- MI355X-native (gfx950) implementation of the QLA site-loop hot path.

The product is the C-ABI shared library ``qla_amd/libqla_mi355x.so`` (declared
in ``include/qla_mi355x.h``): the reference QLA symbols with their argument
lists, running hand-written HIP kernels on device-resident arrays.  This
Python package is a thin host-side convenience used by the tests and the
benchmark: it loads the library (raising if it is not built -- there is no
CPU fallback) and passes torch device tensors to the C entry points.

    import torch, qla_amd
    qla_amd.call("QLA_F3_H_veq_spproj_M_times_pD", r, a, bptr, mu, sign, n)

Tensor arguments are passed as their device addresses; pointer arrays are
int64 device tensors of record addresses (``pointer_array``).
"""
from __future__ import annotations

from .lib import LIB_PATH, check, header_symbols, load  # noqa: F401

REC_REALS = {"V": 6, "H": 12, "D": 24, "M": 18}


def _arg(x):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return x.data_ptr()
    except ImportError:  # pragma: no cover
        pass
    return x


def call(name: str, *args):
    """Call C entry point `name` with tensors replaced by their data pointers."""
    lib = load()
    return getattr(lib, name)(*[_arg(a) for a in args])


def pointer_array(field, idx):
    """int64 device tensor: address of record idx[i] of `field` (dim 0 = sites)."""
    import torch
    rec = field.stride(0) * field.element_size()
    idx = torch.as_tensor(idx, device=field.device, dtype=torch.int64)
    return field.data_ptr() + rec * idx


def set_stream(stream) -> None:
    """Launch on a torch stream (or raw hipStream_t int); None = default stream."""
    s = getattr(stream, "cuda_stream", stream)
    load().QLA_MI355X_set_stream(s or None)


def set_sync(on: bool) -> None:
    load().QLA_MI355X_set_sync(1 if on else 0)


def version() -> str:
    return load().QLA_MI355X_version().decode()
