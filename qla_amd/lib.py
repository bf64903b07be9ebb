"""ctypes binding of ``libqla_mi355x.so``, the gfx950 QLA hot-path library.

The binding mirrors the C ABI one for one (names and argument order of the
reference QLA symbols); it only converts Python integers/pointers.  There is
no CPU fallback: if the shared library is missing or cannot be loaded this
module raises, so nothing can silently run a different code path.

torch is imported before the library is opened so that the HIP runtime torch
already loaded (same SONAME, libamdhip64.so.7) is the one the library binds
to; device pointers from torch allocations are then valid in library calls.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

from .abi_gen import all_symbols

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libqla_mi355x.so"
HEADERS = tuple(HERE.parent / "include" / h for h in ("qla_mi355x.h", "qla_mi355x_sites.h", "qla_mi355x_fields.h",
                                                     "qla_mi355x_matfuncs.h", "qla_mi355x_vecops.h"))

vp = C.c_void_p
i32 = C.c_int

# argument kinds of each exported symbol: 'p' pointer, 'i' int, 'r' void return
_SIGS: dict[str, tuple[type, list]] = {
    # compressed-precision extension
    "QLA_MI355X_F3_M_pack12": (None, [vp, i32, vp, i32]),
    "QLA_MI355X_F3_M_unpack12": (None, [vp, i32, vp, i32]),
    "QLA_MI355X_F3_D_packh": (None, [vp, vp, i32]),
    "QLA_MI355X_F3_D_unpackh": (None, [vp, vp, i32]),
    "QLA_MI355X_F3_H_packh": (None, [vp, vp, i32]),
    "QLA_MI355X_F3_H_unpackh": (None, [vp, vp, i32]),
    "QLA_MI355X_F3_Hh_veq_spproj_M12_times_pDh": (None, [vp, vp, i32, vp, i32, i32, i32]),
    "QLA_MI355X_F3_Dh_vpeq_sprecon_M12_times_pHh": (None, [vp, vp, i32, vp, i32, i32, i32]),
    "QLA_MI355X_DF3_r_veq_norm2_Dh": (None, [vp, vp, i32]),
    "QLA_MI355X_DF3_c_veq_Dh_dot_Dh": (None, [vp, vp, vp, i32]),
    # Cayley-Hamilton matrix functions with FP64 coefficients
    "QLA_MI355X_F3_M_veq_expA_M_ch": (None, [vp, vp, i32]),
    "QLA_MI355X_F3_M_xeq_expA_M_ch": (None, [vp, vp, vp, i32]),
    "QLA_MI355X_F3_M_veq_sqrtPH_M_ch": (None, [vp, vp, i32]),
    "QLA_MI355X_F3_M_xeq_sqrtPH_M_ch": (None, [vp, vp, vp, i32]),
    "QLA_MI355X_F3_M_veq_invsqrtPH_M_ch": (None, [vp, vp, i32]),
    "QLA_MI355X_F3_M_xeq_invsqrtPH_M_ch": (None, [vp, vp, vp, i32]),
    # allocation-free reduction
    "QLA_MI355X_DF3_r_veq_norm2_D_noalloc": (None, [vp, vp, i32]),
    # ragged batches
    "QLA_MI355X_F3_batch_create": (vp, [vp, i32]),
    "QLA_MI355X_batch_run": (None, [vp]),
    "QLA_MI355X_batch_capture": (i32, [vp]),
    "QLA_MI355X_batch_launch_graph": (None, [vp]),
    "QLA_MI355X_batch_autotune": (C.c_float, [vp, i32]),
    "QLA_MI355X_batch_get_shape": (None, [vp, vp, vp, vp]),
    "QLA_MI355X_batch_set_shape": (i32, [vp, i32, i32, i32]),
    "QLA_MI355X_batch_destroy": (None, [vp]),
    # control
    "QLA_MI355X_set_stream": (None, [vp]),
    "QLA_MI355X_get_stream": (vp, []),
    "QLA_MI355X_set_sync": (None, [i32]),
    "QLA_MI355X_set_arith_order": (None, [i32]),
    "QLA_MI355X_get_arith_order": (i32, []),
    "QLA_MI355X_synchronize": (None, []),
    "QLA_MI355X_last_error": (i32, []),
    "QLA_MI355X_last_error_string": (C.c_char_p, []),
    "QLA_MI355X_version": (C.c_char_p, []),
    "QLA_MI355X_set_tiles_per_wave": (None, [i32]),
    "QLA_MI355X_finalize": (None, []),
    "QLA_MI355X_scratch_buffers": (i32, []),
}
# the generated families: the Wilson-hop building blocks (qla_mi355x_sites.h),
# the reductions and conversions (qla_mi355x_fields.h), the colour-matrix
# functions (qla_mi355x_matfuncs.h) and the field linear algebra
# (qla_mi355x_vecops.h)
_SIGS.update({v.name: (None, [vp if k == "p" else i32 for k in v.ctypes_kinds()]) for v in all_symbols()})

class F3Job(C.Structure):
    """QLA_MI355X_F3_job (include/qla_mi355x.h)."""
    _fields_ = [("op", C.c_int), ("mu", C.c_int), ("sign", C.c_int), ("n", C.c_int),
                ("r", C.c_void_p), ("a", C.c_void_p), ("b", C.c_void_p)]


_lib: C.CDLL | None = None


def header_symbols() -> list[str]:
    """Names of every function declared in the public headers."""
    decl = r"^\s*(?:const\s+)?(?:void|int|char|float|double|QLA_MI355X_batch)\s*\**\s*(QLA_\w+)\s*\("
    out = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        out.update(re.findall(decl, text, flags=re.M))
    return sorted(out)


def load() -> C.CDLL:
    """Open the library (once) and attach argument types.  Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (bind to the HIP runtime torch loaded)

    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -m qla_amd.build` (there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check() -> None:
    """Raise if the calling thread saw a HIP error inside the library."""
    lib = load()
    e = lib.QLA_MI355X_last_error()
    if e != 0:
        raise RuntimeError(f"qla_mi355x HIP error {e}")
