"""The oracle: CPU restatement of the QLA 1.9.0 hot path (``liboracle.so``).

TEST INFRASTRUCTURE ONLY.  Imported by the tests, by ``__graft_entry__.smoke``
and by the CPU-baseline leg of ``bench.py`` -- as the checker, never as the
thing measured on the GPU or shipped.  The product library (qla_amd) never
imports this package.

Arrays are numpy, site-major AoS with a trailing (re, im) axis; see
``qla_amd/lattice.py`` for the shapes.  Pointer-array arguments of the QLA
functions are formed here from integer site lists: record i of ``b`` is at
``b.ctypes.data + rec_bytes * idx[i]``.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

EQ, PEQ, MEQ, EQM = 0, 1, 2, 3

_lib = None


def build() -> Path:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    vp, i32 = C.c_void_p, C.c_int
    for pre in ("F3", "D3"):
        getattr(lib, f"orc_{pre}_V_eqop_M_times_V").argtypes = [i32, i32, vp, vp, vp, vp, vp, i32]
        getattr(lib, f"orc_{pre}_H_eqop_spproj_M_times_pD").argtypes = [i32, i32, vp, vp, vp, i32, i32, vp, i32]
        getattr(lib, f"orc_{pre}_D_eqop_sprecon_M_times_pH").argtypes = [i32, i32, vp, vp, vp, i32, i32, vp, i32]
    for p in ("F", "D"):
        f = getattr(lib, f"orc_{p}_norm2")
        f.argtypes, f.restype = [vp, i32, i32], C.c_double
        f = getattr(lib, f"orc_{p}_norm2_ld")
        f.argtypes, f.restype = [vp, i32, i32], C.c_longdouble
        getattr(lib, f"orc_{p}_dot").argtypes = [vp, vp, vp, vp, i32, i32]
        getattr(lib, f"orc_{p}_dot_ld").argtypes = [vp, vp, vp, vp, i32, i32]
        for fn in ("exp", "sqrt", "invsqrt"):
            getattr(lib, f"orc_{p}_M_veq_{fn}_M").argtypes = [i32, vp, vp, i32]
    lib.orc_D_spproj.argtypes = [vp, vp, i32, i32]
    lib.orc_D_sprecon.argtypes = [vp, vp, i32, i32]
    lib.orc_set_threads.argtypes = [i32]
    lib.orc_get_threads.restype = i32
    _lib = lib
    return lib


def _prefix(a: np.ndarray) -> str:
    if a.dtype == np.float32:
        return "F3"
    if a.dtype == np.float64:
        return "D3"
    raise TypeError(a.dtype)


def _ptr(a):
    return None if a is None else a.ctypes.data


def _need(a: np.ndarray):
    assert a.flags.c_contiguous
    return a


def pointer_array(b: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """Host pointer array: address of record idx[i] of b (uint64)."""
    rec = b.strides[0]
    return (np.uint64(b.ctypes.data) + np.asarray(idx, dtype=np.uint64) * np.uint64(rec)).astype(np.uint64)


def set_threads(n: int) -> None:
    load().orc_set_threads(int(n))


def threads() -> int:
    return load().orc_get_threads()


def V_eqop_M_times_V(eqop, r, a, b=None, b_idx=None, index=None, adj=False, n=None):
    """r eqop a * b (b contiguous) or a * (*pb) with pb = &b[b_idx[i]]."""
    lib = load()
    n = a.shape[0] if n is None else n
    pb = None if b_idx is None else pointer_array(b, b_idx)
    ix = None if index is None else np.ascontiguousarray(index, dtype=np.int32)
    getattr(lib, f"orc_{_prefix(a)}_V_eqop_M_times_V")(
        eqop, int(adj), _ptr(_need(r)), _ptr(a), None if pb is not None else _ptr(b), _ptr(pb), _ptr(ix),
        len(ix) if ix is not None else n)
    return r


def H_eqop_spproj_M_times_pD(eqop, r, a, d, d_idx, mu, sign, index=None, adj=False):
    lib = load()
    pb = pointer_array(d, d_idx)
    ix = None if index is None else np.ascontiguousarray(index, dtype=np.int32)
    n = len(ix) if ix is not None else len(d_idx)
    getattr(lib, f"orc_{_prefix(a)}_H_eqop_spproj_M_times_pD")(
        eqop, int(adj), _ptr(_need(r)), _ptr(a), _ptr(pb), mu, sign, _ptr(ix), n)
    return r


def D_eqop_sprecon_M_times_pH(eqop, r, a, h, h_idx, mu, sign, index=None, adj=False):
    lib = load()
    pb = pointer_array(h, h_idx)
    ix = None if index is None else np.ascontiguousarray(index, dtype=np.int32)
    n = len(ix) if ix is not None else len(h_idx)
    getattr(lib, f"orc_{_prefix(a)}_D_eqop_sprecon_M_times_pH")(
        eqop, int(adj), _ptr(_need(r)), _ptr(a), _ptr(pb), mu, sign, _ptr(ix), n)
    return r


def norm2(a: np.ndarray) -> float:
    """Promoted (double) norm2 sum over all complex components of a field."""
    lib = load()
    n = a.shape[0]
    ncomp = a[0].size // 2
    p = "F" if a.dtype == np.float32 else "D"
    return getattr(lib, f"orc_{p}_norm2")(_ptr(_need(a)), ncomp, n)


def norm2_ld(a: np.ndarray) -> float:
    lib = load()
    p = "F" if a.dtype == np.float32 else "D"
    return getattr(lib, f"orc_{p}_norm2_ld")(_ptr(_need(a)), a[0].size // 2, a.shape[0])


def dot(a: np.ndarray, b: np.ndarray) -> complex:
    lib = load()
    p = "F" if a.dtype == np.float32 else "D"
    re, im = C.c_double(), C.c_double()
    getattr(lib, f"orc_{p}_dot")(C.byref(re), C.byref(im), _ptr(_need(a)), _ptr(_need(b)),
                                  a[0].size // 2, a.shape[0])
    return complex(re.value, im.value)


def dot_ld(a: np.ndarray, b: np.ndarray):
    lib = load()
    p = "F" if a.dtype == np.float32 else "D"
    re, im = C.c_longdouble(), C.c_longdouble()
    getattr(lib, f"orc_{p}_dot_ld")(C.byref(re), C.byref(im), _ptr(_need(a)), _ptr(_need(b)),
                                     a[0].size // 2, a.shape[0])
    return re.value, im.value


def spproj1(d: np.ndarray, mu: int, sign: int) -> np.ndarray:
    """Single-site spin projection in double: d (3,4,2) -> h (3,2,2)."""
    h = np.zeros((3, 2, 2))
    load().orc_D_spproj(_ptr(h), _ptr(np.ascontiguousarray(d, dtype=np.float64)), mu, sign)
    return h


def sprecon1(h: np.ndarray, mu: int, sign: int) -> np.ndarray:
    d = np.zeros((3, 4, 2))
    load().orc_D_sprecon(_ptr(d), _ptr(np.ascontiguousarray(h, dtype=np.float64)), mu, sign)
    return d


def matfn(fn: str, a: np.ndarray) -> np.ndarray:
    """M_veq_{exp,sqrt,invsqrt}_M on (n, nc, nc, 2) matrices."""
    lib = load()
    nc = a.shape[1]
    p = "F" if a.dtype == np.float32 else "D"
    r = np.empty_like(a)
    getattr(lib, f"orc_{p}_M_veq_{fn}_M")(nc, _ptr(r), _ptr(_need(a)), a.shape[0])
    return r
