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
        getattr(lib, f"orc_{pre}_site").argtypes = [i32, i32, i32, vp, vp, vp, vp, i32, i32, vp, i32]
        getattr(lib, f"orc_{pre}_multi").argtypes = [i32, i32, vp, vp, vp, vp, i32, vp, i32]
    lib.orc_F3_sse_site.argtypes = [i32, i32, i32, vp, vp, vp, i32, i32, i32, i32, i32, i32]
    lib.orc_D3_sse2_site.argtypes = [i32, vp, vp, vp, i32]
    for p in ("F", "D"):
        f = getattr(lib, f"orc_{p}_norm2")
        f.argtypes, f.restype = [vp, i32, i32], C.c_double
        f = getattr(lib, f"orc_{p}_norm2_ld")
        f.argtypes, f.restype = [vp, i32, i32], C.c_longdouble
        f = getattr(lib, f"orc_{p}_norm2_terms_ld")
        f.argtypes, f.restype = [vp, i32, i32], C.c_longdouble
        getattr(lib, f"orc_{p}_dot_terms_ld").argtypes = [vp, vp, vp, vp, vp, i32, i32]
        getattr(lib, f"orc_{p}_dot").argtypes = [vp, vp, vp, vp, i32, i32]
        getattr(lib, f"orc_{p}_dot_ld").argtypes = [vp, vp, vp, vp, i32, i32]
        for fn in ("exp", "sqrt", "invsqrt"):
            getattr(lib, f"orc_{p}_M_veq_{fn}_M").argtypes = [i32, vp, vp, i32]
        getattr(lib, f"orc_{p}_matfunc").argtypes = [i32, i32, vp, vp, vp, i32]
        real = C.c_float if p == "F" else C.c_double
        getattr(lib, f"orc_{p}_vecop").argtypes = [i32, i32, i32, i32, i32, vp, vp, vp, vp, real, real, vp, i32]
        getattr(lib, f"orc_{p}_fill").argtypes = [i32, i32, i32, vp, vp, vp, i32]
        getattr(lib, f"orc_{p}_gamma").argtypes = [i32, i32, i32, vp, vp, vp, vp, i32]
        getattr(lib, f"orc_{p}_local").argtypes = [i32, i32, i32, i32, vp, vp, vp, vp, vp, i32]
        getattr(lib, f"orc_{p}_sum").argtypes = [i32, vp, vp, vp, i32]
        getattr(lib, f"orc_{p}_fieldmul").argtypes = [i32, i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, i32]
        getattr(lib, f"orc_{p}_outer").argtypes = [i32, vp, vp, vp, vp, vp, i32]
        getattr(lib, f"orc_{p}_pmul").argtypes = [i32, i32, i32, i32, vp, vp, vp, vp, vp, i32]
        getattr(lib, f"orc_{p}_comp").argtypes = [i32, i32, vp, vp, vp, vp, vp, i32]
        getattr(lib, f"orc_{p}_sconv").argtypes = [i32, vp, vp, vp, vp, vp, i32]
        getattr(lib, f"orc_{p}_sfn").argtypes = [i32, vp, vp, vp, vp, vp, i32]
        getattr(lib, f"orc_{p}_mask").argtypes = [i32, i32, vp, vp, vp, vp, vp, i32]
    lib.orc_int.argtypes = [i32, i32, vp, vp, vp, vp, i32, vp, i32]
    lib.orc_convert_DF.argtypes = [vp, vp, C.c_size_t]
    lib.orc_convert_FD.argtypes = [vp, vp, C.c_size_t]
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


MATVEC, SPPROJ, SPRECON, WILSON = 0, 1, 2, 3
PROJ, RECON, PROJRECON = 5, 6, 7        # matrix-free: H = spproj(b), D = sprecon(b), D = sprecon(spproj(b))
MAT_H, MAT_D, MAT_M, MAT_MA = 8, 9, 10, 11   # H, D, M eqop A*b; M eqop A*b^dagger
REC_REALS = {MATVEC: 6, SPPROJ: 12, SPRECON: 24, WILSON: 24, PROJ: 12, RECON: 24, PROJRECON: 24,
             MAT_H: 12, MAT_D: 24, MAT_M: 18, MAT_MA: 18}


def _records(x: np.ndarray, idx) -> np.ndarray:
    """Pointer array to records of x: identity for a field operand (idx None),
    else record idx[i] (a pointer-array operand)."""
    return pointer_array(x, np.arange(x.shape[0]) if idx is None else idx)


def _index(index):
    return None if index is None else np.ascontiguousarray(index, dtype=np.int32)


def site(kind, eqop, r, a, b, mu=0, sign=1, index=None, adj=False, a_idx=None, b_idx=None, n=None,
         ir=None, ia=None, ib=None):
    """One QLA site call on host arrays (orc_*_site).

    kind 0 V eqop A*B, 1 H eqop A*spproj(B), 2 D eqop sprecon(A*B),
    3 D eqop sprecon(A*spproj(B)); A = a or a^dagger (adj); the matrix-free
    kinds 5 H eqop spproj(B), 6 D eqop sprecon(B), 7 D eqop
    sprecon(spproj(B)) take a = None.  a_idx / b_idx
    None: the operand is a field (record i, or index[i] in the x-form); else
    the operand is a pointer array whose entry s points at record a_idx[s].
    ir / ia / ib: per-operand index arrays of the QLA_INDEX_ALL forms (step i
    uses r[ir[i]], a[ia[i]], b[ib[i]]); an indexed field operand is the
    pointer array of the records it names.  n defaults to len(index)
    (x-form) or the number of destination records.
    """
    lib = load()
    pa = None if a is None else _records(a, ia if ia is not None else a_idx)
    pb = _records(b, ib if ib is not None else b_idx)
    ix = _index(index)
    rix = _index(ir)
    if n is None:
        n = len(ix) if ix is not None else (len(rix) if rix is not None else r.shape[0])
    getattr(lib, f"orc_{_prefix(b)}_site")(kind, eqop, int(adj), _ptr(_need(r)), _ptr(rix), _ptr(pa), _ptr(pb),
                                           mu, sign, _ptr(ix), n)
    return r


def multi(eqop, r, a_list, b_list, index=None, adj=False, a_idx=None, b_idx=None, n=None,
          ir=None, ia=None, ib=None):
    """Multi-source r eqop sum_d A_d * B_d (orc_*_multi); a_idx / b_idx are
    per-term index lists (pointer-array operands) or None (fields); ir / ia /
    ib the per-operand index arrays of the QLA_INDEX_ALL forms (shared by
    the terms)."""
    lib = load()
    nd = len(a_list)
    pas = [_records(a, ia if ia is not None else (None if a_idx is None else a_idx[d]))
           for d, a in enumerate(a_list)]
    pbs = [_records(b, ib if ib is not None else (None if b_idx is None else b_idx[d]))
           for d, b in enumerate(b_list)]
    pa = np.array([x.ctypes.data for x in pas], dtype=np.uint64)
    pb = np.array([x.ctypes.data for x in pbs], dtype=np.uint64)
    ix = _index(index)
    rix = _index(ir)
    if n is None:
        n = len(ix) if ix is not None else (len(rix) if rix is not None else r.shape[0])
    getattr(lib, f"orc_{_prefix(r)}_multi")(eqop, int(adj), _ptr(_need(r)), _ptr(rix), _ptr(pa), _ptr(pb), nd,
                                            _ptr(ix), n)
    return r


def V_eqop_M_times_V(eqop, r, a, b=None, b_idx=None, index=None, adj=False, n=None):
    """r eqop a * b (b a field) or a * (*pb) with pb[s] = &b[b_idx[s]]."""
    if n is None and index is None:
        n = a.shape[0]
    return site(MATVEC, eqop, r, a, b, index=index, adj=adj, b_idx=b_idx, n=n)


def H_eqop_spproj_M_times_pD(eqop, r, a, d, d_idx, mu, sign, index=None, adj=False):
    n = None if index is not None else len(d_idx)
    return site(SPPROJ, eqop, r, a, d, mu, sign, index=index, adj=adj, b_idx=d_idx, n=n)


def D_eqop_sprecon_M_times_pH(eqop, r, a, h, h_idx, mu, sign, index=None, adj=False):
    n = None if index is not None else len(h_idx)
    return site(SPRECON, eqop, r, a, h, mu, sign, index=index, adj=adj, b_idx=h_idx, n=n)


def D_eqop_spproj_M_times_pD(eqop, r, a, d, d_idx, mu, sign, index=None, adj=False):
    n = None if index is not None else len(d_idx)
    return site(WILSON, eqop, r, a, d, mu, sign, index=index, adj=adj, b_idx=d_idx, n=n)


# (kind, eqop, adj) of the symbols whose --enable-sse3 build rounds in
# another order (qla_oracle_sse.inc)
SSE_SYMBOLS = {"V_vpeq_M_times_pV": (MATVEC, PEQ, 0), "V_veq_M_times_pV": (MATVEC, EQ, 0),
               "V_veq_Ma_times_V": (MATVEC, EQ, 1), "D_vpeq_sprecon_M_times_pH": (SPRECON, PEQ, 0),
               "D_vpeq_spproj_M_times_pD": (WILSON, PEQ, 0)}


def sse_site(symbol, r, a, b, b_idx=None, mu=0, sign=1, r_al=-1, a_al=-1, b0_al=-1):
    """QLA_F3_<symbol> (a key of SSE_SYMBOLS) in the operation order of a QLA
    built --enable-sse3 (qla_oracle_sse.inc), in place on r.  b is the field
    whose records b_idx[i] the pointer array names (a field operand: b_idx
    None).  r_al / a_al / b0_al are the byte offsets mod 16 of r, a and b[0]
    in the call being modelled (the reference picks its order from them); -1
    takes these host arrays' own."""
    kind, eqop, adj = SSE_SYMBOLS[symbol]
    assert r.dtype == np.float32
    lib = load()
    pa = _records(a, None)
    pb = _records(b, b_idx)
    n = r.shape[0] if b_idx is None else len(b_idx)
    lib.orc_F3_sse_site(kind, eqop, adj, _ptr(_need(r)), _ptr(pa), _ptr(pb), mu, sign, n, r_al, a_al, b0_al)
    return r


# the D3 symbols an --enable-sse3 build takes from its SSE2 versions, with
# their eqop (qla_oracle_sse.inc)
SSE2_D3_SYMBOLS = {"V_veq_M_times_V": EQ, "V_veq_M_times_pV": EQ, "V_vpeq_M_times_V": PEQ,
                   "V_vpeq_M_times_pV": PEQ}


def sse2_d3_site(symbol, r, a, b, b_idx=None):
    """QLA_D3_<symbol> (a key of SSE2_D3_SYMBOLS) in the operation order of
    the SSE2 version an --enable-sse3 build links, in place on r.  b is the
    field whose records b_idx[i] the pointer array names (field operand:
    b_idx None)."""
    eqop = SSE2_D3_SYMBOLS[symbol]
    assert r.dtype == np.float64
    lib = load()
    pa = _records(a, None)
    pb = _records(b, b_idx)
    n = r.shape[0] if b_idx is None else len(b_idx)
    lib.orc_D3_sse2_site(eqop, _ptr(_need(r)), _ptr(pa), _ptr(pb), n)
    return r


def records(a: np.ndarray, index=None, a_idx=None, own=None) -> np.ndarray:
    """The records a QLA reduction or conversion reads, in step order i:
    a[s] (field) or a[a_idx[s]] (pointer array to records a_idx), s = index[i]
    (x-form) or i; with a per-operand index (QLA_INDEX_ALL forms) a[own[i]]."""
    if own is not None:
        return np.ascontiguousarray(a[np.asarray(own)])
    s = np.arange(a.shape[0]) if index is None else np.asarray(index)
    if a_idx is not None:
        s = np.asarray(a_idx)[s]
    return np.ascontiguousarray(a[s])


def convert(a: np.ndarray, dest: np.ndarray, index=None, a_idx=None, ir=None, ia=None) -> np.ndarray:
    """QLA_DF3 / QLA_FD3 T_{v,x}eq_[p]T: the destination record of step i
    (r[ir[i]], r[index[i]] or r[i]) = the source record of step i converted
    to the destination precision (orc_convert_DF / orc_convert_FD)."""
    lib = load()
    src = records(a, index, a_idx, own=ia)
    conv = np.empty(src.shape, dtype=dest.dtype)
    f = lib.orc_convert_DF if dest.dtype == np.float64 else lib.orc_convert_FD
    f(_ptr(conv), _ptr(src), src.size)
    if ir is not None:
        s = np.asarray(ir)
    else:
        s = np.arange(src.shape[0]) if index is None else np.asarray(index)
    dest[s] = conv
    return dest


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


def norm2_terms_ld(a: np.ndarray):
    """The reference's component norms (formed in the field's precision)
    summed in long double: numpy longdouble."""
    lib = load()
    p = "F" if a.dtype == np.float32 else "D"
    return np.longdouble(getattr(lib, f"orc_{p}_norm2_terms_ld")(_ptr(_need(a)), a[0].size // 2, a.shape[0]))


def dot_terms_ld(a: np.ndarray, b: np.ndarray):
    """The reference's double dot terms summed in long double:
    (re, im, sum of |terms|) as numpy longdouble."""
    lib = load()
    p = "F" if a.dtype == np.float32 else "D"
    re, im, sa = C.c_longdouble(), C.c_longdouble(), C.c_longdouble()
    getattr(lib, f"orc_{p}_dot_terms_ld")(C.byref(re), C.byref(im), C.byref(sa), _ptr(_need(a)), _ptr(_need(b)),
                                           a[0].size // 2, a.shape[0])
    return np.longdouble(re.value), np.longdouble(im.value), np.longdouble(sa.value)


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


# function codes of orc_{F,D}_matfunc (qla_oracle.h)
MATFUNC = {"exp": 0, "sqrt": 1, "invsqrt": 2, "det": 3, "inverse": 4, "inverse_V": 5, "inverse_M": 6,
           "expA": 7, "expTA": 8, "sqrtPH": 9, "invsqrtPH": 10, "log": 11, "eigenvalsH": 12}


def matfunc(fn: str, a: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    """One of the matrix functions on (n, nc, nc, 2) matrices ``a``.

    det -> (n, 2); inverse_V (b: (n, nc, 2) vectors) and eigenvalsH ->
    (n, nc, 2); the others (inverse_M with b: (n, nc, nc, 2)) -> matrices."""
    lib = load()
    n, nc = a.shape[0], a.shape[1]
    p = "F" if a.dtype == np.float32 else "D"
    if fn == "det":
        shape = (n, 2)
    elif fn in ("inverse_V", "eigenvalsH"):
        shape = (n, nc, 2)
    else:
        shape = a.shape
    if fn in ("inverse_V", "inverse_M"):
        if b is None or b.dtype != a.dtype or b.shape[0] != n:
            raise ValueError(f"{fn} needs a second operand of {n} records")
    if nc > 16:
        raise ValueError("nc <= 16")
    r = np.empty(shape, dtype=a.dtype)
    getattr(lib, f"orc_{p}_matfunc")(MATFUNC[fn], nc, _ptr(r), _ptr(_need(a)),
                                     _ptr(_need(b)) if b is not None else None, n)
    return r


# field linear algebra (orc_{F,D}_vecop, qla_oracle.h)
COPY, TIMES_I, RSCALE, CSCALE, PLUS, MINUS, R_PLUS, R_MINUS, C_PLUS, C_MINUS = range(10)
# transposition order of the square record types (M: 3 x 3; P: 12 x 12 over (colour, spin))
TRANSPOSE_DIM = {"M": 3, "P": 12}


def vecop(op, eqop, r, x, y=None, s=0.0, conj=False, dim=0, index=None, x_idx=None, y_idx=None,
          ir=None, ix=None, iy=None, n=None):
    """One field linear-algebra call on host arrays (orc_*_vecop).

    r eqop op(x, y): op COPY (conj: the conjugated source; with dim the
    transposed one, so conj + dim is the Hermitian conjugate), TIMES_I,
    RSCALE / CSCALE (times the real / complex constant s, conj and dim as
    for COPY), PLUS / MINUS (x +- y), R_PLUS / R_MINUS / C_PLUS / C_MINUS
    (s*x +- y).  x_idx / y_idx make the operand a pointer array to those
    records; ir / ix / iy are the per-operand indices of the QLA_INDEX_ALL
    forms; index the gang index.  Records are (n, ...) arrays with a
    trailing (re, im) axis."""
    lib = load()
    ncomp = r[0].size // 2
    px = _records(x, ix if ix is not None else x_idx)
    py = None if y is None else _records(y, iy if iy is not None else y_idx)
    ixx = _index(index)
    rix = _index(ir)
    if n is None:
        n = len(ixx) if ixx is not None else (len(rix) if rix is not None else r.shape[0])
    s = complex(s)
    p = "F" if r.dtype == np.float32 else "D"
    getattr(lib, f"orc_{p}_vecop")(op, eqop, int(conj), int(dim), ncomp, _ptr(_need(r)), _ptr(rix), _ptr(px),
                                   _ptr(py), s.real, s.imag, _ptr(ixx), n)
    return r


FILL_ZERO, FILL_REC, FILL_DIAG = 0, 1, 2


def fill(kind, eqop, r, a=None, index=None, n=None):
    """Fill on host arrays (orc_*_fill): FILL_ZERO r = 0; FILL_REC r eqop a
    (a: one record, broadcast); FILL_DIAG r eqop a times the unit matrix (a:
    one complex as a (2,) array; r colour matrices).  index: the x-form."""
    lib = load()
    ncomp = r[0].size // 2
    ix = _index(index)
    if n is None:
        n = len(ix) if ix is not None else r.shape[0]
    p = "F" if r.dtype == np.float32 else "D"
    av = None if a is None else np.ascontiguousarray(a, dtype=r.dtype)
    getattr(lib, f"orc_{p}_fill")(kind, eqop, ncomp, _ptr(_need(r)), _ptr(av), _ptr(ix), n)
    return r


def gamma(right, mu, r, a, index=None, a_idx=None, ir=None, ia=None, n=None):
    """r = Gamma_mu a (right False; D or P records) or a Gamma_mu (right True;
    P) on host arrays (orc_*_gamma), in the addressing of site()."""
    lib = load()
    ncomp = r[0].size // 2
    pa = _records(a, ia if ia is not None else a_idx)
    ix = _index(index)
    rix = _index(ir)
    if n is None:
        n = len(ix) if ix is not None else (len(rix) if rix is not None else r.shape[0])
    p = "F" if r.dtype == np.float32 else "D"
    getattr(lib, f"orc_{p}_gamma")(int(right), ncomp, mu, _ptr(_need(r)), _ptr(rix), _ptr(pa), _ptr(ix), n)
    return r


L_NORM2, L_DOT, L_REDOT, L_RETRACE, L_IMTRACE, L_TRACE, L_ANTIHERM = range(7)


def local(kind, eqop, r, a, b=None, index=None, a_idx=None, b_idx=None, ir=None, ia=None, ib=None, n=None):
    """Site-local reduction or part on host arrays (orc_*_local): r is (n,)
    reals (L_NORM2, L_REDOT, L_RETRACE, L_IMTRACE), (n, 2) complexes (L_DOT,
    L_TRACE) or (n, 3, 3, 2) matrices (L_ANTIHERM); addressing as site()."""
    lib = load()
    ncomp = a[0].size // 2
    rreals = r[0].size if r.ndim > 1 else 1
    pa = _records(a, ia if ia is not None else a_idx)
    pb = None if b is None else _records(b, ib if ib is not None else b_idx)
    ix = _index(index)
    rix = _index(ir)
    if n is None:
        n = len(ix) if ix is not None else (len(rix) if rix is not None else r.shape[0])
    p = "F" if r.dtype == np.float32 else "D"
    getattr(lib, f"orc_{p}_local")(kind, eqop, ncomp, rreals, _ptr(_need(r)), _ptr(rix), _ptr(pa), _ptr(pb),
                                   _ptr(ix), n)
    return r


def field_sum(a, index=None, a_idx=None, n=None) -> np.ndarray:
    """Global sum of the records of a (orc_*_sum): the record's reals summed
    in double over the steps (x-form index, pointer-array targets a_idx);
    returns a double array of one record's shape."""
    lib = load()
    ncr = a[0].size
    pa = _records(a, a_idx)
    ix = _index(index)
    if n is None:
        n = len(ix) if ix is not None else a.shape[0]
    out = np.zeros(a.shape[1:], dtype=np.float64)
    p = "F" if a.dtype == np.float32 else "D"
    getattr(lib, f"orc_{p}_sum")(ncr, _ptr(out), _ptr(pa), _ptr(ix), n)
    return out


def fieldmul(eqop, r, f, b, fa=False, cb=False, dim=0, index=None, f_idx=None, b_idx=None, ir=None, i_f=None,
             ib=None, n=None):
    """r eqop f * b with f a real field ((n,) reals) or complex field ((n, 2));
    fa conjugates f, cb with dim the adjoint of b (orc_*_fieldmul)."""
    lib = load()
    ncomp = r[0].size // 2
    fcomplex = 1 if f.ndim == 2 else 0
    fv = f if f.ndim == 2 else f.reshape(-1, 1)
    pf = _records(np.ascontiguousarray(fv), i_f if i_f is not None else f_idx)
    pb = _records(b, ib if ib is not None else b_idx)
    ix = _index(index)
    rix = _index(ir)
    if n is None:
        n = len(ix) if ix is not None else (len(rix) if rix is not None else r.shape[0])
    p = "F" if r.dtype == np.float32 else "D"
    getattr(lib, f"orc_{p}_fieldmul")(fcomplex, int(fa), int(cb), int(dim), eqop, ncomp, _ptr(_need(r)), _ptr(rix),
                                      _ptr(pf), _ptr(pb), _ptr(ix), n)
    return r


def outer(eqop, r, a, b, index=None, a_idx=None, b_idx=None, ir=None, ia=None, ib=None, n=None):
    """M eqop a * b^dagger for colour vectors a, b (orc_*_outer)."""
    lib = load()
    pa = _records(a, ia if ia is not None else a_idx)
    pb = _records(b, ib if ib is not None else b_idx)
    ix = _index(index)
    rix = _index(ir)
    if n is None:
        n = len(ix) if ix is not None else (len(rix) if rix is not None else r.shape[0])
    p = "F" if r.dtype == np.float32 else "D"
    getattr(lib, f"orc_{p}_outer")(eqop, _ptr(_need(r)), _ptr(rix), _ptr(pa), _ptr(pb), _ptr(ix), n)
    return r


PM_MP, PM_PM, PM_PP = 0, 1, 2


def pmul(kind, eqop, r, a, b, ca=False, cb=False, index=None, a_idx=None, b_idx=None, ir=None, ia=None, ib=None,
         n=None):
    """P eqop A * B (orc_*_pmul): PM_MP A colour matrices, B propagators;
    PM_PM A propagators, B colour matrices; PM_PP both propagators; ca / cb
    take the adjoint of A / B."""
    lib = load()
    pa = _records(a, ia if ia is not None else a_idx)
    pb = _records(b, ib if ib is not None else b_idx)
    ix = _index(index)
    rix = _index(ir)
    if n is None:
        n = len(ix) if ix is not None else (len(rix) if rix is not None else r.shape[0])
    p = "F" if r.dtype == np.float32 else "D"
    getattr(lib, f"orc_{p}_pmul")(kind, eqop, int(ca), int(cb), _ptr(_need(r)), _ptr(rix), _ptr(pa), _ptr(pb),
                                  _ptr(ix), n)
    return r


GET_ELEM, SET_ELEM, GET_COLORVEC, SET_COLORVEC, GET_DIRACVEC, SET_DIRACVEC, SPINTRACE = range(7)
COMP_TYPES = {"V": 0, "H": 1, "D": 2, "M": 3, "P": 4}


def comp(kind, t, extra, r, a, index=None, a_idx=None, ir=None, ia=None, n=None):
    """component access and the spin trace (orc_*_comp): t the array type
    (V, H, D, M, P), extra the reference's extra arguments in order."""
    lib = load()
    pa = _records(a, ia if ia is not None else a_idx)
    ix = _index(index)
    rix = _index(ir)
    if n is None:
        n = len(ix) if ix is not None else (len(rix) if rix is not None else a.shape[0])
    e = np.zeros(4, dtype=np.int32)
    e[:len(extra)] = extra
    p = "F" if r.dtype == np.float32 else "D"
    getattr(lib, f"orc_{p}_comp")(kind, COMP_TYPES[t], _ptr(e), _ptr(_need(r)), _ptr(rix), _ptr(pa), _ptr(ix), n)
    return r


def sconv(kind, r, a, b=None, index=None, a_idx=None, b_idx=None, ir=None, ia=None, ib=None, n=None):
    """real / complex conversions and scalar-field division (orc_*_sconv):
    kind 0 C = R, 1 C = R + iR, 2 R = re C, 3 R = im C, 4 R = R/R, 5 C = C/C."""
    lib = load()
    pa = _records(a, ia if ia is not None else a_idx)
    pb = None if b is None else _records(b, ib if ib is not None else b_idx)
    ix = _index(index)
    rix = _index(ir)
    if n is None:
        n = len(ix) if ix is not None else (len(rix) if rix is not None else r.shape[0])
    fl = r.dtype if r.dtype.kind == "f" else a.dtype      # the floating operand fixes the precision
    p = "F" if fl == np.float32 else "D"
    getattr(lib, f"orc_{p}_sconv")(kind, _ptr(_need(r)), _ptr(rix), _ptr(pa), _ptr(pb), _ptr(ix), n)
    return r


def sfn(fn, r, a, b=None, index=None, a_idx=None, b_idx=None, ir=None, ia=None, ib=None, n=None):
    """elementary functions of real and complex fields (orc_*_sfn): fn as
    there; b a real field (20..24) or an int32 field (25, ldexp)."""
    lib = load()
    pa = _records(a, ia if ia is not None else a_idx)
    pb = None if b is None else _records(b, ib if ib is not None else b_idx)
    ix = _index(index)
    rix = _index(ir)
    if n is None:
        n = len(ix) if ix is not None else (len(rix) if rix is not None else r.shape[0])
    p = "F" if r.dtype == np.float32 else "D"
    getattr(lib, f"orc_{p}_sfn")(fn, _ptr(_need(r)), _ptr(rix), _ptr(pa), _ptr(pb), _ptr(ix), n)
    return r


def mask(op, r, a, b, index=None, a_idx=None, b_idx=None, ir=None, ia=None, ib=None, n=None, real=None):
    """copymask (op < 0; r, a records of any floating type, b an int32 mask
    field) or the comparison op 0..5 (== != > < >= <=) of real fields a, b
    into the int32 field r (orc_*_mask).  real: the field's float dtype
    (defaults to a's)."""
    lib = load()
    pa = _records(a, ia if ia is not None else a_idx)
    pb = _records(b, ib if ib is not None else b_idx)
    ix = _index(index)
    rix = _index(ir)
    if n is None:
        n = len(ix) if ix is not None else (len(rix) if rix is not None else a.shape[0] if ia is None else len(ia))
    dt = real or a.dtype
    p = "F" if dt == np.float32 else "D"
    nw = 0 if op >= 0 else int(np.prod(a.shape[1:], dtype=np.int64)) if a.ndim > 1 else 1
    getattr(lib, f"orc_{p}_mask")(op, nw, _ptr(_need(r)), _ptr(rix), _ptr(pa), _ptr(pb), _ptr(ix), n)
    return r


def int_op(op, eqop, r, a=None, b=None, sc=0, index=None, ir=None, ia=None, ib=None, n=None):
    """integer fields (orc_int): op and eqop as there; int32 arrays."""
    lib = load()
    pa = None if a is None else _records(a, ia)
    pb = None if b is None else _records(b, ib)
    ix = _index(index)
    rix = _index(ir)
    if n is None:
        n = len(ix) if ix is not None else (len(rix) if rix is not None else (a.shape[0] if a is not None
                                                                              else r.shape[0]))
    lib.orc_int(op, eqop, _ptr(_need(r)), _ptr(rix), _ptr(pa), _ptr(pb), int(sc), _ptr(ix), n)
    return r
