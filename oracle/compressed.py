"""Numpy model of the compressed-precision forms (no reference counterpart).

TEST INFRASTRUCTURE ONLY (the checker for the compressed-form kernels; also
loaded by the benchmark to build its inputs on the host).

Record formats (defined by this project, little-endian):
  link, 24 B   rows 0 and 1 of an SU(3) matrix, 12 reals [row][col][re,im]
               as bfloat16 ("bf16") or IEEE binary16 ("f16"); row 2 is
               rebuilt in float32 as conj(row0 x row1).
  Dh,   52 B   DiracFermion: 24 binary16 v[3][4][2] then a float32 scale s;
               value = float(v) * s.  Encoding: s = max |x| over the 24 reals,
               v = f16(x * (1/s)) with round-to-nearest-even (s = 0: v = 0).
  Hh,   28 B   HalfFermion: 12 binary16 v[3][2][2] then a float32 scale.
All arithmetic after decoding is float32 in the same order as the plain
kernels (spin projection, then x = 0; x += a*t over colour), so this model is
bit-exact with respect to the GPU kernels; accuracy against the FP64 oracle is
a separate, tolerance-based check.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


# ---------------------------------------------------------------- conversions
def f32_to_bf16_bits(x: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 bit patterns, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(x, dtype=F32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.astype(np.uint16)


def bf16_bits_to_f32(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint32) << 16).view(F32)


def pack_link(u: np.ndarray, fmt: str) -> np.ndarray:
    """(n,3,3,2) float32 SU(3) -> (n,12) uint16 bit patterns of rows 0, 1."""
    rows = np.ascontiguousarray(u[:, 0:2], dtype=F32).reshape(u.shape[0], 12)
    if fmt == "bf16":
        return f32_to_bf16_bits(rows)
    if fmt == "f16":
        return rows.astype(np.float16).view(np.uint16)
    raise ValueError(fmt)


def unpack_link(p: np.ndarray, fmt: str) -> np.ndarray:
    """(n,12) uint16 -> (n,3,3,2) float32 with row 2 = conj(row0 x row1)."""
    if fmt == "bf16":
        rows = bf16_bits_to_f32(p)
    else:
        rows = p.view(np.float16).astype(F32)
    r = rows.reshape(-1, 2, 3, 2)
    a_re, a_im = r[:, 0, :, 0], r[:, 0, :, 1]
    b_re, b_im = r[:, 1, :, 0], r[:, 1, :, 1]
    out = np.empty((p.shape[0], 3, 3, 2), dtype=F32)
    out[:, 0:2] = r
    for j in range(3):
        j1, j2 = (j + 1) % 3, (j + 2) % 3
        # p = a[j1]*b[j2], q = a[j2]*b[j1], c = conj(p - q)
        p_re = a_re[:, j1] * b_re[:, j2] - a_im[:, j1] * b_im[:, j2]
        p_im = a_re[:, j1] * b_im[:, j2] + a_im[:, j1] * b_re[:, j2]
        q_re = a_re[:, j2] * b_re[:, j1] - a_im[:, j2] * b_im[:, j1]
        q_im = a_re[:, j2] * b_im[:, j1] + a_im[:, j2] * b_re[:, j1]
        out[:, 2, j, 0] = p_re - q_re
        out[:, 2, j, 1] = -(p_im - q_im)
    return out


def pack_fermion(x: np.ndarray):
    """(n, ...) float32 records -> (halves (n,k) float16, scale (n,) float32)."""
    n = x.shape[0]
    flat = np.ascontiguousarray(x, dtype=F32).reshape(n, -1)
    s = np.abs(flat).max(axis=1).astype(F32)
    with np.errstate(divide="ignore"):
        inv = np.where(s > 0, F32(1) / np.where(s > 0, s, F32(1)), F32(0)).astype(F32)
    v = (flat * inv[:, None]).astype(F32).astype(np.float16)
    return v, s


def unpack_fermion(v: np.ndarray, s: np.ndarray, shape) -> np.ndarray:
    return (v.astype(F32) * s[:, None]).astype(F32).reshape(shape)


def to_records(v: np.ndarray, s: np.ndarray) -> np.ndarray:
    """Interleave halves and scale into the packed byte records."""
    n, k = v.shape
    rec = np.empty((n, 2 * k + 4), dtype=np.uint8)
    rec[:, : 2 * k] = np.ascontiguousarray(v, dtype=np.float16).view(np.uint8).reshape(n, 2 * k)
    rec[:, 2 * k:] = np.ascontiguousarray(s, dtype=F32).view(np.uint8).reshape(n, 4)
    return rec


def from_records(rec: np.ndarray, k: int):
    rec = np.ascontiguousarray(rec)
    v = rec[:, : 2 * k].copy().view(np.float16).reshape(-1, k)
    s = rec[:, 2 * k: 2 * k + 4].copy().view(F32).reshape(-1)
    return v, s


# ---------------------------------------------------------------- operations
def _spproj_rows(d, mu, sign):
    """d (n,3,4,2) float32 -> t (n,3,2,2), expressions_gamma.pl:244-276."""
    re, im = d[..., 0], d[..., 1]

    def plus(a, b):
        return re[..., a] + re[..., b], im[..., a] + im[..., b]

    def minus(a, b):
        return re[..., a] - re[..., b], im[..., a] - im[..., b]

    def plus_i(a, b):   # a + i b
        return re[..., a] - im[..., b], im[..., a] + re[..., b]

    def minus_i(a, b):  # a - i b
        return re[..., a] + im[..., b], im[..., a] - re[..., b]

    def ident(a):
        return re[..., a], im[..., a]

    s = 1 if sign == 1 else -1
    table = {
        (0, 1): (plus_i(0, 3), plus_i(1, 2)), (0, -1): (minus_i(0, 3), minus_i(1, 2)),
        (1, 1): (minus(0, 3), plus(1, 2)), (1, -1): (plus(0, 3), minus(1, 2)),
        (2, 1): (plus_i(0, 2), minus_i(1, 3)), (2, -1): (minus_i(0, 2), plus_i(1, 3)),
        (3, 1): (plus(0, 2), plus(1, 3)), (3, -1): (minus(0, 2), minus(1, 3)),
        (4, 1): (ident(0), ident(1)), (4, -1): (ident(2), ident(3)),
    }
    h0, h1 = table[(mu, s)]
    t = np.empty(d.shape[:2] + (2, 2), dtype=F32)
    t[..., 0, 0], t[..., 0, 1] = h0
    t[..., 1, 0], t[..., 1, 1] = h1
    return t


def _mat_times_cols(a, t):
    """r[ic][is] = sum_kc a[ic][kc] t[kc][is], x = 0 start, kc ascending (float32)."""
    n, ns = t.shape[0], t.shape[2]
    r = np.zeros((n, 3, ns, 2), dtype=F32)
    for ic in range(3):
        for is_ in range(ns):
            xr = np.zeros(n, dtype=F32)
            xi = np.zeros(n, dtype=F32)
            for kc in range(3):
                ar, ai = a[:, ic, kc, 0], a[:, ic, kc, 1]
                br, bi = t[:, kc, is_, 0], t[:, kc, is_, 1]
                xr = xr + (ar * br - ai * bi)
                xi = xi + (ar * bi + ai * br)
            r[:, ic, is_, 0], r[:, ic, is_, 1] = xr, xi
    return r


RECON = {  # (mu, sign) -> ops, srcs; ops 0 zero, 1 +h, 2 -h, 3 +ih, 4 -ih
    (0, 1): ((1, 1, 4, 4), (0, 1, 1, 0)), (1, 1): ((1, 1, 1, 2), (0, 1, 1, 0)),
    (2, 1): ((1, 1, 4, 3), (0, 1, 0, 1)), (3, 1): ((1, 1, 1, 1), (0, 1, 0, 1)),
    (4, 1): ((1, 1, 0, 0), (0, 1, 0, 1)),
    (0, -1): ((1, 1, 3, 3), (0, 1, 1, 0)), (1, -1): ((1, 1, 2, 1), (0, 1, 1, 0)),
    (2, -1): ((1, 1, 3, 4), (0, 1, 0, 1)), (3, -1): ((1, 1, 2, 2), (0, 1, 0, 1)),
    (4, -1): ((0, 0, 1, 1), (0, 1, 0, 1)),
}


def spproj(link: np.ndarray, fmt: str, d_rec: np.ndarray, d_idx, mu: int, sign: int) -> np.ndarray:
    """Compressed H = spproj(M * pD); returns packed Hh records (n, 28) uint8."""
    u = unpack_link(link, fmt)
    v, s = from_records(d_rec, 24)
    d = unpack_fermion(v[d_idx], s[d_idx], (len(d_idx), 3, 4, 2))
    h = _mat_times_cols(u, _spproj_rows(d, mu, sign))
    return to_records(*pack_fermion(h))


def sprecon(link: np.ndarray, fmt: str, h_rec: np.ndarray, h_idx, d_rec: np.ndarray, mu: int, sign: int):
    """Compressed D += sprecon(M * pH); returns new packed Dh records."""
    u = unpack_link(link, fmt)
    hv, hs = from_records(h_rec, 12)
    h = unpack_fermion(hv[h_idx], hs[h_idx], (len(h_idx), 3, 2, 2))
    dv, ds = from_records(d_rec, 24)
    d = unpack_fermion(dv, ds, (dv.shape[0], 3, 4, 2))
    t = _mat_times_cols(u, h)
    ops, srcs = RECON[(mu, 1 if sign == 1 else -1)]
    for sp in range(4):
        op, src = ops[sp], srcs[sp]
        if op == 0:
            continue
        tr, ti = t[:, :, src, 0], t[:, :, src, 1]
        if op in (3, 4):
            tr, ti = -ti, tr
        if op in (2, 4):
            d[:, :, sp, 0] = d[:, :, sp, 0] - tr
            d[:, :, sp, 1] = d[:, :, sp, 1] - ti
        else:
            d[:, :, sp, 0] = d[:, :, sp, 0] + tr
            d[:, :, sp, 1] = d[:, :, sp, 1] + ti
    return to_records(*pack_fermion(d))
