"""Lattice geometry and seeded synthetic fields in QLA's site-major AoS layout.

Geometry: an L0 x L1 x L2 x L3 periodic lattice with lexicographic site index
``x + L0*(y + L1*(z + L2*t))``.  ``neighbours(dims)`` gives ``fw[mu][i]`` (site
of x + mu-hat) and ``bw[mu][i]`` (site of x - mu-hat); callers build QLA pointer
arrays as ``base_address + record_bytes * fw[mu]``.  The even (odd) checkerboard
is the set of sites with x+y+z+t even (odd), in increasing site order.

Random numbers: numpy's PCG64 generator seeded with the integer seed, and its
standard normal transform (``Generator.standard_normal``).  A complex
Gaussian(0,1) number has independent N(0,1) real and imaginary parts.
Random SU(3) matrices: 3x3 complex Gaussian entries, rows 0 and 1
Gram-Schmidt orthonormalised, row 2 = conj(row0 x row1).

Array shapes (numpy, last axis = (re, im)):
  ColorVector (n,3,2)  HalfFermion (n,3,2,2)  DiracFermion (n,3,4,2)
  ColorMatrix (n,3,3,2)
"""
from __future__ import annotations

import numpy as np

REC_REALS = {"V": 6, "H": 12, "D": 24, "M": 18}


def volume(dims) -> int:
    v = 1
    for d in dims:
        v *= int(d)
    return v


def coords(dims) -> np.ndarray:
    """(n, 4) integer coordinates of every site in lexicographic order."""
    n = volume(dims)
    i = np.arange(n, dtype=np.int64)
    out = np.empty((n, 4), dtype=np.int64)
    for mu, L in enumerate(dims):
        out[:, mu] = i % L
        i //= L
    return out


def site_index(c, dims) -> np.ndarray:
    idx = np.zeros(c.shape[0], dtype=np.int64)
    for mu in reversed(range(4)):
        idx = idx * dims[mu] + c[:, mu]
    return idx


def neighbours(dims):
    """fw, bw: (4, n) int64 arrays of forward / backward neighbour sites."""
    c = coords(dims)
    fw = np.empty((4, c.shape[0]), dtype=np.int64)
    bw = np.empty_like(fw)
    for mu in range(4):
        cf = c.copy()
        cf[:, mu] = (cf[:, mu] + 1) % dims[mu]
        fw[mu] = site_index(cf, dims)
        cb = c.copy()
        cb[:, mu] = (cb[:, mu] - 1) % dims[mu]
        bw[mu] = site_index(cb, dims)
    return fw, bw


def parity_sites(dims, parity: int) -> np.ndarray:
    """Sites with (x+y+z+t) % 2 == parity, increasing (int32, QLA index type)."""
    c = coords(dims)
    return np.nonzero(c.sum(axis=1) % 2 == parity)[0].astype(np.int32)


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def gaussian(shape, seed: int, dtype=np.float64, chunk: int = 1 << 24) -> np.ndarray:
    """N(0,1) array of `shape`, drawn in row-major order from PCG64(seed)."""
    rng = _rng(seed)
    out = np.empty(shape, dtype=dtype)
    flat = out.reshape(-1)
    for s in range(0, flat.size, chunk):
        e = min(flat.size, s + chunk)
        flat[s:e] = rng.standard_normal(e - s, dtype=np.float64 if dtype == np.float64 else np.float32)
    return out


def su3_from_gaussian(g: np.ndarray) -> np.ndarray:
    """(n,3,3,2) real Gaussian draws -> (n,3,3,2) SU(3) matrices (float64)."""
    z = g[..., 0].astype(np.float64) + 1j * g[..., 1].astype(np.float64)
    r0 = z[:, 0, :]
    r0 = r0 / np.linalg.norm(r0, axis=1, keepdims=True)
    r1 = z[:, 1, :]
    r1 = r1 - np.sum(np.conj(r0) * r1, axis=1, keepdims=True) * r0
    r1 = r1 / np.linalg.norm(r1, axis=1, keepdims=True)
    r2 = np.conj(np.cross(r0, r1))
    u = np.stack([r0, r1, r2], axis=1)
    return np.stack([u.real, u.imag], axis=-1)


def random_su3(n: int, seed: int, dtype=np.float64, chunk: int = 1 << 20) -> np.ndarray:
    """n random SU(3) matrices (n,3,3,2) from PCG64(seed)."""
    rng = _rng(seed)
    out = np.empty((n, 3, 3, 2), dtype=dtype)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        g = rng.standard_normal((e - s, 3, 3, 2))
        out[s:e] = su3_from_gaussian(g)
    return out


def random_field(kind: str, n: int, seed: int, dtype=np.float64) -> np.ndarray:
    """Gaussian(0,1) field of record kind 'V', 'H', 'D', 'M' or 'P'
    (ColorVector, HalfFermion, DiracFermion, ColorMatrix, DiracPropagator),
    'C' (one complex per site) or 'R' (one real per site)."""
    shape = {"V": (n, 3, 2), "H": (n, 3, 2, 2), "D": (n, 3, 4, 2), "M": (n, 3, 3, 2),
             "P": (n, 3, 4, 3, 4, 2), "C": (n, 2), "R": (n,)}[kind]
    return gaussian(shape, seed, dtype=dtype)


def random_antihermitian(n: int, nc: int, seed: int, radius: float = 4.0) -> np.ndarray:
    """A = i*H with H Gaussian Hermitian, traceless, scaled so its spectral
    radius is radius*u (u uniform in (0, 1]), i.e. eigenvalues of H in
    [-radius, radius] across the batch.  Returns (n, nc, nc, 2) float64."""
    rng = _rng(seed)
    g = rng.standard_normal((n, nc, nc)) + 1j * rng.standard_normal((n, nc, nc))
    h = 0.5 * (g + np.conj(np.swapaxes(g, 1, 2)))
    tr = np.trace(h, axis1=1, axis2=2).real / nc
    h = h - tr[:, None, None] * np.eye(nc)
    rad = np.abs(np.linalg.eigvalsh(h)).max(axis=1)
    u = 1.0 - rng.random(n)
    h = h * (radius * u / np.where(rad > 0, rad, 1))[:, None, None]
    a = 1j * h
    return np.stack([a.real, a.imag], axis=-1)


def random_hpd(n: int, nc: int, seed: int, eps_lo: float = 1e-4, eps_hi: float = 1.0) -> np.ndarray:
    """A = B^dagger B + eps*I, B complex Gaussian, eps log-uniform in
    [eps_lo, eps_hi].  Returns (n, nc, nc, 2) float64."""
    rng = _rng(seed)
    b = rng.standard_normal((n, nc, nc)) + 1j * rng.standard_normal((n, nc, nc))
    eps = np.exp(rng.uniform(np.log(eps_lo), np.log(eps_hi), n))
    a = np.einsum("nki,nkj->nij", np.conj(b), b) + eps[:, None, None] * np.eye(nc)
    return np.stack([a.real, a.imag], axis=-1)
