"""Numpy model of the Cayley-Hamilton matrix functions with an FP64
coefficient step (an addition of this library with no reference counterpart).

TEST INFRASTRUCTURE ONLY: the checker of the QLA_MI355X_F3_M_*_ch kernels, also
used by the benchmark for their error figures.

For a 3x3 complex matrix X the functions below return f(X) = f0 + f1 X + f2 X^2
with coefficients computed in double from X's invariants:

* expA: X = A anti-Hermitian (any trace).  With h = Im tr A / 3 and the
  traceless Hermitian Q = -i (A - i h), exp(A) = e^{ih} exp(iQ), and exp(iQ) =
  f0 + f1 Q + f2 Q^2 with the coefficients of Morningstar and Peardon,
  Phys. Rev. D 69, 054501 (2004), eqs. (23)-(34): c0 = det Q, c1 = tr Q^2 / 2,
  theta = arccos(c0 / c0max), u = sqrt(c1/3) cos(theta/3),
  w = sqrt(c1) sin(theta/3), and for c0 < 0 the symmetry
  f_j(-c0) = (-1)^j conj(f_j(c0)).
* sqrtPH / invsqrtPH: X = A Hermitian positive (semi-)definite.  Eigenvalues
  l_k from the trigonometric solution of the characteristic cubic (Smith,
  CACM 4 (1961) 168), s_k = sqrt(l_k), u = sum s_k, v = sum_{j<k} s_j s_k,
  w = prod s_k, d = (s0 + s1)(s0 + s2)(s1 + s2).  The square root S satisfies
  S^3 - u S^2 + v S - w = 0 and S^2 = A, hence S (A + v) = u A + w, which with
  Cayley-Hamilton for A gives sqrt(A) = (u w + (sum l + v) A - A^2) / d; and
  A^{-1/2} = (A + v - u S) / w gives the coefficients of invsqrt below.
  A rank-one A (d = 0) has sqrt(A) = A / u; a zero eigenvalue gives +inf on
  the diagonal of A^{-1/2} and 0 elsewhere.

The device evaluates the same steps in double and rounds each component of
the result to float once, so its results lie within 2^-24 (normwise) of
these.
"""
from __future__ import annotations

import numpy as np


def _cplx(a: np.ndarray) -> np.ndarray:
    """(n,3,3,2) reals -> (n,3,3) complex128."""
    a = np.asarray(a, dtype=np.float64)
    return a[..., 0] + 1j * a[..., 1]


def _reals(z: np.ndarray, dtype) -> np.ndarray:
    return np.stack([z.real, z.imag], axis=-1).astype(dtype)


def _sinc(w: np.ndarray) -> np.ndarray:
    """sin(w)/w, with its Taylor series below |w| = 0.05."""
    w2 = w * w
    series = 1.0 - w2 / 6.0 * (1.0 - w2 / 20.0 * (1.0 - w2 / 42.0))
    with np.errstate(invalid="ignore", divide="ignore"):
        direct = np.sin(w) / w
    return np.where(np.abs(w) < 0.05, series, direct)


def exp_coeffs(c0: np.ndarray, c1: np.ndarray):
    """f0, f1, f2 (complex128) of exp(iQ) = f0 + f1 Q + f2 Q^2 for a traceless
    Hermitian Q with c0 = det Q, c1 = tr Q^2 / 2 (Morningstar-Peardon)."""
    c0 = np.asarray(c0, dtype=np.float64)
    c1 = np.asarray(c1, dtype=np.float64)
    neg = c0 < 0
    ac0 = np.abs(c0)
    zero = c1 <= 0
    c1s = np.where(zero, 1.0, c1)
    c13 = c1s / 3.0
    c0max = 2.0 * c13 * np.sqrt(c13)
    t = np.minimum(ac0 / c0max, 1.0)
    theta = np.arccos(t)
    u = np.sqrt(c13) * np.cos(theta / 3.0)
    w = np.sqrt(c1s) * np.sin(theta / 3.0)
    xi0 = _sinc(w)
    cw = np.cos(w)
    u2, w2 = u * u, w * w
    e2iu = np.exp(2j * u)
    emiu = np.exp(-1j * u)
    h0 = (u2 - w2) * e2iu + emiu * (8 * u2 * cw + 2j * u * (3 * u2 + w2) * xi0)
    h1 = 2 * u * e2iu - emiu * (2 * u * cw - 1j * (3 * u2 - w2) * xi0)
    h2 = e2iu - emiu * (cw + 3j * u * xi0)
    den = 9 * u2 - w2
    f0, f1, f2 = h0 / den, h1 / den, h2 / den
    f0 = np.where(neg, np.conj(f0), f0)
    f1 = np.where(neg, -np.conj(f1), f1)
    f2 = np.where(neg, np.conj(f2), f2)
    f0 = np.where(zero, 1.0 + 0j, f0)
    f1 = np.where(zero, 0j, f1)
    f2 = np.where(zero, 0j, f2)
    return f0, f1, f2


def hermitian_eigenvalues(a: np.ndarray) -> np.ndarray:
    """(n,3,3) complex Hermitian -> (n,3) real eigenvalues, trigonometric form."""
    m = np.real(np.trace(a, axis1=1, axis2=2)) / 3.0
    k = a - m[:, None, None] * np.eye(3)
    p2 = np.sum(np.abs(k) ** 2, axis=(1, 2)) / 6.0
    q = np.real(np.linalg.det(k)) / 2.0
    p = np.sqrt(p2)
    ok = p > 0
    ps = np.where(ok, p, 1.0)
    r = np.clip(q / (ps * ps * ps), -1.0, 1.0)
    phi = np.arccos(r) / 3.0
    l0 = m + 2 * ps * np.cos(phi)
    l2 = m + 2 * ps * np.cos(phi + 2 * np.pi / 3)
    l1 = 3 * m - l0 - l2
    lam = np.stack([l0, l1, l2], axis=1)
    return np.where(ok[:, None], lam, m[:, None])


def sqrt_coeffs(lam: np.ndarray, inverse: bool):
    """Real f0, f1, f2 of A^{1/2} (or A^{-1/2}) = f0 + f1 A + f2 A^2 from the
    eigenvalues (n,3) of A."""
    lam = np.maximum(lam, 0.0)
    s = np.sqrt(lam)
    u = s.sum(axis=1)
    v = s[:, 0] * s[:, 1] + s[:, 0] * s[:, 2] + s[:, 1] * s[:, 2]
    w = s[:, 0] * s[:, 1] * s[:, 2]
    d = (s[:, 0] + s[:, 1]) * (s[:, 0] + s[:, 2]) * (s[:, 1] + s[:, 2])
    tr = lam.sum(axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        if not inverse:
            # rank <= 1 (d = 0): sqrt(A) = A / u, or 0
            rank1 = d == 0
            ds = np.where(rank1, 1.0, d)
            f0 = np.where(rank1, 0.0, u * w / ds)
            f1 = np.where(rank1, np.where(u > 0, 1.0 / np.where(u > 0, u, 1.0), 0.0), (tr + v) / ds)
            f2 = np.where(rank1, 0.0, -1.0 / ds)
            return f0, f1, f2
        # a zero eigenvalue (w = 0): +inf on the diagonal
        sing = w == 0
        wd = np.where(sing, 1.0, w * d)
        # (v d - u^2 w) written as a sum of non-negative terms
        num0 = w * u * u + sum(lam[:, k] * s[:, k] * (tr - lam[:, k]) for k in range(3))
        f0 = np.where(sing, np.inf, num0 / wd)
        f1 = np.where(sing, 0.0, -(u * tr + w) / wd)
        f2 = np.where(sing, 0.0, u / wd)
        return f0, f1, f2


def _combine(x: np.ndarray, f0, f1, f2) -> np.ndarray:
    """f0 + f1 X + f2 X^2 for (n,3,3) complex X and (n,) coefficients."""
    out = f1[:, None, None] * x + f2[:, None, None] * (x @ x)
    idx = np.arange(3)
    out[:, idx, idx] += f0[:, None]
    return out


def expA(a: np.ndarray) -> np.ndarray:
    """exp of the anti-Hermitian part of (n,3,3,2) records -> (n,3,3,2) float64."""
    z = _cplx(a)
    h = np.imag(np.trace(z, axis1=1, axis2=2)) / 3.0
    q = -1j * (z - 1j * h[:, None, None] * np.eye(3))
    q = 0.5 * (q + np.conj(np.swapaxes(q, 1, 2)))         # exactly Hermitian
    c1 = np.sum(np.abs(q) ** 2, axis=(1, 2)) / 2.0
    c0 = np.real(np.linalg.det(q))
    f0, f1, f2 = exp_coeffs(c0, c1)
    ph = np.exp(1j * h)
    return _reals(_combine(q, ph * f0, ph * f1, ph * f2), np.float64)


def sqrtPH(a: np.ndarray, inverse: bool = False) -> np.ndarray:
    """A^{1/2} (or A^{-1/2}) of the Hermitian part of positive records
    (n,3,3,2) -> float64."""
    z = _cplx(a)
    z = 0.5 * (z + np.conj(np.swapaxes(z, 1, 2)))
    f0, f1, f2 = sqrt_coeffs(hermitian_eigenvalues(z), inverse)
    out = _combine(z, f0.astype(np.complex128), f1.astype(np.complex128), f2.astype(np.complex128))
    return _reals(out, np.float64)
