"""CPU checks of oracle/closed_forms.py, the model of the Cayley-Hamilton
matrix functions with an FP64 coefficient step (QLA_MI355X_F3_M_*_ch, an
addition with no reference counterpart).

The model's formulas are pinned by independent mathematics: scipy's expm and
sqrtm on the same matrices, unitarity of exp(A) and det exp(A) = e^{tr A},
sqrt(A)^2 = A and invsqrt(A) A invsqrt(A) = 1, and the branch cases where the
formulas divide by something small or zero: A = 0, multiples of 1, double
and near-double eigenvalues with both signs of det Q, c0 at +-c0max, rank-one
and singular positive matrices.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import closed_forms as cf
from qla_amd import lattice


def c(a):
    return a[..., 0] + 1j * a[..., 1]


def pack(z):
    return np.stack([z.real, z.imag], axis=-1)


def rel(got, ref):
    return np.linalg.norm(got - ref, axis=(1, 2)) / np.linalg.norm(ref, axis=(1, 2))


def unitary_frame(seed):
    rng = np.random.default_rng(seed)
    return np.linalg.qr(rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3)))[0]


def rotate(diags, seed):
    u = unitary_frame(seed)
    return np.array([u @ np.diag(d) @ u.conj().T for d in diags])


@pytest.mark.parametrize("trace", [False, True])
def test_expA_against_expm(trace):
    a = lattice.random_antihermitian(600, 3, 41)
    if trace:
        t = np.random.default_rng(5).uniform(-3, 3, 600)
        for i in range(3):
            a[:, i, i, 1] += t
    z = c(a)
    ref = np.stack([sl.expm(m) for m in z])
    got = c(cf.expA(a))
    assert rel(got, ref).max() < 1e-13
    # exp of an anti-Hermitian matrix is unitary with det = e^{tr A}
    eye = np.eye(3)
    assert np.abs(got @ np.conj(np.swapaxes(got, 1, 2)) - eye).max() < 1e-13
    assert np.abs(np.linalg.det(got) - np.exp(np.trace(z, axis1=1, axis2=2))).max() < 1e-13


def test_expA_branches():
    """Double and triple eigenvalues (w -> 0, c1 -> 0), both signs of
    c0 = det Q, c0 = +-c0max, a 1e-10 splitting, the zero matrix."""
    diags = [[1.0, 1.0, -2.0], [-1.0, -1.0, 2.0], [0.0, 0.0, 0.0], [1e-9, -1e-9, 0.0],
             [0.3, 0.3 + 1e-10, -0.6 - 1e-10], [2.0, -1.0, -1.0], [0.7, 0.7, 0.7], [3.9, -0.1, -3.8],
             [-3.9, 0.1, 3.8]]
    h = rotate(diags, 7)
    a = pack(1j * h)
    ref = np.stack([sl.expm(1j * m) for m in h])
    got = c(cf.expA(a))
    assert np.abs(got - ref).max() < 1e-14
    assert np.array_equal(c(cf.expA(np.zeros((1, 3, 3, 2))))[0], np.eye(3))


def test_exp_coeffs_symmetry():
    """f_j(-c0) = (-1)^j conj f_j(c0): exp(-iQ) = conj(exp(iQ^T)) in coefficient form."""
    rng = np.random.default_rng(3)
    c1 = rng.uniform(0.1, 10, 200)
    c0 = rng.uniform(-1, 1, 200) * 2 * (c1 / 3) ** 1.5
    p = cf.exp_coeffs(c0, c1)
    m = cf.exp_coeffs(-c0, c1)
    for j in range(3):
        assert np.allclose(m[j], (-1) ** j * np.conj(p[j]), rtol=0, atol=1e-14)


@pytest.mark.parametrize("inverse", [False, True])
def test_sqrtPH_against_sqrtm(inverse):
    a = lattice.random_hpd(600, 3, 42)
    z = c(a)
    ref = np.stack([sl.sqrtm(m) for m in z])
    if inverse:
        ref = np.linalg.inv(ref)
    got = c(cf.sqrtPH(a, inverse=inverse))
    cond = np.linalg.cond(z)
    assert (rel(got, ref) / cond).max() < 1e-15
    if not inverse:
        assert rel(got @ got, z).max() < 1e-12
    else:
        assert np.abs(got @ z @ got - np.eye(3)).max() < 1e-10


def test_sqrtPH_branches():
    """Multiples of 1 (p = 0), double eigenvalues, a 1e-4 eigenvalue, a
    rank-one matrix (d = 0: sqrt A = A / u), the zero matrix, and a singular
    matrix for the inverse (+inf on the diagonal)."""
    diags = [[4.0, 4.0, 4.0], [1.0, 1.0, 1e-4], [2.0, 3.0, 3.0], [5.0, 0.0, 0.0], [0.0, 0.0, 0.0]]
    h = rotate(diags, 11)
    a = pack(h)
    got = c(cf.sqrtPH(a))
    for i, d in enumerate(diags):
        u = unitary_frame(11)
        ref = u @ np.diag(np.sqrt(d)) @ u.conj().T
        # an exactly zero eigenvalue of the rotated matrix comes out as
        # +-delta ~ 1e-16 and contributes sqrt(delta) ~ 1e-8 at most
        bar = 1e-7 if 0.0 in d else 1e-12
        assert np.abs(got[i] - ref).max() < bar, d
    inv = c(cf.sqrtPH(a[:3], inverse=True))
    for i, d in enumerate(diags[:3]):
        u = unitary_frame(11)
        ref = u @ np.diag(1 / np.sqrt(d)) @ u.conj().T
        assert np.abs(inv[i] - ref).max() < 1e-10, d
    sing = c(cf.sqrtPH(a[3:], inverse=True))
    for m in sing:
        assert np.all(np.isinf(np.diag(m).real)) and np.all(np.diag(m).imag == 0)
        assert np.all(m[~np.eye(3, dtype=bool)] == 0)


def test_hermitian_eigenvalues():
    rng = np.random.default_rng(9)
    g = rng.standard_normal((400, 3, 3)) + 1j * rng.standard_normal((400, 3, 3))
    h = g + np.conj(np.swapaxes(g, 1, 2))
    lam = np.sort(cf.hermitian_eigenvalues(h), axis=1)
    ref = np.linalg.eigvalsh(h)
    assert np.abs(lam - ref).max() < 1e-13
