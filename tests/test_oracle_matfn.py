"""Pins the oracle's matrix functions (restated from
lib/perl/code-templates/M_eq_{exp,sqrt,invsqrt,inverse,M_inverse}_M.c)
against mathematics: scipy's expm / sqrtm on the same matrices, unitarity of
exp of an anti-Hermitian matrix, sqrt(A)^2 = A, invsqrt(A) A invsqrt(A) = 1,
and the templates' special cases (A = 0, the s = 1 branch, the s <= 1024
cap).  The double-precision restatement must agree with scipy to ~1e-12;
the float one to float accuracy."""
import numpy as np
import pytest
import scipy.linalg as sl

from qla_amd import lattice


def c(a):
    return a[..., 0] + 1j * a[..., 1]


def pack(z):
    return np.stack([z.real, z.imag], axis=-1)


def rel(x, y):
    return np.abs(x - y).max(axis=(1, 2)) / np.abs(y).max(axis=(1, 2))


@pytest.mark.parametrize("nc", [3, 4, 8])
@pytest.mark.parametrize("prec", ["F", "D"])
def test_exp(orc, nc, prec):
    n = 64
    a = lattice.random_antihermitian(n, nc, 100 + nc)
    dt = np.float32 if prec == "F" else np.float64
    r = c(orc.matfn("exp", a.astype(dt)).astype(np.float64))
    ref = np.stack([sl.expm(m) for m in c(a.astype(dt).astype(np.float64))])
    tol = 2e-12 if prec == "D" else 3e-5
    assert rel(r, ref).max() < tol
    uu = np.einsum("nij,nkj->nik", r, r.conj())
    assert np.abs(uu - np.eye(nc)).max() < (1e-12 if prec == "D" else 1e-5)


@pytest.mark.parametrize("prec", ["F", "D"])
def test_exp_branches(orc, prec):
    dt = np.float32 if prec == "F" else np.float64
    z = np.zeros((1, 3, 3, 2), dtype=dt)
    np.testing.assert_array_equal(c(orc.matfn("exp", z)), np.eye(3)[None])           # A = 0
    small = lattice.random_antihermitian(8, 3, 5, radius=0.2).astype(dt)               # s = 1
    ref = np.stack([sl.expm(m) for m in c(small.astype(np.float64))])
    assert rel(c(orc.matfn("exp", small).astype(np.float64)), ref).max() < (1e-12 if prec == "D" else 1e-5)
    big = lattice.random_antihermitian(4, 3, 6, radius=700.0)                          # s capped at 1024
    big = big * (700.0 / np.sqrt((big ** 2).sum(axis=(1, 2, 3))))[:, None, None, None]
    r = c(orc.matfn("exp", big.astype(dt)).astype(np.float64))
    uu = np.einsum("nij,nkj->nik", r, r.conj())
    assert np.isfinite(r).all() and np.abs(uu - np.eye(3)).max() < (1e-6 if prec == "D" else 1e-2)


@pytest.mark.parametrize("nc", [3, 4, 8])
@pytest.mark.parametrize("prec", ["F", "D"])
def test_sqrt_invsqrt(orc, nc, prec):
    n = 64
    a = lattice.random_hpd(n, nc, 200 + nc)
    dt = np.float32 if prec == "F" else np.float64
    a = a.astype(dt)
    A = c(a.astype(np.float64))
    s = c(orc.matfn("sqrt", a).astype(np.float64))
    y = c(orc.matfn("invsqrt", a).astype(np.float64))
    ref = np.stack([sl.sqrtm(m) for m in A])
    cond = np.linalg.cond(A)
    tol = (1e-11 if prec == "D" else 1e-4)
    assert (rel(s, ref) <= tol * np.sqrt(cond)).all()
    assert (rel(s @ s, A) <= tol * np.sqrt(cond)).all()
    eye = y @ A @ y
    assert (np.abs(eye - np.eye(nc)).max(axis=(1, 2)) <= tol * cond).all()


@pytest.mark.parametrize("prec", ["F", "D"])
def test_sqrt_zero(orc, prec):
    dt = np.float32 if prec == "F" else np.float64
    z = np.zeros((1, 3, 3, 2), dtype=dt)
    assert not orc.matfn("sqrt", z).any()
    y = c(orc.matfn("invsqrt", z))[0]
    assert np.isinf(np.diag(y).real).all() and not y[~np.eye(3, dtype=bool)].any()


def similar(n, nc, seed, kind):
    """M = P D P^-1 with a Gaussian complex P and a diagonal D drawn as in the
    reference's own tests (protocol_tensor_sng.m4:481-596): Gaussian complex
    eigenvalues, with Re <= 0 for sqrt / invsqrt (an eigenvalue of positive
    real part is negated there) and purely imaginary ones for exp; returns
    (M, P, d) in complex128."""
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((n, nc, nc)) + 1j * rng.standard_normal((n, nc, nc))
    d = rng.standard_normal((n, nc)) + 1j * rng.standard_normal((n, nc))
    if kind == "exp":
        d = 1j * d.real
    else:
        d = np.where(d.real > 0, -d, d)
    m = p @ (d[:, :, None] * np.linalg.inv(p))
    return m, p, d


@pytest.mark.parametrize("fn", ["exp", "sqrt", "invsqrt"])
@pytest.mark.parametrize("nc", [3, 4, 8])
def test_similarity_transform(orc, fn, nc):
    """f(P D P^-1) = P f(D) P^-1 on general (non-normal) matrices, the
    reference's own construction; principal branches (numpy's complex sqrt),
    in double, within 1e-9 times cond(P)^2."""
    m, p, d = similar(40, nc, 300 + nc, fn)
    f = {"exp": np.exp, "sqrt": np.sqrt, "invsqrt": lambda z: 1 / np.sqrt(z)}[fn]
    ref = p @ (f(d)[:, :, None] * np.linalg.inv(p))
    got = c(orc.matfn(fn, pack(m)))
    cond = np.linalg.cond(p)
    err = np.abs(got - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2))
    assert (err <= 1e-9 * cond ** 2).all(), (fn, nc, (err / cond ** 2).max())
