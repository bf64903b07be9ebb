"""Pins the oracle's other colour-matrix functions (restated from
lib/perl/code-templates/{C_eq_det_M, M_eq_inverse_M, V_eq_M_inverse_V,
M_eq_M_inverse_M, M_eq_expA_M, M_eq_expTA_M, M_eq_sqrtPH_M, M_eq_invsqrtPH_M,
M_eq_log_M, V_eq_eigenvalsH_M}.c) against mathematics:

* numpy's det / inv / solve / eigvalsh and scipy's expm / sqrtm / logm on the
  same matrices (double to ~1e-12, float to float accuracy, scaled by the
  conditioning where it matters);
* exact cases: integer matrices whose determinant the nc = 3 closed form
  evaluates without rounding, permutation matrices (the LU's swap parity),
  singular matrices, diagonal inputs that take the closed forms' degenerate
  branches (a double root) and their r = 0 branches;
* identities: exp(A) unitary with det exp(tr A), sqrtPH(M)^2 = M,
  invsqrtPH(M) M invsqrtPH(M) = 1, exp(log(A)) = A;
* the template's branch structure for nc != 3 (expA, expTA take the exp
  branch, sqrtPH / invsqrtPH the sqrt / invsqrt branch);
* the eigenvalues' QR iteration for nc != 3 against eigvalsh, with the
  rare matrices where an exceptional shift makes the accumulated shift
  complex identified by their signature rather than held to the bar.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from qla_amd import lattice


def c(a):
    return a[..., 0] + 1j * a[..., 1]


def pack(z):
    return np.stack([z.real, z.imag], axis=-1)


def as_dt(a, prec):
    return np.ascontiguousarray(a, dtype=np.float32 if prec == "F" else np.float64)


def gaussian(n, nc, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, nc, nc, 2))


def rel(x, y):
    return np.abs(x - y).max(axis=tuple(range(1, x.ndim))) / np.abs(y).max(axis=tuple(range(1, y.ndim)))


EPS = {"F": float(np.finfo(np.float32).eps), "D": float(np.finfo(np.float64).eps)}


# --------------------------------------------------------------------- det
@pytest.mark.parametrize("nc", [3, 4, 5, 8])
@pytest.mark.parametrize("prec", ["F", "D"])
def test_det_vs_numpy(orc, nc, prec):
    a = as_dt(gaussian(200, nc, 10 + nc), prec)
    A = c(a.astype(np.float64))
    got = c(orc.matfunc("det", a).astype(np.float64))
    ref = np.linalg.det(A)
    # Hadamard bound: |det| <= prod of row norms; rounding error scales with it
    bound = np.prod(np.linalg.norm(A, axis=2), axis=1)
    assert (np.abs(got - ref) <= 40 * nc * EPS[prec] * bound).all()


@pytest.mark.parametrize("prec", ["F", "D"])
def test_det_exact_cases(orc, prec):
    rng = np.random.default_rng(3)
    # nc = 3 closed form on small integers: every product and sum is exact
    ints = rng.integers(-3, 4, size=(500, 3, 3, 2)).astype(np.float64)
    got = c(orc.matfunc("det", as_dt(ints, prec)).astype(np.float64))
    np.testing.assert_array_equal(got, np.round(np.linalg.det(c(ints))))
    # LU: permutation matrices give +-1 exactly (the swap parity)
    for nc in (4, 5, 8):
        perms = np.stack([np.eye(nc)[rng.permutation(nc)] for _ in range(50)])
        got = c(orc.matfunc("det", as_dt(pack(perms + 0j), prec)).astype(np.float64))
        np.testing.assert_array_equal(got, np.round(np.linalg.det(perms)))
        # a zero column is singular: the pivot search finds rmax = 0
        z = gaussian(20, nc, nc)
        z[:, :, 1] = 0
        assert not orc.matfunc("det", as_dt(z, prec)).any()
    z3 = gaussian(20, 3, 1)
    z3[:, :, 2] = 0
    assert not orc.matfunc("det", as_dt(z3, prec)).any()


# ----------------------------------------------------------------- inverses
@pytest.mark.parametrize("nc", [3, 4, 8])
@pytest.mark.parametrize("prec", ["F", "D"])
def test_inverse_and_solves(orc, nc, prec):
    a = as_dt(gaussian(200, nc, 20 + nc), prec)
    A = c(a.astype(np.float64))
    cond = np.linalg.cond(A)
    tol = 20 * nc * EPS[prec] * cond
    inv = c(orc.matfunc("inverse", a).astype(np.float64))
    assert (rel(inv, np.linalg.inv(A)) <= tol).all()
    bv = as_dt(np.random.default_rng(nc).standard_normal((200, nc, 2)), prec)
    x = c(orc.matfunc("inverse_V", a, bv).astype(np.float64))
    ref = np.linalg.solve(A, c(bv.astype(np.float64))[..., None])[..., 0]
    assert (rel(x, ref) <= tol).all()
    bm = as_dt(gaussian(200, nc, 99), prec)
    xm = c(orc.matfunc("inverse_M", a, bm).astype(np.float64))
    assert (rel(xm, np.linalg.solve(A, c(bm.astype(np.float64)))) <= tol).all()


def test_inverse_v_exact_on_permutations(orc):
    """a permutation matrix inverts exactly: x = P^T b, bit for bit."""
    rng = np.random.default_rng(8)
    for nc in (3, 4, 7):
        p = np.stack([np.eye(nc)[rng.permutation(nc)] for _ in range(30)])
        b = rng.standard_normal((30, nc, 2))
        x = orc.matfunc("inverse_V", pack(p + 0j), np.ascontiguousarray(b))
        np.testing.assert_array_equal(c(x), np.einsum("nji,nj->ni", p, c(b)))


# ------------------------------------------------------------- expA, expTA
def with_trace(a, phase):
    z = c(a) + 1j * phase[:, None, None] * np.eye(a.shape[1])
    return pack(z)


@pytest.mark.parametrize("prec", ["F", "D"])
def test_expA_expTA_vs_expm(orc, prec):
    n = 300
    q = lattice.random_antihermitian(n, 3, 31)
    phase = np.random.default_rng(1).uniform(-3, 3, n)
    a = as_dt(with_trace(q, phase), prec)
    tol = 1e-12 if prec == "D" else 4e-6
    A = c(a.astype(np.float64))
    e = c(orc.matfunc("expA", a).astype(np.float64))
    ref = np.stack([sl.expm(m) for m in A])
    assert rel(e, ref).max() < tol
    uu = np.einsum("nij,nkj->nik", e, e.conj())
    assert np.abs(uu - np.eye(3)).max() < 10 * tol
    np.testing.assert_allclose(np.linalg.det(e), np.exp(np.trace(A, axis1=1, axis2=2)), atol=10 * tol)
    t = as_dt(q, prec)
    et = c(orc.matfunc("expTA", t).astype(np.float64))
    assert rel(et, np.stack([sl.expm(m) for m in c(t.astype(np.float64))])).max() < tol


def double_root_formula(q, a):
    """The template's closed form for a double root (M_eq_expA_M.c:199-214):
    f2 = 0, f1 = i e^{ia/2} sin(3a/2)/(3a/2), f0 = e^{ia/2} (cos(3a/2) -
    i sin(3a/2)/3), result f0 + f1 Q.  f0 is the Cayley-Hamilton coefficient;
    f1 is i times it (the template's comment drops the 1/i of
    (e^{2ia} - e^{-ia})/(3ia)), so for such inputs the reference's result is
    not exp(Q).  The oracle and the kernels keep the reference's values."""
    e = np.exp(0.5j * a)
    f1 = 1j * e * np.sin(1.5 * a) / (1.5 * a)
    f0 = e * (np.cos(1.5 * a) - 1j * np.sin(1.5 * a) / 3)
    return f0 * np.eye(3) + f1 * q


@pytest.mark.parametrize("prec", ["F", "D"])
def test_expA_branches(orc, prec):
    tol = 1e-6 if prec == "F" else 1e-14
    # i*diag(2c, -c, -c): a double root, sq^3 == |r| exactly -> the cbrt
    # branch with a = c
    for cval in (0.5, 0.25, 1.5):
        d = np.array([2 * cval, -cval, -cval])
        q = 1j * np.diag(d)
        got = c(orc.matfunc("expTA", as_dt(pack(q[None]), prec)).astype(np.float64))[0]
        np.testing.assert_allclose(got, double_root_formula(q, cval), atol=tol)
        got = c(orc.matfunc("expA", as_dt(pack(q[None] + 0.75j * np.eye(3)), prec)).astype(np.float64))[0]
        np.testing.assert_allclose(got, np.exp(0.75j) * double_root_formula(q, cval), atol=tol)
        # the same spectrum a hair apart takes the regular branch: exp(Q)
        q2 = 1j * np.diag(d + np.array([0, 1e-3, -1e-3]))
        got = c(orc.matfunc("expTA", as_dt(pack(q2[None]), prec)).astype(np.float64))[0]
        np.testing.assert_allclose(got, np.diag(np.exp(np.diag(q2))), atol=1e3 * tol)
    # the templates' r == 0 "zero matrix" branch sets f0 = f1 = f2 = 0, so a
    # matrix whose traceless part is zero maps to the zero matrix (kept as is)
    z = np.zeros((1, 3, 3, 2))
    assert not orc.matfunc("expTA", as_dt(z, prec)).any()
    assert not orc.matfunc("expA", as_dt(pack(0.5j * np.eye(3)[None]), prec)).any()   # s = 0.5 exactly


@pytest.mark.parametrize("fn,general", [("expA", "exp"), ("expTA", "exp"), ("sqrtPH", "sqrt"),
                                        ("invsqrtPH", "invsqrt")])
def test_general_branch_is_the_plain_function(orc, fn, general):
    """nc != 3: M_eq_expA_M.c and M_eq_sqrtPH_M.c fall through to the general
    exp / sqrt code, so the results are the plain function's bit for bit."""
    gen = lattice.random_antihermitian if general == "exp" else lattice.random_hpd
    for nc in (4, 6):
        a = as_dt(gen(40, nc, 50 + nc), "F")
        np.testing.assert_array_equal(orc.matfunc(fn, a), orc.matfunc(general, a))


# ------------------------------------------------------ sqrtPH, invsqrtPH
@pytest.mark.parametrize("prec", ["F", "D"])
def test_sqrtPH_invsqrtPH(orc, prec):
    a = as_dt(lattice.random_hpd(300, 3, 41, eps_lo=1e-2), prec)
    A = c(a.astype(np.float64))
    cond = np.linalg.cond(A)
    tol = (1e-11 if prec == "D" else 2e-4) * cond
    s = c(orc.matfunc("sqrtPH", a).astype(np.float64))
    y = c(orc.matfunc("invsqrtPH", a).astype(np.float64))
    ref = np.stack([sl.sqrtm(m) for m in A])
    assert (rel(s, ref) <= tol).all()
    assert (rel(s @ s, A) <= tol).all()
    assert (np.abs(y @ A @ y - np.eye(3)).max(axis=(1, 2)) <= tol).all()
    assert (rel(y, np.linalg.inv(ref)) <= tol).all()


@pytest.mark.parametrize("prec", ["F", "D"])
def test_sqrtPH_branches(orc, prec):
    tol = 1e-6 if prec == "F" else 1e-14
    # diag(1, 1, 4): double root, exact q and r -> the cbrt branch
    m = as_dt(pack(np.diag([1.0, 1.0, 4.0])[None] + 0j), prec)
    np.testing.assert_allclose(c(orc.matfunc("sqrtPH", m))[0], np.diag([1, 1, 2]), atol=tol)
    np.testing.assert_allclose(c(orc.matfunc("invsqrtPH", m))[0], np.diag([1, 1, 0.5]), atol=tol)
    # 2 * 1: r == 0 -> sqrt(tr/3) * 1
    m = as_dt(pack(2.0 * np.eye(3)[None] + 0j), prec)
    np.testing.assert_allclose(c(orc.matfunc("sqrtPH", m))[0], np.sqrt(2) * np.eye(3), atol=tol)
    np.testing.assert_allclose(c(orc.matfunc("invsqrtPH", m))[0], np.eye(3) / np.sqrt(2), atol=tol)


# --------------------------------------------------------------- eigenvalsH
@pytest.mark.parametrize("prec", ["F", "D"])
def test_eigenvalsH(orc, prec):
    h = c(lattice.random_antihermitian(400, 3, 61)) / 1j
    h = h + np.random.default_rng(2).uniform(-2, 2, 400)[:, None, None] * np.eye(3)
    a = as_dt(pack(h), prec)
    H = c(a.astype(np.float64))
    ev = c(orc.matfunc("eigenvalsH", a).astype(np.float64))
    assert not ev.imag.any()
    ref = np.linalg.eigvalsh(H)
    scale = np.abs(ref).max(axis=1)
    tol = (1e-12 if prec == "D" else 2e-5) * np.maximum(scale, 1)
    assert (np.abs(np.sort(ev.real, axis=1) - ref).max(axis=1) <= tol).all()
    # the template's order: e0 = s1 - 2 sq cos(t/3), then the +- sqrt(3) pair
    m = as_dt(pack(np.diag([1.0, 1.0, 4.0])[None] + 0j), prec)
    np.testing.assert_allclose(c(orc.matfunc("eigenvalsH", m))[0].real, [4, 1, 1], atol=1e-6)


# ---------------------------------------------------------------------- log
@pytest.mark.parametrize("nc", [3, 4])
@pytest.mark.parametrize("prec", ["F", "D"])
def test_log_hpd_and_unitary(orc, nc, prec):
    a = as_dt(lattice.random_hpd(120, nc, 71 + nc, eps_lo=1e-2), prec)
    A = c(a.astype(np.float64))
    lg = c(orc.matfunc("log", a).astype(np.float64))
    ref = np.stack([sl.logm(m) for m in A])
    cond = np.linalg.cond(A)
    tol = (1e-10 if prec == "D" else 1e-3)
    assert (np.abs(lg - ref).max(axis=(1, 2)) <= tol * cond).all()
    # unitary matrices exp(Q), spectrum of -iQ inside (-pi/2, pi/2): log = Q
    q = lattice.random_antihermitian(120, nc, 81 + nc, radius=1.5)
    u = as_dt(pack(np.stack([sl.expm(m) for m in c(q)])), prec)
    lu = c(orc.matfunc("log", u).astype(np.float64))
    assert np.abs(lu - c(q)).max() < (1e-11 if prec == "D" else 5e-5)
    back = np.stack([sl.expm(m) for m in lu])
    assert np.abs(back - c(u.astype(np.float64))).max() < (1e-11 if prec == "D" else 5e-5)


@pytest.mark.parametrize("prec", ["F", "D"])
def test_log_zero(orc, prec):
    r = c(orc.matfunc("log", as_dt(np.zeros((1, 3, 3, 2)), prec)))[0]
    assert (np.diag(r).real == -np.inf).all() and not r[~np.eye(3, dtype=bool)].any()


def hermitian(n, nc, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, nc, nc)) + 1j * rng.standard_normal((n, nc, nc))
    return np.ascontiguousarray(g + np.conj(np.swapaxes(g, 1, 2)))


@pytest.mark.parametrize("nc", [4, 5, 6, 7, 8])
def test_eigenvalsH_qr(orc, nc):
    """The template's QR iteration for nc != 3 (V_eq_eigenvalsH_M.c:180-484)
    in double against numpy's eigvalsh.  Its 2x2 deflation solves from the
    block's real trace and real determinant, exact while the accumulated
    shift t is real; an exceptional shift (iterations 10 and 20) can make t
    complex, and the roots the 2x2 formula then returns carry exactly Im t and
    an error of order (Im t)^2.  So: every matrix whose roots have imaginary
    parts at rounding level agrees with eigvalsh to 1e-13 of the spectral
    radius, with its trace to the same; the others are rare and show that
    signature (two or more roots with one shared imaginary part); and the
    batch contains such a matrix, so the exceptional-shift branch ran."""
    n = 20000
    h = hermitian(n, nc, 200 + nc)
    v = c(orc.matfunc("eigenvalsH", pack(h)))
    ref = np.linalg.eigvalsh(h)
    scale = np.abs(ref).max(axis=1)
    err = np.abs(np.sort(v.real, axis=1) - ref).max(axis=1) / scale
    imag = np.abs(v.imag).max(axis=1) / scale
    clean = imag <= 1e-12
    assert clean.mean() >= 0.995
    assert err[clean].max() <= 1e-13
    trace = np.abs(v.real.sum(axis=1) - np.trace(h, axis1=1, axis2=2).real) / scale
    assert trace[clean].max() <= 1e-13
    shifted = np.nonzero(~clean)[0]
    assert shifted.size >= 1
    for i in shifted:
        im = v[i].imag[np.abs(v[i].imag) > 1e-12 * scale[i]]
        assert im.size >= 2 and np.unique(im).size < im.size, (i, v[i])


@pytest.mark.parametrize("nc", [4, 8])
def test_eigenvalsH_qr_float(orc, nc):
    """The float restatement (the QLA_FN arithmetic) tracks eigvalsh to float
    accuracy on all but the rare complex-shift matrices."""
    h = hermitian(2000, nc, 400 + nc)
    a = pack(h).astype(np.float32)
    v = c(orc.matfunc("eigenvalsH", a).astype(np.float64))
    ref = np.linalg.eigvalsh(c(a.astype(np.float64)))
    err = np.abs(np.sort(v.real, axis=1) - ref).max(axis=1) / np.abs(ref).max(axis=1)
    assert (err <= 1e-4).mean() >= 0.998


@pytest.mark.parametrize("prec", ["F", "D"])
def test_eigenvalsH_qr_exact_cases(orc, prec):
    """Inputs the iteration solves without rounding: diagonal matrices (no
    pivot, every subdiagonal already zero: the roots in diagonal order), a
    multiple of 1, the zero matrix, nc = 1; a block-diagonal matrix of 2x2
    Hermitian blocks deflates each block through the 2x2 formula; a NaN
    input ends (the iteration is bounded) with NaN roots."""
    d = np.array([[3.0, -1.0, 0.5, 2.0, -4.0, 0.25, 1.0, -2.5]])
    for nc in (4, 6, 8):
        m = as_dt(pack(np.diag(d[0, :nc])[None] + 0j), prec)
        v = orc.matfunc("eigenvalsH", m)
        assert np.array_equal(v[0, :, 0], d[0, :nc].astype(m.dtype)) and not v[..., 1].any()
        for z in (np.zeros((1, nc, nc)), 2.5 * np.eye(nc)[None]):
            v = orc.matfunc("eigenvalsH", as_dt(pack(z + 0j), prec))
            assert np.array_equal(v[0, :, 0], np.diag(z[0]).astype(v.dtype)) and not v[..., 1].any()
    v = orc.matfunc("eigenvalsH", as_dt(pack(np.array([[[2.0 - 1.0j]]])), prec))
    assert v[0, 0, 0] == 2.0 and v[0, 0, 1] == -1.0
    blocks = [np.array([[1.0, 2.0 - 1.0j], [2.0 + 1.0j, -3.0]]), np.array([[4.0, 0.5j], [-0.5j, 4.5]])]
    b = np.zeros((4, 4), dtype=complex)
    b[:2, :2], b[2:, 2:] = blocks
    v = c(orc.matfunc("eigenvalsH", as_dt(pack(b[None]), prec)).astype(np.float64))[0]
    tol = 1e-14 if prec == "D" else 1e-5
    np.testing.assert_allclose(np.sort(v.real), np.linalg.eigvalsh(b), atol=tol * 6)
    assert not v.imag.any()
    h = hermitian(1, 6, 5)
    h[0, 2, 3] = np.nan
    v = orc.matfunc("eigenvalsH", as_dt(pack(h), prec))
    assert np.isnan(v).any()


@pytest.mark.parametrize("prec", ["F", "D"])
def test_eigenvalsH_2x2(orc, prec):
    """nc = 2 (V_eq_eigenvalsH_M.c:80-101): from the real trace and the real
    determinant, det = 0 gives (0, tr); the larger-magnitude root first
    formed, the other as det / l1."""
    h = hermitian(500, 2, 9)
    v = c(orc.matfunc("eigenvalsH", as_dt(pack(h), prec)).astype(np.float64))
    ref = np.linalg.eigvalsh(h)
    tol = (1e-14 if prec == "D" else 1e-6) * np.abs(ref).max(axis=1)
    assert (np.abs(np.sort(v.real, axis=1) - ref).max(axis=1) <= tol).all()
    m = as_dt(pack(np.array([[[2.0, 1.0], [1.0, 0.5]]]) + 0j), prec)     # det = 0
    v = orc.matfunc("eigenvalsH", m)
    assert v[0, 0, 0] == 0 and v[0, 1, 0] == 2.5


# ------------------------------------------------------------ nc = 1 and 2
def small_inputs(fn, n, nc, seed):
    """(a, b) complex inputs for fn at nc = 1, 2 and the function of scipy /
    numpy they must equal."""
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, nc, nc)) + 1j * rng.standard_normal((n, nc, nc))
    hpd = np.einsum("nki,nkj->nij", np.conj(g), g) + 0.1 * np.eye(nc)
    ah = (g - np.conj(np.swapaxes(g, 1, 2))) / 2
    b = None
    if fn == "det":
        return g, None, np.linalg.det
    if fn == "inverse":
        return g, None, np.linalg.inv
    if fn == "inverse_V":
        b = rng.standard_normal((n, nc)) + 1j * rng.standard_normal((n, nc))
        return g, b, None
    if fn == "inverse_M":
        b = rng.standard_normal((n, nc, nc)) + 1j * rng.standard_normal((n, nc, nc))
        return g, b, None
    if fn == "exp":
        return g, None, sl.expm
    if fn in ("sqrt", "sqrtPH"):
        return hpd, None, sl.sqrtm
    if fn in ("invsqrt", "invsqrtPH"):
        return hpd, None, lambda m: np.linalg.inv(sl.sqrtm(m))
    if fn == "log":
        return hpd, None, sl.logm
    if fn == "expA":
        return ah, None, sl.expm
    if fn == "expTA":
        return ah - np.trace(ah, axis1=1, axis2=2)[:, None, None] / nc * np.eye(nc), None, sl.expm
    raise ValueError(fn)


SMALL_FNS = ["det", "inverse", "inverse_V", "inverse_M", "exp", "sqrt", "invsqrt", "log", "expA", "expTA",
             "sqrtPH", "invsqrtPH"]


@pytest.mark.parametrize("fn", SMALL_FNS)
@pytest.mark.parametrize("nc", [1, 2])
@pytest.mark.parametrize("prec", ["F", "D"])
def test_small_nc_against_scipy(orc, fn, nc, prec):
    """The templates' own nc = 1, 2 branches (closed forms through lib/cmath)
    against scipy / numpy on the same matrices: within 1e-12 (double) or
    2e-5 (float) of the result's largest component (of 1 for log, whose
    absolute error is the input's relative one), times the condition number
    for the inverses and solves."""
    n = 400
    z, b, ref = small_inputs(fn, n, nc, 11 * nc + len(fn))
    a = as_dt(pack(z), prec)
    zz = c(a.astype(np.float64))
    if b is not None:
        bb = as_dt(pack(b), prec)
        got = c(orc.matfunc(fn, a, bb).astype(np.float64))
        exp = np.linalg.solve(zz, c(bb.astype(np.float64))[..., None])[..., 0] if fn == "inverse_V" \
            else np.linalg.solve(zz, c(bb.astype(np.float64)))
    else:
        got = c(orc.matfunc(fn, a).astype(np.float64))
        exp = np.stack([ref(m) for m in zz])
    got, exp = got.reshape(n, -1), exp.reshape(n, -1)
    scale = np.abs(exp).max(axis=1)
    if fn == "log":
        scale = np.maximum(scale, 1.0)       # log's absolute error is the input's relative one
    tol = (1e-12 if prec == "D" else 2e-5) * scale
    if fn in ("inverse", "inverse_V", "inverse_M", "invsqrt", "invsqrtPH"):
        tol = tol * np.linalg.cond(zz)
    err = np.abs(got - exp).max(axis=1)
    assert (err <= tol).all(), (fn, nc, (err / scale).max())


@pytest.mark.parametrize("prec", ["F", "D"])
def test_small_nc_branch_cases(orc, prec):
    """Inputs that take the nc = 2 branches' special cases, each against its
    exact value: exp of a Jordan block (t = 0: st = ch = 1), sqrt / invsqrt
    of a singular and of a scalar matrix (det = 0: l0 = 0, l1 = tr; equal
    roots), log of s*1 (d = 0: c1 = log(s)/s), sqrtPH of a rank-one and of the
    zero matrix, expTA of a matrix with det = 0 (t = 0); nc = 1 functions of a
    positive number."""
    tol = 1e-14 if prec == "D" else 2e-6
    f = lambda fn, m: c(orc.matfunc(fn, as_dt(pack(np.asarray(m, dtype=complex)[None]), prec)))[0]   # noqa: E731
    s = 0.7
    np.testing.assert_allclose(f("exp", [[s, 1.0], [0.0, s]]), np.exp(s) * np.array([[1, 1], [0, 1]]), atol=tol * 4)
    np.testing.assert_allclose(f("sqrt", [[4.0, 0.0], [0.0, 0.0]]), [[2, 0], [0, 0]], atol=tol)
    np.testing.assert_allclose(f("sqrt", [[2.25, 0.0], [0.0, 2.25]]), 1.5 * np.eye(2), atol=tol)
    np.testing.assert_allclose(f("invsqrt", [[2.25, 0.0], [0.0, 2.25]]), np.eye(2) / 1.5, atol=tol)
    np.testing.assert_allclose(f("log", [[2.0, 0.0], [0.0, 2.0]]), np.log(2.0) * np.eye(2), atol=tol)
    np.testing.assert_allclose(f("sqrtPH", [[1.0, 2.0], [2.0, 4.0]]), np.array([[1, 2], [2, 4]]) / np.sqrt(5), atol=tol)
    assert not f("sqrtPH", np.zeros((2, 2))).any()
    np.testing.assert_allclose(f("expTA", [[0.0, 1.0], [0.0, 0.0]]), [[1, 1], [0, 1]], atol=tol)
    for fn, v in (("exp", np.exp(2.0)), ("sqrt", np.sqrt(2.0)), ("invsqrt", 1 / np.sqrt(2.0)), ("log", np.log(2.0)),
                  ("sqrtPH", np.sqrt(2.0)), ("invsqrtPH", 1 / np.sqrt(2.0)), ("inverse", 0.5), ("det", 2.0)):
        got = c(orc.matfunc(fn, as_dt(pack(np.full((1, 1, 1), 2.0 + 0j)), prec)).astype(np.float64)).ravel()[0]
        assert abs(got - v) <= tol * 4, (fn, got, v)
    got = c(orc.matfunc("expA", as_dt(pack(np.full((1, 1, 1), 0.3 + 0.5j)), prec)).astype(np.float64)).ravel()[0]
    assert abs(got - np.exp(0.5j)) <= tol * 4
