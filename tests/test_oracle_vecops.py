"""Pins the oracle's field linear algebra (orc_*_vecop) against numpy.

Each case is checked against an independent evaluation: numpy's own complex
conjugate, axis transposes of the record shapes (the P transpose as a swap
of the (colour, spin) axis pairs, not the 12 x 12 index formula the oracle
uses), its IEEE single / double arithmetic on whole arrays (one rounding per
operation, as C without contraction), exact complex products on small
integers, long-double bounds on random data, and algebraic identities
(i*(i*x) = -x, (x^dagger)^dagger = x, 1*x = x, 2*x = x + x, x + y - y).
A dropped term, a sign slip, an unconjugated or untransposed source or a
wrong eqop fails these by whole units, not ulps.  The eqm forms keep the
reference macros' zero signs (include/qla_complex.h:149, 167, 268, 334,
371), pinned on cases whose exact result is zero.
"""
import numpy as np
import pytest

from qla_amd import lattice

TYPES = ("V", "H", "D", "M", "P")
L = np.clongdouble


def cx(a):
    return a[..., 0] + 1j * a[..., 1]


def re_im(c, dtype):
    return np.stack([c.real, c.imag], axis=-1).astype(dtype)


def field(t, n, seed, dtype):
    return lattice.random_field(t, n, seed, dtype=dtype)


def transpose(t, c):
    """Complex records (n, ...) transposed as matrices: M (n,3,3) swaps its two
    axes; P (n,3,4,3,4) swaps (ic,is) with (jc,js)."""
    if t == "M":
        return np.swapaxes(c, 1, 2)
    return np.transpose(c, (0, 3, 4, 1, 2))


def eqop_np(eqop, r0, y):
    """eq / peq / meq / eqm on the (re, im) arrays, one IEEE rounding each."""
    return {0: y, 1: r0 + y, 2: r0 - y, 3: -y}[eqop]


DT = {"F": np.float32, "D": np.float64}


# (type, form): transposes exist for the square records M and P only
COPY_FORMS = [(t, f) for t in TYPES for f in ("copy", "conj", "transpose", "adjoint")
              if t in ("M", "P") or f in ("copy", "conj")]
SCALE_FORMS = [(t, adj) for t in TYPES for adj in (False, True) if t in ("M", "P") or not adj]


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("eqop", range(4))
@pytest.mark.parametrize("t,form", COPY_FORMS)
def test_copy_forms_vs_numpy(orc, prec, t, eqop, form):
    dt = DT[prec]
    n = 37
    x = field(t, n, 1, dt)
    r0 = field(t, n, 2, dt)
    r = r0.copy()
    conj = form in ("conj", "adjoint")
    dim = orc.TRANSPOSE_DIM[t] if form in ("transpose", "adjoint") else 0
    orc.vecop(orc.COPY, eqop, r, x, conj=conj, dim=dim)
    c = cx(x)
    if dim:
        c = transpose(t, c)
    if conj:
        c = np.conj(c)
    want = eqop_np(eqop, r0, re_im(c, dt))
    np.testing.assert_array_equal(r, want)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("eqop", range(4))
def test_times_i(orc, prec, eqop):
    dt = DT[prec]
    x = field("D", 29, 3, dt)
    r0 = field("D", 29, 4, dt)
    r = r0.copy()
    orc.vecop(orc.TIMES_I, eqop, r, x)
    ix = np.stack([-x[..., 1], x[..., 0]], axis=-1)
    np.testing.assert_array_equal(r, eqop_np(eqop, r0, ix))
    # i * (i * x) = -x exactly
    t1 = np.zeros_like(x)
    t2 = np.zeros_like(x)
    orc.vecop(orc.TIMES_I, 0, t1, x)
    orc.vecop(orc.TIMES_I, 0, t2, t1)
    np.testing.assert_array_equal(t2, -x)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("eqop", range(4))
@pytest.mark.parametrize("t,adj", SCALE_FORMS)
def test_real_scale_vs_numpy(orc, prec, t, eqop, adj):
    dt = DT[prec]
    n = 31
    x = field(t, n, 5, dt)
    r0 = field(t, n, 6, dt)
    s = dt(0.7310585786300049)
    r = r0.copy()
    dim = orc.TRANSPOSE_DIM[t] if adj else 0
    orc.vecop(orc.RSCALE, eqop, r, x, s=float(s), conj=adj, dim=dim)
    c = cx(x)
    if adj:
        c = np.conj(transpose(t, c))
    src = re_im(c, dt)
    np.testing.assert_array_equal(r, eqop_np(eqop, r0, s * src))


@pytest.mark.parametrize("prec", ["F", "D"])
def test_real_scale_identities(orc, prec):
    dt = DT[prec]
    x = field("P", 9, 7, dt)
    r = np.zeros_like(x)
    orc.vecop(orc.RSCALE, 0, r, x, s=1.0)
    np.testing.assert_array_equal(r, x)
    orc.vecop(orc.RSCALE, 0, r, x, s=2.0)
    s2 = np.zeros_like(x)
    orc.vecop(orc.PLUS, 0, s2, x, x)
    np.testing.assert_array_equal(r, s2)
    orc.vecop(orc.RSCALE, 3, r, x, s=1.0)
    m = np.zeros_like(x)
    orc.vecop(orc.COPY, 3, m, x)
    np.testing.assert_array_equal(r, m)


def _int_field(t, n, seed, dt):
    rng = np.random.default_rng(seed)
    return rng.integers(-9, 10, size=field(t, n, 0, dt).shape).astype(dt)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("eqop", range(4))
@pytest.mark.parametrize("t,adj", SCALE_FORMS)
def test_complex_scale_exact_on_integers(orc, prec, t, eqop, adj):
    """Small integers: every product and sum is exact, so the result must
    equal the complex128 product z*x (conj / transposed for the adjoint)."""
    dt = DT[prec]
    n = 23
    x = _int_field(t, n, 8, dt)
    r0 = _int_field(t, n, 9, dt)
    z = complex(3, -5)
    r = r0.copy()
    dim = orc.TRANSPOSE_DIM[t] if adj else 0
    orc.vecop(orc.CSCALE, eqop, r, x, s=z, conj=adj, dim=dim)
    c = cx(x.astype(np.float64))
    if adj:
        c = np.conj(transpose(t, c))
    y = z * c
    c0 = cx(r0.astype(np.float64))
    want = {0: y, 1: c0 + y, 2: c0 - y, 3: -y}[eqop]
    np.testing.assert_array_equal(cx(r.astype(np.float64)), want)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("eqop", range(4))
@pytest.mark.parametrize("adj", [False, True])
def test_complex_scale_vs_long_double(orc, prec, eqop, adj):
    dt = DT[prec]
    u = np.finfo(dt).eps / 2
    n = 40
    x = field("M", n, 10, dt)
    r0 = field("M", n, 11, dt)
    z = complex(dt(-0.3141592653589793), dt(1.4142135623730951))
    r = r0.copy()
    orc.vecop(orc.CSCALE, eqop, r, x, s=z, conj=adj, dim=3 if adj else 0)
    c = cx(x.astype(np.longdouble)).astype(L)
    if adj:
        c = np.conj(np.swapaxes(c, 1, 2))
    y = L(z) * c
    c0 = cx(r0.astype(np.longdouble)).astype(L)
    want = {0: y, 1: c0 + y, 2: c0 - y, 3: -y}[eqop]
    bound = 4 * u * (abs(z.real) + abs(z.imag)) * (np.abs(c.real) + np.abs(c.imag)) + 2 * u * np.abs(c0)
    got = cx(r.astype(np.longdouble))
    assert np.all(np.abs(got.real - want.real) <= bound + 1e-300)
    assert np.all(np.abs(got.imag - want.imag) <= bound + 1e-300)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("t", TYPES)
def test_sums_vs_numpy(orc, prec, t):
    dt = DT[prec]
    x = field(t, 17, 12, dt)
    y = field(t, 17, 13, dt)
    r = np.zeros_like(x)
    orc.vecop(orc.PLUS, 0, r, x, y)
    np.testing.assert_array_equal(r, x + y)
    orc.vecop(orc.MINUS, 0, r, x, y)
    np.testing.assert_array_equal(r, x - y)
    s = dt(-1.6180339887498949)
    orc.vecop(orc.R_PLUS, 0, r, x, y, s=float(s))
    np.testing.assert_array_equal(r, s * x + y)
    orc.vecop(orc.R_MINUS, 0, r, x, y, s=float(s))
    np.testing.assert_array_equal(r, s * x - y)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("t", TYPES)
@pytest.mark.parametrize("op", ["C_PLUS", "C_MINUS"])
def test_complex_axpy(orc, prec, t, op):
    dt = DT[prec]
    # exact on integers
    xi = _int_field(t, 19, 14, dt)
    yi = _int_field(t, 19, 15, dt)
    z = complex(-7, 2)
    r = np.zeros_like(xi)
    orc.vecop(getattr(orc, op), 0, r, xi, yi, s=z)
    sgn = 1 if op == "C_PLUS" else -1
    want = z * cx(xi.astype(np.float64)) + sgn * cx(yi.astype(np.float64))
    np.testing.assert_array_equal(cx(r.astype(np.float64)), want)
    # random data: within the rounding of (zr x - zi x) +- y
    u = np.finfo(dt).eps / 2
    x = field(t, 19, 16, dt)
    y = field(t, 19, 17, dt)
    z = complex(dt(0.25 + 1 / 3), dt(-2 / 7))
    orc.vecop(getattr(orc, op), 0, r, x, y, s=z)
    c = cx(x.astype(np.longdouble))
    w = L(z) * c + sgn * cx(y.astype(np.longdouble))
    bound = 4 * u * ((abs(z.real) + abs(z.imag)) * (np.abs(c.real) + np.abs(c.imag)) +
                     np.abs(cx(y.astype(np.longdouble))))
    got = cx(r.astype(np.longdouble))
    assert np.all(np.abs(got.real - w.real) <= bound)
    assert np.all(np.abs(got.imag - w.imag) <= bound)


@pytest.mark.parametrize("prec", ["F", "D"])
def test_adjoint_involution_and_sum_inverse(orc, prec):
    dt = DT[prec]
    x = field("P", 11, 18, dt)
    y = field("P", 11, 19, dt)
    a = np.zeros_like(x)
    b = np.zeros_like(x)
    orc.vecop(orc.COPY, 0, a, x, conj=True, dim=12)
    orc.vecop(orc.COPY, 0, b, a, conj=True, dim=12)
    np.testing.assert_array_equal(b, x)
    orc.vecop(orc.COPY, 0, a, x, dim=12)
    orc.vecop(orc.COPY, 0, b, a, dim=12)
    np.testing.assert_array_equal(b, x)
    # integers: (x + y) - y = x exactly
    xi = _int_field("D", 13, 20, dt)
    yi = _int_field("D", 13, 21, dt)
    s = np.zeros_like(xi)
    d = np.zeros_like(xi)
    orc.vecop(orc.PLUS, 0, s, xi, yi)
    orc.vecop(orc.MINUS, 0, d, s, yi)
    np.testing.assert_array_equal(d, xi)


@pytest.mark.parametrize("prec", ["F", "D"])
def test_eqm_zero_signs_follow_the_macros(orc, prec):
    """eqm results whose exact value is zero: the macro forms give the signs
    -(+0) = -0 for copy / i / real scale (the negation of eq) but +0 for the
    complex scale, whose eqm real part is -zr*xr + zi*xi, not -(zr*xr - zi*xi)."""
    dt = DT[prec]
    x = np.zeros((1, 3, 2), dtype=dt)
    x[0, 0] = (1.0, 1.0)          # zr*xr == zi*xi for z = (1, 1)
    r = np.full_like(x, 5.0)
    orc.vecop(orc.CSCALE, 3, r, x, s=complex(1, 1))
    assert r[0, 0, 0] == 0 and not np.signbit(r[0, 0, 0])
    orc.vecop(orc.CSCALE, 0, r, x, s=complex(1, 1))
    assert r[0, 0, 0] == 0 and not np.signbit(r[0, 0, 0])
    r[:] = 5.0
    orc.vecop(orc.COPY, 3, r, np.zeros_like(x))
    assert np.all(np.signbit(r))
    r[:] = 5.0
    orc.vecop(orc.RSCALE, 3, r, np.zeros_like(x), s=2.0)
    assert np.all(np.signbit(r))
    # the conjugated complex scale: eqm imaginary part -zi*xr + zr*xi
    y = np.zeros((1, 3, 2), dtype=dt)
    y[0, 0] = (1.0, -1.0)        # z = x = (1, -1): zi*xr == zr*xi == -1
    r[:] = 5.0
    orc.vecop(orc.CSCALE, 3, r, y, s=complex(1, -1), conj=True)
    assert r[0, 0, 1] == 0 and not np.signbit(r[0, 0, 1])


@pytest.mark.parametrize("prec", ["F", "D"])
def test_index_forms(orc, prec):
    """x-form, pointer operands and the per-operand indices against the
    v-form on gathered records."""
    dt = DT[prec]
    n = 40
    x = field("D", n, 22, dt)
    y = field("D", n, 23, dt)
    rng = np.random.default_rng(24)
    perm = rng.permutation(n)
    sub = np.sort(rng.choice(n, size=n // 2, replace=False))
    s = 0.375
    # gang index: r[idx] = s*x[idx] + y[idx]
    r = field("D", n, 25, dt)
    want = r.copy()
    want[sub] = dt(s) * x[sub] + y[sub]
    orc.vecop(orc.R_PLUS, 0, r, x, y, s=s, index=sub)
    np.testing.assert_array_equal(r, want)
    # pointer sources: r[i] = s * x[perm[i]] - y[perm2[i]]
    perm2 = rng.permutation(n)
    r = np.zeros_like(x)
    orc.vecop(orc.R_MINUS, 0, r, x, y, s=s, x_idx=perm, y_idx=perm2)
    np.testing.assert_array_equal(r, dt(s) * x[perm] - y[perm2])
    # per-operand: r[ir[i]] += x[ix[i]] for i < m
    ir = rng.permutation(n)[:25]
    ix = rng.permutation(n)[:25]
    r = field("D", n, 26, dt)
    want = r.copy()
    want[ir] = want[ir] + x[ix]
    orc.vecop(orc.COPY, 1, r, x, ir=ir, ix=ix)
    np.testing.assert_array_equal(r, want)
    # per-operand on the second source only: r[i] = x[i] + y[iy[i]]
    iy = rng.permutation(n)
    r = np.zeros_like(x)
    orc.vecop(orc.PLUS, 0, r, x, y, iy=iy)
    np.testing.assert_array_equal(r, x + y[iy])


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("t", TYPES)
@pytest.mark.parametrize("eqop", range(4))
def test_record_fill_vs_numpy(orc, prec, t, eqop):
    """r eqop a with one record a for every site (x-form on a subset)."""
    dt = DT[prec]
    n = 21
    r0 = field(t, n, 40, dt)
    a = field(t, 1, 41, dt)[0]
    r = r0.copy()
    orc.fill(orc.FILL_REC, eqop, r, a)
    np.testing.assert_array_equal(r, eqop_np(eqop, r0, np.broadcast_to(a, r0.shape)))
    sub = np.array([3, 7, 11, 20])
    r = r0.copy()
    orc.fill(orc.FILL_REC, eqop, r, a, index=sub)
    want = r0.copy()
    want[sub] = eqop_np(eqop, r0[sub], np.broadcast_to(a, r0[sub].shape))
    np.testing.assert_array_equal(r, want)


@pytest.mark.parametrize("prec", ["F", "D"])
def test_zero_fill(orc, prec):
    dt = DT[prec]
    for t in TYPES:
        r = field(t, 9, 42, dt)
        r[0] = -0.0
        orc.fill(orc.FILL_ZERO, 0, r)
        assert np.all(r == 0) and not np.any(np.signbit(r))
        r = field(t, 9, 43, dt)
        keep = r.copy()
        orc.fill(orc.FILL_ZERO, 0, r, index=np.array([2, 5]))
        keep[[2, 5]] = 0
        np.testing.assert_array_equal(r, keep)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("eqop", range(4))
def test_diagonal_fill(orc, prec, eqop):
    """M eqop c times the unit matrix: values against numpy's c * eye(3)."""
    dt = DT[prec]
    r0 = field("M", 15, 44, dt)
    c = np.array([dt(1.25), dt(-0.375)])
    r = r0.copy()
    orc.fill(orc.FILL_DIAG, eqop, r, c)
    y = re_im(complex(c[0], c[1]) * np.eye(3)[None].repeat(15, 0), dt)
    np.testing.assert_array_equal(r, eqop_np(eqop, r0, y))


@pytest.mark.parametrize("prec", ["F", "D"])
def test_diagonal_fill_zero_signs(orc, prec):
    """The off-diagonal elements take QLA_c_<eqop>_r(e, 0.): peq turns a -0
    real part into +0 and leaves the imaginary part alone, meq keeps -0,
    eqm writes (-0, +0) (include/qla_complex.h:140-145)."""
    dt = DT[prec]
    c = np.array([dt(2.0), dt(3.0)])
    for eqop, want_re_neg, want_im_neg in ((1, False, True), (2, True, True), (3, True, False), (0, False, False)):
        r = np.full((2, 3, 3, 2), -0.0, dtype=dt)
        orc.fill(orc.FILL_DIAG, eqop, r, c)
        off = ~np.eye(3, dtype=bool)
        assert np.all(r[:, off, 0] == 0) and np.all(np.signbit(r[:, off, 0]) == want_re_neg), eqop
        assert np.all(r[:, off, 1] == 0) and np.all(np.signbit(r[:, off, 1]) == want_im_neg), eqop


def gamma_doc(mu):
    """Gamma_mu from the documented basis: the product of G[b] over the set
    bits b of mu, left to right."""
    from gamma_basis import G
    m = np.eye(4, dtype=complex)
    for b in range(4):
        if mu >> b & 1:
            m = m @ G[b]
    return m


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("mu", range(16))
def test_gamma_left_right_vs_documented_basis(orc, prec, mu):
    """r = Gamma_mu a for D and P and r = a Gamma_mu for P against numpy with
    the gamma matrices of doc/qla.tex (exact: entries are +-1, +-i)."""
    dt = DT[prec]
    g = gamma_doc(mu)
    d = field("D", 7, 50 + mu, dt)
    r = np.zeros_like(d)
    orc.gamma(False, mu, r, d)
    np.testing.assert_array_equal(cx(r), np.einsum("st,nct->ncs", g, cx(d)))
    pr = field("P", 5, 70 + mu, dt)
    r = np.zeros_like(pr)
    orc.gamma(False, mu, r, pr)
    np.testing.assert_array_equal(cx(r), np.einsum("st,nctdu->ncsdu", g, cx(pr)))
    r = np.zeros_like(pr)
    orc.gamma(True, mu, r, pr)
    np.testing.assert_array_equal(cx(r), np.einsum("ncsdt,tu->ncsdu", cx(pr), g))


def test_gamma_identities(orc):
    """Gamma_15 = gamma_5 = diag(1, 1, -1, -1); each gamma_mu squares to one;
    a zero record keeps the macro signs (-1 rows give -0 through QLA_c_eqm_c)."""
    d = field("D", 4, 90, np.float64)
    t1 = np.zeros_like(d)
    t2 = np.zeros_like(d)
    for b in range(4):
        orc.gamma(False, 1 << b, t1, d)
        orc.gamma(False, 1 << b, t2, t1)
        np.testing.assert_array_equal(t2, d)
    orc.gamma(False, 15, t1, d)
    np.testing.assert_array_equal(t1[:, :, :2], d[:, :, :2])
    np.testing.assert_array_equal(t1[:, :, 2:], -d[:, :, 2:])
    z = np.zeros_like(d)
    orc.gamma(False, 15, t1, z)
    assert not np.any(np.signbit(t1[:, :, :2])) and np.all(np.signbit(t1[:, :, 2:]))
    t1[:] = 7.0
    orc.gamma(False, 16, t1, d)
    assert np.all(t1 == 7.0)


def test_gamma_index_forms(orc):
    d = field("D", 30, 91, np.float32)
    rng = np.random.default_rng(92)
    perm = rng.permutation(30)
    r = np.zeros_like(d)
    orc.gamma(False, 6, r, d, a_idx=perm)
    want = np.zeros_like(d)
    orc.gamma(False, 6, want, np.ascontiguousarray(d[perm]))
    np.testing.assert_array_equal(r, want)
    ir = rng.permutation(30)[:9]
    ia = rng.permutation(30)[:9]
    r = field("D", 30, 93, np.float32)
    want = r.copy()
    sub = np.zeros((9,) + d.shape[1:], dtype=d.dtype)
    orc.gamma(False, 11, sub, np.ascontiguousarray(d[ia]))
    want[ir] = sub
    orc.gamma(False, 11, r, d, ir=ir, ia=ia)
    np.testing.assert_array_equal(r, want)


# ---- site-local reductions and colour-matrix parts (orc_*_local) ----------

def _dot_exact(a, b):
    """per-record sum of conj(a) b in long double: (n,) complex."""
    ca = cx(a.astype(np.longdouble)).reshape(a.shape[0], -1)
    cb = cx(b.astype(np.longdouble)).reshape(b.shape[0], -1)
    return np.sum(np.conj(ca) * cb, axis=1)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("t", TYPES)
@pytest.mark.parametrize("eqop", range(4))
def test_local_norm2_and_dots_vs_long_double(orc, prec, t, eqop):
    """R eqop norm2, C eqop a.b, R eqop Re a.b per record: the double sum is
    within (n_comp + 2) double roundings of sum |a||b| of the long-double value,
    and the destination is that rounded to the field precision (one more
    rounding of the result and of r0 +- x)."""
    dt = DT[prec]
    n = 25
    a = field(t, n, 60, dt)
    b = field(t, n, 61, dt)
    ld = np.longdouble
    ncomp = a[0].size // 2
    sab = np.sum(np.abs(cx(a.astype(ld))) * np.abs(cx(b.astype(ld))), axis=tuple(range(1, a.ndim - 1)))
    saa = np.sum(np.abs(cx(a.astype(ld))) ** 2, axis=tuple(range(1, a.ndim - 1)))
    acc = {orc.L_NORM2: saa, orc.L_REDOT: sab, orc.L_DOT: sab}
    eps_d = np.finfo(np.float64).eps
    tol_k = {}
    for k, s in acc.items():
        tol_k[k] = (ncomp + 2) * eps_d * s
    tol = lambda v, k=orc.L_REDOT: 2 * np.finfo(dt).eps * np.abs(v) + tol_k[k] + 4 * np.finfo(dt).tiny   # noqa: E731
    r0 = field("V", n, 62, dt)[:, 0, 0].copy()             # n reals
    c0 = np.ascontiguousarray(field("V", n, 63, dt)[:, 0])  # n complexes (n, 2)
    for kind, want in ((orc.L_NORM2, _dot_exact(a, a).real), (orc.L_REDOT, _dot_exact(a, b).real)):
        r = r0.copy()
        orc.local(kind, eqop, r, a, None if kind == orc.L_NORM2 else b)
        w = {0: want, 1: r0.astype(ld) + want, 2: r0.astype(ld) - want, 3: -want}[eqop]
        assert np.all(np.abs(r.astype(ld) - w) <= tol(w, kind)), (kind, eqop)
    c = c0.copy()
    orc.local(orc.L_DOT, eqop, c, a, b)
    want = _dot_exact(a, b)
    c0c = c0[:, 0].astype(ld) + 1j * c0[:, 1].astype(ld)
    w = {0: want, 1: c0c + want, 2: c0c - want, 3: -want}[eqop]
    assert np.all(np.abs(c[:, 0].astype(ld) - w.real) <= 2 * tol(w.real, orc.L_DOT) + tol(np.abs(want), orc.L_DOT))
    assert np.all(np.abs(c[:, 1].astype(ld) - w.imag) <= 2 * tol(w.imag, orc.L_DOT) + tol(np.abs(want), orc.L_DOT))


def test_local_norm2_exact_on_integers(orc):
    """Small-integer records: every step is exact, so norm2 and the dots equal
    the integer sums; the F3 norm2 of the sum sits on the float grid."""
    rng = np.random.default_rng(64)
    for dt in (np.float32, np.float64):
        a = rng.integers(-9, 10, size=(11, 3, 4, 3, 4, 2)).astype(dt)
        b = rng.integers(-9, 10, size=a.shape).astype(dt)
        r = np.zeros(11, dtype=dt)
        orc.local(orc.L_NORM2, 0, r, a)
        np.testing.assert_array_equal(r, (a.astype(np.int64) ** 2).reshape(11, -1).sum(1))
        c = np.zeros((11, 2), dtype=dt)
        orc.local(orc.L_DOT, 0, c, a, b)
        ia, ib = cx(a.astype(np.int64)).reshape(11, -1), cx(b.astype(np.int64)).reshape(11, -1)
        want = np.sum(np.conj(ia) * ib, axis=1)
        np.testing.assert_array_equal(c[:, 0], want.real)
        np.testing.assert_array_equal(c[:, 1], want.imag)


@pytest.mark.parametrize("prec", ["F", "D"])
def test_traces(orc, prec):
    dt = DT[prec]
    m = field("M", 30, 65, dt)
    tr = np.trace(cx(m.astype(np.float64)), axis1=1, axis2=2)
    u = np.finfo(dt).eps
    bound = 4 * u * np.abs(cx(m)).reshape(30, -1).sum(1)
    r = np.zeros(30, dtype=dt)
    orc.local(orc.L_RETRACE, 0, r, m)
    assert np.all(np.abs(r - tr.real) <= bound)
    orc.local(orc.L_IMTRACE, 0, r, m)
    assert np.all(np.abs(r - tr.imag) <= bound)
    c = np.zeros((30, 2), dtype=dt)
    orc.local(orc.L_TRACE, 0, c, m)
    assert np.all(np.abs(c[:, 0] - tr.real) <= bound) and np.all(np.abs(c[:, 1] - tr.imag) <= bound)
    # the diagonal sum in order, in the field precision: exactly numpy's left-to-right float sum
    d = cx(m)[:, [0, 1, 2], [0, 1, 2]]
    np.testing.assert_array_equal(c[:, 0], (d.real[:, 0] + d.real[:, 1]) + d.real[:, 2])


@pytest.mark.parametrize("prec", ["F", "D"])
def test_antiherm_part(orc, prec):
    """D is exactly anti-Hermitian (D_ji = -conj D_ij, real diagonal zero),
    nearly traceless, equals (A - A^dagger)/2 - i Im tr(A)/3 within rounding,
    and maps an anti-Hermitian traceless matrix with exactly zero trace to itself."""
    dt = DT[prec]
    m = field("M", 40, 66, dt)
    d = np.zeros_like(m)
    orc.local(orc.L_ANTIHERM, 0, d, m)
    D = cx(d)
    np.testing.assert_array_equal(D, -np.conj(np.swapaxes(D, 1, 2)))
    A = cx(m.astype(np.float64))
    want = 0.5 * (A - np.conj(np.swapaxes(A, 1, 2))) - 1j * (np.trace(A, axis1=1, axis2=2).imag / 3)[:, None, None] * np.eye(3)
    u = np.finfo(dt).eps
    assert np.all(np.abs(D - want) <= 8 * u * (1 + np.abs(A).max(axis=(1, 2)))[:, None, None])
    assert np.all(np.abs(np.trace(D, axis1=1, axis2=2)) <= 8 * u * np.abs(A).max(axis=(1, 2)))
    # anti-Hermitian input whose diagonal imaginary parts are integers summing to zero
    q = np.zeros((3, 3), dtype=complex)
    q[0, 1], q[0, 2], q[1, 2] = 0.5 + 0.25j, -1.0 + 2j, 0.75 - 0.5j
    q = q - np.conj(q.T)
    q[np.diag_indices(3)] = 1j * np.array([3.0, -1.0, -2.0])
    qa = re_im(q[None], dt)
    out = np.zeros_like(qa)
    orc.local(orc.L_ANTIHERM, 0, out, qa)
    np.testing.assert_array_equal(out, qa)


def test_local_index_forms(orc):
    a = field("D", 30, 67, np.float32)
    b = field("D", 30, 68, np.float32)
    rng = np.random.default_rng(69)
    pa, pb = rng.permutation(30), rng.permutation(30)
    c = np.zeros((30, 2), dtype=np.float32)
    orc.local(orc.L_DOT, 0, c, a, b, a_idx=pa, b_idx=pb)
    want = np.zeros((30, 2), dtype=np.float32)
    orc.local(orc.L_DOT, 0, want, np.ascontiguousarray(a[pa]), np.ascontiguousarray(b[pb]))
    np.testing.assert_array_equal(c, want)
    ir = rng.permutation(30)[:8]
    ia = rng.permutation(30)[:8]
    r = np.full(30, 3.0, dtype=np.float32)
    orc.local(orc.L_NORM2, 1, r, a, ir=ir, ia=ia)
    sub = np.full(8, 3.0, dtype=np.float32)
    orc.local(orc.L_NORM2, 1, sub, np.ascontiguousarray(a[ia]))
    want = np.full(30, 3.0, dtype=np.float32)
    want[ir] = sub
    np.testing.assert_array_equal(r, want)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("t", TYPES)
def test_field_sum(orc, prec, t):
    """Global sums: exact integer sums on small integers; within n double
    roundings of the long-double sum on Gaussian data; x-form and pointer
    targets sum the records they name."""
    dt = DT[prec]
    rng = np.random.default_rng(80)
    a = rng.integers(-50, 51, size=field(t, 300, 0, dt).shape).astype(dt)
    np.testing.assert_array_equal(orc.field_sum(a), a.astype(np.int64).sum(axis=0))
    g = field(t, 300, 81, dt)
    want = g.astype(np.longdouble).sum(axis=0)
    bound = 300 * np.finfo(np.float64).eps * np.abs(g.astype(np.longdouble)).sum(axis=0)
    assert np.all(np.abs(orc.field_sum(g) - want) <= bound)
    idx = np.sort(rng.choice(300, 77, replace=False))
    np.testing.assert_array_equal(orc.field_sum(a, index=idx), a[idx].astype(np.int64).sum(axis=0))
    perm = rng.permutation(300)[:50]
    np.testing.assert_array_equal(orc.field_sum(a, a_idx=perm, n=50), a[perm].astype(np.int64).sum(axis=0))
    assert not np.any(orc.field_sum(a, n=0))


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("eqop", range(4))
@pytest.mark.parametrize("t,adj", SCALE_FORMS)
def test_real_field_times_vs_numpy(orc, prec, eqop, t, adj):
    """r eqop f(x) b(x) for a real field f: numpy's products, one rounding each."""
    dt = DT[prec]
    n = 19
    b = field(t, n, 100, dt)
    r0 = field(t, n, 101, dt)
    f = np.random.default_rng(102).standard_normal(n).astype(dt)
    r = r0.copy()
    dim = orc.TRANSPOSE_DIM[t] if adj else 0
    orc.fieldmul(eqop, r, f, b, cb=adj, dim=dim)
    c = cx(b)
    if adj:
        c = np.conj(transpose(t, c))
    src = re_im(c, dt)
    fb = f.reshape((n,) + (1,) * (src.ndim - 1)) * src
    np.testing.assert_array_equal(r, eqop_np(eqop, r0, fb))


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("eqop", range(4))
@pytest.mark.parametrize("fa", [False, True])
@pytest.mark.parametrize("t,adj", SCALE_FORMS)
def test_complex_field_times_exact(orc, prec, eqop, fa, t, adj):
    """r eqop (conj) f(x) * (adjoint) b(x) on small integers equals the
    complex128 product."""
    dt = DT[prec]
    n = 17
    b = _int_field(t, n, 103, dt)
    r0 = _int_field(t, n, 104, dt)
    f = np.random.default_rng(105).integers(-9, 10, size=(n, 2)).astype(dt)
    r = r0.copy()
    dim = orc.TRANSPOSE_DIM[t] if adj else 0
    orc.fieldmul(eqop, r, f, b, fa=fa, cb=adj, dim=dim)
    c = cx(b.astype(np.float64))
    if adj:
        c = np.conj(transpose(t, c))
    fz = f[:, 0].astype(np.float64) + 1j * f[:, 1]
    if fa:
        fz = np.conj(fz)
    y = fz.reshape((n,) + (1,) * (c.ndim - 1)) * c
    c0 = cx(r0.astype(np.float64))
    want = {0: y, 1: c0 + y, 2: c0 - y, 3: -y}[eqop]
    np.testing.assert_array_equal(cx(r.astype(np.float64)), want)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("eqop", range(4))
def test_outer_product_exact(orc, prec, eqop):
    dt = DT[prec]
    a = _int_field("V", 13, 106, dt)
    b = _int_field("V", 13, 107, dt)
    r0 = _int_field("M", 13, 108, dt)
    r = r0.copy()
    orc.outer(eqop, r, a, b)
    y = cx(a.astype(np.float64))[:, :, None] * np.conj(cx(b.astype(np.float64)))[:, None, :]
    c0 = cx(r0.astype(np.float64))
    want = {0: y, 1: c0 + y, 2: c0 - y, 3: -y}[eqop]
    np.testing.assert_array_equal(cx(r.astype(np.float64)), want)


def _as12(c):
    """(n, 3, 4, 3, 4) complex propagators as (n, 12, 12) matrices over (colour, spin)"""
    return c.reshape(c.shape[0], 12, 12)


def _embed_m(m):
    """colour matrix acting on colour only: M (x) 1_spin as a 12 x 12 matrix"""
    return np.einsum("nab,st->nasbt", m, np.eye(4)).reshape(m.shape[0], 12, 12)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("eqop", range(4))
@pytest.mark.parametrize("ca,cb", [(False, False), (True, False), (False, True), (True, True)])
def test_propagator_products_exact(orc, prec, kind, eqop, ca, cb):
    """P eqop A*B for M*P, P*M and P*P, with either factor's adjoint, exactly
    against complex128 products of the 12 x 12 matrices (colour matrices as
    M (x) 1 over spin) on small integers."""
    dt = DT[prec]
    n = 5
    ta, tb = {0: ("M", "P"), 1: ("P", "M"), 2: ("P", "P")}[kind]
    a = _int_field(ta, n, 110 + kind, dt)
    b = _int_field(tb, n, 120 + kind, dt)
    r0 = _int_field("P", n, 130, dt)
    r = r0.copy()
    orc.pmul(kind, eqop, r, a, b, ca=ca, cb=cb)
    A = cx(a.astype(np.float64))
    B = cx(b.astype(np.float64))
    A = _embed_m(A) if ta == "M" else _as12(A)
    B = _embed_m(B) if tb == "M" else _as12(B)
    if ca:
        A = np.conj(np.swapaxes(A, 1, 2))
    if cb:
        B = np.conj(np.swapaxes(B, 1, 2))
    y = np.einsum("nij,njk->nik", A, B)
    c0 = _as12(cx(r0.astype(np.float64)))
    want = {0: y, 1: c0 + y, 2: c0 - y, 3: -y}[eqop]
    np.testing.assert_array_equal(_as12(cx(r.astype(np.float64))), want)


def test_propagator_product_matches_colour_products(orc):
    """M*P for each fixed (is, js) spin pair is the colour-matrix product of M
    with that 3 x 3 colour block, bit for bit (the same kc order and macros)."""
    m = field("M", 6, 140, np.float32)
    pr = field("P", 6, 141, np.float32)
    r = np.zeros_like(pr)
    orc.pmul(0, 0, r, m, pr)
    for is_ in range(4):
        for js in range(4):
            blk = np.ascontiguousarray(pr[:, :, is_, :, js])
            want = np.zeros_like(blk)
            orc.site(orc.MAT_M, 0, want, m, blk)
            np.testing.assert_array_equal(r[:, :, is_, :, js], want)


# (type, reference extra-argument count for elem / colorvec); the element is
# indexed as the numpy record shape (ic[, is][, jc][, js]) reads it
ELEM_ARGS = {"V": 1, "H": 2, "D": 2, "M": 2, "P": 4}
CVEC_ARGS = {"M": 1, "H": 1, "D": 1, "P": 3}
EXTENT = {"V": (3,), "H": (3, 2), "D": (3, 4), "M": (3, 3), "P": (3, 4, 3, 4)}


def _cvec_slice(t, e):
    """numpy index of the colour column of a record (every ic, the rest fixed)"""
    if t == "M":
        return (slice(None), slice(None), e[0])
    if t in ("H", "D"):
        return (slice(None), slice(None), e[0])
    return (slice(None), slice(None), e[0], e[1], e[2])


def _all_args(shape):
    return [tuple(int(v) for v in ix) for ix in np.ndindex(*shape)]


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("t", list(ELEM_ARGS))
def test_elem_get_set(orc, prec, t):
    """C = elem(a, ic, ...) and its insertion, for every element of the record,
    against numpy indexing of the record shape (an exact copy; the other
    elements of an inserted record untouched)."""
    dt = DT[prec]
    n = 4
    a = field(t, n, 150, dt)
    for e in _all_args(EXTENT[t]):
        c = np.zeros((n, 2), dt)
        orc.comp(orc.GET_ELEM, t, e, c, a)
        np.testing.assert_array_equal(c, a[(slice(None),) + e])
        r = field(t, n, 151, dt)
        want = r.copy()
        src = field("V", n, 152, dt)[:, 0]
        want[(slice(None),) + e] = src
        orc.comp(orc.SET_ELEM, t, e, r, np.ascontiguousarray(src))
        np.testing.assert_array_equal(r, want)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("t", list(CVEC_ARGS))
def test_colorvec_get_set(orc, prec, t):
    """V = colorvec(a, ...) and its insertion for every choice of the fixed
    indices, against numpy slicing."""
    dt = DT[prec]
    n = 4
    a = field(t, n, 160, dt)
    for e in _all_args(EXTENT[t][1:] if t != "M" else (3,)):
        if t in ("H", "D"):
            e = e[:1]
        sl = _cvec_slice(t, e)
        v = np.zeros((n, 3, 2), dt)
        orc.comp(orc.GET_COLORVEC, t, e, v, a)
        np.testing.assert_array_equal(v, a[sl])
        r = field(t, n, 161, dt)
        src = field("V", n, 162, dt)
        want = r.copy()
        want[sl] = src
        orc.comp(orc.SET_COLORVEC, t, e, r, src)
        np.testing.assert_array_equal(r, want)


@pytest.mark.parametrize("prec", ["F", "D"])
def test_diracvec_get_set(orc, prec):
    """D = diracvec(P, jc, js) and its insertion for all 12 columns, against
    numpy slicing; inserting all 12 columns of a propagator rebuilds it."""
    dt = DT[prec]
    n = 5
    a = field("P", n, 170, dt)
    rebuilt = np.zeros_like(a)
    for jc in range(3):
        for js in range(4):
            d = np.zeros((n, 3, 4, 2), dt)
            orc.comp(orc.GET_DIRACVEC, "P", (jc, js), d, a)
            np.testing.assert_array_equal(d, a[:, :, :, jc, js])
            orc.comp(orc.SET_DIRACVEC, "P", (jc, js), rebuilt, d)
    np.testing.assert_array_equal(rebuilt, a)


@pytest.mark.parametrize("prec", ["F", "D"])
def test_spintrace(orc, prec):
    """M = spintrace(P): exactly the spin-diagonal sum on small integers, and
    on random values the ordered sum ((0 + p0) + p1) + p2 + p3 in the field
    precision; tr_colour(spintrace P) = the full 12 x 12 trace."""
    dt = DT[prec]
    n = 6
    a = _int_field("P", n, 180, dt)
    m = np.zeros((n, 3, 3, 2), dt)
    orc.comp(orc.SPINTRACE, "P", (), m, a)
    np.testing.assert_array_equal(cx(m.astype(np.float64)), np.einsum("nasbs->nab", cx(a.astype(np.float64))))
    a = field("P", n, 181, dt)
    orc.comp(orc.SPINTRACE, "P", (), m, a)
    want = np.zeros((n, 3, 3, 2), dt)
    for s in range(4):
        want = (want + a[:, :, s, :, s]).astype(dt)
    np.testing.assert_array_equal(m, want)
    full = np.trace(cx(a.astype(np.float64)).reshape(n, 12, 12), axis1=1, axis2=2)
    np.testing.assert_allclose(np.trace(cx(m.astype(np.float64)), axis1=1, axis2=2), full,
                               rtol=0, atol=1e-5 if prec == "F" else 1e-13)


R_OPS = ["copy", "r_times", "plus", "minus", "r_times_plus", "r_times_minus"]


R_TWO = ("plus", "minus", "r_times_plus", "r_times_minus")
R_CASES = [(f, e) for f in R_OPS for e in (range(1) if f in R_TWO else range(4))]


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("fam,eqop", R_CASES)
def test_real_field_ops(orc, prec, fam, eqop):
    """Real fields (type R): r eqop a, r eqop s*a (every eqop), a +- b and
    s*a +- b (eq) against numpy in the field precision (the same IEEE
    operations one by one), signed zeros included."""
    dt = DT[prec]
    n = 64
    a = field("R", n, 200, dt)
    b = field("R", n, 201, dt)
    a[:4] = [0.0, -0.0, 0.0, -0.0]
    b[:4] = [-0.0, 0.0, 0.0, -0.0]
    r0 = field("R", n, 202, dt)
    s = dt(-0.7071067811865476)
    op = {"copy": orc.COPY, "r_times": orc.RSCALE, "plus": orc.PLUS, "minus": orc.MINUS,
          "r_times_plus": orc.R_PLUS, "r_times_minus": orc.R_MINUS}[fam]
    r = r0.copy()
    orc.vecop(op, eqop, r, a, b if fam in ("plus", "minus", "r_times_plus", "r_times_minus") else None, s=float(s))
    v = {"copy": a, "r_times": s * a, "plus": a + b, "minus": a - b, "r_times_plus": s * a + b,
         "r_times_minus": s * a - b}[fam]
    want = {0: v, 1: r0 + v, 2: r0 - v, 3: -v}[eqop]
    np.testing.assert_array_equal(r.view(np.uint32 if prec == "F" else np.uint64),
                                  want.astype(dt).view(np.uint32 if prec == "F" else np.uint64))


C_OPS = ["copy", "conj", "adjoint", "i", "r_times", "c_times", "plus", "minus", "r_times_plus", "r_times_minus",
         "c_times_plus", "c_times_minus"]


C_TWO = ("plus", "minus", "r_times_plus", "r_times_minus", "c_times_plus", "c_times_minus")
C_CASES = [(f, e) for f in C_OPS for e in (range(1) if f in C_TWO else range(4))]


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("fam,eqop", C_CASES)
def test_complex_field_ops(orc, prec, fam, eqop):
    """Complex fields (type C, one complex per record) exactly against
    complex128 on small integers, every family and eqop it has; the adjoint
    of a complex is its conjugate."""
    two = fam in C_TWO
    dt = DT[prec]
    n = 40
    a = _int_field("C", n, 210, dt)
    b = _int_field("C", n, 211, dt)
    r0 = _int_field("C", n, 212, dt)
    sr, sc = -3.0, complex(2.0, -5.0)
    op = {"copy": orc.COPY, "conj": orc.COPY, "adjoint": orc.COPY, "i": orc.TIMES_I, "r_times": orc.RSCALE,
          "c_times": orc.CSCALE, "plus": orc.PLUS, "minus": orc.MINUS, "r_times_plus": orc.R_PLUS,
          "r_times_minus": orc.R_MINUS, "c_times_plus": orc.C_PLUS, "c_times_minus": orc.C_MINUS}[fam]
    s = sc if fam.startswith("c_times") else sr
    r = r0.copy()
    orc.vecop(op, eqop, r, a, b if two else None, s=s, conj=fam in ("conj", "adjoint"))
    A, B, R0 = cx(a.astype(np.float64)), cx(b.astype(np.float64)), cx(r0.astype(np.float64))
    v = {"copy": A, "conj": np.conj(A), "adjoint": np.conj(A), "i": 1j * A, "r_times": sr * A, "c_times": sc * A,
         "plus": A + B, "minus": A - B, "r_times_plus": sr * A + B, "r_times_minus": sr * A - B,
         "c_times_plus": sc * A + B, "c_times_minus": sc * A - B}[fam]
    want = {0: v, 1: R0 + v, 2: R0 - v, 3: -v}[eqop]
    np.testing.assert_array_equal(cx(r.astype(np.float64)), want)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("t", ["R", "C"])
def test_scalar_field_fills(orc, prec, t):
    """R and C fills: zero, and r eqop a with one constant, every eqop."""
    dt = DT[prec]
    n = 20
    r0 = field(t, n, 220, dt)
    a = np.ascontiguousarray(field(t, 1, 221, dt)[0])
    r = r0.copy()
    orc.fill(orc.FILL_ZERO, 0, r)
    np.testing.assert_array_equal(r, np.zeros_like(r))
    for eqop in range(4):
        r = r0.copy()
        orc.fill(orc.FILL_REC, eqop, r, a)
        want = {0: np.broadcast_to(a, r0.shape), 1: r0 + a, 2: r0 - a, 3: np.broadcast_to(-a, r0.shape)}[eqop]
        np.testing.assert_array_equal(r, want.astype(dt))


@pytest.mark.parametrize("prec", ["F", "D"])
def test_real_complex_conversions(orc, prec):
    """C = R, C = R + iR, R = re C, R = im C as exact copies; R / R as
    numpy's correctly rounded quotient in the field precision."""
    dt = DT[prec]
    n = 50
    ra, rb = field("R", n, 230, dt), field("R", n, 231, dt)
    c = field("C", n, 232, dt)
    out = np.full((n, 2), 7, dt)
    orc.sconv(0, out, ra)
    np.testing.assert_array_equal(out, np.stack([ra, np.zeros_like(ra)], -1))
    orc.sconv(1, out, ra, rb)
    np.testing.assert_array_equal(out, np.stack([ra, rb], -1))
    rr = np.zeros(n, dt)
    orc.sconv(2, rr, c)
    np.testing.assert_array_equal(rr, c[:, 0])
    orc.sconv(3, rr, c)
    np.testing.assert_array_equal(rr, c[:, 1])
    orc.sconv(4, rr, ra, rb)
    np.testing.assert_array_equal(rr, (ra / rb).astype(dt))


@pytest.mark.parametrize("prec", ["F", "D"])
def test_complex_division(orc, prec):
    """C / C exactly against complex128 where every step is exact (integer
    numerators, denominators with |b|^2 a power of two), and within a few
    units in the last place of complex128 division on random values."""
    dt = DT[prec]
    rng = np.random.default_rng(240)
    n = 64
    a = rng.integers(-50, 51, size=(n, 2)).astype(dt)
    b = np.array([[1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [1, -1], [-2, 2], [4, 0]] * 8, dt)
    r = np.zeros((n, 2), dt)
    orc.sconv(5, r, a, b)
    np.testing.assert_array_equal(cx(r.astype(np.float64)), cx(a.astype(np.float64)) / cx(b.astype(np.float64)))
    a = field("C", 4096, 241, dt)
    b = field("C", 4096, 242, dt)
    r = np.zeros_like(a)
    orc.sconv(5, r, a, b)
    want = cx(a.astype(np.float64)) / cx(b.astype(np.float64))
    eps = np.finfo(dt).eps
    err = np.abs(cx(r.astype(np.float64)) - want) / np.abs(want)
    assert err.max() < 8 * eps


SFN_UNARY = ["cos", "sin", "tan", "acos", "asin", "atan", "sqrt", "fabs", "exp", "log", "sign", "cosh", "sinh",
             "tanh", "log10", "floor", "ceil"]
_NP = {"cos": np.cos, "sin": np.sin, "tan": np.tan, "acos": np.arccos, "asin": np.arcsin, "atan": np.arctan,
       "sqrt": np.sqrt, "fabs": np.abs, "exp": np.exp, "log": np.log, "sign": lambda x: np.where(x >= 0, 1.0, -1.0),
       "cosh": np.cosh, "sinh": np.sinh, "tanh": np.tanh, "log10": np.log10, "floor": np.floor, "ceil": np.ceil}
SFN_EXACT = ("sqrt", "fabs", "sign", "floor", "ceil")


def sfn_domain(name, n, seed, dt):
    """inputs inside the function's domain, with signed zeros where it has them"""
    rng = np.random.default_rng(seed)
    if name in ("acos", "asin"):
        x = rng.uniform(-1, 1, n)
    elif name in ("sqrt", "log", "log10"):
        x = rng.uniform(1e-3, 50, n)
    elif name in ("exp", "cosh", "sinh"):
        x = rng.uniform(-20, 20, n)
    else:
        x = rng.uniform(-10, 10, n)
    x = x.astype(dt)
    if name not in ("sqrt", "log", "log10"):
        x[:2] = [0.0, -0.0]
    return x


def ulps(a, b, dt):
    """distance in units in the last place between same-typed arrays"""
    it = np.int32 if dt == np.float32 else np.int64
    ia, ib = a.astype(dt).view(it).astype(np.int64), b.astype(dt).view(it).astype(np.int64)
    ia = np.where(ia < 0, np.iinfo(it).min - ia, ia)
    ib = np.where(ib < 0, np.iinfo(it).min - ib, ib)
    return np.abs(ia - ib)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("k,name", list(enumerate(SFN_UNARY)))
def test_sfn_unary(orc, prec, k, name):
    """R = f(R) against numpy in float64 rounded to the field type: exact for
    sqrt, fabs, sign, floor and ceil, within 2 units in the last place for
    the transcendental functions (the C library's own accuracy)."""
    dt = DT[prec]
    x = sfn_domain(name, 300, 250 + k, dt)
    r = np.zeros_like(x)
    orc.sfn(k, r, x)
    want = _NP[name](x.astype(np.float64)).astype(dt)
    if name in SFN_EXACT:
        np.testing.assert_array_equal(r, want)
    else:
        assert ulps(r, want, dt).max() <= 2


@pytest.mark.parametrize("prec", ["F", "D"])
def test_sfn_binary_and_complex(orc, prec):
    """fmod, max, min, ldexp exactly; pow and atan2 in double rounded to the
    field type; cexpi, norm, arg, cexp, csqrt, clog against complex128."""
    dt = DT[prec]
    rng = np.random.default_rng(260)
    n = 300
    a = rng.uniform(-10, 10, n).astype(dt)
    b = rng.uniform(-10, 10, n).astype(dt)
    r = np.zeros(n, dt)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    for fn, want in ((20, np.fmod(a64, b64)), (21, np.where(a > b, a, b)), (22, np.where(a > b, b, a)),
                     (24, np.arctan2(a64, b64))):
        orc.sfn(fn, r, a, b)
        if fn == 24:
            assert ulps(r, want.astype(dt), dt).max() <= 1
        else:
            np.testing.assert_array_equal(r, want.astype(dt))
    pa = np.abs(a) + dt(0.5)
    orc.sfn(23, r, pa, b)
    assert ulps(r, np.power(pa.astype(np.float64), b64).astype(dt), dt).max() <= 1
    e = rng.integers(-20, 20, n).astype(np.int32)
    orc.sfn(25, r, a, e)
    np.testing.assert_array_equal(r, np.ldexp(a64, e).astype(dt))
    c = np.zeros((n, 2), dt)
    orc.sfn(30, c, a)
    z = np.exp(1j * a64)
    assert ulps(c[:, 0], z.real.astype(dt), dt).max() <= 2 and ulps(c[:, 1], z.imag.astype(dt), dt).max() <= 2
    zc = field("C", n, 261, dt)
    Z = cx(zc.astype(np.float64))
    orc.sfn(31, r, zc)
    np.testing.assert_array_equal(r, np.sqrt((zc[:, 0] * zc[:, 0] + zc[:, 1] * zc[:, 1]).astype(dt)))
    orc.sfn(32, r, zc)
    assert ulps(r, np.angle(Z).astype(dt), dt).max() <= 2
    eps = np.finfo(dt).eps
    for fn, f in ((33, np.exp), (34, np.sqrt), (35, np.log)):
        orc.sfn(fn, c, zc)
        got = cx(c.astype(np.float64))
        assert (np.abs(got - f(Z)) / np.abs(f(Z))).max() < 16 * eps


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("t", ["R", "C", "V", "D", "P"])
def test_copymask(orc, prec, t):
    """r = a exactly where the mask is nonzero (any nonzero int), untouched elsewhere."""
    dt = DT[prec]
    n = 30
    a = field(t, n, 270, dt)
    r0 = field(t, n, 271, dt)
    m = np.random.default_rng(272).integers(-2, 3, n).astype(np.int32)
    r = r0.copy()
    orc.mask(-1, r, a, m)
    sel = (m != 0).reshape((n,) + (1,) * (a.ndim - 1))
    np.testing.assert_array_equal(r, np.where(sel, a, r0))


@pytest.mark.parametrize("prec", ["F", "D"])
def test_real_comparisons(orc, prec):
    """I = (a op b) as numpy's comparisons, with ties, signed zeros and NaN."""
    dt = DT[prec]
    rng = np.random.default_rng(273)
    a = rng.integers(-3, 4, 60).astype(dt)
    b = rng.integers(-3, 4, 60).astype(dt)
    a[:4] = [0.0, -0.0, np.nan, 1.0]
    b[:4] = [-0.0, 0.0, 1.0, np.nan]
    for op, f in enumerate((np.equal, np.not_equal, np.greater, np.less, np.greater_equal, np.less_equal)):
        r = np.full(60, 7, np.int32)
        orc.mask(op, r, a, b)
        np.testing.assert_array_equal(r, f(a, b).astype(np.int32))


def _c_div(a, b):
    """C integer division and remainder (truncation toward zero)"""
    q = np.abs(a) // np.abs(b) * np.sign(a) * np.sign(b)
    return q.astype(np.int32), (a - q * b).astype(np.int32)


INT_REF = {
    0: lambda a, b, s: a, 1: lambda a, b, s: s * a, 2: lambda a, b, s: (a == 0).astype(np.int32),
    10: lambda a, b, s: a + b, 11: lambda a, b, s: a - b, 12: lambda a, b, s: _c_div(a, b)[0],
    13: lambda a, b, s: a * b, 14: lambda a, b, s: s * a + b, 15: lambda a, b, s: s * a - b,
    16: lambda a, b, s: np.left_shift(np.abs(a), np.abs(b) % 8) if False else None,
    18: lambda a, b, s: _c_div(a, b)[1], 19: lambda a, b, s: np.maximum(a, b), 20: lambda a, b, s: np.minimum(a, b),
    21: lambda a, b, s: a == b, 22: lambda a, b, s: a != b, 23: lambda a, b, s: a > b, 24: lambda a, b, s: a < b,
    25: lambda a, b, s: a >= b, 26: lambda a, b, s: a <= b, 27: lambda a, b, s: a | b, 28: lambda a, b, s: a & b,
    29: lambda a, b, s: a ^ b}


@pytest.mark.parametrize("op", [0, 1, 2, 10, 11, 12, 13, 14, 15, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29])
@pytest.mark.parametrize("eqop", range(4))
def test_integer_fields(orc, op, eqop):
    """Integer-field operations against numpy's int32 arithmetic (C division
    and remainder truncating toward zero), every eqop."""
    rng = np.random.default_rng(280 + op)
    n = 80
    a = rng.integers(-50, 51, n).astype(np.int32)
    b = rng.integers(-9, 10, n).astype(np.int32)
    b[b == 0] = 3
    a[:3] = [0, 0, -7]
    r0 = rng.integers(-20, 21, n).astype(np.int32)
    r = r0.copy()
    orc.int_op(op, eqop, r, a, b if op >= 10 else None, sc=-3)
    v = np.asarray(INT_REF[op](a, b, -3)).astype(np.int32)
    want = {0: v, 1: r0 + v, 2: r0 - v, 3: -v}[eqop]
    np.testing.assert_array_equal(r, want)


def test_integer_shifts_mask_and_fills(orc):
    """<< and >> on non-negative counts below 31 (arithmetic right shift of
    negatives), copymask with any nonzero mask, and the fills."""
    rng = np.random.default_rng(290)
    n = 60
    a = rng.integers(-1000, 1000, n).astype(np.int32)
    b = rng.integers(0, 20, n).astype(np.int32)
    r = np.zeros(n, np.int32)
    orc.int_op(16, 0, r, np.abs(a).astype(np.int32), b)
    np.testing.assert_array_equal(r, np.left_shift(np.abs(a).astype(np.int32), b))
    orc.int_op(17, 0, r, a, b)
    np.testing.assert_array_equal(r, np.right_shift(a, b))
    m = rng.integers(-2, 3, n).astype(np.int32)
    r0 = rng.integers(-5, 5, n).astype(np.int32)
    r = r0.copy()
    orc.int_op(30, 0, r, a, m)
    np.testing.assert_array_equal(r, np.where(m != 0, a, r0))
    orc.int_op(40, 0, r, n=n)
    np.testing.assert_array_equal(r, 0)
    for eqop, want in enumerate((lambda x: np.full_like(x, 9), lambda x: x + 9, lambda x: x - 9,
                                 lambda x: np.full_like(x, -9))):
        r = r0.copy()
        orc.int_op(41, eqop, r, sc=9, n=n)
        np.testing.assert_array_equal(r, want(r0))


@pytest.mark.parametrize("prec", ["F", "D"])
def test_real_integer_conversions(orc, prec):
    """R = I exactly; I = trunc R and I = round R as numpy's trunc and
    round-half-away-from-zero."""
    dt = DT[prec]
    rng = np.random.default_rng(295)
    i = rng.integers(-1000, 1000, 50).astype(np.int32)
    r = np.zeros(50, dt)
    orc.sconv(6, r, i)
    np.testing.assert_array_equal(r, i.astype(dt))
    x = (rng.uniform(-100, 100, 50)).astype(dt)
    x[:4] = [2.5, -2.5, 0.5, -1.5]
    out = np.zeros(50, np.int32)
    orc.sconv(7, out, x)
    np.testing.assert_array_equal(out, np.trunc(x).astype(np.int32))
    orc.sconv(8, out, x)
    np.testing.assert_array_equal(out, (np.sign(x) * np.floor(np.abs(x) + dt(0.5))).astype(np.int32))
