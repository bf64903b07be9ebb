"""Pins the oracle's site operations in every eqop, adjoint and operand form.

Each case compares the oracle (orc_*_site / orc_*_multi) with a long-double
evaluation built independently from numpy and the gamma basis of the
reference documentation (doc/qla.tex:1063-1113, typed in gamma_basis.py):

  kind 0   V  eqop A v
  kind 1   H  eqop A (P psi)            P: rows of (1 + s gamma_mu)
  kind 2   D  eqop R (A h)              R: (1 + s gamma_mu) = R P
  kind 3   D  eqop A (1 + s gamma_mu) psi   (mu = 0..3; (1 + s gamma_5)/2 for mu = 4)
  multi    V  eqop sum_d A_d v_d

with A = U or U^dagger and eqop eq (r = y), peq (r0 + y), meq (r0 - y),
eqm (-y), per component within a rounding bound that a dropped term, a
wrong sign, a transposed or unconjugated matrix, or a wrong eqop exceeds by
orders of magnitude.  Pointer operands (pM, pV/pD/pH) are exercised through
permuted record lists, the x-form through a parity subset.
"""
import numpy as np
import pytest

from gamma_basis import G, G5
from qla_amd import lattice
from test_oracle_pin import U32, U64, cplx, proj_recon_mats

EQOPS = ("eq", "peq", "meq", "eqm")
DEST = {0: "V", 1: "H", 2: "D", 3: "D"}
SRC2 = {0: "V", 1: "D", 2: "H", 3: "D"}
L = np.clongdouble


def apply_eqop(eqop, r0, y):
    return {0: y, 1: r0 + y, 2: r0 - y, 3: -y}[eqop]


def exact_site(kind, A, b, mu, sign):
    """(y, |y| bound) in long double; A (n,3,3), b records (n, 3[, s])."""
    if kind == 0:
        return np.einsum("nij,nj->ni", A, b), np.einsum("nij,nj->ni", np.abs(A), np.abs(b))
    P, R = proj_recon_mats(mu, sign)
    if kind == 1:
        t = np.einsum("sk,nck->ncs", P.astype(L), b)
        tb = np.einsum("sk,nck->ncs", np.abs(P), np.abs(b))
        return np.einsum("nab,nbs->nas", A, t), np.einsum("nab,nbs->nas", np.abs(A), tb)
    if kind == 2:
        w = np.einsum("nab,nbs->nas", A, b)
        wb = np.einsum("nab,nbs->nas", np.abs(A), np.abs(b))
        return np.einsum("ks,nas->nak", R.astype(L), w), np.einsum("ks,nas->nak", np.abs(R), wb)
    # kind 3: the full spin matrix (1 + s gamma), independent of the P/R split
    g = G[mu] if mu < 4 else G5
    S = (np.eye(4) + sign * g) if mu < 4 else 0.5 * (np.eye(4) + sign * g)
    sp = np.einsum("kl,ncl->nck", S.astype(L), b)
    spb = np.einsum("kl,ncl->nck", np.abs(S), np.abs(b))
    return np.einsum("nab,nbk->nak", A, sp), 2 * np.einsum("nab,nbk->nak", np.abs(A), spb)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("eqop", range(4))
@pytest.mark.parametrize("adj", [False, True])
def test_site_vs_exact(orc, prec, kind, eqop, adj):
    n = 160
    dt = np.float32 if prec == "F" else np.float64
    u_ = U32 if prec == "F" else U64
    rng = np.random.default_rng(100 * kind + 10 * eqop + adj)
    U = lattice.random_su3(n, 31 + kind).astype(dt)
    b = lattice.random_field(SRC2[kind], n, 32 + kind).astype(dt)
    r0 = lattice.random_field(DEST[kind], n, 33 + eqop).astype(dt)
    a_idx = rng.permutation(n)
    b_idx = rng.permutation(n)
    for ms in range(10 if kind else 1):
        mu, sign = ms % 5, (1 if ms < 5 else -1)
        for pa in (False, True):
            r = r0.copy()
            orc.site(kind, eqop, r, U, b, mu, sign, adj=adj, a_idx=a_idx if pa else None, b_idx=b_idx)
            Uc = cplx(U)[a_idx] if pa else cplx(U)
            A = np.conj(np.swapaxes(Uc, 1, 2)) if adj else Uc
            y, yb = exact_site(kind, A, cplx(b)[b_idx], mu, sign)
            exact = apply_eqop(eqop, cplx(r0), y)
            bound = 8 * u_ * (np.abs(cplx(r0)) + yb) + 1e-300
            err = np.abs(cplx(r) - exact)
            assert (err <= bound).all(), (EQOPS[eqop], mu, sign, pa, float(err.max()))


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("kind", [1, 2, 3])
def test_site_bad_mu_and_x_form(orc, prec, kind):
    dims = (4, 4, 2, 2)
    n = lattice.volume(dims)
    dt = np.float32 if prec == "F" else np.float64
    U = lattice.random_su3(n, 3).astype(dt)
    b = lattice.random_field(SRC2[kind], n, 4).astype(dt)
    r0 = lattice.random_field(DEST[kind], n, 5).astype(dt)
    fw, _ = lattice.neighbours(dims)
    for mu in (-1, 5):
        r = r0.copy()
        orc.site(kind, 0, r, U, b, mu, 1, b_idx=fw[0])
        np.testing.assert_array_equal(r, r0)
    even, odd = lattice.parity_sites(dims, 0), lattice.parity_sites(dims, 1)
    for eqop in range(4):
        r = r0.copy()
        orc.site(kind, eqop, r, U, b, 2, -1, index=even, b_idx=fw[1])
        np.testing.assert_array_equal(r[odd], r0[odd])
        full = r0.copy()
        orc.site(kind, eqop, full, U, b, 2, -1, b_idx=fw[1])
        np.testing.assert_array_equal(r[even], full[even])


@pytest.mark.parametrize("prec", ["F", "D"])
def test_wilson_is_recon_of_projected(orc, prec):
    """kind 3 = kind 2 applied to the projection of B, bit for bit: the
    projection into a temporary by eq (make_code_binary.pl:285-341)."""
    n = 200
    dt = np.float32 if prec == "F" else np.float64
    U = lattice.random_su3(n, 41).astype(dt)
    d = lattice.random_field("D", n, 42).astype(dt)
    r0 = lattice.random_field("D", n, 43).astype(dt)
    eye = np.zeros((n, 3, 3, 2), dtype=dt)
    eye[:, [0, 1, 2], [0, 1, 2], 0] = 1
    idx = np.random.default_rng(4).permutation(n)
    for ms in range(10):
        mu, sign = ms % 5, (1 if ms < 5 else -1)
        h = np.zeros((n, 3, 2, 2), dtype=dt)
        orc.site(1, 0, h, eye, d, mu, sign, b_idx=idx)          # h = 1 * spproj(d[idx]) exactly
        for eqop in range(4):
            for adj in (False, True):
                w = r0.copy()
                orc.site(3, eqop, w, U, d, mu, sign, adj=adj, b_idx=idx)
                v = r0.copy()
                orc.site(2, eqop, v, U, h, mu, sign, adj=adj)
                np.testing.assert_array_equal(w, v)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("eqop", range(4))
@pytest.mark.parametrize("adj", [False, True])
@pytest.mark.parametrize("nd", [1, 4, 17])
def test_multi_vs_exact(orc, prec, eqop, adj, nd):
    n = 120
    dt = np.float32 if prec == "F" else np.float64
    u_ = U32 if prec == "F" else U64
    rng = np.random.default_rng(nd * 10 + eqop)
    As = [lattice.random_su3(n, 50 + d).astype(dt) for d in range(nd)]
    Vs = [lattice.random_field("V", n, 80 + d).astype(dt) for d in range(nd)]
    a_idx = [rng.permutation(n) for _ in range(nd)]
    b_idx = [rng.permutation(n) for _ in range(nd)]
    r0 = lattice.random_field("V", n, 7).astype(dt)
    for pa, pb in ((False, False), (True, True), (False, True)):
        r = r0.copy()
        orc.multi(eqop, r, As, Vs, adj=adj, a_idx=a_idx if pa else None, b_idx=b_idx if pb else None)
        y = np.zeros((n, 3), dtype=L)
        yb = np.zeros((n, 3), dtype=L)
        for d in range(nd):
            Uc = cplx(As[d])[a_idx[d]] if pa else cplx(As[d])
            A = np.conj(np.swapaxes(Uc, 1, 2)) if adj else Uc
            v = cplx(Vs[d])[b_idx[d]] if pb else cplx(Vs[d])
            y += np.einsum("nij,nj->ni", A, v)
            yb += np.einsum("nij,nj->ni", np.abs(A), np.abs(v))
        exact = apply_eqop(eqop, cplx(r0), y)
        bound = 4 * (3 * nd + 2) * u_ * (np.abs(cplx(r0)) + yb)
        assert (np.abs(cplx(r) - exact) <= bound).all()


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("eqop", range(4))
def test_multi_is_chained_matvec(orc, prec, eqop):
    """Term d = 0 with the call's eqop, then peq (eq/peq) or meq (meq/eqm),
    r stored between terms (expressions_tensor.pl:983-1002): bit for bit."""
    n, nd = 150, 5
    dt = np.float32 if prec == "F" else np.float64
    As = [lattice.random_su3(n, 60 + d).astype(dt) for d in range(nd)]
    Vs = [lattice.random_field("V", n, 70 + d).astype(dt) for d in range(nd)]
    r0 = lattice.random_field("V", n, 8).astype(dt)
    idx = np.arange(1, n, 2)
    for adj in (False, True):
        for index in (None, idx):
            got = r0.copy()
            orc.multi(eqop, got, As, Vs, adj=adj, index=index)
            exp = r0.copy()
            later = 2 if eqop in (2, 3) else 1
            for d in range(nd):
                orc.site(0, eqop if d == 0 else later, exp, As[d], Vs[d], adj=adj, index=index)
            np.testing.assert_array_equal(got, exp)
    nothing = r0.copy()
    orc.multi(eqop, nothing, [], [])
    np.testing.assert_array_equal(nothing, r0)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_per_operand_index_forms(orc, prec, kind):
    """QLA_INDEX_ALL forms (indirection_all.pl:66-84): step i uses
    r[indexr[i]], a[indexa[i]], b[indexb[i]]; equal bit for bit to the v-form
    on the gathered operands, scattered back (the reference's own check,
    test/src/protocol_idx.m4:2110-2160, loops single-site calls the same way)."""
    n_sites, n = 90, 57
    dt = np.float32 if prec == "F" else np.float64
    rng = np.random.default_rng(kind)
    U = lattice.random_su3(n_sites, 91).astype(dt)
    b = lattice.random_field(SRC2[kind], n_sites, 92).astype(dt)
    r0 = lattice.random_field(DEST[kind], n_sites, 93).astype(dt)
    ir, ia, ib = (rng.choice(n_sites, n, replace=False) for _ in range(3))
    for eqop in range(4):
        for adj in (False, True):
            for use in ((1, 0, 0), (0, 1, 1), (1, 1, 1)):
                got = r0.copy()
                orc.site(kind, eqop, got, U, b, 3, -1, adj=adj, n=n, ir=ir if use[0] else None,
                         ia=ia if use[1] else None, ib=ib if use[2] else None)
                rs = ir if use[0] else np.arange(n)
                sub = np.ascontiguousarray(r0[rs])
                orc.site(kind, eqop, sub, np.ascontiguousarray(U[ia if use[1] else np.arange(n)]),
                         np.ascontiguousarray(b[ib if use[2] else np.arange(n)]), 3, -1, adj=adj)
                exp = r0.copy()
                exp[rs] = sub
                np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("prec", ["F", "D"])
def test_multi_per_operand_index_forms(orc, prec):
    """V_eq_nxM_times_nV and relatives: one indexr / indexa / indexb shared by
    all terms; bit for bit the v-form on gathered terms, scattered back."""
    n_sites, n, nd = 70, 41, 3
    dt = np.float32 if prec == "F" else np.float64
    rng = np.random.default_rng(5)
    As = [lattice.random_su3(n_sites, 110 + d).astype(dt) for d in range(nd)]
    Vs = [lattice.random_field("V", n_sites, 120 + d).astype(dt) for d in range(nd)]
    r0 = lattice.random_field("V", n_sites, 130).astype(dt)
    ir, ia, ib = (rng.choice(n_sites, n, replace=False) for _ in range(3))
    for eqop in range(4):
        for use in ((1, 0, 0), (0, 1, 0), (1, 1, 1)):
            got = r0.copy()
            orc.multi(eqop, got, As, Vs, n=n, ir=ir if use[0] else None, ia=ia if use[1] else None,
                      ib=ib if use[2] else None)
            rs = ir if use[0] else np.arange(n)
            sub = np.ascontiguousarray(r0[rs])
            orc.multi(eqop, sub, [np.ascontiguousarray(a[ia if use[1] else np.arange(n)]) for a in As],
                      [np.ascontiguousarray(v[ib if use[2] else np.arange(n)]) for v in Vs])
            exp = r0.copy()
            exp[rs] = sub
            np.testing.assert_array_equal(got, exp)


# ---- matrix-free spin operations (kinds 5, 6, 7) ---------------------------
UDEST = {5: "H", 6: "D", 7: "D"}
USRC = {5: "D", 6: "H", 7: "D"}


def exact_unary(kind, b, mu, sign):
    """(y, |y| bound) in long double from the gamma basis."""
    P, R = proj_recon_mats(mu, sign)
    if kind == 5:
        return np.einsum("sk,nck->ncs", P.astype(L), b), np.einsum("sk,nck->ncs", np.abs(P), np.abs(b))
    if kind == 6:
        return np.einsum("ks,ncs->nck", R.astype(L), b), np.einsum("ks,ncs->nck", np.abs(R), np.abs(b))
    g = G[mu] if mu < 4 else G5
    S = (np.eye(4) + sign * g) if mu < 4 else 0.5 * (np.eye(4) + sign * g)
    return np.einsum("kl,ncl->nck", S.astype(L), b), 2 * np.einsum("kl,ncl->nck", np.abs(S), np.abs(b))


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("kind", [5, 6, 7])
@pytest.mark.parametrize("eqop", range(4))
def test_unary_vs_exact(orc, prec, kind, eqop):
    """H eqop P psi, D eqop R h, D eqop (1 + s gamma_mu) psi against the gamma
    basis, all ten (mu, sign), field and pointer sources."""
    n = 160
    dt = np.float32 if prec == "F" else np.float64
    u_ = U32 if prec == "F" else U64
    b = lattice.random_field(USRC[kind], n, 60 + kind).astype(dt)
    r0 = lattice.random_field(UDEST[kind], n, 70 + eqop).astype(dt)
    b_idx = np.random.default_rng(kind + 10 * eqop).permutation(n)
    for ms in range(10):
        mu, sign = ms % 5, (1 if ms < 5 else -1)
        for pb in (False, True):
            r = r0.copy()
            orc.site(kind, eqop, r, None, b, mu, sign, b_idx=b_idx if pb else None)
            y, yb = exact_unary(kind, cplx(b)[b_idx] if pb else cplx(b), mu, sign)
            exact = apply_eqop(eqop, cplx(r0), y)
            err = np.abs(cplx(r) - exact)
            assert (err <= 8 * u_ * (np.abs(cplx(r0)) + yb) + 1e-300).all(), (kind, eqop, mu, sign, pb)


@pytest.mark.parametrize("prec", ["F", "D"])
def test_unary_equals_identity_matrix_forms(orc, prec):
    """With A = 1 the matrix forms reduce to the matrix-free ones bit for bit
    (nonzero data): spproj against H eqop 1*spproj(D), sprecon against
    D eqop sprecon(1*H), and the projection with reconstruction against
    D eqop sprecon(1*spproj(D)).  peq / meq are also r0 +/- the eq result."""
    n = 120
    dt = np.float32 if prec == "F" else np.float64
    eye = np.zeros((n, 3, 3, 2), dtype=dt)
    eye[:, [0, 1, 2], [0, 1, 2], 0] = 1
    for kind, mk in ((5, 1), (6, 2), (7, 3)):
        b = lattice.random_field(USRC[kind], n, 80 + kind).astype(dt)
        r0 = lattice.random_field(UDEST[kind], n, 90 + kind).astype(dt)
        for ms in range(10):
            mu, sign = ms % 5, (1 if ms < 5 else -1)
            eqv = r0.copy()
            orc.site(kind, 0, eqv, None, b, mu, sign)
            for eqop in range(4):
                got = r0.copy()
                orc.site(kind, eqop, got, None, b, mu, sign)
                ref = r0.copy()
                orc.site(mk, eqop, ref, eye, b, mu, sign)
                np.testing.assert_array_equal(got, ref, err_msg=f"kind {kind} eqop {eqop} mu {mu} sign {sign}")
                if kind == 5 and eqop in (1, 2):
                    np.testing.assert_array_equal(got, r0 + eqv if eqop == 1 else r0 - eqv)


@pytest.mark.parametrize("prec", ["F", "D"])
def test_unary_eqm_keeps_macro_zero_sign(orc, prec):
    """eqm is the macro -a.re + b.im (QLA_c_eqm_c_plus_ic), not -(a.re -
    b.im): with a.re == b.im the reference gives +0, a negated eq gives -0."""
    dt = np.float32 if prec == "F" else np.float64
    d = np.zeros((1, 3, 4, 2), dtype=dt)
    d[0, :, 0, 0] = 1.5          # d0.re
    d[0, :, 3, 1] = 1.5          # d3.im -> mu 0, sign +: h0.re = d0.re - d3.im = 0
    d[0, :, 1, 1] = -2.25        # d1.im
    d[0, :, 2, 0] = 2.25         # d2.re -> h1.im = d1.im + d2.re = 0
    for eqop in (0, 3):
        h = np.full((1, 3, 2, 2), 7.0, dtype=dt)
        orc.site(5, eqop, h, None, d, 0, 1)
        assert (h[0, :, 0, 0] == 0).all() and not np.signbit(h[0, :, 0, 0]).any(), eqop
        assert (h[0, :, 1, 1] == 0).all() and not np.signbit(h[0, :, 1, 1]).any(), eqop


# ---- colour-matrix products of H, D, M (kinds 8-11) ------------------------
MT = {8: "H", 9: "D", 10: "M", 11: "M"}


def exact_prod(kind, A, b):
    """(y, bound) in long double: A b for H/D/M records, A b^dagger for kind 11."""
    B = np.conj(np.swapaxes(b, 1, 2)) if kind == 11 else b
    return np.einsum("nik,nk...->ni...", A, B), np.einsum("nik,nk...->ni...", np.abs(A), np.abs(B))


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("kind", [8, 9, 10, 11])
@pytest.mark.parametrize("eqop", range(4))
@pytest.mark.parametrize("adj", [False, True])
def test_matrix_products_vs_exact(orc, prec, kind, eqop, adj):
    n = 120
    dt = np.float32 if prec == "F" else np.float64
    u_ = U32 if prec == "F" else U64
    U = lattice.random_su3(n, 11 + kind).astype(dt)
    b = lattice.random_field(MT[kind], n, 12 + kind).astype(dt)
    r0 = lattice.random_field(MT[kind], n, 13 + eqop).astype(dt)
    idx = np.random.default_rng(kind * 8 + eqop).permutation(n)
    for pb in (False, True):
        r = r0.copy()
        orc.site(kind, eqop, r, U, b, adj=adj, b_idx=idx if pb else None)
        Uc = cplx(U)
        A = np.conj(np.swapaxes(Uc, 1, 2)) if adj else Uc
        y, yb = exact_prod(kind, A, cplx(b)[idx] if pb else cplx(b))
        exact = apply_eqop(eqop, cplx(r0), y)
        err = np.abs(cplx(r) - exact)
        assert (err <= 8 * u_ * (np.abs(cplx(r0)) + yb) + 1e-300).all(), (kind, eqop, adj, pb)


@pytest.mark.parametrize("prec", ["F", "D"])
def test_matrix_products_are_column_matvecs(orc, prec):
    """Each destination column is the matvec of the matching source column,
    bit for bit (the emitter's row-column loop is the same for every type);
    for M b^dagger that column is conj(row jc of b)."""
    n = 100
    dt = np.float32 if prec == "F" else np.float64
    U = lattice.random_su3(n, 5).astype(dt)
    for kind in (8, 9, 10, 11):
        b = lattice.random_field(MT[kind], n, 6 + kind).astype(dt)
        r0 = lattice.random_field(MT[kind], n, 20 + kind).astype(dt)
        ncol = b.shape[2]
        for eqop in range(4):
            for adj in (False, True):
                r = r0.copy()
                orc.site(kind, eqop, r, U, b, adj=adj)
                for c in range(ncol):
                    if kind == 11:
                        col = np.ascontiguousarray(b[:, c, :, :] * np.array([1, -1], dtype=dt))
                    else:
                        col = np.ascontiguousarray(b[:, :, c, :])
                    v = np.ascontiguousarray(r0[:, :, c, :])
                    orc.site(0, eqop, v, U, col, adj=adj)
                    np.testing.assert_array_equal(r[:, :, c, :], v, err_msg=f"kind {kind} eqop {eqop} adj {adj} col {c}")
