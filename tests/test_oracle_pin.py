"""Pins the oracle (the CPU restatement) against facts fixed independently of it.

The reference's hot-path C exists only as Perl-generated code, so it cannot
be compiled here; these checks are what anchors the oracle instead:
  * spin projection / reconstruction against the gamma basis printed in the
    reference documentation (doc/qla.tex:1063-1113, typed in gamma_basis.py):
    recon(proj(psi)) = (1 + s gamma_mu) psi for mu = 0..3 and
    (1 + s gamma_5)/2 psi for mu = 4, with gamma_5 = g1 g2 g3 g4;
  * the fused multiply ops against a long-double evaluation built from the
    gamma matrices and numpy, within a rounding bound per component that a
    dropped term, a sign error or a transposed operand exceeds by orders of
    magnitude;
  * SU(3) invariants (unitarity preserves the norm; U^dagger U v = v);
  * reductions against long-double sums (the reference tests' own method,
    test/src/protocol_idx.m4:128-183).
The exact float rounding order is the one read from the emitters
(expressions_tensor.pl:544-630); no executable reference pins it.
"""
import numpy as np
import pytest

from gamma_basis import G, G5
from qla_amd import lattice

U32 = np.finfo(np.float32).eps / 2
U64 = np.finfo(np.float64).eps / 2


def cplx(a):
    return a[..., 0].astype(np.longdouble) + 1j * a[..., 1].astype(np.longdouble)


def proj_recon_mats(mu, sign):
    """(P, R): 2x4 projection rows and 4x2 reconstruction, from the gamma basis."""
    s = 1 if sign == 1 else -1
    if mu < 4:
        one_plus = np.eye(4) + s * G[mu]
        P = one_plus[0:2, :]
        # (1 + s g) psi = [h ; s X^dagger h] with X^dagger = g[2:4, 0:2] (chiral basis)
        R = np.vstack([np.eye(2), s * G[mu][2:4, 0:2]])
    elif s == 1:
        P = np.eye(4)[0:2, :]
        R = np.vstack([np.eye(2), np.zeros((2, 2))])
    else:
        P = np.eye(4)[2:4, :]
        R = np.vstack([np.zeros((2, 2)), np.eye(2)])
    return P, R


def test_gamma_basis_facts():
    for mu in range(4):
        np.testing.assert_allclose(G[mu] @ G[mu], np.eye(4), atol=0)
        np.testing.assert_allclose(G[mu].conj().T, G[mu], atol=0)
        for nu in range(mu + 1, 4):
            np.testing.assert_allclose(G[mu] @ G[nu] + G[nu] @ G[mu], 0, atol=0)
    np.testing.assert_allclose(G5, np.diag([1, 1, -1, -1]), atol=0)


@pytest.mark.parametrize("mu", range(5))
@pytest.mark.parametrize("sign", [1, -1])
def test_spin_tables(orc, mu, sign):
    rng = np.random.default_rng(7 + mu)
    for _ in range(4):
        d = rng.standard_normal((3, 4, 2))
        psi = d[..., 0] + 1j * d[..., 1]
        h = orc.spproj1(d, mu, sign)
        r = orc.sprecon1(h, mu, sign)
        g = G[mu] if mu < 4 else G5
        proj = (np.eye(4) + sign * g) if mu < 4 else 0.5 * (np.eye(4) + sign * g)
        np.testing.assert_array_equal(r[..., 0] + 1j * r[..., 1], psi @ proj.T)
        P, R = proj_recon_mats(mu, sign)
        np.testing.assert_array_equal(h[..., 0] + 1j * h[..., 1], psi @ P.T)
        np.testing.assert_array_equal(r[..., 0] + 1j * r[..., 1], (h[..., 0] + 1j * h[..., 1]) @ R.T)


def test_spin_bad_mu_is_noop(orc):
    d = np.ones((3, 4, 2))
    assert not orc.spproj1(d, 5, 1).any()
    assert not orc.sprecon1(np.ones((3, 2, 2)), -1, 1).any()


@pytest.mark.parametrize("prec", ["F", "D"])
def test_matvec_vs_exact(orc, prec):
    n = 500
    dt = np.float32 if prec == "F" else np.float64
    u_ = U32 if prec == "F" else U64
    U = lattice.random_su3(n, 3).astype(dt)
    v = lattice.random_field("V", n, 4).astype(dt)
    r = np.zeros_like(v)
    orc.V_eqop_M_times_V(orc.EQ, r, U, v)
    Uc, vc = cplx(U), cplx(v)
    exact = np.einsum("nij,nj->ni", Uc, vc)
    bound = 8 * u_ * np.einsum("nij,nj->ni", np.abs(Uc), np.abs(vc))
    err = np.abs(cplx(r) - exact)
    assert (err <= bound).all(), err.max()
    # unitary: |U v| = |v|
    nr = np.linalg.norm(cplx(r).astype(complex), axis=1)
    nv = np.linalg.norm(vc.astype(complex), axis=1)
    assert np.allclose(nr, nv, rtol=20 * u_)
    # peq through a pointer gather, and the adjoint: U^dagger (U v) = v
    idx = np.random.default_rng(1).permutation(n)
    r0 = lattice.random_field("V", n, 9).astype(dt)
    r2 = r0.copy()
    orc.V_eqop_M_times_V(orc.PEQ, r2, U, v, b_idx=idx)
    exact2 = cplx(r0) + np.einsum("nij,nj->ni", Uc, vc[idx])
    bound2 = 8 * u_ * (np.abs(cplx(r0)) + np.einsum("nij,nj->ni", np.abs(Uc), np.abs(vc[idx])))
    assert (np.abs(cplx(r2) - exact2) <= bound2).all()
    back = np.zeros_like(v)
    orc.V_eqop_M_times_V(orc.EQ, back, U, r, adj=True)
    assert np.allclose(cplx(back).astype(complex), vc.astype(complex), atol=50 * u_ * 3)


@pytest.mark.parametrize("prec", ["F", "D"])
@pytest.mark.parametrize("mu", range(5))
@pytest.mark.parametrize("sign", [1, -1])
def test_spproj_sprecon_vs_exact(orc, prec, mu, sign):
    n = 300
    dt = np.float32 if prec == "F" else np.float64
    u_ = U32 if prec == "F" else U64
    U = lattice.random_su3(n, 13).astype(dt)
    d = lattice.random_field("D", n, 14).astype(dt)
    h = lattice.random_field("H", n, 15).astype(dt)
    idx = np.random.default_rng(mu).permutation(n)
    P, R = proj_recon_mats(mu, sign)
    Uc = cplx(U)
    # spproj: r = U (P psi)
    r = np.zeros((n, 3, 2, 2), dtype=dt)
    orc.H_eqop_spproj_M_times_pD(orc.EQ, r, U, d, idx, mu, sign)
    psi = cplx(d)[idx]                                       # (n,3,4)
    t = np.einsum("sk,nck->ncs", P.astype(np.clongdouble), psi)
    exact = np.einsum("nab,nbs->nas", Uc, t)
    bound = 8 * u_ * np.einsum("nab,nbs->nas", np.abs(Uc), np.einsum("sk,nck->ncs", np.abs(P), np.abs(psi)))
    assert (np.abs(cplx(r) - exact) <= bound + 1e-300).all()
    # sprecon: r = d0 + R (U h)
    d0 = lattice.random_field("D", n, 16).astype(dt)
    r = d0.copy()
    orc.D_eqop_sprecon_M_times_pH(orc.PEQ, r, U, h, idx, mu, sign)
    hc = cplx(h)[idx]
    w = np.einsum("nab,nbs->nas", Uc, hc)
    exact = cplx(d0) + np.einsum("ks,nas->nak", R.astype(np.clongdouble), w)
    wb = np.einsum("nab,nbs->nas", np.abs(Uc), np.abs(hc))
    bound = 8 * u_ * (np.abs(cplx(d0)) + np.einsum("ks,nas->nak", np.abs(R), wb))
    assert (np.abs(cplx(r) - exact) <= bound + 1e-300).all()


@pytest.mark.parametrize("prec", ["F", "D"])
def test_x_form_touches_only_subset(orc, prec):
    dims = (4, 4, 2, 2)
    n = lattice.volume(dims)
    dt = np.float32 if prec == "F" else np.float64
    U = lattice.random_su3(n, 3).astype(dt)
    d = lattice.random_field("D", n, 4).astype(dt)
    h0 = lattice.random_field("H", n, 5).astype(dt)
    even = lattice.parity_sites(dims, 0)
    odd = lattice.parity_sites(dims, 1)
    fw, _ = lattice.neighbours(dims)
    r = h0.copy()
    orc.H_eqop_spproj_M_times_pD(orc.EQ, r, U, d, fw[0], 0, 1, index=even)
    np.testing.assert_array_equal(r[odd], h0[odd])
    full = np.zeros_like(h0)
    orc.H_eqop_spproj_M_times_pD(orc.EQ, full, U, d, fw[0], 0, 1)
    np.testing.assert_array_equal(r[even], full[even])


@pytest.mark.parametrize("prec", ["F", "D"])
def test_reductions_vs_long_double(orc, prec):
    n = 20000
    dt = np.float32 if prec == "F" else np.float64
    for kind in ("V", "D"):
        a = lattice.random_field(kind, n, 5).astype(dt)
        b = lattice.random_field(kind, n, 6).astype(dt)
        ref = orc.norm2_ld(a)
        tol = 1e-7 if prec == "F" else 1e-13
        assert abs(orc.norm2(a) - float(ref)) <= tol * float(ref)
        re, im = orc.dot_ld(a, b)
        got = orc.dot(a, b)
        scale = float(np.sqrt(ref * orc.norm2_ld(b)))
        assert abs(got.real - float(re)) <= 1e-13 * scale
        assert abs(got.imag - float(im)) <= 1e-13 * scale
        # exact dot of the promoted values, independent of the oracle's loop
        ac, bc = cplx(a).reshape(-1), cplx(b).reshape(-1)
        ex = np.sum(np.conj(ac) * bc)
        assert abs(got - complex(ex)) <= 1e-12 * scale


@pytest.mark.parametrize("prec", ["F", "D"])
def test_term_sums_vs_exact(orc, prec):
    """The long-double sums of the reference's rounded terms (the GPU
    reductions' bar) against exact sums (math.fsum) of the same terms formed
    independently in numpy: the component norm re*re + im*im rounded to the
    field precision (two numpy operations of that precision), the dot terms
    ar*br + ai*bi and ar*bi - ai*br in float64.  A dropped term, a missing
    rounding step or a wrong sign fails this far above the long-double
    error (~1e-19 relative)."""
    import math
    n = 5000
    dt = np.float32 if prec == "F" else np.float64
    a = lattice.random_field("D", n, 15).astype(dt)
    b = lattice.random_field("D", n, 16).astype(dt)
    ar, ai = a[..., 0].reshape(-1), a[..., 1].reshape(-1)
    br, bi = b[..., 0].reshape(-1), b[..., 1].reshape(-1)
    terms = ar * ar + ai * ai                     # in the field precision
    ex = math.fsum(terms.astype(np.float64))
    got = orc.norm2_terms_ld(a)
    assert abs(float(got) - ex) <= 1e-15 * ex
    # the rounding to the field precision is part of the semantics
    if prec == "F":
        exact_sq = math.fsum((ar.astype(np.float64) ** 2 + ai.astype(np.float64) ** 2))
        assert abs(float(got) - exact_sq) > 1e-12 * ex
    a64 = [x.astype(np.float64) for x in (ar, ai, br, bi)]
    tr = a64[0] * a64[2] + a64[1] * a64[3]
    ti = a64[0] * a64[3] - a64[1] * a64[2]
    re, im, sa = orc.dot_terms_ld(a, b)
    assert abs(float(re) - math.fsum(tr)) <= 1e-15 * float(sa)
    assert abs(float(im) - math.fsum(ti)) <= 1e-15 * float(sa)
    assert abs(float(sa) - math.fsum(np.abs(tr)) - math.fsum(np.abs(ti))) <= 1e-15 * float(sa)
