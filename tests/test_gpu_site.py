"""GPU parity of the site kernels against the oracle, through the C ABI.

Bar: bit-exact (F3 and D3).  The HIP kernels evaluate the reference macros
in the reference order without contraction, exactly as the oracle does, so
every component must agree bit for bit, -0.0 and all.
Cases follow the reference's own idx tests (test/src/protocol_idx.m4:2087-2215:
v- and x-variants, mu 0..4, sign +-1) plus ragged tile tails, random gather
targets, mu out of range (no-op), sign values other than +-1 (treated as -1)
and n = 0.
"""
import numpy as np
import pytest

import qla_amd
from qla_amd import lattice
from gpu_util import assert_bitwise, dev, host, inputs, say

pytestmark = pytest.mark.gpu

PRECS = ["F", "D"]
SMALL_DIMS = [(4, 4, 4, 4), (6, 4, 2, 3), (1, 1, 1, 1), (3, 5, 7, 1)]


def gather_targets(dims, mu, sign, kind, seed=5):
    n = lattice.volume(dims)
    if kind == "nbr":
        fw, bw = lattice.neighbours(dims)
        return (fw if sign == 1 else bw)[mu % 4]
    return np.random.default_rng(seed + mu).permutation(n)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("dims", SMALL_DIMS)
@pytest.mark.parametrize("kind", ["nbr", "perm"])
def test_spproj_v(orc, prec, dims, kind):
    u, d, _, _ = inputs(dims, prec)
    n = u.shape[0]
    du, dd = dev(u), dev(d)
    name = f"QLA_{prec}3_H_veq_spproj_M_times_pD"
    for mu in range(5):
        for sign in (1, -1, 0):
            say(f"{name} dims={dims} kind={kind} mu={mu} sign={sign}")
            idx = gather_targets(dims, mu, sign, kind)
            r = dev(np.zeros((n, 3, 2, 2), dtype=u.dtype))
            qla_amd.call(name, r, du, qla_amd.pointer_array(dd, idx), mu, sign, n)
            qla_amd.check()
            exp = np.zeros((n, 3, 2, 2), dtype=u.dtype)
            orc.H_eqop_spproj_M_times_pD(orc.EQ, exp, u, d, idx, mu, sign)
            assert_bitwise(host(r), exp, f"{name} mu={mu} sign={sign}")


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (6, 4, 2, 3)])
@pytest.mark.parametrize("parity", [0, 1])
def test_spproj_x(orc, prec, dims, parity):
    u, d, h0, _ = inputs(dims, prec)
    n = u.shape[0]
    du, dd = dev(u), dev(d)
    index = lattice.parity_sites(dims, parity)
    name = f"QLA_{prec}3_H_xeq_spproj_M_times_pD"
    for mu in range(5):
        for sign in (1, -1):
            say(f"{name} dims={dims} parity={parity} mu={mu} sign={sign}")
            idx = gather_targets(dims, mu, sign, "nbr")
            r = dev(h0)     # sites outside the subset must stay untouched
            qla_amd.call(name, r, du, qla_amd.pointer_array(dd, idx), mu, sign, dev(index), len(index))
            qla_amd.check()
            exp = h0.copy()
            orc.H_eqop_spproj_M_times_pD(orc.EQ, exp, u, d, idx, mu, sign, index=index)
            assert_bitwise(host(r), exp, f"{name} mu={mu} sign={sign}")


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("dims", SMALL_DIMS)
@pytest.mark.parametrize("kind", ["nbr", "perm"])
def test_sprecon_v(orc, prec, dims, kind):
    u, d0, h, _ = inputs(dims, prec)
    n = u.shape[0]
    du, dh = dev(u), dev(h)
    name = f"QLA_{prec}3_D_vpeq_sprecon_M_times_pH"
    for mu in range(5):
        for sign in (1, -1, 2):
            say(f"{name} dims={dims} kind={kind} mu={mu} sign={sign}")
            idx = gather_targets(dims, mu, sign, kind)
            r = dev(d0)
            qla_amd.call(name, r, du, qla_amd.pointer_array(dh, idx), mu, sign, n)
            qla_amd.check()
            exp = d0.copy()
            orc.D_eqop_sprecon_M_times_pH(orc.PEQ, exp, u, h, idx, mu, sign)
            assert_bitwise(host(r), exp, f"{name} mu={mu} sign={sign}")


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (6, 4, 2, 3)])
@pytest.mark.parametrize("parity", [0, 1])
def test_sprecon_x(orc, prec, dims, parity):
    u, d0, h, _ = inputs(dims, prec)
    du, dh = dev(u), dev(h)
    index = lattice.parity_sites(dims, parity)
    name = f"QLA_{prec}3_D_xpeq_sprecon_M_times_pH"
    for mu in range(5):
        for sign in (1, -1):
            say(f"{name} dims={dims} parity={parity} mu={mu} sign={sign}")
            idx = gather_targets(dims, mu, sign, "nbr")
            r = dev(d0)
            qla_amd.call(name, r, du, qla_amd.pointer_array(dh, idx), mu, sign, dev(index), len(index))
            qla_amd.check()
            exp = d0.copy()
            orc.D_eqop_sprecon_M_times_pH(orc.PEQ, exp, u, h, idx, mu, sign, index=index)
            assert_bitwise(host(r), exp, f"{name} mu={mu} sign={sign}")


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("dims", SMALL_DIMS)
def test_matvec(orc, prec, dims):
    u, _, _, v = inputs(dims, prec)
    n = u.shape[0]
    du, dv = dev(u), dev(v)
    name = f"QLA_{prec}3_V_veq_M_times_V"
    say(f"{name} dims={dims}")
    r = dev(np.zeros_like(v))
    qla_amd.call(name, r, du, dv, n)
    qla_amd.check()
    exp = np.zeros_like(v)
    orc.V_eqop_M_times_V(orc.EQ, exp, u, v)
    assert_bitwise(host(r), exp, name)

    name = f"QLA_{prec}3_V_vpeq_M_times_pV"
    for kind in ("nbr", "perm"):
        say(f"{name} dims={dims} kind={kind}")
        idx = gather_targets(dims, 2, 1, kind)
        r0 = lattice.random_field("V", n, 99).astype(u.dtype)
        r = dev(r0)
        qla_amd.call(name, r, du, qla_amd.pointer_array(dv, idx), n)
        qla_amd.check()
        exp = r0.copy()
        orc.V_eqop_M_times_V(orc.PEQ, exp, u, v, b_idx=idx)
        assert_bitwise(host(r), exp, name)


@pytest.mark.parametrize("prec", PRECS)
def test_noop_cases(prec):
    """mu outside 0..4 and n <= 0 leave the destination untouched."""
    dims = (4, 4, 2, 2)
    u, d, h, _ = inputs(dims, prec)
    n = u.shape[0]
    du, dd, dh = dev(u), dev(d), dev(h)
    idx = np.arange(n)
    for mu in (-1, 5, 17):
        say(f"no-op mu={mu} prec={prec}")
        r = dev(h)
        qla_amd.call(f"QLA_{prec}3_H_veq_spproj_M_times_pD", r, du, qla_amd.pointer_array(dd, idx), mu, 1, n)
        assert_bitwise(host(r), h, "spproj mu out of range")
        r = dev(d)
        qla_amd.call(f"QLA_{prec}3_D_vpeq_sprecon_M_times_pH", r, du, qla_amd.pointer_array(dh, idx), mu, -1, n)
        assert_bitwise(host(r), d, "sprecon mu out of range")
    say(f"no-op n=0 prec={prec}")
    r = dev(h)
    qla_amd.call(f"QLA_{prec}3_H_veq_spproj_M_times_pD", r, du, qla_amd.pointer_array(dd, idx), 0, 1, 0)
    assert_bitwise(host(r), h, "spproj n=0")
    qla_amd.check()


@pytest.mark.parametrize("prec", PRECS)
def test_full_lattice_32(orc, prec):
    """Config-1 shape: 32^4 periodic lattice, all 8 (mu, sign) of both ops,
    plus the x-forms on the even checkerboard, bit-exact against the oracle."""
    dims = (32, 32, 32, 32)
    n = lattice.volume(dims)
    dt = np.float32 if prec == "F" else np.float64
    u = lattice.random_su3(n, 1).astype(dt)
    d = lattice.random_field("D", n, 2).astype(dt)
    h = lattice.random_field("H", n, 3).astype(dt)
    fw, bw = lattice.neighbours(dims)
    even = lattice.parity_sites(dims, 0)
    du, dd, dh, deven = dev(u), dev(d), dev(h), dev(even)
    for mu in range(4):
        for sign in (1, -1):
            idx = (fw if sign == 1 else bw)[mu]
            say(f"32^4 {prec}3 spproj mu={mu} sign={sign}")
            r = dev(np.zeros((n, 3, 2, 2), dtype=dt))
            qla_amd.call(f"QLA_{prec}3_H_veq_spproj_M_times_pD", r, du, qla_amd.pointer_array(dd, idx), mu, sign, n)
            qla_amd.check()
            exp = np.zeros((n, 3, 2, 2), dtype=dt)
            orc.H_eqop_spproj_M_times_pD(orc.EQ, exp, u, d, idx, mu, sign)
            assert_bitwise(host(r), exp, f"32^4 spproj mu={mu} sign={sign}")
            say(f"32^4 {prec}3 sprecon mu={mu} sign={sign}")
            r = dev(d)
            qla_amd.call(f"QLA_{prec}3_D_vpeq_sprecon_M_times_pH", r, du, qla_amd.pointer_array(dh, idx), mu, sign, n)
            qla_amd.check()
            exp = d.copy()
            orc.D_eqop_sprecon_M_times_pH(orc.PEQ, exp, u, h, idx, mu, sign)
            assert_bitwise(host(r), exp, f"32^4 sprecon mu={mu} sign={sign}")
    mu, sign = 1, -1
    idx = bw[mu]
    say(f"32^4 {prec}3 x-forms even mu={mu} sign={sign}")
    r = dev(np.zeros((n, 3, 2, 2), dtype=dt))
    qla_amd.call(f"QLA_{prec}3_H_xeq_spproj_M_times_pD", r, du, qla_amd.pointer_array(dd, idx), mu, sign, deven,
                 len(even))
    exp = np.zeros((n, 3, 2, 2), dtype=dt)
    orc.H_eqop_spproj_M_times_pD(orc.EQ, exp, u, d, idx, mu, sign, index=even)
    assert_bitwise(host(r), exp, "32^4 x spproj")
    r = dev(d)
    qla_amd.call(f"QLA_{prec}3_D_xpeq_sprecon_M_times_pH", r, du, qla_amd.pointer_array(dh, idx), mu, sign, deven,
                 len(even))
    exp = d.copy()
    orc.D_eqop_sprecon_M_times_pH(orc.PEQ, exp, u, h, idx, mu, sign, index=even)
    assert_bitwise(host(r), exp, "32^4 x sprecon")
    qla_amd.check()


@pytest.mark.parametrize("prec", ["F"])
@pytest.mark.parametrize("shift", [8, 16, 24])
def test_misaligned_fields(orc, prec, shift):
    """QLA only guarantees 8-byte alignment of float records (qla_complex.h:30).
    Fields placed 8/16/24 bytes past a 16-byte boundary exercise the 4-byte
    chunk fallbacks of the contiguous and the gathered staging paths."""
    dims = (6, 4, 4, 5)
    u, d, h, v = inputs(dims, prec)
    n = u.shape[0]
    fw, bw = lattice.neighbours(dims)
    fl = shift // 4

    def shifted(a):
        flat = np.zeros(a.size + 16, dtype=a.dtype)
        t = dev(flat)
        view = t[fl:fl + a.size].view(a.shape)
        view.copy_(dev(a))
        return t, view

    _, du = shifted(u)
    _, dd = shifted(d)
    _, dh = shifted(h)
    for mu, sign in ((0, 1), (3, -1), (4, 1)):
        idx = (fw if sign == 1 else bw)[mu % 4]
        say(f"misaligned shift={shift} spproj mu={mu} sign={sign}")
        _, r = shifted(np.zeros((n, 3, 2, 2), dtype=u.dtype))
        qla_amd.call("QLA_F3_H_veq_spproj_M_times_pD", r, du, qla_amd.pointer_array(dd, idx), mu, sign, n)
        qla_amd.check()
        exp = np.zeros((n, 3, 2, 2), dtype=u.dtype)
        orc.H_eqop_spproj_M_times_pD(orc.EQ, exp, u, d, idx, mu, sign)
        assert_bitwise(host(r), exp, f"misaligned spproj shift={shift}")
        say(f"misaligned shift={shift} sprecon mu={mu} sign={sign}")
        _, r = shifted(d)
        qla_amd.call("QLA_F3_D_vpeq_sprecon_M_times_pH", r, du, qla_amd.pointer_array(dh, idx), mu, sign, n)
        qla_amd.check()
        exp = d.copy()
        orc.D_eqop_sprecon_M_times_pH(orc.PEQ, exp, u, h, idx, mu, sign)
        assert_bitwise(host(r), exp, f"misaligned sprecon shift={shift}")
    say(f"misaligned shift={shift} matvec")
    _, dv = shifted(v)
    _, r = shifted(np.zeros_like(v))
    qla_amd.call("QLA_F3_V_veq_M_times_V", r, du, dv, n)
    exp = np.zeros_like(v)
    orc.V_eqop_M_times_V(orc.EQ, exp, u, v)
    assert_bitwise(host(r), exp, "misaligned matvec")
    idx = np.random.default_rng(3).permutation(n)
    _, r = shifted(v)
    qla_amd.call("QLA_F3_V_vpeq_M_times_pV", r, du, qla_amd.pointer_array(dv, idx), n)
    exp = v.copy()
    orc.V_eqop_M_times_V(orc.PEQ, exp, u, v, b_idx=idx)
    assert_bitwise(host(r), exp, "misaligned matvec pV")
    qla_amd.check()
