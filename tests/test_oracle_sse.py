"""Pins of the oracle's restatement of a QLA built --enable-sse3
(oracle/qla_oracle_sse.inc): QLA_F3_V_vpeq_M_times_pV, V_veq_M_times_pV,
V_veq_Ma_times_V, D_vpeq_sprecon_M_times_pH and D_vpeq_spproj_M_times_pD in
the operation orders of lib/optimized/sse (the SSE sources need the
configure-generated qla_config.h, so the reference build is not available
here).

* Small-integer inputs: every product and partial sum is exact in float, so
  every association order gives the exact result -- the SSE restatement must
  equal the generic oracle bit for bit on all paths and all mu/sign.  A
  dropped term, a wrong sign, a wrong index or a transposed operand fails.
* Known answers derived by hand from the documented orders, on inputs where
  float rounding tells the orders apart (2^25 + 1 rounds to 2^25): the
  aligned SSE multiply, the unaligned scalar fallback, the V 1-site row-2
  expressions, the adjoint orders and the generic order each give a
  different, predicted value.  A restatement in the wrong order fails.
* The V dispatch: which sites take the 1-site body for r 16-byte or only
  8-byte aligned and n odd or even.
* Accuracy: every component within 4 ulp of the cancellation-aware scale
  (|r| + sum |a_k||x_k|) of the exact (float64) result, as is the generic
  order, and the two orders do differ on some components.
"""
import numpy as np
import pytest

from qla_amd import lattice

P25 = float(2 ** 25)


def small_int(shape, rng):
    return rng.integers(-8, 9, size=shape).astype(np.float32)


@pytest.mark.parametrize("kind", ["sprecon", "spproj"])
@pytest.mark.parametrize("al", ["aligned", "unaligned"])
def test_small_integers_equal_generic(orc, kind, al):
    rng = np.random.default_rng(3)
    n = 257
    a = small_int((n, 3, 3, 2), rng)
    src = small_int((n, 3, 2, 2) if kind == "sprecon" else (n, 3, 4, 2), rng)
    r0 = small_int((n, 3, 4, 2), rng)
    idx = rng.permutation(n)
    sym = "D_vpeq_sprecon_M_times_pH" if kind == "sprecon" else "D_vpeq_spproj_M_times_pD"
    k = orc.SSE_SYMBOLS[sym][0]
    r_al = 0 if al == "aligned" else 8
    for mu in range(5):
        for sign in (1, -1):
            exp = r0.copy()
            orc.site(k, orc.PEQ, exp, a, src, mu, sign, b_idx=idx)
            got = orc.sse_site(sym, r0.copy(), a, src, idx, mu, sign, r_al=r_al, a_al=0, b0_al=0)
            np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("sym", ["V_vpeq_M_times_pV", "V_veq_M_times_pV", "V_veq_Ma_times_V"])
@pytest.mark.parametrize("r_al,n", [(0, 64), (0, 65), (8, 64), (8, 65)])
def test_small_integers_equal_generic_v(orc, sym, r_al, n):
    rng = np.random.default_rng(4)
    a = small_int((n, 3, 3, 2), rng)
    v = small_int((n, 3, 2), rng)
    r0 = small_int((n, 3, 2), rng)
    kind, eqop, adj = orc.SSE_SYMBOLS[sym]
    idx = None if adj else rng.permutation(n)
    exp = r0.copy()
    orc.site(orc.MATVEC, eqop, exp, a, v, b_idx=idx, adj=bool(adj), n=n)
    got = orc.sse_site(sym, r0.copy(), a, v, idx, r_al=r_al, a_al=0)
    np.testing.assert_array_equal(got, exp)


def _order_rows():
    """Two colour rows with all x_k = 1 + i: row 0 has (Re, Im) of a_k =
    (2^25, 2^25), (1, 0), (1, 0); row 1 has (1, 0), (2^25, 2^25), (0, 0).
    Re t = sum_k (Re a_k - Im a_k) is 2 and 1 exactly.  Hand evaluation of
    the orders (float: 2^25 + 1 -> 2^25, 2^25 - 2^25 = 0):
      aligned SSE  ((A0 + A1) + A2) - ((B0 + B1) + B2):   row 0: 0   row 1: 0
      scalar       ((((0 + A0) - B0) + A1) - B1) ...:      row 0: 2   row 1: 0
      generic      x += (A_k - B_k), k = 0, 1, 2:          row 0: 2   row 1: 1
    """
    a = np.zeros((3, 3, 2), np.float32)
    a[0, 0] = (P25, P25)
    a[0, 1] = (1, 0)
    a[0, 2] = (1, 0)
    a[1, 0] = (1, 0)
    a[1, 1] = (P25, P25)
    return a


@pytest.mark.parametrize("kind", ["sprecon", "spproj"])
def test_known_answers_half_scalar_generic(orc, kind):
    a = _order_rows()[None]
    if kind == "sprecon":
        src = np.ones((1, 3, 2, 2), np.float32)          # H[k][s] = 1 + i
        sym = "D_vpeq_sprecon_M_times_pH"
    else:
        # spproj mu = 3, sign = +1: h0 = d0 + d2, h1 = d1 + d3 -> 1 + i from d = (1+i)/2
        src = np.full((1, 3, 4, 2), 0.5, np.float32)
        sym = "D_vpeq_spproj_M_times_pD"
    k = orc.SSE_SYMBOLS[sym][0]
    idx = np.array([0])
    mu, sign = 3, 1                                      # r0 += t0, r1 += t1, r2 += t0, r3 += t1
    want = {"aligned": (0.0, 0.0), "unaligned": (2.0, 0.0), "generic": (2.0, 1.0)}
    for how, (w0, w1) in want.items():
        r = np.zeros((1, 3, 4, 2), np.float32)
        if how == "generic":
            orc.site(k, orc.PEQ, r, a, src, mu, sign, b_idx=idx)
        else:
            orc.sse_site(sym, r, a, src, idx, mu, sign, r_al=0 if how == "aligned" else 8, a_al=0, b0_al=0)
        assert (r[0, 0, 0, 0], r[0, 1, 0, 0]) == (w0, w1), how
        assert (r[0, 0, 2, 0], r[0, 1, 2, 0]) == (w0, w1), how
    # b[0] (not r) misaligned also selects the fallback
    r = np.zeros((1, 3, 4, 2), np.float32)
    orc.sse_site(sym, r, a, src, idx, mu, sign, r_al=0, a_al=0, b0_al=8)
    assert (r[0, 0, 0, 0], r[0, 1, 0, 0]) == (2.0, 0.0)


@pytest.mark.parametrize("sym", ["V_vpeq_M_times_pV", "V_veq_M_times_pV"])
def test_known_answers_v_row2(orc, sym):
    """Row 2 of V_vpeq / V_veq: the 1-site body's scalar expression
    ((((A0 - B0) + A1) - B1) + A2) - B2 (added to r = 0 for vpeq) gives 2 on
    row 0 of _order_rows placed in row 2, the 2-site body's vector order
    gives 0; rows 0 and 1 take the vector order in both bodies.  Dispatch: r
    16-byte aligned, n = 5: sites 0-3 in the 2-site body, 4 alone; r 8-byte
    aligned, n = 6: site 0 alone, 1-4 paired, 5 alone."""
    rows = _order_rows()
    a1 = np.zeros((3, 3, 2), np.float32)
    a1[0], a1[1], a1[2] = rows[0], rows[0], rows[0]
    v = np.ones((1, 3, 2), np.float32)                   # b_k = 1 + i
    for r_al, n, single in ((0, 5, {4}), (8, 6, {0, 5}), (0, 6, set()), (8, 5, {0})):
        a = np.repeat(a1[None], n, axis=0)
        r = np.zeros((n, 3, 2), np.float32)
        orc.sse_site(sym, r, a, v, np.zeros(n, dtype=np.int64), r_al=r_al, a_al=0)
        for i in range(n):
            assert r[i, 0, 0] == 0.0 and r[i, 1, 0] == 0.0, (r_al, n, i)
            assert r[i, 2, 0] == (2.0 if i in single else 0.0), (r_al, n, i)
    r = np.zeros((1, 3, 2), np.float32)
    orc.site(orc.MATVEC, orc.SSE_SYMBOLS[sym][1], r, a1[None], v, b_idx=np.array([0]))
    assert r[0, 0, 0] == 2.0                             # generic: x += (A_k - B_k)


def test_known_answers_adjoint(orc):
    """V_veq_Ma_times_V, r[ic] = sum_k conj(a[k][ic]) b[k], with every
    column of a equal to row 0 of _order_rows and b_k = 1 + i:
    Im r = sum_k (Re a_k - Im a_k) = 2 exactly.  By hand:
      adjoint vector  (ar0 + ar1 + ar2) - (ai0 + ai1 + ai2) = 2^25 - 2^25 = 0
      adjoint row 2   ((((ar0 - ai0) + ar1) - ai1) + ar2) - ai2 = 2
      generic         x += (ar_k - ai_k): 2
    (Re r = 2^26 + 2 rounds to 2^26 in all three).  r 8-byte aligned, n = 2:
    site 0 alone (row 2 in the row-2 order), site 1 alone."""
    a1 = np.zeros((3, 3, 2), np.float32)
    col = _order_rows()[0]
    for k in range(3):
        a1[k, :] = col[k]                                # a[k][ic] = col[k] for every ic
    v = np.ones((1, 3, 2), np.float32)
    for r_al, n, single in ((0, 2, set()), (8, 2, {0, 1}), (8, 3, {0})):
        a = np.repeat(a1[None], n, axis=0)
        b = np.repeat(v, n, axis=0)
        r = np.full((n, 3, 2), 7.0, np.float32)
        orc.sse_site("V_veq_Ma_times_V", r, a, b, None, r_al=r_al, a_al=0)
        for i in range(n):
            assert tuple(r[i, 0]) == (2.0 ** 26, 0.0) and tuple(r[i, 1]) == (2.0 ** 26, 0.0), (r_al, n, i)
            assert tuple(r[i, 2]) == (2.0 ** 26, 2.0 if i in single else 0.0), (r_al, n, i)
    r = np.zeros((1, 3, 2), np.float32)
    orc.site(orc.MATVEC, orc.EQ, r, a1[None], v, adj=True, n=1)
    assert tuple(r[0, 0]) == (2.0 ** 26, 2.0)


def _exact(kind, eqop, adj, r0, a, src, idx, mu, sign, orc):
    """float64 evaluation (products exact, sums to ~1e-16) and the
    cancellation-aware scale (|r| for peq) + sum_k |a_k| |x_k| per component."""
    a64 = a.astype(np.float64)
    ac = a64[..., 0] + 1j * a64[..., 1]
    if adj:
        ac = np.conj(np.swapaxes(ac, 1, 2))
    s64 = src.astype(np.float64)[idx]
    sc = s64[..., 0] + 1j * s64[..., 1]
    r64 = r0.astype(np.float64) if eqop == orc.PEQ else np.zeros(r0.shape)
    if kind == orc.MATVEC:
        t = np.einsum("nik,nk->ni", ac, sc)
        mag = np.einsum("nik,nk->ni", np.abs(ac), np.abs(sc))
        out = r64 + np.stack([t.real, t.imag], axis=-1)
        return out, np.abs(r64) + np.stack([mag, mag], axis=-1)
    if kind == orc.WILSON:
        h = np.stack([orc.spproj1(x, mu, sign) for x in src.astype(np.float64)[idx]])
        hc = h[..., 0] + 1j * h[..., 1]
    else:
        hc = sc
    t = np.einsum("nik,nks->nis", ac, hc)
    mag = np.einsum("nik,nks->nis", np.abs(ac), np.abs(hc))
    # the reconstruction is a signed permutation of (re, im) pairs: apply it
    # to t and, for the scale, to (|t| bound, |t| bound)
    tr = np.stack([t.real, t.imag], axis=-1)
    mg = np.stack([mag, mag], axis=-1)
    full = np.stack([orc.sprecon1(x, mu, sign) for x in tr])
    fmag = np.abs(np.stack([orc.sprecon1(x, mu, sign) for x in mg]))
    return r64 + full, np.abs(r64) + fmag


def ulp32(x):
    return np.spacing(np.abs(x).astype(np.float32)).astype(np.float64)


@pytest.mark.parametrize("sym", ["V_vpeq_M_times_pV", "V_veq_M_times_pV", "V_veq_Ma_times_V",
                                 "D_vpeq_sprecon_M_times_pH", "D_vpeq_spproj_M_times_pD"])
def test_within_4ulp_of_exact(orc, sym):
    n = 4096
    dims = (8, 8, 8, 8)
    fw, bw = lattice.neighbours(dims)
    u = lattice.random_su3(n, 1).astype(np.float32)
    kind, eqop, adj = orc.SSE_SYMBOLS[sym]
    worst = 0.0
    differ = 0
    spin = kind != orc.MATVEC
    for mu in range(4 if spin else 1):
        for sign in ((1, -1) if spin else (1,)):
            idx = (fw if sign == 1 else bw)[mu]
            if not spin:
                src = lattice.random_field("V", n, 2).astype(np.float32)
                r0 = lattice.random_field("V", n, 3).astype(np.float32)
            else:
                src = lattice.random_field("H" if kind == orc.SPRECON else "D", n, 2).astype(np.float32)
                r0 = lattice.random_field("D", n, 3).astype(np.float32)
            bidx = None if adj else idx
            gen = r0.copy()
            orc.site(kind, eqop, gen, u, src, mu, sign, b_idx=bidx, adj=bool(adj), n=n)
            for r_al in (0, 8):
                got = orc.sse_site(sym, r0.copy(), u, src, bidx, mu, sign, r_al=r_al, a_al=0, b0_al=0)
                differ += int(np.count_nonzero(got != gen))
                ex, scale = _exact(kind, eqop, adj, r0, u, src, np.arange(n) if adj else idx, mu, sign, orc)
                for res in (got, gen):
                    err = np.abs(res.astype(np.float64) - ex) / ulp32(scale)
                    worst = max(worst, float(err.max()))
                    assert err.max() <= 4.0, (sym, mu, sign, r_al, float(err.max()))
    assert differ > 0, "the SSE and generic orders never differed"
    print(f"{sym}: worst {worst:.2f} ulp of the scale")


def test_bench_loops_equal_oracle(orc):
    """The CPU baseline's loops (mu / sign constant per loop, a as a field)
    give the generic oracle's spproj and the SSE-order sprecon bit for bit,
    every mu / sign, aligned and unaligned destination."""
    import ctypes as C
    lib = orc.load()
    vp, i32 = C.c_void_p, C.c_int
    for f in (lib.orc_F3_bench_spproj, lib.orc_F3_bench_sprecon_sse):
        f.argtypes = [vp, vp, vp, i32, i32, i32]
    n = 1000
    rng = np.random.default_rng(12)
    u = lattice.random_su3(n, 3).astype(np.float32)
    d = lattice.random_field("D", n, 4).astype(np.float32)
    h = lattice.random_field("H", n, 5).astype(np.float32)
    r0 = lattice.random_field("D", n + 1, 6).astype(np.float32)
    idx = rng.permutation(n)
    pd, ph = orc.pointer_array(d, idx), orc.pointer_array(h, idx)     # kept alive across the calls
    for mu in range(5):
        for sign in (1, -1):
            got = np.zeros((n, 3, 2, 2), np.float32)
            lib.orc_F3_bench_spproj(got.ctypes.data, u.ctypes.data, pd.ctypes.data, mu, sign, n)
            exp = np.zeros_like(got)
            orc.H_eqop_spproj_M_times_pD(orc.EQ, exp, u, d, idx, mu, sign)
            np.testing.assert_array_equal(got, exp)
            for off in (0, 1):       # a D record view 0 or 8 bytes past the array start
                flat = r0.reshape(-1).copy()
                view = flat[2 * off:2 * off + n * 24].reshape(n, 3, 4, 2)
                exp = np.ascontiguousarray(view).copy()
                orc.sse_site("D_vpeq_sprecon_M_times_pH", exp, u, h, idx, mu, sign,
                             r_al=view.ctypes.data % 16, a_al=u.ctypes.data % 16, b0_al=h[idx[0]].ctypes.data % 16)
                lib.orc_F3_bench_sprecon_sse(view.ctypes.data, u.ctypes.data, ph.ctypes.data, mu, sign, n)
                np.testing.assert_array_equal(view, exp)


# ---------------------------------------------------------------------------
# The D3 matvecs an --enable-sse3 build takes from its SSE2 versions
# (qla_oracle_sse.inc, orc_D3_sse2_site).
# ---------------------------------------------------------------------------
P53 = float(2 ** 53)
D3_SYMS = ["V_veq_M_times_V", "V_veq_M_times_pV", "V_vpeq_M_times_V", "V_vpeq_M_times_pV"]


@pytest.mark.parametrize("sym", D3_SYMS)
def test_d3_small_integers_equal_generic(orc, sym):
    """Small integers keep every product and partial sum exact in double, so
    the SSE2 order must equal the generic oracle bit for bit."""
    rng = np.random.default_rng(21)
    n = 301
    a = rng.integers(-8, 9, size=(n, 3, 3, 2)).astype(np.float64)
    v = rng.integers(-8, 9, size=(n, 3, 2)).astype(np.float64)
    r0 = rng.integers(-8, 9, size=(n, 3, 2)).astype(np.float64)
    eqop = orc.SSE2_D3_SYMBOLS[sym]
    idx = rng.permutation(n) if sym.endswith("pV") else None
    exp = r0.copy()
    orc.V_eqop_M_times_V(eqop, exp, a, v, b_idx=idx, n=n)
    got = orc.sse2_d3_site(sym, r0.copy(), a, v, idx)
    np.testing.assert_array_equal(got, exp)


def _d3_order_sites():
    """Sites whose results tell the orders apart (double: 2^53 + 1 -> 2^53).
    Each entry: (a, b, r0, row, lane, SSE2 value, generic value); b = 1 + i
    or 1, hand-evaluated:
      A  row 0, Re a = (2^53, 1, 1), Im a = (2^53, 0, 0), b = 1+i, re lane:
         SSE2 ((2^53 + 1) + 1) - 2^53 - 0 - 0 = 0;   generic sum_k (Re - Im) = 2
      B  row 1, Re a = (2^53, 1, -2^53), b = 1+i, re lane: SSE2 columns 0, 2, 1:
         (2^53 - 2^53) + 1 = 1;   generic (2^53 + 1) - 2^53 = 0
      C  row 1, Re a = (2^53, 0, 0), Im a = (1, -2^53, 0), b = 1+i, im lane:
         SSE2 Re a * Im b first: 2^53, then Im columns 1, 0, 2:
         (2^53 - 2^53) + 1 + 0 = 1;   generic (2^53 + 1) + (0 - 2^53) = 0
      D  row 2, Re a = (2^53, 0, 0), Im a = (1, 0, -2^53), b = 1+i, im lane:
         SSE2 2^53, then Im columns 2, 0, 1: (2^53 - 2^53) + 1 + 0 = 1;  generic 0
      E  vpeq, row 0, Re a = (2^53, -2^53, 0), b = 1, r = 1, re lane:
         SSE2 (2^53 - 2^53 + 0) + r = 1 (r added last);  generic r first: 0
    """
    sites = []
    one_i = np.array([[1.0, 1.0]] * 3)
    one = np.array([[1.0, 0.0]] * 3)
    a = np.zeros((3, 3, 2))
    a[0, :, 0] = (P53, 1.0, 1.0)
    a[0, 0, 1] = P53
    sites.append((a, one_i, np.zeros((3, 2)), 0, 0, 0.0, 2.0, "EQ"))
    a = np.zeros((3, 3, 2))
    a[1, :, 0] = (P53, 1.0, -P53)
    sites.append((a, one_i, np.zeros((3, 2)), 1, 0, 1.0, 0.0, "EQ"))
    a = np.zeros((3, 3, 2))
    a[1, :, 0] = (P53, 0.0, 0.0)
    a[1, :, 1] = (1.0, -P53, 0.0)
    sites.append((a, one_i, np.zeros((3, 2)), 1, 1, 1.0, 0.0, "EQ"))
    a = np.zeros((3, 3, 2))
    a[2, :, 0] = (P53, 0.0, 0.0)
    a[2, :, 1] = (1.0, 0.0, -P53)
    sites.append((a, one_i, np.zeros((3, 2)), 2, 1, 1.0, 0.0, "EQ"))
    a = np.zeros((3, 3, 2))
    a[0, :, 0] = (P53, -P53, 0.0)
    r0 = np.zeros((3, 2))
    r0[0, 0] = 1.0
    sites.append((a, one, r0, 0, 0, 1.0, 0.0, "PEQ"))
    return sites


def test_d3_known_answers(orc):
    for k, (a, b, r0, row, lane, sse2, gen, eq) in enumerate(_d3_order_sites()):
        sym = "V_veq_M_times_V" if eq == "EQ" else "V_vpeq_M_times_V"
        eqop = orc.EQ if eq == "EQ" else orc.PEQ
        got = orc.sse2_d3_site(sym, r0[None].copy(), a[None], b[None])
        assert got[0, row, lane] == sse2, (k, got[0, row, lane], sse2)
        exp = r0[None].copy()
        orc.V_eqop_M_times_V(eqop, exp, a[None], b[None])
        assert exp[0, row, lane] == gen, (k, exp[0, row, lane], gen)


@pytest.mark.parametrize("sym", D3_SYMS)
def test_d3_within_4ulp_of_exact(orc, sym):
    """Every component within 4 ulp (double) of the long-double result on
    the scale |r| + sum_k |a_k||x_k|, as is the generic order; the two
    orders do differ on some components."""
    n = 4096
    rng = np.random.default_rng(33)
    u = lattice.random_su3(n, 5)
    v = lattice.random_field("V", n, 6)
    r0 = lattice.random_field("V", n, 7)
    eqop = orc.SSE2_D3_SYMBOLS[sym]
    idx = rng.permutation(n) if sym.endswith("pV") else None
    got = orc.sse2_d3_site(sym, r0.copy(), u, v, idx)
    gen = r0.copy()
    orc.V_eqop_M_times_V(eqop, gen, u, v, b_idx=idx, n=n)
    src = v if idx is None else v[idx]
    A = u[..., 0].astype(np.longdouble) + 1j * u[..., 1].astype(np.longdouble)
    X = src[..., 0].astype(np.longdouble) + 1j * src[..., 1].astype(np.longdouble)
    prod = np.einsum("nij,nj->ni", A, X)
    mag = np.einsum("nij,nj->ni", np.abs(A), np.abs(X))
    if eqop == orc.PEQ:
        prod = prod + (r0[..., 0] + 1j * r0[..., 1])
        mag = mag + np.abs(r0[..., 0] + 1j * r0[..., 1])
    ex = np.stack([prod.real, prod.imag], axis=-1)
    scale = np.stack([mag, mag], axis=-1).astype(np.float64)
    ulp = np.spacing(scale)
    for res in (got, gen):
        err = (np.abs(res.astype(np.longdouble) - ex) / ulp).astype(np.float64)
        assert err.max() <= 4.0, (sym, float(err.max()))
    assert np.count_nonzero(got != gen) > 0
