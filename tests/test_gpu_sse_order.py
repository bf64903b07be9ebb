"""Parity with a QLA built --enable-sse3 (the build the CPU baseline
mirrors).  That build replaces five QLA_F3 symbols with re-associated SSE
code (lib/optimized/sse/src): V_vpeq_M_times_pV, V_veq_M_times_pV,
V_veq_Ma_times_V, D_vpeq_sprecon_M_times_pH and D_vpeq_spproj_M_times_pD.

* QLA_MI355X_set_arith_order(QLA_MI355X_ORDER_SSE3): the GPU result equals
  the oracle's restatement of the SSE build (oracle/qla_oracle_sse.inc) bit
  for bit on the 32^4 lattice with neighbour pointer arrays, every mu/sign,
  with 16-byte aligned operands (SSE multiply, 2-site V loop) and with r or
  b[0] 8 bytes off (scalar fallback, 1-site V head).  The oracle is told the
  device call's address offsets, since the reference picks its order from
  them.
* The four QLA_D3 matvecs that --enable-sse3 takes from its SSE2 versions
  (lib/optimized/sse2/src: V_veq_M_times_V, V_veq_M_times_pV,
  V_vpeq_M_times_V, V_vpeq_M_times_pV) likewise in SSE3 mode, bit-exact
  against the oracle's restatement on the 32^4 lattice, field and pointer
  sources; their default order within 4 ulp (double) of it.
* Default order: bit-exact with the generic oracle (the other tests), and
  against the SSE build within 4 ulp (the F3 parity bar), measured on the
  cancellation-aware scale |r| + sum_k |a_k| |x_k| of each component.  (Each
  order is within 3 ulp of the exact result, tests/test_oracle_sse.py, so two
  orders could in principle differ by 6; on these inputs the largest
  difference is 3, reported by the test.)
"""
import numpy as np
import pytest

import oracle
import qla_amd
from qla_amd import lattice
from gpu_util import assert_bitwise, dev, host, say

pytestmark = pytest.mark.gpu

ORDER_GENERIC, ORDER_SSE3 = 0, 1
DIMS = (32, 32, 32, 32)
SYMS = list(oracle.SSE_SYMBOLS)
_data = {}


def data():
    if not _data:
        n = lattice.volume(DIMS)
        fw, bw = lattice.neighbours(DIMS)
        # one spare record in front of every field: the views at +0 and at
        # one complex (8 bytes) in are 16-byte and 8-byte aligned
        _data.update(n=n, fw=fw, bw=bw,
                     u=lattice.random_su3(n, 1).astype(np.float32),
                     v=lattice.random_field("V", n + 1, 2).astype(np.float32),
                     h=lattice.random_field("H", n + 1, 3).astype(np.float32),
                     d=lattice.random_field("D", n + 1, 4).astype(np.float32),
                     rv=lattice.random_field("V", n + 1, 5).astype(np.float32),
                     rd=lattice.random_field("D", n + 1, 6).astype(np.float32))
        for k in ("u", "v", "h", "d"):
            _data["d" + k] = dev(_data[k])
    return _data


def view(t, off_complex, n, shape):
    """n records of `shape` starting off_complex complexes into flat tensor t"""
    flat = t.reshape(-1)
    c = 2 * int(np.prod(shape[:-1]))
    return flat[2 * off_complex:2 * off_complex + n * c].view((n,) + shape)


def host_view(a, off_complex, n, shape):
    c = 2 * int(np.prod(shape[:-1]))
    return np.ascontiguousarray(a.reshape(-1)[2 * off_complex:2 * off_complex + n * c].reshape((n,) + shape))


def run(sym, mu, sign, r_off, b_off, order):
    """GPU call of QLA_F3_<sym> with r and the gathered records offset by
    r_off / b_off complexes; returns (gpu result, r0, a, src, idx, offsets)"""
    import torch
    D = data()
    n = D["n"]
    kind, eqop, adj = oracle.SSE_SYMBOLS[sym]
    spin = kind != oracle.MATVEC
    rshape = (3, 4, 2) if spin else (3, 2)
    sshape = {oracle.MATVEC: (3, 2), oracle.SPRECON: (3, 2, 2), oracle.WILSON: (3, 4, 2)}[kind]
    src_h = {oracle.MATVEC: D["v"], oracle.SPRECON: D["h"], oracle.WILSON: D["d"]}[kind]
    src_d = {oracle.MATVEC: D["dv"], oracle.SPRECON: D["dh"], oracle.WILSON: D["dd"]}[kind]
    r_host = D["rd"] if spin else D["rv"]
    r0 = host_view(r_host, r_off, n, rshape)
    src = host_view(src_h, b_off, n, sshape)
    r = view(dev(r_host), r_off, n, rshape)
    sv = view(src_d, b_off, n, sshape)
    idx = None if adj else (D["fw"] if sign == 1 else D["bw"])[mu % 4]     # mu = 4: the x neighbours
    lib = qla_amd.load()
    lib.QLA_MI355X_set_arith_order(order)
    try:
        f = "QLA_F3_" + sym
        say(f"{f} mu={mu} sign={sign} r_off={r_off} b_off={b_off} order={order}")
        b_arg = sv if adj else qla_amd.pointer_array(sv, idx)
        args = [r, D["du"], b_arg] + ([mu, sign] if spin else []) + [n]
        qla_amd.call(f, *args)
        qla_amd.check()
    finally:
        lib.QLA_MI355X_set_arith_order(ORDER_GENERIC)
    als = (r.data_ptr() % 16, D["du"].data_ptr() % 16, sv.data_ptr() % 16)
    return host(r), r0, src, idx, als


CASES = [(s, mu, sign) for s in SYMS for mu, sign in
         (((m, g) for m in range(5) for g in (1, -1)) if oracle.SSE_SYMBOLS[s][0] != oracle.MATVEC else [(0, 1)])]


@pytest.mark.parametrize("sym,mu,sign", CASES, ids=lambda x: str(x))
def test_sse3_order_bit_exact(sym, mu, sign):
    D = data()
    offs = [(0, 0), (1, 0), (0, 1)] if mu in (0, 4) or oracle.SSE_SYMBOLS[sym][0] == oracle.MATVEC else [(0, 0)]
    for r_off, b_off in offs:
        got, r0, src, idx, (ra, aa, ba) = run(sym, mu, sign, r_off, b_off, ORDER_SSE3)
        exp = oracle.sse_site(sym, r0.copy(), D["u"], src, idx, mu, sign, r_al=ra, a_al=aa, b0_al=ba)
        assert_bitwise(got, exp, f"{sym} mu={mu} sign={sign} offsets=({r_off},{b_off}) SSE3 order")


def _scale(sym, r0, u, src, idx, mu, sign):
    kind, eqop, adj = oracle.SSE_SYMBOLS[sym]
    a = u.astype(np.float64)
    am = np.hypot(a[..., 0], a[..., 1])
    if adj:
        am = np.swapaxes(am, 1, 2)
    s = src.astype(np.float64)[np.arange(len(r0)) if idx is None else idx]
    if kind == oracle.WILSON:
        s = np.stack([oracle.spproj1(x, mu, sign) for x in s[:4096]])     # the subset checked below
        a, am = a[:4096], am[:4096]
    xm = np.hypot(s[..., 0], s[..., 1])
    rabs = np.abs(r0.astype(np.float64)) if eqop == oracle.PEQ else 0.0
    if kind == oracle.MATVEC:
        m = np.einsum("nik,nk->ni", am, xm)
        return rabs + m[..., None]
    m = np.einsum("nik,nks->nis", am, xm).max(axis=-1)           # per colour row, either spin
    rabs = rabs[:len(m)] if np.ndim(rabs) else rabs
    return rabs + m[:, :, None, None]


@pytest.mark.parametrize("sym", SYMS)
def test_generic_order_close_to_sse3_build(sym):
    D = data()
    worst = 0.0
    spin = oracle.SSE_SYMBOLS[sym][0] != oracle.MATVEC
    for mu, sign in ([(m, g) for m in range(4) for g in (1, -1)] if spin else [(0, 1)]):
        got, r0, src, idx, (ra, aa, ba) = run(sym, mu, sign, 0, 0, ORDER_GENERIC)
        exp = oracle.sse_site(sym, r0.copy(), D["u"], src, idx, mu, sign, r_al=ra, a_al=aa, b0_al=ba)
        scale = _scale(sym, r0, D["u"], src, idx, mu, sign)
        m = len(scale)
        d = np.abs(got[:m].astype(np.float64) - exp[:m]) / np.spacing(scale.astype(np.float32)).astype(np.float64)
        worst = max(worst, float(d.max()))
        assert d.max() <= 4.0, (sym, mu, sign, float(d.max()))
    print(f"{sym}: generic vs SSE3 build, worst {worst:.2f} ulp of the scale")


D3_SYMS = list(oracle.SSE2_D3_SYMBOLS)
_d3 = {}


def d3_data():
    if not _d3:
        n = lattice.volume(DIMS)
        fw, _ = lattice.neighbours(DIMS)
        _d3.update(n=n, idx=fw[0], u=lattice.random_su3(n, 11), v=lattice.random_field("V", n, 12),
                   r=lattice.random_field("V", n, 13))
        _d3["du"], _d3["dv"] = dev(_d3["u"]), dev(_d3["v"])
    return _d3


def run_d3(sym, order):
    D = d3_data()
    n = D["n"]
    r = dev(D["r"])
    pointer = sym.endswith("pV")
    b_arg = qla_amd.pointer_array(D["dv"], D["idx"]) if pointer else D["dv"]
    lib = qla_amd.load()
    lib.QLA_MI355X_set_arith_order(order)
    try:
        say(f"QLA_D3_{sym} 32^4 order={order}")
        qla_amd.call("QLA_D3_" + sym, r, D["du"], b_arg, n)
        qla_amd.check()
    finally:
        lib.QLA_MI355X_set_arith_order(ORDER_GENERIC)
    exp = oracle.sse2_d3_site(sym, D["r"].copy(), D["u"], D["v"], D["idx"] if pointer else None)
    return host(r), exp


@pytest.mark.parametrize("sym", D3_SYMS)
def test_d3_sse3_order_bit_exact(sym):
    got, exp = run_d3(sym, ORDER_SSE3)
    assert_bitwise(got, exp, f"QLA_D3_{sym} SSE3 order")


@pytest.mark.parametrize("sym", D3_SYMS)
def test_d3_generic_order_close_to_sse3_build(sym):
    D = d3_data()
    got, exp = run_d3(sym, ORDER_GENERIC)
    src = D["v"][D["idx"]] if sym.endswith("pV") else D["v"]
    am = np.hypot(D["u"][..., 0], D["u"][..., 1])
    xm = np.hypot(src[..., 0], src[..., 1])
    scale = np.einsum("nik,nk->ni", am, xm)[..., None]
    if oracle.SSE2_D3_SYMBOLS[sym] == oracle.PEQ:
        scale = scale + np.abs(D["r"])
    d = np.abs(got - exp) / np.spacing(scale)
    assert d.max() <= 4.0, (sym, float(d.max()))
    assert np.count_nonzero(got != exp) > 0
    print(f"QLA_D3_{sym}: generic vs SSE3 build, worst {float(d.max()):.2f} ulp of the scale")
