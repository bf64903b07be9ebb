"""Every benchmark configuration at its full size, through the C ABI, against
the oracle on the same seeded inputs.  (Configuration 1, the 32^4 lattice,
is checked at full size in test_gpu_site.py and test_gpu_sse_order.py.)

* c0, 16^4 = 65,536 sites: V_veq_M_times_V bit-exact; r_veq_norm2_V within
  one float ulp of the long-double sum of the reference's terms; the F3
  results' error against the FP64 (D3) results of the same inputs.
* c2, 48^4 = 5,308,416 sites: spproj (mu 0, +1) and sprecon (mu 3, -1)
  bit-exact, F3 and D3; DF3 norm2 / dot within 1e-14 of the scale; both
  compressed link formats (binary16, bfloat16) bit-exact against the format
  model on the first 65,536 sites and, over the whole lattice, within their
  accuracy bar of the FP64 result: binary16 links 1e-3 (the compressed
  form's target); bfloat16 links 5e-3 -- bfloat16 keeps 8 significand bits, a unit
  roundoff of 2^-8 = 3.9e-3 on every stored link real, so no 12-real
  bfloat16 link form can reach 1e-3 (measured 1.8e-3).
* c3: the 64 jobs of draw_jobs(64, seed=7), L in {8..32}, in one batch
  launch: every job bit-exact against the oracle.
* c4: 2,097,152 Nc = 3 matrices (exp on i*H with spectra in [-4, 4], sqrt and
  invsqrt on B^dagger B + eps, eps log-uniform in [1e-4, 1]) and 262,144 at
  Nc = 4 and Nc = 8: bit-exact against the oracle (the reference templates'
  operation order), and the FP32 results' error against FP64 (the same
  functions in double precision on the same inputs) -- the reference
  algorithm's own error, which the bit-exact GPU result shares: normwise
  over the batch about 2e-7 for exp and sqrt and 1e-4 (Nc = 3) to 4e-4
  (Nc = 8) for invsqrt, whose B^dagger B + 1e-4 inputs have condition
  numbers near 1e4 (per-matrix up to 3e-3); bars C4_TOL.
"""
import ctypes as C

import numpy as np
import pytest

import qla_amd
from oracle import compressed as cm
from qla_amd import batch as qb
from qla_amd import lattice
from gpu_util import assert_bitwise, dev, host, say

pytestmark = pytest.mark.gpu

C2 = (48, 48, 48, 48)
_c2 = {}


def rel_err(x, ref):
    return float(np.linalg.norm((np.asarray(x, np.float64) - ref).ravel()) / np.linalg.norm(ref.ravel()))


def test_c0_full_size(orc):
    n = 65536
    u = lattice.random_su3(n, 1)
    v = lattice.random_field("V", n, 2)
    u32, v32 = u.astype(np.float32), v.astype(np.float32)
    du, dv = dev(u32), dev(v32)
    r = dev(np.zeros_like(v32))
    say("c0 QLA_F3_V_veq_M_times_V n=65536")
    qla_amd.call("QLA_F3_V_veq_M_times_V", r, du, dv, n)
    qla_amd.check()
    exp = np.zeros_like(v32)
    orc.V_eqop_M_times_V(orc.EQ, exp, u32, v32)
    got = host(r)
    assert_bitwise(got, exp, "c0 V_veq_M_times_V")
    r64 = dev(np.zeros_like(v))
    qla_amd.call("QLA_D3_V_veq_M_times_V", r64, dev(u32.astype(np.float64)), dev(v32.astype(np.float64)), n)
    qla_amd.check()
    assert rel_err(got, host(r64)) <= 1e-6
    out = (C.c_float * 1)()
    say("c0 QLA_F3_r_veq_norm2_V n=65536")
    qla_amd.call("QLA_F3_r_veq_norm2_V", C.addressof(out), dv, n)
    qla_amd.check()
    ref = orc.norm2_terms_ld(v32)
    assert abs(out[0] - np.float32(ref)) <= np.spacing(np.float32(ref))


def c2_data():
    if not _c2:
        n = lattice.volume(C2)
        fw, bw = lattice.neighbours(C2)
        u64 = lattice.random_su3(n, 51)
        d64 = lattice.random_field("D", n, 52)
        h64 = lattice.random_field("H", n, 53)
        _c2.update(n=n, fw=fw, bw=bw, u64=u64, d64=d64, h64=h64, u=u64.astype(np.float32),
                   d=d64.astype(np.float32), h=h64.astype(np.float32))
    return _c2


def test_c2_plain_48(orc):
    D = c2_data()
    n = D["n"]
    du, dd, dh = dev(D["u"]), dev(D["d"]), dev(D["h"])
    idx = D["fw"][0]
    say("c2 QLA_F3_H_veq_spproj_M_times_pD 48^4 mu=0 sign=+1")
    r = dev(np.zeros((n, 3, 2, 2), np.float32))
    qla_amd.call("QLA_F3_H_veq_spproj_M_times_pD", r, du, qla_amd.pointer_array(dd, idx), 0, 1, n)
    qla_amd.check()
    exp = np.zeros((n, 3, 2, 2), np.float32)
    orc.H_eqop_spproj_M_times_pD(orc.EQ, exp, D["u"], D["d"], idx, 0, 1)
    got = host(r)
    assert_bitwise(got, exp, "c2 spproj F3")
    ref = np.zeros((n, 3, 2, 2))
    orc.H_eqop_spproj_M_times_pD(orc.EQ, ref, D["u64"], D["d64"], idx, 0, 1)
    assert rel_err(got, ref) <= 1e-6
    del r, got, exp
    idx = D["bw"][3]
    say("c2 QLA_F3_D_vpeq_sprecon_M_times_pH 48^4 mu=3 sign=-1")
    r = dev(D["d"])
    qla_amd.call("QLA_F3_D_vpeq_sprecon_M_times_pH", r, du, qla_amd.pointer_array(dh, idx), 3, -1, n)
    qla_amd.check()
    exp = D["d"].copy()
    orc.D_eqop_sprecon_M_times_pH(orc.PEQ, exp, D["u"], D["h"], idx, 3, -1)
    assert_bitwise(host(r), exp, "c2 sprecon F3")
    del r, exp, dd, dh
    # D3: the same calls in double precision on the same (double) inputs
    du64, dd64 = dev(D["u64"]), dev(D["d64"])
    say("c2 QLA_D3_H_veq_spproj_M_times_pD 48^4")
    r = dev(np.zeros((n, 3, 2, 2)))
    qla_amd.call("QLA_D3_H_veq_spproj_M_times_pD", r, du64, qla_amd.pointer_array(dd64, D["fw"][0]), 0, 1, n)
    qla_amd.check()
    assert_bitwise(host(r), ref, "c2 spproj D3")


def test_c2_reductions_48(orc):
    D = c2_data()
    n = D["n"]
    dd, dh2 = dev(D["d"]), dev(D["d"][::-1].copy())
    b = D["d"][::-1].copy()
    out = (C.c_double * 2)()
    say("c2 QLA_DF3_r_veq_norm2_D 48^4")
    qla_amd.call("QLA_DF3_r_veq_norm2_D", C.addressof(out), dd, n)
    qla_amd.check()
    ref = orc.norm2_terms_ld(D["d"])
    assert abs(np.longdouble(out[0]) - ref) <= 1e-14 * ref
    say("c2 QLA_DF3_c_veq_D_dot_D 48^4")
    qla_amd.call("QLA_DF3_c_veq_D_dot_D", C.addressof(out), dd, dh2, n)
    qla_amd.check()
    re, im, scale = orc.dot_terms_ld(D["d"], b)
    assert abs(np.longdouble(out[0]) - re) <= 1e-14 * scale
    assert abs(np.longdouble(out[1]) - im) <= 1e-14 * scale
    f = (C.c_float * 2)()
    qla_amd.call("QLA_F3_r_veq_norm2_D", C.addressof(f), dd, n)
    qla_amd.check()
    assert abs(f[0] - np.float32(ref)) <= np.spacing(np.float32(ref))


# accuracy bars of the compressed forms against FP64 (see the module text)
CTOL = {"f16": 1e-3, "bf16": 5e-3}
FMTS = {"bf16": 0, "f16": 1}
SUB = 65536


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
def test_c2_compressed_48(orc, fmt):
    import torch
    D = c2_data()
    n = D["n"]
    lib = qla_amd.load()
    dl = torch.empty((n, 12), dtype=torch.int16, device="cuda")
    lib.QLA_MI355X_F3_M_pack12(dl.data_ptr(), FMTS[fmt], dev(D["u"]).data_ptr(), n)
    ddrec = torch.empty((n, 52), dtype=torch.uint8, device="cuda")
    lib.QLA_MI355X_F3_D_packh(ddrec.data_ptr(), dev(D["d"]).data_ptr(), n)
    dhrec = torch.empty((n, 28), dtype=torch.uint8, device="cuda")
    lib.QLA_MI355X_F3_H_packh(dhrec.data_ptr(), dev(D["h"]).data_ptr(), n)
    qla_amd.check()
    link, drec, hrec = host(dl).view(np.uint16), host(ddrec), host(dhrec)
    assert np.array_equal(link[:SUB], cm.pack_link(D["u"][:SUB], fmt))
    # spproj, mu = 0, sign = +1
    idx = D["fw"][0]
    r = torch.zeros((n, 28), dtype=torch.uint8, device="cuda")
    say(f"c2 compressed spproj {fmt} 48^4")
    lib.QLA_MI355X_F3_Hh_veq_spproj_M12_times_pDh(r.data_ptr(), dl.data_ptr(), FMTS[fmt],
                                                 qla_amd.pointer_array(ddrec, idx).data_ptr(), 0, 1, n)
    qla_amd.check()
    got = host(r)
    np.testing.assert_array_equal(got[:SUB], cm.spproj(link[:SUB], fmt, drec, idx[:SUB], 0, 1))
    ref = np.zeros((n, 3, 2, 2))
    orc.H_eqop_spproj_M_times_pD(orc.EQ, ref, D["u64"], D["d64"], idx, 0, 1)
    v, s = cm.from_records(got, 12)
    e = rel_err(cm.unpack_fermion(v, s, ref.shape), ref)
    print(f"c2 compressed spproj {fmt}: rel err vs FP64 {e:.3e}")
    assert e <= CTOL[fmt]
    del r, got, ref
    # sprecon, mu = 3, sign = -1
    idx = D["bw"][3]
    r = ddrec.clone()
    say(f"c2 compressed sprecon {fmt} 48^4")
    lib.QLA_MI355X_F3_Dh_vpeq_sprecon_M12_times_pHh(r.data_ptr(), dl.data_ptr(), FMTS[fmt],
                                                   qla_amd.pointer_array(dhrec, idx).data_ptr(), 3, -1, n)
    qla_amd.check()
    got = host(r)
    np.testing.assert_array_equal(got[:SUB], cm.sprecon(link[:SUB], fmt, hrec, idx[:SUB], drec[:SUB], 3, -1))
    ref = D["d64"].copy()
    orc.D_eqop_sprecon_M_times_pH(orc.PEQ, ref, D["u64"], D["h64"], idx, 3, -1)
    v, s = cm.from_records(got, 24)
    e = rel_err(cm.unpack_fermion(v, s, ref.shape), ref)
    print(f"c2 compressed sprecon {fmt}: rel err vs FP64 {e:.3e}")
    assert e <= CTOL[fmt]
    # the reductions of the compressed fermions (decoded terms, DF3 bar)
    out = (C.c_double * 2)()
    lib.QLA_MI355X_DF3_r_veq_norm2_Dh(C.addressof(out), ddrec.data_ptr(), n)
    qla_amd.check()
    dv, dsc = cm.from_records(drec, 24)
    dec = cm.unpack_fermion(dv, dsc, D["d"].shape)
    t = orc.norm2_terms_ld(dec)
    assert abs(np.longdouble(out[0]) - t) <= 1e-14 * t
    e = abs(out[0] - float(orc.norm2_ld(D["d64"]))) / float(orc.norm2_ld(D["d64"]))
    print(f"c2 compressed norm2 {fmt}: rel err vs FP64 {e:.3e}")
    assert e <= 1e-3


def test_c3_exact_draw(orc):
    """The benchmark's 64-job batch (draw_jobs(64, seed=7)), every job against
    the oracle after one batch launch."""
    desc = qb.draw_jobs(64, seed=7)
    jobs, keep = qb.checkerboard_jobs(desc)
    assert sum(kp["n"] for kp in keep) > 4_000_000
    b = qb.Batch(jobs)
    say(f"c3 batch of 64 jobs, {sum(kp['n'] for kp in keep)} sites")
    b.run()
    qla_amd.check()
    for k, kp in enumerate(keep):
        u, src, d0 = host(kp["u"]), host(kp["src"]), host(kp["dst0"])
        rec = kp["src"].stride(0) * 4
        idx = ((host(kp["ptr"]) - kp["src"].data_ptr()) // rec).astype(np.int64)
        if kp["op"] == 0:
            exp = d0.copy()
            orc.V_eqop_M_times_V(orc.PEQ, exp, u, src, b_idx=idx)
        elif kp["op"] == 1:
            exp = np.zeros_like(d0)
            orc.H_eqop_spproj_M_times_pD(orc.EQ, exp, u, src, idx, kp["mu"], kp["sign"])
        else:
            exp = d0.copy()
            orc.D_eqop_sprecon_M_times_pH(orc.PEQ, exp, u, src, idx, kp["mu"], kp["sign"])
        assert_bitwise(host(kp["dst"]), exp, f"c3 job {k} op {kp['op']} n={kp['n']}")
    b.close()


C4_INPUTS = {"exp": (lattice.random_antihermitian, 41), "sqrt": (lattice.random_hpd, 42),
             "invsqrt": (lattice.random_hpd, 42)}
C4_TOL = {"exp": 1e-6, "sqrt": 1e-6, "invsqrt": 1e-3}     # normwise FP32 vs FP64, see the module text


@pytest.mark.parametrize("fn", ["exp", "sqrt", "invsqrt"])
def test_c4_nc3_full(orc, fn):
    n = 2097152
    gen, seed = C4_INPUTS[fn]
    a64 = gen(n, 3, seed)
    a = a64.astype(np.float32)
    r = dev(np.zeros_like(a))
    say(f"c4 QLA_F3_M_veq_{fn}_M n={n}")
    qla_amd.call(f"QLA_F3_M_veq_{fn}_M", r, dev(a), n)
    qla_amd.check()
    got = host(r)
    assert_bitwise(got, orc.matfn(fn, a), f"c4 F3 {fn}")
    r64 = dev(np.zeros(a.shape))
    qla_amd.call(f"QLA_D3_M_veq_{fn}_M", r64, dev(a.astype(np.float64)), n)
    qla_amd.check()
    e = rel_err(got, host(r64))
    print(f"c4 F3 {fn}: rel err vs D3 {e:.3e}")
    assert e <= C4_TOL[fn]


@pytest.mark.parametrize("fn", ["exp", "sqrt", "invsqrt"])
@pytest.mark.parametrize("nc", [4, 8])
def test_c4_fn_full(orc, fn, nc):
    n = 262144
    gen, seed = C4_INPUTS[fn]
    a = gen(n, nc, seed + 2).astype(np.float32)
    r = dev(np.zeros_like(a))
    say(f"c4 QLA_FN_M_veq_{fn}_M nc={nc} n={n}")
    qla_amd.call("QLA_FN_M_veq_{}_M".format(fn), nc, r, dev(a), n)
    qla_amd.check()
    got = host(r)
    assert_bitwise(got, orc.matfn(fn, a), f"c4 FN{nc} {fn}")
    e = rel_err(got, orc.matfn(fn, a.astype(np.float64)))
    print(f"c4 FN{nc} {fn}: rel err vs FP64 {e:.3e}")
    assert e <= C4_TOL[fn]


C4_CLOSED = {
    # fn: (input, symbol suffix)
    "expA": ("antiherm_phase", "M_veq_expA_M"),
    "expTA": ("antiherm", "M_veq_expTA_M"),
    "sqrtPH": ("hpd", "M_veq_sqrtPH_M"),
    "invsqrtPH": ("hpd", "M_veq_invsqrtPH_M"),
    "eigenvalsH": ("herm", "V_veq_eigenvalsH_M"),
}


@pytest.mark.parametrize("fn", sorted(C4_CLOSED))
def test_c4_closed_forms_full(orc, fn):
    """The Nc = 3 closed forms on 2,097,152 F3 matrices, bit-exact against
    the oracle: exp of i*H (spectra in [-4, 4]; expA with a phase on the
    diagonal), the square roots of B^dagger B + eps, the eigenvalues of H.
    Their acosf / cbrtf / sincosf are the host library's algorithms on the
    device, so every branch and every rounding is the reference's."""
    n = 2097152
    kind, sym = C4_CLOSED[fn]
    if kind == "hpd":
        a64 = lattice.random_hpd(n, 3, 51)
    else:
        a64 = lattice.random_antihermitian(n, 3, 52)
        if kind == "antiherm_phase":
            ph = np.random.default_rng(53).uniform(-2.0, 2.0, n)
            a64[:, [0, 1, 2], [0, 1, 2], 1] += ph[:, None]
        elif kind == "herm":
            # H = -i * (i H): (re, im) of i*H -> (im, -re)
            a64 = np.stack([a64[..., 1], -a64[..., 0]], axis=-1)
    a = np.ascontiguousarray(a64.astype(np.float32))
    out_shape = (n, 3, 2) if fn == "eigenvalsH" else a.shape
    r = dev(np.zeros(out_shape, np.float32))
    say(f"c4 QLA_F3_{sym} n={n}")
    qla_amd.call(f"QLA_F3_{sym}", r, dev(a), n)
    qla_amd.check()
    assert_bitwise(host(r), orc.matfunc(fn, a), f"c4 F3 {fn}")
