"""GPU parity of the global reductions and precision conversions of every
type (V, H, D, M, P) through the C ABI (include/qla_mi355x_fields.h).

Reductions (norm2, dot, real part of dot; F3, D3, DF3; v/x forms, field and
pointer operands): the reference's summation order depends on its OpenMP
thread count, so no order is privileged; the terms are fixed (component
norms in the field precision, dot terms in double), and the oracle sums
exactly those terms in long double.  The GPU's result must agree with that
sum to 1e-14 of the sum's scale (sum |terms|), the QLA_D3 accuracy bar; a
float result (F3) may add one float ulp of rounding.  Conversions (DF3 float -> double, FD3 double -> float) are
bit-exact.  Sizes give ragged last tiles; x-forms use a parity subset,
pointer operands permuted targets; host and device result pointers and
n = 0 are exercised.
"""
import ctypes as C

import numpy as np
import pytest

import qla_amd
from qla_amd import lattice
from qla_amd.abi_gen import conversions, reductions
from gpu_util import assert_bitwise, dev, host, say

pytestmark = pytest.mark.gpu

N = {"V": 1000, "H": 700, "D": 500, "M": 600, "P": 150}
_cache = {}


def setup(t, prec):
    key = (t, prec)
    if key in _cache:
        return _cache[key]
    n = N[t]
    dt = np.float64 if prec == "D3" else np.float32
    rng = np.random.default_rng(len(t) * 7 + n)
    s = dict(n=n, a=lattice.random_field(t, n, 11 + n).astype(dt), b=lattice.random_field(t, n, 12 + n).astype(dt),
             a_idx=rng.permutation(n), b_idx=rng.permutation(n),
             index=np.sort(rng.choice(n, n // 2, replace=False)).astype(np.int32),
             own={k: rng.choice(n, n // 3, replace=False).astype(np.int32) for k in "rab"})
    s["da"], s["db"] = dev(s["a"]), dev(s["b"])
    s["pa"] = qla_amd.pointer_array(s["da"], s["a_idx"])
    s["pb"] = qla_amd.pointer_array(s["db"], s["b_idx"])
    s["dindex"] = dev(s["index"])
    s["down"] = {k: dev(v) for k, v in s["own"].items()}
    _cache[key] = s
    return s


def steps(v, s):
    if v.xf:
        return len(s["index"])
    if v.ia or v.ib:
        return len(s["own"]["a"])
    return s["n"]


def expected(orc, v, s):
    index = s["index"] if v.xf else None
    n = steps(v, s)
    # v-form records with a per-operand index on the other operand: the first n
    ga = orc.records(s["a"], index, s["a_idx"] if v.pa else None,
                     own=s["own"]["a"] if v.ia else (np.arange(n) if v.ib else None))
    if v.kind == "norm2":
        ref = orc.norm2_terms_ld(ga)
        return ref, ref
    gb = orc.records(s["b"], index, s["b_idx"] if v.pb else None,
                     own=s["own"]["b"] if v.ib else (np.arange(n) if v.ia else None))
    re, im, scale = orc.dot_terms_ld(ga, gb)
    return (complex(re, im) if v.kind == "dot" else re), scale


def call(v, s, out_ptr, n=None):
    args = [out_ptr, s["pa"] if v.pa else s["da"]]
    if v.ia:
        args.append(s["down"]["a"])
    if v.kind != "norm2":
        args.append(s["pb"] if v.pb else s["db"])
    if v.ib:
        args.append(s["down"]["b"])
    if v.xf:
        args.append(s["dindex"])
    args.append(steps(v, s) if n is None else n)
    qla_amd.call(v.name, *args)
    qla_amd.check()


def check_value(v, got, exp, scale):
    f3 = v.prec == "F3"
    parts = [(got.real, exp.real), (got.imag, exp.imag)] if v.kind == "dot" else [(got, exp)]
    for g, e in parts:
        tol = 1e-14 * scale
        if f3:
            tol += float(np.spacing(np.float32(abs(e))))
        assert abs(g - e) <= tol, (v.name, g, e, tol)


@pytest.mark.parametrize("v", reductions(), ids=lambda v: v.name)
def test_reduction_symbol(orc, v):
    s = setup(v.t, "D3" if v.prec == "D3" else "F3")
    exp, scale = expected(orc, v, s)
    if v.kind == "norm2":
        scale = float(scale)
    real_t = C.c_float if v.prec == "F3" else C.c_double
    out = (real_t * 2)(-7.0, -7.0)
    say(f"{v.name} n={s['n']} host result")
    call(v, s, C.cast(out, C.c_void_p).value)
    got = complex(out[0], out[1]) if v.kind == "dot" else out[0]
    check_value(v, got, exp, scale)
    if v.kind != "dot":
        assert out[1] == -7.0          # a real result writes one value


@pytest.mark.parametrize("prec", ["F3", "D3", "DF3"])
@pytest.mark.parametrize("kind", ["norm2", "dot", "redot"])
def test_reduction_device_result_and_empty(orc, prec, kind):
    import torch
    for v in [x for x in reductions() if x.prec == prec and x.kind == kind and x.t in ("D", "P")
              and x.xf and x.pa and (x.pb or kind == "norm2") and not (x.ia or x.ib)]:
        s = setup(v.t, "D3" if prec == "D3" else "F3")
        exp, scale = expected(orc, v, s)
        tt = torch.float32 if prec == "F3" else torch.float64
        out = torch.full((2,), -7.0, dtype=tt, device="cuda")
        say(f"{v.name} device result")
        call(v, s, out)
        o = host(out)
        got = complex(o[0], o[1]) if kind == "dot" else o[0]
        check_value(v, got, exp, float(scale))
        say(f"{v.name} n=0")
        call(v, s, out, n=0)
        o = host(out)
        assert o[0] == 0 and (kind != "dot" or o[1] == 0)


@pytest.mark.parametrize("v", conversions(), ids=lambda v: v.name)
def test_conversion_symbol(orc, v):
    import torch
    src_prec = "F3" if v.lib == "DF3" else "D3"
    s = setup(v.t, src_prec)
    dest_dt = np.float64 if v.lib == "DF3" else np.float32
    r0 = lattice.random_field(v.t, s["n"], 99).astype(dest_dt)
    r = dev(r0)
    own, down = s["own"], s["down"]
    args = ([r] + ([down["r"]] if v.ir else []) + [s["pa"] if v.pa else s["da"]] + ([down["a"]] if v.ia else [])
            + ([s["dindex"]] if v.xf else []))
    n = len(s["index"]) if v.xf else (len(own["r"]) if (v.ir or v.ia) else s["n"])
    say(f"{v.name} n={n}")
    qla_amd.call(v.name, *args, n)
    qla_amd.check()
    exp = r0.copy()
    orc.convert(s["a"], exp, index=s["index"] if v.xf else None, a_idx=s["a_idx"] if v.pa else None,
                ir=own["r"] if v.ir else (np.arange(n) if v.ia else None),
                ia=own["a"] if v.ia else (np.arange(n) if v.ir else None))
    assert_bitwise(host(r), exp, v.name)
    say(f"{v.name} n=0")
    qla_amd.call(v.name, *args, 0)
    qla_amd.check()
    assert_bitwise(host(r), exp, f"{v.name} n=0 leaves r")
    del torch


@pytest.mark.parametrize("lib", ["DF3", "FD3"])
def test_conversion_large_and_misaligned(orc, lib):
    """A 2^20-site D conversion (the streaming path at full size) and an
    8-byte-offset float field."""
    import torch
    n = 1 << 20
    src_dt, dst_dt = (np.float32, np.float64) if lib == "DF3" else (np.float64, np.float32)
    a = lattice.random_field("D", n, 5).astype(src_dt) * src_dt(1e20)
    da = dev(a)
    r = torch.empty(a.shape, dtype=torch.float64 if dst_dt == np.float64 else torch.float32, device="cuda")
    say(f"QLA_{lib}_D_veq_D n={n}")
    qla_amd.call(f"QLA_{lib}_D_veq_D", r, da, n)
    qla_amd.check()
    exp = np.zeros(a.shape, dtype=dst_dt)
    orc.convert(a, exp)
    assert_bitwise(host(r), exp, f"QLA_{lib}_D_veq_D large")
    if lib == "DF3":
        flat = torch.cat([torch.zeros(2, dtype=torch.float32, device="cuda"), da.reshape(-1)[: 24 * 1000]])
        r2 = torch.empty((1000, 3, 4, 2), dtype=torch.float64, device="cuda")
        say("QLA_DF3_D_veq_D misaligned")
        qla_amd.call("QLA_DF3_D_veq_D", r2, flat.data_ptr() + 8, 1000)
        qla_amd.check()
        exp2 = np.zeros((1000, 3, 4, 2))
        orc.convert(a[:1000], exp2)
        assert_bitwise(host(r2), exp2, "misaligned")
