"""GPU parity of the global reductions (norm2, dot) through the C ABI.

Bar (the reference's own summation order depends on its OpenMP thread count,
so no order is privileged): the GPU's double-precision sum must agree with
the oracle's double sum to 1e-12 relative (it is a sum of the same rounded
terms in another order), which makes a float result (F3) equal to the
oracle's within one float ulp and a double result (D3, DF3) within 1e-12.
Both host and device result pointers are exercised, and n = 0 must give 0.
"""
import ctypes as C

import numpy as np
import pytest

import qla_amd
from qla_amd import lattice
from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu


def ulp_f32(x):
    return np.spacing(np.float32(abs(x)))


def run_host_result(name, ctype, *args):
    out = (ctype * 2)()
    qla_amd.call(name, C.cast(out, C.c_void_p).value, *args)
    qla_amd.check()
    return [out[0], out[1]]


def run_device_result(name, torch_dtype, *args):
    import torch
    out = torch.full((2,), -7.0, dtype=torch_dtype, device="cuda")
    qla_amd.call(name, out, *args)
    qla_amd.check()
    return host(out).tolist()


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 65536, 5 ** 7])
def test_norm2(orc, n):
    import torch
    for kind, ncomp in (("V", 3), ("D", 12)):
        a64 = lattice.random_field(kind, n, 21)
        for prec in ("F", "D"):
            dt = np.float32 if prec == "F" else np.float64
            a = a64.astype(dt)
            ref = orc.norm2(a)
            da = dev(a)
            name = f"QLA_{prec}3_r_veq_norm2_{kind}"
            say(f"{name} n={n}")
            ct = C.c_float if prec == "F" else C.c_double
            tt = torch.float32 if prec == "F" else torch.float64
            for got in (run_host_result(name, ct, da, n)[0], run_device_result(name, tt, da, n)[0]):
                if prec == "F":
                    assert abs(got - np.float32(ref)) <= ulp_f32(ref), (name, got, ref)
                else:
                    assert abs(got - ref) <= 1e-12 * abs(ref), (name, got, ref)
            if kind == "D" and prec == "F":
                name = "QLA_DF3_r_veq_norm2_D"
                say(f"{name} n={n}")
                for got in (run_host_result(name, C.c_double, da, n)[0],
                            run_device_result(name, torch.float64, da, n)[0]):
                    assert abs(got - ref) <= 1e-12 * abs(ref), (name, got, ref)


@pytest.mark.parametrize("n", [1, 63, 1000, 65536, 5 ** 7])
def test_dot(orc, n):
    import torch
    for kind in ("V", "D"):
        a64 = lattice.random_field(kind, n, 31)
        b64 = lattice.random_field(kind, n, 32)
        for prec in ("F", "D"):
            if kind == "V" and prec == "D":
                continue  # only the F3 V dot is exported
            dt = np.float32 if prec == "F" else np.float64
            a, b = a64.astype(dt), b64.astype(dt)
            ref = orc.dot(a, b)
            scale = np.sqrt(orc.norm2(a) * orc.norm2(b))
            da, db = dev(a), dev(b)
            name = f"QLA_{prec}3_c_veq_{kind}_dot_{kind}"
            say(f"{name} n={n}")
            ct = C.c_float if prec == "F" else C.c_double
            tt = torch.float32 if prec == "F" else torch.float64
            for got in (run_host_result(name, ct, da, db, n), run_device_result(name, tt, da, db, n)):
                for g, e in ((got[0], ref.real), (got[1], ref.imag)):
                    if prec == "F":
                        assert abs(g - np.float32(e)) <= ulp_f32(e), (name, g, e)
                    else:
                        assert abs(g - e) <= 1e-12 * scale, (name, g, e)
            if kind == "D" and prec == "F":
                name = "QLA_DF3_c_veq_D_dot_D"
                say(f"{name} n={n}")
                for got in (run_host_result(name, C.c_double, da, db, n),
                            run_device_result(name, torch.float64, da, db, n)):
                    assert abs(got[0] - ref.real) <= 1e-12 * scale
                    assert abs(got[1] - ref.imag) <= 1e-12 * scale


def test_reduce_empty():
    """n = 0: the reference's loop does not run and the result is 0."""
    import torch
    a = torch.ones((8, 3, 4, 2), device="cuda")
    say("norm2/dot n=0")
    assert run_host_result("QLA_F3_r_veq_norm2_D", C.c_float, a, 0)[0] == 0.0
    assert run_device_result("QLA_F3_r_veq_norm2_D", torch.float32, a, 0)[0] == 0.0
    assert run_host_result("QLA_F3_c_veq_D_dot_D", C.c_float, a, a, 0) == [0.0, 0.0]


def test_misaligned_field(orc):
    """A float field starting 8 bytes past a 16-byte boundary (legal: float
    complex is 8-byte aligned) takes the per-complex path and must agree."""
    import torch
    n = 4097
    a = lattice.random_field("D", n + 1, 41).astype(np.float32)
    flat = dev(a.reshape(-1))
    shifted = flat[2:2 + n * 24]   # 8-byte shift
    ref = orc.norm2(np.ascontiguousarray(a.reshape(-1)[2:2 + n * 24]).reshape(n, 3, 4, 2))
    say("norm2 misaligned")
    got = run_host_result("QLA_DF3_r_veq_norm2_D", C.c_double, shifted, n)[0]
    assert abs(got - ref) <= 1e-12 * ref
