"""GPU parity of the global reductions (norm2, dot) through the C ABI.

The reference's own summation order depends on its OpenMP thread count, so
no order is privileged.  The terms are not free, though: each is rounded as
the generated code forms it (component norms in the field precision, dot
terms in double).  The oracle sums exactly those terms in long double
(norm2_terms_ld / dot_terms_ld), and the bars are:
  * D3 / DF3 (double results): within 1e-14 x scale, scale = sum |terms|
    (for norm2 the sum itself) -- the QLA_D3 accuracy bar; a double sum
    of n terms in any order is within ~log2(n) * 2^-53 * scale of it;
  * F3 (float results): within one float ulp of the long-double sum (the
    double sum rounded to float can land on either side of a rounding
    boundary, never further).
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


D_BAR = 1e-14


def assert_f32(got, ref, what):
    assert abs(got - np.float32(ref)) <= ulp_f32(ref), (what, got, float(ref))


def assert_d(got, ref, scale, what):
    err = abs(np.longdouble(got) - ref)
    assert err <= D_BAR * scale, (what, got, float(ref), float(err / scale))


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 65536, 5 ** 7])
def test_norm2(orc, n):
    import torch
    for kind, ncomp in (("V", 3), ("D", 12)):
        a64 = lattice.random_field(kind, n, 21)
        for prec in ("F", "D"):
            dt = np.float32 if prec == "F" else np.float64
            a = a64.astype(dt)
            ref = orc.norm2_terms_ld(a)
            da = dev(a)
            name = f"QLA_{prec}3_r_veq_norm2_{kind}"
            say(f"{name} n={n}")
            ct = C.c_float if prec == "F" else C.c_double
            tt = torch.float32 if prec == "F" else torch.float64
            for got in (run_host_result(name, ct, da, n)[0], run_device_result(name, tt, da, n)[0]):
                if prec == "F":
                    assert_f32(got, ref, name)
                else:
                    assert_d(got, ref, ref, name)
            if kind == "D" and prec == "F":
                name = "QLA_DF3_r_veq_norm2_D"
                say(f"{name} n={n}")
                for got in (run_host_result(name, C.c_double, da, n)[0],
                            run_device_result(name, torch.float64, da, n)[0]):
                    assert_d(got, ref, ref, name)


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
            re, im, scale = orc.dot_terms_ld(a, b)
            da, db = dev(a), dev(b)
            name = f"QLA_{prec}3_c_veq_{kind}_dot_{kind}"
            say(f"{name} n={n}")
            ct = C.c_float if prec == "F" else C.c_double
            tt = torch.float32 if prec == "F" else torch.float64
            for got in (run_host_result(name, ct, da, db, n), run_device_result(name, tt, da, db, n)):
                for g, e in ((got[0], re), (got[1], im)):
                    if prec == "F":
                        assert_f32(g, e, name)
                    else:
                        assert_d(g, e, scale, name)
            if kind == "D" and prec == "F":
                name = "QLA_DF3_c_veq_D_dot_D"
                say(f"{name} n={n}")
                for got in (run_host_result(name, C.c_double, da, db, n),
                            run_device_result(name, torch.float64, da, db, n)):
                    assert_d(got[0], re, scale, name)
                    assert_d(got[1], im, scale, name)


def test_reduce_empty():
    """n = 0: the reference's loop does not run and the result is 0."""
    import torch
    a = torch.ones((8, 3, 4, 2), device="cuda")
    say("norm2/dot n=0")
    assert run_host_result("QLA_F3_r_veq_norm2_D", C.c_float, a, 0)[0] == 0.0
    assert run_device_result("QLA_F3_r_veq_norm2_D", torch.float32, a, 0)[0] == 0.0
    assert run_host_result("QLA_F3_c_veq_D_dot_D", C.c_float, a, a, 0) == [0.0, 0.0]


def test_result_memory_kinds(orc):
    """The result pointer may be pageable host memory, page-locked host
    memory the caller allocated, or device memory; a host result comes back
    through the library's own page-locked slot.  All three give the same
    bits, for the reductions and for a global sum (up to 288 doubles), on
    consecutive calls that reuse the slot."""
    import torch
    n = 3000
    a = lattice.random_field("D", n, 90).astype(np.float32)     # an F3 field
    da = dev(a)
    pinned = torch.zeros(288, dtype=torch.float64, pin_memory=True)
    for rep in range(2):
        say(f"norm2 / dot / sum result kinds, call {rep}")
        ref = run_host_result("QLA_DF3_r_veq_norm2_D", C.c_double, da, n)[0]
        qla_amd.call("QLA_DF3_r_veq_norm2_D", pinned.data_ptr(), da, n)
        qla_amd.check()
        assert pinned[0].item() == ref
        assert run_device_result("QLA_DF3_r_veq_norm2_D", torch.float64, da, n)[0] == ref
        dref = run_host_result("QLA_DF3_c_veq_D_dot_D", C.c_double, da, da, n)
        qla_amd.call("QLA_DF3_c_veq_D_dot_D", pinned.data_ptr(), da, da, n)
        qla_amd.check()
        assert pinned[:2].tolist() == dref
        out = (C.c_double * 24)()
        qla_amd.call("QLA_DF3_d_veq_sum_D", C.cast(out, C.c_void_p).value, da, n)
        qla_amd.check()
        qla_amd.call("QLA_DF3_d_veq_sum_D", pinned.data_ptr(), da, n)
        qla_amd.check()
        assert pinned[:24].tolist() == list(out)
        dsum = torch.zeros(24, dtype=torch.float64, device="cuda")
        qla_amd.call("QLA_DF3_d_veq_sum_D", dsum, da, n)
        qla_amd.check()
        assert host(dsum).tolist() == list(out)
        # each component's sum against a long-double sum, within 1e-14 of
        # that component's sum |terms|
        terms = a.astype(np.longdouble).reshape(n, 24)
        err = np.abs(np.array(list(out), dtype=np.longdouble) - terms.sum(axis=0))
        assert np.all(err <= 1e-14 * np.abs(terms).sum(axis=0)), err
        da.mul_(0.5)
        a = a * np.float32(0.5)


def test_misaligned_field(orc):
    """A float field starting 8 bytes past a 16-byte boundary (legal: float
    complex is 8-byte aligned) takes the per-complex path and must agree."""
    import torch
    n = 4097
    a = lattice.random_field("D", n + 1, 41).astype(np.float32)
    flat = dev(a.reshape(-1))
    shifted = flat[2:2 + n * 24]   # 8-byte shift
    ref = orc.norm2_terms_ld(np.ascontiguousarray(a.reshape(-1)[2:2 + n * 24]).reshape(n, 3, 4, 2))
    say("norm2 misaligned")
    got = run_host_result("QLA_DF3_r_veq_norm2_D", C.c_double, shifted, n)[0]
    assert_d(got, ref, ref, "QLA_DF3_r_veq_norm2_D misaligned")


HIP_ERROR_STREAM_CAPTURE_UNSUPPORTED = 900


def _norm2_dev(lib, out, a, n):
    lib.QLA_DF3_r_veq_norm2_D(out.data_ptr(), a.data_ptr(), n)


def test_two_streams_async(orc):
    """One host thread, asynchronous calls, DF3 norm2 and dot alternating
    between two streams with device results, at a size with a full stage-1
    grid and one with a few blocks: each stream's
    reductions have their own scratch, so every result equals the one a
    synchronous call gives, bit for bit (same kernels, same order)."""
    import torch
    lib = qla_amd.load()
    sizes = [300_000, 2_000]
    f = {n: [dev(lattice.random_field("D", n, 60 + 2 * k + j).astype(np.float32)) for k in range(2)
             for j in range(2)] for n in sizes}
    exp = {}
    for n in sizes:
        for k in range(2):
            a, b = f[n][2 * k], f[n][2 * k + 1]
            out = torch.zeros(3, dtype=torch.float64, device="cuda")
            say(f"sync DF3 norm2/dot n={n} set {k}")
            lib.QLA_DF3_r_veq_norm2_D(out[0:1].data_ptr(), a.data_ptr(), n)
            lib.QLA_DF3_c_veq_D_dot_D(out[1:3].data_ptr(), a.data_ptr(), b.data_ptr(), n)
            qla_amd.check()
            exp[(n, k)] = host(out)
    streams = (torch.cuda.Stream(), torch.cuda.Stream())
    reps = 6
    outs = torch.zeros((reps, len(sizes), 2, 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    qla_amd.set_sync(False)
    try:
        say("async DF3 norm2/dot on two streams")
        for r in range(reps):
            for j, n in enumerate(sizes):
                for k in range(2):
                    qla_amd.set_stream(streams[k])
                    a, b = f[n][2 * k], f[n][2 * k + 1]
                    lib.QLA_DF3_r_veq_norm2_D(outs[r, j, k, 0:1].data_ptr(), a.data_ptr(), n)
                    lib.QLA_DF3_c_veq_D_dot_D(outs[r, j, k, 1:3].data_ptr(), a.data_ptr(), b.data_ptr(), n)
        torch.cuda.synchronize()
    finally:
        qla_amd.set_stream(None)
        qla_amd.set_sync(True)
    qla_amd.check()
    got = host(outs)
    for r in range(reps):
        for j, n in enumerate(sizes):
            for k in range(2):
                assert np.array_equal(got[r, j, k].view(np.uint64), exp[(n, k)].view(np.uint64)), (r, n, k)


def test_threads_own_scratch_and_release(orc):
    """Two host threads, each on its own stream, run reductions with host
    results at the same time: both get the synchronous results, and the
    scratch a thread allocated is released when the thread ends."""
    import threading
    import torch
    lib = qla_amd.load()
    n = 150_000
    fields = [dev(lattice.random_field("D", n, 70 + k).astype(np.float32)) for k in range(2)]
    exp = [run_host_result("QLA_DF3_r_veq_norm2_D", C.c_double, fields[k], n)[0] for k in range(2)]
    before = lib.QLA_MI355X_scratch_buffers()
    results = [[], []]
    errors = []

    def work(k):
        try:
            st = torch.cuda.Stream()
            qla_amd.set_stream(st)
            for _ in range(20):
                out = C.c_double()
                lib.QLA_DF3_r_veq_norm2_D(C.addressof(out), fields[k].data_ptr(), n)
                results[k].append(out.value)
            if lib.QLA_MI355X_last_error() != 0:
                errors.append(k)
            qla_amd.set_stream(None)
        except Exception as e:   # noqa: BLE001
            errors.append(repr(e))

    say("two threads, DF3 norm2 with host results")
    th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for k in range(2):
        assert results[k] == [exp[k]] * 20, k
    assert lib.QLA_MI355X_scratch_buffers() == before


def test_reduction_capture(orc):
    """A stream's first reduction inside a stream capture is refused (it
    would allocate) and allocates nothing; once the stream has run one
    reduction, reductions with device results capture into a hipGraph whose
    replays -- with eager work between them that changes the input -- each
    give the eager call's bits (a few-block and a full-grid size).  A host
    result is refused under capture."""
    import torch
    lib = qla_amd.load()
    sizes = [2_000, 300_000]
    fields = {n: dev(lattice.random_field("D", n, 80).astype(np.float32)) for n in sizes}
    st = torch.cuda.Stream()
    out = torch.zeros(len(sizes), dtype=torch.float64, device="cuda")
    x = torch.zeros(4, device="cuda")
    before = lib.QLA_MI355X_scratch_buffers()
    qla_amd.check()
    g0 = torch.cuda.CUDAGraph()
    say("first reduction inside a capture (refused)")
    qla_amd.set_sync(False)
    try:
        with torch.cuda.graph(g0, stream=st, capture_error_mode="thread_local"):
            x.add_(1.0)
            qla_amd.set_stream(torch.cuda.current_stream())
            _norm2_dev(lib, out[0:1], fields[sizes[0]], sizes[0])
            qla_amd.set_stream(None)
    finally:
        qla_amd.set_stream(None)
        qla_amd.set_sync(True)
    assert lib.QLA_MI355X_last_error() == HIP_ERROR_STREAM_CAPTURE_UNSUPPORTED
    assert lib.QLA_MI355X_scratch_buffers() == before
    del g0
    # warm the stream's scratch outside capture, then capture both sizes
    qla_amd.set_stream(st)
    _norm2_dev(lib, out[0:1], fields[sizes[0]], sizes[0])
    qla_amd.check()
    qla_amd.set_stream(None)
    g = torch.cuda.CUDAGraph()
    qla_amd.set_sync(False)
    say("capture DF3 norm2 (device results)")
    try:
        with torch.cuda.graph(g, stream=st, capture_error_mode="thread_local"):
            qla_amd.set_stream(torch.cuda.current_stream())
            for j, n in enumerate(sizes):
                _norm2_dev(lib, out[j:j + 1], fields[n], n)
            hres = C.c_double()
            lib.QLA_DF3_r_veq_norm2_D(C.addressof(hres), fields[sizes[0]].data_ptr(), sizes[0])
            refused = lib.QLA_MI355X_last_error()
            qla_amd.set_stream(None)
    finally:
        qla_amd.set_stream(None)
        qla_amd.set_sync(True)
    assert refused == HIP_ERROR_STREAM_CAPTURE_UNSUPPORTED
    qla_amd.check()
    for rep in range(3):
        say(f"graph replay {rep}")
        g.replay()
        torch.cuda.synchronize()
        got = host(out).copy()
        for j, n in enumerate(sizes):
            e = run_host_result("QLA_DF3_r_veq_norm2_D", C.c_double, fields[n], n)[0]
            assert got[j] == e, (rep, n, got[j], e)
            fields[n].mul_(2.0)      # eager work between replays: the next norm is exactly 4x
