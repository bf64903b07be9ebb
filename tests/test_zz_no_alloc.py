"""The allocation-free entry point QLA_MI355X_DF3_r_veq_norm2_D_noalloc
against the allocating QLA_DF3_r_veq_norm2_D (the same kernels, so the bar
is bit-exact): two calls in a row, the second with a smaller input (fewer
stage-1 blocks, so the array holds the larger call's stale partials beyond
the ones the second call writes and reads), host and device results, n <= 0, and
a capture on a stream that never ran a reduction (where the allocating
entry point is refused) replayed with eager work between the replays."""
import ctypes as C

import numpy as np
import pytest

import qla_amd
from qla_amd import lattice
from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu


def _field(n, seed):
    return dev(lattice.random_field("D", n, seed).astype(np.float32))


def test_noalloc_matches_allocating_entry_point():
    import torch
    lib = qla_amd.load()
    for k, n in enumerate((300_000, 2_000)):
        a = _field(n, 90 + k)
        ref = torch.full((1,), -7.0, dtype=torch.float64, device="cuda")
        out = torch.full((1,), -7.0, dtype=torch.float64, device="cuda")
        say(f"QLA_DF3_r_veq_norm2_D n={n}")
        lib.QLA_DF3_r_veq_norm2_D(ref.data_ptr(), a.data_ptr(), n)
        qla_amd.check()
        before = lib.QLA_MI355X_scratch_buffers()
        say(f"QLA_MI355X_DF3_r_veq_norm2_D_noalloc n={n} device result")
        lib.QLA_MI355X_DF3_r_veq_norm2_D_noalloc(out.data_ptr(), a.data_ptr(), n)
        qla_amd.check()
        hres = C.c_double(-7.0)
        say(f"QLA_MI355X_DF3_r_veq_norm2_D_noalloc n={n} host result")
        lib.QLA_MI355X_DF3_r_veq_norm2_D_noalloc(C.addressof(hres), a.data_ptr(), n)
        qla_amd.check()
        assert lib.QLA_MI355X_scratch_buffers() == before
        r, o = host(ref), host(out)
        assert r.view(np.uint64)[0] == o.view(np.uint64)[0], (n, r[0], o[0])
        assert hres.value == r[0]
    hres = C.c_double(-7.0)
    say("noalloc n=0")
    lib.QLA_MI355X_DF3_r_veq_norm2_D_noalloc(C.addressof(hres), a.data_ptr(), 0)
    qla_amd.check()
    assert hres.value == 0.0


def test_noalloc_captures_on_a_fresh_stream():
    import torch
    lib = qla_amd.load()
    sizes = (2_000, 300_000)
    fields = {n: _field(n, 95) for n in sizes}
    st = torch.cuda.Stream()
    out = torch.zeros(len(sizes), dtype=torch.float64, device="cuda")
    before = lib.QLA_MI355X_scratch_buffers()
    qla_amd.check()
    g = torch.cuda.CUDAGraph()
    say("capture of the noalloc norm2 on a fresh stream")
    qla_amd.set_sync(False)
    try:
        with torch.cuda.graph(g, stream=st, capture_error_mode="thread_local"):
            qla_amd.set_stream(torch.cuda.current_stream())
            for j, n in enumerate(sizes):
                lib.QLA_MI355X_DF3_r_veq_norm2_D_noalloc(out[j:j + 1].data_ptr(), fields[n].data_ptr(), n)
            qla_amd.set_stream(None)
    finally:
        qla_amd.set_stream(None)
        qla_amd.set_sync(True)
    qla_amd.check()
    assert lib.QLA_MI355X_scratch_buffers() == before
    for rep in range(3):
        say(f"noalloc graph replay {rep}")
        g.replay()
        torch.cuda.synchronize()
        got = host(out).copy()
        for j, n in enumerate(sizes):
            e = C.c_double()
            lib.QLA_DF3_r_veq_norm2_D(C.addressof(e), fields[n].data_ptr(), n)
            qla_amd.check()
            assert got[j] == e.value, (rep, n, got[j], e.value)
            fields[n].mul_(2.0)
