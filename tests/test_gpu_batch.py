"""GPU tests of the ragged batch extension: one persistent launch over a
table of independent QLA_F3 v-form calls.

Bar: every job's destination is bit-identical to the same job run as its own
plain QLA call (the plain kernels are themselves bit-exact with the oracle;
one job per op is also checked against the oracle directly).  The captured
hipGraph is replayed several times with eager work in between and every
replay is checked; the autotuner's choice must not change results.
"""
import numpy as np
import pytest

import qla_amd
from qla_amd import batch as qb
from gpu_util import assert_bitwise, host, say

pytestmark = pytest.mark.gpu


def snapshot(keep):
    return [host(kp["dst"]).copy() for kp in keep]


def expected(keep):
    qb.reset(keep)
    qb.run_sequential(keep)
    out = snapshot(keep)
    qb.reset(keep)
    return out


def check(keep, exp, what):
    got = snapshot(keep)
    for k, (g, e) in enumerate(zip(got, exp)):
        assert_bitwise(g, e, f"{what} job {k} ({qb.OP_NAMES[keep[k]['op']]})")


def test_batch_matches_plain_calls(orc):
    desc = qb.draw_jobs(24, seed=7, Ls=(2, 4, 6, 8))
    jobs, keep = qb.checkerboard_jobs(desc)
    exp = expected(keep)
    b = qb.Batch(jobs)
    say(f"batch run shape={b.shape()}")
    b.run()
    qla_amd.check()
    check(keep, exp, "batch run")
    # every launch shape gives the same bits
    for shape in ((1, 1, 2), (4, 8, 16), (2, 3, 1), (2, 1, 0), (4, 3, 0), (1, 8, 0)):
        qb.reset(keep)
        b.set_shape(*shape)
        say(f"batch run shape={shape}")
        b.run()
        check(keep, exp, f"shape {shape}")
    b.close()


def test_batch_against_oracle(orc):
    """One job of each op compared with the oracle directly."""
    desc = [dict(L=4, parity=0, op=0, mu=1, sign=1), dict(L=6, parity=1, op=1, mu=2, sign=-1),
            dict(L=4, parity=1, op=2, mu=3, sign=1)]
    jobs, keep = qb.checkerboard_jobs(desc, seed=3)
    b = qb.Batch(jobs)
    say("batch vs oracle")
    b.run()
    qla_amd.check()
    for kp in keep:
        u, src, d0 = host(kp["u"]), host(kp["src"]), host(kp["dst0"])
        # pointer array back to source-record indices
        rec = kp["src"].stride(0) * 4
        idx = ((host(kp["ptr"]) - kp["src"].data_ptr()) // rec).astype(np.int64)
        exp = d0.copy()
        if kp["op"] == 0:
            orc.V_eqop_M_times_V(orc.PEQ, exp, u, src, b_idx=idx)
        elif kp["op"] == 1:
            exp = np.zeros_like(d0)
            orc.H_eqop_spproj_M_times_pD(orc.EQ, exp, u, src, idx, kp["mu"], kp["sign"])
        else:
            orc.D_eqop_sprecon_M_times_pH(orc.PEQ, exp, u, src, idx, kp["mu"], kp["sign"])
        assert_bitwise(host(kp["dst"]), exp, f"oracle op {kp['op']}")
    b.close()


def test_batch_graph_replays():
    import torch
    desc = qb.draw_jobs(16, seed=9, Ls=(2, 4, 6))
    jobs, keep = qb.checkerboard_jobs(desc, seed=5)
    exp = expected(keep)
    b = qb.Batch(jobs)
    say("batch capture")
    b.capture()
    for rep in range(3):
        qb.reset(keep)
        torch.cuda.synchronize()
        junk = torch.randn(1 << 16, device="cuda").sum()     # eager work between replays
        say(f"batch graph replay {rep}")
        b.launch_graph()
        qla_amd.check()
        check(keep, exp, f"graph replay {rep}")
        del junk
    b.close()


def test_batch_autotune_keeps_results():
    desc = qb.draw_jobs(12, seed=13, Ls=(4, 6, 8))
    jobs, keep = qb.checkerboard_jobs(desc, seed=8)
    exp = expected(keep)
    b = qb.Batch(jobs)
    say("batch autotune")
    ms = b.autotune(reps=1)
    qla_amd.check()
    assert ms > 0
    wpb, tpc, bpc = b.shape()
    assert wpb in (1, 2, 4) and tpc >= 1 and bpc >= 0
    qb.reset(keep)
    b.run()
    check(keep, exp, "after autotune")
    b.capture()
    qb.reset(keep)
    b.launch_graph()
    check(keep, exp, "graph after autotune")
    b.close()


def test_batch_noop_jobs():
    """n = 0 and mu outside 0..4 are no-ops inside a batch too."""
    desc = [dict(L=4, parity=0, op=1, mu=1, sign=1), dict(L=4, parity=1, op=2, mu=0, sign=-1)]
    jobs, keep = qb.checkerboard_jobs(desc, seed=4)
    jobs[0] = dict(jobs[0], mu=7)
    jobs[1] = dict(jobs[1], n=0)
    b = qb.Batch(jobs)
    say("batch no-op jobs")
    b.run()
    qla_amd.check()
    for kp in keep:
        assert_bitwise(host(kp["dst"]), host(kp["dst0"]), "no-op job")
    b.close()


def test_batch_many_jobs_uncached_lookup(orc):
    """A long job table (2100 jobs, binary-searched per tile): 2100 jobs that each
    take a slice of one spproj call must reproduce the single plain call."""
    import torch
    from qla_amd import lattice
    from gpu_util import dev
    dims = (8, 8, 8, 6)
    n = lattice.volume(dims)
    fw, _ = lattice.neighbours(dims)
    u = lattice.random_su3(n, 21).astype(np.float32)
    d = lattice.random_field("D", n, 22).astype(np.float32)
    du, dd = dev(u), dev(d)
    ptr = qla_amd.pointer_array(dd, fw[2])
    r_plain = torch.zeros((n, 3, 2, 2), dtype=torch.float32, device="cuda")
    qla_amd.call("QLA_F3_H_veq_spproj_M_times_pD", r_plain, du, ptr, 2, 1, n)
    r = torch.zeros_like(r_plain)
    bounds = np.linspace(0, n, 2101).astype(int)
    jobs = []
    for k in range(2100):
        s0, s1 = int(bounds[k]), int(bounds[k + 1])
        jobs.append(dict(op=qb.OP_H_VEQ_SPPROJ_M_TIMES_PD, mu=2, sign=1, n=s1 - s0,
                         r=r.data_ptr() + 48 * s0, a=du.data_ptr() + 72 * s0, b=ptr.data_ptr() + 8 * s0))
    b = qb.Batch(jobs)
    say("batch 2100 jobs")
    b.run()
    qla_amd.check()
    assert_bitwise(host(r), host(r_plain), "2100-job batch")
    b.close()
