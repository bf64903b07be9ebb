"""Benchmark of the QLA hot path on one MI355X.

Workload (the metric's configuration): H_veq_spproj_M_times_pD on a 32^4
periodic lattice (1,048,576 sites), plain FP32 (QLA_F3 symbol through the C
ABI), for the 8 (mu, sign) combinations mu = 0..3, sign = +-1, with pointer
arrays to forward / backward neighbours.  One step = those 8 calls.  Inputs
are resident in HBM before timing; three independent input/output sets
(3 x 235 MB) rotate between calls so the 256 MB Infinity Cache cannot serve
them.

Reported (one JSON line): value = algorithmic GB/s over the timed steps
(224 B/site per call: M 72 + D 96 + pointer 8 + H 48), ms_per_step, sites/s,
the roofline of the spproj kernel (per-launch HIP-event durations on the
library's stream), and the CPU baseline: the oracle (the C restatement of the
reference) timed on this host's cores on a bounded sample of the same calls.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
BYTES_PER_SITE = 72 + 96 + 8 + 48   # M + gathered D + pointer + H (FP32)
DIMS = (32, 32, 32, 32)
NSETS = 3


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="approximate CPU work for the oracle baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="also run 5 device-to-device copies of this many MiB (HBM counter calibration)")
    return p.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design in DESIGN.md, not a measured path)", file=sys.stderr)
        sys.exit(2)
    import torch
    import qla_amd
    from qla_amd import lattice

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    lib = qla_amd.load()
    n = lattice.volume(DIMS)
    fw, bw = lattice.neighbours(DIMS)
    sets = []
    for k in range(NSETS):
        u = torch.from_numpy(lattice.random_su3(n, 1 + 10 * k).astype(np.float32)).cuda()
        d = torch.from_numpy(lattice.random_field("D", n, 2 + 10 * k).astype(np.float32)).cuda()
        h = torch.zeros((n, 3, 2, 2), dtype=torch.float32, device="cuda")
        ptr = {(mu, s): qla_amd.pointer_array(d, (fw if s == 1 else bw)[mu]) for mu in range(4) for s in (1, -1)}
        sets.append((u, d, h, ptr))
    combos = [(mu, s) for mu in range(4) for s in (1, -1)]
    stream = torch.cuda.current_stream()
    qla_amd.set_stream(stream)
    qla_amd.set_sync(False)
    f = lib.QLA_F3_H_veq_spproj_M_times_pD

    def step(it):
        for j, (mu, s) in enumerate(combos):
            u, d, h, ptr = sets[(it * len(combos) + j) % NSETS]
            f(h.data_ptr(), u.data_ptr(), ptr[(mu, s)].data_ptr(), mu, s, n)

    for it in range(args.warmup):
        step(it)
    torch.cuda.synchronize()
    qla_amd.check()

    copy_gbs = None
    if args.copy_mib:
        src = torch.empty(args.copy_mib << 18, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record(stream)
        for _ in range(5):
            dst.copy_(src)
        c1.record(stream)
        torch.cuda.synchronize()
        copy_gbs = 2 * 5 * src.numel() * 4 / (c0.elapsed_time(c1) * 1e-3) / 1e9
        del src, dst

    # whole-job timing
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    t0 = time.perf_counter()
    for it in range(args.steps):
        step(it)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    qla_amd.check()
    ms_total = ev0.elapsed_time(ev1)
    ms_per_step = ms_total / args.steps
    calls = args.steps * len(combos)
    alg_bytes = float(BYTES_PER_SITE) * n * calls
    gbs = alg_bytes / (ms_total * 1e-3) / 1e9
    sites_per_s = n * calls / (ms_total * 1e-3)

    # the timed region holds only back-to-back spproj launches on one stream,
    # so its length over the launch count is the kernel's average duration
    # (rocprofv3 --kernel-trace on the same command agrees, see profiles/)
    avg_ms = ms_total / calls
    achieved = BYTES_PER_SITE * n / (avg_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_seconds)

    line = {
        "metric": "GB/s (%HBM peak), sites/s: spproj*M*pD 32^4, plain FP32 vs BF16 links",
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (PCG64-seeded SU(3) links, Gaussian fermions)",
        "config": {"workload": "QLA_F3_H_veq_spproj_M_times_pD, 32^4 periodic, mu 0..3 x sign +-1 per step",
                   "sites": n, "bytes_per_site": BYTES_PER_SITE, "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                   "sites_per_s": round(sites_per_s, 1), "wall_s": round(wall, 4),
                   "copy_gbs": None if copy_gbs is None else round(copy_gbs, 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                     "kernel": "qla::k_spproj<float,...>", "avg_launch_ms": round(avg_ms, 5),
                     "alg_bytes_per_launch": BYTES_PER_SITE * n},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def measured_traffic():
    """HBM bytes per spproj launch from the committed rocprofv3 counter runs
    (FETCH_SIZE and WRITE_SIZE, each scaled by its copy-kernel calibration)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "spproj_f3_32x4_traffic.json")
    try:
        with open(path) as fh:
            return json.load(fh)["traffic_bytes"]
    except (OSError, ValueError, KeyError):
        return None


def cpu_baseline(seconds: float):
    """Oracle (C restatement, OpenMP over the host's cores) on a bounded sample:
    repeated full-lattice H_veq_spproj_M_times_pD calls for the same 8 combos."""
    import oracle
    from qla_amd import lattice
    lib = oracle.load()
    threads = oracle.threads()
    n = lattice.volume(DIMS)
    fw, bw = lattice.neighbours(DIMS)
    u = lattice.random_su3(n, 1).astype(np.float32)
    d = lattice.random_field("D", n, 2).astype(np.float32)
    h = np.zeros((n, 3, 2, 2), dtype=np.float32)
    combos = [(mu, s) for mu in range(4) for s in (1, -1)]
    ptrs = {(mu, s): oracle.pointer_array(d, (fw if s == 1 else bw)[mu]) for mu, s in combos}
    oracle.H_eqop_spproj_M_times_pD(oracle.EQ, h, u, d, fw[0], 0, 1)   # warm-up
    calls = 0
    t0 = time.perf_counter()
    while True:
        for mu, s in combos:
            lib.orc_F3_H_eqop_spproj_M_times_pD(0, 0, h.ctypes.data, u.ctypes.data, ptrs[(mu, s)].ctypes.data,
                                                mu, s, None, n)
            calls += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    gbs = BYTES_PER_SITE * n * calls / dt / 1e9
    return {"value": round(gbs, 2), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{calls} full 32^4 F3 spproj calls (8 mu/sign combos) in {dt:.1f} s, "
                      f"OpenMP {threads} threads, sites/s {n * calls / dt:.3e}"}


if __name__ == "__main__":
    main()
