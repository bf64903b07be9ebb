"""CPU baseline of the benchmark: the oracle (the C restatement of the
reference, built with the flags of a reference configured --enable-sse3
--enable-openmp: -O3 -msse3 -fopenmp -ffp-contract=off) timed on this
host's cores.  Test infrastructure, run by bench.py's CPU-baseline leg in a
child process, so that the OpenMP settings below are in place before the
OpenMP runtime starts:

    python -m oracle.cpu_bench [--seconds S] [--l L]

OMP_NUM_THREADS = the physical cores this process may use (the smaller of
lscpu's cores and the CPU affinity, capped by an OMP_NUM_THREADS already set
by the environment), OMP_PROC_BIND=close, OMP_PLACES=cores.

Workload: the benchmark's own calls on an L^4 periodic lattice (L = 32), with
the link field a contiguous field and the gathered records through QLA
pointer arrays, as the v-form QLA calls take them, each site loop with mu and
sign constant as in the generated code (orc_F3_bench_*):
  * QLA_F3_H_veq_spproj_M_times_pD, the 8 (mu, sign) combinations with mu
    0..3 (generic C order: the reference has no SSE version of it);
  * QLA_F3_D_vpeq_sprecon_M_times_pH, the same combinations, in the SSE
    build's operation order (the reference's lib/optimized/sse version).
One repetition is the 8 calls; after one warm-up repetition, repetitions
run until at least 10 are done and S/2 seconds have passed for each op; the
median repetition gives the rate.  Algorithmic bytes per site: 224 (M 72 +
D 96 + pointer 8 + H 48) and 320 (M 72 + H 48 + pointer 8 + D read and
written 192).  Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time


def lscpu() -> dict:
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            info[k.strip()] = v.strip()
    return info


def physical_cores(info: dict) -> int:
    try:
        return int(info["Core(s) per socket"]) * int(info["Socket(s)"])
    except (KeyError, ValueError):
        return os.cpu_count() or 1


def omp_env() -> dict:
    """The environment of the timed child: thread count and placement."""
    info = lscpu()
    cores = physical_cores(info)
    threads = min(cores, len(os.sched_getaffinity(0)))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    env = dict(os.environ)
    env.update(OMP_NUM_THREADS=str(max(1, threads)), OMP_PROC_BIND="close", OMP_PLACES="cores")
    return env


def run(seconds: float, L: int) -> dict:
    import ctypes as C

    import numpy as np

    import oracle
    from qla_amd import lattice

    lib = oracle.load()
    vp, i32 = C.c_void_p, C.c_int
    for f in (lib.orc_F3_bench_spproj, lib.orc_F3_bench_sprecon_sse):
        f.argtypes = [vp, vp, vp, i32, i32, i32]
    dims = (L, L, L, L)
    n = lattice.volume(dims)
    fw, bw = lattice.neighbours(dims)
    combos = [(mu, s) for mu in range(4) for s in (1, -1)]
    u = lattice.random_su3(n, 1).astype(np.float32)
    d = lattice.random_field("D", n, 2).astype(np.float32)
    hsrc = lattice.random_field("H", n, 5).astype(np.float32)
    h = np.zeros((n, 3, 2, 2), dtype=np.float32)
    dst = np.zeros((n, 3, 4, 2), dtype=np.float32)
    dptr = {(mu, s): oracle.pointer_array(d, (fw if s == 1 else bw)[mu]) for mu, s in combos}
    hptr = {(mu, s): oracle.pointer_array(hsrc, (fw if s == 1 else bw)[mu]) for mu, s in combos}

    def spproj_rep():
        for mu, s in combos:
            lib.orc_F3_bench_spproj(h.ctypes.data, u.ctypes.data, dptr[(mu, s)].ctypes.data, mu, s, n)

    def sprecon_rep():
        for mu, s in combos:
            lib.orc_F3_bench_sprecon_sse(dst.ctypes.data, u.ctypes.data, hptr[(mu, s)].ctypes.data, mu, s, n)

    def timed(rep):
        rep()
        times = []
        t_end = time.perf_counter() + seconds / 2
        while len(times) < 10 or time.perf_counter() < t_end:
            t0 = time.perf_counter()
            rep()
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), len(times)

    t_sp, reps_sp = timed(spproj_rep)
    t_sr, reps_sr = timed(sprecon_rep)
    calls = len(combos)
    info = lscpu()
    return {
        "threads": oracle.threads(), "model": info.get("Model name", "?"),
        "physical_cores": physical_cores(info), "logical_cpus": int(info.get("CPU(s)", "0") or 0),
        "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
        "lattice": f"{L}^4", "sites": n,
        "spproj": {"median_rep_s": t_sp, "reps": reps_sp, "GBs": 224 * n * calls / t_sp / 1e9,
                   "sites_per_s": n * calls / t_sp, "order": "generic C (no SSE version exists)"},
        "sprecon_sse_order": {"median_rep_s": t_sr, "reps": reps_sr, "GBs": 320 * n * calls / t_sr / 1e9,
                              "sites_per_s": n * calls / t_sr, "order": "SSE build"},
    }


def main() -> None:
    p = argparse.ArgumentParser()
    p.add_argument("--seconds", type=float, default=10.0)
    p.add_argument("--l", type=int, default=32)
    p.add_argument("--child", action="store_true", help=argparse.SUPPRESS)
    a = p.parse_args()
    if not a.child:
        # start the timed run with its OpenMP environment in place
        cmd = [sys.executable, "-m", "oracle.cpu_bench", "--child", "--seconds", str(a.seconds), "--l", str(a.l)]
        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        res = subprocess.run(cmd, env=omp_env(), cwd=root)
        sys.exit(res.returncode)
    print(json.dumps(run(a.seconds, a.l)))


if __name__ == "__main__":
    main()
