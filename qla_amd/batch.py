"""Ragged batches of QLA_F3 v-form calls (the batch extension of the C ABI).

``Batch(jobs)`` wraps a plan: a table of independent calls of
V_vpeq_M_times_pV, H_veq_spproj_M_times_pD or D_vpeq_sprecon_M_times_pH run
in one launch (optionally replayed from a captured hipGraph).

``checkerboard_jobs`` builds the even/odd-checkerboard workload used to
exercise and measure it: each job is one sub-lattice L^4, its fields stored
in checkerboard order (even sites, then odd), the destination and the links
are the job's parity half, and the pointer array points each site at its
mu-neighbour (opposite parity) in the full source field.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lattice
from .lib import F3Job, load

OP_V_VPEQ_M_TIMES_PV = 0
OP_H_VEQ_SPPROJ_M_TIMES_PD = 1
OP_D_VPEQ_SPRECON_M_TIMES_PH = 2
OP_NAMES = {0: "QLA_F3_V_vpeq_M_times_pV", 1: "QLA_F3_H_veq_spproj_M_times_pD",
            2: "QLA_F3_D_vpeq_sprecon_M_times_pH"}
SRC_KIND = {0: "V", 1: "D", 2: "H"}
DST_KIND = {0: "V", 1: "H", 2: "D"}
SHAPE = {"V": (3, 2), "H": (3, 2, 2), "D": (3, 4, 2)}
# algorithmic bytes per site (FP32): matvec pV 152, spproj 224, sprecon 320
BYTES = {0: 152, 1: 224, 2: 320}


class Batch:
    def __init__(self, jobs: list[dict]):
        self.lib = load()
        self.njobs = len(jobs)
        arr = (F3Job * self.njobs)()
        for k, j in enumerate(jobs):
            arr[k] = F3Job(j["op"], j.get("mu", 0), j.get("sign", 1), j["n"], j["r"], j["a"], j["b"])
        self._arr = arr
        self.plan = self.lib.QLA_MI355X_F3_batch_create(C.cast(arr, C.c_void_p), self.njobs)
        if not self.plan:
            raise RuntimeError("QLA_MI355X_F3_batch_create failed")

    def run(self):
        self.lib.QLA_MI355X_batch_run(self.plan)

    def capture(self):
        rc = self.lib.QLA_MI355X_batch_capture(self.plan)
        if rc != 0:
            raise RuntimeError(f"QLA_MI355X_batch_capture failed ({rc})")

    def launch_graph(self):
        self.lib.QLA_MI355X_batch_launch_graph(self.plan)

    def autotune(self, reps: int = 3) -> float:
        return self.lib.QLA_MI355X_batch_autotune(self.plan, reps)

    def shape(self):
        w, t, b = C.c_int(), C.c_int(), C.c_int()
        self.lib.QLA_MI355X_batch_get_shape(self.plan, C.byref(w), C.byref(t), C.byref(b))
        return w.value, t.value, b.value

    def set_shape(self, wpb: int, tpc: int, bpc: int):
        if self.lib.QLA_MI355X_batch_set_shape(self.plan, wpb, tpc, bpc) != 0:
            raise ValueError((wpb, tpc, bpc))

    def close(self):
        if self.plan:
            self.lib.QLA_MI355X_batch_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def draw_jobs(njobs: int, seed: int, Ls=(8, 12, 16, 20, 24, 32)):
    """Seeded job descriptions: L, parity, op, mu, sign for each job."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return [dict(L=int(rng.choice(Ls)), parity=int(rng.integers(2)), op=int(rng.integers(3)),
                 mu=int(rng.integers(4)), sign=int(rng.choice([1, -1]))) for _ in range(njobs)]


def checkerboard_jobs(desc: list[dict], seed: int = 11):
    """Device fields and job dicts for `desc`; returns (jobs, keep) where keep
    holds the tensors (the plan references them) and per-job host metadata."""
    import torch
    jobs, keep = [], []
    for k, dj in enumerate(desc):
        L = dj["L"]
        dims = (L, L, L, L)
        vol = lattice.volume(dims)
        half = vol // 2
        par = lattice.coords(dims).sum(axis=1) % 2
        cb = np.concatenate([np.nonzero(par == 0)[0], np.nonzero(par == 1)[0]])
        pos = np.empty(vol, dtype=np.int64)
        pos[cb] = np.arange(vol)
        fw, bw = lattice.neighbours(dims)
        sites = cb[dj["parity"] * half:(dj["parity"] + 1) * half]
        nbr = (fw if dj["sign"] == 1 else bw)[dj["mu"]][sites]
        op = dj["op"]
        u = lattice.random_su3(half, seed + 97 * k).astype(np.float32)
        src = lattice.random_field(SRC_KIND[op], vol, seed + 97 * k + 1).astype(np.float32)
        dst = lattice.random_field(DST_KIND[op], half, seed + 97 * k + 2).astype(np.float32)
        du, dsrc, ddst = (torch.from_numpy(x).cuda() for x in (u, src, dst))
        rec = dsrc.stride(0) * 4
        ptr = dsrc.data_ptr() + rec * torch.from_numpy(pos[nbr]).cuda()
        jobs.append(dict(op=op, mu=dj["mu"], sign=dj["sign"], n=half, r=ddst.data_ptr(), a=du.data_ptr(),
                         b=ptr.data_ptr()))
        keep.append(dict(u=du, src=dsrc, dst=ddst, ptr=ptr, dst0=ddst.clone(), n=half, op=op,
                         mu=dj["mu"], sign=dj["sign"]))
    return jobs, keep


def run_sequential(keep):
    """The same jobs as 64 plain QLA calls, on the thread's stream."""
    lib = load()
    for kp in keep:
        name = OP_NAMES[kp["op"]]
        if kp["op"] == OP_V_VPEQ_M_TIMES_PV:
            getattr(lib, name)(kp["dst"].data_ptr(), kp["u"].data_ptr(), kp["ptr"].data_ptr(), kp["n"])
        else:
            getattr(lib, name)(kp["dst"].data_ptr(), kp["u"].data_ptr(), kp["ptr"].data_ptr(), kp["mu"],
                               kp["sign"], kp["n"])


def reset(keep):
    for kp in keep:
        kp["dst"].copy_(kp["dst0"])
