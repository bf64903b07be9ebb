"""Helpers for the GPU parity tests: device copies, bitwise comparison and
seeded inputs (the oracle is the checker; nothing here computes results)."""
from __future__ import annotations

import sys

import numpy as np

from qla_amd import lattice


def dev(a: np.ndarray):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t) -> np.ndarray:
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def bits(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint64)


def assert_bitwise(got: np.ndarray, exp: np.ndarray, what: str = ""):
    gb, eb = bits(got), bits(exp)
    if not np.array_equal(gb, eb):
        bad = np.nonzero(gb.reshape(-1) != eb.reshape(-1))[0]
        g = got.reshape(-1)[bad[:5]]
        e = exp.reshape(-1)[bad[:5]]
        raise AssertionError(f"{what}: {bad.size} of {gb.size} components differ; first at {bad[:5]}: "
                             f"got {g} expected {e}")


def say(msg: str) -> None:
    """Name the case before a GPU step (flushed, so a fault still shows it)."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def inputs(dims, prec, seed=11):
    """SU(3) links per site, D/H/V Gaussian fields, in numpy of `prec`."""
    n = lattice.volume(dims)
    dt = np.float32 if prec == "F" else np.float64
    u = lattice.random_su3(n, seed).astype(dt)
    d = lattice.random_field("D", n, seed + 1).astype(dt)
    h = lattice.random_field("H", n, seed + 2).astype(dt)
    v = lattice.random_field("V", n, seed + 3).astype(dt)
    return u, d, h, v
