"""Build the gfx950 shared library ``qla_amd/libqla_mi355x.so`` in-tree.

The library is plain HIP C++ compiled by ``hipcc`` for gfx950 only; it exports
the QLA C ABI declared in ``include/qla_mi355x.h`` and links nothing but the
HIP runtime.  Every source is compiled with ``-ffp-contract=off`` (the
reference's complex macros are evaluated without fused multiply-adds) and
with packed FP32 instructions disabled.

Usage: ``python -m qla_amd.build`` (or ``qla_amd.build.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
LIB = HERE / "libqla_mi355x.so"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CXXFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-I", str(REPO / "include"),
]


# Sources compiled more than once, each time with its own defines: the site
# kernels once per (precision, kind), the multi-source kernels and the field
# linear algebra once per precision and the matrix functions once per
# (precision, nc), so the kernel instantiations spread over parallel compiles.
MULTI_BUILD = {
    "qla_site.hip": [{"QLA_SITE_PREC": p, "QLA_SITE_KIND": k} for p in (0, 1) for k in (0, 1, 2, 3, 5, 6, 7, 8, 9, 10, 11)],
    "qla_multi.hip": [{"QLA_SITE_PREC": p} for p in (0, 1)],
    "qla_matfn.hip": [{"QLA_MATFN_UNIT": u} for u in range(9)],
    "qla_vecops.hip": [{"QLA_VEC_PREC": p} for p in (0, 1)],
}


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def units() -> list[tuple[Path, dict]]:
    """(source, defines) of every object of the library."""
    out = []
    for src in sources():
        for defs in MULTI_BUILD.get(src.name, [{}]):
            out.append((src, defs))
    return out


def _compile(unit: tuple[Path, dict]) -> Path:
    src, defs = unit
    OBJ.mkdir(exist_ok=True)
    tag = "".join(f"_{k.rsplit('_', 1)[-1].lower()}{v}" for k, v in defs.items())
    obj = OBJ / (src.stem + tag + ".o")
    cmd = [HIPCC, *CXXFLAGS, *[f"-D{k}={v}" for k, v in defs.items()], "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every kernel source for gfx950 and link the shared library."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    us = units()
    jobs = jobs or min(len(us), os.cpu_count() or 8, 16)
    # longest compiles first
    us.sort(key=lambda u: 0 if u[1] else 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, us))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
