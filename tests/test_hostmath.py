"""The single-precision acosf, cbrtf and sincosf restated from the host C
library for the device (qla_amd/csrc/qla_hostmath.hpp) against the library
itself: the same header compiled for the host must return the library's
bits on a strided sample of every input domain and on the boundary values
of each algorithm (|x| = 0.5, 1, 2^-57 for acosf; pi/4, 2^-12, 120 for
sincosf; zeros, subnormals and infinities).  The F3 closed forms on the GPU
(expA, expTA, sqrtPH, invsqrtPH, eigenvalsH) then match the reference's
results bit for bit (tests/test_gpu_matfunc.py).

The library picks its sincosf build by processor: the one with fused
multiply-adds on processors with AVX2 and FMA, which the restatement
follows; elsewhere the sincosf comparison is skipped."""
import pathlib
import shutil
import subprocess

import pytest

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE.parent / "qla_amd" / "csrc"


def _cpu_has_fma():
    try:
        flags = pathlib.Path("/proc/cpuinfo").read_text()
    except OSError:
        return False
    return " fma" in flags and " avx2" in flags


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path_factory.mktemp("hostmath") / "hostmath_check"
    subprocess.run([cxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", str(CSRC),
                    str(HERE / "hostmath_check.cpp"), "-o", str(exe), "-lm"], check=True)
    return exe


def test_restated_functions_match_the_library(checker):
    sc = _cpu_has_fma()
    out = subprocess.run([str(checker), "97", "1" if sc else "0"], check=True, capture_output=True,
                         text=True).stdout.strip().splitlines()
    last = out[-1].split()
    counts = dict(zip(last[0::2], (int(v) for v in last[1::2])))
    assert counts["checked"] > 4 * 10 ** 7
    assert counts["acosf"] == 0, out
    assert counts["cbrtf"] == 0, out
    if sc:
        assert counts["sincosf"] == 0, out
