"""GPU checks of the Cayley-Hamilton matrix functions with an FP64 coefficient
step (QLA_MI355X_F3_M_{v,x}eq_{expA,sqrtPH,invsqrtPH}_M_ch; no reference
counterpart) through the C ABI.

Bars:
* against the exact function, i.e. the all-double model of
  oracle/closed_forms.py, which the CPU tests pin against scipy's expm and
  sqrtm: the kernels evaluate the same steps in double and round each result
  component to float once, so the normwise relative difference is at most
  2^-24 plus the double evaluation's own error; the bar is 1e-7;
* the x-form writes exactly the indexed records, each equal bit for bit to the
  v-form's result for that record;
* special inputs: A = 0 (exp 1, sqrt 0, 1/sqrt +inf on the diagonal),
  multiples of 1, double eigenvalues, a rank-one positive matrix, n = 0.
"""
import numpy as np
import pytest

import qla_amd
from oracle import closed_forms as cf
from qla_amd import lattice
from gpu_util import assert_bitwise, dev, host, say

pytestmark = pytest.mark.gpu

SYM = {"expA": "expA_M", "sqrtPH": "sqrtPH_M", "invsqrtPH": "invsqrtPH_M"}
BAR = 1e-7


def c(a):
    a = np.asarray(a, dtype=np.float64)
    return a[..., 0] + 1j * a[..., 1]


def rel(got, ref):
    return np.linalg.norm(got - ref, axis=(1, 2)) / np.linalg.norm(ref, axis=(1, 2))


def inputs(fn, n, seed):
    if fn == "expA":
        a = lattice.random_antihermitian(n, 3, seed)
        # half of them with a trace (expA is exp of any anti-Hermitian matrix)
        t = np.random.default_rng(seed + 1).uniform(-2, 2, n)
        a[: n // 2, [0, 1, 2], [0, 1, 2], 1] += t[: n // 2, None]
        return a.astype(np.float32)
    return lattice.random_hpd(n, 3, seed).astype(np.float32)


def model(fn, a):
    if fn == "expA":
        return cf.expA(a)
    return cf.sqrtPH(a, inverse=fn == "invsqrtPH")


def run(fn, a, index=None, r0=None):
    n = a.shape[0] if index is None else len(index)
    da = dev(a)
    dr = dev(np.zeros_like(a) if r0 is None else r0)
    if index is None:
        qla_amd.call(f"QLA_MI355X_F3_M_veq_{SYM[fn]}_ch", dr, da, n)
    else:
        qla_amd.call(f"QLA_MI355X_F3_M_xeq_{SYM[fn]}_ch", dr, da, dev(np.asarray(index, dtype=np.int32)), n)
    qla_amd.check()
    return host(dr)


@pytest.mark.parametrize("fn", list(SYM))
@pytest.mark.parametrize("n", [333, 65536])
def test_against_model_and_exact(fn, n):
    a = inputs(fn, n, 31 + n % 7)
    say(f"ch {fn} v-form n={n}")
    got = c(run(fn, a))
    err = rel(got, c(model(fn, a)))
    assert err.max() < BAR, (fn, err.max())


@pytest.mark.parametrize("fn", list(SYM))
def test_xform(fn):
    n = 4099
    a = inputs(fn, n, 5)
    idx = np.random.default_rng(6).permutation(n)[: n // 2].astype(np.int32)
    r0 = np.full_like(a, 7.0)
    say(f"ch {fn} x-form n={len(idx)}")
    got = run(fn, a, index=idx, r0=r0)
    full = run(fn, a)
    mask = np.zeros(n, dtype=bool)
    mask[idx] = True
    assert_bitwise(got[mask], full[mask], f"{fn} x-form indexed records")
    assert np.all(got[~mask] == 7.0)


def test_special_inputs():
    eye = np.eye(3)
    u = np.linalg.qr(np.random.default_rng(2).standard_normal((3, 3)) +
                     1j * np.random.default_rng(3).standard_normal((3, 3)))[0]
    rot = lambda d: u @ np.diag(d) @ u.conj().T          # noqa: E731
    pack = lambda z: np.stack([z.real, z.imag], -1).astype(np.float32)   # noqa: E731
    # exp: zero, a multiple of i*1, double eigenvalues of both signs of det
    herm = np.array([np.zeros((3, 3)), 0.7 * eye, rot([1.0, 1.0, -2.0]), rot([-1.0, -1.0, 2.0])])
    a = pack(1j * herm)
    say("ch expA special inputs")
    got = c(run("expA", a))
    assert np.array_equal(got[0], eye)
    assert rel(got, c(cf.expA(a))).max() < BAR
    # sqrt / invsqrt: zero, a multiple of 1, a double eigenvalue, rank one
    # (diagonal, so that it stays rank one in float)
    pos = np.array([np.zeros((3, 3)), 4.0 * eye, rot([2.0, 3.0, 3.0]), np.diag([5.0, 0.0, 0.0])])
    a = pack(pos)
    say("ch sqrtPH special inputs")
    s = c(run("sqrtPH", a))
    assert np.array_equal(s[0], np.zeros((3, 3)))
    refs = [2.0 * eye, rot(np.sqrt([2.0, 3.0, 3.0])), np.diag([np.sqrt(5.0), 0.0, 0.0])]
    for i, ref in enumerate(refs):
        assert np.abs(s[i + 1] - ref).max() < 1e-5, i
    say("ch invsqrtPH special inputs")
    v = c(run("invsqrtPH", a[:3]))
    assert np.all(np.isinf(np.diag(v[0]).real)) and np.all(v[0][~np.eye(3, dtype=bool)] == 0)
    assert np.abs(v[1] - 0.5 * eye).max() < 1e-6
    assert np.abs(v[2] - rot(1 / np.sqrt([2.0, 3.0, 3.0]))).max() < 1e-5


def test_nan_propagates():
    """A NaN anywhere in the input gives NaN results, not a wrong finite matrix."""
    for fn in SYM:
        a = inputs(fn, 64, 3)
        a[5, 1, 2, 0] = np.nan
        say(f"ch {fn} NaN input")
        got = run(fn, a)
        assert np.isnan(got[5]).any(), fn
        clean = np.ones(64, dtype=bool)
        clean[5] = False
        assert np.isfinite(got[clean]).all(), fn


def test_n_zero_is_noop():
    a = inputs("expA", 8, 1)
    r0 = np.full_like(a, 3.0)
    for fn in SYM:
        say(f"ch {fn} n=0")
        dr = dev(r0)
        qla_amd.call(f"QLA_MI355X_F3_M_veq_{SYM[fn]}_ch", dr, dev(a), 0)
        qla_amd.check()
        assert np.all(host(dr) == 3.0)


@pytest.mark.parametrize("fn", list(SYM))
def test_full_size_config4(fn):
    """2,097,152 matrices of the config-4 distributions against the exact
    function (the all-double model)."""
    n = 2097152
    a = lattice.random_antihermitian(n, 3, 41) if fn == "expA" else lattice.random_hpd(n, 3, 42)
    a = a.astype(np.float32)
    say(f"ch {fn} full size n={n}")
    got = c(run(fn, a))
    err = rel(got, c(model(fn, a)))
    print(f"{fn}: max normwise error {err.max():.3g}, mean {err.mean():.3g}")
    assert err.max() < BAR, (fn, err.max())
