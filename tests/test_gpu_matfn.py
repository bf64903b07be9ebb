"""GPU parity of the batched matrix functions (config-4 inputs, reduced n).

Bar: bit-exact against the oracle, which restates the reference templates
(lib/perl/code-templates/M_eq_{exp,sqrt,invsqrt}_M.c) in the same operation
order; the kernels use IEEE-rounded divide and square root and no
contraction, so each of the n matrices must agree bit for bit.  Inputs:
exp on i*H (H Hermitian traceless, spectrum within [-4, 4]), sqrt/invsqrt on
B^dagger B + eps*1 with eps log-uniform in [1e-4, 1]; plus the templates'
special cases (A = 0, the s = 1 branch, the s <= 1024 cap).
"""
import numpy as np
import pytest

import qla_amd
from qla_amd import lattice
from gpu_util import assert_bitwise, dev, host, say

pytestmark = pytest.mark.gpu


def run(name, nc, a, fn_args_nc):
    r = dev(np.zeros_like(a))
    if fn_args_nc:
        qla_amd.call(name, nc, r, dev(a), a.shape[0])
    else:
        qla_amd.call(name, r, dev(a), a.shape[0])
    qla_amd.check()
    return host(r)


@pytest.mark.parametrize("fn", ["exp", "sqrt", "invsqrt"])
@pytest.mark.parametrize("prec", ["F", "D"])
def test_matfn_nc3(orc, fn, prec):
    n = 4099
    dt = np.float32 if prec == "F" else np.float64
    gen = lattice.random_antihermitian if fn == "exp" else lattice.random_hpd
    a = gen(n, 3, 300).astype(dt)
    name = f"QLA_{prec}3_M_veq_{fn}_M"
    say(f"{name} n={n}")
    assert_bitwise(run(name, 3, a, False), orc.matfn(fn, a), name)


@pytest.mark.parametrize("fn", ["exp", "sqrt", "invsqrt"])
@pytest.mark.parametrize("nc", [1, 2, 3, 4, 5, 6, 7, 8])
def test_matfn_fn(orc, fn, nc):
    """Bit-exact at every nc; at nc = 1, 2 exp takes the template's closed
    form, whose expf and sinhf the device and the host library may round
    apart, so there it agrees to 8 ulp of the larger of 1 and the matrix's
    largest component."""
    n = 777
    gen = lattice.random_antihermitian if fn == "exp" else lattice.random_hpd
    a = gen(n, nc, 400 + nc).astype(np.float32)
    name = f"QLA_FN_M_veq_{fn}_M"
    say(f"{name} nc={nc} n={n}")
    got, exp = run(name, nc, a, True), orc.matfn(fn, a)
    if fn == "exp" and nc <= 2:
        scale = np.maximum(np.abs(exp).reshape(n, -1).max(axis=1), 1.0)
        err = np.abs(got - exp).reshape(n, -1).max(axis=1)
        assert (err <= 8 * np.finfo(np.float32).eps * scale).all(), f"{name} nc={nc}"
    else:
        assert_bitwise(got, exp, f"{name} nc={nc}")


@pytest.mark.parametrize("prec", ["F", "D"])
def test_matfn_special_cases(orc, prec):
    dt = np.float32 if prec == "F" else np.float64
    z = np.zeros((5, 3, 3, 2), dtype=dt)
    small = lattice.random_antihermitian(70, 3, 5, radius=0.2).astype(dt)
    big = lattice.random_antihermitian(70, 3, 6, radius=700.0)
    big = (big * (700.0 / np.sqrt((big ** 2).sum(axis=(1, 2, 3))))[:, None, None, None]).astype(dt)
    for fn, a in (("exp", z), ("exp", small), ("exp", big), ("sqrt", z), ("invsqrt", z)):
        name = f"QLA_{prec}3_M_veq_{fn}_M"
        say(f"{name} special n={a.shape[0]}")
        assert_bitwise(run(name, 3, a, False), orc.matfn(fn, a), f"{name} special")


def test_matfn_bad_nc_is_error():
    """nc outside the supported range leaves r untouched and records an error."""
    for nc in (0, 9):
        a = np.ones((4, max(nc, 1), max(nc, 1), 2), dtype=np.float32)
        r = dev(a)
        say(f"QLA_FN_M_veq_exp_M nc={nc} (unsupported)")
        qla_amd.call("QLA_FN_M_veq_exp_M", nc, r, dev(a), 4)
        assert qla_amd.load().QLA_MI355X_last_error() != 0
        np.testing.assert_array_equal(host(r), a)


def general(n, nc, seed, fn, dt):
    """P D P^-1 with a Gaussian complex P and Gaussian complex eigenvalues
    (Re <= 0 for the roots, imaginary for exp), the construction of the
    reference's own tests (protocol_tensor_sng.m4:481-596): general,
    non-normal matrices."""
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((n, nc, nc)) + 1j * rng.standard_normal((n, nc, nc))
    d = rng.standard_normal((n, nc)) + 1j * rng.standard_normal((n, nc))
    d = 1j * d.real if fn == "exp" else np.where(d.real > 0, -d, d)
    m = p @ (d[:, :, None] * np.linalg.inv(p))
    return np.stack([m.real, m.imag], axis=-1).astype(dt)


@pytest.mark.parametrize("fn", ["exp", "sqrt", "invsqrt"])
@pytest.mark.parametrize("lib", ["F3", "D3", "FN4", "FN8"])
def test_matfn_general_matrices(orc, fn, lib):
    """Non-Hermitian inputs: bit-exact against the oracle (the kernels take
    no shortcut that assumes Hermitian or anti-Hermitian structure)."""
    n = 1031
    nc = 3 if lib in ("F3", "D3") else int(lib[2:])
    dt = np.float64 if lib == "D3" else np.float32
    a = general(n, nc, 500 + nc, fn, dt)
    if lib.startswith("FN"):
        name = f"QLA_FN_M_veq_{fn}_M"
        say(f"{name} nc={nc} general n={n}")
        got = run(name, nc, a, True)
    else:
        name = f"QLA_{lib}_M_veq_{fn}_M"
        say(f"{name} general n={n}")
        got = run(name, 3, a, False)
    assert_bitwise(got, orc.matfn(fn, a), f"{name} nc={nc} general")
