"""GPU parity of every colour-matrix function symbol (include/qla_mi355x_matfuncs.h)
through the C ABI: det, inverse, a^-1 b (vector and matrix b), exp, expA,
expTA, sqrt, sqrtPH, invsqrt, invsqrtPH, log and the Hermitian eigenvalues,
for QLA_F3, QLA_D3 and QLA_FN, in every indexing form (v, x, pointer
operands, per-operand indices).

Bars (against the oracle's restatement of the same template, same inputs):
* bit-exact where the template uses only +, -, *, IEEE divide and square
  root: det, inverse, the two solves, exp, sqrt, invsqrt, and expA / expTA /
  sqrtPH / invsqrtPH at nc != 3 (the templates' general branches);
* bit-exact for the single-precision nc = 3 closed forms of expA, expTA,
  sqrtPH, invsqrtPH and eigenvalsH (F3, and FN at nc = 3): they pass
  through acosf, cbrtf and sincosf, which the device evaluates with the
  host C library's own algorithms (qla_amd/csrc/qla_hostmath.hpp,
  checked against the library in tests/test_hostmath.py);
* log: its scale log(|A|_F) is a double log (device library vs glibc), so
  components agree to 4 ulp of the matrix's largest component;
* exp and log at nc = 1, 2 (QLA_FN): the templates' closed forms pass
  through lib/cmath's cexp / csinhcosh / clog, i.e. expf, sinhf, atan2f and a
  double log, which the device library and glibc round differently by an ulp
  or so; the 2x2 formulas multiply and divide those values a few times, so
  components agree to 8 ulp of the larger of 1 and the matrix's largest
  component (measured on 4,096 Gaussian matrices: exp 1.1 / 2.4 ulp, log
  1.0 / 3.7 ulp at nc = 1 / 2; every other function at nc = 1, 2 is
  bit-exact);
* the D3 closed forms pass through the double acos, sincos and cbrt, which
  the device library and glibc round differently by an ulp or so:
  components agree to 64 ulp of the matrix's largest component, times
  cond(M) for the (inverse) square roots.
Inputs: Gaussian matrices for det / inverses, i*H (H Hermitian traceless,
plus a phase for expA) for the exponentials, B^dagger B + eps for the
roots and log, Hermitian matrices for the eigenvalues; special matrices that
take the closed forms' double-root and r = 0 branches, whose branch tests
use exact arithmetic only, so GPU and oracle take the same branch.
"""
import zlib

import numpy as np
import pytest

import qla_amd
from qla_amd import lattice
from qla_amd.abi_gen import MATFUNCS, matfuncs
from gpu_util import assert_bitwise, dev, host, say

pytestmark = pytest.mark.gpu

N = 333
TRANSCENDENTAL = {"expA", "expTA", "sqrtPH", "invsqrtPH", "eigenvalsH"}
EXACT = {"det", "inverse", "inverse_V", "inverse_M", "exp", "sqrt", "invsqrt"}


def c(a):
    return a[..., 0] + 1j * a[..., 1]


def pack(z):
    return np.stack([z.real, z.imag], axis=-1)


def gen_input(fn, n, nc, seed):
    """(a, b) float64 inputs suited to fn; b is None for the unary functions."""
    rng = np.random.default_rng(seed)
    b = None
    if fn in ("det", "inverse", "inverse_V", "inverse_M"):
        a = rng.standard_normal((n, nc, nc, 2))
        if fn == "inverse_V":
            b = rng.standard_normal((n, nc, 2))
        elif fn == "inverse_M":
            b = rng.standard_normal((n, nc, nc, 2))
    elif fn in ("exp", "expTA"):
        a = lattice.random_antihermitian(n, nc, seed)
    elif fn == "expA":
        q = c(lattice.random_antihermitian(n, nc, seed))
        a = pack(q + 1j * rng.uniform(-3, 3, n)[:, None, None] * np.eye(nc))
    elif fn == "eigenvalsH":
        h = c(lattice.random_antihermitian(n, nc, seed)) / 1j
        a = pack(h + rng.uniform(-2, 2, n)[:, None, None] * np.eye(nc))
    else:
        a = lattice.random_hpd(n, nc, seed, eps_lo=1e-2)
    return a, b


def out_shape(fn, n, nc):
    dest = MATFUNCS[fn][1]
    return (n, 2) if dest == "C" else (n, nc, 2) if dest == "V" else (n, nc, nc, 2)


def expected(orc, fn, a, b):
    return orc.matfunc(fn, np.ascontiguousarray(a), None if b is None else np.ascontiguousarray(b))


def check(fn, nc, got, exp, a, what, gap_scale=True):
    """got / exp: (m, ...) result records of the steps actually written."""
    libm = fn == "log" or (fn == "exp" and nc <= 2)
    if not libm and (fn in EXACT or (fn in TRANSCENDENTAL and (nc != 3 or got.dtype == np.float32))):
        assert_bitwise(got, exp, what)
        return
    eps = float(np.finfo(got.dtype).eps)
    m = got.shape[0]
    g = got.reshape(m, -1).astype(np.float64)
    e = exp.reshape(m, -1).astype(np.float64)
    scale = np.abs(e).max(axis=1)
    if libm and nc <= 2:
        tol = 8 * eps * np.maximum(scale, 1.0)
        print(f"{what}: max error {(np.abs(g - e).max(axis=1) / (eps * np.maximum(scale, 1.0))).max():.2f} ulp")
    elif fn == "log":
        tol = 4 * eps * scale
    else:
        tol = 64 * eps * scale
        A = c(a.astype(np.float64))
        if fn in ("sqrtPH", "invsqrtPH"):
            tol = tol * np.linalg.cond(A)
        elif fn in ("expA", "expTA") and gap_scale:
            # the divided differences of exp over the eigenvalues amplify by
            # 1/gap when two eigenvalues are close (-iA is Hermitian)
            ev = np.linalg.eigvalsh(A / 1j)
            gap = np.diff(ev, axis=1).min(axis=1)
            tol = tol * np.maximum(1.0, 1.0 / np.maximum(gap, 1e-30))
    err = np.abs(g - e).max(axis=1)
    bad = np.nonzero(~(err <= tol))[0]
    assert bad.size == 0, (what, bad[:5], err[bad[:5]], tol[bad[:5]])


class Case:
    """Inputs, device copies, pointer arrays and index lists of one (lib, fn, nc)."""

    def __init__(self, lib, fn, nc, n=N):
        self.lib, self.fn, self.nc, self.n = lib, fn, nc, n
        dt = np.float64 if lib == "D3" else np.float32
        rng = np.random.default_rng(zlib.crc32(f"{lib} {fn} {nc}".encode()))
        a, b = gen_input(fn, n, nc, 1000 + nc * 17 + len(fn))
        self.a = a.astype(dt)
        self.b = None if b is None else b.astype(dt)
        self.da = dev(self.a)
        self.db = None if b is None else dev(self.b)
        self.a_idx = rng.permutation(n)
        self.b_idx = rng.permutation(n)
        self.pa = qla_amd.pointer_array(self.da, self.a_idx)
        self.pb = None if b is None else qla_amd.pointer_array(self.db, self.b_idx)
        self.index = np.sort(rng.choice(n, n // 2, replace=False)).astype(np.int32)
        self.dindex = dev(self.index)
        self.own = {k: rng.choice(n, n // 3, replace=False).astype(np.int32) for k in "rab"}
        self.down = {k: dev(v) for k, v in self.own.items()}
        self.r0 = np.random.default_rng(7).standard_normal(out_shape(fn, n, nc)).astype(dt)

    def run(self, v):
        """Call symbol v; returns (result field, expected field)."""
        r = dev(self.r0)
        args = [self.nc] if v.lib == "FN" else []
        args.append(r)
        if v.ir:
            args.append(self.down["r"])
        args.append(self.pa if v.pa else self.da)
        if v.ia:
            args.append(self.down["a"])
        if v.src2:
            args.append(self.pb if v.pb else self.db)
            if v.ib:
                args.append(self.down["b"])
        if v.xf:
            args.append(self.dindex)
        m = len(self.index) if v.xf else (len(self.own["r"]) if v.own_index else self.n)
        say(f"{v.name} nc={self.nc} n={m}")
        qla_amd.call(v.name, *args, m)
        qla_amd.check()
        # expected: step i reads a at its site (through the permutation for a
        # pointer operand) and writes r at its destination
        step = np.arange(m)
        site = self.index if v.xf else step
        rs = self.own["r"] if v.ir else site
        as_ = self.own["a"] if v.ia else site
        bs = self.own["b"] if v.ib else site
        ga = self.a[self.a_idx[as_] if v.pa else as_]
        gb = None if self.b is None else self.b[self.b_idx[bs] if v.pb else bs]
        exp = self.r0.copy()
        exp[rs] = expected(self._orc, self.fn, ga, gb)
        return host(r), exp, rs, ga


_cases = {}


def case(orc, lib, fn, nc):
    key = (lib, fn, nc)
    if key not in _cases:
        _cases[key] = Case(lib, fn, nc)
        _cases[key]._orc = orc
    return _cases[key]


def fn_nc(v):
    # nc = 7 runs exp, sqrt, invsqrt, log (and expA, expTA, sqrtPH,
    # invsqrtPH, which take their branches) on the row kernel, one matrix per
    # lane group, and det, the inverses and the eigenvalues' QR iteration on
    # one thread per matrix
    return 3 if v.lib != "FN" else 7


@pytest.mark.parametrize("v", matfuncs(), ids=lambda v: v.name)
def test_matfunc_symbol(orc, v):
    nc = fn_nc(v)
    cs = case(orc, v.lib, v.fn, nc)
    got, exp, rs, ga = cs.run(v)
    untouched = np.ones(cs.n, dtype=bool)
    untouched[rs] = False
    assert_bitwise(got[untouched], exp[untouched], f"{v.name} untouched records")
    check(v.fn, nc, got[rs], exp[rs], ga, v.name)


@pytest.mark.parametrize("v", [x for x in matfuncs() if x.lib == "FN"], ids=lambda v: v.name)
def test_matfunc_symbol_nc2(orc, v):
    """Every QLA_FN symbol again at nc = 2, where each function takes the
    template's own closed form, in every indexing form."""
    nc = 2
    cs = case(orc, v.lib, v.fn, nc)
    got, exp, rs, ga = cs.run(v)
    untouched = np.ones(cs.n, dtype=bool)
    untouched[rs] = False
    assert_bitwise(got[untouched], exp[untouched], f"{v.name} untouched records")
    check(v.fn, nc, got[rs], exp[rs], ga, v.name)


@pytest.mark.parametrize("nc", [1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("fn", list(MATFUNCS))
def test_fn_colours(orc, fn, nc):
    """Every QLA_FN v-form at each supported colour count."""
    v = [x for x in matfuncs() if x.lib == "FN" and x.fn == fn and not (x.xf or x.pa or x.pb or x.own_index)][0]
    cs = Case("FN", fn, nc, n=150)
    cs._orc = orc
    got, exp, rs, ga = cs.run(v)
    check(fn, nc, got, exp, ga, f"{v.name} nc={nc}")


@pytest.mark.parametrize("lib", ["F3", "D3"])
def test_closed_form_branches(orc, lib):
    """Double-root (cbrt) and r = 0 branches of the nc = 3 closed forms.  Their
    branch decisions use exact arithmetic on these inputs, so the GPU and the
    oracle take the same branch; values are then bit-exact in F3 and agree
    to the transcendental tolerance in D3."""
    dt = np.float64 if lib == "D3" else np.float32
    q = np.stack([1j * np.diag([2 * x, -x, -x]) for x in (0.5, 0.25, 1.5)] + [np.zeros((3, 3))])
    ph = np.stack([1j * np.diag([1.0, 1.0, 4.0]), 1j * 0.5 * np.eye(3)])
    hp = np.stack([np.diag([1.0, 1.0, 4.0]), 2.0 * np.eye(3), np.diag([0.25, 0.25, 9.0])])
    cases = [("expTA", q), ("expA", np.concatenate([q + 0.75j * np.eye(3), ph])), ("sqrtPH", hp),
             ("invsqrtPH", hp), ("eigenvalsH", hp)]
    for fn, z in cases:
        a = pack(z + 0j).astype(dt)
        name = f"QLA_{lib}_{MATFUNCS[fn][1]}_veq_{fn}_M"
        r = dev(np.zeros(out_shape(fn, len(a), 3), dtype=dt))
        say(f"{name} branch cases n={len(a)}")
        qla_amd.call(name, r, dev(a), len(a))
        qla_amd.check()
        check(fn, 3, host(r), orc.matfunc(fn, a), a, f"{name} branches", gap_scale=False)


def test_log_special(orc):
    """log(0) fills the diagonal with -inf (M_eq_d(r, log(0)))."""
    for lib in ("F3", "D3"):
        dt = np.float64 if lib == "D3" else np.float32
        z = np.zeros((3, 3, 3, 2), dtype=dt)
        r = dev(np.ones_like(z))
        say(f"QLA_{lib}_M_veq_log_M zero")
        qla_amd.call(f"QLA_{lib}_M_veq_log_M", r, dev(z), 3)
        qla_amd.check()
        assert_bitwise(host(r), orc.matfunc("log", z), f"QLA_{lib}_M_veq_log_M zero")
    # nc = 6 runs one thread per matrix, nc = 7 and 8 one matrix per lane
    # group; a zero matrix between nonzero ones leaves its neighbours untouched
    for nc in (6, 7, 8):
        a = lattice.random_hpd(16, nc, 3, eps_lo=1e-2).astype(np.float32)
        a[[0, 9]] = 0
        r = dev(np.ones_like(a))
        say(f"QLA_FN_M_veq_log_M zero nc={nc}")
        qla_amd.call("QLA_FN_M_veq_log_M", nc, r, dev(a), len(a))
        qla_amd.check()
        got, exp = host(r), orc.matfunc("log", a)
        assert_bitwise(got[[0, 9]], exp[[0, 9]], f"QLA_FN_M_veq_log_M zero nc={nc}")
        rest = np.setdiff1d(np.arange(len(a)), [0, 9])
        check("log", nc, got[rest], exp[rest], a[rest], f"QLA_FN_M_veq_log_M nc={nc}")


def test_errors_and_empty():
    """Unsupported colour counts record an error and write nothing; n = 0 is
    a no-op."""
    lib = qla_amd.load()
    a = np.ones((4, 9, 9, 2), dtype=np.float32)
    r = dev(a)
    say("QLA_FN_M_veq_expA_M nc=9 (unsupported)")
    qla_amd.call("QLA_FN_M_veq_expA_M", 9, r, dev(a), 4)
    assert lib.QLA_MI355X_last_error() != 0
    np.testing.assert_array_equal(host(r), a)
    a9 = np.ones((4, 9, 9, 2), dtype=np.float32)
    v = dev(np.zeros((4, 9, 2), dtype=np.float32))
    say("QLA_FN_V_veq_eigenvalsH_M nc=9 (unsupported)")
    qla_amd.call("QLA_FN_V_veq_eigenvalsH_M", 9, v, dev(a9), 4)
    assert lib.QLA_MI355X_last_error() != 0
    assert not host(v).any()
    a3 = np.ones((4, 3, 3, 2), dtype=np.float64)
    r3 = dev(a3)
    say("QLA_D3_M_veq_log_M n=0")
    qla_amd.call("QLA_D3_M_veq_log_M", r3, dev(np.zeros_like(a3)), 0)
    qla_amd.check()
    np.testing.assert_array_equal(host(r3), a3)


@pytest.mark.parametrize("nc", [4, 5, 6, 7, 8])
def test_eigenvalsH_qr_many(orc, nc):
    """QLA_FN_V_veq_eigenvalsH_M at nc = 4..8, the template's QR iteration,
    on 20,000 Hermitian matrices (the batch includes matrices that take the
    exceptional shifts), bit-exact against the oracle; then inputs with known
    roots: diagonal and zero matrices (exact), a NaN matrix between finite
    ones (the iteration ends; its neighbours are unaffected)."""
    rng = np.random.default_rng(500 + nc)
    g = rng.standard_normal((20000, nc, nc)) + 1j * rng.standard_normal((20000, nc, nc))
    a = pack(g + np.conj(np.swapaxes(g, 1, 2))).astype(np.float32)
    a = np.ascontiguousarray(a)
    r = dev(np.zeros((len(a), nc, 2), dtype=np.float32))
    say(f"QLA_FN_V_veq_eigenvalsH_M nc={nc} n={len(a)}")
    qla_amd.call("QLA_FN_V_veq_eigenvalsH_M", nc, r, dev(a), len(a))
    qla_amd.check()
    assert_bitwise(host(r), orc.matfunc("eigenvalsH", a), f"eigenvalsH nc={nc}")
    d = np.arange(1, nc + 1, dtype=np.float64)[::-1] - 2.5
    sp = np.stack([np.diag(d), np.zeros((nc, nc)), np.diag(d), np.diag(d)]) + 0j
    sp = np.ascontiguousarray(pack(sp).astype(np.float32))
    sp[2, 2, 1, 0] = np.nan         # below the diagonal, where the iteration reads
    r = dev(np.zeros((4, nc, 2), dtype=np.float32))
    say(f"QLA_FN_V_veq_eigenvalsH_M nc={nc} special inputs")
    qla_amd.call("QLA_FN_V_veq_eigenvalsH_M", nc, r, dev(sp), 4)
    qla_amd.check()
    got = host(r)
    for i in (0, 3):
        assert np.array_equal(got[i, :, 0], d.astype(np.float32)) and not got[i, :, 1].any()
    assert not got[1].any()
    assert np.isnan(got[2]).any()
    np.testing.assert_array_equal(got, orc.matfunc("eigenvalsH", sp))   # NaN where the oracle has NaN


@pytest.mark.parametrize("nc", [1, 2])
def test_small_nc_general_inputs(orc, nc):
    """QLA_FN at nc = 1, 2 on general complex Gaussian matrices (exp, log,
    sqrt, invsqrt, the inverses and det, whose closed forms then take complex
    eigenvalues and both signs of Re(conj(s) sqrt(d))), 4,096 matrices each,
    against the oracle: bit-exact but for exp and log (see the module
    docstring)."""
    rng = np.random.default_rng(70 + nc)
    n = 4096
    for fn in ("exp", "log", "sqrt", "invsqrt", "det", "inverse"):
        a = np.ascontiguousarray(rng.standard_normal((n, nc, nc, 2)).astype(np.float32))
        dest = MATFUNCS[fn][1]
        shape = (n, 2) if dest == "C" else a.shape
        r = dev(np.zeros(shape, dtype=np.float32))
        name = f"QLA_FN_{dest}_veq_{fn}_M"
        say(f"{name} nc={nc} general inputs")
        qla_amd.call(name, nc, r, dev(a), n)
        qla_amd.check()
        check(fn, nc, host(r), orc.matfunc(fn, a), a, f"{name} nc={nc} general")
