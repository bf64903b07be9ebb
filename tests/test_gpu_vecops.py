"""GPU parity of every field linear-algebra symbol (include/qla_mi355x_vecops.h)
through the C ABI: copy / increment, Hermitian conjugate, transpose, complex
conjugate, times i, times a real or complex constant, sums, differences and
a*b +- c, for V, H, D, M, P in QLA_F3 and QLA_D3, every eqop and indexing
form (v, x, pointer operands, per-operand indices).

Bar: bit-exact against the oracle (orc_*_vecop) on the same inputs -- every
element is one or two IEEE operations in a fixed order, so any difference is
a bug -- including the records a call must leave untouched.  97 records
make the float v-form run of a V field odd (291 complexes), so the last
thread's single-complex tail is exercised.
"""
import zlib

import numpy as np
import pytest

import qla_amd
from qla_amd import lattice
from qla_amd.abi_gen import (CMP_OPS, COMP_FAMS, LOCALS, NCOMP, PMUL_KINDS, SCONV, SFN, VECFAMS, comps, fieldmuls, fills,
                             gammas, ints, locals_, masks, pmuls, sconvs, sfns, sums, vecops)
from gpu_util import assert_bitwise, dev, host, say

pytestmark = pytest.mark.gpu

N = 97
TDIM = {"M": 3, "P": 12}
S_REAL = -0.7071067811865476
S_CPLX = complex(0.6, -1.3)


class Case:
    """Fields, pointer arrays and index lists of one (precision, type)."""

    def __init__(self, prec, t, n=N):
        self.prec, self.t, self.n = prec, t, n
        self.dt = np.float64 if prec == "D3" else np.float32
        rng = np.random.default_rng(zlib.crc32(f"vec {prec} {t}".encode()))
        self.x = lattice.random_field(t, n, 31, dtype=self.dt)
        self.y = lattice.random_field(t, n, 32, dtype=self.dt)
        self.r0 = lattice.random_field(t, n, 33, dtype=self.dt)
        self.dx, self.dy = dev(self.x), dev(self.y)
        self.x_idx = rng.permutation(n)
        self.y_idx = rng.permutation(n)
        self.px = qla_amd.pointer_array(self.dx, self.x_idx)
        self.py = qla_amd.pointer_array(self.dy, self.y_idx)
        self.index = np.sort(rng.choice(n, n // 2, replace=False)).astype(np.int32)
        self.dindex = dev(self.index)
        m = n // 3
        self.own = {"r": rng.choice(n, m, replace=False).astype(np.int32),
                    "x": rng.choice(n, m, replace=False).astype(np.int32),
                    "y": rng.choice(n, m, replace=False).astype(np.int32)}
        self.down = {k: dev(v) for k, v in self.own.items()}
        self.sr = np.array([S_REAL], dtype=self.dt)
        self.sc = np.array([S_CPLX.real, S_CPLX.imag], dtype=self.dt)

    def scalar(self, v):
        """(host pointer of the constant, its value as the oracle takes it)."""
        if v.scalar == "r":
            return self.sr.ctypes.data, float(self.sr[0])
        if v.scalar == "c":
            return self.sc.ctypes.data, complex(float(self.sc[0]), float(self.sc[1]))
        return None, 0.0

    def run(self, orc, v, sptr=None):
        r = dev(self.r0)
        args = [r]
        if v.ir:
            args.append(self.down["r"])
        sp, sval = self.scalar(v)
        if v.scalar:
            args.append(sptr if sptr is not None else sp)
        args.append(self.px if v.px else self.dx)
        if v.ix:
            args.append(self.down["x"])
        if v.nsrc == 2:
            args.append(self.py if v.py else self.dy)
            if v.iy:
                args.append(self.down["y"])
        if v.xf:
            args.append(self.dindex)
        m = len(self.index) if v.xf else (len(self.own["r"]) if v.own_index else self.n)
        say(f"{v.name} n={m}")
        qla_amd.call(v.name, *args, m)
        qla_amd.check()
        # expected: the oracle on the records each step reads and writes
        step = np.arange(m)
        site = self.index if v.xf else step
        rs = self.own["r"] if v.ir else site
        xs = self.own["x"] if v.ix else site
        ys = self.own["y"] if v.iy else site
        gx = np.ascontiguousarray(self.x[self.x_idx[xs] if v.px else xs])
        gy = np.ascontiguousarray(self.y[self.y_idx[ys] if v.py else ys]) if v.nsrc == 2 else None
        sub = np.ascontiguousarray(self.r0[rs])
        orc.vecop(VECFAMS[v.fam][0], v.eqop, sub, gx, gy, s=sval, conj=bool(v.xform & 1),
                  dim=TDIM[v.t] if v.xform & 2 else 0)
        exp = self.r0.copy()
        exp[rs] = sub
        return host(r), exp


_cases = {}


def case(prec, t):
    if (prec, t) not in _cases:
        _cases[(prec, t)] = Case(prec, t)
    return _cases[(prec, t)]


@pytest.mark.parametrize("v", vecops(), ids=lambda v: v.name)
def test_vecop_symbol(orc, v):
    got, exp = case(v.prec, v.t).run(orc, v)
    assert_bitwise(got, exp, v.name)


@pytest.mark.parametrize("prec", ["F3", "D3"])
@pytest.mark.parametrize("fam", ["r_times", "c_times", "r_times_minus", "c_times_plus"])
def test_device_constant(orc, prec, fam):
    """The constant read from device memory gives the host-constant result."""
    import torch
    v = [x for x in vecops() if x.prec == prec and x.fam == fam and x.t == "D" and x.eqop == (1 if fam.endswith(
        "times") else 0) and not (x.xf or x.px or x.py or x.own_index or x.adj)][0]
    cs = case(prec, "D")
    dsc = torch.from_numpy(cs.sr if v.scalar == "r" else cs.sc).to("cuda")
    got, exp = cs.run(orc, v, sptr=dsc.data_ptr())
    assert_bitwise(got, exp, f"{v.name} device constant")


@pytest.mark.parametrize("prec", ["F3", "D3"])
@pytest.mark.parametrize("t", ["V", "D", "P"])
def test_unaligned_fields(orc, prec, t):
    """v-form fields starting one complex past a 16-byte boundary (the float
    run takes its one-complex-per-thread path)."""
    import torch
    v = [x for x in vecops() if x.prec == prec and x.fam == "c_times_plus" and x.t == t
         and not (x.xf or x.px or x.py or x.own_index)][0]
    cs = case(prec, t)
    n = cs.n - 1
    flat = lambda a: torch.from_numpy(np.ascontiguousarray(a).reshape(-1)).to("cuda")   # noqa: E731
    c2 = 2 * NCOMP[t]
    bx, by, br = flat(cs.x), flat(cs.y), flat(cs.r0)
    # n records of views that start one complex (2 reals) into each array
    x, y, r = bx[2:2 + n * c2], by[2:2 + n * c2], br[2:2 + n * c2]
    sp, sval = cs.scalar(v)
    say(f"{v.name} unaligned n={n}")
    qla_amd.call(v.name, r, sp, x, y, n)
    qla_amd.check()
    hx = cs.x.reshape(-1)[2:2 + n * c2].reshape((n,) + cs.x.shape[1:])
    hy = cs.y.reshape(-1)[2:2 + n * c2].reshape((n,) + cs.y.shape[1:])
    sub = np.ascontiguousarray(cs.r0.reshape(-1)[2:2 + n * c2].reshape((n,) + cs.r0.shape[1:]))
    orc.vecop(VECFAMS[v.fam][0], 0, sub, np.ascontiguousarray(hx), np.ascontiguousarray(hy), s=sval)
    exp = cs.r0.reshape(-1).copy()
    exp[2:2 + n * c2] = sub.reshape(-1)
    assert_bitwise(host(br), exp, f"{v.name} unaligned")


@pytest.mark.parametrize("prec", ["F3", "D3"])
@pytest.mark.parametrize("t", ["V", "D", "M"])
def test_unaligned_zero_fill(orc, prec, t):
    """The v-form zero fill of a field starting one complex past a 16-byte
    boundary writes exactly its n records (8-byte stores for float fields)
    and leaves the complex before and after them alone."""
    cs = case(prec, t)
    n = cs.n - 1
    c2 = 2 * NCOMP[t]
    import torch
    flat = torch.from_numpy(np.ascontiguousarray(cs.r0).reshape(-1)).to("cuda")
    view = flat[2:2 + n * c2]
    name = f"QLA_{prec}_{t}_veq_zero"
    say(f"{name} unaligned n={n}")
    qla_amd.call(name, view, n)
    qla_amd.check()
    exp = cs.r0.reshape(-1).copy()
    exp[2:2 + n * c2] = 0
    assert_bitwise(host(flat), exp, f"{name} unaligned")


def test_noop_and_errors(orc):
    """n <= 0 writes nothing; a null constant records an error, writes nothing."""
    import ctypes
    cs = case("F3", "D")
    r = dev(cs.r0)
    say("QLA_F3_D_veq_r_times_D n=0 / null constant")
    qla_amd.call("QLA_F3_D_veq_r_times_D", r, cs.sr.ctypes.data, cs.dx, 0)
    qla_amd.call("QLA_F3_D_veq_D_plus_D", r, cs.dx, cs.dy, -3)
    qla_amd.check()
    qla_amd.call("QLA_F3_D_veq_r_times_D", r, ctypes.c_void_p(0), cs.dx, cs.n)
    assert qla_amd.load().QLA_MI355X_last_error() != 0
    assert_bitwise(host(r), cs.r0, "no-op calls")


def test_full_volume_axpy(orc):
    """The solver updates x += a p and r = a p - x on a 32^4 lattice (F3
    DiracFermion, 1048576 records), bit-exact against the oracle."""
    n = 32 ** 4
    p = lattice.random_field("D", n, 41, dtype=np.float32)
    x = lattice.random_field("D", n, 42, dtype=np.float32)
    a = np.array([0.123456789], dtype=np.float32)
    dp, dx = dev(p), dev(x)
    say("QLA_F3_D_vpeq_r_times_D n=32^4")
    qla_amd.call("QLA_F3_D_vpeq_r_times_D", dx, a.ctypes.data, dp, n)
    qla_amd.check()
    exp = x.copy()
    orc.vecop(orc.RSCALE, 1, exp, p, s=float(a[0]))
    assert_bitwise(host(dx), exp, "x += a p")
    dr = dev(x)
    say("QLA_F3_D_veq_r_times_D_minus_D n=32^4")
    qla_amd.call("QLA_F3_D_veq_r_times_D_minus_D", dr, a.ctypes.data, dp, dev(x), n)
    qla_amd.check()
    exp = np.empty_like(x)
    orc.vecop(orc.R_MINUS, 0, exp, p, x, s=float(a[0]))
    assert_bitwise(host(dr), exp, "r = a p - x")


def test_wide_transpose_of_propagators():
    """P transposes over more than 2^32 complexes (the 64-bit step numbering
    of the gather kernel): 29826162 DiracPropagator records, checked on the
    device against torch's own axis swap (a transpose moves data only)."""
    import torch
    n = (1 << 32) // 144 + 7
    need = 3 * n * 1152 + (1 << 30)
    free, _ = torch.cuda.mem_get_info()
    if free < need:
        pytest.skip(f"needs {need >> 30} GiB of free device memory")
    g = torch.Generator(device="cuda").manual_seed(5)
    a = torch.randn((n, 12, 12, 2), dtype=torch.float32, device="cuda", generator=g)
    r = torch.empty_like(a)
    say(f"QLA_F3_P_veq_transpose_P n={n}")
    qla_amd.call("QLA_F3_P_veq_transpose_P", r, a, n)
    qla_amd.check()
    ok = torch.equal(r, a.transpose(1, 2))
    del a, r
    torch.cuda.empty_cache()
    assert ok


FILL_KIND = {"zero": 0, "rec": 1, "diag": 2}


FILL_CASES = [(v, w) for v in fills() for w in (("host",) if v.kind == "zero" else ("host", "device"))]


@pytest.mark.parametrize("v,where", FILL_CASES, ids=lambda x: x if isinstance(x, str) else x.name)
def test_fill_symbol(orc, v, where):
    """Every fill with the record / constant in host or device memory, against
    the oracle (x-forms write only the indexed records)."""
    import torch
    cs = case(v.prec, v.t)
    r = dev(cs.r0)
    if v.kind == "rec":
        a = np.ascontiguousarray(cs.x[5])
    elif v.kind == "diag":
        a = np.array([0.8125, -1.75], dtype=cs.dt)
    else:
        a = None
    args = [r]
    keep = None
    if a is not None:
        if where == "device":
            keep = torch.from_numpy(a).to("cuda")
            args.append(keep.data_ptr())
        else:
            args.append(a.ctypes.data)
    if v.xf:
        args.append(cs.dindex)
    m = len(cs.index) if v.xf else cs.n
    say(f"{v.name} {where} n={m}")
    qla_amd.call(v.name, *args, m)
    qla_amd.check()
    exp = cs.r0.copy()
    orc.fill(FILL_KIND[v.kind], v.eqop, exp, a, index=cs.index if v.xf else None)
    assert_bitwise(host(r), exp, f"{v.name} {where}")



def tposed_cases():
    """(symbol, byte offset): the contiguous single-precision transposed and
    adjoint forms of one source, on 16-byte aligned fields (two complexes an
    access) and 8 bytes off (one complex an access)."""
    vs = [v for v in vecops() if v.prec == "F3" and v.xform & 2 and v.nsrc == 1
          and not (v.ir or v.ix or v.iy or v.px or v.py or v.xf)]
    return [(v, o) for v in vs for o in (0, 8)]


@pytest.mark.parametrize("v,offset", tposed_cases(), ids=lambda x: x if isinstance(x, int) else x.name)
def test_tposed_blocks(orc, v, offset):
    """Transposed forms over 1,001 records (several LDS block steps and an
    odd complex count for colour matrices) against the oracle, bit for bit."""
    import torch
    cs = Case(v.prec, v.t, n=1001)
    keep = []

    def place(a):
        flat = torch.cat([torch.zeros(offset // 4), torch.from_numpy(np.ascontiguousarray(a)).reshape(-1)]).to("cuda")
        keep.append(flat)
        return flat.data_ptr() + offset

    args = [place(cs.r0)]
    sp, sval = cs.scalar(v)
    if v.scalar:
        args.append(sp)
    args.append(place(cs.x))
    say(f"{v.name} n={cs.n} offset={offset}")
    qla_amd.call(v.name, *args, cs.n)
    qla_amd.check()
    exp = np.ascontiguousarray(cs.r0.copy())
    orc.vecop(VECFAMS[v.fam][0], v.eqop, exp, np.ascontiguousarray(cs.x), None, s=sval, conj=bool(v.xform & 1),
              dim=TDIM[v.t])
    got = host(keep[0])[offset // 4:].reshape(exp.shape)
    assert_bitwise(got, exp, f"{v.name} offset={offset}")


GAMMA_CASES = [(v, mu) for v in gammas() for mu in (range(16) if not (v.xf or v.pa or v.ir or v.ia) else (5, 15))]


@pytest.mark.parametrize("v,mu", GAMMA_CASES, ids=lambda x: x if isinstance(x, int) else x.name)
def test_gamma_symbol(orc, v, mu):
    """Every gamma multiplication symbol (all 16 mu on the v-forms) bit for bit
    against the oracle, untouched records included."""
    cs = case(v.prec, v.t)
    r = dev(cs.r0)
    args = [r]
    if v.ir:
        args.append(cs.down["r"])
    args.append(cs.px if v.pa else cs.dx)
    if v.ia:
        args.append(cs.down["x"])
    args.append(mu)
    if v.xf:
        args.append(cs.dindex)
    m = len(cs.index) if v.xf else (len(cs.own["r"]) if (v.ir or v.ia) else cs.n)
    say(f"{v.name} mu={mu} n={m}")
    qla_amd.call(v.name, *args, m)
    qla_amd.check()
    site = cs.index if v.xf else np.arange(m)
    rs = cs.own["r"] if v.ir else site
    xs = cs.own["x"] if v.ia else site
    gx = np.ascontiguousarray(cs.x[cs.x_idx[xs] if v.pa else xs])
    sub = np.zeros((m,) + cs.r0.shape[1:], dtype=cs.dt)
    orc.gamma(v.right, mu, sub, gx)
    exp = cs.r0.copy()
    exp[rs] = sub
    assert_bitwise(host(r), exp, f"{v.name} mu={mu}")



GAMMA_OFFSET_CASES = [(v, mu, off) for v in gammas() if not (v.xf or v.pa or v.ir or v.ia) and v.prec == "F3"
                      for mu in (3, 10, 15) for off in (0, 8)]


@pytest.mark.parametrize("v,mu,offset", GAMMA_OFFSET_CASES, ids=lambda x: x if isinstance(x, int) else x.name)
def test_gamma_flat_blocks(orc, v, mu, offset):
    """The contiguous single-precision gamma forms over 1,001 sites on
    16-byte aligned fields (two complexes an access) and on fields 8 bytes
    off that alignment (one complex an access)."""
    import torch
    n = 1001
    x = lattice.random_field(v.t, n, 51, dtype=np.float32)
    r0 = lattice.random_field(v.t, n, 52, dtype=np.float32)
    keep = []

    def place(a):
        flat = torch.cat([torch.zeros(offset // 4), torch.from_numpy(np.ascontiguousarray(a)).reshape(-1)]).to("cuda")
        keep.append(flat)
        return flat.data_ptr() + offset

    rp, xp = place(r0), place(x)
    say(f"{v.name} mu={mu} n={n} offset={offset}")
    qla_amd.call(v.name, rp, xp, mu, n)
    qla_amd.check()
    exp = np.zeros_like(r0)
    orc.gamma(v.right, mu, exp, x)
    got = host(keep[0])[offset // 4:].reshape(exp.shape)
    assert_bitwise(got, exp, f"{v.name} mu={mu} offset={offset}")


def test_gamma_out_of_range_is_a_noop():
    cs = case("F3", "D")
    r = dev(cs.r0)
    say("QLA_F3_D_veq_gamma_times_D mu=16 / -1")
    qla_amd.call("QLA_F3_D_veq_gamma_times_D", r, cs.dx, 16, cs.n)
    qla_amd.call("QLA_F3_D_veq_gamma_times_D", r, cs.dx, -1, cs.n)
    qla_amd.check()
    assert_bitwise(host(r), cs.r0, "mu out of range")


LOCAL_DEST = {"R": (), "C": (2,), "M": (3, 3, 2)}


@pytest.mark.parametrize("v", locals_(), ids=lambda v: v.name)
def test_local_symbol(orc, v):
    """Every site-local reduction and colour-matrix part bit for bit against
    the oracle (the per-record double sums run in the same order)."""
    code, dest = LOCALS[v.kind][0], LOCALS[v.kind][1]
    cs = case(v.prec, v.t)
    r0 = np.random.default_rng(zlib.crc32(v.name.encode())).standard_normal((cs.n,) + LOCAL_DEST[dest]).astype(cs.dt)
    r = dev(r0)
    args = [r]
    if v.ir:
        args.append(cs.down["r"])
    args.append(cs.px if v.pa else cs.dx)
    if v.ia:
        args.append(cs.down["x"])
    if v.binary:
        args.append(cs.py if v.pb else cs.dy)
        if v.ib:
            args.append(cs.down["y"])
    if v.xf:
        args.append(cs.dindex)
    m = len(cs.index) if v.xf else (len(cs.own["r"]) if v.own_index else cs.n)
    say(f"{v.name} n={m}")
    qla_amd.call(v.name, *args, m)
    qla_amd.check()
    site = cs.index if v.xf else np.arange(m)
    rs = cs.own["r"] if v.ir else site
    xs = cs.own["x"] if v.ia else site
    ys = cs.own["y"] if v.ib else site
    gx = np.ascontiguousarray(cs.x[cs.x_idx[xs] if v.pa else xs])
    gy = np.ascontiguousarray(cs.y[cs.y_idx[ys] if v.pb else ys]) if v.binary else None
    sub = np.ascontiguousarray(r0[rs])
    orc.local(code, v.eqop, sub, gx, gy)
    exp = r0.copy()
    exp[rs] = sub
    assert_bitwise(host(r), exp, v.name)



STAGE_N = 805


def local_blocks_cases():
    """(symbol, byte offset of the fields): the contiguous forms at offset 0
    and, in single precision, 8 bytes off 16-byte alignment."""
    flat = [v for v in locals_() if not (v.ir or v.pa or v.pb or v.ia or v.ib or v.xf)]
    return [(v, 0) for v in flat] + [(v, 8) for v in flat if v.prec != "D3"]


@pytest.mark.parametrize("v,offset", local_blocks_cases(), ids=lambda c: getattr(c, "name", str(c)))
def test_local_symbol_blocks(orc, v, offset):
    """The contiguous site-local forms over several block steps with a ragged
    last one (805 sites), on 16-byte aligned fields and, in single
    precision, on fields 8 bytes off that alignment (the 8-byte staging
    loads)."""
    import torch
    code, dest = LOCALS[v.kind][0], LOCALS[v.kind][1]
    n = STAGE_N
    dt = np.float64 if v.prec == "D3" else np.float32
    x = lattice.random_field(v.t, n, 41, dtype=dt)
    y = lattice.random_field(v.t, n, 42, dtype=dt) if v.binary else None
    r0 = np.random.default_rng(zlib.crc32(v.name.encode())).standard_normal((n,) + LOCAL_DEST[dest]).astype(dt)
    keep = []

    def place(a):
        flat = torch.cat([torch.zeros(offset // a.itemsize, dtype=torch.float64 if dt == np.float64 else torch.float32),
                          torch.from_numpy(np.ascontiguousarray(a)).reshape(-1)]).to("cuda")
        keep.append(flat)
        return flat.data_ptr() + offset

    args = [place(r0), place(x)] + ([place(y)] if v.binary else [])
    say(f"{v.name} n={n} offset={offset}")
    qla_amd.call(v.name, *args, n)
    qla_amd.check()
    exp = np.ascontiguousarray(r0.copy())
    orc.local(code, v.eqop, exp, np.ascontiguousarray(x), None if y is None else np.ascontiguousarray(y))
    got = host(keep[0])[offset // r0.itemsize:].reshape(exp.shape)
    assert_bitwise(got, exp, f"{v.name} offset {offset}")


SUM_N = 1000


@pytest.mark.parametrize("v", sums(), ids=lambda v: v.name)
def test_sum_symbol(orc, v):
    """Every global field sum against the oracle's double sum: the summation
    order differs (the reference's OpenMP reduction fixes none), so a double
    result agrees within n double roundings of sum |a| per real column and a
    float result within one float rounding more."""
    import ctypes
    in_d = v.lib in ("D3", "D")
    out_d = v.lib in ("D3", "DF3", "D", "DF")
    dt = np.float64 if in_d else np.float32
    shape = {"R": (SUM_N,), "C": (SUM_N, 2)}.get(v.t)
    rng = np.random.default_rng(zlib.crc32(v.name.encode()))
    a = (rng.standard_normal(shape) if shape else lattice.random_field(v.t, SUM_N, 90, dtype=np.float64)).astype(dt)
    da = dev(a)
    idx = np.sort(rng.choice(SUM_N, SUM_N // 3, replace=False)).astype(np.int32)
    tgt = rng.permutation(SUM_N)
    src = qla_amd.pointer_array(da, tgt) if v.pa else da
    rec_shape = a.shape[1:]
    out = np.zeros(rec_shape, dtype=np.float64 if out_d else np.float32)
    args = [out.ctypes.data, src]
    if v.xf:
        args.append(dev(idx))
    m = len(idx) if v.xf else SUM_N
    say(f"{v.name} n={m}")
    qla_amd.call(v.name, *args, m)
    qla_amd.check()
    site = idx if v.xf else np.arange(m)
    recs = a[tgt[site]] if v.pa else a[site]
    want = orc.field_sum(np.ascontiguousarray(recs))
    bound = m * np.finfo(np.float64).eps * np.abs(recs.astype(np.float64)).sum(axis=0)
    if not out_d:
        bound = bound + np.finfo(np.float32).eps * np.abs(want)
    assert np.all(np.abs(out.astype(np.float64) - want) <= bound + 1e-300), v.name



@pytest.mark.parametrize("name,t,n,offset", [
    ("QLA_F3_d_veq_sum_D", "D", 100003, 0), ("QLA_F3_d_veq_sum_D", "D", 100003, 8),
    ("QLA_DF3_d_veq_sum_D", "D", 100003, 0), ("QLA_D3_d_veq_sum_D", "D", 100003, 0),
    ("QLA_F3_m_veq_sum_M", "M", 50001, 0), ("QLA_F3_m_veq_sum_M", "M", 50001, 4),
    ("QLA_F3_p_veq_sum_P", "P", 3001, 0), ("QLA_D3_p_veq_sum_P", "P", 3001, 0)])
def test_sum_contiguous_blocks(orc, name, t, n, offset):
    """Contiguous sums over many blocks through the vector reads (16 bytes a
    lane; 8 for float pairs on fields 8 bytes off 16-byte alignment; 4-byte
    offset: the per-column reads), within the double-rounding bound."""
    import torch
    in_d = name.startswith("QLA_D3")
    dt = np.float64 if in_d else np.float32
    a = lattice.random_field(t, n, 93, dtype=np.float64).astype(dt)
    flat = torch.cat([torch.zeros(offset // a.itemsize, dtype=torch.float64 if in_d else torch.float32),
                      torch.from_numpy(np.ascontiguousarray(a)).reshape(-1)]).to("cuda")
    out_d = not name.startswith("QLA_F3")
    out = np.zeros(a.shape[1:], dtype=np.float64 if out_d else np.float32)
    say(f"{name} n={n} offset={offset}")
    qla_amd.call(name, out.ctypes.data, flat.data_ptr() + offset, n)
    qla_amd.check()
    want = orc.field_sum(np.ascontiguousarray(a))
    bound = n * np.finfo(np.float64).eps * np.abs(a.astype(np.float64)).sum(axis=0)
    if not out_d:
        bound = bound + np.finfo(np.float32).eps * np.abs(want)
    assert np.all(np.abs(out.astype(np.float64) - want) <= bound + 1e-300), name


def test_sum_device_result_and_empty(orc):
    import torch
    a = lattice.random_field("D", SUM_N, 91, dtype=np.float32)
    out = torch.zeros((3, 4, 2), dtype=torch.float64, device="cuda")
    say("QLA_DF3_d_veq_sum_D device result")
    qla_amd.call("QLA_DF3_d_veq_sum_D", out, dev(a), SUM_N)
    qla_amd.check()
    want = orc.field_sum(a)
    got = host(out)
    assert np.all(np.abs(got - want) <= SUM_N * np.finfo(np.float64).eps * np.abs(a.astype(np.float64)).sum(axis=0))
    h = np.full((3, 4, 2), 7.0, dtype=np.float32)
    qla_amd.call("QLA_F3_d_veq_sum_D", h.ctypes.data, dev(a), 0)
    qla_amd.check()
    assert not np.any(h)


_factors = {}


def factors(prec):
    """real and complex factor fields of N sites with device copies and pointer arrays"""
    if prec not in _factors:
        dt = np.float64 if prec == "D3" else np.float32
        rng = np.random.default_rng(zlib.crc32(f"factors {prec}".encode()))
        fr = rng.standard_normal(N).astype(dt)
        fc = rng.standard_normal((N, 2)).astype(dt)
        perm = rng.permutation(N)
        dfr, dfc = dev(fr), dev(fc)
        _factors[prec] = {"R": (fr, dfr, qla_amd.pointer_array(dfr.view(N, 1), perm)),
                          "C": (fc, dfc, qla_amd.pointer_array(dfc, perm)), "perm": perm}
    return _factors[prec]


@pytest.mark.parametrize("v", fieldmuls(), ids=lambda v: v.name)
def test_fieldmul_symbol(orc, v):
    """Every multiplication by a real or complex field and every outer product,
    bit for bit against the oracle (untouched records included)."""
    cs = case(v.prec, v.t)
    if v.kind == "outer":
        vc = case(v.prec, "V")
        a_h, a_d, a_p, a_perm = vc.x, vc.dx, vc.px, vc.x_idx
        b_h, b_d, b_p, b_perm = vc.y, vc.dy, vc.py, vc.y_idx
    else:
        fz = factors(v.prec)
        a_h, a_d, a_p = fz[v.kind]
        a_perm = fz["perm"]
        b_h, b_d, b_p, b_perm = cs.x, cs.dx, cs.px, cs.x_idx
    r = dev(cs.r0)
    args = [r]
    if v.ir:
        args.append(cs.down["r"])
    args.append(a_p if v.pa else a_d)
    if v.ia:
        args.append(cs.down["x"])
    args.append(b_p if v.pb else b_d)
    if v.ib:
        args.append(cs.down["y"])
    if v.xf:
        args.append(cs.dindex)
    m = len(cs.index) if v.xf else (len(cs.own["r"]) if v.own_index else cs.n)
    say(f"{v.name} n={m}")
    qla_amd.call(v.name, *args, m)
    qla_amd.check()
    site = cs.index if v.xf else np.arange(m)
    rs = cs.own["r"] if v.ir else site
    as_ = cs.own["x"] if v.ia else site
    bs = cs.own["y"] if v.ib else site
    ga = np.ascontiguousarray(a_h[a_perm[as_] if v.pa else as_])
    gb = np.ascontiguousarray(b_h[b_perm[bs] if v.pb else bs])
    sub = np.ascontiguousarray(cs.r0[rs])
    if v.kind == "outer":
        orc.outer(v.eqop, sub, ga, gb)
    else:
        orc.fieldmul(v.eqop, sub, ga, gb, fa=v.fa, cb=v.adj, dim=TDIM[v.t] if v.adj else 0)
    exp = cs.r0.copy()
    exp[rs] = sub
    assert_bitwise(host(r), exp, v.name)


@pytest.mark.parametrize("v", pmuls(), ids=lambda v: v.name)
def test_pmul_symbol(orc, v):
    """Every product with a propagator (M*P, P*M, P*P, either factor's
    adjoint), bit for bit against the oracle (untouched records included)."""
    cs = case(v.prec, "P")
    ta, tb = PMUL_KINDS[v.kind]
    ca_, cb_ = case(v.prec, ta), case(v.prec, tb)
    a_h, a_d, a_p, a_perm = ca_.x, ca_.dx, ca_.px, ca_.x_idx
    b_h, b_d, b_p, b_perm = cb_.y, cb_.dy, cb_.py, cb_.y_idx
    r = dev(cs.r0)
    args = [r]
    if v.ir:
        args.append(cs.down["r"])
    args.append(a_p if v.pa else a_d)
    if v.ia:
        args.append(cs.down["x"])
    args.append(b_p if v.pb else b_d)
    if v.ib:
        args.append(cs.down["y"])
    if v.xf:
        args.append(cs.dindex)
    m = len(cs.index) if v.xf else (len(cs.own["r"]) if v.own_index else cs.n)
    say(f"{v.name} n={m}")
    qla_amd.call(v.name, *args, m)
    qla_amd.check()
    site = cs.index if v.xf else np.arange(m)
    rs = cs.own["r"] if v.ir else site
    as_ = cs.own["x"] if v.ia else site
    bs = cs.own["y"] if v.ib else site
    ga = np.ascontiguousarray(a_h[a_perm[as_] if v.pa else as_])
    gb = np.ascontiguousarray(b_h[b_perm[bs] if v.pb else bs])
    sub = np.ascontiguousarray(cs.r0[rs])
    orc.pmul(v.kind, v.eqop, sub, ga, gb, ca=v.ca, cb=v.cb)
    exp = cs.r0.copy()
    exp[rs] = sub
    assert_bitwise(host(r), exp, v.name)


def test_propagator_products_large(orc):
    """P = P * P^dagger over 2^16 sites (4096 block steps) against the oracle, bit for bit."""
    n = 1 << 16
    a = lattice.random_field("P", n, 71, dtype=np.float32)
    b = lattice.random_field("P", n, 72, dtype=np.float32)
    r = dev(np.zeros_like(a))
    say(f"QLA_F3_P_veq_P_times_Pa n={n}")
    qla_amd.call("QLA_F3_P_veq_P_times_Pa", r, dev(a), dev(b), n)
    qla_amd.check()
    exp = np.zeros_like(a)
    orc.pmul(2, 0, exp, a, b, cb=True)
    assert_bitwise(host(r), exp, "QLA_F3_P_veq_P_times_Pa")


_cfields = {}


def _c_field(prec, seed):
    """a random complex field of N sites (host, device)"""
    if (prec, seed) not in _cfields:
        dt = np.float64 if prec == "D3" else np.float32
        h = np.random.default_rng(seed).standard_normal((N, 2)).astype(dt)
        _cfields[(prec, seed)] = (h, dev(h))
    return _cfields[(prec, seed)]


def _comp_extra(v):
    """the symbol's element / column arguments, a seeded choice in range"""
    rng = np.random.default_rng(zlib.crc32(v.name.encode()))
    lim = {"ic": 3, "is": 2 if v.t == "H" else 4, "jc": 3, "js": 4}
    return [int(rng.integers(lim[x])) for x in v.extra]


@pytest.mark.parametrize("v", comps(), ids=lambda v: v.name)
def test_comp_symbol(orc, v):
    """Every element / colour-column / Dirac-column access and the spin trace,
    bit for bit against the oracle (untouched records and elements included)."""
    d, s = v.types
    ix = case(v.prec, "P")     # index arrays (every case holds N records)
    if s == "C":
        fz = factors(v.prec)
        a_h, a_d, a_p = fz["C"]
        a_perm = fz["perm"]
    else:
        sc = case(v.prec, s)
        a_h, a_d, a_p, a_perm = sc.x, sc.dx, sc.px, sc.x_idx
    r0 = _c_field(v.prec, 41)[0] if d == "C" else case(v.prec, d).r0
    extra = _comp_extra(v)
    r = dev(r0)
    args = [r]
    if v.ir:
        args.append(ix.down["r"])
    args.append(a_p if v.pa else a_d)
    if v.ia:
        args.append(ix.down["x"])
    args += extra
    if v.xf:
        args.append(ix.dindex)
    m = len(ix.index) if v.xf else (len(ix.own["r"]) if (v.ir or v.ia) else N)
    say(f"{v.name} {extra} n={m}")
    qla_amd.call(v.name, *args, m)
    qla_amd.check()
    site = ix.index if v.xf else np.arange(m)
    rs = ix.own["r"] if v.ir else site
    as_ = ix.own["x"] if v.ia else site
    ga = np.ascontiguousarray(a_h[a_perm[as_] if v.pa else as_])
    sub = np.ascontiguousarray(r0[rs])
    kind = COMP_FAMS[v.qual][2] + (0 if v.get else 1)
    orc.comp(kind, v.t, extra, sub, ga)
    exp = r0.copy()
    exp[rs] = sub
    assert_bitwise(host(r), exp, v.name)


def test_comp_out_of_range_is_an_error():
    """An element index outside its range records an error and writes nothing."""
    cs = case("F3", "M")
    c0 = _c_field("F3", 42)[0]
    r = dev(c0)
    say("QLA_F3_C_veq_elem_M out of range")
    qla_amd.call("QLA_F3_C_veq_elem_M", r, cs.dx, 3, 0, N)
    lib = qla_amd.load()
    lib.QLA_MI355X_synchronize()
    assert lib.QLA_MI355X_last_error() != 0
    assert_bitwise(host(r), c0, "unchanged")
    qla_amd.call("QLA_F3_V_veq_colorvec_H", dev(case("F3", "V").r0), case("F3", "H").dx, 2, N)
    lib.QLA_MI355X_synchronize()
    assert lib.QLA_MI355X_last_error() != 0
    assert lib.QLA_MI355X_last_error() == 0


def test_propagator_from_dirac_columns(orc):
    """Twelve Dirac-column insertions build a propagator at 2^16 sites that
    the spin trace and extraction read back exactly (oracle, bit for bit)."""
    n = 1 << 16
    p = lattice.random_field("P", n, 81, dtype=np.float32)
    cols = [np.ascontiguousarray(p[:, :, :, jc, js]) for jc in range(3) for js in range(4)]
    r = dev(np.zeros_like(p))
    say(f"QLA_F3_P_veq_diracvec_D x12 n={n}")
    for k, c in enumerate(cols):
        qla_amd.call("QLA_F3_P_veq_diracvec_D", r, dev(c), k // 4, k % 4, n)
    qla_amd.check()
    assert_bitwise(host(r), p, "P rebuilt")
    m = dev(np.zeros((n, 3, 3, 2), np.float32))
    qla_amd.call("QLA_F3_M_veq_spintrace_P", m, r, n)
    qla_amd.check()
    exp = np.zeros((n, 3, 3, 2), np.float32)
    orc.comp(orc.SPINTRACE, "P", (), exp, p)
    assert_bitwise(host(m), exp, "spintrace")


@pytest.mark.parametrize("name,kind", [("QLA_F3_P_vpeq_P_times_Pa", 2), ("QLA_F3_P_veq_Ma_times_P", 0),
                                       ("QLA_F3_P_vmeq_P_times_M", 1)])
def test_propagator_products_unaligned(orc, name, kind):
    """Fields starting one complex past a 16-byte boundary (the record staging
    takes its one-complex-per-thread path), bit for bit against the oracle."""
    import torch
    ta, tb = PMUL_KINDS[kind]
    n = N - 1
    ca_, cb_, cs = case("F3", ta), case("F3", tb), case("F3", "P")
    flat = lambda a: torch.from_numpy(np.ascontiguousarray(a).reshape(-1)).to("cuda")   # noqa: E731
    sub_of = lambda a, t: a.reshape(-1)[2:2 + n * 2 * NCOMP[t]].reshape((n,) + a.shape[1:])   # noqa: E731
    ba, bb, br = flat(ca_.x), flat(cb_.y), flat(cs.r0)
    say(f"{name} unaligned n={n}")
    qla_amd.call(name, br[2:], ba[2:], bb[2:], n)
    qla_amd.check()
    sub = np.ascontiguousarray(sub_of(cs.r0, "P"))
    eqop = {"veq": 0, "vpeq": 1, "vmeq": 2}[name.split("_")[3]]
    orc.pmul(kind, eqop, sub, np.ascontiguousarray(sub_of(ca_.x, ta)), np.ascontiguousarray(sub_of(cb_.y, tb)),
             ca=name.split("_")[4].endswith("a"), cb=name.endswith("a"))
    exp = cs.r0.reshape(-1).copy()
    exp[2:2 + n * 288] = sub.reshape(-1)
    assert_bitwise(host(br), exp, f"{name} unaligned")


@pytest.mark.parametrize("prec", ["F3", "D3"])
@pytest.mark.parametrize("t,fam", [("R", "r_times"), ("R", "r_times_minus"), ("C", "c_times_plus")])
def test_scalar_field_device_constant(orc, prec, t, fam):
    """R and C fields with the constant read from device memory give the
    host-constant result."""
    import torch
    v = [x for x in vecops() if x.prec == prec and x.fam == fam and x.t == t and not (
        x.xf or x.px or x.py or x.own_index or x.adj)][0]
    cs = case(prec, t)
    dsc = torch.from_numpy(cs.sr if v.scalar == "r" else cs.sc).to("cuda")
    got, exp = cs.run(orc, v, sptr=dsc.data_ptr())
    assert_bitwise(got, exp, f"{v.name} device constant")


@pytest.mark.parametrize("v", [v for v in sconvs() if v.kind < 6], ids=lambda v: v.name)
def test_sconv_symbol(orc, v):
    """Every real / complex conversion and scalar-field division, bit for bit
    against the oracle (untouched records included)."""
    dest, srcs, _ = SCONV[v.kind]
    cs = case(v.prec, dest)
    sa = case(v.prec, srcs[0])
    sb = case(v.prec, srcs[1]) if v.binary else None
    r = dev(cs.r0)
    args = [r]
    if v.ir:
        args.append(cs.down["r"])
    args.append(sa.px if v.pa else sa.dx)
    if v.ia:
        args.append(cs.down["x"])
    if v.binary:
        args.append(sb.py if v.pb else sb.dy)
        if v.ib:
            args.append(cs.down["y"])
    if v.xf:
        args.append(cs.dindex)
    m = len(cs.index) if v.xf else (len(cs.own["r"]) if (v.ir or v.ia or v.ib) else cs.n)
    say(f"{v.name} n={m}")
    qla_amd.call(v.name, *args, m)
    qla_amd.check()
    site = cs.index if v.xf else np.arange(m)
    rs = cs.own["r"] if v.ir else site
    as_ = cs.own["x"] if v.ia else site
    bs = cs.own["y"] if v.ib else site
    ga = np.ascontiguousarray(sa.x[sa.x_idx[as_] if v.pa else as_])
    gb = np.ascontiguousarray(sb.y[sb.y_idx[bs] if v.pb else bs]) if v.binary else None
    sub = np.ascontiguousarray(cs.r0[rs])
    orc.sconv(v.kind, sub, ga, gb)
    exp = cs.r0.copy()
    exp[rs] = sub
    assert_bitwise(host(r), exp, v.name)


# exact on both sides (IEEE operations, exact conversions); the rest are the
# device library's transcendental functions against the C library's
SFN_EXACT = ("sqrt", "fabs", "sign", "floor", "ceil", "mod", "max", "min", "ldexp", "norm", "csqrt")
SFN_ULPS = 6

_sfn_in = {}


def _sfn_inputs(prec, fn):
    """(a, b) host fields of N records inside the function's domain, and
    their device copies and pointer arrays (a permutation)"""
    key = (prec, fn)
    if key not in _sfn_in:
        dt = np.float64 if prec == "D3" else np.float32
        rng = np.random.default_rng(zlib.crc32(f"sfn {prec} {fn}".encode()))
        _, _, srcs, _ = SFN[fn]
        if srcs[0] == "C":
            a = rng.uniform(-4, 4, (N, 2)).astype(dt)
        elif fn in ("acos", "asin"):
            a = rng.uniform(-1, 1, N).astype(dt)
        elif fn in ("sqrt", "log", "log10", "pow"):
            a = rng.uniform(1e-2, 20, N).astype(dt)
        else:
            a = rng.uniform(-8, 8, N).astype(dt)
        if len(srcs) == 2:
            b = rng.integers(-30, 30, N).astype(np.int32) if srcs[1] == "I" else rng.uniform(-6, 6, N).astype(dt)
        else:
            b = None
        perm = rng.permutation(N)
        da = dev(a)
        db = None if b is None else dev(b)
        _sfn_in[key] = (a, b, da, db, qla_amd.pointer_array(da, perm),
                        None if b is None else qla_amd.pointer_array(db, perm), perm)
    return _sfn_in[key]


@pytest.mark.parametrize("v", sfns(), ids=lambda v: v.name)
def test_sfn_symbol(orc, v):
    """Every elementary function of real and complex fields against the oracle
    (the C library): bit for bit for the exact ones, within a few units in
    the last place for the transcendental ones; untouched records exact."""
    _, dest, srcs, _ = SFN[v.fn]
    cs = case(v.prec, dest)
    a, b, da, db, pa, pb, perm = _sfn_inputs(v.prec, v.fn)
    r = dev(cs.r0)
    args = [r]
    if v.ir:
        args.append(cs.down["r"])
    args.append(pa if v.pa else da)
    if v.ia:
        args.append(cs.down["x"])
    if v.binary:
        args.append(pb if v.pb else db)
        if v.ib:
            args.append(cs.down["y"])
    if v.xf:
        args.append(cs.dindex)
    m = len(cs.index) if v.xf else (len(cs.own["r"]) if (v.ir or v.ia or v.ib) else cs.n)
    say(f"{v.name} n={m}")
    qla_amd.call(v.name, *args, m)
    qla_amd.check()
    site = cs.index if v.xf else np.arange(m)
    rs = cs.own["r"] if v.ir else site
    as_ = cs.own["x"] if v.ia else site
    bs = cs.own["y"] if v.ib else site
    ga = np.ascontiguousarray(a[perm[as_] if v.pa else as_])
    gb = np.ascontiguousarray(b[perm[bs] if v.pb else bs]) if v.binary else None
    sub = np.ascontiguousarray(cs.r0[rs])
    orc.sfn(SFN[v.fn][0], sub, ga, gb)
    exp = cs.r0.copy()
    exp[rs] = sub
    got = host(r)
    if v.fn in SFN_EXACT:
        assert_bitwise(got, exp, v.name)
        return
    untouched = np.ones(cs.n, bool)
    untouched[rs] = False
    assert_bitwise(got[untouched], exp[untouched], v.name + " untouched")
    it = np.int32 if cs.dt == np.float32 else np.int64
    gi = got[rs].astype(cs.dt).view(it).astype(np.int64)
    ei = exp[rs].astype(cs.dt).view(it).astype(np.int64)
    gi = np.where(gi < 0, np.iinfo(it).min - gi, gi)
    ei = np.where(ei < 0, np.iinfo(it).min - ei, ei)
    assert np.abs(gi - ei).max() <= SFN_ULPS, (v.name, np.abs(gi - ei).max())


_mask_in = {}


def _int_mask(prec):
    """an int mask field of N sites (zeros and nonzeros), device copy and a pointer array"""
    if prec not in _mask_in:
        rng = np.random.default_rng(zlib.crc32(f"mask {prec}".encode()))
        m = rng.integers(-2, 3, N).astype(np.int32)
        dm = dev(m)
        perm = rng.permutation(N)
        _mask_in[prec] = (m, dm, qla_amd.pointer_array(dm, perm), perm)
    return _mask_in[prec]


@pytest.mark.parametrize("v", masks(), ids=lambda v: v.name)
def test_mask_symbol(orc, v):
    """Every copymask and real-field comparison, bit for bit against the oracle
    (untouched records included)."""
    if v.t == "cmp":
        src = case(v.prec, "R")
        b_h, b_d, b_p, b_perm = src.y, src.dy, src.py, src.y_idx
        dest = np.random.default_rng(5).integers(-9, 9, N).astype(np.int32)
    else:
        src = case(v.prec, v.t)
        b_h, b_d, b_p, b_perm = _int_mask(v.prec)
        dest = src.r0
    ix = case(v.prec, "P")
    r = dev(dest)
    args = [r]
    if v.ir:
        args.append(ix.down["r"])
    args.append(src.px if v.pa else src.dx)
    if v.ia:
        args.append(ix.down["x"])
    args.append(b_p if v.pb else b_d)
    if v.ib:
        args.append(ix.down["y"])
    if v.xf:
        args.append(ix.dindex)
    m = len(ix.index) if v.xf else (len(ix.own["r"]) if (v.ir or v.ia or v.ib) else N)
    say(f"{v.name} n={m}")
    qla_amd.call(v.name, *args, m)
    qla_amd.check()
    site = ix.index if v.xf else np.arange(m)
    rs = ix.own["r"] if v.ir else site
    as_ = ix.own["x"] if v.ia else site
    bs = ix.own["y"] if v.ib else site
    ga = np.ascontiguousarray(src.x[src.x_idx[as_] if v.pa else as_])
    gb = np.ascontiguousarray(b_h[b_perm[bs] if v.pb else bs])
    sub = np.ascontiguousarray(dest[rs])
    orc.mask(CMP_OPS.index(v.op) if v.t == "cmp" else -1, sub, ga, gb, real=src.dt)
    exp = dest.copy()
    exp[rs] = sub
    assert_bitwise(host(r), exp, v.name)


class IntCase:
    """int fields of N records (a, b, r0), device copies and pointer arrays"""

    def __init__(self, op):
        rng = np.random.default_rng(zlib.crc32(f"int {op}".encode()))
        self.a = rng.integers(-500, 500, N).astype(np.int32)
        self.b = rng.integers(-9, 10, N).astype(np.int32)
        if op in ("lshift", "rshift"):
            self.b = rng.integers(0, 20, N).astype(np.int32)
            if op == "lshift":
                self.a = np.abs(self.a).astype(np.int32)
        if op in ("divide", "mod"):
            self.b[self.b == 0] = 7
        self.r0 = rng.integers(-100, 100, N).astype(np.int32)
        self.perm = rng.permutation(N)
        self.da, self.db = dev(self.a), dev(self.b)
        self.pa = qla_amd.pointer_array(self.da, self.perm)
        self.pb = qla_amd.pointer_array(self.db, self.perm)


_icases = {}


@pytest.mark.parametrize("v", ints(), ids=lambda v: v.name)
def test_int_symbol(orc, v):
    """Every integer-field function, exactly against the oracle (untouched
    records included)."""
    if v.op not in _icases:
        _icases[v.op] = IntCase(v.op)
    ic = _icases[v.op]
    ix = case("F3", "P")
    sc = np.array([-3], np.int32)
    r = dev(ic.r0)
    args = [r]
    if v.ir:
        args.append(ix.down["r"])
    if v.scalar:
        args.append(sc.ctypes.data)
    if not v.fill:
        args.append(ic.pa if v.pa else ic.da)
        if v.ia:
            args.append(ix.down["x"])
        if v.binary:
            args.append(ic.pb if v.pb else ic.db)
            if v.ib:
                args.append(ix.down["y"])
    if v.xf:
        args.append(ix.dindex)
    m = len(ix.index) if v.xf else (len(ix.own["r"]) if (v.ir or v.ia or v.ib) else N)
    say(f"{v.name} n={m}")
    qla_amd.call(v.name, *args, m)
    qla_amd.check()
    site = ix.index if v.xf else np.arange(m)
    rs = ix.own["r"] if v.ir else site
    as_ = ix.own["x"] if v.ia else site
    bs = ix.own["y"] if v.ib else site
    ga = None if v.fill else np.ascontiguousarray(ic.a[ic.perm[as_] if v.pa else as_])
    gb = np.ascontiguousarray(ic.b[ic.perm[bs] if v.pb else bs]) if v.binary else None
    sub = np.ascontiguousarray(ic.r0[rs])
    from qla_amd.abi_gen import INT_OPS
    code = 40 if v.op == "zero" else 41 if v.op == "fill" else INT_OPS[v.op][0]
    orc.int_op(code, v.eqop, sub, ga, gb, sc=-3, n=m)
    exp = ic.r0.copy()
    exp[rs] = sub
    assert_bitwise(host(r), exp, v.name)


@pytest.mark.parametrize("v", [v for v in sconvs() if v.kind >= 6], ids=lambda v: v.name)
def test_real_int_conversion_symbol(orc, v):
    """R = I, I = trunc R, I = round R in every form, exactly against the oracle."""
    dt = np.float64 if v.prec == "D3" else np.float32
    rng = np.random.default_rng(zlib.crc32(v.name.encode()))
    ix = case(v.prec, "P")
    if v.kind == 6:
        a = rng.integers(-10000, 10000, N).astype(np.int32)
        r0 = rng.standard_normal(N).astype(dt)
    else:
        a = rng.uniform(-1000, 1000, N).astype(dt)
        a[:4] = [2.5, -2.5, 0.5, -1.5]
        r0 = rng.integers(-9, 9, N).astype(np.int32)
    perm = rng.permutation(N)
    da = dev(a)
    r = dev(r0)
    args = [r]
    if v.ir:
        args.append(ix.down["r"])
    args.append(qla_amd.pointer_array(da, perm) if v.pa else da)
    if v.ia:
        args.append(ix.down["x"])
    if v.xf:
        args.append(ix.dindex)
    m = len(ix.index) if v.xf else (len(ix.own["r"]) if (v.ir or v.ia) else N)
    say(f"{v.name} n={m}")
    qla_amd.call(v.name, *args, m)
    qla_amd.check()
    site = ix.index if v.xf else np.arange(m)
    rs = ix.own["r"] if v.ir else site
    as_ = ix.own["x"] if v.ia else site
    ga = np.ascontiguousarray(a[perm[as_] if v.pa else as_])
    sub = np.ascontiguousarray(r0[rs])
    orc.sconv(v.kind, sub, ga)
    exp = r0.copy()
    exp[rs] = sub
    assert_bitwise(host(r), exp, v.name)
