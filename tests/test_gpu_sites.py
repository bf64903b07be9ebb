"""GPU parity of every Wilson-hop symbol of the C ABI against the oracle.

Every symbol of include/qla_mi355x_sites.h (QLA_F3_* and QLA_D3_*: the four
eqops, M and Ma, field and pointer operands, v- and x-forms, for V = M*V,
H = M*spproj(D), D = sprecon(M*H), D = sprecon(M*spproj(D)), the
multi-source V = sum_d M_d*V_d, the matrix-free H = spproj(D),
D = sprecon(H), D = sprecon(spproj(D)), and the products H = M*H, D = M*D,
M = M*M, M = M*M^dagger) is called through the C ABI and compared
with the oracle bit for bit.  Each call uses a ragged site count (the last
tile is partial), permuted pointer targets and, for the x-forms, a parity
subset; the spin forms run two mu/sign pairs per symbol and a full
mu/sign sweep per (kind, eqop, adjoint).  The per-operand index forms of
QLA_INDEX_ALL (r[indexr[i]], a[indexa[i]], b[indexb[i]]) use distinct random
subsets per operand.  Also covered: misaligned fields
(the 4-byte load path), mu out of range and n = 0 (no-ops), multi-source
sums longer than one launch (nd > 16), nd = 0, and term lists in device
memory.
"""
import ctypes
import zlib

import numpy as np
import pytest

import qla_amd
from qla_amd import lattice
from qla_amd.abi_gen import variants
from gpu_util import assert_bitwise, dev, host, say

pytestmark = pytest.mark.gpu

KIND = {"matvec": 0, "spproj": 1, "sprecon": 2, "wilson": 3, "proj": 5, "recon": 6, "projrecon": 7,
        "mat_h": 8, "mat_d": 9, "mat_m": 10, "mat_ma": 11}
DIMS = (4, 4, 5, 6)             # 480 sites
N = 453                          # v-form site count: 7 full tiles and a ragged one
N_OWN = 333                      # steps of the per-operand index forms
ND = 3

_cache = {}


def setup(prec, src2, dest, nd=ND):
    """Host and device inputs for one (precision, source, destination) triple."""
    key = (prec, src2, dest, nd)
    if key in _cache:
        return _cache[key]
    nsites = lattice.volume(DIMS)
    dt = np.float32 if prec == "F3" else np.float64
    seed = zlib.crc32(repr(key).encode()) % 1000
    rng = np.random.default_rng(seed)
    s = dict(
        u=[lattice.random_su3(nsites, seed + d).astype(dt) for d in range(nd)],
        b=[lattice.random_field(src2, nsites, seed + 50 + d).astype(dt) for d in range(nd)],
        r0=lattice.random_field(dest, nsites, seed + 99).astype(dt),
        a_idx=[rng.permutation(nsites) for _ in range(nd)],
        b_idx=[rng.permutation(nsites) for _ in range(nd)],
        index=lattice.parity_sites(DIMS, 1)[:200].astype(np.int32),
        own={k: rng.choice(nsites, N_OWN, replace=False).astype(np.int32) for k in "rab"},
    )
    s["du"] = [dev(x) for x in s["u"]]
    s["db"] = [dev(x) for x in s["b"]]
    s["pa"] = [qla_amd.pointer_array(du, ix) for du, ix in zip(s["du"], s["a_idx"])]
    s["pb"] = [qla_amd.pointer_array(db, ix) for db, ix in zip(s["db"], s["b_idx"])]
    s["dindex"] = dev(s["index"])
    s["down"] = {k: dev(v) for k, v in s["own"].items()}
    _cache[key] = s
    return s


def ms_pairs(name):
    k = zlib.crc32(name.encode()) % 10
    return [(k % 5, 1 if k < 5 else -1), ((k + 3) % 5, -1 if k < 5 else 1)]


def run_unary(orc, v, s, mu, sign, n):
    """Matrix-free spin symbols: the one source (`a` of the QLA name) is the
    B data of the set-up; its own index list is own["a"]."""
    r = dev(s["r0"])
    src = s["pb"][0] if v.pa else s["db"][0]
    own, o = s["down"], s["own"]
    args = ([r] + ([own["r"]] if v.ir else []) + [src] + ([own["a"]] if v.ia else []) + [mu, sign]
            + ([s["dindex"]] if v.xf else []) + [n])
    say(f"{v.name} mu={mu} sign={sign} n={n}")
    qla_amd.call(v.name, *args)
    qla_amd.check()
    exp = s["r0"].copy()
    orc.site(KIND[v.kind], v.eqop, exp, None, s["b"][0], mu, sign, index=s["index"] if v.xf else None,
             b_idx=s["b_idx"][0] if v.pa else None, n=n, ir=o["r"] if v.ir else None,
             ib=o["a"] if v.ia else None)
    assert_bitwise(host(r), exp, f"{v.name} mu={mu} sign={sign}")


def run_site(orc, v, s, mu, sign, n=None):
    kind = KIND[v.kind]
    if n is None:
        n = len(s["index"]) if v.xf else (N_OWN if v.own_index else N)
    if v.unary:
        run_unary(orc, v, s, mu, sign, n)
        return
    r = dev(s["r0"])
    a = s["pa"][0] if v.pa else s["du"][0]
    b = s["pb"][0] if v.pb else s["db"][0]
    own = s["down"]
    args = ([r] + ([own["r"]] if v.ir else []) + [a] + ([own["a"]] if v.ia else []) + [b]
            + ([own["b"]] if v.ib else []) + ([mu, sign] if v.spin else []) + ([s["dindex"]] if v.xf else []) + [n])
    say(f"{v.name} mu={mu} sign={sign} n={n}")
    qla_amd.call(v.name, *args)
    qla_amd.check()
    exp = s["r0"].copy()
    o = s["own"]
    orc.site(kind, v.eqop, exp, s["u"][0], s["b"][0], mu, sign, index=s["index"] if v.xf else None,
             adj=v.adj, a_idx=s["a_idx"][0] if v.pa else None, b_idx=s["b_idx"][0] if v.pb else None,
             n=n, ir=o["r"] if v.ir else None, ia=o["a"] if v.ia else None, ib=o["b"] if v.ib else None)
    assert_bitwise(host(r), exp, f"{v.name} mu={mu} sign={sign}")


def term_list(ts):
    """Host array of nd device addresses (kept alive by the caller)."""
    return (ctypes.c_uint64 * len(ts))(*[t.data_ptr() for t in ts])


def run_multi(orc, v, s, nd, device_lists=False):
    n = len(s["index"]) if v.xf else (N_OWN if v.own_index else N)
    r = dev(s["r0"])
    a_terms = s["pa"][:nd] if v.pa else s["du"][:nd]
    b_terms = s["pb"][:nd] if v.pb else s["db"][:nd]
    if device_lists:
        import torch
        al = torch.tensor([t.data_ptr() for t in a_terms], dtype=torch.int64, device="cuda")
        bl = torch.tensor([t.data_ptr() for t in b_terms], dtype=torch.int64, device="cuda")
        a_arg, b_arg = al.data_ptr(), bl.data_ptr()
    else:
        al, bl = term_list(a_terms), term_list(b_terms)
        a_arg, b_arg = ctypes.addressof(al), ctypes.addressof(bl)
    own, o = s["down"], s["own"]
    args = ([r] + ([own["r"]] if v.ir else []) + [a_arg] + ([own["a"]] if v.ia else []) + [b_arg]
            + ([own["b"]] if v.ib else []) + ([s["dindex"]] if v.xf else []) + [n, nd])
    say(f"{v.name} nd={nd} device_lists={device_lists}")
    qla_amd.call(v.name, *args)
    qla_amd.check()
    exp = s["r0"].copy()
    orc.multi(v.eqop, exp, s["u"][:nd], s["b"][:nd], index=s["index"] if v.xf else None, adj=v.adj,
              a_idx=s["a_idx"][:nd] if v.pa else None, b_idx=s["b_idx"][:nd] if v.pb else None, n=n,
              ir=o["r"] if v.ir else None, ia=o["a"] if v.ia else None, ib=o["b"] if v.ib else None)
    assert_bitwise(host(r), exp, f"{v.name} nd={nd}")


@pytest.mark.parametrize("v", variants(), ids=lambda v: v.name)
def test_symbol(orc, v):
    s = setup(v.prec, v.src2, v.dest)
    if v.multi:
        run_multi(orc, v, s, ND)
    elif v.spin:
        for mu, sign in ms_pairs(v.name):
            run_site(orc, v, s, mu, sign)
    else:
        run_site(orc, v, s, 0, 1)


SWEEP = [v for v in variants() if v.spin and not v.xf and not v.own_index
         and ((not v.unary and not v.pa and v.pb) or (v.unary and v.pa))]


@pytest.mark.parametrize("v", SWEEP, ids=lambda v: v.name)
def test_all_mu_sign(orc, v):
    s = setup(v.prec, v.src2, v.dest)
    for ms in range(10):
        run_site(orc, v, s, ms % 5, 1 if ms < 5 else -1)


@pytest.mark.parametrize("prec", ["F3", "D3"])
@pytest.mark.parametrize("kind", ["spproj", "sprecon", "wilson", "proj", "recon", "projrecon"])
def test_noop_cases(orc, prec, kind):
    """mu outside 0..4 and n = 0 leave the destination untouched."""
    unary = kind in ("proj", "recon", "projrecon")
    v = [x for x in variants() if x.prec == prec and x.kind == kind and x.eqop == 0 and (x.pa if unary else x.pb)
         and (unary or not x.pa) and not x.adj and not x.xf and not x.own_index][0]
    s = setup(v.prec, v.src2, v.dest)
    for mu, n in ((5, N), (-1, N), (2, 0)):
        r = dev(s["r0"])
        say(f"{v.name} mu={mu} n={n}")
        if unary:
            qla_amd.call(v.name, r, s["pb"][0], mu, 1, n)
        else:
            qla_amd.call(v.name, r, s["du"][0], s["pb"][0], mu, 1, n)
        qla_amd.check()
        assert_bitwise(host(r), s["r0"], f"{v.name} no-op mu={mu} n={n}")


@pytest.mark.parametrize("prec", ["F3", "D3"])
@pytest.mark.parametrize("kind", ["matvec", "spproj", "sprecon", "wilson", "proj", "recon", "projrecon",
                                  "mat_h", "mat_d", "mat_m", "mat_ma"])
@pytest.mark.parametrize("eqop", [0, 1])
def test_misaligned_fields(orc, prec, kind, eqop):
    """Fields at an 8-byte (F3) offset from 16-byte alignment: the 4-byte path."""
    import torch
    v = [x for x in variants() if x.prec == prec and x.kind == kind and x.eqop == eqop and not x.pa
         and not x.pb and not x.adj and not x.xf and not x.own_index][0]
    s = setup(v.prec, v.src2, v.dest)
    dt = s["u"][0].dtype
    off = 1 if prec == "F3" else 0        # one complex: 8 B (F3) / 16 B (D3, always 16-aligned)
    if off == 0:
        pytest.skip("double complex records are always 16-byte aligned")

    def shifted(x):
        flat = np.concatenate([np.zeros(2, dtype=dt), x.reshape(-1)])
        t = torch.from_numpy(flat).to("cuda")
        return t, t.data_ptr() + 2 * t.element_size()

    tu, au = shifted(s["u"][0])
    tb, ab = shifted(s["b"][0])
    tr, ar = shifted(s["r0"])
    assert au % 16 == 8 and ab % 16 == 8 and ar % 16 == 8
    mu, sign = 3, -1
    args = ([ar] if v.unary else [ar, au]) + [ab] + ([mu, sign] if v.spin else []) + [N]
    say(f"{v.name} misaligned")
    qla_amd.call(v.name, *args)
    qla_amd.check()
    got = host(tr)[2:].reshape(s["r0"].shape)
    exp = s["r0"].copy()
    orc.site(KIND[kind], eqop, exp, None if v.unary else s["u"][0], s["b"][0], mu, sign, n=N)
    assert_bitwise(got, exp, f"{v.name} misaligned")


@pytest.mark.parametrize("prec", ["F3", "D3"])
def test_multi_long_and_empty(orc, prec):
    """nd = 20 (two launches: 16 terms, then 4 accumulated), nd = 0 (no-op),
    term lists in device memory."""
    for eqop in range(4):
        v = [x for x in variants() if x.prec == prec and x.multi and x.eqop == eqop and x.pb and not x.pa
             and not x.xf and not x.adj and not x.own_index][0]
        s = setup(v.prec, v.src2, v.dest, nd=20)
        run_multi(orc, v, s, 20)
        run_multi(orc, v, s, 5, device_lists=True)
        r = dev(s["r0"])
        lst = term_list(s["du"][:1])
        say(f"{v.name} nd=0")
        qla_amd.call(v.name, r, ctypes.addressof(lst), ctypes.addressof(lst), N, 0)
        qla_amd.check()
        assert_bitwise(host(r), s["r0"], f"{v.name} nd=0")


@pytest.mark.parametrize("prec", ["F3", "D3"])
@pytest.mark.parametrize("dims", [(32, 32, 8, 8)])
def test_wilson_full_lattice(orc, prec, dims):
    """D_vpeq_spproj_M_times_pD and its adjoint on a 65536-site lattice with
    neighbour gathers in all eight directions, accumulated into one field
    (the Wilson hopping term), against the oracle."""
    n = lattice.volume(dims)
    dt = np.float32 if prec == "F3" else np.float64
    fw, bw = lattice.neighbours(dims)
    u = [lattice.random_su3(n, 200 + m).astype(dt) for m in range(4)]
    psi = lattice.random_field("D", n, 300).astype(dt)
    du = [dev(x) for x in u]
    dpsi = dev(psi)
    r = dev(np.zeros((n, 3, 4, 2), dtype=dt))
    exp = np.zeros((n, 3, 4, 2), dtype=dt)
    for mu in range(4):
        # forward: U_mu(x) (1 - gamma_mu) psi(x + mu); backward: U_mu(x - mu)^dagger (1 + gamma_mu) psi(x - mu)
        say(f"{prec} wilson hop mu={mu}")
        qla_amd.call(f"QLA_{prec}_D_vpeq_spproj_M_times_pD", r, du[mu], qla_amd.pointer_array(dpsi, fw[mu]),
                     mu, -1, n)
        bwu = qla_amd.pointer_array(du[mu], bw[mu])
        qla_amd.call(f"QLA_{prec}_D_vpeq_spproj_pMa_times_pD", r, bwu, qla_amd.pointer_array(dpsi, bw[mu]),
                     mu, 1, n)
        qla_amd.check()
        orc.site(3, 1, exp, u[mu], psi, mu, -1, b_idx=fw[mu])
        orc.site(3, 1, exp, u[mu], psi, mu, 1, adj=True, a_idx=bw[mu], b_idx=bw[mu])
    assert_bitwise(host(r), exp, f"{prec} wilson hop")


@pytest.mark.parametrize("prec", ["F3", "D3"])
def test_split_wilson_hop_equals_fused(orc, prec):
    """The Wilson hopping term as a QDP-style caller builds it from the
    matrix-free forms -- forward: h = spproj(psi) on every site, then
    D += sprecon(U_mu(x) h(x + mu)); backward: h = U_mu^dagger spproj(psi),
    then D += sprecon(h(x - mu)) -- does the same arithmetic as the fused
    D_vpeq_spproj_M_times_pD / _pMa_times_pD, so the fields agree bit for bit
    (and with the oracle)."""
    dims = (16, 16, 8, 8)
    n = lattice.volume(dims)
    dt = np.float32 if prec == "F3" else np.float64
    fw, bw = lattice.neighbours(dims)
    u = [lattice.random_su3(n, 400 + m).astype(dt) for m in range(4)]
    psi = lattice.random_field("D", n, 500).astype(dt)
    du = [dev(x) for x in u]
    dpsi = dev(psi)
    split = dev(np.zeros((n, 3, 4, 2), dtype=dt))
    fused = dev(np.zeros((n, 3, 4, 2), dtype=dt))
    h = dev(np.zeros((n, 3, 2, 2), dtype=dt))
    for mu in range(4):
        say(f"{prec} split hop mu={mu}")
        qla_amd.call(f"QLA_{prec}_H_veq_spproj_D", h, dpsi, mu, -1, n)
        qla_amd.call(f"QLA_{prec}_D_vpeq_sprecon_M_times_pH", split, du[mu], qla_amd.pointer_array(h, fw[mu]),
                     mu, -1, n)
        qla_amd.call(f"QLA_{prec}_H_veq_spproj_Ma_times_D", h, du[mu], dpsi, mu, 1, n)
        qla_amd.call(f"QLA_{prec}_D_vpeq_sprecon_pH", split, qla_amd.pointer_array(h, bw[mu]), mu, 1, n)
        qla_amd.call(f"QLA_{prec}_D_vpeq_spproj_M_times_pD", fused, du[mu], qla_amd.pointer_array(dpsi, fw[mu]),
                     mu, -1, n)
        qla_amd.call(f"QLA_{prec}_D_vpeq_spproj_pMa_times_pD", fused, qla_amd.pointer_array(du[mu], bw[mu]),
                     qla_amd.pointer_array(dpsi, bw[mu]), mu, 1, n)
        qla_amd.check()
    got = host(split)
    assert_bitwise(got, host(fused), f"{prec} split hop vs fused")
    exp = np.zeros((n, 3, 4, 2), dtype=dt)
    for mu in range(4):
        orc.site(3, 1, exp, u[mu], psi, mu, -1, b_idx=fw[mu])
        orc.site(3, 1, exp, u[mu], psi, mu, 1, adj=True, a_idx=bw[mu], b_idx=bw[mu])
    assert_bitwise(got, exp, f"{prec} split hop vs oracle")
