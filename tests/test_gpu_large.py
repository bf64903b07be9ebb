"""Maximum-size checks through the C ABI: fields whose byte offsets pass 4 GiB,
and site counts up to INT_MAX, the largest an int argument holds.

The site kernels, the reductions and the matrix functions address records
with 64-bit offsets; these tests run them on 48,000,037 sites (a D field of
4.6 GB) and 50,000,017 colour matrices (3.6 GB) and check a sample of 4,096
records that includes the last 1,024 against the oracle bit for bit (the
sampled records' inputs are gathered back to the host and evaluated as a
small call), and the global norm against a double sum formed by torch.
Inputs are drawn on the device; neighbour maps jump by about n/2 so the
gathers reach far addresses.  The INT_MAX tests run 2,147,483,647 sites
(33,554,432 tiles of 64, the last holding 63) with every pointer naming one
record, so each result holds one value everywhere; outputs start as NaN or
0xFF bytes, so a site left unwritten fails.  Each test is skipped when the
device has too little free memory.
"""
import numpy as np
import pytest

import qla_amd
from oracle import compressed as cm
from qla_amd import lattice
from gpu_util import assert_bitwise, dev, say

pytestmark = pytest.mark.gpu

N_SITES = 48_000_037
N_MATS = 50_000_017
N_INT_MAX = 2**31 - 1


def need(nbytes):
    import torch
    free, _ = torch.cuda.mem_get_info()
    if free < nbytes + (2 << 30):
        pytest.skip(f"needs {(nbytes + (2 << 30)) >> 30} GiB of free device memory")


def sample(n, seed):
    rng = np.random.default_rng(seed)
    s = np.concatenate([[0], rng.integers(0, n, 3071), np.arange(n - 1024, n)])
    return np.unique(s)


def test_spproj_sprecon_past_4GiB(orc):
    import torch
    n = N_SITES
    need(n * (72 + 96 + 96 + 48 + 8 + 8))
    g = torch.Generator(device="cuda").manual_seed(11)
    u = torch.randn((n, 3, 3, 2), dtype=torch.float32, device="cuda", generator=g)
    d = torch.randn((n, 3, 4, 2), dtype=torch.float32, device="cuda", generator=g)
    step = n // 2 + 12345
    nbr = (torch.arange(n, device="cuda", dtype=torch.int64) + step) % n
    h = torch.zeros((n, 3, 2, 2), dtype=torch.float32, device="cuda")
    mu, sign = 2, -1
    say(f"QLA_F3_H_veq_spproj_M_times_pD n={n}")
    qla_amd.call("QLA_F3_H_veq_spproj_M_times_pD", h, u, qla_amd.pointer_array(d, nbr), mu, sign, n)
    qla_amd.check()
    s = sample(n, 3)
    st = torch.from_numpy(s).cuda()
    nb = nbr[st]
    us, ds, hs = u[st].cpu().numpy(), d[nb].cpu().numpy(), h[st].cpu().numpy()
    exp = np.zeros_like(hs)
    orc.H_eqop_spproj_M_times_pD(orc.EQ, exp, us, ds, np.arange(len(s)), mu, sign)
    assert_bitwise(hs, exp, "spproj past 4 GiB")

    # D += sprecon(M * H[nbr]) with the destination past 4 GiB, in place on d
    d0s = d[st].cpu().numpy()
    hn = h[nb].cpu().numpy()
    say(f"QLA_F3_D_vpeq_sprecon_M_times_pH n={n}")
    qla_amd.call("QLA_F3_D_vpeq_sprecon_M_times_pH", d, u, qla_amd.pointer_array(h, nbr), mu, sign, n)
    qla_amd.check()
    exp = d0s.copy()
    orc.D_eqop_sprecon_M_times_pH(orc.PEQ, exp, us, hn, np.arange(len(s)), mu, sign)
    assert_bitwise(d[st].cpu().numpy(), exp, "sprecon past 4 GiB")
    del u, d, h, nbr
    torch.cuda.empty_cache()


def test_norm2_past_4GiB():
    import torch
    n = N_SITES
    need(n * 96 * 2)
    g = torch.Generator(device="cuda").manual_seed(12)
    d = torch.randn((n, 3, 4, 2), dtype=torch.float32, device="cuda", generator=g)
    r = torch.zeros(1, dtype=torch.float64, device="cuda")
    say(f"QLA_DF3_r_veq_norm2_D n={n}")
    qla_amd.call("QLA_DF3_r_veq_norm2_D", r, d, n)
    qla_amd.check()
    # the float norm of each complex, summed in double
    terms = d[..., 0] * d[..., 0]
    terms += d[..., 1] * d[..., 1]
    ref = float(terms.double().sum())
    got = float(r.cpu()[0])
    assert abs(got - ref) <= 1e-14 * ref, (got, ref)
    del d, terms
    torch.cuda.empty_cache()


def test_matrix_exp_past_4GiB(orc):
    import torch
    n = N_MATS
    need(n * 72 * 3)
    g = torch.Generator(device="cuda").manual_seed(13)
    x = torch.randn((n, 3, 3, 2), dtype=torch.float32, device="cuda", generator=g)
    # anti-Hermitian A = X - X^dagger (re antisymmetric, im symmetric)
    a = torch.empty_like(x)
    a[..., 0] = x[..., 0] - x[..., 0].transpose(1, 2)
    a[..., 1] = x[..., 1] + x[..., 1].transpose(1, 2)
    del x
    r = torch.empty_like(a)
    say(f"QLA_F3_M_veq_exp_M n={n}")
    qla_amd.call("QLA_F3_M_veq_exp_M", r, a, n)
    qla_amd.check()
    s = sample(n, 4)
    st = torch.from_numpy(s).cuda()
    exp = orc.matfn("exp", a[st].cpu().numpy())
    assert_bitwise(r[st].cpu().numpy(), exp, "exp past 4 GiB")
    del a, r
    torch.cuda.empty_cache()


def _differs(r, rec):
    """whether any record of r (n, ...) differs from rec (1, ...), bytewise,
    compared in slices of 2^28 records"""
    import torch
    flat, ref = r.view(r.shape[0], -1).view(torch.uint8), rec.reshape(1, -1).view(torch.uint8)
    step = 1 << 28
    return any(bool((flat[i:i + step] != ref).any()) for i in range(0, r.shape[0], step))


def test_matvec_at_int_max(orc):
    """V = M V with M and V through pointer arrays (QLA_F3_V_veq_pM_times_pV)
    on INT_MAX sites: the site kernel's tile loop and its last, partial tile."""
    import torch
    n = N_INT_MAX
    need(n * (8 + 8 + 24) + (1 << 28) * 24)
    m = lattice.random_su3(1, 21).astype(np.float32)
    v = np.random.default_rng(22).standard_normal((1, 3, 2)).astype(np.float32)
    dm, dv = dev(m), dev(v)
    pm = torch.full((n,), dm.data_ptr(), dtype=torch.int64, device="cuda")
    pv = torch.full((n,), dv.data_ptr(), dtype=torch.int64, device="cuda")
    r = torch.full((n, 3, 2), float("nan"), dtype=torch.float32, device="cuda")
    say(f"QLA_F3_V_veq_pM_times_pV n={n}")
    qla_amd.call("QLA_F3_V_veq_pM_times_pV", r, pm, pv, n)
    qla_amd.check()
    exp = orc.V_eqop_M_times_V(orc.EQ, np.zeros((1, 3, 2), np.float32), m, v)
    assert not _differs(r, torch.from_numpy(exp).cuda())
    del pm, pv, r
    torch.cuda.empty_cache()


def test_norm2_at_int_max():
    """norm2 through a pointer array (QLA_F3_r_veq_norm2_pV, the gathered
    reduction) on INT_MAX sites: n times the record's norm, the float norms
    added in double, rounded to float."""
    import torch
    n = N_INT_MAX
    need(n * 8)
    v = np.random.default_rng(23).standard_normal((1, 3, 2)).astype(np.float32)
    dv = dev(v)
    pv = torch.full((n,), dv.data_ptr(), dtype=torch.int64, device="cuda")
    r = torch.zeros(1, dtype=torch.float32, device="cuda")
    say(f"QLA_F3_r_veq_norm2_pV n={n}")
    qla_amd.call("QLA_F3_r_veq_norm2_pV", r, pv, n)
    qla_amd.check()
    terms = v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]           # float32, as the kernel forms them
    ref = np.longdouble(n) * terms.astype(np.longdouble).sum()
    got = np.longdouble(float(r.cpu()[0]))
    assert abs(got - ref) <= ref * 2.0 ** -23, (got, ref)
    del pv
    torch.cuda.empty_cache()


@pytest.mark.parametrize("fmt", ["f16"])
def test_compressed_spproj_at_int_max(fmt):
    """Hh = spproj(M12 * pDh) (QLA_MI355X_F3_Hh_veq_spproj_M12_times_pDh) on
    INT_MAX sites: 51.5 GB of contiguous links, one Dh record through the
    pointer array, 60 GB of Hh out, each record equal to the format model's."""
    import torch
    n = N_INT_MAX
    need(n * (24 + 8 + 28) + (1 << 28) * 28)
    u = lattice.random_su3(1, 24).astype(np.float32)
    d = lattice.random_field("D", 1, 25).astype(np.float32)
    link = cm.pack_link(u, fmt)
    drec = cm.to_records(*cm.pack_fermion(d))
    mu, sign = 3, -1
    exp = cm.spproj(link, fmt, drec, np.zeros(1, dtype=np.int64), mu, sign)
    links = torch.from_numpy(np.ascontiguousarray(link).view(np.uint8)).cuda().reshape(1, -1).expand(n, -1).contiguous()
    dd = torch.from_numpy(np.ascontiguousarray(drec).view(np.uint8)).cuda()
    pd = torch.full((n,), dd.data_ptr(), dtype=torch.int64, device="cuda")
    r = torch.full((n, 28), 0xFF, dtype=torch.uint8, device="cuda")
    say(f"QLA_MI355X_F3_Hh_veq_spproj_M12_times_pDh {fmt} n={n}")
    qla_amd.call("QLA_MI355X_F3_Hh_veq_spproj_M12_times_pDh", r, links, {"bf16": 0, "f16": 1}[fmt], pd, mu, sign, n)
    qla_amd.check()
    assert not _differs(r, torch.from_numpy(np.ascontiguousarray(exp).view(np.uint8)).cuda())
    del links, pd, r
    torch.cuda.empty_cache()
