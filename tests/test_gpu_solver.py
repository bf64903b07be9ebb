"""A Wilson-fermion solve driven only through the C ABI, as a QDP/QOP-style
caller would drive QLA: conjugate gradient on the normal equations
D^dagger D x = D^dagger b with

    (D psi)(x) = psi(x) - kappa sum_mu [ U_mu(x) (1 - gamma_mu) psi(x+mu)
                                        + U_mu(x-mu)^dagger (1 + gamma_mu) psi(x-mu) ],

D^dagger the same with the projector signs exchanged.  Each application of
D is four D_v{eq,peq}_spproj_M_times_pD calls (forward neighbours), four
D_vpeq_spproj_pMa_times_pD calls (backward neighbours, links through a
pointer array) and one D_veq_r_times_D_plus_D; each iteration adds the
global reductions (norm2, the real part of the dot) and the axpy forms
D_vpeq_r_times_D / D_vmeq_r_times_D / D_veq_r_times_D_plus_D.

Check: the solution returned by the device solve satisfies the oracle's own
operator on the host, |b - D x| / |b| below the solver target (plus the
rounding of one application of D), so the hop kernels, the reductions and
the field algebra compose into a working solver.
"""
import ctypes

import numpy as np
import pytest

import qla_amd
from qla_amd import lattice
from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu

DIMS = (8, 8, 8, 6)
KAPPA = 0.11


class Wilson:
    """Links, neighbour tables and the pointer arrays of the hop on `DIMS`."""

    def __init__(self, prec, seed=3):
        import torch
        self.prec = prec
        self.dt = np.float64 if prec == "D3" else np.float32
        self.real = ctypes.c_double if prec == "D3" else ctypes.c_float
        self.n = lattice.volume(DIMS)
        self.fw, self.bw = lattice.neighbours(DIMS)
        self.u = [lattice.random_su3(self.n, seed + 10 * mu).astype(self.dt) for mu in range(4)]
        self.du = [dev(u) for u in self.u]
        # backward links U_mu(x - mu) through a pointer array
        self.bl = [qla_amd.pointer_array(self.du[mu], self.bw[mu]) for mu in range(4)]
        self.shape = (self.n, 3, 4, 2)
        self.tdt = torch.float64 if prec == "D3" else torch.float32
        self.tmp = self.field()
        self._ptr = {}

    def field(self):
        import torch
        return torch.zeros(self.shape, dtype=self.tdt, device="cuda")

    def ptrs(self, psi):
        """forward / backward neighbour pointer arrays into psi (cached)."""
        key = psi.data_ptr()
        if key not in self._ptr:
            self._ptr[key] = ([qla_amd.pointer_array(psi, self.fw[mu]) for mu in range(4)],
                              [qla_amd.pointer_array(psi, self.bw[mu]) for mu in range(4)])
        return self._ptr[key]

    def apply(self, r, psi, dagger=False):
        """r = D psi (or D^dagger psi) on the device."""
        lib = qla_amd.load()
        P = self.prec
        s = 1 if dagger else -1
        fwp, bwp = self.ptrs(psi)
        t = self.tmp
        for mu in range(4):
            f = getattr(lib, f"QLA_{P}_D_v{'eq' if mu == 0 else 'peq'}_spproj_M_times_pD")
            f(t.data_ptr(), self.du[mu].data_ptr(), fwp[mu].data_ptr(), mu, s, self.n)
            getattr(lib, f"QLA_{P}_D_vpeq_spproj_pMa_times_pD")(t.data_ptr(), self.bl[mu].data_ptr(),
                                                                bwp[mu].data_ptr(), mu, -s, self.n)
        mk = self.real(-KAPPA)
        getattr(lib, f"QLA_{P}_D_veq_r_times_D_plus_D")(r.data_ptr(), ctypes.addressof(mk), t.data_ptr(),
                                                       psi.data_ptr(), self.n)

    def apply_oracle(self, orc, psi, dagger=False):
        """D psi on the host with the oracle's site operation (kind 3)."""
        s = 1 if dagger else -1
        t = np.zeros_like(psi)
        for mu in range(4):
            orc.site(orc.WILSON, 0 if mu == 0 else 1, t, self.u[mu], psi, mu, s, b_idx=self.fw[mu])
            orc.site(orc.WILSON, 1, t, self.u[mu], psi, mu, -s, adj=True, a_idx=self.bw[mu], b_idx=self.bw[mu])
        r = np.empty_like(psi)
        orc.vecop(orc.R_PLUS, 0, r, t, psi, s=-KAPPA)
        return r


def cgne(w, b, tol2, maxit=1000):
    """x with D^dagger D x = D^dagger b; every field operation is a library call."""
    lib = qla_amd.load()
    P = w.prec
    red = "D3" if P == "D3" else "DF3"
    norm2 = getattr(lib, f"QLA_{red}_r_veq_norm2_D")
    redot = getattr(lib, f"QLA_{red}_r_veq_re_D_dot_D")
    axpy = getattr(lib, f"QLA_{P}_D_vpeq_r_times_D")
    axmy = getattr(lib, f"QLA_{P}_D_vmeq_r_times_D")
    xpay = getattr(lib, f"QLA_{P}_D_veq_r_times_D_plus_D")
    n = w.n
    x, r, p, pn, q, ap = (w.field() for _ in range(6))
    w.apply(r, b, dagger=True)                        # r = D^dagger b (x = 0)
    getattr(lib, f"QLA_{P}_D_veq_D")(p.data_ptr(), r.data_ptr(), n)
    v = ctypes.c_double()
    norm2(ctypes.addressof(v), r.data_ptr(), n)
    rr = v.value
    target = tol2 * rr
    it = 0
    while rr > target and it < maxit:
        w.apply(q, p)
        w.apply(ap, q, dagger=True)
        redot(ctypes.addressof(v), p.data_ptr(), ap.data_ptr(), n)
        alpha = w.real(rr / v.value)
        axpy(x.data_ptr(), ctypes.addressof(alpha), p.data_ptr(), n)
        axmy(r.data_ptr(), ctypes.addressof(alpha), ap.data_ptr(), n)
        norm2(ctypes.addressof(v), r.data_ptr(), n)
        beta = w.real(v.value / rr)
        rr = v.value
        xpay(pn.data_ptr(), ctypes.addressof(beta), p.data_ptr(), r.data_ptr(), n)
        p, pn = pn, p
        it += 1
    qla_amd.check()
    return x, it, rr / (target / tol2)


@pytest.mark.parametrize("prec,tol2,bound", [("D3", 1e-24, 1e-10), ("F3", 1e-12, 1e-4)])
def test_wilson_cgne_solves_the_oracle_operator(orc, prec, tol2, bound):
    qla_amd.set_sync(True)
    w = Wilson(prec)
    b = lattice.random_field("D", w.n, 77).astype(w.dt)
    say(f"{prec} Wilson CGNE on {DIMS}, kappa {KAPPA}")
    x, iters, rel2 = cgne(w, dev(b), tol2)
    assert rel2 <= tol2, (iters, rel2)
    xs = host(x)
    res = b - w.apply_oracle(orc, xs)
    rel = float(np.linalg.norm(res.astype(np.float64)) / np.linalg.norm(b.astype(np.float64)))
    print(f"{prec}: {iters} iterations, |b - Dx|/|b| = {rel:.3e}")
    assert rel < bound, (iters, rel)


def test_device_and_oracle_operators_agree(orc):
    """One application of D and of D^dagger: the device hop sequence equals
    the oracle's bit for bit (the same calls in the same order)."""
    w = Wilson("F3")
    psi = lattice.random_field("D", w.n, 5).astype(np.float32)
    dpsi = dev(psi)
    for dagger in (False, True):
        r = w.field()
        say(f"F3 Wilson D apply dagger={dagger}")
        w.apply(r, dpsi, dagger=dagger)
        qla_amd.check()
        got = host(r)
        exp = w.apply_oracle(orc, psi, dagger=dagger)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), dagger


def test_plaquette_two_ways(orc):
    """sum_x sum_{mu<nu} Re tr U_mu(x) U_nu(x+mu) U_mu(x+nu)^dagger U_nu(x)^dagger
    on 8^3 x 6 through the C ABI, twice:
    (a) X = U_mu * U_nu(x+mu), Y = U_nu * U_mu(x+nu) (M_veq_M_times_pM with a
        pointer array of links), then the global Re X.Y (QLA_DF3_r_veq_re_M_dot_M:
        Re tr Y^dagger X summed);
    (b) Z = X Y^dagger (M_veq_M_times_Ma), the site-local Re tr (R_veq_re_trace_M)
        and the global sum of the real field (QLA_DF_r_veq_sum_R);
    against a complex128 numpy evaluation of the same products."""
    import torch
    lib = qla_amd.load()
    qla_amd.set_sync(True)
    w = Wilson("F3")
    n = w.n
    U = [u.astype(np.float64) for u in w.u]
    Uc = [u[..., 0] + 1j * u[..., 1] for u in U]
    want = 0.0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            x = np.einsum("nij,njk->nik", Uc[mu], Uc[nu][w.fw[mu]])
            y = np.einsum("nij,njk->nik", Uc[nu], Uc[mu][w.fw[nu]])
            want += np.einsum("nij,nij->", x, np.conj(y)).real
    fw_links = [[qla_amd.pointer_array(w.du[nu], w.fw[mu]) for nu in range(4)] for mu in range(4)]
    X = torch.empty((n, 3, 3, 2), dtype=torch.float32, device="cuda")
    Y = torch.empty_like(X)
    Z = torch.empty_like(X)
    tr = torch.empty(n, dtype=torch.float32, device="cuda")
    va = ctypes.c_double()
    vb = ctypes.c_double()
    tot_a = tot_b = 0.0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            say(f"plaquette plane {mu}{nu}")
            lib.QLA_F3_M_veq_M_times_pM(X.data_ptr(), w.du[mu].data_ptr(), fw_links[mu][nu].data_ptr(), n)
            lib.QLA_F3_M_veq_M_times_pM(Y.data_ptr(), w.du[nu].data_ptr(), fw_links[nu][mu].data_ptr(), n)
            lib.QLA_DF3_r_veq_re_M_dot_M(ctypes.addressof(va), Y.data_ptr(), X.data_ptr(), n)
            lib.QLA_F3_M_veq_M_times_Ma(Z.data_ptr(), X.data_ptr(), Y.data_ptr(), n)
            lib.QLA_F3_R_veq_re_trace_M(tr.data_ptr(), Z.data_ptr(), n)
            lib.QLA_DF_r_veq_sum_R(ctypes.addressof(vb), tr.data_ptr(), n)
            tot_a += va.value
            tot_b += vb.value
    qla_amd.check()
    scale = 6 * 3 * n
    print(f"plaquette: {want / scale:.9f} (numpy), {tot_a / scale:.9f} (dot), {tot_b / scale:.9f} (trace + sum)")
    # float products: |error| per site-plane ~ a few float ulps of |tr| <= 3
    assert abs(tot_a - want) <= 1e-5 * scale
    assert abs(tot_b - want) <= 1e-5 * scale
