// qla_matfn_rows.hpp -- exp, sqrt, invsqrt and log of nc x nc complex matrices
// (nc = 5..8; the launcher picks this form where it measured faster) with one
// matrix per group of nc lanes, one row per lane.
//
// The per-thread form (qla_matfn.hpp) holds a whole matrix in each thread:
// at nc = 8 the Pade / Newton temporaries are six 64-complex matrices, more
// than the 512 VGPRs a lane can address, so it spills.  Here lane r of a
// group holds row r of every matrix (8 complexes at nc = 8) and the group
// shares what a row needs of the other rows through its own LDS area.  A
// wavefront holds 64 / nc whole groups (12, 10, 9, 8 at nc = 5..8); its
// last 64 mod nc lanes hold none.  A group's lanes:
//   * a product x = a b: each lane publishes its row of b, then forms
//     x[r][j] = sum_k a[r][k] b[k][j] with x = 0 and k ascending, the
//     reference's order (M_eq_M_times_M, expressions_tensor.pl:544-630);
//   * the Frobenius norm (maxev): each lane publishes its element norms
//     |a[r][j]|^2 (formed in T) and every lane adds all nc^2 of them in
//     double in row-major order, the single-thread order;
//   * Gauss-Jordan with partial pivoting (M_eq_M_inverse_M.c:93-135): the
//     pivot keys |c[i][k]|^2 are published, every lane runs the template's
//     search (largest key, first of equals, rows k.. in order), and instead
//     of exchanging two rows' contents the two lanes exchange their logical
//     row labels; the pivot row is published and every other row eliminates
//     with the template's operations; at the end each lane writes its
//     logical row back to its own lane through LDS.
// Every element therefore sees the same operations in the same order as in
// the per-thread form and the templates: results are bit-identical.
//
// A group's control flow (scaling s, the number of Newton steps) depends
// only on its matrix, so it is uniform within the group; the groups of one
// wave may diverge, and each has its own LDS area.  All lanes of a group are
// in one wavefront, so wave_sync() orders the LDS traffic.
#pragma once
#include "qla_matfn.hpp"

namespace qla {

template <class T, int NC> struct rrow { cpx<T> e[NC]; };

// one group's LDS: a published matrix (product operand / row reorder), the
// pivot row of c and d, the element norms and the pivot keys
template <class T, int NC> struct rgrp {
    cpx<T> m[NC * NC];
    cpx<T> pc[NC], pd[NC];
    T nrm[NC * NC];
    T key[NC];
};

template <class T, int NC> struct rctx {
    rgrp<T, NC> *s;
    int r;          // this lane's row
};

template <class T, int NC> __device__ __forceinline__ void r_eq_d(rrow<T, NC> &x, T d, int r)
{
    LOOP(j, NC) { x.e[j].re = (j == r) ? d : T(0); x.e[j].im = T(0); }
}
template <class T, int NC> __device__ __forceinline__ void r_eq_r_times(rrow<T, NC> &x, T s, const rrow<T, NC> &a)
{
    LOOP(j, NC) { x.e[j].re = s * a.e[j].re; x.e[j].im = s * a.e[j].im; }
}
template <class T, int NC> __device__ __forceinline__ void r_peq_r_times(rrow<T, NC> &x, T s, const rrow<T, NC> &a)
{
    LOOP(j, NC) { x.e[j].re = x.e[j].re + s * a.e[j].re; x.e[j].im = x.e[j].im + s * a.e[j].im; }
}
template <class T, int NC> __device__ __forceinline__ void r_peq(rrow<T, NC> &x, const rrow<T, NC> &a)
{
    LOOP(j, NC) { x.e[j].re = x.e[j].re + a.e[j].re; x.e[j].im = x.e[j].im + a.e[j].im; }
}
template <class T, int NC>
__device__ __forceinline__ void r_eq_minus(rrow<T, NC> &x, const rrow<T, NC> &a, const rrow<T, NC> &b)
{
    LOOP(j, NC) { x.e[j].re = a.e[j].re - b.e[j].re; x.e[j].im = a.e[j].im - b.e[j].im; }
}

// x = a b (x may alias a or b)
template <class T, int NC>
__device__ __forceinline__ void r_times(rrow<T, NC> &x, const rrow<T, NC> &a, const rrow<T, NC> &b,
                                        const rctx<T, NC> &c)
{
    LOOP(j, NC) c.s->m[c.r * NC + j] = b.e[j];
    wave_sync();
    cpx<T> acc[NC];
    LOOP(j, NC) acc[j] = { T(0), T(0) };
    LOOP(k, NC) {
        cpx<T> bk[NC];
        LOOP(j, NC) bk[j] = c.s->m[k * NC + j];
        LOOP(j, NC) acc_mul<false, false>(acc[j], a.e[k], bk[j]);
    }
    wave_sync();
    LOOP(j, NC) x.e[j] = acc[j];
}

// sum over the matrix of |a[i][j]|^2 (formed in T), added in double in
// row-major order (maxev of the templates)
template <class T, int NC> __device__ __forceinline__ double r_maxev2(const rrow<T, NC> &a, const rctx<T, NC> &c)
{
    LOOP(j, NC) c.s->nrm[c.r * NC + j] = norm2c(a.e[j]);
    wave_sync();
    double f = 0;
    LOOP(k, NC * NC) f += (double)c.s->nrm[k];
    wave_sync();
    return f;
}

// x = a^-1 b (IDENT: b = 1), Gauss-Jordan with partial pivoting on |.|^2
template <class T, int NC, bool IDENT>
__device__ __forceinline__ void r_inverse(rrow<T, NC> &x, const rrow<T, NC> &a, const rrow<T, NC> &b,
                                          const rctx<T, NC> &ctx)
{
    rrow<T, NC> c = a, d;
    if (IDENT) r_eq_d(d, T(1), ctx.r);
    else d = b;
    int lr = ctx.r;     // the logical row this lane holds
    LOOP(k, NC) {
        ctx.s->key[lr] = norm2c(c.e[k]);
        wave_sync();
        T rmax = ctx.s->key[k];
        int imax = k;
        LOOP(i, NC) {
            if (i > k) {
                const T rv = ctx.s->key[i];
                if (rv > rmax) { rmax = rv; imax = i; }
            }
        }
        // rows k and imax trade places
        lr = lr == imax ? k : (lr == k ? imax : lr);
        rmax = T(1) / rmax;
        if (lr == k) {
            cpx<T> s;
            s.re = rmax * c.e[k].re;        // QLA_c_eq_r_times_ca
            s.im = -(rmax * c.e[k].im);
            c.e[k] = s;
            LOOP(j, NC) { ctx.s->pc[j] = c.e[j]; ctx.s->pd[j] = d.e[j]; }
        }
        wave_sync();
        if (lr != k) {
            const cpx<T> t = cmul(ctx.s->pc[k], c.e[k]);
            LOOP(j, NC) if (j > k) acc_mul<true, false>(c.e[j], t, ctx.s->pc[j]);
            LOOP(j, NC) acc_mul<true, false>(d.e[j], t, ctx.s->pd[j]);
        }
        wave_sync();
    }
    // x[lr][j] = c[lr][lr] d[lr][j], then every lane takes its own row back
    cpx<T> cii = c.e[0];
    LOOP(j, NC) if (j == lr) cii = c.e[j];
    LOOP(j, NC) ctx.s->m[lr * NC + j] = cmul(cii, d.e[j]);
    wave_sync();
    LOOP(j, NC) x.e[j] = ctx.s->m[ctx.r * NC + j];
    wave_sync();
}

// code-templates/M_eq_exp_M.c:300-367, the Pade(3,3) scaling and squaring of
// f_exp; the last squaring of the binary powering, whose result the template
// discards, is skipped
template <class T, int NC> __device__ __forceinline__ void r_exp(rrow<T, NC> &r, const rrow<T, NC> &a,
                                                                 const rctx<T, NC> &c)
{
    double ds = sqrt(r_maxev2(a, c));
    if (ds == 0) { r_eq_d(r, T(1), c.r); return; }
    unsigned int s = (unsigned int)ceil(2 * ds);
    if (s > 1024) s = 1024;
    ds = (double)s;
    rrow<T, NC> bs, bs2, bs4, bs6, pb, qb;
    const T dsi = (T)(1.0 / ds);
    r_eq_r_times(bs, dsi, a);
    r_eq_d(pb, (T)PADE_P[0], c.r);
    r_eq_d(qb, (T)PADE_Q[0], c.r);
    r_times(bs2, bs, bs, c);
    r_peq_r_times(pb, (T)PADE_P[1], bs2);
    r_peq_r_times(qb, (T)PADE_Q[1], bs2);
    r_times(bs4, bs2, bs2, c);
    r_peq_r_times(pb, (T)PADE_P[2], bs4);
    r_peq_r_times(qb, (T)PADE_Q[2], bs4);
    r_times(bs6, bs4, bs2, c);
    r_peq_r_times(qb, (T)PADE_Q[3], bs6);
    r_times(bs6, bs, pb, c);
    r_eq_minus(pb, qb, bs6);
    r_inverse<T, NC, false>(qb, pb, bs6, c);
    r_eq_d(pb, T(1), c.r);
    r_peq_r_times(pb, T(2), qb);
    if (s == 1) { r = pb; return; }
    rrow<T, NC> va;
    r_eq_d(va, T(1), c.r);
    for (; s > 0; s >>= 1) {
        if (s & 1) r_times(va, va, pb, c);
        if (s > 1) r_times(pb, pb, pb, c);
    }
    r = va;
}

// code-templates/M_eq_sqrt_M.c:136-173 (f_sqrt)
template <class T, int NC> __device__ __forceinline__ void r_sqrt(rrow<T, NC> &r, const rrow<T, NC> &a,
                                                                  const rctx<T, NC> &c)
{
    double ds = r_maxev2(a, c);
    if (ds == 0) { r_eq_d(r, T(0), c.r); return; }
    ds = sqrt(ds);
    rrow<T, NC> x, e, t;
    r_eq_r_times(x, (T)(1 / ds), a);
    r_eq_d(e, (T)0.5, c.r);
    r_peq_r_times(e, (T)-0.5, x);
    r_peq(x, e);
    double enorm;
    const double estop = eps_of<T>::v;
    int nit = 0;
    do {
        nit++;
        r_inverse<T, NC, false>(t, x, e, c);
        r_times(t, e, t, c);
        r_eq_r_times(e, (T)-0.5, t);
        r_peq(x, e);
        enorm = r_maxev2(e, c);
    } while (nit < 20 && enorm > estop);
    r_eq_r_times(r, (T)sqrt(ds), x);
}

// code-templates/M_eq_invsqrt_M.c:135-175 (f_invsqrt)
template <class T, int NC> __device__ __forceinline__ void r_invsqrt(rrow<T, NC> &r, const rrow<T, NC> &a,
                                                                     const rctx<T, NC> &c)
{
    double ds = r_maxev2(a, c);
    if (ds == 0) { r_eq_d(r, (T)(1. / 0.), c.r); return; }
    ds = sqrt(ds);
    rrow<T, NC> x, e, t;
    r_eq_d(x, T(1), c.r);
    r_inverse<T, NC, true>(t, a, a, c);
    r_eq_d(e, (T)-0.5, c.r);
    r_peq_r_times(e, (T)(0.5 * ds), t);
    r_peq(x, e);
    double enorm;
    const double estop = eps_of<T>::v;
    int nit = 0;
    do {
        nit++;
        r_inverse<T, NC, false>(t, x, e, c);
        r_times(t, e, t, c);
        r_eq_r_times(e, (T)-0.5, t);
        r_peq(x, e);
        enorm = r_maxev2(e, c);
    } while (nit < 20 && enorm > estop);
    r_eq_r_times(r, (T)(1 / sqrt(ds)), x);
}

// code-templates/M_eq_log_M.c:49-70, 158-200 (f_log): LOG_NROOTS square
// roots of the scaled matrix by the row form of the Newton iteration, the
// Pade form of log(1 + X) with one Gauss-Jordan inverse, then the scale
template <class T, int NC> __device__ __forceinline__ void r_log(rrow<T, NC> &r, const rrow<T, NC> &a,
                                                                 const rctx<T, NC> &c)
{
    double ds = sqrt(r_maxev2(a, c));
    if (ds == 0) { r_eq_d(r, (T)(-1. / 0.), c.r); return; }
    rrow<T, NC> as, as2, as3, as4, pa, qa;
    r_eq_r_times(as, (T)(1 / ds), a);
    for (int i = 0; i < LOG_NROOTS; i++) {
        as2 = as;
        r_sqrt(as, as2, c);
    }
    LOOP(j, NC) if (j == c.r) as.e[j].re = as.e[j].re - 1;
    r_eq_d(pa, (T)LOG_P[0], c.r);
    r_eq_d(qa, (T)LOG_Q[0], c.r);
    r_peq_r_times(pa, (T)LOG_P[1], as);
    r_peq_r_times(qa, (T)LOG_Q[1], as);
    r_times(as2, as, as, c);
    r_peq_r_times(pa, (T)LOG_P[2], as2);
    r_peq_r_times(qa, (T)LOG_Q[2], as2);
    r_times(as3, as, as2, c);
    r_peq_r_times(pa, (T)LOG_P[3], as3);
    r_peq_r_times(qa, (T)LOG_Q[3], as3);
    r_times(as4, as2, as2, c);
    r_peq_r_times(pa, (T)LOG_P[4], as4);
    r_peq_r_times(qa, (T)LOG_Q[4], as4);
    r_times(as4, as, as4, c);               // as5 of the template
    r_peq_r_times(pa, (T)LOG_P[5], as4);
    r_peq(qa, as4);
    r_inverse<T, NC, true>(as4, qa, qa, c);
    r_times(qa, pa, as4, c);
    r_times(pa, as3, qa, c);
    r_peq_r_times(pa, (T)-0.5, as2);
    r_peq(pa, as);
    r_eq_r_times(r, (T)(1 << LOG_NROOTS), pa);
    ds = log(ds);
    LOOP(j, NC) if (j == c.r) r.e[j].re = (T)((double)r.e[j].re + ds);
}

} // namespace qla
