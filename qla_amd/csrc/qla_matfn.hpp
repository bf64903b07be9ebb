// qla_matfn.hpp -- per-thread nc x nc complex matrix functions, each the
// reference template's algorithm in its operation order
// (lib/perl/code-templates/*.c, cited per function).  Everything is
// evaluated in the operand precision T except where the templates use double
// (scaling norms and factors, the Pade coefficients, the double literals of
// the closed forms), with IEEE-rounded divide and square root and no
// contraction, so the functions that use only these operations agree bit for
// bit with a CPU evaluation of the same order.  The Gauss-Jordan and LU
// pivoting of the reference permutes rows through an index array; here rows
// are swapped in registers with compile-time indices (same values at the same
// logical positions).
#pragma once
#include "qla_dev.hpp"
#include "qla_hostmath.hpp"
#include <float.h>

namespace qla {

template <class T> struct eps_of;
template <> struct eps_of<float> { static constexpr double v = FLT_EPSILON; };
template <> struct eps_of<double> { static constexpr double v = DBL_EPSILON; };

template <class T, int NC> struct mat { cpx<T> e[NC * NC]; };

#define LOOP(i, N) _Pragma("unroll") for (int i = 0; i < (N); i++)

template <class T, int NC> __device__ __forceinline__ void m_eq_d(mat<T, NC> &x, T d)
{
    LOOP(i, NC) LOOP(j, NC) { x.e[i * NC + j].re = (i == j) ? d : T(0); x.e[i * NC + j].im = T(0); }
}
template <class T, int NC> __device__ __forceinline__ void m_eq_r_times_m(mat<T, NC> &r, T s, const mat<T, NC> &a)
{
    LOOP(k, NC * NC) { r.e[k].re = s * a.e[k].re; r.e[k].im = s * a.e[k].im; }
}
template <class T, int NC> __device__ __forceinline__ void m_peq_r_times_m(mat<T, NC> &r, T s, const mat<T, NC> &a)
{
    LOOP(k, NC * NC) { r.e[k].re = r.e[k].re + s * a.e[k].re; r.e[k].im = r.e[k].im + s * a.e[k].im; }
}
template <class T, int NC> __device__ __forceinline__ void m_peq_m(mat<T, NC> &r, const mat<T, NC> &a)
{
    LOOP(k, NC * NC) { r.e[k].re = r.e[k].re + a.e[k].re; r.e[k].im = r.e[k].im + a.e[k].im; }
}
template <class T, int NC>
__device__ __forceinline__ void m_eq_m_minus_m(mat<T, NC> &r, const mat<T, NC> &a, const mat<T, NC> &b)
{
    LOOP(k, NC * NC) { r.e[k].re = a.e[k].re - b.e[k].re; r.e[k].im = a.e[k].im - b.e[k].im; }
}
// r = a b; x = 0, x += a[i][k] b[k][j] for k ascending
template <class T, int NC>
__device__ __forceinline__ void m_eq_m_times_m(mat<T, NC> &r, const mat<T, NC> &a, const mat<T, NC> &b)
{
    LOOP(i, NC) LOOP(j, NC) {
        cpx<T> x = { T(0), T(0) };
        LOOP(k, NC) acc_mul<false, false>(x, a.e[i * NC + k], b.e[k * NC + j]);
        r.e[i * NC + j] = x;
    }
}
template <class T> __device__ __forceinline__ T norm2c(cpx<T> a) { return a.re * a.re + a.im * a.im; }
template <class T> __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
}
template <class T, int NC> __device__ __forceinline__ void swap_rows(mat<T, NC> &m, int i, int k)
{
    LOOP(j, NC) { cpx<T> t = m.e[i * NC + j]; m.e[i * NC + j] = m.e[k * NC + j]; m.e[k * NC + j] = t; }
}

// x = a^-1 b (IDENT: b = identity) by Gauss-Jordan with partial pivoting on |.|^2
template <class T, int NC, bool IDENT>
__device__ __forceinline__ void m_eq_m_inverse_m(mat<T, NC> &x, const mat<T, NC> &a, const mat<T, NC> &b)
{
    mat<T, NC> c = a, d;
    if (IDENT) m_eq_d(d, T(1));
    else d = b;
    LOOP(k, NC) {
        T rmax = norm2c(c.e[k * NC + k]);
        int imax = k;
        LOOP(i, NC) {
            if (i > k) {
                T r = norm2c(c.e[i * NC + k]);
                if (r > rmax) { rmax = r; imax = i; }
            }
        }
        LOOP(i, NC) {
            if (i > k && imax == i) { swap_rows(c, i, k); swap_rows(d, i, k); }
        }
        rmax = T(1) / rmax;
        cpx<T> s;
        s.re = rmax * c.e[k * NC + k].re;       // QLA_c_eq_r_times_ca
        s.im = -(rmax * c.e[k * NC + k].im);
        c.e[k * NC + k] = s;
        LOOP(i, NC) {
            if (i != k) {
                const cpx<T> t = cmul(s, c.e[i * NC + k]);
                LOOP(j, NC) if (j > k) acc_mul<true, false>(c.e[i * NC + j], t, c.e[k * NC + j]);
                LOOP(j, NC) acc_mul<true, false>(d.e[i * NC + j], t, d.e[k * NC + j]);
            }
        }
    }
    LOOP(i, NC) LOOP(j, NC) x.e[i * NC + j] = cmul(c.e[i * NC + i], d.e[i * NC + j]);
}

// M_eq_inverse_M: cofactor formula for nc = 3, Gauss-Jordan otherwise
template <class T, int NC> __device__ __forceinline__ void m_eq_inverse_m(mat<T, NC> &x, const mat<T, NC> &a)
{
    if constexpr (NC != 3) {
        m_eq_m_inverse_m<T, NC, true>(x, a, a);
    } else {
#define A_(i, j) a.e[(i) * 3 + (j)]
        cpx<T> det0, det1, det2, det, idet;
        det2 = cmul(A_(0, 0), A_(1, 1)); acc_mul<true, false>(det2, A_(0, 1), A_(1, 0));
        det1 = cmul(A_(0, 2), A_(1, 0)); acc_mul<true, false>(det1, A_(0, 0), A_(1, 2));
        det0 = cmul(A_(0, 1), A_(1, 2)); acc_mul<true, false>(det0, A_(0, 2), A_(1, 1));
        det = cmul(det2, A_(2, 2)); acc_mul<false, false>(det, det1, A_(2, 1)); acc_mul<false, false>(det, det0, A_(2, 0));
        {   // QLA_c_eq_r_div_c(idet, 1, det)
            const T t = T(1) / norm2c(det);
            idet.re = t * det.re;
            idet.im = -t * det.im;
        }
#define SET(i0, j0, i1, j1, i2, j2)                                                    \
        det = cmul(A_(i1, j1), A_(i2, j2)); acc_mul<true, false>(det, A_(i1, j2), A_(i2, j1)); \
        x.e[(i0) * 3 + (j0)] = cmul(det, idet);
        SET(0, 0, 1, 1, 2, 2);
        SET(0, 1, 2, 1, 0, 2);
        x.e[0 * 3 + 2] = cmul(det0, idet);
        SET(1, 0, 1, 2, 2, 0);
        SET(1, 1, 2, 2, 0, 0);
        x.e[1 * 3 + 2] = cmul(det1, idet);
        SET(2, 0, 1, 0, 2, 1);
        SET(2, 1, 2, 0, 0, 1);
        x.e[2 * 3 + 2] = cmul(det2, idet);
#undef SET
#undef A_
    }
}

// Frobenius norm squared: element norms in T, added in double (maxev2)
template <class T, int NC> __device__ __forceinline__ double maxev2(const mat<T, NC> &a)
{
    double f = 0;
    LOOP(k, NC * NC) f += (double)norm2c(a.e[k]);
    return f;
}

// code-templates/M_eq_exp_M.c:61-72 (double); rounded to T where used
constexpr double PADE_P[3] = { 9.99999999999999999910E-1, 3.02994407707441961300E-2,
                                  1.26177193074810590878E-4 };
constexpr double PADE_Q[4] = { 2.00000000000000000009E0, 2.27265548208155028766E-1,
                                  2.52448340349684104192E-3, 3.00198505138664455042E-6 };

template <class T, int NC> __device__ void f_exp(mat<T, NC> &r, const mat<T, NC> &a)
{
    double ds = sqrt(maxev2(a));
    if (ds == 0) { m_eq_d(r, T(1)); return; }
    unsigned int s = (unsigned int)ceil(2 * ds);
    if (s > 1024) s = 1024;
    ds = (double)s;
    mat<T, NC> bs, bs2, bs4, bs6, pb, qb;
    const T dsi = (T)(1.0 / ds);
    m_eq_r_times_m(bs, dsi, a);
    m_eq_d(pb, (T)PADE_P[0]);
    m_eq_d(qb, (T)PADE_Q[0]);
    m_eq_m_times_m(bs2, bs, bs);
    m_peq_r_times_m(pb, (T)PADE_P[1], bs2);
    m_peq_r_times_m(qb, (T)PADE_Q[1], bs2);
    m_eq_m_times_m(bs4, bs2, bs2);
    m_peq_r_times_m(pb, (T)PADE_P[2], bs4);
    m_peq_r_times_m(qb, (T)PADE_Q[2], bs4);
    m_eq_m_times_m(bs6, bs4, bs2);
    m_peq_r_times_m(qb, (T)PADE_Q[3], bs6);
    m_eq_m_times_m(bs6, bs, pb);
    m_eq_m_minus_m(pb, qb, bs6);
    m_eq_m_inverse_m<T, NC, false>(qb, pb, bs6);
    m_eq_d(pb, T(1));
    m_peq_r_times_m(pb, T(2), qb);
    if (s == 1) { r = pb; return; }
    // binary powering: va = prod of the squares vb selected by the bits of s
    // (M_eq_exp_M.c:354-365; the square formed after the top bit is never
    // used, so it is skipped)
    mat<T, NC> va, tmp;
    m_eq_d(va, T(1));
    for (; s > 0; s >>= 1) {
        if (s & 1) { m_eq_m_times_m(tmp, va, pb); va = tmp; }
        if (s > 1) { m_eq_m_times_m(tmp, pb, pb); pb = tmp; }
    }
    r = va;
}

template <class T, int NC> __device__ void f_sqrt(mat<T, NC> &r, const mat<T, NC> &a)
{
    double ds = maxev2(a);
    if (ds == 0) { m_eq_d(r, T(0)); return; }
    ds = sqrt(ds);
    mat<T, NC> x, e, t1, t2;
    m_eq_r_times_m(x, (T)(1 / ds), a);
    m_eq_d(e, (T)0.5);
    m_peq_r_times_m(e, (T)-0.5, x);
    m_peq_m(x, e);
    double enorm;
    const double estop = eps_of<T>::v;
    int nit = 0;
    do {
        nit++;
        m_eq_m_inverse_m<T, NC, false>(t1, x, e);
        m_eq_m_times_m(t2, e, t1);
        m_eq_r_times_m(e, (T)-0.5, t2);
        m_peq_m(x, e);
        enorm = maxev2(e);
    } while (nit < 20 && enorm > estop);
    m_eq_r_times_m(r, (T)sqrt(ds), x);
}

template <class T, int NC> __device__ void f_invsqrt(mat<T, NC> &r, const mat<T, NC> &a)
{
    double ds = maxev2(a);
    if (ds == 0) { m_eq_d(r, (T)(1. / 0.)); return; }
    ds = sqrt(ds);
    mat<T, NC> x, e, xi, t1, t2;
    m_eq_d(x, T(1));
    m_eq_inverse_m(xi, a);
    m_eq_d(e, (T)-0.5);
    m_peq_r_times_m(e, (T)(0.5 * ds), xi);
    m_peq_m(x, e);
    double enorm;
    const double estop = eps_of<T>::v;
    int nit = 0;
    do {
        nit++;
        m_eq_m_inverse_m<T, NC, false>(t1, x, e);
        m_eq_m_times_m(t2, e, t1);
        m_eq_r_times_m(e, (T)-0.5, t2);
        m_peq_m(x, e);
        enorm = maxev2(e);
    } while (nit < 20 && enorm > estop);
    m_eq_r_times_m(r, (T)(1 / sqrt(ds)), x);
}

// ---------------------------------------------------------------------------
// Scalar helpers of the closed forms: the C99 functions the templates call
// (sqrtP, fabsP, acosP, cbrtP and the sincos of QLA_*_cexpi,
// lib/cmath/QLA_cexpi.c:12-31).  sqrt is IEEE-rounded on both sides.  In
// single precision acos, cbrt and sincos are the host C library's own
// algorithms restated (qla_hostmath.hpp), so the F3 closed forms return the
// reference's bits; sincos of |x| >= 120 and the double-precision functions
// are the device library's (ocml), faithful to an ulp or two, so D3 results
// that pass through them agree with a CPU evaluation to rounding.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float m_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double m_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float m_fabs(float x) { return fabsf(x); }
__device__ __forceinline__ double m_fabs(double x) { return fabs(x); }
__device__ __forceinline__ float m_acos(float x) { return hm_acosf(x); }
__device__ __forceinline__ double m_acos(double x) { return acos(x); }
__device__ __forceinline__ float m_cbrt(float x) { return hm_cbrtf(x); }
__device__ __forceinline__ double m_cbrt(double x) { return cbrt(x); }
__device__ __forceinline__ void m_sincos(float x, float *s, float *c)
{
    if (!hm_sincosf(x, s, c)) sincosf(x, s, c);
}
__device__ __forceinline__ void m_sincos(double x, double *s, double *c) { sincos(x, s, c); }
// QLA_*_cexpi(t) = (cos t, sin t); the argument is rounded to T first
template <class T> __device__ __forceinline__ cpx<T> cexpi(T t)
{
    T s, c;
    m_sincos(t, &s, &c);
    return { c, s };
}
template <class T> __device__ __forceinline__ void cmsub(cpx<T> &c, cpx<T> a, cpx<T> b) { acc_mul<true, false>(c, a, b); }
template <class T> __device__ __forceinline__ void cmadd(cpx<T> &c, cpx<T> a, cpx<T> b) { acc_mul<false, false>(c, a, b); }
// QLA_c_eq_r_div_c(c, 1, b): t = 1/|b|^2 in T, c = (t b.re, -t b.im)
template <class T> __device__ __forceinline__ cpx<T> rdivc1(cpx<T> b)
{
    const T t = T(1) / norm2c(b);
    return { t * b.re, -t * b.im };
}

// The three 2x2 minors of rows 0, 1 shared by the nc = 3 closed forms
// (C_eq_det_M.c:71-76, V_eq_M_inverse_V.c:103-108, M_eq_expA_M.c:262-267).
template <class T>
__device__ __forceinline__ void minors3(cpx<T> &det0, cpx<T> &det1, cpx<T> &det2, cpx<T> a00, cpx<T> a01, cpx<T> a02,
                                        cpx<T> a10, cpx<T> a11, cpx<T> a12)
{
    det2 = cmul(a00, a11); cmsub(det2, a01, a10);
    det1 = cmul(a02, a10); cmsub(det1, a00, a12);
    det0 = cmul(a01, a12); cmsub(det0, a02, a11);
}

// ---- C_eq_det_M (code-templates/C_eq_det_M.c:70-141) ----------------------
// nc = 3: cofactor expansion along row 2; otherwise LU with partial pivoting
// on |.|^2 (the reference pivots through a row index; rows are swapped here,
// which leaves every logical element where the index would find it).
template <class T, int NC> __device__ void f_det(cpx<T> &x, const mat<T, NC> &a)
{
    if constexpr (NC == 3) {
#define A_(i, j) a.e[(i) * 3 + (j)]
        cpx<T> det0, det1, det2;
        minors3(det0, det1, det2, A_(0, 0), A_(0, 1), A_(0, 2), A_(1, 0), A_(1, 1), A_(1, 2));
        x = cmul(det2, A_(2, 2));
        cmadd(x, det1, A_(2, 1));
        cmadd(x, det0, A_(2, 0));
#undef A_
    } else {
        mat<T, NC> c = a;
        int nswaps = 0;
        x = { T(1), T(0) };
        LOOP(j, NC) {
            LOOP(i, NC) {
                if (j > 0 && i >= j) {
                    cpx<T> t2 = c.e[i * NC + j];
                    LOOP(k, NC) if (k < j) cmsub(t2, c.e[i * NC + k], c.e[k * NC + j]);
                    c.e[i * NC + j] = t2;
                }
            }
            T rmax = norm2c(c.e[j * NC + j]);
            int kmax = j;
            LOOP(k, NC) {
                if (k > j) {
                    const T r = norm2c(c.e[k * NC + j]);
                    if (r > rmax) { rmax = r; kmax = k; }
                }
            }
            if (rmax == T(0)) { x = { T(0), T(0) }; return; }
            LOOP(k, NC) if (k > j && kmax == k) { swap_rows(c, j, k); nswaps++; }
            x = cmul(x, c.e[j * NC + j]);
            const T ri = T(1) / rmax;
            const cpx<T> cjji = { ri * c.e[j * NC + j].re, -(ri * c.e[j * NC + j].im) };   // QLA_c_eq_r_times_ca
            LOOP(i, NC) {
                if (i > j) {
                    cpx<T> t2 = c.e[j * NC + i];
                    LOOP(k, NC) if (k < j) cmsub(t2, c.e[j * NC + k], c.e[k * NC + i]);
                    c.e[j * NC + i] = cmul(t2, cjji);
                }
            }
        }
        if (nswaps & 1) { x.re = -x.re; x.im = -x.im; }
    }
}

// ---- V_eq_M_inverse_V (code-templates/V_eq_M_inverse_V.c:86-166) ---------
// nc = 3: adjugate times b over the determinant; otherwise Gauss-Jordan with
// partial pivoting on |.|^2 carried along the right-hand side.
template <class T, int NC> __device__ void f_inverse_v(cpx<T> *x, const mat<T, NC> &a, const cpx<T> *b)
{
    if constexpr (NC == 3) {
#define A_(i, j) a.e[(i) * 3 + (j)]
        cpx<T> r[3], det0, det1, det2, det, idet;
        minors3(det0, det1, det2, A_(0, 0), A_(0, 1), A_(0, 2), A_(1, 0), A_(1, 1), A_(1, 2));
        det = cmul(det2, A_(2, 2));
        cmadd(det, det1, A_(2, 1));
        cmadd(det, det0, A_(2, 0));
        idet = rdivc1(det);
#define SET_(i0, j0, i1, j1, i2, j2)                                                           \
        det = cmul(A_(i1, j1), A_(i2, j2)); cmsub(det, A_(i1, j2), A_(i2, j1)); cmadd(r[i0], det, b[j0]);
        r[0] = cmul(det0, b[2]);
        SET_(0, 0, 1, 1, 2, 2);
        SET_(0, 1, 2, 1, 0, 2);
        r[1] = cmul(det1, b[2]);
        SET_(1, 0, 1, 2, 2, 0);
        SET_(1, 1, 2, 2, 0, 0);
        r[2] = cmul(det2, b[2]);
        SET_(2, 0, 1, 0, 2, 1);
        SET_(2, 1, 2, 0, 0, 1);
#undef SET_
#undef A_
        LOOP(i, 3) x[i] = cmul(r[i], idet);
    } else {
        mat<T, NC> c = a;
        cpx<T> d[NC];
        LOOP(i, NC) d[i] = b[i];
        LOOP(k, NC) {
            T rmax = norm2c(c.e[k * NC + k]);
            int imax = k;
            LOOP(i, NC) {
                if (i > k) {
                    const T r = norm2c(c.e[i * NC + k]);
                    if (r > rmax) { rmax = r; imax = i; }
                }
            }
            LOOP(i, NC) {
                if (i > k && imax == i) {
                    swap_rows(c, i, k);
                    const cpx<T> t = d[i]; d[i] = d[k]; d[k] = t;
                }
            }
            rmax = T(1) / rmax;
            cpx<T> s;
            s.re = rmax * c.e[k * NC + k].re;
            s.im = -(rmax * c.e[k * NC + k].im);
            c.e[k * NC + k] = s;
            LOOP(i, NC) {
                if (i != k) {
                    const cpx<T> t = cmul(s, c.e[i * NC + k]);
                    LOOP(j, NC) if (j > k) cmsub(c.e[i * NC + j], t, c.e[k * NC + j]);
                    cmsub(d[i], t, d[k]);
                }
            }
        }
        LOOP(i, NC) x[i] = cmul(c.e[i * NC + i], d[i]);
    }
}

// ---- expA / expTA, nc = 3 (code-templates/M_eq_expA_M.c:135-297,
// M_eq_expTA_M.c:219-286): exp(Q) = f0 + f1 Q + f2 Q^2 from the eigenvalues
// of the traceless anti-Hermitian Q (roots of x^3 - p2/2 x - det = 0 with
// x = -i lambda).  The double literals of the template are kept: each such
// product is formed in double and rounded to T, as C's conversions do there.
// ---------------------------------------------------------------------------
template <class T> __device__ void getfs_exp(cpx<T> &f0, cpx<T> &f1, cpx<T> &f2, T p2, T det)
{
    const T q = (1. / 6.) * p2;
    const T r = -0.5 * det;
    const T sq = m_sqrt(m_fabs(q));
    const T sq3 = sq * sq * sq;
    if (sq3 > m_fabs(r)) {
        const T t = m_acos(r / sq3);
        const cpx<T> sct = cexpi<T>((1. / 3.) * t);
        const T sqc = sq * sct.re;
        const T sqs = 1.73205080756887729352 * sq * sct.im;
        const T l0 = -2 * sqc;
        const T l1 = sqc + sqs;
        const T l2 = sqc - sqs;
        const T d10 = 0.5 * (l1 - l0);
        const T d20 = 0.5 * (l2 - l0);
        const T d21 = l2 - l1;
        cpx<T> eid10, g10, eid20, eis02, g20, ff0;
        const cpx<T> eis01 = cexpi<T>(0.5 * (l0 + l1));
        if (d10 == 0) {
            eid10 = { T(1), T(0) };
            g10 = eis01;
        } else {
            eid10 = cexpi<T>(d10);
            const T sc10 = eid10.im / d10;
            g10 = { sc10 * eis01.re, sc10 * eis01.im };
        }
        ff0.re = eis01.re * eid10.re + eis01.im * eid10.im;     // QLA_c_eq_c_times_ca
        ff0.im = eis01.im * eid10.re - eis01.re * eid10.im;
        if (d20 == 0) {
            eid20 = { T(1), T(0) };
            eis02 = ff0;
            g20 = eis02;
        } else {
            eid20 = cexpi<T>(d20);
            eis02 = cmul(ff0, eid20);
            const T sc20 = eid20.im / d20;
            g20 = { sc20 * eis02.re, sc20 * eis02.im };
        }
        const cpx<T> g2010 = { g20.re - g10.re, g20.im - g10.im };
        cpx<T> b2;
        if (d21 == 0) {
            b2 = { T(0), T(0) };
        } else {
            const T d21i = 1 / d21;
            b2 = { d21i * g2010.re, d21i * g2010.im };
        }
        f2 = { b2.im, -b2.re };
        const T m01 = -(l0 + l1);
        f1 = { m01 * b2.re + g10.re, m01 * b2.im + g10.im };
        const cpx<T> f10 = { -g10.im, g10.re };
        const cpx<T> z = { l1 * f2.re + f10.re, l1 * f2.im + f10.im };
        const T ml0 = -l0;
        f0 = { ml0 * z.re + ff0.re, ml0 * z.im + ff0.im };
    } else if (r == 0) {
        // the template's "zero matrix" branch: all three coefficients zero
        f0 = { T(0), T(0) };
        f1 = { T(0), T(0) };
        f2 = { T(0), T(0) };
    } else {
        const T a = -m_cbrt(r);
        const cpx<T> eia2 = cexpi<T>(0.5 * a);
        const cpx<T> eia = cmul(eia2, eia2);
        const cpx<T> ei3a2 = cmul(eia, eia2);
        const T snc = ei3a2.im / (1.5 * a);
        cpx<T> t = { snc * eia2.re, snc * eia2.im };
        f1 = { -t.im, t.re };
        f2 = { T(0), T(0) };
        t.re = ei3a2.re;
        t.im = (-1. / 3.) * ei3a2.im;
        f0 = cmul(eia2, t);
    }
}

template <class T, bool TRACELESS> __device__ void f_expA3(mat<T, 3> &e, const mat<T, 3> &a)
{
    mat<T, 3> iq = a, mqq;
    cpx<T> f;
    if constexpr (!TRACELESS) {
        const T tr = iq.e[0].im + iq.e[4].im + iq.e[8].im;
        const T s = (1. / 3.) * tr;
        iq.e[0].im = iq.e[0].im - s;
        iq.e[4].im = iq.e[4].im - s;
        iq.e[8].im = iq.e[8].im - s;
        f = cexpi<T>(s);
    }
    LOOP(i, 3) LOOP(j, 3) {
        cpx<T> x = cmul(iq.e[i * 3 + 0], iq.e[0 * 3 + j]);
        cmadd(x, iq.e[i * 3 + 1], iq.e[1 * 3 + j]);
        cmadd(x, iq.e[i * 3 + 2], iq.e[2 * 3 + j]);
        mqq.e[i * 3 + j] = x;
    }
    const T p2 = -mqq.e[0].re - mqq.e[4].re - mqq.e[8].re;
    cpx<T> det0, det1, det2;
    minors3(det0, det1, det2, iq.e[0], iq.e[1], iq.e[2], iq.e[3], iq.e[4], iq.e[5]);
    T det = -det2.re * iq.e[8].im - det2.im * iq.e[8].re;          // QLA_r_eqm_Im_c_times_c
    det = det - (det1.re * iq.e[7].im + det1.im * iq.e[7].re);    // QLA_r_meq_Im_c_times_c
    det = det - (det0.re * iq.e[6].im + det0.im * iq.e[6].re);
    cpx<T> a0, a1, a2;
    if constexpr (!TRACELESS) {
        cpx<T> f0, f1, f2;
        getfs_exp(f0, f1, f2, p2, det);
        a0 = cmul(f, f0);
        a1 = cmul(f, f1);
        a2 = cmul(f, f2);
    } else {
        getfs_exp(a0, a1, a2, p2, det);
    }
    LOOP(k, 9) {
        e.e[k] = cmul(a2, mqq.e[k]);
        cmadd(e.e[k], a1, iq.e[k]);
    }
    LOOP(i, 3) { e.e[i * 4].re = e.e[i * 4].re + a0.re; e.e[i * 4].im = e.e[i * 4].im + a0.im; }
}

// ---- sqrtPH / invsqrtPH, nc = 3 (code-templates/M_eq_sqrtPH_M.c:83-206,
// M_eq_invsqrtPH_M.c:83-210): f(M) = f0 + f1 M + f2 M^2 from the eigenvalues
// of the positive Hermitian M.
// ---------------------------------------------------------------------------
template <class T, bool INV> __device__ void getfs_sqrt(T &f0, T &f1, T &f2, T tr, T p2, T det)
{
    const T tr3 = (1. / 3.) * tr;
    const T p23 = (1. / 3.) * p2;
    const T tr32 = tr3 * tr3;
    const T q = 0.5 * (p23 - tr32);
    const T r = 0.25 * tr3 * (5 * tr32 - p2) - 0.5 * det;
    const T sq = m_sqrt(m_fabs(q));
    const T sq3 = sq * sq * sq;
    if (sq3 > m_fabs(r)) {
        const T t = m_acos(r / sq3);
        const cpx<T> sct = cexpi<T>((1. / 3.) * t);
        const T sqc = sq * sct.re;
        const T sqs = 1.73205080756887729352 * sq * sct.im;
        const T l0 = tr3 - 2 * sqc;
        const T ll = tr3 + sqc;
        const T l1 = ll + sqs;
        const T l2 = ll - sqs;
        const T sl0 = m_sqrt(m_fabs(l0));
        const T sl1 = m_sqrt(m_fabs(l1));
        const T sl2 = m_sqrt(m_fabs(l2));
        const T u = sl0 + sl1 + sl2;
        const T w = sl0 * sl1 * sl2;
        if constexpr (!INV) {
            const T v = sl0 * (sl1 + sl2) + sl1 * sl2;
            const T d = (sl0 + sl1) * (sl0 + sl2) * (sl1 + sl2);
            const T di = 1. / d;
            f0 = u * w * di;
            f1 = (tr + v) * di;
            f2 = -di;
        } else {
            const T d = w * (sl0 + sl1) * (sl0 + sl2) * (sl1 + sl2);
            const T di = 1. / d;
            f0 = (w * u * u + l0 * sl0 * (l1 + l2) + l1 * sl1 * (l0 + l2) + l2 * sl2 * (l0 + l1)) * di;
            f1 = -(tr * u + w) * di;
            f2 = u * di;
        }
    } else if (r == 0) {
        if constexpr (!INV) f0 = m_sqrt(m_fabs(tr3));
        else f0 = 1 / m_sqrt(m_fabs(tr3));
        f1 = 0;
        f2 = 0;
    } else {
        const T a = -m_cbrt(r);
        const T l0 = tr3 + 2 * a;
        const T l1 = tr3 - a;
        const T sl0 = m_sqrt(m_fabs(l0));
        const T sl1 = m_sqrt(m_fabs(l1));
        const T sl01 = sl0 + sl1;
        if constexpr (!INV) {
            f2 = 0;
            f1 = (sl01 == 0) ? T(0) : 1 / sl01;
            f0 = sl0 * sl1 * f1;
        } else {
            const T p01 = sl0 * sl1;
            const T t = p01 * sl01;
            const T ti = 1 / t;
            f2 = 0;
            f1 = -ti;
            f0 = (p01 + l0 + l1) * ti;
        }
    }
}

template <class T, bool INV> __device__ void f_sqrtPH3(mat<T, 3> &fn, const mat<T, 3> &m)
{
    mat<T, 3> mm;
    const T tr = m.e[0].re + m.e[4].re + m.e[8].re;
    LOOP(i, 3) LOOP(j, 3) {
        cpx<T> x = cmul(m.e[i * 3 + 0], m.e[0 * 3 + j]);
        cmadd(x, m.e[i * 3 + 1], m.e[1 * 3 + j]);
        cmadd(x, m.e[i * 3 + 2], m.e[2 * 3 + j]);
        mm.e[i * 3 + j] = x;
    }
    const T p2 = mm.e[0].re + mm.e[4].re + mm.e[8].re;
    cpx<T> det0, det1, det2;
    minors3(det0, det1, det2, m.e[0], m.e[1], m.e[2], m.e[3], m.e[4], m.e[5]);
    T det = det2.re * m.e[8].re - det2.im * m.e[8].im;              // QLA_r_eq_Re_c_times_c
    det = det + (det1.re * m.e[7].re - det1.im * m.e[7].im);        // QLA_r_peq_Re_c_times_c
    det = det + (det0.re * m.e[6].re - det0.im * m.e[6].im);
    T f0, f1, f2;
    getfs_sqrt<T, INV>(f0, f1, f2, tr, p2, det);
    LOOP(k, 9) {
        fn.e[k] = { f2 * mm.e[k].re, f2 * mm.e[k].im };                  // QLA_c_eq_r_times_c
        fn.e[k].re = fn.e[k].re + f1 * m.e[k].re;                         // QLA_c_peq_r_times_c
        fn.e[k].im = fn.e[k].im + f1 * m.e[k].im;
    }
    LOOP(i, 3) fn.e[i * 4].re = fn.e[i * 4].re + f0;                      // QLA_c_peq_r
}

// ---- V_eq_eigenvalsH_M, nc = 3 (code-templates/V_eq_eigenvalsH_M.c:103-161)
// The three real roots of the shifted characteristic polynomial, in the
// template's order (not sorted), imaginary parts zero.
template <class T> __device__ void f_eigenvalsH3(cpx<T> *v, const mat<T, 3> &m)
{
    cpx<T> a00 = m.e[0], a11 = m.e[4], a22 = m.e[8], det0, det1, det2;
    const T s1 = (1. / 3.) * (a00.re + a11.re + a22.re);
    a00.re = a00.re - s1;
    a11.re = a11.re - s1;
    a22.re = a22.re - s1;
    det2 = cmul(a00, a11); cmsub(det2, m.e[1], m.e[3]);
    det1 = cmul(m.e[2], m.e[3]); cmsub(det1, a00, m.e[5]);
    det0 = cmul(m.e[1], m.e[5]); cmsub(det0, m.e[2], a11);
    T det = det2.re * a22.re - det2.im * a22.im;
    det = det + (det1.re * m.e[7].re - det1.im * m.e[7].im);
    det = det + (det0.re * m.e[6].re - det0.im * m.e[6].im);
    T s2 = det2.re;
    s2 = s2 + (a00.re * a22.re - a00.im * a22.im);
    s2 = s2 - (m.e[2].re * m.e[6].re - m.e[2].im * m.e[6].im);
    s2 = s2 + (a11.re * a22.re - a11.im * a22.im);
    s2 = s2 - (m.e[5].re * m.e[7].re - m.e[5].im * m.e[7].im);
    const T q = (-1. / 3.) * s2;
    const T r = -0.5 * det;
    const T sq = m_sqrt(m_fabs(q));
    const T sq3 = sq * sq * sq;
    if (sq3 > m_fabs(r)) {
        const T t = m_acos(r / sq3);
        const cpx<T> sct = cexpi<T>((1. / 3.) * t);
        const T sqc = sq * sct.re;
        const T sqs = 1.73205080756887729352 * sq * sct.im;
        const T e0 = s1 - 2 * sqc;
        const T ee = s1 + sqc;
        v[0] = { e0, T(0) };
        v[1] = { ee + sqs, T(0) };
        v[2] = { ee - sqs, T(0) };
    } else {
        const T a = -m_cbrt(r);
        v[0] = { s1 + 2 * a, T(0) };
        v[1] = { s1 - a, T(0) };
        v[2] = { s1 - a, T(0) };
    }
}

// ---- V_eq_eigenvalsH_M, nc >= 4 (code-templates/V_eq_eigenvalsH_M.c:80-101,
// 180-484) ----
// Hessenberg form by Gaussian elimination with row-and-column pivoting, then
// the template's complex implicit double-shift QR on rows l..nn: a small
// subdiagonal element deflates one root, a 2x2 block two (solved from its
// real trace and the real part of its determinant), exceptional shifts at
// iterations 10 and 20.  Where the template, after 100 iterations on one
// root, copies the remaining diagonal plus the shift into the result and
// searches again without bound, the search is repeated at most EIGH_RESTARTS
// times and then stops with that copy; the template's store of a zero one row
// past the matrix (element (k+2, k-1) for k = nn-1 = nc-2) is dropped.
// The row and column indices of the QR steps are data-dependent, so the
// working matrix lives in the thread's private memory.
constexpr int EIGH_RESTARTS = 8;

template <class T> __device__ __forceinline__ cpx<T> cdiv_q(cpx<T> a, cpx<T> b)
{
    const T t = (T)(1.0 / (b.re * b.re + b.im * b.im));           // QLA_c_eq_c_div_c
    return { (a.re * b.re + a.im * b.im) * t, (a.im * b.re - a.re * b.im) * t };
}
template <class T> __device__ __forceinline__ T cabs_q(cpx<T> z) { return m_sqrt(z.re * z.re + z.im * z.im); }

template <class T>
__device__ __forceinline__ void eig2x2(cpx<T> &v0, cpx<T> &v1, cpx<T> m00, cpx<T> m01, cpx<T> m10, cpx<T> m11)
{
    T det = m00.re * m11.re - m00.im * m11.im;
    det = det - (m01.re * m10.re - m01.im * m10.im);
    const T tr = m00.re + m11.re;
    if (det == 0) {
        v0 = { T(0), T(0) };
        v1 = { tr, T(0) };
    } else {
        const T s = (T)(0.5 * tr);
        const T d = s * s - det;
        const T sd = m_sqrt(m_fabs(d));
        const T l1 = s + copysign(sd, s);
        v0 = { det / l1, T(0) };
        v1 = { l1, T(0) };
    }
}

template <class T, int NC> __device__ void f_eigenvalsH_qr(cpx<T> *v, const mat<T, NC> &a)
{
    cpx<T> h[NC * NC];
#define H_(i, j) h[(i) * NC + (j)]
    LOOP(k, NC * NC) h[k] = a.e[k];
    for (int k = 0; k < NC - 2; k++) {
        T rmax = norm2c(H_(k + 1, k));
        int imax = k + 1;
        for (int i = k + 2; i < NC; i++) {
            const T n2 = norm2c(H_(i, k));
            if (n2 > rmax) { rmax = n2; imax = i; }
        }
        if (rmax == 0) continue;
        for (int j = 0; j < NC; j++) { const cpx<T> t = H_(k + 1, j); H_(k + 1, j) = H_(imax, j); H_(imax, j) = t; }
        for (int i = 0; i < NC; i++) { const cpx<T> t = H_(i, k + 1); H_(i, k + 1) = H_(i, imax); H_(i, imax) = t; }
        rmax = T(1) / rmax;
        const cpx<T> s = { rmax * H_(k + 1, k).re, -(rmax * H_(k + 1, k).im) };   // QLA_c_eq_r_times_ca
        for (int i = k + 2; i < NC; i++) {
            const cpx<T> t = cmul(s, H_(i, k));
            H_(i, k) = { T(0), T(0) };
            for (int j = k + 1; j < NC; j++) acc_mul<true, false>(H_(i, j), t, H_(k + 1, j));
            for (int j = 0; j < NC; j++) acc_mul<false, false>(H_(j, k + 1), t, H_(j, i));
        }
    }
    T anorm = 0;
    for (int i = 0; i < NC; i++)
        for (int j = (i == 0 ? 0 : i - 1); j < NC; j++) anorm = anorm + cabs_q(H_(i, j));
    int nn = NC - 1, restarts = 0;
    cpx<T> t = { T(0), T(0) };
    while (nn >= 0) {
        int l, its = 0;
        do {
            for (l = nn; l > 0; l--) {
                T s = cabs_q(H_(l - 1, l - 1)) + cabs_q(H_(l, l));
                if (s == 0) s = anorm;
                const T o = cabs_q(H_(l, l - 1));
                if (s + o == s) { H_(l, l - 1) = { T(0), T(0) }; break; }
            }
            cpx<T> x = H_(nn, nn);
            if (l == nn) {
                v[nn] = { x.re + t.re, x.im + t.im };
                nn--;
            } else if (l == nn - 1) {
                eig2x2(v[nn - 1], v[nn], H_(nn - 1, nn - 1), H_(nn - 1, nn), H_(nn, nn - 1), x);
                v[nn - 1].re = v[nn - 1].re + t.re; v[nn - 1].im = v[nn - 1].im + t.im;
                v[nn].re = v[nn].re + t.re; v[nn].im = v[nn].im + t.im;
                nn -= 2;
            } else {
                if (its >= 100) {
                    for (int i = 0; i <= nn; i++) v[i] = { H_(i, i).re + t.re, H_(i, i).im + t.im };
                    if (++restarts > EIGH_RESTARTS) nn = -1;
                    break;
                }
                cpx<T> y = H_(nn - 1, nn - 1), w = cmul(H_(nn, nn - 1), H_(nn - 1, nn));
                if (its == 10 || its == 20) {           // exceptional shift
                    t.re = t.re + x.re; t.im = t.im + x.im;
                    for (int i = 0; i < nn; i++) { H_(i, i).re = H_(i, i).re - x.re; H_(i, i).im = H_(i, i).im - x.im; }
                    H_(nn, nn) = { T(0), T(0) };
                    const T s = cabs_q(H_(nn, nn - 1)) + cabs_q(H_(nn - 1, nn - 2));
                    x = { (T)(0.75 * s), T(0) };
                    y = x;
                    w = { (T)(-0.4375 * s * s), T(0) };
                }
                its++;
                const int m = l;
                cpx<T> p, q, r, rr, ss, t1;
                const cpx<T> z = H_(m, m);
                rr = { x.re - z.re, x.im - z.im };
                ss = { y.re - z.re, y.im - z.im };
                p = cmul(rr, ss);
                p.re = p.re - w.re; p.im = p.im - w.im;
                acc_mul<false, false>(p, H_(m + 1, m), H_(m, m + 1));
                t1 = { H_(m + 1, m + 1).re - z.re, H_(m + 1, m + 1).im - z.im };
                t1.re = t1.re - (rr.re + ss.re); t1.im = t1.im - (rr.im + ss.im);
                q = cmul(H_(m + 1, m), t1);
                r = cmul(H_(m + 1, m), H_(m + 2, m + 1));
                for (int k = m; k < nn; k++) {
                    T xx = 0;
                    if (k != m) {
                        p = H_(k, k - 1);
                        q = H_(k + 1, k - 1);
                        r = { T(0), T(0) };
                        if (k + 1 != nn) r = H_(k + 2, k - 1);
                        xx = cabs_q(p) + cabs_q(q) + cabs_q(r);
                        if (xx != 0) {
                            const T xxi = T(1) / xx;
                            p = { xxi * p.re, xxi * p.im };
                            q = { xxi * q.re, xxi * q.im };
                            r = { xxi * r.re, xxi * r.im };
                        }
                    }
                    const T p2 = norm2c(p);
                    const T s2 = p2 + norm2c(q) + norm2c(r);
                    if (s2 == 0) continue;
                    cpx<T> s;
                    if (p2 > 0) {
                        const double f = sqrt((double)(s2 / p2));      // C's double sqrt of a T
                        s = { (T)(f * p.re), (T)(f * p.im) };
                    } else {
                        s = { (T)sqrt((double)s2), T(0) };
                    }
                    if (k != m) {
                        H_(k, k - 1) = { -xx * s.re, -xx * s.im };
                        H_(k + 1, k - 1) = { T(0), T(0) };
                        if (k + 2 < NC) H_(k + 2, k - 1) = { T(0), T(0) };
                    }
                    p.re = p.re + s.re; p.im = p.im + s.im;
                    x = cdiv_q(p, s);
                    y = cdiv_q(q, s);
                    w = cdiv_q(r, s);
                    q = cdiv_q(q, p);
                    r = cdiv_q(r, p);
                    for (int j = k; j <= nn; j++) {             // row modification
                        cpx<T> u = H_(k, j);
                        const cpx<T> b1 = H_(k + 1, j);
                        u.re = u.re + (q.re * b1.re + q.im * b1.im);        // QLA_c_peq_ca_times_c
                        u.im = u.im + (q.re * b1.im - q.im * b1.re);
                        if (k + 1 != nn) {
                            const cpx<T> b2 = H_(k + 2, j);
                            u.re = u.re + (r.re * b2.re + r.im * b2.im);
                            u.im = u.im + (r.re * b2.im - r.im * b2.re);
                            acc_mul<true, false>(H_(k + 2, j), u, w);
                        }
                        acc_mul<true, false>(H_(k + 1, j), u, y);
                        acc_mul<true, false>(H_(k, j), u, x);
                    }
                    const int ie = nn < k + 3 ? nn : k + 3;
                    for (int i = l; i <= ie; i++) {             // column modification
                        cpx<T> u = cmul(x, H_(i, k));
                        acc_mul<false, false>(u, y, H_(i, k + 1));
                        if (k + 1 != nn) {
                            acc_mul<false, false>(u, w, H_(i, k + 2));
                            H_(i, k + 2).re = H_(i, k + 2).re - (u.re * r.re + u.im * r.im);   // QLA_c_meq_c_times_ca
                            H_(i, k + 2).im = H_(i, k + 2).im - (u.im * r.re - u.re * r.im);
                        }
                        H_(i, k + 1).re = H_(i, k + 1).re - (u.re * q.re + u.im * q.im);
                        H_(i, k + 1).im = H_(i, k + 1).im - (u.im * q.re - u.re * q.im);
                        H_(i, k).re = H_(i, k).re - u.re;
                        H_(i, k).im = H_(i, k).im - u.im;
                    }
                }
            }
        } while (l + 1 < nn);
    }
#undef H_
}

// ---- M_eq_log_M (code-templates/M_eq_log_M.c:49-70, 158-200), nc >= 3 ----
// Scale by the Frobenius norm, take NROOTS = 4 square roots (the M_eq_sqrt_M
// iteration), then log(1 + X) ~ X - X^2/2 + X^3 P(X)/Q(X) with the template's
// Pade coefficients (double, rounded to T where used), times 2^NROOTS, plus
// log of the scale on the diagonal (double, added with C's conversions).
constexpr double LOG_P[6] = { 7.70838733755885391666E0, 1.79368678507819816313E1, 1.44989225341610930846E1,
                                 4.70579119878881725854E0, 4.97494994976747001425E-1, 1.01875663804580931796E-4 };
constexpr double LOG_Q[5] = { 2.31251620126765340583E1, 7.11544750618563894466E1, 8.29875266912776603211E1,
                                 4.52279145837532221105E1, 1.12873587189167450590E1 };
constexpr int LOG_NROOTS = 4;

template <class T, int NC> __device__ void f_log(mat<T, NC> &r, const mat<T, NC> &a)
{
    double ds = sqrt(maxev2(a));
    if (ds == 0) { m_eq_d(r, (T)(-1. / 0.)); return; }       // M_eq_d(r, log(0))
    mat<T, NC> as, as2, as3, as4, pa, qa;
    m_eq_r_times_m(as, (T)(1 / ds), a);
    for (int i = 0; i < LOG_NROOTS; i++) {
        as2 = as;
        f_sqrt(as, as2);
    }
    LOOP(i, NC) as.e[i * NC + i].re = as.e[i * NC + i].re - 1;
    m_eq_d(pa, (T)LOG_P[0]);
    m_eq_d(qa, (T)LOG_Q[0]);
    m_peq_r_times_m(pa, (T)LOG_P[1], as);
    m_peq_r_times_m(qa, (T)LOG_Q[1], as);
    m_eq_m_times_m(as2, as, as);
    m_peq_r_times_m(pa, (T)LOG_P[2], as2);
    m_peq_r_times_m(qa, (T)LOG_Q[2], as2);
    m_eq_m_times_m(as3, as, as2);
    m_peq_r_times_m(pa, (T)LOG_P[3], as3);
    m_peq_r_times_m(qa, (T)LOG_Q[3], as3);
    m_eq_m_times_m(as4, as2, as2);
    m_peq_r_times_m(pa, (T)LOG_P[4], as4);
    m_peq_r_times_m(qa, (T)LOG_Q[4], as4);
    {
        mat<T, NC> as5;
        m_eq_m_times_m(as5, as, as4);
        m_peq_r_times_m(pa, (T)LOG_P[5], as5);
        m_peq_m(qa, as5);
    }
    m_eq_inverse_m(as4, qa);          // as5 of the template; as4 is dead here
    m_eq_m_times_m(qa, pa, as4);
    m_eq_m_times_m(pa, as3, qa);
    m_peq_r_times_m(pa, (T)-0.5, as2);
    m_peq_m(pa, as);
    m_eq_r_times_m(r, (T)(1 << LOG_NROOTS), pa);
    ds = log(ds);
    LOOP(i, NC) r.e[i * NC + i].re = (T)((double)r.e[i * NC + i].re + ds);
}

} // namespace qla
