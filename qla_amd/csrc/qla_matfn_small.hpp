// qla_matfn_small.hpp -- the colour-matrix functions at nc = 1 and 2 (QLA_FN
// with one or two colours), where every template branches to its own closed
// form before the general algorithm: C_eq_det_M.c:59-69, M_eq_inverse_M.c:
// 60-78, V_eq_M_inverse_V.c:61-83, M_eq_M_inverse_M.c:61-92, M_eq_exp_M.c:
// 252-292, M_eq_sqrt_M.c:82-135, M_eq_invsqrt_M.c:82-132, M_eq_expA_M.c:
// 307-348, M_eq_expTA_M.c:296-326, M_eq_sqrtPH_M.c:218-261,
// M_eq_invsqrtPH_M.c:222-250, M_eq_log_M.c:98-150, V_eq_eigenvalsH_M.c:
// 80-101, 165-175.  The complex functions they call are lib/cmath's:
// csqrt (QLA_csqrt.c: from |z| and sqrt of the half sums, IEEE operations
// only), cexpi (the host library's sincosf, restated in qla_hostmath.hpp),
// cexp (QLA_cexp.c: expf and sincosf), csinhcosh (QLA_csinhcosh.c: sinhf,
// sqrtf(1 + sh^2), sincosf) and clog (QLA_clog.c: half the double log of
// |z|^2, atan2f).  Everything but expf, sinhf, the double log and atan2f is
// therefore the reference's arithmetic bit for bit; exp and log at nc = 1, 2
// pass through those four device-library functions.
#pragma once
#include "qla_matfn.hpp"

namespace qla {

template <class T> __device__ __forceinline__ cpx<T> csqrt_q(cpx<T> z)
{
    const T n = m_sqrt(z.re * z.re + z.im * z.im);                 // QLA_norm_c
    const T sr = m_sqrt(m_fabs((T)(0.5 * (n + z.re))));
    const T si = copysign(m_sqrt(m_fabs((T)(0.5 * (n - z.re)))), z.im);
    return { sr, si };
}
template <class T> __device__ __forceinline__ cpx<T> cexp_q(cpx<T> z)
{
    const T mag = exp(z.re);
    const cpx<T> e = cexpi<T>(z.im);
    return { mag * e.re, mag * e.im };
}
template <class T> __device__ __forceinline__ void csinhcosh_q(cpx<T> z, cpx<T> &sh, cpx<T> &ch)
{
    const T sx = sinh(z.re);
    const T cx = m_sqrt(1 + sx * sx);
    const cpx<T> e = cexpi<T>(z.im);                                 // (cos y, sin y)
    sh = { e.re * sx, e.im * cx };
    ch = { e.re * cx, e.im * sx };
}
template <class T> __device__ __forceinline__ cpx<T> clog_q(cpx<T> z)
{
    return { (T)(0.5 * log((double)(z.re * z.re + z.im * z.im))), atan2(z.im, z.re) };
}
// QLA_c_eq_r_div_c(c, x, b)
template <class T> __device__ __forceinline__ cpx<T> rdivc_q(T x, cpx<T> b)
{
    const T t = x / (b.re * b.re + b.im * b.im);
    return { t * b.re, -t * b.im };
}
// QLA_c_eq_c_times_c_plus_c(c, a, x, b)
template <class T> __device__ __forceinline__ cpx<T> cmuladd_q(cpx<T> a, cpx<T> x, cpx<T> b)
{
    return { a.re * x.re - a.im * x.im + b.re, a.re * x.im + a.im * x.re + b.im };
}
// QLA_c_eqm_c_times_c
template <class T> __device__ __forceinline__ cpx<T> cnegmul_q(cpx<T> a, cpx<T> b)
{
    return { -a.re * b.re + a.im * b.im, -a.re * b.im - a.im * b.re };
}
template <class T> __device__ __forceinline__ cpx<T> csub_q(cpx<T> a, cpx<T> b) { return { a.re - b.re, a.im - b.im }; }
template <class T> __device__ __forceinline__ cpx<T> cadd_q(cpx<T> a, cpx<T> b) { return { a.re + b.re, a.im + b.im }; }

// c0 + c1 a on a 2x2 (the templates' closing four lines)
template <class T>
__device__ __forceinline__ void c0_plus_c1a(cpx<T> *r, cpx<T> c0, cpx<T> c1, cpx<T> a00, cpx<T> a01, cpx<T> a10,
                                            cpx<T> a11)
{
    r[0] = cmuladd_q(c1, a00, c0);
    r[1] = cmul(c1, a01);
    r[2] = cmul(c1, a10);
    r[3] = cmuladd_q(c1, a11, c0);
}
// real c0, c1 (QLA_c_eq_c_times_r_plus_r, QLA_c_eq_c_times_r)
template <class T> __device__ __forceinline__ void rc0_plus_rc1a(cpx<T> *r, T c0, T c1, const cpx<T> *a)
{
    r[0] = { a[0].re * c1 + c0, a[0].im * c1 };
    r[1] = { a[1].re * c1, a[1].im * c1 };
    r[2] = { a[2].re * c1, a[2].im * c1 };
    r[3] = { a[3].re * c1 + c0, a[3].im * c1 };
}

// out: the result record (1 complex for det, nc for inverse_V and the
// eigenvalues, nc*nc otherwise); b: the second operand of the solves
template <class T, int NC, int FN>
__device__ __forceinline__ void f_small(cpx<T> *out, const mat<T, NC> &m, const cpx<T> *b)
{
    // function codes as in qla_matfn.hip
    constexpr int EXP = 0, SQRT = 1, INVSQRT = 2, DET = 3, INVERSE = 4, INV_V = 5, INV_M = 6, EXPA = 7, EXPTA = 8,
                  SQRTPH = 9, INVSQRTPH = 10, LOG = 11, EIGH = 12;
    if constexpr (NC == 1) {
        const cpx<T> a = m.e[0];
        if constexpr (FN == DET || FN == EIGH) out[0] = a;
        else if constexpr (FN == INVERSE) out[0] = rdivc_q(T(1), a);
        else if constexpr (FN == INV_V || FN == INV_M) out[0] = cdiv_q(b[0], a);
        else if constexpr (FN == EXP) out[0] = cexp_q(a);
        else if constexpr (FN == SQRT) out[0] = csqrt_q(a);
        else if constexpr (FN == INVSQRT) out[0] = rdivc_q(T(1), csqrt_q(a));
        else if constexpr (FN == EXPA) out[0] = cexpi<T>(a.im);
        else if constexpr (FN == EXPTA) out[0] = { T(1), T(0) };
        else if constexpr (FN == SQRTPH) out[0] = { m_sqrt(m_fabs(a.re)), T(0) };
        else if constexpr (FN == INVSQRTPH) out[0] = { T(1) / m_sqrt(m_fabs(a.re)), T(0) };
        else if constexpr (FN == LOG) out[0] = clog_q(a);
    } else {
        static_assert(NC == 2, "nc = 1 or 2");
        cpx<T> a00 = m.e[0], a01 = m.e[1], a10 = m.e[2], a11 = m.e[3];
        if constexpr (FN == DET) {
            cpx<T> det = cmul(a00, a11);
            acc_mul<true, false>(det, a01, a10);
            out[0] = det;
        } else if constexpr (FN == INVERSE || FN == INV_V || FN == INV_M) {
            cpx<T> det = cmul(a00, a11);
            acc_mul<true, false>(det, a01, a10);
            const cpx<T> idet = rdivc_q(T(1), det);
            if constexpr (FN == INVERSE) {
                out[0] = cmul(a11, idet);
                out[1] = cnegmul_q(a01, idet);
                out[2] = cnegmul_q(a10, idet);
                out[3] = cmul(a00, idet);
            } else if constexpr (FN == INV_V) {
                cpx<T> r0 = cmul(a11, b[0]);
                acc_mul<true, false>(r0, a01, b[1]);
                cpx<T> r1 = cmul(a00, b[1]);
                acc_mul<true, false>(r1, a10, b[0]);
                out[0] = cmul(r0, idet);
                out[1] = cmul(r1, idet);
            } else {
                cpx<T> r00 = cmul(a11, b[0]); acc_mul<true, false>(r00, a01, b[2]);
                cpx<T> r01 = cmul(a11, b[1]); acc_mul<true, false>(r01, a01, b[3]);
                cpx<T> r10 = cmul(a00, b[2]); acc_mul<true, false>(r10, a10, b[0]);
                cpx<T> r11 = cmul(a00, b[3]); acc_mul<true, false>(r11, a10, b[1]);
                out[0] = cmul(r00, idet);
                out[1] = cmul(r01, idet);
                out[2] = cmul(r10, idet);
                out[3] = cmul(r11, idet);
            }
        } else if constexpr (FN == EXP) {
            const cpx<T> tr = cadd_q(a00, a11);
            const cpx<T> s = { (T)(0.5 * tr.re), (T)(0.5 * tr.im) };
            a00 = csub_q(a00, s);
            a11 = csub_q(a11, s);
            cpx<T> mdet = cmul(a01, a10);
            acc_mul<true, false>(mdet, a00, a11);
            const cpx<T> t = csqrt_q(mdet);
            cpx<T> st, ch;
            if (t.re == 0 && t.im == 0) {
                st = { T(1), T(0) };
                ch = { T(1), T(0) };
            } else {
                cpx<T> sh;
                csinhcosh_q(t, sh, ch);
                st = cdiv_q(sh, t);
            }
            const cpx<T> es = cexp_q(s);
            c0_plus_c1a(out, cmul(es, ch), cmul(es, st), a00, a01, a10, a11);
        } else if constexpr (FN == SQRT || FN == INVSQRT) {
            cpx<T> det = cmul(a00, a11);
            acc_mul<true, false>(det, a01, a10);
            const cpx<T> tr = cadd_q(a00, a11);
            cpx<T> l0, l1, c0, c1;
            if (det.re == 0 && det.im == 0) {
                l0 = { T(0), T(0) };
                l1 = tr;
            } else {
                const cpx<T> s = { (T)(0.5 * tr.re), (T)(0.5 * tr.im) };
                const cpx<T> d = csub_q(cmul(s, s), det);
                const cpx<T> sd = csqrt_q(d);
                const T ts = s.re * sd.re + s.im * sd.im;                  // QLA_r_eq_Re_ca_times_c
                l1 = ts >= 0 ? cadd_q(s, sd) : csub_q(s, sd);
                l0 = cdiv_q(det, l1);
            }
            const cpx<T> sl0 = csqrt_q(l0), sl1 = csqrt_q(l1);
            const cpx<T> ss = cadd_q(sl0, sl1);
            cpx<T> ps = cmul(sl0, sl1);
            if constexpr (FN == SQRT) {
                if (ss.re == 0 && ss.im == 0) {
                    c0 = sl0;
                    c1 = { T(0), T(0) };
                } else {
                    c1 = rdivc_q(T(1), ss);
                    c0 = cmul(ps, c1);
                }
            } else {
                c1 = rdivc_q(T(-1), cmul(ps, ss));
                ps = cadd_q(ps, tr);
                c0 = cnegmul_q(ps, c1);
            }
            c0_plus_c1a(out, c0, c1, a00, a01, a10, a11);
        } else if constexpr (FN == EXPA) {
            const cpx<T> tr = cadd_q(a00, a11);
            const cpx<T> s = { (T)(0.5 * tr.re), (T)(0.5 * tr.im) };
            a00 = csub_q(a00, s);
            a11 = csub_q(a11, s);
            T det = a00.re * a11.re - a00.im * a11.im;
            det = det - (a01.re * a10.re - a01.im * a10.im);
            const T t = m_sqrt(m_fabs(det));
            T st, cs;
            if (t == 0) {
                st = 1;
                cs = 1;
            } else {
                const cpx<T> sc = cexpi<T>(t);
                st = sc.im / t;
                cs = sc.re;
            }
            const cpx<T> es = cexpi<T>(s.im);
            c0_plus_c1a(out, cpx<T>{ es.re * cs, es.im * cs }, cpx<T>{ es.re * st, es.im * st }, a00, a01, a10, a11);
        } else if constexpr (FN == EXPTA) {
            T det = a00.re * a11.re - a00.im * a11.im;
            det = det - (a01.re * a10.re - a01.im * a10.im);
            const T t = m_sqrt(m_fabs(det));
            T c0, c1;
            if (t == 0) {
                c0 = 1;
                c1 = 1;
            } else {
                const cpx<T> sc = cexpi<T>(t);
                c0 = sc.re;
                c1 = sc.im / t;
            }
            rc0_plus_rc1a(out, c0, c1, m.e);
        } else if constexpr (FN == SQRTPH || FN == INVSQRTPH) {
            T det = a00.re * a11.re - a00.im * a11.im;
            det = det - (a01.re * a10.re - a01.im * a10.im);
            const T tr = a00.re + a11.re;
            if constexpr (FN == SQRTPH) {
                if (det == 0) {
                    if (tr == 0) {
                        LOOP(k, 4) out[k] = { T(0), T(0) };
                    } else {
                        const T c1 = T(1) / m_sqrt(m_fabs(tr));                 // c1 a, no c0 term
                        LOOP(k, 4) out[k] = { m.e[k].re * c1, m.e[k].im * c1 };
                    }
                } else {
                    const T sdet = m_sqrt(m_fabs(det));
                    const T c1 = T(1) / m_sqrt(m_fabs(tr + 2 * sdet));
                    rc0_plus_rc1a(out, sdet * c1, c1, m.e);
                }
            } else {
                const T sdet = m_sqrt(m_fabs(det));
                const T c1 = T(1) / (sdet * m_sqrt(m_fabs(tr + 2 * sdet)));
                const T c0 = (tr + sdet) * c1;
                rc0_plus_rc1a(out, c0, -c1, m.e);
            }
        } else if constexpr (FN == LOG) {
            const cpx<T> tr = cadd_q(a00, a11);
            const cpx<T> s = { (T)(0.5 * tr.re), (T)(0.5 * tr.im) };
            cpx<T> det = cmul(a00, a11);
            acc_mul<true, false>(det, a01, a10);
            const cpx<T> d = csub_q(cmul(s, s), det);
            cpx<T> c0, c1;
            if (d.re == 0 && d.im == 0) {
                c0 = { T(0), T(0) };
                c1 = cdiv_q(clog_q(s), s);
            } else {
                const cpx<T> sd = csqrt_q(d);
                const T ts = s.re * sd.re + s.im * sd.im;
                cpx<T> e1, de;
                if (ts >= 0) {
                    e1 = cadd_q(s, sd);
                    de = { 2 * sd.re, 2 * sd.im };
                } else {
                    e1 = csub_q(s, sd);
                    de = { -2 * sd.re, -2 * sd.im };
                }
                const cpx<T> e0 = cdiv_q(det, e1);
                const cpx<T> le0 = clog_q(e0), le1 = clog_q(e1);
                const cpx<T> dei = rdivc_q(T(1), de);
                const cpx<T> dl = csub_q(le1, le0);
                cpx<T> dl0 = cmul(e1, le0);
                acc_mul<true, false>(dl0, e0, le1);
                c1 = cmul(dl, dei);
                c0 = cmul(dl0, dei);
            }
            c0_plus_c1a(out, c0, c1, a00, a01, a10, a11);
        } else if constexpr (FN == EIGH) {
            eig2x2(out[0], out[1], a00, a01, a10, a11);
        }
    }
}

} // namespace qla
