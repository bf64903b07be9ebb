// qla_matfn.hip -- batched matrix functions M_veq_{exp,sqrt,invsqrt}_M for
// nc x nc complex matrices (QLA_F3 / QLA_D3 with nc = 3, QLA_FN with nc = 3..8).
//
// Each thread owns one matrix and runs the reference algorithm with the
// reference's operation order (lib/perl/code-templates/M_eq_exp_M.c:300-367,
// M_eq_sqrt_M.c:136-173, M_eq_invsqrt_M.c:135-175, M_eq_M_inverse_M.c:93-135,
// M_eq_inverse_M.c nc = 3 cofactor branch), so results agree bit for bit
// with a CPU evaluation of the same order: scaling norm and scale factor in
// double, everything else in the operand precision, IEEE-rounded divide and
// square root, no contraction.  The Gauss-Jordan pivoting of the reference
// permutes rows through an index array; here rows are swapped in registers
// with compile-time indices (same values at the same logical positions).
//
// The per-matrix work is data dependent (the number of squarings in exp, the
// Newton iteration count in sqrt/invsqrt), so matrices are handed out in
// wave-sized batches from a global counter: a wavefront that finishes a
// cheap batch takes the next one instead of idling until a fixed grid slice
// is done.
#include "qla_dev.hpp"
#include "qla_launch.hpp"
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
__constant__ double PADE_P[3] = { 9.99999999999999999910E-1, 3.02994407707441961300E-2,
                                  1.26177193074810590878E-4 };
__constant__ double PADE_Q[4] = { 2.00000000000000000009E0, 2.27265548208155028766E-1,
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
    mat<T, NC> va, tmp;
    m_eq_d(va, T(1));
    for (; s > 0; s >>= 1) {
        if (s & 1) { m_eq_m_times_m(tmp, va, pb); va = tmp; }
        m_eq_m_times_m(tmp, pb, pb); pb = tmp;
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

enum { FN_EXP = 0, FN_SQRT = 1, FN_INVSQRT = 2 };

// Work distribution: batches of 64 matrices (one per lane) are claimed by
// whole wavefronts from `counter` until the n matrices are exhausted; every
// wave leaves the loop when its claim is past the end.
template <class T, int NC, int FN>
__global__ void __launch_bounds__(256)
k_matfn(cpx<T> *__restrict__ r, const cpx<T> *__restrict__ a, int n, unsigned int *__restrict__ counter)
{
    const int lane = lane_id();
    for (;;) {
        unsigned int batch = 0;
        if (lane == 0) batch = atomicAdd(counter, 1u);
        batch = __shfl(batch, 0, WAVE);
        const long long i = (long long)batch * WAVE + lane;
        if ((long long)batch * WAVE >= n) break;
        if (i < n) {
            mat<T, NC> m, out;
            const cpx<T> *src = a + (size_t)i * NC * NC;
            LOOP(k, NC * NC) m.e[k] = src[k];
            if (FN == FN_EXP) f_exp(out, m);
            else if (FN == FN_SQRT) f_sqrt(out, m);
            else f_invsqrt(out, m);
            cpx<T> *dst = r + (size_t)i * NC * NC;
            LOOP(k, NC * NC) dst[k] = out.e[k];
        }
    }
}

__global__ void k_zero_counter(unsigned int *c) { *c = 0u; }

template <class T, int NC, int FN>
static hipError_t launch_matfn_t(void *r, const void *a, int n, unsigned int *counter, hipStream_t st)
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // enough resident waves to cover the chip; the counter balances the rest
    const long long waves_needed = ((long long)n + WAVE - 1) / WAVE;
    long long blocks = (waves_needed + 3) / 4;
    const long long cap = (long long)cus * 8;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, st, counter);
    hipLaunchKernelGGL((k_matfn<T, NC, FN>), dim3((unsigned)blocks), dim3(256), 0, st, (cpx<T> *)r,
                       (const cpx<T> *)a, n, counter);
    return hipGetLastError();
}

template <class T, int FN>
static hipError_t matfn_nc(int nc, void *r, const void *a, int n, unsigned int *counter, hipStream_t st)
{
    switch (nc) {
    case 3: return launch_matfn_t<T, 3, FN>(r, a, n, counter, st);
    case 4: return launch_matfn_t<T, 4, FN>(r, a, n, counter, st);
    case 5: return launch_matfn_t<T, 5, FN>(r, a, n, counter, st);
    case 6: return launch_matfn_t<T, 6, FN>(r, a, n, counter, st);
    case 7: return launch_matfn_t<T, 7, FN>(r, a, n, counter, st);
    case 8: return launch_matfn_t<T, 8, FN>(r, a, n, counter, st);
    default: return hipErrorInvalidValue;
    }
}

} // namespace qla

using namespace qla;

hipError_t qla_launch_matfn(int prec, int fn, int nc, void *r, const void *a, int n, unsigned int *counter,
                            hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (prec == 0) {
        if (fn == FN_EXP) return matfn_nc<float, FN_EXP>(nc, r, a, n, counter, st);
        if (fn == FN_SQRT) return matfn_nc<float, FN_SQRT>(nc, r, a, n, counter, st);
        return matfn_nc<float, FN_INVSQRT>(nc, r, a, n, counter, st);
    }
    // double precision is exported for nc = 3 only (QLA_D3)
    if (nc != 3) return hipErrorInvalidValue;
    if (fn == FN_EXP) return launch_matfn_t<double, 3, FN_EXP>(r, a, n, counter, st);
    if (fn == FN_SQRT) return launch_matfn_t<double, 3, FN_SQRT>(r, a, n, counter, st);
    return launch_matfn_t<double, 3, FN_INVSQRT>(r, a, n, counter, st);
}
