// qla_hostmath.hpp -- the single-precision acosf, cbrtf and sincosf of the
// host C library (GNU libc 2.35 on x86-64), restated step for step so that a
// device evaluation returns the same bits as the call the reference makes.
//
// The F3 closed forms of the reference (expA, expTA, sqrtPH, invsqrtPH,
// eigenvalsH; code-templates/M_eq_expA_M.c:16-19 and the others) call
// acosf, cbrtf and, through QLA_F_cexpi (lib/cmath/QLA_cexpi.c:13-20),
// sincosf.  These library functions are not correctly rounded (cbrtf
// differs from the correctly rounded cube root on about 10% of inputs,
// acosf on 0.25%, sinf/cosf on 0.03-0.09%), and those differences grow
// through the eigenvalue formulas, so no other implementation reproduces
// the reference's F3 results.  The algorithms restated here:
//   * acosf: the fdlibm rational approximation of asin/acos on three
//     intervals, float arithmetic;
//   * cbrtf: a quadratic first guess on the mantissa in [0.5, 1), one
//     Halley step in double, the exponent split into a multiple of 3 and a
//     table factor 2^(k/3);
//   * sincosf: argument in double, reduction by the nearest multiple of
//     pi/2 for |x| < 120, degree-7 (sin) and degree-8 (cos) polynomials in
//     double evaluated with fused multiply-adds where the x86-64 build of
//     the library contracts them (the build it selects on processors with
//     AVX2 and FMA).  |x| >= 120 is not restated: hm_sincosf returns false
//     and the caller takes the device library's sincosf.
// Every function was compared with the host library on every float input
// of its domain (acosf on [-1, 1], cbrtf on all floats, sincosf on
// |x| < 120): no difference.  tests/test_hostmath.py repeats the comparison
// on a strided sample of inputs.
//
// Compiled for the device by hipcc and, for that test, for the host by a
// C++ compiler with -ffp-contract=off.
#pragma once
#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define QLA_HM __host__ __device__ __forceinline__
#else
#define QLA_HM static inline
#endif

namespace qla {

QLA_HM uint32_t hm_bits(float x) { return __builtin_bit_cast(uint32_t, x); }
QLA_HM float hm_float(uint32_t u) { return __builtin_bit_cast(float, u); }

// acos(x) for float x
QLA_HM float hm_acosf(float x)
{
    const float one = 1.0f, pi = 3.1415925026e+00f, pio2_hi = 1.5707962513e+00f, pio2_lo = 7.5497894159e-08f,
                pS0 = 1.6666667163e-01f, pS1 = -3.2556581497e-01f, pS2 = 2.0121252537e-01f,
                pS3 = -4.0055535734e-02f, pS4 = 7.9153501429e-04f, pS5 = 3.4793309169e-05f,
                qS1 = -2.4033949375e+00f, qS2 = 2.0209457874e+00f, qS3 = -6.8828397989e-01f,
                qS4 = 7.7038154006e-02f;
    const int32_t hx = (int32_t)hm_bits(x), ix = hx & 0x7fffffff;
    if (ix == 0x3f800000) return hx > 0 ? 0.0f : pi + 2.0f * pio2_lo;     // |x| = 1
    if (ix > 0x3f800000) return (x - x) / (x - x);                      // |x| > 1, NaN
    float z, p, q, r, w, s;
    if (ix < 0x3f000000) {                                              // |x| < 0.5
        if (ix <= 0x23000000) return pio2_hi + pio2_lo;                 // |x| < 2^-57
        z = x * x;
        p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
        q = one + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
        r = p / q;
        return pio2_hi - (x - (pio2_lo - x * r));
    }
    if (hx < 0) {                                                       // x < -0.5
        z = (one + x) * 0.5f;
        p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
        q = one + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
        s = sqrtf(z);
        r = p / q;
        w = r * s - pio2_lo;
        return pi - 2.0f * (s + w);
    }
    z = (one - x) * 0.5f;                                               // x > 0.5
    s = sqrtf(z);
    const float df = hm_float(hm_bits(s) & 0xfffff000u);
    const float c = (z - df * df) / (s + df);
    p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    q = one + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    r = p / q;
    w = r * s + c;
    return 2.0f * (df + w);
}

// cube root of float x
QLA_HM float hm_cbrtf(float x)
{
    if (!isfinite(x) || x == 0.0f) return x + x;
    const double cbrt2 = 1.2599210498948731648, sqr_cbrt2 = 1.5874010519681994748;
    int xe;
    const float xm = frexpf(fabsf(x), &xe);
    const int k = xe % 3;       // -2 .. 2
    const double factor = k == -2 ? 1.0 / sqr_cbrt2 : k == -1 ? 1.0 / cbrt2 : k == 0 ? 1.0 : k == 1 ? cbrt2 : sqr_cbrt2;
    const float u = (float)(0.492659620528969547 + (0.697570460207922770 - 0.191502161678719066 * xm) * xm);
    const float t2 = u * u * u;
    const float ym = (float)(u * (t2 + 2.0 * xm) / (2.0 * t2 + xm) * factor);
    return ldexpf(x > 0.0f ? ym : -ym, xe / 3);
}

// sin and cos of float y for |y| < 120 (returns false, writing nothing,
// otherwise)
struct hm_sincos_tab {
    double c0, c1, c2, c3, c4, s1, s2, s3;
};
QLA_HM void hm_sincos_poly(double x, double x2, const hm_sincos_tab &p, int n, float *sinp, float *cosp)
{
    const double x4 = x2 * x2, x3 = x2 * x;
    const double c2 = fma(x2, p.c4, p.c3);
    const double s1 = fma(x2, p.s3, p.s2);
    float *tmp = (n & 1) ? cosp : sinp;     // odd quadrants swap the two results
    cosp = (n & 1) ? sinp : cosp;
    sinp = tmp;
    const double c1 = fma(x2, p.c1, p.c0);
    const double x5 = x3 * x2, x6 = x4 * x2;
    const double s = fma(x3, p.s1, x);
    const double c = fma(x4, p.c2, c1);
    *sinp = (float)fma(x5, s1, s);
    *cosp = (float)fma(x6, c2, c);
}
QLA_HM bool hm_sincosf(float y, float *sinp, float *cosp)
{
    // the second table yields -cos, so quadrants 2 and 3 negate for free
    const hm_sincos_tab t0 = { 0x1p0, -0x1.ffffffd0c621cp-2, 0x1.55553e1068f19p-5, -0x1.6c087e89a359dp-10,
                               0x1.99343027bf8c3p-16, -0x1.555545995a603p-3, 0x1.1107605230bc4p-7,
                               -0x1.994eb3774cf24p-13 };
    const hm_sincos_tab t1 = { -0x1p0, 0x1.ffffffd0c621cp-2, -0x1.55553e1068f19p-5, 0x1.6c087e89a359dp-10,
                               -0x1.99343027bf8c3p-16, -0x1.555545995a603p-3, 0x1.1107605230bc4p-7,
                               -0x1.994eb3774cf24p-13 };
    const double hpi_inv = 0x1.45F306DC9C883p+23, hpi = 0x1.921FB54442D18p0;   // 2^24 * 2/pi, pi/2
    const uint32_t top = (hm_bits(y) >> 20) & 0x7ff;
    double x = y;
    if (top < ((hm_bits(0x1.921fb6p-1f) >> 20) & 0x7ff)) {           // |y| < pi/4
        if (top < ((hm_bits(0x1p-12f) >> 20) & 0x7ff)) {
            *sinp = y;
            *cosp = 1.0f;
            return true;
        }
        hm_sincos_poly(x, x * x, t0, 0, sinp, cosp);
        return true;
    }
    if (top >= ((hm_bits(120.0f) >> 20) & 0x7ff)) return false;
    const double r = x * hpi_inv;
    const int n = ((int32_t)r + 0x800000) >> 24;                      // nearest multiple of pi/2
    x = fma(-(double)n, hpi, x);
    const double s = (n & 3) == 1 || (n & 3) == 2 ? -1.0 : 1.0;
    hm_sincos_poly(x * s, x * x, (n & 2) ? t1 : t0, n, sinp, cosp);
    return true;
}

} // namespace qla
