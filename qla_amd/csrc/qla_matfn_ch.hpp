// qla_matfn_ch.hpp -- Cayley-Hamilton forms of exp, sqrt and 1/sqrt for 3x3
// single-precision matrices with the coefficient step in double: an addition
// of this library with no counterpart in QLA (whose expA / sqrtPH templates
// evaluate their eigenvalues in the operand precision).
//
// f(X) = f0 + f1 X + f2 X^2.  Everything after the load is evaluated in
// double (every float is exact in double): X's invariants, its eigenvalues,
// the coefficients, X^2 and the sum, which is rounded to float once per
// component.  The sum cancels: its terms exceed the result by the ratio
// (|f0| + |f1| |X| + |f2| |X|^2) / |f(X)|, about 100 for a square root with two
// eigenvalues 1e-4 of the largest and unbounded as two of them approach zero,
// so a float sum would lose that factor times 2^-24.  The kernels are
// HBM-bound and gfx950 issues double adds and multiplies at the float rate,
// so the double sum costs no measurable time.
//   * exp of an anti-Hermitian A (its anti-Hermitian part is used): with
//     h = Im tr A / 3 and the traceless Hermitian Q = -i (A - i h 1),
//     exp(A) = e^{ih} exp(iQ) and exp(iQ) = f0 + f1 Q + f2 Q^2 with the
//     coefficients of Morningstar and Peardon, Phys. Rev. D 69, 054501
//     (2004), eqs. 23-34 (c0 = det Q, c1 = tr Q^2 / 2, u, w, and the symmetry
//     f_j(-c0) = (-1)^j conj f_j(c0) for c0 < 0).
//   * sqrt and 1/sqrt of a positive (semi-)definite Hermitian A (its
//     Hermitian part is used): eigenvalues l_k by the trigonometric solution
//     of the characteristic cubic (O. K. Smith, CACM 4 (1961) 168), s_k =
//     sqrt(max(l_k, 0)), u = sum s_k, v = sum_{j<k} s_j s_k, w = prod s_k,
//     d = (s0 + s1)(s0 + s2)(s1 + s2) = u v - w.  S = sqrt(A) satisfies
//     S^3 - u S^2 + v S - w = 0 and S^2 = A, so S (A + v) = u A + w, which
//     with Cayley-Hamilton for A gives sqrt(A) = (u w + (tr A + v) A - A^2)/d;
//     A^{-1/2} = (A + v - u S) / w gives the coefficients of the inverse.
//     A rank-one A (d = 0, u > 0) has sqrt(A) = A / u; a zero A gives zero,
//     and an A with a zero eigenvalue gives +inf on the diagonal for 1/sqrt
//     (0 elsewhere), as the reference's invsqrt does for a zero matrix.
#pragma once
#include "qla_matfn.hpp"

namespace qla {

// Double complex helpers.  This code has no reference operation order to
// follow, so it uses fused multiply-adds (explicit fma; the library is built
// without contraction).
struct zd { double re, im; };
__device__ __forceinline__ zd zmul(zd a, zd b)
{
    return { fma(a.re, b.re, -a.im * b.im), fma(a.re, b.im, a.im * b.re) };
}
__device__ __forceinline__ zd zmadd(zd a, zd b, zd c)     // a b + c
{
    return { fma(a.re, b.re, fma(-a.im, b.im, c.re)), fma(a.re, b.im, fma(a.im, b.re, c.im)) };
}
__device__ __forceinline__ zd zadd(zd a, zd b) { return { a.re + b.re, a.im + b.im }; }
__device__ __forceinline__ zd zscale(double s, zd a) { return { s * a.re, s * a.im }; }

// sin(w) / w from sin(w); Taylor series (error below 1e-16) for |w| < 0.05
__device__ __forceinline__ double ch_sinc(double w, double sw)
{
    const double w2 = w * w;
    if (fabs(w) < 0.05) return 1.0 - w2 / 6.0 * (1.0 - w2 / 20.0 * (1.0 - w2 / 42.0));
    return sw / w;
}

// determinant of a Hermitian 3x3 given by its real diagonal d0..d2 and upper
// off-diagonal a = x01, b = x02, c = x12
__device__ __forceinline__ double herm_det3(double d0, double d1, double d2, zd a, zd b, zd c)
{
    const double acb = (a.re * c.re - a.im * c.im) * b.re + (a.re * c.im + a.im * c.re) * b.im;    // Re(a c conj b)
    const double na = a.re * a.re + a.im * a.im, nb = b.re * b.re + b.im * b.im, nc = c.re * c.re + c.im * c.im;
    return d0 * d1 * d2 + 2.0 * acb - d0 * nc - d1 * nb - d2 * na;
}

// cos(theta / 3) for t = cos(theta) in [0, 1]: the root in [cos(pi/6), 1] of
// 4 x^3 - 3 x = t, where the derivative 12 x^2 - 3 >= 6, from a
// single-precision seed and two Newton steps in double (the seed's error
// 1e-7 falls below double rounding in two quadratic steps)
__device__ __forceinline__ double cos_third(double t)
{
    double x = (double)cosf(acosf((float)t) * (1.0f / 3.0f));
    LOOP(k, 2) {
        const double x2 = x * x;
        x = x - fma(fma(4.0, x2, -3.0), x, -t) / fma(12.0, x2, -3.0);
    }
    return x;
}

// f0, f1, f2 of exp(iQ) = f0 + f1 Q + f2 Q^2, Q traceless Hermitian with
// c0 = det Q and c1 = tr Q^2 / 2 >= 0
__device__ __forceinline__ void ch_exp_coeffs(zd f[3], double c0, double c1)
{
    if (c1 == 0.0) {
        f[0] = { 1.0, 0.0 };
        f[1] = f[2] = { 0.0, 0.0 };
        return;
    }
    const bool neg = c0 < 0.0;
    const double c13 = c1 / 3.0;
    const double sc13 = sqrt(c13);
    const double c0max = 2.0 * c13 * sc13;
    const double t = fmin(fabs(c0) / c0max, 1.0);
    // cos and sin of theta / 3; 1 - cth is exact (Sterbenz), and an error of
    // sth near theta = 0, where the two eigenvalues -u +- w nearly meet,
    // moves them symmetrically about their mean and changes exp only at
    // second order
    const double cth = cos_third(t);
    const double sth = sqrt(fmax(0.0, (1.0 - cth) * (1.0 + cth)));
    const double u = sc13 * cth, w = sqrt(c1) * sth;
    const double u2 = u * u, w2 = w * w;
    double sw, cw, su, cu;
    sincos(w, &sw, &cw);
    sincos(u, &su, &cu);
    const double xi0 = ch_sinc(w, sw);
    const zd e2iu = { fma(cu, cu, -su * su), 2.0 * su * cu }, emiu = { cu, -su };
    const zd h0 = zadd(zscale(u2 - w2, e2iu), zmul(emiu, { 8.0 * u2 * cw, 2.0 * u * (3.0 * u2 + w2) * xi0 }));
    const zd h1 = zadd(zscale(2.0 * u, e2iu), zscale(-1.0, zmul(emiu, { 2.0 * u * cw, -(3.0 * u2 - w2) * xi0 })));
    const zd h2 = zadd(e2iu, zscale(-1.0, zmul(emiu, { cw, 3.0 * u * xi0 })));
    const double di = 1.0 / (9.0 * u2 - w2);
    f[0] = zscale(di, h0);
    f[1] = zscale(di, h1);
    f[2] = zscale(di, h2);
    if (neg) {
        f[0].im = -f[0].im;
        f[1].re = -f[1].re;
        f[2].im = -f[2].im;
    }
}

// eigenvalues of the Hermitian matrix with real diagonal d0..d2 and upper
// off-diagonal a = x01, b = x02, c = x12 (trigonometric solution)
__device__ __forceinline__ void herm_eig3(double l[3], double d0, double d1, double d2, zd a, zd b, zd c)
{
    const double m = (d0 + d1 + d2) / 3.0;
    const double k0 = d0 - m, k1 = d1 - m, k2 = d2 - m;
    const double p2 = (k0 * k0 + k1 * k1 + k2 * k2 + 2.0 * (a.re * a.re + a.im * a.im + b.re * b.re + b.im * b.im +
                                                            c.re * c.re + c.im * c.im)) / 6.0;
    if (p2 == 0.0) {
        l[0] = l[1] = l[2] = m;
        return;
    }
    const double p = sqrt(p2);
    const double q = 0.5 * herm_det3(k0, k1, k2, a, b, c);
    const double r = fmax(-1.0, fmin(1.0, q / (p * p * p)));
    double sp, cp;
    sincos(acos(r) / 3.0, &sp, &cp);
    l[0] = m + 2.0 * p * cp;
    // cos(phi + 2 pi / 3) = -cos(phi) / 2 - sin(phi) sqrt(3) / 2
    l[2] = m - p * (cp + 1.7320508075688772935 * sp);
    l[1] = 3.0 * m - l[0] - l[2];
}

// real coefficients of A^{1/2} (INV: A^{-1/2}) = f0 + f1 A + f2 A^2
template <bool INV> __device__ __forceinline__ void ch_sqrt_coeffs(double f[3], const double lam[3])
{
    double l[3], s[3];
    LOOP(k, 3) {
        l[k] = fmax(lam[k], 0.0);
        s[k] = sqrt(l[k]);
    }
    const double u = s[0] + s[1] + s[2];
    const double v = s[0] * s[1] + s[0] * s[2] + s[1] * s[2];
    const double w = s[0] * s[1] * s[2];
    const double d = (s[0] + s[1]) * (s[0] + s[2]) * (s[1] + s[2]);
    const double tr = l[0] + l[1] + l[2];
    if constexpr (!INV) {
        if (d == 0.0) {             // rank <= 1: sqrt(A) = A / u, or 0
            f[0] = f[2] = 0.0;
            f[1] = u > 0.0 ? 1.0 / u : 0.0;
            return;
        }
        const double di = 1.0 / d;
        f[0] = u * w * di;
        f[1] = (tr + v) * di;
        f[2] = -di;
    } else {
        if (w == 0.0) {             // singular: +inf on the diagonal
            f[0] = __builtin_huge_val();
            f[1] = f[2] = 0.0;
            return;
        }
        const double di = 1.0 / (w * d);
        // v d - u^2 w as a sum of non-negative terms
        const double n0 = w * u * u + l[0] * s[0] * (tr - l[0]) + l[1] * s[1] * (tr - l[1]) + l[2] * s[2] * (tr - l[2]);
        f[0] = n0 * di;
        f[1] = -(u * tr + w) * di;
        f[2] = u * di;
    }
}

// x^2 of a Hermitian x (real diagonal): the diagonal sum_k |x_ik|^2 and the
// upper triangle; the lower triangle is its conjugate
__device__ __forceinline__ void herm_square(zd x2[9], const zd x[9])
{
    LOOP(i, 3) {
        double d = 0.0;
        LOOP(k, 3) d = fma(x[i * 3 + k].re, x[i * 3 + k].re, fma(x[i * 3 + k].im, x[i * 3 + k].im, d));
        x2[i * 4] = { d, 0.0 };
    }
    LOOP(i, 3) LOOP(j, 3) {
        if (j > i) {
            zd t = zmul(x[i * 3], x[j]);
            LOOP(k, 2) t = zmadd(x[i * 3 + k + 1], x[(k + 1) * 3 + j], t);
            x2[i * 3 + j] = t;
            x2[j * 3 + i] = { t.re, -t.im };
        }
    }
}

// e = c0 1 + c1 x + c2 x^2 for a Hermitian x, evaluated in double and rounded
// to float
__device__ __forceinline__ void ch_combine(mat<float, 3> &e, const zd x[9], zd c0, zd c1, zd c2)
{
    zd x2[9];
    herm_square(x2, x);
    LOOP(k, 9) {
        const zd r = zmadd(c2, x2[k], zmadd(c1, x[k], k % 4 == 0 ? c0 : zd{ 0.0, 0.0 }));
        e.e[k] = { (float)r.re, (float)r.im };
    }
}

// the same with real coefficients: the result is Hermitian, its diagonal real
__device__ __forceinline__ void ch_combine_real(mat<float, 3> &e, const zd x[9], double c0, double c1, double c2)
{
    zd x2[9];
    herm_square(x2, x);
    LOOP(i, 3) LOOP(j, 3) {
        if (j == i) {
            e.e[i * 4] = { (float)fma(c2, x2[i * 4].re, fma(c1, x[i * 4].re, c0)), 0.0f };
        } else if (j > i) {
            const int k = i * 3 + j;
            const zd r = { fma(c2, x2[k].re, c1 * x[k].re), fma(c2, x2[k].im, c1 * x[k].im) };
            e.e[k] = { (float)r.re, (float)r.im };
            e.e[j * 3 + i] = { (float)r.re, (float)-r.im };
        }
    }
}

// exp of the anti-Hermitian part of a
__device__ void f_expA_ch(mat<float, 3> &e, const mat<float, 3> &a)
{
    // B = (A - A^dagger) / 2, h = Im tr B / 3, Q = -i B - h 1 (Hermitian, traceless)
    zd b[9];
    LOOP(i, 3) LOOP(j, 3) {
        const cpx<float> x = a.e[i * 3 + j], y = a.e[j * 3 + i];
        b[i * 3 + j] = { 0.5 * ((double)x.re - (double)y.re), 0.5 * ((double)x.im + (double)y.im) };
    }
    const double h = (b[0].im + b[4].im + b[8].im) / 3.0;
    zd q[9];
    LOOP(k, 9) q[k] = { b[k].im, -b[k].re };        // -i (x + iy) = y - ix
    LOOP(i, 3) q[i * 4] = { b[i * 4].im - h, 0.0 };
    double c1 = 0.0;
    LOOP(k, 9) c1 += q[k].re * q[k].re + q[k].im * q[k].im;
    c1 *= 0.5;
    const double c0 = herm_det3(q[0].re, q[4].re, q[8].re, q[1], q[2], q[5]);
    zd f[3];
    ch_exp_coeffs(f, c0, c1);
    zd ph = { 1.0, 0.0 };       // e^{ih}; no sincos for the traceless inputs
    if (h != 0.0) sincos(h, &ph.im, &ph.re);
    ch_combine(e, q, zmul(ph, f[0]), zmul(ph, f[1]), zmul(ph, f[2]));
}

// A^{1/2} (INV: A^{-1/2}) of the Hermitian part of a
template <bool INV> __device__ void f_sqrtPH_ch(mat<float, 3> &e, const mat<float, 3> &a)
{
    zd x[9];
    LOOP(i, 3) LOOP(j, 3) {
        const cpx<float> p = a.e[i * 3 + j], q = a.e[j * 3 + i];
        x[i * 3 + j] = { 0.5 * ((double)p.re + (double)q.re), 0.5 * ((double)p.im - (double)q.im) };
    }
    double lam[3], f[3];
    herm_eig3(lam, x[0].re, x[4].re, x[8].re, x[1], x[2], x[5]);
    ch_sqrt_coeffs<INV>(f, lam);
    ch_combine_real(e, x, f[0], f[1], f[2]);
}

} // namespace qla
