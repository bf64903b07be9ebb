/*
 * qla_oracle.c -- CPU restatement of the QLA 1.9.0 hot path (the oracle).
 * TEST INFRASTRUCTURE ONLY: loaded by the tests, the smoke check and the
 * CPU-baseline leg of the benchmark; never by the product library.
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#define _GNU_SOURCE /* sincosf / sincos, as lib/cmath/QLA_cexpi.c:4 selects */
#include <math.h>
#include <float.h>
#include <string.h>
#include <stddef.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "qla_oracle.h"

/* the small per-site helpers are inlined wherever they are called, so a
 * loop that passes constant mu / sign / eqop gets them folded (the CPU
 * baseline's loops, below) */
#define ORC_HOT static inline __attribute__((always_inline))

/* single precision instantiation */
#define REAL float
#define P(x) F_##x
#define F_EPS FLT_EPSILON
#define SQRTP sqrtf
#define FABSP fabsf
#define ACOSP acosf
#define CBRTP cbrtf
#define SINCOSP sincosf
#define MATHP(fn) fn##f
#include "qla_oracle_impl.inc"
#include "qla_oracle_matfunc.inc"
#include "qla_oracle_vecops.inc"
#include "qla_oracle_sse.inc"
#undef REAL
#undef P
#undef SQRTP
#undef FABSP
#undef ACOSP
#undef CBRTP
#undef SINCOSP
#undef MATHP

/* double precision instantiation */
#define REAL double
#define P(x) D_##x
#define D_EPS DBL_EPSILON
#define SQRTP sqrt
#define FABSP fabs
#define ACOSP acos
#define CBRTP cbrt
#define SINCOSP sincos
#define MATHP(fn) fn
#include "qla_oracle_impl.inc"
#include "qla_oracle_matfunc.inc"
#include "qla_oracle_vecops.inc"
#undef REAL
#undef P

void orc_F3_site(int kind, int eqop, int adj, float *r, const int *rindex, const float *const *pa,
                 const float *const *pb, int mu, int sign, const int *index, int n)
{ F_orc_3_site(kind, eqop, adj, r, rindex, pa, pb, mu, sign, index, n); }
void orc_F3_site_vf(int kind, int eqop, int adj, float *r, const float *af, const float *const *pb, int mu, int sign,
                    const int *index, int n)
{ F_orc_3_site_vf(kind, eqop, adj, r, af, pb, mu, sign, index, n); }
void orc_D3_site_vf(int kind, int eqop, int adj, double *r, const double *af, const double *const *pb, int mu,
                    int sign, const int *index, int n)
{ D_orc_3_site_vf(kind, eqop, adj, r, af, pb, mu, sign, index, n); }
void orc_D3_site(int kind, int eqop, int adj, double *r, const int *rindex, const double *const *pa,
                 const double *const *pb, int mu, int sign, const int *index, int n)
{ D_orc_3_site(kind, eqop, adj, r, rindex, pa, pb, mu, sign, index, n); }
void orc_F3_multi(int eqop, int adj, float *r, const int *rindex, const float *const *const *pa,
                  const float *const *const *pb, int nd, const int *index, int n)
{ F_orc_3_multi(eqop, adj, r, rindex, pa, pb, nd, index, n); }
void orc_D3_multi(int eqop, int adj, double *r, const int *rindex, const double *const *const *pa,
                  const double *const *const *pb, int nd, const int *index, int n)
{ D_orc_3_multi(eqop, adj, r, rindex, pa, pb, nd, index, n); }

double orc_F_norm2(const float *a, int ncomp, int n) { return F_orc_norm2(a, ncomp, n); }
double orc_D_norm2(const double *a, int ncomp, int n) { return D_orc_norm2(a, ncomp, n); }
long double orc_F_norm2_ld(const float *a, int ncomp, int n) { return F_orc_norm2_ld(a, ncomp, n); }
long double orc_F_norm2_terms_ld(const float *a, int ncomp, int n) { return F_orc_norm2_terms_ld(a, ncomp, n); }
long double orc_D_norm2_terms_ld(const double *a, int ncomp, int n) { return D_orc_norm2_terms_ld(a, ncomp, n); }
void orc_F_dot_terms_ld(long double *re, long double *im, long double *sa, const float *a, const float *b, int ncomp,
                        int n)
{ F_orc_dot_terms_ld(re, im, sa, a, b, ncomp, n); }
void orc_D_dot_terms_ld(long double *re, long double *im, long double *sa, const double *a, const double *b,
                        int ncomp, int n)
{ D_orc_dot_terms_ld(re, im, sa, a, b, ncomp, n); }
long double orc_D_norm2_ld(const double *a, int ncomp, int n) { return D_orc_norm2_ld(a, ncomp, n); }
void orc_F_dot(double *re, double *im, const float *a, const float *b, int ncomp, int n)
{ F_orc_dot(re, im, a, b, ncomp, n); }
void orc_D_dot(double *re, double *im, const double *a, const double *b, int ncomp, int n)
{ D_orc_dot(re, im, a, b, ncomp, n); }
void orc_F_dot_ld(long double *re, long double *im, const float *a, const float *b, int ncomp, int n)
{ F_orc_dot_ld(re, im, a, b, ncomp, n); }
void orc_D_dot_ld(long double *re, long double *im, const double *a, const double *b, int ncomp, int n)
{ D_orc_dot_ld(re, im, a, b, ncomp, n); }

/* precision conversion (build_qla.pl:530-585: the "identity" unary op with
 * a cast between the precisions): C's float -> double and double -> float */
void orc_convert_DF(double *r, const float *a, size_t m)
{
    for (size_t i = 0; i < m; i++) r[i] = (double)a[i];
}
void orc_convert_FD(float *r, const double *a, size_t m)
{
    for (size_t i = 0; i < m; i++) r[i] = (float)a[i];
}

void orc_D_spproj(double *h, const double *d, int mu, int sign)
{
    D_cpx *H = (D_cpx *)h;
    const D_cpx *Dd = (const D_cpx *)d;
    for (int ic = 0; ic < 3; ic++) D_proj_row(&H[2 * ic], &H[2 * ic + 1], Dd + 4 * ic, mu, sign);
}
void orc_D_sprecon(double *d, const double *h, int mu, int sign)
{
    D_recon_t tab;
    if (!D_recon_table(&tab, mu, sign)) return;
    D_cpx *Dd = (D_cpx *)d;
    const D_cpx *H = (const D_cpx *)h;
    for (int ic = 0; ic < 3; ic++)
        for (int sp = 0; sp < 4; sp++)
            D_recon_apply(ORC_EQ, tab.op[sp], &Dd[4 * ic + sp], H[2 * ic + tab.src[sp]]);
}

void orc_F_M_veq_exp_M(int nc, float *r, const float *a, int n) { F_orc_M_veq_exp_M(nc, r, a, n); }
void orc_D_M_veq_exp_M(int nc, double *r, const double *a, int n) { D_orc_M_veq_exp_M(nc, r, a, n); }
void orc_F_M_veq_sqrt_M(int nc, float *r, const float *a, int n) { F_orc_M_veq_sqrt_M(nc, r, a, n); }
void orc_D_M_veq_sqrt_M(int nc, double *r, const double *a, int n) { D_orc_M_veq_sqrt_M(nc, r, a, n); }
void orc_F_M_veq_invsqrt_M(int nc, float *r, const float *a, int n) { F_orc_M_veq_invsqrt_M(nc, r, a, n); }
void orc_D_M_veq_invsqrt_M(int nc, double *r, const double *a, int n) { D_orc_M_veq_invsqrt_M(nc, r, a, n); }

void orc_F_matfunc(int fn, int nc, float *r, const float *a, const float *b, int n)
{ F_orc_matfunc(fn, nc, r, a, b, n); }
void orc_D_matfunc(int fn, int nc, double *r, const double *a, const double *b, int n)
{ D_orc_matfunc(fn, nc, r, a, b, n); }

void orc_F_vecop(int op, int eqop, int cj, int dim, int ncomp, float *r, const int *rindex,
                 const float *const *px, const float *const *py, float sr, float si, const int *index, int n)
{ F_orc_vecop(op, eqop, cj, dim, ncomp, r, rindex, px, py, sr, si, index, n); }
void orc_D_vecop(int op, int eqop, int cj, int dim, int ncomp, double *r, const int *rindex,
                 const double *const *px, const double *const *py, double sr, double si, const int *index, int n)
{ D_orc_vecop(op, eqop, cj, dim, ncomp, r, rindex, px, py, sr, si, index, n); }

void orc_F_fill(int kind, int eqop, int ncomp, float *r, const float *a, const int *index, int n)
{ F_orc_fill(kind, eqop, ncomp, r, a, index, n); }
void orc_D_fill(int kind, int eqop, int ncomp, double *r, const double *a, const int *index, int n)
{ D_orc_fill(kind, eqop, ncomp, r, a, index, n); }

void orc_F_gamma(int right, int ncomp, int mu, float *r, const int *rindex, const float *const *pa,
                 const int *index, int n)
{ F_orc_gamma(right, ncomp, mu, r, rindex, pa, index, n); }
void orc_D_gamma(int right, int ncomp, int mu, double *r, const int *rindex, const double *const *pa,
                 const int *index, int n)
{ D_orc_gamma(right, ncomp, mu, r, rindex, pa, index, n); }

void orc_F_local(int kind, int eqop, int ncomp, int rreals, float *r, const int *rindex, const float *const *pa,
                 const float *const *pb, const int *index, int n)
{ F_orc_local(kind, eqop, ncomp, rreals, r, rindex, pa, pb, index, n); }
void orc_D_local(int kind, int eqop, int ncomp, int rreals, double *r, const int *rindex, const double *const *pa,
                 const double *const *pb, const int *index, int n)
{ D_orc_local(kind, eqop, ncomp, rreals, r, rindex, pa, pb, index, n); }

void orc_F_sum(int ncr, double *out, const float *const *pa, const int *index, int n)
{ F_orc_sum(ncr, out, pa, index, n); }
void orc_D_sum(int ncr, double *out, const double *const *pa, const int *index, int n)
{ D_orc_sum(ncr, out, pa, index, n); }

void orc_F_fieldmul(int fcomplex, int fa, int cb, int dim, int eqop, int ncomp, float *r, const int *rindex,
                    const float *const *pf, const float *const *pb, const int *index, int n)
{ F_orc_fieldmul(fcomplex, fa, cb, dim, eqop, ncomp, r, rindex, pf, pb, index, n); }
void orc_D_fieldmul(int fcomplex, int fa, int cb, int dim, int eqop, int ncomp, double *r, const int *rindex,
                    const double *const *pf, const double *const *pb, const int *index, int n)
{ D_orc_fieldmul(fcomplex, fa, cb, dim, eqop, ncomp, r, rindex, pf, pb, index, n); }
void orc_F_outer(int eqop, float *r, const int *rindex, const float *const *pa, const float *const *pb,
                 const int *index, int n)
{ F_orc_outer(eqop, r, rindex, pa, pb, index, n); }
void orc_D_outer(int eqop, double *r, const int *rindex, const double *const *pa, const double *const *pb,
                 const int *index, int n)
{ D_orc_outer(eqop, r, rindex, pa, pb, index, n); }

void orc_F_pmul(int kind, int eqop, int ca, int cb, float *r, const int *rindex, const float *const *pa,
                const float *const *pb, const int *index, int n)
{ F_orc_pmul(kind, eqop, ca, cb, r, rindex, pa, pb, index, n); }
void orc_D_pmul(int kind, int eqop, int ca, int cb, double *r, const int *rindex, const double *const *pa,
                const double *const *pb, const int *index, int n)
{ D_orc_pmul(kind, eqop, ca, cb, r, rindex, pa, pb, index, n); }

void orc_F_comp(int kind, int t, const int *e, float *r, const int *rindex, const float *const *pa, const int *index,
                int n)
{ F_orc_comp(kind, t, e, r, rindex, pa, index, n); }
void orc_D_comp(int kind, int t, const int *e, double *r, const int *rindex, const double *const *pa,
                const int *index, int n)
{ D_orc_comp(kind, t, e, r, rindex, pa, index, n); }

void orc_F_sconv(int kind, float *r, const int *rindex, const float *const *pa, const float *const *pb,
                 const int *index, int n)
{ F_orc_sconv(kind, r, rindex, pa, pb, index, n); }
void orc_D_sconv(int kind, double *r, const int *rindex, const double *const *pa, const double *const *pb,
                 const int *index, int n)
{ D_orc_sconv(kind, r, rindex, pa, pb, index, n); }

void orc_F_sfn(int fn, float *r, const int *rindex, const float *const *pa, const void *const *pb,
               const int *index, int n)
{ F_orc_sfn(fn, r, rindex, pa, pb, index, n); }
void orc_D_sfn(int fn, double *r, const int *rindex, const double *const *pa, const void *const *pb,
               const int *index, int n)
{ D_orc_sfn(fn, r, rindex, pa, pb, index, n); }

void orc_F_mask(int op, int nw, void *r, const int *rindex, const float *const *pa, const void *const *pb,
                const int *index, int n)
{ F_orc_mask(op, nw, r, rindex, pa, pb, index, n); }
void orc_D_mask(int op, int nw, void *r, const int *rindex, const double *const *pa, const void *const *pb,
                const int *index, int n)
{ D_orc_mask(op, nw, r, rindex, pa, pb, index, n); }

/* integer fields, as the reference's generated C writes them (build_qla.pl:
 * 356-402, 1256-1294, 1421-1516, 1837-1888, 1936-2088, 2561-2620):
 * v = a, i*a, !a, a + b, a - b, a / b, a * b, i*a + b, i*a - b, a << b,
 * a >> b, a % b, max, min, ==, !=, >, <, >=, <=, |, &, ^, copymask, then
 * r = v, r += v, r -= v or r = -v; op 40 / 41 the fills r = 0 / r eqop i */
void orc_int(int op, int eqop, int *r, const int *rindex, const int *const *px, const int *const *py, int sc,
             const int *index, int n)
{
    for (int i = 0; i < n; i++) {
        const int s = index ? index[i] : i;
        int *R = r + (rindex ? rindex[i] : s);
        const int a = (op < 40) ? *px[s] : 0;
        const int b = (op >= 10 && op < 40) ? *py[s] : 0;
        int v;
        switch (op) {
        case 0: v = a; break;
        case 1: v = sc * a; break;
        case 2: v = !a; break;
        case 10: v = a + b; break;
        case 11: v = a - b; break;
        case 12: v = a / b; break;
        case 13: v = a * b; break;
        case 14: v = sc * a + b; break;
        case 15: v = sc * a - b; break;
        case 16: v = a << b; break;
        case 17: v = a >> b; break;
        case 18: v = a % b; break;
        case 19: v = a > b ? a : b; break;
        case 20: v = a > b ? b : a; break;
        case 21: v = ( a == b ); break;
        case 22: v = ( a != b ); break;
        case 23: v = ( a > b ); break;
        case 24: v = ( a < b ); break;
        case 25: v = ( a >= b ); break;
        case 26: v = ( a <= b ); break;
        case 27: v = ( a | b ); break;
        case 28: v = ( a & b ); break;
        case 29: v = ( a ^ b ); break;
        case 30: if (!b) continue; v = a; break;
        case 40: v = 0; break;
        default: v = sc; break;
        }
        switch (eqop) {
        case 0: *R = v; break;
        case 1: *R += v; break;
        case 2: *R -= v; break;
        default: *R = -v; break;
        }
    }
}

void orc_set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}
int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- benchmark loops of the CPU baseline (oracle/cpu_bench.py) ------------
 * QLA_F3_H_veq_spproj_M_times_pD in the v-form (a a field, b a QLA pointer
 * array) and QLA_F3_D_vpeq_sprecon_M_times_pH in the SSE build's order, each
 * with mu and sign constant inside its site loop, as the generated code has
 * them (one loop per case of the switch outside it, expressions_gamma.pl:
 * 349-381), so the spin tables fold away.  Same helpers, same order, same
 * results as orc_F3_site_vf / orc_F3_sse_site_vf (tests/test_oracle_sse.py). */
ORC_HOT void bench_spproj_site(float *r, const float *af, const float *b, int mu, int sign)
{
    F_cpx t[6];
    for (int ic = 0; ic < 3; ic++) F_proj_row(&t[2 * ic], &t[2 * ic + 1], (const F_cpx *)b + 4 * ic, mu, sign);
    F_mat_times_cols(ORC_EQ, 0, (F_cpx *)r, (const F_cpx *)af, t, 2);
}

ORC_HOT void bench_sprecon_sse_site(float *r, const float *af, const float *b, int mu, int sign, int aligned)
{
    F_recon_t tab;
    F_recon_table(&tab, mu, sign);
    const F_cpx *A = (const F_cpx *)af, *B = (const F_cpx *)b;
    F_cpx *R = (F_cpx *)r, tv[6];
    for (int ic = 0; ic < 3; ic++)
        for (int s = 0; s < 2; s++)
            tv[2 * ic + s] = aligned ? F_sse_half(A + 3 * ic, B + s, 2) : F_sse_scalar(A + 3 * ic, B + s, 2);
    for (int ic = 0; ic < 3; ic++)
        for (int sp = 0; sp < 4; sp++) F_recon_apply(ORC_PEQ, tab.op[sp], &R[4 * ic + sp], tv[2 * ic + tab.src[sp]]);
}

#define BENCH_CASES(X) X(0, 1) X(1, 1) X(2, 1) X(3, 1) X(4, 1) X(0, -1) X(1, -1) X(2, -1) X(3, -1) X(4, -1)

void orc_F3_bench_spproj(float *r, const float *af, const float *const *pb, int mu, int sign, int n)
{
    if (mu < 0 || mu > 4 || n <= 0) return;
    switch (mu + (sign == 1 ? 0 : 5)) {
#define X(MU, SG)                                                                                              \
    case MU + (SG == 1 ? 0 : 5):                                                                               \
        _Pragma("omp parallel for schedule(static)") for (int i = 0; i < n; i++)                              \
            bench_spproj_site(r + 12 * (size_t)i, af + 18 * (size_t)i, pb[i], MU, SG);                         \
        break;
        BENCH_CASES(X)
#undef X
    }
}

void orc_F3_bench_sprecon_sse(float *r, const float *af, const float *const *pb, int mu, int sign, int n)
{
    if (mu < 0 || mu > 4 || n <= 0) return;
    const int aligned = (size_t)r % 16 == 0 && (size_t)af % 8 == 0 && (size_t)pb[0] % 16 == 0;
    switch (mu + (sign == 1 ? 0 : 5)) {
#define X(MU, SG)                                                                                              \
    case MU + (SG == 1 ? 0 : 5):                                                                               \
        if (aligned) {                                                                                         \
            _Pragma("omp parallel for schedule(static)") for (int i = 0; i < n; i++)                          \
                bench_sprecon_sse_site(r + 24 * (size_t)i, af + 18 * (size_t)i, pb[i], MU, SG, 1);             \
        } else {                                                                                               \
            _Pragma("omp parallel for schedule(static)") for (int i = 0; i < n; i++)                          \
                bench_sprecon_sse_site(r + 24 * (size_t)i, af + 18 * (size_t)i, pb[i], MU, SG, 0);             \
        }                                                                                                      \
        break;
        BENCH_CASES(X)
#undef X
    }
}
