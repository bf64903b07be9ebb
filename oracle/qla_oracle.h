/*
 * qla_oracle.h -- CPU restatement of the QLA 1.9.0 site-loop hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP library, not a
 * product path: only the tests, the smoke check and the CPU-baseline leg of
 * the benchmark may load it.  The product library never links or calls it.
 *
 * Each function restates the arithmetic, operation order and edge behaviour
 * of the generated reference function named in its comment (generated code
 * is specified by the Perl emitters under lib/perl; file:line citations are relative to the
 * reference tree).  Records are site-major AoS exactly as in
 * include/qla_types.h:29-33: complex {re,im}; V[3], H[3][2], D[3][4], M[3][3].
 * Pointer-array arguments hold host addresses of records.
 *
 * eqop codes: 0 = eq, 1 = peq, 2 = meq, 3 = eqm (lib/perl/operatortypes.pl).
 */
#ifndef QLA_ORACLE_H
#define QLA_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_EQ = 0, ORC_PEQ = 1, ORC_MEQ = 2, ORC_EQM = 3 };

/* matrix times colour vector, generic order (expressions_tensor.pl:544-630) */
void orc_F3_V_eqop_M_times_V(int eqop, int adj, float *r, const float *a, const float *b,
                             const float *const *pb, const int *index, int n);
void orc_D3_V_eqop_M_times_V(int eqop, int adj, double *r, const double *a, const double *b,
                             const double *const *pb, const int *index, int n);

/* norm2 / dot (make_code_unary.pl:318-366, make_code_binary.pl:110-170).
 * ncomp = complex components per site (3 for V, 12 for D).  Results are the
 * promoted (double) sums; the F3 entry points round them to float. */
double orc_F_norm2(const float *a, int ncomp, int n);
double orc_D_norm2(const double *a, int ncomp, int n);
void orc_F_dot(double *re, double *im, const float *a, const float *b, int ncomp, int n);
void orc_D_dot(double *re, double *im, const double *a, const double *b, int ncomp, int n);
/* long-double references (the reference test suite's own method,
 * test/src/protocol_idx.m4:128-183) */
long double orc_F_norm2_ld(const float *a, int ncomp, int n);
long double orc_D_norm2_ld(const double *a, int ncomp, int n);
void orc_F_dot_ld(long double *re, long double *im, const float *a, const float *b, int ncomp, int n);
void orc_D_dot_ld(long double *re, long double *im, const double *a, const double *b, int ncomp, int n);

/* spin projection + matrix multiply (make_code_binary.pl:189-222), index may be
 * NULL (v-form) or the gang index (x-form). */
void orc_F3_H_eqop_spproj_M_times_pD(int eqop, int adj, float *r, const float *a,
                                     const float *const *b, int mu, int sign,
                                     const int *index, int n);
void orc_D3_H_eqop_spproj_M_times_pD(int eqop, int adj, double *r, const double *a,
                                     const double *const *b, int mu, int sign,
                                     const int *index, int n);
/* matrix multiply + spin reconstruction (make_code_binary.pl:224-257) */
void orc_F3_D_eqop_sprecon_M_times_pH(int eqop, int adj, float *r, const float *a,
                                      const float *const *b, int mu, int sign,
                                      const int *index, int n);
void orc_D3_D_eqop_sprecon_M_times_pH(int eqop, int adj, double *r, const double *a,
                                      const double *const *b, int mu, int sign,
                                      const int *index, int n);
/* single-site spin projection / reconstruction (eq), used to pin the tables */
void orc_D_spproj(double *h, const double *d, int mu, int sign);
void orc_D_sprecon(double *d, const double *h, int mu, int sign);

/* matrix functions (lib/perl/code-templates/M_eq_{exp,sqrt,invsqrt}_M.c);
 * nc = colours, matrices are nc x nc complex, row-major. */
void orc_F_M_veq_exp_M(int nc, float *r, const float *a, int n);
void orc_D_M_veq_exp_M(int nc, double *r, const double *a, int n);
void orc_F_M_veq_sqrt_M(int nc, float *r, const float *a, int n);
void orc_D_M_veq_sqrt_M(int nc, double *r, const double *a, int n);
void orc_F_M_veq_invsqrt_M(int nc, float *r, const float *a, int n);
void orc_D_M_veq_invsqrt_M(int nc, double *r, const double *a, int n);

/* thread count the OpenMP site loops use (0 = OpenMP default) */
void orc_set_threads(int nthreads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
