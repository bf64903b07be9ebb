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
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_EQ = 0, ORC_PEQ = 1, ORC_MEQ = 2, ORC_EQM = 3 };

/* Site operations (qla_oracle_impl.inc): the A and B records of site s are
 * *pa[s] and *pb[s] (a field operand is passed as the pointer array of its
 * own records); index NULL = v-form, else the gang index (x-form).
 * kind 0: V eqop A*B (expressions_tensor.pl:544-630)
 * kind 1: H eqop A*spproj(B) (make_code_binary.pl:189-222)
 * kind 2: D eqop sprecon(A*B) (make_code_binary.pl:224-257)
 * kind 3: D eqop sprecon(A*spproj(B)) (make_code_binary.pl:285-341)
 * kind 5: H eqop spproj(B) (make_code_unary.pl:183-202; pa unused, may be NULL)
 * kind 6: D eqop sprecon(B) (make_code_unary.pl:204-223)
 * kind 7: D eqop sprecon(spproj(B)) (make_code_unary.pl:243-283)
 * kinds 8, 9, 10: H, D, M eqop A*B; kind 11: M eqop A*B^dagger
 *   (build_qla.pl:1587-1622, expressions_tensor.pl:544-630)
 * adj selects A = a^dagger (QLA_c_peq_ca_times_c).  rindex != NULL puts the
 * result of step i at r[rindex[i]] (the per-operand destination index of the
 * QLA_INDEX_ALL forms, indirection_all.pl:66-84).  mu outside 0..4 is a
 * no-op for kinds 1-3. */
/* the operation order of a QLA built --enable-sse3 for V_vpeq_M_times_pV
 * (kind 0, PEQ), V_veq_M_times_pV (0, EQ), V_veq_Ma_times_V (0, EQ, adj),
 * D_vpeq_sprecon_M_times_pH (2, PEQ) and D_vpeq_spproj_M_times_pD (3, PEQ);
 * r_al / a_al / b0_al: byte offsets mod 16 of r, a, b[0] in the modelled call
 * (-1: these arrays' own) -- qla_oracle_sse.inc */
void orc_F3_sse_site(int kind, int eqop, int adj, float *r, const float *const *pa, const float *const *pb, int mu,
                     int sign, int n, int r_al, int a_al, int b0_al);
/* orc_F3_site / orc_D3_site with the colour matrices a contiguous field af */
void orc_F3_site_vf(int kind, int eqop, int adj, float *r, const float *af, const float *const *pb, int mu, int sign,
                    const int *index, int n);
void orc_D3_site_vf(int kind, int eqop, int adj, double *r, const double *af, const double *const *pb, int mu,
                    int sign, const int *index, int n);
/* QLA_D3_V_{veq,vpeq}_M_times_{V,pV} in the order of the SSE2 versions a
 * QLA built --enable-sse3 links (eqop ORC_EQ or ORC_PEQ) -- qla_oracle_sse.inc */
void orc_D3_sse2_site(int eqop, double *r, const double *const *pa, const double *const *pb, int n);
void orc_F3_sse_site_vf(int kind, int eqop, int adj, float *r, const float *af, const float *const *pb, int mu,
                        int sign, int n, int r_al, int a_al, int b0_al);
/* the CPU baseline's loops (mu / sign constant per loop): v-form
 * H = spproj(M pD) and D += sprecon(M pH) in the SSE build's order */
void orc_F3_bench_spproj(float *r, const float *af, const float *const *pb, int mu, int sign, int n);
void orc_F3_bench_sprecon_sse(float *r, const float *af, const float *const *pb, int mu, int sign, int n);
void orc_F3_site(int kind, int eqop, int adj, float *r, const int *rindex, const float *const *pa,
                 const float *const *pb, int mu, int sign, const int *index, int n);
void orc_D3_site(int kind, int eqop, int adj, double *r, const int *rindex, const double *const *pa,
                 const double *const *pb, int mu, int sign, const int *index, int n);
/* multi-source r eqop sum_{d<nd} A_d*B_d, A_d = *pa[d][s], B_d = *pb[d][s]
 * (build_qla.pl:1693-1725, expressions_tensor.pl:983-1002); destination
 * r[rindex[i]] when rindex != NULL, else r[s] */
void orc_F3_multi(int eqop, int adj, float *r, const int *rindex, const float *const *const *pa,
                  const float *const *const *pb, int nd, const int *index, int n);
void orc_D3_multi(int eqop, int adj, double *r, const int *rindex, const double *const *const *pa,
                  const double *const *const *pb, int nd, const int *index, int n);

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
/* the reference's rounded terms (component norms / double dot terms) summed
 * in long double; *sa = sum |term| */
long double orc_F_norm2_terms_ld(const float *a, int ncomp, int n);
long double orc_D_norm2_terms_ld(const double *a, int ncomp, int n);
void orc_F_dot_terms_ld(long double *re, long double *im, long double *sa, const float *a, const float *b, int ncomp,
                        int n);
void orc_D_dot_terms_ld(long double *re, long double *im, long double *sa, const double *a, const double *b,
                        int ncomp, int n);
void orc_F_dot_ld(long double *re, long double *im, const float *a, const float *b, int ncomp, int n);
void orc_D_dot_ld(long double *re, long double *im, const double *a, const double *b, int ncomp, int n);

/* precision conversion of m reals: float -> double, double -> float
 * (build_qla.pl:530-585, the C conversions) */
void orc_convert_DF(double *r, const float *a, size_t m);
void orc_convert_FD(float *r, const double *a, size_t m);

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
/* the other matrix functions (qla_oracle_matfunc.inc), n records each:
 * fn 0 exp, 1 sqrt, 2 invsqrt (as above), 3 det (C_eq_det_M.c), 4 inverse
 * (M_eq_inverse_M.c), 5 a^-1 b for a vector b (V_eq_M_inverse_V.c), 6 a^-1 b
 * for a matrix b (M_eq_M_inverse_M.c), 7 expA, 8 expTA, 9 sqrtPH,
 * 10 invsqrtPH, 11 log, 12 Hermitian eigenvalues (nc = 3 only).  Results:
 * one complex for det, nc complexes for 5 and 12, an nc x nc matrix else.
 * b is read by fn 5 and 6 only. */
enum { OF_EXP = 0, OF_SQRT, OF_INVSQRT, OF_DET, OF_INVERSE, OF_INV_V, OF_INV_M, OF_EXPA, OF_EXPTA, OF_SQRTPH,
       OF_INVSQRTPH, OF_LOG, OF_EIGH };
void orc_F_matfunc(int fn, int nc, float *r, const float *a, const float *b, int n);
void orc_D_matfunc(int fn, int nc, double *r, const double *a, const double *b, int n);

/* field linear algebra (qla_oracle_vecops.inc): n steps of r eqop op(x, y)
 * over records of ncomp complexes.  op 0 copy (cj: Hermitian / complex
 * conjugate; build_qla.pl:356-494), 1 times i (:1328-1354), 2 times the real
 * constant sr (:1256-1293), 3 times the complex constant (sr, si)
 * (:1295-1326), 4 x + y, 5 x - y (:1421-1456; eq only), 6 sr*x + y,
 * 7 sr*x - y (:1837-1887), 8 (sr,si)*x + y, 9 (sr,si)*x - y (:1889-1935).
 * dim != 0 reads the source transposed as a dim x dim matrix (M: 3, P: 12).
 * Records: destination r[rindex[i]] or r[s], sources *px[s], *py[s] with
 * s = index[i] or i, as for the site operations. */
enum { ORC_VOP_COPY = 0, ORC_VOP_I, ORC_VOP_RSCALE, ORC_VOP_CSCALE, ORC_VOP_PLUS, ORC_VOP_MINUS, ORC_VOP_R_PLUS,
       ORC_VOP_R_MINUS, ORC_VOP_C_PLUS, ORC_VOP_C_MINUS };
void orc_F_vecop(int op, int eqop, int cj, int dim, int ncomp, float *r, const int *rindex,
                 const float *const *px, const float *const *py, float sr, float si, const int *index, int n);
void orc_D_vecop(int op, int eqop, int cj, int dim, int ncomp, double *r, const int *rindex,
                 const double *const *px, const double *const *py, double sr, double si, const int *index, int n);

/* fills (qla_oracle_vecops.inc): kind 0 zero, 1 the record a broadcast,
 * 2 (ncomp 9) a[0] on the diagonal; n steps writing r[index[i]] or r[i] */
void orc_F_fill(int kind, int eqop, int ncomp, float *r, const float *a, const int *index, int n);
void orc_D_fill(int kind, int eqop, int ncomp, double *r, const double *a, const int *index, int n);

/* gamma-matrix multiplication (qla_oracle_vecops.inc): r = Gamma_mu a
 * (right 0; ncomp 12 or 144) or a Gamma_mu (right 1; 144) */
void orc_F_gamma(int right, int ncomp, int mu, float *r, const int *rindex, const float *const *pa,
                 const int *index, int n);
void orc_D_gamma(int right, int ncomp, int mu, double *r, const int *rindex, const double *const *pa,
                 const int *index, int n);

/* site-local reductions and parts (qla_oracle_vecops.inc): kind 0 R eqop
 * norm2(a), 1 C eqop a.b, 2 R eqop Re a.b (a, b records of ncomp
 * complexes), 3 R = Re tr a, 4 R = Im tr a, 5 C = tr a, 6 M = the traceless
 * anti-Hermitian part of a (a colour matrices); rreals = reals per
 * destination record (1, 2 or 18) */
void orc_F_local(int kind, int eqop, int ncomp, int rreals, float *r, const int *rindex, const float *const *pa,
                 const float *const *pb, const int *index, int n);
void orc_D_local(int kind, int eqop, int ncomp, int rreals, double *r, const int *rindex, const double *const *pa,
                 const double *const *pb, const int *index, int n);

/* global sum of a field (qla_oracle_vecops.inc): the ncr real columns of
 * the records *pa[s] summed in double over the n steps */
void orc_F_sum(int ncr, double *out, const float *const *pa, const int *index, int n);
void orc_D_sum(int ncr, double *out, const double *const *pa, const int *index, int n);

/* multiplication by a real (fcomplex 0) or complex field and the colour
 * outer product (qla_oracle_vecops.inc) */
void orc_F_fieldmul(int fcomplex, int fa, int cb, int dim, int eqop, int ncomp, float *r, const int *rindex,
                    const float *const *pf, const float *const *pb, const int *index, int n);
void orc_D_fieldmul(int fcomplex, int fa, int cb, int dim, int eqop, int ncomp, double *r, const int *rindex,
                    const double *const *pf, const double *const *pb, const int *index, int n);
void orc_F_outer(int eqop, float *r, const int *rindex, const float *const *pa, const float *const *pb,
                 const int *index, int n);
void orc_D_outer(int eqop, double *r, const int *rindex, const double *const *pa, const double *const *pb,
                 const int *index, int n);

/* products with propagators (qla_oracle_vecops.inc): kind 0 P eqop M*P,
 * 1 P eqop P*M, 2 P eqop P*P; ca / cb the adjoint of either factor */
void orc_F_pmul(int kind, int eqop, int ca, int cb, float *r, const int *rindex, const float *const *pa,
                const float *const *pb, const int *index, int n);
void orc_D_pmul(int kind, int eqop, int ca, int cb, double *r, const int *rindex, const double *const *pa,
                const double *const *pb, const int *index, int n);

/* component access and the spin trace (qla_oracle_vecops.inc) */
enum { ORC_T_V = 0, ORC_T_H = 1, ORC_T_D = 2, ORC_T_M = 3, ORC_T_P = 4 };
enum { ORC_GET_ELEM = 0, ORC_SET_ELEM = 1, ORC_GET_COLORVEC = 2, ORC_SET_COLORVEC = 3, ORC_GET_DIRACVEC = 4,
       ORC_SET_DIRACVEC = 5, ORC_SPINTRACE = 6 };
void orc_F_comp(int kind, int t, const int *e, float *r, const int *rindex, const float *const *pa, const int *index,
                int n);
void orc_D_comp(int kind, int t, const int *e, double *r, const int *rindex, const double *const *pa,
                const int *index, int n);

/* real / complex conversions and scalar-field division (qla_oracle_vecops.inc) */
void orc_F_sconv(int kind, float *r, const int *rindex, const float *const *pa, const float *const *pb,
                 const int *index, int n);
void orc_D_sconv(int kind, double *r, const int *rindex, const double *const *pa, const double *const *pb,
                 const int *index, int n);

/* elementary functions of real and complex fields (qla_oracle_vecops.inc) */
void orc_F_sfn(int fn, float *r, const int *rindex, const float *const *pa, const void *const *pb,
               const int *index, int n);
void orc_D_sfn(int fn, double *r, const int *rindex, const double *const *pa, const void *const *pb,
               const int *index, int n);

/* copymask (op < 0) and the comparisons of real fields (qla_oracle_vecops.inc) */
void orc_F_mask(int op, int nw, void *r, const int *rindex, const float *const *pa, const void *const *pb,
                const int *index, int n);
void orc_D_mask(int op, int nw, void *r, const int *rindex, const double *const *pa, const void *const *pb,
                const int *index, int n);

/* integer fields (qla_oracle.c) */
void orc_int(int op, int eqop, int *r, const int *rindex, const int *const *px, const int *const *py, int sc,
             const int *index, int n);

/* thread count the OpenMP site loops use (0 = OpenMP default) */
void orc_set_threads(int nthreads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
