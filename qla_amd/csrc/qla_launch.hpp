// qla_launch.hpp -- internal launch interface between the C ABI layer
// (qla_capi.cpp) and the kernel translation units.  prec: 0 = float, 1 = double.
#pragma once
#include <hip/hip_runtime.h>

// ---------------------------------------------------------------------------
// Site operations (qla_site.hip, qla_multi.hip).  kind: 0 V = M*V, 1 H =
// M*spproj(D), 2 D = sprecon(M*H), 3 D = sprecon(M*spproj(D)); eqop 0 eq,
// 1 peq, 2 meq, 3 eqm; adj selects M^dagger; xf the gang-indexed x-form.
// ---------------------------------------------------------------------------
struct qla_site_args {
    void *r;            // destination field
    const void *a;      // ColorMatrix field, or pointer array when pa
    const void *b;      // second source field, or pointer array when pb
    const int *index;   // gang index (x-form)
    const int *ir;      // per-operand index arrays (QLA_INDEX_ALL forms): the
    const int *ia;      // record of site i is r[ir[i]], a[ia[i]], b[ib[i]]; an
    const int *ib;      // operand without one uses the gang index or i
    int n;              // sites (v-form) / index entries (x-form)
    int al;             // 16-byte alignment of the contiguous fields: 1 a, 2 b, 4 r;
                        // 8-byte alignment: 8 r, 16 a
    int order;          // 1: the SSE3-build order (F3 V_vpeq_M_times_pV,
                        // V_veq_M_times_pV, V_veq_Ma_times_V, the vpeq sprecon
                        // and spproj-multiply-reconstruct), else 0 (generic)
    int ms;             // mu + 5 * (sign != 1), 0..9
    int pa, pb;         // a / b are pointer arrays
};
// xf: per-lane addressing (a gang index or any per-operand index);
// tpw: tiles per wave of the persistent grid (launch shape only)
hipError_t qla_launch_site(int prec, int kind, int eqop, int adj, int xf, const qla_site_args &p, int tpw,
                           hipStream_t st);

// Multi-source matrix times vector, r eqop sum_d A_d * B_d for d < nd <= QLA_MULTI_MAXND
// (the C ABI layer splits longer sums into peq/meq launches).
#define QLA_MULTI_MAXND 16
struct qla_multi_args {
    void *r;
    const int *index;                 // gang index (x-form)
    const int *ir, *ia, *ib;          // per-operand indices (QLA_INDEX_ALL forms), shared by the terms
    int n, al, nd, pa, pb;
    const void *a[QLA_MULTI_MAXND];   // per term: ColorMatrix field or pointer array
    const void *b[QLA_MULTI_MAXND];   // per term: ColorVector field or pointer array
};
hipError_t qla_launch_multi(int prec, int eqop, int adj, int xf, const qla_multi_args &p, int tpw, hipStream_t st);

// Global reductions over m complex elements.  kind: 0 = norm2 (b unused),
// 1 = dot conj(a).b.  The result is written to the DEVICE address `out` as
// out_kind 0 double, 1 float, 2 double complex, 3 float complex.  `scratch`
// must hold qla_reduce_scratch_doubles() + 4 = QLA_REDUCE_SCRATCH_DOUBLES + 4
// doubles: the per-block partials, a result slot of two doubles, and in the
// fourth double of the slot the single-launch form's ticket, which must be 0
// before the first call (each call leaves it at 0).
#define QLA_REDUCE_SCRATCH_DOUBLES (2 * 2048)
int qla_reduce_scratch_doubles(void);
// partial slots stage 1 of qla_launch_reduce2 writes for these operands
int qla_reduce2_partials(int prec, int kind, const void *a, const void *b, long long m);
hipError_t qla_launch_reduce2(int prec, int kind, const void *a, const void *b, long long m, void *out,
                              int out_kind, double *scratch, hipStream_t st);

// The same reductions over gathered records: record of site i at *a[s]
// (pa) or a + s*ncomp complexes, s = ia[i] (per-operand index), index[i]
// (gang index) or i, likewise for b with ib; ncomp one of 3, 6, 9, 12, 144.  out_kind 4 / 5: real part of a dot as double / float.
hipError_t qla_launch_reduce_gather(int prec, int kind, int ncomp, const void *a, const void *b, const int *index,
                                    const int *ia, const int *ib, int n, int pa, int pb, void *out, int out_kind,
                                    double *scratch, hipStream_t st);

// Precision conversion of n records of ncomp complexes: to_float 0 float ->
// double, 1 double -> float (round to nearest).  Destination record of site
// i at ir[i], index[i] (x-form) or i; source record at ia[i], index[i] or i,
// read through the pointer array when pa.
hipError_t qla_launch_precision(int to_float, int ncomp, void *r, const void *a, const int *index, const int *ir,
                                const int *ia, int n, int pa, hipStream_t st);

// Colour-matrix functions on n nc x nc matrices.  fn: 0 exp, 1 sqrt,
// 2 invsqrt, 3 det, 4 inverse, 5 a^-1 b (vector b), 6 a^-1 b (matrix b),
// 7 expA, 8 expTA, 9 sqrtPH, 10 invsqrtPH, 11 log, 12 Hermitian eigenvalues
// (nc = 3 only).  Step i reads a at ia[i], index[i] (x-form) or i, through
// the pointer array when pa (likewise b), and writes r at ir[i], index[i] or
// i.  prec 0: nc = 3..8; prec 1: nc = 3.  n < 2^31 (one thread per matrix).
struct qla_matfunc_args {
    void *r;
    const void *a, *b;
    const int *index, *ir, *ia, *ib;
    int n, pa, pb;
};
hipError_t qla_launch_matfunc(int prec, int fn, int nc, const qla_matfunc_args &p, hipStream_t st);
hipError_t qla_matfunc_launch_u0(int fn, const qla_matfunc_args &p, hipStream_t st);
hipError_t qla_matfunc_launch_u1(int fn, const qla_matfunc_args &p, hipStream_t st);
hipError_t qla_matfunc_launch_u2(int fn, const qla_matfunc_args &p, hipStream_t st);
hipError_t qla_matfunc_launch_u3(int fn, const qla_matfunc_args &p, hipStream_t st);
hipError_t qla_matfunc_launch_u4(int fn, const qla_matfunc_args &p, hipStream_t st);
hipError_t qla_matfunc_launch_u5(int fn, const qla_matfunc_args &p, hipStream_t st);
hipError_t qla_matfunc_launch_u6(int fn, const qla_matfunc_args &p, hipStream_t st);
hipError_t qla_matfunc_launch_u7(int fn, const qla_matfunc_args &p, hipStream_t st);
hipError_t qla_matfunc_launch_u8(int fn, const qla_matfunc_args &p, hipStream_t st);

// Field linear algebra (qla_vecops.hip): n steps of r eqop op(x, y) over
// records of ncomp complexes (3, 6, 9, 12 or 144).  op: 0 copy, 1 times i,
// 2 times the real constant, 3 times the complex constant (eqop 0..3);
// 4 x + y, 5 x - y, 6 s*x + y, 7 s*x - y with a real s, 8 / 9 the same with
// a complex s (eqop 0 only); 10 / 11 times a real / complex factor field
// (eqop 0..3).  conj conjugates and trans transposes (records
// of 9 or 144 complexes, as 3 x 3 or 12 x 12 matrices) the source x.  Step i
// writes r at ir[i], index[i] (x-form) or i and reads x at ix[i], index[i] or
// i, through the pointer array when px (likewise y).
struct qla_vec_args {
    void *r;
    const void *x, *y;                  // field sources, or pointer arrays when px / py
    const void *sp;                     // the constant in device memory (real, or re and im), or null
    double s_re, s_im;                  // the constant's value when sp is null
    const int *index, *ir, *ix, *iy;
    int n, px, py;
    int conj, trans;
    int al;                             // r, x and y 16-byte aligned (the contiguous v-form path)
    int wide;                           // n * ncomp >= 2^32 (set by the launcher)
    // ops 10 / 11: a real / complex factor field, one value per step at
    // ifa[i], index[i] or i, through the pointer array when pf; fconj
    // conjugates the complex factor
    const void *f;
    const int *ifa;
    int pf, fconj;
};
hipError_t qla_launch_vec(int prec, int op, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st);
hipError_t qla_vec_launch_p0(int op, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st);
hipError_t qla_vec_launch_p1(int op, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st);

// Colour outer product (qla_vecops.hip): M eqop x y^dagger for colour-vector
// records x, y, addressed as in qla_vec_args.
hipError_t qla_launch_outer(int prec, int eqop, const qla_vec_args &p, hipStream_t st);
hipError_t qla_outer_launch_p0(int eqop, const qla_vec_args &p, hipStream_t st);
hipError_t qla_outer_launch_p1(int eqop, const qla_vec_args &p, hipStream_t st);
// Products with propagators (qla_vecops.hip): P eqop A * B, kind 0 M*P,
// 1 P*M, 2 P*P; ca / cb the adjoint of A / B
hipError_t qla_launch_pmul(int prec, int kind, int eqop, int ca, int cb, const qla_vec_args &p, hipStream_t st);
hipError_t qla_pmul_launch_p0(int kind, int eqop, int ca, int cb, const qla_vec_args &p, hipStream_t st);
hipError_t qla_pmul_launch_p1(int kind, int eqop, int ca, int cb, const qla_vec_args &p, hipStream_t st);
// Component access and the spin trace (qla_vecops.hip): mode 0 extract, 1
// insert, 2 spin trace; the source in x (px, ix), the destination in r (ir),
// index the x-form gang index.  Element c of the moved part sits at
// off + (c / inner) * s1 + (c % inner) * s2 of the array record; nsrc / ndst
// are the source / destination record sizes in complexes.
struct qla_comp_map {
    int nd, nsrc, ndst, off, inner, s1, s2;
};
hipError_t qla_launch_comp(int prec, int mode, const qla_vec_args &p, const qla_comp_map &m, hipStream_t st);
hipError_t qla_comp_launch_p0(int mode, const qla_vec_args &p, const qla_comp_map &m, hipStream_t st);
hipError_t qla_comp_launch_p1(int mode, const qla_vec_args &p, const qla_comp_map &m, hipStream_t st);
// Real / complex conversions and scalar-field division (qla_vecops.hip):
// kind 0 C = R, 1 C = R + iR, 2 R = re C, 3 R = im C, 4 R = R/R, 5 C = C/C
hipError_t qla_launch_sconv(int prec, int kind, const qla_vec_args &p, hipStream_t st);
hipError_t qla_sconv_launch_p0(int kind, const qla_vec_args &p, hipStream_t st);
hipError_t qla_sconv_launch_p1(int kind, const qla_vec_args &p, hipStream_t st);
// Elementary functions of real and complex fields (qla_vecops.hip): fn 0..16
// R = f(R), 20..25 R = a op b (25: b an int field), 30 C = cexpi R, 31 norm,
// 32 arg, 33 cexp, 34 csqrt, 35 clog
hipError_t qla_launch_sfn(int prec, int fn, const qla_vec_args &p, hipStream_t st);
hipError_t qla_sfn_launch_p0(int fn, const qla_vec_args &p, hipStream_t st);
hipError_t qla_sfn_launch_p1(int fn, const qla_vec_args &p, hipStream_t st);
// Integer fields (qla_vecops.hip): op and eqop as k_int documents; the int
// constant in sp (device) or s_re
hipError_t qla_int_launch(int op, int eqop, const qla_vec_args &p, hipStream_t st);
// Copymask and the comparisons of real fields (qla_vecops.hip): r = a where
// the int field y is nonzero (nw words of the field type per record); the
// int r = (a op b) for op 0 ==, 1 !=, 2 >, 3 <, 4 >=, 5 <=
hipError_t qla_launch_mask(int prec, int nw, const qla_vec_args &p, hipStream_t st);
hipError_t qla_mask_launch_p0(int nw, const qla_vec_args &p, hipStream_t st);
hipError_t qla_mask_launch_p1(int nw, const qla_vec_args &p, hipStream_t st);
hipError_t qla_launch_cmp(int prec, int op, const qla_vec_args &p, hipStream_t st);
hipError_t qla_cmp_launch_p0(int op, const qla_vec_args &p, hipStream_t st);
hipError_t qla_cmp_launch_p1(int op, const qla_vec_args &p, hipStream_t st);

// Fills (qla_vecops.hip), n steps writing r at index[i] or i, records of
// ncomp complexes: kind 0 r = 0 (eqop 0 only); 1 r eqop a, one record a
// broadcast to every step; 2 (ncomp 9) r eqop c on the diagonal and eqop 0
// elsewhere.  The record or constant is read from src (device memory) or,
// when src is null, copied from host into the launch at the call.
struct qla_fill_args {
    void *r;
    const void *src;        // device record / constant, or null
    const void *host;       // host record / constant when src is null
    const int *index;
    int n;
};
hipError_t qla_launch_fill(int prec, int kind, int eqop, int ncomp, const qla_fill_args &p, hipStream_t st);
hipError_t qla_fill_launch_p0(int kind, int eqop, int ncomp, const qla_fill_args &p, hipStream_t st);
hipError_t qla_fill_launch_p1(int kind, int eqop, int ncomp, const qla_fill_args &p, hipStream_t st);

// Gamma-matrix multiplication (qla_vecops.hip): r = Gamma_g a (right == 0;
// records of 12 or 144 complexes) or r = a Gamma_g (right == 1; 144), with
// Gamma_g = gamma_0^b0 gamma_1^b1 gamma_2^b2 gamma_3^b3 for the bits b of
// g = 0..15.  Addressing as qla_vec_args (x, px, ix; no y, no constant).
hipError_t qla_launch_gamma(int prec, int right, int g, int ncomp, const qla_vec_args &p, hipStream_t st);
hipError_t qla_gamma_launch_p0(int right, int g, int ncomp, const qla_vec_args &p, hipStream_t st);
hipError_t qla_gamma_launch_p1(int right, int g, int ncomp, const qla_vec_args &p, hipStream_t st);

// Site-local reductions and parts (qla_vecops.hip), one destination record
// per step: kind 0 R eqop norm2(x), 1 C eqop x.y, 2 R eqop Re x.y (records of
// ncomp complexes), 3 R = Re tr x, 4 R = Im tr x, 5 C = tr x, 6 M = the
// traceless anti-Hermitian part of x (colour matrices; kinds 3-6 eq only).
// Addressing as qla_vec_args; r holds Real, Complex or ColorMatrix records.
hipError_t qla_launch_local(int prec, int kind, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st);
hipError_t qla_local_launch_p0(int kind, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st);
hipError_t qla_local_launch_p1(int kind, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st);

// Global sums of fields (qla_reduce.hip): the ncr real columns of the records
// (ncr <= 288) summed over n steps (record index[i] or i, through the pointer
// array when pa) in double, written to the DEVICE address out as float
// (out_double 0) or double; prec_in 0 float, 1 double fields.  scratch holds
// qla_sum_scratch_doubles() doubles (its last 288 are free for a result).
int qla_sum_scratch_doubles(void);
hipError_t qla_launch_sum(int prec_in, int out_double, int ncr, const void *a, const int *index, int n, int pa,
                          void *out, double *scratch, hipStream_t st);

// Compressed-precision forms.  fmt: 0 bfloat16 links, 1 binary16 links.
hipError_t qla_launch_spproj_c(int fmt, void *r, const void *a, const void *b, int mu, int sign, int n,
                               hipStream_t st);
hipError_t qla_launch_sprecon_c(int fmt, void *r, const void *a, const void *b, int mu, int sign, int n,
                                hipStream_t st);
void qla_set_cmp_tiles_per_wave(int tpw);
// kind: 0 link pack, 1 link unpack, 2 D pack, 3 D unpack, 4 H pack, 5 H unpack
hipError_t qla_launch_convert(int kind, int fmt, void *r, const void *a, int n, hipStream_t st);
// norm2 (b == NULL) or dot of decoded Dh fields; doubles to device `out`
hipError_t qla_launch_reduce_dh(const void *a, const void *b, int n, double *out, double *scratch,
                                hipStream_t st);

// Ragged batch of F3 v-form jobs in one persistent launch (qla_batch.hip).
enum { QLA_BOP_MATVEC_PV = 0, QLA_BOP_SPPROJ = 1, QLA_BOP_SPRECON = 2 };
struct qla_batch_job {
    void *r;                 // destination field (V, H or D records)
    const void *a;           // ColorMatrix field
    const void *const *b;    // pointer array (V, D or H records)
    int n, op;               // as in the QLA call
    int ms;                  // mu + 5 * (sign != 1)
    int tile0;               // first global tile of this job
    int al;                  // alignment flags of a and r (bit 0: a, bit 2: r, 16-byte)
    int spare;
};
int qla_batch_lds_bytes_per_wave(void);
// zero `counter`, then one persistent launch of `blocks` x (wpb waves)
hipError_t qla_launch_batch(const qla_batch_job *jobs, int njobs, int total_tiles, int wpb, int tpc, int blocks,
                            unsigned int *counter, hipStream_t st);
