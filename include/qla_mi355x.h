/*
 * qla_mi355x.h -- drop-in C ABI for the QLA site-loop hot path on MI355X (gfx950).
 *
 * Every QLA_* prototype below has exactly the name, argument order and
 * argument meaning of the symbol that the reference QLA 1.9.0 generator emits
 * (lib/perl/prototype.pl:590-688 builds the argument list: dest, src1, src2,
 * src2 extras `int mu, int sign`, gang `int *index`, `int n`; the leading
 * `int nc` appears only for the N-colour library, prototype.pl:594-600).
 * Type layouts are the reference's site-major AoS records
 * (include/qla_types.h:29-33): ColorVector V[3], HalfFermion V[3][2],
 * DiracFermion V[3][4], ColorMatrix V[3][3], complex = {re, im} with float
 * complex aligned to 8 B and double complex to 16 B (include/qla_complex.h:27-35).
 *
 * Device semantics (the one deliberate difference from the CPU library):
 *   - every field pointer, every pointer array, every pointee of a pointer
 *     array and every index array is a DEVICE address;
 *   - the result pointer of a global reduction (r_veq_norm2_*, c_veq_*_dot_*)
 *     may be a host or a device address;
 *   - by default each call returns with its results complete (QLA's
 *     synchronous contract); QLA_MI355X_set_sync(0) turns a host thread's
 *     calls into stream-ordered launches on the stream set by
 *     QLA_MI355X_set_stream (reductions into host memory still synchronise).
 * Behaviour kept from the reference: n <= 0 is a no-op (a reduction then
 * writes 0); mu outside 0..4 is a silent no-op (expressions_gamma.pl:349-381
 * emits a switch without default); any sign != 1 means -1; mu = 4 is the
 * chiral projection.  Arguments are restrict: destinations must not overlap
 * sources.  No 16 B alignment is required of any argument.
 */
#ifndef QLA_MI355X_H
#define QLA_MI355X_H

#ifdef __cplusplus
extern "C" {
#define QLA_MI355X_RESTRICT __restrict__
#else
#define QLA_MI355X_RESTRICT restrict
#endif

/* ------------------------------------------------------------------------ */
/* Types (layout-identical to include/qla_complex.h:27-35, qla_types.h:29-50) */
/* ------------------------------------------------------------------------ */
#ifndef _QLA_TYPES_H
typedef int QLA_Int;
typedef float QLA_F_Real;
typedef double QLA_D_Real;
typedef struct { QLA_F_Real real; QLA_F_Real imag; } __attribute__((aligned(8))) QLA_F_Complex;
typedef struct { QLA_D_Real real; QLA_D_Real imag; } __attribute__((aligned(16))) QLA_D_Complex;
typedef QLA_F_Complex QLA_F3_ColorVector[3];
typedef QLA_F_Complex QLA_F3_HalfFermion[3][2];
typedef QLA_F_Complex QLA_F3_DiracFermion[3][4];
typedef QLA_F_Complex QLA_F3_ColorMatrix[3][3];
typedef QLA_F_Complex QLA_F3_DiracPropagator[3][4][3][4];
typedef QLA_D_Complex QLA_D3_ColorVector[3];
typedef QLA_D_Complex QLA_D3_HalfFermion[3][2];
typedef QLA_D_Complex QLA_D3_DiracFermion[3][4];
typedef QLA_D_Complex QLA_D3_ColorMatrix[3][3];
typedef QLA_D_Complex QLA_D3_DiracPropagator[3][4][3][4];
#endif

/* ------------------------------------------------------------------------ */
/* The Wilson-hop building blocks: every eqop (eq, peq,                    */
/* meq, eqm), M or Ma, field or pointer operands, v- and x-forms, for       */
/* V = M*V, H = M*spproj(D), D = sprecon(M*H), D = sprecon(M*spproj(D)) and */
/* the multi-source V = sum_d M_d*V_d.  Reference: generated from           */
/* lib/perl/expressions_tensor.pl:544-630 (matvec accumulation x=0 or x=r,  */
/* then x +=/-= a[ic][kc]*b[kc], kc = 0,1,2, include/qla_complex.h:330-341),*/
/* build_qla.pl:1162-1250 (spin forms), :1693-1725 (multi-source).  The SSE */
/* variants lib/optimized/sse/src/QLA_F3_V_vpeq_M_times_pV.c,               */
/* ..._D_vpeq_sprecon_M_times_pH.c, ..._D_vpeq_spproj_M_times_pD.c (and    */
/* V_veq_M_times_pV, V_veq_Ma_times_V) re-associate; this library keeps the */
/* generic C order unless QLA_MI355X_set_arith_order selects theirs.        */
/* ------------------------------------------------------------------------ */
#include "qla_mi355x_sites.h"

/* ------------------------------------------------------------------------ */
/* Global reductions.  Reference: make_code_unary.pl:318-366               */
/* (norm2: per component the square is formed in the operand precision and */
/* added to a sum promoted per datatypes.pl:355-359, F -> D, D -> D) and    */
/* make_code_binary.pl:110-170 (dot: sum += conj(a)*b with operands         */
/* promoted first, QLA_c_peq_ca_times_c, qla_complex.h:386-389).            */
/* ------------------------------------------------------------------------ */
/* Every type (V, H, D, M, P), F3 / D3 / DF3, v- and x-forms, field or     */
/* pointer operands, plus the real part of the inner product and the DF3 /  */
/* FD3 precision conversions (build_qla.pl:530-585, 2089-2360).             */
#include "qla_mi355x_fields.h"

/* ------------------------------------------------------------------------ */
/* Colour-matrix functions: det, inverse, a^-1 b for vector / matrix b,     */
/* exp, expA, expTA, sqrt, sqrtPH, invsqrt, invsqrtPH, log and the          */
/* Hermitian eigenvalues, QLA_F3 / QLA_D3 / QLA_FN, every indexing form.    */
/* Reference: build_qla.pl:1002-1066 and 1663-1690 around the algorithms of */
/* the lib/perl/code-templates sources.  QLA_FN_* take the colour count    */
/* first (prototype.pl:594-600); this library supports nc = 1..8 (other nc: */
/* no-op and an error is recorded).                                         */
/* ------------------------------------------------------------------------ */
#include "qla_mi355x_matfuncs.h"

/* ------------------------------------------------------------------------ */
/* Field linear algebra: copies and increments, Hermitian conjugate,        */
/* transpose, complex conjugate, times i, times a real or complex constant, */
/* sums and differences, a*b +- c with a real or complex constant a; types  */
/* V, H, D, M, P, QLA_F3 / QLA_D3, every eqop the reference generates and   */
/* every indexing form.  Reference: build_qla.pl:356-494, 1256-1354,        */
/* 1421-1456, 1837-1935.  The constant is read from host or device memory.  */
/* ------------------------------------------------------------------------ */
#include "qla_mi355x_vecops.h"

/* ------------------------------------------------------------------------ */
/* Compressed-precision extension (no counterpart in the reference, which    */
/* stores FP32/FP64 records only).  Record formats, 4-byte aligned:          */
/*   link 24 B: rows 0,1 of an SU(3) matrix as 12 bfloat16 (fmt 0) or IEEE    */
/*              binary16 (fmt 1) reals; row 2 = conj(row0 x row1) in FP32.     */
/*   Dh   52 B: DiracFermion as 24 binary16 then a float scale s (x = v*s).   */
/*   Hh   28 B: HalfFermion as 12 binary16 then a float scale.                 */
/* Encoding: s = max |x| of the record, v = binary16(x/s) rounded to nearest  */
/* even.  Arithmetic after decoding is FP32 in the order of the plain forms.  */
/* Gathers take pointer arrays to Dh / Hh records, like the QLA pD / pH args. */
/* ------------------------------------------------------------------------ */
enum { QLA_MI355X_LINK_BF16 = 0, QLA_MI355X_LINK_F16 = 1 };
void QLA_MI355X_F3_M_pack12(void *r, int fmt, QLA_F3_ColorMatrix *a, int n);
void QLA_MI355X_F3_M_unpack12(QLA_F3_ColorMatrix *r, int fmt, void *a, int n);
void QLA_MI355X_F3_D_packh(void *r, QLA_F3_DiracFermion *a, int n);
void QLA_MI355X_F3_D_unpackh(QLA_F3_DiracFermion *r, void *a, int n);
void QLA_MI355X_F3_H_packh(void *r, QLA_F3_HalfFermion *a, int n);
void QLA_MI355X_F3_H_unpackh(QLA_F3_HalfFermion *r, void *a, int n);
/* Hh = spproj(M12 * pDh) and Dh += sprecon(M12 * pHh); mu/sign as in QLA. */
void QLA_MI355X_F3_Hh_veq_spproj_M12_times_pDh(void *r, void *a, int fmt, void **b, int mu, int sign, int n);
void QLA_MI355X_F3_Dh_vpeq_sprecon_M12_times_pHh(void *r, void *a, int fmt, void **b, int mu, int sign, int n);
/* norm2 and dot of decoded Dh fields with the QLA_DF3 semantics (float
 * component norms / double-promoted products, double sum). */
void QLA_MI355X_DF3_r_veq_norm2_Dh(QLA_D_Real *r, void *a, int n);
void QLA_MI355X_DF3_c_veq_Dh_dot_Dh(QLA_D_Complex *r, void *a, void *b, int n);

/* ------------------------------------------------------------------------ */
/* Cayley-Hamilton matrix functions with an FP64 coefficient step (no        */
/* counterpart in the reference; QLA_F3_M_veq_{expA,sqrtPH,invsqrtPH}_M are  */
/* the closest, and evaluate their eigenvalues in single precision).  F3    */
/* records in and out: f(X) = f0 + f1 X + f2 X^2, with X's invariants, its  */
/* eigenvalues, the coefficients, X^2 and the sum in double and each result */
/* component rounded to float once.                                          */
/*   expA:      exp of the anti-Hermitian part of a (a itself when a is      */
/*              anti-Hermitian, as for QLA's expA); exp(0) = 1.               */
/*   sqrtPH:    square root of the Hermitian part of a, which must be        */
/*              positive semi-definite (eigenvalues below 0 count as 0).      */
/*   invsqrtPH: its inverse; a zero eigenvalue gives +inf on the diagonal.   */
/* Results are not bit-identical to the QLA symbols: they agree with the     */
/* exact function to 1e-7 relative to the result's norm (one float rounding */
/* per component).  The x-forms act on the records index[0..n) as QLA's     */
/* x-forms do.                                                               */
/* ------------------------------------------------------------------------ */
void QLA_MI355X_F3_M_veq_expA_M_ch(QLA_F3_ColorMatrix *r, QLA_F3_ColorMatrix *a, int n);
void QLA_MI355X_F3_M_xeq_expA_M_ch(QLA_F3_ColorMatrix *r, QLA_F3_ColorMatrix *a, int *index, int n);
void QLA_MI355X_F3_M_veq_sqrtPH_M_ch(QLA_F3_ColorMatrix *r, QLA_F3_ColorMatrix *a, int n);
void QLA_MI355X_F3_M_xeq_sqrtPH_M_ch(QLA_F3_ColorMatrix *r, QLA_F3_ColorMatrix *a, int *index, int n);
void QLA_MI355X_F3_M_veq_invsqrtPH_M_ch(QLA_F3_ColorMatrix *r, QLA_F3_ColorMatrix *a, int n);
void QLA_MI355X_F3_M_xeq_invsqrtPH_M_ch(QLA_F3_ColorMatrix *r, QLA_F3_ColorMatrix *a, int *index, int n);

/* ------------------------------------------------------------------------ */
/* Allocation-free reduction (no counterpart in the reference).  For callers */
/* that cannot allocate device memory: code under HIP graph capture, tight   */
/* serving loops, harnesses that forbid hipMalloc.  Same result, bit for     */
/* bit, as QLA_DF3_r_veq_norm2_D (same kernels), but its scratch is a       */
/* static 32 KiB device array (one copy per GPU, held until the process      */
/* ends) instead of the per-stream buffer the other reductions allocate on  */
/* first use, and it calls no hipMalloc / hipFree.  Size limit: none below  */
/* INT_MAX sites (the first stage runs at most 1024 blocks whatever n is,   */
/* and the array holds their partials; a launch needing more would be       */
/* refused with hipErrorInvalidValue).  r may be host or device memory (a   */
/* host result waits on the host, so only a device r can be captured).     */
/* Every call on one GPU shares that GPU's array: calls that overlap on one */
/* GPU (two streams, or two threads) race; order them.                      */
/* ------------------------------------------------------------------------ */
void QLA_MI355X_DF3_r_veq_norm2_D_noalloc(QLA_D_Real *r, QLA_F3_DiracFermion *a, int n);

/* ------------------------------------------------------------------------ */
/* Ragged batch extension (no counterpart in the reference, where each call  */
/* is its own site loop).  A plan holds a table of independent QLA_F3 v-form */
/* calls; running it executes all of them in one launch that cuts every    */
/* job into 64-site tiles, so jobs of very different sizes share the chip.  */
/* Each job's result is bit-identical to its plain call.  The plan can be   */
/* captured into a hipGraph.                                                */
/* ------------------------------------------------------------------------ */
enum {
    QLA_MI355X_OP_V_VPEQ_M_TIMES_PV = 0,      /* QLA_F3_V_vpeq_M_times_pV         */
    QLA_MI355X_OP_H_VEQ_SPPROJ_M_TIMES_PD = 1,  /* QLA_F3_H_veq_spproj_M_times_pD   */
    QLA_MI355X_OP_D_VPEQ_SPRECON_M_TIMES_PH = 2 /* QLA_F3_D_vpeq_sprecon_M_times_pH */
};
typedef struct {
    int op;        /* QLA_MI355X_OP_* */
    int mu, sign;  /* spin ops only; as in the QLA call */
    int n;         /* sites */
    void *r;       /* destination field (device) */
    void *a;       /* ColorMatrix field (device) */
    void **b;      /* device pointer array of the gathered records */
} QLA_MI355X_F3_job;
typedef struct QLA_MI355X_batch QLA_MI355X_batch;
/* Copies the job table to the device (the operands are referenced, not
 * copied); returns NULL and records an error for an unknown op. */
QLA_MI355X_batch *QLA_MI355X_F3_batch_create(const QLA_MI355X_F3_job *jobs, int njobs);
/* Runs every job (one launch on the thread's stream). */
void QLA_MI355X_batch_run(QLA_MI355X_batch *p);
/* Captures one run into a hipGraph (0 on success); launch it with
 * QLA_MI355X_batch_launch_graph.  Re-capture after a shape change. */
int QLA_MI355X_batch_capture(QLA_MI355X_batch *p);
void QLA_MI355X_batch_launch_graph(QLA_MI355X_batch *p);
/* Launch shape: waves per block (1, 2, 4), tiles of 64 sites per claim
 * (>= 1) and blocks per CU of the persistent grid whose waves claim tiles
 * from a counter (>= 1), or 0 for the static schedule, whose grid covers
 * every tile with no counter.  Results do not depend on the shape.
 * QLA_MI355X_batch_autotune times every shape `reps` times on the plan's own
 * operands, keeps the fastest, caches it for plans of the same size class
 * and returns its time in ms.  The jobs run (1 + reps) x 60 times:
 * accumulating destinations (vpeq) change. */
float QLA_MI355X_batch_autotune(QLA_MI355X_batch *p, int reps);
void QLA_MI355X_batch_get_shape(const QLA_MI355X_batch *p, int *waves_per_block, int *tiles_per_claim,
                                int *blocks_per_cu);
int QLA_MI355X_batch_set_shape(QLA_MI355X_batch *p, int waves_per_block, int tiles_per_claim,
                               int blocks_per_cu);
void QLA_MI355X_batch_destroy(QLA_MI355X_batch *p);

/* ------------------------------------------------------------------------ */
/* Library control (no counterpart in the reference, which has no device   */
/* and no stream; these only select where and how the calls above run).    */
/* ------------------------------------------------------------------------ */
/* Per host thread: the HIP stream (hipStream_t) the calls launch on; NULL
 * selects the default stream. */
void QLA_MI355X_set_stream(void *stream);
void *QLA_MI355X_get_stream(void);
/* Per host thread: the association order of the floating-point sums.
 * QLA_MI355X_ORDER_GENERIC (default): every symbol rounds exactly like the
 * reference's generated C (a default build), bit for bit.
 * QLA_MI355X_ORDER_SSE3: the symbols that a reference built --enable-sse3
 * replaces with re-associated SSE code round exactly like that build
 * instead: the five QLA_F3 ones -- V_vpeq_M_times_pV, V_veq_M_times_pV,
 * V_veq_Ma_times_V, D_vpeq_sprecon_M_times_pH and D_vpeq_spproj_M_times_pD
 * (lib/optimized/sse/src/) -- including that build's choice of order by the
 * alignment of r, a and b[0] and, for the V symbols, by the position of a
 * site in its 1- and 2-site loops; and the four QLA_D3 ones that
 * --enable-sse3 takes from its SSE2 versions -- V_veq_M_times_V,
 * V_veq_M_times_pV, V_vpeq_M_times_V and V_vpeq_M_times_pV
 * (lib/optimized/sse2/src/).  Every other symbol, and the batches, keep the
 * generic order.  Both orders are within 3 ulp (of the sum of the term
 * magnitudes) of the exact result. */
enum { QLA_MI355X_ORDER_GENERIC = 0, QLA_MI355X_ORDER_SSE3 = 1 };
void QLA_MI355X_set_arith_order(int order);
int QLA_MI355X_get_arith_order(void);
/* Per host thread: 1 (default) = each call returns with results complete,
 * 0 = calls are stream-ordered and return after the launch. */
void QLA_MI355X_set_sync(int on);
/* Wait for the calling thread's stream. */
void QLA_MI355X_synchronize(void);
/* 0 if no HIP error has been seen on this thread since the last call, else
 * the hipError_t value; QLA_MI355X_last_error_string describes it. The QLA
 * entry points themselves return void, like the reference. */
int QLA_MI355X_last_error(void);
const char *QLA_MI355X_last_error_string(void);
/* Device scratch of the global reductions and sums: one small buffer (32 KiB
 * of partials, or 4.5 MiB for the field sums) per (host thread, device,
 * stream) that calls them, allocated and zeroed on that triple's first
 * reduction (which therefore must not be inside a stream capture: the call
 * then records hipErrorStreamCaptureUnsupported and does nothing; later
 * reductions on that stream can be captured when their result pointer is
 * device memory -- a host result waits on the host and is refused under
 * capture the same way).  Reductions on different streams never share a buffer.  A
 * thread's buffers are released when it exits; QLA_MI355X_finalize releases
 * every buffer of every thread (no library call may be in flight), and
 * QLA_MI355X_scratch_buffers counts the live ones. */
void QLA_MI355X_finalize(void);
int QLA_MI355X_scratch_buffers(void);
/* Launch-shape knob of the site kernels (spproj, sprecon,
 * matvec): tiles of 64 sites per wavefront in the grid (default 1).  Set by
 * the autotuner; results do not depend on it. */
void QLA_MI355X_set_tiles_per_wave(int tpw);
/* Library version string. */
const char *QLA_MI355X_version(void);

#ifdef __cplusplus
}
#endif

/* QLA's generic names (QLA_V_veq_M_times_V, QLA_F_..., QLA_3_..., generic
 * types), selected by QLA_Precision / QLA_Colors as in the reference. */
#include "qla_mi355x_generic.h"

#endif /* QLA_MI355X_H */
