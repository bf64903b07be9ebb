// qla_launch.hpp -- internal launch interface between the C ABI layer
// (qla_capi.cpp) and the kernel translation units.  prec: 0 = float, 1 = double.
#pragma once
#include <hip/hip_runtime.h>

// V = M*V (eqop eq, b contiguous) or V += M*(*pV) (eqop peq, b pointer array);
// index != NULL selects the gang-indexed x-form.
hipError_t qla_launch_matvec(int prec, int eqop_peq, int pb, void *r, const void *a, const void *b,
                             const int *index, int n, hipStream_t st);
// launch-shape knob of the persistent site kernels
void qla_set_tiles_per_wave(int tpw);
// H = spproj(M * pD)
hipError_t qla_launch_spproj(int prec, void *r, const void *a, const void *b, int mu, int sign,
                             const int *index, int n, hipStream_t st);
// D += sprecon(M * pH)
hipError_t qla_launch_sprecon(int prec, void *r, const void *a, const void *b, int mu, int sign,
                              const int *index, int n, hipStream_t st);

// Global reductions over m complex elements.  kind: 0 = norm2 (b unused),
// 1 = dot conj(a).b.  The result is written to the DEVICE address `out` as
// out_kind 0 double, 1 float, 2 double complex, 3 float complex.  `scratch`
// must hold qla_reduce_scratch_doubles() doubles.
int qla_reduce_scratch_doubles(void);
hipError_t qla_launch_reduce2(int prec, int kind, const void *a, const void *b, long long m, void *out,
                              int out_kind, double *scratch, hipStream_t st);

// Matrix functions on n nc x nc matrices.  fn: 0 exp, 1 sqrt, 2 invsqrt.
// prec 0: nc = 3..8; prec 1: nc = 3.  `counter` is one device unsigned int
// the launch uses as its work counter (reset on the stream first).
hipError_t qla_launch_matfn(int prec, int fn, int nc, void *r, const void *a, int n, unsigned int *counter,
                            hipStream_t st);

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
    int n, op, mu, sign;     // as in the QLA call
    int tile0;               // first global tile of this job
    int al;                  // alignment flags of a and r (bit 0: a, bit 2: r, 16-byte)
};
int qla_batch_lds_bytes_per_wave(void);
// zero `counter`, then one persistent launch of `blocks` x (wpb waves)
hipError_t qla_launch_batch(const qla_batch_job *jobs, int njobs, int total_tiles, int wpb, int tpc, int blocks,
                            unsigned int *counter, hipStream_t st);
