// qla_matfn.hip -- batched colour-matrix functions: det, inverse, a^-1 b for
// vector and matrix b, exp, expA, expTA, sqrt, sqrtPH, invsqrt, invsqrtPH,
// log and the Hermitian eigenvalues (lib/perl/build_qla.pl:1002-1066 and
// 1663-1690 generate them for every indexing form; the algorithms are the
// code-templates restated in qla_matfn.hpp).
//
// Compiled once per (precision, nc) unit: QLA_MATFN_UNIT 0 is float nc = 3
// (QLA_F3 and QLA_FN at nc = 3), 1 double nc = 3 (QLA_D3), 2..6 float
// nc = 4..8 and 7, 8 float nc = 1, 2 (QLA_FN).  At nc = 1, 2 every function
// takes the template's own closed form (qla_matfn_small.hpp).  For nc != 3 the templates' expA / expTA take the exp
// branch and sqrtPH / invsqrtPH the sqrt / invsqrt branch, so those codes run
// the general kernels; the Hermitian eigenvalues take the closed form at
// nc = 3 and the template's QR iteration at nc = 4..8.
//
// One thread owns one matrix (for exp, sqrt, invsqrt and log at the larger
// nc, whose per-thread temporaries would not fit in registers, a group of nc
// lanes owns one, a row per lane: qla_matfn_rows.hpp); the grid covers every matrix, so
// the hardware dispatcher balances the data-dependent work of exp, sqrt,
// invsqrt and log (squarings, Newton steps) by starting a new workgroup
// wherever one retires.  (A global work counter claimed per wave costs one
// same-address atomic per 64 matrices, which serialises at the L2 and capped
// every function at about 4.4e9 matrices/s.)
#include <type_traits>
#include "qla_matfn.hpp"
#include "qla_matfn_ch.hpp"
#include "qla_matfn_rows.hpp"
#include "qla_matfn_small.hpp"
#include "qla_launch.hpp"

#ifndef QLA_MATFN_UNIT
#define QLA_MATFN_UNIT 0
#endif

namespace qla {

// function codes (qla_launch.hpp)
enum {
    MF_EXP = 0, MF_SQRT, MF_INVSQRT, MF_DET, MF_INVERSE, MF_INV_V, MF_INV_M, MF_EXPA, MF_EXPTA, MF_SQRTPH,
    MF_INVSQRTPH, MF_LOG, MF_EIGH,
    // Cayley-Hamilton forms with the coefficient step in double (F3 only;
    // no QLA counterpart, qla_matfn_ch.hpp)
    MF_EXPA_CH, MF_SQRTPH_CH, MF_INVSQRTPH_CH
};

// complexes per result record and per second-operand record
template <int FN, int NC> struct fn_shape {
    static constexpr int OUT = FN == MF_DET ? 1 : (FN == MF_INV_V || FN == MF_EIGH) ? NC : NC * NC;
    static constexpr int B = FN == MF_INV_V ? NC : FN == MF_INV_M ? NC * NC : 0;
};

template <class T, int NC, int FN>
__device__ __forceinline__ void evaluate(cpx<T> *out, const mat<T, NC> &a, const cpx<T> *b)
{
    if constexpr (NC <= 2) {
        f_small<T, NC, FN>(out, a, b);
    } else if constexpr (FN == MF_DET) {
        f_det(out[0], a);
    } else if constexpr (FN == MF_INV_V) {
        f_inverse_v(out, a, b);
    } else if constexpr (FN == MF_EIGH) {
        if constexpr (NC == 3) f_eigenvalsH3(out, a);
        else f_eigenvalsH_qr(out, a);
    } else {
        mat<T, NC> o;
        if constexpr (FN == MF_EXP) f_exp(o, a);
        else if constexpr (FN == MF_SQRT) f_sqrt(o, a);
        else if constexpr (FN == MF_INVSQRT) f_invsqrt(o, a);
        else if constexpr (FN == MF_INVERSE) m_eq_inverse_m(o, a);
        else if constexpr (FN == MF_INV_M) {
            mat<T, NC> bm;
            LOOP(k, NC * NC) bm.e[k] = b[k];
            m_eq_m_inverse_m<T, NC, false>(o, a, bm);
        } else if constexpr (FN == MF_EXPA) f_expA3<T, false>(o, a);
        else if constexpr (FN == MF_EXPTA) f_expA3<T, true>(o, a);
        else if constexpr (FN == MF_SQRTPH) f_sqrtPH3<T, false>(o, a);
        else if constexpr (FN == MF_INVSQRTPH) f_sqrtPH3<T, true>(o, a);
        else if constexpr (FN == MF_EXPA_CH) f_expA_ch(o, a);
        else if constexpr (FN == MF_SQRTPH_CH) f_sqrtPH_ch<false>(o, a);
        else if constexpr (FN == MF_INVSQRTPH_CH) f_sqrtPH_ch<true>(o, a);
        else f_log(o, a);
        LOOP(k, NC * NC) out[k] = o.e[k];
    }
}

// Record addresses of step i: the site is index[i] (x-form) or i; a
// per-operand index (ir / ia / ib) replaces it for its own operand; a pointer
// operand is read through its pointer array at that site.
template <class T>
__device__ __forceinline__ const cpx<T> *operand(const void *base, int pointer, long long site, int ncomp)
{
    if (pointer) return ((const cpx<T> *const *)base)[site];
    return (const cpx<T> *)base + (size_t)site * ncomp;
}

constexpr int MF_BLOCK = 256;

// Block-wide copies of `bytes` (a multiple of 8, at most PER * 8 * MF_BLOCK)
// with consecutive threads on consecutive 8-byte words, so each wave
// instruction moves 512 contiguous bytes.  All of a thread's loads are issued
// before the first is waited for.  HBM is touched once per call, so the
// global side carries the non-temporal hint.
template <int PER> __device__ __forceinline__ void block_load8(char *lds, const char *src, int bytes)
{
    const int words = bytes >> 3;
    const unsigned long long *g = (const unsigned long long *)src;
    unsigned long long v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int w = threadIdx.x + k * MF_BLOCK;
        if (w < words) v[k] = __builtin_nontemporal_load(g + w);
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int w = threadIdx.x + k * MF_BLOCK;
        if (w < words) ((unsigned long long *)lds)[w] = v[k];
    }
}
template <int PER> __device__ __forceinline__ void block_store8(char *dst, const char *lds, int bytes)
{
    const int words = bytes >> 3;
    unsigned long long *g = (unsigned long long *)dst;
    unsigned long long v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int w = threadIdx.x + k * MF_BLOCK;
        if (w < words) v[k] = ((const unsigned long long *)lds)[w];
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int w = threadIdx.x + k * MF_BLOCK;
        if (w < words) __builtin_nontemporal_store(v[k], g + w);
    }
}

// Records of at most this many bytes are staged through LDS by the
// contiguous v-form kernel (F3, D3 and FN with nc = 4).
constexpr int MF_STAGE_MAX = 144;

// Staging pays for the functions whose per-matrix work is uniform and short
// (closed forms, cofactor inverse: 1.3-1.5x measured at 2M matrices); the
// iterative ones (exp, sqrt, invsqrt, log) and the solves run faster without
// the block-wide barriers, which make every thread of a block wait for its
// slowest matrix before any result is stored.
template <class T, int NC, int FN> struct stage_of {
    typedef fn_shape<FN, NC> S;
    static constexpr int RA = NC * NC * (int)sizeof(cpx<T>);
    static constexpr int RB = S::B * (int)sizeof(cpx<T>);
    static constexpr int RO = S::OUT * (int)sizeof(cpx<T>);
    static constexpr bool UNIFORM = FN == MF_DET || FN == MF_INVERSE || FN == MF_EXPA || FN == MF_EXPTA ||
                                    FN == MF_SQRTPH || FN == MF_INVSQRTPH || (FN == MF_EIGH && NC == 3) || FN == MF_EXPA_CH ||
                                    FN == MF_SQRTPH_CH || FN == MF_INVSQRTPH_CH;
    static constexpr bool OK = RA <= MF_STAGE_MAX && UNIFORM;
};

// Contiguous v-form (fields, no index): the block's 256 input records are
// copied into LDS with coalesced loads, each thread evaluates its matrix from
// LDS, and the results leave through LDS the same way.  Reading one strided
// record per lane straight from HBM touches ~37 cache lines per wave
// instruction instead of 4.
template <class T, int NC, int FN>
__global__ void __launch_bounds__(MF_BLOCK) k_matfunc_stage(qla_matfunc_args p)
{
    typedef fn_shape<FN, NC> S;
    typedef stage_of<T, NC, FN> G;
    constexpr int SLOT = G::RA > G::RO ? G::RA : G::RO;
    __shared__ __attribute__((aligned(16))) char lds_a[MF_BLOCK * SLOT];
    __shared__ __attribute__((aligned(16))) char lds_b[G::RB > 0 ? MF_BLOCK * G::RB : 16];
    const long long base = (long long)blockIdx.x * MF_BLOCK;
    const long long left = (long long)p.n - base;
    const int cnt = left < MF_BLOCK ? (int)left : MF_BLOCK;
    block_load8<G::RA / 8>(lds_a, (const char *)p.a + base * G::RA, cnt * G::RA);
    if constexpr (G::RB > 0) block_load8<G::RB / 8>(lds_b, (const char *)p.b + base * G::RB, cnt * G::RB);
    __syncthreads();
    const int t = threadIdx.x;
    mat<T, NC> m;
    cpx<T> bv[S::B > 0 ? S::B : 1];
    if (t < cnt) {
        const cpx<T> *src = (const cpx<T> *)(lds_a + t * G::RA);
        LOOP(k, NC * NC) m.e[k] = src[k];
        if constexpr (S::B > 0) {
            const cpx<T> *bsrc = (const cpx<T> *)(lds_b + t * G::RB);
            LOOP(k, S::B) bv[k] = bsrc[k];
        }
    }
    __syncthreads();            // every input is in registers before results overwrite lds_a
    if (t < cnt) {
        cpx<T> out[S::OUT];
        evaluate<T, NC, FN>(out, m, bv);
        cpx<T> *dst = (cpx<T> *)(lds_a + t * G::RO);
        LOOP(k, S::OUT) dst[k] = out[k];
    }
    __syncthreads();
    block_store8<G::RO / 8>((char *)p.r + base * G::RO, lds_a, cnt * G::RO);
}

template <class T, int NC, int FN>
__global__ void __launch_bounds__(MF_BLOCK) k_matfunc(qla_matfunc_args p)
{
    typedef fn_shape<FN, NC> S;
    const long long i = (long long)blockIdx.x * MF_BLOCK + threadIdx.x;
    if (i >= p.n) return;
    const long long site = p.index ? (long long)p.index[i] : i;
    const long long rs = p.ir ? (long long)p.ir[i] : site;
    const long long as = p.ia ? (long long)p.ia[i] : site;
    mat<T, NC> m;
    const cpx<T> *src = operand<T>(p.a, p.pa, as, NC * NC);
    LOOP(k, NC * NC) m.e[k] = src[k];
    cpx<T> bv[S::B > 0 ? S::B : 1];
    if constexpr (S::B > 0) {
        const long long bs = p.ib ? (long long)p.ib[i] : site;
        const cpx<T> *bsrc = operand<T>(p.b, p.pb, bs, S::B);
        LOOP(k, S::B) bv[k] = bsrc[k];
    }
    cpx<T> out[S::OUT];
    evaluate<T, NC, FN>(out, m, bv);
    cpx<T> *dst = (cpx<T> *)p.r + (size_t)rs * S::OUT;
    LOOP(k, S::OUT) dst[k] = out[k];
}

// exp / sqrt / invsqrt at nc = 6..8 and log at nc = 7, 8: one matrix per
// group of nc lanes, one row per lane (qla_matfn_rows.hpp); every indexing
// form.  Measured on 262,144 matrices against one thread per matrix, the row
// form runs exp 1.30 / 1.41 / 1.90x, sqrt 1.01 / 1.10 / 1.31x, invsqrt
// 1.08 / 1.17 / 1.32x and log 0.92 / 1.08 / 1.20x at nc = 6 / 7 / 8; at
// nc = 5, where one thread's temporaries still fit in its registers, it runs
// 0.74-1.01x.
template <int NC, int FN> struct by_rows {
    static constexpr bool v = (NC >= 6 && (FN == MF_EXP || FN == MF_SQRT || FN == MF_INVSQRT)) ||
                              (NC >= 7 && FN == MF_LOG);
};

// groups per wavefront and per block of the row kernel
template <int NC> struct rows_shape {
    static constexpr int GPW = WAVE / NC;
    static constexpr int GPB = GPW * (MF_BLOCK / WAVE);
};

template <class T, int NC, int FN>
__global__ void __launch_bounds__(MF_BLOCK) k_matfunc_rows(qla_matfunc_args p)
{
    constexpr int GPW = rows_shape<NC>::GPW;
    __shared__ rgrp<T, NC> sh[rows_shape<NC>::GPB];
    const int lane = (int)(threadIdx.x % WAVE);
    const int gw = lane / NC;
    if (gw >= GPW) return;          // the wave's last 64 mod nc lanes
    const int g = (int)(threadIdx.x / WAVE) * GPW + gw;
    const long long i = (long long)blockIdx.x * rows_shape<NC>::GPB + g;
    if (i >= p.n) return;           // whole groups
    const int row = lane - gw * NC;
    const rctx<T, NC> c = { &sh[g], row };
    const long long site = p.index ? (long long)p.index[i] : i;
    const long long rs = p.ir ? (long long)p.ir[i] : site;
    const long long as = p.ia ? (long long)p.ia[i] : site;
    rrow<T, NC> a, r;
    const cpx<T> *src = operand<T>(p.a, p.pa, as, NC * NC) + row * NC;
    LOOP(j, NC) a.e[j] = src[j];
    if constexpr (FN == MF_EXP) r_exp(r, a, c);
    else if constexpr (FN == MF_SQRT) r_sqrt(r, a, c);
    else if constexpr (FN == MF_INVSQRT) r_invsqrt(r, a, c);
    else r_log(r, a, c);
    cpx<T> *dst = (cpx<T> *)p.r + (size_t)rs * (NC * NC) + row * NC;
    LOOP(j, NC) dst[j] = r.e[j];
}

template <class T, int NC, int FN>
static hipError_t launch_t(const qla_matfunc_args &p, hipStream_t st)
{
    if constexpr (by_rows<NC, FN>::v) {
        constexpr int GPB = rows_shape<NC>::GPB;
        const long long gblocks = ((long long)p.n + GPB - 1) / GPB;
        hipLaunchKernelGGL((k_matfunc_rows<T, NC, FN>), dim3((unsigned)gblocks), dim3(MF_BLOCK), 0, st, p);
        return hipGetLastError();
    } else {
        const long long blocks = ((long long)p.n + MF_BLOCK - 1) / MF_BLOCK;
        const bool contiguous = !p.index && !p.ir && !p.ia && !p.ib && !p.pa && !p.pb;
        if constexpr (stage_of<T, NC, FN>::OK) {
            if (contiguous) {
                hipLaunchKernelGGL((k_matfunc_stage<T, NC, FN>), dim3((unsigned)blocks), dim3(MF_BLOCK), 0, st, p);
                return hipGetLastError();
            }
        }
        hipLaunchKernelGGL((k_matfunc<T, NC, FN>), dim3((unsigned)blocks), dim3(MF_BLOCK), 0, st, p);
        return hipGetLastError();
    }
}

template <class T, int NC>
static hipError_t launch_nc(int fn, const qla_matfunc_args &p, hipStream_t st)
{
    switch (fn) {
    case MF_EXP: return launch_t<T, NC, MF_EXP>(p, st);
    case MF_SQRT: return launch_t<T, NC, MF_SQRT>(p, st);
    case MF_INVSQRT: return launch_t<T, NC, MF_INVSQRT>(p, st);
    case MF_DET: return launch_t<T, NC, MF_DET>(p, st);
    case MF_INVERSE: return launch_t<T, NC, MF_INVERSE>(p, st);
    case MF_INV_V: return launch_t<T, NC, MF_INV_V>(p, st);
    case MF_INV_M: return launch_t<T, NC, MF_INV_M>(p, st);
    case MF_LOG: return launch_t<T, NC, MF_LOG>(p, st);
    default: break;
    }
    if constexpr (NC == 3) {
        switch (fn) {
        case MF_EXPA: return launch_t<T, 3, MF_EXPA>(p, st);
        case MF_EXPTA: return launch_t<T, 3, MF_EXPTA>(p, st);
        case MF_SQRTPH: return launch_t<T, 3, MF_SQRTPH>(p, st);
        case MF_INVSQRTPH: return launch_t<T, 3, MF_INVSQRTPH>(p, st);
        case MF_EIGH: return launch_t<T, 3, MF_EIGH>(p, st);
        default: break;
        }
        if constexpr (std::is_same<T, float>::value) {
            switch (fn) {
            case MF_EXPA_CH: return launch_t<float, 3, MF_EXPA_CH>(p, st);
            case MF_SQRTPH_CH: return launch_t<float, 3, MF_SQRTPH_CH>(p, st);
            case MF_INVSQRTPH_CH: return launch_t<float, 3, MF_INVSQRTPH_CH>(p, st);
            default: break;
            }
        }
        return hipErrorInvalidValue;
    } else if constexpr (NC <= 2) {
        // every function has its own closed form at nc = 1, 2 (qla_matfn_small.hpp)
        switch (fn) {
        case MF_EXPA: return launch_t<T, NC, MF_EXPA>(p, st);
        case MF_EXPTA: return launch_t<T, NC, MF_EXPTA>(p, st);
        case MF_SQRTPH: return launch_t<T, NC, MF_SQRTPH>(p, st);
        case MF_INVSQRTPH: return launch_t<T, NC, MF_INVSQRTPH>(p, st);
        case MF_EIGH: return launch_t<T, NC, MF_EIGH>(p, st);
        default: return hipErrorInvalidValue;
        }
    } else {
        // the templates' general branches (M_eq_expA_M.c:357-420 and
        // M_eq_sqrtPH_M.c:270-310 are M_eq_exp_M / M_eq_sqrt_M's)
        switch (fn) {
        case MF_EXPA:
        case MF_EXPTA: return launch_t<T, NC, MF_EXP>(p, st);
        case MF_SQRTPH: return launch_t<T, NC, MF_SQRT>(p, st);
        case MF_INVSQRTPH: return launch_t<T, NC, MF_INVSQRT>(p, st);
        case MF_EIGH: return launch_t<T, NC, MF_EIGH>(p, st);        // the template's QR iteration
        default: return hipErrorInvalidValue;
        }
    }
}

} // namespace qla

using namespace qla;

#if QLA_MATFN_UNIT == 0
hipError_t qla_matfunc_launch_u0(int fn, const qla_matfunc_args &p, hipStream_t st)
{ return launch_nc<float, 3>(fn, p, st); }
#elif QLA_MATFN_UNIT == 1
hipError_t qla_matfunc_launch_u1(int fn, const qla_matfunc_args &p, hipStream_t st)
{ return launch_nc<double, 3>(fn, p, st); }
#elif QLA_MATFN_UNIT == 2
hipError_t qla_matfunc_launch_u2(int fn, const qla_matfunc_args &p, hipStream_t st)
{ return launch_nc<float, 4>(fn, p, st); }
#elif QLA_MATFN_UNIT == 3
hipError_t qla_matfunc_launch_u3(int fn, const qla_matfunc_args &p, hipStream_t st)
{ return launch_nc<float, 5>(fn, p, st); }
#elif QLA_MATFN_UNIT == 4
hipError_t qla_matfunc_launch_u4(int fn, const qla_matfunc_args &p, hipStream_t st)
{ return launch_nc<float, 6>(fn, p, st); }
#elif QLA_MATFN_UNIT == 5
hipError_t qla_matfunc_launch_u5(int fn, const qla_matfunc_args &p, hipStream_t st)
{ return launch_nc<float, 7>(fn, p, st); }
#elif QLA_MATFN_UNIT == 6
hipError_t qla_matfunc_launch_u6(int fn, const qla_matfunc_args &p, hipStream_t st)
{ return launch_nc<float, 8>(fn, p, st); }
#elif QLA_MATFN_UNIT == 7
hipError_t qla_matfunc_launch_u7(int fn, const qla_matfunc_args &p, hipStream_t st)
{ return launch_nc<float, 1>(fn, p, st); }
#elif QLA_MATFN_UNIT == 8
hipError_t qla_matfunc_launch_u8(int fn, const qla_matfunc_args &p, hipStream_t st)
{ return launch_nc<float, 2>(fn, p, st); }
#endif
