// qla_vecops.hip -- field linear algebra around the Wilson hop: copies and
// increments, Hermitian conjugate, transpose, complex conjugate, times i,
// times a real or complex constant, sums and differences and a*b +- c with a
// real or complex constant a (build_qla.pl:356-494, 1256-1354, 1421-1456,
// 1837-1935), for records of V, H, D, M and P, and the products with
// propagators (build_qla.pl:1587-1662; k_vec_pmul), component access and the
// spin trace (build_qla.pl:767-932, 1066-1086; k_vec_comp).
//
// Every operation is element-wise on the complex components, so the kernels
// stream components, not records:
//
//  * k_vec_flat: v-forms over fields (no index, no pointer array, no
//    transpose).  The three arrays are one run of n*ncomp complexes; each
//    work item moves 16 bytes per operand (two float complexes or one double
//    complex) with non-temporal loads and stores.  The record type does not
//    enter, so one kernel serves V, H, D, M and P.
//  * k_vec_gather: x-forms, pointer-array sources, the per-operand index
//    forms and the transposes.  One work item per destination complex
//    (record i = g / NC, component g % NC); each resolves its own record
//    addresses (the lanes of one record read the same index and pointer
//    entries, which the memory pipeline merges), so every record is still
//    read and written as a contiguous run.
//
// Both grids are capped at 2^20 blocks of 256 and stride over their work:
// an HSA dispatch counts its work items in 32 bits, and a P field of 2^25
// records already holds more than 2^32 complexes.
//
// Arithmetic.  The reference macros (include/qla_complex.h) are restated per
// element with the conjugated-source forms (_ca) computed as the plain form
// on the conjugated source, which is bit-identical: negating an operand is
// exact and IEEE multiplication and addition are sign-symmetric and
// commutative, so a*(-b) = -(a*b), x - (-y) = x + y and (-p) + q = q - p hold
// bit for bit, signed zeros included.  The eqm form of the complex constant
// (QLA_c_eqm_c_times_c, include/qla_complex.h:334-337) is not the negation
// of its eq form and is kept as written.
#include <string.h>
#include "qla_dev.hpp"
#include "qla_launch.hpp"

namespace qla {
namespace {

constexpr int VBLOCK = 256;

enum { VOP_COPY = 0, VOP_I, VOP_RSCALE, VOP_CSCALE, VOP_PLUS, VOP_MINUS, VOP_R_PLUS, VOP_R_MINUS, VOP_C_PLUS,
       VOP_C_MINUS, VOP_RFIELD, VOP_CFIELD };

// times a per-site real / complex factor field (build_qla.pl:1517-1586): the
// element arithmetic of the real / complex constant with the factor of the step
template <int OP> constexpr bool is_field() { return OP == VOP_RFIELD || OP == VOP_CFIELD; }
template <int OP> constexpr int elem_op() { return OP == VOP_RFIELD ? VOP_RSCALE : OP == VOP_CFIELD ? VOP_CSCALE : OP; }
template <int OP> constexpr bool has_y() { return OP >= VOP_PLUS && !is_field<OP>(); }
template <int OP> constexpr bool has_scalar()
{
    return OP == VOP_RSCALE || OP == VOP_CSCALE || (OP >= VOP_R_PLUS && OP <= VOP_C_MINUS);
}
template <int OP> constexpr bool complex_scalar() { return OP == VOP_CSCALE || OP == VOP_C_PLUS || OP == VOP_C_MINUS; }
template <int EQOP> constexpr bool reads_r() { return EQOP == 1 || EQOP == 2; }

template <class T, int V> struct vtype;
template <> struct vtype<float, 1> { typedef float v __attribute__((ext_vector_type(2))); };
template <> struct vtype<float, 2> { typedef float v __attribute__((ext_vector_type(4))); };
template <> struct vtype<double, 1> { typedef double v __attribute__((ext_vector_type(2))); };

// the constant: from device memory (sp) or the value passed by the host
template <class T, int OP>
__device__ __forceinline__ void load_scalar(const qla_vec_args &p, T &sr, T &si)
{
    sr = T(0);
    si = T(0);
    if constexpr (has_scalar<OP>()) {
        if (p.sp) {
            const T *s = (const T *)p.sp;
            sr = s[0];
            if constexpr (complex_scalar<OP>()) si = s[1];
        } else {
            sr = (T)p.s_re;
            si = (T)p.s_im;
        }
    }
}

// r eqop op(x, y) on one complex; x already conjugated for the _ca forms
template <int OP, int EQOP, class T>
__device__ __forceinline__ void vop(T &rr, T &ri, T xr, T xi, T yr, T yi, T sr, T si)
{
    T vr, vi;
    if constexpr (OP == VOP_PLUS) {                    // QLA_c_eq_c_plus_c (qla_complex.h:287)
        rr = xr + yr;
        ri = xi + yi;
        return;
    } else if constexpr (OP == VOP_MINUS) {            // QLA_c_eq_c_minus_c (:305)
        rr = xr - yr;
        ri = xi - yi;
        return;
    } else if constexpr (OP == VOP_R_PLUS) {           // QLA_c_eq_r_times_c_plus_c (:445)
        rr = sr * xr + yr;
        ri = sr * xi + yi;
        return;
    } else if constexpr (OP == VOP_R_MINUS) {          // QLA_c_eq_r_times_c_minus_c (:449)
        rr = sr * xr - yr;
        ri = sr * xi - yi;
        return;
    } else if constexpr (OP == VOP_C_PLUS) {           // QLA_c_eq_c_times_c_plus_c (:453)
        rr = sr * xr - si * xi + yr;
        ri = sr * xi + si * xr + yi;
        return;
    } else if constexpr (OP == VOP_C_MINUS) {          // QLA_c_eq_c_times_c_minus_c (:459)
        rr = sr * xr - si * xi - yr;
        ri = sr * xi + si * xr - yi;
        return;
    } else if constexpr (OP == VOP_COPY) {             // QLA_c_<eqop>_c (:147-154)
        vr = xr;
        vi = xi;
    } else if constexpr (OP == VOP_I) {                // QLA_c_<eqop>_ic (:165-172)
        vr = -xi;
        vi = xr;
    } else if constexpr (OP == VOP_RSCALE) {           // QLA_c_<eqop>_c_times_r (:259-266)
        vr = sr * xr;
        vi = sr * xi;
    } else {                                           // QLA_c_<eqop>_c_times_c (:326-341)
        if constexpr (EQOP == 3) {
            rr = -sr * xr + si * xi;
            ri = -sr * xi - si * xr;
            return;
        }
        vr = sr * xr - si * xi;
        vi = sr * xi + si * xr;
    }
    if constexpr (EQOP == 0) {
        rr = vr;
        ri = vi;
    } else if constexpr (EQOP == 1) {
        rr = rr + vr;
        ri = ri + vi;
    } else if constexpr (EQOP == 2) {
        rr = rr - vr;
        ri = ri - vi;
    } else {
        rr = -vr;
        ri = -vi;
    }
}

// Work per thread: each thread of a block handles ITEMS work items spaced
// VBLOCK apart inside the block's chunk of ITEMS * VBLOCK, loading all of
// them before it computes and stores any, so each lane keeps ITEMS
// independent loads (x, y and r of every item) in flight.  An item's store
// follows its own loads, so a destination that is also a source at the same
// element (r += a r) still reads the old value.  Measured on MI355X (32^4
// fields): the v-forms run fastest with one item (16 bytes per operand per
// lane); the transposes gain 15-25% with four, whose strided reads inside a
// record then overlap; the index and pointer forms lose 5-10% with more
// than one (their loads are chains of index, address and element).
#ifndef QLA_VEC_FLAT_ITEMS
#define QLA_VEC_FLAT_ITEMS 1
#endif
constexpr int FITEMS = QLA_VEC_FLAT_ITEMS;
constexpr int TRANS_ITEMS = 4;

// v-forms over fields: work item t moves complexes [t*V, t*V + V) of the run
// of m; V = 2 (float, all three bases 16-byte aligned) or 1.  An odd float
// run ends with one single complex in the last item.  The grid is capped
// (a dispatch holds fewer than 2^32 work items) and strides over the chunks.
template <class T, int OP, int EQOP, int V>
__global__ void __launch_bounds__(VBLOCK) k_vec_flat(qla_vec_args p, long long m)
{
    typedef typename vtype<T, V>::v VT;
    typedef typename vtype<T, 1>::v V1;
    T sr, si;
    load_scalar<T, OP>(p, sr, si);
    const bool cj = p.conj != 0;
    const long long items = (m + V - 1) / V;
    constexpr long long CHUNK = (long long)VBLOCK * FITEMS;
    const long long stride = (long long)gridDim.x * CHUNK;
    for (long long t0 = (long long)blockIdx.x * CHUNK + threadIdx.x; t0 < items; t0 += stride) {
        VT x[FITEMS], y[FITEMS], r[FITEMS];
#pragma unroll
        for (int k = 0; k < FITEMS; k++) {
            const long long t = t0 + (long long)k * VBLOCK;
            x[k] = VT{};
            y[k] = VT{};
            r[k] = VT{};
            if (t < items && (V == 1 || t * V + V <= m)) {
                x[k] = __builtin_nontemporal_load((const VT *)p.x + t);
                if constexpr (has_y<OP>()) y[k] = __builtin_nontemporal_load((const VT *)p.y + t);
                if constexpr (reads_r<EQOP>()) r[k] = __builtin_nontemporal_load((const VT *)p.r + t);
            }
        }
#pragma unroll
        for (int k = 0; k < FITEMS; k++) {
            const long long t = t0 + (long long)k * VBLOCK;
            if (t >= items) break;
            const long long e = t * V;
            if (V == 1 || e + V <= m) {
#pragma unroll
                for (int j = 0; j < V; j++) {
                    T rr = r[k][2 * j], ri = r[k][2 * j + 1];
                    const T xr = x[k][2 * j], xi = cj ? -x[k][2 * j + 1] : x[k][2 * j + 1];
                    vop<OP, EQOP, T>(rr, ri, xr, xi, y[k][2 * j], y[k][2 * j + 1], sr, si);
                    r[k][2 * j] = rr;
                    r[k][2 * j + 1] = ri;
                }
                __builtin_nontemporal_store(r[k], (VT *)p.r + t);
            } else {
                // the last complex of an odd float run
                const V1 x1 = __builtin_nontemporal_load((const V1 *)p.x + e);
                V1 y1 = {}, r1 = {};
                if constexpr (has_y<OP>()) y1 = __builtin_nontemporal_load((const V1 *)p.y + e);
                if constexpr (reads_r<EQOP>()) r1 = __builtin_nontemporal_load((const V1 *)p.r + e);
                T rr = r1[0], ri = r1[1];
                vop<OP, EQOP, T>(rr, ri, x1[0], cj ? -x1[1] : x1[1], y1[0], y1[1], sr, si);
                r1[0] = rr;
                r1[1] = ri;
                __builtin_nontemporal_store(r1, (V1 *)p.r + e);
            }
        }
    }
}

// one step's record of an operand: the pointer-array entry or the field record
template <class T, int NC>
__device__ __forceinline__ const cpx<T> *record_of(const void *base, int ptr, int s)
{
    if (ptr) return ((const cpx<T> *const *)base)[s];
    return (const cpx<T> *)base + (size_t)s * NC;
}

// every other form: one work item per destination complex g = i*NC + c,
// strided over a capped grid; the record number is a 32-bit quotient unless
// the run holds 2^32 complexes or more (wide)
template <class T, int NC, int OP, int EQOP, int GITEMS>
__global__ void __launch_bounds__(VBLOCK) k_vec_gather(qla_vec_args p)
{
    typedef typename vtype<T, 1>::v V1;
    T sr, si;
    load_scalar<T, OP>(p, sr, si);
    const unsigned long long tot = (unsigned long long)p.n * NC;
    constexpr unsigned long long CHUNK = (unsigned long long)VBLOCK * GITEMS;
    const unsigned long long stride = (unsigned long long)gridDim.x * CHUNK;
    for (unsigned long long g0 = (unsigned long long)blockIdx.x * CHUNK + threadIdx.x; g0 < tot; g0 += stride) {
        V1 x[GITEMS], y[GITEMS], r[GITEMS];
        V1 *R[GITEMS];
        T fr[GITEMS], fi[GITEMS];
#pragma unroll
        for (int k = 0; k < GITEMS; k++) {
            const unsigned long long g = g0 + (unsigned long long)k * VBLOCK;
            x[k] = V1{};
            y[k] = V1{};
            r[k] = V1{};
            R[k] = nullptr;
            fr[k] = sr;
            fi[k] = si;
            if (g >= tot) continue;
            int i, c;
            if (p.wide) {
                i = (int)(g / NC);
                c = (int)(g - (unsigned long long)i * NC);
            } else {
                const unsigned int g32 = (unsigned int)g;
                i = (int)(g32 / NC);
                c = (int)(g32 - (unsigned int)i * NC);
            }
            const int s = p.index ? p.index[i] : i;
            const int sd = p.ir ? p.ir[i] : s;
            const cpx<T> *X = record_of<T, NC>(p.x, p.px, p.ix ? p.ix[i] : s);
            int cx = c;
            if constexpr (NC == 9 || NC == 144) {
                constexpr int D = NC == 9 ? 3 : 12;     // M: 3 x 3; P: 12 x 12 over (colour, spin)
                if (p.trans) cx = (c % D) * D + c / D;
            }
            x[k] = __builtin_nontemporal_load((const V1 *)(X + cx));
            if constexpr (has_y<OP>()) {
                const cpx<T> *Y = record_of<T, NC>(p.y, p.py, p.iy ? p.iy[i] : s);
                y[k] = __builtin_nontemporal_load((const V1 *)(Y + c));
            }
            if constexpr (is_field<OP>()) {
                constexpr int FW = OP == VOP_CFIELD ? 2 : 1;
                const int fs = p.ifa ? p.ifa[i] : s;
                const T *F = p.pf ? ((const T *const *)p.f)[fs] : (const T *)p.f + (size_t)fs * FW;
                fr[k] = F[0];
                if constexpr (OP == VOP_CFIELD) fi[k] = p.fconj ? -F[1] : F[1];
            }
            R[k] = (V1 *)((cpx<T> *)p.r + (size_t)sd * NC + c);
            if constexpr (reads_r<EQOP>()) r[k] = __builtin_nontemporal_load(R[k]);
        }
#pragma unroll
        for (int k = 0; k < GITEMS; k++) {
            if (!R[k]) continue;
            T rr = r[k][0], ri = r[k][1];
            vop<elem_op<OP>(), EQOP, T>(rr, ri, x[k][0], p.conj ? -x[k][1] : x[k][1], y[k][0], y[k][1], fr[k], fi[k]);
            r[k][0] = rr;
            r[k][1] = ri;
            __builtin_nontemporal_store(r[k], R[k]);
        }
    }
}

// v-form transposes (M: 3 x 3, P: 12 x 12 records of contiguous fields): a
// block stages RB whole source records in LDS with coalesced loads, then each
// work item takes its destination complex's transposed source element from
// LDS, so every source line is fetched once (reading the transposed element
// straight from memory fetched the lines of a P record 1.36 times; staged,
// the P transpose runs at 6.15 instead of 4.79 TB/s).
template <class T, int NC> struct tpose_shape {
    static constexpr int D = NC == 9 ? 3 : 12;
#ifndef QLA_VEC_TPOSE_M_RB
#define QLA_VEC_TPOSE_M_RB 256
#endif
    // records per block step: 256 colour matrices (2304 complexes; measured
    // 4.96 / 5.21 / 5.47 / 4.34 TB/s for 64 / 128 / 256 / 512 on MI355X) or
    // 8 propagators (1152 complexes)
    static constexpr int RB = NC == 9 ? QLA_VEC_TPOSE_M_RB : 8;
};

// W complexes per global access: 2 (16 bytes) for 16-byte aligned
// single-precision fields, else 1; RB is even, so a block's first complex
// e0 is a multiple of W, and an odd complex count leaves one complex to
// thread 0
template <class T, int NC, int OP, int EQOP, int W>
__global__ void __launch_bounds__(VBLOCK) k_vec_tposed(qla_vec_args p)
{
    typedef tpose_shape<T, NC> S;
    typedef typename vtype<T, 1>::v V1;
    typedef typename vtype<T, W>::v VW;
    __shared__ VW tile[S::RB * NC / W];
    const V1 *t1 = (const V1 *)tile;
    T sr, si;
    load_scalar<T, OP>(p, sr, si);
    const VW *X = (const VW *)p.x;
    VW *R = (VW *)p.r;
    // destination complex kk of the block: element (c / D, c % D) of its
    // record is source element (c % D, c / D)
    auto one = [&](int kk, T &rr, T &ri) {
        const int rec = kk / NC, c = kk - rec * NC;
        const V1 x = t1[rec * NC + (c % S::D) * S::D + c / S::D];
        vop<OP, EQOP, T>(rr, ri, x[0], p.conj ? -x[1] : x[1], T(0), T(0), sr, si);
    };
    for (long long base = (long long)blockIdx.x * S::RB; base < p.n; base += (long long)gridDim.x * S::RB) {
        const int nrec = (int)(p.n - base < S::RB ? p.n - base : S::RB);
        const int tot = nrec * NC;
        const long long e0 = base * NC;
        for (int k = threadIdx.x; k < tot / W; k += VBLOCK) tile[k] = __builtin_nontemporal_load(X + e0 / W + k);
        if constexpr (W == 2) {
            if ((tot & 1) && threadIdx.x == 0)
                ((V1 *)tile)[tot - 1] = __builtin_nontemporal_load((const V1 *)p.x + e0 + tot - 1);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < tot / W; k += VBLOCK) {
            VW r = {};
            if constexpr (reads_r<EQOP>()) r = __builtin_nontemporal_load(R + e0 / W + k);
#pragma unroll
            for (int w = 0; w < W; w++) {
                T rr = r[2 * w], ri = r[2 * w + 1];
                one(k * W + w, rr, ri);
                r[2 * w] = rr;
                r[2 * w + 1] = ri;
            }
            __builtin_nontemporal_store(r, R + e0 / W + k);
        }
        if constexpr (W == 2) {
            if ((tot & 1) && threadIdx.x == 0) {
                V1 *r1 = (V1 *)p.r + e0 + tot - 1;
                V1 r = {};
                if constexpr (reads_r<EQOP>()) r = __builtin_nontemporal_load(r1);
                T rr = r[0], ri = r[1];
                one(tot - 1, rr, ri);
                r[0] = rr;
                r[1] = ri;
                __builtin_nontemporal_store(r, r1);
            }
        }
        __syncthreads();
    }
}

// blocks for `items` work items, at most VMAXBLOCKS (2^28 work items per dispatch)
constexpr long long VMAXBLOCKS = 1 << 20;
unsigned blocks_for(long long items)
{
    const long long b = (items + VBLOCK - 1) / VBLOCK;
    return (unsigned)(b < 1 ? 1 : (b > VMAXBLOCKS ? VMAXBLOCKS : b));
}

// Real fields (QLA_F_Real / QLA_D_Real records, the type R of
// build_qla.pl:356-402, 1256-1294, 1421-1456, 1837-1888): the reference's
// real forms r = a, r += s*a, r = s*a + b ... are the real part of the
// complex element arithmetic, which for the operations R has (copies, real
// constants, sums) reads only real parts.  One work item per record; every
// indexing form.  ncomp 0 selects it.
template <class T, int OP, int EQOP>
__global__ void __launch_bounds__(VBLOCK) k_vec_real(qla_vec_args p)
{
    T sr, si;
    load_scalar<T, OP>(p, sr, si);
    const long long stride = (long long)gridDim.x * VBLOCK;
    for (long long g = (long long)blockIdx.x * VBLOCK + threadIdx.x; g < p.n; g += stride) {
        const int i = (int)g;
        const int s = p.index ? p.index[i] : i;
        const int xs = p.ix ? p.ix[i] : s;
        const T x = p.px ? *((const T *const *)p.x)[xs] : ((const T *)p.x)[xs];
        T y = T(0);
        if constexpr (has_y<OP>()) {
            const int ys = p.iy ? p.iy[i] : s;
            y = p.py ? *((const T *const *)p.y)[ys] : ((const T *)p.y)[ys];
        }
        T *R = (T *)p.r + (p.ir ? p.ir[i] : s);
        T rr = T(0), ri = T(0);
        if constexpr (reads_r<EQOP>()) rr = *R;
        vop<OP, EQOP, T>(rr, ri, x, T(0), y, T(0), sr, si);
        *R = rr;
    }
}

template <class T, int OP, int EQOP>
hipError_t launch_op(int ncomp, const qla_vec_args &p, hipStream_t st)
{
    if (ncomp == 0) {
        if constexpr (is_field<OP>() || complex_scalar<OP>() || OP == VOP_I) {
            return hipErrorInvalidValue;
        } else {
            if (p.trans || p.conj) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_vec_real<T, OP, EQOP>), dim3(blocks_for(p.n)), dim3(VBLOCK), 0, st, p);
            return hipGetLastError();
        }
    }
    const bool flat = !p.index && !p.ir && !p.ix && !p.iy && !p.px && !p.py && !p.trans;
    if constexpr (!is_field<OP>()) {
        if (flat) {
            const long long m = (long long)p.n * ncomp;
            if constexpr (sizeof(T) == 4) {
                if (p.al) {
                    hipLaunchKernelGGL((k_vec_flat<T, OP, EQOP, 2>),
                                       dim3(blocks_for(((m + 1) / 2 + FITEMS - 1) / FITEMS)), dim3(VBLOCK), 0, st, p,
                                       m);
                    return hipGetLastError();
                }
            }
            hipLaunchKernelGGL((k_vec_flat<T, OP, EQOP, 1>), dim3(blocks_for((m + FITEMS - 1) / FITEMS)),
                               dim3(VBLOCK), 0, st, p, m);
            return hipGetLastError();
        }
    }
    qla_vec_args q = p;
    const long long tot = (long long)p.n * ncomp;
    q.wide = tot > 0xffffffffLL ? 1 : 0;
    if constexpr (OP <= VOP_CSCALE) {
        if (p.trans && !p.index && !p.ir && !p.ix && !p.px) {
            const long long rb = ncomp == 9 ? tpose_shape<T, 9>::RB : tpose_shape<T, 144>::RB;
            const long long want = (p.n + rb - 1) / rb;
            const dim3 grid((unsigned)(want > VMAXBLOCKS ? VMAXBLOCKS : want));
            constexpr int W = sizeof(T) == 4 ? 2 : 1;
            if (sizeof(T) == 4 && p.al) {
                if (ncomp == 9) hipLaunchKernelGGL((k_vec_tposed<T, 9, OP, EQOP, W>), grid, dim3(VBLOCK), 0, st, q);
                else hipLaunchKernelGGL((k_vec_tposed<T, 144, OP, EQOP, W>), grid, dim3(VBLOCK), 0, st, q);
            } else {
                if (ncomp == 9) hipLaunchKernelGGL((k_vec_tposed<T, 9, OP, EQOP, 1>), grid, dim3(VBLOCK), 0, st, q);
                else hipLaunchKernelGGL((k_vec_tposed<T, 144, OP, EQOP, 1>), grid, dim3(VBLOCK), 0, st, q);
            }
            return hipGetLastError();
        }
    }
    if constexpr (OP <= VOP_CSCALE || is_field<OP>()) {
        if (p.trans) {
            const dim3 grid(blocks_for((tot + TRANS_ITEMS - 1) / TRANS_ITEMS));
            if (ncomp == 9) hipLaunchKernelGGL((k_vec_gather<T, 9, OP, EQOP, TRANS_ITEMS>), grid, dim3(VBLOCK), 0, st, q);
            else hipLaunchKernelGGL((k_vec_gather<T, 144, OP, EQOP, TRANS_ITEMS>), grid, dim3(VBLOCK), 0, st, q);
            return hipGetLastError();
        }
    }
    const dim3 grid(blocks_for(tot));
#define QLA_VEC_CASE(NC)                                                                          \
    case NC:                                                                                     \
        hipLaunchKernelGGL((k_vec_gather<T, NC, OP, EQOP, 1>), grid, dim3(VBLOCK), 0, st, q);   \
        return hipGetLastError();
    switch (ncomp) {
        QLA_VEC_CASE(1)
        QLA_VEC_CASE(3)
        QLA_VEC_CASE(6)
        QLA_VEC_CASE(9)
        QLA_VEC_CASE(12)
        QLA_VEC_CASE(144)
    }
#undef QLA_VEC_CASE
    return hipErrorInvalidValue;
}

template <class T, int OP>
hipError_t launch_eqop(int eqop, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    if constexpr (OP >= VOP_PLUS && !is_field<OP>()) {
        if (eqop != 0) return hipErrorInvalidValue;        // sums and a*b +- c: eq only
        return launch_op<T, OP, 0>(ncomp, p, st);
    } else {
        switch (eqop) {
        case 0: return launch_op<T, OP, 0>(ncomp, p, st);
        case 1: return launch_op<T, OP, 1>(ncomp, p, st);
        case 2: return launch_op<T, OP, 2>(ncomp, p, st);
        case 3: return launch_op<T, OP, 3>(ncomp, p, st);
        }
        return hipErrorInvalidValue;
    }
}

template <class T>
hipError_t launch_t(int op, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    if (p.trans && ncomp != 9 && ncomp != 144) return hipErrorInvalidValue;
    switch (op) {
    case VOP_COPY: return launch_eqop<T, VOP_COPY>(eqop, ncomp, p, st);
    case VOP_I: return launch_eqop<T, VOP_I>(eqop, ncomp, p, st);
    case VOP_RSCALE: return launch_eqop<T, VOP_RSCALE>(eqop, ncomp, p, st);
    case VOP_CSCALE: return launch_eqop<T, VOP_CSCALE>(eqop, ncomp, p, st);
    case VOP_PLUS: return launch_eqop<T, VOP_PLUS>(eqop, ncomp, p, st);
    case VOP_MINUS: return launch_eqop<T, VOP_MINUS>(eqop, ncomp, p, st);
    case VOP_R_PLUS: return launch_eqop<T, VOP_R_PLUS>(eqop, ncomp, p, st);
    case VOP_R_MINUS: return launch_eqop<T, VOP_R_MINUS>(eqop, ncomp, p, st);
    case VOP_C_PLUS: return launch_eqop<T, VOP_C_PLUS>(eqop, ncomp, p, st);
    case VOP_C_MINUS: return launch_eqop<T, VOP_C_MINUS>(eqop, ncomp, p, st);
    case VOP_RFIELD: return launch_eqop<T, VOP_RFIELD>(eqop, ncomp, p, st);
    case VOP_CFIELD: return launch_eqop<T, VOP_CFIELD>(eqop, ncomp, p, st);
    }
    return hipErrorInvalidValue;
}

// Colour outer product (build_qla.pl:1730-1756): destination complex (ic, jc)
// of step i is QLA_c_<eqop>_c_times_ca(r, x[ic], y[jc]), computed as the
// plain product with the conjugated y element (bit-identical, as above).
template <class T, int EQOP>
__global__ void __launch_bounds__(VBLOCK) k_vec_outer(qla_vec_args p)
{
    const long long tot = (long long)p.n * 9;
    const long long stride = (long long)gridDim.x * VBLOCK;
    for (long long g = (long long)blockIdx.x * VBLOCK + threadIdx.x; g < tot; g += stride) {
        const int i = (int)(g / 9), c = (int)(g - (long long)i * 9);
        const int s = p.index ? p.index[i] : i;
        const int sd = p.ir ? p.ir[i] : s;
        const cpx<T> a = record_of<T, 3>(p.x, p.px, p.ix ? p.ix[i] : s)[c / 3];
        const cpx<T> b = record_of<T, 3>(p.y, p.py, p.iy ? p.iy[i] : s)[c % 3];
        cpx<T> *R = (cpx<T> *)p.r + (size_t)sd * 9 + c;
        T rr = T(0), ri = T(0);
        if constexpr (reads_r<EQOP>()) {
            rr = R->re;
            ri = R->im;
        }
        vop<VOP_CSCALE, EQOP, T>(rr, ri, b.re, -b.im, T(0), T(0), a.re, a.im);
        R->re = rr;
        R->im = ri;
    }
}

template <class T>
hipError_t launch_outer_t(int eqop, const qla_vec_args &p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    const dim3 grid(blocks_for((long long)p.n * 9));
    switch (eqop) {
    case 0: hipLaunchKernelGGL((k_vec_outer<T, 0>), grid, dim3(VBLOCK), 0, st, p); break;
    case 1: hipLaunchKernelGGL((k_vec_outer<T, 1>), grid, dim3(VBLOCK), 0, st, p); break;
    case 2: hipLaunchKernelGGL((k_vec_outer<T, 2>), grid, dim3(VBLOCK), 0, st, p); break;
    case 3: hipLaunchKernelGGL((k_vec_outer<T, 3>), grid, dim3(VBLOCK), 0, st, p); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Products with propagators (build_qla.pl:1587-1662): P eqop A * B with
// KIND 0 A a colour matrix and B a propagator (M acting on colour only),
// KIND 1 A a propagator and B a colour matrix, KIND 2 both propagators (the
// 12 x 12 product over (colour, spin) pairs); CA / CB read A / B as adjoints
// (transposed, with the conjugated-factor macros).  Each output element is
// x = 0 (eq, eqm) or r (peq, meq), then x +=/-= a*b with kc outer and ks
// inner (expressions_tensor.pl:544-630), as acc_mul2 writes the macros.
//
// A block step stages PSB steps' A and B records in LDS (a straight 16-byte
// copy of the run when the operand is a plain field, else each step's
// record from its index or pointer entry) and then computes from LDS, one
// work item per (step, row, PMUL_OUTS adjacent columns): the row of A (or
// the colour matrix) is read once for all of them and their columns of B
// are adjacent, so the LDS reads merge into wide ones.  Every record is
// fetched from memory once.
// ---------------------------------------------------------------------------
// outputs per work item, steps per block step (single, double precision)
// and block size; measured on MI355X over 2^20 single-precision P*P, P*P^dagger
// and M*P: (2, 8, 192) 4.61 / 4.93 / 5.92 TB/s, (4, 16, 192) 4.13 / 3.92 /
// 5.41, (1, 16, 256) 3.72 / 3.84 / 5.70, (3, 8, 192) 4.17 / 5.05 / 5.77,
// (2, 4, 192) 4.04 / 4.80 / 5.59.
#ifndef QLA_PMUL_OUTS
#define QLA_PMUL_OUTS 2
#endif
#ifndef QLA_PMUL_PSB
#define QLA_PMUL_PSB 8
#endif
#ifndef QLA_PMUL_PSB_D
#define QLA_PMUL_PSB_D 8
#endif
#ifndef QLA_PMUL_BLOCK
#define QLA_PMUL_BLOCK 192
#endif
constexpr int PBLOCK = QLA_PMUL_BLOCK;
constexpr int POUTS = QLA_PMUL_OUTS;

template <class T> constexpr int pmul_steps() { return sizeof(T) == 4 ? QLA_PMUL_PSB : QLA_PMUL_PSB_D; }

// nrec records of NC complexes for steps base.. into LDS
template <class T, int NC>
__device__ __forceinline__ void pmul_stage(cpx<T> *t, const void *X, int ptr, const int *index, const int *own,
                                           long long base, int nrec)
{
    const int tot = nrec * NC;
    if (!ptr && !index && !own) {
        const cpx<T> *src = (const cpx<T> *)X + base * NC;
        if constexpr (sizeof(T) == 4) {
            if ((((uintptr_t)src) & 15) == 0) {
                typedef float v4 __attribute__((ext_vector_type(4)));
                const v4 *s4 = (const v4 *)src;
                v4 *t4 = (v4 *)t;
                for (int k = threadIdx.x; k < tot / 2; k += PBLOCK) t4[k] = __builtin_nontemporal_load(s4 + k);
                if ((tot & 1) && threadIdx.x == 0) t[tot - 1] = src[tot - 1];
                return;
            }
        }
        for (int k = threadIdx.x; k < tot; k += PBLOCK) t[k] = src[k];
        return;
    }
    for (int k = threadIdx.x; k < tot; k += PBLOCK) {
        const int q = k / NC, c = k - q * NC;
        const long long i = base + q;
        const int s = index ? index[i] : (int)i;
        t[k] = record_of<T, NC>(X, ptr, own ? own[i] : s)[c];
    }
}

template <class T, int KIND, int EQOP, bool CA, bool CB>
__global__ void __launch_bounds__(PBLOCK) k_vec_pmul(qla_vec_args p)
{
    constexpr int PSB = pmul_steps<T>();
    constexpr int NA = KIND == 0 ? 9 : 144, NB = KIND == 1 ? 9 : 144;
    constexpr int CG = 12 / POUTS;     // column groups per row
    constexpr bool SUB = EQOP == 2 || EQOP == 3;
    typedef typename vtype<T, 1>::v V1;
    __shared__ __attribute__((aligned(16))) cpx<T> ta[PSB * NA];
    __shared__ __attribute__((aligned(16))) cpx<T> tb[PSB * NB];
    for (long long base = (long long)blockIdx.x * PSB; base < p.n; base += (long long)gridDim.x * PSB) {
        const int nrec = (int)(p.n - base < PSB ? p.n - base : PSB);
        pmul_stage<T, NA>(ta, p.x, p.px, p.index, p.ix, base, nrec);
        pmul_stage<T, NB>(tb, p.y, p.py, p.index, p.iy, base, nrec);
        __syncthreads();
        for (int k = threadIdx.x; k < nrec * 12 * CG; k += PBLOCK) {
            const int q = k / (12 * CG), w = k - q * (12 * CG);
            const int row = w / CG, c0 = (w - row * CG) * POUTS;
            const long long i = base + q;
            const int s = p.index ? p.index[i] : (int)i;
            const int sd = p.ir ? p.ir[i] : s;
            V1 *R = (V1 *)((cpx<T> *)p.r + (size_t)sd * 144 + row * 12 + c0);
            cpx<T> x[POUTS];
#pragma unroll
            for (int j = 0; j < POUTS; j++) {
                x[j] = { T(0), T(0) };
                if constexpr (reads_r<EQOP>()) {
                    const V1 v = __builtin_nontemporal_load(R + j);
                    x[j] = { v[0], v[1] };
                }
            }
            const cpx<T> *A = ta + q * NA, *B = tb + q * NB;
#pragma unroll
            for (int j = 0; j < POUTS; j++) {
                const int col = c0 + j;
                if constexpr (KIND == 0) {
                    const int ic = row >> 2, is = row & 3;
#pragma unroll
                    for (int kc = 0; kc < 3; kc++)
                        acc_mul2<SUB, CA, CB>(x[j], CA ? A[kc * 3 + ic] : A[ic * 3 + kc],
                                              CB ? B[col * 12 + kc * 4 + is] : B[(kc * 4 + is) * 12 + col]);
                } else if constexpr (KIND == 1) {
                    const int jc = col >> 2, js = col & 3;
#pragma unroll
                    for (int kc = 0; kc < 3; kc++)
                        acc_mul2<SUB, CA, CB>(x[j], CA ? A[(kc * 4 + js) * 12 + row] : A[row * 12 + kc * 4 + js],
                                              CB ? B[jc * 3 + kc] : B[kc * 3 + jc]);
                } else {
#pragma unroll
                    for (int kk = 0; kk < 12; kk++)
                        acc_mul2<SUB, CA, CB>(x[j], CA ? A[kk * 12 + row] : A[row * 12 + kk],
                                              CB ? B[col * 12 + kk] : B[kk * 12 + col]);
                }
            }
#pragma unroll
            for (int j = 0; j < POUTS; j++) __builtin_nontemporal_store(V1{ x[j].re, x[j].im }, R + j);
        }
        __syncthreads();
    }
}

template <class T, int KIND, int EQOP>
hipError_t launch_pmul_k(int ca, int cb, dim3 grid, const qla_vec_args &p, hipStream_t st)
{
    switch (ca * 2 + cb) {
    case 0: hipLaunchKernelGGL((k_vec_pmul<T, KIND, EQOP, false, false>), grid, dim3(PBLOCK), 0, st, p); break;
    case 1: hipLaunchKernelGGL((k_vec_pmul<T, KIND, EQOP, false, true>), grid, dim3(PBLOCK), 0, st, p); break;
    case 2: hipLaunchKernelGGL((k_vec_pmul<T, KIND, EQOP, true, false>), grid, dim3(PBLOCK), 0, st, p); break;
    case 3: hipLaunchKernelGGL((k_vec_pmul<T, KIND, EQOP, true, true>), grid, dim3(PBLOCK), 0, st, p); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class T, int KIND>
hipError_t launch_pmul_e(int eqop, int ca, int cb, dim3 grid, const qla_vec_args &p, hipStream_t st)
{
    switch (eqop) {
    case 0: return launch_pmul_k<T, KIND, 0>(ca, cb, grid, p, st);
    case 1: return launch_pmul_k<T, KIND, 1>(ca, cb, grid, p, st);
    case 2: return launch_pmul_k<T, KIND, 2>(ca, cb, grid, p, st);
    case 3: return launch_pmul_k<T, KIND, 3>(ca, cb, grid, p, st);
    default: return hipErrorInvalidValue;
    }
}

template <class T>
hipError_t launch_pmul_t(int kind, int eqop, int ca, int cb, const qla_vec_args &p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    if ((ca | cb) & ~1) return hipErrorInvalidValue;
    constexpr long long PSB = pmul_steps<T>();
    const long long steps = (p.n + PSB - 1) / PSB;
    const dim3 grid((unsigned)(steps > VMAXBLOCKS ? VMAXBLOCKS : steps));
    switch (kind) {
    case 0: return launch_pmul_e<T, 0>(eqop, ca, cb, grid, p, st);
    case 1: return launch_pmul_e<T, 1>(eqop, ca, cb, grid, p, st);
    case 2: return launch_pmul_e<T, 2>(eqop, ca, cb, grid, p, st);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// Component access (build_qla.pl:767-932, expressions_tensor.pl:297-347) and
// the spin trace (build_qla.pl:1066-1086, expressions_tensor.pl:394-430),
// eq only.  The host turns the element, colour-column or Dirac-column
// arguments into a position map (qla_comp_map): element c of the moved part
// sits at off + (c / inner) * s1 + (c % inner) * s2 of the array record.
// MODE 0 extracts into records of nd complexes, MODE 1 inserts (only those nd
// positions of each destination record are written), MODE 2 is the spin
// trace M[ic][jc] = x, x = 0 then x += P[ic][is][jc][is] for is = 0..3.  One
// work item per moved complex.
// ---------------------------------------------------------------------------
template <class T, int MODE>
__global__ void __launch_bounds__(VBLOCK) k_vec_comp(qla_vec_args p, qla_comp_map m)
{
    const long long tot = (long long)p.n * m.nd;
    const long long stride = (long long)gridDim.x * VBLOCK;
    for (long long g = (long long)blockIdx.x * VBLOCK + threadIdx.x; g < tot; g += stride) {
        const int i = (int)(g / m.nd), c = (int)(g - (long long)i * m.nd);
        const int s = p.index ? p.index[i] : i;
        const int sd = p.ir ? p.ir[i] : s;
        const int sa = p.ix ? p.ix[i] : s;
        const cpx<T> *A = p.px ? ((const cpx<T> *const *)p.x)[sa] : (const cpx<T> *)p.x + (size_t)sa * m.nsrc;
        cpx<T> *R = (cpx<T> *)p.r + (size_t)sd * m.ndst;
        if constexpr (MODE == 2) {
            const int ic = c / 3, jc = c - ic * 3;
            cpx<T> x = { T(0), T(0) };
#pragma unroll
            for (int is = 0; is < 4; is++) {
                const cpx<T> v = A[ic * 48 + is * 12 + jc * 4 + is];
                x.re = x.re + v.re;
                x.im = x.im + v.im;
            }
            R[c] = x;
        } else {
            const int pos = m.off + (c / m.inner) * m.s1 + (c % m.inner) * m.s2;
            if constexpr (MODE == 0) R[c] = A[pos];
            else R[pos] = A[c];
        }
    }
}

template <class T>
hipError_t launch_comp_t(int mode, const qla_vec_args &p, const qla_comp_map &m, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    const dim3 grid(blocks_for((long long)p.n * m.nd));
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_vec_comp<T, 0>), grid, dim3(VBLOCK), 0, st, p, m); break;
    case 1: hipLaunchKernelGGL((k_vec_comp<T, 1>), grid, dim3(VBLOCK), 0, st, p, m); break;
    case 2: hipLaunchKernelGGL((k_vec_comp<T, 2>), grid, dim3(VBLOCK), 0, st, p, m); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Conversions between real and complex fields and the division of scalar
// fields, eq only (build_qla.pl:623-708, 1457-1482): KIND 0 C = R
// (QLA_c_eq_r_plus_ir(c, a, 0.)), 1 C = R + i R, 2 R = re C, 3 R = im C,
// 4 R = R / R, 6 R = I, 7 I = trunc R ((int) a), 8 I = round R ((int) roundf / round,
// build_qla.pl:709-766), 5 C = C / C (QLA_c_eq_c_div_c, include/qla_complex.h:417-427:
// t = 1.0 / |b|^2 in double rounded to the field type, then
// (a.re b.re + a.im b.im) t and (a.im b.re - a.re b.im) t).  Single-precision
// quotients are taken in double and rounded, which is the correctly rounded
// quotient as the reference computes it whatever the compiler's division
// flags.  One work item per record; every indexing form.
// ---------------------------------------------------------------------------
template <class T> __device__ __forceinline__ T sc_real(const void *base, int ptr, int k)
{
    return ptr ? *((const T *const *)base)[k] : ((const T *)base)[k];
}
template <class T> __device__ __forceinline__ cpx<T> sc_cpx(const void *base, int ptr, int k)
{
    return ptr ? *((const cpx<T> *const *)base)[k] : ((const cpx<T> *)base)[k];
}

template <class T, int KIND>
__global__ void __launch_bounds__(VBLOCK) k_vec_sconv(qla_vec_args p)
{
    const long long stride = (long long)gridDim.x * VBLOCK;
    for (long long g = (long long)blockIdx.x * VBLOCK + threadIdx.x; g < p.n; g += stride) {
        const int i = (int)g;
        const int s = p.index ? p.index[i] : i;
        const int xs = p.ix ? p.ix[i] : s;
        const int ys = p.iy ? p.iy[i] : s;
        const int d = p.ir ? p.ir[i] : s;
        if constexpr (KIND == 0) {
            ((cpx<T> *)p.r)[d] = { sc_real<T>(p.x, p.px, xs), T(0) };
        } else if constexpr (KIND == 1) {
            ((cpx<T> *)p.r)[d] = { sc_real<T>(p.x, p.px, xs), sc_real<T>(p.y, p.py, ys) };
        } else if constexpr (KIND == 2) {
            ((T *)p.r)[d] = sc_cpx<T>(p.x, p.px, xs).re;
        } else if constexpr (KIND == 3) {
            ((T *)p.r)[d] = sc_cpx<T>(p.x, p.px, xs).im;
        } else if constexpr (KIND == 6) {
            ((T *)p.r)[d] = (T)(p.px ? *((const int *const *)p.x)[xs] : ((const int *)p.x)[xs]);
        } else if constexpr (KIND == 7) {
            ((int *)p.r)[d] = (int)sc_real<T>(p.x, p.px, xs);
        } else if constexpr (KIND == 8) {
            const T a = sc_real<T>(p.x, p.px, xs);
            if constexpr (sizeof(T) == 4) ((int *)p.r)[d] = (int)roundf(a);
            else ((int *)p.r)[d] = (int)round(a);
        } else if constexpr (KIND == 4) {
            ((T *)p.r)[d] = (T)((double)sc_real<T>(p.x, p.px, xs) / (double)sc_real<T>(p.y, p.py, ys));
        } else {
            const cpx<T> a = sc_cpx<T>(p.x, p.px, xs), b = sc_cpx<T>(p.y, p.py, ys);
            const T n2 = b.re * b.re + b.im * b.im;
            const T t = (T)(1.0 / (double)n2);
            ((cpx<T> *)p.r)[d] = { (a.re * b.re + a.im * b.im) * t, (a.im * b.re - a.re * b.im) * t };
        }
    }
}

template <class T>
hipError_t launch_sconv_t(int kind, const qla_vec_args &p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    const dim3 grid(blocks_for(p.n));
    switch (kind) {
    case 0: hipLaunchKernelGGL((k_vec_sconv<T, 0>), grid, dim3(VBLOCK), 0, st, p); break;
    case 1: hipLaunchKernelGGL((k_vec_sconv<T, 1>), grid, dim3(VBLOCK), 0, st, p); break;
    case 2: hipLaunchKernelGGL((k_vec_sconv<T, 2>), grid, dim3(VBLOCK), 0, st, p); break;
    case 3: hipLaunchKernelGGL((k_vec_sconv<T, 3>), grid, dim3(VBLOCK), 0, st, p); break;
    case 4: hipLaunchKernelGGL((k_vec_sconv<T, 4>), grid, dim3(VBLOCK), 0, st, p); break;
    case 5: hipLaunchKernelGGL((k_vec_sconv<T, 5>), grid, dim3(VBLOCK), 0, st, p); break;
    case 6: hipLaunchKernelGGL((k_vec_sconv<T, 6>), grid, dim3(VBLOCK), 0, st, p); break;
    case 7: hipLaunchKernelGGL((k_vec_sconv<T, 7>), grid, dim3(VBLOCK), 0, st, p); break;
    case 8: hipLaunchKernelGGL((k_vec_sconv<T, 8>), grid, dim3(VBLOCK), 0, st, p); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Elementary functions of real and complex fields, eq only (build_qla.pl:
// 194-355; make_code_unary.pl:103-176, expressions_tensor.pl:1021-1075,
// lib/cmath).  FN 0..16 R = f(R) (cos sin tan acos asin atan sqrt fabs exp
// log sign cosh sinh tanh log10 floor ceil; the single-precision function
// for float fields, sign as a >= 0 ? 1 : -1); 20..25 R = a op b (fmod, max
// as a > b ? a : b, min as a > b ? b : a, pow, atan2, ldexp with an int
// field b) taken in double and rounded, as the reference's double calls;
// 30 C = cexpi(R); 31 R = |C|; 32 R = arg C; 33 cexp; 34 csqrt
// (lib/cmath/QLA_csqrt.c); 35 clog (0.5 log|z|^2 in double, arg z).  The
// transcendental functions are the device library's, within a few units in
// the last place of the C library's; the rest are exact.
// ---------------------------------------------------------------------------
#define QLA_SFN_MATH(NAME)                                                               \
    template <class T> __device__ __forceinline__ T m_##NAME(T x)                        \
    {                                                                                    \
        if constexpr (sizeof(T) == 4) return NAME##f(x);                                 \
        else return NAME(x);                                                             \
    }
QLA_SFN_MATH(cos)
QLA_SFN_MATH(sin)
QLA_SFN_MATH(tan)
QLA_SFN_MATH(acos)
QLA_SFN_MATH(asin)
QLA_SFN_MATH(atan)
QLA_SFN_MATH(sqrt)
QLA_SFN_MATH(fabs)
QLA_SFN_MATH(exp)
QLA_SFN_MATH(log)
QLA_SFN_MATH(cosh)
QLA_SFN_MATH(sinh)
QLA_SFN_MATH(tanh)
QLA_SFN_MATH(log10)
QLA_SFN_MATH(floor)
QLA_SFN_MATH(ceil)
#undef QLA_SFN_MATH
template <class T> __device__ __forceinline__ T m_atan2(T y, T x)
{
    if constexpr (sizeof(T) == 4) return atan2f(y, x);
    else return atan2(y, x);
}
template <class T> __device__ __forceinline__ T m_copysign(T x, T y)
{
    if constexpr (sizeof(T) == 4) return copysignf(x, y);
    else return copysign(x, y);
}

template <class T, int FN> __device__ __forceinline__ T sfn1(T a)
{
    if constexpr (FN == 0) return m_cos(a);
    else if constexpr (FN == 1) return m_sin(a);
    else if constexpr (FN == 2) return m_tan(a);
    else if constexpr (FN == 3) return m_acos(a);
    else if constexpr (FN == 4) return m_asin(a);
    else if constexpr (FN == 5) return m_atan(a);
    else if constexpr (FN == 6) return m_sqrt(a);
    else if constexpr (FN == 7) return m_fabs(a);
    else if constexpr (FN == 8) return m_exp(a);
    else if constexpr (FN == 9) return m_log(a);
    else if constexpr (FN == 10) return a >= T(0) ? T(1) : T(-1);
    else if constexpr (FN == 11) return m_cosh(a);
    else if constexpr (FN == 12) return m_sinh(a);
    else if constexpr (FN == 13) return m_tanh(a);
    else if constexpr (FN == 14) return m_log10(a);
    else if constexpr (FN == 15) return m_floor(a);
    else return m_ceil(a);
}

template <class T, int FN>
__global__ void __launch_bounds__(VBLOCK) k_vec_sfn(qla_vec_args p)
{
    const long long stride = (long long)gridDim.x * VBLOCK;
    for (long long g = (long long)blockIdx.x * VBLOCK + threadIdx.x; g < p.n; g += stride) {
        const int i = (int)g;
        const int s = p.index ? p.index[i] : i;
        const int xs = p.ix ? p.ix[i] : s;
        const int ys = p.iy ? p.iy[i] : s;
        const int d = p.ir ? p.ir[i] : s;
        if constexpr (FN < 20) {
            ((T *)p.r)[d] = sfn1<T, FN>(sc_real<T>(p.x, p.px, xs));
        } else if constexpr (FN < 30) {
            const double a = sc_real<T>(p.x, p.px, xs);
            double v;
            if constexpr (FN == 25) {
                v = ldexp(a, p.py ? *((const int *const *)p.y)[ys] : ((const int *)p.y)[ys]);
            } else {
                const double b = sc_real<T>(p.y, p.py, ys);
                if constexpr (FN == 20) v = fmod(a, b);
                else if constexpr (FN == 21) v = a > b ? a : b;
                else if constexpr (FN == 22) v = a > b ? b : a;
                else if constexpr (FN == 23) v = pow(a, b);
                else v = atan2(a, b);
            }
            ((T *)p.r)[d] = (T)v;
        } else if constexpr (FN == 30) {
            const T a = sc_real<T>(p.x, p.px, xs);
            ((cpx<T> *)p.r)[d] = { m_cos(a), m_sin(a) };
        } else {
            const cpx<T> z = sc_cpx<T>(p.x, p.px, xs);
            if constexpr (FN == 31) {
                ((T *)p.r)[d] = m_sqrt(z.re * z.re + z.im * z.im);
            } else if constexpr (FN == 32) {
                ((T *)p.r)[d] = m_atan2(z.im, z.re);
            } else if constexpr (FN == 33) {
                const T mag = m_exp(z.re);
                ((cpx<T> *)p.r)[d] = { mag * m_cos(z.im), mag * m_sin(z.im) };
            } else if constexpr (FN == 34) {
                const T nz = m_sqrt(z.re * z.re + z.im * z.im);
                const T sr = m_sqrt(m_fabs((T)(0.5 * (double)(nz + z.re))));
                const T si = m_copysign(m_sqrt(m_fabs((T)(0.5 * (double)(nz - z.re)))), z.im);
                ((cpx<T> *)p.r)[d] = { sr, si };
            } else {
                const T n2 = z.re * z.re + z.im * z.im;
                ((cpx<T> *)p.r)[d] = { (T)(0.5 * log((double)n2)), m_atan2(z.im, z.re) };
            }
        }
    }
}

template <class T>
hipError_t launch_sfn_t(int fn, const qla_vec_args &p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    const dim3 grid(blocks_for(p.n));
#define QLA_SFN_CASE(F)                                                                   \
    case F: hipLaunchKernelGGL((k_vec_sfn<T, F>), grid, dim3(VBLOCK), 0, st, p); break;
    switch (fn) {
        QLA_SFN_CASE(0) QLA_SFN_CASE(1) QLA_SFN_CASE(2) QLA_SFN_CASE(3) QLA_SFN_CASE(4) QLA_SFN_CASE(5)
        QLA_SFN_CASE(6) QLA_SFN_CASE(7) QLA_SFN_CASE(8) QLA_SFN_CASE(9) QLA_SFN_CASE(10) QLA_SFN_CASE(11)
        QLA_SFN_CASE(12) QLA_SFN_CASE(13) QLA_SFN_CASE(14) QLA_SFN_CASE(15) QLA_SFN_CASE(16)
        QLA_SFN_CASE(20) QLA_SFN_CASE(21) QLA_SFN_CASE(22) QLA_SFN_CASE(23) QLA_SFN_CASE(24) QLA_SFN_CASE(25)
        QLA_SFN_CASE(30) QLA_SFN_CASE(31) QLA_SFN_CASE(32) QLA_SFN_CASE(33) QLA_SFN_CASE(34) QLA_SFN_CASE(35)
    default: return hipErrorInvalidValue;
    }
#undef QLA_SFN_CASE
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Copymask (build_qla.pl:2058-2088, expressions_tensor.pl:1065-1070): the
// record of step i is copied, r = a, where the int mask b is nonzero and left
// as it is elsewhere; nw words of type T per record (2 * ncomp for the
// complex types, 1 for R).  One work item per word.  Comparisons of real
// fields (build_qla.pl:1976-2018): the int r = (a op b), 1 or 0, for op
// ==, !=, >, <, >=, <= (OP 0..5, expressions_tensor.pl:1009-1019).
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(VBLOCK) k_vec_mask(qla_vec_args p, int nw)
{
    const long long tot = (long long)p.n * nw;
    const long long stride = (long long)gridDim.x * VBLOCK;
    for (long long g = (long long)blockIdx.x * VBLOCK + threadIdx.x; g < tot; g += stride) {
        const int i = (int)(g / nw), w = (int)(g - (long long)i * nw);
        const int s = p.index ? p.index[i] : i;
        const int ys = p.iy ? p.iy[i] : s;
        const int mask = p.py ? *((const int *const *)p.y)[ys] : ((const int *)p.y)[ys];
        if (!mask) continue;
        const int xs = p.ix ? p.ix[i] : s;
        const T *A = p.px ? ((const T *const *)p.x)[xs] : (const T *)p.x + (size_t)xs * nw;
        ((T *)p.r)[(size_t)(p.ir ? p.ir[i] : s) * nw + w] = A[w];
    }
}

template <class T, int OP>
__global__ void __launch_bounds__(VBLOCK) k_vec_cmp(qla_vec_args p)
{
    const long long stride = (long long)gridDim.x * VBLOCK;
    for (long long g = (long long)blockIdx.x * VBLOCK + threadIdx.x; g < p.n; g += stride) {
        const int i = (int)g;
        const int s = p.index ? p.index[i] : i;
        const T a = sc_real<T>(p.x, p.px, p.ix ? p.ix[i] : s);
        const T b = sc_real<T>(p.y, p.py, p.iy ? p.iy[i] : s);
        int v;
        if constexpr (OP == 0) v = a == b;
        else if constexpr (OP == 1) v = a != b;
        else if constexpr (OP == 2) v = a > b;
        else if constexpr (OP == 3) v = a < b;
        else if constexpr (OP == 4) v = a >= b;
        else v = a <= b;
        ((int *)p.r)[p.ir ? p.ir[i] : s] = v;
    }
}

template <class T>
hipError_t launch_mask_t(int nw, const qla_vec_args &p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    if (nw < 1 || nw > 288) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_vec_mask<T>, dim3(blocks_for((long long)p.n * nw)), dim3(VBLOCK), 0, st, p, nw);
    return hipGetLastError();
}

template <class T>
hipError_t launch_cmp_t(int op, const qla_vec_args &p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    const dim3 grid(blocks_for(p.n));
    switch (op) {
    case 0: hipLaunchKernelGGL((k_vec_cmp<T, 0>), grid, dim3(VBLOCK), 0, st, p); break;
    case 1: hipLaunchKernelGGL((k_vec_cmp<T, 1>), grid, dim3(VBLOCK), 0, st, p); break;
    case 2: hipLaunchKernelGGL((k_vec_cmp<T, 2>), grid, dim3(VBLOCK), 0, st, p); break;
    case 3: hipLaunchKernelGGL((k_vec_cmp<T, 3>), grid, dim3(VBLOCK), 0, st, p); break;
    case 4: hipLaunchKernelGGL((k_vec_cmp<T, 4>), grid, dim3(VBLOCK), 0, st, p); break;
    case 5: hipLaunchKernelGGL((k_vec_cmp<T, 5>), grid, dim3(VBLOCK), 0, st, p); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Integer fields (QLA_Int records; the integer cases of build_qla.pl:356-402,
// 1256-1294, 1421-1516, 1837-1888, 1936-2088, 2561-2620): r eqop v with v
// one of (OP) 0 a, 1 i*a (the constant), 2 !a, 10 a + b, 11 a - b, 12 a / b,
// 13 a * b, 14 i*a + b, 15 i*a - b, 16 a << b, 17 a >> b, 18 a % b, 19 max
// (a > b ? a : b), 20 min (a > b ? b : a), 21..26 a == != > < >= <= b,
// 27 a | b, 28 a & b, 29 a ^ b, 30 a where b is nonzero (copymask); eqop eq
// r = v, peq r += v, meq r -= v, eqm r = -v.  C integer semantics (division
// and % truncate toward zero).  OP 40 / 41 are the fills r = 0 and
// r eqop i.  One work item per record.
// ---------------------------------------------------------------------------
template <int OP, int EQOP>
__global__ void __launch_bounds__(VBLOCK) k_int(qla_vec_args p)
{
    int sc = 0;
    if constexpr (OP == 1 || OP == 14 || OP == 15 || OP == 41) sc = p.sp ? *(const int *)p.sp : (int)p.s_re;
    const long long stride = (long long)gridDim.x * VBLOCK;
    for (long long g = (long long)blockIdx.x * VBLOCK + threadIdx.x; g < p.n; g += stride) {
        const int i = (int)g;
        const int s = p.index ? p.index[i] : i;
        int *R = (int *)p.r + (p.ir ? p.ir[i] : s);
        int v = 0;
        if constexpr (OP == 41) {
            v = sc;
        } else if constexpr (OP != 40) {
            const int xs = p.ix ? p.ix[i] : s;
            const int a = p.px ? *((const int *const *)p.x)[xs] : ((const int *)p.x)[xs];
            int b = 0;
            if constexpr (OP >= 10) {
                const int ys = p.iy ? p.iy[i] : s;
                b = p.py ? *((const int *const *)p.y)[ys] : ((const int *)p.y)[ys];
            }
            if constexpr (OP == 0) v = a;
            else if constexpr (OP == 1) v = sc * a;
            else if constexpr (OP == 2) v = !a;
            else if constexpr (OP == 10) v = a + b;
            else if constexpr (OP == 11) v = a - b;
            else if constexpr (OP == 12) v = a / b;
            else if constexpr (OP == 13) v = a * b;
            else if constexpr (OP == 14) v = sc * a + b;
            else if constexpr (OP == 15) v = sc * a - b;
            else if constexpr (OP == 16) v = a << b;
            else if constexpr (OP == 17) v = a >> b;
            else if constexpr (OP == 18) v = a % b;
            else if constexpr (OP == 19) v = a > b ? a : b;
            else if constexpr (OP == 20) v = a > b ? b : a;
            else if constexpr (OP == 21) v = a == b;
            else if constexpr (OP == 22) v = a != b;
            else if constexpr (OP == 23) v = a > b;
            else if constexpr (OP == 24) v = a < b;
            else if constexpr (OP == 25) v = a >= b;
            else if constexpr (OP == 26) v = a <= b;
            else if constexpr (OP == 27) v = a | b;
            else if constexpr (OP == 28) v = a & b;
            else if constexpr (OP == 29) v = a ^ b;
            else {
                if (!b) continue;
                v = a;
            }
        }
        if constexpr (EQOP == 0) *R = v;
        else if constexpr (EQOP == 1) *R += v;
        else if constexpr (EQOP == 2) *R -= v;
        else *R = -v;
    }
}

template <int OP>
hipError_t launch_int_e(int eqop, const qla_vec_args &p, hipStream_t st)
{
    const dim3 grid(blocks_for(p.n));
    switch (eqop) {
    case 0: hipLaunchKernelGGL((k_int<OP, 0>), grid, dim3(VBLOCK), 0, st, p); break;
    case 1: hipLaunchKernelGGL((k_int<OP, 1>), grid, dim3(VBLOCK), 0, st, p); break;
    case 2: hipLaunchKernelGGL((k_int<OP, 2>), grid, dim3(VBLOCK), 0, st, p); break;
    case 3: hipLaunchKernelGGL((k_int<OP, 3>), grid, dim3(VBLOCK), 0, st, p); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_int(int op, int eqop, const qla_vec_args &p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
#define QLA_INT_CASE(O) \
    case O: return launch_int_e<O>(eqop, p, st);
    switch (op) {
        QLA_INT_CASE(0) QLA_INT_CASE(1) QLA_INT_CASE(2) QLA_INT_CASE(10) QLA_INT_CASE(11) QLA_INT_CASE(12)
        QLA_INT_CASE(13) QLA_INT_CASE(14) QLA_INT_CASE(15) QLA_INT_CASE(16) QLA_INT_CASE(17) QLA_INT_CASE(18)
        QLA_INT_CASE(19) QLA_INT_CASE(20) QLA_INT_CASE(21) QLA_INT_CASE(22) QLA_INT_CASE(23) QLA_INT_CASE(24)
        QLA_INT_CASE(25) QLA_INT_CASE(26) QLA_INT_CASE(27) QLA_INT_CASE(28) QLA_INT_CASE(29) QLA_INT_CASE(30)
        QLA_INT_CASE(40) QLA_INT_CASE(41)
    }
#undef QLA_INT_CASE
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Fills (build_qla.pl:2561-2620).  FILL_ZERO: QLA_c_eq_r(e, 0.) on every
// element.  FILL_REC: QLA_c_<eqop>_c(e, a.e) with one record a for every
// step.  FILL_DIAG (colour matrices): QLA_c_<eqop>_r(e, 0.) on every element
// -- eq (0, 0), peq re += 0, meq re -= 0, eqm (-0, 0) -- then
// QLA_c_<eqop>_c(e, c) on the diagonal (expressions_tensor.pl:77-128).  A
// record or constant from host memory arrives by value in the kernel
// arguments (at most 144 complexes: 2304 bytes in double precision).
// ---------------------------------------------------------------------------
enum { FILL_ZERO = 0, FILL_REC = 1, FILL_DIAG = 2 };

template <class T, int N> struct rec_val { cpx<T> e[N]; };

// v-form zero fill: 16 bytes of zeros per work item over the run of m
// complexes (r 16-byte aligned)
template <class T>
__global__ void __launch_bounds__(VBLOCK) k_vec_zero(cpx<T> *r, long long m)
{
    typedef typename vtype<T, 1>::v V1;
    const long long stride = (long long)gridDim.x * VBLOCK;
    if constexpr (sizeof(T) == 4) {
        typedef typename vtype<T, 2>::v V2;
        const long long items = (m + 1) / 2;
        for (long long t = (long long)blockIdx.x * VBLOCK + threadIdx.x; t < items; t += stride) {
            if (2 * t + 1 < m) __builtin_nontemporal_store(V2{}, (V2 *)r + t);
            else __builtin_nontemporal_store(V1{}, (V1 *)r + 2 * t);
        }
    } else {
        for (long long t = (long long)blockIdx.x * VBLOCK + threadIdx.x; t < m; t += stride)
            __builtin_nontemporal_store(V1{}, (V1 *)r + t);
    }
}

template <class T, int NC, int KIND, int EQOP, int HN>
__global__ void __launch_bounds__(VBLOCK) k_vec_fill(qla_fill_args p, rec_val<T, HN> hv, int wide)
{
    typedef typename vtype<T, 1>::v V1;
    const unsigned long long tot = (unsigned long long)p.n * NC;
    const unsigned long long stride = (unsigned long long)gridDim.x * VBLOCK;
    const cpx<T> *src = (const cpx<T> *)p.src;
    for (unsigned long long g = (unsigned long long)blockIdx.x * VBLOCK + threadIdx.x; g < tot; g += stride) {
        int i, c;
        if (wide) {
            i = (int)(g / NC);
            c = (int)(g - (unsigned long long)i * NC);
        } else {
            const unsigned int g32 = (unsigned int)g;
            i = (int)(g32 / NC);
            c = (int)(g32 - (unsigned int)i * NC);
        }
        const int s = p.index ? p.index[i] : i;
        V1 *R = (V1 *)((cpx<T> *)p.r + (size_t)s * NC + c);
        V1 r = {};
        if constexpr (reads_r<EQOP>()) r = __builtin_nontemporal_load(R);
        T rr = r[0], ri = r[1];
        if constexpr (KIND == FILL_ZERO) {
            rr = T(0);
            ri = T(0);
        } else if constexpr (KIND == FILL_REC) {
            const cpx<T> a = src ? src[c] : hv.e[c];
            vop<VOP_COPY, EQOP, T>(rr, ri, a.re, a.im, T(0), T(0), T(0), T(0));
        } else {
            if constexpr (EQOP == 0) {
                rr = T(0);
                ri = T(0);
            } else if constexpr (EQOP == 1) {
                rr = rr + T(0);
            } else if constexpr (EQOP == 2) {
                rr = rr - T(0);
            } else {
                rr = -T(0);
                ri = T(0);
            }
            if (c / 3 == c % 3) {
                const cpx<T> a = src ? src[0] : hv.e[0];
                vop<VOP_COPY, EQOP, T>(rr, ri, a.re, a.im, T(0), T(0), T(0), T(0));
            }
        }
        r[0] = rr;
        r[1] = ri;
        __builtin_nontemporal_store(r, R);
    }
}

template <class T, int NC, int KIND, int EQOP>
hipError_t launch_fill_nc(const qla_fill_args &p, hipStream_t st)
{
    constexpr int HN = KIND == FILL_REC ? NC : 1;
    rec_val<T, HN> hv = {};
    if (KIND != FILL_ZERO && !p.src) memcpy(&hv, p.host, sizeof(hv));
    const long long tot = (long long)p.n * NC;
    const int wide = tot > 0xffffffffLL ? 1 : 0;
    hipLaunchKernelGGL((k_vec_fill<T, NC, KIND, EQOP, HN>), dim3(blocks_for(tot)), dim3(VBLOCK), 0, st, p, hv,
                       wide);
    return hipGetLastError();
}

template <class T, int KIND, int EQOP>
hipError_t launch_fill_k(int ncomp, const qla_fill_args &p, hipStream_t st)
{
    if constexpr (KIND == FILL_DIAG) {
        return ncomp == 9 ? launch_fill_nc<T, 9, KIND, EQOP>(p, st) : hipErrorInvalidValue;
    } else {
        switch (ncomp) {
        case 1: return launch_fill_nc<T, 1, KIND, EQOP>(p, st);
        case 3: return launch_fill_nc<T, 3, KIND, EQOP>(p, st);
        case 6: return launch_fill_nc<T, 6, KIND, EQOP>(p, st);
        case 9: return launch_fill_nc<T, 9, KIND, EQOP>(p, st);
        case 12: return launch_fill_nc<T, 12, KIND, EQOP>(p, st);
        case 144: return launch_fill_nc<T, 144, KIND, EQOP>(p, st);
        }
        return hipErrorInvalidValue;
    }
}

// fills of real fields (ncomp 0): r = 0 (QLA_r_eq_r(e, 0.)) or r eqop a with
// the one real a from device memory or the value the host passes
template <class T, int KIND, int EQOP>
__global__ void __launch_bounds__(VBLOCK) k_vec_rfill(qla_fill_args p, T hv)
{
    const T a = KIND == FILL_ZERO ? T(0) : (p.src ? *(const T *)p.src : hv);
    const long long stride = (long long)gridDim.x * VBLOCK;
    for (long long g = (long long)blockIdx.x * VBLOCK + threadIdx.x; g < p.n; g += stride) {
        T *R = (T *)p.r + (p.index ? p.index[g] : g);
        if constexpr (KIND == FILL_ZERO) {
            *R = T(0);
        } else {
            T rr = T(0), ri = T(0);
            if constexpr (reads_r<EQOP>()) rr = *R;
            vop<VOP_COPY, EQOP, T>(rr, ri, a, T(0), T(0), T(0), T(0), T(0));
            *R = rr;
        }
    }
}

template <class T, int KIND, int EQOP>
hipError_t launch_rfill(const qla_fill_args &p, hipStream_t st)
{
    T hv = T(0);
    if (KIND != FILL_ZERO && !p.src) memcpy(&hv, p.host, sizeof(T));
    hipLaunchKernelGGL((k_vec_rfill<T, KIND, EQOP>), dim3(blocks_for(p.n)), dim3(VBLOCK), 0, st, p, hv);
    return hipGetLastError();
}

template <class T, int KIND>
hipError_t launch_fill_e(int eqop, int ncomp, const qla_fill_args &p, hipStream_t st)
{
    if (ncomp == 0) {
        if constexpr (KIND == FILL_DIAG) {
            return hipErrorInvalidValue;
        } else {
            switch (eqop) {
            case 0: return launch_rfill<T, KIND, 0>(p, st);
            case 1: return launch_rfill<T, KIND, 1>(p, st);
            case 2: return launch_rfill<T, KIND, 2>(p, st);
            case 3: return launch_rfill<T, KIND, 3>(p, st);
            }
            return hipErrorInvalidValue;
        }
    }
    switch (eqop) {
    case 0: return launch_fill_k<T, KIND, 0>(ncomp, p, st);
    case 1: return launch_fill_k<T, KIND, 1>(ncomp, p, st);
    case 2: return launch_fill_k<T, KIND, 2>(ncomp, p, st);
    case 3: return launch_fill_k<T, KIND, 3>(ncomp, p, st);
    }
    return hipErrorInvalidValue;
}

template <class T>
hipError_t launch_fill_t(int kind, int eqop, int ncomp, const qla_fill_args &p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    if (kind == FILL_ZERO) {
        if (eqop != 0) return hipErrorInvalidValue;
        if (ncomp == 0) return launch_rfill<T, FILL_ZERO, 0>(p, st);
        // the 16-byte stores of k_vec_zero need a 16-byte aligned float
        // field; an 8-byte aligned one takes the per-complex fill
        if (!p.index && (sizeof(T) == 8 || (uintptr_t)p.r % 16 == 0)) {
            const long long m = (long long)p.n * ncomp;
            const long long items = sizeof(T) == 4 ? (m + 1) / 2 : m;
            hipLaunchKernelGGL(k_vec_zero<T>, dim3(blocks_for(items)), dim3(VBLOCK), 0, st, (cpx<T> *)p.r, m);
            return hipGetLastError();
        }
        return launch_fill_k<T, FILL_ZERO, 0>(ncomp, p, st);
    }
    if (kind == FILL_REC) return launch_fill_e<T, FILL_REC>(eqop, ncomp, p, st);
    if (kind == FILL_DIAG) return launch_fill_e<T, FILL_DIAG>(eqop, ncomp, p, st);
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Multiplication by gamma matrices (build_qla.pl:1356-1415, eq only).  Every
// row and column of a product of gamma matrices holds one entry i^k, so
// Gamma_g a is a signed permutation of spin components: destination spin s
// takes the source spin src[s] times i^ph[s], element by element with
// QLA_c_eq_c (k = 0), QLA_c_eq_ic (1), QLA_c_eqm_c (2) and QLA_c_eqm_ic (3)
// (expressions_gamma.pl:387-441, carith1 of headers.pl:135-170).
// ---------------------------------------------------------------------------

// gamma_mu (mu = 0..3) as rows: row r has the entry i^ph at column col.  The
// chiral basis of the reference (expressions_gamma.pl:27-92, doc/qla.tex):
//   g0 = [[0,0,0,i],[0,0,i,0],[0,-i,0,0],[-i,0,0,0]]
//   g1 = [[0,0,0,-1],[0,0,1,0],[0,1,0,0],[-1,0,0,0]]
//   g2 = [[0,0,i,0],[0,0,0,-i],[-i,0,0,0],[0,i,0,0]]
//   g3 = [[0,0,1,0],[0,0,0,1],[1,0,0,0],[0,1,0,0]]
struct gamma_row { int col, ph; };
constexpr gamma_row GAMMA_ROWS[4][4] = {
    { { 3, 1 }, { 2, 1 }, { 1, 3 }, { 0, 3 } },
    { { 3, 2 }, { 2, 0 }, { 1, 0 }, { 0, 2 } },
    { { 2, 1 }, { 3, 3 }, { 0, 3 }, { 1, 1 } },
    { { 2, 0 }, { 3, 0 }, { 0, 0 }, { 1, 0 } },
};

struct gamma_tab { int src[4], ph[4]; };

// Gamma_g = product over the set bits i = 0..3 of gamma_i, left to right
// (expressions_gamma.pl:123-134); left: the source spin and phase of each
// destination row; right: of each destination column.
gamma_tab gamma_table(int g, int right)
{
    gamma_row m[4] = { { 0, 0 }, { 1, 0 }, { 2, 0 }, { 3, 0 } };
    for (int i = 0; i < 4; i++) {
        if (!(g & (1 << i))) continue;
        for (int r = 0; r < 4; r++) {
            const gamma_row b = GAMMA_ROWS[i][m[r].col];
            m[r] = { b.col, (m[r].ph + b.ph) & 3 };
        }
    }
    gamma_tab t;
    for (int r = 0; r < 4; r++) {
        if (!right) {
            t.src[r] = m[r].col;
            t.ph[r] = m[r].ph;
        } else {
            t.src[m[r].col] = r;
            t.ph[m[r].col] = m[r].ph;
        }
    }
    return t;
}

// one work item per destination complex; SPOS: position of the multiplied
// spin index in the record ((colour, spin) for D; (ic, is, jc, js) for P,
// 1 from the left, 3 from the right)
template <class T, int NC, int SPOS>
__global__ void __launch_bounds__(VBLOCK) k_vec_gamma(qla_vec_args p, gamma_tab tab)
{
    typedef typename vtype<T, 1>::v V1;
    const unsigned long long tot = (unsigned long long)p.n * NC;
    const unsigned long long stride = (unsigned long long)gridDim.x * VBLOCK;
    // element stride of the spin index at SPOS
    constexpr int SSTRIDE = NC == 12 ? 1 : (SPOS == 1 ? 12 : 1);
    for (unsigned long long g = (unsigned long long)blockIdx.x * VBLOCK + threadIdx.x; g < tot; g += stride) {
        int i, c;
        if (p.wide) {
            i = (int)(g / NC);
            c = (int)(g - (unsigned long long)i * NC);
        } else {
            const unsigned int g32 = (unsigned int)g;
            i = (int)(g32 / NC);
            c = (int)(g32 - (unsigned int)i * NC);
        }
        const int s = p.index ? p.index[i] : i;
        const int sd = p.ir ? p.ir[i] : s;
        const cpx<T> *X = record_of<T, NC>(p.x, p.px, p.ix ? p.ix[i] : s);
        const int spin = (c / SSTRIDE) & 3;
        const int cx = c + (tab.src[spin] - spin) * SSTRIDE;
        const V1 x = __builtin_nontemporal_load((const V1 *)(X + cx));
        V1 r;
        switch (tab.ph[spin]) {
        case 0: r[0] = x[0]; r[1] = x[1]; break;         // QLA_c_eq_c
        case 1: r[0] = -x[1]; r[1] = x[0]; break;        // QLA_c_eq_ic
        case 2: r[0] = -x[0]; r[1] = -x[1]; break;       // QLA_c_eqm_c
        default: r[0] = x[1]; r[1] = -x[0]; break;       // QLA_c_eqm_ic
        }
        __builtin_nontemporal_store(r, (V1 *)((cpx<T> *)p.r + (size_t)sd * NC + c));
    }
}

// v-form over 16-byte aligned single-precision fields: a work item takes the
// four spin rows one destination group needs, in 16-byte loads of LW = 2
// complexes, and permutes them in registers (F3 D: 5.3 TB/s at 32^4, against
// 4.6-4.9 for one complex per work item; two groups per work item measured
// 0.96x, four 0.92x).
// SSTRIDE 1 (D, or P times Gamma from the right): the group is the 4 adjacent
// spin components; SSTRIDE 12 (Gamma times P): LW adjacent complexes in each
// of the 4 spin rows, 12 complexes apart in a 48-complex (ic, is) block.
template <class T, int SSTRIDE, int LW>
__global__ void __launch_bounds__(VBLOCK) k_vec_gamma_flat(cpx<T> *r, const cpx<T> *x, long long items, gamma_tab tab)
{
    typedef typename vtype<T, LW>::v VL;
    constexpr int WW = SSTRIDE == 1 ? 1 : LW;       // complexes of one spin row in an item
    constexpr int NCH = 4 * WW / LW;                // accesses of LW complexes per item
    constexpr int IPB = 12 / LW;                    // items per 48-complex block (SSTRIDE 12)
    const long long stride = (long long)gridDim.x * VBLOCK;
    for (long long g = (long long)blockIdx.x * VBLOCK + threadIdx.x; g < items; g += stride) {
        const long long e = SSTRIDE == 1 ? g * 4 : (g / IPB) * 48 + (g % IPB) * LW;
        T re[4 * WW], im[4 * WW];
#pragma unroll
        for (int j = 0; j < NCH; j++) {
            const VL v = __builtin_nontemporal_load((const VL *)(x + e + (SSTRIDE == 1 ? j * LW : j * 12)));
#pragma unroll
            for (int w = 0; w < LW; w++) {
                re[j * LW + w] = v[2 * w];
                im[j * LW + w] = v[2 * w + 1];
            }
        }
        T ore[4 * WW], oim[4 * WW];
#pragma unroll
        for (int sp = 0; sp < 4; sp++) {
            const int src = tab.src[sp], ph = tab.ph[sp];
#pragma unroll
            for (int w = 0; w < WW; w++) {
                const T a = src == 0 ? re[w] : src == 1 ? re[WW + w] : src == 2 ? re[2 * WW + w] : re[3 * WW + w];
                const T b = src == 0 ? im[w] : src == 1 ? im[WW + w] : src == 2 ? im[2 * WW + w] : im[3 * WW + w];
                T orr, oi;
                if (ph == 0) {          // QLA_c_eq_c
                    orr = a;
                    oi = b;
                } else if (ph == 1) {   // QLA_c_eq_ic
                    orr = -b;
                    oi = a;
                } else if (ph == 2) {   // QLA_c_eqm_c
                    orr = -a;
                    oi = -b;
                } else {                // QLA_c_eqm_ic
                    orr = b;
                    oi = -a;
                }
                ore[sp * WW + w] = orr;
                oim[sp * WW + w] = oi;
            }
        }
#pragma unroll
        for (int j = 0; j < NCH; j++) {
            VL v;
#pragma unroll
            for (int w = 0; w < LW; w++) {
                v[2 * w] = ore[j * LW + w];
                v[2 * w + 1] = oim[j * LW + w];
            }
            __builtin_nontemporal_store(v, (VL *)(r + e + (SSTRIDE == 1 ? j * LW : j * 12)));
        }
    }
}

// One complex per work item, each reading its source complex: every wave
// access is contiguous.  Double fields (whose 16-byte complexes would put a
// group's four accesses 64 bytes apart across lanes: 2.8 TB/s against 6.3 for
// this form at 32^4) and single-precision fields off 16-byte alignment.
template <class T, int SSTRIDE>
__global__ void __launch_bounds__(VBLOCK) k_vec_gamma_flat1(cpx<T> *r, const cpx<T> *x, long long m, gamma_tab tab)
{
    typedef typename vtype<T, 1>::v V1;
    const long long stride = (long long)gridDim.x * VBLOCK;
    for (long long g = (long long)blockIdx.x * VBLOCK + threadIdx.x; g < m; g += stride) {
        const int spin = (int)((g / SSTRIDE) & 3);
        const int src = spin == 0 ? tab.src[0] : spin == 1 ? tab.src[1] : spin == 2 ? tab.src[2] : tab.src[3];
        const int ph = spin == 0 ? tab.ph[0] : spin == 1 ? tab.ph[1] : spin == 2 ? tab.ph[2] : tab.ph[3];
        const V1 v = __builtin_nontemporal_load((const V1 *)(x + g + (long long)(src - spin) * SSTRIDE));
        V1 o;
        switch (ph) {
        case 0: o[0] = v[0]; o[1] = v[1]; break;        // QLA_c_eq_c
        case 1: o[0] = -v[1]; o[1] = v[0]; break;       // QLA_c_eq_ic
        case 2: o[0] = -v[0]; o[1] = -v[1]; break;      // QLA_c_eqm_c
        default: o[0] = v[1]; o[1] = -v[0]; break;      // QLA_c_eqm_ic
        }
        __builtin_nontemporal_store(o, (V1 *)(r + g));
    }
}

template <class T>
hipError_t launch_gamma_t(int right, int g, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    const gamma_tab tab = gamma_table(g, right);
    if (!p.index && !p.ir && !p.ix && !p.px) {
        if (ncomp != 12 && ncomp != 144) return hipErrorInvalidValue;
        // m is a multiple of 12 (D) or 144 (P)
        const long long m = (long long)p.n * ncomp;
        cpx<T> *r = (cpx<T> *)p.r;
        const cpx<T> *x = (const cpx<T> *)p.x;
        const bool w2 = sizeof(T) == 4 && p.al;
        if (!right && ncomp == 144) {
            if (w2)
                hipLaunchKernelGGL((k_vec_gamma_flat<T, 12, sizeof(T) == 4 ? 2 : 1>), dim3(blocks_for(m / 8)),
                                   dim3(VBLOCK), 0, st, r, x, m / 8, tab);
            else
                hipLaunchKernelGGL((k_vec_gamma_flat1<T, 12>), dim3(blocks_for(m)), dim3(VBLOCK), 0, st, r, x, m, tab);
        } else {
            if (w2)
                hipLaunchKernelGGL((k_vec_gamma_flat<T, 1, sizeof(T) == 4 ? 2 : 1>), dim3(blocks_for(m / 4)),
                                   dim3(VBLOCK), 0, st, r, x, m / 4, tab);
            else
                hipLaunchKernelGGL((k_vec_gamma_flat1<T, 1>), dim3(blocks_for(m)), dim3(VBLOCK), 0, st, r, x, m, tab);
        }
        return hipGetLastError();
    }
    qla_vec_args q = p;
    const long long tot = (long long)p.n * ncomp;
    q.wide = tot > 0xffffffffLL ? 1 : 0;
    const dim3 grid(blocks_for(tot));
    if (!right && ncomp == 12) hipLaunchKernelGGL((k_vec_gamma<T, 12, 1>), grid, dim3(VBLOCK), 0, st, q, tab);
    else if (!right && ncomp == 144) hipLaunchKernelGGL((k_vec_gamma<T, 144, 1>), grid, dim3(VBLOCK), 0, st, q, tab);
    else if (right && ncomp == 144) hipLaunchKernelGGL((k_vec_gamma<T, 144, 3>), grid, dim3(VBLOCK), 0, st, q, tab);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Site-local reductions and colour-matrix parts.  One work item per
// destination record; each sums its record's components in order, as the
// generated code does (a sequential sum fixes the rounding):
//  * norm2 / dot / Re dot (make_code_binary.pl make_code_binary_dot): the sum
//    runs in double with the operands promoted (F and D both promote to D,
//    datatypes.pl:355-359), x += a.re*b.re + a.im*b.im per element (the dot
//    adds x.im += a.re*b.im - a.im*b.re); a real destination then takes C's
//    r op= x in double, a complex one QLA_c_<eqop>_c with x converted
//    (expressions_scalar.pl:36-80);
//  * traces: x = 0, x += the diagonal in order, in the field precision;
//  * the traceless anti-Hermitian part (expressions_tensor.pl:138-232):
//    x = Im tr / 3; D_ii = (0, Im S_ii - x); for i < j, t = (S_ij - conj S_ji)
//    * 0.5, D_ij = t, D_ji = -conj t.
// ---------------------------------------------------------------------------
enum { L_NORM2 = 0, L_DOT, L_REDOT, L_RETRACE, L_IMTRACE, L_TRACE, L_ANTIHERM };

// One site of the site-local reductions and parts: A and B are the site's
// records (B unused but for the dots), sd its destination step; reductions
// and traces write p.r[sd], the anti-Hermitian part writes its 9 complexes
// to `out`.
template <class T, int NC, int KIND, int EQOP>
__device__ __forceinline__ void local_site(const qla_vec_args &p, const cpx<T> *A, const cpx<T> *B, long long sd,
                                           cpx<T> *out)
{
    if constexpr (KIND <= L_REDOT) {
        double xr = 0, xi = 0;
#pragma unroll 4
        for (int e = 0; e < NC; e++) {
            const cpx<T> a = A[e], b = B[e];
            const double ar = a.re, ai = a.im, br = b.re, bi = b.im;
            xr = xr + (ar * br + ai * bi);
            if constexpr (KIND == L_DOT) xi = xi + (ar * bi - ai * br);
        }
        if constexpr (KIND != L_DOT) {
            T *R = (T *)p.r + sd;
            if constexpr (EQOP == 0) *R = (T)xr;
            else if constexpr (EQOP == 1) *R = (T)((double)*R + xr);
            else if constexpr (EQOP == 2) *R = (T)((double)*R - xr);
            else *R = (T)(-xr);
        } else {
            cpx<T> *R = (cpx<T> *)p.r + sd;
            T rr = T(0), ri = T(0);
            if constexpr (reads_r<EQOP>()) {
                rr = R->re;
                ri = R->im;
            }
            vop<VOP_COPY, EQOP, T>(rr, ri, (T)xr, (T)xi, T(0), T(0), T(0), T(0));
            R->re = rr;
            R->im = ri;
        }
    } else if constexpr (KIND <= L_TRACE) {
        T xr = T(0), xi = T(0);
#pragma unroll
        for (int ic = 0; ic < 3; ic++) {
            const cpx<T> a = A[4 * ic];
            if constexpr (KIND == L_RETRACE) xr = xr + a.re;
            else if constexpr (KIND == L_IMTRACE) xr = xr + a.im;
            else {
                xr = xr + a.re;
                xi = xi + a.im;
            }
        }
        if constexpr (KIND == L_TRACE) {
            cpx<T> *R = (cpx<T> *)p.r + sd;
            R->re = xr;
            R->im = xi;
        } else {
            ((T *)p.r)[sd] = xr;
        }
    } else {
        cpx<T> a[9], d[9];
#pragma unroll
        for (int e = 0; e < 9; e++) a[e] = A[e];
        T x = T(0);
#pragma unroll
        for (int ic = 0; ic < 3; ic++) x = x + a[4 * ic].im;
        x = x / T(3);
#pragma unroll
        for (int ic = 0; ic < 3; ic++) {
            d[4 * ic].re = T(0);
            d[4 * ic].im = a[4 * ic].im - x;
        }
#pragma unroll
        for (int ic = 0; ic < 2; ic++)
#pragma unroll
            for (int jc = ic + 1; jc < 3; jc++) {
                const cpx<T> u = a[3 * ic + jc], v = a[3 * jc + ic];
                const T tr = u.re - v.re, ti = u.im + v.im;
                const T hr = T(0.5) * tr, hi = T(0.5) * ti;
                d[3 * ic + jc].re = hr;
                d[3 * ic + jc].im = hi;
                d[3 * jc + ic].re = -hr;
                d[3 * jc + ic].im = hi;
            }
#pragma unroll
        for (int e = 0; e < 9; e++) out[e] = d[e];
    }
}

template <class T, int NC, int KIND, int EQOP>
__global__ void __launch_bounds__(VBLOCK) k_vec_local(qla_vec_args p)
{
    const long long stride = (long long)gridDim.x * VBLOCK;
    for (long long i = (long long)blockIdx.x * VBLOCK + threadIdx.x; i < p.n; i += stride) {
        const int s = p.index ? p.index[i] : (int)i;
        const int sd = p.ir ? p.ir[i] : s;
        const cpx<T> *A = record_of<T, NC>(p.x, p.px, p.ix ? p.ix[i] : s);
        const cpx<T> *B = (KIND == L_DOT || KIND == L_REDOT) ? record_of<T, NC>(p.y, p.py, p.iy ? p.iy[i] : s) : A;
        local_site<T, NC, KIND, EQOP>(p, A, B, sd, (cpx<T> *)p.r + (size_t)sd * 9);
    }
}

// The v-form over contiguous fields (no index, no pointer operand) with 6 to
// 12 complexes a record: a block stages RB whole records of each operand in
// LDS with coalesced non-temporal loads (16 bytes a lane: two single
// complexes when the fields are 16-byte aligned, one double complex), each
// work item evaluates one site from LDS, and the anti-Hermitian part's
// records leave through the same LDS tile with coalesced stores. Read one
// strided record per lane straight from memory, a wave instruction touches
// one cache line per lane and a record's lines are fetched again on each
// visit; at 32^4 sites the staged form runs 1.28-1.61x faster, bit for bit
// the same. A 3-complex record (24 or 48 bytes) is read per lane.
template <class T> struct local_stage_shape {
    static constexpr int RB = sizeof(T) == 4 ? 256 : 128;     // records per block step
};

template <class T, int NC, int KIND, int EQOP, int W>
__global__ void __launch_bounds__(VBLOCK) k_vec_local_stage(qla_vec_args p)
{
    // W complexes per lane access: 2 (16 bytes) for 16-byte aligned single
    // precision fields, else 1
    typedef typename vtype<T, 1>::v V1;
    typedef typename vtype<T, W>::v VW;
    constexpr int RB = local_stage_shape<T>::RB;
    constexpr bool TWO = KIND == L_DOT || KIND == L_REDOT;
    __shared__ VW tx[RB * NC / W];
    __shared__ VW ty[TWO ? RB * NC / W : 1];
    const VW *X = (const VW *)p.x, *Y = (const VW *)p.y;
    for (long long base = (long long)blockIdx.x * RB; base < p.n; base += (long long)gridDim.x * RB) {
        const int nrec = (int)(p.n - base < RB ? p.n - base : RB);
        const int tot = nrec * NC;
        // base is a multiple of RB, so e0 is a multiple of W
        const long long e0 = base * NC / W;
        for (int k = threadIdx.x; k < tot / W; k += VBLOCK) {
            tx[k] = __builtin_nontemporal_load(X + e0 + k);
            if constexpr (TWO) ty[k] = __builtin_nontemporal_load(Y + e0 + k);
        }
        if constexpr (W == 2) {
            // an odd last complex (odd NC and nrec)
            if ((tot & 1) && threadIdx.x == 0) {
                ((V1 *)tx)[tot - 1] = __builtin_nontemporal_load((const V1 *)p.x + base * NC + tot - 1);
                if constexpr (TWO) ((V1 *)ty)[tot - 1] = __builtin_nontemporal_load((const V1 *)p.y + base * NC + tot - 1);
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < nrec; t += VBLOCK) {
            cpx<T> *A = (cpx<T> *)tx + t * NC;
            const cpx<T> *B = TWO ? (const cpx<T> *)ty + t * NC : A;
            // the anti-Hermitian part overwrites the site's own record in the
            // tile: only this work item reads it
            local_site<T, NC, KIND, EQOP>(p, A, B, base + t, A);
        }
        __syncthreads();
        if constexpr (KIND == L_ANTIHERM) {
            VW *R = (VW *)p.r;
            for (int k = threadIdx.x; k < tot / W; k += VBLOCK) __builtin_nontemporal_store(tx[k], R + e0 + k);
            if constexpr (W == 2) {
                if ((tot & 1) && threadIdx.x == 0)
                    __builtin_nontemporal_store(((V1 *)tx)[tot - 1], (V1 *)p.r + base * NC + tot - 1);
            }
            __syncthreads();
        }
    }
}

template <class T, int NC, int KIND, int EQOP>
hipError_t launch_local(const qla_vec_args &p, hipStream_t st)
{
    const bool flat = !p.index && !p.ir && !p.ix && !p.iy && !p.px && !p.py;
    if constexpr (NC >= 6 && NC <= 12) {
        // double records are staged with 16-byte loads: aligned fields only
        if (flat && (sizeof(T) == 4 || p.al)) {
            constexpr int RB = local_stage_shape<T>::RB;
            const long long b = (p.n + RB - 1) / RB;
            const dim3 grid((unsigned)(b > VMAXBLOCKS ? VMAXBLOCKS : b));
            if (sizeof(T) == 4 && p.al)
                hipLaunchKernelGGL((k_vec_local_stage<T, NC, KIND, EQOP, sizeof(T) == 4 ? 2 : 1>), grid, dim3(VBLOCK),
                                   0, st, p);
            else
                hipLaunchKernelGGL((k_vec_local_stage<T, NC, KIND, EQOP, 1>), grid, dim3(VBLOCK), 0, st, p);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_vec_local<T, NC, KIND, EQOP>), dim3(blocks_for(p.n)), dim3(VBLOCK), 0, st, p);
    return hipGetLastError();
}

template <class T, int KIND, int EQOP>
hipError_t launch_local_nc(int ncomp, const qla_vec_args &p, hipStream_t st)
{
#define QLA_LOCAL_CASE(NC)                                                                      \
    case NC: return launch_local<T, NC, KIND, EQOP>(p, st);
    if constexpr (KIND >= L_RETRACE) {
        if (ncomp != 9) return hipErrorInvalidValue;
        return launch_local<T, 9, KIND, EQOP>(p, st);
    } else {
        switch (ncomp) {
            QLA_LOCAL_CASE(3)
            QLA_LOCAL_CASE(6)
            QLA_LOCAL_CASE(9)
            QLA_LOCAL_CASE(12)
            QLA_LOCAL_CASE(144)
        }
        return hipErrorInvalidValue;
    }
#undef QLA_LOCAL_CASE
}

template <class T, int KIND>
hipError_t launch_local_e(int eqop, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    if constexpr (KIND >= L_RETRACE) {
        return eqop == 0 ? launch_local_nc<T, KIND, 0>(ncomp, p, st) : hipErrorInvalidValue;
    } else {
        switch (eqop) {
        case 0: return launch_local_nc<T, KIND, 0>(ncomp, p, st);
        case 1: return launch_local_nc<T, KIND, 1>(ncomp, p, st);
        case 2: return launch_local_nc<T, KIND, 2>(ncomp, p, st);
        case 3: return launch_local_nc<T, KIND, 3>(ncomp, p, st);
        }
        return hipErrorInvalidValue;
    }
}

template <class T>
hipError_t launch_local_t(int kind, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    if (p.n <= 0) return hipSuccess;
    switch (kind) {
    case L_NORM2: return launch_local_e<T, L_NORM2>(eqop, ncomp, p, st);
    case L_DOT: return launch_local_e<T, L_DOT>(eqop, ncomp, p, st);
    case L_REDOT: return launch_local_e<T, L_REDOT>(eqop, ncomp, p, st);
    case L_RETRACE: return launch_local_e<T, L_RETRACE>(eqop, ncomp, p, st);
    case L_IMTRACE: return launch_local_e<T, L_IMTRACE>(eqop, ncomp, p, st);
    case L_TRACE: return launch_local_e<T, L_TRACE>(eqop, ncomp, p, st);
    case L_ANTIHERM: return launch_local_e<T, L_ANTIHERM>(eqop, ncomp, p, st);
    }
    return hipErrorInvalidValue;
}

} // namespace
} // namespace qla

#ifndef QLA_VEC_PREC
#define QLA_VEC_PREC 0
#endif

#if QLA_VEC_PREC == 0
hipError_t qla_vec_launch_p0(int op, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_t<float>(op, eqop, ncomp, p, st);
}
hipError_t qla_fill_launch_p0(int kind, int eqop, int ncomp, const qla_fill_args &p, hipStream_t st)
{
    return qla::launch_fill_t<float>(kind, eqop, ncomp, p, st);
}
hipError_t qla_gamma_launch_p0(int right, int g, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_gamma_t<float>(right, g, ncomp, p, st);
}
hipError_t qla_local_launch_p0(int kind, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_local_t<float>(kind, eqop, ncomp, p, st);
}
hipError_t qla_outer_launch_p0(int eqop, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_outer_t<float>(eqop, p, st);
}
hipError_t qla_pmul_launch_p0(int kind, int eqop, int ca, int cb, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_pmul_t<float>(kind, eqop, ca, cb, p, st);
}
hipError_t qla_comp_launch_p0(int mode, const qla_vec_args &p, const qla_comp_map &m, hipStream_t st)
{
    return qla::launch_comp_t<float>(mode, p, m, st);
}
hipError_t qla_sconv_launch_p0(int kind, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_sconv_t<float>(kind, p, st);
}
hipError_t qla_int_launch(int op, int eqop, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_int(op, eqop, p, st);
}
hipError_t qla_sfn_launch_p0(int fn, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_sfn_t<float>(fn, p, st);
}
hipError_t qla_mask_launch_p0(int nw, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_mask_t<float>(nw, p, st);
}
hipError_t qla_cmp_launch_p0(int op, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_cmp_t<float>(op, p, st);
}
#else
hipError_t qla_vec_launch_p1(int op, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_t<double>(op, eqop, ncomp, p, st);
}
hipError_t qla_fill_launch_p1(int kind, int eqop, int ncomp, const qla_fill_args &p, hipStream_t st)
{
    return qla::launch_fill_t<double>(kind, eqop, ncomp, p, st);
}
hipError_t qla_gamma_launch_p1(int right, int g, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_gamma_t<double>(right, g, ncomp, p, st);
}
hipError_t qla_local_launch_p1(int kind, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_local_t<double>(kind, eqop, ncomp, p, st);
}
hipError_t qla_outer_launch_p1(int eqop, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_outer_t<double>(eqop, p, st);
}
hipError_t qla_pmul_launch_p1(int kind, int eqop, int ca, int cb, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_pmul_t<double>(kind, eqop, ca, cb, p, st);
}
hipError_t qla_comp_launch_p1(int mode, const qla_vec_args &p, const qla_comp_map &m, hipStream_t st)
{
    return qla::launch_comp_t<double>(mode, p, m, st);
}
hipError_t qla_sconv_launch_p1(int kind, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_sconv_t<double>(kind, p, st);
}
hipError_t qla_sfn_launch_p1(int fn, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_sfn_t<double>(fn, p, st);
}
hipError_t qla_mask_launch_p1(int nw, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_mask_t<double>(nw, p, st);
}
hipError_t qla_cmp_launch_p1(int op, const qla_vec_args &p, hipStream_t st)
{
    return qla::launch_cmp_t<double>(op, p, st);
}
#endif
