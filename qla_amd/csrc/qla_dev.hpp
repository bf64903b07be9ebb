// qla_dev.hpp -- device building blocks for the gfx950 site-loop kernels.
//
// Layout facts these helpers rely on (reference include/qla_types.h:29-33):
// every site record is an array of {re, im} complex numbers, site-major, with
// no padding, so a field of n records is n*BPS contiguous bytes and a pointer
// array holds the start address of each record.
//
// Data movement model.  A wavefront owns a tile of 64 sites.  Whole records
// are moved between HBM and LDS with gfx950 direct-to-LDS loads
// (global_load_lds_dword{,x3,x4}): consecutive lanes of one instruction read
// consecutive chunks of the same or the next record -- coalesced for
// contiguous fields and for pointer gathers whose targets are adjacent.
// After the loads each lane reads its own site's record out of LDS (the
// transpose), computes in registers, writes the result record to LDS and the
// wave stores it with the same chunk mapping, again coalesced.
//
// Arithmetic.  Everything is compiled with -ffp-contract=off so each complex
// multiply-accumulate rounds exactly like the reference macro it restates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qla {

constexpr int WAVE = 64;

template <class T> struct cpx { T re, im; };

typedef __attribute__((address_space(3))) void *lds_void_ptr;

// Wave-scope barrier: orders LDS accesses of the lanes of one wavefront
// (the direct-to-LDS loads are waited for by the compiler's vmcnt tracking).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Tiles of WAVE sites over n >= 0 sites, formed so that nothing overflows
// int for any n up to INT_MAX: a tile below tiles_of(n) starts at
// tile * WAVE <= n - 1, and lane_site forms a lane's site only when it is
// below n (the tile's first site otherwise).
__device__ __forceinline__ int tiles_of(int n) { return n / WAVE + (n % WAVE != 0 ? 1 : 0); }
__device__ __forceinline__ int lane_site(int tile, int lane, int n, bool &act)
{
    const int base = tile * WAVE;
    act = lane < n - base;
    return act ? base + lane : base;
}

// Wait until every vector-memory operation of this wave has completed
// (s_waitcnt vmcnt(0): simm16 0x0F70 on gfx9).  Issued through the builtin so
// the compiler's wait-count pass sees it: loads issued before it count as
// complete afterwards, and the pass inserts no second full wait of its own
// when their destination registers are read or reused later (for example a
// loop-carried address register loaded in a prologue).
__device__ __forceinline__ void vm_wait() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// 64-bit value of lane `src` (ds_bpermute on both halves).
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src)
{
    int lo = (int)(uint32_t)v, hi = (int)(uint32_t)(v >> 32);
    lo = __builtin_amdgcn_ds_bpermute(src << 2, lo);
    hi = __builtin_amdgcn_ds_bpermute(src << 2, hi);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// ---------------------------------------------------------------------------
// HBM -> LDS staging.  CH is the chunk size in bytes moved per lane by one
// direct-to-LDS load: 4, 12 or 16 (gfx950 global_load_lds_dword{,x3,x4}).
// Lane l's chunk lands at base + l*LS(CH) where the LDS lane stride LS is 4
// for dword and 16 for both dwordx3 and dwordx4 (a 12-byte chunk leaves a
// 4-byte hole).  A record of BPS bytes = NPS chunks therefore occupies
// NPS*LS bytes of LDS; chunk c of the tile sits at c*LS.  Records must be
// 4-byte aligned for CH = 4 and 12 and 16-byte aligned for CH = 16.
// ---------------------------------------------------------------------------
template <int CH> struct lds_stride { static constexpr int v = (CH == 4) ? 4 : 16; };
template <int BPS, int CH> struct rec_lds { static constexpr int v = (BPS / CH) * lds_stride<CH>::v; };

// Cache policy of the streaming loads and stores: every field record is
// touched once per call, so loads and stores carry the non-temporal hint
// (the "nt" bit, cache-policy value 2 on gfx950).
#ifndef QLA_STREAM_CPOL
#define QLA_STREAM_CPOL 2
#endif

// One direct-to-LDS load of CH bytes per lane (the builtin needs a literal width).
template <int CH> __device__ __forceinline__ void dma_load(const void *g, char *lds_base);
template <> __device__ __forceinline__ void dma_load<4>(const void *g, char *lds_base)
{
    __builtin_amdgcn_global_load_lds(g, (lds_void_ptr)lds_base, 4, 0, QLA_STREAM_CPOL);
}
template <> __device__ __forceinline__ void dma_load<12>(const void *g, char *lds_base)
{
    __builtin_amdgcn_global_load_lds(g, (lds_void_ptr)lds_base, 12, 0, QLA_STREAM_CPOL);
}
template <> __device__ __forceinline__ void dma_load<16>(const void *g, char *lds_base)
{
    __builtin_amdgcn_global_load_lds(g, (lds_void_ptr)lds_base, 16, 0, QLA_STREAM_CPOL);
}

// Contiguous records g[0 .. nsites) -> lds (site s at s*rec_lds).
template <int BPS, int CH>
__device__ __forceinline__ void stage_contig(char *lds, const char *g, int nsites, int lane)
{
    static_assert(BPS % CH == 0, "chunk must divide the record");
    constexpr int NPS = BPS / CH, LS = lds_stride<CH>::v;
    const int nch = nsites * NPS;
#pragma unroll
    for (int k = 0; k < NPS; k++) {
        const int c = k * WAVE + lane;
        if (c < nch) dma_load<CH>(g + (size_t)c * CH, lds + k * WAVE * LS);
    }
}

// Records at per-site addresses: lane s holds the address of site s's record
// (sites >= nsites are skipped) -> lds (site s at s*rec_lds).
template <int BPS, int CH>
__device__ __forceinline__ void stage_gather(char *lds, uint64_t my_addr, int nsites, int lane)
{
    static_assert(BPS % CH == 0, "chunk must divide the record");
    constexpr int NPS = BPS / CH, LS = lds_stride<CH>::v;
    if constexpr (CH == 4) {
        // 4-byte chunks (the misaligned fallback): a rolled loop, to keep the
        // fast path's register budget
#pragma unroll 1
        for (int k = 0; k < NPS; k++) {
            const int c = k * WAVE + lane;
            const int s = c / NPS;
            const int p = c - s * NPS;
            const uint64_t a = shfl64(my_addr, s);
            if (s < nsites) dma_load<CH>((const void *)(a + (uint64_t)p * CH), lds + k * WAVE * LS);
        }
    } else {
        // wide chunks: every chunk address is fetched from its record's lane
        // before the first load is issued, so the shuffles overlap instead of
        // each load waiting for its own
        uint64_t a[NPS];
#pragma unroll
        for (int k = 0; k < NPS; k++) {
            const int c = k * WAVE + lane;
            const int s = c / NPS;
            a[k] = shfl64(my_addr, s) + (uint64_t)(c - s * NPS) * CH;     // s < WAVE: c < NPS * WAVE
        }
#pragma unroll
        for (int k = 0; k < NPS; k++) {
            const int s = (k * WAVE + lane) / NPS;
            if (s < nsites) dma_load<CH>((const void *)a[k], lds + k * WAVE * LS);
        }
    }
}

// A contiguous byte range g[0 .. nbytes) -> lds[0 .. nbytes) (nbytes % 4 == 0,
// g aligned to CH, CH = 4 or 16 so the LDS copy is packed; MAXB bounds nbytes
// at compile time): CH-byte chunks for the bulk, then 4-byte chunks.
template <int CH, int MAXB>
__device__ __forceinline__ void stage_bytes(char *lds, const char *g, int nbytes, int lane)
{
    static_assert(CH == 4 || CH == 16, "packed LDS layout needs CH = 4 or 16");
    const int full = nbytes / CH;
#pragma unroll(CH == 4 ? 1 : (MAXB / CH + WAVE - 1) / WAVE)
    for (int k = 0; k < (MAXB / CH + WAVE - 1) / WAVE; k++) {
        const int c = k * WAVE + lane;
        if (c < full) dma_load<CH>(g + (size_t)c * CH, lds + k * WAVE * CH);
    }
    if (CH > 4) {
        const int done = full * CH;
        const int tail = (nbytes - done) / 4;     // < CH/4 <= WAVE
        if (lane < tail) dma_load<4>(g + done + lane * 4, lds + done);
    }
}

// ---------------------------------------------------------------------------
// Per-lane access to its own record in the LDS chunk layout: element e of
// size ES bytes at record byte b = e*ES lies in chunk b/CH at offset b%CH
// (ES divides CH, so an element never straddles two chunks).
// ---------------------------------------------------------------------------
template <int CH, int ES> struct elem_off {
    static_assert(CH % ES == 0, "element must not straddle chunks");
    static __device__ __forceinline__ constexpr int at(int e)
    {
        return (e * ES / CH) * lds_stride<CH>::v + (e * ES) % CH;
    }
};

template <int CH, class T, int N>
__device__ __forceinline__ void lds_read_rec(cpx<T> (&v)[N], const char *rec)
{
#pragma unroll
    for (int i = 0; i < N; i++) {
        v[i].re = *(const T *)(rec + elem_off<CH, sizeof(T)>::at(2 * i));
        v[i].im = *(const T *)(rec + elem_off<CH, sizeof(T)>::at(2 * i + 1));
    }
}
template <int CH, class T, int N>
__device__ __forceinline__ void lds_write_rec(char *rec, const cpx<T> (&v)[N])
{
#pragma unroll
    for (int i = 0; i < N; i++) {
        *(T *)(rec + elem_off<CH, sizeof(T)>::at(2 * i)) = v[i].re;
        *(T *)(rec + elem_off<CH, sizeof(T)>::at(2 * i + 1)) = v[i].im;
    }
}

// ---------------------------------------------------------------------------
// LDS -> HBM.  Chunk types for the stores.
// ---------------------------------------------------------------------------
typedef uint32_t chunk12 __attribute__((ext_vector_type(3), aligned(4)));
typedef uint32_t chunk16 __attribute__((ext_vector_type(4), aligned(16)));
template <int CH> struct chunk_of;
template <> struct chunk_of<12> { typedef chunk12 type; };
template <> struct chunk_of<16> { typedef chunk16 type; };
template <> struct chunk_of<4> { typedef uint32_t type; };

// store of one chunk with the streaming cache policy
template <class C> __device__ __forceinline__ void stream_store(C *p, C v)
{
#if QLA_STREAM_CPOL == 2
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

template <int BPS, int CH>
__device__ __forceinline__ void store_contig(char *g, const char *lds, int nsites, int lane)
{
    typedef typename chunk_of<CH>::type C;
    constexpr int NPS = BPS / CH, LS = lds_stride<CH>::v;
    const int nch = nsites * NPS;
#pragma unroll
    for (int k = 0; k < NPS; k++) {
        const int c = k * WAVE + lane;
        if (c < nch) stream_store((C *)(g + (size_t)c * CH), *(const C *)(lds + c * LS));
    }
}

// Store of a piece of a record whose cache line the call may write only in
// part: the destination of an x-form (every other record of a parity
// subset) or of an index form.  A write-back store, not a non-temporal one:
// the L2 then merges the pieces of a line before the line goes to HBM,
// where non-temporal stores leave as separate partial-line writes.  On the
// even sites of 32^4 (three input sets rotating between calls) the x-forms
// H_xeq_spproj_M_times_pD and D_xpeq_sprecon_M_times_pH took 40.0 and
// 56.7 us with write-back scatters against 46.4 and 68.6 us with
// non-temporal ones.  Contiguous destinations, written in whole lines, keep
// the non-temporal stores, and so do all loads: write-back loads made the
// x-form slower (52.7 us) and the v-forms 13% slower.
template <class C> __device__ __forceinline__ void scatter_store(C *p, C v) { *p = v; }

template <int BPS, int CH>
__device__ __forceinline__ void store_scatter(uint64_t my_addr, const char *lds, int nsites, int lane)
{
    typedef typename chunk_of<CH>::type C;
    constexpr int NPS = BPS / CH, LS = lds_stride<CH>::v;
#pragma unroll(CH == 4 ? 1 : NPS)
    for (int k = 0; k < NPS; k++) {
        const int c = k * WAVE + lane;
        const int s = c / NPS;
        const int p = c - s * NPS;
        const uint64_t a = shfl64(my_addr, s);
        if (s < nsites) scatter_store((C *)(a + (uint64_t)p * CH), *(const C *)(lds + c * LS));
    }
}

// lds[0 .. nbytes) (packed) -> contiguous g[0 .. nbytes), CH = 4 or 16
template <int CH, int MAXB>
__device__ __forceinline__ void store_bytes(char *g, const char *lds, int nbytes, int lane)
{
    static_assert(CH == 4 || CH == 16, "packed LDS layout needs CH = 4 or 16");
    typedef typename chunk_of<CH>::type C;
    const int full = nbytes / CH;
#pragma unroll(CH == 4 ? 1 : (MAXB / CH + WAVE - 1) / WAVE)
    for (int k = 0; k < (MAXB / CH + WAVE - 1) / WAVE; k++) {
        const int c = k * WAVE + lane;
        if (c < full) stream_store((C *)(g + (size_t)c * CH), *(const C *)(lds + c * CH));
    }
    if (CH > 4) {
        const int done = full * CH;
        const int tail = (nbytes - done) / 4;
        if (lane < tail) stream_store((uint32_t *)(g + done + lane * 4), *(const uint32_t *)(lds + done + lane * 4));
    }
}

// ---------------------------------------------------------------------------
// Complex arithmetic in reference macro order (include/qla_complex.h).
// ---------------------------------------------------------------------------
// x += a*b   (QLA_c_peq_c_times_c, qla_complex.h:330-333)
// x -= a*b   (QLA_c_meq_c_times_c)
// with ADJ: a is conjugated (QLA_c_peq_ca_times_c, qla_complex.h:386-389)
template <bool MEQ, bool ADJ, class T>
__device__ __forceinline__ void acc_mul(cpx<T> &x, cpx<T> a, cpx<T> b)
{
    T tr, ti;
    if (!ADJ) { tr = a.re * b.re - a.im * b.im; ti = a.re * b.im + a.im * b.re; }
    else      { tr = a.re * b.re + a.im * b.im; ti = a.re * b.im - a.im * b.re; }
    if (!MEQ) { x.re = x.re + tr; x.im = x.im + ti; }
    else      { x.re = x.re - tr; x.im = x.im - ti; }
}

// x +=/-= a*b with either factor conjugated (CA, CB), each case as its macro
// writes it: QLA_c_{peq,meq}_c_times_c, _ca_times_c, _c_times_ca (re a.re*b.re
// + a.im*b.im, im a.im*b.re - a.re*b.im) and _ca_times_ca (re a.re*b.re -
// a.im*b.im, im -a.re*b.im - a.im*b.re); qla_complex.h:330-403.
template <bool MEQ, bool CA, bool CB, class T>
__device__ __forceinline__ void acc_mul2(cpx<T> &x, cpx<T> a, cpx<T> b)
{
    T tr, ti;
    if (!CA && !CB) { tr = a.re * b.re - a.im * b.im; ti = a.re * b.im + a.im * b.re; }
    if (CA && !CB)  { tr = a.re * b.re + a.im * b.im; ti = a.re * b.im - a.im * b.re; }
    if (!CA && CB)  { tr = a.re * b.re + a.im * b.im; ti = a.im * b.re - a.re * b.im; }
    if (CA && CB)   { tr = a.re * b.re - a.im * b.im; ti = -a.re * b.im - a.im * b.re; }
    if (!MEQ) { x.re = x.re + tr; x.im = x.im + ti; }
    else      { x.re = x.re - tr; x.im = x.im - ti; }
}

// eqop codes, as in the reference's operatortypes.pl: eq, peq, meq, eqm
enum { EQ = 0, PEQ = 1, MEQ = 2, EQM = 3 };

// r[ic][is] eqop sum_kc A[ic][kc] t[kc][is]; A = a or a^dagger (ADJ).
// Accumulator start and sign follow expressions_tensor.pl:544-630:
// eq: x=0, +=; peq: x=r, +=; meq: x=r, -=; eqm: x=0, -=.
template <int EQOP, bool ADJ, int NS, class T>
__device__ __forceinline__ void mat_times_cols(cpx<T> *r, const cpx<T> *a, const cpx<T> *t)
{
    constexpr bool SUB = (EQOP == MEQ || EQOP == EQM);
#pragma unroll
    for (int ic = 0; ic < 3; ic++) {
#pragma unroll
        for (int is = 0; is < NS; is++) {
            cpx<T> x;
            if (EQOP == PEQ || EQOP == MEQ) x = r[ic * NS + is];
            else { x.re = T(0); x.im = T(0); }
#pragma unroll
            for (int kc = 0; kc < 3; kc++)
                acc_mul<SUB, ADJ>(x, ADJ ? a[kc * 3 + ic] : a[ic * 3 + kc], t[kc * NS + is]);
            r[ic * NS + is] = x;
        }
    }
}

// r[ic][jc] eqop sum_kc A[ic][kc] B[kc][jc] for colour matrices, A = a or
// a^dagger (ADJ), B = b or b^dagger (ADJB: B[kc][jc] = conj(b[jc][kc])), in
// the row-column order of expressions_tensor.pl:544-630.
template <int EQOP, bool ADJ, bool ADJB, class T>
__device__ __forceinline__ void mat_times_mat(cpx<T> *r, const cpx<T> *a, const cpx<T> *b)
{
    constexpr bool SUB = (EQOP == MEQ || EQOP == EQM);
#pragma unroll
    for (int ic = 0; ic < 3; ic++) {
#pragma unroll
        for (int jc = 0; jc < 3; jc++) {
            cpx<T> x;
            if (EQOP == PEQ || EQOP == MEQ) x = r[ic * 3 + jc];
            else { x.re = T(0); x.im = T(0); }
#pragma unroll
            for (int kc = 0; kc < 3; kc++)
                acc_mul2<SUB, ADJ, ADJB>(x, ADJ ? a[kc * 3 + ic] : a[ic * 3 + kc], ADJB ? b[jc * 3 + kc] : b[kc * 3 + jc]);
            r[ic * 3 + jc] = x;
        }
    }
}

// ---------------------------------------------------------------------------
// The association orders of the reference built --enable-sse3 (its
// lib/optimized/sse versions of V_vpeq_M_times_pV, V_veq_M_times_pV,
// V_veq_Ma_times_V, D_vpeq_sprecon_M_times_pH and D_vpeq_spproj_M_times_pD).
// SSE float lanes round like scalar float operations and x + (-y) == x - y,
// so each is a fixed sequence of float operations; ar / ai are the real and
// imaginary parts of the matrix elements of one row (or, for the adjoint,
// one column), x_k the source column with stride XS.
// ---------------------------------------------------------------------------
// vector order of the V bodies (QLA_F3_V_vpeq_M_times_pV_{1,2}_sse.c):
// re = ((r.re + ar0 x0r) + ar1 x1r) + ar2 x2r - ((ai0 x0i + ai1 x1i) + ai2 x2i),
// im = ((r.im + ai0 x0r) + ai1 x1r) + ai2 x2r + ((ar0 x0i + ar1 x1i) + ar2 x2i);
// without r (V_veq) the first product starts each sum (sse_half)
template <class T>
__device__ __forceinline__ cpx<T> sse_pair(cpx<T> r, const cpx<T> *a, const cpx<T> *x)
{
    T R = r.re + a[0].re * x[0].re, I = r.im + a[0].im * x[0].re;
    T Qr = a[0].re * x[0].im, Qi = a[0].im * x[0].im;
    R = R + a[1].re * x[1].re;
    I = I + a[1].im * x[1].re;
    Qr = Qr + a[1].re * x[1].im;
    Qi = Qi + a[1].im * x[1].im;
    R = R + a[2].re * x[2].re;
    I = I + a[2].im * x[2].re;
    Qr = Qr + a[2].re * x[2].im;
    Qi = Qi + a[2].im * x[2].im;
    return { R - Qi, I + Qr };
}
// the multiply of the sprecon / spproj bodies (and V_veq's vector order):
// t.re = ((ar0 x0r + ar1 x1r) + ar2 x2r) - ((ai0 x0i + ai1 x1i) + ai2 x2i),
// t.im = ((ar0 x0i + ar1 x1i) + ar2 x2i) + ((ai0 x0r + ai1 x1r) + ai2 x2r)
template <int XS, class T>
__device__ __forceinline__ cpx<T> sse_half(const cpx<T> *a, const cpx<T> *x)
{
    T R = a[0].re * x[0].re, I = a[0].re * x[0].im;
    T Qr = a[0].im * x[0].re, Qi = a[0].im * x[0].im;
    R = R + a[1].re * x[XS].re;
    I = I + a[1].re * x[XS].im;
    Qr = Qr + a[1].im * x[XS].re;
    Qi = Qi + a[1].im * x[XS].im;
    R = R + a[2].re * x[2 * XS].re;
    I = I + a[2].re * x[2 * XS].im;
    Qr = Qr + a[2].im * x[2 * XS].re;
    Qi = Qi + a[2].im * x[2 * XS].im;
    return { R - Qi, I + Qr };
}
// the unaligned scalar fallback of the sprecon / spproj files:
// t.re = (((((0 + ar0 x0r) - ai0 x0i) + ar1 x1r) - ai1 x1i) + ar2 x2r) - ai2 x2i,
// t.im = (((((0 + ai0 x0r) + ar0 x0i) + ai1 x1r) + ar1 x1i) + ai2 x2r) + ar2 x2i
template <int XS, class T>
__device__ __forceinline__ cpx<T> sse_scalar(const cpx<T> *a, const cpx<T> *x)
{
    T tr = T(0), ti = T(0);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const T xr = x[k * XS].re, xi = x[k * XS].im;
        tr = tr + a[k].re * xr;
        ti = ti + a[k].im * xr;
        tr = tr - a[k].im * xi;
        ti = ti + a[k].re * xi;
    }
    return { tr, ti };
}
// row 2 of the 1-site V bodies, left to right, added to r (vpeq) or stored:
// ((((ar0 x0r - ai0 x0i) + ar1 x1r) - ai1 x1i) + ar2 x2r) - ai2 x2i and
// ((((ar0 x0i + ai0 x0r) + ar1 x1i) + ai1 x1r) + ar2 x2i) + ai2 x2r
template <bool HAVE_R, class T>
__device__ __forceinline__ cpx<T> sse_row2(cpx<T> r, const cpx<T> *a, const cpx<T> *x)
{
    T er = a[0].re * x[0].re;
    er = er - a[0].im * x[0].im;
    er = er + a[1].re * x[1].re;
    er = er - a[1].im * x[1].im;
    er = er + a[2].re * x[2].re;
    er = er - a[2].im * x[2].im;
    T ei = a[0].re * x[0].im;
    ei = ei + a[0].im * x[0].re;
    ei = ei + a[1].re * x[1].im;
    ei = ei + a[1].im * x[1].re;
    ei = ei + a[2].re * x[2].im;
    ei = ei + a[2].im * x[2].re;
    if (HAVE_R) return { r.re + er, r.im + ei };
    return { er, ei };
}
// V_veq_Ma_times_V, a_k = a[k][ic] (column stride 3):
// vector order  re = ((ai0 x0i + ai1 x1i) + ai2 x2i) + ((ar0 x0r + ar1 x1r) + ar2 x2r),
//               im = ((ar0 x0i + ar1 x1i) + ar2 x2i) - ((ai0 x0r + ai1 x1r) + ai2 x2r);
// row 2         re = ((((ar0 x0r + ai0 x0i) + ar1 x1r) + ai1 x1i) + ar2 x2r) + ai2 x2i,
//               im = ((((ar0 x0i - ai0 x0r) + ar1 x1i) - ai1 x1r) + ar2 x2i) - ai2 x2r
template <class T>
__device__ __forceinline__ cpx<T> sse_adj_pair(const cpx<T> *a, const cpx<T> *x)
{
    T Rr = a[0].re * x[0].re, Ri = a[0].im * x[0].re;
    T Ir = a[0].re * x[0].im, Ii = a[0].im * x[0].im;
    Rr = Rr + a[3].re * x[1].re;
    Ri = Ri + a[3].im * x[1].re;
    Ir = Ir + a[3].re * x[1].im;
    Ii = Ii + a[3].im * x[1].im;
    Rr = Rr + a[6].re * x[2].re;
    Ri = Ri + a[6].im * x[2].re;
    Ir = Ir + a[6].re * x[2].im;
    Ii = Ii + a[6].im * x[2].im;
    return { Ii + Rr, Ir - Ri };
}
template <class T>
__device__ __forceinline__ cpx<T> sse_adj_row2(const cpx<T> *a, const cpx<T> *x)
{
    T er = a[0].re * x[0].re;
    er = er + a[0].im * x[0].im;
    er = er + a[3].re * x[1].re;
    er = er + a[3].im * x[1].im;
    er = er + a[6].re * x[2].re;
    er = er + a[6].im * x[2].im;
    T ei = a[0].re * x[0].im;
    ei = ei - a[0].im * x[0].re;
    ei = ei + a[3].re * x[1].im;
    ei = ei - a[3].im * x[1].re;
    ei = ei + a[6].re * x[2].im;
    ei = ei - a[6].im * x[2].re;
    return { er, ei };
}

// ---------------------------------------------------------------------------
// Spin projection (expressions_gamma.pl:244-276) for one colour row d[0..3].
// ---------------------------------------------------------------------------
// QLA_D3_V_{eq,peq}_M_times_V of the SSE2 versions an --enable-sse3 build
// links (lib/optimized/sse2/include/nas/QLA_D3_V_peq_M_times_V.s): per row
// the (re, im) pair starts from Re a[ic][k] * x_k for the first column of
// the row's "re" order, adds the other two, then adds Im a[ic][k] * i x_k
// = (-Im a x.im, Im a x.re) in the row's "im" order, and (peq) r last.
//   re order: rows 0, 1, 2: k = 0 1 2 | 0 2 1 | 0 1 2
//   im order: rows 0, 1, 2: k = 0 1 2 | 1 0 2 | 2 0 1
template <int EQOP>
__device__ __forceinline__ void sse2_d3_matvec(cpx<double> *r, const cpx<double> *a, const cpx<double> *x)
{
    constexpr int RE[3][3] = { { 0, 1, 2 }, { 0, 2, 1 }, { 0, 1, 2 } };
    constexpr int IM[3][3] = { { 0, 1, 2 }, { 1, 0, 2 }, { 2, 0, 1 } };
#pragma unroll
    for (int ic = 0; ic < 3; ic++) {
        const cpx<double> *row = a + 3 * ic;
        double tr = row[RE[ic][0]].re * x[RE[ic][0]].re, ti = row[RE[ic][0]].re * x[RE[ic][0]].im;
#pragma unroll
        for (int j = 1; j < 3; j++) {
            const int k = RE[ic][j];
            tr = tr + row[k].re * x[k].re;
            ti = ti + row[k].re * x[k].im;
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int k = IM[ic][j];
            tr = tr + row[k].im * (-x[k].im);
            ti = ti + row[k].im * x[k].re;
        }
        if (EQOP == PEQ) {
            tr = tr + r[ic].re;
            ti = ti + r[ic].im;
        }
        r[ic].re = tr;
        r[ic].im = ti;
    }
}

template <class T> __device__ __forceinline__ cpx<T> c_plus(cpx<T> a, cpx<T> b)  { return { a.re + b.re, a.im + b.im }; }
template <class T> __device__ __forceinline__ cpx<T> c_minus(cpx<T> a, cpx<T> b) { return { a.re - b.re, a.im - b.im }; }
// QLA_c_eq_c_plus_ic: (a.re - b.im, a.im + b.re)
template <class T> __device__ __forceinline__ cpx<T> c_plus_i(cpx<T> a, cpx<T> b)  { return { a.re - b.im, a.im + b.re }; }
// QLA_c_eq_c_minus_ic: (a.re + b.im, a.im - b.re)
template <class T> __device__ __forceinline__ cpx<T> c_minus_i(cpx<T> a, cpx<T> b) { return { a.re + b.im, a.im - b.re }; }

template <int MU, int SIGN, class T>
__device__ __forceinline__ void spproj_row(cpx<T> &h0, cpx<T> &h1, const cpx<T> *d)
{
    if (SIGN == 1) {
        if (MU == 0) { h0 = c_plus_i(d[0], d[3]);  h1 = c_plus_i(d[1], d[2]); }
        if (MU == 1) { h0 = c_minus(d[0], d[3]);   h1 = c_plus(d[1], d[2]); }
        if (MU == 2) { h0 = c_plus_i(d[0], d[2]);  h1 = c_minus_i(d[1], d[3]); }
        if (MU == 3) { h0 = c_plus(d[0], d[2]);    h1 = c_plus(d[1], d[3]); }
        if (MU == 4) { h0 = d[0];                  h1 = d[1]; }
    } else {
        if (MU == 0) { h0 = c_minus_i(d[0], d[3]); h1 = c_minus_i(d[1], d[2]); }
        if (MU == 1) { h0 = c_plus(d[0], d[3]);    h1 = c_minus(d[1], d[2]); }
        if (MU == 2) { h0 = c_minus_i(d[0], d[2]); h1 = c_plus_i(d[1], d[3]); }
        if (MU == 3) { h0 = c_minus(d[0], d[2]);   h1 = c_minus(d[1], d[3]); }
        if (MU == 4) { h0 = d[2];                  h1 = d[3]; }
    }
}

// r eqop (a op b) for the projection's complex pairs, each eqop/op pair
// written as the carith2 macro it names (lib/perl/headers.pl:175-200,
// include/qla_complex.h:287-324): op 0 = c+c, 1 = c-c, 2 = c+ic, 3 = c-ic.
// eqm is formed as in its macro (e.g. re = -a.re + b.im), not as a negated
// eq, so zeros keep the reference's sign.
template <int EQOP, int OP, class T>
__device__ __forceinline__ void proj_pair(cpx<T> &c, cpx<T> a, cpx<T> b)
{
    T xr, xi;       // the eq value, for eq / peq / meq
    if (OP == 0) { xr = a.re + b.re; xi = a.im + b.im; }
    if (OP == 1) { xr = a.re - b.re; xi = a.im - b.im; }
    if (OP == 2) { xr = a.re - b.im; xi = a.im + b.re; }
    if (OP == 3) { xr = a.re + b.im; xi = a.im - b.re; }
    if (EQOP == EQ) { c.re = xr; c.im = xi; }
    if (EQOP == PEQ) { c.re = c.re + xr; c.im = c.im + xi; }
    if (EQOP == MEQ) { c.re = c.re - xr; c.im = c.im - xi; }
    if (EQOP == EQM) {
        if (OP == 0) { c.re = -a.re - b.re; c.im = -a.im - b.im; }
        if (OP == 1) { c.re = -a.re + b.re; c.im = -a.im + b.im; }
        if (OP == 2) { c.re = -a.re + b.im; c.im = -a.im - b.re; }
        if (OP == 3) { c.re = -a.re - b.im; c.im = -a.im + b.re; }
    }
}
// r eqop a (carith1 eqc / peqc / meqc / eqmc)
template <int EQOP, class T> __device__ __forceinline__ void copy_eqop(cpx<T> &c, cpx<T> a)
{
    if (EQOP == EQ) c = a;
    if (EQOP == PEQ) { c.re = c.re + a.re; c.im = c.im + a.im; }
    if (EQOP == MEQ) { c.re = c.re - a.re; c.im = c.im - a.im; }
    if (EQOP == EQM) { c.re = -a.re; c.im = -a.im; }
}

// H eqop spproj(D) for one colour row (the unary projection,
// make_code_unary.pl:183-202 with expressions_gamma.pl:244-276): the same
// pairs as spproj_row, each combined into h with the call's eqop.
template <int EQOP, int MU, int SIGN, class T>
__device__ __forceinline__ void spproj_row_eqop(cpx<T> &h0, cpx<T> &h1, const cpx<T> *d)
{
    if (SIGN == 1) {
        if (MU == 0) { proj_pair<EQOP, 2>(h0, d[0], d[3]); proj_pair<EQOP, 2>(h1, d[1], d[2]); }
        if (MU == 1) { proj_pair<EQOP, 1>(h0, d[0], d[3]); proj_pair<EQOP, 0>(h1, d[1], d[2]); }
        if (MU == 2) { proj_pair<EQOP, 2>(h0, d[0], d[2]); proj_pair<EQOP, 3>(h1, d[1], d[3]); }
        if (MU == 3) { proj_pair<EQOP, 0>(h0, d[0], d[2]); proj_pair<EQOP, 0>(h1, d[1], d[3]); }
        if (MU == 4) { copy_eqop<EQOP>(h0, d[0]); copy_eqop<EQOP>(h1, d[1]); }
    } else {
        if (MU == 0) { proj_pair<EQOP, 3>(h0, d[0], d[3]); proj_pair<EQOP, 3>(h1, d[1], d[2]); }
        if (MU == 1) { proj_pair<EQOP, 0>(h0, d[0], d[3]); proj_pair<EQOP, 1>(h1, d[1], d[2]); }
        if (MU == 2) { proj_pair<EQOP, 3>(h0, d[0], d[2]); proj_pair<EQOP, 2>(h1, d[1], d[3]); }
        if (MU == 3) { proj_pair<EQOP, 1>(h0, d[0], d[2]); proj_pair<EQOP, 1>(h1, d[1], d[3]); }
        if (MU == 4) { copy_eqop<EQOP>(h0, d[2]); copy_eqop<EQOP>(h1, d[3]); }
    }
}

// ---------------------------------------------------------------------------
// Spin reconstruction (expressions_gamma.pl:311-337): spin component sp of
// the result is op(h[src]) with op in {0, +h, -h, +i h, -i h} = {0,1,2,3,4}.
// ---------------------------------------------------------------------------
template <int MU, int SIGN> struct recon_tab;
#define QLA_RECON(MU_, SIGN_, O0, O1, O2, O3, S0, S1, S2, S3)                          \
    template <> struct recon_tab<MU_, SIGN_> {                                         \
        static constexpr int op[4] = { O0, O1, O2, O3 };                               \
        static constexpr int src[4] = { S0, S1, S2, S3 };                              \
    };
QLA_RECON(0, 1, 1, 1, 4, 4, 0, 1, 1, 0)
QLA_RECON(1, 1, 1, 1, 1, 2, 0, 1, 1, 0)
QLA_RECON(2, 1, 1, 1, 4, 3, 0, 1, 0, 1)
QLA_RECON(3, 1, 1, 1, 1, 1, 0, 1, 0, 1)
QLA_RECON(4, 1, 1, 1, 0, 0, 0, 1, 0, 1)
QLA_RECON(0, -1, 1, 1, 3, 3, 0, 1, 1, 0)
QLA_RECON(1, -1, 1, 1, 2, 1, 0, 1, 1, 0)
QLA_RECON(2, -1, 1, 1, 3, 4, 0, 1, 0, 1)
QLA_RECON(3, -1, 1, 1, 2, 2, 0, 1, 0, 1)
QLA_RECON(4, -1, 0, 0, 1, 1, 0, 1, 0, 1)
#undef QLA_RECON

// r eqop op(h) for one complex component (carith1 table, lib/perl/headers.pl:136-170):
// '0' under eq/eqm writes zero and under peq/meq leaves r alone; the sign of
// the eqop folds into the sign of the operand (e.g. peq -ic = QLA_c_meq_ic).
template <int EQOP, int OP, class T>
__device__ __forceinline__ void recon_apply(cpx<T> &r, cpx<T> h)
{
    if (OP == 0) {
        if (EQOP == EQ || EQOP == EQM) { r.re = T(0); r.im = T(0); }
        return;
    }
    constexpr bool NEG = (OP == 2 || OP == 4) != (EQOP == MEQ || EQOP == EQM);
    constexpr bool IMAG = (OP == 3 || OP == 4);
    cpx<T> v;
    if (IMAG) { v.re = -h.im; v.im = h.re; } else v = h;
    if (EQOP == EQ || EQOP == EQM) {
        if (NEG) { r.re = -v.re; r.im = -v.im; } else r = v;
    } else {
        if (NEG) { r.re = r.re - v.re; r.im = r.im - v.im; }
        else     { r.re = r.re + v.re; r.im = r.im + v.im; }
    }
}

} // namespace qla
