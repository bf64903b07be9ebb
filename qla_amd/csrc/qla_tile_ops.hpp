// qla_tile_ops.hpp -- one 64-site tile of each site operation, split into
// a stage step (issue the direct-to-LDS loads of the tile's records) and a
// finish step (wait, transpose out of LDS, compute in the reference order,
// write the result back through LDS).  The plain per-op kernels and the
// batched persistent kernel are loops over these steps; between the two
// steps a loop can issue the next tile's gather-address loads so they travel
// with the current tile's records.
//
// Per-lane operand addressing: `site` is the lane's site index (v-form:
// base + lane; x-form: index[base + lane]), `act` whether the lane holds a
// site of the tile, `pb` the lane's pointer-array entry (gathered record).
#pragma once
#include "qla_dev.hpp"

namespace qla {

template <class T> struct rec {
    static constexpr int C = (int)sizeof(cpx<T>);
    static constexpr int V = 3 * C, H = 6 * C, D = 12 * C, M = 9 * C;
};

// Wait for this wave's direct-to-LDS loads, then order LDS across lanes.
__device__ __forceinline__ void dma_wait()
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    wave_sync();
}

// ---------------------------------------------------------------------------
// Tile movers.  LDS always holds packed records (site s at s*BPS): 16-byte
// chunks where the addresses allow, 4-byte chunks otherwise (both land
// packed in LDS).  `al16` flags of contiguous operands come from the host;
// gathered addresses are checked per wave.
// ---------------------------------------------------------------------------
template <int BPS>
__device__ __forceinline__ void load_tile(char *lds, const char *g, int nsites, int lane, bool al16)
{
    if (al16) stage_bytes<16, WAVE * BPS>(lds, g, nsites * BPS, lane);
    else stage_bytes<4, WAVE * BPS>(lds, g, nsites * BPS, lane);
}
template <int BPS>
__device__ __forceinline__ void gather_tile(char *lds, uint64_t addr, int nsites, int lane)
{
    if constexpr (BPS % 16 == 0) {
        if (__all((addr & 15) == 0)) { stage_gather<BPS, 16>(lds, addr, nsites, lane); return; }
    }
    stage_gather<BPS, 4>(lds, addr, nsites, lane);
}
template <int BPS>
__device__ __forceinline__ void store_tile(char *g, const char *lds, int nsites, int lane, bool al16)
{
    if (al16) store_bytes<16, WAVE * BPS>(g, lds, nsites * BPS, lane);
    else store_bytes<4, WAVE * BPS>(g, lds, nsites * BPS, lane);
}
template <int BPS>
__device__ __forceinline__ void scatter_tile(uint64_t addr, const char *lds, int nsites, int lane)
{
    if constexpr (BPS % 16 == 0) {
        if (__all((addr & 15) == 0)) { store_scatter<BPS, 16>(addr, lds, nsites, lane); return; }
    }
    store_scatter<BPS, 4>(addr, lds, nsites, lane);
}

template <class T, int N>
__device__ __forceinline__ void lds_read(cpx<T> (&v)[N], const char *p)
{
    const cpx<T> *q = (const cpx<T> *)p;
#pragma unroll
    for (int i = 0; i < N; i++) v[i] = q[i];
}
template <class T, int N>
__device__ __forceinline__ void lds_write(char *p, const cpx<T> (&v)[N])
{
    cpx<T> *q = (cpx<T> *)p;
#pragma unroll
    for (int i = 0; i < N; i++) q[i] = v[i];
}

// alignment flags of the contiguous operands (bit set: 16-byte aligned)
enum { AL_A = 1, AL_B = 2, AL_R = 4 };


// ---------------------------------------------------------------------------
// V eqop M * V : b contiguous (PB = false) or pointer array (PB = true)
// ---------------------------------------------------------------------------
template <class T, int EQOP, bool ADJ, bool PB, bool XF> struct matvec_op {
    typedef rec<T> R;
    static constexpr bool READ_R = (EQOP == PEQ || EQOP == MEQ);
    static constexpr int BYTES = WAVE * (R::M + 2 * R::V);
    static __device__ __forceinline__ void stage(char *lds, cpx<T> *r, const cpx<T> *a, const void *b, uint64_t pb,
                                                 int base, int ns, int lane, int site, bool act, int al)
    {
        char *lm = lds, *lb = lm + WAVE * R::M, *lr = lb + WAVE * R::V;
        if (!XF) {
            load_tile<R::M>(lm, (const char *)(a + 9 * (size_t)base), ns, lane, al & AL_A);
            if (!PB) load_tile<R::V>(lb, (const char *)b + (size_t)base * R::V, ns, lane, al & AL_B);
            if (READ_R) load_tile<R::V>(lr, (const char *)(r + 3 * (size_t)base), ns, lane, al & AL_R);
        } else {
            gather_tile<R::M>(lm, act ? (uint64_t)(a + 9 * (size_t)site) : 0, ns, lane);
            if (!PB) gather_tile<R::V>(lb, act ? (uint64_t)((const char *)b + (size_t)site * R::V) : 0, ns, lane);
            if (READ_R) gather_tile<R::V>(lr, act ? (uint64_t)(r + 3 * (size_t)site) : 0, ns, lane);
        }
        if (PB) gather_tile<R::V>(lb, pb, ns, lane);
    }
    static __device__ __forceinline__ void finish(char *lds, cpx<T> *r, int base, int ns, int lane, int site, bool act,
                                                  int al)
    {
        char *lm = lds, *lb = lm + WAVE * R::M, *lr = lb + WAVE * R::V;
        dma_wait();
        cpx<T> m[9], v[3], x[3];
        lds_read(m, lm + lane * R::M);
        lds_read(v, lb + lane * R::V);
        if (READ_R) lds_read(x, lr + lane * R::V);
        mat_times_cols<EQOP, ADJ, 1>(x, m, v);
        wave_sync();
        lds_write(lr + lane * R::V, x);
        wave_sync();
        if (!XF) store_tile<R::V>((char *)(r + 3 * (size_t)base), lr, ns, lane, al & AL_R);
        else scatter_tile<R::V>(act ? (uint64_t)(r + 3 * (size_t)site) : 0, lr, ns, lane);
        wave_sync();
    }
};

// ---------------------------------------------------------------------------
// H eqop spproj_{MU,SIGN}(M * pD)
// ---------------------------------------------------------------------------
template <class T, int EQOP, bool ADJ, int MU, int SIGN, bool XF> struct spproj_op {
    typedef rec<T> R;
    static constexpr bool READ_R = (EQOP == PEQ || EQOP == MEQ);
    // the M region is reused for the H result (H <= M)
    static constexpr int BYTES = WAVE * (R::M + R::D + (READ_R ? R::H : 0));
    static __device__ __forceinline__ char *hreg(char *lds) { return READ_R ? lds + WAVE * (R::M + R::D) : lds; }
    static __device__ __forceinline__ void stage(char *lds, cpx<T> *r, const cpx<T> *a, const void *, uint64_t pb,
                                                 int base, int ns, int lane, int site, bool act, int al)
    {
        char *lm = lds, *ld = lm + WAVE * R::M, *lh = hreg(lds);
        if (!XF) {
            load_tile<R::M>(lm, (const char *)(a + 9 * (size_t)base), ns, lane, al & AL_A);
            if (READ_R) load_tile<R::H>(lh, (const char *)(r + 6 * (size_t)base), ns, lane, al & AL_R);
        } else {
            gather_tile<R::M>(lm, act ? (uint64_t)(a + 9 * (size_t)site) : 0, ns, lane);
            if (READ_R) gather_tile<R::H>(lh, act ? (uint64_t)(r + 6 * (size_t)site) : 0, ns, lane);
        }
        gather_tile<R::D>(ld, pb, ns, lane);
    }
    static __device__ __forceinline__ void finish(char *lds, cpx<T> *r, int base, int ns, int lane, int site, bool act,
                                                  int al)
    {
        char *lm = lds, *ld = lm + WAVE * R::M, *lh = hreg(lds);
        dma_wait();
        cpx<T> m[9], d[12], h[6];
        lds_read(m, lm + lane * R::M);
        lds_read(d, ld + lane * R::D);
        if (READ_R) lds_read(h, lh + lane * R::H);
        cpx<T> tp[6];
#pragma unroll
        for (int ic = 0; ic < 3; ic++) spproj_row<MU, SIGN>(tp[2 * ic], tp[2 * ic + 1], d + 4 * ic);
        mat_times_cols<EQOP, ADJ, 2>(h, m, tp);
        wave_sync();
        lds_write(lh + lane * R::H, h);
        wave_sync();
        if (!XF) store_tile<R::H>((char *)(r + 6 * (size_t)base), lh, ns, lane, al & AL_R);
        else scatter_tile<R::H>(act ? (uint64_t)(r + 6 * (size_t)site) : 0, lh, ns, lane);
        wave_sync();
    }
};

// ---------------------------------------------------------------------------
// D eqop sprecon_{MU,SIGN}(M * pH)
// ---------------------------------------------------------------------------
template <class T, int EQOP, bool ADJ, int MU, int SIGN, bool XF> struct sprecon_op {
    typedef rec<T> R;
    typedef recon_tab<MU, SIGN> TAB;
    // eq/eqm overwrite every component; peq/meq read the destination first
    static constexpr bool READ_R = (EQOP == PEQ || EQOP == MEQ);
    static constexpr int BYTES = WAVE * (R::M + R::H + R::D);
    static __device__ __forceinline__ void stage(char *lds, cpx<T> *r, const cpx<T> *a, const void *, uint64_t pb,
                                                 int base, int ns, int lane, int site, bool act, int al)
    {
        char *lm = lds, *lh = lm + WAVE * R::M, *ld = lh + WAVE * R::H;
        if (!XF) {
            load_tile<R::M>(lm, (const char *)(a + 9 * (size_t)base), ns, lane, al & AL_A);
            if (READ_R) load_tile<R::D>(ld, (const char *)(r + 12 * (size_t)base), ns, lane, al & AL_R);
        } else {
            gather_tile<R::M>(lm, act ? (uint64_t)(a + 9 * (size_t)site) : 0, ns, lane);
            if (READ_R) gather_tile<R::D>(ld, act ? (uint64_t)(r + 12 * (size_t)site) : 0, ns, lane);
        }
        gather_tile<R::H>(lh, pb, ns, lane);
    }
    static __device__ __forceinline__ void finish(char *lds, cpx<T> *r, int base, int ns, int lane, int site, bool act,
                                                  int al)
    {
        char *lm = lds, *lh = lm + WAVE * R::M, *ld = lh + WAVE * R::H;
        dma_wait();
        cpx<T> m[9], h[6], d[12];
        lds_read(m, lm + lane * R::M);
        lds_read(h, lh + lane * R::H);
        if (READ_R) lds_read(d, ld + lane * R::D);
        cpx<T> tv[6];
        mat_times_cols<EQ, ADJ, 2>(tv, m, h);
#pragma unroll
        for (int ic = 0; ic < 3; ic++) {
            recon_apply<EQOP, TAB::op[0]>(d[4 * ic + 0], tv[2 * ic + TAB::src[0]]);
            recon_apply<EQOP, TAB::op[1]>(d[4 * ic + 1], tv[2 * ic + TAB::src[1]]);
            recon_apply<EQOP, TAB::op[2]>(d[4 * ic + 2], tv[2 * ic + TAB::src[2]]);
            recon_apply<EQOP, TAB::op[3]>(d[4 * ic + 3], tv[2 * ic + TAB::src[3]]);
        }
        wave_sync();
        lds_write(ld + lane * R::D, d);
        wave_sync();
        if (!XF) store_tile<R::D>((char *)(r + 12 * (size_t)base), ld, ns, lane, al & AL_R);
        else scatter_tile<R::D>(act ? (uint64_t)(r + 12 * (size_t)site) : 0, ld, ns, lane);
        wave_sync();
    }
};

} // namespace qla
