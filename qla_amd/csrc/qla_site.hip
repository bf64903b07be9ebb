// qla_site.hip -- gfx950 kernels for the QLA site-loop hot path:
//   V_{eqop}_M_times_{V,pV}            (matrix times colour vector)
//   H_{eqop}_spproj_M_times_pD          (gather D, spin-project, multiply)
//   D_{eqop}_sprecon_M_times_pH         (gather H, multiply, spin-reconstruct)
// each in the v-form (site i) and the gang-indexed x-form (site index[i]).
//
// One wavefront owns one tile of 64 sites.  Records move HBM -> LDS with
// direct-to-LDS loads (coalesced for contiguous fields and for adjacent
// gather targets), each lane then takes its own site's records from LDS,
// computes in registers in the reference's operation order, writes the
// result record back to LDS and the wave stores it coalesced.  All kernels
// are HBM-bandwidth bound (0.3-0.7 flop/byte), so no MFMA is involved.
#include "qla_dev.hpp"
#include "qla_launch.hpp"

namespace qla {

template <class T> struct rec {
    static constexpr int C = (int)sizeof(cpx<T>);
    static constexpr int V = 3 * C, H = 6 * C, D = 12 * C, M = 9 * C;
    // LDS DMA chunk: dwordx3 needs only the 4-byte alignment of float
    // records; double records are 16-byte aligned (qla_complex.h:35).
    static constexpr int CH = sizeof(T) == 4 ? 12 : 16;
    // LDS footprint of one staged record (12-byte chunks sit at a 16-byte stride)
    static constexpr int LV = rec_lds<V, CH>::v, LH = rec_lds<H, CH>::v;
    static constexpr int LD = rec_lds<D, CH>::v, LM = rec_lds<M, CH>::v;
};

// Wait for this wave's direct-to-LDS loads, then order LDS across lanes.
__device__ __forceinline__ void dma_wait()
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    wave_sync();
}

struct tile_t {
    int base;    // first site (v-form) or first index slot (x-form)
    int nsites;  // valid sites in this tile (1..64)
    int lane;
};

template <int WPB>
__device__ __forceinline__ bool get_tile(tile_t &t, int n)
{
    const int wave = threadIdx.x / WAVE;
    const int tile = blockIdx.x * WPB + wave;
    t.lane = lane_id();
    t.base = tile * WAVE;
    if (t.base >= n) return false;
    t.nsites = min(WAVE, n - t.base);
    return true;
}

// ---------------------------------------------------------------------------
// V eqop M * V  (b contiguous: PB = false; b pointer array: PB = true)
// ---------------------------------------------------------------------------
template <class T, int EQOP, bool ADJ, bool PB, bool XF, int WPB>
__global__ void __launch_bounds__(WPB * WAVE)
k_matvec(cpx<T> *__restrict__ r, const cpx<T> *__restrict__ a, const void *__restrict__ b,
         const int *__restrict__ index, int n)
{
    typedef rec<T> R;
    constexpr int CH = R::CH;
    constexpr bool READ_R = (EQOP == PEQ || EQOP == MEQ);
    constexpr int BYTES = WAVE * (R::LM + 2 * R::LV);
    __shared__ __attribute__((aligned(16))) char smem[WPB * BYTES];
    tile_t t;
    if (!get_tile<WPB>(t, n)) return;
    char *lm = smem + (threadIdx.x / WAVE) * BYTES;
    char *lb = lm + WAVE * R::LM;
    char *lr = lb + WAVE * R::LV;
    const bool act = t.lane < t.nsites;

    int site = t.base + t.lane;
    if (XF) site = act ? index[t.base + t.lane] : 0;
    if (!XF) {
        stage_contig<R::M, CH>(lm, (const char *)(a + 9 * (size_t)t.base), t.nsites, t.lane);
        if (!PB) stage_contig<R::V, CH>(lb, (const char *)b + (size_t)t.base * R::V, t.nsites, t.lane);
        if (READ_R) stage_contig<R::V, CH>(lr, (const char *)(r + 3 * (size_t)t.base), t.nsites, t.lane);
    } else {
        stage_gather<R::M, CH>(lm, (uint64_t)(a + 9 * (size_t)site), t.nsites, t.lane);
        if (!PB) stage_gather<R::V, CH>(lb, (uint64_t)((const char *)b + (size_t)site * R::V), t.nsites, t.lane);
        if (READ_R) stage_gather<R::V, CH>(lr, (uint64_t)(r + 3 * (size_t)site), t.nsites, t.lane);
    }
    if (PB) {
        const uint64_t p = act ? (uint64_t)((const void *const *)b)[site] : 0;
        stage_gather<R::V, CH>(lb, p, t.nsites, t.lane);
    }
    dma_wait();
    cpx<T> m[9], v[3], x[3];
    lds_read_rec<CH>(m, lm + t.lane * R::LM);
    lds_read_rec<CH>(v, lb + t.lane * R::LV);
    if (READ_R) lds_read_rec<CH>(x, lr + t.lane * R::LV);
    mat_times_cols<EQOP, ADJ, 1>(x, m, v);
    wave_sync();
    lds_write_rec<CH>(lr + t.lane * R::LV, x);
    wave_sync();
    if (!XF) store_contig<R::V, CH>((char *)(r + 3 * (size_t)t.base), lr, t.nsites, t.lane);
    else store_scatter<R::V, CH>((uint64_t)(r + 3 * (size_t)site), lr, t.nsites, t.lane);
}

// ---------------------------------------------------------------------------
// H eqop spproj_{MU,SIGN}(M * pD)
// ---------------------------------------------------------------------------
template <class T, int EQOP, bool ADJ, int MU, int SIGN, bool XF, int WPB>
__global__ void __launch_bounds__(WPB * WAVE)
k_spproj(cpx<T> *__restrict__ r, const cpx<T> *__restrict__ a, const cpx<T> *const *__restrict__ b,
         const int *__restrict__ index, int n)
{
    typedef rec<T> R;
    constexpr int CH = R::CH;
    constexpr bool READ_R = (EQOP == PEQ || EQOP == MEQ);
    constexpr int BYTES = WAVE * (R::LM + R::LD + R::LH);
    __shared__ __attribute__((aligned(16))) char smem[WPB * BYTES];
    tile_t t;
    if (!get_tile<WPB>(t, n)) return;
    char *lm = smem + (threadIdx.x / WAVE) * BYTES;
    char *ld = lm + WAVE * R::LM;
    char *lh = ld + WAVE * R::LD;
    const bool act = t.lane < t.nsites;

    int site = t.base + t.lane;
    if (XF) site = act ? index[t.base + t.lane] : 0;
    const uint64_t pd = act ? (uint64_t)b[site] : 0;
    if (!XF) {
        stage_contig<R::M, CH>(lm, (const char *)(a + 9 * (size_t)t.base), t.nsites, t.lane);
        if (READ_R) stage_contig<R::H, CH>(lh, (const char *)(r + 6 * (size_t)t.base), t.nsites, t.lane);
    } else {
        stage_gather<R::M, CH>(lm, (uint64_t)(a + 9 * (size_t)site), t.nsites, t.lane);
        if (READ_R) stage_gather<R::H, CH>(lh, (uint64_t)(r + 6 * (size_t)site), t.nsites, t.lane);
    }
    stage_gather<R::D, CH>(ld, pd, t.nsites, t.lane);
    dma_wait();
    cpx<T> m[9], d[12], h[6];
    lds_read_rec<CH>(m, lm + t.lane * R::LM);
    lds_read_rec<CH>(d, ld + t.lane * R::LD);
    if (READ_R) lds_read_rec<CH>(h, lh + t.lane * R::LH);
    cpx<T> tp[6];
#pragma unroll
    for (int ic = 0; ic < 3; ic++) spproj_row<MU, SIGN>(tp[2 * ic], tp[2 * ic + 1], d + 4 * ic);
    mat_times_cols<EQOP, ADJ, 2>(h, m, tp);
    wave_sync();
    lds_write_rec<CH>(lh + t.lane * R::LH, h);
    wave_sync();
    if (!XF) store_contig<R::H, CH>((char *)(r + 6 * (size_t)t.base), lh, t.nsites, t.lane);
    else store_scatter<R::H, CH>((uint64_t)(r + 6 * (size_t)site), lh, t.nsites, t.lane);
}

// ---------------------------------------------------------------------------
// D eqop sprecon_{MU,SIGN}(M * pH)
// ---------------------------------------------------------------------------
template <class T, int EQOP, bool ADJ, int MU, int SIGN, bool XF, int WPB>
__global__ void __launch_bounds__(WPB * WAVE)
k_sprecon(cpx<T> *__restrict__ r, const cpx<T> *__restrict__ a, const cpx<T> *const *__restrict__ b,
          const int *__restrict__ index, int n)
{
    typedef rec<T> R;
    typedef recon_tab<MU, SIGN> TAB;
    constexpr int CH = R::CH;
    // eq/eqm overwrite every component; peq/meq read the destination first
    constexpr bool READ_R = (EQOP == PEQ || EQOP == MEQ);
    constexpr int BYTES = WAVE * (R::LM + R::LH + R::LD);
    __shared__ __attribute__((aligned(16))) char smem[WPB * BYTES];
    tile_t t;
    if (!get_tile<WPB>(t, n)) return;
    char *lm = smem + (threadIdx.x / WAVE) * BYTES;
    char *lh = lm + WAVE * R::LM;
    char *ld = lh + WAVE * R::LH;
    const bool act = t.lane < t.nsites;

    int site = t.base + t.lane;
    if (XF) site = act ? index[t.base + t.lane] : 0;
    const uint64_t ph = act ? (uint64_t)b[site] : 0;
    if (!XF) {
        stage_contig<R::M, CH>(lm, (const char *)(a + 9 * (size_t)t.base), t.nsites, t.lane);
        if (READ_R) stage_contig<R::D, CH>(ld, (const char *)(r + 12 * (size_t)t.base), t.nsites, t.lane);
    } else {
        stage_gather<R::M, CH>(lm, (uint64_t)(a + 9 * (size_t)site), t.nsites, t.lane);
        if (READ_R) stage_gather<R::D, CH>(ld, (uint64_t)(r + 12 * (size_t)site), t.nsites, t.lane);
    }
    stage_gather<R::H, CH>(lh, ph, t.nsites, t.lane);
    dma_wait();
    cpx<T> m[9], h[6], d[12];
    lds_read_rec<CH>(m, lm + t.lane * R::LM);
    lds_read_rec<CH>(h, lh + t.lane * R::LH);
    if (READ_R) lds_read_rec<CH>(d, ld + t.lane * R::LD);
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
    lds_write_rec<CH>(ld + t.lane * R::LD, d);
    wave_sync();
    if (!XF) store_contig<R::D, CH>((char *)(r + 12 * (size_t)t.base), ld, t.nsites, t.lane);
    else store_scatter<R::D, CH>((uint64_t)(r + 12 * (size_t)site), ld, t.nsites, t.lane);
}

// ---------------------------------------------------------------------------
// Host-side launchers (called from the C ABI layer).
// ---------------------------------------------------------------------------
template <class T> struct wpb_of { static constexpr int v = 2; };

template <class T, int EQOP, bool ADJ, bool PB, bool XF>
static hipError_t launch_matvec_t(void *r, const void *a, const void *b, const int *index, int n,
                                  hipStream_t st)
{
    constexpr int W = wpb_of<T>::v;
    const int ntiles = (n + WAVE - 1) / WAVE;
    dim3 grid((ntiles + W - 1) / W), block(W * WAVE);
    hipLaunchKernelGGL((k_matvec<T, EQOP, ADJ, PB, XF, W>), grid, block, 0, st, (cpx<T> *)r,
                       (const cpx<T> *)a, b, index, n);
    return hipGetLastError();
}

template <class T, int EQOP, bool ADJ, int MU, int SIGN, bool XF>
static hipError_t launch_spproj_t(void *r, const void *a, const void *b, const int *index, int n,
                                  hipStream_t st)
{
    constexpr int W = wpb_of<T>::v;
    const int ntiles = (n + WAVE - 1) / WAVE;
    dim3 grid((ntiles + W - 1) / W), block(W * WAVE);
    hipLaunchKernelGGL((k_spproj<T, EQOP, ADJ, MU, SIGN, XF, W>), grid, block, 0, st, (cpx<T> *)r,
                       (const cpx<T> *)a, (const cpx<T> *const *)b, index, n);
    return hipGetLastError();
}

template <class T, int EQOP, bool ADJ, int MU, int SIGN, bool XF>
static hipError_t launch_sprecon_t(void *r, const void *a, const void *b, const int *index, int n,
                                   hipStream_t st)
{
    constexpr int W = wpb_of<T>::v;
    const int ntiles = (n + WAVE - 1) / WAVE;
    dim3 grid((ntiles + W - 1) / W), block(W * WAVE);
    hipLaunchKernelGGL((k_sprecon<T, EQOP, ADJ, MU, SIGN, XF, W>), grid, block, 0, st, (cpx<T> *)r,
                       (const cpx<T> *)a, (const cpx<T> *const *)b, index, n);
    return hipGetLastError();
}

// mu/sign dispatch: mu outside 0..4 is a no-op (reference switch without
// default); any sign other than +1 selects the -1 table.
#define QLA_MUSIGN_DISPATCH(FN, ...)                                                   \
    if (sign == 1) {                                                                   \
        switch (mu) {                                                                  \
        case 0: return FN<0, 1, XF>(__VA_ARGS__);                                      \
        case 1: return FN<1, 1, XF>(__VA_ARGS__);                                      \
        case 2: return FN<2, 1, XF>(__VA_ARGS__);                                      \
        case 3: return FN<3, 1, XF>(__VA_ARGS__);                                      \
        case 4: return FN<4, 1, XF>(__VA_ARGS__);                                      \
        default: return hipSuccess;                                                    \
        }                                                                              \
    } else {                                                                           \
        switch (mu) {                                                                  \
        case 0: return FN<0, -1, XF>(__VA_ARGS__);                                     \
        case 1: return FN<1, -1, XF>(__VA_ARGS__);                                     \
        case 2: return FN<2, -1, XF>(__VA_ARGS__);                                     \
        case 3: return FN<3, -1, XF>(__VA_ARGS__);                                     \
        case 4: return FN<4, -1, XF>(__VA_ARGS__);                                     \
        default: return hipSuccess;                                                    \
        }                                                                              \
    }

template <class T, int EQOP, bool ADJ> struct proj_disp {
    template <int MU, int SIGN, bool XF>
    static hipError_t spproj(void *r, const void *a, const void *b, const int *ix, int n, hipStream_t s)
    { return launch_spproj_t<T, EQOP, ADJ, MU, SIGN, XF>(r, a, b, ix, n, s); }
    template <int MU, int SIGN, bool XF>
    static hipError_t sprecon(void *r, const void *a, const void *b, const int *ix, int n, hipStream_t s)
    { return launch_sprecon_t<T, EQOP, ADJ, MU, SIGN, XF>(r, a, b, ix, n, s); }
};

template <class T, int EQOP, bool ADJ, bool XF>
static hipError_t spproj_mu(void *r, const void *a, const void *b, int mu, int sign, const int *ix, int n,
                            hipStream_t s)
{
    typedef proj_disp<T, EQOP, ADJ> P;
    QLA_MUSIGN_DISPATCH(P::template spproj, r, a, b, ix, n, s)
}
template <class T, int EQOP, bool ADJ, bool XF>
static hipError_t sprecon_mu(void *r, const void *a, const void *b, int mu, int sign, const int *ix, int n,
                             hipStream_t s)
{
    typedef proj_disp<T, EQOP, ADJ> P;
    QLA_MUSIGN_DISPATCH(P::template sprecon, r, a, b, ix, n, s)
}
#undef QLA_MUSIGN_DISPATCH

} // namespace qla

using namespace qla;

// Entry points used by the C ABI layer (qla_capi.cpp).  prec: 0 = float, 1 = double.
hipError_t qla_launch_matvec(int prec, int eqop_peq, int pb, void *r, const void *a, const void *b,
                             const int *index, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (prec == 0) {
        if (!eqop_peq && !pb && !index) return launch_matvec_t<float, EQ, false, false, false>(r, a, b, index, n, st);
        if (eqop_peq && pb && !index) return launch_matvec_t<float, PEQ, false, true, false>(r, a, b, index, n, st);
        if (!eqop_peq && !pb && index) return launch_matvec_t<float, EQ, false, false, true>(r, a, b, index, n, st);
        if (eqop_peq && pb && index) return launch_matvec_t<float, PEQ, false, true, true>(r, a, b, index, n, st);
    } else {
        if (!eqop_peq && !pb && !index) return launch_matvec_t<double, EQ, false, false, false>(r, a, b, index, n, st);
        if (eqop_peq && pb && !index) return launch_matvec_t<double, PEQ, false, true, false>(r, a, b, index, n, st);
        if (!eqop_peq && !pb && index) return launch_matvec_t<double, EQ, false, false, true>(r, a, b, index, n, st);
        if (eqop_peq && pb && index) return launch_matvec_t<double, PEQ, false, true, true>(r, a, b, index, n, st);
    }
    return hipErrorInvalidValue;
}

hipError_t qla_launch_spproj(int prec, void *r, const void *a, const void *b, int mu, int sign,
                             const int *index, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (prec == 0) {
        if (index) return spproj_mu<float, EQ, false, true>(r, a, b, mu, sign, index, n, st);
        return spproj_mu<float, EQ, false, false>(r, a, b, mu, sign, index, n, st);
    }
    if (index) return spproj_mu<double, EQ, false, true>(r, a, b, mu, sign, index, n, st);
    return spproj_mu<double, EQ, false, false>(r, a, b, mu, sign, index, n, st);
}

hipError_t qla_launch_sprecon(int prec, void *r, const void *a, const void *b, int mu, int sign,
                              const int *index, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (prec == 0) {
        if (index) return sprecon_mu<float, PEQ, false, true>(r, a, b, mu, sign, index, n, st);
        return sprecon_mu<float, PEQ, false, false>(r, a, b, mu, sign, index, n, st);
    }
    if (index) return sprecon_mu<double, PEQ, false, true>(r, a, b, mu, sign, index, n, st);
    return sprecon_mu<double, PEQ, false, false>(r, a, b, mu, sign, index, n, st);
}
