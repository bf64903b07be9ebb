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

// Persistent tile loop: wave w of the grid takes tiles w, w + W, w + 2W, ...
// (W = waves in the grid).  The per-site gather address of the next tile
// (pointer array entry, or index entry for the x-form) is loaded while the
// current tile's records are in flight, so a tile costs one memory latency.
struct tiles {
    int tile, stride, ntiles, n, lane;
    __device__ __forceinline__ tiles(int wpb, int n_) : n(n_)
    {
        lane = lane_id();
        tile = blockIdx.x * wpb + (int)(threadIdx.x / WAVE);
        stride = gridDim.x * wpb;
        ntiles = (n + WAVE - 1) / WAVE;
    }
    __device__ __forceinline__ int base() const { return tile * WAVE; }
    __device__ __forceinline__ int nsites() const { return min(WAVE, n - tile * WAVE); }
};

// ---------------------------------------------------------------------------
// V eqop M * V  (b contiguous: PB = false; b pointer array: PB = true)
// ---------------------------------------------------------------------------
template <class T, int EQOP, bool ADJ, bool PB, bool XF, int WPB>
__global__ void __launch_bounds__(WPB * WAVE)
k_matvec(cpx<T> *__restrict__ r, const cpx<T> *__restrict__ a, const void *__restrict__ b,
         const int *__restrict__ index, int n, int al)
{
    typedef rec<T> R;
    constexpr bool READ_R = (EQOP == PEQ || EQOP == MEQ);
    constexpr int BYTES = WAVE * (R::M + 2 * R::V);
    __shared__ __attribute__((aligned(16))) char smem[WPB * BYTES];
    char *lm = smem + (threadIdx.x / WAVE) * BYTES;
    char *lb = lm + WAVE * R::M;
    char *lr = lb + WAVE * R::V;
    tiles t(WPB, n);
    for (; t.tile < t.ntiles; t.tile += t.stride) {
        const int base = t.base(), ns = t.nsites();
        const bool act = t.lane < ns;
        int site = base + t.lane;
        if (XF) site = act ? index[base + t.lane] : 0;
        if (!XF) {
            load_tile<R::M>(lm, (const char *)(a + 9 * (size_t)base), ns, t.lane, al & AL_A);
            if (!PB) load_tile<R::V>(lb, (const char *)b + (size_t)base * R::V, ns, t.lane, al & AL_B);
            if (READ_R) load_tile<R::V>(lr, (const char *)(r + 3 * (size_t)base), ns, t.lane, al & AL_R);
        } else {
            gather_tile<R::M>(lm, act ? (uint64_t)(a + 9 * (size_t)site) : 0, ns, t.lane);
            if (!PB) gather_tile<R::V>(lb, act ? (uint64_t)((const char *)b + (size_t)site * R::V) : 0, ns, t.lane);
            if (READ_R) gather_tile<R::V>(lr, act ? (uint64_t)(r + 3 * (size_t)site) : 0, ns, t.lane);
        }
        if (PB) {
            const uint64_t p = act ? (uint64_t)((const void *const *)b)[site] : 0;
            gather_tile<R::V>(lb, p, ns, t.lane);
        }
        dma_wait();
        cpx<T> m[9], v[3], x[3];
        lds_read(m, lm + t.lane * R::M);
        lds_read(v, lb + t.lane * R::V);
        if (READ_R) lds_read(x, lr + t.lane * R::V);
        mat_times_cols<EQOP, ADJ, 1>(x, m, v);
        wave_sync();
        lds_write(lr + t.lane * R::V, x);
        wave_sync();
        if (!XF) store_tile<R::V>((char *)(r + 3 * (size_t)base), lr, ns, t.lane, al & AL_R);
        else scatter_tile<R::V>(act ? (uint64_t)(r + 3 * (size_t)site) : 0, lr, ns, t.lane);
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// H eqop spproj_{MU,SIGN}(M * pD)
// ---------------------------------------------------------------------------
template <class T, int EQOP, bool ADJ, int MU, int SIGN, bool XF, int WPB, int MINB>
__global__ void __launch_bounds__(WPB * WAVE, MINB)
k_spproj(cpx<T> *__restrict__ r, const cpx<T> *__restrict__ a, const cpx<T> *const *__restrict__ b,
         const int *__restrict__ index, int n, int al)
{
    typedef rec<T> R;
    constexpr bool READ_R = (EQOP == PEQ || EQOP == MEQ);
    // M region is reused for the H result (H <= M)
    constexpr int BYTES = WAVE * (R::M + R::D + (READ_R ? R::H : 0));
    __shared__ __attribute__((aligned(16))) char smem[WPB * BYTES];
    char *lm = smem + (threadIdx.x / WAVE) * BYTES;
    char *ld = lm + WAVE * R::M;
    char *lh = READ_R ? ld + WAVE * R::D : lm;
    tiles t(WPB, n);
    if (t.tile >= t.ntiles) return;
    // gather address of the first tile
    int site = t.base() + t.lane;
    bool act = t.lane < t.nsites();
    if (XF) site = act ? index[site] : 0;
    uint64_t pd = act ? (uint64_t)b[site] : 0;
    for (;;) {
        const int base = t.base(), ns = t.nsites();
        if (!XF) {
            load_tile<R::M>(lm, (const char *)(a + 9 * (size_t)base), ns, t.lane, al & AL_A);
            if (READ_R) load_tile<R::H>(lh, (const char *)(r + 6 * (size_t)base), ns, t.lane, al & AL_R);
        } else {
            gather_tile<R::M>(lm, act ? (uint64_t)(a + 9 * (size_t)site) : 0, ns, t.lane);
            if (READ_R) gather_tile<R::H>(lh, act ? (uint64_t)(r + 6 * (size_t)site) : 0, ns, t.lane);
        }
        gather_tile<R::D>(ld, pd, ns, t.lane);
        // prefetch the next tile's gather address while this tile is in flight
        const int next = t.tile + t.stride;
        int nsite = 0;
        uint64_t npd = 0;
        bool nact = false;
        if (next < t.ntiles) {
            nsite = next * WAVE + t.lane;
            nact = t.lane < min(WAVE, n - next * WAVE);
            if (XF) nsite = nact ? index[nsite] : 0;
            else npd = nact ? (uint64_t)b[nsite] : 0;
        }
        dma_wait();
        cpx<T> m[9], d[12], h[6];
        lds_read(m, lm + t.lane * R::M);
        lds_read(d, ld + t.lane * R::D);
        if (READ_R) lds_read(h, lh + t.lane * R::H);
        cpx<T> tp[6];
#pragma unroll
        for (int ic = 0; ic < 3; ic++) spproj_row<MU, SIGN>(tp[2 * ic], tp[2 * ic + 1], d + 4 * ic);
        mat_times_cols<EQOP, ADJ, 2>(h, m, tp);
        wave_sync();
        lds_write(lh + t.lane * R::H, h);
        wave_sync();
        if (!XF) store_tile<R::H>((char *)(r + 6 * (size_t)base), lh, ns, t.lane, al & AL_R);
        else scatter_tile<R::H>(act ? (uint64_t)(r + 6 * (size_t)site) : 0, lh, ns, t.lane);
        wave_sync();
        if (next >= t.ntiles) break;
        t.tile = next;
        site = nsite;
        act = nact;
        if (XF) npd = nact ? (uint64_t)b[nsite] : 0;
        pd = npd;
    }
}

// ---------------------------------------------------------------------------
// D eqop sprecon_{MU,SIGN}(M * pH)
// ---------------------------------------------------------------------------
template <class T, int EQOP, bool ADJ, int MU, int SIGN, bool XF, int WPB, int MINB>
__global__ void __launch_bounds__(WPB * WAVE, MINB)
k_sprecon(cpx<T> *__restrict__ r, const cpx<T> *__restrict__ a, const cpx<T> *const *__restrict__ b,
          const int *__restrict__ index, int n, int al)
{
    typedef rec<T> R;
    typedef recon_tab<MU, SIGN> TAB;
    // eq/eqm overwrite every component; peq/meq read the destination first
    constexpr bool READ_R = (EQOP == PEQ || EQOP == MEQ);
    constexpr int BYTES = WAVE * (R::M + R::H + R::D);
    __shared__ __attribute__((aligned(16))) char smem[WPB * BYTES];
    char *lm = smem + (threadIdx.x / WAVE) * BYTES;
    char *lh = lm + WAVE * R::M;
    char *ld = lh + WAVE * R::H;
    tiles t(WPB, n);
    if (t.tile >= t.ntiles) return;
    int site = t.base() + t.lane;
    bool act = t.lane < t.nsites();
    if (XF) site = act ? index[site] : 0;
    uint64_t ph = act ? (uint64_t)b[site] : 0;
    for (;;) {
        const int base = t.base(), ns = t.nsites();
        if (!XF) {
            load_tile<R::M>(lm, (const char *)(a + 9 * (size_t)base), ns, t.lane, al & AL_A);
            if (READ_R) load_tile<R::D>(ld, (const char *)(r + 12 * (size_t)base), ns, t.lane, al & AL_R);
        } else {
            gather_tile<R::M>(lm, act ? (uint64_t)(a + 9 * (size_t)site) : 0, ns, t.lane);
            if (READ_R) gather_tile<R::D>(ld, act ? (uint64_t)(r + 12 * (size_t)site) : 0, ns, t.lane);
        }
        gather_tile<R::H>(lh, ph, ns, t.lane);
        const int next = t.tile + t.stride;
        int nsite = 0;
        uint64_t nph = 0;
        bool nact = false;
        if (next < t.ntiles) {
            nsite = next * WAVE + t.lane;
            nact = t.lane < min(WAVE, n - next * WAVE);
            if (XF) nsite = nact ? index[nsite] : 0;
            else nph = nact ? (uint64_t)b[nsite] : 0;
        }
        dma_wait();
        cpx<T> m[9], h[6], d[12];
        lds_read(m, lm + t.lane * R::M);
        lds_read(h, lh + t.lane * R::H);
        if (READ_R) lds_read(d, ld + t.lane * R::D);
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
        lds_write(ld + t.lane * R::D, d);
        wave_sync();
        if (!XF) store_tile<R::D>((char *)(r + 12 * (size_t)base), ld, ns, t.lane, al & AL_R);
        else scatter_tile<R::D>(act ? (uint64_t)(r + 12 * (size_t)site) : 0, ld, ns, t.lane);
        wave_sync();
        if (next >= t.ntiles) break;
        t.tile = next;
        site = nsite;
        act = nact;
        if (XF) nph = nact ? (uint64_t)b[nsite] : 0;
        ph = nph;
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers (called from the C ABI layer).
// Launch shape: WPB waves per block, MINB blocks per CU requested from the
// register allocator; the grid covers the tiles with about TPW tiles per
// wave (the persistent loop balances the rest).
// ---------------------------------------------------------------------------
template <class T> struct shape_of;
template <> struct shape_of<float> { static constexpr int WPB = 2, MINB = 5; };
template <> struct shape_of<double> { static constexpr int WPB = 1, MINB = 5; };

static int g_tiles_per_wave = 2;

static inline int align_flags(const void *a, const void *b, const void *r)
{
    return (((uintptr_t)a % 16 == 0) ? AL_A : 0) | (((uintptr_t)b % 16 == 0) ? AL_B : 0) |
           (((uintptr_t)r % 16 == 0) ? AL_R : 0);
}

static inline dim3 grid_for(int n, int wpb)
{
    const long long ntiles = ((long long)n + WAVE - 1) / WAVE;
    const long long per_block = (long long)wpb * g_tiles_per_wave;
    long long g = (ntiles + per_block - 1) / per_block;
    return dim3((unsigned)(g < 1 ? 1 : g));
}

template <class T, int EQOP, bool ADJ, bool PB, bool XF>
static hipError_t launch_matvec_t(void *r, const void *a, const void *b, const int *index, int n,
                                  hipStream_t st)
{
    constexpr int W = shape_of<T>::WPB;
    hipLaunchKernelGGL((k_matvec<T, EQOP, ADJ, PB, XF, W>), grid_for(n, W), dim3(W * WAVE), 0, st, (cpx<T> *)r,
                       (const cpx<T> *)a, b, index, n, align_flags(a, b, r));
    return hipGetLastError();
}

template <class T, int EQOP, bool ADJ, int MU, int SIGN, bool XF>
static hipError_t launch_spproj_t(void *r, const void *a, const void *b, const int *index, int n,
                                  hipStream_t st)
{
    constexpr int W = shape_of<T>::WPB, MB = shape_of<T>::MINB;
    hipLaunchKernelGGL((k_spproj<T, EQOP, ADJ, MU, SIGN, XF, W, MB>), grid_for(n, W), dim3(W * WAVE), 0, st,
                       (cpx<T> *)r, (const cpx<T> *)a, (const cpx<T> *const *)b, index, n, align_flags(a, b, r));
    return hipGetLastError();
}

template <class T, int EQOP, bool ADJ, int MU, int SIGN, bool XF>
static hipError_t launch_sprecon_t(void *r, const void *a, const void *b, const int *index, int n,
                                   hipStream_t st)
{
    constexpr int W = shape_of<T>::WPB, MB = shape_of<T>::MINB;
    hipLaunchKernelGGL((k_sprecon<T, EQOP, ADJ, MU, SIGN, XF, W, MB>), grid_for(n, W), dim3(W * WAVE), 0, st,
                       (cpx<T> *)r, (const cpx<T> *)a, (const cpx<T> *const *)b, index, n, align_flags(a, b, r));
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

// Tiles per wave of the persistent site kernels (launch-shape knob).
void qla_set_tiles_per_wave(int tpw) { g_tiles_per_wave = tpw < 1 ? 1 : tpw; }

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
