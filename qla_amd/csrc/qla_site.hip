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
#include "qla_tile_ops.hpp"
#include "qla_launch.hpp"

namespace qla {

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
// Per-op kernels: persistent loops over tiles.  The gather address (pointer
// array entry, or index entry for the x-form) of the next tile is loaded
// between the current tile's stage and finish steps.
// ---------------------------------------------------------------------------
template <class OP, class T, bool XF, bool PB, int WPB>
__device__ __forceinline__ void run_tiles(cpx<T> *__restrict__ r, const cpx<T> *__restrict__ a,
                                          const void *__restrict__ b, const int *__restrict__ index, int n, int al)
{
    __shared__ __attribute__((aligned(16))) char smem[WPB * OP::BYTES];
    char *lds = smem + (threadIdx.x / WAVE) * OP::BYTES;
    tiles t(WPB, n);
    if (t.tile >= t.ntiles) return;
    int site = t.base() + t.lane;
    bool act = t.lane < t.nsites();
    if (XF) site = act ? index[site] : 0;
    uint64_t pb = (PB && act) ? (uint64_t)((const void *const *)b)[site] : 0;
    for (;;) {
        const int base = t.base(), ns = t.nsites();
        OP::stage(lds, r, a, b, pb, base, ns, t.lane, site, act, al);
        const int next = t.tile + t.stride;
        int nsite = 0;
        uint64_t npb = 0;
        bool nact = false;
        if (next < t.ntiles) {
            nsite = next * WAVE + t.lane;
            nact = t.lane < min(WAVE, n - next * WAVE);
            if (XF) nsite = nact ? index[nsite] : 0;
            else if (PB) npb = nact ? (uint64_t)((const void *const *)b)[nsite] : 0;
        }
        OP::finish(lds, r, base, ns, t.lane, site, act, al);
        if (next >= t.ntiles) break;
        t.tile = next;
        site = nsite;
        act = nact;
        if (XF && PB) npb = nact ? (uint64_t)((const void *const *)b)[nsite] : 0;
        pb = npb;
    }
}

template <class T, int EQOP, bool ADJ, bool PB, bool XF, int WPB>
__global__ void __launch_bounds__(WPB * WAVE)
k_matvec(cpx<T> *__restrict__ r, const cpx<T> *__restrict__ a, const void *__restrict__ b,
         const int *__restrict__ index, int n, int al)
{
    run_tiles<matvec_op<T, EQOP, ADJ, PB, XF>, T, XF, PB, WPB>(r, a, b, index, n, al);
}

template <class T, int EQOP, bool ADJ, int MU, int SIGN, bool XF, int WPB, int MINB>
__global__ void __launch_bounds__(WPB * WAVE, MINB)
k_spproj(cpx<T> *__restrict__ r, const cpx<T> *__restrict__ a, const cpx<T> *const *__restrict__ b,
         const int *__restrict__ index, int n, int al)
{
    run_tiles<spproj_op<T, EQOP, ADJ, MU, SIGN, XF>, T, XF, true, WPB>(r, a, b, index, n, al);
}

template <class T, int EQOP, bool ADJ, int MU, int SIGN, bool XF, int WPB, int MINB>
__global__ void __launch_bounds__(WPB * WAVE, MINB)
k_sprecon(cpx<T> *__restrict__ r, const cpx<T> *__restrict__ a, const cpx<T> *const *__restrict__ b,
          const int *__restrict__ index, int n, int al)
{
    run_tiles<sprecon_op<T, EQOP, ADJ, MU, SIGN, XF>, T, XF, true, WPB>(r, a, b, index, n, al);
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
