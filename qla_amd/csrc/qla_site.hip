// qla_site.hip -- gfx950 kernels for the QLA Wilson-hop building blocks:
//   V_{eqop}_M[a]_times_V          (colour matrix times colour vector)
//   H_{eqop}_spproj_M[a]_times_D   (spin-project, then multiply)
//   D_{eqop}_sprecon_M[a]_times_H  (multiply, then spin-reconstruct)
//   D_{eqop}_spproj_M[a]_times_D   (project, multiply, reconstruct)
// in every eqop (eq, peq, meq, eqm), with M or its adjoint, with either
// source a field or a pointer array (pM, pV/pD/pH), in the v-form (site i),
// the gang-indexed x-form (site index[i]) and the per-operand index forms
// of QLA_INDEX_ALL (r[indexr[i]], a[indexa[i]], b[indexb[i]]).
//
// and the matrix-free spin operations H_{eqop}_spproj_D,
// D_{eqop}_sprecon_H and D_{eqop}_spproj_D (build_qla.pl:1087-1160).
//
// This file is compiled once per (precision, kind): QLA_SITE_PREC 0 float /
// 1 double, QLA_SITE_KIND 0..3 (K_MATVEC, K_SPPROJ, K_SPRECON, K_WILSON) or
// 5..7 (K_PROJ, K_RECON, K_PROJRECON) or 8..11 (the colour-matrix products
// of H, D, M and M^dagger); each object holds the eqop x adjoint
// x form kernels of its pair (no adjoint for the matrix-free kinds) and
// exports qla_site_launch_p<PREC>_k<KIND>.
//
// One wavefront owns one tile of 64 sites.  Records move HBM -> LDS with
// direct-to-LDS loads (coalesced for contiguous fields and for adjacent
// gather targets), each lane then takes its own site's records from LDS,
// computes in registers in the reference's operation order, writes the
// result record back to LDS and the wave stores it coalesced.  All kernels
// are HBM-bandwidth bound (0.3-0.7 flop/byte), so no MFMA is involved.
#include "qla_tile_ops.hpp"

#ifndef QLA_SITE_PREC
#define QLA_SITE_PREC 0
#endif
#ifndef QLA_SITE_KIND
#define QLA_SITE_KIND 1
#endif

#define QLA_CAT4_(a, b, c, d) a##b##c##d
#define QLA_CAT4(a, b, c, d) QLA_CAT4_(a, b, c, d)
#define QLA_SITE_LAUNCH QLA_CAT4(qla_site_launch_p, QLA_SITE_PREC, _k, QLA_SITE_KIND)

namespace qla {

typedef std::conditional<QLA_SITE_PREC == 0, float, double>::type site_real;
constexpr int SITE_KIND = QLA_SITE_KIND;

// Persistent tile loop: wave w of the grid takes tiles w, w + W, w + 2W, ...
// (W = waves in the grid).  The next tile's gather addresses (pointer-array
// entries) are loaded between the current tile's stage and finish steps so
// they travel with its records; in the XF forms (gang index, per-operand
// indices) the index entries run one tile further ahead, so the pointer
// loads never wait on an index load.
template <class OP, class T, int KIND, bool XF, int WPB>
__device__ __forceinline__ void run_tiles(const site_args &p)
{
    __shared__ __attribute__((aligned(16))) char smem[WPB * OP::BYTES];
    char *lds = smem + (threadIdx.x / WAVE) * OP::BYTES;
    const int lane = lane_id();
    const int stride = gridDim.x * WPB;
    const int ntiles = tiles_of(p.n);
    int tile = blockIdx.x * WPB + (int)(threadIdx.x / WAVE);
    if (tile >= ntiles) return;

    bool act0;
    const int i0 = lane_site(tile, lane, p.n, act0);
    lane_addr cur = addresses_of<T, KIND, XF>(p, sites_of<XF>(p, i0, act0), act0);
    int next = tile + stride;
    lane_sites nsite = { 0, 0, 0 };     // XF forms: this lane's records in the next tile
    bool nact = false;
    if (XF && next < ntiles) {
        const int i1 = lane_site(next, lane, p.n, nact);
        nsite = sites_of<XF>(p, i1, nact);
    }
    // the loop body then never waits on a prologue load: its only memory
    // wait is the one in finish.  (Issuing the first tile's contiguous
    // records before this wait measured 0.5% faster for spproj but 1.2%
    // slower for sprecon and 2.2% for the D3 spproj, so it is not done.)
    vm_wait();
    for (;;) {
        const int base = tile * WAVE, ns = min(WAVE, p.n - base);
        OP::stage(lds, p, cur, base, ns, lane);
        lane_addr nxt = { 0, 0, 0 };
        lane_sites nnsite = { 0, 0, 0 };
        bool nnact = false;
        if (next < ntiles) {
            if (!XF) {
                const int i1 = lane_site(next, lane, p.n, nact);
                nsite = sites_of<XF>(p, i1, nact);
            }
            nxt = addresses_of<T, KIND, XF>(p, nsite, nact);
            const int nn = next + stride;
            if (XF && nn < ntiles) {
                const int i2 = lane_site(nn, lane, p.n, nnact);
                nnsite = sites_of<XF>(p, i2, nnact);
            }
        }
        OP::finish(lds, p, cur, base, ns, lane);
        if (next >= ntiles) break;
        tile = next;
        next = tile + stride;
        cur = nxt;
        if (XF) {
            nsite = nnsite;
            nact = nnact;
        }
    }
}

template <class T, int KIND, int EQOP, bool ADJ, bool XF, int WPB, bool SSEO = false>
__global__ void __launch_bounds__(WPB * WAVE) k_site(site_args p)
{
    run_tiles<site_op<T, KIND, EQOP, ADJ, XF, -1, SSEO>, T, KIND, XF, WPB>(p);
}

// one wave per block: a block's LDS is released when its last wave ends, so
// single-wave blocks return each tile's LDS as soon as that tile is done.
// Measured on one device against two-wave blocks: on 32^4 F3, spproj
// 37.31 against 37.67 us and sprecon 53.24 against 53.38 us; on 48^4 the
// other way round (195.4 against 194.1 us, 276.4 against 271.7 us).  One
// wave per block serves the 32^4 calls.  The grid covers the tiles with
// about `tpw` tiles per wave.
template <class T> struct site_shape { static constexpr int WPB = 1; };

template <int EQOP, bool ADJ, bool XF>
static hipError_t launch_one(const site_args &p, int tpw, hipStream_t st)
{
    typedef site_real T;
    constexpr int W = site_shape<T>::WPB;
    const long long ntiles = ((long long)p.n + WAVE - 1) / WAVE;
    const long long per_block = (long long)W * (tpw < 1 ? 1 : tpw);
    const long long g = (ntiles + per_block - 1) / per_block;
    const dim3 grid((unsigned)(g < 1 ? 1 : g)), block(W * WAVE);
    if constexpr (sse_order<T, SITE_KIND, EQOP, ADJ, XF>::v) {
        if (p.order) {
            hipLaunchKernelGGL((k_site<T, SITE_KIND, EQOP, ADJ, XF, W, true>), grid, block, 0, st, p);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_site<T, SITE_KIND, EQOP, ADJ, XF, W>), grid, block, 0, st, p);
    return hipGetLastError();
}

template <int EQOP>
static hipError_t launch_eqop(int adj, int xf, const site_args &p, int tpw, hipStream_t st)
{
    if constexpr (kind_unary<SITE_KIND>::v) {
        if (adj) return hipErrorInvalidValue;
    } else {
        if (adj) return xf ? launch_one<EQOP, true, true>(p, tpw, st) : launch_one<EQOP, true, false>(p, tpw, st);
    }
    return xf ? launch_one<EQOP, false, true>(p, tpw, st) : launch_one<EQOP, false, false>(p, tpw, st);
}

} // namespace qla

using namespace qla;

hipError_t QLA_SITE_LAUNCH(int eqop, int adj, int xf, const qla_site_args &p, int tpw, hipStream_t st)
{
    switch (eqop) {
    case EQ: return launch_eqop<EQ>(adj, xf, p, tpw, st);
    case PEQ: return launch_eqop<PEQ>(adj, xf, p, tpw, st);
    case MEQ: return launch_eqop<MEQ>(adj, xf, p, tpw, st);
    case EQM: return launch_eqop<EQM>(adj, xf, p, tpw, st);
    }
    return hipErrorInvalidValue;
}
