// qla_multi.hip -- multi-source colour matrix times vector,
//   V_{eqop}_n{M,Ma,pM,pMa}_times_n{V,pV}:  r eqop sum_{d<nd} A_d * B_d
// (build_qla.pl:1693-1725; generic loop expressions_tensor.pl:983-1002; the
// D3 override lib/optimized/c/src/QLA_D3_V_vpeq_nM_times_npV.c keeps the same
// per-site order: d ascending, kc ascending inside each term).
//
// The x-forms (gang index) and the per-operand index forms of QLA_INDEX_ALL
// (one indexr / indexa / indexb array shared by all terms) address every
// record per lane.
//
// Compiled once per precision (QLA_SITE_PREC 0 float / 1 double); exports
// qla_multi_launch_p<PREC>.  A tile of 64 sites runs nd phases: each phase
// stages term d's matrix and vector records into LDS (direct-to-LDS loads),
// loads the next phase's gather addresses while they are in flight, then
// every lane adds its site's term to a register accumulator.  The result is
// written once per tile, coalesced through LDS.
#include "qla_tile_ops.hpp"

#ifndef QLA_SITE_PREC
#define QLA_SITE_PREC 0
#endif

#define QLA_CAT2_(a, b) a##b
#define QLA_CAT2(a, b) QLA_CAT2_(a, b)
#define QLA_MULTI_LAUNCH QLA_CAT2(qla_multi_launch_p, QLA_SITE_PREC)

namespace qla {

typedef std::conditional<QLA_SITE_PREC == 0, float, double>::type multi_real;

template <class T, int EQOP, bool ADJ, bool XF, int WPB>
__global__ void __launch_bounds__(WPB * WAVE) k_multi(multi_args p)
{
    typedef multi_op<T, EQOP, ADJ, XF> OP;
    typedef rec<T> R;
    __shared__ __attribute__((aligned(16))) char smem[WPB * OP::BYTES];
    char *lds = smem + (threadIdx.x / WAVE) * OP::BYTES;
    const int lane = lane_id();
    const int stride = gridDim.x * WPB;
    const int ntiles = tiles_of(p.n);
    int tile = blockIdx.x * WPB + (int)(threadIdx.x / WAVE);
    if (tile >= ntiles) return;

    bool act;
    const int i0 = lane_site(tile, lane, p.n, act);
    lane_sites site = OP::sites(p, i0, act);
    uint64_t ga, gb;
    OP::term_addresses(p, 0, site, act, ga, gb);
    vm_wait();      // the loop's only memory wait is the one before each phase's LDS reads
    for (;;) {
        const int base = tile * WAVE, ns = min(WAVE, p.n - base);
        const int next = tile + stride;
        const bool more = next < ntiles;
        const uint64_t gr = (XF && act) ? (uint64_t)((char *)p.r + (size_t)site.r * R::V) : 0;
        lane_sites nsite = { 0, 0, 0 };
        bool nact = false;
        cpx<T> x[3];
        for (int d = 0; d < p.nd; d++) {
            OP::stage_term(lds, p, d, ga, gb, base, ns, lane);
            if (OP::READ_R && d == 0) {
                if (!XF) load_tile<R::V>(lds + OP::OFF_R, (const char *)p.r + (size_t)base * R::V, ns, lane, p.al & AL_R);
                else gather_tile<R::V>(lds + OP::OFF_R, gr, ns, lane);
            }
            if (d == 0 && more) {
                const int i1 = lane_site(next, lane, p.n, nact);
                nsite = OP::sites(p, i1, nact);
            }
            // gather addresses of the next phase: term d + 1, or term 0 of the next tile
            if (d + 1 < p.nd) OP::term_addresses(p, d + 1, site, act, ga, gb);
            else if (more) OP::term_addresses(p, 0, nsite, nact, ga, gb);
            dma_wait();
            if (d == 0) {
                if (OP::READ_R) {
                    lds_read(x, lds + OP::OFF_R + lane * R::V);
                } else {
#pragma unroll
                    for (int c = 0; c < 3; c++) x[c].re = x[c].im = T(0);
                }
            }
            OP::accumulate(x, lds, lane);
            wave_sync();
        }
        char *lr = lds + (OP::READ_R ? OP::OFF_R : 0);
        lds_write(lr + lane * R::V, x);
        wave_sync();
        if (!XF) store_tile<R::V>((char *)p.r + (size_t)base * R::V, lr, ns, lane, p.al & AL_R);
        else scatter_tile<R::V>(gr, lr, ns, lane);
        wave_sync();
        if (!more) break;
        tile = next;
        site = nsite;
        act = nact;
    }
}

template <class T> struct multi_shape;
template <> struct multi_shape<float> { static constexpr int WPB = 2; };
template <> struct multi_shape<double> { static constexpr int WPB = 1; };

template <int EQOP, bool ADJ, bool XF>
static hipError_t launch_one(const multi_args &p, int tpw, hipStream_t st)
{
    typedef multi_real T;
    constexpr int W = multi_shape<T>::WPB;
    const long long ntiles = ((long long)p.n + WAVE - 1) / WAVE;
    const long long per_block = (long long)W * (tpw < 1 ? 1 : tpw);
    const long long g = (ntiles + per_block - 1) / per_block;
    hipLaunchKernelGGL((k_multi<T, EQOP, ADJ, XF, W>), dim3((unsigned)(g < 1 ? 1 : g)), dim3(W * WAVE), 0, st, p);
    return hipGetLastError();
}

template <int EQOP>
static hipError_t launch_eqop(int adj, int xf, const multi_args &p, int tpw, hipStream_t st)
{
    if (adj) return xf ? launch_one<EQOP, true, true>(p, tpw, st) : launch_one<EQOP, true, false>(p, tpw, st);
    return xf ? launch_one<EQOP, false, true>(p, tpw, st) : launch_one<EQOP, false, false>(p, tpw, st);
}

} // namespace qla

using namespace qla;

hipError_t QLA_MULTI_LAUNCH(int eqop, int adj, int xf, const qla_multi_args &p, int tpw, hipStream_t st)
{
    if (p.nd < 1 || p.nd > MULTI_MAXND) return hipErrorInvalidValue;
    switch (eqop) {
    case EQ: return launch_eqop<EQ>(adj, xf, p, tpw, st);
    case PEQ: return launch_eqop<PEQ>(adj, xf, p, tpw, st);
    case MEQ: return launch_eqop<MEQ>(adj, xf, p, tpw, st);
    case EQM: return launch_eqop<EQM>(adj, xf, p, tpw, st);
    }
    return hipErrorInvalidValue;
}
