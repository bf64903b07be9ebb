// qla_precision.hip -- precision conversion of whole fields, the QLA_DF3 /
// QLA_FD3 "identity" functions T_{v,x}eq_[p]T (build_qla.pl:530-585): every
// complex component converted float -> double (exact) or double -> float
// (the C conversion: round to nearest even), records of V, H, D, M or P.
//
// v-form with a field source: one complex per thread over the contiguous
// arrays, non-temporal 8/16-byte loads and stores (each element is touched
// once).  x-forms, per-operand index forms and pointer sources: the lanes of
// a wave walk the 64 records of a tile one complex at a time, reading the
// source record through its gathered address and writing the destination
// record (r[indexr[i]], r[index[i]] or r[i]).
#include "qla_dev.hpp"
#include "qla_launch.hpp"

namespace qla {

constexpr int PBLOCK = 256;

template <class T> struct vec2;
template <> struct vec2<float> { typedef float v __attribute__((ext_vector_type(2))); };
template <> struct vec2<double> { typedef double v __attribute__((ext_vector_type(2))); };

template <class TI, class TO>
__global__ void __launch_bounds__(PBLOCK) k_precision_contig(TO *__restrict__ r, const TI *__restrict__ a, long long m)
{
    typedef typename vec2<TI>::v VI;
    typedef typename vec2<TO>::v VO;
    const long long stride = (long long)gridDim.x * PBLOCK;
    for (long long i = (long long)blockIdx.x * PBLOCK + threadIdx.x; i < m; i += stride) {
        const VI x = __builtin_nontemporal_load((const VI *)a + i);
        VO y;
        y.x = (TO)x.x;
        y.y = (TO)x.y;
        __builtin_nontemporal_store(y, (VO *)r + i);
    }
}

template <class TI, class TO, int NC>
__global__ void __launch_bounds__(PBLOCK)
k_precision_gather(TO *__restrict__ r, const void *__restrict__ a, const int *__restrict__ index,
                   const int *__restrict__ ir, const int *__restrict__ ia, int n, int pa)
{
    const int lane = threadIdx.x % WAVE;
    const int nwaves = gridDim.x * (PBLOCK / WAVE);
    const int ntiles = tiles_of(n);
    for (int tile = blockIdx.x * (PBLOCK / WAVE) + (int)(threadIdx.x / WAVE); tile < ntiles; tile += nwaves) {
        bool act;
        const int i = lane_site(tile, lane, n, act);
        const int ns = min(WAVE, n - tile * WAVE);
        uint64_t ga = 0, gr = 0;
        if (act) {
            const int g = index ? index[i] : i;
            const int sa = ia ? ia[i] : g;
            const int sr = ir ? ir[i] : g;
            ga = pa ? (uint64_t)((const void *const *)a)[sa] : (uint64_t)((const cpx<TI> *)a + (size_t)sa * NC);
            gr = (uint64_t)(r + (size_t)sr * NC * 2);
        }
        const int tot = ns * NC;
        const int iters = (tot + WAVE - 1) / WAVE;
        for (int k = 0; k < iters; k++) {
            const int c = k * WAVE + lane;
            const int sl = c / NC;
            const int comp = c - sl * NC;
            const int src = sl < WAVE ? sl : 0;
            const uint64_t xa = shfl64(ga, src);
            const uint64_t xr = shfl64(gr, src);
            if (c < tot) {
                const cpx<TI> x = ((const cpx<TI> *)xa)[comp];
                cpx<TO> y;
                y.re = (TO)x.re;
                y.im = (TO)x.im;
                ((cpx<TO> *)xr)[comp] = y;
            }
        }
    }
}

template <class TI, class TO>
static hipError_t launch_gather(int ncomp, void *r, const void *a, const int *index, const int *ir, const int *ia,
                                int n, int pa, hipStream_t st)
{
    const long long ntiles = ((long long)n + WAVE - 1) / WAVE;
    const long long want = (ntiles + 2 * (PBLOCK / WAVE) - 1) / (2 * (PBLOCK / WAVE));
    const int g = (int)(want < 1 ? 1 : (want > 65535 ? 65535 : want));
#define QLA_PREC_CASE(NC)                                                                                    \
    case NC:                                                                                                \
        hipLaunchKernelGGL((k_precision_gather<TI, TO, NC>), dim3(g), dim3(PBLOCK), 0, st, (TO *)r, a, \
                           index, ir, ia, n, pa);                                                       \
        return hipGetLastError();
    switch (ncomp) {
        QLA_PREC_CASE(3)
        QLA_PREC_CASE(6)
        QLA_PREC_CASE(9)
        QLA_PREC_CASE(12)
        QLA_PREC_CASE(144)
    }
#undef QLA_PREC_CASE
    return hipErrorInvalidValue;
}

template <class TI, class TO>
static hipError_t launch_t(int ncomp, void *r, const void *a, const int *index, const int *ir, const int *ia, int n,
                           int pa, hipStream_t st)
{
    if (!index && !ir && !ia && !pa) {
        const long long m = (long long)n * ncomp;
        const long long want = (m + 4LL * PBLOCK - 1) / (4LL * PBLOCK);
        const int g = (int)(want < 1 ? 1 : (want > 65535 ? 65535 : want));
        hipLaunchKernelGGL((k_precision_contig<TI, TO>), dim3(g), dim3(PBLOCK), 0, st, (TO *)r, (const TI *)a, m);
        return hipGetLastError();
    }
    return launch_gather<TI, TO>(ncomp, r, a, index, ir, ia, n, pa, st);
}

} // namespace qla

using namespace qla;

hipError_t qla_launch_precision(int to_float, int ncomp, void *r, const void *a, const int *index, const int *ir,
                                const int *ia, int n, int pa, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (to_float) return launch_t<double, float>(ncomp, r, a, index, ir, ia, n, pa, st);
    return launch_t<float, double>(ncomp, r, a, index, ir, ia, n, pa, st);
}
