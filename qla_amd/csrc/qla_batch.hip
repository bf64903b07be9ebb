// qla_batch.hip -- ragged batches of site operations in one persistent launch.
//
// A batch is a table of independent jobs, each one call of
//   V_vpeq_M_times_pV, H_veq_spproj_M_times_pD or D_vpeq_sprecon_M_times_pH
// (QLA_F3 semantics, v-form) with its own n, mu, sign and operands.  Every
// job is cut into 64-site tiles; the tiles of all jobs form one index space
// (job j owns tiles [tile0_j, tile0_{j+1})).  Wavefronts of a persistent
// grid claim `tpc` consecutive tiles at a time from a global counter, find
// the owning job by binary search over tile0, and run the same per-tile
// stage/finish steps as the single-op kernels (qla_tile_ops.hpp), so each
// job's result is bit-identical to its plain call.  Dynamic claiming balances
// jobs of uneven size; the launch shape (waves per block, tiles per claim,
// blocks per CU) is chosen by the autotuner in qla_capi.hip.
#include "qla_tile_ops.hpp"
#include "qla_launch.hpp"

namespace qla {

// device job table entry (matches qla_batch_job in qla_launch.hpp)
typedef qla_batch_job batch_job;

typedef matvec_op<float, PEQ, false, true, false> bop_matvec;
template <int MU, int SIGN> using bop_spproj = spproj_op<float, EQ, false, MU, SIGN, false>;
template <int MU, int SIGN> using bop_sprecon = sprecon_op<float, PEQ, false, MU, SIGN, false>;

constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int BATCH_BYTES = cmax(bop_matvec::BYTES, cmax(bop_spproj<0, 1>::BYTES, bop_sprecon<0, 1>::BYTES));

template <class OP>
__device__ __forceinline__ void run_op(char *lds, const batch_job &j, uint64_t pb, int base, int ns, int lane)
{
    OP::stage(lds, (cpx<float> *)j.r, (const cpx<float> *)j.a, (const void *)j.b, pb, base, ns, lane, base + lane,
              lane < ns, j.al);
    OP::finish(lds, (cpx<float> *)j.r, base, ns, lane, base + lane, lane < ns, j.al);
}

template <template <int, int> class OPT>
__device__ __forceinline__ void run_musign(char *lds, const batch_job &j, uint64_t pb, int base, int ns, int lane)
{
    if (j.sign == 1) {
        switch (j.mu) {
        case 0: run_op<OPT<0, 1>>(lds, j, pb, base, ns, lane); break;
        case 1: run_op<OPT<1, 1>>(lds, j, pb, base, ns, lane); break;
        case 2: run_op<OPT<2, 1>>(lds, j, pb, base, ns, lane); break;
        case 3: run_op<OPT<3, 1>>(lds, j, pb, base, ns, lane); break;
        case 4: run_op<OPT<4, 1>>(lds, j, pb, base, ns, lane); break;
        }
    } else {
        switch (j.mu) {
        case 0: run_op<OPT<0, -1>>(lds, j, pb, base, ns, lane); break;
        case 1: run_op<OPT<1, -1>>(lds, j, pb, base, ns, lane); break;
        case 2: run_op<OPT<2, -1>>(lds, j, pb, base, ns, lane); break;
        case 3: run_op<OPT<3, -1>>(lds, j, pb, base, ns, lane); break;
        case 4: run_op<OPT<4, -1>>(lds, j, pb, base, ns, lane); break;
        }
    }
}

// owning job of `tile`: the last j with tile0_j <= tile (jobs with no tiles
// have tile0_j == tile0_{j+1} and are never selected)
__device__ __forceinline__ int find_job(const batch_job *jobs, int njobs, int tile)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int WPB>
__global__ void __launch_bounds__(WPB * WAVE)
k_batch(const batch_job *__restrict__ jobs, int njobs, int total_tiles, int tpc, unsigned int *__restrict__ counter)
{
    __shared__ __attribute__((aligned(16))) char smem[WPB * BATCH_BYTES];
    char *lds = smem + (threadIdx.x / WAVE) * BATCH_BYTES;
    const int lane = lane_id();
    for (;;) {
        unsigned int claim = 0;
        if (lane == 0) claim = atomicAdd(counter, 1u);
        claim = __shfl(claim, 0, WAVE);
        const long long first = (long long)claim * tpc;
        if (first >= total_tiles) break;                 // every wave ends here
        const int last = (int)min((long long)total_tiles, first + tpc);
        for (int tile = (int)first; tile < last; tile++) {
            const batch_job j = jobs[find_job(jobs, njobs, tile)];
            const int base = (tile - j.tile0) * WAVE;
            const int ns = min(WAVE, j.n - base);
            const uint64_t pb = lane < ns ? (uint64_t)j.b[base + lane] : 0;
            switch (j.op) {
            case QLA_BOP_MATVEC_PV: run_op<bop_matvec>(lds, j, pb, base, ns, lane); break;
            case QLA_BOP_SPPROJ: run_musign<bop_spproj>(lds, j, pb, base, ns, lane); break;
            case QLA_BOP_SPRECON: run_musign<bop_sprecon>(lds, j, pb, base, ns, lane); break;
            }
        }
    }
}

__global__ void k_batch_zero(unsigned int *c) { *c = 0u; }

} // namespace qla

using namespace qla;

int qla_batch_lds_bytes_per_wave(void) { return BATCH_BYTES; }

hipError_t qla_launch_batch(const qla_batch_job *jobs, int njobs, int total_tiles, int wpb, int tpc, int blocks,
                            unsigned int *counter, hipStream_t st)
{
    if (total_tiles <= 0) return hipSuccess;
    if (tpc < 1 || blocks < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_batch_zero, dim3(1), dim3(1), 0, st, counter);
    switch (wpb) {
    case 1: hipLaunchKernelGGL(k_batch<1>, dim3(blocks), dim3(1 * WAVE), 0, st, jobs, njobs, total_tiles, tpc, counter); break;
    case 2: hipLaunchKernelGGL(k_batch<2>, dim3(blocks), dim3(2 * WAVE), 0, st, jobs, njobs, total_tiles, tpc, counter); break;
    case 4: hipLaunchKernelGGL(k_batch<4>, dim3(blocks), dim3(4 * WAVE), 0, st, jobs, njobs, total_tiles, tpc, counter); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
