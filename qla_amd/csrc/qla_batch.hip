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

// stage / finish of the job's op for one tile (the switch is wave-uniform)
template <class OP>
__device__ __forceinline__ void op_step(bool fin, char *lds, const batch_job &j, uint64_t pb, int base, int ns,
                                        int lane)
{
    if (!fin)
        OP::stage(lds, (cpx<float> *)j.r, (const cpx<float> *)j.a, (const void *)j.b, pb, base, ns, lane, base + lane,
                  lane < ns, j.al);
    else
        OP::finish(lds, (cpx<float> *)j.r, base, ns, lane, base + lane, lane < ns, j.al);
}

template <template <int, int> class OPT>
__device__ __forceinline__ void musign_step(bool fin, char *lds, const batch_job &j, uint64_t pb, int base, int ns,
                                            int lane)
{
    if (j.sign == 1) {
        switch (j.mu) {
        case 0: op_step<OPT<0, 1>>(fin, lds, j, pb, base, ns, lane); break;
        case 1: op_step<OPT<1, 1>>(fin, lds, j, pb, base, ns, lane); break;
        case 2: op_step<OPT<2, 1>>(fin, lds, j, pb, base, ns, lane); break;
        case 3: op_step<OPT<3, 1>>(fin, lds, j, pb, base, ns, lane); break;
        case 4: op_step<OPT<4, 1>>(fin, lds, j, pb, base, ns, lane); break;
        }
    } else {
        switch (j.mu) {
        case 0: op_step<OPT<0, -1>>(fin, lds, j, pb, base, ns, lane); break;
        case 1: op_step<OPT<1, -1>>(fin, lds, j, pb, base, ns, lane); break;
        case 2: op_step<OPT<2, -1>>(fin, lds, j, pb, base, ns, lane); break;
        case 3: op_step<OPT<3, -1>>(fin, lds, j, pb, base, ns, lane); break;
        case 4: op_step<OPT<4, -1>>(fin, lds, j, pb, base, ns, lane); break;
        }
    }
}

__device__ __forceinline__ void job_step(bool fin, char *lds, const batch_job &j, uint64_t pb, int base, int ns,
                                         int lane)
{
    switch (j.op) {
    case QLA_BOP_MATVEC_PV: op_step<bop_matvec>(fin, lds, j, pb, base, ns, lane); break;
    case QLA_BOP_SPPROJ: musign_step<bop_spproj>(fin, lds, j, pb, base, ns, lane); break;
    case QLA_BOP_SPRECON: musign_step<bop_sprecon>(fin, lds, j, pb, base, ns, lane); break;
    }
}

// owning job of `tile`: the last j with tile0_j <= tile (jobs with no tiles
// have tile0_j == tile0_{j+1} and are never selected)
// wave-uniform value in a scalar register (the argument is uniform already;
// this lets the compiler keep job fields and tile indices in SGPRs)
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// t0[k * stride] is job k's first tile
__device__ __forceinline__ int find_job(const int *t0, int stride, int njobs, int tile)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (uni(t0[mid * stride]) <= tile) lo = mid;
        else hi = mid - 1;
    }
    return uni(lo);
}

struct tile_ref {
    batch_job j;
    int base, ns;
    uint64_t pb;
};

__device__ __forceinline__ void load_tile_ref(tile_ref &t, const batch_job *jobs, const int *t0, int stride, int njobs,
                                              int tile, int lane)
{
    t.j = jobs[find_job(t0, stride, njobs, tile)];
    t.base = (tile - t.j.tile0) * WAVE;
    t.ns = min(WAVE, t.j.n - t.base);
    t.pb = lane < t.ns ? (uint64_t)t.j.b[t.base + lane] : 0;
}

// Waves claim `tpc` consecutive tiles at a time.  Each wave holds one claim
// ahead (issued when it starts on the current one), and the next tile's job
// entry and pointer-array entries are loaded while the current tile's
// records are in flight.  A wave stops when a claim starts past the last
// tile; claims are handed out in increasing order, so every later claim is
// past it too and no work is dropped.
template <int WPB>
__global__ void __launch_bounds__(WPB * WAVE)
k_batch(const batch_job *__restrict__ jobs, int njobs, int total_tiles, int tpc, unsigned int *__restrict__ counter)
{
    __shared__ __attribute__((aligned(16))) char smem[WPB * BATCH_BYTES];
    // job lookup: tile0 fields of the job table, read with wave-uniform
    // indices (scalar loads through the scalar cache)
    const int *t0 = &jobs[0].tile0;
    const int t0_stride = (int)(sizeof(batch_job) / sizeof(int));
    char *lds = smem + (threadIdx.x / WAVE) * BATCH_BYTES;
    const int lane = lane_id();

    unsigned int c0 = 0, pend = 0;
    if (lane == 0) {
        c0 = atomicAdd(counter, 1u);
        pend = atomicAdd(counter, 1u);
    }
    const long long first = (long long)uni((int)c0) * tpc;
    if (first >= total_tiles) return;
    int tile = (int)first;
    int last = (int)min((long long)total_tiles, first + tpc);
    tile_ref cur, nxt;
    load_tile_ref(cur, jobs, t0, t0_stride, njobs, tile, lane);
    for (;;) {
        job_step(false, lds, cur.j, cur.pb, cur.base, cur.ns, lane);
        int ntile = tile + 1, nlast = last;
        if (ntile >= last) {                       // move on to the claim held ahead
            const long long nf = (long long)uni((int)pend) * tpc;
            if (lane == 0) pend = atomicAdd(counter, 1u);
            ntile = nf < total_tiles ? (int)nf : total_tiles;
            nlast = (int)min((long long)total_tiles, nf + tpc);
        }
        const bool more = ntile < total_tiles;
        if (more) load_tile_ref(nxt, jobs, t0, t0_stride, njobs, ntile, lane);
        job_step(true, lds, cur.j, cur.pb, cur.base, cur.ns, lane);
        if (!more) break;
        cur = nxt;
        tile = ntile;
        last = nlast;
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
