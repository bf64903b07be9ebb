// qla_batch.hip -- ragged batches of site operations in one persistent launch.
//
// A batch is a table of independent jobs, each one call of
//   V_vpeq_M_times_pV, H_veq_spproj_M_times_pD or D_vpeq_sprecon_M_times_pH
// (QLA_F3 semantics, v-form) with its own n, mu, sign and operands.  Every
// job is cut into 64-site tiles; the tiles of all jobs form one index space
// (job j owns tiles [tile0_j, tile0_{j+1})).  Each wavefront takes `tpc`
// consecutive tiles at a time -- from a global claim counter in a persistent
// grid (dynamic schedule), or by its position in a grid that covers all
// tiles (static schedule) -- finds the owning job by binary search over
// tile0, and runs the same per-tile stage/finish steps as the single-op
// kernels (qla_tile_ops.hpp), so each job's result is bit-identical to its
// plain call.  The launch shape (waves per block, tiles per claim, blocks per
// CU, 0 for the static schedule) is chosen by the autotuner in qla_capi.hip.
#include "qla_tile_ops.hpp"
#include "qla_launch.hpp"

namespace qla {

// device job table entry (matches qla_batch_job in qla_launch.hpp)
typedef qla_batch_job batch_job;

typedef site_op<float, K_MATVEC, PEQ, false, false> bop_matvec;
typedef site_op<float, K_SPPROJ, EQ, false, false> bop_spproj;
typedef site_op<float, K_SPRECON, PEQ, false, false> bop_sprecon;

constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int BATCH_BYTES = cmax(bop_matvec::BYTES, cmax(bop_spproj::BYTES, bop_sprecon::BYTES));

// operands of a job in the form of the per-call kernels (v-form, b a pointer array)
__device__ __forceinline__ site_args job_args(const batch_job &j)
{
    site_args p;
    p.r = j.r;
    p.a = j.a;
    p.b = (const void *)j.b;
    p.index = nullptr;
    p.ir = p.ia = p.ib = nullptr;
    p.n = j.n;
    p.al = j.al;
    p.order = 0;
    p.ms = j.ms;
    p.pa = 0;
    p.pb = 1;
    return p;
}

// stage / finish of the job's op for one tile (the switch is wave-uniform)
template <class OP>
__device__ __forceinline__ void op_step(bool fin, char *lds, const site_args &p, const lane_addr &la, int base,
                                        int ns, int lane)
{
    if (!fin) OP::stage(lds, p, la, base, ns, lane);
    else OP::finish(lds, p, la, base, ns, lane);
}

__device__ __forceinline__ void job_step(bool fin, char *lds, int op, const site_args &p, const lane_addr &la,
                                         int base, int ns, int lane)
{
    switch (op) {
    case QLA_BOP_MATVEC_PV: op_step<bop_matvec>(fin, lds, p, la, base, ns, lane); break;
    case QLA_BOP_SPPROJ: op_step<bop_spproj>(fin, lds, p, la, base, ns, lane); break;
    case QLA_BOP_SPRECON: op_step<bop_sprecon>(fin, lds, p, la, base, ns, lane); break;
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
    site_args p;
    lane_addr la;
    int op, base, ns;
};

__device__ __forceinline__ void load_tile_ref(tile_ref &t, const batch_job *jobs, const int *t0, int stride, int njobs,
                                              int tile, int lane)
{
    const batch_job j = jobs[find_job(t0, stride, njobs, tile)];
    t.p = job_args(j);
    t.op = j.op;
    t.base = (tile - j.tile0) * WAVE;
    t.ns = min(WAVE, j.n - t.base);
    t.la.ga = 0;
    t.la.gr = 0;
    t.la.gb = lane < t.ns ? (uint64_t)j.b[t.base + lane] : 0;
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
    vm_wait();      // the loop's only memory wait is the one in finish
    for (;;) {
        job_step(false, lds, cur.op, cur.p, cur.la, cur.base, cur.ns, lane);
        int ntile = tile + 1, nlast = last;
        if (ntile >= last) {                       // move on to the claim held ahead
            const long long nf = (long long)uni((int)pend) * tpc;
            if (lane == 0) pend = atomicAdd(counter, 1u);
            ntile = nf < total_tiles ? (int)nf : total_tiles;
            nlast = (int)min((long long)total_tiles, nf + tpc);
        }
        const bool more = ntile < total_tiles;
        if (more) load_tile_ref(nxt, jobs, t0, t0_stride, njobs, ntile, lane);
        job_step(true, lds, cur.op, cur.p, cur.la, cur.base, cur.ns, lane);
        if (!more) break;
        cur = nxt;
        tile = ntile;
        last = nlast;
    }
}

// Static schedule (no counter): the grid covers every tile, wave w runs the
// `tpc` consecutive tiles starting at w * tpc, and the hardware dispatcher
// balances uneven jobs by starting a block wherever one retires.  A claim
// counter costs one same-address atomic per claim; those serialise across the
// XCDs at roughly 70 million per second, which bounds the dynamic schedule
// on large batches.
template <int WPB>
__global__ void __launch_bounds__(WPB * WAVE)
k_batch_static(const batch_job *__restrict__ jobs, int njobs, int total_tiles, int tpc)
{
    __shared__ __attribute__((aligned(16))) char smem[WPB * BATCH_BYTES];
    const int *t0 = &jobs[0].tile0;
    const int t0_stride = (int)(sizeof(batch_job) / sizeof(int));
    char *lds = smem + (threadIdx.x / WAVE) * BATCH_BYTES;
    const int lane = lane_id();
    const long long first = ((long long)blockIdx.x * WPB + (long long)(threadIdx.x / WAVE)) * tpc;
    if (first >= total_tiles) return;
    int tile = uni((int)first);
    const int last = (int)min((long long)total_tiles, first + tpc);
    tile_ref cur, nxt;
    load_tile_ref(cur, jobs, t0, t0_stride, njobs, tile, lane);
    vm_wait();
    for (;;) {
        job_step(false, lds, cur.op, cur.p, cur.la, cur.base, cur.ns, lane);
        const int ntile = tile + 1;
        const bool more = ntile < last;
        if (more) load_tile_ref(nxt, jobs, t0, t0_stride, njobs, ntile, lane);
        job_step(true, lds, cur.op, cur.p, cur.la, cur.base, cur.ns, lane);
        if (!more) break;
        cur = nxt;
        tile = ntile;
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
    if (tpc < 1 || blocks < 0) return hipErrorInvalidValue;
    if (blocks == 0) {          // static schedule
        const long long per_block = (long long)wpb * tpc;
        const unsigned g = (unsigned)(((long long)total_tiles + per_block - 1) / per_block);
        switch (wpb) {
        case 1: hipLaunchKernelGGL(k_batch_static<1>, dim3(g), dim3(1 * WAVE), 0, st, jobs, njobs, total_tiles, tpc); break;
        case 2: hipLaunchKernelGGL(k_batch_static<2>, dim3(g), dim3(2 * WAVE), 0, st, jobs, njobs, total_tiles, tpc); break;
        case 4: hipLaunchKernelGGL(k_batch_static<4>, dim3(g), dim3(4 * WAVE), 0, st, jobs, njobs, total_tiles, tpc); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_batch_zero, dim3(1), dim3(1), 0, st, counter);
    switch (wpb) {
    case 1: hipLaunchKernelGGL(k_batch<1>, dim3(blocks), dim3(1 * WAVE), 0, st, jobs, njobs, total_tiles, tpc, counter); break;
    case 2: hipLaunchKernelGGL(k_batch<2>, dim3(blocks), dim3(2 * WAVE), 0, st, jobs, njobs, total_tiles, tpc, counter); break;
    case 4: hipLaunchKernelGGL(k_batch<4>, dim3(blocks), dim3(4 * WAVE), 0, st, jobs, njobs, total_tiles, tpc, counter); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
