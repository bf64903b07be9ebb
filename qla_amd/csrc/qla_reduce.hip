// qla_reduce.hip -- global norm2 and dot reductions with an FP64 combine.
//
// Semantics restated from the reference (make_code_unary.pl:318-366,
// make_code_binary.pl:110-170, datatypes.pl:355-359):
//   norm2: for every complex component the norm re*re + im*im is formed in
//          the operand precision (float for F3/DF3) and added to a double sum;
//   dot:   both operands are promoted to double, sum += conj(a)*b in double.
// The reference's OpenMP reduction order depends on the thread count; here
// the order is fixed by the launch shape (deterministic run to run):
// stage 1 -- each block owns one contiguous chunk of the array, its threads
// accumulate a block-stride slice of the chunk in double, the block combines
// its lanes with a butterfly over the wavefront and LDS across waves and
// writes one double (pair) per block; stage 2 -- one block adds the
// per-block partials in a fixed tree and writes the result.  Contiguous
// chunks keep each block's reads within few DRAM pages (6.7 TB/s measured
// for a 510 MB norm2, against 5.4 TB/s with a grid-stride walk).
// Arrays of at most 65,536 complexes take a single launch instead (the last
// block to finish, found with an atomic ticket, adds the partials): below
// that size it beats the two launches, above it the device-scope fence each
// block needs before its ticket costs more than the second launch.
#include "qla_dev.hpp"
#include "qla_launch.hpp"

namespace qla {

constexpr int RBLOCK = 256;
constexpr int RMAXGRID = 2048;      // partial slots in the scratch (gathered forms)
constexpr int RCHUNKS = 1024;       // blocks of the contiguous stage 1

template <class T> struct vec2_of;
template <> struct vec2_of<float> { typedef float2 type; };
template <> struct vec2_of<double> { typedef double2 type; };

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// sum over the block; result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < RBLOCK / WAVE; i++) s += sh[i];
    }
    return s;
}

// This thread's first element and the end of its block's chunk: the m
// elements are split into gridDim.x chunks of a whole number of block widths.
__device__ __forceinline__ void block_chunk(long long m, long long &i, long long &end)
{
    long long per = (m + gridDim.x - 1) / gridDim.x;
    per = (per + RBLOCK - 1) / RBLOCK * RBLOCK;
    const long long beg = (long long)blockIdx.x * per;
    end = beg + per < m ? beg + per : m;
    i = beg + threadIdx.x;
}

template <class T>
__device__ __forceinline__ void acc_norm2(double &s, T re, T im)
{
    T nr = re * re + im * im;   // operand precision, no contraction
    s += (double)nr;
}
template <class T>
__device__ __forceinline__ void acc_dot(double &sr, double &si, T ar, T ai, T br, T bi)
{
    const double xr = ar, xi = ai, yr = br, yi = bi;
    sr += xr * yr + xi * yi;
    si += xr * yi - xi * yr;
}

// stage 1, one complex per load (any record alignment): this thread's sums
// over its block's chunk
template <class T, bool DOT>
__device__ __forceinline__ void stage1(const cpx<T> *__restrict__ a, const cpx<T> *__restrict__ b, long long m,
                                       double &sr, double &si)
{
    long long i, end;
    block_chunk(m, i, end);
    for (; i + 3 * RBLOCK < end; i += 4 * RBLOCK) {
        cpx<T> x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            x[u] = a[i + u * RBLOCK];
            if (DOT) y[u] = b[i + u * RBLOCK];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (DOT) acc_dot(sr, si, x[u].re, x[u].im, y[u].re, y[u].im);
            else acc_norm2(sr, x[u].re, x[u].im);
        }
    }
    for (; i < end; i += RBLOCK) {
        cpx<T> x = a[i];
        if (DOT) { cpx<T> y = b[i]; acc_dot(sr, si, x.re, x.im, y.re, y.im); }
        else acc_norm2(sr, x.re, x.im);
    }
}

template <class T, bool DOT>
__global__ void __launch_bounds__(RBLOCK)
k_reduce1(const cpx<T> *__restrict__ a, const cpx<T> *__restrict__ b, long long m,
          double *__restrict__ part)
{
    __shared__ double sh[RBLOCK / WAVE];
    double sr = 0, si = 0;
    stage1<T, DOT>(a, b, m, sr, si);
    double tr = block_sum(sr, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tr;
    if (DOT) {
        double ti = block_sum(si, sh);
        if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = ti;
    }
}

typedef float v4f __attribute__((ext_vector_type(4)));

// streaming 16-byte load with the non-temporal hint (each element is read once)
__device__ __forceinline__ float4 ld_stream(const float4 *p)
{
    const v4f v = __builtin_nontemporal_load((const v4f *)p);
    return make_float4(v.x, v.y, v.z, v.w);
}

// stage 1, two float complexes per 16-byte load (a, b 16-byte aligned, m even)
template <bool DOT>
__device__ __forceinline__ void stage1_f4(const float4 *__restrict__ a, const float4 *__restrict__ b,
                                          long long m2, double &sr, double &si)
{
    long long i, end;
    block_chunk(m2, i, end);
    for (; i + 3 * RBLOCK < end; i += 4 * RBLOCK) {
        float4 x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            x[u] = ld_stream(a + i + u * RBLOCK);
            if (DOT) y[u] = ld_stream(b + i + u * RBLOCK);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (DOT) {
                acc_dot(sr, si, x[u].x, x[u].y, y[u].x, y[u].y);
                acc_dot(sr, si, x[u].z, x[u].w, y[u].z, y[u].w);
            } else {
                acc_norm2(sr, x[u].x, x[u].y);
                acc_norm2(sr, x[u].z, x[u].w);
            }
        }
    }
    for (; i < end; i += RBLOCK) {
        float4 x = a[i];
        if (DOT) {
            float4 y = b[i];
            acc_dot(sr, si, x.x, x.y, y.x, y.y);
            acc_dot(sr, si, x.z, x.w, y.z, y.w);
        } else {
            acc_norm2(sr, x.x, x.y);
            acc_norm2(sr, x.z, x.w);
        }
    }
}

template <bool DOT>
__global__ void __launch_bounds__(RBLOCK)
k_reduce1_f4(const float4 *__restrict__ a, const float4 *__restrict__ b, long long m2,
             double *__restrict__ part)
{
    __shared__ double sh[RBLOCK / WAVE];
    double sr = 0, si = 0;
    stage1_f4<DOT>(a, b, m2, sr, si);
    double tr = block_sum(sr, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tr;
    if (DOT) {
        double ti = block_sum(si, sh);
        if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = ti;
    }
}

// stage 1 over gathered records: the record of site i is *pa[s] (pointer
// array) or a[s] (field), s = ia[i] (per-operand index), index[i] (x-form)
// or i, NC complexes each; likewise for b with ib.
// The lanes of a wave walk the 64 records of a tile one complex at a time
// (coalesced where the records are adjacent); the shuffles that hand each
// lane its record's address run with the whole wave active.
template <class T, bool DOT, int NC>
__global__ void __launch_bounds__(RBLOCK)
k_reduce1_gather(const void *__restrict__ a, const void *__restrict__ b, const int *__restrict__ index,
                 const int *__restrict__ ia, const int *__restrict__ ib, int n, int pa, int pb,
                 double *__restrict__ part)
{
    __shared__ double sh[RBLOCK / WAVE];
    double sr = 0, si = 0;
    const int lane = threadIdx.x % WAVE;
    const int nwaves = gridDim.x * (RBLOCK / WAVE);
    const int ntiles = tiles_of(n);
    for (int tile = blockIdx.x * (RBLOCK / WAVE) + (int)(threadIdx.x / WAVE); tile < ntiles; tile += nwaves) {
        bool act;
        const int i = lane_site(tile, lane, n, act);
        const int ns = min(WAVE, n - tile * WAVE);
        uint64_t ga = 0, gb = 0;
        if (act) {
            const int g = index ? index[i] : i;
            const int sa = ia ? ia[i] : g;
            ga = pa ? (uint64_t)((const void *const *)a)[sa] : (uint64_t)((const cpx<T> *)a + (size_t)sa * NC);
            if (DOT) {
                const int sb = ib ? ib[i] : g;
                gb = pb ? (uint64_t)((const void *const *)b)[sb] : (uint64_t)((const cpx<T> *)b + (size_t)sb * NC);
            }
        }
        const int tot = ns * NC;
        const int iters = (tot + WAVE - 1) / WAVE;
        for (int k = 0; k < iters; k++) {
            const int c = k * WAVE + lane;
            const int sl = c / NC;
            const int comp = c - sl * NC;
            const int src = sl < WAVE ? sl : 0;
            const uint64_t xa = shfl64(ga, src);
            const uint64_t xb = DOT ? shfl64(gb, src) : 0;
            if (c < tot) {
                const cpx<T> x = ((const cpx<T> *)xa)[comp];
                if (DOT) {
                    const cpx<T> y = ((const cpx<T> *)xb)[comp];
                    acc_dot(sr, si, x.re, x.im, y.re, y.im);
                } else {
                    acc_norm2(sr, x.re, x.im);
                }
            }
        }
    }
    double tr = block_sum(sr, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tr;
    if (DOT) {
        double ti = block_sum(si, sh);
        if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = ti;
    }
}

// stage 2: fixed-order combine of g partials; out_kind: 0 double, 1 float,
// 2 double complex, 3 float complex, 4 double real part of a dot, 5 float
// real part of a dot
__device__ __forceinline__ void combine_partials(const double *__restrict__ part, int g, bool dot, void *out,
                                                 int out_kind, double *sh)
{
    double sr = 0, si = 0;
    for (int i = threadIdx.x; i < g; i += RBLOCK) {
        sr += part[i];
        if (dot) si += part[g + i];
    }
    double tr = block_sum(sr, sh);
    double ti = dot ? block_sum(si, sh) : 0.0;
    if (threadIdx.x == 0) {
        switch (out_kind) {
        case 0: ((double *)out)[0] = tr; break;
        case 1: ((float *)out)[0] = (float)tr; break;
        case 2: ((double *)out)[0] = tr; ((double *)out)[1] = ti; break;
        case 3: ((float *)out)[0] = (float)tr; ((float *)out)[1] = (float)ti; break;
        case 4: ((double *)out)[0] = tr; break;
        case 5: ((float *)out)[0] = (float)tr; break;
        }
    }
}

__global__ void __launch_bounds__(RBLOCK)
k_reduce2(const double *__restrict__ part, int g, int dot, void *out, int out_kind)
{
    __shared__ double sh[RBLOCK / WAVE];
    combine_partials(part, g, dot != 0, out, out_kind, sh);
}

// Arrays of at most RSMALL_MAX complexes in one launch of at most RSMALL_G
// blocks: each block runs stage 1; its thread 0 stores the block's partials,
// fences them device-wide and draws a ticket; the block with the last ticket
// fences, adds the partials exactly as k_reduce2 does and puts the ticket
// back to 0 for the next call on the same scratch.  The result depends only
// on n (the grid is fixed by it).  Measured in async calls of F3 norm2_V
// against the two launches: 4,096 sites 4.45 against 5.76 us, 16,384 sites
// 4.87 against 5.75 us; at 65,536 sites (196,608 complexes) the one launch
// lost at every grid of 8 to 64 blocks (5.72-6.70 against 5.56 us), the
// fences of its blocks costing more than the second launch.
constexpr int RSMALL_G = 64;
constexpr long long RSMALL_MAX = 1 << 16;

template <class T, bool DOT, bool F4>
__global__ void __launch_bounds__(RBLOCK)
k_reduce_small(const void *__restrict__ a, const void *__restrict__ b, long long m, double *__restrict__ part,
               unsigned int *__restrict__ ticket, void *out, int out_kind)
{
    __shared__ double sh[RBLOCK / WAVE];
    __shared__ int last;
    double sr = 0, si = 0;
    if constexpr (F4) stage1_f4<DOT>((const float4 *)a, (const float4 *)b, m / 2, sr, si);
    else stage1<T, DOT>((const cpx<T> *)a, (const cpx<T> *)b, m, sr, si);
    const double tr = block_sum(sr, sh);
    const double ti = DOT ? block_sum(si, sh) : 0.0;
    if (threadIdx.x == 0) {
        part[blockIdx.x] = tr;
        if (DOT) part[gridDim.x + blockIdx.x] = ti;
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    combine_partials(part, gridDim.x, DOT, out, out_kind, sh);
    if (threadIdx.x == 0) atomicExch(ticket, 0u);
}

} // namespace qla

using namespace qla;

// scratch layout: 2 * RMAXGRID partial slots (real parts, then imaginary)
static_assert(QLA_REDUCE_SCRATCH_DOUBLES == 2 * RMAXGRID, "scratch layout");
int qla_reduce_scratch_doubles(void) { return QLA_REDUCE_SCRATCH_DOUBLES; }

static int grid_for(long long work)
{
    long long g = (work + (long long)RBLOCK * 4 - 1) / ((long long)RBLOCK * 4);
    if (g < 1) g = 1;
    if (g > RCHUNKS) g = RCHUNKS;
    return (int)g;
}

static bool use_f4(int prec, bool dot, const void *a, const void *b, long long m)
{
    return prec == 0 && ((uintptr_t)a % 16 == 0) && (!dot || (uintptr_t)b % 16 == 0) && (m % 2 == 0);
}

int qla_reduce2_partials(int prec, int kind, const void *a, const void *b, long long m)
{
    if (m <= 0) return 0;
    const bool dot = (kind == 1);
    const int g = grid_for(use_f4(prec, dot, a, b, m) ? m / 2 : m);
    return dot ? 2 * g : g;
}

// prec 0/1; kind 0 norm2 / 1 dot; m complex elements; out device address.
hipError_t qla_launch_reduce2(int prec, int kind, const void *a, const void *b, long long m, void *out,
                              int out_kind, double *scratch, hipStream_t st)
{
    const bool dot = (kind == 1);
    if (m <= 0) {
        hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(RBLOCK), 0, st, (const double *)scratch, 0, (int)dot,
                           out, out_kind);
        return hipGetLastError();
    }
    const bool f4 = use_f4(prec, dot, a, b, m);
    const int g = grid_for(f4 ? m / 2 : m);
    if (m <= RSMALL_MAX) {
        const int gs = g < RSMALL_G ? g : RSMALL_G;
        unsigned int *ticket = (unsigned int *)(scratch + QLA_REDUCE_SCRATCH_DOUBLES + 3);
#define QLA_SMALL(T, D, F)                                                                                  \
        hipLaunchKernelGGL((k_reduce_small<T, D, F>), dim3(gs), dim3(RBLOCK), 0, st, a, b, m, scratch, ticket, \
                           out, out_kind)
        if (f4) { if (dot) QLA_SMALL(float, true, true); else QLA_SMALL(float, false, true); }
        else if (prec == 0) { if (dot) QLA_SMALL(float, true, false); else QLA_SMALL(float, false, false); }
        else { if (dot) QLA_SMALL(double, true, false); else QLA_SMALL(double, false, false); }
#undef QLA_SMALL
        return hipGetLastError();
    }
    if (f4) {
        if (dot)
            hipLaunchKernelGGL(k_reduce1_f4<true>, dim3(g), dim3(RBLOCK), 0, st, (const float4 *)a,
                               (const float4 *)b, m / 2, scratch);
        else
            hipLaunchKernelGGL(k_reduce1_f4<false>, dim3(g), dim3(RBLOCK), 0, st, (const float4 *)a,
                               (const float4 *)b, m / 2, scratch);
    } else if (prec == 0) {
        if (dot)
            hipLaunchKernelGGL((k_reduce1<float, true>), dim3(g), dim3(RBLOCK), 0, st, (const cpx<float> *)a,
                               (const cpx<float> *)b, m, scratch);
        else
            hipLaunchKernelGGL((k_reduce1<float, false>), dim3(g), dim3(RBLOCK), 0, st, (const cpx<float> *)a,
                               (const cpx<float> *)b, m, scratch);
    } else {
        if (dot)
            hipLaunchKernelGGL((k_reduce1<double, true>), dim3(g), dim3(RBLOCK), 0, st, (const cpx<double> *)a,
                               (const cpx<double> *)b, m, scratch);
        else
            hipLaunchKernelGGL((k_reduce1<double, false>), dim3(g), dim3(RBLOCK), 0, st, (const cpx<double> *)a,
                               (const cpx<double> *)b, m, scratch);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(RBLOCK), 0, st, (const double *)scratch, g, (int)dot, out,
                       out_kind);
    return hipGetLastError();
}

// Gathered records (x-forms and pointer operands): ncomp complexes per
// record, one of 3, 6, 9, 12, 144 (V, H, M, D, P).
template <class T, bool DOT>
static hipError_t launch_gather_t(int ncomp, const void *a, const void *b, const int *index, const int *ia,
                                  const int *ib, int n, int pa, int pb, int g, double *scratch, hipStream_t st)
{
#define QLA_GATHER_CASE(NC)                                                                                     \
    case NC:                                                                                                   \
        hipLaunchKernelGGL((k_reduce1_gather<T, DOT, NC>), dim3(g), dim3(RBLOCK), 0, st, a, b, index, ia, ib, n, \
                           pa, pb, scratch);                                                                   \
        return hipGetLastError();
    switch (ncomp) {
        QLA_GATHER_CASE(3)
        QLA_GATHER_CASE(6)
        QLA_GATHER_CASE(9)
        QLA_GATHER_CASE(12)
        QLA_GATHER_CASE(144)
    }
#undef QLA_GATHER_CASE
    return hipErrorInvalidValue;
}

hipError_t qla_launch_reduce_gather(int prec, int kind, int ncomp, const void *a, const void *b, const int *index,
                                    const int *ia, const int *ib, int n, int pa, int pb, void *out, int out_kind,
                                    double *scratch, hipStream_t st)
{
    const bool dot = (kind == 1);
    int g = 0;
    if (n > 0) {
        // about two tiles of 64 records per wave
        const long long ntiles = ((long long)n + WAVE - 1) / WAVE;
        const long long want = (ntiles + 2 * (RBLOCK / WAVE) - 1) / (2 * (RBLOCK / WAVE));
        g = (int)(want < 1 ? 1 : (want > RMAXGRID ? RMAXGRID : want));
        hipError_t e;
        if (prec == 0) e = dot ? launch_gather_t<float, true>(ncomp, a, b, index, ia, ib, n, pa, pb, g, scratch, st)
                               : launch_gather_t<float, false>(ncomp, a, b, index, ia, ib, n, pa, pb, g, scratch, st);
        else e = dot ? launch_gather_t<double, true>(ncomp, a, b, index, ia, ib, n, pa, pb, g, scratch, st)
                     : launch_gather_t<double, false>(ncomp, a, b, index, ia, ib, n, pa, pb, g, scratch, st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(RBLOCK), 0, st, (const double *)scratch, g, (int)dot, out, out_kind);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Global sums of fields (build_qla.pl:2449-2521, make_code_unary.pl:318-366
// with qualifier sum): every real column of the record (ncr = reals per
// record, at most 288) summed over the sites in double (each element
// promoted; F and D both accumulate in D), then converted to the result
// precision.  Stage 1: each block owns a contiguous chunk of the steps;
// rpi = RBLOCK / ncr records are read per iteration, thread t taking column
// t % ncr of record t / ncr (coalesced), and the rpi partial sums of a column
// are added in order through LDS; records wider than the block (P: 288
// reals) give each thread the columns t and t + RBLOCK (contiguous fields
// of even width read 16 bytes a lane instead: k_sum1v).  Stage 2: one block
// per column adds that column's per-block partials in a fixed tree.
// ---------------------------------------------------------------------------
namespace qla {

constexpr int SUM_BLOCKS = 2048;
constexpr int SUM_MAXNCR = 288;

template <class T>
__device__ __forceinline__ const T *sum_record(const void *a, const int *index, int pa, int ncr, long long i)
{
    const long long s = index ? index[i] : i;
    if (pa) return ((const T *const *)a)[s];
    return (const T *)a + s * ncr;
}

template <class T>
__global__ void __launch_bounds__(RBLOCK)
k_sum1(const void *a, const int *index, int pa, long long n, int ncr, double *partials)
{
    __shared__ double sh[RBLOCK];
    const int t = threadIdx.x;
    const long long per = (n + gridDim.x - 1) / gridDim.x;
    const long long r0 = (long long)blockIdx.x * per;
    const long long r1 = r0 + per < n ? r0 + per : n;
    double *out = partials + (size_t)blockIdx.x * ncr;
    if (ncr <= RBLOCK) {
        const int rpi = RBLOCK / ncr;
        const int sub = t / ncr, col = t - sub * ncr;
        // four independent records per iteration keep loads in flight; the
        // partial sums are added in record order
        double acc = 0.0;
        if (sub < rpi) {
            long long r = r0 + sub;
            for (; r + 3 * rpi < r1; r += 4 * rpi) {
                const double v0 = (double)sum_record<T>(a, index, pa, ncr, r)[col];
                const double v1 = (double)sum_record<T>(a, index, pa, ncr, r + rpi)[col];
                const double v2 = (double)sum_record<T>(a, index, pa, ncr, r + 2 * rpi)[col];
                const double v3 = (double)sum_record<T>(a, index, pa, ncr, r + 3 * rpi)[col];
                acc += v0;
                acc += v1;
                acc += v2;
                acc += v3;
            }
            for (; r < r1; r += rpi) acc += (double)sum_record<T>(a, index, pa, ncr, r)[col];
        }
        sh[t] = acc;
        __syncthreads();
        if (t < ncr) {
            double s = 0.0;
            for (int k = 0; k < rpi; k++) s += sh[k * ncr + t];
            out[t] = s;
        }
    } else {
        double acc0 = 0.0, acc1 = 0.0;
        const int c1 = t + RBLOCK;
        for (long long r = r0; r < r1; r++) {
            const T *rec = sum_record<T>(a, index, pa, ncr, r);
            acc0 += (double)rec[t];
            if (c1 < ncr) acc1 += (double)rec[c1];
        }
        out[t] = acc0;
        if (c1 < ncr) out[c1] = acc1;
    }
}

// Stage 1 over a contiguous field whose records are a whole number of
// W-element vectors (16 bytes, or 8 for float pairs): the block's chunk
// starts on a record, and thread t of the first TA = (RBLOCK / (ncr / W)) *
// (ncr / W) threads reads vectors t, t + TA, ... of it, so it always holds the
// same W columns (vector t % (ncr / W) of a record); the TA / (ncr / W)
// partials of a column are added in thread order through LDS.
template <class T, int W>
__global__ void __launch_bounds__(RBLOCK)
k_sum1v(const T *a, long long n, int ncr, double *partials)
{
    typedef T VT __attribute__((ext_vector_type(W)));
    __shared__ double sh[RBLOCK * W];
    const int t = threadIdx.x;
    const int vpr = ncr / W;                       // vectors per record
    const int ta = (RBLOCK / vpr) * vpr;
    const long long per = (n + gridDim.x - 1) / gridDim.x;
    const long long r0 = (long long)blockIdx.x * per;
    const long long r1 = r0 + per < n ? r0 + per : n;
    const VT *v = (const VT *)a;
    double acc[W];
#pragma unroll
    for (int w = 0; w < W; w++) acc[w] = 0.0;
    if (t < ta && r0 < r1) {
        const long long k1 = r1 * vpr;
        long long k = r0 * vpr + t;
        for (; k + 3 * ta < k1; k += 4 * ta) {
            const VT x0 = __builtin_nontemporal_load(v + k), x1 = __builtin_nontemporal_load(v + k + ta);
            const VT x2 = __builtin_nontemporal_load(v + k + 2 * ta), x3 = __builtin_nontemporal_load(v + k + 3 * ta);
#pragma unroll
            for (int w = 0; w < W; w++) {
                acc[w] += (double)x0[w];
                acc[w] += (double)x1[w];
                acc[w] += (double)x2[w];
                acc[w] += (double)x3[w];
            }
        }
        for (; k < k1; k += ta) {
            const VT x0 = __builtin_nontemporal_load(v + k);
#pragma unroll
            for (int w = 0; w < W; w++) acc[w] += (double)x0[w];
        }
    }
#pragma unroll
    for (int w = 0; w < W; w++) sh[t * W + w] = acc[w];
    __syncthreads();
    // thread g * vpr + j holds columns j * W .. j * W + W - 1 at sh[g * ncr + j * W + w]
    for (int c = t; c < ncr; c += RBLOCK) {
        double s = 0.0;
        for (int g = 0; g < ta / vpr; g++) s += sh[g * ncr + c];
        partials[(size_t)blockIdx.x * ncr + c] = s;
    }
}

// stage 2: block c adds column c's per-block partials in a fixed tree
template <class TO>
__global__ void __launch_bounds__(RBLOCK) k_sum2(const double *partials, int nb, int ncr, TO *out)
{
    __shared__ double sh[RBLOCK / WAVE];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += RBLOCK) s += partials[(size_t)b * ncr + c];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[c] = (TO)s;
}

} // namespace qla

int qla_sum_scratch_doubles(void) { return qla::SUM_BLOCKS * qla::SUM_MAXNCR + qla::SUM_MAXNCR; }

hipError_t qla_launch_sum(int prec_in, int out_double, int ncr, const void *a, const int *index, int n, int pa,
                          void *out, double *scratch, hipStream_t st)
{
    using namespace qla;
    if (ncr < 1 || ncr > SUM_MAXNCR) return hipErrorInvalidValue;
    const long long nn = n < 0 ? 0 : n;
    const int rpi = ncr <= RBLOCK ? RBLOCK / ncr : 1;
    long long want = (nn + (long long)rpi * 16 - 1) / ((long long)rpi * 16);
    const int nb = (int)(want < 1 ? 1 : (want > SUM_BLOCKS ? SUM_BLOCKS : want));
    const uintptr_t al = (uintptr_t)a;
    const bool flat = !index && !pa;
    if (flat && prec_in == 0 && ncr % 4 == 0 && al % 16 == 0)
        hipLaunchKernelGGL((k_sum1v<float, 4>), dim3(nb), dim3(RBLOCK), 0, st, (const float *)a, nn, ncr, scratch);
    else if (flat && prec_in == 0 && ncr % 2 == 0 && al % 8 == 0)
        hipLaunchKernelGGL((k_sum1v<float, 2>), dim3(nb), dim3(RBLOCK), 0, st, (const float *)a, nn, ncr, scratch);
    else if (flat && prec_in != 0 && ncr % 2 == 0 && al % 16 == 0)
        hipLaunchKernelGGL((k_sum1v<double, 2>), dim3(nb), dim3(RBLOCK), 0, st, (const double *)a, nn, ncr, scratch);
    else if (prec_in == 0)
        hipLaunchKernelGGL(k_sum1<float>, dim3(nb), dim3(RBLOCK), 0, st, a, index, pa, nn, ncr, scratch);
    else hipLaunchKernelGGL(k_sum1<double>, dim3(nb), dim3(RBLOCK), 0, st, a, index, pa, nn, ncr, scratch);
    if (out_double) hipLaunchKernelGGL(k_sum2<double>, dim3(ncr), dim3(RBLOCK), 0, st, scratch, nb, ncr, (double *)out);
    else hipLaunchKernelGGL(k_sum2<float>, dim3(ncr), dim3(RBLOCK), 0, st, scratch, nb, ncr, (float *)out);
    return hipGetLastError();
}

