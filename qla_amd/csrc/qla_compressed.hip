// qla_compressed.hip -- compressed-precision forms of the Dirac building
// blocks (no reference counterpart; the reference stores FP32/FP64 only).
//
// Record formats (little-endian, 4-byte aligned):
//   link  24 B : rows 0 and 1 of an SU(3) matrix, 12 reals [row][col][re,im]
//                in bfloat16 (FMT_BF16) or IEEE binary16 (FMT_F16); row 2 is
//                rebuilt in FP32 as conj(row0 x row1).
//   Dh    52 B : DiracFermion, 24 binary16 values v then a float32 scale s;
//                component = float(v) * s.
//   Hh    28 B : HalfFermion, 12 binary16 values then a float32 scale.
// Encoding (pack kernels and the outputs of the fused kernels): s = max |x|
// over the record's reals, v = f16(x * (1/s)) rounded to nearest even
// (s = 0 gives v = 0).  All arithmetic after decoding is FP32 in the order of
// the plain kernels (spin projection, x = 0; x += a*t over colour).
//
// Staging: links are contiguous (24 B/site, 16-byte chunks when aligned);
// the gathered Dh/Hh records are moved as 12-byte chunks for the binary16
// part plus one 4-byte chunk for the scale; outputs are written back
// contiguously through LDS.
#include <atomic>
#include "qla_dev.hpp"
#include "qla_launch.hpp"

namespace qla {

enum { FMT_BF16 = 0, FMT_F16 = 1 };
constexpr int LINKB = 24, DHB = 52, HHB = 28;

__device__ __forceinline__ float bf16_to_f32(uint32_t bits) { return __uint_as_float(bits << 16); }
__device__ __forceinline__ uint32_t f32_to_bf16(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;     // round to nearest even (finite x)
}
__device__ __forceinline__ float f16_to_f32(uint32_t bits)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)bits);
}
__device__ __forceinline__ uint32_t f32_to_f16(float x)
{
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x);
}

template <int FMT> __device__ __forceinline__ float half_to_f32(uint32_t b)
{
    return FMT == FMT_BF16 ? bf16_to_f32(b) : f16_to_f32(b);
}

// 12 packed 16-bit reals (rows 0, 1) -> full 3x3 matrix, row 2 = conj(row0 x row1)
template <int FMT>
__device__ __forceinline__ void link_decode(cpx<float> (&m)[9], const uint32_t (&w)[6])
{
#pragma unroll
    for (int k = 0; k < 6; k++) {
        m[k].re = half_to_f32<FMT>(w[k] & 0xFFFFu);
        m[k].im = half_to_f32<FMT>(w[k] >> 16);
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        const cpx<float> a1 = m[j1], a2 = m[j2], b1 = m[3 + j1], b2 = m[3 + j2];
        const float pr = a1.re * b2.re - a1.im * b2.im, pi = a1.re * b2.im + a1.im * b2.re;
        const float qr = a2.re * b1.re - a2.im * b1.im, qi = a2.re * b1.im + a2.im * b1.re;
        m[6 + j].re = pr - qr;
        m[6 + j].im = -(pi - qi);
    }
}

// N complex values from N/2... : w holds 2 binary16 per word; value = float(v)*s
template <int NW>
__device__ __forceinline__ void fermion_decode(cpx<float> *x, const uint32_t (&w)[NW], float s)
{
#pragma unroll
    for (int k = 0; k < NW; k++) {
        x[k].re = f16_to_f32(w[k] & 0xFFFFu) * s;
        x[k].im = f16_to_f32(w[k] >> 16) * s;
    }
}
// The product x*(1/s) is rounded to FP32 before the conversion to binary16:
// the empty asm keeps the backend from fusing multiply and conversion into one
// mixed-precision FMA (v_fma_mix), which would round once instead of twice.
__device__ __forceinline__ float f32_rounded(float x)
{
    asm volatile("" : "+v"(x));
    return x;
}

template <int NW>
__device__ __forceinline__ float fermion_encode(uint32_t (&w)[NW], const cpx<float> *x)
{
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < NW; k++) s = fmaxf(s, fmaxf(fabsf(x[k].re), fabsf(x[k].im)));
    const float inv = s > 0.0f ? 1.0f / s : 0.0f;
#pragma unroll
    for (int k = 0; k < NW; k++)
        w[k] = f32_to_f16(f32_rounded(x[k].re * inv)) | (f32_to_f16(f32_rounded(x[k].im * inv)) << 16);
    return s;
}

__device__ __forceinline__ void cmp_dma_wait()
{
    vm_wait();
    wave_sync();
}

// ---------------------------------------------------------------------------
// Hh = spproj_{MU,SIGN}(Mc * pDh)
// ---------------------------------------------------------------------------
template <int FMT, int MU, int SIGN, int WPB>
__global__ void __launch_bounds__(WPB * WAVE)
k_spproj_c(char *__restrict__ r, const char *__restrict__ a, const char *const *__restrict__ b, int n, int al)
{
    // per wave: links 64x24 packed, Dh binary16 part 64x64 (12-byte chunks at a
    // 16-byte stride), Dh scales 64x4, Hh out 64x28 packed
    constexpr int LB = WAVE * LINKB, DB = WAVE * rec_lds<48, 12>::v, SB = WAVE * 4, OB = WAVE * HHB;
    constexpr int BYTES = LB + DB + SB + OB;
    __shared__ __attribute__((aligned(16))) char smem[WPB * BYTES];
    char *ll = smem + (threadIdx.x / WAVE) * BYTES;
    char *ld = ll + LB, *ls = ld + DB, *lo = ls + SB;
    const int lane = lane_id();
    const int ntiles = tiles_of(n), stride = gridDim.x * WPB;
    int tile = blockIdx.x * WPB + threadIdx.x / WAVE;
    if (tile >= ntiles) return;
    bool act;
    int i = lane_site(tile, lane, n, act);
    uint64_t pd = act ? (uint64_t)b[i] : 0;
    auto stage_links = [&](int base, int ns) {
        if (al & 1) stage_bytes<16, WAVE * LINKB>(ll, a + (size_t)base * LINKB, ns * LINKB, lane);
        else stage_bytes<4, WAVE * LINKB>(ll, a + (size_t)base * LINKB, ns * LINKB, lane);
    };
    // the first tile's links travel with its pointers; then the loop's only
    // memory wait is the one before its LDS reads
    stage_links(tile * WAVE, min(WAVE, n - tile * WAVE));
    vm_wait();
    bool first = true;
    for (;;) {
        const int base = tile * WAVE, ns = min(WAVE, n - base);
        if (!first) stage_links(base, ns);
        first = false;
        stage_gather<48, 12>(ld, pd, ns, lane);
        stage_gather<4, 4>(ls, pd + 48, ns, lane);
        // next tile's pointers travel with this tile's records
        const int next = tile + stride;
        uint64_t npd = 0;
        if (next < ntiles) {
            i = lane_site(next, lane, n, act);
            if (act) npd = (uint64_t)b[i];
        }
        cmp_dma_wait();
        uint32_t lw[6], dw[12];
#pragma unroll
        for (int k = 0; k < 6; k++) lw[k] = *(const uint32_t *)(ll + lane * LINKB + 4 * k);
#pragma unroll
        for (int k = 0; k < 12; k++) dw[k] = *(const uint32_t *)(ld + lane * rec_lds<48, 12>::v + elem_off<12, 4>::at(k));
        const float ds = *(const float *)(ls + lane * 4);
        cpx<float> m[9], d[12], t[6], h[6];
        link_decode<FMT>(m, lw);
        fermion_decode(d, dw, ds);
#pragma unroll
        for (int ic = 0; ic < 3; ic++) spproj_row<MU, SIGN>(t[2 * ic], t[2 * ic + 1], d + 4 * ic);
        mat_times_cols<EQ, false, 2>(h, m, t);
        uint32_t hw[6];
        const float hs = fermion_encode(hw, h);
#pragma unroll
        for (int k = 0; k < 6; k++) *(uint32_t *)(lo + lane * HHB + 4 * k) = hw[k];
        *(float *)(lo + lane * HHB + 24) = hs;
        wave_sync();
        if (al & 4) store_bytes<16, WAVE * HHB>(r + (size_t)base * HHB, lo, ns * HHB, lane);
        else store_bytes<4, WAVE * HHB>(r + (size_t)base * HHB, lo, ns * HHB, lane);
        wave_sync();
        if (next >= ntiles) break;
        tile = next;
        pd = npd;
    }
}

// ---------------------------------------------------------------------------
// Dh += sprecon_{MU,SIGN}(Mc * pHh)   (decode, accumulate in FP32, re-encode)
// ---------------------------------------------------------------------------
template <int FMT, int MU, int SIGN, int WPB>
__global__ void __launch_bounds__(WPB * WAVE)
k_sprecon_c(char *__restrict__ r, const char *__restrict__ a, const char *const *__restrict__ b, int n, int al)
{
    typedef recon_tab<MU, SIGN> TAB;
    constexpr int LB = WAVE * LINKB, HB = WAVE * rec_lds<24, 12>::v, SB = WAVE * 4, RB = WAVE * DHB;
    constexpr int BYTES = LB + HB + SB + RB;
    __shared__ __attribute__((aligned(16))) char smem[WPB * BYTES];
    char *ll = smem + (threadIdx.x / WAVE) * BYTES;
    char *lh = ll + LB, *ls = lh + HB, *lr = ls + SB;
    const int lane = lane_id();
    const int ntiles = tiles_of(n), stride = gridDim.x * WPB;
    int tile = blockIdx.x * WPB + threadIdx.x / WAVE;
    if (tile >= ntiles) return;
    bool act;
    int i = lane_site(tile, lane, n, act);
    uint64_t ph = act ? (uint64_t)b[i] : 0;
    auto stage_contig_ops = [&](int base, int ns) {
        if (al & 1) stage_bytes<16, WAVE * LINKB>(ll, a + (size_t)base * LINKB, ns * LINKB, lane);
        else stage_bytes<4, WAVE * LINKB>(ll, a + (size_t)base * LINKB, ns * LINKB, lane);
        if (al & 4) stage_bytes<16, WAVE * DHB>(lr, r + (size_t)base * DHB, ns * DHB, lane);
        else stage_bytes<4, WAVE * DHB>(lr, r + (size_t)base * DHB, ns * DHB, lane);
    };
    // the first tile's links and destination records travel with its
    // pointers; then the loop's only memory wait is the one before its LDS reads
    stage_contig_ops(tile * WAVE, min(WAVE, n - tile * WAVE));
    vm_wait();
    bool first = true;
    for (;;) {
        const int base = tile * WAVE, ns = min(WAVE, n - base);
        if (!first) stage_contig_ops(base, ns);
        first = false;
        stage_gather<24, 12>(lh, ph, ns, lane);
        stage_gather<4, 4>(ls, ph + 24, ns, lane);
        const int next = tile + stride;
        uint64_t nph = 0;
        if (next < ntiles) {
            i = lane_site(next, lane, n, act);
            if (act) nph = (uint64_t)b[i];
        }
        cmp_dma_wait();
        uint32_t lw[6], hw[6], dw[12];
#pragma unroll
        for (int k = 0; k < 6; k++) lw[k] = *(const uint32_t *)(ll + lane * LINKB + 4 * k);
#pragma unroll
        for (int k = 0; k < 6; k++) hw[k] = *(const uint32_t *)(lh + lane * rec_lds<24, 12>::v + elem_off<12, 4>::at(k));
#pragma unroll
        for (int k = 0; k < 12; k++) dw[k] = *(const uint32_t *)(lr + lane * DHB + 4 * k);
        const float hsc = *(const float *)(ls + lane * 4);
        const float dsc = *(const float *)(lr + lane * DHB + 48);
        cpx<float> m[9], h[6], d[12], tv[6];
        link_decode<FMT>(m, lw);
        fermion_decode(h, hw, hsc);
        fermion_decode(d, dw, dsc);
        mat_times_cols<EQ, false, 2>(tv, m, h);
#pragma unroll
        for (int ic = 0; ic < 3; ic++) {
            recon_apply<PEQ, TAB::op[0]>(d[4 * ic + 0], tv[2 * ic + TAB::src[0]]);
            recon_apply<PEQ, TAB::op[1]>(d[4 * ic + 1], tv[2 * ic + TAB::src[1]]);
            recon_apply<PEQ, TAB::op[2]>(d[4 * ic + 2], tv[2 * ic + TAB::src[2]]);
            recon_apply<PEQ, TAB::op[3]>(d[4 * ic + 3], tv[2 * ic + TAB::src[3]]);
        }
        const float s = fermion_encode(dw, d);
        wave_sync();
#pragma unroll
        for (int k = 0; k < 12; k++) *(uint32_t *)(lr + lane * DHB + 4 * k) = dw[k];
        *(float *)(lr + lane * DHB + 48) = s;
        wave_sync();
        if (al & 4) store_bytes<16, WAVE * DHB>(r + (size_t)base * DHB, lr, ns * DHB, lane);
        else store_bytes<4, WAVE * DHB>(r + (size_t)base * DHB, lr, ns * DHB, lane);
        wave_sync();
        if (next >= ntiles) break;
        tile = next;
        ph = nph;
    }
}

// ---------------------------------------------------------------------------
// pack / unpack (one thread per site; conversion kernels, not on the hot path)
// ---------------------------------------------------------------------------
template <int FMT>
__global__ void k_pack_link(uint16_t *__restrict__ r, const float *__restrict__ a, int n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const float x = a[i * 18 + k];
        r[i * 12 + k] = (uint16_t)(FMT == FMT_BF16 ? f32_to_bf16(x) : f32_to_f16(x));
    }
}
template <int FMT>
__global__ void k_unpack_link(float *__restrict__ r, const uint16_t *__restrict__ a, int n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t w[6];
#pragma unroll
    for (int k = 0; k < 6; k++) w[k] = (uint32_t)a[i * 12 + 2 * k] | ((uint32_t)a[i * 12 + 2 * k + 1] << 16);
    cpx<float> m[9];
    link_decode<FMT>(m, w);
#pragma unroll
    for (int k = 0; k < 9; k++) { r[i * 18 + 2 * k] = m[k].re; r[i * 18 + 2 * k + 1] = m[k].im; }
}
// NC complex per record; record bytes = 4*NC + 4
template <int NC>
__global__ void k_pack_fermion(char *__restrict__ r, const float *__restrict__ a, int n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cpx<float> x[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) { x[k].re = a[i * 2 * NC + 2 * k]; x[k].im = a[i * 2 * NC + 2 * k + 1]; }
    uint32_t w[NC];
    const float s = fermion_encode(w, x);
    uint32_t *o = (uint32_t *)(r + (size_t)i * (4 * NC + 4));
#pragma unroll
    for (int k = 0; k < NC; k++) o[k] = w[k];
    ((float *)o)[NC] = s;
}
template <int NC>
__global__ void k_unpack_fermion(float *__restrict__ r, const char *__restrict__ a, int n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t *o = (const uint32_t *)(a + (size_t)i * (4 * NC + 4));
    uint32_t w[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) w[k] = o[k];
    cpx<float> x[NC];
    fermion_decode(x, w, ((const float *)o)[NC]);
#pragma unroll
    for (int k = 0; k < NC; k++) { r[i * 2 * NC + 2 * k] = x[k].re; r[i * 2 * NC + 2 * k + 1] = x[k].im; }
}

// ---------------------------------------------------------------------------
// Reductions over the decoded Dh field: norm2 (float component norms added
// in double) and conj(a).b in double -- the DF3 semantics on decoded values.
// Stage 1 writes one double (pair) per block; k_reduce2 style combine follows.
// ---------------------------------------------------------------------------
constexpr int CRB = 256;
__device__ __forceinline__ double c_wave_sum(double v)
{
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}
__device__ __forceinline__ double c_block_sum(double v, double *sh)
{
    v = c_wave_sum(v);
    __syncthreads();
    if (threadIdx.x % WAVE == 0) sh[threadIdx.x / WAVE] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < CRB / WAVE; i++) s += sh[i];
    return s;
}

template <bool DOT>
__global__ void __launch_bounds__(CRB)
k_reduce_dh(const char *__restrict__ a, const char *__restrict__ b, int n, double *__restrict__ part)
{
    __shared__ double sh[CRB / WAVE];
    double sr = 0, si = 0;
    for (long long i = (long long)blockIdx.x * CRB + threadIdx.x; i < n; i += (long long)gridDim.x * CRB) {
        const uint32_t *pa = (const uint32_t *)(a + (size_t)i * DHB);
        uint32_t wa[12];
#pragma unroll
        for (int k = 0; k < 12; k++) wa[k] = pa[k];
        cpx<float> x[12];
        fermion_decode(x, wa, ((const float *)pa)[12]);
        if (!DOT) {
#pragma unroll
            for (int k = 0; k < 12; k++) sr += (double)(x[k].re * x[k].re + x[k].im * x[k].im);
        } else {
            const uint32_t *pb = (const uint32_t *)(b + (size_t)i * DHB);
            uint32_t wb[12];
#pragma unroll
            for (int k = 0; k < 12; k++) wb[k] = pb[k];
            cpx<float> y[12];
            fermion_decode(y, wb, ((const float *)pb)[12]);
#pragma unroll
            for (int k = 0; k < 12; k++) {
                const double xr = x[k].re, xi = x[k].im, yr = y[k].re, yi = y[k].im;
                sr += xr * yr + xi * yi;
                si += xr * yi - xi * yr;
            }
        }
    }
    const double tr = c_block_sum(sr, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tr;
    if (DOT) {
        const double ti = c_block_sum(si, sh);
        if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = ti;
    }
}

__global__ void __launch_bounds__(CRB)
k_reduce_dh2(const double *__restrict__ part, int g, int dot, double *__restrict__ out)
{
    __shared__ double sh[CRB / WAVE];
    double sr = 0, si = 0;
    for (int i = threadIdx.x; i < g; i += CRB) {
        sr += part[i];
        if (dot) si += part[g + i];
    }
    const double tr = c_block_sum(sr, sh);
    const double ti = dot ? c_block_sum(si, sh) : 0.0;
    if (threadIdx.x == 0) {
        out[0] = tr;
        if (dot) out[1] = ti;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// one wave per block (a block's LDS is released when its last wave ends):
// 32^4 binary16-link spproj 21.96 us against 22.31 us with four-wave blocks
constexpr int CWPB = 1;
static std::atomic<int> g_cmp_tiles_per_wave{ 1 };     // set from any host thread

static dim3 cgrid(int n)
{
    const long long ntiles = ((long long)n + WAVE - 1) / WAVE;
    const long long per_block = (long long)CWPB * g_cmp_tiles_per_wave.load(std::memory_order_relaxed);
    return dim3((unsigned)((ntiles + per_block - 1) / per_block));
}

template <int FMT, int MU, int SIGN>
static hipError_t spproj_c_t(void *r, const void *a, const void *b, int n, int al, hipStream_t st)
{
    hipLaunchKernelGGL((k_spproj_c<FMT, MU, SIGN, CWPB>), cgrid(n), dim3(CWPB * WAVE), 0, st, (char *)r,
                       (const char *)a, (const char *const *)b, n, al);
    return hipGetLastError();
}
template <int FMT, int MU, int SIGN>
static hipError_t sprecon_c_t(void *r, const void *a, const void *b, int n, int al, hipStream_t st)
{
    hipLaunchKernelGGL((k_sprecon_c<FMT, MU, SIGN, CWPB>), cgrid(n), dim3(CWPB * WAVE), 0, st, (char *)r,
                       (const char *)a, (const char *const *)b, n, al);
    return hipGetLastError();
}

#define QLA_CMP_DISPATCH(FN)                                                                      \
    template <int FMT>                                                                             \
    static hipError_t FN##_mu(void *r, const void *a, const void *b, int mu, int sign, int n, int al, \
                              hipStream_t st)                                                      \
    {                                                                                              \
        if (sign == 1) {                                                                           \
            switch (mu) {                                                                          \
            case 0: return FN##_t<FMT, 0, 1>(r, a, b, n, al, st);                                  \
            case 1: return FN##_t<FMT, 1, 1>(r, a, b, n, al, st);                                  \
            case 2: return FN##_t<FMT, 2, 1>(r, a, b, n, al, st);                                  \
            case 3: return FN##_t<FMT, 3, 1>(r, a, b, n, al, st);                                  \
            case 4: return FN##_t<FMT, 4, 1>(r, a, b, n, al, st);                                  \
            }                                                                                      \
        } else {                                                                                   \
            switch (mu) {                                                                          \
            case 0: return FN##_t<FMT, 0, -1>(r, a, b, n, al, st);                                 \
            case 1: return FN##_t<FMT, 1, -1>(r, a, b, n, al, st);                                 \
            case 2: return FN##_t<FMT, 2, -1>(r, a, b, n, al, st);                                 \
            case 3: return FN##_t<FMT, 3, -1>(r, a, b, n, al, st);                                 \
            case 4: return FN##_t<FMT, 4, -1>(r, a, b, n, al, st);                                 \
            }                                                                                      \
        }                                                                                          \
        return hipSuccess;                                                                         \
    }
QLA_CMP_DISPATCH(spproj_c)
QLA_CMP_DISPATCH(sprecon_c)
#undef QLA_CMP_DISPATCH

} // namespace qla

using namespace qla;

void qla_set_cmp_tiles_per_wave(int tpw) { g_cmp_tiles_per_wave.store(tpw < 1 ? 1 : tpw, std::memory_order_relaxed); }

static inline int cmp_align(const void *a, const void *r)
{
    return (((uintptr_t)a % 16 == 0) ? 1 : 0) | (((uintptr_t)r % 16 == 0) ? 4 : 0);
}

hipError_t qla_launch_spproj_c(int fmt, void *r, const void *a, const void *b, int mu, int sign, int n,
                               hipStream_t st)
{
    if (n <= 0 || mu < 0 || mu > 4) return hipSuccess;
    if (fmt == FMT_BF16) return spproj_c_mu<FMT_BF16>(r, a, b, mu, sign, n, cmp_align(a, r), st);
    if (fmt == FMT_F16) return spproj_c_mu<FMT_F16>(r, a, b, mu, sign, n, cmp_align(a, r), st);
    return hipErrorInvalidValue;
}

hipError_t qla_launch_sprecon_c(int fmt, void *r, const void *a, const void *b, int mu, int sign, int n,
                                hipStream_t st)
{
    if (n <= 0 || mu < 0 || mu > 4) return hipSuccess;
    if (fmt == FMT_BF16) return sprecon_c_mu<FMT_BF16>(r, a, b, mu, sign, n, cmp_align(a, r), st);
    if (fmt == FMT_F16) return sprecon_c_mu<FMT_F16>(r, a, b, mu, sign, n, cmp_align(a, r), st);
    return hipErrorInvalidValue;
}

// kind: 0 link pack, 1 link unpack, 2 D pack, 3 D unpack, 4 H pack, 5 H unpack
hipError_t qla_launch_convert(int kind, int fmt, void *r, const void *a, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const dim3 g((unsigned)(((long long)n + 255) / 256)), b(256);
    if ((kind == 0 || kind == 1) && fmt != FMT_BF16 && fmt != FMT_F16) return hipErrorInvalidValue;
    switch (kind) {
    case 0:
        if (fmt == FMT_BF16) hipLaunchKernelGGL(k_pack_link<FMT_BF16>, g, b, 0, st, (uint16_t *)r, (const float *)a, n);
        else hipLaunchKernelGGL(k_pack_link<FMT_F16>, g, b, 0, st, (uint16_t *)r, (const float *)a, n);
        break;
    case 1:
        if (fmt == FMT_BF16) hipLaunchKernelGGL(k_unpack_link<FMT_BF16>, g, b, 0, st, (float *)r, (const uint16_t *)a, n);
        else hipLaunchKernelGGL(k_unpack_link<FMT_F16>, g, b, 0, st, (float *)r, (const uint16_t *)a, n);
        break;
    case 2: hipLaunchKernelGGL(k_pack_fermion<12>, g, b, 0, st, (char *)r, (const float *)a, n); break;
    case 3: hipLaunchKernelGGL(k_unpack_fermion<12>, g, b, 0, st, (float *)r, (const char *)a, n); break;
    case 4: hipLaunchKernelGGL(k_pack_fermion<6>, g, b, 0, st, (char *)r, (const float *)a, n); break;
    case 5: hipLaunchKernelGGL(k_unpack_fermion<6>, g, b, 0, st, (float *)r, (const char *)a, n); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Dh norm2 (b == NULL) or dot; result doubles written to device `out`.
hipError_t qla_launch_reduce_dh(const void *a, const void *b, int n, double *out, double *scratch,
                                hipStream_t st)
{
    const bool dot = b != nullptr;
    long long g = ((long long)n + CRB * 4 - 1) / (CRB * 4);
    if (g < 1) g = 1;
    if (g > 2048) g = 2048;
    if (n > 0) {
        if (dot)
            hipLaunchKernelGGL(k_reduce_dh<true>, dim3((unsigned)g), dim3(CRB), 0, st, (const char *)a,
                               (const char *)b, n, scratch);
        else
            hipLaunchKernelGGL(k_reduce_dh<false>, dim3((unsigned)g), dim3(CRB), 0, st, (const char *)a,
                               (const char *)b, n, scratch);
    } else {
        g = 0;
    }
    hipLaunchKernelGGL(k_reduce_dh2, dim3(1), dim3(CRB), 0, st, (const double *)scratch, (int)g, (int)dot, out);
    return hipGetLastError();
}
