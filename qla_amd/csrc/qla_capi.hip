// qla_capi.hip -- the C ABI: QLA symbols and argument lists (declared in
// include/qla_mi355x.h) on device-resident arrays.
//
// Each entry point validates nothing the reference does not (QLA has no error
// returns), launches its kernel on the calling thread's stream and, in the
// default synchronous mode, waits for it, so a call returns with its results
// complete as the CPU library's does.  Reductions accept a host or a device
// result pointer (hipPointerGetAttributes tells which).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>
#include "qla_launch.hpp"
#include "../../include/qla_mi355x.h"

#define QLA_MI355X_VERSION_STRING "qla_mi355x 0.1 (QLA 1.9.0 hot-path ABI, gfx950)"

namespace {

struct thread_state {
    hipStream_t stream = nullptr;
    int sync = 1;
    int order = QLA_MI355X_ORDER_GENERIC;
    hipError_t err = hipSuccess;
};

thread_local thread_state ts;

void record(hipError_t e, const char *where)
{
    if (e == hipSuccess) return;
    ts.err = e;
    fprintf(stderr, "qla_mi355x: %s: %s\n", where, hipGetErrorString(e));
}

void finish(hipError_t e, const char *where)
{
    record(e, where);
    if (e == hipSuccess && ts.sync) record(hipStreamSynchronize(ts.stream), where);
}

// the calling thread's stream is being captured into a graph
bool capturing()
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(ts.stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// ----------------------------------------------------------- reduction scratch
// Device scratch of the global reductions and sums: one buffer of each kind
// per (host thread, device, stream), so reductions that run concurrently on
// different streams, or that two threads issue on one stream (each copies
// its own result slot back), never share partials or a result slot.  A
// buffer is allocated on its first use, which must not fall inside a
// stream capture; thread exit releases the thread's buffers and
// QLA_MI355X_finalize releases all of them.
//
// Each buffer comes with a result slot in pinned (page-locked, coherent)
// host memory.  A reduction whose result pointer is in host memory has its
// last kernel write the result there, waits for the stream and copies it
// out on the host: no copy command is queued, and the caller's pageable
// memory is never the target of a device transfer (a 16^4 F3 norm2 into
// pageable host memory took 24.4 us that way against 17.0 us into pinned
// memory and 15.0 us into device memory).
enum { SCRATCH_REDUCE = 0, SCRATCH_SUM = 1 };
constexpr int HOST_SLOT_DOUBLES = 288;     // the largest result: a propagator sum, 144 complexes

struct scratch_buf {
    double *dev = nullptr;      // partials (and the device result slot)
    double *host = nullptr;     // pinned result slot, HOST_SLOT_DOUBLES
};

struct scratch_key {
    std::thread::id tid;
    int dev;
    hipStream_t stream;
    int kind;
    bool operator<(const scratch_key &o) const
    {
        return std::tie(tid, dev, stream, kind) < std::tie(o.tid, o.dev, o.stream, o.kind);
    }
};

std::mutex g_scratch_mutex;
std::map<scratch_key, scratch_buf> g_scratch;
const std::thread::id g_main_thread = std::this_thread::get_id();

// frees the calling thread's buffers when the thread ends (the main thread's
// live until the process ends or QLA_MI355X_finalize)
struct scratch_owner {
    bool used = false;
    ~scratch_owner()
    {
        if (!used || std::this_thread::get_id() == g_main_thread) return;
        std::lock_guard<std::mutex> lk(g_scratch_mutex);
        for (auto it = g_scratch.begin(); it != g_scratch.end();) {
            if (it->first.tid == std::this_thread::get_id()) {
                int cur = 0;
                if (hipGetDevice(&cur) == hipSuccess && cur != it->first.dev) (void)hipSetDevice(it->first.dev);
                (void)hipFree(it->second.dev);
                (void)hipHostFree(it->second.host);
                if (cur != it->first.dev) (void)hipSetDevice(cur);
                it = g_scratch.erase(it);
            } else {
                ++it;
            }
        }
    }
};
thread_local scratch_owner t_scratch_owner;

size_t scratch_doubles(int kind)
{
    // reductions: partials, then a 4-double result slot
    return kind == SCRATCH_REDUCE ? (size_t)qla_reduce_scratch_doubles() + 4 : (size_t)qla_sum_scratch_doubles();
}

// the calling thread's buffer of this kind for its device and stream
// (.dev == nullptr after an error, which is recorded)
scratch_buf scratch_for(int kind, const char *name)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { record(e, name); return {}; }
    const scratch_key key = { std::this_thread::get_id(), dev, ts.stream, kind };
    {
        std::lock_guard<std::mutex> lk(g_scratch_mutex);
        auto it = g_scratch.find(key);
        if (it != g_scratch.end()) return it->second;
    }
    if (capturing()) {
        fprintf(stderr, "qla_mi355x: %s: the first reduction on a stream allocates its scratch and cannot be "
                        "captured; run one reduction on this stream before capturing\n", name);
        record(hipErrorStreamCaptureUnsupported, name);
        return {};
    }
    scratch_buf b;
    const size_t bytes = sizeof(double) * scratch_doubles(kind);
    e = hipMalloc((void **)&b.dev, bytes);
    // the reduction's ticket (in its result slot) starts at 0; every call
    // that draws it leaves it at 0
    if (e == hipSuccess && kind == SCRATCH_REDUCE)
        e = hipMemset(b.dev + qla_reduce_scratch_doubles(), 0, 4 * sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc((void **)&b.host, sizeof(double) * HOST_SLOT_DOUBLES, hipHostMallocCoherent);
    if (e != hipSuccess) {
        if (b.dev) (void)hipFree(b.dev);
        if (b.host) (void)hipHostFree(b.host);
        record(e, name);
        return {};
    }
    t_scratch_owner.used = true;
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    g_scratch[key] = b;
    return b;
}

// a result delivered to host memory waits on the host, which a stream
// capture cannot contain
bool refuse_host_result_in_capture(const char *name)
{
    if (!capturing()) return false;
    fprintf(stderr, "qla_mi355x: %s: a host result pointer cannot be captured; pass device memory\n", name);
    record(hipErrorStreamCaptureUnsupported, name);
    return true;
}

bool is_device_pointer(const void *p)
{
    hipPointerAttribute_t at;
    hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess) { (void)hipGetLastError(); return false; }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

// run a reduction with device scratch sc (partials, then a 4-double
// result slot) and deliver `bytes` of result to r (host or device);
// launch(out, scratch) enqueues the kernels that write the result to out.
// A host result goes through the pinned slot `pinned` when there is one,
// else through sc's device result slot and a copy.
template <class L>
void reduce_with(double *sc, double *pinned, L launch, void *r, size_t bytes, const char *name)
{
    if (is_device_pointer(r)) {
        finish(launch(r, sc), name);
        return;
    }
    if (refuse_host_result_in_capture(name)) return;
    if (pinned) {
        hipError_t e = launch((void *)pinned, sc);
        if (e == hipSuccess) e = hipStreamSynchronize(ts.stream);
        if (e == hipSuccess) memcpy(r, pinned, bytes);
        record(e, name);
        return;
    }
    double *slot = sc + qla_reduce_scratch_doubles();
    hipError_t e = launch((void *)slot, sc);
    if (e == hipSuccess) e = hipMemcpyAsync(r, slot, bytes, hipMemcpyDeviceToHost, ts.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ts.stream);
    record(e, name);
}

template <class L>
void reduce_to(L launch, void *r, size_t bytes, const char *name)
{
    const scratch_buf sc = scratch_for(SCRATCH_REDUCE, name);
    if (sc.dev) reduce_with(sc.dev, sc.host, launch, r, bytes, name);
}

// ------------------------------------------------ allocation-free reduction
// Scratch of QLA_MI355X_DF3_r_veq_norm2_D_noalloc: the partials and the
// result slot, laid out as the allocated reduction scratch, in one static
// device array (32 KiB; every GPU the library runs on holds its own copy
// until the process ends).  Stage 1 writes every partial slot stage 2
// reads, and the single-launch form's ticket starts at 0 (static storage)
// and is left at 0 by every call, so nothing in it needs clearing.  Host
// code never names the array in a launch or copy: it takes each device's
// address once with hipGetSymbolAddress and passes that.
constexpr int NOALLOC_DOUBLES = QLA_REDUCE_SCRATCH_DOUBLES + 4;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
std::atomic<double *> g_noalloc_ptr[NOALLOC_MAX_DEVICES];

// this device's address of `scratch`, looked up on the first call on the device
double *noalloc_scratch(const char *name)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { record(e, name); return nullptr; }
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) { record(hipErrorInvalidDevice, name); return nullptr; }
    double *p = g_noalloc_ptr[dev].load(std::memory_order_acquire);
    if (p) return p;
    e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        record(e != hipSuccess ? e : hipErrorInvalidSymbol, name);
        return nullptr;
    }
    g_noalloc_ptr[dev].store(p, std::memory_order_release);
    return p;
}

// One QLA global reduction (qla_fields_abi.inc): kind 0 norm2, 1 dot (the
// real-part forms take out_kind 4 / 5); ncomp complexes per record.  The
// v-form over fields streams the arrays as one run of n*ncomp complexes;
// x-forms, per-operand index forms and pointer operands gather records.
void reduce_entry(int prec, int kind, int ncomp, int pa, int pb, void *r, const void *a, const void *b,
                  const int *index, const int *ia, const int *ib, int n, int out_kind, size_t bytes,
                  const char *name)
{
    if (!index && !ia && !ib && !pa && !pb) {
        const long long m = n > 0 ? (long long)ncomp * n : 0;
        reduce_to([&](void *out, double *sc) {
            return qla_launch_reduce2(prec, kind, a, b, m, out, out_kind, sc, ts.stream);
        }, r, bytes, name);
        return;
    }
    reduce_to([&](void *out, double *sc) {
        return qla_launch_reduce_gather(prec, kind, ncomp, a, b, index, ia, ib, n, pa, pb, out, out_kind, sc,
                                        ts.stream);
    }, r, bytes, name);
}

// One QLA precision conversion: to_float 0 DF3 (float -> double), 1 FD3.
void convert_entry(int to_float, int ncomp, int pa, void *r, const void *a, const int *index, const int *ir,
                   const int *ia, int n, const char *name)
{
    if (n <= 0) return;
    finish(qla_launch_precision(to_float, ncomp, r, a, index, ir, ia, n, pa, ts.stream), name);
}

// One colour-matrix function (qla_matfuncs_abi.inc): fn codes of
// qla_launch_matfunc; prec 0 takes nc = 1..8 (QLA_F3 passes 3), prec 1 nc = 3.
// An unsupported nc or function records hipErrorInvalidValue and writes
// nothing.
void matfunc_entry(int prec, int fn, int nc, int pa, int pb, void *r, const void *a, const void *b,
                   const int *index, const int *ir, const int *ia, const int *ib, int n, const char *name)
{
    if (n <= 0) return;
    if (nc < 1 || nc > 8 || (prec == 1 && nc != 3)) {
        record(hipErrorInvalidValue, name);
        return;
    }
    const qla_matfunc_args p = { r, a, b, index, ir, ia, ib, n, pa, pb };
    finish(qla_launch_matfunc(prec, fn, nc, p, ts.stream), name);
}

// One field linear-algebra call (qla_vecops_abi.inc): op and eqop codes of
// qla_launch_vec; xform bit 0 conjugates and bit 1 transposes the source x;
// scalar 0 none, 1 a real constant, 2 a complex constant at s, in host or
// device memory (a host value is read here, a device one by the kernel when
// it runs, so it may be the result of earlier work on the stream).
void vec_entry(int prec, int op, int eqop, int ncomp, int xform, int px, int py, void *r, const void *x,
               const void *y, const void *s, int scalar, const int *index, const int *ir, const int *ix,
               const int *iy, int n, const char *name)
{
    if (n <= 0) return;
    qla_vec_args p;
    memset(&p, 0, sizeof(p));
    p.r = r;
    p.x = x;
    p.y = y;
    p.index = index;
    p.ir = ir;
    p.ix = ix;
    p.iy = iy;
    p.n = n;
    p.px = px;
    p.py = py;
    p.conj = xform & 1;
    p.trans = (xform >> 1) & 1;
    if (scalar) {
        if (!s) {
            record(hipErrorInvalidValue, name);
            return;
        }
        if (is_device_pointer(s)) {
            p.sp = s;
        } else if (prec == 0) {
            p.s_re = ((const float *)s)[0];
            p.s_im = scalar == 2 ? ((const float *)s)[1] : 0.0;
        } else {
            p.s_re = ((const double *)s)[0];
            p.s_im = scalar == 2 ? ((const double *)s)[1] : 0.0;
        }
    }
    p.al = ((uintptr_t)r % 16 == 0 && (uintptr_t)x % 16 == 0 && (uintptr_t)y % 16 == 0) ? 1 : 0;
    finish(qla_launch_vec(prec, op, eqop, ncomp, p, ts.stream), name);
}

// One fill (qla_vecops_abi.inc): kind 0 zero, 1 a record broadcast, 2 a
// complex constant on the diagonal of colour matrices; the record or
// constant a may be in host memory (copied into the launch at the call) or
// device memory (read by the kernel).
void fill_entry(int prec, int kind, int eqop, int ncomp, void *r, const void *a, const int *index, int n,
                const char *name)
{
    if (n <= 0) return;
    qla_fill_args p = { r, nullptr, nullptr, index, n };
    if (kind != 0) {
        if (!a) {
            record(hipErrorInvalidValue, name);
            return;
        }
        if (is_device_pointer(a)) p.src = a;
        else p.host = a;
    }
    finish(qla_launch_fill(prec, kind, eqop, ncomp, p, ts.stream), name);
}

// One gamma-matrix multiplication (qla_vecops_abi.inc): right 0 r = Gamma_mu a,
// 1 r = a Gamma_mu.  mu outside 0..15 does nothing but print the reference's
// message (expressions_gamma.pl:464-470, the switch default).
void gamma_entry(int prec, int right, int ncomp, int pa, void *r, const void *a, int mu, const int *index,
                 const int *ir, const int *ia, int n, const char *name)
{
    if (n <= 0) return;
    if (mu < 0 || mu > 15) {
        fprintf(stderr, "Gamma matrix %d not supported\n", mu);
        return;
    }
    qla_vec_args p;
    memset(&p, 0, sizeof(p));
    p.r = r;
    p.x = a;
    p.index = index;
    p.ir = ir;
    p.ix = ia;
    p.n = n;
    p.px = pa;
    p.al = ((uintptr_t)r % 16 == 0 && (uintptr_t)a % 16 == 0) ? 1 : 0;
    finish(qla_launch_gamma(prec, right, mu, ncomp, p, ts.stream), name);
}

// One site-local reduction or colour-matrix part (qla_vecops_abi.inc): kind
// codes of qla_launch_local; a, b sources (b for the dots only).
void local_entry(int prec, int kind, int eqop, int ncomp, int pa, int pb, void *r, const void *a, const void *b,
                 const int *index, const int *ir, const int *ia, const int *ib, int n, const char *name)
{
    if (n <= 0) return;
    qla_vec_args p;
    memset(&p, 0, sizeof(p));
    p.r = r;
    p.x = a;
    p.y = b;
    p.index = index;
    p.ir = ir;
    p.ix = ia;
    p.iy = ib;
    p.n = n;
    p.px = pa;
    p.py = pb;
    p.al = ((uintptr_t)r % 16 == 0 && (uintptr_t)a % 16 == 0 && (uintptr_t)b % 16 == 0) ? 1 : 0;
    finish(qla_launch_local(prec, kind, eqop, ncomp, p, ts.stream), name);
}

// One global sum of a field (qla_vecops_abi.inc): ncr reals per record,
// prec_in the field precision, out_double the result precision; the result
// record (bytes) goes to r in host or device memory; n <= 0 gives zeros.
void sum_entry(int prec_in, int out_double, int ncr, int pa, void *r, const void *a, const int *index, int n,
               size_t bytes, const char *name)
{
    const scratch_buf sc = scratch_for(SCRATCH_SUM, name);
    if (!sc.dev) return;
    if (n < 0) n = 0;
    if (is_device_pointer(r)) {
        finish(qla_launch_sum(prec_in, out_double, ncr, a, index, n, pa, r, sc.dev, ts.stream), name);
        return;
    }
    if (refuse_host_result_in_capture(name)) return;
    // the last kernel writes the result into the pinned slot (bytes <= 288 doubles)
    hipError_t e = qla_launch_sum(prec_in, out_double, ncr, a, index, n, pa, sc.host, sc.dev, ts.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ts.stream);
    if (e == hipSuccess) memcpy(r, sc.host, bytes);
    record(e, name);
}

// One multiplication by a real (fcomplex 0) or complex factor field
// (qla_vecops_abi.inc): r eqop f * b; xform bit 0 conjugates and bit 1
// transposes b (its adjoint), fconj conjugates f.
void field_entry(int prec, int fcomplex, int eqop, int ncomp, int xform, int fconj, int pf, int pb, void *r,
                 const void *f, const void *b, const int *index, const int *ir, const int *ifa, const int *ib, int n,
                 const char *name)
{
    if (n <= 0) return;
    qla_vec_args p;
    memset(&p, 0, sizeof(p));
    p.r = r;
    p.x = b;
    p.px = pb;
    p.ix = ib;
    p.f = f;
    p.pf = pf;
    p.ifa = ifa;
    p.fconj = fconj;
    p.index = index;
    p.ir = ir;
    p.n = n;
    p.conj = xform & 1;
    p.trans = (xform >> 1) & 1;
    finish(qla_launch_vec(prec, fcomplex ? 11 : 10, eqop, ncomp, p, ts.stream), name);
}

// One colour outer product (qla_vecops_abi.inc): M eqop a b^dagger.
void outer_entry(int prec, int eqop, int pa, int pb, void *r, const void *a, const void *b, const int *index,
                 const int *ir, const int *ia, const int *ib, int n, const char *name)
{
    if (n <= 0) return;
    qla_vec_args p;
    memset(&p, 0, sizeof(p));
    p.r = r;
    p.x = a;
    p.y = b;
    p.px = pa;
    p.py = pb;
    p.index = index;
    p.ir = ir;
    p.ix = ia;
    p.iy = ib;
    p.n = n;
    finish(qla_launch_outer(prec, eqop, p, ts.stream), name);
}

// products with propagators: P eqop A * B (kind 0 M*P, 1 P*M, 2 P*P), ca / cb
// the adjoint of A / B; every binary form
void pmul_entry(int prec, int kind, int eqop, int ca, int cb, int pa, int pb, void *r, const void *a,
                const void *b, const int *index, const int *ir, const int *ia, const int *ib, int n, const char *name)
{
    if (n <= 0) return;
    qla_vec_args p;
    memset(&p, 0, sizeof(p));
    p.r = r;
    p.x = a;
    p.y = b;
    p.px = pa;
    p.py = pb;
    p.index = index;
    p.ir = ir;
    p.ix = ia;
    p.iy = ib;
    p.n = n;
    finish(qla_launch_pmul(prec, kind, eqop, ca, cb, p, ts.stream), name);
}

// component access and the spin trace.  kind: 0 / 1 get / set an element,
// 2 / 3 a colour column, 4 / 5 a Dirac column, 6 the spin trace; t the array
// type (0 V, 1 H, 2 D, 3 M, 4 P); e0.. the reference's extra arguments in
// order (ic, is, jc, js, those the type has; a colour column fixes jc of M,
// is of H and D, (is, jc, js) of P; a Dirac column (jc, js)).  An index
// outside its range records hipErrorInvalidValue and writes nothing (the
// reference does not check them).
void comp_entry(int prec, int kind, int t, int e0, int e1, int e2, int e3, int pa, void *r, const void *a,
                const int *index, const int *ir, const int *ia, int n, const char *name)
{
    static const int rec_size[5] = { 3, 6, 12, 9, 144 };
    // position of (ic, is, jc, js) in a record: V [ic], H [ic][2], D [ic][4], M [ic][jc], P [ic][is][jc][js]
    auto pos = [t](int ic, int is, int jc, int js) {
        switch (t) {
        case 0: return ic;
        case 1: return ic * 2 + is;
        case 2: return ic * 4 + is;
        case 3: return ic * 3 + jc;
        default: return ((ic * 4 + is) * 3 + jc) * 4 + js;
        }
    };
    if (t < 0 || t > 4 || kind < 0 || kind > 6) {
        record(hipErrorInvalidValue, name);
        return;
    }
    const int e[4] = { e0, e1, e2, e3 };
    int ic = 0, is = 0, jc = 0, js = 0;
    if (kind <= 1) {
        int k = 0;
        ic = e[k++];
        if (t == 1 || t == 2 || t == 4) is = e[k++];
        if (t == 3 || t == 4) jc = e[k++];
        if (t == 4) js = e[k++];
    } else if (kind <= 3) {
        if (t == 3) jc = e[0];
        else if (t == 1 || t == 2) is = e[0];
        else { is = e[0]; jc = e[1]; js = e[2]; }
    } else if (kind <= 5) {
        jc = e[0];
        js = e[1];
    }
    const int ns = t == 1 ? 2 : 4;
    if (ic < 0 || ic >= 3 || is < 0 || is >= ns || jc < 0 || jc >= 3 || js < 0 || js >= 4) {
        record(hipErrorInvalidValue, name);
        return;
    }
    if (n <= 0) return;
    qla_comp_map m;
    memset(&m, 0, sizeof(m));
    int mode = 2;
    if (kind == 6) {
        m.nd = 9;
        m.nsrc = 144;
        m.ndst = 9;
    } else {
        const bool get = (kind & 1) == 0;
        mode = get ? 0 : 1;
        if (kind <= 1) {
            m.nd = 1;
            m.off = pos(ic, is, jc, js);
            m.inner = 1;
        } else if (kind <= 3) {
            m.nd = 3;
            m.off = pos(0, is, jc, js);
            m.inner = 3;
            m.s2 = pos(1, is, jc, js) - m.off;
        } else {
            m.nd = 12;
            m.off = pos(0, 0, jc, js);
            m.inner = 4;
            m.s1 = pos(1, 0, jc, js) - m.off;
            m.s2 = pos(0, 1, jc, js) - m.off;
        }
        m.nsrc = get ? rec_size[t] : m.nd;
        m.ndst = get ? m.nd : rec_size[t];
    }
    qla_vec_args p;
    memset(&p, 0, sizeof(p));
    p.r = r;
    p.x = a;
    p.px = pa;
    p.index = index;
    p.ir = ir;
    p.ix = ia;
    p.n = n;
    finish(qla_launch_comp(prec, mode, p, m, ts.stream), name);
}

// real / complex conversions and scalar-field division (kind as
// qla_launch_sconv), eq only, every form
void sconv_entry(int prec, int kind, int pa, int pb, void *r, const void *a, const void *b, const int *index,
                 const int *ir, const int *ia, const int *ib, int n, const char *name)
{
    if (n <= 0) return;
    qla_vec_args p;
    memset(&p, 0, sizeof(p));
    p.r = r;
    p.x = a;
    p.y = b;
    p.px = pa;
    p.py = pb;
    p.index = index;
    p.ir = ir;
    p.ix = ia;
    p.iy = ib;
    p.n = n;
    finish(qla_launch_sconv(prec, kind, p, ts.stream), name);
}

// elementary functions of real and complex fields (fn as qla_launch_sfn),
// eq only, every form
void sfn_entry(int prec, int fn, int pa, int pb, void *r, const void *a, const void *b, const int *index,
               const int *ir, const int *ia, const int *ib, int n, const char *name)
{
    if (n <= 0) return;
    qla_vec_args p;
    memset(&p, 0, sizeof(p));
    p.r = r;
    p.x = a;
    p.y = b;
    p.px = pa;
    p.py = pb;
    p.index = index;
    p.ir = ir;
    p.ix = ia;
    p.iy = ib;
    p.n = n;
    finish(qla_launch_sfn(prec, fn, p, ts.stream), name);
}

// copymask (op < 0: nw words per record) and the comparisons of real fields
// (op 0..5), eq only, every form
void mask_entry(int prec, int op, int nw, int pa, int pb, void *r, const void *a, const void *b, const int *index,
                const int *ir, const int *ia, const int *ib, int n, const char *name)
{
    if (n <= 0) return;
    qla_vec_args p;
    memset(&p, 0, sizeof(p));
    p.r = r;
    p.x = a;
    p.y = b;
    p.px = pa;
    p.py = pb;
    p.index = index;
    p.ir = ir;
    p.ix = ia;
    p.iy = ib;
    p.n = n;
    finish(op < 0 ? qla_launch_mask(prec, nw, p, ts.stream) : qla_launch_cmp(prec, op, p, ts.stream), name);
}

// integer fields (op and eqop as qla_int_launch); a: the int constant of the
// i_times forms and of the fill r eqop i, in host or device memory
void int_entry(int op, int eqop, int pa, int pb, void *r, const void *s, const void *a, const void *b,
               const int *index, const int *ir, const int *ia, const int *ib, int n, const char *name)
{
    if (n <= 0) return;
    qla_vec_args p;
    memset(&p, 0, sizeof(p));
    p.r = r;
    p.x = a;
    p.y = b;
    p.px = pa;
    p.py = pb;
    p.index = index;
    p.ir = ir;
    p.ix = ia;
    p.iy = ib;
    p.n = n;
    if (op == 1 || op == 14 || op == 15 || op == 41) {
        if (!s) {
            record(hipErrorInvalidValue, name);
            return;
        }
        if (is_device_pointer(s)) p.sp = s;
        else p.s_re = *(const int *)s;
    }
    finish(qla_int_launch(op, eqop, p, ts.stream), name);
}

// compressed-field reductions: result doubles to r (host or device)
void reduce_dh(const void *a, const void *b, int n, void *r, size_t bytes, const char *name)
{
    const scratch_buf sc = scratch_for(SCRATCH_REDUCE, name);
    if (!sc.dev) return;
    if (n < 0) n = 0;
    if (is_device_pointer(r)) {
        finish(qla_launch_reduce_dh(a, b, n, (double *)r, sc.dev, ts.stream), name);
        return;
    }
    if (refuse_host_result_in_capture(name)) return;
    hipError_t e = qla_launch_reduce_dh(a, b, n, sc.host, sc.dev, ts.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ts.stream);
    if (e == hipSuccess) memcpy(r, sc.host, bytes);
    record(e, name);
}


// ------------------------------------------------------------- site calls
// tiles per wave of the persistent site kernels (launch shape only)
std::atomic<int> g_tiles_per_wave{ 1 };

// One QLA site call (qla_sites_abi.inc): the reference's no-op cases, then
// one launch.  kind 0 matvec, 1 spproj, 2 sprecon, 3 project-multiply-
// reconstruct, 5-7 the matrix-free projection, reconstruction and
// projection-with-reconstruction (their source in b, a unused); mu/sign for
// every kind but 0.  index is the gang index of an x-form, ir / ia / ib the
// per-operand indices of the QLA_INDEX_ALL forms.
void site_entry(int prec, int kind, int eqop, int adj, int pa, int pb, void *r, const void *a, const void *b,
                int mu, int sign, const int *index, const int *ir, const int *ia, const int *ib, int n,
                const char *name)
{
    if (n <= 0) return;
    const bool spin = (kind >= 1 && kind <= 3) || (kind >= 5 && kind <= 7);
    if (spin && (mu < 0 || mu > 4)) return;          // switch without default (expressions_gamma.pl:349-381)
    qla_site_args p;
    p.r = r;
    p.a = a;
    p.b = b;
    p.index = index;
    p.ir = ir;
    p.ia = ia;
    p.ib = ib;
    p.n = n;
    p.al = ((!pa && (uintptr_t)a % 16 == 0) ? 1 : 0) | ((!pb && (uintptr_t)b % 16 == 0) ? 2 : 0) |
           (((uintptr_t)r % 16 == 0) ? 4 : 0) | (((uintptr_t)r % 8 == 0) ? 8 : 0) |
           ((!pa && (uintptr_t)a % 8 == 0) ? 16 : 0);
    p.ms = spin ? mu + (sign == 1 ? 0 : 5) : 0;
    // the SSE3-build order exists for the five F3 v-form symbols that build
    // overrides with a different association (QLA_F3_V_vpeq_M_times_pV,
    // V_veq_M_times_pV, V_veq_Ma_times_V, D_vpeq_sprecon_M_times_pH,
    // D_vpeq_spproj_M_times_pD)
    const bool v_sym = kind == 0 && !pa && ((!adj && pb && (eqop == 0 || eqop == 1)) || (adj && !pb && eqop == 0));
    const bool spin_sym = (kind == 2 || kind == 3) && !adj && !pa && pb && eqop == 1;
    // and the four D3 matvecs it takes from its SSE2 versions (QLA_D3_V_veq_M_times_V,
    // V_veq_M_times_pV, V_vpeq_M_times_V, V_vpeq_M_times_pV)
    const bool d3_sym = kind == 0 && !adj && !pa && (eqop == 0 || eqop == 1);
    p.order = (ts.order == QLA_MI355X_ORDER_SSE3 && !index && !ir && !ia && !ib &&
               ((prec == 0 && (v_sym || spin_sym)) || (prec == 1 && d3_sym)))
                  ? 1 : 0;
    p.pa = pa;
    p.pb = pb;
    const int xf = (index || ir || ia || ib) ? 1 : 0;
    finish(qla_launch_site(prec, kind, eqop, adj, xf, p, g_tiles_per_wave.load(), ts.stream), name);
}

// nd term addresses from host or device memory.  A device-resident list is
// read on the calling thread's stream (after the work already queued there,
// which may have written it); that copy waits on the host, so it cannot be
// part of a stream capture.
hipError_t term_list(const void **dst, const void *src, int nd)
{
    if (!is_device_pointer(src)) {
        memcpy(dst, src, sizeof(void *) * (size_t)nd);
        return hipSuccess;
    }
    if (capturing()) return hipErrorStreamCaptureUnsupported;
    hipError_t e = hipMemcpyAsync(dst, src, sizeof(void *) * (size_t)nd, hipMemcpyDeviceToHost, ts.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ts.stream);
    return e;
}

// One multi-source call: terms in launches of at most QLA_MULTI_MAXND, the
// first with the call's eqop and the rest accumulating (peq for eq/peq, meq
// for meq/eqm) -- the reference's own d >= 1 rule (expressions_tensor.pl:985-1000).
void multi_entry(int prec, int eqop, int adj, int pa, int pb, void *r, const void *a, const void *b,
                 const int *index, const int *ir, const int *ia, const int *ib, int n, int nd, const char *name)
{
    const int xf = (index || ir || ia || ib) ? 1 : 0;
    if (n <= 0 || nd <= 0) return;
    std::vector<const void *> av((size_t)nd), bv((size_t)nd);
    hipError_t e = term_list(av.data(), a, nd);
    if (e == hipSuccess) e = term_list(bv.data(), b, nd);
    for (int d0 = 0; e == hipSuccess && d0 < nd; d0 += QLA_MULTI_MAXND) {
        qla_multi_args p;
        memset(&p, 0, sizeof(p));
        p.r = r;
        p.index = index;
        p.ir = ir;
        p.ia = ia;
        p.ib = ib;
        p.n = n;
        p.nd = std::min(QLA_MULTI_MAXND, nd - d0);
        p.pa = pa;
        p.pb = pb;
        int al_a = 1, al_b = 1;
        for (int d = 0; d < p.nd; d++) {
            p.a[d] = av[(size_t)(d0 + d)];
            p.b[d] = bv[(size_t)(d0 + d)];
            al_a &= (!pa && (uintptr_t)p.a[d] % 16 == 0) ? 1 : 0;
            al_b &= (!pb && (uintptr_t)p.b[d] % 16 == 0) ? 1 : 0;
        }
        p.al = al_a | (al_b << 1) | (((uintptr_t)r % 16 == 0) ? 4 : 0);
        const int op = d0 == 0 ? eqop : ((eqop == 2 || eqop == 3) ? 2 : 1);
        e = qla_launch_multi(prec, op, adj, xf, p, g_tiles_per_wave.load(), ts.stream);
    }
    finish(e, name);
}

// ------------------------------------------------------------- batch plans
struct batch_shape {
    int wpb, tpc, bpc;   // waves per block, tiles per claim, blocks per CU
};

// autotuned shapes, keyed by (log2 of the total tile count, job count bucket)
std::mutex g_tune_mutex;
std::map<std::pair<int, int>, batch_shape> g_tune_cache;

int ilog2(long long v)
{
    int k = 0;
    while (v > 1) { v >>= 1; k++; }
    return k;
}

} // namespace

struct QLA_MI355X_batch {
    int device = 0;
    int njobs = 0;
    int total_tiles = 0;
    int cus = 256;
    qla_batch_job *djobs = nullptr;     // device job table
    unsigned int *counter = nullptr;    // device claim counter
    batch_shape shape = { 1, 1, 0 };   // static schedule, one tile per wave
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

namespace {

hipError_t batch_launch(QLA_MI355X_batch *p, hipStream_t st)
{
    const int blocks = p->shape.bpc == 0 ? 0 : std::max(1, p->cus * p->shape.bpc);   // 0: static schedule
    return qla_launch_batch(p->djobs, p->njobs, p->total_tiles, p->shape.wpb, p->shape.tpc, blocks, p->counter, st);
}

void batch_drop_graph(QLA_MI355X_batch *p)
{
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

} // namespace


// Per-object launchers: qla_site.hip is compiled once per (precision, kind)
// and qla_multi.hip once per precision.
#define QLA_SITE_LAUNCHER(P, K) \
    hipError_t qla_site_launch_p##P##_k##K(int, int, int, const qla_site_args &, int, hipStream_t);
QLA_SITE_LAUNCHER(0, 0) QLA_SITE_LAUNCHER(0, 1) QLA_SITE_LAUNCHER(0, 2) QLA_SITE_LAUNCHER(0, 3)
QLA_SITE_LAUNCHER(1, 0) QLA_SITE_LAUNCHER(1, 1) QLA_SITE_LAUNCHER(1, 2) QLA_SITE_LAUNCHER(1, 3)
QLA_SITE_LAUNCHER(0, 5) QLA_SITE_LAUNCHER(0, 6) QLA_SITE_LAUNCHER(0, 7)
QLA_SITE_LAUNCHER(1, 5) QLA_SITE_LAUNCHER(1, 6) QLA_SITE_LAUNCHER(1, 7)
QLA_SITE_LAUNCHER(0, 8) QLA_SITE_LAUNCHER(0, 9) QLA_SITE_LAUNCHER(0, 10) QLA_SITE_LAUNCHER(0, 11)
QLA_SITE_LAUNCHER(1, 8) QLA_SITE_LAUNCHER(1, 9) QLA_SITE_LAUNCHER(1, 10) QLA_SITE_LAUNCHER(1, 11)
#undef QLA_SITE_LAUNCHER
hipError_t qla_multi_launch_p0(int, int, int, const qla_multi_args &, int, hipStream_t);
hipError_t qla_multi_launch_p1(int, int, int, const qla_multi_args &, int, hipStream_t);

hipError_t qla_launch_site(int prec, int kind, int eqop, int adj, int xf, const qla_site_args &p, int tpw,
                           hipStream_t st)
{
    typedef hipError_t (*launcher)(int, int, int, const qla_site_args &, int, hipStream_t);
    // kinds 0-3 with a matrix, 5-7 matrix-free, 8-11 the colour-matrix
    // products of H, D, M, M^dagger (4 is the multi-source sum)
    static const launcher tab[2][12] = {
        { qla_site_launch_p0_k0, qla_site_launch_p0_k1, qla_site_launch_p0_k2, qla_site_launch_p0_k3, nullptr,
          qla_site_launch_p0_k5, qla_site_launch_p0_k6, qla_site_launch_p0_k7, qla_site_launch_p0_k8,
          qla_site_launch_p0_k9, qla_site_launch_p0_k10, qla_site_launch_p0_k11 },
        { qla_site_launch_p1_k0, qla_site_launch_p1_k1, qla_site_launch_p1_k2, qla_site_launch_p1_k3, nullptr,
          qla_site_launch_p1_k5, qla_site_launch_p1_k6, qla_site_launch_p1_k7, qla_site_launch_p1_k8,
          qla_site_launch_p1_k9, qla_site_launch_p1_k10, qla_site_launch_p1_k11 },
    };
    if (prec < 0 || prec > 1 || kind < 0 || kind > 11 || !tab[prec][kind]) return hipErrorInvalidValue;
    return tab[prec][kind](eqop, adj, xf, p, tpw, st);
}

hipError_t qla_launch_matfunc(int prec, int fn, int nc, const qla_matfunc_args &p, hipStream_t st)
{
    if (prec == 1) return nc == 3 ? qla_matfunc_launch_u1(fn, p, st) : hipErrorInvalidValue;
    switch (nc) {
    case 3: return qla_matfunc_launch_u0(fn, p, st);
    case 4: return qla_matfunc_launch_u2(fn, p, st);
    case 5: return qla_matfunc_launch_u3(fn, p, st);
    case 6: return qla_matfunc_launch_u4(fn, p, st);
    case 7: return qla_matfunc_launch_u5(fn, p, st);
    case 8: return qla_matfunc_launch_u6(fn, p, st);
    case 1: return qla_matfunc_launch_u7(fn, p, st);
    case 2: return qla_matfunc_launch_u8(fn, p, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t qla_launch_multi(int prec, int eqop, int adj, int xf, const qla_multi_args &p, int tpw, hipStream_t st)
{
    if (prec == 0) return qla_multi_launch_p0(eqop, adj, xf, p, tpw, st);
    if (prec == 1) return qla_multi_launch_p1(eqop, adj, xf, p, tpw, st);
    return hipErrorInvalidValue;
}

hipError_t qla_launch_vec(int prec, int op, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    if (prec == 0) return qla_vec_launch_p0(op, eqop, ncomp, p, st);
    if (prec == 1) return qla_vec_launch_p1(op, eqop, ncomp, p, st);
    return hipErrorInvalidValue;
}

hipError_t qla_launch_gamma(int prec, int right, int g, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    if (prec == 0) return qla_gamma_launch_p0(right, g, ncomp, p, st);
    if (prec == 1) return qla_gamma_launch_p1(right, g, ncomp, p, st);
    return hipErrorInvalidValue;
}

hipError_t qla_launch_local(int prec, int kind, int eqop, int ncomp, const qla_vec_args &p, hipStream_t st)
{
    if (prec == 0) return qla_local_launch_p0(kind, eqop, ncomp, p, st);
    if (prec == 1) return qla_local_launch_p1(kind, eqop, ncomp, p, st);
    return hipErrorInvalidValue;
}

hipError_t qla_launch_outer(int prec, int eqop, const qla_vec_args &p, hipStream_t st)
{
    if (prec == 0) return qla_outer_launch_p0(eqop, p, st);
    if (prec == 1) return qla_outer_launch_p1(eqop, p, st);
    return hipErrorInvalidValue;
}

hipError_t qla_launch_mask(int prec, int nw, const qla_vec_args &p, hipStream_t st)
{
    if (prec == 0) return qla_mask_launch_p0(nw, p, st);
    if (prec == 1) return qla_mask_launch_p1(nw, p, st);
    return hipErrorInvalidValue;
}

hipError_t qla_launch_cmp(int prec, int op, const qla_vec_args &p, hipStream_t st)
{
    if (prec == 0) return qla_cmp_launch_p0(op, p, st);
    if (prec == 1) return qla_cmp_launch_p1(op, p, st);
    return hipErrorInvalidValue;
}

hipError_t qla_launch_sfn(int prec, int fn, const qla_vec_args &p, hipStream_t st)
{
    if (prec == 0) return qla_sfn_launch_p0(fn, p, st);
    if (prec == 1) return qla_sfn_launch_p1(fn, p, st);
    return hipErrorInvalidValue;
}

hipError_t qla_launch_sconv(int prec, int kind, const qla_vec_args &p, hipStream_t st)
{
    if (prec == 0) return qla_sconv_launch_p0(kind, p, st);
    if (prec == 1) return qla_sconv_launch_p1(kind, p, st);
    return hipErrorInvalidValue;
}

hipError_t qla_launch_comp(int prec, int mode, const qla_vec_args &p, const qla_comp_map &m, hipStream_t st)
{
    if (prec == 0) return qla_comp_launch_p0(mode, p, m, st);
    if (prec == 1) return qla_comp_launch_p1(mode, p, m, st);
    return hipErrorInvalidValue;
}

hipError_t qla_launch_pmul(int prec, int kind, int eqop, int ca, int cb, const qla_vec_args &p, hipStream_t st)
{
    if (prec == 0) return qla_pmul_launch_p0(kind, eqop, ca, cb, p, st);
    if (prec == 1) return qla_pmul_launch_p1(kind, eqop, ca, cb, p, st);
    return hipErrorInvalidValue;
}

hipError_t qla_launch_fill(int prec, int kind, int eqop, int ncomp, const qla_fill_args &p, hipStream_t st)
{
    if (prec == 0) return qla_fill_launch_p0(kind, eqop, ncomp, p, st);
    if (prec == 1) return qla_fill_launch_p1(kind, eqop, ncomp, p, st);
    return hipErrorInvalidValue;
}

extern "C" {

// ----------------------------------------------------------------- control
void QLA_MI355X_set_stream(void *stream) { ts.stream = (hipStream_t)stream; }
void *QLA_MI355X_get_stream(void) { return (void *)ts.stream; }
void QLA_MI355X_set_sync(int on) { ts.sync = on ? 1 : 0; }
void QLA_MI355X_set_arith_order(int order)
{
    if (order != QLA_MI355X_ORDER_GENERIC && order != QLA_MI355X_ORDER_SSE3) {
        record(hipErrorInvalidValue, "QLA_MI355X_set_arith_order");
        return;
    }
    ts.order = order;
}
int QLA_MI355X_get_arith_order(void) { return ts.order; }
void QLA_MI355X_synchronize(void) { record(hipStreamSynchronize(ts.stream), "synchronize"); }
int QLA_MI355X_last_error(void)
{
    int e = (int)ts.err;
    ts.err = hipSuccess;
    return e;
}
const char *QLA_MI355X_last_error_string(void) { return hipGetErrorString(ts.err); }
void QLA_MI355X_finalize(void)
{
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    int cur = 0;
    const bool have_cur = hipGetDevice(&cur) == hipSuccess;
    for (auto &kv : g_scratch) {
        (void)hipSetDevice(kv.first.dev);
        record(hipFree(kv.second.dev), "QLA_MI355X_finalize");
        record(hipHostFree(kv.second.host), "QLA_MI355X_finalize");
    }
    if (have_cur) (void)hipSetDevice(cur);
    g_scratch.clear();
}
int QLA_MI355X_scratch_buffers(void)
{
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    return (int)g_scratch.size();
}
const char *QLA_MI355X_version(void) { return QLA_MI355X_VERSION_STRING; }
void QLA_MI355X_set_tiles_per_wave(int tpw)
{
    g_tiles_per_wave.store(tpw < 1 ? 1 : tpw);
    qla_set_cmp_tiles_per_wave(tpw);
}

// ------------------------------------------------ the Wilson-hop building blocks
#include "qla_sites_abi.inc"

// ---------------------------------------------- reductions and conversions
#include "qla_fields_abi.inc"

// ------------------------------------------------------ colour-matrix functions
#include "qla_matfuncs_abi.inc"

// ------------------------------------------------------ field linear algebra
#include "qla_vecops_abi.inc"

// ----------------------------------------------------------------- compressed
void QLA_MI355X_F3_M_pack12(void *r, int fmt, QLA_F3_ColorMatrix *a, int n)
{ finish(qla_launch_convert(0, fmt, r, a, n, ts.stream), "QLA_MI355X_F3_M_pack12"); }
void QLA_MI355X_F3_M_unpack12(QLA_F3_ColorMatrix *r, int fmt, void *a, int n)
{ finish(qla_launch_convert(1, fmt, r, a, n, ts.stream), "QLA_MI355X_F3_M_unpack12"); }
void QLA_MI355X_F3_D_packh(void *r, QLA_F3_DiracFermion *a, int n)
{ finish(qla_launch_convert(2, 0, r, a, n, ts.stream), "QLA_MI355X_F3_D_packh"); }
void QLA_MI355X_F3_D_unpackh(QLA_F3_DiracFermion *r, void *a, int n)
{ finish(qla_launch_convert(3, 0, r, a, n, ts.stream), "QLA_MI355X_F3_D_unpackh"); }
void QLA_MI355X_F3_H_packh(void *r, QLA_F3_HalfFermion *a, int n)
{ finish(qla_launch_convert(4, 0, r, a, n, ts.stream), "QLA_MI355X_F3_H_packh"); }
void QLA_MI355X_F3_H_unpackh(QLA_F3_HalfFermion *r, void *a, int n)
{ finish(qla_launch_convert(5, 0, r, a, n, ts.stream), "QLA_MI355X_F3_H_unpackh"); }
void QLA_MI355X_F3_Hh_veq_spproj_M12_times_pDh(void *r, void *a, int fmt, void **b, int mu, int sign, int n)
{
    finish(qla_launch_spproj_c(fmt, r, a, (const void *)b, mu, sign, n, ts.stream),
           "QLA_MI355X_F3_Hh_veq_spproj_M12_times_pDh");
}
void QLA_MI355X_F3_Dh_vpeq_sprecon_M12_times_pHh(void *r, void *a, int fmt, void **b, int mu, int sign, int n)
{
    finish(qla_launch_sprecon_c(fmt, r, a, (const void *)b, mu, sign, n, ts.stream),
           "QLA_MI355X_F3_Dh_vpeq_sprecon_M12_times_pHh");
}
// --------------------------------------- Cayley-Hamilton forms, FP64 coefficients
// function codes 13..15 of qla_launch_matfunc (F3, nc = 3 only)
void QLA_MI355X_F3_M_veq_expA_M_ch(QLA_F3_ColorMatrix *r, QLA_F3_ColorMatrix *a, int n)
{ matfunc_entry(0, 13, 3, 0, 0, r, a, nullptr, nullptr, nullptr, nullptr, nullptr, n, "QLA_MI355X_F3_M_veq_expA_M_ch"); }
void QLA_MI355X_F3_M_xeq_expA_M_ch(QLA_F3_ColorMatrix *r, QLA_F3_ColorMatrix *a, int *index, int n)
{ matfunc_entry(0, 13, 3, 0, 0, r, a, nullptr, index, nullptr, nullptr, nullptr, n, "QLA_MI355X_F3_M_xeq_expA_M_ch"); }
void QLA_MI355X_F3_M_veq_sqrtPH_M_ch(QLA_F3_ColorMatrix *r, QLA_F3_ColorMatrix *a, int n)
{ matfunc_entry(0, 14, 3, 0, 0, r, a, nullptr, nullptr, nullptr, nullptr, nullptr, n, "QLA_MI355X_F3_M_veq_sqrtPH_M_ch"); }
void QLA_MI355X_F3_M_xeq_sqrtPH_M_ch(QLA_F3_ColorMatrix *r, QLA_F3_ColorMatrix *a, int *index, int n)
{ matfunc_entry(0, 14, 3, 0, 0, r, a, nullptr, index, nullptr, nullptr, nullptr, n, "QLA_MI355X_F3_M_xeq_sqrtPH_M_ch"); }
void QLA_MI355X_F3_M_veq_invsqrtPH_M_ch(QLA_F3_ColorMatrix *r, QLA_F3_ColorMatrix *a, int n)
{
    matfunc_entry(0, 15, 3, 0, 0, r, a, nullptr, nullptr, nullptr, nullptr, nullptr, n,
                  "QLA_MI355X_F3_M_veq_invsqrtPH_M_ch");
}
void QLA_MI355X_F3_M_xeq_invsqrtPH_M_ch(QLA_F3_ColorMatrix *r, QLA_F3_ColorMatrix *a, int *index, int n)
{
    matfunc_entry(0, 15, 3, 0, 0, r, a, nullptr, index, nullptr, nullptr, nullptr, n,
                  "QLA_MI355X_F3_M_xeq_invsqrtPH_M_ch");
}
void QLA_MI355X_DF3_r_veq_norm2_D_noalloc(QLA_D_Real *r, QLA_F3_DiracFermion *a, int n)
{
    static const char name[] = "QLA_MI355X_DF3_r_veq_norm2_D_noalloc";
    double *sc = noalloc_scratch(name);
    if (!sc) return;
    const long long m = n > 0 ? 12LL * n : 0;
    // the first stage's grid is capped (1024 blocks) whatever n is, so its
    // partials always fit; checked all the same, before anything launches
    if (qla_reduce2_partials(0, 0, a, nullptr, m) > QLA_REDUCE_SCRATCH_DOUBLES) {
        fprintf(stderr, "qla_mi355x: %s: n = %d needs more partial slots than the static scratch holds (%d)\n", name,
                n, QLA_REDUCE_SCRATCH_DOUBLES);
        record(hipErrorInvalidValue, name);
        return;
    }
    reduce_with(sc, nullptr, [&](void *out, double *s) {
        return qla_launch_reduce2(0, 0, a, nullptr, m, out, 0, s, ts.stream);
    }, r, sizeof(double), name);
}
void QLA_MI355X_DF3_r_veq_norm2_Dh(QLA_D_Real *r, void *a, int n)
{
    reduce_dh(a, nullptr, n, r, sizeof(double), "QLA_MI355X_DF3_r_veq_norm2_Dh");
}
void QLA_MI355X_DF3_c_veq_Dh_dot_Dh(QLA_D_Complex *r, void *a, void *b, int n)
{
    reduce_dh(a, b, n, r, 2 * sizeof(double), "QLA_MI355X_DF3_c_veq_Dh_dot_Dh");
}
// ----------------------------------------------------------------- batches
QLA_MI355X_batch *QLA_MI355X_F3_batch_create(const QLA_MI355X_F3_job *jobs, int njobs)
{
    if (njobs <= 0 || !jobs) { record(hipErrorInvalidValue, "QLA_MI355X_F3_batch_create"); return nullptr; }
    std::vector<qla_batch_job> hj((size_t)njobs);
    long long tiles = 0;
    for (int k = 0; k < njobs; k++) {
        const QLA_MI355X_F3_job &j = jobs[k];
        if (j.op != QLA_MI355X_OP_V_VPEQ_M_TIMES_PV && j.op != QLA_MI355X_OP_H_VEQ_SPPROJ_M_TIMES_PD &&
            j.op != QLA_MI355X_OP_D_VPEQ_SPRECON_M_TIMES_PH) {
            record(hipErrorInvalidValue, "QLA_MI355X_F3_batch_create: unknown op");
            return nullptr;
        }
        qla_batch_job &d = hj[(size_t)k];
        d.r = j.r;
        d.a = j.a;
        d.b = (const void *const *)j.b;
        d.op = j.op;
        d.ms = j.op == QLA_MI355X_OP_V_VPEQ_M_TIMES_PV ? 0 : j.mu + (j.sign == 1 ? 0 : 5);
        d.spare = 0;
        // no-op jobs (n <= 0, or mu outside 0..4 for the spin ops) own no tiles
        const bool noop = j.n <= 0 || (j.op != QLA_MI355X_OP_V_VPEQ_M_TIMES_PV && (j.mu < 0 || j.mu > 4));
        d.n = noop ? 0 : j.n;
        d.tile0 = (int)tiles;
        d.al = (((uintptr_t)j.a % 16 == 0) ? 1 : 0) | (((uintptr_t)j.r % 16 == 0) ? 4 : 0);
        tiles += ((long long)d.n + 63) / 64;
    }
    if (tiles > 0x7fffffffLL) { record(hipErrorInvalidValue, "QLA_MI355X_F3_batch_create: too many sites"); return nullptr; }
    QLA_MI355X_batch *p = new QLA_MI355X_batch;
    p->njobs = njobs;
    p->total_tiles = (int)tiles;
    hipError_t e = hipGetDevice(&p->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->device);
    if (e == hipSuccess) e = hipMalloc((void **)&p->djobs, sizeof(qla_batch_job) * (size_t)njobs);
    if (e == hipSuccess) e = hipMalloc((void **)&p->counter, 256);
    if (e == hipSuccess)
        e = hipMemcpy(p->djobs, hj.data(), sizeof(qla_batch_job) * (size_t)njobs, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        record(e, "QLA_MI355X_F3_batch_create");
        QLA_MI355X_batch_destroy(p);
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        auto it = g_tune_cache.find({ ilog2(p->total_tiles), ilog2(njobs) });
        if (it != g_tune_cache.end()) p->shape = it->second;
    }
    return p;
}

void QLA_MI355X_batch_run(QLA_MI355X_batch *p)
{
    if (!p) return;
    finish(batch_launch(p, ts.stream), "QLA_MI355X_batch_run");
}

int QLA_MI355X_batch_capture(QLA_MI355X_batch *p)
{
    if (!p) return (int)hipErrorInvalidValue;
    batch_drop_graph(p);
    hipStream_t st = ts.stream;
    bool own = false;
    if (!st) {  // the legacy default stream cannot be captured
        hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) { record(e, "QLA_MI355X_batch_capture"); return (int)e; }
        own = true;
    }
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t le = batch_launch(p, st);
        e = hipStreamEndCapture(st, &p->graph);
        if (le != hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (own) (void)hipStreamDestroy(st);
    if (e != hipSuccess) {
        batch_drop_graph(p);
        record(e, "QLA_MI355X_batch_capture");
    }
    return (int)e;
}

void QLA_MI355X_batch_launch_graph(QLA_MI355X_batch *p)
{
    if (!p) return;
    if (!p->exec) { record(hipErrorInvalidValue, "QLA_MI355X_batch_launch_graph: not captured"); return; }
    finish(hipGraphLaunch(p->exec, ts.stream), "QLA_MI355X_batch_launch_graph");
}

float QLA_MI355X_batch_autotune(QLA_MI355X_batch *p, int reps)
{
    if (!p || p->total_tiles == 0) return 0.0f;
    if (reps < 1) reps = 1;
    static const int wpbs[] = { 1, 2, 4 };
    static const int tpcs[] = { 1, 2, 4, 8 };
    static const int bpcs[] = { 0, 2, 4, 8, 16 };     // 0: static schedule
    const int lds = qla_batch_lds_bytes_per_wave();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float best_ms = 1e30f;
    batch_shape best = p->shape;
    for (int wpb : wpbs) {
        if (wpb * lds > 64 * 1024) continue;
        for (int tpc : tpcs) {
            for (int bpc : bpcs) {
                if (e != hipSuccess) break;
                p->shape = { wpb, tpc, bpc };
                e = batch_launch(p, ts.stream);                       // warm-up
                if (e == hipSuccess) e = hipEventRecord(e0, ts.stream);
                for (int r = 0; r < reps && e == hipSuccess; r++) e = batch_launch(p, ts.stream);
                if (e == hipSuccess) e = hipEventRecord(e1, ts.stream);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e == hipSuccess && ms / reps < best_ms) { best_ms = ms / reps; best = p->shape; }
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    p->shape = best;
    record(e, "QLA_MI355X_batch_autotune");
    batch_drop_graph(p);   // a captured graph holds the old shape
    {
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        g_tune_cache[{ ilog2(p->total_tiles), ilog2(p->njobs) }] = best;
    }
    return best_ms;
}

void QLA_MI355X_batch_get_shape(const QLA_MI355X_batch *p, int *waves_per_block, int *tiles_per_claim,
                                int *blocks_per_cu)
{
    if (!p) return;
    if (waves_per_block) *waves_per_block = p->shape.wpb;
    if (tiles_per_claim) *tiles_per_claim = p->shape.tpc;
    if (blocks_per_cu) *blocks_per_cu = p->shape.bpc;
}

int QLA_MI355X_batch_set_shape(QLA_MI355X_batch *p, int waves_per_block, int tiles_per_claim, int blocks_per_cu)
{
    if (!p || (waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4) || tiles_per_claim < 1 ||
        blocks_per_cu < 0)
        return (int)hipErrorInvalidValue;
    p->shape = { waves_per_block, tiles_per_claim, blocks_per_cu };
    batch_drop_graph(p);
    return 0;
}

void QLA_MI355X_batch_destroy(QLA_MI355X_batch *p)
{
    if (!p) return;
    batch_drop_graph(p);
    if (p->djobs) (void)hipFree(p->djobs);
    if (p->counter) (void)hipFree(p->counter);
    delete p;
}

} // extern "C"
