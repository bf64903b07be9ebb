// qla_tile_ops.hpp -- one 64-site tile of a site operation, split into a
// stage step (issue the direct-to-LDS loads of the tile's records) and a
// finish step (wait, transpose out of LDS, compute in the reference order,
// write the result back through LDS).  The per-call kernels and the batched
// persistent kernel are loops over these steps; between the two steps a loop
// issues the next tile's gather-address loads so they travel with the
// current tile's records.
//
// One generic operation covers the QLA Wilson-hop building blocks
// (build_qla.pl:1162-1250, make_code_binary.pl:189-350):
//   K_MATVEC   V  eqop A * B               B: V      (expressions_tensor.pl:544-630)
//   K_SPPROJ   H  eqop A * spproj(B)       B: D      (make_code_binary.pl:189-222)
//   K_SPRECON  D  eqop sprecon(A * B)      B: H      (make_code_binary.pl:224-257)
//   K_WILSON   D  eqop sprecon(A * spproj(B))  B: D  (make_code_binary.pl:285-341)
// and the matrix-free spin operations (build_qla.pl:1087-1160), whose one
// source takes the B slot:
//   K_PROJ      H  eqop spproj(B)          B: D      (make_code_unary.pl:183-202)
//   K_RECON     D  eqop sprecon(B)         B: H      (make_code_unary.pl:204-223)
//   K_PROJRECON D  eqop sprecon(spproj(B)) B: D      (make_code_unary.pl:243-283)
// and the colour-matrix products of the other left-multiplied types
// (build_qla.pl:1587-1622, expressions_tensor.pl:544-630):
//   K_MAT_H  H eqop A * B   K_MAT_D  D eqop A * B   K_MAT_M  M eqop A * B
//   K_MAT_MA M eqop A * B^dagger
// with A = M or M^dagger (the `Ma` forms), A and B either fields or pointer
// arrays (the `pM` / `pB` forms), and the v-form (site i) or the gang-indexed
// x-form (site index[i]).  Precision, kind, eqop, adjoint and x-form are
// compile-time; mu/sign and the pointer flags are wave-uniform run-time
// values (a scalar branch per tile).
#pragma once
#include <type_traits>
#include "qla_dev.hpp"
#include "qla_launch.hpp"

namespace qla {

template <class T> struct rec {
    static constexpr int C = (int)sizeof(cpx<T>);
    static constexpr int V = 3 * C, H = 6 * C, D = 12 * C, M = 9 * C;
};

// Wait for this wave's direct-to-LDS loads, then order LDS across lanes.
__device__ __forceinline__ void dma_wait()
{
    vm_wait();
    wave_sync();
}

// ---------------------------------------------------------------------------
// Tile movers.  LDS always holds packed records (site s at s*BPS): 16-byte
// chunks where the addresses allow, 4-byte chunks otherwise (both land
// packed in LDS).  `al16` flags of contiguous operands come from the host;
// gathered addresses are checked per wave.
// ---------------------------------------------------------------------------
template <int BPS>
__device__ __forceinline__ void load_tile(char *lds, const char *g, int nsites, int lane, bool al16)
{
    if (al16) stage_bytes<16, WAVE * BPS>(lds, g, nsites * BPS, lane);
    else stage_bytes<4, WAVE * BPS>(lds, g, nsites * BPS, lane);
}
template <int BPS>
__device__ __forceinline__ void gather_tile(char *lds, uint64_t addr, int nsites, int lane)
{
    if constexpr (BPS % 16 == 0) {
        if (__all((addr & 15) == 0)) { stage_gather<BPS, 16>(lds, addr, nsites, lane); return; }
    }
    stage_gather<BPS, 4>(lds, addr, nsites, lane);
}
template <int BPS>
__device__ __forceinline__ void store_tile(char *g, const char *lds, int nsites, int lane, bool al16)
{
    if (al16) store_bytes<16, WAVE * BPS>(g, lds, nsites * BPS, lane);
    else store_bytes<4, WAVE * BPS>(g, lds, nsites * BPS, lane);
}
template <int BPS>
__device__ __forceinline__ void scatter_tile(uint64_t addr, const char *lds, int nsites, int lane)
{
    if constexpr (BPS % 16 == 0) {
        if (__all((addr & 15) == 0)) { store_scatter<BPS, 16>(addr, lds, nsites, lane); return; }
    }
    store_scatter<BPS, 4>(addr, lds, nsites, lane);
}

template <class T, int N>
__device__ __forceinline__ void lds_read(cpx<T> (&v)[N], const char *p)
{
    const cpx<T> *q = (const cpx<T> *)p;
#pragma unroll
    for (int i = 0; i < N; i++) v[i] = q[i];
}
template <class T, int N>
__device__ __forceinline__ void lds_write(char *p, const cpx<T> (&v)[N])
{
    cpx<T> *q = (cpx<T> *)p;
#pragma unroll
    for (int i = 0; i < N; i++) q[i] = v[i];
}

// alignment flags of the contiguous operands (bit set: 16-byte aligned;
// AL_R8 / AL_A8: r / a 8-byte aligned, which the SSE-order dispatch reads)
enum { AL_A = 1, AL_B = 2, AL_R = 4, AL_R8 = 8, AL_A8 = 16 };

// kinds of site operation
enum {
    K_MATVEC = 0, K_SPPROJ = 1, K_SPRECON = 2, K_WILSON = 3, K_PROJ = 5, K_RECON = 6, K_PROJRECON = 7,
    K_MAT_H = 8, K_MAT_D = 9, K_MAT_M = 10, K_MAT_MA = 11
};
template <int KIND> struct kind_unary { static constexpr bool v = KIND >= K_PROJ && KIND <= K_PROJRECON; };

// record bytes of the second source (B) and of the destination (R)
template <class T, int KIND> struct kind_rec;
template <class T> struct kind_rec<T, K_MATVEC>  { static constexpr int B = rec<T>::V, R = rec<T>::V; };
template <class T> struct kind_rec<T, K_SPPROJ>  { static constexpr int B = rec<T>::D, R = rec<T>::H; };
template <class T> struct kind_rec<T, K_SPRECON> { static constexpr int B = rec<T>::H, R = rec<T>::D; };
template <class T> struct kind_rec<T, K_WILSON>  { static constexpr int B = rec<T>::D, R = rec<T>::D; };
template <class T> struct kind_rec<T, K_PROJ>    { static constexpr int B = rec<T>::D, R = rec<T>::H; };
template <class T> struct kind_rec<T, K_RECON>   { static constexpr int B = rec<T>::H, R = rec<T>::D; };
template <class T> struct kind_rec<T, K_PROJRECON> { static constexpr int B = rec<T>::D, R = rec<T>::D; };
template <class T> struct kind_rec<T, K_MAT_H>   { static constexpr int B = rec<T>::H, R = rec<T>::H; };
template <class T> struct kind_rec<T, K_MAT_D>   { static constexpr int B = rec<T>::D, R = rec<T>::D; };
template <class T> struct kind_rec<T, K_MAT_M>   { static constexpr int B = rec<T>::M, R = rec<T>::M; };
template <class T> struct kind_rec<T, K_MAT_MA>  { static constexpr int B = rec<T>::M, R = rec<T>::M; };

// mu/sign code: ms = mu + 5 * (sign != 1), 0..9
template <int K> struct ms_c {
    static constexpr int MU = K % 5, SIGN = K < 5 ? 1 : -1;
};
// f(ms_c<ms>) with ms compile-time (MS >= 0) or run-time (MS < 0, one
// scalar branch; codes outside 0..9 never reach a kernel)
template <int MS, class F>
__device__ __forceinline__ void with_ms(int ms, F &&f)
{
    if constexpr (MS >= 0) {
        f(ms_c<MS>{});
    } else {
        switch (ms) {
        case 0: f(ms_c<0>{}); break;
        case 1: f(ms_c<1>{}); break;
        case 2: f(ms_c<2>{}); break;
        case 3: f(ms_c<3>{}); break;
        case 4: f(ms_c<4>{}); break;
        case 5: f(ms_c<5>{}); break;
        case 6: f(ms_c<6>{}); break;
        case 7: f(ms_c<7>{}); break;
        case 8: f(ms_c<8>{}); break;
        default: f(ms_c<9>{}); break;
        }
    }
}

// Operands of one site call (a kernel argument: wave-uniform, in SGPRs).
typedef qla_site_args site_args;

// Per-lane record addresses of one tile: gathered A and B records (pointer
// operands, or every operand of the per-lane-addressed forms) and the
// destination of those forms.
struct lane_addr {
    uint64_t ga, gb, gr;
};

// Per-lane record indices of the three operands (XF forms): the operand's
// own index entry (QLA_INDEX_ALL forms), else the gang index entry, else i.
struct lane_sites {
    int r, a, b;
};

template <bool XF>
__device__ __forceinline__ lane_sites sites_of(const site_args &p, int i, bool act)
{
    lane_sites s = { i, i, i };
    if (XF && act) {
        const int g = p.index ? p.index[i] : i;
        s.r = p.ir ? p.ir[i] : g;
        s.a = p.ia ? p.ia[i] : g;
        s.b = p.ib ? p.ib[i] : g;
    }
    return s;
}

template <class T, int KIND, bool XF>
__device__ __forceinline__ lane_addr addresses_of(const site_args &p, const lane_sites &s, bool act)
{
    typedef kind_rec<T, KIND> K;
    lane_addr la = { 0, 0, 0 };
    if (act) {
        if constexpr (!kind_unary<KIND>::v) {
            if (p.pa) la.ga = (uint64_t)((const void *const *)p.a)[s.a];
            else if (XF) la.ga = (uint64_t)((const char *)p.a + (size_t)s.a * rec<T>::M);
        }
        if (p.pb) la.gb = (uint64_t)((const void *const *)p.b)[s.b];
        else if (XF) la.gb = (uint64_t)((const char *)p.b + (size_t)s.b * K::B);
        if (XF) la.gr = (uint64_t)((char *)p.r + (size_t)s.r * K::R);
    }
    return la;
}

// ---------------------------------------------------------------------------
// The SSE3-build order (SSEO): which order site i of a v-form call gets.
// The V symbols split the call as the reference's dispatch does
// (QLA_F3_V_vpeq_M_times_pV.c:59-103): nb1 = 1 when r and a are 8-byte but r
// not 16-byte aligned; sites [nb1, nb1 + n1), n1 = n - nb1 rounded down to
// even, run its 2-site body (vector order in every row), the others its
// 1-site body (vector order in rows 0-1, the row-2 expression in row 2).
// sprecon / spproj take the SSE multiply when r is 16-, a 8- and b[0]
// 16-byte aligned (QLA_F3_D_vpeq_sprecon_M_times_pH.c:108), else the scalar
// fallback's.
// ---------------------------------------------------------------------------
template <class T, int KIND, int EQOP, bool ADJ, bool XF> struct sse_order {
    static constexpr bool v = !XF && ((std::is_same<T, float>::value &&
                                       ((KIND == K_MATVEC && !ADJ && (EQOP == PEQ || EQOP == EQ)) ||
                                        (KIND == K_MATVEC && ADJ && EQOP == EQ) ||
                                        ((KIND == K_SPRECON || KIND == K_WILSON) && !ADJ && EQOP == PEQ))) ||
                                      // the D3 matvecs of the SSE2 versions (field or pointer source)
                                      (std::is_same<T, double>::value && KIND == K_MATVEC && !ADJ &&
                                       (EQOP == PEQ || EQOP == EQ)));
};

__device__ __forceinline__ bool sse_two_site(const site_args &p, int i)
{
    const int nb1 = ((p.al & AL_R8) && (p.al & AL_A8) && !(p.al & AL_R)) ? 1 : 0;
    const int n1 = (p.n - nb1) & ~1;
    return i >= nb1 && i < nb1 + n1;
}

// r[ic] (eqop) A x for the three V symbols in the SSE build's order
template <int EQOP, bool ADJ, class T>
__device__ __forceinline__ void sse_matvec(cpx<T> *r, const cpx<T> *a, const cpx<T> *x, bool two)
{
    if constexpr (ADJ) {
        r[0] = sse_adj_pair(a + 0, x);
        r[1] = sse_adj_pair(a + 1, x);
        r[2] = two ? sse_adj_pair(a + 2, x) : sse_adj_row2(a + 2, x);
    } else if constexpr (EQOP == PEQ) {
        r[0] = sse_pair(r[0], a, x);
        r[1] = sse_pair(r[1], a + 3, x);
        r[2] = two ? sse_pair(r[2], a + 6, x) : sse_row2<true>(r[2], a + 6, x);
    } else {
        r[0] = sse_half<1>(a, x);
        r[1] = sse_half<1>(a + 3, x);
        r[2] = two ? sse_half<1>(a + 6, x) : sse_row2<false>(r[2], a + 6, x);
    }
}

// t[ic][s] = sum_k a[ic][k] x[k][s] (2 spins) in the SSE multiply or the
// scalar fallback order
template <class T>
__device__ __forceinline__ void sse_half_cols(cpx<T> *t, const cpx<T> *a, const cpx<T> *x, bool aligned)
{
#pragma unroll
    for (int ic = 0; ic < 3; ic++) {
#pragma unroll
        for (int is = 0; is < 2; is++)
            t[2 * ic + is] = aligned ? sse_half<2>(a + 3 * ic, x + is) : sse_scalar<2>(a + 3 * ic, x + is);
    }
}

// ---------------------------------------------------------------------------
// The tile operation.  LDS per wave: A tile (none for the matrix-free
// kinds), B tile, and (peq/meq) the destination tile; eq/eqm write their
// result over the A/B region when it fits there, else after it.  SSEO
// selects the SSE3-build order for the symbols sse_order covers.
// ---------------------------------------------------------------------------
template <class T, int KIND, int EQOP, bool ADJ, bool XF, int MS = -1, bool SSEO = false> struct site_op {
    static_assert(!SSEO || sse_order<T, KIND, EQOP, ADJ, XF>::v, "no SSE-build order for this form");
    typedef rec<T> R;
    typedef kind_rec<T, KIND> K;
    static constexpr bool UNARY = kind_unary<KIND>::v;
    static constexpr int A_B = UNARY ? 0 : R::M;          // A record bytes held in LDS
    static constexpr bool READ_R = (EQOP == PEQ || EQOP == MEQ);
    static constexpr bool R_APART = READ_R || K::R > A_B + K::B;
    static constexpr int OFF_B = WAVE * A_B;
    static constexpr int OFF_R = R_APART ? WAVE * (A_B + K::B) : 0;
    static constexpr int BYTES = WAVE * (A_B + K::B + (R_APART ? K::R : 0));

    static __device__ __forceinline__ void stage(char *lds, const site_args &p, const lane_addr &la, int base,
                                                 int ns, int lane)
    {
        if constexpr (!UNARY) {
            if (!XF && !p.pa) load_tile<R::M>(lds, (const char *)p.a + (size_t)base * R::M, ns, lane, p.al & AL_A);
            else gather_tile<R::M>(lds, la.ga, ns, lane);
        }
        if (!XF && !p.pb) load_tile<K::B>(lds + OFF_B, (const char *)p.b + (size_t)base * K::B, ns, lane, p.al & AL_B);
        else gather_tile<K::B>(lds + OFF_B, la.gb, ns, lane);
        if (READ_R) {
            if (!XF) load_tile<K::R>(lds + OFF_R, (const char *)p.r + (size_t)base * K::R, ns, lane, p.al & AL_R);
            else gather_tile<K::R>(lds + OFF_R, la.gr, ns, lane);
        }
    }

    static __device__ __forceinline__ void finish(char *lds, const site_args &p, const lane_addr &la, int base,
                                                  int ns, int lane)
    {
        constexpr int NR = K::R / R::C;
        dma_wait();
        cpx<T> m[9], o[NR];
        if constexpr (!UNARY) lds_read(m, lds + lane * R::M);
        const char *lb = lds + OFF_B + lane * K::B;
        if (READ_R) lds_read(o, lds + OFF_R + lane * K::R);
        if constexpr (KIND == K_PROJ) {
            cpx<T> d[12];
            lds_read(d, lb);
            with_ms<MS>(p.ms, [&](auto c) {
#pragma unroll
                for (int ic = 0; ic < 3; ic++)
                    spproj_row_eqop<EQOP, decltype(c)::MU, decltype(c)::SIGN>(o[2 * ic], o[2 * ic + 1], d + 4 * ic);
            });
        } else if constexpr (KIND == K_RECON || KIND == K_PROJRECON) {
            cpx<T> tv[6];
            if constexpr (KIND == K_RECON) {
                lds_read(tv, lb);
            } else {
                cpx<T> d[12];
                lds_read(d, lb);
                with_ms<MS>(p.ms, [&](auto c) {
#pragma unroll
                    for (int ic = 0; ic < 3; ic++)
                        spproj_row<decltype(c)::MU, decltype(c)::SIGN>(tv[2 * ic], tv[2 * ic + 1], d + 4 * ic);
                });
            }
            with_ms<MS>(p.ms, [&](auto c) {
                typedef recon_tab<decltype(c)::MU, decltype(c)::SIGN> TAB;
#pragma unroll
                for (int ic = 0; ic < 3; ic++) {
                    recon_apply<EQOP, TAB::op[0]>(o[4 * ic + 0], tv[2 * ic + TAB::src[0]]);
                    recon_apply<EQOP, TAB::op[1]>(o[4 * ic + 1], tv[2 * ic + TAB::src[1]]);
                    recon_apply<EQOP, TAB::op[2]>(o[4 * ic + 2], tv[2 * ic + TAB::src[2]]);
                    recon_apply<EQOP, TAB::op[3]>(o[4 * ic + 3], tv[2 * ic + TAB::src[3]]);
                }
            });
        } else if constexpr (KIND == K_MATVEC) {
            cpx<T> v[3];
            lds_read(v, lb);
            if constexpr (SSEO && std::is_same<T, double>::value) sse2_d3_matvec<EQOP>(o, m, v);
            else if constexpr (SSEO) sse_matvec<EQOP, ADJ>(o, m, v, sse_two_site(p, base + lane));
            else mat_times_cols<EQOP, ADJ, 1>(o, m, v);
        } else if constexpr (KIND == K_MAT_H || KIND == K_MAT_D) {
            constexpr int NS = KIND == K_MAT_H ? 2 : 4;
            cpx<T> v[3 * NS];
            lds_read(v, lb);
            mat_times_cols<EQOP, ADJ, NS>(o, m, v);
        } else if constexpr (KIND == K_MAT_M || KIND == K_MAT_MA) {
            cpx<T> bm[9];
            lds_read(bm, lb);
            mat_times_mat<EQOP, ADJ, KIND == K_MAT_MA>(o, m, bm);
        } else if constexpr (KIND == K_SPPROJ) {
            cpx<T> d[12], tp[6];
            lds_read(d, lb);
            with_ms<MS>(p.ms, [&](auto c) {
#pragma unroll
                for (int ic = 0; ic < 3; ic++)
                    spproj_row<decltype(c)::MU, decltype(c)::SIGN>(tp[2 * ic], tp[2 * ic + 1], d + 4 * ic);
            });
            mat_times_cols<EQOP, ADJ, 2>(o, m, tp);
        } else {
            cpx<T> tv[6];
            // SSE order: the reference's alignment test, with b[0] the first
            // pointer-array entry (a wave-uniform load)
            bool sse_al = false;
            if constexpr (SSEO)
                sse_al = (p.al & AL_R) && (p.al & AL_A8) && ((*(const uint64_t *)p.b) & 15) == 0;
            if constexpr (KIND == K_SPRECON) {
                cpx<T> h[6];
                lds_read(h, lb);
                if constexpr (SSEO) sse_half_cols(tv, m, h, sse_al);
                else mat_times_cols<EQ, ADJ, 2>(tv, m, h);
            } else {
                cpx<T> d[12], tp[6];
                lds_read(d, lb);
                with_ms<MS>(p.ms, [&](auto c) {
#pragma unroll
                    for (int ic = 0; ic < 3; ic++)
                        spproj_row<decltype(c)::MU, decltype(c)::SIGN>(tp[2 * ic], tp[2 * ic + 1], d + 4 * ic);
                });
                if constexpr (SSEO) sse_half_cols(tv, m, tp, sse_al);
                else mat_times_cols<EQ, ADJ, 2>(tv, m, tp);
            }
            with_ms<MS>(p.ms, [&](auto c) {
                typedef recon_tab<decltype(c)::MU, decltype(c)::SIGN> TAB;
#pragma unroll
                for (int ic = 0; ic < 3; ic++) {
                    recon_apply<EQOP, TAB::op[0]>(o[4 * ic + 0], tv[2 * ic + TAB::src[0]]);
                    recon_apply<EQOP, TAB::op[1]>(o[4 * ic + 1], tv[2 * ic + TAB::src[1]]);
                    recon_apply<EQOP, TAB::op[2]>(o[4 * ic + 2], tv[2 * ic + TAB::src[2]]);
                    recon_apply<EQOP, TAB::op[3]>(o[4 * ic + 3], tv[2 * ic + TAB::src[3]]);
                }
            });
        }
        char *lr = lds + OFF_R;
        wave_sync();
        lds_write(lr + lane * K::R, o);
        wave_sync();
        if (!XF) store_tile<K::R>((char *)p.r + (size_t)base * K::R, lr, ns, lane, p.al & AL_R);
        else scatter_tile<K::R>(la.gr, lr, ns, lane);
        wave_sync();
    }
};

// ---------------------------------------------------------------------------
// Multi-source colour matrix times vector:  r eqop sum_d A_d * B_d
// (V_eqop_nM_times_nV and its pointer forms; build_qla.pl:1693-1725,
// expressions_tensor.pl:983-1002).  The reference applies eqop for d = 0
// and peq (eq) / meq (eqm) for d >= 1, storing r between terms; r holds the
// same-precision accumulator, so one register accumulator over d and kc
// gives identical bits.
// ---------------------------------------------------------------------------
constexpr int MULTI_MAXND = QLA_MULTI_MAXND;
typedef qla_multi_args multi_args;

template <class T, int EQOP, bool ADJ, bool XF> struct multi_op {
    typedef rec<T> R;
    static constexpr bool READ_R = (EQOP == PEQ || EQOP == MEQ);
    static constexpr bool SUB = (EQOP == MEQ || EQOP == EQM);
    static constexpr int OFF_B = WAVE * R::M;
    static constexpr int OFF_R = WAVE * (R::M + R::V);
    static constexpr int BYTES = WAVE * (R::M + R::V + (READ_R ? R::V : 0));

    // this lane's record indices of step i (XF forms: per-operand index,
    // else gang index, else i)
    static __device__ __forceinline__ lane_sites sites(const multi_args &p, int i, bool act)
    {
        lane_sites s = { i, i, i };
        if (XF && act) {
            const int g = p.index ? p.index[i] : i;
            s.r = p.ir ? p.ir[i] : g;
            s.a = p.ia ? p.ia[i] : g;
            s.b = p.ib ? p.ib[i] : g;
        }
        return s;
    }

    // gathered addresses of term d's records for one lane
    static __device__ __forceinline__ void term_addresses(const multi_args &p, int d, const lane_sites &s, bool act,
                                                          uint64_t &ga, uint64_t &gb)
    {
        ga = 0;
        gb = 0;
        if (!act) return;
        if (p.pa) ga = (uint64_t)((const void *const *)p.a[d])[s.a];
        else if (XF) ga = (uint64_t)((const char *)p.a[d] + (size_t)s.a * R::M);
        if (p.pb) gb = (uint64_t)((const void *const *)p.b[d])[s.b];
        else if (XF) gb = (uint64_t)((const char *)p.b[d] + (size_t)s.b * R::V);
    }

    static __device__ __forceinline__ void stage_term(char *lds, const multi_args &p, int d, uint64_t ga,
                                                      uint64_t gb, int base, int ns, int lane)
    {
        if (!XF && !p.pa) load_tile<R::M>(lds, (const char *)p.a[d] + (size_t)base * R::M, ns, lane, p.al & AL_A);
        else gather_tile<R::M>(lds, ga, ns, lane);
        if (!XF && !p.pb) load_tile<R::V>(lds + OFF_B, (const char *)p.b[d] + (size_t)base * R::V, ns, lane, p.al & AL_B);
        else gather_tile<R::V>(lds + OFF_B, gb, ns, lane);
    }

    static __device__ __forceinline__ void accumulate(cpx<T> (&x)[3], const char *lds, int lane)
    {
        cpx<T> m[9], v[3];
        lds_read(m, lds + lane * R::M);
        lds_read(v, lds + OFF_B + lane * R::V);
#pragma unroll
        for (int ic = 0; ic < 3; ic++)
#pragma unroll
            for (int kc = 0; kc < 3; kc++)
                acc_mul<SUB, ADJ>(x[ic], ADJ ? m[kc * 3 + ic] : m[ic * 3 + kc], v[kc]);
    }
};

} // namespace qla
