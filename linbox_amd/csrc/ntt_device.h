// Device-side building blocks of the number-theoretic transforms, shared by
// the per-call kernels (ntt.hip) and the batch scheduler (batch.hip).
//
// FFT_direct = DIF, natural order in, bit-reversed order out, and
// FFT_inverse = DIT, bit-reversed in, natural out (fft.h:65-89), with the
// butterflies of Butterfly_DIF / Butterfly_DIT (fft-integral.inl:366-558)
// computed on canonical [0,p) words (see modarith.h).
//
// One polynomial of N = 2^L <= 2^15 words stays in LDS for the whole
// transform.  Each thread owns R = 16 words (32 for L = 15) and applies
// log2(R) radix-2 stages in registers between two LDS exchanges ("passes"), so
// a 2^14-point transform is 4+4+4+2 stages with three LDS round trips.  The
// first pass of the forward transform reads the input with stride N/R (each
// load instruction covers consecutive words across the wave) and the last
// pass writes R consecutive words per thread, so global traffic is one
// coalesced read of the input and one coalesced write of the output.  For
// N < 256*R several polynomials share a 256-thread group.  Inputs of at
// most N/2 coefficients (the zero-padded operands of a product) skip the
// subtraction of the first stage, whose right half is zero.
#pragma once

#include "lbx_internal.h"

namespace lbx {
namespace ntt {


template <int L>
struct Cfg {
    static constexpr int N = 1 << L;
    static constexpr int RLOG = L <= 3 ? L : (L <= 14 ? 4 : 5);
    static constexpr int R = 1 << RLOG;
    static constexpr int TP = N / R;                      // threads per polynomial
    static constexpr int G = TP >= 256 ? 1 : 256 / TP;    // polynomials per workgroup
    static constexpr int BLOCK = G * TP;
    static constexpr int STRIDE = N + (N >> 4);           // padded LDS words per polynomial
    static constexpr size_t LDS_BYTES = L <= 3 ? 0 : (size_t)G * STRIDE * 4;
};

// One pad word after every 16: a wave's accesses at word strides 1, 4, 16 or
// multiples of 64 then fall in distinct banks.
__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }

// Stores of data that only a later kernel reads, once: non-temporal hints
// (streaming stores on gfx950).  Measured at 2^14 points: the forward head's
// and tail's outputs written this way make the tail and the pointwise stage
// that read them 1-3 % faster; the inverse head's own stores got slower and
// the inverse tail's no faster, so those stay ordinary.
__device__ __forceinline__ void store_stream(uint32_t *p, uint32_t v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_stream4(uint32_t *p, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<u32x4 *>(p));
}

// Element owned by slot (u,q) of a thread in a pass whose butterfly groups have
// 2^RP elements spaced 2^HL apart inside blocks of 2^(HL+RP) words.
template <int L, int HL, int RP>
__device__ __forceinline__ int elem(int tp, int u, int q)
{
    constexpr int TP = Cfg<L>::TP;
    const int g = tp + u * TP;
    const int blk = g >> HL, off = g & ((1 << HL) - 1);
    return (blk << (HL + RP)) + off + (q << HL);
}

// Padded LDS word of slot (u,q): for HL >= 4, pad(base + q*2^HL) =
// pad(base) + q*(2^HL + 2^(HL-4)), so one address per group and immediate
// offsets for its 2^RP words.
template <int L, int HL, int RP>
__device__ __forceinline__ int slot_base(int tp, int u)
{
    return pad(elem<L, HL, RP>(tp, u, 0));
}
template <int HL>
constexpr int slot_step()
{
    return (1 << HL) + (HL >= 4 ? (1 << (HL - 4)) : 0);
}

template <int L, int HL, int RP>
__device__ __forceinline__ void scatter(const uint32_t (&x)[Cfg<L>::R], uint32_t *lds, int tp)
{
    constexpr int NG = Cfg<L>::R >> RP;
#pragma unroll
    for (int u = 0; u < NG; ++u) {
        if constexpr (HL >= 4) {
            uint32_t *b = lds + slot_base<L, HL, RP>(tp, u);
#pragma unroll
            for (int q = 0; q < (1 << RP); ++q) b[q * slot_step<HL>()] = x[u * (1 << RP) + q];
        } else {
#pragma unroll
            for (int q = 0; q < (1 << RP); ++q) lds[pad(elem<L, HL, RP>(tp, u, q))] = x[u * (1 << RP) + q];
        }
    }
}

template <int L, int HL, int RP>
__device__ __forceinline__ void gather(uint32_t (&x)[Cfg<L>::R], const uint32_t *lds, int tp)
{
    constexpr int NG = Cfg<L>::R >> RP;
#pragma unroll
    for (int u = 0; u < NG; ++u) {
        if constexpr (HL >= 4) {
            const uint32_t *b = lds + slot_base<L, HL, RP>(tp, u);
#pragma unroll
            for (int q = 0; q < (1 << RP); ++q) x[u * (1 << RP) + q] = b[q * slot_step<HL>()];
        } else {
#pragma unroll
            for (int q = 0; q < (1 << RP); ++q) x[u * (1 << RP) + q] = lds[pad(elem<L, HL, RP>(tp, u, q))];
        }
    }
}

// DIF stages S0..S0+RP-1 (span N>>(s+1) at stage s) on the registers.
// tw_off: the twiddle table starts tw_off entries into the pow_w layout (an
// LDS copy of the tail of the table).
template <int L, int S0, int RP, bool HALFZERO>
__device__ __forceinline__ void dif_pass(uint32_t (&x)[Cfg<L>::R], int tp, const uint2 *__restrict__ tw, uint32_t p,
                                         int tw_off = 0)
{
    constexpr int N = Cfg<L>::N, TP = Cfg<L>::TP, NG = Cfg<L>::R >> RP;
    constexpr int HL = L - S0 - RP;
#pragma unroll
    for (int u = 0; u < NG; ++u) {
        const int off = (tp + u * TP) & ((1 << HL) - 1);
#pragma unroll
        for (int st = 0; st < RP; ++st) {
            const int s = S0 + st;
            const int half = 1 << (RP - 1 - st);
            const uint2 *tws = tw + (N - (N >> s) - tw_off);
#pragma unroll
            for (int q = 0; q < (1 << RP); ++q) {
                if (q & half) continue;
                uint32_t &A = x[u * (1 << RP) + q];
                uint32_t &B = x[u * (1 << RP) + q + half];
                if (s == L - 1) {  // span 1: the only twiddle is w^0 = 1
                    const uint32_t a = A, b = B;
                    A = add_mod(a, b, p);
                    B = sub_mod(a, b, p);
                } else {
                    const uint2 t = tws[off + ((q & (half - 1)) << HL)];
                    if (HALFZERO && s == 0) {  // right half of the input is zero
                        B = mont_mul((int32_t)A, t.x, t.y, p);
                    } else {
                        const uint32_t a = A, b = B;
                        A = add_mod(a, b, p);
                        B = mont_mul((int32_t)(a - b), t.x, t.y, p);
                    }
                }
            }
        }
    }
}

// DIT stages D0..D0+RP-1 (span 2^d at stage d).
template <int L, int D0, int RP>
__device__ __forceinline__ void dit_pass(uint32_t (&x)[Cfg<L>::R], int tp, const uint2 *__restrict__ tw, uint32_t p,
                                         int tw_off = 0)
{
    constexpr int N = Cfg<L>::N, TP = Cfg<L>::TP, NG = Cfg<L>::R >> RP;
#pragma unroll
    for (int u = 0; u < NG; ++u) {
        const int off = (tp + u * TP) & ((1 << D0) - 1);
#pragma unroll
        for (int st = 0; st < RP; ++st) {
            const int d = D0 + st;
            const int half = 1 << st;
            const uint2 *tws = tw + (N - (2 << d) - tw_off);
#pragma unroll
            for (int q = 0; q < (1 << RP); ++q) {
                if (q & half) continue;
                uint32_t &A = x[u * (1 << RP) + q];
                uint32_t &B = x[u * (1 << RP) + q + half];
                const uint32_t a = A;
                uint32_t t = B;
                if (d != 0) {
                    const uint2 w = tws[off + ((q & (half - 1)) << D0)];
                    t = shoup_mul(t, w.x, w.y, p);
                }
                A = add_mod(a, t, p);
                B = sub_mod(a, t, p);
            }
        }
    }
}

// DIF passes over stages [S0, SEND): full passes first, then the remainder.
// On entry x holds the mapping of pass S0.
template <int L, int S0, int SEND, bool HALFZERO>
__device__ __forceinline__ void dif_passes(uint32_t (&x)[Cfg<L>::R], uint32_t *lds, int tp, const uint2 *tw,
                                           uint32_t p, int tw_off = 0)
{
    constexpr int RLOG = Cfg<L>::RLOG;
    constexpr int RP = (SEND - S0) < RLOG ? (SEND - S0) : RLOG;
    dif_pass<L, S0, RP, HALFZERO>(x, tp, tw, p, tw_off);
    if constexpr (S0 + RP < SEND) {
        constexpr int S1 = S0 + RP;
        constexpr int NRP = (SEND - S1) < RLOG ? (SEND - S1) : RLOG;
        if constexpr (S0 > 0) __syncthreads();  // previous gather finished everywhere
        scatter<L, L - S0 - RP, RP>(x, lds, tp);
        __syncthreads();
        gather<L, L - S1 - NRP, NRP>(x, lds, tp);
        dif_passes<L, S1, SEND, false>(x, lds, tp, tw, p, tw_off);
    }
}

// Stages in a pass that starts at D0 and ends at L: the remainder first.
template <int L, int D0>
constexpr int dit_rp()
{
    return ((L - D0) % Cfg<L>::RLOG) ? ((L - D0) % Cfg<L>::RLOG) : Cfg<L>::RLOG;
}

// DIT passes over spans 2^D0 .. 2^(L-1): remainder first, then full passes.
template <int L, int D0>
__device__ __forceinline__ void dit_passes(uint32_t (&x)[Cfg<L>::R], uint32_t *lds, int tp, const uint2 *tw,
                                           uint32_t p)
{
    constexpr int RP = dit_rp<L, D0>();
    dit_pass<L, D0, RP>(x, tp, tw, p);
    if constexpr (D0 + RP < L) {
        constexpr int D1 = D0 + RP;
        if constexpr (D0 > 0) __syncthreads();
        scatter<L, D0, RP>(x, lds, tp);
        __syncthreads();
        gather<L, D1, dit_rp<L, D1>()>(x, lds, tp);
        dit_passes<L, D1>(x, lds, tp, tw, p);
    }
}

// Stages in the last DIF pass of [0, SEND).
template <int L, int SEND>
constexpr int dif_last_rp()
{
    return (SEND % Cfg<L>::RLOG) ? (SEND % Cfg<L>::RLOG) : Cfg<L>::RLOG;
}

// DIF stages [0, SEND) of one polynomial per TP threads.  With SEND == L the
// result is the full transform (R consecutive words per thread are stored as
// 16-byte vectors); with SEND < L the data is left in natural storage order
// as 2^(L-SEND) independent blocks for the remaining stages.
template <int L, int SEND, bool HALFZERO>
__device__ __forceinline__ void dif_body(const DevView &in, size_t entries, const EntryMap &map,
                                         const uint2 *__restrict__ tw, uint32_t p, uint32_t *__restrict__ out,
                                         size_t e_base, int tid, uint32_t *lds_all)
{
    using C = Cfg<L>;
    const int pib = tid / C::TP;
    const int tp = tid % C::TP;
    const size_t e = e_base + pib;  // output row
    const bool active = e < entries;
    const size_t ein = map(e);                           // entry of the input view
    uint32_t *lds = lds_all + pib * C::STRIDE;

    uint32_t x[C::R];
    // first pass mapping (words tp + q*TP when the pass is a full radix-R pass)
    constexpr int RP0 = SEND < C::RLOG ? SEND : C::RLOG;
    if (in.layout == 0) {  // polfirst: one row base, word offsets t
        const uint32_t *src = in.base + (active ? ein : 0) * in.stor + in.shift;
#pragma unroll
        for (int u = 0; u < (C::R >> RP0); ++u)
#pragma unroll
            for (int q = 0; q < (1 << RP0); ++q) {
                const int t = elem<L, L - RP0, RP0>(tp, u, q);
                x[u * (1 << RP0) + q] = (active && (size_t)t < in.size) ? src[t] : 0u;
            }
    } else {
#pragma unroll
        for (int u = 0; u < (C::R >> RP0); ++u)
#pragma unroll
            for (int q = 0; q < (1 << RP0); ++q) {
                const size_t t = (size_t)elem<L, L - RP0, RP0>(tp, u, q);
                x[u * (1 << RP0) + q] = (active && t < in.size) ? in.base[view_index(in, ein, t)] : 0u;
            }
    }
    dif_passes<L, 0, SEND, HALFZERO>(x, lds, tp, tw, p);
    constexpr int RP = dif_last_rp<L, SEND>();
    constexpr int NG = C::R >> RP;
    if (!active) return;
    uint32_t *o = out + e * C::N;
    if constexpr (SEND == L) {
#pragma unroll
        for (int u = 0; u < NG; ++u) {
            const int g = tp + u * C::TP;
            if constexpr (RP >= 2) {
#pragma unroll
                for (int q = 0; q < (1 << RP); q += 4)
                    *reinterpret_cast<uint4 *>(o + (g << RP) + q) =
                        make_uint4(x[u * (1 << RP) + q], x[u * (1 << RP) + q + 1], x[u * (1 << RP) + q + 2],
                                   x[u * (1 << RP) + q + 3]);
            } else {
#pragma unroll
                for (int q = 0; q < (1 << RP); ++q) o[(g << RP) + q] = x[u * (1 << RP) + q];
            }
        }
    } else {  // block-major: [block][entry][2^(L-SEND) words]
#pragma unroll
        for (int u = 0; u < NG; ++u)
#pragma unroll
            for (int q = 0; q < (1 << RP); ++q) {
                const size_t i = (size_t)elem<L, L - SEND, RP>(tp, u, q);
                store_stream(&out[(((i >> (L - SEND)) * entries + e) << (L - SEND)) | (i & ((1u << (L - SEND)) - 1))],
                             x[u * (1 << RP) + q]);
            }
    }
}

// DIT stages [DSTART, L) of one polynomial per TP threads.  DSTART == 0 reads
// bit-reversed input with 16-byte vector loads; DSTART > 0 continues a
// transform whose first DSTART stages were applied blockwise.
template <int L, int DSTART, bool SCALE>
__device__ __forceinline__ void dit_body(const uint32_t *__restrict__ in, size_t entries,
                                         const uint2 *__restrict__ tw, uint32_t p, const DevView &out,
                                         size_t out_first, size_t count, uint32_t scale, uint32_t scale_p,
                                         size_t e_base, int tid, uint32_t *lds_all)
{
    using C = Cfg<L>;
    const int pib = tid / C::TP;
    const int tp = tid % C::TP;
    const size_t e = e_base + pib;
    const bool active = e < entries;
    uint32_t *lds = lds_all + pib * C::STRIDE;

    uint32_t x[C::R];
    constexpr int RP0 = dit_rp<L, DSTART>();
    constexpr int NG0 = C::R >> RP0;
    const uint32_t *src = in + (active ? e : 0) * C::N;
    if constexpr (DSTART == 0 && RP0 >= 2) {
#pragma unroll
        for (int u = 0; u < NG0; ++u) {
            const int g = tp + u * C::TP;
#pragma unroll
            for (int q = 0; q < (1 << RP0); q += 4) {
                const uint4 v = *reinterpret_cast<const uint4 *>(src + (g << RP0) + q);
                x[u * (1 << RP0) + q] = v.x;
                x[u * (1 << RP0) + q + 1] = v.y;
                x[u * (1 << RP0) + q + 2] = v.z;
                x[u * (1 << RP0) + q + 3] = v.w;
            }
        }
    } else {
#pragma unroll
        for (int u = 0; u < NG0; ++u)
#pragma unroll
            for (int q = 0; q < (1 << RP0); ++q) x[u * (1 << RP0) + q] = src[elem<L, DSTART, RP0>(tp, u, q)];
    }
    dit_passes<L, DSTART>(x, lds, tp, tw, p);
    // last pass: a full radix-R pass leaves the thread words tp + q*TP; a
    // transform shorter than one pass leaves groups of 2^(L-DSTART)
    constexpr int RPL = (L - DSTART) >= C::RLOG ? C::RLOG : (L - DSTART);
    if (active && out.layout == 0) {  // polfirst: one row base, word offsets t
        uint32_t *o = out.base + (out_first + e) * out.stor + out.shift;
#pragma unroll
        for (int u = 0; u < (C::R >> RPL); ++u)
#pragma unroll
            for (int q = 0; q < (1 << RPL); ++q) {
                const int t = elem<L, L - RPL, RPL>(tp, u, q);
                if ((size_t)t < count) {
                    uint32_t v = x[u * (1 << RPL) + q];
                    if (SCALE) v = shoup_mul(v, scale, scale_p, p);
                    o[t] = v;
                }
            }
    } else if (active) {
#pragma unroll
        for (int u = 0; u < (C::R >> RPL); ++u)
#pragma unroll
            for (int q = 0; q < (1 << RPL); ++q) {
                const size_t t = (size_t)elem<L, L - RPL, RPL>(tp, u, q);
                if (t < count) {
                    uint32_t v = x[u * (1 << RPL) + q];
                    if (SCALE) v = shoup_mul(v, scale, scale_p, p);
                    out.base[view_index(out, out_first + e, t)] = v;
                }
            }
    }
}

// ---- persistent forward head ------------------------------------------------
// Stages [0, SEND) of many polynomials of N = 2^L >= 2^12 words (one per
// workgroup at a time, polfirst input).  Each thread's twiddles of the first
// pass (they depend only on the thread index) are loaded into registers once;
// the twiddle tables of the later stages (N/2^RLOG - N/2^SEND entries) are
// copied to LDS once.  Inside the loop the only global loads are then the
// next polynomial's input, issued before the current one is transformed, so
// they overlap its butterflies (an in-order wait for a twiddle load would also
// wait for them).
template <int RP>
__device__ __forceinline__ constexpr int first_tw_base(int st)
{
    return (1 << RP) - (1 << (RP - st));  // offsets of stage st's 2^(RP-1-st) twiddles
}

template <int L, int SEND>
struct HeadShape {
    static constexpr int RLOG = Cfg<L>::RLOG, R = Cfg<L>::R, TP = Cfg<L>::TP, N = Cfg<L>::N;
    static constexpr int RP0 = SEND < RLOG ? SEND : RLOG;  // stages of the first pass
    static constexpr int HL0 = L - RP0;
    static constexpr int NT = (1 << RP0) - 1;               // register twiddles per thread (one group)
    static constexpr int LDS_TW = (N >> RP0) - (N >> SEND);  // later stages' table entries
    static constexpr size_t LDS_BYTES = Cfg<L>::LDS_BYTES + (size_t)LDS_TW * 8;
    static_assert(Cfg<L>::G == 1 && (R >> RP0) == 1, "one polynomial per group, one group per thread");
};

template <int L, int SEND, bool HALFZERO>
__device__ __forceinline__ void dif_head_persistent(const DevView &in, size_t entries, const EntryMap &map,
                                                    const uint2 *__restrict__ tw, uint32_t p,
                                                    uint32_t *__restrict__ out, int tp, uint32_t *lds)
{
    using H = HeadShape<L, SEND>;
    using C = Cfg<L>;
    uint2 *tw_lds = reinterpret_cast<uint2 *>(lds + C::STRIDE);
    const int tw_off = H::N - (H::N >> H::RP0);
    for (int i = tp; i < H::LDS_TW; i += C::BLOCK) tw_lds[i] = tw[tw_off + i];
    uint2 t0[H::NT];
#pragma unroll
    for (int st = 0; st < H::RP0; ++st)
#pragma unroll
        for (int j = 0; j < (1 << (H::RP0 - 1 - st)); ++j)
            t0[first_tw_base<H::RP0>(st) + j] = tw[H::N - (H::N >> st) + tp + (j << H::HL0)];
    // words tp + q*TP of the first pass; with HALFZERO only q < R/2 can be nonzero
    constexpr int NLOAD = HALFZERO ? H::R / 2 : H::R;
    uint32_t nxt[NLOAD];
    auto fetch = [&](size_t e) {
        const uint32_t *src = in.base + map(e) * in.stor + in.shift;
#pragma unroll
        for (int q = 0; q < NLOAD; ++q) {
            const size_t t = (size_t)tp + (size_t)q * H::TP;
            nxt[q] = t < in.size ? src[t] : 0u;
        }
    };
    size_t e = blockIdx.x;
    if (e < entries) fetch(e);
    for (; e < entries; e += gridDim.x) {
        uint32_t x[H::R];
#pragma unroll
        for (int q = 0; q < H::R; ++q) x[q] = q < NLOAD ? nxt[q] : 0u;
        if (e + gridDim.x < entries) fetch(e + gridDim.x);
        // first pass from the register twiddles
#pragma unroll
        for (int st = 0; st < H::RP0; ++st) {
            const int half = 1 << (H::RP0 - 1 - st);
#pragma unroll
            for (int q = 0; q < H::R; ++q) {
                if (q & half) continue;
                const uint2 w = t0[first_tw_base<H::RP0>(st) + (q & (half - 1))];
                if (HALFZERO && st == 0) {
                    x[q + half] = mont_mul((int32_t)x[q], w.x, w.y, p);
                } else {
                    const uint32_t a = x[q], b = x[q + half];
                    x[q] = add_mod(a, b, p);
                    x[q + half] = mont_mul((int32_t)(a - b), w.x, w.y, p);
                }
            }
        }
        if constexpr (H::RP0 < SEND) {
            constexpr int S1 = H::RP0;
            constexpr int NRP = (SEND - S1) < H::RLOG ? (SEND - S1) : H::RLOG;
            scatter<L, L - H::RP0, H::RP0>(x, lds, tp);
            __syncthreads();  // also publishes the LDS twiddle table on the first entry
            gather<L, L - S1 - NRP, NRP>(x, lds, tp);
            dif_passes<L, S1, SEND, false>(x, lds, tp, tw_lds, p, tw_off);
        }
        constexpr int RP = dif_last_rp<L, SEND>();
        constexpr int NG = H::R >> RP;
        // block-major output: [block][entry][2^(L-SEND) words]
#pragma unroll
        for (int u = 0; u < NG; ++u)
#pragma unroll
            for (int q = 0; q < (1 << RP); ++q) {
                const size_t i = (size_t)elem<L, L - SEND, RP>(tp, u, q);
                store_stream(&out[(((i >> (L - SEND)) * entries + e) << (L - SEND)) | (i & ((1u << (L - SEND)) - 1))],
                             x[u * (1 << RP) + q]);
            }
        __syncthreads();  // LDS is rewritten by the next polynomial
    }
}

// ---- persistent inverse tail ------------------------------------------------
// DIT stages [DSTART, L) of many polynomials of N = 2^L >= 2^12 words in two
// passes: the first (spans 2^DSTART .. 2^(L-RLOG-1)) reads an LDS copy of its
// twiddle tables (2^(L-RLOG) - 2^DSTART entries), the second (the RLOG largest
// spans) the thread's 2^RLOG - 1 twiddles held in registers.  The next
// polynomial's words are loaded while the current one is transformed.
template <int L, int DSTART>
struct TailShape {
    static constexpr int RLOG = Cfg<L>::RLOG, R = Cfg<L>::R, TP = Cfg<L>::TP, N = Cfg<L>::N;
    static constexpr int D1 = L - RLOG;              // first stage of the last pass
    static constexpr int RP0 = D1 - DSTART;          // stages of the first pass
    static constexpr int LDS_TW = (1 << D1) - (1 << DSTART);
    static constexpr size_t LDS_BYTES = Cfg<L>::LDS_BYTES + (size_t)LDS_TW * 8;
    static_assert(Cfg<L>::G == 1 && RP0 >= 1 && RP0 <= RLOG, "one polynomial per group, two passes");
};

template <int L, int DSTART>
__device__ __forceinline__ void dit_tail_persistent(const uint32_t *__restrict__ in, size_t entries,
                                                    const uint2 *__restrict__ tw, uint32_t p, const DevView &out,
                                                    size_t out_first, size_t count, int tp, uint32_t *lds)
{
    using T = TailShape<L, DSTART>;
    using C = Cfg<L>;
    uint2 *tw_lds = reinterpret_cast<uint2 *>(lds + C::STRIDE);
    const int tw_off = T::N - (1 << T::D1);
    for (int i = tp; i < T::LDS_TW; i += C::BLOCK) tw_lds[i] = tw[tw_off + i];
    __syncthreads();  // the first pass reads the table before any other barrier
    uint2 t1[(1 << T::RLOG) - 1];
#pragma unroll
    for (int st = 0; st < T::RLOG; ++st)
#pragma unroll
        for (int j = 0; j < (1 << st); ++j)
            t1[(1 << st) - 1 + j] = tw[T::N - (2 << (T::D1 + st)) + tp + (j << T::D1)];
    constexpr int NG0 = T::R >> T::RP0;
    uint32_t nxt[T::R];
    auto fetch = [&](size_t e) {
        const uint32_t *src = in + e * T::N;
#pragma unroll
        for (int u = 0; u < NG0; ++u)
#pragma unroll
            for (int q = 0; q < (1 << T::RP0); ++q) nxt[u * (1 << T::RP0) + q] = src[elem<L, DSTART, T::RP0>(tp, u, q)];
    };
    size_t e = blockIdx.x;
    if (e < entries) fetch(e);
    for (; e < entries; e += gridDim.x) {
        uint32_t x[T::R];
#pragma unroll
        for (int q = 0; q < T::R; ++q) x[q] = nxt[q];
        if (e + gridDim.x < entries) fetch(e + gridDim.x);
        dit_pass<L, DSTART, T::RP0>(x, tp, tw_lds, p, tw_off);
        scatter<L, DSTART, T::RP0>(x, lds, tp);
        __syncthreads();
        gather<L, T::D1, T::RLOG>(x, lds, tp);
        // last pass: spans 2^D1 .. 2^(L-1), one group per thread (offset tp)
#pragma unroll
        for (int st = 0; st < T::RLOG; ++st) {
            const int half = 1 << st;
#pragma unroll
            for (int q = 0; q < T::R; ++q) {
                if (q & half) continue;
                const uint2 w = t1[half - 1 + (q & (half - 1))];
                const uint32_t a = x[q];
                const uint32_t t = shoup_mul(x[q + half], w.x, w.y, p);
                x[q] = add_mod(a, t, p);
                x[q + half] = sub_mod(a, t, p);
            }
        }
        uint32_t *o = out.base + (out_first + e) * out.stor + out.shift;  // polfirst output
#pragma unroll
        for (int q = 0; q < T::R; ++q) {
            const int t = tp + q * T::TP;
            if ((size_t)t < count) o[t] = x[q];
        }
        __syncthreads();  // LDS is rewritten by the next polynomial
    }
}

}  // namespace ntt
}  // namespace lbx
