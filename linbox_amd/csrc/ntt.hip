// Number-theoretic transforms for gfx950.
//
// FFT_direct = DIF, natural order in, bit-reversed order out, and
// FFT_inverse = DIT, bit-reversed in, natural out (fft.h:65-89), with the
// butterflies of Butterfly_DIF / Butterfly_DIT (fft-integral.inl:366-558)
// computed on canonical [0,p) words (see modarith.h).
//
// Design: one polynomial of N = 2^L <= 2^15 words stays in LDS for the whole
// transform.  Each thread owns R = 16 words (32 for L = 15) and applies
// log2(R) radix-2 stages in registers between two LDS exchanges ("passes"), so
// a 2^14-point transform is 4+4+4+2 stages with three LDS round trips.  The
// first pass of the forward transform reads the input with stride N/R (each
// load instruction covers consecutive words across the wave) and the last
// pass writes R consecutive words per thread, so global traffic is one
// coalesced read of the input and one coalesced write of the output.  For
// N < 256*R several polynomials share a 256-thread workgroup.  Inputs of at
// most N/2 coefficients (the zero-padded operands of a product) skip the
// subtraction of the first stage, whose right half is zero.
#include "lbx_internal.h"

namespace lbx {
namespace {

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

template <int L, int HL, int RP>
__device__ __forceinline__ void scatter(const uint32_t (&x)[Cfg<L>::R], uint32_t *lds, int tp)
{
    constexpr int NG = Cfg<L>::R >> RP;
#pragma unroll
    for (int u = 0; u < NG; ++u)
#pragma unroll
        for (int q = 0; q < (1 << RP); ++q) lds[pad(elem<L, HL, RP>(tp, u, q))] = x[u * (1 << RP) + q];
}

template <int L, int HL, int RP>
__device__ __forceinline__ void gather(uint32_t (&x)[Cfg<L>::R], const uint32_t *lds, int tp)
{
    constexpr int NG = Cfg<L>::R >> RP;
#pragma unroll
    for (int u = 0; u < NG; ++u)
#pragma unroll
        for (int q = 0; q < (1 << RP); ++q) x[u * (1 << RP) + q] = lds[pad(elem<L, HL, RP>(tp, u, q))];
}

// DIF stages S0..S0+RP-1 (span N>>(s+1) at stage s) on the registers.
template <int L, int S0, int RP, bool HALFZERO>
__device__ __forceinline__ void dif_pass(uint32_t (&x)[Cfg<L>::R], int tp, const uint2 *__restrict__ tw, uint32_t p)
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
            const uint2 *tws = tw + (N - (N >> s));
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
                        B = shoup_mul(A, t.x, t.y, p);
                    } else {
                        const uint32_t a = A, b = B;
                        A = add_mod(a, b, p);
                        B = shoup_mul(a + (p - b), t.x, t.y, p);
                    }
                }
            }
        }
    }
}

// DIT stages D0..D0+RP-1 (span 2^d at stage d).
template <int L, int D0, int RP>
__device__ __forceinline__ void dit_pass(uint32_t (&x)[Cfg<L>::R], int tp, const uint2 *__restrict__ tw, uint32_t p)
{
    constexpr int N = Cfg<L>::N, TP = Cfg<L>::TP, NG = Cfg<L>::R >> RP;
#pragma unroll
    for (int u = 0; u < NG; ++u) {
        const int off = (tp + u * TP) & ((1 << D0) - 1);
#pragma unroll
        for (int st = 0; st < RP; ++st) {
            const int d = D0 + st;
            const int half = 1 << st;
            const uint2 *tws = tw + (N - (2 << d));
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

// Remaining DIF passes starting at stage S0 (full passes first, then the
// remainder).  On entry x holds the mapping of pass S0.
template <int L, int S0, bool HALFZERO>
__device__ __forceinline__ void dif_passes(uint32_t (&x)[Cfg<L>::R], uint32_t *lds, int tp, const uint2 *tw,
                                           uint32_t p)
{
    constexpr int RLOG = Cfg<L>::RLOG;
    constexpr int RP = (L - S0) < RLOG ? (L - S0) : RLOG;
    dif_pass<L, S0, RP, HALFZERO>(x, tp, tw, p);
    if constexpr (S0 + RP < L) {
        constexpr int S1 = S0 + RP;
        constexpr int NRP = (L - S1) < RLOG ? (L - S1) : RLOG;
        if constexpr (S0 > 0) __syncthreads();  // previous gather finished everywhere
        scatter<L, L - S0 - RP, RP>(x, lds, tp);
        __syncthreads();
        gather<L, L - S1 - NRP, NRP>(x, lds, tp);
        dif_passes<L, S1, false>(x, lds, tp, tw, p);
    }
}

// Remaining DIT passes starting at span 2^D0 (remainder first, then full).
template <int L, int D0>
__device__ __forceinline__ void dit_passes(uint32_t (&x)[Cfg<L>::R], uint32_t *lds, int tp, const uint2 *tw,
                                           uint32_t p)
{
    constexpr int RLOG = Cfg<L>::RLOG;
    constexpr int RP = (D0 == 0 && (L % RLOG) != 0) ? (L % RLOG) : RLOG;
    dit_pass<L, D0, RP>(x, tp, tw, p);
    if constexpr (D0 + RP < L) {
        constexpr int D1 = D0 + RP;
        if constexpr (D0 > 0) __syncthreads();
        scatter<L, D0, RP>(x, lds, tp);
        __syncthreads();
        gather<L, D1, RLOG>(x, lds, tp);
        dit_passes<L, D1>(x, lds, tp, tw, p);
    }
}

template <int L>
constexpr int dif_last_rp()
{
    return (L % Cfg<L>::RLOG) ? (L % Cfg<L>::RLOG) : Cfg<L>::RLOG;
}

template <int L>
constexpr int dit_first_rp()
{
    return (L % Cfg<L>::RLOG) ? (L % Cfg<L>::RLOG) : Cfg<L>::RLOG;
}

template <int L, bool HALFZERO>
__global__ void __launch_bounds__(Cfg<L>::BLOCK) ntt_dif_kernel(DevView in, size_t entries, const uint2 *__restrict__ tw,
                                                               uint32_t p, uint32_t *__restrict__ out)
{
    using C = Cfg<L>;
    extern __shared__ uint32_t lds_all[];
    const int pib = threadIdx.x / C::TP;
    const int tp = threadIdx.x % C::TP;
    const size_t e = (size_t)blockIdx.x * C::G + pib;
    const bool active = e < entries;
    uint32_t *lds = lds_all + pib * C::STRIDE;

    uint32_t x[C::R];
    // first pass: the thread owns words tp + q*TP
#pragma unroll
    for (int q = 0; q < C::R; ++q) {
        const size_t t = (size_t)tp + (size_t)q * C::TP;
        x[q] = (active && t < in.size) ? in.base[view_index(in, e, t)] : 0u;
    }
    dif_passes<L, 0, HALFZERO>(x, lds, tp, tw, p);
    // last pass: groups of 2^RP consecutive words
    constexpr int RP = dif_last_rp<L>();
    constexpr int NG = C::R >> RP;
    if (active) {
        uint32_t *o = out + e * C::N;
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
    }
}

template <int L, bool SCALE>
__global__ void __launch_bounds__(Cfg<L>::BLOCK) ntt_dit_kernel(const uint32_t *__restrict__ in, size_t entries,
                                                               const uint2 *__restrict__ tw, uint32_t p, DevView out,
                                                               size_t count, uint32_t scale, uint32_t scale_p)
{
    using C = Cfg<L>;
    extern __shared__ uint32_t lds_all[];
    const int pib = threadIdx.x / C::TP;
    const int tp = threadIdx.x % C::TP;
    const size_t e = (size_t)blockIdx.x * C::G + pib;
    const bool active = e < entries;
    uint32_t *lds = lds_all + pib * C::STRIDE;

    uint32_t x[C::R];
    constexpr int RP0 = dit_first_rp<L>();
    constexpr int NG0 = C::R >> RP0;
    {
        const uint32_t *src = in + (active ? e : 0) * C::N;
#pragma unroll
        for (int u = 0; u < NG0; ++u) {
            const int g = tp + u * C::TP;
            if constexpr (RP0 >= 2) {
#pragma unroll
                for (int q = 0; q < (1 << RP0); q += 4) {
                    const uint4 v = *reinterpret_cast<const uint4 *>(src + (g << RP0) + q);
                    x[u * (1 << RP0) + q] = v.x;
                    x[u * (1 << RP0) + q + 1] = v.y;
                    x[u * (1 << RP0) + q + 2] = v.z;
                    x[u * (1 << RP0) + q + 3] = v.w;
                }
            } else {
#pragma unroll
                for (int q = 0; q < (1 << RP0); ++q) x[u * (1 << RP0) + q] = src[(g << RP0) + q];
            }
        }
    }
    dit_passes<L, 0>(x, lds, tp, tw, p);
    // last pass: the thread owns words tp + q*TP
    if (active) {
#pragma unroll
        for (int q = 0; q < C::R; ++q) {
            const size_t t = (size_t)tp + (size_t)q * C::TP;
            if (t < count) {
                uint32_t v = x[q];
                if (SCALE) v = shoup_mul(v, scale, scale_p, p);
                out.base[view_index(out, e, t)] = v;
            }
        }
    }
}

template <int L>
hipError_t fwd_launch(const DevView &in, size_t entries, const uint2 *tw, uint32_t p, uint32_t *out, hipStream_t s)
{
    using C = Cfg<L>;
    const unsigned grid = (unsigned)((entries + C::G - 1) / C::G);
    if (in.size * 2 <= (size_t)C::N) {
        auto k = ntt_dif_kernel<L, true>;
        if (C::LDS_BYTES > 65536) {
            hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS_BYTES);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k, dim3(grid), dim3(C::BLOCK), C::LDS_BYTES, s, in, entries, tw, p, out);
    } else {
        auto k = ntt_dif_kernel<L, false>;
        if (C::LDS_BYTES > 65536) {
            hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS_BYTES);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k, dim3(grid), dim3(C::BLOCK), C::LDS_BYTES, s, in, entries, tw, p, out);
    }
    return hipGetLastError();
}

template <int L>
hipError_t inv_launch(const uint32_t *in, size_t entries, const uint2 *tw, uint32_t p, const DevView &out,
                      size_t count, uint32_t scale, hipStream_t s)
{
    using C = Cfg<L>;
    const unsigned grid = (unsigned)((entries + C::G - 1) / C::G);
    const uint32_t sp = shoup_pre(scale, p);
    if (scale != 1) {
        auto k = ntt_dit_kernel<L, true>;
        if (C::LDS_BYTES > 65536) {
            hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS_BYTES);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k, dim3(grid), dim3(C::BLOCK), C::LDS_BYTES, s, in, entries, tw, p, out, count, scale, sp);
    } else {
        auto k = ntt_dit_kernel<L, false>;
        if (C::LDS_BYTES > 65536) {
            hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS_BYTES);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k, dim3(grid), dim3(C::BLOCK), C::LDS_BYTES, s, in, entries, tw, p, out, count, scale, sp);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_ntt_forward(const DevView &in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                              uint32_t *out, hipStream_t s)
{
    if (entries == 0) return hipSuccess;
    switch (L) {
#define LBX_CASE(l) \
    case l: return fwd_launch<l>(in, entries, tw, p, out, s);
        LBX_CASE(1) LBX_CASE(2) LBX_CASE(3) LBX_CASE(4) LBX_CASE(5) LBX_CASE(6) LBX_CASE(7) LBX_CASE(8)
        LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ntt_inverse(const uint32_t *in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                              const DevView &out, size_t count, uint32_t scale, hipStream_t s)
{
    if (entries == 0) return hipSuccess;
    switch (L) {
#define LBX_CASE(l) \
    case l: return inv_launch<l>(in, entries, tw, p, out, count, scale, s);
        LBX_CASE(1) LBX_CASE(2) LBX_CASE(3) LBX_CASE(4) LBX_CASE(5) LBX_CASE(6) LBX_CASE(7) LBX_CASE(8)
        LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

}  // namespace lbx
