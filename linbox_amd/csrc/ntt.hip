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

// DIF passes over stages [S0, SEND): full passes first, then the remainder.
// On entry x holds the mapping of pass S0.
template <int L, int S0, int SEND, bool HALFZERO>
__device__ __forceinline__ void dif_passes(uint32_t (&x)[Cfg<L>::R], uint32_t *lds, int tp, const uint2 *tw,
                                           uint32_t p)
{
    constexpr int RLOG = Cfg<L>::RLOG;
    constexpr int RP = (SEND - S0) < RLOG ? (SEND - S0) : RLOG;
    dif_pass<L, S0, RP, HALFZERO>(x, tp, tw, p);
    if constexpr (S0 + RP < SEND) {
        constexpr int S1 = S0 + RP;
        constexpr int NRP = (SEND - S1) < RLOG ? (SEND - S1) : RLOG;
        if constexpr (S0 > 0) __syncthreads();  // previous gather finished everywhere
        scatter<L, L - S0 - RP, RP>(x, lds, tp);
        __syncthreads();
        gather<L, L - S1 - NRP, NRP>(x, lds, tp);
        dif_passes<L, S1, SEND, false>(x, lds, tp, tw, p);
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
    // first pass mapping (words tp + q*TP when the pass is a full radix-R pass)
    constexpr int RP0 = SEND < C::RLOG ? SEND : C::RLOG;
#pragma unroll
    for (int u = 0; u < (C::R >> RP0); ++u)
#pragma unroll
        for (int q = 0; q < (1 << RP0); ++q) {
            const size_t t = (size_t)elem<L, L - RP0, RP0>(tp, u, q);
            x[u * (1 << RP0) + q] = (active && t < in.size) ? in.base[view_index(in, e, t)] : 0u;
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
    } else {
#pragma unroll
        for (int u = 0; u < NG; ++u)
#pragma unroll
            for (int q = 0; q < (1 << RP); ++q) o[elem<L, L - SEND, RP>(tp, u, q)] = x[u * (1 << RP) + q];
    }
}

// DIT stages [DSTART, L) of one polynomial per TP threads.  DSTART == 0 reads
// bit-reversed input with 16-byte vector loads; DSTART > 0 continues a
// transform whose first DSTART stages were applied blockwise.
template <int L, int DSTART, bool SCALE>
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
    if (active) {
#pragma unroll
        for (int u = 0; u < (C::R >> RPL); ++u)
#pragma unroll
            for (int q = 0; q < (1 << RPL); ++q) {
                const size_t t = (size_t)elem<L, L - RPL, RPL>(tp, u, q);
                if (t < count) {
                    uint32_t v = x[u * (1 << RPL) + q];
                    if (SCALE) v = shoup_mul(v, scale, scale_p, p);
                    out.base[view_index(out, e, t)] = v;
                }
            }
    }
}

template <typename K>
hipError_t allow_lds(K k, size_t bytes)
{
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <int L, int SEND>
hipError_t fwd_launch(const DevView &in, size_t entries, const uint2 *tw, uint32_t p, uint32_t *out, hipStream_t s)
{
    using C = Cfg<L>;
    const unsigned grid = (unsigned)((entries + C::G - 1) / C::G);
    auto k = (in.size * 2 <= (size_t)C::N) ? ntt_dif_kernel<L, SEND, true> : ntt_dif_kernel<L, SEND, false>;
    hipError_t e = allow_lds(k, C::LDS_BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(grid), dim3(C::BLOCK), C::LDS_BYTES, s, in, entries, tw, p, out);
    return hipGetLastError();
}

template <int L, int DSTART>
hipError_t inv_launch(const uint32_t *in, size_t entries, const uint2 *tw, uint32_t p, const DevView &out,
                      size_t count, uint32_t scale, hipStream_t s)
{
    using C = Cfg<L>;
    const unsigned grid = (unsigned)((entries + C::G - 1) / C::G);
    const uint32_t sp = shoup_pre(scale, p);
    auto k = (scale != 1) ? ntt_dit_kernel<L, DSTART, true> : ntt_dit_kernel<L, DSTART, false>;
    hipError_t e = allow_lds(k, C::LDS_BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(grid), dim3(C::BLOCK), C::LDS_BYTES, s, in, entries, tw, p, out, count, scale, sp);
    return hipGetLastError();
}

// ---- blockwise stages on 64-point blocks of 64 entries ---------------------
// After DIF stages [0, L-6) every polynomial is 2^(L-6) independent blocks of
// 64 consecutive words, and DIT stages [0, 6) act inside such blocks.  These
// kernels run those six stages for 64 entries at once, which lets them write
// (forward) or read (inverse) the point-major layout of the pointwise stage
// with 256-byte runs: the transposition rides along with the butterflies.
constexpr int kBlk = 64;
constexpr int kBlkStride = 65;  // LDS row stride (words) of a [64 entries][64 points] tile

// DIF on the 6 stages of block `blk` held as x[16] per thread: thread owns
// entry q = tid>>2 and local positions sub + 4v (pass 1) then 4(sub+4u)+w (pass 2).
template <int L>
__device__ __forceinline__ void dif_block_stages(uint32_t (*sm)[kBlkStride], int q, int sub, const uint2 *tw, uint32_t p)
{
    constexpr int N = 1 << L;
    uint32_t x[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) x[v] = sm[q][sub + 4 * v];
#pragma unroll
    for (int ls = 0; ls < 4; ++ls) {  // local spans 32, 16, 8, 4
        const int s = L - 6 + ls;
        const int half = 8 >> ls;
        const uint2 *tws = tw + (N - (N >> s));
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            if (v & half) continue;
            const uint2 t = tws[sub + 4 * (v & (half - 1))];
            const uint32_t a = x[v], b = x[v + half];
            x[v] = add_mod(a, b, p);
            x[v + half] = shoup_mul(a + (p - b), t.x, t.y, p);
        }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) sm[q][sub + 4 * v] = x[v];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int w = 0; w < 4; ++w) x[4 * u + w] = sm[q][4 * (sub + 4 * u) + w];
    const uint2 t1 = tw[N - 4 + 1];  // span 2: twiddles {1, w^(N/4)}
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        uint32_t *y = x + 4 * u;
        // span 2
        uint32_t a = y[0], b = y[2];
        y[0] = add_mod(a, b, p);
        y[2] = sub_mod(a, b, p);
        a = y[1], b = y[3];
        y[1] = add_mod(a, b, p);
        y[3] = shoup_mul(a + (p - b), t1.x, t1.y, p);
        // span 1
#pragma unroll
        for (int w = 0; w < 4; w += 2) {
            a = y[w], b = y[w + 1];
            y[w] = add_mod(a, b, p);
            y[w + 1] = sub_mod(a, b, p);
        }
    }
    // each thread rewrites only the positions it read in this pass
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int w = 0; w < 4; ++w) sm[q][4 * (sub + 4 * u) + w] = to_limbs(x[4 * u + w], p);
}

// Last six DIF stages of block blockIdx.x for the 64 entries of limb row
// blockIdx.z, inner chunk blockIdx.y; writes that 256-byte run of every
// point's limb block.
template <int L>
__global__ void __launch_bounds__(256) dif_tail_limbs_kernel(const uint32_t *__restrict__ X, uint8_t *__restrict__ Y,
                                                             int rows, int kdim, int row_major,
                                                             const uint2 *__restrict__ tw, uint32_t p,
                                                             size_t block_bytes, int kchunks)
{
    constexpr int N = 1 << L;
    __shared__ uint32_t sm[kBlk][kBlkStride];
    const int tid = threadIdx.x;
    const int b = blockIdx.x, c = blockIdx.y, r = blockIdx.z;
    {   // 16 lanes per entry row read its 256-byte block run
        const int ch = tid & 15;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = (tid >> 4) + 16 * u;
            const int kk = c * kBlk + q;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (r < rows && kk < kdim) {
                const size_t e = row_major ? (size_t)r * kdim + kk : (size_t)kk * rows + r;
                v = *reinterpret_cast<const uint4 *>(X + e * N + (size_t)b * kBlk + ch * 4);
            }
            sm[q][ch * 4] = v.x;
            sm[q][ch * 4 + 1] = v.y;
            sm[q][ch * 4 + 2] = v.z;
            sm[q][ch * 4 + 3] = v.w;
        }
    }
    __syncthreads();
    dif_block_stages<L>(sm, tid >> 2, tid & 3, tw, p);
    __syncthreads();
    // point t = b*64 + lam: word w of the run = plane w>>4, entries 4(w&15)..+3
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {
        const int lam = it * 4 + (tid >> 6), w = tid & 63;
        const int sft = 8 * (w >> 4), q4 = (w & 15) * 4;
        const uint32_t word = ((sm[q4][lam] >> sft) & 0xffu) | (((sm[q4 + 1][lam] >> sft) & 0xffu) << 8) |
                              (((sm[q4 + 2][lam] >> sft) & 0xffu) << 16) | (((sm[q4 + 3][lam] >> sft) & 0xffu) << 24);
        const size_t t = (size_t)b * kBlk + lam;
        *reinterpret_cast<uint32_t *>(Y + t * block_bytes + ((size_t)r * kchunks + c) * 256 + (w >> 4) * 64 +
                                      (w & 15) * 4) = word;
    }
}

// First six DIT stages of block blockIdx.x for entries 64*blockIdx.y .. +63,
// read from the point-major result C[t][e]; writes polfirst scratch.
template <int L>
__global__ void __launch_bounds__(256) dit_head_kernel(const uint32_t *__restrict__ Cm, size_t entries,
                                                       uint32_t *__restrict__ Y, const uint2 *__restrict__ tw,
                                                       uint32_t p)
{
    constexpr int N = 1 << L;
    __shared__ uint32_t sm[kBlk][kBlkStride];
    const int tid = threadIdx.x;
    const int b = blockIdx.x;
    const size_t e0 = (size_t)blockIdx.y * kBlk;
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {  // 64 lanes read 64 consecutive entries of one point
        const int q = tid & 63, lam = (tid >> 6) + 4 * it;
        const size_t t = (size_t)b * kBlk + lam;
        sm[q][lam] = (e0 + q < entries) ? Cm[t * entries + e0 + q] : 0u;
    }
    __syncthreads();
    const int q = tid >> 2, sub = tid & 3;
    uint32_t x[16];
    // pass 1: spans 1 and 2 on groups of four consecutive positions
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int w = 0; w < 4; ++w) x[4 * u + w] = sm[q][4 * (sub + 4 * u) + w];
    const uint2 t1 = tw[N - 4 + 1];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        uint32_t *y = x + 4 * u;
#pragma unroll
        for (int w = 0; w < 4; w += 2) {
            const uint32_t a = y[w], bb = y[w + 1];
            y[w] = add_mod(a, bb, p);
            y[w + 1] = sub_mod(a, bb, p);
        }
        uint32_t a = y[0], tt = y[2];
        y[0] = add_mod(a, tt, p);
        y[2] = sub_mod(a, tt, p);
        a = y[1];
        tt = shoup_mul(y[3], t1.x, t1.y, p);
        y[1] = add_mod(a, tt, p);
        y[3] = sub_mod(a, tt, p);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int w = 0; w < 4; ++w) sm[q][4 * (sub + 4 * u) + w] = x[4 * u + w];
    __syncthreads();
    // pass 2: spans 4..32 on positions sub + 4v
#pragma unroll
    for (int v = 0; v < 16; ++v) x[v] = sm[q][sub + 4 * v];
#pragma unroll
    for (int d = 2; d < 6; ++d) {
        const int half = 1 << (d - 2);
        const uint2 *tws = tw + (N - (2 << d));
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            if (v & half) continue;
            const uint2 w = tws[sub + 4 * (v & (half - 1))];
            const uint32_t a = x[v];
            const uint32_t tt = shoup_mul(x[v + half], w.x, w.y, p);
            x[v] = add_mod(a, tt, p);
            x[v + half] = sub_mod(a, tt, p);
        }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) sm[q][sub + 4 * v] = x[v];
    __syncthreads();
    {   // 16 lanes per entry write its 256-byte run
        const int ch = tid & 15;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int qq = (tid >> 4) + 16 * u;
            if (e0 + qq < entries)
                *reinterpret_cast<uint4 *>(Y + (e0 + qq) * N + (size_t)b * kBlk + ch * 4) =
                    make_uint4(sm[qq][ch * 4], sm[qq][ch * 4 + 1], sm[qq][ch * 4 + 2], sm[qq][ch * 4 + 3]);
        }
    }
}

}  // namespace

hipError_t launch_ntt_forward(const DevView &in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                              uint32_t *out, hipStream_t s)
{
    if (entries == 0) return hipSuccess;
    switch (L) {
#define LBX_CASE(l) \
    case l: return fwd_launch<l, l>(in, entries, tw, p, out, s);
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
    case l: return inv_launch<l, 0>(in, entries, tw, p, out, count, scale, s);
        LBX_CASE(1) LBX_CASE(2) LBX_CASE(3) LBX_CASE(4) LBX_CASE(5) LBX_CASE(6) LBX_CASE(7) LBX_CASE(8)
        LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ntt_forward_head(const DevView &in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                                   uint32_t *out, hipStream_t s)
{
    if (entries == 0) return hipSuccess;
    switch (L) {
#define LBX_CASE(l) \
    case l: return fwd_launch<l, l - 6>(in, entries, tw, p, out, s);
        LBX_CASE(8) LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ntt_inverse_tail(const uint32_t *in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                                   const DevView &out, size_t count, uint32_t scale, hipStream_t s)
{
    if (entries == 0) return hipSuccess;
    switch (L) {
#define LBX_CASE(l) \
    case l: return inv_launch<l, 6>(in, entries, tw, p, out, count, scale, s);
        LBX_CASE(8) LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_dif_tail_limbs(const uint32_t *X, uint8_t *Y, size_t rows, size_t kdim, bool row_major, unsigned L,
                                 const uint2 *tw, uint32_t p, hipStream_t s)
{
    const size_t rpad = (rows + 31) & ~(size_t)31, kchunks = (kdim + 63) / 64;
    if (rpad > 65535 || kchunks > 65535 || L < 8 || L > 15) return hipErrorInvalidValue;
    const size_t block_bytes = rpad * kchunks * 256;
    dim3 grid((unsigned)((size_t)1 << (L - 6)), (unsigned)kchunks, (unsigned)rpad);
    switch (L) {
#define LBX_CASE(l)                                                                                                  \
    case l:                                                                                                          \
        hipLaunchKernelGGL(dif_tail_limbs_kernel<l>, grid, dim3(256), 0, s, X, Y, (int)rows, (int)kdim,             \
                           row_major ? 1 : 0, tw, p, block_bytes, (int)kchunks);                                    \
        break;
        LBX_CASE(8) LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    }
    return hipGetLastError();
}

hipError_t launch_dit_head(const uint32_t *Cm, size_t entries, unsigned L, const uint2 *tw, uint32_t p, uint32_t *Y,
                           hipStream_t s)
{
    if (entries == 0) return hipSuccess;
    const size_t groups = (entries + 63) / 64;
    if (groups > 65535 || L < 8 || L > 15) return hipErrorInvalidValue;
    dim3 grid((unsigned)((size_t)1 << (L - 6)), (unsigned)groups);
    switch (L) {
#define LBX_CASE(l)                                                                                    \
    case l: hipLaunchKernelGGL(dit_head_kernel<l>, grid, dim3(256), 0, s, Cm, entries, Y, tw, p); break;
        LBX_CASE(8) LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    }
    return hipGetLastError();
}

}  // namespace lbx
