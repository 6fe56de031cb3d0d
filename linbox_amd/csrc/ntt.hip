// Number-theoretic transform kernels and their launchers: whole transforms
// for the plain path and the NTT entry points, and the split head / tail
// transforms around the point-major pointwise stage (see ntt_device.h).
#include <algorithm>
#include <vector>

#include "ntt_device.h"

namespace lbx {
namespace {
using namespace ntt;

template <int L, int SEND, bool HALFZERO>
__global__ void __launch_bounds__(Cfg<L>::BLOCK) ntt_dif_kernel(DevView in, size_t entries, EntryMap map,
                                                               const uint2 *__restrict__ tw, uint32_t p,
                                                               uint32_t *__restrict__ out)
{
    extern __shared__ uint32_t lds_all[];
    dif_body<L, SEND, HALFZERO>(in, entries, map, tw, p, out, (size_t)blockIdx.x * Cfg<L>::G, threadIdx.x, lds_all);
}

template <int L, int SEND, bool HALFZERO>
__global__ void __launch_bounds__(Cfg<L>::BLOCK) dif_head_persistent_kernel(DevView in, size_t entries, EntryMap map,
                                                                           const uint2 *__restrict__ tw, uint32_t p,
                                                                           uint32_t *__restrict__ out)
{
    extern __shared__ uint32_t lds_all[];
    dif_head_persistent<L, SEND, HALFZERO>(in, entries, map, tw, p, out, threadIdx.x, lds_all);
}

template <int L, int DSTART>
__global__ void __launch_bounds__(Cfg<L>::BLOCK) dit_tail_persistent_kernel(const uint32_t *__restrict__ in,
                                                                           size_t entries,
                                                                           const uint2 *__restrict__ tw, uint32_t p,
                                                                           DevView out, size_t out_first, size_t count)
{
    extern __shared__ uint32_t lds_all[];
    dit_tail_persistent<L, DSTART>(in, entries, tw, p, out, out_first, count, threadIdx.x, lds_all);
}

template <int L, int DSTART, bool SCALE>
__global__ void __launch_bounds__(Cfg<L>::BLOCK) ntt_dit_kernel(const uint32_t *__restrict__ in, size_t entries,
                                                               const uint2 *__restrict__ tw, uint32_t p, DevView out,
                                                               size_t out_first, size_t count, uint32_t scale,
                                                               uint32_t scale_p)
{
    extern __shared__ uint32_t lds_all[];
    dit_body<L, DSTART, SCALE>(in, entries, tw, p, out, out_first, count, scale, scale_p,
                               (size_t)blockIdx.x * Cfg<L>::G, threadIdx.x, lds_all);
}

template <typename K>
hipError_t allow_lds(K k, size_t bytes)
{
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <int L, int SEND>
hipError_t fwd_launch(const DevView &in, size_t entries, const EntryMap &map, const uint2 *tw, uint32_t p,
                      uint32_t *out, hipStream_t s)
{
    using C = Cfg<L>;
    const unsigned grid = (unsigned)((entries + C::G - 1) / C::G);
    auto k = (in.size * 2 <= (size_t)C::N) ? ntt_dif_kernel<L, SEND, true> : ntt_dif_kernel<L, SEND, false>;
    hipError_t e = allow_lds(k, C::LDS_BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(grid), dim3(C::BLOCK), C::LDS_BYTES, s, in, entries, map, tw, p, out);
    return hipGetLastError();
}


// Forward head: the persistent form (grid of two groups per CU) for polfirst
// inputs of 2^12..2^14 points, the one-group-per-polynomial form otherwise (at
// 2^15 the register twiddles of a 32-word pass would spill).
template <int L, int SEND>
hipError_t head_launch(const DevView &in, size_t entries, const EntryMap &map, const uint2 *tw, uint32_t p,
                       uint32_t *out, hipStream_t s)
{
    if constexpr (L >= 12 && L <= 14) {
        if (in.layout == 0) {
            using H = HeadShape<L, SEND>;
            const bool hz = in.size * 2 <= (size_t)H::N;
            auto k = hz ? dif_head_persistent_kernel<L, SEND, true> : dif_head_persistent_kernel<L, SEND, false>;
            hipError_t e = allow_lds(k, H::LDS_BYTES);
            if (e != hipSuccess) return e;
            const size_t grid = std::min<size_t>(entries, (size_t)device_cus() * 2);
            hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(Cfg<L>::BLOCK), H::LDS_BYTES, s, in, entries, map, tw, p,
                               out);
            return hipGetLastError();
        }
    }
    return fwd_launch<L, SEND>(in, entries, map, tw, p, out, s);
}

template <int L, int DSTART>
hipError_t inv_launch(const uint32_t *in, size_t entries, const uint2 *tw, uint32_t p, const DevView &out,
                      size_t out_first, size_t count, uint32_t scale, hipStream_t s);

// Inverse tail: the persistent form for 2^12..2^14 points, polfirst output and
// no final scale; the one-group-per-polynomial form otherwise.
template <int L, int DSTART>
hipError_t tail_inv_launch(const uint32_t *in, size_t entries, const uint2 *tw, uint32_t p, const DevView &out,
                           size_t out_first, size_t count, uint32_t scale, hipStream_t s)
{
    if constexpr (L >= 12 && L <= 14) {
        if (scale == 1 && out.layout == 0) {
            using T = TailShape<L, DSTART>;
            auto k = dit_tail_persistent_kernel<L, DSTART>;
            hipError_t e = allow_lds(k, T::LDS_BYTES);
            if (e != hipSuccess) return e;
            const size_t grid = std::min<size_t>(entries, (size_t)device_cus() * 2);
            hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(Cfg<L>::BLOCK), T::LDS_BYTES, s, in, entries, tw, p, out,
                               out_first, count);
            return hipGetLastError();
        }
    }
    return inv_launch<L, DSTART>(in, entries, tw, p, out, out_first, count, scale, s);
}

template <int L, int DSTART>
hipError_t inv_launch(const uint32_t *in, size_t entries, const uint2 *tw, uint32_t p, const DevView &out,
                      size_t out_first, size_t count, uint32_t scale, hipStream_t s)
{
    using C = Cfg<L>;
    const unsigned grid = (unsigned)((entries + C::G - 1) / C::G);
    const uint32_t sp = shoup_pre(scale, p);
    auto k = (scale != 1) ? ntt_dit_kernel<L, DSTART, true> : ntt_dit_kernel<L, DSTART, false>;
    hipError_t e = allow_lds(k, C::LDS_BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(grid), dim3(C::BLOCK), C::LDS_BYTES, s, in, entries, tw, p, out, out_first, count,
                       scale, sp);
    return hipGetLastError();
}


// ---- blockwise stages on 2^BL-point blocks of 64 entries -------------------
// After DIF stages [0, L-BL) every polynomial is 2^(L-BL) independent blocks
// of 2^BL consecutive words, and DIT stages [0, BL) act inside such blocks.
// These kernels run those BL stages for 64 entries at once, which lets them
// write (forward) or read (inverse) the point-major layout of the pointwise
// stage with 256-byte runs: the transposition rides along with the
// butterflies.  TPE = 2^(BL-4) threads own one entry row, 16 words each: in
// one pass the words sub + TPE*v (the four stages of spans >= TPE), in the
// other TPE-word groups TPE*(sub + TPE*u) + w (the stages of spans < TPE).
template <int BL>
struct Blk {
    static constexpr int P = 1 << BL;         // points per block
    static constexpr int TPE = 1 << (BL - 4);  // threads per entry row
    static constexpr int T = 64 * TPE;         // threads per workgroup
    static constexpr int STRIDE = P + 1;       // LDS words per entry row
    static constexpr size_t LDS_BYTES = (size_t)64 * STRIDE * 4;
};

// DIF butterfly on canonical words with Montgomery twiddle pairs; tw index 0
// of every span table is w^0 = 1.
__device__ __forceinline__ void dif_bfly(uint32_t &A, uint32_t &B, const uint2 t, bool trivial, uint32_t p)
{
    const uint32_t a = A, b = B;
    A = add_mod(a, b, p);
    B = trivial ? sub_mod(a, b, p) : mont_mul((int32_t)(a - b), t.x, t.y, p);
}

__device__ __forceinline__ void dit_bfly(uint32_t &A, uint32_t &B, const uint2 t, bool trivial, uint32_t p)
{
    const uint32_t a = A;
    const uint32_t u = trivial ? B : shoup_mul(B, t.x, t.y, p);
    A = add_mod(a, u, p);
    B = sub_mod(a, u, p);
}

// The last BL DIF stages of one block row (spans 2^(BL-1) .. 1).  On return x
// holds the group mapping; the row in LDS still holds the stride mapping.
template <int L, int BL>
__device__ __forceinline__ void dif_block_stages(uint32_t *row, int sub, const uint2 *__restrict__ tw, uint32_t p,
                                                 uint32_t (&x)[16])
{
    constexpr int N = 1 << L, TPE = Blk<BL>::TPE;
#pragma unroll
    for (int v = 0; v < 16; ++v) x[v] = row[sub + TPE * v];
#pragma unroll
    for (int ls = 0; ls < 4; ++ls) {  // spans 8*TPE .. TPE
        const int s = L - BL + ls;
        const int half = 8 >> ls;
        const uint2 *tws = tw + (N - (N >> s));
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            if (v & half) continue;
            dif_bfly(x[v], x[v + half], tws[sub + TPE * (v & (half - 1))], false, p);
        }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) row[sub + TPE * v] = x[v];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16 / TPE; ++u)
#pragma unroll
        for (int w = 0; w < TPE; ++w) x[u * TPE + w] = row[TPE * (sub + TPE * u) + w];
#pragma unroll
    for (int h = TPE / 2; h >= 1; h >>= 1) {  // spans TPE/2 .. 1, twiddles uniform across threads
        const uint2 *tws = tw + (N - 2 * h);
#pragma unroll
        for (int u = 0; u < 16 / TPE; ++u)
#pragma unroll
            for (int w = 0; w < TPE; ++w) {
                if (w & h) continue;
                const int ti = w & (h - 1);
                dif_bfly(x[u * TPE + w], x[u * TPE + w + h], ti ? tws[ti] : make_uint2(1, 0), ti == 0, p);
            }
    }
}

// Last BL DIF stages of block blockIdx.z for the 64 entries of limb row
// blockIdx.y, inner chunk blockIdx.x; writes that 256-byte run of every
// point's block: four planes of 64 signed 8-bit digits (LIMBS), or the 64
// centred values as int32 words (the point-major operands of the plain form).
template <int L, int BL, bool LIMBS>
__global__ void __launch_bounds__(Blk<BL>::T) dif_tail_limbs_kernel(const uint32_t *__restrict__ X,
                                                                    uint8_t *__restrict__ Y, int rows_valid, int row0,
                                                                    int kdim, int row_major,
                                                                    const uint2 *__restrict__ tw, uint32_t p,
                                                                    size_t block_bytes, int kchunks)
{
    using B = Blk<BL>;
    constexpr int P = B::P, T = B::T, CPR = P / 4;  // 16-byte chunks per block row
    extern __shared__ uint32_t sm[];
    const int tid = threadIdx.x;
    // the chunk and row vary fastest: groups that run together fill adjacent
    // 256-byte runs of the same points
    const int c = blockIdx.x, r = blockIdx.y, b = blockIdx.z;  // r: row within the chunk
#pragma unroll
    for (int u = 0; u < 4; ++u) {  // CPR lanes per entry row read its 4P-byte block run
        const int id = tid + T * u;
        const int q = id / CPR, ch = id % CPR;
        const int kk = c * 64 + q;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (r < rows_valid && kk < kdim) {
            const size_t e = row_major ? (size_t)r * kdim + kk : (size_t)kk * rows_valid + r;
            // block-major head output: [block][entry][P words]
            v = *reinterpret_cast<const uint4 *>(X + ((size_t)b * rows_valid * kdim + e) * P + ch * 4);
        }
        uint32_t *d = sm + q * B::STRIDE + ch * 4;
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        d[3] = v.w;
    }
    __syncthreads();
    const int q = tid / B::TPE, sub = tid % B::TPE;
    uint32_t *row = sm + q * B::STRIDE;
    uint32_t x[16];
    dif_block_stages<L, BL>(row, sub, tw, p, x);
    // each thread rewrites only the positions it read in the group mapping
#pragma unroll
    for (int u = 0; u < 16 / B::TPE; ++u)
#pragma unroll
        for (int w = 0; w < B::TPE; ++w)
            row[B::TPE * (sub + B::TPE * u) + w] =
                LIMBS ? to_limbs(x[u * B::TPE + w], p) : (uint32_t)centre(x[u * B::TPE + w], p);
    __syncthreads();
    // point t = b*P + lam: with LIMBS, 16-byte chunk ch of the run = plane
    // ch>>2, entries 16(ch&3)..+15 (byte j of word i holds entry 16(ch&3) + 4i + j);
    // without, chunk ch = the words of entries 4ch..4ch+3
#pragma unroll 2
    for (int it = 0; it < 4; ++it) {
        const int id = tid + T * it;
        const int lam = id >> 4, ch = id & 15;
        uint32_t w4[4];
        if constexpr (LIMBS) {
            const int sft = 8 * (ch >> 2);
            const uint32_t *col = sm + (16 * (ch & 3)) * B::STRIDE + lam;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                w4[i] = ((col[(4 * i) * B::STRIDE] >> sft) & 0xffu) |
                        (((col[(4 * i + 1) * B::STRIDE] >> sft) & 0xffu) << 8) |
                        (((col[(4 * i + 2) * B::STRIDE] >> sft) & 0xffu) << 16) |
                        (((col[(4 * i + 3) * B::STRIDE] >> sft) & 0xffu) << 24);
        } else {
            const uint32_t *col = sm + (4 * ch) * B::STRIDE + lam;
#pragma unroll
            for (int i = 0; i < 4; ++i) w4[i] = col[i * B::STRIDE];
        }
        const size_t t = (size_t)b * P + lam;
        store_stream4(reinterpret_cast<uint32_t *>(Y + t * block_bytes + ((size_t)(row0 + r) * kchunks + c) * 256 +
                                                   ch * 16),
                      w4[0], w4[1], w4[2], w4[3]);
    }
}

// First BL DIT stages of block blockIdx.y for entries 64*blockIdx.x .. +63,
// read from the point-major result C[t][e]; writes polfirst scratch.  The
// entry group varies fastest, so groups that run together read adjacent
// 256-byte runs of the same points.
template <int L, int BL>
__global__ void __launch_bounds__(Blk<BL>::T) dit_head_kernel(const uint32_t *__restrict__ Cm, size_t entries,
                                                              size_t e_first, size_t e_count,
                                                              uint32_t *__restrict__ Y, const uint2 *__restrict__ tw,
                                                              uint32_t p)
{
    using B = Blk<BL>;
    constexpr int N = 1 << L, P = B::P, T = B::T, TPE = B::TPE, CPR = P / 4;
    extern __shared__ uint32_t sm[];
    const int tid = threadIdx.x;
    const int b = blockIdx.y;
    const size_t e0 = (size_t)blockIdx.x * 64;  // relative to e_first
    if (((entries | e_first) & 3) == 0 && e0 + 64 <= e_count) {
        // 16 lanes read one point's 64 consecutive entries as 16-byte chunks
#pragma unroll 2
        for (int it = 0; it < 4; ++it) {
            const int id = tid + T * it;
            const int lam = id >> 4, q = (id & 15) * 4;
            const size_t t = (size_t)b * P + lam;
            const uint4 v = *reinterpret_cast<const uint4 *>(Cm + t * entries + e_first + e0 + q);
            sm[q * B::STRIDE + lam] = v.x;
            sm[(q + 1) * B::STRIDE + lam] = v.y;
            sm[(q + 2) * B::STRIDE + lam] = v.z;
            sm[(q + 3) * B::STRIDE + lam] = v.w;
        }
    } else {
#pragma unroll 4
        for (int it = 0; it < 16; ++it) {  // 64 lanes read 64 consecutive entries of one point
            const int id = tid + T * it;
            const int q = id & 63, lam = id >> 6;
            const size_t t = (size_t)b * P + lam;
            sm[q * B::STRIDE + lam] = (e0 + q < e_count) ? Cm[t * entries + e_first + e0 + q] : 0u;
        }
    }
    __syncthreads();
    const int q = tid / TPE, sub = tid % TPE;
    uint32_t *row = sm + q * B::STRIDE;
    uint32_t x[16];
    // spans 1 .. TPE/2 on TPE-word groups, twiddles uniform across threads
#pragma unroll
    for (int u = 0; u < 16 / TPE; ++u)
#pragma unroll
        for (int w = 0; w < TPE; ++w) x[u * TPE + w] = row[TPE * (sub + TPE * u) + w];
#pragma unroll
    for (int h = 1; h < TPE; h <<= 1) {
        const uint2 *tws = tw + (N - 2 * h);
#pragma unroll
        for (int u = 0; u < 16 / TPE; ++u)
#pragma unroll
            for (int w = 0; w < TPE; ++w) {
                if (w & h) continue;
                const int ti = w & (h - 1);
                dit_bfly(x[u * TPE + w], x[u * TPE + w + h], ti ? tws[ti] : make_uint2(1, 0), ti == 0, p);
            }
    }
#pragma unroll
    for (int u = 0; u < 16 / TPE; ++u)
#pragma unroll
        for (int w = 0; w < TPE; ++w) row[TPE * (sub + TPE * u) + w] = x[u * TPE + w];
    __syncthreads();
    // spans TPE .. 8*TPE on the words sub + TPE*v
#pragma unroll
    for (int v = 0; v < 16; ++v) x[v] = row[sub + TPE * v];
#pragma unroll
    for (int ds = 0; ds < 4; ++ds) {
        const int h = TPE << ds, half = 1 << ds;
        const uint2 *tws = tw + (N - 2 * h);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            if (v & half) continue;
            dit_bfly(x[v], x[v + half], tws[sub + TPE * (v & (half - 1))], false, p);
        }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) row[sub + TPE * v] = x[v];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {  // CPR lanes per entry write its 4P-byte run
        const int id = tid + T * u;
        const int qq = id / CPR, ch = id % CPR;
        if (e0 + qq < e_count) {
            const uint32_t *src = sm + qq * B::STRIDE + ch * 4;
            *reinterpret_cast<uint4 *>(Y + (e0 + qq) * N + (size_t)b * P + ch * 4) =
                make_uint4(src[0], src[1], src[2], src[3]);
        }
    }
}

// ---- batch transforms ---------------------------------------------------------
// Persistent work-queue kernel for the transforms of one length 2^L across a
// batch: each 512-thread workgroup repeatedly takes an item (entries
// [e0, e0+count) of one operand of one job) from a device counter.  A 2^13
// transform uses all 512 threads; shorter ones run as two independent
// 256-thread groups of Cfg<L>::G polynomials each.
// Waves per SIMD the register budget targets: 8 (four groups per CU) where the
// forward body fits in 64 VGPRs without spilling, else 4.
template <int L, bool INVERSE>
constexpr int batch_waves() { return (INVERSE || L == 12) ? 4 : 8; }

template <int L, bool INVERSE>
__global__ void __launch_bounds__(kBatchNttThreads, (batch_waves<L, INVERSE>()))
    batch_ntt_kernel(const BatchJob *__restrict__ jobs, const NttItem *__restrict__ items, uint32_t nitems,
                     uint32_t *counter)
{
    using C = Cfg<L>;
    constexpr bool whole = C::BLOCK == kBatchNttThreads;
    extern __shared__ uint32_t lds_all[];
    __shared__ uint32_t item_sh;
    const int sub = whole ? 0 : (int)(threadIdx.x >> 8);
    uint32_t *lds = lds_all + sub * C::G * C::STRIDE;
    for (;;) {
        // thread index re-derived every item: otherwise the per-slot LDS and
        // global offsets are hoisted out of the loop and held in registers
        int tid = whole ? (int)threadIdx.x : (int)(threadIdx.x & 255);
        asm volatile("" : "+v"(tid));
        if (threadIdx.x == 0) item_sh = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t it = __builtin_amdgcn_readfirstlane(item_sh);  // uniform: scalar loads of the item
        __syncthreads();
        if (it >= nitems) return;
        const NttItem w = items[it];
        const BatchJob &J = jobs[w.job];
        const size_t e_base = (size_t)w.e0 + (size_t)sub * C::G;
        const size_t e_end = (size_t)w.e0 + w.count;
        if constexpr (!INVERSE) {
            const bool isa = w.op == 0;
            DevView in;
            in.base = global_ptr(const_cast<uint32_t *>(isa ? J.a : J.b));
            in.layout = 0;
            in.rows = isa ? J.m : J.k;
            in.cols = isa ? J.k : J.n;
            in.stor = in.size = isa ? J.sa : J.sb;
            in.shift = 0;
            const EntryMap id{0, ~(size_t)0, 0};
            uint32_t *out = global_ptr(isa ? J.ahat : J.bhat);
            const uint2 *tw = global_ptr(J.twf);
            if (in.size * 2 <= (size_t)C::N)
                dif_body<L, L, true>(in, e_end, id, tw, J.p, out, e_base, tid, lds);
            else
                dif_body<L, L, false>(in, e_end, id, tw, J.p, out, e_base, tid, lds);
        } else {
            DevView out;
            out.base = global_ptr(J.c);
            out.layout = 0;
            out.rows = J.m;
            out.cols = J.n;
            out.stor = out.size = J.sc;
            out.shift = 0;
            dit_body<L, 0, false>(global_ptr(J.chat), e_end, global_ptr(J.twi), J.p, out, 0, J.sc, 1u, 0u, e_base, tid,
                                  lds);
        }
        __syncthreads();  // LDS is reused by the next item
    }
}

// ---- transforms longer than the LDS-resident size (2^16 .. 2^kMaxLog) ------
// After DIF stages [0, s) a transform of N = 2^L words is 2^s independent
// blocks of N/2^s consecutive words, each needing a DIF of that length with
// root w^(2^s), whose pow_w table is the tail of the big table from offset
// N - N/2^s (the subarray of span h sits at N - 2h in both).  DIT stages
// [0, d) likewise act inside blocks of 2^d words.  So a long forward transform
// is a few HBM passes of RC <= 4 "column" stages (each thread holds 2^RC words
// of one block, 2^RC apart in units of the block's 1/2^RC, and applies RC
// radix-2 stages in registers) followed by the LDS-resident DIF of the blocks
// of 2^kLongLds words; the inverse is the LDS-resident DIT of those blocks
// followed by column passes for the remaining spans.  Consecutive threads take
// consecutive words of a block, so every load and store instruction of a wave
// covers 256 contiguous bytes, and so does every twiddle load.
constexpr unsigned kLongLds = 14;

// DIF stages [s0, s0+RC) of `entries` polynomials.  Input: the view `in`
// (first in.size coefficients, zero beyond) when FROM_VIEW, else the
// contiguous out[e*N ...] (in place).  Output: out[e*N ...].
template <int RC, bool FROM_VIEW>
__global__ void __launch_bounds__(256) dif_cols_kernel(DevView in, size_t entries, unsigned L, unsigned s0,
                                                       const uint2 *__restrict__ tw, uint32_t p, uint32_t *out)
{
    constexpr int R = 1 << RC;
    const size_t N = (size_t)1 << L;
    const size_t B0 = N >> s0;    // block length at stage s0
    const size_t J = B0 >> RC;    // distance between a thread's words
    const size_t per_entry = N >> RC;
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= entries * per_entry) return;
    const size_t j = g % J, rest = g / J;
    const size_t e = rest % entries, blk = rest / entries;
    const size_t base = blk * B0 + j;
    uint32_t x[R];
    if constexpr (FROM_VIEW) {
        if (in.layout == 0) {
            const uint32_t *src = in.base + e * in.stor + in.shift;
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const size_t t = base + (size_t)i * J;
                x[i] = t < in.size ? src[t] : 0u;
            }
        } else {
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const size_t t = base + (size_t)i * J;
                x[i] = t < in.size ? in.base[view_index(in, e, t)] : 0u;
            }
        }
    } else {
        const uint32_t *src = out + e * N;
#pragma unroll
        for (int i = 0; i < R; ++i) x[i] = src[base + (size_t)i * J];
    }
#pragma unroll
    for (int t = 0; t < RC; ++t) {
        const int half = 1 << (RC - 1 - t);
        const unsigned st = s0 + t;
        const uint2 *tws = tw + (N - (N >> st));
#pragma unroll
        for (int i = 0; i < R; ++i) {
            if (i & half) continue;
            const uint32_t a = x[i], b = x[i + half];
            if (st == L - 1) {  // span 1: the twiddle is w^0 = 1
                x[i] = add_mod(a, b, p);
                x[i + half] = sub_mod(a, b, p);
            } else {
                const uint2 w = tws[j + (size_t)(i & (half - 1)) * J];
                x[i] = add_mod(a, b, p);
                x[i + half] = mont_mul((int32_t)(a - b), w.x, w.y, p);
            }
        }
    }
    uint32_t *dst = out + e * N;
#pragma unroll
    for (int i = 0; i < R; ++i) dst[base + (size_t)i * J] = x[i];
}

// DIT stages [d0, d0+RC) (spans 2^d0 .. 2^(d0+RC-1)) of `entries` contiguous
// polynomials buf[e*N ...], in place; with LAST the words are written to
// the view `out` instead (coefficients t < count, times scale when SCALE).
template <int RC, bool LAST, bool SCALE>
__global__ void __launch_bounds__(256) dit_cols_kernel(uint32_t *buf, size_t entries, unsigned L, unsigned d0,
                                                       const uint2 *__restrict__ tw, uint32_t p, DevView out,
                                                       size_t count, uint32_t scale, uint32_t scale_p)
{
    constexpr int R = 1 << RC;
    const size_t N = (size_t)1 << L;
    const size_t J = (size_t)1 << d0;  // distance between a thread's words
    const size_t per_entry = N >> RC;
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= entries * per_entry) return;
    const size_t j = g % J, rest = g / J;
    const size_t e = rest % entries, blk = rest / entries;
    const size_t base = (blk << (d0 + RC)) + j;
    uint32_t x[R];
    uint32_t *src = buf + e * N;
#pragma unroll
    for (int i = 0; i < R; ++i) x[i] = src[base + (size_t)i * J];
#pragma unroll
    for (int t = 0; t < RC; ++t) {
        const int half = 1 << t;
        const unsigned d = d0 + t;
        const uint2 *tws = tw + (N - ((size_t)2 << d));
#pragma unroll
        for (int i = 0; i < R; ++i) {
            if (i & half) continue;
            const uint32_t a = x[i];
            uint32_t b = x[i + half];
            if (d != 0) {
                const uint2 w = tws[j + (size_t)(i & (half - 1)) * J];
                b = shoup_mul(b, w.x, w.y, p);
            }
            x[i] = add_mod(a, b, p);
            x[i + half] = sub_mod(a, b, p);
        }
    }
    if constexpr (LAST) {
        if (out.layout == 0) {
            uint32_t *o = out.base + e * out.stor + out.shift;
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const size_t t = base + (size_t)i * J;
                if (t < count) o[t] = SCALE ? shoup_mul(x[i], scale, scale_p, p) : x[i];
            }
        } else {
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const size_t t = base + (size_t)i * J;
                if (t < count) out.base[view_index(out, e, t)] = SCALE ? shoup_mul(x[i], scale, scale_p, p) : x[i];
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < R; ++i) src[base + (size_t)i * J] = x[i];
    }
}

// Stage counts of the column passes covering `stages` stages: as even as
// possible, at most four per pass.
std::vector<unsigned> col_split(unsigned stages)
{
    std::vector<unsigned> r;
    const unsigned npass = (stages + 3) / 4;
    for (unsigned i = 0; i < npass; ++i) r.push_back(stages / npass + (i < stages % npass ? 1u : 0u));
    return r;
}

template <int RC, bool FROM_VIEW>
hipError_t dif_cols_launch(const DevView &in, size_t entries, unsigned L, unsigned s0, const uint2 *tw, uint32_t p,
                           uint32_t *out, hipStream_t s)
{
    const size_t threads = entries * (((size_t)1 << L) >> RC);
    const size_t grid = (threads + 255) / 256;
    if (grid > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((dif_cols_kernel<RC, FROM_VIEW>), dim3((unsigned)grid), dim3(256), 0, s, in, entries, L, s0,
                       tw, p, out);
    return hipGetLastError();
}

template <bool FROM_VIEW>
hipError_t dif_cols(unsigned rc, const DevView &in, size_t entries, unsigned L, unsigned s0, const uint2 *tw,
                    uint32_t p, uint32_t *out, hipStream_t s)
{
    switch (rc) {
    case 1: return dif_cols_launch<1, FROM_VIEW>(in, entries, L, s0, tw, p, out, s);
    case 2: return dif_cols_launch<2, FROM_VIEW>(in, entries, L, s0, tw, p, out, s);
    case 3: return dif_cols_launch<3, FROM_VIEW>(in, entries, L, s0, tw, p, out, s);
    case 4: return dif_cols_launch<4, FROM_VIEW>(in, entries, L, s0, tw, p, out, s);
    default: return hipErrorInvalidValue;
    }
}

template <int RC, bool LAST, bool SCALE>
hipError_t dit_cols_launch(uint32_t *buf, size_t entries, unsigned L, unsigned d0, const uint2 *tw, uint32_t p,
                           const DevView &out, size_t count, uint32_t scale, hipStream_t s)
{
    const size_t threads = entries * (((size_t)1 << L) >> RC);
    const size_t grid = (threads + 255) / 256;
    if (grid > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((dit_cols_kernel<RC, LAST, SCALE>), dim3((unsigned)grid), dim3(256), 0, s, buf, entries, L,
                       d0, tw, p, out, count, scale, shoup_pre(scale, p));
    return hipGetLastError();
}

template <bool LAST, bool SCALE>
hipError_t dit_cols(unsigned rc, uint32_t *buf, size_t entries, unsigned L, unsigned d0, const uint2 *tw, uint32_t p,
                    const DevView &out, size_t count, uint32_t scale, hipStream_t s)
{
    switch (rc) {
    case 1: return dit_cols_launch<1, LAST, SCALE>(buf, entries, L, d0, tw, p, out, count, scale, s);
    case 2: return dit_cols_launch<2, LAST, SCALE>(buf, entries, L, d0, tw, p, out, count, scale, s);
    case 3: return dit_cols_launch<3, LAST, SCALE>(buf, entries, L, d0, tw, p, out, count, scale, s);
    case 4: return dit_cols_launch<4, LAST, SCALE>(buf, entries, L, d0, tw, p, out, count, scale, s);
    default: return hipErrorInvalidValue;
    }
}

// Forward DIF of length 2^L > 2^15: column passes over stages [0, L-14)
// from the view into out, then the LDS-resident DIF of the 2^(L-14) blocks of
// every entry, in place.
hipError_t ntt_forward_long(const DevView &in, size_t entries, unsigned L, const uint2 *tw, uint32_t p, uint32_t *out,
                            hipStream_t s)
{
    const size_t N = (size_t)1 << L, M = (size_t)1 << kLongLds;
    unsigned s0 = 0;
    for (unsigned rc : col_split(L - kLongLds)) {
        hipError_t e = s0 == 0 ? dif_cols<true>(rc, in, entries, L, s0, tw, p, out, s)
                               : dif_cols<false>(rc, in, entries, L, s0, tw, p, out, s);
        if (e != hipSuccess) return e;
        s0 += rc;
    }
    const size_t blocks = entries << (L - kLongLds);
    const DevView bv{out, 0, blocks, 1, M, 0, M};
    const EntryMap id{0, blocks, 0};
    return fwd_launch<kLongLds, kLongLds>(bv, blocks, id, tw + (N - M), p, out, s);
}

// Inverse DIT of length 2^L > 2^15: the LDS-resident DIT of the 2^(L-14)
// blocks of every entry in place on `in`, then column passes for spans
// 2^14 .. 2^(L-1), the last writing the view.
hipError_t ntt_inverse_long(uint32_t *in, size_t entries, unsigned L, const uint2 *tw, uint32_t p, const DevView &out,
                            size_t count, uint32_t scale, hipStream_t s)
{
    const size_t N = (size_t)1 << L, M = (size_t)1 << kLongLds;
    const size_t blocks = entries << (L - kLongLds);
    const DevView bv{in, 0, blocks, 1, M, 0, M};
    hipError_t e = inv_launch<kLongLds, 0>(in, blocks, tw + (N - M), p, bv, 0, M, 1, s);
    if (e != hipSuccess) return e;
    const std::vector<unsigned> split = col_split(L - kLongLds);
    unsigned d0 = kLongLds;
    for (size_t i = 0; i < split.size(); ++i) {
        const bool last = i + 1 == split.size();
        if (!last)
            e = dit_cols<false, false>(split[i], in, entries, L, d0, tw, p, out, count, 1, s);
        else if (scale != 1)
            e = dit_cols<true, true>(split[i], in, entries, L, d0, tw, p, out, count, scale, s);
        else
            e = dit_cols<true, false>(split[i], in, entries, L, d0, tw, p, out, count, 1, s);
        if (e != hipSuccess) return e;
        d0 += split[i];
    }
    return hipSuccess;
}

}  // namespace

hipError_t launch_ntt_forward(const DevView &in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                              uint32_t *out, hipStream_t s)
{
    if (entries == 0) return hipSuccess;
    if (L > (unsigned)kMaxLdsLog && L <= (unsigned)kMaxLog) return ntt_forward_long(in, entries, L, tw, p, out, s);
    const EntryMap id{0, entries, 0};
    switch (L) {
#define LBX_CASE(l) \
    case l: return fwd_launch<l, l>(in, entries, id, tw, p, out, s);
        LBX_CASE(1) LBX_CASE(2) LBX_CASE(3) LBX_CASE(4) LBX_CASE(5) LBX_CASE(6) LBX_CASE(7) LBX_CASE(8)
        LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ntt_inverse(uint32_t *in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                              const DevView &out, size_t count, uint32_t scale, hipStream_t s)
{
    if (entries == 0) return hipSuccess;
    if (L > (unsigned)kMaxLdsLog && L <= (unsigned)kMaxLog)
        return ntt_inverse_long(in, entries, L, tw, p, out, count, scale, s);
    switch (L) {
#define LBX_CASE(l) \
    case l: return inv_launch<l, 0>(in, entries, tw, p, out, 0, count, scale, s);
        LBX_CASE(1) LBX_CASE(2) LBX_CASE(3) LBX_CASE(4) LBX_CASE(5) LBX_CASE(6) LBX_CASE(7) LBX_CASE(8)
        LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

// Block sizes 2^BL of the split transforms: the kernels are written for
// BL in [6, 8]; BL = 6 is instantiated (7 and 8 measured slower at 2^14 points:
// what the per-entry kernels save, the blockwise ones lose), with L in [8, 15].
static bool split_ok(unsigned L, unsigned BL) { return BL == kSplitBL && L >= BL + 2 && L <= 15; }

hipError_t launch_ntt_forward_head(const DevView &in, size_t entries, const EntryMap &map, unsigned L, unsigned BL,
                                   const uint2 *tw, uint32_t p, uint32_t *out, hipStream_t s)
{
    if (entries == 0) return hipSuccess;
    if (!split_ok(L, BL)) return hipErrorInvalidValue;
    switch (L * 16 + BL) {
#define LBX_CASE(l, bl) \
    case l * 16 + bl: return head_launch<l, l - bl>(in, entries, map, tw, p, out, s);
        LBX_CASE(8, 6) LBX_CASE(9, 6) LBX_CASE(10, 6) LBX_CASE(11, 6) LBX_CASE(12, 6) LBX_CASE(13, 6) LBX_CASE(14, 6)
        LBX_CASE(15, 6)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ntt_inverse_tail(const uint32_t *in, size_t entries, unsigned L, unsigned BL, const uint2 *tw,
                                   uint32_t p, const DevView &out, size_t out_first, size_t count, uint32_t scale,
                                   hipStream_t s)
{
    if (entries == 0) return hipSuccess;
    if (!split_ok(L, BL)) return hipErrorInvalidValue;
    switch (L * 16 + BL) {
#define LBX_CASE(l, bl) \
    case l * 16 + bl: return tail_inv_launch<l, bl>(in, entries, tw, p, out, out_first, count, scale, s);
        LBX_CASE(8, 6) LBX_CASE(9, 6) LBX_CASE(10, 6) LBX_CASE(11, 6) LBX_CASE(12, 6) LBX_CASE(13, 6) LBX_CASE(14, 6)
        LBX_CASE(15, 6)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

template <int L, int BL>
static hipError_t tail_launch(dim3 grid, const uint32_t *X, uint8_t *Y, int rows_valid, int row0, int kdim,
                              int row_major, const uint2 *tw, uint32_t p, size_t block_bytes, int kchunks, bool limbs,
                              hipStream_t s)
{
    auto k = limbs ? dif_tail_limbs_kernel<L, BL, true> : dif_tail_limbs_kernel<L, BL, false>;
    hipError_t e = allow_lds(k, Blk<BL>::LDS_BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, grid, dim3(Blk<BL>::T), Blk<BL>::LDS_BYTES, s, X, Y, rows_valid, row0, kdim, row_major,
                       tw, p, block_bytes, kchunks);
    return hipGetLastError();
}

template <int L, int BL>
static hipError_t dit_head_launch(dim3 grid, const uint32_t *Cm, size_t entries, size_t e_first, size_t e_count,
                                  uint32_t *Y, const uint2 *tw, uint32_t p, hipStream_t s)
{
    auto k = dit_head_kernel<L, BL>;
    hipError_t e = allow_lds(k, Blk<BL>::LDS_BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, grid, dim3(Blk<BL>::T), Blk<BL>::LDS_BYTES, s, Cm, entries, e_first, e_count, Y, tw, p);
    return hipGetLastError();
}

hipError_t launch_dif_tail_limbs(const uint32_t *X, uint8_t *Y, size_t rows_total, size_t row0, size_t rows_chunk,
                                 size_t rows_valid, size_t kdim, bool row_major, unsigned L, unsigned BL,
                                 const uint2 *tw, uint32_t p, hipStream_t s, bool limbs)
{
    const size_t rpad = (rows_total + 31) & ~(size_t)31, kchunks = (kdim + 63) / 64;
    if (rows_chunk == 0) return hipSuccess;
    if (row0 + rows_chunk > rpad || rows_valid > rows_chunk || rows_chunk > 65535 || kchunks > 65535 ||
        !split_ok(L, BL))
        return hipErrorInvalidValue;
    const size_t block_bytes = rpad * kchunks * 256;
    dim3 grid((unsigned)kchunks, (unsigned)rows_chunk, (unsigned)((size_t)1 << (L - BL)));
    switch (L * 16 + BL) {
#define LBX_CASE(l, bl)                                                                                            \
    case l * 16 + bl:                                                                                              \
        return tail_launch<l, bl>(grid, X, Y, (int)rows_valid, (int)row0, (int)kdim, row_major ? 1 : 0, tw, p,     \
                                  block_bytes, (int)kchunks, limbs, s);
        LBX_CASE(8, 6) LBX_CASE(9, 6) LBX_CASE(10, 6) LBX_CASE(11, 6) LBX_CASE(12, 6) LBX_CASE(13, 6) LBX_CASE(14, 6)
        LBX_CASE(15, 6)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_dit_head(const uint32_t *Cm, size_t entries, size_t e_first, size_t e_count, unsigned L,
                           unsigned BL, const uint2 *tw, uint32_t p, uint32_t *Y, hipStream_t s)
{
    if (e_count == 0) return hipSuccess;
    const size_t groups = (e_count + 63) / 64;
    if (groups > 65535 || !split_ok(L, BL) || e_first + e_count > entries) return hipErrorInvalidValue;
    dim3 grid((unsigned)groups, (unsigned)((size_t)1 << (L - BL)));
    switch (L * 16 + BL) {
#define LBX_CASE(l, bl) \
    case l * 16 + bl: return dit_head_launch<l, bl>(grid, Cm, entries, e_first, e_count, Y, tw, p, s);
        LBX_CASE(8, 6) LBX_CASE(9, 6) LBX_CASE(10, 6) LBX_CASE(11, 6) LBX_CASE(12, 6) LBX_CASE(13, 6) LBX_CASE(14, 6)
        LBX_CASE(15, 6)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

size_t batch_ntt_lds_bytes()
{
    // 2 groups x G x padded stride for L <= 12, one padded 2^13 row for L = 13: all equal
    return (size_t)Cfg<13>::STRIDE * 4;
}

hipError_t launch_batch_ntt(unsigned L, const BatchJob *jobs, const NttItem *items, uint32_t nitems,
                            uint32_t *counter, bool inverse, unsigned grid, hipStream_t s)
{
    static_assert(2 * Cfg<12>::G * Cfg<12>::STRIDE == Cfg<13>::STRIDE, "batch LDS layout");
    static_assert(2 * Cfg<4>::G * Cfg<4>::STRIDE == Cfg<13>::STRIDE, "batch LDS layout");
    switch (L) {
#define LBX_CASE(l)                                                                                          \
    case l:                                                                                                  \
        if (inverse)                                                                                         \
            hipLaunchKernelGGL((batch_ntt_kernel<l, true>), dim3(grid), dim3(kBatchNttThreads),              \
                               batch_ntt_lds_bytes(), s, jobs, items, nitems, counter);                       \
        else                                                                                                 \
            hipLaunchKernelGGL((batch_ntt_kernel<l, false>), dim3(grid), dim3(kBatchNttThreads),             \
                               batch_ntt_lds_bytes(), s, jobs, items, nitems, counter);                       \
        break;
        LBX_CASE(4) LBX_CASE(5) LBX_CASE(6) LBX_CASE(7) LBX_CASE(8) LBX_CASE(9) LBX_CASE(10) LBX_CASE(11)
        LBX_CASE(12) LBX_CASE(13)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int device_cus()
{
    // initialised once, thread-safely, on first use
    static const int n = [] {
        int dev = 0, c = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                     hipSuccess || c <= 0)
            c = 256;
        return c;
    }();
    return n;
}

}  // namespace lbx
