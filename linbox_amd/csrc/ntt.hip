// Number-theoretic transform kernels and their launchers: whole transforms
// for the plain path and the NTT entry points, and the split head / tail
// transforms around the point-major pointwise stage (see ntt_device.h).
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
                                                             int rows_valid, int row0, int kdim,
                                                             int row_major, const uint2 *__restrict__ tw, uint32_t p,
                                                             size_t block_bytes, int kchunks)
{
    constexpr int N = 1 << L;
    __shared__ uint32_t sm[kBlk][kBlkStride];
    const int tid = threadIdx.x;
    const int b = blockIdx.x, c = blockIdx.y, r = blockIdx.z;  // r: row within the chunk
    {   // 16 lanes per entry row read its 256-byte block run
        const int ch = tid & 15;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = (tid >> 4) + 16 * u;
            const int kk = c * kBlk + q;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (r < rows_valid && kk < kdim) {
                const size_t e = row_major ? (size_t)r * kdim + kk : (size_t)kk * rows_valid + r;
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
        *reinterpret_cast<uint32_t *>(Y + t * block_bytes + ((size_t)(row0 + r) * kchunks + c) * 256 + (w >> 4) * 64 +
                                      (w & 15) * 4) = word;
    }
}

// First six DIT stages of block blockIdx.x for entries 64*blockIdx.y .. +63,
// read from the point-major result C[t][e]; writes polfirst scratch.
template <int L>
__global__ void __launch_bounds__(256) dit_head_kernel(const uint32_t *__restrict__ Cm, size_t entries,
                                                       size_t e_first, size_t e_count, uint32_t *__restrict__ Y,
                                                       const uint2 *__restrict__ tw, uint32_t p)
{
    constexpr int N = 1 << L;
    __shared__ uint32_t sm[kBlk][kBlkStride];
    const int tid = threadIdx.x;
    const int b = blockIdx.x;
    const size_t e0 = (size_t)blockIdx.y * kBlk;  // relative to e_first
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {  // 64 lanes read 64 consecutive entries of one point
        const int q = tid & 63, lam = (tid >> 6) + 4 * it;
        const size_t t = (size_t)b * kBlk + lam;
        sm[q][lam] = (e0 + q < e_count) ? Cm[t * entries + e_first + e0 + q] : 0u;
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
            if (e0 + qq < e_count)
                *reinterpret_cast<uint4 *>(Y + (e0 + qq) * N + (size_t)b * kBlk + ch * 4) =
                    make_uint4(sm[qq][ch * 4], sm[qq][ch * 4 + 1], sm[qq][ch * 4 + 2], sm[qq][ch * 4 + 3]);
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
            in.base = const_cast<uint32_t *>(isa ? J.a : J.b);
            in.layout = 0;
            in.rows = isa ? J.m : J.k;
            in.cols = isa ? J.k : J.n;
            in.stor = in.size = isa ? J.sa : J.sb;
            in.shift = 0;
            const EntryMap id{0, ~(size_t)0, 0};
            uint32_t *out = isa ? J.ahat : J.bhat;
            if (in.size * 2 <= (size_t)C::N)
                dif_body<L, L, true>(in, e_end, id, J.twf, J.p, out, e_base, tid, lds);
            else
                dif_body<L, L, false>(in, e_end, id, J.twf, J.p, out, e_base, tid, lds);
        } else {
            DevView out;
            out.base = J.c;
            out.layout = 0;
            out.rows = J.m;
            out.cols = J.n;
            out.stor = out.size = J.sc;
            out.shift = 0;
            dit_body<L, 0, false>(J.chat, e_end, J.twi, J.p, out, 0, J.sc, 1u, 0u, e_base, tid, lds);
        }
        __syncthreads();  // LDS is reused by the next item
    }
}

}  // namespace

hipError_t launch_ntt_forward(const DevView &in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                              uint32_t *out, hipStream_t s)
{
    if (entries == 0) return hipSuccess;
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

hipError_t launch_ntt_inverse(const uint32_t *in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                              const DevView &out, size_t count, uint32_t scale, hipStream_t s)
{
    if (entries == 0) return hipSuccess;
    switch (L) {
#define LBX_CASE(l) \
    case l: return inv_launch<l, 0>(in, entries, tw, p, out, 0, count, scale, s);
        LBX_CASE(1) LBX_CASE(2) LBX_CASE(3) LBX_CASE(4) LBX_CASE(5) LBX_CASE(6) LBX_CASE(7) LBX_CASE(8)
        LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ntt_forward_head(const DevView &in, size_t entries, const EntryMap &map, unsigned L,
                                   const uint2 *tw, uint32_t p, uint32_t *out, hipStream_t s)
{
    if (entries == 0) return hipSuccess;
    switch (L) {
#define LBX_CASE(l) \
    case l: return fwd_launch<l, l - 6>(in, entries, map, tw, p, out, s);
        LBX_CASE(8) LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ntt_inverse_tail(const uint32_t *in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                                   const DevView &out, size_t out_first, size_t count, uint32_t scale, hipStream_t s)
{
    if (entries == 0) return hipSuccess;
    switch (L) {
#define LBX_CASE(l) \
    case l: return inv_launch<l, 6>(in, entries, tw, p, out, out_first, count, scale, s);
        LBX_CASE(8) LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_dif_tail_limbs(const uint32_t *X, uint8_t *Y, size_t rows_total, size_t row0, size_t rows_chunk,
                                 size_t rows_valid, size_t kdim, bool row_major, unsigned L, const uint2 *tw,
                                 uint32_t p, hipStream_t s)
{
    const size_t rpad = (rows_total + 31) & ~(size_t)31, kchunks = (kdim + 63) / 64;
    if (rows_chunk == 0) return hipSuccess;
    if (row0 + rows_chunk > rpad || rows_valid > rows_chunk || rows_chunk > 65535 || kchunks > 65535 || L < 8 ||
        L > 15)
        return hipErrorInvalidValue;
    const size_t block_bytes = rpad * kchunks * 256;
    dim3 grid((unsigned)((size_t)1 << (L - 6)), (unsigned)kchunks, (unsigned)rows_chunk);
    switch (L) {
#define LBX_CASE(l)                                                                                                  \
    case l:                                                                                                          \
        hipLaunchKernelGGL(dif_tail_limbs_kernel<l>, grid, dim3(256), 0, s, X, Y, (int)rows_valid, (int)row0,       \
                           (int)kdim, row_major ? 1 : 0, tw, p, block_bytes, (int)kchunks);                          \
        break;
        LBX_CASE(8) LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    }
    return hipGetLastError();
}

hipError_t launch_dit_head(const uint32_t *Cm, size_t entries, size_t e_first, size_t e_count, unsigned L,
                           const uint2 *tw, uint32_t p, uint32_t *Y, hipStream_t s)
{
    if (e_count == 0) return hipSuccess;
    const size_t groups = (e_count + 63) / 64;
    if (groups > 65535 || L < 8 || L > 15 || e_first + e_count > entries) return hipErrorInvalidValue;
    dim3 grid((unsigned)((size_t)1 << (L - 6)), (unsigned)groups);
    switch (L) {
#define LBX_CASE(l)                                                                                               \
    case l: hipLaunchKernelGGL(dit_head_kernel<l>, grid, dim3(256), 0, s, Cm, entries, e_first, e_count, Y, tw, p); \
        break;
        LBX_CASE(8) LBX_CASE(9) LBX_CASE(10) LBX_CASE(11) LBX_CASE(12) LBX_CASE(13) LBX_CASE(14) LBX_CASE(15)
#undef LBX_CASE
    }
    return hipGetLastError();
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

}  // namespace lbx
