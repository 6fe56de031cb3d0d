// Layout conversion on the device: PolynomialMatrix::copy(M, beg, end, start)
// between any two of the three storage layouts and SubPolynomialMatrix views
// (polynomial-matrix.h:206-243 into polfirst, :508-532 into matfirst,
// :763-790 into matrowfirst, :988-992 through a view): for every entry e and
// i in [beg, end], dst[e][start + i - beg] = src[e][i].
//
// The copy is pure data movement, bound by HBM.  Each 256-thread group moves
// 64-entry x 64-coefficient tiles through LDS: it reads along the source's
// contiguous axis (coefficients for polfirst, entries for matfirst and
// matrowfirst) and writes along the destination's, so both sides are
// coalesced whichever layouts meet.  The 65-word LDS rows keep both the
// row-wise and the column-wise LDS accesses of a wave in distinct banks.
// Aligned polfirst <-> matfirst pairs move 16 bytes per lane access
// (copy_vec_kernel); every other case moves 4.
#include "lbx_ctx.h"

namespace lbx {
namespace {

// Word address of coefficient t of entry e, with the row/column split of the
// matrowfirst layout in 32-bit arithmetic (entries < 2^32 is checked on the host).
__device__ __forceinline__ size_t copy_addr(const DevView &v, uint32_t e, size_t t)
{
    const size_t tt = v.shift + t;
    if (v.layout == 0) return (size_t)e * v.stor + tt;
    if (v.layout == 1) return (size_t)e + tt * v.rows * v.cols;
    const uint32_t cols = (uint32_t)v.cols, i = e / cols, j = e - i * cols;
    return (size_t)i * v.cols * v.stor + j + tt * v.cols;
}

template <bool SRC_T, bool DST_T>  // true: coefficients are the contiguous axis of that side
__global__ void __launch_bounds__(256) copy_tile_kernel(DevView src, DevView dst, uint32_t entries, size_t cnt,
                                                        size_t beg, size_t start, size_t etiles, size_t ntiles)
{
    __shared__ uint32_t tile[64][65];  // [entry][coefficient] of the tile
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (size_t tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
        const uint32_t e0 = (uint32_t)(tix % etiles) * 64;
        const size_t t0 = (tix / etiles) * 64;
#pragma unroll 4
        for (int r = w; r < 64; r += 4) {
            const uint32_t el = SRC_T ? r : lane, tl = SRC_T ? lane : r;
            const uint32_t e = e0 + el;
            const size_t t = t0 + tl;
            if (e < entries && t < cnt) tile[el][tl] = src.base[copy_addr(src, e, beg + t)];
        }
        __syncthreads();
#pragma unroll 4
        for (int r = w; r < 64; r += 4) {
            const uint32_t el = DST_T ? r : lane, tl = DST_T ? lane : r;
            const uint32_t e = e0 + el;
            const size_t t = t0 + tl;
            if (e < entries && t < cnt) dst.base[copy_addr(dst, e, start + t)] = tile[el][tl];
        }
        __syncthreads();  // the tile is refilled by the next iteration
    }
}

// The polfirst <-> matfirst pairs with 16-byte global accesses: 16 lanes move
// one 64-word run (an entry's 64 coefficients, or a coefficient's 64 entries),
// four words to and from the LDS tile each.  The host checks that every run
// starts 16-byte aligned: bases, the polfirst stride and coefficient offset,
// and the entry count (the matfirst stride) are multiples of four words.
// Coefficient runs that end inside a chunk (cnt not a multiple of 4) take
// their last words one at a time.
template <bool TO_MATFIRST>
__global__ void __launch_bounds__(256) copy_vec_kernel(const uint32_t *__restrict__ src, size_t src_stride,
                                                       size_t src_t0, uint32_t *__restrict__ dst, size_t dst_stride,
                                                       size_t dst_t0, uint32_t entries, size_t cnt, size_t etiles,
                                                       size_t ntiles)
{
    // polfirst word (e, t) at base + e*stride + t0 + t; matfirst at base + (t0 + t)*stride + e
    __shared__ uint32_t tile[64][65];  // [entry][coefficient]
    const int c = threadIdx.x & 15, r0 = threadIdx.x >> 4;
    for (size_t tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
        const uint32_t e0 = (uint32_t)(tix % etiles) * 64;
        const size_t t0 = (tix / etiles) * 64;
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int row = r0 + 16 * it;  // entry (polfirst source) or coefficient (matfirst source)
            const uint32_t e = TO_MATFIRST ? e0 + row : e0 + 4 * c;
            const size_t t = TO_MATFIRST ? t0 + 4 * c : t0 + row;
            if (e >= entries || t >= cnt) continue;
            const uint32_t *g = TO_MATFIRST ? src + (size_t)e * src_stride + src_t0 + t
                                            : src + (src_t0 + t) * src_stride + e;
            uint32_t w[4];
            if (!TO_MATFIRST || t + 3 < cnt) {
                const uint4 v = *reinterpret_cast<const uint4 *>(g);
                w[0] = v.x, w[1] = v.y, w[2] = v.z, w[3] = v.w;
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) w[i] = t + i < cnt ? g[i] : 0u;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (TO_MATFIRST)
                    tile[row][4 * c + i] = w[i];
                else
                    tile[4 * c + i][row] = w[i];
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int row = r0 + 16 * it;  // coefficient (matfirst destination) or entry (polfirst destination)
            const uint32_t e = TO_MATFIRST ? e0 + 4 * c : e0 + row;
            const size_t t = TO_MATFIRST ? t0 + row : t0 + 4 * c;
            if (e >= entries || t >= cnt) continue;
            uint32_t w[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) w[i] = TO_MATFIRST ? tile[4 * c + i][row] : tile[row][4 * c + i];
            uint32_t *g = TO_MATFIRST ? dst + (dst_t0 + t) * dst_stride + e : dst + (size_t)e * dst_stride + dst_t0 + t;
            if (TO_MATFIRST || t + 3 < cnt) {
                *reinterpret_cast<uint4 *>(g) = make_uint4(w[0], w[1], w[2], w[3]);
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (t + i < cnt) g[i] = w[i];
            }
        }
        __syncthreads();  // the tile is refilled by the next iteration
    }
}

template <bool SRC_T, bool DST_T>
hipError_t copy_launch(const DevView &src, const DevView &dst, size_t entries, size_t cnt, size_t beg, size_t start,
                       hipStream_t s)
{
    const size_t etiles = (entries + 63) / 64, ntiles = etiles * ((cnt + 63) / 64);
    const unsigned grid = (unsigned)(ntiles < 16384 ? ntiles : 16384);
    hipLaunchKernelGGL((copy_tile_kernel<SRC_T, DST_T>), dim3(grid), dim3(256), 0, s, src, dst, (uint32_t)entries,
                       cnt, beg, start, etiles, ntiles);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_copy(const DevView &src, const DevView &dst, size_t entries, size_t cnt, size_t beg, size_t start,
                       hipStream_t s)
{
    const bool st = src.layout == 0, dt = dst.layout == 0;
    const DevView &pf = st ? src : dst;  // the polfirst side of a polfirst <-> matfirst pair
    const size_t pf_t0 = pf.shift + (st ? beg : start), mf_t0 = st ? dst.shift + start : src.shift + beg;
    const bool pair = (st && dst.layout == 1) || (dt && src.layout == 1);
    const bool aligned = ((uintptr_t)src.base & 15) == 0 && ((uintptr_t)dst.base & 15) == 0 && pf.stor % 4 == 0 &&
                         pf_t0 % 4 == 0 && entries % 4 == 0;
    if (pair && aligned) {
        const size_t etiles = (entries + 63) / 64, ntiles = etiles * ((cnt + 63) / 64);
        const unsigned grid = (unsigned)(ntiles < 16384 ? ntiles : 16384);
        if (st)
            hipLaunchKernelGGL(copy_vec_kernel<true>, dim3(grid), dim3(256), 0, s, src.base, src.stor, pf_t0, dst.base,
                               entries, mf_t0, (uint32_t)entries, cnt, etiles, ntiles);
        else
            hipLaunchKernelGGL(copy_vec_kernel<false>, dim3(grid), dim3(256), 0, s, src.base, entries, mf_t0,
                               dst.base, dst.stor, pf_t0, (uint32_t)entries, cnt, etiles, ntiles);
        return hipGetLastError();
    }
    if (st && dt) return copy_launch<true, true>(src, dst, entries, cnt, beg, start, s);
    if (st) return copy_launch<true, false>(src, dst, entries, cnt, beg, start, s);
    if (dt) return copy_launch<false, true>(src, dst, entries, cnt, beg, start, s);
    return copy_launch<false, false>(src, dst, entries, cnt, beg, start, s);
}

}  // namespace lbx

using namespace lbx;

extern "C" lbx_status lbx_polymat_copy_u32(lbx_ctx *ctx, lbx_polmat *dst, const lbx_polmat *src, size_t beg,
                                           size_t end, size_t start, void *stream)
{
    return guarded([&]() -> lbx_status {
        if (!ctx || !dst || !src) return LBX_ERR_INVALID_ARG;
        LBX_TRY(check_view(src));
        LBX_TRY(check_view(dst));
        if (src->rows != dst->rows || src->cols != dst->cols) return LBX_ERR_SIZE_MISMATCH;
        if (beg > end || end >= src->size) return LBX_ERR_INVALID_ARG;
        const size_t cnt = end - beg + 1;
        if (start > dst->size || cnt > dst->size - start) return LBX_ERR_SIZE_MISMATCH;  // _size >= end-beg+1
        if (src->data == dst->data) return LBX_ERR_INVALID_ARG;                          // overlapping storage
        const size_t entries = src->rows * src->cols;
        if (entries >= ((size_t)1 << 32)) return LBX_ERR_UNSUPPORTED;
        std::lock_guard<std::mutex> lock(ctx->mu);
        DeviceGuard guard(ctx->device);
        hipStream_t s = (hipStream_t)stream;
        StreamOrder stream_order(ctx, s);
        Staged ss, sd;
        lbx_status st = stage_in(ctx, "stage_a", src, 4, true, &ss, s);
        if (st == LBX_OK) st = stage_in(ctx, "stage_c", dst, 4, true, &sd, s);  // the rest of dst's storage is kept
        if (st == LBX_OK) st = from_hip(launch_copy(devview(src, ss.dev), devview(dst, sd.dev), entries, cnt, beg, start, s));
        if (st == LBX_OK && sd.owned) {
            st = device_to_host(ctx, dst->data, sd.dev, storage_words(dst) * 4, s);
        } else if (st == LBX_OK && ss.owned) {
            st = from_hip(hipStreamSynchronize(s));
        }
        if (ss.owned || sd.owned) (void)hipStreamSynchronize(s);
        release(ss);
        release(sd);
        return st;
    });
}
