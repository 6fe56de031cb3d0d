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
    Staged ss, sd;
    lbx_status st = stage_in(src, 4, true, &ss);
    if (st == LBX_OK) st = stage_in(dst, 4, true, &sd);  // the rest of dst's storage is kept
    if (st == LBX_OK) st = from_hip(launch_copy(devview(src, ss.dev), devview(dst, sd.dev), entries, cnt, beg, start, s));
    if (st == LBX_OK && sd.owned) {
        hipError_t e = hipMemcpyAsync(dst->data, sd.dev, storage_words(dst) * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        st = from_hip(e);
    } else if (st == LBX_OK && ss.owned) {
        st = from_hip(hipStreamSynchronize(s));
    }
    if (ss.owned || sd.owned) (void)hipStreamSynchronize(s);
    release(ss);
    release(sd);
    return st;
}
