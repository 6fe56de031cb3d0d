// Copies between PolynomialMatrix views and contiguous entry-major buffers.
// Used where no transform is involved: the constant-matrix product that
// PolynomialMatrixMulDomain::mul sends to the naive domain when
// a.size()+b.size() <= 2 (polynomial-matrix-domain.h:61-70).
#include "lbx_internal.h"

namespace lbx {
namespace {

__global__ void __launch_bounds__(256) gather_kernel(DevView in, size_t entries, size_t count, uint32_t *out,
                                                     size_t stride)
{
    const size_t total = entries * count;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const size_t e = i / count, t = i - e * count;
        out[e * stride + t] = t < in.size ? in.base[view_index(in, e, t)] : 0u;
    }
}

__global__ void __launch_bounds__(256) scatter_kernel(const uint32_t *in, size_t stride, size_t entries, size_t count,
                                                      DevView out)
{
    const size_t total = entries * count;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const size_t e = i / count, t = i - e * count;
        out.base[view_index(out, e, t)] = in[e * stride + t];
    }
}

__global__ void __launch_bounds__(256) fill_kernel(DevView out, size_t entries, size_t t0, size_t count, uint32_t value)
{
    const size_t total = entries * count;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const size_t e = i / count, t = t0 + (i - e * count);
        out.base[view_index(out, e, t)] = value;
    }
}

// pow_w table of init_powers (fft-integral.inl:1614-1650) built on the
// device: entry idx in [N-2h, N-h) is w^((idx-(N-2h)) N/(2h)), as the pair of
// words `fmt` asks for, the same words the host-built cached table holds.
__global__ void __launch_bounds__(256) build_twiddles_kernel(uint2 *tw, uint32_t p, unsigned L, uint32_t w,
                                                             TwFormat fmt)
{
    const uint32_t pinv = inv_mod32(p);
    const size_t n = (size_t)1 << L;
    const size_t count = n > 1 ? n - 1 : 1;
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < count; idx += (size_t)gridDim.x * 256) {
        uint64_t e = 0;
        if (n > 1) {
            const size_t r = n - idx - 1;  // in [h, 2h)
            const size_t h = (size_t)1 << (63 - __builtin_clzll((unsigned long long)r));
            e = (uint64_t)(idx - (n - 2 * h)) * (n / (2 * h));
        }
        uint64_t acc = 1 % p, base = w % p;
        while (e) {
            if (e & 1) acc = acc * base % p;
            base = base * base % p;
            e >>= 1;
        }
        if (fmt == kTwMont) {
            const uint32_t wm = mont_wm((uint32_t)acc, p);
            tw[idx] = make_uint2(wm, mont_wq(wm, pinv));
        } else {
            tw[idx] = make_uint2((uint32_t)acc, (uint32_t)((acc << 32) / p));
        }
    }
}

unsigned grid_for(size_t total)
{
    size_t g = (total + 255) / 256;
    return (unsigned)(g < 65536 ? (g ? g : 1) : 65536);
}

}  // namespace

hipError_t launch_gather(const DevView &in, size_t entries, size_t count, uint32_t *out, size_t stride, hipStream_t s)
{
    hipLaunchKernelGGL(gather_kernel, dim3(grid_for(entries * count)), dim3(256), 0, s, in, entries, count, out, stride);
    return hipGetLastError();
}

hipError_t launch_scatter(const uint32_t *in, size_t stride, size_t entries, size_t count, const DevView &out,
                          hipStream_t s)
{
    hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(entries * count)), dim3(256), 0, s, in, stride, entries, count, out);
    return hipGetLastError();
}

hipError_t launch_fill(const DevView &out, size_t entries, size_t t0, size_t count, uint32_t value, hipStream_t s)
{
    if (!entries || !count) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(entries * count)), dim3(256), 0, s, out, entries, t0, count, value);
    return hipGetLastError();
}

hipError_t launch_build_twiddles(uint2 *tw, uint32_t p, unsigned L, uint32_t w, TwFormat fmt, hipStream_t s)
{
    const size_t n = (size_t)1 << L;
    hipLaunchKernelGGL(build_twiddles_kernel, dim3(grid_for(n > 1 ? n - 1 : 1)), dim3(256), 0, s, tw, p, L, w, fmt);
    return hipGetLastError();
}

}  // namespace lbx
