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

}  // namespace lbx
