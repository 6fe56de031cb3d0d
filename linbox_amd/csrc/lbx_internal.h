// Internal declarations shared by the kernels and the C ABI implementation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "modarith.h"

namespace lbx {

// A polynomial matrix (or view) in device memory, in one of the three
// PolynomialMatrix layouts (polynomial-matrix.h:54-77).
struct DevView {
    uint32_t *base;
    int layout;  // 0 polfirst, 1 matfirst, 2 matrowfirst
    size_t rows, cols;
    size_t stor;   // storage coefficients per entry
    size_t shift;  // first coefficient of the view
    size_t size;   // coefficients visible in the view
};

// Address of coefficient t of entry e (e = i*cols + j) of a view.
__host__ __device__ __forceinline__ size_t view_index(const DevView &v, size_t e, size_t t)
{
    const size_t tt = v.shift + t;
    if (v.layout == 0) return e * v.stor + tt;
    if (v.layout == 1) return e + tt * v.rows * v.cols;
    const size_t i = e / v.cols, j = e - i * v.cols;
    return i * v.cols * v.stor + j + tt * v.cols;
}

constexpr int kMaxLdsLog = 15;  // LDS-resident transforms up to 2^15 points

// Twiddle table of one root w of order 2^L: the reference's pow_w layout
// (fft-integral.inl:64-77): subarray for butterfly span h at offset N-2h,
// entry j = w^(j*N/(2h)), j < h; interleaved with Shoup constants.
// N-1 entries of {w^e, floor(w^e 2^32/p)}.

// ---- NTT launchers (ntt.hip) ----
// Forward DIF (natural in, bit-reversed out) of `entries` polynomials read
// from `in` (first in.size coefficients, zero-padded to 2^L), written
// contiguously to out[e*2^L ...].
hipError_t launch_ntt_forward(const DevView &in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                              uint32_t *out, hipStream_t s);
// Inverse DIT (bit-reversed in, natural out) of `entries` contiguous
// polynomials in[e*2^L ...]; coefficients 0..count-1 of entry e are written
// to the view `out` after multiplication by `scale` (scale==1: none).
hipError_t launch_ntt_inverse(const uint32_t *in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                              const DevView &out, size_t count, uint32_t scale, hipStream_t s);

// Output row l of a forward transform reads input entry
// (l / cw) * stride + off + (l % cw): a contiguous range (cw = count), or a
// set of column windows of a matrix (cw columns starting at off, stride = cols).
struct EntryMap {
    size_t off, cw, stride;
    __host__ __device__ __forceinline__ size_t operator()(size_t l) const
    {
        const size_t q = l / cw;
        return q * stride + off + (l - q * cw);
    }
};

// Split transforms for the point-major pointwise form (L in [8,15]):
// forward head = DIF stages [0, L-6) per entry, polfirst out (row l <- entry map(l));
// forward tail = the last six DIF stages on 64-point blocks of 64 entries,
//                written straight into rows [row0, row0+rows_chunk) of the
//                per-point limb blocks (rows >= rows_valid are zero padding;
//                input entry of (row r, inner q) is r*kdim+q if row_major,
//                else q*rows_valid+r);
// inverse head = the first six DIT stages on 64-point blocks of the entries
//                [e_first, e_first+e_count) of the point-major product C[t][e],
//                polfirst out (row = e - e_first);
// inverse tail = DIT stages [6, L) per entry, written to entries
//                out_first + row of the view.
hipError_t launch_ntt_forward_head(const DevView &in, size_t entries, const EntryMap &map, unsigned L,
                                   const uint2 *tw, uint32_t p, uint32_t *out, hipStream_t s);
hipError_t launch_dif_tail_limbs(const uint32_t *X, uint8_t *Y, size_t rows_total, size_t row0, size_t rows_chunk,
                                 size_t rows_valid, size_t kdim, bool row_major, unsigned L, const uint2 *tw,
                                 uint32_t p, hipStream_t s);
hipError_t launch_dit_head(const uint32_t *Cm, size_t entries, size_t e_first, size_t e_count, unsigned L,
                           const uint2 *tw, uint32_t p, uint32_t *Y, hipStream_t s);
hipError_t launch_ntt_inverse_tail(const uint32_t *in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                                   const DevView &out, size_t out_first, size_t count, uint32_t scale, hipStream_t s);

// ---- pointwise stage (pointwise.hip) ----
// C[e'][t] = scale * sum_q A[i*k+q][t] B[q*n+j][t] mod p at every point t < npts,
// all operands polfirst with row stride npts (e' = i*n+j).
hipError_t launch_pointwise_plain(const uint32_t *A, const uint32_t *B, uint32_t *C, size_t m, size_t k,
                                  size_t n, size_t npts, uint32_t p, uint32_t scale, hipStream_t s);

// Bytes of one point's limb block for an operand with `rows` rows and inner
// dimension kdim (rows padded to 32, inner dimension to 64, four planes).
size_t limb_block_bytes(size_t rows, size_t kdim);
// Polfirst NTT data -> per-point limb blocks.  Entry of (row r, inner q) is
// r*kdim+q when row_major, q*rows+r otherwise (the k x n operand).
hipError_t launch_to_limbs(const uint32_t *X, uint8_t *Y, size_t rows, size_t kdim, bool row_major,
                           size_t npts, uint32_t p, hipStream_t s);
// INT8-MFMA pointwise: per point, C[t][i*n+j] (point-major) from limb blocks.
hipError_t launch_pointwise_mfma(const uint8_t *Al, const uint8_t *Bl, uint32_t *Cmat, size_t m, size_t k,
                                 size_t n, size_t npts, uint32_t p, uint32_t scale, hipStream_t s);
// Point-major [t][e] -> polfirst [e][t] transpose of `entries` x npts words.
hipError_t launch_point_to_entry(const uint32_t *X, uint32_t *Y, size_t entries, size_t npts, hipStream_t s);

// ---- layout helpers (layout.hip) ----
// out[e*stride + t] = view(e, t) for t < count (zero for t >= view.size).
hipError_t launch_gather(const DevView &in, size_t entries, size_t count, uint32_t *out, size_t stride,
                         hipStream_t s);
// view(e, t) = in[e*stride + t] for t < count.
hipError_t launch_scatter(const uint32_t *in, size_t stride, size_t entries, size_t count, const DevView &out,
                          hipStream_t s);

}  // namespace lbx
