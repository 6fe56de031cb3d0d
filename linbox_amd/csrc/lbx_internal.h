// Internal declarations shared by the kernels and the C ABI implementation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "modarith.h"
#include "pw_math.h"

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

// A pointer that a kernel loads from memory (a job descriptor's field), for
// use as a global-memory pointer.  The compiler cannot tell where such a
// pointer points and emits flat instructions for its accesses, and flat loads
// count against the same wait counter as LDS operations, so every LDS wait
// would also wait for them.  The cast to the global address space, held
// behind an empty asm so it is not folded away, lets every access derived
// from the pointer use global instructions.  p must point to global memory
// and be the same across the wave.
template <class T>
__device__ __forceinline__ T *global_ptr(T *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    typedef __attribute__((address_space(1))) T global_t;
    global_t *g = (global_t *)p;
    asm("" : "+s"(g));
    return (T *)g;
#else
    return p;
#endif
}

constexpr int kMaxLdsLog = 15;  // LDS-resident transforms up to 2^15 points
constexpr int kMaxLog = 27;     // longest transform: 2^27 points (the 2-adic valuation of 2013265921 - 1)

// Twiddle table of one root w of order 2^L: the reference's pow_w layout
// (fft-integral.inl:64-77): subarray for butterfly span h at offset N-2h,
// entry j = w^(j*N/(2h)), j < h; interleaved with Shoup constants.
// N-1 entries of {w^e, floor(w^e 2^32/p)}.

// ---- NTT launchers (ntt.hip) ----
// Forward DIF (natural in, bit-reversed out) of `entries` polynomials read
// from `in` (first in.size coefficients, zero-padded to 2^L), written
// contiguously to out[e*2^L ...].  Lengths 2^1..2^kMaxLdsLog stay in LDS
// (one HBM read and write); longer ones, up to 2^kMaxLog, add HBM passes of
// up to four stages each.  `in` and `out` may be the same polfirst storage.
hipError_t launch_ntt_forward(const DevView &in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
                              uint32_t *out, hipStream_t s);
// Inverse DIT (bit-reversed in, natural out) of `entries` contiguous
// polynomials in[e*2^L ...]; coefficients 0..count-1 of entry e are written
// to the view `out` after multiplication by `scale` (scale==1: none).
// Transforms longer than 2^kMaxLdsLog run in several passes and overwrite
// `in`; shorter ones leave it intact.
hipError_t launch_ntt_inverse(uint32_t *in, size_t entries, unsigned L, const uint2 *tw, uint32_t p,
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

// Split transforms for the point-major pointwise form, with blocks of 2^BL
// points (BL = kSplitBL, BL + 2 <= L <= 15):
// forward head = DIF stages [0, L-BL) per entry, written block-major:
//                word w of block j of output row l (input entry map(l)) at
//                (j*entries + l)*2^BL + w, so a tail tile is one contiguous run;
// forward tail = the last BL DIF stages on 2^BL-point blocks of 64 entries,
//                written straight into rows [row0, row0+rows_chunk) of the
//                per-point limb blocks (rows >= rows_valid are zero padding;
//                input entry of (row r, inner q) is r*kdim+q if row_major,
//                else q*rows_valid+r);
// inverse head = the first BL DIT stages on 2^BL-point blocks of the entries
//                [e_first, e_first+e_count) of the point-major product C[t][e],
//                polfirst out (row = e - e_first);
// inverse tail = DIT stages [BL, L) per entry, written to entries
//                out_first + row of the view.
constexpr unsigned kSplitBL = 6;
// Compute units of the device current at the first call (256 on MI355X; 256 if
// the query fails), determined once per process.
int device_cus();
hipError_t launch_ntt_forward_head(const DevView &in, size_t entries, const EntryMap &map, unsigned L, unsigned BL,
                                   const uint2 *tw, uint32_t p, uint32_t *out, hipStream_t s);
// With limbs = false the tail writes the centred values (int32) instead of
// their digits, in the same block geometry ([rpad][kpad] words per point).
hipError_t launch_dif_tail_limbs(const uint32_t *X, uint8_t *Y, size_t rows_total, size_t row0, size_t rows_chunk,
                                 size_t rows_valid, size_t kdim, bool row_major, unsigned L, unsigned BL,
                                 const uint2 *tw, uint32_t p, hipStream_t s, bool limbs = true);
hipError_t launch_dit_head(const uint32_t *Cm, size_t entries, size_t e_first, size_t e_count, unsigned L,
                           unsigned BL, const uint2 *tw, uint32_t p, uint32_t *Y, hipStream_t s);
hipError_t launch_ntt_inverse_tail(const uint32_t *in, size_t entries, unsigned L, unsigned BL, const uint2 *tw,
                                   uint32_t p, const DevView &out, size_t out_first, size_t count, uint32_t scale,
                                   hipStream_t s);

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
// Plain-form pointwise on point-major blocks of centred int32 values (the
// geometry of limb_block_bytes: A as [rpad(m)][kpad(k)], B^T as [rpad(n)][kpad(k)]
// words per point, padding zero): C[t][i*n+j] = scale * sum_q A[i][q] B^T[j][q] mod p.
hipError_t launch_pointwise_plain_pm(const uint32_t *Aw, const uint32_t *Bw, uint32_t *Cmat, size_t m, size_t k,
                                     size_t n, size_t npts, uint32_t p, uint32_t scale, hipStream_t s);
// Point-major [t][e] -> polfirst [e][t] transpose of `entries` x npts words.
hipError_t launch_point_to_entry(const uint32_t *X, uint32_t *Y, size_t entries, size_t npts, hipStream_t s);

// ---- batch scheduler (batch.hip, with kernels in ntt.hip / pointwise.hip) ----

// One product of a batch as the device sees it: operands polfirst, per-job
// slices of the batch workspace for the transforms.
struct BatchJob {
    uint32_t p, L, m, k, n, sa, sb, sc;
    uint32_t tile;            // plain form: pointwise register tile, 0 = 2x2, 1 = 4x4, 2 = 8x8
    uint32_t form;            // pointwise form: 0 plain, 1 INT8 MFMA
    const uint32_t *a, *b;
    uint32_t *c;
    uint32_t *ahat, *bhat, *chat;  // m*k, k*n, m*n rows of 2^L words
    const uint2 *twf, *twi;        // root: forward (Montgomery) table; its inverse: inverse (Shoup) table
    PlainConst pc;                 // plain epilogue constants with scale 2^-L
    uint8_t *alimb, *blimb;        // INT8: per-point limb blocks of A and B^T (limb_block_bytes each)
    uint32_t *cmat;                // INT8: point-major product [t][i*n+j]
    MfmaConst mc;                  // INT8 epilogue constants with scale 2^-L
};
// Transform work item: entries [e0, e0+count) of operand op (0 a, 1 b, 2 c).
struct NttItem {
    uint32_t job, op, e0, count;
};
// INT8 batch work items.  Limbs: the 64-point x 64-entry tile of operand op
// (0 a, 1 b) of one job at points t0.., limb row r and inner chunk c
// (rc = r * kchunks + c).  MFMA: points [t0, t0+count) of one job.
// Transpose: entries [e0, e0+64) x points [t0, t0+64) of one job's
// point-major product into its polfirst chat.
struct LimbItem {
    uint32_t job, op, t0, rc;
};
struct MfmaItem {
    uint32_t job, t0, count, pad_;
};
struct TrItem {
    uint32_t job, e0, t0, pad_;
};
// Pointwise work item: 256 points from pt0 of output tile (ti, tj).
struct PwItem {
    uint32_t job, ti, tj, pt0;
};
constexpr int kBatchMinL = 4, kBatchMaxL = 13;
constexpr int kBatchNttThreads = 512;
// Polynomials of length 2^L one 512-thread transform work item covers.
inline uint32_t batch_polys_per_item(unsigned L) { return L >= 13 ? 1u : (1u << (13 - L)); }
size_t batch_ntt_lds_bytes();
// Persistent work-queue kernels; counter[0] must be zero at launch.  The
// transform kernel serves the items of one length 2^L.
hipError_t launch_batch_ntt(unsigned L, const BatchJob *jobs, const NttItem *items, uint32_t nitems,
                            uint32_t *counter, bool inverse, unsigned grid, hipStream_t s);
hipError_t launch_batch_pointwise(const BatchJob *jobs, const PwItem *items, uint32_t nitems, uint32_t *counter,
                                  unsigned grid, hipStream_t s);
// The INT8 stage of a batch's INT8 jobs, each a persistent work-queue kernel
// over its items (counter at zero on entry): NTT output -> limb blocks, the
// MFMA products (inner dimension at most 4096, one exact int32 chunk), and the
// point-major -> polfirst transpose of the result for the inverse transforms.
hipError_t launch_batch_limbs(const BatchJob *jobs, const LimbItem *items, uint32_t nitems, uint32_t *counter,
                              unsigned grid, hipStream_t s);
hipError_t launch_batch_mfma(const BatchJob *jobs, const MfmaItem *items, uint32_t nitems, uint32_t *counter,
                             unsigned grid, hipStream_t s);
hipError_t launch_batch_point_to_entry(const BatchJob *jobs, const TrItem *items, uint32_t nitems,
                                       uint32_t *counter, unsigned grid, hipStream_t s);
// Zero `n` counters with a kernel (graph-safe, no memset node).
hipError_t launch_zero_counters(uint32_t *counters, int n, hipStream_t s);

// ---- layout helpers (layout.hip) ----
// out[e*stride + t] = view(e, t) for t < count (zero for t >= view.size).
hipError_t launch_gather(const DevView &in, size_t entries, size_t count, uint32_t *out, size_t stride,
                         hipStream_t s);
// out[e*hdeg + i] = in(e, hdeg-1-i) (0 past the view's size), i < hdeg: the
// reversal of the small midproduct operand (wordsize-fast.inl:228-235).
hipError_t launch_reverse_gather(const DevView &in, size_t entries, size_t hdeg, uint32_t *out, hipStream_t s);
// Word pairs of a twiddle table entry w^e: the inverse (DIT) transforms take
// {w^e, Shoup constant floor(w^e 2^32 / p)}, the forward (DIF) ones the signed
// Montgomery pair {w^e 2^32 mod p, that times p^-1 mod 2^32} (modarith.h).
enum TwFormat : int { kTwShoup = 0, kTwMont = 1 };
// The pow_w twiddle table of root w of order 2^L (N-1 entries) computed on
// the device, identical to the host-built one of the same format.
hipError_t launch_build_twiddles(uint2 *tw, uint32_t p, unsigned L, uint32_t w, TwFormat fmt, hipStream_t s);
// view(e, t) = value for t0 <= t < t0+count (a graph-safe fill, no memset node).
hipError_t launch_fill(const DevView &out, size_t entries, size_t t0, size_t count, uint32_t value, hipStream_t s);
// view(e, t) = in[e*stride + t] for t < count.
hipError_t launch_scatter(const uint32_t *in, size_t stride, size_t entries, size_t count, const DevView &out,
                          hipStream_t s);

}  // namespace lbx
