// Pointwise stage of the FFT product: at every evaluation point t,
// C_t = A_t . B_t (m x k . k x n) mod p, then the 1/2^L scaling of
// mul_fft (matpoly-mult-fft-wordsize-fast.inl:160-164,183-189) folded in.
//
// The reference runs one fflas fgemm per point.  Here all points are one
// launch, in one of two forms:
//
//  * plain: vector ALUs, one point per lane, a TIxTJ register tile of 64-bit
//    accumulators per lane.  Each product a*b < p^2 < 2^62 is added with
//    carry into (acc:u64, hi:u32), i.e. reduction is delayed to the end of the
//    k loop, where (hi,acc) is reduced with three Shoup products.  Operands
//    are read straight from the polfirst NTT output: consecutive lanes read
//    consecutive points, so every load is coalesced.
//
//  * INT8 MFMA: every residue is centred (|c| < p/2 < 2^30) and split into
//    four balanced signed 8-bit digits, c = sum_s d_s 2^(8s).  Then
//    C = sum_{s,t} 2^(8(s+t)) (A_s . B_t), and the 16 digit products are
//    accumulated on v_mfma_i32_16x16x64_i8 directly into the seven
//    anti-diagonals D_d = sum_{s+t=d} A_s.B_t (int32, exact: |D_d| <= 4 k 2^14,
//    k is processed in chunks of 4096 so |D_d| < 2^28).  The epilogue forms
//    sum_d D_d (2^(8d) / 2^L mod p) in int64 (|.| < 2^62) and reduces it.
//    The digit split happens once per element in launch_to_limbs, which also
//    transposes the NTT output into per-point blocks laid out for direct
//    16-byte operand loads.
#include "lbx_internal.h"
#include "pw_math.h"

namespace lbx {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

template <int TI, int TJ>
__device__ __forceinline__ void pw_plain_body(const uint32_t *__restrict__ A, const uint32_t *__restrict__ B,
                                              uint32_t *__restrict__ C, int m, int k, int n, size_t npts,
                                              uint32_t p, const PlainConst &pc, size_t t, int i0, int j0)
{
    // acc += a*b is one v_mad_u64_u32; every pc.fold products the bits above
    // 2^62 are folded back (plain_fold), which keeps acc below 2^64
    uint64_t acc[TI][TJ];
#pragma unroll
    for (int ii = 0; ii < TI; ++ii)
#pragma unroll
        for (int jj = 0; jj < TJ; ++jj) acc[ii][jj] = 0;
    for (int q0 = 0; q0 < k; q0 += (int)pc.fold) {
        const int q1 = min(k, q0 + (int)pc.fold);
        for (int q = q0; q < q1; ++q) {
            uint32_t a[TI], b[TJ];
#pragma unroll
            for (int ii = 0; ii < TI; ++ii)
                a[ii] = (i0 + ii < m) ? A[((size_t)(i0 + ii) * k + q) * npts + t] : 0u;
#pragma unroll
            for (int jj = 0; jj < TJ; ++jj)
                b[jj] = (j0 + jj < n) ? B[((size_t)q * n + j0 + jj) * npts + t] : 0u;
#pragma unroll
            for (int ii = 0; ii < TI; ++ii)
#pragma unroll
                for (int jj = 0; jj < TJ; ++jj) acc[ii][jj] += (uint64_t)a[ii] * b[jj];
        }
#pragma unroll
        for (int ii = 0; ii < TI; ++ii)
#pragma unroll
            for (int jj = 0; jj < TJ; ++jj)
                acc[ii][jj] = plain_fold(acc[ii][jj], pc);
    }
#pragma unroll
    for (int ii = 0; ii < TI; ++ii)
#pragma unroll
        for (int jj = 0; jj < TJ; ++jj) {
            if (i0 + ii >= m || j0 + jj >= n) continue;
            C[((size_t)(i0 + ii) * n + j0 + jj) * npts + t] = plain_finish(acc[ii][jj], p, pc);
        }
}

template <int TI, int TJ>
__global__ void __launch_bounds__(256) pw_plain_kernel(const uint32_t *__restrict__ A, const uint32_t *__restrict__ B,
                                                       uint32_t *__restrict__ C, int m, int k, int n, size_t npts,
                                                       uint32_t p, PlainConst pc)
{
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= npts) return;
    const int tiles_j = (n + TJ - 1) / TJ;
    pw_plain_body<TI, TJ>(A, B, C, m, k, n, npts, p, pc, t, (blockIdx.y / tiles_j) * TI, (blockIdx.y % tiles_j) * TJ);
}

// Persistent work-queue form over the items of a batch: each workgroup takes
// 256 points of one output tile of one job at a time.
__global__ void __launch_bounds__(256) batch_pointwise_kernel(const BatchJob *__restrict__ jobs,
                                                              const PwItem *__restrict__ items, uint32_t nitems,
                                                              uint32_t *counter)
{
    __shared__ uint32_t item_sh;
    for (;;) {
        if (threadIdx.x == 0) item_sh = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t it = __builtin_amdgcn_readfirstlane(item_sh);  // uniform: scalar loads of the item
        __syncthreads();
        if (it >= nitems) return;
        const PwItem w = items[it];
        const BatchJob &J = jobs[w.job];
        const size_t npts = (size_t)1 << J.L;
        const size_t t = (size_t)w.pt0 + threadIdx.x;
        const uint32_t *A = global_ptr(J.ahat), *B = global_ptr(J.bhat);
        uint32_t *Cp = global_ptr(J.chat);
        if (t < npts) {
            if (J.tile == 0)
                pw_plain_body<2, 2>(A, B, Cp, J.m, J.k, J.n, npts, J.p, J.pc, t, w.ti * 2, w.tj * 2);
            else if (J.tile == 1)
                pw_plain_body<4, 4>(A, B, Cp, J.m, J.k, J.n, npts, J.p, J.pc, t, w.ti * 4, w.tj * 4);
            else
                pw_plain_body<8, 8>(A, B, Cp, J.m, J.k, J.n, npts, J.p, J.pc, t, w.ti * 8, w.tj * 8);
        }
    }
}

__global__ void zero_counters_kernel(uint32_t *c, int n)
{
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

// ---- limb blocks -------------------------------------------------------
// Per point, an operand with `rows` rows and inner dimension kdim is stored as
// [rpad][kpad/64][4 planes][64 bytes]: the 64 inner-dimension digits of plane s
// of one row for one 64-wide k step are contiguous, so a lane's MFMA operand
// (16 consecutive digits) is one 16-byte load.
__host__ __device__ inline size_t rpad_of(size_t rows) { return (rows + 31) & ~(size_t)31; }
__host__ __device__ inline size_t kpad_of(size_t kdim) { return (kdim + 63) & ~(size_t)63; }
__device__ __forceinline__ size_t limb_block_bytes_dev(size_t rows, size_t kdim) { return rpad_of(rows) * kpad_of(kdim) * 4; }

// One 64-point x 64-entry tile: points t0.., inner chunk c, row r of the
// padded range; reads 64 consecutive points of 64 entries (coalesced) and
// writes, for each point, that row's 256-byte run of the limb block.
__device__ __forceinline__ void limbs_tile(const uint32_t *__restrict__ X, uint8_t *__restrict__ Y, int rows,
                                           int kdim, int row_major, size_t npts, uint32_t p, size_t block_bytes,
                                           int kchunks, size_t t0, int c, int r, int tid, uint32_t (&tile)[64][65])
{
    const int q0 = c * 64;
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {
        const int ql = it * 4 + tid / 64, tl = tid % 64;
        const int q = q0 + ql;
        const size_t t = t0 + tl;
        uint32_t v = 0;
        if (r < rows && q < kdim && t < npts) {
            const size_t e = row_major ? (size_t)r * kdim + q : (size_t)q * rows + r;
            v = to_limbs(X[e * npts + t], p);
        }
        tile[ql][tl] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {
        const int tl = it * 4 + tid / 64, w = tid % 64;
        const size_t t = t0 + tl;
        if (t >= npts) continue;
        const int s = w >> 4, qb = (w & 15) * 4;
        const uint32_t sh = 8 * s;
        const uint32_t word = ((tile[qb][tl] >> sh) & 0xffu) | (((tile[qb + 1][tl] >> sh) & 0xffu) << 8) |
                              (((tile[qb + 2][tl] >> sh) & 0xffu) << 16) | (((tile[qb + 3][tl] >> sh) & 0xffu) << 24);
        uint8_t *dst = Y + t * block_bytes + ((size_t)r * kchunks + c) * 256 + s * 64 + (w & 15) * 4;
        *reinterpret_cast<uint32_t *>(dst) = word;
    }
}

__global__ void __launch_bounds__(256) to_limbs_kernel(const uint32_t *__restrict__ X, uint8_t *__restrict__ Y,
                                                       int rows, int kdim, int row_major, size_t npts, uint32_t p,
                                                       size_t block_bytes, int kchunks)
{
    __shared__ uint32_t tile[64][65];
    limbs_tile(X, Y, rows, kdim, row_major, npts, p, block_bytes, kchunks, (size_t)blockIdx.x * 64, blockIdx.y,
               blockIdx.z, threadIdx.x, tile);
}

__device__ __forceinline__ uint32_t mfma_combine(const v4i (&acc)[7], int r, uint32_t p, const MfmaConst &mc)
{
    int32_t D[7];
#pragma unroll
    for (int d = 0; d < 7; ++d) D[d] = acc[d][r];
    return combine_diagonals(D, p, mc);
}

constexpr int kMaxKbPerChunk = 64;  // 64 steps of 64 = 4096 inner products per exact int32 chunk

// One wave: the 32x32 output tile (i0, j0) of one point from its limb
// blocks A and B (read from HBM/L2), as 2x2 16x16 MFMA tiles for each of the
// 16 digit pairs, k in exact int32 chunks of 4096.
__device__ __forceinline__ void mfma_tile_global(const uint8_t *__restrict__ A, const uint8_t *__restrict__ B,
                                                 uint32_t *__restrict__ Ct, int m, int n, int kchunks, int i0,
                                                 int j0, uint32_t p, const MfmaConst &mc, int l16, int lg)
{
    const size_t row_bytes = (size_t)kchunks * 256;
    {
        uint32_t res[2][2][4];
#pragma unroll
        for (int ai = 0; ai < 2; ++ai)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int r = 0; r < 4; ++r) res[ai][bj][r] = 0;
        for (int kc0 = 0; kc0 < kchunks; kc0 += kMaxKbPerChunk) {
            const int kc1 = min(kchunks, kc0 + kMaxKbPerChunk);
            v4i acc[2][2][7];
#pragma unroll
            for (int ai = 0; ai < 2; ++ai)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                    for (int d = 0; d < 7; ++d) acc[ai][bj][d] = v4i{0, 0, 0, 0};
            for (int kb = kc0; kb < kc1; ++kb) {
                v4i a[2][4], b[2][4];
#pragma unroll
                for (int sub = 0; sub < 2; ++sub)
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        a[sub][s] = *reinterpret_cast<const v4i *>(A + (size_t)(i0 + sub * 16 + l16) * row_bytes +
                                                                   (size_t)kb * 256 + s * 64 + lg * 16);
                        b[sub][s] = *reinterpret_cast<const v4i *>(B + (size_t)(j0 + sub * 16 + l16) * row_bytes +
                                                                   (size_t)kb * 256 + s * 64 + lg * 16);
                    }
#pragma unroll
                for (int ai = 0; ai < 2; ++ai)
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                        for (int s = 0; s < 4; ++s)
#pragma unroll
                            for (int u = 0; u < 4; ++u)
                                acc[ai][bj][s + u] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                    a[ai][s], b[bj][u], acc[ai][bj][s + u], 0, 0, 0);
            }
#pragma unroll
            for (int ai = 0; ai < 2; ++ai)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        res[ai][bj][r] = add_mod(res[ai][bj][r], mfma_combine(acc[ai][bj], r, p, mc), p);
        }
#pragma unroll
        for (int ai = 0; ai < 2; ++ai)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = i0 + ai * 16 + lg * 4 + r, j = j0 + bj * 16 + l16;
                    if (i < m && j < n) Ct[(size_t)i * n + j] = res[ai][bj][r];
                }
    }
}

// One workgroup per point; each wave computes 32x32 output tiles.
__global__ void __launch_bounds__(256) pw_mfma_kernel(const uint8_t *__restrict__ Al, const uint8_t *__restrict__ Bl,
                                                      uint32_t *__restrict__ Cmat, int m, int n, int kchunks,
                                                      int mpad, int npad, size_t a_bytes, size_t b_bytes, uint32_t p,
                                                      MfmaConst mc)
{
    const size_t t = blockIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tiles_n = npad / 32, ntiles = (mpad / 32) * tiles_n;
    for (int tile = wave; tile < ntiles; tile += 4)
        mfma_tile_global(Al + t * a_bytes, Bl + t * b_bytes, Cmat + t * (size_t)m * n, m, n, kchunks,
                         (tile / tiles_n) * 32, (tile % tiles_n) * 32, p, mc, lane & 15, lane >> 4);
}

// LDS-staged form, used for the other shapes whose A limb block fits in 64 KiB
// (m*k <= 16384 after padding): persistent workgroups, two per CU, each takes a point, copies
// its A block to LDS once, then streams B through LDS 32 columns at a time; every
// wave computes the 32x32 tiles of one row band.  16-byte chunks are stored
// XOR-swizzled within each row (chunk c of row r at c ^ (r & 15)) so the 16 rows
// a ds_read_b128 touches fall in distinct banks.
constexpr int kLdsAMax = 65536;

__device__ __forceinline__ void mfma_tile_lds(const uint4 *__restrict__ lA, const uint4 *__restrict__ lB, int cpr,
                                              int kchunks, int arow0, int l16, int lg, v4i (&acc)[2][2][7])
{
#pragma unroll
    for (int ai = 0; ai < 2; ++ai)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int d = 0; d < 7; ++d) acc[ai][bj][d] = v4i{0, 0, 0, 0};
    for (int kb = 0; kb < kchunks; ++kb) {
        v4i a[2][4], b[2][4];
#pragma unroll
        for (int sub = 0; sub < 2; ++sub)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int c = (kb * 16 + s * 4 + lg) ^ l16;
                const uint4 va = lA[(arow0 + sub * 16 + l16) * cpr + c];
                const uint4 vb = lB[(sub * 16 + l16) * cpr + c];
                a[sub][s] = v4i{(int)va.x, (int)va.y, (int)va.z, (int)va.w};
                b[sub][s] = v4i{(int)vb.x, (int)vb.y, (int)vb.z, (int)vb.w};
            }
#pragma unroll
        for (int ai = 0; ai < 2; ++ai)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        acc[ai][bj][s + u] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(a[ai][s], b[bj][u], acc[ai][bj][s + u], 0, 0, 0);
    }
}

typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) const void global_cvoid;

// Global -> LDS copy of `chunks` 16-byte chunks (a multiple of 256) with the
// row swizzle, using global_load_lds_dwordx4: each wave instruction fills 64
// consecutive LDS chunks (1 KiB at M0 + 16*lane) with no register staging, so
// every load of the copy is in flight at once.  Lane i of a wave's j-th
// instruction fills LDS chunk P = 64*(4j + wave) + i, i.e. row r = P / cpr,
// swizzled column P % cpr, which holds global chunk r*cpr + ((P % cpr) ^ (r & 15)).
__device__ __forceinline__ void copy_lds(const uint4 *__restrict__ g, uint4 *l, int chunks, int cpr, int wave,
                                         int lane)
{
    for (int j = 0; j * 256 < chunks; ++j) {
        const int base = 64 * (4 * j + wave);
        const int P = base + lane;
        const int r = P / cpr, c = P - r * cpr;
        __builtin_amdgcn_global_load_lds((global_cvoid *)(g + r * cpr + (c ^ (r & 15))), (lds_void *)(l + base), 16,
                                         0, 0);
    }
}

__global__ void __launch_bounds__(256, 2) pw_mfma_lds_kernel(const uint8_t *__restrict__ Al,
                                                             const uint8_t *__restrict__ Bl, uint32_t *__restrict__ Cmat,
                                                             int m, int n, int kchunks, int mpad, int npad,
                                                             size_t npts, uint32_t p, MfmaConst mc)
{
    extern __shared__ uint4 smem4[];
    const int cpr = kchunks * 16;  // 16-byte chunks per row
    uint4 *lA = smem4;
    uint4 *lB = smem4 + (size_t)mpad * cpr;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l16 = lane & 15, lg = lane >> 4;
    const int a_chunks = mpad * cpr, b_chunks = 32 * cpr;
    const int nrt = mpad / 32, ncb = npad / 32;
    for (size_t t = blockIdx.x; t < npts; t += gridDim.x) {
        const uint4 *gA = reinterpret_cast<const uint4 *>(Al) + t * (size_t)a_chunks;
        const uint4 *gB = reinterpret_cast<const uint4 *>(Bl) + t * (size_t)npad * cpr;
        __syncthreads();  // the previous point is done with lA and lB
        copy_lds(gA, lA, a_chunks, cpr, wave, lane);
        uint32_t *Ct = Cmat + t * (size_t)m * n;
        for (int cb = 0; cb < ncb; ++cb) {
            if (cb) __syncthreads();
            const uint4 *gBb = gB + (size_t)cb * b_chunks;
            copy_lds(gBb, lB, b_chunks, cpr, wave, lane);
            __builtin_amdgcn_s_waitcnt(0);  // this thread's LDS-direct loads have landed
            __syncthreads();
            for (int rt = wave; rt < nrt; rt += 4) {
                v4i acc[2][2][7];
                mfma_tile_lds(lA, lB, cpr, kchunks, rt * 32, l16, lg, acc);
#pragma unroll
                for (int ai = 0; ai < 2; ++ai)
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int i = rt * 32 + ai * 16 + lg * 4 + r, j = cb * 32 + bj * 16 + l16;
                            if (i < m && j < n) Ct[(size_t)i * n + j] = mfma_combine(acc[ai][bj], r, p, mc);
                        }
            }
        }
    }
}

// A in registers, B in LDS: for mpad in {32, 64, 128}, k <= 128 and a B block
// of at most 64 KiB.  Each wave multiplies one 32-row band of A, which it
// loads straight into registers (64 VGPRs at k = 128); only B, which every
// wave needs, goes through LDS, whole, with one LDS-direct copy.  A point
// therefore costs one wait and one barrier, with A and B (128 KiB for 128x128 operands) in
// flight together; two groups per CU overlap one's loads with the other's
// multiplies.  With nrt = mpad/32 < 4 bands, wave w takes band w % nrt and
// every (4/nrt)-th column block from w / nrt.
template <int KC>
__global__ void __launch_bounds__(256, 2) pw_mfma_rega_kernel(const uint8_t *__restrict__ Al,
                                                              const uint8_t *__restrict__ Bl,
                                                              uint32_t *__restrict__ Cmat, int m, int n, int mpad,
                                                              int npad, size_t npts, uint32_t p, MfmaConst mc)
{
    extern __shared__ uint4 smem4[];
    constexpr int cpr = KC * 16;  // 16-byte chunks per row
    uint4 *lB = smem4;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l16 = lane & 15, lg = lane >> 4;
    const int a_chunks = mpad * cpr, b_chunks = npad * cpr;
    const int nrt = mpad / 32, ncb = npad / 32;
    const int band = wave % nrt, cb0 = wave / nrt, cbs = 4 / nrt;
    for (size_t t = blockIdx.x; t < npts; t += gridDim.x) {
        const uint4 *gA = reinterpret_cast<const uint4 *>(Al) + t * (size_t)a_chunks;
        const uint4 *gB = reinterpret_cast<const uint4 *>(Bl) + t * (size_t)b_chunks;
        __syncthreads();  // the previous point is done with lB
        copy_lds(gB, lB, b_chunks, cpr, wave, lane);
        v4i a[2][KC][4];
#pragma unroll
        for (int sub = 0; sub < 2; ++sub)
#pragma unroll
            for (int kb = 0; kb < KC; ++kb)
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const uint4 v = gA[(band * 32 + sub * 16 + l16) * cpr + kb * 16 + s * 4 + lg];
                    a[sub][kb][s] = v4i{(int)v.x, (int)v.y, (int)v.z, (int)v.w};
                }
        __builtin_amdgcn_s_waitcnt(0);  // A in registers, this thread's B chunks in LDS
        __syncthreads();
        uint32_t *Ct = Cmat + t * (size_t)m * n;
        for (int cb = cb0; cb < ncb; cb += cbs) {
            v4i acc[2][2][7];
#pragma unroll
            for (int ai = 0; ai < 2; ++ai)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                    for (int d = 0; d < 7; ++d) acc[ai][bj][d] = v4i{0, 0, 0, 0};
#pragma unroll
            for (int kb = 0; kb < KC; ++kb) {
                v4i b[2][4];
#pragma unroll
                for (int sub = 0; sub < 2; ++sub)
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const int c = (kb * 16 + s * 4 + lg) ^ l16;
                        const uint4 vb = lB[(cb * 32 + sub * 16 + l16) * cpr + c];
                        b[sub][s] = v4i{(int)vb.x, (int)vb.y, (int)vb.z, (int)vb.w};
                    }
#pragma unroll
                for (int ai = 0; ai < 2; ++ai)
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                        for (int s = 0; s < 4; ++s)
#pragma unroll
                            for (int u = 0; u < 4; ++u)
                                acc[ai][bj][s + u] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[ai][kb][s], b[bj][u],
                                                                                          acc[ai][bj][s + u], 0, 0, 0);
            }
#pragma unroll
            for (int ai = 0; ai < 2; ++ai)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int i = band * 32 + ai * 16 + lg * 4 + r, j = cb * 32 + bj * 16 + l16;
                        if (i < m && j < n) Ct[(size_t)i * n + j] = mfma_combine(acc[ai][bj], r, p, mc);
                    }
        }
    }
}

// Plain form on point-major operands (the split transforms' centred int32
// blocks, [rpad][kpad] words per point for A and for B^T): per point and per
// output tile of (16 RI) x (16 RJ), the 256 threads stage 64-deep slices of
// the A and B^T rows in LDS (row stride 68 words, 16-byte aligned) and each
// thread accumulates RI x RJ outputs i = i0 + ti + 16 ii, j = j0 + tj + 16 jj
// in signed 64-bit registers (one v_mad_i64_i32 per product), reading four
// k-steps of its rows as one 16-byte LDS load and folding after those four
// products (plain_s_fold allows six).  Every operand word is read from HBM once
// per output tile.
constexpr int kPmKc = 64;       // k-depth of one LDS stage
constexpr int kPmStride = 68;   // LDS words per staged row

template <int RI, int RJ>
__global__ void __launch_bounds__(256, 2) pw_plain_pm_kernel(const uint32_t *__restrict__ Aw,
                                                             const uint32_t *__restrict__ Bw,
                                                             uint32_t *__restrict__ Cmat, int m, int n, int mpad,
                                                             int npad, int kpad, size_t npts, uint32_t p, PlainSConst pc)
{
    constexpr int TM = 16 * RI, TN = 16 * RJ;
    __shared__ uint4 lds4[(TM + TN) * kPmStride / 4];
    uint32_t *lA = reinterpret_cast<uint32_t *>(lds4);
    uint32_t *lB = lA + TM * kPmStride;
    const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
    const int tiles_m = (mpad + TM - 1) / TM, tiles_n = (npad + TN - 1) / TN, tiles = tiles_m * tiles_n;
    const size_t a_words = (size_t)mpad * kpad, b_words = (size_t)npad * kpad;
    for (size_t work = blockIdx.x; work < npts * (size_t)tiles; work += gridDim.x) {
        const size_t t = work / tiles;
        const int tile = (int)(work - t * tiles);
        const int i0 = (tile / tiles_n) * TM, j0 = (tile % tiles_n) * TN;
        const uint32_t *A = Aw + t * a_words, *B = Bw + t * b_words;
        int64_t acc[RI][RJ];
#pragma unroll
        for (int ii = 0; ii < RI; ++ii)
#pragma unroll
            for (int jj = 0; jj < RJ; ++jj) acc[ii][jj] = 0;
        for (int k0 = 0; k0 < kpad; k0 += kPmKc) {
            __syncthreads();  // the previous stage (or tile) is done with the LDS rows
            // 16 chunks of 16 bytes per staged row; rows past the padded operand read as zero
            for (int c = tid; c < (TM + TN) * 16; c += 256) {
                const int r = c >> 4, ch = c & 15;
                uint4 v = make_uint4(0, 0, 0, 0);
                if (r < TM) {
                    if (i0 + r < mpad)
                        v = *reinterpret_cast<const uint4 *>(A + (size_t)(i0 + r) * kpad + k0 + 4 * ch);
                    *reinterpret_cast<uint4 *>(lA + r * kPmStride + 4 * ch) = v;
                } else {
                    const int rb = r - TM;
                    if (j0 + rb < npad)
                        v = *reinterpret_cast<const uint4 *>(B + (size_t)(j0 + rb) * kpad + k0 + 4 * ch);
                    *reinterpret_cast<uint4 *>(lB + rb * kPmStride + 4 * ch) = v;
                }
            }
            __syncthreads();
#pragma unroll 1
            for (int q4 = 0; q4 < kPmKc / 4; ++q4) {
                uint4 a[RI], b[RJ];
#pragma unroll
                for (int ii = 0; ii < RI; ++ii)
                    a[ii] = *reinterpret_cast<const uint4 *>(lA + (ti + 16 * ii) * kPmStride + 4 * q4);
#pragma unroll
                for (int jj = 0; jj < RJ; ++jj)
                    b[jj] = *reinterpret_cast<const uint4 *>(lB + (tj + 16 * jj) * kPmStride + 4 * q4);
                static_assert(kPlainSFold >= 4, "four products between folds");
#pragma unroll
                for (int ii = 0; ii < RI; ++ii) {
                    const int32_t a0 = (int32_t)a[ii].x, a1 = (int32_t)a[ii].y, a2 = (int32_t)a[ii].z,
                                  a3 = (int32_t)a[ii].w;
#pragma unroll
                    for (int jj = 0; jj < RJ; ++jj) {
                        int64_t v = acc[ii][jj];
                        v += (int64_t)a0 * (int32_t)b[jj].x;
                        v += (int64_t)a1 * (int32_t)b[jj].y;
                        v += (int64_t)a2 * (int32_t)b[jj].z;
                        v += (int64_t)a3 * (int32_t)b[jj].w;
                        acc[ii][jj] = plain_s_fold(v, pc);
                    }
                }
            }
        }
        uint32_t *Ct = Cmat + t * (size_t)m * n;
#pragma unroll
        for (int ii = 0; ii < RI; ++ii)
#pragma unroll
            for (int jj = 0; jj < RJ; ++jj) {
                const int i = i0 + ti + 16 * ii, j = j0 + tj + 16 * jj;
                if (i < m && j < n) Ct[(size_t)i * n + j] = plain_s_finish(acc[ii][jj], p, pc);
            }
    }
}

// Entries e0.. x points t0.. of a point-major [t][e] array into polfirst [e][t].
__device__ __forceinline__ void point_to_entry_tile(const uint32_t *__restrict__ X, uint32_t *__restrict__ Y,
                                                    size_t entries, size_t npts, size_t e0, size_t t0, int tid,
                                                    uint32_t (&tile)[64][65])
{
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {
        const int tl = it * 4 + tid / 64, el = tid % 64;
        const size_t t = t0 + tl, e = e0 + el;
        tile[tl][el] = (t < npts && e < entries) ? X[t * entries + e] : 0u;
    }
    __syncthreads();
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {
        const int el = it * 4 + tid / 64, tl = tid % 64;
        const size_t t = t0 + tl, e = e0 + el;
        if (t < npts && e < entries) Y[e * npts + t] = tile[tl][el];
    }
}

__global__ void __launch_bounds__(256) point_to_entry_kernel(const uint32_t *__restrict__ X, uint32_t *__restrict__ Y,
                                                             size_t entries, size_t npts)
{
    __shared__ uint32_t tile[64][65];
    point_to_entry_tile(X, Y, entries, npts, (size_t)blockIdx.y * 64, (size_t)blockIdx.x * 64, threadIdx.x, tile);
}

// ---- INT8 stage of a batch: persistent work-queue kernels -------------------
// Each 256-thread group takes an item index from the counter, makes it
// wave-uniform (scalar loads of the job descriptor) and processes it.
__device__ __forceinline__ uint32_t next_item(uint32_t *counter, uint32_t &slot)
{
    if (threadIdx.x == 0) slot = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t it = __builtin_amdgcn_readfirstlane(slot);
    __syncthreads();  // everyone has the index (and is done with the previous item's LDS)
    return it;
}

__global__ void __launch_bounds__(256) batch_limbs_kernel(const BatchJob *__restrict__ jobs,
                                                          const LimbItem *__restrict__ items, uint32_t nitems,
                                                          uint32_t *counter)
{
    __shared__ uint32_t tile[64][65];
    __shared__ uint32_t slot;
    for (;;) {
        const uint32_t it = next_item(counter, slot);
        if (it >= nitems) return;
        const LimbItem w = items[it];
        const BatchJob &J = jobs[w.job];
        const bool isa = w.op == 0;
        const int rows = isa ? J.m : J.n, kdim = J.k;
        const int kchunks = (kdim + 63) / 64;
        const size_t block_bytes = limb_block_bytes_dev(rows, kdim);
        limbs_tile(global_ptr(isa ? J.ahat : J.bhat), global_ptr(isa ? J.alimb : J.blimb), rows, kdim, isa ? 1 : 0,
                   (size_t)1 << J.L, J.p,
                   block_bytes, kchunks, w.t0, (int)(w.rc % kchunks), (int)(w.rc / kchunks), threadIdx.x, tile);
    }
}

__global__ void __launch_bounds__(256) batch_mfma_kernel(const BatchJob *__restrict__ jobs,
                                                         const MfmaItem *__restrict__ items, uint32_t nitems,
                                                         uint32_t *counter)
{
    __shared__ uint32_t slot;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (;;) {
        const uint32_t it = next_item(counter, slot);
        if (it >= nitems) return;
        const MfmaItem w = items[it];
        const BatchJob &J = jobs[w.job];
        const int mpad = (J.m + 31) & ~31, npad = (J.n + 31) & ~31, kchunks = (J.k + 63) / 64;
        const int tiles_n = npad / 32, tiles = (mpad / 32) * tiles_n;
        const size_t a_bytes = limb_block_bytes_dev(J.m, J.k), b_bytes = limb_block_bytes_dev(J.n, J.k);
        const uint8_t *al = global_ptr(J.alimb), *bl = global_ptr(J.blimb);
        uint32_t *cm = global_ptr(J.cmat);
        for (uint32_t u = wave; u < w.count * (uint32_t)tiles; u += 4) {
            const size_t t = w.t0 + u / tiles;
            const int tile = (int)(u % tiles);
            mfma_tile_global(al + t * a_bytes, bl + t * b_bytes, cm + t * (size_t)J.m * J.n, J.m, J.n,
                             kchunks, (tile / tiles_n) * 32, (tile % tiles_n) * 32, J.p, J.mc, lane & 15, lane >> 4);
        }
    }
}

__global__ void __launch_bounds__(256) batch_point_to_entry_kernel(const BatchJob *__restrict__ jobs,
                                                                   const TrItem *__restrict__ items, uint32_t nitems,
                                                                   uint32_t *counter)
{
    __shared__ uint32_t tile[64][65];
    __shared__ uint32_t slot;
    for (;;) {
        const uint32_t it = next_item(counter, slot);
        if (it >= nitems) return;
        const TrItem w = items[it];
        const BatchJob &J = jobs[w.job];
        point_to_entry_tile(global_ptr(J.cmat), global_ptr(J.chat), (size_t)J.m * J.n, (size_t)1 << J.L, w.e0, w.t0,
                            threadIdx.x, tile);
    }
}

}  // namespace


hipError_t launch_batch_pointwise(const BatchJob *jobs, const PwItem *items, uint32_t nitems, uint32_t *counter,
                                  unsigned grid, hipStream_t s)
{
    hipLaunchKernelGGL(batch_pointwise_kernel, dim3(grid), dim3(256), 0, s, jobs, items, nitems, counter);
    return hipGetLastError();
}

hipError_t launch_batch_limbs(const BatchJob *jobs, const LimbItem *items, uint32_t nitems, uint32_t *counter,
                              unsigned grid, hipStream_t s)
{
    hipLaunchKernelGGL(batch_limbs_kernel, dim3(grid), dim3(256), 0, s, jobs, items, nitems, counter);
    return hipGetLastError();
}

hipError_t launch_batch_mfma(const BatchJob *jobs, const MfmaItem *items, uint32_t nitems, uint32_t *counter,
                             unsigned grid, hipStream_t s)
{
    hipLaunchKernelGGL(batch_mfma_kernel, dim3(grid), dim3(256), 0, s, jobs, items, nitems, counter);
    return hipGetLastError();
}

hipError_t launch_batch_point_to_entry(const BatchJob *jobs, const TrItem *items, uint32_t nitems,
                                       uint32_t *counter, unsigned grid, hipStream_t s)
{
    hipLaunchKernelGGL(batch_point_to_entry_kernel, dim3(grid), dim3(256), 0, s, jobs, items, nitems, counter);
    return hipGetLastError();
}

hipError_t launch_zero_counters(uint32_t *counters, int n, hipStream_t s)
{
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, counters, n);
    return hipGetLastError();
}

hipError_t launch_pointwise_plain(const uint32_t *A, const uint32_t *B, uint32_t *C, size_t m, size_t k, size_t n,
                                  size_t npts, uint32_t p, uint32_t scale, hipStream_t s)
{
    constexpr int TI = 8, TJ = 8;
    const PlainConst pc = plain_const(p, scale);
    const size_t tiles = ((m + TI - 1) / TI) * ((n + TJ - 1) / TJ);
    if (tiles > 65535) return hipErrorInvalidValue;
    dim3 grid((unsigned)((npts + 255) / 256), (unsigned)tiles);
    hipLaunchKernelGGL((pw_plain_kernel<TI, TJ>), grid, dim3(256), 0, s, A, B, C, (int)m, (int)k, (int)n, npts, p, pc);
    return hipGetLastError();
}

size_t limb_block_bytes(size_t rows, size_t kdim) { return rpad_of(rows) * kpad_of(kdim) * 4; }

hipError_t launch_to_limbs(const uint32_t *X, uint8_t *Y, size_t rows, size_t kdim, bool row_major, size_t npts,
                           uint32_t p, hipStream_t s)
{
    const size_t rpad = rpad_of(rows), kchunks = kpad_of(kdim) / 64;
    if (rpad > 65535 || kchunks > 65535) return hipErrorInvalidValue;
    dim3 grid((unsigned)((npts + 63) / 64), (unsigned)kchunks, (unsigned)rpad);
    hipLaunchKernelGGL(to_limbs_kernel, grid, dim3(256), 0, s, X, Y, (int)rows, (int)kdim, row_major ? 1 : 0, npts, p,
                       limb_block_bytes(rows, kdim), (int)kchunks);
    return hipGetLastError();
}

hipError_t launch_pointwise_mfma(const uint8_t *Al, const uint8_t *Bl, uint32_t *Cmat, size_t m, size_t k, size_t n,
                                 size_t npts, uint32_t p, uint32_t scale, hipStream_t s)
{
    const MfmaConst mc = mfma_const(p, scale);
    const size_t mpad = rpad_of(m), npad = rpad_of(n), kchunks = kpad_of(k) / 64;
    if (npts > 0x7fffffff) return hipErrorInvalidValue;
    const size_t a_bytes = limb_block_bytes(m, k);
    const int cus = device_cus();
    if ((mpad == 32 || mpad == 64 || mpad == 128) && kchunks <= 2 && npad * kchunks * 256 <= 65536) {
        const size_t lds = npad * kchunks * 256;
        const size_t grid = npts < (size_t)cus * 2 ? npts : (size_t)cus * 2;
        auto k = kchunks == 1 ? pw_mfma_rega_kernel<1> : pw_mfma_rega_kernel<2>;
        hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(256), lds, s, Al, Bl, Cmat, (int)m, (int)n, (int)mpad,
                           (int)npad, npts, p, mc);
        return hipGetLastError();
    }
    if (a_bytes <= (size_t)kLdsAMax) {
        const size_t lds = a_bytes + (size_t)32 * kchunks * 256;
        hipError_t e = hipFuncSetAttribute((const void *)pw_mfma_lds_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        const size_t grid = npts < (size_t)cus * 2 ? npts : (size_t)cus * 2;
        hipLaunchKernelGGL(pw_mfma_lds_kernel, dim3((unsigned)grid), dim3(256), lds, s, Al, Bl, Cmat, (int)m, (int)n,
                           (int)kchunks, (int)mpad, (int)npad, npts, p, mc);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(pw_mfma_kernel, dim3((unsigned)npts), dim3(256), 0, s, Al, Bl, Cmat, (int)m, (int)n,
                       (int)kchunks, (int)mpad, (int)npad, limb_block_bytes(m, k), limb_block_bytes(n, k), p, mc);
    return hipGetLastError();
}

template <int RI, int RJ>
static hipError_t pm_launch(const uint32_t *Aw, const uint32_t *Bw, uint32_t *Cmat, size_t m, size_t n, size_t mpad,
                            size_t npad, size_t kpad, size_t npts, uint32_t p, const PlainSConst &pc, hipStream_t s)
{
    const int cus = device_cus();
    const size_t tiles = ((mpad + 16 * RI - 1) / (16 * RI)) * ((npad + 16 * RJ - 1) / (16 * RJ));
    const size_t work = npts * tiles;
    const size_t grid = work < (size_t)cus * 2 ? work : (size_t)cus * 2;
    hipLaunchKernelGGL((pw_plain_pm_kernel<RI, RJ>), dim3((unsigned)grid), dim3(256), 0, s, Aw, Bw, Cmat, (int)m,
                       (int)n, (int)mpad, (int)npad, (int)kpad, npts, p, pc);
    return hipGetLastError();
}

hipError_t launch_pointwise_plain_pm(const uint32_t *Aw, const uint32_t *Bw, uint32_t *Cmat, size_t m, size_t k,
                                     size_t n, size_t npts, uint32_t p, uint32_t scale, hipStream_t s)
{
    const size_t mpad = rpad_of(m), npad = rpad_of(n), kpad = kpad_of(k);
    if (mpad > 0x7fffffff || npad > 0x7fffffff || kpad > 0x7fffffff) return hipErrorInvalidValue;
    const PlainSConst pc = plain_s_const(p, scale);
    // the widest tile each dimension fills: 16 x {2, 4, 8} rows and columns
    const int ri = mpad >= 128 ? 8 : (mpad >= 64 ? 4 : 2), rj = npad >= 128 ? 8 : (npad >= 64 ? 4 : 2);
    switch (ri * 16 + rj) {
    case 8 * 16 + 8: return pm_launch<8, 8>(Aw, Bw, Cmat, m, n, mpad, npad, kpad, npts, p, pc, s);
    case 8 * 16 + 4: return pm_launch<8, 4>(Aw, Bw, Cmat, m, n, mpad, npad, kpad, npts, p, pc, s);
    case 4 * 16 + 8: return pm_launch<4, 8>(Aw, Bw, Cmat, m, n, mpad, npad, kpad, npts, p, pc, s);
    case 4 * 16 + 4: return pm_launch<4, 4>(Aw, Bw, Cmat, m, n, mpad, npad, kpad, npts, p, pc, s);
    case 8 * 16 + 2: return pm_launch<8, 2>(Aw, Bw, Cmat, m, n, mpad, npad, kpad, npts, p, pc, s);
    case 2 * 16 + 8: return pm_launch<2, 8>(Aw, Bw, Cmat, m, n, mpad, npad, kpad, npts, p, pc, s);
    case 4 * 16 + 2: return pm_launch<4, 2>(Aw, Bw, Cmat, m, n, mpad, npad, kpad, npts, p, pc, s);
    case 2 * 16 + 4: return pm_launch<2, 4>(Aw, Bw, Cmat, m, n, mpad, npad, kpad, npts, p, pc, s);
    default: return pm_launch<2, 2>(Aw, Bw, Cmat, m, n, mpad, npad, kpad, npts, p, pc, s);
    }
}

hipError_t launch_point_to_entry(const uint32_t *X, uint32_t *Y, size_t entries, size_t npts, hipStream_t s)
{
    if (!entries || !npts) return hipSuccess;
    if ((npts + 63) / 64 > 0x7fffffff || (entries + 63) / 64 > 65535) return hipErrorInvalidValue;
    dim3 grid((unsigned)((npts + 63) / 64), (unsigned)((entries + 63) / 64));
    hipLaunchKernelGGL(point_to_entry_kernel, grid, dim3(256), 0, s, X, Y, entries, npts);
    return hipGetLastError();
}

}  // namespace lbx
