// Scalar arithmetic of the two pointwise forms, shared by the gfx950 kernels
// and host code (so it can be checked on the host at edge values).
//
// INT8 form: a residue x in [0,p) is centred (c = x or x-p, |c| < p/2 < 2^30)
// and written as four balanced signed 8-bit digits d_s (to_limbs in
// modarith.h), c = sum_s d_s 2^(8s).  The MFMA accumulates the seven
// anti-diagonals D_d = sum_{s+t=d} A_s.B_t of a k-term inner product in int32
// (|D_d| <= 4 k 2^14, so k <= 4096 per chunk keeps |D_d| <= 2^28), and
//   sum_q A_q B_q * scale = sum_d D_d (2^(8d) scale mod p)  (mod p).
// combine_diagonals forms that sum in int64 (|.| <= 7 * 2^28 * 2^31 < 2^62),
// adds a multiple of p >= 2^62 to make it non-negative (< 2^63), and reduces
// hi*2^32 + lo with two Shoup products.
#pragma once

#include <cstdint>

#include "modarith.h"

namespace lbx {

// Constants of the MFMA epilogue.
struct MfmaConst {
    int32_t cd[7];       // 2^(8d) * scale mod p (< p < 2^31, so a signed 32-bit value)
    uint64_t off;        // multiple of p >= 2^62
    uint32_t r32, r32p;  // 2^32 mod p and its Shoup constant
    uint32_t onep;       // floor(2^32/p): Shoup constant of 1
};

// For p < 2^31 and scale < p.
inline MfmaConst mfma_const(uint32_t p, uint32_t scale)
{
    MfmaConst mc;
    uint64_t pw = scale % p;
    for (int d = 0; d < 7; ++d) {
        mc.cd[d] = (int32_t)pw;
        pw = pw * 256 % p;
    }
    const uint64_t two62 = (uint64_t)1 << 62;
    mc.off = ((two62 + p - 1) / p) * p;
    mc.r32 = (uint32_t)(((uint64_t)1 << 32) % p);
    mc.r32p = shoup_pre(mc.r32, p);
    mc.onep = shoup_pre(1, p);
    return mc;
}

// sum_d D[d] * cd[d] mod p, in [0,p), for |D[d]| <= 2^28.
LBX_HD uint32_t combine_diagonals(const int32_t (&D)[7], uint32_t p, const MfmaConst &mc)
{
    int64_t S = 0;
#pragma unroll
    // 32 x 32 -> 64-bit signed products (one v_mad_i64_i32 each)
    for (int d = 0; d < 7; ++d) S += (int64_t)D[d] * (int64_t)mc.cd[d];
    const uint64_t u = (uint64_t)S + mc.off;
    const uint32_t hi = (uint32_t)(u >> 32), lo = (uint32_t)u;
    return add_mod(shoup_mul(hi, mc.r32, mc.r32p, p), shoup_mul(lo, 1u, mc.onep, p), p);
}

// Plain (uint64 VALU) form: acc += a*b with a, b < p < 2^31, one 32x32->64
// multiply-add per product.  Every `fold` products the bits above 2^62 are
// folded back with f62 = 2^62 mod p (plain_fold): after a fold
// acc < 2^62 + 3*2^31 < 2^62 + 2^33, and `fold` is the largest n <= 64 with
// 2^62 + 2^33 + n (p-1)^2 < 2^64, so acc never wraps.  At the end
// acc * scale mod p = hi(acc)*c1 + lo(acc)*c0 with c0 = scale, c1 = 2^32 scale
// (plain_finish).
struct PlainConst {
    uint32_t c0, c0p, c1, c1p;
    uint32_t f62;   // 2^62 mod p
    uint32_t fold;  // products added between folds
};

inline PlainConst plain_const(uint32_t p, uint32_t scale)
{
    PlainConst pc;
    const uint32_t r32 = (uint32_t)(((uint64_t)1 << 32) % p);
    pc.c0 = scale % p;
    pc.c1 = (uint32_t)((uint64_t)r32 * pc.c0 % p);
    pc.c0p = shoup_pre(pc.c0, p);
    pc.c1p = shoup_pre(pc.c1, p);
    pc.f62 = (uint32_t)(((uint64_t)1 << 62) % p);
    const unsigned __int128 room = ((unsigned __int128)1 << 64) - ((uint64_t)1 << 62) - ((uint64_t)1 << 33) - 1;
    const uint64_t sq = (uint64_t)(p - 1) * (p - 1);
    const unsigned __int128 nmax = sq ? room / sq : 64;
    pc.fold = (uint32_t)(nmax > 64 ? 64 : nmax);
    return pc;
}

LBX_HD uint64_t plain_fold(uint64_t acc, const PlainConst &pc)
{
    constexpr uint64_t kLow62 = ((uint64_t)1 << 62) - 1;
    return (acc & kLow62) + (uint64_t)(uint32_t)(acc >> 62) * pc.f62;
}

LBX_HD uint32_t plain_finish(uint64_t acc, uint32_t p, const PlainConst &pc)
{
    const uint32_t al = (uint32_t)acc, ah = (uint32_t)(acc >> 32);
    return add_mod(shoup_mul(ah, pc.c1, pc.c1p, p), shoup_mul(al, pc.c0, pc.c0p, p), p);
}

// Plain form on centred operands: a, b in (-p/2, p/2] as int32, so |a b| < 2^60
// and acc += a*b is one v_mad_i64_i32.  After at most kPlainSFold = 6 products
// the bits above 2^60 are folded back (plain_s_fold: acc mod 2^60 +
// floor(acc / 2^60) (2^60 mod p), one v_mad_i64_i32 more, which leaves
// |acc| < 2^60 + 2^34, and 2^60 + 2^34 + 6 * 2^60 < 2^63); at the end
// acc + off (off a multiple of p >= 2^63, so the sum is in [0, 2^64)) is
// reduced as in plain_finish.
struct PlainSConst {
    uint32_t c0, c0p, c1, c1p;  // scale and 2^32 scale mod p, with Shoup constants
    uint32_t f60;               // 2^60 mod p
    uint64_t off;               // multiple of p, >= 2^63
};
constexpr uint32_t kPlainSFold = 6;

inline PlainSConst plain_s_const(uint32_t p, uint32_t scale)
{
    PlainSConst pc;
    const uint32_t r32 = (uint32_t)(((uint64_t)1 << 32) % p);
    pc.c0 = scale % p;
    pc.c1 = (uint32_t)((uint64_t)r32 * pc.c0 % p);
    pc.c0p = shoup_pre(pc.c0, p);
    pc.c1p = shoup_pre(pc.c1, p);
    pc.f60 = (uint32_t)(((uint64_t)1 << 60) % p);
    const uint64_t two63 = (uint64_t)1 << 63;
    pc.off = ((two63 + p - 1) / p) * p;
    return pc;
}

// x in [0,p) -> its centred representative in (-p/2, p/2].
LBX_HD int32_t centre(uint32_t x, uint32_t p) { return (int32_t)(x > (p >> 1) ? x - p : x); }

LBX_HD int64_t plain_s_fold(int64_t acc, const PlainSConst &pc)
{
    constexpr int64_t kLow60 = ((int64_t)1 << 60) - 1;
    // floor(acc / 2^60) is in [-8, 7] and f60 < 2^31: a 32 x 32 -> 64-bit signed product
    return (acc & kLow60) + (int64_t)(int32_t)(acc >> 60) * (int64_t)(int32_t)pc.f60;
}

LBX_HD uint32_t plain_s_finish(int64_t acc, uint32_t p, const PlainSConst &pc)
{
    const uint64_t u = (uint64_t)acc + pc.off;
    const uint32_t al = (uint32_t)u, ah = (uint32_t)(u >> 32);
    return add_mod(shoup_mul(ah, pc.c1, pc.c1p, p), shoup_mul(al, pc.c0, pc.c0p, p), p);
}

}  // namespace lbx
