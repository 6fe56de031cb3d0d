// Modular arithmetic over Z/pZ for word-size primes p < 2^31, shared by the
// host set-up code and the gfx950 kernels.
//
// Representatives are kept canonical in [0,p) between butterflies: with
// p up to 2^31 (p = 2013265921 = 15*2^27+1 is 31 bits) a 32-bit word has only
// one bit of headroom, so the reference's lazy [0,2p)/[0,4p) bounds
// (fft-integral.inl:366-377,463-474, valid for p < 2^30) do not fit.  Every
// addition is followed by a conditional subtraction (two VALU ops).  The
// inverse transforms' twiddle products use Shoup's precomputed quotient (the
// reference's mul_precomp_b_without_reduction, fft-simd.h:123-132), whose
// [0,2p) result is brought back to [0,p) the same way; the forward ones use a
// signed Montgomery product, which takes the difference a - b of a DIF
// butterfly as it is, without adding p first.
#pragma once

#include <cstdint>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define LBX_HD __host__ __device__ __forceinline__
#else
#define LBX_HD inline
#endif

namespace lbx {

// x in [0,2p) -> x mod p.  x - p wraps above x exactly when x < p, so the
// unsigned minimum selects the reduced value (v_sub_u32 + v_min_u32).
LBX_HD uint32_t csub(uint32_t x, uint32_t p)
{
    uint32_t y = x - p;
    return y < x ? y : x;
}

LBX_HD uint32_t mulhi32(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

// Shoup: for any 32-bit x, w < p and wp = floor(w 2^32 / p),
// x*w - floor(x*wp/2^32)*p lies in [0,2p) and fits 32 bits since 2p < 2^32.
LBX_HD uint32_t shoup_lazy(uint32_t x, uint32_t w, uint32_t wp, uint32_t p)
{
    uint32_t q = mulhi32(x, wp);
    return x * w - q * p;
}

LBX_HD uint32_t shoup_mul(uint32_t x, uint32_t w, uint32_t wp, uint32_t p)
{
    return csub(shoup_lazy(x, w, wp, p), p);
}

LBX_HD uint32_t shoup_pre(uint32_t w, uint32_t p) { return (uint32_t)(((uint64_t)w << 32) / p); }

LBX_HD uint32_t add_mod(uint32_t a, uint32_t b, uint32_t p) { return csub(a + b, p); }
// a, b < p < 2^31: a - b wraps to above p exactly when a < b, and then
// a - b + p (which wraps back) is the reduced value (sub, add, min).
LBX_HD uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p)
{
    const uint32_t d = a - b, e = d + p;
    return e < d ? e : d;
}

LBX_HD int32_t mulhi_s32(int32_t a, int32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __mulhi(a, b);
#else
    return (int32_t)(((int64_t)a * b) >> 32);
#endif
}

// Signed Montgomery product x*w mod p in [0,p), for |x| < p < 2^31 and a
// twiddle w stored as wm = w 2^32 mod p and wq = wm p^-1 mod 2^32.  With
// m = x wq mod 2^32 (as a signed word), x wm and m p agree modulo 2^32, so
// hi(x wm) - hi(m p) is exactly (x wm - m p) / 2^32 = x w (mod p), and
// |x wm| < p^2, |m p| <= 2^31 p put it in (-p, p); a negative value wraps to
// above p and the minimum with r + p picks the reduced one.
LBX_HD uint32_t mont_mul(int32_t x, uint32_t wm, uint32_t wq, uint32_t p)
{
    const int32_t m = (int32_t)((uint32_t)x * wq);
    const uint32_t r = (uint32_t)(mulhi_s32(x, (int32_t)wm) - mulhi_s32(m, (int32_t)p));
    const uint32_t r2 = r + p;
    return r2 < r ? r2 : r;
}

// p^-1 mod 2^32 for odd p (Newton: each step doubles the correct low bits).
LBX_HD uint32_t inv_mod32(uint32_t p)
{
    uint32_t x = p;  // correct to 3 bits: p*p = 1 mod 8
    for (int i = 0; i < 4; ++i) x *= 2u - p * x;
    return x;
}

// The Montgomery pair {w 2^32 mod p, (w 2^32 mod p) p^-1 mod 2^32} of w < p.
LBX_HD uint32_t mont_wm(uint32_t w, uint32_t p) { return (uint32_t)(((uint64_t)w << 32) % p); }
LBX_HD uint32_t mont_wq(uint32_t wm, uint32_t pinv) { return wm * pinv; }

// Balanced signed base-256 digits of the centred representative of x:
// c = x or x-p, |c| < p/2 < 2^30, and c = sum_s d_s 2^(8s) with d_s in
// [-128,127].  Adding 0x80808080 offsets every digit by 128 without
// borrows; xor-ing 0x80 back per byte gives the two's-complement digits.
// Byte s of the result is digit d_s (s=0 least significant).
LBX_HD uint32_t to_limbs(uint32_t x, uint32_t p)
{
    uint32_t c = x > (p >> 1) ? x - p : x;
    return (c + 0x80808080u) ^ 0x80808080u;
}

// Host-side helpers.
inline uint32_t powmod_h(uint32_t a, uint64_t e, uint32_t p)
{
    uint64_t r = 1 % p, b = a % p;
    while (e) {
        if (e & 1) r = r * b % p;
        b = b * b % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

inline uint32_t invmod_h(uint32_t a, uint32_t p) { return powmod_h(a, (uint64_t)p - 2, p); }

}  // namespace lbx
