// Remainder of a multiword accumulator by a normalized multiword modulus,
// shared by the CRT kernels and a host build of the same code.
//
// The CRT recombination forms U = sum_i v_i (M_i mod P) 2^sh with v_i < 2^31
// and r <= 32 terms, so U < 2^36 Q where Q = P 2^sh.  sh = 32 off + s places P
// at the top of NL words with its top bit set (off = NL - limbs(P),
// s = leading zeros of P's top word), which makes Q a normalized divisor of
// exactly NL words.  U then has NL + 2 words, its quotient by Q is below 2^36,
// and two base-2^32 quotient digits of Knuth's algorithm D (TAOCP vol. 2,
// 4.3.1) leave U mod Q in the low NL words.  Every array index is a
// compile-time constant, so the arrays stay in registers.
#pragma once

#include <cstdint>

#include "modarith.h"

namespace lbx {

// Reciprocal of a normalized word d (top bit set): floor((2^64 - 1) / d) - 2^32.
LBX_HD uint32_t div21_inverse(uint32_t d) { return (uint32_t)(~(uint64_t)0 / d - ((uint64_t)1 << 32)); }

// Quotient and remainder of (u1 2^32 + u0) / d for a normalized d and u1 < d,
// from dinv = div21_inverse(d): Moller and Granlund, "Improved division by
// invariant integers" (IEEE Trans. Computers 60(2), 2011), Algorithm 4.  One
// 32x32 -> 64 product and two corrections, in place of a 64-by-32-bit
// division.
LBX_HD uint32_t div21(uint32_t u1, uint32_t u0, uint32_t d, uint32_t dinv, uint32_t &r)
{
    const uint64_t q = (uint64_t)dinv * u1 + (((uint64_t)u1 << 32) | u0);  // < 2^64 since u1 < d
    uint32_t q1 = (uint32_t)(q >> 32) + 1u;
    const uint32_t q0 = (uint32_t)q;
    uint32_t rr = u0 - q1 * d;
    if (rr > q0) {
        --q1;
        rr += d;
    }
    if (rr >= d) {
        ++q1;
        rr -= d;
    }
    r = rr;
    return q1;
}

template <int NL>
LBX_HD void knuth_rem2(uint32_t (&u)[NL + 2], const uint32_t (&v)[NL], uint32_t vinv)
{
    static_assert(NL >= 2, "two-word divisor at least");
    const uint32_t d = v[NL - 1];
#pragma unroll
    for (int j = 1; j >= 0; --j) {
        // estimate the digit from the top two words of the partial remainder
        // (step D3); the top word never exceeds d, and equals it only when the
        // estimate is 2^32 or more, which starts at 2^32 - 1 instead
        const uint32_t u1 = u[j + NL], u0 = u[j + NL - 1];
        uint64_t qhat, rhat;
        if (u1 >= d) {
            qhat = 0xffffffffu;
            rhat = (uint64_t)u0 + d;
        } else {
            uint32_t r32;
            qhat = div21(u1, u0, d, vinv, r32);
            rhat = r32;
        }
        while (!(rhat >> 32) && qhat * v[NL - 2] > ((rhat << 32) | u[j + NL - 2])) {
            --qhat;
            rhat += d;
        }
        // u[j .. j+NL] -= qhat v
        int64_t k = 0;
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const uint64_t pr = qhat * v[i];
            const int64_t t = (int64_t)u[i + j] - k - (int64_t)(pr & 0xffffffffu);
            u[i + j] = (uint32_t)t;
            k = (int64_t)(pr >> 32) - (t >> 32);
        }
        const int64_t t = (int64_t)u[j + NL] - k;
        u[j + NL] = (uint32_t)t;
        if (t < 0) {  // qhat was one too large: add v back
            uint64_t c = 0;
#pragma unroll
            for (int i = 0; i < NL; ++i) {
                const uint64_t s2 = (uint64_t)u[i + j] + v[i] + c;
                u[i + j] = (uint32_t)s2;
                c = s2 >> 32;
            }
            u[j + NL] += (uint32_t)c;
        }
    }
}

}  // namespace lbx
