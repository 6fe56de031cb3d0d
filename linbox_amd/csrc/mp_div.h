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

template <int NL>
LBX_HD void knuth_rem2(uint32_t (&u)[NL + 2], const uint32_t (&v)[NL])
{
    static_assert(NL >= 2, "two-word divisor at least");
#pragma unroll
    for (int j = 1; j >= 0; --j) {
        // estimate the digit from the top two words of the partial remainder
        const uint64_t num = ((uint64_t)u[j + NL] << 32) | u[j + NL - 1];
        uint64_t qhat = num / v[NL - 1];
        uint64_t rhat = num - qhat * v[NL - 1];
        while (qhat >> 32 || qhat * v[NL - 2] > ((rhat << 32) | u[j + NL - 2])) {
            --qhat;
            rhat += v[NL - 1];
            if (rhat >> 32) break;
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
