// Host build of the two-digit Knuth remainder of linbox_amd/csrc/mp_div.h for
// tests/test_mp_div.py: u (NL + 2 words) is reduced in place modulo the
// normalized divisor v (NL words).
#include "../../linbox_amd/csrc/mp_div.h"

template <int NL>
static void run(uint32_t *u, const uint32_t *v)
{
    uint32_t uu[NL + 2], vv[NL];
    for (int i = 0; i < NL + 2; ++i) uu[i] = u[i];
    for (int i = 0; i < NL; ++i) vv[i] = v[i];
    lbx::knuth_rem2<NL>(uu, vv, lbx::div21_inverse(vv[NL - 1]));
    for (int i = 0; i < NL + 2; ++i) u[i] = uu[i];
}

extern "C" void lbx_test_rem2(uint32_t *u, const uint32_t *v) { run<2>(u, v); }
extern "C" void lbx_test_rem4(uint32_t *u, const uint32_t *v) { run<4>(u, v); }
extern "C" void lbx_test_rem8(uint32_t *u, const uint32_t *v) { run<8>(u, v); }
// Quotient of (u1 2^32 + u0) / d and its remainder through the reciprocal.
extern "C" uint32_t lbx_test_div21(uint32_t u1, uint32_t u0, uint32_t d, uint32_t *r)
{
    return lbx::div21(u1, u0, d, lbx::div21_inverse(d), *r);
}
extern "C" void lbx_test_rem16(uint32_t *u, const uint32_t *v) { run<16>(u, v); }
