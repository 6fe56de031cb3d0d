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
    lbx::knuth_rem2<NL>(uu, vv);
    for (int i = 0; i < NL + 2; ++i) u[i] = uu[i];
}

extern "C" void lbx_test_rem8(uint32_t *u, const uint32_t *v) { run<8>(u, v); }
extern "C" void lbx_test_rem16(uint32_t *u, const uint32_t *v) { run<16>(u, v); }
