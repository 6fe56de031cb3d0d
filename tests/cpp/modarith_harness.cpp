// Host build of the word-size modular arithmetic the kernels use
// (linbox_amd/csrc/modarith.h, linbox_amd/csrc/pw_math.h) for
// tests/test_modarith_edges.py, which checks every result against Python
// integers.  Each entry point computes one operation on plain integers, or
// (lbx_test_dif) one forward transform through a twiddle table.
#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../linbox_amd/csrc/pw_math.h"

using namespace lbx;

extern "C" {

uint32_t lbx_test_csub(uint32_t x, uint32_t p) { return csub(x, p); }
uint32_t lbx_test_shoup_pre(uint32_t w, uint32_t p) { return shoup_pre(w, p); }
uint32_t lbx_test_shoup_lazy(uint32_t x, uint32_t w, uint32_t p) { return shoup_lazy(x, w, shoup_pre(w, p), p); }
uint32_t lbx_test_shoup_mul(uint32_t x, uint32_t w, uint32_t p) { return shoup_mul(x, w, shoup_pre(w, p), p); }
uint32_t lbx_test_add_mod(uint32_t a, uint32_t b, uint32_t p) { return add_mod(a, b, p); }
uint32_t lbx_test_sub_mod(uint32_t a, uint32_t b, uint32_t p) { return sub_mod(a, b, p); }
uint32_t lbx_test_mont_mul(int32_t x, uint32_t w, uint32_t p)
{
    const uint32_t wm = mont_wm(w, p);
    return mont_mul(x, wm, mont_wq(wm, inv_mod32(p)), p);
}
uint32_t lbx_test_inv_mod32(uint32_t p) { return inv_mod32(p); }
uint32_t lbx_test_to_limbs(uint32_t x, uint32_t p) { return to_limbs(x, p); }
uint32_t lbx_test_powmod(uint32_t a, uint64_t e, uint32_t p) { return powmod_h(a, e, p); }
uint32_t lbx_test_invmod(uint32_t a, uint32_t p) { return invmod_h(a, p); }

// combine_diagonals over seven int32 anti-diagonal sums.
uint32_t lbx_test_combine(const int32_t *D, uint32_t p, uint32_t scale)
{
    const MfmaConst mc = mfma_const(p, scale);
    int32_t d[7];
    for (int i = 0; i < 7; ++i) d[i] = D[i];
    return combine_diagonals(d, p, mc);
}

// The plain form's accumulation of n products a[i]*b[i] with the fold every
// pc.fold products and the final scaled reduction, as in the kernel.
uint32_t lbx_test_plain_dot(const uint32_t *a, const uint32_t *b, int n, uint32_t p, uint32_t scale)
{
    const PlainConst pc = plain_const(p, scale);
    uint64_t acc = 0;
    for (int q0 = 0; q0 < n; q0 += (int)pc.fold) {
        const int q1 = n < q0 + (int)pc.fold ? n : q0 + (int)pc.fold;
        for (int q = q0; q < q1; ++q) acc += (uint64_t)a[q] * b[q];
        acc = plain_fold(acc, pc);
    }
    return plain_finish(acc, p, pc);
}

uint32_t lbx_test_plain_fold_count(uint32_t p) { return plain_const(p, 1).fold; }

// The point-major plain form: centred operands, signed 64-bit accumulation
// folded every kPlainSFold products, as the kernel does.
uint32_t lbx_test_plain_s_dot(const uint32_t *a, const uint32_t *b, int n, uint32_t p, uint32_t scale)
{
    const PlainSConst pc = plain_s_const(p, scale);
    int64_t acc = 0;
    uint32_t cnt = 0;
    for (int q = 0; q < n; ++q) {
        acc += (int64_t)centre(a[q], p) * centre(b[q], p);
        if (++cnt == kPlainSFold) {
            cnt = 0;
            acc = plain_s_fold(acc, pc);
        }
    }
    return plain_s_finish(acc, p, pc);
}

int32_t lbx_test_centre(uint32_t x, uint32_t p) { return centre(x, p); }
}

extern "C" {

// A length-2^L DIF transform (natural order in, bit-reversed out) of x in
// place, as the forward kernels compute it: pow_w twiddle table (the subarray
// of span h at N - 2h, entry j = w^(j N/2h)), canonical butterflies, twiddle
// products in the signed Montgomery form the forward tables hold when mont is
// set, else in Shoup's form.  Returns 0, or 1 if L is out of range.
int lbx_test_dif(uint32_t *x, unsigned L, uint32_t p, uint32_t w, int mont)
{
    if (L < 1 || L > 16) return 1;
    const size_t n = (size_t)1 << L;
    std::vector<uint32_t> pw(n - 1);
    pw[0] = 1;
    for (size_t i = 1; i < n / 2; ++i) pw[i] = (uint32_t)((uint64_t)pw[i - 1] * w % p);
    size_t idx = n / 2;
    for (size_t kk = 2; kk <= n / 2; kk <<= 1)
        for (size_t i = 0; i < n / 2; i += kk, idx++) pw[idx] = pw[i];
    const uint32_t pinv = inv_mod32(p);
    for (size_t h = n / 2; h >= 1; h >>= 1) {
        for (size_t blk = 0; blk < n; blk += 2 * h)
            for (size_t j = 0; j < h; ++j) {
                const uint32_t a = x[blk + j], b = x[blk + j + h], t = pw[n - 2 * h + j];
                x[blk + j] = add_mod(a, b, p);
                if (mont) {
                    const uint32_t wm = mont_wm(t, p);
                    x[blk + j + h] = mont_mul((int32_t)(a - b), wm, mont_wq(wm, pinv), p);
                } else {
                    x[blk + j + h] = shoup_mul(a + (p - b), t, shoup_pre(t, p), p);
                }
            }
        if (h == 1) break;
    }
    return 0;
}
}  // extern "C"
