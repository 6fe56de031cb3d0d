// Multiprecision modulus route (placeholder until the CRT kernels land).
#include "lbx_mp.h"

namespace lbx {
void lbx_mp_release(lbx_ctx *) {}
}  // namespace lbx

extern "C" lbx_status lbx_polymat_mul_mp(lbx_ctx *ctx, const uint32_t *P, int nlimbs, lbx_polmat *c,
                                         const lbx_polmat *a, const lbx_polmat *b, int method, void *stream)
{
    (void)ctx; (void)P; (void)nlimbs; (void)c; (void)a; (void)b; (void)method; (void)stream;
    return LBX_ERR_UNSUPPORTED;
}

extern "C" lbx_status lbx_polymat_mul_batch(lbx_ctx *ctx, const lbx_job *jobs, size_t njobs, int method, void *stream)
{
    (void)ctx; (void)jobs; (void)njobs; (void)method; (void)stream;
    return LBX_ERR_UNSUPPORTED;
}
