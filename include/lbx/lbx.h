/*
 * lbx.h -- C ABI of the MI355X (gfx950) FFT polynomial-matrix multiplication.
 *
 * The reference (LinBox 1.7.1) is header-only C++ and has no C ABI; the
 * seam this library replaces is the class template
 *   PolynomialMatrixFFTMulDomain<Givaro::Modular<uint32_t,uint64_t>>::mul(c,a,b,max_rowdeg)
 *   (linbox/algorithms/polynomial-matrix/matpoly-mult-fft-wordsize.inl:57-88)
 * reached from PolynomialMatrixMulDomain::mul (polynomial-matrix-domain.h:56-75).
 * Each entry point below names the reference interface it stands for.
 *
 * Conventions
 *  - Every polynomial matrix is described by an lbx_polmat: a base pointer,
 *    one of the three PolynomialMatrix storage layouts (polynomial-matrix.h:54-77)
 *    and the view window of a SubPolynomialMatrix (polynomial-matrix.h:940-1036).
 *  - Pointers may be host or device memory (detected per call).  With device
 *    pointers the call is stream-ordered on `stream` (a hipStream_t, NULL for
 *    the default stream) and returns once the work is enqueued; with any host
 *    pointer the call copies through device buffers and returns when done.
 *  - Coefficients are uint32_t in [0,p) (Modular<uint32_t,uint64_t>::Element).
 *    The multiprecision entry point uses nlimbs little-endian uint32_t words
 *    per coefficient.
 *  - Errors are returned as lbx_status; the codes map one-to-one onto the
 *    reference's throw sites (see lbx_status).  lbx_strerror gives the text.
 *  - A context serialises the calls made on it; use one context per host
 *    thread for concurrency.  Its workspace is shared by all its calls: a
 *    call on a stream other than the previous call's stream first makes its
 *    stream wait for the previous call's work, so calls on different streams
 *    through one context are correct but do not overlap on the device.
 */
#ifndef LBX_LBX_H
#define LBX_LBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum lbx_status {
    LBX_OK = 0,
    LBX_ERR_BAD_FFT_PRIME = 1,      /* "LinBox ERROR: bad FFT Prime" (matpoly-mult-fft-wordsize-fast.inl:107-112) */
    LBX_ERR_K_TOO_LARGE = 2,        /* "FFT: k is larger that log2(p-1)" (fft-utils.h:92-93) */
    LBX_ERR_K_NOT_POSITIVE = 3,     /* "FFT: k must be positive" (fft.h:135-136) */
    LBX_ERR_NOT_PRIMITIVE_ROOT = 4, /* "w is not a primitive k-root of unity" (fft.h:140-143) */
    LBX_ERR_SIZE_MISMATCH = 5,      /* a.coldim()!=b.rowdim() (wordsize-fast.inl:65), or c too small: a view
                                       cannot be resized larger (polynomial-matrix.h:983-989) */
    LBX_ERR_NOT_ENOUGH_PRIMES = 6,  /* "not enough FFT primes" (three-primes.inl:122-126, matpoly-mult-fft.h:321-324) */
    LBX_ERR_INVALID_ARG = 7,        /* null pointer, zero dimension, unknown layout or method */
    LBX_ERR_PRIME_TOO_LARGE = 8,    /* modulus outside the supported range (three-primes.inl:57-64 analogue) */
    LBX_ERR_HIP = 9,                /* a HIP runtime call failed */
    LBX_ERR_OUT_OF_MEMORY = 10,     /* a device, pinned or host allocation failed */
    LBX_ERR_UNSUPPORTED = 11,       /* shape beyond what the library implements */
    LBX_ERR_NO_DEVICE = 12,         /* no gfx950 device at the requested index */
    LBX_ERR_WORKSPACE_TOO_SMALL = 13, /* the call needs more device memory than its fixed workspace holds */
    LBX_ERR_CHECK_FAILED = 14       /* the product check at a random point failed: the "error" report of
                                       check_mul (matpoly-mult-fft.h:105-117) under CHECK_MATPOL_MUL */
} lbx_status;

typedef enum lbx_layout {
    LBX_POLFIRST = 0,   /* element (e,t) at data[e*storage_size + t], e = i*cols+j (polynomial-matrix.h:131-137) */
    LBX_MATFIRST = 1,   /* element (e,t) at data[e + t*rows*cols] (polynomial-matrix.h:422-428) */
    LBX_MATROWFIRST = 2 /* element (i,j,t) at data[i*cols*storage_size + j + t*cols] (polynomial-matrix.h:660-683) */
} lbx_layout;

typedef enum lbx_method {
    LBX_METHOD_AUTO = 0,      /* INT8 MFMA pointwise when m,k,n >= 16 and (m*k*n >= 2^18 or at most 2^9
                                 points), else PLAIN_U64 (the faster form on gfx950); both are bit-exact */
    LBX_METHOD_PLAIN_U64 = 1, /* pointwise products on vector ALUs, 64-bit accumulation, delayed reduction */
    LBX_METHOD_INT8_MFMA = 2  /* pointwise products on INT8 MFMA over balanced signed 8-bit limbs */
} lbx_method;

/* A polynomial matrix or a SubPolynomialMatrix view of one.  Coefficient t
 * (0 <= t < size) of entry e lives at storage coefficient shift+t. */
typedef struct lbx_polmat {
    void *data;          /* storage base, host or device */
    int layout;          /* lbx_layout */
    size_t rows, cols;   /* rowdim(), coldim() */
    size_t size;         /* size() of the view: degree+1 */
    size_t storage_size; /* size() of the underlying PolynomialMatrix (>= shift+size) */
    size_t shift;        /* SubPolynomialMatrix::_shift (0 for a whole matrix) */
} lbx_polmat;

typedef struct lbx_ctx lbx_ctx;

/* Context on one device: stream-ordered workspace and cached twiddle tables. */
lbx_status lbx_init(lbx_ctx **ctx, int device);
lbx_status lbx_destroy(lbx_ctx *ctx);
const char *lbx_strerror(lbx_status status);
/* Release the cached device workspace and the pinned host buffers of host-operand
 * copies (kept between calls otherwise). */
lbx_status lbx_trim(lbx_ctx *ctx);

/* PolynomialMatrixFFTMulDomain<Modular<uint32_t,uint64_t>>::mul(c,a,b,max_rowdeg)
 * (matpoly-mult-fft-wordsize.inl:57-88), together with the d<=2 naive route of
 * PolynomialMatrixMulDomain::mul (polynomial-matrix-domain.h:56-75).
 * deg = max_rowdeg ? max_rowdeg : a.size+b.size-2; coefficients 0..deg of c are
 * written and c->size is set to deg+1 (the reference's c.resize(deg+1));
 * c->size must be >= deg+1 on entry.  Any prime p < 2^32 is accepted: FFT
 * primes with 2^ceil(log2(deg+1)) | p-1 and p < 2^31 take the single-prime
 * NTT product; every other p goes through lbx_polymat_mul_mp with one limb,
 * which is the few-primes route of PolynomialMatrixThreePrimesFFTMulDomain
 * (three-primes.inl:67-191; the reference takes it for p < 2^29 and its
 * multiprecision domain above, with the same result). */
lbx_status lbx_polymat_mul_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *c, const lbx_polmat *a,
                               const lbx_polmat *b, size_t max_rowdeg, int method, void *stream);

/* check_mul(c, a, b, deg) (matpoly-mult-fft.h:87-163) at the point x instead
 * of 1: a and b are evaluated over all their coefficients and c over its
 * first deg (c->size when deg is 0) at x mod p, and *equal is set to 1 when
 * C(x) = A(x).B(x) mod p, else to 0.  Any layout or view, host or device
 * storage, 2 <= p < 2^32.  Returns when the answer is known. */
lbx_status lbx_polymat_check_mul_u32(lbx_ctx *ctx, uint32_t p, const lbx_polmat *c, const lbx_polmat *a,
                                     const lbx_polmat *b, size_t deg, uint32_t x, int *equal, void *stream);

/* check_midproduct (matpoly-mult-fft.h:168-262), by random projections
 * instead of evaluation at 1: with random r (m words) and v (n words),
 * r^T c v against sum over q of the scalar middle products of (r^T a)_q and
 * (b v)_q, coefficients 0..min(c->size, deg-hdeg+1)-1 of c (hdeg = n0 ? n0 :
 * c->size, deg = n1 ? n1 : 2*hdeg-1) against coefficients hdeg-1.. of a.b.
 * A wrong c passes with probability at most 2/p.  *equal = 1 when they
 * agree.  Any p < 2^32; waits for the check. */
lbx_status lbx_polymat_check_midproduct_u32(lbx_ctx *ctx, uint32_t p, const lbx_polmat *c, const lbx_polmat *a,
                                           const lbx_polmat *b, size_t n0, size_t n1, int *equal, void *stream);

/* CHECK_MATPOL_MUL and CHECK_MATPOL_MIDP (polynomial-matrix-domain.h:72-74,
 * 88-90) as a run-time switch: with on != 0, every later lbx_polymat_mul_u32
 * on the context whose c receives the whole product (max_rowdeg 0 or >=
 * a.size+b.size-2) is followed by lbx_polymat_check_mul_u32 at a
 * pseudo-random point, and every lbx_polymat_midproduct_u32 and
 * lbx_polymat_midproduct_naive_u32 by lbx_polymat_check_midproduct_u32 over
 * the coefficients it computes; a mismatch returns LBX_ERR_CHECK_FAILED (c
 * keeps the computed result).  The check waits for the work it checks, so the
 * call no longer returns before it is done.  LBX_CHECK_MUL=1 in the
 * environment turns it on for every new context. */
lbx_status lbx_set_check(lbx_ctx *ctx, int on);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): lbx_polymat_mul_u32 restricted to device pointers and to
 * its single-prime FFT and constant-matrix routes, with every device buffer
 * (transform scratch, limb blocks, twiddle tables) cut from one fixed 64 MiB
 * device array per GPU, held until the process ends; the call makes no
 * device allocation.  A product needing more than 64 MiB of workspace returns
 * LBX_ERR_WORKSPACE_TOO_SMALL before any work is queued (e.g. the plain form
 * needs 4 (mk+kn+mn) 2^L bytes plus twiddles), host pointers
 * LBX_ERR_INVALID_ARG, and moduli that need the few-primes or multiprecision
 * route LBX_ERR_UNSUPPORTED.  All calls on one GPU share the array: calls
 * that overlap on the device (other streams, other contexts, other host
 * threads) race; calls queued in order on one stream are correct.  Results
 * equal lbx_polymat_mul_u32's. */
lbx_status lbx_polymat_mul_u32_noalloc(lbx_ctx *ctx, uint32_t p, lbx_polmat *c, const lbx_polmat *a,
                                       const lbx_polmat *b, size_t max_rowdeg, int method, void *stream);

/* PolynomialMatrixFFTPrimeMulDomain<F>::mul_fft(lpts, c, a, b)
 * (matpoly-mult-fft-wordsize-fast.inl:86-190): the narrower seam on padded
 * polfirst buffers of 2^lpts coefficients per entry; a and b are left intact
 * (the reference transforms them in place).  Returns LBX_ERR_BAD_FFT_PRIME
 * when 2^lpts does not divide p-1. */
lbx_status lbx_polymat_mul_fft_u32(lbx_ctx *ctx, uint32_t p, unsigned lpts, size_t m, size_t k, size_t n,
                                   const uint32_t *a, const uint32_t *b, uint32_t *c, int method,
                                   void *stream);

/* FFT<Modular<uint32_t,uint64_t>>(F, lpts, root) then FFT_direct (DIF: natural
 * order in, bit-reversed out) / FFT_inverse (DIT: bit-reversed in, natural out,
 * no scaling) on `count` contiguous polynomials of 2^lpts words, in place
 * (fft.h:57-89,133-144).  root==0 selects compute_primitive_root
 * (fft-utils.h:83-104); *root_used (may be NULL) receives the root. */
lbx_status lbx_ntt_forward_u32(lbx_ctx *ctx, uint32_t p, unsigned lpts, uint32_t root, uint32_t *data,
                               size_t count, uint32_t *root_used, void *stream);
lbx_status lbx_ntt_inverse_u32(lbx_ctx *ctx, uint32_t p, unsigned lpts, uint32_t root, uint32_t *data,
                               size_t count, uint32_t *root_used, void *stream);

/* PolynomialMatrixFFTMulDomain<Modular<integer>>::mul (matpoly-mult-fft-multiprecision.inl:611-630,
 * CRT in ZRing<integer>::mul_crtla :141-277) for a multiprecision modulus P
 * given as nlimbs (1..64) little-endian uint32 words; coefficients hold
 * nlimbs words each, in [0,P).  The integer product is computed modulo FFT
 * primes below 2^31 (RandomFFTPrime::generatePrimes order, random-fftprime.h:107-133)
 * covering 2 P^2 k min(sa,sb), recombined by CRT on the device, and reduced mod P.
 * A one-limb P takes the few-primes route instead (three-primes.inl:67-191):
 * primes covering (P-1)^2 k min(sa,sb), the operands used as they are when
 * every prime exceeds P (reduced first otherwise), and the mixed-radix
 * reconstruction mod P (:164-181); the product is the same.
 * s = max_rowdeg ? max_rowdeg+1 : a.size+b.size-1 coefficients are written
 * (mul_crtla :150-152, three-primes.inl:69-73): with s < a.size+b.size-1 the
 * cyclic product mod x^pts - 1, pts the least power of two >= s, truncated;
 * a.size, b.size <= pts.  c->size must be >= s and is set to s. */
lbx_status lbx_polymat_mul_mp(lbx_ctx *ctx, const uint32_t *P, int nlimbs, lbx_polmat *c,
                              const lbx_polmat *a, const lbx_polmat *b, size_t max_rowdeg, int method,
                              void *stream);

/* PolynomialMatrixFFTMulDomain<Givaro::ZRing<integer>>::mul(c, a, b, max_rowdeg)
 * (matpoly-mult-fft-multiprecision.inl:43-89, mul_crtla :141-277): the product
 * over the integers.  Coefficients of a and b are ab_limbs (1..32) and those of
 * c c_limbs (1..80) little-endian uint32 words each, two's complement.  As in
 * the reference, maxA = 2^(largest bit size of a's coefficients), likewise
 * maxB, and bound = 2 maxA maxB k min(sa,sb) (:78-86); the product is computed
 * modulo FFT primes below 2^31 (generatePrimes order) whose product exceeds the
 * bound and reconstructed in the symmetric range.  c_limbs words must hold
 * maxA maxB k min(sa,sb) with its sign, else LBX_ERR_SIZE_MISMATCH; more than
 * 96 primes gives LBX_ERR_NOT_ENOUGH_PRIMES.  The coefficient count s and the
 * max_rowdeg wrap are those of lbx_polymat_mul_mp.  The call reads the inputs'
 * bit sizes back before it queues the products. */
lbx_status lbx_polymat_mul_zz(lbx_ctx *ctx, lbx_polmat *c, int c_limbs, const lbx_polmat *a, const lbx_polmat *b,
                              int ab_limbs, size_t max_rowdeg, int method, void *stream);

/* One product of a batch (device pointers, polfirst storage, whole matrices). */
typedef struct lbx_job {
    uint32_t p;          /* prime of this product */
    size_t m, k, n;      /* a is m x k, b is k x n */
    size_t sa, sb;       /* a.size(), b.size(); c receives sa+sb-1 coefficients */
    const uint32_t *a;   /* polfirst, stride sa */
    const uint32_t *b;   /* polfirst, stride sb */
    uint32_t *c;         /* polfirst, stride sa+sb-1 */
} lbx_job;

/* Independent products of uneven shapes, degrees and primes, the mix that
 * OrderBasis::PM_Basis issues through PolynomialMatrixMulDomain::mul
 * (order-basis.h:219-326 -> :297).  Jobs with an FFT prime below 2^31 whose
 * transform length 2^ceil(log2(sa+sb-1)) is in [2^4, 2^13] run in the batch
 * kernels: persistent work-queue launches for the forward transforms (one per
 * length), the pointwise products (plain form, and for jobs with every
 * dimension >= 16 the INT8 MFMA form when the autotuner measured it faster)
 * and the inverse transforms, replayed from a cached hipGraph; other jobs take
 * the single-product routes after it.  Results equal one lbx_polymat_mul_u32
 * per job.  method is used for the jobs outside the batch kernels. */
lbx_status lbx_polymat_mul_batch(lbx_ctx *ctx, const lbx_job *jobs, size_t njobs, int method, void *stream);

/* No counterpart in the reference (which has no autotuner): the batch
 * scheduler's tuned choices on this context as a JSON object -- the plain
 * form's register tile per size class, and per class of INT8-capable jobs the
 * pointwise form, the points per MFMA work item and the measured times of
 * both forms (null where not tuned yet).  The string belongs to the context
 * and stays valid until the next call of this function on it. */
const char *lbx_batch_tuning(lbx_ctx *ctx);

/* Middle product, PolynomialMatrixFFTMulDomain::midproduct
 * (matpoly-mult-fft-wordsize.inl:90-127 -> matpoly-mult-fft-wordsize-fast.inl:209-316;
 * multiprecision route :93-117,427-520): hdeg = n0 ? n0 : c->size,
 * deg = n1 ? n1 : 2*hdeg-1, pts = the least power of two >= deg.  c->size
 * coefficients are written: the cyclic correlation, over pts points, of the
 * small operand (a if small_left, else b) with its first hdeg coefficients
 * reversed, and the other operand; coefficients 0..deg-hdeg of it are
 * coefficients hdeg-1..deg-1 of a.b.  c->size is not changed.  Requires
 * c->size >= deg-hdeg, small operand size <= hdeg, other operand size
 * < hdeg+deg (wordsize-fast.inl:214-218), every size <= pts (else
 * LBX_ERR_SIZE_MISMATCH); deg <= 1 gives LBX_ERR_K_NOT_POSITIVE.  Every p < 2^32:
 * single FFT prime when p < 2^31 and pts | p-1, FFT primes and CRT otherwise. */
lbx_status lbx_polymat_midproduct_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *c, const lbx_polmat *a,
                                      const lbx_polmat *b, int small_left, size_t n0, size_t n1, int method,
                                      void *stream);

/* PolynomialMatrixNaiveMulDomain::midproduct (matpoly-mult-naive.h:85-125),
 * the route of PolynomialMatrixMulDomain::midproductgen for c.size() <= 4
 * (polynomial-matrix-domain.h:95-103): c[kk-hdeg+1] = coefficient kk of a.b
 * for hdeg-1 <= kk < min(a.size+b.size-1, deg); other coefficients of c are
 * left unchanged. */
lbx_status lbx_polymat_midproduct_naive_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *c, const lbx_polmat *a,
                                            const lbx_polmat *b, size_t n0, size_t n1, void *stream);

/* OrderBasis<Modular<uint32_t,uint64_t>>::PM_Basis(sigma, serie, order, shift)
 * (linbox/algorithms/polynomial-matrix/order-basis.h:219-326): an order basis
 * sigma (m x m) of the matrix power series serie (m x n, coefficients
 * 0..order-1 used; serie->size >= order), i.e. sigma.serie = 0 mod x^order,
 * with shifted row degrees tracked in shift (m entries, in/out, the
 * reference's std::vector<size_t>&).  Below order 32 (MBASIS_THRESHOLD,
 * :47-48,232) it is M_Basis; above, the halves are joined by a middle product
 * (:274, lbx_polymat_midproduct_u32) and a product (:297, lbx_polymat_mul_u32)
 * on device buffers.  sigma->size must be >= order+1 on entry; coefficients
 * 0..*degree are written and sigma->size is set to *degree+1 (the
 * reference's sigma.resize).  Limits: m <= 128, n <= 64, m*n <= 4096, order < 2^27 (transform lengths up to 2^27),
 * shift entries < 2^48 (else LBX_ERR_UNSUPPORTED).  Any p < 2^32.  The call
 * returns when the basis is complete (the recursion reads each half's degree).
 * The discrepancy's PLUQ (:470-473) is taken in its rank-profile form: pivot
 * rows = row rank profile, placed first, every group kept in row order. */
lbx_status lbx_order_basis_pm_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *sigma, const lbx_polmat *serie,
                                  size_t order, size_t *shift, size_t *degree, int method, void *stream);
/* OrderBasis<Modular<uint32_t,uint64_t>>::M_Basis(sigma, serie, order, shift)
 * (order-basis.h:346-600) for any order < 2^15, in one device workgroup;
 * arguments and results as lbx_order_basis_pm_u32. */
lbx_status lbx_order_basis_m_u32(lbx_ctx *ctx, uint32_t p, lbx_polmat *sigma, const lbx_polmat *serie,
                                 size_t order, size_t *shift, size_t *degree, void *stream);

/* PolynomialMatrix<Field,T>::copy(M, beg, end, start) for any pair of
 * layouts and views (polynomial-matrix.h:206-243 into polfirst, :508-532 into
 * matfirst, :763-790 into matrowfirst; SubPolynomialMatrix::copy :988-992
 * writes at the view's shift, here dst->shift): for every entry e and i in
 * [beg, end], coefficient start+i-beg of dst = coefficient i of src.  Same
 * dimensions (else LBX_ERR_SIZE_MISMATCH); end < src->size and
 * start+end-beg < dst->size; the rest of dst's storage is left as it is.
 * src and dst must not share storage (LBX_ERR_INVALID_ARG when their data
 * pointers are equal).  Host or device pointers; device work is stream-ordered. */
lbx_status lbx_polymat_copy_u32(lbx_ctx *ctx, lbx_polmat *dst, const lbx_polmat *src, size_t beg, size_t end,
                                size_t start, void *stream);

/* FFT_PROFILER analogue (matpoly-mult-fft.h:43-77): with enable = 1, calls
 * record device events between their stages, labelled as in mul_fft
 * ("direct FFT_DIF", "Polfirst to Matfirst", "Pointwise mult",
 * "Matfirst to Polfirst", "inverse FFT_DIT"; wordsize-fast.inl:116-189),
 * without any host synchronisation.  enable = 2 also brackets every kernel
 * launch of the single-prime product with events on the stream it runs on,
 * labelled "kernel <name>[ a|b]" (a and b name the operand); enable = 0 stops.
 * lbx_profile_get waits for the recorded events, returns the average time (ms)
 * per label over the records made since the previous get (per call for a
 * stage, per launch for a kernel), copies up to `max` (label, ms) pairs and
 * returns their count. */
lbx_status lbx_profile_enable(lbx_ctx *ctx, int enable);
int lbx_profile_get(lbx_ctx *ctx, const char **labels, float *ms, int max);

#ifdef __cplusplus
}
#endif
#endif /* LBX_LBX_H */
