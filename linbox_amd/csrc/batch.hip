// Batch scheduler: many independent products of uneven shapes, degrees and
// primes, the mix that OrderBasis::PM_Basis issues through
// PolynomialMatrixMulDomain::mul (order-basis.h:219-326 -> :297).
//
// All forward transforms, all pointwise products and all inverse transforms
// of a batch run as persistent work-queue kernels fed from device item lists:
// one transform kernel per length 2^L present (launched as parallel graph
// branches), one pointwise kernel for every job.  The whole sequence is
// captured once per batch shape into a hipGraph and replayed; a call only
// uploads the job descriptors (the caller's pointers) and launches the graph.
// An autotuner picks, on the first batch of each size class, the plain form's
// register tile (2x2, 4x4, 8x8) and, for jobs whose three dimensions are all
// at least 16, the pointwise form: the plain form, or the INT8 MFMA form
// (a limb-conversion kernel, the MFMA products with 1, 4 or 16 points per
// work item, and a transpose back to the polfirst layout the inverse
// transforms read, on a branch parallel to the plain jobs' pointwise kernel).
// Jobs the batch kernels do not cover (transform length outside [2^4, 2^13],
// no 2^L-th root of unity mod p, p >= 2^31, constant matrices) run through the
// single-product routes after the graph, on the same stream.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lbx_ctx.h"

namespace lbx {

struct BatchPlan {
    std::vector<BatchJob> jobs;            // host copy: workspace slices, twiddles, tiles, forms
    BatchJob *jobs_d = nullptr;
    NttItem *ntt_d = nullptr;              // forward items grouped by L, then inverse items grouped by L
    PwItem *pw_d = nullptr;                // plain-form jobs
    LimbItem *limb_d = nullptr;            // INT8 jobs
    MfmaItem *mfma_d = nullptr;
    TrItem *tr_d = nullptr;
    uint32_t *counters = nullptr;          // one per kernel
    void *ws = nullptr;                    // transform workspace of every job, then INT8 blocks
    std::vector<unsigned> Ls;              // lengths present
    std::vector<uint32_t> fwd_off, fwd_cnt, inv_off, inv_cnt;
    uint32_t npw = 0, nlimb = 0, nmfma = 0, ntr = 0;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
};

namespace {

int size_class(const lbx_job &j)
{
    const size_t d = std::max(j.m, j.n);
    return d <= 4 ? 0 : (d <= 16 ? 1 : 2);
}

unsigned tile_dim(uint32_t tile) { return 2u << tile; }  // 0 -> 2, 1 -> 4, 2 -> 8

unsigned log2_len(size_t deg)
{
    unsigned L = 0;
    size_t pts = 1;
    while (pts <= deg) {
        pts <<= 1;
        ++L;
    }
    return L;
}

bool batchable(const lbx_job &j, unsigned *Lout)
{
    if (!j.a || !j.b || !j.c || !j.m || !j.k || !j.n || !j.sa || !j.sb) return false;
    if (j.sa + j.sb <= 2 || j.p >= (1u << 31) || j.p < 3) return false;
    const unsigned L = log2_len(j.sa + j.sb - 2);
    if (L < (unsigned)kBatchMinL || L > (unsigned)kBatchMaxL) return false;
    if ((j.p - 1) % ((uint32_t)1 << L) != 0) return false;
    if (j.m > 0xffffu || j.k > 0xffffu || j.n > 0xffffu) return false;
    *Lout = L;
    return true;
}

int cus() { return device_cus(); }

// Class of an INT8-capable job (every dimension >= 16, one exact int32 chunk
// of inner products): 0 for a smallest dimension below 32, 1 above; -1 when
// the job stays in the plain form.
int int8_class(const lbx_job &j)
{
    const size_t d = std::min(j.m, std::min(j.k, j.n));
    if (d < 16 || j.k > 4096) return -1;
    return d < 32 ? 0 : 1;
}

size_t rpad32(size_t x) { return (x + 31) & ~(size_t)31; }
size_t limb_bytes(size_t rows, size_t kdim) { return rpad32(rows) * ((kdim + 63) & ~(size_t)63) * 4; }

void limb_items_for(const BatchJob &J, uint32_t job, std::vector<LimbItem> &out)
{
    const uint32_t npts = 1u << J.L, kchunks = (J.k + 63) / 64;
    for (uint32_t op = 0; op < 2; ++op) {
        const uint32_t rows = (uint32_t)rpad32(op == 0 ? J.m : J.n);
        for (uint32_t t0 = 0; t0 < npts; t0 += 64)
            for (uint32_t rc = 0; rc < rows * kchunks; ++rc) out.push_back(LimbItem{job, op, t0, rc});
    }
}

void mfma_items_for(const BatchJob &J, uint32_t job, uint32_t ppi, std::vector<MfmaItem> &out)
{
    const uint32_t npts = 1u << J.L;
    for (uint32_t t0 = 0; t0 < npts; t0 += ppi) out.push_back(MfmaItem{job, t0, std::min(ppi, npts - t0), 0});
}

void tr_items_for(const BatchJob &J, uint32_t job, std::vector<TrItem> &out)
{
    const uint32_t npts = 1u << J.L, entries = J.m * J.n;
    for (uint32_t t0 = 0; t0 < npts; t0 += 64)
        for (uint32_t e0 = 0; e0 < entries; e0 += 64) out.push_back(TrItem{job, e0, t0, 0});
}

void pw_items_for(const BatchJob &J, uint32_t job, std::vector<PwItem> &out)
{
    const unsigned T = tile_dim(J.tile);
    const uint32_t ti = (J.m + T - 1) / T, tj = (J.n + T - 1) / T, blocks = (((uint32_t)1 << J.L) + 255) / 256;
    for (uint32_t b = 0; b < blocks; ++b)
        for (uint32_t i = 0; i < ti; ++i)
            for (uint32_t jj = 0; jj < tj; ++jj) out.push_back(PwItem{job, i, jj, b * 256});
}

lbx_status upload(void **dst, const void *src, size_t bytes, hipStream_t s)
{
    LBX_TRY_HIP(hipMalloc(dst, bytes ? bytes : 16));
    if (bytes) LBX_TRY_HIP(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, s));
    return LBX_OK;
}

// Times the pointwise kernel of the jobs of each size class with each
// candidate tile (the workspace contents do not affect the time) and keeps
// the fastest per class.  LBX_BATCH_TILE=0|1|2 forces a tile instead.
lbx_status tune_tiles(lbx_ctx *ctx, BatchPlan *plan, const std::vector<int> &cls, hipStream_t s)
{
    const char *force = std::getenv("LBX_BATCH_TILE");
    if (force && *force >= '0' && *force <= '2') {
        for (auto &J : plan->jobs) J.tile = (uint32_t)(*force - '0');
        return LBX_OK;
    }
    static const std::vector<uint32_t> cand[3] = {{0, 1}, {0, 1, 2}, {1, 2}};
    for (int c = 0; c < 3; ++c) {
        std::vector<uint32_t> members;
        for (uint32_t j = 0; j < plan->jobs.size(); ++j)
            if (cls[j] == c) members.push_back(j);
        if (members.empty()) continue;
        if (ctx->tile_tuned[c] < 0) {
            float best = 1e30f;
            uint32_t best_tile = cand[c][0];
            for (uint32_t t : cand[c]) {
                std::vector<BatchJob> jobs = plan->jobs;
                for (uint32_t j : members) jobs[j].tile = t;
                std::vector<PwItem> items;
                for (uint32_t j : members) pw_items_for(jobs[j], j, items);
                void *jd = nullptr, *id = nullptr, *cnt = nullptr;
                lbx_status st = upload(&jd, jobs.data(), jobs.size() * sizeof(BatchJob), s);
                if (st == LBX_OK) st = upload(&id, items.data(), items.size() * sizeof(PwItem), s);
                if (st == LBX_OK) st = from_hip(hipMalloc(&cnt, 16));
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (st == LBX_OK) st = from_hip(hipEventCreate(&e0));
                if (st == LBX_OK) st = from_hip(hipEventCreate(&e1));
                const unsigned grid = (unsigned)std::min<size_t>(items.size(), (size_t)cus() * 8);
                float ms = 0;
                for (int rep = 0; rep < 2 && st == LBX_OK; ++rep) {  // first run warms up
                    st = from_hip(launch_zero_counters((uint32_t *)cnt, 1, s));
                    if (st == LBX_OK) st = from_hip(hipEventRecord(e0, s));
                    if (st == LBX_OK)
                        st = from_hip(launch_batch_pointwise((const BatchJob *)jd, (const PwItem *)id,
                                                             (uint32_t)items.size(), (uint32_t *)cnt, grid, s));
                    if (st == LBX_OK) st = from_hip(hipEventRecord(e1, s));
                    if (st == LBX_OK) st = from_hip(hipEventSynchronize(e1));
                    if (st == LBX_OK) st = from_hip(hipEventElapsedTime(&ms, e0, e1));
                }
                if (e0) (void)hipEventDestroy(e0);
                if (e1) (void)hipEventDestroy(e1);
                (void)hipStreamSynchronize(s);
                if (jd) (void)hipFree(jd);
                if (id) (void)hipFree(id);
                if (cnt) (void)hipFree(cnt);
                if (st != LBX_OK) return st;
                if (ms < best) {
                    best = ms;
                    best_tile = t;
                }
            }
            ctx->tile_tuned[c] = (int)best_tile;
        }
        for (uint32_t j : members) plan->jobs[j].tile = (uint32_t)ctx->tile_tuned[c];
    }
    return LBX_OK;
}

// Times a sequence of work-queue launches (each with its own counter) twice
// and returns the second run's time.
template <class F>
lbx_status time_launches(F launches, int nctr, hipStream_t s, float *ms)
{
    void *cnt = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    lbx_status st = from_hip(hipMalloc(&cnt, 4 * nctr + 16));
    if (st == LBX_OK) st = from_hip(hipEventCreate(&e0));
    if (st == LBX_OK) st = from_hip(hipEventCreate(&e1));
    for (int rep = 0; rep < 2 && st == LBX_OK; ++rep) {  // first run warms up
        st = from_hip(launch_zero_counters((uint32_t *)cnt, nctr, s));
        if (st == LBX_OK) st = from_hip(hipEventRecord(e0, s));
        if (st == LBX_OK) st = launches((uint32_t *)cnt);
        if (st == LBX_OK) st = from_hip(hipEventRecord(e1, s));
        if (st == LBX_OK) st = from_hip(hipEventSynchronize(e1));
        if (st == LBX_OK) st = from_hip(hipEventElapsedTime(ms, e0, e1));
    }
    (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (cnt) (void)hipFree(cnt);
    return st;
}

// For each class of INT8-capable jobs: times their pointwise stage in the plain
// form (tuned register tiles) and in the INT8 form (limb conversion, MFMA
// products with 1, 4 or 16 points per work item, transpose back) and keeps
// the fastest form and item size.  LBX_BATCH_FORM=plain|int8 forces the form.
lbx_status tune_forms(lbx_ctx *ctx, BatchPlan *plan, const std::vector<int> &icls, hipStream_t s)
{
    const char *force = std::getenv("LBX_BATCH_FORM");
    for (int c = 0; c < 2; ++c) {
        std::vector<uint32_t> members;
        for (uint32_t j = 0; j < plan->jobs.size(); ++j)
            if (icls[j] == c) members.push_back(j);
        if (members.empty()) continue;
        if (force && (!std::strcmp(force, "plain") || !std::strcmp(force, "int8"))) {
            for (uint32_t j : members) {
                plan->jobs[j].form = std::strcmp(force, "int8") == 0 ? 1u : 0u;
            }
            if (ctx->ppi_tuned[c] < 0) ctx->ppi_tuned[c] = 4;
            continue;
        }
        if (ctx->form_tuned[c] < 0) {
            std::vector<BatchJob> jobs = plan->jobs;
            void *jd = nullptr;
            lbx_status st = upload(&jd, jobs.data(), jobs.size() * sizeof(BatchJob), s);
            // plain
            std::vector<PwItem> pw;
            for (uint32_t j : members) pw_items_for(jobs[j], j, pw);
            void *pwd = nullptr;
            if (st == LBX_OK) st = upload(&pwd, pw.data(), pw.size() * sizeof(PwItem), s);
            float plain_ms = 1e30f;
            if (st == LBX_OK)
                st = time_launches([&](uint32_t *cnt) {
                    return from_hip(launch_batch_pointwise((const BatchJob *)jd, (const PwItem *)pwd,
                                                           (uint32_t)pw.size(), cnt,
                                                           (unsigned)std::min<size_t>(pw.size(), (size_t)cus() * 8),
                                                           s));
                }, 1, s, &plain_ms);
            // INT8 with each item size
            std::vector<LimbItem> li;
            std::vector<TrItem> tr;
            for (uint32_t j : members) {
                limb_items_for(jobs[j], j, li);
                tr_items_for(jobs[j], j, tr);
            }
            void *lid = nullptr, *trd = nullptr;
            if (st == LBX_OK) st = upload(&lid, li.data(), li.size() * sizeof(LimbItem), s);
            if (st == LBX_OK) st = upload(&trd, tr.data(), tr.size() * sizeof(TrItem), s);
            float best = 1e30f;
            int best_ppi = 4;
            for (uint32_t ppi : {1u, 4u, 16u}) {
                std::vector<MfmaItem> mi;
                for (uint32_t j : members) mfma_items_for(jobs[j], j, ppi, mi);
                void *mid = nullptr;
                if (st == LBX_OK) st = upload(&mid, mi.data(), mi.size() * sizeof(MfmaItem), s);
                float ms = 1e30f;
                if (st == LBX_OK)
                    st = time_launches([&](uint32_t *cnt) {
                        lbx_status r = from_hip(launch_batch_limbs(
                            (const BatchJob *)jd, (const LimbItem *)lid, (uint32_t)li.size(), cnt,
                            (unsigned)std::min<size_t>(li.size(), (size_t)cus() * 8), s));
                        if (r == LBX_OK)
                            r = from_hip(launch_batch_mfma((const BatchJob *)jd, (const MfmaItem *)mid,
                                                           (uint32_t)mi.size(), cnt + 1,
                                                           (unsigned)std::min<size_t>(mi.size(), (size_t)cus() * 4),
                                                           s));
                        if (r == LBX_OK)
                            r = from_hip(launch_batch_point_to_entry(
                                (const BatchJob *)jd, (const TrItem *)trd, (uint32_t)tr.size(), cnt + 2,
                                (unsigned)std::min<size_t>(tr.size(), (size_t)cus() * 8), s));
                        return r;
                    }, 3, s, &ms);
                if (mid) (void)hipFree(mid);
                if (ms < best) {
                    best = ms;
                    best_ppi = (int)ppi;
                }
            }
            (void)hipStreamSynchronize(s);
            for (void *x : {jd, pwd, lid, trd})
                if (x) (void)hipFree(x);
            if (st != LBX_OK) return st;
            ctx->form_tuned[c] = best < plain_ms ? 1 : 0;
            ctx->ppi_tuned[c] = best_ppi;
            ctx->form_ms[c][0] = plain_ms;
            ctx->form_ms[c][1] = best;
        }
        for (uint32_t j : members) plan->jobs[j].form = (uint32_t)ctx->form_tuned[c];
    }
    return LBX_OK;
}

// Capture: zero counters; forward kernels of every length as parallel
// branches; pointwise (plain jobs, and INT8 jobs on a parallel branch:
// limbs, MFMA, transpose); inverse kernels as parallel branches.
lbx_status capture(lbx_ctx *ctx, BatchPlan *plan)
{
    if (!ctx->aux) LBX_TRY_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
    const size_t nb = plan->Ls.size();
    std::vector<hipStream_t> br(nb, nullptr);
    std::vector<hipEvent_t> evs;
    auto new_event = [&](hipEvent_t *e) -> lbx_status {
        LBX_TRY_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
        evs.push_back(*e);
        return LBX_OK;
    };
    lbx_status st = LBX_OK;
    for (size_t i = 0; i < nb && st == LBX_OK; ++i) st = from_hip(hipStreamCreateWithFlags(&br[i], hipStreamNonBlocking));
    hipStream_t cs = ctx->aux;
    bool capturing = false;
    if (st == LBX_OK) {
        st = from_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        capturing = st == LBX_OK;
    }
    const uint32_t nctr = (uint32_t)(2 * nb + 4);
    if (st == LBX_OK) st = from_hip(launch_zero_counters(plan->counters, (int)nctr, cs));
    hipStream_t i8 = nullptr;
    if (st == LBX_OK && plan->nmfma) st = from_hip(hipStreamCreateWithFlags(&i8, hipStreamNonBlocking));
    for (int phase = 0; phase < 3 && st == LBX_OK; ++phase) {
        if (phase == 1) {
            uint32_t *ctr = plan->counters + 2 * nb;
            hipEvent_t fork = nullptr, join = nullptr;
            if (plan->nmfma) {  // INT8 jobs on a parallel branch
                if ((st = new_event(&fork)) != LBX_OK) break;
                if ((st = from_hip(hipEventRecord(fork, cs))) != LBX_OK) break;
                st = from_hip(hipStreamWaitEvent(i8, fork, 0));
                if (st == LBX_OK)
                    st = from_hip(launch_batch_limbs(plan->jobs_d, plan->limb_d, plan->nlimb, ctr + 1,
                                                     (unsigned)std::min<size_t>(plan->nlimb, (size_t)cus() * 8), i8));
                if (st == LBX_OK)
                    st = from_hip(launch_batch_mfma(plan->jobs_d, plan->mfma_d, plan->nmfma, ctr + 2,
                                                    (unsigned)std::min<size_t>(plan->nmfma, (size_t)cus() * 4), i8));
                if (st == LBX_OK)
                    st = from_hip(launch_batch_point_to_entry(plan->jobs_d, plan->tr_d, plan->ntr, ctr + 3,
                                                              (unsigned)std::min<size_t>(plan->ntr, (size_t)cus() * 8),
                                                              i8));
            }
            const unsigned grid = (unsigned)std::min<size_t>(plan->npw, (size_t)cus() * 8);
            if (st == LBX_OK && plan->npw)
                st = from_hip(launch_batch_pointwise(plan->jobs_d, plan->pw_d, plan->npw, ctr, grid, cs));
            if (st == LBX_OK && plan->nmfma) {
                if ((st = new_event(&join)) != LBX_OK) break;
                st = from_hip(hipEventRecord(join, i8));
                if (st == LBX_OK) st = from_hip(hipStreamWaitEvent(cs, join, 0));
            }
            continue;
        }
        const bool inv = phase == 2;
        hipEvent_t fork;
        if ((st = new_event(&fork)) != LBX_OK) break;
        if ((st = from_hip(hipEventRecord(fork, cs))) != LBX_OK) break;
        for (size_t i = 0; i < nb && st == LBX_OK; ++i) {
            const uint32_t off = inv ? plan->inv_off[i] : plan->fwd_off[i];
            const uint32_t cnt = inv ? plan->inv_cnt[i] : plan->fwd_cnt[i];
            st = from_hip(hipStreamWaitEvent(br[i], fork, 0));
            const unsigned grid = (unsigned)std::max<size_t>(1, std::min<size_t>(cnt, (size_t)cus() * 4));
            if (st == LBX_OK && cnt)
                st = from_hip(launch_batch_ntt(plan->Ls[i], plan->jobs_d, plan->ntt_d + off, cnt,
                                               plan->counters + (inv ? nb : 0) + i, inv, grid, br[i]));
            hipEvent_t join;
            if (st == LBX_OK) st = new_event(&join);
            if (st == LBX_OK) st = from_hip(hipEventRecord(join, br[i]));
            if (st == LBX_OK) st = from_hip(hipStreamWaitEvent(cs, join, 0));
        }
    }
    if (capturing) {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(cs, &g);
        if (st == LBX_OK) st = from_hip(e);
        if (st == LBX_OK) st = from_hip(hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0));
        plan->graph = g;
    }
    for (auto e : evs) (void)hipEventDestroy(e);
    for (auto s : br)
        if (s) (void)hipStreamDestroy(s);
    if (i8) (void)hipStreamDestroy(i8);
    return st;
}

lbx_status build_plan(lbx_ctx *ctx, const std::vector<lbx_job> &jobs, const std::vector<unsigned> &Ls,
                      hipStream_t s, BatchPlan **out)
{
    BatchPlan *plan = new BatchPlan();
    auto fail = [&](lbx_status st) {
        if (plan->jobs_d) (void)hipFree(plan->jobs_d);
        if (plan->ntt_d) (void)hipFree(plan->ntt_d);
        if (plan->pw_d) (void)hipFree(plan->pw_d);
        if (plan->limb_d) (void)hipFree(plan->limb_d);
        if (plan->mfma_d) (void)hipFree(plan->mfma_d);
        if (plan->tr_d) (void)hipFree(plan->tr_d);
        if (plan->counters) (void)hipFree(plan->counters);
        if (plan->ws) (void)hipFree(plan->ws);
        delete plan;
        return st;
    };
    // workspace slices and descriptors: polfirst transforms of every job, then
    // the limb blocks and point-major product of every INT8-capable job
    size_t words = 0;
    std::vector<size_t> off(jobs.size()), ioff(jobs.size(), 0);
    std::vector<int> icls(jobs.size());
    for (size_t j = 0; j < jobs.size(); ++j) {
        off[j] = words;
        words += (jobs[j].m * jobs[j].k + jobs[j].k * jobs[j].n + jobs[j].m * jobs[j].n) << Ls[j];
    }
    size_t bytes = words * 4;
    for (size_t j = 0; j < jobs.size(); ++j) {
        icls[j] = int8_class(jobs[j]);
        if (icls[j] < 0) continue;
        bytes = (bytes + 255) & ~(size_t)255;
        ioff[j] = bytes;
        bytes += ((limb_bytes(jobs[j].m, jobs[j].k) + limb_bytes(jobs[j].n, jobs[j].k)) << Ls[j]) +
                 ((jobs[j].m * jobs[j].n * 4) << Ls[j]);
    }
    lbx_status st = from_hip(hipMalloc(&plan->ws, bytes + 256));
    if (st != LBX_OK) return fail(st);
    std::vector<int> cls(jobs.size());
    for (size_t j = 0; j < jobs.size(); ++j) {
        const lbx_job &x = jobs[j];
        const unsigned L = Ls[j];
        BatchJob J;
        std::memset(&J, 0, sizeof(J));
        J.p = x.p;
        J.L = L;
        J.m = (uint32_t)x.m;
        J.k = (uint32_t)x.k;
        J.n = (uint32_t)x.n;
        J.sa = (uint32_t)x.sa;
        J.sb = (uint32_t)x.sb;
        J.sc = (uint32_t)(x.sa + x.sb - 1);
        J.tile = 1;
        uint32_t *w = (uint32_t *)plan->ws + off[j];
        J.ahat = w;
        J.bhat = w + ((x.m * x.k) << L);
        J.chat = w + ((x.m * x.k + x.k * x.n) << L);
        uint32_t root;
        if ((st = primitive_root(x.p, L, &root)) != LBX_OK) return fail(st);
        if ((st = twiddles(ctx, x.p, L, root, kTwMont, &J.twf)) != LBX_OK) return fail(st);
        if ((st = twiddles(ctx, x.p, L, invroot(x.p, L, root), kTwShoup, &J.twi)) != LBX_OK) return fail(st);
        const uint32_t inv_n = invmod_h((uint32_t)(((uint64_t)1 << L) % x.p), x.p);
        J.pc = plain_const(x.p, inv_n);
        J.form = 0;
        if (icls[j] >= 0) {
            uint8_t *base = (uint8_t *)plan->ws + ioff[j];
            J.alimb = base;
            J.blimb = base + (limb_bytes(x.m, x.k) << L);
            J.cmat = (uint32_t *)(J.blimb + (limb_bytes(x.n, x.k) << L));
            J.mc = mfma_const(x.p, inv_n);
        }
        plan->jobs.push_back(J);
        cls[j] = size_class(x);
    }
    if ((st = tune_tiles(ctx, plan, cls, s)) != LBX_OK) return fail(st);
    if ((st = tune_forms(ctx, plan, icls, s)) != LBX_OK) return fail(st);
    // transform items grouped by length, biggest operands first within a length
    for (unsigned L = kBatchMaxL; L >= (unsigned)kBatchMinL; --L)
        if (std::find(Ls.begin(), Ls.end(), L) != Ls.end()) plan->Ls.push_back(L);
    std::vector<NttItem> items;
    for (int inv = 0; inv < 2; ++inv) {
        for (unsigned L : plan->Ls) {
            (inv ? plan->inv_off : plan->fwd_off).push_back((uint32_t)items.size());
            const uint32_t per = batch_polys_per_item(L);
            for (uint32_t j = 0; j < jobs.size(); ++j) {
                if (Ls[j] != L) continue;
                for (uint32_t op = inv ? 2 : 0; op < (inv ? 3u : 2u); ++op) {
                    const size_t E = op == 0 ? jobs[j].m * jobs[j].k : (op == 1 ? jobs[j].k * jobs[j].n : jobs[j].m * jobs[j].n);
                    for (size_t e0 = 0; e0 < E; e0 += per)
                        items.push_back(NttItem{j, op, (uint32_t)e0, (uint32_t)std::min<size_t>(per, E - e0)});
                }
            }
            (inv ? plan->inv_cnt : plan->fwd_cnt).push_back((uint32_t)items.size() - (inv ? plan->inv_off : plan->fwd_off).back());
        }
    }
    // pointwise items, largest jobs first
    std::vector<uint32_t> order(jobs.size());
    for (uint32_t j = 0; j < order.size(); ++j) order[j] = j;
    std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        const double cx = (double)jobs[x].m * jobs[x].k * jobs[x].n * ((size_t)1 << Ls[x]);
        const double cy = (double)jobs[y].m * jobs[y].k * jobs[y].n * ((size_t)1 << Ls[y]);
        return cx > cy;
    });
    std::vector<PwItem> pw;
    std::vector<LimbItem> li;
    std::vector<MfmaItem> mi;
    std::vector<TrItem> tr;
    for (uint32_t j : order) {
        const BatchJob &J = plan->jobs[j];
        if (J.form == 1) {
            limb_items_for(J, j, li);
            mfma_items_for(J, j, (uint32_t)std::max(1, ctx->ppi_tuned[icls[j]]), mi);
            tr_items_for(J, j, tr);
        } else {
            pw_items_for(J, j, pw);
        }
    }
    plan->npw = (uint32_t)pw.size();
    plan->nlimb = (uint32_t)li.size();
    plan->nmfma = (uint32_t)mi.size();
    plan->ntr = (uint32_t)tr.size();
    void *p;
    if ((st = upload(&p, plan->jobs.data(), plan->jobs.size() * sizeof(BatchJob), s)) != LBX_OK) return fail(st);
    plan->jobs_d = (BatchJob *)p;
    if ((st = upload(&p, items.data(), items.size() * sizeof(NttItem), s)) != LBX_OK) return fail(st);
    plan->ntt_d = (NttItem *)p;
    if ((st = upload(&p, pw.data(), pw.size() * sizeof(PwItem), s)) != LBX_OK) return fail(st);
    plan->pw_d = (PwItem *)p;
    if ((st = upload(&p, li.data(), li.size() * sizeof(LimbItem), s)) != LBX_OK) return fail(st);
    plan->limb_d = (LimbItem *)p;
    if ((st = upload(&p, mi.data(), mi.size() * sizeof(MfmaItem), s)) != LBX_OK) return fail(st);
    plan->mfma_d = (MfmaItem *)p;
    if ((st = upload(&p, tr.data(), tr.size() * sizeof(TrItem), s)) != LBX_OK) return fail(st);
    plan->tr_d = (TrItem *)p;
    if ((st = from_hip(hipMalloc(&p, 4 * (2 * plan->Ls.size() + 4) + 16))) != LBX_OK) return fail(st);
    plan->counters = (uint32_t *)p;
    if ((st = from_hip(hipStreamSynchronize(s))) != LBX_OK) return fail(st);
    if ((st = capture(ctx, plan)) != LBX_OK) {
        if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
        if (plan->graph) (void)hipGraphDestroy(plan->graph);
        return fail(st);
    }
    *out = plan;
    return LBX_OK;
}

std::string signature(const std::vector<lbx_job> &jobs)
{
    std::string sig;
    sig.reserve(jobs.size() * 24);
    for (const auto &j : jobs) {
        const uint32_t v[6] = {j.p, (uint32_t)j.m, (uint32_t)j.k, (uint32_t)j.n, (uint32_t)j.sa, (uint32_t)j.sb};
        sig.append((const char *)v, sizeof(v));
    }
    return sig;
}

}  // namespace

void batch_release(lbx_ctx *ctx)
{
    for (auto &kv : ctx->batch_plans) {
        BatchPlan *p = kv.second;
        if (p->exec) (void)hipGraphExecDestroy(p->exec);
        if (p->graph) (void)hipGraphDestroy(p->graph);
        if (p->jobs_d) (void)hipFree(p->jobs_d);
        if (p->ntt_d) (void)hipFree(p->ntt_d);
        if (p->pw_d) (void)hipFree(p->pw_d);
        if (p->limb_d) (void)hipFree(p->limb_d);
        if (p->mfma_d) (void)hipFree(p->mfma_d);
        if (p->tr_d) (void)hipFree(p->tr_d);
        if (p->counters) (void)hipFree(p->counters);
        if (p->ws) (void)hipFree(p->ws);
        delete p;
    }
    ctx->batch_plans.clear();
}

// The batch kernels on jobs that are all batchable (Ls their lengths), with
// device operands; the context lock is held by the caller.
lbx_status run_batch_plan(lbx_ctx *ctx, const std::vector<lbx_job> &bj, const std::vector<unsigned> &Ls,
                          hipStream_t s)
{
    const std::string sig = signature(bj);
    BatchPlan *plan = nullptr;
    auto it = ctx->batch_plans.find(sig);
    if (it != ctx->batch_plans.end()) plan = it->second;
    if (!plan) {
        if (ctx->batch_plans.size() >= 8) {  // bounded cache
            (void)hipDeviceSynchronize();      // earlier launches of these graphs may still run
            batch_release(ctx);
        }
        LBX_TRY(build_plan(ctx, bj, Ls, s, &plan));
        ctx->batch_plans[sig] = plan;
    }
    for (size_t j = 0; j < bj.size(); ++j) {  // this call's operands
        plan->jobs[j].a = bj[j].a;
        plan->jobs[j].b = bj[j].b;
        plan->jobs[j].c = bj[j].c;
    }
    LBX_TRY_HIP(hipMemcpyAsync(plan->jobs_d, plan->jobs.data(), plan->jobs.size() * sizeof(BatchJob),
                               hipMemcpyHostToDevice, s));
    LBX_TRY_HIP(hipGraphLaunch(plan->exec, s));
    return LBX_OK;
}

lbx_status batch_products(lbx_ctx *ctx, const lbx_job *jobs, size_t njobs, hipStream_t s)
{
    std::vector<lbx_job> bj(jobs, jobs + njobs);
    std::vector<unsigned> Ls(njobs);
    for (size_t i = 0; i < njobs; ++i)
        if (!batchable(jobs[i], &Ls[i])) return LBX_ERR_UNSUPPORTED;
    return run_batch_plan(ctx, bj, Ls, s);
}

}  // namespace lbx

using namespace lbx;

extern "C" const char *lbx_batch_tuning(lbx_ctx *ctx)
{
    // no C++ exception may reach a C caller (only a failed allocation throws here)
    try {
        if (!ctx) return "{}";
        std::lock_guard<std::mutex> lock(ctx->mu);
        static const char *tiles[3] = {"2x2", "4x4", "8x8"};
        std::string js = "{\"plain_tile\": {";
        const char *cls[3] = {"max(m,n)<=4", "max(m,n)<=16", "max(m,n)>16"};
        for (int c = 0; c < 3; ++c) {
            js += std::string(c ? ", " : "") + "\"" + cls[c] + "\": ";
            js += ctx->tile_tuned[c] < 0 ? std::string("null") : std::string("\"") + tiles[ctx->tile_tuned[c]] + "\"";
        }
        js += "}, \"int8_capable\": {";
        const char *icl[2] = {"min(m,k,n) in [16,32)", "min(m,k,n)>=32"};
        for (int c = 0; c < 2; ++c) {
            js += std::string(c ? ", " : "") + "\"" + icl[c] + "\": ";
            if (ctx->form_tuned[c] < 0) {
                js += "null";
                continue;
            }
            char buf[256];
            std::snprintf(buf, sizeof buf,
                          "{\"form\": \"%s\", \"points_per_mfma_item\": %d, \"plain_ms\": %.4f, \"int8_ms\": %.4f}",
                          ctx->form_tuned[c] ? "int8_mfma" : "plain_u64", ctx->ppi_tuned[c], ctx->form_ms[c][0],
                          ctx->form_ms[c][1]);
            js += buf;
        }
        js += "}}";
        ctx->tuning_json = js;
        return ctx->tuning_json.c_str();
    } catch (...) {
        return "{}";
    }
}

extern "C" lbx_status lbx_polymat_mul_batch(lbx_ctx *ctx, const lbx_job *jobs, size_t njobs, int method, void *stream)
{
    return guarded([&]() -> lbx_status {
        if (!ctx || (!jobs && njobs)) return LBX_ERR_INVALID_ARG;
        if (method < 0 || method > 2) return LBX_ERR_INVALID_ARG;
        for (size_t i = 0; i < njobs; ++i) {
            const lbx_job &j = jobs[i];
            if (!j.a || !j.b || !j.c || !j.m || !j.k || !j.n || !j.sa || !j.sb || j.p < 2) return LBX_ERR_INVALID_ARG;
        }
        if (!njobs) return LBX_OK;
        std::lock_guard<std::mutex> lock(ctx->mu);
        // operands must be device memory; pointers already seen by the cached plan
        // of this shape are not looked up again
        auto pointers_ok = [&](const BatchPlan *plan, size_t j, const lbx_job &x) {
            if (plan && j < plan->jobs.size() && plan->jobs[j].a == x.a && plan->jobs[j].b == x.b &&
                plan->jobs[j].c == x.c)
                return true;
            return is_device_ptr(x.a) && is_device_ptr(x.b) && is_device_ptr(x.c);
        };
        DeviceGuard guard(ctx->device);
        hipStream_t s = (hipStream_t)stream;
        StreamOrder stream_order(ctx, s);
        std::vector<lbx_job> bj;
        std::vector<unsigned> Ls;
        std::vector<size_t> rest;
        for (size_t i = 0; i < njobs; ++i) {
            unsigned L;
            if (batchable(jobs[i], &L)) {
                bj.push_back(jobs[i]);
                Ls.push_back(L);
            } else {
                rest.push_back(i);
            }
        }
        if (!bj.empty()) {
            const std::string sig = signature(bj);
            auto it = ctx->batch_plans.find(sig);
            const BatchPlan *plan = it != ctx->batch_plans.end() ? it->second : nullptr;
            for (size_t j = 0; j < bj.size(); ++j)
                if (!pointers_ok(plan, j, bj[j])) return LBX_ERR_INVALID_ARG;
            LBX_TRY(run_batch_plan(ctx, bj, Ls, s));
        }
        for (size_t i : rest) {  // outside the batch kernels' range: single-product routes
            const lbx_job &j = jobs[i];
            if (!pointers_ok(nullptr, 0, j)) return LBX_ERR_INVALID_ARG;
            lbx_polmat a{(void *)j.a, LBX_POLFIRST, j.m, j.k, j.sa, j.sa, 0};
            lbx_polmat b{(void *)j.b, LBX_POLFIRST, j.k, j.n, j.sb, j.sb, 0};
            lbx_polmat c{(void *)j.c, LBX_POLFIRST, j.m, j.n, j.sa + j.sb - 1, j.sa + j.sb - 1, 0};
            const size_t deg = j.sa + j.sb - 2;
            const unsigned L = log2_len(deg);
            const bool single = j.sa + j.sb > 2 && j.p < (1u << 31) && L <= (unsigned)kMaxLog &&
                                (j.p - 1) % ((uint32_t)1 << L) == 0;
            if (single) {
                LBX_TRY(fft_prime_product(ctx, j.p, L, devview(&a, (void *)j.a), devview(&b, (void *)j.b),
                                          devview(&c, j.c), j.m, j.k, j.n, deg + 1, method, s));
            } else if (j.sa + j.sb <= 2 && j.p < (1u << 31)) {
                LBX_TRY(constant_product(ctx, j.p, devview(&a, (void *)j.a), devview(&b, (void *)j.b), devview(&c, j.c),
                                         j.m, j.k, j.n, s));
            } else {
                ctx->mu.unlock();  // the multiprecision entry point takes the lock itself
                lbx_status st = lbx_polymat_mul_mp(ctx, &j.p, 1, &c, &a, &b, 0, method, stream);
                ctx->mu.lock();
                LBX_TRY(st);
            }
        }
        return LBX_OK;
    });
}
