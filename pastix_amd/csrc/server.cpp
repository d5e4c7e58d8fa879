// Serving loop: device-resident factors, continuous batching of solves,
// refactorisations of the same pattern overlapped with solves.
//
// The server drives K slots, each a pastix instance analysed and factored on
// the same pattern (so K factors fit the HBM budget the caller sized them
// for).  Requests are taken in arrival order:
//   - a refactorisation (A + shift I) creates matrix version v+1; it runs on
//     a slot whose current version has no queued or running solve left, so
//     older solves keep their factor while the new one is built on another
//     slot's stream;
//   - a solve binds, at submission, to the newest version submitted; queued
//     solves of the version a slot holds are merged, in arrival order and up
//     to max_rhs right-hand sides, into one multi-RHS solve + batched CG
//     refinement (one CG recurrence per column, so every column's result does
//     not depend on its batch mates beyond rounding).
// One worker thread per slot issues that slot's work on its own HIP stream:
// slots overlap on the device.  No request waits for a later one, and a slot
// holding version v is freed as soon as v's solves drain, so the loop always
// makes progress.
//
// HBM-budget admission: the work blocks of a merged solve (right-hand sides,
// solutions and refinement vectors: Engine::work_bytes plus the transfer
// block) count against a budget, by default IPARM_GPU_MEMORY_PERCENTAGE of
// the device memory left free once the slots hold their factors.  A slot
// keeps its blocks between batches; a batch is admitted only while the
// blocks of all slots fit the budget, shorter if need be.  When the oldest
// request does not fit, it waits, and idle slots give their blocks back; a
// request that cannot fit even alone ends with PASTIX_ERR_OUTOFMEMORY.
//
// The scheduling (queue, versions, batching, admission, worker threads) is
// ServeCore (serve_core.cpp); this file binds its slot operations to pastix
// instances on the GPU and exports the C API.
#include "pastix_data.h"
#include "serve_core.h"

using namespace pxa_api;

namespace {

// The slots as pastix instances: merged solves through the engine's device
// refinement, refactorisations of the resident input plus a shift.
struct PastixSlotOps final : pxa::ServeOps {
    std::vector<pastix_data_t *> slots;
    int64_t n = 0;

    void bind(int k) override { (void)hipSetDevice((int)slots[k]->ext[PXA_IPARM_DEVICE]); }

    // device bytes a slot holds for a merged solve of R columns: the
    // engine's refinement blocks and the transfer block (B and X, FP64)
    int64_t work_bytes(int k, int64_t R) override {
        if (R <= 0) return 0;
        return slots[k]->engine->work_bytes((int)R) + 16 * n * R;
    }

    int refactor(int k, double shift) override { return pastix_amd_refactorize_shift(slots[k], shift); }

    // One merged solve: the requests' columns are copied side by side into
    // a device block, solved and refined there (permutations on the
    // device), and copied back into each request's b.
    int solve(int k, const std::vector<pxa::SolveCols> &batch) override {
        pastix_data_t *d = slots[k];
        int64_t R = 0;
        for (const pxa::SolveCols &r : batch) R += r.nrhs;
        int rc = PASTIX_SUCCESS;
        double *dB = device_block(d, 2 * n * R);
        if (!dB) rc = PASTIX_ERR_OUTOFMEMORY;
        double *dX = dB ? dB + (size_t)n * R : nullptr;
        const hipStream_t st = d->engine->stream();
        const size_t col = (size_t)n * sizeof(double);
        int64_t c0 = 0;
        for (const pxa::SolveCols &r : batch) {
            if (!rc && hipMemcpy2DAsync(dB + (size_t)c0 * n, col, r.b, (size_t)r.ldb * sizeof(double), col,
                                        (size_t)r.nrhs, hipMemcpyHostToDevice, st) != hipSuccess)
                rc = PASTIX_ERR_INTERNAL;
            c0 += r.nrhs;
        }
        int it = 0;
        double res = 0;
        if (!rc && d->engine->refine_user(dB, n, dX, n, (int)R, std::vector<int64_t>{0, n}, 1, true, true,
                                          (int)d->iparm[IPARM_ITERMAX], 1, refine_eps(d), &it, &res))
            rc = PASTIX_ERR_INTERNAL;
        c0 = 0;
        for (const pxa::SolveCols &r : batch) {
            if (!rc && hipMemcpy2DAsync(r.b, (size_t)r.ldb * sizeof(double), dX + (size_t)c0 * n, col, col,
                                        (size_t)r.nrhs, hipMemcpyDeviceToHost, st) != hipSuccess)
                rc = PASTIX_ERR_INTERNAL;
            c0 += r.nrhs;
        }
        if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = PASTIX_ERR_INTERNAL;
        return rc;
    }

    void release(int k) override {
        release_device_block(slots[k]);
        slots[k]->engine->release_work();
    }
};

}  // namespace

struct pastix_amd_server_s {
    PastixSlotOps ops;
    std::unique_ptr<pxa::ServeCore> core;
};

extern "C" int pastix_amd_server_create(pastix_amd_server_t **srv, pastix_data_t **slots, int nslots,
                                        pastix_int_t max_rhs) {
    if (!srv || !slots || nslots < 1 || max_rhs < 1) return PASTIX_ERR_BADPARAMETER;
    for (int k = 0; k < nslots; k++) {
        pastix_data_t *d = slots[k];
        if (!d || !d->engine || !(d->steps & STEP_NUMFACT) || !d->ext[PXA_IPARM_RESIDENT_INPUT] ||
            d->n != slots[0]->n || d->symbol.cblknbr() != slots[0]->symbol.cblknbr() ||
            d->iparm[IPARM_FACTORIZATION] == PastixFactGETRF || is_ldlt(d) || d->schurfcol < d->n)
            return PASTIX_ERR_BADPARAMETER;
    }
    auto *S = new pastix_amd_server_s();
    S->ops.slots.assign(slots, slots + nslots);
    S->ops.n = slots[0]->n;
    // the work budget: IPARM_GPU_MEMORY_PERCENTAGE of the memory left free
    size_t freeb = 0, totb = 0;
    (void)hipSetDevice((int)slots[0]->ext[PXA_IPARM_DEVICE]);
    if (hipMemGetInfo(&freeb, &totb) != hipSuccess) freeb = 0;
    const pastix_int_t pct =
        std::max<pastix_int_t>(1, std::min<pastix_int_t>(100, slots[0]->iparm[IPARM_GPU_MEMORY_PERCENTAGE]));
    const int64_t budget = (int64_t)((double)freeb * (double)pct / 100.0);
    S->core = std::make_unique<pxa::ServeCore>(&S->ops, nslots, (int64_t)max_rhs, budget, PASTIX_ERR_OUTOFMEMORY);
    *srv = S;
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_server_submit_solve(pastix_amd_server_t *S, pastix_int_t nrhs, double *b, pastix_int_t ldb,
                                              int64_t *ticket) {
    if (!S || !b || nrhs < 1 || ldb < S->ops.n || !ticket) return PASTIX_ERR_BADPARAMETER;
    *ticket = S->core->submit_solve(nrhs, b, ldb);
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_server_submit_refactor(pastix_amd_server_t *S, double shift, int64_t *ticket) {
    if (!S || !ticket) return PASTIX_ERR_BADPARAMETER;
    *ticket = S->core->submit_refactor(shift);
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_server_wait(pastix_amd_server_t *S, int64_t ticket, double *latency, int64_t *version) {
    if (!S) return PASTIX_ERR_BADPARAMETER;
    int rc = 0;
    if (!S->core->wait(ticket, &rc, latency, version)) return PASTIX_ERR_BADPARAMETER;
    return rc;
}

extern "C" int pastix_amd_server_set_budget(pastix_amd_server_t *S, int64_t bytes) {
    if (!S || bytes < 0) return PASTIX_ERR_BADPARAMETER;
    S->core->set_budget(bytes);
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_server_memory(pastix_amd_server_t *S, int64_t out[4]) {
    if (!S || !out) return PASTIX_ERR_BADPARAMETER;
    S->core->memory(out);
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_server_stats(pastix_amd_server_t *S, int64_t out[4]) {
    if (!S || !out) return PASTIX_ERR_BADPARAMETER;
    S->core->stats(out);
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_server_destroy(pastix_amd_server_t **srv) {
    if (!srv || !*srv) return PASTIX_ERR_BADPARAMETER;
    pastix_amd_server_s *S = *srv;
    S->core->shutdown();
    delete S;
    *srv = nullptr;
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_device_bytes(const pastix_data_t *d, int64_t *bytes) {
    if (!d || !d->engine || !bytes) return PASTIX_ERR_BADPARAMETER;
    *bytes = d->engine->device_bytes();
    return PASTIX_SUCCESS;
}
