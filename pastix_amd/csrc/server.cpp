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
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "pastix_data.h"

using namespace pxa_api;

namespace {
using Clock = std::chrono::steady_clock;

struct Req {
    int64_t id = 0;
    int kind = 0;  // 0 solve, 1 refactorisation
    int64_t version = 0;
    pastix_int_t nrhs = 0;
    double *b = nullptr;
    pastix_int_t ldb = 0;
    double shift = 0;
    int state = 0;  // 0 queued, 1 running, 2 done
    int rc = 0;
    int slot = -1;
    Clock::time_point t_submit, t_done;
};
}  // namespace

struct pastix_amd_server_s {
    std::vector<pastix_data_t *> slots;
    pastix_int_t max_rhs = 64;
    int64_t n = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req *> queue;
    std::unordered_map<int64_t, std::unique_ptr<Req>> reqs;
    std::map<int64_t, int64_t> live_solves;  // version -> queued or running solves
    std::map<int64_t, int> failed;            // version -> status of its failed refactorisation
    int64_t next_id = 1, latest = 0;
    std::vector<int64_t> held;
    std::vector<char> ready;
    std::vector<std::thread> workers;
    bool stop = false;
    // HBM budget of the work blocks: bytes, and per slot the columns its
    // blocks are sized for, the bytes they hold and whether it is running
    int64_t budget = 0, peak = 0, waits = 0;
    std::vector<int64_t> cap_cols, held_bytes;
    std::vector<char> busy;
    bool pressure = false;  // a slot waits for memory: pressure_slot
    int pressure_slot = -1;
    // statistics
    int64_t batches = 0, batched_rhs = 0;
};

namespace {

// device bytes a slot holds for a merged solve of R columns: the engine's
// refinement blocks and the transfer block (B and X, FP64)
int64_t work_need(pastix_amd_server_s *S, int k, int64_t R) {
    if (R <= 0) return 0;
    return S->slots[k]->engine->work_bytes((int)R) + 16 * S->n * R;
}

int64_t held_total(pastix_amd_server_s *S) {
    int64_t t = 0;
    for (int64_t b : S->held_bytes) t += b;
    return t;
}

// One merged solve: the requests' columns are copied side by side into a
// device block, solved and refined there (permutations on the device), and
// copied back into each request's b.
void run_solves(pastix_amd_server_s *S, int k, std::vector<Req *> &batch) {
    pastix_data_t *d = S->slots[k];
    const int64_t n = S->n;
    pastix_int_t R = 0;
    for (Req *r : batch) R += r->nrhs;
    int rc = PASTIX_SUCCESS;
    double *dB = device_block(d, 2 * n * R);
    if (!dB) rc = PASTIX_ERR_OUTOFMEMORY;
    double *dX = dB ? dB + (size_t)n * R : nullptr;
    const hipStream_t st = d->engine->stream();
    const size_t col = (size_t)n * sizeof(double);
    pastix_int_t c0 = 0;
    for (Req *r : batch) {
        if (!rc && hipMemcpy2DAsync(dB + (size_t)c0 * n, col, r->b, (size_t)r->ldb * sizeof(double), col, r->nrhs,
                                    hipMemcpyHostToDevice, st) != hipSuccess)
            rc = PASTIX_ERR_INTERNAL;
        c0 += r->nrhs;
    }
    int it = 0;
    double res = 0;
    if (!rc && d->engine->refine_user(dB, n, dX, n, (int)R, std::vector<int64_t>{0, n}, 1, true, true,
                                      (int)d->iparm[IPARM_ITERMAX], 1, refine_eps(d), &it, &res))
        rc = PASTIX_ERR_INTERNAL;
    c0 = 0;
    for (Req *r : batch) {
        if (!rc && hipMemcpy2DAsync(r->b, (size_t)r->ldb * sizeof(double), dX + (size_t)c0 * n, col, col, r->nrhs,
                                    hipMemcpyDeviceToHost, st) != hipSuccess)
            rc = PASTIX_ERR_INTERNAL;
        c0 += r->nrhs;
    }
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = PASTIX_ERR_INTERNAL;
    for (Req *r : batch) r->rc = rc;
}

void worker(pastix_amd_server_s *S, int k) {
    (void)hipSetDevice((int)S->slots[k]->ext[PXA_IPARM_DEVICE]);
    std::unique_lock<std::mutex> lk(S->mu);
    while (!S->stop) {
        // 1. the oldest queued refactorisation, if this slot's version is drained
        Req *ref = nullptr;
        for (Req *r : S->queue)
            if (r->kind == 1 && r->state == 0) {
                ref = r;
                break;
            }
        const bool free_slot = S->held[k] < 0 || S->live_solves[S->held[k]] == 0;
        if (ref && free_slot) {
            ref->state = 1;
            ref->slot = k;
            S->held[k] = ref->version;
            S->ready[k] = 0;
            lk.unlock();
            const int rc = pastix_amd_refactorize_shift(S->slots[k], ref->shift);
            lk.lock();
            ref->rc = rc;
            ref->state = 2;
            ref->t_done = Clock::now();
            S->ready[k] = rc == 0;
            for (auto it = S->queue.begin(); it != S->queue.end(); ++it)
                if (*it == ref) {
                    S->queue.erase(it);
                    break;
                }
            if (rc != 0) {
                // no factor of this version exists: its solves end with the
                // refactorisation's status, later submissions bound to it
                // too, and the slot is free again
                S->failed[ref->version] = rc;
                for (auto it = S->queue.begin(); it != S->queue.end();) {
                    Req *r = *it;
                    if (r->kind == 0 && r->state == 0 && r->version == ref->version) {
                        r->rc = rc;
                        r->state = 2;
                        r->t_done = ref->t_done;
                        S->live_solves[r->version]--;
                        it = S->queue.erase(it);
                    } else {
                        ++it;
                    }
                }
                S->held[k] = -1;
            }
            S->cv.notify_all();
            continue;
        }
        // 2. a batch of queued solves of the version this slot holds, as
        // long as the slots' work blocks fit the HBM budget
        std::vector<Req *> batch;
        Req *first = nullptr;
        // while another slot waits for memory, this one takes no new batch
        const bool yield = S->pressure && S->pressure_slot != k;
        if (S->held[k] >= 0 && S->ready[k] && !yield) {
            const int64_t avail = S->budget - (held_total(S) - S->held_bytes[k]);
            pastix_int_t R = 0;
            for (Req *r : S->queue) {
                if (r->kind != 0 || r->state != 0 || r->version != S->held[k]) continue;
                if (!first) first = r;
                if (!batch.empty() && R + r->nrhs > S->max_rhs) break;
                const int64_t cols = std::max<int64_t>(S->cap_cols[k], R + r->nrhs);
                if (work_need(S, k, cols) > avail) break;
                batch.push_back(r);
                R += r->nrhs;
            }
            if (batch.empty() && first) {
                if (work_need(S, k, first->nrhs) > S->budget) {
                    // cannot fit even alone
                    first->rc = PASTIX_ERR_OUTOFMEMORY;
                    first->state = 2;
                    first->t_done = Clock::now();
                    S->live_solves[first->version]--;
                    for (auto it = S->queue.begin(); it != S->queue.end(); ++it)
                        if (*it == first) {
                            S->queue.erase(it);
                            break;
                        }
                    S->cv.notify_all();
                    continue;
                }
                // wait for the other slots' blocks
                if (!S->pressure) {
                    S->waits++;
                    S->pressure = true;
                    S->pressure_slot = k;
                    S->cv.notify_all();
                }
            }
            if (S->pressure && S->pressure_slot == k && (!first || !batch.empty())) {
                S->pressure = false;
                S->pressure_slot = -1;
                S->cv.notify_all();
            }
        }
        if (!batch.empty()) {
            pastix_int_t R = 0;
            for (Req *r : batch) {
                r->state = 1;
                r->slot = k;
                R += r->nrhs;
            }
            S->cap_cols[k] = std::max<int64_t>(S->cap_cols[k], R);
            S->held_bytes[k] = work_need(S, k, S->cap_cols[k]);
            S->peak = std::max(S->peak, held_total(S));
            S->busy[k] = 1;
            lk.unlock();
            run_solves(S, k, batch);
            lk.lock();
            S->busy[k] = 0;
            const auto now = Clock::now();
            for (Req *r : batch) {
                r->state = 2;
                r->t_done = now;
                S->live_solves[r->version]--;
                for (auto it = S->queue.begin(); it != S->queue.end(); ++it)
                    if (*it == r) {
                        S->queue.erase(it);
                        break;
                    }
                S->batched_rhs += r->nrhs;
            }
            S->batches++;
            S->cv.notify_all();
            continue;
        }
        // an idle slot gives its work blocks back when another one waits for memory
        if (S->pressure && S->pressure_slot != k && S->held_bytes[k] > 0) {
            S->busy[k] = 1;
            lk.unlock();
            release_device_block(S->slots[k]);
            S->slots[k]->engine->release_work();
            lk.lock();
            S->busy[k] = 0;
            S->held_bytes[k] = 0;
            S->cap_cols[k] = 0;
            S->cv.notify_all();
            continue;
        }
        S->cv.wait(lk);
    }
}

}  // namespace

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
    S->slots.assign(slots, slots + nslots);
    S->max_rhs = max_rhs;
    S->n = slots[0]->n;
    S->held.assign(nslots, -1);
    S->ready.assign(nslots, 0);
    S->cap_cols.assign(nslots, 0);
    S->held_bytes.assign(nslots, 0);
    S->busy.assign(nslots, 0);
    {
        // the work budget: IPARM_GPU_MEMORY_PERCENTAGE of the memory left free
        size_t freeb = 0, totb = 0;
        (void)hipSetDevice((int)slots[0]->ext[PXA_IPARM_DEVICE]);
        if (hipMemGetInfo(&freeb, &totb) != hipSuccess) freeb = 0;
        const pastix_int_t pct = std::max<pastix_int_t>(1, std::min<pastix_int_t>(100, slots[0]->iparm[IPARM_GPU_MEMORY_PERCENTAGE]));
        S->budget = (int64_t)((double)freeb * (double)pct / 100.0);
    }
    S->held[0] = 0;  // slot 0 serves the matrix it was factored on
    S->ready[0] = 1;
    for (int k = 0; k < nslots; k++) S->workers.emplace_back(worker, S, k);
    *srv = S;
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_server_submit_solve(pastix_amd_server_t *S, pastix_int_t nrhs, double *b, pastix_int_t ldb,
                                              int64_t *ticket) {
    if (!S || !b || nrhs < 1 || ldb < S->n || !ticket) return PASTIX_ERR_BADPARAMETER;
    auto r = std::make_unique<Req>();
    r->kind = 0;
    r->nrhs = nrhs;
    r->b = b;
    r->ldb = ldb;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(S->mu);
    r->id = S->next_id++;
    r->version = S->latest;
    *ticket = r->id;
    const auto f = S->failed.find(r->version);
    if (f != S->failed.end()) {
        // the newest version failed to factor: the request ends at once
        r->rc = f->second;
        r->state = 2;
        r->t_done = Clock::now();
        S->reqs.emplace(r->id, std::move(r));
        S->cv.notify_all();
        return PASTIX_SUCCESS;
    }
    S->live_solves[r->version]++;
    S->queue.push_back(r.get());
    S->reqs.emplace(r->id, std::move(r));
    S->cv.notify_all();
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_server_submit_refactor(pastix_amd_server_t *S, double shift, int64_t *ticket) {
    if (!S || !ticket) return PASTIX_ERR_BADPARAMETER;
    auto r = std::make_unique<Req>();
    r->kind = 1;
    r->shift = shift;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(S->mu);
    r->id = S->next_id++;
    r->version = ++S->latest;
    *ticket = r->id;
    S->queue.push_back(r.get());
    S->reqs.emplace(r->id, std::move(r));
    S->cv.notify_all();
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_server_wait(pastix_amd_server_t *S, int64_t ticket, double *latency, int64_t *version) {
    if (!S) return PASTIX_ERR_BADPARAMETER;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return PASTIX_ERR_BADPARAMETER;
    Req *r = it->second.get();
    S->cv.wait(lk, [&] { return r->state == 2; });
    if (latency) *latency = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    if (version) *version = r->version;
    const int rc = r->rc;
    S->reqs.erase(it);
    return rc;
}

extern "C" int pastix_amd_server_set_budget(pastix_amd_server_t *S, int64_t bytes) {
    if (!S || bytes < 0) return PASTIX_ERR_BADPARAMETER;
    std::lock_guard<std::mutex> g(S->mu);
    S->budget = bytes;
    S->cv.notify_all();
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_server_memory(pastix_amd_server_t *S, int64_t out[4]) {
    if (!S || !out) return PASTIX_ERR_BADPARAMETER;
    std::lock_guard<std::mutex> g(S->mu);
    out[0] = S->budget;
    out[1] = held_total(S);
    out[2] = S->peak;
    out[3] = S->waits;
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_server_stats(pastix_amd_server_t *S, int64_t out[4]) {
    if (!S || !out) return PASTIX_ERR_BADPARAMETER;
    std::lock_guard<std::mutex> g(S->mu);
    out[0] = S->batches;
    out[1] = S->batched_rhs;
    out[2] = S->latest;
    out[3] = (int64_t)S->queue.size();
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_server_destroy(pastix_amd_server_t **srv) {
    if (!srv || !*srv) return PASTIX_ERR_BADPARAMETER;
    pastix_amd_server_s *S = *srv;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        // finish what was admitted
        S->cv.wait(lk, [&] { return S->queue.empty(); });
        S->stop = true;
        S->cv.notify_all();
    }
    for (std::thread &t : S->workers) t.join();
    delete S;
    *srv = nullptr;
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_device_bytes(const pastix_data_t *d, int64_t *bytes) {
    if (!d || !d->engine || !bytes) return PASTIX_ERR_BADPARAMETER;
    *bytes = d->engine->device_bytes();
    return PASTIX_SUCCESS;
}
