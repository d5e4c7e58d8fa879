// Scheduling core of the serving loop; see serve_core.h and the overview
// at the top of server.cpp.
#include "serve_core.h"

#include <algorithm>

namespace pxa {

ServeCore::ServeCore(ServeOps *ops, int nslots, int64_t max_rhs, int64_t budget, int err_oom)
    : ops_(ops), max_rhs_(max_rhs), err_oom_(err_oom), budget_(budget) {
    held_.assign(nslots, -1);
    ready_.assign(nslots, 0);
    cap_cols_.assign(nslots, 0);
    held_bytes_.assign(nslots, 0);
    busy_.assign(nslots, 0);
    held_[0] = 0;  // slot 0 serves the matrix it was factored on
    ready_[0] = 1;
    for (int k = 0; k < nslots; k++) workers_.emplace_back(&ServeCore::worker, this, k);
}

ServeCore::~ServeCore() { shutdown(); }

int64_t ServeCore::held_total() const {
    int64_t t = 0;
    for (int64_t b : held_bytes_) t += b;
    return t;
}

void ServeCore::unqueue(Req *r) {
    for (auto it = queue_.begin(); it != queue_.end(); ++it)
        if (*it == r) {
            queue_.erase(it);
            return;
        }
}

void ServeCore::worker(int k) {
    ops_->bind(k);
    std::unique_lock<std::mutex> lk(mu_);
    while (!stop_) {
        // 1. the oldest queued refactorisation, if this slot's version is drained
        Req *ref = nullptr;
        for (Req *r : queue_)
            if (r->kind == 1 && r->state == 0) {
                ref = r;
                break;
            }
        const bool free_slot = held_[k] < 0 || live_solves_[held_[k]] == 0;
        if (ref && free_slot) {
            ref->state = 1;
            ref->slot = k;
            held_[k] = ref->version;
            ready_[k] = 0;
            const double shift = ref->shift;
            lk.unlock();
            const int rc = ops_->refactor(k, shift);
            lk.lock();
            ref->rc = rc;
            ref->state = 2;
            ref->t_done = Clock::now();
            ready_[k] = rc == 0;
            unqueue(ref);
            if (rc != 0) {
                // no factor of this version exists: its solves end with the
                // refactorisation's status, later submissions bound to it
                // too, and the slot is free again
                failed_[ref->version] = rc;
                for (auto it = queue_.begin(); it != queue_.end();) {
                    Req *r = *it;
                    if (r->kind == 0 && r->state == 0 && r->version == ref->version) {
                        r->rc = rc;
                        r->state = 2;
                        r->t_done = ref->t_done;
                        live_solves_[r->version]--;
                        it = queue_.erase(it);
                    } else {
                        ++it;
                    }
                }
                held_[k] = -1;
            }
            cv_.notify_all();
            continue;
        }
        // 2. a batch of queued solves of the version this slot holds, as
        // long as the slots' work blocks fit the budget
        std::vector<Req *> batch;
        Req *first = nullptr;
        // while another slot waits for memory, this one takes no new batch
        const bool yield = pressure_ && pressure_slot_ != k;
        if (held_[k] >= 0 && ready_[k] && !yield) {
            const int64_t avail = budget_ - (held_total() - held_bytes_[k]);
            int64_t R = 0;
            for (Req *r : queue_) {
                if (r->kind != 0 || r->state != 0 || r->version != held_[k]) continue;
                if (!first) first = r;
                if (!batch.empty() && R + r->nrhs > max_rhs_) break;
                const int64_t cols = std::max<int64_t>(cap_cols_[k], R + r->nrhs);
                if (ops_->work_bytes(k, cols) > avail) break;
                batch.push_back(r);
                R += r->nrhs;
            }
            if (batch.empty() && first) {
                if (ops_->work_bytes(k, first->nrhs) > budget_) {
                    // cannot fit even alone
                    first->rc = err_oom_;
                    first->state = 2;
                    first->t_done = Clock::now();
                    live_solves_[first->version]--;
                    unqueue(first);
                    cv_.notify_all();
                    continue;
                }
                // wait for the other slots' blocks
                if (!pressure_) {
                    waits_++;
                    pressure_ = true;
                    pressure_slot_ = k;
                    cv_.notify_all();
                }
            }
            if (pressure_ && pressure_slot_ == k && (!first || !batch.empty())) {
                pressure_ = false;
                pressure_slot_ = -1;
                cv_.notify_all();
            }
        }
        if (!batch.empty()) {
            int64_t R = 0;
            std::vector<SolveCols> cols;
            for (Req *r : batch) {
                r->state = 1;
                r->slot = k;
                R += r->nrhs;
                cols.push_back(SolveCols{r->b, r->ldb, r->nrhs});
            }
            cap_cols_[k] = std::max<int64_t>(cap_cols_[k], R);
            held_bytes_[k] = ops_->work_bytes(k, cap_cols_[k]);
            peak_ = std::max(peak_, held_total());
            busy_[k] = 1;
            lk.unlock();
            const int rc = ops_->solve(k, cols);
            lk.lock();
            busy_[k] = 0;
            const auto now = Clock::now();
            for (Req *r : batch) {
                r->rc = rc;
                r->state = 2;
                r->t_done = now;
                live_solves_[r->version]--;
                unqueue(r);
                batched_rhs_ += r->nrhs;
            }
            batches_++;
            cv_.notify_all();
            continue;
        }
        // an idle slot gives its work blocks back when another one waits for memory
        if (pressure_ && pressure_slot_ != k && held_bytes_[k] > 0) {
            busy_[k] = 1;
            lk.unlock();
            ops_->release(k);
            lk.lock();
            busy_[k] = 0;
            held_bytes_[k] = 0;
            cap_cols_[k] = 0;
            cv_.notify_all();
            continue;
        }
        cv_.wait(lk);
    }
}

int64_t ServeCore::submit_solve(int64_t nrhs, double *b, int64_t ldb) {
    auto r = std::make_unique<Req>();
    r->kind = 0;
    r->nrhs = nrhs;
    r->b = b;
    r->ldb = ldb;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(mu_);
    r->id = next_id_++;
    r->version = latest_;
    const int64_t id = r->id;
    const auto f = failed_.find(r->version);
    if (f != failed_.end()) {
        // the newest version failed to factor: the request ends at once
        r->rc = f->second;
        r->state = 2;
        r->t_done = Clock::now();
        reqs_.emplace(id, std::move(r));
        cv_.notify_all();
        return id;
    }
    live_solves_[r->version]++;
    queue_.push_back(r.get());
    reqs_.emplace(id, std::move(r));
    cv_.notify_all();
    return id;
}

int64_t ServeCore::submit_refactor(double shift) {
    auto r = std::make_unique<Req>();
    r->kind = 1;
    r->shift = shift;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(mu_);
    r->id = next_id_++;
    r->version = ++latest_;
    const int64_t id = r->id;
    queue_.push_back(r.get());
    reqs_.emplace(id, std::move(r));
    cv_.notify_all();
    return id;
}

bool ServeCore::wait(int64_t ticket, int *rc, double *latency, int64_t *version) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return false;
    Req *r = it->second.get();
    cv_.wait(lk, [&] { return r->state == 2; });
    if (latency) *latency = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    if (version) *version = r->version;
    if (rc) *rc = r->rc;
    reqs_.erase(it);
    return true;
}

void ServeCore::set_budget(int64_t bytes) {
    std::lock_guard<std::mutex> g(mu_);
    budget_ = bytes;
    cv_.notify_all();
}

void ServeCore::memory(int64_t out[4]) {
    std::lock_guard<std::mutex> g(mu_);
    out[0] = budget_;
    out[1] = held_total();
    out[2] = peak_;
    out[3] = waits_;
}

void ServeCore::stats(int64_t out[4]) {
    std::lock_guard<std::mutex> g(mu_);
    out[0] = batches_;
    out[1] = batched_rhs_;
    out[2] = latest_;
    out[3] = (int64_t)queue_.size();
}

void ServeCore::shutdown() {
    {
        std::unique_lock<std::mutex> lk(mu_);
        if (stop_) return;
        // finish what was admitted
        cv_.wait(lk, [&] { return queue_.empty(); });
        stop_ = true;
        cv_.notify_all();
    }
    for (std::thread &t : workers_) t.join();
    workers_.clear();
}

}  // namespace pxa
