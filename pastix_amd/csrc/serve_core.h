// Scheduling core of the serving loop (pastix_amd_server_*, server.cpp):
// request queue, slot versions, continuous batching of solves, HBM-budget
// admission and one worker thread per slot.  It does no device work itself:
// the slots' operations come through ServeOps (server.cpp binds them to
// pastix instances on the GPU; the ThreadSanitizer test binds them to a
// host model), so the core builds and runs with or without a GPU.
#pragma once
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace pxa {

// The columns of one request in a merged solve: nrhs columns of b, leading
// dimension ldb, overwritten with the solution.
struct SolveCols {
    double *b;
    int64_t ldb, nrhs;
};

// What the loop needs of its slots.  Every call about slot k but the
// constructor's is made by slot k's worker thread, outside the core's lock.
struct ServeOps {
    virtual ~ServeOps() = default;
    virtual void bind(int k) = 0;                         // the worker's start (device selection)
    virtual int64_t work_bytes(int k, int64_t cols) = 0;  // memory a merged solve of cols columns holds
    virtual int refactor(int k, double shift) = 0;        // A + shift I on slot k: 0 or an error code
    virtual int solve(int k, const std::vector<SolveCols> &batch) = 0;  // 0 or an error code
    virtual void release(int k) = 0;                      // give slot k's work memory back
};

class ServeCore {
public:
    // slot 0 serves version 0 (the matrix the slots were factored on);
    // err_oom ends a request that cannot fit the budget even alone
    ServeCore(ServeOps *ops, int nslots, int64_t max_rhs, int64_t budget, int err_oom);
    ~ServeCore();
    int64_t submit_solve(int64_t nrhs, double *b, int64_t ldb);
    int64_t submit_refactor(double shift);
    // blocks until the request ends; false for an unknown ticket
    bool wait(int64_t ticket, int *rc, double *latency, int64_t *version);
    void set_budget(int64_t bytes);
    void memory(int64_t out[4]);  // budget, held, peak, waits for memory
    void stats(int64_t out[4]);   // batches, their columns, newest version, unfinished requests
    void shutdown();              // finish what was admitted, then join the workers

private:
    using Clock = std::chrono::steady_clock;
    struct Req {
        int64_t id = 0;
        int kind = 0;  // 0 solve, 1 refactorisation
        int64_t version = 0;
        int64_t nrhs = 0;
        double *b = nullptr;
        int64_t ldb = 0;
        double shift = 0;
        int state = 0;  // 0 queued, 1 running, 2 done
        int rc = 0;
        int slot = -1;
        Clock::time_point t_submit, t_done;
    };
    void worker(int k);
    int64_t held_total() const;
    void unqueue(Req *r);

    ServeOps *ops_;
    int64_t max_rhs_;
    int err_oom_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Req *> queue_;
    std::unordered_map<int64_t, std::unique_ptr<Req>> reqs_;
    std::map<int64_t, int64_t> live_solves_;  // version -> queued or running solves
    std::map<int64_t, int> failed_;            // version -> status of its failed refactorisation
    int64_t next_id_ = 1, latest_ = 0;
    std::vector<int64_t> held_;   // version each slot holds (-1 none)
    std::vector<char> ready_;     // its factor is built
    std::vector<std::thread> workers_;
    bool stop_ = false;
    // work memory: the budget, and per slot the columns its blocks are
    // sized for, the bytes they hold and whether it is running
    int64_t budget_ = 0, peak_ = 0, waits_ = 0;
    std::vector<int64_t> cap_cols_, held_bytes_;
    std::vector<char> busy_;
    bool pressure_ = false;  // a slot waits for memory: pressure_slot_
    int pressure_slot_ = -1;
    int64_t batches_ = 0, batched_rhs_ = 0;
};

}  // namespace pxa
