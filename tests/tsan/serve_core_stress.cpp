// ThreadSanitizer driver of the serving loop's scheduling core
// (pastix_amd/csrc/serve_core.cpp, built with -fsanitize=thread): the slots
// are a host model of the device work (a diagonal matrix D + shift I,
// "factored" by a refactorisation that takes a moment, solved column by
// column), six client threads submit solves and refactorisations, one of
// which fails, under a work budget small enough that slots wait for memory
// and give their blocks back.  Checks: every request ends; every solve
// returns the solution of the version it was bound to, or the failed
// refactorisation's status for that version.  Exit status 0 when all
// checks hold; ThreadSanitizer ends the run with its own status on a report.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "serve_core.h"

namespace {

constexpr int64_t N = 256;       // unknowns
constexpr int ERR_NEG = 7;       // status of a refactorisation with a non-positive pivot
constexpr int ERR_OOM = 5;       // status of a request larger than the budget
constexpr int64_t COL_BYTES = 1 << 20;

double diag(int64_t i) { return 2.0 + (double)(i % 17); }

struct HostSlots final : pxa::ServeOps {
    std::vector<double> shift;  // shift of each slot's factor; touched by its worker only
    explicit HostSlots(int nslots) : shift(nslots, 0.0) {}
    void bind(int) override {}
    int64_t work_bytes(int, int64_t cols) override { return cols * COL_BYTES; }
    int refactor(int k, double s) override {
        std::this_thread::sleep_for(std::chrono::microseconds(300));
        if (diag(0) + s <= 0) return ERR_NEG;
        shift[k] = s;
        return 0;
    }
    int solve(int k, const std::vector<pxa::SolveCols> &batch) override {
        std::this_thread::sleep_for(std::chrono::microseconds(100));
        for (const pxa::SolveCols &c : batch)
            for (int64_t j = 0; j < c.nrhs; j++)
                for (int64_t i = 0; i < N; i++) c.b[i + j * c.ldb] /= diag(i) + shift[k];
        return 0;
    }
    void release(int) override {}
};

}  // namespace

int main() {
    const int nslots = 3;
    HostSlots ops(nslots);
    // 12 columns of budget: merged batches up to 8 columns on three slots
    // make the slots wait for each other's blocks
    pxa::ServeCore core(&ops, nslots, 8, 12 * COL_BYTES, ERR_OOM);
    // client 0 submits the refactorisations one after the other: versions
    // 1, 2 (fails: pivot 2 - 5 < 0), 3
    const std::vector<double> shifts{0.0, 0.5, -5.0, 0.25};
    std::mutex mu;
    int failures = 0, solved = 0, ended_failed = 0;
    auto client = [&](int id) {
        std::mt19937_64 rng(7 + id);
        std::uniform_real_distribution<double> U(-1.0, 1.0);
        for (int r = 0; r < 25; r++) {
            if (id == 0 && (r == 3 || r == 9 || r == 15)) {
                const int64_t v = (r == 3) ? 1 : (r == 9) ? 2 : 3;
                const int64_t t = core.submit_refactor(shifts[v]);
                int rc = 0;
                int64_t ver = -1;
                const bool ok = core.wait(t, &rc, nullptr, &ver);
                std::lock_guard<std::mutex> g(mu);
                if (!ok || ver != v || (v == 2 ? rc != ERR_NEG : rc != 0)) failures++;
                continue;
            }
            // four solves in flight per client, then their results
            struct Sub {
                int64_t t, nrhs, ld;
                std::vector<double> b, x;
            };
            std::vector<Sub> subs(4);
            for (Sub &q : subs) {
                q.nrhs = 1 + (int64_t)(rng() % 6);
                q.ld = N + 3;
                q.b.resize((size_t)(q.ld * q.nrhs));
                for (double &v : q.b) v = U(rng);
                q.x = q.b;
                q.t = core.submit_solve(q.nrhs, q.x.data(), q.ld);
            }
            for (Sub &q : subs) {
                int rc = 0;
                int64_t ver = -1;
                const bool ok = core.wait(q.t, &rc, nullptr, &ver);
                std::lock_guard<std::mutex> g(mu);
                if (!ok || ver < 0 || ver >= (int64_t)shifts.size()) {
                    failures++;
                    continue;
                }
                if (ver == 2) {
                    ended_failed++;
                    if (rc != ERR_NEG) failures++;
                    continue;
                }
                if (rc != 0) {
                    failures++;
                    continue;
                }
                solved++;
                for (int64_t j = 0; j < q.nrhs; j++)
                    for (int64_t i = 0; i < N; i++) {
                        const double want = q.b[i + j * q.ld] / (diag(i) + shifts[ver]);
                        if (std::fabs(q.x[i + j * q.ld] - want) > 1e-15 * std::fabs(want) + 1e-300) failures++;
                    }
            }
        }
    };
    std::vector<std::thread> th;
    for (int id = 0; id < 6; id++) th.emplace_back(client, id);
    for (std::thread &t : th) t.join();
    // a request larger than the budget ends with its status
    {
        std::vector<double> big((size_t)(N * 13), 1.0);
        const int64_t t = core.submit_solve(13, big.data(), N);
        int rc = 0;
        core.wait(t, &rc, nullptr, nullptr);
        if (rc != ERR_OOM) failures++;
    }
    int64_t mem[4], st[4];
    core.memory(mem);
    core.stats(st);
    core.shutdown();
    printf("serve core stress: %d failures, %d solves checked, %d ended with the failed version, %lld batches, "
           "%lld waits for memory, peak %lld of %lld bytes\n",
           failures, solved, ended_failed, (long long)st[0], (long long)mem[3], (long long)mem[2], (long long)mem[0]);
    return failures ? 1 : 0;
}
