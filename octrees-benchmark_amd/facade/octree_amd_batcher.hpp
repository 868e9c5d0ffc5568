// Coalescing of concurrent per-query searches into batched device launches.
//
// The reference's harness calls LinearOctree's searches one query at a time from an OpenMP loop
// (neighbor_benchmarks.hpp:507-561). Each such call here becomes a request; the first caller that
// finds no batch running takes every pending request, runs them as one oct_radius / oct_radius_box /
// oct_knn call per (kernel, algorithm, shape) group, hands each caller its own result and wakes the
// others. Requests that arrive while a batch runs form the next batch, so with T calling threads a
// launch serves up to T queries instead of one. A single caller sees batches of one, as before.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <exception>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "octree_amd_runtime.hpp"

namespace octamd {

// One per-query search: inputs, and the result the batch writes back.
struct SearchRequest {
    enum Kind { RADIUS, BOX, KNN } kind = RADIUS;
    int kernel = 0, algo = 0, precision = 0;
    double c[3] = {0, 0, 0};
    double radius = 0.0;           // RADIUS
    double r3[3] = {0, 0, 0};      // BOX: per-axis half extents
    uint32_t k = 0;                // KNN
    std::vector<uint32_t> entries; // RADIUS / BOX: indices, or [first, last) pairs for Struct
    std::vector<int32_t> knn_idx;  // KNN: k indices (-1 past the cloud)
    std::vector<double> knn_d2;    // KNN: k squared distances
    bool done = false;
    std::exception_ptr error;
};

class SearchBatcher {
public:
    SearchBatcher(oct_ctx *ctx, std::mutex &ctx_mutex) : ctx_(ctx), ctx_mu_(ctx_mutex) {}

    // Runs r in some caller's batch and returns when its result is in place (or rethrows its error).
    void run(SearchRequest &r) {
        std::unique_lock<std::mutex> lk(mu_);
        pending_.push_back(&r);
        while (!r.done) {
            if (leader_) {
                cv_.wait(lk);
                continue;
            }
            leader_ = true;
            std::vector<SearchRequest *> batch;
            batch.swap(pending_);
            lk.unlock();
            execute(batch);
            lk.lock();
            for (SearchRequest *q : batch) q->done = true;
            leader_ = false;
            cv_.notify_all();
        }
        if (r.error) std::rethrow_exception(r.error);
    }

private:
    oct_ctx *ctx_;
    std::mutex &ctx_mu_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::vector<SearchRequest *> pending_;
    bool leader_ = false;

    void execute(std::vector<SearchRequest *> &batch) {
        std::map<std::tuple<int, int, int, int, uint32_t>, std::vector<SearchRequest *>> groups;
        for (SearchRequest *q : batch) groups[{(int)q->kind, q->kernel, q->algo, q->precision, q->k}].push_back(q);
        std::lock_guard<std::mutex> g(ctx_mu_);
        for (auto &kv : groups) {
            std::vector<SearchRequest *> &qs = kv.second;
            try {
                if (qs[0]->kind == SearchRequest::KNN)
                    run_knn(qs);
                else
                    run_lists(qs);
            } catch (...) {
                const std::exception_ptr e = std::current_exception();
                for (SearchRequest *q : qs) q->error = e;
            }
        }
    }

    void run_lists(std::vector<SearchRequest *> &qs) {
        const size_t nq = qs.size();
        const SearchRequest &f = *qs[0];
        std::vector<double> c(3 * nq), r(nq), r3(3 * nq);
        for (size_t i = 0; i < nq; i++)
            for (int a = 0; a < 3; a++) c[3 * i + a] = qs[i]->c[a], r3[3 * i + a] = qs[i]->r3[a];
        for (size_t i = 0; i < nq; i++) r[i] = qs[i]->radius;
        if (f.kind == SearchRequest::BOX)
            check(ctx_, oct_radius_box(ctx_, f.kernel, f.algo, c.data(), nullptr, r3.data(), nullptr, nq, 1, nullptr, nullptr),
                  "oct_radius_box");
        else
            check(ctx_, oct_radius(ctx_, f.kernel, f.algo, f.precision, c.data(), nullptr, r.data(), 0.0, nq, 1, nullptr, nullptr),
                  "oct_radius");
        std::vector<uint64_t> offs(nq + 1);
        check(ctx_, oct_radius_fetch(ctx_, offs.data(), nullptr, 0), "oct_radius_fetch");
        const size_t per = f.algo == OCT_ALGO_STRUCT ? 2 : 1;
        std::vector<uint32_t> ent((size_t)offs[nq] * per);
        if (!ent.empty()) check(ctx_, oct_radius_fetch(ctx_, offs.data(), ent.data(), ent.size()), "oct_radius_fetch");
        for (size_t i = 0; i < nq; i++)
            qs[i]->entries.assign(ent.begin() + (ptrdiff_t)(offs[i] * per), ent.begin() + (ptrdiff_t)(offs[i + 1] * per));
    }

    void run_knn(std::vector<SearchRequest *> &qs) {
        const size_t nq = qs.size();
        const uint32_t k = qs[0]->k;
        std::vector<double> c(3 * nq);
        for (size_t i = 0; i < nq; i++)
            for (int a = 0; a < 3; a++) c[3 * i + a] = qs[i]->c[a];
        std::vector<int32_t> idx(nq * k);
        std::vector<double> d2(nq * k);
        check(ctx_, oct_knn(ctx_, qs[0]->precision, c.data(), nullptr, nq, k, idx.data(), d2.data(), nullptr), "oct_knn");
        for (size_t i = 0; i < nq; i++) {
            qs[i]->knn_idx.assign(idx.begin() + (ptrdiff_t)(i * k), idx.begin() + (ptrdiff_t)((i + 1) * k));
            qs[i]->knn_d2.assign(d2.begin() + (ptrdiff_t)(i * k), d2.begin() + (ptrdiff_t)((i + 1) * k));
        }
    }
};

}  // namespace octamd
