// Batched and served partitioning: independent requests run concurrently on
// one GPU, each in its own context (stream + HBM arena), so the launch and
// host latencies of one request overlap with the device work of others.
// Requests are admitted while the estimated HBM footprint of the requests
// in flight stays within a budget.  Every request runs the same
// deterministic pipeline as gp_partition / gp_refine_partition, so its
// result does not depend on what else is in flight.
#include "gpart_mi355x.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

// Device bytes of one request: the input graph plus its hierarchy (levels
// shrink geometrically; the sum stays below twice the input) and the
// refinement state (about 32 bytes per vertex), with sort scratch on top.
std::uint64_t footprint_of(const gp_request* r) {
    const std::uint64_t n = static_cast<std::uint64_t>(r->n), m = static_cast<std::uint64_t>(r->m);
    const std::uint64_t graph = 4 * (n + 1) + 4 * m + (r->edge_weights ? 4 * m : 0) + 4 * n;
    return 3 * graph + 48 * n + 24 * m + (64ULL << 20);
}

struct Job {
    gp_request* req;
    std::uint64_t bytes;
    std::chrono::steady_clock::time_point submitted;
    bool done = false;
};

}  // namespace

struct gp_server {
    int device = 0;
    std::uint64_t budget = 0;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Job>> queue;
    // Tickets handed out and not yet waited for (released by destroy).
    std::unordered_set<std::shared_ptr<Job>*> tickets;
    std::uint64_t in_flight = 0;
    int running = 0;
    bool stopping = false;
    std::string init_error;
    std::atomic<int> init_failures{0};

    void worker_main() {
        gp_context* ctx = nullptr;
        if (gp_context_create(device, &ctx) != GP_OK) {
            std::lock_guard<std::mutex> g(mu);
            init_error = gp_last_error();
            init_failures.fetch_add(1);
            cv_done.notify_all();
            return;
        }
        // Requests overlap one another; each context keeps a single stream.
        gp_context_set_overlap(ctx, 0);
        for (;;) {
            std::shared_ptr<Job> job;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stopping || admissible(); });
                if (queue.empty() && stopping) break;
                job = queue.front();
                queue.pop_front();
                job->req->queue_seconds =
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - job->submitted).count();
                in_flight += job->bytes;
                ++running;
            }
            run(ctx, job->req);
            {
                std::lock_guard<std::mutex> g(mu);
                in_flight -= job->bytes;
                --running;
                job->done = true;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
        gp_context_destroy(ctx);
    }

    // Front request fits the budget, or nothing else is running.
    bool admissible() const {
        if (queue.empty()) return false;
        return running == 0 || in_flight + queue.front()->bytes <= budget;
    }

    static void run(gp_context* ctx, gp_request* r) {
        const auto t0 = std::chrono::steady_clock::now();
        gp_graph* g = nullptr;
        int rc = gp_graph_create(ctx, r->n, r->m, r->offsets, r->neighbors, r->edge_weights, r->vertex_weights, &g);
        if (rc == GP_OK) {
            if (r->warm_labels) {
                std::memcpy(r->labels, r->warm_labels, static_cast<std::size_t>(r->n) * sizeof(int32_t));
                gp_refine_options o;
                gp_default_refine_options(r->parts, r->imbalance, &o);
                gp_level_stats ls;
                rc = gp_refine_partition(ctx, g, r->labels, &o, &ls);
                if (rc == GP_OK) {
                    std::memset(&r->stats, 0, sizeof r->stats);
                    r->stats.final_cut = ls.final_cut;
                    r->stats.levels = 1;
                }
            } else {
                rc = gp_partition(ctx, g, r->parts, r->imbalance, r->seed, r->labels, &r->stats);
            }
            gp_graph_destroy(g);
        }
        if (rc != GP_OK) {
            std::strncpy(r->error, gp_last_error(), sizeof r->error - 1);
            r->error[sizeof r->error - 1] = 0;
        }
        r->status = rc;
        r->device_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

extern "C" {

int gp_server_create(int device, int workers, uint64_t hbm_budget_bytes, gp_server** out) {
    if (!out || workers < 1 || workers > 64) return GP_ERR_INVALID;
    auto s = std::make_unique<gp_server>();
    s->device = device;
    s->budget = hbm_budget_bytes ? hbm_budget_bytes : (64ULL << 30);
    for (int i = 0; i < workers; ++i) s->workers.emplace_back([p = s.get()] { p->worker_main(); });
    *out = s.release();
    return GP_OK;
}

int gp_server_submit(gp_server* s, gp_request* r, gp_ticket** ticket) {
    if (!s || !r || !r->offsets || !r->labels || !ticket || r->n <= 0) return GP_ERR_INVALID;
    auto job = std::make_shared<Job>();
    job->req = r;
    job->bytes = footprint_of(r);
    job->submitted = std::chrono::steady_clock::now();
    r->status = -1;
    auto* handle = new std::shared_ptr<Job>(job);
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->init_failures.load() == static_cast<int>(s->workers.size())) {
            delete handle;
            return GP_ERR_HIP;
        }
        s->queue.push_back(job);
        s->tickets.insert(handle);
    }
    s->cv_work.notify_one();
    *ticket = reinterpret_cast<gp_ticket*>(handle);
    return GP_OK;
}

int gp_server_wait(gp_server* s, gp_ticket* ticket) {
    if (!s || !ticket) return GP_ERR_INVALID;
    auto* jp = reinterpret_cast<std::shared_ptr<Job>*>(ticket);
    std::unique_lock<std::mutex> lk(s->mu);
    if (!s->tickets.erase(jp)) return GP_ERR_INVALID;
    std::shared_ptr<Job> job = *jp;
    delete jp;
    s->cv_done.wait(lk, [&] {
        return job->done || s->init_failures.load() == static_cast<int>(s->workers.size());
    });
    if (!job->done) return GP_ERR_HIP;
    return job->req->status;
}

void gp_server_destroy(gp_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    // Queued requests have run; tickets never waited for are released here.
    for (auto* jp : s->tickets) delete jp;
    s->tickets.clear();
    delete s;
}

int gp_partition_batch(int device, gp_request* items, int64_t count, int workers, uint64_t hbm_budget_bytes) {
    if ((!items && count) || count < 0) return GP_ERR_INVALID;
    gp_server* s = nullptr;
    int rc = gp_server_create(device, workers, hbm_budget_bytes, &s);
    if (rc != GP_OK) return rc;
    // Largest first keeps the tail short.
    std::vector<int64_t> order(static_cast<std::size_t>(count));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return items[a].m + items[a].n > items[b].m + items[b].n; });
    std::vector<gp_ticket*> tickets(static_cast<std::size_t>(count), nullptr);
    for (int64_t i : order) {
        const int r = gp_server_submit(s, &items[i], &tickets[static_cast<std::size_t>(i)]);
        if (r != GP_OK) {
            items[i].status = r;
            if (rc == GP_OK) rc = r;
        }
    }
    for (int64_t i = 0; i < count; ++i) {
        if (!tickets[static_cast<std::size_t>(i)]) continue;
        const int r = gp_server_wait(s, tickets[static_cast<std::size_t>(i)]);
        if (r != GP_OK && rc == GP_OK) rc = r;
    }
    gp_server_destroy(s);
    return rc;
}

}  // extern "C"
