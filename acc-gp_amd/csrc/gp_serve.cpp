// Batched and served partitioning: independent requests run concurrently on
// one GPU, each in its own context (stream + HBM arena), so the launch and
// host latencies of one request overlap with the device work of others.
// Requests are admitted while the estimated HBM footprint of the requests
// in flight stays within a budget.  Every request runs the same
// deterministic pipeline as gp_partition / gp_refine_partition, so its
// result does not depend on what else is in flight.
//
// Resident graphs stay in HBM (in a context of the server) between
// requests, with the partition of the last request run on them: a
// repartition request ships only its edge changes, which are applied on the
// device (gp_graph_apply_delta), and refines from that partition.  Requests
// on one resident graph run one at a time in submission order; requests on
// other graphs overlap with them.
#include "gpart_mi355x.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
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

// CSR entries per concurrency slot of a request (GP_SHARE_ENTRIES; 0: every
// request takes one slot).
constexpr std::int64_t kShareEntries = 0;

struct Job {
    gp_request* req;
    std::uint64_t bytes;
    int share = 1;  // concurrency slots the request takes while it runs
    std::chrono::steady_clock::time_point submitted;
    bool done = false;
};

// A graph kept in HBM by the server, with its current partition.
struct Resident {
    gp_graph* g = nullptr;
    std::int64_t n = 0;
    std::vector<std::int32_t> labels;  // partition of the last completed request (empty: none yet)
    bool busy = false;
    int pending = 0;  // queued or running requests on it
};

}  // namespace

struct gp_server {
    int device = 0;
    std::uint64_t budget = 0;
    gp_context* home = nullptr;  // owns the resident graphs (created on first use)
    std::mutex home_mu;          // serialises uses of home's stream
    std::unordered_map<std::uint64_t, std::unique_ptr<Resident>> graphs;
    std::uint64_t next_graph = 1;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Job>> queue;
    // Tickets handed out and not yet waited for (released by destroy).
    std::unordered_set<std::shared_ptr<Job>*> tickets;
    std::uint64_t in_flight = 0;
    int running = 0;
    // Concurrency slots (one per worker) and the CSR entries that make a
    // request take one more: a large request shares the GPU with fewer
    // others, so its latency stays near its latency alone, while small
    // requests, bound by their launch chains, still overlap fully.
    int slots = 1;
    int used = 0;
    std::int64_t share_entries = 0;

    int share_of(std::int64_t m) const {
        if (share_entries <= 0) return 1;
        const std::int64_t k = 1 + m / share_entries;
        return static_cast<int>(std::min<std::int64_t>(k, slots));
    }
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
            Resident* res = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stopping || pick() >= 0; });
                const long at = pick();
                if (at < 0) break;  // stopping with nothing runnable: the queue is empty
                job = queue[static_cast<std::size_t>(at)];
                queue.erase(queue.begin() + at);
                job->req->queue_seconds =
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - job->submitted).count();
                in_flight += job->bytes;
                used += job->share;
                ++running;
                if (job->req->graph_id) {
                    res = graphs.at(job->req->graph_id).get();
                    res->busy = true;
                }
            }
            if (res)
                run_resident(ctx, job->req, *res);
            else
                run(ctx, job->req);
            {
                std::lock_guard<std::mutex> g(mu);
                in_flight -= job->bytes;
                used -= job->share;
                --running;
                job->done = true;
                if (res) {
                    res->busy = false;
                    --res->pending;
                }
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
        gp_context_destroy(ctx);
    }

    // Index of the request to start next, or -1: the first queued request
    // whose resident graph (if any) is idle and has no earlier queued
    // request, provided it fits the budget or nothing else runs.
    long pick() const {
        std::unordered_set<std::uint64_t> seen;
        for (std::size_t i = 0; i < queue.size(); ++i) {
            const std::uint64_t id = queue[i]->req->graph_id;
            if (id) {
                const bool blocked = graphs.at(id)->busy || seen.count(id);
                seen.insert(id);
                if (blocked) continue;
            }
            const bool fits = in_flight + queue[i]->bytes <= budget && used + queue[i]->share <= slots;
            return running == 0 || fits ? static_cast<long>(i) : -1;
        }
        return -1;
    }

    // A request on a resident graph: edge changes first, then a cold
    // partition, a refinement from the caller's labels, or a refinement
    // from the graph's current partition; the result becomes the graph's
    // partition.
    static void run_resident(gp_context* ctx, gp_request* r, Resident& res) {
        const auto t0 = std::chrono::steady_clock::now();
        int rc = GP_OK;
        if (r->delta_count > 0)
            rc = gp_graph_apply_delta(ctx, res.g, r->delta_count, r->delta_u, r->delta_v, r->delta_w, r->delta_insert,
                                      nullptr);
        if (rc == GP_OK) {
            int64_t n = 0, m = 0;
            gp_graph_info(res.g, &n, &m);
            r->n = n;
            r->m = m;
            const int32_t* warm = r->warm_resident ? (res.labels.empty() ? nullptr : res.labels.data()) : r->warm_labels;
            if (r->warm_resident && !warm) {
                rc = GP_ERR_INVALID;
                std::strncpy(r->error, "resident graph has no partition to start from", sizeof r->error - 1);
            } else {
                rc = partition_or_refine(ctx, res.g, r, warm);
            }
            if (rc == GP_OK) res.labels.assign(r->labels, r->labels + n);
        }
        if (rc != GP_OK && !r->error[0]) {
            std::strncpy(r->error, gp_last_error(), sizeof r->error - 1);
            r->error[sizeof r->error - 1] = 0;
        }
        r->status = rc;
        r->device_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    static int partition_or_refine(gp_context* ctx, gp_graph* g, gp_request* r, const int32_t* warm) {
        int rc;
        if (warm) {
            std::memcpy(r->labels, warm, static_cast<std::size_t>(r->n) * sizeof(int32_t));
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
        return rc;
    }

    static void run(gp_context* ctx, gp_request* r) {
        const auto t0 = std::chrono::steady_clock::now();
        gp_graph* g = nullptr;
        int rc = gp_graph_create(ctx, r->n, r->m, r->offsets, r->neighbors, r->edge_weights, r->vertex_weights, &g);
        if (rc == GP_OK) {
            rc = partition_or_refine(ctx, g, r, r->warm_labels);
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
    s->slots = workers;
    s->share_entries = kShareEntries;
    if (const char* e = std::getenv("GP_SHARE_ENTRIES")) s->share_entries = std::atoll(e);
    for (int i = 0; i < workers; ++i) s->workers.emplace_back([p = s.get()] { p->worker_main(); });
    *out = s.release();
    return GP_OK;
}

int gp_server_graph_add(gp_server* s, int64_t n, int64_t m, const int32_t* offsets, const int32_t* neighbors,
                        const uint32_t* edge_weights, const uint32_t* vertex_weights, uint64_t* graph_id) {
    if (!s || !graph_id) return GP_ERR_INVALID;
    std::lock_guard<std::mutex> h(s->home_mu);
    if (!s->home) {
        const int rc = gp_context_create(s->device, &s->home);
        if (rc != GP_OK) return rc;
        gp_context_set_overlap(s->home, 0);
    }
    gp_graph* g = nullptr;
    const int rc = gp_graph_create(s->home, n, m, offsets, neighbors, edge_weights, vertex_weights, &g);
    if (rc != GP_OK) return rc;
    auto res = std::make_unique<Resident>();
    res->g = g;
    res->n = n;
    std::lock_guard<std::mutex> lk(s->mu);
    *graph_id = s->next_graph++;
    s->graphs.emplace(*graph_id, std::move(res));
    return GP_OK;
}

int gp_server_graph_download(gp_server* s, uint64_t graph_id, int64_t* n, int64_t* m, int32_t* offsets,
                             int32_t* neighbors, uint32_t* edge_weights, int32_t* labels) {
    if (!s) return GP_ERR_INVALID;
    Resident* res = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->graphs.find(graph_id);
        if (it == s->graphs.end() || it->second->pending) return GP_ERR_INVALID;
        res = it->second.get();
    }
    int64_t n0 = 0, m0 = 0;
    gp_graph_info(res->g, &n0, &m0);
    if (n) *n = n0;
    if (m) *m = m0;
    if (labels && !res->labels.empty()) std::memcpy(labels, res->labels.data(), res->labels.size() * sizeof(int32_t));
    if (!offsets && !neighbors && !edge_weights) return GP_OK;
    std::lock_guard<std::mutex> h(s->home_mu);
    return gp_graph_download(s->home, res->g, offsets, neighbors, edge_weights, nullptr);
}

int gp_server_graph_release(gp_server* s, uint64_t graph_id) {
    if (!s) return GP_ERR_INVALID;
    std::unique_ptr<Resident> res;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->graphs.find(graph_id);
        if (it == s->graphs.end() || it->second->pending) return GP_ERR_INVALID;
        res = std::move(it->second);
        s->graphs.erase(it);
    }
    std::lock_guard<std::mutex> h(s->home_mu);
    gp_graph_destroy(res->g);
    return GP_OK;
}

int gp_server_submit(gp_server* s, gp_request* r, gp_ticket** ticket) {
    if (!s || !r || !r->labels || !ticket) return GP_ERR_INVALID;
    if (!r->graph_id && (!r->offsets || r->n <= 0)) return GP_ERR_INVALID;
    r->error[0] = 0;
    auto job = std::make_shared<Job>();
    job->req = r;
    job->submitted = std::chrono::steady_clock::now();
    r->status = -1;
    auto* handle = new std::shared_ptr<Job>(job);
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->init_failures.load() == static_cast<int>(s->workers.size())) {
            delete handle;
            return GP_ERR_HIP;
        }
        if (r->graph_id) {
            auto it = s->graphs.find(r->graph_id);
            if (it == s->graphs.end()) {
                delete handle;
                return GP_ERR_INVALID;
            }
            // The footprint of the graph as it stands (edge changes are small).
            int64_t n = 0, m = 0;
            gp_graph_info(it->second->g, &n, &m);
            gp_request shape = *r;
            shape.n = n;
            shape.m = m;
            shape.edge_weights = nullptr;
            job->bytes = footprint_of(&shape);
            job->share = s->share_of(m);
            ++it->second->pending;
        } else {
            job->bytes = footprint_of(r);
            job->share = s->share_of(r->m);
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
    for (auto& kv : s->graphs) gp_graph_destroy(kv.second->g);
    s->graphs.clear();
    if (s->home) gp_context_destroy(s->home);
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
