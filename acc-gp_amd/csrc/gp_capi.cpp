// C ABI (include/gpart_mi355x.h) over the gfx950 partitioner.
#include "gpart_mi355x.h"

#include "gp_coarsen.hpp"
#include "gp_host.hpp"
#include "gp_metrics.hpp"
#include "gp_refine.hpp"
#include "gp_update.hpp"

#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <new>
#include <string>

struct gp_context {
    explicit gp_context(int device) : ctx(device) {}
    gp::Context ctx;
};

namespace gp {
std::uint64_t& thread_launches() {
    thread_local std::uint64_t count = 0;
    return count;
}
}  // namespace gp

namespace {
std::atomic<std::uint64_t> g_graph_serial{0};
}

struct gp_graph {
    gp_graph() : serial(++g_graph_serial) {}
    std::unique_ptr<gp::DGraph> owned;
    gp::DGraph* view = nullptr;
    const std::uint64_t serial;  // identifies the graph a workspace describes
};

// SclpWorkspace (include/coarsen.hpp:55-69): the coarsening state carried
// from one level to the next (row flags, weight bounds, per-vertex row
// bounds of the current level) and the graph it describes.
struct gp_workspace {
    gp::CoarsenState st;
    std::uint64_t current = 0;  // serial of the graph st describes (0: none)
};

// DeviceAggregateResult (include/coarsen.hpp:40-46).
struct gp_aggregate {
    gp::AggregateResult r;
    gp::DBuf<std::int32_t> map;
    gp::DBuf<std::uint32_t> cvw;
    std::uint64_t graph = 0;  // serial of the aggregated graph
    std::uint64_t state = 0;  // serial the workspace state must still describe
    bool consumed = false;
};

struct gp_hierarchy {
    gp::Hierarchy h;
    std::vector<std::unique_ptr<gp_graph>> views;
    gp_graph* input = nullptr;
    bool owns_input = false;  // assembled hierarchies own their level 0
    ~gp_hierarchy() {
        views.clear();
        h = gp::Hierarchy();
        if (owns_input) delete input;
    }
};

struct gp_result {
    gp::UncoarsenResult r;
};

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& what) {
    g_error = what;
    return code;
}

template <typename Fn>
int guarded(Fn&& fn) {
    try {
        fn();
        return GP_OK;
    } catch (const gp::HipError& e) {
        return fail(GP_ERR_HIP, e.what());
    } catch (const std::invalid_argument& e) {
        return fail(GP_ERR_INVALID, e.what());
    } catch (const std::bad_alloc& e) {
        return fail(GP_ERR_MEMORY, e.what());
    } catch (const std::exception& e) {
        return fail(GP_ERR_RUNTIME, e.what());
    } catch (...) {
        return fail(GP_ERR_RUNTIME, "unknown error");
    }
}

gp::CoarsenOptions to_options(const gp_coarsen_options& o) {
    gp::CoarsenOptions c;
    c.parts = o.parts;
    c.seed = o.seed;
    c.stop_ratio = o.stop_ratio;
    c.max_levels = o.max_levels;
    c.cluster_size_factor = o.cluster_size_factor;
    c.pack_isolated = o.pack_isolated != 0;
    c.progressive = o.progressive != 0;
    c.fine_factor = o.fine_factor;
    c.capacity_growth = o.capacity_growth;
    c.seed_matching = o.seed_matching != 0;
    c.fast_path = o.fast_path != 0;
    c.hem_degree_limit = o.hem_degree_limit;
    return c;
}

gp::RefineOptions to_options(const gp_refine_options& o) {
    gp::RefineOptions r;
    r.parts = o.parts;
    r.imbalance = o.imbalance;
    r.max_rounds = o.max_rounds;
    r.stale_rounds = o.stale_rounds;
    r.growth_ratio = o.growth_ratio;
    r.exchange_rounds = o.exchange_rounds;
    r.strict_verify = o.strict_verify != 0;
    r.device_loop = o.gpu_lp_loop != 0;
    return r;
}

void check_refine_options(const gp::RefineOptions& o) {
    if (o.parts < 2 || o.parts > gp::kMaxParts || o.max_rounds < 0 || o.max_rounds > 64 || o.stale_rounds < 1 ||
        o.stale_rounds > 64 || o.exchange_rounds < 0 || o.exchange_rounds > 8 || !(o.growth_ratio >= 1.0) ||
        !(o.imbalance >= 1.0))
        throw std::invalid_argument("invalid refinement options");
}

gp_level_stats to_c(const gp::LevelStats& s) {
    gp_level_stats o;
    o.n = s.n;
    o.m = s.m;
    o.refined = s.refined;
    o.initial_repair_failed = s.initial_repair_failed;
    o.rounds = s.rounds;
    o.initial_cut = s.initial_cut;
    o.final_cut = s.final_cut;
    o.lp_moves = s.lp_moves;
    o.repair_moves = s.repair_moves;
    o.group_moves = s.group_moves;
    o.repair_iterations = s.repair_iterations;
    o.strict_checks = s.strict_checks;
    return o;
}

template <typename T>
T* host_copy(const std::vector<T>& v) {
    T* p = static_cast<T*>(std::malloc(std::max<std::size_t>(1, v.size()) * sizeof(T)));
    if (!p) throw std::bad_alloc();
    if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

// Host copy of a device graph (copy_device_weighted).
gp::HostGraph download_graph(gp::Context& c, const gp::DGraph& g) {
    gp::HostGraph h;
    h.n = g.n;
    h.off.resize(static_cast<std::size_t>(g.n) + 1);
    h.nbr.resize(static_cast<std::size_t>(g.m));
    h.vw.resize(static_cast<std::size_t>(g.n));
    c.download(h.off.data(), g.off.get(), h.off.size() * sizeof(std::int32_t));
    c.download(h.nbr.data(), g.nbr.get(), h.nbr.size() * sizeof(std::int32_t));
    c.download(h.vw.data(), g.vw.get(), h.vw.size() * sizeof(std::uint32_t));
    if (g.ew) {
        h.ew.resize(static_cast<std::size_t>(g.m));
        c.download(h.ew.data(), g.ew.get(), h.ew.size() * sizeof(std::uint32_t));
    }
    return h;
}

double seconds_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace

extern "C" {

const char* gp_last_error(void) { return g_error.c_str(); }

int gp_context_create(int device, gp_context** out) {
    if (!out) return fail(GP_ERR_INVALID, "null output");
    return guarded([&] {
        auto c = std::make_unique<gp_context>(device);
        // Set-up of gp_edge_cut_noalloc: the device address of its static
        // scratch on this device, resolved once per device (device current).
        gp::resolve_noalloc_scratch(device);
        *out = c.release();
    });
}
void gp_context_destroy(gp_context* ctx) { delete ctx; }
uint64_t gp_context_reserved_bytes(const gp_context* ctx) {
    return ctx ? const_cast<gp_context*>(ctx)->ctx.arena()->reserved_bytes() : 0;
}
int gp_context_set_overlap(gp_context* ctx, int on) {
    if (!ctx) return fail(GP_ERR_INVALID, "null context");
    ctx->ctx.set_overlap(on != 0);
    return GP_OK;
}

int gp_graph_create(gp_context* ctx, int64_t n, int64_t m, const int32_t* offsets, const int32_t* neighbors,
                    const uint32_t* edge_weights, const uint32_t* vertex_weights, gp_graph** out) {
    if (!ctx || !out || !offsets || (m > 0 && !neighbors)) return fail(GP_ERR_INVALID, "null graph input");
    return guarded([&] {
        if (n <= 0 || n > 0x7fffffffLL || m < 0 || m > 0x7fffffffLL)
            throw std::invalid_argument("graph exceeds supported 32-bit CSR limits");
        // validate_weighted_shape (src/graph.cpp:118-155)
        if (offsets[0] != 0 || offsets[n] != m) throw std::invalid_argument("invalid weighted CSR shape");
        for (int64_t v = 0; v < n; ++v)
            if (offsets[v] > offsets[v + 1]) throw std::invalid_argument("weighted CSR offsets are not monotone");
        for (int64_t e = 0; e < m; ++e)
            if (neighbors[e] < 0 || neighbors[e] >= n) throw std::invalid_argument("weighted CSR contains an invalid neighbor");
        if (vertex_weights)
            for (int64_t v = 0; v < n; ++v)
                if (!vertex_weights[v]) throw std::invalid_argument("vertex weights must be positive");
        gp::Context& c = ctx->ctx;
        auto g = std::make_unique<gp::DGraph>();
        g->n = n;
        g->m = m;
        g->off = c.buf<std::int32_t>(static_cast<std::size_t>(n + 1));
        g->nbr = c.buf<std::int32_t>(static_cast<std::size_t>(std::max<int64_t>(m, 1)));
        g->vw = c.buf<std::uint32_t>(static_cast<std::size_t>(n));
        c.upload(g->off.get(), offsets, static_cast<std::size_t>(n + 1) * 4);
        if (m) c.upload(g->nbr.get(), neighbors, static_cast<std::size_t>(m) * 4);
        bool unit = true;
        if (edge_weights)
            for (int64_t e = 0; e < m && unit; ++e) unit = edge_weights[e] == 1;
        if (!unit) {
            g->ew = c.buf<std::uint32_t>(static_cast<std::size_t>(m));
            c.upload(g->ew.get(), edge_weights, static_cast<std::size_t>(m) * 4);
        }
        if (vertex_weights)
            c.upload(g->vw.get(), vertex_weights, static_cast<std::size_t>(n) * 4);
        else
            gp::fill<std::uint32_t>(c, g->vw.get(), n, 1u);
        c.sync();
        auto* h = new gp_graph;
        h->owned = std::move(g);
        h->view = h->owned.get();
        *out = h;
    });
}

int gp_graph_info(const gp_graph* g, int64_t* n, int64_t* m) {
    if (!g) return fail(GP_ERR_INVALID, "null graph");
    if (n) *n = g->view->n;
    if (m) *m = g->view->m;
    return GP_OK;
}

int gp_graph_download(gp_context* ctx, const gp_graph* g, int32_t* offsets, int32_t* neighbors,
                      uint32_t* edge_weights, uint32_t* vertex_weights) {
    if (!ctx || !g) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        gp::Context& c = ctx->ctx;
        const gp::DGraph& d = *g->view;
        if (offsets) c.download(offsets, d.off.get(), static_cast<std::size_t>(d.n + 1) * 4);
        if (neighbors && d.m) c.download(neighbors, d.nbr.get(), static_cast<std::size_t>(d.m) * 4);
        if (vertex_weights) c.download(vertex_weights, d.vw.get(), static_cast<std::size_t>(d.n) * 4);
        if (edge_weights && d.m) {
            if (d.ew)
                c.download(edge_weights, d.ew.get(), static_cast<std::size_t>(d.m) * 4);
            else
                for (int64_t e = 0; e < d.m; ++e) edge_weights[e] = 1;
        }
    });
}

int gp_graph_metadata(const gp_graph* g, uint64_t* max_degree, int* has_heavy, uint64_t* incident_bound,
                      int* implicit_unit_weights) {
    if (!g) return fail(GP_ERR_INVALID, "null graph");
    const gp::DGraph& d = *g->view;
    if (max_degree) *max_degree = d.max_degree;
    if (has_heavy) *has_heavy = d.has_heavy;
    if (incident_bound) *incident_bound = d.incident_bound;
    if (implicit_unit_weights) *implicit_unit_weights = d.unit_edges();
    return GP_OK;
}

int gp_graph_apply_delta(gp_context* ctx, gp_graph* g, int64_t count, const int32_t* u, const int32_t* v,
                         const uint32_t* weights, const uint8_t* insert, int64_t* m) {
    if (!ctx || !g || !g->view) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        if (!g->owned) throw std::invalid_argument("only graphs made by gp_graph_create or gp_contract can change");
        ctx->ctx.activate();
        const std::int64_t m2 = gp::apply_delta(ctx->ctx, *g->view, count, u, v, weights, insert);
        if (m) *m = m2;
    });
}

void gp_graph_destroy(gp_graph* g) { delete g; }

int gp_debug_below_0999(uint64_t cut, uint64_t best) { return gp::debug_below_0999(cut, best) ? 1 : 0; }

int gp_debug_radix_sort(gp_context* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int begin_bit, int end_bit) {
    return guarded([&] {
        if (!ctx || (n > 0 && !keys)) throw std::invalid_argument("null argument");
        ctx->ctx.activate();
        gp::debug_radix_sort(ctx->ctx, keys, vals, n, begin_bit, end_bit);
    });
}

int gp_capacity(uint64_t total, int parts, double imbalance, uint64_t* out) {
    if (!out) return fail(GP_ERR_INVALID, "null output");
    return guarded([&] { *out = gp::capacity_of(total, parts, imbalance); });
}

void gp_default_coarsen_options(int parts, uint32_t seed, gp_coarsen_options* out) {
    if (!out) return;
    const gp::CoarsenOptions o = gp::CoarsenOptions::production(parts, seed);
    out->parts = o.parts;
    out->seed = o.seed;
    out->stop_ratio = o.stop_ratio;
    out->max_levels = o.max_levels;
    out->cluster_size_factor = o.cluster_size_factor;
    out->pack_isolated = o.pack_isolated;
    out->progressive = o.progressive;
    out->fine_factor = o.fine_factor;
    out->capacity_growth = o.capacity_growth;
    out->seed_matching = o.seed_matching;
    out->fast_path = o.fast_path;
    out->hem_degree_limit = o.hem_degree_limit;
}

int gp_coarsen(gp_context* ctx, gp_graph* input, const gp_coarsen_options* opts, gp_hierarchy** out) {
    if (!ctx || !input || !opts || !out) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        auto h = std::make_unique<gp_hierarchy>();
        h->h = gp::coarsen(ctx->ctx, *input->view, to_options(*opts));
        h->input = input;
        for (int i = 1; i < h->h.levels(); ++i) {
            auto v = std::make_unique<gp_graph>();
            v->view = h->h.coarse[static_cast<std::size_t>(i - 1)].get();
            h->views.push_back(std::move(v));
        }
        ctx->ctx.sync();
        *out = h.release();
    });
}

int gp_hierarchy_levels(const gp_hierarchy* h) { return h ? h->h.levels() : 0; }

const gp_graph* gp_hierarchy_level(const gp_hierarchy* h, int level) {
    if (!h || level < 0 || level >= h->h.levels()) return nullptr;
    return level == 0 ? h->input : h->views[static_cast<std::size_t>(level - 1)].get();
}

int gp_hierarchy_map_download(gp_context* ctx, const gp_hierarchy* h, int level, int32_t* map) {
    if (!ctx || !h || !map || level < 0 || level + 1 >= h->h.levels()) return fail(GP_ERR_INVALID, "invalid level");
    return guarded([&] {
        const gp::DGraph& g = h->h.level(level);
        ctx->ctx.download(map, h->h.maps[static_cast<std::size_t>(level)].get(), static_cast<std::size_t>(g.n) * 4);
    });
}

int gp_hierarchy_level_info(const gp_hierarchy* h, int level, uint64_t* max_degree, int* has_heavy,
                            uint64_t* incident_bound) {
    if (!h || level < 0 || level >= h->h.levels()) return fail(GP_ERR_INVALID, "invalid level");
    const gp::DGraph& g = h->h.level(level);
    if (max_degree) *max_degree = g.max_degree;
    if (has_heavy) *has_heavy = g.has_heavy;
    if (incident_bound) *incident_bound = g.incident_bound;
    return GP_OK;
}

int gp_hierarchy_assemble(gp_context* ctx, int levels, gp_graph* const* level_graphs, const int32_t* const* maps,
                          gp_hierarchy** out) {
    if (!ctx || !level_graphs || !out || levels < 1 || (levels > 1 && !maps)) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        gp::Context& c = ctx->ctx;
        for (int i = 0; i < levels; ++i)
            if (!level_graphs[i] || !level_graphs[i]->owned)
                throw std::invalid_argument("hierarchy levels must be graphs made by gp_graph_create or gp_contract");
        for (int i = 0; i + 1 < levels; ++i) {
            const std::int64_t nf = level_graphs[i]->view->n, nc = level_graphs[i + 1]->view->n;
            if (!maps[i] || nc > nf) throw std::invalid_argument("resident refinement requires a complete hierarchy");
            for (std::int64_t v = 0; v < nf; ++v)
                if (maps[i][v] < 0 || maps[i][v] >= nc) throw std::invalid_argument("hierarchy map contains an invalid coarse id");
        }
        auto h = std::make_unique<gp_hierarchy>();
        for (int i = 0; i < levels; ++i) {
            gp::DGraph& d = *level_graphs[i]->view;
            gp::CoarsenState st;
            gp::compute_level0_metadata(c, d, st);  // exact bounds of a level built elsewhere
            if (i + 1 < levels) {
                gp::DBuf<std::int32_t> m = c.buf<std::int32_t>(static_cast<std::size_t>(d.n));
                c.upload(m.get(), maps[i], static_cast<std::size_t>(d.n) * sizeof(std::int32_t));
                h->h.maps.push_back(std::move(m));
                h->h.caps.push_back(std::numeric_limits<std::uint64_t>::max());
            }
        }
        // Ownership moves in only once everything above has succeeded.
        h->input = level_graphs[0];
        h->h.input = level_graphs[0]->view;
        h->h.stop_reason = "assembled";
        for (int i = 1; i < levels; ++i) {
            h->h.coarse.push_back(std::move(level_graphs[i]->owned));
            auto v = std::make_unique<gp_graph>();
            v->view = h->h.coarse.back().get();
            h->views.push_back(std::move(v));
            delete level_graphs[i];
        }
        h->owns_input = true;
        c.sync();
        *out = h.release();
    });
}

int gp_graph_edge_weight(gp_context* ctx, const gp_graph* g, uint64_t* total) {
    if (!ctx || !g || !total) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        const std::uint64_t directed = gp::device_edge_weight_sum(ctx->ctx, *g->view);
        if (directed & 1ULL) throw std::runtime_error("symmetric graph has odd directed edge weight");
        *total = directed / 2;
    });
}

int gp_hierarchy_level_capacity(const gp_hierarchy* h, int level, uint64_t* capacity) {
    if (!h || !capacity || level < 0 || level + 1 >= h->h.levels()) return fail(GP_ERR_INVALID, "invalid level");
    *capacity = h->h.caps[static_cast<std::size_t>(level)];
    return GP_OK;
}

const char* gp_hierarchy_stop_reason(const gp_hierarchy* h) { return h ? h->h.stop_reason.c_str() : ""; }

void gp_hierarchy_destroy(gp_hierarchy* h) { delete h; }

int gp_workspace_create(gp_workspace** out) {
    if (!out) return fail(GP_ERR_INVALID, "null output");
    return guarded([&] { *out = new gp_workspace; });
}

void gp_workspace_destroy(gp_workspace* ws) { delete ws; }

int gp_aggregate_level(gp_context* ctx, gp_workspace* ws, gp_graph* g, int parts, uint32_t seed, int level,
                       int cluster_size_factor, int pack_isolated, int seed_matching, int fast_path,
                       gp_aggregate** out) {
    if (!ctx || !ws || !g || !out) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        if (parts < 2 || parts > gp::kMaxParts || cluster_size_factor < 2 || level < 0)
            throw std::invalid_argument("invalid aggregation options");
        gp::Context& c = ctx->ctx;
        gp::DGraph& d = *g->view;
        if (ws->current != g->serial) {
            // A graph the workspace has not contracted into: level-0 metadata.
            ws->st = gp::CoarsenState();
            gp::compute_level0_metadata(c, d, ws->st);
            ws->current = g->serial;
        }
        auto a = std::make_unique<gp_aggregate>();
        a->r = gp::aggregate(c, d, ws->st, parts, seed, level, cluster_size_factor, pack_isolated != 0,
                             seed_matching != 0, fast_path != 0, a->map, a->cvw);
        // The aggregation pass tightens the level's degree metadata.
        d.max_degree = a->r.stats_max_degree;
        d.has_heavy = a->r.stats_has_heavy;
        a->graph = g->serial;
        a->state = ws->current;
        c.sync();
        *out = a.release();
    });
}

int gp_aggregate_info(const gp_aggregate* a, int64_t* coarse_vertices, uint64_t* maximum_weight, uint64_t* capacity) {
    if (!a) return fail(GP_ERR_INVALID, "null aggregate");
    if (coarse_vertices) *coarse_vertices = a->r.coarse_n;
    if (maximum_weight) *maximum_weight = a->r.max_weight;
    if (capacity) *capacity = a->r.capacity;
    return GP_OK;
}

int gp_aggregate_map_download(gp_context* ctx, const gp_aggregate* a, int32_t* map) {
    if (!ctx || !a || !map) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] { ctx->ctx.download(map, a->map.get(), a->map.size() * sizeof(std::int32_t)); });
}

int gp_aggregate_vertex_weights_download(gp_context* ctx, const gp_aggregate* a, uint32_t* vertex_weights) {
    if (!ctx || !a || !vertex_weights) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        if (a->consumed) throw std::invalid_argument("aggregate was consumed by gp_contract");
        ctx->ctx.download(vertex_weights, a->cvw.get(), static_cast<std::size_t>(a->r.coarse_n) * sizeof(std::uint32_t));
    });
}

void gp_aggregate_destroy(gp_aggregate* a) { delete a; }

int gp_contract(gp_context* ctx, gp_workspace* ws, const gp_graph* fine, gp_aggregate* a, gp_graph** out) {
    if (!ctx || !ws || !fine || !a || !out) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        if (a->consumed) throw std::invalid_argument("aggregate was already contracted");
        if (a->graph != fine->serial || ws->current != a->state || a->state != fine->serial)
            throw std::invalid_argument("aggregate, workspace and fine graph do not belong together");
        gp::Context& c = ctx->ctx;
        auto coarse = gp::contract(c, *fine->view, ws->st, a->map.get(), a->r.coarse_n, std::move(a->cvw));
        a->consumed = true;
        ws->st.max_vertex_weight = a->r.max_weight;
        ws->st.sorted_unique_rows = true;
        auto h = std::make_unique<gp_graph>();
        h->owned = std::move(coarse);
        h->view = h->owned.get();
        ws->current = h->serial;
        c.sync();
        *out = h.release();
    });
}

int gp_initial_partition_host(int64_t n, const int32_t* offsets, const int32_t* neighbors,
                              const uint32_t* edge_weights, const uint32_t* vertex_weights, int parts,
                              double imbalance, uint32_t seed, int require_feasible, int32_t* labels) {
    if (!offsets || !labels) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        gp::initial_partition(n, offsets, neighbors, edge_weights, vertex_weights, parts, imbalance, seed,
                              require_feasible != 0, labels);
    });
}

int gp_initial_partition(gp_context* ctx, const gp_graph* g, int parts, double imbalance, uint32_t seed,
                         int require_feasible, int32_t* labels) {
    if (!ctx || !g || !labels) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        gp::HostGraph h = download_graph(ctx->ctx, *g->view);
        gp::initial_partition(h.n, h.off.data(), h.nbr.data(), h.ew.empty() ? nullptr : h.ew.data(), h.vw.data(),
                              parts, imbalance, seed, require_feasible != 0, labels);
    });
}

int gp_repair_capacity_host(int64_t n, const int32_t* offsets, const int32_t* neighbors,
                            const uint32_t* edge_weights, const uint32_t* vertex_weights, int parts, uint64_t capacity,
                            int32_t* labels, uint64_t* moved) {
    if (!offsets || !labels || (offsets[n] > 0 && !neighbors)) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        if (n <= 0 || parts < 2 || parts > gp::kMaxParts) throw std::invalid_argument("invalid repair options");
        std::vector<std::uint64_t> loads(static_cast<std::size_t>(parts), 0);
        for (int64_t v = 0; v < n; ++v) {
            if (labels[v] < 0 || labels[v] >= parts) throw std::invalid_argument("labels outside [0, parts)");
            loads[static_cast<std::size_t>(labels[v])] += vertex_weights ? vertex_weights[v] : 1u;
        }
        const std::size_t k = gp::repair_capacity(n, offsets, neighbors, edge_weights, vertex_weights, parts, capacity,
                                                  labels, loads);
        if (moved) *moved = k;
    });
}

const char* gp_initial_partition_backend(void) { return gp::initial_partition_backend(); }

void gp_default_refine_options(int parts, double imbalance, gp_refine_options* out) {
    if (!out) return;
    gp::RefineOptions o;
    out->parts = parts;
    out->imbalance = imbalance;
    out->max_rounds = o.max_rounds;
    out->stale_rounds = o.stale_rounds;
    out->growth_ratio = o.growth_ratio;
    out->exchange_rounds = o.exchange_rounds;
    out->strict_verify = 0;
    out->gpu_lp_loop = o.device_loop ? 1 : 0;
}

int gp_uncoarsen(gp_context* ctx, const gp_hierarchy* h, const int32_t* coarsest_labels,
                 const gp_refine_options* opts, gp_result** out) {
    if (!ctx || !h || !coarsest_labels || !opts || !out) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        gp::RefineOptions o = to_options(*opts);
        check_refine_options(o);
        gp::Context& c = ctx->ctx;
        const gp::DGraph& cg = h->h.level(h->h.levels() - 1);
        for (int64_t v = 0; v < cg.n; ++v)
            if (coarsest_labels[v] < 0 || coarsest_labels[v] >= o.parts)
                throw std::invalid_argument("invalid coarsest device partition");
        gp::DBuf<std::int32_t> lab = c.buf<std::int32_t>(static_cast<std::size_t>(cg.n));
        c.upload(lab.get(), coarsest_labels, static_cast<std::size_t>(cg.n) * 4);
        auto r = std::make_unique<gp_result>();
        r->r = gp::uncoarsen(c, h->h, lab.get(), o);
        c.sync();
        *out = r.release();
    });
}

int gp_result_labels(gp_context* ctx, const gp_result* r, int32_t* labels) {
    if (!ctx || !r || !labels) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        ctx->ctx.download(labels, r->r.labels.get(), r->r.labels.size() * sizeof(std::int32_t));
    });
}
uint64_t gp_result_cut(const gp_result* r) { return r ? r->r.cut : 0; }
int gp_result_loads(const gp_result* r, uint64_t* loads) {
    if (!r || !loads) return fail(GP_ERR_INVALID, "null argument");
    std::memcpy(loads, r->r.loads.data(), r->r.loads.size() * sizeof(uint64_t));
    return GP_OK;
}
int gp_result_levels(const gp_result* r) { return r ? static_cast<int>(r->r.levels.size()) : 0; }
int gp_result_level_stats(const gp_result* r, int index, gp_level_stats* out) {
    if (!r || !out || index < 0 || index >= static_cast<int>(r->r.levels.size())) return fail(GP_ERR_INVALID, "bad index");
    *out = to_c(r->r.levels[static_cast<std::size_t>(index)]);
    return GP_OK;
}
void gp_result_destroy(gp_result* r) { delete r; }
const int32_t* gp_result_device_labels(const gp_result* r) { return r ? r->r.labels.get() : nullptr; }

int gp_edge_cut(gp_context* ctx, const gp_graph* g, const int32_t* labels, int parts, uint64_t* cut,
                uint64_t* loads) {
    return guarded([&] {
        if (!ctx || !g || !g->view || !cut || !loads) throw std::invalid_argument("null argument");
        ctx->ctx.activate();
        gp::edge_cut(ctx->ctx, *g->view, labels, parts, cut, loads);
    });
}

int gp_edge_cut_noalloc(gp_context* ctx, const gp_graph* g, const int32_t* labels, int parts, uint64_t* cut,
                        uint64_t* loads) {
    return guarded([&] {
        if (!ctx || !g || !g->view || !cut || !loads) throw std::invalid_argument("null argument");
        ctx->ctx.activate();
        gp::edge_cut_noalloc(ctx->ctx, *g->view, labels, parts, cut, loads);
    });
}

int gp_refine_partition(gp_context* ctx, const gp_graph* g, int32_t* labels, const gp_refine_options* opts,
                        gp_level_stats* stats) {
    if (!ctx || !g || !labels || !opts) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        gp::RefineOptions o = to_options(*opts);
        check_refine_options(o);
        gp::Context& c = ctx->ctx;
        const gp::DGraph& d = *g->view;
        for (int64_t v = 0; v < d.n; ++v)
            if (labels[v] < 0 || labels[v] >= o.parts) throw std::invalid_argument("refinement partition has an invalid part id");
        gp::DBuf<std::int32_t> lab = c.buf<std::int32_t>(static_cast<std::size_t>(d.n));
        c.upload(lab.get(), labels, static_cast<std::size_t>(d.n) * 4);
        std::vector<std::uint64_t> loads(static_cast<std::size_t>(o.parts));
        gp::device_loads(c, d, lab.get(), o.parts, loads.data());
        std::uint64_t total = 0;
        for (auto x : loads) total += x;
        const std::uint64_t cap = gp::capacity_of(total, o.parts, o.imbalance);
        gp::LevelStats st = gp::refine_level(c, d, lab.get(), loads, o, cap, o.max_rounds, false, 0);
        c.download(labels, lab.get(), static_cast<std::size_t>(d.n) * 4);
        if (stats) *stats = to_c(st);
    });
}

int gp_partition(gp_context* ctx, gp_graph* g, int parts, double imbalance, uint32_t seed, int32_t* labels,
                 gp_partition_stats* stats) {
    if (!ctx || !g) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        gp::Context& c = ctx->ctx;
        gp::DGraph& d = *g->view;
        gp::capacity_of(static_cast<std::uint64_t>(d.n), parts, imbalance);
        const std::uint64_t launches0 = gp::thread_launches(), syncs0 = c.syncs();
        const auto t0 = std::chrono::steady_clock::now();
        gp::Hierarchy h = gp::coarsen(c, d, gp::CoarsenOptions::production(parts, seed));
        c.sync();
        const double t_coarsen = seconds_since(t0);
        const auto t1 = std::chrono::steady_clock::now();
        const gp::DGraph& cg = h.level(h.levels() - 1);
        gp::HostGraph hc = download_graph(c, cg);
        std::vector<std::int32_t> coarse_labels(static_cast<std::size_t>(cg.n));
        if (parts > cg.n) throw std::invalid_argument("more parts than coarsest vertices");
        // Production defers coarse imbalance to the repair of the refinement
        // (app/partition_main.cu:111-112: require_feasible = false).
        gp::initial_partition(hc.n, hc.off.data(), hc.nbr.data(), hc.ew.empty() ? nullptr : hc.ew.data(),
                              hc.vw.data(), parts, imbalance, seed, false, coarse_labels.data());
        gp::DBuf<std::int32_t> lab = c.buf<std::int32_t>(static_cast<std::size_t>(cg.n));
        c.upload(lab.get(), coarse_labels.data(), coarse_labels.size() * 4);
        const double t_initial = seconds_since(t1);
        const auto t2 = std::chrono::steady_clock::now();
        gp::RefineOptions ro;
        ro.parts = parts;
        ro.imbalance = imbalance;
        gp::UncoarsenResult r = gp::uncoarsen(c, h, lab.get(), ro);
        c.sync();
        const double t_uncoarsen = seconds_since(t2);
        const int levels = h.levels();
        const std::int64_t coarsest_n = cg.n;
        h = gp::Hierarchy();
        lab.reset();
        c.sync();
        const double t_total = seconds_since(t0);
        const std::uint64_t launches = gp::thread_launches() - launches0, syncs = c.syncs() - syncs0;
        if (labels) c.download(labels, r.labels.get(), static_cast<std::size_t>(d.n) * 4);
        if (stats) {
            stats->coarsen_seconds = t_coarsen;
            stats->initial_partition_seconds = t_initial;
            stats->uncoarsen_seconds = t_uncoarsen;
            stats->algorithm_total_seconds = t_total;
            stats->levels = levels;
            stats->coarsest_vertices = coarsest_n;
            stats->final_cut = r.cut;
            std::uint64_t mx = 0, total = 0;
            for (auto x : r.loads) {
                mx = std::max(mx, x);
                total += x;
            }
            stats->max_part_weight = mx;
            stats->capacity = gp::capacity_of(total, parts, imbalance);
            stats->total_edge_weight = gp::device_edge_weight_sum(c, d) / 2;
            stats->launches = launches;
            stats->host_syncs = syncs;
        }
    });
}

static int emit_graph(const gp::HostGraph& h, int64_t* n, int64_t* m, int32_t** off, int32_t** nbr) {
    *n = h.n;
    *m = static_cast<int64_t>(h.nbr.size());
    *off = host_copy(h.off);
    *nbr = host_copy(h.nbr);
    return GP_OK;
}

int gp_generate_grid(int64_t rows, int64_t cols, int64_t* n, int64_t* m, int32_t** offsets, int32_t** neighbors) {
    if (!n || !m || !offsets || !neighbors || rows <= 0 || cols <= 0) return fail(GP_ERR_INVALID, "invalid grid");
    return guarded([&] { emit_graph(gp::generate_grid(rows, cols), n, m, offsets, neighbors); });
}
int gp_generate_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, int64_t* n, int64_t* m,
                     int32_t** offsets, int32_t** neighbors) {
    if (!n || !m || !offsets || !neighbors) return fail(GP_ERR_INVALID, "null output");
    return guarded([&] { emit_graph(gp::generate_rmat(scale, edge_factor, a, b, c, seed), n, m, offsets, neighbors); });
}
int gp_generate_rgg2d(int64_t n_points, double radius, uint64_t seed, int64_t* n, int64_t* m, int32_t** offsets,
                      int32_t** neighbors) {
    if (!n || !m || !offsets || !neighbors || n_points <= 0 || !(radius > 0)) return fail(GP_ERR_INVALID, "invalid rgg");
    return guarded([&] { emit_graph(gp::generate_rgg2d(n_points, radius, seed), n, m, offsets, neighbors); });
}
int gp_generate_knn3d(int64_t n_points, int knn, uint64_t seed, int64_t* n, int64_t* m, int32_t** offsets,
                      int32_t** neighbors, uint32_t** edge_weights, uint32_t** vertex_weights) {
    if (!n || !m || !offsets || !neighbors || !edge_weights || !vertex_weights || n_points <= knn || knn < 1)
        return fail(GP_ERR_INVALID, "invalid knn");
    return guarded([&] {
        gp::HostGraph h = gp::generate_knn3d(n_points, knn, seed);
        emit_graph(h, n, m, offsets, neighbors);
        *edge_weights = host_copy(h.ew);
        *vertex_weights = host_copy(h.vw);
    });
}
void gp_free_host(void* p) { std::free(p); }

int gp_load_csr(const char* indptr_path, const char* indices_path, int64_t* n, int64_t* m, int32_t** offsets,
                int32_t** neighbors) {
    if (!indptr_path || !indices_path || !n || !m || !offsets || !neighbors) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] { emit_graph(gp::load_csr(indptr_path, indices_path), n, m, offsets, neighbors); });
}
int gp_save_csr(const char* indptr_path, const char* indices_path, int64_t n, int64_t m, const int32_t* offsets,
                const int32_t* neighbors) {
    if (!indptr_path || !indices_path || !offsets || (m && !neighbors)) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        gp::HostGraph h;
        h.n = n;
        h.off.assign(offsets, offsets + n + 1);
        h.nbr.assign(neighbors, neighbors + m);
        gp::save_csr(h, indptr_path, indices_path);
    });
}

int gp_profile_enable(gp_context* ctx, int enable) {
    if (!ctx) return fail(GP_ERR_INVALID, "null context");
    ctx->ctx.set_profiling(enable != 0);
    return GP_OK;
}
int gp_profile_query(gp_context* ctx, const char* name, double* ms, double* bytes, int64_t* launches) {
    if (!ctx || !name) return fail(GP_ERR_INVALID, "null argument");
    return guarded([&] {
        auto& p = ctx->ctx.profile();
        auto it = p.find(name);
        if (ms) *ms = it == p.end() ? 0.0 : it->second.ms;
        if (bytes) *bytes = it == p.end() ? 0.0 : it->second.bytes;
        if (launches) *launches = it == p.end() ? 0 : it->second.launches;
    });
}
int gp_profile_reset(gp_context* ctx) {
    if (!ctx) return fail(GP_ERR_INVALID, "null context");
    return guarded([&] { ctx->ctx.profile().clear(); });
}

}  // extern "C"
