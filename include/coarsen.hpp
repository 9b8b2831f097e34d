// Coarsening (include/coarsen.hpp:10-112 of the reference): options,
// one-level aggregate()/contract() with an SclpWorkspace, and the full level
// loop coarsen() for host and device graphs, as header-only C++ over the C
// ABI of libgpart_mi355x.
//
// Options the reference keeps and this device does not need:
// row_local_contraction (contraction is always exact and row-sorted here;
// both of the reference's paths give the same CSR) and detailed_timings
// (no sub-phase timers).  strict_verify runs validate_coarsening_step
// (src/coarsen.cu:1692-1758) on host copies of every level.
#pragma once

#include "graph.hpp"
#include "hierarchy.hpp"

#include <chrono>
#include <cstdint>
#include <iostream>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace gpart {

inline constexpr std::int64_t kBeta = 256;

struct CoarsenOptions {
    int parts = 4;
    std::uint32_t seed = 0;
    double stop_contraction_ratio = 0.90;
    int max_levels = 24;
    bool strict_verify = false;
    // Cluster capacity is ceil(total_weight / (cluster_size_factor * parts)).
    int cluster_size_factor = kBeta;
    bool pack_isolated = false;
    // Start with fine clusters, then relax the bound as graph size falls.
    bool progressive_cluster_capacity = false;
    int progressive_fine_factor = 64;
    int progressive_capacity_growth = 8;
    bool detailed_timings = true;
    bool row_local_contraction = true;
    bool seed_matching = true;
    bool capacity_safe_fast_path = true;
    int heavy_edge_degree_limit = 64;
};

// Map and coarse vertex weights of one aggregation (held by the library
// until contract() consumes the weights).
struct DeviceAggregateResult {
    std::shared_ptr<detail::Session> session;
    std::shared_ptr<gp_aggregate> handle;
    gpart::GraphTypes::VertexT coarse_vertices = 0;
    std::uint64_t maximum_weight = 0;
    std::uint64_t capacity = 0;

    std::vector<gpart::GraphTypes::VertexT> map_to_host(std::int64_t fine_vertices) const {
        std::vector<gpart::GraphTypes::VertexT> out(static_cast<std::size_t>(fine_vertices));
        detail::check(gp_aggregate_map_download(session->get(), handle.get(), out.data()));
        return out;
    }
};

// Coarsening state carried from a level to the level contract() makes.
class SclpWorkspace {
public:
    SclpWorkspace() : impl(new Impl) { detail::check(gp_workspace_create(&impl->ws)); }
    ~SclpWorkspace() {
        if (impl) gp_workspace_destroy(impl->ws);
    }
    SclpWorkspace(SclpWorkspace&&) noexcept = default;
    SclpWorkspace& operator=(SclpWorkspace&& o) noexcept {
        if (this != &o) {
            if (impl) gp_workspace_destroy(impl->ws);
            impl = std::move(o.impl);
        }
        return *this;
    }
    SclpWorkspace(const SclpWorkspace&) = delete;
    SclpWorkspace& operator=(const SclpWorkspace&) = delete;

    // The arena grows on demand; nothing to reserve ahead.
    void reserve_contract_buffers(std::int64_t) {}

    struct Impl {
        gp_workspace* ws = nullptr;
    };
    std::unique_ptr<Impl> impl;
};

struct SclpStats {
    std::uint64_t maximum_degree = std::numeric_limits<std::uint64_t>::max();
    bool has_heavy_rows = true;
    std::uint64_t capacity = 0;
    std::uint64_t lp_accepted = 0;
    std::uint64_t capacity_rejected = 0;
    std::uint64_t two_hop_merged = 0;
    std::uint64_t singleton_count = 0;
    std::uint64_t proposal_count = 0;
    std::uint64_t positive_gain_vertices = 0;
    std::uint64_t role_blocked_vertices = 0;
    std::uint64_t role_blocked_gain = 0;
    std::uint64_t singleton_with_favorite = 0;
    std::uint64_t pairable_singletons = 0;
    int rounds = 0;
    double affinity_seconds = 0.0;
    double admission_seconds = 0.0;
    double two_hop_seconds = 0.0;
    double compact_seconds = 0.0;
    double aggregate_seconds = 0.0;
};

struct ContractionTimings {
    double total_seconds = 0.0;
    double compact_seconds = 0.0;
    double sort_seconds = 0.0;
    double reduce_seconds = 0.0;
    double csr_build_seconds = 0.0;
    double vertex_weight_seconds = 0.0;
};

namespace detail {

inline void refresh_metadata(DeviceWeightedGraph& g) {
    int heavy = 1, unit = 0;
    check(gp_graph_metadata(g.handle.get(), &g.maximum_degree, &heavy, &g.incident_weight_bound, &unit));
    g.has_heavy_rows = heavy != 0;
    g.implicit_unit_weights = unit != 0;
}

inline gp_coarsen_options to_c(const CoarsenOptions& o) {
    gp_coarsen_options c;
    c.parts = o.parts;
    c.seed = o.seed;
    c.stop_ratio = o.stop_contraction_ratio;
    c.max_levels = o.max_levels;
    c.cluster_size_factor = o.cluster_size_factor;
    c.pack_isolated = o.pack_isolated;
    c.progressive = o.progressive_cluster_capacity;
    c.fine_factor = o.progressive_fine_factor;
    c.capacity_growth = o.progressive_capacity_growth;
    c.seed_matching = o.seed_matching;
    c.fast_path = o.capacity_safe_fast_path;
    c.hem_degree_limit = o.heavy_edge_degree_limit;
    return c;
}

// validate_coarsening_step (src/coarsen.cu:1692-1758): map length and
// range, every coarse vertex covered, cluster weights equal to the sums of
// their members and within the cap, vertex weight conserved, the coarse
// CSR valid (strict: symmetric with equal reverse weights, no self loops),
// and the cut of a hashed 23-colouring equal on both levels.
inline void validate_coarsening_step(const WeightedGraph& fine, const WeightedGraph& coarse,
                                     const std::vector<GraphTypes::VertexT>& map, std::uint64_t cluster_cap, int level,
                                     bool strict_verify) {
    if (map.size() != static_cast<std::size_t>(fine.vertices()))
        throw std::runtime_error("coarsening map has the wrong length");
    if (coarse.vertices() <= 0 || coarse.vertices() > fine.vertices())
        throw std::runtime_error("coarsening produced an invalid vertex count");
    std::vector<std::uint64_t> sums(static_cast<std::size_t>(coarse.vertices()), 0);
    std::vector<std::uint64_t> members(static_cast<std::size_t>(coarse.vertices()), 0);
    for (std::int64_t v = 0; v < fine.vertices(); ++v) {
        const auto c = map[static_cast<std::size_t>(v)];
        if (c < 0 || c >= coarse.vertices()) throw std::runtime_error("coarsening map contains an out-of-range id");
        sums[static_cast<std::size_t>(c)] += fine.vertex_weights[static_cast<std::size_t>(v)];
        ++members[static_cast<std::size_t>(c)];
    }
    const auto total = [](const std::vector<GraphTypes::WeightT>& w) {
        return std::accumulate(w.begin(), w.end(), std::uint64_t{0},
                               [](std::uint64_t s, GraphTypes::WeightT x) { return s + x; });
    };
    if (total(fine.vertex_weights) != total(coarse.vertex_weights))
        throw std::runtime_error("coarsening did not conserve vertex weight");
    for (std::int64_t c = 0; c < coarse.vertices(); ++c) {
        const std::uint64_t w = coarse.vertex_weights[static_cast<std::size_t>(c)];
        if (!members[static_cast<std::size_t>(c)] || sums[static_cast<std::size_t>(c)] != w)
            throw std::runtime_error("coarsening cluster weight mismatch");
        if (w > cluster_cap) throw std::runtime_error("coarse point exceeds configured weight cap");
    }
    if (strict_verify)
        validate_weighted_csr(coarse, false);
    else
        validate_weighted_shape(coarse, false);
    std::vector<GraphTypes::VertexT> coarse_labels(static_cast<std::size_t>(coarse.vertices()));
    for (std::int64_t c = 0; c < coarse.vertices(); ++c)
        coarse_labels[static_cast<std::size_t>(c)] =
            static_cast<GraphTypes::VertexT>((static_cast<std::uint64_t>(c) * 2654435761ULL + 17) % 23);
    std::vector<GraphTypes::VertexT> fine_labels(map.size());
    for (std::size_t v = 0; v < map.size(); ++v) fine_labels[v] = coarse_labels[static_cast<std::size_t>(map[v])];
    if (host_cut(fine, fine_labels) != host_cut(coarse, coarse_labels))
        throw std::runtime_error("cut is not conserved by coarse projection");
    std::cout << "ml_layer_verify level=" << level << " map=ok weights=ok csr=ok projection_cut=ok\n";
}

}  // namespace detail

// One aggregation level (src/coarsen.cu:467-1380): the fine-to-coarse map
// and the coarse vertex weights of `graph` at this level's seed and
// cluster-size factor.  The workspace must be the one whose contract()
// produced `graph` (or fresh, for an input graph).
inline DeviceAggregateResult aggregate(const DeviceWeightedGraph& graph, SclpWorkspace& workspace, int parts,
                                       std::uint32_t seed, int level, SclpStats& stats, bool diagnostics,
                                       int cluster_size_factor = kBeta, bool pack_isolated = false,
                                       bool seed_matching = false, bool capacity_safe_fast_path = true) {
    (void)diagnostics;
    if (!graph.handle) throw std::invalid_argument("aggregate of an empty device graph");
    const auto t0 = std::chrono::steady_clock::now();
    gp_aggregate* a = nullptr;
    detail::check(gp_aggregate_level(graph.session->get(), workspace.impl->ws, graph.handle.get(), parts, seed, level,
                                     cluster_size_factor, pack_isolated, seed_matching, capacity_safe_fast_path, &a));
    DeviceAggregateResult r;
    r.session = graph.session;
    r.handle = std::shared_ptr<gp_aggregate>(a, [](gp_aggregate* p) { gp_aggregate_destroy(p); });
    std::int64_t coarse_n = 0;
    detail::check(gp_aggregate_info(a, &coarse_n, &r.maximum_weight, &r.capacity));
    r.coarse_vertices = static_cast<gpart::GraphTypes::VertexT>(coarse_n);
    int heavy = 1;
    detail::check(gp_graph_metadata(graph.handle.get(), &stats.maximum_degree, &heavy, nullptr, nullptr));
    stats.has_heavy_rows = heavy != 0;
    stats.capacity = r.capacity;
    stats.aggregate_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return r;
}

// Exact contraction of `fine` by an aggregation (src/coarsen.cu:1394-1688):
// the coarse CSR with sorted unique rows and summed weights.
inline DeviceWeightedGraph contract(DeviceWeightedGraph fine, DeviceAggregateResult&& aggregate,
                                    SclpWorkspace& workspace, ContractionTimings& timings) {
    if (!fine.handle || !aggregate.handle) throw std::invalid_argument("contract of an empty graph or aggregate");
    const auto t0 = std::chrono::steady_clock::now();
    gp_graph* c = nullptr;
    detail::check(gp_contract(fine.session->get(), workspace.impl->ws, fine.handle.get(), aggregate.handle.get(), &c));
    DeviceAggregateResult consumed = std::move(aggregate);
    DeviceWeightedGraph out;
    out.session = fine.session;
    out.handle = detail::own_graph(c);
    detail::refresh_metadata(out);
    timings.total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

// The level loop (src/coarsen.cu:1762-2060) on a device graph, which it
// takes over as level 0.
inline DeviceHierarchy coarsen(DeviceWeightedGraph&& graph, const CoarsenOptions& options) {
    if (!graph.handle) throw std::invalid_argument("coarsen of an empty device graph");
    const auto t0 = std::chrono::steady_clock::now();
    std::cout << "ml_input_verify_seconds=0 status=ok mode=" << (options.strict_verify ? "strict" : "fast") << '\n';
    DeviceWeightedGraph input = std::move(graph);
    const gp_coarsen_options o = detail::to_c(options);
    gp_hierarchy* raw = nullptr;
    detail::check(gp_coarsen(input.session->get(), input.handle.get(), &o, &raw));
    // The library borrows level 0: the hierarchy object keeps it alive.
    std::shared_ptr<gp_graph> level0 = input.handle;
    std::shared_ptr<gp_hierarchy> h(raw, [level0](gp_hierarchy* p) { gp_hierarchy_destroy(p); });
    DeviceHierarchy out;
    out.handle = h;
    out.stop_reason = gp_hierarchy_stop_reason(raw);
    const int levels = gp_hierarchy_levels(raw);
    for (int i = 0; i < levels; ++i) {
        DeviceWeightedGraph level;
        level.session = input.session;
        level.handle = std::shared_ptr<gp_graph>(h, const_cast<gp_graph*>(gp_hierarchy_level(raw, i)));
        detail::refresh_metadata(level);
        out.levels.push_back(std::move(level));
        if (i + 1 < levels) {
            DeviceLevelMap map;
            map.session = input.session;
            map.hierarchy = h;
            map.level = i;
            map.length = out.levels.back().vertices();
            out.fine_to_coarse.push_back(std::move(map));
            std::uint64_t cap = 0;
            detail::check(gp_hierarchy_level_capacity(raw, i, &cap));
            out.cluster_caps.push_back(cap);
        }
    }
    input.handle.reset();
    if (options.strict_verify) {
        WeightedGraph fine = copy_device_weighted(out.levels.front());
        for (int i = 0; i + 1 < levels; ++i) {
            WeightedGraph coarse = copy_device_weighted(out.levels[static_cast<std::size_t>(i + 1)]);
            detail::validate_coarsening_step(fine, coarse, out.fine_to_coarse[static_cast<std::size_t>(i)].to_host(),
                                             out.cluster_caps[static_cast<std::size_t>(i)], i, true);
            fine = std::move(coarse);
        }
    }
    out.hierarchy_loop_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "ml_hierarchy_loop_seconds=" << out.hierarchy_loop_seconds
              << " includes_snapshot_and_verify=" << (options.strict_verify ? 1 : 0) << " method=regret_capacity\n";
    return out;
}

// The same on a host graph: validated (strict: validate_weighted_csr),
// uploaded, coarsened, and every level and map copied back.
inline Hierarchy coarsen(const WeightedGraph& graph, const CoarsenOptions& options) {
    if (options.strict_verify)
        validate_weighted_csr(graph, true);
    else
        validate_weighted_shape(graph, true);
    DeviceHierarchy device = coarsen(make_device_weighted(graph), options);
    return copy_device_hierarchy(device);
}

}  // namespace gpart
