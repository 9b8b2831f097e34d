// Refinement (include/refine.hpp:11-98 of the reference): options, per-level
// results, refine_partition on a host graph and refine_hierarchy_device on
// host or device hierarchies, as header-only C++ over the C ABI of
// libgpart_mi355x.
//
// Options the reference keeps that select among its own implementations
// (affinity_layout, repair_enumeration, sparse_cut_bound) have one
// implementation here and are accepted and ignored; they do not change
// results in the reference either.  gpu_lp_loop selects, as in the
// reference, whether the LP round loop is decided on the device (true) or
// by the host after a read-back per round (false); both give the same
// partition.  strict_verify recounts the cut and the loads after every
// projection and committed batch on the device (with the host loop).
#pragma once

#include "graph.hpp"
#include "hierarchy.hpp"
#include "partition_capacity.hpp"

#include <thrust/device_vector.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace gpart {

enum class AffinityLayout { Automatic, Dense, Sparse, OnDemand };

struct RefineOptions {
    int parts = 4;
    double imbalance_ratio = 1.03;
    int max_rounds = 32;
    bool strict_verify = false;
    bool allow_initial_imbalance = false;
    int stale_rounds = 3;
    double refinement_growth_ratio = 2.0;
    int exchange_rounds = 2;
    AffinityLayout affinity_layout = AffinityLayout::Automatic;
    bool gpu_lp_loop = true;
    bool repair_enumeration = true;
    bool sparse_cut_bound = true;
};

struct RefineStats {
    bool initial_repair_failed = false;
    int rounds = 0;
    std::uint64_t proposals = 0;
    std::uint64_t accepted = 0;
    std::uint64_t initial_cut = 0;
    std::uint64_t final_cut = 0;
    std::uint64_t final_max_part_weight = 0;
    double seconds = 0.0;
    int admission_iterations = 0;
    int admission_fallbacks = 0;
    int affinity_counter_bits = 0;
    bool sparse_affinity = false;
    bool on_demand_affinity = false;
    std::uint64_t frontier_vertices_initial = 0;
    std::uint64_t affinity_bytes = 0;
    std::uint64_t lp_cut_gain = 0;
    std::uint64_t lp_cut_loss = 0;
    std::uint64_t rebalance_cut_gain = 0;
    std::uint64_t rebalance_cut_loss = 0;
    std::uint64_t checkpoint_discarded_cut = 0;
    int checkpoint_updates = 0;
    std::uint64_t group_cut_gain = 0;
    std::uint64_t group_accepted = 0;
    // This build: LP, repair and exchange moves apart, and the strict recounts.
    std::uint64_t lp_moves = 0;
    std::uint64_t repair_moves = 0;
    std::uint64_t strict_checks = 0;
};

struct RefineLevelTimings {
    double workspace_resize_seconds = 0.0;
    double graph_h2d_seconds = 0.0;
    double projection_gpu_seconds = 0.0;
    double refinement_seconds = 0.0;
    double verification_seconds = 0.0;
};

struct RefineLevelResult {
    std::size_t level = 0;
    std::int64_t vertices = 0;
    std::int64_t edge_entries = 0;
    std::uint64_t projection_cut = 0;
    std::uint64_t projection_max_part_weight = 0;
    double projection_imbalance = 0.0;
    RefineStats refinement;
    RefineLevelTimings timings;
};

struct DeviceUncoarsenResult {
    std::vector<gpart::GraphTypes::VertexT> partition;
    thrust::device_vector<gpart::GraphTypes::VertexT> device_partition;
    std::vector<RefineLevelResult> levels;
    std::vector<std::uint64_t> final_part_weights;
    std::uint64_t final_cut = 0;
    std::uint64_t total_edge_weight = 0;
    double final_d2h_seconds = 0.0;
};

namespace detail {

inline gp_refine_options to_c(const RefineOptions& o) {
    gp_refine_options c;
    gp_default_refine_options(o.parts, o.imbalance_ratio, &c);
    c.max_rounds = o.max_rounds;
    c.stale_rounds = o.stale_rounds;
    c.growth_ratio = o.refinement_growth_ratio;
    c.exchange_rounds = o.exchange_rounds;
    c.strict_verify = o.strict_verify ? 1 : 0;
    c.gpu_lp_loop = o.gpu_lp_loop ? 1 : 0;
    return c;
}

inline RefineStats to_stats(const gp_level_stats& s) {
    RefineStats r;
    r.initial_repair_failed = s.initial_repair_failed != 0;
    r.rounds = s.rounds;
    r.initial_cut = s.initial_cut;
    r.final_cut = s.final_cut;
    r.accepted = s.lp_moves + s.repair_moves + s.group_moves;
    r.admission_iterations = s.repair_iterations;
    r.group_accepted = s.group_moves;
    r.lp_moves = s.lp_moves;
    r.repair_moves = s.repair_moves;
    r.strict_checks = s.strict_checks;
    return r;
}

inline void check_refine_options(const RefineOptions& o) {
    if (o.parts < 2 || o.parts > kMaxParts || o.max_rounds < 0 || o.max_rounds > 64 || o.stale_rounds < 1 ||
        o.stale_rounds > 64 || o.exchange_rounds < 0 || o.exchange_rounds > 8 ||
        !(o.refinement_growth_ratio >= 1.0) || !(o.imbalance_ratio >= 1.0))
        throw std::invalid_argument("invalid refinement options");
}

// Part loads of labels on a host graph, with the label range checked.
inline std::vector<std::uint64_t> host_loads(const std::vector<GraphTypes::WeightT>& vertex_weights,
                                             const std::vector<GraphTypes::VertexT>& labels, int parts,
                                             const char* what) {
    if (labels.size() != vertex_weights.size()) throw std::invalid_argument(what);
    std::vector<std::uint64_t> loads(static_cast<std::size_t>(parts), 0);
    for (std::size_t v = 0; v < labels.size(); ++v) {
        if (labels[v] < 0 || labels[v] >= parts) throw std::invalid_argument(what);
        loads[static_cast<std::size_t>(labels[v])] += vertex_weights[v];
    }
    return loads;
}

// Runs gp_uncoarsen on a library hierarchy and fills the reference's result.
inline DeviceUncoarsenResult run_uncoarsen(gp_context* ctx, gp_hierarchy* h, const gp_graph* level0,
                                           const std::vector<GraphTypes::VertexT>& coarsest,
                                           const RefineOptions& options, bool device_output) {
    check_refine_options(options);
    const int levels = gp_hierarchy_levels(h);
    const gp_graph* coarse = gp_hierarchy_level(h, levels - 1);
    std::int64_t nc = 0;
    check(gp_graph_info(coarse, &nc, nullptr));
    if (coarsest.size() != static_cast<std::size_t>(nc)) throw std::invalid_argument("invalid coarsest device partition");
    if (!options.allow_initial_imbalance) {
        std::vector<GraphTypes::WeightT> vw(static_cast<std::size_t>(nc));
        check(gp_graph_download(ctx, coarse, nullptr, nullptr, nullptr, vw.data()));
        const auto loads = host_loads(vw, coarsest, options.parts, "invalid coarsest device partition");
        std::uint64_t total = 0;
        for (auto x : loads) total += x;
        if (*std::max_element(loads.begin(), loads.end()) > partition_capacity(total, options.parts, options.imbalance_ratio))
            throw std::invalid_argument("initial refinement partition is imbalanced");
    }
    const gp_refine_options o = to_c(options);
    gp_result* raw = nullptr;
    check(gp_uncoarsen(ctx, h, coarsest.data(), &o, &raw));
    std::unique_ptr<gp_result, void (*)(gp_result*)> r(raw, gp_result_destroy);
    DeviceUncoarsenResult out;
    out.final_cut = gp_result_cut(raw);
    out.final_part_weights.resize(static_cast<std::size_t>(options.parts));
    check(gp_result_loads(raw, out.final_part_weights.data()));
    check(gp_graph_edge_weight(ctx, level0, &out.total_edge_weight));
    const int count = gp_result_levels(raw);
    for (int i = 0; i < count; ++i) {
        gp_level_stats s;
        check(gp_result_level_stats(raw, i, &s));
        RefineLevelResult lr;
        lr.level = static_cast<std::size_t>(levels - 1 - i);
        lr.vertices = s.n;
        lr.edge_entries = s.m;
        lr.projection_cut = s.initial_cut;
        lr.refinement = to_stats(s);
        out.levels.push_back(lr);
    }
    if (!out.levels.empty())
        out.levels.back().refinement.final_max_part_weight =
            *std::max_element(out.final_part_weights.begin(), out.final_part_weights.end());
    std::int64_t n0 = 0;
    check(gp_graph_info(level0, &n0, nullptr));
    if (device_output) {
        out.device_partition.resize(static_cast<std::size_t>(n0));
        if (hipMemcpy(thrust::raw_pointer_cast(out.device_partition.data()), gp_result_device_labels(raw),
                      static_cast<std::size_t>(n0) * sizeof(GraphTypes::VertexT), hipMemcpyDeviceToDevice) != hipSuccess)
            throw std::runtime_error("copy of the device partition failed");
    } else {
        const auto t0 = std::chrono::steady_clock::now();
        out.partition.resize(static_cast<std::size_t>(n0));
        check(gp_result_labels(ctx, raw, out.partition.data()));
        out.final_d2h_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return out;
}

// Uploads a host hierarchy into a library hierarchy that owns the levels.
inline std::shared_ptr<gp_hierarchy> upload_hierarchy(gp_context* ctx, const Hierarchy& hierarchy) {
    if (hierarchy.levels.empty() || hierarchy.fine_to_coarse.size() + 1 != hierarchy.levels.size())
        throw std::invalid_argument("resident refinement requires a complete hierarchy");
    std::vector<gp_graph*> graphs;
    try {
        for (const auto& level : hierarchy.levels) {
            validate_weighted_shape(level, false);
            const bool unit = std::all_of(level.edge_weights.begin(), level.edge_weights.end(),
                                          [](GraphTypes::WeightT w) { return w == 1; });
            gp_graph* g = nullptr;
            check(gp_graph_create(ctx, level.vertices(), level.edges(), level.offsets.data(), level.neighbors.data(),
                                  unit ? nullptr : level.edge_weights.data(), level.vertex_weights.data(), &g));
            graphs.push_back(g);
        }
        std::vector<const std::int32_t*> maps;
        for (const auto& m : hierarchy.fine_to_coarse) maps.push_back(m.data());
        gp_hierarchy* h = nullptr;
        check(gp_hierarchy_assemble(ctx, static_cast<int>(graphs.size()), graphs.data(), maps.data(), &h));
        return std::shared_ptr<gp_hierarchy>(h, [](gp_hierarchy* p) { gp_hierarchy_destroy(p); });
    } catch (...) {
        for (gp_graph* g : graphs) gp_graph_destroy(g);
        throw;
    }
}

}  // namespace detail

// One refinement level on a host graph (src/refine.cu:470-481): labels
// in/out; throws when the initial capacity repair fails.
inline void refine_partition(const WeightedGraph& graph, std::vector<gpart::GraphTypes::VertexT>& partition,
                             const RefineOptions& options, RefineStats* stats = nullptr) {
    detail::check_refine_options(options);
    validate_weighted_shape(graph, false);
    const auto loads = detail::host_loads(graph.vertex_weights, partition, options.parts,
                                          "refinement partition has an invalid part id");
    if (!options.allow_initial_imbalance) {
        std::uint64_t total = 0;
        for (auto x : loads) total += x;
        if (*std::max_element(loads.begin(), loads.end()) > partition_capacity(total, options.parts, options.imbalance_ratio))
            throw std::invalid_argument("initial refinement partition is imbalanced");
    }
    DeviceWeightedGraph device = make_device_weighted(graph);
    const gp_refine_options o = detail::to_c(options);
    gp_level_stats s;
    const auto t0 = std::chrono::steady_clock::now();
    detail::check(gp_refine_partition(device.session->get(), device.handle.get(), partition.data(), &o, &s));
    if (s.initial_repair_failed) throw std::runtime_error("initial refinement capacity repair failed");
    if (stats) {
        *stats = detail::to_stats(s);
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
}

// Uncoarsening of a host hierarchy: the partition comes back on the host.
inline DeviceUncoarsenResult refine_hierarchy_device(const Hierarchy& hierarchy,
                                                     const std::vector<gpart::GraphTypes::VertexT>& coarsest_partition,
                                                     const RefineOptions& options) {
    auto session = detail::session();
    auto h = detail::upload_hierarchy(session->get(), hierarchy);
    return detail::run_uncoarsen(session->get(), h.get(), gp_hierarchy_level(h.get(), 0), coarsest_partition, options,
                                 false);
}

// Uncoarsening of a device hierarchy (src/refine.cu:483-730): consumes the
// hierarchy's levels and maps and the coarsest labels; the partition stays
// on the device (device_partition).
inline DeviceUncoarsenResult refine_hierarchy_device(
    DeviceHierarchy& hierarchy, thrust::device_vector<gpart::GraphTypes::VertexT>&& coarsest_partition,
    const RefineOptions& options) {
    if (!hierarchy.handle || hierarchy.levels.empty() || hierarchy.fine_to_coarse.size() + 1 != hierarchy.levels.size())
        throw std::invalid_argument("resident refinement requires a complete hierarchy");
    std::vector<gpart::GraphTypes::VertexT> coarsest(coarsest_partition.size());
    thrust::copy(coarsest_partition.begin(), coarsest_partition.end(), coarsest.begin());
    thrust::device_vector<gpart::GraphTypes::VertexT>{}.swap(coarsest_partition);
    auto session = hierarchy.levels.front().session;
    std::shared_ptr<gp_hierarchy> h = hierarchy.handle;
    DeviceUncoarsenResult out =
        detail::run_uncoarsen(session->get(), h.get(), gp_hierarchy_level(h.get(), 0), coarsest, options, true);
    hierarchy.levels.clear();
    hierarchy.fine_to_coarse.clear();
    hierarchy.handle.reset();
    return out;
}

}  // namespace gpart
