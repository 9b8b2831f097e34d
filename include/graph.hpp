// Host and device weighted CSR graphs (include/graph.hpp:15-76 of the
// reference), as header-only C++ over the C ABI of libgpart_mi355x.
//
// WeightedGraph is the reference's host CSR.  DeviceWeightedGraph is a
// handle to a graph resident in HBM (gp_graph): its arrays live in the
// context arena of the thread that made it, and the handle is shared, so a
// hierarchy that took the graph over (coarsen(DeviceWeightedGraph&&)) keeps
// it alive as its level 0.  The metadata fields carry what the reference's
// carry: implicit unit weights, the degree bound, heavy rows and the
// incident-weight bound.
#pragma once

#include "graph_types.hpp"
#include "gpart_session.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace gpart {

struct WeightedGraph {
    using VertexT = gpart::GraphTypes::VertexT;
    using OffsetT = gpart::GraphTypes::OffsetT;
    using WeightT = gpart::GraphTypes::WeightT;

    std::vector<OffsetT> offsets;
    std::vector<VertexT> neighbors;
    std::vector<WeightT> edge_weights;
    std::vector<WeightT> vertex_weights;

    std::int64_t vertices() const { return static_cast<std::int64_t>(vertex_weights.size()); }
    std::int64_t edges() const { return static_cast<std::int64_t>(neighbors.size()); }
};

struct DeviceWeightedGraph {
    using VertexT = gpart::GraphTypes::VertexT;
    using OffsetT = gpart::GraphTypes::OffsetT;
    using WeightT = gpart::GraphTypes::WeightT;

    std::shared_ptr<detail::Session> session;  // context that owns the arrays
    std::shared_ptr<gp_graph> handle;          // empty once moved into coarsen()
    bool implicit_unit_weights = false;
    std::uint64_t maximum_degree = std::numeric_limits<std::uint64_t>::max();
    bool has_heavy_rows = true;
    std::uint64_t incident_weight_bound = std::numeric_limits<std::uint64_t>::max();

    std::int64_t vertices() const {
        std::int64_t n = 0;
        if (handle) detail::check(gp_graph_info(handle.get(), &n, nullptr));
        return n;
    }
    std::int64_t edges() const {
        std::int64_t m = 0;
        if (handle) detail::check(gp_graph_info(handle.get(), nullptr, &m));
        return m;
    }
    // Entries of the stored edge-weight array (0 for implicit unit weights).
    std::int64_t stored_edge_weight_entries() const { return implicit_unit_weights ? 0 : edges(); }
};

// Reads the reference's binary CSR (src/graph.cpp:47-116): little-endian
// int64 indptr[n+1] and indices[m], with its checks (front 0, back m,
// monotone offsets, neighbours in [0, n), 32-bit limits); unit edge and
// vertex weights; prints the reference's "vertices=" line.
inline WeightedGraph load_weighted_graph(const std::string& indptr_path, const std::string& indices_path) {
    std::int64_t n = 0, m = 0;
    std::int32_t *off = nullptr, *nbr = nullptr;
    detail::check(gp_load_csr(indptr_path.c_str(), indices_path.c_str(), &n, &m, &off, &nbr));
    WeightedGraph g;
    g.offsets.assign(off, off + n + 1);
    g.neighbors.assign(nbr, nbr + m);
    gp_free_host(off);
    gp_free_host(nbr);
    g.edge_weights.assign(static_cast<std::size_t>(m), 1u);
    g.vertex_weights.assign(static_cast<std::size_t>(n), 1u);
    std::cout << "vertices=" << n << " edge_entries=" << m << " device_count=1\n";
    return g;
}

// Shape checks of src/graph.cpp:118-155: array sizes, front 0 / back m,
// monotone offsets, neighbours in range, and no self loops unless allowed.
inline void validate_weighted_shape(const WeightedGraph& graph, bool allow_self_loops) {
    if (graph.offsets.size() != graph.vertex_weights.size() + 1 || graph.offsets.empty() ||
        graph.offsets.front() != 0 || static_cast<std::int64_t>(graph.offsets.back()) != graph.edges() ||
        graph.neighbors.size() != graph.edge_weights.size())
        throw std::runtime_error("invalid weighted CSR shape");
    const std::int64_t n = graph.vertices();
    for (std::int64_t v = 0; v < n; ++v)
        if (graph.offsets[static_cast<std::size_t>(v)] > graph.offsets[static_cast<std::size_t>(v + 1)])
            throw std::runtime_error("weighted CSR offsets are not monotone");
    bool bad = false;
    for (const auto u : graph.neighbors) bad |= u < 0 || u >= n;
    if (bad) throw std::runtime_error("weighted CSR contains an invalid neighbor");
    if (allow_self_loops) return;
    for (std::int64_t v = 0; v < n; ++v)
        for (auto e = graph.offsets[static_cast<std::size_t>(v)]; e < graph.offsets[static_cast<std::size_t>(v + 1)]; ++e)
            if (graph.neighbors[static_cast<std::size_t>(e)] == v)
                throw std::runtime_error("contracted graph contains a self loop");
}

// Strict checks of src/graph.cpp:157-186 on top of the shape: every entry
// (v, u) has its reverse (u, v), found by binary search in u's sorted row,
// with the same weight.
inline void validate_weighted_csr(const WeightedGraph& graph, bool allow_self_loops) {
    validate_weighted_shape(graph, allow_self_loops);
    const std::int64_t n = graph.vertices();
    for (std::int64_t v = 0; v < n; ++v) {
        for (auto e = graph.offsets[static_cast<std::size_t>(v)]; e < graph.offsets[static_cast<std::size_t>(v + 1)]; ++e) {
            const auto u = graph.neighbors[static_cast<std::size_t>(e)];
            const auto first = graph.neighbors.begin() + graph.offsets[static_cast<std::size_t>(u)];
            const auto last = graph.neighbors.begin() + graph.offsets[static_cast<std::size_t>(u) + 1];
            const auto back = std::lower_bound(first, last, static_cast<GraphTypes::VertexT>(v));
            if (back == last || *back != v) throw std::runtime_error("weighted CSR is not bidirectional");
            if (graph.edge_weights[static_cast<std::size_t>(e)] !=
                graph.edge_weights[static_cast<std::size_t>(back - graph.neighbors.begin())])
                throw std::runtime_error("weighted CSR reverse edge weights differ");
        }
    }
}

// Exact weighted edge cut of host labels (src/graph.cpp:188-207): the sum of
// the weights of entries whose ends differ, halved; an odd directed sum
// means the CSR is not symmetric.
inline std::uint64_t host_cut(const WeightedGraph& graph, const std::vector<GraphTypes::VertexT>& labels) {
    if (labels.size() != static_cast<std::size_t>(graph.vertices()))
        throw std::invalid_argument("labels do not match the graph");
    std::uint64_t directed = 0;
    for (std::int64_t v = 0; v < graph.vertices(); ++v)
        for (auto e = graph.offsets[static_cast<std::size_t>(v)]; e < graph.offsets[static_cast<std::size_t>(v + 1)]; ++e)
            if (labels[static_cast<std::size_t>(v)] !=
                labels[static_cast<std::size_t>(graph.neighbors[static_cast<std::size_t>(e)])])
                directed += graph.edge_weights[static_cast<std::size_t>(e)];
    if (directed & 1ULL) throw std::runtime_error("weighted CSR cut is not symmetric");
    return directed / 2;
}

namespace detail {

inline std::shared_ptr<gp_graph> own_graph(gp_graph* g) {
    return std::shared_ptr<gp_graph>(g, [](gp_graph* p) { gp_graph_destroy(p); });
}

}  // namespace detail

// Uploads a host CSR into HBM (src/graph.cpp:209-224).  All-ones edge
// weights are always stored implicitly on this device (an all-ones array
// carries nothing the kernels use), so omit_unit_weights only decides what
// implicit_unit_weights reports, as in the reference.
inline DeviceWeightedGraph make_device_weighted(const WeightedGraph& graph, bool omit_unit_weights = true) {
    if (graph.edge_weights.size() != graph.neighbors.size())
        throw std::invalid_argument("device graph requires one weight per edge");
    DeviceWeightedGraph out;
    out.session = detail::session();
    const bool unit = std::all_of(graph.edge_weights.begin(), graph.edge_weights.end(),
                                  [](GraphTypes::WeightT w) { return w == 1; });
    out.implicit_unit_weights = omit_unit_weights && unit;
    gp_graph* g = nullptr;
    detail::check(gp_graph_create(out.session->get(), graph.vertices(), graph.edges(), graph.offsets.data(),
                                  graph.neighbors.data(), unit ? nullptr : graph.edge_weights.data(),
                                  graph.vertex_weights.data(), &g));
    out.handle = detail::own_graph(g);
    return out;
}

// Downloads a device graph into a host CSR (src/graph.cpp:226-246); implicit
// unit weights come back as ones.
inline WeightedGraph copy_device_weighted(const DeviceWeightedGraph& graph) {
    if (!graph.handle || !graph.session) throw std::invalid_argument("invalid device edge-weight storage");
    WeightedGraph out;
    const std::int64_t n = graph.vertices(), m = graph.edges();
    out.offsets.resize(static_cast<std::size_t>(n + 1));
    out.neighbors.resize(static_cast<std::size_t>(m));
    out.edge_weights.resize(static_cast<std::size_t>(m));
    out.vertex_weights.resize(static_cast<std::size_t>(n));
    detail::check(gp_graph_download(graph.session->get(), graph.handle.get(), out.offsets.data(),
                                    out.neighbors.data(), out.edge_weights.data(), out.vertex_weights.data()));
    return out;
}

}  // namespace gpart
