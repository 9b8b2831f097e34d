// Coarsening hierarchies (include/hierarchy.hpp:15-37 of the reference) and
// the Jet-compatible int32 hierarchy file (write_hierarchy,
// include/hierarchy.hpp:40-123; the on-disk layout is the reference's, so
// files interchange with tools that read it):
//
//   int32 level_count, then per level i:
//     int32 n, int32 m, int32 offsets[n+1], int32 neighbors[m],
//     int32 edge_weights[m], int32 vertex_weights[n],
//     and for i > 0: int32 fine_to_coarse[n_{i-1}] (level i-1 -> level i).
//
// DeviceHierarchy holds the device levels as DeviceWeightedGraph handles
// that share ownership of the library's hierarchy object; level 0 is the
// graph coarsen() took over.
#pragma once

#include "graph.hpp"

#include <cstdio>
#include <cstdint>
#include <fstream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace gpart {

struct Hierarchy {
    using VertexT = gpart::GraphTypes::VertexT;

    std::vector<WeightedGraph> levels;
    std::vector<std::vector<VertexT>> fine_to_coarse;
    std::string stop_reason = "capacity_floor";
    double hierarchy_loop_seconds = 0.0;
    double snapshot_seconds = 0.0;

    std::size_t layer_count() const { return levels.size(); }
};

// fine_to_coarse map of one device level, readable until the hierarchy is gone.
struct DeviceLevelMap {
    std::shared_ptr<detail::Session> session;
    std::shared_ptr<gp_hierarchy> hierarchy;
    int level = 0;
    std::int64_t length = 0;

    std::size_t size() const { return static_cast<std::size_t>(length); }
    std::vector<GraphTypes::VertexT> to_host() const {
        std::vector<GraphTypes::VertexT> out(size());
        if (length) detail::check(gp_hierarchy_map_download(session->get(), hierarchy.get(), level, out.data()));
        return out;
    }
};

struct DeviceHierarchy {
    std::vector<DeviceWeightedGraph> levels;
    std::vector<DeviceLevelMap> fine_to_coarse;
    std::string stop_reason = "capacity_floor";
    double hierarchy_loop_seconds = 0.0;
    double snapshot_seconds = 0.0;
    // Cluster weight caps of each level's aggregation (levels - 1 entries).
    std::vector<std::uint64_t> cluster_caps;
    std::shared_ptr<gp_hierarchy> handle;

    std::size_t layer_count() const { return levels.size(); }
};

// Host copy of every device level and map.
inline Hierarchy copy_device_hierarchy(const DeviceHierarchy& device) {
    Hierarchy h;
    for (const auto& level : device.levels) h.levels.push_back(copy_device_weighted(level));
    for (const auto& map : device.fine_to_coarse) h.fine_to_coarse.push_back(map.to_host());
    h.stop_reason = device.stop_reason;
    h.hierarchy_loop_seconds = device.hierarchy_loop_seconds;
    return h;
}

namespace detail {

inline void put_i32(std::ofstream& out, std::int64_t v, const char* what) {
    if (v < 0 || v > std::numeric_limits<std::int32_t>::max()) throw std::runtime_error(what);
    const auto x = static_cast<std::int32_t>(v);
    out.write(reinterpret_cast<const char*>(&x), sizeof x);
}

inline std::int32_t get_i32(std::ifstream& in) {
    std::int32_t x = 0;
    in.read(reinterpret_cast<char*>(&x), sizeof x);
    if (!in) throw std::runtime_error("truncated hierarchy file");
    return x;
}

}  // namespace detail

// Writes the hierarchy through a temporary file renamed into place (the
// same checks as the reference: complete hierarchy, 32-bit dimensions,
// offsets and weights within int32, maps of the fine level's length with
// ids inside the coarse level).
inline void write_hierarchy(const Hierarchy& hierarchy, const std::string& path) {
    const auto& levels = hierarchy.levels;
    const auto& maps = hierarchy.fine_to_coarse;
    if (levels.empty() || maps.size() + 1 != levels.size())
        throw std::runtime_error("cannot write an incomplete hierarchy");
    const std::string temporary = path + ".tmp";
    try {
        std::ofstream out(temporary, std::ios::binary | std::ios::trunc);
        if (!out) throw std::runtime_error("cannot open hierarchy temporary file");
        detail::put_i32(out, static_cast<std::int64_t>(levels.size()), "hierarchy export requires 32-bit dimensions");
        for (std::size_t i = 0; i < levels.size(); ++i) {
            const WeightedGraph& g = levels[i];
            if (g.vertices() > std::numeric_limits<std::int32_t>::max() ||
                g.edges() > std::numeric_limits<std::int32_t>::max())
                throw std::runtime_error("hierarchy export requires 32-bit dimensions");
            const std::int64_t n = g.vertices(), m = g.edges();
            detail::put_i32(out, n, "hierarchy export requires 32-bit dimensions");
            detail::put_i32(out, m, "hierarchy export requires 32-bit dimensions");
            for (const auto o : g.offsets) detail::put_i32(out, o, "hierarchy row offset exceeds int32");
            out.write(reinterpret_cast<const char*>(g.neighbors.data()),
                      static_cast<std::streamsize>(m * static_cast<std::int64_t>(sizeof(std::int32_t))));
            for (const auto w : g.edge_weights) detail::put_i32(out, w, "hierarchy edge weight exceeds int32");
            for (const auto w : g.vertex_weights) detail::put_i32(out, w, "hierarchy vertex weight exceeds int32");
            if (i > 0) {
                const auto& map = maps[i - 1];
                if (map.size() != static_cast<std::size_t>(levels[i - 1].vertices()))
                    throw std::runtime_error("hierarchy map has the wrong fine-level length");
                for (const auto c : map) {
                    if (c < 0 || c >= n) throw std::runtime_error("hierarchy map contains an invalid coarse id");
                    out.write(reinterpret_cast<const char*>(&c), sizeof c);
                }
            }
        }
        out.flush();
        if (!out) throw std::runtime_error("failed while writing hierarchy");
        out.close();
        if (std::rename(temporary.c_str(), path.c_str()) != 0) throw std::runtime_error("cannot commit hierarchy output");
    } catch (...) {
        std::remove(temporary.c_str());
        throw;
    }
}

// Reads a file written by write_hierarchy (the inverse; no counterpart in
// the reference, which only writes).
inline Hierarchy read_hierarchy(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    Hierarchy h;
    const std::int32_t count = detail::get_i32(in);
    if (count <= 0) throw std::runtime_error("hierarchy file has no levels");
    for (std::int32_t i = 0; i < count; ++i) {
        WeightedGraph g;
        const std::int32_t n = detail::get_i32(in), m = detail::get_i32(in);
        if (n < 0 || m < 0) throw std::runtime_error("hierarchy file has negative dimensions");
        g.offsets.resize(static_cast<std::size_t>(n) + 1);
        g.neighbors.resize(static_cast<std::size_t>(m));
        g.edge_weights.resize(static_cast<std::size_t>(m));
        g.vertex_weights.resize(static_cast<std::size_t>(n));
        for (auto& o : g.offsets) o = detail::get_i32(in);
        for (auto& x : g.neighbors) x = detail::get_i32(in);
        for (auto& w : g.edge_weights) w = static_cast<GraphTypes::WeightT>(detail::get_i32(in));
        for (auto& w : g.vertex_weights) w = static_cast<GraphTypes::WeightT>(detail::get_i32(in));
        if (i > 0) {
            std::vector<GraphTypes::VertexT> map(static_cast<std::size_t>(h.levels.back().vertices()));
            for (auto& c : map) c = detail::get_i32(in);
            h.fine_to_coarse.push_back(std::move(map));
        }
        h.levels.push_back(std::move(g));
    }
    return h;
}

}  // namespace gpart
