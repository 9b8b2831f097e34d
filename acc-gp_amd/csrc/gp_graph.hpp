// Device-resident weighted CSR, hierarchy and option structs.
// Graph type: int32 vertices/offsets, uint32 vertex/edge weights, 64-bit sums
// (include/graph_types.hpp:4-10 in the reference).
#pragma once
#include "gp_prims.hpp"

#include <limits>
#include <memory>
#include <vector>

namespace gp {

struct DGraph {
    std::int64_t n = 0, m = 0;
    DBuf<std::int32_t> off, nbr;
    DBuf<std::uint32_t> ew;  // empty: implicit unit edge weights
    DBuf<std::uint32_t> vw;
    // Metadata carried by coarsening (include/graph.hpp:34-57).
    std::uint64_t max_degree = std::numeric_limits<std::uint64_t>::max();
    bool has_heavy = true;
    std::uint64_t incident_bound = std::numeric_limits<std::uint64_t>::max();
    // Upper bound on the entries of any row (exact at the finest level,
    // the longest raw row for coarse levels); picks the row tiers to launch.
    std::uint64_t row_len_bound = std::numeric_limits<std::uint64_t>::max();

    const std::uint32_t* ew_ptr() const { return ew ? ew.get() : nullptr; }
    bool unit_edges() const { return !ew; }
};

struct CoarsenOptions {
    int parts = 4;
    std::uint32_t seed = 0;
    double stop_ratio = 0.90;
    int max_levels = 24;
    int cluster_size_factor = 256;
    bool pack_isolated = false;
    bool progressive = false;
    int fine_factor = 64;
    int capacity_growth = 8;
    bool seed_matching = true;
    bool fast_path = true;
    int hem_degree_limit = 64;

    // Production schedule of the CLI (app/partition_main.cu:52-87).
    static CoarsenOptions production(int parts, std::uint32_t seed) {
        CoarsenOptions o;
        o.parts = parts;
        o.seed = seed;
        o.cluster_size_factor = parts <= 8 ? std::max(64, (512 + parts - 1) / parts)
                                           : std::max(2, (32 + parts - 1) / parts);
        o.pack_isolated = true;
        o.progressive = parts >= 16;
        return o;
    }
};

struct RefineOptions {
    int parts = 4;
    double imbalance = 1.03;
    int max_rounds = 32;
    int stale_rounds = 3;
    double growth_ratio = 2.0;
    int exchange_rounds = 2;
    std::int64_t graph_loop_threshold = 65536;
    std::int64_t exchange_max_n = 65536;
    // Recount the cut and the loads after every committed batch and every
    // projection (include/refine.hpp:17, the CLI's --verify).
    bool strict_verify = false;
    // LP round loop decided on the device (incremental.cuh:911-961 as
    // lp_graph.cuh records it); false: on the host, one read-back per round.
    bool device_loop = true;
};

// Coarsening state carried across levels (SclpWorkspace::Impl, src/coarsen.cu:299-396).
struct CoarsenState {
    bool sorted_unique_rows = false, positive_ew = true, unit_ew = true;
    std::uint64_t total_weight = 0, max_vertex_weight = 0, max_row_weight = 0;
    DBuf<std::uint64_t> row_bound;  // per vertex of the current level
};

struct LevelStats {
    std::int64_t n = 0, m = 0;
    int refined = 0;
    int initial_repair_failed = 0;
    int rounds = 0;
    std::uint64_t initial_cut = 0, final_cut = 0;
    std::uint64_t lp_moves = 0, repair_moves = 0, group_moves = 0;
    int repair_iterations = 0;
    std::uint64_t strict_checks = 0;  // recounts that matched (strict_verify)
};

class Hierarchy {
public:
    const DGraph* input = nullptr;          // level 0, borrowed
    std::vector<std::unique_ptr<DGraph>> coarse;  // levels 1..
    std::vector<DBuf<std::int32_t>> maps;         // maps[i]: level i -> level i+1
    std::vector<std::uint64_t> caps;              // caps[i]: cluster weight cap of level i's aggregation
    std::string stop_reason = "capacity_floor";

    int levels() const { return 1 + static_cast<int>(coarse.size()); }
    const DGraph& level(int i) const { return i == 0 ? *input : *coarse[static_cast<std::size_t>(i - 1)]; }
};

// Kernel-time accounting for the roofline report (HIP events on the stream).
struct KernelStat {
    double ms = 0.0;
    double bytes = 0.0;
    long launches = 0;
};

}  // namespace gp
