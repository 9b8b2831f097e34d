// Refinement entry points (include/refine.hpp:84-98, include/uncoarsen.hpp:17-20).
#pragma once
#include "gp_graph.hpp"

namespace gp {

struct UncoarsenResult {
    DBuf<std::int32_t> labels;  // original-graph labels on device
    std::vector<std::uint64_t> loads;
    std::uint64_t cut = 0;
    std::vector<LevelStats> levels;  // coarsest first
};

// Projection fine[v] = coarse[map[v]] (src/refine.cu:100-111).
void project(Context& c, std::int64_t n, const std::int32_t* map, const std::int32_t* coarse, std::int32_t* fine);

// Exact directed cut / 2 and part loads.
std::uint64_t device_cut(Context& c, const DGraph& g, const std::int32_t* labels);
// Directed edge-weight sum (m for implicit unit weights).
std::uint64_t device_edge_weight_sum(Context& c, const DGraph& g);
void device_loads(Context& c, const DGraph& g, const std::int32_t* labels, int k, std::uint64_t* loads_host);

// One refinement level on device labels (run_incremental_refinement).
LevelStats refine_level(Context& c, const DGraph& g, std::int32_t* labels, std::vector<std::uint64_t>& loads,
                        const RefineOptions& o, std::uint64_t cap, int rounds, bool has_known_cut,
                        std::uint64_t known_cut);

// Coarsest refinement, then project and refine level by level.
UncoarsenResult uncoarsen(Context& c, const Hierarchy& h, const std::int32_t* coarsest_labels, const RefineOptions& o);

// cut < 0.999 * best as x87 long double evaluates it, in integers (the
// device loop control's form of the stale rule of small levels).
bool debug_below_0999(std::uint64_t cut, std::uint64_t best);

}  // namespace gp
