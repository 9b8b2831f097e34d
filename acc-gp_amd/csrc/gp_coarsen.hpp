// Coarsening entry points (include/coarsen.hpp:94-112 in the reference).
#pragma once
#include "gp_graph.hpp"

namespace gp {

struct AggregateResult {
    std::int64_t coarse_n = 0;
    std::uint64_t capacity = 0;
    std::uint64_t max_weight = 0;
    std::uint64_t stats_max_degree = 0;
    bool stats_has_heavy = false;
};

// Level-0 metadata pass (record bounds, row flags, weight totals).
void compute_level0_metadata(Context& c, DGraph& g, CoarsenState& st);

// One aggregation level; fills map (n) and the coarse vertex weights.
AggregateResult aggregate(Context& c, const DGraph& g, CoarsenState& st, int parts, std::uint32_t seed, int level,
                          int level_factor, bool pack_isolated, bool seed_matching, bool fast_path,
                          DBuf<std::int32_t>& map, DBuf<std::uint32_t>& coarse_vw);

// Exact contraction of g by map into nc coarse vertices.
std::unique_ptr<DGraph> contract(Context& c, const DGraph& fine, CoarsenState& st, const std::int32_t* map,
                                 std::int64_t nc, DBuf<std::uint32_t>&& coarse_vw);

// Full level loop; the input graph stays owned by the caller (level 0).
Hierarchy coarsen(Context& c, DGraph& input, const CoarsenOptions& o);

// The device radix sort of the admission and contraction orders on host
// arrays (stable, by bits [begin_bit, end_bit) of the keys, taken in whole
// 8-bit digits from begin_bit, so bits past end_bit in the last digit must
// be zero; vals may be null).  For tests.
void debug_radix_sort(Context& c, std::uint64_t* keys, std::uint32_t* vals, std::int64_t n, int begin_bit,
                      int end_bit);

}  // namespace gp
