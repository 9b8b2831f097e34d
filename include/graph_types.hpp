// The single graph type of the partitioner (include/graph_types.hpp:1-10 of
// the reference, kept as is: callers name these types): int32 vertices and
// offsets, uint32 vertex and edge weights; cut, gain and load sums are
// 64-bit; at most 128 parts.
#pragma once
#include <cstdint>

namespace gpart {

inline constexpr int kMaxParts = 128;

struct GraphTypes {
    using VertexT = std::int32_t;
    using OffsetT = std::int32_t;
    using WeightT = std::uint32_t;
};

inline constexpr const char* kActiveGraphName = "standard";

}  // namespace gpart
