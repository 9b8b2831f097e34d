// partition_capacity (include/partition_capacity.hpp:14-25 of the reference):
// the integer part capacity floor(total * imbalance / parts), through the C
// ABI (gp_capacity).  std::invalid_argument for parts outside [2, 128], an
// imbalance below 1 or not finite, or a capacity that cannot hold
// ceil(total / parts); std::overflow_error when it exceeds uint64.
#pragma once

#include "graph_types.hpp"
#include "gpart_session.hpp"

#include <cstdint>

namespace gpart {

inline std::uint64_t partition_capacity(std::uint64_t total, int parts, double imbalance_ratio) {
    std::uint64_t capacity = 0;
    detail::check(gp_capacity(total, parts, imbalance_ratio, &capacity), /*runtime_is_overflow=*/true);
    return capacity;
}

}  // namespace gpart
