// Device update of a CSR by an edge delta (resident graphs of the server).
#pragma once
#include "gp_graph.hpp"

namespace gp {

// Applies count undirected edge changes (u[i], v[i]): insert[i] != 0 inserts
// with weight w[i] (w null: 1), else deletes.  Replaces g's arrays (new ones
// come from the arena of g's own arrays); returns the new entry count.
std::int64_t apply_delta(Context& c, DGraph& g, std::int64_t count, const std::int32_t* u, const std::int32_t* v,
                         const std::uint32_t* w, const std::uint8_t* insert);

}  // namespace gp
