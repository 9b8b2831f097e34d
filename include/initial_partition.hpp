// Initial partition of the coarsest graph on the host
// (include/initial_partition.hpp:1-8 of the reference): the device graph is
// downloaded, partitioned by METIS 5 when the library was built with it and
// by the in-repo recursive bisection otherwise (gp_initial_partition_backend()
// says which), and the labels come back as a device vector.  With
// require_feasible the strict single-move capacity repair follows
// (src/initial_partition_metis.cu:35-83); production passes false and
// leaves coarse imbalance to the refinement's repair.
#pragma once

#include "graph.hpp"

#include <thrust/device_vector.h>

#include <cstdint>
#include <vector>

namespace gpart {

inline thrust::device_vector<gpart::GraphTypes::VertexT> initial_partition_metis(const DeviceWeightedGraph& graph,
                                                                                 int parts, double imbalance_ratio,
                                                                                 std::uint32_t seed = 0,
                                                                                 bool require_feasible = true) {
    if (!graph.handle || !graph.session) throw std::invalid_argument("initial partition of an empty device graph");
    std::vector<gpart::GraphTypes::VertexT> labels(static_cast<std::size_t>(graph.vertices()));
    detail::check(gp_initial_partition(graph.session->get(), graph.handle.get(), parts, imbalance_ratio, seed,
                                       require_feasible ? 1 : 0, labels.data()));
    return thrust::device_vector<gpart::GraphTypes::VertexT>(labels.begin(), labels.end());
}

}  // namespace gpart
