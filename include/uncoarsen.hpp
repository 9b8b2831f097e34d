// Uncoarsening (include/uncoarsen.hpp:9-20 of the reference): project the
// coarsest partition level by level and refine, printing the reference's
// per-level "uncoarsen_level=" lines and the final "final_edge_cut=" block
// (src/uncoarsen.cu:50-110) that run harnesses parse.
#pragma once

#include "hierarchy.hpp"
#include "refine.hpp"

#include <chrono>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <vector>

namespace gpart {

namespace detail {

inline void report_uncoarsen(const DeviceUncoarsenResult& result, std::chrono::steady_clock::time_point start) {
    for (const auto& level : result.levels) {
        const auto& s = level.refinement;
        std::cout << std::setprecision(10) << "uncoarsen_level=" << level.level << " vertices=" << level.vertices
                  << " edge_entries=" << level.edge_entries << " projection_cut=" << level.projection_cut
                  << " refine_rounds=" << s.rounds << " refine_accepted=" << s.accepted
                  << " lp_moves=" << s.lp_moves << " repair_moves=" << s.repair_moves
                  << " group_accepted=" << s.group_accepted << " admission_iterations=" << s.admission_iterations
                  << " initial_repair_failed=" << s.initial_repair_failed << " strict_checks=" << s.strict_checks
                  << " refined_cut=" << s.final_cut << '\n';
    }
    const std::uint64_t total = std::accumulate(result.final_part_weights.begin(), result.final_part_weights.end(),
                                                std::uint64_t{0});
    const std::uint64_t maximum =
        result.final_part_weights.empty()
            ? 0
            : *std::max_element(result.final_part_weights.begin(), result.final_part_weights.end());
    const double imbalance =
        total == 0 ? 0.0 : static_cast<double>(maximum) * result.final_part_weights.size() / static_cast<double>(total);
    std::cout << std::setprecision(10) << "final_edge_cut=" << result.final_cut << '\n'
              << "final_cut_ratio="
              << (result.total_edge_weight == 0 ? 0.0
                                                : static_cast<double>(result.final_cut) / result.total_edge_weight)
              << '\n'
              << "final_max_part_weight=" << maximum << '\n'
              << "final_imbalance=" << imbalance << '\n'
              << "final_d2h_seconds=" << result.final_d2h_seconds << '\n'
              << "uncoarsen_total_seconds="
              << std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count() << '\n';
}

}  // namespace detail

// Host hierarchy in, host partition out.
inline std::vector<gpart::GraphTypes::VertexT> uncoarsen(const Hierarchy& hierarchy,
                                                          const std::vector<gpart::GraphTypes::VertexT>& coarsest_partition,
                                                          const RefineOptions& options) {
    const auto start = std::chrono::steady_clock::now();
    if (hierarchy.levels.empty() || hierarchy.fine_to_coarse.size() + 1 != hierarchy.levels.size())
        throw std::invalid_argument("uncoarsen requires a complete hierarchy");
    DeviceUncoarsenResult result = refine_hierarchy_device(hierarchy, coarsest_partition, options);
    detail::report_uncoarsen(result, start);
    return std::move(result.partition);
}

// Consumes the device levels, maps and coarsest labels; the result keeps the
// final partition on the device.
inline DeviceUncoarsenResult uncoarsen(DeviceHierarchy& hierarchy,
                                       thrust::device_vector<gpart::GraphTypes::VertexT>&& coarsest_partition,
                                       const RefineOptions& options) {
    const auto start = std::chrono::steady_clock::now();
    if (hierarchy.levels.empty() || hierarchy.fine_to_coarse.size() + 1 != hierarchy.levels.size())
        throw std::invalid_argument("uncoarsen requires a complete hierarchy");
    DeviceUncoarsenResult result = refine_hierarchy_device(hierarchy, std::move(coarsest_partition), options);
    detail::report_uncoarsen(result, start);
    return result;
}

}  // namespace gpart
