// gpart_partition_standard: the reference's command-line driver
// (app/partition_main.cu:21-183) on the C++ interface headers of this
// library (include/graph.hpp, coarsen.hpp, initial_partition.hpp,
// uncoarsen.hpp, partition_capacity.hpp), which call the gfx950 kernels
// through the C ABI.
//
// Usage: <indptr.bin> <indices.bin> <parts> <partition.out> [imbalance=1.03] [seed=0] [--verify]
// Exit codes: 0 success, 1 exception, 2 usage.  Parsed stdout keys:
// coarsen_mode=, final_edge_cut=, partition_timing ... algorithm_total_seconds=.
// --verify: strict input and per-level coarsening checks, device recounts of
// the cut and loads after every projection and committed refinement batch,
// and a host recount of the final partition.
#include "coarsen.hpp"
#include "graph.hpp"
#include "graph_types.hpp"
#include "initial_partition.hpp"
#include "partition_capacity.hpp"
#include "uncoarsen.hpp"

#include <thrust/copy.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

int main(int argc, char** argv) {
    try {
        const auto wall_start = std::chrono::steady_clock::now();
        bool verify = false;
        int positional_argc = argc;
        while (positional_argc > 1 && std::string(argv[positional_argc - 1]) == "--verify") {
            verify = true;
            --positional_argc;
        }
        if (positional_argc < 5 || positional_argc > 7) {
            std::cerr << "Usage: " << argv[0] << " <indptr.bin> <indices.bin> <parts> <partition.out>"
                      << " [imbalance_ratio=1.03] [seed] [--verify]\n";
            return 2;
        }
        for (const char* retired : {"GPART_REFINE_MODE", "GPART_JOINT_ROUNDS", "GPART_JOINT_CONFLICT_ROUNDS",
                                    "GPART_JOINT_DEDUPLICATE", "GPART_JOINT_COMPLEMENT", "GPART_JOINT_DIAGNOSTICS",
                                    "GPART_SPARSE_DELTA", "GPART_REPAIR_ENUMERATION", "GPART_SPARSE_CUT_BOUND"})
            if (std::getenv(retired))
                throw std::invalid_argument(std::string(retired) + " was retired; unset it. Production uses capacity-exchange.");
        const int parts = std::stoi(argv[3]);
        const double imbalance = positional_argc >= 6 ? std::stod(argv[5]) : 1.03;
        const auto seed = positional_argc >= 7 ? static_cast<std::uint32_t>(std::stoul(argv[6])) : 0U;
        const auto load_start = std::chrono::steady_clock::now();
        auto graph = gpart::load_weighted_graph(argv[1], argv[2]);
        const std::uint64_t capacity =
            gpart::partition_capacity(static_cast<std::uint64_t>(graph.vertices()), parts, imbalance);
        const auto load_end = std::chrono::steady_clock::now();
        std::cout << "verify=" << (verify ? 1 : 0) << '\n';

        gpart::CoarsenOptions coarsen_options;
        coarsen_options.parts = parts;
        coarsen_options.seed = seed;
        coarsen_options.strict_verify = verify;
        coarsen_options.detailed_timings = std::getenv("GPART_GPU_TIMINGS") != nullptr;
        const char* matching = std::getenv("GPART_HEM_COARSEN");
        coarsen_options.seed_matching = !matching || std::string(matching) != "0";
        const char* force_regret = std::getenv("GPART_REGRET_FORCE_FULL");
        coarsen_options.capacity_safe_fast_path = !force_regret || std::string(force_regret) == "0";
        const char* requested_mode = std::getenv("GPART_COARSEN_MODE");
        const std::string coarsen_mode = requested_mode ? requested_mode : "regret";
        if (coarsen_mode != "regret") throw std::runtime_error("Only regret coarsening is maintained");
        coarsen_options.heavy_edge_degree_limit = 64;
        if (const char* limit = std::getenv("GPART_HEM_DEGREE_LIMIT")) {
            coarsen_options.heavy_edge_degree_limit = std::stoi(limit);
            if (coarsen_options.heavy_edge_degree_limit < 2 || coarsen_options.heavy_edge_degree_limit > 256)
                throw std::runtime_error("GPART_HEM_DEGREE_LIMIT must be between 2 and 256");
        }
        // Production schedule (app/partition_main.cu:75-78): beta by k, packing
        // of isolated vertices, progressive capacity for k >= 16.
        coarsen_options.cluster_size_factor =
            parts <= 8 ? std::max(64, (512 + parts - 1) / parts) : std::max(2, (32 + parts - 1) / parts);
        coarsen_options.pack_isolated = true;
        coarsen_options.progressive_cluster_capacity = parts >= 16;
        if (const char* factor = std::getenv("GPART_FINE_FACTOR"))
            coarsen_options.progressive_fine_factor = std::stoi(factor);
        if (const char* growth = std::getenv("GPART_CAPACITY_GROWTH"))
            coarsen_options.progressive_capacity_growth = std::stoi(growth);
        if (const char* beta = std::getenv("GPART_COARSE_BETA")) {
            coarsen_options.cluster_size_factor = std::stoi(beta);
            if (coarsen_options.cluster_size_factor < 2 || coarsen_options.cluster_size_factor > 256)
                throw std::runtime_error("GPART_COARSE_BETA must be between 2 and 256");
        }
        std::cout << "pipeline_mode=incremental_bucket cluster_size_factor=" << coarsen_options.cluster_size_factor
                  << " progressive_cluster_capacity=" << coarsen_options.progressive_cluster_capacity
                  << " progressive_fine_factor=" << coarsen_options.progressive_fine_factor
                  << " progressive_capacity_growth=" << coarsen_options.progressive_capacity_growth
                  << " pack_isolated=" << coarsen_options.pack_isolated << '\n';
        std::cout << "seed_matching=" << coarsen_options.seed_matching
                  << " heavy_edge_degree_limit=" << coarsen_options.heavy_edge_degree_limit
                  << " coarsen_mode=" << coarsen_mode
                  << " initial_partition_backend=" << gp_initial_partition_backend() << '\n';

        const auto prepare_start = std::chrono::steady_clock::now();
        if (verify)
            gpart::validate_weighted_csr(graph, true);
        else
            gpart::validate_weighted_shape(graph, true);
        auto device_graph = gpart::make_device_weighted(graph);
        std::cout << "input_implicit_unit_weights=" << device_graph.implicit_unit_weights
                  << " stored_edge_weight_entries=" << device_graph.stored_edge_weight_entries() << '\n';

        const auto coarsen_start = std::chrono::steady_clock::now();
        auto hierarchy = gpart::coarsen(std::move(device_graph), coarsen_options);
        const auto coarsen_end = std::chrono::steady_clock::now();
        const auto initial_start = coarsen_end;
        auto coarse_partition = gpart::initial_partition_metis(hierarchy.levels.back(), parts, imbalance, seed, false);
        const auto initial_end = std::chrono::steady_clock::now();

        gpart::RefineOptions refine_options;
        std::cout << "refine_mode=capacity-exchange exchange_rounds=" << refine_options.exchange_rounds << '\n';
        refine_options.parts = parts;
        refine_options.allow_initial_imbalance = true;
        refine_options.imbalance_ratio = imbalance;
        refine_options.max_rounds = 32;
        refine_options.strict_verify = verify;
        refine_options.refinement_growth_ratio = 2.0;
        const auto uncoarsen_start = std::chrono::steady_clock::now();
        auto result = gpart::uncoarsen(hierarchy, std::move(coarse_partition), refine_options);
        const auto uncoarsen_end = std::chrono::steady_clock::now();
        hierarchy.levels.clear();
        hierarchy.fine_to_coarse.clear();
        const auto free_end = std::chrono::steady_clock::now();

        std::vector<gpart::GraphTypes::VertexT> partition(result.device_partition.size());
        thrust::copy(result.device_partition.begin(), result.device_partition.end(), partition.begin());
        const auto final_d2h_end = std::chrono::steady_clock::now();
        if (verify) {
            // Independent host recount of the written partition.
            const std::uint64_t recount = gpart::host_cut(graph, partition);
            std::vector<std::uint64_t> loads(static_cast<std::size_t>(parts), 0);
            for (std::size_t v = 0; v < partition.size(); ++v) {
                if (partition[v] < 0 || partition[v] >= parts) throw std::runtime_error("invalid output label");
                loads[static_cast<std::size_t>(partition[v])] += graph.vertex_weights[v];
            }
            if (recount != result.final_cut) throw std::runtime_error("final cut differs from an independent recount");
            if (*std::max_element(loads.begin(), loads.end()) > capacity)
                throw std::runtime_error("final partition violates capacity");
            std::cout << "verify_status=ok cut_recount=" << recount << '\n';
        }

        const auto output_start = std::chrono::steady_clock::now();
        std::ofstream output(argv[4], std::ios::binary | std::ios::trunc);
        if (!output) throw std::runtime_error("cannot open partition output");
        output.write(reinterpret_cast<const char*>(partition.data()),
                     static_cast<std::streamsize>(partition.size() * sizeof(gpart::GraphTypes::VertexT)));
        output.close();
        if (!output) throw std::runtime_error("cannot write partition output");
        const auto output_end = std::chrono::steady_clock::now();
        const auto secs = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
        const double algorithm_seconds = secs(coarsen_start, free_end);
        std::cout << std::setprecision(10) << "partition_timing input_load_seconds=" << secs(load_start, load_end)
                  << " coarsen_seconds=" << secs(coarsen_start, coarsen_end)
                  << " initial_partition_seconds=" << secs(initial_start, initial_end)
                  << " uncoarsen_seconds=" << secs(uncoarsen_start, uncoarsen_end)
                  << " input_prepare_seconds=" << secs(prepare_start, coarsen_start)
                  << " hierarchy_free_seconds=" << secs(uncoarsen_end, free_end)
                  << " final_partition_d2h_seconds=" << secs(free_end, final_d2h_end)
                  << " timing_scope=device_input_to_device_partition"
                  << " partition_output_seconds=" << secs(output_start, output_end)
                  << " partition_total_seconds=" << algorithm_seconds
                  << " algorithm_total_seconds=" << algorithm_seconds
                  << " wall_total_seconds=" << secs(wall_start, output_end) << '\n';
        return 0;
    } catch (const std::exception& error) {
        std::cerr << "gpart_partition: " << error.what() << '\n';
        return 1;
    }
}
