// gpart_partition_standard: drop-in CLI of the reference (app/partition_main.cu:21-183).
// Usage: <indptr.bin> <indices.bin> <parts> <partition.out> [imbalance=1.03] [seed=0] [--verify]
// Exit codes: 0 success, 1 exception, 2 usage.
#include "gpart_mi355x.h"

#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {
void check(int rc) {
    if (rc != GP_OK) throw std::runtime_error(gp_last_error());
}
}  // namespace

int main(int argc, char** argv) {
    try {
        const auto wall_start = std::chrono::steady_clock::now();
        bool verify = false;
        int positional = argc;
        while (positional > 1 && std::string(argv[positional - 1]) == "--verify") {
            verify = true;
            --positional;
        }
        if (positional < 5 || positional > 7) {
            std::cerr << "Usage: " << argv[0]
                      << " <indptr.bin> <indices.bin> <parts> <partition.out> [imbalance_ratio=1.03] [seed] [--verify]\n";
            return 2;
        }
        for (const char* retired : {"GPART_REFINE_MODE", "GPART_JOINT_ROUNDS", "GPART_JOINT_CONFLICT_ROUNDS",
                                    "GPART_JOINT_DEDUPLICATE", "GPART_JOINT_COMPLEMENT", "GPART_JOINT_DIAGNOSTICS",
                                    "GPART_SPARSE_DELTA", "GPART_REPAIR_ENUMERATION", "GPART_SPARSE_CUT_BOUND"})
            if (std::getenv(retired)) throw std::invalid_argument(std::string(retired) + " was retired; unset it.");
        const int parts = std::stoi(argv[3]);
        const double imbalance = positional >= 6 ? std::stod(argv[5]) : 1.03;
        const auto seed = positional >= 7 ? static_cast<std::uint32_t>(std::stoul(argv[6])) : 0U;
        const auto load_start = std::chrono::steady_clock::now();
        int64_t n = 0, m = 0;
        int32_t *off = nullptr, *nbr = nullptr;
        check(gp_load_csr(argv[1], argv[2], &n, &m, &off, &nbr));
        std::uint64_t cap = 0;
        check(gp_capacity(static_cast<std::uint64_t>(n), parts, imbalance, &cap));
        const auto load_end = std::chrono::steady_clock::now();
        std::cout << "vertices=" << n << " edge_entries=" << m << " device_count=1\n";
        std::cout << "verify=" << (verify ? 1 : 0) << '\n';
        std::cout << "coarsen_mode=regret initial_partition_backend=" << gp_initial_partition_backend() << '\n';
        gp_context* ctx = nullptr;
        check(gp_context_create(0, &ctx));
        const auto prepare_start = std::chrono::steady_clock::now();
        gp_graph* g = nullptr;
        check(gp_graph_create(ctx, n, m, off, nbr, nullptr, nullptr, &g));
        std::vector<int32_t> labels(static_cast<std::size_t>(n));
        gp_partition_stats st{};
        const auto part_start = std::chrono::steady_clock::now();
        check(gp_partition(ctx, g, parts, imbalance, seed, labels.data(), &st));
        const auto part_end = std::chrono::steady_clock::now();
        std::uint64_t cut_check = 0;
        if (verify) {
            std::vector<std::uint64_t> loads(static_cast<std::size_t>(parts), 0);
            std::uint64_t directed = 0;
            for (int64_t v = 0; v < n; ++v) {
                if (labels[v] < 0 || labels[v] >= parts) throw std::runtime_error("invalid output label");
                ++loads[static_cast<std::size_t>(labels[v])];
                for (int32_t e = off[v]; e < off[v + 1]; ++e) directed += labels[v] != labels[nbr[e]];
            }
            cut_check = directed / 2;
            if (cut_check != st.final_cut) throw std::runtime_error("final cut differs from an independent recount");
            for (auto l : loads)
                if (l > cap) throw std::runtime_error("final partition violates capacity");
            std::cout << "verify_status=ok cut_recount=" << cut_check << '\n';
        }
        const auto output_start = std::chrono::steady_clock::now();
        std::ofstream out(argv[4], std::ios::binary | std::ios::trunc);
        if (!out) throw std::runtime_error("cannot open partition output");
        out.write(reinterpret_cast<const char*>(labels.data()), static_cast<std::streamsize>(labels.size() * 4));
        out.close();
        if (!out) throw std::runtime_error("cannot write partition output");
        const auto output_end = std::chrono::steady_clock::now();
        auto secs = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
        std::cout << std::setprecision(10) << "final_edge_cut=" << st.final_cut << '\n'
                  << "final_cut_ratio=" << (m ? (double)st.final_cut / (double)(m / 2) : 0.0) << '\n'
                  << "final_max_part_weight=" << st.max_part_weight << '\n'
                  << "final_imbalance=" << (double)st.max_part_weight * parts / (double)n << '\n'
                  << "partition_timing input_load_seconds=" << secs(load_start, load_end)
                  << " coarsen_seconds=" << st.coarsen_seconds
                  << " initial_partition_seconds=" << st.initial_partition_seconds
                  << " uncoarsen_seconds=" << st.uncoarsen_seconds
                  << " input_prepare_seconds=" << secs(prepare_start, part_start)
                  << " timing_scope=device_input_to_device_partition"
                  << " partition_output_seconds=" << secs(output_start, output_end)
                  << " partition_total_seconds=" << st.algorithm_total_seconds
                  << " algorithm_total_seconds=" << st.algorithm_total_seconds
                  << " wall_total_seconds=" << secs(wall_start, output_end) << '\n';
        (void)part_end;
        gp_graph_destroy(g);
        gp_context_destroy(ctx);
        gp_free_host(off);
        gp_free_host(nbr);
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "gpart_partition: " << e.what() << '\n';
        return 1;
    }
}
