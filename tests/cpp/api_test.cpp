// Compiled test of the C++ interface headers (include/*.hpp), the drop-in
// surface of the reference's include/ directory.
//
//   gpart_api_test cpu <dir>   host-side API: capacity rule and errors, the
//                              int64 CSR reader and its checks, shape and
//                              strict CSR validation, host_cut, and the
//                              hierarchy file round trip
//   gpart_api_test gpu <dir>   device API: coarsen (strict) -> initial
//                              partition -> uncoarsen against gp_partition;
//                              aggregate/contract level by level against
//                              coarsen; host-hierarchy uncoarsening and
//                              refine_partition against the device forms
//
// Prints one "ok <check>" line per check; exit 0 when all pass.
#include "coarsen.hpp"
#include "graph.hpp"
#include "hierarchy.hpp"
#include "initial_partition.hpp"
#include "partition_capacity.hpp"
#include "refine.hpp"
#include "uncoarsen.hpp"

#include <thrust/copy.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iostream>
#include <string>
#include <vector>

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
    std::cout << (ok ? "ok " : "FAIL ") << what << std::endl;
    if (!ok) ++failures;
}

template <typename E>
void expect_throw(const std::function<void()>& fn, const std::string& what, const std::string& needle = "") {
    try {
        fn();
    } catch (const E& e) {
        expect(needle.empty() || std::string(e.what()).find(needle) != std::string::npos,
               what + " (" + e.what() + ")");
        return;
    } catch (const std::exception& e) {
        expect(false, what + " threw the wrong type: " + e.what());
        return;
    }
    expect(false, what + " did not throw");
}

void write_i64(const std::string& path, const std::vector<std::int64_t>& v) {
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    out.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * 8));
}

gpart::WeightedGraph grid(int rows, int cols) {
    std::int64_t n = 0, m = 0;
    std::int32_t *off = nullptr, *nbr = nullptr;
    gpart::detail::check(gp_generate_grid(rows, cols, &n, &m, &off, &nbr));
    gpart::WeightedGraph g;
    g.offsets.assign(off, off + n + 1);
    g.neighbors.assign(nbr, nbr + m);
    g.edge_weights.assign(static_cast<std::size_t>(m), 1u);
    g.vertex_weights.assign(static_cast<std::size_t>(n), 1u);
    gp_free_host(off);
    gp_free_host(nbr);
    return g;
}

gpart::WeightedGraph knn(int points, int k, std::uint64_t seed) {
    std::int64_t n = 0, m = 0;
    std::int32_t *off = nullptr, *nbr = nullptr;
    std::uint32_t *ew = nullptr, *vw = nullptr;
    gpart::detail::check(gp_generate_knn3d(points, k, seed, &n, &m, &off, &nbr, &ew, &vw));
    gpart::WeightedGraph g;
    g.offsets.assign(off, off + n + 1);
    g.neighbors.assign(nbr, nbr + m);
    g.edge_weights.assign(ew, ew + m);
    g.vertex_weights.assign(vw, vw + n);
    gp_free_host(off);
    gp_free_host(nbr);
    gp_free_host(ew);
    gp_free_host(vw);
    return g;
}

void cpu_checks(const std::string& dir) {
    // partition_capacity (include/partition_capacity.hpp:14-25)
    expect(gpart::partition_capacity(65536, 8, 1.03) == 8437, "capacity grid256 k=8 is 8437");
    expect(gpart::partition_capacity(4194304, 64, 1.03) == 67502, "capacity rmat22 k=64 is 67502");
    expect_throw<std::invalid_argument>([] { gpart::partition_capacity(100, 1, 1.03); }, "capacity parts=1",
                                        "invalid partition capacity options");
    expect_throw<std::invalid_argument>([] { gpart::partition_capacity(100, 129, 1.03); }, "capacity parts=129");
    expect_throw<std::invalid_argument>([] { gpart::partition_capacity(100, 4, 0.5); }, "capacity imbalance<1");
    expect_throw<std::invalid_argument>([] { gpart::partition_capacity(10, 4, 1.0); }, "capacity cannot hold",
                                        "cannot hold total weight");
    expect_throw<std::overflow_error>([] { gpart::partition_capacity(~0ULL, 2, 3.0); }, "capacity overflow",
                                      "exceeds uint64");

    // int64 CSR reader (src/graph.cpp:47-116): a 4-cycle
    const std::string ip = dir + "/c4.indptr", ix = dir + "/c4.indices";
    write_i64(ip, {0, 2, 4, 6, 8});
    write_i64(ix, {1, 3, 0, 2, 1, 3, 0, 2});
    gpart::WeightedGraph c4 = gpart::load_weighted_graph(ip, ix);
    expect(c4.vertices() == 4 && c4.edges() == 8 && c4.edge_weights == std::vector<std::uint32_t>(8, 1u) &&
               c4.vertex_weights == std::vector<std::uint32_t>(4, 1u),
           "load 4-cycle with unit weights");
    gpart::validate_weighted_csr(c4, false);
    expect(true, "4-cycle passes strict validation");
    expect(gpart::host_cut(c4, {0, 0, 1, 1}) == 2 && gpart::host_cut(c4, {0, 1, 0, 1}) == 4,
           "host_cut of two 4-cycle colourings");
    write_i64(ix, {1, 3, 0, 2, 1, 3, 0, 4});
    expect_throw<std::runtime_error>([&] { gpart::load_weighted_graph(ip, ix); }, "reader rejects neighbour >= n");
    write_i64(ix, {1, 3, 0, 2, 1, 3, 0});
    expect_throw<std::runtime_error>([&] { gpart::load_weighted_graph(ip, ix); }, "reader rejects back != m");
    write_i64(ip, {0, 2, 1, 6, 7});
    expect_throw<std::runtime_error>([&] { gpart::load_weighted_graph(ip, ix); }, "reader rejects non-monotone offsets");

    // validation (src/graph.cpp:118-186)
    gpart::WeightedGraph asym = c4;
    asym.edge_weights[0] = 5;
    expect_throw<std::runtime_error>([&] { gpart::validate_weighted_csr(asym, false); }, "strict: reverse weight differs",
                                     "reverse edge weights differ");
    gpart::validate_weighted_shape(asym, false);
    expect(true, "shape check accepts asymmetric weights");
    gpart::WeightedGraph loop = c4;
    loop.neighbors[0] = 0;
    expect_throw<std::runtime_error>([&] { gpart::validate_weighted_shape(loop, false); }, "self loop refused",
                                     "self loop");
    gpart::validate_weighted_shape(loop, true);
    expect(true, "self loop allowed when asked");
    gpart::WeightedGraph oneway = c4;
    oneway.neighbors[1] = 2;  // 0 -> 2 has no 2 -> 0
    expect_throw<std::runtime_error>([&] { gpart::validate_weighted_csr(oneway, false); }, "strict: missing reverse",
                                     "not bidirectional");

    // Jet-compatible hierarchy file (include/hierarchy.hpp:40-123)
    gpart::Hierarchy h;
    h.levels.push_back(c4);
    gpart::WeightedGraph pair;
    pair.offsets = {0, 1, 2};
    pair.neighbors = {1, 0};
    pair.edge_weights = {2, 2};
    pair.vertex_weights = {2, 2};
    h.levels.push_back(pair);
    h.fine_to_coarse.push_back({0, 0, 1, 1});
    const std::string hp = dir + "/h.bin";
    gpart::write_hierarchy(h, hp);
    std::ifstream f(hp, std::ios::binary | std::ios::ate);
    // 4 + level 0 (8 + 20 + 32 + 32 + 16) + level 1 (8 + 12 + 8 + 8 + 8 + map 16)
    expect(static_cast<long>(f.tellg()) == 4 + 108 + 60, "hierarchy file has the reference's byte layout size");
    gpart::Hierarchy back = gpart::read_hierarchy(hp);
    expect(back.layer_count() == 2 && back.levels[1].neighbors == pair.neighbors &&
               back.levels[1].edge_weights == pair.edge_weights && back.fine_to_coarse[0] == h.fine_to_coarse[0] &&
               back.levels[0].offsets == c4.offsets,
           "hierarchy file round trip");
    h.fine_to_coarse[0][3] = 2;
    expect_throw<std::runtime_error>([&] { gpart::write_hierarchy(h, hp); }, "hierarchy map id out of range",
                                     "invalid coarse id");
    std::ifstream still(hp + ".tmp");
    expect(!still, "failed write leaves no temporary file");
}

std::vector<std::int32_t> c_partition(const gpart::WeightedGraph& g, int k, double eps, std::uint32_t seed,
                                      std::uint64_t* cut) {
    gp_context* ctx = nullptr;
    gpart::detail::check(gp_context_create(0, &ctx));
    gp_graph* dg = nullptr;
    const bool unit = std::all_of(g.edge_weights.begin(), g.edge_weights.end(), [](std::uint32_t w) { return w == 1; });
    gpart::detail::check(gp_graph_create(ctx, g.vertices(), g.edges(), g.offsets.data(), g.neighbors.data(),
                                         unit ? nullptr : g.edge_weights.data(), g.vertex_weights.data(), &dg));
    std::vector<std::int32_t> labels(static_cast<std::size_t>(g.vertices()));
    gp_partition_stats st;
    gpart::detail::check(gp_partition(ctx, dg, k, eps, seed, labels.data(), &st));
    *cut = st.final_cut;
    gp_graph_destroy(dg);
    gp_context_destroy(ctx);
    return labels;
}

gpart::CoarsenOptions production(int k, std::uint32_t seed) {
    gpart::CoarsenOptions o;
    o.parts = k;
    o.seed = seed;
    o.cluster_size_factor = k <= 8 ? std::max(64, (512 + k - 1) / k) : std::max(2, (32 + k - 1) / k);
    o.pack_isolated = true;
    o.progressive_cluster_capacity = k >= 16;
    return o;
}

void gpu_case(const std::string& name, const gpart::WeightedGraph& g, int k) {
    std::cout << "case " << name << std::endl;
    std::uint64_t c_cut = 0;
    const std::vector<std::int32_t> want = c_partition(g, k, 1.03, 0, &c_cut);

    // coarsen(DeviceWeightedGraph&&) -> initial_partition_metis -> uncoarsen, strict
    gpart::CoarsenOptions co = production(k, 0);
    co.strict_verify = true;
    auto device = gpart::make_device_weighted(g);
    auto hierarchy = gpart::coarsen(std::move(device), co);
    expect(!device.handle, name + ": coarsen took the input graph over");
    const gpart::Hierarchy host = gpart::copy_device_hierarchy(hierarchy);
    auto coarse = gpart::initial_partition_metis(hierarchy.levels.back(), k, 1.03, 0, false);
    std::vector<std::int32_t> coarse_host(coarse.size());
    thrust::copy(coarse.begin(), coarse.end(), coarse_host.begin());
    gpart::RefineOptions ro;
    ro.parts = k;
    ro.imbalance_ratio = 1.03;
    ro.allow_initial_imbalance = true;
    ro.strict_verify = true;
    auto result = gpart::uncoarsen(hierarchy, std::move(coarse), ro);
    expect(hierarchy.levels.empty() && !hierarchy.handle, name + ": uncoarsen consumed the hierarchy");
    std::vector<std::int32_t> got(result.device_partition.size());
    thrust::copy(result.device_partition.begin(), result.device_partition.end(), got.begin());
    expect(got == want && result.final_cut == c_cut, name + ": C++ pipeline (strict) equals gp_partition");
    expect(gpart::host_cut(g, got) == result.final_cut, name + ": host_cut equals the reported cut");
    std::uint64_t checks = 0;
    for (const auto& l : result.levels) checks += l.refinement.strict_checks;
    expect(checks >= result.levels.size(), name + ": strict recounts ran on every level");

    // host hierarchy through refine_hierarchy_device / uncoarsen(const Hierarchy&)
    gpart::RefineOptions plain = ro;
    plain.strict_verify = false;
    const std::vector<std::int32_t> host_labels = gpart::uncoarsen(host, coarse_host, plain);
    expect(host_labels == want, name + ": host-hierarchy uncoarsen equals gp_partition");

    // aggregate/contract level by level reproduce coarsen's levels
    gpart::SclpWorkspace ws;
    gpart::DeviceWeightedGraph cur = gpart::make_device_weighted(g);
    // The level loop classifies the input's rows before its first seed-matching
    // decision (src/coarsen.cu:1806-1848); a caller of aggregate() does the same.
    std::uint64_t max_degree = 0;
    bool sorted_unique = true, positive = true;
    for (std::int64_t v = 0; v < g.vertices(); ++v) {
        const auto b = g.offsets[static_cast<std::size_t>(v)], e = g.offsets[static_cast<std::size_t>(v) + 1];
        max_degree = std::max<std::uint64_t>(max_degree, static_cast<std::uint64_t>(e - b));
        for (auto x = b; x + 1 < e; ++x)
            sorted_unique &= g.neighbors[static_cast<std::size_t>(x)] < g.neighbors[static_cast<std::size_t>(x) + 1];
    }
    for (auto w : g.edge_weights) positive &= w > 0;
    cur.maximum_degree = max_degree;
    cur.has_heavy_rows = max_degree >= 4096;
    bool same = true;
    std::size_t level = 0;
    const std::int64_t cutoff = std::max<std::int64_t>(32, static_cast<std::int64_t>(co.cluster_size_factor) * k);
    for (; level + 1 < host.layer_count(); ++level) {
        int factor = co.cluster_size_factor;
        if (co.progressive_cluster_capacity) {
            const std::int64_t ff = std::max<std::int64_t>(co.progressive_fine_factor, (512 + k - 1) / k);
            const std::int64_t den = static_cast<std::int64_t>(co.progressive_capacity_growth) * k;
            factor = static_cast<int>(std::max<std::int64_t>(factor, std::min(ff, (cur.vertices() + den - 1) / den)));
        }
        (void)cutoff;
        const std::int64_t n = cur.vertices(), m = cur.edges();
        // Contraction emits sorted unique rows of positive sums.
        const bool rows_ok = level == 0 ? sorted_unique && positive : true;
        const bool hem = co.seed_matching && n >= 65536 && !cur.has_heavy_rows &&
                         (m / n < 8 || cur.maximum_degree <= 256) && m / n < co.heavy_edge_degree_limit && rows_ok;
        gpart::SclpStats stats;
        auto agg = gpart::aggregate(cur, ws, k, co.seed + static_cast<std::uint32_t>(level), static_cast<int>(level),
                                    stats, false, factor, co.pack_isolated, hem, co.capacity_safe_fast_path);
        same = same && agg.map_to_host(n) == host.fine_to_coarse[level];
        gpart::ContractionTimings t;
        cur = gpart::contract(std::move(cur), std::move(agg), ws, t);
        const gpart::WeightedGraph c = gpart::copy_device_weighted(cur);
        const gpart::WeightedGraph& w = host.levels[level + 1];
        same = same && c.offsets == w.offsets && c.neighbors == w.neighbors && c.edge_weights == w.edge_weights &&
               c.vertex_weights == w.vertex_weights;
    }
    expect(same && level + 1 == host.layer_count(), name + ": aggregate/contract chain equals coarsen's " +
                                                         std::to_string(host.layer_count()) + " levels");

    // refine_partition on the finest graph from a projected coarse partition
    std::vector<std::int32_t> start(static_cast<std::size_t>(g.vertices()));
    std::vector<std::int32_t> proj = coarse_host;
    for (std::size_t l = host.fine_to_coarse.size(); l-- > 0;) {
        std::vector<std::int32_t> fine(host.fine_to_coarse[l].size());
        for (std::size_t v = 0; v < fine.size(); ++v) fine[v] = proj[static_cast<std::size_t>(host.fine_to_coarse[l][v])];
        proj.swap(fine);
    }
    start = proj;
    gpart::RefineOptions rp = plain;
    gpart::RefineStats rs;
    std::vector<std::int32_t> refined = start;
    gpart::refine_partition(g, refined, rp, &rs);
    std::uint64_t total = 0, worst = 0;
    std::vector<std::uint64_t> loads(static_cast<std::size_t>(k), 0);
    for (std::size_t v = 0; v < refined.size(); ++v) loads[static_cast<std::size_t>(refined[v])] += g.vertex_weights[v];
    for (auto x : loads) {
        total += x;
        worst = std::max(worst, x);
    }
    expect(gpart::host_cut(g, refined) == rs.final_cut && rs.final_cut <= gpart::host_cut(g, start) &&
               worst <= gpart::partition_capacity(total, k, 1.03),
           name + ": refine_partition result is feasible and no worse");
}

void gpu_checks() {
    gpu_case("grid64_k4", grid(64, 64), 4);
    gpu_case("knn4000_k16_weighted", knn(4000, 8, 3), 16);
    gpu_case("grid256_k8", grid(256, 256), 8);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "usage: gpart_api_test cpu|gpu <scratch dir>\n";
        return 2;
    }
    try {
        if (std::string(argv[1]) == "cpu")
            cpu_checks(argv[2]);
        else
            gpu_checks();
    } catch (const std::exception& e) {
        std::cout << "FAIL exception: " << e.what() << std::endl;
        return 1;
    }
    std::cout << (failures ? "FAILED " : "PASSED ") << failures << std::endl;
    return failures ? 1 : 0;
}
