// Host initial partition of the coarsest graph: the boundary of
// initial_partition_metis (src/initial_partition_metis.cu:86-183).
//
// The boundary checks and normalisation are the reference's: positive vertex
// weights, no vertex above the integer capacity, exact common-factor (GCD)
// normalisation of vertex and edge weights (the weighted objective and the
// capacities are unchanged by it), self loops and zero-weight edges left out
// of the partitioner's input, and in strict mode a feasibility check of the
// normalised weights plus the single-move capacity repair of
// initial_partition_metis.cu:35-83.  The partitioner behind the boundary is
// METIS 5 (METIS_PartGraphKway, seed & 0x7fffffff, 0-based numbering, the
// imbalance as ubvec) when the build finds metis.h and libmetis
// (GP_WITH_METIS), and otherwise the in-repo recursive bisection of
// gp_host.cpp, which is not METIS.
#include "gp_host.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef GP_WITH_METIS
#include <metis.h>
#endif

namespace gp {

namespace {

std::uint64_t checked_sum(const std::uint32_t* w, std::size_t count, const char* what) {
    std::uint64_t total = 0;
    for (std::size_t i = 0; i < count; ++i) {
        const std::uint64_t x = w ? w[i] : 1u;
        if (x > std::numeric_limits<std::uint64_t>::max() - total)
            throw std::overflow_error(std::string(what) + " weight sum exceeds uint64");
        total += x;
    }
    return total;
}

}  // namespace

const char* initial_partition_backend() {
#ifdef GP_WITH_METIS
    return "metis";
#else
    return "bisection";
#endif
}

std::size_t repair_capacity(std::int64_t n, const std::int32_t* off, const std::int32_t* nbr, const std::uint32_t* ew,
                            const std::uint32_t* vw, int parts, std::uint64_t capacity, std::int32_t* labels,
                            std::vector<std::uint64_t>& loads) {
    // initial_partition_metis.cu:35-83.  While a part is overloaded, move the
    // single vertex of an overloaded part to a part with room for all of it
    // whose (A(v,p) - A(v,s)) / min(w, load_s - cap) is largest (long double;
    // the first such pair in (vertex, part) order on ties).
    if (*std::max_element(loads.begin(), loads.end()) <= capacity) return 0;
    const std::size_t k = static_cast<std::size_t>(parts);
    std::vector<std::uint64_t> affinity(static_cast<std::size_t>(n) * k, 0);
    for (std::int64_t v = 0; v < n; ++v)
        for (std::int64_t e = off[v]; e < off[v + 1]; ++e) {
            const std::int32_t u = nbr[e];
            if (u != v) affinity[static_cast<std::size_t>(v) * k + static_cast<std::size_t>(labels[u])] += ew ? ew[e] : 1u;
        }
    std::size_t moved = 0;
    while (*std::max_element(loads.begin(), loads.end()) > capacity) {
        std::int64_t chosen = -1;
        int target = -1;
        long double best = -std::numeric_limits<long double>::infinity();
        for (std::int64_t v = 0; v < n; ++v) {
            const int source = labels[v];
            if (loads[static_cast<std::size_t>(source)] <= capacity) continue;
            const std::uint64_t weight = vw ? vw[v] : 1u;
            const std::uint64_t removable = std::min(weight, loads[static_cast<std::size_t>(source)] - capacity);
            for (std::size_t p = 0; p < k; ++p) {
                if (static_cast<int>(p) == source || loads[p] > capacity || weight > capacity - loads[p]) continue;
                const long double gain =
                    static_cast<long double>(affinity[static_cast<std::size_t>(v) * k + p]) -
                    static_cast<long double>(affinity[static_cast<std::size_t>(v) * k + static_cast<std::size_t>(source)]);
                const long double score = gain / static_cast<long double>(removable);
                if (score > best) {
                    best = score;
                    chosen = v;
                    target = static_cast<int>(p);
                }
            }
        }
        if (target < 0) throw std::runtime_error("initial partition cannot meet strict capacity by feasible single moves");
        const int source = labels[chosen];
        const std::uint64_t weight = vw ? vw[chosen] : 1u;
        loads[static_cast<std::size_t>(source)] -= weight;
        loads[static_cast<std::size_t>(target)] += weight;
        for (std::int64_t e = off[chosen]; e < off[chosen + 1]; ++e) {
            const std::int32_t u = nbr[e];
            if (u == chosen) continue;
            const std::uint64_t w = ew ? ew[e] : 1u;
            affinity[static_cast<std::size_t>(u) * k + static_cast<std::size_t>(source)] -= w;
            affinity[static_cast<std::size_t>(u) * k + static_cast<std::size_t>(target)] += w;
        }
        labels[chosen] = target;
        ++moved;
    }
    return moved;
}

void initial_partition(std::int64_t n, const std::int32_t* off, const std::int32_t* nbr, const std::uint32_t* ew,
                       const std::uint32_t* vw, int parts, double imbalance, std::uint32_t seed, bool require_feasible,
                       std::int32_t* labels, InitialReport* report) {
    const auto start = std::chrono::steady_clock::now();
    if (n <= 0 || parts < 2 || parts > 128 || parts > n || !std::isfinite(imbalance) || imbalance < 1.0)
        throw std::invalid_argument("invalid initial partition options");
    const std::int64_t m = off[n];
    const std::uint64_t total = checked_sum(vw, static_cast<std::size_t>(n), "vertex");
    const std::uint64_t edge_total = checked_sum(ew, static_cast<std::size_t>(m), "edge");
    const std::uint64_t capacity = capacity_of(total, parts, imbalance);
    std::uint64_t vertex_gcd = 0, edge_gcd = 0;
    for (std::int64_t v = 0; v < n; ++v) {
        const std::uint64_t w = vw ? vw[v] : 1u;
        if (!w) throw std::invalid_argument("initial partition requires positive vertex weights");
        if (w > capacity) throw std::invalid_argument("a vertex exceeds partition capacity");
        vertex_gcd = std::gcd(vertex_gcd, w);
    }
    for (std::int64_t e = 0; e < m; ++e) edge_gcd = std::gcd(edge_gcd, static_cast<std::uint64_t>(ew ? ew[e] : 1u));
    if (!edge_gcd) edge_gcd = 1;
    const std::uint64_t scaled_total = total / vertex_gcd;
#ifdef GP_WITH_METIS
    const std::uint64_t limit = static_cast<std::uint64_t>(std::numeric_limits<idx_t>::max());
#else
    const std::uint64_t limit = static_cast<std::uint64_t>(std::numeric_limits<std::int64_t>::max());
#endif
    if (scaled_total > limit || edge_total / edge_gcd > limit)
        throw std::overflow_error("weighted coarse graph exceeds the partitioner's signed weight range");
    const std::uint64_t minimum_load = scaled_total / static_cast<unsigned>(parts) +
                                       (scaled_total % static_cast<unsigned>(parts) != 0);
    if (require_feasible && capacity / vertex_gcd < minimum_load)
        throw std::runtime_error("indivisible weights cannot satisfy partition capacity");

    // Normalised input: weights divided by their common factors, self loops
    // and zero-weight edges dropped.
    std::vector<std::int32_t> noff(static_cast<std::size_t>(n) + 1, 0), nnbr;
    std::vector<std::uint32_t> new_, nvw(static_cast<std::size_t>(n));
    nnbr.reserve(static_cast<std::size_t>(m));
    new_.reserve(static_cast<std::size_t>(m));
    for (std::int64_t v = 0; v < n; ++v) {
        nvw[static_cast<std::size_t>(v)] = static_cast<std::uint32_t>((vw ? vw[v] : 1u) / vertex_gcd);
        for (std::int64_t e = off[v]; e < off[v + 1]; ++e) {
            const std::uint32_t w = ew ? ew[e] : 1u;
            if (nbr[e] == v || !w) continue;
            nnbr.push_back(nbr[e]);
            new_.push_back(static_cast<std::uint32_t>(w / edge_gcd));
        }
        noff[static_cast<std::size_t>(v) + 1] = static_cast<std::int32_t>(nnbr.size());
    }
    const char* backend = "bisection";
#ifdef GP_WITH_METIS
    {
        backend = "metis";
        idx_t nn = static_cast<idx_t>(n), ncon = 1, k = parts, objective = 0;
        std::vector<idx_t> xadj(noff.begin(), noff.end()), adjncy(nnbr.begin(), nnbr.end()),
            adjwgt(new_.begin(), new_.end()), vwgt(nvw.begin(), nvw.end()), part(static_cast<std::size_t>(n));
        idx_t options[METIS_NOPTIONS];
        if (METIS_SetDefaultOptions(options) != METIS_OK) throw std::runtime_error("METIS_SetDefaultOptions failed");
        options[METIS_OPTION_SEED] = static_cast<idx_t>(seed & 0x7fffffffU);
        options[METIS_OPTION_NUMBERING] = 0;
        real_t upper = static_cast<real_t>(imbalance);
        if (upper <= real_t{1}) upper = std::nextafter(real_t{1}, std::numeric_limits<real_t>::infinity());
        const int status = METIS_PartGraphKway(&nn, &ncon, xadj.data(), adjncy.data(), vwgt.data(), nullptr,
                                               adjwgt.data(), &k, nullptr, &upper, options, &objective, part.data());
        if (status != METIS_OK) throw std::runtime_error("METIS_PartGraphKway failed: " + std::to_string(status));
        for (std::int64_t v = 0; v < n; ++v) labels[v] = static_cast<std::int32_t>(part[static_cast<std::size_t>(v)]);
    }
#else
    initial_partition_host(n, noff.data(), nnbr.data(), new_.data(), nvw.data(), parts, imbalance, seed, labels);
#endif
    std::vector<std::uint64_t> loads(static_cast<std::size_t>(parts), 0);
    for (std::int64_t v = 0; v < n; ++v) {
        if (labels[v] < 0 || labels[v] >= parts) throw std::runtime_error("initial partitioner returned an invalid label");
        loads[static_cast<std::size_t>(labels[v])] += vw ? vw[v] : 1u;
    }
    const std::uint64_t initial_max = *std::max_element(loads.begin(), loads.end());
    const std::size_t repaired =
        require_feasible ? repair_capacity(n, off, nbr, ew, vw, parts, capacity, labels, loads) : 0;
    if (report) {
        report->backend = backend;
        report->vertex_gcd = vertex_gcd;
        report->edge_gcd = edge_gcd;
        report->max_load = initial_max;
        report->repaired = repaired;
        std::uint64_t d = 0;
        for (std::int64_t v = 0; v < n; ++v)
            for (std::int64_t e = off[v]; e < off[v + 1]; ++e)
                if (labels[v] != labels[nbr[e]]) d += ew ? ew[e] : 1u;
        report->edgecut = d / 2;
        report->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    }
}

}  // namespace gp
