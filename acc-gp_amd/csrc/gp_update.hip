// Device update of a resident CSR by a batch of undirected edge deletions
// and insertions (repartition requests of a server's resident graphs; no
// counterpart in the reference, whose graphs are read from disk once).
//
// Semantics: deletions apply first, then insertions; deleting an absent
// edge changes nothing, and an insertion of an edge present after the
// deletions, or given twice, is dropped (the first weight is kept).  Each
// row keeps its surviving entries in their old order, followed by its new
// neighbours in ascending order.  Inserted edges carry the given weights (1
// when none are given); a graph with implicit unit weights gets an explicit
// weight array when an inserted weight is not 1.
//
// Layout on the device: the directed delta records (both directions) sorted
// by (source, kind, neighbour), with one record range per changed row.  Rows
// without a record move by an entry-parallel copy (each entry finds its row
// by a binary search over the old offsets); a changed row is scanned and
// rebuilt by one wave, each entry looked up by binary search among the row's
// sorted deletions and insertions (hub rows carry many records).  Bound:
// HBM, one read and one write of the CSR.
#include "gp_update.hpp"

#include <algorithm>
#include <stdexcept>
#include <tuple>
#include <vector>

namespace gp {

namespace {

struct DeltaRows {
    const std::int32_t* row;      // changed rows, ascending
    const std::int32_t* first;    // records of row i: [first[i], first[i + 1])
    const std::int32_t* mid;      // deletions [first[i], mid[i]), insertions [mid[i], first[i + 1])
    const std::int32_t* dst;      // record neighbour, ascending within each kind
    const std::uint32_t* weight;  // insert weight
    std::uint8_t* present;        // per insertion: its edge is in the row after the deletions
    std::int64_t count;           // changed rows
};

// Position of u in the ascending range [lo, hi) of a, or -1.
__device__ inline std::int32_t find_sorted(const std::int32_t* a, std::int32_t lo, std::int32_t hi, std::int32_t u) {
    const std::int32_t end = hi;
    while (lo < hi) {
        const std::int32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    return lo < end && a[lo] == u ? lo : -1;
}

__global__ void k_delta_degree(std::int64_t n, const std::int32_t* off, std::int32_t* deg) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x)
        deg[v] = off[v + 1] - off[v];
}

// One wave per changed row: the entries that survive the row's deletions
// (a binary search over the sorted deletions each), and the insertions whose
// edge is among them (flagged present).  deg[v] receives the survivors.
__global__ void k_delta_scan_rows(const std::int32_t* off, const std::int32_t* nbr, DeltaRows d, std::int32_t* deg) {
    const int lane = lane_id();
    const std::int64_t waves = (std::int64_t)gridDim.x * blockDim.x / kWave;
    for (std::int64_t i = (blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x) / kWave; i < d.count; i += waves) {
        const std::int32_t v = d.row[i];
        const std::int64_t b = off[v], e = off[v + 1];
        const std::int32_t f0 = d.first[i], fm = d.mid[i], f1 = d.first[i + 1];
        std::int64_t kept = 0;
        for (std::int64_t x = b + lane; x < e; x += kWave) {
            const std::int32_t u = nbr[x];
            if (fm > f0 && find_sorted(d.dst, f0, fm, u) >= 0) continue;
            ++kept;
            const std::int32_t r = f1 > fm ? find_sorted(d.dst, fm, f1, u) : -1;
            if (r >= 0) d.present[r] = 1;
        }
        kept = wave_sum(kept);
        if (lane == 0) deg[v] = (std::int32_t)kept;
    }
}

// One thread per changed row: survivors plus the insertions not present.
__global__ void k_delta_row_degree(DeltaRows d, std::int32_t* deg, std::uint8_t* changed) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < d.count;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t v = d.row[i];
        std::int32_t added = 0;
        for (std::int32_t r = d.mid[i]; r < d.first[i + 1]; ++r) added += !d.present[r];
        deg[v] += added;
        changed[v] = 1;
    }
}

// Entries of unchanged rows, one thread per old entry.
__global__ void k_delta_copy(std::int64_t n, std::int64_t m, const std::int32_t* off, const std::int32_t* nbr,
                             const std::uint32_t* ew, const std::uint8_t* changed, const std::int32_t* noff,
                             std::int32_t* nnbr, std::uint32_t* new_ew) {
    for (std::int64_t x = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; x < m;
         x += (std::int64_t)gridDim.x * blockDim.x) {
        std::int64_t lo = 0, hi = n - 1;  // last row whose first entry is <= x
        while (lo < hi) {
            const std::int64_t mid = (lo + hi + 1) >> 1;
            if (off[mid] <= x) lo = mid;
            else hi = mid - 1;
        }
        if (changed[lo]) continue;
        const std::int64_t y = noff[lo] + (x - off[lo]);
        nnbr[y] = nbr[x];
        if (new_ew) new_ew[y] = ew ? ew[x] : 1u;
    }
}

// One wave per changed row: the survivors in their old order (ballot
// compaction), then the insertions not present, ascending.
__global__ void k_delta_rows(const std::int32_t* off, const std::int32_t* nbr, const std::uint32_t* ew, DeltaRows d,
                             const std::int32_t* noff, std::int32_t* nnbr, std::uint32_t* new_ew) {
    const int lane = lane_id();
    const unsigned long long below = (1ULL << lane) - 1ULL;
    const std::int64_t waves = (std::int64_t)gridDim.x * blockDim.x / kWave;
    for (std::int64_t i = (blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x) / kWave; i < d.count; i += waves) {
        const std::int32_t v = d.row[i];
        const std::int64_t b = off[v], e = off[v + 1];
        const std::int32_t f0 = d.first[i], fm = d.mid[i], f1 = d.first[i + 1];
        std::int64_t y = noff[v];
        for (std::int64_t x0 = b; x0 < e; x0 += kWave) {
            const std::int64_t x = x0 + lane;
            std::int32_t u = 0;
            bool keep = false;
            if (x < e) {
                u = nbr[x];
                keep = fm == f0 || find_sorted(d.dst, f0, fm, u) < 0;
            }
            const unsigned long long mask = __ballot(keep);
            if (keep) {
                const std::int64_t at = y + __popcll(mask & below);
                nnbr[at] = u;
                if (new_ew) new_ew[at] = ew ? ew[x] : 1u;
            }
            y += __popcll(mask);
        }
        for (std::int32_t r0 = fm; r0 < f1; r0 += kWave) {
            const std::int32_t r = r0 + lane;
            const bool add = r < f1 && !d.present[r];
            const unsigned long long mask = __ballot(add);
            if (add) {
                const std::int64_t at = y + __popcll(mask & below);
                nnbr[at] = d.dst[r];
                if (new_ew) new_ew[at] = d.weight[r];
            }
            y += __popcll(mask);
        }
    }
}

inline int blocks_for(std::int64_t n) {
    return static_cast<int>(std::max<std::int64_t>(1, std::min<std::int64_t>((n + 255) / 256, 4096)));
}

}  // namespace

std::int64_t apply_delta(Context& c, DGraph& g, std::int64_t count, const std::int32_t* u, const std::int32_t* v,
                         const std::uint32_t* w, const std::uint8_t* insert) {
    if (count < 0 || (count > 0 && (!u || !v || !insert))) throw std::invalid_argument("invalid edge delta");
    // Directed records, validated, sorted by (source, kind, neighbour), with
    // repeated records dropped (the first insertion's weight is kept).
    struct Rec {
        std::int32_t src, dst;
        std::uint8_t ins;
        std::uint32_t w;
        std::int64_t order;
    };
    std::vector<Rec> recs;
    recs.reserve(static_cast<std::size_t>(2 * count));
    bool unit = true;
    for (std::int64_t i = 0; i < count; ++i) {
        if (u[i] < 0 || u[i] >= g.n || v[i] < 0 || v[i] >= g.n || u[i] == v[i])
            throw std::invalid_argument("edge delta names an invalid vertex or a self loop");
        const std::uint32_t wi = (insert[i] && w) ? w[i] : 1u;
        if (insert[i] && !wi) throw std::invalid_argument("inserted edge weights must be positive");
        unit &= !insert[i] || wi == 1u;
        recs.push_back(Rec{u[i], v[i], (std::uint8_t)(insert[i] ? 1 : 0), wi, 2 * i});
        recs.push_back(Rec{v[i], u[i], (std::uint8_t)(insert[i] ? 1 : 0), wi, 2 * i + 1});
    }
    std::sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) {
        return std::tie(a.src, a.ins, a.dst, a.order) < std::tie(b.src, b.ins, b.dst, b.order);
    });
    recs.erase(std::unique(recs.begin(), recs.end(),
                           [](const Rec& a, const Rec& b) { return a.src == b.src && a.ins == b.ins && a.dst == b.dst; }),
               recs.end());
    std::vector<std::int32_t> rows, first, dst;
    std::vector<std::uint8_t> kind;
    std::vector<std::uint32_t> weight;
    for (std::size_t r = 0; r < recs.size(); ++r) {
        if (rows.empty() || rows.back() != recs[r].src) {
            rows.push_back(recs[r].src);
            first.push_back(static_cast<std::int32_t>(r));
        }
        dst.push_back(recs[r].dst);
        kind.push_back(recs[r].ins);
        weight.push_back(recs[r].w);
    }
    first.push_back(static_cast<std::int32_t>(recs.size()));
    const std::int64_t nrows = static_cast<std::int64_t>(rows.size());
    if (!nrows) return g.m;
    // Within a row the deletions come first: mid[i] is its first insertion.
    std::vector<std::int32_t> mid(static_cast<std::size_t>(nrows));
    for (std::int64_t i = 0; i < nrows; ++i) {
        std::int32_t r = first[static_cast<std::size_t>(i)];
        while (r < first[static_cast<std::size_t>(i) + 1] && !kind[static_cast<std::size_t>(r)]) ++r;
        mid[static_cast<std::size_t>(i)] = r;
    }

    Arena* home = g.off.arena();  // the new arrays belong where the graph's do
    const hipStream_t s = c.stream();
    DBuf<std::int32_t> d_rows(home, rows.size()), d_first(home, first.size()), d_mid(home, mid.size()),
        d_dst(home, dst.size());
    DBuf<std::uint8_t> d_present(home, dst.size());
    DBuf<std::uint32_t> d_weight(home, weight.size());
    c.upload(d_rows.get(), rows.data(), rows.size() * 4);
    c.upload(d_first.get(), first.data(), first.size() * 4);
    c.upload(d_mid.get(), mid.data(), mid.size() * 4);
    c.upload(d_dst.get(), dst.data(), dst.size() * 4);
    c.upload(d_weight.get(), weight.data(), weight.size() * 4);
    fill<std::uint8_t>(c, d_present.get(), static_cast<std::int64_t>(dst.size()), 0);
    const DeltaRows d{d_rows.get(), d_first.get(), d_mid.get(), d_dst.get(), d_weight.get(), d_present.get(), nrows};

    DBuf<std::int32_t> deg(home, static_cast<std::size_t>(g.n) + 1);
    DBuf<std::uint8_t> changed(home, static_cast<std::size_t>(g.n));
    fill<std::uint8_t>(c, changed.get(), g.n, 0);
    hipLaunchKernelGGL(k_delta_degree, dim3(blocks_for(g.n)), dim3(256), 0, s, g.n, g.off.get(), deg.get());
    GP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_delta_scan_rows, dim3(blocks_for(nrows * kWave)), dim3(256), 0, s, g.off.get(), g.nbr.get(),
                       d, deg.get());
    GP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_delta_row_degree, dim3(blocks_for(nrows)), dim3(256), 0, s, d, deg.get(), changed.get());
    GP_LAUNCH_CHECK();
    fill<std::int32_t>(c, deg.get() + g.n, 1, 0);
    DBuf<unsigned long long> total(home, 1);
    DBuf<std::int32_t> noff(home, static_cast<std::size_t>(g.n) + 1);
    exclusive_sum<unsigned long long>(c, deg.get(), noff.get(), g.n + 1, total.get());
    const unsigned long long m2 = c.read_scalar(total.get());
    if (m2 > 0x7fffffffULL) throw std::invalid_argument("updated graph exceeds 32-bit CSR limits");
    DBuf<std::int32_t> nnbr(home, static_cast<std::size_t>(std::max<unsigned long long>(m2, 1)));
    const bool weighted = g.ew || !unit;
    DBuf<std::uint32_t> new_ew;
    if (weighted) new_ew = DBuf<std::uint32_t>(home, static_cast<std::size_t>(std::max<unsigned long long>(m2, 1)));
    if (g.m > 0) {
        hipLaunchKernelGGL(k_delta_copy, dim3(blocks_for(g.m)), dim3(256), 0, s, g.n, g.m, g.off.get(), g.nbr.get(),
                           g.ew_ptr(), changed.get(), noff.get(), nnbr.get(), weighted ? new_ew.get() : nullptr);
        GP_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_delta_rows, dim3(blocks_for(nrows * kWave)), dim3(256), 0, s, g.off.get(), g.nbr.get(), g.ew_ptr(),
                       d, noff.get(), nnbr.get(), weighted ? new_ew.get() : nullptr);
    GP_LAUNCH_CHECK();
    c.sync();
    g.off = std::move(noff);
    g.nbr = std::move(nnbr);
    g.ew = std::move(new_ew);
    g.m = static_cast<std::int64_t>(m2);
    // Bounds carried by coarsening are unknown for the new rows.
    g.max_degree = std::numeric_limits<std::uint64_t>::max();
    g.has_heavy = true;
    g.incident_bound = std::numeric_limits<std::uint64_t>::max();
    g.row_len_bound = std::numeric_limits<std::uint64_t>::max();
    return g.m;
}

}  // namespace gp
