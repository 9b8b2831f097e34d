// Host components: initial partitioner (not METIS), generators, CSR I/O.
#include "gp_host.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <thread>
#include <tuple>

namespace gp {
namespace {

inline std::uint32_t hmix32(std::uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    return x ^ (x >> 16);
}
inline std::uint64_t splitmix64(std::uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
inline double unit_double(std::uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

int worker_count() {
    unsigned h = std::thread::hardware_concurrency();
    if (!h) h = 4;
    return static_cast<int>(std::min(16u, h));
}

template <typename Fn>
void parallel_for(std::int64_t n, Fn fn) {
    const int t = worker_count();
    if (n < 4096 || t == 1) {
        fn(0, n, 0);
        return;
    }
    std::vector<std::thread> th;
    for (int i = 0; i < t; ++i) {
        const std::int64_t b = n * i / t, e = n * (i + 1) / t;
        th.emplace_back([=] { fn(b, e, i); });
    }
    for (auto& x : th) x.join();
}

// Build a sorted, deduplicated, loop-free symmetric CSR from directed pairs.
HostGraph build_csr(std::int64_t n, std::vector<std::vector<std::uint64_t>>& parts_edges) {
    HostGraph g;
    g.n = n;
    std::vector<std::atomic<std::uint32_t>> deg(static_cast<std::size_t>(n));
    for (auto& d : deg) d.store(0, std::memory_order_relaxed);
    const int T = static_cast<int>(parts_edges.size());
    {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                for (auto key : parts_edges[static_cast<std::size_t>(t)])
                    deg[static_cast<std::size_t>(key >> 32)].fetch_add(1, std::memory_order_relaxed);
            });
        for (auto& x : th) x.join();
    }
    std::vector<std::int64_t> start(static_cast<std::size_t>(n) + 1, 0);
    for (std::int64_t v = 0; v < n; ++v) start[v + 1] = start[v] + deg[static_cast<std::size_t>(v)].load();
    std::vector<std::int32_t> raw(static_cast<std::size_t>(start[n]));
    std::vector<std::atomic<std::int64_t>> cur(static_cast<std::size_t>(n));
    for (std::int64_t v = 0; v < n; ++v) cur[static_cast<std::size_t>(v)].store(start[v], std::memory_order_relaxed);
    {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                for (auto key : parts_edges[static_cast<std::size_t>(t)]) {
                    const std::int64_t u = static_cast<std::int64_t>(key >> 32);
                    const std::int64_t p = cur[static_cast<std::size_t>(u)].fetch_add(1, std::memory_order_relaxed);
                    raw[static_cast<std::size_t>(p)] = static_cast<std::int32_t>(key & 0xffffffffULL);
                }
            });
        for (auto& x : th) x.join();
    }
    for (auto& v : parts_edges) std::vector<std::uint64_t>().swap(v);
    std::vector<std::int64_t> ucount(static_cast<std::size_t>(n));
    parallel_for(n, [&](std::int64_t b, std::int64_t e, int) {
        for (std::int64_t v = b; v < e; ++v) {
            auto first = raw.begin() + start[v], last = raw.begin() + start[v + 1];
            std::sort(first, last);
            auto end = std::unique(first, last);
            ucount[static_cast<std::size_t>(v)] = end - first;
        }
    });
    g.off.assign(static_cast<std::size_t>(n) + 1, 0);
    std::int64_t total = 0;
    for (std::int64_t v = 0; v < n; ++v) {
        g.off[static_cast<std::size_t>(v)] = static_cast<std::int32_t>(total);
        total += ucount[static_cast<std::size_t>(v)];
        if (total > std::numeric_limits<std::int32_t>::max()) throw std::runtime_error("generated graph exceeds int32");
    }
    g.off[static_cast<std::size_t>(n)] = static_cast<std::int32_t>(total);
    g.nbr.resize(static_cast<std::size_t>(total));
    parallel_for(n, [&](std::int64_t b, std::int64_t e, int) {
        for (std::int64_t v = b; v < e; ++v)
            std::copy(raw.begin() + start[v], raw.begin() + start[v] + ucount[static_cast<std::size_t>(v)],
                      g.nbr.begin() + g.off[static_cast<std::size_t>(v)]);
    });
    return g;
}

// ------------------------------------------------------------ partitioner
// In-repo initial partitioner of the coarsest graph (the role METIS 5 has in
// the reference, src/initial_partition_metis.cu:86-183; it is not METIS):
// recursive bisection.  Each bisection grows breadth-first regions from
// hashed start vertices (kGrowths, or kWideGrowths on sparse coarse graphs),
// refines the best of them (kRefined, or kWideRefined; ranked by cut, then
// deviation from the target weight) with Fiduccia-Mattheyses passes and
// keeps the best.  The whole recursion runs from kRecursions seeds on
// threads of their own and the labels of least capacity excess, then least
// cut, are kept.  Deterministic for a given seed and independent of timing.
// The wider search pays on sparse coarse graphs (meshes: a bisection's cut
// depends on where its regions start) and is cheap there; coarse graphs of
// at most kWideEntries row entries and average degree at most kWideDegree
// take it.  Dense coarse cores (power-law graphs) keep the narrow search:
// on R-MAT scale 22 the wide one moved the cut by 0.01 % and, through a
// different start, added two refinement rounds at the finest level.
constexpr int kGrowths = 8;
constexpr int kRefined = 2;
constexpr int kWideGrowths = 16;
constexpr int kWideRefined = 4;
constexpr std::int64_t kWideEntries = 1 << 14;
constexpr std::int64_t kWideDegree = 16;
constexpr int kRecursions = 8;
constexpr int kPasses = 8;

struct Bisector {
    std::int64_t n;
    const std::int32_t* off;
    const std::int32_t* nbr;
    const std::uint32_t* ew;
    const std::uint32_t* vw;
    std::uint32_t seed;
    std::int32_t* labels;
    std::vector<std::int32_t> pos;  // subset slot of each vertex, -1 outside
    int growths_per_bisection;
    int refined_per_bisection;

    std::uint64_t w(std::int32_t v) const { return vw ? vw[v] : 1u; }
    std::int64_t e(std::int64_t x) const { return ew ? static_cast<std::int64_t>(ew[x]) : 1; }
    static std::uint64_t dev(std::uint64_t a, std::uint64_t b) { return a > b ? a - b : b - a; }

    // Indexed binary max-heap of subset slots keyed by (gain, lower vertex id).
    struct GainHeap {
        std::vector<std::int32_t> heap, where;  // where[i] = heap slot of i, -1 outside
        const std::vector<std::int64_t>* gain = nullptr;
        const std::vector<std::int32_t>* ids = nullptr;
        bool above(std::int32_t x, std::int32_t y) const {
            const std::int64_t gx = (*gain)[static_cast<std::size_t>(x)], gy = (*gain)[static_cast<std::size_t>(y)];
            return gx != gy ? gx > gy : (*ids)[static_cast<std::size_t>(x)] < (*ids)[static_cast<std::size_t>(y)];
        }
        void place(std::size_t slot, std::int32_t x) {
            heap[slot] = x;
            where[static_cast<std::size_t>(x)] = static_cast<std::int32_t>(slot);
        }
        void up(std::size_t slot) {
            const std::int32_t x = heap[slot];
            while (slot && above(x, heap[(slot - 1) / 2])) {
                place(slot, heap[(slot - 1) / 2]);
                slot = (slot - 1) / 2;
            }
            place(slot, x);
        }
        void down(std::size_t slot) {
            const std::int32_t x = heap[slot];
            for (;;) {
                std::size_t c = 2 * slot + 1;
                if (c >= heap.size()) break;
                if (c + 1 < heap.size() && above(heap[c + 1], heap[c])) ++c;
                if (!above(heap[c], x)) break;
                place(slot, heap[c]);
                slot = c;
            }
            place(slot, x);
        }
        bool has(std::int32_t x) const { return where[static_cast<std::size_t>(x)] >= 0; }
        void push(std::int32_t x) {
            heap.push_back(x);
            up(heap.size() - 1);
        }
        void update(std::int32_t x) {
            up(static_cast<std::size_t>(where[static_cast<std::size_t>(x)]));
            down(static_cast<std::size_t>(where[static_cast<std::size_t>(x)]));
        }
        std::int32_t pop() {
            const std::int32_t top = heap[0];
            where[static_cast<std::size_t>(top)] = -1;
            const std::int32_t last = heap.back();
            heap.pop_back();
            if (!heap.empty()) {
                place(0, last);
                down(0);
            }
            return top;
        }
    };

    // Flip slot i to the other side and bring its neighbours' gains (weight
    // to the other side minus weight to their own) up to date: a neighbour
    // that i joined loses 2w, one that i left gains 2w.  Calls touch(j) for
    // each neighbour slot j.
    template <typename Touch>
    void flip(const std::vector<std::int32_t>& S, std::vector<std::uint8_t>& side, std::vector<std::int64_t>& gain,
              std::int32_t i, Touch touch) {
        const std::size_t ti = static_cast<std::size_t>(i);
        side[ti] ^= 1;
        gain[ti] = -gain[ti];
        const std::int32_t v = S[ti];
        for (std::int64_t x = off[v]; x < off[v + 1]; ++x) {
            const std::int32_t j = pos[static_cast<std::size_t>(nbr[x])];
            if (j < 0 || j == i) continue;
            gain[static_cast<std::size_t>(j)] += side[static_cast<std::size_t>(j)] == side[ti] ? -2 * e(x) : 2 * e(x);
            touch(j);
        }
    }

    // Fiduccia-Mattheyses passes over the bisection of the subset S with
    // cut `cut`: each pass moves, one at a time, the unlocked vertex of
    // largest gain among the boundary and the neighbours of moved vertices
    // (lowest id on ties) whose move keeps the side-0 weight within tol of
    // the target or brings it closer, and locks it.  A pass ends after
    // min(max(m/100, 15), 100) moves without a new best and keeps its best
    // prefix (balanced before unbalanced, then smaller cut, then smaller
    // deviation), flipping the rest back; passes repeat while they improve,
    // at most kPasses.  Gains carry over between passes.  Returns the cut.
    std::int64_t refine(const std::vector<std::int32_t>& S, std::vector<std::uint8_t>& side, std::uint64_t& wa,
                        std::uint64_t target, std::uint64_t tol, std::int64_t cut) {
        const std::int32_t m = static_cast<std::int32_t>(S.size());
        const int stall = std::min(std::max(m / 100, 15), 100);
        std::vector<std::int64_t> gain(static_cast<std::size_t>(m), 0), internal(static_cast<std::size_t>(m), 0);
        for (std::int32_t i = 0; i < m; ++i) {
            const std::int32_t v = S[static_cast<std::size_t>(i)];
            for (std::int64_t x = off[v]; x < off[v + 1]; ++x) {
                const std::int32_t j = pos[static_cast<std::size_t>(nbr[x])];
                if (j < 0 || j == i) continue;
                const bool other = side[static_cast<std::size_t>(j)] != side[static_cast<std::size_t>(i)];
                gain[static_cast<std::size_t>(i)] += other ? e(x) : -e(x);
                internal[static_cast<std::size_t>(i)] += e(x);
            }
        }
        std::vector<std::uint8_t> locked(static_cast<std::size_t>(m));
        std::vector<std::int32_t> moves;
        GainHeap H;
        H.gain = &gain;
        H.ids = &S;
        H.where.assign(static_cast<std::size_t>(m), -1);
        const auto key = [&](std::uint64_t x, std::int64_t c) {
            return std::make_tuple(dev(x, target) > tol ? 1 : 0, c, dev(x, target));
        };
        for (int pass = 0; pass < kPasses; ++pass) {
            H.heap.clear();
            for (std::int32_t i = 0; i < m; ++i) {
                locked[static_cast<std::size_t>(i)] = 0;
                H.where[static_cast<std::size_t>(i)] = -1;
                // on the boundary: some weight to the other side
                if (gain[static_cast<std::size_t>(i)] > -internal[static_cast<std::size_t>(i)]) H.push(i);
            }
            moves.clear();
            auto best = key(wa, cut);
            std::size_t best_len = 0;
            std::uint64_t cur_wa = wa;
            std::int64_t cur_cut = cut;
            int since_best = 0;
            while (!H.heap.empty() && since_best <= stall) {
                const std::int32_t i = H.pop();
                const std::size_t ti = static_cast<std::size_t>(i);
                const std::uint64_t wv = w(S[ti]);
                const std::uint64_t nwa = side[ti] == 0 ? cur_wa - wv : cur_wa + wv;
                if (dev(nwa, target) > tol && dev(nwa, target) >= dev(cur_wa, target)) continue;
                locked[ti] = 1;
                cur_wa = nwa;
                cur_cut -= gain[ti];
                moves.push_back(i);
                flip(S, side, gain, i, [&](std::int32_t j) {
                    if (locked[static_cast<std::size_t>(j)]) return;
                    if (H.has(j)) H.update(j);
                    else H.push(j);
                });
                const auto now = key(cur_wa, cur_cut);
                if (now < best) {
                    best = now;
                    best_len = moves.size();
                    since_best = 0;
                } else {
                    ++since_best;
                }
            }
            for (std::size_t q = moves.size(); q > best_len; --q) {
                const std::int32_t i = moves[q - 1];
                const std::uint64_t wv = w(S[static_cast<std::size_t>(i)]);
                cur_wa = side[static_cast<std::size_t>(i)] == 0 ? cur_wa - wv : cur_wa + wv;
                flip(S, side, gain, i, [](std::int32_t) {});
            }
            wa = cur_wa;
            cut = std::get<1>(best);
            if (best_len == 0) break;
        }
        return cut;
    }

    void bisect(std::vector<std::int32_t>& S, int p0, int kk) {
        if (S.empty()) return;
        if (kk == 1) {
            for (auto v : S) labels[v] = p0;
            return;
        }
        const int k1 = kk / 2;
        std::uint64_t W = 0, maxw = 0;
        for (auto v : S) {
            W += w(v);
            maxw = std::max(maxw, w(v));
        }
        const std::uint64_t target = (W / kk) * k1 + ((W % kk) * k1) / kk;
        const std::uint64_t tol = std::max<std::uint64_t>(maxw, W / 100);
        const std::int32_t m = static_cast<std::int32_t>(S.size());
        for (std::int32_t i = 0; i < m; ++i) pos[static_cast<std::size_t>(S[static_cast<std::size_t>(i)])] = i;
        // Breadth-first growths of side 0 from hashed starts (restarting at
        // the lowest unvisited slot when a component is exhausted) until it
        // holds the target weight; the cut is kept up to date as vertices
        // join side 0.
        std::vector<std::vector<std::uint8_t>> grown;
        std::vector<std::tuple<std::int64_t, std::uint64_t, int, std::uint64_t>> rank;  // cut, dev, growth, weight
        std::vector<std::uint8_t> visited;
        std::vector<std::int32_t> queue;
        const int growths = std::min<std::int32_t>(growths_per_bisection, m);
        for (int t = 0; t < growths; ++t) {
            const std::uint32_t h = hmix32(seed ^ ((std::uint32_t)p0 * 0x9e3779b9U) ^ ((std::uint32_t)kk * 0x85ebca6bU) ^
                                           ((std::uint32_t)t * 0xc2b2ae35U));
            std::vector<std::uint8_t> side(static_cast<std::size_t>(m), 1);
            visited.assign(static_cast<std::size_t>(m), 0);
            queue.clear();
            std::size_t head = 0;
            std::int32_t next_unvisited = 0;
            const std::int32_t s0 = static_cast<std::int32_t>(h % static_cast<std::uint32_t>(m));
            queue.push_back(s0);
            visited[static_cast<std::size_t>(s0)] = 1;
            std::uint64_t wa = 0;
            std::int64_t cut = 0;
            while (wa < target) {
                if (head == queue.size()) {
                    while (next_unvisited < m && visited[static_cast<std::size_t>(next_unvisited)]) ++next_unvisited;
                    if (next_unvisited == m) break;
                    visited[static_cast<std::size_t>(next_unvisited)] = 1;
                    queue.push_back(next_unvisited);
                }
                const std::int32_t i = queue[head++];
                side[static_cast<std::size_t>(i)] = 0;
                const std::int32_t v = S[static_cast<std::size_t>(i)];
                wa += w(v);
                for (std::int64_t x = off[v]; x < off[v + 1]; ++x) {
                    const std::int32_t j = pos[static_cast<std::size_t>(nbr[x])];
                    if (j < 0 || j == i) continue;
                    cut += side[static_cast<std::size_t>(j)] == 0 ? -e(x) : e(x);
                    if (!visited[static_cast<std::size_t>(j)]) {
                        visited[static_cast<std::size_t>(j)] = 1;
                        queue.push_back(j);
                    }
                }
            }
            rank.emplace_back(cut, dev(wa, target), t, wa);
            grown.push_back(std::move(side));
        }
        std::sort(rank.begin(), rank.end());
        std::int64_t best_cut = std::numeric_limits<std::int64_t>::max();
        std::uint64_t best_dev = std::numeric_limits<std::uint64_t>::max();
        std::vector<std::uint8_t> best_side;
        for (std::size_t r = 0; r < rank.size() && r < static_cast<std::size_t>(refined_per_bisection); ++r) {
            std::vector<std::uint8_t>& side = grown[static_cast<std::size_t>(std::get<2>(rank[r]))];
            std::uint64_t wa = std::get<3>(rank[r]);
            const std::int64_t cut = refine(S, side, wa, target, tol, std::get<0>(rank[r]));
            const std::uint64_t d = dev(wa, target);
            if (cut < best_cut || (cut == best_cut && d < best_dev)) {
                best_cut = cut;
                best_dev = d;
                best_side = side;
            }
        }
        std::vector<std::int32_t> A, B;
        for (std::int32_t i = 0; i < m; ++i) {
            (best_side[static_cast<std::size_t>(i)] == 0 ? A : B).push_back(S[static_cast<std::size_t>(i)]);
            pos[static_cast<std::size_t>(S[static_cast<std::size_t>(i)])] = -1;
        }
        std::vector<std::int32_t>().swap(S);
        bisect(A, p0, k1);
        bisect(B, p0 + k1, kk - k1);
    }
};

}  // namespace

void initial_partition_host(std::int64_t n, const std::int32_t* off, const std::int32_t* nbr, const std::uint32_t* ew,
                            const std::uint32_t* vw, int parts, double imbalance, std::uint32_t seed,
                            std::int32_t* labels) {
    if (n <= 0 || parts < 2) throw std::invalid_argument("invalid initial partition input");
    std::uint64_t total = 0;
    for (std::int64_t v = 0; v < n; ++v) total += vw ? vw[v] : 1u;
    const long double raw = (long double)total * (imbalance >= 1.0 ? imbalance : 1.0) / parts;
    const std::uint64_t cap = (std::uint64_t)std::floor(raw);
    const bool wide = off[n] <= kWideEntries && off[n] <= kWideDegree * n;
    std::vector<std::vector<std::int32_t>> trial(static_cast<std::size_t>(kRecursions));
    std::vector<std::pair<std::uint64_t, std::uint64_t>> score(static_cast<std::size_t>(kRecursions));
    auto run = [&](int t) {
        const std::uint32_t s = t == 0 ? seed : hmix32(seed ^ (0x68e31da4U * (std::uint32_t)t));
        std::vector<std::int32_t>& lab = trial[static_cast<std::size_t>(t)];
        lab.assign(static_cast<std::size_t>(n), 0);
        Bisector b{n,   off, nbr, ew, vw, s, lab.data(), std::vector<std::int32_t>(static_cast<std::size_t>(n), -1),
                   wide ? kWideGrowths : kGrowths, wide ? kWideRefined : kRefined};
        std::vector<std::int32_t> all(static_cast<std::size_t>(n));
        for (std::int64_t v = 0; v < n; ++v) all[static_cast<std::size_t>(v)] = static_cast<std::int32_t>(v);
        b.bisect(all, 0, parts);
        std::vector<std::uint64_t> loads(static_cast<std::size_t>(parts), 0);
        std::uint64_t directed = 0;
        for (std::int64_t v = 0; v < n; ++v) {
            loads[static_cast<std::size_t>(lab[static_cast<std::size_t>(v)])] += vw ? vw[v] : 1u;
            for (std::int64_t x = off[v]; x < off[v + 1]; ++x)
                if (lab[static_cast<std::size_t>(nbr[x])] != lab[static_cast<std::size_t>(v)]) directed += ew ? ew[x] : 1u;
        }
        std::uint64_t over = 0;
        for (auto l : loads) over += l > cap ? l - cap : 0;
        score[static_cast<std::size_t>(t)] = {over, directed / 2};
    };
    std::vector<std::thread> threads;
    for (int t = 1; t < kRecursions; ++t) threads.emplace_back(run, t);
    run(0);
    for (auto& th : threads) th.join();
    int best = 0;
    for (int t = 1; t < kRecursions; ++t)
        if (score[static_cast<std::size_t>(t)] < score[static_cast<std::size_t>(best)]) best = t;
    std::copy(trial[static_cast<std::size_t>(best)].begin(), trial[static_cast<std::size_t>(best)].end(), labels);
}

// ------------------------------------------------------------- generators
HostGraph generate_grid(std::int64_t rows, std::int64_t cols) {
    HostGraph g;
    g.n = rows * cols;
    g.off.reserve(static_cast<std::size_t>(g.n) + 1);
    g.nbr.reserve(static_cast<std::size_t>(4 * g.n));
    g.off.push_back(0);
    for (std::int64_t r = 0; r < rows; ++r)
        for (std::int64_t c = 0; c < cols; ++c) {
            if (r > 0) g.nbr.push_back(static_cast<std::int32_t>((r - 1) * cols + c));
            if (c > 0) g.nbr.push_back(static_cast<std::int32_t>(r * cols + c - 1));
            if (c + 1 < cols) g.nbr.push_back(static_cast<std::int32_t>(r * cols + c + 1));
            if (r + 1 < rows) g.nbr.push_back(static_cast<std::int32_t>((r + 1) * cols + c));
            g.off.push_back(static_cast<std::int32_t>(g.nbr.size()));
        }
    return g;
}

HostGraph generate_rmat(int scale, int edge_factor, double a, double b, double c, std::uint64_t seed) {
    if (scale < 1 || scale > 30) throw std::invalid_argument("rmat scale out of range");
    const std::int64_t n = std::int64_t(1) << scale;
    const std::int64_t samples = n * edge_factor;
    const int T = worker_count();
    std::vector<std::vector<std::uint64_t>> parts(static_cast<std::size_t>(T));
    std::vector<std::thread> th;
    const double ab = a + b, abc = a + b + c;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            const std::int64_t lo = samples * t / T, hi = samples * (t + 1) / T;
            auto& out = parts[static_cast<std::size_t>(t)];
            out.reserve(static_cast<std::size_t>(2 * (hi - lo)));
            for (std::int64_t i = lo; i < hi; ++i) {
                std::uint64_t u = 0, v = 0;
                const std::uint64_t base = splitmix64(seed * 0x9e3779b97f4a7c15ULL ^ (std::uint64_t)i);
                for (int bit = 0; bit < scale; ++bit) {
                    const double r = unit_double(splitmix64(base + (std::uint64_t)bit * 0xd1b54a32d192ed03ULL));
                    const std::uint64_t qu = r >= ab, qv = (r >= a && r < ab) || r >= abc;
                    u = (u << 1) | qu;
                    v = (v << 1) | qv;
                }
                if (u == v) continue;
                out.push_back((u << 32) | v);
                out.push_back((v << 32) | u);
            }
        });
    for (auto& x : th) x.join();
    return build_csr(n, parts);
}

HostGraph generate_rgg2d(std::int64_t n, double radius, std::uint64_t seed) {
    std::vector<double> x(static_cast<std::size_t>(n)), y(static_cast<std::size_t>(n));
    for (std::int64_t i = 0; i < n; ++i) {
        x[static_cast<std::size_t>(i)] = unit_double(splitmix64(seed ^ (2 * (std::uint64_t)i)));
        y[static_cast<std::size_t>(i)] = unit_double(splitmix64(seed ^ (2 * (std::uint64_t)i + 1)));
    }
    const std::int64_t cells = std::max<std::int64_t>(1, (std::int64_t)std::floor(1.0 / radius));
    std::vector<std::int64_t> cstart(static_cast<std::size_t>(cells * cells) + 1, 0);
    auto cell_of = [&](std::int64_t i) {
        std::int64_t cx = std::min<std::int64_t>(cells - 1, (std::int64_t)(x[static_cast<std::size_t>(i)] * cells));
        std::int64_t cy = std::min<std::int64_t>(cells - 1, (std::int64_t)(y[static_cast<std::size_t>(i)] * cells));
        return cx * cells + cy;
    };
    for (std::int64_t i = 0; i < n; ++i) ++cstart[static_cast<std::size_t>(cell_of(i)) + 1];
    for (std::size_t i = 1; i < cstart.size(); ++i) cstart[i] += cstart[i - 1];
    std::vector<std::int32_t> cell_pts(static_cast<std::size_t>(n));
    {
        std::vector<std::int64_t> cur(cstart.begin(), cstart.end() - 1);
        for (std::int64_t i = 0; i < n; ++i)
            cell_pts[static_cast<std::size_t>(cur[static_cast<std::size_t>(cell_of(i))]++)] = static_cast<std::int32_t>(i);
    }
    const int T = worker_count();
    std::vector<std::vector<std::uint64_t>> parts(static_cast<std::size_t>(T));
    const double r2 = radius * radius;
    parallel_for(n, [&](std::int64_t b, std::int64_t e, int t) {
        auto& out = parts[static_cast<std::size_t>(t)];
        for (std::int64_t i = b; i < e; ++i) {
            const std::int64_t c = cell_of(i), cx = c / cells, cy = c % cells;
            for (std::int64_t dx = -1; dx <= 1; ++dx)
                for (std::int64_t dy = -1; dy <= 1; ++dy) {
                    const std::int64_t nx = cx + dx, ny = cy + dy;
                    if (nx < 0 || ny < 0 || nx >= cells || ny >= cells) continue;
                    const std::int64_t cc = nx * cells + ny;
                    for (std::int64_t p = cstart[static_cast<std::size_t>(cc)]; p < cstart[static_cast<std::size_t>(cc) + 1]; ++p) {
                        const std::int32_t j = cell_pts[static_cast<std::size_t>(p)];
                        if (j == i) continue;
                        const double ddx = x[static_cast<std::size_t>(i)] - x[static_cast<std::size_t>(j)];
                        const double ddy = y[static_cast<std::size_t>(i)] - y[static_cast<std::size_t>(j)];
                        if (ddx * ddx + ddy * ddy <= r2) out.push_back(((std::uint64_t)i << 32) | (std::uint32_t)j);
                    }
                }
        }
    });
    return build_csr(n, parts);
}

HostGraph generate_knn3d(std::int64_t n, int knn, std::uint64_t seed) {
    std::vector<double> P(static_cast<std::size_t>(3 * n));
    for (std::int64_t i = 0; i < 3 * n; ++i) P[static_cast<std::size_t>(i)] = unit_double(splitmix64(seed ^ (std::uint64_t)i));
    const std::int64_t side = std::max<std::int64_t>(1, (std::int64_t)std::cbrt((double)n / 2.0));
    auto cell_of = [&](std::int64_t i) {
        std::int64_t c[3];
        for (int d = 0; d < 3; ++d)
            c[d] = std::min<std::int64_t>(side - 1, (std::int64_t)(P[static_cast<std::size_t>(3 * i + d)] * side));
        return (c[0] * side + c[1]) * side + c[2];
    };
    const std::int64_t ncell = side * side * side;
    std::vector<std::int64_t> cstart(static_cast<std::size_t>(ncell) + 1, 0);
    for (std::int64_t i = 0; i < n; ++i) ++cstart[static_cast<std::size_t>(cell_of(i)) + 1];
    for (std::size_t i = 1; i < cstart.size(); ++i) cstart[i] += cstart[i - 1];
    std::vector<std::int32_t> pts(static_cast<std::size_t>(n));
    {
        std::vector<std::int64_t> cur(cstart.begin(), cstart.end() - 1);
        for (std::int64_t i = 0; i < n; ++i)
            pts[static_cast<std::size_t>(cur[static_cast<std::size_t>(cell_of(i))]++)] = static_cast<std::int32_t>(i);
    }
    const int T = worker_count();
    std::vector<std::vector<std::uint64_t>> parts(static_cast<std::size_t>(T));
    parallel_for(n, [&](std::int64_t b, std::int64_t e, int t) {
        auto& out = parts[static_cast<std::size_t>(t)];
        std::vector<std::pair<double, std::int32_t>> best;
        for (std::int64_t i = b; i < e; ++i) {
            const std::int64_t c = cell_of(i);
            const std::int64_t cz = c % side, cy = (c / side) % side, cx = c / (side * side);
            best.clear();
            for (std::int64_t ring = 0;; ++ring) {
                for (std::int64_t dx = -ring; dx <= ring; ++dx)
                    for (std::int64_t dy = -ring; dy <= ring; ++dy)
                        for (std::int64_t dz = -ring; dz <= ring; ++dz) {
                            if (std::max({std::llabs(dx), std::llabs(dy), std::llabs(dz)}) != ring) continue;
                            const std::int64_t nx = cx + dx, ny = cy + dy, nz = cz + dz;
                            if (nx < 0 || ny < 0 || nz < 0 || nx >= side || ny >= side || nz >= side) continue;
                            const std::int64_t cc = (nx * side + ny) * side + nz;
                            for (std::int64_t p = cstart[static_cast<std::size_t>(cc)]; p < cstart[static_cast<std::size_t>(cc) + 1]; ++p) {
                                const std::int32_t j = pts[static_cast<std::size_t>(p)];
                                if (j == i) continue;
                                double d2 = 0;
                                for (int d = 0; d < 3; ++d) {
                                    const double q = P[static_cast<std::size_t>(3 * i + d)] - P[static_cast<std::size_t>(3 * j + d)];
                                    d2 += q * q;
                                }
                                best.emplace_back(d2, j);
                            }
                        }
                if ((std::int64_t)best.size() >= knn) {
                    std::nth_element(best.begin(), best.begin() + (knn - 1), best.end());
                    const double kth = best[static_cast<std::size_t>(knn - 1)].first;
                    const double reach = (double)ring / (double)side;
                    if (kth <= reach * reach || ring > side) break;
                }
                if (ring > side) break;
            }
            std::sort(best.begin(), best.end());
            for (int q = 0; q < knn && q < (int)best.size(); ++q) {
                const std::int32_t j = best[static_cast<std::size_t>(q)].second;
                out.push_back(((std::uint64_t)i << 32) | (std::uint32_t)j);
                out.push_back(((std::uint64_t)j << 32) | (std::uint32_t)i);
            }
        }
    });
    HostGraph g = build_csr(n, parts);
    g.ew.resize(g.nbr.size());
    for (std::int64_t v = 0; v < n; ++v)
        for (std::int64_t x = g.off[static_cast<std::size_t>(v)]; x < g.off[static_cast<std::size_t>(v) + 1]; ++x) {
            const std::uint64_t u = static_cast<std::uint64_t>(g.nbr[static_cast<std::size_t>(x)]);
            const std::uint64_t lo = std::min<std::uint64_t>(u, v), hi = std::max<std::uint64_t>(u, v);
            g.ew[static_cast<std::size_t>(x)] =
                1 + (std::uint32_t)(splitmix64((lo << 32 | hi) ^ (seed + 3) * 0x9e3779b97f4a7c15ULL) % 100);
        }
    g.vw.resize(static_cast<std::size_t>(n));
    for (std::int64_t v = 0; v < n; ++v)
        g.vw[static_cast<std::size_t>(v)] = 1 + (std::uint32_t)(splitmix64((std::uint64_t)v ^ (seed + 4) * 0xbf58476d1ce4e5b9ULL) % 4);
    return g;
}

// ------------------------------------------------------------------ I/O
// Reads and validates the reference's int64 CSR (src/graph.cpp:47-116).
HostGraph load_csr(const std::string& indptr_path, const std::string& indices_path) {
    auto count = [](const std::string& path) {
        std::ifstream in(path, std::ios::binary | std::ios::ate);
        if (!in) throw std::runtime_error("cannot open " + path);
        const auto bytes = in.tellg();
        if (bytes < 0 || bytes % 8 != 0) throw std::runtime_error("invalid binary size " + path);
        return static_cast<std::int64_t>(bytes / 8);
    };
    const std::int64_t offs = count(indptr_path), m = count(indices_path);
    if (offs < 1) throw std::runtime_error("empty CSR indptr");
    const std::int64_t n = offs - 1;
    if (n > std::numeric_limits<std::int32_t>::max() || m > std::numeric_limits<std::int32_t>::max())
        throw std::runtime_error("graph exceeds supported 32-bit CSR limits; dataset excluded");
    std::vector<std::int64_t> o(static_cast<std::size_t>(offs));
    {
        std::ifstream in(indptr_path, std::ios::binary);
        in.read(reinterpret_cast<char*>(o.data()), offs * 8);
        if (!in) throw std::runtime_error("cannot read " + indptr_path);
    }
    if (o.front() != 0 || o.back() != m) throw std::runtime_error("CSR offsets do not match indices");
    HostGraph g;
    g.n = n;
    g.off.resize(static_cast<std::size_t>(offs));
    for (std::int64_t i = 0; i < offs; ++i) {
        if (o[static_cast<std::size_t>(i)] < 0 || (i > 0 && o[static_cast<std::size_t>(i)] < o[static_cast<std::size_t>(i - 1)]))
            throw std::runtime_error("CSR offsets are not monotone");
        g.off[static_cast<std::size_t>(i)] = static_cast<std::int32_t>(o[static_cast<std::size_t>(i)]);
    }
    g.nbr.resize(static_cast<std::size_t>(m));
    std::ifstream in(indices_path, std::ios::binary);
    std::vector<std::int64_t> raw(std::size_t(1) << 20);
    std::int64_t done = 0;
    while (done < m) {
        const std::int64_t take = std::min<std::int64_t>(static_cast<std::int64_t>(raw.size()), m - done);
        in.read(reinterpret_cast<char*>(raw.data()), take * 8);
        if (!in) throw std::runtime_error("cannot read " + indices_path);
        for (std::int64_t i = 0; i < take; ++i) {
            const std::int64_t u = raw[static_cast<std::size_t>(i)];
            if (u < 0 || u >= n) throw std::runtime_error("CSR contains an invalid neighbor");
            g.nbr[static_cast<std::size_t>(done + i)] = static_cast<std::int32_t>(u);
        }
        done += take;
    }
    return g;
}

void save_csr(const HostGraph& g, const std::string& indptr_path, const std::string& indices_path) {
    std::ofstream a(indptr_path, std::ios::binary | std::ios::trunc), b(indices_path, std::ios::binary | std::ios::trunc);
    if (!a || !b) throw std::runtime_error("cannot open CSR output");
    for (auto x : g.off) {
        const std::int64_t v = x;
        a.write(reinterpret_cast<const char*>(&v), 8);
    }
    for (auto x : g.nbr) {
        const std::int64_t v = x;
        b.write(reinterpret_cast<const char*>(&v), 8);
    }
    if (!a || !b) throw std::runtime_error("cannot write CSR output");
}

std::uint64_t host_cut(const HostGraph& g, const std::int32_t* labels) {
    std::uint64_t d = 0;
    for (std::int64_t v = 0; v < g.n; ++v)
        for (std::int64_t x = g.off[static_cast<std::size_t>(v)]; x < g.off[static_cast<std::size_t>(v) + 1]; ++x)
            if (labels[v] != labels[g.nbr[static_cast<std::size_t>(x)]]) d += g.ew.empty() ? 1 : g.ew[static_cast<std::size_t>(x)];
    return d / 2;
}

std::vector<std::uint64_t> host_loads(const HostGraph& g, const std::int32_t* labels, int parts) {
    std::vector<std::uint64_t> l(static_cast<std::size_t>(parts), 0);
    for (std::int64_t v = 0; v < g.n; ++v) l[static_cast<std::size_t>(labels[v])] += g.vw.empty() ? 1 : g.vw[static_cast<std::size_t>(v)];
    return l;
}

std::uint64_t capacity_of(std::uint64_t total, int parts, double imbalance) {
    if (parts < 2 || parts > 128 || !std::isfinite(imbalance) || imbalance < 1.0)
        throw std::invalid_argument("invalid partition capacity options");
    const long double raw = (long double)total * imbalance / parts;
    if (!std::isfinite((double)raw) || raw >= (long double)std::numeric_limits<std::uint64_t>::max())
        throw std::overflow_error("partition capacity exceeds uint64");
    const std::uint64_t cap = (std::uint64_t)std::floor(raw);
    if (cap < total / (unsigned)parts + (total % (unsigned)parts != 0))
        throw std::invalid_argument("integer partition capacity cannot hold total weight");
    return cap;
}

}  // namespace gp
