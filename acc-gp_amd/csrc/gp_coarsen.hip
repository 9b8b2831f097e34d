// Coarsening on gfx950: stable-receiver aggregation (seed matching, frozen
// roles, exact top-two cluster affinity, regret-ordered weight-prefix
// admission, isolated packing, two-hop pairing, compaction) and exact
// coarse-CSR contraction.  Restates src/coarsen.cu and src/coarsening/*.cuh
// of the reference with wave64 kernels and a canonical (vertex-id) order.
#include "gp_coarsen.hpp"
#include "gp_segsort.hpp"
#include "gp_rows.hpp"

#include <algorithm>
#include <cstdio>

namespace gp {
namespace {

constexpr int kLowDegree = 16;     // group of 16 lanes
constexpr int kWaveDegree = 512;   // one wave, up to 8 items per lane
constexpr int kBlockDegree = 4096; // one block, row held in LDS
static_assert(kBlockDegree <= kBlockSortMax, "block-class rows fit one LDS tile");

__device__ inline bool sclp_better(std::uint64_t a, std::uint64_t t, std::uint64_t oa, std::uint64_t ot) {
    return a > oa || (a == oa && a != 0 && t > ot);  // aggregation.cuh:79-84
}

// ------------------------------------------------------------ metadata
// record_bound_kernel (src/coarsen.cu:51-99): saturating incident weight
// sums, unsorted/zero/non-unit flags and the maximum raw degree.  Each lane
// owns a vertex; rows longer than 64 are scanned cooperatively by the wave.
struct MetaOut {
    unsigned long long max_row;   // 0
    unsigned long long unsorted;  // 1
    unsigned long long flags;     // 2: bit0 zero weight, bit1 non-unit
    unsigned long long max_deg;   // 3
    unsigned long long heavy;     // 4: any degree >= 4096
    unsigned long long total_vw;  // 5
    unsigned long long max_vw;    // 6
};

__device__ inline std::uint64_t sat_add(std::uint64_t a, std::uint64_t b) {
    return ~0ULL - a < b ? ~0ULL : a + b;
}

// Row sums (exact: u32 weights over int32 rows stay below 2^63) and the
// sorted-row check; the body's accumulator counts unsorted rows.
struct MetaBody {
    const std::int32_t* off;
    const std::int32_t* nbr;
    const std::uint32_t* ew;
    std::uint64_t* row_bound;
    __device__ std::int64_t degree(std::int32_t v) const { return off[v + 1] - off[v]; }
    template <class Team>
    __device__ RowOut operator()(const Team& t, std::int32_t v, unsigned long long*) const {
        const std::int64_t b = off[v], e = off[v + 1];
        unsigned long long d = 0;
        unsigned long long unsorted = 0;
        for (std::int64_t x = b + t.rank(); x < e; x += Team::size) {
            d += ew_at(ew, x);
            if (x > b && nbr[x] <= nbr[x - 1]) unsorted = 1;
        }
        d = t.sum(d);
        unsorted = t.sum(unsorted);
        if (t.leader()) row_bound[v] = d;
        return RowOut{false, unsorted ? 1 : 0};
    }
};

// Per-vertex maxima and sums; one atomic per block and statistic.
__global__ void __launch_bounds__(256) k_vertex_meta(std::int64_t n, const std::int32_t* off, const std::uint32_t* vw,
                                                     const std::uint64_t* row_bound, MetaOut* out) {
    __shared__ unsigned long long part[4][4];
    unsigned long long maxrow = 0, maxdeg = 0, totalvw = 0, maxvw = 0;
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x) {
        const unsigned long long d = (unsigned long long)(off[v + 1] - off[v]);
        const unsigned long long w = vw[v];
        maxdeg = maxdeg > d ? maxdeg : d;
        maxvw = maxvw > w ? maxvw : w;
        totalvw += w;
        maxrow = maxrow > row_bound[v] ? maxrow : row_bound[v];
    }
    maxrow = wave_max(maxrow);
    maxdeg = wave_max(maxdeg);
    maxvw = wave_max(maxvw);
    totalvw = wave_sum(totalvw);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        part[w][0] = maxrow;
        part[w][1] = maxdeg;
        part[w][2] = maxvw;
        part[w][3] = totalvw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 4; ++i) {
            maxrow = maxrow > part[i][0] ? maxrow : part[i][0];
            maxdeg = maxdeg > part[i][1] ? maxdeg : part[i][1];
            maxvw = maxvw > part[i][2] ? maxvw : part[i][2];
            totalvw += part[i][3];
        }
        if (maxrow) atomicMax(&out->max_row, maxrow);
        if (maxdeg) atomicMax(&out->max_deg, maxdeg);
        if (maxvw) atomicMax(&out->max_vw, maxvw);
        if (totalvw) atomicAdd(&out->total_vw, totalvw);
    }
}

// Edge-weight flags: bit0 some weight is zero, bit1 some weight is not one.
__global__ void __launch_bounds__(256) k_edge_flags(std::int64_t m, const std::uint32_t* ew, MetaOut* out) {
    bool zero = false, non_unit = false;
    for (std::int64_t x = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; x < m;
         x += (std::int64_t)gridDim.x * blockDim.x) {
        const std::uint32_t w = ew[x];
        zero |= w == 0;
        non_unit |= w != 1;
    }
    const unsigned long long z = __ballot(zero), nu = __ballot(non_unit);
    if (lane_id() == 0 && (z || nu)) atomicOr(&out->flags, (z ? 1ULL : 0ULL) | (nu ? 2ULL : 0ULL));
}

// Degree classes for this level (src/coarsening/aggregation.cuh:5-71).
// Queue order is irrelevant: every later write is per vertex.
struct ClassCounts {
    unsigned long long c[8];  // 0: <=16, 1: <=64, 2: <=128, 3: <=256, 4: <=512, 5: <=4096, 6: >4096, 7: max degree
};

__device__ inline int degree_class(std::int64_t d) {
    return d <= kLowDegree ? 0 : d <= 64 ? 1 : d <= 128 ? 2 : d <= 256 ? 3 : d <= kWaveDegree ? 4 : d <= kBlockDegree ? 5 : 6;
}

// Row-length classification into seven lists.  A block classifies a tile
// of 256 x kClassItems rows: ranks come from wave ballots and an LDS
// counter per class, and each class costs one global atomic per tile.
constexpr int kClassItems = 8;

template <typename Off>
__global__ void __launch_bounds__(256) k_classify(std::int64_t rows, const Off* off, std::int32_t* lists,
                                                  std::int64_t stride, ClassCounts* counts) {
    __shared__ unsigned int lc[8];
    __shared__ unsigned long long gb[8];
    const int lane = lane_id();
    const unsigned long long lt = (1ULL << lane) - 1ULL;
    unsigned long long maxdeg = 0;
    for (std::int64_t base = blockIdx.x * (std::int64_t)(256 * kClassItems); base < rows;
         base += (std::int64_t)gridDim.x * (256 * kClassItems)) {
        if (threadIdx.x < 8) lc[threadIdx.x] = 0;
        __syncthreads();
        int cls[kClassItems];
        unsigned int rank[kClassItems];
#pragma unroll
        for (int it = 0; it < kClassItems; ++it) {
            const std::int64_t v = base + it * 256 + threadIdx.x;
            const std::int64_t d = v < rows ? (std::int64_t)(off[v + 1] - off[v]) : 0;
            maxdeg = maxdeg > (unsigned long long)d ? maxdeg : (unsigned long long)d;
            cls[it] = v < rows ? degree_class(d) : -1;
            rank[it] = 0;
            for (int c = 0; c < 7; ++c) {
                const unsigned long long mask = __ballot(cls[it] == c);
                if (!mask) continue;
                const int leader = __ffsll(mask) - 1;
                unsigned int wb = 0;
                if (lane == leader) wb = atomicAdd(&lc[c], (unsigned int)__popcll(mask));
                wb = __shfl(wb, leader);
                if (cls[it] == c) rank[it] = wb + (unsigned int)__popcll(mask & lt);
            }
        }
        __syncthreads();
        if (threadIdx.x < 7)
            gb[threadIdx.x] = lc[threadIdx.x] ? atomicAdd(&counts->c[threadIdx.x], (unsigned long long)lc[threadIdx.x]) : 0;
        __syncthreads();
#pragma unroll
        for (int it = 0; it < kClassItems; ++it)
            if (cls[it] >= 0)
                lists[cls[it] * stride + gb[cls[it]] + rank[it]] = (std::int32_t)(base + it * 256 + threadIdx.x);
        __syncthreads();
    }
    maxdeg = wave_max(maxdeg);
    if (lane == 0 && maxdeg) atomicMax(&counts->c[7], maxdeg);
}

template <typename Off>
void classify(Context& c, std::int64_t rows, const Off* off, std::int32_t* lists, std::int64_t stride,
              ClassCounts* counts) {
    if (rows <= 0) return;
    const int blocks = static_cast<int>(std::min<std::int64_t>(ceil_div(rows, 256 * kClassItems), 2048));
    hipLaunchKernelGGL((k_classify<Off>), dim3(blocks), dim3(256), 0, c.stream(), rows, off, lists, stride, counts);
    GP_LAUNCH_CHECK();
}

// -------------------------------------------------------- seed matching
// seed_matching.cuh:12-91: mutual heaviest admissible edge; order
// (weight, symmetric hash, id); rows with degree > 1024 are skipped.
__global__ void k_match_choose(std::int64_t n, const std::int32_t* off, const std::int32_t* nbr,
                               const std::uint32_t* ew, const std::uint32_t* vw, std::uint64_t cap, std::uint32_t salt,
                               std::int32_t* choices) {
    const int lane = lane_id();
    for (std::int64_t base = (blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x) - lane; base < n;
         base += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int64_t v = base + lane;
        std::int64_t b = 0, e = 0;
        if (v < n) {
            b = off[v];
            e = off[v + 1];
        }
        const bool heavy = v < n && e - b > 32 && e - b <= 1024;
        if (v < n && !heavy) {
            std::int32_t best = -1;
            std::uint64_t bw = 0;
            std::uint32_t bh = 0;
            const std::uint64_t wv = vw[v];
            if (e - b <= 1024) {
                for (std::int64_t x = b; x < e; ++x) {
                    const std::int32_t u = nbr[x];
                    if (u == v || wv > cap || (std::uint64_t)vw[u] > cap - wv) continue;
                    const std::uint32_t lo = (std::uint32_t)(u < v ? u : v), hi = (std::uint32_t)(u < v ? v : u);
                    const std::uint32_t h = mix32(lo ^ mix32(hi ^ salt));
                    const std::uint64_t w = ew_at(ew, x);
                    if (best < 0 || w > bw || (w == bw && (h > bh || (h == bh && u > best)))) {
                        best = u;
                        bw = w;
                        bh = h;
                    }
                }
            }
            choices[v] = best;
        }
        unsigned long long hv = __ballot(heavy);
        while (hv) {
            const int src = __ffsll(hv) - 1;
            hv &= hv - 1;
            const std::int64_t hvx = __shfl(v, src);
            const std::int64_t hb = off[hvx], he = off[hvx + 1];
            const std::uint64_t wv = vw[hvx];
            std::int32_t best = -1;
            std::uint64_t bw = 0;
            std::uint32_t bh = 0;
            for (std::int64_t x = hb + lane; x < he; x += kWave) {
                const std::int32_t u = nbr[x];
                if (u == hvx || wv > cap || (std::uint64_t)vw[u] > cap - wv) continue;
                const std::uint32_t lo = (std::uint32_t)(u < hvx ? u : hvx), hi = (std::uint32_t)(u < hvx ? hvx : u);
                const std::uint32_t h = mix32(lo ^ mix32(hi ^ salt));
                const std::uint64_t w = ew_at(ew, x);
                if (best < 0 || w > bw || (w == bw && (h > bh || (h == bh && u > best)))) {
                    best = u;
                    bw = w;
                    bh = h;
                }
            }
            for (int d = kWave / 2; d; d >>= 1) {
                const std::int32_t ob = __shfl_xor(best, d);
                const std::uint64_t ow = __shfl_xor(bw, d);
                const std::uint32_t oh = __shfl_xor(bh, d);
                if (ob >= 0 && (best < 0 || ow > bw || (ow == bw && (oh > bh || (oh == bh && ob > best))))) {
                    best = ob;
                    bw = ow;
                    bh = oh;
                }
            }
            if (lane == src) choices[hvx] = best;
        }
    }
}

__global__ void k_match_roots(std::int64_t n, const std::int32_t* choices, const std::uint32_t* vw,
                              std::int32_t* clusters, unsigned long long* cw, unsigned long long* roots) {
    unsigned long long count = 0;
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t u = choices[v];
        const bool mutual = u >= 0 && choices[u] == v;
        const std::int32_t root = mutual ? (u < v ? u : (std::int32_t)v) : (std::int32_t)v;
        clusters[v] = root;
        count += root == v;
        atomicAdd(&cw[root], (unsigned long long)vw[v]);
    }
    count = wave_sum(count);
    if (lane_id() == 0 && count) atomicAdd(roots, count);
}

__global__ void k_init_singletons(std::int64_t n, const std::uint32_t* vw, std::int32_t* clusters,
                                  unsigned long long* cw) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x) {
        clusters[v] = (std::int32_t)v;
        cw[v] = vw[v];
    }
}

// -------------------------------------------------------------- roles
// residual_freeze_roles (singleton_growth.cuh:27-40); also resets proposals.
__global__ void k_roles(std::int64_t n, const std::int32_t* clusters, const std::uint32_t* vw,
                        const unsigned long long* cw, std::uint32_t salt, std::uint32_t threshold, bool all,
                        std::uint8_t* roles, std::int32_t* proposals) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x) {
        roles[v] = clusters[v] == v && cw[v] == (unsigned long long)vw[v] &&
                   (all || mix32((std::uint32_t)v ^ salt) < threshold);
        proposals[v] = kInvalid;
    }
}

struct AffinityArgs {
    const std::int32_t* off;
    const std::int32_t* nbr;
    const std::uint32_t* ew;
    const std::int32_t* clusters;
    const unsigned long long* cw;
    const std::uint8_t* roles;
    std::uint64_t cap;
    std::uint32_t tie_salt;
    bool filter_roles;
    bool use_second;
    std::int32_t* proposals;
    std::uint64_t* gains;
    std::uint64_t* regret;
    std::uint32_t* ties;
};

// Candidate accumulator: best (value, tie) and the runner-up value of the
// multiset (aggregation.cuh:97-106); independent of evaluation order.
struct Top2 {
    std::uint64_t current, best, tie, second;
};

__device__ inline void top2_add(const AffinityArgs& a, std::int32_t v, std::int32_t src, std::int32_t c,
                                std::uint64_t val, Top2& t) {
    if (c == src) {
        t.current = val;
        return;
    }
    if (a.filter_roles) {
        const std::uint64_t base = a.cw[c];
        if (a.roles[c] || !(base != 0 && base <= a.cap && a.cw[src] <= a.cap - base)) return;
    }
    const std::uint32_t h = mix32((std::uint32_t)v ^ mix32((std::uint32_t)c) ^ a.tie_salt);
    const std::uint64_t key = ((std::uint64_t)h << 32) | (std::uint64_t)(0xffffffffU - (std::uint32_t)c);
    if (a.use_second) {
        if (sclp_better(val, key, t.best, t.tie)) {
            t.second = t.second > t.best ? t.second : t.best;
            t.best = val;
            t.tie = key;
        } else
            t.second = t.second > val ? t.second : val;
    } else if (sclp_better(val, key, t.best, t.tie)) {
        t.best = val;
        t.tie = key;
    }
}

__device__ inline Top2 top2_merge(const Top2& x, const Top2& y) {
    Top2 r;
    r.current = x.current > y.current ? x.current : y.current;
    if (sclp_better(y.best, y.tie, x.best, x.tie)) {
        r.best = y.best;
        r.tie = y.tie;
        std::uint64_t s = x.second > y.second ? x.second : y.second;
        r.second = s > x.best ? s : x.best;
    } else {
        r.best = x.best;
        r.tie = x.tie;
        std::uint64_t s = x.second > y.second ? x.second : y.second;
        r.second = s > y.best ? s : y.best;
    }
    return r;
}

template <int W>
__device__ inline Top2 top2_group_reduce(Top2 t) {
    for (int d = W / 2; d; d >>= 1) {
        Top2 o;
        o.current = __shfl_xor(t.current, d, W);
        o.best = __shfl_xor(t.best, d, W);
        o.tie = __shfl_xor(t.tie, d, W);
        o.second = __shfl_xor(t.second, d, W);
        t = top2_merge(t, o);
    }
    return t;
}

// sclp_decode_gain_kernel (aggregation.cuh:513-557) for one mover.
__device__ inline void decode_proposal(const AffinityArgs& a, std::int32_t v, std::int32_t src, const Top2& t) {
    if (!(t.best > t.current)) return;
    const std::int64_t target = (std::int64_t)(0xffffffffULL - (t.tie & 0xffffffffULL));
    if (target < 0 || target > 0x7fffffffLL || target == src) return;
    a.proposals[v] = (std::int32_t)target;
    a.gains[v] = t.best - t.current;
    if (a.use_second) a.regret[v] = t.best - t.second;
    a.ties[v] = (std::uint32_t)(t.tie >> 32);
}

// Group kernel: W lanes per mover, rows of at most W*I entries.
template <int W, int I>
__global__ void k_affinity_group(AffinityArgs a, const std::int32_t* list, std::int64_t count) {
    const int lg = threadIdx.x % W;
    const int group_base = (threadIdx.x % kWave) / W * W;
    const std::int64_t groups = (std::int64_t)gridDim.x * (blockDim.x / W);
    for (std::int64_t idx = blockIdx.x * (std::int64_t)(blockDim.x / W) + threadIdx.x / W; idx < count;
         idx += groups) {
        const std::int32_t v = list[idx];
        const std::int32_t src = a.clusters[v];
        if (!a.roles[src]) continue;
        const std::int64_t b = a.off[v], d = a.off[v + 1] - b;
        std::int32_t k[I];
        std::uint32_t w[I];
#pragma unroll
        for (int it = 0; it < I; ++it) {
            const std::int64_t x = it * W + lg;
            if (x < d) {
                const std::int32_t u = a.nbr[b + x];
                k[it] = u == v ? kSentinel : a.clusters[u];
                w[it] = (std::uint32_t)ew_at(a.ew, b + x);
            } else {
                k[it] = kSentinel;
                w[it] = 0;
            }
        }
        group_bitonic_sort<W, I>(k, w, lg);
        Top2 t{0, 0, 0, 0};
        group_reduce_runs<W, I>(k, w, lg, group_base, [&](std::int32_t c, std::uint64_t s, int) {
            top2_add(a, v, src, c, s, t);
        });
        t = top2_group_reduce<W>(t);
        if (lg == 0) decode_proposal(a, v, src, t);
    }
}

// Block kernel: one 256-thread block per mover with <= 4096 entries.
__global__ void __launch_bounds__(kBlockSortThreads) k_affinity_block(AffinityArgs a, const std::int32_t* list,
                                                                      std::int64_t count) {
    __shared__ BlockSortSmem s;
    __shared__ Top2 part[kBlockSortThreads / kWave];
    for (std::int64_t idx = blockIdx.x; idx < count; idx += gridDim.x) {
        const std::int32_t v = list[idx];
        const std::int32_t src = a.clusters[v];
        if (!a.roles[src]) continue;
        const std::int64_t b = a.off[v];
        const int d = (int)(a.off[v + 1] - b);
        for (int x = threadIdx.x; x < d; x += kBlockSortThreads) {
            const std::int32_t u = a.nbr[b + x];
            s.key[x] = u == v ? kSentinel : a.clusters[u];
            s.val[x] = (std::uint32_t)ew_at(a.ew, b + x);
        }
        __syncthreads();
        Top2 t{0, 0, 0, 0};
        block_sort_reduce(s, d, [&](std::int32_t c, std::uint64_t sum, int) { top2_add(a, v, src, c, sum, t); });
        t = top2_group_reduce<kWave>(t);
        if (lane_id() == 0) part[threadIdx.x / kWave] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            Top2 r = part[0];
            for (int i = 1; i < kBlockSortThreads / kWave; ++i) r = top2_merge(r, part[i]);
            decode_proposal(a, v, src, r);
        }
        __syncthreads();
    }
}

// Hash variant of the block kernel (row weights below 2^32): per-cluster
// sums in an LDS hash table, then the order-independent top-2 over slots.
// Rows of at most MaxRow entries, one block each, summed per cluster in an
// LDS hash of H::slots slots (exact while every row weighs less than 2^32).
template <class H, int MaxRow, int Threads>
__global__ void __launch_bounds__(Threads) k_affinity_block_hash(AffinityArgs a, const std::int32_t* list,
                                                                 std::int64_t count) {
    static_assert(4 * MaxRow <= 3 * H::slots, "a row fills at most three quarters of the table");
    __shared__ H s;
    __shared__ Top2 part[Threads / kWave];
    for (std::int64_t idx = blockIdx.x; idx < count; idx += gridDim.x) {
        const std::int32_t v = list[idx];
        const std::int32_t src = a.clusters[v];
        if (!a.roles[src]) continue;
        const std::int64_t b = a.off[v];
        const int d = (int)(a.off[v + 1] - b);
        const int slots = hash_slots_for<H>(d);
        block_hash_clear(s, slots);
        __syncthreads();
        // Sixteen (eight in 1024-thread blocks, within their register
        // budget) gathers per thread are issued before their inserts.
        constexpr int kPer = Threads >= 1024 ? 8 : 16;
        for (int x0 = 0; x0 < d; x0 += kPer * Threads) {
            std::int32_t kk[kPer];
            std::uint32_t vv[kPer];
#pragma unroll
            for (int r = 0; r < kPer; ++r) {
                const int x = x0 + r * Threads + threadIdx.x;
                const std::int32_t u = x < d ? a.nbr[b + x] : v;
                kk[r] = u != v ? a.clusters[u] : -1;
                vv[r] = x < d ? (std::uint32_t)ew_at(a.ew, b + x) : 0u;
            }
#pragma unroll
            for (int r = 0; r < kPer; ++r)
                if (kk[r] >= 0) block_hash_add(s, slots, kk[r], vv[r]);
        }
        __syncthreads();
        Top2 t{0, 0, 0, 0};
        for (int i = threadIdx.x; i < slots; i += Threads)
            if (s.key[i] >= 0) top2_add(a, v, src, s.key[i], s.val[i], t);
        t = top2_group_reduce<kWave>(t);
        if (lane_id() == 0) part[threadIdx.x / kWave] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            Top2 r = part[0];
            for (int i = 1; i < Threads / kWave; ++i) r = top2_merge(r, part[i]);
            decode_proposal(a, v, src, r);
        }
        __syncthreads();
    }
}

// Hub rows of at most kBigRowDegree entries take one block each with a
// larger table (one block fills a CU's LDS, so it brings 16 waves);
// longer rows go through records and the global sort.
constexpr int kBigRowDegree = 12288;
constexpr int kBlockHashThreads = 512;
using BigHashSmem = BlockHashT<16384>;

// Splits the hub list by row length (order inside each part is arbitrary:
// each row's proposal is its own).
__global__ void k_split_hubs(const std::int32_t* off, const std::int32_t* list, std::int64_t count, std::int32_t* big,
                             std::int32_t* rest, unsigned long long* counts) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t v = list[i];
        if (off[v + 1] - off[v] <= kBigRowDegree)
            big[atomicAdd(&counts[0], 1ULL)] = v;
        else
            rest[atomicAdd(&counts[1], 1ULL)] = v;
    }
}

// Hub rows (> 4096 entries): records (row, cluster, w), radix sort, reduce.
__global__ void k_hub_records(AffinityArgs a, const std::int32_t* list, std::int64_t count,
                              const std::int64_t* rec_off, int key_bits, std::uint64_t* keys, std::uint32_t* vals) {
    for (std::int64_t r = blockIdx.x; r < count; r += gridDim.x) {
        const std::int32_t v = list[r];
        const std::int64_t b = a.off[v], d = a.off[v + 1] - b;
        const std::int64_t o = rec_off[r];
        const bool mover = a.roles[a.clusters[v]] != 0;
        for (std::int64_t x = threadIdx.x; x < d; x += blockDim.x) {
            const std::int32_t u = a.nbr[b + x];
            const bool skip = !mover || u == v;
            keys[o + x] = skip ? ~0ULL : (((std::uint64_t)r << key_bits) | (std::uint64_t)a.clusters[u]);
            vals[o + x] = skip ? 0u : (std::uint32_t)ew_at(a.ew, b + x);
        }
    }
}

// Record pre-reduction for long rows: a block sort-reduces one chunk of
// up to kBlockSortMax entries of one listed row in LDS and appends its
// unique (list index, key) records with their sums.  Chunk order in the
// output is arbitrary; the global sort that follows makes it canonical.
// chunk_off[i] is the first chunk of list row i (count + 1 entries).
// Partial sums above 2^32 - 1 raise *overflow (the caller then reduces the
// raw records instead).
template <class Src>
__global__ void __launch_bounds__(kBlockSortThreads) k_chunk_records(Src src, const std::int64_t* chunk_off,
                                                                     std::int64_t count, int key_bits,
                                                                     std::uint64_t* keys, std::uint32_t* vals,
                                                                     unsigned long long* out_count,
                                                                     unsigned long long* overflow) {
    __shared__ BlockSortSmem s;
    __shared__ unsigned long long obase;
    __shared__ std::int64_t row_of;
    const std::int64_t chunks = chunk_off[count];
    for (std::int64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
        if (threadIdx.x == 0) {
            std::int64_t lo = 0, hi = count;  // last i with chunk_off[i] <= ch
            while (hi - lo > 1) {
                const std::int64_t mid = lo + (hi - lo) / 2;
                if (chunk_off[mid] <= ch)
                    lo = mid;
                else
                    hi = mid;
            }
            row_of = lo;
        }
        __syncthreads();
        const std::int64_t li = row_of;
        const std::int64_t b = src.begin(li), e = src.end(li);
        const std::int64_t x0 = b + (ch - chunk_off[li]) * kBlockSortMax;
        const int d = (int)((e - x0) < kBlockSortMax ? (e - x0) : kBlockSortMax);
        for (int x = threadIdx.x; x < d; x += kBlockSortThreads) {
            s.key[x] = src.key(li, x0 + x);
            s.val[x] = src.val(li, x0 + x);
        }
        __syncthreads();
        block_sort_reduce(
            s, d,
            [&](std::int32_t key, std::uint64_t sum, int i) {
                if (sum > 0xffffffffULL) atomicOr(overflow, 1ULL);
                keys[obase + i] = ((std::uint64_t)li << key_bits) | (std::uint64_t)key;
                vals[obase + i] = (std::uint32_t)sum;
            },
            [&](int total) {
                if (threadIdx.x == 0) obase = total ? atomicAdd(out_count, (unsigned long long)total) : 0ULL;
                __syncthreads();
            });
        __syncthreads();
    }
}

// Hash variant of k_chunk_records for row weights below 2^32: a chunk's
// unique keys are gathered from an LDS hash table (output order is
// arbitrary; the global sort makes it canonical).
template <class Src>
__global__ void __launch_bounds__(kBlockSortThreads) k_chunk_records_hash(Src src, const std::int64_t* chunk_off,
                                                                          std::int64_t count, int key_bits,
                                                                          std::uint64_t* keys, std::uint32_t* vals,
                                                                          unsigned long long* out_count) {
    __shared__ BlockHashSmem s;
    __shared__ int wsum[kBlockSortThreads / kWave];
    __shared__ unsigned long long obase;
    __shared__ std::int64_t row_of;
    const std::int64_t chunks = chunk_off[count];
    for (std::int64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
        if (threadIdx.x == 0) {
            std::int64_t lo = 0, hi = count;  // last i with chunk_off[i] <= ch
            while (hi - lo > 1) {
                const std::int64_t mid = lo + (hi - lo) / 2;
                if (chunk_off[mid] <= ch)
                    lo = mid;
                else
                    hi = mid;
            }
            row_of = lo;
        }
        __syncthreads();
        const std::int64_t li = row_of;
        const std::int64_t b = src.begin(li), e = src.end(li);
        const std::int64_t x0 = b + (ch - chunk_off[li]) * kBlockSortMax;
        const int d = (int)((e - x0) < kBlockSortMax ? (e - x0) : kBlockSortMax);
        const int slots = hash_slots_for(d);
        block_hash_clear(s, slots);
        __syncthreads();
        // The chunk's loads are all issued before the inserts.
        constexpr int kPer = kBlockSortMax / kBlockSortThreads;
        std::int32_t kk[kPer];
        std::uint32_t vv[kPer];
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const int x = r * kBlockSortThreads + threadIdx.x;
            kk[r] = kSentinel;
            vv[r] = 0;
            if (x < d) {
                kk[r] = src.key(li, x0 + x);
                vv[r] = src.val(li, x0 + x);
            }
        }
#pragma unroll
        for (int r = 0; r < kPer; ++r)
            if (kk[r] != kSentinel) block_hash_add(s, slots, kk[r], vv[r]);
        __syncthreads();
        // Each wave compacts a contiguous quarter of the slots with ballots,
        // so consecutive lanes write consecutive records.
        const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
        const int q = slots / (kBlockSortThreads / kWave);
        int cnt = 0;
        for (int i0 = w * q; i0 < (w + 1) * q; i0 += kWave) cnt += __popcll(__ballot(s.key[i0 + lane] >= 0));
        int inc, before, total;
        block_scan<PlusOp>(lane == 0 ? cnt : 0, wsum, inc, before, total);
        if (threadIdx.x == 0) obase = total ? atomicAdd(out_count, (unsigned long long)total) : 0ULL;
        __syncthreads();
        unsigned long long o = obase + (unsigned long long)__shfl(before, 0);
        const unsigned long long lt = (1ULL << lane) - 1ULL;
        for (int i0 = w * q; i0 < (w + 1) * q; i0 += kWave) {
            const std::int32_t key = s.key[i0 + lane];
            const unsigned long long mask = __ballot(key >= 0);
            if (key >= 0) {
                const unsigned long long j = o + __popcll(mask & lt);
                keys[j] = ((std::uint64_t)li << key_bits) | (std::uint64_t)key;
                vals[j] = s.val[i0 + lane];
            }
            o += __popcll(mask);
        }
        __syncthreads();
    }
}

// Chunks per listed row (zero for rows the source skips).
template <class Src>
__global__ void k_chunk_counts(Src src, std::int64_t count, std::int64_t* out) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x)
        out[i] = src.active(i) ? (src.end(i) - src.begin(i) + kBlockSortMax - 1) / kBlockSortMax : 0;
}

// Hub affinity records: neighbour clusters of mover hubs.
struct HubSrc {
    AffinityArgs a;
    const std::int32_t* list;
    __device__ bool active(std::int64_t i) const { return a.roles[a.clusters[list[i]]] != 0; }
    __device__ std::int64_t begin(std::int64_t i) const { return a.off[list[i]]; }
    __device__ std::int64_t end(std::int64_t i) const { return a.off[list[i] + 1]; }
    __device__ std::int32_t key(std::int64_t i, std::int64_t x) const {
        const std::int32_t u = a.nbr[x];
        return u == list[i] ? kSentinel : a.clusters[u];
    }
    __device__ std::uint32_t val(std::int64_t, std::int64_t x) const { return (std::uint32_t)ew_at(a.ew, x); }
};

// Unique (row, key) runs of sorted records: heads, scan, atomic sums.
// Run heads of sorted records: a record whose key differs from its
// predecessor's, padding keys (all ones) excluded.
__device__ inline bool run_head(const std::uint64_t* keys, std::int64_t i) {
    const std::uint64_t k = keys[i];
    return k != ~0ULL && (i == 0 || k != keys[i - 1]);
}

// Heads per tile of kScanTile records (one block per tile).
__global__ void __launch_bounds__(kScanThreads) k_run_tile_heads(const std::uint64_t* keys, std::int64_t n,
                                                                std::int64_t* tile_heads) {
    __shared__ std::int64_t wsum[kScanThreads / kWave];
    const std::int64_t tile = blockIdx.x * (std::int64_t)kScanTile;
    std::int64_t c = 0;
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        const std::int64_t i = tile + r * kScanThreads + threadIdx.x;
        c += i < n && run_head(keys, i);
    }
    c = wave_sum(c);
    if (lane_id() == 0) wsum[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        std::int64_t t = 0;
        for (int w = 0; w < kScanThreads / kWave; ++w) t += wsum[w];
        tile_heads[blockIdx.x] = t;
    }
}

// Sums of runs of equal keys in sorted records.  A block reduces one tile;
// runs inside the tile are stored, runs crossing a tile edge add their
// partial sums atomically into the zero-filled usums.  A run's index is the
// tile's head offset (exclusive scan of k_run_tile_heads) plus the heads
// before it inside the tile.
__global__ void __launch_bounds__(kScanThreads) k_run_reduce(const std::uint64_t* keys, const std::uint32_t* vals,
                                                            std::int64_t n, const std::int64_t* tile_base,
                                                            std::uint64_t* ukeys, unsigned long long* usums) {
    __shared__ SegU64 wsum[kScanThreads / kWave];
    __shared__ unsigned hsum[kScanThreads / kWave];
    const int lane = lane_id(), w = threadIdx.x / kWave;
    const std::int64_t tile = blockIdx.x * (std::int64_t)kScanTile;
    const std::int64_t end = tile + kScanTile < n ? tile + kScanTile : n;
    const std::int64_t base = tile + w * (kScanItems * kWave);
    SegU64 inc[kScanItems];
    unsigned hinc[kScanItems];
    SegU64 acc{0, 0};
    unsigned hacc = 0;
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        const std::int64_t idx = base + r * kWave + lane;
        SegU64 x{0, 0};
        unsigned h = 0;
        if (idx < end) {
            x = SegU64{vals[idx], (idx == tile || keys[idx] != keys[idx - 1]) ? 1u : 0u};
            h = run_head(keys, idx) ? 1u : 0u;
        }
        inc[r] = wave_inclusive_scan<SegPlusOp>(x);
        hinc[r] = wave_inclusive_scan<PlusOp>(h);
        acc = SegPlusOp::op<SegU64>(acc, shfl_val(inc[r], kWave - 1));
        hacc += __shfl(hinc[r], kWave - 1);
    }
    if (lane == 0) {
        wsum[w] = acc;
        hsum[w] = hacc;
    }
    __syncthreads();
    SegU64 carry{0, 0};
    std::int64_t hcarry = tile_base[blockIdx.x];
    for (int i = 0; i < w; ++i) {
        carry = SegPlusOp::op<SegU64>(carry, wsum[i]);
        hcarry += hsum[i];
    }
    const std::uint64_t left = tile > 0 ? keys[tile - 1] : ~0ULL;
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        const std::int64_t idx = base + r * kWave + lane;
        const SegU64 full = SegPlusOp::op<SegU64>(carry, inc[r]);
        const std::int64_t heads = hcarry + hinc[r];
        carry = SegPlusOp::op<SegU64>(carry, shfl_val(inc[r], kWave - 1));
        hcarry += __shfl(hinc[r], kWave - 1);
        if (idx >= end) continue;
        const std::uint64_t key = keys[idx];
        if (key == ~0ULL) continue;
        const bool tail = idx + 1 == n || keys[idx + 1] != key;
        if (!tail && idx + 1 != end) continue;
        const std::int64_t u = heads - 1;  // heads up to idx -> run id
        if (tail) ukeys[u] = key;
        if (tail && key != left)
            usums[u] = full.v;
        else
            atomicAdd(&usums[u], (unsigned long long)full.v);
    }
}

// Record range [first, last) of list row li in keys sorted by (li, key).
// Record range [ranges[2r], ranges[2r+1]) of every row r of sorted unique
// records keyed (row << key_bits | key): one pass marks where rows start
// and end (rows without records keep the zero-filled empty range).
__global__ void k_row_ranges(const std::uint64_t* ukeys, std::int64_t nu, int key_bits, std::int64_t* ranges) {
    for (std::int64_t j = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; j < nu;
         j += (std::int64_t)gridDim.x * blockDim.x) {
        const std::uint64_t row = ukeys[j] >> key_bits;
        if (j == 0 || (ukeys[j - 1] >> key_bits) != row) ranges[2 * row] = j;
        if (j + 1 == nu || (ukeys[j + 1] >> key_bits) != row) ranges[2 * row + 1] = j + 1;
    }
}

struct HubTop2 {
    unsigned long long current, best, tie, second;
};

// Per hub row (one block): the weight toward its own cluster, the best and
// second-best admissible weights (order statistics, so the result does not
// depend on visiting order) and the hashed tie key among the best.
__global__ void __launch_bounds__(256) k_hub_reduce(AffinityArgs a, const std::int32_t* list, std::int64_t count,
                                                   const std::uint64_t* ukeys, const unsigned long long* usums,
                                                   const std::int64_t* ranges, int key_bits, HubTop2* acc) {
    __shared__ unsigned long long part[4][3];
    const std::uint64_t mask = (1ULL << key_bits) - 1ULL;
    const int w = threadIdx.x / kWave;
    for (std::int64_t r = blockIdx.x; r < count; r += gridDim.x) {
        __syncthreads();
        const std::int64_t first = ranges[2 * r], last = ranges[2 * r + 1];
        const std::int32_t v = list[r];
        const std::int32_t src = a.clusters[v];
        auto admissible = [&](std::int32_t c) {
            if (!a.filter_roles) return true;
            const std::uint64_t base = a.cw[c];
            return !a.roles[c] && base != 0 && base <= a.cap && a.cw[src] <= a.cap - base;
        };
        unsigned long long cur = 0, b1 = 0, b2 = 0;
        // Four records per thread in flight: the records, then their
        // clusters' weights and roles, then the order statistics.
        const std::uint64_t srcw = a.cw[src];
        constexpr int kHubBatch = 4;
        for (std::int64_t j0 = first + threadIdx.x; j0 < last; j0 += kHubBatch * blockDim.x) {
            std::int32_t c[kHubBatch];
            unsigned long long val[kHubBatch], cwt[kHubBatch];
            std::uint8_t role[kHubBatch];
#pragma unroll
            for (int q = 0; q < kHubBatch; ++q) {
                const std::int64_t j = j0 + q * (std::int64_t)blockDim.x;
                c[q] = j < last ? (std::int32_t)(ukeys[j] & mask) : -1;
                val[q] = j < last ? usums[j] : 0ULL;
            }
#pragma unroll
            for (int q = 0; q < kHubBatch; ++q) {
                const bool look = c[q] >= 0 && c[q] != src && a.filter_roles;
                cwt[q] = look ? a.cw[c[q]] : 0ULL;
                role[q] = look ? a.roles[c[q]] : 0;
            }
#pragma unroll
            for (int q = 0; q < kHubBatch; ++q) {
                if (c[q] < 0) continue;
                if (c[q] == src) {
                    cur = val[q];
                    continue;
                }
                if (a.filter_roles && !(!role[q] && cwt[q] != 0 && cwt[q] <= a.cap && srcw <= a.cap - cwt[q]))
                    continue;
                if (val[q] > b1) {
                    b2 = b1;
                    b1 = val[q];
                } else if (val[q] > b2)
                    b2 = val[q];
            }
        }
        for (int d = kWave / 2; d; d >>= 1) {
            const unsigned long long o1 = __shfl_xor(b1, d), o2 = __shfl_xor(b2, d), oc = __shfl_xor(cur, d);
            const unsigned long long lo = b1 < o1 ? b1 : o1, hi2 = b2 > o2 ? b2 : o2;
            b1 = b1 > o1 ? b1 : o1;
            b2 = lo > hi2 ? lo : hi2;
            cur = cur > oc ? cur : oc;
        }
        if (lane_id() == 0) {
            part[w][0] = b1;
            part[w][1] = b2;
            part[w][2] = cur;
        }
        __syncthreads();
        b1 = part[0][0];
        b2 = part[0][1];
        cur = part[0][2];
        for (int i = 1; i < 4; ++i) {
            const unsigned long long o1 = part[i][0], o2 = part[i][1];
            const unsigned long long lo = b1 < o1 ? b1 : o1, hi2 = b2 > o2 ? b2 : o2;
            b1 = b1 > o1 ? b1 : o1;
            b2 = lo > hi2 ? lo : hi2;
            cur = cur > part[i][2] ? cur : part[i][2];
        }
        unsigned long long tie = 0;
        if (b1)
            for (std::int64_t j = first + threadIdx.x; j < last; j += blockDim.x) {
                const std::int32_t c = (std::int32_t)(ukeys[j] & mask);
                if (c == src || usums[j] != b1 || !admissible(c)) continue;
                const std::uint32_t h = mix32((std::uint32_t)v ^ mix32((std::uint32_t)c) ^ a.tie_salt);
                const unsigned long long key =
                    ((unsigned long long)h << 32) | (unsigned long long)(0xffffffffU - (std::uint32_t)c);
                tie = tie > key ? tie : key;
            }
        tie = wave_max(tie);
        __syncthreads();
        if (lane_id() == 0) part[w][0] = tie;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = 1; i < 4; ++i) tie = tie > part[i][0] ? tie : part[i][0];
            acc[r] = HubTop2{cur, b1, tie, a.use_second ? b2 : 0ULL};
        }
        __syncthreads();
    }
}
__global__ void k_hub_decode(AffinityArgs a, const std::int32_t* list, std::int64_t count, const HubTop2* acc) {
    for (std::int64_t r = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; r < count;
         r += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t v = list[r];
        const std::int32_t src = a.clusters[v];
        if (!a.roles[src]) continue;
        // A zero best never wins a comparison; keep that rule for the tie too.
        Top2 t{acc[r].current, acc[r].best, acc[r].best ? acc[r].tie : 0, acc[r].second};
        decode_proposal(a, v, src, t);
    }
}

// ------------------------------------------------------------ admission
__global__ void k_commit_all(std::int64_t n, const std::int32_t* proposals, const std::uint32_t* vw,
                             std::int32_t* clusters, unsigned long long* cw) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t t = proposals[v];
        if (t == kInvalid) continue;
        clusters[v] = t;
        cw[v] = 0;
        atomicAdd(&cw[t], (unsigned long long)vw[v]);
    }
}

__global__ void k_valid_flags(std::int64_t n, const std::int32_t* targets, std::uint8_t* flags) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x)
        flags[v] = targets[v] != kInvalid;
}

// Admission needs the ordered weight prefix only where the requests of a
// target can exceed its room: with base = frozen weight of the target and
// total = weight of all its requests, base <= cap and total <= cap - base
// admits every request (whatever the order), base > cap admits none.
// flags: 1 contended (ordered admission), 2 admitted without ordering.
__global__ void k_target_totals(std::int64_t n, const std::int32_t* targets, const std::uint32_t* vw,
                                unsigned long long* tot) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x)
        if (targets[v] != kInvalid) atomicAdd(&tot[targets[v]], (unsigned long long)vw[v]);
}
__global__ void k_contended_flags(std::int64_t n, const std::int32_t* targets, const unsigned long long* cw,
                                  const unsigned long long* tot, std::uint64_t cap, std::uint8_t* flags) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t t = targets[v];
        std::uint8_t f = 0;
        if (t != kInvalid) {
            const std::uint64_t base = cw[t];
            if (base <= cap) f = tot[t] > cap - base ? 1 : 2;
        }
        flags[v] = f;
    }
}
__global__ void k_commit_uncontended(std::int64_t n, const std::int32_t* targets, const std::uint8_t* flags,
                                     const std::uint32_t* vw, std::int32_t* clusters, unsigned long long* cw) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x) {
        if (flags[v] != 2) continue;
        const std::int32_t s = clusters[v], t = targets[v];
        const unsigned long long w = vw[v];
        atomicAdd(&cw[s], 0ULL - w);
        atomicAdd(&cw[t], w);
        clusters[v] = t;
    }
}
__global__ void k_flag_is(std::int64_t n, const std::uint8_t* in, std::uint8_t value, std::uint8_t* out) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x)
        out[v] = in[v] == value;
}

// Composite key, least significant first: (~tie) | (~score masked) << 32 | target << (32+sb).
__global__ void k_admission_keys(std::int64_t count, const std::int32_t* order, const std::int32_t* targets,
                                 const std::uint64_t* score, const std::uint32_t* ties, int score_bits, bool with_target,
                                 std::uint64_t* keys, std::uint32_t* vals) {
    const std::uint64_t mask = score_bits >= 32 ? 0xffffffffULL : ((1ULL << score_bits) - 1);
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t v = order[i];
        std::uint64_t key = (std::uint64_t)(~ties[v]);
        key |= ((~score[v]) & mask) << 32;
        if (with_target) key |= (std::uint64_t)targets[v] << (32 + score_bits);
        keys[i] = key;
        vals[i] = (std::uint32_t)v;
    }
}
// Second LSD stage keys when the composite does not fit 64 bits.
__global__ void k_key_field(std::int64_t count, const std::uint32_t* vals, const std::int32_t* targets,
                            const std::uint64_t* score, int mode, int shift, std::uint64_t* keys) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t v = (std::int32_t)vals[i];
        keys[i] = mode == 0 ? (std::uint64_t)targets[v] : ((~score[v]) >> shift);
    }
}

// Segment heads by target in sorted order, per-item weight.
__global__ void k_seg_items(std::int64_t count, const std::uint32_t* vals, const std::int32_t* targets,
                            const std::uint32_t* vw, bool ones, SegU64* items) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t v = (std::int32_t)vals[i];
        const bool head = i == 0 || targets[(std::int32_t)vals[i - 1]] != targets[v];
        items[i] = SegU64{ones ? 1ULL : (std::uint64_t)vw[v], head ? 1u : 0u};
    }
}

__global__ void k_gather_base(std::int64_t count, const std::uint32_t* vals, const std::int32_t* targets,
                              const unsigned long long* cw, std::uint64_t* base) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x)
        base[i] = cw[targets[(std::int32_t)vals[i]]];
}

// sclp_commit_kernel (aggregation.cuh:624-652)
__global__ void k_admission_commit(std::int64_t count, const std::uint32_t* vals, const std::int32_t* targets,
                                   const std::uint32_t* vw, const SegU64* prefix, const std::uint64_t* base,
                                   std::uint64_t cap, std::int32_t* clusters, unsigned long long* cw) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t v = (std::int32_t)vals[i];
        const std::int32_t t = targets[v];
        const std::uint64_t w = vw[v];
        if (base[i] <= cap && prefix[i].v <= cap - base[i]) {
            const std::int32_t s = clusters[v];
            atomicAdd(&cw[s], 0ULL - (unsigned long long)w);
            atomicAdd(&cw[t], (unsigned long long)w);
            clusters[v] = t;
        }
    }
}

__global__ void k_count_nonzero(std::int64_t n, const unsigned long long* cw, unsigned long long* out) {
    unsigned long long c = 0;
    const std::int64_t stride = (std::int64_t)gridDim.x * blockDim.x;
    // Four independent loads per step.
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n; v += 4 * stride) {
        unsigned long long x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) x[j] = v + j * stride < n ? cw[v + j * stride] : 0ULL;
#pragma unroll
        for (int j = 0; j < 4; ++j) c += x[j] != 0;
    }
    c = wave_sum(c);
    if (lane_id() == 0 && c) atomicAdd(out, c);
}

// ------------------------------------------------------------ isolated
__global__ void k_isolated_flags(std::int64_t n, const std::int32_t* off, std::uint8_t* flags) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x)
        flags[v] = off[v] == off[v + 1];
}
__global__ void k_gather_vw(std::int64_t count, const std::int32_t* ids, const std::uint32_t* vw, std::uint64_t* out) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x)
        out[i] = vw[ids[i]];
}
// isolated_representative_kernel / isolated_commit_kernel (aggregation.cuh:713-745)
__global__ void k_isolated_rep(std::int64_t count, const std::int32_t* iso, const std::uint64_t* prefix,
                               const std::uint32_t* vw, std::uint64_t cap, std::int32_t* rep) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t v = iso[i];
        const std::uint64_t fb = (prefix[i] - vw[v]) / cap, lb = (prefix[i] - 1) / cap;
        if (fb == lb) atomicMin(&rep[fb], v);
    }
}
__global__ void k_isolated_commit(std::int64_t count, const std::int32_t* iso, const std::uint64_t* prefix,
                                  const std::uint32_t* vw, std::uint64_t cap, const std::int32_t* rep,
                                  std::int32_t* clusters, unsigned long long* cw) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t v = iso[i];
        const std::uint64_t fb = (prefix[i] - vw[v]) / cap, lb = (prefix[i] - 1) / cap;
        if (fb != lb) continue;
        const std::int32_t t = rep[fb];
        if (t == v) continue;
        clusters[v] = t;
        cw[v] = 0;
        atomicAdd(&cw[t], (unsigned long long)vw[v]);
    }
}

// ------------------------------------------------------------- two-hop
__global__ void k_favorites(std::int64_t n, const std::uint32_t* vw, const unsigned long long* cw,
                            const std::int32_t* clusters, std::int32_t* proposals) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x)
        if (cw[clusters[v]] != (unsigned long long)vw[v]) proposals[v] = kInvalid;
}
// sclp_pair_flags_kernel (aggregation.cuh:666-676); ranks are 1-based.
__global__ void k_pair_flags(std::int64_t count, const std::uint32_t* vals, const std::uint32_t* vw,
                             const SegU64* ranks, std::uint64_t cap, std::uint8_t* flags) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        const bool second = (ranks[i].v & 1ULL) == 0ULL;
        flags[i] = second && i > 0 && (std::uint64_t)vw[vals[i - 1]] <= cap - (std::uint64_t)vw[vals[i]];
    }
}
__global__ void k_pair_commit(std::int64_t count, const std::uint32_t* vals, const std::uint8_t* flags,
                              const std::int64_t* pos, std::int64_t budget, const std::uint32_t* vw,
                              std::int32_t* clusters, unsigned long long* cw) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        if (!flags[i] || pos[i] >= budget) continue;
        const std::int32_t v = (std::int32_t)vals[i];
        const std::int32_t s = clusters[v], t = clusters[vals[i - 1]];
        if (s == t) continue;
        const unsigned long long w = vw[v];
        atomicAdd(&cw[s], 0ULL - w);
        atomicAdd(&cw[t], w);
        clusters[v] = t;
    }
}

// ----------------------------------------------------------- compaction
__global__ void k_nonempty(std::int64_t n, const unsigned long long* cw, std::uint8_t* flags) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x)
        flags[v] = cw[v] > 0;
}
__global__ void k_compact_map(std::int64_t n, const std::int32_t* clusters, const std::int64_t* ids,
                              const std::uint32_t* vw, std::int32_t* map, unsigned long long* sums) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t c = (std::int32_t)ids[clusters[v]];
        map[v] = c;
        atomicAdd(&sums[c], (unsigned long long)vw[v]);
    }
}
__global__ void k_narrow_weights(std::int64_t n, const unsigned long long* sums, std::uint32_t* out,
                                 unsigned long long* maxw) {
    unsigned long long m = 0;
    for (std::int64_t c = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; c < n;
         c += (std::int64_t)gridDim.x * blockDim.x) {
        out[c] = (std::uint32_t)sums[c];
        m = m > sums[c] ? m : sums[c];
    }
    m = wave_max(m);
    if (lane_id() == 0 && m) atomicMax(maxw, m);
}

// ----------------------------------------------------------- contraction
// contract_record_bounds (src/coarsen.cu:101-119) plus raw coarse-row sizes.
// plain: the host has shown that no sum can saturate (n * max bound fits
// 64 bits), so the saturating compare-and-swap loop reduces to an add.
__global__ void k_contract_bounds(std::int64_t n, const std::int32_t* map, const std::int32_t* off,
                                  const std::uint64_t* fine_bounds, unsigned long long* coarse_bounds,
                                  unsigned long long* raw_counts, bool plain) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t c = map[v];
        atomicAdd(&raw_counts[c], (unsigned long long)(off[v + 1] - off[v]));
        unsigned long long* target = &coarse_bounds[c];
        const unsigned long long amount = fine_bounds[v];
        if (plain) {
            atomicAdd(target, amount);
            continue;
        }
        unsigned long long prev = *target;
        while (true) {
            const unsigned long long sum = ~0ULL - prev < amount ? ~0ULL : prev + amount;
            const unsigned long long seen = atomicCAS(target, prev, sum);
            if (seen == prev) break;
            prev = seen;
        }
    }
}

// Scatter fine edges into their coarse rows (bucket order is arbitrary; the
// row sort makes the output canonical).  Internal edges become kSentinel.
struct ScatterBody {
    const std::int32_t* off;
    const std::int32_t* nbr;
    const std::uint32_t* ew;
    const std::int32_t* map;
    unsigned long long* cursor;
    std::int32_t* rkey;
    std::uint32_t* rval;
    __device__ std::int64_t degree(std::int32_t v) const { return off[v + 1] - off[v]; }
    template <class Team>
    __device__ RowOut operator()(const Team& t, std::int32_t v, unsigned long long*) const {
        const std::int64_t b = off[v], e = off[v + 1];
        if (e == b) return RowOut{false, 0};
        const std::int32_t s = map[v];
        unsigned long long o = 0;
        if (t.leader()) o = atomicAdd(&cursor[s], (unsigned long long)(e - b));
        o = t.bcast(o);
        row_batches(
            t, b, e, [&](std::int64_t x) { return map[nbr[x]]; },
            [&](std::int64_t x, std::int32_t tc) {
                rkey[o + (x - b)] = tc == s ? kSentinel : tc;
                rval[o + (x - b)] = (std::uint32_t)ew_at(ew, x);
            });
        return RowOut{false, 0};
    }
};

struct RowArgs {
    const std::int64_t* raw_off;
    std::int32_t* rkey;
    std::uint32_t* rval;
    unsigned long long* ucount;
    unsigned long long* overflow;
};

// Raw long coarse rows as record sources.
struct LongSrc {
    RowArgs r;
    const std::int32_t* list;
    __device__ bool active(std::int64_t) const { return true; }
    __device__ std::int64_t begin(std::int64_t i) const { return r.raw_off[list[i]]; }
    __device__ std::int64_t end(std::int64_t i) const { return r.raw_off[list[i] + 1]; }
    __device__ std::int32_t key(std::int64_t, std::int64_t x) const { return r.rkey[x]; }
    __device__ std::uint32_t val(std::int64_t, std::int64_t x) const { return r.rval[x]; }
};

__device__ inline void emit_row_entry(const RowArgs& r, std::int64_t start, std::int32_t key, std::uint64_t sum,
                                      int idx) {
    if (sum > 0xffffffffULL) atomicOr(r.overflow, 1ULL);
    r.rkey[start + idx] = key;
    r.rval[start + idx] = (std::uint32_t)sum;
}

// Sort-reduce coarse rows in place: unique entries written at the row start.
template <int W, int I>
__global__ void k_rows_group(RowArgs r, const std::int32_t* list, std::int64_t count) {
    const int lg = threadIdx.x % W;
    const int group_base = (threadIdx.x % kWave) / W * W;
    const std::int64_t groups = (std::int64_t)gridDim.x * (blockDim.x / W);
    for (std::int64_t idx = blockIdx.x * (std::int64_t)(blockDim.x / W) + threadIdx.x / W; idx < count;
         idx += groups) {
        const std::int32_t row = list[idx];
        const std::int64_t b = r.raw_off[row], d = r.raw_off[row + 1] - b;
        std::int32_t k[I];
        std::uint32_t w[I];
#pragma unroll
        for (int it = 0; it < I; ++it) {
            const std::int64_t x = it * W + lg;
            k[it] = x < d ? r.rkey[b + x] : kSentinel;
            w[it] = x < d ? r.rval[b + x] : 0u;
        }
        group_bitonic_sort<W, I>(k, w, lg);
        const int u = group_reduce_runs<W, I>(
            k, w, lg, group_base, [&](std::int32_t key, std::uint64_t s, int i) { emit_row_entry(r, b, key, s, i); });
        if (lg == 0) r.ucount[row] = (unsigned long long)u;
    }
}

// Coarse keys are below 2^key_bits, so an LDS radix sort needs
// ceil((key_bits + 1) / 8) passes.
__global__ void __launch_bounds__(kBlockSortThreads) k_rows_block(RowArgs r, const std::int32_t* list,
                                                                  std::int64_t count, int key_bits) {
    __shared__ BlockSortSmem s;
    __shared__ BlockRadixSmem x;
    for (std::int64_t idx = blockIdx.x; idx < count; idx += gridDim.x) {
        const std::int32_t row = list[idx];
        const std::int64_t b = r.raw_off[row];
        const int d = (int)(r.raw_off[row + 1] - b);
        for (int i = threadIdx.x; i < d; i += kBlockSortThreads) {
            s.key[i] = r.rkey[b + i];
            s.val[i] = r.rval[b + i];
        }
        __syncthreads();
        block_radix_sort(s, x, d, key_bits);
        const int N = block_pad(s, d);
        const int u = block_reduce_sorted(s, N, [&](std::int32_t key, std::uint64_t sum, int i) {
            emit_row_entry(r, b, key, sum, i);
        });
        if (threadIdx.x == 0) r.ucount[row] = (unsigned long long)u;
        __syncthreads();
    }
}

// Long rows: records keyed by (long-row index, target).
__global__ void k_long_records(RowArgs r, const std::int32_t* list, std::int64_t count, const std::int64_t* rec_off,
                               int key_bits, std::uint64_t* keys, std::uint32_t* vals) {
    for (std::int64_t i = blockIdx.x; i < count; i += gridDim.x) {
        const std::int32_t row = list[i];
        const std::int64_t b = r.raw_off[row], d = r.raw_off[row + 1] - b;
        const std::int64_t o = rec_off[i];
        for (std::int64_t x = threadIdx.x; x < d; x += blockDim.x) {
            const std::int32_t key = r.rkey[b + x];
            keys[o + x] = key == kSentinel ? ~0ULL : (((std::uint64_t)i << key_bits) | (std::uint64_t)key);
            vals[o + x] = r.rval[b + x];
        }
    }
}
// Write reduced long-row entries back at their row start; one block per row.
__global__ void k_long_writeback(RowArgs r, const std::int32_t* list, std::int64_t count, const std::uint64_t* ukeys,
                                 const unsigned long long* usums, const std::int64_t* ranges, int key_bits) {
    for (std::int64_t li = blockIdx.x; li < count; li += gridDim.x) {
        const std::int64_t first = ranges[2 * li], last = ranges[2 * li + 1];
        const std::int32_t row = list[li];
        const std::int64_t start = r.raw_off[row];
        if (threadIdx.x == 0) r.ucount[row] = (unsigned long long)(last - first);
        for (std::int64_t j = first + threadIdx.x; j < last; j += blockDim.x)
            emit_row_entry(r, start, (std::int32_t)(ukeys[j] & ((1ULL << key_bits) - 1)), usums[j], (int)(j - first));
    }
}

// Lengths of listed rows (offsets of type O), for record layouts.
template <typename O>
__global__ void k_list_lengths(const std::int32_t* list, std::int64_t count, const O* off, std::int64_t* out) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x)
        out[i] = (std::int64_t)(off[list[i] + 1] - off[list[i]]);
}

// Compacts reduced rows into the final CSR; coff has rows + 1 entries.
struct CopyBody {
    const std::int64_t* raw_off;
    const std::int32_t* rkey;
    const std::uint32_t* rval;
    const std::int64_t* coff;
    std::int32_t* off;
    std::int32_t* nbr;
    std::uint32_t* ew;
    std::int64_t rows;
    __device__ std::int64_t degree(std::int32_t r) const { return coff[r + 1] - coff[r]; }
    template <class Team>
    __device__ RowOut operator()(const Team& t, std::int32_t row, unsigned long long*) const {
        const std::int64_t src = raw_off[row], dst = coff[row], cnt = coff[row + 1] - dst;
        if (t.leader()) {
            off[row] = (std::int32_t)dst;
            if (row == rows - 1) off[rows] = (std::int32_t)coff[rows];
        }
        for (std::int64_t x = t.rank(); x < cnt; x += Team::size) {
            nbr[dst + x] = rkey[src + x];
            ew[dst + x] = rval[src + x];
        }
        return RowOut{false, 0};
    }
};

inline int grid_for(std::int64_t n, int per_block = 256, int cap = 8192) {
    return static_cast<int>(std::max<std::int64_t>(1, std::min<std::int64_t>(ceil_div(n, per_block), cap)));
}

// Launch group/block/hub kernels over degree-class lists.
template <typename GroupFn, typename BlockFn>
void launch_classes(Context& c, const std::int32_t* lists, std::int64_t stride, const unsigned long long* counts,
                    GroupFn group, BlockFn block) {
    (void)c;
    group(0, lists + 0 * stride, (std::int64_t)counts[0]);
    group(1, lists + 1 * stride, (std::int64_t)counts[1]);
    group(2, lists + 2 * stride, (std::int64_t)counts[2]);
    group(3, lists + 3 * stride, (std::int64_t)counts[3]);
    group(4, lists + 4 * stride, (std::int64_t)counts[4]);
    block(lists + 5 * stride, (std::int64_t)counts[5]);
}

}  // namespace

// ====================================================================== host

// Global reduce of (row,key) records: sort, unique, per-run sums.
static std::int64_t reduce_records(Context& c, DBuf<std::uint64_t>& keys, DBuf<std::uint32_t>& vals, std::int64_t count,
                                   int total_bits, DBuf<std::uint64_t>& ukeys, DBuf<unsigned long long>& usums) {
    radix_sort_pairs(c, keys, vals, count, 0, total_bits);
    const std::int64_t tiles = std::max<std::int64_t>(1, ceil_div(count, kScanTile));
    DBuf<std::int64_t> tile_base = c.buf<std::int64_t>(static_cast<std::size_t>(tiles));
    DBuf<std::int64_t> total = c.buf<std::int64_t>(1);
    if (count > 0) {
        hipLaunchKernelGGL(k_run_tile_heads, dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0, c.stream(),
                           keys.get(), count, tile_base.get());
        GP_LAUNCH_CHECK();
    } else {
        fill<std::int64_t>(c, tile_base.get(), 1, 0);
    }
    exclusive_sum<std::int64_t>(c, tile_base.get(), tile_base.get(), tiles, total.get());
    const std::int64_t nu = c.read_scalar(total.get());
    ukeys = c.buf<std::uint64_t>(static_cast<std::size_t>(std::max<std::int64_t>(nu, 1)));
    usums = c.buf<unsigned long long>(static_cast<std::size_t>(std::max<std::int64_t>(nu, 1)));
    fill<unsigned long long>(c, usums.get(), nu, 0ULL);
    if (count > 0 && nu > 0)
        hipLaunchKernelGGL(k_run_reduce, dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0, c.stream(),
                           keys.get(), vals.get(), count, tile_base.get(), ukeys.get(), usums.get());
    GP_LAUNCH_CHECK();
    return nu;
}

// Pre-reduces the records of the listed long rows chunk by chunk, then
// reduces them globally.  Returns false (nothing reduced) when a partial
// sum does not fit the 32-bit record weight; the caller then reduces the
// raw records.
template <class Src>
static bool reduce_long_rows(Context& c, const Src& src, std::int64_t count, std::int64_t raw_records, int key_bits,
                             int total_bits, std::uint64_t row_weight_bound, DBuf<std::uint64_t>& ukeys,
                             DBuf<unsigned long long>& usums, std::int64_t& nu) {
    DBuf<std::int64_t> chunk_off = c.buf<std::int64_t>(static_cast<std::size_t>(count + 1));
    DBuf<std::int64_t> ctot = c.buf<std::int64_t>(1);
    hipLaunchKernelGGL((k_chunk_counts<Src>), dim3(grid_for(count)), dim3(256), 0, c.stream(), src, count,
                       chunk_off.get());
    GP_LAUNCH_CHECK();
    exclusive_sum<std::int64_t>(c, chunk_off.get(), chunk_off.get(), count, ctot.get());
    c.copy_d2d(chunk_off.get() + count, ctot.get(), sizeof(std::int64_t));
    DBuf<unsigned long long> ctr = c.buf<unsigned long long>(2);
    fill<unsigned long long>(c, ctr.get(), 2, 0ULL);
    DBuf<std::uint64_t> keys = c.buf<std::uint64_t>(static_cast<std::size_t>(std::max<std::int64_t>(raw_records, 1)));
    DBuf<std::uint32_t> vals = c.buf<std::uint32_t>(static_cast<std::size_t>(std::max<std::int64_t>(raw_records, 1)));
    const std::int64_t max_chunks = raw_records / kBlockSortMax + count;
    if (row_weight_bound <= 0xffffffffULL)
        hipLaunchKernelGGL((k_chunk_records_hash<Src>), dim3(grid_for(max_chunks, 1, 4096)), dim3(kBlockSortThreads),
                           0, c.stream(), src, chunk_off.get(), count, key_bits, keys.get(), vals.get(), ctr.get());
    else
        hipLaunchKernelGGL((k_chunk_records<Src>), dim3(grid_for(max_chunks, 1, 4096)), dim3(kBlockSortThreads), 0,
                           c.stream(), src, chunk_off.get(), count, key_bits, keys.get(), vals.get(), ctr.get(),
                           ctr.get() + 1);
    GP_LAUNCH_CHECK();
    unsigned long long h[2];
    c.read(h, ctr.get(), sizeof h);
    if (h[1]) return false;
    nu = h[0] ? reduce_records(c, keys, vals, (std::int64_t)h[0], total_bits, ukeys, usums) : 0;
    if (!h[0]) {
        ukeys = c.buf<std::uint64_t>(1);
        usums = c.buf<unsigned long long>(1);
    }
    return true;
}

void compute_level0_metadata(Context& c, DGraph& g, CoarsenState& st) {
    st.row_bound = c.buf<std::uint64_t>(static_cast<std::size_t>(g.n));
    DBuf<MetaOut> out = c.buf<MetaOut>(1);
    fill<unsigned long long>(c, reinterpret_cast<unsigned long long*>(out.get()), sizeof(MetaOut) / 8, 0ULL);
    if (g.n > 0) {
        DBuf<std::int32_t> ta = c.buf<std::int32_t>(static_cast<std::size_t>(g.n));
        DBuf<std::int32_t> tb = c.buf<std::int32_t>(static_cast<std::size_t>(g.n));
        DBuf<unsigned long long> tn = c.buf<unsigned long long>(2);
        const MetaBody body{g.off.get(), g.nbr.get(), g.ew_ptr(), st.row_bound.get()};
        run_tiers(c, body, nullptr, g.n, nullptr, ta.get(), tb.get(), tn.get(), nullptr, nullptr, reinterpret_cast<unsigned long long*>(out.get()) + 1,
                  0, g.n);
        hipLaunchKernelGGL(k_vertex_meta, dim3(grid_for(g.n, 256, 1024)), dim3(256), 0, c.stream(), g.n, g.off.get(),
                           g.vw.get(), st.row_bound.get(), out.get());
        GP_LAUNCH_CHECK();
        if (g.ew && g.m > 0) {
            hipLaunchKernelGGL(k_edge_flags, dim3(grid_for(g.m, 256, 1024)), dim3(256), 0, c.stream(), g.m, g.ew.get(),
                               out.get());
            GP_LAUNCH_CHECK();
        }
    }
    MetaOut h = c.read_scalar(out.get());
    st.max_row_weight = h.max_row;
    g.incident_bound = h.max_row;
    st.sorted_unique_rows = h.unsorted == 0;
    st.positive_ew = (h.flags & 1ULL) == 0;
    st.unit_ew = (h.flags & 2ULL) == 0;
    g.max_degree = h.max_deg;
    g.row_len_bound = h.max_deg;
    g.has_heavy = h.max_deg >= 4096;
    st.total_weight = h.total_vw;
    st.max_vertex_weight = h.max_vw;
}

AggregateResult aggregate(Context& c, const DGraph& g, CoarsenState& st, int parts, std::uint32_t seed, int level,
                          int level_factor, bool pack_isolated, bool seed_matching, bool fast_path,
                          DBuf<std::int32_t>& map, DBuf<std::uint32_t>& coarse_vw) {
    const std::int64_t n = g.n;
    const hipStream_t s = c.stream();
    const std::uint64_t total = st.total_weight, maxv = st.max_vertex_weight;
    const std::uint64_t denom = (std::uint64_t)level_factor * (std::uint64_t)parts;
    const std::uint64_t cap = std::max(maxv, total / denom + (total % denom != 0));
    AggregateResult res;
    res.capacity = cap;

    DBuf<std::int32_t> clusters = c.buf<std::int32_t>(n);
    DBuf<unsigned long long> cw = c.buf<unsigned long long>(n);
    DBuf<std::uint8_t> roles = c.buf<std::uint8_t>(n);
    DBuf<std::int32_t> proposals = c.buf<std::int32_t>(n);
    DBuf<std::uint64_t> gains = c.buf<std::uint64_t>(n);
    DBuf<std::uint64_t> regret = c.buf<std::uint64_t>(n);
    DBuf<std::uint32_t> ties = c.buf<std::uint32_t>(n);
    DBuf<unsigned long long> scalar = c.buf<unsigned long long>(4);
    // scalar[0] accumulates one count or maximum at a time; reading it back
    // also clears it, so only its first use needs a fill.
    bool scalar_zero = false;
    auto scalar_ready = [&] {
        if (!scalar_zero) fill<unsigned long long>(c, scalar.get(), 1, 0ULL);
        scalar_zero = false;
    };
    auto scalar_take = [&]() -> unsigned long long {
        unsigned long long v = 0;
        c.read_clear(&v, scalar.get(), 1, 1);
        scalar_zero = true;
        return v;
    };
    const int vb = grid_for(n);

    std::int64_t cluster_count = n;
    if (seed_matching) {
        DBuf<std::int32_t> choices = c.buf<std::int32_t>(n);
        const std::uint32_t salt = seed ^ ((std::uint32_t)level * 0x9e3779b9U);
        hipLaunchKernelGGL(k_match_choose, dim3(vb), dim3(256), 0, s, n, g.off.get(), g.nbr.get(), g.ew_ptr(),
                           g.vw.get(), cap, salt, choices.get());
        GP_LAUNCH_CHECK();
        fill<unsigned long long>(c, cw.get(), n, 0ULL);
        scalar_ready();
        hipLaunchKernelGGL(k_match_roots, dim3(vb), dim3(256), 0, s, n, choices.get(), g.vw.get(), clusters.get(),
                           cw.get(), scalar.get());
        GP_LAUNCH_CHECK();
        cluster_count = (std::int64_t)scalar_take();
    } else {
        hipLaunchKernelGGL(k_init_singletons, dim3(vb), dim3(256), 0, s, n, g.vw.get(), clusters.get(), cw.get());
        GP_LAUNCH_CHECK();
    }

    // Degree classes (lists built once per level).
    DBuf<std::int32_t> lists = c.buf<std::int32_t>(static_cast<std::size_t>(7 * n));
    DBuf<ClassCounts> ccount = c.buf<ClassCounts>(1);
    fill<unsigned long long>(c, reinterpret_cast<unsigned long long*>(ccount.get()), 8, 0ULL);
    classify(c, n, g.off.get(), lists.get(), n, ccount.get());
    GP_LAUNCH_CHECK();
    ClassCounts cc = c.read_scalar(ccount.get());
    std::uint64_t stats_maxdeg = cc.c[7];
    bool stats_heavy = cc.c[7] >= 4096;
    if (g.max_degree <= 8) {  // a bound <= 8 is trusted (src/coarsen.cu:84-90)
        stats_maxdeg = g.max_degree;
        stats_heavy = false;
    }
    res.stats_max_degree = stats_maxdeg;
    res.stats_has_heavy = stats_heavy;

    // Hub record layout (rows > 4096) prepared once per level; with row
    // weights below 2^32 the rows of at most kBigRowDegree entries are split
    // off for the one-block path.
    std::int64_t hub_count = (std::int64_t)cc.c[6];
    const std::int32_t* hub_list = lists.get() + 6 * n;
    DBuf<std::int32_t> big_hubs, rest_hubs;
    std::int64_t big_count = 0;
    if (hub_count && st.max_row_weight <= 0xffffffffULL) {
        big_hubs = c.buf<std::int32_t>(static_cast<std::size_t>(hub_count));
        rest_hubs = c.buf<std::int32_t>(static_cast<std::size_t>(hub_count));
        DBuf<unsigned long long> hc = c.buf<unsigned long long>(2);
        fill<unsigned long long>(c, hc.get(), 2, 0ULL);
        hipLaunchKernelGGL(k_split_hubs, dim3(grid_for(hub_count)), dim3(256), 0, s, g.off.get(), hub_list, hub_count,
                           big_hubs.get(), rest_hubs.get(), hc.get());
        GP_LAUNCH_CHECK();
        unsigned long long h2[2];
        c.read(h2, hc.get(), sizeof h2);
        big_count = (std::int64_t)h2[0];
        hub_count = (std::int64_t)h2[1];
        hub_list = rest_hubs.get();
    }
    DBuf<std::int64_t> hub_off;
    std::int64_t hub_records = 0;
    if (hub_count) {
        hub_off = c.buf<std::int64_t>(static_cast<std::size_t>(hub_count));
        DBuf<std::int64_t> tot = c.buf<std::int64_t>(1);
        hipLaunchKernelGGL((k_list_lengths<std::int32_t>), dim3(grid_for(hub_count)), dim3(256), 0, s, hub_list,
                           hub_count, g.off.get(), hub_off.get());
        GP_LAUNCH_CHECK();
        exclusive_sum<std::int64_t>(c, hub_off.get(), hub_off.get(), hub_count, tot.get());
        hub_records = c.read_scalar(tot.get());
    }
    const int cluster_bits = std::max(1, bit_width_u64(static_cast<std::uint64_t>(n)));
    const int hub_bits = std::max(1, bit_width_u64(static_cast<std::uint64_t>(hub_count)));

    auto form_proposals = [&](int round, bool filter_roles, std::uint32_t threshold, bool need_regret) {
        const std::uint32_t role_salt =
            seed ^ ((std::uint32_t)level * 0x9e3779b9U) ^ ((std::uint32_t)round * 0x85ebca6bU) ^ 0x27d4eb2dU;
        const std::uint32_t tie_salt =
            seed ^ ((std::uint32_t)level * 0xc2b2ae35U) ^ ((std::uint32_t)round * 0x165667b1U) ^ 0xd3a2646cU;
        hipLaunchKernelGGL(k_roles, dim3(vb), dim3(256), 0, s, n, clusters.get(), g.vw.get(), cw.get(), role_salt,
                           threshold, !filter_roles, roles.get(), proposals.get());
        GP_LAUNCH_CHECK();
        AffinityArgs a{g.off.get(), g.nbr.get(), g.ew_ptr(), clusters.get(), cw.get(), roles.get(), cap, tie_salt,
                       filter_roles, filter_roles && need_regret, proposals.get(), gains.get(), regret.get(),
                       ties.get()};
        // The hub rows' records and the other classes touch disjoint
        // vertices: with hub records (and tier overlap on) the other classes
        // run on the side stream while the records are sorted here.
        const bool side = hub_count > 0 && c.overlap();
        if (side) c.fork();
        const hipStream_t as = side ? c.aux() : s;
        auto group = [&](int cls, const std::int32_t* list, std::int64_t count) {
            if (!count) return;
            switch (cls) {
                case 0:
                    hipLaunchKernelGGL((k_affinity_group<16, 1>), dim3(grid_for(count, 16)), dim3(256), 0, as, a, list,
                                       count);
                    break;
                case 1:
                    hipLaunchKernelGGL((k_affinity_group<64, 1>), dim3(grid_for(count, 4)), dim3(256), 0, as, a, list,
                                       count);
                    break;
                case 2:
                    hipLaunchKernelGGL((k_affinity_group<64, 2>), dim3(grid_for(count, 4)), dim3(256), 0, as, a, list,
                                       count);
                    break;
                case 3:
                    hipLaunchKernelGGL((k_affinity_group<64, 4>), dim3(grid_for(count, 4)), dim3(256), 0, as, a, list,
                                       count);
                    break;
                default:
                    hipLaunchKernelGGL((k_affinity_group<64, 8>), dim3(grid_for(count, 4)), dim3(256), 0, as, a, list,
                                       count);
                    break;
            }
            GP_LAUNCH_CHECK();
        };
        auto block = [&](const std::int32_t* list, std::int64_t count) {
            if (!count) return;
            // The 64 KiB table allows two blocks per CU, so each block
            // brings eight waves.
            if (st.max_row_weight <= 0xffffffffULL)
                hipLaunchKernelGGL((k_affinity_block_hash<BlockHashSmem, kBlockDegree, kBlockHashThreads>),
                                   dim3(grid_for(count, 1, 4096)), dim3(kBlockHashThreads), 0, as, a, list, count);
            else
                hipLaunchKernelGGL(k_affinity_block, dim3(grid_for(count, 1, 4096)), dim3(kBlockSortThreads), 0, as, a,
                                   list, count);
            GP_LAUNCH_CHECK();
        };
        launch_classes(c, lists.get(), n, cc.c, group, block);
        if (big_count) {
            hipLaunchKernelGGL((k_affinity_block_hash<BigHashSmem, kBigRowDegree, 1024>),
                               dim3(grid_for(big_count, 1, 4096)), dim3(1024), 0, as, a, big_hubs.get(), big_count);
            GP_LAUNCH_CHECK();
        }
        if (hub_count) {
            DBuf<std::uint64_t> ukeys;
            DBuf<unsigned long long> usums;
            std::int64_t nu = 0;
            if (!reduce_long_rows(c, HubSrc{a, hub_list}, hub_count, hub_records, cluster_bits, cluster_bits + hub_bits,
                                  st.max_row_weight, ukeys, usums, nu)) {
                DBuf<std::uint64_t> keys = c.buf<std::uint64_t>(static_cast<std::size_t>(hub_records));
                DBuf<std::uint32_t> vals = c.buf<std::uint32_t>(static_cast<std::size_t>(hub_records));
                hipLaunchKernelGGL(k_hub_records, dim3(grid_for(hub_count, 1, 4096)), dim3(256), 0, s, a, hub_list,
                                   hub_count, hub_off.get(), cluster_bits, keys.get(), vals.get());
                GP_LAUNCH_CHECK();
                nu = reduce_records(c, keys, vals, hub_records, cluster_bits + hub_bits, ukeys, usums);
            }
            DBuf<HubTop2> acc = c.buf<HubTop2>(static_cast<std::size_t>(hub_count));
            DBuf<std::int64_t> ranges = c.buf<std::int64_t>(static_cast<std::size_t>(2 * hub_count));
            fill<std::int64_t>(c, ranges.get(), 2 * hub_count, 0);
            if (nu > 0) {
                hipLaunchKernelGGL(k_row_ranges, dim3(grid_for(nu)), dim3(256), 0, s, ukeys.get(), nu, cluster_bits,
                                   ranges.get());
                GP_LAUNCH_CHECK();
            }
            hipLaunchKernelGGL(k_hub_reduce, dim3(grid_for(hub_count, 1, 4096)), dim3(256), 0, s, a, hub_list,
                               hub_count, ukeys.get(), usums.get(), ranges.get(), cluster_bits, acc.get());
            GP_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_hub_decode, dim3(grid_for(hub_count)), dim3(256), 0, s, a, hub_list, hub_count,
                               acc.get());
            GP_LAUNCH_CHECK();
        }
        if (side) c.join();
    };

    auto count_clusters = [&]() {
        scalar_ready();
        hipLaunchKernelGGL(k_count_nonzero, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, n, cw.get(), scalar.get());
        GP_LAUNCH_CHECK();
        return (std::int64_t)scalar_take();
    };

    const int score_bits = std::max(1, std::min(64, bit_width_u64(st.max_row_weight)));
    const int target_bits = std::max(1, bit_width_u64(static_cast<std::uint64_t>(n - 1)));
    // Stable selection of valid proposals then LSD sort by
    // (target asc, score desc, tie desc); vertex order breaks ties.
    // select: vertices to order (null: every valid proposal).
    auto admission_order = [&](const std::uint64_t* score, DBuf<std::uint64_t>& keys, DBuf<std::uint32_t>& vals,
                               const std::uint8_t* select) -> std::int64_t {
        DBuf<std::uint8_t> flags;
        DBuf<std::int32_t> order = c.buf<std::int32_t>(n);
        DBuf<std::int64_t> cnt = c.buf<std::int64_t>(1);
        if (!select) {
            flags = c.buf<std::uint8_t>(n);
            hipLaunchKernelGGL(k_valid_flags, dim3(vb), dim3(256), 0, s, n, proposals.get(), flags.get());
            GP_LAUNCH_CHECK();
        }
        stable_select(c, select ? select : flags.get(), n, order.get(), cnt.get());
        const std::int64_t count = c.read_scalar(cnt.get());
        if (!count) return 0;
        keys = c.buf<std::uint64_t>(static_cast<std::size_t>(count));
        vals = c.buf<std::uint32_t>(static_cast<std::size_t>(count));
        const bool fits = 32 + score_bits + target_bits <= 64;
        const int low_bits = std::min(score_bits, 32);
        hipLaunchKernelGGL(k_admission_keys, dim3(grid_for(count)), dim3(256), 0, s, count, order.get(),
                           proposals.get(), score, ties.get(), fits ? score_bits : low_bits, fits, keys.get(),
                           vals.get());
        GP_LAUNCH_CHECK();
        if (fits) {
            radix_sort_pairs(c, keys, vals, count, 0, 32 + score_bits + target_bits);
        } else {
            radix_sort_pairs(c, keys, vals, count, 0, 32 + low_bits);
            if (score_bits > low_bits) {
                hipLaunchKernelGGL(k_key_field, dim3(grid_for(count)), dim3(256), 0, s, count, vals.get(),
                                   proposals.get(), score, 1, low_bits, keys.get());
                GP_LAUNCH_CHECK();
                radix_sort_pairs(c, keys, vals, count, 0, score_bits - low_bits);
            }
            hipLaunchKernelGGL(k_key_field, dim3(grid_for(count)), dim3(256), 0, s, count, vals.get(),
                               proposals.get(), score, 0, 0, keys.get());
            GP_LAUNCH_CHECK();
            radix_sort_pairs(c, keys, vals, count, 0, target_bits);
        }
        return count;
    };

    const std::int64_t desired = (n + 1) / 2;
    std::uint64_t bound = seed_matching ? (maxv > cap / 2 ? cap : 2 * maxv) : maxv;
    for (int round = 0; round < 4; ++round) {
        if ((double)cluster_count <= 0.60 * (double)n) break;
        const std::int64_t removable = std::max<std::int64_t>(0, cluster_count - desired);
        const double frac = std::min(0.50, (double)removable / (double)cluster_count);
        const std::uint32_t threshold = (std::uint32_t)(frac * 4294967296.0);
        const std::uint64_t room = cap - bound;
        const std::uint64_t degree = stats_maxdeg;
        const bool fit = fast_path && (degree == 0 || bound <= room / maxv / degree);
        const std::uint64_t next_bound = fit ? bound + bound * degree * maxv : cap;
        form_proposals(round, true, threshold, !fit);
        if (fit) {
            hipLaunchKernelGGL(k_commit_all, dim3(vb), dim3(256), 0, s, n, proposals.get(), g.vw.get(),
                               clusters.get(), cw.get());
            GP_LAUNCH_CHECK();
        } else {
            DBuf<std::uint64_t> keys;
            DBuf<std::uint32_t> vals;
            DBuf<unsigned long long> tot = c.buf<unsigned long long>(n);
            DBuf<std::uint8_t> aflags = c.buf<std::uint8_t>(n);
            DBuf<std::uint8_t> contended = c.buf<std::uint8_t>(n);
            fill<unsigned long long>(c, tot.get(), n, 0ULL);
            hipLaunchKernelGGL(k_target_totals, dim3(vb), dim3(256), 0, s, n, proposals.get(), g.vw.get(), tot.get());
            hipLaunchKernelGGL(k_contended_flags, dim3(vb), dim3(256), 0, s, n, proposals.get(), cw.get(), tot.get(),
                               cap, aflags.get());
            hipLaunchKernelGGL(k_flag_is, dim3(vb), dim3(256), 0, s, n, aflags.get(), (std::uint8_t)1,
                               contended.get());
            GP_LAUNCH_CHECK();
            const std::int64_t count = admission_order(regret.get(), keys, vals, contended.get());
            if (count) {
                DBuf<SegU64> items = c.buf<SegU64>(static_cast<std::size_t>(count));
                DBuf<std::uint64_t> base = c.buf<std::uint64_t>(static_cast<std::size_t>(count));
                hipLaunchKernelGGL(k_seg_items, dim3(grid_for(count)), dim3(256), 0, s, count, vals.get(),
                                   proposals.get(), g.vw.get(), false, items.get());
                GP_LAUNCH_CHECK();
                scan<SegPlusOp, SegU64, SegU64, SegU64>(c, items.get(), items.get(), count, false);
                hipLaunchKernelGGL(k_gather_base, dim3(grid_for(count)), dim3(256), 0, s, count, vals.get(),
                                   proposals.get(), cw.get(), base.get());
                GP_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_admission_commit, dim3(grid_for(count)), dim3(256), 0, s, count, vals.get(),
                                   proposals.get(), g.vw.get(), items.get(), base.get(), cap, clusters.get(),
                                   cw.get());
                GP_LAUNCH_CHECK();
            }
            // Sources (movers) and targets (non-movers) are disjoint, so the
            // frozen decisions of uncontended targets stay valid here.
            hipLaunchKernelGGL(k_commit_uncontended, dim3(vb), dim3(256), 0, s, n, proposals.get(), aflags.get(),
                               g.vw.get(), clusters.get(), cw.get());
            GP_LAUNCH_CHECK();
        }
        cluster_count = count_clusters();
        bound = next_bound;
        if ((double)cluster_count <= 0.60 * (double)n) break;
    }

    if (pack_isolated) {
        DBuf<std::uint8_t> flags = c.buf<std::uint8_t>(n);
        DBuf<std::int32_t> iso = c.buf<std::int32_t>(n);
        DBuf<std::int64_t> cnt = c.buf<std::int64_t>(1);
        hipLaunchKernelGGL(k_isolated_flags, dim3(vb), dim3(256), 0, s, n, g.off.get(), flags.get());
        GP_LAUNCH_CHECK();
        stable_select(c, flags.get(), n, iso.get(), cnt.get());
        const std::int64_t count = c.read_scalar(cnt.get());
        if (count > 1) {
            DBuf<std::uint64_t> prefix = c.buf<std::uint64_t>(static_cast<std::size_t>(count));
            DBuf<std::uint64_t> tot = c.buf<std::uint64_t>(1);
            hipLaunchKernelGGL(k_gather_vw, dim3(grid_for(count)), dim3(256), 0, s, count, iso.get(), g.vw.get(),
                               prefix.get());
            GP_LAUNCH_CHECK();
            inclusive_sum<std::uint64_t>(c, prefix.get(), prefix.get(), count, tot.get());
            const std::uint64_t t = c.read_scalar(tot.get());
            const std::uint64_t bins = t / cap + (t % cap != 0);
            if (!bins) throw std::runtime_error("isolated vertices carry no weight");
            DBuf<std::int32_t> rep = c.buf<std::int32_t>(static_cast<std::size_t>(bins));
            fill<std::int32_t>(c, rep.get(), (std::int64_t)bins, 0x7fffffff);
            hipLaunchKernelGGL(k_isolated_rep, dim3(grid_for(count)), dim3(256), 0, s, count, iso.get(), prefix.get(),
                               g.vw.get(), cap, rep.get());
            GP_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_isolated_commit, dim3(grid_for(count)), dim3(256), 0, s, count, iso.get(),
                               prefix.get(), g.vw.get(), cap, rep.get(), clusters.get(), cw.get());
            GP_LAUNCH_CHECK();
            cluster_count = count_clusters();
        }
    }

    if ((double)cluster_count > 0.60 * (double)n) {
        form_proposals(4, false, 0, false);
        hipLaunchKernelGGL(k_favorites, dim3(vb), dim3(256), 0, s, n, g.vw.get(), cw.get(), clusters.get(),
                           proposals.get());
        GP_LAUNCH_CHECK();
        DBuf<std::uint64_t> keys;
        DBuf<std::uint32_t> vals;
        const std::int64_t count = admission_order(gains.get(), keys, vals, nullptr);
        if (count) {
            DBuf<SegU64> ranks = c.buf<SegU64>(static_cast<std::size_t>(count));
            hipLaunchKernelGGL(k_seg_items, dim3(grid_for(count)), dim3(256), 0, s, count, vals.get(), proposals.get(),
                               g.vw.get(), true, ranks.get());
            GP_LAUNCH_CHECK();
            scan<SegPlusOp, SegU64, SegU64, SegU64>(c, ranks.get(), ranks.get(), count, false);
            DBuf<std::uint8_t> pflags = c.buf<std::uint8_t>(static_cast<std::size_t>(count));
            DBuf<std::int64_t> pos = c.buf<std::int64_t>(static_cast<std::size_t>(count));
            hipLaunchKernelGGL(k_pair_flags, dim3(grid_for(count)), dim3(256), 0, s, count, vals.get(), g.vw.get(),
                               ranks.get(), cap, pflags.get());
            GP_LAUNCH_CHECK();
            exclusive_sum<std::int64_t>(c, pflags.get(), pos.get(), count);
            const std::int64_t budget = std::max<std::int64_t>(0, cluster_count - desired);
            if (budget > 0) {
                hipLaunchKernelGGL(k_pair_commit, dim3(grid_for(count)), dim3(256), 0, s, count, vals.get(),
                                   pflags.get(), pos.get(), budget, g.vw.get(), clusters.get(), cw.get());
                GP_LAUNCH_CHECK();
            }
        }
        cluster_count = count_clusters();
    }

    // Compaction.
    {
        DBuf<std::uint8_t> flags = c.buf<std::uint8_t>(n);
        DBuf<std::int64_t> ids = c.buf<std::int64_t>(n);
        DBuf<std::int64_t> cnt = c.buf<std::int64_t>(1);
        hipLaunchKernelGGL(k_nonempty, dim3(vb), dim3(256), 0, s, n, cw.get(), flags.get());
        GP_LAUNCH_CHECK();
        exclusive_sum<std::int64_t>(c, flags.get(), ids.get(), n, cnt.get());
        const std::int64_t nc = c.read_scalar(cnt.get());
        map = c.buf<std::int32_t>(n);
        DBuf<unsigned long long> sums = c.buf<unsigned long long>(static_cast<std::size_t>(nc));
        fill<unsigned long long>(c, sums.get(), nc, 0ULL);
        hipLaunchKernelGGL(k_compact_map, dim3(vb), dim3(256), 0, s, n, clusters.get(), ids.get(), g.vw.get(),
                           map.get(), sums.get());
        GP_LAUNCH_CHECK();
        coarse_vw = c.buf<std::uint32_t>(static_cast<std::size_t>(nc));
        scalar_ready();
        hipLaunchKernelGGL(k_narrow_weights, dim3(grid_for(nc)), dim3(256), 0, s, nc, sums.get(), coarse_vw.get(),
                           scalar.get());
        GP_LAUNCH_CHECK();
        res.max_weight = scalar_take();
        res.coarse_n = nc;
        if (res.max_weight > 0xffffffffULL) throw std::runtime_error("coarse vertex weight exceeds uint32");
        if (res.max_weight > cap) throw std::runtime_error("coarsening exceeded cluster capacity");
    }
    return res;
}

std::unique_ptr<DGraph> contract(Context& c, const DGraph& fine, CoarsenState& st, const std::int32_t* map,
                                 std::int64_t nc, DBuf<std::uint32_t>&& coarse_vw) {
    const std::int64_t n = fine.n;
    const hipStream_t s = c.stream();
    auto out = std::make_unique<DGraph>();
    out->n = nc;
    // Bounds carried through the map (src/coarsen.cu:1470-1513).
    DBuf<unsigned long long> next_rb = c.buf<unsigned long long>(static_cast<std::size_t>(nc));
    DBuf<unsigned long long> rawc = c.buf<unsigned long long>(static_cast<std::size_t>(nc));
    fill<unsigned long long>(c, next_rb.get(), nc, 0ULL);
    fill<unsigned long long>(c, rawc.get(), nc, 0ULL);
    const bool plain_bounds = st.max_row_weight == 0 || (std::uint64_t)n <= ~0ULL / st.max_row_weight;
    hipLaunchKernelGGL(k_contract_bounds, dim3(grid_for(n)), dim3(256), 0, s, n, map, fine.off.get(), st.row_bound.get(),
                       next_rb.get(), rawc.get(), plain_bounds);
    GP_LAUNCH_CHECK();
    // The level's small counters, cleared by one fill: the two maxima, the
    // tier counters of the scatter and of the copy, the class counts and
    // the weight-overflow flag.
    constexpr int kSmallRed = 0, kSmallTierScatter = 2, kSmallTierCopy = 4, kSmallClasses = 6, kSmallOverflow = 14,
                  kSmallWords = 15;
    static_assert(sizeof(ClassCounts) == 8 * sizeof(unsigned long long), "class counts are eight words");
    DBuf<unsigned long long> small = c.buf<unsigned long long>(kSmallWords);
    fill<unsigned long long>(c, small.get(), kSmallWords, 0ULL);
    unsigned long long* red = small.get() + kSmallRed;
    hipLaunchKernelGGL((k_reduce_max<unsigned long long>), dim3(grid_for(nc, 256, 1024)), dim3(256), 0, s, next_rb.get(), nc,
                       red);
    hipLaunchKernelGGL((k_reduce_max<unsigned long long>), dim3(grid_for(nc, 256, 1024)), dim3(256), 0, s, rawc.get(), nc,
                       red + 1);
    GP_LAUNCH_CHECK();
    unsigned long long rb[2];
    c.read(rb, red, sizeof rb);
    const bool count_raw = st.positive_ew && n >= 65536 && fine.max_degree <= 512;
    const std::uint64_t crb = rb[0];
    const std::uint64_t craw = count_raw ? rb[1] : ~0ULL;
    const std::uint64_t next_degree_bound = count_raw && fine.max_degree <= 128 ? craw : crb;

    // Raw coarse rows.
    DBuf<std::int64_t> raw_off = c.buf<std::int64_t>(static_cast<std::size_t>(nc + 1));
    DBuf<std::int64_t> rawtot = c.buf<std::int64_t>(1);
    exclusive_sum<std::int64_t>(c, rawc.get(), raw_off.get(), nc, rawtot.get());
    c.copy_d2d(raw_off.get() + nc, rawtot.get(), sizeof(std::int64_t));
    DBuf<unsigned long long> cursor = c.buf<unsigned long long>(static_cast<std::size_t>(nc));
    c.copy_d2d(cursor.get(), raw_off.get(), static_cast<std::size_t>(nc) * sizeof(std::int64_t));
    const std::int64_t m = fine.m;
    DBuf<std::int32_t> tier_a = c.buf<std::int32_t>(static_cast<std::size_t>(std::max(n, nc)));
    DBuf<std::int32_t> tier_b = c.buf<std::int32_t>(static_cast<std::size_t>(std::max(n, nc)));
    DBuf<std::int32_t> rkey = c.buf<std::int32_t>(static_cast<std::size_t>(std::max<std::int64_t>(m, 1)));
    DBuf<std::uint32_t> rval = c.buf<std::uint32_t>(static_cast<std::size_t>(std::max<std::int64_t>(m, 1)));
    if (m > 0) {
        // Algorithmic traffic: offsets, map[v], neighbours, map[neighbour],
        // optional edge weights, and the (key, weight) record written per entry.
        const double bytes = 8.0 * (double)n + 4.0 + (double)m * (8.0 + (fine.ew ? 4.0 : 0.0) + 8.0);
        const ScatterBody body{fine.off.get(), fine.nbr.get(), fine.ew_ptr(), map, cursor.get(), rkey.get(), rval.get()};
        c.timed("contract_scatter", bytes,
                [&] { run_tiers(c, body, nullptr, n, nullptr, tier_a.get(), tier_b.get(), small.get() + kSmallTierScatter,
                                nullptr, nullptr, nullptr, 0, n, fine.row_len_bound, true); });
    }
    cursor.reset();
    // Row classes by raw length.
    DBuf<std::int32_t> lists = c.buf<std::int32_t>(static_cast<std::size_t>(7 * nc));
    ClassCounts* ccount = reinterpret_cast<ClassCounts*>(small.get() + kSmallClasses);
    classify(c, nc, raw_off.get(), lists.get(), nc, ccount);
    ClassCounts cc = c.read_scalar(ccount);
    DBuf<unsigned long long> ucount = c.buf<unsigned long long>(static_cast<std::size_t>(nc));
    fill<unsigned long long>(c, ucount.get(), nc, 0ULL);
    unsigned long long* overflow = small.get() + kSmallOverflow;
    RowArgs r{raw_off.get(), rkey.get(), rval.get(), ucount.get(), overflow};
    // Rows of the short and block classes and the long rows are disjoint:
    // with long rows present (and tier overlap on) the short classes run on
    // the side stream while the long rows' record sort runs here, and the
    // stream waits for them before the unique counts are scanned.
    const bool side = cc.c[6] > 0 && c.overlap();
    if (side) c.fork();
    const hipStream_t rs = side ? c.aux() : s;
    auto group = [&](int cls, const std::int32_t* list, std::int64_t count) {
        if (!count) return;
        switch (cls) {
            case 0:
                hipLaunchKernelGGL((k_rows_group<16, 1>), dim3(grid_for(count, 16)), dim3(256), 0, rs, r, list, count);
                break;
            case 1:
                hipLaunchKernelGGL((k_rows_group<64, 1>), dim3(grid_for(count, 4)), dim3(256), 0, rs, r, list, count);
                break;
            case 2:
                hipLaunchKernelGGL((k_rows_group<64, 2>), dim3(grid_for(count, 4)), dim3(256), 0, rs, r, list, count);
                break;
            case 3:
                hipLaunchKernelGGL((k_rows_group<64, 4>), dim3(grid_for(count, 4)), dim3(256), 0, rs, r, list, count);
                break;
            default:
                hipLaunchKernelGGL((k_rows_group<64, 8>), dim3(grid_for(count, 4)), dim3(256), 0, rs, r, list, count);
                break;
        }
        GP_LAUNCH_CHECK();
    };
    auto block = [&](const std::int32_t* list, std::int64_t count) {
        if (!count) return;
        hipLaunchKernelGGL(k_rows_block, dim3(grid_for(count, 1, 4096)), dim3(kBlockSortThreads), 0, rs, r, list, count,
                           std::max(1, bit_width_u64(static_cast<std::uint64_t>(nc))));
        GP_LAUNCH_CHECK();
    };
    launch_classes(c, lists.get(), nc, cc.c, group, block);
    const std::int64_t long_count = (std::int64_t)cc.c[6];
    if (long_count) {
        const std::int32_t* long_list = lists.get() + 6 * nc;
        DBuf<std::int64_t> rec_off = c.buf<std::int64_t>(static_cast<std::size_t>(long_count));
        DBuf<std::int64_t> rtot = c.buf<std::int64_t>(1);
        hipLaunchKernelGGL((k_list_lengths<std::int64_t>), dim3(grid_for(long_count)), dim3(256), 0, s, long_list,
                           long_count, raw_off.get(), rec_off.get());
        GP_LAUNCH_CHECK();
        exclusive_sum<std::int64_t>(c, rec_off.get(), rec_off.get(), long_count, rtot.get());
        const std::int64_t records = c.read_scalar(rtot.get());
        const int key_bits = std::max(1, bit_width_u64(static_cast<std::uint64_t>(nc)));
        const int row_bits = std::max(1, bit_width_u64(static_cast<std::uint64_t>(long_count)));
        DBuf<std::uint64_t> ukeys;
        DBuf<unsigned long long> usums;
        std::int64_t nu = 0;
        if (!reduce_long_rows(c, LongSrc{r, long_list}, long_count, records, key_bits, key_bits + row_bits, crb,
                              ukeys, usums, nu)) {
            DBuf<std::uint64_t> keys = c.buf<std::uint64_t>(static_cast<std::size_t>(records));
            DBuf<std::uint32_t> vals = c.buf<std::uint32_t>(static_cast<std::size_t>(records));
            hipLaunchKernelGGL(k_long_records, dim3(grid_for(long_count, 1, 4096)), dim3(256), 0, s, r, long_list,
                               long_count, rec_off.get(), key_bits, keys.get(), vals.get());
            GP_LAUNCH_CHECK();
            nu = reduce_records(c, keys, vals, records, key_bits + row_bits, ukeys, usums);
        }
        if (nu) {
            DBuf<std::int64_t> ranges = c.buf<std::int64_t>(static_cast<std::size_t>(2 * long_count));
            fill<std::int64_t>(c, ranges.get(), 2 * long_count, 0);
            hipLaunchKernelGGL(k_row_ranges, dim3(grid_for(nu)), dim3(256), 0, s, ukeys.get(), nu, key_bits,
                               ranges.get());
            GP_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_long_writeback, dim3(grid_for(long_count, 1, 4096)), dim3(256), 0, s, r, long_list,
                               long_count, ukeys.get(), usums.get(), ranges.get(), key_bits);
            GP_LAUNCH_CHECK();
        }
    }
    if (side) c.join();
    if (c.read_scalar(overflow)) throw std::runtime_error("coarse edge weight exceeds uint32");
    // Final offsets and copy.
    DBuf<std::int64_t> coff64 = c.buf<std::int64_t>(static_cast<std::size_t>(nc + 1));
    DBuf<std::int64_t> mtot = c.buf<std::int64_t>(1);
    exclusive_sum<std::int64_t>(c, ucount.get(), coff64.get(), nc, mtot.get());
    c.copy_d2d(coff64.get() + nc, mtot.get(), sizeof(std::int64_t));
    const std::int64_t mc = c.read_scalar(mtot.get());
    if (mc > 0x7fffffffLL) throw std::runtime_error("coarse graph exceeds int32 offsets");
    out->m = mc;
    out->off = c.buf<std::int32_t>(static_cast<std::size_t>(nc + 1));
    out->nbr = c.buf<std::int32_t>(static_cast<std::size_t>(std::max<std::int64_t>(mc, 1)));
    out->ew = c.buf<std::uint32_t>(static_cast<std::size_t>(std::max<std::int64_t>(mc, 1)));
    const CopyBody copy{raw_off.get(), rkey.get(), rval.get(), coff64.get(), out->off.get(), out->nbr.get(),
                        out->ew.get(), nc};
    // Unique coarse rows are no longer than the raw ones (max rb[1]).
    run_tiers(c, copy, nullptr, nc, nullptr, tier_a.get(), tier_b.get(), small.get() + kSmallTierCopy, nullptr, nullptr,
              nullptr, 0, nc, rb[1], true);
    out->vw = std::move(coarse_vw);
    out->incident_bound = crb;
    if (st.positive_ew) {
        out->max_degree = next_degree_bound;
        out->has_heavy = next_degree_bound >= 4096;
    } else {
        out->max_degree = ~0ULL;
        out->has_heavy = true;
    }
    // Both bound the row length: the longest raw row, and (positive edge
    // weights, each unique entry weighing at least 1) the degree bound.
    out->row_len_bound = std::min<std::uint64_t>(rb[1], out->max_degree);
    st.max_row_weight = crb;
    st.row_bound = c.buf<std::uint64_t>(static_cast<std::size_t>(nc));
    c.copy_d2d(st.row_bound.get(), next_rb.get(), static_cast<std::size_t>(nc) * sizeof(std::uint64_t));
    return out;
}

Hierarchy coarsen(Context& c, DGraph& input, const CoarsenOptions& o) {
    if (input.n <= 0 || input.n > 0x7fffffffLL) throw std::invalid_argument("unsupported graph size");
    if (o.parts < 2 || o.parts > kMaxParts || o.cluster_size_factor < 2 || o.max_levels < 0)
        throw std::invalid_argument("invalid coarsening options");
    Hierarchy h;
    h.input = &input;
    CoarsenState st;
    compute_level0_metadata(c, input, st);
    const int k = o.parts;
    const std::int64_t cutoff = std::max<std::int64_t>(32, (std::int64_t)o.cluster_size_factor * k);
    const DGraph* cur = &input;
    std::uint64_t cur_max_degree = input.max_degree;
    bool cur_heavy = input.has_heavy;
    int level = 0;
    for (; level < o.max_levels && cur->n > cutoff; ++level) {
        const std::uint32_t level_seed = o.seed + (std::uint32_t)level;
        int level_factor = o.cluster_size_factor;
        if (o.progressive) {
            const std::int64_t ff = std::max<std::int64_t>(o.fine_factor, (512 + k - 1) / k);
            const std::int64_t den = (std::int64_t)o.capacity_growth * k;
            const std::int64_t remaining = (cur->n + den - 1) / den;
            level_factor = (int)std::max<std::int64_t>(level_factor, std::min(ff, remaining));
        }
        const bool try_matching = o.seed_matching && cur->n >= 65536 && !cur_heavy &&
                                  (cur->m / cur->n < 8 || cur_max_degree <= 256) &&
                                  cur->m / cur->n < o.hem_degree_limit && st.sorted_unique_rows && st.positive_ew;
        DBuf<std::int32_t> map;
        DBuf<std::uint32_t> cvw;
        // Metadata of the level as seen by aggregate (bound, not rescanned).
        DGraph* mut = const_cast<DGraph*>(cur);
        mut->max_degree = cur_max_degree;
        mut->has_heavy = cur_heavy;
        AggregateResult r = aggregate(c, *cur, st, k, level_seed, level, level_factor, o.pack_isolated, try_matching,
                                      o.fast_path, map, cvw);
        cur_max_degree = r.stats_max_degree;
        cur_heavy = r.stats_has_heavy;
        mut->max_degree = cur_max_degree;
        mut->has_heavy = cur_heavy;
        const double ratio = (double)r.coarse_n / (double)cur->n;
        if (r.coarse_n >= cur->n || ratio > o.stop_ratio) {
            h.stop_reason = "insufficient_contraction";
            break;
        }
        auto coarse = contract(c, *cur, st, map.get(), r.coarse_n, std::move(cvw));
        st.max_vertex_weight = r.max_weight;
        h.maps.push_back(std::move(map));
        h.caps.push_back(r.capacity);
        h.coarse.push_back(std::move(coarse));
        cur = h.coarse.back().get();
        cur_max_degree = cur->max_degree;
        cur_heavy = cur->has_heavy;
        st.sorted_unique_rows = true;
    }
    if (level == o.max_levels && cur->n > cutoff) h.stop_reason = "max_levels";
    return h;
}

void debug_radix_sort(Context& c, std::uint64_t* keys, std::uint32_t* vals, std::int64_t n, int begin_bit,
                      int end_bit) {
    if (n < 0 || begin_bit < 0 || end_bit > 64 || begin_bit > end_bit) throw std::invalid_argument("bad sort range");
    if (n == 0) return;
    DBuf<std::uint64_t> k = c.buf<std::uint64_t>(static_cast<std::size_t>(n));
    DBuf<std::uint32_t> v;
    if (vals) v = c.buf<std::uint32_t>(static_cast<std::size_t>(n));
    c.upload(k.get(), keys, static_cast<std::size_t>(n) * sizeof(std::uint64_t));
    if (vals) c.upload(v.get(), vals, static_cast<std::size_t>(n) * sizeof(std::uint32_t));
    radix_sort_pairs(c, k, v, n, begin_bit, end_bit);
    c.download(keys, k.get(), static_cast<std::size_t>(n) * sizeof(std::uint64_t));
    if (vals) c.download(vals, v.get(), static_cast<std::size_t>(n) * sizeof(std::uint32_t));
}

}  // namespace gp
