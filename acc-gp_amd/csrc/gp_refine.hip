// Refinement on gfx950: projection, label propagation with per-vertex
// part-connectivity tables in LDS, afterburner priority filter, exact
// incremental cut/load updates, deterministic capacity repair and the
// capacity-complementary exchange.  Restates src/refine.cu and
// src/refinement/*.cuh of the reference; order-dependent steps follow the
// canonical vertex-id order.
#include "gp_refine.hpp"
#include "gp_rows.hpp"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <type_traits>

namespace gp {
namespace {

constexpr int kBalanceBuckets = 50;
constexpr int kBalanceSections = 16;
constexpr int kBalanceWidth = kBalanceBuckets * kBalanceSections;

// Slot of bucket b = (gain bucket, section) in a part's row of the counts
// array: section-major, so the sections of one gain bucket (the donors of a
// part with similar gains, often most of them) sit kBalanceBuckets words
// apart, in different cache lines, and their atomics spread over memory
// channels instead of queueing on one line.
__host__ __device__ constexpr int count_slot(int b) {
    return (b % kBalanceSections) * kBalanceBuckets + b / kBalanceSections;
}
constexpr int kJointSeeds = 128;
constexpr int kJointSide = 4;
constexpr int kJointArchive = 4;

// Counter slots (device array of u64).
enum Ctr {
    kFrontier = 0,
    kCand = 1,
    kTier = 2,  // two slots: medium, heavy
    kMovers = 4,
    kDelta = 5,
    kDonors = 6,
    kActive = 7,
    kRowSum = 8,  // entries of the rows scanned by the last LP proposal sweep
    // Tier list counters of the second and third sweep of a round, so that
    // one clear of the whole block at the round start serves all three.
    kTierB = 9,
    kTierC = 11,
    // (entries, rows) handled by the afterburner, update and repair-proposal
    // sweeps, counted only when the context profiles kernel classes.
    kEntAfter = 13,
    kEntUpdate = 15,
    kEntRepair = 17,
    kCtrCount = 19
};

// Part ids inside refinement: k <= 128 fits a byte, which keeps the label
// array of a level small enough for the neighbour gathers to hit in cache.
using Label = std::uint8_t;

struct RepairSample {
    unsigned long long loss;    // modular sum of -gain
    unsigned long long weight;
};

// 1.5^j for j = -23..24 at index j + 23, by repeated exact-IEEE products
// and quotients (compile-time evaluation rounds like the run-time ops).
struct Pow15 {
    double p[48];
};
constexpr Pow15 make_pow15() {
    Pow15 t{};
    t.p[23] = 1.0;
    for (int j = 1; j <= 24; ++j) t.p[23 + j] = t.p[22 + j] * 1.5;
    for (int j = 1; j <= 23; ++j) t.p[23 - j] = t.p[24 - j] / 1.5;
    return t;
}
__constant__ Pow15 kPow15 = make_pow15();

// One byte lane of the packed per-vertex state word (indexing like an
// array of T; the four lanes of a vertex share one 32-bit word).
template <typename T>
struct StateLane {
    T* base;
    __host__ __device__ T& operator[](std::int64_t v) const { return base[4 * v]; }
};

// Per-vertex refinement state, one 32-bit word per vertex: target part
// (signed byte, -1 none) in byte 0, moving flag in byte 1, frontier
// membership in byte 2, lock in byte 3.  The update sweep reads a
// neighbour's whole state with one gather.
constexpr int kStateTarget = 0, kStateFlag = 1, kStateMember = 2, kStateLock = 3;

struct RefArgs {
    std::int64_t n;
    const std::int32_t* off;
    const std::int32_t* nbr;
    const std::uint32_t* ew;
    const std::uint32_t* vw;
    int k;
    std::uint64_t cap;
    double filter;
    Label* labels;
    unsigned long long* loads;
    const double* prices;  // null unless vertex weights exceed n
    std::uint32_t* state;  // packed lanes below
    StateLane<std::int8_t> targets;  // -1: none
    std::int64_t* gains;
    // When every row of the level weighs less than 2^31, the gains are kept
    // in 32 bits here instead (gains is null): the afterburner gathers half
    // the bytes from an array half the size.
    std::int32_t* gains32;
    StateLane<std::uint8_t> locks;
    StateLane<std::uint8_t> flags;
    StateLane<std::uint8_t> member;
    unsigned long long* ctr;
    RepairSample* samples;
    // repair
    const unsigned long long* loads0;
    std::int32_t* bucket;
    unsigned long long* counts;
};

// Part-connectivity table of one vertex in team-private LDS; returns the
// incident weight excluding self loops (src/refinement/affinity.cuh:9-66).
// Bin: the table's word (32 bits when no row of the level weighs 2^32).
template <class Team, class Bin = unsigned long long>
__device__ inline std::uint64_t team_histogram(const Team& t, const RefArgs& a, std::int64_t v, Bin* hist) {
    t.sync();  // earlier readers of the table are done
    for (int p = t.rank(); p < a.k; p += Team::size) hist[p] = 0;
    t.sync();
    std::uint64_t total = 0;
    struct Rec {
        int label;  // -1 for a self loop
        std::uint32_t w;
    };
    row_batches(
        t, a.off[v], a.off[v + 1],
        [&](std::int64_t x) {
            const std::int32_t u = a.nbr[(std::uint32_t)x];
            return Rec{u == v ? -1 : a.labels[(std::uint32_t)u], (std::uint32_t)ew_at(a.ew, x)};
        },
        [&](std::int64_t, const Rec& r) {
            if (r.label < 0) return;
            atomicAdd(&hist[r.label], (Bin)r.w);
            total += r.w;
        });
    t.sync();
    return t.sum(total);
}

// Heavy rows: one block, so each wave adds into its own table (hub rows
// put most of their entries into a few parts, and one shared table would
// serialise the whole block on those bins); the wave tables are then summed
// into hist[0, k).  The block's scratch holds 1 + waves tables.
constexpr int kHeavyTables = 1 + BlockTeam::size / kWave;
template <class Bin = unsigned long long>
__device__ inline std::uint64_t team_histogram(const BlockTeam& t, const RefArgs& a, std::int64_t v, Bin* hist) {
    constexpr int kWaves = BlockTeam::size / kWave;
    Bin* mine = hist + (threadIdx.x / kWave + 1) * a.k;
    t.sync();  // earlier readers of the table are done
    for (int i = t.rank(); i < kHeavyTables * a.k; i += BlockTeam::size) hist[i] = 0;
    t.sync();
    std::uint64_t total = 0;
    struct Rec {
        int label;  // -1 for a self loop
        std::uint32_t w;
    };
    row_batches(
        t, a.off[v], a.off[v + 1],
        [&](std::int64_t x) {
            const std::int32_t u = a.nbr[(std::uint32_t)x];
            return Rec{u == v ? -1 : a.labels[(std::uint32_t)u], (std::uint32_t)ew_at(a.ew, x)};
        },
        [&](std::int64_t, const Rec& r) {
            if (r.label < 0) return;
            atomicAdd(&mine[r.label], (Bin)r.w);
            total += r.w;
        });
    t.sync();
    for (int p = t.rank(); p < a.k; p += BlockTeam::size) {
        Bin sum = 0;
        for (int w = 1; w <= kWaves; ++w) sum += hist[w * a.k + p];
        hist[p] = sum;
    }
    t.sync();
    return t.sum(total);
}

// Short rows (light teams of at most 16 lanes, at most kTierLight entries):
// the row's labels stay in registers and only their bins are touched, so a
// table that starts zero is left zero and no per-row clear of all k bins is
// needed.  The light kernel zeroes the tables once at start.
template <int W>
struct LightRow {
    static constexpr int kPer = kTierLight / W;
    static constexpr int kWords = (kPer + 3) / 4;
    unsigned packed[kWords];  // four byte labels per word; 0xff: no entry or a self loop
    __device__ int lab(int j) const { return (int)((packed[j >> 2] >> (8 * (j & 3))) & 0xffu); }
};

template <int W, class Bin>
__device__ inline std::uint64_t light_histogram(const GroupTeam<W>& t, const RefArgs& a, std::int64_t v, Bin* hist,
                                                LightRow<W>& r) {
    const std::int64_t b = a.off[v], e = a.off[v + 1];
    std::uint64_t total = 0;
#pragma unroll
    for (int q = 0; q < LightRow<W>::kWords; ++q) r.packed[q] = 0xffffffffu;
    t.sync();  // the previous row's clean-up is done
    // Four entries per lane in flight; iterations beyond the longest row of
    // the wave are skipped as a whole (t.wave_rows_max is wave-uniform).
#pragma unroll
    for (int q = 0; q < LightRow<W>::kWords; ++q) {
        if (q * 4 * W >= t.wave_rows_max) break;
        int lab[4];
        std::uint32_t w[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const std::int64_t x = b + t.rank() + (q * 4 + i) * W;
            const bool in = x < e;
            const std::int32_t u = in ? a.nbr[(std::uint32_t)x] : (std::int32_t)v;
            lab[i] = u != v ? (int)a.labels[(std::uint32_t)u] : 0xff;
            w[i] = in ? (std::uint32_t)ew_at(a.ew, x) : 0u;
        }
        unsigned word = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (lab[i] != 0xff) {
                atomicAdd(&hist[lab[i]], static_cast<Bin>(w[i]));
                total += w[i];
            }
            word |= (unsigned)lab[i] << (8 * i);
        }
        r.packed[q] = word;
    }
    t.sync();
    return t.sum(total);
}

template <int W, class Bin>
__device__ inline void light_clear(const GroupTeam<W>& t, Bin* hist, const LightRow<W>& r) {
    t.sync();  // all reads of the row's bins are done
#pragma unroll
    for (int j = 0; j < LightRow<W>::kPer; ++j) {
        const int l = r.lab(j);
        if (l != 0xff) hist[l] = 0;
    }
}

template <class Team>
struct IsLight : std::false_type {};
template <int W>
struct IsLight<GroupTeam<W>> : std::integral_constant<bool, (W <= 16)> {};

// estimated_repair_cost (src/refinement/incremental.cuh:156-159).
__device__ inline double repair_cost(const RefArgs& a, int part, std::uint64_t weight) {
    const std::uint64_t slack = a.cap - a.loads[part];
    return weight > slack ? a.prices[part] * (double)(weight - slack) : 0.0;
}

struct Cand {
    int p;
    std::uint64_t conn;
    double score;
};

__device__ inline bool cand_better(const Cand& x, const Cand& y, bool priced) {
    if (x.p < 0) return false;
    if (y.p < 0) return true;
    if (priced) return x.score > y.score || (x.score == y.score && x.p < y.p);
    return x.conn > y.conn || (x.conn == y.conn && x.p < y.p);
}

template <int W>
__device__ inline Cand cand_group_reduce(Cand c, bool priced) {
    for (int d = W / 2; d; d >>= 1) {
        Cand o;
        o.p = __shfl_xor(c.p, d, W);
        o.conn = __shfl_xor(c.conn, d, W);
        o.score = __shfl_xor(c.score, d, W);
        if (cand_better(o, c, priced)) c = o;
    }
    return c;
}
template <int W>
__device__ inline Cand team_reduce(const GroupTeam<W>&, Cand c, bool priced) {
    return cand_group_reduce<W>(c, priced);
}
__device__ inline Cand team_reduce(const BlockTeam& t, Cand c, bool priced) {
    c = cand_group_reduce<kWave>(c, priced);
    __syncthreads();
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        t.red[3 * w] = (unsigned long long)(long long)c.p;
        t.red[3 * w + 1] = c.conn;
        t.red[3 * w + 2] = __double_as_longlong(c.score);
    }
    __syncthreads();
    Cand r{(int)(long long)t.red[0], t.red[1], __longlong_as_double((long long)t.red[2])};
    for (int i = 1; i < BlockTeam::size / kWave; ++i) {
        Cand o{(int)(long long)t.red[3 * i], t.red[3 * i + 1], __longlong_as_double((long long)t.red[3 * i + 2])};
        if (cand_better(o, r, priced)) r = o;
    }
    __syncthreads();
    return r;
}

// Stores a vertex's proposal gain in the level's gain array (32 or 64 bits;
// the "none" value INT64_MIN is stored as INT32_MIN in 32 bits).
__device__ inline void set_gain(const RefArgs& a, std::int64_t v, std::int64_t g) {
    if (a.gains32)
        a.gains32[v] = g == INT64_MIN ? INT32_MIN : (std::int32_t)g;
    else
        a.gains[v] = g;
}

struct RowBody {
    RefArgs a;
    __device__ std::int64_t degree(std::int32_t v) const { return a.off[v + 1] - a.off[v]; }
};

// LP proposal (incremental.cuh:166-198, 229-296).  Bin is the bin type of
// the part tables: 32 bits when no row of the level weighs 2^32 or more
// (short rows: half the LDS per team, twice the resident teams; longer rows:
// one bank per LDS atomic instead of two), else 64.
template <class Bin>
struct LpProposeBody : RowBody {
    using LightBin = Bin;
    // With prices, the scored candidates read (cap - load, price) of their
    // part from an LDS copy of the part table.
    __host__ __device__ int part_words() const { return a.prices ? 2 * a.k : 0; }
    __device__ void stage(unsigned long long* t) const {
        for (int p = threadIdx.x; p < a.k; p += blockDim.x) {
            t[2 * p] = a.cap - a.loads[p];
            t[2 * p + 1] = (unsigned long long)__double_as_longlong(a.prices[p]);
        }
    }
    // estimated_repair_cost (src/refinement/incremental.cuh:156-159).
    __device__ double cost(const unsigned long long* part, int p, std::uint64_t w) const {
        const std::uint64_t slack = part ? part[2 * p] : a.cap - a.loads[p];
        const double price = part ? __longlong_as_double((long long)part[2 * p + 1]) : a.prices[p];
        return w > slack ? price * (double)(w - slack) : 0.0;
    }
    template <class Team>
    __device__ RowOut operator()(const Team& t, std::int32_t v, unsigned long long* hist) const {
        const int locked = a.locks[v];
        t.sync();
        if (t.leader()) {
            a.targets[v] = -1;
            set_gain(a, v, INT64_MIN);
            a.flags[v] = 0;
            a.locks[v] = 0;
        }
        if (locked) return RowOut{false, 0};
        const bool priced = a.prices != nullptr;
        const std::uint64_t w = a.vw[v];
        const int s = a.labels[v];
        Cand best{-1, 0, 0.0};
        std::uint64_t total, current;
        auto consider = [&](int p, std::uint64_t c) {
            if (p == s || !c) return;
            Cand x{p, c, priced ? (double)c - cost(t.part, p, w) : 0.0};
            if (cand_better(x, best, priced)) best = x;
        };
        if constexpr (IsLight<Team>::value) {
            LightRow<Team::size> r;
            Bin* lh = reinterpret_cast<Bin*>(hist);
            total = light_histogram(t, a, v, lh, r);
            current = lh[s];
            if (current != total)
#pragma unroll
                for (int j = 0; j < LightRow<Team::size>::kPer; ++j) {
                    const int l = r.lab(j);
                    if (l != 0xff) consider(l, lh[l]);
                }
            best = team_reduce(t, best, priced);
            light_clear(t, lh, r);
        } else {
            Bin* bh = reinterpret_cast<Bin*>(hist);
            total = team_histogram(t, a, v, bh);
            current = bh[s];
            if (current != total)
                for (int p = t.rank(); p < a.k; p += Team::size) consider(p, bh[p]);
            best = team_reduce(t, best, priced);
        }
        if (current == total) return RowOut{false, degree(v)};
        const bool accept = best.p >= 0 && (best.conn >= current ||
                                            current - best.conn < (std::uint64_t)(a.filter * (double)current));
        if (accept && t.leader()) {
            a.targets[v] = best.p;
            set_gain(a, v, best.conn >= current ? (std::int64_t)(best.conn - current) : -(std::int64_t)(current - best.conn));
        }
        return RowOut{accept, degree(v)};
    }
};

// Afterburner (incremental.cuh:361-438).  G: the level's gain word
// (std::int32_t: gains32, else gains).
template <class G>
struct AfterBody : RowBody {
    __device__ std::int64_t gain(std::int64_t u) const {
        if constexpr (sizeof(G) == 4) {
            const std::int32_t g = a.gains32[u];
            return g == INT32_MIN ? INT64_MIN : (std::int64_t)g;
        } else {
            return a.gains[u];
        }
    }
    template <class Team>
    __device__ RowOut operator()(const Team& t, std::int32_t v, unsigned long long*) const {
        const int s = a.labels[v], tv = a.targets[v];
        const std::int64_t gv = gain(v);
        std::int64_t corr = 0;
        // Neighbour state packed in one word: target (signed byte, -1 none)
        // in bits 0-7, label in bits 8-15.
        struct Rec {
            std::int32_t u;
            int tl;
            std::int64_t gu;
        };
        row_batches(
            t, a.off[v], a.off[v + 1],
            [&](std::int64_t x) {
                const std::int32_t u = a.nbr[(std::uint32_t)x];
                const std::uint32_t uu = (std::uint32_t)u;
                // The gain is gathered beside the target rather than after
                // it: one round trip per batch instead of two.
                const int tu = a.targets[uu];
                return Rec{u, (tu & 0xff) | ((int)a.labels[uu] << 8), gain(uu)};
            },
            [&](std::int64_t x, const Rec& r) {
                const int tu = (int)(std::int8_t)(r.tl & 0xff), lu = r.tl >> 8;
                if (r.u == v || tu < 0) return;
                if (!(r.gu > gv || (r.gu == gv && r.u < v))) return;
                corr += (std::int64_t)ew_at(a.ew, x) * ((tu == tv) - (tu == s) - (lu == tv) + (lu == s));
            });
        corr = t.sum(corr);
        bool keep;
        if (a.prices) {
            const std::uint64_t w = a.vw[v];
            const double adj = repair_cost(a, s, w) - repair_cost(a, tv, w);
            keep = (double)(gv + corr) + adj >= 0;
        } else
            keep = gv + corr >= 0;
        if (t.leader()) a.flags[v] = keep;
        return RowOut{keep, 0};
    }
};

// Exact cut delta (incremental.cuh:440-547): an undirected edge is owned by
// the endpoint that moves alone, or by the larger id when both move.
struct UpdateBody : RowBody {
    bool samples;
    template <class Team>
    __device__ RowOut operator()(const Team& t, std::int32_t v, unsigned long long*) const {
        const int s = a.labels[v], tv = a.targets[v];
        std::int64_t gain = 0;
        // Neighbour label and state word (two gathers per entry).
        struct Rec {
            std::int32_t u;
            int label;
            std::uint32_t st;
        };
        row_batches(
            t, a.off[v], a.off[v + 1],
            [&](std::int64_t x) {
                const std::int32_t u = a.nbr[(std::uint32_t)x];
                const std::uint32_t uu = (std::uint32_t)u;
                return Rec{u, (int)a.labels[uu], a.state[uu]};
            },
            [&](std::int64_t x, const Rec& r) {
                if (r.u == v) return;
                const int old_u = r.label, tu = (int)(std::int8_t)((r.st >> (8 * kStateTarget)) & 0xff);
                const bool moves = ((r.st >> (8 * kStateFlag)) & 0xff) != 0;
                // Most neighbours are already in the frontier: skip the
                // scattered byte store for them.
                if (!((r.st >> (8 * kStateMember)) & 0xff)) a.member[r.u] = 1;
                const int new_u = moves ? tu : old_u;
                if (!moves || r.u > v) gain += (std::int64_t)ew_at(a.ew, x) * ((s != old_u) - (tv != new_u));
            });
        gain = t.sum(gain);
        if (samples && a.samples && t.leader()) {
            atomicAdd(&a.samples[s].loss, (unsigned long long)(-gain));
            atomicAdd(&a.samples[s].weight, (unsigned long long)a.vw[v]);
        }
        return RowOut{false, gain};
    }
};

// Boundary test for the frontier (incremental.cuh:214-225).
struct BoundaryBody : RowBody {
    template <class Team>
    __device__ RowOut operator()(const Team& t, std::int32_t v, unsigned long long*) const {
        const int s = a.labels[v];
        // Batches of kBatch entries per lane, with a team-wide test after
        // each: most rows of a cut graph meet another part in their first
        // batch and stop there.  The loop bound is the same for the whole
        // team, so every lane takes part in each test.
        const std::int64_t b = a.off[v], e = a.off[v + 1];
        int any = 0;
        for (std::int64_t x0 = b; x0 < e && !any; x0 += kBatch * Team::size) {
            bool other[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                const std::int64_t x = x0 + t.rank() + j * Team::size;
                other[j] = false;
                if (x < e) {
                    const std::int32_t u = a.nbr[(std::uint32_t)x];
                    other[j] = u != v && a.labels[(std::uint32_t)u] != s;
                }
            }
            int mine = 0;
#pragma unroll
            for (int j = 0; j < kBatch; ++j) mine |= other[j];
            any = t.sum(mine);
        }
        if (t.leader()) {
            a.member[v] = any != 0;
            a.targets[v] = -1;
            set_gain(a, v, INT64_MIN);
            a.flags[v] = 0;
            a.locks[v] = 0;
        }
        return RowOut{false, 0};
    }
};

// incremental_commit (incremental.cuh:532-547) with per-block load deltas.
__global__ void k_commit(RefArgs a, const std::int32_t* list, const unsigned long long* count_ptr, bool lp) {
    __shared__ unsigned long long delta[kMaxParts];
    for (int p = threadIdx.x; p < a.k; p += blockDim.x) delta[p] = 0;
    __syncthreads();
    const std::int64_t count = (std::int64_t)*count_ptr;
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t v = list[i];
        const int s = a.labels[v], t = a.targets[v];
        const unsigned long long w = a.vw[v];
        atomicAdd(&delta[s], 0ULL - w);
        atomicAdd(&delta[t], w);
        a.labels[v] = (Label)t;
        if (lp) a.locks[v] = 1;
        a.member[v] = 1;
    }
    __syncthreads();
    for (int p = threadIdx.x; p < a.k; p += blockDim.x)
        if (delta[p]) atomicAdd(&a.loads[p], delta[p]);
}

// Sets byte lanes of every state word: state = (state & keep) | put.
__global__ void k_set_lanes(std::uint32_t* state, std::int64_t n, std::uint32_t keep, std::uint32_t put) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x)
        state[v] = (state[v] & keep) | put;
}

// Frontier list from member flags (order irrelevant: per-vertex decisions).
// Each wave reads kScanRounds x 64 consecutive flags before appending them.
constexpr int kScanRounds = 8;
// gate: null, or a device word; when it is zero the launch does nothing (a
// round of a device-controlled loop that has ended).
template <typename F>
__global__ void k_compact_flags(std::int64_t n, F flags, std::int32_t* out, unsigned long long* count,
                                const unsigned long long* gate) {
    if (gate && !*gate) return;
    __shared__ std::int32_t sb[kTierThreads / kWave][kStage];
    Stage st{sb[threadIdx.x / kWave], out, count, 0};
    const std::int64_t wave = (blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x) / kWave;
    const std::int64_t waves = (std::int64_t)gridDim.x * blockDim.x / kWave;
    for (std::int64_t base = wave * (kWave * kScanRounds); base < n; base += waves * (kWave * kScanRounds)) {
        bool f[kScanRounds];
#pragma unroll
        for (int r = 0; r < kScanRounds; ++r) {
            const std::int64_t v = base + r * kWave + lane_id();
            f[r] = v < n && flags[v] != 0;
        }
#pragma unroll
        for (int r = 0; r < kScanRounds; ++r) st.push(f[r], (std::int32_t)(base + r * kWave + lane_id()));
    }
    st.flush();
}

// ------------------------------------------------------------- repair
__device__ inline int gain_bucket(const RefArgs& a, double gain) {
    if (gain > 0) return 0;
    if (gain == 0) return 1;
    // First j in -23..23 with x <= 1.5^j gives bucket 25 + j (at least 2);
    // beyond the table, 49.  Equivalently 2 + #{table entries below x}.
    const double x = -gain;
    int lo = 0, hi = 47;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (kPow15.p[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return 2 + lo;
}

// Donor flags, resets (flags, bucket) on frozen loads.
// With counts set it also starts the repair iteration: the bucket weights
// cleared and loads0 = loads, as k_repair_begin does but for the counter
// block, which the caller knows to be zero.  The donor test then reads the
// part loads themselves, the values loads0 is given.
__global__ void k_repair_donors(RefArgs a, std::int32_t* donors, unsigned long long* count,
                                unsigned long long* loads0, unsigned long long* counts, std::int64_t ncounts) {
    __shared__ std::int32_t sb[kTierThreads / kWave][kStage];
    if (counts)
        for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < ncounts;
             i += (std::int64_t)gridDim.x * blockDim.x) {
            counts[i] = 0;
            if (i < a.k) loads0[i] = a.loads[i];
        }
    const unsigned long long* frozen = counts ? a.loads : a.loads0;
    Stage st{sb[threadIdx.x / kWave], donors, count, 0};
    const std::int64_t wave = (blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x) / kWave;
    const std::int64_t waves = (std::int64_t)gridDim.x * blockDim.x / kWave;
    for (std::int64_t base = wave * (kWave * kScanRounds); base < a.n; base += waves * (kWave * kScanRounds)) {
        int lab[kScanRounds];
#pragma unroll
        for (int r = 0; r < kScanRounds; ++r) {
            const std::int64_t v = base + r * kWave + lane_id();
            lab[r] = v < a.n ? (int)a.labels[v] : -1;
        }
#pragma unroll
        for (int r = 0; r < kScanRounds; ++r) {
            const std::int64_t v = base + r * kWave + lane_id();
            bool d = false;
            if (lab[r] >= 0) {
                a.flags[v] = 0;
                a.bucket[v] = -1;
                d = frozen[lab[r]] > a.cap;
            }
            st.push(d, (std::int32_t)v);
        }
    }
    st.flush();
}

struct RepairPick {
    int p;
    std::uint64_t conn, slack;
};
__device__ inline bool pick_better(const RepairPick& x, const RepairPick& y) {
    if (x.p < 0) return false;
    if (y.p < 0) return true;
    return x.conn > y.conn || (x.conn == y.conn && (x.slack > y.slack || (x.slack == y.slack && x.p < y.p)));
}
template <int W>
__device__ inline RepairPick pick_group_reduce(RepairPick c) {
    for (int d = W / 2; d; d >>= 1) {
        RepairPick o;
        o.p = __shfl_xor(c.p, d, W);
        o.conn = __shfl_xor(c.conn, d, W);
        o.slack = __shfl_xor(c.slack, d, W);
        if (pick_better(o, c)) c = o;
    }
    return c;
}
template <int W>
__device__ inline RepairPick team_reduce(const GroupTeam<W>&, RepairPick c) {
    return pick_group_reduce<W>(c);
}
__device__ inline RepairPick team_reduce(const BlockTeam& t, RepairPick c) {
    c = pick_group_reduce<kWave>(c);
    __syncthreads();
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        t.red[3 * w] = (unsigned long long)(long long)c.p;
        t.red[3 * w + 1] = c.conn;
        t.red[3 * w + 2] = c.slack;
    }
    __syncthreads();
    RepairPick r{(int)(long long)t.red[0], t.red[1], t.red[2]};
    for (int i = 1; i < BlockTeam::size / kWave; ++i) {
        RepairPick o{(int)(long long)t.red[3 * i], t.red[3 * i + 1], t.red[3 * i + 2]};
        if (pick_better(o, r)) r = o;
    }
    __syncthreads();
    return r;
}

// enumerate_repair_destination (rebalance_select.cuh:8-75) on frozen loads,
// then the gain bucket (incremental.cuh:559-626).
// load0: the frozen part loads (the block's LDS copy where it has one).
// Block-level table of repair bucket sums in LDS: kAggSlots keys (u32 count
// slots, all ones when free) packed two per word, then kAggSlots u64 sums.
// Linear probing over at most kAggProbes slots; a sum that finds no slot goes
// to global memory directly.  Sums are order-independent, so the counts are
// the same with or without the table.
constexpr int kAggSlots = 512;
constexpr int kAggProbes = 8;
constexpr int kAggWords = kAggSlots / 2 + kAggSlots;

__device__ inline bool agg_add(unsigned long long* agg, std::uint32_t key, unsigned long long w) {
    std::uint32_t* keys = reinterpret_cast<std::uint32_t*>(agg);
    unsigned long long* sums = agg + kAggSlots / 2;
    std::uint32_t h = mix32(key) & (kAggSlots - 1);
    for (int probe = 0; probe < kAggProbes; ++probe) {
        const std::uint32_t old = atomicCAS(&keys[h], 0xffffffffu, key);
        if (old == 0xffffffffu || old == key) {
            atomicAdd(&sums[h], w);
            return true;
        }
        h = (h + 1) & (kAggSlots - 1);
    }
    return false;
}

__device__ inline void repair_finish(const RefArgs& a, std::int64_t v, int s, std::uint64_t w, std::uint64_t current,
                                     RepairPick best, const unsigned long long* load0, unsigned long long* agg) {
    if (best.p < 0) {
        // Slack-weighted hash routing in part order.
        std::uint64_t tot = 0;
        for (int p = 0; p < a.k; ++p) {
            const std::uint64_t l = load0[p];
            if (p != s && l < a.cap && w <= a.cap - l) {
                const std::uint64_t sl = a.cap - l;
                tot = ~0ULL - tot < sl ? ~0ULL : tot + sl;
            }
        }
        best.conn = 0;
        if (tot) {
            const std::uint32_t h = mix32((std::uint32_t)v);
            std::uint64_t ticket = tot <= 0xffffffffULL ? (std::uint64_t)(h % (std::uint32_t)tot) : (std::uint64_t)h % tot;
            for (int p = 0; p < a.k; ++p) {
                const std::uint64_t l = load0[p];
                if (p != s && l < a.cap && w <= a.cap - l) {
                    const std::uint64_t sl = a.cap - l;
                    if (ticket < sl) {
                        best.p = p;
                        break;
                    }
                    ticket -= sl;
                }
            }
        }
    }
    a.targets[v] = best.p;
    if (best.p < 0) return;
    const double gain = best.conn >= current ? (double)(best.conn - current) : -(double)(current - best.conn);
    const std::uint64_t excess = load0[s] - a.cap;
    const std::uint64_t relief = w < excess ? w : excess;
    const double gpw = gain / (double)relief;
    const int lb = gain_bucket(a, gpw) * kBalanceSections + (int)(mix32((std::uint32_t)v) % kBalanceSections);
    a.bucket[v] = s * kBalanceWidth + lb;
    const std::uint32_t slot = (std::uint32_t)(s * kBalanceWidth + count_slot(lb));
    if (!agg || !agg_add(agg, slot, (unsigned long long)w)) atomicAdd(&a.counts[slot], (unsigned long long)w);
}

template <class Bin>
struct RepairProposeBody : RowBody {
    using LightBin = Bin;
    // Frozen part loads in LDS for the destination scan, then the block's
    // table of bucket sums: the donors of one part mostly share a few
    // buckets, so each block adds its sums to the global counts once.
    __host__ __device__ int part_words() const { return a.k + kAggWords; }
    __device__ void stage(unsigned long long* t) const {
        for (int p = threadIdx.x; p < a.k; p += blockDim.x) t[p] = a.loads0[p];
        for (int i = threadIdx.x; i < kAggWords; i += blockDim.x) t[a.k + i] = i < kAggSlots / 2 ? ~0ULL : 0ULL;
    }
    __device__ void flush(const unsigned long long* t) const {
        const std::uint32_t* keys = reinterpret_cast<const std::uint32_t*>(t + a.k);
        const unsigned long long* sums = t + a.k + kAggSlots / 2;
        for (int i = threadIdx.x; i < kAggSlots; i += blockDim.x)
            if (keys[i] != 0xffffffffu && sums[i]) atomicAdd(&a.counts[keys[i]], sums[i]);
    }
    template <class Team>
    __device__ RowOut operator()(const Team& t, std::int32_t v, unsigned long long* hist) const {
        const int s = a.labels[v];
        const std::uint64_t w = a.vw[v];
        RepairPick best{-1, 0, 0};
        const unsigned long long* load0 = t.part ? t.part : a.loads0;
        auto consider = [&](int p, std::uint64_t c) {
            if (p == s || !c || load0[p] >= a.cap || w > a.cap - load0[p]) return;
            RepairPick x{p, c, a.cap - load0[p]};
            if (pick_better(x, best)) best = x;
        };
        std::uint64_t total, current;
        if constexpr (IsLight<Team>::value) {
            LightRow<Team::size> r;
            Bin* lh = reinterpret_cast<Bin*>(hist);
            total = light_histogram(t, a, v, lh, r);
            current = lh[s];
            if (current != total)
#pragma unroll
                for (int j = 0; j < LightRow<Team::size>::kPer; ++j) {
                    const int l = r.lab(j);
                    if (l != 0xff) consider(l, lh[l]);
                }
            best = team_reduce(t, best);
            light_clear(t, lh, r);
        } else {
            Bin* bh = reinterpret_cast<Bin*>(hist);
            total = team_histogram(t, a, v, bh);
            current = bh[s];
            if (current != total)
                for (int p = t.rank(); p < a.k; p += Team::size) consider(p, bh[p]);
            best = team_reduce(t, best);
        }
        unsigned long long* agg = t.part ? const_cast<unsigned long long*>(t.part) + a.k : nullptr;
        if (t.leader()) repair_finish(a, v, s, w, current, best, load0, agg);
        return RowOut{false, 0};
    }
};

// Per overloaded part: exclusive prefix over the 800 buckets and the bucket
// that contains the excess threshold (balance_scan_buckets, incremental.cuh:628-652).
// One block per part; the first bucket whose range contains the excess is
// the minimum over buckets that satisfy the test.
__global__ void __launch_bounds__(kScanThreads) k_bucket_scan(RefArgs a, unsigned long long* prefix,
                                                             std::int32_t* bstar, unsigned long long* excess) {
    __shared__ unsigned long long wsum[kScanThreads / kWave];
    __shared__ int star;
    const int p = blockIdx.x;
    if (!(a.loads0[p] > a.cap)) {
        if (threadIdx.x == 0) {
            bstar[p] = -1;
            excess[p] = 0;
        }
        return;
    }
    const unsigned long long ex = a.loads0[p] - a.cap;
    if (threadIdx.x == 0) star = kBalanceWidth;
    __syncthreads();
    unsigned long long carry = 0;
    for (int b0 = 0; b0 < kBalanceWidth; b0 += kScanThreads) {
        const int b = b0 + threadIdx.x;
        const unsigned long long cnt = b < kBalanceWidth ? a.counts[p * kBalanceWidth + count_slot(b)] : 0;
        unsigned long long inc, exc, tot;
        block_scan<PlusOp>(cnt, wsum, inc, exc, tot);
        const unsigned long long sum = carry + exc;
        if (b < kBalanceWidth) {
            prefix[p * kBalanceWidth + b] = sum;
            if (cnt && sum < ex && ex <= sum + cnt) atomicMin(&star, b);
        }
        carry += tot;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bstar[p] = star;
        excess[p] = ex;
    }
}

__global__ void k_repair_flag(RefArgs a, const std::int32_t* bstar, std::uint8_t* f) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < a.n;
         v += (std::int64_t)gridDim.x * blockDim.x) {
        const int b = a.bucket[v];
        f[v] = b >= 0 && (b % kBalanceWidth) <= bstar[b / kBalanceWidth];
    }
}

// Position of the calling lane among the team's lanes with pred set, and
// their number (every lane of the team calls it).
template <int W>
__device__ inline int team_ballot_prefix(const GroupTeam<W>& t, bool pred, int& total) {
    const unsigned long long m = __ballot(pred);
    const unsigned long long bits = W == kWave ? m : (m >> (lane_id() & ~(W - 1))) & ((1ULL << W) - 1ULL);
    total = __popcll(bits);
    return __popcll(bits & ((1ULL << t.lg) - 1ULL));
}
__device__ inline int team_ballot_prefix(const BlockTeam& t, bool pred, int& total) {
    constexpr int kWaves = BlockTeam::size / kWave;
    const unsigned long long m = __ballot(pred);
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if (lane_id() == 0) t.red[w] = (unsigned long long)__popcll(m);
    __syncthreads();
    int before = 0;
    total = 0;
    for (int i = 0; i < kWaves; ++i) {
        if (i < w) before += (int)t.red[i];
        total += (int)t.red[i];
    }
    __syncthreads();
    return before + __popcll(m & ((1ULL << lane_id()) - 1ULL));
}

// Sparse connection rows: the parts a listed donor connects to, in part
// order, as (weight, part) pairs in slots [0, nnz) of its k slots, by the
// row tiers over the positions [0, ne) of the donor list E (the tier
// kernels hand the position over as the row id): short rows take 4-lane
// teams, hub rows whole blocks; team ballots give each nonzero part's slot.
template <class Bin>
struct ConnBody : RowBody {
    using LightBin = Bin;
    const std::int32_t* E;
    const std::uint8_t* skip;  // null, or nonzero for donors whose row is not needed
    unsigned long long* vals;
    std::uint8_t* parts;
    std::uint8_t* nnz;
    __device__ std::int64_t degree(std::int32_t i) const {
        const std::int32_t v = E[i];
        return a.off[v + 1] - a.off[v];
    }
    template <class Team>
    __device__ void emit(const Team& t, std::int64_t i, const Bin* bh) const {
        int base = 0;
        for (int p0 = 0; p0 < a.k; p0 += Team::size) {
            const int p = p0 + t.rank();
            const unsigned long long c = p < a.k ? (unsigned long long)bh[p] : 0ULL;
            int cnt = 0;
            const int pos = team_ballot_prefix(t, c != 0, cnt);
            if (c) {
                const std::int64_t j = i * a.k + base + pos;
                vals[j] = c;
                parts[j] = static_cast<std::uint8_t>(p);
            }
            base += cnt;
        }
        if (t.leader()) nnz[i] = static_cast<std::uint8_t>(base);
    }
    template <class Team>
    __device__ RowOut operator()(const Team& t, std::int32_t i, unsigned long long* hist) const {
        if (skip && skip[i]) return RowOut{false, 0};
        const std::int32_t v = E[i];
        Bin* bh = reinterpret_cast<Bin*>(hist);
        if constexpr (IsLight<Team>::value) {
            LightRow<Team::size> r;
            light_histogram(t, a, v, bh, r);
            emit(t, i, bh);
            light_clear(t, bh, r);
        } else {
            team_histogram(t, a, v, bh);
            emit(t, i, bh);
        }
        return RowOut{false, 0};
    }
};


// Receiver reservation (rebalance_graph.cuh:86-156) in canonical order:
// requesters of one part are ranked by vertex id on top of the part's
// request total at the start of the attempt.  A rank is an exclusive
// weighted prefix among earlier items with the same key, computed without
// sorting: per-tile key sums, a per-key scan over tiles, then in-tile
// prefixes from wave scans per distinct key.
constexpr int kKeyTile = 1024;
constexpr int kKeyRounds = kKeyTile / kTierThreads;

// Exclusive weighted prefix among lanes with the same key (lower lanes
// first); wt[key] receives the wave's total for every key present.  key < 0
// marks a lane without a key.  All lanes of the wave call it.
__device__ inline std::uint64_t wave_keyed_exclusive(int key, std::uint64_t w, unsigned long long* wt) {
    const int lane = lane_id();
    unsigned long long pending = __ballot(key >= 0);
    std::uint64_t excl = 0;
    while (pending) {
        const int leader = __ffsll(pending) - 1;
        const int kk = __shfl(key, leader);
        const bool mine = key == kk;
        const unsigned long long peers = __ballot(mine);
        std::uint64_t x = mine ? w : 0;
        for (int d = 1; d < kWave; d <<= 1) {
            const std::uint64_t y = __shfl_up(x, d);
            if (lane >= d) x += y;
        }
        if (mine) excl = x - w;
        const std::uint64_t total = __shfl(x, kWave - 1);
        if (lane == 0) wt[kk] = total;
        pending &= ~peers;
    }
    return excl;
}

// Per-key exclusive scan over tiles (one wave per key); with requests,
// ranks start at the request totals and the totals advance by this
// attempt's requests.
constexpr int kKeyScanThreads = 1024;

__global__ void __launch_bounds__(kKeyScanThreads) k_keyed_scan(unsigned long long* S, int nb, int k,
                                                               unsigned long long* requests) {
    const int lane = lane_id();
    for (int c = threadIdx.x / kWave; c < k; c += kKeyScanThreads / kWave) {
        unsigned long long carry = requests ? requests[c] : 0;
        for (int b0 = 0; b0 < nb; b0 += kWave) {
            const int b = b0 + lane;
            const unsigned long long own = b < nb ? S[(std::int64_t)c * nb + b] : 0;
            unsigned long long x = own;
            for (int d = 1; d < kWave; d <<= 1) {
                const unsigned long long y = __shfl_up(x, d);
                if (lane >= d) x += y;
            }
            if (b < nb) S[(std::int64_t)c * nb + b] = carry + x - own;
            carry += __shfl(x, kWave - 1);
        }
        if (requests && lane == 0) requests[c] = carry;
    }
}

// In-tile ranks in index order; f(i, key, weight, rank) returns whether i
// is appended to the stage list.
template <class F>
__global__ void __launch_bounds__(kTierThreads) k_keyed_apply(const std::int32_t* keys, std::int64_t n,
                                                             const unsigned long long* S, int nb, F f,
                                                             std::int32_t* list, unsigned long long* count) {
    __shared__ unsigned long long wt[kTierThreads / kWave][kMaxParts];
    __shared__ unsigned long long run[kMaxParts];
    __shared__ std::int32_t sb[kTierThreads / kWave][kStage];
    const int w = threadIdx.x / kWave;
    Stage st{sb[w], list, count, 0};
    for (int c = threadIdx.x; c < kMaxParts; c += kTierThreads) {
        run[c] = 0;
        for (int i = 0; i < kTierThreads / kWave; ++i) wt[i][c] = 0;
    }
    __syncthreads();
    const std::int64_t base = blockIdx.x * (std::int64_t)kKeyTile;
    for (int r = 0; r < kKeyRounds; ++r) {
        const std::int64_t i = base + r * kTierThreads + threadIdx.x;
        const int key = i < n ? keys[i] : -1;
        const std::uint64_t wgt = key >= 0 ? f.weight(i) : 0;
        const std::uint64_t ex = wave_keyed_exclusive(key, wgt, wt[w]);
        __syncthreads();
        bool append = false;
        if (key >= 0) {
            std::uint64_t rank = S[(std::int64_t)key * nb + blockIdx.x] + run[key] + ex;
            for (int j = 0; j < w; ++j) rank += wt[j][key];
            append = f(i, key, wgt, rank);
        }
        st.push(append, append ? f.vertex(i) : 0);
        __syncthreads();
        for (int c = threadIdx.x; c < kMaxParts; c += kTierThreads) {
            unsigned long long add = 0;
            for (int j = 0; j < kTierThreads / kWave; ++j) {
                add += wt[j][c];
                wt[j][c] = 0;
            }
            run[c] += add;
        }
        __syncthreads();
    }
    st.flush();
}

// Per-tile key weight sums (S is key-major: S[key * nb + tile]).
template <class F>
__device__ inline void tile_key_sums(const std::int32_t* keys, std::int64_t n, int k, int nb, const F& f,
                                     unsigned long long* acc, unsigned long long* S) {
    __syncthreads();
    const std::int64_t base = blockIdx.x * (std::int64_t)kKeyTile;
    for (int r = 0; r < kKeyRounds; ++r) {
        const std::int64_t i = base + r * kTierThreads + threadIdx.x;
        if (i < n && keys[i] >= 0) atomicAdd(&acc[keys[i]], (unsigned long long)f.weight(i));
    }
    __syncthreads();
    for (int c = threadIdx.x; c < k; c += kTierThreads) S[(std::int64_t)c * nb + blockIdx.x] = acc[c];
}

// Eligibility (rebalance_graph.cuh:86-110): donors below the boundary
// bucket are eligible; inside it, while their rank stays below the excess.
struct EligFn {
    RefArgs a;
    const std::int32_t* L;
    const unsigned long long* prefix;
    const unsigned long long* excess;
    std::uint8_t* elig;
    __device__ std::uint64_t weight(std::int64_t i) const { return a.vw[L[i]]; }
    __device__ std::int32_t vertex(std::int64_t i) const { return L[i]; }
    __device__ bool operator()(std::int64_t i, int key, std::uint64_t, std::uint64_t rank) const {
        elig[i] = prefix[a.bucket[L[i]]] + rank < excess[key];
        return false;
    }
};

__global__ void __launch_bounds__(kTierThreads) k_elig_keys(RefArgs a, const std::int32_t* L, std::int64_t nl,
                                                           const std::int32_t* bstar, int nb, std::int32_t* keys,
                                                           std::uint8_t* elig, unsigned long long* S) {
    __shared__ unsigned long long acc[kMaxParts];
    for (int c = threadIdx.x; c < kMaxParts; c += kTierThreads) acc[c] = 0;
    const std::int64_t base = blockIdx.x * (std::int64_t)kKeyTile;
    for (int r = 0; r < kKeyRounds; ++r) {
        const std::int64_t i = base + r * kTierThreads + threadIdx.x;
        if (i >= nl) break;
        const int b = a.bucket[L[i]];
        const int s = b / kBalanceWidth, lb = b % kBalanceWidth;
        keys[i] = lb == bstar[s] ? s : -1;
        elig[i] = lb < bstar[s];
    }
    EligFn f{a, L, nullptr, nullptr, nullptr};
    tile_key_sums(keys, nl, a.k, nb, f, acc, S);
}

// Donor selection of a repair iteration on a level of at most
// kSelectSmallMax vertices, in one block: the flagged donors (bucket at or
// below the part's boundary bucket) in vertex order, then their eligibility
// (below the boundary bucket, or inside it while the weighted rank among
// earlier donors of that bucket stays under the excess), compacted into E
// with their number in *ne.  One pass in index order gives the same ranks
// as the tiled form (per-tile sums, a scan over tiles, in-tile ranks), which
// larger levels use.
constexpr int kDonorThreads = 1024;
constexpr int kDonorWaves = kDonorThreads / kWave;

__global__ void __launch_bounds__(kDonorThreads) k_donors_small(RefArgs a, const std::int32_t* bstar,
                                                               const unsigned long long* prefix,
                                                               const unsigned long long* excess, std::int32_t* E,
                                                               std::int64_t* ne) {
    __shared__ std::int32_t L[kSelectSmallMax];
    __shared__ unsigned long long wt[kDonorWaves][kMaxParts];
    __shared__ unsigned long long run[kMaxParts];
    __shared__ int wsum[kDonorWaves];
    __shared__ int base;
    const int t = threadIdx.x, w = t / kWave, lane = lane_id();
    const unsigned long long lt = (1ULL << lane) - 1ULL;
    for (int c = t; c < kMaxParts; c += kDonorThreads) {
        run[c] = 0;
        for (int j = 0; j < kDonorWaves; ++j) wt[j][c] = 0;
    }
    if (t == 0) base = 0;
    __syncthreads();
    // Stable compaction of `mine` over one round: writes dst[base + rank].
    auto compact = [&](bool mine, std::int32_t value, std::int32_t* dst) {
        const unsigned long long m = __ballot(mine);
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int pre = base, total = 0;
        for (int j = 0; j < kDonorWaves; ++j) {
            pre += j < w ? wsum[j] : 0;
            total += wsum[j];
        }
        if (mine) dst[pre + __popcll(m & lt)] = value;
        __syncthreads();
        if (t == 0) base += total;
        __syncthreads();
    };
    for (std::int64_t r = 0; r < a.n; r += kDonorThreads) {
        const std::int64_t v = r + t;
        bool f = false;
        if (v < a.n) {
            const int b = a.bucket[v];
            f = b >= 0 && (b % kBalanceWidth) <= bstar[b / kBalanceWidth];
        }
        compact(f, static_cast<std::int32_t>(v), L);
    }
    const int nl = base;
    __syncthreads();
    if (t == 0) base = 0;
    __syncthreads();
    for (int r = 0; r < nl; r += kDonorThreads) {
        const int i = r + t;
        int key = -1, b = 0;
        bool el = false;
        std::int32_t v = 0;
        std::uint64_t wgt = 0;
        if (i < nl) {
            v = L[i];
            b = a.bucket[v];
            const int s = b / kBalanceWidth, lb = b % kBalanceWidth;
            if (lb == bstar[s]) {
                key = s;
                wgt = a.vw[v];
            }
            el = lb < bstar[s];
        }
        const std::uint64_t ex = wave_keyed_exclusive(key, wgt, wt[w]);
        __syncthreads();
        if (key >= 0) {
            std::uint64_t rank = run[key] + ex;
            for (int j = 0; j < w; ++j) rank += wt[j][key];
            el = prefix[b] + rank < excess[key];
        }
        __syncthreads();
        for (int c = t; c < kMaxParts; c += kDonorThreads) {
            unsigned long long add = 0;
            for (int j = 0; j < kDonorWaves; ++j) {
                add += wt[j][c];
                wt[j][c] = 0;
            }
            run[c] += add;
        }
        compact(el, v, E);
    }
    if (t == 0) *ne = base;
}

// A live donor as the attempts read it: index into the eligible list, the
// vertex, its weight and label (one 16-byte load per donor).
struct DonorInfo {
    std::int32_t e;
    std::int32_t v;
    std::uint32_t w;
    std::int32_t label;
};

struct ReserveState {
    const std::int32_t* E;           // eligible donors in vertex order
    std::int64_t ne;
    const unsigned long long* conn;  // ne x k sparse rows: weights
    const std::uint8_t* conn_part;   // ne x k: their parts
    const std::uint8_t* conn_nnz;    // ne: entries per row
    unsigned long long* requests;    // k, request totals
    const unsigned long long* slack; // k
    std::uint8_t* done;              // 1 admitted, 2 without a part left
    unsigned long long* tried;       // ne x 2
    const DonorInfo* live;           // donors still choosing, ascending
    std::int64_t nlive;
};

// Choices of one attempt, one 4-lane group per live donor (many donors in
// flight per wave: the per-donor loads form a dependent chain).  Attempt 0
// requests the proposed target; later attempts the untried part with the
// most connection, then the most room left at the start of the attempt,
// then the lowest id.  Room only shrinks (request totals only grow) and
// tried parts stay tried, so a donor without a choice in some attempt has
// none in any later one: it is marked done (2) and leaves the live list.
struct Choice {
    int p;
    unsigned long long conn, room;
};
__device__ inline bool choice_better(const Choice& x, const Choice& y) {
    if (x.p < 0) return false;
    if (y.p < 0) return true;
    return x.conn > y.conn || (x.conn == y.conn && (x.room > y.room || (x.room == y.room && x.p < y.p)));
}

// A part without connection to the donor (conn 0) loses to every connected
// one, and among unconnected parts the order is (room desc, id asc): the
// scan walks the connected parts of the sparse row, and only when none is
// open falls back to the first untried part in that block-wide order (open
// when its room admits the donor; rooms after it are no larger).
constexpr int kChooseLanes = 4;
__global__ void __launch_bounds__(kTierThreads) k_choose(RefArgs a, ReserveState r, int attempt, std::int32_t* keys,
                                                        unsigned long long* active) {
    constexpr int W = kChooseLanes;
    __shared__ unsigned long long room[kMaxParts];  // start-of-attempt room per part
    __shared__ std::uint8_t order[kMaxParts];       // parts by (room desc, id asc)
    const int lg = threadIdx.x % W;
    const int gbase = (threadIdx.x % kWave) / W * W;
    if (attempt > 0) {
        for (int p = threadIdx.x; p < a.k; p += kTierThreads)
            room[p] = r.requests[p] < r.slack[p] ? r.slack[p] - r.requests[p] : 0;
        __syncthreads();
        for (int p = threadIdx.x; p < a.k; p += kTierThreads) {
            int rank = 0;
            for (int q = 0; q < a.k; ++q) rank += room[q] > room[p] || (room[q] == room[p] && q < p);
            order[rank] = static_cast<std::uint8_t>(p);
        }
        __syncthreads();
    }
    const std::int64_t groups = (std::int64_t)gridDim.x * (kTierThreads / W);
    unsigned long long cnt = 0;
    for (std::int64_t i = (blockIdx.x * (std::int64_t)kTierThreads + threadIdx.x) / W; i < r.nlive; i += groups) {
        const DonorInfo d = r.live[i];
        const std::int64_t e = d.e;
        Choice best{-1, 0, 0};
        bool open = !r.done[e];
        if (open) {
            const std::int32_t v = d.v;
            if (attempt == 0)
                best.p = a.targets[v];
            else {
                const std::uint64_t w = d.w;
                const int lv = d.label;
                const unsigned long long tried0 = r.tried[2 * e], tried1 = r.tried[2 * e + 1];
                auto closed = [&](int p) { return p == lv || ((((p >> 6) ? tried1 : tried0) >> (p & 63)) & 1ULL); };
                const int cnt = r.conn_nnz[e];
                for (int j = lg; j < cnt; j += W) {
                    const int p = r.conn_part[e * a.k + j];
                    if (closed(p)) continue;
                    const unsigned long long avail = room[p];
                    if (avail < w) continue;
                    const Choice x{p, r.conn[e * a.k + j], avail};
                    if (choice_better(x, best)) best = x;
                }
                for (int d = W / 2; d; d >>= 1) {
                    const Choice o{__shfl_xor(best.p, d, W), __shfl_xor(best.conn, d, W), __shfl_xor(best.room, d, W)};
                    if (choice_better(o, best)) best = o;
                }
                if (best.p < 0) {
                    for (int q0 = 0; q0 < a.k; q0 += W) {
                        const int q = q0 + lg;
                        const bool open = q < a.k && !closed(order[q]);
                        const unsigned sub = static_cast<unsigned>((__ballot(open) >> gbase) & ((1ULL << W) - 1));
                        if (sub) {
                            const int p = order[q0 + __ffs(sub) - 1];
                            if (room[p] >= w) best = Choice{p, 0, room[p]};
                            break;
                        }
                    }
                }
            }
        }
        if (lg == 0) {
            keys[i] = best.p;
            cnt += best.p >= 0;
            if (open && attempt > 0 && best.p < 0) r.done[e] = 2;
        }
    }
    cnt = wave_sum(cnt);
    if (lane_id() == 0 && cnt) atomicAdd(active, cnt);
}

struct DonorWeight {
    const DonorInfo* live;
    __device__ std::uint64_t weight(std::int64_t i) const { return live[i].w; }
};

template <class F>
__global__ void __launch_bounds__(kTierThreads) k_keyed_sums(const std::int32_t* keys, std::int64_t n, int k, int nb,
                                                            F f, unsigned long long* S) {
    __shared__ unsigned long long acc[kMaxParts];
    for (int c = threadIdx.x; c < kMaxParts; c += kTierThreads) acc[c] = 0;
    tile_key_sums(keys, n, k, nb, f, acc, S);
}

// Admission: the request fits when it ends within the receiver's slack.
struct AdmitFn {
    RefArgs a;
    ReserveState r;
    __device__ std::uint64_t weight(std::int64_t i) const { return r.live[i].w; }
    __device__ std::int32_t vertex(std::int64_t i) const { return r.live[i].v; }
    __device__ bool operator()(std::int64_t i, int c, std::uint64_t w, std::uint64_t begin) const {
        const std::int64_t e = r.live[i].e;
        r.tried[2 * e + (c >> 6)] |= 1ULL << (c & 63);
        if (begin <= r.slack[c] && w <= r.slack[c] - begin) {
            const std::int32_t v = r.live[i].v;
            a.targets[v] = c;
            a.flags[v] = 1;
            r.done[e] = 1;
            return true;
        }
        return false;
    }
};

// The remaining attempts of a reservation whose live donors fit one block
// (at most kTailDonors, in rounds of one donor per thread), looped inside one launch: per
// attempt the same start-of-attempt room and part order, the same choice
// per donor (attempt 0: the proposed target), the same ranks (exclusive
// weight prefix per requested part in donor order on top of the part's
// request total) and admissions as the per-attempt launches, and the loop
// stops after an attempt with no chooser.  Connection rows must be built
// before the launch.
constexpr int kTailThreads = 1024;
constexpr int kTailDonors = 4096;  // donors per launch: rounds of kTailThreads
__global__ void __launch_bounds__(kTailThreads) k_reserve_tail(RefArgs a, ReserveState r, int attempt0,
                                                               std::int32_t* movers, unsigned long long* mcount) {
    constexpr int kWaves = kTailThreads / kWave;
    __shared__ unsigned long long room[kMaxParts];
    __shared__ unsigned long long run[kMaxParts];  // this attempt's requests of earlier rounds
    __shared__ std::uint8_t order[kMaxParts];
    __shared__ unsigned long long wt[kWaves][kMaxParts];
    __shared__ std::int32_t sb[kWaves][kStage];
    // Positions in r.live of the donors still choosing, ascending: donors
    // that were admitted or found no part are dropped after each attempt
    // (they would choose nothing again), so later attempts visit only the
    // waiting ones.  Ranks are weight prefixes over choosing donors in
    // list order, which dropping non-choosers leaves unchanged.
    __shared__ std::uint16_t pos[kTailDonors];
    __shared__ int wcount[kWaves];
    __shared__ int active;
    const int t = threadIdx.x, w = t / kWave;
    Stage st{sb[w], movers, mcount, 0};
    for (int i = t; i < kWaves * kMaxParts; i += kTailThreads) wt[i / kMaxParts][i % kMaxParts] = 0;
    int nl = static_cast<int>(r.nlive);
    for (int i = t; i < nl; i += kTailThreads) pos[i] = static_cast<std::uint16_t>(i);
    for (int attempt = attempt0; attempt < a.k; ++attempt) {
        for (int p = t; p < a.k; p += kTailThreads) {
            room[p] = r.requests[p] < r.slack[p] ? r.slack[p] - r.requests[p] : 0;
            run[p] = 0;
        }
        if (t == 0) active = 0;
        __syncthreads();
        for (int p = t; p < a.k; p += kTailThreads) {
            int rank = 0;
            for (int q = 0; q < a.k; ++q) rank += room[q] > room[p] || (room[q] == room[p] && q < p);
            order[rank] = static_cast<std::uint8_t>(p);
        }
        __syncthreads();
        for (int base = 0; base < nl; base += kTailThreads) {
            const int i = base + t;
            int key = -1;
            DonorInfo d{-1, 0, 0, 0};
            if (i < nl) {
                d = r.live[pos[i]];
                const std::int64_t e = d.e;
                if (!r.done[e] && attempt == 0) {
                    key = a.targets[d.v];
                } else if (!r.done[e]) {
                    const std::uint64_t wd = d.w;
                    const int lv = d.label;
                    const unsigned long long tried0 = r.tried[2 * e], tried1 = r.tried[2 * e + 1];
                    auto closed = [&](int p) {
                        return p == lv || ((((p >> 6) ? tried1 : tried0) >> (p & 63)) & 1ULL);
                    };
                    Choice best{-1, 0, 0};
                    const int cnt = r.conn_nnz[e];
                    for (int j = 0; j < cnt; ++j) {
                        const int p = r.conn_part[e * a.k + j];
                        if (closed(p) || room[p] < wd) continue;
                        const Choice x{p, r.conn[e * a.k + j], room[p]};
                        if (choice_better(x, best)) best = x;
                    }
                    if (best.p < 0)
                        for (int q = 0; q < a.k; ++q) {
                            const int p = order[q];
                            if (closed(p)) continue;
                            if (room[p] >= wd) best = Choice{p, 0, room[p]};
                            break;
                        }
                    if (best.p < 0) r.done[e] = 2;
                    key = best.p;
                }
            }
            const std::uint64_t wgt = key >= 0 ? d.w : 0;
            const std::uint64_t ex = wave_keyed_exclusive(key, wgt, wt[w]);
            __syncthreads();
            bool admit = false;
            if (key >= 0) {
                atomicAdd(&active, 1);
                std::uint64_t rank = r.requests[key] + run[key] + ex;
                for (int j = 0; j < w; ++j) rank += wt[j][key];
                const std::int64_t e = d.e;
                r.tried[2 * e + (key >> 6)] |= 1ULL << (key & 63);
                if (rank <= r.slack[key] && wgt <= r.slack[key] - rank) {
                    a.targets[d.v] = key;
                    a.flags[d.v] = 1;
                    r.done[e] = 1;
                    admit = true;
                }
            }
            st.push(admit, d.v);
            __syncthreads();
            for (int p = t; p < a.k; p += kTailThreads) {
                unsigned long long add = 0;
                for (int j = 0; j < kWaves; ++j) {
                    add += wt[j][p];
                    wt[j][p] = 0;
                }
                run[p] += add;
            }
            __syncthreads();
        }
        const int act = active;
        for (int p = t; p < a.k; p += kTailThreads) r.requests[p] += run[p];
        __syncthreads();
        if (act == 0) break;
        // Stable compaction of the waiting donors: thread t keeps its
        // kPer consecutive positions in registers, a block scan of the
        // counts gives where they go.
        constexpr int kPer = kTailDonors / kTailThreads;
        std::uint16_t q[kPer];
        bool waiting[kPer];
        int keep = 0;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int i = t * kPer + j;
            q[j] = i < nl ? pos[i] : 0;
            waiting[j] = i < nl && !r.done[r.live[q[j]].e];
            keep += waiting[j];
        }
        const int inc = wave_inclusive_scan<PlusOp>(keep);
        if (lane_id() == kWave - 1) wcount[w] = inc;
        __syncthreads();  // every position has been read
        int at = inc - keep, total = 0;
        for (int j = 0; j < kWaves; ++j) {
            if (j < w) at += wcount[j];
            total += wcount[j];
        }
#pragma unroll
        for (int j = 0; j < kPer; ++j)
            if (waiting[j]) pos[at++] = q[j];
        nl = total;
        __syncthreads();
        if (nl == 0) break;
    }
    st.flush();
}

__global__ void k_live_flags(const DonorInfo* live, std::int64_t count, const std::uint8_t* done, std::uint8_t* flags) {
    for (std::int64_t j = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; j < count;
         j += (std::int64_t)gridDim.x * blockDim.x)
        flags[j] = done[live[j].e] == 0;
}

__global__ void k_gather_info(const DonorInfo* src, const std::int32_t* idx, std::int64_t count, DonorInfo* dst) {
    for (std::int64_t j = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; j < count;
         j += (std::int64_t)gridDim.x * blockDim.x)
        dst[j] = src[idx[j]];
}

// Set-up of a reservation in one launch: the eligible donors E[j] =
// L[idx[j]] in vertex order (E already holds them when idx is null; then
// L is E) with their packed records, cleared per-donor
// state (done, tried bits) and each part's room and request total.
__global__ void k_reserve_setup(RefArgs a, const std::int32_t* L, const std::int32_t* idx, std::int64_t ne,
                                std::int32_t* E, DonorInfo* live, std::uint8_t* done, unsigned long long* tried,
                                unsigned long long* slack, unsigned long long* requests) {
    for (std::int64_t j = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; j < ne;
         j += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t v = idx ? L[idx[j]] : L[j];
        if (idx) E[j] = v;
        live[j] = DonorInfo{static_cast<std::int32_t>(j), v, a.vw[v], (std::int32_t)a.labels[v]};
        done[j] = 0;
        tried[2 * j] = 0;
        tried[2 * j + 1] = 0;
    }
    if (blockIdx.x == 0)
        for (int p = threadIdx.x; p < a.k; p += blockDim.x) {
            slack[p] = a.loads0[p] < a.cap ? a.cap - a.loads0[p] : 0;
            requests[p] = 0;
        }
}

// Start of a repair iteration in one launch: freezes the part loads
// (loads0 = loads), clears the bucket weights and the counter block.
__global__ void k_repair_begin(RefArgs a, unsigned long long* loads0, unsigned long long* counts, std::int64_t ncounts,
                               unsigned long long* ctr, int nctr) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < ncounts;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        counts[i] = 0;
        if (i < a.k) loads0[i] = a.loads[i];
        if (i < nctr) ctr[i] = 0;
    }
}

// New repair prices from the samples of the last repair (parts without a
// sample keep their price): loss / weight for a positive loss, else 0.  The
// samples are cleared for the next repair.
__global__ void k_publish_prices(RepairSample* samples, double* prices, int k) {
    for (int p = threadIdx.x; p < k; p += blockDim.x) {
        const RepairSample x = samples[p];
        samples[p] = RepairSample{0, 0};
        if (!x.weight) continue;
        const long long loss = (long long)x.loss;
        prices[p] = loss > 0 ? (double)loss / (double)x.weight : 0.0;
    }
}

// ------------------------------------------------- device LP loop control
// The LP round loop of one level decided on the device (the rules of the
// host loop, incremental.cuh:911-961, which lp_graph.cuh:57-92 records as a
// GPU-controlled loop), so the host can issue rounds ahead instead of reading
// counters back after each.  Its words sit behind the part loads in the
// counter block; the best loads (k words) follow them.
enum LoopWord {
    kRun = 0,      // 1 while rounds run; a round issued after the stop does nothing
    kPhase,        // rounds run
    kStale,
    kEmpty,
    kCut,
    kBestCut,
    kLpMoves,
    kReason,       // LoopReason of the stop
    kImproved,     // the last round set a new best cut: checkpoint its labels
    kAccFrontier,  // profiled runs: rows and entries of the three sweeps, summed
    kAccRowSum,
    kAccAfterE,
    kAccAfterR,
    kAccUpdateE,
    kAccUpdateR,
    kLoopWords
};
enum LoopReason { kRunning = 0, kEnded = 1, kNeedRepair = 2, kCutUnderflow = 3, kDeferred = 4 };
// Words of the loop state mirrored into the host mailbox after each round.
constexpr int kMailWords = kImproved + 1;

struct LoopParams {
    int k;
    int rounds;
    int stale_rounds;
    int graph_mode;  // stale rule in integers (graph loop) or long double
    int light_only;  // no medium launch: a deferred row is an error
    int profiling;
    std::uint64_t cap;
    std::uint64_t m0999;  // 64-bit significand of the long double 0.999, scaled by 2^64
};

// (long double)cut < 0.999L * (long double)best exactly as x87 extended
// precision evaluates it (64-bit significands, round to nearest even), in
// integers: 0.999L = m0999 / 2^64, so the product is best * m0999 / 2^64
// rounded to 64 significant bits.
__host__ __device__ inline bool below_0999(std::uint64_t cut, std::uint64_t best, std::uint64_t m0999) {
    using u128 = unsigned __int128;
    u128 p = (u128)best * m0999;
    const std::uint64_t hi = (std::uint64_t)(p >> 64);
    const int len = hi ? 128 - __builtin_clzll(hi) : 0;
    if (len > 64) {
        const int s = len - 64;
        const u128 q = p >> s;
        const u128 rem = p - (q << s);
        const u128 half = (u128)1 << (s - 1);
        const u128 r = (rem > half || (rem == half && (q & 1))) ? q + 1 : q;
        p = r << s;
    }
    return ((u128)cut << 64) < p;
}

// 0.999L as m / 2^64 (0.999L lies in [1/2, 1), so m has 64 bits).
inline std::uint64_t significand_0999() { return static_cast<std::uint64_t>(std::ldexp(0.999L, 64)); }

// Loop state at the start of a run of device-controlled rounds; clears the
// counter block; copy_best: the best loads take the current loads.
__global__ void k_lp_arm(unsigned long long* ctr, int k, unsigned long long run, unsigned long long phase,
                         unsigned long long stale, unsigned long long empty, unsigned long long cut,
                         unsigned long long best, unsigned long long lp_moves, int copy_best) {
    unsigned long long* loads = ctr + kCtrCount;
    unsigned long long* L = loads + k;
    for (int i = threadIdx.x; i < kCtrCount; i += blockDim.x) ctr[i] = 0;
    if (copy_best)
        for (int p = threadIdx.x; p < k; p += blockDim.x) L[kLoopWords + p] = loads[p];
    if (threadIdx.x == 0) {
        L[kRun] = run;
        L[kPhase] = phase;
        L[kStale] = stale;
        L[kEmpty] = empty;
        L[kCut] = cut;
        L[kBestCut] = best;
        L[kLpMoves] = lp_moves;
        L[kReason] = kRunning;
        L[kImproved] = 0;
        for (int i = kAccFrontier; i < kLoopWords; ++i) L[i] = 0;
    }
}

// End of one LP round: the movers, cut delta and part loads of the round
// decide whether the loop goes on; the counter block is cleared for the next
// round and the state is mirrored to the host mailbox.
// keep: non-null on small levels, where this block also checkpoints the
// labels of an improving round (n bytes from labels to keep) instead of a
// separate k_lp_keep launch.
__global__ void __launch_bounds__(256) k_lp_decide(unsigned long long* ctr, LoopParams P,
                                                   unsigned long long* mailbox, const Label* labels, Label* keep,
                                                   std::int64_t n) {
    unsigned long long* loads = ctr + kCtrCount;
    unsigned long long* L = loads + P.k;
    __shared__ int s_run, s_improved, s_stop;
    __shared__ unsigned long long s_words[kLoopWords];
    __shared__ unsigned long long s_max[256 / kWave];
    // Largest part load, by every thread.
    unsigned long long mx = 0;
    for (int p = threadIdx.x; p < P.k; p += blockDim.x) mx = loads[p] > mx ? loads[p] : mx;
    mx = wave_max(mx);
    if (lane_id() == 0) s_max[threadIdx.x / kWave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (unsigned i = 1; i < blockDim.x / kWave; ++i) mx = s_max[i] > mx ? s_max[i] : mx;
        s_run = L[kRun] != 0;
        s_improved = 0;
        s_stop = 0;
        L[kImproved] = 0;
        if (s_run) {
            const std::uint64_t phase = L[kPhase] + 1;
            std::uint64_t stale = L[kStale], empty = L[kEmpty], cut = L[kCut], best = L[kBestCut];
            const std::uint64_t movers = ctr[kMovers];
            const std::int64_t delta = (std::int64_t)ctr[kDelta];
            int reason = kRunning;
            if (P.profiling) {
                L[kAccFrontier] += ctr[kFrontier];
                L[kAccRowSum] += ctr[kRowSum];
                L[kAccAfterE] += ctr[kEntAfter];
                L[kAccAfterR] += ctr[kEntAfter + 1];
                L[kAccUpdateE] += ctr[kEntUpdate];
                L[kAccUpdateR] += ctr[kEntUpdate + 1];
            }
            if (P.light_only && (ctr[kTier] || ctr[kTierB] || ctr[kTierC])) {
                reason = kDeferred;
            } else if (!movers) {
                if (++empty >= 2)
                    reason = kEnded;
                else
                    ++stale;
            } else {
                empty = 0;
                if (delta >= 0 && (std::uint64_t)delta > cut) {
                    reason = kCutUnderflow;
                } else {
                    cut = delta >= 0 ? cut - (std::uint64_t)delta : cut + (std::uint64_t)(-delta);
                    L[kLpMoves] += movers;
                    if (mx > P.cap) {
                        reason = kNeedRepair;  // the host repairs and finishes the round
                    } else {
                        ++stale;
                        if (cut < best) {
                            if (P.graph_mode ? best - cut > best / 1000 : below_0999(cut, best, P.m0999)) stale = 0;
                            best = cut;
                            s_improved = 1;
                            L[kImproved] = 1;
                        }
                    }
                }
            }
            if (reason == kRunning && !(phase < (std::uint64_t)P.rounds && stale < (std::uint64_t)P.stale_rounds))
                reason = kEnded;
            L[kPhase] = phase;
            L[kStale] = stale;
            L[kEmpty] = empty;
            L[kCut] = cut;
            L[kBestCut] = best;
            L[kReason] = reason;
            if (reason != kRunning) {
                L[kRun] = 0;
                s_stop = 1;
            }
            for (int i = 0; i < kLoopWords; ++i) s_words[i] = L[i];
        }
    }
    __syncthreads();
    if (!s_run) return;
    // Running: the first kMailWords loop words to the mailbox; stopped:
    // every loop word, then the loads and the best loads.
    if (mailbox)
        for (int i = threadIdx.x; i < (s_stop ? (int)kLoopWords : kMailWords); i += blockDim.x) mailbox[i] = s_words[i];
    for (int p = threadIdx.x; p < P.k; p += blockDim.x) {
        const unsigned long long load = loads[p];
        const unsigned long long best = s_improved ? load : L[kLoopWords + p];
        if (s_improved) L[kLoopWords + p] = load;
        if (mailbox && s_stop) {
            mailbox[kLoopWords + p] = load;
            mailbox[kLoopWords + P.k + p] = best;
        }
    }
    for (int i = threadIdx.x; i < kCtrCount; i += blockDim.x) ctr[i] = 0;
    if (keep && s_improved)
        for (std::int64_t v = threadIdx.x; v < n; v += blockDim.x) keep[v] = labels[v];
}

// Checkpoint of the labels of a round that set a new best cut.
constexpr std::int64_t kKeepInline = 16384;
// vec: both arrays are 16-byte aligned (copied in 16-byte words).
__global__ void k_lp_keep(const unsigned long long* flag, const Label* labels, Label* best, std::int64_t n, bool vec) {
    if (!*flag) return;
    const std::int64_t words = vec ? n / 16 : 0;
    const uint4* src = reinterpret_cast<const uint4*>(labels);
    uint4* dst = reinterpret_cast<uint4*>(best);
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < words;
         i += (std::int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
    for (std::int64_t v = words * 16 + blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x)
        best[v] = labels[v];
}

// Donors that were not admitted carry no target (rebalance_graph.cuh:155).
__global__ void k_repair_clear(RefArgs a, const std::int32_t* donors, const unsigned long long* count_ptr) {
    const std::int64_t count = (std::int64_t)*count_ptr;
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < count;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int32_t v = donors[i];
        if (!a.flags[v]) a.targets[v] = -1;
    }
}

// ------------------------------------------------------------ exchange
struct JMove {
    std::int32_t v[2 * kJointSide];
    std::int32_t size, a, b, id;
    std::int64_t gain, net;
};

__host__ __device__ inline JMove jmove_empty() {
    JMove m;
    for (int j = 0; j < 2 * kJointSide; ++j) m.v[j] = -1;
    m.size = 0;
    m.a = -1;
    m.b = -1;
    m.id = 0;
    m.gain = 0;
    m.net = 0;
    return m;
}

__device__ inline std::int64_t jgap(std::int64_t net, std::int64_t lo, std::int64_t hi) {
    return net < lo ? lo - net : net > hi ? net - hi : 0;
}

template <int W>
__global__ void k_dense_affinity(RefArgs a, unsigned long long* A) {
    extern __shared__ unsigned long long smem[];
    const int lg = threadIdx.x % W, gi = threadIdx.x / W;
    unsigned long long* hist = smem + gi * a.k;
    const std::int64_t gpb = blockDim.x / W;
    for (std::int64_t v = blockIdx.x * gpb + gi; v < a.n; v += (std::int64_t)gridDim.x * gpb) {
        team_histogram(GroupTeam<W>{lg}, a, v, hist);
        for (int p = lg; p < a.k; p += W) A[v * a.k + p] = hist[p];
        wave_sync();
    }
}

__global__ void k_quotient(RefArgs a, const unsigned long long* A, unsigned long long* Q) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < a.n;
         v += (std::int64_t)gridDim.x * blockDim.x) {
        const int la = a.labels[v];
        for (int b = la + 1; b < a.k; ++b) {
            const unsigned long long x = A[v * a.k + b];
            if (x) atomicAdd(&Q[la * a.k + b], x);
        }
    }
}

// joint_match (exchange.cuh:55-83): greedy mutual-best pairing of parts in
// rounds, one block.  The pair scores Q / (1 + visits)^2 are computed once
// into LDS (a pair's visit count changes only when the pair is matched, and
// matched parts are not scored again); each part's best partner (largest
// score, then smallest part) is found by a group of W lanes over
// q = sub, sub + W, ...
constexpr int kMatchThreads = 1024;
// The block also resets the pass's seed slots (seed_words of ~0) for
// k_joint_seeds, which runs next.
__global__ void __launch_bounds__(kMatchThreads) k_joint_match(int k, const unsigned long long* Q,
                                                               std::int32_t* partner, std::int32_t* visits,
                                                               unsigned long long* seeds, int seed_words) {
    __shared__ double score_of[kMaxParts * kMaxParts];
    __shared__ int choice[kMaxParts];
    __shared__ int mate[kMaxParts];
    __shared__ int progress;
    for (int x = threadIdx.x; x < seed_words; x += blockDim.x) seeds[x] = ~0ULL;
    for (int x = threadIdx.x; x < k * k; x += blockDim.x) {
        const int p = x / k, q = x % k;
        double v = 0.0;
        if (p < q && Q[x]) {
            const double penalty = 1.0 + visits[x];
            v = (double)Q[x] / (penalty * penalty);
        }
        score_of[x] = v;
    }
    for (int p = threadIdx.x; p < k; p += blockDim.x) mate[p] = -1;
    __syncthreads();
    int W = kWave;
    while (W > 1 && W * k > kMatchThreads) W >>= 1;
    const int p = threadIdx.x / W, sub = threadIdx.x % W;
    for (int round = 0; round < k / 2; ++round) {
        if (threadIdx.x == 0) progress = 0;
        double score = -1;
        int best = -1;
        if (p < k && mate[p] < 0)
            for (int q = sub; q < k; q += W) {
                if (q == p || mate[q] >= 0) continue;
                const double v = score_of[p < q ? p * k + q : q * k + p];
                if (v > 0.0 && v > score) {
                    score = v;
                    best = q;
                }
            }
        for (int d = W / 2; d; d >>= 1) {
            const double os = __shfl_xor(score, d, W);
            const int ob = __shfl_xor(best, d, W);
            if (os > score || (os == score && ob < best)) {
                score = os;
                best = ob;
            }
        }
        if (p < k && sub == 0) choice[p] = best;
        __syncthreads();
        if (p < k && sub == 0 && best >= 0 && choice[best] == p) {
            mate[p] = best;
            if (p < best) {
                ++visits[p * k + best];
                atomicAdd(&progress, 1);
            }
        }
        __syncthreads();
        const int gained = progress;
        __syncthreads();
        if (!gained) break;
    }
    for (int q = threadIdx.x; q < k; q += blockDim.x) partner[q] = mate[q];
}

__global__ void k_joint_seeds(RefArgs a, const unsigned long long* A, const std::int32_t* partner,
                              unsigned long long* seeds, int sw, int pass) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < a.n;
         v += (std::int64_t)gridDim.x * blockDim.x) {
        const int p = a.labels[v], q = partner[p];
        if (q < 0 || !A[v * a.k + q]) continue;
        const std::uint32_t h = mix32((std::uint32_t)v ^ mix32((std::uint32_t)pass + 17));
        atomicMin(&seeds[p * sw + h % (std::uint32_t)sw], ((unsigned long long)h << 32) | (std::uint32_t)v);
    }
}

// ------------------------------------------------ exchange: transaction growth
// Growth of capacity-complementary transactions (the semantics of
// exchange.cuh:96-204).  A slot is a (partner seed, policy) pair; its
// transaction starts at the seed vertex and adds, up to eight members and at
// most four per side, the admissible vertex that ranks first by (marginal
// gain, balance gap, id), or by (gap, gain, id) under policy 0 while the net
// weight lies outside the feasible window.  Candidates are the vertices of a
// member's first 128 row entries that share the member's part, and, while no
// member sits on the partner side, the partner part's seeds; vertices of
// degree above 1024 never join.  Each step records the transaction as an
// archive fragment (positive gain, at most four members) and as the slot's
// best feasible transaction.
//
// CDNA design: one wave per slot with a candidate pool in LDS.  When a
// member joins, its row is read once and the entries that pass the static
// tests are appended to the pool by ballot compaction, together with their
// static marginal gain A(u, other) - A(u, own) (+2w for a row entry) and
// weight.  A step then scans at most 1024 pool records in LDS instead of
// re-reading every member row and the affinity rows behind it; records of a
// vertex that joins are retired.  The step's winner comes from a wave
// butterfly whose in-row exchanges are DPP moves, and the joining vertex's
// correction term (its edges to the members) is a DPP sum over its row.
constexpr int kGrowNominate = 128;                          // row entries that nominate
constexpr int kGrowPool = 7 * kGrowNominate + kJointSeeds;  // members 1..7 nominate; seeds once
constexpr int kGrowWaves = 2;                               // slots per block
constexpr std::uint8_t kPoolPartner = 1, kPoolSeed = 2, kPoolRetired = 4;

struct GrowPool {
    std::int64_t marginal[kGrowPool];
    std::int32_t u[kGrowPool];
    std::uint32_t w[kGrowPool];
    std::uint8_t kind[kGrowPool];
    std::int32_t members[8];
    std::int32_t member_label[8];
    int count;
};

struct GrowPick {
    std::int64_t marginal, gap;
    std::int32_t u, pad;
};

// x ranks before y (the step order of exchange.cuh:166-176).
__device__ inline bool grow_before(const GrowPick& x, const GrowPick& y, bool gap_first) {
    if (gap_first) {
        if (x.gap != y.gap) return x.gap < y.gap;
        if (x.marginal != y.marginal) return x.marginal > y.marginal;
    } else {
        if (x.marginal != y.marginal) return x.marginal > y.marginal;
        if (x.gap != y.gap) return x.gap < y.gap;
    }
    return x.u < y.u;
}

struct GrowBest {
    bool gap_first;
    __device__ GrowPick operator()(const GrowPick& x, const GrowPick& y) const {
        return grow_before(y, x, gap_first) ? y : x;
    }
};

// Appends what one new member nominates (member >= 0: its first 128 row
// entries in its own part) or, with member < 0, the partner part's seeds.
// Wave-uniform; every lane takes part.
__device__ inline void grow_nominate(const RefArgs& a, const unsigned long long* A, const unsigned long long* seeds,
                                     int sw, int source, int target, std::int32_t member, int size, GrowPool& P,
                                     int lane) {
    const int k = a.k;
    std::int64_t begin = 0, end = sw;
    int own = -1;
    if (member >= 0) {
        begin = a.off[member];
        end = a.off[member + 1];
        if (end > begin + kGrowNominate) end = begin + kGrowNominate;
        own = a.labels[member];
    }
    for (std::int64_t base = begin; base < end; base += kWave) {
        const std::int64_t e = base + lane;
        bool keep = false;
        std::int32_t u = -1;
        int p = -1;
        if (e < end) {
            if (member < 0) {
                const unsigned long long val = seeds[(std::int64_t)target * sw + e];
                if (val != ~0ULL) u = (std::int32_t)(std::uint32_t)val;
            } else {
                u = a.nbr[e];
            }
            if (u >= 0) {
                p = a.labels[u];
                keep = (p == source || p == target) && (member < 0 || p == own) && a.off[u + 1] - a.off[u] <= 1024;
                for (int j = 0; j < size; ++j) keep &= u != P.members[j];
            }
        }
        const unsigned long long ballot = __ballot(keep);
        if (keep) {
            const int r = P.count + __popcll(ballot & ((1ULL << lane) - 1ULL));
            const int other = p == source ? target : source;
            std::int64_t marginal = (std::int64_t)A[(std::int64_t)u * k + other] - (std::int64_t)A[(std::int64_t)u * k + p];
            if (member >= 0) marginal += 2 * (std::int64_t)ew_at(a.ew, e);
            P.marginal[r] = marginal;
            P.u[r] = u;
            P.w[r] = a.vw[u];
            P.kind[r] = (std::uint8_t)((p == source ? 0 : kPoolPartner) | (member < 0 ? kPoolSeed : 0));
        }
        wave_sync();
        if (lane == 0) P.count += __popcll(ballot);
        wave_sync();
    }
}

__global__ void __launch_bounds__(kGrowWaves* kWave) k_joint_grow(RefArgs a, const unsigned long long* A, int slots,
                                                                   int sw, const unsigned long long* seeds,
                                                                   const std::int32_t* partner, JMove* moves,
                                                                   JMove* archive) {
    __shared__ GrowPool pools[kGrowWaves];
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    const int slot = blockIdx.x * kGrowWaves + wave;
    if (slot >= slots) return;
    GrowPool& P = pools[wave];
    const int policy = slot % 2;
    const int k = a.k;
    JMove best = jmove_empty();
    best.id = slot;
    if (lane == 0) {  // the same lane writes these slots again below
        moves[slot] = best;
        for (int j = 0; j < kJointArchive; ++j) archive[slot * kJointArchive + j] = jmove_empty();
    }
    const unsigned long long enc = seeds[slot / 2];
    if (enc == ~0ULL) return;
    const std::int32_t root = (std::int32_t)(std::uint32_t)enc;
    const int source = a.labels[root], target = partner[source];
    if (target < 0) return;
    if (lane == 0) {
        P.count = 0;
        P.members[0] = root;
        P.member_label[0] = source;
    }
    wave_sync();
    int size = 1, c0 = 1, c1 = 0;
    std::int64_t net = a.vw[root];
    const std::int64_t lo = (std::int64_t)a.loads[source] - (std::int64_t)a.cap;
    const std::int64_t hi = (std::int64_t)a.cap - (std::int64_t)a.loads[target];
    std::int64_t gain = (std::int64_t)A[(std::int64_t)root * k + target] - (std::int64_t)A[(std::int64_t)root * k + source];
    const int pa = source < target ? source : target, pb = source < target ? target : source;
    grow_nominate(a, A, seeds, sw, source, target, -1, size, P, lane);
    grow_nominate(a, A, seeds, sw, source, target, root, size, P, lane);
    for (int step = 0;; ++step) {
        if (lane == 0 && gain > 0 && size <= kJointArchive) {
            JMove st = jmove_empty();
            st.size = size;
            st.a = pa;
            st.b = pb;
            st.gain = gain;
            st.net = source == pa ? net : -net;
            st.id = slot * kJointArchive + step;
            for (int j = 0; j < size; ++j) {  // members in ascending id order
                int pos = j;
                while (pos && st.v[pos - 1] > P.members[j]) {
                    st.v[pos] = st.v[pos - 1];
                    --pos;
                }
                st.v[pos] = P.members[j];
            }
            archive[slot * kJointArchive + step] = st;
        }
        if (gain > best.gain && !jgap(net, lo, hi)) {
            best.size = size;
            best.a = pa;
            best.b = pb;
            best.gain = gain;
            best.net = source == pa ? net : -net;
            for (int j = 0; j < 8; ++j) best.v[j] = j < size ? P.members[j] : -1;
        }
        if (size == 8) break;
        const bool gap_first = policy == 0 && jgap(net, lo, hi) != 0;
        GrowPick pick{LLONG_MIN, LLONG_MAX, INT_MAX, 0};
        for (int r = lane; r < P.count; r += kWave) {
            const int kind = P.kind[r];
            const int side = kind & kPoolPartner;
            if ((kind & kPoolRetired) || ((kind & kPoolSeed) && c1) || (side ? c1 : c0) >= kJointSide) continue;
            if ((net > hi && !side) || (net < lo && side)) continue;
            const std::int64_t cnet = net + (side ? -(std::int64_t)P.w[r] : (std::int64_t)P.w[r]);
            const GrowPick c{P.marginal[r], jgap(cnet, lo, hi), P.u[r], 0};
            if (grow_before(c, pick, gap_first)) pick = c;
        }
        pick = wave_allreduce(pick, GrowBest{gap_first});
        if (pick.u == INT_MAX) break;
        const std::int32_t chosen = pick.u;
        const int lc = a.labels[chosen];
        const int side = lc == source ? 0 : 1;
        std::int64_t corr = 0;
        for (std::int64_t e = a.off[chosen] + lane; e < a.off[chosen + 1]; e += kWave) {
            const std::int32_t nb = a.nbr[e];
            for (int j = 0; j < size; ++j)
                if (nb == P.members[j]) corr += (P.member_label[j] == lc ? 2 : -2) * (std::int64_t)ew_at(a.ew, e);
        }
        corr = wave_allreduce(corr, PlusI64{});
        gain += (std::int64_t)A[(std::int64_t)chosen * k + (side ? source : target)] -
                (std::int64_t)A[(std::int64_t)chosen * k + lc] + corr;
        net += side ? -(std::int64_t)a.vw[chosen] : (std::int64_t)a.vw[chosen];
        if (side) ++c1;
        else ++c0;
        // Retire the newcomer's records, then let it nominate.
        for (int r = lane; r < P.count; r += kWave)
            if (P.u[r] == chosen) P.kind[r] |= kPoolRetired;
        if (lane == 0) {
            P.members[size] = chosen;
            P.member_label[size] = lc;
        }
        wave_sync();
        ++size;
        if (size < 8) grow_nominate(a, A, seeds, sw, source, target, chosen, size, P, lane);
    }
    if (lane == 0) moves[slot] = best;
}

// Orders for the archive and the transaction pool (complementary_moves.cuh:5-14, exchange.cuh:33-40).
struct NetOrder {
    __device__ bool operator()(const JMove& x, const JMove& y) const {  // x before y
        if (x.a != y.a) return x.a < y.a;
        if (x.net != y.net) return x.net < y.net;
        if (x.gain != y.gain) return x.gain > y.gain;
        if (x.size != y.size) return x.size < y.size;
        for (int j = 0; j < x.size; ++j)
            if (x.v[j] != y.v[j]) return x.v[j] < y.v[j];
        return x.id < y.id;
    }
};
struct JointOrder {
    __device__ bool operator()(const JMove& x, const JMove& y) const {
        if ((x.size != 0) != (y.size != 0)) return x.size != 0;
        if (x.a != y.a) return x.a < y.a;
        if (x.gain != y.gain) return x.gain > y.gain;
        return x.id < y.id;
    }
};

template <typename Cmp>
__device__ inline bool idx_after(int x, int y, const JMove* m, Cmp cmp) {  // x sorts after y
    if (x < 0) return y >= 0;
    if (y < 0) return false;
    return cmp(m[y], m[x]);
}

template <typename Cmp>
__global__ void k_bitonic_global(int* idx, int N, int j, int kk, const JMove* m, Cmp cmp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int l = i ^ j;
    if (l <= i) return;
    const int x = idx[i], y = idx[l];
    const bool asc = (i & kk) == 0;
    if (asc ? idx_after(x, y, m, cmp) : idx_after(y, x, m, cmp)) {
        idx[i] = y;
        idx[l] = x;
    }
}
// x sorts after y, for local slots of moves staged in LDS (g holds each
// slot's global index, -1 for padding, which sorts last).
template <typename Cmp>
__device__ inline bool slot_after(int x, int y, const int* g, const JMove* mv, Cmp cmp) {
    if (g[x] < 0) return g[y] >= 0;
    if (g[y] < 0) return false;
    return cmp(mv[y], mv[x]);
}

// Bitonic stages over C indices held by one block, with the moves they
// name copied into LDS once so that every comparison reads LDS: stages
// (kk, j) for kk from kk_first to kk_last and j from min(kk, C) / 2 down to
// 1.  Directions use the global position base + t.
template <int C, typename Cmp>
__device__ inline void bitonic_lds(int* idx, int base, int kk_first, int kk_last, const JMove* m, Cmp cmp,
                                   JMove* mv, int* g, short* s) {
    for (int i = threadIdx.x; i < C; i += blockDim.x) {
        const int x = idx[base + i];
        g[i] = x;
        s[i] = (short)i;
        if (x >= 0) mv[i] = m[x];
    }
    __syncthreads();
    for (int kk = kk_first; kk <= kk_last; kk <<= 1)
        for (int j = (kk < C ? kk : C) >> 1; j > 0; j >>= 1) {
            // One compare-exchange pair (t, t + j) per thread.
            for (int q = threadIdx.x; q < C / 2; q += blockDim.x) {
                const int t = ((q & ~(j - 1)) << 1) | (q & (j - 1)), l = t + j;
                const bool asc = ((base + t) & kk) == 0;
                const int x = s[t], y = s[l];
                if (asc ? slot_after(x, y, g, mv, cmp) : slot_after(y, x, g, mv, cmp)) {
                    s[t] = (short)y;
                    s[l] = (short)x;
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < C; i += blockDim.x) idx[base + i] = g[s[i]];
}

// The stages of a bitonic network whose partners lie in one aligned chunk
// of kBitonicChunk indices, run in LDS by one block per chunk: every stage
// of the merges up to the chunk size (kk == 0), or the stages j < chunk of
// the merge of size kk.  The chunk's moves (64 bytes each) take 128 KiB of
// LDS.
constexpr int kBitonicChunk = 2048;
template <typename Cmp>
__global__ void __launch_bounds__(1024) k_bitonic_chunk(int* idx, int kk_fixed, const JMove* m, Cmp cmp) {
    __shared__ JMove mv[kBitonicChunk];
    __shared__ int g[kBitonicChunk];
    __shared__ short s[kBitonicChunk];
    const int base = blockIdx.x * kBitonicChunk;
    if (kk_fixed == 0)
        bitonic_lds<kBitonicChunk>(idx, base, 2, kBitonicChunk, m, cmp, mv, g, s);
    else
        bitonic_lds<kBitonicChunk>(idx, base, kk_fixed, kk_fixed, m, cmp, mv, g, s);
}

// Whole sort of N <= 1024 indices (N a power of two) in one block.
template <typename Cmp>
__global__ void __launch_bounds__(1024) k_bitonic_local(int* idx, int N, const JMove* m, Cmp cmp) {
    __shared__ JMove mv[1024];
    __shared__ int g[1024];
    __shared__ short s[1024];
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const int x = idx[i];
        g[i] = x;
        s[i] = (short)i;
        if (x >= 0) mv[i] = m[x];
    }
    __syncthreads();
    for (int kk = 2; kk <= N; kk <<= 1)
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int q = threadIdx.x; q < N / 2; q += blockDim.x) {
                const int t = ((q & ~(j - 1)) << 1) | (q & (j - 1)), l = t + j;
                const bool asc = (t & kk) == 0;
                const int x = s[t], y = s[l];
                if (asc ? slot_after(x, y, g, mv, cmp) : slot_after(y, x, g, mv, cmp)) {
                    s[t] = (short)y;
                    s[l] = (short)x;
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < N; i += blockDim.x) idx[i] = g[s[i]];
}

__global__ void k_nonempty_moves(const JMove* m, int count, std::uint8_t* f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) f[i] = m[i].size != 0;
}
__global__ void k_gather_moves(const JMove* src, const int* idx, int count, JMove* dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = src[idx[i]];
}
// exchange.cuh:19-32 equality of consecutive sorted archive entries.
__global__ void k_unique_flags(const JMove* m, int count, std::uint8_t* f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    bool keep = true;
    if (i > 0) {
        const JMove& x = m[i - 1];
        const JMove& y = m[i];
        bool same = x.size && x.size == y.size && x.a == y.a && x.b == y.b && x.gain == y.gain && x.net == y.net;
        for (int j = 0; same && j < x.size; ++j) {
            bool found = false;
            for (int t = 0; t < y.size; ++t) found |= x.v[j] == y.v[t];
            same = found;
        }
        keep = !same;
    }
    f[i] = keep;
}

__device__ inline int jnet_bound(const JMove* r, int count, int pa, std::int64_t net, bool upper) {
    int lo = 0, hi = count;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        const JMove& x = r[mid];
        if (x.a < pa || (x.a == pa && (x.net < net || (upper && x.net == net)))) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// joint_combine (complementary_moves.cuh:31-75): one wave per fragment.
__global__ void k_joint_combine(RefArgs a, int count, const JMove* archive, int id_base, JMove* out) {
    const int lane = lane_id();
    const int i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (i >= count) return;
    // Member loops run to the constant bound 8 with a size test, so the
    // members stay in registers.
    const JMove x = archive[i];
    const std::int64_t lo = (std::int64_t)a.loads[x.a] - (std::int64_t)a.cap;
    const std::int64_t hi = (std::int64_t)a.cap - (std::int64_t)a.loads[x.b];
    if (lane == 0) {
        JMove o = x;
        o.id = id_base + i;
        if (jgap(x.net, lo, hi)) o.size = 0;
        out[i] = o;
        out[count + i] = jmove_empty();
    }
    const int begin = jnet_bound(archive, count, x.a, lo - x.net, false);
    const int end = jnet_bound(archive, count, x.a, hi - x.net, true);
    // Up to four candidates (quartiles of the feasible range), checked
    // first; the member rows of x are then scanned once for all of them.
    int cand[4] = {-1, -1, -1, -1};
    std::int32_t cv[4][8];
    int nc = 0, previous = -1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int t = 0; t < 8; ++t) cv[r][t] = -1;
        if (begin >= end) continue;
        const int j = begin + (int)(((std::int64_t)(end - begin - 1) * r) / 3);
        if (j == previous) continue;
        previous = j;
        if (j == i) continue;
        const JMove y = archive[j];
        if (y.a != x.a || y.b != x.b || x.size + y.size > 8) continue;
        bool overlap = false;
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int t = 0; t < 8; ++t) overlap |= u < x.size && t < y.size && x.v[u] == y.v[t];
        if (overlap) continue;
        cand[r] = j;
#pragma unroll
        for (int t = 0; t < 8; ++t) cv[r][t] = t < y.size ? y.v[t] : -1;
        ++nc;
    }
    std::int64_t corr[4] = {0, 0, 0, 0};
    if (nc)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (u >= x.size) break;
            const std::int32_t v = x.v[u];
            const int lv = a.labels[v];
            for (std::int64_t e = a.off[v] + lane; e < a.off[v + 1]; e += kWave) {
                const std::int32_t nb = a.nbr[e];
                bool seen = false;
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int t = 0; t < 8; ++t) seen |= nb == cv[c][t];
                if (!seen) continue;
                const std::int64_t w = (lv == a.labels[nb] ? 2 : -2) * (std::int64_t)ew_at(a.ew, e);
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int t = 0; t < 8; ++t)
                        if (nb == cv[c][t]) corr[c] += w;
            }
        }
    std::int64_t best_gain = 0;
    int best = -1;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        if (cand[c] < 0) continue;
        const std::int64_t gain = x.gain + archive[cand[c]].gain + wave_sum(corr[c]);
        if (gain > best_gain || (gain == best_gain && gain > 0 && cand[c] < best)) {
            best_gain = gain;
            best = cand[c];
        }
    }
    if (lane == 0 && best >= 0) {
        JMove res = x;
        const JMove y = archive[best];
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if (t < y.size && q == x.size + t) res.v[q] = y.v[t];
        res.size = x.size + y.size;
        res.gain = best_gain;
        res.net += y.net;
        res.id = id_base + count + i;
        out[count + i] = res;
    }
}

__global__ void k_joint_claim(const JMove* m, int count, std::int32_t* owners) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        for (int j = 0; j < m[i].size; ++j) atomicMin(&owners[m[i].v[j]], i);
}
__global__ void k_joint_resolve(JMove* m, int count, const std::int32_t* owners) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    for (int j = 0; j < m[i].size; ++j)
        if (owners[m[i].v[j]] != i) {
            m[i].size = 0;
            break;
        }
}
// idx[0, m) then -1 up to N (the power-of-two length a bitonic sort takes).
__global__ void k_pad_index(const std::int32_t* idx, int m, std::int32_t* out, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) out[i] = i < m ? idx[i] : -1;
}

// Clears the owner of every vertex and its moving flag before the claimed
// transactions are marked and accepted (nothing in between reads the flag).
__global__ void k_exchange_reset(std::int64_t n, std::int32_t* owners, std::uint32_t* state) {
    const std::uint32_t keep = ~(0xffu << (8 * kStateFlag));
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x) {
        owners[v] = INT_MAX;
        state[v] &= keep;
    }
}

// Also clears each transaction's event total for k_joint_events.
__global__ void k_joint_mark(RefArgs a, const JMove* m, int count, std::int32_t* owners, unsigned long long* events) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    events[i] = 0;
    for (int j = 0; j < m[i].size; ++j) {
        const std::int32_t v = m[i].v[j];
        owners[v] = i;
        a.targets[v] = a.labels[v] == m[i].a ? m[i].b : m[i].a;
    }
}
// Cut events of the claimed transactions (the semantics of
// exchange.cuh:205-262).  Every undirected edge with an end in a transaction
// contributes the change of its cut state once, counted from its lower end
// when both ends move: an edge inside one transaction, or to a vertex that
// stays, is charged whole to that transaction; an edge between transactions
// i < j is split into the change when i applies (charged to i) and the rest
// when j applies (charged to j), so prefix sums in list order are exact.
// Edges to a vertex of another transaction on other parts do not change.
//
// CDNA design: a block takes 32 consecutive transactions, stages their
// member rows (start, length, transaction) in LDS and walks them as one
// flattened edge range, so lanes stay busy whatever the row lengths.  Each
// wave reduces its lanes' charges to their own transactions by masked DPP
// butterflies, one per distinct transaction among its lanes (found by
// ballot), into 64-bit LDS totals; a charge to another transaction goes to
// global memory at once.  The totals are flushed once per block.  All sums
// are modulo 2^64, as the signed contributions of the reference.
constexpr int kEventTx = 32;
constexpr int kEventRows = kEventTx * 8;
constexpr int kEventThreads = 256;

__global__ void __launch_bounds__(kEventThreads) k_joint_events(RefArgs a, const JMove* m, int count,
                                                                const std::int32_t* owners,
                                                                unsigned long long* events) {
    __shared__ std::int32_t row_start[kEventRows];
    __shared__ std::int32_t row_first[kEventRows + 1];  // exclusive prefix of row lengths
    __shared__ std::int32_t row_vertex[kEventRows];
    __shared__ std::int32_t tx_a[kEventTx], tx_b[kEventTx];
    __shared__ unsigned long long totals[kEventTx];
    __shared__ std::int32_t wave_sums[kEventThreads / kWave];
    const int t = threadIdx.x, lane = lane_id(), w = t / kWave;
    const int tx0 = blockIdx.x * kEventTx;
    // Row t is member t % 8 of transaction t / 8 (length 0 if absent).
    {
        const int tx = t / 8, j = t % 8, i = tx0 + tx;
        std::int32_t len = 0, v = -1, start = 0;
        if (i < count && j < m[i].size) {
            v = m[i].v[j];
            start = a.off[v];
            len = a.off[v + 1] - start;
        }
        row_start[t] = start;
        row_vertex[t] = v;
        if (j == 0) {
            tx_a[tx] = i < count ? m[i].a : -1;
            tx_b[tx] = i < count ? m[i].b : -1;
            totals[tx] = 0;
        }
        const std::int32_t inc = wave_inclusive_scan<PlusOp>(len);
        if (lane == kWave - 1) wave_sums[w] = inc;
        __syncthreads();
        std::int32_t before = 0;
        for (int x = 0; x < w; ++x) before += wave_sums[x];
        row_first[t] = before + inc - len;
        if (t == kEventThreads - 1) row_first[kEventRows] = before + inc;
    }
    __syncthreads();
    const std::int32_t total = row_first[kEventRows];
    for (std::int32_t base = 0; base < total; base += kEventThreads) {
        const std::int32_t f = base + t;
        int key = -1;
        unsigned long long own = 0;
        if (f < total) {
            int lo = 0, hi = kEventRows - 1;  // last row whose first entry is <= f
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (row_first[mid] <= f) lo = mid;
                else hi = mid - 1;
            }
            const int tx = lo / 8, i = tx0 + tx;
            const std::int32_t v = row_vertex[lo];
            const std::int64_t e = (std::int64_t)row_start[lo] + (f - row_first[lo]);
            const std::int32_t u = a.nbr[e];
            const std::int32_t other = owners[u];
            const int lu = a.labels[u], lv = a.labels[v];
            const bool counted = u != v && (other == INT_MAX || (v < u && (lu == tx_a[tx] || lu == tx_b[tx])));
            if (counted) {
                const std::int64_t we = (std::int64_t)ew_at(a.ew, e);
                const int tv = a.targets[v];
                const int before = lv != lu;
                if (other == i || other == INT_MAX) {
                    const int after = tv != (other == i ? (int)a.targets[u] : lu);
                    own = (unsigned long long)((before - after) * we);
                } else {
                    const int tu = a.targets[u];
                    const int middle = i < other ? tv != lu : lv != tu;
                    const int after = tv != tu;
                    const unsigned long long first = (unsigned long long)((before - middle) * we);
                    const unsigned long long last = (unsigned long long)((middle - after) * we);
                    own = i < other ? first : last;
                    const unsigned long long theirs = i < other ? last : first;
                    if (theirs) atomicAdd(&events[other], theirs);
                }
                if (own) key = tx;
            }
        }
        // Per distinct transaction among the lanes: one masked butterfly sum.
        unsigned long long pending = __ballot(key >= 0);
        while (pending) {
            const int leader = __ffsll(pending) - 1;
            const int lk = __shfl(key, leader);
            const bool mine = key == lk;
            const unsigned long long s = wave_allreduce(mine ? own : 0ULL, PlusU64{});
            if (lane == leader) atomicAdd(&totals[lk], s);
            pending &= ~__ballot(mine);
        }
    }
    __syncthreads();
    if (t < kEventTx && tx0 + t < count && totals[t]) atomicAdd(&events[tx0 + t], totals[t]);
}
// joint_prefix / joint_accept (exchange.cuh:263-289).  One wave per part
// pa over its moves (the list is sorted by part): gain and net prefixes by
// wave scans, 64 moves per step.  The serial rule (walk the moves, keep the
// move where a feasible prefix gain first exceeds the best so far, starting
// from 0) keeps the first move of the largest feasible positive prefix
// gain, which is what the per-step maximum and first-lane ballot take.
__global__ void k_joint_prefix(RefArgs a, const JMove* m, int count, const unsigned long long* events,
                               std::int32_t* stops) {
    const int lane = lane_id();
    const int pa = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (pa >= a.k) return;
    int low = 0, high = count;
    while (low < high) {
        const int mid = (low + high) / 2;
        if (m[mid].a < pa) low = mid + 1;
        else high = mid;
    }
    int end = low;
    high = count;
    while (end < high) {
        const int mid = (end + high) / 2;
        if (m[mid].a <= pa) end = mid + 1;
        else high = mid;
    }
    const std::int64_t lo = (std::int64_t)a.loads[pa] - (std::int64_t)a.cap;
    std::int64_t gain = 0, net = 0, best = 0;
    int stop = -1;
    for (int base = low; base < end; base += kWave) {
        const int i = base + lane;
        std::int64_t g = 0, d = 0;
        int b = 0;
        bool live = false;
        if (i < end) {
            live = m[i].size != 0;
            if (live) {
                g = (std::int64_t)events[i];
                d = m[i].net;
                b = m[i].b;
            }
        }
        const std::int64_t pg = gain + wave_inclusive_scan<PlusOp>(g);
        const std::int64_t pn = net + wave_inclusive_scan<PlusOp>(d);
        std::int64_t cand = LLONG_MIN;
        if (live && !jgap(pn, lo, (std::int64_t)a.cap - (std::int64_t)a.loads[b])) cand = pg;
        const std::int64_t top = wave_max(cand);
        if (top > best) {
            best = top;
            stop = base + __ffsll((unsigned long long)__ballot(cand == top)) - 1;
        }
        gain = __shfl(pg, kWave - 1);
        net = __shfl(pn, kWave - 1);
    }
    if (lane == 0) stops[pa] = stop;
}
// Block 0 also clears the counter words [first, last) for the compaction,
// update and commit of the accepted moves.
__global__ void k_joint_accept(RefArgs a, const JMove* m, int count, const std::int32_t* stops, int first, int last) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0)
        for (int x = first + (int)threadIdx.x; x < last; x += blockDim.x) a.ctr[x] = 0;
    if (i < count && m[i].size && i <= stops[m[i].a])
        for (int j = 0; j < m[i].size; ++j) a.flags[m[i].v[j]] = 1;
}

// ---------------------------------------------------------- basics
template <typename L>
__global__ void k_project(std::int64_t n, const std::int32_t* map, const L* coarse, L* fine) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x)
        fine[v] = coarse[map[v]];
}
template <typename L>
__global__ void k_cut(std::int64_t n, const std::int32_t* off, const std::int32_t* nbr, const std::uint32_t* ew,
                      const L* labels, unsigned long long* out) {
    const int lane = lane_id();
    unsigned long long local = 0;
    for (std::int64_t base = (blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x) - lane; base < n;
         base += (std::int64_t)gridDim.x * blockDim.x) {
        const std::int64_t v = base + lane;
        std::int64_t b = 0, e = 0;
        if (v < n) {
            b = off[v];
            e = off[v + 1];
        }
        const bool heavy = v < n && e - b > 64;
        if (v < n && !heavy) {
            const int s = labels[v];
            for (std::int64_t x = b; x < e; ++x)
                if (labels[nbr[x]] != s) local += ew_at(ew, x);
        }
        unsigned long long hv = __ballot(heavy);
        while (hv) {
            const int src = __ffsll(hv) - 1;
            hv &= hv - 1;
            const std::int64_t hvx = __shfl(v, src);
            const int s = labels[hvx];
            for (std::int64_t x = off[hvx] + lane; x < off[hvx + 1]; x += kWave)
                if (labels[nbr[x]] != s) local += ew_at(ew, x);
        }
    }
    local = wave_sum(local);
    if (lane == 0 && local) atomicAdd(out, local);
}
template <typename L>
__global__ void k_loads(std::int64_t n, const std::uint32_t* vw, const L* labels, unsigned long long* loads) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x)
        atomicAdd(&loads[labels[v]], (unsigned long long)vw[v]);
}
template <typename A, typename B>
__global__ void k_convert(std::int64_t n, const A* in, B* out) {
    for (std::int64_t v = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; v < n;
         v += (std::int64_t)gridDim.x * blockDim.x)
        out[v] = (B)in[v];
}

inline int grid_for(std::int64_t n, int per_block = 256, int cap = 8192) {
    return static_cast<int>(std::max<std::int64_t>(1, std::min<std::int64_t>(ceil_div(n, per_block), cap)));
}

template <typename A, typename B>
void convert(Context& c, std::int64_t n, const A* in, B* out) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_convert<A, B>), dim3(grid_for(n, 256, 4096)), dim3(256), 0, c.stream(), n, in, out);
    GP_LAUNCH_CHECK();
}

template <typename L>
std::uint64_t cut_of(Context& c, const DGraph& g, const L* labels) {
    DBuf<unsigned long long> out = c.buf<unsigned long long>(1);
    fill<unsigned long long>(c, out.get(), 1, 0ULL);
    if (g.n > 0) {
        hipLaunchKernelGGL((k_cut<L>), dim3(grid_for(g.n)), dim3(256), 0, c.stream(), g.n, g.off.get(), g.nbr.get(),
                           g.ew_ptr(), labels, out.get());
        GP_LAUNCH_CHECK();
    }
    const unsigned long long d = c.read_scalar(out.get());
    if (d & 1ULL) throw std::runtime_error("refinement cut is not symmetric");
    return d / 2;
}

template <typename L>
void loads_of(Context& c, const DGraph& g, const L* labels, int k, std::uint64_t* loads_host) {
    DBuf<unsigned long long> out = c.buf<unsigned long long>(static_cast<std::size_t>(k));
    fill<unsigned long long>(c, out.get(), k, 0ULL);
    hipLaunchKernelGGL((k_loads<L>), dim3(grid_for(g.n)), dim3(256), 0, c.stream(), g.n, g.vw.get(), labels, out.get());
    GP_LAUNCH_CHECK();
    c.read(loads_host, out.get(), static_cast<std::size_t>(k) * sizeof(unsigned long long));
}

}  // namespace

// ====================================================================== host

void project(Context& c, std::int64_t n, const std::int32_t* map, const std::int32_t* coarse, std::int32_t* fine) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_project<std::int32_t>), dim3(grid_for(n)), dim3(256), 0, c.stream(), n, map, coarse, fine);
    GP_LAUNCH_CHECK();
}

std::uint64_t device_cut(Context& c, const DGraph& g, const std::int32_t* labels) { return cut_of(c, g, labels); }

std::uint64_t device_edge_weight_sum(Context& c, const DGraph& g) {
    if (!g.ew) return static_cast<std::uint64_t>(g.m);
    DBuf<unsigned long long> out = c.buf<unsigned long long>(1);
    fill<unsigned long long>(c, out.get(), 1, 0ULL);
    if (g.m > 0) {
        hipLaunchKernelGGL((k_reduce_sum<std::uint32_t>), dim3(grid_for(g.m)), dim3(256), 0, c.stream(), g.ew.get(), g.m,
                           out.get());
        GP_LAUNCH_CHECK();
    }
    return c.read_scalar(out.get());
}

void device_loads(Context& c, const DGraph& g, const std::int32_t* labels, int k, std::uint64_t* loads_host) {
    loads_of(c, g, labels, k, loads_host);
}

namespace {

class LevelRefiner {
public:
    LevelRefiner(Context& c, const DGraph& g, Label* labels, std::vector<std::uint64_t>& loads,
                 const RefineOptions& o, std::uint64_t cap, LevelStats& st)
        : c_(c), g_(g), loads_host_(loads), o_(o), st_(st) {
        const std::int64_t n = g.n;
        const int k = o.parts;
        std::memset(&a_, 0, sizeof a_);
        a_.n = n;
        a_.off = g.off.get();
        a_.nbr = g.nbr.get();
        a_.ew = g.ew_ptr();
        a_.vw = g.vw.get();
        a_.k = k;
        a_.cap = cap;
        a_.labels = labels;
        state_ = c.buf<std::uint32_t>(n);
        // Gains of at most the row weight: 32 bits below 2^31.
        if (g.incident_bound < (1ULL << 31))
            gains32_ = c.buf<std::int32_t>(n);
        else
            gains_ = c.buf<std::int64_t>(n);
        frontier_ = c.buf<std::int32_t>(n);
        cand_ = c.buf<std::int32_t>(n);
        movers_ = c.buf<std::int32_t>(n);
        medium_ = c.buf<std::int32_t>(n);
        heavy_ = c.buf<std::int32_t>(n);
        // Counters and part loads share one allocation: one read returns both.
        // Counters, part loads, then the device loop words and best loads.
        ctr_ = c.buf<unsigned long long>(static_cast<std::size_t>(kCtrCount + k + kLoopWords + k));
        loads0_ = c.buf<unsigned long long>(static_cast<std::size_t>(k));
        bucket_ = c.buf<std::int32_t>(n);
        counts_ = c.buf<unsigned long long>(static_cast<std::size_t>(k) * kBalanceWidth);
        std::uint8_t* lanes = reinterpret_cast<std::uint8_t*>(state_.get());
        a_.state = state_.get();
        a_.targets = StateLane<std::int8_t>{reinterpret_cast<std::int8_t*>(lanes + kStateTarget)};
        a_.flags = StateLane<std::uint8_t>{lanes + kStateFlag};
        a_.member = StateLane<std::uint8_t>{lanes + kStateMember};
        a_.locks = StateLane<std::uint8_t>{lanes + kStateLock};
        a_.gains = gains_.get();
        a_.gains32 = gains32_.get();
        a_.ctr = ctr_.get();
        a_.loads = ctr_.get() + kCtrCount;
        a_.loads0 = loads0_.get();
        a_.bucket = bucket_.get();
        a_.counts = counts_.get();
        c.upload(a_.loads, loads.data(), static_cast<std::size_t>(k) * sizeof(std::uint64_t));
        fill<unsigned long long>(c, ctr_.get(), kCtrCount, 0ULL);
        std::uint64_t total = 0;
        for (auto x : loads) total += x;
        if (total > static_cast<std::uint64_t>(n)) {
            prices_ = c.buf<double>(static_cast<std::size_t>(k));
            samples_ = c.buf<RepairSample>(static_cast<std::size_t>(k));
            fill<double>(c, prices_.get(), k, 0.0);
            a_.prices = prices_.get();
            a_.samples = samples_.get();
        }
        // Boundary vertices seed the frontier; per-level resets (its tier
        // counters were cleared with the counter block above).
        tiers(BoundaryBody{{a_}}, nullptr, n, nullptr, nullptr, nullptr, nullptr, 0, kTier, true);
        // filter constant (incremental.cuh:897-899)
        std::uint64_t max_edge = 1;
        if (g.ew) {
            DBuf<unsigned long long> mx = c.buf<unsigned long long>(1);
            fill<unsigned long long>(c, mx.get(), 1, 0ULL);
            if (g.m > 0) {
                hipLaunchKernelGGL((k_reduce_max<std::uint32_t>), dim3(grid_for(g.m, 256, 2048)), dim3(256), 0, c.stream(),
                                   g.ew.get(), g.m, mx.get());
                GP_LAUNCH_CHECK();
            }
            max_edge = c.read_scalar(mx.get());
        }
        a_.filter = max_edge <= 1 ? 0.25 : 0.75;
    }

    std::uint64_t max_load() const {
        std::uint64_t m = 0;
        for (auto x : loads_host_) m = std::max(m, x);
        return m;
    }

    unsigned long long* loads_dev() const { return a_.loads; }

    // Clears the moving flag and the target (-1) of every vertex, the state
    // the exchange starts from.
    void clear_moves(Context& c) {
        const std::uint32_t keep = ~((0xffu << (8 * kStateFlag)) | (0xffu << (8 * kStateTarget)));
        const std::uint32_t put = 0xffu << (8 * kStateTarget);
        hipLaunchKernelGGL(k_set_lanes, dim3(grid_for(a_.n, 256, 2048)), dim3(256), 0, c.stream(), state_.get(), a_.n,
                           keep, put);
        GP_LAUNCH_CHECK();
    }

    // Short-row bins fit 32 bits when every row of the level weighs less
    // than 2^32 (the level's incident-weight bound).
    bool narrow_bins() const { return g_.incident_bound <= 0xffffffffULL; }

    // slot: the two tier list counters; zeroed: they are known to be zero
    // (cleared with the counter block since their last use).
    template <class Body>
    void tiers(const Body& body, const std::int32_t* list, std::int64_t n, const unsigned long long* count,
               std::int32_t* out, unsigned long long* out_count, unsigned long long* acc, int scratch,
               int slot = kTier, bool zeroed = false, int entries_slot = -1) {
        run_tiers(c_, body, list, n, count, medium_.get(), heavy_.get(), ctr_.get() + slot, out, out_count, acc,
                  scratch, a_.n, g_.row_len_bound, zeroed,
                  entries_slot >= 0 && c_.profiling() ? ctr_.get() + entries_slot : nullptr);
    }

    // One LP round over the frontier; returns the number of movers and the cut delta.
    std::int64_t lp_round(std::int64_t& delta) {
        const hipStream_t s = c_.stream();
        const std::int64_t n = a_.n;
        if (!ctr_clear_) fill<unsigned long long>(c_, ctr_.get(), kCtrCount, 0ULL);
        ctr_clear_ = false;
        hipLaunchKernelGGL(k_compact_flags<StateLane<std::uint8_t>>, dim3(grid_for(n, 256 * kScanRounds, 2048)), dim3(256), 0, s, n,
                           a_.member,
                           frontier_.get(), ctr_.get() + kFrontier, nullptr);
        GP_LAUNCH_CHECK();
        // Algorithmic bytes per scanned row: offsets 8, label and weight 8,
        // target and gain 12, and per entry neighbour 4, neighbour label 4
        // and edge weight 4 when weighted; added once the row sums are read.
        propose_sweeps();
        return apply(true, false, delta, true);
    }

    // The launches of one LP round of a device-controlled loop: the frontier
    // compaction does nothing once the loop word kRun is zero, so every later
    // launch of the round finds empty lists; k_lp_decide ends the round.
    void round_issue(const LoopParams& P, Label* best_labels, bool vec) {
        const hipStream_t s = c_.stream();
        const std::int64_t n = a_.n;
        unsigned long long* L = ctr_.get() + kCtrCount + a_.k;
        hipLaunchKernelGGL(k_compact_flags<StateLane<std::uint8_t>>, dim3(grid_for(n, 256 * kScanRounds, 2048)),
                           dim3(256), 0, s, n, a_.member, frontier_.get(), ctr_.get() + kFrontier, L + kRun);
        GP_LAUNCH_CHECK();
        propose_sweeps();
        apply_launch(true, false, true);
        // Levels of at most kKeepInline vertices checkpoint inside the
        // decision kernel; larger ones in a launch of their own.
        const bool inline_keep = n <= kKeepInline;
        hipLaunchKernelGGL(k_lp_decide, dim3(1), dim3(256), 0, s, ctr_.get(), P, c_.mailbox(), a_.labels,
                           inline_keep ? best_labels : nullptr, n);
        GP_LAUNCH_CHECK();
        if (!inline_keep) {
            hipLaunchKernelGGL(k_lp_keep, dim3(grid_for(ceil_div(n, 16), 256, 1024)), dim3(256), 0, s, L + kImproved,
                               a_.labels, best_labels, n, vec);
            GP_LAUNCH_CHECK();
        }
    }

    // The LP proposal and afterburner sweeps of a round (counter block clear).
    void propose_sweeps() {
        c_.timed("lp_propose", 0.0, [&] {
            if (narrow_bins())
                tiers(LpProposeBody<unsigned>{{a_}}, frontier_.get(), 0, ctr_.get() + kFrontier, cand_.get(),
                      ctr_.get() + kCand, ctr_.get() + kRowSum, a_.k, kTier, true);
            else
                tiers(LpProposeBody<unsigned long long>{{a_}}, frontier_.get(), 0, ctr_.get() + kFrontier, cand_.get(),
                      ctr_.get() + kCand, ctr_.get() + kRowSum, a_.k, kTier, true);
        });
        c_.timed("after", 0.0, [&] {
            if (a_.gains32)
                tiers(AfterBody<std::int32_t>{{a_}}, cand_.get(), 0, ctr_.get() + kCand, movers_.get(),
                      ctr_.get() + kMovers, nullptr, 0, kTierB, true, kEntAfter);
            else
                tiers(AfterBody<std::int64_t>{{a_}}, cand_.get(), 0, ctr_.get() + kCand, movers_.get(),
                      ctr_.get() + kMovers, nullptr, 0, kTierB, true, kEntAfter);
        });
    }

    // The LP round loop of the level with its control on the device
    // (k_lp_decide applies the host loop's rules after each round).  The
    // host issues each round before it has seen the previous one end
    // (ahead) or right after, watching only the mailbox; it takes over when
    // a round leaves a part over capacity, for the repair and the rest of
    // that round, and then re-arms the loop.  Same rounds, moves and
    // checkpoints as the host loop.
    void lp_loop(int rounds, bool graph_mode, std::uint64_t& cut, std::uint64_t& best_cut, Label* best_labels,
                 std::vector<std::uint64_t>& best_loads) {
        const hipStream_t s = c_.stream();
        const int k = a_.k;
        const bool vec = reinterpret_cast<std::uintptr_t>(a_.labels) % 16 == 0 &&
                         reinterpret_cast<std::uintptr_t>(best_labels) % 16 == 0;
        const LoopParams P{k,
                           rounds,
                           o_.stale_rounds,
                           graph_mode ? 1 : 0,
                           g_.row_len_bound <= static_cast<std::uint64_t>(kTierLight) ? 1 : 0,
                           c_.profiling() ? 1 : 0,
                           a_.cap,
                           significand_0999()};
        const int ahead = c_.loop_ahead();
        volatile unsigned long long* mail = c_.mailbox();
        std::uint64_t phase = 0, stale = 0, empty = 0, lp_moves = 0;
        bool copy_best = true;
        while (phase < static_cast<std::uint64_t>(rounds) && stale < static_cast<std::uint64_t>(o_.stale_rounds)) {
            reset_samples();
            const std::uint64_t phase0 = phase;
            hipLaunchKernelGGL(k_lp_arm, dim3(1), dim3(256), 0, s, ctr_.get(), k, 1ULL, phase, stale, empty, cut,
                               best_cut, lp_moves, copy_best ? 1 : 0);
            GP_LAUNCH_CHECK();
            mail[kRun] = 1;  // the stream is idle: no round of an earlier run is in flight
            std::vector<std::size_t> marks;
            int issued = 0;
            auto issue = [&] {
                marks.push_back(c_.pending_mark());
                round_issue(P, best_labels, vec);
                c_.mark(issued++);
            };
            issue();
            for (;;) {
                if (ahead) issue();
                c_.wait_mark(issued - 1 - ahead);
                if (!mail[kRun]) break;
                if (!ahead) issue();
            }
            // Issuing ahead, the stop may come from the round still in
            // flight: wait for the stream before reading the full state.
            if (ahead) c_.sync();
            std::vector<unsigned long long> h(static_cast<std::size_t>(kLoopWords + 2 * k));
            for (std::size_t i = 0; i < h.size(); ++i) h[i] = mail[i];
            ctr_clear_ = true;
            phase = h[kPhase];
            stale = h[kStale];
            empty = h[kEmpty];
            cut = h[kCut];
            best_cut = h[kBestCut];
            lp_moves = h[kLpMoves];
            for (int p = 0; p < k; ++p) {
                loads_host_[static_cast<std::size_t>(p)] = h[static_cast<std::size_t>(kLoopWords + p)];
                best_loads[static_cast<std::size_t>(p)] = h[static_cast<std::size_t>(kLoopWords + k + p)];
            }
            // Rounds issued after the stop did nothing: drop their intervals.
            const std::size_t ran = static_cast<std::size_t>(phase - phase0);
            if (ran < marks.size()) c_.drop_pending(marks[ran]);
            if (c_.profiling()) {
                const double w = a_.ew ? 4.0 : 0.0;
                c_.add_bytes("lp_propose", 28.0 * (double)h[kAccFrontier] + (8.0 + w) * (double)h[kAccRowSum]);
                const double g = a_.gains32 ? 4.0 : 0.0;  // 32-bit gains: 4 B less per row and entry
                c_.add_bytes("after", (16.0 - g) * (double)h[kAccAfterR] + (20.0 - g + w) * (double)h[kAccAfterE]);
                c_.add_bytes("update", 24.0 * (double)h[kAccUpdateR] + (16.0 + w) * (double)h[kAccUpdateE]);
            }
            const int reason = static_cast<int>(h[kReason]);
            if (reason == kCutUnderflow) throw std::runtime_error("incremental cut underflow");
            if (reason == kDeferred)
                throw std::logic_error("rows longer than the level's row-length bound were deferred and not processed");
            if (reason != kNeedRepair) break;
            // The rest of the round as the host loop runs it.
            if (!rebalance(cut)) break;
            publish_prices();
            ++stale;
            copy_best = false;
            if (cut < best_cut) {
                if (graph_mode) {
                    if (best_cut - cut > best_cut / 1000) stale = 0;
                } else if ((long double)cut < 0.999L * (long double)best_cut)
                    stale = 0;
                best_cut = cut;
                c_.copy_d2d(best_labels, a_.labels, static_cast<std::size_t>(a_.n) * sizeof(Label));
                best_loads = loads_host_;
                copy_best = true;
            }
        }
        st_.rounds = static_cast<int>(phase);
        st_.lp_moves += lp_moves;
    }

    // Update + commit of the current mover list.  zeroed: the delta and
    // third-sweep tier counters are still zero from the counter block's
    // clear at the start of the round or repair iteration.
    void apply_launch(bool lp, bool samples, bool zeroed) {
        const hipStream_t s = c_.stream();
        if (!zeroed) fill<unsigned long long>(c_, ctr_.get() + kDelta, 1, 0ULL);
        if (!zeroed && c_.profiling()) fill<unsigned long long>(c_, ctr_.get() + kEntUpdate, 2, 0ULL);
        c_.timed("update", 0.0, [&] {
            tiers(UpdateBody{{a_}, samples}, movers_.get(), 0, ctr_.get() + kMovers, nullptr, nullptr,
                  ctr_.get() + kDelta, 0, kTierC, zeroed, kEntUpdate);
            hipLaunchKernelGGL(k_commit, dim3(grid_for(a_.n, 256, 1024)), dim3(256), 0, s, a_, movers_.get(),
                               ctr_.get() + kMovers, lp);
            GP_LAUNCH_CHECK();
        });
        if (samples) samples_clear_ = false;
    }

    std::int64_t apply(bool lp, bool samples, std::int64_t& delta, bool zeroed = false) {
        apply_launch(lp, samples, zeroed);
        std::vector<unsigned long long> h(static_cast<std::size_t>(kCtrCount + a_.k));
        // The read also clears the counters (not the part loads behind them).
        c_.read_clear(h.data(), ctr_.get(), h.size(), kCtrCount);
        ctr_clear_ = true;
        // With no row above kTierLight entries no medium launch was made; a
        // row the light launch deferred would have gone unprocessed.
        if (g_.row_len_bound <= static_cast<std::uint64_t>(kTierLight) && (h[kTier] || h[kTierB] || h[kTierC]))
            throw std::logic_error("rows longer than the level's row-length bound were deferred and not processed");
        for (int p = 0; p < a_.k; ++p) loads_host_[static_cast<std::size_t>(p)] = h[static_cast<std::size_t>(kCtrCount + p)];
        delta = (std::int64_t)h[kDelta];
        if (c_.profiling()) {
            // Algorithmic bytes per row and per entry of each sweep (weighted
            // graphs add 4 B of edge weight per entry): LP proposal 28 + 8;
            // afterburner 16 + 20 (neighbour, target, gain, label; 4 B less
            // each with 32-bit gains); update and
            // commit 24 + 16 (neighbour, label, target, flag); repair proposal
            // 24 + 8 (offsets, label, weight, bucket and target out; neighbour
            // and its label per entry).
            const double w = a_.ew ? 4.0 : 0.0;
            if (lp) {
                c_.add_bytes("lp_propose", 28.0 * (double)h[kFrontier] + (8.0 + w) * (double)h[kRowSum]);
                const double g = a_.gains32 ? 4.0 : 0.0;
                c_.add_bytes("after", (16.0 - g) * (double)h[kEntAfter + 1] + (20.0 - g + w) * (double)h[kEntAfter]);
            }
            c_.add_bytes("update", 24.0 * (double)h[kEntUpdate + 1] + (16.0 + w) * (double)h[kEntUpdate]);
            if (samples)
                c_.add_bytes("repair_propose", 24.0 * (double)h[kEntRepair + 1] + (8.0 + w) * (double)h[kEntRepair]);
        }
        return (std::int64_t)h[kMovers];
    }

    bool update_cut(std::uint64_t& cut, std::int64_t delta) {
        if (delta >= 0) {
            if ((std::uint64_t)delta > cut) throw std::runtime_error("incremental cut underflow");
            cut -= (std::uint64_t)delta;
        } else
            cut += (std::uint64_t)(-delta);
        if (o_.strict_verify) verify(cut);
        return true;
    }

    // Strict mode (incremental.cuh:846-853): after every committed batch the
    // incremental cut must equal a full CSR recount of the labels and the
    // part loads a full recount of the vertex weights.
    void verify(std::uint64_t cut) {
        if (cut_of(c_, g_, a_.labels) != cut) throw std::runtime_error("incremental cut differs from CSR scan");
        std::vector<std::uint64_t> loads(static_cast<std::size_t>(a_.k));
        loads_of(c_, g_, a_.labels, a_.k, loads.data());
        if (loads != loads_host_) throw std::runtime_error("incremental part weights differ from a recount");
        ++st_.strict_checks;
    }

    // Sparse connection rows of the listed donors (ConnBody by the row
    // tiers over the positions [0, ne) of E).
    void conn_rows(const std::int32_t* E, std::int64_t ne, const std::uint8_t* skip, unsigned long long* vals,
                   std::uint8_t* parts, std::uint8_t* nnz) {
        // Within a repair iteration the second sweep's tier counters are
        // still zero from the iteration's counter clear (k_repair_begin).
        if (narrow_bins())
            tiers(ConnBody<unsigned>{{a_}, E, skip, vals, parts, nnz}, nullptr, ne, nullptr, nullptr, nullptr, nullptr,
                  a_.k, kTierB, true);
        else
            tiers(ConnBody<unsigned long long>{{a_}, E, skip, vals, parts, nnz}, nullptr, ne, nullptr, nullptr, nullptr,
                  nullptr, a_.k, kTierB, true);
    }

    // Receiver reservation over the flagged donors L (vertex order).
    void reserve(const std::int32_t* L, std::int64_t nl, const unsigned long long* prefix, const std::int32_t* bstar,
                 const unsigned long long* excess) {
        const hipStream_t s = c_.stream();
        const int k = a_.k;
        const int nbl = static_cast<int>(ceil_div(nl, kKeyTile));
        DBuf<std::int32_t> keys = c_.buf<std::int32_t>(static_cast<std::size_t>(nl));
        DBuf<unsigned long long> S = c_.buf<unsigned long long>(static_cast<std::size_t>(nbl) * k);
        DBuf<std::uint8_t> elig = c_.buf<std::uint8_t>(static_cast<std::size_t>(nl));
        hipLaunchKernelGGL(k_elig_keys, dim3(nbl), dim3(kTierThreads), 0, s, a_, L, nl, bstar, nbl, keys.get(),
                           elig.get(), S.get());
        GP_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_keyed_scan, dim3(1), dim3(kKeyScanThreads), 0, s, S.get(), nbl, k, nullptr);
        GP_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_keyed_apply<EligFn>), dim3(nbl), dim3(kTierThreads), 0, s, keys.get(), nl, S.get(), nbl,
                           EligFn{a_, L, prefix, excess, elig.get()}, nullptr, nullptr);
        GP_LAUNCH_CHECK();
        DBuf<std::int32_t> idx = c_.buf<std::int32_t>(static_cast<std::size_t>(nl));
        DBuf<std::int64_t> cnt = c_.buf<std::int64_t>(1);
        stable_select(c_, elig.get(), nl, idx.get(), cnt.get());
        const std::int64_t ne = c_.read_scalar(cnt.get());
        if (ne <= 0) return;
        DBuf<std::int32_t> E = c_.buf<std::int32_t>(static_cast<std::size_t>(ne));
        reserve_donors(E.get(), L, idx.get(), ne);
    }

    // Reservation over the ne eligible donors E[j] = L[idx[j]] (E = L when
    // idx is null: the list is already the eligible donors).
    void reserve_donors(std::int32_t* E, const std::int32_t* L, const std::int32_t* idx, std::int64_t ne) {
        const hipStream_t s = c_.stream();
        const int k = a_.k;
        DBuf<std::int32_t> keys = c_.buf<std::int32_t>(static_cast<std::size_t>(ne));
        DBuf<unsigned long long> S =
            c_.buf<unsigned long long>(static_cast<std::size_t>(ceil_div(ne, kKeyTile)) * k);
        // Connection rows are read from the second attempt on, and only for
        // donors still waiting; they are built once attempt 0 is done.
        DBuf<unsigned long long> conn = c_.buf<unsigned long long>(static_cast<std::size_t>(ne) * k);
        DBuf<std::uint8_t> conn_part = c_.buf<std::uint8_t>(static_cast<std::size_t>(ne) * k);
        DBuf<std::uint8_t> conn_nnz = c_.buf<std::uint8_t>(static_cast<std::size_t>(ne));
        DBuf<unsigned long long> requests = c_.buf<unsigned long long>(static_cast<std::size_t>(k));
        DBuf<unsigned long long> slack = c_.buf<unsigned long long>(static_cast<std::size_t>(k));
        DBuf<std::uint8_t> done = c_.buf<std::uint8_t>(static_cast<std::size_t>(ne));
        DBuf<unsigned long long> tried = c_.buf<unsigned long long>(static_cast<std::size_t>(ne) * 2);
        DBuf<DonorInfo> live = c_.buf<DonorInfo>(static_cast<std::size_t>(ne));
        hipLaunchKernelGGL(k_reserve_setup, dim3(grid_for(ne)), dim3(256), 0, s, a_, L, idx, ne, E, live.get(),
                           done.get(), tried.get(), slack.get(), requests.get());
        GP_LAUNCH_CHECK();
        ReserveState r{E,           ne,         conn.get(), conn_part.get(), conn_nnz.get(), requests.get(),
                       slack.get(), done.get(), tried.get(), live.get(),     ne};
        // Attempts run in groups between host checks: an attempt without a
        // chooser changes nothing (no tried bits, no requests, no admissions),
        // so the attempts after the last active one are no-ops.  Between
        // groups the live list drops donors that are admitted or have no
        // part left; it stays in ascending order, so ranks are unchanged.
        constexpr int kAttemptGroup = 4;
        constexpr std::int64_t kCompactMin = 4096;
        DBuf<unsigned long long> active = c_.buf<unsigned long long>(kAttemptGroup);
        if (ne <= kTailDonors) {
            // Every donor's connection row first (the rows of donors admitted
            // in attempt 0 go unused), then all attempts in one launch.
            conn_rows(E, ne, nullptr, conn.get(), conn_part.get(), conn_nnz.get());
            GP_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_reserve_tail, dim3(1), dim3(kTailThreads), 0, s, a_, r, 0, movers_.get(),
                               ctr_.get() + kMovers);
            GP_LAUNCH_CHECK();
            return;
        }
        for (int attempt = 0; attempt < k;) {
            const int group = std::min(kAttemptGroup, k - attempt);
            const std::int64_t nlv = r.nlive;
            const int nb = static_cast<int>(ceil_div(nlv, kKeyTile));
            fill<unsigned long long>(c_, active.get(), kAttemptGroup, 0ULL);
            for (int j = 0; j < group; ++j, ++attempt) {
                if (attempt == 1) {
                    conn_rows(E, ne, done.get(), conn.get(), conn_part.get(), conn_nnz.get());
                    GP_LAUNCH_CHECK();
                }
                // Blocks stride over the donors: the per-block room table and
                // part order are set up once per block.
                hipLaunchKernelGGL(k_choose, dim3(grid_for(nlv, kTierThreads / kChooseLanes, 1024)), dim3(kTierThreads), 0, s, a_,
                                   r, attempt, keys.get(), active.get() + j);
                GP_LAUNCH_CHECK();
                hipLaunchKernelGGL((k_keyed_sums<DonorWeight>), dim3(nb), dim3(kTierThreads), 0, s, keys.get(), nlv, k,
                                   nb, DonorWeight{r.live}, S.get());
                GP_LAUNCH_CHECK();
                // Ranks start at the start-of-attempt totals; totals advance.
                hipLaunchKernelGGL(k_keyed_scan, dim3(1), dim3(kKeyScanThreads), 0, s, S.get(), nb, k, requests.get());
                GP_LAUNCH_CHECK();
                hipLaunchKernelGGL((k_keyed_apply<AdmitFn>), dim3(nb), dim3(kTierThreads), 0, s, keys.get(), nlv,
                                   S.get(), nb, AdmitFn{a_, r}, movers_.get(), ctr_.get() + kMovers);
                GP_LAUNCH_CHECK();
            }
            if (!c_.read_scalar(active.get() + group - 1)) break;
            if (nlv >= kCompactMin && attempt < k) {
                DBuf<std::uint8_t> lf = c_.buf<std::uint8_t>(static_cast<std::size_t>(nlv));
                DBuf<std::int32_t> li = c_.buf<std::int32_t>(static_cast<std::size_t>(nlv));
                DBuf<std::int64_t> lc = c_.buf<std::int64_t>(1);
                hipLaunchKernelGGL(k_live_flags, dim3(grid_for(nlv)), dim3(256), 0, s, r.live, nlv, done.get(), lf.get());
                GP_LAUNCH_CHECK();
                stable_select(c_, lf.get(), nlv, li.get(), lc.get());
                const std::int64_t next = c_.read_scalar(lc.get());
                if (next == 0) break;
                DBuf<DonorInfo> live2 = c_.buf<DonorInfo>(static_cast<std::size_t>(next));
                hipLaunchKernelGGL(k_gather_info, dim3(grid_for(next)), dim3(256), 0, s, r.live, li.get(), next,
                                   live2.get());
                GP_LAUNCH_CHECK();
                live.swap(live2);
                r.live = live.get();
                r.nlive = next;
            }
            // A donor list that fits one block finishes in one launch.
            if (r.nlive <= kTailDonors && attempt < k) {
                hipLaunchKernelGGL(k_reserve_tail, dim3(1), dim3(kTailThreads), 0, s, a_, r, attempt, movers_.get(),
                                   ctr_.get() + kMovers);
                GP_LAUNCH_CHECK();
                break;
            }
        }
    }

    // Capacity repair loop (rebalance_graph.cuh:160-280); returns feasibility.
    bool rebalance(std::uint64_t& cut) {
        const hipStream_t s = c_.stream();
        const std::int64_t n = a_.n;
        const int k = a_.k;
        int iterations = 0;
        bool feasible = false;
        std::int64_t count = 0;
        DBuf<unsigned long long> prefix = c_.buf<unsigned long long>(static_cast<std::size_t>(k) * kBalanceWidth);
        DBuf<std::int32_t> bstar = c_.buf<std::int32_t>(static_cast<std::size_t>(k));
        DBuf<unsigned long long> excess = c_.buf<unsigned long long>(static_cast<std::size_t>(k));
        DBuf<std::uint8_t> f = c_.buf<std::uint8_t>(n);
        DBuf<std::int32_t> list = c_.buf<std::int32_t>(n);
        DBuf<std::int64_t> lcount = c_.buf<std::int64_t>(1);
        const std::int64_t ncounts = (std::int64_t)k * kBalanceWidth;
        do {
            static_assert(kCtrCount <= kBalanceWidth, "the counter block clears inside the bucket clear");
            // The counter block is zero after a read-back that cleared it
            // (every iteration's last); then one launch starts the iteration.
            const bool zeroed = ctr_clear_;
            ctr_clear_ = false;
            const int donor_grid = grid_for(n, 256 * kScanRounds, 2048);
            if (zeroed) {
                hipLaunchKernelGGL(k_repair_donors, dim3(std::max(donor_grid, grid_for(ncounts))), dim3(256), 0, s, a_,
                                   cand_.get(), ctr_.get() + kDonors, loads0_.get(), counts_.get(), ncounts);
            } else {
                hipLaunchKernelGGL(k_repair_begin, dim3(grid_for(ncounts)), dim3(256), 0, s, a_, loads0_.get(),
                                   counts_.get(), ncounts, ctr_.get(), (int)kCtrCount);
                GP_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_repair_donors, dim3(donor_grid), dim3(256), 0, s, a_, cand_.get(),
                                   ctr_.get() + kDonors, nullptr, nullptr, 0);
            }
            GP_LAUNCH_CHECK();
            c_.timed("repair_propose", 0.0, [&] {
                if (narrow_bins())
                    tiers(RepairProposeBody<unsigned>{{a_}}, cand_.get(), 0, ctr_.get() + kDonors, nullptr, nullptr,
                          nullptr, a_.k, kTier, true, kEntRepair);
                else
                    tiers(RepairProposeBody<unsigned long long>{{a_}}, cand_.get(), 0, ctr_.get() + kDonors, nullptr,
                          nullptr, nullptr, a_.k, kTier, true, kEntRepair);
            });
            hipLaunchKernelGGL(k_bucket_scan, dim3(k), dim3(kScanThreads), 0, s, a_, prefix.get(), bstar.get(),
                               excess.get());
            GP_LAUNCH_CHECK();
            if (n <= kSelectSmallMax) {
                // Flagging, eligibility and both compactions in one block.
                hipLaunchKernelGGL(k_donors_small, dim3(1), dim3(kDonorThreads), 0, s, a_, bstar.get(), prefix.get(),
                                   excess.get(), list.get(), lcount.get());
                GP_LAUNCH_CHECK();
                const std::int64_t ne = c_.read_scalar(lcount.get());
                if (ne > 0) reserve_donors(list.get(), list.get(), nullptr, ne);
            } else {
                hipLaunchKernelGGL(k_repair_flag, dim3(grid_for(n)), dim3(256), 0, s, a_, bstar.get(), f.get());
                GP_LAUNCH_CHECK();
                stable_select(c_, f.get(), n, list.get(), lcount.get());
                const std::int64_t nl = c_.read_scalar(lcount.get());
                if (nl > 0) reserve(list.get(), nl, prefix.get(), bstar.get(), excess.get());
            }
            hipLaunchKernelGGL(k_repair_clear, dim3(grid_for(n)), dim3(256), 0, s, a_, cand_.get(),
                               ctr_.get() + kDonors);
            GP_LAUNCH_CHECK();
            std::int64_t delta = 0;
            count = apply(false, true, delta, true);
            ++iterations;
            if (count) {
                update_cut(cut, delta);
                st_.repair_moves += static_cast<std::uint64_t>(count);
            }
            feasible = max_load() <= a_.cap;
        } while (!feasible && count && iterations < 2 * k + 4);
        st_.repair_iterations += iterations;
        return feasible;
    }

    // Capacity-complementary exchange (incremental.cuh:974-1041).
    void exchange(int passes, std::uint64_t& cut, std::uint64_t& best_cut) {
        const hipStream_t s = c_.stream();
        ctr_clear_ = false;
        const std::int64_t n = a_.n;
        const int k = a_.k;
        int sw = std::min(kJointSeeds, std::max(32, 4096 / k));
        const int slots = k * sw * 2;
        DBuf<unsigned long long> A = c_.buf<unsigned long long>(static_cast<std::size_t>(n) * k);
        DBuf<unsigned long long> Q = c_.buf<unsigned long long>(static_cast<std::size_t>(k) * k);
        DBuf<std::int32_t> visits = c_.buf<std::int32_t>(static_cast<std::size_t>(k) * k);
        DBuf<std::int32_t> partner = c_.buf<std::int32_t>(static_cast<std::size_t>(k));
        DBuf<unsigned long long> seeds = c_.buf<unsigned long long>(static_cast<std::size_t>(k) * sw);
        DBuf<JMove> grown = c_.buf<JMove>(static_cast<std::size_t>(slots));
        DBuf<JMove> archive = c_.buf<JMove>(static_cast<std::size_t>(slots) * kJointArchive);
        DBuf<std::int32_t> owners = c_.buf<std::int32_t>(n);
        DBuf<std::int32_t> stops = c_.buf<std::int32_t>(static_cast<std::size_t>(k));
        auto dense = [&] {
            hipLaunchKernelGGL((k_dense_affinity<64>), dim3(grid_for(n, 4, 4096)), dim3(256), 4 * k * sizeof(unsigned long long), s, a_,
                               A.get());
            GP_LAUNCH_CHECK();
        };
        dense();
        fill<unsigned long long>(c_, Q.get(), (std::int64_t)k * k, 0ULL);
        fill<std::int32_t>(c_, visits.get(), (std::int64_t)k * k, 0);
        hipLaunchKernelGGL(k_quotient, dim3(grid_for(n)), dim3(256), 0, s, a_, A.get(), Q.get());
        GP_LAUNCH_CHECK();
        int empty = 0;
        for (int pass = 0; pass < passes; ++pass) {
            hipLaunchKernelGGL(k_joint_match, dim3(1), dim3(kMatchThreads), 0, s, k, Q.get(), partner.get(), visits.get(),
                               seeds.get(), k * sw);
            GP_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_joint_seeds, dim3(grid_for(n)), dim3(256), 0, s, a_, A.get(), partner.get(),
                               seeds.get(), sw, pass);
            GP_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_joint_grow, dim3(ceil_div(slots, kGrowWaves)), dim3(kGrowWaves * kWave), 0, s, a_,
                               A.get(), slots, sw,
                               seeds.get(), partner.get(), grown.get(), archive.get());
            GP_LAUNCH_CHECK();
            // Archive: non-empty, sorted by net order, unique.
            const int an = slots * kJointArchive;
            DBuf<JMove> frag;
            const int fragments = compact_sort(archive.get(), an, NetOrder{}, frag, true);
            const int pool_cap = slots + 2 * fragments;
            DBuf<JMove> pool = c_.buf<JMove>(static_cast<std::size_t>(std::max(pool_cap, 1)));
            c_.copy_d2d(pool.get(), grown.get(), static_cast<std::size_t>(slots) * sizeof(JMove));
            if (fragments) {
                hipLaunchKernelGGL(k_joint_combine, dim3(ceil_div(fragments, 4)), dim3(256), 0, s, a_, fragments,
                                   frag.get(), slots, pool.get() + slots);
                GP_LAUNCH_CHECK();
            }
            DBuf<JMove> moves;
            const int active = compact_sort(pool.get(), pool_cap, JointOrder{}, moves, false);
            if (!active) {
                if (++empty >= 2) break;
                continue;
            }
            fill<std::int32_t>(c_, owners.get(), n, INT_MAX);
            hipLaunchKernelGGL(k_joint_claim, dim3(ceil_div(active, 256)), dim3(256), 0, s, moves.get(), active,
                               owners.get());
            hipLaunchKernelGGL(k_joint_resolve, dim3(ceil_div(active, 256)), dim3(256), 0, s, moves.get(), active,
                               owners.get());
            hipLaunchKernelGGL(k_exchange_reset, dim3(grid_for(n)), dim3(256), 0, s, n, owners.get(), state_.get());
            DBuf<unsigned long long> events = c_.buf<unsigned long long>(static_cast<std::size_t>(active));
            hipLaunchKernelGGL(k_joint_mark, dim3(ceil_div(active, 256)), dim3(256), 0, s, a_, moves.get(), active,
                               owners.get(), events.get());
            GP_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_joint_events, dim3(ceil_div(active, kEventTx)), dim3(kEventThreads), 0, s, a_,
                               moves.get(), active,
                               owners.get(), events.get());
            hipLaunchKernelGGL(k_joint_prefix, dim3(ceil_div(k, 4)), dim3(256), 0, s, a_, moves.get(), active,
                               events.get(), stops.get());
            // The accept also clears the mover count and every counter the
            // update and commit of this batch use (delta, its tier lists,
            // profiled entries).
            hipLaunchKernelGGL(k_joint_accept, dim3(ceil_div(active, 256)), dim3(256), 0, s, a_, moves.get(), active,
                               stops.get(), (int)kMovers, (int)kCtrCount);
            hipLaunchKernelGGL(k_compact_flags<StateLane<std::uint8_t>>, dim3(grid_for(n, 256 * kScanRounds, 2048)), dim3(256), 0, s,
                               n, a_.flags,
                               movers_.get(), ctr_.get() + kMovers, nullptr);
            GP_LAUNCH_CHECK();
            const unsigned long long cnt = c_.read_scalar(ctr_.get() + kMovers);
            if (!cnt) {
                if (++empty >= 2) break;
                continue;
            }
            empty = 0;
            std::int64_t delta = 0;
            apply(false, false, delta, true);
            if (delta <= 0) throw std::runtime_error("non-improving boundary group batch");
            update_cut(cut, delta);
            st_.group_moves += cnt;
            if (max_load() > a_.cap || cut >= best_cut) throw std::runtime_error("capacity exchange violated cut/capacity");
            best_cut = cut;
            dense();
        }
        // Later passes wrote the mover counter after the last read.
        ctr_clear_ = false;
    }

    // Stable compaction of non-empty moves, sort by cmp, optional unique.
    template <typename Cmp>
    int compact_sort(const JMove* src, int count, Cmp cmp, DBuf<JMove>& out, bool unique) {
        const hipStream_t s = c_.stream();
        if (count <= 0) return 0;
        DBuf<std::uint8_t> f = c_.buf<std::uint8_t>(static_cast<std::size_t>(count));
        DBuf<std::int32_t> idx = c_.buf<std::int32_t>(static_cast<std::size_t>(count));
        DBuf<std::int64_t> cnt = c_.buf<std::int64_t>(1);
        hipLaunchKernelGGL(k_nonempty_moves, dim3(ceil_div(count, 256)), dim3(256), 0, s, src, count, f.get());
        GP_LAUNCH_CHECK();
        stable_select(c_, f.get(), count, idx.get(), cnt.get());
        const int m = (int)c_.read_scalar(cnt.get());
        if (!m) return 0;
        int N = 1;
        while (N < m) N <<= 1;
        DBuf<std::int32_t> sidx = c_.buf<std::int32_t>(static_cast<std::size_t>(N));
        hipLaunchKernelGGL(k_pad_index, dim3(ceil_div(N, 256)), dim3(256), 0, s, idx.get(), m, sidx.get(), N);
        GP_LAUNCH_CHECK();
        if (N <= 1024) {
            hipLaunchKernelGGL((k_bitonic_local<Cmp>), dim3(1), dim3(1024), 0, s, sidx.get(), N, src, cmp);
            GP_LAUNCH_CHECK();
        } else {
            // N is a power of two above 1024: chunk-local stages in LDS, one
            // launch per cross-chunk stage.
            const int chunks = N / kBitonicChunk;
            hipLaunchKernelGGL((k_bitonic_chunk<Cmp>), dim3(chunks), dim3(1024), 0, s, sidx.get(), 0, src, cmp);
            GP_LAUNCH_CHECK();
            for (int kk = 2 * kBitonicChunk; kk <= N; kk <<= 1) {
                for (int j = kk >> 1; j >= kBitonicChunk; j >>= 1) {
                    hipLaunchKernelGGL((k_bitonic_global<Cmp>), dim3(ceil_div(N, 256)), dim3(256), 0, s, sidx.get(), N,
                                       j, kk, src, cmp);
                    GP_LAUNCH_CHECK();
                }
                hipLaunchKernelGGL((k_bitonic_chunk<Cmp>), dim3(chunks), dim3(1024), 0, s, sidx.get(), kk, src, cmp);
                GP_LAUNCH_CHECK();
            }
        }
        DBuf<JMove> sorted = c_.buf<JMove>(static_cast<std::size_t>(m));
        hipLaunchKernelGGL(k_gather_moves, dim3(ceil_div(m, 256)), dim3(256), 0, s, src, sidx.get(), m, sorted.get());
        GP_LAUNCH_CHECK();
        if (!unique) {
            out = std::move(sorted);
            return m;
        }
        DBuf<std::uint8_t> uf = c_.buf<std::uint8_t>(static_cast<std::size_t>(m));
        DBuf<std::int32_t> uidx = c_.buf<std::int32_t>(static_cast<std::size_t>(m));
        hipLaunchKernelGGL(k_unique_flags, dim3(ceil_div(m, 256)), dim3(256), 0, s, sorted.get(), m, uf.get());
        GP_LAUNCH_CHECK();
        stable_select(c_, uf.get(), m, uidx.get(), cnt.get());
        const int u = (int)c_.read_scalar(cnt.get());
        out = c_.buf<JMove>(static_cast<std::size_t>(u));
        hipLaunchKernelGGL(k_gather_moves, dim3(ceil_div(u, 256)), dim3(256), 0, s, sorted.get(), uidx.get(), u,
                           out.get());
        GP_LAUNCH_CHECK();
        return u;
    }

    void publish_prices() {
        if (!prices_) return;
        hipLaunchKernelGGL(k_publish_prices, dim3(1), dim3(kMaxParts), 0, c_.stream(), samples_.get(), prices_.get(),
                           a_.k);
        GP_LAUNCH_CHECK();
        samples_clear_ = true;
    }

    void reset_samples() {
        if (samples_ && !samples_clear_)
            fill<unsigned long long>(c_, reinterpret_cast<unsigned long long*>(samples_.get()), 2 * a_.k, 0ULL);
        samples_clear_ = true;
    }

    Context& c_;
    const DGraph& g_;
    std::vector<std::uint64_t>& loads_host_;
    const RefineOptions& o_;
    LevelStats& st_;
    RefArgs a_;
    DBuf<std::uint32_t> state_;
    DBuf<std::int32_t> frontier_, cand_, movers_, medium_, heavy_, bucket_;
    DBuf<std::int64_t> gains_;
    DBuf<std::int32_t> gains32_;
    DBuf<unsigned long long> ctr_, loads0_, counts_;
    DBuf<double> prices_;
    DBuf<RepairSample> samples_;
    // Host knowledge that the counter block (first kCtrCount words) or the
    // repair samples are zero on the device, so their clear can be skipped.
    bool ctr_clear_ = false;
    bool samples_clear_ = false;
};

}  // namespace

namespace {

LevelStats refine_level_impl(Context& c, const DGraph& g, Label* labels, std::vector<std::uint64_t>& loads,
                             const RefineOptions& o, std::uint64_t cap, int rounds, bool has_known_cut,
                             std::uint64_t known_cut) {
    LevelStats st;
    st.n = g.n;
    st.m = g.m;
    st.refined = 1;
    const int k = o.parts;
    LevelRefiner R(c, g, labels, loads, o, cap, st);
    std::uint64_t cut = has_known_cut ? known_cut : cut_of(c, g, labels);
    if (o.strict_verify && has_known_cut) {
        // incremental.cuh:777-778: the carried cut must equal a recount.
        if (cut_of(c, g, labels) != known_cut) throw std::runtime_error("known projection cut differs from CSR scan");
        ++st.strict_checks;
    }
    st.initial_cut = cut;
    if (R.max_load() > cap) {
        R.reset_samples();
        if (!R.rebalance(cut)) {
            st.initial_repair_failed = 1;
            st.final_cut = cut;
            return st;
        }
    }
    const std::uint64_t feasible_start = cut;
    std::uint64_t best_cut = cut;
    DBuf<Label> best_labels = c.buf<Label>(static_cast<std::size_t>(g.n));
    std::vector<std::uint64_t> best_loads = loads;
    c.copy_d2d(best_labels.get(), labels, static_cast<std::size_t>(g.n) * sizeof(Label));
    const bool graph_mode = g.n >= o.graph_loop_threshold;
    int stale = 0, empty = 0;
    // Strict mode recounts on the host after every batch: host loop.
    const bool device_loop = o.device_loop && c.device_loop() && !o.strict_verify;
    if (device_loop) R.lp_loop(rounds, graph_mode, cut, best_cut, best_labels.get(), best_loads);
    const int host_rounds = device_loop ? 0 : rounds;
    for (int phase = 0; phase < host_rounds && stale < o.stale_rounds; ++phase) {
        st.rounds = phase + 1;
        R.reset_samples();
        std::int64_t delta = 0;
        const std::int64_t count = R.lp_round(delta);
        if (!count) {
            if (++empty >= 2) break;
            ++stale;
            continue;
        }
        empty = 0;
        R.update_cut(cut, delta);
        st.lp_moves += static_cast<std::uint64_t>(count);
        if (R.max_load() > cap && !R.rebalance(cut)) break;
        R.publish_prices();
        ++stale;
        if (cut < best_cut) {
            if (graph_mode) {
                if (best_cut - cut > best_cut / 1000) stale = 0;
            } else if ((long double)cut < 0.999L * (long double)best_cut)
                stale = 0;
            best_cut = cut;
            c.copy_d2d(best_labels.get(), labels, static_cast<std::size_t>(g.n) * sizeof(Label));
            best_loads = loads;
        }
    }
    c.copy_d2d(labels, best_labels.get(), static_cast<std::size_t>(g.n) * sizeof(Label));
    loads = best_loads;
    c.upload(R.loads_dev(), loads.data(), static_cast<std::size_t>(k) * sizeof(std::uint64_t));
    cut = best_cut;
    if (rounds > 0 && o.exchange_rounds > 0 && g.n <= o.exchange_max_n) {
        R.clear_moves(c);
        R.exchange(o.exchange_rounds, cut, best_cut);
    }
    st.final_cut = cut;
    std::uint64_t mx = 0;
    for (auto x : loads) mx = std::max(mx, x);
    if (cut != best_cut || cut > feasible_start || mx > cap)
        throw std::runtime_error("incremental refinement postcondition failed");
    if (o.strict_verify) {
        // The restored checkpoint: its cut and loads against recounts.
        R.verify(cut);
    }
    return st;
}

}  // namespace

bool debug_below_0999(std::uint64_t cut, std::uint64_t best) { return below_0999(cut, best, significand_0999()); }

LevelStats refine_level(Context& c, const DGraph& g, std::int32_t* labels, std::vector<std::uint64_t>& loads,
                        const RefineOptions& o, std::uint64_t cap, int rounds, bool has_known_cut,
                        std::uint64_t known_cut) {
    DBuf<Label> lab = c.buf<Label>(static_cast<std::size_t>(g.n));
    convert(c, g.n, labels, lab.get());
    LevelStats st = refine_level_impl(c, g, lab.get(), loads, o, cap, rounds, has_known_cut, known_cut);
    convert(c, g.n, lab.get(), labels);
    return st;
}

UncoarsenResult uncoarsen(Context& c, const Hierarchy& h, const std::int32_t* coarsest_labels, const RefineOptions& o) {
    const int k = o.parts;
    UncoarsenResult res;
    const DGraph& cg = h.level(h.levels() - 1);
    DBuf<Label> labels = c.buf<Label>(static_cast<std::size_t>(cg.n));
    convert(c, cg.n, coarsest_labels, labels.get());
    std::vector<std::uint64_t> loads(static_cast<std::size_t>(k), 0);
    loads_of(c, cg, labels.get(), k, loads.data());
    std::uint64_t total = 0;
    for (auto x : loads) total += x;
    const long double raw = (long double)total * o.imbalance / k;
    const std::uint64_t cap = (std::uint64_t)std::floor(raw);
    if (cap < total / (unsigned)k + (total % (unsigned)k != 0))
        throw std::invalid_argument("integer partition capacity cannot hold total weight");
    LevelStats st = refine_level_impl(c, cg, labels.get(), loads, o, cap, o.max_rounds, false, 0);
    res.levels.push_back(st);
    std::uint64_t last_cut = st.final_cut;
    std::int64_t last_refined = cg.n;
    for (int lvl = h.levels() - 2; lvl >= 0; --lvl) {
        const DGraph& fg = h.level(lvl);
        DBuf<Label> fl = c.buf<Label>(static_cast<std::size_t>(fg.n));
        if (fg.n > 0) {
            hipLaunchKernelGGL((k_project<Label>), dim3(grid_for(fg.n)), dim3(256), 0, c.stream(), fg.n,
                               h.maps[static_cast<std::size_t>(lvl)].get(), labels.get(), fl.get());
            GP_LAUNCH_CHECK();
        }
        labels = std::move(fl);
        LevelStats ls;
        ls.n = fg.n;
        ls.m = fg.m;
        if (o.strict_verify) {
            // refine.cu:641-659: projection preserves the cut and the part loads.
            std::vector<std::uint64_t> projected(static_cast<std::size_t>(k));
            loads_of(c, fg, labels.get(), k, projected.data());
            if (cut_of(c, fg, labels.get()) != last_cut || projected != loads)
                throw std::runtime_error("GPU projection failed cut or part-weight preservation");
            ++ls.strict_checks;
        }
        std::uint64_t maxload = 0;
        for (auto x : loads) maxload = std::max(maxload, x);
        const bool refine = lvl == 0 || maxload > cap || (double)fg.n >= o.growth_ratio * (double)last_refined;
        if (refine) {
            const std::uint64_t projection_checks = ls.strict_checks;
            ls = refine_level_impl(c, fg, labels.get(), loads, o, cap, o.max_rounds, true, last_cut);
            ls.strict_checks += projection_checks;
            if (!ls.initial_repair_failed) last_refined = fg.n;
        } else {
            ls.initial_cut = ls.final_cut = last_cut;
        }
        last_cut = ls.final_cut;
        res.levels.push_back(ls);
    }
    std::uint64_t maxload = 0;
    for (auto x : loads) maxload = std::max(maxload, x);
    if (maxload > cap) throw std::runtime_error("final original partition violates capacity");
    const DGraph& g0 = h.level(0);
    res.labels = c.buf<std::int32_t>(static_cast<std::size_t>(g0.n));
    convert(c, g0.n, labels.get(), res.labels.get());
    res.loads = loads;
    res.cut = last_cut;
    return res;
}

}  // namespace gp
