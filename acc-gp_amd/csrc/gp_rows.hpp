// Three-tier row processing for skewed degree distributions.
//
// A "body" processes one CSR row with a team of lanes: a 16-lane group
// (rows <= kTierLight entries), a full wave (rows <= kTierMedium) or a
// 256-thread block (longer rows).  Light kernels defer longer rows to a
// medium list and medium kernels defer to a heavy list, with wave-aggregated
// appends (one atomic per wave).  Loops are wave-uniform so every lane of a
// wave reaches the ballots.
#pragma once
#include "gp_prims.hpp"

#include <type_traits>
#include <utility>

namespace gp {

constexpr int kTierLight = 64;
constexpr int kTierMedium = 4096;
constexpr int kTierThreads = 256;
constexpr int kHeavyThreads = 512;  // one block per heavy row: many entries in flight
constexpr int kBlockRed = 3 * (kHeavyThreads / kWave);  // u64 words of block-team reduction scratch

// Appends v for every lane with pred set; all lanes of the wave must call it.
__device__ inline void wave_append(bool pred, std::int32_t v, std::int32_t* list, unsigned long long* count) {
    const unsigned long long mask = __ballot(pred);
    if (!mask) return;
    const int lane = lane_id();
    const int leader = __ffsll(mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    if (pred) list[base + __popcll(mask & ((1ULL << lane) - 1ULL))] = v;
}

// Wave-private LDS staging of list appends: one global atomic per kStage
// entries instead of one per wave step.  List order is unspecified; every
// list built this way is consumed by per-vertex work.  All lanes of the
// wave call push() and flush() together.
constexpr int kStage = 256;

struct Stage {
    std::int32_t* buf;  // kStage LDS entries owned by one wave
    std::int32_t* list;
    unsigned long long* count;
    int fill;

    __device__ void flush() {
        if (!fill) return;
        wave_sync();
        unsigned long long base = 0;
        if (lane_id() == 0) base = atomicAdd(count, (unsigned long long)fill);
        base = __shfl(base, 0);
        for (int i = lane_id(); i < fill; i += kWave) list[base + i] = buf[i];
        wave_sync();
        fill = 0;
    }
    __device__ void push(bool pred, std::int32_t v) {
        if (!list) return;
        const unsigned long long mask = __ballot(pred);
        if (!mask) return;
        const int cnt = __popcll(mask);
        if (fill + cnt > kStage) flush();
        if (pred) buf[fill + __popcll(mask & ((1ULL << lane_id()) - 1ULL))] = v;
        fill += cnt;
    }
};

template <int W>
struct GroupTeam {
    int lg;
    int wave_rows_max = 1 << 30;           // longest row any team of this wave holds (wave-uniform)
    const unsigned long long* part = nullptr;  // block-shared staged table (or null)
    static constexpr int size = W;
    __device__ int rank() const { return lg; }
    template <typename T>
    __device__ T sum(T x) const { return group_sum<W>(x); }
    __device__ void sync() const { wave_sync(); }
    __device__ bool leader() const { return lg == 0; }
    template <typename T>
    __device__ T bcast(T x) const { return __shfl(x, 0, W); }
};

struct BlockTeam {
    unsigned long long* red;  // LDS scratch of kBlockRed entries
    const unsigned long long* part = nullptr;  // block-shared staged table (or null)
    static constexpr int size = kHeavyThreads;
    __device__ int rank() const { return threadIdx.x; }
    template <typename T>
    __device__ T sum(T x) const {
        x = wave_sum(x);
        __syncthreads();
        if (lane_id() == 0) red[threadIdx.x / kWave] = (unsigned long long)x;
        __syncthreads();
        T s = 0;
        for (int i = 0; i < size / kWave; ++i) s += (T)red[i];
        __syncthreads();
        return s;
    }
    __device__ void sync() const { __syncthreads(); }
    __device__ bool leader() const { return threadIdx.x == 0; }
    template <typename T>
    __device__ T bcast(T x) const {
        static_assert(sizeof(T) <= sizeof(unsigned long long), "bcast of a word");
        __syncthreads();
        if (threadIdx.x == 0) red[0] = (unsigned long long)x;
        __syncthreads();
        const T r = (T)red[0];
        __syncthreads();
        return r;
    }
};

// Visits entries [b, e) of a row with the team's stride and kBatch entries
// per lane in flight: every load(x) of a batch is issued before the first
// use(x, record), so dependent gathers overlap instead of serialising.
constexpr int kBatch = 4;

template <class Team, class Load, class Use>
__device__ inline void row_batches(const Team& t, std::int64_t b, std::int64_t e, Load load, Use use) {
    for (std::int64_t x0 = b + t.rank(); x0 < e; x0 += kBatch * Team::size) {
        using Rec = decltype(load(x0));
        Rec rec[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const std::int64_t x = x0 + j * Team::size;
            rec[j] = load(x < e ? x : e - 1);
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const std::int64_t x = x0 + j * Team::size;
            if (x < e) use(x, rec[j]);
        }
    }
}

// Bodies may stage a block-shared, read-only table at kernel start:
// part_words() u64 words, filled cooperatively by stage(); teams carry a
// pointer to it.  Bodies without part_words() stage nothing.
template <class B, class = void>
struct HasParts : std::false_type {};
template <class B>
struct HasParts<B, std::void_t<decltype(std::declval<const B&>().part_words())>> : std::true_type {};

template <class Body>
__host__ __device__ inline int part_words_of(const Body& body) {
    if constexpr (HasParts<Body>::value)
        return body.part_words();
    else
        return 0;
}

template <class Body>
__device__ inline const unsigned long long* stage_parts(const Body& body, unsigned long long* part) {
    if constexpr (HasParts<Body>::value) {
        if (body.part_words()) {
            body.stage(part);
            __syncthreads();
            return part;
        }
    }
    return nullptr;
}

// Bodies may also keep block-level sums in their staged table (updated with
// LDS atomics while rows run) and add them to global memory once per block:
// flush(part) runs after the block's last row.
template <class B, class = void>
struct HasFlush : std::false_type {};
template <class B>
struct HasFlush<B, std::void_t<decltype(std::declval<const B&>().flush(std::declval<const unsigned long long*>()))>>
    : std::true_type {};

template <class Body>
__device__ inline void flush_parts(const Body& body, const unsigned long long* part) {
    if constexpr (HasFlush<Body>::value) {
        if (part) {
            __syncthreads();
            body.flush(part);
        }
    }
}

struct RowOut {
    bool append;
    std::int64_t acc;
};

// Short-row tables of bodies that declare LightBin use that bin width; the
// scratch argument of run_tiers counts 64-bit bins.
template <class B, class = void>
struct LightBinOf {
    using type = unsigned long long;
};
template <class B>
struct LightBinOf<B, std::void_t<typename B::LightBin>> {
    using type = typename B::LightBin;
};
template <class Body>
constexpr int light_scratch_words(int scratch) {
    return static_cast<int>((static_cast<std::size_t>(scratch) * sizeof(typename LightBinOf<Body>::type) + 7) / 8);
}

struct TierLists {
    std::int32_t* medium;
    unsigned long long* medium_count;
    std::int32_t* heavy;
    unsigned long long* heavy_count;
    std::int32_t* out;          // rows whose body asked for an append
    unsigned long long* out_count;
    unsigned long long* acc;    // sum of body accumulators (modular)
    // Null, or two words that receive the entries and the number of the rows
    // handled (algorithmic-byte accounting of the profiled runs).
    unsigned long long* entries;
};

// Light tier: list (or the range [0, n) when list is null) of count rows,
// count read from count_ptr when non-null.
template <class Body, int W>
__global__ void __launch_bounds__(kTierThreads) k_tier_light(Body body, const std::int32_t* list, std::int64_t n,
                                                             const unsigned long long* count_ptr, TierLists t,
                                                             int scratch_per_group, bool route_heavy) {
    extern __shared__ unsigned long long smem[];
    __shared__ std::int32_t stage_buf[kTierThreads / kWave][2][kStage];
    constexpr int G = kWave / W;
    const std::int64_t count = count_ptr ? (std::int64_t)*count_ptr : n;
    // Blocks past the listed rows leave before their set-up (the grid is
    // sized for the longest list the caller may pass).
    if ((std::int64_t)blockIdx.x * (kTierThreads / kWave) * G >= count) return;
    const int lane = lane_id(), gw = lane / W, lg = lane % W;
    const unsigned long long* part = stage_parts(body, smem);
    unsigned long long* team_smem = smem + part_words_of(body);
    unsigned long long* scratch = team_smem + (threadIdx.x / W) * scratch_per_group;
    Stage defer_stage{stage_buf[threadIdx.x / kWave][0], t.medium, t.medium_count, 0};
    Stage out_stage{stage_buf[threadIdx.x / kWave][1], t.out, t.out_count, 0};
    // Team tables start zero; bodies that use them leave them zero.
    for (int i = threadIdx.x; i < (kTierThreads / W) * scratch_per_group; i += kTierThreads) team_smem[i] = 0;
    __syncthreads();
    const std::int64_t wave = (blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x) / kWave;
    const std::int64_t waves = (std::int64_t)gridDim.x * blockDim.x / kWave;
    std::int64_t acc = 0, rows_entries = 0, rows_done = 0;
    for (std::int64_t base = wave * G; base < count; base += waves * G) {
        const std::int64_t idx = base + gw;
        std::int32_t v = 0;
        bool defer = false;
        RowOut r{false, 0};
        std::int64_t deg = 0;
        bool heavy = false;
        if (idx < count) {
            v = list ? list[idx] : (std::int32_t)idx;
            deg = body.degree(v);
            defer = deg > kTierLight;
            heavy = route_heavy && deg > kTierMedium;
        }
        // Heavy rows (rare) go straight to the heavy list, so the heavy
        // launch does not wait for the medium one; without route_heavy the
        // medium launch takes every longer row.
        if (lg == 0 && heavy) t.heavy[atomicAdd(t.heavy_count, 1ULL)] = v;
        // Wave-uniform bound on the entries of the rows handled together,
        // so that bodies with unrolled row loops skip empty iterations.
        const int wmax = (int)wave_max((std::uint64_t)(idx < count && !defer ? deg : 0));
        if (idx < count && !defer) r = body(GroupTeam<W>{lg, wmax, part}, v, scratch);
        if (lg == 0) acc += r.acc;
        if (lg == 0 && idx < count && !defer) {
            rows_entries += deg;
            ++rows_done;
        }
        defer_stage.push(lg == 0 && defer && !heavy, v);
        out_stage.push(lg == 0 && r.append, v);
    }
    defer_stage.flush();
    out_stage.flush();
    if (t.acc) {
        acc = wave_sum(acc);
        if (lane == 0 && acc) atomicAdd(t.acc, (unsigned long long)acc);
    }
    if (t.entries) {
        rows_entries = wave_sum(rows_entries);
        rows_done = wave_sum(rows_done);
        if (lane == 0 && rows_done) {
            atomicAdd(t.entries, (unsigned long long)rows_entries);
            atomicAdd(t.entries + 1, (unsigned long long)rows_done);
        }
    }
    flush_parts(body, part);
}

// Medium tier: teams of W (32 or 64) lanes, G = 64 / W rows per wave step.
template <class Body, int W>
__global__ void __launch_bounds__(kTierThreads) k_tier_medium(Body body, TierLists t, int scratch_per_group) {
    static_assert(W == 32 || W == 64, "medium teams are half or whole waves");
    extern __shared__ unsigned long long smem[];
    __shared__ std::int32_t stage_buf[kTierThreads / kWave][kStage];
    constexpr int G = kWave / W;
    const std::int64_t count = (std::int64_t)*t.medium_count;
    if ((std::int64_t)blockIdx.x * (kTierThreads / kWave) * G >= count) return;
    const int lane = lane_id(), gw = lane / W, lg = lane % W;
    const unsigned long long* part = stage_parts(body, smem);
    unsigned long long* scratch = smem + part_words_of(body) + (threadIdx.x / W) * scratch_per_group;
    Stage out_stage{stage_buf[threadIdx.x / kWave], t.out, t.out_count, 0};
    const std::int64_t wave = (blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x) / kWave;
    const std::int64_t waves = (std::int64_t)gridDim.x * blockDim.x / kWave;
    std::int64_t acc = 0, rows_entries = 0, rows_done = 0;
    for (std::int64_t base = wave * G; base < count; base += waves * G) {
        const std::int64_t idx = base + gw;
        std::int32_t v = 0;
        RowOut r{false, 0};
        // The light launch routes rows longer than kTierMedium to the heavy
        // list; every row listed here is handled here.
        if (idx < count) {
            v = t.medium[idx];
            r = body(GroupTeam<W>{lg, 1 << 30, part}, v, scratch);
            if (t.entries && lg == 0) {
                rows_entries += body.degree(v);
                ++rows_done;
            }
        }
        if (lg == 0) acc += r.acc;
        out_stage.push(lg == 0 && r.append, v);
    }
    out_stage.flush();
    if (t.acc) {
        acc = wave_sum(acc);
        if (lane == 0 && acc) atomicAdd(t.acc, (unsigned long long)acc);
    }
    if (t.entries) {
        rows_entries = wave_sum(rows_entries);
        rows_done = wave_sum(rows_done);
        if (lane == 0 && rows_done) {
            atomicAdd(t.entries, (unsigned long long)rows_entries);
            atomicAdd(t.entries + 1, (unsigned long long)rows_done);
        }
    }
    flush_parts(body, part);
}

// More blocks than stay resident: with about one row per block the
// dispatcher hands the next row to whichever block slot frees first, which
// balances rows of very different lengths.  Blocks beyond the heavy row
// count exit at once.
constexpr int kHeavyGrid = 4096;

template <class Body>
__global__ void __launch_bounds__(kHeavyThreads) k_tier_heavy(Body body, TierLists t) {
    extern __shared__ unsigned long long smem[];
    __shared__ unsigned long long red[kBlockRed];
    const std::int64_t count = (std::int64_t)*t.heavy_count;
    if ((std::int64_t)blockIdx.x >= count) return;
    const unsigned long long* part = stage_parts(body, smem);
    unsigned long long* scratch = smem + part_words_of(body);
    std::int64_t acc = 0, rows_entries = 0, rows_done = 0;
    for (std::int64_t idx = blockIdx.x; idx < count; idx += gridDim.x) {
        const std::int32_t v = t.heavy[idx];
        RowOut r = body(BlockTeam{red, part}, v, scratch);
        if (threadIdx.x == 0) {
            acc += r.acc;
            if (t.out && r.append) t.out[atomicAdd(t.out_count, 1ULL)] = v;
            if (t.entries) {
                rows_entries += body.degree(v);
                ++rows_done;
            }
        }
        __syncthreads();
    }
    if (t.acc && threadIdx.x == 0 && acc) atomicAdd(t.acc, (unsigned long long)acc);
    if (t.entries && threadIdx.x == 0 && rows_done) {
        atomicAdd(t.entries, (unsigned long long)rows_entries);
        atomicAdd(t.entries + 1, (unsigned long long)rows_done);
    }
    flush_parts(body, part);
}

// Light launches stop at 1024 blocks (four per CU, all resident): each block
// then loops over many rows, so its set-up (part table in LDS, team tables)
// is paid once per block rather than once per few rows.  Measured against
// 512, 2048 and 4096 blocks on R-MAT scale 22 at k=64.
constexpr int kLightGrid = 1024;

// Host launcher: three launches, counters in dev[0] (medium) and dev[1]
// (heavy) are reset here unless the caller knows them to be zero
// (counts_zeroed).  scratch is per-team LDS in u64 units.
// row_bound: an upper bound on the entries of the listed rows.  With no
// row above kTierLight no medium launch is made; with none above
// kTierMedium no heavy launch is made (longer rows would run in the medium
// launch, which is correct for any length).
template <class Body>
void run_tiers(Context& c, const Body& body, const std::int32_t* list, std::int64_t n, const unsigned long long* count_ptr,
               std::int32_t* medium, std::int32_t* heavy, unsigned long long* tier_counts, std::int32_t* out,
               unsigned long long* out_count, unsigned long long* acc, int scratch, std::int64_t max_rows,
               std::uint64_t row_bound = ~0ULL, bool counts_zeroed = false, unsigned long long* entries = nullptr) {
    const hipStream_t s = c.stream();
    const bool heavy_rows = row_bound > static_cast<std::uint64_t>(kTierMedium);
    const bool medium_rows = row_bound > static_cast<std::uint64_t>(kTierLight);
    if (!counts_zeroed) fill<unsigned long long>(c, tier_counts, 2, 0ULL);
    TierLists t{medium, tier_counts, heavy, tier_counts + 1, out, out_count, acc, entries};
    const int lscratch = light_scratch_words<Body>(scratch);
    int lanes = c.light_lanes();
    while (lanes < 16 && (kTierThreads / lanes) * lscratch * sizeof(unsigned long long) > 48 * 1024) lanes *= 2;
    const int light_blocks = static_cast<int>(
        std::max<std::int64_t>(1, std::min<std::int64_t>(ceil_div(max_rows, kTierThreads / lanes), kLightGrid)));
    const std::size_t part = static_cast<std::size_t>(part_words_of(body)) * sizeof(unsigned long long);
    const std::size_t light_lds = part + (kTierThreads / lanes) * lscratch * sizeof(unsigned long long);
    if (lanes == 4)
        hipLaunchKernelGGL((k_tier_light<Body, 4>), dim3(light_blocks), dim3(kTierThreads), light_lds, s, body, list, n,
                           count_ptr, t, lscratch, heavy_rows);
    else if (lanes == 8)
        hipLaunchKernelGGL((k_tier_light<Body, 8>), dim3(light_blocks), dim3(kTierThreads), light_lds, s, body, list, n,
                           count_ptr, t, lscratch, heavy_rows);
    else
        hipLaunchKernelGGL((k_tier_light<Body, 16>), dim3(light_blocks), dim3(kTierThreads), light_lds, s, body, list, n,
                           count_ptr, t, lscratch, heavy_rows);
    GP_LAUNCH_CHECK();
    // The light launch fills the heavy list: with overlap on, the heavy
    // launch (few long rows, few blocks) runs on the side stream beside the
    // medium launch.
    // Heavy teams keep one table per wave plus the merged one.
    const std::size_t heavy_lds = part + static_cast<std::size_t>(scratch) * (1 + kHeavyThreads / kWave) *
                                             sizeof(unsigned long long);
    const bool side = heavy_rows && c.overlap();
    if (side) {
        c.fork();
        hipLaunchKernelGGL((k_tier_heavy<Body>), dim3(kHeavyGrid), dim3(kHeavyThreads), heavy_lds, c.aux(), body, t);
        GP_LAUNCH_CHECK();
    }
    if (medium_rows) {
        const int mlanes = c.medium_lanes();
        const int medium_blocks = static_cast<int>(
            std::max<std::int64_t>(1, std::min<std::int64_t>(ceil_div(max_rows, kTierThreads / mlanes), 2048)));
        const std::size_t medium_lds = part + (kTierThreads / mlanes) * scratch * sizeof(unsigned long long);
        if (mlanes == 32)
            hipLaunchKernelGGL((k_tier_medium<Body, 32>), dim3(medium_blocks), dim3(kTierThreads), medium_lds, s, body,
                               t, scratch);
        else
            hipLaunchKernelGGL((k_tier_medium<Body, 64>), dim3(medium_blocks), dim3(kTierThreads), medium_lds, s, body,
                               t, scratch);
        GP_LAUNCH_CHECK();
    }
    // Heavy rows are few but long; blocks beyond the row count exit at once.
    if (side) {
        c.join();
    } else if (heavy_rows) {
        hipLaunchKernelGGL((k_tier_heavy<Body>), dim3(kHeavyGrid), dim3(kHeavyThreads), heavy_lds, s, body, t);
        GP_LAUNCH_CHECK();
    }
}

}  // namespace gp
