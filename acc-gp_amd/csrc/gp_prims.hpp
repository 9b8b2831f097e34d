// Device primitives used by the partitioner: fills, reduce-then-scan scans
// (plain and segmented), stable index selection and an LSD radix sort of
// (u64 key, u32 value) pairs with wave64 multi-split ranking.  All launches go
// to the context stream; nothing here allocates outside the arena.
#pragma once
#include "gp_runtime.hpp"

namespace gp {

// ------------------------------------------------------------------ fill
template <typename T>
__global__ void k_fill(T* p, std::int64_t n, T v) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < n;
         i += (std::int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
template <typename T>
inline void fill(Context& c, T* p, std::int64_t n, T v) {
    if (n <= 0) return;
    const int blocks = static_cast<int>(std::min<std::int64_t>(ceil_div(n, 256), 4096));
    hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, c.stream(), p, n, v);
    GP_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_iota(T* p, std::int64_t n) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < n;
         i += (std::int64_t)gridDim.x * blockDim.x)
        p[i] = static_cast<T>(i);
}
template <typename T>
inline void iota(Context& c, T* p, std::int64_t n) {
    if (n <= 0) return;
    const int blocks = static_cast<int>(std::min<std::int64_t>(ceil_div(n, 256), 4096));
    hipLaunchKernelGGL(k_iota<T>, dim3(blocks), dim3(256), 0, c.stream(), p, n);
    GP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ scans
struct PlusOp {
    template <typename T>
    __device__ static T id() { return T(0); }
    template <typename T>
    __device__ static T op(T a, T b) { return a + b; }
};
struct MaxOp {
    template <typename T>
    __device__ static T id() { return T(0); }
    template <typename T>
    __device__ static T op(T a, T b) { return a > b ? a : b; }
};

// Segmented value: head marks the first element of a segment.
struct SegU64 {
    std::uint64_t v;
    std::uint32_t head;
};
struct SegPlusOp {
    template <typename T>
    __device__ static T id() { return T{0, 0}; }
    template <typename T>
    __device__ static T op(T a, T b) { return b.head ? b : T{a.v + b.v, a.head | b.head}; }
};

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// Lane shuffles of scan values.
template <typename T>
__device__ inline T shfl_up_val(T x, int d) {
    return __shfl_up(x, d);
}
__device__ inline SegU64 shfl_up_val(SegU64 x, int d) {
    return SegU64{__shfl_up(x.v, d), __shfl_up(x.head, d)};
}

// Block-wide scan of one value per thread: wave scans with shuffles, then
// the wave totals through LDS (two barriers).  wsum holds one T per wave.
template <typename Op, typename T>
__device__ inline void block_scan(T x, T* wsum, T& inclusive, T& exclusive, T& total) {
    const int lane = lane_id(), w = threadIdx.x / kWave;
    for (int d = 1; d < kWave; d <<= 1) {
        const T y = shfl_up_val(x, d);
        if (lane >= d) x = Op::template op<T>(y, x);
    }
    T wex = shfl_up_val(x, 1);
    if (lane == kWave - 1) wsum[w] = x;
    __syncthreads();
    T pre = Op::template id<T>(), tot = Op::template id<T>();
    for (int i = 0; i < kScanThreads / kWave; ++i) {
        if (i < w) pre = Op::template op<T>(pre, wsum[i]);
        tot = Op::template op<T>(tot, wsum[i]);
    }
    __syncthreads();
    inclusive = Op::template op<T>(pre, x);
    exclusive = lane ? Op::template op<T>(pre, wex) : pre;
    total = tot;
}

// Inclusive scan across the 64 lanes of a wave.
template <typename Op, typename T>
__device__ inline T wave_inclusive_scan(T x) {
    const int lane = lane_id();
    for (int d = 1; d < kWave; d <<= 1) {
        const T y = shfl_up_val(x, d);
        if (lane >= d) x = Op::template op<T>(y, x);
    }
    return x;
}
template <typename T>
__device__ inline T shfl_val(T x, int src) {
    return __shfl(x, src);
}
__device__ inline SegU64 shfl_val(SegU64 x, int src) {
    return SegU64{__shfl(x.v, src), __shfl(x.head, src)};
}

// Tile layout: wave w owns the contiguous kScanItems * 64 elements at
// base + w * kScanItems * 64, read in rounds of 64 consecutive elements so
// that every load and store instruction covers contiguous memory.
template <typename Op, typename T, typename In>
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const In* in, std::int64_t n, T* partial) {
    __shared__ T wsum[kScanThreads / kWave];
    const int lane = lane_id(), w = threadIdx.x / kWave;
    const std::int64_t base = blockIdx.x * (std::int64_t)kScanTile + w * (kScanItems * kWave);
    T acc = Op::template id<T>();
    for (int r = 0; r < kScanItems; ++r) {
        const std::int64_t idx = base + r * kWave + lane;
        const T x = idx < n ? static_cast<T>(in[idx]) : Op::template id<T>();
        acc = Op::template op<T>(acc, shfl_val(wave_inclusive_scan<Op>(x), kWave - 1));
    }
    if (lane == 0) wsum[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T tot = wsum[0];
        for (int i = 1; i < kScanThreads / kWave; ++i) tot = Op::template op<T>(tot, wsum[i]);
        partial[blockIdx.x] = tot;
    }
}

// Exclusive scan of the block partials by one block (loops over chunks).
template <typename Op, typename T>
__global__ void __launch_bounds__(kScanThreads) k_scan_partials(T* partial, std::int64_t nb, T* total) {
    __shared__ T wsum[kScanThreads / kWave];
    T carry = Op::template id<T>();
    for (std::int64_t base = 0; base < nb; base += kScanThreads) {
        const std::int64_t idx = base + threadIdx.x;
        const T x = idx < nb ? partial[idx] : Op::template id<T>();
        T inc, exc, tot;
        block_scan<Op>(x, wsum, inc, exc, tot);
        if (idx < nb) partial[idx] = Op::template op<T>(carry, exc);
        carry = Op::template op<T>(carry, tot);
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

// Final pass; exclusive=true writes out[i] = prefix before i.
template <typename Op, typename T, typename In, typename Out>
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const In* in, Out* out, std::int64_t n, const T* partial,
                                                             bool exclusive) {
    __shared__ T wsum[kScanThreads / kWave];
    const int lane = lane_id(), w = threadIdx.x / kWave;
    const std::int64_t base = blockIdx.x * (std::int64_t)kScanTile + w * (kScanItems * kWave);
    T inc[kScanItems], own[kScanItems];
    T acc = Op::template id<T>();
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        const std::int64_t idx = base + r * kWave + lane;
        own[r] = idx < n ? static_cast<T>(in[idx]) : Op::template id<T>();
        inc[r] = wave_inclusive_scan<Op>(own[r]);
        acc = Op::template op<T>(acc, shfl_val(inc[r], kWave - 1));
    }
    if (lane == 0) wsum[w] = acc;
    __syncthreads();
    T carry = partial[blockIdx.x];
    for (int i = 0; i < w; ++i) carry = Op::template op<T>(carry, wsum[i]);
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        const std::int64_t idx = base + r * kWave + lane;
        const T full = Op::template op<T>(carry, inc[r]);
        // exclusive: the inclusive value of the previous lane (or carry)
        T prev = shfl_up_val(full, 1);
        if (lane == 0) prev = carry;
        if (idx < n) out[idx] = static_cast<Out>(exclusive ? prev : full);
        carry = Op::template op<T>(carry, shfl_val(inc[r], kWave - 1));
    }
}

// Scan n elements of In into Out (in == out allowed).  Accumulates in T.
// total (device, optional) receives the reduction of all elements.
template <typename Op, typename T, typename In, typename Out>
inline void scan(Context& c, const In* in, Out* out, std::int64_t n, bool exclusive, T* total = nullptr) {
    if (n <= 0) {
        if (total) fill<T>(c, total, 1, T{});
        return;
    }
    const std::int64_t nb = (n + kScanTile - 1) / kScanTile;
    DBuf<T> partial = c.buf<T>(static_cast<std::size_t>(nb));
    hipLaunchKernelGGL((k_scan_reduce<Op, T, In>), dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0,
                       c.stream(), in, n, partial.get());
    GP_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_scan_partials<Op, T>), dim3(1), dim3(kScanThreads), 0, c.stream(), partial.get(), nb,
                       total);
    GP_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_scan_apply<Op, T, In, Out>), dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0,
                       c.stream(), in, out, n, partial.get(), exclusive);
    GP_LAUNCH_CHECK();
}

template <typename T, typename In, typename Out>
inline void exclusive_sum(Context& c, const In* in, Out* out, std::int64_t n, T* total = nullptr) {
    scan<PlusOp, T, In, Out>(c, in, out, n, true, total);
}
template <typename T, typename In, typename Out>
inline void inclusive_sum(Context& c, const In* in, Out* out, std::int64_t n, T* total = nullptr) {
    scan<PlusOp, T, In, Out>(c, in, out, n, false, total);
}

// ------------------------------------------------------ stable selection
// out[pos[i]] = i for flagged i, pos = exclusive scan of flags; count -> *count.
static __global__ void k_select_scatter(const std::uint8_t* flags, const std::int64_t* pos, std::int64_t n,
                                        std::int32_t* out) {
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < n;
         i += (std::int64_t)gridDim.x * blockDim.x)
        if (flags[i]) out[pos[i]] = static_cast<std::int32_t>(i);
}
// Short inputs: one block walks the flags in rounds of 1024 (ballot ranks
// inside waves, wave counts through LDS), one launch instead of four.
constexpr int kSelectSmallMax = 8192;
static __global__ void __launch_bounds__(1024) k_select_small(const std::uint8_t* flags, std::int64_t n,
                                                              std::int32_t* out, std::int64_t* count_dev) {
    __shared__ int wsum[1024 / kWave];
    __shared__ int base;
    const int t = threadIdx.x, w = t / kWave, lane = t % kWave;
    const unsigned long long lt = (1ULL << lane) - 1ULL;
    if (t == 0) base = 0;
    __syncthreads();
    for (std::int64_t r = 0; r < n; r += 1024) {
        const std::int64_t i = r + t;
        const bool f = i < n && flags[i] != 0;
        const unsigned long long mask = __ballot(f);
        if (lane == 0) wsum[w] = __popcll(mask);
        __syncthreads();
        int pre = base, total = 0;
        for (int j = 0; j < 1024 / kWave; ++j) {
            pre += j < w ? wsum[j] : 0;
            total += wsum[j];
        }
        if (f) out[pre + __popcll(mask & lt)] = static_cast<std::int32_t>(i);
        __syncthreads();
        if (t == 0) base += total;
        __syncthreads();
    }
    if (t == 0) *count_dev = base;
}

inline void stable_select(Context& c, const std::uint8_t* flags, std::int64_t n, std::int32_t* out,
                          std::int64_t* count_dev) {
    if (n > 0 && n <= kSelectSmallMax) {
        hipLaunchKernelGGL(k_select_small, dim3(1), dim3(1024), 0, c.stream(), flags, n, out, count_dev);
        GP_LAUNCH_CHECK();
        return;
    }
    DBuf<std::int64_t> pos = c.buf<std::int64_t>(static_cast<std::size_t>(n > 0 ? n : 1));
    exclusive_sum<std::int64_t>(c, flags, pos.get(), n, count_dev);
    if (n > 0) {
        const int blocks = static_cast<int>(std::min<std::int64_t>(ceil_div(n, 256), 8192));
        hipLaunchKernelGGL(k_select_scatter, dim3(blocks), dim3(256), 0, c.stream(), flags, pos.get(), n, out);
        GP_LAUNCH_CHECK();
    }
}

// ---------------------------------------------------------------- radix sort
constexpr int kRadixThreads = 256;
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kRadixThreads * kRadixItems;
constexpr int kRadixBins = 256;

static __global__ void k_radix_hist(const std::uint64_t* keys, std::int64_t n, int shift, std::uint32_t* hist,
                                    int nblocks) {
    __shared__ std::uint32_t h[kRadixBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const std::int64_t base = blockIdx.x * (std::int64_t)kRadixTile;
    for (int r = 0; r < kRadixItems; ++r) {
        const std::int64_t i = base + r * kRadixThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xff], 1u);
    }
    __syncthreads();
    hist[(std::int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile.  Each wave ranks a contiguous quarter of the
// tile (16 rounds of 64 keys, peers by 8 ballots) against wave-private digit
// counters, so ranking needs no block barrier; the tile is then reordered by
// digit in LDS and written out in index order, which makes the stores of
// one digit run consecutive.
static __global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(
    const std::uint64_t* keys_in, const std::uint32_t* vals_in, std::uint64_t* keys_out, std::uint32_t* vals_out,
    std::int64_t n, int shift, const std::uint32_t* offsets, int nblocks) {
    constexpr int kWaves = kRadixThreads / kWave;
    constexpr int kPerWave = kRadixTile / kWaves;
    constexpr int kRounds = kPerWave / kWave;
    __shared__ std::uint32_t wcount[kWaves][kRadixBins];
    __shared__ std::uint32_t start[kRadixBins];
    __shared__ std::uint32_t gbase[kRadixBins];
    __shared__ std::uint32_t wsum[kWaves];
    __shared__ std::uint64_t skeys[kRadixTile];
    __shared__ std::uint32_t svals[kRadixTile];
    const int t = threadIdx.x, w = t / kWave, lane = t % kWave;
    for (int i = 0; i < kWaves; ++i) wcount[i][t] = 0;
    __syncthreads();
    const std::int64_t tile = blockIdx.x * (std::int64_t)kRadixTile;
    const std::int64_t wbase = tile + (std::int64_t)w * kPerWave;
    const unsigned long long lt = (1ULL << lane) - 1ULL;
    std::uint64_t key[kRounds];
    std::uint32_t val[kRounds];
    std::uint32_t off[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const std::int64_t i = wbase + r * kWave + lane;
        key[r] = i < n ? keys_in[i] : 0;
        val[r] = i < n && vals_in ? vals_in[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const bool valid = wbase + r * kWave + lane < n;
        const std::uint32_t d = static_cast<std::uint32_t>((key[r] >> shift) & 0xff);
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        // Program order of one wave's LDS accesses keeps the read of the
        // running count ahead of the leader's update.
        const std::uint32_t before = valid ? wcount[w][d] : 0;
        off[r] = before + static_cast<std::uint32_t>(__popcll(peers & lt));
        if (valid && (peers & lt) == 0) wcount[w][d] = before + static_cast<std::uint32_t>(__popcll(peers));
    }
    __syncthreads();
    std::uint32_t run = 0;
    for (int j = 0; j < kWaves; ++j) {
        const std::uint32_t c = wcount[j][t];
        wcount[j][t] = run;
        run += c;
    }
    std::uint32_t inc, exc, tot;
    block_scan<PlusOp>(run, wsum, inc, exc, tot);
    start[t] = exc;
    gbase[t] = offsets[(std::int64_t)t * nblocks + blockIdx.x] - exc;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        if (wbase + r * kWave + lane < n) {
            const std::uint32_t d = static_cast<std::uint32_t>((key[r] >> shift) & 0xff);
            const std::uint32_t p = start[d] + wcount[w][d] + off[r];
            skeys[p] = key[r];
            svals[p] = val[r];
        }
    }
    __syncthreads();
    const int count = static_cast<int>(std::min<std::int64_t>(kRadixTile, n - tile));
    for (int i = t; i < count; i += kRadixThreads) {
        const std::uint64_t k = skeys[i];
        const std::uint32_t g = gbase[(k >> shift) & 0xff] + static_cast<std::uint32_t>(i);
        keys_out[g] = k;
        if (vals_out) vals_out[g] = svals[i];
    }
}

// Whole sort of at most kSortTileMax keys in one block: the keys stay in
// registers and LDS across all digit passes (same ranking as the scatter
// pass), so a short sort costs one launch instead of three per digit.  A
// pass whose digit is equal for every key is skipped (stable identity).
// The block has Waves waves of kSortTileRounds x 64 keys each, the fewest
// of 4, 8 and 16 that hold the input: the ranking rounds, the per-wave
// count scan and the barriers of every pass scale with the waves.
constexpr int kSortTileThreads = 1024;
constexpr int kSortTileRounds = 8;
constexpr int kSortTileMax = kSortTileThreads * kSortTileRounds;

template <int Waves>
static __global__ __launch_bounds__(Waves* kWave) void k_radix_sort_tile(std::uint64_t* keys, std::uint32_t* vals, int n,
                                                                   int begin_bit, int end_bit) {
    constexpr int kWaves = Waves;
    constexpr int kThreads = Waves * kWave;
    constexpr int kPerWave = kSortTileRounds * kWave;
    constexpr int kScanWaves = kRadixBins / kWave;
    static_assert(kThreads >= kRadixBins, "one thread per digit in the count scan");
    __shared__ std::uint32_t wcount[kWaves][kRadixBins];
    __shared__ std::uint32_t start[kRadixBins];
    __shared__ std::uint32_t wsum[kScanWaves];
    __shared__ int same;
    __shared__ std::uint64_t skeys[kThreads * kSortTileRounds];
    __shared__ std::uint32_t svals[kThreads * kSortTileRounds];
    const int t = threadIdx.x, w = t / kWave, lane = t % kWave;
    const int wbase = w * kPerWave;
    const unsigned long long lt = (1ULL << lane) - 1ULL;
    std::uint64_t key[kSortTileRounds];
    std::uint32_t val[kSortTileRounds];
    std::uint32_t off[kSortTileRounds];
#pragma unroll
    for (int r = 0; r < kSortTileRounds; ++r) {
        const int i = wbase + r * kWave + lane;
        key[r] = i < n ? keys[i] : 0;
        val[r] = i < n && vals ? vals[i] : 0;
    }
    // same holds shift + 1 of a pass whose digit is constant; it is written
    // after the first barrier of a pass, which every thread passes only
    // after reading it in the previous pass, so it needs no reset.
    if (t == 0) same = -1;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        for (int i = t; i < kWaves * kRadixBins; i += kThreads) wcount[i / kRadixBins][i % kRadixBins] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSortTileRounds; ++r) {
            // A round past the input has no key in this wave (uniform).
            if (wbase + r * kWave >= n) break;
            const bool valid = wbase + r * kWave + lane < n;
            const std::uint32_t d = static_cast<std::uint32_t>((key[r] >> shift) & 0xff);
            unsigned long long peers = __ballot(valid);
            for (int b = 0; b < 8; ++b) {
                const bool bit = (d >> b) & 1u;
                const unsigned long long bal = __ballot(bit);
                peers &= bit ? bal : ~bal;
            }
            const std::uint32_t before = valid ? wcount[w][d] : 0;
            off[r] = before + static_cast<std::uint32_t>(__popcll(peers & lt));
            if (valid && (peers & lt) == 0) wcount[w][d] = before + static_cast<std::uint32_t>(__popcll(peers));
        }
        __syncthreads();
        std::uint32_t run = 0, x = 0;
        if (t < kRadixBins) {
            for (int j = 0; j < kWaves; ++j) {
                const std::uint32_t c = wcount[j][t];
                wcount[j][t] = run;
                run += c;
            }
            if (run == static_cast<std::uint32_t>(n)) same = shift + 1;
            x = run;
            for (int d = 1; d < kWave; d <<= 1) {
                const std::uint32_t y = __shfl_up(x, d);
                if (lane >= d) x += y;
            }
            if (lane == kWave - 1) wsum[w] = x;
        }
        __syncthreads();
        if (same == shift + 1) continue;
        if (t < kRadixBins) {
            std::uint32_t pre = 0;
            for (int j = 0; j < w; ++j) pre += wsum[j];
            start[t] = pre + x - run;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSortTileRounds; ++r) {
            if (wbase + r * kWave + lane < n) {
                const std::uint32_t d = static_cast<std::uint32_t>((key[r] >> shift) & 0xff);
                const std::uint32_t p = start[d] + wcount[w][d] + off[r];
                skeys[p] = key[r];
                svals[p] = val[r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSortTileRounds; ++r) {
            const int i = wbase + r * kWave + lane;
            if (i < n) {
                key[r] = skeys[i];
                val[r] = svals[i];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kSortTileRounds; ++r) {
        const int i = wbase + r * kWave + lane;
        if (i < n) {
            keys[i] = key[r];
            if (vals) vals[i] = val[r];
        }
    }
}

// Sorts (keys, vals) by bits [begin_bit, end_bit), stably.  The sorted
// data ends in whichever of the two buffer pairs the last pass wrote; the
// DBufs are exchanged rather than copied back.  vals may be empty.
inline void radix_sort_pairs(Context& c, DBuf<std::uint64_t>& keys, DBuf<std::uint32_t>& vals, std::int64_t n,
                             int begin_bit, int end_bit) {
    if (n <= 1 || end_bit <= begin_bit) return;
    if (n > 0xffffffffLL) throw std::invalid_argument("radix sort size exceeds 32-bit offsets");
    if (n <= kSortTileMax) {
        constexpr int kPerWave = kSortTileRounds * kWave;
        if (n <= 4 * kPerWave)
            hipLaunchKernelGGL(k_radix_sort_tile<4>, dim3(1), dim3(4 * kWave), 0, c.stream(), keys.get(), vals.get(),
                               static_cast<int>(n), begin_bit, end_bit);
        else if (n <= 8 * kPerWave)
            hipLaunchKernelGGL(k_radix_sort_tile<8>, dim3(1), dim3(8 * kWave), 0, c.stream(), keys.get(), vals.get(),
                               static_cast<int>(n), begin_bit, end_bit);
        else
            hipLaunchKernelGGL(k_radix_sort_tile<16>, dim3(1), dim3(16 * kWave), 0, c.stream(), keys.get(),
                               vals.get(), static_cast<int>(n), begin_bit, end_bit);
        GP_LAUNCH_CHECK();
        return;
    }
    const int nblocks = ceil_div(n, kRadixTile);
    DBuf<std::uint64_t> k2 = c.buf<std::uint64_t>(static_cast<std::size_t>(n));
    DBuf<std::uint32_t> v2;
    if (vals.get()) v2 = c.buf<std::uint32_t>(static_cast<std::size_t>(n));
    DBuf<std::uint32_t> hist = c.buf<std::uint32_t>(static_cast<std::size_t>(kRadixBins) * nblocks);
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kRadixThreads), 0, c.stream(), keys.get(), n, shift,
                           hist.get(), nblocks);
        GP_LAUNCH_CHECK();
        exclusive_sum<std::uint32_t>(c, hist.get(), hist.get(), static_cast<std::int64_t>(kRadixBins) * nblocks);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kRadixThreads), 0, c.stream(), keys.get(), vals.get(),
                           k2.get(), v2.get(), n, shift, hist.get(), nblocks);
        GP_LAUNCH_CHECK();
        keys.swap(k2);
        vals.swap(v2);
    }
}

inline int bit_width_u64(std::uint64_t v) {
    int b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}

// ---------------------------------------------------------- reductions
// One global atomic per block (blocks of at most 1024 threads).
template <typename T>
__global__ void k_reduce_max(const T* in, std::int64_t n, unsigned long long* out) {
    __shared__ unsigned long long part[16];
    unsigned long long m = 0;
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < n;
         i += (std::int64_t)gridDim.x * blockDim.x)
        m = m > (unsigned long long)in[i] ? m : (unsigned long long)in[i];
    m = wave_max(m);
    if (lane_id() == 0) part[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (unsigned w = 1; w < blockDim.x / kWave; ++w) m = m > part[w] ? m : part[w];
        if (m) atomicMax(out, m);
    }
}
template <typename T>
__global__ void k_reduce_sum(const T* in, std::int64_t n, unsigned long long* out) {
    __shared__ unsigned long long part[16];
    unsigned long long s = 0;
    for (std::int64_t i = blockIdx.x * (std::int64_t)blockDim.x + threadIdx.x; i < n;
         i += (std::int64_t)gridDim.x * blockDim.x)
        s += (unsigned long long)in[i];
    s = wave_sum(s);
    if (lane_id() == 0) part[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (unsigned w = 1; w < blockDim.x / kWave; ++w) s += part[w];
        if (s) atomicAdd(out, s);
    }
}

}  // namespace gp
