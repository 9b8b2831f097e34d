// Segment sort-and-reduce building blocks shared by aggregation (cluster
// affinities of a row) and contraction (coarse rows).  A segment of at most
// W*I (key:int32, value:uint32) pairs is sorted in registers by a bitonic
// network over aligned groups of W lanes, then equal keys are merged by a
// segmented scan.  Longer segments use an LDS block sort (<= 4096) or the
// global radix path.
#pragma once
#include "gp_prims.hpp"

namespace gp {

constexpr std::int32_t kSentinel = 0x7fffffff;

template <int W>
__device__ inline unsigned long long group_mask_bits() {
    return W == 64 ? ~0ULL : ((1ULL << W) - 1ULL);
}

// Bitonic sort of W*I elements; element index = item*W + lane-in-group.
template <int W, int I>
__device__ inline void group_bitonic_sort(std::int32_t (&k)[I], std::uint32_t (&v)[I], int lg) {
    constexpr int N = W * I;
#pragma unroll
    for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
        for (int dist = size >> 1; dist > 0; dist >>= 1) {
            std::int32_t nk[I];
            std::uint32_t nv[I];
#pragma unroll
            for (int it = 0; it < I; ++it) {
                const int idx = it * W + lg;
                std::int32_t ok;
                std::uint32_t ov;
                if (dist >= W) {
                    ok = k[it ^ (dist / W)];
                    ov = v[it ^ (dist / W)];
                } else {
                    ok = __shfl_xor(k[it], dist, W);
                    ov = __shfl_xor(v[it], dist, W);
                }
                const bool lower = ((idx & dist) == 0) == ((idx & size) == 0);
                const bool take = lower ? ok < k[it] : ok > k[it];
                nk[it] = take ? ok : k[it];
                nv[it] = take ? ov : v[it];
            }
#pragma unroll
            for (int it = 0; it < I; ++it) {
                k[it] = nk[it];
                v[it] = nv[it];
            }
        }
    }
}

// After group_bitonic_sort: calls emit(key, sum, index) once per distinct
// key (excluding kSentinel), in ascending key order; index counts emissions.
// Returns the number of emitted keys (uniform across the group).
template <int W, int I, typename Emit>
__device__ inline int group_reduce_runs(const std::int32_t (&k)[I], const std::uint32_t (&v)[I], int lg,
                                        int group_base, Emit&& emit) {
    std::int32_t carry_key = kSentinel;
    std::uint64_t carry_sum = 0;
    int emitted = 0;
    const unsigned long long lt = (1ULL << lg) - 1ULL;
#pragma unroll
    for (int it = 0; it < I; ++it) {
        const std::int32_t key = k[it];
        std::uint64_t sum = v[it];
        if (lg == 0 && key == carry_key) sum += carry_sum;
#pragma unroll
        for (int d = 1; d < W; d <<= 1) {
            const std::int32_t ok = __shfl_up(key, d, W);
            const std::uint64_t os = __shfl_up(sum, d, W);
            if (lg >= d && ok == key) sum += os;
        }
        std::int32_t next_first = kSentinel;
        if (it + 1 < I) next_first = __shfl(k[it + 1 < I ? it + 1 : it], 0, W);
        const std::int32_t following = __shfl_down(key, 1, W);
        const std::int32_t nxt = lg < W - 1 ? following : next_first;
        const bool e = key != kSentinel && key != nxt;
        const unsigned long long mask = (__ballot(e) >> group_base) & group_mask_bits<W>();
        if (e) emit(key, sum, emitted + __popcll(mask & lt));
        emitted += __popcll(mask);
        carry_key = __shfl(key, W - 1, W);
        carry_sum = __shfl(sum, W - 1, W);
    }
    return emitted;
}

// Block-wide (256 threads) LDS bitonic sort of n <= 4096 pairs followed by
// run reduction.  emit(key, sum, index) is called by the thread owning the
// last element of each run; returns the emission count to every thread.
constexpr int kBlockSortThreads = 256;
constexpr int kBlockSortMax = 4096;
static_assert(kBlockSortThreads == kScanThreads, "block scans assume one block shape");

struct BlockSortSmem {
    std::int32_t key[kBlockSortMax];
    std::uint32_t val[kBlockSortMax];
    SegU64 scan[kBlockSortThreads];
    int count[kBlockSortThreads];
};

struct NoBegin {
    __device__ void operator()(int) const {}
};

// begin(total) is called by every thread once the emission count is known,
// before the first emit.
template <typename Emit, typename Begin = NoBegin>
__device__ inline int block_sort_reduce(BlockSortSmem& s, int n, Emit&& emit, Begin&& begin = Begin{}) {
    int N = 2;
    while (N < n) N <<= 1;
    if (N < kBlockSortThreads * 2) N = kBlockSortThreads * 2;
    for (int i = n + threadIdx.x; i < N; i += kBlockSortThreads) {
        s.key[i] = kSentinel;
        s.val[i] = 0;
    }
    __syncthreads();
    for (int size = 2; size <= N; size <<= 1) {
        for (int dist = size >> 1; dist > 0; dist >>= 1) {
            for (int p = threadIdx.x; p < N / 2; p += kBlockSortThreads) {
                const int i = (p / dist) * dist * 2 + (p % dist);
                const int j = i + dist;
                const bool up = (i & size) == 0;
                const std::int32_t a = s.key[i], b = s.key[j];
                if ((a > b) == up) {
                    s.key[i] = b;
                    s.key[j] = a;
                    const std::uint32_t t = s.val[i];
                    s.val[i] = s.val[j];
                    s.val[j] = t;
                }
            }
            __syncthreads();
        }
    }
    // Segmented scan: each thread owns E consecutive elements.
    const int E = N / kBlockSortThreads;
    const int b0 = threadIdx.x * E;
    std::uint64_t run = 0;
    std::uint32_t head_seen = 0;
    for (int i = 0; i < E; ++i) {
        const int idx = b0 + i;
        const bool head = idx == 0 || s.key[idx] != s.key[idx - 1];
        run = head ? s.val[idx] : run + s.val[idx];
        head_seen |= head ? 1u : 0u;
    }
    SegU64 agg{run, head_seen};
    SegU64 inc, prev, all;
    block_scan<SegPlusOp>(agg, s.scan, inc, prev, all);
    // Count run ends per thread for the emission index.
    int ends = 0;
    for (int i = 0; i < E; ++i) {
        const int idx = b0 + i;
        const std::int32_t key = s.key[idx];
        const std::int32_t nxt = idx + 1 < N ? s.key[idx + 1] : kSentinel;
        ends += key != kSentinel && key != nxt;
    }
    int ends_inc, before, total;
    block_scan<PlusOp>(ends, s.count, ends_inc, before, total);
    begin(total);
    std::uint64_t carry = prev.v;
    bool continuing = true;  // the thread's first run continues from previous threads
    std::uint64_t sum = 0;
    for (int i = 0; i < E; ++i) {
        const int idx = b0 + i;
        const bool head = idx == 0 || s.key[idx] != s.key[idx - 1];
        if (head) continuing = false;
        sum = head ? s.val[idx] : sum + s.val[idx];
        const std::uint64_t full = continuing ? carry + sum : sum;
        const std::int32_t key = s.key[idx];
        const std::int32_t nxt = idx + 1 < N ? s.key[idx + 1] : kSentinel;
        if (key != kSentinel && key != nxt) emit(key, full, before++);
    }
    __syncthreads();
    return total;
}

// Block-wide aggregation of (key >= 0, u32 weight) pairs in an LDS hash
// table with linear probing.  Sums are exact while every row's total weight
// stays below 2^32 (callers check the level's row-weight bound).  Slot order
// is arbitrary: callers use it only where the result is order-independent.
constexpr int kHashSlots = 8192;

struct BlockHashSmem {
    std::int32_t key[kHashSlots];
    std::uint32_t val[kHashSlots];
};

// Table size for n entries: a power of two >= 2n (at least 512).
__device__ inline int hash_slots_for(int n) {
    int h = 512;
    while (h < 2 * n) h <<= 1;
    return h;
}

__device__ inline void block_hash_clear(BlockHashSmem& s, int slots) {
    for (int i = threadIdx.x; i < slots; i += blockDim.x) {
        s.key[i] = -1;
        s.val[i] = 0;
    }
}

__device__ inline void block_hash_add(BlockHashSmem& s, int slots, std::int32_t key, std::uint32_t w) {
    unsigned slot = mix32((std::uint32_t)key) & (unsigned)(slots - 1);
    while (true) {
        const std::int32_t prev = atomicCAS(&s.key[slot], -1, key);
        if (prev == -1 || prev == key) {
            atomicAdd(&s.val[slot], w);
            return;
        }
        slot = (slot + 1) & (unsigned)(slots - 1);
    }
}

}  // namespace gp
