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

// Pads [n, N) with sentinels; N is a power of two >= max(n, 512).
__device__ inline int block_pad(BlockSortSmem& s, int n) {
    int N = 2;
    while (N < n) N <<= 1;
    if (N < kBlockSortThreads * 2) N = kBlockSortThreads * 2;
    for (int i = n + threadIdx.x; i < N; i += kBlockSortThreads) {
        s.key[i] = kSentinel;
        s.val[i] = 0;
    }
    __syncthreads();
    return N;
}

__device__ inline void block_bitonic(BlockSortSmem& s, int N) {
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
}

// Scratch of block_radix_sort.
struct BlockRadixSmem {
    std::uint32_t wcount[kBlockSortThreads / kWave][256];
    std::uint32_t wsum[kBlockSortThreads / kWave];
    std::uint32_t start[256];
    int same;
};

// Stable LSD sort of s.key[0, n) (values along) for keys in [0, 2^key_bits)
// or kSentinel, which orders last (ranked as 2^key_bits): ceil((key_bits +
// 1) / 8) passes of 8 bits instead of the log^2 stages of a bitonic network.
// Each wave ranks a contiguous slice in registers against wave-private
// digit counters; a pass whose digit is equal for all keys is skipped.
__device__ inline void block_radix_sort(BlockSortSmem& s, BlockRadixSmem& x, int n, int key_bits) {
    constexpr int kRounds = kBlockSortMax / kBlockSortThreads;
    const int t = threadIdx.x, w = t / kWave, lane = t % kWave;
    const int rounds = (n + kBlockSortThreads - 1) / kBlockSortThreads;
    const int wbase = w * rounds * kWave;
    const unsigned long long lt = (1ULL << lane) - 1ULL;
    const std::uint32_t top = 1u << key_bits;
    std::uint32_t key[kRounds], val[kRounds], off[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int i = wbase + r * kWave + lane;
        key[r] = 0;
        val[r] = 0;
        if (r < rounds && i < n) {
            const std::int32_t k = s.key[i];
            key[r] = k == kSentinel ? top : static_cast<std::uint32_t>(k);
            val[r] = s.val[i];
        }
    }
    // same holds shift + 1 of a skipped pass; it is written after the first
    // barrier of a pass, which every thread passes only after reading it.
    if (t == 0) x.same = -1;
    bool moved = false;
    for (int shift = 0; shift <= key_bits; shift += 8) {
        for (int i = t; i < (kBlockSortThreads / kWave) * 256; i += kBlockSortThreads) x.wcount[i / 256][i % 256] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const bool valid = r < rounds && wbase + r * kWave + lane < n;
            const std::uint32_t d = (key[r] >> shift) & 0xffu;
            unsigned long long peers = __ballot(valid);
            for (int b = 0; b < 8; ++b) {
                const bool bit = (d >> b) & 1u;
                const unsigned long long bal = __ballot(bit);
                peers &= bit ? bal : ~bal;
            }
            const std::uint32_t before = valid ? x.wcount[w][d] : 0;
            off[r] = before + static_cast<std::uint32_t>(__popcll(peers & lt));
            if (valid && (peers & lt) == 0) x.wcount[w][d] = before + static_cast<std::uint32_t>(__popcll(peers));
        }
        __syncthreads();
        std::uint32_t run = 0;
        for (int j = 0; j < kBlockSortThreads / kWave; ++j) {
            const std::uint32_t c = x.wcount[j][t];
            x.wcount[j][t] = run;
            run += c;
        }
        if (run == static_cast<std::uint32_t>(n)) x.same = shift + 1;
        std::uint32_t inc, exc, tot;
        block_scan<PlusOp>(run, x.wsum, inc, exc, tot);
        if (x.same == shift + 1) continue;
        x.start[t] = exc;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            if (r < rounds && wbase + r * kWave + lane < n) {
                const std::uint32_t d = (key[r] >> shift) & 0xffu;
                const std::uint32_t p = x.start[d] + x.wcount[w][d] + off[r];
                s.key[p] = static_cast<std::int32_t>(key[r]);
                s.val[p] = val[r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const int i = wbase + r * kWave + lane;
            if (r < rounds && i < n) {
                key[r] = static_cast<std::uint32_t>(s.key[i]);
                val[r] = s.val[i];
            }
        }
        moved = true;
    }
    if (moved) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const int i = wbase + r * kWave + lane;
            if (r < rounds && i < n) s.key[i] = key[r] == top ? kSentinel : static_cast<std::int32_t>(key[r]);
        }
    }
    __syncthreads();
}

// Run reduction of sorted s.key[0, N) (N from block_pad).
// begin(total) is called by every thread once the emission count is known,
// before the first emit.
template <typename Emit, typename Begin = NoBegin>
__device__ inline int block_reduce_sorted(BlockSortSmem& s, int N, Emit&& emit, Begin&& begin = Begin{}) {
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

template <typename Emit, typename Begin = NoBegin>
__device__ inline int block_sort_reduce(BlockSortSmem& s, int n, Emit&& emit, Begin&& begin = Begin{}) {
    const int N = block_pad(s, n);
    block_bitonic(s, N);
    return block_reduce_sorted(s, N, static_cast<Emit&&>(emit), static_cast<Begin&&>(begin));
}

// Block-wide aggregation of (key >= 0, u32 weight) pairs in an LDS hash
// table with linear probing.  Sums are exact while every row's total weight
// stays below 2^32 (callers check the level's row-weight bound).  Slot order
// is arbitrary: callers use it only where the result is order-independent.
constexpr int kHashSlots = 8192;

template <int Slots>
struct BlockHashT {
    static constexpr int slots = Slots;
    std::int32_t key[Slots];
    std::uint32_t val[Slots];
};
using BlockHashSmem = BlockHashT<kHashSlots>;

// Table size for n entries: a power of two >= 2n (at least 512), at most
// the table's capacity (rows of a class fill at most three quarters).
template <class H = BlockHashSmem>
__device__ inline int hash_slots_for(int n) {
    int h = 512;
    while (h < 2 * n && h < H::slots) h <<= 1;
    return h;
}

template <class H>
__device__ inline void block_hash_clear(H& s, int slots) {
    for (int i = threadIdx.x; i < slots; i += blockDim.x) {
        s.key[i] = -1;
        s.val[i] = 0;
    }
}

template <class H>
__device__ inline void block_hash_add(H& s, int slots, std::int32_t key, std::uint32_t w) {
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
