// Edge cut and part loads of device-resident labels: the verification
// counts of the reference (host_cut, include/graph.hpp:67-69; the device cut
// and load recounts of src/refine.cu:63-98 and the final check of
// src/uncoarsen.cu:90-100), in two passes with no atomics to global memory.
//
// Pass 1: each block walks chunks of 64 consecutive vertices per wave.  A
// wave stages the chunk's 65 offsets and 64 labels in LDS, adds the vertex
// weights into a wave-private LDS table of part loads, then its lanes walk
// the chunk's entries in coalesced 64-entry strides, find each entry's
// source row by a binary search over the staged offsets and compare the
// source label with the gathered neighbour label.  The block writes one row
// of partials: [directed cut weight, labels outside [0, k), load_0 ..
// load_{k-1}].  Pass 2: one block sums the rows column by column.
//
// Bound: HBM.  Algorithmic bytes per call: 4(n+1) offsets + 4n labels + 4n
// vertex weights + m (4 neighbour + 4 neighbour label [+ 4 edge weight]).
#include "gp_metrics.hpp"

#include <mutex>
#include <stdexcept>
#include <unordered_map>

// Scratch of the allocation-free evaluation (edge_cut_noalloc): room for the
// partial rows and the totals row of the largest evaluation, 1024 partial
// blocks + 1 totals row of 130 words (cut, out-of-range count, 128 loads),
// 1,066,000 bytes.  Every GPU holds its own copy until the process ends.
// Host code never passes this name to a kernel: it is a host-side stand-in,
// and the device address comes from hipGetSymbolAddress per device.
static __device__ __align__(256) unsigned long long scratch[gp::kNoAllocScratchWords];

namespace gp {

namespace {

constexpr int kCutThreads = 256;
constexpr int kCutWaves = kCutThreads / kWave;

template <bool Weighted>
__global__ void __launch_bounds__(kCutThreads) k_cut_partials(std::int64_t n, const std::int32_t* off,
                                                              const std::int32_t* nbr, const std::uint32_t* ew,
                                                              const std::uint32_t* vw, const std::int32_t* labels,
                                                              int k, unsigned long long* partials) {
    __shared__ unsigned long long loads[kCutWaves][kMaxParts];
    __shared__ std::int32_t soff[kCutWaves][kWave + 1];
    __shared__ std::int32_t slab[kCutWaves][kWave];
    __shared__ unsigned long long red[kCutWaves][2];
    const int lane = lane_id(), w = threadIdx.x / kWave;
    for (int p = lane; p < k; p += kWave) loads[w][p] = 0;
    wave_sync();
    unsigned long long cut = 0, bad = 0;
    const std::int64_t chunks = (n + kWave - 1) / kWave;
    for (std::int64_t ch = (std::int64_t)blockIdx.x * kCutWaves + w; ch < chunks;
         ch += (std::int64_t)gridDim.x * kCutWaves) {
        const std::int64_t v0 = ch * kWave;
        const int cnt = static_cast<int>(n - v0 < kWave ? n - v0 : kWave);
        if (lane < cnt) {
            const std::int32_t lab = labels[v0 + lane];
            soff[w][lane] = off[v0 + lane];
            slab[w][lane] = lab;
            if (static_cast<unsigned>(lab) < static_cast<unsigned>(k))
                atomicAdd(&loads[w][lab], static_cast<unsigned long long>(vw[v0 + lane]));
            else
                ++bad;
        }
        if (lane == 0) soff[w][cnt] = off[v0 + cnt];
        wave_sync();
        const std::int32_t e0 = soff[w][0], e1 = soff[w][cnt];
        for (std::int32_t e = e0 + lane; e < e1; e += kWave) {
            int lo = 0, hi = cnt - 1;  // last row whose first entry is <= e
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (soff[w][mid] <= e) lo = mid;
                else hi = mid - 1;
            }
            if (slab[w][lo] != labels[nbr[e]]) cut += Weighted ? static_cast<unsigned long long>(ew[e]) : 1ULL;
        }
        wave_sync();
    }
    cut = wave_sum(cut);
    bad = wave_sum(bad);
    if (lane == 0) {
        red[w][0] = cut;
        red[w][1] = bad;
    }
    __syncthreads();
    unsigned long long* row = partials + static_cast<std::size_t>(blockIdx.x) * (k + 2);
    if (threadIdx.x < 2) {
        unsigned long long s = 0;
        for (int x = 0; x < kCutWaves; ++x) s += red[x][threadIdx.x];
        row[threadIdx.x] = s;
    }
    for (int p = threadIdx.x; p < k; p += kCutThreads) {
        unsigned long long s = 0;
        for (int x = 0; x < kCutWaves; ++x) s += loads[x][p];
        row[2 + p] = s;
    }
}

__global__ void __launch_bounds__(kCutThreads) k_cut_finish(const unsigned long long* partials, int blocks, int cols,
                                                            unsigned long long* out) {
    for (int c = threadIdx.x; c < cols; c += kCutThreads) {
        unsigned long long s = 0;
        for (int b = 0; b < blocks; ++b) s += partials[static_cast<std::size_t>(b) * cols + c];
        out[c] = s;
    }
}

}  // namespace

int edge_cut_blocks(std::int64_t n) {
    const std::int64_t chunks = (n + kWave - 1) / kWave;
    const std::int64_t b = (chunks + kCutWaves - 1) / kCutWaves;
    return static_cast<int>(b < 1 ? 1 : (b > kEdgeCutMaxBlocks ? kEdgeCutMaxBlocks : b));
}

std::size_t edge_cut_scratch_words(std::int64_t n, int k) {
    return static_cast<std::size_t>(edge_cut_blocks(n) + 1) * static_cast<std::size_t>(k + 2);
}

void edge_cut_into(Context& c, const DGraph& g, const std::int32_t* labels, int k, unsigned long long* scratch,
                   std::uint64_t* cut, std::uint64_t* loads) {
    if (k < 2 || k > kMaxParts) throw std::invalid_argument("parts must be in [2, 128]");
    if (!labels && g.n > 0) throw std::invalid_argument("labels must be a device pointer");
    const int blocks = edge_cut_blocks(g.n);
    const int cols = k + 2;
    unsigned long long* partials = scratch;
    unsigned long long* out = scratch + static_cast<std::size_t>(blocks) * cols;
    const hipStream_t s = c.stream();
    if (g.ew)
        hipLaunchKernelGGL(k_cut_partials<true>, dim3(blocks), dim3(kCutThreads), 0, s, g.n, g.off.get(), g.nbr.get(),
                           g.ew.get(), g.vw.get(), labels, k, partials);
    else
        hipLaunchKernelGGL(k_cut_partials<false>, dim3(blocks), dim3(kCutThreads), 0, s, g.n, g.off.get(),
                           g.nbr.get(), nullptr, g.vw.get(), labels, k, partials);
    GP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cut_finish, dim3(1), dim3(kCutThreads), 0, s, partials, blocks, cols, out);
    GP_LAUNCH_CHECK();
    unsigned long long host[kMaxParts + 2];
    c.read(host, out, static_cast<std::size_t>(cols) * sizeof(unsigned long long));
    if (host[1]) throw std::invalid_argument("labels outside [0, parts)");
    if (host[0] & 1ULL) throw std::runtime_error("directed cut is odd: the graph is not symmetric");
    *cut = host[0] / 2;
    for (int p = 0; p < k; ++p) loads[p] = host[2 + p];
}

void edge_cut(Context& c, const DGraph& g, const std::int32_t* labels, int k, std::uint64_t* cut,
              std::uint64_t* loads) {
    DBuf<unsigned long long> buf = c.buf<unsigned long long>(edge_cut_scratch_words(g.n, k));
    edge_cut_into(c, g, labels, k, buf.get(), cut, loads);
}

namespace {

std::mutex g_scratch_mutex;
std::unordered_map<int, unsigned long long*> g_scratch_address;  // device -> address of `scratch`

}  // namespace

unsigned long long* resolve_noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    auto it = g_scratch_address.find(device);
    if (it != g_scratch_address.end()) return it->second;
    int current = -1;
    GP_HIP(hipGetDevice(&current));
    if (current != device) throw std::logic_error("resolve_noalloc_scratch: device is not current");
    void* ptr = nullptr;
    GP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw HipError("hipGetSymbolAddress returned no address for the evaluation scratch");
    auto* p = static_cast<unsigned long long*>(ptr);
    g_scratch_address.emplace(device, p);
    return p;
}

void edge_cut_noalloc(Context& c, const DGraph& g, const std::int32_t* labels, int k, std::uint64_t* cut,
                      std::uint64_t* loads) {
    if (k < 2 || k > kMaxParts) throw std::invalid_argument("parts must be in [2, 128]");
    const std::size_t words = edge_cut_scratch_words(g.n, k);
    if (words > kNoAllocScratchWords)
        throw std::invalid_argument("evaluation needs " + std::to_string(words) + " scratch words, above the " +
                                    std::to_string(kNoAllocScratchWords) + " of the static scratch");
    unsigned long long* p = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        auto it = g_scratch_address.find(c.device());
        if (it != g_scratch_address.end()) p = it->second;
    }
    if (!p) throw std::logic_error("evaluation scratch of this device was not resolved at context creation");
    // No word of the scratch has to start at zero: the partial pass writes
    // every row the finishing pass reads, and the finishing pass writes the
    // totals row that is read back.
    edge_cut_into(c, g, labels, k, p, cut, loads);
}

}  // namespace gp
