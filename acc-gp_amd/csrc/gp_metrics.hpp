// Edge cut and part loads of device-resident labels (host_cut,
// include/graph.hpp:67-69; refine.cu:63-98).
#pragma once
#include "gp_graph.hpp"

namespace gp {

// Blocks of the partial pass: one per four 64-vertex chunks, at most 1024.
constexpr int kEdgeCutMaxBlocks = 1024;
int edge_cut_blocks(std::int64_t n);
// Scratch words of one evaluation: (blocks + 1) rows of k + 2 partials.
std::size_t edge_cut_scratch_words(std::int64_t n, int k);

// Cut and loads of `labels` (device, n entries in [0, k)) with the partials
// in caller scratch of edge_cut_scratch_words(n, k) words; loads_host gets k
// words.  Synchronizes the context stream.
void edge_cut_into(Context& c, const DGraph& g, const std::int32_t* labels, int k, unsigned long long* scratch,
                   std::uint64_t* cut, std::uint64_t* loads_host);
// The same with the scratch from the context arena.
void edge_cut(Context& c, const DGraph& g, const std::int32_t* labels, int k, std::uint64_t* cut,
              std::uint64_t* loads_host);

// Words of the static device scratch of edge_cut_noalloc: the largest
// evaluation, kEdgeCutMaxBlocks partial rows + 1 totals row of k + 2 words.
constexpr std::size_t kNoAllocScratchWords =
    static_cast<std::size_t>(kEdgeCutMaxBlocks + 1) * static_cast<std::size_t>(kMaxParts + 2);
// Resolves (once per device, that device current) and caches the device
// address of the static scratch; called at context creation.
unsigned long long* resolve_noalloc_scratch(int device);
// edge_cut with its partials in the static scratch of the context's device:
// allocates nothing (no arena, no hipMalloc).  Calls that overlap on one GPU
// share the scratch and race.
void edge_cut_noalloc(Context& c, const DGraph& g, const std::int32_t* labels, int k, std::uint64_t* cut,
                      std::uint64_t* loads_host);

}  // namespace gp
