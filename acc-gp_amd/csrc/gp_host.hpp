// Host-side components: initial partitioner of the coarsest graph, synthetic
// graph generators, binary CSR I/O and an exact cut/load checker.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gp {

struct HostGraph {
    std::int64_t n = 0;
    std::vector<std::int32_t> off, nbr;
    std::vector<std::uint32_t> ew;  // empty: unit weights
    std::vector<std::uint32_t> vw;  // empty: unit weights
};

// Deterministic recursive-bisection partitioner for the coarsest graph.
// This is not METIS: breadth-first region growing from hashed seeds followed
// by greedy boundary refinement, best of four trials per bisection.
void initial_partition_host(std::int64_t n, const std::int32_t* off, const std::int32_t* nbr, const std::uint32_t* ew,
                            const std::uint32_t* vw, int parts, double imbalance, std::uint32_t seed,
                            std::int32_t* labels);

struct InitialReport {
    const char* backend = "";
    std::uint64_t vertex_gcd = 0, edge_gcd = 0;
    std::uint64_t max_load = 0;   // largest part load before any repair
    std::uint64_t repaired = 0;   // vertices moved by the strict repair
    std::uint64_t edgecut = 0;    // exact cut of the returned labels
    double seconds = 0.0;
};

// The initial-partition boundary of initial_partition_metis
// (src/initial_partition_metis.cu:86-183): checks, exact GCD normalisation,
// METIS when built with GP_WITH_METIS and the recursive bisection otherwise,
// and with require_feasible the strict single-move repair.
void initial_partition(std::int64_t n, const std::int32_t* off, const std::int32_t* nbr, const std::uint32_t* ew,
                       const std::uint32_t* vw, int parts, double imbalance, std::uint32_t seed, bool require_feasible,
                       std::int32_t* labels, InitialReport* report = nullptr);
// repair_capacity (src/initial_partition_metis.cu:35-83): single moves until
// no part exceeds capacity; loads in/out; returns the vertices moved; throws
// when no feasible move is left.
std::size_t repair_capacity(std::int64_t n, const std::int32_t* off, const std::int32_t* nbr, const std::uint32_t* ew,
                            const std::uint32_t* vw, int parts, std::uint64_t capacity, std::int32_t* labels,
                            std::vector<std::uint64_t>& loads);
// "metis" or "bisection" (the build's initial partitioner).
const char* initial_partition_backend();

// Generators (rows sorted, symmetric, no self loops, no duplicates).
HostGraph generate_grid(std::int64_t rows, std::int64_t cols);
HostGraph generate_rmat(int scale, int edge_factor, double a, double b, double c, std::uint64_t seed);
HostGraph generate_rgg2d(std::int64_t n, double radius, std::uint64_t seed);
// Union-symmetrised k-nearest-neighbour graph of uniform points in [0,1)^3
// with hashed edge weights 1..100 and vertex weights 1..4.
HostGraph generate_knn3d(std::int64_t n, int knn, std::uint64_t seed);

// Reference on-disk format: little-endian int64 indptr[n+1], indices[m].
HostGraph load_csr(const std::string& indptr_path, const std::string& indices_path);
void save_csr(const HostGraph& g, const std::string& indptr_path, const std::string& indices_path);

// Exact undirected cut and part loads.
std::uint64_t host_cut(const HostGraph& g, const std::int32_t* labels);
std::vector<std::uint64_t> host_loads(const HostGraph& g, const std::int32_t* labels, int parts);

std::uint64_t capacity_of(std::uint64_t total, int parts, double imbalance);

}  // namespace gp
