/*
 * gpart_mi355x.h -- C ABI of the MI355X multilevel k-way graph partitioner.
 *
 * The reference (Shixg-cnic/Acc-GP) exposes a C++ API over thrust device
 * vectors and has no C binding; each entry point below names the reference
 * interface it replaces.  Plain pointers and sizes only.  Every function
 * returns 0 (GP_OK) on success or a gp_status code, with a message available
 * from gp_last_error() on the calling thread (the reference throws
 * std::runtime_error / invalid_argument / overflow_error instead,
 * include/check.hpp:8-18).
 *
 * Graph type: int32 vertices and offsets, uint32 vertex and edge weights,
 * 64-bit cut / load sums, 2 <= parts <= 128 (include/graph_types.hpp:4-10).
 */
#ifndef GPART_MI355X_H
#define GPART_MI355X_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gp_status {
    GP_OK = 0,
    GP_ERR_INVALID = 1,   /* std::invalid_argument in the reference */
    GP_ERR_RUNTIME = 2,   /* std::runtime_error / overflow_error */
    GP_ERR_HIP = 3,       /* device failure (CUDA_CHECK in the reference) */
    GP_ERR_MEMORY = 4
} gp_status;

typedef struct gp_context gp_context;     /* one GPU: stream + HBM arena */
typedef struct gp_graph gp_graph;         /* device-resident weighted CSR */
typedef struct gp_hierarchy gp_hierarchy; /* coarsening levels + maps */
typedef struct gp_result gp_result;       /* uncoarsening result (labels on device) */

typedef struct gp_coarsen_options {       /* include/coarsen.hpp:15-38 */
    int parts;
    uint32_t seed;
    double stop_ratio;
    int max_levels;
    int cluster_size_factor;
    int pack_isolated;
    int progressive;
    int fine_factor;
    int capacity_growth;
    int seed_matching;
    int fast_path;
    int hem_degree_limit;
} gp_coarsen_options;

typedef struct gp_refine_options {        /* include/refine.hpp:13-26 */
    int parts;
    double imbalance;
    int max_rounds;
    int stale_rounds;
    double growth_ratio;
    int exchange_rounds;
    int strict_verify;                    /* recount cut and loads after every batch and projection */
    int gpu_lp_loop;                      /* 1: LP round loop decided on the device (default); 0: on the host
                                             after a counter read-back per round (include/refine.hpp:23) */
} gp_refine_options;

typedef struct gp_level_stats {           /* RefineLevelResult, include/refine.hpp:60-68 */
    int64_t n, m;
    int refined;
    int initial_repair_failed;
    int rounds;
    uint64_t initial_cut, final_cut;
    uint64_t lp_moves, repair_moves, group_moves;
    int repair_iterations;
    uint64_t strict_checks;               /* recounts that matched under strict_verify */
} gp_level_stats;

typedef struct gp_partition_stats {       /* partition_timing line, app/partition_main.cu:163-177 */
    double coarsen_seconds;
    double initial_partition_seconds;
    double uncoarsen_seconds;
    double algorithm_total_seconds;
    int levels;
    int64_t coarsest_vertices;
    uint64_t final_cut;
    uint64_t max_part_weight;
    uint64_t capacity;
    uint64_t total_edge_weight;
    uint64_t launches;      /* no counterpart: kernel launches and copies issued by the partition */
    uint64_t host_syncs;    /* no counterpart: host waits on the stream (read-backs included) */
} gp_partition_stats;

const char* gp_last_error(void);

/* ---- context -------------------------------------------------------- */
/* Replaces the thread-local ScratchPoolSession (include/detail/scratch_pool.hpp:115-130). */
int gp_context_create(int device, gp_context** out);
void gp_context_destroy(gp_context* ctx);
/* Cached HBM bytes held by the context arena. */
uint64_t gp_context_reserved_bytes(const gp_context* ctx);
/* Tuning, no counterpart in the reference: on (default) runs the launch of
 * the longest rows of each row sweep on a second stream beside the launch of
 * medium rows; off keeps every launch on the context stream (one hardware
 * queue per context, which the concurrent contexts of gp_server use).
 * GP_TIER_OVERLAP=0 in the environment sets off at context creation. */
int gp_context_set_overlap(gp_context* ctx, int on);

/* ---- graphs (include/graph.hpp:59-76) -------------------------------- */
/* make_device_weighted: uploads a host CSR.  ew==NULL or all-ones edge weights
 * are stored implicitly; vw==NULL means unit vertex weights. */
int gp_graph_create(gp_context* ctx, int64_t n, int64_t m, const int32_t* offsets, const int32_t* neighbors,
                    const uint32_t* edge_weights, const uint32_t* vertex_weights, gp_graph** out);
int gp_graph_info(const gp_graph* g, int64_t* n, int64_t* m);
/* copy_device_weighted: any output pointer may be NULL; ew expands unit weights. */
int gp_graph_download(gp_context* ctx, const gp_graph* g, int32_t* offsets, int32_t* neighbors,
                      uint32_t* edge_weights, uint32_t* vertex_weights);
/* DeviceWeightedGraph metadata (include/graph.hpp:34-57): degree bound,
 * heavy rows, incident-weight bound (unknown: UINT64_MAX / 1 until the
 * graph has been coarsened or aggregated) and implicit unit edge weights. */
int gp_graph_metadata(const gp_graph* g, uint64_t* max_degree, int* has_heavy, uint64_t* incident_bound,
                      int* implicit_unit_weights);
/* Changes a device graph in place by count undirected edge changes (u[i],
 * v[i]): insert[i] != 0 inserts the edge with weights[i] (weights NULL: 1),
 * else deletes it.  No counterpart in the reference (its graphs are read
 * once); the request server's resident graphs use it for repartitioning.
 * Deletions apply first, then insertions; deleting an absent edge changes
 * nothing, inserting a present or repeated edge is dropped.  Each row keeps
 * its surviving entries in their old order followed by its new neighbours in
 * ascending order.  m (optional) gets the new entry count.  GP_ERR_INVALID
 * for a vertex out of range, a self loop or a zero weight (nothing changes). */
int gp_graph_apply_delta(gp_context* ctx, gp_graph* g, int64_t count, const int32_t* u, const int32_t* v,
                         const uint32_t* weights, const uint8_t* insert, int64_t* m);
void gp_graph_destroy(gp_graph* g);

/* partition_capacity (include/partition_capacity.hpp:14-25). */
int gp_capacity(uint64_t total, int parts, double imbalance, uint64_t* out);

/* ---- coarsening (include/coarsen.hpp:94-112) -------------------------- */
void gp_default_coarsen_options(int parts, uint32_t seed, gp_coarsen_options* out);
/* coarsen(DeviceWeightedGraph&&, CoarsenOptions): the input graph becomes
 * level 0 and must outlive the hierarchy (it is borrowed, not consumed). */
int gp_coarsen(gp_context* ctx, gp_graph* input, const gp_coarsen_options* opts, gp_hierarchy** out);
int gp_hierarchy_levels(const gp_hierarchy* h);
/* Borrowed view of level i (0 = input). */
const gp_graph* gp_hierarchy_level(const gp_hierarchy* h, int level);
/* fine_to_coarse map of level i (length = vertices of level i), i < levels-1. */
int gp_hierarchy_map_download(gp_context* ctx, const gp_hierarchy* h, int level, int32_t* map);
/* Level metadata carried by contraction: max degree bound, heavy rows, incident weight bound. */
int gp_hierarchy_level_info(const gp_hierarchy* h, int level, uint64_t* max_degree, int* has_heavy,
                            uint64_t* incident_bound);
/* A hierarchy from levels built elsewhere (the host-hierarchy overloads of
 * refine_hierarchy_device / uncoarsen, include/refine.hpp:90-94 and
 * include/uncoarsen.hpp:11-14): level_graphs[0..levels) made by
 * gp_graph_create or gp_contract, finest first, and host maps[i] of
 * level i's length with ids in [0, n_{i+1}).  On success the hierarchy owns
 * every level graph (their handles are no longer the caller's; level 0 is
 * released with the hierarchy); on failure nothing is taken.  Level
 * metadata is recomputed exactly; the cluster caps are unknown (UINT64_MAX). */
int gp_hierarchy_assemble(gp_context* ctx, int levels, gp_graph* const* level_graphs, const int32_t* const* maps,
                          gp_hierarchy** out);
/* Undirected edge weight of a device graph (DeviceUncoarsenResult::total_edge_weight,
 * src/refine.cu:704-712); GP_ERR_RUNTIME when the directed sum is odd. */
int gp_graph_edge_weight(gp_context* ctx, const gp_graph* g, uint64_t* total);
/* Cluster weight cap of the aggregation that produced level+1 (the cap
 * checked by validate_coarsening_step, src/coarsen.cu:1692-1758), level < levels-1. */
int gp_hierarchy_level_capacity(const gp_hierarchy* h, int level, uint64_t* capacity);
/* DeviceHierarchy::stop_reason: "capacity_floor", "insufficient_contraction" or "max_levels". */
const char* gp_hierarchy_stop_reason(const gp_hierarchy* h);
void gp_hierarchy_destroy(gp_hierarchy* h);

/* One level at a time: aggregate() and contract() with an SclpWorkspace
 * (include/coarsen.hpp:55-104).  The workspace carries the coarsening state
 * from a graph to the graph gp_contract makes from it, so aggregating level 0,
 * contracting, aggregating the result, ... reproduces gp_coarsen's levels
 * when the caller passes the level loop's per-level seed, factor and
 * seed-matching decision; a graph the workspace did not produce starts from
 * its level-0 metadata. */
typedef struct gp_workspace gp_workspace;
typedef struct gp_aggregate gp_aggregate; /* DeviceAggregateResult: map + coarse vertex weights */
int gp_workspace_create(gp_workspace** out);
void gp_workspace_destroy(gp_workspace* ws);
int gp_aggregate_level(gp_context* ctx, gp_workspace* ws, gp_graph* g, int parts, uint32_t seed, int level,
                       int cluster_size_factor, int pack_isolated, int seed_matching, int fast_path,
                       gp_aggregate** out);
int gp_aggregate_info(const gp_aggregate* a, int64_t* coarse_vertices, uint64_t* maximum_weight, uint64_t* capacity);
int gp_aggregate_map_download(gp_context* ctx, const gp_aggregate* a, int32_t* map);
int gp_aggregate_vertex_weights_download(gp_context* ctx, const gp_aggregate* a, uint32_t* vertex_weights);
void gp_aggregate_destroy(gp_aggregate* a);
/* contract(fine, std::move(aggregate), workspace): the coarse graph; the
 * aggregate's vertex weights move into it (its map stays readable). */
int gp_contract(gp_context* ctx, gp_workspace* ws, const gp_graph* fine, gp_aggregate* a, gp_graph** out);

/* ---- initial partition (include/initial_partition.hpp:6-8) ------------ */
/* initial_partition_metis(graph, parts, eps, seed, require_feasible) on a host
 * CSR (src/initial_partition_metis.cu:86-183): positive vertex weights, none
 * above the integer capacity, exact GCD normalisation of vertex and edge
 * weights, self loops and zero-weight edges left out; METIS 5 when the
 * library was built with it (gp_initial_partition_backend() == "metis"),
 * the in-repo recursive bisection otherwise ("bisection", not METIS).  With
 * require_feasible, the normalised weights must admit the capacity and the
 * strict single-move repair of gp_repair_capacity_host follows; production
 * (gp_partition) passes 0, as the reference CLI does. */
int gp_initial_partition_host(int64_t n, const int32_t* offsets, const int32_t* neighbors,
                              const uint32_t* edge_weights, const uint32_t* vertex_weights, int parts,
                              double imbalance, uint32_t seed, int require_feasible, int32_t* labels);
/* The same on a device graph (downloaded to the host first); labels are host output. */
int gp_initial_partition(gp_context* ctx, const gp_graph* g, int parts, double imbalance, uint32_t seed,
                         int require_feasible, int32_t* labels);
/* repair_capacity (src/initial_partition_metis.cu:35-83): while a part is
 * above capacity, move the vertex of an overloaded part to a part with room
 * for all of it whose (A(v,p) - A(v,s)) / min(w, load_s - capacity) is
 * largest (first in vertex, part order on ties).  labels in/out; moved gets
 * the count; GP_ERR_RUNTIME when no feasible move is left. */
int gp_repair_capacity_host(int64_t n, const int32_t* offsets, const int32_t* neighbors,
                            const uint32_t* edge_weights, const uint32_t* vertex_weights, int parts, uint64_t capacity,
                            int32_t* labels, uint64_t* moved);
/* "metis" or "bisection". */
const char* gp_initial_partition_backend(void);

/* ---- uncoarsening / refinement (include/uncoarsen.hpp:17-20, refine.hpp:84-98) */
void gp_default_refine_options(int parts, double imbalance, gp_refine_options* out);
/* uncoarsen(DeviceHierarchy&, coarsest labels, RefineOptions). */
int gp_uncoarsen(gp_context* ctx, const gp_hierarchy* h, const int32_t* coarsest_labels,
                 const gp_refine_options* opts, gp_result** out);
int gp_result_labels(gp_context* ctx, const gp_result* r, int32_t* labels);
uint64_t gp_result_cut(const gp_result* r);
int gp_result_loads(const gp_result* r, uint64_t* loads);
int gp_result_levels(const gp_result* r);
int gp_result_level_stats(const gp_result* r, int index, gp_level_stats* out);
void gp_result_destroy(gp_result* r);
/* device_partition of DeviceUncoarsenResult (include/refine.hpp:74-82): the
 * result's labels in device memory, valid until gp_result_destroy. */
const int32_t* gp_result_device_labels(const gp_result* r);
/* Edge cut and part loads of device-resident labels (host_cut,
 * include/graph.hpp:67-69; the device recounts of src/refine.cu:63-98):
 * labels is a device pointer to n entries in [0, parts); cut and loads
 * (parts words) are host outputs.  GP_ERR_INVALID for labels out of range.
 * Partials come from the context arena; synchronizes the context stream. */
int gp_edge_cut(gp_context* ctx, const gp_graph* g, const int32_t* labels, int parts, uint64_t* cut,
                uint64_t* loads);
/* gp_edge_cut without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as gp_edge_cut; the per-block partials live
 * in a static device array of 1025 x 130 words (1,066,000 bytes, held by
 * each GPU the library runs on until the process ends) instead of the
 * context arena, so the call makes no hipMalloc and touches no arena.  It
 * covers every graph size at parts <= 128 (GP_ERR_INVALID above).  Calls
 * that overlap on one GPU share that array and race; the array keeps the
 * last call's partials. */
int gp_edge_cut_noalloc(gp_context* ctx, const gp_graph* g, const int32_t* labels, int parts, uint64_t* cut,
                        uint64_t* loads);
/* refine_partition(graph, labels, options): one level, labels in/out (host). */
int gp_refine_partition(gp_context* ctx, const gp_graph* g, int32_t* labels, const gp_refine_options* opts,
                        gp_level_stats* stats);

/* ---- whole pipeline: the CLI's algorithm_total_seconds scope ---------- */
/* coarsen -> host initial partition -> uncoarsen, input resident on device.
 * labels may be NULL (result stays on device and is released). */
int gp_partition(gp_context* ctx, gp_graph* g, int parts, double imbalance, uint32_t seed, int32_t* labels,
                 gp_partition_stats* stats);

/* ---- synthetic inputs and checks (host only) -------------------------- */
/* Generated CSR arrays are allocated by the library; release with gp_free_host. */
int gp_generate_grid(int64_t rows, int64_t cols, int64_t* n, int64_t* m, int32_t** offsets, int32_t** neighbors);
int gp_generate_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, int64_t* n, int64_t* m,
                     int32_t** offsets, int32_t** neighbors);
int gp_generate_rgg2d(int64_t n_points, double radius, uint64_t seed, int64_t* n, int64_t* m, int32_t** offsets,
                      int32_t** neighbors);
int gp_generate_knn3d(int64_t n_points, int knn, uint64_t seed, int64_t* n, int64_t* m, int32_t** offsets,
                      int32_t** neighbors, uint32_t** edge_weights, uint32_t** vertex_weights);
void gp_free_host(void* p);
/* Reference int64 binary CSR reader/writer (src/graph.cpp:47-116). */
int gp_load_csr(const char* indptr_path, const char* indices_path, int64_t* n, int64_t* m, int32_t** offsets,
                int32_t** neighbors);
int gp_save_csr(const char* indptr_path, const char* indices_path, int64_t n, int64_t m, const int32_t* offsets,
                const int32_t* neighbors);

/* ---- batched and served partitioning ----------------------------------
 * No counterpart in the reference (its CLI partitions one graph per process).
 * Independent requests run concurrently on one GPU, each in its own context;
 * a request is admitted while the estimated HBM footprint of the requests in
 * flight stays within the budget.  Each result equals the same request run
 * alone through gp_partition (cold) or gp_refine_partition (warm start from
 * warm_labels).  The caller keeps every pointer valid until the request is
 * waited for. */
typedef struct gp_request {
    int64_t n, m;
    const int32_t* offsets;
    const int32_t* neighbors;
    const uint32_t* edge_weights;     /* NULL: unit */
    const uint32_t* vertex_weights;   /* NULL: unit */
    int parts;
    double imbalance;
    uint32_t seed;
    const int32_t* warm_labels;       /* NULL: cold partition; else refine from these labels */
    int32_t* labels;                  /* out: n labels */
    /* Resident graphs (gp_server_graph_add): graph_id != 0 runs the request on
     * that graph instead of n/m/offsets/...; its delta_count edge changes
     * (gp_graph_apply_delta) apply first; warm_resident refines from the
     * graph's current partition (the result of the last request on it). */
    uint64_t graph_id;
    int64_t delta_count;
    const int32_t* delta_u;
    const int32_t* delta_v;
    const uint32_t* delta_w;          /* NULL: 1 */
    const uint8_t* delta_insert;      /* 1 insert, 0 delete */
    int warm_resident;
    gp_partition_stats stats;         /* out */
    int status;                       /* out: gp_status (-1 while pending) */
    double queue_seconds;             /* out: submission to admission */
    double device_seconds;            /* out: upload + partition + download */
    char error[256];                  /* out */
} gp_request;

typedef struct gp_server gp_server;
typedef struct gp_ticket gp_ticket;

/* workers: concurrent requests (contexts); hbm_budget_bytes 0 = 64 GiB. */
int gp_server_create(int device, int workers, uint64_t hbm_budget_bytes, gp_server** out);
/* Keeps a graph in HBM for the server's requests (uploaded once); graph_id
 * out.  Requests on one resident graph run one at a time in submission
 * order. */
int gp_server_graph_add(gp_server* s, int64_t n, int64_t m, const int32_t* offsets, const int32_t* neighbors,
                        const uint32_t* edge_weights, const uint32_t* vertex_weights, uint64_t* graph_id);
/* The resident graph as it stands (any output may be NULL; call with NULL
 * arrays first for n and m) and its current partition; GP_ERR_INVALID while
 * requests on it are queued or running. */
int gp_server_graph_download(gp_server* s, uint64_t graph_id, int64_t* n, int64_t* m, int32_t* offsets,
                             int32_t* neighbors, uint32_t* edge_weights, int32_t* labels);
/* Frees a resident graph; GP_ERR_INVALID while requests on it are pending. */
int gp_server_graph_release(gp_server* s, uint64_t graph_id);
int gp_server_submit(gp_server* s, gp_request* r, gp_ticket** ticket);
/* Blocks until the request is done; releases the ticket; returns its status
 * (GP_ERR_INVALID for a ticket this server does not hold). */
int gp_server_wait(gp_server* s, gp_ticket* ticket);
/* Finishes queued requests, then stops the workers and releases the tickets
 * not waited for (their requests have completed; the caller's request
 * buffers must stay valid until this returns). */
void gp_server_destroy(gp_server* s);
/* Whole batch through a temporary server, largest request first. */
int gp_partition_batch(int device, gp_request* items, int64_t count, int workers, uint64_t hbm_budget_bytes);

/* Kernel-time accounting (HIP events on the context stream) for roofline reports. */
int gp_profile_enable(gp_context* ctx, int enable);
/* name: kernel class; returns accumulated ms, algorithmic bytes and launches. */
int gp_profile_query(gp_context* ctx, const char* name, double* ms, double* bytes, int64_t* launches);
int gp_profile_reset(gp_context* ctx);

/* The stale-round rule of small levels (incremental.cuh:911-961: a new best
 * cut resets the stale count when cut < 0.999 * best in long double) as the
 * device-side loop control evaluates it, in integers; 1 or 0.  For tests. */
int gp_debug_below_0999(uint64_t cut, uint64_t best);

/* The device radix sort behind the admission order (capacity_admission.hpp:14-33:
 * proposals ordered by score, target and vertex) and the contraction's record
 * order, on host arrays: stable by bits [begin_bit, end_bit) of keys, vals
 * (may be NULL) carried along, sorted in place.  The passes take whole 8-bit
 * digits from begin_bit, so key bits from end_bit up to the end of the last
 * digit must be zero (the callers' keys are exactly end_bit wide).  For
 * tests. */
int gp_debug_radix_sort(gp_context* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int begin_bit, int end_bit);

#ifdef __cplusplus
}
#endif
#endif
