/*
 * gp_oracle.h -- CPU oracle for the multilevel k-way partitioner hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a single-threaded plain-C restatement of
 * the reference partitioner's coarsening, contraction, projection and
 * refinement schedule (Shixg-cnic/Acc-GP, files cited per function in
 * gp_oracle.c).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.  The product library
 * never links it.
 *
 * Where the reference depends on atomic arrival order, the oracle defines the
 * canonical (vertex-id) order that the GPU implementation also follows, so the
 * two produce bit-identical labels.
 */
#ifndef GP_ORACLE_H
#define GP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct og_hier og_hier;

typedef struct og_coarsen_opts {
    int parts;                 /* k, 2..128 */
    uint32_t seed;
    double stop_ratio;         /* 0.90 */
    int max_levels;            /* 24 */
    int cluster_size_factor;   /* beta */
    int pack_isolated;
    int progressive;           /* progressive cluster capacity */
    int fine_factor;           /* 64 */
    int capacity_growth;       /* 8 */
    int seed_matching;
    int fast_path;             /* capacity-safe commit-all */
    int hem_degree_limit;      /* 64 */
} og_coarsen_opts;

typedef struct og_refine_opts {
    int parts;
    double imbalance;          /* 1.03 */
    int max_rounds;            /* 32 */
    int stale_rounds;          /* 3 */
    double growth_ratio;       /* 2.0 */
    int exchange_rounds;       /* 2 */
    int graph_loop_threshold;  /* 65536: levels with n >= this use the integer stale rule */
    int exchange_max_n;        /* 65536 */
    int exchange_only;         /* checks only: skip LP rounds, run the exchange passes */
} og_refine_opts;

typedef struct og_level_stats {
    int64_t n, m;
    int refined;
    int initial_repair_failed;
    int rounds;
    uint64_t initial_cut, final_cut;
    uint64_t lp_moves, repair_moves, group_moves;
    int repair_iterations;
} og_level_stats;

/* Threads of the OpenMP loops (0 or less: unchanged); results do not
 * depend on it.  og_get_threads: the current maximum. */
void og_set_threads(int threads);
int og_get_threads(void);

/* Production option defaults (app/partition_main.cu:52-87). */
void og_default_coarsen_opts(int parts, uint32_t seed, og_coarsen_opts* out);
void og_default_refine_opts(int parts, double imbalance, og_refine_opts* out);

/* floor(total*eps/k) in long double with the feasibility check; 0 on error. */
uint64_t og_capacity(uint64_t total, int parts, double imbalance);

/* Coarsen a CSR graph.  ew==NULL means unit edge weights, vw==NULL unit
 * vertex weights.  Returns NULL on error (og_last_error()). */
og_hier* og_coarsen(int64_t n, int64_t m, const int32_t* off, const int32_t* nbr,
                    const uint32_t* ew, const uint32_t* vw, const og_coarsen_opts* opts);
int og_hier_levels(const og_hier* h);
void og_hier_level_info(const og_hier* h, int level, int64_t* n, int64_t* m,
                        uint64_t* max_degree, int* has_heavy, uint64_t* incident_bound);
/* ew/vw outputs are always written (unit weights expanded). */
void og_hier_level_copy(const og_hier* h, int level, int32_t* off, int32_t* nbr,
                        uint32_t* ew, uint32_t* vw);
void og_hier_map_copy(const og_hier* h, int level, int32_t* map);
void og_hier_free(og_hier* h);

/* Refine the hierarchy from the coarsest labels to the original graph.
 * level_stats (optional) receives one entry per level, coarsest first.
 * Returns 0 on success. */
int og_uncoarsen(const og_hier* h, const int32_t* coarsest_labels, const og_refine_opts* opts,
                 int32_t* out_labels, uint64_t* out_cut, uint64_t* out_loads,
                 og_level_stats* level_stats);

/* Refine one graph in place (refine_partition).  rounds = LP round budget. */
int og_refine_graph(int64_t n, int64_t m, const int32_t* off, const int32_t* nbr,
                    const uint32_t* ew, const uint32_t* vw, int32_t* labels,
                    const og_refine_opts* opts, int has_known_cut, uint64_t known_cut,
                    og_level_stats* stats);

/* Directed-cut / 2 and part loads of a labelled graph. */
uint64_t og_cut(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew,
                const int32_t* labels);

/* Exact contraction of a graph by a map (stage check). */
int og_contract_graph(int64_t n, int64_t m, const int32_t* off, const int32_t* nbr,
                      const uint32_t* ew, const int32_t* map, int64_t nc,
                      int32_t* coff, int32_t* cnbr, uint32_t* cew, int64_t* cm);

/* One coarsening admission batch on explicit requests (checks only). */
int og_admission_batch(int64_t cnt, const int32_t* targets, const uint64_t* keys, const uint32_t* ties,
                       const int32_t* vertices, const uint64_t* weights, const uint64_t* bases, uint64_t cap,
                       uint8_t* accepted);

/* ---- single stages of the schedule (the oracle's own checks) ----------
 * Each runs one stage on caller arrays through the functions the level loops
 * use.  vw is required (no implicit unit vertex weights); ew may be NULL. */
/* Seed matching with the given salt: roots[v] = cluster root; returns the cluster count. */
int64_t og_t_seed_match(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew, const uint32_t* vw,
                        uint64_t cap, uint32_t salt, int32_t* roots);
/* Mover threshold of a growth round with `clusters` clusters over n vertices. */
uint32_t og_t_mover_threshold(int64_t clusters, int64_t n);
int og_t_below_0999(uint64_t cut, uint64_t best);
/* Roles and exact top-two proposals of one growth round (filter_roles = 1) or
 * the unrestricted favourites of two-hop pairing (filter_roles = 0). */
void og_t_proposals(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew, const uint32_t* vw,
                    const int32_t* clusters, const uint64_t* cw, uint64_t cap, uint32_t seed, int level, int round,
                    int filter_roles, uint32_t threshold, int need_regret, uint8_t* roles, int32_t* proposals,
                    uint64_t* gains, uint64_t* regret, uint32_t* ties);
/* Isolated packing on clusters/cw in place; returns 1 when it ran (>1 isolated vertex). */
int og_t_pack_isolated(int64_t n, const int32_t* off, const uint32_t* vw, uint64_t cap, int32_t* clusters,
                       uint64_t* cw);
/* Two-hop pairing on clusters/cw in place with a pair budget. */
void og_t_two_hop(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew, const uint32_t* vw,
                  int32_t* clusters, uint64_t* cw, uint64_t cap, uint32_t seed, int level, int64_t budget);
/* One LP proposal sweep; locks in/out; prices NULL = off; returns the candidate count. */
int64_t og_t_lp_propose(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew, const uint32_t* vw,
                        int k, const int32_t* labels, int32_t* locks, const uint64_t* loads, uint64_t cap,
                        const double* prices, double filter, int32_t* targets, int64_t* gains);
/* Afterburner over the candidates (targets >= 0); keep[v] = 1 for kept moves; returns their count. */
int64_t og_t_afterburner(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew, const uint32_t* vw,
                         int k, const int32_t* labels, const uint64_t* loads, uint64_t cap, const double* prices,
                         const int32_t* targets, const int64_t* gains, uint8_t* keep);
/* Repair gain bucket of a gain per unit of relief. */
int og_t_gain_bucket(double gain);
/* Repair destination of vertex v (-1 none), with A(v, dest) and A(v, own). */
int og_t_repair_destination(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew,
                            const uint32_t* vw, int k, const int32_t* labels, const uint64_t* loads, uint64_t cap,
                            int32_t v, uint64_t* best_conn, uint64_t* current);
/* One repair iteration without applying it: first-choice destinations,
 * donor buckets (-1 none), eligibility, and the admitted receiver (-1 not
 * admitted); returns the number admitted. */
int64_t og_t_repair_iteration(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew,
                              const uint32_t* vw, int k, const int32_t* labels, const uint64_t* loads, uint64_t cap,
                              int32_t* first_choice, int32_t* bucket, uint8_t* eligible, int32_t* receiver);

const char* og_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
