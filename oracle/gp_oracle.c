/*
 * gp_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the multilevel
 * k-way partitioner hot path: coarsening, contraction, projection and
 * refinement.  Plain C; every function cites the reference file:line it
 * restates (paths relative to the Acc-GP source tree).  Built with -fopenmp,
 * the loops whose iterations are independent (per-vertex proposals, seed
 * choices, coarse-row merges, LP proposals, afterburner, cut and affinity
 * sums) run on all threads: each iteration writes only its own outputs, the
 * order-carrying steps (compactions, bucket counts, commits) stay serial, and
 * sums are integer, so results do not depend on the thread count.
 *
 * Canonical orders (the reference's atomic arrival order is replaced by vertex
 * id order; the GPU library follows the same rule):
 *   - repair bucket rank: sum of weights of same-bucket donors with smaller id
 *     (reference: atomicAdd order, src/refinement/incremental.cuh:618);
 *   - receiver tickets: per attempt, requesters of one part in vertex order on
 *     top of the part's request total at the start of the attempt, and
 *     alternative destinations priced on that start-of-attempt total
 *     (reference: warp arrival order and a relaxed racy read,
 *     src/refinement/rebalance_graph.cuh:65-68,87-156);
 *   - gain bucket of a repair move: smallest j with 1.5^j >= -gain from an
 *     exactly-defined power table instead of ceil(log(-g)/log(1.5))
 *     (src/refinement/incremental.cuh:553-557).
 * Everything else in the reference is already order independent (integer
 * atomic sums, atomicMin/atomicMax, keyed sorts with a vertex tie-break).
 *
 * Compile with -ffp-contract=off: several decisions compare IEEE doubles.
 */
#include "gp_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OG_INVALID 0x7fffffff
#define OG_KMAX 128
#define OG_BALANCE_BUCKETS 50
#define OG_BALANCE_SECTIONS 16
#define OG_BALANCE_WIDTH (OG_BALANCE_BUCKETS * OG_BALANCE_SECTIONS)
#define OG_JOINT_SEEDS 128
#define OG_JOINT_SIDE 4
#define OG_JOINT_ARCHIVE 4

static char og_error_text[512];
const char* og_last_error(void) { return og_error_text; }
static int og_fail(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(og_error_text, sizeof og_error_text, fmt, ap);
    va_end(ap);
    return -1;
}

static void* og_alloc(size_t count, size_t size) {
    void* p = calloc(count ? count : 1, size);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

/* src/coarsen.cu:173-179, src/coarsening/seed_matching.cuh:7-10, src/refine.cu:50-56 */
#ifdef _OPENMP
static int og_threads(void) { return omp_get_max_threads(); }
static int og_thread(void) { return omp_get_thread_num(); }
#else
static int og_threads(void) { return 1; }
static int og_thread(void) { return 0; }
#endif

void og_set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

int og_get_threads(void) { return og_threads(); }

static inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU;
    return x ^ (x >> 16);
}
static inline uint64_t sat_add(uint64_t a, uint64_t b) { return UINT64_MAX - a < b ? UINT64_MAX : a + b; }
static inline uint64_t ew_at(const uint32_t* ew, int64_t e) { return ew ? ew[e] : 1u; }
static inline uint64_t u64max(uint64_t a, uint64_t b) { return a > b ? a : b; }
static inline uint64_t u64min(uint64_t a, uint64_t b) { return a < b ? a : b; }

/* ------------------------------------------------------------------ graph */
typedef struct {
    int64_t n, m;
    int32_t* off;
    int32_t* nbr;
    uint32_t* ew;      /* NULL: implicit unit weights (include/detail/edge_weight.hpp:8-11) */
    uint32_t* vw;
    uint64_t* row_bound; /* carried incident-weight bound, src/coarsen.cu:51-99,101-119 */
    uint64_t max_degree;
    int has_heavy;
    uint64_t incident_bound;
} og_graph;

static void og_graph_free(og_graph* g) {
    free(g->off); free(g->nbr); free(g->ew); free(g->vw); free(g->row_bound);
    memset(g, 0, sizeof *g);
}

struct og_hier {
    int levels;
    og_graph* g;      /* levels */
    int32_t** maps;   /* levels-1 */
};

typedef struct {
    int sorted_unique_rows, positive_ew, unit_ew;
    uint64_t total_weight, max_vertex_weight, max_row_weight;
} og_cstate;

void og_default_coarsen_opts(int parts, uint32_t seed, og_coarsen_opts* o) {
    /* include/coarsen.hpp:15-38 with the CLI overrides app/partition_main.cu:52-87 */
    memset(o, 0, sizeof *o);
    o->parts = parts; o->seed = seed; o->stop_ratio = 0.90; o->max_levels = 24;
    o->cluster_size_factor = parts <= 8 ? (64 > (512 + parts - 1) / parts ? 64 : (512 + parts - 1) / parts)
                                        : (2 > (32 + parts - 1) / parts ? 2 : (32 + parts - 1) / parts);
    o->pack_isolated = 1; o->progressive = parts >= 16; o->fine_factor = 64; o->capacity_growth = 8;
    o->seed_matching = 1; o->fast_path = 1; o->hem_degree_limit = 64;
}

void og_default_refine_opts(int parts, double imbalance, og_refine_opts* o) {
    /* include/refine.hpp:13-26, app/partition_main.cu:116-123 */
    memset(o, 0, sizeof *o);
    o->parts = parts; o->imbalance = imbalance; o->max_rounds = 32; o->stale_rounds = 3;
    o->growth_ratio = 2.0; o->exchange_rounds = 2; o->graph_loop_threshold = 65536;
    o->exchange_max_n = 65536;
}

/* include/partition_capacity.hpp:14-25 */
uint64_t og_capacity(uint64_t total, int parts, double imbalance) {
    if (parts < 2 || parts > OG_KMAX || !isfinite(imbalance) || imbalance < 1.0) return 0;
    long double raw = (long double)total * imbalance / parts;
    if (!isfinite((double)raw) || raw >= (long double)UINT64_MAX) return 0;
    uint64_t cap = (uint64_t)floorl(raw);
    if (cap < total / (unsigned)parts + (total % (unsigned)parts != 0)) return 0;
    return cap;
}

/* Level-0 metadata pass: src/coarsen.cu:51-99 (record_bound_kernel) and its
 * consumer 1806-1848. */
static void og_level0_metadata(og_graph* g, og_cstate* st) {
    int unsorted = 0, zero = 0, nonunit = 0;
    uint64_t maxrb = 0, maxdeg = 0;
    g->row_bound = og_alloc((size_t)g->n, sizeof(uint64_t));
    for (int64_t v = 0; v < g->n; ++v) {
        uint64_t d = 0;
        for (int64_t e = g->off[v]; e < g->off[v + 1]; ++e) {
            uint64_t w = ew_at(g->ew, e);
            d = sat_add(d, w);
            if (e > g->off[v] && g->nbr[e] <= g->nbr[e - 1]) unsorted = 1;
            if (!w) zero = 1;
            if (w != 1) nonunit = 1;
        }
        g->row_bound[v] = d;
        maxrb = u64max(maxrb, d);
        maxdeg = u64max(maxdeg, (uint64_t)(g->off[v + 1] - g->off[v]));
    }
    st->max_row_weight = maxrb;
    g->incident_bound = maxrb;
    st->sorted_unique_rows = !unsorted;
    st->positive_ew = !zero;
    st->unit_ew = !nonunit;
    g->max_degree = maxdeg;
    g->has_heavy = maxdeg >= 4096;
    uint64_t total = 0, maxv = 0;
    for (int64_t v = 0; v < g->n; ++v) { total += g->vw[v]; maxv = u64max(maxv, g->vw[v]); }
    st->total_weight = total;
    st->max_vertex_weight = maxv;
}

/* ------------------------------------------------------------- aggregate */
typedef struct { int32_t c; uint64_t w; } og_cw_pair;
static int cmp_cw_pair(const void* a, const void* b) {
    const og_cw_pair* x = a; const og_cw_pair* y = b;
    return x->c < y->c ? -1 : x->c > y->c;
}

/* Admission record: target asc, key desc, tie desc, vertex asc.
 * include/detail/capacity_admission.hpp:14-33, src/coarsen.cu:666-718. */
typedef struct { int32_t target; uint64_t key; uint32_t tie; int32_t v; uint64_t w, base; int64_t idx; } og_adm;
static int cmp_adm(const void* a, const void* b) {
    const og_adm* x = a; const og_adm* y = b;
    if (x->target != y->target) return x->target < y->target ? -1 : 1;
    if (x->key != y->key) return x->key > y->key ? -1 : 1;
    if (x->tie != y->tie) return x->tie > y->tie ? -1 : 1;
    return x->v < y->v ? -1 : x->v > y->v;
}

/* One admission batch (include/detail/capacity_admission.hpp:14-33,
 * src/coarsen.cu:1031-1075): requests sorted by cmp_adm; request i is
 * accepted iff the inclusive weight prefix of its target's requests up to i
 * is at most cap - base, base being the target's weight frozen at the start
 * of the batch.  A prefix, not bin packing: a rejected request still counts
 * toward the prefix of later ones.  accept[] is indexed by adm[].idx. */
static void og_admission(og_adm* adm, int64_t cnt, uint64_t cap, uint8_t* accept) {
    qsort(adm, (size_t)cnt, sizeof(og_adm), cmp_adm);
    uint64_t prefix = 0;
    for (int64_t i = 0; i < cnt; ++i) {
        if (i == 0 || adm[i].target != adm[i - 1].target) prefix = 0;
        prefix += adm[i].w;
        accept[adm[i].idx] = adm[i].base <= cap && prefix <= cap - adm[i].base;
    }
}

int og_admission_batch(int64_t cnt, const int32_t* targets, const uint64_t* keys, const uint32_t* ties,
                       const int32_t* vertices, const uint64_t* weights, const uint64_t* bases, uint64_t cap,
                       uint8_t* accepted) {
    og_adm* adm = og_alloc((size_t)(cnt > 0 ? cnt : 1), sizeof(og_adm));
    for (int64_t i = 0; i < cnt; ++i) {
        adm[i].target = targets[i]; adm[i].key = keys[i]; adm[i].tie = ties[i]; adm[i].v = vertices[i];
        adm[i].w = weights[i]; adm[i].base = bases[i]; adm[i].idx = i;
    }
    og_admission(adm, cnt, cap, accepted);
    free(adm);
    return 0;
}

typedef struct {
    const og_graph* g;
    int64_t n;
    int32_t* clusters;
    uint64_t* cw;
    uint8_t* roles;
    int32_t* proposals;
    uint64_t* gains;
    uint64_t* regret;
    uint32_t* ties;
    uint64_t cap;
    og_cw_pair* scratch;     /* scratch_stride entries per thread */
    int64_t scratch_stride;
} og_agg;

/* src/coarsening/aggregation.cuh:79-106 */
static inline int sclp_better(uint64_t a, uint64_t t, uint64_t oa, uint64_t ot) {
    return a > oa || (a == oa && a != 0 && t > ot);
}

/* One form_exact_proposals pass: src/coarsen.cu:739-970 with
 * src/coarsening/singleton_growth.cuh:4-40 and the decode in
 * src/coarsening/aggregation.cuh:513-557.  All degree-binned kernels compute
 * the same per-vertex exact top-two; this restates their common result. */
static void og_form_proposals(og_agg* A, uint32_t seed, int level, int round,
                              int filter_roles, uint32_t threshold, int need_regret) {
    const og_graph* g = A->g;
    const int64_t n = A->n;
    const uint32_t role_salt = seed ^ ((uint32_t)level * 0x9e3779b9U) ^ ((uint32_t)round * 0x85ebca6bU) ^ 0x27d4eb2dU;
    const uint32_t tie_salt = seed ^ ((uint32_t)level * 0xc2b2ae35U) ^ ((uint32_t)round * 0x165667b1U) ^ 0xd3a2646cU;
    const int use_second = filter_roles && need_regret;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n; ++v)
        A->roles[v] = A->clusters[v] == v && A->cw[v] == (uint64_t)g->vw[v] &&
                      (!filter_roles || mix32((uint32_t)v ^ role_salt) < threshold);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t v = 0; v < n; ++v) {
        og_cw_pair* scratch = A->scratch + (size_t)og_thread() * (size_t)A->scratch_stride;
        A->proposals[v] = OG_INVALID; A->gains[v] = 0; A->ties[v] = 0;
        const int32_t src = A->clusters[v];
        if (!A->roles[src]) continue;
        int cnt = 0;
        for (int64_t e = g->off[v]; e < g->off[v + 1]; ++e) {
            const int32_t u = g->nbr[e];
            if (u == v) continue;
            scratch[cnt].c = A->clusters[u];
            scratch[cnt].w = ew_at(g->ew, e);
            ++cnt;
        }
        qsort(scratch, (size_t)cnt, sizeof(og_cw_pair), cmp_cw_pair);
        uint64_t current = 0, best = 0, tie = 0, second = 0;
        for (int i = 0; i < cnt;) {
            const int32_t c = scratch[i].c;
            uint64_t val = 0;
            while (i < cnt && scratch[i].c == c) val += scratch[i++].w;
            if (c == src) { current = val; continue; }
            if (filter_roles) {
                const uint64_t base = A->cw[c];
                if (A->roles[c] || !(base != 0 && base <= A->cap && A->cw[src] <= A->cap - base)) continue;
            }
            const uint32_t h = mix32((uint32_t)v ^ mix32((uint32_t)c) ^ tie_salt);
            const uint64_t key = ((uint64_t)h << 32) | (uint64_t)(0xffffffffU - (uint32_t)c);
            if (use_second) {
                if (sclp_better(val, key, best, tie)) { second = u64max(second, best); best = val; tie = key; }
                else second = u64max(second, val);
            } else if (sclp_better(val, key, best, tie)) { best = val; tie = key; }
        }
        if (!(best > current)) continue;
        const int64_t target = (int64_t)(0xffffffffULL - (tie & 0xffffffffULL));
        if (target < 0 || target > INT32_MAX || target == src) continue;
        A->proposals[v] = (int32_t)target;
        A->gains[v] = best - current;
        if (use_second) A->regret[v] = best - second;
        A->ties[v] = (uint32_t)(tie >> 32);
    }
}

static int64_t og_count_clusters(const uint64_t* cw, int64_t n) {
    int64_t c = 0;
    for (int64_t v = 0; v < n; ++v) c += cw[v] != 0;
    return c;
}

/* seed_matching.cuh:12-91, call coarsen.cu:505-539.  Every vertex of degree at
 * most 1024 picks its admissible neighbour of largest (edge weight, pair hash
 * mix32(lo ^ mix32(hi ^ salt)), id); a neighbour is admissible when the two
 * vertex weights fit the cluster capacity together.  Mutual picks merge into
 * a cluster rooted at the smaller id.  Writes clusters and cluster weights
 * (cw must start zero); returns the number of clusters. */
static int64_t og_seed_match(og_agg* A, uint32_t salt) {
    const og_graph* g = A->g;
    const int64_t n = A->n;
    const uint64_t cap = A->cap;
    int32_t* choices = og_alloc((size_t)n, sizeof(int32_t));
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t v = 0; v < n; ++v) {
        choices[v] = -1;
        const int64_t deg = g->off[v + 1] - g->off[v];
        if (deg > 1024) continue;
        const uint64_t wv = g->vw[v];
        int32_t best = -1; uint64_t bw = 0; uint32_t bh = 0;
        for (int64_t e = g->off[v]; e < g->off[v + 1]; ++e) {
            const int32_t u = g->nbr[e];
            if (u == v || wv > cap || (uint64_t)g->vw[u] > cap - wv) continue;
            const uint32_t lo = (uint32_t)(u < v ? u : v), hi = (uint32_t)(u < v ? v : u);
            const uint32_t h = mix32(lo ^ mix32(hi ^ salt));
            const uint64_t w = ew_at(g->ew, e);
            if (best < 0 || w > bw || (w == bw && (h > bh || (h == bh && u > best)))) {
                best = u; bw = w; bh = h;
            }
        }
        choices[v] = best;
    }
    int32_t* matched = og_alloc((size_t)n, sizeof(int32_t));
    for (int64_t v = 0; v < n; ++v) matched[v] = -1;
    for (int64_t v = 0; v < n; ++v) {
        const int32_t u = choices[v];
        if (u > v && choices[u] == v) { matched[v] = u; matched[u] = (int32_t)v; }
    }
    int64_t count = 0;
    for (int64_t v = 0; v < n; ++v) {
        const int32_t p = matched[v];
        const int32_t root = p < 0 ? (int32_t)v : (p < v ? p : (int32_t)v);
        A->clusters[v] = root;
        count += root == v;
    }
    for (int64_t v = 0; v < n; ++v) A->cw[A->clusters[v]] += g->vw[v];
    free(choices); free(matched);
    return count;
}

/* Mover threshold of one growth round (coarsen.cu:984-990): a singleton root
 * moves when mix32(v ^ role_salt) is below floor(2^32 * min(1/2,
 * max(0, clusters - ceil(n/2)) / clusters)), taken in double. */
static uint32_t og_mover_threshold(int64_t cluster_count, int64_t n) {
    const int64_t desired = (n + 1) / 2;
    const int64_t removable = cluster_count - desired > 0 ? cluster_count - desired : 0;
    double frac = (double)removable / (double)cluster_count;
    if (frac > 0.50) frac = 0.50;
    return (uint32_t)(frac * 4294967296.0);
}

/* Isolated packing: coarsen.cu:1130-1161, aggregation.cuh:705-745.  Zero-degree
 * vertices in id order are laid on the weight line; bin b covers
 * [b*cap, (b+1)*cap) of the running weight prefix.  A vertex whose interval
 * [P-w, P) lies inside one bin joins the smallest such vertex of that bin;
 * a vertex straddling a bin edge stays alone.  Returns 1 when it ran. */
static int og_pack_isolated(og_agg* A) {
    const og_graph* g = A->g;
    const int64_t n = A->n;
    const uint64_t cap = A->cap;
    int64_t count = 0;
    for (int64_t v = 0; v < n; ++v) count += g->off[v] == g->off[v + 1];
    if (count <= 1) return 0;
    int32_t* iso = og_alloc((size_t)count, sizeof(int32_t));
    uint64_t* prefix = og_alloc((size_t)count, sizeof(uint64_t));
    int64_t c = 0; uint64_t run = 0;
    for (int64_t v = 0; v < n; ++v)
        if (g->off[v] == g->off[v + 1]) { iso[c] = (int32_t)v; run += g->vw[v]; prefix[c] = run; ++c; }
    const uint64_t tot = run;
    const uint64_t bins = tot / cap + (tot % cap != 0);
    int32_t* rep = og_alloc((size_t)bins, sizeof(int32_t));
    for (uint64_t b = 0; b < bins; ++b) rep[b] = INT32_MAX;
    for (int64_t i = 0; i < count; ++i) {
        const uint64_t w = g->vw[iso[i]];
        const uint64_t fb = (prefix[i] - w) / cap, lb = (prefix[i] - 1) / cap;
        if (fb == lb && iso[i] < rep[fb]) rep[fb] = iso[i];
    }
    for (int64_t i = 0; i < count; ++i) {
        const int32_t v = iso[i];
        const uint64_t w = g->vw[v];
        const uint64_t fb = (prefix[i] - w) / cap, lb = (prefix[i] - 1) / cap;
        if (fb != lb) continue;
        const int32_t t = rep[fb];
        if (t == v) continue;
        A->clusters[v] = t; A->cw[v] = 0; A->cw[t] += w;
    }
    free(iso); free(prefix); free(rep);
    return 1;
}

/* Two-hop pairing: coarsen.cu:1166-1259, aggregation.cuh:654-696.  Every
 * vertex still alone in its cluster takes its unrestricted favourite cluster
 * (the exact top-1 of og_form_proposals without roles or capacity); vertices
 * sharing a favourite are ordered by (gain desc, tie desc, id asc) and the
 * 2nd, 4th, ... of each group joins the cluster of the one before it when
 * their weights fit the capacity.  Only the first `budget` such pairs merge;
 * adm is scratch of n records. */
static void og_two_hop(og_agg* A, uint32_t seed, int level, int64_t budget, og_adm* adm) {
    const og_graph* g = A->g;
    const int64_t n = A->n;
    const uint64_t cap = A->cap;
    og_form_proposals(A, seed, level, 4, 0, 0, 0);
    int64_t cnt = 0;
    for (int64_t v = 0; v < n; ++v) {
        const int32_t fav = A->cw[A->clusters[v]] == (uint64_t)g->vw[v] ? A->proposals[v] : OG_INVALID;
        if (fav == OG_INVALID) continue;
        adm[cnt].target = fav; adm[cnt].key = A->gains[v]; adm[cnt].tie = A->ties[v];
        adm[cnt].v = (int32_t)v; ++cnt;
    }
    qsort(adm, (size_t)cnt, sizeof(og_adm), cmp_adm);
    int64_t position = 0, rank = 0;
    for (int64_t i = 0; i < cnt; ++i) {
        rank = (i == 0 || adm[i].target != adm[i - 1].target) ? 1 : rank + 1;
        const uint64_t wi = g->vw[adm[i].v];
        const int flag = (rank % 2 == 0) && i > 0 && (uint64_t)g->vw[adm[i - 1].v] <= cap - wi;
        if (!flag) continue;
        if (position < budget) {
            const int32_t v = adm[i].v;
            const int32_t s = A->clusters[v], t = A->clusters[adm[i - 1].v];
            if (s != t) { A->cw[s] -= wi; A->cw[t] += wi; A->clusters[v] = t; }
        }
        ++position;
    }
}

typedef struct {
    int64_t coarse_n;
    uint64_t capacity;
    uint64_t max_weight;
    uint64_t stats_max_degree;
    int stats_has_heavy;
} og_agg_result;

/* aggregate_typed: src/coarsen.cu:467-1380. */
static int og_aggregate(const og_graph* g, og_cstate* st, int parts, uint32_t seed, int level,
                        int level_factor, int pack_isolated, int seed_matching, int fast_path,
                        int32_t* map, uint32_t** coarse_vw, og_agg_result* res) {
    const int64_t n = g->n;
    const uint64_t total = st->total_weight, maxv = st->max_vertex_weight;
    const uint64_t denom = (uint64_t)level_factor * (uint64_t)parts;
    const uint64_t average_cap = total / denom + (total % denom != 0);
    const uint64_t cap = u64max(maxv, average_cap);                       /* coarsen.cu:494-498 */
    og_agg A;
    memset(&A, 0, sizeof A);
    A.g = g; A.n = n; A.cap = cap;
    A.clusters = og_alloc((size_t)n, sizeof(int32_t));
    A.cw = og_alloc((size_t)n, sizeof(uint64_t));
    A.roles = og_alloc((size_t)n, 1);
    A.proposals = og_alloc((size_t)n, sizeof(int32_t));
    A.gains = og_alloc((size_t)n, sizeof(uint64_t));
    A.regret = og_alloc((size_t)n, sizeof(uint64_t));
    A.ties = og_alloc((size_t)n, sizeof(uint32_t));
    int64_t maxdeg_row = 0;
    for (int64_t v = 0; v < n; ++v) {
        int64_t d = g->off[v + 1] - g->off[v];
        if (d > maxdeg_row) maxdeg_row = d;
    }
    A.scratch_stride = maxdeg_row + 1;
    A.scratch = og_alloc((size_t)A.scratch_stride * (size_t)og_threads(), sizeof(og_cw_pair));

    int64_t initial_count = n;
    if (seed_matching) initial_count = og_seed_match(&A, seed ^ ((uint32_t)level * 0x9e3779b9U));
    else
        for (int64_t v = 0; v < n; ++v) { A.clusters[v] = (int32_t)v; A.cw[v] = g->vw[v]; }

    /* Degree statistics: coarsen.cu:547-579.  A bound <= 8 is trusted. */
    uint64_t stats_maxdeg; int stats_heavy;
    if (g->max_degree <= 8) { stats_maxdeg = g->max_degree; stats_heavy = 0; }
    else {
        stats_maxdeg = (uint64_t)maxdeg_row;
        stats_heavy = maxdeg_row >= 4096;
    }

    int64_t cluster_count = initial_count;
    const int64_t desired = (n + 1) / 2;
    uint64_t bound = seed_matching ? (maxv > cap / 2 ? cap : 2 * maxv) : maxv;  /* coarsen.cu:513-515 */
    og_adm* adm = og_alloc((size_t)n, sizeof(og_adm));
    for (int round = 0; round < 4; ++round) {                                /* coarsen.cu:516-1128 */
        if ((double)cluster_count <= 0.60 * (double)n) break;
        const uint32_t threshold = og_mover_threshold(cluster_count, n);
        const uint64_t room = cap - bound;
        const uint64_t degree = stats_maxdeg;
        const int fit = fast_path && (degree == 0 || bound <= room / maxv / degree);
        const uint64_t next_bound = fit ? bound + bound * degree * maxv : cap;
        og_form_proposals(&A, seed, level, round, 1, threshold, !fit);
        if (fit) {
            /* singleton_growth.cuh:44-60 */
            for (int64_t v = 0; v < n; ++v) {
                const int32_t t = A.proposals[v];
                if (t == OG_INVALID) continue;
                A.clusters[v] = t; A.cw[v] = 0; A.cw[t] += g->vw[v];
            }
        } else {
            /* Admission: coarsen.cu:1031-1075, aggregation.cuh:561-652 */
            int64_t cnt = 0;
            for (int64_t v = 0; v < n; ++v) {
                if (A.proposals[v] == OG_INVALID) continue;
                adm[cnt].target = A.proposals[v]; adm[cnt].key = A.regret[v];
                adm[cnt].tie = A.ties[v]; adm[cnt].v = (int32_t)v;
                adm[cnt].w = g->vw[v]; adm[cnt].base = A.cw[A.proposals[v]]; adm[cnt].idx = cnt;
                ++cnt;
            }
            uint8_t* accept = og_alloc((size_t)(cnt > 0 ? cnt : 1), 1);
            og_admission(adm, cnt, cap, accept);
            for (int64_t i = 0; i < cnt; ++i) {
                if (!accept[adm[i].idx]) continue;
                const int32_t v = adm[i].v;
                const uint64_t w = g->vw[v];
                const int32_t s = A.clusters[v];
                A.cw[s] -= w; A.cw[adm[i].target] += w; A.clusters[v] = adm[i].target;
            }
            free(accept);
        }
        cluster_count = og_count_clusters(A.cw, n);
        bound = next_bound;
        if ((double)cluster_count <= 0.60 * (double)n) break;
    }

    if (pack_isolated && og_pack_isolated(&A)) cluster_count = og_count_clusters(A.cw, n);

    if ((double)cluster_count > 0.60 * (double)n) {
        og_two_hop(&A, seed, level, cluster_count - desired > 0 ? cluster_count - desired : 0, adm);
        cluster_count = og_count_clusters(A.cw, n);
    }

    /* Compaction: coarsen.cu:1261-1299, aggregation.cuh:698-703,747-752 */
    int32_t* ids = og_alloc((size_t)n, sizeof(int32_t));
    int64_t nc = 0;
    for (int64_t c = 0; c < n; ++c) { ids[c] = (int32_t)nc; nc += A.cw[c] > 0; }
    uint64_t* sums = og_alloc((size_t)nc, sizeof(uint64_t));
    for (int64_t v = 0; v < n; ++v) { map[v] = ids[A.clusters[v]]; sums[map[v]] += g->vw[v]; }
    uint32_t* cvw = og_alloc((size_t)nc, sizeof(uint32_t));
    uint64_t maxw = 0;
    for (int64_t c = 0; c < nc; ++c) {
        if (sums[c] > UINT32_MAX) return og_fail("coarse vertex weight exceeds uint32");
        cvw[c] = (uint32_t)sums[c];
        maxw = u64max(maxw, sums[c]);
    }
    free(ids); free(sums); free(adm);
    free(A.clusters); free(A.cw); free(A.roles); free(A.proposals); free(A.gains); free(A.regret);
    free(A.ties); free(A.scratch);
    if (maxw > cap) return og_fail("coarsening exceeded cluster capacity");
    *coarse_vw = cvw;
    res->coarse_n = nc; res->capacity = cap; res->max_weight = maxw;
    res->stats_max_degree = stats_maxdeg; res->stats_has_heavy = stats_heavy;
    return 0;
}

/* ------------------------------------------------------------- contract */
typedef struct { int32_t t; uint32_t w; } og_tw;
static int cmp_tw(const void* a, const void* b) {
    const og_tw* x = a; const og_tw* y = b;
    return x->t < y->t ? -1 : x->t > y->t;
}

/* Exact coarse CSR: sorted unique rows, summed weights, internal edges
 * removed (src/coarsen.cu:1394-1673, contraction_rows.cuh:18-154,
 * docs/GPU_ARCHITECTURE_ANALYSIS.md:867-873). */
static int og_contract_csr(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew,
                           const int32_t* map, int64_t nc, int32_t** out_off, int32_t** out_nbr,
                           uint32_t** out_ew, int64_t* out_m) {
    int64_t* raw = og_alloc((size_t)nc + 1, sizeof(int64_t));
    for (int64_t v = 0; v < n; ++v) raw[map[v] + 1] += off[v + 1] - off[v];
    for (int64_t c = 0; c < nc; ++c) raw[c + 1] += raw[c];
    og_tw* rec = og_alloc((size_t)raw[nc], sizeof(og_tw));
    int64_t* cur = og_alloc((size_t)nc, sizeof(int64_t));
    for (int64_t c = 0; c < nc; ++c) cur[c] = raw[c];
    for (int64_t v = 0; v < n; ++v) {
        const int32_t s = map[v];
        for (int64_t e = off[v]; e < off[v + 1]; ++e) {
            rec[cur[s]].t = map[nbr[e]] == s ? INT32_MAX : map[nbr[e]];
            rec[cur[s]].w = (uint32_t)ew_at(ew, e);
            ++cur[s];
        }
    }
    int32_t* coff = og_alloc((size_t)nc + 1, sizeof(int32_t));
    int64_t total = 0;
    int overflow = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total) reduction(| : overflow)
    for (int64_t c = 0; c < nc; ++c) {
        qsort(rec + raw[c], (size_t)(raw[c + 1] - raw[c]), sizeof(og_tw), cmp_tw);
        int64_t out = raw[c];
        for (int64_t i = raw[c]; i < raw[c + 1];) {
            const int32_t t = rec[i].t;
            uint64_t sum = 0;
            while (i < raw[c + 1] && rec[i].t == t) sum += rec[i++].w;
            if (t == INT32_MAX) continue;
            if (sum > UINT32_MAX) overflow = 1;
            rec[out].t = t; rec[out].w = (uint32_t)sum; ++out;
        }
        cur[c] = out - raw[c];
        total += cur[c];
    }
    if (overflow) { free(raw); free(rec); free(cur); free(coff); return og_fail("coarse edge weight exceeds uint32"); }
    if (total > INT32_MAX) return og_fail("coarse graph exceeds int32 offsets");
    int32_t* cn = og_alloc((size_t)total, sizeof(int32_t));
    uint32_t* cwt = og_alloc((size_t)total, sizeof(uint32_t));
    int64_t pos = 0;
    for (int64_t c = 0; c < nc; ++c) {
        coff[c] = (int32_t)pos;
        for (int64_t i = 0; i < cur[c]; ++i) { cn[pos] = rec[raw[c] + i].t; cwt[pos] = rec[raw[c] + i].w; ++pos; }
    }
    coff[nc] = (int32_t)pos;
    free(raw); free(rec); free(cur);
    *out_off = coff; *out_nbr = cn; *out_ew = cwt; *out_m = total;
    return 0;
}

/* contract_typed metadata: src/coarsen.cu:1458-1513,1619-1632,1664-1668 */
static int og_contract(const og_graph* fine, og_cstate* st, const int32_t* map, int64_t nc,
                       uint32_t* coarse_vw, og_graph* out) {
    const int64_t n = fine->n;
    const int count_raw = st->positive_ew && n >= 65536 && fine->max_degree <= 512;
    uint64_t* next_rb = og_alloc((size_t)nc, sizeof(uint64_t));
    uint64_t* rawc = og_alloc((size_t)nc, sizeof(uint64_t));
    for (int64_t v = 0; v < n; ++v) {
        next_rb[map[v]] = sat_add(next_rb[map[v]], fine->row_bound[v]);
        rawc[map[v]] += (uint64_t)(fine->off[v + 1] - fine->off[v]);
    }
    uint64_t crb = 0, craw = 0;
    for (int64_t c = 0; c < nc; ++c) { crb = u64max(crb, next_rb[c]); craw = u64max(craw, rawc[c]); }
    if (!count_raw) craw = UINT64_MAX;
    const uint64_t next_degree_bound = count_raw && fine->max_degree <= 128 ? craw : crb;
    free(rawc);
    memset(out, 0, sizeof *out);
    out->n = nc;
    if (og_contract_csr(n, fine->off, fine->nbr, fine->ew, map, nc, &out->off, &out->nbr, &out->ew, &out->m))
        return -1;
    out->vw = coarse_vw;
    out->row_bound = next_rb;
    out->incident_bound = crb;
    if (st->positive_ew) { out->max_degree = next_degree_bound; out->has_heavy = next_degree_bound >= 4096; }
    else { out->max_degree = UINT64_MAX; out->has_heavy = 1; }
    st->max_row_weight = crb;
    return 0;
}

/* ---------------------------------------------------------------- coarsen */
og_hier* og_coarsen(int64_t n, int64_t m, const int32_t* off, const int32_t* nbr,
                    const uint32_t* ew, const uint32_t* vw, const og_coarsen_opts* opts) {
    if (n <= 0 || n > INT32_MAX || m < 0 || m > INT32_MAX) { og_fail("unsupported graph size"); return NULL; }
    og_coarsen_opts o;
    if (opts) o = *opts; else og_default_coarsen_opts(2, 0, &o);
    const int k = o.parts;
    og_hier* h = og_alloc(1, sizeof(og_hier));
    h->g = og_alloc((size_t)o.max_levels + 1, sizeof(og_graph));
    h->maps = og_alloc((size_t)o.max_levels + 1, sizeof(int32_t*));
    og_graph cur;
    memset(&cur, 0, sizeof cur);
    cur.n = n; cur.m = m;
    cur.off = og_alloc((size_t)n + 1, sizeof(int32_t)); memcpy(cur.off, off, ((size_t)n + 1) * sizeof(int32_t));
    cur.nbr = og_alloc((size_t)m, sizeof(int32_t)); memcpy(cur.nbr, nbr, (size_t)m * sizeof(int32_t));
    int unit = 1;
    if (ew) for (int64_t e = 0; e < m; ++e) if (ew[e] != 1) { unit = 0; break; }
    if (!unit) { cur.ew = og_alloc((size_t)m, sizeof(uint32_t)); memcpy(cur.ew, ew, (size_t)m * sizeof(uint32_t)); }
    cur.vw = og_alloc((size_t)n, sizeof(uint32_t));
    for (int64_t v = 0; v < n; ++v) cur.vw[v] = vw ? vw[v] : 1u;
    og_cstate st;
    memset(&st, 0, sizeof st);
    og_level0_metadata(&cur, &st);
    /* coarsen.cu:1859-1879 */
    int64_t cutoff = (int64_t)o.cluster_size_factor * k;
    if (cutoff < 32) cutoff = 32;
    int level = 0;
    for (; level < o.max_levels && cur.n > cutoff; ++level) {
        const uint32_t level_seed = o.seed + (uint32_t)level;
        int level_factor = o.cluster_size_factor;
        if (o.progressive) {                                                  /* coarsen.cu:1887-1894 */
            const int64_t ff = o.fine_factor > (512 + k - 1) / k ? o.fine_factor : (512 + k - 1) / k;
            const int64_t den = (int64_t)o.capacity_growth * k;
            const int64_t remaining = (cur.n + den - 1) / den;
            const int64_t mn = ff < remaining ? ff : remaining;
            level_factor = (int)(level_factor > mn ? level_factor : mn);
        }
        const int try_matching = o.seed_matching && cur.n >= 65536 && !cur.has_heavy &&   /* 1895-1899 */
            (cur.m / cur.n < 8 || cur.max_degree <= 256) && cur.m / cur.n < o.hem_degree_limit &&
            st.sorted_unique_rows && st.positive_ew;
        int32_t* map = og_alloc((size_t)cur.n, sizeof(int32_t));
        uint32_t* cvw = NULL;
        og_agg_result r;
        memset(&r, 0, sizeof r);
        if (og_aggregate(&cur, &st, k, level_seed, level, level_factor, o.pack_isolated, try_matching,
                         o.fast_path, map, &cvw, &r)) {
            free(map); og_graph_free(&cur); og_hier_free(h); return NULL;
        }
        cur.max_degree = r.stats_max_degree;
        cur.has_heavy = r.stats_has_heavy;
        const double ratio = (double)r.coarse_n / (double)cur.n;
        if (r.coarse_n >= cur.n || ratio > o.stop_ratio) { free(map); free(cvw); break; }   /* 1925-1938 */
        og_graph coarse;
        if (og_contract(&cur, &st, map, r.coarse_n, cvw, &coarse)) {
            free(map); og_graph_free(&cur); og_hier_free(h); return NULL;
        }
        st.max_vertex_weight = r.max_weight;
        h->g[h->levels] = cur;
        h->maps[h->levels] = map;
        h->levels++;
        cur = coarse;
        st.sorted_unique_rows = 1;                                             /* 1997-1998 */
    }
    h->g[h->levels++] = cur;
    return h;
}

int og_hier_levels(const og_hier* h) { return h->levels; }
void og_hier_level_info(const og_hier* h, int level, int64_t* n, int64_t* m, uint64_t* max_degree,
                        int* has_heavy, uint64_t* incident_bound) {
    const og_graph* g = &h->g[level];
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (max_degree) *max_degree = g->max_degree;
    if (has_heavy) *has_heavy = g->has_heavy;
    if (incident_bound) *incident_bound = g->incident_bound;
}
void og_hier_level_copy(const og_hier* h, int level, int32_t* off, int32_t* nbr, uint32_t* ew, uint32_t* vw) {
    const og_graph* g = &h->g[level];
    if (off) memcpy(off, g->off, ((size_t)g->n + 1) * sizeof(int32_t));
    if (nbr) memcpy(nbr, g->nbr, (size_t)g->m * sizeof(int32_t));
    if (ew) for (int64_t e = 0; e < g->m; ++e) ew[e] = (uint32_t)ew_at(g->ew, e);
    if (vw) memcpy(vw, g->vw, (size_t)g->n * sizeof(uint32_t));
}
void og_hier_map_copy(const og_hier* h, int level, int32_t* map) {
    memcpy(map, h->maps[level], (size_t)h->g[level].n * sizeof(int32_t));
}
void og_hier_free(og_hier* h) {
    if (!h) return;
    for (int i = 0; i < h->levels; ++i) og_graph_free(&h->g[i]);
    for (int i = 0; i + 1 < h->levels; ++i) free(h->maps[i]);
    free(h->g); free(h->maps); free(h);
}

int og_contract_graph(int64_t n, int64_t m, const int32_t* off, const int32_t* nbr, const uint32_t* ew,
                      const int32_t* map, int64_t nc, int32_t* coff, int32_t* cnbr, uint32_t* cew, int64_t* cm) {
    (void)m;
    int32_t *o = NULL, *nb = NULL; uint32_t* w = NULL; int64_t mm = 0;
    if (og_contract_csr(n, off, nbr, ew, map, nc, &o, &nb, &w, &mm)) return -1;
    if (coff) memcpy(coff, o, ((size_t)nc + 1) * sizeof(int32_t));
    if (cnbr) memcpy(cnbr, nb, (size_t)mm * sizeof(int32_t));
    if (cew) memcpy(cew, w, (size_t)mm * sizeof(uint32_t));
    *cm = mm;
    free(o); free(nb); free(w);
    return 0;
}

/* ============================================================ refinement */
uint64_t og_cut(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew, const int32_t* labels) {
    uint64_t d = 0;                                                        /* refine.cu:63-98 */
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : d)
    for (int64_t v = 0; v < n; ++v)
        for (int64_t e = off[v]; e < off[v + 1]; ++e)
            if (labels[v] != labels[nbr[e]]) d += ew_at(ew, e);
    return d / 2;
}

typedef struct { int64_t loss; uint64_t weight; } og_sample;

typedef struct {
    const og_graph* g;
    int64_t n;
    int k;
    uint64_t cap;
    int32_t* labels;
    uint64_t* loads;
    int32_t* targets;
    int64_t* gains;
    int32_t* locks;
    int32_t* flags;
    uint64_t* conn;      /* k scratch per thread */
    double* prices;      /* NULL when disabled */
    og_sample* samples;
    double pow15[48];    /* 1.5^j, j=-23..24, index j+23 */
    uint64_t cut;
    og_level_stats* st;
} og_ref;

/* A(v,.) and the row total excluding self loops (src/refinement/affinity.cuh:9-66). */
static uint64_t og_affinity(const og_ref* R, int64_t v, uint64_t* conn) {
    const og_graph* g = R->g;
    uint64_t total = 0;
    memset(conn, 0, (size_t)R->k * sizeof(uint64_t));
    for (int64_t e = g->off[v]; e < g->off[v + 1]; ++e) {
        const int32_t u = g->nbr[e];
        if (u == v) continue;
        const uint64_t w = ew_at(g->ew, e);
        conn[R->labels[u]] += w;
        total += w;
    }
    return total;
}

/* incremental.cuh:156-159 */
static double og_repair_cost(const og_ref* R, int part, uint64_t weight) {
    const uint64_t slack = R->cap - R->loads[part];
    return weight > slack ? R->prices[part] * (double)(weight - slack) : 0.0;
}

static uint64_t og_max_load(const og_ref* R) {
    uint64_t m = 0;
    for (int p = 0; p < R->k; ++p) m = u64max(m, R->loads[p]);
    return m;
}

/* Exact incremental cut update + commit: incremental.cuh:440-547.  Returns the
 * signed cut reduction; moves are the vertices with flags set. */
static int64_t og_apply(og_ref* R, const int32_t* movers, int64_t count, int lp, int with_samples) {
    const og_graph* g = R->g;
    uint64_t delta = 0;
    const int sampled = with_samples && R->samples;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : delta) if (!sampled)
    for (int64_t i = 0; i < count; ++i) {
        const int32_t v = movers[i];
        const int32_t s = R->labels[v], t = R->targets[v];
        uint64_t gain = 0;
        for (int64_t e = g->off[v]; e < g->off[v + 1]; ++e) {
            const int32_t u = g->nbr[e];
            if (u == v) continue;
            const int32_t old_u = R->labels[u];
            const int moves = R->flags[u] != 0;
            const int32_t new_u = moves ? R->targets[u] : old_u;
            if (!moves || u > v) {
                const int64_t d = (int64_t)(s != old_u) - (int64_t)(t != new_u);
                gain += (uint64_t)((int64_t)ew_at(g->ew, e) * d);
            }
        }
        delta += gain;
        if (sampled) {
            R->samples[s].loss += -(int64_t)gain;
            R->samples[s].weight += g->vw[v];
        }
    }
    for (int64_t i = 0; i < count; ++i) {
        const int32_t v = movers[i];
        const int32_t s = R->labels[v], t = R->targets[v];
        R->loads[s] -= g->vw[v];
        R->loads[t] += g->vw[v];
        R->labels[v] = t;
        if (lp) R->locks[v] = 1;
    }
    return (int64_t)delta;
}

static int og_update_cut(og_ref* R, int64_t delta) {
    if (delta >= 0) {
        if ((uint64_t)delta > R->cut) return og_fail("incremental cut underflow");
        R->cut -= (uint64_t)delta;
    } else R->cut += (uint64_t)(-delta);
    return 0;
}

/* LP proposals: incremental.cuh:166-198,229-296.  A vertex locked by the
 * previous round is skipped (and unlocked); an interior vertex (all incident
 * weight in its own part) is skipped.  The target is the part p != s of
 * largest A(v,p) > 0, the lowest id on ties; with prices on, the largest
 * score A(v,p) - pi_p * max(0, w - slack_p), the first on ties.  The vertex
 * proposes when best >= A(v,s) or A(v,s) - best < floor(filter * A(v,s)).
 * Candidates are listed in vertex order in movers; returns their count. */
static int64_t og_lp_propose(og_ref* R, double filter, int32_t* movers) {
    const og_graph* g = R->g;
    const int64_t n = R->n;
    const int k = R->k;
    int64_t ncand = 0;
    /* Per-vertex choices in parallel (flags marks the candidates), then the
     * candidate list in vertex order. */
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t v = 0; v < n; ++v) {
        uint64_t* conn = R->conn + (size_t)og_thread() * (size_t)k;
        R->targets[v] = -1; R->gains[v] = INT64_MIN; R->flags[v] = 0;
        if (R->locks[v]) { R->locks[v] = 0; continue; }
        const int s = R->labels[v];
        const uint64_t total = og_affinity(R, v, conn);
        const uint64_t current = conn[s];
        if (current == total) continue;
        const uint64_t w = g->vw[v];
        uint64_t best = 0; int target = -1; double best_score = -1.7976931348623157e308;
        for (int p = 0; p < k; ++p) {
            const uint64_t c = conn[p];
            if (p == s || !c) continue;
            if (R->prices) {
                const double score = (double)c - og_repair_cost(R, p, w);
                if (target < 0 || score > best_score) { best_score = score; best = c; target = p; }
            } else if (target < 0 || c > best) { best = c; target = p; }
        }
        if (target >= 0 && (best >= current || current - best < (uint64_t)(filter * (double)current))) {
            R->targets[v] = target;
            R->gains[v] = best >= current ? (int64_t)(best - current) : -(int64_t)(current - best);
            R->flags[v] = 1;
        }
    }
    for (int64_t v = 0; v < n; ++v)
        if (R->flags[v]) { R->flags[v] = 0; movers[ncand++] = (int32_t)v; }
    return ncand;
}

/* Afterburner: incremental.cuh:361-438.  Candidate v keeps its move when
 * g_v + sum over candidate neighbours u ranked before v (g_u > g_v, or equal
 * and u < v) of w_vu([t_u=t_v] - [t_u=s_v] - [p(u)=t_v] + [p(u)=s_v]) is >= 0,
 * plus the price difference of leaving s_v for t_v when prices are on: the
 * exact cut change of moving v after every candidate ranked before it.  Sets
 * flags, compacts movers to the kept candidates and returns their count. */
static int64_t og_afterburner(og_ref* R, int32_t* movers, int64_t ncand) {
    const og_graph* g = R->g;
    int64_t count = 0;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < ncand; ++i) {
        const int32_t v = movers[i];
        const int s = R->labels[v], t = R->targets[v];
        const int64_t gv = R->gains[v];
        int64_t corr = 0;
        for (int64_t e = g->off[v]; e < g->off[v + 1]; ++e) {
            const int32_t u = g->nbr[e];
            if (u == v || R->targets[u] < 0 || !(R->gains[u] > gv || (R->gains[u] == gv && u < v))) continue;
            corr += (int64_t)ew_at(g->ew, e) *
                    ((R->targets[u] == t) - (R->targets[u] == s) - (R->labels[u] == t) + (R->labels[u] == s));
        }
        int keep;
        if (R->prices) {
            const uint64_t w = g->vw[v];
            const double adj = og_repair_cost(R, s, w) - og_repair_cost(R, t, w);
            keep = (double)(gv + corr) + adj >= 0;
        } else keep = gv + corr >= 0;
        R->flags[v] = keep;
    }
    for (int64_t i = 0; i < ncand; ++i) if (R->flags[movers[i]]) movers[count++] = movers[i];
    return count;
}

/* 1.5^j for j = -23..24 at index j + 23: exact products upward from 1, and
 * quotients by 1.5 downward (both directions in IEEE double). */
static void og_pow15_table(double* t) {
    t[23] = 1.0;
    for (int j = 1; j <= 24; ++j) t[23 + j] = t[22 + j] * 1.5;
    for (int j = 1; j <= 23; ++j) t[23 - j] = t[24 - j] / 1.5;
}

/* balance_gain_bucket (incremental.cuh:553-557): 0 for a positive gain per
 * unit of relief, 1 for zero, else max(2, min(49, 25 + ceil(log(-g)/log 1.5))),
 * taken as the smallest j with -g <= 1.5^j on the table above instead of the
 * libm logarithms (same bucket wherever log is correctly rounded). */
static int og_gain_bucket_table(const double* pow15, double gain) {
    if (gain > 0) return 0;
    if (gain == 0) return 1;
    const double x = -gain;
    for (int j = -23; j <= 23; ++j)
        if (x <= pow15[j + 23]) return 25 + j < 2 ? 2 : 25 + j;
    return 49;
}
static int og_gain_bucket(const og_ref* R, double gain) { return og_gain_bucket_table(R->pow15, gain); }

/* enumerate_repair_destination: rebalance_select.cuh:8-75 (with the dense
 * ascending scan and the slack-weighted hash fallback). */
static int og_repair_destination(const og_ref* R, int64_t v, int source, uint64_t weight,
                                 const uint64_t* conn, uint64_t total, uint64_t* best_conn, uint64_t* current) {
    const int k = R->k;
    *current = conn[source];
    *best_conn = 0;
    int best = -1; uint64_t best_slack = 0;
    if (*current != total) {
        for (int p = 0; p < k; ++p) {
            if (p == source || R->loads[p] >= R->cap || weight > R->cap - R->loads[p]) continue;
            const uint64_t value = conn[p], slack = R->cap - R->loads[p];
            if (!value) continue;
            if (best < 0 || value > *best_conn || (value == *best_conn && slack > best_slack)) {
                best = p; *best_conn = value; best_slack = slack;
            }
        }
    }
    if (best >= 0) return best;
    *best_conn = 0;
    uint64_t tot = 0;
    for (int p = 0; p < k; ++p)
        if (p != source && R->loads[p] < R->cap && weight <= R->cap - R->loads[p])
            tot = sat_add(tot, R->cap - R->loads[p]);
    if (!tot) return -1;
    uint64_t ticket = (uint64_t)mix32((uint32_t)v) % tot;
    for (int p = 0; p < k; ++p)
        if (p != source && R->loads[p] < R->cap && weight <= R->cap - R->loads[p]) {
            const uint64_t slack = R->cap - R->loads[p];
            if (ticket < slack) return p;
            ticket -= slack;
        }
    return -1;
}

/* Scratch of one capacity-repair iteration. */
typedef struct {
    int32_t* bucket;     /* n: donor bucket index (part*800 + gain bucket*16 + section), -1 none */
    uint64_t* counts;    /* k*800 bucket weights */
    uint64_t* prefix;    /* k*800 exclusive prefix of counts within each part */
    uint64_t* run;       /* k*800 running same-bucket weight (canonical rank) */
    uint64_t* loads0;    /* k: loads frozen at the start of the iteration */
    uint64_t* slack;     /* k: cap - loads0 (0 when overloaded) */
    uint64_t* requests;  /* k: admitted request weight per receiver */
    uint64_t* snapshot;  /* k */
    uint64_t* batch;     /* k */
    uint8_t* eligible;   /* n */
    uint8_t* done;       /* n */
    uint8_t* tried;      /* n*k */
    int32_t* choice;     /* n */
    int32_t* elist;      /* n: eligible donors in vertex order */
} og_repair_ws;

static void og_repair_ws_alloc(og_repair_ws* W, int64_t n, int k) {
    W->bucket = og_alloc((size_t)n, sizeof(int32_t));
    W->counts = og_alloc((size_t)k * OG_BALANCE_WIDTH, sizeof(uint64_t));
    W->prefix = og_alloc((size_t)k * OG_BALANCE_WIDTH, sizeof(uint64_t));
    W->run = og_alloc((size_t)k * OG_BALANCE_WIDTH, sizeof(uint64_t));
    W->loads0 = og_alloc((size_t)k, sizeof(uint64_t));
    W->slack = og_alloc((size_t)k, sizeof(uint64_t));
    W->requests = og_alloc((size_t)k, sizeof(uint64_t));
    W->snapshot = og_alloc((size_t)k, sizeof(uint64_t));
    W->batch = og_alloc((size_t)k, sizeof(uint64_t));
    W->eligible = og_alloc((size_t)n, 1);
    W->done = og_alloc((size_t)n, 1);
    W->tried = og_alloc((size_t)n * (size_t)k, 1);
    W->choice = og_alloc((size_t)n, sizeof(int32_t));
    W->elist = og_alloc((size_t)n, sizeof(int32_t));
}

static void og_repair_ws_free(og_repair_ws* W) {
    free(W->bucket); free(W->counts); free(W->prefix); free(W->run); free(W->loads0); free(W->slack);
    free(W->requests); free(W->snapshot); free(W->batch); free(W->eligible); free(W->done); free(W->tried);
    free(W->choice); free(W->elist);
}

/* Donor proposals and eligibility of one repair iteration on the loads frozen
 * at its start: incremental.cuh:559-652.  Every vertex of an overloaded part
 * takes its repair destination (og_repair_destination); its gain per unit of
 * relief min(w, excess) picks one of 50 gain buckets and mix32(v) % 16 one of
 * 16 sections.  Buckets are ordered within the part; a donor is eligible when
 * the weight of the part's donors before it -- earlier buckets, then smaller
 * ids in its own bucket (the reference ranks by atomic arrival instead,
 * incremental.cuh:618) -- is below the part's excess.  Sets targets to the
 * first-choice destination and returns the eligible donors in vertex order
 * in W->elist. */
static int64_t og_repair_select(og_ref* R, og_repair_ws* W) {
    const og_graph* g = R->g;
    const int64_t n = R->n;
    const int k = R->k;
    const uint64_t cap = R->cap;
    memcpy(W->loads0, R->loads, (size_t)k * sizeof(uint64_t));
    memset(W->counts, 0, (size_t)k * OG_BALANCE_WIDTH * sizeof(uint64_t));
    for (int p = 0; p < k; ++p) { W->slack[p] = W->loads0[p] < cap ? cap - W->loads0[p] : 0; W->requests[p] = 0; }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t v = 0; v < n; ++v) {
        uint64_t* conn = R->conn + (size_t)og_thread() * (size_t)k;
        W->bucket[v] = -1; R->targets[v] = -1; R->flags[v] = 0;
        const int s = R->labels[v];
        if (!(W->loads0[s] > cap)) continue;
        const uint64_t w = g->vw[v];
        const uint64_t total = og_affinity(R, v, conn);
        uint64_t best_conn, current;
        const int best = og_repair_destination(R, v, s, w, conn, total, &best_conn, &current);
        if (best < 0) continue;
        const double gain = best_conn >= current ? (double)(best_conn - current) : -(double)(current - best_conn);
        const uint64_t excess = W->loads0[s] - cap;
        const uint64_t relief = w < excess ? w : excess;
        const double gpw = gain / (double)relief;
        R->targets[v] = best;
        W->bucket[v] = s * OG_BALANCE_WIDTH + og_gain_bucket(R, gpw) * OG_BALANCE_SECTIONS +
                       (int)(mix32((uint32_t)v) % OG_BALANCE_SECTIONS);
    }
    for (int64_t v = 0; v < n; ++v)
        if (W->bucket[v] >= 0) W->counts[W->bucket[v]] += g->vw[v];
    for (int p = 0; p < k; ++p) {
        uint64_t sum = 0;
        for (int b = 0; b < OG_BALANCE_WIDTH; ++b) {
            W->prefix[p * OG_BALANCE_WIDTH + b] = sum;
            sum += W->counts[p * OG_BALANCE_WIDTH + b];
        }
    }
    memset(W->run, 0, (size_t)k * OG_BALANCE_WIDTH * sizeof(uint64_t));
    int64_t ecount = 0;
    for (int64_t v = 0; v < n; ++v) {
        W->eligible[v] = 0; W->done[v] = 0;
        const int b = W->bucket[v];
        if (b < 0) continue;
        const int s = b / OG_BALANCE_WIDTH;
        const uint64_t rank = W->run[b];
        W->run[b] += g->vw[v];
        if (W->prefix[b] + rank < W->loads0[s] - cap) { W->eligible[v] = 1; W->elist[ecount++] = (int32_t)v; }
    }
    return ecount;
}

/* Receiver reservation: rebalance_graph.cuh:86-156, at most k attempts.  In
 * attempt 0 every eligible donor requests its first-choice destination; in
 * later attempts a donor not yet admitted requests the untried part p != s of
 * largest (A(v,p), room) whose room -- slack minus the requests admitted
 * before the attempt -- holds it.  Per attempt, requests to one receiver are
 * ranked in vertex order on top of its admitted total; a request is admitted
 * when it ends within the receiver's slack.  Admitted donors get flags = 1
 * and their receiver in targets; an attempt without any request ends it. */
static void og_repair_reserve(og_ref* R, og_repair_ws* W, int64_t ecount) {
    const og_graph* g = R->g;
    const int k = R->k;
    for (int64_t i = 0; i < ecount; ++i) memset(W->tried + (size_t)W->elist[i] * k, 0, (size_t)k);
    for (int attempt = 0; attempt < k; ++attempt) {
        memcpy(W->snapshot, W->requests, (size_t)k * sizeof(uint64_t));
        int any = 0;
        for (int64_t i = 0; i < ecount; ++i) {
            const int32_t v = W->elist[i];
            W->choice[v] = -1;
            if (W->done[v]) continue;
            if (attempt == 0) W->choice[v] = R->targets[v];
            else {
                const uint64_t w = g->vw[v];
                og_affinity(R, v, R->conn);
                int t = -1; uint64_t best = 0, room = 0;
                for (int p = 0; p < k; ++p) {
                    if (p == R->labels[v] || W->tried[(size_t)v * k + p]) continue;
                    const uint64_t avail = W->snapshot[p] < W->slack[p] ? W->slack[p] - W->snapshot[p] : 0;
                    if (avail < w) continue;
                    if (t < 0 || R->conn[p] > best || (R->conn[p] == best && avail > room)) {
                        t = p; best = R->conn[p]; room = avail;
                    }
                }
                W->choice[v] = t;
            }
            if (W->choice[v] >= 0) any = 1;
        }
        if (!any) break;
        memset(W->batch, 0, (size_t)k * sizeof(uint64_t));
        for (int64_t i = 0; i < ecount; ++i) {
            const int32_t v = W->elist[i];
            const int t = W->choice[v];
            if (t < 0) continue;
            const uint64_t w = g->vw[v];
            W->tried[(size_t)v * k + t] = 1;
            const uint64_t begin = W->requests[t] + W->batch[t];
            W->batch[t] += w;
            if (begin <= W->slack[t] && w <= W->slack[t] - begin) { R->targets[v] = t; R->flags[v] = 1; W->done[v] = 1; }
        }
        for (int p = 0; p < k; ++p) W->requests[p] += W->batch[p];
    }
}

/* One capacity-repair loop: rebalance_graph.cuh:160-280 (body 215-262, loop
 * control 48-63): select, reserve, apply; repeated while infeasible, moves
 * were made and fewer than 2k+4 iterations ran.  Returns feasibility. */
static int og_rebalance(og_ref* R, int* ok) {
    const int64_t n = R->n;
    const int k = R->k;
    const uint64_t cap = R->cap;
    *ok = 0;
    og_repair_ws W;
    og_repair_ws_alloc(&W, n, k);
    int32_t* movers = og_alloc((size_t)n, sizeof(int32_t));
    int iterations = 0, feasible = 0;
    int64_t count = 0;
    do {
        const int64_t ecount = og_repair_select(R, &W);
        og_repair_reserve(R, &W, ecount);
        count = 0;
        for (int64_t v = 0; v < n; ++v) {
            if (R->flags[v]) movers[count++] = (int32_t)v;
            else R->targets[v] = -1;
        }
        ++iterations;
        if (count) {
            const int64_t delta = og_apply(R, movers, count, 0, 1);
            if (og_update_cut(R, delta)) goto fail;
            R->st->repair_moves += (uint64_t)count;
        }
        feasible = og_max_load(R) <= cap;
    } while (!feasible && count && iterations < 2 * k + 4);
    R->st->repair_iterations += iterations;
    *ok = 1;
fail:
    og_repair_ws_free(&W);
    free(movers);
    return feasible;
}

/* ------------------------------------------------------------- exchange */
typedef struct {
    int32_t vertices[2 * OG_JOINT_SIDE];
    int32_t size, a, b, id;
    int64_t gain, net;
} og_move;

static void og_move_clear(og_move* x) {
    for (int j = 0; j < 2 * OG_JOINT_SIDE; ++j) x->vertices[j] = -1;
    x->size = 0; x->a = -1; x->b = -1; x->id = 0; x->gain = 0; x->net = 0;
}

/* exchange.cuh:19-40 */
static int og_same_move(const og_move* x, const og_move* y) {
    if (!x->size || x->size != y->size || x->a != y->a || x->b != y->b || x->gain != y->gain || x->net != y->net)
        return 0;
    for (int j = 0; j < x->size; ++j) {
        int found = 0;
        for (int t = 0; t < y->size; ++t) found |= x->vertices[j] == y->vertices[t];
        if (!found) return 0;
    }
    return 1;
}
static int cmp_joint_order(const void* pa, const void* pb) {
    const og_move* x = pa; const og_move* y = pb;
    if ((x->size != 0) != (y->size != 0)) return x->size != 0 ? -1 : 1;
    if (x->a != y->a) return x->a < y->a ? -1 : 1;
    if (x->gain != y->gain) return x->gain > y->gain ? -1 : 1;
    return x->id < y->id ? -1 : x->id > y->id;
}
/* complementary_moves.cuh:5-14 */
static int cmp_joint_net(const void* pa, const void* pb) {
    const og_move* x = pa; const og_move* y = pb;
    if (x->a != y->a) return x->a < y->a ? -1 : 1;
    if (x->net != y->net) return x->net < y->net ? -1 : 1;
    if (x->gain != y->gain) return x->gain > y->gain ? -1 : 1;
    if (x->size != y->size) return x->size < y->size ? -1 : 1;
    for (int j = 0; j < x->size; ++j)
        if (x->vertices[j] != y->vertices[j]) return x->vertices[j] < y->vertices[j] ? -1 : 1;
    return x->id < y->id ? -1 : x->id > y->id;
}
static inline int64_t og_gap(int64_t net, int64_t lo, int64_t hi) {
    return net < lo ? lo - net : net > hi ? net - hi : 0;
}
/* complementary_moves.cuh:20-28 */
static int og_net_bound(const og_move* r, int count, int a, int64_t net, int upper) {
    int lo = 0, hi = count;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        const og_move* x = &r[mid];
        if (x->a < a || (x->a == a && (x->net < net || (upper && x->net == net)))) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* joint_grow: exchange.cuh:96-204 */
static void og_joint_grow(og_ref* R, const uint64_t* Adense, int slot, int sw, const uint64_t* seeds,
                          const int32_t* partner, og_move* moves, og_move* archive) {
    const og_graph* g = R->g;
    const int k = R->k;
    const int seed_slot = slot / 2, policy = slot % 2;
    og_move best; og_move_clear(&best); best.id = slot;
    moves[slot] = best;
    for (int j = 0; j < OG_JOINT_ARCHIVE; ++j) og_move_clear(&archive[slot * OG_JOINT_ARCHIVE + j]);
    const uint64_t enc = seeds[seed_slot];
    if (enc == UINT64_MAX) return;
    const int32_t root = (int32_t)(uint32_t)enc;
    const int source = R->labels[root], target = partner[source];
    if (target < 0) return;
    int32_t members[8] = {root, -1, -1, -1, -1, -1, -1, -1};
    int size = 1, counts[2] = {1, 0};
    int64_t net = g->vw[root];
    const int64_t lo = (int64_t)R->loads[source] - (int64_t)R->cap;
    const int64_t hi = (int64_t)R->cap - (int64_t)R->loads[target];
    int64_t gain = (int64_t)Adense[(size_t)root * k + target] - (int64_t)Adense[(size_t)root * k + source];
    for (int step = 0; step < 8; ++step) {
        if (gain > 0 && size <= OG_JOINT_ARCHIVE) {
            og_move s; og_move_clear(&s);
            s.size = size; s.a = source < target ? source : target; s.b = source < target ? target : source;
            s.gain = gain; s.net = source == s.a ? net : -net; s.id = slot * OG_JOINT_ARCHIVE + step;
            for (int j = 0; j < size; ++j) {
                int pos = j;
                while (pos && s.vertices[pos - 1] > members[j]) { s.vertices[pos] = s.vertices[pos - 1]; --pos; }
                s.vertices[pos] = members[j];
            }
            archive[slot * OG_JOINT_ARCHIVE + step] = s;
        }
        if (gain > best.gain && !og_gap(net, lo, hi)) {
            best.size = size; best.a = source < target ? source : target; best.b = source < target ? target : source;
            best.gain = gain; best.net = source == best.a ? net : -net;
            for (int j = 0; j < size; ++j) best.vertices[j] = members[j];
        }
        if (size == 8) break;
        int64_t cg = LLONG_MIN, cgap = LLONG_MAX;
        int32_t chosen = INT32_MAX;
        const int gap_first = policy == 0 && og_gap(net, lo, hi) != 0;
        for (int origin = -1; origin < size; ++origin) {
            const int rev = origin < 0;
            if (rev && counts[1]) continue;
            const int32_t mm = rev ? -1 : members[origin];
            const int64_t begin = rev ? 0 : g->off[mm];
            int64_t end = rev ? sw : g->off[mm + 1];
            if (!rev && end > begin + 128) end = begin + 128;
            for (int64_t e = begin; e < end; ++e) {
                int32_t u;
                if (rev) {
                    const uint64_t val = seeds[(size_t)target * sw + e];
                    if (val == UINT64_MAX) continue;
                    u = (int32_t)(uint32_t)val;
                } else u = g->nbr[e];
                const int p = R->labels[u];
                if (p != source && p != target) continue;
                const int side = p == source ? 0 : 1;
                if (counts[side] >= OG_JOINT_SIDE) continue;
                if (!rev && p != R->labels[mm]) continue;
                int dup = 0;
                for (int j = 0; j < size; ++j) dup |= u == members[j];
                if (dup || g->off[u + 1] - g->off[u] > 1024) continue;
                const int64_t cnet = net + (side ? -(int64_t)g->vw[u] : (int64_t)g->vw[u]);
                const int64_t gp = og_gap(cnet, lo, hi);
                if (net > hi && !side) continue;
                if (net < lo && side) continue;
                int64_t marginal = (int64_t)Adense[(size_t)u * k + (side ? source : target)] -
                                   (int64_t)Adense[(size_t)u * k + p];
                if (!rev) marginal += 2 * (int64_t)ew_at(g->ew, e);
                const int better = gap_first
                    ? (gp < cgap || (gp == cgap && (marginal > cg || (marginal == cg && u < chosen))))
                    : (marginal > cg || (marginal == cg && (gp < cgap || (gp == cgap && u < chosen))));
                if (better) { chosen = u; cg = marginal; cgap = gp; }
            }
        }
        if (chosen == INT32_MAX) break;
        const int side = R->labels[chosen] == source ? 0 : 1;
        int64_t corr = 0;
        for (int64_t e = g->off[chosen]; e < g->off[chosen + 1]; ++e)
            for (int j = 0; j < size; ++j)
                if (g->nbr[e] == members[j])
                    corr += (R->labels[members[j]] == R->labels[chosen] ? 2 : -2) * (int64_t)ew_at(g->ew, e);
        const int64_t cgain = (int64_t)Adense[(size_t)chosen * k + (side ? source : target)] -
                              (int64_t)Adense[(size_t)chosen * k + R->labels[chosen]] + corr;
        members[size++] = chosen; ++counts[side]; gain += cgain;
        net += side ? -(int64_t)g->vw[chosen] : (int64_t)g->vw[chosen];
    }
    moves[slot] = best;
}

/* joint_combine: complementary_moves.cuh:31-75 */
static void og_joint_combine(og_ref* R, int count, const og_move* archive, int id_base, og_move* out, int i) {
    const og_graph* g = R->g;
    const og_move x = archive[i];
    const int64_t lo = (int64_t)R->loads[x.a] - (int64_t)R->cap;
    const int64_t hi = (int64_t)R->cap - (int64_t)R->loads[x.b];
    out[i] = x; out[i].id = id_base + i;
    if (og_gap(x.net, lo, hi)) out[i].size = 0;
    og_move_clear(&out[count + i]);
    const int begin = og_net_bound(archive, count, x.a, lo - x.net, 0);
    const int end = og_net_bound(archive, count, x.a, hi - x.net, 1);
    int64_t best_gain = 0; int best = -1, previous = -1;
    for (int r = 0; r < 4 && begin < end; ++r) {
        const int j = begin + (int)(((int64_t)(end - begin - 1) * r) / 3);
        if (j == previous) continue;
        previous = j;
        if (j == i) continue;
        const og_move* y = &archive[j];
        if (y->a != x.a || y->b != x.b || x.size + y->size > 8) continue;
        int overlap = 0;
        for (int u = 0; u < x.size; ++u) for (int v = 0; v < y->size; ++v) overlap |= x.vertices[u] == y->vertices[v];
        if (overlap) continue;
        int64_t corr = 0;
        for (int u = 0; u < x.size; ++u) {
            const int32_t v = x.vertices[u];
            for (int64_t e = g->off[v]; e < g->off[v + 1]; ++e)
                for (int t = 0; t < y->size; ++t)
                    if (g->nbr[e] == y->vertices[t])
                        corr += (R->labels[v] == R->labels[g->nbr[e]] ? 2 : -2) * (int64_t)ew_at(g->ew, e);
        }
        const int64_t gain = x.gain + y->gain + corr;
        if (gain > best_gain || (gain == best_gain && gain > 0 && j < best)) { best_gain = gain; best = j; }
    }
    if (best >= 0) {
        og_move res = x;
        const og_move* y = &archive[best];
        for (int j = 0; j < y->size; ++j) res.vertices[res.size++] = y->vertices[j];
        res.gain = best_gain; res.net += y->net; res.id = id_base + count + i;
        out[count + i] = res;
    }
}

static void og_dense_affinity(og_ref* R, uint64_t* A) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t v = 0; v < R->n; ++v) og_affinity(R, v, A + (size_t)v * R->k);
}

/* Capacity-complementary exchange driver: incremental.cuh:974-1041. */
static int og_exchange(og_ref* R, int passes, uint64_t* best_cut) {
    const og_graph* g = R->g;
    const int64_t n = R->n;
    const int k = R->k;
    int sw = 4096 / k; if (sw < 32) sw = 32; if (sw > OG_JOINT_SEEDS) sw = OG_JOINT_SEEDS;
    const int slots = k * sw * 2;
    uint64_t* A = og_alloc((size_t)n * k, sizeof(uint64_t));
    uint64_t* Q = og_alloc((size_t)k * k, sizeof(uint64_t));
    int32_t* visits = og_alloc((size_t)k * k, sizeof(int32_t));
    int32_t* partner = og_alloc((size_t)k, sizeof(int32_t));
    int32_t* choice = og_alloc((size_t)k, sizeof(int32_t));
    uint64_t* seeds = og_alloc((size_t)k * sw, sizeof(uint64_t));
    og_move* grown = og_alloc((size_t)slots, sizeof(og_move));
    og_move* archive = og_alloc((size_t)slots * OG_JOINT_ARCHIVE, sizeof(og_move));
    og_move* pool = og_alloc((size_t)slots + 2 * (size_t)slots * OG_JOINT_ARCHIVE, sizeof(og_move));
    uint64_t* events = og_alloc((size_t)slots + 2 * (size_t)slots * OG_JOINT_ARCHIVE, sizeof(uint64_t));
    int32_t* owners = og_alloc((size_t)n, sizeof(int32_t));
    int32_t* stops = og_alloc((size_t)k, sizeof(int32_t));
    int32_t* movers = og_alloc((size_t)n, sizeof(int32_t));
    int rc = 0;
    og_dense_affinity(R, A);
    for (int64_t v = 0; v < n; ++v) {                                       /* exchange.cuh:41-54 */
        const int a = R->labels[v];
        for (int b = a + 1; b < k; ++b) Q[(size_t)a * k + b] += A[(size_t)v * k + b];
    }
    int empty = 0;
    for (int pass = 0; pass < passes; ++pass) {
        /* joint_match: exchange.cuh:55-83 */
        for (int p = 0; p < k; ++p) partner[p] = -1;
        for (int round = 0; round < k / 2; ++round) {
            int progress = 0;
            for (int p = 0; p < k; ++p) {
                int best = -1; double score = -1;
                if (partner[p] < 0)
                    for (int q = 0; q < k; ++q) {
                        if (q == p || partner[q] >= 0) continue;
                        const int pair = (p < q ? p : q) * k + (p < q ? q : p);
                        if (!Q[pair]) continue;
                        const double penalty = 1.0 + visits[pair];
                        const double value = (double)Q[pair] / (penalty * penalty);
                        if (value > score || (value == score && q < best)) { score = value; best = q; }
                    }
                choice[p] = best;
            }
            for (int p = 0; p < k; ++p) {
                const int b = choice[p];
                if (b >= 0 && choice[b] == p) {
                    partner[p] = b;
                    if (p < b) { ++visits[p * k + b]; ++progress; }
                }
            }
            if (!progress) break;
        }
        /* joint_seeds: exchange.cuh:84-94 */
        for (int i = 0; i < k * sw; ++i) seeds[i] = UINT64_MAX;
        for (int64_t v = 0; v < n; ++v) {
            const int p = R->labels[v], q = partner[p];
            if (q < 0 || !A[(size_t)v * k + q]) continue;
            const uint32_t h = mix32((uint32_t)v ^ mix32((uint32_t)pass + 17));
            const uint64_t val = ((uint64_t)h << 32) | (uint32_t)v;
            uint64_t* slot = &seeds[(size_t)p * sw + h % (uint32_t)sw];
            if (val < *slot) *slot = val;
        }
        for (int s = 0; s < slots; ++s) og_joint_grow(R, A, s, sw, seeds, partner, grown, archive);
        /* archive: remove empty, sort, unique (incremental.cuh:1002-1005) */
        int fr = 0;
        for (int i = 0; i < slots * OG_JOINT_ARCHIVE; ++i) if (archive[i].size) archive[fr++] = archive[i];
        qsort(archive, (size_t)fr, sizeof(og_move), cmp_joint_net);
        int uq = 0;
        for (int i = 0; i < fr; ++i) {
            if (uq && og_same_move(&archive[uq - 1], &archive[i])) continue;
            archive[uq++] = archive[i];
        }
        const int fragments = uq;
        for (int s = 0; s < slots; ++s) pool[s] = grown[s];
        for (int i = 0; i < fragments; ++i) og_joint_combine(R, fragments, archive, slots, pool + slots, i);
        int active = 0;
        for (int i = 0; i < slots + 2 * fragments; ++i) if (pool[i].size) pool[active++] = pool[i];
        if (!active) { if (++empty >= 2) break; continue; }
        qsort(pool, (size_t)active, sizeof(og_move), cmp_joint_order);
        /* claim / resolve / mark: exchange.cuh:205-226 */
        for (int64_t v = 0; v < n; ++v) owners[v] = INT32_MAX;
        for (int i = 0; i < active; ++i)
            for (int j = 0; j < pool[i].size; ++j)
                if (i < owners[pool[i].vertices[j]]) owners[pool[i].vertices[j]] = i;
        for (int i = 0; i < active; ++i)
            for (int j = 0; j < pool[i].size; ++j)
                if (owners[pool[i].vertices[j]] != i) { pool[i].size = 0; break; }
        for (int64_t v = 0; v < n; ++v) owners[v] = INT32_MAX;
        for (int i = 0; i < active; ++i)
            for (int j = 0; j < pool[i].size; ++j) {
                const int32_t v = pool[i].vertices[j];
                owners[v] = i;
                R->targets[v] = R->labels[v] == pool[i].a ? pool[i].b : pool[i].a;
            }
        /* joint_events: exchange.cuh:227-262 */
        memset(events, 0, (size_t)active * sizeof(uint64_t));
        for (int i = 0; i < active; ++i) {
            if (!pool[i].size) continue;
            for (int j = 0; j < pool[i].size; ++j) {
                const int32_t v = pool[i].vertices[j];
                for (int64_t e = g->off[v]; e < g->off[v + 1]; ++e) {
                    const int32_t u = g->nbr[e];
                    const int32_t other = owners[u];
                    if (u == v || (other != INT32_MAX && v > u)) continue;
                    if (other != INT32_MAX && R->labels[u] != pool[i].a && R->labels[u] != pool[i].b) continue;
                    const int64_t w = (int64_t)ew_at(g->ew, e);
                    const int before = R->labels[v] != R->labels[u];
                    if (other == i || other == INT32_MAX) {
                        const int after = R->targets[v] != (other == i ? R->targets[u] : R->labels[u]);
                        events[i] += (uint64_t)((before - after) * w);
                    } else {
                        const int middle = i < other ? R->targets[v] != R->labels[u] : R->labels[v] != R->targets[u];
                        const int after = R->targets[v] != R->targets[u];
                        const uint64_t first = (uint64_t)((before - middle) * w);
                        const uint64_t last = (uint64_t)((middle - after) * w);
                        events[i] += i < other ? first : last;
                        events[other] += i < other ? last : first;
                    }
                }
            }
        }
        /* joint_prefix / joint_accept: exchange.cuh:263-289 */
        for (int a = 0; a < k; ++a) {
            int64_t gain = 0, net = 0, best = 0; int stop = -1;
            int low = 0, high = active;
            while (low < high) { const int mid = (low + high) / 2; if (pool[mid].a < a) low = mid + 1; else high = mid; }
            for (int i = low; i < active && pool[i].a == a; ++i) {
                if (!pool[i].size) continue;
                gain += (int64_t)events[i]; net += pool[i].net;
                const int b = pool[i].b;
                const int64_t lo = (int64_t)R->loads[a] - (int64_t)R->cap;
                const int64_t hi = (int64_t)R->cap - (int64_t)R->loads[b];
                if (!og_gap(net, lo, hi) && gain > best) { best = gain; stop = i; }
            }
            stops[a] = stop;
        }
        for (int64_t v = 0; v < n; ++v) R->flags[v] = 0;
        for (int i = 0; i < active; ++i)
            if (pool[i].size && i <= stops[pool[i].a])
                for (int j = 0; j < pool[i].size; ++j) R->flags[pool[i].vertices[j]] = 1;
        int64_t count = 0;
        for (int64_t v = 0; v < n; ++v) if (R->flags[v]) movers[count++] = (int32_t)v;
        if (!count) { if (++empty >= 2) break; continue; }
        empty = 0;
        const int64_t delta = og_apply(R, movers, count, 0, 0);
        if (delta <= 0) { rc = og_fail("non-improving boundary group batch"); break; }
        if (og_update_cut(R, delta)) { rc = -1; break; }
        R->st->group_moves += (uint64_t)count;
        if (og_max_load(R) > R->cap || R->cut >= *best_cut) { rc = og_fail("capacity exchange violated cut/capacity"); break; }
        *best_cut = R->cut;
        og_dense_affinity(R, A);
    }
    free(A); free(Q); free(visits); free(partner); free(choice); free(seeds); free(grown); free(archive);
    free(pool); free(events); free(owners); free(stops); free(movers);
    return rc;
}

/* run_incremental_refinement_typed: incremental.cuh:660-1051 (host loop
 * 911-961; GPU-graph loop control lp_graph.cuh:33-92). */
static int og_refine_level(const og_graph* g, int32_t* labels, uint64_t* loads, const og_refine_opts* o,
                           uint64_t cap, int rounds, int has_known_cut, uint64_t known_cut, og_level_stats* st) {
    const int64_t n = g->n;
    const int k = o->parts;
    og_ref R;
    memset(&R, 0, sizeof R);
    R.g = g; R.n = n; R.k = k; R.cap = cap; R.labels = labels; R.loads = loads; R.st = st;
    R.targets = og_alloc((size_t)n, sizeof(int32_t));
    R.gains = og_alloc((size_t)n, sizeof(int64_t));
    R.locks = og_alloc((size_t)n, sizeof(int32_t));
    R.flags = og_alloc((size_t)n, sizeof(int32_t));
    R.conn = og_alloc((size_t)k * (size_t)og_threads(), sizeof(uint64_t));
    og_pow15_table(R.pow15);
    uint64_t total_mass = 0;
    for (int p = 0; p < k; ++p) total_mass += loads[p];
    if (total_mass > (uint64_t)n) {
        R.prices = og_alloc((size_t)k, sizeof(double));
        R.samples = og_alloc((size_t)k, sizeof(og_sample));
    }
    int32_t* movers = og_alloc((size_t)n, sizeof(int32_t));
    int32_t* best_labels = og_alloc((size_t)n, sizeof(int32_t));
    uint64_t* best_loads = og_alloc((size_t)k, sizeof(uint64_t));
    int rc = 0;
    st->initial_cut = has_known_cut ? known_cut : og_cut(n, g->off, g->nbr, g->ew, labels);
    R.cut = st->initial_cut;
    if (og_max_load(&R) > cap) {
        int ok;
        if (R.samples) memset(R.samples, 0, (size_t)k * sizeof(og_sample));
        const int feasible = og_rebalance(&R, &ok);
        if (!ok) { rc = -1; goto done; }
        if (!feasible) { st->initial_repair_failed = 1; st->final_cut = R.cut; goto done; }
    }
    {
        const uint64_t feasible_start = R.cut;
        uint64_t best_cut = R.cut;
        memcpy(best_labels, labels, (size_t)n * sizeof(int32_t));
        memcpy(best_loads, loads, (size_t)k * sizeof(uint64_t));
        uint64_t maximum_edge = 0;
        if (!g->ew) maximum_edge = 1;
        else for (int64_t e = 0; e < g->m; ++e) maximum_edge = u64max(maximum_edge, g->ew[e]);
        const double filter = maximum_edge <= 1 ? 0.25 : 0.75;              /* incremental.cuh:897-899 */
        const int graph_mode = n >= o->graph_loop_threshold;
        int stale = 0, empty = 0;
        const int lp_rounds = o->exchange_only ? 0 : rounds;
        for (int phase = 0; phase < lp_rounds && stale < o->stale_rounds; ++phase) {
            st->rounds = phase + 1;
            if (R.samples) memset(R.samples, 0, (size_t)k * sizeof(og_sample));
            const int64_t ncand = og_lp_propose(&R, filter, movers);
            const int64_t count = og_afterburner(&R, movers, ncand);
            if (!count) { if (++empty >= 2) break; ++stale; continue; }
            empty = 0;
            const int64_t delta = og_apply(&R, movers, count, 1, 0);
            if (og_update_cut(&R, delta)) { rc = -1; goto done; }
            st->lp_moves += (uint64_t)count;
            for (int64_t i = 0; i < count; ++i) R.flags[movers[i]] = 0;
            if (og_max_load(&R) > cap) {
                int ok;
                const int feasible = og_rebalance(&R, &ok);
                if (!ok) { rc = -1; goto done; }
                if (!feasible) break;
            }
            if (R.prices)                                                   /* incremental.cuh:161-164 */
                for (int p = 0; p < k; ++p)
                    if (R.samples[p].weight)
                        R.prices[p] = R.samples[p].loss > 0 ? (double)R.samples[p].loss / (double)R.samples[p].weight : 0.0;
            ++stale;
            if (R.cut < best_cut) {
                if (graph_mode) { if (best_cut - R.cut > best_cut / 1000) stale = 0; }
                else if ((long double)R.cut < 0.999L * (long double)best_cut) stale = 0;
                best_cut = R.cut;
                memcpy(best_labels, labels, (size_t)n * sizeof(int32_t));
                memcpy(best_loads, loads, (size_t)k * sizeof(uint64_t));
            }
        }
        memcpy(labels, best_labels, (size_t)n * sizeof(int32_t));
        memcpy(loads, best_loads, (size_t)k * sizeof(uint64_t));
        R.cut = best_cut;
        if (rounds > 0 && o->exchange_rounds > 0 && n <= o->exchange_max_n) {
            for (int64_t v = 0; v < n; ++v) { R.flags[v] = 0; R.targets[v] = -1; }
            if (og_exchange(&R, o->exchange_rounds, &best_cut)) { rc = -1; goto done; }
        }
        st->final_cut = R.cut;
        if (R.cut != best_cut || R.cut > feasible_start || og_max_load(&R) > cap) {
            rc = og_fail("incremental refinement postcondition failed");
            goto done;
        }
    }
done:
    free(R.targets); free(R.gains); free(R.locks); free(R.flags); free(R.conn); free(R.prices); free(R.samples);
    free(movers); free(best_labels); free(best_loads);
    return rc;
}

/* refine_hierarchy_impl: src/refine.cu:483-730. */
int og_uncoarsen(const og_hier* h, const int32_t* coarsest_labels, const og_refine_opts* o,
                 int32_t* out_labels, uint64_t* out_cut, uint64_t* out_loads, og_level_stats* level_stats) {
    const int k = o->parts;
    const og_graph* cg = &h->g[h->levels - 1];
    int32_t* labels = og_alloc((size_t)cg->n, sizeof(int32_t));
    uint64_t* loads = og_alloc((size_t)k, sizeof(uint64_t));
    uint64_t total = 0;
    for (int64_t v = 0; v < cg->n; ++v) {
        if (coarsest_labels[v] < 0 || coarsest_labels[v] >= k) { free(labels); free(loads); return og_fail("invalid coarsest partition"); }
        labels[v] = coarsest_labels[v];
        loads[labels[v]] += cg->vw[v];
        total += cg->vw[v];
    }
    const uint64_t cap = og_capacity(total, k, o->imbalance);
    if (!cap) { free(labels); free(loads); return og_fail("invalid capacity"); }
    og_level_stats dummy;
    og_level_stats* st = level_stats ? &level_stats[0] : &dummy;
    memset(st, 0, sizeof *st);
    st->n = cg->n; st->m = cg->m; st->refined = 1;
    if (og_refine_level(cg, labels, loads, o, cap, o->max_rounds, 0, 0, st)) { free(labels); free(loads); return -1; }
    uint64_t last_cut = st->final_cut;
    int64_t last_refined = cg->n;
    int idx = 1;
    for (int lvl = h->levels - 2; lvl >= 0; --lvl, ++idx) {
        const og_graph* fg = &h->g[lvl];
        const int32_t* map = h->maps[lvl];
        int32_t* fl = og_alloc((size_t)fg->n, sizeof(int32_t));
        for (int64_t v = 0; v < fg->n; ++v) fl[v] = labels[map[v]];        /* refine.cu:100-111 */
        free(labels);
        labels = fl;
        st = level_stats ? &level_stats[idx] : &dummy;
        memset(st, 0, sizeof *st);
        st->n = fg->n; st->m = fg->m;
        uint64_t maxload = 0;
        for (int p = 0; p < k; ++p) maxload = u64max(maxload, loads[p]);
        const int refine = lvl == 0 || maxload > cap || (double)fg->n >= o->growth_ratio * (double)last_refined;
        if (refine) {
            st->refined = 1;
            if (og_refine_level(fg, labels, loads, o, cap, o->max_rounds, 1, last_cut, st)) { free(labels); free(loads); return -1; }
            if (!st->initial_repair_failed) last_refined = fg->n;
        } else {
            st->initial_cut = st->final_cut = last_cut;
        }
        last_cut = st->final_cut;
    }
    uint64_t maxload = 0;
    for (int p = 0; p < k; ++p) maxload = u64max(maxload, loads[p]);
    int rc = 0;
    if (maxload > cap) rc = og_fail("final original partition violates capacity");
    memcpy(out_labels, labels, (size_t)h->g[0].n * sizeof(int32_t));
    if (out_cut) *out_cut = last_cut;
    if (out_loads) memcpy(out_loads, loads, (size_t)k * sizeof(uint64_t));
    free(labels); free(loads);
    return rc;
}

int og_refine_graph(int64_t n, int64_t m, const int32_t* off, const int32_t* nbr, const uint32_t* ew,
                    const uint32_t* vw, int32_t* labels, const og_refine_opts* o, int has_known_cut,
                    uint64_t known_cut, og_level_stats* stats) {
    og_graph g;
    memset(&g, 0, sizeof g);
    g.n = n; g.m = m;
    g.off = (int32_t*)off; g.nbr = (int32_t*)nbr;
    int unit = 1;
    if (ew) for (int64_t e = 0; e < m; ++e) if (ew[e] != 1) { unit = 0; break; }
    g.ew = unit ? NULL : (uint32_t*)ew;
    uint32_t* vv = og_alloc((size_t)n, sizeof(uint32_t));
    for (int64_t v = 0; v < n; ++v) vv[v] = vw ? vw[v] : 1u;
    g.vw = vv;
    const int k = o->parts;
    uint64_t* loads = og_alloc((size_t)k, sizeof(uint64_t));
    uint64_t total = 0;
    for (int64_t v = 0; v < n; ++v) {
        if (labels[v] < 0 || labels[v] >= k) { free(vv); free(loads); return og_fail("invalid partition"); }
        loads[labels[v]] += vv[v]; total += vv[v];
    }
    const uint64_t cap = og_capacity(total, k, o->imbalance);
    og_level_stats dummy;
    og_level_stats* st = stats ? stats : &dummy;
    memset(st, 0, sizeof *st);
    st->n = n; st->m = m; st->refined = 1;
    const int rc = cap ? og_refine_level(&g, labels, loads, o, cap, o->max_rounds, has_known_cut, known_cut, st)
                       : og_fail("invalid capacity");
    free(vv); free(loads);
    return rc;
}

/* ====================================================== single stages
 * Entry points for the oracle's own checks: each runs one stage of the
 * restated schedule on caller arrays through the same functions the level
 * loops call. */
static og_graph og_view(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew, const uint32_t* vw) {
    og_graph g;
    memset(&g, 0, sizeof g);
    g.n = n; g.m = off[n];
    g.off = (int32_t*)off; g.nbr = (int32_t*)nbr; g.ew = (uint32_t*)ew; g.vw = (uint32_t*)vw;
    return g;
}

static void og_agg_view(og_agg* A, const og_graph* g, uint64_t cap, int32_t* clusters, uint64_t* cw) {
    memset(A, 0, sizeof *A);
    A->g = g; A->n = g->n; A->cap = cap; A->clusters = clusters; A->cw = cw;
    A->roles = og_alloc((size_t)g->n, 1);
    A->proposals = og_alloc((size_t)g->n, sizeof(int32_t));
    A->gains = og_alloc((size_t)g->n, sizeof(uint64_t));
    A->regret = og_alloc((size_t)g->n, sizeof(uint64_t));
    A->ties = og_alloc((size_t)g->n, sizeof(uint32_t));
    int64_t maxdeg = 0;
    for (int64_t v = 0; v < g->n; ++v) if (g->off[v + 1] - g->off[v] > maxdeg) maxdeg = g->off[v + 1] - g->off[v];
    A->scratch_stride = maxdeg + 1;
    A->scratch = og_alloc((size_t)A->scratch_stride * (size_t)og_threads(), sizeof(og_cw_pair));
}

static void og_agg_view_free(og_agg* A) {
    free(A->roles); free(A->proposals); free(A->gains); free(A->regret); free(A->ties); free(A->scratch);
}

int64_t og_t_seed_match(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew, const uint32_t* vw,
                        uint64_t cap, uint32_t salt, int32_t* roots) {
    og_graph g = og_view(n, off, nbr, ew, vw);
    uint64_t* cw = og_alloc((size_t)n, sizeof(uint64_t));
    og_agg A;
    og_agg_view(&A, &g, cap, roots, cw);
    const int64_t count = og_seed_match(&A, salt);
    og_agg_view_free(&A);
    free(cw);
    return count;
}

uint32_t og_t_mover_threshold(int64_t clusters, int64_t n) { return og_mover_threshold(clusters, n); }
/* The small-level stale rule of the LP loop above (incremental.cuh:911-961). */
int og_t_below_0999(uint64_t cut, uint64_t best) { return (long double)cut < 0.999L * (long double)best; }

void og_t_proposals(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew, const uint32_t* vw,
                    const int32_t* clusters, const uint64_t* cw, uint64_t cap, uint32_t seed, int level, int round,
                    int filter_roles, uint32_t threshold, int need_regret, uint8_t* roles, int32_t* proposals,
                    uint64_t* gains, uint64_t* regret, uint32_t* ties) {
    og_graph g = og_view(n, off, nbr, ew, vw);
    og_agg A;
    og_agg_view(&A, &g, cap, (int32_t*)clusters, (uint64_t*)cw);
    og_form_proposals(&A, seed, level, round, filter_roles, threshold, need_regret);
    memcpy(roles, A.roles, (size_t)n);
    memcpy(proposals, A.proposals, (size_t)n * sizeof(int32_t));
    memcpy(gains, A.gains, (size_t)n * sizeof(uint64_t));
    memcpy(regret, A.regret, (size_t)n * sizeof(uint64_t));
    memcpy(ties, A.ties, (size_t)n * sizeof(uint32_t));
    og_agg_view_free(&A);
}

int og_t_pack_isolated(int64_t n, const int32_t* off, const uint32_t* vw, uint64_t cap, int32_t* clusters,
                       uint64_t* cw) {
    og_graph g = og_view(n, off, NULL, NULL, vw);
    og_agg A;
    og_agg_view(&A, &g, cap, clusters, cw);
    const int ran = og_pack_isolated(&A);
    og_agg_view_free(&A);
    return ran;
}

void og_t_two_hop(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew, const uint32_t* vw,
                  int32_t* clusters, uint64_t* cw, uint64_t cap, uint32_t seed, int level, int64_t budget) {
    og_graph g = og_view(n, off, nbr, ew, vw);
    og_agg A;
    og_agg_view(&A, &g, cap, clusters, cw);
    og_adm* adm = og_alloc((size_t)n, sizeof(og_adm));
    og_two_hop(&A, seed, level, budget, adm);
    free(adm);
    og_agg_view_free(&A);
}

static void og_ref_view(og_ref* R, const og_graph* g, int k, uint64_t cap, int32_t* labels, uint64_t* loads,
                        const double* prices, og_level_stats* st) {
    memset(R, 0, sizeof *R);
    R->g = g; R->n = g->n; R->k = k; R->cap = cap; R->labels = labels; R->loads = loads; R->st = st;
    R->targets = og_alloc((size_t)g->n, sizeof(int32_t));
    R->gains = og_alloc((size_t)g->n, sizeof(int64_t));
    R->locks = og_alloc((size_t)g->n, sizeof(int32_t));
    R->flags = og_alloc((size_t)g->n, sizeof(int32_t));
    R->conn = og_alloc((size_t)k * (size_t)og_threads(), sizeof(uint64_t));
    og_pow15_table(R->pow15);
    if (prices) {
        R->prices = og_alloc((size_t)k, sizeof(double));
        memcpy(R->prices, prices, (size_t)k * sizeof(double));
    }
}

static void og_ref_view_free(og_ref* R) {
    free(R->targets); free(R->gains); free(R->locks); free(R->flags); free(R->conn); free(R->prices);
}

int64_t og_t_lp_propose(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew, const uint32_t* vw,
                        int k, const int32_t* labels, int32_t* locks, const uint64_t* loads, uint64_t cap,
                        const double* prices, double filter, int32_t* targets, int64_t* gains) {
    og_graph g = og_view(n, off, nbr, ew, vw);
    og_ref R;
    og_level_stats st;
    memset(&st, 0, sizeof st);
    og_ref_view(&R, &g, k, cap, (int32_t*)labels, (uint64_t*)loads, prices, &st);
    memcpy(R.locks, locks, (size_t)n * sizeof(int32_t));
    int32_t* movers = og_alloc((size_t)n, sizeof(int32_t));
    const int64_t ncand = og_lp_propose(&R, filter, movers);
    memcpy(locks, R.locks, (size_t)n * sizeof(int32_t));
    memcpy(targets, R.targets, (size_t)n * sizeof(int32_t));
    memcpy(gains, R.gains, (size_t)n * sizeof(int64_t));
    free(movers);
    og_ref_view_free(&R);
    return ncand;
}

int64_t og_t_afterburner(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew, const uint32_t* vw,
                         int k, const int32_t* labels, const uint64_t* loads, uint64_t cap, const double* prices,
                         const int32_t* targets, const int64_t* gains, uint8_t* keep) {
    og_graph g = og_view(n, off, nbr, ew, vw);
    og_ref R;
    og_level_stats st;
    memset(&st, 0, sizeof st);
    og_ref_view(&R, &g, k, cap, (int32_t*)labels, (uint64_t*)loads, prices, &st);
    memcpy(R.targets, targets, (size_t)n * sizeof(int32_t));
    memcpy(R.gains, gains, (size_t)n * sizeof(int64_t));
    int32_t* movers = og_alloc((size_t)n, sizeof(int32_t));
    int64_t ncand = 0;
    for (int64_t v = 0; v < n; ++v) if (targets[v] >= 0) movers[ncand++] = (int32_t)v;
    const int64_t count = og_afterburner(&R, movers, ncand);
    for (int64_t v = 0; v < n; ++v) keep[v] = (uint8_t)(targets[v] >= 0 && R.flags[v]);
    free(movers);
    og_ref_view_free(&R);
    return count;
}

int og_t_gain_bucket(double gain) {
    double t[48];
    og_pow15_table(t);
    return og_gain_bucket_table(t, gain);
}

int og_t_repair_destination(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew,
                            const uint32_t* vw, int k, const int32_t* labels, const uint64_t* loads, uint64_t cap,
                            int32_t v, uint64_t* best_conn, uint64_t* current) {
    og_graph g = og_view(n, off, nbr, ew, vw);
    og_ref R;
    og_level_stats st;
    memset(&st, 0, sizeof st);
    og_ref_view(&R, &g, k, cap, (int32_t*)labels, (uint64_t*)loads, NULL, &st);
    const uint64_t total = og_affinity(&R, v, R.conn);
    const int dest = og_repair_destination(&R, v, labels[v], vw[v], R.conn, total, best_conn, current);
    og_ref_view_free(&R);
    return dest;
}

int64_t og_t_repair_iteration(int64_t n, const int32_t* off, const int32_t* nbr, const uint32_t* ew,
                              const uint32_t* vw, int k, const int32_t* labels, const uint64_t* loads, uint64_t cap,
                              int32_t* first_choice, int32_t* bucket, uint8_t* eligible, int32_t* receiver) {
    og_graph g = og_view(n, off, nbr, ew, vw);
    og_ref R;
    og_level_stats st;
    memset(&st, 0, sizeof st);
    og_ref_view(&R, &g, k, cap, (int32_t*)labels, (uint64_t*)loads, NULL, &st);
    og_repair_ws W;
    og_repair_ws_alloc(&W, n, k);
    const int64_t ecount = og_repair_select(&R, &W);
    memcpy(first_choice, R.targets, (size_t)n * sizeof(int32_t));
    memcpy(bucket, W.bucket, (size_t)n * sizeof(int32_t));
    memcpy(eligible, W.eligible, (size_t)n);
    og_repair_reserve(&R, &W, ecount);
    int64_t admitted = 0;
    for (int64_t v = 0; v < n; ++v) {
        receiver[v] = R.flags[v] ? R.targets[v] : -1;
        admitted += R.flags[v] != 0;
    }
    og_repair_ws_free(&W);
    og_ref_view_free(&R);
    return admitted;
}
