/*
 * CPU oracle for the NSG top-M / PSS_v5 search path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * algorithm used as the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (parallel_anns_amd) never links or calls it.
 *
 * Every function names the reference file:line it follows.  Tags:
 *   PSS5LG = core/Searching.202102031939.PSS_v5.large_graph.dist_thresh.profiling.{cpp,h}
 *   TOPMLG = core/Searching.202011241615.top_m_pure_large_graph.{cpp,h}
 *   PSS5   = core/Searching.202101301609.PSS_v5.profiling.{cpp,h}   (opt-graph family)
 *   TOPM   = core/Searching.202011241533.top_m_pure.{cpp,h}         (opt-graph family)
 *
 * The opt-graph family runs the same searches with the distance
 * norm(v) - 2 v.q (PSS5.cpp:336-390) instead of squared L2; a po_graph with
 * non-NULL norms selects it.
 *
 * Pinning (what checks each function against something other than itself):
 *   po_distance            AVX2 intrinsics build with GCC contraction (tests),
 *                          fp64 closeness, v==q -> +0, one-hot exact values.
 *   po_norm, po_distance_with_norm   the same AVX2 route for the reference's
 *                          norm and dot kernels; exact values on integer
 *                          vectors (norm - 2 v.q == |v-q|^2 - |q|^2); fp64
 *                          closeness.  Searches with norms: on integer data
 *                          equal the squared-L2 searches (same ids, counts;
 *                          distances shifted by |q|^2 exactly).
 *   po_add_into_queue      brute force over tiny queues (sorted insert model).
 *   po_prepare_init_ids    hand-made graphs with known answers.
 *   po_search_top_m        L == N reduces to exact brute force; complete graph;
 *                          an independent set-semantics (top-L of union)
 *                          restatement in numpy; invariants
 *                          (#dist = L + |visited \ init|, final queue all checked).
 *   po_search_pss_v5_t1    equals po_search_top_m with M = 1 for every X.
 *   po_merge_two_queues_fixed / _incr   hand-made queues (duplicates with
 *                          every checked/unchecked pair, the last-index and
 *                          past-the-end cases) against a set-merge model.
 *   po_search_pss_v5       T = 1 equals po_search_pss_v5_t1; for T > 1 an
 *                          independent list-based restatement of the lockstep
 *                          schedule in Python (ids, distances, checked
 *                          flags, counts); invariants (final queue all
 *                          checked, #dist = L + |visited \ init|); recall
 *                          parity with the top-M search.
 *   po_subsearch_with_top_m   scale-M equals the same set semantics with the
 *                          doubling M schedule.
 *   po_brute_force_knn     tiny inputs against numpy float64 ranking with ties
 *                          checked by id; equals L == N search.
 *   po_search_pss_v5_par   the CPU baseline's timed reference design (its T
 *                          threads race as the reference's do, so no parity
 *                          claim at T > 1): T = 1 equals po_search_top_m
 *                          with M = 1 for every X; T > 1 by properties (K
 *                          distinct ids, #dist >= L, recall within 0.02 of
 *                          po_search_pss_v5).
 * The reference itself cannot be built here (it needs Boost, which the image
 * lacks), so no golden vector comes from a compiled reference.
 */
#ifndef PANNS_ORACLE_H
#define PANNS_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/Candidate.h:11-28 : {id, distance, is_checked}, ordered by (distance, id). */
typedef struct {
    uint32_t id;
    float dist;
    uint8_t checked;
    uint8_t pad_[3];
} po_cand;

/* CSR graph as built by load_common_nsg_graph (PSS5LG.cpp:138-251):
 * vertex_base[v] indexes deg_ngbrs, where deg_ngbrs[vertex_base[v]] = degree
 * followed by the neighbour ids in file order. */
typedef struct {
    uint32_t num_v;
    uint32_t width;
    uint32_t ep;
    uint64_t num_e;
    const uint64_t *vertex_base;
    const uint32_t *deg_ngbrs;
    /* NULL: squared-L2 distances (large-graph family).  Else norms[v] =
     * po_norm of base row v, and distances are norm - 2 v.q (opt-graph
     * family, load_nsg_graph PSS5.cpp:56-132 stores it in each record). */
    const float *norms;
} po_graph;

typedef struct {
    uint64_t distance_computations; /* count_distance_computation_ (PSS5LG.h:170) */
    uint64_t merges;                /* count_merge_ (PSS5LG.h:171) */
    uint64_t expansions;            /* candidates expanded (hops) */
} po_counters;

/* PSS5LG.cpp:525-560 compute_distance as compiled with -O3 -march=native:
 * vsubps + vfmadd per 8-lane accumulator, tail block first, then a sequential
 * horizontal sum.  dim must be a multiple of 8. */
float po_distance(const float *v, const float *q, uint32_t dim);

/* PSS5.cpp:293-334 compute_norm: the same chain with fma(v, v, acc). */
float po_norm(const float *v, uint32_t dim);

/* PSS5.cpp:336-390 compute_distance_with_norm: the same chain with
 * fma(v, q, acc), then -2 * dot + norm. */
float po_distance_with_norm(const float *v, const float *q, uint32_t dim, float norm);

/* include/Candidate.h:19-27 operator< */
int po_cand_less(const po_cand *a, const po_cand *b);

/* PSS5LG.cpp:565-608 add_into_queue.  Returns the insert index relative to
 * queue_start, or queue_capacity when rejected (duplicate or full). */
uint32_t po_add_into_queue(po_cand *queue, uint32_t queue_start, uint32_t *queue_size,
                           uint32_t queue_capacity, po_cand cand);

/* PSS5LG.cpp:356-410 prepare_init_ids.  Returns 0, or -1 when L > num_v. */
int po_prepare_init_ids(const po_graph *g, uint32_t L, uint32_t *init_ids);

/* TOPMLG.cpp:2022-2207 para_search_with_top_m_pure_large_graph with
 * num_threads_ == 1 (TOPM.cpp:1881-2070 para_search_with_top_m_pure, its
 * num_threads_ == 1 branch at :1922, when g->norms != NULL).  visited must hold num_v bits, all clear; it is cleared
 * again on return (PSS5LG.cpp:1691).  set_L receives the final queue (L
 * entries), set_K the first K ids. */
void po_search_top_m(const po_graph *g, const float *base, uint32_t dim, const float *query,
                     uint32_t K, uint32_t L, uint32_t M, const uint32_t *init_ids,
                     po_cand *set_L, uint32_t *set_K, uint64_t *visited, po_counters *cnt);

/* PSS5LG.cpp:1376-1724 para_search_PSS_v5_large_graph_dist_thresh_profiling
 * with num_threads_ == 1 (X = subsearch_iterations); PSS5.cpp:1211-1462
 * para_search_PSS_v5_profiling when g->norms != NULL. */
void po_search_pss_v5_t1(const po_graph *g, const float *base, uint32_t dim, const float *query,
                         uint32_t K, uint32_t L, uint32_t X, const uint32_t *init_ids,
                         po_cand *set_L, uint32_t *set_K, uint64_t *visited, po_counters *cnt);

/* PSS5LG.cpp:648-719 merge_two_queues_into_1st_queue_seq_fixed: merges the
 * sorted queue2 into queue1 (queue1_size entries, size kept).  A duplicate
 * id keeps queue1's copy, unchecked when queue2's copy is; an element that
 * lands on queue1's last index replaces it.  Returns the lower-bound index
 * of queue2's first element in queue1 (queue1_size: nothing merged). */
uint32_t po_merge_two_queues_fixed(po_cand *queue1, uint32_t queue1_start, uint32_t queue1_size,
                                   po_cand *queue2, uint32_t queue2_start, uint32_t queue2_size);

/* PSS5LG.cpp:725-805 merge_two_queues_into_1st_queue_seq_incr: the same
 * merge into a queue that may grow up to queue1_length entries. */
uint32_t po_merge_two_queues_incr(po_cand *queue1, uint32_t queue1_start, uint32_t *queue1_size,
                                  uint32_t queue1_length, po_cand *queue2, uint32_t queue2_start,
                                  uint32_t queue2_size);

/* PSS5LG.cpp:1376-1724 para_search_PSS_v5_large_graph_dist_thresh_profiling
 * with num_threads_ = T: pick_top_m_to_workers (PSS5LG.cpp:917-947), up to
 * X expansions per queue (PSS5LG.cpp:1584-1646) and
 * merge_all_queues_to_master (PSS5LG.cpp:811-842).  The reference's threads
 * race; this runs them in lockstep in a fixed order (step s expands the
 * s-th candidate of worker 0, 1, ..., T-2, then of the master), one of the
 * schedules the reference can take, so results are reproducible.  set_L
 * holds (T-1)*local_L + L entries, the master last; set_K gets its first K
 * ids.  T = 1 is po_search_pss_v5_t1. */
void po_search_pss_v5(const po_graph *g, const float *base, uint32_t dim, const float *query, uint32_t K,
                      uint32_t L, uint32_t X, uint32_t T, uint32_t local_L, const uint32_t *init_ids, po_cand *set_L,
                      uint32_t *set_K, uint64_t *visited, po_counters *cnt);

/* The reference design for the CPU baseline (timing, not parity): queries
 * one after another, each by PSS_v5 with T OpenMP threads expanding at once
 * (PSS5LG.cpp:1376-1724, prepare_init_ids inside), the shared visited bits
 * claimed with atomic fetch-or and the master's L-th distance read with
 * relaxed atomics.  Results depend on the schedule at T > 1.  out_ids gets
 * nq x K ids; returns the distance computations. */
uint64_t po_search_pss_v5_par(const po_graph *g, const float *base, uint32_t dim, const float *queries,
                              uint32_t nq, uint32_t K, uint32_t L, uint32_t X, uint32_t T, uint32_t local_L,
                              uint32_t *out_ids, uint32_t *out_ndist);

/* The compiler that built the library (__VERSION__). */
const char *po_build_info(void);

/* TOPMLG.cpp:1023-1071 subsearch_with_top_m + :1077-1131
 * subsearch_top_m_for_one_iteration.  Works in place on set_L[set_L_start ..
 * set_L_start + *set_L_size) with capacity local_L and the caller's visited
 * bits (left set on return, as in the reference). */
void po_subsearch_with_top_m(const po_graph *g, const float *base, uint32_t dim,
                             const float *query, uint32_t M_max, uint32_t local_L,
                             po_cand *set_L, uint32_t set_L_start, uint32_t *set_L_size,
                             uint64_t *visited, uint64_t *count_distance);

/* PSS5LG.cpp:1055-1096 initialize_set_L_para (marks init ids visited,
 * computes L distances, sorts); PSS5.cpp:890-927 when norms != NULL.
 * Exposed for subsearch tests. */
void po_initialize_set_L(const float *base, uint32_t dim, const float *query, uint32_t L,
                         po_cand *set_L, uint32_t set_L_start, uint32_t *set_L_size,
                         const uint32_t *init_ids, uint64_t *visited, uint64_t *count_distance,
                         const float *norms);

/* Batch driver used for the CPU baseline: queries [q0, q0+nq) searched
 * independently, one per OpenMP thread (inter-query parallel, num_t = 1 per
 * query, the PSS_v5_multi_queries pattern with inner = 1).  mode 0 = top-M,
 * 1 = PSS_v5 T=1 with X = M.  out_ids is nq*K.  Returns total distances. */
uint64_t po_search_batch(const po_graph *g, const float *base, uint32_t dim,
                         const float *queries, uint32_t nq, uint32_t K, uint32_t L,
                         uint32_t M, int mode, const uint32_t *init_ids, uint32_t *out_ids,
                         float *out_dists, uint32_t *out_ndist, int nthreads);

/* tools/save_true_nearest_neighbors.cpp:16-35 get_true_NN: every base point
 * offered to add_into_queue with capacity K, so ties go to the lower id. */
void po_brute_force_knn(const float *base, uint32_t n, uint32_t dim, const float *queries,
                        uint32_t nq, uint32_t K, uint32_t *out_ids, float *out_dists,
                        int nthreads);

/* PSS5LG.cpp:309-348 get_recall_for_all_queries, recall@100 only, generalised
 * to recall@R for R <= min(K_true, K_res):  sum_q |GT_R ∩ res[0..min(L,R))| / (R*Q). */
double po_recall(const uint32_t *true_ids, uint32_t true_k, const uint32_t *res_ids,
                 uint32_t res_k, uint32_t nq, uint32_t L, uint32_t R);

#ifdef __cplusplus
}
#endif

#endif
