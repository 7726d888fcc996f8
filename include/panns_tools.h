/*
 * Host-side tools that surround the search path: deterministic synthetic
 * inputs, the reference's on-disk formats, exact brute-force k-NN (the
 * recall judge) and NSG-style graph construction.  Plain C ABI; every entry
 * point returns 0 on success and a negative code on failure (no exit()).
 *
 * Formats follow the reference:
 *   fvecs   include/utils.h:95-142 (per row int32 D + D float32)
 *   .nsg    include/efanna2e/index_nsg.cpp:20-32 and PSS5LG.cpp:138-251
 *           (uint32 width, uint32 ep, per vertex uint32 deg + deg ids)
 *   truth   tools/save_true_nearest_neighbors.cpp:37-64
 *           (uint32 Q, uint32 K, then Q*K {uint32 id, float32 dist})
 *   result  include/utils.h:146-163 (per query uint32 K + K ids)
 */
#ifndef PANNS_TOOLS_H
#define PANNS_TOOLS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_OK 0
#define PT_ERR_ARG (-1)
#define PT_ERR_IO (-2)
#define PT_ERR_FORMAT (-3)
#define PT_ERR_ALLOC (-4)
#define PT_ERR_RANGE (-5)

/* Thread-local message for the last failing call. */
const char *pt_last_error(void);

/* Exact squared L2 with the reference's summation order
 * (PSS5LG.cpp:525-560 as compiled: per-lane fma chain, tail block first,
 * sequential horizontal sum).  dim % 8 == 0. */
float pt_l2(const float *v, const float *q, uint32_t dim);

/* Seeded Gaussian mixture, bit-reproducible on any x86-64 host: integer
 * RNG per row, Marsaglia polar normals with a project-written log, fp64
 * arithmetic without contraction.  Centres c_j ~ N(0, I) from seed_centers;
 * row i: label uniform in [0, n_comp), x = c_label + sigma * N(0, I), then
 * divided by its L2 norm when normalize != 0.  Rows [row0, row0 + n). */
int pt_gen_mixture(uint64_t seed_centers, uint64_t seed_rows, uint64_t row0, uint32_t n,
                   uint32_t dim, uint32_t n_comp, float sigma, int normalize, float *out,
                   int nthreads);
/* The mixture component of rows [row0, row0+n) of pt_gen_mixture's stream
 * seed_rows (the first draw of each row's generator). */
int pt_mixture_labels(uint64_t seed_rows, uint64_t row0, uint32_t n, uint32_t n_comp, uint32_t *labels,
                      int nthreads);

/* fvecs */
int pt_fvecs_info(const char *path, uint32_t *n, uint32_t *dim);
int pt_fvecs_read(const char *path, float *out, uint32_t n, uint32_t dim);
int pt_fvecs_write(const char *path, const float *data, uint32_t n, uint32_t dim);

/* .nsg, from / to a padded adjacency: adj[v*stride + i], i < deg[v]. */
int pt_nsg_info(const char *path, uint32_t *width, uint32_t *ep, uint32_t *n,
                uint64_t *num_e, uint32_t *max_deg);
int pt_nsg_read(const char *path, uint32_t n, uint32_t stride, uint32_t *adj, uint32_t *deg);
int pt_nsg_write(const char *path, uint32_t width, uint32_t ep, uint32_t n, uint32_t stride,
                 const uint32_t *adj, const uint32_t *deg);

/* ground truth and result files */
int pt_truth_info(const char *path, uint32_t *nq, uint32_t *k);
int pt_truth_read(const char *path, uint32_t nq, uint32_t k, uint32_t *ids, float *dists);
int pt_truth_write(const char *path, uint32_t nq, uint32_t k, const uint32_t *ids,
                   const float *dists);
int pt_result_write(const char *path, uint32_t nq, uint32_t k, const uint32_t *ids);

/* Exact k-NN by brute force with pt_l2 and ties to the lower id (the
 * semantics of tools/save_true_nearest_neighbors.cpp:16-35).  When
 * exclude_self != 0, query i is base row i and is skipped (graph k-NN). */
int pt_exact_knn(const float *base, uint32_t n, uint32_t dim, const float *queries,
                 uint32_t nq, uint32_t k, int exclude_self, uint32_t *out_ids, float *out_dists,
                 int nthreads);

/* recall@R as get_recall_for_all_queries (PSS5LG.cpp:309-348) computes
 * recall@100, generalised to any R <= min(true_k, res_k). */
double pt_recall(const uint32_t *true_ids, uint32_t true_k, const uint32_t *res_ids,
                 uint32_t res_k, uint32_t nq, uint32_t L, uint32_t R);
/* per-query recall@R into out[nq] */
int pt_recall_per_query(const uint32_t *true_ids, uint32_t true_k, const uint32_t *res_ids,
                        uint32_t res_k, uint32_t nq, uint32_t L, uint32_t R, float *out);

/* Navigating node: the row nearest (pt_l2, ties to the lower id) to the
 * fp64 mean of all rows, rounded to fp32 (index_nsg.cpp:211-231 uses the
 * search result for the centroid). */
int pt_navigating_node(const float *base, uint32_t n, uint32_t dim, uint32_t *ep, int nthreads);

/* NSG-style graph from exact k-NN lists (knn[v*knn_k + i], sorted by
 * (dist, id), self excluded):
 *   1. pool(v) = kNN(v) plus every vertex expanded by a best-first search for
 *      v over the k-NN graph from ep with queue length L_build
 *      (index_nsg.cpp:354-401 Link uses the search's visited set);
 *   2. RNG occlusion prune of the pool to <= R (index_nsg.cpp:233-283);
 *   3. reverse edges, re-pruned when over R (index_nsg.cpp:285-352);
 *   4. DFS from ep; each unreached vertex is attached to the nearest reached
 *      vertex among its k-NN that still has degree < R (else the first
 *      reached vertex with spare degree), so degree never exceeds R
 *      (index_nsg.cpp:638-718 tree_grow, whose findroot can exceed R).
 * L_build = 0 skips step 1's search (the pool is the k-NN list alone) and
 * options without PT_GRAPH_REVERSE_EDGES skip step 3, which together give the
 * plain "k-NN, RNG prune, DFS stitch" graph.
 * Output adj[v*R + i] for i < deg[v] (sorted by (dist, id) from v), ep. */
#define PT_GRAPH_REVERSE_EDGES 1u
int pt_build_graph(const float *base, uint32_t n, uint32_t dim, const uint32_t *knn,
                   uint32_t knn_k, uint32_t R, uint32_t L_build, uint32_t C, uint32_t ep,
                   uint32_t options, uint32_t *adj, uint32_t *deg, int nthreads);

/* Vertices reachable from ep (BFS); returns the count in *reached. */
int pt_reachable(const uint32_t *adj, const uint32_t *deg, uint32_t n, uint32_t stride,
                 uint32_t ep, uint32_t *reached);

#ifdef __cplusplus
}
#endif

#endif
