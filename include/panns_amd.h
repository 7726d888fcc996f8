/*
 * C-ABI of the MI355X NSG top-M search (the drop-in boundary).
 *
 * The reference boundary is the C++ class PANNS::Searching (no FFI); these
 * entry points replace its search-path methods one for one and are what a
 * foreign binding (ctypes, cgo, JNI) or the C++ drop-in in
 * include/panns/Searching.h call.  Tags:
 *   PSS5LG = core/Searching.202102031939.PSS_v5.large_graph.dist_thresh.profiling.{cpp,h}
 *   TOPMLG = core/Searching.202011241615.top_m_pure_large_graph.{cpp,h}
 *
 * Conventions: every call returns PANNS_OK (0) or a negative panns_status;
 * nothing calls exit().  panns_last_error() gives the message of the last
 * failing call on the calling thread.  One handle owns one HIP stream and its
 * device memory; calls on one handle are serialised by the caller; distinct
 * handles are independent.  No torch or HIP types appear in the signatures.
 */
#ifndef PANNS_AMD_H
#define PANNS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PANNS_OK = 0,
    PANNS_ERR_ARG = -1,         /* bad argument (K > L, L > N, M == 0, null pointer, ...) */
    PANNS_ERR_IO = -2,          /* file cannot be opened / read */
    PANNS_ERR_FORMAT = -3,      /* truncated or inconsistent file */
    PANNS_ERR_DIM = -4,         /* base / query dimension mismatch or unsupported D */
    PANNS_ERR_RANGE = -5,       /* neighbour id >= N, vertex-count mismatch, id space */
    PANNS_ERR_ALLOC = -6,       /* host or device allocation failed */
    PANNS_ERR_DEVICE = -7,      /* HIP runtime error */
    PANNS_ERR_UNSUPPORTED = -8, /* valid in the reference but not offered here */
    PANNS_ERR_STATE = -9        /* call order: e.g. search before the graph is loaded */
} panns_status;

typedef enum {
    PANNS_MODE_TOP_M = 0,   /* para_search_with_top_m_pure_large_graph (TOPMLG.cpp:2022-2207) */
    PANNS_MODE_PSS_V5 = 1,  /* para_search_PSS_v5_large_graph_dist_thresh_profiling (PSS5LG.cpp:1376-1724) */
    PANNS_MODE_SCALE_M = 2  /* M = 1, 2, 4, ... up to M (the subsearch_with_top_m schedule, TOPMLG.cpp:1023-1071) */
} panns_mode;

/* Distance family of a handle, chosen by the graph loader:
 * load_common_nsg_graph (the large-graph searches, PSS5LG / TOPMLG) -> L2,
 * load_nsg_graph (the opt-graph searches, PSS5.cpp / TOPM.cpp) -> NORM_DOT. */
typedef enum {
    PANNS_DISTANCE_L2 = 0,       /* squared L2, compute_distance (PSS5LG.cpp:525-560) */
    PANNS_DISTANCE_NORM_DOT = 1  /* norm(v) - 2 v.q, compute_distance_with_norm (PSS5.cpp:336-390);
                                    q.q is left out, so values may be negative */
} panns_distance;

typedef enum {
    PANNS_STOP_FIXED = 0,   /* run until no unchecked candidate (the reference) */
    PANNS_STOP_LEARNED = 1  /* stop when the loaded MLP predicts the top-K final */
} panns_stop;

typedef struct {
    uint32_t K;            /* results per query, K <= L */
    uint32_t L;            /* queue length, L <= N */
    uint32_t M;            /* TOP_M: picks per iteration; SCALE_M: M_max; PSS_V5: X (sub-search steps) */
    uint32_t mode;         /* panns_mode */
    uint32_t num_threads;  /* PSS_V5: the reference's num_threads_, 1 to 16 queues.  Above 1 the
                              queues run in lockstep in a fixed order (see panns_search); the
                              other modes take 1 */
    uint32_t stop;         /* panns_stop */
    float stop_threshold;  /* LEARNED: stop when sigmoid output >= threshold */
    uint32_t hash_slots;   /* 0 = automatic; else a multiple of 64 in [64, 32768]: 16-bit slots of
                              the per-query LDS visited table (raised when the id space needs).
                              Automatic: 8,192 up to L = 512, 10,240 above, scaled by the launch's
                              largest L when no plan options give a size (see
                              panns_plan_options.table_eighths = 0); an automatic table then
                              grows into the LDS that the 1,280-byte allocation granule leaves
                              free at the launch's workgroups per CU (not for buckets side by
                              side).  Results never depend on the size */
    uint32_t waves;        /* wavefronts per query: 0 or 1, else 2 or 4 (TOP_M and SCALE_M with the
                              fixed stop): the picks of an iteration are expanded by several waves
                              and merged into the master queue, the deterministic form of the
                              reference's multi-thread merge (PSS5LG.cpp:811-842, 1543-1673);
                              results equal the single-wave search */
    uint32_t stop_min_iter;  /* LEARNED: the model runs from this iteration on (0: from the first) */
    uint32_t stop_every;     /* LEARNED: ... at every stop_every-th iteration (0 or 1: every one) */
    uint32_t bucket_slots;   /* visited-table buckets of 8, 4 or 2 slots; 0 = automatic (8, or 4 when the
                                id space needs more buckets than an 8-slot table of that size has).
                                The id space sets the least bucket count (ids / 2^15), so narrower
                                buckets allow smaller tables on large id spaces */
} panns_params;

typedef struct {
    uint64_t distance_computations; /* count_distance_computation_ (PSS5LG.h:170) */
    uint64_t merges;                /* count_merge_ (PSS5LG.h:171) as the reference counts it */
    uint64_t expansions;            /* candidates expanded */
    uint64_t adjacency_slots;       /* neighbour ids read (sum of expanded degrees) */
    uint64_t visited_spills;        /* vertices kept in the global bitmap, not the LDS hash */
    double kernel_ms;               /* device time of the search kernel(s) */
} panns_counters;

/* Outputs of a search, host memory.  ids is required; the rest may be NULL.
 * Per-query arrays are indexed by the position of the query in the call. */
typedef struct {
    uint32_t *ids;           /* nq * K: set_K */
    float *dists;            /* nq * K */
    uint32_t *queue_ids;     /* nq * L: final set_L ids */
    float *queue_dists;      /* nq * L: final set_L distances */
    uint8_t *queue_checked;  /* nq * L: final set_L is_checked_ */
    uint32_t *ndist;         /* nq: distance computations per query */
    uint32_t *nexp;          /* nq: expansions per query */
    uint32_t *niter;         /* nq: iterations (picks rounds) per query */
} panns_outputs;

/* A handle and the plans made from it are used from one host thread at a
 * time (its scratch buffers, stream and plan registry are not locked);
 * separate handles may be used from separate threads. */
typedef struct panns_index panns_index;
typedef struct panns_plan panns_plan;

const char *panns_last_error(void);

/* Handle on HIP device `device` (ordinal).  Replaces PANNS::Searching's
 * constructor; destroy replaces ~Searching (PSS5LG.h:211-219).  Plans still
 * held when their handle is destroyed are detached: running or fetching
 * them returns PANNS_ERR_STATE, and panns_plan_destroy frees them. */
int panns_create(int device, panns_index **out);
int panns_destroy(panns_index *h);

/* PSS5LG.cpp:14-30 load_data_load: fvecs base set; errors on a dimension
 * that differs from already-loaded queries. */
int panns_load_data_load(panns_index *h, const char *fvecs_path);
/* PSS5LG.cpp:35-49 load_queries_load. */
int panns_load_queries_load(panns_index *h, const char *fvecs_path);
/* In-memory forms of the two loaders (row-major n x dim float32). */
int panns_set_base(panns_index *h, const float *data, uint32_t n, uint32_t dim);
int panns_set_queries(panns_index *h, const float *data, uint32_t nq, uint32_t dim);

/* PSS5LG.cpp:138-251 load_common_nsg_graph: .nsg file; errors when the
 * vertex count differs from the base set or a neighbour id is >= N.  Sets
 * the distance to PANNS_DISTANCE_L2. */
int panns_load_common_nsg_graph(panns_index *h, const char *nsg_path);
/* PSS5.cpp:56-132 load_nsg_graph (the opt-graph loader): the same file and
 * checks, plus every vertex's norm (compute_norm, PSS5.cpp:293-334, computed
 * on the device); sets the distance to PANNS_DISTANCE_NORM_DOT.  Like the
 * reference, it needs the base set loaded first; unlike it, the base rows
 * stay where they are (the reference copies them into per-vertex records
 * and frees data_load_). */
int panns_load_nsg_graph(panns_index *h, const char *nsg_path);
/* In-memory form: adj[v*stride + i] for i < deg[v].  Keeps the distance. */
int panns_set_graph(panns_index *h, uint32_t width, uint32_t ep, uint32_t n, uint32_t stride,
                    const uint32_t *adj, const uint32_t *deg);
/* Switches the distance family of later searches (panns_distance); the
 * norms are computed when first needed and again after panns_set_base.
 * The learned stop and traces run with PANNS_DISTANCE_L2 only. */
int panns_set_distance(panns_index *h, int distance);
int panns_get_distance(panns_index *h, int *distance);
/* The N per-vertex norms (PSS5.cpp:293-334 compute_norm), bit-exact. */
int panns_get_norms(panns_index *h, float *norms);

/* num_v_, dimension_, num_queries_, width_, ep_, num_e_ (PSS5LG.h:32-45). */
int panns_get_info(panns_index *h, uint32_t *num_v, uint32_t *dim, uint32_t *num_queries,
                   uint32_t *width, uint32_t *ep, uint64_t *num_e);

/* PSS5LG.cpp:356-410 prepare_init_ids: L ids into init_ids. */
int panns_prepare_init_ids(panns_index *h, uint32_t L, uint32_t *init_ids);

/* One batch of queries (ids into the loaded query set) searched with the same
 * parameters and the same init_ids (n_init >= L).  Covers, per query:
 *   TOP_M   -> para_search_with_top_m_pure_large_graph (TOPMLG.cpp:2022-2207)
 *   PSS_V5  -> para_search_PSS_v5_large_graph_dist_thresh_profiling
 *              (PSS5LG.cpp:1376-1724) with num_threads_ = p->num_threads
 *              queues and X = p->M sub-search steps
 *   SCALE_M -> initialize_set_L_para then subsearch_with_top_m
 * Results are identical in ids to the reference's num_threads_ = 1 run.
 * PSS_V5 with T > 1 queues: the reference's T threads race on the visited
 * bitset and the master's last distance, so its ids vary between runs;
 * here pick_top_m_to_workers (PSS5LG.cpp:917-947) deals the master's
 * unchecked entries to the workers, then the T expansion loops
 * (PSS5LG.cpp:1584-1646) run in lockstep -- step s expands the s-th
 * candidate of worker 0, 1, ..., T-2, then of the master -- and
 * merge_all_queues_to_master (PSS5LG.cpp:811-842) folds the worker queues
 * (L entries each) into the master.  That is one schedule the reference can
 * take, and every run gives the same ids; count_merge_ counts the rounds. */
int panns_search(panns_index *h, const panns_params *p, const uint32_t *query_ids, uint32_t nq,
                 const uint32_t *init_ids, uint32_t n_init, panns_outputs *out,
                 panns_counters *cnt);

/* panns_search without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops that must not call hipMalloc, harnesses that forbid it).  Same
 * arguments, results and errors as panns_search; every device buffer the
 * search needs beyond the handle's resident base rows, queries and graph
 * (query ids, init ids, outputs, work counters, visited bitmaps and spill
 * logs) comes from one 64 MiB array in the library, one copy per GPU, held
 * until the process ends.  A batch whose buffers do not fit (with at least
 * one workgroup's N-bit bitmap) returns PANNS_ERR_ARG; within the limit the
 * number of persistent workgroups is what the rest of the array holds.
 * Queries run in the caller's order (no locality reordering).  The opt-graph
 * distance needs its norms computed first (panns_set_distance).  All calls
 * on one GPU share the array: calls that overlap on one GPU, from any handle
 * or thread, race and must not be made. */
int panns_search_noalloc(panns_index *h, const panns_params *p, const uint32_t *query_ids, uint32_t nq,
                         const uint32_t *init_ids, uint32_t n_init, panns_outputs *out, panns_counters *cnt);

/* TOPMLG.cpp:1023-1071 subsearch_with_top_m on a caller-held, already
 * initialised queue.  queue_* hold local_L entries sorted by (dist, id);
 * *queue_size must equal local_L (a full queue; the reference's callers
 * always pass one).  visited_words holds num_v bits (boost::dynamic_bitset
 * block order: bit v in word v/64); it is read and updated, and left set on
 * return as the reference leaves it.  *count_distance is incremented. */
int panns_subsearch_with_top_m(panns_index *h, uint32_t M_max, uint32_t query_id,
                               uint32_t local_L, uint32_t *queue_ids, float *queue_dists,
                               uint8_t *queue_checked, uint32_t *queue_size,
                               uint64_t *visited_words, uint64_t *count_distance);

/* Resident batches: inputs uploaded once, outputs kept in HBM until fetched,
 * so repeated runs time only the device work.  L_per_query / K_per_query may
 * be NULL (uniform p->L / p->K); otherwise each entry overrides them for its
 * query (uneven batches) and init_ids must hold max L entries.
 * A plan reads the handle's base rows, queries and graph where they were
 * when it was built: after panns_set_base, panns_set_queries, a graph load
 * or their file forms, panns_plan_run returns PANNS_ERR_STATE for plans
 * built before (create them again); fetching results of an earlier run
 * still works. */
int panns_plan_create(panns_index *h, const panns_params *p, const uint32_t *query_ids,
                      uint32_t nq, const uint32_t *L_per_query, const uint32_t *K_per_query,
                      const uint32_t *init_ids, uint32_t n_init, panns_plan **out);
/* Enqueue the batch; when sync != 0 wait for it.  *kernel_ms (nullable)
 * receives the device time of this run when sync != 0. */
int panns_plan_run(panns_plan *plan, int sync, float *kernel_ms);
/* Copy results of the last run to host (waits for it).  Per-query K is the
 * plan's max K: row q holds K_q valid ids followed by 0xFFFFFFFF. */
int panns_plan_fetch(panns_plan *plan, panns_outputs *out, panns_counters *cnt);
int panns_plan_destroy(panns_plan *plan);
/* Launch geometry chosen for the plan: persistent workgroups, the LDS
 * bytes each one uses and its visited-table slots (0: no table). */
int panns_plan_describe(panns_plan *plan, uint32_t *workgroups, uint32_t *lds_bytes,
                        uint32_t *hash_slots, uint32_t *max_L);

/* Uneven batches (no reference counterpart; the reference runs one query
 * per call).  Queries are grouped by L into queue-size buckets, each bucket
 * one persistent launch sized for its own largest L (LDS, visited table,
 * occupancy), its queries taken longest-first from its device work counter.
 * The buckets run one after another on the handle's stream or, with
 * `concurrent`, side by side on streams of their own; with `use_graph` the
 * whole batch (counter resets and launches) is captured once as a hipGraph
 * and replayed by panns_plan_run.  Results come back in the caller's order
 * and equal those of a single-bucket plan. */
#define PANNS_MAX_BUCKETS 8
typedef struct {
    uint32_t n_buckets;                        /* 1..PANNS_MAX_BUCKETS (0 means 1) */
    uint32_t bucket_max_L[PANNS_MAX_BUCKETS];  /* ascending upper bounds of L; all 0: split the
                                                  queries into buckets of equal summed L */
    uint32_t table_eighths;                    /* visited tables at this many eighths (1-8) of the
                                                  automatic size, for every bucket; 0: each bucket's
                                                  own by its largest L (6/8 up to L = 200, 7/8 to 512,
                                                  6/8 to 800, 3/8 above; 256 overflow entries up to
                                                  L = 512, overflow_slots ignored) */
    uint32_t use_graph;                        /* capture and replay as one hipGraph */
    uint32_t concurrent;                       /* buckets side by side */
    uint32_t want_queue;                       /* keep the final queues for panns_plan_fetch */
    uint32_t waves;                            /* 0: panns_params.waves; else wavefronts per query */
    uint32_t rows_per_wave;                    /* neighbour rows a wave gathers per distance step:
                                                  32, 64 or 96 (0: 96, or 64 in the learned-stop
                                                  kernel, whose registers hold the model) */
    uint32_t overflow_slots;                   /* LDS overflow entries (full ids) for visited ids whose
                                                  buckets are full, before the global bitmap: 0 or
                                                  a power of two in [8, 4096] */
    uint32_t deferred_merge;                   /* 0: automatic (on for buckets whose largest L is above
                                                  1,000, and above 512 where its 2 KB of LDS cost no
                                                  workgroup per CU and the bucket does not run side by
                                                  side with others); 1: merge every iteration's survivors into the
                                                  queue; 2: always defer.  Deferred: survivors wait in a
                                                  sorted buffer of 256 keys and are merged into the queue
                                                  when it fills, a pick could rank outside the top L or
                                                  the stop model runs (one wave, M <= 64, no traces);
                                                  results are the same */
    uint32_t visited_table;                    /* 0 or 1: an LDS table with exact spill to the bitmap;
                                                  2: none, every visited id claimed in the bitmap.
                                                  Results are the same */
} panns_plan_options;

int panns_plan_create_ex(panns_index *h, const panns_params *p, const uint32_t *query_ids,
                         uint32_t nq, const uint32_t *L_per_query, const uint32_t *K_per_query,
                         const uint32_t *init_ids, uint32_t n_init, const panns_plan_options *opt,
                         panns_plan **out);
/* Per bucket b < *n_buckets: largest L, queries, workgroups, LDS bytes,
 * visited-table slots, 0 for none (arrays of PANNS_MAX_BUCKETS, each nullable). */
int panns_plan_describe_buckets(panns_plan *plan, uint32_t *n_buckets, uint32_t *max_L, uint32_t *nq,
                                uint32_t *workgroups, uint32_t *lds_bytes, uint32_t *hash_slots);
/* The autotuner: times candidate bucketings (1, 2, 3, 4, 6, 8 buckets of equal
 * summed L, each bucket's table by its own largest L; side by side or not;
 * then one table size for all buckets, 8/8 down to 3/8 of the automatic
 * size, each with 0, 256 or 512 overflow entries; 1, 2 or 4 waves
 * per query where the search allows it;
 * 96, 64 or 32 neighbour rows per wave and distance step;
 * always as a graph) on every `sample_stride`-th query of the batch (0: a stride that
 * leaves about 20,000 queries) and returns the fastest in *best, with its
 * bucket bounds filled in, and its device time on the sample in *best_ms
 * (nullable).  Results do not depend on the choice. */
int panns_plan_autotune(panns_index *h, const panns_params *p, const uint32_t *query_ids, uint32_t nq,
                        const uint32_t *L_per_query, const uint32_t *K_per_query,
                        const uint32_t *init_ids, uint32_t n_init, uint32_t sample_stride,
                        panns_plan_options *best, float *best_ms);

/* Exact k-NN by brute force on the GPU with the reference's distance and
 * (distance, id) order: the semantics of tools/save_true_nearest_neighbors.cpp:16-35
 * (get_true_NN).  source 0: the base rows are the queries and each row's own
 * id is skipped (the k-NN lists a graph is built from); source 1: the loaded
 * queries (ground truth).  k <= 128.  out_* hold (number of queries) * k. */
int panns_exact_knn(panns_index *h, int source, uint32_t k, uint32_t *out_ids, float *out_dists);

typedef enum {
    PANNS_KNN_AUTO = 0,        /* MFMA from 8,192 base rows, EXACT_VALU below */
    PANNS_KNN_EXACT_VALU = 1,  /* the exact distance chain for every pair */
    PANNS_KNN_MFMA = 2         /* FP32 MFMA filter with a proven bound, exact re-rank of the candidates */
} panns_knn_method;

/* panns_exact_knn with a choice of method; all give the same ids and
 * distances.  MFMA: the exact chain over a sample of the rows (the first
 * max(4k + 64, min(4096, n / 8))) bounds each query's k-th distance; FP32
 * matrix instructions then estimate |v|^2 - 2 q.v for the rows of a larger
 * prefix and keep the rows an fp32 error bound cannot exclude from that
 * distance; their distances are recomputed with the reference's chain and
 * the k smallest kept, which bounds the next, larger prefix, up to all n
 * rows (prefixes grow by at most 32x).  Queries whose candidate list overflows (or where the sample is too
 * small to use) are recomputed with the exact chain over every row (their
 * count in *n_recomputed, nullable). */
int panns_exact_knn_ex(panns_index *h, int source, uint32_t k, int method, uint32_t *out_ids, float *out_dists,
                       uint32_t *n_recomputed);

/* Learned early termination (no reference counterpart; the reference's
 * studies of when the top-K stops changing are
 * core/Searching.202012041559.top_m_pure_large_graph.last_update_step.* and
 * ...num_unchecked.*).
 *
 * panns_search_trace: the fixed-L search of panns_search, also recording
 * after every iteration s < trace_cap the 8 stop features
 *   [iteration, top-1 dist, K-th dist, K-th/top-1, new keys that beat the
 *    queue's last key in the iteration, best initial dist, query norm, unchecked entries]
 * into feat[(q*trace_cap + s)*8 ...] and an order-free hash of the top-K id
 * set into topk_hash[q*trace_cap + s]; niter[q] receives the iteration count,
 * and the entries of steps s >= niter[q] are zero.  A step is labelled
 * "final" when its hash equals the last step's.
 *
 * panns_set_stop_model: the 8-64-64-1 MLP (ReLU, ReLU, sigmoid) evaluated by
 * one wavefront after every iteration when panns_params.stop = LEARNED; the
 * search stops once its output reaches stop_threshold.  params holds
 * PANNS_STOP_MODEL_FLOATS fp32 values: W1[64][8], b1[64], W2[64][64], b2[64],
 * w3[64], b3, 3 padding, mu[8], sd[8] (features enter as (f - mu) / sd). */
#define PANNS_STOP_MODEL_FLOATS 4820
int panns_search_trace(panns_index *h, const panns_params *p, const uint32_t *query_ids,
                       uint32_t nq, const uint32_t *init_ids, uint32_t n_init, uint32_t trace_cap,
                       float *feat, uint32_t *topk_hash, uint32_t *niter, uint32_t *out_ids);
/* panns_search_trace that also records, after every iteration s, how many
 * of the first min(K, 100) queue entries are among the first min(gt_k, 100)
 * true neighbours of the query (recall@100's count, PSS5LG.cpp:309-348):
 * hits[q*trace_cap + s], and (ndist_trace, nullable) the distance
 * computations so far.  gt_ids holds gt_k ids per loaded query (indexed
 * by query id).  A step whose count equals the last step's has reached its
 * final recall, a label that also covers queries whose top-K keeps changing
 * without gaining recall. */
int panns_search_trace_hits(panns_index *h, const panns_params *p, const uint32_t *query_ids, uint32_t nq,
                            const uint32_t *init_ids, uint32_t n_init, uint32_t trace_cap, const uint32_t *gt_ids,
                            uint32_t gt_k, float *feat, uint32_t *topk_hash, uint16_t *hits,
                            uint32_t *ndist_trace, uint32_t *niter, uint32_t *out_ids);
int panns_set_stop_model(panns_index *h, const float *params, uint32_t n_floats);

/* Training of the stop model on the GPU: BF16 MFMA forward and backward
 * kernels with FP32 accumulation, Adam on FP32 master weights, binary
 * cross-entropy.  X holds n raw feature rows (8 floats), Y the 0/1 labels;
 * features are normalised with params' mu/sd (left unchanged).  params
 * (PANNS_STOP_MODEL_FLOATS) is read as the initial model and receives the
 * trained one.  Samples are visited in a seeded random order each epoch;
 * epoch_loss[e] (nullable) gets the mean loss of epoch e. */
int panns_mlp_train(panns_index *h, const float *X, const float *Y, uint64_t n, float *params,
                    uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *epoch_loss);
/* Gradient of the mean loss over one batch of n <= 65536 samples (for
 * checks): grad receives 4801 floats in the parameter layout, *loss the mean. */
int panns_mlp_grad(panns_index *h, const float *params, const float *X, const float *Y, uint32_t n,
                   float *grad, float *loss);
/* One 16x16x16 BF16 MFMA tile, C = A * B (row-major fp32 in and out), to
 * check the fragment layout the training kernels assume. */
int panns_mfma_probe(panns_index *h, const float *A, const float *B, float *C);
/* One 16x16x4 FP32 MFMA tile, C (16x16) = A (16x4) * B (4x16), row-major,
 * to check the fragment layout panns_exact_knn_ex assumes. */
int panns_mfma_f32_probe(panns_index *h, const float *A, const float *B, float *C);

/* Measurement helpers (no reference counterpart).
 * panns_plan_phase_cycles: shader-clock cycles the last run spent per search
 * phase, summed over workgroups: [init, pick, claim, distance, merge, stop,
 * output, idle].  Only libraries built with PANNS_PHASE_TIMERS record them;
 * others return PANNS_ERR_UNSUPPORTED.
 * panns_device_copy: copies `bytes` (a multiple of 16) between two device
 * buffers with a plain streaming kernel, so hardware byte counters can be
 * scaled against a kernel of known traffic; *ms gets its device time. */
int panns_plan_phase_cycles(panns_plan *plan, uint64_t *out8);
int panns_device_copy(panns_index *h, uint64_t bytes, float *ms);

#ifdef __cplusplus
}
#endif

#endif
