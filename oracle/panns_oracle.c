/*
 * CPU oracle for the NSG top-M / PSS_v5 search path -- TEST INFRASTRUCTURE.
 * See panns_oracle.h for scope, citations and how each function is pinned.
 * Built with -ffp-contract=off so that the only fused operations are the
 * explicit fmaf() calls that mirror the reference's compiled FMA chain.
 */
#include "panns_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline int bit_get(const uint64_t *bits, uint32_t i) { return (int)((bits[i >> 6] >> (i & 63)) & 1u); }
static inline void bit_set(uint64_t *bits, uint32_t i) { bits[i >> 6] |= (uint64_t)1 << (i & 63); }

/* The 8-accumulator chain shared by the reference's three AVX kernels
 * (PSS5LG.cpp:525-560 compute_distance, PSS5.cpp:293-334 compute_norm,
 * PSS5.cpp:336-390 compute_distance_with_norm).  D8 = (D+7)&~7,
 * DR = D8 % 16, DD = D8 - DR.  The tail block [DD, DD+8) is added first when
 * DR != 0, then blocks of 16 as two 8-wide halves; lane j accumulates with
 * one fma per element:  L2: fma(v-q, v-q, acc)   DOT: fma(v, q, acc)
 * NORM: fma(v, v, acc).  Horizontal sum left to right. */
enum { CH_L2 = 0, CH_DOT = 1, CH_NORM = 2 };

static inline float chain_term(float a, float b, float acc, int kind)
{
    if (kind == CH_L2) {
        float t = a - b;
        return fmaf(t, t, acc);
    }
    if (kind == CH_DOT) return fmaf(a, b, acc);
    return fmaf(a, a, acc);
}

static float chain8(const float *v, const float *q, uint32_t dim, int kind)
{
    const uint32_t D8 = (dim + 7u) & ~7u;
    const uint32_t DR = D8 % 16u;
    const uint32_t DD = D8 - DR;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (DR) {
        for (int j = 0; j < 8; ++j) acc[j] = chain_term(v[DD + j], q[DD + j], acc[j], kind);
    }
    for (uint32_t i = 0; i < DD; i += 16) {
        for (int j = 0; j < 8; ++j) acc[j] = chain_term(v[i + j], q[i + j], acc[j], kind);
        for (int j = 0; j < 8; ++j) acc[j] = chain_term(v[i + 8 + j], q[i + 8 + j], acc[j], kind);
    }
    float r = acc[0];
    for (int j = 1; j < 8; ++j) r = r + acc[j];
    return r;
}

/* PSS5LG.cpp:525-560 */
float po_distance(const float *v, const float *q, uint32_t dim) { return chain8(v, q, dim, CH_L2); }

/* PSS5.cpp:293-334 */
float po_norm(const float *v, uint32_t dim) { return chain8(v, v, dim, CH_NORM); }

/* PSS5.cpp:336-390: result = -2 * dot + vertex_norm (one rounding whether
 * or not it is fused, -2 * dot being exact). */
float po_distance_with_norm(const float *v, const float *q, uint32_t dim, float norm)
{
    return fmaf(-2.0f, chain8(v, q, dim, CH_DOT), norm);
}

/* Distance of base row v to the query under the graph's layout: the CSR
 * (large-graph) family uses squared L2, the opt-graph family, whose vertex
 * records carry the norm (PSS5.cpp:86-91), uses norm - 2 * dot. */
static inline float vdist(const float *norms, const float *base, uint32_t dim, uint32_t v,
                          const float *query)
{
    const float *row = base + (size_t)v * dim;
    return norms ? po_distance_with_norm(row, query, dim, norms[v]) : po_distance(row, query, dim);
}

int po_cand_less(const po_cand *a, const po_cand *b)
{
    if (a->dist != b->dist) return a->dist < b->dist;
    return a->id < b->id;
}

/* std::lower_bound over [lo, hi) with Candidate::operator<. */
static uint32_t lower_bound_cand(const po_cand *q, uint32_t lo, uint32_t hi, const po_cand *c)
{
    uint32_t count = hi - lo;
    while (count > 0) {
        uint32_t step = count / 2;
        uint32_t it = lo + step;
        if (po_cand_less(&q[it], c)) {
            lo = it + 1;
            count -= step + 1;
        } else {
            count = step;
        }
    }
    return lo;
}

/* PSS5LG.cpp:565-608 */
uint32_t po_add_into_queue(po_cand *queue, uint32_t queue_start, uint32_t *queue_size,
                           uint32_t queue_capacity, po_cand cand)
{
    if (*queue_size == 0) {
        queue[queue_start + (*queue_size)++] = cand;
        return 0;
    }
    uint32_t queue_end = queue_start + *queue_size;
    uint32_t insert_loc = lower_bound_cand(queue, queue_start, queue_end, &cand);
    if (insert_loc != queue_end) {
        if (cand.id == queue[insert_loc].id) return queue_capacity; /* duplicate */
        if (*queue_size >= queue_capacity) { /* full: drop the last */
            --*queue_size;
            --queue_end;
        }
    } else {
        if (*queue_size < queue_capacity) {
            queue[insert_loc] = cand;
            ++*queue_size;
            return *queue_size - 1;
        }
        return queue_capacity;
    }
    memmove(queue + insert_loc + 1, queue + insert_loc, (size_t)(queue_end - insert_loc) * sizeof(po_cand));
    queue[insert_loc] = cand;
    ++*queue_size;
    return insert_loc - queue_start;
}

/* PSS5LG.cpp:356-410: ep's neighbours in adjacency order (de-duplicated),
 * then ep+1, ep+2, ... wrapping at num_v, skipping ids already chosen. */
int po_prepare_init_ids(const po_graph *g, uint32_t L, uint32_t *init_ids)
{
    if (L > g->num_v) return -1;
    uint64_t *sel = (uint64_t *)calloc((g->num_v + 63) / 64, sizeof(uint64_t));
    if (!sel) return -2;
    const uint32_t *out_edges = g->deg_ngbrs + g->vertex_base[g->ep];
    uint32_t out_degree = *out_edges++;
    uint32_t end = 0;
    for (uint32_t e = 0; e < out_degree && end < L; ++e) {
        uint32_t v = out_edges[e];
        if (bit_get(sel, v)) continue;
        bit_set(sel, v);
        init_ids[end++] = v;
    }
    uint32_t tmp = g->ep + 1;
    while (end < L) {
        if (tmp == g->num_v) tmp = 0;
        uint32_t v = tmp++;
        if (bit_get(sel, v)) continue;
        bit_set(sel, v);
        init_ids[end++] = v;
    }
    free(sel);
    return 0;
}

static int cmp_cand(const void *a, const void *b)
{
    const po_cand *x = (const po_cand *)a, *y = (const po_cand *)b;
    if (po_cand_less(x, y)) return -1;
    if (po_cand_less(y, x)) return 1;
    return 0;
}

/* PSS5LG.cpp:1055-1096 (PSS5.cpp:890-927 with norms) */
void po_initialize_set_L(const float *base, uint32_t dim, const float *query, uint32_t L,
                         po_cand *set_L, uint32_t set_L_start, uint32_t *set_L_size,
                         const uint32_t *init_ids, uint64_t *visited, uint64_t *count_distance,
                         const float *norms)
{
    for (uint32_t c = 0; c < L; ++c) bit_set(visited, init_ids[c]);
    for (uint32_t i = 0; i < L; ++i) {
        uint32_t v = init_ids[i];
        po_cand c;
        memset(&c, 0, sizeof(c));
        c.id = v;
        c.dist = vdist(norms, base, dim, v, query);
        c.checked = 0;
        set_L[set_L_start + i] = c;
    }
    *count_distance += L;
    qsort(set_L + set_L_start, L, sizeof(po_cand), cmp_cand);
    *set_L_size = L;
}

/* TOPMLG.cpp:1262-1330 (and PSS5LG.cpp:984-1053 whose extra dist_thresh
 * test is implied by the live bound, which never increases).  dist_bound
 * points into the queue so that it is read live, as the reference's
 * `const distf &` parameter is.  Returns the lowest insert index. */
static uint32_t expand_one_candidate(const po_graph *g, const float *base, uint32_t dim,
                                     uint32_t cand_id, const float *query,
                                     const float *dist_bound, po_cand *set_L,
                                     uint32_t queue_start, uint32_t *queue_size,
                                     uint32_t queue_capacity, uint64_t *visited,
                                     uint64_t *count_distance)
{
    const uint32_t *out_edges = g->deg_ngbrs + g->vertex_base[cand_id];
    uint32_t out_degree = *out_edges++;
    uint32_t nk = queue_capacity;
    uint64_t cnt = 0;
    for (uint32_t e = 0; e < out_degree; ++e) {
        uint32_t nb = out_edges[e];
        if (bit_get(visited, nb)) continue;
        bit_set(visited, nb);
        ++cnt;
        float d = vdist(g->norms, base, dim, nb, query);
        if (d > *dist_bound) continue;
        po_cand c;
        memset(&c, 0, sizeof(c));
        c.id = nb;
        c.dist = d;
        c.checked = 0;
        uint32_t r = po_add_into_queue(set_L, queue_start, queue_size, queue_capacity, c);
        if (r < nk) nk = r;
    }
    *count_distance += cnt;
    return nk;
}

/* TOPMLG.cpp:1172-1198 */
static void pick_top_m_unchecked(uint32_t M, uint32_t k_uc, po_cand *set_L, uint32_t start,
                                 uint32_t size, uint32_t *top_m, uint32_t *top_m_size,
                                 uint32_t *last_k)
{
    uint32_t tmp_last_k = size;
    uint32_t n = 0;
    for (uint32_t c = start + k_uc; c < start + size && n < M; ++c) {
        if (set_L[c].checked) continue;
        tmp_last_k = c - start;
        set_L[c].checked = 1;
        top_m[n++] = set_L[c].id;
    }
    *last_k = tmp_last_k;
    *top_m_size = n;
}

/* TOPMLG.cpp:2022-2207 at num_threads_ = 1: the only queue is the master,
 * so merge_all_queues_to_master (PSS5LG.cpp:811-842) merges nothing. */
void po_search_top_m(const po_graph *g, const float *base, uint32_t dim, const float *query,
                     uint32_t K, uint32_t L, uint32_t M, const uint32_t *init_ids,
                     po_cand *set_L, uint32_t *set_K, uint64_t *visited, po_counters *cnt)
{
    uint32_t size = 0;
    uint64_t ndist = 0;
    po_initialize_set_L(base, dim, query, L, set_L, 0, &size, init_ids, visited, &ndist, g->norms);
    const float *last_dist = &set_L[L - 1].dist;
    uint32_t *top_m = (uint32_t *)malloc(sizeof(uint32_t) * (M ? M : 1));
    uint32_t k = 0, last_k = 0, n_top = 0;
    uint64_t merges = 0, exps = 0;
    while (1) {
        last_k = L;
        pick_top_m_unchecked(M, k, set_L, 0, size, top_m, &n_top, &last_k);
        if (!n_top) break;
        uint32_t nk = L;
        for (uint32_t c = 0; c < n_top; ++c) {
            uint32_t r = expand_one_candidate(g, base, dim, top_m[c], query, last_dist, set_L, 0,
                                              &size, L, visited, &ndist);
            if (r < nk) nk = r;
        }
        exps += n_top;
        ++merges;
        /* merge_all_queues_to_master returns L here (no worker queues) */
        k = (nk <= last_k) ? nk : last_k + 1;
    }
    for (uint32_t i = 0; i < K; ++i) set_K[i] = set_L[i].id;
    memset(visited, 0, sizeof(uint64_t) * ((g->num_v + 63) / 64));
    free(top_m);
    if (cnt) {
        cnt->distance_computations += ndist;
        cnt->merges += merges;
        cnt->expansions += exps;
    }
}

/* PSS5LG.cpp:1376-1724 at num_threads_ = 1.  pick_top_m_to_workers
 * (PSS5LG.cpp:917-947) then only counts unchecked entries (there is no
 * worker queue), the master runs up to X expansions from k_master, and the
 * merge loop (PSS5LG.cpp:821) is empty. */
void po_search_pss_v5_t1(const po_graph *g, const float *base, uint32_t dim, const float *query,
                         uint32_t K, uint32_t L, uint32_t X, const uint32_t *init_ids,
                         po_cand *set_L, uint32_t *set_K, uint64_t *visited, po_counters *cnt)
{
    uint32_t size = 0;
    uint64_t ndist = 0, merges = 0, exps = 0;
    po_initialize_set_L(base, dim, query, L, set_L, 0, &size, init_ids, visited, &ndist, g->norms);
    const float *last_dist = &set_L[size - 1].dist;
    uint32_t k_master = 0;
    int no_need = 0;
    /* sequential start: one step (seq_iter_bound = 1), PSS5LG.cpp:1478-1534 */
    if (k_master == L) {
        no_need = 1;
    } else {
        uint32_t r;
        po_cand *c = &set_L[k_master];
        if (!c->checked) {
            c->checked = 1;
            r = expand_one_candidate(g, base, dim, c->id, query, last_dist, set_L, 0, &size, L,
                                     visited, &ndist);
            ++exps;
        } else {
            r = L;
        }
        k_master = (r <= k_master) ? r : k_master + 1;
    }
    while (!no_need) {
        /* pick_top_m_to_workers: counts unchecked from k_master */
        uint32_t count = 0;
        for (uint32_t c = k_master; c < size; ++c)
            if (!set_L[c].checked) ++count;
        if (!count) break;
        /* master expands up to X candidates, PSS5LG.cpp:1584-1646 */
        uint32_t k_uc = k_master, it = 0;
        while (it < X && k_uc < size) {
            po_cand *c = &set_L[k_uc];
            if (!c->checked) {
                c->checked = 1;
                ++it;
                uint32_t r = expand_one_candidate(g, base, dim, c->id, query, last_dist, set_L, 0,
                                                  &size, L, visited, &ndist);
                ++exps;
                k_uc = (r <= k_uc) ? r : k_uc + 1;
            } else {
                ++k_uc;
            }
        }
        k_master = k_uc;
        ++merges;
        /* merge_all_queues_to_master returns L > k_master: no change */
    }
    for (uint32_t i = 0; i < K; ++i) set_K[i] = set_L[i].id;
    memset(visited, 0, sizeof(uint64_t) * ((g->num_v + 63) / 64));
    if (cnt) {
        cnt->distance_computations += ndist;
        cnt->merges += merges;
        cnt->expansions += exps;
    }
}

/* PSS5LG.cpp:630-643 insert_one_element_at: shift [insert_index, size-1)
 * right by one (the last element falls off) and place cand. */
static void insert_one_element_at(po_cand cand, po_cand *queue, uint32_t insert_index, uint32_t queue_start,
                                  uint32_t queue_size)
{
    const uint32_t dest = queue_start + insert_index;
    memmove(queue + dest + 1, queue + dest, (size_t)(queue_size - insert_index - 1) * sizeof(po_cand));
    queue[dest] = cand;
}

/* PSS5LG.cpp:611-628 add_into_queue_at: insert at insert_index, growing the
 * queue unless it is at queue_length (then the last element falls off). */
static void add_into_queue_at(po_cand cand, po_cand *queue, uint32_t insert_index, uint32_t queue_start,
                              uint32_t *queue_size, uint32_t queue_length)
{
    const uint32_t dest = queue_start + insert_index;
    if (*queue_size == queue_length) --*queue_size;
    memmove(queue + dest + 1, queue + dest, (size_t)(*queue_size - insert_index) * sizeof(po_cand));
    queue[dest] = cand;
    ++*queue_size;
}

/* PSS5LG.cpp:648-719 merge_two_queues_into_1st_queue_seq_fixed. */
uint32_t po_merge_two_queues_fixed(po_cand *queue1, uint32_t queue1_start, uint32_t queue1_size,
                                   po_cand *queue2, uint32_t queue2_start, uint32_t queue2_size)
{
    const uint32_t it_loc = lower_bound_cand(queue1, queue1_start, queue1_start + queue1_size,
                                             &queue2[queue2_start]);
    const uint32_t insert_index = it_loc - queue1_start;
    if (insert_index == queue1_size) return insert_index;
    if (insert_index == queue1_size - 1) {
        queue1[queue1_start + insert_index] = queue2[queue2_start];
        return insert_index;
    }
    /* the first element of queue2 */
    if (queue2[queue2_start].id != queue1[it_loc].id) {
        insert_one_element_at(queue2[queue2_start], queue1, insert_index, queue1_start, queue1_size);
    } else if (!queue2[queue2_start].checked && queue1[it_loc].checked) {
        queue1[it_loc].checked = 0;
    }
    if (queue2_size == 1) return insert_index;
    uint32_t q_i_1 = insert_index + 1 + queue1_start;
    uint32_t q_i_2 = queue2_start + 1;
    const uint32_t q_i_1_bound = queue1_start + queue1_size;
    const uint32_t q_i_2_bound = queue2_start + queue2_size;
    for (uint32_t insert_i = insert_index + 1; insert_i < queue1_size; ++insert_i) {
        if (q_i_1 >= q_i_1_bound || q_i_2 >= q_i_2_bound) {
            break;
        } else if (po_cand_less(&queue1[q_i_1], &queue2[q_i_2])) {
            ++q_i_1;
        } else if (po_cand_less(&queue2[q_i_2], &queue1[q_i_1])) {
            insert_one_element_at(queue2[q_i_2++], queue1, insert_i, queue1_start, queue1_size);
            ++q_i_1;
        } else {
            /* duplicate: the first queue keeps its copy, unchecked if the
             * second queue's copy is */
            if (!queue2[q_i_2].checked && queue1[q_i_1].checked) queue1[q_i_1].checked = 0;
            ++q_i_2;
            ++q_i_1;
        }
    }
    return insert_index;
}

/* PSS5LG.cpp:725-805 merge_two_queues_into_1st_queue_seq_incr (the merge of
 * the reference's tree-shaped merge_all_queues_to_master, kept at
 * PSS5LG.cpp:844-909 as a backup; the live merge uses the fixed form). */
uint32_t po_merge_two_queues_incr(po_cand *queue1, uint32_t queue1_start, uint32_t *queue1_size,
                                  uint32_t queue1_length, po_cand *queue2, uint32_t queue2_start,
                                  uint32_t queue2_size)
{
    const uint32_t it_loc = lower_bound_cand(queue1, queue1_start, queue1_start + *queue1_size,
                                             &queue2[queue2_start]);
    const uint32_t insert_index = it_loc - queue1_start;
    if (insert_index == *queue1_size) {
        const uint32_t copy_count = (*queue1_size + queue2_size > queue1_length) ? queue1_length - *queue1_size
                                                                                 : queue2_size;
        memmove(queue1 + queue1_start + *queue1_size, queue2 + queue2_start, (size_t)copy_count * sizeof(po_cand));
        *queue1_size += copy_count;
        return insert_index;
    }
    if (queue2[queue2_start].id != queue1[it_loc].id) {
        add_into_queue_at(queue2[queue2_start], queue1, insert_index, queue1_start, queue1_size, queue1_length);
    } else if (!queue2[queue2_start].checked && queue1[it_loc].checked) {
        queue1[it_loc].checked = 0;
    }
    if (queue2_size == 1) return insert_index;
    uint32_t q_i_1 = insert_index + 1 + queue1_start;
    uint32_t q_i_2 = queue2_start + 1;
    uint32_t q_i_1_bound = queue1_start + *queue1_size;
    const uint32_t q_i_2_bound = queue2_start + queue2_size;
    for (uint32_t insert_i = insert_index + 1; insert_i < queue1_length; ++insert_i) {
        if (q_i_1 >= q_i_1_bound) {
            const uint32_t a = queue1_length - insert_i, b = q_i_2_bound - q_i_2;
            *queue1_size += a < b ? a : b;
            for (; insert_i < *queue1_size; ++insert_i) queue1[queue1_start + insert_i] = queue2[q_i_2++];
            break;
        } else if (q_i_2 >= q_i_2_bound) {
            break;
        } else if (po_cand_less(&queue1[q_i_1], &queue2[q_i_2])) {
            ++q_i_1;
        } else if (po_cand_less(&queue2[q_i_2], &queue1[q_i_1])) {
            add_into_queue_at(queue2[q_i_2++], queue1, insert_i, queue1_start, queue1_size, queue1_length);
            ++q_i_1;
            q_i_1_bound = queue1_start + *queue1_size;
        } else {
            if (!queue2[q_i_2].checked && queue1[q_i_1].checked) queue1[q_i_1].checked = 0;
            ++q_i_2;
            ++q_i_1;
        }
    }
    return insert_index;
}

/* PSS5LG.cpp:811-842 merge_all_queues_to_master: worker queues 0..T-2 in
 * order into the master (the last queue, fixed length L); returns the
 * lowest insert index (L when nothing was merged); worker sizes reset. */
static uint32_t merge_all_queues_to_master(po_cand *set_L, const uint32_t *starts, uint32_t *sizes, uint32_t T,
                                           uint32_t L)
{
    uint32_t nk = L;
    const uint32_t master_start = starts[T - 1];
    for (uint32_t w = 0; w + 1 < T; ++w) {
        if (sizes[w] == 0) continue;
        const uint32_t r = po_merge_two_queues_fixed(set_L, master_start, L, set_L, starts[w], sizes[w]);
        if (r < nk) nk = r;
        sizes[w] = 0;
    }
    return nk;
}

/* PSS5LG.cpp:917-947 pick_top_m_to_workers: every unchecked master entry
 * from k_uc on, dealt round-robin to workers 0..T-2 with every T-th left
 * to the master; a dealt entry is copied (unchecked) and marked checked in
 * the master.  Stops when a worker queue reaches local_capacity.  Returns
 * the number of unchecked entries seen. */
static uint32_t pick_top_m_to_workers(po_cand *set_L, const uint32_t *starts, uint32_t *sizes, uint32_t T,
                                      uint32_t local_capacity, uint32_t k_uc)
{
    const uint32_t last_start = starts[T - 1];
    const uint32_t c_bound = last_start + sizes[T - 1];
    uint32_t count = 0, dest = 0;
    for (uint32_t c = k_uc + last_start; c < c_bound; ++c) {
        if (set_L[c].checked) continue;
        ++count;
        if (T - 1 != dest) {
            set_L[starts[dest] + sizes[dest]++] = set_L[c];
            set_L[c].checked = 1;
            if (sizes[dest] == local_capacity) break;
            ++dest;
        } else {
            dest = 0;
        }
    }
    return count;
}

/* PSS5LG.cpp:1376-1724 with T = num_threads_ queues.  The reference runs
 * the T expansion loops (PSS5LG.cpp:1584-1646) as OpenMP threads that race
 * on the shared visited bitset and on the master's L-th distance; here they
 * run in lockstep in a fixed order, which is one schedule the reference can
 * take: step s = 1..X expands the s-th candidate of worker 0, then worker
 * 1, ..., then the master (queue T-1), each to completion, so a neighbour
 * reached by two queues in one step goes to the lower-numbered one, and every
 * worker's bound is the master's L-th distance as the previous step left it.
 * set_L holds (T-1)*local_L + L entries: worker w's queue at w*local_L, the
 * master at (T-1)*local_L (APP_PSS:136-141 uses local_L = L).  At T = 1 this
 * is po_search_pss_v5_t1. */
void po_search_pss_v5(const po_graph *g, const float *base, uint32_t dim, const float *query, uint32_t K,
                      uint32_t L, uint32_t X, uint32_t T, uint32_t local_L, const uint32_t *init_ids, po_cand *set_L,
                      uint32_t *set_K, uint64_t *visited, po_counters *cnt)
{
    uint32_t *starts = (uint32_t *)malloc(sizeof(uint32_t) * T);
    uint32_t *sizes = (uint32_t *)calloc(T, sizeof(uint32_t));
    uint32_t *kuc = (uint32_t *)malloc(sizeof(uint32_t) * T);
    for (uint32_t w = 0; w < T; ++w) starts[w] = w * local_L;
    const uint32_t ms = starts[T - 1];
    uint64_t ndist = 0, merges = 0, exps = 0;
    po_initialize_set_L(base, dim, query, L, set_L, ms, &sizes[T - 1], init_ids, visited, &ndist, g->norms);
    /* the live bound: the master's last entry (PSS5LG.cpp:1412) */
    const float *last_dist = &set_L[ms + sizes[T - 1] - 1].dist;
    uint32_t k_master = 0;
    int no_need = 0;
    /* sequential start: one master step (PSS5LG.cpp:1478-1534) */
    if (k_master == L) {
        no_need = 1;
    } else {
        uint32_t r;
        po_cand *c = &set_L[ms + k_master];
        if (!c->checked) {
            c->checked = 1;
            r = expand_one_candidate(g, base, dim, c->id, query, last_dist, set_L, ms, &sizes[T - 1], L, visited,
                                     &ndist);
            ++exps;
        } else {
            r = L;
        }
        k_master = (r <= k_master) ? r : k_master + 1;
    }
    while (!no_need) {
        if (!pick_top_m_to_workers(set_L, starts, sizes, T, local_L, k_master)) break;
        /* the expansion loops (PSS5LG.cpp:1584-1646), in lockstep */
        for (uint32_t w = 0; w < T; ++w) kuc[w] = (w == T - 1) ? k_master : 0;
        for (uint32_t step = 0; step < X; ++step) {
            int any = 0;
            for (uint32_t w = 0; w < T; ++w) {
                const uint32_t st = starts[w];
                const uint32_t cap = (w == T - 1) ? L : local_L;
                while (kuc[w] < sizes[w] && set_L[st + kuc[w]].checked) ++kuc[w];
                if (kuc[w] >= sizes[w]) continue;
                set_L[st + kuc[w]].checked = 1;
                const uint32_t id = set_L[st + kuc[w]].id;
                const uint32_t r = expand_one_candidate(g, base, dim, id, query, last_dist, set_L, st, &sizes[w],
                                                        cap, visited, &ndist);
                ++exps;
                kuc[w] = (r <= kuc[w]) ? r : kuc[w] + 1;
                any = 1;
            }
            if (!any) break;
        }
        k_master = kuc[T - 1];
        /* PSS5LG.cpp:1655-1667 */
        ++merges;
        const uint32_t r = merge_all_queues_to_master(set_L, starts, sizes, T, L);
        if (r <= k_master) k_master = r;
    }
    for (uint32_t i = 0; i < K; ++i) set_K[i] = set_L[ms + i].id;
    memset(visited, 0, sizeof(uint64_t) * ((g->num_v + 63) / 64));
    free(starts);
    free(sizes);
    free(kuc);
    if (cnt) {
        cnt->distance_computations += ndist;
        cnt->merges += merges;
        cnt->expansions += exps;
    }
}

/* ---- the reference design for the CPU baseline (timed, not a parity
 * function): PSS_v5 with its T loops running at once as OpenMP threads.
 *
 * Per query as PSS5LG.cpp:1376-1724: initialise the master queue (the L
 * init distances in an OpenMP loop, :1055-1096), one sequential master
 * step, then iterations of pick_top_m_to_workers, a parallel region in
 * which thread w runs up to X expansions on its own queue (the master,
 * w = T-1, on the master queue; :1584-1646), and merge_all_queues_to_master.
 * The reference's shared state is made race-free without changing what is
 * shared: the visited bitset is claimed with an atomic fetch-or (the
 * reference's unsynchronised read-modify-write of bitset blocks,
 * :1017-1020), and the master's L-th distance, which the workers read live
 * through `const distf &last_dist` (:1412), is published by the master
 * thread after each of its insertions and read with relaxed atomics.  At
 * T > 1 the result depends on the schedule, as the reference's does. */
static inline float load_bound(const uint32_t *bits)
{
    uint32_t b = __atomic_load_n(bits, __ATOMIC_RELAXED);
    float f;
    memcpy(&f, &b, sizeof(f));
    return f;
}

static uint32_t expand_one_candidate_par(const po_graph *g, const float *base, uint32_t dim, uint32_t cand_id,
                                         const float *query, int is_master, uint32_t *bound_bits,
                                         float dist_thresh, po_cand *set_L, uint32_t queue_start,
                                         uint32_t *queue_size, uint32_t queue_capacity, uint64_t *visited,
                                         uint64_t *count_distance)
{
    const uint32_t *out_edges = g->deg_ngbrs + g->vertex_base[cand_id];
    uint32_t out_degree = *out_edges++;
    uint32_t nk = queue_capacity;
    uint64_t cnt = 0;
    for (uint32_t e = 0; e < out_degree; ++e) {
        const uint32_t nb = out_edges[e];
        const uint64_t m = (uint64_t)1 << (nb & 63);
        if (__atomic_load_n(&visited[nb >> 6], __ATOMIC_RELAXED) & m) continue;
        if (__atomic_fetch_or(&visited[nb >> 6], m, __ATOMIC_RELAXED) & m) continue;
        ++cnt;
        const float d = vdist(g->norms, base, dim, nb, query);
        const float bound = is_master ? set_L[queue_start + *queue_size - 1].dist : load_bound(bound_bits);
        if (d > bound || d > dist_thresh) continue;
        po_cand c;
        memset(&c, 0, sizeof(c));
        c.id = nb;
        c.dist = d;
        c.checked = 0;
        const uint32_t r = po_add_into_queue(set_L, queue_start, queue_size, queue_capacity, c);
        if (r < nk) nk = r;
        if (is_master && r < queue_capacity) {
            uint32_t b;
            memcpy(&b, &set_L[queue_start + *queue_size - 1].dist, sizeof(b));
            __atomic_store_n(bound_bits, b, __ATOMIC_RELAXED);
        }
    }
    *count_distance += cnt;
    return nk;
}

uint64_t po_search_pss_v5_par(const po_graph *g, const float *base, uint32_t dim, const float *queries,
                              uint32_t nq, uint32_t K, uint32_t L, uint32_t X, uint32_t T, uint32_t local_L,
                              uint32_t *out_ids, uint32_t *out_ndist)
{
    const size_t words = (g->num_v + 63) / 64;
    po_cand *set_L = (po_cand *)malloc(sizeof(po_cand) * ((size_t)(T - 1) * local_L + L));
    uint64_t *visited = (uint64_t *)calloc(words, sizeof(uint64_t));
    uint32_t *starts = (uint32_t *)malloc(sizeof(uint32_t) * T);
    uint32_t *sizes = (uint32_t *)calloc(T, sizeof(uint32_t));
    uint32_t *init_ids = (uint32_t *)malloc(sizeof(uint32_t) * L);
    for (uint32_t w = 0; w < T; ++w) starts[w] = w * local_L;
    const uint32_t ms = starts[T - 1];
    uint64_t total = 0;
    po_prepare_init_ids(g, L, init_ids);
    for (uint32_t qi = 0; qi < nq; ++qi) {
        const float *query = queries + (size_t)qi * dim;
        uint64_t ndist = L;
        for (uint32_t c = 0; c < L; ++c) bit_set(visited, init_ids[c]);
#ifdef _OPENMP
#pragma omp parallel for num_threads(T)
#endif
        for (int64_t i = 0; i < (int64_t)L; ++i) {
            po_cand c;
            memset(&c, 0, sizeof(c));
            c.id = init_ids[i];
            c.dist = vdist(g->norms, base, dim, c.id, query);
            set_L[ms + i] = c;
        }
        qsort(set_L + ms, L, sizeof(po_cand), cmp_cand);
        for (uint32_t w = 0; w + 1 < T; ++w) sizes[w] = 0;
        sizes[T - 1] = L;
        uint32_t bound_bits;
        memcpy(&bound_bits, &set_L[ms + L - 1].dist, sizeof(bound_bits));
        uint32_t k_master = 0;
        {
            po_cand *c = &set_L[ms];
            c->checked = 1;
            const uint32_t r = expand_one_candidate_par(g, base, dim, c->id, query, 1, &bound_bits,
                                                        set_L[ms + L - 1].dist, set_L, ms, &sizes[T - 1], L,
                                                        visited, &ndist);
            k_master = r <= k_master ? r : k_master + 1;
        }
        while (pick_top_m_to_workers(set_L, starts, sizes, T, local_L, k_master)) {
            const float dist_thresh = load_bound(&bound_bits);
            uint32_t km = k_master;
#ifdef _OPENMP
#pragma omp parallel num_threads(T) reduction(+ : ndist)
#endif
            {
#ifdef _OPENMP
                const uint32_t w = (uint32_t)omp_get_thread_num();
#else
                for (uint32_t w = 0; w < T; ++w)
#endif
                {
                    const int master = w == T - 1;
                    const uint32_t st = starts[w], cap = master ? L : local_L;
                    uint32_t k_uc = master ? km : 0u, it = 0;
                    uint64_t nd = 0;
                    while (it < X && k_uc < sizes[w]) {
                        po_cand *c = &set_L[st + k_uc];
                        if (c->checked) {
                            ++k_uc;
                            continue;
                        }
                        c->checked = 1;
                        ++it;
                        const uint32_t r = expand_one_candidate_par(g, base, dim, c->id, query, master, &bound_bits,
                                                                    dist_thresh, set_L, st, &sizes[w], cap, visited,
                                                                    &nd);
                        k_uc = r <= k_uc ? r : k_uc + 1;
                    }
                    if (master) km = k_uc;
                    ndist += nd;
                }
            }
            k_master = km;
            const uint32_t r = merge_all_queues_to_master(set_L, starts, sizes, T, L);
            if (r <= k_master) k_master = r;
            memcpy(&bound_bits, &set_L[ms + L - 1].dist, sizeof(bound_bits));
        }
        for (uint32_t i = 0; i < K; ++i) out_ids[(size_t)qi * K + i] = set_L[ms + i].id;
        if (out_ndist) out_ndist[qi] = (uint32_t)ndist;
        total += ndist;
        memset(visited, 0, sizeof(uint64_t) * words);
    }
    free(set_L);
    free(visited);
    free(starts);
    free(sizes);
    free(init_ids);
    return total;
}

/* The compiler that built this library (its __VERSION__ string). */
const char *po_build_info(void) { return __VERSION__; }

/* TOPMLG.cpp:1077-1131: the bound passed to expand_one_candidate is the
 * element at index set_L_size-1 at the time of each call (a live reference). */
static void subsearch_one_iteration(const po_graph *g, const float *base, uint32_t dim,
                                    const float *query, uint32_t *k_uc, uint32_t M, uint32_t L,
                                    po_cand *set_L, uint32_t start, uint32_t *size,
                                    uint32_t *top_m, uint64_t *visited, uint64_t *ndist,
                                    uint64_t *exps)
{
    uint32_t n_top = 0, last_k = L;
    pick_top_m_unchecked(M, *k_uc, set_L, start, *size, top_m, &n_top, &last_k);
    uint32_t nk = L;
    for (uint32_t c = 0; c < n_top; ++c) {
        const float *bound = &set_L[*size - 1 + start].dist;
        uint32_t r = expand_one_candidate(g, base, dim, top_m[c], query, bound, set_L, start, size,
                                          L, visited, ndist);
        if (r < nk) nk = r;
    }
    *exps += n_top;
    *k_uc = (nk <= last_k) ? nk : last_k + 1;
}

/* TOPMLG.cpp:1023-1071: M = 1, doubled after every iteration while
 * M < value_M_max; loop while k < local_L. */
void po_subsearch_with_top_m(const po_graph *g, const float *base, uint32_t dim,
                             const float *query, uint32_t M_max, uint32_t local_L,
                             po_cand *set_L, uint32_t set_L_start, uint32_t *set_L_size,
                             uint64_t *visited, uint64_t *count_distance)
{
    uint32_t cap = 1;
    while (cap < M_max) cap <<= 1;
    uint32_t *top_m = (uint32_t *)malloc(sizeof(uint32_t) * (cap + 1));
    uint32_t k = 0, M = 1;
    uint64_t exps = 0;
    while (k < local_L) {
        subsearch_one_iteration(g, base, dim, query, &k, M, local_L, set_L, set_L_start,
                                set_L_size, top_m, visited, count_distance, &exps);
        if (M < M_max) M <<= 1;
    }
    free(top_m);
}

uint64_t po_search_batch(const po_graph *g, const float *base, uint32_t dim,
                         const float *queries, uint32_t nq, uint32_t K, uint32_t L,
                         uint32_t M, int mode, const uint32_t *init_ids, uint32_t *out_ids,
                         float *out_dists, uint32_t *out_ndist, int nthreads)
{
    uint64_t total = 0;
    const size_t words = (g->num_v + 63) / 64;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(+ : total)
#endif
    {
        po_cand *set_L = (po_cand *)malloc(sizeof(po_cand) * L);
        uint64_t *visited = (uint64_t *)calloc(words, sizeof(uint64_t));
        uint32_t *set_K = (uint32_t *)malloc(sizeof(uint32_t) * K);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
            po_counters c = {0, 0, 0};
            const float *q = queries + (size_t)qi * dim;
            if (mode == 1)
                po_search_pss_v5_t1(g, base, dim, q, K, L, M, init_ids, set_L, set_K, visited, &c);
            else
                po_search_top_m(g, base, dim, q, K, L, M, init_ids, set_L, set_K, visited, &c);
            for (uint32_t i = 0; i < K; ++i) {
                out_ids[(size_t)qi * K + i] = set_L[i].id;
                if (out_dists) out_dists[(size_t)qi * K + i] = set_L[i].dist;
            }
            if (out_ndist) out_ndist[qi] = (uint32_t)c.distance_computations;
            total += c.distance_computations;
        }
        free(set_L);
        free(visited);
        free(set_K);
    }
    return total;
}

void po_brute_force_knn(const float *base, uint32_t n, uint32_t dim, const float *queries,
                        uint32_t nq, uint32_t K, uint32_t *out_ids, float *out_dists,
                        int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        po_cand *q = (po_cand *)malloc(sizeof(po_cand) * K);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
            uint32_t size = 0;
            const float *qd = queries + (size_t)qi * dim;
            for (uint32_t v = 0; v < n; ++v) {
                po_cand c;
                memset(&c, 0, sizeof(c));
                c.id = v;
                c.dist = po_distance(base + (size_t)v * dim, qd, dim);
                po_add_into_queue(q, 0, &size, K, c);
            }
            for (uint32_t i = 0; i < K; ++i) {
                out_ids[(size_t)qi * K + i] = i < size ? q[i].id : 0xFFFFFFFFu;
                if (out_dists) out_dists[(size_t)qi * K + i] = i < size ? q[i].dist : INFINITY;
            }
        }
        free(q);
    }
}

double po_recall(const uint32_t *true_ids, uint32_t true_k, const uint32_t *res_ids,
                 uint32_t res_k, uint32_t nq, uint32_t L, uint32_t R)
{
    const uint32_t set_K_size = L < R ? L : R;
    double hits = 0.0;
    for (uint32_t q = 0; q < nq; ++q) {
        for (uint32_t t = 0; t < R; ++t) {
            uint32_t tid = true_ids[(size_t)q * true_k + t];
            for (uint32_t n = 0; n < set_K_size && n < res_k; ++n)
                if (res_ids[(size_t)q * res_k + n] == tid) hits += 1.0;
        }
    }
    return hits / ((double)R * (double)nq);
}
