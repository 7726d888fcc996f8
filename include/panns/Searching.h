/* C++ host drop-in for PANNS::Searching over the C-ABI of panns_amd.h.
 * Header-only: include it and link libpanns_amd.so.
 *
 * It keeps the names, argument order and meaning of the reference classes
 * of both layouts: the large-graph CSR form
 * (core/Searching.202102031939.PSS_v5.large_graph.dist_thresh.profiling.h:30-268,
 * core/Searching.202011241615.top_m_pure_large_graph.h:30-352) and the
 * opt-graph form (core/Searching.202101301609.PSS_v5.profiling.h:30-270,
 * core/Searching.202011241533.top_m_pure.h:30-360), so a driver written
 * against the reference, such as
 * app/profile_find_L_para_single_query_top_m_search_pure_large_graph.cpp,
 * compiles against this header once its include line points here.  The
 * graph loader picks the distance as the reference's class does:
 * load_common_nsg_graph -> squared L2, load_nsg_graph -> norm - 2 v.q.
 *
 * What differs, and why:
 *  - The base set, queries and graph live in GPU memory owned by the handle;
 *    data_load_, queries_load_ and the CSR members stay nullptr.
 *  - is_visited is a PANNS::dynamic_bitset with the part of
 *    boost::dynamic_bitset<> the reference uses (operator[], reset, size,
 *    resize); the search keeps its visited set on the GPU and leaves the
 *    caller's bitset clear, which is how the reference leaves it.
 *  - Errors print "Error: ..." to stderr and exit(EXIT_FAILURE) as the
 *    reference does; define PANNS_SEARCHING_THROW before including this
 *    header to get std::runtime_error instead.
 *  - num_threads_: the top-M searches need 1 (the GPU reproduces the
 *    reference's single-thread run).  The PSS_v5 searches take 1 to 16:
 *    the reference's threads race, so its T > 1 ids vary from run to run;
 *    here the T queues run in lockstep in a fixed order (step s expands the
 *    s-th candidate of worker 0, 1, ..., then of the master), one of the
 *    schedules the reference can take, and worker queues hold
 *    local_queue_capacity = L entries as APP_PSS sets them.
 *  - time_expand_ receives the device time of the search (expansion and
 *    merging are one kernel); time_merge_ and time_seq_ stay 0.
 */
#ifndef PANNS_SEARCHING_DROP_IN_H
#define PANNS_SEARCHING_DROP_IN_H

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "panns_amd.h"

namespace PANNS {

using dataf = float;
using distf = float;
using idi = uint32_t;
using edgei = uint64_t;

/* include/Candidate.h:11-28: ordered by (distance, id). */
struct Candidate {
    idi id_;
    distf distance_;
    bool is_checked_;

    Candidate() = default;
    Candidate(idi id, distf dist, bool is_checked) : id_(id), distance_(dist), is_checked_(is_checked) {}
    bool operator<(const Candidate &b) const
    {
        return distance_ != b.distance_ ? distance_ < b.distance_ : id_ < b.id_;
    }
};

/* The members of boost::dynamic_bitset<> that the reference's search path
 * uses.  Bit v lives in 64-bit block v / 64 at position v % 64, Boost's
 * block order. */
class dynamic_bitset {
public:
    class reference {
    public:
        reference(uint64_t &w, uint64_t m) : w_(w), m_(m) {}
        operator bool() const { return (w_ & m_) != 0; }
        reference &operator=(bool v)
        {
            w_ = v ? (w_ | m_) : (w_ & ~m_);
            return *this;
        }
        reference &operator=(const reference &o) { return *this = static_cast<bool>(o); }

    private:
        uint64_t &w_;
        uint64_t m_;
    };

    dynamic_bitset() = default;
    explicit dynamic_bitset(size_t n) { resize(n); }
    void resize(size_t n, bool value = false)
    {
        const size_t old = n_;
        words_.resize((n + 63) / 64, 0);
        n_ = n;
        if (value)
            for (size_t i = old; i < n; ++i) words_[i / 64] |= 1ull << (i % 64);
        if (n % 64) words_.back() &= (1ull << (n % 64)) - 1;
    }
    size_t size() const { return n_; }
    size_t num_blocks() const { return words_.size(); }
    bool operator[](size_t i) const { return (words_[i / 64] >> (i % 64)) & 1u; }
    reference operator[](size_t i) { return reference(words_[i / 64], 1ull << (i % 64)); }
    bool test(size_t i) const { return (*this)[i]; }
    dynamic_bitset &set(size_t i, bool v = true)
    {
        (*this)[i] = v;
        return *this;
    }
    dynamic_bitset &reset()
    {
        for (auto &w : words_) w = 0;
        return *this;
    }
    size_t count() const
    {
        size_t c = 0;
        for (auto w : words_) c += (size_t)__builtin_popcountll(w);
        return c;
    }
    bool none() const
    {
        for (auto w : words_)
            if (w) return false;
        return true;
    }
    uint64_t *blocks() { return words_.data(); }
    const uint64_t *blocks() const { return words_.data(); }

private:
    std::vector<uint64_t> words_;
    size_t n_ = 0;
};

class Searching {
public:
    /* PSS5LG.h:32-48 */
    idi num_v_ = 0;
    edgei num_e_ = 0;
    idi num_queries_ = 0;
    uint64_t dimension_ = 0;
    idi width_ = 0;
    idi ep_ = 0;
    dataf *data_load_ = nullptr;
    dataf *queries_load_ = nullptr;
    idi *common_nsg_deg_ngbrs_ = nullptr;
    edgei *common_nsg_vertex_base_ = nullptr;
    int num_threads_ = 1;

    /* PSS5LG.h:170-175 */
    uint64_t count_distance_computation_ = 0;
    uint64_t count_merge_ = 0;
    double time_expand_ = 0.0;
    double time_merge_ = 0.0;
    double time_seq_ = 0.0;

    explicit Searching(int device = 0) { check(panns_create(device, &h_), "panns_create"); }
    ~Searching()
    {
        if (h_) panns_destroy(h_);
    }
    Searching(const Searching &) = delete;
    Searching &operator=(const Searching &) = delete;

    /* The C-ABI handle, for batched calls (panns_plan_*). */
    panns_index *handle() const { return h_; }

    /* PSS5LG.cpp:14-30 */
    void load_data_load(char *filename)
    {
        check(panns_load_data_load(h_, filename), "load_data_load");
        refresh_info();
    }
    /* PSS5LG.cpp:35-49 */
    void load_queries_load(char *filename)
    {
        check(panns_load_queries_load(h_, filename), "load_queries_load");
        refresh_info();
    }
    /* Alias of load_queries_load under the shorter name. */
    void load_queries(char *filename) { load_queries_load(filename); }
    /* PSS5LG.cpp:138-251 */
    void load_common_nsg_graph(char *filename)
    {
        check(panns_load_common_nsg_graph(h_, filename), "load_common_nsg_graph");
        refresh_info();
    }

    /* PSS5.cpp:56-132 (opt-graph form): the same .nsg file; later searches
     * use norm(v) - 2 v.q (PSS5.cpp:336-390) with the norms computed on the
     * GPU (PSS5.cpp:293-334).  The base set must be loaded first. */
    void load_nsg_graph(char *filename)
    {
        check(panns_load_nsg_graph(h_, filename), "load_nsg_graph");
        refresh_info();
    }

    /* PSS5LG.cpp:356-410: L ids; init_ids is resized to L when smaller. */
    void prepare_init_ids(std::vector<unsigned> &init_ids, const unsigned L) const
    {
        if (init_ids.size() < L) init_ids.resize(L);
        check(panns_prepare_init_ids(h_, L, init_ids.data()), "prepare_init_ids");
    }

    /* PSS5LG.cpp:259-307: [uint32 Q][uint32 K] then Q*K (uint32 id, float dist). */
    void load_true_NN(const char *filename, std::vector<std::vector<idi>> &true_nn_list)
    {
        std::ifstream fin(filename, std::ios::binary);
        if (!fin.is_open()) fail(std::string("cannot open file ") + filename);
        idi t_query_num = 0, t_K = 0;
        fin.read(reinterpret_cast<char *>(&t_query_num), sizeof(t_query_num));
        fin.read(reinterpret_cast<char *>(&t_K), sizeof(t_K));
        if (!fin) fail(std::string("cannot read the header of ") + filename);
        if (t_query_num < num_queries_)
            fail("t_query_num " + std::to_string(t_query_num) + " is smaller than num_queries_ " +
                 std::to_string(num_queries_));
        if (t_K < 100) fail("t_K " + std::to_string(t_K) + " is smaller than 100.");
        true_nn_list.assign(t_query_num, std::vector<idi>(t_K));
        std::vector<uint32_t> row(2 * (size_t)t_K);
        for (idi q = 0; q < t_query_num; ++q) {
            fin.read(reinterpret_cast<char *>(row.data()), (std::streamsize)(row.size() * 4));
            if (!fin) fail(std::string("truncated ground-truth file ") + filename);
            for (idi i = 0; i < t_K; ++i) true_nn_list[q][i] = row[2 * (size_t)i];
        }
    }

    /* PSS5LG.cpp:309-348: recalls[x] for x in {1,5,10,20,50,100}: true
     * top-100 ids found in set_K[0..min(L,100)) at a position below x,
     * divided by x * num_queries_.  set_K lists must hold min(L,100) ids. */
    void get_recall_for_all_queries(const std::vector<std::vector<idi>> &true_nn_list,
                                    const std::vector<std::vector<unsigned>> &set_K_list,
                                    std::unordered_map<unsigned, double> &recalls, const idi L) const
    {
        if (true_nn_list.empty() || true_nn_list[0].size() < 100)
            fail("Number of true nearest neighbors of a query is smaller than 100.");
        static const unsigned ats[6] = {1, 5, 10, 20, 50, 100};
        double hits[6] = {0, 0, 0, 0, 0, 0};
        const idi n = L < 100 ? L : 100;
        for (idi q = 0; q < num_queries_; ++q) {
            const std::vector<idi> &truth = true_nn_list[q];
            const std::vector<unsigned> &got = set_K_list[q];
            for (idi t = 0; t < 100; ++t)
                for (idi i = 0; i < n; ++i)
                    if (got[i] == truth[t])
                        for (int a = 0; a < 6; ++a)
                            if (i < ats[a]) hits[a] += 1;
        }
        for (int a = 0; a < 6; ++a) recalls[ats[a]] = hits[a] / (ats[a] * (double)num_queries_);
    }

    /* PSS5LG.cpp:1376-1724 with num_threads_ = T queues (1 to 16).  The
     * master queue is set_L[local_queues_starts[T-1], +L); worker queues
     * hold local_queue_capacity entries, which must be L when T > 1 (the
     * reference's drivers set it so, APP_PSS:136-141).  At one thread there
     * are no worker queues, and the ids do not depend on
     * subsearch_iterations. */
    void para_search_PSS_v5_large_graph_dist_thresh_profiling(
        const idi query_id, const idi K, const idi L, std::vector<Candidate> &set_L,
        const std::vector<idi> &init_ids, std::vector<idi> &set_K, const idi local_queue_capacity,
        const std::vector<idi> &local_queues_starts, std::vector<idi> &local_queues_sizes,
        dynamic_bitset &is_visited, const idi subsearch_iterations)
    {
        if (num_threads_ > 1) {
            if (local_queue_capacity != L)
                fail("local_queue_capacity is " + std::to_string(local_queue_capacity) + "; with num_threads_ > 1 "
                     "the GPU search keeps worker queues of L entries (set it to L)");
            if (local_queues_starts.size() < (size_t)num_threads_ || local_queues_sizes.size() < (size_t)num_threads_)
                fail("local_queues_starts and local_queues_sizes need num_threads_ entries");
        }
        // M carries the subsearch count X (at T = 1 it also sets count_merge_
        // as the reference counts merges, ceil((expansions - 1) / X))
        panns_params p = params(PANNS_MODE_PSS_V5, K, L, subsearch_iterations);
        one_query(p, query_id, set_L, init_ids, set_K, local_queues_starts, local_queues_sizes, is_visited);
    }

    /* TOPMLG.cpp:2022-2207 with num_threads_ = 1. */
    void para_search_with_top_m_pure_large_graph(
        const idi M, const idi query_id, const idi K, const idi L, std::vector<Candidate> &set_L,
        const std::vector<idi> &init_ids, std::vector<idi> &set_K, const idi local_queue_capacity,
        const std::vector<idi> &local_queues_starts, std::vector<idi> &local_queues_sizes,
        std::vector<idi> &top_m_candidates, dynamic_bitset &is_visited)
    {
        (void)local_queue_capacity;
        (void)top_m_candidates;
        panns_params p = params(PANNS_MODE_TOP_M, K, L, M);
        one_query(p, query_id, set_L, init_ids, set_K, local_queues_starts, local_queues_sizes, is_visited);
    }

    /* Opt-graph form, PSS5.cpp:1211-1462 at num_threads_ = 1: the same
     * search as the large-graph one, with the distance of the loaded graph. */
    void para_search_PSS_v5_profiling(const idi query_id, const idi K, const idi L, std::vector<Candidate> &set_L,
                                      const std::vector<idi> &init_ids, std::vector<idi> &set_K,
                                      const idi local_queue_capacity, const std::vector<idi> &local_queues_starts,
                                      std::vector<idi> &local_queues_sizes, dynamic_bitset &is_visited,
                                      const idi subsearch_iterations)
    {
        para_search_PSS_v5_large_graph_dist_thresh_profiling(query_id, K, L, set_L, init_ids, set_K,
                                                             local_queue_capacity, local_queues_starts,
                                                             local_queues_sizes, is_visited, subsearch_iterations);
    }

    /* Opt-graph form, TOPM.cpp:1881-2070 (its num_threads_ = 1 branch, :1922). */
    void para_search_with_top_m_pure(const idi M, const idi query_id, const idi K, const idi L,
                                     std::vector<Candidate> &set_L, const std::vector<idi> &init_ids,
                                     std::vector<idi> &set_K, const idi local_queue_capacity,
                                     const std::vector<idi> &local_queues_starts,
                                     std::vector<idi> &local_queues_sizes, std::vector<idi> &top_m_candidates,
                                     dynamic_bitset &is_visited)
    {
        para_search_with_top_m_pure_large_graph(M, query_id, K, L, set_L, init_ids, set_K, local_queue_capacity,
                                                local_queues_starts, local_queues_sizes, top_m_candidates,
                                                is_visited);
    }

    /* TOPM.cpp:1382-1484 seq_search_with_top_m_pure: top-M with the queue
     * at set_L[0, L); count_merge_ is not touched, as in the reference. */
    void seq_search_with_top_m_pure(const idi M, const idi query_id, const idi K, const idi L,
                                    std::vector<Candidate> &set_L, const std::vector<idi> &init_ids,
                                    std::vector<idi> &set_K, std::vector<idi> &top_m_candidates,
                                    dynamic_bitset &is_visited)
    {
        (void)top_m_candidates;
        const std::vector<idi> starts;
        std::vector<idi> sizes;
        one_query(params(PANNS_MODE_TOP_M, K, L, M), query_id, set_L, init_ids, set_K, starts, sizes, is_visited,
                  false);
    }

    /* TOPM.cpp:1273-1380 seq_search_with_top_m_scale_m: M = 1, 2, 4, ... up
     * to M_max from the initial queue at set_L[0, L); count_merge_ is not
     * touched, as in the reference. */
    void seq_search_with_top_m_scale_m(const idi M_max, const idi query_id, const idi K, const idi L,
                                       std::vector<Candidate> &set_L, const std::vector<idi> &init_ids,
                                       std::vector<idi> &set_K, std::vector<idi> &top_m_candidates,
                                       dynamic_bitset &is_visited)
    {
        (void)top_m_candidates;
        const std::vector<idi> starts;
        std::vector<idi> sizes;
        one_query(params(PANNS_MODE_SCALE_M, K, L, M_max), query_id, set_L, init_ids, set_K, starts, sizes,
                  is_visited, false);
    }

    /* TOPMLG.cpp:1023-1071: scale-M search on the initialised queue
     * set_L[set_L_start, +set_L_size), which must be full (set_L_size ==
     * local_L, as the reference's callers pass it).  is_visited is read and
     * left updated, and local_count_distance_computation is incremented, as
     * in the reference. */
    void subsearch_with_top_m(const idi value_M_max, const idi query_id, const idi local_L,
                              std::vector<Candidate> &set_L, const idi set_L_start, idi &set_L_size,
                              std::vector<idi> &local_top_m_candidates, dynamic_bitset &is_visited,
                              uint64_t &local_count_distance_computation)
    {
        (void)local_top_m_candidates;
        if (set_L.size() < (size_t)set_L_start + set_L_size) fail("set_L is smaller than set_L_start + set_L_size");
        if (is_visited.size() < num_v_) fail("is_visited holds fewer than num_v_ bits");
        std::vector<uint32_t> ids(set_L_size);
        std::vector<float> dists(set_L_size);
        std::vector<uint8_t> chk(set_L_size);
        for (idi i = 0; i < set_L_size; ++i) {
            const Candidate &c = set_L[set_L_start + i];
            ids[i] = c.id_;
            dists[i] = c.distance_;
            chk[i] = c.is_checked_ ? 1 : 0;
        }
        uint32_t size = set_L_size;
        check(panns_subsearch_with_top_m(h_, value_M_max, query_id, local_L, ids.data(), dists.data(), chk.data(),
                                         &size, is_visited.blocks(), &local_count_distance_computation),
              "subsearch_with_top_m");
        set_L_size = size;
        for (idi i = 0; i < size; ++i) set_L[set_L_start + i] = Candidate(ids[i], dists[i], chk[i] != 0);
    }

    /* Batched form (no reference counterpart; the reference loops over
     * queries one call each): all query_ids in one launch with the same
     * K, L, M and init_ids.  set_K_list[i] receives K ids of query_ids[i]. */
    void para_search_with_top_m_batch(const idi M, const std::vector<idi> &query_ids, const idi K, const idi L,
                                      const std::vector<idi> &init_ids,
                                      std::vector<std::vector<idi>> &set_K_list)
    {
        panns_params p = params(PANNS_MODE_TOP_M, K, L, M);
        const size_t nq = query_ids.size();
        std::vector<uint32_t> ids(nq * K), ndist(nq);
        panns_outputs out{};
        out.ids = ids.data();
        out.ndist = ndist.data();
        panns_counters cnt{};
        check(panns_search(h_, &p, query_ids.data(), (uint32_t)nq, init_ids.data(), (uint32_t)init_ids.size(), &out,
                           &cnt),
              "para_search_with_top_m_batch");
        set_K_list.resize(nq);
        for (size_t q = 0; q < nq; ++q) set_K_list[q].assign(ids.begin() + q * K, ids.begin() + (q + 1) * K);
        count_distance_computation_ += cnt.distance_computations;
        time_expand_ += cnt.kernel_ms * 1e-3;
    }

private:
    panns_index *h_ = nullptr;

    [[noreturn]] static void fail(const std::string &msg)
    {
#ifdef PANNS_SEARCHING_THROW
        throw std::runtime_error("Error: " + msg);
#else
        fprintf(stderr, "Error: %s\n", msg.c_str());
        exit(EXIT_FAILURE);
#endif
    }
    static void check(int st, const char *what)
    {
        if (st != PANNS_OK) fail(std::string(what) + ": " + panns_last_error());
    }

    void refresh_info()
    {
        uint32_t nv = 0, d = 0, nq = 0, w = 0, ep = 0;
        uint64_t ne = 0;
        check(panns_get_info(h_, &nv, &d, &nq, &w, &ep, &ne), "get_info");
        num_v_ = nv;
        dimension_ = d;
        num_queries_ = nq;
        width_ = w;
        ep_ = ep;
        num_e_ = ne;
    }

    panns_params params(uint32_t mode, idi K, idi L, idi M) const
    {
        if (mode == PANNS_MODE_PSS_V5 ? (num_threads_ < 1 || num_threads_ > 16) : num_threads_ != 1)
            fail("num_threads_ is " + std::to_string(num_threads_) +
                 (mode == PANNS_MODE_PSS_V5 ? "; the PSS_v5 search runs with 1 to 16"
                                            : "; the GPU top-M search reproduces the single-thread run only "
                                              "(set num_threads_ = 1)"));
        panns_params p{};
        p.K = K;
        p.L = L;
        p.M = M;
        p.mode = mode;
        p.num_threads = mode == PANNS_MODE_PSS_V5 ? (uint32_t)num_threads_ : 1u;
        p.stop = PANNS_STOP_FIXED;
        return p;
    }

    void one_query(const panns_params &p, idi query_id, std::vector<Candidate> &set_L,
                   const std::vector<idi> &init_ids, std::vector<idi> &set_K,
                   const std::vector<idi> &local_queues_starts, std::vector<idi> &local_queues_sizes,
                   dynamic_bitset &is_visited, bool count_merges = true)
    {
        const idi master = local_queues_starts.empty() ? 0 : local_queues_starts[num_threads_ - 1];
        if (set_L.size() < (size_t)master + p.L) fail("set_L is smaller than the master queue start + L");
        if (init_ids.size() < p.L) fail("init_ids holds fewer than L ids");
        std::vector<uint32_t> qids(p.L);
        std::vector<float> qd(p.L);
        std::vector<uint8_t> qc(p.L);
        if (set_K.size() < p.K) set_K.resize(p.K);
        uint32_t nd = 0;
        panns_outputs out{};
        out.ids = set_K.data();
        out.queue_ids = qids.data();
        out.queue_dists = qd.data();
        out.queue_checked = qc.data();
        out.ndist = &nd;
        panns_counters cnt{};
        check(panns_search(h_, &p, &query_id, 1, init_ids.data(), (uint32_t)init_ids.size(), &out, &cnt),
              "search");
        for (idi i = 0; i < p.L; ++i) set_L[master + i] = Candidate(qids[i], qd[i], qc[i] != 0);
        // the workers' queues end merged (empty), the master full
        if (local_queues_sizes.size() >= (size_t)num_threads_) {
            for (int w = 0; w + 1 < num_threads_; ++w) local_queues_sizes[w] = 0;
            local_queues_sizes[num_threads_ - 1] = p.L;
        }
        count_distance_computation_ += cnt.distance_computations;
        if (count_merges) count_merge_ += cnt.merges;
        time_expand_ += cnt.kernel_ms * 1e-3;
        is_visited.reset();
    }
};

} // namespace PANNS

#endif
