// Find-L driver over the PANNS::Searching drop-in: for each recall@100
// target, binary-search L in [L_lower, L_upper] with one top-M search call
// per query, print the operating point and save the last set_K lists.
//
// Command line and output fields follow
// app/profile_find_L_para_single_query_top_m_search_pure_large_graph.cpp
// (main: :171-300, search_one_time: :33-149):
//   find_L_top_m <data_file> <query_file> <nsg_path> <L_lower> <K> <result_file>
//                <M> <true_NN_file> <num_threads> <L_upper> <P@100> [<P@100> ...]
// Set PANNS_BATCH=1 to search all queries of one L in a single batched call
// instead of one call per query, or PANNS_PSS=1 to call
// para_search_PSS_v5_large_graph_dist_thresh_profiling per query (M unused;
// the single-thread PSS_v5 search equals top-M with M = 1; with
// <num_threads> above 1 it runs that many queues, PANNS_PSS_X sub-search
// steps each, default 1).  PANNS_OPT_GRAPH=1
// loads the graph with load_nsg_graph (the opt-graph family, distance
// norm - 2 v.q) and calls that family's para_search_with_top_m_pure
// (TOPM.cpp:1881-2070) or para_search_PSS_v5_profiling (PSS5.cpp:1211-1462).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <unordered_map>
#include <vector>

#include "panns/Searching.h"

namespace {

struct Run {
    double runtime = 0;
    uint64_t compt = 0;
    double recall100 = 0, recall1 = 0, avg_merge = 0;
};

// include/utils.h:146-163: per query a uint32 count then the ids.
void save_result(const char *filename, const std::vector<std::vector<unsigned>> &results)
{
    std::ofstream fout(filename, std::ios::binary);
    for (const auto &r : results) {
        const unsigned n = (unsigned)r.size();
        fout.write(reinterpret_cast<const char *>(&n), sizeof(n));
        fout.write(reinterpret_cast<const char *>(r.data()), (std::streamsize)(n * sizeof(unsigned)));
    }
}

Run search_one_time(PANNS::Searching &engine, unsigned M, unsigned L, unsigned K, bool batch, bool pss, bool opt,
                    unsigned X, const std::vector<std::vector<PANNS::idi>> &true_nn_list,
                    std::vector<std::vector<unsigned>> &set_K_list)
{
    const unsigned nq = engine.num_queries_;
    set_K_list.assign(nq, std::vector<unsigned>(K));
    std::vector<PANNS::idi> init_ids(L);
    PANNS::dynamic_bitset is_visited(engine.num_v_);
    const unsigned T = (unsigned)engine.num_threads_;
    std::vector<PANNS::Candidate> set_L((T - 1) * L + L);
    std::vector<PANNS::idi> local_queues_sizes(T, 0), local_queues_starts(T);
    for (unsigned t = 0; t < T; ++t) local_queues_starts[t] = t * L;
    std::vector<PANNS::idi> top_m_candidates(M);

    const auto s = std::chrono::high_resolution_clock::now();
    engine.prepare_init_ids(init_ids, L);
    if (batch) {
        std::vector<PANNS::idi> qids(nq);
        for (unsigned q = 0; q < nq; ++q) qids[q] = q;
        engine.para_search_with_top_m_batch(M, qids, K, L, init_ids, set_K_list);
    } else if (pss && opt) {
        for (unsigned q = 0; q < nq; ++q)
            engine.para_search_PSS_v5_profiling(q, K, L, set_L, init_ids, set_K_list[q], L, local_queues_starts,
                                                local_queues_sizes, is_visited, X);
    } else if (pss) {
        for (unsigned q = 0; q < nq; ++q)
            engine.para_search_PSS_v5_large_graph_dist_thresh_profiling(q, K, L, set_L, init_ids, set_K_list[q], L,
                                                                        local_queues_starts, local_queues_sizes,
                                                                        is_visited, X);
    } else if (opt) {
        for (unsigned q = 0; q < nq; ++q)
            engine.para_search_with_top_m_pure(M, q, K, L, set_L, init_ids, set_K_list[q], L, local_queues_starts,
                                               local_queues_sizes, top_m_candidates, is_visited);
    } else {
        for (unsigned q = 0; q < nq; ++q)
            engine.para_search_with_top_m_pure_large_graph(M, q, K, L, set_L, init_ids, set_K_list[q], L,
                                                           local_queues_starts, local_queues_sizes,
                                                           top_m_candidates, is_visited);
    }
    const auto e = std::chrono::high_resolution_clock::now();
    Run r;
    r.runtime = std::chrono::duration<double>(e - s).count();
    std::unordered_map<unsigned, double> recalls;
    engine.get_recall_for_all_queries(true_nn_list, set_K_list, recalls, L);
    r.recall100 = recalls[100];
    r.recall1 = recalls[1];
    r.compt = engine.count_distance_computation_;
    r.avg_merge = engine.count_merge_ * 1.0 / nq;
    const double d = (double)engine.dimension_;
    printf("num_threads: %d M: %u L: %u runtime(s.): %f computation: %lu K: %u Volume: %u Dimension: %u "
           "query_num: %u query_per_sec: %f avg_latency(ms.): %f P@100: %f P@1: %f G/s: %f GFLOPS: %f "
           "avg_merge: %f\n",
           engine.num_threads_, M, L, r.runtime, (unsigned long)r.compt, K, engine.num_v_, (unsigned)d, nq,
           nq / r.runtime, r.runtime * 1000 / nq, r.recall100, r.recall1,
           d * 4.0 * r.compt / (1u << 30) / r.runtime, d * 3.0 * r.compt / (1u << 30) / r.runtime, r.avg_merge);
    engine.count_distance_computation_ = 0;
    engine.count_merge_ = 0;
    return r;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 12) {
        fprintf(stderr,
                "Usage: %s <data_file> <query_file> <nsg_path> <L_lower> <K> <result_file> <M> <true_NN_file> "
                "<num_threads> <L_upper> <P@100> [<P@100> ...]\n",
                argv[0]);
        return EXIT_FAILURE;
    }
    setbuf(stdout, nullptr);
    const char *b = getenv("PANNS_BATCH");
    const bool batch = b && b[0] == '1';
    const char *ps = getenv("PANNS_PSS");
    const bool pss = ps && ps[0] == '1';
    const char *og = getenv("PANNS_OPT_GRAPH");
    const bool opt = og && og[0] == '1';
    const char *px = getenv("PANNS_PSS_X");
    const unsigned X = px ? (unsigned)strtoul(px, nullptr, 0) : 1u;

    PANNS::Searching engine;
    engine.load_data_load(argv[1]);
    engine.load_queries_load(argv[2]);
    if (opt)
        engine.load_nsg_graph(argv[3]);
    else
        engine.load_common_nsg_graph(argv[3]);
    const unsigned L_lower_origin = (unsigned)strtoul(argv[4], nullptr, 0);
    const unsigned K = (unsigned)strtoul(argv[5], nullptr, 0);
    const unsigned M = (unsigned)strtoul(argv[7], nullptr, 0);
    std::vector<std::vector<PANNS::idi>> true_nn_list;
    engine.load_true_NN(argv[8], true_nn_list);
    engine.num_threads_ = (int)strtol(argv[9], nullptr, 0);
    const unsigned L_upper_origin = (unsigned)strtoul(argv[10], nullptr, 0);

    for (int a = 11; a < argc; ++a) {
        const double P_dest = strtod(argv[a], nullptr);
        std::vector<std::vector<unsigned>> set_K_list;
        unsigned L_lower = L_lower_origin, L_upper = L_upper_origin, L = L_upper;
        Run cur, last;
        unsigned last_L = 0;
        bool have_last = false;
        while (L_lower <= L_upper) {
            printf("L: %u L_lower: %u L_upper: %u\n", L, L_lower, L_upper);
            cur = search_one_time(engine, M, L, K, batch, pss, opt, X, true_nn_list, set_K_list);
            if (cur.recall100 < P_dest) {
                L_lower = L + 1;
            } else if (cur.recall100 > P_dest) {
                L_upper = L - 1;
                last = cur;
                last_L = L;
                have_last = true;
            } else {
                break;
            }
            if (L_lower <= L_upper) L = (L_lower + L_upper) / 2;
        }
        if (cur.recall100 < P_dest && L < L_upper_origin && have_last) {
            cur = last;
            L = last_L;
        }
        save_result(argv[6], set_K_list);
        printf("---- FINAL ----\n");
        printf("P_dest: %f runtime(s.): %f compt.: %lu P@100: %f latency(ms.): %f L: %u M: %u avg_merge: %f\n",
               P_dest, cur.runtime, (unsigned long)cur.compt, cur.recall100,
               cur.runtime / engine.num_queries_ * 1000.0, L, M, cur.avg_merge);
    }
    return 0;
}
