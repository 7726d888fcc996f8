// C-ABI of the MI355X NSG top-M search: handle, loaders, plans, searches.
// Device memory is owned by the handle; each handle has its own HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "panns_amd.h"
#include "panns_tools.h"
#include "search_kernel.h"
#include "search_kernel_mw.h"
#include "search_kernel_pss.h"
#include "knn_kernel.h"
#include "mlp_kernels.h"

#include <random>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(PANNS_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t ensure(size_t count)
    {
        if (count <= n && p) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
};

// Visited-table shape by a launch's largest L, measured on the 1M and 10M
// workloads (M = 4, every table size from 3/8 to 8/8 with 0, 256 or 1,024
// overflow entries, per L range): up to L = 200 the table holds almost every
// visited id at 6/8 with 256 overflow entries (8 workgroups per CU);
// 7/8 with 256 up to L = 512; 6/8 up to L = 800; above that the visited
// set outgrows any table that leaves room for occupancy, and 3/8 (more
// workgroups, more ids in the global bitmap) is fastest.
void table_by_L(uint32_t L, uint32_t *eighths, uint32_t *overflow)
{
    *eighths = L <= 200 ? 6u : L <= 512 ? 7u : L <= 800 ? 6u : 3u;
    *overflow = L <= 512 ? 256u : 0u;
}

uint32_t next_pow2(uint32_t x)
{
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}
uint32_t log2u(uint32_t x)
{
    uint32_t l = 0;
    while ((1u << l) < x) ++l;
    return l;
}

constexpr uint32_t kMaxL = 4096;
constexpr uint32_t kMaxM = 1024;
constexpr uint32_t kLogCap = 32768;
constexpr uint32_t kRanges = 8;         // work ranges (one per XCD)
constexpr uint32_t kPivots = 256;       // locality pivots
constexpr uint32_t kLocalityMin = 512;  // smallest batch that is reordered
constexpr uint32_t kPendCap = 256;      // pending keys of the deferred merge
// LDS allocation granule of a workgroup on gfx950, measured: one-wave
// workgroups of 32,672 bytes run 4 per CU (160 KiB of LDS) and of 31,776
// bytes 5 per CU (18 % faster on the same queries), while the runtime's
// occupancy query reports 5 for both; 1,280 bytes (160 KiB / 128) is the
// granule both fit
constexpr uint32_t kLdsGranule = 1280;
// exact k-NN: the automatic method takes the matrix-instruction path from
// this many base rows (measured, 64-NN of every row: 1.9x faster at 8,192
// rows, 2.0x at 49,000, 2.1x at 100,000)
constexpr uint32_t kKnnMfmaMinRows = 8192;
// automatic deferred merge above this L: measured on 1M x 96 (M = 4) it
// loses 8-19 % at L = 720 and 900, where its 2 KB of LDS cost a workgroup
// per CU, and gains 2 % at L = 1100 and 7 % at 1400; on 50M x 96 it gains
// 8 % for a bucket of L 579-1049
constexpr uint32_t kDeferMinL = 1000;
// ... and from this L where its buffer costs no workgroup per CU
constexpr uint32_t kDeferFreeMinL = 512;

}  // namespace

// Device memory of panns_search_noalloc: every buffer that search needs
// beyond the handle's resident base rows, queries and graph (query ids, init
// ids, outputs, work counters, visited bitmaps and spill logs) is carved
// from this file-scope array.  Each GPU the library runs on holds its own
// copy until the process ends; host code must use the per-device address
// from hipGetSymbolAddress, never the array's name.
constexpr size_t kNoAllocBytes = 64ull << 20;
constexpr uint32_t kNoAllocLogCap = 1024;  // spill log entries per workgroup (more: the whole bitmap is cleared)
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocDevices] = {};

// The address of this device's copy of `scratch`, looked up once per device
// with that device current (the caller has made it current).
int scratch_address(int device, uint8_t **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kNoAllocDevices) return fail(PANNS_ERR_DEVICE, "device ordinal above 63");
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(PANNS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    *out = static_cast<uint8_t *>(g_scratch_ptr[device]);
    return PANNS_OK;
}
}  // namespace

struct panns_index {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int num_cu = 0;
    uint32_t lds_per_cu = 0;
    // base and queries
    uint32_t n = 0, dim = 0, nq = 0;
    DevBuf<float> base, queries;
    // graph
    bool has_graph = false;
    uint32_t width = 0, ep = 0, max_deg = 0, stride_log2 = 5;
    uint64_t num_e = 0;
    DevBuf<uint32_t> adj;
    std::vector<uint32_t> ep_row;
    // scratch for visited spills: one bitmap and log per persistent workgroup
    DevBuf<uint32_t> bitmaps, spill_log;
    uint32_t bitmap_slots = 0;
    uint64_t bitmap_words = 0;
    // bumped whenever the scratch buffers are reallocated (a captured graph
    // holds their addresses and must be captured again)
    uint64_t scratch_gen = 0;
    // plans handed to the caller: panns_destroy detaches them, so a plan
    // destroyed after its index frees its own resources only
    std::unordered_set<panns_plan *> plans;
    // bumped whenever base rows, queries, graph or norms are replaced: a
    // plan holds their device addresses and its query ids were checked
    // against the query count of its build
    uint64_t input_gen = 0;
    DevBuf<uint32_t> counter;
    // learned stop model (device copy), 8-64-64-1
    DevBuf<float> model;
    bool has_model = false;
    // locality order of batches: pivot rows and their position on a
    // nearest-neighbour tour (computed once per base set)
    DevBuf<float> pivots;
    std::vector<uint32_t> tourpos;
    bool has_pivots = false;
    // distance family: squared L2 (large-graph searches) or norm - 2 v.q
    // (opt-graph searches), with the per-vertex norms the latter reads
    uint32_t distance = PANNS_DISTANCE_L2;
    DevBuf<float> norms;
    bool norms_valid = false;
    // device buffers of the last panns_search call, lent to the next one so
    // that a stream of calls of the same or smaller sizes allocates nothing
    struct OneShot {
        DevBuf<uint32_t> qids, init_ids, out_ids, out_stats, out_qids;
        DevBuf<float> out_dists, out_qdists;
        DevBuf<uint8_t> out_qchk;
    } oneshot;
};

struct panns_plan {
    panns_index *h = nullptr;  // null once the index is destroyed
    int device = 0;
    panns_params p{};
    uint32_t nq = 0, maxL = 0, maxK = 0;
    DevBuf<uint32_t> qids, qL, qK, init_ids, out_ids, out_stats, out_qids;
    DevBuf<float> out_dists, out_qdists;
    DevBuf<uint8_t> out_qchk;
    DevBuf<uint64_t> init_queue;
    DevBuf<float> trace_feat;
    DevBuf<uint32_t> trace_topk;
    uint32_t trace_cap = 0;
    DevBuf<uint32_t> trace_gt;   // true neighbours of the loaded queries (trace with hits)
    DevBuf<uint16_t> trace_hits;
    DevBuf<uint32_t> trace_ndist;
    uint32_t trace_gt_k = 0;
    bool want_queue = false;
    uint32_t *shared_bitmap = nullptr;
    uint32_t grid = 0, lds = 0, hash_slots = 0, bucket_slots = 8;
    uint32_t table_eighths = 8;
    uint32_t rows_per_wave = 0;  // neighbour rows per distance step (0: 96, as registers allow)
    uint32_t overflow_slots = 0;  // LDS overflow entries of the visited set (0: none)
    uint32_t deferred_merge = 0;  // panns_plan_options.deferred_merge
    uint32_t visited_table = 0;   // panns_plan_options.visited_table
    bool no_table = false;        // every visited id claimed in the bitmap
    uint32_t pend_cap = 0;        // pending-key buffer of the deferred merge (0: none)
    bool shape_given = false;     // launch shape from panns_plan_options (else automatic)
    uint32_t waves = 1;  // wavefronts per query
    bool pss = false;    // PSS_v5 with num_threads > 1 (search_kernel_pss)
    // work item i is the caller's query perm[i] (empty: the caller's order)
    std::vector<uint32_t> perm;
    bool locality = true;
    // a bucket launched side by side with others: its workgroups share the
    // CUs' LDS with theirs
    bool shares_cus = false;
    // scratch placement: bitmap/log slots [slot0, slot0 + grid) and the work
    // counter index, bound to the handle's buffers before every run
    uint32_t slot0 = 0, counter_idx = 0;
    DevBuf<uint64_t> phase_cycles;
    panns::KArgs args{};
    bool ran = false;
    // bucketed batches: parts[b] holds bucket b's queries (largest L first),
    // part_rows[b][i] the caller's row of its work item i
    std::vector<std::unique_ptr<panns_plan>> parts;
    std::vector<std::vector<uint32_t>> part_rows;
    panns_plan_options opt{};
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> side_ev;
    hipEvent_t fork_ev = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint64_t graph_gen = 0;  // the handle's scratch_gen when the graph was captured
    uint64_t input_gen = 0;  // the handle's input_gen when the plan was built

    ~panns_plan()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : side_ev) (void)hipEventDestroy(e);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        for (hipStream_t st : side) (void)hipStreamDestroy(st);
    }
};

// A plan the caller will hold: registered with its index.
static panns_plan *hand_out(panns_index *h, std::unique_ptr<panns_plan> &pl)
{
    pl->device = h->device;
    h->plans.insert(pl.get());
    return pl.release();
}

namespace {

template <int D, bool kLearned, bool kNorm, bool kTrace = false, bool kDefer = false>
hipError_t launch_search(const panns::KArgs &a, uint32_t grid, uint32_t lds, hipStream_t s)
{
    hipLaunchKernelGGL((panns::search_kernel<D, kLearned, kNorm, kTrace, kDefer>), dim3(grid), dim3(64), lds, s, a);
    return hipGetLastError();
}

template <int D, int W, bool kNorm>
hipError_t launch_search_mw(const panns::KArgs &a, uint32_t grid, uint32_t lds, hipStream_t s)
{
    hipLaunchKernelGGL((panns::search_kernel_mw<D, W, kNorm>), dim3(grid), dim3(64 * W), lds, s, a);
    return hipGetLastError();
}

template <int D, bool kNorm>
hipError_t launch_search_pss(const panns::KArgs &a, uint32_t grid, uint32_t lds, hipStream_t s)
{
    hipLaunchKernelGGL((panns::search_kernel_pss<D, kNorm>), dim3(grid), dim3(64), lds, s, a);
    return hipGetLastError();
}

using LaunchFn = hipError_t (*)(const panns::KArgs &, uint32_t, uint32_t, hipStream_t);
using KernelPtr = const void *;

// The search kernel for a dimension; `learned` selects the variant that
// evaluates the stop model (squared L2 only), `norm` the opt-graph distance,
// `defer` the fixed-L variant with the deferred merge.
bool kernel_for_dim(uint32_t dim, bool learned, bool norm, bool trace, bool defer, LaunchFn *fn, KernelPtr *kp)
{
    switch (dim) {
#define PANNS_DIM(DD)                                                                                   \
    case DD:                                                                                            \
        if (defer) {                                                                                    \
            *fn = norm      ? launch_search<DD, false, true, false, true>                               \
                  : learned ? launch_search<DD, true, false, false, true>                               \
                            : launch_search<DD, false, false, false, true>;                             \
            *kp = norm      ? reinterpret_cast<const void *>(&panns::search_kernel<DD, false, true, false, true>) \
                  : learned ? reinterpret_cast<const void *>(&panns::search_kernel<DD, true, false, false, true>) \
                            : reinterpret_cast<const void *>(&panns::search_kernel<DD, false, false, false, true>); \
            return true;                                                                                \
        }                                                                                               \
        if (trace) {                                                                                    \
            *fn = launch_search<DD, false, false, true>;                                                \
            *kp = reinterpret_cast<const void *>(&panns::search_kernel<DD, false, false, true>);        \
            return true;                                                                                \
        }                                                                                               \
        *fn = norm ? launch_search<DD, false, true>                                                     \
                   : (learned ? launch_search<DD, true, false> : launch_search<DD, false, false>);       \
        *kp = norm ? reinterpret_cast<const void *>(&panns::search_kernel<DD, false, true>)             \
                   : (learned ? reinterpret_cast<const void *>(&panns::search_kernel<DD, true, false>)  \
                              : reinterpret_cast<const void *>(&panns::search_kernel<DD, false, false>)); \
        return true;
        PANNS_DIM(16)
        PANNS_DIM(32)
        PANNS_DIM(40)
        PANNS_DIM(64)
        PANNS_DIM(96)
        PANNS_DIM(128)
#undef PANNS_DIM
    default:
        return false;
    }
}

// The multi-wave kernel for a dimension and W waves per query (2 or 4).
bool kernel_for_dim_mw(uint32_t dim, uint32_t waves, bool norm, LaunchFn *fn, KernelPtr *kp)
{
    switch (dim) {
#define PANNS_DIM(DD)                                                                                     \
    case DD:                                                                                              \
        if (norm) {                                                                                       \
            *fn = waves == 2 ? launch_search_mw<DD, 2, true> : launch_search_mw<DD, 4, true>;              \
            *kp = waves == 2 ? reinterpret_cast<const void *>(&panns::search_kernel_mw<DD, 2, true>)       \
                             : reinterpret_cast<const void *>(&panns::search_kernel_mw<DD, 4, true>);      \
        } else {                                                                                          \
            *fn = waves == 2 ? launch_search_mw<DD, 2, false> : launch_search_mw<DD, 4, false>;            \
            *kp = waves == 2 ? reinterpret_cast<const void *>(&panns::search_kernel_mw<DD, 2, false>)      \
                             : reinterpret_cast<const void *>(&panns::search_kernel_mw<DD, 4, false>);     \
        }                                                                                                 \
        return true;
        PANNS_DIM(16)
        PANNS_DIM(32)
        PANNS_DIM(40)
        PANNS_DIM(64)
        PANNS_DIM(96)
        PANNS_DIM(128)
#undef PANNS_DIM
    default:
        return false;
    }
}

// The PSS_v5 kernel with T > 1 queues for a dimension.
bool kernel_for_dim_pss(uint32_t dim, bool norm, LaunchFn *fn, KernelPtr *kp)
{
    switch (dim) {
#define PANNS_DIM(DD)                                                                                      \
    case DD:                                                                                               \
        *fn = norm ? launch_search_pss<DD, true> : launch_search_pss<DD, false>;                            \
        *kp = norm ? reinterpret_cast<const void *>(&panns::search_kernel_pss<DD, true>)                    \
                   : reinterpret_cast<const void *>(&panns::search_kernel_pss<DD, false>);                  \
        return true;
        PANNS_DIM(16)
        PANNS_DIM(32)
        PANNS_DIM(40)
        PANNS_DIM(64)
        PANNS_DIM(96)
        PANNS_DIM(128)
#undef PANNS_DIM
    default:
        return false;
    }
}

bool select_kernel(uint32_t dim, bool learned, bool norm, bool trace, uint32_t waves, bool pss, bool defer,
                   LaunchFn *fn, KernelPtr *kp)
{
    if (pss) return kernel_for_dim_pss(dim, norm, fn, kp);
    return waves > 1 ? kernel_for_dim_mw(dim, waves, norm, fn, kp)
                     : kernel_for_dim(dim, learned, norm, trace, defer, fn, kp);
}

using KnnLaunch = void (*)(const float *, uint32_t, const float *, uint32_t, uint32_t, int, uint32_t,
                          const uint32_t *, uint32_t *, float *, uint32_t, uint32_t, hipStream_t);
// The three launches of the matrix-instruction k-NN (knn_kernel.h): per-query
// bounds from a sample's exact k-th distance, the filter over every row, the
// exact re-rank of the candidates.
struct KnnMfmaLaunch {
    void (*bound)(const float *, uint32_t, const float *, uint32_t, float, float *, hipStream_t);
    void (*filter)(const float *, const float *, uint32_t, const float *, uint32_t, const float *, int, uint32_t,
                   uint32_t, uint32_t *, uint32_t, uint32_t *, uint32_t, hipStream_t);
    void (*rerank)(const float *, const float *, uint32_t, uint32_t, const uint32_t *, uint32_t, const uint32_t *,
                   uint32_t *, float *, uint32_t *, hipStream_t);
};

template <int D>
void launch_knn(const float *b, uint32_t n, const float *q, uint32_t nq, uint32_t k, int ex, uint32_t qoff,
                const uint32_t *self, uint32_t *oi, float *od, uint32_t grid, uint32_t lds, hipStream_t s)
{
    hipLaunchKernelGGL(panns::knn_kernel<D>, dim3(grid), dim3(256), lds, s, b, n, q, nq, k, ex, qoff, self, oi, od);
}

template <int D>
void launch_knn_bound(const float *q, uint32_t nq, const float *sd, uint32_t k, float vmax2, float *bound,
                      hipStream_t s)
{
    hipLaunchKernelGGL(panns::knn_bound_kernel<D>, dim3((nq + 255) / 256), dim3(256), 0, s, q, nq, sd, k, vmax2,
                       bound);
}

template <int D>
void launch_knn_filter(const float *b, const float *norms, uint32_t n, const float *q, uint32_t nq,
                       const float *bound, int ex, uint32_t qoff, uint32_t rows_per_split, uint32_t *cand,
                       uint32_t cap, uint32_t *cnt, uint32_t splits, hipStream_t s)
{
    hipLaunchKernelGGL(panns::knn_filter_kernel<D>, dim3((nq + 63) / 64, splits), dim3(256), 0, s, b, norms, n, q,
                       nq, bound, ex, qoff, rows_per_split, cand, cap, cnt);
}

template <int D>
void launch_knn_rerank(const float *b, const float *q, uint32_t nq, uint32_t k, const uint32_t *cand, uint32_t cap,
                       const uint32_t *cnt, uint32_t *oi, float *od, uint32_t *fl, hipStream_t s)
{
    hipLaunchKernelGGL(panns::knn_rerank_kernel<D>, dim3(nq), dim3(64), 0, s, b, q, k, cand, cap, cnt, oi, od, fl);
}

// PSS5.cpp:293-334 compute_norm for every base row, once per base set.
int ensure_norms(panns_index *h)
{
    if (h->norms_valid) return PANNS_OK;
    HIP_TRY(h->norms.ensure(h->n));
    const dim3 grid((h->n + 255) / 256), block(256);
    switch (h->dim) {
#define PANNS_DIM(DD)                                                                                  \
    case DD:                                                                                           \
        hipLaunchKernelGGL(panns::norm_kernel<DD>, grid, block, 0, h->stream, h->base.p, h->n, h->norms.p); \
        break;
        PANNS_DIM(16)
        PANNS_DIM(32)
        PANNS_DIM(40)
        PANNS_DIM(64)
        PANNS_DIM(96)
        PANNS_DIM(128)
#undef PANNS_DIM
    default:
        return fail(PANNS_ERR_DIM, "dimension " + std::to_string(h->dim) + " has no kernel (16, 32, 40, 64, 96, 128)");
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->norms_valid = true;
    return PANNS_OK;
}

int check_ready(panns_index *h)
{
    if (!h) return fail(PANNS_ERR_ARG, "null handle");
    if (!h->n) return fail(PANNS_ERR_STATE, "base set not loaded");
    if (!h->has_graph) return fail(PANNS_ERR_STATE, "graph not loaded");
    return PANNS_OK;
}

int check_params(panns_index *h, const panns_params *p, uint32_t L, uint32_t K)
{
    if (!p) return fail(PANNS_ERR_ARG, "null params");
    if (p->mode > 2) return fail(PANNS_ERR_ARG, "unknown mode");
    if (p->mode == PANNS_MODE_PSS_V5 && (p->num_threads == 0 || p->num_threads > panns::kPssMaxQueues))
        return fail(PANNS_ERR_UNSUPPORTED, "PSS_v5 runs with num_threads_ from 1 to 16");
    if (p->mode != PANNS_MODE_PSS_V5 && p->num_threads > 1)
        return fail(PANNS_ERR_UNSUPPORTED, "the top-M searches run with num_threads_ = 1 (PSS_v5 takes more)");
    if (p->mode == PANNS_MODE_PSS_V5 && p->num_threads > 1) {
        if (p->waves > 1) return fail(PANNS_ERR_UNSUPPORTED, "PSS_v5 with num_threads_ > 1 runs with waves = 1");
        if (p->stop != PANNS_STOP_FIXED)
            return fail(PANNS_ERR_UNSUPPORTED, "PSS_v5 with num_threads_ > 1 runs with the fixed stop");
        if (h->stride_log2 > 8)
            return fail(PANNS_ERR_UNSUPPORTED, "PSS_v5 with num_threads_ > 1 needs a graph degree of at most 256");
    }
    if (L == 0 || K == 0) return fail(PANNS_ERR_ARG, "K and L must be positive");
    if (K > L) return fail(PANNS_ERR_ARG, "K > L");
    if (L > h->n) return fail(PANNS_ERR_ARG, "L > number of base vectors");
    if (L > kMaxL) return fail(PANNS_ERR_UNSUPPORTED, "L above 4096 is not offered");
    if (p->M == 0) return fail(PANNS_ERR_ARG, "M must be positive");
    if (p->M > kMaxM) return fail(PANNS_ERR_UNSUPPORTED, "M above 1024 is not offered");
    if (p->stop == PANNS_STOP_LEARNED && !h->has_model) return fail(PANNS_ERR_STATE, "no stop model loaded");
    if (p->stop == PANNS_STOP_LEARNED && h->distance != PANNS_DISTANCE_L2)
        return fail(PANNS_ERR_UNSUPPORTED, "the learned stop's features are squared-L2 distances; use the fixed stop");
    if (p->hash_slots && (p->hash_slots % 64 || p->hash_slots < 64 || p->hash_slots > 32768))
        return fail(PANNS_ERR_ARG, "hash_slots must be a multiple of 64 in [64, 32768]");
    if (p->bucket_slots && p->bucket_slots != 2 && p->bucket_slots != 4 && p->bucket_slots != 8)
        return fail(PANNS_ERR_ARG, "bucket_slots must be 0 (automatic), 2, 4 or 8");
    if (p->waves > 1) {
        if (p->waves != 2 && p->waves != 4) return fail(PANNS_ERR_ARG, "waves must be 1, 2 or 4");
        if (p->mode == PANNS_MODE_PSS_V5) return fail(PANNS_ERR_UNSUPPORTED, "PSS_V5 expands one candidate per step; use waves = 1");
        if (p->stop != PANNS_STOP_FIXED) return fail(PANNS_ERR_UNSUPPORTED, "the learned stop runs with waves = 1");
    }
    return PANNS_OK;
}

// Layout of one workgroup's LDS; must match search_kernel's carving:
// Q[Lcap] cand[nb_cap] bb[64] P[pend_cap] (8 B each; the caller adds P's
// 8 * pend_cap bytes), query row (4*dim B),
// visited table (2 B per slot), newids[nb_cap] picks[Mcap] counts[4]
// (4 B each).  The stop model's hidden layer shares cand.
uint32_t lds_bytes(uint32_t Lcap, uint32_t nb_cap, uint32_t tbl_slots, uint32_t Mcap, uint32_t dim)
{
    return 8u * (Lcap + nb_cap + 64u) + 4u * dim + 2u * tbl_slots + 4u * (nb_cap + Mcap + 4u);
}

// search_kernel_mw's carving: Q[Lcap] cand[W][nb_cap] bb[64] (8 B each),
// query row, visited table, newids[W][nb_cap] picks[Mcap] ctrl (4 B each).
uint32_t lds_bytes_mw(uint32_t Lcap, uint32_t nb_cap, uint32_t tbl_slots, uint32_t Mcap, uint32_t dim, uint32_t W)
{
    return 8u * (Lcap + W * nb_cap + 64u) + 4u * dim + 2u * tbl_slots + 4u * (W * nb_cap + Mcap) +
           4u * panns::kCtrlWords;
}

// search_kernel_pss's carving: Q[T][Lcap] cand[nb_cap] bb[64] (8 B each),
// query row, visited table, newids/newown/candown[nb_cap], per-queue words
// and counts (4 B each).
uint32_t lds_bytes_pss(uint32_t Lcap, uint32_t nb_cap, uint32_t tbl_slots, uint32_t dim, uint32_t T)
{
    return 8u * (T * Lcap + nb_cap + 64u) + 4u * dim + 2u * tbl_slots + 4u * 3u * nb_cap +
           4u * (panns::kWsN * panns::kPssMaxQueues + 4u);
}

// One N-bit bitmap per persistent workgroup, at least 64 words long, so even
// for a tiny id space no two workgroups' bitmaps share a cache line.
uint64_t bitmap_words_for(uint32_t n) { return std::max<uint64_t>(((uint64_t)n + 31) / 32, 64); }

int ensure_scratch(panns_index *h, uint32_t grid)
{
    const uint64_t words = bitmap_words_for(h->n);
    if (h->bitmap_slots < grid || h->bitmap_words != words) {
        ++h->scratch_gen;
        h->bitmaps.release();
        h->spill_log.release();
        HIP_TRY(h->bitmaps.ensure((size_t)grid * words));
        HIP_TRY(hipMemsetAsync(h->bitmaps.p, 0, (size_t)grid * words * 4, h->stream));
        HIP_TRY(h->spill_log.ensure((size_t)grid * kLogCap));
        HIP_TRY(hipStreamSynchronize(h->stream));
        h->bitmap_slots = grid;
        h->bitmap_words = words;
    }
    if (!h->counter.p) HIP_TRY(h->counter.ensure((PANNS_MAX_BUCKETS + 1) * kRanges));
    return PANNS_OK;
}

// Locality order of a batch: each query's nearest of kPivots evenly spaced
// base rows (computed on the GPU), queries sorted by that pivot's position
// on a greedy nearest-neighbour tour of the pivots.  Queries that run
// together then touch the same part of the graph, and the kernel's
// per-XCD work ranges keep them on one L2.  Results do not depend on it.
int locality_order(panns_index *h, const uint32_t *query_ids, uint32_t nq, std::vector<uint32_t> &order)
{
    const uint32_t D = h->dim;
    const uint32_t P = std::min<uint32_t>(kPivots, h->n);
    if (!h->has_pivots) {
        std::vector<float> piv((size_t)P * D);
        for (uint32_t i = 0; i < P; ++i) {
            const uint64_t row = (uint64_t)i * h->n / P;
            HIP_TRY(hipMemcpy(piv.data() + (size_t)i * D, h->base.p + row * D, D * 4ull, hipMemcpyDeviceToHost));
        }
        HIP_TRY(h->pivots.ensure((size_t)P * D));
        HIP_TRY(hipMemcpy(h->pivots.p, piv.data(), (size_t)P * D * 4, hipMemcpyHostToDevice));
        // greedy tour from pivot 0
        std::vector<char> used(P, 0);
        h->tourpos.assign(P, 0);
        uint32_t cur = 0;
        used[0] = 1;
        for (uint32_t t = 1; t < P; ++t) {
            double best = 1e300;
            uint32_t bi = 0;
            for (uint32_t j = 0; j < P; ++j) {
                if (used[j]) continue;
                double d = 0;
                for (uint32_t k = 0; k < D; ++k) {
                    const double x = (double)piv[(size_t)cur * D + k] - (double)piv[(size_t)j * D + k];
                    d += x * x;
                }
                if (d < best) {
                    best = d;
                    bi = j;
                }
            }
            used[bi] = 1;
            h->tourpos[bi] = t;
            cur = bi;
        }
        h->has_pivots = true;
    }
    DevBuf<uint32_t> dq, dn;
    HIP_TRY(dq.ensure(nq));
    HIP_TRY(dn.ensure(nq));
    HIP_TRY(hipMemcpy(dq.p, query_ids, nq * 4ull, hipMemcpyHostToDevice));
    const dim3 pg((nq + 31) / 32), pb(256);
    switch (D) {
#define PANNS_DIM(DD)                                                                                        \
    case DD:                                                                                                 \
        hipLaunchKernelGGL(panns::nearest_pivot_kernel<DD>, pg, pb, 0, h->stream, h->queries.p, dq.p, nq,     \
                           h->pivots.p, P, dn.p);                                                            \
        break;
        PANNS_DIM(16)
        PANNS_DIM(32)
        PANNS_DIM(40)
        PANNS_DIM(64)
        PANNS_DIM(96)
        PANNS_DIM(128)
#undef PANNS_DIM
    default:
        return fail(PANNS_ERR_DIM, "dimension " + std::to_string(D) + " has no kernel");
    }
    HIP_TRY(hipGetLastError());
    std::vector<uint32_t> near(nq);
    HIP_TRY(hipMemcpyAsync(near.data(), dn.p, nq * 4ull, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    order.resize(nq);
    for (uint32_t i = 0; i < nq; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return h->tourpos[near[x]] < h->tourpos[near[y]]; });
    return PANNS_OK;
}

// Launch geometry of a plan (no device memory): the parameter checks, the
// kernel, the LDS carving and the persistent grid.  Fills *geo.
struct Geometry {
    uint32_t Lcap = 0, nb_cap = 0, Mcap = 0, nb = 0, bslots = 8, id_bits = 1;
    bool norm = false;
};

int plan_geometry(panns_plan *pl, const uint32_t *query_ids, const uint32_t *Lq, const uint32_t *Kq,
                  const uint32_t *init_ids, uint32_t n_init, Geometry *geo)
{
    panns_index *h = pl->h;
    const panns_params &p = pl->p;
    pl->waves = std::max<uint32_t>(1u, p.waves);
    if (pl->waves > 1 && (pl->init_queue.p || pl->trace_cap || pl->shared_bitmap))
        return fail(PANNS_ERR_UNSUPPORTED, "traces and caller-held queues run with waves = 1");
    pl->pss = p.mode == PANNS_MODE_PSS_V5 && p.num_threads > 1;
    if (pl->pss && (pl->init_queue.p || pl->trace_cap || pl->shared_bitmap))
        return fail(PANNS_ERR_UNSUPPORTED, "traces and caller-held queues run with num_threads_ = 1");
    const bool norm = h->distance == PANNS_DISTANCE_NORM_DOT;
    if (norm && pl->trace_cap) return fail(PANNS_ERR_UNSUPPORTED, "traces are taken with squared-L2 distances");
    LaunchFn fn;
    KernelPtr kp;
    if (!select_kernel(h->dim, p.stop == PANNS_STOP_LEARNED, norm, pl->trace_cap != 0, pl->waves, pl->pss, false,
                       &fn, &kp))
        return fail(PANNS_ERR_DIM, "dimension " + std::to_string(h->dim) + " has no kernel (16, 32, 40, 64, 96, 128)");
    pl->maxL = p.L;
    pl->maxK = p.K;
    if (Lq || Kq) {
        pl->maxL = 0;
        pl->maxK = 0;
        for (uint32_t i = 0; i < pl->nq; ++i) {
            const uint32_t L = Lq ? Lq[i] : p.L, K = Kq ? Kq[i] : p.K;
            int rc = check_params(h, &p, L, K);
            if (rc) return rc;
            pl->maxL = std::max(pl->maxL, L);
            pl->maxK = std::max(pl->maxK, K);
        }
    }
    int rc = check_params(h, &p, pl->maxL, pl->maxK);
    if (rc) return rc;
    if (!pl->init_queue.p && n_init < pl->maxL) return fail(PANNS_ERR_ARG, "init_ids shorter than L");
    for (uint32_t i = 0; i < pl->nq; ++i)
        if (query_ids[i] >= h->nq) return fail(PANNS_ERR_RANGE, "query id out of range");
    if (!pl->init_queue.p)
        for (uint32_t i = 0; i < pl->maxL; ++i)
            if (init_ids[i] >= h->n) return fail(PANNS_ERR_RANGE, "init id out of range");

    const uint32_t stride = 1u << h->stride_log2;
    // multi-wave: per-wave capacity for its share of the picks' neighbours
    const uint32_t W = pl->waves;
    // PSS_v5 with T queues: one step's neighbours (T rows), and a worker
    // queue's entries staged for the merge into the master
    // one wave: a pick group's slots (M rows, at most kGroupSlots) plus room
    // for the compaction staging (128 keys) -- the LDS it saves at small M
    // lets long queues fit one more workgroup per CU
    uint32_t nb_cap =
        pl->pss ? (std::max<uint32_t>({panns::kGroupSlots, p.num_threads * stride, (pl->maxL + 7u) & ~7u}) + 7u) & ~7u
        : W > 1 ? std::max<uint32_t>(panns::kGroupSlots, (p.M + W - 1) / W * stride)
                : std::max<uint32_t>({128u, stride, std::min<uint32_t>(panns::kGroupSlots, p.M * stride)});
    const uint32_t Mcap = std::max<uint32_t>(4u, (p.M + 3u) & ~3u);
    // visited table: hs 16-bit slots in nb buckets of 8; the tag of an id
    // is its scrambled value divided by nb and must stay below 2^15 (0xFFFF
    // marks an empty slot, bit 15 a second-bucket entry), so nb grows with N
    uint32_t hs = p.hash_slots;
    // automatic size: 8,192 slots up to L = 512, 10,240 above (measured:
    // larger tables lose more to occupancy than they save in spills), taken
    // whole or in eighths.  A plan built without launch options or a table
    // size, or with table_eighths = 0, sizes it by its largest L
    // (table_by_L).
    const bool auto_size = (!pl->shape_given || pl->table_eighths == 0) && !p.hash_slots;
    if (auto_size)
        table_by_L(pl->maxL, &pl->table_eighths, &pl->overflow_slots);
    else if (pl->table_eighths == 0)
        pl->table_eighths = 8;
    if (!hs) hs = ((pl->maxL <= 512 ? 8192u : 10240u) * pl->table_eighths / 8u + 63u) & ~63u;
    const uint32_t id_bits = std::max<uint32_t>(1u, log2u(h->n));
    const uint64_t min_nb = (((1ull << id_bits) - 1ull) / 0x7FFEull) + 1ull;
    // buckets of 8 slots, or of 4 when an 8-slot table of this size has
    // fewer buckets than the id space needs (ids above 2^25 at 8,192 slots):
    // the table then keeps its size instead of growing with N.  A table of
    // at most half the least 4-slot one (the small tables long queues ask
    // for, on large id spaces) takes buckets of 2 slots: measured on 50M x
    // 96, 4,112 two-slot slots instead of 8,224 four-slot ones fit 6 and 5
    // workgroups per CU instead of 4 for L 1,368-1,683 and 1,684-2,000, -8
    // and -6 %, though 60 % more ids spill to the bitmap.
    const uint32_t nb_min = (uint32_t)((min_nb + 7ull) & ~7ull);
    uint32_t bslots = p.bucket_slots ? p.bucket_slots
                      : (hs / 8u) >= min_nb ? 8u
                      : 2u * hs <= 4u * nb_min ? 2u
                                               : 4u;
    uint32_t nb = std::max<uint32_t>(hs / bslots, nb_min);
    nb = std::max<uint32_t>(nb, 8u);
    hs = nb * bslots;
    // No table (visited_table = 2 only): every visited id is claimed in the
    // bitmap.  Measured on 50M x 96 over L 100-2000, a table is faster at
    // every L even where the id space forces it to 8,224 slots (16-bit tags
    // hold 15 bits of the scrambled id) and the LDS allows 4 workgroups per
    // CU: 4 to 48 % per L range, since a claim in the bitmap is a memory
    // atomic on a random line and the bitmap must be cleared again.
    pl->no_table = !pl->pss && !pl->shared_bitmap && pl->visited_table == 2;
    if (pl->no_table) {
        hs = 64;
        nb = 8;
        bslots = 8;
        pl->overflow_slots = 0;
    }
    if (pl->shared_bitmap) {
        hs = 64;
        nb = 8;
        bslots = 8;
    }
    pl->bucket_slots = bslots;
    pl->hash_slots = hs;
    // Lcap rounded to 8 keys keeps every LDS region 16-byte aligned.
    const uint32_t Lcap = (pl->maxL + 7u) & ~7u;
    const uint32_t ovf = pl->shared_bitmap ? 0u : pl->overflow_slots;
    // the overflow entries are 32-bit, i.e. two 16-bit slots each in the carving
    // deferred merge: the one-wave search (fixed or learned stop) without
    // traces or a caller-held queue, at most 64 picks per iteration
    const bool can_defer = !pl->pss && W == 1 && !pl->trace_cap && !pl->init_queue.p && !pl->shared_bitmap &&
                           p.M <= 64;
    pl->pend_cap = can_defer && (pl->deferred_merge == 2 || (pl->deferred_merge == 0 && pl->maxL > kDeferMinL))
                       ? kPendCap : 0u;
    // Below that, from L = kDeferFreeMinL, a plan that has the CUs to itself
    // defers when the pending keys fit in the LDS the allocation granule
    // leaves free, i.e. cost no workgroup per CU (measured on 50M x 96 at
    // equal occupancy: +2.8 % for L 418-734, +8.6 % for L 735-1049; where
    // they cost a workgroup, L 100-417, -14 %).
    if (!pl->pend_cap && can_defer && pl->deferred_merge == 0 && pl->maxL > kDeferFreeMinL && !pl->shares_cus &&
        h->lds_per_cu) {
        const auto fit = [&](uint32_t x) {
            return h->lds_per_cu / ((x + kLdsGranule - 1u) / kLdsGranule * kLdsGranule);
        };
        const uint32_t nbd = std::max<uint32_t>(nb_cap, 3u * Mcap);
        if (fit(lds_bytes(Lcap, nbd, hs + 2u * ovf, Mcap, h->dim) + 8u * kPendCap) >=
            fit(lds_bytes(Lcap, nb_cap, hs + 2u * ovf, Mcap, h->dim)))
            pl->pend_cap = kPendCap;
    }
    // the deferred pick stages its candidates and positions in 3 * Mcap keys
    // of cand
    if (pl->pend_cap) nb_cap = std::max<uint32_t>(nb_cap, 3u * Mcap);
    pl->lds = pl->pss ? lds_bytes_pss(Lcap, nb_cap, hs + 2u * ovf, h->dim, p.num_threads)
              : W > 1 ? lds_bytes_mw(Lcap, nb_cap, hs + 2u * ovf, Mcap, h->dim, W)
                      : lds_bytes(Lcap, nb_cap, hs + 2u * ovf, Mcap, h->dim) + 8u * pl->pend_cap;
    if (pl->pend_cap)
        select_kernel(h->dim, p.stop == PANNS_STOP_LEARNED, norm, false, 1u, false, true, &fn,
                      &kp);  // the deferred-merge variant
    if (pl->lds > 65536) {
        HIP_TRY(hipFuncSetAttribute(kp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl->lds));
    }
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kp, 64 * W, pl->lds));
    if (per_cu <= 0) return fail(PANNS_ERR_UNSUPPORTED, "kernel does not fit on a CU with this L");
    // the CU allocates LDS to a workgroup in granules of kLdsGranule bytes,
    // which the occupancy query does not count: the grid holds the
    // workgroups that are resident together
    if (h->lds_per_cu) {
        const uint32_t g = (pl->lds + kLdsGranule - 1u) / kLdsGranule * kLdsGranule;
        per_cu = std::max(1, std::min(per_cu, (int)(h->lds_per_cu / std::max(g, 1u))));
        // A table sized here (no hash_slots given) grows into the LDS that
        // this many workgroups per CU leave free: more slots, fewer bitmap
        // claims (memory-side atomics), the same resident workgroups.  Whole
        // groups of 8 buckets, so the table keeps its alignment.  Not for a
        // bucket that runs side by side with others, whose workgroups share
        // the CUs with theirs (measured on 50M x 96, L 100-2000 in six
        // buckets: 2-4 % faster per bucket alone, 1.5 % slower side by side).
        const uint32_t budget = h->lds_per_cu / (uint32_t)per_cu / kLdsGranule * kLdsGranule;
        if (!p.hash_slots && !pl->no_table && !pl->shared_bitmap && !pl->shares_cus && budget > pl->lds) {
            const uint32_t step = 8u * bslots;
            const uint32_t extra = (budget - pl->lds) / 2u / step * step;
            if (extra) {
                hs += extra;
                nb = hs / bslots;
                pl->hash_slots = hs;
                pl->lds += 2u * extra;
                if (pl->lds > 65536)
                    HIP_TRY(hipFuncSetAttribute(kp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl->lds));
            }
        }
    }
    pl->grid = std::min<uint32_t>(pl->nq, (uint32_t)per_cu * (uint32_t)h->num_cu);
    if (pl->grid == 0) pl->grid = 1;
    geo->Lcap = Lcap;
    geo->nb_cap = nb_cap;
    geo->Mcap = Mcap;
    geo->nb = nb;
    geo->bslots = bslots;
    geo->id_bits = id_bits;
    geo->norm = norm;
    if (norm) {
        rc = ensure_norms(h);
        if (rc) return rc;
    }
    return PANNS_OK;
}

// Device buffers a plan's kernel reads and writes.
struct PlanBufs {
    const uint32_t *qids = nullptr, *qL = nullptr, *qK = nullptr, *init_ids = nullptr;
    uint32_t *work_counter = nullptr;
    uint32_t *out_ids = nullptr, *out_stats = nullptr, *out_qids = nullptr;
    float *out_dists = nullptr, *out_qdists = nullptr;
    uint8_t *out_qchk = nullptr;
    uint32_t *bitmaps = nullptr, *spill_log = nullptr;
    uint64_t bitmap_words = 0;
    uint32_t log_cap = 0;
};

void fill_args(panns_plan *pl, const Geometry &geo, const PlanBufs &b);

int build_plan(panns_plan *pl, const uint32_t *query_ids, const uint32_t *Lq, const uint32_t *Kq,
               const uint32_t *init_ids, uint32_t n_init)
{
    panns_index *h = pl->h;
    Geometry geo;
    int rc = plan_geometry(pl, query_ids, Lq, Kq, init_ids, n_init, &geo);
    if (rc) return rc;
    rc = ensure_scratch(h, pl->slot0 + pl->grid);
    if (rc) return rc;

    // work items in locality order (plain searches of batches big enough)
    std::vector<uint32_t> oq, oL, oK;
    pl->perm.clear();
    if (pl->locality && pl->nq >= kLocalityMin && !pl->init_queue.p && !pl->trace_cap) {
        rc = locality_order(h, query_ids, pl->nq, pl->perm);
        if (rc) return rc;
        oq.resize(pl->nq);
        for (uint32_t i = 0; i < pl->nq; ++i) oq[i] = query_ids[pl->perm[i]];
        query_ids = oq.data();
        if (Lq) {
            oL.resize(pl->nq);
            for (uint32_t i = 0; i < pl->nq; ++i) oL[i] = Lq[pl->perm[i]];
            Lq = oL.data();
        }
        if (Kq) {
            oK.resize(pl->nq);
            for (uint32_t i = 0; i < pl->nq; ++i) oK[i] = Kq[pl->perm[i]];
            Kq = oK.data();
        }
    }
    HIP_TRY(pl->qids.ensure(pl->nq));
    HIP_TRY(hipMemcpyAsync(pl->qids.p, query_ids, pl->nq * 4ull, hipMemcpyHostToDevice, h->stream));
    if (Lq) {
        HIP_TRY(pl->qL.ensure(pl->nq));
        HIP_TRY(hipMemcpyAsync(pl->qL.p, Lq, pl->nq * 4ull, hipMemcpyHostToDevice, h->stream));
    }
    if (Kq) {
        HIP_TRY(pl->qK.ensure(pl->nq));
        HIP_TRY(hipMemcpyAsync(pl->qK.p, Kq, pl->nq * 4ull, hipMemcpyHostToDevice, h->stream));
    }
    if (!pl->init_queue.p) {
        HIP_TRY(pl->init_ids.ensure(pl->maxL));
        HIP_TRY(hipMemcpyAsync(pl->init_ids.p, init_ids, pl->maxL * 4ull, hipMemcpyHostToDevice, h->stream));
    }
    HIP_TRY(pl->out_ids.ensure((size_t)pl->nq * pl->maxK));
    HIP_TRY(pl->out_dists.ensure((size_t)pl->nq * pl->maxK));
    HIP_TRY(pl->out_stats.ensure((size_t)pl->nq * 5));
    if (pl->trace_cap) {
        HIP_TRY(pl->trace_feat.ensure((size_t)pl->nq * pl->trace_cap * panns::kFeat));
        HIP_TRY(pl->trace_topk.ensure((size_t)pl->nq * pl->trace_cap));
        if (pl->trace_gt_k) {
            HIP_TRY(pl->trace_hits.ensure((size_t)pl->nq * pl->trace_cap));
            HIP_TRY(pl->trace_ndist.ensure((size_t)pl->nq * pl->trace_cap));
        }
    }
#ifdef PANNS_PHASE_TIMERS
    HIP_TRY(pl->phase_cycles.ensure((size_t)pl->grid * panns::kPhases));
    HIP_TRY(hipMemsetAsync(pl->phase_cycles.p, 0, (size_t)pl->grid * panns::kPhases * 8, h->stream));
#endif
    if (pl->want_queue) {
        HIP_TRY(pl->out_qids.ensure((size_t)pl->nq * pl->maxL));
        HIP_TRY(pl->out_qdists.ensure((size_t)pl->nq * pl->maxL));
        HIP_TRY(pl->out_qchk.ensure((size_t)pl->nq * pl->maxL));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));

    PlanBufs b;
    b.qids = pl->qids.p;
    b.qL = Lq ? pl->qL.p : nullptr;
    b.qK = Kq ? pl->qK.p : nullptr;
    b.init_ids = pl->init_ids.p;
    b.work_counter = h->counter.p + (size_t)pl->counter_idx * kRanges;
    b.out_ids = pl->out_ids.p;
    b.out_dists = pl->out_dists.p;
    b.out_stats = pl->out_stats.p;
    if (pl->want_queue) {
        b.out_qids = pl->out_qids.p;
        b.out_qdists = pl->out_qdists.p;
        b.out_qchk = pl->out_qchk.p;
    }
    b.bitmaps = h->bitmaps.p + (size_t)pl->slot0 * h->bitmap_words;
    b.bitmap_words = h->bitmap_words;
    b.spill_log = h->spill_log.p + (size_t)pl->slot0 * kLogCap;
    b.log_cap = kLogCap;
    fill_args(pl, geo, b);
    return PANNS_OK;
}

// The kernel arguments of a plan from its geometry and device buffers.
void fill_args(panns_plan *pl, const Geometry &geo, const PlanBufs &b)
{
    panns_index *h = pl->h;
    const panns_params &p = pl->p;
    panns::KArgs &a = pl->args;
    a = panns::KArgs{};
    a.base = h->base.p;
    a.norms = geo.norm ? h->norms.p : nullptr;
    a.adj = h->adj.p;
    a.n = h->n;
    a.stride_log2 = h->stride_log2;
    a.queries = h->queries.p;
    a.qids = b.qids;
    a.qL = b.qL;
    a.qK = b.qK;
    a.n_work = pl->nq;
    a.L = p.L;
    a.K = p.K;
    a.M = p.M;
    a.mode = p.mode;
    a.Kstride = pl->maxK;
    a.Lstride = pl->maxL;
    a.init_ids = b.init_ids;
    a.work_counter = b.work_counter;
    a.n_ranges = kRanges;
    a.out_ids = b.out_ids;
    a.out_dists = b.out_dists;
    a.out_qids = b.out_qids;
    a.out_qdists = b.out_qdists;
    a.out_qchk = b.out_qchk;
    a.out_stats = b.out_stats;
    a.bitmaps = b.bitmaps;
    a.bitmap_words = b.bitmap_words;
    a.spill_log = b.spill_log;
    a.log_cap = b.log_cap;
    a.Lcap = geo.Lcap;
    a.ovf_slots = pl->shared_bitmap ? 0u : pl->overflow_slots;
    a.tbl_words = pl->hash_slots / 2u + a.ovf_slots;
    a.tbl_buckets = geo.nb;
    a.tbl_bshift = geo.bslots == 8u ? 2u : geo.bslots == 4u ? 1u : 0u;
    a.tbl_magic = (uint32_t)((1ull << 32) / geo.nb);
    a.id_bits = geo.id_bits;
    a.use_table = pl->shared_bitmap || pl->no_table ? 0u : 1u;
    a.Mcap = geo.Mcap;
    a.nb_cap = geo.nb_cap;
    a.group_picks = std::max<uint32_t>(1u, geo.nb_cap >> h->stride_log2);
    a.group_rows = pl->rows_per_wave ? pl->rows_per_wave : 96u;
    a.num_threads = p.num_threads;
    a.pend_cap = pl->pend_cap;
    a.init_queue = pl->init_queue.p;
    a.shared_bitmap = pl->shared_bitmap;
    a.stop_learned = p.stop == PANNS_STOP_LEARNED ? 1u : 0u;
    a.stop_threshold = p.stop_threshold;
    a.stop_min_iter = p.stop_min_iter;
    a.stop_every = std::max<uint32_t>(1u, p.stop_every);
    if (h->has_model) {
        const float *m = h->model.p;
        a.model.w1 = m;
        a.model.b1 = m + 512;
        a.model.w2 = m + 576;
        a.model.b2 = m + 576 + 4096;
        a.model.w3 = m + 576 + 4096 + 64;
        a.model.b3 = m + 576 + 4096 + 128;
        a.model.mu = m + 4804;
        a.model.sd = m + 4812;
    }
    a.trace_feat = pl->trace_cap ? pl->trace_feat.p : nullptr;
    a.trace_topk = pl->trace_cap ? pl->trace_topk.p : nullptr;
    a.trace_cap = pl->trace_cap;
    a.trace_gt = pl->trace_gt_k ? pl->trace_gt.p : nullptr;
    a.trace_gt_k = pl->trace_gt_k;
    a.trace_hits = pl->trace_gt_k ? pl->trace_hits.p : nullptr;
    a.trace_ndist = pl->trace_gt_k ? pl->trace_ndist.p : nullptr;
    a.phase_cycles = pl->phase_cycles.p;
    pl->input_gen = h->input_gen;
}

// Points a plan's kernel arguments at the handle's current scratch (another
// plan may have grown, and so moved, it since this one was built).
int bind_scratch(panns_plan *pl)
{
    panns_index *h = pl->h;
    int rc = ensure_scratch(h, pl->slot0 + pl->grid);
    if (rc) return rc;
    panns::KArgs &a = pl->args;
    a.bitmaps = h->bitmaps.p + (size_t)pl->slot0 * h->bitmap_words;
    a.spill_log = h->spill_log.p + (size_t)pl->slot0 * kLogCap;
    a.work_counter = h->counter.p + (size_t)pl->counter_idx * kRanges;
    return PANNS_OK;
}

hipError_t launch_one(const panns_plan *pl, hipStream_t st)
{
    LaunchFn fn;
    KernelPtr kp;
    select_kernel(pl->h->dim, pl->args.stop_learned != 0, pl->args.norms != nullptr, pl->args.trace_feat != nullptr,
                  pl->waves, pl->pss, pl->pend_cap != 0, &fn, &kp);
    return fn(pl->args, pl->grid, pl->lds, st);
}

// Enqueues a bucketed batch on the handle's stream (which may be capturing):
// one counter reset for all buckets, then the bucket launches, in order or
// as side-by-side branches joined back into the stream.
int launch_parts(panns_plan *pl)
{
    panns_index *h = pl->h;
    const uint32_t nb = (uint32_t)pl->parts.size();
    hipLaunchKernelGGL(panns::reset_counter_kernel, dim3(1), dim3(64), 0, h->stream, h->counter.p, nb * kRanges);
    HIP_TRY(hipGetLastError());
    if (!pl->opt.concurrent) {
        for (auto &part : pl->parts) HIP_TRY(launch_one(part.get(), h->stream));
        return PANNS_OK;
    }
    HIP_TRY(hipEventRecord(pl->fork_ev, h->stream));
    for (uint32_t b = 0; b < nb; ++b) {
        HIP_TRY(hipStreamWaitEvent(pl->side[b], pl->fork_ev, 0));
        HIP_TRY(launch_one(pl->parts[b].get(), pl->side[b]));
        HIP_TRY(hipEventRecord(pl->side_ev[b], pl->side[b]));
    }
    for (uint32_t b = 0; b < nb; ++b) HIP_TRY(hipStreamWaitEvent(h->stream, pl->side_ev[b], 0));
    return PANNS_OK;
}

int run_plan(panns_plan *pl, int sync, float *ms)
{
    panns_index *h = pl->h;
    if (pl->input_gen != h->input_gen)
        return fail(PANNS_ERR_STATE, "the plan was built before the base set, queries, graph or distance changed; "
                                     "create it again");
    if (pl->parts.empty()) {
        int rc = bind_scratch(pl);
        if (rc) return rc;
        hipLaunchKernelGGL(panns::reset_counter_kernel, dim3(1), dim3(64), 0, h->stream,
                           h->counter.p + (size_t)pl->counter_idx * kRanges, kRanges);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(h->ev0, h->stream));
        HIP_TRY(launch_one(pl, h->stream));
        HIP_TRY(hipEventRecord(h->ev1, h->stream));
    } else {
        for (auto &part : pl->parts) {
            int rc = bind_scratch(part.get());
            if (rc) return rc;
        }
        if (pl->opt.use_graph && (!pl->gexec || pl->graph_gen != h->scratch_gen)) {
            if (pl->gexec) HIP_TRY(hipGraphExecDestroy(pl->gexec));
            if (pl->graph) HIP_TRY(hipGraphDestroy(pl->graph));
            pl->gexec = nullptr;
            pl->graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            int rc = launch_parts(pl);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(h->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            HIP_TRY(e);
            pl->graph = g;
            HIP_TRY(hipGraphInstantiate(&pl->gexec, pl->graph, nullptr, nullptr, 0));
            pl->graph_gen = h->scratch_gen;
        }
        HIP_TRY(hipEventRecord(h->ev0, h->stream));
        if (pl->opt.use_graph) {
            HIP_TRY(hipGraphLaunch(pl->gexec, h->stream));
        } else {
            int rc = launch_parts(pl);
            if (rc) return rc;
        }
        HIP_TRY(hipEventRecord(h->ev1, h->stream));
        for (auto &part : pl->parts) part->ran = true;
    }
    pl->ran = true;
    if (sync) {
        HIP_TRY(hipEventSynchronize(h->ev1));
        if (ms) HIP_TRY(hipEventElapsedTime(ms, h->ev0, h->ev1));
    }
    return PANNS_OK;
}

// Adds n queries' per-query stats ([ndist, nexp, niter, spills, adjacency
// slots] each) to the counters, count_merge_ as the reference counts it.
void add_counters(const panns_plan *pl, const uint32_t *st, uint32_t n, panns_counters *cnt)
{
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t *s = &st[(size_t)i * 5];
        cnt->distance_computations += s[0];
        cnt->expansions += s[1];
        if (pl->p.mode == PANNS_MODE_TOP_M) {
            cnt->merges += s[2];
        } else if (pl->pss) {
            // rounds of pick, expansion and merge, counted by the kernel
            cnt->merges += s[2];
        } else if (pl->p.mode == PANNS_MODE_PSS_V5) {
            // PSS5LG.cpp:1543-1673 at T=1: one sequential expansion, then
            // rounds of up to X expansions, one count_merge_ per round.
            const uint32_t X = pl->p.M;
            if (s[1] > 1) cnt->merges += (s[1] - 1 + X - 1) / X;
        }
        cnt->visited_spills += s[3];
        cnt->adjacency_slots += s[4];
    }
}

int fetch_single(panns_plan *pl, panns_outputs *out, panns_counters *cnt, float kernel_ms)
{
    panns_index *h = pl->h;
    if (!pl->ran) return fail(PANNS_ERR_STATE, "plan has not run");
    HIP_TRY(hipStreamSynchronize(h->stream));
    std::vector<uint32_t> st((size_t)pl->nq * 5);
    HIP_TRY(hipMemcpy(st.data(), pl->out_stats.p, st.size() * 4, hipMemcpyDeviceToHost));
    if (out) {
        // rows come back in work-item order; perm maps them to the caller's
        const std::vector<uint32_t> &pm = pl->perm;
        auto row = [&](uint32_t i) -> size_t { return pm.empty() ? i : pm[i]; };
        auto fetch_rows = [&](void *dst, const void *src, size_t row_bytes) -> hipError_t {
            if (pm.empty()) return hipMemcpy(dst, src, (size_t)pl->nq * row_bytes, hipMemcpyDeviceToHost);
            std::vector<char> tmp((size_t)pl->nq * row_bytes);
            hipError_t e = hipMemcpy(tmp.data(), src, tmp.size(), hipMemcpyDeviceToHost);
            if (e != hipSuccess) return e;
            for (uint32_t i = 0; i < pl->nq; ++i)
                std::memcpy(static_cast<char *>(dst) + row(i) * row_bytes, tmp.data() + (size_t)i * row_bytes,
                            row_bytes);
            return hipSuccess;
        };
        if (out->ids) HIP_TRY(fetch_rows(out->ids, pl->out_ids.p, (size_t)pl->maxK * 4));
        if (out->dists) HIP_TRY(fetch_rows(out->dists, pl->out_dists.p, (size_t)pl->maxK * 4));
        if (pl->want_queue) {
            if (out->queue_ids) HIP_TRY(fetch_rows(out->queue_ids, pl->out_qids.p, (size_t)pl->maxL * 4));
            if (out->queue_dists) HIP_TRY(fetch_rows(out->queue_dists, pl->out_qdists.p, (size_t)pl->maxL * 4));
            if (out->queue_checked) HIP_TRY(fetch_rows(out->queue_checked, pl->out_qchk.p, (size_t)pl->maxL));
        }
        for (uint32_t i = 0; i < pl->nq; ++i) {
            if (out->ndist) out->ndist[row(i)] = st[(size_t)i * 5 + 0];
            if (out->nexp) out->nexp[row(i)] = st[(size_t)i * 5 + 1];
            if (out->niter) out->niter[row(i)] = st[(size_t)i * 5 + 2];
        }
    }
    if (cnt) {
        add_counters(pl, st.data(), pl->nq, cnt);
        cnt->kernel_ms += kernel_ms;
    }
    return PANNS_OK;
}

// Results of a bucketed batch in the caller's order: each bucket's rows are
// fetched and scattered; ids past a bucket's largest K read 0xFFFFFFFF and
// distances +inf, as in a single plan.
int fetch_parts(panns_plan *pl, panns_outputs *out, panns_counters *cnt, float kernel_ms)
{
    if (!pl->ran) return fail(PANNS_ERR_STATE, "plan has not run");
    HIP_TRY(hipStreamSynchronize(pl->h->stream));
    const uint32_t K = pl->maxK, L = pl->maxL;
    const float inf = std::numeric_limits<float>::infinity();
    for (size_t b = 0; b < pl->parts.size(); ++b) {
        panns_plan *pt = pl->parts[b].get();
        const std::vector<uint32_t> &rows = pl->part_rows[b];
        const uint32_t n = pt->nq, kb = pt->maxK, lb = pt->maxL;
        std::vector<uint32_t> ids, qids, nd, ne, ni;
        std::vector<float> d, qd;
        std::vector<uint8_t> qc;
        panns_outputs o{};
        const bool q = out && pl->opt.want_queue;
        if (out) {
            if (out->ids) { ids.resize((size_t)n * kb); o.ids = ids.data(); }
            if (out->dists) { d.resize((size_t)n * kb); o.dists = d.data(); }
            if (q && out->queue_ids) { qids.resize((size_t)n * lb); o.queue_ids = qids.data(); }
            if (q && out->queue_dists) { qd.resize((size_t)n * lb); o.queue_dists = qd.data(); }
            if (q && out->queue_checked) { qc.resize((size_t)n * lb); o.queue_checked = qc.data(); }
            if (out->ndist) { nd.resize(n); o.ndist = nd.data(); }
            if (out->nexp) { ne.resize(n); o.nexp = ne.data(); }
            if (out->niter) { ni.resize(n); o.niter = ni.data(); }
        }
        int rc = fetch_single(pt, out ? &o : nullptr, cnt, 0.f);
        if (rc) return rc;
        if (!out) continue;
        for (uint32_t i = 0; i < n; ++i) {
            const size_t r = rows[i];
            for (uint32_t j = 0; j < K; ++j) {
                if (out->ids) out->ids[r * K + j] = j < kb ? ids[(size_t)i * kb + j] : 0xFFFFFFFFu;
                if (out->dists) out->dists[r * K + j] = j < kb ? d[(size_t)i * kb + j] : inf;
            }
            if (q)
                for (uint32_t j = 0; j < L; ++j) {
                    if (out->queue_ids) out->queue_ids[r * L + j] = j < lb ? qids[(size_t)i * lb + j] : 0xFFFFFFFFu;
                    if (out->queue_dists) out->queue_dists[r * L + j] = j < lb ? qd[(size_t)i * lb + j] : inf;
                    if (out->queue_checked) out->queue_checked[r * L + j] = j < lb ? qc[(size_t)i * lb + j] : 0;
                }
            if (out->ndist) out->ndist[r] = nd[i];
            if (out->nexp) out->nexp[r] = ne[i];
            if (out->niter) out->niter[r] = ni[i];
        }
    }
    if (cnt) cnt->kernel_ms += kernel_ms;
    return PANNS_OK;
}

int fetch_plan(panns_plan *pl, panns_outputs *out, panns_counters *cnt, float kernel_ms)
{
    return pl->parts.empty() ? fetch_single(pl, out, cnt, kernel_ms) : fetch_parts(pl, out, cnt, kernel_ms);
}

int upload_graph(panns_index *h, uint32_t width, uint32_t ep, uint32_t n, uint32_t stride_in,
                 const uint32_t *adj, const uint32_t *deg)
{
    if (n != h->n)
        return fail(PANNS_ERR_RANGE, "graph has " + std::to_string(n) + " vertices, base set has " +
                                         std::to_string(h->n));
    if (ep >= n) return fail(PANNS_ERR_RANGE, "entry point out of range");
    uint32_t md = 0;
    uint64_t ne = 0;
    for (uint32_t v = 0; v < n; ++v) {
        if (deg[v] > stride_in) return fail(PANNS_ERR_RANGE, "degree exceeds the row stride");
        md = std::max(md, deg[v]);
        ne += deg[v];
        for (uint32_t i = 0; i < deg[v]; ++i)
            if (adj[(size_t)v * stride_in + i] >= n) return fail(PANNS_ERR_RANGE, "neighbour id >= N");
    }
    const uint32_t stride = next_pow2(std::max<uint32_t>(32u, md));
    std::vector<uint32_t> rows((size_t)n * stride, panns::kSent);
    for (uint32_t v = 0; v < n; ++v)
        std::memcpy(&rows[(size_t)v * stride], adj + (size_t)v * stride_in, deg[v] * 4ull);
    ++h->input_gen;
    HIP_TRY(h->adj.ensure(rows.size()));
    HIP_TRY(hipMemcpy(h->adj.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice));
    h->width = width;
    h->ep = ep;
    h->max_deg = md;
    h->num_e = ne;
    h->stride_log2 = log2u(stride);
    h->ep_row.assign(adj + (size_t)ep * stride_in, adj + (size_t)ep * stride_in + deg[ep]);
    h->has_graph = true;
    return PANNS_OK;
}

}  // namespace

namespace {
template <typename T>
void swap_buf(DevBuf<T> &a, DevBuf<T> &b)
{
    std::swap(a.p, b.p);
    std::swap(a.n, b.n);
}
void swap_oneshot(panns_plan *pl, panns_index::OneShot &o)
{
    swap_buf(pl->qids, o.qids);
    swap_buf(pl->init_ids, o.init_ids);
    swap_buf(pl->out_ids, o.out_ids);
    swap_buf(pl->out_stats, o.out_stats);
    swap_buf(pl->out_qids, o.out_qids);
    swap_buf(pl->out_dists, o.out_dists);
    swap_buf(pl->out_qdists, o.out_qdists);
    swap_buf(pl->out_qchk, o.out_qchk);
}
// Lends the handle's one-shot buffers to a plan and takes them back when the
// call ends, on every return path.  Work left on the handle's stream by a
// failed call stays ordered before the next call's use of the buffers.
struct OneShotLease {
    panns_plan *pl;
    panns_index *h;
    OneShotLease(panns_plan *p, panns_index *x) : pl(p), h(x) { swap_oneshot(pl, h->oneshot); }
    ~OneShotLease() { swap_oneshot(pl, h->oneshot); }
};
}  // namespace

extern "C" {

const char *panns_last_error(void) { return g_err.c_str(); }

int panns_create(int device, panns_index **out)
{
    if (!out) return fail(PANNS_ERR_ARG, "null out");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(PANNS_ERR_DEVICE, "no such HIP device");
    HIP_TRY(hipSetDevice(device));
    std::unique_ptr<panns_index> h(new panns_index());
    h->device = device;
    HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&h->ev0));
    HIP_TRY(hipEventCreate(&h->ev1));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    h->num_cu = prop.multiProcessorCount;
    h->lds_per_cu = (uint32_t)prop.maxSharedMemoryPerMultiProcessor;
    // this device's allocation-free scratch (panns_search_noalloc), looked
    // up once per device; a failure here is reported by that entry point
    uint8_t *scr = nullptr;
    (void)scratch_address(device, &scr);
    g_err.clear();
    *out = h.release();
    return PANNS_OK;
}

int panns_destroy(panns_index *h)
{
    if (!h) return PANNS_OK;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    for (panns_plan *pl : h->plans) {
        pl->h = nullptr;
        for (auto &part : pl->parts) part->h = nullptr;
    }
    h->plans.clear();
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PANNS_OK;
}

int panns_set_base(panns_index *h, const float *data, uint32_t n, uint32_t dim)
{
    if (!h || !data || !n || !dim) return fail(PANNS_ERR_ARG, "bad base set");
    if (h->queries.p && h->dim && dim != h->dim)
        return fail(PANNS_ERR_DIM, "data dimension " + std::to_string(dim) + " is not equal to query dimension " +
                                       std::to_string(h->dim));
    if (dim % 8) return fail(PANNS_ERR_DIM, "dimension must be a multiple of 8 (the reference reads 8-float blocks)");
    if (n >= 0x80000000u) return fail(PANNS_ERR_RANGE, "N must be below 2^31");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(h->base.ensure((size_t)n * dim));
    HIP_TRY(hipMemcpy(h->base.p, data, (size_t)n * dim * 4, hipMemcpyHostToDevice));
    if (h->n != n) h->has_graph = false;
    ++h->input_gen;
    h->n = n;
    h->dim = dim;
    h->has_pivots = false;
    h->norms_valid = false;
    return PANNS_OK;
}

int panns_set_queries(panns_index *h, const float *data, uint32_t nq, uint32_t dim)
{
    if (!h || !data || !nq || !dim) return fail(PANNS_ERR_ARG, "bad query set");
    if (h->n && dim != h->dim)
        return fail(PANNS_ERR_DIM, "query dimension " + std::to_string(dim) + " is not equal to data dimension " +
                                       std::to_string(h->dim));
    if (dim % 8) return fail(PANNS_ERR_DIM, "dimension must be a multiple of 8");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(h->queries.ensure((size_t)nq * dim));
    HIP_TRY(hipMemcpy(h->queries.p, data, (size_t)nq * dim * 4, hipMemcpyHostToDevice));
    ++h->input_gen;
    h->nq = nq;
    h->dim = dim;
    return PANNS_OK;
}

int panns_load_data_load(panns_index *h, const char *path)
{
    if (!h || !path) return fail(PANNS_ERR_ARG, "null argument");
    uint32_t n = 0, d = 0;
    if (pt_fvecs_info(path, &n, &d)) return fail(PANNS_ERR_IO, pt_last_error());
    std::vector<float> buf((size_t)n * d);
    if (pt_fvecs_read(path, buf.data(), n, d)) return fail(PANNS_ERR_FORMAT, pt_last_error());
    return panns_set_base(h, buf.data(), n, d);
}

int panns_load_queries_load(panns_index *h, const char *path)
{
    if (!h || !path) return fail(PANNS_ERR_ARG, "null argument");
    uint32_t n = 0, d = 0;
    if (pt_fvecs_info(path, &n, &d)) return fail(PANNS_ERR_IO, pt_last_error());
    std::vector<float> buf((size_t)n * d);
    if (pt_fvecs_read(path, buf.data(), n, d)) return fail(PANNS_ERR_FORMAT, pt_last_error());
    return panns_set_queries(h, buf.data(), n, d);
}

int panns_load_common_nsg_graph(panns_index *h, const char *path)
{
    if (!h || !path) return fail(PANNS_ERR_ARG, "null argument");
    if (!h->n) return fail(PANNS_ERR_STATE, "load the base set before the graph");
    uint32_t width, ep, n, md;
    uint64_t ne;
    const int rc = pt_nsg_info(path, &width, &ep, &n, &ne, &md);
    if (rc == PT_ERR_IO) return fail(PANNS_ERR_IO, pt_last_error());
    if (rc) return fail(PANNS_ERR_FORMAT, pt_last_error());
    const uint32_t stride = std::max<uint32_t>(md, 1u);
    std::vector<uint32_t> adj((size_t)n * stride), deg(n);
    if (pt_nsg_read(path, n, stride, adj.data(), deg.data())) return fail(PANNS_ERR_FORMAT, pt_last_error());
    int rc2;
    HIP_TRY(hipSetDevice(h->device));
    rc2 = upload_graph(h, width, ep, n, stride, adj.data(), deg.data());
    if (rc2 == PANNS_OK) h->distance = PANNS_DISTANCE_L2;
    return rc2;
}

int panns_load_nsg_graph(panns_index *h, const char *path)
{
    int rc = panns_load_common_nsg_graph(h, path);
    if (rc) return rc;
    return panns_set_distance(h, PANNS_DISTANCE_NORM_DOT);
}

int panns_set_distance(panns_index *h, int distance)
{
    if (!h) return fail(PANNS_ERR_ARG, "null handle");
    if (distance != PANNS_DISTANCE_L2 && distance != PANNS_DISTANCE_NORM_DOT)
        return fail(PANNS_ERR_ARG, "unknown distance");
    if (distance == PANNS_DISTANCE_NORM_DOT) {
        if (!h->n) return fail(PANNS_ERR_STATE, "load the base set before choosing the norm distance");
        HIP_TRY(hipSetDevice(h->device));
        int rc = ensure_norms(h);
        if (rc) return rc;
    }
    h->distance = (uint32_t)distance;
    return PANNS_OK;
}

int panns_get_distance(panns_index *h, int *distance)
{
    if (!h || !distance) return fail(PANNS_ERR_ARG, "null argument");
    *distance = (int)h->distance;
    return PANNS_OK;
}

int panns_get_norms(panns_index *h, float *out)
{
    if (!h || !out) return fail(PANNS_ERR_ARG, "null argument");
    if (!h->n) return fail(PANNS_ERR_STATE, "base set not loaded");
    HIP_TRY(hipSetDevice(h->device));
    int rc = ensure_norms(h);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(out, h->norms.p, (size_t)h->n * 4, hipMemcpyDeviceToHost));
    return PANNS_OK;
}

int panns_set_graph(panns_index *h, uint32_t width, uint32_t ep, uint32_t n, uint32_t stride,
                    const uint32_t *adj, const uint32_t *deg)
{
    if (!h || !adj || !deg) return fail(PANNS_ERR_ARG, "null argument");
    if (!h->n) return fail(PANNS_ERR_STATE, "load the base set before the graph");
    HIP_TRY(hipSetDevice(h->device));
    return upload_graph(h, width, ep, n, stride, adj, deg);
}

int panns_get_info(panns_index *h, uint32_t *num_v, uint32_t *dim, uint32_t *num_queries,
                   uint32_t *width, uint32_t *ep, uint64_t *num_e)
{
    if (!h) return fail(PANNS_ERR_ARG, "null handle");
    if (num_v) *num_v = h->n;
    if (dim) *dim = h->dim;
    if (num_queries) *num_queries = h->nq;
    if (width) *width = h->width;
    if (ep) *ep = h->ep;
    if (num_e) *num_e = h->num_e;
    return PANNS_OK;
}

int panns_prepare_init_ids(panns_index *h, uint32_t L, uint32_t *init_ids)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!init_ids) return fail(PANNS_ERR_ARG, "null init_ids");
    if (L > h->n) return fail(PANNS_ERR_ARG, "L > number of base vectors");
    // PSS5LG.cpp:356-410
    std::vector<uint64_t> sel((h->n + 63) / 64, 0);
    auto get = [&](uint32_t v) { return (sel[v >> 6] >> (v & 63)) & 1u; };
    auto set = [&](uint32_t v) { sel[v >> 6] |= 1ull << (v & 63); };
    uint32_t end = 0;
    for (size_t e = 0; e < h->ep_row.size() && end < L; ++e) {
        const uint32_t v = h->ep_row[e];
        if (get(v)) continue;
        set(v);
        init_ids[end++] = v;
    }
    uint32_t tmp = h->ep + 1;
    while (end < L) {
        if (tmp == h->n) tmp = 0;
        const uint32_t v = tmp++;
        if (get(v)) continue;
        set(v);
        init_ids[end++] = v;
    }
    return PANNS_OK;
}

int panns_plan_create(panns_index *h, const panns_params *p, const uint32_t *query_ids, uint32_t nq,
                      const uint32_t *L_per_query, const uint32_t *K_per_query, const uint32_t *init_ids,
                      uint32_t n_init, panns_plan **out)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!p || !query_ids || !init_ids || !out || nq == 0) return fail(PANNS_ERR_ARG, "bad plan arguments");
    if (!h->nq) return fail(PANNS_ERR_STATE, "queries not loaded");
    HIP_TRY(hipSetDevice(h->device));
    std::unique_ptr<panns_plan> pl(new panns_plan());
    pl->h = h;
    pl->p = *p;
    pl->nq = nq;
    pl->want_queue = true;
    rc = build_plan(pl.get(), query_ids, L_per_query, K_per_query, init_ids, n_init);
    if (rc) return rc;
    *out = hand_out(h, pl);
    return PANNS_OK;
}

int panns_plan_create_ex(panns_index *h, const panns_params *p, const uint32_t *query_ids, uint32_t nq,
                         const uint32_t *L_per_query, const uint32_t *K_per_query, const uint32_t *init_ids,
                         uint32_t n_init, const panns_plan_options *opt, panns_plan **out)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!p || !query_ids || !init_ids || !out || nq == 0) return fail(PANNS_ERR_ARG, "bad plan arguments");
    if (!h->nq) return fail(PANNS_ERR_STATE, "queries not loaded");
    panns_plan_options o{};
    if (opt) o = *opt;
    if (o.n_buckets == 0) o.n_buckets = 1;
    if (o.n_buckets > PANNS_MAX_BUCKETS) return fail(PANNS_ERR_ARG, "at most 8 buckets");
    if (o.table_eighths > 8) return fail(PANNS_ERR_ARG, "table_eighths must be 0..8");
    if (o.rows_per_wave && o.rows_per_wave != 32 && o.rows_per_wave != 64 && o.rows_per_wave != 96)
        return fail(PANNS_ERR_ARG, "rows_per_wave must be 0, 32, 64 or 96");
    if (o.overflow_slots &&
        ((o.overflow_slots & (o.overflow_slots - 1u)) || o.overflow_slots < 8u || o.overflow_slots > 4096u))
        return fail(PANNS_ERR_ARG, "overflow_slots must be 0 or a power of two in [8, 4096]");
    if (o.deferred_merge > 2) return fail(PANNS_ERR_ARG, "deferred_merge must be 0, 1 or 2");
    if (o.visited_table > 2) return fail(PANNS_ERR_ARG, "visited_table must be 0, 1 or 2");
    HIP_TRY(hipSetDevice(h->device));
    std::unique_ptr<panns_plan> pl(new panns_plan());
    pl->h = h;
    pl->p = *p;
    pl->nq = nq;
    if (o.waves) pl->p.waves = o.waves;
    pl->opt = o;
    pl->want_queue = o.want_queue != 0;
    pl->table_eighths = o.table_eighths;
    pl->shape_given = opt != nullptr;
    pl->rows_per_wave = o.rows_per_wave;
    pl->overflow_slots = o.overflow_slots;
    pl->deferred_merge = o.deferred_merge;
    pl->visited_table = o.visited_table;
    if (o.n_buckets == 1 && !o.use_graph && !o.concurrent) {
        rc = build_plan(pl.get(), query_ids, L_per_query, K_per_query, init_ids, n_init);
        if (rc) return rc;
        *out = hand_out(h, pl);
        return PANNS_OK;
    }
    // L and K of every query, checked as a single plan would
    std::vector<uint32_t> Lall(nq), Kall(nq);
    pl->maxL = 0;
    pl->maxK = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        Lall[i] = L_per_query ? L_per_query[i] : p->L;
        Kall[i] = K_per_query ? K_per_query[i] : p->K;
        rc = check_params(h, &pl->p, Lall[i], Kall[i]);
        if (rc) return rc;
        pl->maxL = std::max(pl->maxL, Lall[i]);
        pl->maxK = std::max(pl->maxK, Kall[i]);
    }
    // bucket bounds: given, or equal summed L over the queries sorted by L
    std::vector<uint32_t> bounds;
    bool given = false;
    for (uint32_t b = 0; b < o.n_buckets; ++b) given = given || o.bucket_max_L[b] != 0;
    if (given) {
        for (uint32_t b = 0; b < o.n_buckets; ++b) {
            if (b > 0 && o.bucket_max_L[b] <= o.bucket_max_L[b - 1])
                return fail(PANNS_ERR_ARG, "bucket_max_L must be ascending");
            bounds.push_back(o.bucket_max_L[b]);
        }
        if (bounds.back() < pl->maxL) return fail(PANNS_ERR_ARG, "the last bucket bound is below the largest L");
    } else {
        std::vector<uint32_t> sorted(Lall);
        std::sort(sorted.begin(), sorted.end());
        double total = 0;
        for (uint32_t v : sorted) total += v;
        double acc = 0;
        uint32_t next = 1;
        for (uint32_t v : sorted) {
            acc += v;
            while (next < o.n_buckets && acc >= total * next / o.n_buckets) {
                if (bounds.empty() || bounds.back() < v) bounds.push_back(v);
                ++next;
            }
        }
        if (bounds.empty() || bounds.back() < pl->maxL) bounds.push_back(pl->maxL);
    }
    // queries per bucket, longest first; buckets with the largest L first
    std::vector<std::vector<uint32_t>> members(bounds.size());
    std::vector<uint32_t> order(nq);
    for (uint32_t i = 0; i < nq; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return Lall[x] > Lall[y]; });
    for (uint32_t i : order) {
        const size_t b = std::lower_bound(bounds.begin(), bounds.end(), Lall[i]) - bounds.begin();
        members[b].push_back(i);
    }
    uint32_t slot = 0, counter = 0;
    const size_t n_parts =
        std::count_if(members.begin(), members.end(), [](const std::vector<uint32_t> &m) { return !m.empty(); });
    for (size_t bi = bounds.size(); bi-- > 0;) {
        const std::vector<uint32_t> &m = members[bi];
        if (m.empty()) continue;
        std::unique_ptr<panns_plan> part(new panns_plan());
        part->h = h;
        part->p = pl->p;
        part->nq = (uint32_t)m.size();
        part->want_queue = pl->want_queue;
        part->table_eighths = o.table_eighths;
        part->rows_per_wave = o.rows_per_wave;
        part->overflow_slots = o.overflow_slots;
        part->deferred_merge = o.deferred_merge;
        part->visited_table = o.visited_table;
        part->shape_given = pl->shape_given;
        part->slot0 = o.concurrent ? slot : 0u;
        part->shares_cus = o.concurrent && n_parts > 1;
        part->counter_idx = counter++;
        std::vector<uint32_t> q(m.size()), lq(m.size()), kq(m.size());
        for (size_t i = 0; i < m.size(); ++i) {
            q[i] = query_ids[m[i]];
            lq[i] = Lall[m[i]];
            kq[i] = Kall[m[i]];
        }
        rc = build_plan(part.get(), q.data(), lq.data(), kq.data(), init_ids, n_init);
        if (rc) return rc;
        slot += part->grid;
        pl->parts.push_back(std::move(part));
        pl->part_rows.push_back(m);
    }
    if (o.concurrent) {
        HIP_TRY(hipEventCreateWithFlags(&pl->fork_ev, hipEventDisableTiming));
        for (size_t b = 0; b < pl->parts.size(); ++b) {
            hipStream_t st = nullptr;
            HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            pl->side.push_back(st);
            hipEvent_t e = nullptr;
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            pl->side_ev.push_back(e);
        }
    }
    pl->input_gen = h->input_gen;
    pl->grid = 0;
    for (auto &part : pl->parts) {
        pl->grid += part->grid;
        pl->lds = std::max(pl->lds, part->lds);
        pl->hash_slots = std::max(pl->hash_slots, part->no_table ? 0u : part->hash_slots);
    }
    *out = hand_out(h, pl);
    return PANNS_OK;
}

int panns_plan_describe_buckets(panns_plan *plan, uint32_t *n_buckets, uint32_t *max_L, uint32_t *nq,
                                uint32_t *workgroups, uint32_t *lds_bytes, uint32_t *hash_slots)
{
    if (!plan) return fail(PANNS_ERR_ARG, "null plan");
    std::vector<const panns_plan *> ps;
    if (plan->parts.empty()) ps.push_back(plan);
    for (auto &part : plan->parts) ps.push_back(part.get());
    if (n_buckets) *n_buckets = (uint32_t)ps.size();
    for (size_t b = 0; b < ps.size(); ++b) {
        if (max_L) max_L[b] = ps[b]->maxL;
        if (nq) nq[b] = ps[b]->nq;
        if (workgroups) workgroups[b] = ps[b]->grid;
        if (lds_bytes) lds_bytes[b] = ps[b]->lds;
        if (hash_slots) hash_slots[b] = ps[b]->no_table ? 0u : ps[b]->hash_slots;
    }
    return PANNS_OK;
}

int panns_plan_autotune(panns_index *h, const panns_params *p, const uint32_t *query_ids, uint32_t nq,
                        const uint32_t *L_per_query, const uint32_t *K_per_query, const uint32_t *init_ids,
                        uint32_t n_init, uint32_t sample_stride, panns_plan_options *best, float *best_ms)
{
    if (!p || !query_ids || !best || nq == 0) return fail(PANNS_ERR_ARG, "bad autotune arguments");
    const uint32_t stride = sample_stride ? sample_stride : std::max<uint32_t>(1u, nq / 20000u);
    std::vector<uint32_t> sq, sl, sk, Lall(nq);
    for (uint32_t i = 0; i < nq; ++i) Lall[i] = L_per_query ? L_per_query[i] : p->L;
    for (uint32_t i = 0; i < nq; i += stride) {
        sq.push_back(query_ids[i]);
        sl.push_back(Lall[i]);
        if (K_per_query) sk.push_back(K_per_query[i]);
    }
    std::vector<uint32_t> sorted(Lall);
    std::sort(sorted.begin(), sorted.end());
    double total = 0;
    for (uint32_t v : sorted) total += v;
    float best_t = std::numeric_limits<float>::infinity();
    // one candidate: build on the sample, one warm run, best of three
    auto trial = [&](const panns_plan_options &o, float *t) -> int {
        panns_plan *pl = nullptr;
        int rc = panns_plan_create_ex(h, p, sq.data(), (uint32_t)sq.size(), L_per_query ? sl.data() : nullptr,
                                      K_per_query ? sk.data() : nullptr, init_ids, n_init, &o, &pl);
        if (rc) return rc;
        float ms = 0.f;
        *t = std::numeric_limits<float>::infinity();
        rc = run_plan(pl, 1, nullptr);
        for (int r = 0; r < 3 && !rc; ++r) {
            rc = run_plan(pl, 1, &ms);
            *t = std::min(*t, ms);
        }
        panns_plan_destroy(pl);
        return rc;
    };
    auto consider = [&](const panns_plan_options &o) -> int {
        float t = 0.f;
        int rc = trial(o, &t);
        if (rc) return rc;
        if (t < best_t) {
            best_t = t;
            *best = o;
        }
        return PANNS_OK;
    };
    // stage 1: buckets of equal summed L over the whole batch (explicit
    // bounds, so they carry over from the sample), in sequence or side by side
    static const uint32_t kCounts[] = {1, 2, 3, 4, 6, 8};
    for (uint32_t nb : kCounts) {
        panns_plan_options o{};
        uint32_t k = 0;
        double acc = 0;
        uint32_t next = 1;
        for (uint32_t v : sorted) {
            acc += v;
            while (next < nb && acc >= total * next / nb) {
                if (k == 0 || o.bucket_max_L[k - 1] < v) o.bucket_max_L[k++] = v;
                ++next;
            }
        }
        if (k == 0 || o.bucket_max_L[k - 1] < sorted.back()) o.bucket_max_L[k++] = sorted.back();
        if (nb > 1 && k < 2) continue;
        o.n_buckets = k;
        o.use_graph = 1;
        o.table_eighths = 0;  // each bucket's table by its own largest L
        o.waves = 1;
        for (uint32_t conc = 0; conc < (k > 1 ? 2u : 1u); ++conc) {
            o.concurrent = conc;
            int rc = consider(o);
            if (rc) return rc;
        }
    }
    // stage 2: one table size for every bucket (more workgroups per CU and
    // more ids past their buckets as it shrinks), with or without an LDS
    // overflow area for those
    const panns_plan_options s1 = *best;
    for (uint32_t eighths : {8u, 7u, 6u, 5u, 4u, 3u}) {
        for (uint32_t ovf : {0u, 256u, 512u}) {
            panns_plan_options o = s1;
            o.table_eighths = eighths;
            o.overflow_slots = ovf;
            int rc = consider(o);
            if (rc) return rc;
        }
    }
    // stage 3: several waves per query, where the search allows it
    if (p->mode != PANNS_MODE_PSS_V5 && p->stop == PANNS_STOP_FIXED) {
        const panns_plan_options s2 = *best;
        for (uint32_t waves : {2u, 4u}) {
            panns_plan_options o = s2;
            o.waves = waves;
            int rc = consider(o);
            if (rc) return rc;
        }
    }
    // stage 4: neighbour rows gathered per wave and distance step (96 is
    // the default where the kernel's registers allow it; fewer rows in
    // flight mean shorter groups at the tail of an iteration)
    {
        const panns_plan_options s3 = *best;
        for (uint32_t rows : {64u, 32u}) {
            panns_plan_options o = s3;
            o.rows_per_wave = rows;
            int rc = consider(o);
            if (rc) return rc;
        }
    }
    if (best_ms) *best_ms = best_t;
    return PANNS_OK;
}

int panns_plan_run(panns_plan *plan, int sync, float *kernel_ms)
{
    if (!plan) return fail(PANNS_ERR_ARG, "null plan");
    if (!plan->h) return fail(PANNS_ERR_STATE, "the plan's index was destroyed");
    HIP_TRY(hipSetDevice(plan->h->device));
    return run_plan(plan, sync, kernel_ms);
}

int panns_plan_fetch(panns_plan *plan, panns_outputs *out, panns_counters *cnt)
{
    if (!plan) return fail(PANNS_ERR_ARG, "null plan");
    if (!plan->h) return fail(PANNS_ERR_STATE, "the plan's index was destroyed");
    HIP_TRY(hipSetDevice(plan->h->device));
    float ms = 0.f;
    if (plan->ran) {
        HIP_TRY(hipEventSynchronize(plan->h->ev1));
        (void)hipEventElapsedTime(&ms, plan->h->ev0, plan->h->ev1);
    }
    return fetch_plan(plan, out, cnt, ms);
}

int panns_plan_destroy(panns_plan *plan)
{
    if (!plan) return PANNS_OK;
    (void)hipSetDevice(plan->device);
    if (plan->h) {
        (void)hipStreamSynchronize(plan->h->stream);
        plan->h->plans.erase(plan);
    }
    delete plan;
    return PANNS_OK;
}

int panns_plan_describe(panns_plan *plan, uint32_t *workgroups, uint32_t *lds, uint32_t *hash_slots,
                        uint32_t *max_L)
{
    if (!plan) return fail(PANNS_ERR_ARG, "null plan");
    if (workgroups) *workgroups = plan->grid;
    if (lds) *lds = plan->lds;
    if (hash_slots) *hash_slots = plan->no_table ? 0u : plan->hash_slots;
    if (max_L) *max_L = plan->maxL;
    return PANNS_OK;
}

int panns_search(panns_index *h, const panns_params *p, const uint32_t *query_ids, uint32_t nq,
                 const uint32_t *init_ids, uint32_t n_init, panns_outputs *out, panns_counters *cnt)
{
    if (!out || !out->ids) return fail(PANNS_ERR_ARG, "outputs.ids is required");
    panns_plan *pl = nullptr;
    int rc = check_ready(h);
    if (rc) return rc;
    if (!p || !query_ids || !init_ids || nq == 0) return fail(PANNS_ERR_ARG, "bad search arguments");
    if (!h->nq) return fail(PANNS_ERR_STATE, "queries not loaded");
    HIP_TRY(hipSetDevice(h->device));
    std::unique_ptr<panns_plan> own(new panns_plan());
    pl = own.get();
    OneShotLease lease(pl, h);
    pl->h = h;
    pl->p = *p;
    pl->nq = nq;
    pl->want_queue = out->queue_ids || out->queue_dists || out->queue_checked;
    rc = build_plan(pl, query_ids, nullptr, nullptr, init_ids, n_init);
    if (rc) return rc;
    float ms = 0.f;
    rc = run_plan(pl, 1, &ms);
    if (rc) return rc;
    return fetch_plan(pl, out, cnt, ms);
}

int panns_search_noalloc(panns_index *h, const panns_params *p, const uint32_t *query_ids, uint32_t nq,
                         const uint32_t *init_ids, uint32_t n_init, panns_outputs *out, panns_counters *cnt)
{
    if (!out || !out->ids) return fail(PANNS_ERR_ARG, "outputs.ids is required");
    int rc = check_ready(h);
    if (rc) return rc;
    if (!p || !query_ids || !init_ids || nq == 0) return fail(PANNS_ERR_ARG, "bad search arguments");
    if (!h->nq) return fail(PANNS_ERR_STATE, "queries not loaded");
    HIP_TRY(hipSetDevice(h->device));
    if (h->distance == PANNS_DISTANCE_NORM_DOT && !h->norms_valid)
        return fail(PANNS_ERR_STATE, "vertex norms not computed: call panns_set_distance(NORM_DOT) first");
    uint8_t *ptr = nullptr;
    rc = scratch_address(h->device, &ptr);
    if (rc) return rc;
    // the plan lives on the stack and owns no device memory
    panns_plan pl;
    pl.h = h;
    pl.p = *p;
    pl.nq = nq;
    pl.locality = false;
    pl.want_queue = out->queue_ids || out->queue_dists || out->queue_checked;
    Geometry geo;
    rc = plan_geometry(&pl, query_ids, nullptr, nullptr, init_ids, n_init, &geo);
    if (rc) return rc;
    // carve the scratch: fixed buffers first, then as many workgroups'
    // bitmaps and spill logs as fit (at most the occupancy grid)
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t K = pl.maxK, L = pl.maxL;
    const size_t o_cnt = take(kRanges * 4), o_q = take((size_t)nq * 4), o_init = take(L * 4);
    const size_t o_ids = take(nq * K * 4), o_d = take(nq * K * 4), o_st = take((size_t)nq * 5 * 4);
    size_t o_qi = 0, o_qd = 0, o_qc = 0;
    if (pl.want_queue) {
        o_qi = take(nq * L * 4);
        o_qd = take(nq * L * 4);
        o_qc = take(nq * L);
    }
    const uint64_t words = bitmap_words_for(h->n);
    const size_t bm_bytes = ((size_t)words * 4 + 255) & ~(size_t)255, log_bytes = (size_t)kNoAllocLogCap * 4;
    if (off + bm_bytes + log_bytes > kNoAllocBytes)
        return fail(PANNS_ERR_ARG, "this batch needs " + std::to_string(off + bm_bytes + log_bytes) +
                                       " bytes of device scratch; the allocation-free search has " +
                                       std::to_string(kNoAllocBytes) +
                                       " (use fewer queries, a smaller K or L, or no queue outputs)");
    pl.grid = std::min<uint32_t>(pl.grid, (uint32_t)((kNoAllocBytes - off) / (bm_bytes + log_bytes)));
    const size_t o_bm = take((size_t)pl.grid * bm_bytes), o_log = take((size_t)pl.grid * log_bytes);
    hipStream_t st = h->stream;
    HIP_TRY(hipMemcpyAsync(ptr + o_q, query_ids, (size_t)nq * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(ptr + o_init, init_ids, L * 4, hipMemcpyHostToDevice, st));
    // the bitmaps start empty (the kernel leaves them so after every query,
    // but an earlier call's layout may have put other data there)
    const uint64_t n16 = (uint64_t)pl.grid * bm_bytes / 16;
    hipLaunchKernelGGL(panns::zero_kernel, dim3((uint32_t)std::min<uint64_t>((n16 + 255) / 256, 2048)), dim3(256), 0,
                       st, reinterpret_cast<uint4 *>(ptr + o_bm), n16);
    HIP_TRY(hipGetLastError());
    PlanBufs b;
    b.qids = reinterpret_cast<const uint32_t *>(ptr + o_q);
    b.init_ids = reinterpret_cast<const uint32_t *>(ptr + o_init);
    b.work_counter = reinterpret_cast<uint32_t *>(ptr + o_cnt);
    b.out_ids = reinterpret_cast<uint32_t *>(ptr + o_ids);
    b.out_dists = reinterpret_cast<float *>(ptr + o_d);
    b.out_stats = reinterpret_cast<uint32_t *>(ptr + o_st);
    if (pl.want_queue) {
        b.out_qids = reinterpret_cast<uint32_t *>(ptr + o_qi);
        b.out_qdists = reinterpret_cast<float *>(ptr + o_qd);
        b.out_qchk = ptr + o_qc;
    }
    b.bitmaps = reinterpret_cast<uint32_t *>(ptr + o_bm);
    b.bitmap_words = bm_bytes / 4;
    b.spill_log = reinterpret_cast<uint32_t *>(ptr + o_log);
    b.log_cap = kNoAllocLogCap;
    fill_args(&pl, geo, b);
    hipLaunchKernelGGL(panns::reset_counter_kernel, dim3(1), dim3(64), 0, st, b.work_counter, kRanges);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(h->ev0, st));
    HIP_TRY(launch_one(&pl, st));
    HIP_TRY(hipEventRecord(h->ev1, st));
    HIP_TRY(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    // results in the caller's order (no locality reordering here)
    HIP_TRY(hipMemcpy(out->ids, ptr + o_ids, nq * K * 4, hipMemcpyDeviceToHost));
    if (out->dists) HIP_TRY(hipMemcpy(out->dists, ptr + o_d, nq * K * 4, hipMemcpyDeviceToHost));
    if (out->queue_ids) HIP_TRY(hipMemcpy(out->queue_ids, ptr + o_qi, nq * L * 4, hipMemcpyDeviceToHost));
    if (out->queue_dists) HIP_TRY(hipMemcpy(out->queue_dists, ptr + o_qd, nq * L * 4, hipMemcpyDeviceToHost));
    if (out->queue_checked) HIP_TRY(hipMemcpy(out->queue_checked, ptr + o_qc, nq * L, hipMemcpyDeviceToHost));
    // per-query stats through a stack buffer, a chunk of queries at a time
    uint32_t stats[256 * 5];
    for (uint32_t q0 = 0; q0 < nq; q0 += 256) {
        const uint32_t n = std::min<uint32_t>(256u, nq - q0);
        HIP_TRY(hipMemcpy(stats, ptr + o_st + (size_t)q0 * 20, (size_t)n * 20, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) {
            if (out->ndist) out->ndist[q0 + i] = stats[i * 5 + 0];
            if (out->nexp) out->nexp[q0 + i] = stats[i * 5 + 1];
            if (out->niter) out->niter[q0 + i] = stats[i * 5 + 2];
        }
        if (cnt) add_counters(&pl, stats, n, cnt);
    }
    if (cnt) cnt->kernel_ms += ms;
    return PANNS_OK;
}

int panns_subsearch_with_top_m(panns_index *h, uint32_t M_max, uint32_t query_id, uint32_t local_L,
                               uint32_t *queue_ids, float *queue_dists, uint8_t *queue_checked,
                               uint32_t *queue_size, uint64_t *visited_words, uint64_t *count_distance)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!queue_ids || !queue_dists || !queue_checked || !queue_size || !visited_words)
        return fail(PANNS_ERR_ARG, "null argument");
    if (*queue_size != local_L)
        return fail(PANNS_ERR_UNSUPPORTED, "subsearch_with_top_m needs a full queue (set_L_size == local_L)");
    if (query_id >= h->nq) return fail(PANNS_ERR_RANGE, "query id out of range");
    panns_params p{};
    p.K = 1;
    p.L = local_L;
    p.M = M_max;
    p.mode = PANNS_MODE_SCALE_M;
    p.num_threads = 1;
    rc = check_params(h, &p, local_L, 1);
    if (rc) return rc;
    const bool norm = h->distance == PANNS_DISTANCE_NORM_DOT;
    std::vector<uint64_t> keys(local_L);
    for (uint32_t i = 0; i < local_L; ++i) {
        if (queue_ids[i] >= h->n) return fail(PANNS_ERR_RANGE, "queue id out of range");
        uint32_t db;
        std::memcpy(&db, &queue_dists[i], 4);
        if (norm) {
            // the kernel's key order for distances of either sign (-0 as +0)
            if (db == 0x80000000u) db = 0;
            db ^= (db & 0x80000000u) ? 0xFFFFFFFFu : 0x80000000u;
        }
        keys[i] = ((uint64_t)db << 32) | queue_ids[i] | (queue_checked[i] ? panns::kChk : 0ull);
        if (i && (keys[i] & ~panns::kChk) < (keys[i - 1] & ~panns::kChk))
            return fail(PANNS_ERR_ARG, "queue is not sorted by (distance, id)");
        if (queue_dists[i] != queue_dists[i] || (!norm && queue_dists[i] < 0.f))
            return fail(PANNS_ERR_ARG, norm ? "queue distances must be numbers"
                                            : "queue distances must be non-negative numbers");
    }
    HIP_TRY(hipSetDevice(h->device));
    const size_t words32 = ((size_t)h->n + 31) / 32;
    // sized as the per-workgroup bitmaps are (bitmap_words_for)
    const size_t bm_words = bitmap_words_for(h->n);
    DevBuf<uint32_t> bm;
    HIP_TRY(bm.ensure(bm_words));
    HIP_TRY(hipMemset(bm.p, 0, bm_words * 4));
    HIP_TRY(hipMemcpy(bm.p, visited_words, words32 * 4, hipMemcpyHostToDevice));
    std::unique_ptr<panns_plan> pl(new panns_plan());
    pl->h = h;
    pl->p = p;
    pl->nq = 1;
    pl->want_queue = true;
    pl->shared_bitmap = bm.p;
    HIP_TRY(pl->init_queue.ensure(local_L));
    HIP_TRY(hipMemcpy(pl->init_queue.p, keys.data(), local_L * 8ull, hipMemcpyHostToDevice));
    rc = build_plan(pl.get(), &query_id, nullptr, nullptr, nullptr, 0);
    if (rc) return rc;
    float ms = 0.f;
    rc = run_plan(pl.get(), 1, &ms);
    if (rc) return rc;
    std::vector<uint32_t> ids(1);
    uint32_t nd = 0;
    panns_outputs o{};
    o.ids = ids.data();
    o.queue_ids = queue_ids;
    o.queue_dists = queue_dists;
    o.queue_checked = queue_checked;
    o.ndist = &nd;
    rc = fetch_plan(pl.get(), &o, nullptr, ms);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(visited_words, bm.p, words32 * 4, hipMemcpyDeviceToHost));
    if (count_distance) *count_distance += nd;
    return PANNS_OK;
}

namespace {
int search_trace(panns_index *h, const panns_params *p, const uint32_t *query_ids, uint32_t nq,
                 const uint32_t *init_ids, uint32_t n_init, uint32_t trace_cap, const uint32_t *gt_ids, uint32_t gt_k,
                 float *feat, uint32_t *topk_hash, uint16_t *hits, uint32_t *ndist_trace, uint32_t *niter,
                 uint32_t *out_ids)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!p || !query_ids || !init_ids || !feat || !topk_hash || !niter || nq == 0 || trace_cap == 0)
        return fail(PANNS_ERR_ARG, "bad trace arguments");
    if (gt_ids && (gt_k == 0 || !hits)) return fail(PANNS_ERR_ARG, "ground truth needs gt_k > 0 and a hits array");
    if (p->stop != PANNS_STOP_FIXED) return fail(PANNS_ERR_ARG, "traces are taken from the fixed-L search");
    if (!h->nq) return fail(PANNS_ERR_STATE, "queries not loaded");
    HIP_TRY(hipSetDevice(h->device));
    std::unique_ptr<panns_plan> pl(new panns_plan());
    pl->h = h;
    pl->p = *p;
    pl->nq = nq;
    pl->trace_cap = trace_cap;
    if (gt_ids) {
        pl->trace_gt_k = gt_k;
        HIP_TRY(pl->trace_gt.ensure((size_t)h->nq * gt_k));
        HIP_TRY(hipMemcpy(pl->trace_gt.p, gt_ids, (size_t)h->nq * gt_k * 4, hipMemcpyHostToDevice));
    }
    rc = build_plan(pl.get(), query_ids, nullptr, nullptr, init_ids, n_init);
    if (rc) return rc;
    // entries past a query's last iteration read as zeros
    const size_t cells = (size_t)nq * trace_cap;
    HIP_TRY(hipMemsetAsync(pl->trace_feat.p, 0, cells * panns::kFeat * 4, h->stream));
    HIP_TRY(hipMemsetAsync(pl->trace_topk.p, 0, cells * 4, h->stream));
    if (gt_ids) {
        HIP_TRY(hipMemsetAsync(pl->trace_hits.p, 0, cells * 2, h->stream));
        HIP_TRY(hipMemsetAsync(pl->trace_ndist.p, 0, cells * 4, h->stream));
    }
    float ms = 0.f;
    rc = run_plan(pl.get(), 1, &ms);
    if (rc) return rc;
    std::vector<uint32_t> ids((size_t)nq * pl->maxK);
    panns_outputs o{};
    o.ids = out_ids ? out_ids : ids.data();
    o.niter = niter;
    rc = fetch_plan(pl.get(), &o, nullptr, ms);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(feat, pl->trace_feat.p, (size_t)nq * trace_cap * panns::kFeat * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(topk_hash, pl->trace_topk.p, (size_t)nq * trace_cap * 4, hipMemcpyDeviceToHost));
    if (gt_ids) HIP_TRY(hipMemcpy(hits, pl->trace_hits.p, (size_t)nq * trace_cap * 2, hipMemcpyDeviceToHost));
    if (gt_ids && ndist_trace)
        HIP_TRY(hipMemcpy(ndist_trace, pl->trace_ndist.p, (size_t)nq * trace_cap * 4, hipMemcpyDeviceToHost));
    return PANNS_OK;
}
}  // namespace

int panns_search_trace(panns_index *h, const panns_params *p, const uint32_t *query_ids, uint32_t nq,
                       const uint32_t *init_ids, uint32_t n_init, uint32_t trace_cap, float *feat,
                       uint32_t *topk_hash, uint32_t *niter, uint32_t *out_ids)
{
    return search_trace(h, p, query_ids, nq, init_ids, n_init, trace_cap, nullptr, 0, feat, topk_hash, nullptr,
                        nullptr, niter, out_ids);
}

int panns_search_trace_hits(panns_index *h, const panns_params *p, const uint32_t *query_ids, uint32_t nq,
                            const uint32_t *init_ids, uint32_t n_init, uint32_t trace_cap, const uint32_t *gt_ids,
                            uint32_t gt_k, float *feat, uint32_t *topk_hash, uint16_t *hits,
                            uint32_t *ndist_trace, uint32_t *niter, uint32_t *out_ids)
{
    if (!gt_ids) return fail(PANNS_ERR_ARG, "null ground truth");
    return search_trace(h, p, query_ids, nq, init_ids, n_init, trace_cap, gt_ids, gt_k, feat, topk_hash, hits,
                        ndist_trace, niter, out_ids);
}

int panns_set_stop_model(panns_index *h, const float *params, uint32_t n_floats)
{
    if (!h || !params) return fail(PANNS_ERR_ARG, "null argument");
    if (n_floats != PANNS_STOP_MODEL_FLOATS)
        return fail(PANNS_ERR_ARG, "stop model must hold " + std::to_string(PANNS_STOP_MODEL_FLOATS) + " floats");
    for (uint32_t i = 0; i < n_floats; ++i)
        if (!std::isfinite(params[i]))
            return fail(PANNS_ERR_ARG, "stop model holds a non-finite value");
    for (uint32_t i = 0; i < panns::kFeat; ++i)
        if (!(params[4812 + i] > 0.f)) return fail(PANNS_ERR_ARG, "feature scales must be positive");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(h->model.ensure(n_floats));
    HIP_TRY(hipMemcpy(h->model.p, params, n_floats * 4ull, hipMemcpyHostToDevice));
    h->has_model = true;
    return PANNS_OK;
}

}  // extern "C"

namespace {

// (x - mu) / sd for n rows of 8 features.
std::vector<float> normalise_features(const float *X, uint64_t n, const float *params)
{
    std::vector<float> out((size_t)n * panns::kMlpIn);
    const float *mu = params + 4804, *sd = params + 4812;
    for (uint64_t r = 0; r < n; ++r)
        for (int f = 0; f < panns::kMlpIn; ++f)
            out[r * panns::kMlpIn + f] = (X[r * panns::kMlpIn + f] - mu[f]) / sd[f];
    return out;
}

int check_model_scales(const float *params)
{
    for (int i = 0; i < panns::kMlpIn; ++i)
        if (!(params[4812 + i] > 0.f) || !std::isfinite(params[4804 + i]))
            return fail(PANNS_ERR_ARG, "feature mu must be finite and sd positive");
    return PANNS_OK;
}

}  // namespace

extern "C" {

int panns_mlp_grad(panns_index *h, const float *params, const float *X, const float *Y, uint32_t n, float *grad,
                   float *loss)
{
    if (!h || !params || !X || !Y || !grad || n == 0 || n > 65536) return fail(PANNS_ERR_ARG, "bad gradient arguments");
    int rc = check_model_scales(params);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(h->device));
    const std::vector<float> xn = normalise_features(X, n, params);
    const uint32_t nwg = (n + panns::kMlpTile - 1) / panns::kMlpTile;
    DevBuf<float> dP, dX, dY, part, dl, dg;
    HIP_TRY(dP.ensure(panns::kMlpParamsPad));
    HIP_TRY(dX.ensure(xn.size()));
    HIP_TRY(dY.ensure(n));
    HIP_TRY(part.ensure((size_t)nwg * panns::kMlpParamsPad));
    HIP_TRY(dl.ensure(nwg));
    HIP_TRY(dg.ensure(panns::kMlpParamsPad));
    HIP_TRY(hipMemcpy(dP.p, params, panns::kMlpParams * 4ull, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dX.p, xn.data(), xn.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dY.p, Y, n * 4ull, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(panns::mlp_fwd_bwd_kernel, dim3(nwg), dim3(256), 0, h->stream, dP.p, dX.p, dY.p, nullptr,
                       (uint64_t)0, n, part.p, dl.p);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(panns::mlp_adam_kernel, dim3((panns::kMlpParams + 255) / 256), dim3(256), 0, h->stream, dP.p,
                       part.p, nwg, nullptr, nullptr, 0.f, 0.f, 0.f, 0.f, 0.f, dg.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(grad, dg.p, panns::kMlpParams * 4ull, hipMemcpyDeviceToHost));
    std::vector<float> ls(nwg);
    HIP_TRY(hipMemcpy(ls.data(), dl.p, nwg * 4ull, hipMemcpyDeviceToHost));
    double tot = 0.0;
    for (float v : ls) tot += v;
    if (loss) *loss = (float)(tot / n);
    return PANNS_OK;
}

int panns_mlp_train(panns_index *h, const float *X, const float *Y, uint64_t n, float *params, uint32_t epochs,
                    uint32_t batch, float lr, uint64_t seed, float *epoch_loss)
{
    if (!h || !X || !Y || !params || n == 0 || epochs == 0 || batch == 0 || batch > 65536 || !(lr > 0.f))
        return fail(PANNS_ERR_ARG, "bad training arguments");
    int rc = check_model_scales(params);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(h->device));
    const std::vector<float> xn = normalise_features(X, n, params);
    const uint32_t nwg = (batch + panns::kMlpTile - 1) / panns::kMlpTile;
    const uint64_t nbatch = (n + batch - 1) / batch;
    DevBuf<float> dP, dX, dY, part, dl, m1, m2;
    DevBuf<uint32_t> dperm;
    HIP_TRY(dP.ensure(panns::kMlpParamsPad));
    HIP_TRY(dX.ensure(xn.size()));
    HIP_TRY(dY.ensure(n));
    HIP_TRY(dperm.ensure(n));
    HIP_TRY(part.ensure((size_t)nwg * panns::kMlpParamsPad));
    HIP_TRY(dl.ensure((size_t)nbatch * nwg));
    HIP_TRY(m1.ensure(panns::kMlpParamsPad));
    HIP_TRY(m2.ensure(panns::kMlpParamsPad));
    HIP_TRY(hipMemcpy(dP.p, params, panns::kMlpParams * 4ull, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dX.p, xn.data(), xn.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dY.p, Y, n * 4ull, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(m1.p, 0, panns::kMlpParamsPad * 4ull));
    HIP_TRY(hipMemset(m2.p, 0, panns::kMlpParamsPad * 4ull));
    std::vector<uint32_t> perm(n);
    for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    std::mt19937_64 rng(seed);
    float t = 0.f;
    for (uint32_t e = 0; e < epochs; ++e) {
        for (uint64_t i = n - 1; i > 0; --i) {
            const uint64_t j = rng() % (i + 1);
            std::swap(perm[i], perm[j]);
        }
        HIP_TRY(hipMemcpyAsync(dperm.p, perm.data(), n * 4ull, hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipMemsetAsync(dl.p, 0, (size_t)nbatch * nwg * 4, h->stream));
        for (uint64_t b = 0; b < nbatch; ++b) {
            const uint64_t first = b * batch;
            const uint32_t cnt = (uint32_t)std::min<uint64_t>(batch, n - first);
            const uint32_t g = (cnt + panns::kMlpTile - 1) / panns::kMlpTile;
            hipLaunchKernelGGL(panns::mlp_fwd_bwd_kernel, dim3(g), dim3(256), 0, h->stream, dP.p, dX.p, dY.p,
                               dperm.p, first, cnt, part.p, dl.p + b * nwg);
            HIP_TRY(hipGetLastError());
            t += 1.f;
            hipLaunchKernelGGL(panns::mlp_adam_kernel, dim3((panns::kMlpParams + 255) / 256), dim3(256), 0,
                               h->stream, dP.p, part.p, g, m1.p, m2.p, lr, 0.9f, 0.999f, 1e-8f, t, nullptr);
            HIP_TRY(hipGetLastError());
        }
        if (epoch_loss) {
            std::vector<float> ls((size_t)nbatch * nwg);
            HIP_TRY(hipMemcpyAsync(ls.data(), dl.p, ls.size() * 4, hipMemcpyDeviceToHost, h->stream));
            HIP_TRY(hipStreamSynchronize(h->stream));
            double tot = 0.0;
            for (float v : ls) tot += v;
            epoch_loss[e] = (float)(tot / (double)n);
        }
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(params, dP.p, panns::kMlpParams * 4ull, hipMemcpyDeviceToHost));
    return PANNS_OK;
}

int panns_mfma_probe(panns_index *h, const float *A, const float *B, float *C)
{
    if (!h || !A || !B || !C) return fail(PANNS_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(h->device));
    DevBuf<float> dA, dB, dC;
    HIP_TRY(dA.ensure(256));
    HIP_TRY(dB.ensure(256));
    HIP_TRY(dC.ensure(256));
    HIP_TRY(hipMemcpy(dA.p, A, 1024, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dB.p, B, 1024, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(panns::mfma_probe_kernel, dim3(1), dim3(64), 0, h->stream, dA.p, dB.p, dC.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(C, dC.p, 1024, hipMemcpyDeviceToHost));
    return PANNS_OK;
}

int panns_plan_phase_cycles(panns_plan *plan, uint64_t *out8)
{
    if (plan && !plan->parts.empty())
        return fail(PANNS_ERR_UNSUPPORTED, "phase cycles are kept for single-bucket plans only");
    if (!plan || !out8) return fail(PANNS_ERR_ARG, "null argument");
    if (!plan->h) return fail(PANNS_ERR_STATE, "the plan's index was destroyed");
#ifdef PANNS_PHASE_TIMERS
    HIP_TRY(hipSetDevice(plan->h->device));
    HIP_TRY(hipStreamSynchronize(plan->h->stream));
    std::vector<uint64_t> c((size_t)plan->grid * panns::kPhases);
    HIP_TRY(hipMemcpy(c.data(), plan->phase_cycles.p, c.size() * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < panns::kPhases; ++i) {
        out8[i] = 0;
        for (uint32_t g = 0; g < plan->grid; ++g) out8[i] += c[(size_t)g * panns::kPhases + i];
    }
    return PANNS_OK;
#else
    return fail(PANNS_ERR_UNSUPPORTED, "library built without PANNS_PHASE_TIMERS");
#endif
}

int panns_device_copy(panns_index *h, uint64_t bytes, float *ms)
{
    if (!h || bytes == 0 || bytes % 16) return fail(PANNS_ERR_ARG, "bytes must be a positive multiple of 16");
    HIP_TRY(hipSetDevice(h->device));
    DevBuf<uint4> src, dst;
    const uint64_t n16 = bytes / 16;
    HIP_TRY(src.ensure(n16));
    HIP_TRY(dst.ensure(n16));
    HIP_TRY(hipMemsetAsync(src.p, 1, bytes, h->stream));
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n16 + 255) / 256, 8192);
    hipLaunchKernelGGL(panns::copy_kernel, dim3(grid), dim3(256), 0, h->stream, src.p, dst.p, n16);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(h->ev0, h->stream));
    hipLaunchKernelGGL(panns::copy_kernel, dim3(grid), dim3(256), 0, h->stream, src.p, dst.p, n16);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(h->ev1, h->stream));
    HIP_TRY(hipEventSynchronize(h->ev1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, h->ev0, h->ev1));
    if (ms) *ms = t;
    return PANNS_OK;
}

int panns_exact_knn(panns_index *h, int source, uint32_t k, uint32_t *out_ids, float *out_dists)
{
    return panns_exact_knn_ex(h, source, k, PANNS_KNN_AUTO, out_ids, out_dists, nullptr);
}

int panns_exact_knn_ex(panns_index *h, int source, uint32_t k, int method, uint32_t *out_ids, float *out_dists,
                       uint32_t *n_recomputed)
{
    if (!h || !out_ids) return fail(PANNS_ERR_ARG, "null argument");
    if (!h->n) return fail(PANNS_ERR_STATE, "base set not loaded");
    if (source != 0 && source != 1) return fail(PANNS_ERR_ARG, "source must be 0 (base rows) or 1 (queries)");
    if (method != PANNS_KNN_AUTO && method != PANNS_KNN_EXACT_VALU && method != PANNS_KNN_MFMA)
        return fail(PANNS_ERR_ARG, "unknown k-NN method");
    if (source == 1 && !h->nq) return fail(PANNS_ERR_STATE, "queries not loaded");
    if (k == 0 || k > (uint32_t)panns::kKnnKCap) return fail(PANNS_ERR_ARG, "k must be in [1, 128]");
    if (k > h->n - (source == 0 ? 1u : 0u)) return fail(PANNS_ERR_ARG, "k exceeds the number of candidates");
    HIP_TRY(hipSetDevice(h->device));
    const float *qsrc = source == 0 ? h->base.p : h->queries.p;
    const uint32_t nq = source == 0 ? h->n : h->nq;
    const int ex = source == 0 ? 1 : 0;
    if (n_recomputed) *n_recomputed = 0;
    const void *kp = nullptr;
    KnnLaunch launch = nullptr;
    KnnMfmaLaunch mf{};
    switch (h->dim) {
#define PANNS_KNN_DIM(DD)                                                                                 \
    case DD:                                                                                              \
        kp = reinterpret_cast<const void *>(&panns::knn_kernel<DD>);                                      \
        launch = launch_knn<DD>;                                                                          \
        mf = {launch_knn_bound<DD>, launch_knn_filter<DD>, launch_knn_rerank<DD>};                        \
        break;
        PANNS_KNN_DIM(16)
        PANNS_KNN_DIM(32)
        PANNS_KNN_DIM(40)
        PANNS_KNN_DIM(64)
        PANNS_KNN_DIM(96)
        PANNS_KNN_DIM(128)
#undef PANNS_KNN_DIM
    default:
        return fail(PANNS_ERR_DIM, "dimension has no k-NN kernel (16, 32, 40, 64, 96, 128)");
    }
    const uint32_t D = h->dim, RS = D + 4u;
    const uint32_t lds = 8u * (panns::kKnnQT * panns::kKnnKCap + panns::kKnnQT * 64u + 4u * 64u) +
                         4u * panns::kKnnRT * RS + 4u * 2u * panns::kKnnQT + 8u * panns::kKnnQT;
    HIP_TRY(hipFuncSetAttribute(kp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    // Matrix-instruction path: the exact chain over the first S0 rows bounds
    // each query's k-th distance; then each pass filters the first S(i+1)
    // rows against the bound from S(i) and re-ranks the candidates, which
    // gives the exact k-th distance over S(i+1), up to S = n.  About
    // k * S(i+1) / S(i) rows per query pass a filter, so the sizes grow by
    // an equal ratio of at most 32 and a list holds four times that ratio's
    // share (at least 1,024 ids, at most 4 * 128 * 32).
    const uint32_t S0 = std::max<uint32_t>(4u * k + 64u, std::min<uint32_t>(4096u, h->n / 8u));
    std::vector<uint32_t> levels;
    uint32_t cap = 1024u;
    if (h->n >= 2u * S0) {
        const double span = (double)h->n / (double)S0;
        const int passes = std::max(1, (int)std::ceil(std::log(span) / std::log(32.0) - 1e-9));
        const double ratio = std::pow(span, 1.0 / passes);
        double cur = S0;
        for (int i = 0; i < passes; ++i) {
            cur = (i + 1 == passes) ? (double)h->n : cur * ratio;
            levels.push_back(std::min<uint32_t>(h->n, (uint32_t)std::ceil(cur)));
        }
        cap = std::max<uint32_t>(cap, 4u * k * (uint32_t)std::ceil(ratio));
    }
    const bool use_mfma = (method == PANNS_KNN_MFMA || (method == PANNS_KNN_AUTO && h->n >= kKnnMfmaMinRows)) &&
                          !levels.empty();
    float vmax2 = 0.f;
    if (use_mfma) {
        int rc = ensure_norms(h);
        if (rc) return rc;
        std::vector<float> nv(h->n);
        HIP_TRY(hipMemcpy(nv.data(), h->norms.p, (size_t)h->n * 4, hipMemcpyDeviceToHost));
        for (float x : nv) vmax2 = std::max(vmax2, x);
    }
    // launches of at most 64k queries (16k with candidate lists of up to
    // 64 KB each) keep each kernel short and the lists within 1 GiB
    const uint32_t chunk = use_mfma ? std::min<uint32_t>(65536u, std::max<uint32_t>(64u, (1u << 28) / cap)) : 65536u;
    DevBuf<uint32_t> di, df, dsel, dsi, dcnt, dcand;
    DevBuf<float> dd, dq, dsd, dbound;
    HIP_TRY(di.ensure((size_t)std::min(nq, chunk) * k));
    HIP_TRY(dd.ensure((size_t)std::min(nq, chunk) * k));
    HIP_TRY(df.ensure(std::min(nq, chunk)));
    if (use_mfma) {
        HIP_TRY(dsi.ensure((size_t)std::min(nq, chunk) * k));
        HIP_TRY(dsd.ensure((size_t)std::min(nq, chunk) * k));
        HIP_TRY(dbound.ensure(std::min(nq, chunk)));
        HIP_TRY(dcnt.ensure(std::min(nq, chunk)));
        HIP_TRY(dcand.ensure((size_t)std::min(nq, chunk) * cap));
    }
    std::vector<uint32_t> flags, redo;
    for (uint32_t q0 = 0; q0 < nq; q0 += chunk) {
        const uint32_t cn = std::min(chunk, nq - q0);
        const uint32_t grid = (cn + panns::kKnnQT - 1) / panns::kKnnQT;
        const float *qc = qsrc + (size_t)q0 * D;
        // q0 offsets the chunk's query index so that exclude_self compares
        // base-row ids with global query ids.
        if (use_mfma) {
            // filter rows [0, rows) against the bounds from sd, re-rank into (oi, od)
            auto pass = [&](uint32_t rows, const float *sd, uint32_t *oi, float *od) -> int {
                mf.bound(qc, cn, sd, k, vmax2, dbound.p, h->stream);
                HIP_TRY(hipGetLastError());
                HIP_TRY(hipMemsetAsync(dcnt.p, 0, (size_t)cn * 4, h->stream));
                // split the rows so that the launch has at least 2,048 workgroups
                const uint32_t qblocks = (cn + 63u) / 64u;
                const uint32_t want = std::max<uint32_t>(1u, (2048u + qblocks - 1u) / qblocks);
                const uint32_t per = std::max<uint32_t>(4096u, ((rows + want - 1u) / want + 63u) & ~63u);
                const uint32_t splits = (rows + per - 1u) / per;
                mf.filter(h->base.p, h->norms.p, rows, qc, cn, dbound.p, ex, q0, per, dcand.p, cap, dcnt.p, splits,
                          h->stream);
                HIP_TRY(hipGetLastError());
                mf.rerank(h->base.p, qc, cn, k, dcand.p, cap, dcnt.p, oi, od, df.p, h->stream);
                HIP_TRY(hipGetLastError());
                return PANNS_OK;
            };
            // each level's exact k-th distances feed the next level's bounds
            // (a query whose pass overflowed gets none, hence unbounded later
            // passes and, overflowing again, the exact recompute); the last
            // pass writes the result into (di, dd)
            uint32_t *li[2] = {dsi.p, di.p};
            float *ld[2] = {dsd.p, dd.p};
            const int nl = (int)levels.size();
            int cur = nl % 2 == 0 ? 1 : 0;  // so that the last pass lands in (di, dd)
            launch(h->base.p, S0, qc, cn, k, ex, q0, nullptr, li[cur], ld[cur], grid, lds, h->stream);
            HIP_TRY(hipGetLastError());
            for (int l = 0; l < nl; ++l) {
                int rc = pass(levels[l], ld[cur], li[1 - cur], ld[1 - cur]);
                if (rc) return rc;
                cur = 1 - cur;
            }
        } else {
            launch(h->base.p, h->n, qc, cn, k, ex, q0, nullptr, di.p, dd.p, grid, lds, h->stream);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out_ids + (size_t)q0 * k, di.p, (size_t)cn * k * 4, hipMemcpyDeviceToHost, h->stream));
        if (out_dists)
            HIP_TRY(hipMemcpyAsync(out_dists + (size_t)q0 * k, dd.p, (size_t)cn * k * 4, hipMemcpyDeviceToHost,
                                   h->stream));
        if (use_mfma) {
            flags.resize(cn);
            HIP_TRY(hipMemcpyAsync(flags.data(), df.p, (size_t)cn * 4, hipMemcpyDeviceToHost, h->stream));
        }
        HIP_TRY(hipStreamSynchronize(h->stream));
        for (uint32_t i = 0; use_mfma && i < cn; ++i)
            if (flags[i]) redo.push_back(q0 + i);
    }
    if (!redo.empty()) {
        // the queries whose candidate list overflowed: the exact kernel on
        // every row, own ids excluded as before
        if (n_recomputed) *n_recomputed = (uint32_t)redo.size();
        for (size_t r0 = 0; r0 < redo.size(); r0 += chunk) {
            const uint32_t cn = (uint32_t)std::min<size_t>(chunk, redo.size() - r0);
            HIP_TRY(dq.ensure((size_t)cn * D));
            HIP_TRY(dsel.ensure(cn));
            HIP_TRY(hipMemcpy(dsel.p, redo.data() + r0, (size_t)cn * 4, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(panns::gather_rows_kernel, dim3((uint32_t)(((uint64_t)cn * (D / 4) + 255) / 256)),
                               dim3(256), 0, h->stream, qsrc, dsel.p, cn, D, dq.p);
            HIP_TRY(hipGetLastError());
            const uint32_t grid = (cn + panns::kKnnQT - 1) / panns::kKnnQT;
            launch(h->base.p, h->n, dq.p, cn, k, ex, 0, dsel.p, di.p, dd.p, grid, lds, h->stream);
            HIP_TRY(hipGetLastError());
            std::vector<uint32_t> ids((size_t)cn * k);
            std::vector<float> ds((size_t)cn * k);
            HIP_TRY(hipMemcpyAsync(ids.data(), di.p, ids.size() * 4, hipMemcpyDeviceToHost, h->stream));
            HIP_TRY(hipMemcpyAsync(ds.data(), dd.p, ds.size() * 4, hipMemcpyDeviceToHost, h->stream));
            HIP_TRY(hipStreamSynchronize(h->stream));
            for (uint32_t i = 0; i < cn; ++i) {
                const size_t dst = (size_t)redo[r0 + i] * k;
                std::memcpy(out_ids + dst, ids.data() + (size_t)i * k, k * 4);
                if (out_dists) std::memcpy(out_dists + dst, ds.data() + (size_t)i * k, k * 4);
            }
        }
    }
    return PANNS_OK;
}

int panns_mfma_f32_probe(panns_index *h, const float *A, const float *B, float *C)
{
    if (!h || !A || !B || !C) return fail(PANNS_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(h->device));
    DevBuf<float> dA, dB, dC;
    HIP_TRY(dA.ensure(64));
    HIP_TRY(dB.ensure(64));
    HIP_TRY(dC.ensure(256));
    HIP_TRY(hipMemcpy(dA.p, A, 256, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dB.p, B, 256, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(panns::mfma_f32_probe_kernel, dim3(1), dim3(64), 0, h->stream, dA.p, dB.p, dC.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(C, dC.p, 1024, hipMemcpyDeviceToHost));
    return PANNS_OK;
}

}  // extern "C"
