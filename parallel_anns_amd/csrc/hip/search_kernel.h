// Device code of the NSG top-M search for CDNA4 (gfx950), one 64-lane
// wavefront per query, persistent workgroups pulling queries from a device
// counter.
//
// Per-query state lives in LDS:
//   Q       sorted candidate queue, L entries of 64-bit keys
//           (float bits << 32 | id) ordered exactly like Candidate::operator<
//           (include/Candidate.h:19-27: distance, then id); bit 31 of the low
//           word is is_checked_ (ids < 2^31), masked off for comparisons.
//   tbl     visited table: buckets of eight 16-bit tags (see Visited); an
//           id whose bucket is full goes to a per-workgroup global bitmap
//           whose touched words are logged and cleared when the query ends.
//           An id is claimed "new" exactly once, so visited = init ids plus
//           every vertex whose distance was computed, as boost::dynamic_bitset
//           gives it in PSS5LG.cpp:1017-1020.
//   qs      the query row
//   newids  neighbour ids newly claimed in the current pick group
//   cand    keys of new neighbours that beat the current L-th key
//   bb      one sorted 64-key batch being merged
//   picks   the ids picked this iteration
//
// One iteration (TOPMLG.cpp:2106-2172): pick the first M unchecked entries
// (pick_top_m_unchecked, TOPMLG.cpp:1172-1198), claim their unvisited
// neighbours, compute distances, and merge the survivors.  The queue after
// the iteration is the L smallest keys of (old queue ∪ new neighbours) with
// old checked flags kept -- the same state the reference's sequential
// add_into_queue calls reach, because the visited set grows by the same union
// and the L-th distance never increases.  Distances use the reference's
// compiled summation order (PSS5LG.cpp:525-560), so every key is bit-equal.
//
// Latency structure: a group of picks has all its adjacency rows loaded in
// one batch, its visited claims resolved in LDS and then with one batch of
// global atomics, and its distances computed 64 rows per pass with both
// 32-row halves in flight together.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace panns {

constexpr uint32_t kSent = 0xFFFFFFFFu;
constexpr uint64_t kChk = 0x80000000ull;
constexpr uint64_t kKeyMax = ~0ull;
constexpr uint32_t kIdMask = 0x7FFFFFFFu;
constexpr int kClaimChunks = 4;        // 64-slot chunks claimed per batch
constexpr uint32_t kGroupSlots = 256;  // neighbour slots per pick group

enum : uint32_t { kModeTopM = 0, kModePssV5 = 1, kModeScaleM = 2 };

// Learned-stop MLP: 8 -> 64 -> 64 -> 1 (ReLU, ReLU, sigmoid), fp32.
constexpr int kFeat = 8;
constexpr int kHid = 64;
struct StopModel {
    const float *w1;  // [64][8]
    const float *b1;  // [64]
    const float *w2;  // [64][64]
    const float *b2;  // [64]
    const float *w3;  // [64]
    const float *b3;  // [1]
    const float *mu;  // [8] feature means
    const float *sd;  // [8] feature scales
};

struct KArgs {
    const float *base;
    const float *norms;   // [n] per-vertex norms (opt-graph distance), or null
    const uint32_t *adj;  // [n][stride], unused slots = kSent
    uint32_t n;
    uint32_t stride_log2;
    const float *queries;
    const uint32_t *qids;  // [n_work]
    const uint32_t *qL;    // [n_work] or null
    const uint32_t *qK;    // [n_work] or null
    uint32_t n_work;
    uint32_t n_ranges;  // work split into this many contiguous ranges, one counter each
    uint32_t L, K, M, mode;
    uint32_t Kstride, Lstride;
    const uint32_t *init_ids;
    uint32_t *work_counter;  // [n_ranges]
    uint32_t *out_ids;
    float *out_dists;
    uint32_t *out_qids;
    float *out_qdists;
    uint8_t *out_qchk;
    uint32_t *out_stats;  // [n_work][5]: ndist, nexp, niter, nspill, nadj
    uint32_t *bitmaps;    // [grid][bitmap_words]
    uint64_t bitmap_words;
    uint32_t *spill_log;  // [grid][log_cap]
    uint32_t log_cap;
    uint32_t Lcap;
    uint32_t tbl_words, tbl_buckets, tbl_magic, id_bits, use_table;  // LDS visited table
    uint32_t tbl_bshift;  // log2 of 32-bit words per bucket: 2 (8 slots) or 1 (4 slots)
    uint32_t ovf_slots;   // LDS overflow entries (full ids) after the buckets: 0 or a power of two
    uint32_t Mcap, nb_cap, group_picks;
    uint32_t group_rows;  // neighbour rows a wave gathers per distance step: 32, 64 or 96
    uint32_t num_threads; // PSS_v5 queues (the reference's num_threads_), search_kernel_pss only
    uint32_t pend_cap;    // pending-key buffer of the deferred merge (multiple of 64, <= 256), 0: merge every step
    // subsearch on a caller-held queue and visited bitmap
    const uint64_t *init_queue;  // [n_work][L] keys, or null
    uint32_t *shared_bitmap;     // used instead of bitmaps when non-null; never cleared
    // learned stop
    uint32_t stop_learned;
    float stop_threshold;
    uint32_t stop_min_iter, stop_every;  // the model runs at iterations >= min_iter, every stop_every-th
    StopModel model;
    // traces for learned-stop training: per iteration 8 features + top-K hash
    float *trace_feat;      // [n_work][trace_cap][8] or null
    uint32_t *trace_topk;   // [n_work][trace_cap] order-free hash of the top-K id set
    uint32_t trace_cap;
    const uint32_t *trace_gt;  // [num_queries][trace_gt_k] true neighbours by query id, or null
    uint32_t trace_gt_k;
    uint16_t *trace_hits;      // [n_work][trace_cap] top-min(K, 100) ids among the first min(gt_k, 100) true ones
    uint32_t *trace_ndist;     // [n_work][trace_cap] distance computations so far, or null
    // per-workgroup cycle counts of the search phases (builds with
    // PANNS_PHASE_TIMERS only): [grid][kPhases]
    uint64_t *phase_cycles;
};

// Phases timed by PANNS_PHASE_TIMERS builds.
enum : int { kPhInit = 0, kPhPick, kPhClaim, kPhDist, kPhMerge, kPhStop, kPhOut, kPhIdle, kPhases };

#ifdef PANNS_PHASE_TIMERS
#define PANNS_PT_DECL                      \
    uint64_t pt_acc_[kPhases] = {};        \
    uint64_t pt_t_ = clock64();
#define PANNS_PT(ph)                       \
    do {                                   \
        const uint64_t n_ = clock64();     \
        pt_acc_[ph] += n_ - pt_t_;         \
        pt_t_ = n_;                        \
    } while (0)
#define PANNS_PT_FLUSH(a)                                                              \
    do {                                                                               \
        if (threadIdx.x == 0 && (a).phase_cycles)                                      \
            for (int i_ = 0; i_ < kPhases; ++i_)                                       \
                (a).phase_cycles[(size_t)blockIdx.x * kPhases + i_] = pt_acc_[i_];     \
    } while (0)
#else
#define PANNS_PT_DECL
#define PANNS_PT(ph) \
    do {             \
    } while (0)
#define PANNS_PT_FLUSH(a) \
    do {                  \
    } while (0)
#endif

__device__ __forceinline__ uint64_t mk_key(float d, uint32_t id)
{
    return ((uint64_t)__float_as_uint(d) << 32) | (uint64_t)id;
}

// Distances as ordered 32-bit key words.  Squared L2 is never negative
// (an fma chain of squares from +0), so its bit pattern already orders like
// its value.  The opt-graph distance norm - 2 v.q is negative whenever
// v.q > norm / 2, so kNorm maps the sign-magnitude bits onto an unsigned
// order: positive values get the top bit set, negative ones are inverted.
// -0 cannot occur (the dot chain starts at +0 and -2 * dot + norm rounds an
// exact zero to +0), so the key order equals the reference's float order.
template <bool kNorm>
__device__ __forceinline__ uint32_t dkey(float d)
{
    const uint32_t u = __float_as_uint(d);
    if constexpr (kNorm) return u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
    return u;
}
template <bool kNorm>
__device__ __forceinline__ float kdist(uint64_t e)
{
    const uint32_t k = (uint32_t)(e >> 32);
    if constexpr (kNorm) return __uint_as_float(k ^ ((k & 0x80000000u) ? 0x80000000u : 0xFFFFFFFFu));
    return __uint_as_float(k);
}
__device__ __forceinline__ uint64_t km(uint64_t e) { return e & ~kChk; }

// Lane exchanges v -> v[lane ^ M] without the LDS crossbar: DPP moves for
// M <= 8 and gfx950's permlane swaps for 16 and 32.  For M = 4 (two row
// rotations) and M = 16, 32 (two halves of a swap) each lane picks one of
// two results; which one is fixed by the hardware's lane order, so it is
// measured once per kernel on the lane ids themselves (make_xor_sel)
// rather than assumed.
struct XorSel {
    bool s4, s16, s32;
};

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, false);
}

constexpr int kDppXor1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int kDppRor4 = 0x124;  // row_ror:4
constexpr int kDppRor8 = 0x128;  // row_ror:8
constexpr int kDppRor12 = 0x12C; // row_ror:12

__device__ __forceinline__ XorSel make_xor_sel(int lane)
{
    const uint32_t l = (uint32_t)lane;
    XorSel s;
    s.s4 = dpp_mov<kDppRor4>(l) == (l ^ 4u);
    const auto p16 = __builtin_amdgcn_permlane16_swap(l, l, false, false);
    s.s16 = p16[0] == (l ^ 16u);
    const auto p32 = __builtin_amdgcn_permlane32_swap(l, l, false, false);
    s.s32 = p32[0] == (l ^ 32u);
    return s;
}

template <int M>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v, const XorSel &s)
{
    if constexpr (M == 1) {
        return dpp_mov<kDppXor1>(v);
    } else if constexpr (M == 2) {
        return dpp_mov<kDppXor2>(v);
    } else if constexpr (M == 4) {
        const uint32_t a = dpp_mov<kDppRor4>(v), b = dpp_mov<kDppRor12>(v);
        return s.s4 ? a : b;
    } else if constexpr (M == 8) {
        return dpp_mov<kDppRor8>(v);
    } else if constexpr (M == 16) {
        const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return s.s16 ? p[0] : p[1];
    } else {
        static_assert(M == 32, "lane exchange distance");
        const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return s.s32 ? p[0] : p[1];
    }
}

template <int M>
__device__ __forceinline__ float xor_lane_f(float v, const XorSel &s)
{
    return __uint_as_float(xor_lane<M>(__float_as_uint(v), s));
}

template <int M>
__device__ __forceinline__ uint64_t xor_lane64(uint64_t x, const XorSel &s)
{
    const uint32_t lo = xor_lane<M>((uint32_t)x, s), hi = xor_lane<M>((uint32_t)(x >> 32), s);
    return ((uint64_t)hi << 32) | lo;
}

// Butterfly sum over the wave (exchange distances 32, 16, ..., 1, in that
// order); every lane gets the total.
template <typename T>
__device__ __forceinline__ T wave_sum(T x, const XorSel &s)
{
    if constexpr (sizeof(T) == 4 && std::is_floating_point<T>::value) {
        x += xor_lane_f<32>(x, s);
        x += xor_lane_f<16>(x, s);
        x += xor_lane_f<8>(x, s);
        x += xor_lane_f<4>(x, s);
        x += xor_lane_f<2>(x, s);
        x += xor_lane_f<1>(x, s);
    } else {
        x += xor_lane<32>(x, s);
        x += xor_lane<16>(x, s);
        x += xor_lane<8>(x, s);
        x += xor_lane<4>(x, s);
        x += xor_lane<2>(x, s);
        x += xor_lane<1>(x, s);
    }
    return x;
}

template <int K, int J>
__device__ __forceinline__ uint64_t bitonic_step(uint64_t x, int lane, const XorSel &s)
{
    const uint64_t y = xor_lane64<J>(x, s);
    const bool up = (lane & K) == 0;
    const bool lower = (lane & J) == 0;
    const uint64_t mn = x < y ? x : y;
    const uint64_t mx = x < y ? y : x;
    return (lower == up) ? mn : mx;
}

template <int K>
__device__ __forceinline__ uint64_t bitonic_merge(uint64_t x, int lane, const XorSel &s)
{
    if constexpr (K >= 64) x = bitonic_step<K, 32>(x, lane, s);
    if constexpr (K >= 32) x = bitonic_step<K, 16>(x, lane, s);
    if constexpr (K >= 16) x = bitonic_step<K, 8>(x, lane, s);
    if constexpr (K >= 8) x = bitonic_step<K, 4>(x, lane, s);
    if constexpr (K >= 4) x = bitonic_step<K, 2>(x, lane, s);
    return bitonic_step<K, 1>(x, lane, s);
}

// The lane id as a value the compiler cannot see through, so the step
// directions derived from it are computed where they are used (two VALU
// instructions a step) instead of being hoisted out of the query loop as 21
// lane masks that outgrow the scalar registers and are spilled.
__device__ __forceinline__ int opaque_lane(int lane)
{
    asm volatile("" : "+v"(lane));
    return lane;
}

// Bitonic sort of one key per lane, ascending across the 64 lanes.
__device__ __forceinline__ uint64_t sort64(uint64_t x, int lane, const XorSel &s)
{
    lane = opaque_lane(lane);
    x = bitonic_merge<2>(x, lane, s);
    x = bitonic_merge<4>(x, lane, s);
    x = bitonic_merge<8>(x, lane, s);
    x = bitonic_merge<16>(x, lane, s);
    x = bitonic_merge<32>(x, lane, s);
    return bitonic_merge<64>(x, lane, s);
}

// sort64 for a batch held in lanes [0, n) with kKeyMax in every lane above:
// the merges of width k sort each k-lane block on its own, and the block
// [0, k) ascending, so once k >= n that block holds the whole batch in order
// and every lane above it still holds kKeyMax -- the wider merges would
// change nothing.  n is wave-uniform.
__device__ __forceinline__ uint64_t sort64_first(uint64_t x, uint32_t n, int lane, const XorSel &s)
{
    lane = opaque_lane(lane);
    x = bitonic_merge<2>(x, lane, s);
    x = bitonic_merge<4>(x, lane, s);
    x = bitonic_merge<8>(x, lane, s);
    if (n <= 8u) return x;
    x = bitonic_merge<16>(x, lane, s);
    if (n <= 16u) return x;
    x = bitonic_merge<32>(x, lane, s);
    if (n <= 32u) return x;
    return bitonic_merge<64>(x, lane, s);
}

// Block order of the reference's 8-wide accumulation: the tail block first
// when (D/8) is odd, then blocks 0, 1, ...
template <int D>
__device__ __forceinline__ constexpr int blk_of(int i)
{
    constexpr int nb = D / 8;
    return (nb & 1) ? (i == 0 ? nb - 1 : i - 1) : i;
}

// Distances of NV base rows to the query, each computed by a lane pair:
// the even lane owns accumulators 0-3, the odd lane 4-7 of every 8-wide
// block; each accumulator is an fma chain in the reference's block order and
// the horizontal sum is ((((((a0+a1)+a2)+a3)+a4)+a5)+a6)+a7.  kNorm = false:
// squared L2, acc = fma(v-q, v-q, acc) (PSS5LG.cpp:525-560); kNorm = true:
// the opt-graph distance, acc = fma(v, q, acc) then -2 * dot + norms[id]
// (PSS5.cpp:336-390).  All rows' loads are issued before any is reduced.
// The exact values are returned in the odd lane.  qs is the query row in
// LDS.  Loads are unconditional: callers pass a valid row (id 0) for
// inactive lanes and discard those results, because exec-masked loads make
// the compiler wait for each one where the branches join.
template <int D, int NV, bool kNorm>
__device__ __forceinline__ void pair_dist(const float *__restrict__ base, const float *__restrict__ norms,
                                          const uint32_t (&id)[NV], int half, const float4 *qs, float (&d)[NV])
{
    constexpr int NB = D / 8;
    float4 v[NV][NB];
    float nrm[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const float4 *r = reinterpret_cast<const float4 *>(base + (size_t)id[j] * D);
#pragma unroll
        for (int i = 0; i < NB; ++i) v[j][i] = r[2 * blk_of<D>(i) + half];
        if constexpr (kNorm) nrm[j] = norms[id[j]];
    }
    float acc[NV][4];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j][0] = acc[j][1] = acc[j][2] = acc[j][3] = 0.f;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const float4 q = qs[2 * blk_of<D>(i) + half];
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            if constexpr (kNorm) {
                acc[j][0] = __builtin_fmaf(v[j][i].x, q.x, acc[j][0]);
                acc[j][1] = __builtin_fmaf(v[j][i].y, q.y, acc[j][1]);
                acc[j][2] = __builtin_fmaf(v[j][i].z, q.z, acc[j][2]);
                acc[j][3] = __builtin_fmaf(v[j][i].w, q.w, acc[j][3]);
            } else {
                const float tx = v[j][i].x - q.x, ty = v[j][i].y - q.y, tz = v[j][i].z - q.z,
                            tw = v[j][i].w - q.w;
                acc[j][0] = __builtin_fmaf(tx, tx, acc[j][0]);
                acc[j][1] = __builtin_fmaf(ty, ty, acc[j][1]);
                acc[j][2] = __builtin_fmaf(tz, tz, acc[j][2]);
                acc[j][3] = __builtin_fmaf(tw, tw, acc[j][3]);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const float sj = ((acc[j][0] + acc[j][1]) + acc[j][2]) + acc[j][3];
        const float o = __uint_as_float(dpp_mov<kDppXor1>(__float_as_uint(sj)));
        d[j] = (((o + acc[j][0]) + acc[j][1]) + acc[j][2]) + acc[j][3];
        if constexpr (kNorm) d[j] = __builtin_fmaf(-2.0f, d[j], nrm[j]);
    }
}

// PSS5.cpp:293-334 compute_norm: one thread per base row, the same chain
// with acc = fma(v, v, acc) over all 8 accumulators.
template <int D>
__global__ void norm_kernel(const float *__restrict__ base, uint32_t n, float *__restrict__ out)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    constexpr int NB = D / 8;
    const float4 *r = reinterpret_cast<const float4 *>(base + (size_t)v * D);
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const float4 lo = r[2 * blk_of<D>(i)], hi = r[2 * blk_of<D>(i) + 1];
        acc[0] = __builtin_fmaf(lo.x, lo.x, acc[0]);
        acc[1] = __builtin_fmaf(lo.y, lo.y, acc[1]);
        acc[2] = __builtin_fmaf(lo.z, lo.z, acc[2]);
        acc[3] = __builtin_fmaf(lo.w, lo.w, acc[3]);
        acc[4] = __builtin_fmaf(hi.x, hi.x, acc[4]);
        acc[5] = __builtin_fmaf(hi.y, hi.y, acc[5]);
        acc[6] = __builtin_fmaf(hi.z, hi.z, acc[6]);
        acc[7] = __builtin_fmaf(hi.w, hi.w, acc[7]);
    }
    float s = acc[0];
#pragma unroll
    for (int j = 1; j < 8; ++j) s = s + acc[j];
    out[v] = s;
}

__device__ __forceinline__ uint32_t lb_queue(const uint64_t *Q, uint32_t S, uint64_t x)
{
    uint32_t lo = 0, cnt = S;
    while (cnt > 0) {
        const uint32_t step = cnt >> 1;
        const uint32_t mid = lo + step;
        if (km(Q[mid]) < x) {
            lo = mid + 1;
            cnt -= step + 1;
        } else {
            cnt = step;
        }
    }
    return lo;
}

// Merge a sorted batch (one key per lane, kKeyMax past nb) into Q[0, S),
// capacity L: Q becomes the L smallest of both.  Old entries move right in
// place, processed from the right end so nothing is overwritten before it is
// read.  Returns the queue position of the smallest inserted key, or kSent;
// adds the number of inserted keys to *ninserted when given.
// Synchronises one wave's LDS accesses when it shares its workgroup with
// waves doing other work (a workgroup barrier would wait for them).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

template <bool kWave>
__device__ __forceinline__ void merge_sync()
{
    if constexpr (kWave) {
        wave_sync();
    } else {
        __syncthreads();
    }
}

// kFlags: the batch's keys carry checked flags (bit 31), which are kept in
// the queue and left out of every comparison.
template <bool kWave = false, bool kFlags = false>
__device__ __forceinline__ uint32_t merge_batch(uint64_t *Q, uint32_t &S, uint32_t L, uint64_t x,
                                                uint32_t nb, uint64_t *bb, int lane,
                                                uint32_t *ninserted = nullptr)
{
    uint32_t pos = kSent;
    bool valid = false;
    if ((uint32_t)lane < nb) {
        pos = lb_queue(Q, S, kFlags ? km(x) : x) + (uint32_t)lane;
        valid = pos < L;
    }
    const uint32_t nv = (uint32_t)__popcll(__ballot(valid));
    const uint32_t p0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)pos);
    if (nv == 0) return kSent;
    if (ninserted) *ninserted += nv;
    if (valid) bb[lane] = x;
    merge_sync<kWave>();
    uint32_t hi = S;
    while (hi > p0) {
        const uint32_t lo = (hi - p0 > 64u) ? hi - 64u : p0;
        const uint32_t j = lo + (uint32_t)lane;
        uint64_t e = 0;
        uint32_t np = kSent;
        if (j < hi) {
            e = Q[j];
            const uint64_t ek = km(e);
            uint32_t c = 0, cnt = nv;
            while (cnt > 0) {
                const uint32_t step = cnt >> 1;
                const uint32_t mid = c + step;
                if ((kFlags ? km(bb[mid]) : bb[mid]) < ek) {
                    c = mid + 1;
                    cnt -= step + 1;
                } else {
                    cnt = step;
                }
            }
            np = j + c;
        }
        merge_sync<kWave>();
        if (np < L) Q[np] = e;
        merge_sync<kWave>();
        hi = lo;
    }
    if (valid) Q[pos] = x;
    S = min(L, S + nv);
    merge_sync<kWave>();
    return p0;
}

// Merges the sorted run R[0, nr) (nr <= 256; keys may carry checked flags)
// into Q[0, S), capacity L: Q becomes the L smallest of both, flags kept.
// Run key i lands at i + (old keys below it), old key j moves right by
// (run keys below it); both are computed on the queue before it changes.
// Returns the queue position of the run's first key, or kSent when no run
// key lands.
__device__ __forceinline__ uint32_t merge_run(uint64_t *Q, uint32_t &S, uint32_t L, const uint64_t *R, uint32_t nr,
                                              int lane)
{
    if (nr == 0) return kSent;
    uint32_t rp[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t i = u * 64u + (uint32_t)lane;
        rp[u] = i < nr ? lb_queue(Q, S, km(R[i])) + i : kSent;
    }
    const uint32_t p0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)rp[0]);
    if (p0 >= L) return kSent;
    __syncthreads();
    uint32_t hi = S;
    while (hi > p0) {
        const uint32_t lo = (hi - p0 > 64u) ? hi - 64u : p0;
        const uint32_t j = lo + (uint32_t)lane;
        uint64_t e = 0;
        uint32_t np = kSent;
        if (j < hi) {
            e = Q[j];
            const uint64_t ek = km(e);
            uint32_t c = 0, cnt = nr;
            while (cnt > 0) {
                const uint32_t step = cnt >> 1;
                const uint32_t mid = c + step;
                if (km(R[mid]) < ek) {
                    c = mid + 1;
                    cnt -= step + 1;
                } else {
                    cnt = step;
                }
            }
            np = j + c;
        }
        __syncthreads();
        if (np < L) Q[np] = e;
        __syncthreads();
        hi = lo;
    }
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t i = u * 64u + (uint32_t)lane;
        if (rp[u] < L) Q[rp[u]] = R[i];
    }
    S = min(L, S + nr);
    __syncthreads();
    return p0;
}

// Visited set of one query: an LDS tag table first, a per-workgroup global
// bitmap for what the table cannot hold.
//
// The table has nb buckets of eight 16-bit slots (one 16-byte LDS read per
// bucket), or four where the id space needs more buckets.  An id is scrambled by a bijection of [0, 2^idbits)
// (multiplication by an odd constant, then an xor-shift); the remainder of
// the result by nb picks the bucket and the quotient (< 2^15) is the tag
// stored in it (0xFFFF marks an empty slot), so bucket and tag together
// identify the id exactly.  An id whose bucket is full goes to one second
// bucket (probe(); with 8-slot buckets both are read at once), then to the
// LDS overflow area of full ids when there is one.  Slots are only ever
// filled during a query, so once both buckets of an id are full (and its
// overflow entries), an id absent from them is decided by the bitmap, and
// an id is never both in the table and in the bitmap.  spills is a
// wave-uniform count kept in registers.
struct Visited {
    uint32_t *tbl;      // LDS: buckets x (1 << bshift) words
    uint32_t nb, magic, idbits, idmask, xshift, use;  // buckets, floor(2^32 / nb)
    uint32_t bshift;    // 2: buckets of 8 slots, 1: of 4 slots, 0: of 2 slots
    uint32_t *ovf;      // LDS: overflow entries (full ids, kSent = empty), or null
    uint32_t ovf_mask;  // entries - 1
    uint32_t *bitmap;   // global
    uint32_t *log;      // global: words touched in the bitmap, may be null
    uint32_t log_cap;
    uint32_t spills;    // ids stored in the bitmap

    __device__ __forceinline__ uint32_t scramble(uint32_t id) const
    {
        uint32_t h = (id * 0x9E3779B1u) & idmask;
        return h ^ (h >> xshift);
    }

    // One bucket of NW words (2 slots each): 0 if `tag` is there, 1 if it
    // was inserted, 2 if the bucket is full without it.  Every failed CAS
    // means another slot of the bucket was filled, so at most 2 * NW retries
    // precede a full bucket or a success.
    // The slot tests are word-parallel and written as selects, so a probe is
    // straight-line code between its LDS read and its CAS: x has a zero
    // 16-bit half iff (x - 0x00010001) & ~x & 0x80008000 != 0 (a borrow into
    // the high half happens only when the low half is zero, which already
    // makes the test true).  Matches are zero halves of word ^ (tag, tag),
    // empty slots (0xFFFF) zero halves of ~word.
    static __device__ __forceinline__ bool has_zero_half(uint32_t x)
    {
        return ((x - 0x00010001u) & ~x & 0x80008000u) != 0u;
    }
    template <int NW>
    __device__ __forceinline__ int probe_words(uint32_t *w, uint32_t tag) const
    {
        const uint32_t t2 = tag * 0x00010001u;
#pragma unroll 1
        for (int attempt = 0; attempt <= 2 * NW; ++attempt) {
            uint32_t words[NW];
            if constexpr (NW == 4) {
                const uint4 v = *reinterpret_cast<const uint4 *>(w);
                words[0] = v.x;
                words[1] = v.y;
                words[2] = v.z;
                words[3] = v.w;
            } else if constexpr (NW == 2) {
                const uint2 v = *reinterpret_cast<const uint2 *>(w);
                words[0] = v.x;
                words[1] = v.y;
            } else {
                words[0] = *w;
            }
            uint32_t hm = 0, ew = NW, eold = 0;  // ew: first word with an empty slot
#pragma unroll
            for (int i = NW - 1; i >= 0; --i) {
                const uint32_t x = words[i] ^ t2;
                hm |= (x - 0x00010001u) & ~x;
                const bool e = has_zero_half(~words[i]);
                ew = e ? (uint32_t)i : ew;
                eold = e ? words[i] : eold;
            }
            if (hm & 0x80008000u) return 0;
            if (ew == NW) return 2;
            const uint32_t eshift = (eold & 0xFFFFu) == 0xFFFFu ? 0u : 16u;
            const uint32_t nw = (eold & ~(0xFFFFu << eshift)) | (tag << eshift);
            if (atomicCAS(&w[ew], eold, nw) == eold) return 1;
        }
        return 2;
    }
    __device__ __forceinline__ int probe_bucket(uint32_t b, uint32_t tag) const
    {
        return bshift == 2   ? probe_words<4>(tbl + (b << 2), tag)
               : bshift == 1 ? probe_words<2>(tbl + (b << 1), tag)
                             : probe_words<1>(tbl + b, tag);
    }

    // 0: visited before; 1: new and now in the table; 2: in neither of its
    // two buckets, both full -- the bitmap decides.  Buckets hold 8 slots,
    // or 4 where the id space needs more buckets than an 8-slot table of the
    // chosen size has (the tag must stay below 2^15).
    // With h = scramble(id) = t * nb + b1, the id's primary bucket is b1 and
    // its tag t (< 2^15: the host keeps nb large enough).  Only when b1 is
    // full does the id go to its second bucket b2 = (b1 + alt(t)) mod nb,
    // alt(t) in [1, nb), as t | 0x8000: slots never empty during a query,
    // so an id can be in b2 only if b1 is full, and (bucket, stored tag)
    // still determines the id (b1 = b2 - alt(t) mod nb, then h, then the id,
    // since the scramble is a bijection).  t = 0x7FFF has no second bucket
    // (its secondary tag would read as an empty slot).
    __device__ __forceinline__ void locate(uint32_t id, uint32_t &b1, uint32_t &tag) const
    {
        const uint32_t h = scramble(id);
        // h / nb by multiply-high with magic = floor(2^32 / nb): the
        // estimate is low by at most one
        tag = __umulhi(h, magic);
        b1 = h - tag * nb;
        if (b1 >= nb) {
            b1 -= nb;
            ++tag;
        }
    }
    __device__ __forceinline__ uint32_t second(uint32_t b1, uint32_t tag) const
    {
        uint32_t b2 = b1 + 1u + __umulhi(tag * 0x85EBCA77u, nb - 1u);
        return b2 >= nb ? b2 - nb : b2;
    }
    // Ids whose buckets are full try the overflow area: full ids (no tag
    // needed) in groups of four entries, one 16-byte read per group, two
    // consecutive groups from a hash of the id.  Exact for the bucket
    // argument: entries only fill and are taken in order within a group, so
    // an id inserted somewhere finds every entry before it taken, and an id
    // sent to the bitmap finds all of its entries taken again.  ovf_mask is
    // the entry count - 1 (entries a multiple of 4).
    static constexpr int kOvfGroups = 2;
    __device__ __forceinline__ int probe_overflow(uint32_t id, uint32_t h) const
    {
        if (!ovf) return 2;
        const uint32_t g0 = h & (ovf_mask >> 2);
#pragma unroll 1
        for (int gi = 0; gi < kOvfGroups; ++gi) {
            uint32_t *w = ovf + (((g0 + (uint32_t)gi) & (ovf_mask >> 2)) << 2);
#pragma unroll 1
            for (int attempt = 0; attempt < 5; ++attempt) {
                const uint4 v = *reinterpret_cast<const uint4 *>(w);
                if (v.x == id || v.y == id || v.z == id || v.w == id) return 0;
                const int e = v.x == kSent ? 0 : v.y == kSent ? 1 : v.z == kSent ? 2 : v.w == kSent ? 3 : -1;
                if (e < 0) break;  // group full: the next one
                const uint32_t old = atomicCAS(w + e, kSent, id);
                if (old == kSent) return 1;
                if (old == id) return 0;
                // another id took entry e: read the group again
            }
        }
        return 2;
    }
    // 8-slot buckets: both buckets of the id read at once and decided in
    // straight-line code.  The tag is looked for in both (as t in b1, as
    // t | 0x8000 in b2); if absent it goes into the first empty slot of b1,
    // or, b1 being full, of b2, by one CAS; a failed CAS means a slot of the
    // two buckets was filled meanwhile, so at most 16 retries precede a
    // success, a hit or two full buckets.  Slots fill in order (every insert
    // takes the lowest empty slot of its snapshot, and a stale snapshot fails
    // its CAS), so a word has an empty slot iff its high half is empty, and
    // the low half is empty too iff the whole word is.  t = 0x7FFF has no
    // second bucket: b1 stands in for it, which finds and inserts the same.
    __device__ __forceinline__ int probe_pair(uint32_t id, uint32_t b1, uint32_t tag) const
    {
        const bool has2 = tag != 0x7FFFu;
        const uint32_t tg2 = has2 ? (tag | 0x8000u) : tag;
        uint32_t *w1 = tbl + (b1 << 2);
        uint32_t *w2 = tbl + ((has2 ? second(b1, tag) : b1) << 2);
        const uint32_t t1 = tag * 0x00010001u, t2 = tg2 * 0x00010001u;
#pragma unroll 1
        for (int attempt = 0; attempt <= 16; ++attempt) {
            const uint4 v1 = *reinterpret_cast<const uint4 *>(w1);
            const uint4 v2 = *reinterpret_cast<const uint4 *>(w2);
            const uint32_t a[8] = {v1.x, v1.y, v1.z, v1.w, v2.x, v2.y, v2.z, v2.w};
            uint32_t hm = 0, ew = 8, eold = 0;
#pragma unroll
            for (int i = 7; i >= 0; --i) {
                const uint32_t x = a[i] ^ (i < 4 ? t1 : t2);
                hm |= (x - 0x00010001u) & ~x;
                const bool e = a[i] >= 0xFFFF0000u;
                ew = e ? (uint32_t)i : ew;
                eold = e ? a[i] : eold;
            }
            if (hm & 0x80008000u) return 0;
            if (ew == 8u) return probe_overflow(id, b1 * 0x9E3779B1u + tag);
            const uint32_t t = ew < 4u ? tag : tg2;
            const uint32_t nw = eold == kSent ? (0xFFFF0000u | t) : ((eold & 0xFFFFu) | (t << 16));
            uint32_t *w = ew < 4u ? w1 + ew : w2 + (ew - 4u);
            if (atomicCAS(w, eold, nw) == eold) return 1;
        }
        return probe_overflow(id, b1 * 0x9E3779B1u + tag);
    }
    __device__ __forceinline__ int probe(uint32_t id) const
    {
        uint32_t b1, tag;
        locate(id, b1, tag);
        if (bshift == 2u) return probe_pair(id, b1, tag);
        int r = probe_bucket(b1, tag);
        if (r != 2) return r;
        if (tag != 0x7FFFu) r = probe_bucket(second(b1, tag), tag | 0x8000u);
        if (r != 2) return r;
        return probe_overflow(id, b1 * 0x9E3779B1u + tag);
    }
};

// Zeroes the bitmap words a query set: the first n entries of its spill log
// (one word index per id stored in the bitmap, duplicates allowed).  Eight
// log entries per lane are read before any word is written, so a query with
// thousands of spills waits for n / 512 log round trips instead of n / 64.
// The log was written by other lanes of the wave with plain stores, so it is
// read at agent scope (past the vector L1).
__device__ __forceinline__ void clear_logged(uint32_t *bitmap, const uint32_t *log, uint32_t n, int lane)
{
    constexpr uint32_t kU = 8;
    for (uint32_t i0 = 0; i0 < n; i0 += 64u * kU) {
        uint32_t wd[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const uint32_t i = i0 + u * 64u + (uint32_t)lane;
            wd[u] = __hip_atomic_load(&log[i < n ? i : n - 1u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u)
            if (i0 + u * 64u + (uint32_t)lane < n) bitmap[wd[u]] = 0u;
    }
}

// Claims up to kClaimChunks ids per lane (kSent = empty slot): table probes
// chunk by chunk, then one batch of bitmap atomics for the ids the table could
// not take.  Appends the newly visited ids to newids[nnew...] in chunk order
// and returns the new count.  An id offered twice is claimed once: within a
// chunk by the table CAS or the atomic, across chunks because table slots
// only fill and the bitmap atomics are exact.
// `okbits` marks the chunks where this lane's value is a slot at all; the
// values themselves are only read here, so the loads that produced them stay
// in flight until the first probe.
__device__ __forceinline__ uint32_t claim_chunks(Visited &vis, uint32_t (&ids)[kClaimChunks],
                                                 uint32_t okbits, uint32_t nch, uint32_t *newids,
                                                 uint32_t nnew, uint64_t lt, uint32_t &nadj)
{
    uint32_t nbits = 0, gbits = 0;
#pragma unroll
    for (int u = 0; u < kClaimChunks; ++u) {
        if ((uint32_t)u < nch) {
            if (!((okbits >> u) & 1u)) ids[u] = kSent;
            int r = 0;
            if (ids[u] != kSent) {
                ++nadj;
                r = vis.use ? vis.probe(ids[u]) : 2;
            }
            if (r == 1) nbits |= 1u << u;
            if (r == 2) gbits |= 1u << u;
        }
    }
    // The bitmap path, only when some lane of the wave has an id there, and
    // only those lanes issue the atomic: the L2 forwards every bitmap atomic
    // to memory (measured: TCC_ATOMIC = TCC_EA0_ATOMIC), so an idle lane's
    // OR of 0 would cost a memory request too.
    if (__ballot(gbits != 0u)) {
        uint32_t old[kClaimChunks];
#pragma unroll
        for (int u = 0; u < kClaimChunks; ++u) {
            old[u] = 0;
            const bool g = (gbits >> u) & 1u;
            if (g) old[u] = atomicOr(&vis.bitmap[ids[u] >> 5], 1u << (ids[u] & 31));
        }
#pragma unroll
        for (int u = 0; u < kClaimChunks; ++u) {
            if ((uint32_t)u < nch) {
                const bool g = ((gbits >> u) & 1u) && !(old[u] & (1u << (ids[u] & 31)));
                const uint64_t gm = __ballot(g);
                if (g) {
                    nbits |= 1u << u;
                    const uint32_t i = vis.spills + (uint32_t)__popcll(gm & lt);
                    if (vis.log && i < vis.log_cap) vis.log[i] = ids[u] >> 5;
                }
                vis.spills += (uint32_t)__popcll(gm);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < kClaimChunks; ++u) {
        if ((uint32_t)u < nch) {
            const bool nw = (nbits >> u) & 1u;
            const uint64_t nm = __ballot(nw);
            if (nw) newids[nnew + (uint32_t)__popcll(nm & lt)] = ids[u];
            nnew += (uint32_t)__popcll(nm);
        }
    }
    return nnew;
}

// The per-lane part of the stop model, kept in registers for the kernel's
// life: lane j's row of W1, its biases and its output weight; the feature
// means and inverse scales are the same for all lanes.  Lane j's row of the
// 64x64 layer (fp32) is read at each evaluation (16 KB shared by every
// workgroup, so it stays in the caches); the model runs a few times per
// query, and holding the row in registers would take 64 of them for the
// kernel's life.
struct LaneModel {
    float w1[kFeat];
    float b1, b2, w3, b3;
    float mu[kFeat], isd[kFeat];
    const float4 *w2row;
};

__device__ __forceinline__ LaneModel load_lane_model(const StopModel &m, int lane)
{
    LaneModel r;
#pragma unroll
    for (int i = 0; i < kFeat; ++i) r.w1[i] = m.w1[lane * kFeat + i];
    r.b1 = m.b1[lane];
    r.b2 = m.b2[lane];
    r.w3 = m.w3[lane];
    r.b3 = m.b3[0];
#pragma unroll
    for (int i = 0; i < kFeat; ++i) {
        r.mu[i] = m.mu[i];
        r.isd[i] = 1.0f / m.sd[i];
    }
    r.w2row = reinterpret_cast<const float4 *>(m.w2 + lane * kHid);
    return r;
}

// One-wave MLP in fp32: lane j computes hidden unit j of both layers;
// features enter as (f - mu) * (1 / sd) (a multiply per feature instead of
// an IEEE division), the sums run in index order.  Returns the sigmoid
// output in every lane.  (Folding the scaling into the first layer would
// cancel catastrophically for a feature of near-zero spread.)
__device__ __forceinline__ float mlp_eval(const LaneModel &lm, const float (&f)[kFeat], float *h1s, int lane,
                                          const XorSel &xs)
{
    // the 64x64 row's loads go out first and land while layer 1 runs
    float4 w2[kHid / 4];
#pragma unroll
    for (int i = 0; i < kHid / 4; ++i) w2[i] = lm.w2row[i];
    float h = lm.b1;
#pragma unroll
    for (int i = 0; i < kFeat; ++i) h = __builtin_fmaf(lm.w1[i], (f[i] - lm.mu[i]) * lm.isd[i], h);
    h = h > 0.f ? h : 0.f;
    h1s[lane] = h;
    __syncthreads();
    float g = lm.b2;
    const float4 *h1v = reinterpret_cast<const float4 *>(h1s);
#pragma unroll
    for (int i = 0; i < kHid / 4; ++i) {
        const float4 x = h1v[i];
        g = __builtin_fmaf(w2[i].x, x.x, g);
        g = __builtin_fmaf(w2[i].y, x.y, g);
        g = __builtin_fmaf(w2[i].z, x.z, g);
        g = __builtin_fmaf(w2[i].w, x.w, g);
    }
    g = g > 0.f ? g : 0.f;
    const float o0 = lm.w3 * g;
    float o = wave_sum(o0, xs);
    __syncthreads();
    o += lm.b3;
    return 1.f / (1.f + __expf(-o));
}

// Distances for ids[cnt0, cnt0 + 32 * NV) of ids[0, cnt) (in LDS), one
// lane pair per id with NV rows per pair in flight; keys below `bound` are
// appended to cand[ncand...].  Returns the new ncand.  With own_in, each
// appended key's owner own_in[i] goes to own_out at the same position.
template <int D, int NV, bool kNorm>
__device__ __forceinline__ uint32_t distance_group(const float *__restrict__ base, const float *__restrict__ norms,
                                                   const uint32_t *ids, uint32_t cnt0, uint32_t cnt,
                                                   const float4 *qs, uint64_t bound, uint64_t *cand,
                                                   uint32_t ncand, int lane, uint64_t lt,
                                                   const uint32_t *own_in = nullptr, uint32_t *own_out = nullptr)
{
    const int half = lane & 1;
    uint32_t id[NV];
    bool ok[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const uint32_t i = cnt0 + (uint32_t)(lane >> 1) + 32u * (uint32_t)j;
        ok[j] = i < cnt;
        const uint32_t r = ids[ok[j] ? i : 0u];
        id[j] = ok[j] ? r : 0u;
    }
    float d[NV];
    pair_dist<D, NV, kNorm>(base, norms, id, half, qs, d);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const uint64_t kj = ((uint64_t)dkey<kNorm>(d[j]) << 32) | (uint64_t)id[j];
        const bool sj = ok[j] && half && kj < bound;
        const uint64_t mj = __ballot(sj);
        if (sj) {
            const uint32_t at = ncand + (uint32_t)__popcll(mj & lt);
            cand[at] = kj;
            if (own_in) own_out[at] = own_in[cnt0 + (uint32_t)(lane >> 1) + 32u * (uint32_t)j];
        }
        ncand += (uint32_t)__popcll(mj);
    }
    return ncand;
}

// Distances for ids[0, cnt) in groups of `rows` neighbour rows per wave
// (one, two or three rows per lane pair in flight): with rows = 96, groups
// of 96 while more than 64 remain, then groups of 64; with 64, groups of 64;
// with 32, groups of 32 -- and a remainder of at most 32 rows always as one
// group of 32, so a short pass does not compute 32 idle rows.  kWide = false
// caps groups at 64 (fewer registers; the learned-stop kernel holds its
// model).
template <int D, bool kWide, bool kNorm>
__device__ __forceinline__ uint32_t distance_pass(const float *__restrict__ base, const float *__restrict__ norms,
                                                  const uint32_t *ids, uint32_t cnt, const float4 *qs,
                                                  uint64_t bound, uint64_t *cand, uint32_t ncand, int lane,
                                                  uint64_t lt, uint32_t rows, const uint32_t *own_in = nullptr,
                                                  uint32_t *own_out = nullptr)
{
    uint32_t p = 0;
    while (p < cnt) {
        if (kWide && rows >= 96u && cnt - p > 64u) {
            ncand = distance_group<D, 3, kNorm>(base, norms, ids, p, cnt, qs, bound, cand, ncand, lane, lt, own_in,
                                                own_out);
            p += 96u;
        } else if (rows >= 64u && cnt - p > 32u) {
            ncand = distance_group<D, 2, kNorm>(base, norms, ids, p, cnt, qs, bound, cand, ncand, lane, lt, own_in,
                                                own_out);
            p += 64u;
        } else {
            ncand = distance_group<D, 1, kNorm>(base, norms, ids, p, cnt, qs, bound, cand, ncand, lane, lt, own_in,
                                                own_out);
            p += 32u;
        }
    }
    return ncand;
}

// Next work item for a workgroup.  The items (queries in locality order)
// are split into n_ranges contiguous ranges, each with its own counter; a
// workgroup starts on range blockIdx.x % n_ranges -- with workgroups dealt
// round-robin to the chip's 8 XCDs, each XCD works through its own range, so
// queries running together on one L2 are neighbours -- and moves on to the
// next range when its own is exhausted.  Returns n_work when every range is.
// r and tried are the caller's state (lane 0 keeps them).
__device__ __forceinline__ uint32_t next_work(const KArgs &a, uint32_t &r, uint32_t &tried)
{
    while (tried < a.n_ranges) {
        const uint32_t lo = (uint32_t)((uint64_t)a.n_work * r / a.n_ranges);
        const uint32_t hi = (uint32_t)((uint64_t)a.n_work * (r + 1) / a.n_ranges);
        const uint32_t i = atomicAdd(a.work_counter + r, 1u);
        if (i < hi - lo) return lo + i;
        r = (r + 1 == a.n_ranges) ? 0u : r + 1u;
        ++tried;
    }
    return a.n_work;
}

// kLearned: the variant that evaluates the stop model (a.stop_learned);
// the other ignores it.  kNorm: the opt-graph distance norm - 2 v.q
// (a.norms) instead of squared L2.  kTrace: the fixed-L search that also
// records the stop features and top-K hashes (a.trace_feat, a.trace_topk);
// the search variants carry none of that code, which keeps their uniform
// state within the scalar registers.  kDefer: the fixed-L search with the
// deferred merge (a.pend_cap > 0; not with kLearned or kTrace); the other
// variants carry none of its code (measured: its branches alone cost the
// plain search 4 %).
template <int D, bool kLearned, bool kNorm, bool kTrace = false, bool kDefer = false>
__global__ __launch_bounds__(64, 2) void search_kernel(KArgs a)
{
    extern __shared__ uint64_t smem[];
    const int lane = threadIdx.x;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const XorSel xs = make_xor_sel(lane);

    uint64_t *Q = smem;
    uint64_t *cand = Q + a.Lcap;
    uint64_t *bb = cand + a.nb_cap;
    // deferred merge (fixed-stop search without traces): survivors wait in
    // P, sorted, and are merged into Q when P fills or a pick needs it
    uint64_t *P = bb + 64;
    float4 *qs = reinterpret_cast<float4 *>(P + a.pend_cap);
    uint32_t *tbl = reinterpret_cast<uint32_t *>(qs + D / 4);
    uint32_t *newids = tbl + a.tbl_words;
    uint32_t *picks = newids + a.nb_cap;
    uint32_t *counts = picks + a.Mcap;     // [2] work item
    // the stop model's hidden layer (64 floats) shares cand: the model runs
    // after the iteration's merge, when cand holds nothing live, and
    // mlp_eval ends with a barrier
    float *h1s = reinterpret_cast<float *>(cand);

    const uint32_t stride = 1u << a.stride_log2;
    Visited vis;
    vis.tbl = tbl;
    vis.nb = a.tbl_buckets;
    vis.magic = a.tbl_magic;
    vis.bshift = a.tbl_bshift;
    vis.ovf = a.ovf_slots ? tbl + (a.tbl_buckets << a.tbl_bshift) : nullptr;
    vis.ovf_mask = a.ovf_slots - 1u;
    vis.idbits = a.id_bits;
    vis.idmask = a.id_bits >= 32 ? ~0u : (1u << a.id_bits) - 1u;
    vis.xshift = a.id_bits / 2 + 1;
    vis.use = a.use_table;
    vis.bitmap = a.shared_bitmap ? a.shared_bitmap : a.bitmaps + (size_t)blockIdx.x * a.bitmap_words;
    vis.log = a.shared_bitmap ? nullptr : a.spill_log + (size_t)blockIdx.x * a.log_cap;
    vis.log_cap = a.log_cap;
    LaneModel lm{};
    if constexpr (kLearned) lm = load_lane_model(a.model, lane);
    PANNS_PT_DECL

    uint32_t wr = blockIdx.x % a.n_ranges, wtried = 0;
    for (;;) {
        if (lane == 0) counts[2] = next_work(a, wr, wtried);
        __syncthreads();
        const uint32_t w = counts[2];
        PANNS_PT(kPhIdle);
        if (w >= a.n_work) break;
        const uint32_t qid = a.qids[w];
        const uint32_t L = a.qL ? a.qL[w] : a.L;
        const uint32_t K = a.qK ? a.qK[w] : a.K;

        if (lane < D / 4) qs[lane] = reinterpret_cast<const float4 *>(a.queries + (size_t)qid * D)[lane];
        for (uint32_t i = lane; i < a.tbl_words; i += 64) tbl[i] = kSent;
        vis.spills = 0;
        __syncthreads();

        uint32_t S = 0, ndist = 0, nexp = 0, niter = 0, nadj = 0;
        uint32_t ins_last = 0;
        float ep_dist = 0.f;
        // the first iteration >= max(1, stop_min_iter) that is a multiple of
        // stop_every (>= 1); later ones follow by addition
        uint32_t next_eval = 0;
        if constexpr (kLearned)
            next_eval = (max(a.stop_min_iter, 1u) + a.stop_every - 1u) / a.stop_every * a.stop_every;
        if (!a.init_queue) {
            // initialize_set_L_para (PSS5LG.cpp:1055-1096): mark every init id
            // visited, compute its distance, sort the L keys.
            // newids holds nb_cap ids, so a claim batch covers at most
            // nb_cap / 64 chunks of init ids
            const uint32_t ich = min((uint32_t)kClaimChunks, a.nb_cap / 64u);
            for (uint32_t c0 = 0; c0 < L; c0 += 64u * ich) {
                uint32_t ids[kClaimChunks];
                uint32_t okbits = 0;
                const uint32_t nch = min(ich, (L - c0 + 63) / 64);
#pragma unroll
                for (int u = 0; u < kClaimChunks; ++u) {
                    const uint32_t i = c0 + (uint32_t)u * 64 + lane;
                    const bool ok = (uint32_t)u < nch && i < L;
                    ids[u] = a.init_ids[ok ? i : 0u];
                    okbits |= ok ? 1u << u : 0u;
                }
                uint32_t dummy = 0;
                claim_chunks(vis, ids, okbits, nch, newids, 0, lt, dummy);
            }
            for (uint32_t c0 = 0; c0 < L; c0 += 64) {
                const uint32_t cnt = min(64u, L - c0);
                const uint32_t n0 = distance_pass<D, !kLearned, kNorm>(a.base, a.norms, a.init_ids + c0, cnt, qs,
                                                                       kKeyMax, cand, 0, lane, lt, a.group_rows);
                (void)n0;
                __syncthreads();
                uint64_t x = (uint32_t)lane < cnt ? cand[lane] : kKeyMax;
                x = sort64(x, lane, xs);
                merge_batch(Q, S, L, x, cnt, bb, lane);
            }
            ndist = L;
        } else {
            const uint64_t *iq = a.init_queue + (size_t)w * a.Lstride;
            for (uint32_t i = lane; i < L; i += 64) Q[i] = iq[i];
            S = L;
            __syncthreads();
        }
        ep_dist = kdist<kNorm>(Q[0]);
        float qnorm = 0.f;
        if constexpr (kLearned || kTrace) {
            if (lane < D / 4) {
                const float4 q = qs[lane];
                qnorm = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
            }
            qnorm = wave_sum(qnorm, xs);
            qnorm = sqrtf(qnorm);
        }
        PANNS_PT(kPhInit);

        uint32_t k = 0;
        uint32_t Mcur = (a.mode == kModeScaleM) ? 1u : (a.mode == kModePssV5 ? 1u : a.M);
        // deferred merge: picks of at most 64, a full queue, no traces (they
        // need the exact queue after every iteration).  The learned stop
        // needs it only where the model runs: the pending keys are merged
        // before such an iteration's pick and its survivors go straight into
        // the queue, so the bound that counts the new keys, and the queue the
        // features are read from, are those of the search without deferral.
        static_assert(!kDefer || !kTrace, "traces need the exact queue after every iteration");
        const bool defer = kDefer && S == L;
        uint32_t SP = 0;
        for (;;) {
            const bool eval_iter = kLearned && niter + 1u == next_eval;
            if (kLearned && defer && eval_iter && SP > 0) {
                const uint32_t p0 = merge_run(Q, S, L, P, SP, lane);
                SP = 0;
                if (p0 != kSent) k = min(k, p0);
                __syncthreads();
            }
            uint32_t m = 0, last_k = 0, pos = k;
            uint32_t nk = kSent, k_next = 0;
            bool picked = false;
            if (defer && SP > 0) {
                // The exact queue is the L smallest of Q and P.  Its first
                // Mcur unchecked entries are the Mcur smallest of Q's first
                // Mcur unchecked ones (from k) and P's, provided those Q
                // entries rank below L in the union (position + SP < L);
                // otherwise P is merged first and the plain pick runs.
                // staging of at most Mcur <= Mcap picks per list (3 * Mcap
                // keys of cand)
                uint64_t *ck = cand, *pk = cand + a.Mcap;
                uint32_t *cp = reinterpret_cast<uint32_t *>(cand + 2u * a.Mcap), *pi = cp + a.Mcap;
                uint32_t mq = 0, q_last = 0;
                while (mq < Mcur && pos < S) {
                    const uint32_t j = pos + (uint32_t)lane;
                    const uint64_t e = j < S ? Q[j] : kChk;
                    const bool unc = j < S && !(e & kChk);
                    const uint64_t msk = __ballot(unc);
                    const uint32_t rank = (uint32_t)__popcll(msk & lt);
                    const bool take = unc && rank < Mcur - mq;
                    if (take) {
                        ck[mq + rank] = e;
                        cp[mq + rank] = j;
                    }
                    const uint64_t tm = __ballot(take);
                    if (tm) {
                        q_last = pos + 63u - (uint32_t)__clzll(tm);
                        mq += (uint32_t)__popcll(tm);
                    }
                    pos += 64;
                }
                if (mq == Mcur && q_last + SP < L) {
                    uint32_t mp = 0;
                    for (uint32_t p = 0; mp < Mcur && p < SP; p += 64) {
                        const uint32_t j = p + (uint32_t)lane;
                        const uint64_t e = j < SP ? P[j] : kChk;
                        const bool unc = j < SP && !(e & kChk);
                        const uint64_t msk = __ballot(unc);
                        const uint32_t rank = (uint32_t)__popcll(msk & lt);
                        const bool take = unc && rank < Mcur - mp;
                        if (take) {
                            pk[mp + rank] = e;
                            pi[mp + rank] = j;
                        }
                        mp += (uint32_t)__popcll(__ballot(take));
                    }
                    __syncthreads();
                    // each candidate's rank among both lists (they are
                    // sorted, and no key is in both)
                    bool tq = false, tp = false;
                    uint32_t rq = kSent, rpk = kSent;
                    if ((uint32_t)lane < mq) {
                        const uint64_t key = km(ck[lane]);
                        uint32_t r = 0;
                        for (uint32_t t = 0; t < mp; ++t) r += km(pk[t]) < key ? 1u : 0u;
                        rq = (uint32_t)lane + r;
                        tq = rq < Mcur;
                    }
                    if ((uint32_t)lane < mp) {
                        const uint64_t key = km(pk[lane]);
                        uint32_t r = 0;
                        for (uint32_t t = 0; t < mq; ++t) r += km(ck[t]) < key ? 1u : 0u;
                        rpk = (uint32_t)lane + r;
                        tp = rpk < Mcur;
                    }
                    __syncthreads();
                    if (tq) {
                        Q[cp[lane]] |= kChk;
                        picks[rq] = (uint32_t)ck[lane] & kIdMask;
                    }
                    if (tp) {
                        P[pi[lane]] |= kChk;
                        picks[rpk] = (uint32_t)pk[lane] & kIdMask;
                    }
                    // Q's picks are a prefix of its unchecked entries: the
                    // scan restarts after the last of them (or where it was)
                    const uint64_t qm = __ballot(tq);
                    k_next = qm ? cp[63 - __clzll(qm)] + 1u : k;
                    m = Mcur;
                    picked = true;
                } else {
                    const uint32_t p0 = merge_run(Q, S, L, P, SP, lane);
                    SP = 0;
                    if (p0 != kSent) k = min(k, p0);
                    pos = k;
                }
                __syncthreads();
            }
            if (!picked) {
            // pick_top_m_unchecked (TOPMLG.cpp:1172-1198)
            while (m < Mcur && pos < S) {
                const uint32_t j = pos + (uint32_t)lane;
                const uint64_t e = j < S ? Q[j] : kChk;
                const bool unc = j < S && !(e & kChk);
                const uint64_t msk = __ballot(unc);
                const uint32_t need = Mcur - m;
                const uint32_t rank = (uint32_t)__popcll(msk & lt);
                const bool take = unc && rank < need;
                if (take) {
                    Q[j] = e | kChk;
                    picks[m + rank] = (uint32_t)e & kIdMask;
                }
                const uint64_t tm = __ballot(take);
                if (tm) {
                    last_k = pos + 63u - (uint32_t)__clzll(tm);
                    m += (uint32_t)__popcll(tm);
                }
                pos += 64;
            }
            }
            __syncthreads();
            PANNS_PT(kPhPick);
            if (m == 0) break;
            ++niter;
            nexp += m;
            uint32_t ins = 0;
            for (uint32_t g0 = 0; g0 < m; g0 += a.group_picks) {
                const uint32_t gs = min(a.group_picks, m - g0);
                const uint32_t slots = gs << a.stride_log2;
                uint32_t nnew = 0;
                for (uint32_t c0 = 0; c0 < slots; c0 += 64 * kClaimChunks) {
                    uint32_t ids[kClaimChunks];
                    uint32_t okbits = 0;
                    const uint32_t nch = min((uint32_t)kClaimChunks, (slots - c0 + 63) / 64);
#pragma unroll
                    for (int u = 0; u < kClaimChunks; ++u) {
                        // unconditional load (slot 0 of the group for unused
                        // lanes) so all chunks are in flight together
                        const uint32_t s = c0 + (uint32_t)u * 64 + lane;
                        const bool ok = (uint32_t)u < nch && s < slots;
                        const uint32_t sc = ok ? s : 0u;
                        ids[u] = a.adj[((size_t)picks[g0 + (sc >> a.stride_log2)] << a.stride_log2) +
                                       (sc & (stride - 1u))];
                        okbits |= ok ? 1u << u : 0u;
                    }
                    uint32_t nch_c = nch;
                    if (nch > 2) {
                        // compact the real neighbour ids (rows are padded to
                        // the stride, mostly with empty slots) through the
                        // idle candidate buffer, so the probes run on fewer,
                        // full chunks
                        uint32_t *stage = reinterpret_cast<uint32_t *>(cand);
                        uint32_t nvalid = 0;
#pragma unroll
                        for (int u = 0; u < kClaimChunks; ++u) {
                            const bool v = ((okbits >> u) & 1u) && ids[u] != kSent;
                            const uint64_t vm = __ballot(v);
                            if (v) stage[nvalid + (uint32_t)__popcll(vm & lt)] = ids[u];
                            nvalid += (uint32_t)__popcll(vm);
                        }
                        __syncthreads();
                        nch_c = (nvalid + 63u) / 64u;
                        okbits = 0;
#pragma unroll
                        for (int u = 0; u < kClaimChunks; ++u) {
                            const uint32_t i = (uint32_t)u * 64 + lane;
                            const bool ok = i < nvalid;
                            ids[u] = stage[ok ? i : 0u];
                            okbits |= ok ? 1u << u : 0u;
                        }
                        __syncthreads();
                    }
                    nnew = claim_chunks(vis, ids, okbits, nch_c, newids, nnew, lt, nadj);
                }
                __syncthreads();
                PANNS_PT(kPhClaim);
                ndist += nnew;
                const uint32_t ncand = distance_pass<D, !kLearned, kNorm>(a.base, a.norms, newids, nnew, qs,
                                                                          km(Q[S - 1]), cand, 0, lane, lt,
                                                                          a.group_rows);
                // stop feature: new keys that beat the queue's last key (a set
                // fixed by the group, unlike the keys a batch-wise merge
                // inserts and a later batch evicts)
                ins += ncand;
                __syncthreads();
                PANNS_PT(kPhDist);
                for (uint32_t b0 = 0; b0 < ncand; b0 += 64) {
                    const uint32_t cnt = min(64u, ncand - b0);
                    uint64_t x = (uint32_t)lane < cnt ? cand[b0 + lane] : kKeyMax;
                    x = sort64_first(x, cnt, lane, xs);
                    if (defer && !eval_iter) {
                        // into P, which is merged into Q first when full
                        if (SP + cnt > a.pend_cap) {
                            const uint32_t p0 = merge_run(Q, S, L, P, SP, lane);
                            SP = 0;
                            if (p0 != kSent) nk = min(nk, p0);
                        }
                        merge_batch(P, SP, a.pend_cap, x, cnt, bb, lane);
                    } else {
                        const uint32_t p0 = merge_batch(Q, S, L, x, cnt, bb, lane);
                        if (p0 != kSent) nk = min(nk, p0);
                    }
                }
                PANNS_PT(kPhMerge);
            }
            // k = nk <= last_k ? nk : last_k + 1 (TOPMLG.cpp:2168-2172)
            if (!picked) k_next = last_k + 1u;
            k = min(nk, k_next);
            ins_last = ins;
            if (a.mode == kModeScaleM && Mcur < a.M) Mcur <<= 1;

            const bool eval_stop = kLearned && niter == next_eval;
            if (eval_stop) next_eval += a.stop_every;
            if (eval_stop || kTrace) {
                // features: step, top-1 dist, K-th dist, ratio, new keys below
                // the queue's last key in the last step, best initial (entry)
                // dist, query norm, unchecked count
                uint32_t unc = 0;
                for (uint32_t j = lane; j < S; j += 64) unc += (Q[j] & kChk) ? 0u : 1u;
                unc = wave_sum(unc, xs);
                float f[kFeat];
                const float d1 = kdist<kNorm>(Q[0]);
                const float dk = kdist<kNorm>(Q[min(K, S) - 1]);
                f[0] = (float)niter;
                f[1] = d1;
                f[2] = dk;
                f[3] = dk / fmaxf(d1, 1e-12f);
                f[4] = (float)ins_last;
                f[5] = ep_dist;
                f[6] = qnorm;
                f[7] = (float)unc;
                if (kTrace && niter <= a.trace_cap) {
                    float *tf = a.trace_feat + ((size_t)w * a.trace_cap + (niter - 1)) * kFeat;
                    if (lane < kFeat) tf[lane] = f[lane];
                    uint32_t hsh = 0;
                    for (uint32_t j = lane; j < K && j < S; j += 64) {
                        uint32_t x = ((uint32_t)Q[j] & kIdMask) * 0x9E3779B1u;
                        x ^= x >> 15;
                        hsh += x * 0x85EBCA77u;
                    }
                    hsh = wave_sum(hsh, xs);
                    if (lane == 0) a.trace_topk[(size_t)w * a.trace_cap + (niter - 1)] = hsh;
                    if (a.trace_gt) {
                        // recall@100's count at this step (get_recall_for_all_queries,
                        // PSS5LG.cpp:309-348): queue entries j < min(K, 100) found
                        // among the first min(gt_k, 100) true neighbours
                        const uint32_t *gq = a.trace_gt + (size_t)qid * a.trace_gt_k;
                        const uint32_t gk = min(a.trace_gt_k, 100u);
                        uint32_t hits = 0;
                        for (uint32_t j = lane; j < K && j < S && j < 100u; j += 64) {
                            const uint32_t id = (uint32_t)Q[j] & kIdMask;
                            for (uint32_t t = 0; t < gk; ++t) hits += gq[t] == id ? 1u : 0u;
                        }
                        hits = wave_sum(hits, xs);
                        if (lane == 0) a.trace_hits[(size_t)w * a.trace_cap + (niter - 1)] = (uint16_t)hits;
                    }
                    if (a.trace_ndist && lane == 0) a.trace_ndist[(size_t)w * a.trace_cap + (niter - 1)] = ndist;
                }
                if constexpr (kLearned) {
                    if (eval_stop) {
                        const float pstop = mlp_eval(lm, f, h1s, lane, xs);
                        if (pstop >= a.stop_threshold) break;
                    }
                }
                PANNS_PT(kPhStop);
            }
        }

        // set_K and the final queue
        for (uint32_t i = lane; i < a.Kstride; i += 64) {
            const bool have = i < K && i < S;
            const uint64_t e = have ? Q[i] : kKeyMax;
            a.out_ids[(size_t)w * a.Kstride + i] = have ? ((uint32_t)e & kIdMask) : kSent;
            if (a.out_dists) a.out_dists[(size_t)w * a.Kstride + i] = have ? kdist<kNorm>(e) : __int_as_float(0x7f800000);
        }
        if (a.out_qids) {
            // rows of queries with L below the batch's largest are padded
            for (uint32_t i = lane; i < a.Lstride; i += 64) {
                const bool have = i < L;
                const uint64_t e = have ? Q[i] : 0ull;
                a.out_qids[(size_t)w * a.Lstride + i] = have ? ((uint32_t)e & kIdMask) : kSent;
                a.out_qdists[(size_t)w * a.Lstride + i] = have ? kdist<kNorm>(e) : __int_as_float(0x7f800000);
                a.out_qchk[(size_t)w * a.Lstride + i] = (have && (e & kChk)) ? 1 : 0;
            }
        }
        nadj = wave_sum(nadj, xs);
        const uint32_t nsp = vis.spills;
        if (lane == 0) {
            uint32_t *st = a.out_stats + (size_t)w * 5;
            st[0] = ndist;
            st[1] = nexp;
            st[2] = niter;
            st[3] = nsp;
            st[4] = nadj;
        }
        // clear the bitmap words this query set, so the next query starts
        // empty; every atomic above has returned, so plain stores suffice
        if (!a.shared_bitmap && nsp) {
            if (nsp <= a.log_cap) {
                clear_logged(vis.bitmap, vis.log, nsp, lane);
            } else {
                for (uint64_t i = lane; i < a.bitmap_words; i += 64) vis.bitmap[i] = 0u;
            }
            __threadfence();
        }
        __syncthreads();
        PANNS_PT(kPhOut);
    }
    PANNS_PT_FLUSH(a);
}

// Streaming copy used to calibrate hardware byte counters against a kernel
// of known traffic (n16 16-byte elements read and written once).
__global__ void copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n16)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Nearest of P pivot rows for each query (fp32, any summation order: used
// only to order a batch's queries by locality, never for results).  Eight
// lanes per query, each holding the query row in registers and scanning
// every eighth pivot (the pivot reads of a lane group are shared); the
// eight partial minima are reduced with lane exchanges, ties to the lower
// pivot.  Blocks of 256 threads serve 32 queries.
template <int D>
__global__ void nearest_pivot_kernel(const float *__restrict__ queries, const uint32_t *__restrict__ qids,
                                     uint32_t nq, const float *__restrict__ piv, uint32_t P,
                                     uint32_t *__restrict__ out)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = t >> 3, slice = t & 7u;
    const bool live = i < nq;
    float q[D];
    const float4 *qr = reinterpret_cast<const float4 *>(queries + (size_t)qids[live ? i : 0u] * D);
#pragma unroll
    for (int k = 0; k < D / 4; ++k) {
        const float4 v = qr[k];
        q[4 * k] = v.x;
        q[4 * k + 1] = v.y;
        q[4 * k + 2] = v.z;
        q[4 * k + 3] = v.w;
    }
    float best = __int_as_float(0x7f800000);
    uint32_t bi = 0xFFFFFFFFu;
    for (uint32_t p = slice; p < P; p += 8) {
        const float *v = piv + (size_t)p * D;
        float d = 0.f;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float x = v[k] - q[k];
            d = __builtin_fmaf(x, x, d);
        }
        if (d < best) {
            best = d;
            bi = p;
        }
    }
#pragma unroll
    for (int m = 1; m < 8; m <<= 1) {
        const float ob = __shfl_xor(best, m);
        const uint32_t oi = (uint32_t)__shfl_xor((int)bi, m);
        if (ob < best || (ob == best && oi < bi)) {
            best = ob;
            bi = oi;
        }
    }
    if (live && slice == 0) out[i] = bi;
}

// Zeroes n16 16-byte words (a kernel, not a memset, like the counter reset
// below).
__global__ void zero_kernel(uint4 *__restrict__ p, uint64_t n16)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Zeroes the n work counters of a launch sequence (a kernel, not a memset,
// so the reset can live inside a captured graph); n <= 64.
__global__ void reset_counter_kernel(uint32_t *c, uint32_t n)
{
    if (threadIdx.x < n) c[threadIdx.x] = 0;
}

}  // namespace panns
