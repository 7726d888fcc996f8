// Multi-wave top-M search: W waves share one query.
//
// The deterministic GPU form of the reference's multi-thread search
// (PSS5LG.cpp:1543-1673 hands candidates to T-1 workers and merges their
// local queues into the master, merge_all_queues_to_master at :811-842):
// per iteration wave 0 picks the M best unchecked entries, every wave
// expands its share of the picks (adjacency, visited claims in the shared
// LDS table, distances into its own candidate region), and wave 0 merges
// all regions into the master queue.  One iteration is "the L smallest keys
// of the old queue and the unvisited neighbours of the M picks" whichever
// wave claims a shared neighbour, so ids, distances and counters equal the
// single-wave search with the same M (unlike the reference, whose shared
// visited bitset is raced).  More waves per query hide more latency when
// the query's LDS (large L) leaves few workgroups per CU.
//
// LDS layout (lds_bytes_mw on the host must match):
//   Q[Lcap]  cand[W][nbw]  bb[64]   (8 B each)
//   query row (4*D B), visited table (2 B per slot),
//   newids[W][nbw]  picks[Mcap]  ctrl[kCtrlWords]   (4 B each)
#pragma once

#include "search_kernel.h"

namespace panns {

constexpr int kCtrlWords = 64;
// ctrl words
enum : int {
    kCtlWork = 0,     // work item
    kCtlM = 1,        // picks this iteration
    kCtlS = 2,        // queue size
    kCtlN = 8,        // [W] candidates per region
    kCtlDist = 16,    // [W] distance computations per wave
    kCtlAdj = 24,     // [W] adjacency slots per wave
    kCtlSpill = 32,   // [W] spills per wave
};

template <int D, int W, bool kNorm>
__global__ __launch_bounds__(64 * W, 2) void search_kernel_mw(KArgs a)
{
    extern __shared__ uint64_t smem[];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const XorSel xs = make_xor_sel(lane);
    const uint32_t nbw = a.nb_cap;  // per-wave candidate / new-id capacity

    uint64_t *Q = smem;
    uint64_t *cand_all = Q + a.Lcap;
    uint64_t *bb = cand_all + (size_t)W * nbw;
    float4 *qs = reinterpret_cast<float4 *>(bb + 64);
    uint32_t *tbl = reinterpret_cast<uint32_t *>(qs + D / 4);
    uint32_t *newids_all = tbl + a.tbl_words;
    uint32_t *picks = newids_all + (size_t)W * nbw;
    uint32_t *ctrl = picks + a.Mcap;
    uint64_t *cand = cand_all + (size_t)wid * nbw;
    uint32_t *newids = newids_all + (size_t)wid * nbw;

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
    vis.bitmap = a.bitmaps + (size_t)blockIdx.x * a.bitmap_words;
    // each wave logs its spills in its own part of the workgroup's log
    vis.log_cap = a.log_cap / W;
    vis.log = a.spill_log + (size_t)blockIdx.x * a.log_cap + (size_t)wid * vis.log_cap;

    uint32_t wr = blockIdx.x % a.n_ranges, wtried = 0;
    for (;;) {
        if (threadIdx.x == 0) ctrl[kCtlWork] = next_work(a, wr, wtried);
        __syncthreads();
        const uint32_t w = ctrl[kCtlWork];
        if (w >= a.n_work) break;
        const uint32_t qid = a.qids[w];
        const uint32_t L = a.qL ? a.qL[w] : a.L;
        const uint32_t K = a.qK ? a.qK[w] : a.K;

        if (threadIdx.x < D / 4) qs[threadIdx.x] = reinterpret_cast<const float4 *>(a.queries + (size_t)qid * D)[threadIdx.x];
        for (uint32_t i = threadIdx.x; i < a.tbl_words; i += 64 * W) tbl[i] = kSent;
        vis.spills = 0;
        __syncthreads();

        uint32_t S = 0, ndist = 0, nexp = 0, niter = 0, nadj = 0;
        // initialize_set_L_para (PSS5LG.cpp:1055-1096), by wave 0
        if (wid == 0) {
            for (uint32_t c0 = 0; c0 < L; c0 += 64 * kClaimChunks) {
                uint32_t ids[kClaimChunks];
                uint32_t okbits = 0;
                const uint32_t nch = min((uint32_t)kClaimChunks, (L - c0 + 63) / 64);
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
                distance_pass<D, true, kNorm>(a.base, a.norms, a.init_ids + c0, cnt, qs, kKeyMax, cand, 0, lane, lt,
                                              a.group_rows);
                wave_sync();
                uint64_t x = (uint32_t)lane < cnt ? cand[lane] : kKeyMax;
                x = sort64(x, lane, xs);
                merge_batch<true>(Q, S, L, x, cnt, bb, lane);
            }
            ndist = L;
            if (lane == 0) ctrl[kCtlS] = S;
        }
        __syncthreads();
        S = ctrl[kCtlS];

        uint32_t k = 0;
        uint32_t Mcur = (a.mode == kModeScaleM) ? 1u : a.M;
        for (;;) {
            // pick_top_m_unchecked (TOPMLG.cpp:1172-1198), by wave 0
            uint32_t last_k = 0;
            if (wid == 0) {
                uint32_t m = 0, pos = k;
                while (m < Mcur && pos < S) {
                    const uint32_t j = pos + (uint32_t)lane;
                    const uint64_t e = j < S ? Q[j] : kChk;
                    const bool unc = j < S && !(e & kChk);
                    const uint64_t msk = __ballot(unc);
                    const uint32_t rank = (uint32_t)__popcll(msk & lt);
                    const bool take = unc && rank < Mcur - m;
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
                if (lane == 0) ctrl[kCtlM] = m;
            }
            __syncthreads();
            const uint32_t m = ctrl[kCtlM];
            if (m == 0) break;
            ++niter;
            nexp += m;

            // this wave's share of the picks
            const uint32_t per = (m + W - 1) / W;
            const uint32_t g0 = min(m, (uint32_t)wid * per);
            const uint32_t gs = min(m, g0 + per) - g0;
            const uint32_t slots = gs << a.stride_log2;
            uint32_t nnew = 0;
            for (uint32_t c0 = 0; c0 < slots; c0 += 64 * kClaimChunks) {
                uint32_t ids[kClaimChunks];
                uint32_t okbits = 0;
                const uint32_t nch = min((uint32_t)kClaimChunks, (slots - c0 + 63) / 64);
#pragma unroll
                for (int u = 0; u < kClaimChunks; ++u) {
                    const uint32_t s = c0 + (uint32_t)u * 64 + lane;
                    const bool ok = (uint32_t)u < nch && s < slots;
                    const uint32_t sc = ok ? s : 0u;
                    ids[u] = a.adj[((size_t)picks[g0 + (sc >> a.stride_log2)] << a.stride_log2) +
                                   (sc & (stride - 1u))];
                    okbits |= ok ? 1u << u : 0u;
                }
                nnew = claim_chunks(vis, ids, okbits, nch, newids, nnew, lt, nadj);
            }
            wave_sync();
            ndist += nnew;
            // every wave tests against the queue's last key before this
            // iteration's merge (one group: top-L of the union)
            const uint32_t ncand =
                distance_pass<D, true, kNorm>(a.base, a.norms, newids, nnew, qs, km(Q[S - 1]), cand, 0, lane, lt,
                                              a.group_rows);
            if (lane == 0) ctrl[kCtlN + wid] = ncand;
            __syncthreads();

            // wave 0 merges every region into the master queue
            if (wid == 0) {
                uint32_t nk = kSent;
                for (int r = 0; r < W; ++r) {
                    const uint64_t *cr = cand_all + (size_t)r * nbw;
                    const uint32_t nr = ctrl[kCtlN + r];
                    for (uint32_t b0 = 0; b0 < nr; b0 += 64) {
                        const uint32_t cnt = min(64u, nr - b0);
                        uint64_t x = (uint32_t)lane < cnt ? cr[b0 + lane] : kKeyMax;
                        x = sort64_first(x, cnt, lane, xs);
                        const uint32_t p0 = merge_batch<true>(Q, S, L, x, cnt, bb, lane);
                        if (p0 != kSent) nk = min(nk, p0);
                    }
                }
                k = (nk <= last_k) ? nk : last_k + 1u;
                if (lane == 0) {
                    ctrl[kCtlS] = S;
                    ctrl[kCtlM] = k;
                }
            }
            __syncthreads();
            S = ctrl[kCtlS];
            k = ctrl[kCtlM];
            if (a.mode == kModeScaleM && Mcur < a.M) Mcur <<= 1;
            __syncthreads();  // ctrl[kCtlM] is rewritten by the next pick
        }

        // per-wave counters to wave 0 (adjacency slots are counted per lane)
        nadj = wave_sum(nadj, xs);
        if (lane == 0) {
            ctrl[kCtlDist + wid] = ndist;
            ctrl[kCtlAdj + wid] = nadj;
            ctrl[kCtlSpill + wid] = vis.spills;
        }
        __syncthreads();
        if (wid == 0) {
            uint32_t nd = 0, na = 0, nsp = 0;
            for (int r = 0; r < W; ++r) {
                nd += ctrl[kCtlDist + r];
                na += ctrl[kCtlAdj + r];
                nsp += ctrl[kCtlSpill + r];
            }
            for (uint32_t i = lane; i < a.Kstride; i += 64) {
                const bool have = i < K && i < S;
                const uint64_t e = have ? Q[i] : kKeyMax;
                a.out_ids[(size_t)w * a.Kstride + i] = have ? ((uint32_t)e & kIdMask) : kSent;
                if (a.out_dists)
                    a.out_dists[(size_t)w * a.Kstride + i] = have ? kdist<kNorm>(e) : __int_as_float(0x7f800000);
            }
            if (a.out_qids) {
                for (uint32_t i = lane; i < a.Lstride; i += 64) {
                    const bool have = i < L;
                    const uint64_t e = have ? Q[i] : 0ull;
                    a.out_qids[(size_t)w * a.Lstride + i] = have ? ((uint32_t)e & kIdMask) : kSent;
                    a.out_qdists[(size_t)w * a.Lstride + i] = have ? kdist<kNorm>(e) : __int_as_float(0x7f800000);
                    a.out_qchk[(size_t)w * a.Lstride + i] = (have && (e & kChk)) ? 1 : 0;
                }
            }
            if (lane == 0) {
                uint32_t *st = a.out_stats + (size_t)w * 5;
                st[0] = nd;
                st[1] = nexp;
                st[2] = niter;
                st[3] = nsp;
                st[4] = na;
            }
        }
        // each wave clears the bitmap words it set
        if (vis.spills) {
            __threadfence();
            if (vis.spills <= vis.log_cap) {
                clear_logged(vis.bitmap, vis.log, vis.spills, lane);
            } else {
                // log overflow: this wave clears the whole bitmap (another
                // wave clearing its own words meanwhile writes zeros too)
                for (uint64_t i = lane; i < a.bitmap_words; i += 64) vis.bitmap[i] = 0u;
            }
        }
        __threadfence();
        __syncthreads();
    }
}

}  // namespace panns
