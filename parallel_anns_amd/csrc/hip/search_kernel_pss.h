// PSS_v5 with T queues (the reference's num_threads_ = T) on one 64-lane
// wavefront per query: PSS5LG.cpp:1376-1724 with pick_top_m_to_workers
// (:917-947), the workers' X-step expansion loops (:1584-1646) and
// merge_all_queues_to_master / merge_two_queues_into_1st_queue_seq_fixed
// (:811-842, :648-719).
//
// The reference runs its T loops as OpenMP threads racing on one visited
// bitset and on the master queue's last distance, so its ids change from run
// to run.  Here the T loops run in lockstep in a fixed order -- step s
// expands the s-th candidate of worker 0, 1, ..., T-2, then of the master
// (queue T-1) -- which is one schedule the reference can take: a neighbour
// reached by two queues in one step belongs to the lower-numbered one, and
// every queue's bound is the master's last distance as the previous step
// left it.  One step's expansions run together like one top-M iteration:
// every adjacency row of the step is loaded at once, visited claims are
// resolved queue by queue (so ownership follows the order above), all new
// neighbours' distances run in one pass, and each queue merges its own
// survivors.  Worker queues hold up to L entries (the reference's driver
// sets local_queue_capacity = L, APP_PSS:136-141).
//
// LDS: Q[T][Lcap] keys (worker w at w*Lcap, the master last), cand[nb_cap]
// keys, bb[64], the query row, the visited table, newids/newown/candown
// [nb_cap], then per-queue words ws[8][kPssMaxQueues] and counts[4].
#pragma once

#include "search_kernel.h"

namespace panns {

constexpr uint32_t kPssMaxQueues = 16;
enum : uint32_t { kWsSize = 0, kWsKuc, kWsPick, kWsAct, kWsLo, kWsHi, kWsR, kWsN };

template <int D, bool kNorm>
__global__ __launch_bounds__(64, 2) void search_kernel_pss(KArgs a)
{
    extern __shared__ uint64_t smem[];
    const int lane = threadIdx.x;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const XorSel xs = make_xor_sel(lane);
    const uint32_t T = a.num_threads;

    uint64_t *Qall = smem;
    uint64_t *cand = Qall + (size_t)T * a.Lcap;
    uint64_t *bb = cand + a.nb_cap;
    float4 *qs = reinterpret_cast<float4 *>(bb + 64);
    uint32_t *tbl = reinterpret_cast<uint32_t *>(qs + D / 4);
    uint32_t *newids = tbl + a.tbl_words;
    uint32_t *newown = newids + a.nb_cap;
    uint32_t *candown = newown + a.nb_cap;
    uint32_t *ws = candown + a.nb_cap;  // [kWsN][kPssMaxQueues]
    uint32_t *counts = ws + kWsN * kPssMaxQueues;
    uint32_t *wsize = ws + kWsSize * kPssMaxQueues, *wkuc = ws + kWsKuc * kPssMaxQueues;
    uint32_t *wpick = ws + kWsPick * kPssMaxQueues, *wact = ws + kWsAct * kPssMaxQueues;
    uint32_t *wlo = ws + kWsLo * kPssMaxQueues, *whi = ws + kWsHi * kPssMaxQueues;

    const uint32_t stride = 1u << a.stride_log2;
    // adjacency chunks (64 slots) per row: 1 for strides up to 64
    const uint32_t cpr = stride > 64u ? stride >> 6 : 1u;
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
    vis.use = 1;
    vis.bitmap = a.bitmaps + (size_t)blockIdx.x * a.bitmap_words;
    vis.log = a.spill_log + (size_t)blockIdx.x * a.log_cap;
    vis.log_cap = a.log_cap;

    uint32_t wr = blockIdx.x % a.n_ranges, wtried = 0;
    for (;;) {
        if (lane == 0) counts[2] = next_work(a, wr, wtried);
        __syncthreads();
        const uint32_t w = counts[2];
        if (w >= a.n_work) break;
        const uint32_t qid = a.qids[w];
        const uint32_t L = a.qL ? a.qL[w] : a.L;
        const uint32_t K = a.qK ? a.qK[w] : a.K;
        uint64_t *Qm = Qall + (size_t)(T - 1) * a.Lcap;

        if (lane < D / 4) qs[lane] = reinterpret_cast<const float4 *>(a.queries + (size_t)qid * D)[lane];
        for (uint32_t i = lane; i < a.tbl_words; i += 64) tbl[i] = kSent;
        if ((uint32_t)lane < kPssMaxQueues) wsize[lane] = 0;
        vis.spills = 0;
        __syncthreads();

        uint32_t ndist = 0, nexp = 0, nmerge = 0, nadj = 0;
        // initialize_set_L_para (PSS5LG.cpp:1055-1096) into the master queue
        uint32_t Sm = 0;
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
            __syncthreads();
            uint64_t x = (uint32_t)lane < cnt ? cand[lane] : kKeyMax;
            x = sort64(x, lane, xs);
            merge_batch(Qm, Sm, L, x, cnt, bb, lane);
        }
        ndist = L;
        if (lane == 0) wsize[T - 1] = L;
        __syncthreads();

        // One lockstep step over the queues listed in wact[0, nact) (ascending),
        // each expanding the candidate wpick[q] found at position wkuc[q] of
        // its own queue (expand_one_candidate, PSS5LG.cpp:984-1053); wkuc[q]
        // then follows the reference's r <= k_uc ? r : k_uc + 1.
        auto step = [&](uint32_t nact) {
            const uint64_t lastk = Qm[L - 1];
            // accepts d <= the master's last distance (dist > dist_bound
            // rejects, PSS5LG.cpp:1032); ids stay below 2^31
            const uint64_t bound = (lastk & 0xFFFFFFFF00000000ull) | 0xFFFFFFFFull;
            uint32_t nnew = 0;
            const uint32_t per = max(1u, (uint32_t)kClaimChunks / cpr);  // rows loaded together
            for (uint32_t i0 = 0; i0 < nact; i0 += per) {
                const uint32_t gs = min(per, nact - i0);
                uint32_t ids[kClaimChunks];
#pragma unroll
                for (int u = 0; u < kClaimChunks; ++u) {
                    const uint32_t r = (uint32_t)u / cpr, ch = (uint32_t)u % cpr;
                    const uint32_t slot = ch * 64u + (uint32_t)lane;
                    const bool ok = r < gs && slot < stride;
                    const uint32_t q = wact[i0 + (r < gs ? r : 0u)];
                    ids[u] = a.adj[((size_t)wpick[q] << a.stride_log2) + (ok ? slot : 0u)];
                    if (!ok) ids[u] = kSent;
                }
                for (uint32_t r = 0; r < gs; ++r) {
                    // this row's chunks, claimed before the next queue's
                    uint32_t one[kClaimChunks];
                    uint32_t okbits = 0;
#pragma unroll
                    for (int u = 0; u < kClaimChunks; ++u) {
                        uint32_t v = kSent;
#pragma unroll
                        for (int k = 0; k < kClaimChunks; ++k) v = ((uint32_t)k == r * cpr + (uint32_t)u) ? ids[k] : v;
                        one[u] = (uint32_t)u < cpr ? v : kSent;
                        okbits |= (uint32_t)u < cpr ? 1u << u : 0u;
                    }
                    const uint32_t before = nnew;
                    nnew = claim_chunks(vis, one, okbits, cpr, newids, nnew, lt, nadj);
                    const uint32_t q = wact[i0 + r];
                    for (uint32_t i = before + (uint32_t)lane; i < nnew; i += 64) newown[i] = q;
                }
            }
            __syncthreads();
            ndist += nnew;
            nexp += nact;
            const uint32_t ncand = distance_pass<D, true, kNorm>(a.base, a.norms, newids, nnew, qs, bound, cand, 0,
                                                                 lane, lt, a.group_rows, newown, candown);
            __syncthreads();
            // each queue's survivors: a run of cand (claims went queue by queue)
            if ((uint32_t)lane < nact) {
                const uint32_t q = wact[lane];
                uint32_t lo = 0, hi = ncand;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (candown[mid] < q) lo = mid + 1; else hi = mid;
                }
                wlo[lane] = lo;
                uint32_t lo2 = lo;
                hi = ncand;
                while (lo2 < hi) {
                    const uint32_t mid = (lo2 + hi) >> 1;
                    if (candown[mid] <= q) lo2 = mid + 1; else hi = mid;
                }
                whi[lane] = lo2;
            }
            __syncthreads();
            for (uint32_t i = 0; i < nact; ++i) {
                const uint32_t q = wact[i];
                uint64_t *Qq = Qall + (size_t)q * a.Lcap;
                uint32_t Sq = wsize[q];
                uint32_t nk = kSent;
                const uint32_t b = wlo[i], e = whi[i];
                for (uint32_t b0 = b; b0 < e; b0 += 64) {
                    const uint32_t cnt = min(64u, e - b0);
                    uint64_t x = (uint32_t)lane < cnt ? cand[b0 + lane] : kKeyMax;
                    x = sort64_first(x, cnt, lane, xs);
                    const uint32_t p0 = merge_batch(Qq, Sq, L, x, cnt, bb, lane);
                    if (p0 != kSent) nk = min(nk, p0);
                }
                const uint32_t r = nk == kSent ? L : nk;
                const uint32_t k = wkuc[q];
                __syncthreads();
                if (lane == 0) {
                    wsize[q] = Sq;
                    wkuc[q] = (r <= k) ? r : k + 1u;
                }
                __syncthreads();
            }
        };

        // sequential start: the master expands its first entry (PSS5LG.cpp:1478-1534)
        if (lane == 0) {
            Qm[0] |= kChk;
            wpick[T - 1] = (uint32_t)Qm[0] & kIdMask;
            wkuc[T - 1] = 0;
            wact[0] = T - 1;
        }
        __syncthreads();
        step(1);
        uint32_t k_master = wkuc[T - 1];

        for (;;) {
            // pick_top_m_to_workers (PSS5LG.cpp:917-947): the j-th unchecked
            // master entry from k_master goes to worker j mod T, every T-th
            // stays with the master; worker queues are empty here, and with
            // capacity L none fills (j < L), so the deal never stops early
            uint32_t seen = 0;
            for (uint32_t p = k_master; p < L; p += 64) {
                const uint32_t j = p + (uint32_t)lane;
                const uint64_t e = j < L ? Qm[j] : kChk;
                const bool unc = j < L && !(e & kChk);
                const uint64_t msk = __ballot(unc);
                if (unc) {
                    const uint32_t rank = seen + (uint32_t)__popcll(msk & lt);
                    const uint32_t q = rank % T;
                    if (q != T - 1) {
                        Qall[(size_t)q * a.Lcap + rank / T] = e;
                        Qm[j] = e | kChk;
                    }
                }
                seen += (uint32_t)__popcll(msk);
            }
            if (seen == 0) break;
            if ((uint32_t)lane + 1u < T) wsize[lane] = (uint32_t)lane < seen ? (seen - (uint32_t)lane + T - 1u) / T : 0u;
            if ((uint32_t)lane < T) wkuc[lane] = (uint32_t)lane == T - 1 ? k_master : 0u;
            __syncthreads();

            // the expansion loops, X steps in lockstep (PSS5LG.cpp:1584-1646)
            for (uint32_t s = 0; s < a.M; ++s) {
                // each queue's first unchecked entry at or after its k_uc
                uint32_t nact = 0;
                for (uint32_t q = 0; q < T; ++q) {
                    uint64_t *Qq = Qall + (size_t)q * a.Lcap;
                    const uint32_t Sq = wsize[q];
                    uint32_t pos = wkuc[q], found = kSent;
                    while (pos < Sq && found == kSent) {
                        const uint32_t j = pos + (uint32_t)lane;
                        const bool unc = j < Sq && !(Qq[j] & kChk);
                        const uint64_t msk = __ballot(unc);
                        if (msk) found = pos + (uint32_t)__ffsll((unsigned long long)msk) - 1u;
                        pos += 64;
                    }
                    __syncthreads();
                    if (found != kSent) {
                        if (lane == 0) {
                            Qq[found] |= kChk;
                            wpick[q] = (uint32_t)Qq[found] & kIdMask;
                            wkuc[q] = found;
                            wact[nact] = q;
                        }
                        ++nact;
                    } else if (lane == 0) {
                        wkuc[q] = Sq;
                    }
                    __syncthreads();
                }
                if (nact == 0) break;
                step(nact);
            }
            k_master = wkuc[T - 1];

            // merge_all_queues_to_master (PSS5LG.cpp:811-842), workers in order
            ++nmerge;
            uint32_t nk = L;
            for (uint32_t q = 0; q + 1 < T; ++q) {
                const uint32_t Sq = wsize[q];
                if (Sq == 0) continue;
                const uint64_t *Qq = Qall + (size_t)q * a.Lcap;
                // merge_two_queues_into_1st_queue_seq_fixed (PSS5LG.cpp:648-719):
                // the lower bound of the worker's first entry is the return
                const uint32_t r = lb_queue(Qm, L, km(Qq[0]));
                nk = min(nk, r);
                if (r == L - 1) {
                    // lands on the last index: replaces it, flag included
                    __syncthreads();
                    if (lane == 0) Qm[L - 1] = Qq[0];
                    __syncthreads();
                } else if (r < L - 1) {
                    // a duplicate keeps the master's copy, unchecked when the
                    // worker's copy is; the others merge as sorted batches
                    uint32_t nst = 0;
                    for (uint32_t c0 = 0; c0 < Sq; c0 += 64) {
                        const uint32_t j = c0 + (uint32_t)lane;
                        bool keep = false;
                        uint64_t e = 0;
                        if (j < Sq) {
                            e = Qq[j];
                            const uint32_t pos = lb_queue(Qm, L, km(e));
                            const bool dup = pos < L && km(Qm[pos]) == km(e);
                            if (dup && !(e & kChk)) Qm[pos] &= ~kChk;
                            keep = !dup;
                        }
                        const uint64_t km_ = __ballot(keep);
                        if (keep) cand[nst + (uint32_t)__popcll(km_ & lt)] = e;
                        nst += (uint32_t)__popcll(km_);
                    }
                    __syncthreads();
                    uint32_t S2 = L;
                    for (uint32_t b0 = 0; b0 < nst; b0 += 64) {
                        const uint32_t cnt = min(64u, nst - b0);
                        const uint64_t x = (uint32_t)lane < cnt ? cand[b0 + lane] : kKeyMax;
                        // sorted already (a worker queue is); the merge keys
                        // carry the worker's checked flags
                        merge_batch<false, true>(Qm, S2, L, x, cnt, bb, lane);
                    }
                }
                __syncthreads();
                if (lane == 0) wsize[q] = 0;
                __syncthreads();
            }
            if (nk <= k_master) k_master = nk;
        }

        // set_K and the final (master) queue
        for (uint32_t i = lane; i < a.Kstride; i += 64) {
            const bool have = i < K;
            const uint64_t e = have ? Qm[i] : kKeyMax;
            a.out_ids[(size_t)w * a.Kstride + i] = have ? ((uint32_t)e & kIdMask) : kSent;
            if (a.out_dists) a.out_dists[(size_t)w * a.Kstride + i] = have ? kdist<kNorm>(e) : __int_as_float(0x7f800000);
        }
        if (a.out_qids) {
            for (uint32_t i = lane; i < a.Lstride; i += 64) {
                const bool have = i < L;
                const uint64_t e = have ? Qm[i] : 0ull;
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
            st[2] = nmerge;
            st[3] = nsp;
            st[4] = nadj;
        }
        if (nsp) {
            if (nsp <= a.log_cap) {
                clear_logged(vis.bitmap, vis.log, nsp, lane);
            } else {
                for (uint64_t i = lane; i < a.bitmap_words; i += 64) vis.bitmap[i] = 0u;
            }
            __threadfence();
        }
        __syncthreads();
    }
}

}  // namespace panns
