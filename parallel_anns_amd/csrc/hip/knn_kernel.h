// Exact brute-force k-NN for gfx950 with the reference's distance and
// (distance, id) order, i.e. the semantics of
// tools/save_true_nearest_neighbors.cpp:16-35 (every base point offered to
// add_into_queue with capacity K; ties go to the lower id).
//
// A workgroup of 256 lanes owns 32 queries, 8 lanes per query; each lane
// holds its query in registers and computes exact distances for every 8th
// row of a 64-row tile staged in LDS (row stride padded to D+4 floats so the
// eight rows read together fall in different banks).  Keys below a query's
// current K-th key go to a small LDS buffer; after each tile one wave per 8
// queries sorts the buffer (bitonic, 64 lanes) and merges it into the query's
// sorted list, then the threshold tightens.  After the first tiles almost no
// row passes, so the cost is the distance arithmetic.
#pragma once

#include "search_kernel.h"

namespace panns {

constexpr int kKnnQT = 32;    // queries per workgroup
constexpr int kKnnTPQ = 8;    // lanes per query
constexpr int kKnnRT = 64;    // rows per tile
constexpr int kKnnKCap = 128; // max k



template <int D>
__global__ __launch_bounds__(256) void knn_kernel(const float *__restrict__ base, uint32_t n,
                                                  const float *__restrict__ queries, uint32_t nq,
                                                  uint32_t k, int exclude_self, uint32_t qoff,
                                                  uint32_t *__restrict__ out_ids,
                                                  float *__restrict__ out_dists)
{
    constexpr int RS = D + 4;  // padded LDS row stride (floats)
    extern __shared__ uint64_t sm[];
    uint64_t *lists = sm;                                   // [QT][KCap]
    uint64_t *cbuf = lists + kKnnQT * kKnnKCap;             // [QT][64]
    uint64_t *bbs = cbuf + kKnnQT * 64;                     // [4 waves][64]
    float *rows = reinterpret_cast<float *>(bbs + 4 * 64);  // [RT][RS]
    uint32_t *ccnt = reinterpret_cast<uint32_t *>(rows + kKnnRT * RS);  // [QT]
    uint32_t *lsz = ccnt + kKnnQT;                                      // [QT]
    uint64_t *thr = reinterpret_cast<uint64_t *>(lsz + kKnnQT);         // [QT]

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const XorSel xs = make_xor_sel(lane);
    const int ql = tid / kKnnTPQ;
    const int sub = tid % kKnnTPQ;
    const uint32_t q = blockIdx.x * kKnnQT + ql;
    const bool qvalid = q < nq;

    float qv[D];
    {
        const float4 *q4 = reinterpret_cast<const float4 *>(queries + (size_t)(qvalid ? q : 0) * D);
#pragma unroll
        for (int i = 0; i < D / 4; ++i) {
            const float4 t = q4[i];
            qv[4 * i] = t.x;
            qv[4 * i + 1] = t.y;
            qv[4 * i + 2] = t.z;
            qv[4 * i + 3] = t.w;
        }
    }
    if (tid < kKnnQT) {
        ccnt[tid] = 0;
        lsz[tid] = 0;
        thr[tid] = kKeyMax;
    }
    __syncthreads();

    uint64_t my_thr = kKeyMax;
    for (uint32_t r0 = 0; r0 < n; r0 += kKnnRT) {
        const uint32_t rn = min((uint32_t)kKnnRT, n - r0);
        // stage the tile: rn rows of D floats, 16-byte chunks
        {
            const float4 *src = reinterpret_cast<const float4 *>(base + (size_t)r0 * D);
            const uint32_t chunks = rn * (D / 4);
            for (uint32_t c = tid; c < chunks; c += 256) {
                const uint32_t r = c / (D / 4), e4 = c % (D / 4);
                *reinterpret_cast<float4 *>(&rows[r * RS + e4 * 4]) = src[c];
            }
        }
        __syncthreads();
        if (qvalid) {
            for (uint32_t rr = sub; rr < rn; rr += kKnnTPQ) {
                const float *rv = &rows[rr * RS];
                float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int i = 0; i < D / 8; ++i) {
                    const int b = blk_of<D>(i);
                    const float4 lo = *reinterpret_cast<const float4 *>(&rv[8 * b]);
                    const float4 hi = *reinterpret_cast<const float4 *>(&rv[8 * b + 4]);
                    const float t0 = lo.x - qv[8 * b + 0], t1 = lo.y - qv[8 * b + 1];
                    const float t2 = lo.z - qv[8 * b + 2], t3 = lo.w - qv[8 * b + 3];
                    const float t4 = hi.x - qv[8 * b + 4], t5 = hi.y - qv[8 * b + 5];
                    const float t6 = hi.z - qv[8 * b + 6], t7 = hi.w - qv[8 * b + 7];
                    acc[0] = __builtin_fmaf(t0, t0, acc[0]);
                    acc[1] = __builtin_fmaf(t1, t1, acc[1]);
                    acc[2] = __builtin_fmaf(t2, t2, acc[2]);
                    acc[3] = __builtin_fmaf(t3, t3, acc[3]);
                    acc[4] = __builtin_fmaf(t4, t4, acc[4]);
                    acc[5] = __builtin_fmaf(t5, t5, acc[5]);
                    acc[6] = __builtin_fmaf(t6, t6, acc[6]);
                    acc[7] = __builtin_fmaf(t7, t7, acc[7]);
                }
                float d = acc[0];
#pragma unroll
                for (int j = 1; j < 8; ++j) d = d + acc[j];
                const uint32_t id = r0 + rr;
                const uint64_t key = mk_key(d, id);
                if (key < my_thr && !(exclude_self && id == qoff + q)) {
                    const uint32_t slot = atomicAdd(&ccnt[ql], 1u);
                    cbuf[ql * 64 + slot] = key;
                }
            }
        }
        __syncthreads();
        // one wave merges the buffers of its 8 queries
        for (int j = 0; j < kKnnQT / 4; ++j) {
            const int qq = wave * (kKnnQT / 4) + j;
            const uint32_t cnt = ccnt[qq];
            if (cnt == 0) continue;
            uint64_t x = (uint32_t)lane < cnt ? cbuf[qq * 64 + lane] : kKeyMax;
            x = sort64(x, lane, xs);
            uint32_t S = lsz[qq];
            merge_batch<true>(lists + qq * kKnnKCap, S, k, x, cnt, bbs + wave * 64, lane);
            if (lane == 0) {
                lsz[qq] = S;
                thr[qq] = (S == k) ? lists[qq * kKnnKCap + k - 1] : kKeyMax;
                ccnt[qq] = 0;
            }
            wave_sync();
        }
        __syncthreads();
        my_thr = thr[ql];
    }
    // write out: lanes of the query's 8-lane group
    if (qvalid) {
        const uint32_t S = lsz[ql];
        for (uint32_t i = sub; i < k; i += kKnnTPQ) {
            const bool have = i < S;
            const uint64_t e = have ? lists[ql * kKnnKCap + i] : kKeyMax;
            out_ids[(size_t)q * k + i] = have ? (uint32_t)e : kSent;
            if (out_dists) out_dists[(size_t)q * k + i] = have ? __uint_as_float((uint32_t)(e >> 32)) : __int_as_float(0x7f800000);
        }
    }
}

}  // namespace panns
