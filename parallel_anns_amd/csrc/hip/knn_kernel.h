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



// exclude_self: skip the base row whose id is the query's own, self_ids[q]
// when given, else qoff + q.
template <int D>
__global__ __launch_bounds__(256) void knn_kernel(const float *__restrict__ base, uint32_t n,
                                                  const float *__restrict__ queries, uint32_t nq,
                                                  uint32_t k, int exclude_self, uint32_t qoff,
                                                  const uint32_t *__restrict__ self_ids,
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
    const uint32_t self = qvalid ? (self_ids ? self_ids[q] : qoff + q) : kSent;

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
                if (key < my_thr && !(exclude_self && id == self)) {
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

// dst[i] = rows src[sel[i]] (D floats each); one thread per float4.
__global__ void gather_rows_kernel(const float *__restrict__ src, const uint32_t *__restrict__ sel, uint32_t n,
                                   uint32_t D, float *__restrict__ dst)
{
    const uint32_t per = D / 4;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)n * per) return;
    const uint32_t i = (uint32_t)(t / per), c = (uint32_t)(t % per);
    reinterpret_cast<float4 *>(dst)[(size_t)i * per + c] =
        reinterpret_cast<const float4 *>(src)[(size_t)sel[i] * per + c];
}

typedef float knn_f32x4 __attribute__((ext_vector_type(4)));

// Layout check of the FP32 matrix instruction used below: C (16x16) =
// A (16x4) * B (4x16), row-major.  Lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and receives C[4 (l / 16) + t][l % 16], t = 0..3.
__global__ void mfma_f32_probe_kernel(const float *A, const float *B, float *C)
{
    const int lane = threadIdx.x;
    knn_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A[(lane & 15) * 4 + (lane >> 4)], B[(lane >> 4) * 16 + (lane & 15)],
                                               acc, 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 4; ++t) C[(4 * (lane >> 4) + t) * 16 + (lane & 15)] = acc[t];
}

// Brute-force k-NN with FP32 matrix instructions, exact by construction.
//
// Three launches per batch of queries:
//  1. knn_kernel over a sample of the rows (the first S) gives each query
//     the exact k-th distance tau among them, an upper bound on its k-th
//     distance over all rows.
//  2. knn_filter_kernel streams every row past the queries: the dot products
//     q.v come from v_mfma_f32_16x16x4_f32 and a = |v|^2 - 2 q.v orders a
//     query's rows as |v - q|^2 does.  A row is appended to the query's
//     candidate list when a <= tau - |q|^2 + 2 err, err bounding
//     |a - (|v - q|^2 - |q|^2)| over the fp32 roundings involved (norm chain,
//     matrix dot product, L2 chain, |q|^2; each at most g = 2 (D + 4) 2^-24
//     of the magnitudes it sums).  Every row whose exact distance is at most
//     tau -- the k nearest among them -- is therefore a candidate.
//  3. knn_rerank_kernel computes every candidate's distance with the
//     reference's chain (knn_kernel's arithmetic, bit-exact) and keeps the k
//     smallest (distance, id) keys.
// A query whose list overflowed (or, were the bound wrong, held fewer than k
// rows) gets flag 1 and the host recomputes it with knn_kernel over every row.

// bound[q] = tau - |q|^2 + 2 err (+inf when the sample held fewer than k rows).
template <int D>
__global__ void knn_bound_kernel(const float *__restrict__ queries, uint32_t nq, const float *__restrict__ sample_d,
                                 uint32_t k, float vmax2, float *__restrict__ bound)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const float *qr = queries + (size_t)q * D;
    float qn2 = 0.f;
    for (int i = 0; i < D; ++i) qn2 = __builtin_fmaf(qr[i], qr[i], qn2);
    const float tau = sample_d[(size_t)q * k + k - 1];
    const float vm = sqrtf(vmax2), qn = sqrtf(qn2);
    const float g = 2.0f * (float)(D + 4) * 5.9604645e-8f;
    const float err = g * (vmax2 + 4.0f * vm * qn + (vm + qn) * (vm + qn) + qn2) +
                      9.5367432e-7f * (vm + qn) * (vm + qn);
    // twice the bound, and a relative slack for the rounding of this sum
    const float b = tau - qn2 + 2.0f * err;
    bound[q] = tau == __int_as_float(0x7f800000) ? tau : b + 1e-6f * fabsf(b) + 1e-30f;
}

// Rows [r_lo, r_hi) of blockIdx.y's share against 64 queries per workgroup
// (16 per wave).  A lane of a wave supplies the A fragment of query
// qb + (lane % 16), dims g * D/4 + s (g = lane / 16; the query's D/4 values
// per slot stay in registers) and the B fragment of tile row
// 16 rb + lane % 16 at the same dims, and receives the products of queries
// qb + 4 g + t (t = 0..3) with row 16 rb + lane % 16 (the layout
// mfma_f32_probe_kernel checks).  The next 64-row tile is loaded into
// registers while the current one is multiplied.
template <int D>
__global__ __launch_bounds__(256) void knn_filter_kernel(const float *__restrict__ base,
                                                         const float *__restrict__ norms, uint32_t n,
                                                         const float *__restrict__ queries, uint32_t nq,
                                                         const float *__restrict__ bound, int exclude_self,
                                                         uint32_t qoff, uint32_t rows_per_split,
                                                         uint32_t *__restrict__ cand, uint32_t cap,
                                                         uint32_t *__restrict__ cnt)
{
    constexpr int DS = D / 4;                      // dims per k-slot
    constexpr int RS = D + 4;                      // padded LDS row stride (floats)
    constexpr int F4 = 64 * (D / 4);               // float4 per 64-row tile
    constexpr int NPF = (F4 + 255) / 256;          // per thread
    __shared__ float rows[64 * RS];
    __shared__ float tn[64];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int i16 = lane & 15, g = lane >> 4;
    const uint32_t qb = blockIdx.x * 64u + 16u * (uint32_t)wave;
    const uint32_t lo = blockIdx.y * rows_per_split;
    if (lo >= n) return;  // the whole workgroup (blockIdx only)
    const uint32_t hi = min(n, lo + rows_per_split);

    float av[DS];
    {
        const uint32_t qa = min(qb + (uint32_t)i16, nq - 1u);
        const float *qr = queries + (size_t)qa * D + g * DS;
#pragma unroll
        for (int s = 0; s < DS; ++s) av[s] = qr[s];
    }
    float bnd[4];
    uint32_t selfid[4];
    bool qok[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const uint32_t qo = qb + 4u * (uint32_t)g + (uint32_t)t;
        qok[t] = qo < nq;
        bnd[t] = qok[t] ? bound[qo] : 0.f;
        selfid[t] = exclude_self ? qoff + qo : kSent;
    }

    float4 pre[NPF];
    float pn = 0.f;
    auto load = [&](uint32_t r0) {
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const uint32_t c = (uint32_t)tid + 256u * (uint32_t)u;
            const uint32_t r = c / (uint32_t)(D / 4), e = c % (uint32_t)(D / 4);
            const uint32_t row = min(r0 + r, hi - 1u);  // rows past hi are never emitted
            pre[u] = c < (uint32_t)F4 ? reinterpret_cast<const float4 *>(base + (size_t)row * D)[e]
                                      : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (tid < 64) pn = norms[min(r0 + (uint32_t)tid, hi - 1u)];
    };
    load(lo);
    for (uint32_t r0 = lo; r0 < hi; r0 += 64) {
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const uint32_t c = (uint32_t)tid + 256u * (uint32_t)u;
            if (c < (uint32_t)F4) {
                const uint32_t r = c / (uint32_t)(D / 4), e = c % (uint32_t)(D / 4);
                *reinterpret_cast<float4 *>(&rows[r * RS + 4u * e]) = pre[u];
            }
        }
        if (tid < 64) tn[tid] = pn;
        __syncthreads();
        if (r0 + 64u < hi) load(r0 + 64u);
        knn_f32x4 acc[4];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) acc[rb] = knn_f32x4{0.f, 0.f, 0.f, 0.f};
        if constexpr (DS % 4 == 0) {
#pragma unroll
            for (int s4 = 0; s4 < DS / 4; ++s4) {
                float4 b[4];
#pragma unroll
                for (int rb = 0; rb < 4; ++rb)
                    b[rb] = *reinterpret_cast<const float4 *>(&rows[(16 * rb + i16) * RS + g * DS + 4 * s4]);
#pragma unroll
                for (int rb = 0; rb < 4; ++rb) {
                    acc[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[4 * s4 + 0], b[rb].x, acc[rb], 0, 0, 0);
                    acc[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[4 * s4 + 1], b[rb].y, acc[rb], 0, 0, 0);
                    acc[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[4 * s4 + 2], b[rb].z, acc[rb], 0, 0, 0);
                    acc[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[4 * s4 + 3], b[rb].w, acc[rb], 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int s = 0; s < DS; ++s) {
#pragma unroll
                for (int rb = 0; rb < 4; ++rb)
                    acc[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s], rows[(16 * rb + i16) * RS + g * DS + s],
                                                                   acc[rb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            const uint32_t row = 16u * (uint32_t)rb + (uint32_t)i16;
            const uint32_t rid = r0 + row;
            const float vn = tn[row];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float a = vn - 2.0f * acc[rb][t];
                if (qok[t] && rid < hi && a <= bnd[t] && rid != selfid[t]) {
                    const uint32_t qo = qb + 4u * (uint32_t)g + (uint32_t)t;
                    const uint32_t slot = atomicAdd(&cnt[qo], 1u);
                    if (slot < cap) cand[(size_t)qo * cap + slot] = rid;
                }
            }
        }
        __syncthreads();
    }
}

// One wave per query: the reference's distance for each candidate (one lane
// per candidate, knn_kernel's chain), 64 at a time sorted and merged into
// the query's k smallest keys.
template <int D>
__global__ __launch_bounds__(64) void knn_rerank_kernel(const float *__restrict__ base,
                                                        const float *__restrict__ queries, uint32_t k,
                                                        const uint32_t *__restrict__ cand, uint32_t cap,
                                                        const uint32_t *__restrict__ cnt,
                                                        uint32_t *__restrict__ out_ids,
                                                        float *__restrict__ out_dists, uint32_t *__restrict__ flags)
{
    __shared__ uint64_t list[kKnnKCap];
    __shared__ uint64_t bb[64];
    __shared__ float qs[D];
    const uint32_t q = blockIdx.x;
    const int lane = threadIdx.x;
    const XorSel xs = make_xor_sel(lane);
    for (int i = lane; i < D; i += 64) qs[i] = queries[(size_t)q * D + i];
    wave_sync();
    const uint32_t c = cnt[q];
    const uint32_t m = c > cap ? 0u : c;
    uint32_t S = 0;
    for (uint32_t b0 = 0; b0 < m; b0 += 64) {
        const uint32_t i = b0 + (uint32_t)lane;
        const bool ok = i < m;
        const uint32_t id = ok ? cand[(size_t)q * cap + i] : 0u;
        const float *rv = base + (size_t)id * D;
        float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ii = 0; ii < D / 8; ++ii) {
            const int b = blk_of<D>(ii);
            const float4 lo4 = *reinterpret_cast<const float4 *>(&rv[8 * b]);
            const float4 hi4 = *reinterpret_cast<const float4 *>(&rv[8 * b + 4]);
            const float t0 = lo4.x - qs[8 * b + 0], t1 = lo4.y - qs[8 * b + 1];
            const float t2 = lo4.z - qs[8 * b + 2], t3 = lo4.w - qs[8 * b + 3];
            const float t4 = hi4.x - qs[8 * b + 4], t5 = hi4.y - qs[8 * b + 5];
            const float t6 = hi4.z - qs[8 * b + 6], t7 = hi4.w - qs[8 * b + 7];
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
        uint64_t x = ok ? mk_key(d, id) : kKeyMax;
        x = sort64(x, lane, xs);
        merge_batch<true>(list, S, k, x, min(64u, m - b0), bb, lane);
        wave_sync();
    }
    for (uint32_t i = lane; i < k; i += 64) {
        const bool have = i < S;
        const uint64_t e = have ? list[i] : kKeyMax;
        out_ids[(size_t)q * k + i] = have ? (uint32_t)e : kSent;
        if (out_dists) out_dists[(size_t)q * k + i] = have ? __uint_as_float((uint32_t)(e >> 32)) : __int_as_float(0x7f800000);
    }
    if (lane == 0) flags[q] = (c > cap || S < k) ? 1u : 0u;
}

}  // namespace panns
