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

// Brute-force k-NN with FP32 MFMA candidates and an exact re-rank.
//
// A workgroup owns 32 queries.  Per 64-row tile, the four waves form the
// 32 x 64 block of dot products q.v with v_mfma_f32_16x16x4_f32 (each wave a
// 16-query by 32-row quarter) and store a = |v|^2 - 2 q.v, which orders the
// rows of a query as |v - q|^2 does.  Eight lanes per query keep the kc
// smallest (a, id) keys exactly as knn_kernel keeps exact ones.  At the end
// every candidate's distance is recomputed with the reference's fixed-order
// chain (knn_kernel's arithmetic, bit-exact), the candidates are sorted by
// (distance, id) and the first k written.
//
// The result is exact when the k-th exact distance E_k satisfies
// E_k - |q|^2 + err < A_c, A_c being the largest kept a: every row left out
// has a >= A_c, so its exact distance exceeds E_k.  err bounds
// |a - (|v - q|^2 - |q|^2)| over the fp32 roundings involved (norm chain,
// matrix dot product, L2 chain, |q|^2), each at most g = 2 (D + 4) 2^-24 of
// the magnitudes it sums.  A query that fails the test (or whose candidate
// list is shorter than k) gets flag 1 and the host recomputes it with
// knn_kernel; a query whose list holds every row is exact by construction.
template <int D>
__global__ __launch_bounds__(256) void knn_mfma_kernel(const float *__restrict__ base,
                                                       const float *__restrict__ norms, uint32_t n,
                                                       const float *__restrict__ queries, uint32_t nq,
                                                       uint32_t kc, uint32_t k, int exclude_self, uint32_t qoff,
                                                       float vmax2, uint32_t *__restrict__ out_ids,
                                                       float *__restrict__ out_dists, uint32_t *__restrict__ flags)
{
    constexpr int RS = D + 4;  // padded LDS row stride (floats)
    extern __shared__ uint64_t sm[];
    uint64_t *lists = sm;                              // [QT][KCap]
    uint64_t *cbuf = lists + kKnnQT * kKnnKCap;        // [QT][64]
    uint64_t *bbs = cbuf + kKnnQT * 64;                // [4 waves][64]
    uint64_t *thr = bbs + 4 * 64;                      // [QT]
    float *rows = reinterpret_cast<float *>(thr + kKnnQT);  // [RT][RS]
    float *qsm = rows + kKnnRT * RS;                   // [QT][RS]
    float *dmat = qsm + kKnnQT * RS;                   // [RT][QT + 1] (padded against bank conflicts)
    float *tnorm = dmat + (kKnnQT + 1) * kKnnRT;       // [RT]
    uint32_t *ccnt = reinterpret_cast<uint32_t *>(tnorm + kKnnRT);  // [QT]
    uint32_t *lsz = ccnt + kKnnQT;                                  // [QT]

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const XorSel xs = make_xor_sel(lane);
    const int ql = tid / kKnnTPQ;
    const int sub = tid % kKnnTPQ;
    const uint32_t q = blockIdx.x * kKnnQT + ql;
    const bool qvalid = q < nq;

    // the workgroup's queries (query 0 stands in past nq)
    for (uint32_t c = tid; c < (uint32_t)kKnnQT * (D / 4); c += 256) {
        const uint32_t qq = c / (D / 4), e4 = c % (D / 4);
        const uint32_t src = blockIdx.x * kKnnQT + qq;
        const float4 v = reinterpret_cast<const float4 *>(queries + (size_t)(src < nq ? src : 0u) * D)[e4];
        *reinterpret_cast<float4 *>(&qsm[qq * RS + 4 * e4]) = v;
    }
    if (tid < kKnnQT) {
        ccnt[tid] = 0;
        lsz[tid] = 0;
        thr[tid] = kKeyMax;
    }
    __syncthreads();

    const int qb = wave >> 1, rh = wave & 1;
    uint64_t my_thr = kKeyMax;
    for (uint32_t r0 = 0; r0 < n; r0 += kKnnRT) {
        const uint32_t rn = min((uint32_t)kKnnRT, n - r0);
        {
            const float4 *src = reinterpret_cast<const float4 *>(base + (size_t)r0 * D);
            const uint32_t chunks = rn * (D / 4);
            for (uint32_t c = tid; c < chunks; c += 256) {
                const uint32_t r = c / (D / 4), e4 = c % (D / 4);
                *reinterpret_cast<float4 *>(&rows[r * RS + e4 * 4]) = src[c];
            }
            if ((uint32_t)tid < rn) tnorm[tid] = norms[r0 + tid];
        }
        __syncthreads();
        {
            // rows past rn hold an earlier tile (or nothing yet): their
            // products are computed and discarded
            knn_f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
            const float *qa = &qsm[(16 * qb + (lane & 15)) * RS + (lane >> 4)];
            const float *rb = &rows[(32 * rh + (lane & 15)) * RS + (lane >> 4)];
#pragma unroll
            for (int s4 = 0; s4 < D / 4; ++s4) {
                const float av = qa[4 * s4];
                acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, rb[4 * s4], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, rb[16 * RS + 4 * s4], acc1, 0, 0, 0);
            }
            const uint32_t j0 = 32u * rh + (lane & 15), j1 = j0 + 16u;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int qi = 16 * qb + 4 * (lane >> 4) + t;
                dmat[j0 * (kKnnQT + 1) + qi] = j0 < rn ? tnorm[j0] - 2.0f * acc0[t] : __int_as_float(0x7f800000);
                dmat[j1 * (kKnnQT + 1) + qi] = j1 < rn ? tnorm[j1] - 2.0f * acc1[t] : __int_as_float(0x7f800000);
            }
        }
        __syncthreads();
        if (qvalid) {
            for (uint32_t rr = sub; rr < rn; rr += kKnnTPQ) {
                const uint32_t id = r0 + rr;
                const uint64_t key = ((uint64_t)dkey<true>(dmat[rr * (kKnnQT + 1) + ql]) << 32) | id;
                if (key < my_thr && !(exclude_self && id == qoff + q)) {
                    const uint32_t slot = atomicAdd(&ccnt[ql], 1u);
                    cbuf[ql * 64 + slot] = key;
                }
            }
        }
        __syncthreads();
        for (int j = 0; j < kKnnQT / 4; ++j) {
            const int qq = wave * (kKnnQT / 4) + j;
            const uint32_t cnt = ccnt[qq];
            if (cnt == 0) continue;
            uint64_t x = (uint32_t)lane < cnt ? cbuf[qq * 64 + lane] : kKeyMax;
            x = sort64(x, lane, xs);
            uint32_t S = lsz[qq];
            merge_batch<true>(lists + qq * kKnnKCap, S, kc, x, cnt, bbs + wave * 64, lane);
            if (lane == 0) {
                lsz[qq] = S;
                thr[qq] = (S == kc) ? lists[qq * kKnnKCap + kc - 1] : kKeyMax;
                ccnt[qq] = 0;
            }
            wave_sync();
        }
        __syncthreads();
        my_thr = thr[ql];
    }
    const bool full = lsz[ql] == kc && my_thr != kKeyMax;
    const float a_c = full ? kdist<true>(my_thr) : 0.f;
    __syncthreads();

    // exact distances of the candidates, in place
    if (qvalid) {
        const uint32_t S = lsz[ql];
        const float *qr = &qsm[ql * RS];
        for (uint32_t i = sub; i < S; i += kKnnTPQ) {
            const uint32_t id = (uint32_t)lists[ql * kKnnKCap + i];
            const float *rv = base + (size_t)id * D;
            float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ii = 0; ii < D / 8; ++ii) {
                const int b = blk_of<D>(ii);
                const float4 lo = *reinterpret_cast<const float4 *>(&rv[8 * b]);
                const float4 hi = *reinterpret_cast<const float4 *>(&rv[8 * b + 4]);
                const float t0 = lo.x - qr[8 * b + 0], t1 = lo.y - qr[8 * b + 1];
                const float t2 = lo.z - qr[8 * b + 2], t3 = lo.w - qr[8 * b + 3];
                const float t4 = hi.x - qr[8 * b + 4], t5 = hi.y - qr[8 * b + 5];
                const float t6 = hi.z - qr[8 * b + 6], t7 = hi.w - qr[8 * b + 7];
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
            lists[ql * kKnnKCap + i] = mk_key(d, id);
        }
    }
    __syncthreads();
    // sort each query's exact keys (two batches of 64) and keep the first k
    for (int j = 0; j < kKnnQT / 4; ++j) {
        const int qq = wave * (kKnnQT / 4) + j;
        const uint32_t S = lsz[qq];
        if (S == 0) continue;
        uint64_t x0 = (uint32_t)lane < S ? lists[qq * kKnnKCap + lane] : kKeyMax;
        uint64_t x1 = 64u + lane < S ? lists[qq * kKnnKCap + 64 + lane] : kKeyMax;
        x0 = sort64(x0, lane, xs);
        x1 = sort64(x1, lane, xs);
        wave_sync();
        uint32_t Sk = 0;
        merge_batch<true>(lists + qq * kKnnKCap, Sk, k, x0, min(S, 64u), bbs + wave * 64, lane);
        if (S > 64u) merge_batch<true>(lists + qq * kKnnKCap, Sk, k, x1, S - 64u, bbs + wave * 64, lane);
        if (lane == 0) lsz[qq] = Sk;
        wave_sync();
    }
    __syncthreads();
    if (qvalid) {
        const uint32_t S = lsz[ql];
        for (uint32_t i = sub; i < k; i += kKnnTPQ) {
            const bool have = i < S;
            const uint64_t e = have ? lists[ql * kKnnKCap + i] : kKeyMax;
            out_ids[(size_t)q * k + i] = have ? (uint32_t)e : kSent;
            if (out_dists) out_dists[(size_t)q * k + i] = have ? __uint_as_float((uint32_t)(e >> 32)) : __int_as_float(0x7f800000);
        }
        if (sub == 0) {
            uint32_t flag = 0;
            if (S < k) {
                flag = 1;
            } else if (full) {
                const float *qr = &qsm[ql * RS];
                float qn2 = 0.f;
                for (int i = 0; i < D; ++i) qn2 = __builtin_fmaf(qr[i], qr[i], qn2);
                const float ek = __uint_as_float((uint32_t)(lists[ql * kKnnKCap + k - 1] >> 32));
                const float vm = sqrtf(vmax2), qn = sqrtf(qn2);
                const float g = 2.0f * (float)(D + 4) * 5.9604645e-8f;
                const float err = g * (vmax2 + 4.0f * vm * qn + (vm + qn) * (vm + qn) + qn2) +
                                  9.5367432e-7f * (vm + qn) * (vm + qn);
                flag = (ek - qn2 + err < a_c) ? 0u : 1u;
            }
            flags[q] = flag;
        }
    }
}

}  // namespace panns
