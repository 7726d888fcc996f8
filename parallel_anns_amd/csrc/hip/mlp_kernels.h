// Training of the learned-stop MLP (8 -> 64 -> 64 -> 1, ReLU, ReLU,
// sigmoid, binary cross-entropy) with BF16 MFMA and FP32 accumulation.
//
// One workgroup (4 waves) takes a tile of 64 samples.  Every GEMM of the
// forward and backward pass is a set of 16x16x16 MFMA tiles
// (v_mfma_f32_16x16x16_bf16): operands are read from fp32 LDS arrays and
// rounded to bf16 (nearest even) as fragments are built, products accumulate
// in fp32.  The 64-wide output layer and the bias sums are plain fp32.
// Each workgroup writes its partial gradient; mlp_adam_kernel sums the
// partials and applies Adam to the fp32 master weights.
//
// Fragment layout of the 16x16x16 bf16 MFMA on CDNA3/4, lane l (0..63):
//   A (16 rows x 16 k):  row l%16, k = 4*(l/16) .. +3
//   B (16 k x 16 cols):  col l%16, k = 4*(l/16) .. +3
//   D (16 rows x 16 cols): col l%16, rows 4*(l/16) .. +3
// (checked on the device by panns_mfma_probe against a host product).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace panns {

constexpr int kMlpTile = 64;      // samples per workgroup
constexpr int kMlpIn = 8;
constexpr int kMlpH = 64;
// parameter layout (floats): W1[64][8] b1[64] W2[64][64] b2[64] w3[64] b3
constexpr int kOffW1 = 0, kOffB1 = 512, kOffW2 = 576, kOffB2 = 4672, kOffW3 = 4736, kOffB3 = 4800;
constexpr int kMlpParams = 4801;
constexpr int kMlpParamsPad = 4864;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// acc += A[m0.., k0..] * B[k0.., n0..] over K (multiple of 16) for one 16x16
// output tile.  A(i, k) = A[i*lda + k] (or A[k*lda + i] when ta), B(k, j) =
// B[k*ldb + j] (or B[j*ldb + k] when tb); rows/cols beyond the real extent are
// handled by the caller padding its LDS arrays with zeros.
__device__ __forceinline__ f32x4 tile_mma(const float *A, int lda, bool ta, const float *B, int ldb,
                                          bool tb, int m0, int n0, int K, f32x4 acc, int lane)
{
    const int r = lane & 15, kb = 4 * (lane >> 4);
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int k = k0 + kb + t;
            const float av = ta ? A[k * lda + (m0 + r)] : A[(m0 + r) * lda + k];
            const float bv = tb ? B[(n0 + r) * ldb + k] : B[k * ldb + (n0 + r)];
            a[t] = to_bf16(av);
            b[t] = to_bf16(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// Stores a 16x16 fp32 tile (D layout) at C[(m0+i)*ldc + n0+j].
__device__ __forceinline__ void tile_store(float *C, int ldc, int m0, int n0, f32x4 d, int lane)
{
    const int j = lane & 15, ib = 4 * (lane >> 4);
#pragma unroll
    for (int t = 0; t < 4; ++t) C[(m0 + ib + t) * ldc + n0 + j] = d[t];
}

// One 16x16x16 product for the layout check: C = A * B (row-major fp32 in,
// bf16-rounded operands, fp32 out).
__global__ void mfma_probe_kernel(const float *A, const float *B, float *C)
{
    const int lane = threadIdx.x;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = tile_mma(A, 16, false, B, 16, false, 0, 0, 16, acc, lane);
    tile_store(C, 16, 0, 0, acc, lane);
}

// Forward + backward of one 64-sample tile.  X: [n][8] normalised features,
// Y: [n] labels, perm: sample order (or null), first: index of this batch's
// first sample in perm, count: samples in the batch.  Writes the workgroup's
// gradient (already divided by the batch size) to part[blockIdx.x][...] and
// its summed loss to loss[blockIdx.x].
__global__ __launch_bounds__(256) void mlp_fwd_bwd_kernel(const float *__restrict__ P,
                                                          const float *__restrict__ X,
                                                          const float *__restrict__ Y,
                                                          const uint32_t *__restrict__ perm,
                                                          uint64_t first, uint32_t count,
                                                          float *__restrict__ part,
                                                          float *__restrict__ loss)
{
    constexpr int T = kMlpTile, H = kMlpH, XI = 16;  // inputs padded to 16 for the MFMA k
    __shared__ float sX[T * XI];      // inputs
    __shared__ float sW1[H * XI];     // W1 padded to 16 columns
    __shared__ float sW2[H * H];
    __shared__ float sH1[T * H];      // relu(Z1)
    __shared__ float sH2[T * H];      // relu(Z2)
    __shared__ float sD2[T * H];      // dZ2
    __shared__ float sD1[T * H];      // dH1 then dZ1
    __shared__ float sDo[T];          // dL/do per sample
    __shared__ float sLoss[T];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t s0 = blockIdx.x * T;
    const float inv_b = 1.f / (float)count;

    for (int i = tid; i < T * XI; i += 256) {
        const int s = i / XI, f = i % XI;
        float v = 0.f;
        if (f < kMlpIn && s0 + s < count) {
            const uint64_t g = first + s0 + s;
            const uint64_t row = perm ? perm[g] : g;
            v = X[row * kMlpIn + f];
        }
        sX[i] = v;
    }
    for (int i = tid; i < H * XI; i += 256) {
        const int j = i / XI, f = i % XI;
        sW1[i] = f < kMlpIn ? P[kOffW1 + j * kMlpIn + f] : 0.f;
    }
    for (int i = tid; i < H * H; i += 256) sW2[i] = P[kOffW2 + i];
    __syncthreads();

    // Z1 = X W1^T + b1 -> H1 (wave w owns sample rows 16w..16w+15)
    const int m0 = 16 * wave;
    for (int n0 = 0; n0 < H; n0 += 16) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = tile_mma(sX, XI, false, sW1, XI, true, m0, n0, XI, acc, lane);
        const int j = n0 + (lane & 15);
        const float b = P[kOffB1 + j];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const float z = acc[t] + b;
            acc[t] = z > 0.f ? z : 0.f;
        }
        tile_store(sH1, H, m0, n0, acc, lane);
    }
    __syncthreads();
    // Z2 = H1 W2^T + b2 -> H2
    for (int n0 = 0; n0 < H; n0 += 16) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = tile_mma(sH1, H, false, sW2, H, true, m0, n0, H, acc, lane);
        const int j = n0 + (lane & 15);
        const float b = P[kOffB2 + j];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const float z = acc[t] + b;
            acc[t] = z > 0.f ? z : 0.f;
        }
        tile_store(sH2, H, m0, n0, acc, lane);
    }
    __syncthreads();
    // o = H2 w3 + b3, p = sigmoid(o), loss and dL/do (fp32, one lane per sample)
    if (tid < T) {
        const int s = tid;
        float o = P[kOffB3];
        for (int j = 0; j < H; ++j) o = __builtin_fmaf(sH2[s * H + j], P[kOffW3 + j], o);
        const bool real = s0 + s < count;
        float y = 0.f;
        if (real) {
            const uint64_t g = first + s0 + s;
            y = Y[perm ? perm[g] : g];
        }
        const float p = 1.f / (1.f + __expf(-o));
        // numerically stable BCE: softplus(o) - y*o
        const float l = fmaxf(o, 0.f) - y * o + log1pf(__expf(-fabsf(o)));
        sLoss[s] = real ? l : 0.f;
        sDo[s] = real ? (p - y) * inv_b : 0.f;
    }
    __syncthreads();
    // dZ2 = (dL/do * w3) masked by ReLU'
    for (int i = tid; i < T * H; i += 256) {
        const int s = i / H, j = i % H;
        sD2[i] = sH2[i] > 0.f ? sDo[s] * P[kOffW3 + j] : 0.f;
    }
    __syncthreads();
    float *g = part + (size_t)blockIdx.x * kMlpParamsPad;
    // dW2 = dZ2^T H1 (64x64, k = samples); wave w owns rows 16w..
    for (int n0 = 0; n0 < H; n0 += 16) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = tile_mma(sD2, H, true, sH1, H, false, m0, n0, T, acc, lane);
        tile_store(g + kOffW2, H, m0, n0, acc, lane);
    }
    // dH1 = dZ2 W2, then masked by ReLU' into dZ1
    for (int n0 = 0; n0 < H; n0 += 16) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = tile_mma(sD2, H, false, sW2, H, false, m0, n0, H, acc, lane);
        tile_store(sD1, H, m0, n0, acc, lane);
    }
    __syncthreads();
    for (int i = tid; i < T * H; i += 256) sD1[i] = sH1[i] > 0.f ? sD1[i] : 0.f;
    __syncthreads();
    // dW1 = dZ1^T X (64 x 16, only 8 columns real): wave w owns rows 16w..
    {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = tile_mma(sD1, H, true, sX, XI, false, m0, 0, T, acc, lane);
        const int j = lane & 15, ib = 4 * (lane >> 4);
        if (j < kMlpIn) {
#pragma unroll
            for (int t = 0; t < 4; ++t) g[kOffW1 + (m0 + ib + t) * kMlpIn + j] = acc[t];
        }
    }
    // bias and output-layer gradients (fp32 sums over the tile)
    if (tid < H) {
        const int j = tid;
        float db1 = 0.f, db2 = 0.f, dw3 = 0.f;
        for (int s = 0; s < T; ++s) {
            db1 += sD1[s * H + j];
            db2 += sD2[s * H + j];
            dw3 = __builtin_fmaf(sDo[s], sH2[s * H + j], dw3);
        }
        g[kOffB1 + j] = db1;
        g[kOffB2 + j] = db2;
        g[kOffW3 + j] = dw3;
    }
    if (tid == 64) {
        float db3 = 0.f, ls = 0.f;
        for (int s = 0; s < T; ++s) {
            db3 += sDo[s];
            ls += sLoss[s];
        }
        g[kOffB3] = db3;
        loss[blockIdx.x] = ls;
    }
}

// Sums nparts partial gradients and applies one Adam step (bias-corrected
// with step t) to the fp32 parameters.
__global__ void mlp_adam_kernel(float *__restrict__ P, const float *__restrict__ part, uint32_t nparts,
                                float *__restrict__ m1, float *__restrict__ m2, float lr, float beta1,
                                float beta2, float eps, float t, float *__restrict__ grad_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    float gsum = 0.f;
    for (uint32_t b = 0; b < nparts; ++b) gsum += part[(size_t)b * kMlpParamsPad + i];
    if (grad_out) grad_out[i] = gsum;
    if (lr == 0.f) return;
    const float a = beta1 * m1[i] + (1.f - beta1) * gsum;
    const float v = beta2 * m2[i] + (1.f - beta2) * gsum * gsum;
    m1[i] = a;
    m2[i] = v;
    const float ah = a / (1.f - __powf(beta1, t));
    const float vh = v / (1.f - __powf(beta2, t));
    P[i] -= lr * ah / (sqrtf(vh) + eps);
}

}  // namespace panns
