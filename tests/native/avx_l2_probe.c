/*
 * Test helper: squared L2 written with 8-wide AVX intrinsics in the shape of
 * the reference's compute_distance (PSS5LG.cpp:525-560: the D%16 tail block
 * first, then 16 floats per step as two 8-wide halves, one running 8-lane sum,
 * horizontal sum left to right).  Built twice by the tests: with GCC's
 * default contraction (-mfma, as the reference's -O3 -march=native build
 * gets) and with -ffp-contract=off.  The first must equal the oracle's fmaf
 * chain bit for bit; the second shows the comparison can tell them apart.
 */
#include <immintrin.h>

static __m256 step8(__m256 acc, const float *a, const float *b)
{
    __m256 d = _mm256_sub_ps(_mm256_loadu_ps(a), _mm256_loadu_ps(b));
    return _mm256_add_ps(acc, _mm256_mul_ps(d, d));
}

float probe_l2(const float *v, const float *q, unsigned dim)
{
    const unsigned padded = (dim + 7u) & ~7u;
    const unsigned tail = padded % 16u;
    const unsigned body = padded - tail;
    float lanes[8] __attribute__((aligned(32))) = {0, 0, 0, 0, 0, 0, 0, 0};
    __m256 acc = _mm256_load_ps(lanes);
    if (tail) acc = step8(acc, v + body, q + body);
    for (unsigned i = 0; i < body; i += 16) {
        acc = step8(acc, v + i, q + i);
        acc = step8(acc, v + i + 8, q + i + 8);
    }
    _mm256_store_ps(lanes, acc);
    return lanes[0] + lanes[1] + lanes[2] + lanes[3] + lanes[4] + lanes[5] + lanes[6] + lanes[7];
}

/* The same structure for the opt-graph kernels: the norm (PSS5.cpp:293-334,
 * an 8-lane sum of v*v) and the dot product with -2 * dot + norm
 * (PSS5.cpp:336-390).  A product followed by an add in the source, which
 * GCC contracts into one fma per element when contraction is on. */
static __m256 dot8(__m256 acc, const float *a, const float *b)
{
    return _mm256_add_ps(acc, _mm256_mul_ps(_mm256_loadu_ps(a), _mm256_loadu_ps(b)));
}

static float dot_chain(const float *v, const float *q, unsigned dim)
{
    const unsigned padded = (dim + 7u) & ~7u;
    const unsigned tail = padded % 16u;
    const unsigned body = padded - tail;
    float lanes[8] __attribute__((aligned(32))) = {0, 0, 0, 0, 0, 0, 0, 0};
    __m256 acc = _mm256_load_ps(lanes);
    if (tail) acc = dot8(acc, v + body, q + body);
    for (unsigned i = 0; i < body; i += 16) {
        acc = dot8(acc, v + i, q + i);
        acc = dot8(acc, v + i + 8, q + i + 8);
    }
    _mm256_store_ps(lanes, acc);
    return lanes[0] + lanes[1] + lanes[2] + lanes[3] + lanes[4] + lanes[5] + lanes[6] + lanes[7];
}

float probe_norm(const float *v, unsigned dim) { return dot_chain(v, v, dim); }

float probe_dot_norm(const float *v, const float *q, unsigned dim, float norm)
{
    float r = dot_chain(v, q, dim);
    return -2 * r + norm;
}
