// Host tools: synthetic inputs, file formats, exact brute force, graph build.
// Built with -ffp-contract=off: the only fused multiply-adds are explicit fmaf
// calls, so pt_l2 reproduces the reference's compiled summation order and the
// generator is bit-reproducible across x86-64 hosts.
#include "panns_tools.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <omp.h>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int nthreads_or_default(int t) { return t > 0 ? t : omp_get_max_threads(); }

// ---------------------------------------------------------------------------
// Deterministic random numbers.
inline uint64_t splitmix64(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Xoshiro256 {
    uint64_t s[4];
    Xoshiro256(uint64_t seed, uint64_t stream)
    {
        uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull) ^ 0x243F6A8885A308D3ull;
        for (auto &v : s) v = splitmix64(x);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform in (0, 1) with 53 random bits
    double uniform() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
};

// Natural log from IEEE basic operations only (no libm): x = m * 2^e with
// m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s), s = (m-1)/(m+1), |s| < 0.172.
double det_log(double x)
{
    uint64_t b;
    std::memcpy(&b, &x, 8);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    b = (b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m;
    std::memcpy(&m, &b, 8);
    if (m > 1.4142135623730951) {
        m *= 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double term = s, sum = 0.0;
    for (int k = 0; k < 16; ++k) {
        sum += term / (double)(2 * k + 1);
        term *= s2;
    }
    return 2.0 * sum + (double)e * 0.69314718055994530942;
}

struct Normal {
    Xoshiro256 rng;
    bool has_spare = false;
    double spare = 0.0;
    Normal(uint64_t seed, uint64_t stream) : rng(seed, stream) {}
    // Marsaglia polar method.
    double next()
    {
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        for (;;) {
            const double u = 2.0 * rng.uniform() - 1.0;
            const double v = 2.0 * rng.uniform() - 1.0;
            const double s = u * u + v * v;
            if (s >= 1.0 || s == 0.0) continue;
            const double f = std::sqrt(-2.0 * det_log(s) / s);
            spare = v * f;
            has_spare = true;
            return u * f;
        }
    }
};

// ---------------------------------------------------------------------------
struct Cand {
    float dist;
    uint32_t id;
    bool operator<(const Cand &o) const { return dist != o.dist ? dist < o.dist : id < o.id; }
    bool operator==(const Cand &o) const { return dist == o.dist && id == o.id; }
};

// Bounded sorted list keeping the k smallest (dist, id); the behaviour of
// add_into_queue (PSS5LG.cpp:565-608) on distinct ids.
struct TopK {
    std::vector<Cand> v;
    uint32_t k;
    explicit TopK(uint32_t k_) : k(k_) { v.reserve(k_ + 1); }
    float bound() const { return v.size() < k ? INFINITY : v.back().dist; }
    void offer(float d, uint32_t id)
    {
        Cand c{d, id};
        if (v.size() == k) {
            if (!(c < v.back())) return;
            v.pop_back();
        }
        v.insert(std::upper_bound(v.begin(), v.end(), c), c);
    }
};

std::vector<uint64_t> new_bits(uint32_t n) { return std::vector<uint64_t>((n + 63) / 64, 0); }
inline bool bget(const std::vector<uint64_t> &b, uint32_t i) { return (b[i >> 6] >> (i & 63)) & 1u; }
inline void bset(std::vector<uint64_t> &b, uint32_t i) { b[i >> 6] |= 1ull << (i & 63); }
inline void bclr(std::vector<uint64_t> &b, uint32_t i) { b[i >> 6] &= ~(1ull << (i & 63)); }

}  // namespace

extern "C" const char *pt_last_error(void) { return g_err.c_str(); }

extern "C" float pt_l2(const float *v, const float *q, uint32_t dim)
{
    const uint32_t D8 = (dim + 7u) & ~7u;
    const uint32_t DR = D8 % 16u;
    const uint32_t DD = D8 - DR;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (DR) {
        for (int j = 0; j < 8; ++j) {
            const float t = v[DD + j] - q[DD + j];
            acc[j] = std::fmaf(t, t, acc[j]);
        }
    }
    for (uint32_t i = 0; i < DD; i += 16) {
        for (int j = 0; j < 8; ++j) {
            const float t = v[i + j] - q[i + j];
            acc[j] = std::fmaf(t, t, acc[j]);
        }
        for (int j = 0; j < 8; ++j) {
            const float t = v[i + 8 + j] - q[i + 8 + j];
            acc[j] = std::fmaf(t, t, acc[j]);
        }
    }
    float r = acc[0];
    for (int j = 1; j < 8; ++j) r = r + acc[j];
    return r;
}

extern "C" int pt_gen_mixture(uint64_t seed_centers, uint64_t seed_rows, uint64_t row0,
                              uint32_t n, uint32_t dim, uint32_t n_comp, float sigma,
                              int normalize, float *out, int nthreads)
{
    if (!out || dim == 0 || n_comp == 0) return fail(PT_ERR_ARG, "pt_gen_mixture: bad argument");
    std::vector<double> centers((size_t)n_comp * dim);
    for (uint32_t c = 0; c < n_comp; ++c) {
        Normal g(seed_centers, c);
        for (uint32_t e = 0; e < dim; ++e) centers[(size_t)c * dim + e] = g.next();
    }
    const double sg = (double)sigma;
#pragma omp parallel for schedule(static) num_threads(nthreads_or_default(nthreads))
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        Normal g(seed_rows, row0 + (uint64_t)i);
        const uint32_t label = g.rng.below(n_comp);
        std::vector<double> x(dim);
        double ss = 0.0;
        for (uint32_t e = 0; e < dim; ++e) {
            const double z = g.next();
            const double prod = sg * z;
            x[e] = centers[(size_t)label * dim + e] + prod;
        }
        for (uint32_t e = 0; e < dim; ++e) {
            const double sq = x[e] * x[e];
            ss = ss + sq;
        }
        const double norm = normalize ? std::sqrt(ss) : 1.0;
        for (uint32_t e = 0; e < dim; ++e) out[(size_t)i * dim + e] = (float)(x[e] / norm);
    }
    return PT_OK;
}

extern "C" int pt_mixture_labels(uint64_t seed_rows, uint64_t row0, uint32_t n, uint32_t n_comp,
                                 uint32_t *labels, int nthreads)
{
    if (!labels || n_comp == 0) return fail(PT_ERR_ARG, "pt_mixture_labels: bad argument");
#pragma omp parallel for schedule(static) num_threads(nthreads_or_default(nthreads))
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        Normal g(seed_rows, row0 + (uint64_t)i);
        labels[i] = g.rng.below(n_comp);
    }
    return PT_OK;
}

// ---------------------------------------------------------------------------
// fvecs: per row int32 D followed by D float32 (include/utils.h:95-142).
extern "C" int pt_fvecs_info(const char *path, uint32_t *n, uint32_t *dim)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(PT_ERR_IO, std::string("cannot open ") + path);
    int32_t d = 0;
    if (std::fread(&d, 4, 1, f) != 1 || d <= 0) {
        std::fclose(f);
        return fail(PT_ERR_FORMAT, std::string("bad fvecs header in ") + path);
    }
    std::fseek(f, 0, SEEK_END);
    const uint64_t sz = (uint64_t)std::ftell(f);
    std::fclose(f);
    const uint64_t row = 4ull + 4ull * (uint64_t)d;
    if (sz % row) return fail(PT_ERR_FORMAT, std::string("fvecs size not a multiple of the row in ") + path);
    if (sz / row > 0xFFFFFFFFull) return fail(PT_ERR_RANGE, "too many rows");
    *n = (uint32_t)(sz / row);
    *dim = (uint32_t)d;
    return PT_OK;
}

extern "C" int pt_fvecs_read(const char *path, float *out, uint32_t n, uint32_t dim)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(PT_ERR_IO, std::string("cannot open ") + path);
    for (uint32_t i = 0; i < n; ++i) {
        int32_t d = 0;
        if (std::fread(&d, 4, 1, f) != 1 || (uint32_t)d != dim ||
            std::fread(out + (size_t)i * dim, 4, dim, f) != dim) {
            std::fclose(f);
            return fail(PT_ERR_FORMAT, std::string("short or inconsistent fvecs row in ") + path);
        }
    }
    std::fclose(f);
    return PT_OK;
}

extern "C" int pt_fvecs_write(const char *path, const float *data, uint32_t n, uint32_t dim)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(PT_ERR_IO, std::string("cannot create ") + path);
    const int32_t d = (int32_t)dim;
    for (uint32_t i = 0; i < n; ++i) {
        if (std::fwrite(&d, 4, 1, f) != 1 || std::fwrite(data + (size_t)i * dim, 4, dim, f) != dim) {
            std::fclose(f);
            return fail(PT_ERR_IO, std::string("write failed: ") + path);
        }
    }
    std::fclose(f);
    return PT_OK;
}

// ---------------------------------------------------------------------------
// .nsg: uint32 width, uint32 ep, then per vertex uint32 deg + deg uint32 ids.
extern "C" int pt_nsg_info(const char *path, uint32_t *width, uint32_t *ep, uint32_t *n,
                           uint64_t *num_e, uint32_t *max_deg)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(PT_ERR_IO, std::string("cannot open ") + path);
    uint32_t hdr[2];
    if (std::fread(hdr, 4, 2, f) != 2) {
        std::fclose(f);
        return fail(PT_ERR_FORMAT, "truncated .nsg header");
    }
    uint64_t cnt = 0, e = 0;
    uint32_t md = 0, deg = 0;
    while (std::fread(&deg, 4, 1, f) == 1) {
        if (std::fseek(f, (long)deg * 4, SEEK_CUR) != 0) break;
        e += deg;
        md = std::max(md, deg);
        ++cnt;
    }
    const long pos = std::ftell(f);
    std::fseek(f, 0, SEEK_END);
    const long end = std::ftell(f);
    std::fclose(f);
    if (pos != end) return fail(PT_ERR_FORMAT, "truncated .nsg adjacency list");
    uint64_t expect = 8 + 4 * cnt + 4 * e;
    if ((uint64_t)end != expect) return fail(PT_ERR_FORMAT, "truncated .nsg adjacency list");
    if (cnt > 0xFFFFFFFFull) return fail(PT_ERR_RANGE, "too many vertices");
    *width = hdr[0];
    *ep = hdr[1];
    *n = (uint32_t)cnt;
    *num_e = e;
    *max_deg = md;
    return PT_OK;
}

extern "C" int pt_nsg_read(const char *path, uint32_t n, uint32_t stride, uint32_t *adj,
                           uint32_t *deg)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(PT_ERR_IO, std::string("cannot open ") + path);
    uint32_t hdr[2];
    if (std::fread(hdr, 4, 2, f) != 2) {
        std::fclose(f);
        return fail(PT_ERR_FORMAT, "truncated .nsg header");
    }
    for (uint32_t v = 0; v < n; ++v) {
        uint32_t d = 0;
        if (std::fread(&d, 4, 1, f) != 1) {
            std::fclose(f);
            return fail(PT_ERR_FORMAT, ".nsg has fewer vertices than expected");
        }
        if (d > stride) {
            std::fclose(f);
            return fail(PT_ERR_RANGE, ".nsg degree exceeds the row stride");
        }
        if (std::fread(adj + (size_t)v * stride, 4, d, f) != d) {
            std::fclose(f);
            return fail(PT_ERR_FORMAT, "truncated .nsg adjacency list");
        }
        deg[v] = d;
    }
    std::fclose(f);
    return PT_OK;
}

extern "C" int pt_nsg_write(const char *path, uint32_t width, uint32_t ep, uint32_t n,
                            uint32_t stride, const uint32_t *adj, const uint32_t *deg)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(PT_ERR_IO, std::string("cannot create ") + path);
    uint32_t hdr[2] = {width, ep};
    bool ok = std::fwrite(hdr, 4, 2, f) == 2;
    for (uint32_t v = 0; ok && v < n; ++v) {
        ok = std::fwrite(&deg[v], 4, 1, f) == 1 &&
             std::fwrite(adj + (size_t)v * stride, 4, deg[v], f) == deg[v];
    }
    std::fclose(f);
    return ok ? PT_OK : fail(PT_ERR_IO, std::string("write failed: ") + path);
}

// ---------------------------------------------------------------------------
extern "C" int pt_truth_info(const char *path, uint32_t *nq, uint32_t *k)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(PT_ERR_IO, std::string("cannot open ") + path);
    uint32_t hdr[2];
    const bool ok = std::fread(hdr, 4, 2, f) == 2;
    std::fclose(f);
    if (!ok) return fail(PT_ERR_FORMAT, "truncated truth header");
    *nq = hdr[0];
    *k = hdr[1];
    return PT_OK;
}

extern "C" int pt_truth_read(const char *path, uint32_t nq, uint32_t k, uint32_t *ids,
                             float *dists)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(PT_ERR_IO, std::string("cannot open ") + path);
    uint32_t hdr[2];
    if (std::fread(hdr, 4, 2, f) != 2 || hdr[0] < nq || hdr[1] != k) {
        std::fclose(f);
        return fail(PT_ERR_FORMAT, "truth header does not match");
    }
    for (size_t i = 0; i < (size_t)nq * k; ++i) {
        uint32_t rec[2];
        if (std::fread(rec, 4, 2, f) != 2) {
            std::fclose(f);
            return fail(PT_ERR_FORMAT, "truncated truth file");
        }
        ids[i] = rec[0];
        if (dists) std::memcpy(&dists[i], &rec[1], 4);
    }
    std::fclose(f);
    return PT_OK;
}

extern "C" int pt_truth_write(const char *path, uint32_t nq, uint32_t k, const uint32_t *ids,
                              const float *dists)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(PT_ERR_IO, std::string("cannot create ") + path);
    uint32_t hdr[2] = {nq, k};
    bool ok = std::fwrite(hdr, 4, 2, f) == 2;
    for (size_t i = 0; ok && i < (size_t)nq * k; ++i) {
        uint32_t rec[2];
        rec[0] = ids[i];
        std::memcpy(&rec[1], &dists[i], 4);
        ok = std::fwrite(rec, 4, 2, f) == 2;
    }
    std::fclose(f);
    return ok ? PT_OK : fail(PT_ERR_IO, "truth write failed");
}

extern "C" int pt_result_write(const char *path, uint32_t nq, uint32_t k, const uint32_t *ids)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(PT_ERR_IO, std::string("cannot create ") + path);
    bool ok = true;
    for (uint32_t q = 0; ok && q < nq; ++q)
        ok = std::fwrite(&k, 4, 1, f) == 1 && std::fwrite(ids + (size_t)q * k, 4, k, f) == k;
    std::fclose(f);
    return ok ? PT_OK : fail(PT_ERR_IO, "result write failed");
}

// ---------------------------------------------------------------------------
extern "C" int pt_exact_knn(const float *base, uint32_t n, uint32_t dim, const float *queries,
                            uint32_t nq, uint32_t k, int exclude_self, uint32_t *out_ids,
                            float *out_dists, int nthreads)
{
    if (dim % 8) return fail(PT_ERR_ARG, "dimension must be a multiple of 8");
    if (k == 0) return fail(PT_ERR_ARG, "k must be positive");
    const uint32_t QB = 8;  // queries per pass over the base, so each row is reused
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads_or_default(nthreads))
    for (int64_t qb = 0; qb < (int64_t)((nq + QB - 1) / QB); ++qb) {
        const uint32_t q0 = (uint32_t)qb * QB;
        const uint32_t qn = std::min(QB, nq - q0);
        std::vector<TopK> tk;
        for (uint32_t j = 0; j < qn; ++j) tk.emplace_back(k);
        for (uint32_t v = 0; v < n; ++v) {
            const float *row = base + (size_t)v * dim;
            for (uint32_t j = 0; j < qn; ++j) {
                const uint32_t q = q0 + j;
                if (exclude_self && v == q) continue;
                const float d = pt_l2(row, queries + (size_t)q * dim, dim);
                if (d <= tk[j].bound()) tk[j].offer(d, v);
            }
        }
        for (uint32_t j = 0; j < qn; ++j) {
            const uint32_t q = q0 + j;
            for (uint32_t i = 0; i < k; ++i) {
                const bool have = i < tk[j].v.size();
                out_ids[(size_t)q * k + i] = have ? tk[j].v[i].id : 0xFFFFFFFFu;
                if (out_dists) out_dists[(size_t)q * k + i] = have ? tk[j].v[i].dist : INFINITY;
            }
        }
    }
    return PT_OK;
}

extern "C" double pt_recall(const uint32_t *true_ids, uint32_t true_k, const uint32_t *res_ids,
                            uint32_t res_k, uint32_t nq, uint32_t L, uint32_t R)
{
    const uint32_t m = std::min(std::min(L, R), res_k);
    double hits = 0.0;
    for (uint32_t q = 0; q < nq; ++q) {
        for (uint32_t t = 0; t < R; ++t) {
            const uint32_t tid = true_ids[(size_t)q * true_k + t];
            for (uint32_t i = 0; i < m; ++i)
                if (res_ids[(size_t)q * res_k + i] == tid) hits += 1.0;
        }
    }
    return hits / ((double)R * (double)nq);
}

extern "C" int pt_recall_per_query(const uint32_t *true_ids, uint32_t true_k,
                                   const uint32_t *res_ids, uint32_t res_k, uint32_t nq,
                                   uint32_t L, uint32_t R, float *out)
{
    if (R > true_k) return fail(PT_ERR_ARG, "R exceeds the truth depth");
    const uint32_t m = std::min(std::min(L, R), res_k);
    for (uint32_t q = 0; q < nq; ++q) {
        uint32_t hits = 0;
        for (uint32_t t = 0; t < R; ++t) {
            const uint32_t tid = true_ids[(size_t)q * true_k + t];
            for (uint32_t i = 0; i < m; ++i)
                if (res_ids[(size_t)q * res_k + i] == tid) ++hits;
        }
        out[q] = (float)hits / (float)R;
    }
    return PT_OK;
}

// ---------------------------------------------------------------------------
extern "C" int pt_navigating_node(const float *base, uint32_t n, uint32_t dim, uint32_t *ep,
                                  int nthreads)
{
    if (n == 0 || dim % 8) return fail(PT_ERR_ARG, "pt_navigating_node: bad argument");
    // Fixed chunking keeps the fp64 sum independent of the thread count.
    const uint32_t CH = 4096;
    const uint32_t nch = (n + CH - 1) / CH;
    std::vector<double> part((size_t)nch * dim, 0.0);
#pragma omp parallel for schedule(static) num_threads(nthreads_or_default(nthreads))
    for (int64_t c = 0; c < (int64_t)nch; ++c) {
        const uint32_t lo = (uint32_t)c * CH, hi = std::min(n, lo + CH);
        for (uint32_t v = lo; v < hi; ++v)
            for (uint32_t e = 0; e < dim; ++e) part[(size_t)c * dim + e] += (double)base[(size_t)v * dim + e];
    }
    std::vector<float> center(dim);
    for (uint32_t e = 0; e < dim; ++e) {
        double s = 0.0;
        for (uint32_t c = 0; c < nch; ++c) s += part[(size_t)c * dim + e];
        center[e] = (float)(s / (double)n);
    }
    uint32_t best = 0;
    float bd = INFINITY;
    for (uint32_t v = 0; v < n; ++v) {
        const float d = pt_l2(base + (size_t)v * dim, center.data(), dim);
        if (d < bd) {
            bd = d;
            best = v;
        }
    }
    *ep = best;
    return PT_OK;
}

namespace {

// RNG occlusion (index_nsg.cpp:233-283): walk the pool in (dist, id) order;
// keep p unless some kept r has d(r, p) < d(v, p).
void rng_prune(const float *base, uint32_t dim, const std::vector<Cand> &pool, uint32_t R,
               uint32_t C, std::vector<Cand> &result)
{
    result.clear();
    for (size_t i = 0; i < pool.size() && i < C && result.size() < R; ++i) {
        const Cand &p = pool[i];
        bool occluded = false;
        for (const Cand &r : result) {
            if (r.id == p.id) {
                occluded = true;
                break;
            }
            const float djk = pt_l2(base + (size_t)r.id * dim, base + (size_t)p.id * dim, dim);
            if (djk < p.dist) {
                occluded = true;
                break;
            }
        }
        if (!occluded) result.push_back(p);
    }
}

// Best-first search for x over the k-NN graph from ep; returns every vertex
// it expanded, with its distance to x.
void knn_graph_search(const float *base, uint32_t dim, const uint32_t *knn, uint32_t knn_k,
                      uint32_t ep, const float *x, uint32_t Lb, std::vector<uint32_t> &stamp,
                      uint32_t epoch, std::vector<Cand> &expanded)
{
    struct QE {
        Cand c;
        bool checked;
    };
    std::vector<QE> q;
    q.reserve(Lb + 1);
    expanded.clear();
    auto visit = [&](uint32_t v) {
        if (stamp[v] == epoch) return false;
        stamp[v] = epoch;
        return true;
    };
    visit(ep);
    q.push_back({{pt_l2(base + (size_t)ep * dim, x, dim), ep}, false});
    size_t k = 0;
    while (k < q.size()) {
        if (q[k].checked) {
            ++k;
            continue;
        }
        q[k].checked = true;
        const Cand cur = q[k].c;
        expanded.push_back(cur);
        size_t nk = q.size();
        for (uint32_t i = 0; i < knn_k; ++i) {
            const uint32_t nb = knn[(size_t)cur.id * knn_k + i];
            if (nb == 0xFFFFFFFFu || !visit(nb)) continue;
            const Cand c{pt_l2(base + (size_t)nb * dim, x, dim), nb};
            if (q.size() == Lb && !(c < q.back().c)) continue;
            if (q.size() == Lb) q.pop_back();
            auto it = std::upper_bound(q.begin(), q.end(), c,
                                       [](const Cand &a, const QE &b) { return a < b.c; });
            const size_t pos = (size_t)(it - q.begin());
            q.insert(it, {c, false});
            nk = std::min(nk, pos);
        }
        k = std::min(nk, k + 1);
    }
}

}  // namespace

extern "C" int pt_build_graph(const float *base, uint32_t n, uint32_t dim, const uint32_t *knn,
                              uint32_t knn_k, uint32_t R, uint32_t L_build, uint32_t C,
                              uint32_t ep, uint32_t options, uint32_t *adj, uint32_t *deg,
                              int nthreads)
{
    const bool reverse_edges = (options & PT_GRAPH_REVERSE_EDGES) != 0;
    if (n == 0 || R == 0 || dim % 8 || ep >= n) return fail(PT_ERR_ARG, "pt_build_graph: bad argument");
    const int nt = nthreads_or_default(nthreads);
    std::vector<std::vector<Cand>> out(n);
    // 1-2: pool = kNN + search path, RNG prune.
#pragma omp parallel num_threads(nt)
    {
        std::vector<uint32_t> stamp(n, 0);
        uint32_t epoch = 0;
        std::vector<Cand> expanded, pool, res;
#pragma omp for schedule(dynamic, 256)
        for (int64_t vi = 0; vi < (int64_t)n; ++vi) {
            const uint32_t v = (uint32_t)vi;
            const float *x = base + (size_t)v * dim;
            if (++epoch == 0) {
                std::fill(stamp.begin(), stamp.end(), 0);
                epoch = 1;
            }
            pool.clear();
            if (L_build > 0) {
                knn_graph_search(base, dim, knn, knn_k, ep, x, L_build, stamp, epoch, expanded);
                pool.insert(pool.end(), expanded.begin(), expanded.end());
            }
            for (uint32_t i = 0; i < knn_k; ++i) {
                const uint32_t nb = knn[(size_t)v * knn_k + i];
                if (nb == 0xFFFFFFFFu) continue;
                pool.push_back({pt_l2(base + (size_t)nb * dim, x, dim), nb});
            }
            std::sort(pool.begin(), pool.end());
            pool.erase(std::unique(pool.begin(), pool.end()), pool.end());
            pool.erase(std::remove_if(pool.begin(), pool.end(), [&](const Cand &c) { return c.id == v; }),
                       pool.end());
            rng_prune(base, dim, pool, R, C, res);
            out[v] = res;
        }
    }
    // 3: reverse edges, collected in increasing source order.
    std::vector<uint32_t> indeg(n + 1, 0);
    for (uint32_t v = 0; v < n; ++v)
        for (const Cand &c : out[v]) ++indeg[c.id + 1];
    for (uint32_t v = 0; v < n; ++v) indeg[v + 1] += indeg[v];
    std::vector<uint32_t> rev(indeg[n]);
    {
        std::vector<uint32_t> fill(indeg.begin(), indeg.end() - 1);
        for (uint32_t v = 0; v < n; ++v)
            for (const Cand &c : out[v]) rev[fill[c.id]++] = v;
    }
    std::vector<std::vector<Cand>> fin(n);
#pragma omp parallel num_threads(nt)
    {
        std::vector<Cand> pool, res;
#pragma omp for schedule(dynamic, 256)
        for (int64_t vi = 0; vi < (int64_t)n; ++vi) {
            const uint32_t v = (uint32_t)vi;
            pool = out[v];
            const float *x = base + (size_t)v * dim;
            for (uint32_t i = reverse_edges ? indeg[v] : indeg[v + 1]; i < indeg[v + 1]; ++i) {
                const uint32_t s = rev[i];
                if (s == v) continue;
                pool.push_back({pt_l2(base + (size_t)s * dim, x, dim), s});
            }
            std::sort(pool.begin(), pool.end());
            pool.erase(std::unique(pool.begin(), pool.end(),
                                   [](const Cand &a, const Cand &b) { return a.id == b.id && a.dist == b.dist; }),
                       pool.end());
            if (pool.size() <= R) {
                fin[v] = pool;
            } else {
                rng_prune(base, dim, pool, R, (uint32_t)pool.size(), res);
                fin[v] = res;
            }
        }
    }
    out.clear();
    out.shrink_to_fit();
    for (uint32_t v = 0; v < n; ++v) {
        deg[v] = (uint32_t)fin[v].size();
        for (uint32_t i = 0; i < deg[v]; ++i) adj[(size_t)v * R + i] = fin[v][i].id;
    }
    fin.clear();
    // 4: DFS stitching from ep with the degree cap (tree_grow,
    // index_nsg.cpp:638-718, without its rand() and its degree overflow):
    // the lowest-id vertex not reached yet gets one edge from a reached
    // vertex with spare degree -- the first such one in its own k-NN list,
    // else the lowest-id one -- and everything it reaches is reached.
    // `spare` holds the reached vertices whose degree is below R.
    std::vector<uint64_t> reached = new_bits(n), spare = new_bits(n);
    std::vector<uint32_t> stack;
    auto mark = [&](uint32_t w) {
        bset(reached, w);
        if (deg[w] < R) bset(spare, w);
    };
    auto dfs = [&](uint32_t root) {
        stack.clear();
        if (bget(reached, root)) return;
        mark(root);
        stack.push_back(root);
        while (!stack.empty()) {
            const uint32_t u = stack.back();
            stack.pop_back();
            for (uint32_t i = 0; i < deg[u]; ++i) {
                const uint32_t w = adj[(size_t)u * R + i];
                if (!bget(reached, w)) {
                    mark(w);
                    stack.push_back(w);
                }
            }
        }
    };
    dfs(ep);
    for (uint32_t u = 0; u < n; ++u) {
        if (bget(reached, u)) continue;
        uint32_t root = 0xFFFFFFFFu;
        for (uint32_t i = 0; i < knn_k; ++i) {
            const uint32_t c = knn[(size_t)u * knn_k + i];
            if (c != 0xFFFFFFFFu && bget(spare, c)) {
                root = c;
                break;
            }
        }
        if (root == 0xFFFFFFFFu) {
            for (size_t wd = 0; wd < spare.size(); ++wd)
                if (spare[wd]) {
                    root = (uint32_t)(wd * 64 + (size_t)__builtin_ctzll(spare[wd]));
                    break;
                }
            if (root == 0xFFFFFFFFu) return fail(PT_ERR_RANGE, "no reached vertex has spare degree");
        }
        adj[(size_t)root * R + deg[root]++] = u;
        if (deg[root] == R) bclr(spare, root);
        dfs(u);
    }
    return PT_OK;
}

extern "C" int pt_reachable(const uint32_t *adj, const uint32_t *deg, uint32_t n,
                            uint32_t stride, uint32_t ep, uint32_t *reached)
{
    if (ep >= n) return fail(PT_ERR_ARG, "ep out of range");
    std::vector<uint64_t> seen = new_bits(n);
    std::vector<uint32_t> frontier{ep}, next;
    bset(seen, ep);
    uint32_t cnt = 1;
    while (!frontier.empty()) {
        next.clear();
        for (uint32_t u : frontier)
            for (uint32_t i = 0; i < deg[u]; ++i) {
                const uint32_t w = adj[(size_t)u * stride + i];
                if (w < n && !bget(seen, w)) {
                    bset(seen, w);
                    next.push_back(w);
                    ++cnt;
                }
            }
        frontier.swap(next);
    }
    *reached = cnt;
    return PT_OK;
}
