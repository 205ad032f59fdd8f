// Seeded synthetic edge-list generators for the benchmark workloads: R-MAT,
// Chung-Lu, Erdos-Renyi G(n,m) and Barabasi-Albert.  They stand in for the
// datasets the reference downloads (Dataset/Cit-Patents/get+preprocess.sh),
// which are unavailable offline.  R-MAT, Chung-Lu and ER draw every pair
// from a counter-based stream (splitmix64 of seed and pair index), so the
// output does not depend on the number of threads.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

namespace {

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Small per-pair stream: splitmix64 over a counter.
struct Stream {
    uint64_t s;
    explicit Stream(uint64_t seed, uint64_t idx) : s(splitmix64(seed ^ splitmix64(idx))) {}
    uint64_t next() { return splitmix64(s += 0x632BE59BD9B4E019ull); }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

using namespace trust;

namespace trust {

// Pairs [e0, e0 + count) of the R-MAT stream (pair e depends only on seed
// and e, so any range can be drawn alone).
void gen_rmat_range(int scale, double a, double b, double c, uint64_t seed, int64_t e0,
                    int64_t count, int32_t *src, int32_t *dst)
{
    const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < count; i++) {
        Stream r(seed, (uint64_t)(e0 + i));
        uint32_t u = 0, v = 0;
        for (int l = 0; l < scale; l++) {
            double x = r.uniform();
            uint32_t bu = x >= ab, bv = (x >= a && x < ab) || x >= abc;
            u = (u << 1) | bu;
            v = (v << 1) | bv;
        }
        src[i] = (int32_t)u;
        dst[i] = (int32_t)v;
    }
}

}  // namespace trust

extern "C" int trust_gen_rmat(int scale, int64_t m, double a, double b, double c, uint64_t seed,
                              int32_t *src, int32_t *dst)
{
    if (scale < 1 || scale > 30 || m < 0 || !src || !dst)
        return fail(TRUST_ERR_ARG, "trust_gen_rmat: bad arguments (scale=%d m=%lld)", scale,
                    (long long)m);
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return fail(TRUST_ERR_ARG, "trust_gen_rmat: bad quadrant probabilities");
    gen_rmat_range(scale, a, b, c, seed, 0, m, src, dst);
    return TRUST_OK;
}

extern "C" int trust_gen_chung_lu(int64_t n, int64_t m, double gamma, double max_degree,
                                  uint64_t seed, int32_t *src, int32_t *dst)
{
    if (n < 2 || n > INT32_MAX || m < 0 || gamma <= 1.0 || max_degree <= 0 || !src || !dst)
        return fail(TRUST_ERR_ARG, "trust_gen_chung_lu: bad arguments");
    const double ex = -1.0 / (gamma - 1.0);
    // expected degree of vertex 0 = 2m * w0 / sum(w); find i0 so it equals max_degree
    auto top_degree = [&](double i0) {
        double s = 0.0;
#pragma omp parallel for reduction(+ : s)
        for (int64_t i = 0; i < n; i++)
            s += std::pow((double)i + i0, ex);
        return 2.0 * (double)m * std::pow(i0, ex) / s;
    };
    double lo = 1.0, hi = 1.0;
    if (top_degree(1.0) <= max_degree) {
        hi = 1.0;
    } else {
        while (top_degree(hi) > max_degree && hi < 1e15)
            hi *= 2.0;
        lo = hi / 2.0;
        for (int it = 0; it < 60; it++) {
            double mid = 0.5 * (lo + hi);
            if (top_degree(mid) > max_degree)
                lo = mid;
            else
                hi = mid;
        }
    }
    const double i0 = hi;
    std::vector<double> cdf((size_t)n);
    double acc = 0.0;
    for (int64_t i = 0; i < n; i++) {
        acc += std::pow((double)i + i0, ex);
        cdf[i] = acc;
    }
    const double total = acc;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m; e++) {
        Stream r(seed, (uint64_t)e);
        int32_t ends[2];
        for (int k = 0; k < 2; k++) {
            double x = r.uniform() * total;
            int64_t i = std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin();
            ends[k] = (int32_t)std::min<int64_t>(i, n - 1);
        }
        src[e] = ends[0];
        dst[e] = ends[1];
    }
    return TRUST_OK;
}

extern "C" int trust_gen_erdos_renyi(int64_t n, int64_t m, uint64_t seed, int32_t *src,
                                     int32_t *dst)
{
    if (n < 1 || n > INT32_MAX || m < 0 || !src || !dst)
        return fail(TRUST_ERR_ARG, "trust_gen_erdos_renyi: bad arguments");
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m; e++) {
        Stream r(seed, (uint64_t)e);
        src[e] = (int32_t)(r.next() % (uint64_t)n);
        dst[e] = (int32_t)(r.next() % (uint64_t)n);
    }
    return TRUST_OK;
}

// Batagelj-Brandes linear-time preferential attachment: pair t of new
// vertex v links v to the endpoint stored at a uniformly chosen earlier
// slot, so endpoints are drawn proportionally to degree.
extern "C" int trust_gen_barabasi_albert(int64_t n, int k, uint64_t seed, int32_t *src,
                                         int32_t *dst, int64_t *m_out)
{
    if (n < 2 || n > INT32_MAX || k < 1 || !src || !dst || !m_out)
        return fail(TRUST_ERR_ARG, "trust_gen_barabasi_albert: bad arguments");
    const int64_t pairs = n * (int64_t)k;
    std::vector<int32_t> M((size_t)(2 * pairs));
    Stream r(seed, 0);
    int64_t w = 0;
    for (int64_t v = 0; v < n; v++) {
        for (int t = 0; t < k; t++) {
            int64_t slot = 2 * (v * k + t);
            M[slot] = (int32_t)v;
            int64_t pick = (int64_t)(r.next() % (uint64_t)(slot + 1));
            M[slot + 1] = M[pick];
            src[w] = (int32_t)v;
            dst[w] = M[slot + 1];
            w++;
        }
    }
    *m_out = w;
    return TRUST_OK;
}
