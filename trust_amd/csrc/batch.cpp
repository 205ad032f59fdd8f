// Batches of many small graphs: seeded generation of a mixed Erdos-Renyi /
// Barabasi-Albert batch, each graph put through the reference's input
// contract (preprocess_edges), concatenated into one CSR whose vertex ids
// are global: graph g owns vertices [vbase[g], vbase[g+1]).  The counting
// kernel then runs once over the concatenation and attributes triangles to
// graphs by vertex range.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

namespace {

inline uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline double unit(uint64_t x) { return (double)(mix64(x) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

using namespace trust;

namespace trust {

int batch_validate(const int64_t *beg, const int32_t *adj, int64_t V, const int64_t *vbase,
                   int64_t ngraphs)
{
    if (ngraphs < 1 || !vbase)
        return fail(TRUST_ERR_ARG, "a batch needs ngraphs >= 1 and vbase");
    if (vbase[0] != 0 || vbase[ngraphs] != V)
        return fail(TRUST_ERR_FORMAT, "vbase must run from 0 to V = %lld", (long long)V);
    for (int64_t g = 0; g < ngraphs; g++)
        if (vbase[g + 1] < vbase[g])
            return fail(TRUST_ERR_FORMAT, "vbase decreases at graph %lld", (long long)g);
    int64_t bad = -1;
#pragma omp parallel for schedule(dynamic, 16) reduction(max : bad)
    for (int64_t g = 0; g < ngraphs; g++) {
        const int64_t lo = vbase[g], hi = vbase[g + 1];
        for (int64_t e = beg[lo]; e < beg[hi]; e++)
            if (adj[e] < lo || adj[e] >= hi) {
                bad = std::max(bad, g);
                break;
            }
    }
    if (bad >= 0)
        return fail(TRUST_ERR_FORMAT, "graph %lld of the batch has an edge leaving it",
                    (long long)bad);
    return TRUST_OK;
}

}  // namespace trust

extern "C" int trust_batch_shape(int64_t ngraphs, uint64_t seed, int64_t nmin, int64_t nmax,
                                 double dmin, double dmax, int64_t *n_out, double *d_out,
                                 int32_t *kind_out)
{
    if (ngraphs < 1 || nmin < 2 || nmax < nmin || nmax > INT32_MAX / 2 || dmin <= 0 ||
        dmax < dmin || !n_out || !d_out || !kind_out)
        return fail(TRUST_ERR_ARG, "bad batch shape arguments");
    const double lo = std::log((double)nmin), hi = std::log((double)nmax);
    for (int64_t g = 0; g < ngraphs; g++) {
        const uint64_t s = mix64(seed * 0x100000001B3ull + (uint64_t)g);
        n_out[g] = (int64_t)std::llround(std::exp(lo + (hi - lo) * unit(s)));
        n_out[g] = std::min<int64_t>(std::max<int64_t>(n_out[g], nmin), nmax);
        d_out[g] = dmin + (dmax - dmin) * unit(s ^ 0xA5A5A5A5A5A5A5A5ull);
        kind_out[g] = (int32_t)(g & 1);  // 0: Erdos-Renyi, 1: Barabasi-Albert
    }
    return TRUST_OK;
}

extern "C" int trust_batch_generate(int64_t ngraphs, uint64_t seed, int64_t nmin, int64_t nmax,
                                    double dmin, double dmax, trust_csr *out, int64_t *vbase)
{
    if (!out || !vbase)
        return fail(TRUST_ERR_ARG, "null output");
    memset(out, 0, sizeof(*out));
    std::vector<int64_t> n((size_t)std::max<int64_t>(ngraphs, 0));
    std::vector<double> d(n.size());
    std::vector<int32_t> kind(n.size());
    int rc = trust_batch_shape(ngraphs, seed, nmin, nmax, dmin, dmax, n.data(), d.data(),
                               kind.data());
    if (rc)
        return rc;
    std::vector<HostCsr> parts((size_t)ngraphs);
    std::vector<int> err((size_t)ngraphs, 0);
    const int saved_nested = omp_get_max_active_levels();
    omp_set_max_active_levels(1);  // one thread per graph; no nested teams
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t g = 0; g < ngraphs; g++) {
        const uint64_t gs = mix64(seed ^ mix64((uint64_t)g + 0x51ED2701ull));
        std::vector<int32_t> a, b;
        int r;
        if (kind[g] == 0) {
            const int64_t m = std::max<int64_t>(1, std::llround((double)n[g] * d[g] / 2.0));
            a.resize((size_t)m);
            b.resize((size_t)m);
            r = trust_gen_erdos_renyi(n[g], m, gs, a.data(), b.data());
        } else {
            const int k = std::max(1, (int)std::lround(d[g] / 2.0));
            a.resize((size_t)(n[g] * k));
            b.resize((size_t)(n[g] * k));
            int64_t m = 0;
            r = trust_gen_barabasi_albert(n[g], k, gs, a.data(), b.data(), &m);
            a.resize((size_t)m);
            b.resize((size_t)m);
        }
        if (!r)
            r = preprocess_edges(a.data(), b.data(), (int64_t)a.size(), parts[g]);
        err[g] = r;
    }
    omp_set_max_active_levels(saved_nested);
    for (int64_t g = 0; g < ngraphs; g++)
        if (err[g])
            return fail(err[g], "graph %lld of the batch failed to generate",
                        (long long)g);
    vbase[0] = 0;
    std::vector<int64_t> ebase((size_t)ngraphs + 1, 0);
    for (int64_t g = 0; g < ngraphs; g++) {
        vbase[g + 1] = vbase[g] + parts[g].V;
        ebase[g + 1] = ebase[g] + parts[g].E;
    }
    const int64_t V = vbase[ngraphs], E = ebase[ngraphs];
    if (V >= (int64_t)INT32_MAX)
        return fail(TRUST_ERR_OVERFLOW, "batch has %lld vertices (int32 ids)", (long long)V);
    out->beg = (int64_t *)malloc(sizeof(int64_t) * (size_t)(V + 1));
    out->adj = (int32_t *)malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(E, 1));
    if (!out->beg || !out->adj) {
        free(out->beg);
        free(out->adj);
        out->beg = nullptr;
        out->adj = nullptr;
        return fail(TRUST_ERR_OOM, "cannot allocate a batch of %lld edges", (long long)E);
    }
    out->vertices = V;
    out->edges = E;
    out->beg[V] = E;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t g = 0; g < ngraphs; g++) {
        const HostCsr &p = parts[g];
        const int64_t vb = vbase[g], eb = ebase[g];
        for (int64_t u = 0; u < p.V; u++)
            out->beg[vb + u] = eb + p.beg[u];
        for (int64_t e = 0; e < p.E; e++)
            out->adj[eb + e] = (int32_t)(vb + p.adj[e]);
        parts[g] = HostCsr();  // release as we go
    }
    return TRUST_OK;
}
