/*
 * Runs every oracle function once on small graphs with known answers, for a
 * build with AddressSanitizer and UndefinedBehaviorSanitizer (make -C oracle
 * san).  Test infrastructure only, like the rest of oracle/.  Exit status 0
 * when every check holds; a sanitizer report aborts the run.
 *
 * Graphs: complete graphs K_n (C(n,3) triangles, W = C(n,3) probes), a
 * complete bipartite graph (0), a wheel with n rim vertices (n), a degree-
 * ordered random graph (hashed = merge = brute force), the 1 x 1 tile job of
 * that graph and a batch of two graphs with an empty one between them.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

uint64_t oracle_count_hash(const int64_t *beg, const int32_t *adj, int64_t V, int nthreads,
                           int64_t *max_bucket);
uint64_t oracle_count_merge(const int64_t *beg, const int32_t *adj, int64_t V, int nthreads);
uint64_t oracle_count_brute(int64_t n, const int32_t *eu, const int32_t *ev, int64_t m);
uint64_t oracle_count_tile_job(const int64_t *ht_beg, const int32_t *ht_adj,
                               const int64_t *walk_beg, const int32_t *walk_adj,
                               const int64_t *probe_beg, const int32_t *probe_adj, int64_t rows,
                               int64_t probe_rows, int nthreads);
uint64_t oracle_probes(const int64_t *beg, const int32_t *adj, int64_t V, int64_t *per_vertex);
uint64_t oracle_probes_short(const int64_t *beg, const int32_t *adj, int64_t V,
                             int64_t *per_vertex);
void oracle_features(const int64_t *beg, const int32_t *adj, int64_t V, double *out);
void oracle_features_short(const int64_t *beg, const int32_t *adj, int64_t V, double *out);
void oracle_count_hash_batch(const int64_t *beg, const int32_t *adj, int64_t V,
                             const int64_t *vbase, int64_t ngraphs, int nthreads, uint64_t *out);
uint64_t oracle_count_hash_sample(const int64_t *beg, const int32_t *adj, int64_t V,
                                  int64_t first, int64_t stride, int nthreads,
                                  uint64_t *probes_out);

static int failures = 0;
#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                 \
            fprintf(stderr, "\n");                        \
            failures++;                                   \
        }                                                 \
    } while (0)

/* Oriented CSR of an undirected edge list: u -> v when u < v (acyclic). */
typedef struct {
    int64_t V, E;
    int64_t *beg;
    int32_t *adj;
} csr;

static csr orient(int64_t V, const int32_t *eu, const int32_t *ev, int64_t m)
{
    csr g;
    g.V = V;
    g.E = m;
    g.beg = calloc((size_t)V + 1, sizeof(int64_t));
    g.adj = malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    for (int64_t i = 0; i < m; i++) {
        int32_t a = eu[i] < ev[i] ? eu[i] : ev[i];
        g.beg[a + 1]++;
    }
    for (int64_t u = 0; u < V; u++)
        g.beg[u + 1] += g.beg[u];
    int64_t *fill = malloc(sizeof(int64_t) * (size_t)(V > 0 ? V : 1));
    memcpy(fill, g.beg, sizeof(int64_t) * (size_t)V);
    for (int64_t i = 0; i < m; i++) {
        int32_t a = eu[i] < ev[i] ? eu[i] : ev[i], b = eu[i] < ev[i] ? ev[i] : eu[i];
        g.adj[fill[a]++] = b;
    }
    free(fill);
    return g;
}

static void release(csr *g)
{
    free(g->beg);
    free(g->adj);
}

static void check_graph(const char *name, int64_t V, const int32_t *eu, const int32_t *ev,
                        int64_t m, uint64_t want, int64_t want_probes)
{
    csr g = orient(V, eu, ev, m);
    int64_t mb = 0;
    const uint64_t h = oracle_count_hash(g.beg, g.adj, g.V, 2, &mb);
    const uint64_t s = oracle_count_merge(g.beg, g.adj, g.V, 2);
    const uint64_t b = oracle_count_brute(V, eu, ev, m);
    CHECK(h == want && s == want && b == want, "%s: hash %llu merge %llu brute %llu want %llu",
          name, (unsigned long long)h, (unsigned long long)s, (unsigned long long)b,
          (unsigned long long)want);
    int64_t *pv = calloc((size_t)V + 1, sizeof(int64_t));
    const uint64_t w = oracle_probes(g.beg, g.adj, g.V, pv);
    uint64_t wsum = 0;
    for (int64_t u = 0; u < V; u++)
        wsum += (uint64_t)pv[u];
    CHECK(w == wsum, "%s: probes %llu != per-vertex sum %llu", name, (unsigned long long)w,
          (unsigned long long)wsum);
    if (want_probes >= 0)
        CHECK(w == (uint64_t)want_probes, "%s: probes %llu want %lld", name,
              (unsigned long long)w, (long long)want_probes);
    const uint64_t ws = oracle_probes_short(g.beg, g.adj, g.V, pv);
    CHECK(ws <= w, "%s: short-side probes %llu > reference %llu", name, (unsigned long long)ws,
          (unsigned long long)w);
    double *f = malloc(sizeof(double) * 3 * (size_t)(V + 1));
    oracle_features(g.beg, g.adj, g.V, f);
    oracle_features_short(g.beg, g.adj, g.V, f);
    free(f);
    uint64_t sp = 0;
    const uint64_t c1 = oracle_count_hash_sample(g.beg, g.adj, g.V, 0, 1, 2, &sp);
    CHECK(c1 == want && sp == w, "%s: stride-1 sample %llu / %llu probes", name,
          (unsigned long long)c1, (unsigned long long)sp);
    uint64_t part = 0;
    for (int64_t first = 0; first < 3; first++)
        part += oracle_count_hash_sample(g.beg, g.adj, g.V, first, 3, 2, NULL);
    CHECK(part == want, "%s: stride-3 samples sum to %llu", name, (unsigned long long)part);
    /* the whole graph as one tile job: hash, walk and probe tiles are the graph */
    const uint64_t t = oracle_count_tile_job(g.beg, g.adj, g.beg, g.adj, g.beg, g.adj, g.V, g.V, 2);
    CHECK(t == want, "%s: 1x1 tile job %llu", name, (unsigned long long)t);
    free(pv);
    release(&g);
}

int main(void)
{
    /* K_n */
    for (int n = 3; n <= 40; n += 37) {
        int64_t m = (int64_t)n * (n - 1) / 2, k = 0;
        int32_t *eu = malloc(sizeof(int32_t) * m), *ev = malloc(sizeof(int32_t) * m);
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) {
                eu[k] = i;
                ev[k++] = j;
            }
        const uint64_t c3 = (uint64_t)n * (n - 1) * (n - 2) / 6;
        check_graph("K_n", n, eu, ev, m, c3, (int64_t)c3);
        free(eu);
        free(ev);
    }
    /* K_{20,30} */
    {
        int64_t m = 600, k = 0;
        int32_t *eu = malloc(sizeof(int32_t) * m), *ev = malloc(sizeof(int32_t) * m);
        for (int i = 0; i < 20; i++)
            for (int j = 0; j < 30; j++) {
                eu[k] = i;
                ev[k++] = 20 + j;
            }
        check_graph("K_20,30", 50, eu, ev, m, 0, -1);
        free(eu);
        free(ev);
    }
    /* wheel: hub 0, rim 1..n */
    {
        const int n = 25;
        int32_t eu[50], ev[50];
        for (int i = 0; i < n; i++) {
            eu[i] = 0;
            ev[i] = 1 + i;
            eu[n + i] = 1 + i;
            ev[n + i] = 1 + (i + 1) % n;
        }
        check_graph("wheel", n + 1, eu, ev, 2 * n, (uint64_t)n, -1);
    }
    /* random graph, no duplicate edges: the three counters must agree */
    {
        const int n = 300;
        const int64_t cap = 6000;
        int32_t *eu = malloc(sizeof(int32_t) * cap), *ev = malloc(sizeof(int32_t) * cap);
        unsigned char *seen = calloc((size_t)n * n, 1);
        uint64_t x = 88172645463325252ull;
        int64_t m = 0;
        while (m < cap) {
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            int a = (int)(x % n), b = (int)((x >> 20) % n);
            if (a == b || seen[a * n + b])
                continue;
            seen[a * n + b] = seen[b * n + a] = 1;
            eu[m] = a;
            ev[m++] = b;
        }
        csr g = orient(n, eu, ev, m);
        const uint64_t want = oracle_count_brute(n, eu, ev, m);
        release(&g);
        check_graph("random", n, eu, ev, m, want, -1);
        /* a batch: the random graph, an empty graph, K_3 */
        csr r = orient(n, eu, ev, m);
        const int64_t V = n + 5 + 3, E = m + 3;
        int64_t *beg = calloc((size_t)V + 1, sizeof(int64_t));
        int32_t *adj = malloc(sizeof(int32_t) * (size_t)E);
        memcpy(beg, r.beg, sizeof(int64_t) * (n + 1));
        memcpy(adj, r.adj, sizeof(int32_t) * (size_t)m);
        for (int64_t u = n; u <= n + 5; u++)
            beg[u] = m;
        const int32_t k3[3] = {n + 6, n + 7, n + 7}; /* (n+5)->(n+6),(n+7); (n+6)->(n+7) */
        beg[n + 6] = m + 2;
        beg[n + 7] = m + 3;
        beg[n + 8] = m + 3;
        memcpy(adj + m, k3, sizeof(k3));
        const int64_t vbase[4] = {0, n, n + 5, n + 8};
        uint64_t out[3] = {9, 9, 9};
        oracle_count_hash_batch(beg, adj, V, vbase, 3, 2, out);
        CHECK(out[0] == want && out[1] == 0 && out[2] == 1, "batch: %llu %llu %llu",
              (unsigned long long)out[0], (unsigned long long)out[1], (unsigned long long)out[2]);
        free(beg);
        free(adj);
        release(&r);
        free(seen);
        free(eu);
        free(ev);
    }
    if (failures)
        fprintf(stderr, "%d check(s) failed\n", failures);
    else
        printf("oracle sanitizer checks: ok\n");
    return failures ? 1 : 0;
}
