/*
 * tc_oracle.c -- CPU oracle for TRUST's vertex-centric hashed triangle count.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library
 * (trust_amd/) never links, loads or calls anything here.
 *
 * Every function restates one piece of the reference algorithm, with the
 * reference file:line it follows (paths relative to the reference repo root):
 *
 *   oracle_count_hash   Without-graph-partition/TC.cu:262-305  (CTA hash build)
 *                       Without-graph-partition/TC.cu:330-359  (CTA probe walk)
 *                       Without-graph-partition/TC.cu:435-545  (warp tier)
 *                       Without-graph-partition/TC.cu:70-110   (linear_search)
 *                       Without-graph-partition/TC.cu:586-599  (uint64 reduce)
 *       Bucket = id & (nbuckets-1), nbuckets = 1024 for d+(u) > 100 and 32
 *       otherwise; vertices with d+(u) < 2 are skipped (TC.cu:440-441).
 *       Unlike the reference, a bucket never drops entries (the reference
 *       stops at 100 per bucket, TC.cu:293-301); the largest bucket seen is
 *       reported so a caller can tell whether the reference would have
 *       dropped anything on the same input.
 *   oracle_count_merge  Without-graph-partition/graph.hpp:121-149 (dead CPU
 *       sort-merge counter): an independent second method.  Each adjacency
 *       list is sorted, then |N+(u) & N+(v)| is a merge per oriented edge.
 *   oracle_count_brute  no reference counterpart: O(n^3/64) bitset count of
 *       all i<j<k triples on the raw undirected edge list, independent of
 *       orientation and preprocessing.
 *   oracle_count_tile_job  With-graph-partition/TC.cu:630-973 and
 *       With-graph-partition/TC.cu:985-1012: one (i,j,k) job of the 2D hash
 *       partition.  Rows u are taken from the walk tile's map order; N+(u)
 *       is hashed from the HT tile, v walks the walk tile and w probes the
 *       neighbour tile.  The map/division files are not needed: every row of
 *       the walk tile with at least one entry is processed, which is what the
 *       reference's map + division select (Preprocess/partition.cpp:131-165).
 *   oracle_probes       W = sum over processed u of sum_{v in N+(u)} d+(v):
 *       the number of linear_search calls the reference makes
 *       (Without-graph-partition/TC.cu:354,534).
 *   oracle_probes_short  the probes of the product's short-side walk (no
 *       reference counterpart; the count is the reference's): each oriented
 *       edge u -> v with d+(u) >= 2 costs min(d+(u), d+(v)) probes, charged
 *       to u when d+(v) <= d+(u) and to v otherwise.
 *   oracle_features_short  the same features over the short-side walk's
 *       rows (no reference counterpart): sum and max of min(d+(u), d+(v)),
 *       charged as oracle_probes_short charges them.
 *   oracle_features     the learned scheduler's per-vertex features
 *       d+(u), sum_{v in N+(u)} d+(v), max_{v in N+(u)} d+(v).
 *   oracle_count_hash_batch  oracle_count_hash per graph of a batch (graph
 *       g owns vertices [vbase[g], vbase[g+1])), parallel over graphs: the
 *       reference counts one graph per run (Without-graph-partition/
 *       mpi_main.cpp:72-96), so the per-graph count is that run's count.
 *
 * Pinning (tests/test_oracle.py): closed forms (K_n, K_{a,b}, C_n, wheels,
 * friendship graphs, complete tripartite), brute force on small random
 * graphs, networkx triangle counts committed as fixtures, agreement of the
 * three independent methods, and W against a direct recount.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_CTA_THRESHOLD 100 /* USE_CTA, Without-graph-partition/TC.cu:19 */
#define ORACLE_MIN_DEGREE 2      /* USE_WARP, Without-graph-partition/TC.cu:20 */
#define ORACLE_CTA_BUCKETS 1024  /* block_bucketnum, TC.cu:24 */
#define ORACLE_WARP_BUCKETS 32   /* warp_bucketnum, TC.cu:25 */

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Bucketed membership table for one apex u: counting-sort N+(u) into
 * nb buckets by id & (nb-1), then a lookup scans one bucket. */
typedef struct {
    int32_t *start; /* nb+1 */
    int32_t *items; /* d */
    int32_t *fill;  /* nb */
    int cap_items;
} bucket_table;

static void bt_build(bucket_table *t, const int32_t *list, int64_t d, int nb)
{
    int mask = nb - 1;
    memset(t->fill, 0, sizeof(int32_t) * nb);
    for (int64_t i = 0; i < d; i++)
        t->fill[list[i] & mask]++;
    t->start[0] = 0;
    for (int b = 0; b < nb; b++)
        t->start[b + 1] = t->start[b] + t->fill[b];
    memset(t->fill, 0, sizeof(int32_t) * nb);
    for (int64_t i = 0; i < d; i++) {
        int b = list[i] & mask;
        t->items[t->start[b] + t->fill[b]++] = list[i];
    }
}

static int bt_contains(const bucket_table *t, int32_t key, int nb)
{
    int b = key & (nb - 1);
    for (int32_t i = t->start[b]; i < t->start[b + 1]; i++)
        if (t->items[i] == key)
            return 1;
    return 0;
}

/*
 * Hashed count over one oriented CSR.  Returns the uint64 triangle count;
 * *max_bucket (nullable) receives the largest bucket occupancy seen.
 */
uint64_t oracle_count_hash(const int64_t *beg, const int32_t *adj, int64_t V,
                           int nthreads, int64_t *max_bucket)
{
    uint64_t total = 0;
    int64_t maxb = 0;
    int64_t maxd = 0;
    for (int64_t u = 0; u < V; u++) {
        int64_t d = beg[u + 1] - beg[u];
        if (d > maxd)
            maxd = d;
    }
    set_threads(nthreads);
#pragma omp parallel reduction(+ : total) reduction(max : maxb)
    {
        bucket_table t;
        t.start = (int32_t *)malloc(sizeof(int32_t) * (ORACLE_CTA_BUCKETS + 1));
        t.fill = (int32_t *)malloc(sizeof(int32_t) * ORACLE_CTA_BUCKETS);
        t.items = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxd + 1));
#pragma omp for schedule(dynamic, 64)
        for (int64_t u = 0; u < V; u++) {
            int64_t d = beg[u + 1] - beg[u];
            if (d < ORACLE_MIN_DEGREE)
                continue;
            int nb = d > ORACLE_CTA_THRESHOLD ? ORACLE_CTA_BUCKETS : ORACLE_WARP_BUCKETS;
            bt_build(&t, adj + beg[u], d, nb);
            for (int b = 0; b < nb; b++) {
                int64_t occ = t.start[b + 1] - t.start[b];
                if (occ > maxb)
                    maxb = occ;
            }
            uint64_t c = 0;
            for (int64_t e = beg[u]; e < beg[u + 1]; e++) {
                int32_t v = adj[e];
                for (int64_t f = beg[v]; f < beg[v + 1]; f++)
                    c += (uint64_t)bt_contains(&t, adj[f], nb);
            }
            total += c;
        }
        free(t.start);
        free(t.fill);
        free(t.items);
    }
    if (max_bucket)
        *max_bucket = maxb;
    return total;
}

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* Sort-merge count: an independent method (no hashing, no buckets). */
uint64_t oracle_count_merge(const int64_t *beg, const int32_t *adj, int64_t V, int nthreads)
{
    int64_t E = beg[V];
    int32_t *s = (int32_t *)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
    memcpy(s, adj, sizeof(int32_t) * (size_t)E);
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t u = 0; u < V; u++)
        qsort(s + beg[u], (size_t)(beg[u + 1] - beg[u]), sizeof(int32_t), cmp_i32);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
    for (int64_t u = 0; u < V; u++) {
        for (int64_t e = beg[u]; e < beg[u + 1]; e++) {
            int32_t v = s[e];
            int64_t i = beg[u], ie = beg[u + 1];
            int64_t j = beg[v], je = beg[v + 1];
            while (i < ie && j < je) {
                if (s[i] < s[j])
                    i++;
                else if (s[i] > s[j])
                    j++;
                else {
                    total++;
                    i++;
                    j++;
                }
            }
        }
    }
    free(s);
    return total;
}

/*
 * Brute force on an undirected edge list over vertex ids [0, n): counts
 * unordered triples {i<j<k} with all three edges present.  Self-loops and
 * duplicate edges are ignored.  Meant for n up to a few thousand.
 */
uint64_t oracle_count_brute(int64_t n, const int32_t *eu, const int32_t *ev, int64_t m)
{
    int64_t words = (n + 63) / 64;
    uint64_t *bits = (uint64_t *)calloc((size_t)(n * words > 0 ? n * words : 1), sizeof(uint64_t));
    for (int64_t e = 0; e < m; e++) {
        int64_t a = eu[e], b = ev[e];
        if (a == b || a < 0 || b < 0 || a >= n || b >= n)
            continue;
        bits[a * words + b / 64] |= 1ull << (b % 64);
        bits[b * words + a / 64] |= 1ull << (a % 64);
    }
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : total)
    for (int64_t i = 0; i < n; i++) {
        const uint64_t *ri = bits + i * words;
        for (int64_t j = i + 1; j < n; j++) {
            if (!((ri[j / 64] >> (j % 64)) & 1))
                continue;
            const uint64_t *rj = bits + j * words;
            /* count k > j adjacent to both */
            int64_t w0 = (j + 1) / 64;
            for (int64_t w = w0; w < words; w++) {
                uint64_t x = ri[w] & rj[w];
                if (w == w0) {
                    int sh = (int)((j + 1) % 64);
                    x = sh ? (x >> sh) << sh : x;
                }
                total += (uint64_t)__builtin_popcountll(x);
            }
        }
    }
    free(bits);
    return total;
}

/*
 * One (i,j,k) job of the 2D partition (With-graph-partition/TC.cu:630-973).
 * ht:    tile (i,j) -- rows u (local), hashed entries w (class j, local ids)
 * walk:  tile (i,k) -- rows u (local), entries v (class k, local ids)
 * probe: tile (k,j) -- rows v (local), entries w (class j, local ids)
 * rows:  number of rows shared by ht and walk (they have the same row set).
 * Rows whose walk degree is 0 are skipped (Preprocess/partition.cpp:159
 * truncates the map at the first degree-0 row).  The hash table uses
 * blockDim-many buckets in the reference (With-graph-partition/TC.cu:667);
 * membership does not depend on the bucket count, so 1024 is used.
 */
uint64_t oracle_count_tile_job(const int64_t *ht_beg, const int32_t *ht_adj,
                               const int64_t *walk_beg, const int32_t *walk_adj,
                               const int64_t *probe_beg, const int32_t *probe_adj,
                               int64_t rows, int64_t probe_rows, int nthreads)
{
    int64_t maxd = 0;
    for (int64_t u = 0; u < rows; u++) {
        int64_t d = ht_beg[u + 1] - ht_beg[u];
        if (d > maxd)
            maxd = d;
    }
    uint64_t total = 0;
    set_threads(nthreads);
#pragma omp parallel reduction(+ : total)
    {
        bucket_table t;
        t.start = (int32_t *)malloc(sizeof(int32_t) * (ORACLE_CTA_BUCKETS + 1));
        t.fill = (int32_t *)malloc(sizeof(int32_t) * ORACLE_CTA_BUCKETS);
        t.items = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxd + 1));
#pragma omp for schedule(dynamic, 64)
        for (int64_t u = 0; u < rows; u++) {
            int64_t dw = walk_beg[u + 1] - walk_beg[u];
            int64_t dh = ht_beg[u + 1] - ht_beg[u];
            if (dw < 1 || dh < 1)
                continue;
            bt_build(&t, ht_adj + ht_beg[u], dh, ORACLE_CTA_BUCKETS);
            uint64_t c = 0;
            for (int64_t e = walk_beg[u]; e < walk_beg[u + 1]; e++) {
                int32_t v = walk_adj[e];
                if (v < 0 || v >= probe_rows)
                    continue;
                for (int64_t f = probe_beg[v]; f < probe_beg[v + 1]; f++)
                    c += (uint64_t)bt_contains(&t, probe_adj[f], ORACLE_CTA_BUCKETS);
            }
            total += c;
        }
        free(t.start);
        free(t.fill);
        free(t.items);
    }
    return total;
}

/*
 * W: number of probes (linear_search calls) over vertices with d+ >= 2.
 * per_vertex (nullable, length V) receives W_u (0 for skipped vertices).
 */
uint64_t oracle_probes(const int64_t *beg, const int32_t *adj, int64_t V, int64_t *per_vertex)
{
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : total)
    for (int64_t u = 0; u < V; u++) {
        int64_t d = beg[u + 1] - beg[u];
        int64_t w = 0;
        if (d >= ORACLE_MIN_DEGREE)
            for (int64_t e = beg[u]; e < beg[u + 1]; e++)
                w += beg[adj[e] + 1] - beg[adj[e]];
        if (per_vertex)
            per_vertex[u] = w;
        total += (uint64_t)w;
    }
    return total;
}

uint64_t oracle_probes_short(const int64_t *beg, const int32_t *adj, int64_t V,
                             int64_t *per_vertex)
{
    if (per_vertex)
        memset(per_vertex, 0, sizeof(int64_t) * (size_t)V);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : total)
    for (int64_t u = 0; u < V; u++) {
        const int64_t du = beg[u + 1] - beg[u];
        if (du < ORACLE_MIN_DEGREE)
            continue;
        for (int64_t e = beg[u]; e < beg[u + 1]; e++) {
            const int64_t v = adj[e], dv = beg[v + 1] - beg[v];
            const int64_t w = dv <= du ? dv : du;
            total += (uint64_t)w;
            if (per_vertex) {
                int64_t *slot = per_vertex + (dv <= du ? u : v);
#pragma omp atomic
                *slot += w;
            }
        }
    }
    return total;
}

/* Features for the cost model: out[3u+0]=d+(u), [3u+1]=sum d+(v), [3u+2]=max d+(v). */
void oracle_features(const int64_t *beg, const int32_t *adj, int64_t V, double *out)
{
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t u = 0; u < V; u++) {
        int64_t d = beg[u + 1] - beg[u];
        double s = 0.0, mx = 0.0;
        for (int64_t e = beg[u]; e < beg[u + 1]; e++) {
            double dv = (double)(beg[adj[e] + 1] - beg[adj[e]]);
            s += dv;
            if (dv > mx)
                mx = dv;
        }
        out[3 * u + 0] = (double)d;
        out[3 * u + 1] = s;
        out[3 * u + 2] = mx;
    }
}

void oracle_features_short(const int64_t *beg, const int32_t *adj, int64_t V, double *out)
{
    for (int64_t u = 0; u < V; u++) {
        out[3 * u + 0] = (double)(beg[u + 1] - beg[u]);
        out[3 * u + 1] = 0.0;
        out[3 * u + 2] = 0.0;
    }
    for (int64_t u = 0; u < V; u++) {
        const int64_t du = beg[u + 1] - beg[u];
        if (du < ORACLE_MIN_DEGREE)
            continue;
        for (int64_t e = beg[u]; e < beg[u + 1]; e++) {
            const int64_t v = adj[e], dv = beg[v + 1] - beg[v];
            const int64_t owner = dv <= du ? u : v;
            const double len = (double)(dv <= du ? dv : du);
            out[3 * owner + 1] += len;
            if (len > out[3 * owner + 2])
                out[3 * owner + 2] = len;
        }
    }
}

/*
 * Per-graph hashed counts of a batch: out[g] = oracle_count_hash of graph g.
 * Graphs run in parallel, the vertices of one graph in sequence.
 */
void oracle_count_hash_batch(const int64_t *beg, const int32_t *adj, int64_t V,
                             const int64_t *vbase, int64_t ngraphs, int nthreads, uint64_t *out)
{
    int64_t maxd = 0;
    for (int64_t u = 0; u < V; u++) {
        int64_t d = beg[u + 1] - beg[u];
        if (d > maxd)
            maxd = d;
    }
    set_threads(nthreads);
#pragma omp parallel
    {
        bucket_table t;
        t.start = (int32_t *)malloc(sizeof(int32_t) * (ORACLE_CTA_BUCKETS + 1));
        t.fill = (int32_t *)malloc(sizeof(int32_t) * ORACLE_CTA_BUCKETS);
        t.items = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxd + 1));
#pragma omp for schedule(dynamic, 1)
        for (int64_t g = 0; g < ngraphs; g++) {
            uint64_t c = 0;
            for (int64_t u = vbase[g]; u < vbase[g + 1]; u++) {
                int64_t d = beg[u + 1] - beg[u];
                if (d < ORACLE_MIN_DEGREE)
                    continue;
                int nb = d > ORACLE_CTA_THRESHOLD ? ORACLE_CTA_BUCKETS : ORACLE_WARP_BUCKETS;
                bt_build(&t, adj + beg[u], d, nb);
                for (int64_t e = beg[u]; e < beg[u + 1]; e++) {
                    int32_t v = adj[e];
                    for (int64_t f = beg[v]; f < beg[v + 1]; f++)
                        c += (uint64_t)bt_contains(&t, adj[f], nb);
                }
            }
            out[g] = c;
        }
        free(t.start);
        free(t.fill);
        free(t.items);
    }
}

/*
 * Sampled hashed count for the CPU baseline: runs oracle_count_hash's inner
 * loop over the vertices u = first, first+stride, ... (< V).  Returns the
 * partial count; *probes_out (nullable) receives the probes performed.
 */
uint64_t oracle_count_hash_sample(const int64_t *beg, const int32_t *adj, int64_t V,
                                  int64_t first, int64_t stride, int nthreads,
                                  uint64_t *probes_out)
{
    int64_t maxd = 0;
    for (int64_t u = 0; u < V; u++) {
        int64_t d = beg[u + 1] - beg[u];
        if (d > maxd)
            maxd = d;
    }
    if (stride < 1)
        stride = 1;
    int64_t n = first < V ? (V - first + stride - 1) / stride : 0;
    uint64_t total = 0, probes = 0;
    set_threads(nthreads);
#pragma omp parallel reduction(+ : total, probes)
    {
        bucket_table t;
        t.start = (int32_t *)malloc(sizeof(int32_t) * (ORACLE_CTA_BUCKETS + 1));
        t.fill = (int32_t *)malloc(sizeof(int32_t) * ORACLE_CTA_BUCKETS);
        t.items = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxd + 1));
#pragma omp for schedule(dynamic, 16)
        for (int64_t k = 0; k < n; k++) {
            int64_t u = first + k * stride;
            int64_t d = beg[u + 1] - beg[u];
            if (d < ORACLE_MIN_DEGREE)
                continue;
            int nb = d > ORACLE_CTA_THRESHOLD ? ORACLE_CTA_BUCKETS : ORACLE_WARP_BUCKETS;
            bt_build(&t, adj + beg[u], d, nb);
            uint64_t c = 0, p = 0;
            for (int64_t e = beg[u]; e < beg[u + 1]; e++) {
                int32_t v = adj[e];
                p += (uint64_t)(beg[v + 1] - beg[v]);
                for (int64_t f = beg[v]; f < beg[v + 1]; f++)
                    c += (uint64_t)bt_contains(&t, adj[f], nb);
            }
            total += c;
            probes += p;
        }
        free(t.start);
        free(t.fill);
        free(t.items);
    }
    if (probes_out)
        *probes_out = probes;
    return total;
}
