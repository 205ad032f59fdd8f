// Input contract: raw edge list -> degree-oriented, class-ordered CSR.
//
// The output is byte-identical to running the reference's two tools on the
// same list, Preprocess/fromDirectToUndirect.cpp (compact ids, undirected,
// dedup) followed by Preprocess/preprocess_outdegree.cpp (orientation, sort,
// relabel).  Byte identity is the drop-in requirement for this format, so
// the *steps* follow the reference where they decide the output order:
//
//   * ids present in any pair are compacted in ascending order
//     (fromDirectToUndirect.cpp:53-85); pairs become (min,max), deduplicated
//     (:86-109); a self-loop survives there and is counted twice in the
//     degree (preprocess_outdegree.cpp:71-72), then dropped by orientation
//     (:95);
//   * vertices are ordered by ascending degree with std::sort and the
//     comparator of preprocess_outdegree.cpp:21-24 (the order of equal
//     degrees is whatever libstdc++'s introsort produces, and std::sort's
//     permutation depends only on the comparator's answers, so sorting
//     (degree, id) records reproduces it);
//   * each edge points from the lower to the higher position (:79-98), and
//     an out-list is ordered by descending original id (the reference pops
//     an ascending list from the back);
//   * vertices are re-ordered by descending out-degree, again with std::sort
//     and the reference comparator (:25-28,236);
//   * ids are reassigned per class (out-degree > 100, 2..100, < 2) in
//     first-touch order of a scan over the out-lists of vertices with
//     out-degree > 2, then untouched vertices in position order (:100-185).
//
// Everything except the two std::sort calls runs in parallel: the
// first-touch relabel is computed as "rank of the first edge index at which
// a vertex appears", which is exactly the order in which the reference's
// sequential scan hands out ids.
//
// The reference's other preprocessing tools are the same pipeline with two
// switches (PreprocessOpts):
//
//   * self_loop_weight: Preprocess/speedupIO/preprocess_out.cpp:101-126
//     deduplicates every adjacency list, so a self-loop adds 1 to the degree
//     where the 1.mmio path adds 2;
//   * order: otherPreprocess/preprocess.cpp:178 keeps the descending
//     out-degree order without the relabel (ids = positions), and
//     otherPreprocess/preprocess_indegree.cpp:131-156 replaces that sort by
//     a stable split of the ascending-degree order into the three classes.
#include <omp.h>

#include <algorithm>
#include <cstring>

#include "internal.h"

namespace trust {

void radix_sort_u64(uint64_t *keys, uint64_t *tmp, size_t n, int key_bits)
{
    if (n < 2 || key_bits <= 0)
        return;
    const int R = 11;
    const size_t B = (size_t)1 << R;
    const int passes = (key_bits + R - 1) / R;
    const int T = std::max(1, std::min(omp_get_max_threads(), (int)((n + 65535) / 65536)));
    std::vector<size_t> hist((size_t)T * B);
    uint64_t *src = keys, *dst = tmp;
    for (int p = 0; p < passes; p++) {
        const int shift = p * R;
#pragma omp parallel num_threads(T)
        {
            const int nt = omp_get_num_threads();
            const int t = omp_get_thread_num();
            const size_t lo = n * (size_t)t / (size_t)nt, hi = n * (size_t)(t + 1) / (size_t)nt;
            size_t *h = &hist[(size_t)t * B];
            std::fill(h, h + B, (size_t)0);
            for (size_t i = lo; i < hi; i++)
                h[(src[i] >> shift) & (B - 1)]++;
#pragma omp barrier
#pragma omp single
            {
                size_t sum = 0;
                for (size_t d = 0; d < B; d++)
                    for (int t2 = 0; t2 < nt; t2++) {
                        size_t c = hist[(size_t)t2 * B + d];
                        hist[(size_t)t2 * B + d] = sum;
                        sum += c;
                    }
            }
            for (size_t i = lo; i < hi; i++) {
                uint64_t k = src[i];
                dst[h[(k >> shift) & (B - 1)]++] = k;
            }
        }
        std::swap(src, dst);
    }
    if (src != keys)
        memcpy(keys, src, n * sizeof(uint64_t));
}

static int bits_for(uint64_t x)
{
    int b = 0;
    while (b < 64 && (x >> b) != 0)
        b++;
    return b;
}

// Record with the fields the reference's comparators look at.
struct SizeRec {
    int64_t size;
    int32_t id;
};

int preprocess_edges(const int32_t *src, const int32_t *dst, int64_t m, HostCsr &out,
                     const PreprocessOpts &o)
{
    if (m < 0 || (m > 0 && (!src || !dst)))
        return fail(TRUST_ERR_ARG, "bad edge list (m=%lld)", (long long)m);
    if (o.order < TRUST_ORDER_OUTDEGREE || o.order > TRUST_ORDER_INDEGREE)
        return fail(TRUST_ERR_ARG, "unknown row order %d", o.order);
    if (o.self_loop_weight < 1 || o.self_loop_weight > 2)
        return fail(TRUST_ERR_ARG, "self-loop weight must be 1 or 2");
    const int64_t n_fixed = o.n_fixed;
    int64_t maxid = -1, neg = 0;
#pragma omp parallel for reduction(max : maxid) reduction(+ : neg)
    for (int64_t e = 0; e < m; e++) {
        int64_t a = src[e], b = dst[e];
        if (a < 0 || b < 0)
            neg++;
        maxid = std::max(maxid, std::max(a, b));
    }
    if (neg)
        return fail(TRUST_ERR_FORMAT, "%lld pairs hold negative ids", (long long)neg);

    // ---- compact ids (fromDirectToUndirect.cpp:53-85) ------------------
    // n_fixed >= 0: ids are already 0..n_fixed-1 and all of them are kept
    // (a 1.mmio header fixes the vertex count, preprocess_outdegree.cpp:58-66).
    if (n_fixed >= 0 && maxid >= n_fixed)
        return fail(TRUST_ERR_FORMAT, "id %lld exceeds the declared %lld vertices",
                    (long long)maxid, (long long)n_fixed);
    const int64_t idspace = n_fixed >= 0 ? n_fixed : maxid + 1;
    std::vector<int32_t> cid((size_t)idspace, n_fixed >= 0 ? 1 : 0);
    if (n_fixed < 0) {
#pragma omp parallel for
        for (int64_t e = 0; e < m; e++) {
            __atomic_store_n(&cid[src[e]], 1, __ATOMIC_RELAXED);
            __atomic_store_n(&cid[dst[e]], 1, __ATOMIC_RELAXED);
        }
    }
    int64_t n = 0;
    for (int64_t x = 0; x < idspace; x++)
        cid[x] = cid[x] ? (int32_t)n++ : -1;
    if (n == 0) {  // no vertices: begin.bin holds the single offset 0
        out.V = out.E = 0;
        out.beg.assign(1, 0);
        out.adj.clear();
        return TRUST_OK;
    }

    // ---- undirected unique pairs (:86-109) -----------------------------
    std::vector<uint64_t> keys((size_t)m), tmp((size_t)m);
#pragma omp parallel for
    for (int64_t e = 0; e < m; e++) {
        uint64_t a = (uint64_t)cid[src[e]], b = (uint64_t)cid[dst[e]];
        if (a > b)
            std::swap(a, b);
        keys[e] = (a << 32) | b;
    }
    std::vector<int32_t>().swap(cid);
    const int nb = bits_for((uint64_t)(n > 0 ? n - 1 : 0));
    radix_sort_u64(keys.data(), tmp.data(), (size_t)m, 32 + nb);
    int64_t K = 0;
    for (int64_t e = 0; e < m; e++)
        if (e == 0 || keys[e] != keys[e - 1])
            keys[K++] = keys[e];

    // ---- degree; a self-loop counts twice on the 1.mmio path
    // (preprocess_outdegree.cpp:66-73), once after speedupIO's list dedup
    std::vector<int64_t> deg((size_t)n, 0);
    const int64_t loop_w = o.self_loop_weight;
#pragma omp parallel for
    for (int64_t e = 0; e < K; e++) {
        int64_t a = (int64_t)(keys[e] >> 32), b = (int64_t)(keys[e] & 0xffffffffu);
        if (a == b) {
            __atomic_fetch_add(&deg[a], loop_w, __ATOMIC_RELAXED);
            continue;
        }
        __atomic_fetch_add(&deg[a], 1, __ATOMIC_RELAXED);
        __atomic_fetch_add(&deg[b], 1, __ATOMIC_RELAXED);
    }

    // ---- ascending-degree order (:21-24,226) ---------------------------
    std::vector<SizeRec> s1((size_t)n);
    for (int64_t x = 0; x < n; x++)
        s1[x] = SizeRec{deg[x], (int32_t)x};
    std::sort(s1.begin(), s1.end(),
              [](const SizeRec &a, const SizeRec &b) { return a.size < b.size; });
    std::vector<int32_t> a1((size_t)n);  // vertex -> position
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++)
        a1[s1[i].id] = (int32_t)i;

    // ---- orientation (:79-98) ------------------------------------------
    std::vector<int64_t> outdeg((size_t)n, 0);
#pragma omp parallel for
    for (int64_t e = 0; e < K; e++) {
        int64_t a = (int64_t)(keys[e] >> 32), b = (int64_t)(keys[e] & 0xffffffffu);
        if (a == b)
            continue;
        int64_t s = a1[a] < a1[b] ? a : b;
        __atomic_fetch_add(&outdeg[s], 1, __ATOMIC_RELAXED);
    }

    std::vector<int32_t> a2((size_t)n);  // vertex -> row position p
    if (o.order == TRUST_ORDER_INDEGREE) {
        // ---- ascending-degree order split stably into the classes
        // out-degree > 100, 2..100, < 2 (preprocess_indegree.cpp:131-156)
        auto cls = [](int64_t d) { return d > 100 ? 0 : (d >= 2 ? 1 : 2); };
        int64_t next[3] = {0, 0, 0};
        for (int64_t i = 0; i < n; i++)
            next[cls(outdeg[s1[i].id])]++;
        next[2] = next[0] + next[1];
        next[1] = next[0];
        next[0] = 0;
        for (int64_t i = 0; i < n; i++) {
            const int32_t x = s1[i].id;
            a2[x] = (int32_t)next[cls(outdeg[x])]++;
        }
        std::vector<SizeRec>().swap(s1);
    } else {
        // ---- descending out-degree order (:25-28,236) ------------------
        std::vector<SizeRec> s2((size_t)n);
        for (int64_t i = 0; i < n; i++)
            s2[i] = SizeRec{outdeg[s1[i].id], (int32_t)i};
        std::sort(s2.begin(), s2.end(),
                  [](const SizeRec &a, const SizeRec &b) { return a.size > b.size; });
#pragma omp parallel for
        for (int64_t p = 0; p < n; p++)
            a2[s1[s2[p].id].id] = (int32_t)p;
        std::vector<SizeRec>().swap(s1);
        std::vector<SizeRec>().swap(s2);
    }

    // ---- out-lists by position, neighbours by descending original id ----
    // key = (position of source) << 32 | (n-1 - original id of target)
    int64_t E = 0;
    {
        int64_t w = 0;
        for (int64_t e = 0; e < K; e++) {  // compact in place, sequential order kept
            uint64_t k = keys[e];
            int64_t a = (int64_t)(k >> 32), b = (int64_t)(k & 0xffffffffu);
            if (a == b)
                continue;
            int64_t s = a1[a] < a1[b] ? a : b, t = s == a ? b : a;
            keys[w++] = ((uint64_t)a2[s] << 32) | (uint64_t)(n - 1 - t);
        }
        E = w;
    }
    std::vector<int32_t>().swap(a1);
    radix_sort_u64(keys.data(), tmp.data(), (size_t)E, 32 + nb);
    std::vector<uint64_t>().swap(tmp);

    std::vector<int64_t> begL((size_t)n + 1, 0);
    std::vector<int32_t> L((size_t)E);
#pragma omp parallel for
    for (int64_t e = 0; e < E; e++) {
        int64_t p = (int64_t)(keys[e] >> 32);
        int64_t t = n - 1 - (int64_t)(keys[e] & 0xffffffffu);
        L[e] = a2[t];
        if (e == E - 1 || (int64_t)(keys[e + 1] >> 32) != p)
            begL[p + 1] = e + 1;  // end of row p
    }
    std::vector<uint64_t>().swap(keys);
    std::vector<int32_t>().swap(a2);
    for (int64_t p = 0; p < n; p++)  // rows with no entries
        begL[p + 1] = std::max(begL[p + 1], begL[p]);
    if (o.order != TRUST_ORDER_OUTDEGREE) {  // ids are the row positions
        out.V = n;
        out.E = E;
        out.beg.swap(begL);
        out.adj.swap(L);
        return TRUST_OK;
    }

    // ---- class-preserving first-touch relabel (:100-185) ---------------
    int64_t k2 = -1, k3 = -1, cutoff = n;
    for (int64_t p = 0; p < n; p++) {
        int64_t sz = begL[p + 1] - begL[p];
        if (k2 == -1 && sz <= 100)  // bounder, preprocess_outdegree.cpp:7
            k2 = p;
        if (k3 == -1 && sz < 2)
            k3 = p;
        if (cutoff == n && sz <= 2)
            cutoff = p;
    }
    if (n > 0 && (k2 < 0 || k3 < 0))
        return fail(TRUST_ERR_FORMAT, "no vertex without out-edges (cannot happen)");
    const int64_t scan_end = begL[cutoff];
    std::vector<int64_t> first((size_t)n, INT64_MAX);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < scan_end; e++) {
        int64_t q = L[e];
        int64_t cur = __atomic_load_n(&first[q], __ATOMIC_RELAXED);
        while (e < cur && !__atomic_compare_exchange_n(&first[q], &cur, e, false,
                                                       __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
    }
    std::vector<int32_t> newid((size_t)n);
    const int64_t cls_lo[3] = {0, k2, k3}, cls_hi[3] = {k2, k3, n};
    for (int c = 0; c < 3; c++) {
        const int64_t lo = cls_lo[c], hi = cls_hi[c];
        if (hi <= lo)
            continue;
        std::vector<uint64_t> touched;
        touched.reserve((size_t)(hi - lo));
        for (int64_t q = lo; q < hi; q++)
            if (first[q] != INT64_MAX)
                touched.push_back((uint64_t)first[q]);
        std::vector<uint64_t> ttmp(touched.size());
        radix_sort_u64(touched.data(), ttmp.data(), touched.size(), bits_for((uint64_t)E));
        const int64_t nt = (int64_t)touched.size();
#pragma omp parallel for
        for (int64_t r = 0; r < nt; r++)
            newid[L[(int64_t)touched[r]]] = (int32_t)(lo + r);
        int64_t next = lo + nt;
        for (int64_t q = lo; q < hi; q++)
            if (first[q] == INT64_MAX)
                newid[q] = (int32_t)next++;
    }
    std::vector<int64_t>().swap(first);

    // ---- final CSR in new ids, list order kept (:186-222) --------------
    out.V = n;
    out.E = E;
    out.beg.assign((size_t)n + 1, 0);
    out.adj.resize((size_t)E);
    for (int64_t p = 0; p < n; p++)
        out.beg[(size_t)newid[p] + 1] = begL[p + 1] - begL[p];
    for (int64_t r = 0; r < n; r++)
        out.beg[r + 1] += out.beg[r];
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t p = 0; p < n; p++) {
        int64_t o = out.beg[newid[p]];
        for (int64_t e = begL[p]; e < begL[p + 1]; e++)
            out.adj[o++] = newid[L[e]];
    }
    return TRUST_OK;
}

}  // namespace trust

extern "C" int trust_preprocess_edges(const int32_t *src, const int32_t *dst, int64_t m,
                                      trust_csr *out)
{
    using namespace trust;
    if (!out)
        return fail(TRUST_ERR_ARG, "null output");
    memset(out, 0, sizeof(*out));
    HostCsr g;
    int rc;
    try {
        rc = preprocess_edges(src, dst, m, g);
    } catch (const std::bad_alloc &) {
        return fail(TRUST_ERR_OOM, "out of host memory preprocessing %lld pairs", (long long)m);
    }
    if (rc)
        return rc;
    return export_csr(std::move(g), out);
}

extern "C" int trust_preprocess_edges_ex(const int32_t *src, const int32_t *dst, int64_t m,
                                         int order, uint32_t flags, trust_csr *out)
{
    using namespace trust;
    if (!out)
        return fail(TRUST_ERR_ARG, "null output");
    memset(out, 0, sizeof(*out));
    if (flags & ~TRUST_PRE_SELF_LOOP_ONCE)
        return fail(TRUST_ERR_ARG, "unknown preprocessing flags 0x%x", flags);
    PreprocessOpts o;
    o.order = order;
    o.self_loop_weight = (flags & TRUST_PRE_SELF_LOOP_ONCE) ? 1 : 2;
    HostCsr g;
    int rc;
    try {
        rc = preprocess_edges(src, dst, m, g, o);
    } catch (const std::bad_alloc &) {
        return fail(TRUST_ERR_OOM, "out of host memory preprocessing %lld pairs", (long long)m);
    }
    if (rc)
        return rc;
    return export_csr(std::move(g), out);
}
