// Out-of-core input contract: a seeded R-MAT edge stream -> the reference's
// degree-oriented CSR files (begin.bin, adjacent.bin), byte-identical to
// preprocess_edges on the same pairs (preprocess.cpp, which follows
// Preprocess/fromDirectToUndirect.cpp and Preprocess/preprocess_outdegree.cpp),
// without ever holding the pairs or the CSR in host memory.
//
// preprocess_edges keeps every pair as a 64-bit key plus a sort buffer (about
// 150 GB at R-MAT scale 28).  Here the pairs live in bucket files under a
// work directory, and memory holds arrays over the vertices plus one bucket:
//
//   0  presence  stream the generator once: which ids occur (a bitmap; the
//                compact id is a rank in it) and a histogram of each pair's
//                smaller raw id, which cuts id ranges into buckets
//   1  pairs     stream the generator again: key (min, max) of compact ids to
//                the bucket of its smaller id
//   2  dedup     per bucket: sort, drop repeats, count degrees, rewrite
//   -            the two std::sort calls of preprocess_edges on the same
//                records (ascending degree, then descending out-degree after
//                orientation), which fix row positions and ties identically;
//                each row's length is known from the out-degrees
//   3  rows      re-bucket the unique pairs by their source's row position
//   4  lists     per row bucket: sort (row, n-1-original target), append the
//                targets' positions to L (a file) and note each vertex's first
//                appearance below the relabel's scan end
//   -            first-touch relabel per class, as preprocess_edges does it
//   5  placement re-bucket L's entries by their new row, then place each
//                bucket's entries in row order and write adjacent.bin in order
//
// Every bucket holds at most bucket_pairs keys unless one histogram cell
// alone holds more.  Only the out-degree row order (the default contract,
// self-loops counted twice as on the 1.mmio path) is produced this way.
#include <omp.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

namespace trust {
namespace {

int bits_for(uint64_t x)
{
    int b = 0;
    while (b < 64 && (x >> b) != 0)
        b++;
    return b;
}

struct SizeRec {  // the fields preprocess_edges' comparators look at
    int64_t size;
    int32_t id;
};

// Bucket files under the work directory; removed when done.
struct Buckets {
    std::vector<std::string> path;
    std::vector<FILE *> f;
    std::vector<int64_t> count;
    ~Buckets() { close_and_remove(); }
    int create(const std::string &dir, const char *tag, int n)
    {
        close_and_remove();
        path.resize((size_t)n);
        f.assign((size_t)n, nullptr);
        count.assign((size_t)n, 0);
        for (int b = 0; b < n; b++) {
            path[(size_t)b] = dir + "/" + tag + std::to_string(b) + ".bin";
            f[(size_t)b] = fopen(path[(size_t)b].c_str(), "wb");
            if (!f[(size_t)b])
                return fail(TRUST_ERR_IO, "cannot create %s: %s", path[(size_t)b].c_str(),
                            strerror(errno));
        }
        return TRUST_OK;
    }
    int append(int b, const uint64_t *x, size_t n)
    {
        if (n && fwrite(x, sizeof(uint64_t), n, f[(size_t)b]) != n)
            return fail(TRUST_ERR_IO, "short write on %s", path[(size_t)b].c_str());
        count[(size_t)b] += (int64_t)n;
        return TRUST_OK;
    }
    int finish()
    {
        for (size_t b = 0; b < f.size(); b++) {
            if (f[b] && fclose(f[b]) != 0)
                return fail(TRUST_ERR_IO, "cannot close %s", path[b].c_str());
            f[b] = nullptr;
        }
        return TRUST_OK;
    }
    int load(int b, std::vector<uint64_t> &x) const
    {
        x.resize((size_t)count[(size_t)b]);
        return read_file(path[(size_t)b], x.data(), sizeof(uint64_t) * x.size());
    }
    int rewrite(int b, const uint64_t *x, size_t n)
    {
        count[(size_t)b] = (int64_t)n;
        return write_file(path[(size_t)b], x, sizeof(uint64_t) * n);
    }
    void remove_one(int b)
    {
        if (b < (int)path.size() && !path[(size_t)b].empty()) {
            unlink(path[(size_t)b].c_str());
            path[(size_t)b].clear();
        }
    }
    void close_and_remove()
    {
        for (size_t b = 0; b < f.size(); b++)
            if (f[b])
                fclose(f[b]);
        f.clear();
        for (auto &p : path)
            if (!p.empty())
                unlink(p.c_str());
        path.clear();
    }
};

// Splits keys among buckets and appends each bucket's run (keys keep their
// order inside a run; a key whose bucket is negative is dropped).
template <typename BucketOf>
int scatter(Buckets &bk, const std::vector<uint64_t> &keys, BucketOf bucket_of,
            std::vector<uint64_t> &tmp)
{
    const int nb = (int)bk.f.size();
    const int64_t nk = (int64_t)keys.size();
    const int nt = omp_get_max_threads();
    std::vector<int> which((size_t)nk);
    // per thread (a static range of keys) and bucket: count, then offset
    std::vector<int64_t> cnt((size_t)nt * nb, 0);
#pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num();
        const int64_t lo = nk * t / nt, hi = nk * (t + 1) / nt;
        int64_t *c = cnt.data() + (size_t)t * nb;
        for (int64_t i = lo; i < hi; i++) {
            const int b = bucket_of(keys[(size_t)i]);
            which[(size_t)i] = b;
            if (b >= 0)
                c[b]++;
        }
    }
    std::vector<int64_t> start((size_t)nb + 1, 0);
    for (int b = 0; b < nb; b++) {
        int64_t run = start[(size_t)b];
        for (int t = 0; t < nt; t++) {
            const int64_t x = cnt[(size_t)t * nb + b];
            cnt[(size_t)t * nb + b] = run;
            run += x;
        }
        start[(size_t)b + 1] = run;
    }
    tmp.resize((size_t)start[(size_t)nb]);
#pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num();
        const int64_t lo = nk * t / nt, hi = nk * (t + 1) / nt;
        int64_t *at = cnt.data() + (size_t)t * nb;
        for (int64_t i = lo; i < hi; i++) {
            const int b = which[(size_t)i];
            if (b >= 0)
                tmp[(size_t)at[b]++] = keys[(size_t)i];
        }
    }
    for (int b = 0; b < nb; b++) {
        int rc = bk.append(b, tmp.data() + start[(size_t)b],
                           (size_t)(start[(size_t)b + 1] - start[(size_t)b]));
        if (rc)
            return rc;
    }
    return TRUST_OK;
}

// Greedy cut of consecutive cells (weights w) into ranges of at most cap;
// returns each range's first cell, plus w.size() at the end.
std::vector<int64_t> cut_ranges(const std::vector<int64_t> &w, int64_t cap)
{
    std::vector<int64_t> first{0};
    int64_t run = 0;
    for (size_t i = 0; i < w.size(); i++) {
        if (run > 0 && run + w[i] > cap) {
            first.push_back((int64_t)i);
            run = 0;
        }
        run += w[i];
    }
    first.push_back((int64_t)w.size());
    return first;
}

// Progress lines on stderr when TRUST_OOC_VERBOSE is set: a scale-28 run
// takes minutes per pass.
struct Progress {
    bool on = getenv("TRUST_OOC_VERBOSE") != nullptr;
    double t0 = omp_get_wtime();
    void mark(const char *what) const
    {
        if (on)
            fprintf(stderr, "[ooc] %s: %.1f s\n", what, omp_get_wtime() - t0);
    }
};

}  // namespace

int rmat_csr_ooc(int scale, int64_t m, double a, double b, double c, uint64_t seed,
                 const std::string &out_dir, const std::string &work_dir, int64_t bucket_pairs)
{
    if (scale < 1 || scale > 30 || m < 0)
        return fail(TRUST_ERR_ARG, "rmat_csr_ooc: bad arguments (scale=%d m=%lld)", scale,
                    (long long)m);
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return fail(TRUST_ERR_ARG, "rmat_csr_ooc: bad quadrant probabilities");
    if (bucket_pairs < 1024)
        return fail(TRUST_ERR_ARG, "rmat_csr_ooc: bucket of %lld pairs is too small",
                    (long long)bucket_pairs);
    const Progress prog;
    const int64_t chunk = std::min<int64_t>(bucket_pairs, 1ll << 24);  // generator pairs per step
    const int64_t idspace = 1ll << scale;
    // Histogram cells of a pair's smaller raw id, which the buckets are cut
    // from: R-MAT's smaller ids crowd near 0 (at scale 28 half the pairs have
    // a smaller id below 2^16), so ids below 2^18 get a cell each and the
    // rest share 2^16 equal ranges.
    const int64_t fine = std::min<int64_t>(idspace, 1ll << 18);
    int cls = 0;
    while (((idspace - fine) >> cls) > (1ll << 16))
        cls++;
    const int64_t ncoarse = fine + ((idspace - fine + (1ll << cls) - 1) >> cls);
    auto cell = [fine, cls](uint32_t x) -> int64_t {
        return (int64_t)x < fine ? (int64_t)x : fine + (((int64_t)x - fine) >> cls);
    };

    // ---- 0: presence bitmap and coarse histogram of the smaller raw id
    std::vector<uint64_t> present((size_t)((idspace + 63) / 64), 0);
    std::vector<int64_t> coarse((size_t)ncoarse, 0);
    std::vector<int32_t> src((size_t)chunk), dst((size_t)chunk);
    for (int64_t e0 = 0; e0 < m; e0 += chunk) {
        const int64_t cnt = std::min(chunk, m - e0);
        gen_rmat_range(scale, a, b, c, seed, e0, cnt, src.data(), dst.data());
#pragma omp parallel
        {
            std::vector<int64_t> h((size_t)ncoarse, 0);
#pragma omp for schedule(static)
            for (int64_t i = 0; i < cnt; i++) {
                const uint32_t u = (uint32_t)src[(size_t)i], v = (uint32_t)dst[(size_t)i];
                // test first: R-MAT repeats its hub ids, whose words are set
                // early, and a read keeps their cache lines shared
                for (const uint32_t x : {u, v})
                    if (!(__atomic_load_n(&present[x >> 6], __ATOMIC_RELAXED) & (1ull << (x & 63))))
                        __atomic_fetch_or(&present[x >> 6], 1ull << (x & 63), __ATOMIC_RELAXED);
                h[(size_t)cell(std::min(u, v))]++;
            }
#pragma omp critical
            for (int64_t i = 0; i < ncoarse; i++)
                coarse[(size_t)i] += h[(size_t)i];
        }
    }
    prog.mark("0 presence");
    std::vector<int64_t> rank_base(present.size() + 1, 0);  // compact id = rank in the bitmap
    for (size_t w = 0; w < present.size(); w++)
        rank_base[w + 1] = rank_base[w] + __builtin_popcountll(present[w]);
    const int64_t n = rank_base.back();
    if (n > INT32_MAX)
        return fail(TRUST_ERR_OVERFLOW, "%lld vertices exceed int32 ids", (long long)n);
    auto cid = [&](uint32_t x) -> uint64_t {
        return (uint64_t)(rank_base[x >> 6] +
                          __builtin_popcountll(present[x >> 6] & ((1ull << (x & 63)) - 1)));
    };
    mkdir(out_dir.c_str(), 0755);
    if (n == 0) {
        const int64_t zero = 0;
        int rc = write_file(out_dir + "/begin.bin", &zero, sizeof(zero));
        return rc ? rc : write_file(out_dir + "/adjacent.bin", nullptr, 0);
    }
    const int nb = bits_for((uint64_t)(n - 1));

    // ---- 1: pairs to buckets of the smaller id's coarse range
    const std::vector<int64_t> cut = cut_ranges(coarse, bucket_pairs);
    const int nbk = (int)cut.size() - 1;
    std::vector<int> bucket_of_coarse((size_t)ncoarse);
    for (int k = 0; k < nbk; k++)
        for (int64_t i = cut[(size_t)k]; i < cut[(size_t)k + 1]; i++)
            bucket_of_coarse[(size_t)i] = k;
    std::vector<int64_t>().swap(coarse);
    Buckets pairs;
    int rc = pairs.create(work_dir, "trust_ooc_pairs_", nbk);
    if (rc)
        return rc;
    std::vector<uint64_t> keys, tmp;
    for (int64_t e0 = 0; e0 < m; e0 += chunk) {
        const int64_t cnt = std::min(chunk, m - e0);
        gen_rmat_range(scale, a, b, c, seed, e0, cnt, src.data(), dst.data());
        keys.resize((size_t)cnt);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < cnt; i++) {
            uint32_t u = (uint32_t)src[(size_t)i], v = (uint32_t)dst[(size_t)i];
            if (u > v)
                std::swap(u, v);
            // the raw id travels in the low bits until the bucket is known
            keys[(size_t)i] = ((uint64_t)u << 32) | v;
        }
        rc = scatter(pairs, keys,
                     [&](uint64_t k) { return bucket_of_coarse[(size_t)cell((uint32_t)(k >> 32))]; },
                     tmp);
        if (rc)
            return rc;
    }
    std::vector<int32_t>().swap(src);
    std::vector<int32_t>().swap(dst);
    if ((rc = pairs.finish()))
        return rc;
    prog.mark("1 pairs to buckets");

    // ---- 2: per bucket: compact ids, sort, drop repeats, degrees
    std::vector<int64_t> deg((size_t)n, 0);
    int64_t K = 0;
    for (int k = 0; k < nbk; k++) {
        if ((rc = pairs.load(k, keys)))
            return rc;
        const int64_t nk = (int64_t)keys.size();
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nk; i++) {
            const uint64_t x = keys[(size_t)i];
            keys[(size_t)i] = (cid((uint32_t)(x >> 32)) << 32) | cid((uint32_t)x);
        }
        tmp.resize((size_t)nk);
        radix_sort_u64(keys.data(), tmp.data(), (size_t)nk, 32 + nb);
        int64_t w = 0;
        for (int64_t i = 0; i < nk; i++)
            if (i == 0 || keys[(size_t)i] != keys[(size_t)i - 1])
                keys[(size_t)w++] = keys[(size_t)i];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < w; i++) {
            const int64_t x = (int64_t)(keys[(size_t)i] >> 32), y = (int64_t)(keys[(size_t)i] & 0xffffffffu);
            if (x == y) {
                __atomic_fetch_add(&deg[(size_t)x], 2, __ATOMIC_RELAXED);  // the 1.mmio path
                continue;
            }
            __atomic_fetch_add(&deg[(size_t)x], 1, __ATOMIC_RELAXED);
            __atomic_fetch_add(&deg[(size_t)y], 1, __ATOMIC_RELAXED);
        }
        if ((rc = pairs.rewrite(k, keys.data(), (size_t)w)))
            return rc;
        K += w;
    }
    std::vector<uint64_t>().swap(present);
    std::vector<int64_t>().swap(rank_base);
    prog.mark("2 dedup and degrees");

    // ---- row positions: the same std::sort calls as preprocess_edges
    std::vector<SizeRec> s1((size_t)n);
    for (int64_t x = 0; x < n; x++)
        s1[(size_t)x] = SizeRec{deg[(size_t)x], (int32_t)x};
    std::vector<int64_t>().swap(deg);
    std::sort(s1.begin(), s1.end(), [](const SizeRec &p, const SizeRec &q) { return p.size < q.size; });
    std::vector<int32_t> a1((size_t)n);
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++)
        a1[(size_t)s1[(size_t)i].id] = (int32_t)i;
    std::vector<int64_t> outdeg((size_t)n, 0);
    for (int k = 0; k < nbk; k++) {
        if ((rc = pairs.load(k, keys)))
            return rc;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)keys.size(); i++) {
            const int64_t x = (int64_t)(keys[(size_t)i] >> 32), y = (int64_t)(keys[(size_t)i] & 0xffffffffu);
            if (x != y)
                __atomic_fetch_add(&outdeg[(size_t)(a1[(size_t)x] < a1[(size_t)y] ? x : y)], 1,
                                   __ATOMIC_RELAXED);
        }
    }
    std::vector<int32_t> a2((size_t)n);  // vertex -> row position
    std::vector<int64_t> begL((size_t)n + 1, 0);
    {
        std::vector<SizeRec> s2((size_t)n);
        for (int64_t i = 0; i < n; i++)
            s2[(size_t)i] = SizeRec{outdeg[(size_t)s1[(size_t)i].id], (int32_t)i};
        std::sort(s2.begin(), s2.end(), [](const SizeRec &p, const SizeRec &q) { return p.size > q.size; });
#pragma omp parallel for
        for (int64_t p = 0; p < n; p++)
            a2[(size_t)s1[(size_t)s2[(size_t)p].id].id] = (int32_t)p;
        for (int64_t p = 0; p < n; p++)
            begL[(size_t)p + 1] = begL[(size_t)p] + s2[(size_t)p].size;
    }
    std::vector<SizeRec>().swap(s1);
    std::vector<int64_t>().swap(outdeg);
    const int64_t E = begL[(size_t)n];
    prog.mark("row positions");

    // ---- 3: unique pairs to buckets of their source's row position
    std::vector<int64_t> rowlen((size_t)n);
    for (int64_t p = 0; p < n; p++)
        rowlen[(size_t)p] = begL[(size_t)p + 1] - begL[(size_t)p];
    const std::vector<int64_t> rcut = cut_ranges(rowlen, bucket_pairs);
    std::vector<int64_t>().swap(rowlen);
    const int nrb = (int)rcut.size() - 1;
    Buckets rows;
    if ((rc = rows.create(work_dir, "trust_ooc_rows_", nrb)))
        return rc;
    for (int k = 0; k < nbk; k++) {
        if ((rc = pairs.load(k, keys)))
            return rc;
        pairs.remove_one(k);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)keys.size(); i++) {
            const int64_t x = (int64_t)(keys[(size_t)i] >> 32), y = (int64_t)(keys[(size_t)i] & 0xffffffffu);
            if (x == y) {
                keys[(size_t)i] = ~0ull;  // a self-loop: no row entry
                continue;
            }
            const int64_t s = a1[(size_t)x] < a1[(size_t)y] ? x : y, t = s == x ? y : x;
            keys[(size_t)i] = ((uint64_t)a2[(size_t)s] << 32) | (uint64_t)(n - 1 - t);
        }
        rc = scatter(rows, keys, [&](uint64_t key) {
            if (key == ~0ull)
                return -1;
            const int64_t p = (int64_t)(key >> 32);
            return (int)(std::upper_bound(rcut.begin(), rcut.end(), p) - rcut.begin()) - 1;
        }, tmp);
        if (rc)
            return rc;
    }
    std::vector<int32_t>().swap(a1);
    if ((rc = rows.finish()))
        return rc;
    prog.mark("3 rows to buckets");

    // ---- relabel classes and scan end (preprocess_outdegree.cpp:100-185)
    int64_t k2 = -1, k3 = -1, cutoff = n;
    for (int64_t p = 0; p < n; p++) {
        const int64_t sz = begL[(size_t)p + 1] - begL[(size_t)p];
        if (k2 == -1 && sz <= 100)
            k2 = p;
        if (k3 == -1 && sz < 2)
            k3 = p;
        if (cutoff == n && sz <= 2)
            cutoff = p;
    }
    if (k2 < 0 || k3 < 0)
        return fail(TRUST_ERR_FORMAT, "no vertex without out-edges (cannot happen)");
    const int64_t scan_end = begL[(size_t)cutoff];

    // ---- 4: rows in order: L (positions of targets) and first touches
    const std::string lpath = work_dir + "/trust_ooc_L.bin";
    struct Unlink {
        std::string p;
        ~Unlink() { unlink(p.c_str()); }
    } l_guard{lpath};
    FILE *lf = fopen(lpath.c_str(), "wb");
    if (!lf)
        return fail(TRUST_ERR_IO, "cannot create %s: %s", lpath.c_str(), strerror(errno));
    std::vector<int64_t> first((size_t)n, INT64_MAX);
    std::vector<int32_t> L;
    int64_t e0 = 0;
    for (int k = 0; k < nrb; k++) {
        if ((rc = rows.load(k, keys))) {
            fclose(lf);
            return rc;
        }
        rows.remove_one(k);
        const int64_t nk = (int64_t)keys.size();
        tmp.resize((size_t)nk);
        radix_sort_u64(keys.data(), tmp.data(), (size_t)nk, 32 + nb);
        L.resize((size_t)nk);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nk; i++) {
            const int64_t t = n - 1 - (int64_t)(keys[(size_t)i] & 0xffffffffu);
            const int32_t q = a2[(size_t)t];
            L[(size_t)i] = q;
            const int64_t e = e0 + i;
            if (e < scan_end) {
                int64_t cur = __atomic_load_n(&first[(size_t)q], __ATOMIC_RELAXED);
                while (e < cur && !__atomic_compare_exchange_n(&first[(size_t)q], &cur, e, false,
                                                               __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                }
            }
        }
        if (nk && fwrite(L.data(), sizeof(int32_t), (size_t)nk, lf) != (size_t)nk) {
            fclose(lf);
            return fail(TRUST_ERR_IO, "short write on %s", lpath.c_str());
        }
        e0 += nk;
    }
    if (fclose(lf) != 0)
        return fail(TRUST_ERR_IO, "cannot close %s", lpath.c_str());
    std::vector<int32_t>().swap(a2);
    prog.mark("4 lists");
    if (e0 != E)
        return fail(TRUST_ERR_FORMAT, "row buckets hold %lld entries, rows %lld", (long long)e0,
                    (long long)E);

    // ---- first-touch relabel per class: rank of the first touch, then
    // untouched vertices in position order
    if (bits_for((uint64_t)E) + 31 > 64)
        return fail(TRUST_ERR_OVERFLOW, "%lld entries: too many for the relabel's keys", (long long)E);
    std::vector<int32_t> newid((size_t)n);
    const int64_t cls_lo[3] = {0, k2, k3}, cls_hi[3] = {k2, k3, n};
    for (int cl = 0; cl < 3; cl++) {
        const int64_t lo = cls_lo[cl], hi = cls_hi[cl];
        if (hi <= lo)
            continue;
        std::vector<uint64_t> touched;
        for (int64_t q = lo; q < hi; q++)
            if (first[(size_t)q] != INT64_MAX)
                touched.push_back(((uint64_t)first[(size_t)q] << 31) | (uint64_t)q);
        std::vector<uint64_t> ttmp(touched.size());
        radix_sort_u64(touched.data(), ttmp.data(), touched.size(), bits_for((uint64_t)E) + 31);
        const int64_t nt = (int64_t)touched.size();
#pragma omp parallel for
        for (int64_t r = 0; r < nt; r++)
            newid[(size_t)(touched[(size_t)r] & 0x7fffffffu)] = (int32_t)(lo + r);
        int64_t next = lo + nt;
        for (int64_t q = lo; q < hi; q++)
            if (first[(size_t)q] == INT64_MAX)
                newid[(size_t)q] = (int32_t)next++;
    }
    std::vector<int64_t>().swap(first);
    prog.mark("relabel");

    // ---- 5: final rows in new ids, list order kept
    std::vector<int64_t> beg((size_t)n + 1, 0);
    for (int64_t p = 0; p < n; p++)
        beg[(size_t)newid[(size_t)p] + 1] = begL[(size_t)p + 1] - begL[(size_t)p];
    for (int64_t r = 0; r < n; r++)
        beg[(size_t)r + 1] += beg[(size_t)r];
    std::vector<int64_t> newlen((size_t)n);
    for (int64_t r = 0; r < n; r++)
        newlen[(size_t)r] = beg[(size_t)r + 1] - beg[(size_t)r];
    const std::vector<int64_t> fcut = cut_ranges(newlen, bucket_pairs);
    std::vector<int64_t>().swap(newlen);
    const int nfb = (int)fcut.size() - 1;
    Buckets fin;
    if ((rc = fin.create(work_dir, "trust_ooc_final_", nfb)))
        return rc;
    {
        FILE *in = fopen(lpath.c_str(), "rb");
        if (!in)
            return fail(TRUST_ERR_IO, "cannot open %s", lpath.c_str());
        int64_t p = 0;
        for (int64_t at = 0; at < E; at += chunk) {
            const int64_t cnt = std::min(chunk, E - at);
            L.resize((size_t)cnt);
            if (fread(L.data(), sizeof(int32_t), (size_t)cnt, in) != (size_t)cnt) {
                fclose(in);
                return fail(TRUST_ERR_IO, "short read on %s", lpath.c_str());
            }
            keys.resize((size_t)cnt);
            for (int64_t i = 0; i < cnt; i++) {  // entry at + i belongs to row p
                while (begL[(size_t)p + 1] <= at + i)
                    p++;
                keys[(size_t)i] = ((uint64_t)(uint32_t)newid[(size_t)p] << 32) |
                                  (uint32_t)newid[(size_t)L[(size_t)i]];
            }
            rc = scatter(fin, keys, [&](uint64_t key) {
                const int64_t r = (int64_t)(key >> 32);
                return (int)(std::upper_bound(fcut.begin(), fcut.end(), r) - fcut.begin()) - 1;
            }, tmp);
            if (rc) {
                fclose(in);
                return rc;
            }
        }
        fclose(in);
    }
    if ((rc = fin.finish()))
        return rc;
    prog.mark("5 entries to buckets");
    std::vector<int64_t>().swap(begL);
    std::vector<int32_t>().swap(newid);
    const std::string apath = out_dir + "/adjacent.bin";
    FILE *af = fopen(apath.c_str(), "wb");
    if (!af)
        return fail(TRUST_ERR_IO, "cannot create %s: %s", apath.c_str(), strerror(errno));
    for (int k = 0; k < nfb; k++) {
        if ((rc = fin.load(k, keys))) {
            fclose(af);
            return rc;
        }
        fin.remove_one(k);
        const int64_t r0 = fcut[(size_t)k], r1 = fcut[(size_t)k + 1];
        const int64_t base = beg[(size_t)r0];
        L.assign((size_t)(beg[(size_t)r1] - base), 0);
        std::vector<int64_t> cur(beg.begin() + r0, beg.begin() + r1);
        for (size_t i = 0; i < keys.size(); i++) {  // file order keeps each row's list order
            const int64_t r = (int64_t)(keys[i] >> 32);
            L[(size_t)(cur[(size_t)(r - r0)]++ - base)] = (int32_t)(keys[i] & 0xffffffffu);
        }
        if (!L.empty() && fwrite(L.data(), sizeof(int32_t), L.size(), af) != L.size()) {
            fclose(af);
            return fail(TRUST_ERR_IO, "short write on %s", apath.c_str());
        }
    }
    if (fclose(af) != 0)
        return fail(TRUST_ERR_IO, "cannot close %s", apath.c_str());
    prog.mark("5 placement");
    (void)K;
    return write_file(out_dir + "/begin.bin", beg.data(), sizeof(int64_t) * beg.size());
}

}  // namespace trust

extern "C" int trust_rmat_csr_ooc(int scale, int64_t m, double a, double b, double c,
                                  uint64_t seed, const char *out_folder, const char *work_dir,
                                  int64_t bucket_pairs)
{
    using namespace trust;
    if (!out_folder || !work_dir)
        return fail(TRUST_ERR_ARG, "null folder");
    try {
        return rmat_csr_ooc(scale, m, a, b, c, seed, out_folder, work_dir, bucket_pairs);
    } catch (const std::bad_alloc &) {
        return fail(TRUST_ERR_OOM, "out of host memory in the out-of-core preprocessing");
    }
}
