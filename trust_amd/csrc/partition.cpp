// 2D hash partition of an oriented CSR into n x n tiles, in the layout of
// Preprocess/partition.cpp:72-165 (the layout has to match for the tile
// folders to be interchangeable):
//   tile (i,j) has one row per vertex u with u % n == i (local row u / n),
//   holding the entries v of N+(u) with v % n == j stored as v / n, in the
//   order they appear in N+(u);
//   `map` lists the first endofprocess/n + 1 local rows sorted by
//   descending tile degree (std::sort with the reference comparator,
//   partition.cpp:127-130, so ties fall as they do there), cut at the first
//   row of degree 0, where endofprocess is the first global vertex with
//   out-degree < 2 (:107);
//   `division` is the text "firstwarpvertex firstnothreadvertex\n" (:146-163).
// Rows past the end of a tile count as degree 0 here; the reference reads
// them from stale memory (:136).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

#include "internal.h"

namespace trust {

// The n tiles (i, 0..n-1) of tile row i; `row` receives them.
static void partition_row(const int64_t *beg, const int32_t *adj, int64_t V, int n, int i,
                          HostCsr *row)
{
    {
        const int64_t rows = V > i ? (V - i + n - 1) / n : 0;
        std::vector<int64_t> cnt((size_t)n * (size_t)(rows + 1), 0);  // [j][r+1]
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t r = 0; r < rows; r++) {
            int64_t u = r * n + i;
            for (int64_t e = beg[u]; e < beg[u + 1]; e++)
                cnt[(size_t)(adj[e] % n) * (rows + 1) + r + 1]++;
        }
        for (int j = 0; j < n; j++) {
            HostCsr &t = row[j];
            t.V = rows;
            t.beg.assign((size_t)rows + 1, 0);
            int64_t *c = &cnt[(size_t)j * (rows + 1)];
            for (int64_t r = 0; r < rows; r++)
                t.beg[r + 1] = t.beg[r] + c[r + 1];
            t.E = t.beg[rows];
            t.adj.resize((size_t)t.E);
        }
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t r = 0; r < rows; r++) {
            int64_t u = r * n + i;
            int64_t pos[1024];
            for (int j = 0; j < n; j++)
                pos[j] = row[j].beg[r];
            for (int64_t e = beg[u]; e < beg[u + 1]; e++) {
                int j = adj[e] % n;
                row[j].adj[pos[j]++] = adj[e] / n;
            }
        }
    }
}

int partition_tiles(const int64_t *beg, const int32_t *adj, int64_t V, int n,
                    std::vector<HostCsr> &tiles)
{
    if (n < 1 || n > 1024)
        return fail(TRUST_ERR_ARG, "partition count n=%d out of range", n);
    tiles.assign((size_t)n * n, HostCsr());
    for (int i = 0; i < n; i++)
        partition_row(beg, adj, V, n, i, &tiles[(size_t)i * n]);
    return TRUST_OK;
}

}  // namespace trust

using namespace trust;

extern "C" int trust_partition_csr(const trust_csr *csr, int n, trust_csr *tiles)
{
    if (!csr || !tiles || !csr->beg)
        return fail(TRUST_ERR_ARG, "null argument");
    std::vector<HostCsr> t;
    int rc = partition_tiles(csr->beg, csr->adj, csr->vertices, n, t);
    if (rc)
        return rc;
    for (size_t k = 0; k < t.size(); k++) {
        rc = export_csr(std::move(t[k]), &tiles[k]);
        if (rc)
            return rc;
    }
    return TRUST_OK;
}

namespace {
struct VertexDegree {  // fields of vertexwithdegree, Preprocess/partition.cpp:122-126
    int vertexid;
    int64_t degree;
};
}  // namespace

// Writes the tile folders of tile row i (map and division as described at
// the top of this file).
static int write_row(const std::string &folder, const HostCsr *row, int n, int i,
                     int64_t endofprocess)
{
    // C integer division as in the reference (-1/n == 0 for n > 1).
    const int64_t considered = endofprocess / n + 1;
    for (int j = 0; j < n; j++) {
        const HostCsr &tile = row[j];
        std::string dir = folder + "/partition" + std::to_string(i) + "_" + std::to_string(j);
        mkdir(dir.c_str(), 0755);
        int rc = write_file(dir + "/begin.bin", tile.beg.data(),
                            sizeof(int64_t) * (size_t)(tile.V + 1));
        if (rc)
            return rc;
        rc = write_file(dir + "/adjacent.bin", tile.adj.data(), sizeof(int32_t) * (size_t)tile.E);
        if (rc)
            return rc;
        std::vector<VertexDegree> ver((size_t)std::max<int64_t>(considered, 0));
        for (int64_t r = 0; r < considered; r++) {
            ver[r].vertexid = (int)r;
            ver[r].degree = r < tile.V ? tile.beg[r + 1] - tile.beg[r] : 0;
        }
        std::sort(ver.begin(), ver.end(), [](const VertexDegree &a, const VertexDegree &b) {
            return a.degree > b.degree;
        });
        std::vector<int32_t> map;
        int64_t firstwarp = -1, firstnothread = -1;
        for (int64_t r = 0; r < considered; r++) {
            if (firstwarp == -1 && ver[r].degree <= 100)
                firstwarp = r;
            if (firstnothread == -1 && ver[r].degree < 1) {
                firstnothread = r;
                break;
            }
            map.push_back(ver[r].vertexid);
        }
        if (firstnothread == -1)
            firstnothread = endofprocess / n;
        rc = write_file(dir + "/map", map.data(), sizeof(int32_t) * map.size());
        if (rc)
            return rc;
        char line[64];
        int len = snprintf(line, sizeof(line), "%lld %lld\n", (long long)firstwarp,
                           (long long)firstnothread);
        rc = write_file(dir + "/division", line, (size_t)len);
        if (rc)
            return rc;
    }
    return TRUST_OK;
}

// All n x n tile folders, one tile row in memory at a time.
static int write_tiles(const std::string &folder, const int64_t *beg, const int32_t *adj,
                       int64_t V, int n)
{
    if (n < 1 || n > 1024)
        return fail(TRUST_ERR_ARG, "partition count n=%d out of range", n);
    int64_t endofprocess = -1;  // first vertex with out-degree < 2 (partition.cpp:107)
    for (int64_t u = 0; u < V; u++)
        if (beg[u + 1] - beg[u] < 2) {
            endofprocess = u;
            break;
        }
    mkdir(folder.c_str(), 0755);
    std::vector<HostCsr> row((size_t)n);
    for (int i = 0; i < n; i++) {
        partition_row(beg, adj, V, n, i, row.data());
        int rc = write_row(folder, row.data(), n, i, endofprocess);
        if (rc)
            return rc;
        for (auto &t : row)
            t = HostCsr();
    }
    return TRUST_OK;
}

extern "C" int trust_partition_write(const char *folder, const trust_csr *csr, int n)
{
    if (!folder || !csr || !csr->beg)
        return fail(TRUST_ERR_ARG, "null argument");
    try {
        return write_tiles(folder, csr->beg, csr->adj, csr->vertices, n);
    } catch (const std::bad_alloc &) {
        return fail(TRUST_ERR_OOM, "out of host memory writing %d x %d tiles", n, n);
    }
}

namespace {
// A read-only file mapping.
struct Mapped {
    void *p = MAP_FAILED;
    size_t bytes = 0;
    ~Mapped()
    {
        if (p != MAP_FAILED && bytes)
            munmap(p, bytes);
    }
    int open_file(const std::string &path)
    {
        int fd = ::open(path.c_str(), O_RDONLY);
        if (fd < 0)
            return fail(TRUST_ERR_IO, "cannot open %s", path.c_str());
        struct stat st;
        if (fstat(fd, &st) != 0) {
            ::close(fd);
            return fail(TRUST_ERR_IO, "cannot stat %s", path.c_str());
        }
        bytes = (size_t)st.st_size;
        if (bytes)
            p = mmap(nullptr, bytes, PROT_READ, MAP_SHARED, fd, 0);
        ::close(fd);
        if (bytes && p == MAP_FAILED)
            return fail(TRUST_ERR_IO, "cannot map %s", path.c_str());
        if (bytes)
            madvise(p, bytes, MADV_SEQUENTIAL);
        return TRUST_OK;
    }
};
}  // namespace

extern "C" int trust_partition_folder(const char *folder, int n)
{
    if (!folder)
        return fail(TRUST_ERR_ARG, "null argument");
    const std::string f(folder);
    Mapped mb, ma;
    int rc = mb.open_file(f + "/begin.bin");
    if (rc)
        return rc;
    rc = ma.open_file(f + "/adjacent.bin");
    if (rc)
        return rc;
    if (mb.bytes < sizeof(int64_t) || mb.bytes % sizeof(int64_t) || ma.bytes % sizeof(int32_t))
        return fail(TRUST_ERR_FORMAT, "%s: begin.bin/adjacent.bin sizes are not a CSR", folder);
    const int64_t V = (int64_t)(mb.bytes / sizeof(int64_t)) - 1;
    const int64_t E = (int64_t)(ma.bytes / sizeof(int32_t));
    const int64_t *beg = (const int64_t *)(mb.bytes ? mb.p : nullptr);
    const int32_t *adj = (const int32_t *)(ma.bytes ? ma.p : nullptr);
    rc = trust_csr_validate(beg, adj, V, E);
    if (rc)
        return rc;
    try {
        return write_tiles(f, beg, adj, V, n);
    } catch (const std::bad_alloc &) {
        return fail(TRUST_ERR_OOM, "out of host memory writing %d x %d tiles", n, n);
    }
}
