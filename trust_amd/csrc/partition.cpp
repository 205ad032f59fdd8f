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
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

#include "internal.h"

namespace trust {

int partition_tiles(const int64_t *beg, const int32_t *adj, int64_t V, int n,
                    std::vector<HostCsr> &tiles)
{
    if (n < 1 || n > 1024)
        return fail(TRUST_ERR_ARG, "partition count n=%d out of range", n);
    tiles.assign((size_t)n * n, HostCsr());
    for (int i = 0; i < n; i++) {
        const int64_t rows = V > i ? (V - i + n - 1) / n : 0;
        std::vector<int64_t> cnt((size_t)n * (size_t)(rows + 1), 0);  // [j][r+1]
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t r = 0; r < rows; r++) {
            int64_t u = r * n + i;
            for (int64_t e = beg[u]; e < beg[u + 1]; e++)
                cnt[(size_t)(adj[e] % n) * (rows + 1) + r + 1]++;
        }
        for (int j = 0; j < n; j++) {
            HostCsr &t = tiles[(size_t)i * n + j];
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
                pos[j] = tiles[(size_t)i * n + j].beg[r];
            for (int64_t e = beg[u]; e < beg[u + 1]; e++) {
                int j = adj[e] % n;
                tiles[(size_t)i * n + j].adj[pos[j]++] = adj[e] / n;
            }
        }
    }
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

extern "C" int trust_partition_write(const char *folder, const trust_csr *csr, int n)
{
    if (!folder || !csr || !csr->beg)
        return fail(TRUST_ERR_ARG, "null argument");
    std::vector<HostCsr> t;
    int rc = partition_tiles(csr->beg, csr->adj, csr->vertices, n, t);
    if (rc)
        return rc;
    int64_t endofprocess = -1;
    for (int64_t u = 0; u < csr->vertices; u++)
        if (csr->beg[u + 1] - csr->beg[u] < 2) {
            endofprocess = u;
            break;
        }
    // C integer division as in the reference (-1/n == 0 for n > 1).
    const int64_t considered = endofprocess / n + 1;
    mkdir(folder, 0755);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            const HostCsr &tile = t[(size_t)i * n + j];
            std::string dir = std::string(folder) + "/partition" + std::to_string(i) + "_" +
                              std::to_string(j);
            mkdir(dir.c_str(), 0755);
            rc = write_file(dir + "/begin.bin", tile.beg.data(),
                            sizeof(int64_t) * (size_t)(tile.V + 1));
            if (rc)
                return rc;
            rc = write_file(dir + "/adjacent.bin", tile.adj.data(),
                            sizeof(int32_t) * (size_t)tile.E);
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
