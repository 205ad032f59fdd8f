// Errors, file I/O and CSR validation for the oriented-CSR input format:
// <folder>/begin.bin holds V+1 int64 offsets and <folder>/adjacent.bin holds
// E int32 ids (writer: Preprocess/preprocess_outdegree.cpp:205-221; reader:
// Without-graph-partition/graph.hpp:18-67, which derives V and E from the
// file sizes and checks nothing).  Here every size and offset is checked and
// a failure is returned, never an exit.
#include <sys/stat.h>

#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "internal.h"

namespace trust {

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

int64_t file_size(const std::string &path)
{
    struct stat st;
    if (stat(path.c_str(), &st) != 0)
        return -1;
    return (int64_t)st.st_size;
}

int read_file(const std::string &path, void *dst, size_t bytes)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f)
        return fail(TRUST_ERR_IO, "cannot open %s: %s", path.c_str(), strerror(errno));
    size_t got = bytes ? fread(dst, 1, bytes, f) : 0;
    fclose(f);
    if (got != bytes)
        return fail(TRUST_ERR_IO, "short read on %s: %zu of %zu bytes", path.c_str(), got, bytes);
    return TRUST_OK;
}

int write_file(const std::string &path, const void *src, size_t bytes)
{
    FILE *f = fopen(path.c_str(), "wb");
    if (!f)
        return fail(TRUST_ERR_IO, "cannot create %s: %s", path.c_str(), strerror(errno));
    size_t put = bytes ? fwrite(src, 1, bytes, f) : 0;
    int rc = fclose(f);
    if (put != bytes || rc != 0)
        return fail(TRUST_ERR_IO, "short write on %s", path.c_str());
    return TRUST_OK;
}

int load_folder(const std::string &folder, HostCsr &out)
{
    const std::string fb = folder + "/begin.bin", fa = folder + "/adjacent.bin";
    int64_t sb = file_size(fb), sa = file_size(fa);
    if (sb < 0)
        return fail(TRUST_ERR_IO, "missing %s", fb.c_str());
    if (sa < 0)
        return fail(TRUST_ERR_IO, "missing %s", fa.c_str());
    if (sb % 8 != 0 || sb < 8)
        return fail(TRUST_ERR_FORMAT, "%s: size %lld is not a positive multiple of 8",
                    fb.c_str(), (long long)sb);
    if (sa % 4 != 0)
        return fail(TRUST_ERR_FORMAT, "%s: size %lld is not a multiple of 4", fa.c_str(),
                    (long long)sa);
    out.V = sb / 8 - 1;
    out.E = sa / 4;
    try {
        out.beg.resize((size_t)out.V + 1);
        out.adj.resize((size_t)out.E);
    } catch (...) {
        return fail(TRUST_ERR_OOM, "cannot allocate %lld bytes for %s", (long long)(sa + sb),
                    folder.c_str());
    }
    int rc = read_file(fb, out.beg.data(), (size_t)sb);
    if (rc)
        return rc;
    rc = read_file(fa, out.adj.data(), (size_t)sa);
    if (rc)
        return rc;
    return trust_csr_validate(out.beg.data(), out.adj.data(), out.V, out.E);
}

int64_t host_max_degree(const int64_t *beg, int64_t V)
{
    int64_t maxd = 0;
#pragma omp parallel for reduction(max : maxd)
    for (int64_t u = 0; u < V; u++)
        maxd = std::max(maxd, beg[u + 1] - beg[u]);
    return maxd;
}

void host_parallel_copy(void *dst, const void *src, size_t bytes)
{
    constexpr size_t kPiece = 4u << 20;
    const long long pieces = (long long)((bytes + kPiece - 1) / kPiece);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < pieces; i++) {
        const size_t o = (size_t)i * kPiece;
        memcpy((char *)dst + o, (const char *)src + o, std::min(kPiece, bytes - o));
    }
}

void host_tile_row_degrees(const trust_csr *tiles, int n, int c, int64_t rows, int32_t *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; r++) {
        int64_t x = 0;
        for (int b = 0; b < n; b++)
            x += tiles[(size_t)c * n + b].beg[r + 1] - tiles[(size_t)c * n + b].beg[r];
        out[r] = (int32_t)std::min<int64_t>(x, INT32_MAX);
    }
}

int export_csr(HostCsr &&src, trust_csr *out)
{
    out->vertices = src.V;
    out->edges = src.E;
    out->beg = (int64_t *)malloc(sizeof(int64_t) * (size_t)(src.V + 1));
    out->adj = (int32_t *)malloc(sizeof(int32_t) * (size_t)(src.E > 0 ? src.E : 1));
    if (!out->beg || !out->adj) {
        free(out->beg);
        free(out->adj);
        out->beg = nullptr;
        out->adj = nullptr;
        return fail(TRUST_ERR_OOM, "cannot allocate exported CSR");
    }
    memcpy(out->beg, src.beg.data(), sizeof(int64_t) * (size_t)(src.V + 1));
    if (src.E)
        memcpy(out->adj, src.adj.data(), sizeof(int32_t) * (size_t)src.E);
    src = HostCsr();
    return TRUST_OK;
}

}  // namespace trust

using namespace trust;

extern "C" const char *trust_last_error(void) { return g_last_error.c_str(); }

extern "C" const char *trust_version(void) { return "trust-mi355x 0.1 (gfx950)"; }

extern "C" void trust_params_init(trust_params *p)
{
    if (!p)
        return;
    memset(p, 0, sizeof(*p));
    p->total_process = 1;
    p->schedule = TRUST_SCHED_DEGREE;
    p->hub_threshold = 100;  // USE_CTA, Without-graph-partition/TC.cu:19
}

extern "C" int trust_csr_validate(const int64_t *beg, const int32_t *adj, int64_t V, int64_t E)
{
    if (V < 0 || E < 0)
        return fail(TRUST_ERR_ARG, "negative size V=%lld E=%lld", (long long)V, (long long)E);
    if (!beg || (E > 0 && !adj))
        return fail(TRUST_ERR_ARG, "null CSR array");
    if (V >= (int64_t)INT32_MAX)
        return fail(TRUST_ERR_OVERFLOW, "V=%lld does not fit int32 vertex ids", (long long)V);
    if (beg[0] != 0)
        return fail(TRUST_ERR_FORMAT, "begin[0]=%lld, expected 0", (long long)beg[0]);
    if (beg[V] != E)
        return fail(TRUST_ERR_FORMAT, "begin[V]=%lld but adjacent holds %lld ids",
                    (long long)beg[V], (long long)E);
    int64_t bad_off = -1;
#pragma omp parallel for schedule(static) reduction(max : bad_off)
    for (int64_t u = 0; u < V; u++)
        if (beg[u + 1] < beg[u])
            bad_off = u > bad_off ? u : bad_off;
    if (bad_off >= 0)
        return fail(TRUST_ERR_FORMAT, "offsets decrease at vertex %lld", (long long)bad_off);
    int64_t bad_id = -1;
#pragma omp parallel for schedule(static) reduction(max : bad_id)
    for (int64_t e = 0; e < E; e++)
        if (adj[e] < 0 || (int64_t)adj[e] >= V)
            bad_id = e > bad_id ? e : bad_id;
    if (bad_id >= 0)
        return fail(TRUST_ERR_FORMAT, "adjacent[%lld]=%d is not a vertex id in [0,%lld)",
                    (long long)bad_id, adj[bad_id], (long long)V);
    return TRUST_OK;
}

extern "C" int trust_csr_load(const char *folder, trust_csr *out)
{
    if (!folder || !out)
        return fail(TRUST_ERR_ARG, "null argument");
    memset(out, 0, sizeof(*out));
    HostCsr g;
    int rc = load_folder(folder, g);
    if (rc)
        return rc;
    return export_csr(std::move(g), out);
}

extern "C" int trust_csr_save(const char *folder, const trust_csr *csr)
{
    if (!folder || !csr || !csr->beg)
        return fail(TRUST_ERR_ARG, "null argument");
    mkdir(folder, 0755);
    std::string f(folder);
    int rc = write_file(f + "/begin.bin", csr->beg, sizeof(int64_t) * (size_t)(csr->vertices + 1));
    if (rc)
        return rc;
    return write_file(f + "/adjacent.bin", csr->adj, sizeof(int32_t) * (size_t)csr->edges);
}

extern "C" void trust_csr_free(trust_csr *csr)
{
    if (!csr)
        return;
    free(csr->beg);
    free(csr->adj);
    csr->beg = nullptr;
    csr->adj = nullptr;
    csr->vertices = csr->edges = 0;
}
