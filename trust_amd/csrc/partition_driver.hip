// 2D hash-partition path on one GPU.
//
// A job (i,j,k) hashes tile (i,j), walks tile (i,k) and probes tile (k,j)
// (With-graph-partition/TC.cu:985-994); summing the n^3 jobs gives the
// whole count because a triangle u->v->w (u->w) has exactly one (u%n, w%n,
// v%n).  The reference runs one job per MPI rank and copies its three
// tiles synchronously each time (TC.cu:1034-1077).  Here all jobs stream
// through one device: tiles live in a bounded device cache, missing tiles
// are copied on a second stream while the previous job counts, and each
// job's uint64 lands in its own control block, summed once at the end.
#include <hip/hip_runtime.h>

#include <sys/types.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "engine.h"
#include "internal.h"

namespace trust {

// Rows of tile (i, *) hash/walk vertices u with u % n == i; entries of a
// walk tile index rows of the probe tile.  HT and probe entries are only
// compared for equality, so they only need to be valid (non-negative) keys.
static int check_job_tiles(const trust_csr &ht, const trust_csr &walk, const trust_csr &probe)
{
    const trust_csr *t[3] = {&ht, &walk, &probe};
    for (int q = 0; q < 3; q++) {
        const trust_csr &c = *t[q];
        if (c.vertices < 0 || !c.beg || (c.edges > 0 && !c.adj))
            return fail(TRUST_ERR_ARG, "tile %d is empty or null", q);
        if (c.beg[0] != 0 || c.beg[c.vertices] != c.edges)
            return fail(TRUST_ERR_FORMAT, "tile %d: offsets do not span its %lld entries", q,
                        (long long)c.edges);
        for (int64_t u = 0; u < c.vertices; u++)
            if (c.beg[u + 1] < c.beg[u])
                return fail(TRUST_ERR_FORMAT, "tile %d: offsets decrease at row %lld", q,
                            (long long)u);
    }
    if (ht.vertices != walk.vertices)
        return fail(TRUST_ERR_FORMAT, "hash tile has %lld rows but walk tile has %lld",
                    (long long)ht.vertices, (long long)walk.vertices);
    int64_t bad = -1;
    for (int64_t e = 0; e < walk.edges && bad < 0; e++)
        if (walk.adj[e] < 0 || walk.adj[e] >= probe.vertices)
            bad = e;
    if (bad >= 0)
        return fail(TRUST_ERR_FORMAT, "walk tile entry %lld = %d is not a row of the probe tile",
                    (long long)bad, walk.adj[bad]);
    for (int q = 0; q < 3; q += 2)
        for (int64_t e = 0; e < t[q]->edges; e++)
            if (t[q]->adj[e] < 0)
                return fail(TRUST_ERR_FORMAT, "tile %d holds a negative id", q);
    return TRUST_OK;
}

namespace {

// Two pinned host buffers that tile uploads pass through.  The host copies a
// tile into one with all its threads while the DMA engine drains the other,
// so an upload costs the caller a parallel memcpy and the transfer itself
// runs asynchronously.  A pageable hipMemcpyAsync instead stages every byte
// through the runtime's own buffers on the calling thread before it returns.
// Allocated with hipHostMalloc.
struct PinnedStage {
    char *buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};  // the DMA out of buf[b] has finished
    int next = 0;
    PinnedStage() = default;
    PinnedStage(const PinnedStage &) = delete;
    PinnedStage &operator=(const PinnedStage &) = delete;
    ~PinnedStage()
    {
        for (int b = 0; b < 2; b++) {
            if (done[b]) {
                (void)hipEventSynchronize(done[b]);
                (void)hipEventDestroy(done[b]);
            }
            if (buf[b])
                (void)hipHostFree(buf[b]);
        }
    }
    // false when pinned memory is not available (uploads then stay pageable)
    bool init(size_t n)
    {
        for (int b = 0; b < 2; b++)
            if (hipHostMalloc((void **)&buf[b], n, hipHostMallocDefault) != hipSuccess ||
                hipEventCreateWithFlags(&done[b], hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                return false;
            }
        return true;
    }
};

// Tile t to the device CSR g through the next pinned buffer, on stream s.
static int upload_pinned(Ctx *ctx, PinnedStage &st, DevCsr &g, const trust_csr &t, bool off64,
                         hipStream_t s)
{
    const size_t ab = sizeof(int32_t) * (size_t)t.edges;
    const size_t boff = (ab + 255) & ~(size_t)255;
    const size_t bb = sizeof(int64_t) * (size_t)(t.vertices + 1);
    const int b = st.next;
    st.next ^= 1;
    HIPCHK(hipEventSynchronize(st.done[b]));
    char *h = st.buf[b];
    host_parallel_copy(h, t.adj, ab);
    host_parallel_copy(h + boff, t.beg, bb);
    g.V = t.vertices;
    g.E = t.edges;
    g.off64 = off64;
    g.max_deg = host_max_degree(t.beg, t.vertices);
    if (ab > 0)
        HIPCHK(hipMemcpyAsync(g.adj.p, h, ab, hipMemcpyHostToDevice, s));
    if (off64) {
        HIPCHK(hipMemcpyAsync(g.beg.p, h + boff, bb, hipMemcpyHostToDevice, s));
    } else {
        HIPCHK(hipMemcpyAsync(ctx->staging.p, h + boff, bb, hipMemcpyHostToDevice, s));
        HIPCHK(launch_offsets_to_i32((const int64_t *)ctx->staging.p, (int32_t *)g.beg.p,
                                     t.vertices + 1, s));
    }
    HIPCHK(hipEventRecord(st.done[b], s));
    return TRUST_OK;
}

struct TileSlot {
    DevCsr g;
    int tile = -1;                 // i*n+j held, -1 when empty
    hipEvent_t ready = nullptr;    // copy finished (copy stream)
    hipEvent_t released = nullptr; // last job reading it finished (compute stream)
    long long last_job = -1;
};

}  // namespace

// ---- job journal (resume after an interrupted stream) -------------------
// A text file: a header that names the partition, then one line per
// finished job, appended and flushed as each job's count lands on the host:
//   trust-tiles-journal 1 n=<n> tiles=<fingerprint>
//   job <id> <i> <j> <k> <triangles> <probes>
// Counts are additive, so a resumed stream counts only the jobs the file
// lacks.  A last line without its newline (a write cut short) is ignored.
struct Journal {
    FILE *f = nullptr;
    std::vector<char> have;                // job id -> already in the file
    std::vector<unsigned long long> cnt, prb;
    ~Journal()
    {
        if (f)
            fclose(f);
    }
};

static uint64_t tiles_fingerprint(const trust_csr *tiles, int n)
{
    uint64_t h = 1469598103934665603ull;  // FNV-1a over sizes and edge entries
    auto mix = [&h](uint64_t x) {
        for (int b = 0; b < 8; b++) {
            h ^= (x >> (8 * b)) & 0xff;
            h *= 1099511628211ull;
        }
    };
    for (int t = 0; t < n * n; t++) {
        const trust_csr &c = tiles[t];
        mix((uint64_t)c.vertices);
        mix((uint64_t)c.edges);
        const int64_t step = std::max<int64_t>(1, c.edges / 64);
        for (int64_t e = 0; e < c.edges; e += step)
            mix((uint64_t)(uint32_t)c.adj[e]);
    }
    return h;
}

static int journal_open(Journal &jr, const char *path, int n, const trust_csr *tiles,
                        long long njobs)
{
    char head[128];
    snprintf(head, sizeof(head), "trust-tiles-journal 1 n=%d tiles=%016llx", n,
             (unsigned long long)tiles_fingerprint(tiles, n));
    jr.have.assign((size_t)njobs, 0);
    jr.cnt.assign((size_t)njobs, 0);
    jr.prb.assign((size_t)njobs, 0);
    std::string text;
    if (FILE *in = fopen(path, "rb")) {
        char buf[1 << 16];
        size_t got;
        while ((got = fread(buf, 1, sizeof(buf), in)) > 0)
            text.append(buf, got);
        fclose(in);
    }
    bool fresh = text.empty();
    if (!fresh) {
        const size_t eol = text.find('\n');
        if (eol == std::string::npos || text.compare(0, eol, head) != 0)
            return fail(TRUST_ERR_FORMAT, "journal %s belongs to another partition (expected \"%s\")",
                        path, head);
        size_t at = eol + 1;
        while (at < text.size()) {
            const size_t end = text.find('\n', at);
            if (end == std::string::npos)
                break;  // a line cut short: that job is counted again
            long long id;
            int i, j, k;
            unsigned long long c, w;
            if (sscanf(text.c_str() + at, "job %lld %d %d %d %llu %llu", &id, &i, &j, &k, &c, &w) != 6 ||
                id < 0 || id >= njobs)
                return fail(TRUST_ERR_FORMAT, "journal %s: bad line at byte %zu", path, at);
            jr.have[(size_t)id] = 1;
            jr.cnt[(size_t)id] = c;
            jr.prb[(size_t)id] = w;
            at = end + 1;
        }
        if (at < text.size()) {  // drop the partial line before appending
            if (truncate(path, (off_t)at) != 0)
                return fail(TRUST_ERR_IO, "journal %s: cannot drop a partial line", path);
        }
    }
    jr.f = fopen(path, "ab");
    if (!jr.f)
        return fail(TRUST_ERR_IO, "cannot open journal %s", path);
    if (fresh && (fprintf(jr.f, "%s\n", head) < 0 || fflush(jr.f) != 0))
        return fail(TRUST_ERR_IO, "cannot write journal %s", path);
    return TRUST_OK;
}

static int journal_add(Journal &jr, long long id, int n, unsigned long long c, unsigned long long w)
{
    const int i = (int)(id / ((long long)n * n)), j = (int)(id / n % n), k = (int)(id % n);
    if (fprintf(jr.f, "job %lld %d %d %d %llu %llu\n", id, i, j, k, c, w) < 0 || fflush(jr.f) != 0)
        return fail(TRUST_ERR_IO, "cannot append to the journal");
    jr.have[(size_t)id] = 1;
    jr.cnt[(size_t)id] = c;
    jr.prb[(size_t)id] = w;
    return TRUST_OK;
}

// Streams all jobs of an n x n partition; tiles[i*n+j] is tile (i,j).
// job_counts (nullable, n^3 entries) receives each job's count.  With a
// journal path, jobs already in the journal are skipped and each new job is
// appended as its count lands; max_new_jobs > 0 stops after that many new
// jobs, and *remaining (nullable) receives the jobs still missing.
static int stream_tiles(Ctx *ctx, const trust_csr *tiles, int n, const trust_params &p,
                        int cache_slots, uint64_t *job_counts, trust_result *r,
                        const char *journal = nullptr, long long max_new_jobs = 0,
                        long long *remaining = nullptr)
{
    const long long njobs = (long long)n * n * n;
    bool off64 = (p.flags & TRUST_FLAG_OFFSETS64) != 0;
    int64_t maxd = 0, Vsum = 0, Esum = 0;
    for (int t = 0; t < n * n; t++) {
        off64 = off64 || tiles[t].edges >= kMaxEdges32;
        for (int64_t u = 0; u < tiles[t].vertices; u++)
            maxd = std::max(maxd, tiles[t].beg[u + 1] - tiles[t].beg[u]);
        Esum += tiles[t].edges;
    }
    for (int i = 0; i < n; i++)
        Vsum += tiles[(size_t)i * n].vertices;
    if (maxd > (int64_t)(INT32_MAX / 64))
        return fail(TRUST_ERR_OVERFLOW, "tile degree %lld too large", (long long)maxd);
    if (!(p.flags & TRUST_FLAG_NO_VALIDATE))
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
                for (int k = 0; k < n; k++) {
                    int rc = check_job_tiles(tiles[i * n + j], tiles[i * n + k], tiles[k * n + j]);
                    if (rc)
                        return rc;
                }
    LaunchPlan L;
    int rc = resolve_launch(ctx, p, maxd, Vsum, off64, L);
    if (rc)
        return rc;
    cache_slots = std::max(3, std::min(cache_slots, n * n));
    std::vector<TileSlot> slots((size_t)cache_slots);
    for (auto &s : slots) {
        HIPCHK(hipEventCreateWithFlags(&s.ready, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&s.released, hipEventDisableTiming));
    }
    // Size every device buffer for the largest tile up front: a buffer is
    // never reallocated while a queued copy or count may still use it.
    int64_t maxV = 0, maxE = 0;
    for (int t = 0; t < n * n; t++) {
        maxV = std::max(maxV, tiles[t].vertices);
        maxE = std::max(maxE, tiles[t].edges);
    }
    rc = ctx->staging.ensure(sizeof(int64_t) * (size_t)(maxV + 1));
    if (rc)
        return rc;
    for (auto &s : slots) {
        rc = s.g.beg.ensure((off64 ? 8 : 4) * (size_t)(maxV + 1));
        if (!rc)
            rc = s.g.adj.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(maxE, 1));
        if (rc)
            return rc;
    }
    // uploads go through two pinned buffers, each sized for the largest tile
    PinnedStage stage;
    const size_t stage_bytes = ((sizeof(int32_t) * (size_t)maxE + 255) & ~(size_t)255) +
                               sizeof(int64_t) * (size_t)(maxV + 1);
    const bool pinned_uploads = stage.init(stage_bytes);
    rc = ctx->list.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(maxV, 1));
    if (!rc)
        rc = ctx->block_counts.ensure(sizeof(long long) * kNumTiers *
                                      (size_t)std::max<int64_t>((maxV + kPlanPerBlock - 1) / kPlanPerBlock, 1));
    if (rc)
        return rc;
    struct SlotGuard {
        std::vector<TileSlot> &s;
        ~SlotGuard()
        {
            for (auto &x : s) {
                if (x.ready)
                    (void)hipEventDestroy(x.ready);
                if (x.released)
                    (void)hipEventDestroy(x.released);
            }
        }
    } guard{slots};
    // The short-side walk by global out-degree (default; REFERENCE_WALK keeps
    // the reference's): job (i,j,k) is two counts, the out-walk of tile (i,k)
    // with tables of tile (i,j) probing tile (k,j), and its in-walk with
    // tables of tile (k,j) probing tile (i,j).  Both read only tiles (i,j)
    // and (k,j); the walk tile is needed only to build its walk, once per i.
    bool short_walk = !(p.flags & TRUST_FLAG_REFERENCE_WALK);
    std::vector<int64_t> rows_of((size_t)n), deg_off((size_t)n + 1, 0);
    for (int c = 0; c < n; c++) {
        rows_of[c] = tiles[(size_t)c * n].vertices;
        for (int b = 1; b < n; b++)
            if (tiles[(size_t)c * n + b].vertices != rows_of[c])
                short_walk = false;  // tiles of one class disagree on its rows: no global degrees
        short_walk = short_walk && tile_walk_bins_ok(rows_of[c]);
        deg_off[c + 1] = deg_off[c] + rows_of[c];
    }
    DevBuf degs;
    std::vector<TileWalk> walks(short_walk ? (size_t)n : 0);
    std::vector<int> walk_row(walks.size(), -1);  // the tile row i each walk was built for
    if (short_walk) {
        // global out-degree of every vertex, class by class (sum over its tile row)
        std::vector<int32_t> d((size_t)deg_off[n]);
        for (int c = 0; c < n; c++)
            host_tile_row_degrees(tiles, n, c, rows_of[c], d.data() + deg_off[c]);
        if ((rc = degs.ensure(sizeof(int32_t) * std::max<size_t>(d.size(), 1))))
            return rc;
        HIPCHK(hipMemcpy(degs.p, d.data(), sizeof(int32_t) * d.size(), hipMemcpyHostToDevice));
    }
    Journal jr;
    if (journal && (rc = journal_open(jr, journal, n, tiles, njobs)))
        return rc;
    std::vector<long long> todo;  // the jobs this call counts, in stream order
    for (long long job = 0; job < njobs; job++)
        if (!journal || !jr.have[(size_t)job])
            todo.push_back(job);
    if (max_new_jobs > 0 && (long long)todo.size() > max_new_jobs)
        todo.resize((size_t)max_new_jobs);
    DevBuf ctls;
    const long long nctl = short_walk ? 2 * njobs : njobs;
    const int per_job = short_walk ? 2 : 1;  // control blocks per job
    // journal: each job's control blocks come back to pinned host memory
    // right after its launches, with an event to poll
    TcCtl *hctl = nullptr;
    std::vector<hipEvent_t> cev;
    struct JournalGuard {
        TcCtl *&h;
        std::vector<hipEvent_t> &e;
        hipStream_t s;
        ~JournalGuard()
        {
            if (h || !e.empty())
                (void)hipStreamSynchronize(s);
            for (auto x : e)
                if (x)
                    (void)hipEventDestroy(x);
            if (h)
                (void)hipHostFree(h);
        }
    } journal_guard{hctl, cev, ctx->stream};
    if (journal) {
        HIPCHK(hipHostMalloc((void **)&hctl, sizeof(TcCtl) * (size_t)nctl, hipHostMallocDefault));
        cev.assign((size_t)njobs, nullptr);
        for (long long job : todo)
            HIPCHK(hipEventCreateWithFlags(&cev[(size_t)job], hipEventDisableTiming));
    }
    size_t journaled = 0;  // todo[0, journaled) are in the file
    auto journal_ready = [&](bool wait) -> int {
        while (journaled < todo.size()) {
            const long long job = todo[journaled];
            if (!cev[(size_t)job])
                break;
            hipError_t e = wait ? hipEventSynchronize(cev[(size_t)job]) : hipEventQuery(cev[(size_t)job]);
            if (e == hipErrorNotReady)
                break;
            if (e != hipSuccess)
                return hip_fail(e, "journal event");
            unsigned long long c = 0, w = 0;
            for (int q = 0; q < per_job; q++) {
                c += hctl[(size_t)(per_job * job + q)].count;
                w += hctl[(size_t)(per_job * job + q)].probes;
            }
            if (int rc2 = journal_add(jr, job, n, c, w))
                return rc2;
            journaled++;
        }
        return TRUST_OK;
    };
    rc = ctls.ensure(sizeof(TcCtl) * (size_t)nctl);
    if (rc)
        return rc;

    // per job: events around its counting launches (the count's share of the
    // streamed time); host time inside the tile uploads; the tile walks' builds
    double copy_host_s = 0;
    // and events around each tile walk's build (at most one per tile (i,k)):
    // the build's time on the stream, host readbacks between its passes included
    std::vector<hipEvent_t> jev((size_t)(2 * njobs + 2 * n * n), nullptr);
    hipEvent_t *wev = jev.data() + 2 * njobs;
    int nbuilds = 0;
    struct EventGuard {
        std::vector<hipEvent_t> &e;
        ~EventGuard()
        {
            for (auto x : e)
                if (x)
                    (void)hipEventDestroy(x);
        }
    } jev_guard{jev};
    for (auto &e : jev)
        HIPCHK(hipEventCreate(&e));
    auto job_tiles = [n, short_walk](long long job, int out[3]) {
        int i = (int)(job / ((long long)n * n)), j = (int)(job / n % n), k = (int)(job % n);
        out[0] = i * n + j;  // hash (the in-walk's probe)
        out[1] = short_walk ? -1 : i * n + k;  // walk (the short walk reads its own index)
        out[2] = k * n + j;  // probe (the in-walk's hash)
    };
    // Make tile t resident for job `job`; never evicts tiles of jobs >= pin_from.
    auto acquire = [&](int t, long long job, const int keep[6], int &slot_out) -> int {
        for (int s = 0; s < cache_slots; s++)
            if (slots[s].tile == t) {
                slot_out = s;
                return TRUST_OK;
            }
        int victim = -1;
        for (int s = 0; s < cache_slots; s++) {
            bool pinned = false;
            for (int q = 0; q < 6; q++)
                pinned = pinned || (keep[q] >= 0 && slots[s].tile == keep[q]);
            if (pinned)
                continue;
            if (victim < 0 || slots[s].last_job < slots[victim].last_job)
                victim = s;
        }
        if (victim < 0)
            return fail(TRUST_ERR_ARG, "tile cache of %d slots too small", cache_slots);
        TileSlot &v = slots[victim];
        if (v.last_job >= 0)
            HIPCHK(hipStreamWaitEvent(ctx->copy_stream, v.released, 0));
        const trust_csr &src = tiles[t];
        const auto c0 = std::chrono::steady_clock::now();
        int rc2 = pinned_uploads
                      ? upload_pinned(ctx, stage, v.g, src, off64, ctx->copy_stream)
                      : upload_csr(ctx, v.g, src.beg, src.adj, src.vertices, src.edges, off64,
                                   ctx->copy_stream);
        if (rc2)
            return rc2;
        // host time of the upload: the wait for a free pinned buffer and the
        // parallel memcpy into it (pageable: the runtime staging every byte)
        copy_host_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
        HIPCHK(hipEventRecord(v.ready, ctx->copy_stream));
        v.tile = t;
        v.last_job = job;
        slot_out = victim;
        (void)job;
        return TRUST_OK;
    };

    auto count_view = [&](const CsrViews &v, int64_t rows, TcCtl *ctl) -> int {
        return off64 ? enqueue_count<int64_t>(ctx, v, rows, false, p, L, ctl, kModeCount, nullptr,
                                              nullptr, false, ctx->stream, nullptr, nullptr)
                     : enqueue_count<int32_t>(ctx, v, rows, false, p, L, ctl, kModeCount, nullptr,
                                              nullptr, false, ctx->stream, nullptr, nullptr);
    };
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    for (size_t t = 0; t < todo.size(); t++) {
        const long long job = todo[t];
        const int ji = (int)(job / ((long long)n * n)), jj = (int)(job / n % n), jk = (int)(job % n);
        int cur[3], nxt[3] = {-1, -1, -1};
        job_tiles(job, cur);
        if (t + 1 < todo.size())
            job_tiles(todo[t + 1], nxt);
        if (short_walk && walk_row[(size_t)jk] != ji) {
            // the walk of tile (i,k), built once for the jobs (i, *, k)
            const int wt = ji * n + jk;
            const int keepw[6] = {wt, cur[0], cur[2], -1, -1, -1};
            int ws;
            if ((rc = acquire(wt, job, keepw, ws)))
                return rc;
            slots[ws].last_job = job;
            HIPCHK(hipStreamWaitEvent(ctx->stream, slots[ws].ready, 0));
            const DevCsr &wg = slots[ws].g;
            const int32_t *dg = (const int32_t *)degs.p;
            HIPCHK(hipEventRecord(wev[2 * nbuilds], ctx->stream));
            rc = off64 ? tile_walk_build<int64_t>(ctx, (const int64_t *)wg.beg.p,
                                                  (const int32_t *)wg.adj.p, rows_of[ji],
                                                  rows_of[jk], dg + deg_off[ji], dg + deg_off[jk],
                                                  walks[(size_t)jk])
                       : tile_walk_build<int32_t>(ctx, (const int32_t *)wg.beg.p,
                                                  (const int32_t *)wg.adj.p, rows_of[ji],
                                                  rows_of[jk], dg + deg_off[ji], dg + deg_off[jk],
                                                  walks[(size_t)jk]);
            if (rc)
                return rc;
            HIPCHK(hipEventRecord(wev[2 * nbuilds + 1], ctx->stream));
            nbuilds++;
            HIPCHK(hipEventRecord(slots[ws].released, ctx->stream));
            walk_row[(size_t)jk] = ji;
        }
        const int keep[6] = {cur[0], cur[1], cur[2], -1, -1, -1};
        int sl[3] = {-1, -1, -1};
        for (int q = 0; q < 3; q++) {
            if (cur[q] < 0)
                continue;
            rc = acquire(cur[q], job, keep, sl[q]);
            if (rc)
                return rc;
        }
        for (int q = 0; q < 3; q++) {
            if (sl[q] < 0)
                continue;
            slots[sl[q]].last_job = job;
            HIPCHK(hipStreamWaitEvent(ctx->stream, slots[sl[q]].ready, 0));
        }
        const DevCsr &th = slots[sl[0]].g, &tp = slots[sl[2]].g;
        HIPCHK(hipEventRecord(jev[(size_t)(2 * job)], ctx->stream));
        if (short_walk) {
            const TileWalk &w = walks[(size_t)jk];
            CsrViews vo{th.beg.p, (const int32_t *)th.adj.p, w.out_beg.p,
                        (const int32_t *)w.out_ids.p, tp.beg.p, (const int32_t *)tp.adj.p};
            if ((rc = count_view(vo, rows_of[ji], (TcCtl *)ctls.p + 2 * job)))
                return rc;
            CsrViews vi{tp.beg.p, (const int32_t *)tp.adj.p, w.in_beg.p,
                        (const int32_t *)w.in_ids.p, th.beg.p, (const int32_t *)th.adj.p};
            if ((rc = count_view(vi, rows_of[jk], (TcCtl *)ctls.p + 2 * job + 1)))
                return rc;
        } else {
            const DevCsr &tw = slots[sl[1]].g;
            CsrViews v{th.beg.p, (const int32_t *)th.adj.p, tw.beg.p, (const int32_t *)tw.adj.p,
                       tp.beg.p, (const int32_t *)tp.adj.p};
            if ((rc = count_view(v, tiles[cur[0]].vertices, (TcCtl *)ctls.p + job)))
                return rc;
        }
        (void)jj;
        HIPCHK(hipEventRecord(jev[(size_t)(2 * job + 1)], ctx->stream));
        if (journal) {
            const size_t c0 = (size_t)(per_job * job);
            HIPCHK(hipMemcpyAsync(hctl + c0, (TcCtl *)ctls.p + c0, sizeof(TcCtl) * per_job,
                                  hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipEventRecord(cev[(size_t)job], ctx->stream));
            if ((rc = journal_ready(false)))
                return rc;
        }
        for (int q = 0; q < 3; q++)
            if (sl[q] >= 0)
                HIPCHK(hipEventRecord(slots[sl[q]].released, ctx->stream));
        // Prefetch the next job's tiles while this one counts.
        if (t + 1 < todo.size()) {
            const int keep2[6] = {cur[0], cur[1], cur[2], nxt[0], nxt[1], nxt[2]};
            for (int q = 0; q < 3; q++) {
                int dummy;
                bool fits = cache_slots >= 6;
                if (!fits)
                    break;  // too few slots to prefetch without evicting the running job
                if (nxt[q] < 0)
                    continue;
                rc = acquire(nxt[q], todo[t + 1], keep2, dummy);
                if (rc)
                    return rc;
            }
        }
    }
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    std::vector<TcCtl> h((size_t)nctl);
    HIPCHK(hipMemcpyAsync(h.data(), ctls.p, sizeof(TcCtl) * (size_t)nctl, hipMemcpyDeviceToHost,
                          ctx->stream));
    if (short_walk)  // fold each job's in-walk count into its out-walk count
        for (long long j : todo) {
            TcCtl &a = h[(size_t)(2 * j)];
            const TcCtl &b = h[(size_t)(2 * j + 1)];
            a.count += b.count;
            a.probes += b.probes;
            a.rebuilds += b.rebuilds;
        }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (journal && (rc = journal_ready(true)))
        return rc;
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    double count_s = 0;
    std::vector<char> counted((size_t)njobs, 0);
    for (long long j : todo) {
        float jm = 0;
        HIPCHK(hipEventElapsedTime(&jm, jev[(size_t)(2 * j)], jev[(size_t)(2 * j + 1)]));
        count_s += jm * 1e-3;
        counted[(size_t)j] = 1;
    }
    double build_s = 0;
    for (int b = 0; b < nbuilds; b++) {
        float bm = 0;
        HIPCHK(hipEventElapsedTime(&bm, wev[2 * b], wev[2 * b + 1]));
        build_s += bm * 1e-3;
    }
    uint64_t total = 0, probes = 0, rebuilds = 0;
    long long missing = 0;
    for (long long j = 0; j < njobs; j++) {
        uint64_t c = 0;
        if (counted[(size_t)j]) {  // this call
            const TcCtl &x = h[(size_t)(short_walk ? 2 * j : j)];
            c = x.count;
            probes += x.probes;
            rebuilds += x.rebuilds;
        } else if (journal && jr.have[(size_t)j]) {  // an earlier call
            c = jr.cnt[(size_t)j];
            probes += jr.prb[(size_t)j];
        } else {
            missing++;
        }
        total += c;
        if (job_counts)
            job_counts[j] = c;
    }
    if (remaining)
        *remaining = missing;
    if (r) {
        memset(r, 0, sizeof(*r));
        r->vertices = Vsum;
        r->edges = Esum;
        r->triangles = total;
        r->probes = probes;
        r->seconds = ms * 1e-3;       // the whole stream: copies, walk builds, counts
        r->count_seconds = count_s;   // the jobs' counting launches alone
        r->index_seconds = build_s;   // the tile walks' builds
        r->h2d_seconds = copy_host_s;  // host time inside the tile uploads
        r->gteps = r->seconds > 0 ? (double)Esum / r->seconds / 1e9 : 0;
        r->alg_bytes = alg_bytes(Vsum, Esum, probes, off64);
        r->threads = L.main.threads;
        r->blocks = L.main.blocks;
        r->hub_slots = 1 << L.main.group_lg;
        r->wave_slots = L.main.wave_slots;
        r->linear_rebuilds = rebuilds;
        r->walk_used = short_walk ? 1 : 0;
    }
    return TRUST_OK;
}

}  // namespace trust

using namespace trust;

extern "C" int trust_count_tiles(const trust_csr *tiles, int n, const trust_params *p,
                                 trust_result *r)
{
    if (!tiles || !p || n < 1)
        return fail(TRUST_ERR_ARG, "bad arguments");
    int rc = check_params(*p);
    if (rc)
        return rc;
    Ctx ctx;
    rc = ctx_init(&ctx, p->device);
    if (rc)
        return rc;
    // reserved[0]: device tile-cache slots (0 = 2n+2, enough to prefetch)
    int slots = p->reserved[0] > 0 ? p->reserved[0] : 2 * n + 2;
    return stream_tiles(&ctx, tiles, n, *p, slots, nullptr, r);
}

extern "C" int trust_count_tiles_journal(const trust_csr *tiles, int n, const trust_params *p,
                                         const char *journal, int64_t max_new_jobs,
                                         uint64_t *job_counts, trust_result *r)
{
    if (!tiles || !p || !journal || n < 1)
        return fail(TRUST_ERR_ARG, "bad arguments");
    int rc = check_params(*p);
    if (rc)
        return rc;
    Ctx ctx;
    rc = ctx_init(&ctx, p->device);
    if (rc)
        return rc;
    int slots = p->reserved[0] > 0 ? p->reserved[0] : 2 * n + 2;
    long long remaining = 0;
    rc = stream_tiles(&ctx, tiles, n, *p, slots, job_counts, r, journal, (long long)max_new_jobs,
                      &remaining);
    if (rc)
        return rc;
    return (int)std::min<long long>(remaining, INT32_MAX);
}

extern "C" int trust_triangle_count_partition(int rank, const char *folder, int total_process,
                                              int threads, int blocks, int chunk_size,
                                              int partition_num, trust_arguments *out)
{
    if (!folder || !out || partition_num < 1)
        return fail(TRUST_ERR_ARG, "bad arguments");
    const int n = partition_num;
    const long long n3 = (long long)n * n * n;
    if (total_process < n3 || total_process % n3 != 0)
        return fail(TRUST_ERR_ARG,
                    "total_process %d must be a positive multiple of n^3 = %lld (one job per "
                    "(i,j,k) tile triple)",
                    total_process, n3);
    if (rank < 0 || rank >= total_process)
        return fail(TRUST_ERR_ARG, "rank %d out of range", rank);
    const int i = rank % n, j = rank / n % n, k = rank / n / n % n;
    auto tile_dir = [&](int a, int b) {
        return std::string(folder) + "/partition" + std::to_string(a) + "_" + std::to_string(b);
    };
    HostCsr g[3];
    const std::string dirs[3] = {tile_dir(i, j), tile_dir(i, k), tile_dir(k, j)};
    for (int q = 0; q < 3; q++) {
        const std::string fb = dirs[q] + "/begin.bin", fa = dirs[q] + "/adjacent.bin";
        int64_t sb = file_size(fb), sa = file_size(fa);
        if (sb < 8 || sb % 8 || sa < 0 || sa % 4)
            return fail(TRUST_ERR_IO, "tile %s is missing or malformed", dirs[q].c_str());
        g[q].V = sb / 8 - 1;
        g[q].E = sa / 4;
        g[q].beg.resize((size_t)g[q].V + 1);
        g[q].adj.resize((size_t)g[q].E);
        int rc = read_file(fb, g[q].beg.data(), (size_t)sb);
        if (!rc)
            rc = read_file(fa, g[q].adj.data(), (size_t)sa);
        if (rc)
            return rc;
    }
    trust_csr t[3];
    for (int q = 0; q < 3; q++)
        t[q] = trust_csr{g[q].V, g[q].E, g[q].beg.data(), g[q].adj.data()};
    int rc = check_job_tiles(t[0], t[1], t[2]);
    if (rc)
        return rc;
    trust_params p;
    trust_params_init(&p);
    p.rank = (int)(rank / n3);
    p.total_process = (int)(total_process / n3);
    p.threads = threads > 0 ? threads : 0;
    p.blocks = blocks > 0 ? blocks : 0;
    p.chunk = chunk_size > 0 ? chunk_size : 0;
    Ctx ctx;
    rc = ctx_init(&ctx, 0);
    if (rc)
        return rc;
    bool off64 = false;
    int64_t maxd = 0;
    for (int q = 0; q < 3; q++)
        off64 = off64 || t[q].edges >= kMaxEdges32;
    for (int64_t u = 0; u < t[0].vertices; u++)
        maxd = std::max(maxd, t[0].beg[u + 1] - t[0].beg[u]);
    DevCsr d[3];
    for (int q = 0; q < 3; q++) {
        rc = upload_csr(&ctx, d[q], t[q].beg, t[q].adj, t[q].vertices, t[q].edges, off64,
                        ctx.stream);
        if (rc)
            return rc;
        HIPCHK(hipStreamSynchronize(ctx.stream));
    }
    LaunchPlan L;
    rc = resolve_launch(&ctx, p, maxd, t[0].vertices, off64, L);
    if (rc)
        return rc;
    rc = ctx.ctl.ensure(sizeof(TcCtl));
    if (rc)
        return rc;
    CsrViews v{d[0].beg.p, (const int32_t *)d[0].adj.p, d[1].beg.p, (const int32_t *)d[1].adj.p,
               d[2].beg.p, (const int32_t *)d[2].adj.p};
    HIPCHK(hipEventRecord(ctx.ev[0], ctx.stream));
    rc = off64 ? enqueue_count<int64_t>(&ctx, v, t[0].vertices, false, p, L, (TcCtl *)ctx.ctl.p,
                                        kModeCount, nullptr, nullptr, false, ctx.stream, ctx.ev[1],
                                        ctx.ev[2])
               : enqueue_count<int32_t>(&ctx, v, t[0].vertices, false, p, L, (TcCtl *)ctx.ctl.p,
                                        kModeCount, nullptr, nullptr, false, ctx.stream, ctx.ev[1],
                                        ctx.ev[2]);
    if (rc)
        return rc;
    TcCtl h;
    HIPCHK(hipMemcpyAsync(&h, ctx.ctl.p, sizeof(TcCtl), hipMemcpyDeviceToHost, ctx.stream));
    HIPCHK(hipStreamSynchronize(ctx.stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx.ev[0], ctx.ev[2]));
    for (int q = 0; q < 3; q++) {
        d[q].beg.release();
        d[q].adj.release();
    }
    out->edge_count = t[2].edges;
    out->count = h.count;
    out->time = ms * 1e-3;
    out->vertices = t[2].vertices - 1;  // With-graph-partition/TC.cu:1152
    out->degree = t[2].vertices > 0 ? t[2].edges / t[2].vertices : 0;
    return TRUST_OK;
}
