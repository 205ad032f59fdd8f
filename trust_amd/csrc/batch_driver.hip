// Batches of graphs on one device: load, one persistent count over the
// whole batch with per-graph attribution, and replay of that count as a
// captured hipGraph.
//
// The reference counts one graph per process (Without-graph-partition/
// mpi_main.cpp:72-96, one Triangle_count call per run).  Many small graphs
// that way cost a process, a load and several launches each; here the batch
// is one concatenated CSR and one planned count, so graphs share the
// persistent workgroups and small graphs fill the gaps of large ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "engine.h"

namespace trust {

void drop_captured(Ctx *ctx)
{
    if (ctx->graph_exec)
        (void)hipGraphExecDestroy(ctx->graph_exec);
    ctx->graph_exec = nullptr;
    ctx->graph_key.clear();
}

// Everything a captured count bakes in: the parameters and every device
// address its launches use.  Buffers only grow, so an unchanged address
// still holds at least the captured size.
static std::vector<uintptr_t> capture_key(const Ctx *ctx, const trust_params &p)
{
    std::vector<uintptr_t> k((sizeof(p) + sizeof(uintptr_t) - 1) / sizeof(uintptr_t), 0);
    memcpy(k.data(), &p, sizeof(p));
    const void *ptrs[] = {ctx->g.beg.p,      ctx->g.adj.p,      ctx->list.p,
                          ctx->block_counts.p, ctx->ctl.p,       ctx->spill.p,
                          ctx->feat.p,       ctx->cost.p,       ctx->tiers.p,
                          ctx->mlp_params.p, ctx->batch_vbase.p, ctx->batch_page.p,
                          ctx->batch_out.p,  ctx->side_cnt.p,   ctx->side_bsum.p,
                          ctx->side_beg.p,   ctx->side_ent.p,   ctx->side_row.p,
                          ctx->walk_acc.p,   ctx->seg_beg.p,
                          ctx->seg_vbase.p,  ctx->seg_page.p};
    for (const void *q : ptrs)
        k.push_back((uintptr_t)q);
    k.push_back((uintptr_t)ctx->g.V);
    k.push_back((uintptr_t)ctx->g.E);
    k.push_back((uintptr_t)ctx->batch_ngraphs);
    k.push_back((uintptr_t)ctx->mlp_ready);
    k.push_back((uintptr_t)ctx->segs.size());
    return k;
}

// out[i] = beg[i] - e0 for i < n: one segment's offsets, relative to its
// first edge (every value is below 2^30).
__global__ void seg_rebase_kernel(const long long *__restrict__ beg, long long n, long long e0,
                                  int *__restrict__ out)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (int)(beg[i] - e0);
}

int build_segments(Ctx *ctx, int64_t max_edges)
{
    drop_captured(ctx);
    ctx->segs.clear();
    const DevCsr &g = ctx->g;
    const int64_t ng = ctx->batch_ngraphs;
    if (max_edges <= 0 || ng < 1 || !g.off64 || g.V < 1)
        return TRUST_OK;
    max_edges = std::min(max_edges, kSegMaxEdges);
    const std::vector<int64_t> &vb = ctx->batch_vbase_h, &eb = ctx->batch_ebase_h;
    if ((int64_t)vb.size() != ng + 1 || (int64_t)eb.size() != ng + 1)
        return fail(TRUST_ERR_ARG, "batch bases missing");
    std::vector<BatchSeg> segs;
    int64_t beg_off = 0, vbase_off = 0, page_off = 0;
    for (int64_t gi = 0; gi < ng;) {
        if (eb[gi + 1] - eb[gi] > max_edges)
            return TRUST_OK;  // a graph that needs int64 offsets: the batch stays whole
        BatchSeg q;
        q.g0 = gi;
        while (gi < ng && eb[gi + 1] - eb[q.g0] <= max_edges)
            gi++;
        q.g1 = gi;
        q.v0 = vb[q.g0];
        q.v1 = vb[gi];
        q.e0 = eb[q.g0];
        q.e1 = eb[gi];
        if (q.v1 == q.v0)
            continue;  // graphs without vertices: nothing to count or attribute
        // segments cover [0, V) in order, so beg_off = v0 + (segments before)
        // and the globally indexed view seg_beg + beg_off - v0 starts in the buffer
        q.beg_off = beg_off;
        q.vbase_off = vbase_off;
        q.page_off = page_off;
        if (q.beg_off < q.v0)
            return fail(TRUST_ERR_ARG, "segment layout");
        beg_off += q.v1 - q.v0 + 1;
        vbase_off += q.g1 - q.g0 + 1;
        page_off += ((q.v1 - q.v0 - 1) >> kGraphPageShift) + 2;
        segs.push_back(q);
    }
    // per segment: local graph bases and the page table of its vertices
    std::vector<int64_t> vloc((size_t)vbase_off);
    std::vector<int32_t> page((size_t)page_off);
    for (const BatchSeg &q : segs) {
        int64_t *vl = vloc.data() + q.vbase_off;
        const int64_t n = q.g1 - q.g0, Vs = q.v1 - q.v0;
        for (int64_t k = 0; k <= n; k++)
            vl[k] = vb[q.g0 + k] - q.v0;
        const int64_t np = ((Vs - 1) >> kGraphPageShift) + 2;
        for (int64_t t = 0; t < np; t++) {
            const int64_t x = std::min<int64_t>(t << kGraphPageShift, Vs - 1);
            page[(size_t)(q.page_off + t)] = (int32_t)(std::upper_bound(vl, vl + n, x) - vl - 1);
        }
    }
    int rc;
    if ((rc = ctx->seg_beg.ensure(sizeof(int32_t) * (size_t)beg_off)) ||
        (rc = ctx->seg_vbase.ensure(sizeof(int64_t) * (size_t)vbase_off)) ||
        (rc = ctx->seg_page.ensure(sizeof(int32_t) * (size_t)page_off)))
        return rc;
    HIPCHK(hipMemcpyAsync(ctx->seg_vbase.p, vloc.data(), sizeof(int64_t) * vloc.size(),
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->seg_page.p, page.data(), sizeof(int32_t) * page.size(),
                          hipMemcpyHostToDevice, ctx->stream));
    for (const BatchSeg &q : segs) {
        const long long n = q.v1 - q.v0 + 1;
        const int blocks = (int)std::min<long long>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(seg_rebase_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           (const long long *)g.beg.p + q.v0, n, (long long)q.e0,
                           (int *)ctx->seg_beg.p + q.beg_off);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->segs = std::move(segs);
    return TRUST_OK;
}

// Plans every segment into its own part of the list, then counts them in
// turn: each segment is a three-view count whose hash and walk rows are
// indexed by segment-local vertices and whose probe rows by the global ids
// stored in adj (seg_beg + beg_off - v0), all over adj + e0.
int enqueue_segments(Ctx *ctx, const trust_params &p, int mode, const ResidentRun &R, bool timed)
{
    const DevCsr &g = ctx->g;
    const int64_t ng = ctx->batch_ngraphs;
    int rc = ctx->list.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(g.V, 1));
    if (rc)
        return rc;
    unsigned long long *out = (unsigned long long *)ctx->batch_out.p;
    const int32_t *adj = (const int32_t *)g.adj.p;
    const int32_t *sb = (const int32_t *)ctx->seg_beg.p;
    TcCtl *ctl = (TcCtl *)ctx->ctl.p;
    if (timed)
        HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    for (int pass = 0; pass < 2; pass++) {  // 0: plans, 1: counts
        for (size_t k = 0; k < ctx->segs.size(); k++) {
            const BatchSeg &q = ctx->segs[k];
            const int32_t *loc = sb + q.beg_off;
            const int32_t *glo = sb + (q.beg_off - q.v0);
            CsrViews v{loc, adj + q.e0, loc, adj + q.e0, glo, adj + q.e0};
            SegView sv;
            sv.list = (const int32_t *)ctx->list.p + q.v0;
            sv.graph_vbase = (const long long *)ctx->seg_vbase.p + q.vbase_off;
            sv.graph_page = (const int *)ctx->seg_page.p + q.page_off;
            sv.ngraphs = (int)(q.g1 - q.g0);
            sv.graph_counts = out + q.g0;
            sv.graph_done = out + ng + q.g0;
            if (pass == 0 && R.learned)
                rc = enqueue_learned_plan_segment(ctx, p, R.L, ctl + k, ctx->stream, v, q.v1 - q.v0,
                                                  q.v0);
            else
                rc = enqueue_count<int32_t>(ctx, v, q.v1 - q.v0, true, p, R.L, ctl + k, mode,
                                            nullptr, nullptr, pass == 1, ctx->stream, nullptr,
                                            nullptr, &sv, pass == 0);
            if (rc)
                return rc;
        }
        if (timed)
            HIPCHK(hipEventRecord(ctx->ev[pass + 1], ctx->stream));
    }
    return TRUST_OK;
}

int read_ctl(Ctx *ctx, const ResidentRun &R, TcCtl &h)
{
    const size_t n = R.segmented ? ctx->segs.size() : 1;
    std::vector<TcCtl> hs(n);
    HIPCHK(hipMemcpyAsync(hs.data(), ctx->ctl.p, sizeof(TcCtl) * n, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    h = hs[0];
    for (size_t k = 1; k < n; k++) {  // list_begin is summed too: tier sizes are its differences
        h.count += hs[k].count;
        h.probes += hs[k].probes;
        h.rebuilds += hs[k].rebuilds;
        for (int t = 0; t < 8; t++)
            h.list_begin[t] += hs[k].list_begin[t];
    }
    return TRUST_OK;
}

static int capture_batch(Ctx *ctx, const trust_params &p, const ResidentRun &R)
{
    drop_captured(ctx);
    // one eager pass first: it sizes every buffer, so that nothing is
    // allocated while the stream is captured
    int rc = enqueue_resident(ctx, p, kModeBatch, R, false);
    if (rc)
        return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const std::vector<uintptr_t> key = capture_key(ctx, p);
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_resident(ctx, p, kModeBatch, R, false);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc) {
        if (graph)
            (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess)
        return hip_fail(e, "hipStreamEndCapture");
    if (capture_key(ctx, p) != key) {
        (void)hipGraphDestroy(graph);
        return fail(TRUST_ERR_HIP, "a buffer moved while the batch count was captured");
    }
    e = hipGraphInstantiate(&ctx->graph_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        ctx->graph_exec = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    ctx->graph_key = key;
    return TRUST_OK;
}

}  // namespace trust

using namespace trust;

extern "C" int trust_ctx_load_batch(trust_ctx *c, const int64_t *beg, const int32_t *adj,
                                    int64_t vertices, int64_t edges, const int64_t *vbase,
                                    int64_t ngraphs, uint32_t flags)
{
    clear_error();
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    if (!ctx || !beg || (!adj && edges > 0) || !vbase)
        return fail(TRUST_ERR_ARG, "null argument");
    if (vertices < 0 || edges < 0)
        return fail(TRUST_ERR_ARG, "negative size");
    int rc;
    if (!(flags & TRUST_FLAG_NO_VALIDATE)) {
        if ((rc = trust_csr_validate(beg, adj, vertices, edges)))
            return rc;
    }
    if ((rc = batch_validate(beg, adj, vertices, vbase, ngraphs)))
        return rc;
    if ((rc = load_ctx_csr(ctx, beg, adj, vertices, edges, flags | TRUST_FLAG_NO_VALIDATE)))
        return rc;
    if (ngraphs > INT32_MAX - 1)
        return fail(TRUST_ERR_OVERFLOW, "%lld graphs in one batch", (long long)ngraphs);
    // page table: page[p] = largest g < ngraphs with vbase[g] <= min(p << shift, V - 1)
    const int64_t npages = vertices > 0 ? ((vertices - 1) >> kGraphPageShift) + 2 : 1;
    std::vector<int32_t> page((size_t)npages, 0);
    for (int64_t q = 0; q < npages && vertices > 0; q++) {
        const int64_t x = std::min<int64_t>(q << kGraphPageShift, vertices - 1);
        page[q] = (int32_t)(std::upper_bound(vbase, vbase + ngraphs, x) - vbase - 1);
    }
    if ((rc = ctx->batch_vbase.ensure(sizeof(int64_t) * (size_t)(ngraphs + 1))) ||
        (rc = ctx->batch_page.ensure(sizeof(int32_t) * (size_t)npages)) ||
        (rc = ctx->batch_out.ensure(sizeof(uint64_t) * (size_t)(2 * ngraphs + 1))))
        return rc;
    HIPCHK(hipMemcpyAsync(ctx->batch_vbase.p, vbase, sizeof(int64_t) * (size_t)(ngraphs + 1),
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->batch_page.p, page.data(), sizeof(int32_t) * (size_t)npages,
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
    if (khz <= 0)
        return fail(TRUST_ERR_HIP, "device reports no wall-clock rate");
    ctx->wall_hz = 1e3 * khz;
    ctx->batch_ngraphs = ngraphs;
    ctx->batch_vbase_h.assign(vbase, vbase + ngraphs + 1);
    ctx->batch_ebase_h.resize((size_t)ngraphs + 1);
    for (int64_t g = 0; g <= ngraphs; g++)
        ctx->batch_ebase_h[(size_t)g] = beg[vbase[g]];
    // too many edges for int32 offsets: count it as int32 segments
    return ctx->g.off64 && !(flags & TRUST_FLAG_OFFSETS64) ? build_segments(ctx, kSegMaxEdges)
                                                           : TRUST_OK;
}

extern "C" int trust_ctx_batch_segments(trust_ctx *c, int64_t max_edges, int64_t *segments)
{
    clear_error();
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    if (!ctx)
        return fail(TRUST_ERR_ARG, "null argument");
    if (ctx->batch_ngraphs < 1)
        return fail(TRUST_ERR_ARG, "context holds no batch");
    if (max_edges < 0)
        return fail(TRUST_ERR_ARG, "negative max_edges");
    HIPCHK(hipSetDevice(ctx->device));
    const int rc = build_segments(ctx, max_edges);
    if (segments)
        *segments = (int64_t)ctx->segs.size();
    return rc;
}

extern "C" int trust_ctx_count_batch(trust_ctx *c, const trust_params *p, uint64_t *counts,
                                     double *done_seconds, trust_result *r)
{
    clear_error();
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    if (!ctx || !p)
        return fail(TRUST_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(ctx->device));
    ResidentRun R;
    int rc = prepare_resident(ctx, *p, kModeBatch, R);
    if (rc)
        return rc;
    float t_plan = 0, t_count = 0;
    if (p->flags & TRUST_FLAG_HIP_GRAPH) {
        if (!ctx->graph_exec || capture_key(ctx, *p) != ctx->graph_key) {
            if ((rc = capture_batch(ctx, *p, R)))
                return rc;
        }
        HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
        HIPCHK(hipGraphLaunch(ctx->graph_exec, ctx->stream));
        HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));
    } else {
        if ((rc = enqueue_resident(ctx, *p, kModeBatch, R, true)))
            return rc;
    }
    const int64_t ng = ctx->batch_ngraphs;
    std::vector<uint64_t> out((size_t)(2 * ng + 1));
    TcCtl h;
    HIPCHK(hipMemcpyAsync(out.data(), ctx->batch_out.p, sizeof(uint64_t) * out.size(),
                          hipMemcpyDeviceToHost, ctx->stream));
    if ((rc = read_ctl(ctx, R, h)))
        return rc;
    h.count = 0;  // the batch mode adds every triangle to its graph only
    for (int64_t g = 0; g < ng; g++)
        h.count += out[(size_t)g];
    if (p->flags & TRUST_FLAG_HIP_GRAPH) {
        HIPCHK(hipEventElapsedTime(&t_count, ctx->ev[1], ctx->ev[2]));
    } else {
        HIPCHK(hipEventElapsedTime(&t_plan, ctx->ev[0], ctx->ev[1]));
        HIPCHK(hipEventElapsedTime(&t_count, ctx->ev[1], ctx->ev[2]));
    }
    const uint64_t t0 = out[2 * ng];
    for (int64_t g = 0; g < ng; g++) {
        if (counts)
            counts[g] = out[g];
        if (done_seconds) {
            const uint64_t d = out[ng + g];
            done_seconds[g] = d > t0 ? (double)(d - t0) / ctx->wall_hz : 0.0;
        }
    }
    return finish_resident(ctx, R, h, t_plan * 1e-3, t_count * 1e-3, r);
}
