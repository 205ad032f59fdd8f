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
                          ctx->side_beg.p,   ctx->side_adj.p,   ctx->side_row.p,
                          ctx->side_deg.p,   ctx->walk_acc.p};
    for (const void *q : ptrs)
        k.push_back((uintptr_t)q);
    k.push_back((uintptr_t)ctx->g.V);
    k.push_back((uintptr_t)ctx->g.E);
    k.push_back((uintptr_t)ctx->batch_ngraphs);
    k.push_back((uintptr_t)ctx->mlp_ready);
    return k;
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
    return TRUST_OK;
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
    HIPCHK(hipMemcpyAsync(&h, ctx->ctl.p, sizeof(TcCtl), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(out.data(), ctx->batch_out.p, sizeof(uint64_t) * out.size(),
                          hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
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
    finish_resident(ctx, R, h, t_plan * 1e-3, t_count * 1e-3, r);
    return TRUST_OK;
}
