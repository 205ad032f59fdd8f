// Host engine of libtrust_amd.so: device contexts, launch-shape choice,
// planning + counting launches, and the C-ABI counting entry points.
//
// Replaces the host half of the reference (Triangle_count,
// Without-graph-partition/TC.cu:602-731 and With-graph-partition/TC.cu:976-1157):
// the reference loads, allocates, copies, launches and frees inside every
// call and exits on any CUDA error; here a context keeps the graph and the
// scratch buffers resident, every failure is an error code, and times come
// from hipEvents on the launch stream.
#include <hip/hip_runtime.h>

#include <atomic>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>

#include "engine.h"

namespace trust {

int hip_fail(hipError_t e, const char *what)
{
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return fail(TRUST_ERR_OOM, "%s: %s", what, hipGetErrorString(e));
    return fail(TRUST_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// Device bytes held by the library's buffers (all contexts and devices of
// the process) and the cap trust_set_device_memory_cap puts on them (0: none).
static std::atomic<unsigned long long> g_dev_held{0};
static std::atomic<unsigned long long> g_dev_cap{0};

int DevBuf::ensure(size_t n)
{
    if (n <= bytes)
        return TRUST_OK;
    release();
    if (n == 0)
        return TRUST_OK;
    const unsigned long long held = g_dev_held.fetch_add(n) + n;
    const unsigned long long cap = g_dev_cap.load();
    if (cap && held > cap) {
        g_dev_held.fetch_sub(n);
        return fail(TRUST_ERR_OOM,
                    "device allocation of %zu bytes would make the library hold %llu bytes, above "
                    "its cap of %llu (trust_set_device_memory_cap)",
                    n, held, cap);
    }
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        g_dev_held.fetch_sub(n);
        p = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    bytes = n;
    return TRUST_OK;
}

void DevBuf::release()
{
    if (p) {
        (void)hipFree(p);
        g_dev_held.fetch_sub(bytes);
    }
    p = nullptr;
    bytes = 0;
}

int upload_csr(Ctx *ctx, DevCsr &g, const int64_t *beg, const int32_t *adj, int64_t V, int64_t E,
               bool force64, hipStream_t s)
{
    g.V = V;
    g.E = E;
    g.off64 = force64 || E >= kMaxEdges32;
    g.max_deg = host_max_degree(beg, V);
    int rc = g.adj.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(E, 1));
    if (rc)
        return rc;
    HIPCHK(hipMemcpyAsync(g.adj.p, adj, sizeof(int32_t) * (size_t)E, hipMemcpyHostToDevice, s));
    if (g.off64) {
        rc = g.beg.ensure(sizeof(int64_t) * (size_t)(V + 1));
        if (rc)
            return rc;
        HIPCHK(hipMemcpyAsync(g.beg.p, beg, sizeof(int64_t) * (size_t)(V + 1),
                              hipMemcpyHostToDevice, s));
    } else {
        rc = g.beg.ensure(sizeof(int32_t) * (size_t)(V + 1));
        if (rc)
            return rc;
        rc = ctx->staging.ensure(sizeof(int64_t) * (size_t)(V + 1));
        if (rc)
            return rc;
        HIPCHK(hipMemcpyAsync(ctx->staging.p, beg, sizeof(int64_t) * (size_t)(V + 1),
                              hipMemcpyHostToDevice, s));
        HIPCHK(launch_offsets_to_i32((const int64_t *)ctx->staging.p, (int32_t *)g.beg.p, V + 1, s));
    }
    return TRUST_OK;
}

static bool is_pow2(long long x) { return x > 0 && (x & (x - 1)) == 0; }
static int floor_lg(long long x)
{
    int lg = 0;
    while ((2ll << lg) <= x)
        lg++;
    return lg;
}

// LDS words (pow2) a workgroup of `threads` may use for its table without
// lowering occupancy below 8 waves per SIMD: 160 KB shared by 32 waves.
static long long fair_table_words(long long lds_bytes, int threads)
{
    const int nwaves = threads / kWave;
    const long long share = lds_bytes * nwaves / 32 / 4 - lds_extra_words(threads);
    return share > 0 ? 1ll << floor_lg(share) : 0;
}

static int shape_occupancy(Ctx *ctx, LaunchShape &s, bool off64, int blocks_req, int mode,
                           bool big)
{
    int per_cu = 0;
    hipError_t e = off64 ? count_occupancy<int64_t>(s.threads, s.lds_bytes, mode, big, &per_cu)
                         : count_occupancy<int32_t>(s.threads, s.lds_bytes, mode, big, &per_cu);
    if (e != hipSuccess)
        return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (per_cu < 1)
        return fail(TRUST_ERR_ARG, "launch of %d threads with %zu LDS bytes cannot run",
                    s.threads, s.lds_bytes);
    s.blocks = blocks_req > 0 ? blocks_req : per_cu * ctx->num_cus;
    return TRUST_OK;
}

int resolve_launch(Ctx *ctx, const trust_params &p, int64_t max_deg_hash, int64_t V,
                   bool off64, LaunchPlan &L, int mode)
{
    L = LaunchPlan();
    LaunchShape &m = L.main;
    m.threads = p.threads > 0 ? p.threads : 256;
    if (m.threads % kWave != 0 || m.threads > 1024)
        return fail(TRUST_ERR_ARG, "threads=%d: must be a multiple of 64 and at most 1024",
                    m.threads);
    const int nwaves = m.threads / kWave;
    m.wave_slots = p.wave_slots > 0 ? p.wave_slots : 512;
    if (!is_pow2(m.wave_slots) || m.wave_slots < 64)
        return fail(TRUST_ERR_ARG, "wave_slots=%d: must be a power of two >= 64", m.wave_slots);
    m.wave_lg = floor_lg(m.wave_slots);
    const long long lds_max = (long long)ctx->lds_max;
    const long long budget = p.lds_budget > 0 ? std::min<long long>(p.lds_budget, lds_max) : lds_max;
    const long long wave_words = (long long)nwaves * m.wave_slots;
    if (budget / 4 - lds_extra_words(m.threads) < wave_words)
        return fail(TRUST_ERR_ARG, "LDS budget %lld bytes cannot hold %d wave tables of %d slots",
                    budget, nwaves, m.wave_slots);
    // main launch: group table as large as the fair LDS share allows (and
    // no larger than the biggest 4d table needs)
    const long long need4 = 1ll << host_table_lg(2 * std::max<int64_t>(max_deg_hash, 1));
    long long main_words = std::min(need4, p.lds_budget > 0
                                               ? 1ll << floor_lg(budget / 4 - lds_extra_words(m.threads))
                                               : fair_table_words(lds_max, m.threads));
    m.table_words = (int)std::max(wave_words, main_words);
    m.group_lg = floor_lg(m.table_words);
    m.lds_bytes = (size_t)(m.table_words + lds_extra_words(m.threads)) * 4;
    int rc = shape_occupancy(ctx, m, off64, p.blocks, mode, false);
    if (rc)
        return rc;
    if (p.chunk > 0) {
        L.chunk_group = L.chunk_wave = p.chunk;
    } else {
        // about 16 grabs per wave, at most 32 vertices each (large chunks of
        // a skewed list leave long chunks for the end)
        L.chunk_group = 1;
        const long long waves = (long long)m.blocks * nwaves;
        L.chunk_wave = (int)std::max<long long>(1, std::min<long long>(32, V / (waves * 16 + 1)));
    }
    // large-LDS launch for tables the main launch cannot hold, plus the
    // HBM-table tier for tables no LDS can hold
    const int need_lg = host_table_lg(std::max<int64_t>(max_deg_hash, 1));
    const bool force_spill = (p.flags & TRUST_FLAG_FORCE_SPILL) != 0;
    if (need_lg > m.group_lg || force_spill) {
        L.has_big = true;
        LaunchShape &b = L.big;
        b.threads = 1024;
        b.wave_slots = m.wave_slots;
        b.wave_lg = m.wave_lg;
        const long long extra = lds_extra_words(b.threads);
        const long long cap_half = 1ll << floor_lg(std::max<long long>(budget / 2 / 4 - extra, 64));
        const long long cap_full = 1ll << floor_lg(std::max<long long>(budget / 4 - extra, 64));
        // two workgroups per CU when the 2d table fits half the LDS
        long long words = (1ll << need_lg) <= cap_half ? std::min(cap_half, need4) : cap_full;
        words = std::min(words, std::max<long long>(need4, 64));
        b.table_words = (int)words;
        b.group_lg = floor_lg(b.table_words);
        b.lds_bytes = (size_t)(b.table_words + extra) * 4;
        if (b.lds_bytes > (size_t)budget) {  // budget too small for a 1024-thread workgroup
            b.threads = 64;
            b.table_words = 64;
            b.group_lg = 6;
            b.lds_bytes = (size_t)(64 + lds_extra_words(64)) * 4;
        }
        rc = shape_occupancy(ctx, b, off64, 0, mode, true);
        if (rc)
            return rc;
        if (need_lg > b.group_lg || force_spill) {
            L.spill_slots = 1ll << (need_lg + 1);  // 4d slots
            const long long spill_budget = 2ll << 30;  // bytes of HBM for spill tables
            long long groups = spill_budget / (L.spill_slots * 4);
            L.spill_groups = (int)std::max<long long>(1, std::min<long long>(groups, b.blocks));
        }
    }
    return TRUST_OK;
}

template <typename OffT>
static void fill_args(TcArgs<OffT> &a, Ctx *ctx, const CsrViews &views, const trust_params &p,
                      const LaunchPlan &L, const LaunchShape &s, TcCtl *ctl_dev,
                      unsigned long long *instr_cycles, unsigned long long *instr_probes,
                      const SegView *seg)
{
    memset(&a, 0, sizeof(a));
    a.begH = (const OffT *)views.begH;
    a.adjH = views.adjH;
    a.begW = (const OffT *)views.begW;
    a.endW = (const OffT *)views.begW + 1;
    a.adjW = views.adjW;
    a.begP = (const OffT *)views.begP;
    a.adjP = views.adjP;
    a.list = (const int32_t *)ctx->list.p;
    a.spill_mem = (int32_t *)ctx->spill.p;
    a.spill_slots = L.spill_slots;
    a.spill_groups = L.spill_groups;
    a.table_words = s.table_words;
    a.group_lg = s.group_lg;
    a.wave_slots = s.wave_slots;
    a.wave_lg = s.wave_lg;
    a.chunk_group = L.chunk_group;
    a.chunk_wave = L.chunk_wave;
    a.rank = p.rank;
    a.nranks = p.total_process;
    a.force_linear = (p.flags & TRUST_FLAG_LINEAR_PROBING) ? 1
                     : (p.flags & TRUST_FLAG_REBUILD_LINEAR) ? 2 : 0;
    a.ctl = ctl_dev;
    a.instr_cycles = instr_cycles;
    a.instr_probes = instr_probes;
    a.entW = (const WalkEntry<OffT> *)views.entW;
    if (ctx->batch_ngraphs > 0) {  // read by the batch mode only
        const int64_t ng = ctx->batch_ngraphs;
        a.graph_vbase = (const long long *)ctx->batch_vbase.p;
        a.graph_page = (const int *)ctx->batch_page.p;
        a.graph_page_shift = kGraphPageShift;
        a.ngraphs = (int)ng;
        a.graph_counts = (unsigned long long *)ctx->batch_out.p;
        a.graph_done = a.graph_counts + ng;
        a.graph_t0 = a.graph_counts + 2 * ng;
    }
    if (seg) {  // one segment of a batch: its own list and graph tables, the shared t0
        a.list = seg->list;
        a.graph_vbase = seg->graph_vbase;
        a.graph_page = seg->graph_page;
        a.ngraphs = seg->ngraphs;
        a.graph_counts = seg->graph_counts;
        a.graph_done = seg->graph_done;
    }
}

template <typename OffT>
int enqueue_count(Ctx *ctx, const CsrViews &views, int64_t V, bool same_csr,
                  const trust_params &p, const LaunchPlan &L, TcCtl *ctl_dev, int mode,
                  unsigned long long *instr_cycles, unsigned long long *instr_probes,
                  bool preplanned, hipStream_t s, hipEvent_t ev_plan, hipEvent_t ev_count,
                  const SegView *seg, bool plan_only, bool with_core)
{
    int rc = ctx->list.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(V, 1));
    if (rc)
        return rc;
    const int per_block = kPlanPerBlock;
    const int nblocks = (int)((V + per_block - 1) / per_block);
    rc = ctx->block_counts.ensure(sizeof(long long) * (size_t)kNumTiers * std::max(nblocks, 1));
    if (rc)
        return rc;
    if (L.spill_groups > 0) {
        rc = ctx->spill.ensure(sizeof(int32_t) * (size_t)L.spill_slots * L.spill_groups);
        if (rc)
            return rc;
    }
    if (!preplanned) {
        HIPCHK(launch_fill_u32((uint32_t *)ctl_dev, 0u, sizeof(TcCtl) / 4, s));
        PlanArgs<OffT> pa;
        memset(&pa, 0, sizeof(pa));
        pa.begH = (const OffT *)views.begH;
        pa.begW = (const OffT *)views.begW;
        pa.endW = (const OffT *)views.begW + 1;
        pa.V = V;
        pa.same_csr = same_csr ? 1 : 0;
        pa.hub_threshold = p.hub_threshold;
        pa.wave_lg = L.main.wave_lg;
        pa.hub_lg = L.main.group_lg;
        pa.big_lg = L.has_big ? L.big.group_lg : 0;
        pa.force_spill = (p.flags & TRUST_FLAG_FORCE_SPILL) ? 1 : 0;
        pa.wave_only = (p.flags & TRUST_FLAG_WAVE_ONLY) ? 1 : 0;
        pa.tier_override = nullptr;
        pa.list = seg ? (int32_t *)seg->list : (int32_t *)ctx->list.p;
        pa.block_counts = (long long *)ctx->block_counts.p;
        pa.per_block = per_block;
        pa.ctl = ctl_dev;
        HIPCHK(launch_plan<OffT>(pa, nblocks, s));
    }
    if (ev_plan)
        HIPCHK(hipEventRecord(ev_plan, s));
    if (plan_only)
        return TRUST_OK;
    // the dense core first: its bit rows have the L2s to themselves
    if (with_core) {
        if (ev_plan)
            HIPCHK(hipEventRecord(ctx->ev_core[0], s));
        if ((rc = core_enqueue(ctx, ctl_dev, s)))
            return rc;
        if (ev_plan)
            HIPCHK(hipEventRecord(ctx->ev_core[1], s));
    }

    // The large-LDS launch (big + spill tiers) and the main launch (hub +
    // wave tiers) share nothing but the result counters, so the big one runs
    // on the auxiliary stream and the main launch fills the CUs it frees.
    // Under stream capture both stay on s.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(s, &cap));
    const bool fork = L.has_big && cap == hipStreamCaptureStatusNone && ctx->aux_stream;
    TcArgs<OffT> a;
    if (L.has_big) {
        fill_args(a, ctx, views, p, L, L.big, ctl_dev, instr_cycles, instr_probes, seg);
        a.spill_tier = L.spill_groups > 0 ? kTierSpill : -1;
        a.group_tier = kTierBig;
        a.wave_tier = -1;
        hipStream_t sb = s;
        if (fork) {
            HIPCHK(hipEventRecord(ctx->ev_fork, s));
            HIPCHK(hipStreamWaitEvent(ctx->aux_stream, ctx->ev_fork, 0));
            sb = ctx->aux_stream;
        }
        HIPCHK(launch_count<OffT>(a, L.big.blocks, L.big.threads, L.big.lds_bytes, mode, true, sb));
        if (fork)
            HIPCHK(hipEventRecord(ctx->ev_join, ctx->aux_stream));
    }
    fill_args(a, ctx, views, p, L, L.main, ctl_dev, instr_cycles, instr_probes, seg);
    a.spill_tier = -1;
    a.group_tier = kTierHub;
    a.wave_tier = kTierWave;
    HIPCHK(launch_count<OffT>(a, L.main.blocks, L.main.threads, L.main.lds_bytes, mode, false, s));
    if (fork)
        HIPCHK(hipStreamWaitEvent(s, ctx->ev_join, 0));
    if (ev_count)
        HIPCHK(hipEventRecord(ev_count, s));
    return TRUST_OK;
}

template int enqueue_count<int32_t>(Ctx *, const CsrViews &, int64_t, bool, const trust_params &,
                                    const LaunchPlan &, TcCtl *, int, unsigned long long *,
                                    unsigned long long *, bool, hipStream_t,
                                    hipEvent_t, hipEvent_t, const SegView *, bool, bool);
template int enqueue_count<int64_t>(Ctx *, const CsrViews &, int64_t, bool, const trust_params &,
                                    const LaunchPlan &, TcCtl *, int, unsigned long long *,
                                    unsigned long long *, bool, hipStream_t,
                                    hipEvent_t, hipEvent_t, const SegView *, bool, bool);

int check_params(const trust_params &p)
{
    if (p.total_process < 1 || p.rank < 0 || p.rank >= p.total_process)
        return fail(TRUST_ERR_ARG, "rank %d of total_process %d is invalid", p.rank,
                    p.total_process);
    if (p.hub_threshold < 0)
        return fail(TRUST_ERR_ARG, "hub_threshold %d < 0", p.hub_threshold);
    if (p.chunk < 0 || p.blocks < 0 || p.threads < 0)
        return fail(TRUST_ERR_ARG, "negative launch parameter");
    if (p.schedule != TRUST_SCHED_DEGREE && p.schedule != TRUST_SCHED_LEARNED &&
        p.schedule != TRUST_SCHED_AUTO)
        return fail(TRUST_ERR_ARG, "unknown schedule %d", p.schedule);
    const uint32_t known = TRUST_FLAG_FORCE_SPILL | TRUST_FLAG_OFFSETS64 | TRUST_FLAG_NO_VALIDATE |
                           TRUST_FLAG_WAVE_ONLY | TRUST_FLAG_LINEAR_PROBING | TRUST_FLAG_HIP_GRAPH |
                           TRUST_FLAG_REFERENCE_WALK | TRUST_FLAG_SHORT_WALK |
                           TRUST_FLAG_OFFSET_LOOKUP | TRUST_FLAG_FRESH_INDEX |
                           TRUST_FLAG_REBUILD_LINEAR;
    if (p.flags & ~known)
        return fail(TRUST_ERR_ARG, "unknown flag bits 0x%x", p.flags & ~known);
    if ((p.flags & TRUST_FLAG_REFERENCE_WALK) && (p.flags & TRUST_FLAG_SHORT_WALK))
        return fail(TRUST_ERR_ARG, "REFERENCE_WALK and SHORT_WALK together");
    const int K = p.core_vertices;
    if (K < -1 || (K > 0 && (K < 256 || K > 16384 || (K & (K - 1)))))
        return fail(TRUST_ERR_ARG, "core_vertices %d: -1, 0 or a power of two in [256, 16384]", K);
    if (p.core_min_degree < 0 || p.core_min_degree == 1)
        return fail(TRUST_ERR_ARG, "core_min_degree %d: 0 or >= 2", p.core_min_degree);
    return TRUST_OK;
}

double alg_bytes(int64_t V, int64_t E, uint64_t W, bool off64)
{
    // offsets once + every N+(u) once + two offsets per walked v + every probe
    const double ob = off64 ? 8.0 : 4.0;
    return ob * (double)(V + 1) + 4.0 * (double)E + 2.0 * ob * (double)E + 4.0 * (double)W;
}

double alg_bytes_starts(int64_t V, int64_t E, uint64_t W, bool off64)
{
    const double ob = off64 ? 8.0 : 4.0;
    return ob * (double)(V + 1) + 4.0 * (double)E + (ob + 4.0) * (double)E + 4.0 * (double)W;
}

double alg_bytes_short(int64_t V, int64_t E, uint64_t W, bool off64, int64_t X)
{
    // hash rows: offsets once + every N+(u); walk rows: offsets once + every
    // entry's list start and length; every probe
    const double ob = off64 ? 8.0 : 4.0;
    return 2.0 * ob * (double)(V + 1) + 4.0 * (double)E + (ob + 4.0) * (double)X +
           4.0 * (double)W;
}

// Marks the start of the per-graph index work of this count (once).
static int index_begin(Ctx *ctx, ResidentRun &R)
{
    if (!R.index_ev) {
        HIPCHK(hipEventRecord(ctx->ev_idx[0], ctx->stream));
        R.index_ev = true;
    }
    return TRUST_OK;
}

// The walk of an unpartitioned count: the short side when forced or when
// it makes at most kShortWalkShare of the reference walk's probes.  The
// probe totals and the in-walk rows depend on the graph alone, so a context
// makes them once per loaded graph (or again under TRUST_FLAG_FRESH_INDEX);
// the count that makes them includes their device time in its seconds.
static int prepare_side(Ctx *ctx, bool force, ResidentRun &R)
{
    DevCsr &g = ctx->g;
    // The binned build needs scratch beyond the index itself (per-edge d+,
    // records, ranks); when HBM cannot hold it, the atomic build and its own
    // statistics pass take over.
    bool binned = walk_index_binned(ctx);
    auto leave_binned = [&]() -> int {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        release_walk_scratch(ctx);
        ctx->wi.counted = false;
        binned = false;
        // the atomic build walks every edge: no dense core
        core_release(ctx);
        ctx->core.built = true;
        return TRUST_OK;
    };
    // The binned count pass leaves the dense core's edges out of the walk,
    // so the core comes first; without HBM for it the walk keeps them.
    auto count_pass = [&]() -> int {
        if (!ctx->core.built) {
            int r = core_build(ctx);
            if (r == TRUST_ERR_OOM) {
                HIPCHK(hipStreamSynchronize(ctx->stream));
                core_release(ctx);
                ctx->core.built = true;
                r = TRUST_OK;
            }
            if (r)
                return r;
        }
        return walk_index_count(ctx);
    };
    int rc;
    if (!force && !ctx->side_stats_ready) {
        if ((rc = index_begin(ctx, R)))
            return rc;
        if (binned) {
            // the binned build's count pass yields the probe totals too
            rc = count_pass();
            if (rc == TRUST_ERR_OOM)
                rc = leave_binned();
            if (rc)
                return rc;
        }
        if (!binned) {
            rc = ctx->side_stats.ensure(4 * sizeof(unsigned long long));
            if (rc)
                return rc;
            unsigned long long *st = (unsigned long long *)ctx->side_stats.p;
            if (g.off64)
                HIPCHK(launch_side_stats<int64_t>((const int64_t *)g.beg.p,
                                                  (const int32_t *)g.adj.p, g.V, st, ctx->stream));
            else
                HIPCHK(launch_side_stats<int32_t>((const int32_t *)g.beg.p,
                                                  (const int32_t *)g.adj.p, g.V, st, ctx->stream));
            unsigned long long h[2];
            HIPCHK(hipMemcpyAsync(h, st, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            ctx->side_w_ref = h[0];
            ctx->side_w_short = h[1];
            ctx->side_stats_ready = true;
        }
    }
    R.side = force || (double)ctx->side_w_short <= kShortWalkShare * (double)ctx->side_w_ref;
    if (R.side && !ctx->side_built) {
        if ((rc = index_begin(ctx, R)))
            return rc;
        const size_t ob = g.off64 ? sizeof(int64_t) : sizeof(int32_t);
        bool done = false;
        if (binned) {
            rc = ctx->wi.counted ? TRUST_OK : count_pass();
            if (!rc)
                rc = walk_index_build(ctx);
            if (rc == TRUST_ERR_OOM)
                rc = leave_binned();
            else if (!rc)
                done = true;
            if (rc)
                return rc;
        }
        if (!done) {
            rc = ctx->side_cnt.ensure(2 * sizeof(int) * (size_t)g.V);
            if (!rc)
                rc = ctx->side_bsum.ensure(sizeof(long long) * (size_t)(scan_blocks(g.V) + 1));
            if (!rc)
                rc = ctx->side_beg.ensure(ob * (size_t)(g.V + 1));
            if (!rc)
                rc = ctx->side_ent.ensure(2 * ob * (size_t)g.E);  // WalkEntry: start, length
            if (rc)
                return rc;
            if (g.off64)
                HIPCHK(launch_side_build<int64_t>((const int64_t *)g.beg.p,
                                                  (const int32_t *)g.adj.p, g.V,
                                                  (int *)ctx->side_cnt.p,
                                                  (long long *)ctx->side_bsum.p,
                                                  (int64_t *)ctx->side_beg.p,
                                                  (WalkEntry<int64_t> *)ctx->side_ent.p,
                                                  ctx->stream));
            else
                HIPCHK(launch_side_build<int32_t>((const int32_t *)g.beg.p,
                                                  (const int32_t *)g.adj.p, g.V,
                                                  (int *)ctx->side_cnt.p,
                                                  (long long *)ctx->side_bsum.p,
                                                  (int32_t *)ctx->side_beg.p,
                                                  (WalkEntry<int32_t> *)ctx->side_ent.p,
                                                  ctx->stream));
            ctx->side_entries = -1;  // read back when needed, outside the timed work
        }
        ctx->side_built = true;
        ctx->side_rows_built = false;
    }
    return TRUST_OK;
}

// Frees the binned build's scratch (per-edge d+, histograms, records,
// ranks, bases, out-part copy).
void release_walk_scratch(Ctx *ctx)
{
    WalkIndexState &w = ctx->wi;
    w.dv.release();
    w.M.release();
    w.O.release();
    w.recs.release();
    w.rel.release();
    w.sbase.release();
    w.tmp_out.release();
    w.dtab.release();
}

int ensure_walk_rows(Ctx *ctx)
{
    const DevCsr &g = ctx->g;
    if (!ctx->side_built)
        return fail(TRUST_ERR_ARG, "short-side walk rows not built");
    if (ctx->side_rows_built || g.V <= 0)
        return TRUST_OK;
    int rc = ctx->side_row.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(g.E, 1));
    if (rc)
        return rc;
    if (g.off64)
        HIPCHK(launch_side_rows<int64_t>((const int64_t *)ctx->side_beg.p, g.V,
                                         (int32_t *)ctx->side_row.p, ctx->stream));
    else
        HIPCHK(launch_side_rows<int32_t>((const int32_t *)ctx->side_beg.p, g.V,
                                         (int32_t *)ctx->side_row.p, ctx->stream));
    ctx->side_rows_built = true;
    return TRUST_OK;
}

int side_entry_count(Ctx *ctx, int64_t *n)
{
    const DevCsr &g = ctx->g;
    if (ctx->side_entries < 0) {
        if (g.off64) {
            int64_t x = 0;
            HIPCHK(hipMemcpyAsync(&x, (const int64_t *)ctx->side_beg.p + g.V, sizeof(x),
                                  hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            ctx->side_entries = x;
        } else {
            int32_t x = 0;
            HIPCHK(hipMemcpyAsync(&x, (const int32_t *)ctx->side_beg.p + g.V, sizeof(x),
                                  hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            ctx->side_entries = x;
        }
    }
    *n = ctx->side_entries;
    return TRUST_OK;
}

int ensure_walk_index(Ctx *ctx)
{
    if (ctx->g.V <= 0 || ctx->g.E <= 0 || ctx->side_built)
        return TRUST_OK;
    ResidentRun R;
    int rc = prepare_side(ctx, true, R);
    if (!rc)
        HIPCHK(hipStreamSynchronize(ctx->stream));
    return rc;
}

// Forgets the per-graph indexes of the resident graph, so that the next
// count makes them again (TRUST_FLAG_FRESH_INDEX: a count of a freshly
// loaded graph without reloading it).
static void drop_indexes(Ctx *ctx)
{
    core_drop(ctx);
    ctx->feat_walk = -1;
    ctx->side_stats_ready = ctx->side_built = ctx->side_rows_built = false;
    ctx->wi.counted = false;
    ctx->side_entries = -1;
    ctx->ref_built = false;
}

int prepare_resident(Ctx *ctx, const trust_params &p, int mode, ResidentRun &R)
{
    int rc = check_params(p);
    if (rc)
        return rc;
    DevCsr &g = ctx->g;
    if (!g.adj.p && g.E > 0)
        return fail(TRUST_ERR_ARG, "context holds no graph");
    if (g.max_deg > (int64_t)(INT32_MAX / 64))
        return fail(TRUST_ERR_OVERFLOW, "max out-degree %lld too large for one probe group",
                    (long long)g.max_deg);
    if (mode == kModeBatch && (ctx->batch_ngraphs < 1 || !ctx->batch_vbase.p))
        return fail(TRUST_ERR_ARG, "context holds no batch");
    if (p.flags & TRUST_FLAG_FRESH_INDEX) {
        // the previous count may still read the indexes being rebuilt
        HIPCHK(hipStreamSynchronize(ctx->stream));
        drop_indexes(ctx);
    }
    // the walk rows leave out the edges of the core they were built with:
    // a count that asks for another core makes the indexes again
    int core_K, core_thr;
    bool core_auto;
    core_config(ctx, p, &core_K, &core_thr, &core_auto);
    if (core_K != ctx->core.want_K || core_thr != ctx->core.want_thr ||
        core_auto != ctx->core.want_auto) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        drop_indexes(ctx);
        ctx->core.want_K = core_K;
        ctx->core.want_thr = core_thr;
        ctx->core.want_auto = core_auto;
    }
    R.side = false;
    R.core = false;
    R.index_ev = false;
    R.index_seconds = 0;
    if (!(p.flags & TRUST_FLAG_REFERENCE_WALK) && g.V > 0 && g.E > 0) {
        rc = prepare_side(ctx, (p.flags & TRUST_FLAG_SHORT_WALK) != 0, R);
        if (rc)
            return rc;
        R.core = R.side && ctx->core.active;
    }
    if (mode == kModeInstr) {
        rc = ctx->instr_cycles.ensure(sizeof(unsigned long long) * (size_t)std::max<int64_t>(g.V, 1));
        if (!rc)
            rc = ctx->instr_probes.ensure(sizeof(unsigned long long) *
                                          (size_t)std::max<int64_t>(g.V, 1));
        if (rc)
            return rc;
    }
    // learned schedule when a cost model is loaded (AUTO: unless the judge
    // found the degree split faster on this graph), else the degree split
    R.learned = ctx->mlp_ready && (p.schedule == TRUST_SCHED_LEARNED ||
                                   (p.schedule == TRUST_SCHED_AUTO && ctx->judge != TRUST_SCHED_DEGREE));
    // An int64 batch with segments runs the int32 kernel per segment, each
    // with its own plan (degree split or learned); the short-side walk is
    // built over the whole batch and keeps the int64 kernel, and so does the
    // instrumented mode, whose outputs are indexed by vertex.
    R.segmented = !ctx->segs.empty() && mode != kModeInstr && !R.side;
    rc = resolve_launch(ctx, p, g.max_deg, g.V, g.off64 && !R.segmented, R.L, mode);
    if (rc)
        return rc;
    rc = ctx->ctl.ensure(sizeof(TcCtl) * (R.segmented ? ctx->segs.size() : 1));
    if (rc)
        return rc;
    // Single graphs with int32 offsets.  Batches keep the lookups: measured
    // on the 10k ER/BA batch (7 int32 segments) the start index made the
    // gather-only pass faster (162 -> 152 ms) but the count slower (207 ->
    // 210 ms): its many small vertices are bound by instruction issue, and
    // the lookup walk reuses the table's first 64 ids as walk ids for free.
    R.starts = !R.side && !(p.flags & TRUST_FLAG_OFFSET_LOOKUP) && g.V > 0 && g.E > 0 &&
               !g.off64 && ctx->batch_ngraphs == 0;
    if (R.starts && (rc = ensure_start_index(ctx, R)))
        return rc;
    // the learned plan's features read each walk entry's row, and size their
    // pass by the walk rows' entries
    if (R.side && R.learned) {
        int64_t n;
        if ((rc = side_entry_count(ctx, &n)) || (rc = ensure_walk_rows(ctx)))
            return rc;
    }
    // the learned plan's per-vertex features depend on the graph and the
    // walk alone: made once, as index work, and kept for later plans
    if (R.learned && ctx->feat_walk != (R.side ? 1 : 0)) {
        if ((rc = index_begin(ctx, R)) || (rc = ensure_features(ctx, R.side)))
            return rc;
    }
    if (R.index_ev)
        HIPCHK(hipEventRecord(ctx->ev_idx[1], ctx->stream));
    return TRUST_OK;
}

int ensure_start_index(Ctx *ctx, ResidentRun &R)
{
    const DevCsr &g = ctx->g;
    if (ctx->ref_built)
        return TRUST_OK;
    if (g.off64 || R.segmented)
        return fail(TRUST_ERR_ARG, "start index of an int64 layout");
    int rc = ctx->ref_ent.ensure(sizeof(WalkEntry<int32_t>) * (size_t)g.E);
    if (rc)
        return rc;
    if ((rc = index_begin(ctx, R)))
        return rc;
    HIPCHK(launch_start_index<int32_t>((const int32_t *)g.beg.p, (const int32_t *)g.adj.p, g.E,
                                       (WalkEntry<int32_t> *)ctx->ref_ent.p, ctx->stream));
    ctx->ref_built = true;
    return TRUST_OK;
}

int enqueue_resident(Ctx *ctx, const trust_params &p, int mode, const ResidentRun &R, bool timed)
{
    DevCsr &g = ctx->g;
    // buffers are cleared by fill kernels, which a captured graph replays
    // correctly (hipMemsetAsync nodes are avoided)
    if (mode == kModeInstr && g.V > 0) {
        HIPCHK(launch_fill_u32((uint32_t *)ctx->instr_cycles.p, 0u, 2 * g.V, ctx->stream));
        HIPCHK(launch_fill_u32((uint32_t *)ctx->instr_probes.p, 0u, 2 * g.V, ctx->stream));
    }
    if (mode == kModeBatch) {
        uint32_t *out = (uint32_t *)ctx->batch_out.p;
        HIPCHK(launch_fill_u32(out, 0u, 4 * ctx->batch_ngraphs, ctx->stream));
        HIPCHK(launch_fill_u32(out + 4 * ctx->batch_ngraphs, 0xffffffffu, 2, ctx->stream));
    }
    if (R.segmented)
        return enqueue_segments(ctx, p, mode, R, timed);
    CsrViews v{g.beg.p, (const int32_t *)g.adj.p, g.beg.p, (const int32_t *)g.adj.p, g.beg.p,
               (const int32_t *)g.adj.p};
    if (timed)
        HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    int rc;
    if (R.side) {  // walk rows built by prepare_side; hash and probe rows stay N+
        v.begW = ctx->side_beg.p;
        v.adjW = nullptr;
        v.entW = ctx->side_ent.p;
    } else if (R.starts) {  // walk rows = N+, each entry's list read from the start index
        v.entW = ctx->ref_ent.p;
    }
    if (R.learned) {
        rc = enqueue_learned_plan(ctx, p, R.L, (TcCtl *)ctx->ctl.p, ctx->stream,
                                  R.side ? v.begW : nullptr);
        if (rc)
            return rc;
    }
    hipEvent_t e1 = timed ? ctx->ev[1] : nullptr, e2 = timed ? ctx->ev[2] : nullptr;
    // the short-side walk probes long lists: six windows in flight per wave
    const int m = R.side && (mode == kModeCount || mode == kModeGather) ? mode | kProbeDeep : mode;
    // the dense core's edges, left out of the walk, are counted beside it
    // (the gather-only pass times the walk alone)
    const bool core = R.core && (mode == kModeCount || mode == kModeInstr);
    if (g.off64)
        rc = enqueue_count<int64_t>(ctx, v, g.V, true, p, R.L, (TcCtl *)ctx->ctl.p, m,
                                    (unsigned long long *)ctx->instr_cycles.p,
                                    (unsigned long long *)ctx->instr_probes.p, R.learned,
                                    ctx->stream, e1, e2, nullptr, false, core);
    else
        rc = enqueue_count<int32_t>(ctx, v, g.V, true, p, R.L, (TcCtl *)ctx->ctl.p, m,
                                    (unsigned long long *)ctx->instr_cycles.p,
                                    (unsigned long long *)ctx->instr_probes.p, R.learned,
                                    ctx->stream, e1, e2, nullptr, false, core);
    return rc;
}

int finish_resident(Ctx *ctx, ResidentRun &R, const TcCtl &h, double t_plan, double t_count,
                    trust_result *r)
{
    if (R.index_ev) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev_idx[0], ctx->ev_idx[1]));
        R.index_seconds = ms * 1e-3;
    }
    if (!r)
        return TRUST_OK;
    const DevCsr &g = ctx->g;
    const LaunchPlan &L = R.L;
    int64_t X = 0;
    if (R.side) {
        int rc = side_entry_count(ctx, &X);
        if (rc)
            return rc;
    }
    memset(r, 0, sizeof(*r));
    r->vertices = g.V;
    r->edges = g.E;
    r->triangles = h.count;
    r->probes = h.probes;
    r->plan_seconds = t_plan;
    r->count_seconds = t_count;
    r->index_seconds = R.index_seconds;
    r->seconds = R.index_seconds + t_plan + t_count;
    r->h2d_seconds = ctx->h2d_seconds;
    r->gteps = r->seconds > 0 ? (double)g.E / r->seconds / 1e9 : 0.0;
    r->alg_bytes = R.side ? alg_bytes_short(g.V, g.E, h.probes, g.off64, X)
                 : R.starts ? alg_bytes_starts(g.V, g.E, h.probes, g.off64 && !R.segmented)
                          : alg_bytes(g.V, g.E, h.probes, g.off64 && !R.segmented);
    r->n_spill = h.list_begin[kTierSpill + 1] - h.list_begin[kTierSpill];
    r->n_hub = h.list_begin[kTierHub + 1] - h.list_begin[kTierBig];  // big + hub
    r->n_wave = h.list_begin[kTierWave + 1] - h.list_begin[kTierWave];
    r->threads = L.main.threads;
    r->blocks = L.main.blocks;
    r->hub_slots = 1 << L.main.group_lg;
    r->wave_slots = L.main.wave_slots;
    r->schedule_used = R.learned ? TRUST_SCHED_LEARNED : TRUST_SCHED_DEGREE;
    r->walk_used = R.side ? 1 : 0;
    r->linear_rebuilds = h.rebuilds;
    if (R.core) {
        r->core_vertices = ctx->core.nC;
        r->dense_edges = ctx->core.dense_edges;
        r->dense_probes = ctx->core.dense_probes;
        r->dense_seconds = R.dense_seconds;
    }
    return TRUST_OK;
}

// Count over the context's resident graph.
int count_resident(Ctx *ctx, const trust_params &p, int mode,
                   unsigned long long *cycles_out, unsigned long long *probes_out,
                   trust_result *r)
{
    ResidentRun R;
    int rc = prepare_resident(ctx, p, mode, R);
    if (rc)
        return rc;
    rc = enqueue_resident(ctx, p, mode, R, true);
    if (rc)
        return rc;
    const DevCsr &g = ctx->g;
    TcCtl h;
    if (mode == kModeInstr) {
        if (cycles_out)
            HIPCHK(hipMemcpyAsync(cycles_out, ctx->instr_cycles.p,
                                  sizeof(unsigned long long) * g.V, hipMemcpyDeviceToHost,
                                  ctx->stream));
        if (probes_out)
            HIPCHK(hipMemcpyAsync(probes_out, ctx->instr_probes.p,
                                  sizeof(unsigned long long) * g.V, hipMemcpyDeviceToHost,
                                  ctx->stream));
    }
    if ((rc = read_ctl(ctx, R, h)))
        return rc;
    float t_plan = 0, t_count = 0;
    HIPCHK(hipEventElapsedTime(&t_plan, ctx->ev[0], ctx->ev[1]));
    HIPCHK(hipEventElapsedTime(&t_count, ctx->ev[1], ctx->ev[2]));
    if (R.core && (mode == kModeCount || mode == kModeInstr)) {
        float t_core = 0;
        HIPCHK(hipEventElapsedTime(&t_core, ctx->ev_core[0], ctx->ev_core[1]));
        R.dense_seconds = t_core * 1e-3;
    }
    return finish_resident(ctx, R, h, t_plan * 1e-3, t_count * 1e-3, r);
}

int ctx_init(Ctx *ctx, int device)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(TRUST_ERR_NODEV, "no HIP device available (%s)",
                    e == hipSuccess ? "count 0" : hipGetErrorString(e));
    if (device < 0 || device >= n)
        return fail(TRUST_ERR_ARG, "device %d out of range [0,%d)", device, n);
    ctx->device = device;
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    ctx->num_cus = prop.multiProcessorCount;
    ctx->lds_max = prop.sharedMemPerBlock;
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    for (auto &ev : ctx->ev)
        HIPCHK(hipEventCreate(&ev));
    for (auto &ev : ctx->ev_idx)
        HIPCHK(hipEventCreate(&ev));
    for (auto &ev : ctx->ev_core)
        HIPCHK(hipEventCreate(&ev));
    ctx->ready = true;
    return TRUST_OK;
}

Ctx::~Ctx()
{
    if (!ready)
        return;
    (void)hipSetDevice(device);
    (void)hipStreamSynchronize(stream);
    (void)hipStreamSynchronize(copy_stream);
    (void)hipStreamSynchronize(aux_stream);
    drop_captured(this);
    g.beg.release();
    g.adj.release();
    for (auto &ev : this->ev)
        (void)hipEventDestroy(ev);
    for (auto &ev : this->ev_idx)
        (void)hipEventDestroy(ev);
    for (auto &ev : this->ev_core)
        (void)hipEventDestroy(ev);
    (void)hipEventDestroy(ev_fork);
    (void)hipEventDestroy(ev_join);
    (void)hipStreamDestroy(stream);
    (void)hipStreamDestroy(copy_stream);
    (void)hipStreamDestroy(aux_stream);
}

int load_ctx_csr(Ctx *ctx, const int64_t *beg, const int32_t *adj, int64_t V, int64_t E,
                 uint32_t flags)
{
    if (!(flags & TRUST_FLAG_NO_VALIDATE)) {
        int rc = trust_csr_validate(beg, adj, V, E);
        if (rc)
            return rc;
    }
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // state derived from the previous graph
    drop_captured(ctx);
    ctx->batch_ngraphs = 0;
    ctx->segs.clear();
    drop_indexes(ctx);  // the walk statistics and indexes belong to the graph
    ctx->judge = -1;    // the judgement belongs to the graph; the model stays
    auto t0 = std::chrono::steady_clock::now();
    int rc = upload_csr(ctx, ctx->g, beg, adj, V, E, (flags & TRUST_FLAG_OFFSETS64) != 0,
                        ctx->stream);
    if (rc)
        return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->h2d_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return TRUST_OK;
}

}  // namespace trust

using namespace trust;

extern "C" int trust_set_device_memory_cap(uint64_t bytes)
{
    clear_error();
    g_dev_cap.store((unsigned long long)bytes);
    return TRUST_OK;
}

extern "C" uint64_t trust_device_memory_held(void) { return (uint64_t)g_dev_held.load(); }

extern "C" int trust_device_count(void)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice)
        return 0;
    if (e != hipSuccess)
        return hip_fail(e, "hipGetDeviceCount");
    return n;
}

extern "C" int trust_ctx_create(int device, trust_ctx **out)
{
    if (!out)
        return fail(TRUST_ERR_ARG, "null output");
    *out = nullptr;
    std::unique_ptr<Ctx> c(new (std::nothrow) Ctx());
    if (!c)
        return fail(TRUST_ERR_OOM, "cannot allocate context");
    int rc = ctx_init(c.get(), device);
    if (rc)
        return rc;
    *out = reinterpret_cast<trust_ctx *>(c.release());
    return TRUST_OK;
}

extern "C" void trust_ctx_destroy(trust_ctx *ctx) { delete reinterpret_cast<Ctx *>(ctx); }

extern "C" int trust_ctx_load_csr(trust_ctx *c, const int64_t *beg, const int32_t *adj,
                                  int64_t V, int64_t E, uint32_t flags, double *h2d_seconds)
{
    if (!c)
        return fail(TRUST_ERR_ARG, "null context");
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    int rc = load_ctx_csr(ctx, beg, adj, V, E, flags);
    if (!rc && h2d_seconds)
        *h2d_seconds = ctx->h2d_seconds;
    return rc;
}

extern "C" int trust_ctx_count(trust_ctx *c, const trust_params *p, trust_result *r)
{
    if (!c || !p)
        return fail(TRUST_ERR_ARG, "null argument");
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    HIPCHK(hipSetDevice(ctx->device));
    return count_resident(ctx, *p, kModeCount, nullptr, nullptr, r);
}

extern "C" int trust_ctx_profile_vertices(trust_ctx *c, const trust_params *p,
                                          uint64_t *cycles, uint64_t *probes, trust_result *r)
{
    if (!c || !p)
        return fail(TRUST_ERR_ARG, "null argument");
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    HIPCHK(hipSetDevice(ctx->device));
    return count_resident(ctx, *p, kModeInstr, (unsigned long long *)cycles,
                          (unsigned long long *)probes, r);
}

extern "C" int trust_count_csr(const int64_t *beg, const int32_t *adj, int64_t V, int64_t E,
                               const trust_params *p, trust_result *r)
{
    if (!p)
        return fail(TRUST_ERR_ARG, "null params");
    Ctx ctx;
    int rc = ctx_init(&ctx, p->device);
    if (rc)
        return rc;
    rc = load_ctx_csr(&ctx, beg, adj, V, E, p->flags);
    if (rc)
        return rc;
    return count_resident(&ctx, *p, kModeCount, nullptr, nullptr, r);
}

extern "C" int trust_triangle_count(int rank, const char *folder, int total_process, int threads,
                                    int blocks, int chunk_size, trust_arguments *out)
{
    if (!folder || !out)
        return fail(TRUST_ERR_ARG, "null argument");
    HostCsr g;
    int rc = load_folder(folder, g);
    if (rc)
        return rc;
    trust_params p;
    trust_params_init(&p);
    p.rank = rank;
    p.total_process = total_process;
    p.threads = threads > 0 ? threads : 0;
    p.blocks = blocks > 0 ? blocks : 0;
    p.chunk = chunk_size > 0 ? chunk_size : 0;
    Ctx ctx;
    rc = ctx_init(&ctx, 0);
    if (rc)
        return rc;
    rc = load_ctx_csr(&ctx, g.beg.data(), g.adj.data(), g.V, g.E, TRUST_FLAG_NO_VALIDATE);
    if (rc)
        return rc;
    trust_result r;
    rc = count_resident(&ctx, p, kModeCount, nullptr, nullptr, &r);
    if (rc)
        return rc;
    out->edge_count = g.E;
    out->count = r.triangles;
    out->time = r.seconds;
    out->degree = g.V > 0 ? g.E / g.V : 0;
    out->vertices = g.V;
    return TRUST_OK;
}

extern "C" int trust_ctx_gather(trust_ctx *c, const trust_params *p, trust_result *r)
{
    if (!c || !p)
        return fail(TRUST_ERR_ARG, "null argument");
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    HIPCHK(hipSetDevice(ctx->device));
    return count_resident(ctx, *p, kModeGather, nullptr, nullptr, r);
}

extern "C" int trust_membench_read(int device, int64_t bytes, int reps, double *bytes_per_second)
{
    if (bytes < 16 || reps < 1 || !bytes_per_second)
        return fail(TRUST_ERR_ARG, "bad arguments");
    Ctx ctx;
    int rc = ctx_init(&ctx, device);
    if (rc)
        return rc;
    const long long n = bytes / 16;
    const int blocks = ctx.num_cus * 8;
    DevBuf a, sink;
    rc = a.ensure((size_t)n * 16);
    if (!rc)
        rc = sink.ensure(sizeof(unsigned) * (size_t)blocks);
    if (rc)
        return rc;
    HIPCHK(hipMemsetAsync(a.p, 1, (size_t)n * 16, ctx.stream));
    HIPCHK(launch_read((const int4 *)a.p, n, 1, blocks, (unsigned *)sink.p, ctx.stream));  // warm-up
    HIPCHK(hipEventRecord(ctx.ev[0], ctx.stream));
    HIPCHK(launch_read((const int4 *)a.p, n, reps, blocks, (unsigned *)sink.p, ctx.stream));
    HIPCHK(hipEventRecord(ctx.ev[1], ctx.stream));
    HIPCHK(hipStreamSynchronize(ctx.stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx.ev[0], ctx.ev[1]));
    *bytes_per_second = (double)blocks * reps * (double)n * 16.0 / (ms * 1e-3);
    return TRUST_OK;
}

extern "C" int trust_membench_copy(int device, int64_t bytes, int reps, double *seconds_per_copy)
{
    if (bytes < 16 || reps < 1 || !seconds_per_copy)
        return fail(TRUST_ERR_ARG, "bad arguments");
    Ctx ctx;
    int rc = ctx_init(&ctx, device);
    if (rc)
        return rc;
    const long long n = bytes / 16;
    DevBuf a, b;
    rc = a.ensure((size_t)n * 16);
    if (!rc)
        rc = b.ensure((size_t)n * 16);
    if (rc)
        return rc;
    HIPCHK(hipMemsetAsync(a.p, 1, (size_t)n * 16, ctx.stream));
    const int blocks = ctx.num_cus * 8;
    HIPCHK(launch_copy((const int4 *)a.p, (int4 *)b.p, n, blocks, ctx.stream));  // warm-up
    HIPCHK(hipEventRecord(ctx.ev[0], ctx.stream));
    for (int i = 0; i < reps; i++)
        HIPCHK(launch_copy((const int4 *)a.p, (int4 *)b.p, n, blocks, ctx.stream));
    HIPCHK(hipEventRecord(ctx.ev[1], ctx.stream));
    HIPCHK(hipStreamSynchronize(ctx.stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx.ev[0], ctx.ev[1]));
    *seconds_per_copy = ms * 1e-3 / reps;
    return TRUST_OK;
}
