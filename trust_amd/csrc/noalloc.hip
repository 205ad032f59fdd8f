// trust_count_csr_noalloc: the count of trust_count_csr for a CSR the caller
// already holds in device memory, with no device allocation at all.
//
// For callers that cannot allocate device memory (code under stream
// capture, serving loops, harnesses that forbid hipMalloc).  Everything the
// kernels need beyond the caller's inputs -- the control block, the
// planner's per-block tier counts, the int32 offsets and the tier list --
// lives in one file-scope device array, `scratch`, of kScratchBytes: each
// GPU the library runs on holds its own copy until the process ends.  Its
// device address comes from hipGetSymbolAddress, once per device, with that
// device current (the array's name in host code is only a host-side
// stand-in); every launch gets that address.  Limits: V <= kNoAllocMaxV,
// E < 2^30, no table beyond the LDS (max out-degree about 16k at 1024
// threads), the reference walk with offset lookups (the per-graph walk
// indexes are as large as adj), the degree split.  Calls on one GPU share
// its array, so calls that overlap on one GPU race; each call clears what
// must start at zero.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "engine.h"

namespace trust {

constexpr int64_t kNoAllocMaxV = int64_t(1) << 22;
constexpr size_t kScrCtl = 0;                      // TcCtl
constexpr size_t kScrMisc = 4096;                  // NoAllocMisc
constexpr size_t kScrBlocks = 8192;                // kNumTiers x planner blocks (long long)
constexpr size_t kScrBlocksBytes = sizeof(long long) * kNumTiers *
                                   ((kNoAllocMaxV + kPlanPerBlock - 1) / kPlanPerBlock);
constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
constexpr size_t kScrBeg = align256(kScrBlocks + kScrBlocksBytes);       // int32 offsets
constexpr size_t kScrList = align256(kScrBeg + 4 * (kNoAllocMaxV + 1));  // tier list
constexpr size_t kScratchBytes = align256(kScrList + 4 * kNoAllocMaxV);
static_assert(sizeof(TcCtl) <= kScrMisc - kScrCtl, "control block does not fit");

struct NoAllocMisc {
    unsigned long long max_deg;  // largest out-degree
    unsigned long long t0, t1;   // device wall clock around the count
    unsigned int bad_offsets;    // beg[0] != 0, beg[V] != E or a decreasing offset
    unsigned int bad_ids;        // an id outside [0, V)
};

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// hipGetSymbolAddress result per device (null: not yet looked up)
constexpr int kMaxDevices = 64;
static void *g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_address(int device, unsigned char **out)
{
    if (device < 0 || device >= kMaxDevices)
        return fail(TRUST_ERR_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
            return fail(TRUST_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch[device] = p;
    }
    *out = (unsigned char *)g_scratch[device];
    return TRUST_OK;
}

// int64 -> int32 offsets, the largest out-degree and the offsets' checks.
__global__ void na_offsets_kernel(const int64_t *__restrict__ beg, long long V, long long E,
                                  int32_t *__restrict__ beg32, NoAllocMisc *__restrict__ m)
{
    unsigned long long dmax = 0;
    unsigned int bad = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i <= V;
         i += (long long)gridDim.x * blockDim.x) {
        const int64_t b = beg[i];
        beg32[i] = (int32_t)b;
        if (i == 0)
            bad |= b != 0;
        if (i == V)
            bad |= b != E;
        if (i < V) {
            const int64_t d = beg[i + 1] - b;
            bad |= d < 0;
            dmax = d > (int64_t)dmax ? (unsigned long long)d : dmax;
        }
    }
    if (dmax)
        atomicMax(&m->max_deg, dmax);
    if (bad)
        atomicOr(&m->bad_offsets, 1u);
}

__global__ void na_ids_kernel(const int32_t *__restrict__ adj, long long E, long long V,
                              NoAllocMisc *__restrict__ m)
{
    unsigned int bad = 0;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (long long)gridDim.x * blockDim.x) {
        const int v = adj[e];
        bad |= v < 0 || (long long)v >= V;
    }
    if (bad)
        atomicOr(&m->bad_ids, 1u);
}

__global__ void na_stamp_kernel(unsigned long long *t) { *t = wall_clock64(); }

}  // namespace trust

using namespace trust;

extern "C" int trust_count_csr_noalloc(const int64_t *beg_dev, const int32_t *adj_dev, int64_t V,
                                       int64_t E, const trust_params *p, trust_result *r,
                                       void *stream)
{
    clear_error();
    if (!p)
        return fail(TRUST_ERR_ARG, "null params");
    if (V < 0 || E < 0 || (V > 0 && !beg_dev) || (E > 0 && !adj_dev))
        return fail(TRUST_ERR_ARG, "bad CSR arguments");
    if (V > kNoAllocMaxV)
        return fail(TRUST_ERR_OVERFLOW,
                    "V=%lld exceeds the allocation-free path's maximum of %lld vertices",
                    (long long)V, (long long)kNoAllocMaxV);
    if (E >= kMaxEdges32)
        return fail(TRUST_ERR_OVERFLOW,
                    "E=%lld: the allocation-free path takes fewer than 2^30 edges", (long long)E);
    int rc = check_params(*p);
    if (rc)
        return rc;
    const uint32_t unsupported = TRUST_FLAG_FORCE_SPILL | TRUST_FLAG_OFFSETS64 |
                                 TRUST_FLAG_HIP_GRAPH | TRUST_FLAG_SHORT_WALK |
                                 TRUST_FLAG_FRESH_INDEX;
    if (p->flags & unsupported)
        return fail(TRUST_ERR_ARG, "flags 0x%x are not available without allocation",
                    p->flags & unsupported);
    if (p->schedule == TRUST_SCHED_LEARNED)
        return fail(TRUST_ERR_ARG, "the learned schedule needs a context (trust_ctx)");
    HIPCHK(hipSetDevice(p->device));
    unsigned char *scr = nullptr;
    if ((rc = scratch_address(p->device, &scr)))
        return rc;
    hipStream_t s = (hipStream_t)stream;
    TcCtl *ctl = (TcCtl *)(scr + kScrCtl);
    NoAllocMisc *misc = (NoAllocMisc *)(scr + kScrMisc);
    long long *block_counts = (long long *)(scr + kScrBlocks);
    int32_t *beg32 = (int32_t *)(scr + kScrBeg);
    int32_t *list = (int32_t *)(scr + kScrList);

    // clear the control block and the checks (fill kernels, no memset nodes)
    HIPCHK(launch_fill_u32((uint32_t *)scr, 0u, (kScrMisc + sizeof(NoAllocMisc)) / 4, s));
    if (V > 0) {
        const int blocks = (int)std::min<long long>((V + 256) / 256, 4096);
        hipLaunchKernelGGL(na_offsets_kernel, dim3(blocks), dim3(256), 0, s, beg_dev, (long long)V,
                           (long long)E, beg32, misc);
        HIPCHK(hipGetLastError());
        if (!(p->flags & TRUST_FLAG_NO_VALIDATE) && E > 0) {
            const int eb = (int)std::min<long long>((E + 255) / 256, 8192);
            hipLaunchKernelGGL(na_ids_kernel, dim3(eb), dim3(256), 0, s, adj_dev, (long long)E,
                               (long long)V, misc);
            HIPCHK(hipGetLastError());
        }
    }
    NoAllocMisc hm;
    HIPCHK(hipMemcpyAsync(&hm, misc, sizeof(hm), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hm.bad_offsets)
        return fail(TRUST_ERR_FORMAT, "offsets: begin[0] != 0, begin[V] != E or not monotone");
    if (hm.bad_ids)
        return fail(TRUST_ERR_FORMAT, "an id of adj lies outside [0, V)");
    if (hm.max_deg > (unsigned long long)(INT32_MAX / 64))
        return fail(TRUST_ERR_OVERFLOW, "max out-degree %llu too large for one probe group",
                    hm.max_deg);

    // launch shapes from the device's properties (a context that owns nothing)
    Ctx props;
    int cus = 0, lds = 0, wall_khz = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device));
    HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device));
    HIPCHK(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, p->device));
    props.device = p->device;
    props.num_cus = cus;
    props.lds_max = (size_t)lds;
    LaunchPlan L;
    if ((rc = resolve_launch(&props, *p, (int64_t)hm.max_deg, V, false, L, kModeCount)))
        return rc;
    if (L.spill_groups > 0)
        return fail(TRUST_ERR_OVERFLOW,
                    "max out-degree %llu needs an HBM hash table, which the allocation-free path "
                    "does not have", hm.max_deg);

    hipLaunchKernelGGL(na_stamp_kernel, dim3(1), dim3(1), 0, s, &misc->t0);
    const int nblocks = (int)((V + kPlanPerBlock - 1) / kPlanPerBlock);
    PlanArgs<int32_t> pa;
    memset(&pa, 0, sizeof(pa));
    pa.begH = pa.begW = beg32;
    pa.endW = beg32 + 1;
    pa.V = V;
    pa.same_csr = 1;
    pa.hub_threshold = p->hub_threshold;
    pa.wave_lg = L.main.wave_lg;
    pa.hub_lg = L.main.group_lg;
    pa.big_lg = L.has_big ? L.big.group_lg : 0;
    pa.wave_only = (p->flags & TRUST_FLAG_WAVE_ONLY) ? 1 : 0;
    pa.list = list;
    pa.block_counts = block_counts;
    pa.per_block = kPlanPerBlock;
    pa.ctl = ctl;
    HIPCHK(launch_plan<int32_t>(pa, nblocks, s));
    TcArgs<int32_t> a;
    memset(&a, 0, sizeof(a));
    a.begH = a.begW = a.begP = beg32;
    a.endW = beg32 + 1;
    a.adjH = a.adjW = a.adjP = adj_dev;
    a.list = list;
    a.spill_tier = -1;
    a.chunk_group = L.chunk_group;
    a.chunk_wave = L.chunk_wave;
    a.rank = p->rank;
    a.nranks = p->total_process;
    a.force_linear = (p->flags & TRUST_FLAG_LINEAR_PROBING) ? 1
                     : (p->flags & TRUST_FLAG_REBUILD_LINEAR) ? 2 : 0;
    a.ctl = ctl;
    if (L.has_big) {  // tables beyond the main launch's LDS share
        a.table_words = L.big.table_words;
        a.group_lg = L.big.group_lg;
        a.wave_slots = L.big.wave_slots;
        a.wave_lg = L.big.wave_lg;
        a.group_tier = kTierBig;
        a.wave_tier = -1;
        HIPCHK(launch_count<int32_t>(a, L.big.blocks, L.big.threads, L.big.lds_bytes, kModeCount,
                                     true, s));
    }
    a.table_words = L.main.table_words;
    a.group_lg = L.main.group_lg;
    a.wave_slots = L.main.wave_slots;
    a.wave_lg = L.main.wave_lg;
    a.group_tier = kTierHub;
    a.wave_tier = kTierWave;
    HIPCHK(launch_count<int32_t>(a, L.main.blocks, L.main.threads, L.main.lds_bytes, kModeCount,
                                 false, s));
    hipLaunchKernelGGL(na_stamp_kernel, dim3(1), dim3(1), 0, s, &misc->t1);
    HIPCHK(hipGetLastError());
    TcCtl h;
    HIPCHK(hipMemcpyAsync(&h, ctl, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&hm, misc, sizeof(hm), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (h.list_begin[kTierSpill + 1] != h.list_begin[kTierSpill])
        return fail(TRUST_ERR_OVERFLOW, "a table needs HBM, which the allocation-free path lacks");
    if (r) {
        memset(r, 0, sizeof(*r));
        r->vertices = V;
        r->edges = E;
        r->triangles = h.count;
        r->probes = h.probes;
        r->seconds = r->count_seconds =
            wall_khz > 0 && hm.t1 > hm.t0 ? (double)(hm.t1 - hm.t0) / (wall_khz * 1e3) : 0.0;
        r->gteps = r->seconds > 0 ? (double)E / r->seconds / 1e9 : 0.0;
        r->alg_bytes = alg_bytes(V, E, h.probes, false);
        r->n_spill = 0;
        r->n_hub = h.list_begin[kTierHub + 1] - h.list_begin[kTierBig];
        r->n_wave = h.list_begin[kTierWave + 1] - h.list_begin[kTierWave];
        r->threads = L.main.threads;
        r->blocks = L.main.blocks;
        r->hub_slots = 1 << L.main.group_lg;
        r->wave_slots = L.main.wave_slots;
        r->schedule_used = TRUST_SCHED_DEGREE;
        r->walk_used = 0;
        r->linear_rebuilds = h.rebuilds;
    }
    return TRUST_OK;
}
