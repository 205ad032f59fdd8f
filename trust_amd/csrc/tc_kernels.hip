// gfx950 kernels for TRUST's vertex-centric hashed triangle count.
//
// Count = sum_u sum_{v in N+(u)} |N+(u) & N+(v)| over a degree-oriented CSR
// (reference semantics: Without-graph-partition/TC.cu:230-600; partitioned
// three-tile form: With-graph-partition/TC.cu:630-973).
//
// Design (MI355X-first, not a translation of the CUDA kernel):
//  * wave64 throughout.  A vertex is counted by one wavefront (wave tier),
//    by a whole workgroup with an LDS table (hub tier) or by a whole
//    workgroup with an HBM table (spill tier) when its table does not fit
//    LDS.  Tables are open-addressing (linear probing, Fibonacci hash) with
//    >= 2 slots per key, so they never overflow (the reference drops entries
//    past 100 per bucket, TC.cu:293-301) and every probe loop ends at an
//    empty slot.
//  * The probe walk processes 64 neighbours v of u at a time: lanes load
//    (beg[v], deg v), an inclusive wave scan (DPP row_shr / row_bcast) lays
//    the 64 lists end to end, and each 64-wide window of that concatenation
//    is one coalesced load per lane.  The segment owning each lane is found
//    with a 64-entry LDS mark array and a wave max-scan; a window inside a
//    single list skips that and reads its base with one readlane.  This
//    replaces the reference's per-lane skip loop and its 32-lane
//    __shfl_sync resynchronisation (TC.cu:513-541) and its 16x64 CTA walk
//    that is only correct at 1024 threads (TC.cu:330-359).
//  * Persistent workgroups pull chunks of vertices from per-tier atomic
//    queues (the reference's G_INDEX counters, TC.cu:405-422,560-577);
//    logical rank r of P takes chunks c with c % P == r (TC.cu:259,433).
//  * The count is reduced per wave in registers and added once per wave to
//    a uint64 in device memory (TC.cu:586-596).
#include <hip/hip_runtime.h>

#include "tc_kernels.h"

namespace trust {

constexpr int kEmpty = -1;

// ---------------------------------------------------------------- wave ops
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive prefix sum over the 64 lanes (all lanes active).
__device__ __forceinline__ int wave_scan_add(int x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

// Inclusive prefix max over the 64 lanes of non-negative values.
__device__ __forceinline__ int wave_scan_max(int x)
{
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));
    return x;
}

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ int shfl_idx(int x, int src) { return __builtin_amdgcn_ds_bpermute(src << 2, x); }
__device__ __forceinline__ int64_t shfl_idx(int64_t x, int src)
{
    int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(x & 0xffffffff));
    int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(x >> 32));
    return ((int64_t)hi << 32) | (unsigned int)lo;
}
__device__ __forceinline__ int readlane(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ int64_t readlane(int64_t x, int l)
{
    int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffff), l);
    int hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return ((int64_t)hi << 32) | (unsigned int)lo;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x)
{
    for (int off = 32; off > 0; off >>= 1)
        x += __shfl_xor(x, off);
    return x;
}

// ---------------------------------------------------------- hash tables
__device__ __forceinline__ int table_lg(int d)
{
    // smallest lg >= kMinTableLg with 2^lg >= 2d
    int need = 2 * d - 1;
    int lg = need > 0 ? 32 - __clz(need) : 0;
    return lg < kMinTableLg ? kMinTableLg : lg;
}

__device__ __forceinline__ unsigned int hash_slot(int key, int lg)
{
    return ((unsigned int)key * 0x9E3779B1u) >> (32 - lg);
}

template <bool kGlobal>
__device__ __forceinline__ int slot_load(const int *p)
{
    if constexpr (kGlobal)
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return *p;
}

template <bool kGlobal>
__device__ __forceinline__ void slot_clear(int *p)
{
    if constexpr (kGlobal)
        __hip_atomic_store(p, kEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        *p = kEmpty;
}

template <bool kGlobal>
__device__ __forceinline__ void table_insert(int *tab, int lg, int key)
{
    const unsigned int mask = (1u << lg) - 1u;
    unsigned int h = hash_slot(key, lg);
    for (;;) {
        int prev = atomicCAS(tab + h, kEmpty, key);
        if (prev == kEmpty || prev == key)
            return;
        h = (h + 1u) & mask;
    }
}

template <bool kGlobal>
__device__ __forceinline__ unsigned int table_contains(const int *tab, int lg, int key)
{
    const unsigned int mask = (1u << lg) - 1u;
    unsigned int h = hash_slot(key, lg);
    for (;;) {
        int s = slot_load<kGlobal>(tab + h);
        if (s == key)
            return 1u;
        if (s == kEmpty)
            return 0u;
        h = (h + 1u) & mask;
    }
}

// ---------------------------------------------------------- probe walk
// Probes every w in N+(v) for the (up to 64) neighbours v held one per
// lane, against the table of N+(u).  Must be called by all 64 lanes.
// `mark` is this wave's 64-int LDS scratch and is all-zero on entry and
// on exit.  Returns this lane's hits; `total` receives the probe count.
template <typename OffT, bool kGlobal>
__device__ __forceinline__ unsigned int probe_group(int v, bool valid, const OffT *__restrict__ begP,
                                                    const int32_t *__restrict__ adjP, const int *tab,
                                                    int lg, int *mark, int &total)
{
    const int lane = lane_id();
    OffT s = 0;
    int deg = 0;
    if (valid) {
        s = begP[v];
        deg = (int)(begP[v + 1] - s);
    }
    const int incl = wave_scan_add(deg);
    const int excl = incl - deg;
    total = readlane(incl, 63);
    const OffT base = s - (OffT)excl;  // element of window position t is adjP[base_j + t]
    unsigned int hits = 0;
    int carry = 0;  // segment that owns the first position of the window
    for (int b = 0; b < total; b += kWave) {
        OffT bj;
        if (readlane(incl, carry) >= b + kWave) {
            bj = readlane(base, carry);  // the window lies inside one list
        } else {
            if (deg > 0 && excl >= b && excl < b + kWave)
                mark[excl - b] = lane;
            wave_sync();
            const int m = mark[lane];
            mark[lane] = 0;
            wave_sync();
            const int j = max(wave_scan_max(m), carry);
            carry = readlane(j, 63);
            bj = shfl_idx(base, j);
        }
        const int t = b + lane;
        if (t < total)
            hits += table_contains<kGlobal>(tab, lg, adjP[bj + (OffT)t]);
    }
    return hits;
}

// ------------------------------------------------------- per-vertex work
// Whole workgroup counts vertex u with a table at `tab` (LDS or HBM).
template <typename OffT, bool kGlobal, bool kInstr>
__device__ __forceinline__ void group_vertex(const TcArgs<OffT> &a, int u, int *tab, int *mark,
                                             int *cursor, unsigned long long &acc,
                                             unsigned long long &probes)
{
    unsigned long long t0 = 0;
    if (kInstr && threadIdx.x == 0)
        t0 = clock64();
    const OffT sH = a.begH[u];
    const int dH = (int)(a.begH[u + 1] - sH);
    const OffT sW = a.begW[u];
    const int dW = (int)(a.begW[u + 1] - sW);
    const int lg = table_lg(dH);
    const int size = 1 << lg;
    for (int i = threadIdx.x; i < size; i += blockDim.x)
        slot_clear<kGlobal>(tab + i);
    if (threadIdx.x == 0)
        *cursor = 0;
    if constexpr (kGlobal)
        __threadfence();  // HBM slots: cleared values visible in L2 before any insert/lookup
    __syncthreads();
    for (int i = threadIdx.x; i < dH; i += blockDim.x)
        table_insert<kGlobal>(tab, lg, a.adjH[sH + (OffT)i]);
    if constexpr (kGlobal)
        __threadfence();
    __syncthreads();
    const int nchunks = (dW + kWave - 1) / kWave;
    unsigned int hits = 0;
    unsigned long long wave_probes = 0;
    for (;;) {
        int c = 0;
        if (lane_id() == 0)
            c = atomicAdd(cursor, 1);
        c = __builtin_amdgcn_readfirstlane(c);
        if (c >= nchunks)
            break;
        const int idx = c * kWave + lane_id();
        const bool valid = idx < dW;
        const int v = valid ? a.adjW[sW + (OffT)idx] : 0;
        int total;
        hits += probe_group<OffT, kGlobal>(v, valid, a.begP, a.adjP, tab, lg, mark, total);
        wave_probes += (unsigned long long)total;
    }
    acc += hits;
    probes += wave_probes;
    __syncthreads();
    if constexpr (kInstr) {
        // per-vertex probe total: every wave adds its share
        if (lane_id() == 0 && wave_probes)
            atomicAdd(a.instr_probes + u, wave_probes);
        if (threadIdx.x == 0)
            a.instr_cycles[u] = clock64() - t0;
    }
}

// One wavefront counts vertex u with its own LDS table slice.
template <typename OffT, bool kInstr>
__device__ __forceinline__ void wave_vertex(const TcArgs<OffT> &a, int u, int *tab, int *mark,
                                            unsigned long long &acc, unsigned long long &probes)
{
    unsigned long long t0 = 0;
    if (kInstr)
        t0 = clock64();
    const int lane = lane_id();
    const OffT sH = a.begH[u];
    const int dH = (int)(a.begH[u + 1] - sH);
    const OffT sW = a.begW[u];
    const int dW = (int)(a.begW[u + 1] - sW);
    const int lg = table_lg(dH);
    const int size = 1 << lg;
    for (int i = lane; i < size; i += kWave)
        tab[i] = kEmpty;
    wave_sync();
    for (int i = lane; i < dH; i += kWave)
        table_insert<false>(tab, lg, a.adjH[sH + (OffT)i]);
    wave_sync();
    unsigned int hits = 0;
    unsigned long long wave_probes = 0;
    for (int c = 0; c < dW; c += kWave) {
        const bool valid = c + lane < dW;
        const int v = valid ? a.adjW[sW + (OffT)(c + lane)] : 0;
        int total;
        hits += probe_group<OffT, false>(v, valid, a.begP, a.adjP, tab, lg, mark, total);
        wave_probes += (unsigned long long)total;
    }
    wave_sync();
    acc += hits;
    probes += wave_probes;
    if constexpr (kInstr) {
        if (lane == 0) {
            a.instr_probes[u] = wave_probes;
            a.instr_cycles[u] = clock64() - t0;
        }
    }
}

// ------------------------------------------------------------ the kernel
template <typename OffT, bool kInstr>
__global__ void __launch_bounds__(1024) tc_count_kernel(TcArgs<OffT> a)
{
    extern __shared__ int lds[];
    const int nwaves = blockDim.x / kWave;
    const int wave = threadIdx.x / kWave;
    const int lane = lane_id();
    int *mark = lds + a.table_words + wave * kWave;
    int *misc = lds + a.table_words + nwaves * kWave;  // [0] chunk cursor, [1] grab
    mark[lane] = 0;
    unsigned long long acc = 0, probes = 0;
    const long long lb0 = a.ctl->list_begin[0], lb1 = a.ctl->list_begin[1];
    const long long lb2 = a.ctl->list_begin[2], lb3 = a.ctl->list_begin[3];

    // Spill tier: tables in HBM, one region per participating workgroup.
    if ((int)blockIdx.x < a.spill_groups && lb1 > lb0) {
        int *tab = a.spill_mem + (long long)blockIdx.x * a.spill_slots;
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0)
                misc[1] = (int)atomicAdd(&a.ctl->next[kTierSpill], 1u);
            __syncthreads();
            const long long c = (long long)(unsigned int)misc[1] * a.nranks + a.rank;
            const long long first = lb0 + c * a.chunk_group;
            if (first >= lb1)
                break;
            const long long last = min(first + (long long)a.chunk_group, lb1);
            for (long long i = first; i < last; i++)
                group_vertex<OffT, true, kInstr>(a, a.list[i], tab, mark, misc, acc, probes);
        }
    }
    // Hub tier: the whole LDS table region is one table.
    if (lb2 > lb1) {
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0)
                misc[1] = (int)atomicAdd(&a.ctl->next[kTierHub], 1u);
            __syncthreads();
            const long long c = (long long)(unsigned int)misc[1] * a.nranks + a.rank;
            const long long first = lb1 + c * a.chunk_group;
            if (first >= lb2)
                break;
            const long long last = min(first + (long long)a.chunk_group, lb2);
            for (long long i = first; i < last; i++)
                group_vertex<OffT, false, kInstr>(a, a.list[i], lds, mark, misc, acc, probes);
        }
    }
    __syncthreads();
    // Wave tier: each wave owns a slice of the table region.
    int *wtab = lds + wave * a.wave_slots;
    for (;;) {
        unsigned int k = 0;
        if (lane == 0)
            k = atomicAdd(&a.ctl->next[kTierWave], 1u);
        k = __builtin_amdgcn_readfirstlane(k);
        const long long c = (long long)k * a.nranks + a.rank;
        const long long first = lb2 + c * a.chunk_wave;
        if (first >= lb3)
            break;
        const long long last = min(first + (long long)a.chunk_wave, lb3);
        for (long long i = first; i < last; i++)
            wave_vertex<OffT, kInstr>(a, a.list[i], wtab, mark, acc, probes);
    }
    acc = wave_sum(acc);
    probes = wave_sum(probes) / kWave;  // every lane carried the wave's probe total
    if (lane == 0) {
        if (acc)
            atomicAdd(&a.ctl->count, acc);
        if (probes)
            atomicAdd(&a.ctl->probes, probes);
    }
}

// ------------------------------------------------------------- planning
template <typename OffT>
__device__ __forceinline__ int plan_tier(const PlanArgs<OffT> &p, long long u)
{
    if (p.tier_override)
        return (int)p.tier_override[u];
    const long long dH = (long long)(p.begH[u + 1] - p.begH[u]);
    const long long dW = (long long)(p.begW[u + 1] - p.begW[u]);
    if (dH == 0 || dW == 0 || (p.same_csr && dW < 2))
        return kTierNone;  // contributes nothing (reference skips d+ < 2, TC.cu:440-441)
    int lg = kMinTableLg;
    while ((1ll << lg) < 2 * dH)
        lg++;
    if (p.force_spill || lg > p.hub_lg)
        return kTierSpill;
    if (!p.wave_only && dW > p.hub_threshold)
        return kTierHub;
    if (lg > p.wave_lg)
        return kTierHub;
    return kTierWave;
}

template <typename OffT>
__global__ void __launch_bounds__(256) plan_count_kernel(PlanArgs<OffT> p, int nblocks)
{
    __shared__ long long cnt[kNumTiers];
    if (threadIdx.x < kNumTiers)
        cnt[threadIdx.x] = 0;
    __syncthreads();
    const long long lo = (long long)blockIdx.x * p.per_block;
    const long long hi = min(lo + p.per_block, p.V);
    int local[kNumTiers] = {0, 0, 0};
    for (long long u = lo + threadIdx.x; u < hi; u += blockDim.x) {
        int t = plan_tier(p, u);
        if (t < kNumTiers)
            local[t]++;
    }
    for (int t = 0; t < kNumTiers; t++)
        if (local[t])
            atomicAdd((unsigned long long *)&cnt[t], (unsigned long long)local[t]);
    __syncthreads();
    if (threadIdx.x < kNumTiers)
        p.block_counts[(long long)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// One workgroup: exclusive scan of the per-block counts of each tier into
// global write offsets; records the tier boundaries in ctl->list_begin.
template <typename OffT>
__global__ void __launch_bounds__(1024) plan_scan_kernel(PlanArgs<OffT> p, int nblocks)
{
    __shared__ long long part[1024];
    __shared__ long long base;
    if (threadIdx.x == 0)
        base = 0;
    __syncthreads();
    for (int t = 0; t < kNumTiers; t++) {
        if (threadIdx.x == 0)
            p.ctl->list_begin[t] = base;
        for (int c0 = 0; c0 < nblocks; c0 += blockDim.x) {
            const int i = c0 + threadIdx.x;
            long long x = i < nblocks ? p.block_counts[(long long)t * nblocks + i] : 0;
            part[threadIdx.x] = x;
            __syncthreads();
            for (int off = 1; off < (int)blockDim.x; off <<= 1) {
                long long y = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
                __syncthreads();
                part[threadIdx.x] += y;
                __syncthreads();
            }
            const long long incl = part[threadIdx.x];
            if (i < nblocks)
                p.block_counts[(long long)t * nblocks + i] = base + incl - x;
            __syncthreads();
            if (threadIdx.x == blockDim.x - 1)
                base += incl;
            __syncthreads();
        }
    }
    if (threadIdx.x == 0)
        p.ctl->list_begin[kNumTiers] = base;
}

template <typename OffT>
__global__ void __launch_bounds__(256) plan_scatter_kernel(PlanArgs<OffT> p, int nblocks)
{
    __shared__ int wcount[kNumTiers][4];
    __shared__ long long run[kNumTiers];
    const int wave = threadIdx.x / kWave, lane = lane_id();
    if (threadIdx.x < kNumTiers)
        run[threadIdx.x] = p.block_counts[(long long)threadIdx.x * nblocks + blockIdx.x];
    __syncthreads();
    const long long lo = (long long)blockIdx.x * p.per_block;
    const long long hi = min(lo + p.per_block, p.V);
    for (long long u0 = lo; u0 < hi; u0 += blockDim.x) {
        const long long u = u0 + threadIdx.x;
        const int tier = u < hi ? plan_tier(p, u) : kTierNone;
        int rank_in_wave[kNumTiers];
        for (int t = 0; t < kNumTiers; t++) {
            unsigned long long m = __ballot(tier == t);
            rank_in_wave[t] = __builtin_amdgcn_mbcnt_hi((unsigned int)(m >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((unsigned int)m, 0));
            if (lane == 0)
                wcount[t][wave] = __popcll(m);
        }
        __syncthreads();
        if (tier < kNumTiers) {
            long long pos = run[tier] + rank_in_wave[tier];
            for (int w = 0; w < wave; w++)
                pos += wcount[tier][w];
            p.list[pos] = (int32_t)u;
        }
        __syncthreads();
        if (threadIdx.x < kNumTiers) {
            long long s = 0;
            for (int w = 0; w < (int)(blockDim.x / kWave); w++)
                s += wcount[threadIdx.x][w];
            run[threadIdx.x] += s;
        }
        __syncthreads();
    }
}

__global__ void offsets_to_i32_kernel(const int64_t *src, int32_t *dst, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = (int32_t)src[i];
}

__global__ void fill_u32_kernel(uint32_t *dst, uint32_t value, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = value;
}

// ------------------------------------------------------------- wrappers
template <typename OffT>
hipError_t launch_plan(const PlanArgs<OffT> &p, int nblocks, hipStream_t stream)
{
    if (nblocks > 0) {
        hipLaunchKernelGGL(plan_count_kernel<OffT>, dim3(nblocks), dim3(256), 0, stream, p, nblocks);
    }
    hipLaunchKernelGGL(plan_scan_kernel<OffT>, dim3(1), dim3(1024), 0, stream, p, nblocks);
    if (nblocks > 0) {
        hipLaunchKernelGGL(plan_scatter_kernel<OffT>, dim3(nblocks), dim3(256), 0, stream, p, nblocks);
    }
    return hipGetLastError();
}

template <typename OffT>
hipError_t launch_count(const TcArgs<OffT> &a, int blocks, int threads, size_t lds_bytes,
                        bool instrument, hipStream_t stream)
{
    if (instrument)
        hipLaunchKernelGGL((tc_count_kernel<OffT, true>), dim3(blocks), dim3(threads), lds_bytes,
                           stream, a);
    else
        hipLaunchKernelGGL((tc_count_kernel<OffT, false>), dim3(blocks), dim3(threads), lds_bytes,
                           stream, a);
    return hipGetLastError();
}

template <typename OffT>
hipError_t count_occupancy(int threads, size_t lds_bytes, bool instrument, int *blocks_per_cu)
{
    if (instrument)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            blocks_per_cu, tc_count_kernel<OffT, true>, threads, lds_bytes);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,
                                                        tc_count_kernel<OffT, false>, threads,
                                                        lds_bytes);
}

hipError_t launch_offsets_to_i32(const int64_t *src, int32_t *dst, long long n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    long long g = min((n + 255) / 256, 65536ll);
    hipLaunchKernelGGL(offsets_to_i32_kernel, dim3((unsigned)g), dim3(256), 0, stream, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, long long n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    long long g = min((n + 255) / 256, 65536ll);
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)g), dim3(256), 0, stream, dst, value, n);
    return hipGetLastError();
}

template hipError_t launch_plan<int32_t>(const PlanArgs<int32_t> &, int, hipStream_t);
template hipError_t launch_plan<int64_t>(const PlanArgs<int64_t> &, int, hipStream_t);
template hipError_t launch_count<int32_t>(const TcArgs<int32_t> &, int, int, size_t, bool, hipStream_t);
template hipError_t launch_count<int64_t>(const TcArgs<int64_t> &, int, int, size_t, bool, hipStream_t);
template hipError_t count_occupancy<int32_t>(int, size_t, bool, int *);
template hipError_t count_occupancy<int64_t>(int, size_t, bool, int *);

}  // namespace trust
