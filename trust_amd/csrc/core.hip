// Dense core: the edges into a small set of high-degree vertices, counted
// with bit rows instead of hashed walks.  No counterpart in the reference;
// the count is the same.
//
// Let C be at most 32768 vertices with ids cid(x) in [0, |C|), and B(x) the
// row of N+(x) ∩ C over those ids.  A dense target is a v in C whose whole
// out-list lies in C, without repeats, with d+(v) >= t; a dense source is a
// u with d+(u) >= t whose core part has no repeats and holds a dense
// target.  For an edge u -> v from a dense source to a dense target every w
// in N+(u) ∩ N+(v) is in C, so
//
//     |N+(u) ∩ N+(v)| = popcount(B(u) & B(v)),
//
// which costs |C|/8 bytes of bit rows instead of min(d+(u), d+(v)) probes of
// 4-byte ids.  The walk index leaves these edges out (walk_index.hip, pass
// A) and one launch counts them.
//
// Choosing C.  Orientation runs from lower to higher degree (the degree the
// preprocessing ordered by), so the set of vertices above any such degree is
// closed under out-neighbours and all of its members with d+ >= t are dense
// targets.  Exact in-degrees would cost a
// global atomic per edge; an estimate (d+ plus 128x the in-count over every
// 128th entry of adj) picks the seed, the K0 highest estimates in a 64-bins-
// per-octave histogram, and the seed is then closed: each step adds the
// out-neighbours of the newest members, until nothing is added (at most
// kCoreSteps steps; a closure past 32768 vertices is retried from a halved
// seed, and below 256 the seed alone is kept).
// R-MAT's degrees come in classes (a vertex's expected degree depends on the
// bit pattern of its id): at scale 26 the 16k highest estimates close to
// 17.9k vertices whose edges carry 52% of the short-side walk's probes
// (43% with both ends' d+ >= 384), where the seed alone, cut inside one
// class by a noisy estimate, keeps 19%; sampling every 16th, 32nd, 64th or
// 128th entry closes to the same core.  The dense tests make the count exact
// for any C.
//
// Layout for the MI355X's eight XCDs: each row is cut into S slices of wd
// words (wd a power of two up to 32: a slice row is at most one 128-byte
// line) and the rows are stored slice-major, rows[s][cid][wd]; one slice of
// all rows is |C| x 128 bytes at most (2.3 MB for R-MAT 26's 17.9k-vertex
// core in 18 slices), within one XCD's 4 MB L2.  The index also stores each
// dense source's row B(u) (bsrc).  A work item is (dense source u, slice s):
// the wave reads u's slice of B(u), stages its dense targets' ids in LDS,
// and per target reads that target's slice row (16 bytes per lane) and adds
// the popcount of the AND.  Each slice has a queue; a wave serves its XCD's
// slices (x, x + 8, ...; hardware XCC_ID) first, then the others, so every
// item is counted once wherever the waves land and an XCD's L2 mostly
// serves its own slices.  (Slices of 36 words, two lines per row, measured
// 41 ms for R-MAT 26's dense edges.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "engine.h"
#include "wave_ops.h"

namespace trust {

constexpr int kCoreSeedMax = 16384;  // K0
constexpr int kCoreSeedMin = 256;
constexpr int kCoreMax = 32768;      // |C| (int16 ids; at most 32 slices of 32 words)
constexpr int kCoreMaxSlices = 32;
constexpr int kCoreGrabLg = 16;      // the lists kernel's grabs: 2^16 entries of adj
constexpr int kCoreSteps = 6;        // closure steps
// default t: R-MAT 26 fresh counts measured 253.6 ms at 256, 236.3 at 384,
// 236.8 at 512; R-MAT 24 47.1, 47.5, 51.8 ms
constexpr int kCoreMinDegree = 384;
constexpr int kCoreSample = 128;     // in-degree sampled every 128th entry
constexpr int kCoreBins = 2048;      // histogram: est < 64 exact, then 64 bins per octave
constexpr int kCorePart = 2048;      // dense targets per work item (longer lists are cut)
constexpr int kCoreThreads = 256;
constexpr int kCoreQueueStride = 32; // one 128-byte line per slice queue
// misc words: [0] items of long lists (front of recs), [1] items of short
// lists (back of recs), [2] closure additions, [3] bsrc slots taken, [4] rows
// with d+ >= t, [5] the lists kernel's grab queue, then the slice queues
constexpr int kCoreMiscWords = 64 + kCoreMaxSlices * kCoreQueueStride;

// A work item: dense targets pool[lo + t0, lo + t1) of the source whose
// row B(u) is bsrc[b].
template <typename OffT>
struct DenseRec {
    OffT lo;
    int32_t t0, t1, b;
};

__device__ __forceinline__ int core_bin(long long est)
{
    if (est < 64)
        return (int)max(est, 0ll);
    const int lg = 63 - __clzll((unsigned long long)est);
    return 64 + (lg - 6) * 64 + (int)((est >> (lg - 6)) & 63);
}

static int core_bin_host(long long est)
{
    if (est < 64)
        return (int)std::max(est, 0ll);
    int lg = 63;
    while (!((unsigned long long)est >> lg))
        lg--;
    return 64 + (lg - 6) * 64 + (int)((est >> (lg - 6)) & 63);
}

__global__ void __launch_bounds__(kCoreThreads)
core_sample_kernel(const int32_t *__restrict__ adj, long long E, int *__restrict__ din)
{
    const long long n = (E + kCoreSample - 1) / kCoreSample;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        atomicAdd(&din[adj[i * kCoreSample]], 1);
}

template <typename OffT>
__device__ __forceinline__ long long core_est(const OffT *beg, const int *din, long long x)
{
    const long long e = (long long)(beg[x + 1] - beg[x]) + (long long)kCoreSample * din[x];
    return min(e, (long long)INT32_MAX);
}

template <typename OffT>
__global__ void __launch_bounds__(kCoreThreads)
core_hist_kernel(const OffT *__restrict__ beg, long long V, const int *__restrict__ din,
                 unsigned *__restrict__ hist)
{
    __shared__ unsigned h[kCoreBins];
    for (int b = threadIdx.x; b < kCoreBins; b += blockDim.x)
        h[b] = 0;
    __syncthreads();
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < V;
         x += (long long)gridDim.x * blockDim.x)
        atomicAdd(&h[core_bin(core_est(beg, din, x))], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < kCoreBins; b += blockDim.x)
        if (h[b])
            atomicAdd(&hist[b], h[b]);
}

// Membership states in flag[x]: 0 outside, 1 member whose out-list has been
// added, 2 member to expand, 3 added by the running step.
template <typename OffT>
__global__ void __launch_bounds__(kCoreThreads)
core_seed_kernel(const OffT *__restrict__ beg, long long V, const int *__restrict__ din,
                 int *__restrict__ flag, int bstar)
{
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < V;
         x += (long long)gridDim.x * blockDim.x)
        flag[x] = core_bin(core_est(beg, din, x)) >= bstar ? 2 : 0;
}

// One closure step: a wave takes 64 vertices and expands each member in
// state 2 with the whole wave (its out-neighbours outside join in state 3).
template <typename OffT>
__global__ void __launch_bounds__(kCoreThreads)
core_expand_kernel(const OffT *__restrict__ beg, const int32_t *__restrict__ adj, long long V,
                   int *__restrict__ flag, unsigned *__restrict__ added)
{
    constexpr int kWaves = kCoreThreads / kWave;
    const int wave = threadIdx.x / kWave, lane = lane_id();
    const long long step = (long long)gridDim.x * kWaves * kWave;
    unsigned n_add = 0;
    for (long long x0 = ((long long)blockIdx.x * kWaves + wave) * kWave; x0 < V; x0 += step) {
        const long long x = x0 + lane;
        const int f = x < V ? flag[x] : 0;
        unsigned long long todo = __ballot(f == 2);
        if (f == 2)
            flag[x] = 1;
        while (todo) {
            const int i = __builtin_ctzll(todo);
            todo &= todo - 1;
            const long long lo = (long long)beg[x0 + i], hi = (long long)beg[x0 + i + 1];
            for (long long p = lo + lane; p < hi; p += kWave) {
                const int w = adj[p];
                if (flag[w] == 0 && atomicCAS(&flag[w], 0, 3) == 0)
                    n_add++;
            }
        }
    }
    n_add = (unsigned)wave_sum((unsigned long long)n_add);
    if (lane == 0 && n_add)
        atomicAdd(added, n_add);
}

// After a step: state 3 -> 2 (the next step expands them).  last: every
// member -> 1 and everything else -> 0 (the scan's input).
__global__ void __launch_bounds__(kCoreThreads)
core_promote_kernel(long long V, int *__restrict__ flag, int last)
{
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < V;
         x += (long long)gridDim.x * blockDim.x) {
        const int f = flag[x];
        if (last)
            flag[x] = f != 0 ? 1 : 0;
        else if (f == 3)
            flag[x] = 2;
    }
}

// flag (0/1) and its exclusive scan -> cid (int16, -1 outside) and coreV
__global__ void __launch_bounds__(kCoreThreads)
core_assign_kernel(long long V, const int *__restrict__ flag, const int *__restrict__ pos,
                   int16_t *__restrict__ cid, int *__restrict__ coreV)
{
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < V;
         x += (long long)gridDim.x * blockDim.x) {
        if (flag[x]) {
            cid[x] = (int16_t)pos[x];
            coreV[pos[x]] = (int)x;
        } else {
            cid[x] = -1;
        }
    }
}

// One workgroup per core vertex v: its Kb-bit row, kept (and v made a dense
// target) when d+(v) >= thr and every out-neighbour is a core vertex, once.
template <typename OffT>
__global__ void __launch_bounds__(kCoreThreads)
core_rows_kernel(const OffT *__restrict__ beg, const int32_t *__restrict__ adj,
                 const int16_t *__restrict__ cid, const int *__restrict__ coreV, int Kb, int S,
                 int thr, uint32_t *__restrict__ rows, uint32_t *__restrict__ dst_bits,
                 uint32_t *__restrict__ dense_cid)
{
    __shared__ uint32_t bits[kCoreMax / 32];
    __shared__ int bad;
    const int r = blockIdx.x;
    const int v = coreV[r];
    const OffT b0 = beg[v];
    const long long d = (long long)(beg[v + 1] - b0);
    if (d < thr || d > Kb)
        return;  // the whole workgroup
    const int words = Kb / 32;
    for (int i = threadIdx.x; i < words; i += blockDim.x)
        bits[i] = 0;
    if (threadIdx.x == 0)
        bad = 0;
    __syncthreads();
    for (long long j = threadIdx.x; j < d; j += blockDim.x) {
        const int c = cid[adj[(long long)b0 + j]];
        if (c < 0) {
            bad = 1;  // outside the core
            break;
        }
        const uint32_t m = 1u << (c & 31);
        if (atomicOr(&bits[c >> 5], m) & m) {
            bad = 1;  // a repeated neighbour
            break;
        }
    }
    __syncthreads();
    if (bad)
        return;
    const int wd = words / S;  // words of one slice
    for (int i = threadIdx.x; i < words; i += blockDim.x) {
        const int s = i / wd;
        rows[((size_t)s * Kb + r) * wd + (i - s * wd)] = bits[i];
    }
    if (threadIdx.x == 0) {
        atomicOr(&dst_bits[v >> 5], 1u << (v & 31));
        atomicOr(&dense_cid[r >> 5], 1u << (r & 31));
    }
}

// The dense sources: a wave takes 64 rows, and each row with d+ >= thr in
// turn: its core ids into an LDS bitmap, which checks for repeats and is the
// row B(u) (stored to bsrc for a dense source), and its dense targets' ids
// from the row's start of pool (pool has adj's indexing); then the row's
// work items (targets cut into kCorePart pieces; rows with more than
// kCorePart targets fill recs from the front, the others from the back, so
// the long ones are handed out first).
template <typename OffT>
__global__ void __launch_bounds__(kCoreThreads)
core_lists_kernel(const OffT *__restrict__ beg, const int32_t *__restrict__ adj, long long V,
                  const int16_t *__restrict__ cid, const uint32_t *__restrict__ dense_cid, int Kb,
                  int thr, uint16_t *__restrict__ pool, DenseRec<OffT> *__restrict__ recs,
                  long long cap, uint32_t *__restrict__ bsrc, long long nsrc_cap,
                  const long long *__restrict__ grab_row, long long ngrab,
                  unsigned *__restrict__ misc, uint32_t *__restrict__ src_bits)
{
    constexpr int kWaves = kCoreThreads / kWave;
    __shared__ uint32_t dc[kCoreMax / 32];
    __shared__ uint32_t seen_s[kWaves][kCoreMax / 32];
    const int words = Kb / 32;
    for (int i = threadIdx.x; i < words; i += blockDim.x)
        dc[i] = dense_cid[i];
    const int wave = threadIdx.x / kWave, lane = lane_id();
    uint32_t *seen = seen_s[wave];
    for (int i = lane; i < words; i += kWave)
        seen[i] = 0;
    __syncthreads();
    // R-MAT's long rows crowd at low ids: the grabs are equal ranges of
    // entries (the rows starting in 2^16 entries of adj), from a queue
    for (;;) {
        unsigned q = 0;
        if (lane == 0)
            q = atomicAdd(&misc[5], 1u);
        q = __builtin_amdgcn_readfirstlane(q);
        if ((long long)q >= ngrab)
            break;
        const long long r0 = grab_row[q], r1 = grab_row[q + 1];
      for (long long u0 = r0; u0 < r1; u0 += kWave) {
        const int n = (int)min((long long)kWave, r1 - u0);
        const long long d = lane < n ? (long long)(beg[u0 + lane + 1] - beg[u0 + lane]) : 0;
        unsigned long long cand = __ballot(d >= thr);
        // a bsrc slot for each candidate row of the group
        unsigned b0 = 0;
        if (lane == 0 && cand)
            b0 = atomicAdd(&misc[3], (unsigned)__popcll(cand));
        b0 = __builtin_amdgcn_readfirstlane(b0);
        while (cand) {
            const int i = __builtin_ctzll(cand);
            const unsigned b = b0++;
            cand &= cand - 1;
            const long long u = u0 + i;
            const long long lo = (long long)beg[u], hi = (long long)beg[u + 1];
            int nd = 0, dup = 0;
            constexpr int kW = 4;  // windows in flight
            for (long long p0 = lo; p0 < hi; p0 += kW * kWave) {
                int cw[kW];
#pragma unroll
                for (int k = 0; k < kW; k++) {
                    const long long at = p0 + k * kWave + lane;
                    cw[k] = at < hi ? adj[at] : -1;
                }
#pragma unroll
                for (int k = 0; k < kW; k++)
                    cw[k] = cw[k] >= 0 ? (int)cid[cw[k]] : -1;
#pragma unroll
                for (int k = 0; k < kW; k++) {
                    const int c = cw[k];
                    const bool core = c >= 0;
                    const bool dense = core && ((dc[c >> 5] >> (c & 31)) & 1u);
                    int rep = 0;
                    if (core) {
                        const uint32_t m = 1u << (c & 31);
                        rep = (atomicOr(&seen[c >> 5], m) & m) ? 1 : 0;
                    }
                    dup |= __ballot(rep) != 0ull;
                    const unsigned long long bd = __ballot(dense);
                    const int rd = (int)__builtin_amdgcn_mbcnt_hi(
                        (unsigned)(bd >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bd, 0));
                    if (dense)
                        pool[lo + nd + rd] = (uint16_t)c;
                    nd += __popcll(bd);
                }
            }
            wave_sync();
            if (!dup && nd > 0 && (long long)b < nsrc_cap) {
                uint32_t *row = bsrc + (size_t)b * words;
                for (int k = lane; k < words; k += kWave)
                    row[k] = seen[k];
            }
            wave_sync();
            for (int k = lane; k < words; k += kWave)
                seen[k] = 0;
            wave_sync();
            if (dup || nd == 0 || (long long)b >= nsrc_cap)
                continue;
            const int parts = (nd + kCorePart - 1) / kCorePart;
            if (lane < parts) {
                long long slot;
                if (parts > 1)
                    slot = (long long)atomicAdd(&misc[0], 1u);
                else
                    slot = cap - 1 - (long long)atomicAdd(&misc[1], 1u);
                DenseRec<OffT> rc;
                rc.lo = (OffT)lo;
                rc.t0 = lane * kCorePart;
                rc.t1 = min(nd, (lane + 1) * kCorePart);
                rc.b = (int32_t)b;
                recs[slot] = rc;
            }
            if (lane == 0)
                atomicOr(&src_bits[u >> 5], 1u << (u & 31));
        }
      }
    }
}

// grab_row[q] = the first row starting at or after entry q 2^kCoreGrabLg
// (q <= ngrab; grab_row[ngrab] = V).
template <typename OffT>
__global__ void __launch_bounds__(kCoreThreads)
core_grabs_kernel(const OffT *__restrict__ beg, long long V, long long ngrab,
                  long long *__restrict__ grab_row)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q > ngrab)
        return;
    if (q == ngrab) {
        grab_row[q] = V;
        return;
    }
    const long long at = q << kCoreGrabLg;
    long long lo = 0, hi = V;  // the first u in [0, V] with beg[u] >= at
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if ((long long)beg[mid] >= at)
            hi = mid;
        else
            lo = mid + 1;
    }
    grab_row[q] = lo;
}

// Rows with d+ >= thr (an upper bound on the dense sources), into misc[4].
template <typename OffT>
__global__ void __launch_bounds__(kCoreThreads)
core_rows_over_kernel(const OffT *__restrict__ beg, long long V, int thr, unsigned *__restrict__ n)
{
    unsigned k = 0;
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < V;
         x += (long long)gridDim.x * blockDim.x)
        k += (long long)(beg[x + 1] - beg[x]) >= thr;
    k = (unsigned)wave_sum((unsigned long long)k);
    if (lane_id() == 0 && k)
        atomicAdd(n, k);
}

// The counting launch.  A wave's item (r, s): u's slice s of B(u) from bsrc
// (16 bytes per lane) and the item's dense targets' ids staged in LDS, then
// one slice row per target: rows of wd = Kb/(32 S) words read as 16-byte
// pieces, wd/4 lanes per target, 64/(wd/4) targets at a time (8 for 32-word
// slices), kIn per lane in flight.
template <typename OffT>
__global__ void __launch_bounds__(kCoreThreads)
core_dense_kernel(const DenseRec<OffT> *__restrict__ recs, long long cap,
                  const unsigned *__restrict__ misc_in, unsigned *__restrict__ queues,
                  const uint16_t *__restrict__ pool, const uint32_t *__restrict__ bsrc,
                  const uint32_t *__restrict__ rows, int Kb, int S, TcCtl *__restrict__ ctl)
{
    constexpr int kWaves = kCoreThreads / kWave;
    constexpr int kIn = 8;
    __shared__ uint16_t tg_s[kWaves][kCorePart];
    const int wave = threadIdx.x / kWave, lane = lane_id();
    uint16_t *tg = tg_s[wave];
    const int words = Kb >> 5;                        // of a whole row
    const int wd = words / S;                         // of a slice row: 4, 8, 16 or 32
    const int L = wd >> 2;                            // lanes per target, 4 words each
    const int G = kWave / L;                          // targets per step
    const int g = lane / L, w4 = lane - g * L;        // lanes with g >= G idle
    const long long n_long = misc_in[0], n_short = misc_in[1];
    const long long n = n_long + n_short;
    unsigned long long acc = 0;
    // slices in the order: this XCD's (x, x + 8, ...), then the others
    const int x = xcc_id() & (kXcds - 1);
    const int n_own = x < S ? (S - 1 - x) / kXcds + 1 : 0;
    auto slice_at = [&](int i) {
        if (i < n_own)
            return x + kXcds * i;
        int k = i - n_own;
        for (int c = 0;; c++)
            if (c % kXcds != x && k-- == 0)
                return c;
    };
    int vis = 0, s = slice_at(0);
    unsigned qn = 0;  // the next item, taken while this one is counted
    if (lane == 0)
        qn = atomicAdd(&queues[s * kCoreQueueStride], 1u);
    for (;;) {
        const unsigned q = __builtin_amdgcn_readfirstlane(qn);
        if ((long long)q >= n) {
            if (++vis == S)
                break;
            s = slice_at(vis);
            if (lane == 0)
                qn = atomicAdd(&queues[s * kCoreQueueStride], 1u);
            continue;
        }
        const long long ri = (long long)q < n_long ? (long long)q : cap - 1 - ((long long)q - n_long);
        const DenseRec<OffT> rc = recs[ri];
        if (lane == 0)
            qn = atomicAdd(&queues[s * kCoreQueueStride], 1u);
        const uint4 bu = ((const uint4 *)(bsrc + (size_t)rc.b * words + (size_t)s * wd))[w4];
        if (__ballot((bu.x | bu.y | bu.z | bu.w) != 0u) == 0ull)
            continue;  // no core neighbour of u in this slice
        const int nt = rc.t1 - rc.t0;
        const uint16_t *src = pool + (long long)rc.lo + rc.t0;
        for (int j = lane; j < nt; j += kWave)
            tg[j] = src[j];
        wave_sync();
        const uint4 *base = (const uint4 *)(rows + (size_t)s * Kb * wd) + w4;
        int t = g < G ? g : nt;
        for (; t + (kIn - 1) * G < nt; t += kIn * G) {
            uint4 x[kIn];
#pragma unroll
            for (int k = 0; k < kIn; k++)
                x[k] = base[(size_t)tg[t + k * G] * L];
#pragma unroll
            for (int k = 0; k < kIn; k++)
                acc += (unsigned long long)(__popc(x[k].x & bu.x) + __popc(x[k].y & bu.y) +
                                            __popc(x[k].z & bu.z) + __popc(x[k].w & bu.w));
        }
        for (; t < nt; t += G) {
            const uint4 x = base[(size_t)tg[t] * L];
            acc += (unsigned long long)(__popc(x.x & bu.x) + __popc(x.y & bu.y) +
                                        __popc(x.z & bu.z) + __popc(x.w & bu.w));
        }
        wave_sync();  // before the next item overwrites tg
    }
    acc = wave_sum(acc);
    if (lane == 0 && acc)
        atomicAdd(&ctl->count, acc);
}

// ------------------------------------------------------------------ host
void core_config(const Ctx *ctx, const trust_params &p, int *K, int *thr, bool *automatic)
{
    *K = 0;
    *thr = 0;
    *automatic = false;
    if (p.core_vertices < 0 || p.total_process > 1 || ctx->batch_ngraphs > 0)
        return;
    const int t = p.core_min_degree > 0 ? p.core_min_degree : kCoreMinDegree;
    if (p.core_vertices > 0) {
        *K = p.core_vertices;
    } else if (ctx->g.E >= (1ll << 23) && ctx->g.max_deg >= 2ll * t) {
        // (a graph without out-degrees of 2t cannot have the heavy tail the
        // automatic core asks for below: Chung-Lu 3.8M's fresh count skips
        // the check, 0.1 ms of its 1.8 ms)
        *K = kCoreSeedMax;
        *automatic = true;
    }
    if (*K)
        *thr = t;
}

void core_drop(Ctx *ctx)
{
    CoreIndex &c = ctx->core;
    c.built = c.active = false;
    c.K = c.thr = c.nC = c.Kb = c.S = 0;
    c.dense_edges = c.dense_probes = 0;
}

void core_release(Ctx *ctx)
{
    CoreIndex &c = ctx->core;
    core_drop(ctx);
    for (DevBuf *b : {&c.flag, &c.cid, &c.pos, &c.coreV, &c.bits, &c.rows, &c.pool, &c.recs,
                      &c.bsrc, &c.grabs, &c.misc, &c.hist})
        b->release();
}

template <typename OffT>
static int core_build_t(Ctx *ctx)
{
    CoreIndex &c = ctx->core;
    const DevCsr &g = ctx->g;
    const long long V = g.V, E = g.E;
    const int K0 = c.want_K, thr = c.want_thr;
    const OffT *beg = (const OffT *)g.beg.p;
    const int32_t *adj = (const int32_t *)g.adj.p;
    hipStream_t st = ctx->stream;
    c.words = (V + 31) / 32;
    c.cap = E / std::max(thr, 1) + E / kCorePart + 2;
    int rc;
    if ((rc = c.flag.ensure(sizeof(int) * (size_t)V)) ||
        (rc = c.pos.ensure(sizeof(int) * (size_t)(V + 1))) ||
        (rc = c.cid.ensure(sizeof(int16_t) * (size_t)V)) ||
        (rc = c.coreV.ensure(sizeof(int) * (size_t)kCoreMax)) ||
        (rc = c.bits.ensure(sizeof(uint32_t) * (size_t)(2 * c.words + kCoreMax / 32))) ||
        (rc = c.misc.ensure(sizeof(unsigned) * kCoreMiscWords)) ||
        (rc = c.hist.ensure(sizeof(unsigned) * kCoreBins)) ||
        (rc = ctx->side_bsum.ensure(sizeof(long long) * (size_t)(scan_blocks(V) + 1))))
        return rc;
    int *flag = (int *)c.flag.p;
    int *din = (int *)c.pos.p;  // the sampled in-counts, until the scan
    uint32_t *bits = (uint32_t *)c.bits.p;
    uint32_t *dst_bits = bits, *src_bits = bits + c.words, *dense_cid = bits + 2 * c.words;
    unsigned *misc = (unsigned *)c.misc.p;
    const int blocks = 4 * ctx->num_cus;
    HIPCHK(launch_fill_u32((uint32_t *)din, 0u, V, st));
    HIPCHK(launch_fill_u32((uint32_t *)c.hist.p, 0u, kCoreBins, st));
    HIPCHK(launch_fill_u32(bits, 0u, 2 * c.words + kCoreMax / 32, st));
    HIPCHK(launch_fill_u32(misc, 0u, kCoreMiscWords, st));
    hipLaunchKernelGGL(core_sample_kernel, dim3(blocks), dim3(kCoreThreads), 0, st, adj, E, din);
    hipLaunchKernelGGL(core_hist_kernel<OffT>, dim3(blocks), dim3(kCoreThreads), 0, st, beg, V,
                       (const int *)din, (unsigned *)c.hist.p);
    HIPCHK(hipGetLastError());
    std::vector<unsigned> h(kCoreBins);
    HIPCHK(hipMemcpyAsync(h.data(), c.hist.p, sizeof(unsigned) * kCoreBins,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    // the seed: the top bins, at most k0 vertices, estimates of at least thr
    const int bmin = core_bin_host(thr);
    auto seed_cut = [&](long long k0, long long *take) {
        int bstar = kCoreBins;
        *take = 0;
        for (int b = kCoreBins - 1; b >= bmin; b--) {
            if (*take + h[(size_t)b] > (unsigned long long)k0)
                break;
            *take += h[(size_t)b];
            bstar = b;
        }
        return bstar;
    };
    long long take = 0;
    int bstar = seed_cut(K0, &take);
    c.K = K0;
    c.thr = thr;
    c.built = true;
    c.active = false;
    // The automatic core needs a heavy tail: its seed's smallest estimates at
    // least 2t (R-MAT 26: about 12k; a Chung-Lu graph capped at degree 800
    // has no such vertices, and takes the reference walk anyway)
    if (take == 0 || (c.want_auto && bstar < core_bin_host(2ll * thr)))
        return TRUST_OK;
    // The closure, while it fits; a seed whose closure passes kCoreMax
    // vertices is halved and closed again (down to kCoreSeedMin), and past
    // that the seed alone is the core (its targets are whichever lists lie
    // inside it).
    const long long xblocks =
        std::min<long long>((V + kCoreThreads - 1) / kCoreThreads, 8ll * ctx->num_cus);
    long long members = take;
    for (long long k0 = K0;;) {
        hipLaunchKernelGGL(core_seed_kernel<OffT>, dim3(blocks), dim3(kCoreThreads), 0, st, beg,
                           V, (const int *)din, flag, bstar);
        HIPCHK(hipGetLastError());
        members = take;
        bool closed = false;
        for (int step = 0; step < kCoreSteps && !closed && members <= kCoreMax; step++) {
            HIPCHK(launch_fill_u32(misc + 2, 0u, 1, st));
            hipLaunchKernelGGL(core_expand_kernel<OffT>, dim3((unsigned)std::max(xblocks, 1ll)),
                               dim3(kCoreThreads), 0, st, beg, adj, V, flag, misc + 2);
            hipLaunchKernelGGL(core_promote_kernel, dim3(blocks), dim3(kCoreThreads), 0, st, V,
                               flag, 0);
            HIPCHK(hipGetLastError());
            unsigned added = 0;
            HIPCHK(hipMemcpyAsync(&added, misc + 2, sizeof(added), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            members += added;
            closed = added == 0;
        }
        if (members <= kCoreMax)
            break;
        long long t2 = 0;
        const int b2 = k0 / 2 >= kCoreSeedMin ? seed_cut(k0 / 2, &t2) : kCoreBins;
        if (t2 == 0) {  // no smaller seed: the seed alone
            hipLaunchKernelGGL(core_seed_kernel<OffT>, dim3(blocks), dim3(kCoreThreads), 0, st,
                               beg, V, (const int *)din, flag, bstar);
            members = take;
            break;
        }
        k0 /= 2;
        bstar = b2;
        take = t2;
    }
    hipLaunchKernelGGL(core_promote_kernel, dim3(blocks), dim3(kCoreThreads), 0, st, V, flag, 1);
    HIPCHK(hipGetLastError());
    // bit rows: S slices of wd words, wd a power of two in [4, 32]: a slice
    // row is at most one 128-byte line (a lane reads 16 bytes of it)
    const int nC = (int)members;
    int wd = 4;
    while (wd < 32 && wd * 32 * kXcds < nC)
        wd <<= 1;
    const int S = (nC + 32 * wd - 1) / (32 * wd);
    const int Kb = wd * 32 * S;
    c.nC = nC;
    c.Kb = Kb;
    c.S = S;
    // the dense sources' rows B(u): at most one per row with d+ >= thr
    hipLaunchKernelGGL(core_rows_over_kernel<OffT>, dim3(blocks), dim3(kCoreThreads), 0, st, beg,
                       V, thr, misc + 4);
    HIPCHK(hipGetLastError());
    unsigned over = 0;
    HIPCHK(hipMemcpyAsync(&over, misc + 4, sizeof(over), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    c.nsrc_cap = std::max<long long>(over, 1);
    if ((rc = c.rows.ensure((size_t)Kb * (size_t)Kb / 8)) ||
        (rc = c.pool.ensure(sizeof(uint16_t) * (size_t)E)) ||
        (rc = c.recs.ensure(sizeof(DenseRec<OffT>) * (size_t)c.cap)) ||
        (rc = c.bsrc.ensure((size_t)c.nsrc_cap * (size_t)Kb / 8)) ||
        (rc = c.grabs.ensure(sizeof(long long) * (size_t)((E >> kCoreGrabLg) + 2))))
        return rc;
    HIPCHK(launch_scan<int32_t>(flag, nullptr, V, (long long *)ctx->side_bsum.p, (int32_t *)c.pos.p,
                                st));
    hipLaunchKernelGGL(core_assign_kernel, dim3(blocks), dim3(kCoreThreads), 0, st, V,
                       (const int *)flag, (const int *)c.pos.p, (int16_t *)c.cid.p,
                       (int *)c.coreV.p);
    hipLaunchKernelGGL(core_rows_kernel<OffT>, dim3((unsigned)nC), dim3(kCoreThreads), 0, st, beg,
                       adj, (const int16_t *)c.cid.p, (const int *)c.coreV.p, Kb, S, thr,
                       (uint32_t *)c.rows.p, dst_bits, dense_cid);
    const long long ngrab = (E >> kCoreGrabLg) + 1;
    hipLaunchKernelGGL(core_grabs_kernel<OffT>,
                       dim3((unsigned)((ngrab + kCoreThreads) / kCoreThreads)), dim3(kCoreThreads),
                       0, st, beg, V, ngrab, (long long *)c.grabs.p);
    hipLaunchKernelGGL(core_lists_kernel<OffT>, dim3((unsigned)std::max(xblocks, 1ll)),
                       dim3(kCoreThreads), 0, st, beg, adj, V, (const int16_t *)c.cid.p,
                       (const uint32_t *)dense_cid, Kb, thr, (uint16_t *)c.pool.p,
                       (DenseRec<OffT> *)c.recs.p, c.cap, (uint32_t *)c.bsrc.p, c.nsrc_cap,
                       (const long long *)c.grabs.p, ngrab, misc, src_bits);
    HIPCHK(hipGetLastError());
    c.active = true;
    return TRUST_OK;
}

int core_build(Ctx *ctx)
{
    CoreIndex &c = ctx->core;
    core_drop(ctx);
    const int K = c.want_K;
    if (K == 0 || ctx->g.V <= 0 || ctx->g.E <= 0) {
        c.built = true;
        return TRUST_OK;
    }
    if (K < kCoreSeedMin || K > kCoreSeedMax || (K & (K - 1)))
        return fail(TRUST_ERR_ARG, "core_vertices %d: a power of two in [%d, %d]", K,
                    kCoreSeedMin, kCoreSeedMax);
    if (c.want_thr < 2)
        return fail(TRUST_ERR_ARG, "core_min_degree %d < 2", c.want_thr);
    return ctx->g.off64 ? core_build_t<int64_t>(ctx) : core_build_t<int32_t>(ctx);
}

int core_enqueue(Ctx *ctx, TcCtl *ctl, hipStream_t s)
{
    CoreIndex &c = ctx->core;
    if (!c.active)
        return TRUST_OK;
    unsigned *misc = (unsigned *)c.misc.p;
    HIPCHK(launch_fill_u32(misc + 64, 0u, kCoreMaxSlices * kCoreQueueStride, s));
    const int blocks = 8 * ctx->num_cus;
    if (ctx->g.off64)
        hipLaunchKernelGGL(core_dense_kernel<int64_t>, dim3(blocks), dim3(kCoreThreads), 0, s,
                           (const DenseRec<int64_t> *)c.recs.p, c.cap, (const unsigned *)misc,
                           misc + 64, (const uint16_t *)c.pool.p, (const uint32_t *)c.bsrc.p,
                           (const uint32_t *)c.rows.p, c.Kb, c.S, ctl);
    else
        hipLaunchKernelGGL(core_dense_kernel<int32_t>, dim3(blocks), dim3(kCoreThreads), 0, s,
                           (const DenseRec<int32_t> *)c.recs.p, c.cap, (const unsigned *)misc,
                           misc + 64, (const uint16_t *)c.pool.p, (const uint32_t *)c.bsrc.p,
                           (const uint32_t *)c.rows.p, c.Kb, c.S, ctl);
    HIPCHK(hipGetLastError());
    return TRUST_OK;
}

}  // namespace trust
