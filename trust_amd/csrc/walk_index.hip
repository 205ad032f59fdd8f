// The short-side walk index, built as a destination-binned transpose.
//
// Row x of the walk lists first the v in N+(x) with d+(v) <= d+(x) (its
// out-part), then the u with u -> x and 2 <= d+(u) < d+(x) (its in-part);
// an edge from a vertex with d+ < 2 closes no triangle at that edge and is
// left out.  Each entry is stored as the {start, length} of the list it
// probes (WalkEntry), which the count reads in place of a vertex id.
//
// The in-parts are a transpose of most of the CSR (70-78% of R-MAT's
// edges).  Scattering each in-entry straight to its row needs a global
// atomic per entry to count the row and another to claim a slot: device-
// scope atomics are performed beyond the per-XCD L2, and at R-MAT scale 26
// the two passes ran at about 25 G atomics/s and wrote 6-9x their bytes with
// random 8-byte stores (89 ms of a 101 ms build).  Here the only atomics on
// the hot path are LDS atomics, and every global store lands either in a
// run that one workgroup owns or inside one bin's rows:
//
//   A   count    one pass over the CSR, 64 rows per wave in coalesced
//                64-wide windows: both walks' probe totals (which decide the
//                walk), each row's kept out-entries (per-wave LDS counters,
//                one row per lane) and, per workgroup, an LDS histogram of
//                in-entries by target bin (kBinSize consecutive targets);
//                the histograms are stored bin-major, M[bin][workgroup]
//   scan         M -> O, each workgroup's run in each bin's region; kept ->
//                the out-parts' offsets in a row-ordered temporary
//   B   scatter  the same pass again with the same row-to-workgroup
//                assignment: in-entry records {start, length, target} go to
//                the workgroup's run of their bin (LDS cursors), out-entries
//                to the temporary
//   C1  rank     one workgroup per slice of a bin, LDS counters for the
//                bin's targets: each record's rank among its target's
//                records, and the targets' in-counts (plain stores when the
//                bin is one slice, one atomic per target and slice otherwise)
//   scan         kept + in-count -> the row offsets begX
//   C2  place    each record to begX[v] + kept[v] + (its slice's base) + rank
//   D   copy     the out-parts to the front of their rows
//
// Graphs with more than kMaxBins bins keep the atomic build (tc_kernels.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>

#include "engine.h"
#include "wave_ops.h"

namespace trust {

// 2^13 targets per bin: a bin's rows (about 1.6 MB of entries at R-MAT
// scale 26) stay in one XCD's 4 MB L2 while C2 fills them, with room for
// the next bin (2^14: 44.0 ms for the whole build at scale 26, 2^13: 40.6)
constexpr int kBinLg = 13;
constexpr int kBinSize = 1 << kBinLg;  // targets per bin (C1's LDS counters: 32 KB)
constexpr int kMaxBins = 16384;        // V <= 2^27 (pass B's LDS cursors: <= 64 KB, 128 with int64)
constexpr int kWin = 4;                // 64-wide windows per step of passes A and B
constexpr int kPassThreads = 1024;
constexpr int kPassWaves = kPassThreads / kWave;
constexpr int kRankThreads = 1024;
constexpr int kPlaceThreads = 1024;
constexpr int kScanPer = 8192;         // elements per scan workgroup (1024 threads x 8)

// An in-entry on its way to row v: the list it probes is adj[start, start + deg).
template <typename OffT>
struct InRec {
    OffT start;
    int32_t deg;
    int32_t v;
};
// A tile walk's in-entry on its way to row v (a vertex of the column class):
// id is the row of the probe tile it names.
struct IdRec {
    int32_t id;
    int32_t v;
};

// What the walk stores.  The resident graph's walk stores each entry as the
// list it probes (WalkEntry).  A tile walk (i,k) serves the jobs (i,j,k) of
// every j, whose probe tiles differ, so it stores ids.
template <typename OffT, bool kTile>
struct WiFmt {
    using Out = WalkEntry<OffT>;
    using Rec = InRec<OffT>;
};
template <typename OffT>
struct WiFmt<OffT, true> {
    using Out = int32_t;
    using Rec = IdRec;
};

template <typename OffT>
using Cursor = typename std::conditional<sizeof(OffT) == 4, unsigned int, unsigned long long>::type;

template <typename OffT>
static size_t pass_lds_bytes(int nbins)
{
    const size_t bins = ((size_t)nbins * sizeof(Cursor<OffT>) + 7) & ~(size_t)7;
    return bins + (size_t)kPassWaves * (kWave * kWin + kWave) * sizeof(int);
}

template <typename OffT, bool kTile>
struct WiPassArgs {
    const OffT *beg;
    const int32_t *adj;
    long long V;  // rows
    int nbins;
    // tiles: global out-degree of each row's vertex and of each column's
    // vertex (the resident graph takes both from beg)
    const int32_t *degR, *degC;
    int *kept;                    // A: out-entries per row
    int *M;                       // A: in-entries per bin and workgroup, bin-major
    unsigned long long *stats;    // A: probes of both walks (resident graph)
    // resident graph: d+ of each entry's vertex (edges of rows with d+ >= 2),
    // written by A, read by B in place of a second pair of offset lookups;
    // -1 marks an edge the dense core counts
    int32_t *dvE;
    // the dense core's bitmaps (resident graph, null without a core): an
    // edge u -> v with u a dense source and v a dense target is counted by
    // bit rows and stays out of the walk
    const uint32_t *core_src, *core_dst;
    // resident graph, when every d+ < 2^15: d+(v) | (v a dense target) << 15,
    // the one 2-byte gather pass A makes per entry (null: the offsets and the
    // target map instead)
    const uint16_t *dtab;
    const long long *O;           // B: each workgroup's first record in each bin
    const OffT *outBeg;           // B: each row's first out-entry in tmpOut
    typename WiFmt<OffT, kTile>::Out *tmpOut;
    typename WiFmt<OffT, kTile>::Rec *recs;
};

// Passes A (kPass 0) and B (kPass 1).  Grid-stride over 64-row groups by
// global wave id: both passes give every row to the same workgroup, which
// is what lets pass B write at the offsets pass A's histograms produced.
template <typename OffT, int kPass, bool kTile>
__global__ void __launch_bounds__(kPassThreads) wi_pass_kernel(WiPassArgs<OffT, kTile> a)
{
    using CurT = Cursor<OffT>;
    extern __shared__ __align__(8) unsigned char smem[];
    CurT *bins = (CurT *)smem;  // A: in-entries per bin; B: the next free record of each bin
    int *words = (int *)(smem + ((a.nbins * sizeof(CurT) + 7) & ~(size_t)7));
    const int wave = threadIdx.x / kWave, lane = lane_id();
    int *mark = words + wave * (kWave * kWin);
    int *rowcnt = words + kPassWaves * kWave * kWin + wave * kWave;
    const int G = (int)gridDim.x;
    const OffT *__restrict__ beg = a.beg;
    const int32_t *__restrict__ adj = a.adj;
    const long long V = a.V;
    for (int b = threadIdx.x; b < a.nbins; b += blockDim.x) {
        if constexpr (kPass == 0)
            bins[b] = 0;
        else
            bins[b] = (CurT)a.O[(long long)b * G + blockIdx.x];
    }
    for (int k = 0; k < kWin; k++)
        mark[k * kWave + lane] = 0;
    __syncthreads();
    unsigned long long w_ref = 0, w_short = 0, w_dense = 0, n_dense = 0;
    const long long step = (long long)G * kPassWaves * kWave;
    for (long long u0 = ((long long)blockIdx.x * kPassWaves + wave) * kWave; u0 < V; u0 += step) {
        const int n = (int)min((long long)kWave, V - u0);
        // lane i < n holds row u0 + i; lane n (if any) the end of the range
        const OffT s = beg[u0 + min(lane, n)];
        const int d = lane < n ? (int)(beg[u0 + lane + 1] - s) : 0;
        int rdeg = d;  // the side rule's out-degree of the row's vertex
        if constexpr (kTile)
            rdeg = lane < n ? a.degR[u0 + lane] : 0;
        OffT ob = 0;  // the row's first out-entry in the temporary
        if constexpr (kPass == 1)
            ob = lane < n ? a.outBeg[u0 + lane] : (OffT)0;
        int src = 0;  // the row's vertex is a dense source (pass A's test)
        if constexpr (kPass == 0 && !kTile)
            if (a.core_src && lane < n)
                src = (int)((a.core_src[(u0 + lane) >> 5] >> ((u0 + lane) & 31)) & 1u);
        const long long e0 = (long long)readlane(s, 0);
        const long long e1 = n == kWave ? (long long)readlane(s, 63) + readlane(d, 63)
                                        : (long long)readlane(s, n);
        rowcnt[lane] = 0;
        wave_sync();
        int carry = 0;
        for (long long p = e0; p < e1; p += kWave * kWin) {
            // a non-empty row starting inside this span marks its position
            if (d > 0 && (long long)s >= p && (long long)s < p + kWave * kWin)
                mark[(int)((long long)s - p)] = lane;
            wave_sync();
            int own[kWin], v[kWin], dvk[kWin];
#pragma unroll
            for (int k = 0; k < kWin; k++) {
                own[k] = max(wave_scan_max(mark[k * kWave + lane]), carry);
                carry = readlane(own[k], 63);
                const long long pos = p + k * kWave + lane;
                v[k] = pos < e1 ? adj[pos] : -1;
                dvk[k] = 0;
                if constexpr (kPass == 1 && !kTile)
                    dvk[k] = pos < e1 ? a.dvE[pos] : 0;
            }
            wave_sync();
#pragma unroll
            for (int k = 0; k < kWin; k++)
                mark[k * kWave + lane] = 0;
#pragma unroll
            for (int k = 0; k < kWin; k++) {
                const int du = shfl_idx(rdeg, own[k]);
                int srcu = 0;
                if constexpr (kPass == 0 && !kTile)
                    if (a.core_src)
                        srcu = shfl_idx(src, own[k]);
                OffT su = 0, obu = 0;  // shuffled while every lane is active
                if constexpr (kPass == 1) {
                    if constexpr (!kTile)
                        su = shfl_idx(s, own[k]);
                    obu = shfl_idx(ob, own[k]);
                }
                if (v[k] >= 0 && du >= 2) {
                    OffT bv = 0;
                    int dv, dstk = 0;
                    if constexpr (kTile) {
                        dv = a.degC[v[k]];
                    } else if constexpr (kPass == 0) {
                        if (a.dtab) {
                            const unsigned x = a.dtab[v[k]];
                            dv = (int)(x & 0x7fffu);
                            dstk = (int)(x >> 15);
                        } else {
                            bv = beg[v[k]];
                            dv = (int)(beg[v[k] + 1] - bv);
                        }
                    } else {
                        dv = dvk[k];  // an out-entry looks up its list start below
                    }
                    if constexpr (kPass == 0) {
                        if constexpr (!kTile) {
                            w_ref += (unsigned long long)dv;
                            w_short += (unsigned long long)min(du, dv);
                            if (srcu && (a.dtab ? dstk != 0
                                                : ((a.core_dst[v[k] >> 5] >> (v[k] & 31)) & 1u))) {
                                w_dense += (unsigned long long)min(du, dv);
                                n_dense++;
                                a.dvE[p + k * kWave + lane] = -1;
                                continue;
                            }
                            a.dvE[p + k * kWave + lane] = dv;
                        }
                        if (dv <= du)
                            atomicAdd(rowcnt + own[k], 1);
                        else
                            atomicAdd(bins + (v[k] >> kBinLg), (CurT)1);
                    } else {
                        if constexpr (!kTile)
                            if (dv < 0)
                                continue;  // counted by the dense core
                        if (dv <= du) {
                            const int slot = atomicAdd(rowcnt + own[k], 1);
                            if constexpr (kTile)
                                a.tmpOut[obu + (OffT)slot] = v[k];
                            else
                                a.tmpOut[obu + (OffT)slot] = WalkEntry<OffT>{beg[v[k]], dv};
                        } else {
                            const CurT at = atomicAdd(bins + (v[k] >> kBinLg), (CurT)1);
                            if constexpr (kTile)
                                a.recs[at] = IdRec{(int32_t)(u0 + own[k]), v[k]};
                            else
                                a.recs[at] = InRec<OffT>{su, du, v[k]};
                        }
                    }
                }
            }
            wave_sync();
        }
        if constexpr (kPass == 0) {
            if (lane < n)
                a.kept[u0 + lane] = rowcnt[lane];
        }
        wave_sync();
    }
    if constexpr (kPass == 0) {
        __syncthreads();
        for (int b = threadIdx.x; b < a.nbins; b += blockDim.x)
            a.M[(long long)b * G + blockIdx.x] = (int)bins[b];
        if constexpr (!kTile) {
            w_ref = wave_sum(w_ref);
            w_short = wave_sum(w_short);
            w_dense = wave_sum(w_dense);
            n_dense = wave_sum(n_dense);
            if (lane == 0 && w_ref) {
                atomicAdd(a.stats, w_ref);
                atomicAdd(a.stats + 1, w_short);
                if (n_dense) {
                    atomicAdd(a.stats + 2, w_dense);
                    atomicAdd(a.stats + 3, n_dense);
                }
            }
        }
    }
}

// C1: ranks within one slice; in-counts of the bin's targets.
template <typename Rec>
__global__ void __launch_bounds__(kRankThreads)
wi_rank_kernel(const WiSlice *__restrict__ sl, const Rec *__restrict__ recs,
               int *__restrict__ rel, int *__restrict__ incnt, int *__restrict__ sbase)
{
    __shared__ int cnt[kBinSize];
    const WiSlice s = sl[blockIdx.x];
    for (int t = threadIdx.x; t < kBinSize; t += blockDim.x)
        cnt[t] = 0;
    __syncthreads();
    // one round of blockDim records at a time: a target's ranks in round r all
    // precede its ranks in round r + 1, so the records of any whole rounds
    // (C2's chunks) hold a contiguous rank range of each target
    for (long long r0 = s.lo; r0 < s.hi; r0 += blockDim.x) {
        const long long i = r0 + threadIdx.x;
        if (i < s.hi)
            rel[i] = atomicAdd(&cnt[recs[i].v & (kBinSize - 1)], 1);
        __syncthreads();
    }
    const long long v0 = (long long)s.bin << kBinLg;
    for (int t = threadIdx.x; t < kBinSize; t += blockDim.x) {
        const int c = cnt[t];
        if (c == 0)
            continue;
        if (s.multi < 0)
            incnt[v0 + t] = c;  // the bin's only slice
        else
            sbase[(long long)s.multi * kBinSize + t] = atomicAdd(&incnt[v0 + t], c);
    }
}

// C2: every record to its slot in its target's in-part: at begX[v] +
// kept[v] (the row's out-part comes first) for the resident graph, at
// begX[v] for a tile walk, whose in-rows are a CSR of their own.
//
// Stored in record order, the slots land anywhere in the bin's rows, one
// 8-byte entry at a time: random stores at about 20 G/s (39 ms at R-MAT
// scale 26).  So a workgroup takes a chunk of whole C1 rounds, whose records
// of any one target hold consecutive slots, sorts them by target in LDS (a
// counting sort over the bin's 8k targets) and stores them in that order:
// the records of a target go out together, to consecutive addresses.  The
// chunks of a bin are queued on one XCD (bins dealt largest first to the
// least loaded XCD); a workgroup reads the hardware XCC_ID and drains its
// own XCD's queue first, then the others', so one XCD's L2 holds the lines
// a bin's concurrent chunks fill.  The register only steers locality: any
// workgroup may take any chunk.

template <typename OffT, bool kTile>
struct PlaceStage {
    OffT slot;
    typename WiFmt<OffT, kTile>::Out val;
};
template <typename OffT, bool kTile>
constexpr int place_chunk()
{
    // LDS: 8k counters (32 KB) + one chunk of staged records (<= 64 KB)
    return sizeof(PlaceStage<OffT, kTile>) <= 8 ? 8192
           : sizeof(PlaceStage<OffT, kTile>) <= 16 ? 4096 : 2048;
}

template <typename OffT, bool kTile>
__global__ void __launch_bounds__(kPlaceThreads)
wi_place_kernel(const WiSlice *__restrict__ sl, const WiChunk *__restrict__ chunks,
                WiQueue *__restrict__ q, const typename WiFmt<OffT, kTile>::Rec *__restrict__ recs,
                const int *__restrict__ rel, const int *__restrict__ sbase,
                const OffT *__restrict__ begX, const int *__restrict__ kept,
                typename WiFmt<OffT, kTile>::Out *__restrict__ entX)
{
    using Out = typename WiFmt<OffT, kTile>::Out;
    constexpr int kC = place_chunk<OffT, kTile>();
    constexpr int kPer = kC / kPlaceThreads;  // records per thread
    constexpr int kCntPer = kBinSize / kPlaceThreads;
    __shared__ int pos[kBinSize];
    __shared__ PlaceStage<OffT, kTile> stage[kC];
    __shared__ int take;
    __shared__ int part[kPlaceThreads / kWave];
    const int home = xcc_id();
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    for (int x = 0; x < kXcds;) {
        const int qx = (home + x) & (kXcds - 1);
        __syncthreads();
        if (threadIdx.x == 0) {
            const int n = q->begin[qx + 1] - q->begin[qx];
            const int c = (int)atomicAdd(&q->next[qx * 32], 1u);
            take = c < n ? q->begin[qx] + c : -1;
        }
        for (int k = 0; k < kCntPer; k++)
            pos[threadIdx.x * kCntPer + k] = 0;
        __syncthreads();
        const int c = take;
        if (c < 0) {
            x++;  // this queue is drained: help the next XCD's
            continue;
        }
        const WiChunk ch = chunks[c];
        const WiSlice s = sl[ch.slice];
        const int *sb = s.multi >= 0 ? sbase + (long long)s.multi * kBinSize : nullptr;
        const int n = (int)(ch.hi - ch.lo);
        OffT slot[kPer];
        Out val[kPer];
        int tgt[kPer];
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int j = k * kPlaceThreads + threadIdx.x;
            tgt[k] = -1;
            if (j < n) {
                const long long i = ch.lo + j;
                const auto r = recs[i];
                OffT at = begX[r.v] + (OffT)rel[i];
                if constexpr (!kTile)
                    at += (OffT)kept[r.v];
                tgt[k] = r.v & (kBinSize - 1);
                if (sb)
                    at += (OffT)sb[tgt[k]];
                slot[k] = at;
                if constexpr (kTile)
                    val[k] = r.id;
                else
                    val[k] = WalkEntry<OffT>{r.start, r.deg};
                atomicAdd(&pos[tgt[k]], 1);
            }
        }
        __syncthreads();
        // exclusive scan of the 8k counts: each thread its kCntPer, then the
        // per-thread totals across the workgroup
        int run = 0, mine[kCntPer];
#pragma unroll
        for (int k = 0; k < kCntPer; k++) {
            mine[k] = pos[threadIdx.x * kCntPer + k];
            run += mine[k];
        }
        const int incl = wave_scan_add(run);
        if (lane == kWave - 1)
            part[wave] = incl;
        __syncthreads();
        int before = incl - run;
        for (int w = 0; w < wave; w++)
            before += part[w];
#pragma unroll
        for (int k = 0; k < kCntPer; k++) {
            pos[threadIdx.x * kCntPer + k] = before;
            before += mine[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kPer; k++)
            if (tgt[k] >= 0)
                stage[atomicAdd(&pos[tgt[k]], 1)] = PlaceStage<OffT, kTile>{slot[k], val[k]};
        __syncthreads();
        for (int j = threadIdx.x; j < n; j += kPlaceThreads)
            entX[stage[j].slot] = stage[j].val;
    }
}

// D: out-part of row u = tmpOut[outBeg[u], outBeg[u+1]) -> entX[begX[u], ...).
// A wave takes 64 consecutive rows, whose out-parts are contiguous, and
// copies them in 64-wide windows; a mark array and a max-scan give each
// position its row.
template <typename OffT>
__global__ void __launch_bounds__(256)
wi_outcopy_kernel(const OffT *__restrict__ outBeg, const OffT *__restrict__ begX, long long V,
                  const WalkEntry<OffT> *__restrict__ tmpOut, WalkEntry<OffT> *__restrict__ entX)
{
    __shared__ int mark_s[4][kWave];
    const int wave = threadIdx.x / kWave, lane = lane_id();
    int *mark = mark_s[wave];
    mark[lane] = 0;
    const long long step = (long long)gridDim.x * (blockDim.x / kWave) * kWave;
    for (long long u0 = ((long long)blockIdx.x * (blockDim.x / kWave) + wave) * kWave; u0 < V;
         u0 += step) {
        const int n = (int)min((long long)kWave, V - u0);
        const OffT s = outBeg[u0 + min(lane, n)];
        const OffT e = lane < n ? outBeg[u0 + lane + 1] : s;
        const OffT bx = lane < n ? begX[u0 + lane] : (OffT)0;
        const long long p0 = (long long)readlane(s, 0);
        const long long p1 = n == kWave ? (long long)readlane(e, 63) : (long long)readlane(s, n);
        wave_sync();
        int carry = 0;
        for (long long w = p0; w < p1; w += kWave) {
            if (lane < n && e > s && (long long)s >= w && (long long)s < w + kWave)
                mark[(int)((long long)s - w)] = lane;
            wave_sync();
            const int own = max(wave_scan_max(mark[lane]), carry);
            carry = readlane(own, 63);
            wave_sync();
            mark[lane] = 0;
            const OffT so = shfl_idx(s, own), bo = shfl_idx(bx, own);
            if (w + lane < p1)
                entX[bo + (OffT)(w + lane - (long long)so)] = tmpOut[w + lane];
            wave_sync();
        }
    }
}

// ---- exclusive scan of a[i] (+ b[i] when b is given), i < n, into
// out[0, n] (out[n] = the total): per-block sums, one workgroup scanning
// those, per-block scans.
__global__ void __launch_bounds__(1024) scan_reduce_kernel(const int *__restrict__ a,
                                                           const int *__restrict__ b, long long n,
                                                           long long *__restrict__ bsum)
{
    __shared__ long long part[1024 / kWave];
    const long long lo = (long long)blockIdx.x * kScanPer;
    const long long hi = min(lo + kScanPer, n);
    long long x = 0;
    for (long long i = lo + threadIdx.x; i < hi; i += blockDim.x)
        x += (long long)a[i] + (b ? b[i] : 0);
    x = (long long)wave_sum((unsigned long long)x);
    if (lane_id() == 0)
        part[threadIdx.x / kWave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < (int)(blockDim.x / kWave); w++)
            t += part[w];
        bsum[blockIdx.x] = t;
    }
}

template <typename OutT>
__global__ void __launch_bounds__(1024) scan_top_kernel(long long *__restrict__ bsum, long long nb,
                                                        long long n, OutT *__restrict__ out)
{
    __shared__ long long part[1024];
    __shared__ long long base;
    if (threadIdx.x == 0)
        base = 0;
    __syncthreads();
    for (long long c0 = 0; c0 < nb; c0 += blockDim.x) {
        const long long i = c0 + threadIdx.x;
        const long long x = i < nb ? bsum[i] : 0;
        part[threadIdx.x] = x;
        __syncthreads();
        for (int off = 1; off < (int)blockDim.x; off <<= 1) {
            const long long y = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += y;
            __syncthreads();
        }
        if (i < nb)
            bsum[i] = base + part[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1)
            base += part[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[n] = (OutT)base;
}

template <typename OutT>
__global__ void __launch_bounds__(1024) scan_out_kernel(const int *__restrict__ a,
                                                        const int *__restrict__ b, long long n,
                                                        const long long *__restrict__ bsum,
                                                        OutT *__restrict__ out)
{
    constexpr int kPer = kScanPer / 1024;  // thread t scans [lo + kPer t, lo + kPer (t + 1))
    __shared__ long long part[1024];
    const long long lo = (long long)blockIdx.x * kScanPer + (long long)threadIdx.x * kPer;
    int x[kPer];
    long long t = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        x[k] = lo + k < n ? a[lo + k] + (b ? b[lo + k] : 0) : 0;
        t += x[k];
    }
    part[threadIdx.x] = t;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        const long long y = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += y;
        __syncthreads();
    }
    long long run = bsum[blockIdx.x] + part[threadIdx.x] - t;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        if (lo + k < n)
            out[lo + k] = (OutT)run;
        run += x[k];
    }
}

long long scan_blocks(long long n) { return std::max<long long>(1, (n + kScanPer - 1) / kScanPer); }

template <typename OutT>
hipError_t launch_scan(const int *a, const int *b, long long n, long long *bsum, OutT *out,
                       hipStream_t stream)
{
    const long long nb = scan_blocks(n);
    if (n > 0) {
        hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nb), dim3(1024), 0, stream, a, b, n,
                           bsum);
        hipLaunchKernelGGL(scan_top_kernel<OutT>, dim3(1), dim3(1024), 0, stream, bsum, nb, n, out);
        hipLaunchKernelGGL(scan_out_kernel<OutT>, dim3((unsigned)nb), dim3(1024), 0, stream, a, b,
                           n, (const long long *)bsum, out);
    } else {
        hipLaunchKernelGGL(scan_top_kernel<OutT>, dim3(1), dim3(1024), 0, stream, bsum, 0ll, 0ll,
                           out);
    }
    return hipGetLastError();
}
template hipError_t launch_scan<int32_t>(const int *, const int *, long long, long long *,
                                         int32_t *, hipStream_t);
template hipError_t launch_scan<int64_t>(const int *, const int *, long long, long long *,
                                         int64_t *, hipStream_t);

__global__ void wi_bin_starts_kernel(const long long *__restrict__ O, int G, int nbins,
                                     long long *__restrict__ out)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b <= nbins)
        out[b] = O[(long long)b * G];  // O[nbins * G] is the total
}

template <typename OffT>
__global__ void __launch_bounds__(256) wi_dtab_kernel(const OffT *__restrict__ beg, long long V,
                                                      const uint32_t *__restrict__ core_dst,
                                                      uint16_t *__restrict__ dtab)
{
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < V;
         x += (long long)gridDim.x * blockDim.x) {
        const unsigned d = (unsigned)(beg[x + 1] - beg[x]);
        const unsigned t = core_dst ? (core_dst[x >> 5] >> (x & 31)) & 1u : 0u;
        dtab[x] = (uint16_t)(d | (t << 15));
    }
}

// ------------------------------------------------------------------ host
int walk_bins(long long V) { return (int)((V + kBinSize - 1) / kBinSize); }

bool walk_index_binned(const Ctx *ctx) { return ctx->g.V > 0 && walk_bins(ctx->g.V) <= kMaxBins; }

// Pass A and the scans over a CSR of `rows` rows whose in-entries target
// [0, cols): kept and outBeg (rows + 1) on the device; the bins' first
// records, the out-entry total and (resident graph) both walks' probe
// totals, then the probes and edges the dense core took, on the host.
// Synchronises.
template <typename OffT, bool kTile>
static int wi_count(Ctx *ctx, const OffT *beg, const int32_t *adj, long long rows, long long cols,
                    const int32_t *degR, const int32_t *degC, int *kept, OffT *outBeg,
                    unsigned long long stats_out[4], int32_t *dvE)
{
    WalkIndexState &w = ctx->wi;
    const int nbins = std::max(1, walk_bins(cols));
    if (nbins > kMaxBins)
        return fail(TRUST_ERR_ARG, "walk index: %d target bins (at most %d)", nbins, kMaxBins);
    const size_t lds = pass_lds_bytes<OffT>(nbins);
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, wi_pass_kernel<OffT, 1, kTile>,
                                                       kPassThreads, lds));
    if (per_cu < 1)
        return fail(TRUST_ERR_ARG, "walk index: %d bins do not fit the LDS", nbins);
    const int G = std::min(per_cu, 2) * ctx->num_cus;
    const long long nM = (long long)nbins * G;
    int rc;
    if ((rc = w.M.ensure(sizeof(int) * (size_t)nM)) ||
        (rc = w.O.ensure(sizeof(long long) * (size_t)(nM + 1))) ||
        (rc = w.bin_start.ensure(sizeof(long long) * (size_t)(nbins + 1))) ||
        (rc = w.bsum.ensure(sizeof(long long) *
                            (size_t)(scan_blocks(std::max(nM, std::max(rows, cols))) + 1))) ||
        (rc = ctx->side_stats.ensure(4 * sizeof(unsigned long long))))
        return rc;
    unsigned long long *st = (unsigned long long *)ctx->side_stats.p;
    HIPCHK(launch_fill_u32((uint32_t *)st, 0u, 8, ctx->stream));
    WiPassArgs<OffT, kTile> pa;
    memset(&pa, 0, sizeof(pa));
    pa.beg = beg;
    pa.adj = adj;
    pa.V = rows;
    pa.nbins = nbins;
    pa.degR = degR;
    pa.degC = degC;
    pa.kept = kept;
    pa.M = (int *)w.M.p;
    pa.stats = st;
    pa.dvE = dvE;
    if (!kTile && ctx->core.active) {
        pa.core_src = ctx->core.src_bits();
        pa.core_dst = ctx->core.dst_bits();
    }
    if (!kTile && rows > 0 && ctx->g.max_deg < (1 << 15)) {
        if ((rc = w.dtab.ensure(sizeof(uint16_t) * (size_t)rows)))
            return rc;
        const long long blocks = std::min<long long>((rows + 255) / 256, 32ll * ctx->num_cus);
        hipLaunchKernelGGL(wi_dtab_kernel<OffT>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                           beg, rows, pa.core_dst, (uint16_t *)w.dtab.p);
        HIPCHK(hipGetLastError());
        pa.dtab = (const uint16_t *)w.dtab.p;
    }
    if (rows > 0) {
        hipLaunchKernelGGL((wi_pass_kernel<OffT, 0, kTile>), dim3(G), dim3(kPassThreads), lds,
                           ctx->stream, pa);
        HIPCHK(hipGetLastError());
    } else {
        HIPCHK(launch_fill_u32((uint32_t *)w.M.p, 0u, nM, ctx->stream));
    }
    HIPCHK(launch_scan<int64_t>((const int *)w.M.p, nullptr, nM, (long long *)w.bsum.p,
                                (int64_t *)w.O.p, ctx->stream));
    hipLaunchKernelGGL(wi_bin_starts_kernel, dim3((unsigned)((nbins + 256) / 256)), dim3(256), 0,
                       ctx->stream, (const long long *)w.O.p, G, nbins, (long long *)w.bin_start.p);
    HIPCHK(hipGetLastError());
    HIPCHK(launch_scan<OffT>(kept, nullptr, rows, (long long *)w.bsum.p, outBeg, ctx->stream));
    w.bin_start_h.resize((size_t)nbins + 1);
    unsigned long long h[4];
    OffT out_total = 0;
    HIPCHK(hipMemcpyAsync(h, st, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(w.bin_start_h.data(), w.bin_start.p, sizeof(long long) * (nbins + 1),
                          hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(&out_total, outBeg + rows, sizeof(OffT), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (stats_out)
        for (int k = 0; k < 4; k++)
            stats_out[k] = h[k];
    w.grid = G;
    w.nbins = nbins;
    w.out_total = (int64_t)out_total;
    return TRUST_OK;
}

// Passes B, C1, C2 (and D for the resident graph) after wi_count on the same
// CSR: tmpOut receives the out-entries in row order (at outBeg), entX the
// in-entries at begX (resident graph: after each row's out-part, copied in
// by D; tile walk: a CSR of the in-rows, begX = its cols + 1 offsets).
// Launches only (one host-to-device copy of the slice table).
template <typename OffT, bool kTile>
static int wi_build(Ctx *ctx, const OffT *beg, const int32_t *adj, long long rows, long long cols,
                    const int32_t *degR, const int32_t *degC, const int *kept, const OffT *outBeg,
                    int *incnt, typename WiFmt<OffT, kTile>::Out *tmpOut, OffT *begX,
                    typename WiFmt<OffT, kTile>::Out *entX, int32_t *dvE)
{
    using Rec = typename WiFmt<OffT, kTile>::Rec;
    WalkIndexState &w = ctx->wi;
    const int nbins = w.nbins, G = w.grid;
    const long long x_in = w.bin_start_h[(size_t)nbins], x_out = w.out_total;
    // slices of about 1/4 of a fair share of the in-records per CU
    const long long S = std::max<long long>(
        1 << 16, std::min<long long>(1 << 22, x_in / (4ll * ctx->num_cus) + 1));
    std::vector<WiSlice> &sl = w.slices_h;
    sl.clear();
    int multi = 0;
    for (int b = 0; b < nbins; b++) {
        const long long lo = w.bin_start_h[(size_t)b], hi = w.bin_start_h[(size_t)b + 1];
        if (hi <= lo)
            continue;
        const long long ns = (hi - lo + S - 1) / S;
        for (long long q = 0; q < ns; q++) {
            WiSlice sc;
            sc.bin = b;
            sc.multi = ns > 1 ? multi++ : -1;
            sc.lo = lo + q * S;
            sc.hi = std::min(hi, sc.lo + S);
            sl.push_back(sc);
        }
    }
    // C2's chunks: bins dealt to the XCDs largest first, each to the least
    // loaded; an XCD's chunks in bin order
    std::vector<int> bins_by_size;
    for (int b = 0; b < nbins; b++)
        if (w.bin_start_h[(size_t)b + 1] > w.bin_start_h[(size_t)b])
            bins_by_size.push_back(b);
    std::sort(bins_by_size.begin(), bins_by_size.end(), [&](int x, int y) {
        const long long nx = w.bin_start_h[(size_t)x + 1] - w.bin_start_h[(size_t)x];
        const long long ny = w.bin_start_h[(size_t)y + 1] - w.bin_start_h[(size_t)y];
        return nx != ny ? nx > ny : x < y;
    });
    std::vector<int> bin_xcd((size_t)nbins, 0);
    long long load[kXcds] = {0};
    for (int b : bins_by_size) {
        int best = 0;
        for (int x = 1; x < kXcds; x++)
            if (load[x] < load[best])
                best = x;
        bin_xcd[(size_t)b] = best;
        load[best] += w.bin_start_h[(size_t)b + 1] - w.bin_start_h[(size_t)b];
    }
    // C2 chunks: whole C1 rounds, so that their ranks of a target are contiguous
    constexpr long long kC = place_chunk<OffT, kTile>();
    static_assert(kC % kRankThreads == 0, "a C2 chunk must hold whole C1 rounds");
    std::vector<WiChunk> &ch = w.chunks_h;
    ch.clear();
    WiQueue qh;
    memset(&qh, 0, sizeof(qh));
    for (int x = 0; x < kXcds; x++) {
        qh.begin[x] = (int)ch.size();
        for (size_t k = 0; k < sl.size(); k++) {
            if (bin_xcd[(size_t)sl[k].bin] != x)
                continue;
            for (long long lo = sl[k].lo; lo < sl[k].hi; lo += kC)
                ch.push_back(WiChunk{(int)k, 0, lo, std::min(sl[k].hi, lo + kC)});
        }
    }
    qh.begin[kXcds] = (int)ch.size();
    int rc;
    if ((rc = w.slices.ensure(sizeof(WiSlice) * std::max<size_t>(sl.size(), 1))) ||
        (rc = w.chunks.ensure(sizeof(WiChunk) * std::max<size_t>(ch.size(), 1) + sizeof(WiQueue))) ||
        (rc = w.recs.ensure(sizeof(Rec) * (size_t)std::max<long long>(x_in, 1))) ||
        (rc = w.rel.ensure(sizeof(int) * (size_t)std::max<long long>(x_in, 1))) ||
        (rc = w.sbase.ensure(sizeof(int) * (size_t)kBinSize * std::max(multi, 1))))
        return rc;
    const size_t lds = pass_lds_bytes<OffT>(nbins);
    WiQueue *qd = (WiQueue *)w.chunks.p;
    WiChunk *chd = (WiChunk *)((char *)w.chunks.p + sizeof(WiQueue));
    if (!sl.empty()) {
        HIPCHK(hipMemcpyAsync(w.slices.p, sl.data(), sizeof(WiSlice) * sl.size(),
                              hipMemcpyHostToDevice, ctx->stream));
        w.queue_h = qh;
        HIPCHK(hipMemcpyAsync(qd, &w.queue_h, sizeof(WiQueue), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(chd, ch.data(), sizeof(WiChunk) * ch.size(), hipMemcpyHostToDevice,
                              ctx->stream));
    }
    WiPassArgs<OffT, kTile> pa;
    memset(&pa, 0, sizeof(pa));
    pa.beg = beg;
    pa.adj = adj;
    pa.V = rows;
    pa.nbins = nbins;
    pa.degR = degR;
    pa.degC = degC;
    pa.O = (const long long *)w.O.p;
    pa.outBeg = outBeg;
    pa.dvE = dvE;
    pa.tmpOut = tmpOut;
    pa.recs = (Rec *)w.recs.p;
    if (rows > 0) {
        hipLaunchKernelGGL((wi_pass_kernel<OffT, 1, kTile>), dim3(G), dim3(kPassThreads), lds,
                           ctx->stream, pa);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(launch_fill_u32((uint32_t *)incnt, 0u, cols, ctx->stream));
    if (!sl.empty()) {
        hipLaunchKernelGGL(wi_rank_kernel<Rec>, dim3((unsigned)sl.size()), dim3(kRankThreads), 0,
                           ctx->stream, (const WiSlice *)w.slices.p, (const Rec *)w.recs.p,
                           (int *)w.rel.p, incnt, (int *)w.sbase.p);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(launch_scan<OffT>(kTile ? incnt : kept, kTile ? nullptr : incnt, cols,
                             (long long *)w.bsum.p, begX, ctx->stream));
    if (!sl.empty()) {
        // persistent: one workgroup per CU (the LDS of a chunk; none waits on another)
        const int blocks = (int)std::min<long long>((long long)ch.size(), (long long)ctx->num_cus);
        hipLaunchKernelGGL((wi_place_kernel<OffT, kTile>), dim3(blocks), dim3(kPlaceThreads), 0,
                           ctx->stream, (const WiSlice *)w.slices.p, (const WiChunk *)chd, qd,
                           (const Rec *)w.recs.p, (const int *)w.rel.p, (const int *)w.sbase.p,
                           (const OffT *)begX, kept, entX);
        HIPCHK(hipGetLastError());
    }
    if constexpr (!kTile) {
        if (x_out > 0) {
            const int blocks = (int)std::min<long long>((rows + 4 * kWave - 1) / (4 * kWave), 16384ll);
            hipLaunchKernelGGL(wi_outcopy_kernel<OffT>, dim3(blocks), dim3(256), 0, ctx->stream,
                               outBeg, (const OffT *)begX, rows, (const WalkEntry<OffT> *)tmpOut,
                               entX);
            HIPCHK(hipGetLastError());
        }
    }
    return TRUST_OK;
}

template <typename OffT>
static int count_pass_t(Ctx *ctx)
{
    DevCsr &g = ctx->g;
    WalkIndexState &w = ctx->wi;
    const long long V = g.V;
    int rc;
    if ((rc = w.out_beg.ensure(sizeof(OffT) * (size_t)(V + 1))) ||
        (rc = w.dv.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(g.E, 1))) ||
        (rc = ctx->side_cnt.ensure(2 * sizeof(int) * (size_t)V)))
        return rc;
    unsigned long long h[4];
    rc = wi_count<OffT, false>(ctx, (const OffT *)g.beg.p, (const int32_t *)g.adj.p, V, V, nullptr,
                               nullptr, (int *)ctx->side_cnt.p + V, (OffT *)w.out_beg.p, h,
                               (int32_t *)w.dv.p);
    if (rc)
        return rc;
    ctx->side_w_ref = h[0];
    ctx->side_w_short = h[1];
    ctx->core.dense_probes = h[2];
    ctx->core.dense_edges = h[3];
    ctx->side_stats_ready = true;
    w.counted = true;
    return TRUST_OK;
}

int walk_index_count(Ctx *ctx)
{
    return ctx->g.off64 ? count_pass_t<int64_t>(ctx) : count_pass_t<int32_t>(ctx);
}

template <typename OffT>
static int build_t(Ctx *ctx)
{
    DevCsr &g = ctx->g;
    WalkIndexState &w = ctx->wi;
    const long long V = g.V;
    const long long x_in = w.bin_start_h[(size_t)w.nbins], x_out = w.out_total;
    const size_t ob = sizeof(OffT);
    int rc;
    if ((rc = w.tmp_out.ensure(2 * ob * (size_t)std::max<long long>(x_out, 1))) ||
        (rc = ctx->side_beg.ensure(ob * (size_t)(V + 1))) ||
        (rc = ctx->side_ent.ensure(2 * ob * (size_t)std::max<long long>(x_in + x_out, 1))))
        return rc;
    int *incnt = (int *)ctx->side_cnt.p, *kept = incnt + V;
    rc = wi_build<OffT, false>(ctx, (const OffT *)g.beg.p, (const int32_t *)g.adj.p, V, V, nullptr,
                               nullptr, kept, (const OffT *)w.out_beg.p, incnt,
                               (WalkEntry<OffT> *)w.tmp_out.p, (OffT *)ctx->side_beg.p,
                               (WalkEntry<OffT> *)ctx->side_ent.p, (int32_t *)w.dv.p);
    if (rc)
        return rc;
    ctx->side_entries = x_in + x_out;
    return TRUST_OK;
}

int walk_index_build(Ctx *ctx)
{
    if (!ctx->wi.counted)
        return fail(TRUST_ERR_ARG, "walk index: count pass missing");
    return ctx->g.off64 ? build_t<int64_t>(ctx) : build_t<int32_t>(ctx);
}

// The walk of tile (i,k) for the short side by global out-degree: rows of
// class i (the tile's rows) keep the v with d+(v) <= d+(u) (out-walk, ids
// of class-k rows), and each class-k row v gets the u -> v with
// 2 <= d+(u) < d+(v) (in-walk, ids of class-i rows).
template <typename OffT>
int tile_walk_build(Ctx *ctx, const OffT *beg, const int32_t *adj, long long rows, long long cols,
                    const int32_t *degR, const int32_t *degC, TileWalk &tw)
{
    int rc;
    const size_t ob = sizeof(OffT);
    tw.rows = rows;
    tw.cols = cols;
    if ((rc = tw.out_beg.ensure(ob * (size_t)(rows + 1))) ||
        (rc = tw.in_beg.ensure(ob * (size_t)(cols + 1))) ||
        (rc = ctx->side_cnt.ensure(sizeof(int) * (size_t)(rows + cols))))
        return rc;
    int *kept = (int *)ctx->side_cnt.p, *incnt = kept + rows;
    rc = wi_count<OffT, true>(ctx, beg, adj, rows, cols, degR, degC, kept, (OffT *)tw.out_beg.p,
                              nullptr, nullptr);
    if (rc)
        return rc;
    WalkIndexState &w = ctx->wi;
    tw.x_out = w.out_total;
    tw.x_in = w.bin_start_h[(size_t)w.nbins];
    if ((rc = tw.out_ids.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(tw.x_out, 1))) ||
        (rc = tw.in_ids.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(tw.x_in, 1))))
        return rc;
    return wi_build<OffT, true>(ctx, beg, adj, rows, cols, degR, degC, kept,
                                (const OffT *)tw.out_beg.p, incnt, (int32_t *)tw.out_ids.p,
                                (OffT *)tw.in_beg.p, (int32_t *)tw.in_ids.p, nullptr);
}
template int tile_walk_build<int32_t>(Ctx *, const int32_t *, const int32_t *, long long,
                                      long long, const int32_t *, const int32_t *, TileWalk &);
template int tile_walk_build<int64_t>(Ctx *, const int64_t *, const int32_t *, long long,
                                      long long, const int32_t *, const int32_t *, TileWalk &);

int tile_walk_bins_ok(long long cols) { return walk_bins(cols) <= kMaxBins; }

}  // namespace trust
