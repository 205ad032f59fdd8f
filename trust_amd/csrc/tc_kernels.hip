// gfx950 kernels for TRUST's vertex-centric hashed triangle count.
//
// Count = sum_u sum_{v in N+(u)} |N+(u) & N+(v)| over a degree-oriented CSR
// (reference semantics: Without-graph-partition/TC.cu:230-600; partitioned
// three-tile form: With-graph-partition/TC.cu:630-973).
//
// Design (MI355X-first, not a translation of the CUDA kernel):
//  * wave64 throughout.  A vertex is counted by one wavefront (wave tier),
//    by a whole workgroup with an LDS table (hub tier in the main launch,
//    big tier in a large-LDS launch) or by a whole workgroup with an HBM
//    table (spill tier).  Tables hold >= 2 slots per key, so they never
//    overflow (the reference drops entries past 100 per bucket,
//    TC.cu:293-301).  They are bucketed two-choice cuckoo tables: a lookup
//    is two 8-byte reads and branch-free compares, so no lane's probe chain
//    stalls its wave (the reference's linear_search scans a bucket,
//    TC.cu:70-110).  A vertex whose cuckoo insertion fails falls back to
//    linear probing, whose loops end at an empty slot.
//  * The probe walk takes 64 neighbours v of u at a time: lanes load
//    (beg[v], deg v), an inclusive wave scan (DPP row_shr / row_bcast) lays
//    the 64 lists end to end, and each 64-wide window of that concatenation
//    is one coalesced load per lane.  A batch of kU windows (four, or six
//    for the short-side walk's long lists) is in flight per wave before
//    any hash lookup, so each lane has kU independent loads outstanding.  The segment owning each lane is found with a
//    64-entry LDS mark array and a wave max-scan; a window inside a single
//    list skips that and reads its base with one readlane.  This replaces
//    the reference's per-lane skip loop with its 32-lane __shfl_sync
//    resynchronisation (TC.cu:513-541) and its 16x64 CTA walk that is only
//    correct at 1024 threads (TC.cu:330-359).
//  * A workgroup vertex splits its probe windows evenly over its waves
//    (per-chunk window counts, a prefix, then units of windows pulled from
//    an LDS counter), so no wave idles while another walks a long list.
//  * Persistent workgroups pull chunks of vertices from per-tier atomic
//    queues (the reference's G_INDEX counters, TC.cu:405-422,560-577);
//    logical rank r of P takes chunks c with c % P == r (TC.cu:259,433).
//  * The count is reduced per wave in registers and added once per wave to
//    a uint64 in device memory (TC.cu:586-596).
#include <hip/hip_runtime.h>

#include "tc_kernels.h"
#include "wave_ops.h"

namespace trust {

constexpr int kEmpty = -1;
constexpr int kNoProbe = -2;  // a window position past the end: matches no slot

// Per-wave attribution of triangles to the graphs of a batch.  The state
// sits in LDS, in the kMarkStride - kWave words after the wave's mark array
// (gs[0], gs[1]) = the vertex range of the graph being counted, gs[2] = that
// graph (-1: none), not in scalar registers, where it would stay live across
// the whole probe walk and push other values out.  The wave's running count
// acc holds the current graph's triangles only: entering a vertex of another
// graph adds acc to the previous graph, with a completion time stamp, and
// clears it.  Called with the whole wave active, u wave-uniform, before the
// vertex's hits are added to acc.
__device__ __forceinline__ unsigned long long uniform(unsigned long long x)
{
    const int lo = __builtin_amdgcn_readfirstlane((int)(x & 0xffffffffu));
    const int hi = __builtin_amdgcn_readfirstlane((int)(x >> 32));
    return ((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo;
}

__device__ __forceinline__ void graph_init(int *gs)
{
    gs[0] = 0;
    gs[1] = 0;  // an empty range: the first vertex enters its graph
    gs[2] = -1;
}

template <typename Args>
__device__ __forceinline__ void graph_flush(const Args &a, int *gs, unsigned long long &acc)
{
    const unsigned long long total = uniform(wave_sum(acc));
    const int g = __builtin_amdgcn_readfirstlane(gs[2]);
    if (g >= 0 && __lane_id() == 0) {
        if (total)
            atomicAdd(a.graph_counts + g, total);
        atomicMax(a.graph_done + g, (unsigned long long)wall_clock64());
    }
    acc = 0;
}

template <typename Args>
__device__ __forceinline__ void graph_enter(const Args &a, int *gs, int u,
                                            unsigned long long &acc)
{
    const int2 range = *(const int2 *)gs;
    if (u >= __builtin_amdgcn_readfirstlane(range.x) && u < __builtin_amdgcn_readfirstlane(range.y))
        return;
    graph_flush(a, gs, acc);
    // the page of u bounds the search: vbase[lo] <= u < vbase[hi]
    const int pg = u >> a.graph_page_shift;
    int lo = a.graph_page[pg], hi = a.graph_page[pg + 1] + 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.graph_vbase[mid] <= u)
            lo = mid;
        else
            hi = mid;
    }
    const int g = __builtin_amdgcn_readfirstlane(lo);
    gs[0] = (int)a.graph_vbase[g];  // every lane stores the same values
    gs[1] = (int)a.graph_vbase[g + 1];
    gs[2] = g;
}

// ---------------------------------------------------------- hash tables
__device__ __forceinline__ int ceil_lg(int x)  // smallest lg with 2^lg >= x (x >= 1)
{
    return x > 1 ? 32 - __clz(x - 1) : 0;
}

// Table size for a hash source of degree d in a tier whose capacity is
// 2^cap_lg: >= 2d slots (the planner only sends vertices whose 2d fits).
__device__ __forceinline__ int table_lg(int d, int cap_lg)
{
    int lg = ceil_lg(2 * d);
    lg = lg < kMinTableLg ? kMinTableLg : lg;
    return lg > cap_lg ? cap_lg : lg;
}

// ---- bucketed two-choice cuckoo table (the lookup path) ----
// 2^lg int slots = 2^(lg-1) buckets of two slots.  A key lives in bucket
// b1 = key & mask or b2 = a multiplicative hash of key, so a lookup is
// exactly two 8-byte reads and branch-free compares: no probe chain whose
// longest lane would stall the whole wave.  Load factor <= 1/2, far below
// the ~0.89 limit of 2-way, 2-slot cuckoo hashing.  b2 multiplies the low
// 24 bits (a full-rate v_mul_u32_u24; the 32-bit multiply is quarter rate
// and the walk is instruction-bound) and takes the top bits of the low word;
// keys equal in their low 24 bits share both buckets, which a bucket pair's
// four slots absorb, and a failed insertion falls back to linear probing.
__device__ __forceinline__ unsigned int bucket1(int key, int lgb)
{
    return (unsigned int)key & ((1u << lgb) - 1u);
}
__device__ __forceinline__ unsigned int bucket2(int key, int lgb)
{
    // __umul24 returns int: shift the unsigned value (a logical shift)
    return (unsigned int)__umul24((unsigned int)key, 0x9E3779u) >> (32 - lgb);
}

template <bool kGlobal>
__device__ __forceinline__ int2 bucket_load(const int *tab, unsigned int b)
{
    if constexpr (kGlobal) {
        unsigned long long v = __hip_atomic_load((const unsigned long long *)tab + b,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return make_int2((int)(v & 0xffffffffu), (int)(v >> 32));
    } else {
        return *((const int2 *)tab + b);
    }
}

// Concurrent insertion: claim an empty slot in either bucket, else evict
// an occupant with an atomic exchange and carry it to its other bucket.
// Returns true when `kMaxKicks` evictions did not place every key (the
// caller then rebuilds the table with linear probing).
constexpr int kMaxKicks = 48;
__device__ __forceinline__ bool cuckoo_insert(int *tab, int lgb, int key)
{
    for (int it = 0; it < kMaxKicks; it++) {
        const unsigned int c1 = bucket1(key, lgb), c2 = bucket2(key, lgb);
        int prev = atomicCAS(tab + 2 * c1, kEmpty, key);
        if (prev == kEmpty || prev == key)
            return false;
        prev = atomicCAS(tab + 2 * c1 + 1, kEmpty, key);
        if (prev == kEmpty || prev == key)
            return false;
        prev = atomicCAS(tab + 2 * c2, kEmpty, key);
        if (prev == kEmpty || prev == key)
            return false;
        prev = atomicCAS(tab + 2 * c2 + 1, kEmpty, key);
        if (prev == kEmpty || prev == key)
            return false;
        const unsigned int slot = 2 * ((it & 1) ? c2 : c1) + (((unsigned int)it >> 1) & 1u);
        key = atomicExch(tab + slot, key);
        if (key == kEmpty)
            return false;
    }
    return true;
}

// Membership count of N keys (kNoProbe matches no slot), summed over
// the wave: the per-key hit masks are counted with scalar popcounts, which
// keeps vector instructions (the walk's bound) for the lookup itself.  All
// 64 lanes must call it.  Both buckets of every key are read at once:
// reading the second bucket only when the first is full (sound, as slots
// are never vacated) saves LDS reads but serialises two LDS round trips, and
// measured 12% slower at R-MAT 24.
template <bool kGlobal, int N>
__device__ __forceinline__ unsigned int cuckoo_count(const int *tab, int lg, const int (&w)[N])
{
    const int lgb = lg - 1;
    int2 x1[N], x2[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
        x1[k] = bucket_load<kGlobal>(tab, bucket1(w[k], lgb));
        x2[k] = bucket_load<kGlobal>(tab, bucket2(w[k], lgb));
    }
    unsigned int hits = 0;
#pragma unroll
    for (int k = 0; k < N; k++) {
        // each compare's lane mask feeds a scalar OR and popcount directly
        const unsigned long long m = __builtin_amdgcn_ballot_w64(x1[k].x == w[k]) |
                                     __builtin_amdgcn_ballot_w64(x1[k].y == w[k]) |
                                     __builtin_amdgcn_ballot_w64(x2[k].x == w[k]) |
                                     __builtin_amdgcn_ballot_w64(x2[k].y == w[k]);
        hits += (unsigned int)__popcll(m);
    }
    return hits;
}

// ---- linear probing (fallback when a cuckoo insertion fails) ----

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

// Membership of up to N keys (key < 0: skip), summed over the wave.
// First slots of all keys are read together, then each key finishes its
// probe sequence.  All 64 lanes must call it.
template <bool kGlobal, int N>
__device__ __forceinline__ unsigned int linear_count(const int *tab, int lg, const int (&w)[N])
{
    const unsigned int mask = (1u << lg) - 1u;
    unsigned int h[N];
    int s[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
        h[k] = hash_slot(w[k], lg);
        s[k] = w[k] >= 0 ? slot_load<kGlobal>(tab + h[k]) : kEmpty;
    }
    unsigned int hits = 0;
#pragma unroll
    for (int k = 0; k < N; k++) {
        if (w[k] < 0)
            continue;
        for (;;) {
            if (s[k] == w[k]) {
                hits++;
                break;
            }
            if (s[k] == kEmpty)
                break;
            h[k] = (h[k] + 1u) & mask;
            s[k] = slot_load<kGlobal>(tab + h[k]);
        }
    }
    return (unsigned int)wave_sum((unsigned long long)hits);
}

// ---------------------------------------------------------- probe walk
// Lane state for a group of up to 64 neighbours v (one per lane): this
// lane's list is positions [excl, incl) of the concatenation, whose element
// at position t is adjP[base + t].
template <typename OffT>
struct Group {
    int incl, excl, deg;
    OffT base;
    int total;  // uniform: sum of the 64 degrees
};

template <typename OffT>
__device__ __forceinline__ Group<OffT> make_group(OffT s, int deg)
{
    Group<OffT> g;
    g.deg = deg;
    g.incl = wave_scan_add(deg);
    g.excl = g.incl - deg;
    g.total = readlane(g.incl, 63);
    g.base = s - (OffT)g.excl;
    return g;
}

template <typename OffT>
__device__ __forceinline__ Group<OffT> load_group(int v, bool valid, const OffT *__restrict__ begP)
{
    OffT s = 0;
    int deg = 0;
    if (valid) {
        s = begP[v];
        deg = (int)(begP[v + 1] - s);
    }
    return make_group(s, deg);
}

// Group of 64 walk entries starting at entry `at` of the walk view.  A walk
// index that stores each entry's list start and length (entW) is read in
// place, coalesced, one load per entry; otherwise the entry is a vertex id whose offsets
// are looked up.
template <typename OffT>
__device__ __forceinline__ Group<OffT> walk_group(const TcArgs<OffT> &a, OffT at, bool valid)
{
    if (a.entW) {
        OffT s = 0;
        int deg = 0;
        if (valid) {
            const WalkEntry<OffT> x = a.entW[at];
            s = x.start;
            deg = x.deg;
        }
        return make_group(s, deg);
    }
    return load_group(valid ? a.adjW[at] : 0, valid, a.begP);
}

// Probes windows [wbeg, wend) (64 positions each) of a group's
// concatenation against the table and returns the hits of the whole wave
// (wave-uniform; gather mode: the wave's checksum).  All 64 lanes must call
// it.  `mark` is this wave's 64-int LDS scratch: all-zero on entry and exit.
template <typename OffT, bool kGlobal, int kMode, int kU>
__device__ __forceinline__ unsigned long long probe_windows(const Group<OffT> &g, int wbeg, int wend,
                                                      const int32_t *__restrict__ adjP,
                                                      const int *tab, int lg, bool linear,
                                                      int *mark)
{
    const int lane = lane_id();
    const int bend = min(wend * kWave, g.total);
    const int b = wbeg * kWave;
    if (b >= bend)
        return 0;
    // segment holding position b = number of lists ending at or before b
    int carry = (int)__popcll(__ballot(g.incl <= b));
    // end and base of list `carry`, kept in scalar registers until it changes
    int carry_incl = readlane(g.incl, carry);
    OffT carry_base = readlane(g.base, carry);
    // Base of the window starting at position bk (< bend): adjP[base + t] is
    // position t.  A window inside one list takes the cached base; otherwise
    // each position's list comes from the mark array and a max-scan.
    auto window_base = [&](int bk) -> OffT {
        if (carry_incl >= bk + kWave)
            return carry_base;
        const unsigned int off = (unsigned int)(g.excl - bk);
        if (g.deg > 0 && off < (unsigned int)kWave)
            mark[off] = lane;
        wave_sync();
        const int m = mark[lane];
        mark[lane] = 0;
        wave_sync();
        const int j = max(wave_scan_max(m), carry);
        carry = readlane(j, 63);
        carry_incl = readlane(g.incl, carry);
        carry_base = readlane(g.base, carry);
        return shfl_idx(g.base, j);
    };
    unsigned long long total = 0;  // wave-uniform hits
    unsigned int checksum = 0;     // gather mode, per lane
    auto lookup = [&](const auto &w) {
        constexpr int N = (int)(sizeof(w) / sizeof(int));
        if constexpr (kMode == kModeGather) {
#pragma unroll
            for (int k = 0; k < N; k++)
                checksum += w[k] >= 0 ? (unsigned int)w[k] : 0u;
        } else if constexpr (N > 4 && N % 2 == 0) {
            // two halves, so that the table reads of one half and their
            // compares need only half the registers
            int lo[N / 2], hi[N / 2];
#pragma unroll
            for (int k = 0; k < N / 2; k++) {
                lo[k] = w[k];
                hi[k] = w[N / 2 + k];
            }
            total += linear ? linear_count<kGlobal, N / 2>(tab, lg, lo)
                            : cuckoo_count<kGlobal, N / 2>(tab, lg, lo);
            total += linear ? linear_count<kGlobal, N / 2>(tab, lg, hi)
                            : cuckoo_count<kGlobal, N / 2>(tab, lg, hi);
        } else {
            total += linear ? linear_count<kGlobal, N>(tab, lg, w)
                            : cuckoo_count<kGlobal, N>(tab, lg, w);
        }
    };
    if (bend - b <= kWave) {
        // a single window (most small vertices): one load, one lookup
        const OffT bj = window_base(b);
        const int t = b + lane;
        const int w1[1] = {t < bend ? adjP[bj + (OffT)t] : kNoProbe};
        lookup(w1);
    } else {
        // Element idx of adjP.  With int32 offsets adj holds fewer than 2^30
        // ids (the engine's rule), so the byte offset fits 32 bits and the
        // load takes a 32-bit offset from the scalar base: no sign extension
        // and no 64-bit address arithmetic per lane.
        auto load = [&](OffT idx) -> int {
            if constexpr (sizeof(OffT) == 4)
                return *(const int *)((const char *)adjP + ((uint32_t)idx << 2));
            else
                return adjP[idx];
        };
        auto bases = [&](int at, OffT (&bj)[kU]) {
#pragma unroll
            for (int k = 0; k < kU; k++)
                bj[k] = at + k * kWave < bend ? window_base(at + k * kWave) : (OffT)0;
        };
        // The loads of a batch lying entirely before bend.  Every lane loads
        // and nothing reads the results here: with no branch around the
        // loads the compiler knows exactly kU loads are younger than the
        // previous batch and waits for that batch alone.
        auto fetch_full = [&](int at, int (&w)[kU]) {
            OffT bj[kU];
            bases(at, bj);
#pragma unroll
            for (int k = 0; k < kU; k++)
                w[k] = load(bj[k] + (OffT)(at + k * kWave + lane));
        };
        // The last batch: positions past bend read adjP[0] (which exists,
        // since bend > 0) and are masked before the lookup.
        auto fetch_tail = [&](int at, int (&w)[kU]) {
            OffT bj[kU];
            bases(at, bj);
#pragma unroll
            for (int k = 0; k < kU; k++) {
                const int t = at + k * kWave + lane;
                w[k] = load(t < bend ? bj[k] + (OffT)t : (OffT)0);
            }
        };
        // Software pipeline over batches of kU windows: the next
        // batch's loads are in flight while the current one is looked up.
        constexpr int kBatch = kWave * kU;
        const int last = b + (bend - b - 1) / kBatch * kBatch;  // start of the last batch
        int wcur[kU];
        if (last == b) {
            fetch_tail(b, wcur);
        } else {
            fetch_full(b, wcur);
            for (int bn = b + kBatch; bn < last; bn += kBatch) {
                int wnext[kU];
                fetch_full(bn, wnext);
                lookup(wcur);
#pragma unroll
                for (int k = 0; k < kU; k++)
                    wcur[k] = wnext[k];
            }
            int wtail[kU];
            fetch_tail(last, wtail);
            lookup(wcur);
#pragma unroll
            for (int k = 0; k < kU; k++)
                wcur[k] = wtail[k];
        }
#pragma unroll
        for (int k = 0; k < kU; k++)
            wcur[k] = last + k * kWave + lane < bend ? wcur[k] : kNoProbe;
        lookup(wcur);
    }
    if constexpr (kMode == kModeGather)
        total += wave_sum((unsigned long long)checksum);
    return total;
}

// ------------------------------------------------------- per-vertex work
// Whole workgroup counts vertex u with a table at `tab` (LDS or HBM) of
// capacity 2^cap_lg.  misc[0]: unit cursor, misc[2]: window total;
// ctot: kMaxChunks+1 ints of LDS.
template <typename OffT, bool kGlobal, int kMode, int kU>
__device__ __forceinline__ void group_vertex(const TcArgs<OffT> &a, int u, int *tab, int cap_lg,
                                             int *mark, int *misc, int *ctot,
                                             unsigned long long &acc, unsigned long long &probes)
{
    constexpr bool kInstr = kMode == kModeInstr;
    constexpr bool kTable = kMode != kModeGather;
    if constexpr (kMode == kModeBatch)
        graph_enter(a, mark + kWave, u, acc);
    unsigned long long t0 = 0;
    if (kInstr && threadIdx.x == 0)
        t0 = clock64();
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const int nwaves = blockDim.x / kWave;
    const OffT sH = a.begH[u];
    const int dH = (int)(a.begH[u + 1] - sH);
    const OffT sW = a.begW[u];
    const int dW = (int)(a.endW[u] - sW);
    const int lg = table_lg(dH, cap_lg);
    const int size = 1 << lg;
    if constexpr (kTable)
        for (int i = threadIdx.x; i < size; i += blockDim.x)
            slot_clear<kGlobal>(tab + i);
    if (threadIdx.x == 0) {
        misc[0] = 0;                       // window-unit cursor
        misc[3] = a.force_linear == 1;     // linear-probing table for this vertex
    }
    if constexpr (kGlobal)
        __threadfence();  // HBM slots: cleared values visible in L2 before any insert/lookup
    __syncthreads();
    bool linear = false;
    if constexpr (kTable) {
        bool failed = a.force_linear == 2;  // forced: rebuild after the cuckoo attempt
        if (a.force_linear != 1)
            for (int i = threadIdx.x; i < dH; i += blockDim.x)
                failed |= cuckoo_insert(tab, lg - 1, a.adjH[sH + (OffT)i]);
        if (failed)
            misc[3] = 1;  // benign race: every writer stores 1
        if constexpr (kGlobal)
            __threadfence();
        __syncthreads();
        linear = misc[3] != 0;
        if (linear) {  // rebuild this vertex's table with linear probing
            if (threadIdx.x == 0 && a.force_linear != 1)
                atomicAdd(&a.ctl->rebuilds, 1ull);
            __syncthreads();
            for (int i = threadIdx.x; i < size; i += blockDim.x)
                slot_clear<kGlobal>(tab + i);
            if constexpr (kGlobal)
                __threadfence();
            __syncthreads();
            for (int i = threadIdx.x; i < dH; i += blockDim.x)
                table_insert(tab, lg, a.adjH[sH + (OffT)i]);
        }
    }
    if constexpr (kGlobal)
        __threadfence();
    __syncthreads();
    const int nchunks = (dW + kWave - 1) / kWave;
    unsigned long long hits = 0;  // wave-uniform
    unsigned long long wave_probes = 0;
    if (nchunks <= kMaxChunks) {
        // window count of every 64-neighbour chunk, then an exclusive prefix
        for (int c = wave; c < nchunks; c += nwaves) {
            const int idx = c * kWave + lane;
            const Group<OffT> g = walk_group(a, sW + (OffT)idx, idx < dW);
            if (lane == 0)
                ctot[c] = (g.total + kWave - 1) / kWave;
            wave_probes += (unsigned long long)g.total;
        }
        __syncthreads();
        if (wave == 0) {
            int run = 0;
            for (int c0 = 0; c0 < nchunks; c0 += kWave) {
                const int x = c0 + lane < nchunks ? ctot[c0 + lane] : 0;
                const int inc = wave_scan_add(x);
                if (c0 + lane < nchunks)
                    ctot[c0 + lane] = run + inc - x;
                run += readlane(inc, 63);
            }
            if (lane == 0) {
                ctot[nchunks] = run;
                misc[2] = run;
            }
        }
        __syncthreads();
        const int TW = misc[2];
        // units of K windows, K a multiple of kU
        int K = TW / (nwaves * 4);
        K = K < kU ? kU : (K > 64 ? 64 : K);
        K = (K + kU - 1) / kU * kU;
        int cur = -1;
        Group<OffT> g = {0, 0, 0, 0, 0};
        for (;;) {
            int k = 0;
            if (lane == 0)
                k = atomicAdd(&misc[0], 1);
            k = __builtin_amdgcn_readfirstlane(k);
            int w0 = k * K;
            if (w0 >= TW)
                break;
            const int w1 = min(w0 + K, TW);
            while (w0 < w1) {
                int lo = 0, hi = nchunks;  // ctot[lo] <= w0 < ctot[hi]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (ctot[mid] <= w0)
                        lo = mid;
                    else
                        hi = mid;
                }
                if (lo != cur) {
                    const int idx = lo * kWave + lane;
                    g = walk_group(a, sW + (OffT)idx, idx < dW);
                    cur = lo;
                }
                const int cend = min(w1, ctot[lo + 1]);
                hits += probe_windows<OffT, kGlobal, kMode, kU>(g, w0 - ctot[lo], cend - ctot[lo], a.adjP,
                                                     tab, lg, linear, mark);
                w0 = cend;
            }
        }
    } else {
        // very long walk lists: whole chunks pulled from the cursor
        for (;;) {
            int c = 0;
            if (lane == 0)
                c = atomicAdd(&misc[0], 1);
            c = __builtin_amdgcn_readfirstlane(c);
            if (c >= nchunks)
                break;
            const int idx = c * kWave + lane;
            const Group<OffT> g = walk_group(a, sW + (OffT)idx, idx < dW);
            wave_probes += (unsigned long long)g.total;
            hits += probe_windows<OffT, kGlobal, kMode, kU>(g, 0, (g.total + kWave - 1) / kWave, a.adjP, tab,
                                                 lg, linear, mark);
        }
    }
    acc += lane == 0 ? hits : 0ull;  // the wave's hits count once
    probes += wave_probes;
    __syncthreads();
    if constexpr (kInstr) {
        if (lane == 0 && wave_probes)
            atomicAdd(a.instr_probes + u, wave_probes);
        if (threadIdx.x == 0)
            a.instr_cycles[u] = clock64() - t0;
    }
}

// One wavefront counts vertex u with its own LDS table slice.
template <typename OffT, int kMode, int kU>
__device__ __forceinline__ void wave_vertex(const TcArgs<OffT> &a, int u, int *tab, int *mark,
                                            unsigned long long &acc, unsigned long long &probes)
{
    constexpr bool kInstr = kMode == kModeInstr;
    constexpr bool kTable = kMode != kModeGather;
    if constexpr (kMode == kModeBatch)
        graph_enter(a, mark + kWave, u, acc);
    unsigned long long t0 = 0;
    if (kInstr)
        t0 = clock64();
    const int lane = lane_id();
    const OffT sH = a.begH[u];
    const int dH = (int)(a.begH[u + 1] - sH);
    const OffT sW = a.begW[u];
    const int dW = (int)(a.endW[u] - sW);
    const int lg = table_lg(dH, a.wave_lg);
    const int size = 1 << lg;
    // On an unpartitioned graph the hashed row and the walked row are one
    // row: its first 64 ids, loaded for the table, also start the walk
    // (one dependent load less per vertex).
    const bool same_row = kTable && a.adjH == a.adjW && sH == sW && dH == dW;
    int head = 0;
    if constexpr (kTable)
        head = lane < dH ? a.adjH[sH + (OffT)lane] : 0;
    bool linear = false;
    if constexpr (kTable) {
        for (int i = lane; i < size; i += kWave)
            tab[i] = kEmpty;
        wave_sync();
        // force_linear 1: linear probing only; 2: rebuild after the cuckoo attempt
        bool failed = a.force_linear != 0;
        if (a.force_linear != 1) {
            if (lane < dH)
                failed |= cuckoo_insert(tab, lg - 1, head);
            for (int i = lane + kWave; i < dH; i += kWave)
                failed |= cuckoo_insert(tab, lg - 1, a.adjH[sH + (OffT)i]);
        }
        wave_sync();
        linear = __ballot(failed) != 0;
        if (linear) {  // rebuild this vertex's table with linear probing
            if (lane == 0 && a.force_linear != 1)
                atomicAdd(&a.ctl->rebuilds, 1ull);
            for (int i = lane; i < size; i += kWave)
                tab[i] = kEmpty;
            wave_sync();
            for (int i = lane; i < dH; i += kWave)
                table_insert(tab, lg, a.adjH[sH + (OffT)i]);
            wave_sync();
        }
    }
    unsigned long long hits = 0;  // wave-uniform
    unsigned long long wave_probes = 0;
    for (int c = 0; c < dW; c += kWave) {
        const bool valid = c + lane < dW;
        Group<OffT> g;
        if (a.entW) {
            g = walk_group(a, sW + (OffT)(c + lane), valid);
        } else {
            const int v = c == 0 && same_row ? head : (valid ? a.adjW[sW + (OffT)(c + lane)] : 0);
            g = load_group(v, valid, a.begP);
        }
        wave_probes += (unsigned long long)g.total;
        hits += probe_windows<OffT, false, kMode, kU>(g, 0, (g.total + kWave - 1) / kWave, a.adjP, tab, lg,
                                           linear, mark);
    }
    wave_sync();
    acc += lane == 0 ? hits : 0ull;  // the wave's hits count once
    probes += wave_probes;
    if constexpr (kInstr) {
        if (lane == 0) {
            a.instr_probes[u] = wave_probes;
            a.instr_cycles[u] = clock64() - t0;
        }
    }
}

// ------------------------------------------------------------ the kernel
template <typename OffT, int kMode, bool kGlobal, int kU>
__device__ __forceinline__ void group_phase(const TcArgs<OffT> &a, int tier, int *tab, int cap_lg,
                                            int *mark, int *misc, int *ctot,
                                            unsigned long long &acc, unsigned long long &probes)
{
    const long long lb = a.ctl->list_begin[tier], le = a.ctl->list_begin[tier + 1];
    if (le <= lb)
        return;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0)
            misc[1] = (int)atomicAdd(&a.ctl->next[tier], 1u);
        __syncthreads();
        const long long c = (long long)(unsigned int)misc[1] * a.nranks + a.rank;
        const long long first = lb + c * a.chunk_group;
        if (first >= le)
            break;
        const long long last = min(first + (long long)a.chunk_group, le);
        for (long long i = first; i < last; i++)
            group_vertex<OffT, kGlobal, kMode, kU>(a, a.list[i], tab, cap_lg, mark, misc, ctot, acc,
                                                probes);
    }
}

// Waves per SIMD: with int32 offsets eight, and the compiler fits the
// kernel in 64 VGPRs (a few spills outside the probe loop).  The walk is
// latency-bound, and the eighth wave measured 15% faster at R-MAT 24 than
// seven waves without spills (1.6x at 1024-thread workgroups).  With int64
// offsets (64-bit window bases and list starts) 64 VGPRs spill inside the
// probe loop: seven waves (72 VGPRs) measured 25.4 ms on a 1000-graph batch
// against 33.4 ms at eight, 26.3 at six.  kBig: the large-LDS launch
// (HBM-table spill tier + big tier); otherwise the main launch (hub + wave
// tiers).  Each is compiled with only its own phases, so the registers of
// one do not constrain the other.
template <typename OffT>
constexpr int kMinWavesPerEU = sizeof(OffT) == 4 ? 8 : 7;

template <typename OffT, int kMode, bool kBig, int kU = kUnroll>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(kMinWavesPerEU<OffT>, 8)))
tc_count_kernel(TcArgs<OffT> a)
{
    extern __shared__ int lds[];
    const int nwaves = blockDim.x / kWave;
    const int wave = threadIdx.x / kWave;
    const int lane = lane_id();
    int *mark = lds + a.table_words + wave * kMarkStride;  // + the graph state (batch mode)
    int *misc = lds + a.table_words + nwaves * kMarkStride;  // 8 words
    int *ctot = misc + 8;                                     // kMaxChunks + 1 words
    mark[lane] = 0;
    unsigned long long acc = 0, probes = 0;
    if constexpr (kMode == kModeBatch) {
        graph_init(mark + kWave);
        if (threadIdx.x == 0)
            atomicMin(a.graph_t0, (unsigned long long)wall_clock64());
    }

    if (kBig && a.spill_tier >= 0 && (int)blockIdx.x < a.spill_groups) {
        int *tab = a.spill_mem + (long long)blockIdx.x * a.spill_slots;
        int cap_lg = 0;
        while ((2ll << cap_lg) <= a.spill_slots)
            cap_lg++;
        group_phase<OffT, kMode, true, kU>(a, a.spill_tier, tab, cap_lg, mark, misc, ctot, acc, probes);
    }
    if (a.group_tier >= 0)
        group_phase<OffT, kMode, false, kU>(a, a.group_tier, lds, a.group_lg, mark, misc, ctot, acc,
                                        probes);
    __syncthreads();  // the wave tables reuse the group table's LDS
    if (!kBig && a.wave_tier >= 0) {
        const long long lb = a.ctl->list_begin[a.wave_tier];
        const long long le = a.ctl->list_begin[a.wave_tier + 1];
        int *wtab = lds + wave * a.wave_slots;
        // Shard s hands out the rank's chunks j with j % kQueueShards == s
        // (chunk c = j * nranks + rank, the reference's interleave), so the
        // list order is kept across shards.  A wave starts on shard
        // blockIdx % kQueueShards and moves on when a shard runs dry; after
        // finding every shard dry once it is done.
        //
        // Batches instead give shard s the s-th contiguous eighth of the
        // rank's chunks and start a wave on the shard of its XCD: the list
        // is in graph order, so each XCD counts its own graphs and their
        // lists stay in its L2 (10k-graph ER/BA batch: 4.8% faster).  A
        // single large graph keeps the interleave, which spreads each
        // vertex range over all eight L2s (R-MAT 26: contiguous shards
        // measured up to 3% slower).
        constexpr bool kContig = kMode == kModeBatch;
        int shard = kContig ? xcc_id() : (int)(blockIdx.x % kQueueShards);
        long long per = 0;  // chunks per contiguous shard
        if constexpr (kContig) {
            const long long nchunk = (le - lb + a.chunk_wave - 1) / a.chunk_wave;
            const long long nmine = nchunk > a.rank ? (nchunk - a.rank + a.nranks - 1) / a.nranks : 0;
            per = (nmine + kQueueShards - 1) / kQueueShards;
        }
        int dry = 0;
        while (dry < kQueueShards) {
            unsigned int k = 0;
            if (lane == 0)
                k = atomicAdd(&a.ctl->shard[shard * kShardStride], 1u);
            k = __builtin_amdgcn_readfirstlane(k);
            const long long j = kContig ? (long long)shard * per + (long long)k
                                        : (long long)k * kQueueShards + shard;
            const long long first = lb + (j * a.nranks + a.rank) * a.chunk_wave;
            if ((kContig && (long long)k >= per) || first >= le) {
                dry++;
                shard = shard + 1 == kQueueShards ? 0 : shard + 1;
                continue;
            }
            const long long last = min(first + (long long)a.chunk_wave, le);
            for (long long i = first; i < last; i++)
                wave_vertex<OffT, kMode, kU>(a, a.list[i], wtab, mark, acc, probes);
        }
    }
    if constexpr (kMode == kModeBatch)
        graph_flush(a, mark + kWave, acc);  // acc is now 0: the batch total is the graphs' sum
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
    const long long dW = (long long)(p.endW[u] - p.begW[u]);
    if (dH == 0 || dW == 0 || (p.same_csr && dH < 2))
        return kTierNone;  // contributes nothing (reference skips d+ < 2, TC.cu:440-441)
    int lg = kMinTableLg;
    while ((1ll << lg) < 2 * dH)
        lg++;
    if (p.force_spill || lg > max(p.hub_lg, p.big_lg))
        return kTierSpill;
    if (lg > p.hub_lg)
        return kTierBig;
    if (!p.wave_only && dW > p.hub_threshold)
        return kTierHub;
    if (lg > p.wave_lg)
        return kTierHub;
    return kTierWave;
}

template <typename OffT>
__global__ void __launch_bounds__(256) plan_count_kernel(PlanArgs<OffT> p, int nblocks)
{
    __shared__ unsigned long long cnt[kNumTiers];
    if (threadIdx.x < kNumTiers)
        cnt[threadIdx.x] = 0;
    __syncthreads();
    const long long lo = (long long)blockIdx.x * p.per_block;
    const long long hi = min(lo + p.per_block, p.V);
    int local[kNumTiers] = {0, 0, 0, 0};
    for (long long u = lo + threadIdx.x; u < hi; u += blockDim.x) {
        int t = plan_tier(p, u);
#pragma unroll
        for (int q = 0; q < kNumTiers; q++)
            local[q] += t == q;
    }
    for (int t = 0; t < kNumTiers; t++)
        if (local[t])
            atomicAdd(&cnt[t], (unsigned long long)local[t]);
    __syncthreads();
    if (threadIdx.x < kNumTiers)
        p.block_counts[(long long)threadIdx.x * nblocks + blockIdx.x] = (long long)cnt[threadIdx.x];
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
#pragma unroll
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

// ------------------------------------------------------ short-side walk
// The walk rows of the short-side walk.  Row x lists first the v in N+(x)
// with d+(v) <= d+(x) (N+(v) is probed against x's table), then the u with
// u -> x and d+(u) < d+(x) (N+(u) is probed against x's table); an edge from
// a vertex with d+ < 2 closes no triangle at that edge and is left out.
// Every oriented edge that can close a triangle lands in exactly one row.
//
// A wave takes 64 consecutive rows u, whose lists are contiguous in adj, and
// walks them in coalesced 64-wide windows; the row of each position comes
// from a mark array and a max-scan.
//   kPass 0: sums the probes of the reference walk and of the short-side
//            walk into stats[0], stats[1];
//   kPass 1: kept[u] = number of kept out-entries of row u (counted in LDS,
//            each row belongs to one wave), incnt[v] += 1 per in-entry;
//   kPass 2: writes the rows: out-entries at begX[u] + (LDS slot), in-entries
//            from the row's end down (incnt counts back to zero).  An entry x
//            is stored as its list's start and length (entX = {beg[x],
//            d+(x)}, one store), which the count reads in place of x.
// The entries' rows (rowX, the learned plan's features) follow from begX
// alone and are written afterwards, in order, by side_rows_kernel: in the
// scatter pass they were a third of the in-entries' random writes.
constexpr int kSideWin = 4;

template <typename OffT, int kPass>
__global__ void __launch_bounds__(256) side_pass_kernel(const OffT *__restrict__ beg,
                                                        const int32_t *__restrict__ adj,
                                                        long long V, int *__restrict__ incnt,
                                                        int *__restrict__ kept,
                                                        const OffT *__restrict__ begX,
                                                        WalkEntry<OffT> *__restrict__ entX,
                                                        unsigned long long *__restrict__ stats)
{
    unsigned long long w_ref = 0, w_short = 0;
    __shared__ int mark_s[4][kWave * kSideWin];
    __shared__ int rowcnt_s[4][kWave];
    const int wave = threadIdx.x / kWave, lane = lane_id();
    int *mark = mark_s[wave];
    int *rowcnt = rowcnt_s[wave];
    for (int k = 0; k < kSideWin; k++)
        mark[k * kWave + lane] = 0;
    const long long step = (long long)gridDim.x * (blockDim.x / kWave) * kWave;
    for (long long u0 = ((long long)blockIdx.x * (blockDim.x / kWave) + wave) * kWave; u0 < V;
         u0 += step) {
        const int n = (int)min((long long)kWave, V - u0);
        // lane i < n holds row u0 + i; lane n (if any) the end of the range
        const OffT s = beg[u0 + min(lane, n)];
        const int d = lane < n ? (int)(beg[u0 + lane + 1] - s) : 0;
        const long long e0 = (long long)readlane(s, 0);
        const long long e1 = n == kWave ? (long long)readlane(s, 63) + readlane(d, 63)
                                        : (long long)readlane(s, n);
        rowcnt[lane] = 0;
        wave_sync();
        int carry = 0;
        for (long long p = e0; p < e1; p += kWave * kSideWin) {
            // a non-empty row starting inside this span marks its position
            if (d > 0 && (long long)s >= p && (long long)s < p + kWave * kSideWin)
                mark[(int)((long long)s - p)] = lane;
            wave_sync();
            int own[kSideWin], v[kSideWin];
#pragma unroll
            for (int k = 0; k < kSideWin; k++) {
                own[k] = max(wave_scan_max(mark[k * kWave + lane]), carry);
                carry = readlane(own[k], 63);
                const long long pos = p + k * kWave + lane;
                v[k] = pos < e1 ? adj[pos] : -1;
            }
            wave_sync();
#pragma unroll
            for (int k = 0; k < kSideWin; k++)
                mark[k * kWave + lane] = 0;
#pragma unroll
            for (int k = 0; k < kSideWin; k++) {
                const int du = shfl_idx(d, own[k]);
                OffT su = 0;  // beg[u], shuffled while every lane is active
                if constexpr (kPass == 2)
                    su = shfl_idx(s, own[k]);
                if (v[k] >= 0 && du >= 2) {
                    const OffT bv = beg[v[k]];
                    const int dv = (int)(beg[v[k] + 1] - bv);
                    if constexpr (kPass == 0) {
                        w_ref += (unsigned long long)dv;
                        w_short += (unsigned long long)min(du, dv);
                    } else if constexpr (kPass == 1) {
                        if (dv <= du)
                            atomicAdd(rowcnt + own[k], 1);
                        else
                            atomicAdd(incnt + v[k], 1);
                    } else {
                        const long long u = u0 + own[k];
                        if (dv <= du) {
                            const int slot = atomicAdd(rowcnt + own[k], 1);
                            const OffT at = begX[u] + (OffT)slot;
                            entX[at] = WalkEntry<OffT>{bv, dv};
                        } else {
                            const int k2 = atomicSub(incnt + v[k], 1);
                            const OffT at = begX[v[k] + 1] - (OffT)k2;
                            entX[at] = WalkEntry<OffT>{su, du};
                        }
                    }
                }
            }
            wave_sync();
        }
        if constexpr (kPass == 1) {
            if (lane < n)
                kept[u0 + lane] = rowcnt[lane];
        }
        wave_sync();
    }
    if constexpr (kPass == 0) {
        w_ref = wave_sum(w_ref);
        w_short = wave_sum(w_short);
        if (lane == 0 && w_ref) {
            atomicAdd(stats, w_ref);
            atomicAdd(stats + 1, w_short);
        }
    }
}

// rowX[at] = the row whose entries [begX[u], begX[u+1]) hold position at.
// A wave takes 64 consecutive rows, whose entries are contiguous, and
// writes them in 64-wide windows: each non-empty row marks its first
// position in LDS, and a max-scan carries the owner across the window.
template <typename OffT>
__global__ void __launch_bounds__(256) side_rows_kernel(const OffT *__restrict__ begX, long long V,
                                                        int32_t *__restrict__ rowX)
{
    __shared__ int mark_s[4][kWave];
    const int wave = threadIdx.x / kWave, lane = lane_id();
    int *mark = mark_s[wave];
    mark[lane] = 0;
    const long long step = (long long)gridDim.x * (blockDim.x / kWave) * kWave;
    for (long long u0 = ((long long)blockIdx.x * (blockDim.x / kWave) + wave) * kWave; u0 < V;
         u0 += step) {
        const int n = (int)min((long long)kWave, V - u0);
        // lane i < n: row u0 + i is [s, e); lane n (if any) holds the range's end in s
        const OffT s = begX[u0 + min(lane, n)];
        const OffT e = lane < n ? begX[u0 + lane + 1] : s;
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
            if (w + lane < p1)
                rowX[w + lane] = (int32_t)(u0 + own);
            wave_sync();
        }
    }
}

template <typename OffT>
hipError_t launch_side_build(const OffT *beg, const int32_t *adj, long long V, int *cnt,
                             long long *bsum, OffT *begX, WalkEntry<OffT> *entX,
                             hipStream_t stream)
{
    if (V <= 0)
        return hipSuccess;
    int *incnt = cnt, *kept = cnt + V;
    const int pass_blocks = (int)min((V + 4 * kWave - 1) / (4 * kWave), 16384ll);
    hipError_t e = launch_fill_u32((uint32_t *)incnt, 0u, V, stream);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL((side_pass_kernel<OffT, 1>), dim3(pass_blocks), dim3(256), 0, stream,
                       beg, adj, V, incnt, kept, (const OffT *)nullptr,
                       (WalkEntry<OffT> *)nullptr, (unsigned long long *)nullptr);
    if ((e = launch_scan<OffT>(incnt, kept, V, bsum, begX, stream)) != hipSuccess)
        return e;
    hipLaunchKernelGGL((side_pass_kernel<OffT, 2>), dim3(pass_blocks), dim3(256), 0, stream,
                       beg, adj, V, incnt, (int *)nullptr, (const OffT *)begX, entX,
                       (unsigned long long *)nullptr);
    return hipGetLastError();
}

template <typename OffT>
hipError_t launch_side_rows(const OffT *begX, long long V, int32_t *rowX, hipStream_t stream)
{
    if (V <= 0)
        return hipSuccess;
    const int row_blocks = (int)min((V + 4 * kWave - 1) / (4 * kWave), 16384ll);
    hipLaunchKernelGGL((side_rows_kernel<OffT>), dim3(row_blocks), dim3(256), 0, stream, begX, V,
                       rowX);
    return hipGetLastError();
}

template hipError_t launch_side_build<int32_t>(const int32_t *, const int32_t *, long long, int *,
                                               long long *, int32_t *, WalkEntry<int32_t> *,
                                               hipStream_t);
template hipError_t launch_side_build<int64_t>(const int64_t *, const int32_t *, long long, int *,
                                               long long *, int64_t *, WalkEntry<int64_t> *,
                                               hipStream_t);
template hipError_t launch_side_rows<int32_t>(const int32_t *, long long, int32_t *, hipStream_t);
template hipError_t launch_side_rows<int64_t>(const int64_t *, long long, int32_t *, hipStream_t);

template <typename OffT>
hipError_t launch_side_stats(const OffT *beg, const int32_t *adj, long long V,
                             unsigned long long *stats, hipStream_t stream)
{
    hipError_t e = launch_fill_u32((uint32_t *)stats, 0u, 4, stream);
    if (e != hipSuccess || V <= 0)
        return e;
    const int pass_blocks = (int)min((V + 4 * kWave - 1) / (4 * kWave), 16384ll);
    hipLaunchKernelGGL((side_pass_kernel<OffT, 0>), dim3(pass_blocks), dim3(256), 0, stream,
                       beg, adj, V, (int *)nullptr, (int *)nullptr, (const OffT *)nullptr,
                       (WalkEntry<OffT> *)nullptr, stats);
    return hipGetLastError();
}

template <typename OffT>
__global__ void __launch_bounds__(256) start_index_kernel(const OffT *__restrict__ begP,
                                                          const int32_t *__restrict__ adj,
                                                          long long E,
                                                          WalkEntry<OffT> *__restrict__ entR)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) {
        const int v = adj[e];
        const OffT s = begP[v];
        entR[e] = WalkEntry<OffT>{s, (int32_t)(begP[v + 1] - s)};
    }
}

template <typename OffT>
hipError_t launch_start_index(const OffT *begP, const int32_t *adj, long long E,
                              WalkEntry<OffT> *entR, hipStream_t stream)
{
    if (E <= 0)
        return hipSuccess;
    const int blocks = (int)min((E + 255) / 256, 32768ll);
    hipLaunchKernelGGL((start_index_kernel<OffT>), dim3(blocks), dim3(256), 0, stream, begP, adj, E,
                       entR);
    return hipGetLastError();
}

template hipError_t launch_start_index<int32_t>(const int32_t *, const int32_t *, long long,
                                                WalkEntry<int32_t> *, hipStream_t);
template hipError_t launch_start_index<int64_t>(const int64_t *, const int32_t *, long long,
                                                WalkEntry<int64_t> *, hipStream_t);

template hipError_t launch_side_stats<int32_t>(const int32_t *, const int32_t *, long long,
                                               unsigned long long *, hipStream_t);
template hipError_t launch_side_stats<int64_t>(const int64_t *, const int32_t *, long long,
                                               unsigned long long *, hipStream_t);

// Plain streaming copy: the known-traffic kernel that calibrates HBM byte
// counters and gives the sequential-bandwidth reference.
__global__ void __launch_bounds__(256) copy_kernel(const int4 *__restrict__ src,
                                                   int4 *__restrict__ dst, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy(const int4 *src, int4 *dst, long long n, int blocks, hipStream_t stream)
{
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, stream, src, dst, n);
    return hipGetLastError();
}

// Every workgroup reads the whole buffer `reps` times from its own starting
// point: a buffer that fits the L2 measures the L2s' read rate (the dense
// core's bit rows are read from there).
__global__ void __launch_bounds__(256) read_kernel(const int4 *__restrict__ src, long long n,
                                                   int reps, unsigned *__restrict__ sink)
{
    unsigned acc = 0;
    const long long start = ((long long)blockIdx.x * 977) % n;
    for (int r = 0; r < reps; r++)
        for (long long i = threadIdx.x; i < n; i += blockDim.x) {
            long long j = i + start;
            if (j >= n)
                j -= n;
            const int4 x = src[j];
            acc += (unsigned)(x.x ^ x.y ^ x.z ^ x.w);
        }
    if (acc == 0x9e3779b9u)  // keeps the loads; practically never stores
        sink[blockIdx.x] = acc;
}

hipError_t launch_read(const int4 *src, long long n, int reps, int blocks, unsigned *sink,
                       hipStream_t stream)
{
    hipLaunchKernelGGL(read_kernel, dim3(blocks), dim3(256), 0, stream, src, n, reps, sink);
    return hipGetLastError();
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

template <typename OffT, bool kBig>
static void launch_count_t(const TcArgs<OffT> &a, int blocks, int threads, size_t lds_bytes,
                           int mode, hipStream_t stream)
{
    const bool deep = (mode & kProbeDeep) != 0;
    mode &= ~kProbeDeep;
    if (mode == kModeInstr)
        hipLaunchKernelGGL((tc_count_kernel<OffT, kModeInstr, kBig>), dim3(blocks), dim3(threads),
                           lds_bytes, stream, a);
    else if (mode == kModeGather && deep)
        hipLaunchKernelGGL((tc_count_kernel<OffT, kModeGather, kBig, kUnrollDeep>), dim3(blocks),
                           dim3(threads), lds_bytes, stream, a);
    else if (mode == kModeGather)
        hipLaunchKernelGGL((tc_count_kernel<OffT, kModeGather, kBig>), dim3(blocks), dim3(threads),
                           lds_bytes, stream, a);
    else if (mode == kModeBatch)
        hipLaunchKernelGGL((tc_count_kernel<OffT, kModeBatch, kBig>), dim3(blocks), dim3(threads),
                           lds_bytes, stream, a);
    else if (deep)
        hipLaunchKernelGGL((tc_count_kernel<OffT, kModeCount, kBig, kUnrollDeep>), dim3(blocks),
                           dim3(threads), lds_bytes, stream, a);
    else
        hipLaunchKernelGGL((tc_count_kernel<OffT, kModeCount, kBig>), dim3(blocks), dim3(threads),
                           lds_bytes, stream, a);
}

template <typename OffT>
hipError_t launch_count(const TcArgs<OffT> &a, int blocks, int threads, size_t lds_bytes,
                        int mode, bool big, hipStream_t stream)
{
    if (big)
        launch_count_t<OffT, true>(a, blocks, threads, lds_bytes, mode, stream);
    else
        launch_count_t<OffT, false>(a, blocks, threads, lds_bytes, mode, stream);
    return hipGetLastError();
}

template <typename OffT, bool kBig>
static hipError_t count_occupancy_t(int threads, size_t lds_bytes, int mode, int *blocks_per_cu)
{
    const bool deep = (mode & kProbeDeep) != 0;
    mode &= ~kProbeDeep;
    if (mode == kModeInstr)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            blocks_per_cu, tc_count_kernel<OffT, kModeInstr, kBig>, threads, lds_bytes);
    if (mode == kModeGather)
        return deep ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          blocks_per_cu, tc_count_kernel<OffT, kModeGather, kBig, kUnrollDeep>,
                          threads, lds_bytes)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          blocks_per_cu, tc_count_kernel<OffT, kModeGather, kBig>, threads,
                          lds_bytes);
    if (mode == kModeBatch)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            blocks_per_cu, tc_count_kernel<OffT, kModeBatch, kBig>, threads, lds_bytes);
    return deep ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                      blocks_per_cu, tc_count_kernel<OffT, kModeCount, kBig, kUnrollDeep>, threads,
                      lds_bytes)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                      blocks_per_cu, tc_count_kernel<OffT, kModeCount, kBig>, threads, lds_bytes);
}

template <typename OffT>
hipError_t count_occupancy(int threads, size_t lds_bytes, int mode, bool big, int *blocks_per_cu)
{
    return big ? count_occupancy_t<OffT, true>(threads, lds_bytes, mode, blocks_per_cu)
               : count_occupancy_t<OffT, false>(threads, lds_bytes, mode, blocks_per_cu);
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
template hipError_t launch_count<int32_t>(const TcArgs<int32_t> &, int, int, size_t, int, bool,
                                          hipStream_t);
template hipError_t launch_count<int64_t>(const TcArgs<int64_t> &, int, int, size_t, int, bool,
                                          hipStream_t);
template hipError_t count_occupancy<int32_t>(int, size_t, int, bool, int *);
template hipError_t count_occupancy<int64_t>(int, size_t, int, bool, int *);

}  // namespace trust
