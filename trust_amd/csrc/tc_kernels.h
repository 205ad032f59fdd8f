// Device-side structures and launch wrappers of the counting kernels.
// Shared by tc_kernels.hip (device code) and the host engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace trust {

// Work tiers.  A vertex u is counted by exactly one tier.
//   spill: whole workgroup, hash table in HBM (table too big for any LDS)
//   big:   whole workgroup of the large-LDS launch, table in LDS
//   hub:   whole workgroup of the main launch, table in LDS
//   wave:  one wavefront, table in its LDS slice
enum Tier : int { kTierSpill = 0, kTierBig = 1, kTierHub = 2, kTierWave = 3, kTierNone = 4 };
constexpr int kNumTiers = 4;
constexpr int kWave = 64;
constexpr int kMinTableLg = 6;       // smallest table: 64 slots
constexpr int kPlanPerBlock = 4096;  // vertices per planning workgroup
constexpr int kMaxChunks = 256;      // 64-neighbour chunks balanced per group vertex
// Probe windows in flight per wave (the probe walk's batch): four, or six
// with kProbeDeep.  Six keep more loads in flight where the probed lists are
// long (the short-side walk: R-MAT 26 count 3% faster, R-MAT 24 2%) and
// waste more lookups on partial batches where they are short (Chung-Lu with
// the reference walk 13% slower, the 10k ER/BA batch 4%).
constexpr int kUnroll = 4;
constexpr int kUnrollDeep = 6;
// int32 offsets (and 32-bit byte offsets into adj) are used when a CSR holds
// fewer than kMaxEdges32 ids; larger ones keep int64 offsets.
constexpr long long kMaxEdges32 = 1ll << 30;           // probe windows in flight per wave

// Kernel modes: count triangles, count + per-vertex instrumentation, or
// gather only (the same loads, every hash lookup replaced by a checksum of
// the loaded ids: the load path alone, used as the roofline denominator).
constexpr int kModeCount = 0, kModeInstr = 1, kModeGather = 2;
// kModeBatch: kModeCount plus per-graph attribution (TcArgs::graph_*)
constexpr int kModeBatch = 3;
// OR-ed into kModeCount or kModeGather at enqueue/launch: the deep probe walk
constexpr int kProbeDeep = 16;

// Device control block: queue cursors, list bounds and results.  Cleared
// by a fill kernel before a run.
// The wave queue is split over kQueueShards counters, each on its own
// 128-byte line: one address sustains only about 1e8 atomics/s, which
// bounds graphs of many small vertices.
constexpr int kQueueShards = 8;
constexpr int kShardStride = 32;  // unsigned ints per counter line
struct TcCtl {
    unsigned long long count;       // triangles
    unsigned long long probes;      // hash lookups (W)
    unsigned int next[8];           // queue cursors, one per tier
    long long list_begin[8];        // tier t occupies list[list_begin[t], list_begin[t+1])
    unsigned long long rebuilds;    // tables rebuilt with linear probing after a cuckoo attempt
    unsigned long long pad[3];
    unsigned int shard[kQueueShards * kShardStride];  // wave-queue cursors
};

// Three CSR views: the hash source (N+(u) that is tabled), the walk (the v
// in N+(u) whose lists are probed) and the probe rows (N+(v)).  For an
// unpartitioned graph all three are the same arrays; for a 2D tile job they
// are tiles (i,j), (i,k) and (k,j).
// A walk entry of an index: the start of the list it probes and that
// list's length, read together with one load and written with one store.
template <typename OffT>
struct alignas(2 * sizeof(OffT)) WalkEntry {
    OffT start;
    int32_t deg;
};

template <typename OffT>
struct TcArgs {
    const OffT *begH;
    const int32_t *adjH;
    const OffT *begW;        // row u of the walk is adjW[begW[u], endW[u])
    const OffT *endW;
    const int32_t *adjW;
    const OffT *begP;
    const int32_t *adjP;
    const int32_t *list;     // vertices grouped by tier (see TcCtl::list_begin)
    int32_t *spill_mem;      // HBM hash tables, spill_slots ints per workgroup
    long long spill_slots;
    int spill_groups;        // workgroups that own an HBM table
    int spill_tier;          // tier counted with HBM tables here, or -1
    int group_tier;          // tier counted by whole workgroups with the LDS table, or -1
    int wave_tier;           // tier counted by single waves, or -1
    int table_words;         // LDS ints for tables (group table = all of it)
    int group_lg;            // log2 of the group table capacity
    int wave_slots;          // LDS ints per wave table
    int wave_lg;
    int chunk_group;         // vertices per grab, spill/group tiers
    int chunk_wave;          // vertices per grab, wave tier
    int rank;
    int nranks;
    int force_linear;        // 1: linear-probing tables only (no cuckoo); 2: every cuckoo
                             // table is rebuilt with linear probing (a forced failure)
    // batch of graphs (nullable): graph g owns vertices [vbase[g], vbase[g+1])
    const long long *graph_vbase;
    // graph_page[p] = the graph holding vertex min(p << shift, V - 1)
    const int *graph_page;
    int graph_page_shift;
    int ngraphs;
    unsigned long long *graph_counts;  // per-graph triangles
    unsigned long long *graph_done;    // per-graph completion (wall-clock ticks)
    unsigned long long *graph_t0;      // earliest workgroup start (wall-clock ticks)
    TcCtl *ctl;
    unsigned long long *instr_cycles;  // per vertex (instrumented kernel only)
    unsigned long long *instr_probes;
    // Walk entries stored as the start and length of the list they probe
    // (nullable; a walk index).  Then adjW is not read.
    const WalkEntry<OffT> *entW;
};

// Planner parameters for the degree split.
template <typename OffT>
struct PlanArgs {
    const OffT *begH;
    const OffT *begW;
    const OffT *endW;
    long long V;
    int same_csr;       // H, W, P are one CSR: skip d+ < 2 as the reference does
    int hub_threshold;  // walk degree above which the workgroup tier is used
    int wave_lg;        // log2 of a wave table
    int hub_lg;         // log2 of the main launch's group table
    int big_lg;         // log2 of the large-LDS launch's table (0: no such launch)
    int force_spill;
    int wave_only;
    const unsigned char *tier_override;  // nullable: per-vertex tier chosen elsewhere
    int32_t *list;
    long long *block_counts;  // [kNumTiers][nblocks]
    int per_block;
    TcCtl *ctl;
};

// Launch wrappers (tc_kernels.hip).  All launches go on `stream`.
template <typename OffT>
hipError_t launch_plan(const PlanArgs<OffT> &p, int nblocks, hipStream_t stream);
template <typename OffT>
// big: the large-LDS launch (spill + big tiers), else the main launch.
hipError_t launch_count(const TcArgs<OffT> &a, int blocks, int threads, size_t lds_bytes,
                        int mode, bool big, hipStream_t stream);
template <typename OffT>
hipError_t count_occupancy(int threads, size_t lds_bytes, int mode, bool big, int *blocks_per_cu);
hipError_t launch_copy(const int4 *src, int4 *dst, long long n, int blocks, hipStream_t stream);
// every one of `blocks` workgroups reads src[0, n) `reps` times (sink: blocks words)
hipError_t launch_read(const int4 *src, long long n, int reps, int blocks, unsigned *sink,
                       hipStream_t stream);
hipError_t launch_offsets_to_i32(const int64_t *src, int32_t *dst, long long n,
                                 hipStream_t stream);
hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, long long n, hipStream_t stream);

// Exclusive scan of a[i] (+ b[i] when b is not null), i < n, into
// out[0, n], out[n] = the total; bsum holds scan_blocks(n) + 1 long longs
// (walk_index.hip).  Launches only.
long long scan_blocks(long long n);
template <typename OutT>
hipError_t launch_scan(const int *a, const int *b, long long n, long long *bsum, OutT *out,
                       hipStream_t stream);

// Builds the walk rows of the short-side walk for an oriented CSR (the walk
// view of the three-view kernel) with one global atomic per in-entry to
// count its row and one to claim its slot: the build for graphs with too
// many target bins for the binned build (walk_index.hip).  Row x = the v in
// N+(x) with d+(v) <= d+(x), then the u with u -> x and d+(u) < d+(x), rows
// with d+ < 2 contributing nothing.  cnt: 2V ints of scratch; bsum:
// scan_blocks(V) + 1 long longs; begX: V + 1 offsets; entX: up to E (the
// start and length d+ of the entry's list in adj).  Launches only.
template <typename OffT>
hipError_t launch_side_build(const OffT *beg, const int32_t *adj, long long V, int *cnt,
                             long long *bsum, OffT *begX, WalkEntry<OffT> *entX,
                             hipStream_t stream);
// rowX[at] = the row x whose entries [begX[x], begX[x+1]) hold position at.
template <typename OffT>
hipError_t launch_side_rows(const OffT *begX, long long V, int32_t *rowX, hipStream_t stream);
// Start index of the reference walk: for e < E, entR[e] = {begP[adj[e]],
// its list length}.  Launches only.
template <typename OffT>
hipError_t launch_start_index(const OffT *begP, const int32_t *adj, long long E,
                              WalkEntry<OffT> *entR, hipStream_t stream);
// stats[0] = probes of the reference walk, stats[1] = of the short-side walk
// (stats: 2 unsigned long longs, cleared here).
template <typename OffT>
hipError_t launch_side_stats(const OffT *beg, const int32_t *adj, long long V,
                             unsigned long long *stats, hipStream_t stream);
// The short-side walk is chosen by default when it makes at most this share
// of the reference walk's probes (its lists are shorter, so a probe costs
// more; measured break-even near 0.45-0.5 on R-MAT).
constexpr double kShortWalkShare = 0.45;

// Per-wave LDS ints after the tables: a 64-int mark array, then the wave's
// graph state of the batch mode.
constexpr int kMarkStride = kWave + 4;
// LDS ints besides the tables: per-wave marks, misc words, chunk totals.
inline int lds_extra_words(int threads)
{
    return (threads / kWave) * kMarkStride + 8 + kMaxChunks + 8;
}

// log2 of the smallest table for a hash source of degree d: >= 2d slots, >= 64.
inline int host_table_lg(long long d)
{
    int lg = kMinTableLg;
    while ((1ll << lg) < 2 * d)
        lg++;
    return lg;
}

}  // namespace trust
