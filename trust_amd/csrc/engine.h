// Host-engine internals shared by engine.hip, partition_driver.hip,
// learned.hip and batch_driver.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "internal.h"
#include "tc_kernels.h"

#define HIPCHK(expr)                                      \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess)                             \
            return ::trust::hip_fail(_e, #expr);          \
    } while (0)

namespace trust {

int hip_fail(hipError_t e, const char *what);

// Grow-only device allocation.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t n);
    void release();
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
};

// A CSR resident in HBM.  Offsets are int32 when E < 2^30 (halves the
// offset traffic of the probe walk) unless int64 is requested.
struct DevCsr {
    DevBuf beg, adj;
    int64_t V = 0, E = 0;
    int64_t max_deg = 0;
    bool off64 = false;
};

struct CsrViews {
    const void *begH;
    const int32_t *adjH;
    const void *begW;
    const int32_t *adjW;
    const void *begP;
    const int32_t *adjP;
    // walk entries as WalkEntry<OffT> (list start and length; null: adjW
    // holds ids)
    const void *entW = nullptr;
};

// Shape of one counting launch.
struct LaunchShape {
    int threads = 0, blocks = 0;
    int wave_slots = 0, wave_lg = 0;
    int table_words = 0, group_lg = 0;
    size_t lds_bytes = 0;
};

// The counting launches of one count: the main launch (hub + wave tiers,
// small LDS, full occupancy) and, when some table is too large for it, a
// large-LDS launch (big tier) that also runs the HBM-table spill tier.
struct LaunchPlan {
    LaunchShape main, big;
    bool has_big = false;
    int chunk_group = 1, chunk_wave = 1;
    long long spill_slots = 0;
    int spill_groups = 0;
};

// Vertices per page of the batch's graph lookup table.
constexpr int kGraphPageShift = 12;

// A batch too large for int32 offsets is counted as segments of whole
// graphs, each with fewer than 2^30 edges (the int32 kernel's limit), whose
// offsets are int32 relative to the segment's first edge.
constexpr int64_t kSegMaxEdges = (int64_t(1) << 30) - 1;
struct BatchSeg {
    int64_t v0 = 0, v1 = 0;   // vertices [v0, v1)
    int64_t e0 = 0, e1 = 0;   // edges [e0, e1)
    int64_t g0 = 0, g1 = 0;   // graphs [g0, g1)
    int64_t beg_off = 0;      // its v1 - v0 + 1 offsets start at Ctx::seg_beg[beg_off]
    int64_t vbase_off = 0;    // its g1 - g0 + 1 local vertex bases in Ctx::seg_vbase
    int64_t page_off = 0;     // its page table in Ctx::seg_page
};
// What a counting launch of one segment reads instead of the whole batch's
// list and graph tables (vertex ids in them are segment-local).
struct SegView {
    const int32_t *list;
    const long long *graph_vbase;
    const int *graph_page;
    int ngraphs;
    unsigned long long *graph_counts, *graph_done;
};

// A slice [lo, hi) of one target bin's in-entry records (walk_index.hip);
// multi >= 0 numbers the slices of bins cut into more than one.
struct WiSlice {
    int bin;
    int multi;
    long long lo, hi;
};

// C2's work queue of the binned walk-index build (walk_index.hip): chunks
// of a slice's records, in one list per XCD.
constexpr int kXcds = 8;
struct WiChunk {
    int slice;
    int pad;
    long long lo, hi;
};
struct WiQueue {
    int begin[kXcds + 1];           // XCD x's chunks are [begin[x], begin[x+1])
    unsigned int next[kXcds * 32];  // cursor of XCD x at next[32 x] (its own 128-byte line)
};

// Scratch and host-side state of the binned walk-index build.
struct WalkIndexState {
    DevBuf M, O, bin_start, bsum, out_beg, slices, chunks, recs, rel, sbase, tmp_out, dv, dtab;
    std::vector<long long> bin_start_h;  // first record of each bin (nbins + 1)
    std::vector<WiSlice> slices_h;
    std::vector<WiChunk> chunks_h;  // host copies stay alive until their copies run
    WiQueue queue_h;
    int grid = 0, nbins = 0;
    int64_t out_total = 0;  // out-entries of all rows
    bool counted = false;   // the count pass ran on the resident graph
};

// The short-side walk of one 2D tile (i,k) (walk_index.hip): out-walk rows
// (class i, ids of class-k rows) and in-walk rows (class k, ids of class-i
// rows), each a CSR.
struct TileWalk {
    DevBuf out_beg, out_ids, in_beg, in_ids;
    int64_t rows = 0, cols = 0, x_out = 0, x_in = 0;
};

// The dense core (core.hip): at most 32768 high-degree vertices (a seed of
// K, closed under out-neighbours) with ids cid; the dense targets (core
// vertices whose out-list lies in the core, d+ >= thr) with their bit rows,
// sliced per XCD; the dense sources
// (d+ >= thr, a repeat-free core part and a dense target) with their core
// ids.  A per-graph index like the walk's, which leaves the edges from a
// dense source to a dense target to the core's launch.
struct CoreIndex {
    DevBuf flag, cid, pos, coreV, bits, rows, pool, recs, bsrc, grabs, misc, hist;
    long long nsrc_cap = 0;        // rows B(u) bsrc holds
    int want_K = 0, want_thr = 0;  // the configuration the next build uses
    bool want_auto = false;        // chosen automatically: only for heavy-tailed graphs
    int K = 0, thr = 0;            // the configuration of the built index
    int nC = 0;                    // core vertices
    int Kb = 0, S = 0;             // bits of a row, slices of the rows
    long long cap = 0;             // record slots
    long long words = 0;           // 32-bit words of one V-bit bitmap
    bool built = false;            // made for the resident graph (maybe empty)
    bool active = false;           // built with dense targets: the walk leaves its edges out
    uint64_t dense_probes = 0;     // the short walk's probes of the edges it took
    uint64_t dense_edges = 0;
    const uint32_t *dst_bits() const { return (const uint32_t *)bits.p; }
    const uint32_t *src_bits() const { return (const uint32_t *)bits.p + words; }
};

struct Ctx {
    int device = 0;
    bool ready = false;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;
    // the large-LDS launch runs here, beside the main launch (fork / join)
    hipStream_t aux_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // around the per-graph index work a count does before its plan
    hipEvent_t ev_idx[2] = {nullptr, nullptr};
    // around the dense core's launch of a timed count
    hipEvent_t ev_core[2] = {nullptr, nullptr};
    int num_cus = 0;
    size_t lds_max = 0;
    DevCsr g;
    double h2d_seconds = 0;
    double learned_plan_seconds = 0;
    DevBuf staging, list, block_counts, ctl, spill, instr_cycles, instr_probes;
    // short-side walk: probe totals of both walks and the walk rows, a
    // per-graph index made the first time a count needs it
    DevBuf side_cnt, side_bsum, side_beg, side_ent, side_row, side_stats;
    bool side_stats_ready = false, side_built = false;
    bool side_rows_built = false;  // side_row: each walk entry's row (the learned features)
    WalkIndexState wi;
    CoreIndex core;
    uint64_t side_w_ref = 0, side_w_short = 0;
    int64_t side_entries = -1;  // entries of the walk rows (-1: not read back yet)
    // learned scheduler state
    DevBuf feat, tiers, cost, mlp_params;
    DevBuf walk_acc;  // per-vertex sum and max of the short-side walk's list lengths
    // the walk whose per-vertex features feat holds (0: N+ rows, 1: the
    // short-side walk's rows; -1: none): they depend on the graph alone, so
    // a context keeps them like the walk indexes
    int feat_walk = -1;
    // start index of the reference walk (single graphs, int32 offsets): per
    // entry of adj, its list start and length
    DevBuf ref_ent;
    bool ref_built = false;
    bool mlp_ready = false;
    // the degree split as judge: which schedule counted the resident graph
    // faster when the model was trained on it (-1: not judged)
    int judge = -1;
    double judge_learned_s = 0, judge_degree_s = 0;
    // batch of graphs: vbase[ngraphs + 1]; out = counts[ng], done[ng], t0
    DevBuf batch_vbase, batch_page, batch_out;
    int64_t batch_ngraphs = 0;
    double wall_hz = 0;
    // host copies of the batch's graph vertex and edge bases (ngraphs + 1)
    std::vector<int64_t> batch_vbase_h, batch_ebase_h;
    // segments of an int64 batch (empty: the int64 kernel counts it whole)
    std::vector<BatchSeg> segs;
    DevBuf seg_beg, seg_vbase, seg_page;
    // captured plan + count of the batch (TRUST_FLAG_HIP_GRAPH)
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uintptr_t> graph_key;
    ~Ctx();
};

int ctx_init(Ctx *ctx, int device);
int upload_csr(Ctx *ctx, DevCsr &g, const int64_t *beg, const int32_t *adj, int64_t V, int64_t E,
               bool force64, hipStream_t s);
int load_ctx_csr(Ctx *ctx, const int64_t *beg, const int32_t *adj, int64_t V, int64_t E,
                 uint32_t flags);
// Launch shapes of one count; `mode` selects the kernel whose occupancy
// sizes the grid.
int resolve_launch(Ctx *ctx, const trust_params &p, int64_t max_deg_hash, int64_t V, bool off64,
                   LaunchPlan &L, int mode = kModeCount);
int check_params(const trust_params &p);
double alg_bytes(int64_t V, int64_t E, uint64_t W, bool off64);
// The same for the short-side walk, whose walk rows hold X entries.
double alg_bytes_short(int64_t V, int64_t E, uint64_t W, bool off64, int64_t X);
// The reference walk with the start index: every entry's start and length
// read in order instead of two offsets per walked vertex.
double alg_bytes_starts(int64_t V, int64_t E, uint64_t W, bool off64);

template <typename OffT>
int enqueue_count(Ctx *ctx, const CsrViews &views, int64_t V, bool same_csr,
                  const trust_params &p, const LaunchPlan &L, TcCtl *ctl_dev, int mode,
                  unsigned long long *instr_cycles, unsigned long long *instr_probes,
                  bool preplanned, hipStream_t s, hipEvent_t ev_plan, hipEvent_t ev_count,
                  const SegView *seg = nullptr, bool plan_only = false,
                  bool with_core = false);

int count_resident(Ctx *ctx, const trust_params &p, int mode,
                   unsigned long long *cycles_out, unsigned long long *probes_out,
                   trust_result *r);

// One count over the resident graph split into its steps, so that the
// enqueue step can also be captured into a hipGraph.
struct ResidentRun {
    LaunchPlan L;
    bool learned = false;
    bool side = false;           // short-side walk
    bool segmented = false;      // an int64 batch counted as int32 segments
    bool starts = false;         // the reference walk reads the start index
    bool core = false;           // the walk rows leave the dense core's edges out
    double dense_seconds = 0;    // device time of the core's launch (timed counts)
    bool index_ev = false;       // this count enqueued index work between ctx->ev_idx[0..1]
    double index_seconds = 0;    // device time of that work (after the count's sync)
};
int prepare_resident(Ctx *ctx, const trust_params &p, int mode, ResidentRun &R);
// Builds the reference walk's start index unless it exists.
int ensure_start_index(Ctx *ctx, ResidentRun &R);
// Entries of the short-side walk rows (read back once after a build).
int side_entry_count(Ctx *ctx, int64_t *n);
// Builds the short-side walk index of the resident graph if it has none.
int ensure_walk_index(Ctx *ctx);
// Writes each walk entry's row (side_row) unless written for this index.
int ensure_walk_rows(Ctx *ctx);
void release_walk_scratch(Ctx *ctx);
// Binned walk-index build (walk_index.hip).  walk_index_binned: the graph
// has few enough target bins; walk_index_count: the count pass, which also
// yields both walks' probe totals (synchronises); walk_index_build: the
// rows from the counted graph (launches only).
bool walk_index_binned(const Ctx *ctx);
// Dense core (core.hip).  core_config: the core a count with p asks for
// (K = 0: none); core_build: the index for ctx->core.want_K/want_thr
// (synchronises once); core_enqueue: the counting launch, adding to
// ctl->count; core_drop: forget the index.
void core_config(const Ctx *ctx, const trust_params &p, int *K, int *thr, bool *automatic);
int core_build(Ctx *ctx);
int core_enqueue(Ctx *ctx, TcCtl *ctl, hipStream_t s);
void core_drop(Ctx *ctx);
// core_drop and free the core's device memory (a synchronised stream)
void core_release(Ctx *ctx);
int walk_index_count(Ctx *ctx);
int walk_index_build(Ctx *ctx);
// Builds the walk of a tile from its device CSR and the global out-degrees
// of its rows' (degR) and columns' (degC) vertices; synchronises once.
template <typename OffT>
int tile_walk_build(Ctx *ctx, const OffT *beg, const int32_t *adj, long long rows, long long cols,
                    const int32_t *degR, const int32_t *degC, TileWalk &tw);
// Whether a tile walk with `cols` column vertices fits the binned build.
int tile_walk_bins_ok(long long cols);

// Releases the captured batch graph (after a new load or at teardown).
void drop_captured(Ctx *ctx);
// Cuts the loaded batch into segments of at most max_edges edges (0: none).
int build_segments(Ctx *ctx, int64_t max_edges);
// Enqueues the plans and counting launches of every segment.
int enqueue_segments(Ctx *ctx, const trust_params &p, int mode, const ResidentRun &R, bool timed);
// Copies the control blocks of the last count back (one, or one per
// segment, summed) and waits for the stream.
int read_ctl(Ctx *ctx, const ResidentRun &R, TcCtl &h);
// Enqueues the fills, plan and counting launches on ctx->stream; `timed`
// records ctx->ev[0..2] around plan and count (not under capture).
int enqueue_resident(Ctx *ctx, const trust_params &p, int mode, const ResidentRun &R, bool timed);
// Fills *r from the control blocks copied back after the count (the stream
// has been synchronised): seconds = index work of this call + plan + count.
int finish_resident(Ctx *ctx, ResidentRun &R, const TcCtl &h, double t_plan, double t_count,
                    trust_result *r);

// Learned scheduler (learned.hip): enqueues MLP inference over the features
// ensure_features made, tier choice and the longest-first counting sort that
// fill ctx->list and ctl->list_begin for the resident graph.
// begWalk: the walk offsets when the count takes the short-side walk, else
// null (the walk is N+(u)).
int enqueue_learned_plan(Ctx *ctx, const trust_params &p, const LaunchPlan &L, TcCtl *ctl,
                         hipStream_t s, const void *begWalk);
// Makes ctx->feat hold the learned plan's per-vertex features of the given
// walk (computed once per graph and walk; the caller times it as index work).
int ensure_features(Ctx *ctx, bool walk);
// The same plan for one int32 segment of a batch (views as the segment's
// counting launch reads them; V its vertices, v0 its first): features, cost
// and list go to the segment's slices of the context's buffers, the tier
// bounds to ctl.
int enqueue_learned_plan_segment(Ctx *ctx, const trust_params &p, const LaunchPlan &L, TcCtl *ctl,
                                 hipStream_t s, const CsrViews &v, int64_t V, int64_t v0);
// Instrumented wave-only pass, GPU features, GPU training of the MLP.
int train_cost_model(Ctx *ctx, const trust_params &p, int epochs, long long max_samples,
                     uint64_t seed, double *loss_out, double *seconds_out);

}  // namespace trust
