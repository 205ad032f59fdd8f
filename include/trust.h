/*
 * trust.h -- C ABI of the MI355X (gfx950) TRUST triangle-counting library.
 *
 * Plain C types only: pointers, sizes and fixed-width integers.  Every entry
 * point returns 0 (TRUST_OK) or a negative TRUST_ERR_* code and never exits
 * the process; trust_last_error() gives the message of the calling thread's
 * last failure.  Citations name the reference interface each entry point
 * replaces (paths relative to the wzbxpy/TRUST repository root).
 */
#ifndef TRUST_H
#define TRUST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define TRUST_OK 0
#define TRUST_ERR_ARG (-1)      /* invalid argument */
#define TRUST_ERR_IO (-2)       /* missing / unreadable / unwritable file */
#define TRUST_ERR_FORMAT (-3)   /* malformed CSR (sizes, offsets, ids) */
#define TRUST_ERR_OOM (-4)      /* host or device allocation failed */
#define TRUST_ERR_HIP (-5)      /* HIP runtime error */
#define TRUST_ERR_OVERFLOW (-6) /* a size exceeds what the representation holds */
#define TRUST_ERR_NODEV (-7)    /* no usable GPU */

/* ---- schedules and flags ---------------------------------------------- */
#define TRUST_SCHED_DEGREE 0  /* plain degree-threshold split (reference USE_CTA) */
#define TRUST_SCHED_LEARNED 1 /* MLP cost model: tiers + longest-first queue */
#define TRUST_SCHED_AUTO 2    /* LEARNED unless the degree split counted this graph faster
                                 when the model was trained on it (the split as judge) */

#define TRUST_FLAG_FORCE_SPILL 1u  /* every hashed vertex uses an HBM table */
#define TRUST_FLAG_OFFSETS64 2u    /* keep int64 offsets on the device even if E < 2^30 */
#define TRUST_FLAG_NO_VALIDATE 4u  /* skip the host-side CSR validation */
#define TRUST_FLAG_WAVE_ONLY 8u    /* no workgroup tier: every vertex that fits a wave table uses one wave */
#define TRUST_FLAG_LINEAR_PROBING 16u /* linear-probing tables instead of two-choice cuckoo buckets */
#define TRUST_FLAG_HIP_GRAPH 64u   /* batch counts: replay plan + count as one captured hipGraph */
/* Walks.  The reference walks every v in N+(u) from u's table
 * (Without-graph-partition/TC.cu:330-359,510-544).  The short-side walk
 * intersects each oriented edge u -> v (d+(u) >= 2) from the shorter side:
 * N+(v) through u's table when d+(v) <= d+(u), else N+(u) through v's table.
 * Same triangles, min(d+(u), d+(v)) probes per edge.  It needs the in-walk
 * rows (for every v, the u -> v with 2 <= d+(u) < d+(v)), a per-graph index
 * that a context builds once, the first time a count uses it; that call
 * reports the build in trust_result.index_seconds and includes it in
 * seconds.
 * Without either flag an unpartitioned count takes the short-side walk when
 * its probes are at most 45% of the reference walk's (a pass over the graph,
 * also once per graph and reported the same way); partitioned counts always
 * take the reference walk.  Both flags together are rejected. */
#define TRUST_FLAG_REFERENCE_WALK 128u
#define TRUST_FLAG_SHORT_WALK 256u
/* The reference walk of a resident single graph with int32 offsets reads
 * each walked v's list start and length from a per-graph start index
 * (start[e] = beg[adj[e]], len[e] = d+(adj[e]), read in order with the walk)
 * instead of looking up beg[v] and beg[v + 1] at random; built once, the
 * first time a count uses it, and reported in trust_result.index_seconds
 * (included in seconds).  Batches and int64 offsets keep the lookups.  This
 * flag keeps them too. */
#define TRUST_FLAG_OFFSET_LOOKUP 512u
/* Discard the resident graph's per-graph indexes (walk statistics, the
 * short-side walk rows, the start index) before counting, so that this
 * count makes the ones it uses again, inside its seconds: the cost of
 * counting a freshly loaded graph, as the reference's timed region covers
 * all device work of its count (Without-graph-partition/TC.cu:686-692),
 * without reloading the graph. */
#define TRUST_FLAG_FRESH_INDEX 1024u
/* Test aid: every hash table is built as a cuckoo table first and then
 * treated as if an insertion had failed, i.e. cleared and rebuilt with
 * linear probing -- the path a real failure takes (counted in
 * trust_result.linear_rebuilds).  Counts are unchanged. */
#define TRUST_FLAG_REBUILD_LINEAR 2048u
/* Any other bit in trust_params.flags is rejected with TRUST_ERR_ARG. */

/*
 * Launch and schedule parameters.  threads / blocks / chunk / rank /
 * total_process carry the meaning of the reference's argv[3], argv[4],
 * argv[5], the MPI rank and argv[2] (Without-graph-partition/mpi_main.cpp:78-81).
 * Zero means "choose automatically" for threads, blocks, chunk, wave_slots
 * and lds_budget.  Fill with trust_params_init() before changing fields.
 */
typedef struct trust_params {
    int32_t threads;       /* work-items per workgroup: multiple of 64, <= 1024 */
    int32_t blocks;        /* persistent workgroups in the launch */
    int32_t chunk;         /* vertices taken per queue grab */
    int32_t rank;          /* logical rank of this call, 0 <= rank < total_process */
    int32_t total_process; /* number of logical ranks the work is split over */
    int32_t schedule;      /* TRUST_SCHED_* */
    int32_t hub_threshold; /* walk degree above which a whole workgroup takes a vertex (reference: 100) */
    int32_t wave_slots;    /* per-wave LDS hash-table slots (power of two) */
    int32_t lds_budget;    /* bytes of LDS per workgroup for the hub table */
    uint32_t flags;        /* TRUST_FLAG_* */
    int32_t device;        /* HIP device index */
    /* (after reserved) Dense core of the short-side walk (no counterpart in the reference,
     * whose counts it leaves unchanged): the edges u -> v with d+(u) and
     * d+(v) >= core_min_degree into a core vertex v whose out-list lies in
     * the core are counted as popcount(B(u) & B(v)) over bit rows of the
     * core, and the walk leaves them out.  The core is the core_vertices
     * highest estimated degrees, closed under out-neighbours while it stays
     * within 32768 vertices.  A per-graph index like the walk rows, made by
     * the first count that uses it.  core_vertices: 0 automatic (16384 for
     * single graphs of at least 2^23 edges counted by one logical rank, else
     * none), -1 none, else a power of two in [256, 16384].
     * core_min_degree: 0 automatic (384), else >= 2. */
    int32_t reserved[3];   /* reserved[0]: a tile stream's device tile-cache slots (0: 2n+2) */
    int32_t core_vertices;
    int32_t core_min_degree;
} trust_params;

/* Result of one counting call. */
typedef struct trust_result {
    int64_t vertices;      /* V (rows of the CSR) */
    int64_t edges;         /* E (entries of adjacent.bin = undirected edges) */
    uint64_t triangles;    /* exact count (this rank's share) */
    uint64_t probes;       /* W: hash lookups performed by this call */
    double seconds;        /* device time of this call's work (hipEvents): per-graph index
                              work it did (index_seconds) + planning + counting */
    double count_seconds;  /* device time of the counting launch alone */
    double plan_seconds;   /* device time of the planning kernels */
    double h2d_seconds;    /* host-to-device copy time (not part of seconds) */
    double gteps;          /* edges / seconds / 1e9 */
    double alg_bytes;      /* algorithmic HBM bytes of the counting launch */
    int64_t n_spill;       /* vertices counted with an HBM hash table */
    int64_t n_hub;         /* vertices counted by a whole workgroup (LDS table) */
    int64_t n_wave;        /* vertices counted by one wavefront (LDS table) */
    int32_t threads;       /* launch shape actually used */
    int32_t blocks;
    int32_t hub_slots;     /* LDS hub-table slots per workgroup */
    int32_t wave_slots;    /* LDS table slots per wavefront */
    int32_t schedule_used; /* TRUST_SCHED_* that ran (LEARNED falls back to DEGREE without a model) */
    int32_t walk_used;     /* 0: the reference walk, 1: the short-side walk */
    double index_seconds;  /* device time of the per-graph index work this call did (walk
                              statistics, the in-walk rows or the start index; 0 when a
                              context reuses them), included in seconds */
    uint64_t linear_rebuilds; /* hash tables rebuilt with linear probing after a cuckoo
                                 insertion failed (or was forced to, TRUST_FLAG_REBUILD_LINEAR) */
    int64_t core_vertices;    /* vertices of the dense core this count used (0: none) */
    uint64_t dense_edges;     /* oriented edges counted by the core's bit rows */
    uint64_t dense_probes;    /* the short-side walk's probes of those edges (not in probes) */
    double dense_seconds;     /* device time of the core's launch (part of count_seconds) */
} trust_result;

/* Library-allocated CSR (trust_csr_free releases it). */
typedef struct trust_csr {
    int64_t vertices;
    int64_t edges;
    int64_t *beg; /* vertices+1 offsets (begin.bin layout) */
    int32_t *adj; /* edges ids (adjacent.bin layout) */
} trust_csr;

/* Drop-in result of Triangle_count (Without-graph-partition/TC.cuh:2-8), widened. */
typedef struct trust_arguments {
    int64_t edge_count; /* reference: int */
    uint64_t count;     /* reference: long long */
    double time;        /* device seconds of the count: index work, planning and counting */
    int64_t degree;     /* edge_count / vertices */
    int64_t vertices;   /* reference: int */
} trust_arguments;

typedef struct trust_ctx trust_ctx;

/* ---- library ---------------------------------------------------------- */
/* Message of this thread's last failed call.  The reference prints and exits
 * on any CUDA error instead (HRR, Without-graph-partition/herror.h:1-8). */
const char *trust_last_error(void);
const char *trust_version(void);
/* Caps the device memory all of the library's buffers may hold together (all
 * contexts and devices of the process; 0 = no cap, the default).  An
 * allocation beyond the cap fails with TRUST_ERR_OOM exactly as a failed
 * hipMalloc does, so the out-of-memory paths (the walk index's fallback to
 * its low-scratch build, clean errors, a context that stays usable) can be
 * exercised, and a caller can bound the library's footprint.  No counterpart
 * in the reference, which exits on any allocation failure
 * (Without-graph-partition/herror.h:1-8). */
int trust_set_device_memory_cap(uint64_t bytes);
/* Device bytes the library's buffers hold now (graphs, indexes, scratch). */
uint64_t trust_device_memory_held(void);
/* The launch fields stand for the reference's argv[3..5] and the MPI rank /
 * argv[2] (Without-graph-partition/mpi_main.cpp:78-81).
 * Defaults: threads 0, blocks 0, chunk 0, rank 0, total_process 1,
 * schedule DEGREE, hub_threshold 100, wave_slots 0, lds_budget 0.  With
 * schedule LEARNED or AUTO and no cost model loaded, the degree split runs;
 * trust_result.schedule_used says which schedule ran. */
void trust_params_init(trust_params *p);
/* Number of HIP devices (0 when none); negative on runtime failure.  The
 * reference picks device (rank + 1) % count itself (TC.cu:615-621); here the
 * device is trust_params.device. */
int trust_device_count(void);

/* ---- drop-in counting entry points ------------------------------------ */
/*
 * Replaces Triangle_count(rank, folder, args, total_process, threads, blocks,
 * chunk_size) of Without-graph-partition/TC.cuh:10 (TC.cu:602-731): loads
 * <folder>/begin.bin and <folder>/adjacent.bin, counts this rank's share of
 * vertex chunks (chunk c is rank c % total_process, TC.cu:259,433) and fills
 * *out.  threads/blocks/chunk_size <= 0 select automatic values.
 */
int trust_triangle_count(int rank, const char *folder, int total_process, int threads,
                         int blocks, int chunk_size, trust_arguments *out);

/*
 * Replaces Triangle_count(..., partition_num) of With-graph-partition/TC.cuh:10
 * (TC.cu:976-1157): rank selects the tile job (i,j,k) = (rank%n, rank/n%n,
 * rank/n/n%n) over <folder>/partition{i}_{j}/ and rank/n^3 of total_process/n^3
 * selects the chunk share inside that job (TC.cu:985-994,1024-1025).
 */
int trust_triangle_count_partition(int rank, const char *folder, int total_process,
                                   int threads, int blocks, int chunk_size,
                                   int partition_num, trust_arguments *out);

/* Triangle_count over a CSR already in host memory instead of a folder
 * (Without-graph-partition/TC.cu:602-731 without its file load); copied to the
 * device inside the call. */
int trust_count_csr(const int64_t *beg, const int32_t *adj, int64_t vertices, int64_t edges,
                    const trust_params *p, trust_result *r);

/* No counterpart in the reference.  trust_count_csr for a CSR the caller
 * already holds in DEVICE memory (beg_dev: vertices+1 int64 offsets in the
 * begin.bin layout, adj_dev: edges int32 ids), on the caller's HIP stream
 * (null: the default stream), that allocates nothing: no hipMalloc, hipFree
 * or allocating library call.  For callers that cannot allocate (code under
 * stream capture, serving loops, harnesses that forbid it).  Its scratch is
 * one fixed device array of the library (about 34 MB per GPU, held until
 * the process ends).  Limits: vertices <= 2^22, edges < 2^30, out-degrees
 * whose tables fit the LDS (about 16k at the default launch), the
 * reference walk with offset lookups, the degree split; FORCE_SPILL,
 * OFFSETS64, HIP_GRAPH, SHORT_WALK, FRESH_INDEX and the learned schedule are
 * rejected.  The CSR is checked on the device (offsets, ids) unless
 * TRUST_FLAG_NO_VALIDATE.  Calls on one GPU share its scratch: calls that
 * overlap on one GPU race.  seconds = device wall clock of plan + count. */
int trust_count_csr_noalloc(const int64_t *beg_dev, const int32_t *adj_dev, int64_t vertices,
                            int64_t edges, const trust_params *p, trust_result *r,
                            void *stream);

/*
 * Sum over all n^3 (i,j,k) jobs of a 2D hash partition held in host memory,
 * streamed through one GPU with asynchronous tile copies.  tiles[i*n+j] is
 * tile (i,j) in the layout of Preprocess/partition.cpp:72-121.
 */
int trust_count_tiles(const trust_csr *tiles, int n, const trust_params *p, trust_result *r);
/* trust_count_tiles with a job journal, so that an interrupted stream resumes
 * where it stopped.  The journal (a text file, created when missing) names
 * the partition on its first line and gets one line per finished job,
 * "job <id> <i> <j> <k> <triangles> <probes>", appended and flushed as each
 * job's count reaches the host; jobs already in it are not counted again (the
 * jobs' counts are additive).  A last line cut short is dropped; a journal of
 * another partition is TRUST_ERR_FORMAT.  max_new_jobs > 0 stops after that
 * many new jobs.  Returns the number of jobs still missing from the journal
 * (0: r holds the whole count) or a negative error code; r sums every job in
 * the journal, its times cover this call's work; job_counts (nullable, n^3)
 * receives each job's count (0 for jobs still missing).  No counterpart in
 * the reference, whose ranks each count one job and keep nothing
 * (With-graph-partition/mpi_main.cpp). */
int trust_count_tiles_journal(const trust_csr *tiles, int n, const trust_params *p,
                              const char *journal, int64_t max_new_jobs, uint64_t *job_counts,
                              trust_result *r);

/* ---- resident graphs -------------------------------------------------- */
/* No counterpart in the reference, whose every Triangle_count call loads,
 * copies and frees the graph (Without-graph-partition/TC.cu:611-711): a
 * context keeps one resident on the device for repeated counts, and keeps
 * its per-graph walk indexes. */
int trust_ctx_create(int device, trust_ctx **out);
void trust_ctx_destroy(trust_ctx *ctx);
/* Copy a CSR to the device once; later counts reuse it.  h2d_seconds nullable. */
int trust_ctx_load_csr(trust_ctx *ctx, const int64_t *beg, const int32_t *adj,
                       int64_t vertices, int64_t edges, uint32_t flags, double *h2d_seconds);
int trust_ctx_count(trust_ctx *ctx, const trust_params *p, trust_result *r);
/*
 * Instrumented count: per-vertex cycle counts (clock64 around each vertex's
 * work) and probe counts.  cycles/probes have `vertices` entries (nullable).
 */
int trust_ctx_profile_vertices(trust_ctx *ctx, const trust_params *p, uint64_t *cycles,
                               uint64_t *probes, trust_result *r);

/* ---- learned cost scheduler (no counterpart in the reference) ---------- */
/*
 * Cost model: an MLP 3 -> H -> H -> 1 (ReLU), H = 32, fp32, predicting
 * log2(1 + single-wave cycles)/32 of a vertex from x = log2(1 + f)/32 of the
 * features f = (d+(u), sum_{v in N+(u)} d+(v), max_{v in N+(u)} d+(v)).
 * Parameter layout: W1[H][3], b1[H], W2[H][H], b2[H], W3[H], b3  (H^2+6H+1).
 * Forward, backward and Adam run as HIP kernels on the device.
 */
int64_t trust_mlp_param_count(int hidden);
int trust_mlp_init(int hidden, uint64_t seed, float *params);
/* y[n] = MLP(x[n][3]) */
int trust_mlp_forward(int device, int hidden, const float *params, const float *x, int64_t n,
                      float *y);
/* grad[P] of mean((y - target)^2); *loss (nullable) = that mean. */
int trust_mlp_loss_grad(int device, int hidden, const float *params, const float *x,
                        const float *target, int64_t n, float *grad, double *loss);
/* Minibatch Adam (batch, lr), seeded shuffling; params updated in place. */
int trust_mlp_train(int device, int hidden, float *params, const float *x, const float *target,
                    int64_t n, int epochs, int batch, float lr, uint64_t seed, double *loss_out);
/* Scaled features of every vertex of the resident graph, feat[V][3]. */
int trust_ctx_features(trust_ctx *ctx, float *feat);
/* The same features over the short-side walk's rows (d+(u), sum and max of
 * the list lengths u's walk probes): what the learned plan uses when a count
 * takes that walk.  Builds the walk index first if the context has none. */
int trust_ctx_walk_features(trust_ctx *ctx, float *feat);
/*
 * Trains the context's cost model: an instrumented wave-only count records
 * per-vertex cycles, then the MLP is trained on a seeded sample of at most
 * max_samples vertices.  *loss: final MSE; *seconds: device time spent.
 * Counts with schedule TRUST_SCHED_LEARNED then use the model.
 */
int trust_ctx_train_cost_model(trust_ctx *ctx, const trust_params *p, int epochs,
                               int64_t max_samples, uint64_t seed, double *loss, double *seconds);
int trust_ctx_set_cost_model(trust_ctx *ctx, int hidden, const float *params);
/*
 * The degree split's verdict on the trained model for the resident graph:
 * trust_ctx_train_cost_model ends with two learned and two degree-split
 * counts of the graph at the caller's launch parameters (their device time is
 * included in its *seconds).  *decision = TRUST_SCHED_LEARNED when the
 * learned count's best of two was more than 1% faster, else
 * TRUST_SCHED_DEGREE; -1 when no judgement exists (no training on this graph,
 * or a model set by trust_ctx_set_cost_model).  TRUST_SCHED_AUTO follows it.
 * Times nullable.
 */
int trust_ctx_schedule_judgement(trust_ctx *ctx, int *decision, double *learned_seconds,
                                 double *degree_seconds);
int trust_ctx_get_cost_model(trust_ctx *ctx, int *hidden, float *params);
/* Predicted single-wave cycles per vertex (0 for vertices that are not counted). */
int trust_ctx_predict_costs(trust_ctx *ctx, float *cost);

/* ---- measurement helpers (no counterpart in the reference) ------------ */
/*
 * Gather-only pass: the same plan, tiers and probe walk as trust_ctx_count,
 * with every hash lookup replaced by a checksum of the loaded ids.  Its
 * count_seconds is the time of the load path alone, the roofline
 * denominator for random segment access on this graph (r->triangles holds
 * the checksum, not a count).
 */
int trust_ctx_gather(trust_ctx *ctx, const trust_params *p, trust_result *r);
/* Device-to-device streaming copy of `bytes`, `reps` times; seconds per copy
 * (the known-traffic kernel that calibrates HBM byte counters). */
int trust_membench_copy(int device, int64_t bytes, int reps, double *seconds_per_copy);
/* Read rate (bytes per second) of 8 workgroups per CU each reading a buffer
 * of `bytes` `reps` times: a buffer that fits the L2 (<= 4 MB) measures the
 * L2s' read rate, the reference for the dense core's bit-row reads. */
int trust_membench_read(int device, int64_t bytes, int reps, double *bytes_per_second);

/* ---- launch autotuner (no counterpart in the reference) --------------- */
/*
 * Replaces choosing argv[3..5] (threads, blocks, chunk) of
 * Without-graph-partition/mpi_main.cpp:78-81 by hand and the compile-time
 * USE_CTA threshold (TC.cu:14): times counts of the resident graph with
 * `base` as the starting point, by coordinate descent over threads
 * {128..1024}, wave_slots {128..2048}, hub_threshold {16..2000} and chunk
 * {auto, 1..64}, keeping a change only when it is more than 1% faster.  Each
 * candidate is the minimum of `reps` device times; every count must equal
 * the first (else TRUST_ERR_HIP).  max_seconds > 0 bounds the search.
 * *best receives the fastest parameters (other fields as in base).
 */
int trust_ctx_autotune(trust_ctx *ctx, const trust_params *base, int reps, double max_seconds,
                       trust_params *best, double *best_seconds, int *tried);

/* ---- batches of graphs (no counterpart in the reference) -------------- */
/*
 * The reference counts one graph per process (Without-graph-partition/
 * mpi_main.cpp:72-96).  A batch is many graphs concatenated into one CSR
 * with global ids: graph g owns vertices [vbase[g], vbase[g+1]) and every
 * edge stays inside its graph.  One persistent count covers the whole batch
 * and attributes each vertex's triangles to its graph.
 *
 * Shapes of a seeded mixed batch: graph g has n[g] vertices (log-uniform in
 * [nmin, nmax]), mean degree d[g] (uniform in [dmin, dmax]) and kind[g] = g & 1
 * (0: Erdos-Renyi G(n, m = n*d/2), 1: Barabasi-Albert with k = round(d/2)).
 */
int trust_batch_shape(int64_t ngraphs, uint64_t seed, int64_t nmin, int64_t nmax, double dmin,
                      double dmax, int64_t *n, double *d, int32_t *kind);
/* Generates that batch, each graph through trust_preprocess_edges, and
 * concatenates it into *out; vbase has ngraphs + 1 entries. */
int trust_batch_generate(int64_t ngraphs, uint64_t seed, int64_t nmin, int64_t nmax, double dmin,
                         double dmax, trust_csr *out, int64_t *vbase);
/* Copies a batch CSR and its vbase[ngraphs + 1] to the device (flags as for
 * trust_ctx_load_csr; the graph is also the context's resident graph). */
int trust_ctx_load_batch(trust_ctx *ctx, const int64_t *beg, const int32_t *adj, int64_t vertices,
                         int64_t edges, const int64_t *vbase, int64_t ngraphs, uint32_t flags);
/*
 * Counts every graph of the resident batch in one pass: counts[g] (uint64)
 * and done_seconds[g] = time from the first workgroup's start to the end of
 * graph g's last vertex (0 for a graph with nothing to count); both nullable.
 * With TRUST_FLAG_HIP_GRAPH in p->flags the planning and counting launches
 * are captured once into a hipGraph (re-captured when the parameters or
 * buffers change) and each call replays it.
 */
int trust_ctx_count_batch(trust_ctx *ctx, const trust_params *p, uint64_t *counts,
                          double *done_seconds, trust_result *r);
/*
 * No counterpart in the reference (which counts one graph per process).
 * A batch with E >= 2^30 needs int64 offsets as a whole, but no edge leaves
 * its graph, so trust_ctx_load_batch cuts it into segments of whole graphs
 * with fewer than 2^30 edges each, whose offsets are int32 relative to the
 * segment's first edge, and counts run the int32 kernel per segment (the
 * short-side walk and the learned plan keep the int64 kernel).  This call
 * re-cuts the loaded batch with at most max_edges edges per segment (capped
 * at 2^30 - 1; 0: no segments).  Segments need int64 offsets on the device
 * (E >= 2^30, or TRUST_FLAG_OFFSETS64 at load, which alone does not cut) and
 * no graph above the cap; otherwise the batch stays whole.  *segments
 * (nullable) receives the segment count, 0 when the batch is whole.
 */
int trust_ctx_batch_segments(trust_ctx *ctx, int64_t max_edges, int64_t *segments);

/* ---- input contract: files, preprocessing, tiles ---------------------- */
/* Read <folder>/begin.bin + adjacent.bin (Without-graph-partition/graph.hpp:18-67)
 * with size and content validation. */
int trust_csr_load(const char *folder, trust_csr *out);
/* Write <folder>/begin.bin (int64 offsets) + adjacent.bin (int32 ids), the
 * files Preprocess/preprocess_outdegree.cpp:205-222 writes. */
int trust_csr_save(const char *folder, const trust_csr *csr);
/* Check offsets (monotone, beg[0]=0, beg[V]=E) and ids (< V).  No
 * counterpart: the reference reads the files unchecked (graph.hpp:18-67). */
int trust_csr_validate(const int64_t *beg, const int32_t *adj, int64_t vertices, int64_t edges);
void trust_csr_free(trust_csr *csr);

/*
 * Edge list -> degree-oriented CSR, byte-identical to running
 * Preprocess/fromDirectToUndirect.cpp then Preprocess/preprocess_outdegree.cpp
 * on the same list: ids compacted in ascending order, undirected,
 * deduplicated, oriented from lower to higher position of an ascending
 * degree sort, rows ordered by out-degree classes (>100, 2..100, <2) with a
 * first-touch relabel inside each class.  src/dst hold m pairs of
 * non-negative ids.
 */
int trust_preprocess_edges(const int32_t *src, const int32_t *dst, int64_t m, trust_csr *out);

/* Row orders of the reference's preprocessing tools.  All three share the
 * ascending-degree sort and the orientation; only the final row order and ids
 * differ, so every order counts the same triangles.
 *   OUTDEGREE  preprocess_outdegree.cpp (and speedupIO/preprocess_out.cpp):
 *              descending out-degree, first-touch relabel per class (:100-185)
 *   PLAIN      otherPreprocess/preprocess.cpp:97-135,178: descending
 *              out-degree, ids = positions, no relabel
 *   INDEGREE   otherPreprocess/preprocess_indegree.cpp:131-156: the
 *              ascending-degree order split stably into the classes
 *              out-degree > 100, 2..100, < 2; ids = positions */
#define TRUST_ORDER_OUTDEGREE 0
#define TRUST_ORDER_PLAIN 1
#define TRUST_ORDER_INDEGREE 2
/* speedupIO/preprocess_out.cpp:101-126 deduplicates each adjacency list, so
 * a self-loop adds 1 to its vertex's degree; the 1.mmio path adds 2
 * (preprocess_outdegree.cpp:71-72).  The degree decides the sort, so the
 * bytes differ when self-loops are present. */
#define TRUST_PRE_SELF_LOOP_ONCE 1u
/* trust_preprocess_edges with a row order and flags (above); any other flag
 * bit is rejected. */
int trust_preprocess_edges_ex(const int32_t *src, const int32_t *dst, int64_t m, int order,
                              uint32_t flags, trust_csr *out);

/* Tile (i,j) of an n x n hash partition (Preprocess/partition.cpp:72-121). */
int trust_partition_csr(const trust_csr *csr, int n, trust_csr *tiles /* n*n */);
/* Write partition{i}_{j}/{begin.bin,adjacent.bin,map,division} under folder
 * (Preprocess/partition.cpp:101-165). */
int trust_partition_write(const char *folder, const trust_csr *csr, int n);
/* Same tile folders, written from <folder>/begin.bin + adjacent.bin mapped
 * read-only from disk, one tile row (n tiles) in host memory at a time:
 * the reference (partition.cpp:72-121) holds the whole graph and every tile
 * in memory and stops at n <= 10 (:74-77); this takes n <= 1024. */
int trust_partition_folder(const char *folder, int n);
/* The R-MAT stream of trust_gen_rmat, preprocessed out of core straight from
 * the generator into out_folder/begin.bin and adjacent.bin: byte-identical to
 * trust_preprocess_edges on trust_gen_rmat's pairs, with the pairs, the
 * oriented rows and the final adjacency passing through bucket files of at
 * most bucket_pairs 8-byte keys under work_dir (removed at the end).  Host
 * memory: arrays over the vertices (about 60 bytes each) plus two buckets.
 * Followed by trust_partition_folder, it writes the 2D tiles of a graph that
 * never sits in host memory whole; the reference holds the whole edge list
 * and CSR (Preprocess/fromDirectToUndirect.cpp, preprocess_outdegree.cpp,
 * partition.cpp:72-121). */
int trust_rmat_csr_ooc(int scale, int64_t m, double a, double b, double c, uint64_t seed,
                       const char *out_folder, const char *work_dir, int64_t bucket_pairs);

/* ---- seeded synthetic generators (edge lists, not yet preprocessed) --- */
/* No counterpart in the reference, which downloads its datasets
 * (Dataset/Cit-Patents/get+preprocess.sh); the output goes through the same
 * preprocessing as a downloaded edge list. */
/* R-MAT: m = edge_factor << scale samples into src/dst (caller-allocated). */
int trust_gen_rmat(int scale, int64_t m, double a, double b, double c, uint64_t seed,
                   int32_t *src, int32_t *dst);
/* Chung-Lu: m samples, weights (i+i0)^(-1/(gamma-1)) scaled so the largest
 * expected degree is max_degree. */
int trust_gen_chung_lu(int64_t n, int64_t m, double gamma, double max_degree, uint64_t seed,
                       int32_t *src, int32_t *dst);
/* Erdos-Renyi G(n, m): m uniform pairs. */
int trust_gen_erdos_renyi(int64_t n, int64_t m, uint64_t seed, int32_t *src, int32_t *dst);
/* Barabasi-Albert (preferential attachment by sampling earlier pair ends):
 * n vertices, k pairs per vertex, n*k pairs in all (caller allocates k*n);
 * self-loops and repeated pairs occur and preprocessing removes them.
 * *m_out receives the pair count. */
int trust_gen_barabasi_albert(int64_t n, int k, uint64_t seed, int32_t *src, int32_t *dst,
                              int64_t *m_out);

#ifdef __cplusplus
}
#endif

#endif /* TRUST_H */
