// Learned cost scheduler.
//
// A small MLP (mlp.hip) predicts, from per-vertex degree features, the
// cycles one wavefront spends on a vertex; it is trained on the GPU on
// cycles that the instrumented counting kernel records itself.  At count
// time (inside the timed region) the scheduler
//   1. computes the features on the GPU: log2(1+d+(u)), log2(1+sum_v d+(v)),
//      log2(1+max_v d+(v)) over v in N+(u)  (the second is the probe count),
//   2. evaluates the MLP per vertex,
//   3. chooses the tier: tables too large for a wave table go to the
//      workgroup / large-LDS / HBM tiers exactly as the degree split does
//      (the LDS-vs-HBM choice is by table size); a vertex that fits a wave
//      table goes to the workgroup tier only when its predicted cost
//      exceeds half of one wave's fair share of the total predicted work,
//   4. orders every tier longest-first: 64 logarithmic cost bins (factor
//      sqrt 2), a stable counting sort so vertices keep their id order
//      (and the locality of the reference relabel) inside a bin.
// Without a trained model the degree-threshold split is used (the
// reference's plain split, Without-graph-partition/TC.cu:19,632) and the
// result reports which schedule ran.
#include <hip/hip_runtime.h>

#include <cstddef>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.h"
#include "mlp.h"
#include "wave_ops.h"

namespace trust {

constexpr int kBins = 64;
constexpr int kKeys = kNumTiers * kBins;  // 256
constexpr unsigned short kKeyNone = 0xFFFF;

__device__ __forceinline__ int wmax_i(int x)
{
    for (int off = 32; off > 0; off >>= 1)
        x = max(x, __shfl_xor(x, off));
    return x;
}

// d+(u), sum and max of d+(v) over v in N+(u), log-scaled.  A wave takes 64
// consecutive vertices; their lists are contiguous in adj, so the wave
// walks the concatenation in coalesced 64-wide windows, kFeatWin windows
// at a time (their loads in flight together).  The vertex owning each
// position comes from a mark array of list starts and a running max; each
// lane adds d+(v) of its position into the owner's LDS slot.
constexpr int kFeatWin = 4;

template <typename OffT>
__global__ void __launch_bounds__(256) features_kernel(const OffT *__restrict__ beg,
                                                       const int32_t *__restrict__ adj,
                                                       const OffT *__restrict__ begD, long long V,
                                                       float *__restrict__ feat)
{
    __shared__ int mark_s[4][64 * kFeatWin];
    __shared__ unsigned long long sum_s[4][64];
    __shared__ int max_s[4][64];
    const int wave = threadIdx.x / 64, lane = lane_id();
    int *mark = mark_s[wave];
    unsigned long long *lsum = sum_s[wave];
    int *lmax = max_s[wave];
    const long long waves = (long long)gridDim.x * (blockDim.x / 64);
    for (long long u0 = ((long long)blockIdx.x * (blockDim.x / 64) + wave) * 64; u0 < V;
         u0 += waves * 64) {
        const int n = (int)min(64ll, V - u0);
        // lane i < n holds vertex u0+i; lanes >= n hold the end of the range
        const OffT s = beg[u0 + min(lane, n)];
        const OffT d = lane < n ? beg[u0 + lane + 1] - s : 0;
        const long long e0 = (long long)__shfl(s, 0);
        const long long e1 = (long long)__shfl(s, n == 64 ? 63 : n) + (n == 64 ? (long long)__shfl(d, 63) : 0);
        lsum[lane] = 0;
        lmax[lane] = 0;
        int carry = 0;
        for (long long p = e0; p < e1; p += 64 * kFeatWin) {
#pragma unroll
            for (int k = 0; k < kFeatWin; k++)
                mark[k * 64 + lane] = 0;
            wave_sync();
            // list starts inside this span (an empty list starts where the
            // next one does, so only non-empty lists mark)
            if (d > 0 && (long long)s >= p && (long long)s < p + 64 * kFeatWin)
                mark[(int)((long long)s - p)] = lane;
            wave_sync();
            int own[kFeatWin], v[kFeatWin];
#pragma unroll
            for (int k = 0; k < kFeatWin; k++) {
                own[k] = max(wave_scan_max(mark[k * 64 + lane]), carry);
                carry = __shfl(own[k], 63);
                const long long pos = p + k * 64 + lane;
                v[k] = pos < e1 ? adj[pos] : -1;
            }
            int dv[kFeatWin];
#pragma unroll
            for (int k = 0; k < kFeatWin; k++)
                dv[k] = v[k] >= 0 ? (int)(begD[v[k] + 1] - begD[v[k]]) : 0;
#pragma unroll
            for (int k = 0; k < kFeatWin; k++) {
                if (v[k] >= 0) {
                    atomicAdd(&lsum[own[k]], (unsigned long long)dv[k]);
                    atomicMax(&lmax[own[k]], dv[k]);
                }
            }
            wave_sync();
        }
        wave_sync();
        if (lane < n) {
            const long long u = u0 + lane;
            const float dd = (float)d;  // table size d+(u)
            feat[3 * u] = log2f(1.f + dd) * kFeatScale;
            feat[3 * u + 1] = log2f(1.f + (float)lsum[lane]) * kFeatScale;
            feat[3 * u + 2] = log2f(1.f + (float)lmax[lane]) * kFeatScale;
        }
        wave_sync();
    }
}

// The same features over the short-side walk's rows, from the per-entry row
// and list length the walk index stores: a flat pass, 64 consecutive entries
// per wave step, so long rows (a hub's in-walk) spread over many waves.  Rows
// are contiguous and in order, so each wave reduces its row segments (an
// add scan and a max scan keyed by row) and adds one atomic per segment into
// wsum / wmax (cleared before).
__global__ void __launch_bounds__(256) walk_feat_kernel(const int32_t *__restrict__ rowX,
                                                        const int32_t *__restrict__ degX,
                                                        int dstride, long long X,
                                                        unsigned long long *__restrict__ wsum,
                                                        unsigned int *__restrict__ wmax)
{
    const int lane = lane_id();
    const long long nw = (long long)gridDim.x * (blockDim.x / 64);
    for (long long w = (long long)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; w * 64 < X;
         w += nw) {
        const long long pos = w * 64 + lane;
        const bool valid = pos < X;
        const int r = valid ? rowX[pos] : -1;
        const int dv = valid ? degX[pos * dstride] : 0;
        const int rprev = __shfl_up(r, 1);
        const int rnext = __shfl_down(r, 1);
        const bool head = valid && (lane == 0 || rprev != r);
        const bool last = valid && (lane == 63 || rnext != r);
        int inc = dv;  // inclusive add scan (a window's sum fits an int)
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(inc, off);
            if (lane >= off)
                inc += y;
        }
        // rows do not decrease along the entries, so a max scan of
        // (row, d+) keys is a max per row segment
        unsigned long long key =
            valid ? ((unsigned long long)(unsigned int)r << 32) | (unsigned int)dv : 0ull;
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned long long y = __shfl_up(key, off);
            if (lane >= off && y > key)
                key = y;
        }
        const int hl = wave_scan_max(head ? lane : 0);  // this lane's segment head
        const int before = __shfl(inc, hl > 0 ? hl - 1 : 0);
        const int seg = hl > 0 ? inc - before : inc;
        if (last) {
            atomicAdd(wsum + r, (unsigned long long)seg);
            atomicMax(wmax + r, (unsigned int)(key & 0xffffffffu));
        }
    }
}

template <typename OffT>
__global__ void __launch_bounds__(256) walk_feat_final_kernel(const OffT *__restrict__ begD,
                                                              long long V,
                                                              const unsigned long long *__restrict__ wsum,
                                                              const unsigned int *__restrict__ wmax,
                                                              float *__restrict__ feat)
{
    for (long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x; u < V;
         u += (long long)gridDim.x * blockDim.x) {
        feat[3 * u] = log2f(1.f + (float)(begD[u + 1] - begD[u])) * kFeatScale;
        feat[3 * u + 1] = log2f(1.f + (float)wsum[u]) * kFeatScale;
        feat[3 * u + 2] = log2f(1.f + (float)wmax[u]) * kFeatScale;
    }
}

// Predicted single-wave cycles per vertex (0 for vertices that are not
// counted); per-block partial sums of the predictions.
//
// The 32 x 32 hidden layer runs on the matrix cores.  A wave takes 32
// vertices at a time and computes H2^T = W2 x H1^T (outputs x vertices) with
// four v_mfma_f32_32x32x8_f16, f16 operands and f32 accumulation; the first
// layer (3 inputs) and the last (one output) are vector code.  Operand
// layout of that instruction, lane L, q = 0..3, r = 0..15:
//   A[row L%32][k 4(L/32)+q],  B[k 4(L/32)+q][col L%32],
//   D[row 8(r/4)+4(L/32)+r%4][col L%32]
// so a column is a vertex: lanes L and L+32 share vertex L%32, each builds
// half of its hidden units as its B operand and sums half of the output
// rows, and one exchange adds the halves.
typedef _Float16 mlp_half4 __attribute__((ext_vector_type(4)));
typedef float mlp_float16 __attribute__((ext_vector_type(16)));

template <typename OffT>
__global__ void __launch_bounds__(256) cost_kernel(const OffT *__restrict__ beg, long long V,
                                                   const float *__restrict__ feat,
                                                   const float *__restrict__ params,
                                                   float *__restrict__ cost,
                                                   double *__restrict__ block_sum)
{
    constexpr int H = kMlpHidden;
    static_assert(H == 32, "one 32 x 32 MFMA tile per hidden layer");
    __shared__ float p[kMlpParams<H>];
    __shared__ double part[4];
    for (int i = threadIdx.x; i < kMlpParams<H>; i += blockDim.x)
        p[i] = params[i];
    __syncthreads();
    const float *W1 = p, *b1 = p + 3 * H, *W2 = p + 4 * H, *b2 = W2 + H * H, *W3 = b2 + H,
                *b3 = W3 + H;
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    const int col = lane & 31, half = lane >> 5;
    // A: W2 (output unit x input unit), the same for every tile
    mlp_half4 a[4];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int q = 0; q < 4; q++)
            a[c][q] = (_Float16)W2[col * H + 8 * c + 4 * half + q];
    // this lane's output units: their bias and last-layer weight
    float b2r[16], w3r[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int n = 8 * (r / 4) + 4 * half + (r % 4);
        b2r[r] = b2[n];
        w3r[r] = W3[n];
    }
    double local = 0;
    const long long tiles = (V + 31) / 32;
    for (long long t = (long long)blockIdx.x * (blockDim.x / kWave) + wave; t < tiles;
         t += (long long)gridDim.x * (blockDim.x / kWave)) {
        const long long u = t * 32 + col;
        const bool in = u < V;
        int d = 0;
        float x0 = 0.f, x1 = 0.f, x2 = 0.f;
        if (in) {
            d = (int)(beg[u + 1] - beg[u]);
            x0 = feat[3 * u];
            x1 = feat[3 * u + 1];
            x2 = feat[3 * u + 2];
        }
        // B: this lane's half of vertex u's first-layer units
        mlp_half4 b[4];
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int k = 8 * c + 4 * half + q;
                const float h = b1[k] + W1[3 * k] * x0 + W1[3 * k + 1] * x1 + W1[3 * k + 2] * x2;
                b[c][q] = (_Float16)(h > 0.f ? h : 0.f);
            }
        mlp_float16 acc;
#pragma unroll
        for (int r = 0; r < 16; r++)
            acc[r] = b2r[r];
#pragma unroll
        for (int c = 0; c < 4; c++)
            acc = __builtin_amdgcn_mfma_f32_32x32x8f16(a[c], b[c], acc, 0, 0, 0);
        float y = 0.f;
#pragma unroll
        for (int r = 0; r < 16; r++)
            y += w3r[r] * (acc[r] > 0.f ? acc[r] : 0.f);
        y += __shfl_xor(y, 32);
        y = fminf(fmaxf(y + b3[0], 0.f), 1.f);  // log2(1+cycles)/32 in [0, 1]
        const float cst = (in && d >= 2) ? exp2f(y * 32.f) - 1.f : 0.f;
        if (in && half == 0) {
            cost[u] = cst;
            local += (double)cst;
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        local += __shfl_xor(local, off);
    if (lane == 0)
        part[wave] = local;
    __syncthreads();
    if (threadIdx.x == 0)
        block_sum[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// total = sum of block partial sums, one wave: lane i adds blocks i, i+64,
// ... in order, then a fixed butterfly (the result does not depend on
// timing).
__global__ void sum_blocks_kernel(const double *block_sum, int nb, double *total)
{
    double a = 0;
    for (int b = threadIdx.x; b < nb; b += 64)
        a += block_sum[b];
    for (int off = 32; off > 0; off >>= 1)
        a += __shfl_xor(a, off);
    if (threadIdx.x == 0)
        *total = a;
}

struct KeyArgs {
    long long V;
    int same_csr;
    int wave_lg, hub_lg, big_lg;
    int force_spill;
    int wave_only;
    int hub_threshold;      // walk length above which a vertex takes a workgroup
    double split_fraction;  // wave -> workgroup when cost > fraction * total / waves
    double waves;
};

template <typename OffT>
__global__ void __launch_bounds__(256) keys_kernel(const OffT *__restrict__ beg,
                                                   const OffT *__restrict__ begWalk, KeyArgs k,
                                                   const float *__restrict__ cost,
                                                   const double *__restrict__ total,
                                                   unsigned short *__restrict__ keys)
{
    const double tau = k.split_fraction * (*total) / k.waves;
    for (long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x; u < k.V;
         u += (long long)gridDim.x * blockDim.x) {
        const long long d = (long long)(beg[u + 1] - beg[u]);
        unsigned short key = kKeyNone;
        // vertices with an empty walk row (short-side walk) have nothing to count
        if (d >= 2 && (!begWalk || begWalk[u + 1] > begWalk[u])) {
            int lg = kMinTableLg;
            while ((1ll << lg) < 2 * d)
                lg++;
            int tier;
            if (k.force_spill || lg > max(k.hub_lg, k.big_lg))
                tier = kTierSpill;
            else if (lg > k.hub_lg)
                tier = kTierBig;
            else if (lg > k.wave_lg)
                tier = kTierHub;
            else if (!k.wave_only &&
                     (begWalk ? (long long)(begWalk[u + 1] - begWalk[u]) : d) > k.hub_threshold)
                tier = kTierHub;  // long walks take a workgroup, as in the degree split
            else
                tier = (double)cost[u] > tau ? kTierHub : kTierWave;
            int bin = (int)(log2f(1.f + cost[u]) * 2.f);
            bin = bin < 0 ? 0 : (bin >= kBins ? kBins - 1 : bin);
            key = (unsigned short)(tier * kBins + (kBins - 1 - bin));
        }
        keys[u] = key;
    }
}

// ---- stable counting sort of vertices by key into the work list ----
__global__ void __launch_bounds__(256) key_hist_kernel(const unsigned short *__restrict__ keys,
                                                       long long V, long long per_block,
                                                       long long *__restrict__ counts, int nb)
{
    __shared__ unsigned int h[kKeys];
    for (int i = threadIdx.x; i < kKeys; i += blockDim.x)
        h[i] = 0;
    __syncthreads();
    const long long lo = (long long)blockIdx.x * per_block, hi = min(lo + per_block, V);
    for (long long u = lo + threadIdx.x; u < hi; u += blockDim.x) {
        const unsigned short key = keys[u];
        if (key != kKeyNone)
            atomicAdd(&h[key], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kKeys; i += blockDim.x)
        counts[(long long)i * nb + blockIdx.x] = h[i];
}

// Exclusive scan over the (key, block) counts, in place, in the list order
// tier -> superblock -> cost bin -> block: with nsb superblocks of
// consecutive vertex blocks, each tier drains the vertex ranges in turn,
// longest-first inside each (nsb = 1: longest-first over the whole tier).
// A batch of graphs uses nsb > 1 so that graphs complete progressively.
// Tier t starts at the first cell of its range.
__device__ __forceinline__ long long scan_cell(long long i, int nb, int nsb)
{
    const long long per_tier = (long long)kBins * nb;
    const long long t = i / per_tier, r = i % per_tier;
    const int B = (nb + nsb - 1) / nsb;  // blocks per superblock
    const int nsb_used = (nb + B - 1) / B;
    long long sb = r / ((long long)kBins * B);
    sb = sb < nsb_used - 1 ? sb : nsb_used - 1;
    const long long rr = r - sb * kBins * B;
    const long long b0 = sb * B;
    const long long sz = (long long)min(nb, (int)(b0 + B)) - b0;
    const long long bin = rr / sz, blk = b0 + rr % sz;
    return (t * kBins + bin) * nb + blk;
}

__global__ void __launch_bounds__(1024) key_scan_kernel(long long *counts, int nb, int nsb,
                                                        TcCtl *ctl)
{
    __shared__ long long part[1024];
    __shared__ long long base;
    if (threadIdx.x == 0)
        base = 0;
    __syncthreads();
    const long long n = (long long)kKeys * nb;
    const long long per_tier = (long long)kBins * nb;
    for (long long c0 = 0; c0 < n; c0 += blockDim.x) {
        const long long i = c0 + threadIdx.x;
        const long long cell = i < n ? scan_cell(i, nb, nsb) : 0;
        const long long x = i < n ? counts[cell] : 0;
        part[threadIdx.x] = x;
        __syncthreads();
        for (int off = 1; off < (int)blockDim.x; off <<= 1) {
            const long long y = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += y;
            __syncthreads();
        }
        const long long incl = part[threadIdx.x];
        if (i < n) {
            const long long off = base + incl - x;
            counts[cell] = off;
            if (i % per_tier == 0)
                ctl->list_begin[i / per_tier] = off;
        }
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1)
            base += incl;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        ctl->list_begin[kNumTiers] = base;
}

__global__ void __launch_bounds__(256) key_scatter_kernel(const unsigned short *__restrict__ keys,
                                                          long long V, long long per_block,
                                                          const long long *__restrict__ offsets,
                                                          int nb, int32_t *__restrict__ list)
{
    __shared__ long long run[kKeys];
    __shared__ unsigned short wcnt[4][kKeys];
    const int wave = threadIdx.x / 64, lane = lane_id();
    for (int i = threadIdx.x; i < kKeys; i += blockDim.x) {
        run[i] = offsets[(long long)i * nb + blockIdx.x];
        wcnt[0][i] = wcnt[1][i] = wcnt[2][i] = wcnt[3][i] = 0;
    }
    __syncthreads();
    const long long lo = (long long)blockIdx.x * per_block, hi = min(lo + per_block, V);
    for (long long u0 = lo; u0 < hi; u0 += blockDim.x) {
        const long long u = u0 + threadIdx.x;
        const int key = u < hi ? (int)keys[u] : (int)kKeyNone;
        int rank = 0;
        unsigned long long remaining = __ballot(key != (int)kKeyNone);
        while (remaining) {
            const int lead = __ffsll((long long)remaining) - 1;
            const int kk = __shfl(key, lead);
            const unsigned long long m = __ballot(key == kk);
            if (key == kk)
                rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned int)(m >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((unsigned int)m, 0));
            if (lane == 0)
                wcnt[wave][kk] = (unsigned short)__popcll(m);
            remaining &= ~m;
        }
        __syncthreads();
        if (key != (int)kKeyNone) {
            long long pos = run[key] + rank;
            for (int w = 0; w < wave; w++)
                pos += wcnt[w][key];
            list[pos] = (int32_t)u;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < kKeys; i += blockDim.x) {
            run[i] += (long long)wcnt[0][i] + wcnt[1][i] + wcnt[2][i] + wcnt[3][i];
            wcnt[0][i] = wcnt[1][i] = wcnt[2][i] = wcnt[3][i] = 0;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------- host side
// Features over the walk the count takes: the rows of N+ (the reference
// walk) or, with walk = true, the short-side walk's rows; table sizes and
// list lengths are always d+.
template <typename OffT>
static int enqueue_features(Ctx *ctx, float *feat, hipStream_t s, bool walk)
{
    const long long V = ctx->g.V;
    if (V == 0)
        return TRUST_OK;
    const int blocks = (int)std::min<long long>((V + 255) / 256, (long long)ctx->num_cus * 32);
    const OffT *begD = (const OffT *)ctx->g.beg.p;
    if (!walk) {
        hipLaunchKernelGGL(features_kernel<OffT>, dim3(blocks), dim3(256), 0, s, begD,
                           (const int32_t *)ctx->g.adj.p, begD, V, feat);
        HIPCHK(hipGetLastError());
        return TRUST_OK;
    }
    int rc = ctx->walk_acc.ensure(12 * (size_t)V);
    if (rc)
        return rc;
    unsigned long long *wsum = (unsigned long long *)ctx->walk_acc.p;
    unsigned int *wmax = (unsigned int *)(wsum + V);
    HIPCHK(launch_fill_u32((uint32_t *)ctx->walk_acc.p, 0u, 3 * V, s));
    int64_t X = 0;
    if ((rc = side_entry_count(ctx, &X)))  // read back already when under capture
        return rc;
    if (X > 0) {
        // the length field of every WalkEntry<OffT> of the walk index
        const int32_t *deg = (const int32_t *)((const char *)ctx->side_ent.p +
                                               offsetof(WalkEntry<OffT>, deg));
        const int dstride = (int)(sizeof(WalkEntry<OffT>) / sizeof(int32_t));
        const int wb = (int)std::min<long long>((X + 255) / 256, (long long)ctx->num_cus * 64);
        hipLaunchKernelGGL(walk_feat_kernel, dim3(wb), dim3(256), 0, s,
                           (const int32_t *)ctx->side_row.p, deg, dstride, X, wsum, wmax);
    }
    hipLaunchKernelGGL(walk_feat_final_kernel<OffT>, dim3(blocks), dim3(256), 0, s, begD, V,
                       (const unsigned long long *)wsum, (const unsigned int *)wmax, feat);
    HIPCHK(hipGetLastError());
    return TRUST_OK;
}

// Features of the given walk into ctx->feat (which then holds that walk's).
static int features_dispatch(Ctx *ctx, hipStream_t s, bool walk = false)
{
    int rc = ctx->feat.ensure(sizeof(float) * 3 * (size_t)std::max<int64_t>(ctx->g.V, 1));
    if (rc)
        return rc;
    ctx->feat_walk = -1;
    if (walk && (rc = ensure_walk_rows(ctx)))  // written before any capture (prepare_resident)
        return rc;
    float *feat = (float *)ctx->feat.p;
    rc = ctx->g.off64 ? enqueue_features<int64_t>(ctx, feat, s, walk)
                      : enqueue_features<int32_t>(ctx, feat, s, walk);
    if (rc)
        return rc;
    ctx->feat_walk = walk ? 1 : 0;
    return TRUST_OK;
}

int ensure_features(Ctx *ctx, bool walk)
{
    if (ctx->feat_walk == (walk ? 1 : 0))
        return TRUST_OK;
    return features_dispatch(ctx, ctx->stream, walk);
}

// Cost prediction, tier keys and the longest-first list of the vertices
// [0, V) of one hash view (beg): list and ctl->list_begin.  feat, cost and
// keys are this range's slices.
template <typename OffT>
static int plan_from_features(Ctx *ctx, const trust_params &p, const LaunchPlan &L, TcCtl *ctl,
                              hipStream_t s, const OffT *beg, const OffT *begWalk, long long V,
                              const float *feat, float *cost, unsigned short *keys, int32_t *list)
{
    const int cb = (int)std::max<long long>(1, std::min<long long>((V + 255) / 256, 4096));
    // block sums (cb doubles) + total (1 double) live in block_counts' space
    const long long per_block = std::max<long long>(4096, (V + 1023) / 1024);
    const int nb = (int)std::max<long long>(1, (V + per_block - 1) / per_block);
    int rc = ctx->block_counts.ensure(sizeof(long long) * ((size_t)kKeys * nb + cb + 1));
    if (rc)
        return rc;
    long long *counts = (long long *)ctx->block_counts.p;
    double *bsum = (double *)(counts + (size_t)kKeys * nb);
    double *total = bsum + cb;
    hipLaunchKernelGGL(cost_kernel<OffT>, dim3(cb), dim3(256), 0, s, beg, V, feat,
                       (const float *)ctx->mlp_params.p, cost, bsum);
    hipLaunchKernelGGL(sum_blocks_kernel, dim3(1), dim3(64), 0, s, bsum, cb, total);
    KeyArgs k;
    k.V = V;
    k.same_csr = 1;
    k.wave_lg = L.main.wave_lg;
    k.hub_lg = L.main.group_lg;
    k.big_lg = L.has_big ? L.big.group_lg : 0;
    k.force_spill = (p.flags & TRUST_FLAG_FORCE_SPILL) ? 1 : 0;
    k.wave_only = (p.flags & TRUST_FLAG_WAVE_ONLY) ? 1 : 0;
    k.hub_threshold = p.hub_threshold;
    k.split_fraction = 0.5;
    k.waves = (double)L.main.blocks * (L.main.threads / kWave);
    hipLaunchKernelGGL(keys_kernel<OffT>, dim3(cb), dim3(256), 0, s, beg, begWalk, k,
                       (const float *)cost, (const double *)total, keys);
    hipLaunchKernelGGL(key_hist_kernel, dim3(nb), dim3(256), 0, s, (const unsigned short *)keys, V,
                       per_block, counts, nb);
    // a batch drains its graphs in 32 vertex ranges (progressive completion)
    const int nsb = ctx->batch_ngraphs > 1 ? 32 : 1;
    hipLaunchKernelGGL(key_scan_kernel, dim3(1), dim3(1024), 0, s, counts, nb, nsb, ctl);
    hipLaunchKernelGGL(key_scatter_kernel, dim3(nb), dim3(256), 0, s, (const unsigned short *)keys,
                       V, per_block, (const long long *)counts, nb, list);
    HIPCHK(hipGetLastError());
    return TRUST_OK;
}

static int ensure_plan_buffers(Ctx *ctx)
{
    const long long V = std::max<long long>(ctx->g.V, 1);
    int rc;
    if ((rc = ctx->feat.ensure(sizeof(float) * 3 * V)) ||
        (rc = ctx->cost.ensure(sizeof(float) * V)) ||
        (rc = ctx->tiers.ensure(sizeof(unsigned short) * V)) ||
        (rc = ctx->list.ensure(sizeof(int32_t) * V)))
        return rc;
    return TRUST_OK;
}

template <typename OffT>
static int enqueue_learned_plan_t(Ctx *ctx, const trust_params &p, const LaunchPlan &L,
                                  TcCtl *ctl, hipStream_t s, const void *begWalk)
{
    if (ctx->feat_walk != (begWalk != nullptr ? 1 : 0))
        return fail(TRUST_ERR_ARG, "learned plan: features of this walk not prepared");
    int rc = ensure_plan_buffers(ctx);
    if (rc)
        return rc;
    HIPCHK(launch_fill_u32((uint32_t *)ctl, 0u, sizeof(TcCtl) / 4, s));
    return plan_from_features<OffT>(ctx, p, L, ctl, s, (const OffT *)ctx->g.beg.p,
                                    (const OffT *)begWalk, ctx->g.V, (const float *)ctx->feat.p,
                                    (float *)ctx->cost.p, (unsigned short *)ctx->tiers.p,
                                    (int32_t *)ctx->list.p);
}

int enqueue_learned_plan_segment(Ctx *ctx, const trust_params &p, const LaunchPlan &L, TcCtl *ctl,
                                 hipStream_t s, const CsrViews &v, int64_t V, int64_t v0)
{
    if (!ctx->mlp_ready)
        return fail(TRUST_ERR_ARG, "no cost model");
    if (ctx->feat_walk != 0)  // segments walk N+: the features of the whole batch, by vertex
        return fail(TRUST_ERR_ARG, "learned plan: features not prepared");
    int rc = ensure_plan_buffers(ctx);
    if (rc)
        return rc;
    HIPCHK(launch_fill_u32((uint32_t *)ctl, 0u, sizeof(TcCtl) / 4, s));
    if (V == 0)
        return TRUST_OK;
    const float *feat = (const float *)ctx->feat.p + 3 * v0;
    return plan_from_features<int32_t>(ctx, p, L, ctl, s, (const int32_t *)v.begH, nullptr, V, feat,
                                       (float *)ctx->cost.p + v0,
                                       (unsigned short *)ctx->tiers.p + v0,
                                       (int32_t *)ctx->list.p + v0);
}

int enqueue_learned_plan(Ctx *ctx, const trust_params &p, const LaunchPlan &L, TcCtl *ctl,
                         hipStream_t s, const void *begWalk)
{
    if (!ctx->mlp_ready)
        return fail(TRUST_ERR_ARG, "no cost model");
    return ctx->g.off64 ? enqueue_learned_plan_t<int64_t>(ctx, p, L, ctl, s, begWalk)
                        : enqueue_learned_plan_t<int32_t>(ctx, p, L, ctl, s, begWalk);
}

// Instrumented wave-only pass -> features -> GPU training.
int train_cost_model(Ctx *ctx, const trust_params &p_in, int epochs, long long max_samples,
                     uint64_t seed, double *loss_out, double *seconds_out)
{
    const long long V = ctx->g.V;
    if (V < 1)
        return fail(TRUST_ERR_ARG, "context holds no graph");
    trust_params p = p_in;
    p.schedule = TRUST_SCHED_DEGREE;
    p.flags |= TRUST_FLAG_WAVE_ONLY;
    p.rank = 0;
    p.total_process = 1;
    std::vector<unsigned long long> cyc((size_t)V);
    trust_result r;
    int rc = count_resident(ctx, p, kModeInstr, cyc.data(), nullptr, &r);
    if (rc)
        return rc;
    const double t0 = r.seconds;
    rc = ctx->feat.ensure(sizeof(float) * 3 * V);
    if (rc)
        return rc;
    // features of the walk the instrumented count took
    rc = features_dispatch(ctx, ctx->stream, r.walk_used != 0);
    if (rc)
        return rc;
    std::vector<float> feat((size_t)3 * V);
    HIPCHK(hipMemcpyAsync(feat.data(), ctx->feat.p, sizeof(float) * 3 * V, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // training set: counted vertices (cycles > 0), a seeded sample
    std::vector<long long> idx;
    for (long long u = 0; u < V; u++)
        if (cyc[u] > 0)
            idx.push_back(u);
    if (idx.empty())
        return fail(TRUST_ERR_ARG, "no counted vertex to train on");
    uint64_t st = seed ^ 0x5851F42D4C957F2Dull;
    auto rnd = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const long long n = std::min<long long>((long long)idx.size(), std::max<long long>(max_samples, 1));
    for (long long i = 0; i < n; i++)
        std::swap(idx[i], idx[i + (long long)(rnd() % (uint64_t)(idx.size() - i))]);
    // vertices too big for a wave table ran on a workgroup: scale to wave-cycles
    const double wg_waves = r.threads / kWave;
    std::vector<float> x((size_t)3 * n), t((size_t)n);
    for (long long i = 0; i < n; i++) {
        const long long u = idx[i];
        x[3 * i] = feat[3 * u];
        x[3 * i + 1] = feat[3 * u + 1];
        x[3 * i + 2] = feat[3 * u + 2];
        const long long d = (long long)std::llround(exp2((double)feat[3 * u] / kFeatScale) - 1.0);
        const bool wave_table = host_table_lg(d) <= 31 && (1ll << host_table_lg(d)) <= r.wave_slots;
        double c = (double)cyc[u] * (wave_table ? 1.0 : wg_waves);
        t[i] = (float)(log2(1.0 + c) * kFeatScale);
    }
    DevBuf dx, dt;
    const int P = kMlpParams<kMlpHidden>;
    if ((rc = dx.ensure(sizeof(float) * 3 * n)) || (rc = dt.ensure(sizeof(float) * n)) ||
        (rc = ctx->mlp_params.ensure(sizeof(float) * P)))
        return rc;
    std::vector<float> init(P);
    mlp_init_host(seed, init.data());
    HIPCHK(hipMemcpyAsync(ctx->mlp_params.p, init.data(), sizeof(float) * P, hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(dx.p, x.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dt.p, t.data(), sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
    rc = mlp_train_device(ctx, (float *)ctx->mlp_params.p, (const float *)dx.p, (const float *)dt.p,
                          n, epochs, 4096, 3e-3f, seed, loss_out);
    if (rc)
        return rc;
    HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev[3], ctx->ev[2]));
    ctx->mlp_ready = true;
    // the degree split judges the trained model: one learned and one degree
    // count of this graph (best of two each) with the caller's launch
    double judge_s = 0;
    for (int sched : {TRUST_SCHED_LEARNED, TRUST_SCHED_DEGREE}) {
        trust_params q = p_in;
        q.schedule = sched;
        double best = 1e300;
        for (int i = 0; i < 2; i++) {
            trust_result rj;
            if ((rc = count_resident(ctx, q, kModeCount, nullptr, nullptr, &rj)))
                return rc;
            best = std::min(best, rj.seconds);
            judge_s += rj.seconds;
        }
        (sched == TRUST_SCHED_LEARNED ? ctx->judge_learned_s : ctx->judge_degree_s) = best;
    }
    // the learned schedule must win by more than run-to-run noise (1%): a
    // tie goes to the degree split, which has no plan to pay for
    ctx->judge = ctx->judge_learned_s < 0.99 * ctx->judge_degree_s ? TRUST_SCHED_LEARNED
                                                                    : TRUST_SCHED_DEGREE;
    if (seconds_out)
        *seconds_out = t0 + ms * 1e-3 + judge_s;
    return TRUST_OK;
}

}  // namespace trust

using namespace trust;

extern "C" int trust_ctx_features(trust_ctx *c, float *feat)
{
    if (!c || !feat)
        return fail(TRUST_ERR_ARG, "null argument");
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    HIPCHK(hipSetDevice(ctx->device));
    const long long V = ctx->g.V;
    int rc = ctx->feat.ensure(sizeof(float) * 3 * std::max<long long>(V, 1));
    if (rc)
        return rc;
    rc = features_dispatch(ctx, ctx->stream);
    if (rc)
        return rc;
    HIPCHK(hipMemcpyAsync(feat, ctx->feat.p, sizeof(float) * 3 * V, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return TRUST_OK;
}

extern "C" int trust_ctx_walk_features(trust_ctx *c, float *feat)
{
    if (!c || !feat)
        return fail(TRUST_ERR_ARG, "null argument");
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    HIPCHK(hipSetDevice(ctx->device));
    const long long V = ctx->g.V;
    int rc = ctx->feat.ensure(sizeof(float) * 3 * std::max<long long>(V, 1));
    if (rc)
        return rc;
    if ((rc = ensure_walk_index(ctx)))
        return rc;
    rc = features_dispatch(ctx, ctx->stream, ctx->g.V > 0 && ctx->g.E > 0);
    if (rc)
        return rc;
    HIPCHK(hipMemcpyAsync(feat, ctx->feat.p, sizeof(float) * 3 * V, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return TRUST_OK;
}

extern "C" int trust_ctx_train_cost_model(trust_ctx *c, const trust_params *p, int epochs,
                                          int64_t max_samples, uint64_t seed, double *loss,
                                          double *seconds)
{
    if (!c || !p || epochs < 0)
        return fail(TRUST_ERR_ARG, "bad arguments");
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    HIPCHK(hipSetDevice(ctx->device));
    return train_cost_model(ctx, *p, epochs, max_samples, seed, loss, seconds);
}

extern "C" int trust_ctx_set_cost_model(trust_ctx *c, int hidden, const float *params)
{
    if (!c || !params || hidden != kMlpHidden)
        return fail(TRUST_ERR_ARG, "bad arguments (hidden must be %d)", kMlpHidden);
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    HIPCHK(hipSetDevice(ctx->device));
    const int P = kMlpParams<kMlpHidden>;
    int rc = ctx->mlp_params.ensure(sizeof(float) * P);
    if (rc)
        return rc;
    HIPCHK(hipMemcpy(ctx->mlp_params.p, params, sizeof(float) * P, hipMemcpyHostToDevice));
    ctx->mlp_ready = true;
    ctx->judge = -1;
    return TRUST_OK;
}

extern "C" int trust_ctx_schedule_judgement(trust_ctx *c, int *decision, double *learned_seconds,
                                            double *degree_seconds)
{
    clear_error();
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    if (!ctx || !decision)
        return fail(TRUST_ERR_ARG, "null argument");
    *decision = ctx->judge;
    if (learned_seconds)
        *learned_seconds = ctx->judge >= 0 ? ctx->judge_learned_s : 0.0;
    if (degree_seconds)
        *degree_seconds = ctx->judge >= 0 ? ctx->judge_degree_s : 0.0;
    return TRUST_OK;
}

extern "C" int trust_ctx_get_cost_model(trust_ctx *c, int *hidden, float *params)
{
    if (!c || !hidden)
        return fail(TRUST_ERR_ARG, "null argument");
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    if (!ctx->mlp_ready)
        return fail(TRUST_ERR_ARG, "no cost model trained or set");
    *hidden = kMlpHidden;
    if (params) {
        HIPCHK(hipSetDevice(ctx->device));
        HIPCHK(hipMemcpy(params, ctx->mlp_params.p, sizeof(float) * kMlpParams<kMlpHidden>,
                         hipMemcpyDeviceToHost));
    }
    return TRUST_OK;
}

extern "C" int trust_ctx_predict_costs(trust_ctx *c, float *cost)
{
    if (!c || !cost)
        return fail(TRUST_ERR_ARG, "null argument");
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    if (!ctx->mlp_ready)
        return fail(TRUST_ERR_ARG, "no cost model trained or set");
    HIPCHK(hipSetDevice(ctx->device));
    const long long V = ctx->g.V;
    int rc;
    const int cb = (int)std::max<long long>(1, std::min<long long>((V + 255) / 256, 4096));
    if ((rc = ctx->feat.ensure(sizeof(float) * 3 * std::max<long long>(V, 1))) ||
        (rc = ctx->cost.ensure(sizeof(float) * std::max<long long>(V, 1))) ||
        (rc = ctx->block_counts.ensure(sizeof(double) * (cb + 1))))
        return rc;
    rc = features_dispatch(ctx, ctx->stream);
    if (rc)
        return rc;
    if (ctx->g.off64)
        hipLaunchKernelGGL(cost_kernel<int64_t>, dim3(cb), dim3(256), 0, ctx->stream,
                           (const int64_t *)ctx->g.beg.p, V, (const float *)ctx->feat.p,
                           (const float *)ctx->mlp_params.p, (float *)ctx->cost.p,
                           (double *)ctx->block_counts.p);
    else
        hipLaunchKernelGGL(cost_kernel<int32_t>, dim3(cb), dim3(256), 0, ctx->stream,
                           (const int32_t *)ctx->g.beg.p, V, (const float *)ctx->feat.p,
                           (const float *)ctx->mlp_params.p, (float *)ctx->cost.p,
                           (double *)ctx->block_counts.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(cost, ctx->cost.p, sizeof(float) * V, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return TRUST_OK;
}
