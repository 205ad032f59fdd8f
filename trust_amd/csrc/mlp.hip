// Cost-model MLP: forward, loss + backward, gradient reduction and Adam,
// hand-written HIP for gfx950 (no BLAS: the matrices are 32x32).
//
// Network: x (3) -> relu(W1 x + b1) (H) -> relu(W2 h1 + b2) (H) -> W3 h2 + b3.
// Parameter layout (floats): W1[H][3], b1[H], W2[H][H], b2[H], W3[H], b3
// = H^2 + 6H + 1 values.  Loss: mean squared error over the samples.
//
// One thread per sample.  In the backward pass each wave reduces every
// gradient component over its 64 lanes with a butterfly of cross-lane
// shuffles, writes the wave's partial to its own LDS row, and the block
// sums its waves in a fixed order into one partial per block; a second
// kernel sums the block partials in a fixed order.  No floating-point
// atomics, so gradients are deterministic run to run.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "engine.h"
#include "mlp.h"

namespace trust {

constexpr int kMlpThreads = 256;

__device__ __forceinline__ float wave_sum_f(float x)
{
    for (int off = 32; off > 0; off >>= 1)
        x += __shfl_xor(x, off);
    return x;
}

template <int H>
struct MlpView {
    const float *W1, *b1, *W2, *b2, *W3, *b3;
    __device__ MlpView(const float *p)
        : W1(p), b1(p + 3 * H), W2(p + 4 * H), b2(p + 4 * H + H * H), W3(p + 5 * H + H * H),
          b3(p + 6 * H + H * H)
    {
    }
};

template <int H>
__device__ __forceinline__ float mlp_eval(const MlpView<H> &m, const float x[3], float h1[H],
                                          float h2[H])
{
#pragma unroll
    for (int i = 0; i < H; i++) {
        float a = m.b1[i] + m.W1[3 * i] * x[0] + m.W1[3 * i + 1] * x[1] + m.W1[3 * i + 2] * x[2];
        h1[i] = a > 0.f ? a : 0.f;
    }
    float y = m.b3[0];
#pragma unroll 4
    for (int i = 0; i < H; i++) {
        float a = m.b2[i];
#pragma unroll
        for (int j = 0; j < H; j++)
            a += m.W2[i * H + j] * h1[j];
        h2[i] = a > 0.f ? a : 0.f;
        y += m.W3[i] * h2[i];
    }
    return y;
}

template <int H>
__global__ void __launch_bounds__(kMlpThreads) mlp_forward_kernel(const float *__restrict__ params,
                                                                  const float *__restrict__ x,
                                                                  long long n, float *__restrict__ y)
{
    __shared__ float p[kMlpParams<H>];
    for (int i = threadIdx.x; i < kMlpParams<H>; i += blockDim.x)
        p[i] = params[i];
    __syncthreads();
    MlpView<H> m(p);
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (long long)gridDim.x * blockDim.x) {
        float xi[3] = {x[3 * s], x[3 * s + 1], x[3 * s + 2]};
        float h1[H], h2[H];
        y[s] = mlp_eval<H>(m, xi, h1, h2);
    }
}

// Per-block partial gradient (and loss sum in the last slot) of
// sum_s (y_s - t_s)^2 / n over samples [blockIdx*256, +256) of [0, n).
template <int H>
__global__ void __launch_bounds__(kMlpThreads) mlp_grad_kernel(const float *__restrict__ params,
                                                               const float *__restrict__ x,
                                                               const float *__restrict__ t,
                                                               long long n, float inv_n,
                                                               float *__restrict__ partial)
{
    constexpr int P = kMlpParams<H>;
    __shared__ float p[P];
    __shared__ float wpart[kMlpThreads / 64][P + 1];
    for (int i = threadIdx.x; i < P; i += blockDim.x)
        p[i] = params[i];
    __syncthreads();
    MlpView<H> m(p);
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = s < n;
    float xi[3] = {0.f, 0.f, 0.f};
    float ti = 0.f;
    if (live) {
        xi[0] = x[3 * s];
        xi[1] = x[3 * s + 1];
        xi[2] = x[3 * s + 2];
        ti = t[s];
    }
    float h1[H], h2[H];
    const float y = mlp_eval<H>(m, xi, h1, h2);
    const float err = live ? y - ti : 0.f;
    const float dy = 2.f * err * inv_n;
    float *wp = wpart[threadIdx.x / 64];
    const int lane = threadIdx.x % 64;
    auto put = [&](int idx, float g) {
        g = wave_sum_f(g);
        if (lane == 0)
            wp[idx] = g;
    };
    // layer 3
    put(6 * H + H * H, dy);  // b3
    float dh2[H];
#pragma unroll
    for (int i = 0; i < H; i++) {
        put(5 * H + H * H + i, dy * h2[i]);  // W3
        dh2[i] = h2[i] > 0.f ? dy * m.W3[i] : 0.f;
    }
    // layer 2
    float dh1[H];
#pragma unroll
    for (int j = 0; j < H; j++)
        dh1[j] = 0.f;
#pragma unroll 4
    for (int i = 0; i < H; i++) {
        put(4 * H + H * H + i, dh2[i]);  // b2
#pragma unroll
        for (int j = 0; j < H; j++) {
            put(4 * H + i * H + j, dh2[i] * h1[j]);  // W2[i][j]
            dh1[j] += m.W2[i * H + j] * dh2[i];
        }
    }
    // layer 1
#pragma unroll
    for (int j = 0; j < H; j++) {
        const float g = h1[j] > 0.f ? dh1[j] : 0.f;
        put(3 * H + j, g);  // b1
        put(3 * j, g * xi[0]);
        put(3 * j + 1, g * xi[1]);
        put(3 * j + 2, g * xi[2]);
    }
    put(P, err * err);  // loss contribution
    __syncthreads();
    for (int i = threadIdx.x; i <= P; i += blockDim.x) {
        float a = 0.f;
        for (int w = 0; w < kMlpThreads / 64; w++)
            a += wpart[w][i];
        partial[(long long)blockIdx.x * (P + 1) + i] = a;
    }
}

// grad[i] = sum over blocks of partial[b][i] (fixed order); slot P = loss sum.
template <int H>
__global__ void mlp_reduce_kernel(const float *__restrict__ partial, int nblocks, float *__restrict__ grad)
{
    constexpr int P = kMlpParams<H>;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= P; i += gridDim.x * blockDim.x) {
        float a = 0.f;
        for (int b = 0; b < nblocks; b++)
            a += partial[(long long)b * (P + 1) + i];
        grad[i] = a;
    }
}

__global__ void adam_kernel(float *__restrict__ params, const float *__restrict__ grad, float *m,
                            float *v, int P, float lr, float b1, float b2, float eps, float c1,
                            float c2)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
        const float g = grad[i];
        const float mi = b1 * m[i] + (1.f - b1) * g;
        const float vi = b2 * v[i] + (1.f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        params[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// ---------------------------------------------------------------- host
hipError_t mlp_forward(const float *params, const float *x, long long n, float *y, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    int blocks = (int)std::min<long long>((n + kMlpThreads - 1) / kMlpThreads, 8192);
    hipLaunchKernelGGL(mlp_forward_kernel<kMlpHidden>, dim3(blocks), dim3(kMlpThreads), 0, s,
                       params, x, n, y);
    return hipGetLastError();
}

// grad_out holds P+1 floats (gradient, then the summed squared error).
hipError_t mlp_loss_grad(const float *params, const float *x, const float *t, long long n,
                         float *partial, float *grad_out, hipStream_t s)
{
    const int nb = (int)((n + kMlpThreads - 1) / kMlpThreads);
    if (nb <= 0)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(mlp_grad_kernel<kMlpHidden>, dim3(nb), dim3(kMlpThreads), 0, s, params, x,
                       t, n, 1.f / (float)n, partial);
    hipLaunchKernelGGL(mlp_reduce_kernel<kMlpHidden>, dim3((kMlpParams<kMlpHidden> + 256) / 256),
                       dim3(256), 0, s, partial, nb, grad_out);
    return hipGetLastError();
}

hipError_t adam_step(float *params, const float *grad, float *m, float *v, int P, float lr, int t,
                     hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(adam_kernel, dim3((P + 255) / 256), dim3(256), 0, s, params, grad, m, v, P,
                       lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

void mlp_init_host(uint64_t seed, float *params)
{
    const int H = kMlpHidden;
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    auto uni = [&]() {  // splitmix64 -> [-1, 1)
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (float)((double)(z >> 11) / 4503599627370496.0 - 1.0);
    };
    float *W1 = params, *b1 = W1 + 3 * H, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H,
          *b3 = W3 + H;
    const float a1 = sqrtf(6.f / 3.f), a2 = sqrtf(6.f / H), a3 = sqrtf(3.f / H);
    for (int i = 0; i < 3 * H; i++)
        W1[i] = a1 * uni();
    for (int i = 0; i < H; i++)
        b1[i] = 0.01f;
    for (int i = 0; i < H * H; i++)
        W2[i] = a2 * uni();
    for (int i = 0; i < H; i++)
        b2[i] = 0.01f;
    for (int i = 0; i < H; i++)
        W3[i] = a3 * uni();
    b3[0] = 0.f;
}

// Minibatch Adam on device-resident samples (x: n*3, t: n).  Shuffles by
// a seeded permutation each epoch.  Returns the final full-data MSE.
int mlp_train_device(Ctx *ctx, float *d_params, const float *d_x, const float *d_t, long long n,
                     int epochs, int batch, float lr, uint64_t seed, double *loss_out)
{
    const int P = kMlpParams<kMlpHidden>;
    if (n < 1 || epochs < 0 || batch < 1)
        return fail(TRUST_ERR_ARG, "bad training arguments");
    batch = (int)std::min<long long>(batch, n);
    DevBuf m, v, grad, partial, bx, bt;
    const long long nb_max = (std::max<long long>(n, batch) + kMlpThreads - 1) / kMlpThreads;
    int rc = m.ensure(sizeof(float) * P);
    if (!rc)
        rc = v.ensure(sizeof(float) * P);
    if (!rc)
        rc = grad.ensure(sizeof(float) * (P + 1));
    if (!rc)
        rc = partial.ensure(sizeof(float) * (P + 1) * nb_max);
    if (!rc)
        rc = bx.ensure(sizeof(float) * 3 * batch);
    if (!rc)
        rc = bt.ensure(sizeof(float) * batch);
    if (rc)
        return rc;
    HIPCHK(hipMemsetAsync(m.p, 0, sizeof(float) * P, ctx->stream));
    HIPCHK(hipMemsetAsync(v.p, 0, sizeof(float) * P, ctx->stream));
    // host-side permutation, uploaded per epoch as gather indices
    std::vector<long long> perm((size_t)n);
    for (long long i = 0; i < n; i++)
        perm[i] = i;
    DevBuf dperm;
    rc = dperm.ensure(sizeof(long long) * n);
    if (rc)
        return rc;
    uint64_t st = seed ^ 0xD1B54A32D192ED03ull;
    auto rnd = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    int step = 0;
    for (int ep = 0; ep < epochs; ep++) {
        for (long long i = n - 1; i > 0; i--)
            std::swap(perm[i], perm[rnd() % (uint64_t)(i + 1)]);
        HIPCHK(hipMemcpyAsync(dperm.p, perm.data(), sizeof(long long) * n, hipMemcpyHostToDevice,
                              ctx->stream));
        for (long long b0 = 0; b0 + batch <= n; b0 += batch) {
            HIPCHK(launch_gather_samples(d_x, d_t, (const long long *)dperm.p + b0, batch,
                                         (float *)bx.p, (float *)bt.p, ctx->stream));
            HIPCHK(mlp_loss_grad(d_params, (const float *)bx.p, (const float *)bt.p, batch,
                                 (float *)partial.p, (float *)grad.p, ctx->stream));
            HIPCHK(adam_step(d_params, (const float *)grad.p, (float *)m.p, (float *)v.p, P, lr,
                             ++step, ctx->stream));
        }
        // keep host permutation buffer alive until the copy has been consumed
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    if (loss_out) {
        HIPCHK(mlp_loss_grad(d_params, d_x, d_t, n, (float *)partial.p, (float *)grad.p,
                             ctx->stream));
        float sse = 0;
        HIPCHK(hipMemcpyAsync(&sse, (float *)grad.p + P, sizeof(float), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *loss_out = (double)sse / (double)n;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return TRUST_OK;
}

__global__ void gather_samples_kernel(const float *__restrict__ x, const float *__restrict__ t,
                                      const long long *__restrict__ idx, long long n,
                                      float *__restrict__ bx, float *__restrict__ bt)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const long long s = idx[i];
        bx[3 * i] = x[3 * s];
        bx[3 * i + 1] = x[3 * s + 1];
        bx[3 * i + 2] = x[3 * s + 2];
        bt[i] = t[s];
    }
}

hipError_t launch_gather_samples(const float *x, const float *t, const long long *idx, long long n,
                                 float *bx, float *bt, hipStream_t s)
{
    int g = (int)std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(gather_samples_kernel, dim3(g), dim3(256), 0, s, x, t, idx, n, bx, bt);
    return hipGetLastError();
}

}  // namespace trust

using namespace trust;

// ------------------------------------------------------------------ C ABI
namespace {
struct HostToDev {  // upload helper for the stand-alone MLP entry points
    Ctx ctx;
    DevBuf params, x, t, y, partial, grad;
};
}  // namespace

extern "C" int64_t trust_mlp_param_count(int hidden)
{
    return hidden == kMlpHidden ? (int64_t)kMlpParams<kMlpHidden> : -1;
}

extern "C" int trust_mlp_init(int hidden, uint64_t seed, float *params)
{
    if (hidden != kMlpHidden || !params)
        return fail(TRUST_ERR_ARG, "hidden must be %d", kMlpHidden);
    mlp_init_host(seed, params);
    return TRUST_OK;
}

extern "C" int trust_mlp_forward(int device, int hidden, const float *params, const float *x,
                                 int64_t n, float *y)
{
    if (hidden != kMlpHidden || !params || !x || !y || n < 0)
        return fail(TRUST_ERR_ARG, "bad arguments");
    HostToDev h;
    int rc = ctx_init(&h.ctx, device);
    if (rc)
        return rc;
    const int P = kMlpParams<kMlpHidden>;
    if ((rc = h.params.ensure(sizeof(float) * P)) || (rc = h.x.ensure(sizeof(float) * 3 * (n + 1))) ||
        (rc = h.y.ensure(sizeof(float) * (n + 1))))
        return rc;
    HIPCHK(hipMemcpyAsync(h.params.p, params, sizeof(float) * P, hipMemcpyHostToDevice, h.ctx.stream));
    HIPCHK(hipMemcpyAsync(h.x.p, x, sizeof(float) * 3 * n, hipMemcpyHostToDevice, h.ctx.stream));
    HIPCHK(mlp_forward((const float *)h.params.p, (const float *)h.x.p, n, (float *)h.y.p, h.ctx.stream));
    HIPCHK(hipMemcpyAsync(y, h.y.p, sizeof(float) * n, hipMemcpyDeviceToHost, h.ctx.stream));
    HIPCHK(hipStreamSynchronize(h.ctx.stream));
    return TRUST_OK;
}

extern "C" int trust_mlp_loss_grad(int device, int hidden, const float *params, const float *x,
                                   const float *target, int64_t n, float *grad, double *loss)
{
    if (hidden != kMlpHidden || !params || !x || !target || !grad || n < 1)
        return fail(TRUST_ERR_ARG, "bad arguments");
    HostToDev h;
    int rc = ctx_init(&h.ctx, device);
    if (rc)
        return rc;
    const int P = kMlpParams<kMlpHidden>;
    const long long nb = (n + kMlpThreads - 1) / kMlpThreads;
    if ((rc = h.params.ensure(sizeof(float) * P)) || (rc = h.x.ensure(sizeof(float) * 3 * n)) ||
        (rc = h.t.ensure(sizeof(float) * n)) || (rc = h.partial.ensure(sizeof(float) * (P + 1) * nb)) ||
        (rc = h.grad.ensure(sizeof(float) * (P + 1))))
        return rc;
    HIPCHK(hipMemcpyAsync(h.params.p, params, sizeof(float) * P, hipMemcpyHostToDevice, h.ctx.stream));
    HIPCHK(hipMemcpyAsync(h.x.p, x, sizeof(float) * 3 * n, hipMemcpyHostToDevice, h.ctx.stream));
    HIPCHK(hipMemcpyAsync(h.t.p, target, sizeof(float) * n, hipMemcpyHostToDevice, h.ctx.stream));
    HIPCHK(mlp_loss_grad((const float *)h.params.p, (const float *)h.x.p, (const float *)h.t.p, n,
                         (float *)h.partial.p, (float *)h.grad.p, h.ctx.stream));
    std::vector<float> g(P + 1);
    HIPCHK(hipMemcpyAsync(g.data(), h.grad.p, sizeof(float) * (P + 1), hipMemcpyDeviceToHost,
                          h.ctx.stream));
    HIPCHK(hipStreamSynchronize(h.ctx.stream));
    memcpy(grad, g.data(), sizeof(float) * P);
    if (loss)
        *loss = (double)g[P] / (double)n;
    return TRUST_OK;
}

extern "C" int trust_mlp_train(int device, int hidden, float *params, const float *x,
                               const float *target, int64_t n, int epochs, int batch, float lr,
                               uint64_t seed, double *loss_out)
{
    if (hidden != kMlpHidden || !params || !x || !target || n < 1)
        return fail(TRUST_ERR_ARG, "bad arguments");
    HostToDev h;
    int rc = ctx_init(&h.ctx, device);
    if (rc)
        return rc;
    const int P = kMlpParams<kMlpHidden>;
    if ((rc = h.params.ensure(sizeof(float) * P)) || (rc = h.x.ensure(sizeof(float) * 3 * n)) ||
        (rc = h.t.ensure(sizeof(float) * n)))
        return rc;
    HIPCHK(hipMemcpyAsync(h.params.p, params, sizeof(float) * P, hipMemcpyHostToDevice, h.ctx.stream));
    HIPCHK(hipMemcpyAsync(h.x.p, x, sizeof(float) * 3 * n, hipMemcpyHostToDevice, h.ctx.stream));
    HIPCHK(hipMemcpyAsync(h.t.p, target, sizeof(float) * n, hipMemcpyHostToDevice, h.ctx.stream));
    rc = mlp_train_device(&h.ctx, (float *)h.params.p, (const float *)h.x.p, (const float *)h.t.p, n,
                          epochs, batch, lr, seed, loss_out);
    if (rc)
        return rc;
    HIPCHK(hipMemcpyAsync(params, h.params.p, sizeof(float) * P, hipMemcpyDeviceToHost, h.ctx.stream));
    HIPCHK(hipStreamSynchronize(h.ctx.stream));
    return TRUST_OK;
}
