// Launch-parameter autotuner for the resident graph.
//
// The reference takes its launch shape from the command line (threads,
// blocks, chunk: Without-graph-partition/mpi_main.cpp:78-81) and fixes the
// hub threshold (USE_CTA 100, TC.cu:14) and bucket sizes at compile time.
// Here the tuner times real counts of the resident graph over a grid of
// workgroup size, wave-table slots (the per-wave bucket array), hub
// threshold and chunk size by coordinate descent, and keeps a candidate only
// when it is faster by more than the noise margin.  Every candidate must
// reproduce the first count exactly.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <vector>

#include "engine.h"

namespace trust {
namespace {

struct Dim {
    int trust_params::*field;
    std::vector<int> values;
};

}  // namespace
}  // namespace trust

using namespace trust;

extern "C" int trust_ctx_autotune(trust_ctx *c, const trust_params *base, int reps,
                                  double max_seconds, trust_params *best, double *best_seconds,
                                  int *tried)
{
    clear_error();
    Ctx *ctx = reinterpret_cast<Ctx *>(c);
    if (!ctx || !base || !best)
        return fail(TRUST_ERR_ARG, "null argument");
    if (reps < 1)
        return fail(TRUST_ERR_ARG, "reps must be >= 1");
    HIPCHK(hipSetDevice(ctx->device));
    const auto t_start = std::chrono::steady_clock::now();
    auto elapsed = [&]() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    };
    uint64_t ref = 0;
    bool have_ref = false;
    int n_tried = 0;
    // min over reps of the device time of one count; ERR_ARG marks a
    // candidate the device cannot launch (e.g. wave tables beyond the LDS)
    auto measure = [&](const trust_params &q, double *sec) -> int {
        double s = 1e300;
        for (int i = 0; i < reps; i++) {
            trust_result r;
            int rc = count_resident(ctx, q, kModeCount, nullptr, nullptr, &r);
            if (rc)
                return rc;
            if (have_ref && r.triangles != ref)
                return fail(TRUST_ERR_HIP,
                            "autotune: count changed from %llu to %llu (threads %d, wave_slots "
                            "%d, hub_threshold %d, chunk %d)",
                            (unsigned long long)ref, (unsigned long long)r.triangles, q.threads,
                            q.wave_slots, q.hub_threshold, q.chunk);
            ref = r.triangles;
            have_ref = true;
            s = std::min(s, r.seconds);
        }
        n_tried++;
        *sec = s;
        return TRUST_OK;
    };
    trust_params cur = *base;
    double cur_s = 0;
    int rc = measure(cur, &cur_s);
    if (rc)
        return rc;
    const std::vector<Dim> dims = {
        {&trust_params::threads, {128, 256, 512, 1024}},
        {&trust_params::wave_slots, {128, 256, 512, 1024, 2048}},
        {&trust_params::hub_threshold, {16, 32, 64, 100, 200, 500, 2000}},
        {&trust_params::chunk, {0, 1, 2, 4, 8, 16, 32, 64}},
    };
    const double margin = 0.01;  // relative gain below which a change is noise
    bool out_of_time = false;
    for (int pass = 0; pass < 2 && !out_of_time; pass++) {
        bool changed = false;
        for (const Dim &d : dims) {
            trust_params dim_best = cur;
            double dim_best_s = cur_s;
            for (int v : d.values) {
                if (max_seconds > 0 && elapsed() > max_seconds) {
                    out_of_time = true;
                    break;
                }
                if (cur.*d.field == v)
                    continue;
                trust_params q = cur;
                q.*d.field = v;
                double s = 0;
                rc = measure(q, &s);
                if (rc == TRUST_ERR_ARG) {  // not launchable: skip
                    clear_error();
                    continue;
                }
                if (rc)
                    return rc;
                if (s < dim_best_s * (1.0 - margin)) {
                    dim_best = q;
                    dim_best_s = s;
                }
            }
            if (dim_best_s < cur_s) {
                changed = changed || memcmp(&dim_best, &cur, sizeof(cur)) != 0;
                cur = dim_best;
                cur_s = dim_best_s;
            }
            if (out_of_time)
                break;
        }
        if (!changed)
            break;
    }
    *best = cur;
    if (best_seconds)
        *best_seconds = cur_s;
    if (tried)
        *tried = n_tried;
    return TRUST_OK;
}
