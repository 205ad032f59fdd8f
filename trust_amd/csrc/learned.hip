// Learned cost scheduler (placeholder until the MLP kernels land).
#include "engine.h"

namespace trust {

int learned_plan(Ctx *ctx, const trust_params &p, const LaunchPlan &L,
                 const unsigned char **tiers)
{
    (void)ctx;
    (void)p;
    (void)L;
    *tiers = nullptr;
    return fail(TRUST_ERR_ARG, "learned schedule not available in this build");
}

}  // namespace trust
