#include "ctx.h"
namespace edhmc { struct HierModel {}; void hier_destroy(HierModel* h) { delete h; } }
using namespace edhmc;
extern "C" {
int edhmc_bind_hier(edhmc_ctx* ctx, const float*, const float*, const int64_t*, int32_t, int64_t, int32_t, void*) { return set_err(ctx, EDHMC_EUNSUPPORTED, "todo"); }
int edhmc_hier_logp_grad(edhmc_ctx* ctx, const float*, int32_t, float*, double*, void*) { return set_err(ctx, EDHMC_EUNSUPPORTED, "todo"); }
int edhmc_hier_set_state(edhmc_ctx* ctx, const float*, int32_t, void*) { return set_err(ctx, EDHMC_EUNSUPPORTED, "todo"); }
int edhmc_hier_transition(edhmc_ctx* ctx, float, int32_t, const float*, const double*, uint64_t, uint64_t, int32_t*, float*, int32_t*, void*) { return set_err(ctx, EDHMC_EUNSUPPORTED, "todo"); }
}
