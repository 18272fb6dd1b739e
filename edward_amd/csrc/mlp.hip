#include "ctx.h"
namespace edhmc { struct Surrogate {}; void surrogate_destroy(Surrogate* s) { delete s; } }
using namespace edhmc;
extern "C" {
int edhmc_sur_create(edhmc_ctx* ctx, int32_t, int32_t, int32_t) { return set_err(ctx, EDHMC_EUNSUPPORTED, "todo"); }
int64_t edhmc_sur_num_params(edhmc_ctx*) { return 0; }
int edhmc_sur_set_params(edhmc_ctx* ctx, const float*, void*) { return set_err(ctx, EDHMC_EUNSUPPORTED, "todo"); }
int edhmc_sur_get_params(edhmc_ctx* ctx, float*, void*) { return set_err(ctx, EDHMC_EUNSUPPORTED, "todo"); }
int edhmc_sur_set_norm(edhmc_ctx* ctx, const float*, const float*, const float*, const float*, void*) { return set_err(ctx, EDHMC_EUNSUPPORTED, "todo"); }
int edhmc_sur_forward(edhmc_ctx* ctx, const float*, int32_t, float*, void*) { return set_err(ctx, EDHMC_EUNSUPPORTED, "todo"); }
int edhmc_sur_loss_grad(edhmc_ctx* ctx, const float*, const float*, int32_t, double*, float*, void*) { return set_err(ctx, EDHMC_EUNSUPPORTED, "todo"); }
int edhmc_sur_train_step(edhmc_ctx* ctx, const float*, const float*, int32_t, float, float, float, float, int32_t, double*, void*) { return set_err(ctx, EDHMC_EUNSUPPORTED, "todo"); }
int edhmc_transition_surrogate(edhmc_ctx* ctx, float, int32_t, const float*, const double*, uint64_t, uint64_t, int32_t*, float*, int32_t*, void*) { return set_err(ctx, EDHMC_EUNSUPPORTED, "todo"); }
}
