// psr_mlp.hip -- learned chemistry (MLP surrogate of step_evolve).
#include <hip/hip_runtime.h>

#include "psr_internal.h"

struct MlpState {};

void psr_mlp_free(psr_ctx *c) {
    delete c->mlp;
    c->mlp = nullptr;
}

int psr_mlp_update_dev(psr_ctx *c, int64_t, double) {
    PSR_FAIL(c, PSR_ERR_NOTSUP, "surrogate chemistry not built yet");
}

extern "C" int psr_mlp_features(const psr_step *, const psr_units *, float *) { return PSR_ERR_NOTSUP; }
extern "C" int psr_mlp_train(psr_ctx *c, const float *, const float *, int64_t, int32_t, int32_t, float,
                             uint64_t, float *) {
    PSR_FAIL(c, PSR_ERR_NOTSUP, "surrogate chemistry not built yet");
}
extern "C" int psr_mlp_forward(psr_ctx *c, const float *, int64_t, float *) {
    PSR_FAIL(c, PSR_ERR_NOTSUP, "surrogate chemistry not built yet");
}
extern "C" int64_t psr_mlp_param_count(void) { return 0; }
extern "C" int psr_mlp_get_params(psr_ctx *c, float *) { PSR_FAIL(c, PSR_ERR_NOTSUP, "not built"); }
extern "C" int psr_mlp_set_params(psr_ctx *c, const float *) { PSR_FAIL(c, PSR_ERR_NOTSUP, "not built"); }
extern "C" int psr_mlp_grad(psr_ctx *c, const float *, const float *, int64_t, float *, float *) {
    PSR_FAIL(c, PSR_ERR_NOTSUP, "not built");
}
extern "C" int psr_mlp_set_residual(psr_ctx *c, double tol) {
    if (!c || !(tol > 0)) return PSR_ERR_ARG;
    c->mlp_residual = tol;
    return PSR_OK;
}
