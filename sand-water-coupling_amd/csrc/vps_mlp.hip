// vps_mlp.hip -- learned CG warm start (placeholder until the MFMA kernels
// land): reports that the network is unavailable.
#include "vps_internal.h"

namespace vps {
int mlp_create(vps_context*) { return VPS_OK; }
void mlp_destroy(vps_context*) {}
int mlp_warm_start(vps_context* c, float* x0) {
    return warm_start(c, VPS_WARM_NODE_KEYED, x0);
}
int mlp_after_solve(vps_context*) { return VPS_OK; }
}  // namespace vps

extern "C" {
int vps_mlp_set_training(vps_context* c, int32_t, int32_t, double, uint64_t) {
    if (c) c->err = "learned warm start not built";
    return VPS_ERR_STATE;
}
int64_t vps_mlp_parameter_count(void) { return 4801; }
int vps_mlp_get_parameters(vps_context* c, float*) {
    if (c) c->err = "learned warm start not built";
    return VPS_ERR_STATE;
}
int vps_mlp_set_parameters(vps_context* c, const float*) {
    if (c) c->err = "learned warm start not built";
    return VPS_ERR_STATE;
}
int vps_mlp_loss_and_grad(vps_context* c, int64_t, const float*, const float*,
                          double*, float*) {
    if (c) c->err = "learned warm start not built";
    return VPS_ERR_STATE;
}
int vps_mlp_forward(vps_context* c, int64_t, const float*, float*) {
    if (c) c->err = "learned warm start not built";
    return VPS_ERR_STATE;
}
int vps_mlp_get_features(vps_context* c, float*) {
    if (c) c->err = "learned warm start not built";
    return VPS_ERR_STATE;
}
}
