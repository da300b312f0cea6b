// Entry points whose device implementations are not in this build yet: they
// fail loudly with KFC_EUNSUPPORTED.
#include "kfc_internal.h"

extern "C" {
int kfc_batch_transmons(kfc_ctx* ctx, int64_t, const int64_t*, const int64_t*, const double*, int32_t, int32_t, double,
                        int32_t, double*, double*, kfc_stats*) {
    return kfc_fail(ctx, KFC_EUNSUPPORTED, "kfc_batch_transmons is not built yet");
}
int kfc_mlp_create(kfc_ctx* ctx, const int32_t*, int32_t, uint64_t, kfc_mlp**) {
    return kfc_fail(ctx, KFC_EUNSUPPORTED, "kfc_mlp is not built yet");
}
int kfc_mlp_destroy(kfc_mlp*) { return KFC_OK; }
int kfc_mlp_set_params(kfc_mlp*, const float*) { return KFC_EUNSUPPORTED; }
int kfc_mlp_get_params(kfc_mlp*, float*) { return KFC_EUNSUPPORTED; }
int64_t kfc_mlp_num_params(const kfc_mlp*) { return 0; }
int kfc_mlp_forward(kfc_mlp*, const float*, int64_t, float*) { return KFC_EUNSUPPORTED; }
int kfc_mlp_backward(kfc_mlp*, const float*, const float*, int64_t, float*) { return KFC_EUNSUPPORTED; }
int kfc_mlp_train(kfc_mlp*, const float*, const float*, int64_t, int64_t, int64_t, float, uint64_t, float*) {
    return KFC_EUNSUPPORTED;
}
}
