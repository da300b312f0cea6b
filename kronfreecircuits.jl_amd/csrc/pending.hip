// Entry points whose device implementations are not in this build yet: they
// fail loudly with KFC_EUNSUPPORTED.
#include "kfc_internal.h"

extern "C" {
int kfc_batch_transmons(kfc_ctx* ctx, int64_t, const int64_t*, const int64_t*, const double*, int32_t, int32_t, double,
                        int32_t, double*, double*, kfc_stats*) {
    return kfc_fail(ctx, KFC_EUNSUPPORTED, "kfc_batch_transmons is not built yet");
}
}
