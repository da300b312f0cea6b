// kfc_apply for callers that cannot allocate device memory: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.
//
// The only device memory kfc_apply may grow is the chain path's pair of
// intermediate blocks (N x ncols scalars each, used when a kept term has two or more
// factors); the fused stencil and the dense kept-basis pairs need none.  Here those two
// blocks live in one file-scope device array of NOALLOC_SCRATCH_BYTES, which every GPU
// the library runs on holds until the process ends.  Its device address differs per
// GPU and is taken once per device, when a context is created on it
// (noalloc_setup, from kfc_ctx_create).  In host code `scratch` names a host-side
// stand-in, never the device memory: only the per-device address from
// hipGetSymbolAddress is handed to kernels.
//
// Every call on one GPU shares that GPU's array, and the array keeps the last call's
// contents; each chain launch writes every entry it later reads, so nothing needs
// clearing.  Calls that overlap on one GPU (different streams or host threads) race.
#include <mutex>

#include "kfc_internal.h"

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;   // 64 MiB per GPU
constexpr int MAX_DEVICES = 64;

std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};   // the array's device address on each GPU

}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

int noalloc_setup(kfc_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= MAX_DEVICES)
        return kfc_fail(ctx, KFC_EUNSUPPORTED, "kfc_apply_noalloc: device index above " + std::to_string(MAX_DEVICES));
    std::lock_guard<std::mutex> g(g_mu);
    if (g_scratch[ctx->device]) return KFC_OK;
    void* p = nullptr;
    // ctx->device is the current device here (kfc_ctx_create set it)
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return kfc_fail(ctx, KFC_EDEVICE, std::string("kfc_apply_noalloc scratch: hipGetSymbolAddress: ") +
                                              hipGetErrorString(e));
    g_scratch[ctx->device] = p;
    return KFC_OK;
}

extern "C" int kfc_apply_noalloc(kfc_op* op, const void* X, void* Y, int64_t ncols, int64_t ld, int32_t precision) {
    if (!op) return kfc_fail(nullptr, KFC_EARG, "kfc_apply_noalloc: null operator");
    if (precision == KFC_PREC_Z64) {
        if (!op->cplx)
            return kfc_fail(op->ctx, KFC_EUNSUPPORTED,
                            "kfc_apply_noalloc: complex blocks on a real operator take kfc_apply (it packs them)");
        precision = KFC_PREC_F64;
    }
    bool real_on_complex = false;
    int rc = check_apply_args(op, X, Y, ncols, ld, precision, real_on_complex, "kfc_apply_noalloc");
    if (rc != KFC_OK || ncols == 0) return rc;
    const size_t need = op_chain_scratch_bytes(op, ncols, precision);
    const size_t half = need / 2, padded = ((half + 255) & ~size_t(255)) + half;
    if (padded > NOALLOC_SCRATCH_BYTES)
        return kfc_fail(op->ctx, KFC_EARG,
                        "kfc_apply_noalloc: the chain path needs " + std::to_string(need) + " bytes of scratch for " +
                            std::to_string(ncols) + " columns, above the " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                            " the library holds per GPU; apply fewer columns per call or use kfc_apply");
    void* p;
    {
        std::lock_guard<std::mutex> g(g_mu);
        p = g_scratch[op->ctx->device];
    }
    if (!p) return kfc_fail(op->ctx, KFC_EDEVICE, "kfc_apply_noalloc: the scratch array was not set up on this device");
    return op_apply_fixed(op, X, Y, ncols, ld, precision, p, op->ctx->stream);
}
