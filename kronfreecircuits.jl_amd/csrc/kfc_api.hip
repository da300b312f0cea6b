// The C ABI: context, operator construction (with the reference's diagonal
// fusion), application and the derived quantities the solvers need.
#include <algorithm>
#include <cmath>
#include <queue>

#include "blas.h"
#include "kfc_internal.h"

static thread_local std::string g_noctx_err;

int kfc_fail(kfc_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    else g_noctx_err = msg;
    return code;
}

int ctx_callback(kfc_ctx* ctx, const kfc_iter_info& info, const char* who) {
    if (!ctx->cb || ctx->cb(&info, ctx->cb_user) == 0) return KFC_OK;
    return kfc_fail(ctx, KFC_EABORT, std::string(who) + ": the callback stopped the solve at iteration " +
                                         std::to_string(info.iteration));
}

extern "C" {

int kfc_ctx_create(kfc_ctx** out, int device) {
    if (!out) return kfc_fail(nullptr, KFC_EARG, "kfc_ctx_create: null output pointer");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return kfc_fail(nullptr, KFC_EDEVICE, std::string("kfc_ctx_create: no HIP device: ") + hipGetErrorString(e));
    if (device < 0 || device >= n) return kfc_fail(nullptr, KFC_EARG, "kfc_ctx_create: device index out of range");
    auto* c = new kfc_ctx();
    c->device = device;
    e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return kfc_fail(nullptr, KFC_EDEVICE, std::string("kfc_ctx_create: ") + hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    // the per-device address of kfc_apply_noalloc's scratch array, taken with the device current
    const int rs = noalloc_setup(c);
    if (rs != KFC_OK) {
        g_noctx_err = c->err;
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rs;
    }
    *out = c;
    return KFC_OK;
}

int kfc_ctx_destroy(kfc_ctx* ctx) {
    if (!ctx) return KFC_OK;
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return KFC_OK;
}

const char* kfc_last_error(const kfc_ctx* ctx) { return ctx ? ctx->err.c_str() : g_noctx_err.c_str(); }

void* kfc_ctx_stream(kfc_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int kfc_ctx_set_callback(kfc_ctx* ctx, kfc_callback cb, void* user) {
    if (!ctx) return kfc_fail(nullptr, KFC_EARG, "kfc_ctx_set_callback: null context");
    ctx->cb = cb;
    ctx->cb_user = cb ? user : nullptr;
    return KFC_OK;
}

int kfc_sync(kfc_ctx* ctx) {
    if (!ctx) return kfc_fail(nullptr, KFC_EARG, "kfc_sync: null context");
    KFC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KFC_OK;
}

static int read_factor(kfc_ctx* ctx, const kfc_factor_desc& fd, int64_t dim, bool cplx_ok, Factor& f) {
    f.axis = fd.axis;
    f.kind = fd.kind;
    f.d = fd.d;
    if (fd.d != dim)
        return kfc_fail(ctx, KFC_EDIM, "KronFreeOperator: factor on axis " + std::to_string(fd.axis) + " has dimension " +
                                           std::to_string(fd.d) + ", expected " + std::to_string(dim));
    auto get = [&](const double* p, int64_t n, std::vector<cd>& out) -> int {
        out.resize(n);
        if (n > 0 && !p) return kfc_fail(ctx, KFC_EARG, "KronFreeOperator: null factor data");
        for (int64_t i = 0; i < n; ++i) {
            out[i] = fd.is_complex ? cd(p[2 * i], p[2 * i + 1]) : cd(p[i], 0.0);
            if (!cplx_ok && out[i].imag() != 0)
                return kfc_fail(ctx, KFC_EARG, "KronFreeOperator: complex factor data on a real (KFC_F64) operator");
        }
        return KFC_OK;
    };
    int rc = KFC_OK;
    switch (fd.kind) {
        case KFC_DIAG: rc = get(fd.diag, fd.d, f.diag); break;
        case KFC_DENSE: rc = get(fd.matrix, fd.d * fd.d, f.matrix); break;
        case KFC_BANDED:
            if (fd.nbands < 0) return kfc_fail(ctx, KFC_EARG, "BandedOp: negative band count");
            if (fd.nbands > 0 && !fd.offsets) return kfc_fail(ctx, KFC_EARG, "BandedOp: null offsets");
            f.offsets.assign(fd.offsets, fd.offsets + fd.nbands);
            rc = get(fd.weights, fd.d * fd.nbands, f.weights);
            break;
        case KFC_SHIFT:
            f.shift = fd.shift;
            f.factor = cd(fd.factor_re, fd.is_complex ? fd.factor_im : 0.0);
            if (!cplx_ok && f.factor.imag() != 0)
                return kfc_fail(ctx, KFC_EARG, "KronFreeOperator: complex shift factor on a real operator");
            break;
        default: return kfc_fail(ctx, KFC_EARG, "KronFreeOperator: unknown factor kind");
    }
    return rc;
}

int kfc_op_create(kfc_ctx* ctx, const kfc_op_desc* desc, kfc_op** out) {
    if (!ctx || !desc || !out) return kfc_fail(ctx, KFC_EARG, "kfc_op_create: null argument");
    if (desc->ndims < 1 || desc->ndims > 16 || !desc->dims)
        return kfc_fail(ctx, KFC_EARG, "KronFreeOperator: ndims must lie in 1:16");
    if (desc->dtype != KFC_F64 && desc->dtype != KFC_C128)
        return kfc_fail(ctx, KFC_EARG, "KronFreeOperator: dtype must be KFC_F64 or KFC_C128");
    if (desc->block_capacity < 1) return kfc_fail(ctx, KFC_EARG, "KronFreeOperator: block_capacity must be positive");
    auto op = std::make_unique<kfc_op>();
    op->ctx = ctx;
    op->ndims = desc->ndims;
    op->N = 1;
    for (int a = 0; a < desc->ndims; ++a) {
        if (desc->dims[a] < 1)
            return kfc_fail(ctx, KFC_EARG, "KronFreeOperator: dims must be positive");
        op->dims.push_back(desc->dims[a]);
        op->N *= desc->dims[a];
    }
    op->cplx = desc->dtype == KFC_C128;
    op->constant = cd(desc->constant_re, op->cplx ? desc->constant_im : 0.0);
    if (!op->cplx && desc->constant_im != 0)
        return kfc_fail(ctx, KFC_EARG, "KronFreeOperator: complex constant on a real operator");
    op->capacity = desc->block_capacity;
    op->hermitian = desc->hermitian != 0;
    if (op->constant != cd(0)) op->diagterms.push_back(Term{op->constant, {}});
    for (int t = 0; t < desc->nterms; ++t) {
        const kfc_term_desc& td = desc->terms[t];
        Term term;
        term.coeff = cd(td.coeff_re, op->cplx ? td.coeff_im : 0.0);
        if (!op->cplx && td.coeff_im != 0)
            return kfc_fail(ctx, KFC_EARG, "KronFreeOperator: complex coefficient on a real operator");
        int prev = -1;
        bool all_diag = true;
        for (int k = 0; k < td.nfactors; ++k) {
            const kfc_factor_desc& fd = td.factors[k];
            if (fd.axis < 0 || fd.axis >= op->ndims)
                return kfc_fail(ctx, KFC_EARG, "KronFreeOperator: factor axis " + std::to_string(fd.axis) + " outside 0:" +
                                                   std::to_string(op->ndims - 1));
            if (fd.axis <= prev)
                return kfc_fail(ctx, KFC_EARG, "AxisTerm: factor axes must be strictly increasing");
            prev = fd.axis;
            Factor f;
            int rc = read_factor(ctx, fd, op->dims[fd.axis], op->cplx, f);
            if (rc != KFC_OK) return rc;
            all_diag &= f.kind == KFC_DIAG;
            term.factors.push_back(std::move(f));
        }
        if (all_diag) {
            if (!term.factors.empty()) op->diagterms.push_back(std::move(term));
            else op->constant += term.coeff, op->diagterms.push_back(std::move(term));
        } else {
            op->terms.push_back(std::move(term));
        }
    }
    int rc = op_build_plan(op.get());
    if (rc == KFC_OK) rc = op_build_pairs(op.get());
    // the FP64 tables go to the device with the operator, as the reference adapts its
    // data to `on` at construction (src/apply.jl:248-250); the FP32 ones follow on the
    // first FP32 apply
    if (rc == KFC_OK) rc = op_ensure_device(op.get(), KFC_PREC_F64);
    if (rc != KFC_OK) return rc;
    *out = op.release();
    return KFC_OK;
}

int kfc_op_destroy(kfc_op* op) {
    if (!op) return KFC_OK;
    if (op->ctx) (void)hipStreamSynchronize(op->ctx->stream);
    delete op;
    return KFC_OK;
}

int64_t kfc_op_size(const kfc_op* op) { return op ? op->N : 0; }
int kfc_op_fused(const kfc_op* op) { return op && op->plan.ok ? 1 : 0; }

int kfc_op_work(const kfc_op* op, double* bytes, double* flops) {
    if (!op || !bytes || !flops) return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_op_work: null argument");
    const double N = double(op->N), s = op->cplx ? 16.0 : 8.0;
    double f = 0;
    if (op->plan.ok) {
        int used = 0;
        for (int p = 0; p < 27; ++p) used += op->plan.pos[p].used ? 1 : 0;
        f = 2.0 * N * (used + 1);   // one multiply-add per stencil position and the diagonal
        *bytes = 2.0 * N * s + (op->plan.need_full_diag ? N * s : 0.0);
    } else if (op->pairs.ok) {
        for (const PairGroup& g : op->pairs.groups) f += double(g.c.size()) * 2.0 * N * double(g.dp + g.dq);
        f += 2.0 * N;   // the diagonal
        *bytes = 3.0 * N * s;   // x, y and the stored diagonal
    } else {
        for (const Term& t : op->terms) f += 2.0 * N * double(t.factors.size());
        *bytes = (2.0 + 2.0 * double(op->terms.size())) * N * s;
    }
    *flops = op->cplx ? 4.0 * f : f;
    return KFC_OK;
}

int kfc_op_stencil_get(const kfc_op* op, int32_t* variant, int32_t* march) {
    if (!op || !variant || !march) return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_op_stencil_get: null argument");
    *variant = op->stencil_variant;
    *march = op->stencil_march;
    return KFC_OK;
}

int kfc_op_stencil_set(kfc_op* op, int32_t variant, int32_t march) {
    if (!op) return kfc_fail(nullptr, KFC_EARG, "kfc_op_stencil_set: null operator");
    if (variant < -1 || variant > 5) return kfc_fail(op->ctx, KFC_EARG, "kfc_op_stencil_set: variant must lie in -1:5");
    if (march < 0 || march > 4096) return kfc_fail(op->ctx, KFC_EARG, "kfc_op_stencil_set: march must lie in 0:4096");
    if (variant < 0) {
        op->tuned_cols[0] = op->tuned_cols[1] = op->tuned_cols[2] = -1;
        op->variant_z = -1;
        op->stencil_pinned = false;
        return KFC_OK;
    }
    op->stencil_variant = variant;
    op->stencil_march = march;
    op->stencil_pinned = true;
    return KFC_OK;
}

}  // extern "C"

namespace {
template <class T>
__global__ void k_fill_scratch(T* p, size_t n, T v) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}

template <class T>
hipError_t fill_scratch(DevBuf& b, T v, hipStream_t st) {
    if (!b.p) return hipSuccess;
    const size_t n = b.bytes / sizeof(T);
    hipLaunchKernelGGL(k_fill_scratch<T>, dim3(unsigned(std::min<size_t>(4096, (n + 255) / 256 + 1))), dim3(256), 0, st,
                       b.as<T>(), n, v);
    return hipGetLastError();
}
}  // namespace

extern "C" {

int kfc_op_fill_scratch(kfc_op* op, double value) {
    if (!op) return kfc_fail(nullptr, KFC_EARG, "kfc_op_fill_scratch: null operator");
    hipStream_t st = op->ctx->stream;
    // FP64 tables' scratch holds doubles (complex: pairs of them), FP32 tables' floats
    KFC_HIP(op->ctx, fill_scratch<double>(op->dd[KFC_PREC_F64].buf1, value, st));
    KFC_HIP(op->ctx, fill_scratch<double>(op->dd[KFC_PREC_F64].buf2, value, st));
    KFC_HIP(op->ctx, fill_scratch<float>(op->dd[KFC_PREC_F32].buf1, float(value), st));
    KFC_HIP(op->ctx, fill_scratch<float>(op->dd[KFC_PREC_F32].buf2, float(value), st));
    KFC_HIP(op->ctx, fill_scratch<double>(op->ddz.buf1, value, st));
    KFC_HIP(op->ctx, fill_scratch<double>(op->ddz.buf2, value, st));
    KFC_HIP(op->ctx, fill_scratch<double>(op->dotpart, value, st));
    for (DevBuf& b : op->evo) KFC_HIP(op->ctx, fill_scratch<double>(b, value, st));
    KFC_HIP(op->ctx, hipStreamSynchronize(st));
    return KFC_OK;
}

}  // extern "C"

int check_apply_args(kfc_op* op, const void* X, const void* Y, int64_t ncols, int64_t ld, int32_t& precision,
                     bool& real_on_complex, const char* who) {
    if (precision != KFC_PREC_F64 && precision != KFC_PREC_F32 && precision != KFC_PREC_Z64)
        return kfc_fail(op->ctx, KFC_EARG, std::string(who) + ": unknown precision");
    // complex FP64 blocks: a complex operator's own FP64 apply, or a real one's
    // _mul_real_operator! (src/apply.jl:397-407)
    real_on_complex = precision == KFC_PREC_Z64 && !op->cplx;
    if (precision == KFC_PREC_Z64 && op->cplx) precision = KFC_PREC_F64;
    if (ncols < 0) return kfc_fail(op->ctx, KFC_EARG, std::string(who) + ": negative column count");
    if (ld < op->N)
        return kfc_fail(op->ctx, KFC_EDIM, "KronFreeOperator of size (" + std::to_string(op->N) + ", " +
                                               std::to_string(op->N) + ") applied to x with leading dimension " +
                                               std::to_string(ld));
    if (ncols == 0) return KFC_OK;
    if (!X || !Y) return kfc_fail(op->ctx, KFC_EARG, std::string(who) + ": null buffer");
    const size_t es = (op->cplx || real_on_complex ? 2 : 1) * (precision == KFC_PREC_F32 ? 4 : 8);
    const char* xb = static_cast<const char*>(X);
    const char* yb = static_cast<const char*>(Y);
    const size_t span = size_t(((ncols - 1) * ld + op->N)) * es;
    if (xb < yb + span && yb < xb + span) return kfc_fail(op->ctx, KFC_EARG, "mul!: y must not alias x");
    return KFC_OK;
}

extern "C" {

int kfc_apply(kfc_op* op, const void* X, void* Y, int64_t ncols, int64_t ld, int32_t precision) {
    if (!op) return kfc_fail(nullptr, KFC_EARG, "kfc_apply: null operator");
    bool real_on_complex = false;
    int rc = check_apply_args(op, X, Y, ncols, ld, precision, real_on_complex, "kfc_apply");
    if (rc != KFC_OK || ncols == 0) return rc;
    if (real_on_complex) return op_apply_real_on_complex(op, X, Y, ncols, ld);
    return op_apply(op, X, Y, ncols, ld, precision, op->ctx->stream);
}

int kfc_apply_cheb_dots(kfc_op* op, const void* B, const void* A, void* C, double a, double b, double g,
                        int32_t precision, double* dots, int32_t* fused) {
    if (!op) return kfc_fail(nullptr, KFC_EARG, "kfc_apply_cheb_dots: null operator");
    if (op->cplx) return kfc_fail(op->ctx, KFC_EUNSUPPORTED, "kfc_apply_cheb_dots: real operators only");
    if (precision != KFC_PREC_F64 && precision != KFC_PREC_F32)
        return kfc_fail(op->ctx, KFC_EARG, "kfc_apply_cheb_dots: unknown precision");
    if (!B || !C || !dots) return kfc_fail(op->ctx, KFC_EARG, "kfc_apply_cheb_dots: null buffer");
    const size_t span = size_t(op->N) * (precision == KFC_PREC_F64 ? 8 : 4);
    const char* bb = static_cast<const char*>(B);
    const char* cb = static_cast<const char*>(C);
    if (bb < cb + span && cb < bb + span) return kfc_fail(op->ctx, KFC_EARG, "mul!: y must not alias x");
    Work w;
    w.ctx = op->ctx;
    KFC_HIP(op->ctx, w.small2.ensure(2 * sizeof(double)));
    bool f = false;
    int rc = op_apply_cheb_dots(op, B, A, C, op->N, precision, a, b, g, w.small2.as<double>(), f, op->ctx->stream);
    if (rc != KFC_OK) return rc;
    if (f) {
        KFC_HIP(op->ctx, hipMemcpyAsync(dots, w.small2.p, 2 * sizeof(double), hipMemcpyDeviceToHost, op->ctx->stream));
        KFC_HIP(op->ctx, hipStreamSynchronize(op->ctx->stream));
    } else {
        std::vector<cd> g0, g1;
        if (precision == KFC_PREC_F64) {
            rc = blas_gram<double>(w, static_cast<const double*>(B), op->N, 1, static_cast<const double*>(B), op->N, 1,
                                   op->N, g0);
            if (rc == KFC_OK)
                rc = blas_gram<double>(w, static_cast<const double*>(B), op->N, 1, static_cast<const double*>(C),
                                       op->N, 1, op->N, g1);
        } else {
            rc = blas_gram<float>(w, static_cast<const float*>(B), op->N, 1, static_cast<const float*>(B), op->N, 1,
                                  op->N, g0);
            if (rc == KFC_OK)
                rc = blas_gram<float>(w, static_cast<const float*>(B), op->N, 1, static_cast<const float*>(C), op->N,
                                      1, op->N, g1);
        }
        if (rc != KFC_OK) return rc;
        dots[0] = g0[0].real();
        dots[1] = g1[0].real();
    }
    if (fused) *fused = f ? 1 : 0;
    return KFC_OK;
}

}  // extern "C"

// ---- derived quantities -------------------------------------------------------------------

static void dense_of(const Factor& f, std::vector<cd>& M) {
    // dense(op) (src/localops.jl:115-130)
    const int64_t d = f.d;
    M.assign(size_t(d) * d, cd(0));
    if (f.kind == KFC_DENSE) M = f.matrix;
    else if (f.kind == KFC_DIAG)
        for (int64_t i = 0; i < d; ++i) M[i + i * d] = f.diag[i];
    else if (f.kind == KFC_BANDED) {
        for (size_t b = 0; b < f.offsets.size(); ++b)
            for (int64_t i = 0; i < d; ++i) {
                const int64_t j = i + f.offsets[b];
                if (j >= 0 && j < d) M[i + j * d] += f.weights[i + b * d];
            }
    } else {
        for (int64_t i = 0; i < d; ++i) {
            const int64_t j = i + f.shift;
            if (j >= 0 && j < d) M[j + i * d] = f.factor;
        }
    }
}

static double infnorm_of(const Factor& f) {
    // infnorm(op) (src/localops.jl:174-188)
    const int64_t d = f.d;
    if (f.kind == KFC_SHIFT) return std::llabs(f.shift) < d ? std::abs(f.factor) : 0.0;
    if (f.kind == KFC_DIAG) {
        double m = 0;
        for (const cd& x : f.diag) m = std::max(m, std::abs(x));
        return m;
    }
    double m = 0;
    for (int64_t i = 0; i < d; ++i) {
        double r = 0;
        if (f.kind == KFC_DENSE)
            for (int64_t j = 0; j < d; ++j) r += std::abs(f.matrix[i + j * d]);
        else
            for (size_t b = 0; b < f.offsets.size(); ++b) {
                const int64_t j = i + f.offsets[b];
                if (j >= 0 && j < d) r += std::abs(f.weights[i + b * d]);
            }
        m = std::max(m, r);
    }
    return m;
}

static bool separable_diag(const kfc_op* op) {
    for (const Term& t : op->diagterms)
        if (t.factors.size() > 1) return false;
    return true;
}

extern "C" int kfc_op_gershgorin(kfc_op* op, double* out) {
    // gershgorin_bound(H) (src/apply.jl:586-593)
    if (!op || !out) return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_op_gershgorin: null argument");
    if (!op->have_gershgorin) {
        double off = 0;
        for (const Term& t : op->terms) {
            double p = std::abs(t.coeff);
            for (const Factor& f : t.factors) p *= infnorm_of(f);
            off += p;
        }
        double dmax, dmin, imax = 0;
        if (!op->cplx && separable_diag(op)) {
            // extrema over a sum of per-axis vectors are the sums of the per-axis extrema
            std::vector<std::vector<double>> ax(op->ndims);
            for (int a = 0; a < op->ndims; ++a) ax[a].assign(op->dims[a], 0.0);
            double c = 0;
            for (const Term& t : op->diagterms) {
                if (t.factors.empty()) { c += t.coeff.real(); continue; }
                const Factor& f = t.factors[0];
                for (int64_t i = 0; i < f.d; ++i) ax[f.axis][i] += (t.coeff * f.diag[i]).real();
            }
            dmax = dmin = c;
            for (auto& v : ax) {
                dmax += *std::max_element(v.begin(), v.end());
                dmin += *std::min_element(v.begin(), v.end());
            }
        } else {
            // the fused diagonal's range, reduced on the device
            const int rc = op_fused_diag_range_dev(op, &dmax, &dmin, &imax);
            if (rc != KFC_OK) return rc;
        }
        op->gershgorin_lo = dmin - imax - off;
        op->gershgorin = dmax + imax + off;
        op->have_gershgorin = true;
    }
    *out = op->gershgorin;
    return KFC_OK;
}

int op_spectral_interval(kfc_op* op, double* lo, double* hi) {
    // Gershgorin discs: every eigenvalue lies in [min d_i - r, max d_i + r]
    int rc = kfc_op_gershgorin(op, hi);
    if (rc == KFC_OK) *lo = op->gershgorin_lo;
    return rc;
}

std::vector<cd> kept_factor_diagonal(const Factor& f) {
    std::vector<cd> M, dd(f.d);
    if (f.kind == KFC_DIAG) return f.diag;
    if (f.kind == KFC_DENSE) {
        for (int64_t i = 0; i < f.d; ++i) dd[i] = f.matrix[i + i * f.d];
        return dd;
    }
    dense_of(f, M);
    for (int64_t i = 0; i < f.d; ++i) dd[i] = M[i + i * f.d];
    return dd;
}

int op_full_diagonal(kfc_op* op, std::vector<double>& D) {
    // _full_diagonal(H) (src/apply.jl:459-478): fused diagonal plus the kept terms' diagonals
    op_host_diag(op);
    D.resize(op->N);
    for (int64_t i = 0; i < op->N; ++i) D[i] = op->hdiag[i].real();
    std::vector<int64_t> stride(op->ndims, 1);
    for (int a = 1; a < op->ndims; ++a) stride[a] = stride[a - 1] * op->dims[a - 1];
    for (const Term& t : op->terms) {
        std::vector<std::vector<cd>> dv;
        bool zero = false;
        for (const Factor& f : t.factors) {
            std::vector<cd> dd = kept_factor_diagonal(f);
            bool allz = true;
            for (const cd& x : dd) allz &= x == cd(0);
            zero |= allz;
            dv.push_back(std::move(dd));
        }
        if (zero) continue;
        for (int64_t idx = 0; idx < op->N; ++idx) {
            cd v = t.coeff;
            for (size_t k = 0; k < t.factors.size(); ++k) {
                const int a = t.factors[k].axis;
                v *= dv[k][(idx / stride[a]) % op->dims[a]];
            }
            D[idx] += v.real();
        }
    }
    return KFC_OK;
}

extern "C" int kfc_op_diagonal(kfc_op* op, double* out) {
    if (!op || !out) return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_op_diagonal: null argument");
    std::vector<double> D;
    int rc = op_full_diagonal(op, D);
    if (rc != KFC_OK) return rc;
    std::copy(D.begin(), D.end(), out);
    return KFC_OK;
}

// separable_eigenbases(H) (src/apply.jl:606-642): per axis, the eigendecomposition of the
// sum of single-axis terms (kept and fused), the constant added to axis 0.
int op_separable(kfc_op* op, std::vector<std::vector<cd>>& U, std::vector<std::vector<double>>& lam,
                 std::vector<bool>& ident) {
    if (op->have_separable) {
        U = op->sep_U;
        lam = op->sep_lam;
        ident = op->sep_ident;
        return KFC_OK;
    }
    const int K = op->ndims;
    std::vector<std::vector<cd>> A(K);
    for (int a = 0; a < K; ++a) A[a].assign(size_t(op->dims[a]) * op->dims[a], cd(0));
    auto add = [&](const Term& t) {
        if (t.factors.size() == 1) {
            const Factor& f = t.factors[0];
            std::vector<cd> M;
            dense_of(f, M);
            for (size_t i = 0; i < M.size(); ++i) A[f.axis][i] += t.coeff * M[i];
        } else if (t.factors.empty()) {
            for (int64_t i = 0; i < op->dims[0]; ++i) A[0][i + i * op->dims[0]] += t.coeff;
        }
    };
    for (const Term& t : op->terms) add(t);
    for (const Term& t : op->diagterms) add(t);
    U.assign(K, {});
    lam.assign(K, {});
    ident.assign(K, false);
    for (int a = 0; a < K; ++a) {
        const int64_t d = op->dims[a];
        std::vector<cd>& M = A[a];
        double asym = 0, nrm = 0;
        bool isdiag = true;
        for (int64_t j = 0; j < d; ++j)
            for (int64_t i = 0; i < d; ++i) {
                asym += std::norm(M[i + j * d] - std::conj(M[j + i * d]));
                nrm += std::norm(M[i + j * d]);
                if (i != j && M[i + j * d] != cd(0)) isdiag = false;
            }
        if (std::sqrt(asym) > std::sqrt(2.220446049250313e-16) * std::sqrt(nrm))
            return kfc_fail(op->ctx, KFC_EARG, "separable_eigenbases: the single-axis terms on axis " + std::to_string(a) +
                                                   " sum to a non-Hermitian matrix, so the operator is not Hermitian");
        if (isdiag) {
            U[a].assign(size_t(d) * d, cd(0));
            for (int64_t i = 0; i < d; ++i) U[a][i + i * d] = 1;
            lam[a].resize(d);
            for (int64_t i = 0; i < d; ++i) lam[a][i] = M[i + i * d].real();
            ident[a] = true;
        } else {
            std::vector<cd> Hm(size_t(d) * d);
            for (int64_t j = 0; j < d; ++j)
                for (int64_t i = 0; i < d; ++i) Hm[i + j * d] = 0.5 * (M[i + j * d] + std::conj(M[j + i * d]));
            herm_eig(int(d), Hm, lam[a], U[a]);
        }
    }
    op->sep_U = U;
    op->sep_lam = lam;
    op->sep_ident = ident;
    op->have_separable = true;
    return KFC_OK;
}

// product_seed(H, k) (src/apply.jl:655-685) into host X[N x k] (complex or real layout).
int op_product_seed(kfc_op* op, int64_t k, std::vector<cd>& X) {
    const int64_t N = op->N;
    if (k < 1 || k > N) return kfc_fail(op->ctx, KFC_EARG, "product_seed: k must lie in 1:" + std::to_string(N));
    std::vector<std::vector<cd>> U;
    std::vector<std::vector<double>> lam;
    std::vector<bool> ident;
    int rc = op_separable(op, U, lam, ident);
    if (rc != KFC_OK) return rc;
    X.assign(size_t(N) * k, cd(0));
    bool all_ident = std::all_of(ident.begin(), ident.end(), [](bool b) { return b; });
    // the k smallest values with the smallest index first among ties (a bounded max-heap),
    // ordered as partialsortperm's isless orders them
    auto k_smallest = [&](auto value_of) {
        std::priority_queue<std::pair<uint64_t, int64_t>> heap;
        for (int64_t i = 0; i < N; ++i) {
            const std::pair<uint64_t, int64_t> p{seed_order_key(value_of(i)), i};
            if ((int64_t)heap.size() < k) heap.push(p);
            else if (p < heap.top()) {
                heap.pop();
                heap.push(p);
            }
        }
        std::vector<std::pair<uint64_t, int64_t>> out;
        while (!heap.empty()) { out.push_back(heap.top()); heap.pop(); }
        std::sort(out.begin(), out.end());
        return out;
    };
    std::vector<int64_t> stride(op->ndims, 1);
    for (int a = 1; a < op->ndims; ++a) stride[a] = stride[a - 1] * op->dims[a - 1];
    if (all_ident) {
        std::vector<double> D;
        op_full_diagonal(op, D);
        auto sel = k_smallest([&](int64_t i) { return D[i]; });
        for (int64_t j = 0; j < k; ++j) X[sel[j].second + j * N] = 1;
        return KFC_OK;
    }
    auto sel = k_smallest([&](int64_t i) {
        double s = 0;
        for (int a = 0; a < op->ndims; ++a) s += lam[a][(i / stride[a]) % op->dims[a]];
        return s;
    });
    for (int64_t j = 0; j < k; ++j) {
        std::vector<int64_t> idx(op->ndims);
        for (int a = 0; a < op->ndims; ++a) idx[a] = (sel[j].second / stride[a]) % op->dims[a];
        for (int64_t t = 0; t < N; ++t) {
            cd v = 1;
            for (int a = 0; a < op->ndims; ++a) {
                const int64_t d = op->dims[a];
                v *= U[a][(t / stride[a]) % d + idx[a] * d];
            }
            X[t + j * N] = v;
        }
    }
    return KFC_OK;
}

extern "C" int kfc_op_product_seed(kfc_op* op, int64_t k, double* Xout) {
    if (!op || !Xout) return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_op_product_seed: null argument");
    std::vector<cd> X;
    int rc = op_product_seed(op, k, X);
    if (rc != KFC_OK) return rc;
    for (size_t i = 0; i < X.size(); ++i) {
        if (op->cplx) {
            Xout[2 * i] = X[i].real();
            Xout[2 * i + 1] = X[i].imag();
        } else {
            Xout[i] = X[i].real();
        }
    }
    return KFC_OK;
}

