// The linear solve on the orthogonal complement of a node's known states, the
// Sternheimer step of the derivatives: _complement_solve (src/resolvent.jl:146-256).
// X[:, j] solves Q (H - s_j) Q x = Q B[:, j] with x orthogonal to the basis Phi
// (Q = 1 - Phi Phi^H) by preconditioned conjugate gradients on all columns in lockstep,
// each with its own scalars and the Jacobi preconditioner 1 / hypot(D - s_j, eta).
// Every block operation runs on the device (the operator apply, the projections as
// FP64-MFMA Grams and tall-skinny products, the per-column updates); the host keeps the
// per-column scalars and the stopping rules of the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "blas.h"
#include "kfc_internal.h"

namespace {

constexpr int CS_MAXC = 64;         // columns per lockstep solve
constexpr double CS_FLOOR_CAP = 1000;   // src/resolvent.jl:129, units of eps |R_j|
constexpr int CS_FLOOR_PATIENCE = 5;    // src/resolvent.jl:131

struct ColScal {
    double v[CS_MAXC];
};

// X[:, j] += a_j P[:, j];  R[:, j] -= a_j AP[:, j]
template <class S>
__global__ __launch_bounds__(256) void k_cs_step(S* __restrict__ X, S* __restrict__ R, const S* __restrict__ P,
                                                 const S* __restrict__ AP, int64_t N, int n, ColScal a) {
    const int64_t tot = N * n;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < tot; e += int64_t(gridDim.x) * 256) {
        const int j = int(e / N);
        const double aj = a.v[j];
        if (aj == 0.0) continue;
        X[e] = s_add(X[e], s_scale(aj, P[e]));
        R[e] = s_add(R[e], s_scale(-aj, AP[e]));
    }
}

// P[:, j] = Z[:, j] + b_j P[:, j]
template <class S>
__global__ __launch_bounds__(256) void k_cs_dir(S* __restrict__ P, const S* __restrict__ Z, int64_t N, int n,
                                                ColScal b) {
    const int64_t tot = N * n;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < tot; e += int64_t(gridDim.x) * 256) {
        const int j = int(e / N);
        P[e] = s_add(Z[e], s_scale(b.v[j], P[e]));
    }
}

// Z[:, j] = R[:, j] / hypot(D - s_j, eta)  (the Jacobi preconditioner, src/resolvent.jl:159)
template <class S>
__global__ __launch_bounds__(256) void k_cs_prec(S* __restrict__ Z, const S* __restrict__ R,
                                                 const double* __restrict__ D, int64_t N, int n, ColScal s,
                                                 double eta) {
    const int64_t tot = N * n;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < tot; e += int64_t(gridDim.x) * 256) {
        const int j = int(e / N);
        const int64_t i = e - int64_t(j) * N;
        const double f = 1.0 / hypot(D[i] - s.v[j], eta);
        Z[e] = s_scale(typename RealOf<S>::type(f), R[e]);
    }
}

int cs_grid(int64_t n) { return int(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192))); }

template <class S>
struct CsWork {
    kfc_op* op;
    Work w;
    const S* Phi;
    int k;
    int64_t N;
    int n;
    std::vector<cd> C, G;
    // Y <- Y - Phi (Phi^H Y)
    int project(S* Y) {
        if (k == 0) return KFC_OK;
        int rc = blas_gram<S>(w, Phi, N, k, Y, N, n, N, C);
        if (rc != KFC_OK) return rc;
        for (auto& c : C) c = -c;
        return blas_gemm<S>(w, Phi, N, k, C, n, Y, N, N, 1.0);
    }
    // Re <A_j, B_j>
    int colredot(const S* A, const S* B, std::vector<double>& out) {
        int rc = blas_gram<S>(w, A, N, n, B, N, n, N, G);
        if (rc != KFC_OK) return rc;
        out.resize(n);
        for (int j = 0; j < n; ++j) out[j] = G[j + size_t(j) * n].real();
        return KFC_OK;
    }
};

template <class S>
int complement_t(kfc_op* op, const double* D, const S* Phi, int k, const S* B, int n, const double* shifts_in,
                 double tol, int64_t maxiter, double eta, const double* scale_in, S* X, int64_t* iters_out,
                 double* residuals, int32_t* flags, double* floor_residuals) {
    const int64_t N = op->N;
    hipStream_t st = op->ctx->stream;
    CsWork<S> cw{op, Work(), Phi, k, N, n, {}, {}};
    cw.w.ctx = op->ctx;
    Work& w = cw.w;
    const std::vector<double> shifts(shifts_in, shifts_in + n);
    DevBuf bQB, bR, bZ, bP, bAP, bT;
    for (DevBuf* b : {&bQB, &bR, &bZ, &bP, &bAP, &bT}) KFC_HIP(op->ctx, b->ensure(sizeof(S) * size_t(N) * n));
    S* QB = bQB.as<S>();
    S* R = bR.as<S>();
    S* Z = bZ.as<S>();
    S* P = bP.as<S>();
    S* AP = bAP.as<S>();
    S* T = bT.as<S>();
    const int nb = cs_grid(N * n);
    std::vector<double> scale;
    int rc = KFC_OK;
    if (scale_in) scale.assign(scale_in, scale_in + n);
    else if ((rc = blas_colnorms<S>(w, B, N, n, N, scale)) != KFC_OK) return rc;
    const double eps = 2.220446049250313e-16;
    std::vector<double> cap(n), b, goal(n);
    for (int j = 0; j < n; ++j) cap[j] = CS_FLOOR_CAP * eps * scale[j];
    KFC_HIP(op->ctx, hipMemcpyAsync(QB, B, sizeof(S) * size_t(N) * n, hipMemcpyDeviceToDevice, st));
    if ((rc = cw.project(QB)) != KFC_OK) return rc;
    if ((rc = blas_colnorms<S>(w, QB, N, n, N, b)) != KFC_OK) return rc;
    std::vector<char> met(n, 0), atfloor(n, 0), from_start(n, 0);
    std::vector<double> keep(n, 1.0);
    for (int j = 0; j < n; ++j) {
        goal[j] = tol * b[j];
        if (b[j] <= cap[j]) from_start[j] = atfloor[j] = 1, keep[j] = 0.0;
    }
    if ((rc = blas_scale_cols<S>(w, QB, N, n, N, keep)) != KFC_OK) return rc;   // x_j = 0 exactly there
    KFC_HIP(op->ctx, hipMemsetAsync(X, 0, sizeof(S) * size_t(N) * n, st));
    KFC_HIP(op->ctx, hipMemcpyAsync(R, QB, sizeof(S) * size_t(N) * n, hipMemcpyDeviceToDevice, st));
    ColScal sh{};
    for (int j = 0; j < n; ++j) sh.v[j] = shifts[j];
    // Z = Q (Minv .* R)
    auto precondition = [&]() -> int {
        hipLaunchKernelGGL(k_cs_prec<S>, dim3(nb), dim3(256), 0, st, Z, R, D, N, n, sh, eta);
        KFC_HIP(op->ctx, hipGetLastError());
        return cw.project(Z);
    };
    // Y = (H - s) X, the shift of each column in the apply's epilogue
    auto shifted_apply = [&](const S* Xs, S* Y) -> int {
        for (int j = 0; j < n; ++j) {
            const int r = op_apply_cheb(op, Xs + size_t(j) * N, nullptr, Y + size_t(j) * N, 1, N, KFC_PREC_F64, 1.0,
                                        -shifts[j], 0.0, st);
            if (r != KFC_OK) return r;
        }
        return KFC_OK;
    };
    if ((rc = precondition()) != KFC_OK) return rc;
    KFC_HIP(op->ctx, hipMemcpyAsync(P, Z, sizeof(S) * size_t(N) * n, hipMemcpyDeviceToDevice, st));
    std::vector<double> rho, kappa, rn = b, tr, rhon;
    if ((rc = cw.colredot(R, Z, rho)) != KFC_OK) return rc;
    // |Q (H - s) X - Q B| per column, with a fresh application of H
    auto true_residuals = [&](std::vector<double>& out) -> int {
        int r = shifted_apply(X, T);
        if (r == KFC_OK) r = cw.project(T);
        if (r == KFC_OK) r = blas_axpby<S>(w, QB, N, T, N, n, N, -1.0, 1.0);
        if (r == KFC_OK) r = blas_colnorms<S>(w, T, N, n, N, out);
        return r;
    };
    std::vector<double> anchor(n, INFINITY);
    std::vector<int> stalled(n, 0);
    auto done = [&]() {
        for (int j = 0; j < n; ++j)
            if (!(met[j] || atfloor[j])) return false;
        return true;
    };
    int64_t it = 0;
    while (!done() && it < maxiter) {
        ++it;
        // AP = Q (H - s) P
        if ((rc = shifted_apply(P, AP)) != KFC_OK) return rc;
        if ((rc = cw.project(AP)) != KFC_OK) return rc;
        if ((rc = cw.colredot(P, AP, kappa)) != KFC_OK) return rc;
        for (int j = 0; j < n; ++j) {
            if (met[j] || atfloor[j] || kappa[j] > 0) continue;
            if (rn[j] <= cap[j]) {
                atfloor[j] = 1;
                continue;
            }
            char buf[512];
            snprintf(buf, sizeof buf,
                     "the linear solve on the complement is not positive on the complement at column %d "
                     "(<p, (H - s) p> = %.6g after %lld iterations, residual %.6g relative to its complement part): "
                     "a state below or degenerate with the kept state was not computed (or was not retained); raise "
                     "`compute` (or retain the node's computed states)",
                     j + 1, kappa[j], (long long)it, rn[j] / b[j]);
            return kfc_fail(op->ctx, KFC_EBREAKDOWN, buf);
        }
        ColScal a{};
        for (int j = 0; j < n; ++j) a.v[j] = (met[j] || atfloor[j]) ? 0.0 : rho[j] / kappa[j];
        std::vector<char> active(n);
        for (int j = 0; j < n; ++j) active[j] = !(met[j] || atfloor[j]);
        hipLaunchKernelGGL(k_cs_step<S>, dim3(nb), dim3(256), 0, st, X, R, P, AP, N, n, a);
        KFC_HIP(op->ctx, hipGetLastError());
        if ((rc = blas_colnorms<S>(w, R, N, n, N, rn)) != KFC_OK) return rc;
        bool any_near = false;
        std::vector<char> near(n, 0);
        for (int j = 0; j < n; ++j) {
            if (active[j] && rn[j] <= goal[j]) met[j] = 1;
            near[j] = !(met[j] || atfloor[j]) && rn[j] <= cap[j];
            any_near |= near[j] != 0;
        }
        if (any_near) {
            if ((rc = true_residuals(tr)) != KFC_OK) return rc;
            for (int j = 0; j < n; ++j) {
                if (!near[j]) continue;
                if (tr[j] < anchor[j] / 2) {
                    anchor[j] = tr[j];
                    stalled[j] = 0;
                } else {
                    ++stalled[j];
                }
                if (stalled[j] >= CS_FLOOR_PATIENCE && tr[j] <= cap[j]) atfloor[j] = 1;
            }
        }
        if (done()) break;
        if ((rc = precondition()) != KFC_OK) return rc;
        if ((rc = cw.colredot(R, Z, rhon)) != KFC_OK) return rc;
        ColScal bb{};
        for (int j = 0; j < n; ++j) bb.v[j] = (met[j] || atfloor[j]) ? 0.0 : rhon[j] / rho[j];
        hipLaunchKernelGGL(k_cs_dir<S>, dim3(nb), dim3(256), 0, st, P, Z, N, n, bb);
        KFC_HIP(op->ctx, hipGetLastError());
        rho = rhon;
    }
    if ((rc = true_residuals(tr)) != KFC_OK) return rc;
    for (int j = 0; j < n; ++j)
        if (!(met[j] || atfloor[j]) && tr[j] <= cap[j]) atfloor[j] = 1;   // at maxiter, at the round-off floor
    if (!done()) {
        int j = 0;
        while (met[j] || atfloor[j]) ++j;
        char buf[400];
        snprintf(buf, sizeof buf,
                 "the linear solve on the complement did not converge at column %d: residual %.6g relative to its "
                 "complement part (true residual %.6g relative to the right-hand side) after %lld iterations "
                 "(tolerance %g, at most %lld iterations)",
                 j + 1, rn[j] / b[j], tr[j] / scale[j], (long long)it, tol, (long long)maxiter);
        return kfc_fail(op->ctx, KFC_ENOCONV, buf);
    }
    for (int j = 0; j < n; ++j) {
        if (from_start[j]) continue;
        const bool met0 = met[j];
        if (met[j] && tr[j] > 10 * goal[j]) {
            met[j] = 0;
            atfloor[j] = 1;   // the recurrence went below the true floor
        }
        if (atfloor[j] && tr[j] > cap[j]) {
            char buf[400];
            snprintf(buf, sizeof buf,
                     "the linear solve on the complement did not converge at column %d: %s%.6g relative to the "
                     "right-hand side, above the round-off cap 1000 eps (after %lld iterations)",
                     j + 1,
                     met0 ? "it met its tolerance by the recurrence, but its true residual is "
                          : "it stopped at the round-off floor, but its final true residual is ",
                     tr[j] / scale[j], (long long)it);
            return kfc_fail(op->ctx, KFC_ENOCONV, buf);
        }
    }
    *iters_out = it;
    for (int j = 0; j < n; ++j) {
        const bool af = atfloor[j] && !from_start[j];
        if (residuals) residuals[j] = met[j] ? tr[j] / b[j] : 0.0;
        if (floor_residuals) floor_residuals[j] = af ? tr[j] / scale[j] : 0.0;
        if (flags) flags[j] = (met[j] ? 1 : 0) | (from_start[j] ? 2 : 0) | (af ? 4 : 0);
    }
    return KFC_OK;
}

}  // namespace

extern "C" int kfc_complement_solve(kfc_op* op, const double* D, const void* basis, int64_t k, const void* B,
                                    int64_t n, const double* shifts, double tol, int64_t maxiter, double eta,
                                    const double* scale, void* X, int64_t* iterations, double* residuals,
                                    int32_t* flags, double* floor_residuals) {
    if (!op || !D || !B || !X || !shifts || !iterations || (k > 0 && !basis))
        return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_complement_solve: null argument");
    if (n < 1 || n > CS_MAXC)
        return kfc_fail(op->ctx, KFC_EARG, "kfc_complement_solve: the number of columns must lie in 1:64");
    if (k < 0 || k >= op->N) return kfc_fail(op->ctx, KFC_EARG, "kfc_complement_solve: the basis must have 0 <= k < N columns");
    if (!(tol > 0 && std::isfinite(tol))) return kfc_fail(op->ctx, KFC_EARG, "kfc_complement_solve: tol must be positive and finite");
    if (maxiter < 1) return kfc_fail(op->ctx, KFC_EARG, "kfc_complement_solve: maxiter must be at least 1");
    if (!(eta >= 0 && std::isfinite(eta))) return kfc_fail(op->ctx, KFC_EARG, "kfc_complement_solve: eta must be finite and nonnegative");
    for (int64_t j = 0; j < n; ++j)
        if (!std::isfinite(shifts[j])) return kfc_fail(op->ctx, KFC_EARG, "kfc_complement_solve: shifts must be finite");
    return op->cplx
               ? complement_t<Cx<double>>(op, D, static_cast<const Cx<double>*>(basis), int(k),
                                          static_cast<const Cx<double>*>(B), int(n), shifts, tol, maxiter, eta, scale,
                                          static_cast<Cx<double>*>(X), iterations, residuals, flags, floor_residuals)
               : complement_t<double>(op, D, static_cast<const double*>(basis), int(k), static_cast<const double*>(B),
                                      int(n), shifts, tol, maxiter, eta, scale, static_cast<double*>(X), iterations,
                                      residuals, flags, floor_residuals);
}
