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

// Re(conj(a) b)
__device__ inline double s_redot(double a, double b) { return a * b; }
__device__ inline double s_redot(Cx<double> a, Cx<double> b) { return a.re * b.re + a.im * b.im; }

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

// X[:, j] += a_j P[:, j];  R[:, j] -= a_j AP[:, j], and the partial sums of |R[:, j]|^2
// after the update (block (bx, j) of an nbx x n grid writes part[j nbx + bx]).  A
// column with a_j == 0 is left as it is (its norm is unchanged).
template <class S>
__global__ __launch_bounds__(256) void k_cs_step_norm(S* __restrict__ X, S* __restrict__ R, const S* __restrict__ P,
                                                      const S* __restrict__ AP, int64_t N, ColScal a,
                                                      double* __restrict__ part) {
    const int j = blockIdx.y;
    const double aj = a.v[j];
    if (aj == 0.0) return;
    const int64_t off = int64_t(j) * N;
    double acc = 0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < N; i += int64_t(gridDim.x) * 256) {
        X[off + i] = s_add(X[off + i], s_scale(aj, P[off + i]));
        const S r = s_add(R[off + i], s_scale(-aj, AP[off + i]));
        R[off + i] = r;
        acc += s_abs2(r);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[int64_t(j) * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Y[:, c0 + c] -= sum_l Phi[:, l] C[l, c0 + c] for c < NC, and the partial sums of
// Re <A[:, c0 + c], Y[:, c0 + c]> after the update (block b writes part[c nblk + b]).
// One thread per row: the row of Phi is read once for the NC columns.
template <class S, int NC>
__global__ __launch_bounds__(256) void k_cs_proj_dot(S* __restrict__ Y, const S* __restrict__ A,
                                                     const S* __restrict__ Phi, const S* __restrict__ Cm, int k,
                                                     int ldc, int64_t N, int nc, double* __restrict__ part) {
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < N; i += int64_t(gridDim.x) * 256) {
        S y[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) y[c] = c < nc ? Y[i + c * N] : S{};
        for (int l = 0; l < k; ++l) {
            const S ph = Phi[i + int64_t(l) * N];
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (c < nc) y[c] = s_sub(y[c], s_mul(ph, Cm[l + c * ldc]));
        }
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if (c < nc) {
                Y[i + c * N] = y[c];
                acc[c] += s_redot(A[i + c * N], y[c]);
            }
    }
    __shared__ double red[4][NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        double v = acc[c];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][c] = v;
    }
    __syncthreads();
    if (int(threadIdx.x) < nc)
        part[int64_t(threadIdx.x) * gridDim.x + blockIdx.x] =
            red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
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
    DevBuf dC, dpart;
    std::vector<double> hpart;
    // Y <- Y - Phi (Phi^H Y) with dots[j] = Re <A_j, Y_j> after it: the Gram, then
    // one pass that updates Y and reduces the dots
    int project_dot(S* Y, const S* A, std::vector<double>& dots) {
        dots.assign(n, 0.0);
        std::vector<S> hc;
        if (k > 0) {
            int rc = blas_gram<S>(w, Phi, N, k, Y, N, n, N, C);
            if (rc != KFC_OK) return rc;
            hc.resize(size_t(k) * n);
            for (size_t i = 0; i < hc.size(); ++i) hc[i] = s_from<S>(C[i]);
        } else {
            hc.assign(1, S{});
        }
        KFC_HIP(op->ctx, dC.ensure(sizeof(S) * hc.size()));
        KFC_HIP(op->ctx, hipMemcpyAsync(dC.p, hc.data(), sizeof(S) * hc.size(), hipMemcpyHostToDevice, op->ctx->stream));
        const int nblk = int(std::max<int64_t>(1, std::min<int64_t>((N + 255) / 256, 2048)));
        KFC_HIP(op->ctx, dpart.ensure(sizeof(double) * size_t(nblk) * 16));
        hpart.resize(size_t(nblk) * 16);
        for (int c0 = 0; c0 < n; c0 += 16) {
            const int nc = std::min(16, n - c0);
#define KFC_PD(NCv)                                                                                       \
    hipLaunchKernelGGL((k_cs_proj_dot<S, NCv>), dim3(nblk), dim3(256), 0, op->ctx->stream, Y + size_t(c0) * N, \
                       A + size_t(c0) * N, Phi, dC.as<S>() + size_t(c0) * std::max(k, 1), k, std::max(k, 1), N, nc, \
                       dpart.as<double>())
            // the column count of the launch rounded up to 4: no idle accumulators
            if (nc <= 4) KFC_PD(4);
            else if (nc <= 8) KFC_PD(8);
            else if (nc <= 12) KFC_PD(12);
            else KFC_PD(16);
#undef KFC_PD
            KFC_HIP(op->ctx, hipGetLastError());
            KFC_HIP(op->ctx, hipMemcpyAsync(hpart.data(), dpart.p, sizeof(double) * size_t(nblk) * nc,
                                            hipMemcpyDeviceToHost, op->ctx->stream));
            KFC_HIP(op->ctx, hipStreamSynchronize(op->ctx->stream));
            for (int c = 0; c < nc; ++c) {
                double sum = 0;
                for (int b = 0; b < nblk; ++b) sum += hpart[size_t(c) * nblk + b];
                dots[c0 + c] = sum;
            }
        }
        return KFC_OK;
    }
    // X += a P, R -= a AP with the norms of R after it (columns with a_j == 0 keep rn_j)
    int step_norm(S* X, S* R, const S* P, const S* AP, const ColScal& a, std::vector<double>& rn) {
        const int nbx = int(std::max<int64_t>(1, std::min<int64_t>((N + 255) / 256, (2048 + n - 1) / n)));
        KFC_HIP(op->ctx, dpart.ensure(sizeof(double) * size_t(nbx) * std::max(n, 16)));
        hipLaunchKernelGGL(k_cs_step_norm<S>, dim3(nbx, n), dim3(256), 0, op->ctx->stream, X, R, P, AP, N, a,
                           dpart.as<double>());
        KFC_HIP(op->ctx, hipGetLastError());
        hpart.resize(size_t(nbx) * n);
        KFC_HIP(op->ctx, hipMemcpyAsync(hpart.data(), dpart.p, sizeof(double) * hpart.size(), hipMemcpyDeviceToHost,
                                        op->ctx->stream));
        KFC_HIP(op->ctx, hipStreamSynchronize(op->ctx->stream));
        for (int j = 0; j < n; ++j) {
            if (a.v[j] == 0.0) continue;
            double sum = 0;
            for (int b = 0; b < nbx; ++b) sum += hpart[size_t(j) * nbx + b];
            rn[j] = std::sqrt(sum);
        }
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
    // Z = Q (Minv .* R) with rho_j = Re <R_j, Z_j>
    auto precondition_dot = [&](std::vector<double>& dots) -> int {
        hipLaunchKernelGGL(k_cs_prec<S>, dim3(nb), dim3(256), 0, st, Z, R, D, N, n, sh, eta);
        KFC_HIP(op->ctx, hipGetLastError());
        return cw.project_dot(Z, R, dots);
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
    std::vector<double> rho, kappa, rn = b, tr, rhon;
    if ((rc = precondition_dot(rho)) != KFC_OK) return rc;
    KFC_HIP(op->ctx, hipMemcpyAsync(P, Z, sizeof(S) * size_t(N) * n, hipMemcpyDeviceToDevice, st));
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
        if ((rc = cw.project_dot(AP, P, kappa)) != KFC_OK) return rc;
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
        if ((rc = cw.step_norm(X, R, P, AP, a, rn)) != KFC_OK) return rc;
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
        if ((rc = precondition_dot(rhon)) != KFC_OK) return rc;
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
