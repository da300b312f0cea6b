// Eigensolver loops on the device: Dense, thick-restart Lanczos and LOBPCG.
// Block data stay on the device; only small projected matrices cross to the
// host, where they are diagonalized in FP64 (herm_eig).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <numeric>
#include <random>
#include <thread>

#include "blas.h"

int op_full_diagonal(kfc_op* op, std::vector<double>& D);
int lobpcg_check_x0(kfc_op* op, const double* X0, int64_t nx0, std::vector<cd>& X0h);
void lobpcg_seed_block(const kfc_op* op, int m, int W, const std::vector<cd>& X0h, int64_t nx0,
                       const std::vector<cd>& P, bool random_only, uint64_t s, std::vector<cd>& X, int& nperturb);
template <class S>
int lobpcg_seed_dev(kfc_op* op, int m, int W, const std::vector<cd>& X0h, int64_t nx0, bool product, bool random_only,
                    uint64_t s, S* Xd, int& nperturb);

namespace {

using clk = std::chrono::steady_clock;
inline double secs(clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); }

template <class S> struct Lower { using type = S; };
template <> struct Lower<double> { using type = float; };
template <> struct Lower<Cx<double>> { using type = Cx<float>; };

template <class S> struct Upper { using type = S; };
template <> struct Upper<float> { using type = double; };
template <> struct Upper<Cx<float>> { using type = Cx<double>; };

template <class S> constexpr int prec_of() { return std::is_same<typename RealOf<S>::type, double>::value ? 0 : 1; }

// Host <-> device block transfer of N x p (host cd array, column-major)
template <class S>
int upload_block(kfc_ctx* ctx, const std::vector<cd>& H, int64_t N, int p, S* D, int64_t ld) {
    std::vector<S> h(size_t(N) * p);
    for (size_t i = 0; i < h.size(); ++i) h[i] = s_from<S>(H[i]);
    if (ld == N) {
        KFC_HIP(ctx, hipMemcpyAsync(D, h.data(), sizeof(S) * h.size(), hipMemcpyHostToDevice, ctx->stream));
    } else {
        for (int j = 0; j < p; ++j)
            KFC_HIP(ctx, hipMemcpyAsync(D + size_t(j) * ld, h.data() + size_t(j) * N, sizeof(S) * N,
                                        hipMemcpyHostToDevice, ctx->stream));
    }
    KFC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KFC_OK;
}

template <class S>
int download_block(kfc_ctx* ctx, const S* D, int64_t ld, int64_t N, int p, std::vector<cd>& H) {
    std::vector<S> h(size_t(N) * p);
    if (ld == N) {
        KFC_HIP(ctx, hipMemcpyAsync(h.data(), D, sizeof(S) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    } else {
        for (int j = 0; j < p; ++j)
            KFC_HIP(ctx, hipMemcpyAsync(h.data() + size_t(j) * N, D + size_t(j) * ld, sizeof(S) * N,
                                        hipMemcpyDeviceToHost, ctx->stream));
    }
    KFC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    H.resize(h.size());
    for (size_t i = 0; i < h.size(); ++i) {
        if constexpr (IsCx<S>::value) H[i] = cd(h[i].re, h[i].im);
        else H[i] = cd(h[i], 0.0);
    }
    return KFC_OK;
}

void write_host_V(const kfc_op* op, const std::vector<cd>& V, double* out) {
    for (size_t i = 0; i < V.size(); ++i) {
        if (op->cplx) {
            out[2 * i] = V[i].real();
            out[2 * i + 1] = V[i].imag();
        } else {
            out[i] = V[i].real();
        }
    }
}

void read_host_V(const kfc_op* op, const double* in, size_t n, std::vector<cd>& V) {
    V.resize(n);
    for (size_t i = 0; i < n; ++i) V[i] = op->cplx ? cd(in[2 * i], in[2 * i + 1]) : cd(in[i], 0.0);
}

// Hermitian eig of a small host matrix (n x n column-major, symmetrized).
void small_eig(int n, const std::vector<cd>& G, std::vector<double>& w, std::vector<cd>& Z) {
    std::vector<cd> A(size_t(n) * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) A[i + size_t(j) * n] = 0.5 * (G[i + size_t(j) * n] + std::conj(G[j + size_t(i) * n]));
    herm_eig(n, A, w, Z);
}

struct Timer {
    double apply = 0, apply_gpu = 0;
    int64_t cols = 0, cols32 = 0, calls = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // event pairs of applies not yet read back (read at the next flush, so that an
    // apply inside a Krylov block does not stall the host)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending, pool;
    Timer() {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
    }
    ~Timer() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        for (auto& pr : pending) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
        for (auto& pr : pool) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    }
    Timer(const Timer&) = delete;
    Timer& operator=(const Timer&) = delete;
    hipError_t pair(std::pair<hipEvent_t, hipEvent_t>& pr) {
        if (!pool.empty()) {
            pr = pool.back();
            pool.pop_back();
            return hipSuccess;
        }
        hipError_t e = hipEventCreate(&pr.first);
        if (e == hipSuccess) e = hipEventCreate(&pr.second);
        return e;
    }
    // add the device time of every pending apply (waits for the last one)
    hipError_t flush() {
        for (auto& pr : pending) {
            hipError_t e = hipEventSynchronize(pr.second);
            if (e != hipSuccess) return e;
            float ms = 0;
            e = hipEventElapsedTime(&ms, pr.first, pr.second);
            if (e != hipSuccess) return e;
            apply_gpu += 1e-3 * ms;
            pool.push_back(pr);
        }
        pending.clear();
        return hipSuccess;
    }
};

template <class S>
int apply_timed(kfc_op* op, const S* X, S* Y, int64_t ncols, int64_t ld, Timer& t, double alpha = 1.0) {
    if (ncols == 0) return KFC_OK;
    auto t0 = clk::now();
    KFC_HIP(op->ctx, hipEventRecord(t.e0, op->ctx->stream));
    int rc = alpha == 1.0 ? op_apply(op, X, Y, ncols, ld, prec_of<S>(), op->ctx->stream)
                          : op_apply_cheb(op, X, nullptr, Y, ncols, ld, prec_of<S>(), alpha, 0.0, 0.0, op->ctx->stream);
    if (rc != KFC_OK) return rc;
    KFC_HIP(op->ctx, hipEventRecord(t.e1, op->ctx->stream));
    KFC_HIP(op->ctx, hipEventSynchronize(t.e1));
    float ms = 0;
    KFC_HIP(op->ctx, hipEventElapsedTime(&ms, t.e0, t.e1));
    t.apply_gpu += 1e-3 * ms;
    t.apply += secs(t0);
    t.cols += ncols;
    if (prec_of<S>() == 1) t.cols32 += ncols;
    t.calls += 1;
    return KFC_OK;
}

// Fresh FP64 true residuals of pairs whose vectors are on the device (N x m, ld N).
template <class S>
int true_residuals_dev(kfc_op* op, Work& w, const std::vector<double>& E, const S* X, int m, std::vector<double>& res,
                       Timer& t) {
    const int64_t N = op->N;
    DevBuf by;
    KFC_HIP(op->ctx, by.ensure(sizeof(S) * size_t(N) * m));
    int rc = apply_timed<S>(op, X, by.as<S>(), m, N, t);
    if (rc == KFC_OK) rc = blas_residual<S>(w, by.as<S>(), N, by.as<S>(), N, X, N, m, N, E, res);
    return rc;
}

// Fresh FP64 true residuals of host pairs.
template <class S>
int true_residuals(kfc_op* op, Work& w, const std::vector<double>& E, const std::vector<cd>& Vh, int m,
                   std::vector<double>& res, Timer& t) {
    const int64_t N = op->N;
    DevBuf bx, by;
    KFC_HIP(op->ctx, bx.ensure(sizeof(S) * size_t(N) * m));
    KFC_HIP(op->ctx, by.ensure(sizeof(S) * size_t(N) * m));
    int rc = upload_block<S>(op->ctx, Vh, N, m, bx.as<S>(), N);
    if (rc == KFC_OK) rc = apply_timed<S>(op, bx.as<S>(), by.as<S>(), m, N, t);
    if (rc == KFC_OK) rc = blas_residual<S>(w, by.as<S>(), N, by.as<S>(), N, bx.as<S>(), N, m, N, E, res);
    return rc;
}

// ------------------------------------------------------------------------------------------
// Dense: Matrix(H) by identity blocks on the device (src/apply.jl:506-531), host eig.

// Matrix(H) of a small operator formed on the host from its terms (src/apply.jl:506-531
// forms it by applying H to identity blocks; for a few hundred states the device round
// trips cost more than the matrix): the fused diagonal plus, per kept term, its
// coefficient times the Kronecker product of the factors' dense forms.
void host_matrix(kfc_op* op, std::vector<cd>& M) {
    const int64_t N = op->N;
    const int K = op->ndims;
    std::vector<int64_t> stride(K, 1);
    for (int a = 1; a < K; ++a) stride[a] = stride[a - 1] * op->dims[a - 1];
    M.assign(size_t(N) * N, cd(0));
    for (int64_t i = 0; i < N; ++i) M[i + size_t(i) * N] = op->hdiag[i];
    for (const Term& t : op->terms) {
        std::vector<std::vector<cd>> F(K);   // empty: identity on that axis
        for (const Factor& f : t.factors) {
            const int64_t d = f.d;
            std::vector<cd>& A = F[f.axis];
            A.assign(size_t(d) * d, cd(0));
            if (f.kind == KFC_DIAG) {
                for (int64_t i = 0; i < d; ++i) A[i + size_t(i) * d] = f.diag[i];
            } else if (f.kind == KFC_BANDED) {   // y[i] = sum_b w_b[i] x[i + o_b] on valid rows
                for (size_t b = 0; b < f.offsets.size(); ++b)
                    for (int64_t i = 0; i < d; ++i) {
                        const int64_t j = i + f.offsets[b];
                        if (j >= 0 && j < d) A[i + size_t(j) * d] += f.weights[i + b * d];
                    }
            } else if (f.kind == KFC_SHIFT) {    // y[i + s] = factor x[i]
                for (int64_t i = 0; i < d; ++i) {
                    const int64_t r = i + f.shift;
                    if (r >= 0 && r < d) A[r + size_t(i) * d] += f.factor;
                }
            } else {
                A = f.matrix;
            }
        }
        for (int64_t c = 0; c < N; ++c)
            for (int64_t r = 0; r < N; ++r) {
                cd v = t.coeff;
                for (int a = 0; a < K && v != cd(0); ++a) {
                    const int64_t d = op->dims[a];
                    const int64_t ra = (r / stride[a]) % d, ca = (c / stride[a]) % d;
                    if (F[a].empty()) {
                        if (ra != ca) v = 0;
                    } else {
                        v *= F[a][ra + size_t(ca) * d];
                    }
                }
                M[r + size_t(c) * N] += v;
            }
    }
}

// Operators up to this many states have Matrix(H) formed on the host
constexpr int64_t DENSE_HOST_MAX = 256;

template <class S>
int dense_t(kfc_op* op, int64_t m, double* Eout, double* Vout, kfc_stats* st) {
    auto t0 = clk::now();
    const int64_t N = op->N;
    if (N > 8192) return kfc_fail(op->ctx, KFC_EUNSUPPORTED, "Dense: prod(dims) above 8192 is not served; use an iterative solver");
    Timer tm;
    std::vector<cd> M, blk;
    if (N <= DENSE_HOST_MAX) {
        const int rc = op_host_diag(op);
        if (rc != KFC_OK) return rc;
        host_matrix(op, M);
    }
    const int B = int(std::min<int64_t>(N, 256));
    DevBuf be, by;
    if (N > DENSE_HOST_MAX) {
        KFC_HIP(op->ctx, be.ensure(sizeof(S) * size_t(N) * B));
        KFC_HIP(op->ctx, by.ensure(sizeof(S) * size_t(N) * B));
        M.assign(size_t(N) * N, cd(0));
    }
    for (int64_t c0 = 0; N > DENSE_HOST_MAX && c0 < N; c0 += B) {
        const int w = int(std::min<int64_t>(B, N - c0));
        std::vector<cd> E(size_t(N) * w, cd(0));
        for (int j = 0; j < w; ++j) E[(c0 + j) + size_t(j) * N] = 1;
        int rc = upload_block<S>(op->ctx, E, N, w, be.as<S>(), N);
        if (rc == KFC_OK) rc = apply_timed<S>(op, be.as<S>(), by.as<S>(), w, N, tm);
        if (rc == KFC_OK) rc = download_block<S>(op->ctx, by.as<S>(), N, N, w, blk);
        if (rc != KFC_OK) return rc;
        std::copy(blk.begin(), blk.end(), M.begin() + size_t(c0) * N);
    }
    std::vector<double> w;
    std::vector<cd> Z;
    small_eig(int(N), M, w, Z);
    const int64_t k = (m > 0 && m < N) ? m : N;
    for (int64_t j = 0; j < k; ++j) Eout[j] = w[j];
    Z.resize(size_t(N) * k);
    write_host_V(op, Z, Vout);
    if (st) {
        *st = kfc_stats{};
        st->n_matvec = tm.cols;
        st->n_apply_calls = tm.calls;
        st->seconds_apply = tm.apply;
        st->seconds_apply_gpu = tm.apply_gpu;
        st->seconds_total = secs(t0);
        st->n_converged = int(k);
    }
    return KFC_OK;
}

// Dense solves of many small operators (the tops of a batched sweep): each Matrix(H)
// formed on the host and its eigendecomposition, on host threads; larger operators
// take dense_t one after another.
int dense_batch(int n, kfc_op* const* ops, int64_t m, double* E, double* const* Vs, kfc_stats* stats) {
    std::vector<int> rcs(n, KFC_OK), big;
    for (int i = 0; i < n; ++i) {
        if (ops[i]->N > DENSE_HOST_MAX) {
            big.push_back(i);
            continue;
        }
        rcs[i] = op_host_diag(ops[i]);   // memoized on the operator, formed before the threads
        if (rcs[i] != KFC_OK) return kfc_fail(ops[0]->ctx, rcs[i], "Dense (batch) of operator " + std::to_string(i) + ": " + ops[i]->ctx->err);
    }
    std::atomic<int> next{0};
    auto work = [&] {
        for (int i = next++; i < n; i = next++) {
            kfc_op* op = ops[i];
            const int64_t N = op->N;
            if (N > DENSE_HOST_MAX) continue;
            auto t0 = clk::now();
            std::vector<cd> M, Z;
            std::vector<double> w;
            host_matrix(op, M);
            small_eig(int(N), M, w, Z);
            const int64_t k = (m > 0 && m < N) ? m : N;
            for (int64_t j = 0; j < k; ++j) E[size_t(i) * m + j] = w[j];
            Z.resize(size_t(N) * k);
            write_host_V(op, Z, Vs[i]);
            if (stats) {
                stats[i] = kfc_stats{};
                stats[i].seconds_total = secs(t0);
                stats[i].n_converged = int(k);
            }
        }
    };
    const int T = std::max(1, std::min({n, 16, int(std::thread::hardware_concurrency())}));
    std::vector<std::thread> th;
    for (int t = 1; t < T; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    for (int i : big) {
        const int rc = ops[i]->cplx ? dense_t<Cx<double>>(ops[i], m, E + size_t(i) * m, Vs[i], stats ? stats + i : nullptr)
                                    : dense_t<double>(ops[i], m, E + size_t(i) * m, Vs[i], stats ? stats + i : nullptr);
        if (rc != KFC_OK) return kfc_fail(ops[0]->ctx, rc, "Dense (batch) of operator " + std::to_string(i) + ": " + ops[i]->ctx->err);
    }
    return KFC_OK;
}

// ------------------------------------------------------------------------------------------
// Thick-restart Lanczos (Wu & Simon's restart, full reorthogonalization), computed
// s steps at a time.
//
// The reference's Lanczos (KrylovKit's, src/solvers.jl:85-106) extends the Krylov
// basis by one vector per matvec and orthogonalizes it against the whole basis,
// which reads the basis twice per matvec.  Here each cycle extends it by s vectors
// at once: K = [q, T1(Hs) q, ..., Ts(Hs) q] with Ts the Chebyshev polynomials of H
// scaled to the spectral interval [lo, hi] (s matvecs), then K is orthogonalized
// against the basis Q in one block pass (block classical Gram-Schmidt, twice) and
// within itself (Cholesky QR, twice).  The projected matrix of the extended basis
// follows from the Chebyshev recurrence coefficients and the two factorizations
// without further matvecs (the s-step Krylov formulation of Hoemmen's thesis,
// 2010, ch. 4), so the basis is read three times per s matvecs.  The subspace is
// the one the one-vector recurrence builds; only the order of the arithmetic
// differs.
//
// Between cycles the Krylov decomposition H Q = Q T + q b^T is kept, with Q
// orthonormal (k columns), q a unit vector orthogonal to Q, T Hermitian.

template <class S>
struct LanczosOut {
    std::vector<double> theta;   // Ritz values of the last projection
    std::vector<double> resid;   // |b^T y|, the Ritz residual norms
    int converged = 0;
    int restarts = 0;
    int cycles = 0;
    bool invariant = false;
};

// Upper Cholesky factor R (n x n) of the leading block of a Hermitian G (leading
// dimension ldg): G = R^H R.  Returns the first column whose pivot is not above
// `dep` times its diagonal entry (n when every pivot passes); the off-diagonal
// entries of that column are filled before its pivot is tested.
int chol_upper(int n, const std::vector<cd>& G, int ldg, std::vector<cd>& R, double dep, double& pivot_rel) {
    R.assign(size_t(n) * n, cd(0));
    pivot_rel = 1.0;
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < j; ++i) {
            cd s = G[i + size_t(j) * ldg];
            for (int l = 0; l < i; ++l) s -= std::conj(R[l + size_t(i) * n]) * R[l + size_t(j) * n];
            R[i + size_t(j) * n] = s / R[i + size_t(i) * n];
        }
        const double gjj = G[j + size_t(j) * ldg].real();
        double d = gjj;
        for (int l = 0; l < j; ++l) d -= std::norm(R[l + size_t(j) * n]);
        if (!(gjj > 0) || !(d > dep * gjj)) {
            pivot_rel = gjj > 0 ? d / gjj : 0.0;
            return j;
        }
        R[j + size_t(j) * n] = std::sqrt(d);
    }
    return n;
}

// Inverse of an upper triangular n x n matrix (column-major, leading dimension n).
std::vector<cd> inv_upper(int n, const std::vector<cd>& R) {
    std::vector<cd> X(size_t(n) * n, cd(0));
    for (int j = 0; j < n; ++j) {
        X[j + size_t(j) * n] = 1.0 / R[j + size_t(j) * n];
        for (int i = j - 1; i >= 0; --i) {
            cd s = 0;
            for (int l = i + 1; l <= j; ++l) s += R[i + size_t(l) * n] * X[l + size_t(j) * n];
            X[i + size_t(j) * n] = -s / R[i + size_t(i) * n];
        }
    }
    return X;
}

// C (r x c) = A (r x t) B (t x c), all column-major with leading dimensions r, r, t.
std::vector<cd> mm(int r, int t, int c, const std::vector<cd>& A, const std::vector<cd>& B) {
    std::vector<cd> C(size_t(r) * c, cd(0));
    for (int j = 0; j < c; ++j)
        for (int l = 0; l < t; ++l) {
            const cd b = B[l + size_t(j) * t];
            if (b == cd(0)) continue;
            for (int i = 0; i < r; ++i) C[i + size_t(j) * r] += A[i + size_t(l) * r] * b;
        }
    return C;
}

// dots (device, real S): also sum B^2 and sum B C into dots[0], dots[1] when the fused
// apply can (*fused says whether it did).
template <class S>
int apply_cheb_timed(kfc_op* op, const S* B, const S* A, S* C, double a, double b, double g, Timer& t,
                     double* dots = nullptr, bool* fused = nullptr) {
    auto t0 = clk::now();
    std::pair<hipEvent_t, hipEvent_t> pr;
    KFC_HIP(op->ctx, t.pair(pr));
    KFC_HIP(op->ctx, hipEventRecord(pr.first, op->ctx->stream));
    bool f = false;
    int rc = dots ? op_apply_cheb_dots(op, B, A, C, op->N, prec_of<S>(), a, b, g, dots, f, op->ctx->stream)
                  : op_apply_cheb(op, B, A, C, 1, op->N, prec_of<S>(), a, b, g, op->ctx->stream);
    if (fused) *fused = f;
    if (rc != KFC_OK) return rc;
    KFC_HIP(op->ctx, hipEventRecord(pr.second, op->ctx->stream));
    t.pending.push_back(pr);
    t.apply += secs(t0);
    t.cols += 1;
    if (prec_of<S>() == 1) t.cols32 += 1;
    t.calls += 1;
    return KFC_OK;
}

constexpr int LZ_SMAX = 12;   // Chebyshev block length: the basis condition stays near 1e3

// |b^T y_i| for the columns of Y (n x n): the Ritz residual norms of H V = V T + q b^T.
std::vector<double> ritz_residuals(int n, const std::vector<cd>& b, const std::vector<cd>& Y) {
    std::vector<double> r(n);
    for (int i = 0; i < n; ++i) {
        cd acc = 0;
        for (int j = 0; j < n; ++j) acc += b[j] * Y[j + size_t(i) * n];
        r[i] = std::abs(acc);
    }
    return r;
}

// max-norm condition estimate of an upper triangular s x s matrix
double cond_upper(int s, const std::vector<cd>& R) {
    if (s <= 0) return 1.0;
    const std::vector<cd> Ri = inv_upper(s, R);
    double a = 0, b = 0;
    for (size_t i = 0; i < R.size(); ++i) {
        a = std::max(a, std::abs(R[i]));
        b = std::max(b, std::abs(Ri[i]));
    }
    return double(s) * a * b;
}

template <class S>
int lanczos_core(kfc_op* op, Work& w, int m, int kd, double tol, int64_t maxiter, const S* start, int64_t start_ld,
                 uint64_t seed, double glo, double ghi, Timer& tm, LanczosOut<S>& out, DevBuf& result,
                 int verbosity, int s_max = LZ_SMAX) {
    using Rl = typename RealOf<S>::type;
    constexpr bool single = std::is_same<Rl, float>::value;
    // a column whose part outside the span of the others is below ~1e-10 (FP64
    // storage) or ~1e-5 (FP32) of its norm counts as dependent: the Krylov space
    // is then invariant
    const double dep = single ? 1e-10 : 1e-20;
    const double cond_floor = single ? 1e-8 : 1e-14;
    const double eps_s = single ? 5.96e-8 : 1.11e-16;
    const int64_t N = op->N;
    DevBuf bufA, bufB, bufR;
    KFC_HIP(op->ctx, bufA.ensure(sizeof(S) * size_t(N) * (kd + 1)));
    KFC_HIP(op->ctx, bufB.ensure(sizeof(S) * size_t(N) * (kd + 1)));
    S* V = bufA.as<S>();
    S* V2 = bufB.as<S>();
    int rc;
    if (start) rc = blas_convert<S, S>(w, start, start_ld, V, N, 1, N);
    else rc = blas_randn<S>(w, V, N, 1, N, seed);
    if (rc != KFC_OK) return rc;
    {
        std::vector<double> nrm;
        rc = blas_colnorms<S>(w, V, N, 1, N, nrm);
        if (rc != KFC_OK) return rc;
        if (!(nrm[0] > 0)) return kfc_fail(op->ctx, KFC_EBREAKDOWN, "Lanczos: zero start vector");
        rc = blas_scale_cols<S>(w, V, N, 1, N, std::vector<double>{1.0 / nrm[0]});
        if (rc != KFC_OK) return rc;
    }
    const double hi = ghi;
    double lo = std::min(glo, ghi - std::max(1e-300, 1e-12 * std::fabs(ghi)));
    bool ritz_known = false;
    int k = 0;
    std::vector<cd> T, b;   // k x k and k
    int64_t restarts = 0;
    int s_limit = std::max(1, std::min(s_max, LZ_SMAX)), s_cap = s_limit, anchors = 0, fresh_starts = 0;
    // H Q = Q T + q b^T holds to rounding (false after an explicit check found the
    // basis polluted: then every first projection is computed from Q itself)
    bool relation_exact = true;
    for (int64_t cyc = 0;; ++cyc) {
        int s = std::min(s_cap, kd - k);
        if (!ritz_known) s = std::min(s, 4);   // the first block runs on the loose interval
        S* K = V + size_t(k) * N;
        const double cc = 0.5 * (hi + lo), ee = 0.5 * (hi - lo);
        // The first projection below needs the moments mu_j = q^H T_j q (j <= s) of the
        // block's start q = K_0.  For real blocks the fused apply returns <K_i, K_i> and
        // <K_i, K_{i+1}> with K_{i+1}, and T_i T_i = (T_2i + T_0) / 2, T_i T_{i+1} =
        // (T_2i+1 + T_1) / 2 give mu_2i = 2 <K_i, K_i> - mu_0, mu_2i+1 = 2 <K_i, K_{i+1}> - mu_1
        // without a pass over the block.
        const bool want_moments = !IsCx<S>::value && k > 0 && relation_exact;
        const int nmom = s / 2 + 1;   // applies 0 .. s/2 carry the dots
        bool moments_fused = want_moments;
        if (want_moments) KFC_HIP(op->ctx, w.small2.ensure(sizeof(double) * 2 * size_t(nmom)));
        for (int i = 0; i < s; ++i) {
            double* dots = want_moments && i < nmom ? w.small2.as<double>() + 2 * i : nullptr;
            bool fz = false;
            rc = i == 0 ? apply_cheb_timed<S>(op, K, nullptr, K + N, 1.0 / ee, -cc / ee, 0.0, tm, dots, &fz)
                        : apply_cheb_timed<S>(op, K + size_t(i) * N, K + size_t(i - 1) * N, K + size_t(i + 1) * N,
                                              2.0 / ee, -2.0 * cc / ee, -1.0, tm, dots, &fz);
            if (rc != KFC_OK) return rc;
            if (dots && !fz) moments_fused = false;
        }
        const int nb0 = s + 1;
        std::vector<cd> pend_ri;   // a pending K <- K pend_ri (pend_n columns), see the Cholesky QR
        int pend_n = 0;
        // block classical Gram-Schmidt against Q, twice (K <- K - Q Cm): C1 = Q^H K;
        // K -= Q C1 with [C2; Gk] = [Q K]^H K in the same pass.  The second
        // subtraction K -= Q C2 is left to the first Cholesky QR scaling, which
        // reads [Q K] anyway; its Gram follows from this pass's: with Q^H Q = I,
        // (K - Q C2)^H (K - Q C2) = K^H K - C2^H C2.
        std::vector<cd> Cm(size_t(k) * nb0, cd(0)), Gk, C2, M1;
        // chained passes: the first subtraction's pass only measures [Q K']^T K'
        // (K' = K - Q C1 is not stored) and the first Cholesky scaling re-forms K' in
        // registers, bit for bit, before subtracting Q C2 (one write of K per block)
        const bool chain = k > 0 && tsmm_chain_ok<S>(k + nb0, k, nb0, nb0);
        if (k > 0) {
            std::vector<cd> C1;
            if (relation_exact) {
                // The first projection follows from the Krylov relation without a pass
                // over Q: Q^H H v = T Q^H v + conj(b) (q^H v), so Q^H K obeys the
                // block's Chebyshev recurrence given q^H K (one short Gram of the block).
                // The explicit second pass then removes what rounding left.
                std::vector<cd> g0;
                if (moments_fused) {
                    std::vector<double> d(2 * size_t(nmom));
                    KFC_HIP(op->ctx, hipMemcpyAsync(d.data(), w.small2.p, d.size() * sizeof(double),
                                                    hipMemcpyDeviceToHost, op->ctx->stream));
                    KFC_HIP(op->ctx, hipStreamSynchronize(op->ctx->stream));
                    g0.assign(nb0, cd(0));
                    g0[0] = d[0];
                    if (nb0 > 1) g0[1] = d[1];
                    for (int j = 2; j < nb0; ++j) {
                        const int i = j / 2;
                        g0[j] = (j & 1) ? 2.0 * d[2 * i + 1] - d[1] : 2.0 * d[2 * i] - d[0];
                    }
                } else {
                    rc = blas_gram<S>(w, K, N, 1, K, N, nb0, N, g0);
                    if (rc != KFC_OK) return rc;
                }
                C1.assign(size_t(k) * nb0, cd(0));
                auto hcol = [&](int i, std::vector<cd>& out) {   // Q^H (H - c) K_i
                    out.assign(k, cd(0));
                    for (int r = 0; r < k; ++r) {
                        cd acc = std::conj(b[r]) * g0[i] - cc * C1[r + size_t(i) * k];
                        for (int j = 0; j < k; ++j) acc += T[r + size_t(j) * k] * C1[j + size_t(i) * k];
                        out[r] = acc;
                    }
                };
                std::vector<cd> hv;
                for (int i = 0; i < s; ++i) {
                    hcol(i, hv);
                    for (int r = 0; r < k; ++r)
                        C1[r + size_t(i + 1) * k] = i == 0 ? hv[r] / ee
                                                           : 2.0 * hv[r] / ee - C1[r + size_t(i - 1) * k];
                }
            } else {
                rc = blas_gram<S>(w, V, N, k, K, N, nb0, N, C1);
                if (rc != KFC_OK) return rc;
            }
            auto upd = [&](const std::vector<cd>& C) {   // [[-C]; [I]] on [Q K]
                std::vector<cd> M(size_t(k + nb0) * nb0, cd(0));
                for (int j = 0; j < nb0; ++j) {
                    for (int i = 0; i < k; ++i) M[i + size_t(j) * (k + nb0)] = -C[i + size_t(j) * k];
                    M[k + j + size_t(j) * (k + nb0)] = 1.0;
                }
                return M;
            };
            std::vector<cd> G1;
            M1 = upd(C1);
            rc = blas_tsmm<S>(w, V, N, k + nb0, M1, nb0, K, N, N, chain ? 5 : 3, k, &G1);
            if (rc != KFC_OK) return rc;
            const int gp = k + nb0;
            C2.assign(size_t(k) * nb0, cd(0));
            Gk.assign(size_t(nb0) * nb0, cd(0));
            for (int j = 0; j < nb0; ++j)
                for (int i = 0; i < k; ++i) C2[i + size_t(j) * k] = G1[i + size_t(j) * gp];
            for (int j = 0; j < nb0; ++j)
                for (int i = 0; i <= j; ++i) {
                    cd acc = 0.5 * (G1[k + i + size_t(j) * gp] + std::conj(G1[k + j + size_t(i) * gp]));
                    for (int l = 0; l < k; ++l) acc -= std::conj(C2[l + size_t(i) * k]) * C2[l + size_t(j) * k];
                    Gk[i + size_t(j) * nb0] = acc;
                    Gk[j + size_t(i) * nb0] = std::conj(acc);
                }
            for (size_t i = 0; i < Cm.size(); ++i) Cm[i] = C1[i] + C2[i];
        } else {
            rc = blas_gram<S>(w, K, N, nb0, K, N, nb0, N, Gk);
            if (rc != KFC_OK) return rc;
        }
        // Cholesky QR, twice, on the independent leading columns:
        // K_orig[:, j] = Q Cm[:, j] + Qn[:, 0:nind] Rt[0:nind, j]
        int nind = nb0;
        bool dep_last = false;   // column s lies in the span of the columns before it
        const int ldr = nb0;
        std::vector<cd> Rt(size_t(nb0) * nb0, cd(0));
        for (int i = 0; i < nb0; ++i) Rt[i + size_t(i) * ldr] = 1.0;
        std::vector<cd> G = Gk;   // Gram of the current K[:, 0:nind] (leading dimension ldg)
        int ldg = nb0;
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<cd> Rp, dcoef;
            for (;;) {
                double prel = 1;
                const int n0 = nind;
                // a first-pass pivot below `cond_floor` (relative, squared) would make the
                // block's condition number too large for the second pass to repair
                // (eps cond^2 >= 1): such a column ends the block
                const int fc = chol_upper(nind, G, ldg, Rp, pass == 0 ? cond_floor : dep, prel);
                if (fc == nind) break;
                if (pass == 0 && fc >= 1 && (std::fabs(prel) <= dep || fc == 1)) {
                    dcoef.assign(Rp.begin() + size_t(fc) * n0, Rp.begin() + size_t(fc) * n0 + fc);
                    dep_last = true;
                    s = fc;
                    nind = fc;
                } else if (fc >= 2) {   // an ill-conditioned block: drop its tail
                    s = fc - 1;
                    nind = fc;
                    dep_last = false;
                } else {
                    return kfc_fail(op->ctx, KFC_EBREAKDOWN,
                                    "Lanczos: the Krylov block lost rank (cycle " + std::to_string(cyc) + ", pass " +
                                        std::to_string(pass) + ", column " + std::to_string(fc) + " of " +
                                        std::to_string(nind) + ", pivot ratio " + std::to_string(prel) + ", basis " +
                                        std::to_string(k) + " of " + std::to_string(kd) + ")");
                }
            }
            const int nb = s + 1;
            // K[:, 0:nind] <- K[:, 0:nind] Rp^{-1}, in place (with the next pass's Gram).
            // The second scaling is left pending when this block fills the basis:
            // the thick restart that follows folds it into its own product.
            const std::vector<cd> Ri = inv_upper(nind, Rp);
            if (pass == 1 && !dep_last && k + s >= kd) {
                pend_ri = Ri;
                pend_n = nind;
            } else if (pass == 0 && k > 0) {
                // K[:, 0:nind] <- (K - Q C2)[:, 0:nind] Ri: [[-C2 Ri]; [Ri]] on [Q K]
                const int pr = k + nind;
                std::vector<cd> Mq(size_t(pr) * nind, cd(0));
                for (int j = 0; j < nind; ++j) {
                    for (int i = 0; i < k; ++i) {
                        cd acc = 0;
                        for (int l = 0; l <= j; ++l) acc -= C2[i + size_t(l) * k] * Ri[l + size_t(j) * nind];
                        Mq[i + size_t(j) * pr] = acc;
                    }
                    for (int l = 0; l <= j; ++l) Mq[k + l + size_t(j) * pr] = Ri[l + size_t(j) * nind];
                }
                std::vector<cd> Gn;
                if (chain) {
                    // the same map on [Q K'] with K' = K M1 re-formed from the unchanged K
                    const int pc = k + nb0;
                    std::vector<cd> M2(size_t(pc) * nind, cd(0));
                    for (int j = 0; j < nind; ++j)
                        for (int i = 0; i < pr; ++i) M2[i + size_t(j) * pc] = Mq[i + size_t(j) * pr];
                    rc = blas_tsmm_chain<S>(w, V, N, pc, M1, nb0, k, M2, nind, K, N, N, Gn);
                } else {
                    rc = blas_tsmm<S>(w, V, N, pr, Mq, nind, K, N, N, 2, 0, &Gn);
                }
                if (rc != KFC_OK) return rc;
                G.swap(Gn);
            } else {
                std::vector<cd> Gn;
                rc = blas_tsmm<S>(w, K, N, nind, Ri, nind, K, N, N, pass == 0 ? 2 : 0, 0, pass == 0 ? &Gn : nullptr);
                if (rc != KFC_OK) return rc;
                G.swap(Gn);
            }
            ldg = nind;
            // Rt <- Rp Rt (the dependent column's coefficients come from the partial factor)
            std::vector<cd> Rn(size_t(nb0) * nb0, cd(0));
            for (int j = 0; j < nb; ++j)
                for (int i = 0; i < nind; ++i) {
                    if (pass == 0 && dep_last && j == nind) {
                        Rn[i + size_t(j) * ldr] = dcoef[i];
                        continue;
                    }
                    cd acc = 0;
                    for (int l = i; l < nind; ++l) acc += Rp[i + size_t(l) * nind] * Rt[l + size_t(j) * ldr];
                    Rn[i + size_t(j) * ldr] = acc;
                }
            Rt.swap(Rn);
        }
        KFC_HIP(op->ctx, tm.flush());   // the block's applies finished before the Grams came back
        const int nb = s + 1, n = k + s, rows = k + nb;
        // coordinates in [Q, Qn] (rows x ...): H Q = [Q Qn] Lm, K = [Q Qn] CR, H K[:, 0:s] = K Bh
        std::vector<cd> Lm(size_t(rows) * k, cd(0)), CR(size_t(rows) * nb, cd(0)), Bh(size_t(nb) * s, cd(0));
        for (int j = 0; j < k; ++j) {
            for (int i = 0; i < k; ++i) Lm[i + size_t(j) * rows] = T[i + size_t(j) * k] + Cm[i] * b[j];
            for (int i = 0; i < nb; ++i) Lm[k + i + size_t(j) * rows] = Rt[i] * b[j];
        }
        for (int j = 0; j < nb; ++j) {
            for (int i = 0; i < k; ++i) CR[i + size_t(j) * rows] = Cm[i + size_t(j) * k];
            for (int i = 0; i < nb; ++i) CR[k + i + size_t(j) * rows] = Rt[i + size_t(j) * ldr];
        }
        for (int i = 0; i < s; ++i) {
            Bh[i + size_t(i) * nb] = cc;
            Bh[i + 1 + size_t(i) * nb] = i == 0 ? ee : 0.5 * ee;
            if (i > 0) Bh[i - 1 + size_t(i) * nb] = 0.5 * ee;
        }
        // right block: (CR Bh - Lm Cm[:, 0:s]) Rss^{-1}
        std::vector<cd> Cs(size_t(k) * s);
        for (int j = 0; j < s; ++j)
            for (int i = 0; i < k; ++i) Cs[i + size_t(j) * k] = Cm[i + size_t(j) * k];
        std::vector<cd> Z = mm(rows, nb, s, CR, Bh);
        if (k > 0) {
            const std::vector<cd> LC = mm(rows, k, s, Lm, Cs);
            for (size_t i = 0; i < Z.size(); ++i) Z[i] -= LC[i];
        }
        std::vector<cd> Rss(size_t(s) * s);
        for (int j = 0; j < s; ++j)
            for (int i = 0; i < s; ++i) Rss[i + size_t(j) * s] = Rt[i + size_t(j) * ldr];
        const std::vector<cd> right = mm(rows, s, s, Z, inv_upper(s, Rss));
        // the (k + s) x (k + s) projection and the residual row
        std::vector<cd> Tn(size_t(n) * n), bnew(n);
        for (int j = 0; j < n; ++j) {
            const cd* col = j < k ? &Lm[size_t(j) * rows] : &right[size_t(j - k) * rows];
            for (int i = 0; i < n; ++i) Tn[i + size_t(j) * n] = col[i];
            bnew[j] = dep_last ? cd(0) : col[n];
        }
        std::vector<double> theta;
        std::vector<cd> Y;
        small_eig(n, Tn, theta, Y);   // symmetrizes Tn
        std::vector<double> resid = ritz_residuals(n, bnew, Y);
        int conv = 0;
        while (conv < n && resid[conv] <= tol) ++conv;
        // rounding in the projection grows with the block's condition number
        // (eps |H| cond(Rss) per cycle): long blocks while the wanted residuals are
        // far above the tolerance, short ones for the last digits
        {
            const double far = resid[std::min(m, n) - 1];
            const double kap = cond_upper(s, Rss);
            const double noise = eps_s * std::max(std::fabs(hi), std::fabs(lo)) * kap;
            int want = far > 1e3 * tol ? LZ_SMAX : (far > 3 * tol ? 6 : 3);
            if (noise > far * 1e-1 || kap > 1e5) want = std::max(1, std::min(want, s - 2));
            s_cap = std::min(want, s_limit);
        }
        bool stop = conv >= m || dep_last || (n >= kd && restarts >= maxiter);
        if (stop && pend_n > 0) {   // the explicit check needs the basis itself
            rc = blas_tsmm<S>(w, K, N, pend_n, pend_ri, pend_n, K, N, N, 0, 0, nullptr);
            if (rc != KFC_OK) return rc;
            pend_n = 0;
        }
        if (stop && !dep_last) {
            // confirm with an explicit Rayleigh-Ritz on the basis: W = H V gives the
            // exact projection and the true residuals of the wanted Ritz pairs
            S* W = V2;
            rc = apply_timed<S>(op, V, W, n, N, tm);
            if (rc != KFC_OK) return rc;
            std::vector<cd> Tx, bx;
            rc = blas_gram<S>(w, V, N, n, W, N, n, N, Tx);
            if (rc == KFC_OK) rc = blas_gram<S>(w, V + size_t(n) * N, N, 1, W, N, n, N, bx);
            if (rc != KFC_OK) return rc;
            small_eig(n, Tx, theta, Y);
            const int q = std::min(m, n);
            std::vector<cd> M(size_t(n) * q);
            for (int cI = 0; cI < q; ++cI)
                for (int i = 0; i < n; ++i) M[i + size_t(cI) * n] = Y[i + size_t(cI) * n];
            KFC_HIP(op->ctx, result.ensure(sizeof(S) * size_t(N) * std::max(q, 1)));
            rc = blas_tsmm<S>(w, V, N, n, M, q, result.as<S>(), N, N, 0, 0, nullptr);
            std::vector<double> tr;
            KFC_HIP(op->ctx, bufR.ensure(sizeof(S) * size_t(N) * std::max(q, 1)));
            if (rc == KFC_OK) rc = blas_tsmm<S>(w, W, N, n, M, q, bufR.as<S>(), N, N, 0, 0, nullptr);
            if (rc == KFC_OK) rc = blas_residual<S>(w, bufR.as<S>(), N, bufR.as<S>(), N, result.as<S>(), N, q, N,
                                                  std::vector<double>(theta.begin(), theta.begin() + q), tr);
            if (rc != KFC_OK) return rc;
            // the convergence test reads the residuals of the exact Krylov relation
            // H V = V T + q b^T, as KrylovKit's normres does, and asks the true
            // residuals of the returned vectors to sit within the rounding floor of
            // H v (10 eps max(|H|, |E_1|), src/solvers.jl:75) of the tolerance: a
            // basis whose Krylov relation was corrupted by an ill-conditioned block
            // fails here and goes on with short blocks
            resid = ritz_residuals(n, bx, Y);
            const double floor_ = 10.0 * eps_s * std::max(std::max(std::fabs(hi), std::fabs(lo)), std::fabs(theta[0]));
            const double tbar = 10.0 * std::max(tol, floor_);
            conv = 0;
            while (conv < n && resid[conv] <= tol && (conv >= q || tr[conv] <= tbar)) ++conv;
            bool polluted = false;
            for (int i = 0; i < q; ++i) polluted |= resid[i] <= tol && tr[i] > tbar;
            if (polluted) {
                s_limit = 1;
                relation_exact = false;
            }
            ++anchors;
            out.theta = theta;
            out.resid = resid;
            out.converged = conv;
            if (verbosity > 0)
                fprintf(stderr, "[kfc lanczos] explicit check %d: converged %d of %d, resid[m-1] %.3e, true max %.3e\n",
                        anchors, conv, m, resid[q - 1], *std::max_element(tr.begin(), tr.end()));
            if (conv >= m || (n >= kd && restarts >= maxiter) || anchors >= 16) return KFC_OK;
            if (polluted && fresh_starts < 3) {
                // H V has components outside [V q]: the Krylov relation a thick restart
                // relies on no longer holds, and restarts from it keep that error.  Start
                // afresh from one vector, a combination of the wanted Ritz vectors (all of
                // them present in it), with one-vector blocks (plain Lanczos with full
                // reorthogonalization), so the relation is exact again from its first step
                ++fresh_starts;
                std::vector<cd> comb(q);
                std::mt19937_64 g(seed ^ (0x9e3779b97f4a7c15ull * uint64_t(fresh_starts)));
                std::normal_distribution<double> nd;
                for (int i = 0; i < q; ++i) comb[i] = 1.0 + 0.25 * nd(g);
                rc = blas_gemm<S>(w, result.as<S>(), N, q, comb, 1, V, N, N, 0.0);
                std::vector<double> nrm;
                if (rc == KFC_OK) rc = blas_colnorms<S>(w, V, N, 1, N, nrm);
                if (rc != KFC_OK) return rc;
                if (!(nrm[0] > 0)) return kfc_fail(op->ctx, KFC_EBREAKDOWN, "Lanczos: zero restart vector");
                rc = blas_scale_cols<S>(w, V, N, 1, N, std::vector<double>{1.0 / nrm[0]});
                if (rc != KFC_OK) return rc;
                k = 0;
                T.clear();
                b.clear();
                relation_exact = true;
                s_limit = 1;
                s_cap = 1;
                ++restarts;
                continue;
            }
            // not there yet: continue from the exact decomposition with shorter blocks
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) Tn[i + size_t(j) * n] = 0.5 * (Tx[i + size_t(j) * n] + std::conj(Tx[j + size_t(i) * n]));
            bnew = bx;
            s_limit = std::max(1, std::min(s_limit, s) / 2);
            s_cap = std::min(s_cap, s_limit);
            stop = false;
        }
        out.theta = theta;
        out.resid = resid;
        out.converged = conv;
        out.restarts = int(restarts);
        out.cycles = int(cyc + 1);
        out.invariant = dep_last;
        if (verbosity > 0 && cyc % verbosity == 0)
            fprintf(stderr, "[kfc lanczos] cycle %lld restart %lld matvec %lld s %d converged %d resid[m-1] %.3e theta0 %.12f\n",
                    (long long)cyc, (long long)restarts, (long long)tm.cols, s, conv, resid[std::min(m, n) - 1],
                    theta[0]);
        if (stop) {   // an invariant subspace: its Ritz pairs are exact
            const int q = std::min(m, n);
            std::vector<cd> M(size_t(n) * q);
            for (int cI = 0; cI < q; ++cI)
                for (int i = 0; i < n; ++i) M[i + size_t(cI) * n] = Y[i + size_t(cI) * n];
            KFC_HIP(op->ctx, result.ensure(sizeof(S) * size_t(N) * std::max(q, 1)));
            return blas_gemm<S>(w, V, N, n, M, q, result.as<S>(), N, N, 0.0);
        }
        ritz_known = true;
        lo = theta[0] - std::max(resid[0], 1e-3 * (hi - theta[0]));
        if (n >= kd) {
            // thick restart: keep the lowest Ritz vectors and the residual direction
            int keep = (3 * kd + 2 * conv) / 5;
            keep = std::max(keep, std::min(m + 1, kd - 2));
            keep = std::min(keep, kd - 2);
            // [V2 Ritz vectors, q] = [V q] [[Y_keep, 0]; [0, 1]] in one pass; with the
            // block's last scaling pending, V[:, k:] holds Qn P^{-1} and the rows of the
            // block (and the q column) are mapped through P^{-1} first
            std::vector<cd> M(size_t(n + 1) * (keep + 1), cd(0));
            for (int cI = 0; cI < keep; ++cI)
                for (int i = 0; i < n; ++i) M[i + size_t(cI) * (n + 1)] = Y[i + size_t(cI) * n];
            M[n + size_t(keep) * (n + 1)] = 1.0;
            if (pend_n > 0) {
                const int k0 = n + 1 - pend_n;   // first block column (the block ends with q)
                std::vector<cd> Mn = M;
                for (int cI = 0; cI <= keep; ++cI)
                    for (int l = 0; l < pend_n; ++l) {
                        cd acc = 0;
                        for (int j = l; j < pend_n; ++j)
                            acc += pend_ri[l + size_t(j) * pend_n] * M[k0 + j + size_t(cI) * (n + 1)];
                        Mn[k0 + l + size_t(cI) * (n + 1)] = acc;
                    }
                M.swap(Mn);
                pend_n = 0;
            }
            rc = blas_tsmm<S>(w, V, N, n + 1, M, keep + 1, V2, N, N, 0, 0, nullptr);
            if (rc != KFC_OK) return rc;
            std::swap(V, V2);
            T.assign(size_t(keep) * keep, cd(0));
            b.assign(keep, cd(0));
            for (int i = 0; i < keep; ++i) {
                T[i + size_t(i) * keep] = theta[i];
                cd r = 0;
                for (int j = 0; j < n; ++j) r += bnew[j] * Y[j + size_t(i) * n];
                b[i] = r;
            }
            k = keep;
            ++restarts;
        } else {
            T.assign(size_t(n) * n, cd(0));
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) T[i + size_t(j) * n] = 0.5 * (Tn[i + size_t(j) * n] + std::conj(Tn[j + size_t(i) * n]));
            b = bnew;
            k = n;
        }
    }
}

template <class S>
int lanczos_t(kfc_op* op, int64_t m, int64_t krylovdim, double tol, int64_t maxiter, uint64_t seed, int mode,
              int verbosity, double* Eout, double* Vout, kfc_stats* st) {
    auto t0 = clk::now();
    const int64_t N = op->N;
    if (m < 1) return kfc_fail(op->ctx, KFC_EARG, "Lanczos requires m >= 1, got " + std::to_string(m));
    if (m > N) return kfc_fail(op->ctx, KFC_EARG, "Lanczos: m = " + std::to_string(m) + " exceeds the operator dimension " + std::to_string(N));
    int kd = int(std::max<int64_t>(2 * m + 10, krylovdim));   // src/solvers.jl:92
    kd = int(std::min<int64_t>(kd, N));
    if (kd > 512) return kfc_fail(op->ctx, KFC_EUNSUPPORTED, "Lanczos: krylovdim above 512 is not served");
    // a Krylov basis that can fill (almost) the whole space: its last blocks are
    // dependent by construction, and the Rayleigh-Ritz of the assembled operator is
    // the exact answer the recurrence would reach (N <= 524 here)
    if (int64_t(kd) + LZ_SMAX + 1 >= N) return dense_t<S>(op, m, Eout, Vout, st);
    Work w;
    w.ctx = op->ctx;
    Timer tm;
    DevBuf res64;
    LanczosOut<S> out;
    int rc;
    double glo = 0, hb = 0;
    rc = op_spectral_interval(op, &glo, &hb);   // the Chebyshev interval of the first block
    if (rc != KFC_OK) return rc;
    if (mode == KFC_MODE_MIXED) {
        using L = typename Lower<S>::type;
        // the FP32 floor: a few units of roundoff of the largest eigenvalue scale
        const double tol32 = std::max(tol, 64.0 * 5.96e-8 * std::max(std::fabs(hb), 1.0));
        LanczosOut<L> o32;
        DevBuf res32;
        rc = lanczos_core<L>(op, w, int(m), kd, tol32, maxiter, nullptr, 0, seed, glo, hb, tm, o32, res32, verbosity);
        if (rc != KFC_OK) return rc;
        // FP64 continuation from the FP32 Ritz vectors combined into one start vector
        const int q = int(std::min<int64_t>(m, N));
        DevBuf s64;
        KFC_HIP(op->ctx, s64.ensure(sizeof(S) * size_t(N) * q));
        rc = blas_convert<L, S>(w, res32.as<L>(), N, s64.as<S>(), N, q, N);
        if (rc != KFC_OK) return rc;
        std::vector<cd> comb(q);
        std::mt19937_64 g(seed ^ 0x5eedull);
        std::normal_distribution<double> nd;
        for (int i = 0; i < q; ++i) comb[i] = 1.0 + 0.25 * nd(g);
        DevBuf st64;
        KFC_HIP(op->ctx, st64.ensure(sizeof(S) * size_t(N)));
        rc = blas_gemm<S>(w, s64.as<S>(), N, q, comb, 1, st64.as<S>(), N, N, 0.0);
        if (rc != KFC_OK) return rc;
        // the start vector lies almost in an invariant subspace: a long Chebyshev block
        // on it would be ill-conditioned, so the continuation runs short blocks
        rc = lanczos_core<S>(op, w, int(m), kd, tol, maxiter, st64.as<S>(), N, seed, glo, hb, tm, out, res64, verbosity,
                             2);
        out.restarts += o32.restarts;
    } else {
        rc = lanczos_core<S>(op, w, int(m), kd, tol, maxiter, nullptr, 0, seed, glo, hb, tm, out, res64, verbosity);
    }
    if (rc != KFC_OK) return rc;
    const int n = int(out.theta.size());
    if (n < m || out.converged < m) {
        char nr[64];
        snprintf(nr, sizeof nr, "%.3e", out.resid.empty() ? 0.0 : out.resid[std::min<int64_t>(m, n) - 1]);
        return kfc_fail(op->ctx, KFC_ENOCONV,
                        "Lanczos: only " + std::to_string(std::min(out.converged, int(m))) + " of " + std::to_string(m) +
                            " eigenpairs converged (normres = " + nr + ", after " + std::to_string(out.restarts) +
                            " restarts)");
    }
    // the FP64 (or complex FP64) result block has the caller's host layout: one copy
    KFC_HIP(op->ctx, hipMemcpyAsync(Vout, res64.p, sizeof(S) * size_t(N) * m, hipMemcpyDeviceToHost, op->ctx->stream));
    KFC_HIP(op->ctx, hipStreamSynchronize(op->ctx->stream));
    std::vector<double> E(out.theta.begin(), out.theta.begin() + m);
    for (int64_t j = 0; j < m; ++j) Eout[j] = E[j];
    if (st) {
        std::vector<double> res;
        rc = true_residuals_dev<S>(op, w, E, res64.as<S>(), int(m), res, tm);
        if (rc != KFC_OK) return rc;
        *st = kfc_stats{};
        st->max_residual = *std::max_element(res.begin(), res.end());
        st->iterations = out.restarts;
        st->n_matvec = tm.cols - m;   // the acceptance check is not part of the solve
        st->n_matvec_f32 = tm.cols32;
        st->n_apply_calls = tm.calls;
        st->seconds_apply = tm.apply;
        st->seconds_apply_gpu = tm.apply_gpu;
        st->n_converged = int(m);
        st->seconds_total = secs(t0);
        st->seconds_ortho_gpu = w.tsmm_gpu;
        st->bytes_ortho = w.tsmm_bytes;
        st->n_ortho_calls = w.tsmm_calls;
    }
    return KFC_OK;
}

// ------------------------------------------------------------------------------------------
// LOBPCG with soft locking (Knyazev; basis [X, P, R] kept orthonormal as in
// Hetmaniuk & Lehoucq), wrapped with the reference's seed, guard,
// preconditioner, retry and acceptance rules (src/lobpcg.jl:217-301).

struct LobpcgCfg {
    int m, w;
    double tol;
    int64_t maxiter;
    int precond;
    double eta;
    bool callback = false;   // hand each iteration to the context's callback
    int64_t it0 = 0;         // iterations of an earlier phase (the mixed path's FP32 run)
};

struct LobpcgRes {
    std::vector<double> theta, resid;
    std::vector<cd> X;   // host N x w
    int64_t iterations = 0;
    bool breakdown = false;
    std::string why;
};

// Orthonormalize the columns [c0, c0+nc) of S against [0, c0) (twice) and among
// themselves (eigendecomposition of the Gram, small directions dropped); the
// surviving columns are compacted to [c0, c0+k).  AS follows when given.
template <class S>
int ortho_block(Work& w, S* Sb, int64_t N, int c0, int nc, S* tmp, int& kout, double drop_rel) {
    int rc;
    for (int pass = 0; pass < 2 && c0 > 0; ++pass) {
        std::vector<cd> G;
        rc = blas_gram<S>(w, Sb, N, c0, Sb + size_t(c0) * N, N, nc, N, G);
        if (rc != KFC_OK) return rc;
        for (auto& x : G) x = -x;
        rc = blas_gemm<S>(w, Sb, N, c0, G, nc, Sb + size_t(c0) * N, N, N, 1.0);
        if (rc != KFC_OK) return rc;
    }
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<cd> G;
        rc = blas_gram<S>(w, Sb + size_t(c0) * N, N, nc, Sb + size_t(c0) * N, N, nc, N, G);
        if (rc != KFC_OK) return rc;
        std::vector<double> d(nc);
        for (int i = 0; i < nc; ++i) d[i] = 1.0 / std::sqrt(std::max(G[i + size_t(i) * nc].real(), 1e-300));
        for (int j = 0; j < nc; ++j)
            for (int i = 0; i < nc; ++i) G[i + size_t(j) * nc] *= d[i] * d[j];
        std::vector<double> lam;
        std::vector<cd> Z;
        small_eig(nc, G, lam, Z);
        const double top = lam.empty() ? 0 : lam.back();
        std::vector<int> keep;
        for (int i = nc - 1; i >= 0; --i)
            if (top > 0 && lam[i] > drop_rel * top) keep.push_back(i);
        const int k = int(keep.size());
        std::vector<cd> M(size_t(nc) * k);
        for (int c = 0; c < k; ++c)
            for (int i = 0; i < nc; ++i) M[i + size_t(c) * nc] = d[i] * Z[i + size_t(keep[c]) * nc] / std::sqrt(lam[keep[c]]);
        rc = blas_gemm<S>(w, Sb + size_t(c0) * N, N, nc, M, k, tmp, N, N, 0.0);
        if (rc != KFC_OK) return rc;
        rc = blas_axpby<S>(w, tmp, N, Sb + size_t(c0) * N, N, k, N, 1.0, 0.0);
        if (rc != KFC_OK) return rc;
        nc = k;
        if (k == 0) break;
    }
    kout = nc;
    return KFC_OK;
}

// Generalized small eigenproblem GA c = theta GB c with GB Hermitian positive definite
// (Cholesky reduction); returns false when GB is not positive definite.
bool gen_eig(int n, const std::vector<cd>& GA, const std::vector<cd>& GB, std::vector<double>& th, std::vector<cd>& C) {
    std::vector<cd> L(size_t(n) * n, cd(0));
    for (int j = 0; j < n; ++j) {
        double s = GB[j + size_t(j) * n].real();
        for (int k = 0; k < j; ++k) s -= std::norm(L[j + size_t(k) * n]);
        if (!(s > 0)) return false;
        const double ljj = std::sqrt(s);
        L[j + size_t(j) * n] = ljj;
        for (int i = j + 1; i < n; ++i) {
            cd t = 0.5 * (GB[i + size_t(j) * n] + std::conj(GB[j + size_t(i) * n]));
            for (int k = 0; k < j; ++k) t -= L[i + size_t(k) * n] * std::conj(L[j + size_t(k) * n]);
            L[i + size_t(j) * n] = t / ljj;
        }
    }
    // Linv by forward substitution
    std::vector<cd> Li(size_t(n) * n, cd(0));
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) {
            cd t = (i == c) ? cd(1) : cd(0);
            for (int k = 0; k < i; ++k) t -= L[i + size_t(k) * n] * Li[k + size_t(c) * n];
            Li[i + size_t(c) * n] = t / L[i + size_t(i) * n];
        }
    }
    // G = Li GA Li^H
    std::vector<cd> T1(size_t(n) * n, cd(0)), G(size_t(n) * n, cd(0));
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) {
            const cd a = GA[k + size_t(j) * n];
            for (int i = 0; i < n; ++i) T1[i + size_t(j) * n] += Li[i + size_t(k) * n] * a;
        }
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) {
            const cd b = std::conj(Li[j + size_t(k) * n]);
            for (int i = 0; i < n; ++i) G[i + size_t(j) * n] += T1[i + size_t(k) * n] * b;
        }
    std::vector<cd> Z;
    small_eig(n, G, th, Z);
    // C = Li^H Z
    C.assign(size_t(n) * n, cd(0));
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) {
            const cd z = Z[k + size_t(j) * n];
            for (int i = 0; i < n; ++i) C[i + size_t(j) * n] += std::conj(Li[k + size_t(i) * n]) * z;
        }
    return true;
}

template <class S>
int lobpcg_core(kfc_op* op, Work& w, const LobpcgCfg& cfg, const std::vector<cd>& X0h, const double* Ddev,
                Timer& tm, LobpcgRes& out, double switch_tol, int nperturb = 0,
                const typename Upper<S>::type* X0d = nullptr) {
    const int64_t N = op->N;
    const int W = cfg.w;
    DevBuf sA, sB, aA, aB, tmpb;
    const size_t blk = sizeof(S) * size_t(N) * 3 * W;
    KFC_HIP(op->ctx, sA.ensure(blk));
    KFC_HIP(op->ctx, sB.ensure(blk));
    KFC_HIP(op->ctx, aA.ensure(blk));
    KFC_HIP(op->ctx, aB.ensure(blk));
    KFC_HIP(op->ctx, tmpb.ensure(sizeof(S) * size_t(N) * W));
    S* Sb = sA.as<S>();
    S* S2 = sB.as<S>();
    S* AS = aA.as<S>();
    S* AS2 = aB.as<S>();
    S* tmp = tmpb.as<S>();
    auto col = [&](S* b, int c) { return b + size_t(c) * N; };
    // the initial block: from the device seed (FP64, converted) or from the host
    int rc = X0d ? blas_convert<typename Upper<S>::type, S>(w, X0d, N, Sb, N, W, N)
                 : upload_block<S>(op->ctx, X0h, N, W, Sb, N);
    if (rc == KFC_OK) rc = blas_perturb<S>(w, Sb, N, nperturb, N, 1, 0, nullptr, 0x4b46);
    if (rc != KFC_OK) return rc;
    int kx = 0;
    rc = ortho_block<S>(w, Sb, N, 0, W, tmp, kx, 1e-14);
    if (rc != KFC_OK) return rc;
    if (kx < W) {
        out.breakdown = true;
        out.why = "the initial block is rank deficient";
        return KFC_OK;
    }
    rc = apply_timed<S>(op, Sb, AS, W, N, tm);
    if (rc != KFC_OK) return rc;
    // Rayleigh-Ritz of X
    std::vector<cd> G;
    rc = blas_gram<S>(w, Sb, N, W, AS, N, W, N, G);
    if (rc != KFC_OK) return rc;
    std::vector<double> theta;
    std::vector<cd> Z;
    small_eig(W, G, theta, Z);
    rc = blas_gemm<S>(w, Sb, N, W, Z, W, S2, N, N, 0.0);
    if (rc == KFC_OK) rc = blas_gemm<S>(w, AS, N, W, Z, W, AS2, N, N, 0.0);
    if (rc != KFC_OK) return rc;
    std::swap(Sb, S2);
    std::swap(AS, AS2);
    int np = 0, nlock = 0;
    std::vector<double> res(W, INFINITY);
    std::vector<double> prec_theta;
    bool first_prep = true;
    int64_t it = 0;
    for (;; ++it) {
        const int na = W - nlock;
        // residuals of the active columns, into the R slot
        const int cR = W + np;
        std::vector<double> th_act(theta.begin() + nlock, theta.end()), rn;
        rc = blas_residual<S>(w, col(Sb, cR), N, col(AS, nlock), N, col(Sb, nlock), N, na, N, th_act, rn);
        if (rc != KFC_OK) return rc;
        for (int j = 0; j < na; ++j) res[nlock + j] = rn[j];
        for (double r : rn)
            if (!std::isfinite(r)) {
                out.breakdown = true;
                out.why = "non-finite residual";
                return KFC_OK;
            }
        int newly = 0;
        while (nlock + newly < W && res[nlock + newly] <= cfg.tol) ++newly;
        out.iterations = it;
        if (cfg.callback) {
            // info after the locking count, lambdas of the columns active during the
            // iteration (src/lobpcg.jl:5-8)
            kfc_iter_info info{};
            info.solver = 0;
            info.precision = prec_of<S>();
            info.iteration = cfg.it0 + it + 1;
            info.n_matvec = tm.cols;
            info.width = W;
            info.n_locked = nlock + newly;
            info.active0 = nlock;
            info.lambda = theta.data();
            info.residuals = res.data();
            info.lambda_cut = NAN;
            info.lambda_max = NAN;
            rc = ctx_callback(op->ctx, info, "LOBPCG");
            if (rc != KFC_OK) return rc;
        }
        bool done = nlock + newly >= W;
        double worst = 0;
        for (int j = nlock; j < W; ++j) worst = std::max(worst, res[j]);
        if (done || it >= cfg.maxiter || (switch_tol > 0 && worst <= switch_tol)) {
            nlock += newly;
            break;
        }
        // drop the newly locked residual columns (they are a prefix of the active ones)
        if (newly > 0) {
            rc = blas_axpby<S>(w, col(Sb, cR + newly), N, tmp, N, na - newly, N, 1.0, 0.0);
            if (rc == KFC_OK) rc = blas_axpby<S>(w, tmp, N, col(Sb, cR), N, na - newly, N, 1.0, 0.0);
            if (rc != KFC_OK) return rc;
        }
        nlock += newly;
        const int nact = W - nlock;
        // preconditioner (precondprep! + ldiv!, src/lobpcg.jl:98-119)
        if (cfg.precond == KFC_PRECOND_RITZ) {
            if (first_prep) {
                rc = blas_diag_expect<S>(w, col(Sb, nlock), N, nact, N, Ddev, prec_theta);
                if (rc != KFC_OK) return rc;
                first_prep = false;
            } else {
                prec_theta.assign(theta.begin() + nlock, theta.end());
            }
            rc = blas_precond<S>(w, col(Sb, cR), N, nact, N, Ddev, prec_theta, cfg.eta);
        } else if (cfg.precond == KFC_PRECOND_JACOBI) {
            rc = blas_precond<S>(w, col(Sb, cR), N, nact, N, Ddev + N, {}, -1.0);
        }
        if (rc != KFC_OK) return rc;
        int nr = 0;
        rc = ortho_block<S>(w, Sb, N, cR, nact, tmp, nr, 1e-12);
        if (rc != KFC_OK) return rc;
        if (nr == 0) {
            out.breakdown = true;
            out.why = "the preconditioned residual block vanished";
            return KFC_OK;
        }
        rc = apply_timed<S>(op, col(Sb, cR), col(AS, cR), nr, N, tm);
        if (rc != KFC_OK) return rc;
        const int ns = W + np + nr;
        std::vector<cd> GA, GB;
        rc = blas_gram<S>(w, Sb, N, ns, AS, N, ns, N, GA);
        if (rc == KFC_OK) rc = blas_gram<S>(w, Sb, N, ns, Sb, N, ns, N, GB);
        if (rc != KFC_OK) return rc;
        std::vector<double> th_all;
        std::vector<cd> C;
        int nsu = ns;
        if (!gen_eig(ns, GA, GB, th_all, C)) {
            // drop P and retry with [X, R]
            if (np == 0) {
                out.breakdown = true;
                out.why = "the Rayleigh-Ritz Gram matrix is not positive definite";
                return KFC_OK;
            }
            rc = blas_axpby<S>(w, col(Sb, cR), N, tmp, N, nr, N, 1.0, 0.0);
            if (rc == KFC_OK) rc = blas_axpby<S>(w, tmp, N, col(Sb, W), N, nr, N, 1.0, 0.0);
            if (rc == KFC_OK) rc = blas_axpby<S>(w, col(AS, cR), N, tmp, N, nr, N, 1.0, 0.0);
            if (rc == KFC_OK) rc = blas_axpby<S>(w, tmp, N, col(AS, W), N, nr, N, 1.0, 0.0);
            if (rc != KFC_OK) return rc;
            np = 0;
            nsu = W + nr;
            std::vector<cd> ga2(size_t(nsu) * nsu), gb2(size_t(nsu) * nsu);
            rc = blas_gram<S>(w, Sb, N, nsu, AS, N, nsu, N, ga2);
            if (rc == KFC_OK) rc = blas_gram<S>(w, Sb, N, nsu, Sb, N, nsu, N, gb2);
            if (rc != KFC_OK) return rc;
            if (!gen_eig(nsu, ga2, gb2, th_all, C)) {
                out.breakdown = true;
                out.why = "the Rayleigh-Ritz Gram matrix is not positive definite";
                return KFC_OK;
            }
        }
        // new X: lowest W Ritz vectors; new P: the active ones' components outside X,
        // orthonormalized in coefficient space against the new X coefficients
        std::vector<cd> M(size_t(nsu) * (2 * W), cd(0));
        for (int c = 0; c < W; ++c)
            for (int i = 0; i < nsu; ++i) M[i + size_t(c) * nsu] = C[i + size_t(c) * nsu];
        // GB-orthonormal coefficient vectors for P: zero the X rows of the active columns,
        // then project out the new X coefficients and orthonormalize (in the GB inner product)
        std::vector<std::vector<cd>> pcols;
        // With S nearly orthonormal, use the Euclidean inner product of coefficients weighted by GB.
        std::vector<cd> GBu(size_t(nsu) * nsu);
        if (nsu == ns) GBu = GB;
        else {
            std::vector<cd> g2;
            rc = blas_gram<S>(w, Sb, N, nsu, Sb, N, nsu, N, g2);
            if (rc != KFC_OK) return rc;
            GBu = g2;
        }
        // GB-products of the vectors the P columns are orthogonalized against, so
        // that each inner product a^H GB v costs O(nsu) (u = GB v is updated along)
        auto gbmul = [&](const cd* x, std::vector<cd>& y) {
            y.assign(nsu, cd(0));
            for (int j = 0; j < nsu; ++j) {
                const cd xj = x[j];
                if (xj == cd(0)) continue;
                for (int i = 0; i < nsu; ++i) y[i] += GBu[i + size_t(j) * nsu] * xj;
            }
        };
        auto dot = [&](const cd* a, const cd* b) {
            cd t = 0;
            for (int i = 0; i < nsu; ++i) t += std::conj(a[i]) * b[i];
            return t;
        };
        std::vector<std::vector<cd>> GX(W), gpcols;
        for (int q = 0; q < W; ++q) gbmul(&C[size_t(q) * nsu], GX[q]);
        int npn = 0;
        for (int c = nlock; c < W; ++c) {
            std::vector<cd> v(nsu, cd(0)), u;
            for (int i = W; i < nsu; ++i) v[i] = C[i + size_t(c) * nsu];
            gbmul(v.data(), u);
            for (int pass = 0; pass < 2; ++pass) {
                for (int q = 0; q < W; ++q) {
                    const cd* x = &C[size_t(q) * nsu];
                    const cd sx = dot(x, u.data());
                    for (int i = 0; i < nsu; ++i) {
                        v[i] -= sx * x[i];
                        u[i] -= sx * GX[q][i];
                    }
                }
                for (size_t k = 0; k < pcols.size(); ++k) {
                    const cd sx = dot(pcols[k].data(), u.data());
                    for (int i = 0; i < nsu; ++i) {
                        v[i] -= sx * pcols[k][i];
                        u[i] -= sx * gpcols[k][i];
                    }
                }
            }
            const double nv = std::sqrt(std::max(dot(v.data(), u.data()).real(), 0.0));
            if (nv < 1e-10) continue;
            for (auto& x : v) x /= nv;
            for (auto& x : u) x /= nv;
            pcols.push_back(v);
            gpcols.push_back(u);
            ++npn;
        }
        for (int c = 0; c < npn; ++c)
            for (int i = 0; i < nsu; ++i) M[i + size_t(W + c) * nsu] = pcols[c][i];
        M.resize(size_t(nsu) * (W + npn));
        rc = blas_gemm<S>(w, Sb, N, nsu, M, W + npn, S2, N, N, 0.0);
        if (rc == KFC_OK) rc = blas_gemm<S>(w, AS, N, nsu, M, W + npn, AS2, N, N, 0.0);
        if (rc != KFC_OK) return rc;
        std::swap(Sb, S2);
        std::swap(AS, AS2);
        np = npn;
        theta.assign(th_all.begin(), th_all.begin() + W);
    }
    out.theta = theta;
    out.resid = res;
    // final residuals of every column (fresh from AX)
    std::vector<double> rn;
    rc = blas_residual<S>(w, tmp, N, AS, N, Sb, N, W, N, theta, rn);
    if (rc != KFC_OK) return rc;
    out.resid = rn;
    return download_block<S>(op->ctx, Sb, N, N, W, out.X);
}

void random_block(std::vector<cd>& X, int64_t N, int p, bool cplx, uint64_t seed) {
    std::mt19937_64 g(seed);
    std::normal_distribution<double> nd;
    X.resize(size_t(N) * p);
    for (auto& x : X) x = cplx ? cd(nd(g), nd(g)) : cd(nd(g), 0);
}

template <class S>
int lobpcg_t(kfc_op* op, int64_t m, int64_t guard, double tol, int64_t maxiter, int precond, double eta, int seed_kind,
             const double* X0, uint64_t seed, int mode, double* Eout, double* Vout, kfc_stats* st,
             bool use_callback = false) {
    auto t0 = clk::now();
    const int64_t N = op->N;
    if (m < 1) return kfc_fail(op->ctx, KFC_EARG, "LOBPCG requires m >= 1, got " + std::to_string(m));
    if (N < 3 * m + 5)
        return kfc_fail(op->ctx, KFC_EARG, "LOBPCG needs N >= 3m + 5, got N = " + std::to_string(N) + " for m = " +
                                               std::to_string(m) + "; the block eigensolver is unreliable on so small a problem, use Dense()");
    if (!(tol > 0) || !std::isfinite(tol)) return kfc_fail(op->ctx, KFC_EARG, "LOBPCG: tol must be positive and finite");
    if (maxiter < 1) return kfc_fail(op->ctx, KFC_EARG, "LOBPCG: maxiter must be at least 1");
    if (precond == KFC_PRECOND_RITZ && !(eta > 0 && std::isfinite(eta)))
        return kfc_fail(op->ctx, KFC_EARG, "eta must be positive and finite");
    if (guard < 0) return kfc_fail(op->ctx, KFC_EARG, "LOBPCG: guard must be nonnegative");
    if (seed_kind == KFC_SEED_WARM_GUARDS && !X0)
        return kfc_fail(op->ctx, KFC_EARG, "LOBPCG: the warm-guards seed needs an X0 block");
    const int g = int(std::max<int64_t>(0, std::min<int64_t>(guard, (N - 5) / 3 - m)));   // src/lobpcg.jl:243
    const int W = int(m) + g;
    // X0 holds m columns, or m + g (the guards too) with the warm-guards seed
    const int64_t nx0 = seed_kind == KFC_SEED_WARM_GUARDS ? W : m;
    std::vector<cd> X0h;
    if (X0) {
        int rc = lobpcg_check_x0(op, X0, nx0, X0h);
        if (rc != KFC_OK) return rc;
    }
    // preconditioner diagonal, formed on the device: D (and 1/(D - sigma) for :jacobi)
    DevBuf dD;
    if (precond != KFC_PRECOND_NONE) {
        int rc = op_precond_diag_dev(op, precond == KFC_PRECOND_JACOBI, dD);
        if (rc != KFC_OK) return rc;
    }
    // the seed block in device memory: the product seed is formed there, warm starts
    // and random blocks are uploaded
    DevBuf dX;
    KFC_HIP(op->ctx, dX.ensure(sizeof(S) * size_t(N) * W));
    int nperturb = 0;
    auto seed_block = [&](bool random_only, uint64_t s) {
        return lobpcg_seed_dev<S>(op, int(m), W, X0h, X0 ? nx0 : 0, seed_kind == KFC_SEED_PRODUCT, random_only, s,
                                  dX.as<S>(), nperturb);
    };
    Work w;
    w.ctx = op->ctx;
    Timer tm;
    LobpcgCfg cfg{int(m), W, tol, maxiter, precond, eta};
    cfg.callback = use_callback && op->ctx->cb != nullptr;
    const double* Dp = dD.as<double>();
    LobpcgRes res;
    bool retried = false;
    const std::vector<cd> none;
    auto run = [&](int rc) -> int {
        if (rc != KFC_OK) return rc;
        res = LobpcgRes();
        if (mode == KFC_MODE_MIXED) {
            using L = typename Lower<S>::type;
            double hb = 0;
            kfc_op_gershgorin(op, &hb);
            const double sw = std::max(tol, 256.0 * 5.96e-8 * std::max(std::fabs(hb), 1.0));
            LobpcgRes r32;
            rc = lobpcg_core<L>(op, w, cfg, none, Dp, tm, r32, sw, nperturb, dX.as<S>());
            if (rc != KFC_OK) return rc;
            if (r32.breakdown) {
                res = r32;
                return KFC_OK;
            }
            LobpcgCfg c2 = cfg;
            c2.maxiter = std::max<int64_t>(1, cfg.maxiter - r32.iterations);
            c2.it0 = r32.iterations + 1;
            rc = lobpcg_core<S>(op, w, c2, r32.X, Dp, tm, res, 0.0);
            res.iterations += r32.iterations;
            return rc;
        }
        return lobpcg_core<S>(op, w, cfg, none, Dp, tm, res, 0.0, nperturb, dX.as<S>());
    };
    int rc = run(seed_block(seed_kind == KFC_SEED_RANDOM && !X0, seed));
    if (rc != KFC_OK) return rc;
    if (res.breakdown) {
        // one retry from a random block (src/lobpcg.jl:272-289)
        if (!X0 && seed_kind != KFC_SEED_PRODUCT)
            return kfc_fail(op->ctx, KFC_EBREAKDOWN, "LOBPCG: the eigensolver failed (" + res.why +
                                                         "). An exactly zero residual block can break its orthogonalization; use Dense() for such operators");
        retried = true;
        rc = run(seed_block(true, seed + 1));
        if (rc != KFC_OK) return rc;
        if (res.breakdown)
            return kfc_fail(op->ctx, KFC_EBREAKDOWN, "LOBPCG: the eigensolver failed (" + res.why +
                                                         "). An exactly zero residual block can break its orthogonalization; use Dense() for such operators");
    }
    int bad = 0;
    double worst = 0;
    for (int64_t j = 0; j < m; ++j) {
        if (!(res.resid[j] <= tol)) ++bad;
        worst = std::max(worst, res.resid[j]);
    }
    if (bad > 0)
        return kfc_fail(op->ctx, KFC_ENOCONV, "LOBPCG: " + std::to_string(bad) + " of " + std::to_string(m) +
                                                  " eigenpairs did not converge within " + std::to_string(maxiter) +
                                                  " iterations (only " + std::to_string(m - bad) +
                                                  " converged; maximum residual " + std::to_string(worst) +
                                                  " > tol = " + std::to_string(tol) + ")");
    for (int64_t j = 0; j < m; ++j) Eout[j] = res.theta[j];
    std::vector<cd> Vh(res.X.begin(), res.X.begin() + size_t(N) * m);
    write_host_V(op, Vh, Vout);
    if (st) {
        std::vector<double> E(res.theta.begin(), res.theta.begin() + m), tr;
        const int64_t cols_solve = tm.cols;
        rc = true_residuals<S>(op, w, E, Vh, int(m), tr, tm);
        if (rc != KFC_OK) return rc;
        *st = kfc_stats{};
        st->iterations = res.iterations;
        st->n_matvec = cols_solve;
        st->n_matvec_f32 = tm.cols32;
        st->n_apply_calls = tm.calls;
        st->seconds_apply = tm.apply;
        st->seconds_apply_gpu = tm.apply_gpu;
        st->max_residual = *std::max_element(tr.begin(), tr.end());
        st->n_converged = int(m);
        st->warm = X0 ? 1 : 0;
        st->retried = retried ? 1 : 0;
        st->seconds_total = secs(t0);
    }
    return KFC_OK;
}

template <class S>
int residuals_t(kfc_op* op, int64_t m, const double* E, const double* V, double* out) {
    Work w;
    w.ctx = op->ctx;
    Timer tm;
    std::vector<cd> Vh;
    read_host_V(op, V, size_t(op->N) * m, Vh);
    std::vector<double> Ev(E, E + m), r;
    int rc = true_residuals<S>(op, w, Ev, Vh, int(m), r, tm);
    if (rc != KFC_OK) return rc;
    std::copy(r.begin(), r.end(), out);
    return KFC_OK;
}

}  // namespace

// X0 (prod(dims) x nx0, the caller's host layout) read and checked as
// lobpcg_solve checks it (src/lobpcg.jl:226-233).
int lobpcg_check_x0(kfc_op* op, const double* X0, int64_t nx0, std::vector<cd>& X0h) {
    const int64_t N = op->N;
    read_host_V(op, X0, size_t(N) * nx0, X0h);
    for (const cd& x : X0h)
        if (!std::isfinite(x.real()) || !std::isfinite(x.imag()))
            return kfc_fail(op->ctx, KFC_EARG, "LOBPCG: X0 has non-finite entries");
    for (int64_t j = 0; j < nx0; ++j) {
        bool z = true;
        for (int64_t i = 0; i < N; ++i) z &= X0h[i + j * N] == cd(0);
        if (z) return kfc_fail(op->ctx, KFC_EARG, "LOBPCG: column " + std::to_string(j + 1) + " of X0 is zero");
    }
    return KFC_OK;
}

// The initial block, N x W (_seed_columns, src/lobpcg.jl:182-190): the warm start
// (nx0 = m, or W when it carries its guards) or the product seed; the guards of an
// m-column warm start from the product seed (else random); a random block when
// neither exists or random_only (the retry).  The seeded columns (the first
// nperturb) are perturbed on the device after upload (blas_perturb, seed 0x4b46).
// The initial LOBPCG block of W columns in device memory (src/lobpcg.jl:236-262): the
// caller's X0 (m columns, or W with the guards), the product seed's guard columns or
// random ones after a partial X0, the product seed, or a random block.  nperturb is
// the number of leading columns the seed perturbation applies to.
template <class S>
int lobpcg_seed_dev(kfc_op* op, int m, int W, const std::vector<cd>& X0h, int64_t nx0, bool product, bool random_only,
                    uint64_t s, S* Xd, int& nperturb) {
    const int64_t N = op->N;
    const int g = W - m;
    nperturb = 0;
    std::vector<cd> R;
    if (random_only) {
        random_block(R, N, W, op->cplx, s);
        return upload_block<S>(op->ctx, R, N, W, Xd, N);
    }
    if (nx0 == W) {   // the warm start carries its guards
        nperturb = W;
        return upload_block<S>(op->ctx, X0h, N, W, Xd, N);
    }
    if (nx0 > 0) {
        nperturb = m;
        int rc = KFC_OK;
        if (product && g > 0) {
            rc = op_product_seed_dev<S>(op, W, Xd, N);
            nperturb = W;
        } else if (g > 0) {
            random_block(R, N, g, op->cplx, s);
            rc = upload_block<S>(op->ctx, R, N, g, Xd + size_t(m) * N, N);
        }
        if (rc != KFC_OK) return rc;
        return upload_block<S>(op->ctx, X0h, N, m, Xd, N);
    }
    if (product) {
        nperturb = W;
        return op_product_seed_dev<S>(op, W, Xd, N);
    }
    random_block(R, N, W, op->cplx, s);
    return upload_block<S>(op->ctx, R, N, W, Xd, N);
}


template int lobpcg_seed_dev<double>(kfc_op*, int, int, const std::vector<cd>&, int64_t, bool, bool, uint64_t, double*,
                                     int&);
template int lobpcg_seed_dev<Cx<double>>(kfc_op*, int, int, const std::vector<cd>&, int64_t, bool, bool, uint64_t,
                                         Cx<double>*, int&);

void lobpcg_seed_block(const kfc_op* op, int m, int W, const std::vector<cd>& X0h, int64_t nx0,
                       const std::vector<cd>& P, bool random_only, uint64_t s, std::vector<cd>& X, int& nperturb) {
    const int64_t N = op->N;
    const int g = W - m;
    X.assign(size_t(N) * W, cd(0));
    nperturb = 0;
    if (random_only) {
        random_block(X, N, W, op->cplx, s);
        return;
    }
    if (nx0 == W) {   // the warm start carries its guards
        X = X0h;
        nperturb = W;
    } else if (nx0 > 0) {
        std::copy(X0h.begin(), X0h.begin() + size_t(N) * m, X.begin());
        nperturb = m;
        if (!P.empty() && g > 0) {
            std::copy(P.begin() + size_t(m) * N, P.end(), X.begin() + size_t(m) * N);
            nperturb = W;
        } else if (g > 0) {
            std::vector<cd> R;
            random_block(R, N, g, op->cplx, s);
            std::copy(R.begin(), R.end(), X.begin() + size_t(m) * N);
        }
    } else if (!P.empty()) {
        X = P;
        nperturb = W;
    } else {
        random_block(X, N, W, op->cplx, s);
    }
}

int lobpcg_solve(kfc_op* op, int64_t m, int64_t guard, double tol, int64_t maxiter, int32_t precond, double eta,
                 int32_t seed_kind, const double* X0, uint64_t seed, int32_t mode, double* E, double* V,
                 kfc_stats* stats, bool use_callback) {
    if (!op || !E || !V) return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_lobpcg: null argument");
    return op->cplx ? lobpcg_t<Cx<double>>(op, m, guard, tol, maxiter, precond, eta, seed_kind, X0, seed, mode, E, V,
                                           stats, use_callback)
                    : lobpcg_t<double>(op, m, guard, tol, maxiter, precond, eta, seed_kind, X0, seed, mode, E, V, stats,
                                       use_callback);
}

extern "C" {

int kfc_dense_batch(int32_t n, kfc_op* const* ops, int64_t m, double* E, double* const* Vs, kfc_stats* stats) {
    if (n < 1 || !ops || !E || !Vs) return kfc_fail(nullptr, KFC_EARG, "kfc_dense_batch: empty batch or null argument");
    for (int i = 0; i < n; ++i) {
        if (!ops[i] || !Vs[i]) return kfc_fail(ops[0] ? ops[0]->ctx : nullptr, KFC_EARG, "kfc_dense_batch: null operator or output");
        if (m < 1 || m > ops[i]->N)
            return kfc_fail(ops[i]->ctx, KFC_EARG, "kfc_dense_batch: m must lie in 1:prod(dims) of every operator");
        if (ops[i]->N > 8192)
            return kfc_fail(ops[i]->ctx, KFC_EUNSUPPORTED, "Dense: prod(dims) above 8192 is not served; use an iterative solver");
    }
    return dense_batch(n, ops, m, E, Vs, stats);
}

int kfc_dense(kfc_op* op, int64_t m, double* E, double* V, kfc_stats* stats) {
    if (!op || !E || !V) return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_dense: null argument");
    if (m < 0) return kfc_fail(op->ctx, KFC_EARG, "Dense requires m >= 0, got " + std::to_string(m));
    return op->cplx ? dense_t<Cx<double>>(op, m, E, V, stats) : dense_t<double>(op, m, E, V, stats);
}

int kfc_lanczos(kfc_op* op, int64_t m, int64_t krylovdim, double tol, int64_t maxiter, uint64_t seed, int32_t mode,
                int32_t verbosity, double* E, double* V, kfc_stats* stats) {
    if (!op || !E || !V) return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_lanczos: null argument");
    return op->cplx ? lanczos_t<Cx<double>>(op, m, krylovdim, tol, maxiter, seed, mode, verbosity, E, V, stats)
                    : lanczos_t<double>(op, m, krylovdim, tol, maxiter, seed, mode, verbosity, E, V, stats);
}

int kfc_lobpcg(kfc_op* op, int64_t m, int64_t guard, double tol, int64_t maxiter, int32_t precond, double eta,
               int32_t seed_kind, const double* X0, uint64_t seed, int32_t mode, double* E, double* V,
               kfc_stats* stats) {
    return lobpcg_solve(op, m, guard, tol, maxiter, precond, eta, seed_kind, X0, seed, mode, E, V, stats, true);
}

int kfc_residuals(kfc_op* op, int64_t m, const double* E, const double* V, double* resid) {
    if (!op || !E || !V || !resid) return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_residuals: null argument");
    return op->cplx ? residuals_t<Cx<double>>(op, m, E, V, resid) : residuals_t<double>(op, m, E, V, resid);
}

}  // extern "C"

namespace {

// G = V^H (op V) for a device block V (k columns, leading dimension ld): the
// operator applied into a scratch block, then one Gram on the matrix cores.
template <class S>
int sandwich_t(kfc_op* op, const void* V, int64_t k, int64_t ld, double* G) {
    const int64_t N = op->N;
    Work w;
    w.ctx = op->ctx;
    DevBuf dW;
    KFC_HIP(op->ctx, dW.ensure(sizeof(S) * size_t(ld) * k));
    int rc = op_apply(op, V, dW.p, k, ld, KFC_PREC_F64, op->ctx->stream);
    if (rc != KFC_OK) return rc;
    std::vector<cd> Gh;
    rc = blas_gram<S>(w, static_cast<const S*>(V), ld, int(k), dW.as<S>(), ld, int(k), N, Gh);
    if (rc != KFC_OK) return rc;
    KFC_HIP(op->ctx, hipStreamSynchronize(op->ctx->stream));
    for (size_t i = 0; i < Gh.size(); ++i) {
        if (IsCx<S>::value) {
            G[2 * i] = Gh[i].real();
            G[2 * i + 1] = Gh[i].imag();
        } else {
            G[i] = Gh[i].real();
        }
    }
    return KFC_OK;
}

}  // namespace

// The projected operator V^H H V of a device block; with H a single product term
// (the node's port operator on its axes) this is _port_matrix / _joint_port
// (src/materialize.jl:441-444, 490-493).
extern "C" int kfc_sandwich(kfc_op* op, const void* V, int64_t k, int64_t ld, double* G) {
    if (!op || !V || !G) return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_sandwich: null argument");
    if (k < 1 || k > 1024) return kfc_fail(op->ctx, KFC_EARG, "kfc_sandwich: k must lie in 1:1024, got " + std::to_string(k));
    if (ld < op->N) return kfc_fail(op->ctx, KFC_EARG, "kfc_sandwich: ld below prod(dims)");
    return op->cplx ? sandwich_t<Cx<double>>(op, V, k, ld, G) : sandwich_t<double>(op, V, k, ld, G);
}

// Independent LOBPCG solves of several operators, one host thread per distinct
// context: each context's stream carries its operators' solves in order, and the
// contexts' streams run concurrently on the device (a parameter sweep).
extern "C" int kfc_lobpcg_batch(int32_t n, kfc_op* const* ops, int64_t m, int64_t guard, double tol, int64_t maxiter,
                                int32_t precond, double eta, int32_t seed_kind, const double* const* X0s,
                                uint64_t seed, int32_t mode, double* E, double* const* Vs, kfc_stats* stats) {
    if (n < 1 || !ops || !E || !Vs) return kfc_fail(n > 0 && ops && ops[0] ? ops[0]->ctx : nullptr, KFC_EARG,
                                                    "kfc_lobpcg_batch: bad arguments");
    std::vector<kfc_ctx*> ctxs;
    for (int i = 0; i < n; ++i) {
        if (!ops[i]) return kfc_fail(ops[0]->ctx, KFC_EARG, "kfc_lobpcg_batch: null operator");
        if (std::find(ctxs.begin(), ctxs.end(), ops[i]->ctx) == ctxs.end()) ctxs.push_back(ops[i]->ctx);
    }
    std::vector<int> rcs(n, KFC_OK);
    std::vector<std::thread> th;
    for (kfc_ctx* c : ctxs) {
        th.emplace_back([&, c]() {
            (void)hipSetDevice(c->device);
            for (int i = 0; i < n; ++i) {
                if (ops[i]->ctx != c) continue;
                kfc_op* op = ops[i];
                rcs[i] = op->cplx
                             ? lobpcg_t<Cx<double>>(op, m, guard, tol, maxiter, precond, eta, seed_kind,
                                                    X0s ? X0s[i] : nullptr, seed + i, mode, E + i * m, Vs[i],
                                                    stats ? stats + i : nullptr)
                             : lobpcg_t<double>(op, m, guard, tol, maxiter, precond, eta, seed_kind,
                                                X0s ? X0s[i] : nullptr, seed + i, mode, E + i * m, Vs[i],
                                                stats ? stats + i : nullptr);
            }
        });
    }
    for (auto& t : th) t.join();
    for (int i = 0; i < n; ++i)
        if (rcs[i] != KFC_OK) {
            std::string msg = "instance " + std::to_string(i) + ": " + ops[i]->ctx->err;
            return kfc_fail(ops[0]->ctx, rcs[i], msg);
        }
    return KFC_OK;
}
