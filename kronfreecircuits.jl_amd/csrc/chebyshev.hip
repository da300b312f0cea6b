// Chebyshev-filtered subspace iteration (Zhou-Saad ChFSI) on the device, with
// the reference's locking, guard, degree-cap, SVQB/Cholesky-QR and acceptance
// rules (src/chebyshev.jl:1-406).  The three-term recurrence step
// C = (2/e)(H B - c B) - A runs inside the apply kernel's epilogue.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <random>

#include "blas.h"


namespace {

using clk = std::chrono::steady_clock;
inline double secs(clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); }

constexpr double SVQB_CUTOFF = 1e-12;       // src/chebyshev.jl:8
constexpr double LOCK_FRACTION = 0.5;       // src/chebyshev.jl:12
constexpr double SEED_ADMIXTURE = 1e-2;     // src/chebyshev.jl:195
constexpr int CHEB_GUARDS = 4;              // src/chebyshev.jl:202
constexpr double GUARD_GAIN_FRACTION = 0.1; // src/chebyshev.jl:226
constexpr double EPS = 2.220446049250313e-16;

template <class S> constexpr int prec_of() { return std::is_same<typename RealOf<S>::type, double>::value ? 0 : 1; }
template <class S> struct Lower { using type = S; };
template <> struct Lower<double> { using type = float; };
template <> struct Lower<Cx<double>> { using type = Cx<float>; };
template <class S> struct Upper { using type = S; };
template <> struct Upper<float> { using type = double; };
template <> struct Upper<Cx<float>> { using type = Cx<double>; };

struct CTimer {
    double apply_gpu = 0, apply = 0;
    int64_t cols = 0, cols32 = 0, calls = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CTimer() { (void)hipEventCreate(&e0); (void)hipEventCreate(&e1); }
    ~CTimer() { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); }
};

// C = a (H B) + b B + g A on nc columns (A may be null)
template <class S>
int cheb_apply(kfc_op* op, const S* B, const S* A, S* C, int nc, double a, double b, double g, CTimer& t) {
    if (nc == 0) return KFC_OK;
    auto t0 = clk::now();
    KFC_HIP(op->ctx, hipEventRecord(t.e0, op->ctx->stream));
    int rc = op_apply_cheb(op, B, A, C, nc, op->N, prec_of<S>(), a, b, g, op->ctx->stream);
    if (rc != KFC_OK) return rc;
    KFC_HIP(op->ctx, hipEventRecord(t.e1, op->ctx->stream));
    KFC_HIP(op->ctx, hipEventSynchronize(t.e1));
    float ms = 0;
    KFC_HIP(op->ctx, hipEventElapsedTime(&ms, t.e0, t.e1));
    t.apply_gpu += 1e-3 * ms;
    t.apply += secs(t0);
    t.cols += nc;
    if (prec_of<S>() == 1) t.cols32 += nc;
    t.calls += 1;
    return KFC_OK;
}

void small_herm_eig(int n, const std::vector<cd>& G, std::vector<double>& w, std::vector<cd>& Z) {
    std::vector<cd> A(size_t(n) * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) A[i + size_t(j) * n] = 0.5 * (G[i + size_t(j) * n] + std::conj(G[j + size_t(i) * n]));
    herm_eig(n, A, w, Z);
}

template <class S>
struct Cheb {
    kfc_op* op;
    Work w;
    CTimer tm;
    int64_t N;
    std::mt19937_64 rng;
    double cutoff = SVQB_CUTOFF;   // relative Gram eigenvalue cutoff (raised for FP32 storage)
    S* col(S* b, int c) { return b + size_t(c) * N; }

    // Z <- Z - L (L^H Z), twice (_project_out!, src/chebyshev.jl:57-63)
    int project_out(S* Z, int nz, const S* L, int nl) {
        if (nl == 0 || nz == 0) return KFC_OK;
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<cd> G;
            int rc = blas_gram<S>(w, L, N, nl, Z, N, nz, N, G);
            if (rc != KFC_OK) return rc;
            for (auto& x : G) x = -x;
            rc = blas_gemm<S>(w, L, N, nl, G, nz, Z, N, N, 1.0);
            if (rc != KFC_OK) return rc;
        }
        return KFC_OK;
    }

    // _normalize_columns! (src/chebyshev.jl:18-21)
    int normalize(S* Z, int nz) {
        std::vector<double> nrm;
        int rc = blas_colnorms<S>(w, Z, N, nz, N, nrm);
        if (rc != KFC_OK) return rc;
        for (double& x : nrm) x = 1.0 / std::max(x, 2.2250738585072014e-308);
        return blas_scale_cols<S>(w, Z, N, nz, N, nrm);
    }

    // _svqb_pass! (src/chebyshev.jl:68-81): returns k
    int svqb_pass(S* W, const S* Z, int nz, int& k) {
        k = 0;
        if (nz == 0) return KFC_OK;
        std::vector<cd> G;
        int rc = blas_gram<S>(w, Z, N, nz, Z, N, nz, N, G);
        if (rc != KFC_OK) return rc;
        std::vector<double> d(nz);
        for (int i = 0; i < nz; ++i) d[i] = 1.0 / std::sqrt(std::max(G[i + size_t(i) * nz].real(), 2.2250738585072014e-308));
        std::vector<cd> Gs(size_t(nz) * nz);
        for (int j = 0; j < nz; ++j)
            for (int i = 0; i < nz; ++i) Gs[i + size_t(j) * nz] = d[i] * G[i + size_t(j) * nz] * d[j];
        std::vector<double> lam;
        std::vector<cd> V;
        small_herm_eig(nz, Gs, lam, V);
        const double top = *std::max_element(lam.begin(), lam.end());
        std::vector<int> keep;
        if (top > 0)
            for (int i = 0; i < nz; ++i)
                if (lam[i] > cutoff * top) keep.push_back(i);
        k = int(keep.size());
        if (k == 0) return KFC_OK;
        std::vector<cd> M(size_t(nz) * k);
        for (int c = 0; c < k; ++c)
            for (int i = 0; i < nz; ++i) M[i + size_t(c) * nz] = d[i] * V[i + size_t(keep[c]) * nz] / std::sqrt(lam[keep[c]]);
        return blas_gemm<S>(w, Z, N, nz, M, k, W, N, N, 0.0);
    }

    // _svqb! (src/chebyshev.jl:88-116)
    int svqb(S* W, S* Z, int na, const S* L, int nl, S* Sb) {
        int rc = project_out(Z, na, L, nl);
        if (rc != KFC_OK) return rc;
        int k1 = 0, k = 0;
        rc = svqb_pass(Sb, Z, na, k1);
        if (rc == KFC_OK) rc = svqb_pass(W, Sb, k1, k);
        if (rc != KFC_OK) return rc;
        int refills = 0;
        while (k < na) {
            if (++refills > 10)
                return kfc_fail(op->ctx, KFC_EBREAKDOWN,
                                "ChebyshevFilter: SVQB could not refill " + std::to_string(na - k) +
                                    " dropped directions with random vectors; the block is wider than the space left beside the locked columns");
            const int r = na - k;
            rc = blas_randn<S>(w, Z, N, r, N, rng());
            if (rc != KFC_OK) return rc;
            std::vector<double> before, after;
            rc = blas_colnorms<S>(w, Z, N, r, N, before);
            if (rc == KFC_OK) rc = project_out(Z, r, L, nl);
            if (rc == KFC_OK) rc = project_out(Z, r, W, k);
            if (rc == KFC_OK) rc = blas_colnorms<S>(w, Z, N, r, N, after);
            if (rc != KFC_OK) return rc;
            std::vector<double> room(r);
            bool all = true;
            for (int j = 0; j < r; ++j) {
                room[j] = after[j] > std::sqrt(EPS) * before[j] ? 1.0 : 0.0;
                all &= room[j] == 1.0;
            }
            if (!all && (rc = blas_scale_cols<S>(w, Z, N, r, N, room)) != KFC_OK) return rc;
            if ((rc = normalize(Z, r)) != KFC_OK) return rc;
            int j1 = 0;
            if ((rc = svqb_pass(Sb, Z, r, j1)) != KFC_OK) return rc;
            if (j1 == 0) continue;
            if ((rc = project_out(Sb, j1, W, k)) != KFC_OK) return rc;
            int kk = 0;
            if ((rc = svqb_pass(col(W, k), Sb, j1, kk)) != KFC_OK) return rc;
            k += kk;
        }
        return project_out(W, na, L, nl);
    }

    // _cholqr! (src/chebyshev.jl:119-131)
    int cholqr(S* W, S* Z, int na, const S* L, int nl, S* Sb) {
        int rc = project_out(Z, na, L, nl);
        if (rc != KFC_OK) return rc;
        S* src = Z;
        for (S* dst : {Sb, W}) {
            std::vector<cd> G;
            rc = blas_gram<S>(w, src, N, na, src, N, na, N, G);
            if (rc != KFC_OK) return rc;
            // Cholesky G = U^H U, then inv(U)
            std::vector<cd> U(size_t(na) * na, cd(0));
            for (int j = 0; j < na; ++j) {
                double s = G[j + size_t(j) * na].real();
                for (int k = 0; k < j; ++k) s -= std::norm(U[k + size_t(j) * na]);
                if (!(s > 0))
                    return kfc_fail(op->ctx, KFC_EBREAKDOWN,
                                    "ChebyshevFilter: Cholesky-QR failed, the filtered block is numerically rank deficient; use ortho=:svqb (it drops and refills dependent directions)");
                const double ujj = std::sqrt(s);
                U[j + size_t(j) * na] = ujj;
                for (int i = j + 1; i < na; ++i) {
                    cd t = 0.5 * (G[j + size_t(i) * na] + std::conj(G[i + size_t(j) * na]));
                    for (int k = 0; k < j; ++k) t -= std::conj(U[k + size_t(j) * na]) * U[k + size_t(i) * na];
                    U[j + size_t(i) * na] = t / ujj;
                }
            }
            std::vector<cd> Ui(size_t(na) * na, cd(0));
            for (int c = 0; c < na; ++c)
                for (int i = na - 1; i >= 0; --i) {
                    cd t = (i == c) ? cd(1) : cd(0);
                    for (int k = i + 1; k < na; ++k) t -= U[i + size_t(k) * na] * Ui[k + size_t(c) * na];
                    Ui[i + size_t(c) * na] = t / U[i + size_t(i) * na];
                }
            rc = blas_gemm<S>(w, src, N, na, Ui, na, dst, N, N, 0.0);
            if (rc != KFC_OK) return rc;
            src = dst;
        }
        return project_out(W, na, L, nl);
    }

    // _rayleigh_ritz! (src/chebyshev.jl:138-144)
    int rayleigh_ritz(S* Xa, S* HXa, const S* W, const S* HW, int n, std::vector<double>& th) {
        std::vector<cd> G, Q;
        int rc = blas_gram<S>(w, W, N, n, HW, N, n, N, G);
        if (rc != KFC_OK) return rc;
        small_herm_eig(n, G, th, Q);
        rc = blas_gemm<S>(w, W, N, n, Q, n, Xa, N, N, 0.0);
        if (rc == KFC_OK) rc = blas_gemm<S>(w, HW, N, n, Q, n, HXa, N, N, 0.0);
        return rc;
    }
};

double log_gain(double t, double lcut, double lmax) {
    const double x = std::fabs((t - (lmax + lcut) / 2) / ((lmax - lcut) / 2));
    return x > 1 ? std::acosh(x) : 0.0;
}

double partial_m(const std::vector<double>& th, int n, int m) {   // partialsort(th[0:n], m)
    std::vector<double> v(th.begin(), th.begin() + n);
    std::nth_element(v.begin(), v.begin() + (m - 1), v.end());
    return v[m - 1];
}

// The amplification cap 1/sqrt(eps) uses the storage precision's eps (src/chebyshev.jl:169-174).
int chebyshev_degree(int degree, double t1, double lcut, double lmax, double eps) {
    const double x = std::fabs((t1 - (lmax + lcut) / 2) / ((lmax - lcut) / 2));
    if (!(x > 1)) return degree;
    const double cap = std::acosh(1.0 / std::sqrt(eps)) / std::acosh(x);
    return cap >= degree ? degree : std::max(int(std::floor(cap)), 1);
}

int counted_guards(const std::vector<double>& th, int m, int g, double lcut, double lmax, int deg) {   // :238-248
    if (!(deg > 0 && std::isfinite(lcut))) return g;
    const double ref = log_gain(partial_m(th, int(th.size()), m), lcut, lmax);
    if (!(ref > 0)) return g;
    int k = 0;
    for (int j = m; j < m + g; ++j) {
        if (!(log_gain(th[j], lcut, lmax) >= GUARD_GAIN_FRACTION * ref)) break;
        ++k;
    }
    return k;
}

bool chebyshev_done(const std::vector<double>& th, int nlock, int m, int need) {   // :297-301
    if (nlock < need) return false;
    if (nlock == int(th.size())) return true;
    return th[nlock] >= partial_m(th, nlock, m);
}

struct ChebCfg {
    int m, degree, extra, rescale_every, ortho;
    int64_t maxiter;
    double bound, tol, hbound, lmax;
    bool callback = false;   // hand each iteration to the context's callback
    int64_t it0 = 0;         // iterations of an earlier phase (the mixed path's FP32 run)
};

// One ChFSI run on scalar S from the seed (N x nb): the FP64 device block X0d, else the
// host block X0h; returns theta, X (host) and the locked count.  switch_tol > 0 stops
// once every unlocked residual is below it.
template <class S>
int cheb_run(kfc_op* op, const ChebCfg& c, const std::vector<cd>& X0h, uint64_t seed, CTimer& tm,
             std::vector<double>& theta_out, std::vector<cd>& Xout, int& nlock_out, int64_t& it_out, int& deg_out,
             double switch_tol, bool seed_locked_ok, const typename Upper<S>::type* X0d = nullptr) {
    const int64_t N = op->N;
    const int nb = int(std::min<int64_t>(c.m + c.extra, N));
    const int m = c.m;
    const double eps_s = prec_of<S>() == 0 ? EPS : 1.1920928955078125e-07;
    const int g = std::max(std::min(std::min(c.extra - 1, CHEB_GUARDS), nb - 1 - m), 0);
    Cheb<S> ch{op, Work(), CTimer(), N, std::mt19937_64(seed)};
    ch.w.ctx = op->ctx;
    if (prec_of<S>() == 1) ch.cutoff = 1e-6;
    DevBuf b0, b1, b2, bX, bHX;
    for (DevBuf* b : {&b0, &b1, &b2, &bX, &bHX}) KFC_HIP(op->ctx, b->ensure(sizeof(S) * size_t(N) * nb));
    S* bufs[3] = {b0.as<S>(), b1.as<S>(), b2.as<S>()};
    S* X = bX.as<S>();
    S* HX = bHX.as<S>();
    if (X0d) {
        const int rc0 = blas_convert<typename Upper<S>::type, S>(ch.w, X0d, N, bufs[0], N, nb, N);
        if (rc0 != KFC_OK) return rc0;
    } else {
        std::vector<S> h(size_t(N) * nb);
        for (size_t i = 0; i < h.size(); ++i) h[i] = s_from<S>(X0h[i]);
        KFC_HIP(op->ctx, hipMemcpy(bufs[0], h.data(), h.size() * sizeof(S), hipMemcpyHostToDevice));
    }
    auto ortho = [&](S* W, S* Z, int na, const S* L, int nl, S* Sb) {
        return c.ortho == 0 ? ch.svqb(W, Z, na, L, nl, Sb) : ch.cholqr(W, Z, na, L, nl, Sb);
    };
    const double lmax = c.lmax;
    int rc = ch.normalize(bufs[0], nb);
    if (rc == KFC_OK) rc = ortho(bufs[1], bufs[0], nb, nullptr, 0, bufs[2]);
    if (rc == KFC_OK) rc = cheb_apply<S>(op, bufs[1], nullptr, bufs[0], nb, 1.0, 0.0, 0.0, tm);
    std::vector<double> th;
    if (rc == KFC_OK) rc = ch.rayleigh_ritz(X, HX, bufs[1], bufs[0], nb, th);
    if (rc != KFC_OK) return rc;
    auto check_bound = [&](const std::vector<double>& t, int from) -> int {   // :179-186
        const double slack = std::sqrt(EPS) * std::max(std::fabs(lmax), 1.0);
        double top = -INFINITY;
        for (size_t i = from; i < t.size(); ++i) top = std::max(top, t[i]);
        if (!(top <= lmax + slack))
            return kfc_fail(op->ctx, KFC_EARG, "ChebyshevFilter: a Ritz value " + std::to_string(top) +
                                                   " exceeds the upper end lambda_max = " + std::to_string(lmax) +
                                                   " of the filter interval, so the bound is below the true largest eigenvalue; use bound=:gershgorin");
        return KFC_OK;
    };
    if ((rc = check_bound(th, 0)) != KFC_OK) return rc;
    std::vector<double> res(nb, INFINITY);
    int nlock = 0, deg = 0;
    int64_t it = 0;
    double lcut = NAN;
    while (true) {
        const int na = nb - nlock;
        std::vector<double> th_act(th.begin() + nlock, th.end()), rn;
        rc = blas_residual<S>(ch.w, bufs[0], N, ch.col(HX, nlock), N, ch.col(X, nlock), N, na, N, th_act, rn);
        if (rc != KFC_OK) return rc;
        for (int j = 0; j < na; ++j) res[nlock + j] = rn[j];
        const bool filtered = it >= 1 || nb == N || seed_locked_ok;
        const double floor = 10 * eps_s * std::max(std::fabs(c.hbound), std::fabs(*std::min_element(th.begin(), th.end())));
        while (filtered && nlock < nb && res[nlock] <= LOCK_FRACTION * std::max(c.tol, floor)) ++nlock;
        if (c.callback) {   // src/chebyshev.jl:362-363
            kfc_iter_info info{};
            info.solver = 1;
            info.precision = prec_of<S>();
            info.iteration = c.it0 + it;
            info.n_matvec = tm.cols;
            info.width = nb;
            info.n_locked = nlock;
            info.active0 = 0;
            info.degree = deg;
            info.lambda = th.data();
            info.residuals = res.data();
            info.lambda_cut = lcut;
            info.lambda_max = lmax;
            rc = ctx_callback(op->ctx, info, "ChebyshevFilter");
            if (rc != KFC_OK) return rc;
        }
        const int need = m + counted_guards(th, m, g, lcut, lmax, deg);
        if (chebyshev_done(th, nlock, m, need)) break;
        if (switch_tol > 0 && it >= 1) {
            double worst = 0;
            for (int j = nlock; j < nb; ++j) worst = std::max(worst, res[j]);
            if (worst <= switch_tol) break;
        }
        if (it >= c.maxiter) {
            std::string why;
            if (nlock < need)
                why = "only " + std::to_string(nlock) + " of " + std::to_string(need) + " eigenpairs converged (" +
                      std::to_string(m) + " wanted and " + std::to_string(need - m) + " guard columns)";
            else
                why = "the lowest " + std::to_string(need) + " columns are locked, but an unlocked Ritz value lies below the " +
                      std::to_string(m) + "-th locked energy, a lower state not yet converged";
            return kfc_fail(op->ctx, KFC_ENOCONV, "ChebyshevFilter: within maxiter = " + std::to_string(c.maxiter) +
                                                      " iterations " + why);
        }
        ++it;
        const int nact = nb - nlock;
        lcut = th[nb - 1];   // _chebyshev_cut (src/chebyshev.jl:157)
        if (!(lcut < lmax))
            return kfc_fail(op->ctx, KFC_EARG, "ChebyshevFilter: the block's largest Ritz value reaches the upper end lambda_max of the filter interval; the block spans the top of the spectrum, use Dense()");
        deg = chebyshev_degree(c.degree, *std::min_element(th.begin(), th.end()), lcut, lmax, eps_s);
        // the filter (src/chebyshev.jl:38-54)
        const double cc = (lmax + lcut) / 2, e = (lmax - lcut) / 2;
        S* A = bufs[0];
        S* B = bufs[1];
        S* C = bufs[2];
        rc = blas_axpby<S>(ch.w, ch.col(X, nlock), N, A, N, nact, N, 1.0, 0.0);
        if (rc == KFC_OK) rc = cheb_apply<S>(op, A, nullptr, B, nact, 1.0 / e, -cc / e, 0.0, tm);
        if (rc != KFC_OK) return rc;
        for (int k = 2; k <= deg; ++k) {
            rc = cheb_apply<S>(op, B, A, C, nact, 2.0 / e, -2.0 * cc / e, -1.0, tm);
            if (rc != KFC_OK) return rc;
            S* t = A;
            A = B;
            B = C;
            C = t;
            if (k % c.rescale_every == 0) {
                std::vector<double> sc;
                if ((rc = blas_colnorms<S>(ch.w, B, N, nact, N, sc)) != KFC_OK) return rc;
                for (double& x : sc) x = 1.0 / std::max(x, 2.2250738585072014e-308);
                if ((rc = blas_scale_cols<S>(ch.w, A, N, nact, N, sc)) != KFC_OK) return rc;
                if ((rc = blas_scale_cols<S>(ch.w, B, N, nact, N, sc)) != KFC_OK) return rc;
            }
        }
        S* Z = B;
        std::vector<double> zn;
        if ((rc = blas_colnorms<S>(ch.w, Z, N, nact, N, zn)) != KFC_OK) return rc;
        for (double x : zn)
            if (!std::isfinite(x))
                return kfc_fail(op->ctx, KFC_EARG, "ChebyshevFilter: the filtered block is not finite: the upper end lambda_max of the filter interval is below the true largest eigenvalue; use bound=:gershgorin");
        S* Wv = (Z == A) ? B : A;
        S* Sv = (Z == C) ? B : C;
        if (Z == A) Sv = C;
        if (Z == C) { Wv = A; Sv = B; }
        if ((rc = ch.normalize(Z, nact)) != KFC_OK) return rc;
        if ((rc = ortho(Wv, Z, nact, X, nlock, Sv)) != KFC_OK) return rc;
        if ((rc = cheb_apply<S>(op, Wv, nullptr, Z, nact, 1.0, 0.0, 0.0, tm)) != KFC_OK) return rc;
        std::vector<double> tha;
        if ((rc = ch.rayleigh_ritz(ch.col(X, nlock), ch.col(HX, nlock), Wv, Z, nact, tha)) != KFC_OK) return rc;
        for (int j = 0; j < nact; ++j) th[nlock + j] = tha[j];
        if ((rc = check_bound(th, nlock)) != KFC_OK) return rc;
    }
    theta_out = th;
    nlock_out = nlock;
    it_out = it;
    deg_out = deg;
    std::vector<S> h(size_t(N) * nb);
    KFC_HIP(op->ctx, hipMemcpy(h.data(), X, h.size() * sizeof(S), hipMemcpyDeviceToHost));
    Xout.resize(h.size());
    for (size_t i = 0; i < h.size(); ++i) {
        if constexpr (IsCx<S>::value) Xout[i] = cd(h[i].re, h[i].im);
        else Xout[i] = cd(h[i], 0.0);
    }
    return KFC_OK;
}

template <class S>
int chebyshev_t(kfc_op* op, int64_t m, int degree, int64_t extra, int64_t maxiter, int rescale_every, int ortho,
                double bound, int seed_kind, double tol, uint64_t seed, int mode, double* Eout, double* Vout,
                kfc_stats* st) {
    auto t0 = clk::now();
    const int64_t N = op->N;
    if (m < 1) return kfc_fail(op->ctx, KFC_EARG, "ChebyshevFilter requires m >= 1, got " + std::to_string(m));
    if (m > N) return kfc_fail(op->ctx, KFC_EARG, "ChebyshevFilter: m exceeds the operator dimension");
    if (!(tol > 0 && std::isfinite(tol))) return kfc_fail(op->ctx, KFC_EARG, "ChebyshevFilter: tol must be positive and finite");
    if (degree < 1 || degree > 200) return kfc_fail(op->ctx, KFC_EARG, "ChebyshevFilter: degree must lie in 1:200");
    if (extra < 2) return kfc_fail(op->ctx, KFC_EARG, "ChebyshevFilter: extra must be at least 2");
    if (maxiter < 1 || rescale_every < 1) return kfc_fail(op->ctx, KFC_EARG, "ChebyshevFilter: maxiter and rescale_every must be at least 1");
    const int nb = int(std::min<int64_t>(m + extra, N));
    if (nb > 256) return kfc_fail(op->ctx, KFC_EUNSUPPORTED, "ChebyshevFilter: blocks above 256 columns are not served");
    double hb = 0;
    int rc = kfc_op_gershgorin(op, &hb);
    if (rc != KFC_OK) return rc;
    ChebCfg c{int(m), degree, int(extra), rescale_every, ortho, maxiter, bound, tol, hb, std::isnan(bound) ? hb : bound};
    c.callback = op->ctx->cb != nullptr;
    // seed, formed on the device: the product seed plus a 1e-2 random admixture
    // (_admixed_seed, src/chebyshev.jl:252-261), or a random block
    DevBuf dX0;
    KFC_HIP(op->ctx, dX0.ensure(sizeof(S) * size_t(N) * nb));
    {
        Work w0;
        w0.ctx = op->ctx;
        if (seed_kind == KFC_SEED_PRODUCT) {
            rc = op_product_seed_dev<S>(op, nb, dX0.as<S>(), N);
            if (rc == KFC_OK) rc = blas_perturb<S>(w0, dX0.as<S>(), N, nb, N, 1, 0, nullptr, 0x4346 ^ seed, SEED_ADMIXTURE);
        } else {
            rc = blas_randn<S>(w0, dX0.as<S>(), N, nb, N, 0x4346 ^ seed);
        }
        if (rc != KFC_OK) return rc;
        KFC_HIP(op->ctx, hipStreamSynchronize(op->ctx->stream));
    }
    const std::vector<cd> X0;
    CTimer tm;
    std::vector<double> th;
    std::vector<cd> X;
    int nlock = 0, deg = 0;
    int64_t it = 0;
    if (mode == KFC_MODE_MIXED) {
        using L = typename Lower<S>::type;
        const double sw = std::max(tol, 256.0 * 5.96e-8 * std::max(std::fabs(hb), 1.0));
        int64_t it32 = 0;
        rc = cheb_run<L>(op, c, X0, seed, tm, th, X, nlock, it32, deg, sw, false, dX0.as<S>());
        if (rc != KFC_OK) return rc;
        ChebCfg c2 = c;
        c2.maxiter = std::max<int64_t>(1, maxiter - it32);
        c2.it0 = it32 + 1;
        std::vector<cd> Xs = X;
        rc = cheb_run<S>(op, c2, Xs, seed + 1, tm, th, X, nlock, it, deg, 0.0, false);
        it += it32;
    } else {
        rc = cheb_run<S>(op, c, X0, seed, tm, th, X, nlock, it, deg, 0.0, false, dX0.as<S>());
    }
    if (rc != KFC_OK) return rc;
    // the lowest m locked pairs, checked with a fresh FP64 application (src/chebyshev.jl:395-405)
    std::vector<int> idx(nlock);
    for (int i = 0; i < nlock; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return th[a] < th[b]; });
    std::vector<cd> Vh(size_t(N) * m);
    std::vector<double> E(m);
    for (int j = 0; j < m; ++j) {
        E[j] = th[idx[j]];
        std::copy(X.begin() + size_t(idx[j]) * N, X.begin() + size_t(idx[j] + 1) * N, Vh.begin() + size_t(j) * N);
    }
    // fresh residuals in FP64
    Work w;
    w.ctx = op->ctx;
    DevBuf bx, by;
    KFC_HIP(op->ctx, bx.ensure(sizeof(S) * size_t(N) * m));
    KFC_HIP(op->ctx, by.ensure(sizeof(S) * size_t(N) * m));
    {
        std::vector<S> h(Vh.size());
        for (size_t i = 0; i < h.size(); ++i) h[i] = s_from<S>(Vh[i]);
        KFC_HIP(op->ctx, hipMemcpy(bx.p, h.data(), h.size() * sizeof(S), hipMemcpyHostToDevice));
    }
    const int64_t cols_solve = tm.cols;
    rc = cheb_apply<S>(op, bx.as<S>(), nullptr, by.as<S>(), int(m), 1.0, 0.0, 0.0, tm);
    std::vector<double> tr;
    if (rc == KFC_OK) rc = blas_residual<S>(w, by.as<S>(), N, by.as<S>(), N, bx.as<S>(), N, int(m), N, E, tr);
    if (rc != KFC_OK) return rc;
    const double worst = *std::max_element(tr.begin(), tr.end());
    const double floor = 10 * EPS * std::max(std::fabs(hb), std::fabs(E[0]));
    if (!(worst <= std::max(tol, floor)))
        return kfc_fail(op->ctx, KFC_ENOCONV, "ChebyshevFilter: the true residual ||H v - E v|| = " + std::to_string(worst) +
                                                  " exceeds " + std::to_string(std::max(tol, floor)) + " on return");
    for (int j = 0; j < m; ++j) Eout[j] = E[j];
    for (size_t i = 0; i < Vh.size(); ++i) {
        if (op->cplx) {
            Vout[2 * i] = Vh[i].real();
            Vout[2 * i + 1] = Vh[i].imag();
        } else {
            Vout[i] = Vh[i].real();
        }
    }
    if (st) {
        *st = kfc_stats{};
        st->iterations = it;
        st->n_matvec = cols_solve;
        st->n_matvec_f32 = tm.cols32;
        st->n_apply_calls = tm.calls;
        st->seconds_apply = tm.apply;
        st->seconds_apply_gpu = tm.apply_gpu;
        st->max_residual = worst;
        st->n_converged = int(m);
        st->degree_last = deg;
        st->seconds_total = secs(t0);
    }
    return KFC_OK;
}

}  // namespace

extern "C" int kfc_chebyshev(kfc_op* op, int64_t m, int32_t degree, int64_t extra, int64_t maxiter,
                             int32_t rescale_every, int32_t ortho, double bound, int32_t seed_kind, double tol,
                             uint64_t seed, int32_t mode, double* E, double* V, kfc_stats* stats) {
    if (!op || !E || !V) return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_chebyshev: null argument");
    if (!std::isnan(bound) && !std::isfinite(bound))
        return kfc_fail(op->ctx, KFC_EARG, "ChebyshevFilter: bound must be :gershgorin or a finite number");
    return op->cplx ? chebyshev_t<Cx<double>>(op, m, degree, extra, maxiter, rescale_every, ortho, bound, seed_kind, tol,
                                              seed, mode, E, V, stats)
                    : chebyshev_t<double>(op, m, degree, extra, maxiter, rescale_every, ortho, bound, seed_kind, tol,
                                          seed, mode, E, V, stats);
}
