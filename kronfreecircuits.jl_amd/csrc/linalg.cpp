// Host Hermitian eigensolver for the small matrices of the solver loops
// (Rayleigh-Ritz, Gram, Krylov projections, Dense nodes): Householder
// reduction to a real symmetric tridiagonal, then implicit QL with Wilkinson
// shifts accumulating the rotations (Golub & Van Loan, "Matrix
// Computations", 4th ed., sections 8.3.1 and 8.3.5).  No LAPACK is linked.
#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <numeric>
#include <vector>

using cd = std::complex<double>;

namespace {

template <class T> inline T tconj(T x);
template <> inline double tconj<double>(double x) { return x; }
template <> inline cd tconj<cd>(cd x) { return std::conj(x); }
inline double tabs(double x) { return std::fabs(x); }
inline double tabs(cd x) { return std::abs(x); }

// A (n x n, column-major, Hermitian, lower triangle used) -> Q T Q^H with T real
// tridiagonal (d, e) and Q accumulated into Z.
template <class T>
void tridiagonalize(int n, std::vector<T>& A, std::vector<double>& d, std::vector<double>& e, std::vector<T>& Z) {
    auto a = [&](int i, int j) -> T& { return A[i + size_t(j) * n]; };
    // symmetrize from the lower triangle
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) a(j, i) = tconj(a(i, j));
    Z.assign(size_t(n) * n, T(0));
    for (int i = 0; i < n; ++i) Z[i + size_t(i) * n] = T(1);
    std::vector<T> v(n), p(n), q(n);
    std::vector<T> sub(std::max(n - 1, 0));
    for (int k = 0; k + 2 < n; ++k) {
        // x = A[k+1:n, k]
        double xn = 0;
        for (int i = k + 1; i < n; ++i) xn += std::norm(cd(a(i, k)));
        xn = std::sqrt(xn);
        if (xn == 0) { sub[k] = T(0); continue; }
        T x0 = a(k + 1, k);
        double ax0 = tabs(x0);
        T ph = ax0 == 0 ? T(1) : x0 / ax0;
        T alpha = -ph * xn;
        // v = x - alpha e1, normalized
        std::fill(v.begin(), v.end(), T(0));
        for (int i = k + 1; i < n; ++i) v[i] = a(i, k);
        v[k + 1] -= alpha;
        double vn = 0;
        for (int i = k + 1; i < n; ++i) vn += std::norm(cd(v[i]));
        vn = std::sqrt(vn);
        if (vn == 0) { sub[k] = a(k + 1, k); continue; }
        for (int i = k + 1; i < n; ++i) v[i] /= vn;
        // p = A22 v, K = v^H p, q = p - K v;  A22 -= 2 v q^H + 2 q v^H
        for (int i = k + 1; i < n; ++i) {
            T s = 0;
            for (int j = k + 1; j < n; ++j) s += a(i, j) * v[j];
            p[i] = s;
        }
        T K = 0;
        for (int i = k + 1; i < n; ++i) K += tconj(v[i]) * p[i];
        for (int i = k + 1; i < n; ++i) q[i] = p[i] - K * v[i];
        for (int j = k + 1; j < n; ++j) {
            const T vj = tconj(v[j]), qj = tconj(q[j]);
            for (int i = k + 1; i < n; ++i) a(i, j) -= T(2) * (v[i] * qj + q[i] * vj);
        }
        a(k + 1, k) = alpha;
        a(k, k + 1) = tconj(alpha);
        for (int i = k + 2; i < n; ++i) a(i, k) = a(k, i) = T(0);
        sub[k] = alpha;
        // Z <- Z (I - 2 v v^H)
        for (int r = 0; r < n; ++r) {
            T s = 0;
            for (int i = k + 1; i < n; ++i) s += Z[r + size_t(i) * n] * v[i];
            s *= T(2);
            for (int i = k + 1; i < n; ++i) Z[r + size_t(i) * n] -= s * tconj(v[i]);
        }
    }
    if (n >= 2) sub[n - 2] = a(n - 1, n - 2);
    d.resize(n);
    e.assign(n, 0.0);
    for (int i = 0; i < n; ++i) d[i] = std::real(cd(a(i, i)));
    // make the subdiagonal real and nonnegative with a diagonal phase: D_{i+1} = D_i e_i/|e_i|
    T D = T(1);
    for (int i = 0; i + 1 < n; ++i) {
        const T s = sub[i];
        const double as = tabs(s);
        e[i] = as;
        T ph = as == 0 ? T(1) : s / as;
        D = D * ph;
        for (int r = 0; r < n; ++r) Z[r + size_t(i + 1) * n] *= D;
    }
}

// sqrt(a^2 + b^2) without the cost of std::hypot where no scaling is needed
inline double fast_hypot(double a, double b) {
    const double m = std::max(std::fabs(a), std::fabs(b));
    if (m < 1e150 && m > 1e-150) return std::sqrt(a * a + b * b);
    return std::hypot(a, b);
}

// Implicit QL with Wilkinson shifts on the symmetric tridiagonal (d, e[i] = T[i+1, i]),
// rotations accumulated into the columns of Z (n_rows x n).
template <class T>
void tql(int n, std::vector<double>& d, std::vector<double>& e, std::vector<T>& Z, int n_rows) {
    if (n == 0) return;
    e[n - 1] = 0.0;
    const double eps = std::numeric_limits<double>::epsilon();
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        while (true) {
            int m = l;
            for (; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) break;   // no convergence (never seen for symmetric input)
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = fast_hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            bool deflated = false;
            for (; i >= l; --i) {
                double f = s * e[i], b = c * e[i];
                r = fast_hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < n_rows; ++k) {
                    T* zi = &Z[k + size_t(i) * n_rows];
                    T* zi1 = &Z[k + size_t(i + 1) * n_rows];
                    const T fz = *zi1;
                    *zi1 = s * *zi + c * fz;
                    *zi = c * *zi - s * fz;
                }
            }
            if (deflated) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
}

template <class T>
void eig_impl(int n, std::vector<T>& A, std::vector<double>& w, std::vector<T>& Z) {
    std::vector<double> d, e;
    tridiagonalize(n, A, d, e, Z);
    tql(n, d, e, Z, n);
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
    w.resize(n);
    std::vector<T> Zs(size_t(n) * n);
    for (int j = 0; j < n; ++j) {
        w[j] = d[idx[j]];
        std::copy(Z.begin() + size_t(idx[j]) * n, Z.begin() + size_t(idx[j] + 1) * n, Zs.begin() + size_t(j) * n);
    }
    Z.swap(Zs);
}

// Solve (T - sigma I) x = b for the symmetric tridiagonal T (diagonal d, off-diagonal
// e) by Gaussian elimination with partial pivoting (the LU of a tridiagonal has two
// superdiagonals); a zero pivot is replaced by a tiny one, as inverse iteration wants.
struct TriLU {
    int n = 0;
    std::vector<double> dl, dg, du, du2;
    std::vector<char> swp;
    void factor(int n_, const std::vector<double>& d, const std::vector<double>& e, double sigma, double tiny) {
        n = n_;
        dg.resize(n);
        dl.assign(std::max(n - 1, 0), 0.0);
        du.assign(std::max(n - 1, 0), 0.0);
        du2.assign(std::max(n - 2, 0), 0.0);
        swp.assign(std::max(n - 1, 0), 0);
        for (int i = 0; i < n; ++i) dg[i] = d[i] - sigma;
        for (int i = 0; i + 1 < n; ++i) dl[i] = du[i] = e[i];
        for (int i = 0; i + 1 < n; ++i) {
            if (std::fabs(dg[i]) >= std::fabs(dl[i])) {
                if (dg[i] == 0.0) dg[i] = tiny;
                const double f = dl[i] / dg[i];
                dl[i] = f;
                dg[i + 1] -= f * du[i];
            } else {   // rows i and i + 1 trade places
                const double f = dg[i] / dl[i];
                dg[i] = dl[i];
                dl[i] = f;
                const double t = du[i];
                du[i] = dg[i + 1];
                dg[i + 1] = t - f * dg[i + 1];
                if (i + 2 < n) {
                    du2[i] = du[i + 1];
                    du[i + 1] = -f * du[i + 1];
                }
                swp[i] = 1;
            }
        }
        if (n > 0 && dg[n - 1] == 0.0) dg[n - 1] = tiny;
    }
    void solve(std::vector<double>& x) const {
        for (int i = 0; i + 1 < n; ++i) {
            if (swp[i]) std::swap(x[i], x[i + 1]);
            x[i + 1] -= dl[i] * x[i];
        }
        x[n - 1] /= dg[n - 1];
        if (n >= 2) x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / dg[n - 2];
        for (int i = n - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / dg[i];
    }
};

}  // namespace

// Lowest k eigenpairs of a real symmetric matrix: Householder tridiagonalization
// (Golub & Van Loan, 4th ed., alg. 8.3.1), all eigenvalues by implicit QL without
// vectors, the k eigenvectors of the tridiagonal by inverse iteration with the
// members of a cluster (eigenvalues closer than 1e-3 of the norm) orthogonalized
// against each other (Wilkinson; the scheme of Peters & Wilkinson, as LAPACK's
// stein uses it), then back-transformed through the reflectors.  O(4/3 n^3 + 2 k n^2)
// against the O(9 n^3) of QL with vectors.
void sym_eig_lowest(int n, int k, std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z) {
    auto a = [&](int i, int j) -> double& { return A[i + size_t(j) * n]; };
    k = std::max(0, std::min(k, n));
    std::vector<double> d(n), e(n, 0.0), beta(n, 0.0), V(size_t(n) * n, 0.0), p(n), q(n);
    double anorm = 0;
    for (int j = 0; j < n; ++j) {
        double s = 0;
        for (int i = 0; i < n; ++i) s += std::fabs(a(i, j));
        anorm = std::max(anorm, s);
    }
    for (int j = 0; j + 2 < n; ++j) {
        const int m = n - j - 1;
        double* x = &A[(j + 1) + size_t(j) * n];
        double sigma = 0;
        for (int i = 1; i < m; ++i) sigma += x[i] * x[i];
        const double x0 = x[0];
        if (sigma == 0.0) {
            e[j] = x0;
            continue;
        }
        const double mu = std::sqrt(x0 * x0 + sigma);
        const double v0 = x0 <= 0 ? x0 - mu : -sigma / (x0 + mu);
        const double b = 2 * v0 * v0 / (sigma + v0 * v0);
        double* v = &V[(j + 1) + size_t(j) * n];
        v[0] = 1.0;
        for (int i = 1; i < m; ++i) v[i] = x[i] / v0;
        beta[j] = b;
        e[j] = mu;
        // A22 <- P A22 P with P = I - b v v^T:  p = b A22 v, w = p - (b/2)(p.v) v,
        // A22 -= v w^T + w v^T
        std::fill(p.begin(), p.begin() + m, 0.0);
        for (int l = 0; l < m; ++l) {
            const double vl = b * v[l];
            const double* col = &A[(j + 1) + size_t(j + 1 + l) * n];
            for (int i = 0; i < m; ++i) p[i] += col[i] * vl;
        }
        double pv = 0;
        for (int i = 0; i < m; ++i) pv += p[i] * v[i];
        const double K = 0.5 * b * pv;
        for (int i = 0; i < m; ++i) q[i] = p[i] - K * v[i];
        for (int l = 0; l < m; ++l) {
            double* col = &A[(j + 1) + size_t(j + 1 + l) * n];
            const double vl = v[l], ql = q[l];
            for (int i = 0; i < m; ++i) col[i] -= v[i] * ql + q[i] * vl;
        }
    }
    for (int i = 0; i < n; ++i) d[i] = a(i, i);
    if (n >= 2) e[n - 2] = a(n - 1, n - 2);
    e[n - 1] = 0.0;
    // eigenvalues
    std::vector<double> dq = d, eq = e, none;
    tql<double>(n, dq, eq, none, 0);
    std::sort(dq.begin(), dq.end());
    w = dq;
    // eigenvectors of the tridiagonal for the lowest k
    const double eps = std::numeric_limits<double>::epsilon();
    const double tnorm = std::max(anorm, 1e-300);
    const double ortol = 1e-3 * tnorm, tiny = eps * tnorm;
    std::vector<double> Y(size_t(n) * k), x(n);
    TriLU lu;
    int c0 = 0;   // first member of the current cluster
    uint64_t rs = 0x9e3779b97f4a7c15ull;
    for (int j = 0; j < k; ++j) {
        if (j > 0 && w[j] - w[j - 1] > ortol) c0 = j;
        // members of a cluster get slightly separated shifts
        double sig = w[j];
        if (j > c0 && w[j] - w[j - 1] < 10 * eps * tnorm) sig = w[j - 1] + 10 * eps * tnorm;
        lu.factor(n, d, e, sig, tiny);
        for (int i = 0; i < n; ++i) {   // a fixed pseudo-random start
            rs ^= rs << 13;
            rs ^= rs >> 7;
            rs ^= rs << 17;
            x[i] = double(int64_t(rs % 2001) - 1000) / 1000.0;
        }
        for (int it = 0; it < 4; ++it) {
            lu.solve(x);
            for (int pass = 0; pass < (j > c0 ? 2 : 1); ++pass)
                for (int l = c0; l < j; ++l) {
                    const double* y = &Y[size_t(l) * n];
                    double s = 0;
                    for (int i = 0; i < n; ++i) s += y[i] * x[i];
                    for (int i = 0; i < n; ++i) x[i] -= s * y[i];
                }
            double nx = 0;
            for (int i = 0; i < n; ++i) nx += x[i] * x[i];
            nx = std::sqrt(nx);
            if (!(nx > 0)) {
                for (int i = 0; i < n; ++i) x[i] = i == j ? 1.0 : 0.0;
                continue;
            }
            for (int i = 0; i < n; ++i) x[i] /= nx;
        }
        std::copy(x.begin(), x.end(), Y.begin() + size_t(j) * n);
    }
    // back-transform z = H_0 H_1 ... H_{n-3} y
    Z = Y;
    for (int c = 0; c < k; ++c) {
        double* z = &Z[size_t(c) * n];
        for (int j = n - 3; j >= 0; --j) {
            if (beta[j] == 0.0) continue;
            const double* v = &V[(j + 1) + size_t(j) * n];
            const int m = n - j - 1;
            double s = 0;
            for (int i = 0; i < m; ++i) s += v[i] * z[j + 1 + i];
            s *= beta[j];
            for (int i = 0; i < m; ++i) z[j + 1 + i] -= s * v[i];
        }
    }
}

void herm_eig(int n, std::vector<cd>& A, std::vector<double>& w, std::vector<cd>& Z) { eig_impl<cd>(n, A, w, Z); }
void sym_eig(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z) {
    eig_impl<double>(n, A, w, Z);
}
