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
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            bool deflated = false;
            for (; i >= l; --i) {
                double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
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

}  // namespace

void herm_eig(int n, std::vector<cd>& A, std::vector<double>& w, std::vector<cd>& Z) { eig_impl<cd>(n, A, w, Z); }
void sym_eig(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z) {
    eig_impl<double>(n, A, w, Z);
}
