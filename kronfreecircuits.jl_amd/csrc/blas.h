// Block linear algebra on tall device blocks (N rows, a few dozen columns):
// Gram matrices, block combinations, column reductions, the Lanczos fused
// orthogonalization passes and the seeded device RNG.  Reductions accumulate
// in FP64 whatever the storage precision.
#pragma once
#include "kfc_internal.h"

template <class S> struct AccOf { using type = double; };
template <class R> struct AccOf<Cx<R>> { using type = Cx<double>; };

// Scratch owned by a solver call: device partials and staging.
struct Work {
    kfc_ctx* ctx = nullptr;
    DevBuf part;      // reduction partials
    DevBuf small;     // small matrices / vectors uploaded to the device
    DevBuf small2;
    DevBuf tmp;       // N x q scratch of the general tall-skinny product
    std::vector<unsigned char> host;
    // the matrix-core tall-skinny products (k_tsmm): HIP-event time on the context
    // stream and algorithmic bytes (block columns read + columns written)
    double tsmm_gpu = 0, tsmm_bytes = 0;
    int64_t tsmm_calls = 0;
    hipEvent_t te0 = nullptr, te1 = nullptr;
    Work() = default;
    Work(const Work&) = delete;
    Work& operator=(const Work&) = delete;
    ~Work() {
        if (te0) (void)hipEventDestroy(te0);
        if (te1) (void)hipEventDestroy(te1);
    }
    hipError_t tsmm_begin() {
        if (!te0) {
            hipError_t e = hipEventCreate(&te0);
            if (e == hipSuccess) e = hipEventCreate(&te1);
            if (e != hipSuccess) return e;
        }
        return hipEventRecord(te0, ctx->stream);
    }
    // after the stream has been synchronized past the kernel
    hipError_t tsmm_end_synced(double bytes) {
        float ms = 0;
        hipError_t e = hipEventElapsedTime(&ms, te0, te1);
        if (e != hipSuccess) return e;
        tsmm_gpu += 1e-3 * ms;
        tsmm_bytes += bytes;
        ++tsmm_calls;
        return hipSuccess;
    }
};

// G (p x q, column-major, host) = A^H B over N rows.
template <class S>
int blas_gram(Work& w, const S* A, int64_t lda, int p, const S* B, int64_t ldb, int q, int64_t N, std::vector<cd>& G);
// C[:, 0:q] = A[:, 0:p] M (+ beta C), M p x q column-major host; C must not alias A.
template <class S>
int blas_gemm(Work& w, const S* A, int64_t lda, int p, const std::vector<cd>& M, int q, S* C, int64_t ldc, int64_t N,
              double beta);
// Y[:, 0:q] = X[:, 0:p] M (M p x q host, column-major), Y may be columns of X (each row
// of X is read before that row of Y is written).  gram = 1 also returns G = X[:, 0:pz]^H Y
// (pz x q), gram = 2 returns G = Y^H Y (q x q), gram = 3 returns G = [X[:, 0:pz] Y]^H Y
// ((pz + q) x q, the new Y in its lower block), over all N rows of the result.
template <class S>
int blas_tsmm(Work& w, const S* X, int64_t ldx, int p, const std::vector<cd>& M, int q, S* Y, int64_t ldy, int64_t N,
              int gram, int pz, std::vector<cd>* G);
// gram = 5 is gram = 3 without storing Y: only G leaves the call (real blocks with
// pz + q = p <= 64, q <= 16).  Chained: Z = X M1 (q1 columns, rounded to S, the very
// values a gram = 5 call forms from the same X and M1) replaces columns pz..pz+q1-1
// of X for each row, then Y[:, 0:q2] = [X with Z] M2 with G = Y^H Y (q2 x q2); X is
// not written.  M1 is p x q1, M2 p x q2; tsmm_chain_ok says whether a shape runs.
template <class S>
bool tsmm_chain_ok(int p, int pz, int q1, int q2);
template <class S>
int blas_tsmm_chain(Work& w, const S* X, int64_t ldx, int p, const std::vector<cd>& M1, int q1, int pz,
                    const std::vector<cd>& M2, int q2, S* Y, int64_t ldy, int64_t N, std::vector<cd>& G);
// Column 2-norms of A[:, 0:p] (host).
template <class S>
int blas_colnorms(Work& w, const S* A, int64_t lda, int p, int64_t N, std::vector<double>& nrm);
// A[:, j] *= s[j]
template <class S>
int blas_scale_cols(Work& w, S* A, int64_t lda, int p, int64_t N, const std::vector<double>& s);
// R[:, j] = AX[:, j] - theta[j] X[:, j]; returns column norms of R.
template <class S>
int blas_residual(Work& w, S* R, int64_t ldr, const S* AX, int64_t ldax, const S* X, int64_t ldx, int p, int64_t N,
                  const std::vector<double>& theta, std::vector<double>& nrm);
// R[:, j] /= hypot(D - theta_j, eta)  (RitzJacobi ldiv!, src/lobpcg.jl:114-119), or
// R[:, j] *= inv (FixedJacobi, src/lobpcg.jl:31-34) when eta < 0 (theta ignored, inv precomputed in D).
template <class S>
int blas_precond(Work& w, S* R, int64_t ldr, int p, int64_t N, const double* D, const std::vector<double>& theta,
                 double eta);
// theta[j] = sum D |X|^2 / sum |X|^2  (_diagonal_expectation!, src/lobpcg.jl:73-96)
template <class S>
int blas_diag_expect(Work& w, const S* X, int64_t ldx, int p, int64_t N, const double* D, std::vector<double>& theta);
// Y = alpha X + beta Y on p columns
template <class S>
int blas_axpby(Work& w, const S* X, int64_t ldx, S* Y, int64_t ldy, int p, int64_t N, double alpha, double beta);
// copy with precision conversion (columns)
template <class S, class T>
int blas_convert(Work& w, const S* X, int64_t ldx, T* Y, int64_t ldy, int p, int64_t N);
// The LOBPCG seed perturbation on the device: the first ncols[b] columns (all p when
// ncols is null) of each of `batch` blocks (stride bstride) get a random vector of
// rel (default sqrt(eps)) times their norm added, deterministic in (seed, column, row).
template <class S>
int blas_perturb(Work& w, S* X, int64_t ldx, int p, int64_t N, int batch, int64_t bstride, const int* ncols,
                 uint64_t seed, double rel = 1.4901161193847656e-08);
// N x p standard normal (complex: both parts), deterministic in (seed, column offset)
template <class S>
int blas_randn(Work& w, S* X, int64_t ldx, int p, int64_t N, uint64_t seed);

// Lanczos pass 1: h[i] = s_i * V[:, i]^H w for i < k (one pass over V), plus |w|^2.
template <class S>
int lz_multidot(Work& w, const S* V, int64_t ldv, int k, const double* scale, const S* wv, int64_t N,
                std::vector<cd>& h, double* wnorm2);
// Lanczos pass 2: w -= sum_i V[:, i] (s_i h_i), returns |w|^2 after.
template <class S>
int lz_update(Work& w, const S* V, int64_t ldv, int k, const std::vector<cd>& hs, S* wv, int64_t N, double* wnorm2);
