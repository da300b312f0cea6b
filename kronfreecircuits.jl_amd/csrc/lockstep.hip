// Lockstep LOBPCG over a sweep of operators of one structure.
//
// A parameter sweep solves many nodes that differ only in their numbers: the
// zero-pi parent [40, 40, 30] at every sweep point has the same dims, the same
// axis pairs and the same number of kept terms.  One such solve at a time is
// latency-bound (48 000 states and a 20-column block: each iteration is a few
// short launches and several host round trips).  Here every point of the sweep
// advances together: each step of the iteration is one launch for all active
// points (residual with the preconditioner, the dense-pair apply, the Gram
// matrices of the basis, the rotation), with one device-to-host and one
// host-to-device transfer per iteration, and the points' small Rayleigh-Ritz
// problems solved on host threads.
//
// The iteration is LOBPCG (Knyazev 2001) with the contract of lobpcg_solve
// (src/lobpcg.jl:217-301): block width w = m + g, every column (guards included)
// must reach the residual tolerance (n_conv_check = w, :263-269), RitzJacobi /
// FixedJacobi / no preconditioning (src/lobpcg.jl:18-135), product or warm
// seeds perturbed as _seed_columns does (:167-190).  Each point's basis is
// S = [X R P] (P absent on the first step).  The Rayleigh-Ritz step takes
// G = S^T S and K = S^T H S, scales the columns to unit norm, factors G by
// Cholesky (or, when G is close to singular, diagonalizes it and drops its
// near-null directions: SVQB, Stathopoulos & Wu, SIAM J. Sci. Comput. 23 (2002))
// and diagonalizes the projected K, as in the robust LOBPCG of Duersch, Shao,
// Yang & Gu (SIAM J. Sci. Comput. 40 (2018)).  The new P is the part of the
// new X that comes from R and P.  There is no soft locking: a point leaves the
// batch when all w of its residuals are below tol.  A point whose basis
// collapses is handed to the one-point path (kfc_lobpcg) after the batch.
//
// mixed mode: the blocks are stored, applied and rotated in FP32 (Gram sums
// in FP64) until every residual of a point is below the FP32 floor; then all
// points continue in FP64 from their FP32 blocks.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <thread>

#include "blas.h"

int lobpcg_check_x0(kfc_op* op, const double* X0, int64_t nx0, std::vector<cd>& X0h);
template <class S>
int lobpcg_seed_dev(kfc_op* op, int m, int W, const std::vector<cd>& X0h, int64_t nx0, bool product, bool random_only,
                    uint64_t s, S* Xd, int& nperturb);
void lobpcg_seed_block(const kfc_op* op, int m, int W, const std::vector<cd>& X0h, int64_t nx0,
                       const std::vector<cd>& P, bool random_only, uint64_t s, std::vector<cd>& X, int& nperturb);

namespace {

using clk = std::chrono::steady_clock;
inline double secs(clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); }

typedef double ld4 __attribute__((ext_vector_type(4)));
__device__ inline ld4 mfma64(double a, double b, ld4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

constexpr int LS_WMAX = 21;                      // w with 3w <= 64
constexpr int LS_CMAX = 64;                      // basis columns, padded
constexpr int LS_NT = 10;                        // upper 16 x 16 tiles of a 64 x 64 symmetric matrix
constexpr int LS_NORM = 2 * LS_NT * 256;         // offset of the residual norms in a point's record
constexpr int LS_E = LS_NORM + 2 * LS_WMAX + 2;  // reduced doubles per point
constexpr int LS_SR = 32;                        // rows per Gram slab
constexpr int LS_MO = 48;                        // padded output columns of the rotation (2w <= 42)

struct LsSlot {
    int p;     // point index
    int par;   // buffer holding its current basis
    int c;     // basis columns of this step ([X R P] prefix)
    int no;    // rotation: output columns (X' then P')
};

// Column layout of a point's block: X at 0, R at w, P at 2w; leading dimension N.
template <class T>
struct LsArgs {
    T* S[2];
    T* A[2];          // H S
    int64_t N, blk;   // rows; elements per point (3 w N)
    int W;
    int nchunk;
    int64_t rows_per;
    const LsSlot* slot;
    double* part;     // [slot][chunk][LS_E]
};

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// R = AX - X diag(theta) on ncol columns, |R_j|^2 into the norm partials, and
// (not fresh) R preconditioned into the R slot: RitzJacobi (pmode 1,
// 1/hypot(D - pth_j, eta), src/lobpcg.jl:114-119), FixedJacobi (pmode 2, D holds
// 1/(D - sigma), :31-34) or none.  fresh: AX from the other buffer's AX slot
// (a fresh apply of X), nothing written but the norms.
template <class T>
__global__ __launch_bounds__(256) void k_ls_resid(const LsArgs<T> a, int ncol, const double* __restrict__ theta,
                                                  const double* __restrict__ pth, const double* __restrict__ Dall,
                                                  int pmode, double eta, int fresh) {
    const int q = blockIdx.y;
    const LsSlot sl = a.slot[q];
    const int64_t N = a.N;
    const T* X = a.S[sl.par] + sl.p * a.blk;
    const T* AX = a.A[fresh ? sl.par ^ 1 : sl.par] + sl.p * a.blk;
    T* R = a.S[sl.par] + sl.p * a.blk + a.W * N;
    const double* D = Dall ? Dall + sl.p * N : nullptr;
    const int64_t r0 = blockIdx.x * a.rows_per, r1 = min(N, r0 + a.rows_per);
    double acc[LS_WMAX];
#pragma unroll
    for (int j = 0; j < LS_WMAX; ++j) {
        acc[j] = 0;
        if (j < ncol) {
            const double th = theta[sl.p * a.W + j];
            const double pt = pmode == 1 ? pth[sl.p * a.W + j] : 0.0;
            double s = 0;
            for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) {
                const double r = double(AX[i + j * N]) - th * double(X[i + j * N]);
                s += r * r;
                if (!fresh) {
                    double f = 1.0;
                    if (pmode == 1) f = 1.0 / hypot(D[i] - pt, eta);
                    else if (pmode == 2) f = D[i];
                    R[i + j * N] = T(f * r);
                }
            }
            acc[j] = s;
        }
    }
    __shared__ double red[4][LS_WMAX];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < LS_WMAX; ++j) {
        const double v = wave_sum(acc[j]);
        if (lane == 0) red[wv][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < ncol) {
        const int j = threadIdx.x;
        a.part[(size_t(q) * a.nchunk + blockIdx.x) * LS_E + LS_NORM + j] = red[0][j] + red[1][j] + red[2][j] + red[3][j];
    }
}

// Diagonal expectations of the columns of X: sum D x^2 and sum x^2 per column
// (_diagonal_expectation!, src/lobpcg.jl:73-96), into the norm partials (2w).
template <class T>
__global__ __launch_bounds__(256) void k_ls_dexp(const LsArgs<T> a, const double* __restrict__ Dall) {
    const int q = blockIdx.y;
    const LsSlot sl = a.slot[q];
    const int64_t N = a.N;
    const T* X = a.S[sl.par] + sl.p * a.blk;
    const double* D = Dall + sl.p * N;
    const int64_t r0 = blockIdx.x * a.rows_per, r1 = min(N, r0 + a.rows_per);
    __shared__ double red[4][2 * LS_WMAX];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = 0; j < a.W; ++j) {
        double num = 0, den = 0;
        for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) {
            const double x = double(X[i + j * N]);
            num += D[i] * x * x;
            den += x * x;
        }
        num = wave_sum(num);
        den = wave_sum(den);
        if (lane == 0) {
            red[wv][2 * j] = num;
            red[wv][2 * j + 1] = den;
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * a.W) {
        const int j = threadIdx.x;
        a.part[(size_t(q) * a.nchunk + blockIdx.x) * LS_E + LS_NORM + j] = red[0][j] + red[1][j] + red[2][j] + red[3][j];
    }
}

// pth[p w + j] = num / den from the reduced expectations
__global__ void k_ls_ratio(const LsSlot* __restrict__ slot, int nact, int W, const double* __restrict__ red,
                           double* __restrict__ pth) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nact * W) return;
    const int q = t / W, j = t - q * W;
    const double* r = red + size_t(q) * LS_E + LS_NORM;
    pth[slot[q].p * W + j] = r[2 * j] / r[2 * j + 1];
}

// G = S^T S and K = S^T (H S) over the first c columns of [X R P], upper 16 x 16
// tiles only, on the FP64 matrix cores.  A block takes one chunk of rows of one
// point; each 32-row slab of S and HS is staged in LDS as FP64 and every wave
// owns whole output tiles (no cross-wave reduction), the next slab prefetched
// into registers during the current one.  Tile t < nt is G's, t >= nt K's, in
// the order (0,0), (0,1), ..., (0,T-1), (1,1), ...
template <class T>
__global__ __launch_bounds__(256) void k_ls_gram(const LsArgs<T> a) {
    constexpr int LD = LS_SR + 1;
    __shared__ double sS[LS_CMAX * LD];
    __shared__ double sA[LS_CMAX * LD];
    const int q = blockIdx.y;
    const LsSlot sl = a.slot[q];
    const int64_t N = a.N;
    const T* S = a.S[sl.par] + sl.p * a.blk;
    const T* AS = a.A[sl.par] + sl.p * a.blk;
    const int c = sl.c, T16 = (c + 15) >> 4, nt = T16 * (T16 + 1) / 2;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int tI[5], tJ[5];
    bool tK[5], tv[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int t = wv + 4 * k;
        tv[k] = t < 2 * nt;
        tK[k] = t >= nt;
        int u = tK[k] ? t - nt : t, I = 0;
        while (I < T16 && u >= T16 - I) {
            u -= T16 - I;
            ++I;
        }
        tI[k] = I;
        tJ[k] = I + u;
    }
    ld4 acc[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) acc[k] = ld4{0, 0, 0, 0};
    const int64_t r0 = blockIdx.x * a.rows_per, r1 = min(N, r0 + a.rows_per);
    const int lr = tid & (LS_SR - 1), cg = tid / LS_SR;   // staging: row in the slab, column group (8)
    double pS[8], pA[8];
    auto fetch = [&](int64_t rb) {
        const int64_t row = rb + lr;
        const bool ok = row < r1;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int col = cg + 8 * u;
            pS[u] = (ok && col < c) ? double(S[row + col * N]) : 0.0;
            pA[u] = (ok && col < c) ? double(AS[row + col * N]) : 0.0;
        }
    };
    if (r0 < r1) fetch(r0);
    for (int64_t rb = r0; rb < r1; rb += LS_SR) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int col = cg + 8 * u;
            sS[col * LD + lr] = pS[u];
            sA[col * LD + lr] = pA[u];
        }
        __syncthreads();
        if (rb + LS_SR < r1) fetch(rb + LS_SR);
#pragma unroll
        for (int ks = 0; ks < LS_SR / 4; ++ks) {
            const int rr = 4 * ks + (lane >> 4);
#pragma unroll
            for (int k = 0; k < 5; ++k)
                if (tv[k]) {
                    const double av = sS[(16 * tI[k] + (lane & 15)) * LD + rr];
                    const double bv = (tK[k] ? sA : sS)[(16 * tJ[k] + (lane & 15)) * LD + rr];
                    acc[k] = mfma64(av, bv, acc[k]);
                }
        }
    }
    double* out = a.part + (size_t(q) * a.nchunk + blockIdx.x) * LS_E;
#pragma unroll
    for (int k = 0; k < 5; ++k)
        if (tv[k]) {
            const int t = wv + 4 * k;
            // f64 accumulator layout: register r of lane l holds row l/16 + 4r, column l%16
#pragma unroll
            for (int r = 0; r < 4; ++r) out[t * 256 + ((lane >> 4) + 4 * r) * 16 + (lane & 15)] = acc[k][r];
        }
}

// red[q][e] = sum over chunks of part[q][chunk][e], e < ne
__global__ __launch_bounds__(256) void k_ls_sum(const double* __restrict__ part, int nchunk, int ne,
                                                double* __restrict__ red) {
    const int q = blockIdx.y;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ne) return;
    const double* p = part + size_t(q) * nchunk * LS_E + e;
    double s = 0;
    for (int c = 0; c < nchunk; ++c) s += p[size_t(c) * LS_E];
    red[size_t(q) * LS_E + e] = s;
}

// Rotation: [X' P'] = S[:, 0:c] M (M c x nout, nout = w + np <= 2w, from the slot)
// into the other buffer, X' at column 0 and P' at column 2w, for S (blockIdx.z = 0)
// and H S (1).
// A block takes a chunk of rows of one point in 64-row slabs staged in LDS; wave
// v forms rows 16 v .. 16 v + 15 of the slab as out^T = M^T S^T on the FP64
// matrix cores, so that each store writes 16 consecutive rows of one column.
template <class T>
__global__ __launch_bounds__(256) void k_ls_rot(const LsArgs<T> a, const double* __restrict__ Mall) {
    constexpr int LD = 65;
    __shared__ double sl[LS_CMAX * LD];
    __shared__ double sM[LS_CMAX * LS_MO];
    const int q = blockIdx.y;
    const LsSlot st = a.slot[q];
    const int64_t N = a.N;
    const int W = a.W;
    const T* Src = (blockIdx.z ? a.A : a.S)[st.par] + st.p * a.blk;
    T* Dst = (blockIdx.z ? a.A : a.S)[st.par ^ 1] + st.p * a.blk;
    const int c = st.c, c4 = (c + 3) & ~3, nout = st.no, nto = (nout + 15) >> 4;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const double* M = Mall + size_t(q) * LS_CMAX * LS_MO;
    for (int e = tid; e < LS_CMAX * LS_MO; e += 256) sM[e] = M[e];
    const int64_t r0 = blockIdx.x * a.rows_per, r1 = min(N, r0 + a.rows_per);
    const int lr = tid & 63, cg = tid >> 6;
    // the next 64-row slab is fetched into registers while the current one is
    // contracted (column cg + 4u of row lr in pf[u]; c4 <= 64 columns)
    double pf[LS_CMAX / 4];
    auto fetch = [&](int64_t rb) {
        const int64_t row = rb + lr;
        const bool ok = row < r1;
#pragma unroll
        for (int u = 0; u < LS_CMAX / 4; ++u) {
            const int col = cg + 4 * u;
            pf[u] = (ok && col < c) ? double(Src[row + col * N]) : 0.0;
        }
    };
    if (r0 < r1) fetch(r0);
    for (int64_t rb = r0; rb < r1; rb += 64) {
        __syncthreads();   // the previous slab (and, first, sM) is consumed / written
#pragma unroll
        for (int u = 0; u < LS_CMAX / 4; ++u) {
            const int col = cg + 4 * u;
            if (col < c4) sl[col * LD + lr] = pf[u];
        }
        __syncthreads();
        if (rb + 64 < r1) fetch(rb + 64);
#pragma unroll
        for (int O = 0; O < 3; ++O) {
            if (O >= nto) break;
            ld4 acc = ld4{0, 0, 0, 0};
            for (int ks = 0; ks < c4; ks += 4) {
                const int k = ks + (lane >> 4);
                const double av = sM[k * LS_MO + 16 * O + (lane & 15)];
                const double bv = sl[k * LD + 16 * wv + (lane & 15)];
                acc = mfma64(av, bv, acc);
            }
            const int64_t orow = rb + 16 * wv + (lane & 15);
            if (orow < r1) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int oc = 16 * O + (lane >> 4) + 4 * r;
                    if (oc < nout) {
                        const int dc = oc < W ? oc : 2 * W + (oc - W);
                        Dst[orow + (int64_t)dc * N] = T(acc[r]);
                    }
                }
            }
        }
    }
}

template <class S, class T>
__global__ void k_ls_convert(const S* __restrict__ X, T* __restrict__ Y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        Y[i] = T(X[i]);
}

// Pinned host memory (hipHostMalloc) for the per-iteration transfers.
struct PinBuf {
    void* p = nullptr;
    size_t bytes = 0;
    PinBuf() = default;
    PinBuf(const PinBuf&) = delete;
    PinBuf& operator=(const PinBuf&) = delete;
    ~PinBuf() {
        if (p) (void)hipHostFree(p);
    }
    hipError_t ensure(size_t n) {
        if (n <= bytes) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipHostMalloc(&p, n, hipHostMallocDefault);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// A fixed team of host threads for the per-point Rayleigh-Ritz problems.
class Team {
  public:
    explicit Team(int n) {
        for (int i = 0; i < n; ++i) th_.emplace_back([this] { loop(); });
    }
    ~Team() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    int size() const { return int(th_.size()); }
    void run(int n, const std::function<void(int)>& fn) {
        if (th_.empty() || n < 2) {
            for (int i = 0; i < n; ++i) fn(i);
            return;
        }
        start(n, fn);
        finish();
    }
    // fn(0 .. n-1) as run() does (tasks are taken in index order); as each consecutive
    // group of `chunk` tasks finishes, the calling thread runs group(lo, hi) for it, in
    // group order, while the team works on later tasks.  The first nonzero return of
    // group() is returned once every task has finished (later groups are not handed on).
    int run_groups(int n, int chunk, const std::function<void(int)>& fn, const std::function<int(int, int)>& group) {
        chunk = std::max(1, chunk);
        const int ng = (n + chunk - 1) / chunk;
        if (th_.empty() || n < 2 || ng < 2) {
            run(n, fn);
            return n > 0 ? group(0, n) : 0;
        }
        std::vector<std::atomic<int>> left(ng);
        for (int g = 0; g < ng; ++g) left[g].store(std::min(chunk, n - g * chunk));
        const std::function<void(int)> task = [&](int i) {
            fn(i);
            if (left[i / chunk].fetch_sub(1) == 1) {
                std::lock_guard<std::mutex> g(mu_);
                group_.notify_one();
            }
        };
        start(n, task);
        int rc = 0;
        for (int g = 0; g < ng; ++g) {
            {
                std::unique_lock<std::mutex> l(mu_);
                group_.wait(l, [&] { return left[g].load() == 0; });
            }
            if (rc == 0) rc = group(g * chunk, std::min(n, (g + 1) * chunk));
        }
        finish();
        return rc;
    }

  private:
    void start(int n, const std::function<void(int)>& fn) {
        {
            std::lock_guard<std::mutex> g(mu_);
            fn_ = &fn;
            n_ = n;
            next_.store(0);
            busy_ = int(th_.size());
            ++gen_;
        }
        cv_.notify_all();
    }
    void finish() {
        std::unique_lock<std::mutex> g(mu_);
        done_.wait(g, [this] { return busy_ == 0; });
        fn_ = nullptr;
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)>* fn;
            int n;
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                fn = fn_;
                n = n_;
            }
            for (int i = next_.fetch_add(1); i < n; i = next_.fetch_add(1)) (*fn)(i);
            {
                std::lock_guard<std::mutex> g(mu_);
                if (--busy_ == 0) done_.notify_one();
            }
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_, group_;
    const std::function<void(int)>* fn_ = nullptr;
    int n_ = 0, busy_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
    std::atomic<int> next_{0};
};

// Lower Cholesky of the n x n SPD matrix A (column-major) in place; false when a
// pivot falls below floor.
bool chol_lower(int n, std::vector<double>& A, double floor) {
    for (int j = 0; j < n; ++j) {
        double d = A[j + j * n];
        for (int k = 0; k < j; ++k) d -= A[j + k * n] * A[j + k * n];
        if (!(d > floor)) return false;
        d = std::sqrt(d);
        A[j + j * n] = d;
        for (int i = j + 1; i < n; ++i) {
            double s = A[i + j * n];
            for (int k = 0; k < j; ++k) s -= A[i + k * n] * A[j + k * n];
            A[i + j * n] = s / d;
        }
    }
    return true;
}

// The Rayleigh-Ritz step of one point from its reduced record, the Grams G and K of
// the first c columns of [X R P] (X: w, R: w, P: c - 2w; c = w on the first step).
// As lobpcg_core does with soft locking, the R columns of the nlock leading
// converged pairs are left out of the projected problem.  Returns theta (w) and the
// rotation M (LS_CMAX x LS_MO, row k = basis column, zero padded): X' in columns
// 0..w-1 and P' in w..w+np-1, where P' is the part of the active new Ritz vectors
// outside X', orthonormalized against X' and itself in the Gram inner product
// (columns below 1e-10, FP32: 1e-4, dropped), so that the next basis is close to orthonormal
// and the implicit H S rotation stays stable.  single: the blocks are FP32, whose
// Grams carry ~1e-7 noise, so near-dependent directions are cut at that level.
// When [X R P] holds fewer than w well-separated directions the step is retried
// without P (the usual LOBPCG restart); false when that fails too.
bool ls_rr_solve(const std::vector<int>& U, int c, int W, bool single, const std::vector<double>& G,
                 const std::vector<double>& K, std::vector<double>& theta, std::vector<double>& C) {
    const int n = int(U.size());
    std::vector<double> Gs(size_t(n) * n), Ks(size_t(n) * n), d(n);
    for (int i = 0; i < n; ++i) {
        const double g = G[U[i] + size_t(U[i]) * c];
        d[i] = g > 0 ? 1.0 / std::sqrt(g) : 0.0;
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            Gs[i + size_t(j) * n] = G[U[i] + size_t(U[j]) * c] * d[i] * d[j];
            Ks[i + size_t(j) * n] = K[U[i] + size_t(U[j]) * c] * d[i] * d[j];
        }
    const double chol_floor = single ? 1e-4 : 1e-10, drop = single ? 1e-5 : 1e-12;
    int k = 0;
    std::vector<double> B;
    std::vector<double> L = Gs;
    // B upper triangular (the Cholesky path): the products below skip its zero half
    bool upper = false;
    if (chol_lower(n, L, chol_floor)) {
        k = n;
        upper = true;
        // B = L^{-T}: back substitution on L^T B = I (B[i, j] = 0 for i > j)
        B.assign(size_t(n) * n, 0.0);
        for (int j = 0; j < n; ++j)
            for (int i = j; i >= 0; --i) {
                double s = i == j ? 1.0 : 0.0;
                for (int l = i + 1; l <= j; ++l) s -= L[l + size_t(i) * n] * B[l + size_t(j) * n];
                B[i + size_t(j) * n] = s / L[i + size_t(i) * n];
            }
    } else {
        std::vector<double> mu, Uv, Gc = Gs;
        sym_eig(n, Gc, mu, Uv);
        const double top = mu.empty() ? 0.0 : mu.back();
        std::vector<int> keep;
        for (int j = 0; j < n; ++j)
            if (mu[j] > drop * top) keep.push_back(j);
        k = int(keep.size());
        B.assign(size_t(n) * k, 0.0);
        for (int jj = 0; jj < k; ++jj) {
            const int j = keep[jj];
            const double f = 1.0 / std::sqrt(mu[j]);
            for (int i = 0; i < n; ++i) B[i + size_t(jj) * n] = Uv[i + size_t(j) * n] * f;
        }
    }
    if (k < W) return false;
    std::vector<double> KB(size_t(n) * k, 0.0), Kp(size_t(k) * k);
    for (int j = 0; j < k; ++j)
        for (int l = 0; l < n; ++l) {
            const double b = B[l + size_t(j) * n];
            if (b == 0.0) continue;
            for (int i = 0; i < n; ++i) KB[i + size_t(j) * n] += Ks[i + size_t(l) * n] * b;
        }
    for (int j = 0; j < k; ++j)
        for (int i = 0; i <= j; ++i) {
            double s = 0;
            const int le = upper ? i + 1 : n;   // B[l, i] = 0 past the diagonal
            for (int l = 0; l < le; ++l) s += B[l + size_t(i) * n] * KB[l + size_t(j) * n];
            Kp[i + size_t(j) * k] = Kp[j + size_t(i) * k] = s;
        }
    std::vector<double> lam, Y;
    sym_eig_lowest(k, W, Kp, lam, Y);
    theta.assign(lam.begin(), lam.begin() + W);
    // C (c x w, full basis index) = D B Y[:, 0:w] on the rows of U
    C.assign(size_t(c) * W, 0.0);
    for (int j = 0; j < W; ++j)
        for (int i = 0; i < n; ++i) {
            double s = 0;
            for (int l = upper ? i : 0; l < k; ++l) s += B[i + size_t(l) * n] * Y[l + size_t(j) * k];   // B[i, l] = 0 for l < i
            C[U[i] + size_t(j) * c] = d[i] * s;
        }
    return true;
}

bool ls_rayleigh_ritz(int c, int W, int nlock, bool single, const double* rec, std::vector<double>& theta, double* M,
                      int& np_out) {
    const int T16 = (c + 15) >> 4, nt = T16 * (T16 + 1) / 2;
    std::vector<double> G(size_t(c) * c), K(size_t(c) * c);
    int t = 0;
    for (int I = 0; I < T16; ++I)
        for (int J = I; J < T16; ++J, ++t)
            for (int r = 0; r < 16; ++r)
                for (int s = 0; s < 16; ++s) {
                    const int i = 16 * I + r, j = 16 * J + s;
                    if (i >= c || j >= c) continue;
                    const double g = rec[t * 256 + r * 16 + s], k = rec[(nt + t) * 256 + r * 16 + s];
                    G[i + size_t(j) * c] = g;
                    K[i + size_t(j) * c] = k;
                    if (I != J) {
                        G[j + size_t(i) * c] = g;
                        K[j + size_t(i) * c] = k;
                    }
                }
    // diagonal tiles hold both triangles: symmetrize (S^T H S is symmetric up to rounding)
    for (int j = 0; j < c; ++j)
        for (int i = j + 1; i < c; ++i)
            if ((i >> 4) == (j >> 4)) {
                const double v = 0.5 * (K[i + size_t(j) * c] + K[j + size_t(i) * c]);
                K[i + size_t(j) * c] = K[j + size_t(i) * c] = v;
                const double g = 0.5 * (G[i + size_t(j) * c] + G[j + size_t(i) * c]);
                G[i + size_t(j) * c] = G[j + size_t(i) * c] = g;
            }
    // the used basis: X, the active R columns, P
    auto basis = [&](bool withP) {
        std::vector<int> U;
        for (int i = 0; i < W; ++i) U.push_back(i);
        if (c > W)
            for (int j = nlock; j < W; ++j) U.push_back(W + j);
        if (withP)
            for (int i = 2 * W; i < c; ++i) U.push_back(i);
        return U;
    };
    std::vector<double> C;
    if (!ls_rr_solve(basis(true), c, W, single, G, K, theta, C)) {
        if (c <= 2 * W || !ls_rr_solve(basis(false), c, W, single, G, K, theta, C)) return false;
    }
    std::fill(M, M + size_t(LS_CMAX) * LS_MO, 0.0);
    for (int j = 0; j < W; ++j)
        for (int i = 0; i < c; ++i) M[i * LS_MO + j] = C[i + size_t(j) * c];
    np_out = 0;
    if (c == W) return true;
    // P': the active new Ritz vectors' parts outside X, G-orthonormalized against X'
    // and each other (u = G v is carried along so every inner product is O(c))
    auto gmul = [&](const double* x, std::vector<double>& y) {
        y.assign(c, 0.0);
        for (int j = 0; j < c; ++j) {
            const double xj = x[j];
            if (xj == 0.0) continue;
            for (int i = 0; i < c; ++i) y[i] += G[i + size_t(j) * c] * xj;
        }
    };
    auto dot = [&](const double* a, const double* b) {
        double s = 0;
        for (int i = 0; i < c; ++i) s += a[i] * b[i];
        return s;
    };
    std::vector<std::vector<double>> GX(W), pc, gpc;
    for (int q = 0; q < W; ++q) gmul(&C[size_t(q) * c], GX[q]);
    for (int col = nlock; col < W; ++col) {
        std::vector<double> v(c, 0.0), u;
        for (int i = W; i < c; ++i) v[i] = C[i + size_t(col) * c];
        gmul(v.data(), u);
        for (int pass = 0; pass < 2; ++pass) {
            for (int q = 0; q < W; ++q) {
                const double* x = &C[size_t(q) * c];
                const double sx = dot(x, u.data());
                for (int i = 0; i < c; ++i) {
                    v[i] -= sx * x[i];
                    u[i] -= sx * GX[q][i];
                }
            }
            for (size_t k2 = 0; k2 < pc.size(); ++k2) {
                const double sx = dot(pc[k2].data(), u.data());
                for (int i = 0; i < c; ++i) {
                    v[i] -= sx * pc[k2][i];
                    u[i] -= sx * gpc[k2][i];
                }
            }
        }
        const double nv = std::sqrt(std::max(dot(v.data(), u.data()), 0.0));
        if (nv < (single ? 1e-4 : 1e-10)) continue;
        for (auto& x : v) x /= nv;
        for (auto& x : u) x /= nv;
        pc.push_back(v);
        gpc.push_back(u);
    }
    np_out = int(pc.size());
    for (int j = 0; j < np_out; ++j)
        for (int i = 0; i < c; ++i) M[i * LS_MO + W + j] = pc[j][i];
    return true;
}

struct LsPoint {
    kfc_op* op = nullptr;
    int par = 0;
    int np = 0;      // P columns of the current basis
    int nlock = 0;   // leading converged pairs (their R leaves the projected problem)
    bool done = false, failed = false;
    int64_t it = 0, matvec = 0;
    std::vector<double> theta, res;
};

struct LsCfg {
    int m, W;
    double tol;
    int64_t maxiter;
    int precond;
    double eta;
};

// One precision phase of the lockstep iteration over every point in `pts`.
// On entry the X slot of buffer 0 holds each point's start block (any basis of
// the right span); on exit, each point's X (Ritz vectors, ascending theta) is in
// the X slot of buffer pts[i].par.
template <class T>
struct Phase {
    kfc_ctx* ctx;
    hipStream_t st;
    std::vector<LsPoint>& pts;
    const LsCfg& cfg;
    int64_t N;
    int nchunk;
    int64_t rows_per;
    bool pair_batch;
    T* S[2];
    T* A[2];
    DevBuf &dpart, &dred, &dslot, &dtheta, &dpth, &dM, &dtab;
    PinBuf &hred, &hslot, &htheta, &hM, &htab;
    const double* Dall;   // preconditioner diagonals (n x N) or null
    Team& team;
    double switch_tol;    // > 0: a point is done once every residual is below it
    double* apply_gpu;    // accumulated apply time (s)
    hipEvent_t e0, e1;
    double t_wait = 0, t_rr = 0;   // host time blocked on the stream / in the Rayleigh-Ritz steps
    int64_t n_iter = 0, n_pointiter = 0, n_collapsed = 0;
    int nchunk0 = 0;               // row chunks per point with every point active (the partials' capacity)
    // The device work of a step (the uploads of theta, the slots and the pair table,
    // the residual, the apply, the Grams, their sums and the records' download) as one
    // graph per active count: captured the second time that count occurs (its first
    // run tunes the pair launch shapes) and replayed while it lasts.  The host refills
    // the pinned tables before each replay.  KFC_GRAPHS=0: separate launches.
    std::map<int, hipGraphExec_t> gexec = {};
    std::set<int> gseen = {};
    int64_t n_graph_steps = 0, n_captures = 0;
    // host time capturing and instantiating step graphs, in the first Rayleigh-Ritz,
    // and in the steps run as separate launches (their waits included)
    double t_capture = 0, t_init = 0, t_eager = 0;
    bool apply_untimed = false;    // some applies ran inside graphs (no events there)

    ~Phase() {
        for (auto& kv : gexec) (void)hipGraphExecDestroy(kv.second);
    }

    static bool graphs_on() {
        const char* e = std::getenv("KFC_GRAPHS");
        return !(e && e[0] == '0');
    }

    // Row chunks per point for a launch over `nact` points: with fewer active points
    // each point's rows are split finer (up to 16x), so that the tail of a sweep, where
    // a few slow points remain, still fills the device.  nact * nchunk stays within the
    // partial-sum buffer (sized for every point at nchunk0 chunks).
    void shape_for(int nact) {
        if (nchunk0 == 0) nchunk0 = nchunk;
        const int f = std::max(1, std::min(16, int(pts.size()) / std::max(1, nact)));
        const int64_t want = int64_t(nchunk0) * f;
        rows_per = std::max<int64_t>(LS_SR, ((N + want - 1) / want + LS_SR - 1) / LS_SR * LS_SR);
        nchunk = int((N + rows_per - 1) / rows_per);
    }

    LsArgs<T> args(int nact) {
        LsArgs<T> a;
        a.S[0] = S[0];
        a.S[1] = S[1];
        a.A[0] = A[0];
        a.A[1] = A[1];
        a.N = N;
        a.blk = int64_t(3) * cfg.W * N;
        a.W = cfg.W;
        a.nchunk = nchunk;
        a.rows_per = rows_per;
        a.slot = dslot.as<LsSlot>();
        a.part = dpart.as<double>();
        (void)nact;
        return a;
    }

    // Slot tables: half 0 for the steps of an iteration, half 1 for the rotation
    // (its upload may still be in flight when the next iteration fills half 0).
    int upload_slots(const std::vector<LsSlot>& sl, int half = 0) {
        fill_slots(sl, half);
        return copy_slots(int(sl.size()), half);
    }
    void fill_slots(const std::vector<LsSlot>& sl, int half) {
        std::copy(sl.begin(), sl.end(), hslot.as<LsSlot>() + size_t(half) * pts.size());
    }
    int copy_slots(int count, int half) {
        LsSlot* h = hslot.as<LsSlot>() + size_t(half) * pts.size();
        LsSlot* d = dslot.as<LsSlot>() + size_t(half) * pts.size();
        KFC_HIP(ctx, hipMemcpyAsync(d, h, sizeof(LsSlot) * count, hipMemcpyHostToDevice, st));
        return KFC_OK;
    }

    // H applied to `ncol` columns at column `c0` of each listed point's buffer
    // `par` (S), written to column `c0` of A[par ^ flip]
    int apply(const std::vector<int>& list, int c0, int ncol, int flip) {
        const int n = int(list.size());
        if (n == 0) return KFC_OK;
        const int prec = std::is_same<T, double>::value ? KFC_PREC_F64 : KFC_PREC_F32;
        KFC_HIP(ctx, hipEventRecord(e0, st));
        std::vector<kfc_op*> ops(n);
        std::vector<const void*> X(n);
        std::vector<void*> Y(n);
        for (int q = 0; q < n; ++q) {
            LsPoint& pt = pts[list[q]];
            ops[q] = pt.op;
            const int64_t off = int64_t(list[q]) * 3 * cfg.W * N + int64_t(c0) * N;
            X[q] = S[pt.par] + off;
            Y[q] = A[pt.par ^ flip] + off;
            pt.matvec += ncol;
        }
        if (pair_batch) {
            int rc = launch_pairs_batch<T>(ops.data(), n, X.data(), Y.data(), ncol, N, st, htab.p, dtab.p);
            if (rc != KFC_OK) return rc;
        } else {
            for (int q = 0; q < n; ++q) {
                int rc = op_apply(ops[q], X[q], Y[q], ncol, N, prec, st);
                if (rc != KFC_OK) return rc;
            }
        }
        KFC_HIP(ctx, hipEventRecord(e1, st));
        return KFC_OK;
    }

    int read_apply_time() {
        float ms = 0;
        KFC_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
        *apply_gpu += 1e-3 * ms;
        return KFC_OK;
    }

    int reduce_and_fetch(int nact, int ne) {
        hipLaunchKernelGGL(k_ls_sum, dim3((ne + 255) / 256, nact), dim3(256), 0, st, dpart.as<double>(), nchunk, ne,
                           dred.as<double>());
        KFC_HIP(ctx, hipGetLastError());
        KFC_HIP(ctx, hipMemcpyAsync(hred.p, dred.p, sizeof(double) * LS_E * nact, hipMemcpyDeviceToHost, st));
        const auto t0 = clk::now();
        KFC_HIP(ctx, hipStreamSynchronize(st));
        t_wait += secs(t0);
        return KFC_OK;
    }

    int upload_theta() {
        fill_theta();
        return copy_theta();
    }
    void fill_theta() {
        double* h = htheta.as<double>();
        for (size_t p = 0; p < pts.size(); ++p)
            for (int j = 0; j < cfg.W; ++j) h[p * cfg.W + j] = pts[p].theta.empty() ? 0.0 : pts[p].theta[j];
    }
    int copy_theta() {
        KFC_HIP(ctx, hipMemcpyAsync(dtheta.p, htheta.p, sizeof(double) * pts.size() * cfg.W, hipMemcpyHostToDevice, st));
        return KFC_OK;
    }

    // The pair table of H applied to the R columns (W of them at column W) of each
    // listed point, written to the R columns of A[par]; the matvecs counted.
    void fill_apply_r(const std::vector<int>& list) {
        const int n = int(list.size());
        std::vector<kfc_op*> ops(n);
        std::vector<const void*> X(n);
        std::vector<void*> Y(n);
        for (int q = 0; q < n; ++q) {
            LsPoint& pt = pts[list[q]];
            ops[q] = pt.op;
            const int64_t off = int64_t(list[q]) * 3 * cfg.W * N + int64_t(cfg.W) * N;
            X[q] = S[pt.par] + off;
            Y[q] = A[pt.par] + off;
            pt.matvec += cfg.W;
        }
        pairs_fill_table<T>(ops.data(), n, X.data(), Y.data(), cfg.W, N, htab.p);
    }

    // The device work of one step over nact points (see gexec): uploads, residual,
    // apply (from the filled pair table), Grams, sums, records to the host.
    int enqueue_step(const std::vector<int>& act, const double* pth, const double* D, int pmode) {
        const int nact = int(act.size());
        int rc = copy_theta();
        if (rc == KFC_OK) rc = copy_slots(nact, 0);
        if (rc != KFC_OK) return rc;
        const LsArgs<T> a = args(nact);
        hipLaunchKernelGGL(k_ls_resid<T>, dim3(nchunk, nact), dim3(256), 0, st, a, cfg.W, dtheta.as<double>(), pth, D,
                           pmode, cfg.eta, 0);
        KFC_HIP(ctx, hipGetLastError());
        std::vector<kfc_op*> ops(nact);
        for (int q = 0; q < nact; ++q) ops[q] = pts[act[q]].op;
        rc = launch_pairs_filled<T>(ops.data(), nact, st, htab.p, dtab.p);
        if (rc != KFC_OK) return rc;
        hipLaunchKernelGGL(k_ls_gram<T>, dim3(nchunk, nact), dim3(256), 0, st, a);
        KFC_HIP(ctx, hipGetLastError());
        const int ne = LS_NORM + cfg.W;
        hipLaunchKernelGGL(k_ls_sum, dim3((ne + 255) / 256, nact), dim3(256), 0, st, dpart.as<double>(), nchunk, ne,
                           dred.as<double>());
        KFC_HIP(ctx, hipGetLastError());
        KFC_HIP(ctx, hipMemcpyAsync(hred.p, dred.p, sizeof(double) * LS_E * nact, hipMemcpyDeviceToHost, st));
        return KFC_OK;
    }

    // One step through its graph (captured or replayed); false in *done when the step
    // has to run as launches (the first time its active count occurs).
    int graph_step(const std::vector<int>& act, const double* pth, const double* D, int pmode, bool* done) {
        const int nact = int(act.size());
        *done = false;
        auto it = gexec.find(nact);
        if (it == gexec.end()) {
            if (!gseen.count(nact)) {
                gseen.insert(nact);
                return KFC_OK;
            }
            const auto tc = clk::now();
            KFC_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_step(act, pth, D, pmode);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(st, &g);
            if (rc != KFC_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return kfc_fail(ctx, KFC_EDEVICE, std::string("lockstep step capture: ") + hipGetErrorString(e));
            hipGraphExec_t x = nullptr;
            const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess)
                return kfc_fail(ctx, KFC_EDEVICE, std::string("lockstep step instantiate: ") + hipGetErrorString(ei));
            it = gexec.emplace(nact, x).first;
            ++n_captures;
            t_capture += secs(tc);
        }
        KFC_HIP(ctx, hipGraphLaunch(it->second, st));
        const auto t0 = clk::now();
        KFC_HIP(ctx, hipStreamSynchronize(st));
        t_wait += secs(t0);
        ++n_graph_steps;
        apply_untimed = true;
        *done = true;
        return KFC_OK;
    }

    // Rayleigh-Ritz of each listed point on its first c columns, then the rotation.
    // Points whose basis collapses are marked failed.  The points go to the host team
    // in groups of its size; each group's rotation is launched as soon as its
    // Rayleigh-Ritz problems are solved, so the device rotates one group while the
    // host solves the next (each point's rows are rotated as in one launch).
    int rr_rotate(const std::vector<int>& list, const std::vector<int>& cs) {
        const int n = int(list.size());
        std::vector<char> ok(n, 1);
        std::vector<int> npn(n, 0);
        double* Mh = hM.as<double>();
        const double* rec = hred.as<double>();
        const auto t0 = clk::now();
        n_pointiter += n;
        ++n_iter;
        constexpr size_t MW = size_t(LS_CMAX) * LS_MO;   // doubles of one rotation matrix
        int rdone = 0;   // rotated points so far: their matrices and slots are packed in front
        auto rotate_group = [&](int lo, int hi) -> int {
            std::vector<int> rl, src;
            std::vector<LsSlot> sl;
            for (int q = lo; q < hi; ++q) {
                LsPoint& pt = pts[list[q]];
                if (!ok[q]) {
                    // an FP32 basis that collapses ends the point's FP32 phase (the FP64
                    // phase restarts from its X); in FP64 the point goes to the one-point path
                    if (switch_tol > 0) pt.done = true;
                    else pt.failed = true;
                    ++n_collapsed;
                    continue;
                }
                rl.push_back(list[q]);
                sl.push_back(LsSlot{list[q], pt.par, cs[q], cfg.W + npn[q]});
                src.push_back(q);
            }
            if (rl.empty()) return KFC_OK;
            // pack this group's matrices after the earlier groups' (rdone <= lo <= src[r],
            // and the team writes only matrices at indices >= hi)
            const int r0 = rdone, nr = int(rl.size());
            for (int r = 0; r < nr; ++r)
                if (r0 + r != src[r]) std::copy(Mh + size_t(src[r]) * MW, Mh + size_t(src[r] + 1) * MW, Mh + size_t(r0 + r) * MW);
            rdone += nr;
            KFC_HIP(ctx, hipMemcpyAsync(dM.as<double>() + size_t(r0) * MW, Mh + size_t(r0) * MW, sizeof(double) * MW * nr,
                                        hipMemcpyHostToDevice, st));
            LsSlot* hs = hslot.as<LsSlot>() + pts.size() + r0;
            LsSlot* ds = dslot.as<LsSlot>() + pts.size() + r0;
            std::copy(sl.begin(), sl.end(), hs);
            KFC_HIP(ctx, hipMemcpyAsync(ds, hs, sizeof(LsSlot) * nr, hipMemcpyHostToDevice, st));
            LsArgs<T> ar = args(nr);
            // rows of 2048 per block with every point in the launch, finer (>= 64 rows) for fewer
            const int f = std::max(1, std::min(16, int(pts.size()) / std::max(1, nr)));
            const unsigned gx = unsigned(std::max<int64_t>(1, std::min<int64_t>((N + 63) / 64, ((N + 2047) / 2048) * f)));
            ar.slot = ds;
            ar.rows_per = ((N + gx - 1) / gx + 63) / 64 * 64;
            hipLaunchKernelGGL(k_ls_rot<T>, dim3(gx, nr, 2), dim3(256), 0, st, ar, dM.as<double>() + size_t(r0) * MW);
            KFC_HIP(ctx, hipGetLastError());
            for (int r = 0; r < nr; ++r) {
                pts[rl[r]].par ^= 1;
                pts[rl[r]].np = npn[src[r]];
            }
            return KFC_OK;
        };
        const int rc = team.run_groups(
            n, std::max(1, team.size()),
            [&](int q) {
                LsPoint& pt = pts[list[q]];
                ok[q] = ls_rayleigh_ritz(cs[q], cfg.W, pt.nlock, std::is_same<T, float>::value, rec + size_t(q) * LS_E,
                                         pt.theta, Mh + size_t(q) * MW, npn[q]);
            },
            rotate_group);
        t_rr += secs(t0);
        return rc;
    }

    int run(int64_t* iters_out) {
        const int W = cfg.W;
        std::vector<int> act;
        for (size_t p = 0; p < pts.size(); ++p)
            if (!pts[p].done && !pts[p].failed) act.push_back(int(p));
        for (int p : act) {
            pts[p].par = 0;
            pts[p].np = 0;
            pts[p].nlock = 0;
        }
        if (act.empty()) return KFC_OK;
        // initial Rayleigh-Ritz of the start blocks: AX, G = X^T X, K = X^T A X
        const auto ti = clk::now();
        shape_for(int(act.size()));
        int rc = apply(act, 0, W, 0);
        if (rc != KFC_OK) return rc;
        {
            std::vector<LsSlot> sl;
            for (int p : act) sl.push_back(LsSlot{p, pts[p].par, W, 0});
            rc = upload_slots(sl);
            if (rc != KFC_OK) return rc;
            hipLaunchKernelGGL(k_ls_gram<T>, dim3(nchunk, int(act.size())), dim3(256), 0, st, args(int(act.size())));
            KFC_HIP(ctx, hipGetLastError());
            rc = reduce_and_fetch(int(act.size()), LS_NORM);
            if (rc != KFC_OK) return rc;
            rc = read_apply_time();
            if (rc != KFC_OK) return rc;
            rc = rr_rotate(act, std::vector<int>(act.size(), W));
            if (rc != KFC_OK) return rc;
        }
        t_init += secs(ti);
        bool first_prep = true;
        for (int64_t it = 0;; ++it) {
            act.clear();
            for (size_t p = 0; p < pts.size(); ++p)
                if (!pts[p].done && !pts[p].failed) act.push_back(int(p));
            if (act.empty()) break;
            const int nact = int(act.size());
            shape_for(nact);
            std::vector<LsSlot> sl;
            std::vector<int> cs;
            for (int p : act) {
                const int c = 2 * W + pts[p].np;
                sl.push_back(LsSlot{p, pts[p].par, c, 0});
                cs.push_back(c);
            }
            // the RitzJacobi shifts: diagonal expectations on the first step, then the Ritz values
            const double* pth = dtheta.as<double>();
            int pmode = cfg.precond == KFC_PRECOND_RITZ ? 1 : cfg.precond == KFC_PRECOND_JACOBI ? 2 : 0;
            const double* D = pmode == 0 ? nullptr : pmode == 2 ? Dall + pts.size() * N : Dall;
            bool graphed = false;
            if (pair_batch && graphs_on() && !(pmode == 1 && first_prep)) {
                fill_theta();
                fill_slots(sl, 0);
                fill_apply_r(act);
                rc = graph_step(act, pth, D, pmode, &graphed);
                if (rc != KFC_OK) return rc;
                if (!graphed) {   // this count's first step: launches (the table is filled)
                    for (int p : act) pts[p].matvec -= W;
                }
            }
            const auto te = clk::now();
            if (!graphed) {
            rc = upload_theta();
            if (rc == KFC_OK) rc = upload_slots(sl);
            if (rc != KFC_OK) return rc;
            const LsArgs<T> a = args(nact);
            if (pmode == 1 && first_prep) {
                hipLaunchKernelGGL(k_ls_dexp<T>, dim3(nchunk, nact), dim3(256), 0, st, a, Dall);
                KFC_HIP(ctx, hipGetLastError());
                hipLaunchKernelGGL(k_ls_sum, dim3(1, nact), dim3(256), 0, st, dpart.as<double>() + LS_NORM, nchunk,
                                   2 * W, dred.as<double>() + LS_NORM);
                KFC_HIP(ctx, hipGetLastError());
                hipLaunchKernelGGL(k_ls_ratio, dim3((nact * W + 255) / 256), dim3(256), 0, st, dslot.as<LsSlot>(), nact,
                                   W, dred.as<double>(), dpth.as<double>());
                KFC_HIP(ctx, hipGetLastError());
                pth = dpth.as<double>();
                first_prep = false;
            }
            hipLaunchKernelGGL(k_ls_resid<T>, dim3(nchunk, nact), dim3(256), 0, st, a, W, dtheta.as<double>(), pth, D,
                               pmode, cfg.eta, 0);
            KFC_HIP(ctx, hipGetLastError());
            rc = apply(act, W, W, 0);
            if (rc != KFC_OK) return rc;
            hipLaunchKernelGGL(k_ls_gram<T>, dim3(nchunk, nact), dim3(256), 0, st, a);
            KFC_HIP(ctx, hipGetLastError());
            rc = reduce_and_fetch(nact, LS_NORM + W);
            if (rc != KFC_OK) return rc;
            rc = read_apply_time();
            if (rc != KFC_OK) return rc;
            t_eager += secs(te);
            }
            // convergence of every column on the residuals of the current X
            std::vector<int> go;
            std::vector<int> gcs;
            std::vector<int> rec_of;   // record index of each continuing point
            const double* rec = hred.as<double>();
            for (int q = 0; q < nact; ++q) {
                LsPoint& pt = pts[act[q]];
                pt.res.assign(W, 0.0);
                const double tl = switch_tol > 0 ? switch_tol : cfg.tol;
                bool all = true;
                for (int j = 0; j < W; ++j) {
                    const double r = std::sqrt(std::max(0.0, rec[size_t(q) * LS_E + LS_NORM + j]));
                    pt.res[j] = r;
                    all &= r <= tl;
                }
                // soft locking: the leading converged pairs (lobpcg_core's prefix rule)
                pt.nlock = 0;
                while (pt.nlock < W && pt.res[pt.nlock] <= tl) ++pt.nlock;
                pt.it = it;
                if (all || it >= cfg.maxiter) {
                    pt.done = true;
                    continue;
                }
                go.push_back(act[q]);
                gcs.push_back(cs[q]);
                rec_of.push_back(q);
            }
            if (go.empty()) break;
            // move the continuing points' records to the front (rr_rotate reads record q for list[q])
            double* hr = hred.as<double>();
            for (size_t r = 0; r < rec_of.size(); ++r)
                if (int(r) != rec_of[r])
                    std::copy(hr + size_t(rec_of[r]) * LS_E, hr + size_t(rec_of[r] + 1) * LS_E, hr + r * LS_E);
            rc = rr_rotate(go, gcs);
            if (rc != KFC_OK) return rc;
            if (iters_out) *iters_out = it + 1;
        }
        return KFC_OK;
    }
};

// The start blocks from a learned subspace (the subspace seed of
// kfc_lobpcg_lockstep_subspace): point i's block spans S_i = P_i B C_i, where B
// (nf x nb) is a basis of a parameter-independent frame of the parent basis, P_i
// places frame row f at parent row rowmap_i[f] (none when -1) and C_i (nb x ns)
// picks ns directions (the identity when absent).  The start block is the lowest
// w Ritz vectors of H_i on span(S_i): K = S^T H S and G = S^T S (FP64 Grams on
// the device), the projected problem on host threads (the generalized form only
// when G is not the identity, i.e. when frame rows were dropped).
struct SubspaceSeed {
    const double* B = nullptr;   // device, nf x nb column-major
    int64_t nf = 0;
    int nb = 0, ns = 0;
    const int64_t* const* rowmaps = nullptr;   // host, n arrays of nf entries
    const double* const* C = nullptr;          // host, n arrays nb x ns, or null
};

__global__ __launch_bounds__(256) void k_frame_scatter(const double* __restrict__ BC, int64_t ldb, int64_t nf, int ns,
                                                       const int64_t* __restrict__ rowmap, double* __restrict__ X,
                                                       int64_t ldx) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nf * ns; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = e % nf, j = e / nf;
        const int64_t r = rowmap[f];
        if (r >= 0) X[r + j * ldx] = BC[f + j * ldb];
    }
}

// S_i into X (N x ns, leading dimension N): zero, then the frame rows scattered
int subspace_block(Work& w, const SubspaceSeed& sub, int i, int64_t N, DevBuf& BC, DevBuf& map, double* X) {
    kfc_ctx* ctx = w.ctx;
    hipStream_t st = ctx->stream;
    const double* src = sub.B;
    if (sub.C && sub.C[i]) {
        std::vector<cd> M(size_t(sub.nb) * sub.ns);
        for (size_t k = 0; k < M.size(); ++k) M[k] = sub.C[i][k];
        KFC_HIP(ctx, BC.ensure(sizeof(double) * size_t(sub.nf) * sub.ns));
        int rc = blas_gemm<double>(w, sub.B, sub.nf, sub.nb, M, sub.ns, BC.as<double>(), sub.nf, sub.nf, 0.0);
        if (rc != KFC_OK) return rc;
        src = BC.as<double>();
    }
    KFC_HIP(ctx, map.ensure(sizeof(int64_t) * size_t(sub.nf)));
    KFC_HIP(ctx, hipMemcpyAsync(map.p, sub.rowmaps[i], sizeof(int64_t) * size_t(sub.nf), hipMemcpyHostToDevice, st));
    KFC_HIP(ctx, hipMemsetAsync(X, 0, sizeof(double) * size_t(N) * sub.ns, st));
    const int64_t tot = sub.nf * sub.ns;
    const unsigned gx = unsigned(std::max<int64_t>(1, std::min<int64_t>((tot + 255) / 256, 8192)));
    hipLaunchKernelGGL(k_frame_scatter, dim3(gx), dim3(256), 0, st, src, sub.nf, sub.nf, sub.ns, map.as<int64_t>(), X, N);
    KFC_HIP(ctx, hipGetLastError());
    return KFC_OK;   // the caller's rowmap outlives the call, hence the copy
}

// The lowest W Ritz vectors of every point on its subspace, written as its start block
// (the X slot of S64, point stride 3 W N).
int subspace_starts(kfc_op* const* ops, int n, int W, const SubspaceSeed& sub, double* S0, Team& team) {
    kfc_ctx* ctx = ops[0]->ctx;
    const int64_t N = ops[0]->N;
    const int ns = sub.ns;
    const bool trace = std::getenv("KFC_TRACE") != nullptr;
    auto t0 = clk::now();
    Work w;
    w.ctx = ctx;
    DevBuf X, HX, BC, map;
    KFC_HIP(ctx, X.ensure(sizeof(double) * size_t(N) * ns));
    KFC_HIP(ctx, HX.ensure(sizeof(double) * size_t(N) * ns));
    std::vector<std::vector<cd>> G(n), K(n);
    // S^T S = (B C)^T (B C) when no frame row is dropped: formed once for such points
    std::vector<char> full(n, 1);
    for (int i = 0; i < n; ++i)
        for (int64_t f = 0; f < sub.nf && full[i]; ++f) full[i] = sub.rowmaps[i][f] >= 0;
    std::vector<cd> GB;
    bool have_gb = false;
    for (int i = 0; i < n; ++i) {
        int rc = subspace_block(w, sub, i, N, BC, map, X.as<double>());
        if (rc == KFC_OK) rc = op_apply(ops[i], X.p, HX.p, ns, N, KFC_PREC_F64, ctx->stream);
        if (rc == KFC_OK && !full[i]) rc = blas_gram<double>(w, X.as<double>(), N, ns, X.as<double>(), N, ns, N, G[i]);
        if (rc == KFC_OK) rc = blas_gram<double>(w, X.as<double>(), N, ns, HX.as<double>(), N, ns, N, K[i]);
        if (rc != KFC_OK) return kfc_fail(ctx, rc, "subspace start of point " + std::to_string(i) + ": " + ops[i]->ctx->err);
        if (full[i] && !sub.C && !have_gb) {   // B^T B, the same for every such point
            rc = blas_gram<double>(w, sub.B, sub.nf, ns, sub.B, sub.nf, ns, sub.nf, GB);
            if (rc != KFC_OK) return kfc_fail(ctx, rc, "subspace start: " + ctx->err);
            have_gb = true;
        }
        if (full[i]) {
            if (sub.C) {   // (B C_i)^T (B C_i), its own Gram
                rc = blas_gram<double>(w, X.as<double>(), N, ns, X.as<double>(), N, ns, N, G[i]);
                if (rc != KFC_OK) return kfc_fail(ctx, rc, "subspace start: " + ctx->err);
            } else {
                G[i] = GB;
            }
        }
    }
    const double t_dev = secs(t0);
    std::vector<std::vector<double>> Y(n);
    std::vector<char> ok(n, 1);
    team.run(n, [&](int i) {
        std::vector<double> g(size_t(ns) * ns), k(size_t(ns) * ns);
        double off = 0;
        for (int b = 0; b < ns; ++b)
            for (int a = 0; a < ns; ++a) {
                g[a + size_t(b) * ns] = 0.5 * (G[i][a + size_t(b) * ns].real() + G[i][b + size_t(a) * ns].real());
                k[a + size_t(b) * ns] = 0.5 * (K[i][a + size_t(b) * ns].real() + K[i][b + size_t(a) * ns].real());
                off = std::max(off, std::fabs(g[a + size_t(b) * ns] - (a == b ? 1.0 : 0.0)));
            }
        std::vector<double> lam, th;
        if (off <= 1e-12) {   // orthonormal S: the standard problem
            sym_eig_lowest(ns, W, k, lam, Y[i]);
        } else {
            std::vector<int> U(ns);
            for (int a = 0; a < ns; ++a) U[a] = a;
            ok[i] = ls_rr_solve(U, ns, W, false, g, k, th, Y[i]);
        }
    });
    const double t_rr = secs(t0) - t_dev;
    for (int i = 0; i < n; ++i)
        if (!ok[i])
            return kfc_fail(ctx, KFC_EBREAKDOWN, "subspace start of point " + std::to_string(i) +
                                                     ": the subspace holds fewer than " + std::to_string(W) +
                                                     " independent directions");
    if (!sub.C) {
        // S_i Y_i = scatter_i(B Y_i), S_i's rows being B's rows placed by the point's row
        // map (zero elsewhere): one product of B with every point's Y_i, then W columns
        // scattered per point (the same sums as S_i Y_i, row by row)
        std::vector<cd> M(size_t(ns) * W * n);
        for (int i = 0; i < n; ++i)
            for (size_t k = 0; k < size_t(ns) * W; ++k) M[size_t(i) * ns * W + k] = Y[i][k];
        DevBuf Z;
        KFC_HIP(ctx, Z.ensure(sizeof(double) * size_t(sub.nf) * W * n));
        int rc = blas_gemm<double>(w, sub.B, sub.nf, ns, M, W * n, Z.as<double>(), sub.nf, sub.nf, 0.0);
        if (rc != KFC_OK) return kfc_fail(ctx, rc, "subspace starts: " + ctx->err);
        KFC_HIP(ctx, map.ensure(sizeof(int64_t) * size_t(sub.nf)));
        const int64_t tot = sub.nf * W;
        const unsigned gx = unsigned(std::max<int64_t>(1, std::min<int64_t>((tot + 255) / 256, 8192)));
        for (int i = 0; i < n; ++i) {
            double* Si = S0 + size_t(i) * 3 * W * N;
            KFC_HIP(ctx, hipMemcpyAsync(map.p, sub.rowmaps[i], sizeof(int64_t) * size_t(sub.nf), hipMemcpyHostToDevice,
                                        ctx->stream));
            KFC_HIP(ctx, hipMemsetAsync(Si, 0, sizeof(double) * size_t(N) * W, ctx->stream));
            hipLaunchKernelGGL(k_frame_scatter, dim3(gx), dim3(256), 0, ctx->stream, Z.as<double>() + size_t(i) * W * sub.nf,
                               sub.nf, sub.nf, W, map.as<int64_t>(), Si, N);
            KFC_HIP(ctx, hipGetLastError());
        }
        KFC_HIP(ctx, hipStreamSynchronize(ctx->stream));   // Z is freed on return
    } else {
        for (int i = 0; i < n; ++i) {
            int rc = subspace_block(w, sub, i, N, BC, map, X.as<double>());
            std::vector<cd> M(Y[i].size());
            for (size_t k = 0; k < M.size(); ++k) M[k] = Y[i][k];
            if (rc == KFC_OK)
                rc = blas_gemm<double>(w, X.as<double>(), N, ns, M, W, S0 + size_t(i) * 3 * W * N, N, N, 0.0);
            if (rc != KFC_OK)
                return kfc_fail(ctx, rc, "subspace start of point " + std::to_string(i) + ": " + ops[i]->ctx->err);
        }
    }
    if (trace)
        std::fprintf(stderr, "[lockstep] subspace starts: projections %.3f s, Rayleigh-Ritz %.3f s, blocks %.3f s\n", t_dev,
                     t_rr, secs(t0) - t_dev - t_rr);
    return KFC_OK;
}

template <class T>
int alloc_phase_buffers(kfc_ctx* ctx, int n, int64_t N, int W, DevBuf* S, DevBuf* A) {
    const size_t bytes = sizeof(T) * size_t(n) * 3 * W * N;
    for (int b = 0; b < 2; ++b) {
        KFC_HIP(ctx, S[b].ensure(bytes));
        KFC_HIP(ctx, A[b].ensure(bytes));
    }
    return KFC_OK;
}

int lockstep_run(int n, kfc_op* const* ops, int64_t m, int64_t guard, double tol, int64_t maxiter, int32_t precond,
                 double eta, int32_t seed_kind, const double* const* X0s, uint64_t seed, int32_t mode, double* E,
                 double* const* Vs, kfc_stats* stats, const SubspaceSeed* sub = nullptr) {
    auto t0 = clk::now();
    // KFC_TRACE set: per-phase host timings on stderr
    const bool trace = std::getenv("KFC_TRACE") != nullptr;
    kfc_op* op0 = ops[0];
    kfc_ctx* ctx = op0->ctx;
    const int64_t N = op0->N;
    const int g = int(std::max<int64_t>(0, std::min<int64_t>(guard, (N - 5) / 3 - m)));   // src/lobpcg.jl:243
    const int W = int(m) + g;
    hipStream_t st = ctx->stream;
    // start blocks (host) and preconditioner diagonals
    std::vector<LsPoint> pts(n);
    const int64_t nx0 = seed_kind == KFC_SEED_WARM_GUARDS ? W : m;
    DevBuf S64[2], A64[2], S32[2], A32[2];
    int rc = alloc_phase_buffers<double>(ctx, n, N, W, S64, A64);
    if (rc != KFC_OK) return rc;
    const int pmode = precond == KFC_PRECOND_RITZ ? 1 : precond == KFC_PRECOND_JACOBI ? 2 : 0;
    DevBuf dD;
    if (pmode) KFC_HIP(ctx, dD.ensure(sizeof(double) * size_t(n) * N * 2));
    {
        std::vector<int> rcs(n, KFC_OK), npert(n, 0);
        std::vector<std::vector<cd>> X0hs(n);
        std::vector<std::vector<double>> Xh(n);
        std::vector<char> devseed(n, 0);
        Team team0(std::max(1, std::min(16, int(std::thread::hardware_concurrency()))));
        // warm starts checked and host-made blocks (random, or a full warm start) formed on
        // host threads; product seeds are formed on the device below
        team0.run(sub ? 0 : n, [&](int i) {
            kfc_op* op = ops[i];
            const bool warm = X0s && X0s[i];
            if (warm) {
                rcs[i] = lobpcg_check_x0(op, X0s[i], nx0, X0hs[i]);
                if (rcs[i] != KFC_OK) return;
            }
            devseed[i] = seed_kind == KFC_SEED_PRODUCT && !(warm && nx0 == W);
            if (devseed[i]) return;
            std::vector<cd> X;
            lobpcg_seed_block(op, int(m), W, X0hs[i], warm ? nx0 : 0, std::vector<cd>(),
                              seed_kind == KFC_SEED_RANDOM && !warm, seed + i, X, npert[i]);
            Xh[i].resize(X.size());
            for (size_t k = 0; k < X.size(); ++k) Xh[i][k] = X[k].real();
        });
        for (int i = 0; i < n; ++i)
            if (rcs[i] != KFC_OK) return kfc_fail(ctx, rcs[i], "instance " + std::to_string(i) + ": " + ops[i]->ctx->err);
        for (int i = 0; i < n; ++i) {
            kfc_op* op = ops[i];
            pts[i].op = op;
            double* Xd = S64[0].as<double>() + size_t(i) * 3 * W * N;
            if (sub) {
                npert[i] = W;   // formed below, perturbed as a warm start with its guards
            } else if (devseed[i]) {
                rc = lobpcg_seed_dev<double>(op, int(m), W, X0hs[i], (X0s && X0s[i]) ? nx0 : 0, true, false, seed + i, Xd,
                                             npert[i]);
                if (rc != KFC_OK) return kfc_fail(ctx, rc, "instance " + std::to_string(i) + ": " + op->ctx->err);
            } else {
                KFC_HIP(ctx, hipMemcpy(Xd, Xh[i].data(), sizeof(double) * N * W, hipMemcpyHostToDevice));
            }
            if (pmode) {
                // D of every point first, then the FixedJacobi inverses 1 / (D - minimum(D) + 1)
                double* Di = dD.as<double>() + size_t(i) * N;
                double dmin = INFINITY;
                rc = op_full_diag_dev(op, Di, &dmin, nullptr);
                if (rc == KFC_OK) rc = op_shift_inverse_dev(op, Di, dmin - 1, dD.as<double>() + size_t(n + i) * N);
                if (rc != KFC_OK) return kfc_fail(ctx, rc, "instance " + std::to_string(i) + ": " + op->ctx->err);
            }
        }
        if (sub) {
            const auto ts = clk::now();
            rc = subspace_starts(ops, n, W, *sub, S64[0].as<double>(), team0);
            if (rc != KFC_OK) return rc;
            if (trace) std::fprintf(stderr, "[lockstep] subspace starts (%d directions): %.3f s\n", sub->ns, secs(ts));
        }
        // the seeded columns of every point perturbed on the device, one launch
        DevBuf dnp;
        KFC_HIP(ctx, dnp.ensure(sizeof(int) * n));
        KFC_HIP(ctx, hipMemcpy(dnp.p, npert.data(), sizeof(int) * n, hipMemcpyHostToDevice));
        Work wp;
        wp.ctx = ctx;
        rc = blas_perturb<double>(wp, S64[0].as<double>(), N, W, N, n, int64_t(3) * W * N, dnp.as<int>(), 0x4b46);
        if (rc != KFC_OK) return rc;
        KFC_HIP(ctx, hipStreamSynchronize(st));
    }
    bool pair_batch = true;
    for (int i = 0; i < n; ++i) {
        for (int prec = 0; prec < 2; ++prec) {
            if (prec == 1 && mode != KFC_MODE_MIXED) continue;
            rc = op_ensure_device(ops[i], prec);
            if (rc != KFC_OK) return rc;
        }
        pair_batch &= !ops[i]->plan.ok && ops[i]->pairs.ok && pairs_same_structure(op0, ops[i]);
    }
    // the split pair products' scratch for every point's W columns, before any step is
    // captured (a table fill allocates nothing)
    if (pair_batch) {
        rc = pairs_reserve(op0, n, W, 0);
        if (rc != KFC_OK) return rc;
    }
    const int64_t rows_per = std::max<int64_t>(LS_SR, ((N + 15) / 16 + LS_SR - 1) / LS_SR * LS_SR);
    const int nchunk = int((N + rows_per - 1) / rows_per);
    DevBuf dpart, dred, dslot, dtheta, dpth, dM, dtab;
    PinBuf hred, hslot, htheta, hM, htab;
    KFC_HIP(ctx, dpart.ensure(sizeof(double) * size_t(n) * nchunk * LS_E));
    KFC_HIP(ctx, dred.ensure(sizeof(double) * size_t(n) * LS_E));
    KFC_HIP(ctx, hred.ensure(sizeof(double) * size_t(n) * LS_E));
    KFC_HIP(ctx, dslot.ensure(sizeof(LsSlot) * 2 * n));
    KFC_HIP(ctx, hslot.ensure(sizeof(LsSlot) * 2 * n));
    KFC_HIP(ctx, dtheta.ensure(sizeof(double) * n * W));
    KFC_HIP(ctx, htheta.ensure(sizeof(double) * n * W));
    KFC_HIP(ctx, dpth.ensure(sizeof(double) * n * W));
    KFC_HIP(ctx, dM.ensure(sizeof(double) * n * LS_CMAX * LS_MO));
    KFC_HIP(ctx, hM.ensure(sizeof(double) * n * LS_CMAX * LS_MO));
    const size_t tabb = std::max(pair_args_bytes(0), pair_args_bytes(1)) * n * std::max<size_t>(1, op0->pairs.groups.size());
    KFC_HIP(ctx, dtab.ensure(tabb));
    KFC_HIP(ctx, htab.ensure(tabb));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    KFC_HIP(ctx, hipEventCreate(&e0));
    KFC_HIP(ctx, hipEventCreate(&e1));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } evg{e0, e1};
    const int nthreads = std::max(1, std::min(16, int(std::thread::hardware_concurrency())));
    Team team(nthreads);
    if (trace)
        std::fprintf(stderr, "[lockstep] %d points, N %lld, w %d, setup %.3f s, %d host threads, pair batch %d\n", n,
                     (long long)N, W, secs(t0), nthreads, int(pair_batch));
    LsCfg cfg{int(m), W, tol, maxiter, precond, eta};
    double apply_gpu = 0;
    bool untimed = false;   // applies ran inside step graphs: no per-apply device time
    int64_t iters32 = 0, iters64 = 0;
    std::vector<int64_t> it32(n, 0), mv32(n, 0);
    const double* Dall = pmode ? dD.as<double>() : nullptr;
    if (mode == KFC_MODE_MIXED) {
        rc = alloc_phase_buffers<float>(ctx, n, N, W, S32, A32);
        if (rc != KFC_OK) return rc;
        // the start blocks in FP32
        for (int i = 0; i < n; ++i) {
            const size_t off = size_t(i) * 3 * W * N;
            hipLaunchKernelGGL((k_ls_convert<double, float>), dim3(256), dim3(256), 0, st, S64[0].as<double>() + off,
                               S32[0].as<float>() + off, int64_t(W) * N);
        }
        KFC_HIP(ctx, hipGetLastError());
        double hb = 0, lo = 0;
        rc = op_spectral_interval(op0, &lo, &hb);
        if (rc != KFC_OK) return rc;
        double hmax = std::max(std::fabs(hb), std::fabs(lo));
        for (int i = 1; i < n; ++i) {
            double l2 = 0, h2 = 0;
            if (op_spectral_interval(ops[i], &l2, &h2) == KFC_OK) hmax = std::max(hmax, std::max(std::fabs(h2), std::fabs(l2)));
        }
        const double sw = std::max(tol, 256.0 * 5.96e-8 * std::max(hmax, 1.0));
        Phase<float> ph{ctx, st, pts, cfg, N, nchunk, rows_per, pair_batch, {S32[0].as<float>(), S32[1].as<float>()},
                        {A32[0].as<float>(), A32[1].as<float>()}, dpart, dred, dslot, dtheta, dpth, dM, dtab,
                        hred, hslot, htheta, hM, htab, Dall, team, sw, &apply_gpu, e0, e1};
        rc = ph.run(&iters32);
        if (rc != KFC_OK) return rc;
        untimed |= ph.apply_untimed;
        if (trace)
            std::fprintf(stderr, "[lockstep] fp32 phase: %lld steps (%lld from %lld graphs), %lld point-steps, %lld collapsed, wait %.3f s, rayleigh-ritz %.3f s, first rayleigh-ritz %.3f s, launched steps %.3f s, captures %.3f s, %.3f s\n",
                         (long long)ph.n_iter, (long long)ph.n_graph_steps, (long long)ph.n_captures,
                         (long long)ph.n_pointiter, (long long)ph.n_collapsed, ph.t_wait, ph.t_rr, ph.t_init, ph.t_eager, ph.t_capture, secs(t0));
        // FP64 continuation from each point's FP32 X
        for (int i = 0; i < n; ++i) {
            it32[i] = pts[i].it;
            mv32[i] = pts[i].matvec;
            if (pts[i].failed) continue;
            const size_t off = size_t(i) * 3 * W * N;
            hipLaunchKernelGGL((k_ls_convert<float, double>), dim3(256), dim3(256), 0, st,
                               S32[pts[i].par].as<float>() + off, S64[0].as<double>() + off, int64_t(W) * N);
            pts[i].done = false;
            pts[i].par = 0;
        }
        KFC_HIP(ctx, hipGetLastError());
        for (int b = 0; b < 2; ++b) {
            S32[b].release();
            A32[b].release();
        }
    }
    {
        Phase<double> ph{ctx, st, pts, cfg, N, nchunk, rows_per, pair_batch, {S64[0].as<double>(), S64[1].as<double>()},
                         {A64[0].as<double>(), A64[1].as<double>()}, dpart, dred, dslot, dtheta, dpth, dM, dtab,
                         hred, hslot, htheta, hM, htab, Dall, team, 0.0, &apply_gpu, e0, e1};
        rc = ph.run(&iters64);
        if (rc != KFC_OK) return rc;
        untimed |= ph.apply_untimed;
        if (trace)
            std::fprintf(stderr, "[lockstep] fp64 phase: %lld steps (%lld from %lld graphs), %lld point-steps, %lld collapsed, wait %.3f s, rayleigh-ritz %.3f s, first rayleigh-ritz %.3f s, launched steps %.3f s, captures %.3f s, %.3f s\n",
                         (long long)ph.n_iter, (long long)ph.n_graph_steps, (long long)ph.n_captures,
                         (long long)ph.n_pointiter, (long long)ph.n_collapsed, ph.t_wait, ph.t_rr, ph.t_init, ph.t_eager, ph.t_capture, secs(t0));
        // fresh FP64 residuals of the returned m pairs: A[par ^ 1] X slot = H X
        std::vector<int> fin;
        for (int i = 0; i < n; ++i)
            if (!pts[i].failed) fin.push_back(i);
        if (!fin.empty()) {
            rc = ph.apply(fin, 0, int(m), 1);
            if (rc != KFC_OK) return rc;
            std::vector<LsSlot> sl;
            for (int p : fin) sl.push_back(LsSlot{p, pts[p].par, W, 0});
            rc = ph.upload_theta();
            if (rc == KFC_OK) rc = ph.upload_slots(sl);
            if (rc != KFC_OK) return rc;
            ph.shape_for(int(fin.size()));
            hipLaunchKernelGGL(k_ls_resid<double>, dim3(ph.nchunk, int(fin.size())), dim3(256), 0, st,
                               ph.args(int(fin.size())), int(m), dtheta.as<double>(), dtheta.as<double>(),
                               (const double*)nullptr, 0, 0.0, 1);
            KFC_HIP(ctx, hipGetLastError());
            rc = ph.reduce_and_fetch(int(fin.size()), LS_NORM + int(m));
            if (rc != KFC_OK) return rc;
            rc = ph.read_apply_time();
            if (rc != KFC_OK) return rc;
            const double* rec = hred.as<double>();
            std::vector<double> fresh(n, 0.0);
            for (size_t q = 0; q < fin.size(); ++q) {
                double w = 0;
                for (int j = 0; j < m; ++j) w = std::max(w, std::sqrt(std::max(0.0, rec[q * LS_E + LS_NORM + j])));
                fresh[fin[q]] = w;
            }
            // results
            for (int p : fin) {
                LsPoint& pt = pts[p];
                int bad = 0;
                double worst = 0;
                for (int j = 0; j < m; ++j) {
                    if (!(pt.res[j] <= tol)) ++bad;
                    worst = std::max(worst, pt.res[j]);
                }
                if (bad > 0)
                    return kfc_fail(ctx, KFC_ENOCONV,
                                    "instance " + std::to_string(p) + ": LOBPCG: " + std::to_string(bad) + " of " +
                                        std::to_string(m) + " eigenpairs did not converge within " +
                                        std::to_string(maxiter) + " iterations (only " + std::to_string(m - bad) +
                                        " converged; maximum residual " + std::to_string(worst) + " > tol = " +
                                        std::to_string(tol) + ")");
                for (int j = 0; j < m; ++j) E[size_t(p) * m + j] = pt.theta[j];
                KFC_HIP(ctx, hipMemcpyAsync(Vs[p], S64[pt.par].as<double>() + size_t(p) * 3 * W * N,
                                            sizeof(double) * N * m, hipMemcpyDeviceToHost, st));
                if (stats) {
                    kfc_stats& s = stats[p];
                    s = kfc_stats{};
                    s.iterations = pt.it + it32[p];
                    s.n_matvec = pt.matvec;
                    s.n_matvec_f32 = mv32[p];
                    s.n_apply_calls = 0;
                    s.max_residual = fresh[p];
                    s.n_converged = int(m);
                    s.warm = ((X0s && X0s[p]) || sub) ? 1 : 0;
                }
            }
            KFC_HIP(ctx, hipStreamSynchronize(st));
        }
    }
    // points whose lockstep basis collapsed: the one-point path, with its retry
    for (int i = 0; i < n; ++i) {
        if (!pts[i].failed) continue;
        rc = lobpcg_solve(ops[i], m, guard, tol, maxiter, precond, eta, seed_kind, X0s ? X0s[i] : nullptr, seed + i,
                          mode, E + size_t(i) * m, Vs[i], stats ? stats + i : nullptr, false);
        if (rc != KFC_OK) return kfc_fail(ctx, rc, "instance " + std::to_string(i) + ": " + ops[i]->ctx->err);
    }
    if (stats) {
        const double tot = secs(t0);
        for (int i = 0; i < n; ++i) {
            if (pts[i].failed) continue;
            stats[i].seconds_total = tot / n;
            stats[i].seconds_apply_gpu = untimed ? -1.0 : apply_gpu / n;
            stats[i].seconds_apply = untimed ? -1.0 : apply_gpu / n;
        }
    }
    (void)iters32;
    (void)iters64;
    return KFC_OK;
}

}  // namespace

// Lockstep LOBPCG of a sweep: n real operators of one size on one context.
extern "C" int kfc_lobpcg_lockstep(int32_t n, kfc_op* const* ops, int64_t m, int64_t guard, double tol,
                                   int64_t maxiter, int32_t precond, double eta, int32_t seed_kind,
                                   const double* const* X0s, uint64_t seed, int32_t mode, double* E,
                                   double* const* Vs, kfc_stats* stats) {
    if (n < 1 || !ops || !ops[0] || !E || !Vs)
        return kfc_fail(n > 0 && ops && ops[0] ? ops[0]->ctx : nullptr, KFC_EARG, "kfc_lobpcg_lockstep: bad arguments");
    kfc_op* op0 = ops[0];
    const int64_t N = op0->N;
    for (int i = 0; i < n; ++i) {
        if (!ops[i] || !Vs[i]) return kfc_fail(op0->ctx, KFC_EARG, "kfc_lobpcg_lockstep: null operator or output");
        if (ops[i]->ctx != op0->ctx)
            return kfc_fail(op0->ctx, KFC_EARG, "kfc_lobpcg_lockstep: every operator must live on one context");
        if (ops[i]->N != N) return kfc_fail(op0->ctx, KFC_EARG, "kfc_lobpcg_lockstep: every operator must have one size");
        if (ops[i]->cplx) return kfc_fail(op0->ctx, KFC_EUNSUPPORTED, "kfc_lobpcg_lockstep: complex operators are not served");
    }
    if (m < 1) return kfc_fail(op0->ctx, KFC_EARG, "LOBPCG requires m >= 1, got " + std::to_string(m));
    if (N < 3 * m + 5)
        return kfc_fail(op0->ctx, KFC_EARG, "LOBPCG needs N >= 3m + 5, got N = " + std::to_string(N) + " for m = " +
                                                std::to_string(m) + "; the block eigensolver is unreliable on so small a problem, use Dense()");
    if (!(tol > 0) || !std::isfinite(tol)) return kfc_fail(op0->ctx, KFC_EARG, "LOBPCG: tol must be positive and finite");
    if (maxiter < 1) return kfc_fail(op0->ctx, KFC_EARG, "LOBPCG: maxiter must be at least 1");
    if (precond == KFC_PRECOND_RITZ && !(eta > 0 && std::isfinite(eta)))
        return kfc_fail(op0->ctx, KFC_EARG, "eta must be positive and finite");
    if (guard < 0) return kfc_fail(op0->ctx, KFC_EARG, "LOBPCG: guard must be nonnegative");
    if (seed_kind == KFC_SEED_WARM_GUARDS && !X0s)
        return kfc_fail(op0->ctx, KFC_EARG, "LOBPCG: the warm-guards seed needs X0 blocks");
    const int g = int(std::max<int64_t>(0, std::min<int64_t>(guard, (N - 5) / 3 - m)));
    if (3 * (m + g) > LS_CMAX)
        return kfc_fail(op0->ctx, KFC_EUNSUPPORTED, "kfc_lobpcg_lockstep: m + guard above 21 is not served");
    (void)hipSetDevice(op0->ctx->device);
    return lockstep_run(n, ops, m, guard, tol, maxiter, precond, eta, seed_kind, X0s, seed, mode, E, Vs, stats);
}

// Lockstep LOBPCG of a sweep started from a learned subspace (no counterpart in the
// reference: it supplies the X0 of src/lobpcg.jl:217 for every sweep point).
extern "C" int kfc_lobpcg_lockstep_subspace(int32_t n, kfc_op* const* ops, int64_t m, int64_t guard, double tol,
                                            int64_t maxiter, int32_t precond, double eta, const double* B, int64_t nf,
                                            int32_t nb, const int64_t* const* rowmaps, const double* const* C,
                                            int32_t ns, uint64_t seed, int32_t mode, double* E, double* const* Vs,
                                            kfc_stats* stats) {
    if (n < 1 || !ops || !ops[0])
        return kfc_fail(nullptr, KFC_EARG, "kfc_lobpcg_lockstep_subspace: bad arguments");
    kfc_op* op0 = ops[0];
    if (!B || nf < 1 || nb < 1 || !rowmaps)
        return kfc_fail(op0->ctx, KFC_EARG, "kfc_lobpcg_lockstep_subspace: the subspace needs B, nf >= 1, nb >= 1 and rowmaps");
    if (!C) ns = nb;
    if (ns < 1 || ns > nb)
        return kfc_fail(op0->ctx, KFC_EARG, "kfc_lobpcg_lockstep_subspace: ns must lie in 1:nb, got " + std::to_string(ns));
    const int64_t N = op0->N;
    const int g = int(std::max<int64_t>(0, std::min<int64_t>(guard, (N - 5) / 3 - m)));
    if (ns < m + g)
        return kfc_fail(op0->ctx, KFC_EARG, "kfc_lobpcg_lockstep_subspace: the subspace has " + std::to_string(ns) +
                                                " directions for a block of " + std::to_string(m + g));
    for (int i = 0; i < n; ++i) {
        if (!rowmaps[i]) return kfc_fail(op0->ctx, KFC_EARG, "kfc_lobpcg_lockstep_subspace: null rowmap");
        std::vector<char> hit;
        for (int64_t f = 0; f < nf; ++f) {
            const int64_t r = rowmaps[i][f];
            if (r < -1 || r >= N)
                return kfc_fail(op0->ctx, KFC_EARG, "kfc_lobpcg_lockstep_subspace: rowmap entry " + std::to_string(r) +
                                                        " outside -1:" + std::to_string(N - 1));
            if (r >= 0) {
                if (hit.empty()) hit.assign(size_t(N), 0);
                if (hit[size_t(r)]) return kfc_fail(op0->ctx, KFC_EARG, "kfc_lobpcg_lockstep_subspace: two frame rows map to row " + std::to_string(r));
                hit[size_t(r)] = 1;
            }
        }
    }
    SubspaceSeed sub;
    sub.B = B;
    sub.nf = nf;
    sub.nb = nb;
    sub.ns = ns;
    sub.rowmaps = rowmaps;
    sub.C = C;
    // the checks of kfc_lobpcg_lockstep, then the run with the subspace starts
    if (!E || !Vs) return kfc_fail(op0->ctx, KFC_EARG, "kfc_lobpcg_lockstep_subspace: bad arguments");
    for (int i = 0; i < n; ++i) {
        if (!ops[i] || !Vs[i]) return kfc_fail(op0->ctx, KFC_EARG, "kfc_lobpcg_lockstep_subspace: null operator or output");
        if (ops[i]->ctx != op0->ctx)
            return kfc_fail(op0->ctx, KFC_EARG, "kfc_lobpcg_lockstep_subspace: every operator must live on one context");
        if (ops[i]->N != N) return kfc_fail(op0->ctx, KFC_EARG, "kfc_lobpcg_lockstep_subspace: every operator must have one size");
        if (ops[i]->cplx) return kfc_fail(op0->ctx, KFC_EUNSUPPORTED, "kfc_lobpcg_lockstep_subspace: complex operators are not served");
    }
    if (m < 1) return kfc_fail(op0->ctx, KFC_EARG, "LOBPCG requires m >= 1, got " + std::to_string(m));
    if (N < 3 * m + 5)
        return kfc_fail(op0->ctx, KFC_EARG, "LOBPCG needs N >= 3m + 5, got N = " + std::to_string(N) + " for m = " +
                                                std::to_string(m) + "; the block eigensolver is unreliable on so small a problem, use Dense()");
    if (!(tol > 0) || !std::isfinite(tol)) return kfc_fail(op0->ctx, KFC_EARG, "LOBPCG: tol must be positive and finite");
    if (maxiter < 1) return kfc_fail(op0->ctx, KFC_EARG, "LOBPCG: maxiter must be at least 1");
    if (precond == KFC_PRECOND_RITZ && !(eta > 0 && std::isfinite(eta)))
        return kfc_fail(op0->ctx, KFC_EARG, "eta must be positive and finite");
    if (guard < 0) return kfc_fail(op0->ctx, KFC_EARG, "LOBPCG: guard must be nonnegative");
    if (3 * (m + g) > LS_CMAX)
        return kfc_fail(op0->ctx, KFC_EUNSUPPORTED, "kfc_lobpcg_lockstep: m + guard above 21 is not served");
    (void)hipSetDevice(op0->ctx->device);
    return lockstep_run(n, ops, m, guard, tol, maxiter, precond, eta, KFC_SEED_PRODUCT, nullptr, seed, mode, E, Vs,
                        stats, &sub);
}
