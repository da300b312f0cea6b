// Batched Chebyshev-filtered subspace iteration over many one-axis operators.
//
// A parameter sweep of small nodes (the leaves of the coupled-transmon sweep:
// periodic phase grids of 256-768 points) is latency-bound when each node is
// solved on its own: a few hundred tiny launches and host round trips per
// solve.  Here every problem of the batch advances in lockstep: one launch
// applies the filter step to every active problem's block, one launch forms
// every problem's Gram matrices, one launch rotates every block, and the small
// per-problem eigenproblems are solved on host threads.  The iteration is the
// Chebyshev-filtered subspace iteration of src/chebyshev.jl:310-406 (scaled
// three-term filter of Zhou and Saad, SVQB orthonormalization, Rayleigh-Ritz,
// convergence of the lowest m on fresh residuals), with a fixed filter degree
// shared by the batch so that the steps batch; FP32 filter steps until the
// residuals reach the FP32 floor, then FP64 to the tolerance.
//
// Each operator is turned into one modular band: y_i = sum_b W_b[i] x_{(i+b-h) mod N}.
// Banded, shift and diagonal terms with any offsets map onto it exactly (a band
// entry that a term leaves out, an invalid row of a banded factor, is zero).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <mutex>
#include <random>
#include <thread>

#include "kfc_internal.h"

namespace {

using clk = std::chrono::steady_clock;

template <class S> struct Dbl { using type = double; };
template <class R> struct Dbl<Cx<R>> { using type = Cx<double>; };

__device__ inline double to_d(float x) { return double(x); }
__device__ inline double to_d(double x) { return x; }
__device__ inline Cx<double> to_d(Cx<float> x) { return Cx<double>{double(x.re), double(x.im)}; }
__device__ inline Cx<double> to_d(Cx<double> x) { return x; }
template <class S> __device__ inline S from_d(typename Dbl<S>::type a);
template <> __device__ inline float from_d<float>(double a) { return float(a); }
template <> __device__ inline double from_d<double>(double a) { return a; }
template <> __device__ inline Cx<float> from_d<Cx<float>>(Cx<double> a) { return Cx<float>{float(a.re), float(a.im)}; }
template <> __device__ inline Cx<double> from_d<Cx<double>>(Cx<double> a) { return a; }

template <class S>
struct BatchDev {
    int nb = 0, B = 0, h = 0;
    const int* act = nullptr;       // active problem indices
    const int64_t* N = nullptr;     // per problem
    const int64_t* xo = nullptr;    // block offset (elements), column-major with ld N_p
    const int64_t* wo = nullptr;    // band-weight offset (elements), B x N_p
    const S* W = nullptr;
};

// Y = a (H X) + b X + g Z per problem (a, b, g from coef[3 p .. 3 p + 2]); Z may be null.
// One block applies one column of one problem (any block size).
template <class S>
__device__ inline void bapply_col(const BatchDev<S>& d, const S* __restrict__ X, const S* __restrict__ Z,
                                  S* __restrict__ Y, const double* __restrict__ coef, int p, int col) {
    using R = typename RealOf<S>::type;
    const int64_t N = d.N[p];
    const S* Xc = X + d.xo[p] + col * N;
    const S* Zc = Z ? Z + d.xo[p] + col * N : nullptr;
    S* Yc = Y + d.xo[p] + col * N;
    const S* Wp = d.W + d.wo[p];
    const R a = R(coef[3 * p]), b = R(coef[3 * p + 1]), g = R(coef[3 * p + 2]);
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
        S acc = s_zero(Xc[0]);
        for (int k = 0; k < d.B; ++k) {
            int64_t j = i + k - d.h;
            j += j < 0 ? N : 0;
            j -= j >= N ? N : 0;
            acc = s_fma(acc, Wp[k * N + i], Xc[j]);
        }
        S y = s_add(s_scale(a, acc), s_scale(b, Xc[i]));
        if (Zc) y = s_add(y, s_scale(g, Zc[i]));
        Yc[i] = y;
    }
}

// Launch shape 0 and 2: a grid of (columns, active problems), 256 or 128 threads.
template <class S>
__global__ __launch_bounds__(256) void k_bapply(const BatchDev<S> d, const S* __restrict__ X, const S* __restrict__ Z,
                                                S* __restrict__ Y, const double* __restrict__ coef) {
    bapply_col<S>(d, X, Z, Y, coef, d.act[blockIdx.y], blockIdx.x);
}

// Launch shape 1: a persistent grid walking the (problem, column) work list in order
// of decreasing problem size (ord: active slots by decreasing N), so that the long
// columns of uneven batches start first and the short ones fill in behind them.
template <class S>
__global__ __launch_bounds__(256) void k_bapply_list(const BatchDev<S> d, const int* __restrict__ ord, int nitems,
                                                     const S* __restrict__ X, const S* __restrict__ Z,
                                                     S* __restrict__ Y, const double* __restrict__ coef) {
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int q = ord[it / d.nb], col = it - (it / d.nb) * d.nb;
        bapply_col<S>(d, X, Z, Y, coef, d.act[q], col);
    }
}

// G_p = X_p^H Y_p (nb x nb, FP64 accumulation), one block per active problem;
// 32-row tiles of both blocks staged in LDS.
template <class S>
__global__ __launch_bounds__(256) void k_bgram(const BatchDev<S> d, const S* __restrict__ X, const S* __restrict__ Y,
                                               typename Dbl<S>::type* __restrict__ G) {
    using A = typename Dbl<S>::type;
    __shared__ A sx[32][33];
    __shared__ A sy[32][33];
    const int q = blockIdx.x, p = d.act[q], nb = d.nb;
    const int64_t N = d.N[p];
    const S* Xp = X + d.xo[p];
    const S* Yp = Y + d.xo[p];
    A acc[4];
    const A zero = to_d(s_zero(Xp[0]));
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] = zero;
    for (int64_t r0 = 0; r0 < N; r0 += 32) {
        __syncthreads();
        for (int e = threadIdx.x; e < 32 * nb; e += 256) {
            const int r = e % 32, c = e / 32;
            const int64_t row = r0 + r;
            sx[c][r] = row < N ? to_d(Xp[row + c * N]) : zero;
            sy[c][r] = row < N ? to_d(Yp[row + c * N]) : zero;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = threadIdx.x + 256 * k;
            if (e < nb * nb) {
                const int i = e % nb, j = e / nb;
                A s = acc[k];
                for (int r = 0; r < 32; ++r) s = s_fma(s, s_conj(sx[i][r]), sy[j][r]);
                acc[k] = s;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int e = threadIdx.x + 256 * k;
        if (e < nb * nb) G[(int64_t)q * nb * nb + e] = acc[k];
    }
}

// Y_p[:, 0:nc] = X_p C_q (C_q nb x nc column-major, FP64), per active problem q.
template <class S>
__global__ __launch_bounds__(256) void k_brot(const BatchDev<S> d, const S* __restrict__ X, S* __restrict__ Y,
                                              const typename Dbl<S>::type* __restrict__ C, int nc) {
    using A = typename Dbl<S>::type;
    __shared__ A sc[32 * 32];
    const int q = blockIdx.y, p = d.act[q], nb = d.nb;
    const int64_t N = d.N[p];
    for (int e = threadIdx.x; e < nb * nc; e += 256) sc[e] = C[(int64_t)q * nb * nc + e];
    __syncthreads();
    const int64_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const S* Xp = X + d.xo[p];
    S* Yp = Y + d.xo[p];
    A x[32];
    for (int c = 0; c < nb; ++c) x[c] = to_d(Xp[i + c * N]);
    for (int j = 0; j < nc; ++j) {
        A s = to_d(s_zero(Xp[0]));
        for (int c = 0; c < nb; ++c) s = s_fma(s, x[c], sc[c + j * nb]);
        Yp[i + j * N] = from_d<S>(s);
    }
}

// r[q * nb + j] = || HX_j - theta_j X_j ||^2 per active problem q
template <class S>
__global__ __launch_bounds__(256) void k_bresid(const BatchDev<S> d, const S* __restrict__ X, const S* __restrict__ HX,
                                                const double* __restrict__ theta, double* __restrict__ r) {
    __shared__ double red[256];
    const int q = blockIdx.y, p = d.act[q], j = blockIdx.x, nb = d.nb;
    const int64_t N = d.N[p];
    const S* Xc = X + d.xo[p] + j * N;
    const S* Hc = HX + d.xo[p] + j * N;
    const double th = theta[q * nb + j];
    double s = 0;
    for (int64_t i = threadIdx.x; i < N; i += 256) {
        const auto v = s_sub(to_d(Hc[i]), s_scale(th, to_d(Xc[i])));
        s += double(s_abs2(v));
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int hh = 128; hh > 0; hh >>= 1) {
        if (threadIdx.x < hh) red[threadIdx.x] += red[threadIdx.x + hh];
        __syncthreads();
    }
    if (threadIdx.x == 0) r[q * nb + j] = red[0];
}

// dst = src on the active problems' blocks only (the other blocks keep their results)
template <class S>
__global__ __launch_bounds__(256) void k_bcopy(const BatchDev<S> d, const S* __restrict__ src, S* __restrict__ dst) {
    const int p = d.act[blockIdx.y];
    const int64_t N = d.N[p], off = d.xo[p] + blockIdx.x * N;
    for (int64_t i = threadIdx.x; i < N; i += 256) dst[off + i] = src[off + i];
}

template <class S, class T>
__global__ void k_bconvert(const S* __restrict__ X, T* __restrict__ Y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        Y[i] = from_d<T>(to_d(X[i]));
}

// ------------------------------------------------------------------------------------------
// Host side

struct Band1D {
    int64_t N = 0;
    int h = 0;
    std::vector<cd> W;   // (2h + 1) x N, band k = offset k - h
    double hi = 0;       // Gershgorin upper bound
    double lo = 0;       // Gershgorin lower bound
};

int band_of(kfc_op* op, Band1D& out) {
    if (op->ndims != 1) return kfc_fail(op->ctx, KFC_EUNSUPPORTED, "kfc_chebyshev_batch: operators must have one axis");
    const int64_t N = op->N;
    int rc = op_host_diag(op);
    if (rc != KFC_OK) return rc;
    std::map<int64_t, std::vector<cd>> bands;
    auto add = [&](int64_t o, int64_t row, cd v) {
        int64_t mo = ((o % N) + N) % N;
        if (mo > N / 2) mo -= N;
        auto& vec = bands[mo];
        if (vec.empty()) vec.assign(N, cd(0));
        vec[row] += v;
    };
    for (int64_t i = 0; i < N; ++i) add(0, i, op->hdiag[i]);
    for (const Term& t : op->terms) {
        if (t.factors.size() != 1)
            return kfc_fail(op->ctx, KFC_EUNSUPPORTED, "kfc_chebyshev_batch: a kept term has more than one factor");
        const Factor& f = t.factors[0];
        if (f.kind == KFC_BANDED) {
            for (size_t b = 0; b < f.offsets.size(); ++b)
                for (int64_t i = 0; i < N; ++i) {
                    const int64_t j = i + f.offsets[b];
                    if (j >= 0 && j < N) add(f.offsets[b], i, t.coeff * f.weights[i + b * N]);
                }
        } else if (f.kind == KFC_SHIFT) {   // y[i + s] += factor x[i]
            for (int64_t i = 0; i < N; ++i) {
                const int64_t r = i + f.shift;
                if (r >= 0 && r < N) add(-f.shift, r, t.coeff * f.factor);
            }
        } else if (f.kind == KFC_DIAG) {
            for (int64_t i = 0; i < N; ++i) add(0, i, t.coeff * f.diag[i]);
        } else {
            return kfc_fail(op->ctx, KFC_EUNSUPPORTED, "kfc_chebyshev_batch: dense factors are not served");
        }
    }
    int h = 0;
    for (auto& kv : bands) h = std::max<int>(h, int(std::abs(kv.first)));
    out.N = N;
    out.h = h;
    out.W.assign(size_t(2 * h + 1) * N, cd(0));
    for (auto& kv : bands)
        for (int64_t i = 0; i < N; ++i) out.W[size_t(kv.first + h) * N + i] = kv.second[i];
    double hi = -INFINITY, lo = INFINITY;
    for (int64_t i = 0; i < N; ++i) {
        double off = 0;
        for (int k = 0; k <= 2 * h; ++k)
            if (k != h) off += std::abs(out.W[size_t(k) * N + i]);
        const double dd = out.W[size_t(h) * N + i].real();
        hi = std::max(hi, dd + off);
        lo = std::min(lo, dd - off);
    }
    out.hi = hi;
    out.lo = lo;
    return KFC_OK;
}

// run fn(i) for i in [0, n) on a few host threads
template <class F>
void parallel_for(int n, F fn) {
    const int T = std::max(1, std::min<int>(n, int(std::thread::hardware_concurrency()) > 0
                                                     ? std::min(16, int(std::thread::hardware_concurrency()))
                                                     : 4));
    if (T <= 1 || n < 8) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            for (int i = t; i < n; i += T) fn(i);
        });
    for (auto& x : th) x.join();
}

template <class S>
struct Batch {
    kfc_ctx* ctx;
    int n, nb, B, h;
    std::vector<int64_t> N, xo, wo;
    int64_t total = 0;
    DevBuf dN, dxo, dwo, dW, dact, bX, bY, bZ, bH, dG, dC, dcoef, dth, dr;
    BatchDev<S> dev;
    std::vector<int> act;
    // The filter of one iteration: every step's coefficients in one table (pinned host
    // copy, device copy), and its launches captured as a graph that is replayed while
    // the active set, the degree and the table stay the same.
    double* hcoefs = nullptr;
    size_t ncoefs = 0;
    DevBuf dcoefs;
    hipGraphExec_t fexec = nullptr;
    std::vector<int> fact;
    int fdeg = 0;
    const void* fcoefs = nullptr;
    // the filter step's launch shape (the tuner's pick for this active set's sizes) and
    // the active slots by decreasing size for the persistent shape
    int shape = 0, grid = 0;
    std::vector<int64_t> shape_key;
    DevBuf dord;
    Batch() = default;
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
    ~Batch() {
        if (fexec) (void)hipGraphExecDestroy(fexec);
        if (hcoefs) (void)hipHostFree(hcoefs);
    }
};

// KFC_GRAPHS=0 runs the filter steps as separate launches instead of one graph.
bool graphs_enabled() {
    const char* e = getenv("KFC_GRAPHS");
    return !(e && e[0] == '0');
}

template <class S>
int batch_init(Batch<S>& b, kfc_ctx* ctx, const std::vector<Band1D>& bands, int nb) {
    b.ctx = ctx;
    b.n = int(bands.size());
    b.nb = nb;
    b.h = 0;
    for (const auto& x : bands) b.h = std::max(b.h, x.h);
    b.B = 2 * b.h + 1;
    b.N.resize(b.n);
    b.xo.resize(b.n);
    b.wo.resize(b.n);
    int64_t w = 0;
    b.total = 0;
    for (int p = 0; p < b.n; ++p) {
        b.N[p] = bands[p].N;
        b.xo[p] = b.total;
        b.wo[p] = w;
        b.total += bands[p].N * nb;
        w += bands[p].N * b.B;
    }
    std::vector<S> Wh(static_cast<size_t>(w));
    for (int p = 0; p < b.n; ++p) {
        const auto& x = bands[p];
        for (int k = 0; k < b.B; ++k) {
            const int kk = k - b.h + x.h;   // this problem's band index
            for (int64_t i = 0; i < x.N; ++i)
                Wh[b.wo[p] + size_t(k) * x.N + i] =
                    s_from<S>((kk >= 0 && kk <= 2 * x.h) ? x.W[size_t(kk) * x.N + i] : cd(0));
        }
    }
    KFC_HIP(ctx, b.dN.ensure(sizeof(int64_t) * b.n));
    KFC_HIP(ctx, b.dxo.ensure(sizeof(int64_t) * b.n));
    KFC_HIP(ctx, b.dwo.ensure(sizeof(int64_t) * b.n));
    KFC_HIP(ctx, b.dW.ensure(sizeof(S) * std::max<size_t>(1, Wh.size())));
    KFC_HIP(ctx, b.dact.ensure(sizeof(int) * b.n));
    for (DevBuf* x : {&b.bX, &b.bY, &b.bZ, &b.bH}) KFC_HIP(ctx, x->ensure(sizeof(S) * size_t(b.total)));
    using A = typename Dbl<S>::type;
    KFC_HIP(ctx, b.dG.ensure(sizeof(A) * size_t(b.n) * nb * nb));
    KFC_HIP(ctx, b.dC.ensure(sizeof(A) * size_t(b.n) * nb * nb));
    KFC_HIP(ctx, b.dcoef.ensure(sizeof(double) * 3 * size_t(b.n)));
    KFC_HIP(ctx, b.dth.ensure(sizeof(double) * size_t(b.n) * nb));
    KFC_HIP(ctx, b.dr.ensure(sizeof(double) * size_t(b.n) * nb));
    hipStream_t st = ctx->stream;
    KFC_HIP(ctx, hipMemcpyAsync(b.dN.p, b.N.data(), sizeof(int64_t) * b.n, hipMemcpyHostToDevice, st));
    KFC_HIP(ctx, hipMemcpyAsync(b.dxo.p, b.xo.data(), sizeof(int64_t) * b.n, hipMemcpyHostToDevice, st));
    KFC_HIP(ctx, hipMemcpyAsync(b.dwo.p, b.wo.data(), sizeof(int64_t) * b.n, hipMemcpyHostToDevice, st));
    if (!Wh.empty()) KFC_HIP(ctx, hipMemcpyAsync(b.dW.p, Wh.data(), sizeof(S) * Wh.size(), hipMemcpyHostToDevice, st));
    KFC_HIP(ctx, hipStreamSynchronize(st));
    b.dev.nb = nb;
    b.dev.B = b.B;
    b.dev.h = b.h;
    b.dev.N = b.dN.template as<int64_t>();
    b.dev.xo = b.dxo.template as<int64_t>();
    b.dev.wo = b.dwo.template as<int64_t>();
    b.dev.W = b.dW.template as<S>();
    b.dev.act = b.dact.template as<int>();
    return KFC_OK;
}

template <class S>
int set_active(Batch<S>& b, const std::vector<int>& act) {
    b.act = act;
    // active slots in order of decreasing problem size (stable), for the persistent shape
    std::vector<int> ord(act.size());
    for (size_t q = 0; q < act.size(); ++q) ord[q] = int(q);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return b.N[act[x]] > b.N[act[y]]; });
    KFC_HIP(b.ctx, b.dord.ensure(sizeof(int) * std::max<size_t>(1, ord.size())));
    if (!act.empty()) {
        KFC_HIP(b.ctx, hipMemcpyAsync(b.dact.p, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, b.ctx->stream));
        KFC_HIP(b.ctx, hipMemcpyAsync(b.dord.p, ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice, b.ctx->stream));
    }
    KFC_HIP(b.ctx, hipStreamSynchronize(b.ctx->stream));
    return KFC_OK;
}

// One filter-step launch over the active set in the batch's launch shape: 0 = a
// block of 256 threads per (column, problem), 2 = the same with 128 threads, 1 = a
// persistent grid of b.grid blocks over the size-ordered work list.
template <class S>
void launch_bapply(Batch<S>& b, const S* X, const S* Z, S* Y, const double* coef) {
    const unsigned na = unsigned(b.act.size());
    hipStream_t st = b.ctx->stream;
    if (b.shape == 1) {
        const int nitems = int(na) * b.nb;
        hipLaunchKernelGGL(k_bapply_list<S>, dim3(unsigned(std::max(1, std::min(b.grid, nitems)))), dim3(256), 0, st,
                           b.dev, b.dord.template as<int>(), nitems, X, Z, Y, coef);
    } else {
        hipLaunchKernelGGL(k_bapply<S>, dim3(b.nb, na), dim3(b.shape == 2 ? 128 : 256), 0, st, b.dev, X, Z, Y, coef);
    }
}

// The tuner of the filter step's launch shape: on the first step of each active set
// of a new size signature (sorted sizes, block width, bands, scalar size) every shape
// is timed with HIP events over a few launches of that step and the fastest is kept
// for every later batch with the same signature.  Every shape computes each output
// entry with the same operations in the same order: the choice changes the speed only.
std::mutex g_bshape_mu;
std::map<std::vector<int64_t>, std::pair<int, int>> g_bshape;

template <class S>
int tune_bapply(Batch<S>& b, const S* X, S* Y, const double* coef) {
    if (b.act.empty()) return KFC_OK;
    std::vector<int64_t> key;
    for (int p : b.act) key.push_back(b.N[p]);
    std::sort(key.begin(), key.end());
    key.push_back(b.nb);
    key.push_back(b.B);
    key.push_back(int64_t(sizeof(S)));
    if (key == b.shape_key) return KFC_OK;
    b.shape_key = key;
    if (const char* env = std::getenv("KFC_BATCH_SHAPE")) {   // pinned for tests and measurements
        b.shape = std::atoi(env);
        b.grid = 8 * std::max(1, b.ctx->num_cus);
        return KFC_OK;
    }
    {
        std::lock_guard<std::mutex> g(g_bshape_mu);
        auto it = g_bshape.find(key);
        if (it != g_bshape.end()) {
            b.shape = it->second.first;
            b.grid = it->second.second;
            return KFC_OK;
        }
    }
    const int cus = std::max(1, b.ctx->num_cus);
    const std::pair<int, int> cands[] = {{0, 0}, {2, 0}, {1, 4 * cus}, {1, 8 * cus}, {1, 16 * cus}};
    hipEvent_t e0, e1;
    KFC_HIP(b.ctx, hipEventCreate(&e0));
    KFC_HIP(b.ctx, hipEventCreate(&e1));
    float best = 1e30f;
    std::pair<int, int> pick = cands[0];
    int rc = KFC_OK;
    for (const auto& c : cands) {
        b.shape = c.first;
        b.grid = c.second;
        launch_bapply<S>(b, X, nullptr, Y, coef);   // warm
        if (hipEventRecord(e0, b.ctx->stream) != hipSuccess) { rc = KFC_EDEVICE; break; }
        for (int r = 0; r < 3; ++r) launch_bapply<S>(b, X, nullptr, Y, coef);
        if (hipEventRecord(e1, b.ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = KFC_EDEVICE; break; }
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = KFC_EDEVICE; break; }
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != KFC_OK || hipGetLastError() != hipSuccess) return kfc_fail(b.ctx, KFC_EDEVICE, "batch filter launch tuning failed");
    b.shape = pick.first;
    b.grid = pick.second;
    std::lock_guard<std::mutex> g(g_bshape_mu);
    g_bshape[key] = pick;
    return KFC_OK;
}

template <class S>
int bcopy(Batch<S>& b, const S* src, S* dst) {
    if (b.act.empty()) return KFC_OK;
    hipLaunchKernelGGL(k_bcopy<S>, dim3(b.nb, unsigned(b.act.size())), dim3(256), 0, b.ctx->stream, b.dev, src, dst);
    KFC_HIP(b.ctx, hipGetLastError());
    return KFC_OK;
}

template <class S>
int bapply(Batch<S>& b, const S* X, const S* Z, S* Y, const std::vector<double>& coef) {
    if (b.act.empty()) return KFC_OK;
    KFC_HIP(b.ctx, hipMemcpyAsync(b.dcoef.p, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice,
                                  b.ctx->stream));
    launch_bapply<S>(b, X, Z, Y, b.dcoef.template as<double>());
    KFC_HIP(b.ctx, hipGetLastError());
    return KFC_OK;
}

// The degree steps of the filter on every active block (the three-term recurrence
// through X, Y, Z; step k's coefficients at row k of the device table), the result
// copied to X.
template <class S>
int filter_launches(Batch<S>& b, S* X, S* Y, S* Z, int degree) {
    const double* tab = b.dcoefs.template as<double>();
    const size_t stride = 3 * size_t(b.n);
    const dim3 grid(b.nb, unsigned(b.act.size()));
    launch_bapply<S>(b, X, nullptr, Y, tab);
    S* x0 = X;
    S* y0 = Y;
    S* z0 = Z;
    for (int k = 2; k <= degree; ++k) {   // z = a H y + b y + g x
        launch_bapply<S>(b, y0, x0, z0, tab + size_t(k - 1) * stride);
        S* t = x0;   // (x, y, z) <- (y, z, x)
        x0 = y0;
        y0 = z0;
        z0 = t;
    }
    if (y0 != X) hipLaunchKernelGGL(k_bcopy<S>, grid, dim3(256), 0, b.ctx->stream, b.dev, y0, X);
    KFC_HIP(b.ctx, hipGetLastError());
    return KFC_OK;
}

// The filter of one iteration, replayed from a graph captured on its first run with this
// active set and degree (the launches read their coefficients from the device table,
// which the caller refreshed).
template <class S>
int run_filter(Batch<S>& b, S* X, S* Y, S* Z, int degree) {
    if (b.act.empty()) return KFC_OK;
    const int shape0 = b.shape, grid0 = b.grid;
    const int trc = tune_bapply<S>(b, X, Y, b.dcoefs.template as<double>());   // outside any capture
    if (trc != KFC_OK) return trc;
    if (!graphs_enabled()) return filter_launches<S>(b, X, Y, Z, degree);
    hipStream_t st = b.ctx->stream;
    if (!b.fexec || b.fact != b.act || b.fdeg != degree || b.fcoefs != b.dcoefs.p || b.shape != shape0 ||
        b.grid != grid0) {
        if (b.fexec) {
            KFC_HIP(b.ctx, hipGraphExecDestroy(b.fexec));
            b.fexec = nullptr;
        }
        KFC_HIP(b.ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = filter_launches<S>(b, X, Y, Z, degree);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc != KFC_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return kfc_fail(b.ctx, KFC_EDEVICE, std::string("filter graph capture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&b.fexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            b.fexec = nullptr;
            return kfc_fail(b.ctx, KFC_EDEVICE, std::string("filter graph instantiate: ") + hipGetErrorString(ei));
        }
        b.fact = b.act;
        b.fdeg = degree;
        b.fcoefs = b.dcoefs.p;
    }
    KFC_HIP(b.ctx, hipGraphLaunch(b.fexec, st));
    return KFC_OK;
}

template <class S>
int bgram(Batch<S>& b, const S* X, const S* Y, std::vector<cd>& G) {
    using A = typename Dbl<S>::type;
    const int na = int(b.act.size()), nb = b.nb;
    G.assign(size_t(na) * nb * nb, cd(0));
    if (na == 0) return KFC_OK;
    hipLaunchKernelGGL(k_bgram<S>, dim3(na), dim3(256), 0, b.ctx->stream, b.dev, X, Y, b.dG.template as<A>());
    KFC_HIP(b.ctx, hipGetLastError());
    std::vector<A> h(size_t(na) * nb * nb);
    KFC_HIP(b.ctx, hipMemcpyAsync(h.data(), b.dG.p, sizeof(A) * h.size(), hipMemcpyDeviceToHost, b.ctx->stream));
    KFC_HIP(b.ctx, hipStreamSynchronize(b.ctx->stream));
    for (size_t i = 0; i < h.size(); ++i) {
        if constexpr (IsCx<S>::value) G[i] = cd(h[i].re, h[i].im);
        else G[i] = cd(h[i], 0.0);
    }
    return KFC_OK;
}

// Y = X C per active problem (C: na blocks of nb x nc)
template <class S>
int brot(Batch<S>& b, const S* X, S* Y, const std::vector<cd>& C, int nc) {
    using A = typename Dbl<S>::type;
    const int na = int(b.act.size()), nb = b.nb;
    if (na == 0) return KFC_OK;
    std::vector<A> h(C.size());
    for (size_t i = 0; i < h.size(); ++i) {
        if constexpr (IsCx<S>::value) h[i] = A{C[i].real(), C[i].imag()};
        else h[i] = C[i].real();
    }
    KFC_HIP(b.ctx, hipMemcpyAsync(b.dC.p, h.data(), sizeof(A) * h.size(), hipMemcpyHostToDevice, b.ctx->stream));
    int64_t nmax = 0;
    for (int p : b.act) nmax = std::max(nmax, b.N[p]);
    hipLaunchKernelGGL(k_brot<S>, dim3(unsigned((nmax + 255) / 256), unsigned(na)), dim3(256), 0, b.ctx->stream, b.dev,
                       X, Y, b.dC.template as<A>(), nc);
    KFC_HIP(b.ctx, hipGetLastError());
    KFC_HIP(b.ctx, hipStreamSynchronize(b.ctx->stream));   // the staged C dies here
    return KFC_OK;
}

// Orthonormalize every active block in place (SVQB, twice: X <- X D U diag(lambda)^-1/2
// with D the inverse column norms).  A direction whose Gram eigenvalue is below
// 1e-14 of the largest is scaled as if it sat at that floor: what survives of it
// is rounding noise, which the next filter treats as a fresh direction.
template <class S>
int bsvqb(Batch<S>& b, S* X, S* T) {
    const int nb = b.nb, na = int(b.act.size());
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<cd> G;
        int rc = bgram<S>(b, X, X, G);
        if (rc != KFC_OK) return rc;
        std::vector<cd> C(size_t(na) * nb * nb, cd(0));
        parallel_for(na, [&](int q) {
            std::vector<cd> g(G.begin() + size_t(q) * nb * nb, G.begin() + size_t(q + 1) * nb * nb);
            std::vector<double> d(nb);
            for (int i = 0; i < nb; ++i) d[i] = 1.0 / std::sqrt(std::max(g[i + i * nb].real(), 1e-300));
            for (int j = 0; j < nb; ++j)
                for (int i = 0; i < nb; ++i) g[i + j * nb] *= d[i] * d[j];
            std::vector<double> lam;
            std::vector<cd> U;
            herm_eig(nb, g, lam, U);
            const double top = std::max(lam.back(), 1e-300);
            for (int j = 0; j < nb; ++j) {
                const double l = std::max(lam[j], 1e-14 * top);
                for (int i = 0; i < nb; ++i) C[size_t(q) * nb * nb + i + j * nb] = d[i] * U[i + j * nb] / std::sqrt(l);
            }
        });
        rc = brot<S>(b, X, T, C, nb);
        if (rc == KFC_OK) rc = bcopy<S>(b, T, X);
        if (rc != KFC_OK) return rc;
    }
    return KFC_OK;
}

struct ProbState {
    std::vector<double> theta;   // nb Ritz values
    std::vector<double> resid;   // nb residual norms
    double hi = 0, lo = 0;
    int iters = 0;
    int64_t matvec = 0;
    bool done = false;
};

// Rayleigh-Ritz of every active block: HX = H X, theta, and the rotated X, HX.
template <class S>
int brr(Batch<S>& b, S* X, S* HX, S* T, std::vector<ProbState>& ps) {
    const int nb = b.nb, na = int(b.act.size());
    std::vector<double> one(3 * size_t(b.n), 0.0);
    for (int p : b.act) one[3 * p] = 1.0;
    int rc = bapply<S>(b, X, nullptr, HX, one);
    if (rc != KFC_OK) return rc;
    std::vector<cd> A;
    rc = bgram<S>(b, X, HX, A);
    if (rc != KFC_OK) return rc;
    std::vector<cd> C(size_t(na) * nb * nb);
    parallel_for(na, [&](int q) {
        const int p = b.act[q];
        std::vector<cd> a(A.begin() + size_t(q) * nb * nb, A.begin() + size_t(q + 1) * nb * nb);
        for (int j = 0; j < nb; ++j)
            for (int i = 0; i < nb; ++i) a[i + j * nb] = 0.5 * (a[i + j * nb] + std::conj(A[size_t(q) * nb * nb + j + i * nb]));
        std::vector<double> th;
        std::vector<cd> U;
        herm_eig(nb, a, th, U);
        ps[p].theta = th;
        ps[p].matvec += nb;
        std::copy(U.begin(), U.end(), C.begin() + size_t(q) * nb * nb);
    });
    rc = brot<S>(b, X, T, C, nb);
    if (rc == KFC_OK) rc = bcopy<S>(b, T, X);
    if (rc == KFC_OK) rc = brot<S>(b, HX, T, C, nb);
    if (rc == KFC_OK) rc = bcopy<S>(b, T, HX);
    if (rc != KFC_OK) return rc;
    // fresh residual norms of the rotated pairs
    std::vector<double> th(size_t(na) * nb);
    for (int q = 0; q < na; ++q)
        for (int j = 0; j < nb; ++j) th[size_t(q) * nb + j] = ps[b.act[q]].theta[j];
    KFC_HIP(b.ctx, hipMemcpyAsync(b.dth.p, th.data(), sizeof(double) * th.size(), hipMemcpyHostToDevice, b.ctx->stream));
    hipLaunchKernelGGL(k_bresid<S>, dim3(nb, na), dim3(256), 0, b.ctx->stream, b.dev, X, HX, b.dth.template as<double>(),
                       b.dr.template as<double>());
    KFC_HIP(b.ctx, hipGetLastError());
    std::vector<double> r(th.size());
    KFC_HIP(b.ctx, hipMemcpyAsync(r.data(), b.dr.p, sizeof(double) * r.size(), hipMemcpyDeviceToHost, b.ctx->stream));
    KFC_HIP(b.ctx, hipStreamSynchronize(b.ctx->stream));
    for (int q = 0; q < na; ++q) {
        auto& s = ps[b.act[q]];
        s.resid.resize(nb);
        for (int j = 0; j < nb; ++j) s.resid[j] = std::sqrt(r[size_t(q) * nb + j]);
    }
    return KFC_OK;
}

// One precision phase: filter, orthonormalize, Rayleigh-Ritz until the lowest m of
// every problem reach tol_p (or maxiter).  X holds the current blocks.
template <class S>
int bphase(Batch<S>& b, std::vector<ProbState>& ps, int m, int degree, int64_t maxiter,
           const std::vector<double>& tolp) {
    S* X = b.bX.template as<S>();
    S* Y = b.bY.template as<S>();
    S* Z = b.bZ.template as<S>();
    S* HX = b.bH.template as<S>();
    const int nb = b.nb;
    auto converged = [&](int p) {
        for (int j = 0; j < m; ++j)
            if (!(ps[p].resid.size() == size_t(nb) && ps[p].resid[j] <= tolp[p])) return false;
        return true;
    };
    std::vector<int> act;
    for (int p = 0; p < b.n; ++p)
        if (!converged(p)) act.push_back(p);
    int rc = set_active<S>(b, act);
    if (rc != KFC_OK) return rc;
    if (act.empty()) return KFC_OK;
    rc = bsvqb<S>(b, X, Y);
    if (rc == KFC_OK) rc = brr<S>(b, X, HX, Y, ps);
    if (rc != KFC_OK) return rc;
    for (int64_t it = 0; it < maxiter; ++it) {
        act.clear();
        for (int p : b.act)
            if (!converged(p)) act.push_back(p);
            else ps[p].done = true;
        rc = set_active<S>(b, act);
        if (rc != KFC_OK || act.empty()) return rc;
        // filter interval per problem: damp [theta_{nb-1}, hi], scale at theta_0
        std::vector<double> e(b.n), c(b.n), sig(b.n), tau(b.n);
        for (int p : act) {
            const double a = ps[p].theta[nb - 1], bb = ps[p].hi, a0 = ps[p].theta[0];
            e[p] = std::max(0.5 * (bb - a), 1e-12 * std::fabs(bb));
            c[p] = 0.5 * (bb + a);
            sig[p] = e[p] / (a0 - c[p]);
            tau[p] = 2.0 / sig[p];
        }
        // every step's coefficients: step 0 Y = (H X - c X) sigma / e, then the
        // three-term recurrence z = a H y + b y + g x (Zhou-Saad scaling)
        const size_t stride = 3 * size_t(b.n), need = stride * size_t(degree);
        if (b.ncoefs < need) {
            if (b.hcoefs) KFC_HIP(b.ctx, hipHostFree(b.hcoefs));
            b.hcoefs = nullptr;
            b.ncoefs = 0;
            KFC_HIP(b.ctx, hipHostMalloc(reinterpret_cast<void**>(&b.hcoefs), sizeof(double) * need, hipHostMallocDefault));
            b.ncoefs = need;
            KFC_HIP(b.ctx, b.dcoefs.ensure(sizeof(double) * need));
        }
        KFC_HIP(b.ctx, hipStreamSynchronize(b.ctx->stream));   // the previous iteration's table is read
        std::fill(b.hcoefs, b.hcoefs + need, 0.0);
        for (int p : act) {
            b.hcoefs[3 * p] = sig[p] / e[p];
            b.hcoefs[3 * p + 1] = -c[p] * sig[p] / e[p];
        }
        for (int k = 2; k <= degree; ++k) {
            double* row = b.hcoefs + size_t(k - 1) * stride;
            for (int p : act) {
                const double sn = 1.0 / (tau[p] - sig[p]);
                row[3 * p] = 2.0 * sn / e[p];
                row[3 * p + 1] = -2.0 * c[p] * sn / e[p];
                row[3 * p + 2] = -sig[p] * sn;
                sig[p] = sn;
            }
        }
        KFC_HIP(b.ctx, hipMemcpyAsync(b.dcoefs.p, b.hcoefs, sizeof(double) * need, hipMemcpyHostToDevice,
                                      b.ctx->stream));
        rc = run_filter<S>(b, X, Y, Z, degree);
        if (rc != KFC_OK) return rc;
        for (int p : act) {
            ps[p].matvec += int64_t(degree) * nb;
            ps[p].iters += 1;
        }
        rc = bsvqb<S>(b, X, Y);
        if (rc == KFC_OK) rc = brr<S>(b, X, HX, Y, ps);
        if (rc != KFC_OK) return rc;
    }
    return KFC_OK;
}

template <class S>
int cheb_batch_t(int n, kfc_op* const* ops, int64_t m, int64_t extra, int degree, double tol, int64_t maxiter,
                 uint64_t seed, int mode, double* E, double* const* V, kfc_stats* stats) {
    using L = typename std::conditional<IsCx<S>::value, Cx<float>, float>::type;
    kfc_ctx* ctx = ops[0]->ctx;
    auto t0 = clk::now();
    std::vector<Band1D> bands(n);
    std::vector<int> brc(n, KFC_OK);
    for (int i = 0; i < n; ++i) {   // the memoized diagonals first (an operator may appear twice)
        const int rc = ops[i]->ndims == 1 ? op_host_diag(ops[i]) : KFC_OK;
        if (rc != KFC_OK) return rc;
    }
    parallel_for(n, [&](int i) { brc[i] = band_of(ops[i], bands[i]); });
    for (int i = 0; i < n; ++i) {
        if (brc[i] != KFC_OK) return brc[i];
        if (m + extra > bands[i].N)
            return kfc_fail(ctx, KFC_EARG, "kfc_chebyshev_batch: m + extra exceeds the size of problem " + std::to_string(i));
    }
    const int nb = int(m + extra);
    if (nb > 32) return kfc_fail(ctx, KFC_EUNSUPPORTED, "kfc_chebyshev_batch: m + extra above 32 is not served");
    std::vector<ProbState> ps(n);
    for (int p = 0; p < n; ++p) {
        ps[p].hi = bands[p].hi;
        ps[p].lo = bands[p].lo;
    }
    int64_t total = 0;
    for (auto& x : bands) total += x.N * nb;
    // random start (seeded per problem), FP64 on the host, the problems on host threads
    std::vector<cd> X0(static_cast<size_t>(total));
    {
        std::vector<int64_t> offs(n + 1, 0);
        for (int p = 0; p < n; ++p) offs[p + 1] = offs[p] + bands[p].N * nb;
        parallel_for(n, [&](int p) {
            std::mt19937_64 g(seed + 0x9e3779b97f4a7c15ull * uint64_t(p + 1));
            std::normal_distribution<double> nd;
            for (int64_t i = 0; i < bands[p].N * nb; ++i)
                X0[offs[p] + i] = IsCx<S>::value ? cd(nd(g), nd(g)) : cd(nd(g), 0.0);
        });
    }
    Batch<S> b64;
    int rc = batch_init<S>(b64, ctx, bands, nb);
    if (rc != KFC_OK) return rc;
    std::vector<double> tol64(n, tol);
    if (mode == KFC_MODE_MIXED) {
        Batch<L> b32;
        rc = batch_init<L>(b32, ctx, bands, nb);
        if (rc != KFC_OK) return rc;
        std::vector<L> h(X0.size());
        for (size_t i = 0; i < h.size(); ++i) h[i] = s_from<L>(X0[i]);
        KFC_HIP(ctx, hipMemcpy(b32.bX.p, h.data(), sizeof(L) * h.size(), hipMemcpyHostToDevice));
        // FP32 to a few units of roundoff of each spectrum's scale
        std::vector<double> tol32(n);
        for (int p = 0; p < n; ++p)
            tol32[p] = std::max(tol, 64.0 * 5.96e-8 * std::max(std::fabs(bands[p].hi), std::fabs(bands[p].lo)));
        rc = bphase<L>(b32, ps, int(m), degree, maxiter, tol32);
        if (rc != KFC_OK) return rc;
        hipLaunchKernelGGL((k_bconvert<L, S>), dim3(1024), dim3(256), 0, ctx->stream, b32.bX.template as<L>(), b64.bX.template as<S>(),
                           total);
        KFC_HIP(ctx, hipGetLastError());
        for (auto& s : ps) s.resid.clear();   // FP64 residuals decide from here on
    } else {
        std::vector<S> h(X0.size());
        for (size_t i = 0; i < h.size(); ++i) h[i] = s_from<S>(X0[i]);
        KFC_HIP(ctx, hipMemcpy(b64.bX.p, h.data(), sizeof(S) * h.size(), hipMemcpyHostToDevice));
    }
    rc = bphase<S>(b64, ps, int(m), degree, maxiter, tol64);
    if (rc != KFC_OK) return rc;
    // every problem's lowest m pairs, on FP64 residuals of its last Rayleigh-Ritz
    int bad = 0;
    double worst = 0;
    for (int p = 0; p < n; ++p)
        for (int j = 0; j < m; ++j) {
            if (!(ps[p].resid[j] <= tol)) ++bad;
            worst = std::max(worst, ps[p].resid[j]);
        }
    std::vector<S> Xh(static_cast<size_t>(total));
    KFC_HIP(ctx, hipMemcpy(Xh.data(), b64.bX.p, sizeof(S) * Xh.size(), hipMemcpyDeviceToHost));
    for (int p = 0; p < n; ++p) {
        for (int j = 0; j < m; ++j) E[size_t(p) * m + j] = ps[p].theta[j];
        const int64_t N = bands[p].N;
        double* out = V[p];
        for (int64_t j = 0; j < m; ++j)
            for (int64_t i = 0; i < N; ++i) {
                const S v = Xh[b64.xo[p] + j * N + i];
                if constexpr (IsCx<S>::value) {
                    out[2 * (i + j * N)] = double(v.re);
                    out[2 * (i + j * N) + 1] = double(v.im);
                } else {
                    out[i + j * N] = double(v);
                }
            }
        if (stats) {
            kfc_stats& st = stats[p];
            st = kfc_stats{};
            st.iterations = ps[p].iters;
            st.n_matvec = ps[p].matvec;
            st.n_apply_calls = ps[p].iters;
            st.max_residual = *std::max_element(ps[p].resid.begin(), ps[p].resid.begin() + m);
            st.n_converged = int(m);
            st.degree_last = degree;
            st.seconds_total = std::chrono::duration<double>(clk::now() - t0).count();
        }
    }
    if (bad > 0)
        return kfc_fail(ctx, KFC_ENOCONV, "ChebyshevFilter (batch): " + std::to_string(bad) + " eigenpairs did not converge within " +
                                              std::to_string(maxiter) + " iterations (maximum residual " + std::to_string(worst) + ")");
    return KFC_OK;
}

}  // namespace

extern "C" int kfc_chebyshev_batch(int32_t n, kfc_op* const* ops, int64_t m, int64_t extra, int32_t degree, double tol,
                                   int64_t maxiter, uint64_t seed, int32_t mode, double* E, double* const* V,
                                   kfc_stats* stats) {
    if (n < 1 || !ops || !E || !V) return kfc_fail(nullptr, KFC_EARG, "kfc_chebyshev_batch: empty batch or null argument");
    for (int i = 0; i < n; ++i)
        if (!ops[i] || !V[i]) return kfc_fail(ops[0] ? ops[0]->ctx : nullptr, KFC_EARG, "kfc_chebyshev_batch: null operator or output");
    kfc_ctx* ctx = ops[0]->ctx;
    for (int i = 1; i < n; ++i) {
        if (ops[i]->ctx != ctx) return kfc_fail(ctx, KFC_EARG, "kfc_chebyshev_batch: every operator must live on one context");
        if (ops[i]->cplx != ops[0]->cplx) return kfc_fail(ctx, KFC_EARG, "kfc_chebyshev_batch: mixed real and complex operators");
    }
    if (m < 1 || extra < 1) return kfc_fail(ctx, KFC_EARG, "kfc_chebyshev_batch: m and extra must be positive");
    if (degree < 1 || degree > 200) return kfc_fail(ctx, KFC_EARG, "kfc_chebyshev_batch: degree must lie in 1:200");
    if (!(tol > 0) || !std::isfinite(tol)) return kfc_fail(ctx, KFC_EARG, "kfc_chebyshev_batch: tol must be positive and finite");
    if (maxiter < 1) return kfc_fail(ctx, KFC_EARG, "kfc_chebyshev_batch: maxiter must be at least 1");
    if (mode != KFC_MODE_F64 && mode != KFC_MODE_MIXED) return kfc_fail(ctx, KFC_EARG, "kfc_chebyshev_batch: unknown mode");
    return ops[0]->cplx ? cheb_batch_t<Cx<double>>(n, ops, m, extra, degree, tol, maxiter, seed, mode, E, V, stats)
                        : cheb_batch_t<double>(n, ops, m, extra, degree, tol, maxiter, seed, mode, E, V, stats);
}
