// Tall-skinny block kernels for the eigensolver loops (wave64, FP64 accumulation).
#include "blas.h"

#include <algorithm>
#include <cmath>

namespace {

template <class A> __device__ inline A acc_zero();
template <> __device__ inline double acc_zero<double>() { return 0.0; }
template <> __device__ inline Cx<double> acc_zero<Cx<double>>() { return Cx<double>{0.0, 0.0}; }

__device__ inline double to_acc(float x) { return double(x); }
__device__ inline double to_acc(double x) { return x; }
__device__ inline Cx<double> to_acc(Cx<float> x) { return Cx<double>{double(x.re), double(x.im)}; }
__device__ inline Cx<double> to_acc(Cx<double> x) { return x; }

template <class S> __device__ inline S from_acc(typename AccOf<S>::type a);
template <> __device__ inline float from_acc<float>(double a) { return float(a); }
template <> __device__ inline double from_acc<double>(double a) { return a; }
template <> __device__ inline Cx<float> from_acc<Cx<float>>(Cx<double> a) { return Cx<float>{float(a.re), float(a.im)}; }
template <> __device__ inline Cx<double> from_acc<Cx<double>>(Cx<double> a) { return a; }

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum of a per-thread value over a 256-thread block; result valid in thread 0.
__device__ inline double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0) s = red[0] + red[1] + red[2] + red[3];
    return s;
}

inline int nchunks_for(int64_t N, int64_t per = 4096, int cap = 1024) {
    return int(std::max<int64_t>(1, std::min<int64_t>(cap, (N + per - 1) / per)));
}

inline unsigned grid1(int64_t total) {
    return unsigned(std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536)));
}

inline void acc_add(double& a, double b) { a += b; }
inline void acc_add(cd& a, cd b) { a += b; }

}  // namespace

// ------------------------------------------------------------------------------------------
// Gram: G = A^H B, 64 x 64 output tiles, 32-row LDS stages, partial per row chunk.

template <class S, class A>
__global__ __launch_bounds__(256) void k_gram(const S* __restrict__ Am, int64_t lda, int p, const S* __restrict__ Bm,
                                              int64_t ldb, int q, int64_t N, int64_t rows_per, A* __restrict__ part) {
    __shared__ A sA[64][33];
    __shared__ A sB[64][33];
    const int tid = threadIdx.x, ti = tid & 15, tj = tid >> 4;
    const int pt = blockIdx.y, qt = blockIdx.z, chunk = blockIdx.x;
    const int64_t r0 = chunk * rows_per, r1 = min(N, r0 + rows_per);
    A acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = acc_zero<A>();
    for (int64_t rb = r0; rb < r1; rb += 32) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int e = tid + k * 256, r = e & 31, c = e >> 5;
            const int64_t row = rb + r;
            const int ca = pt * 64 + c, cb = qt * 64 + c;
            sA[c][r] = (row < r1 && ca < p) ? to_acc(Am[row + ca * lda]) : acc_zero<A>();
            sB[c][r] = (row < r1 && cb < q) ? to_acc(Bm[row + cb * ldb]) : acc_zero<A>();
        }
        __syncthreads();
#pragma unroll 4
        for (int r = 0; r < 32; ++r) {
            A a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = s_conj(sA[ti * 4 + i][r]);
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = sB[tj * 4 + j][r];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = s_fma(acc[i][j], a[i], b[j]);
        }
        __syncthreads();
    }
    A* out = part + ((int64_t)chunk * gridDim.y * gridDim.z + (int64_t)pt * gridDim.z + qt) * 4096;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) out[(ti * 4 + i) + (tj * 4 + j) * 64] = acc[i][j];
}

template <class A>
__global__ void k_gram_reduce(const A* __restrict__ part, int nchunk, int ntp, int ntq, int p, int q, A* __restrict__ G) {
    const int64_t total = (int64_t)p * q;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int i = int(x % p), j = int(x / p);
        const int pt = i >> 6, qt = j >> 6, ii = i & 63, jj = j & 63;
        A s = acc_zero<A>();
        for (int c = 0; c < nchunk; ++c) s = s_add(s, part[((int64_t)c * ntp * ntq + pt * ntq + qt) * 4096 + ii + jj * 64]);
        G[i + (int64_t)j * p] = s;
    }
}

template <class S>
int blas_gram(Work& w, const S* A, int64_t lda, int p, const S* B, int64_t ldb, int q, int64_t N, std::vector<cd>& G) {
    using Acc = typename AccOf<S>::type;
    G.assign(size_t(p) * q, cd(0));
    if (p == 0 || q == 0) return KFC_OK;
    const int ntp = (p + 63) / 64, ntq = (q + 63) / 64;
    const int nchunk = nchunks_for(N, 8192, 256);
    const int64_t rows_per = (N + nchunk - 1) / nchunk;
    KFC_HIP(w.ctx, w.part.ensure(sizeof(Acc) * 4096 * size_t(nchunk) * ntp * ntq));
    KFC_HIP(w.ctx, w.small.ensure(sizeof(Acc) * size_t(p) * q));
    hipLaunchKernelGGL((k_gram<S, Acc>), dim3(nchunk, ntp, ntq), dim3(256), 0, w.ctx->stream, A, lda, p, B, ldb, q, N,
                       rows_per, w.part.as<Acc>());
    hipLaunchKernelGGL(k_gram_reduce<Acc>, dim3(grid1(int64_t(p) * q)), dim3(256), 0, w.ctx->stream,
                       w.part.as<Acc>(), nchunk, ntp, ntq, p, q, w.small.as<Acc>());
    KFC_HIP(w.ctx, hipGetLastError());
    std::vector<Acc> h(size_t(p) * q);
    KFC_HIP(w.ctx, hipMemcpyAsync(h.data(), w.small.p, h.size() * sizeof(Acc), hipMemcpyDeviceToHost, w.ctx->stream));
    KFC_HIP(w.ctx, hipStreamSynchronize(w.ctx->stream));
    for (size_t i = 0; i < h.size(); ++i) {
        if constexpr (IsCx<S>::value) G[i] = cd(reinterpret_cast<Cx<double>&>(h[i]).re, reinterpret_cast<Cx<double>&>(h[i]).im);
        else G[i] = cd(reinterpret_cast<double&>(h[i]), 0.0);
    }
    return KFC_OK;
}

// ------------------------------------------------------------------------------------------
// GEMM: C[:, j] = sum_i A[:, i] M[i, j] (+ beta C): one row per thread, 16 output columns
// per block row; M read through uniform (scalar) loads.

template <class S>
__global__ __launch_bounds__(256) void k_gemm(const S* __restrict__ A, int64_t lda, int p,
                                              const typename AccOf<S>::type* __restrict__ M, int q, S* __restrict__ C,
                                              int64_t ldc, int64_t N, double beta) {
    using Acc = typename AccOf<S>::type;
    const int j0 = blockIdx.y * 16;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
        Acc acc[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] = acc_zero<Acc>();
        for (int i = 0; i < p; ++i) {
            const Acc a = to_acc(A[r + i * lda]);
            const Acc* Mi = M + i;
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (j0 + j < q) acc[j] = s_fma(acc[j], a, Mi[(int64_t)(j0 + j) * p]);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            if (j0 + j < q) {
                S* c = C + r + (int64_t)(j0 + j) * ldc;
                Acc v = acc[j];
                if (beta != 0.0) v = s_fma(v, s_scale(beta, s_one<Acc>()), to_acc(*c));
                *c = from_acc<S>(v);
            }
        }
    }
}

template <class S>
int blas_gemm(Work& w, const S* A, int64_t lda, int p, const std::vector<cd>& M, int q, S* C, int64_t ldc, int64_t N,
              double beta) {
    using Acc = typename AccOf<S>::type;
    if (q == 0 || N == 0) return KFC_OK;
    std::vector<Acc> h(size_t(p) * q);
    for (size_t i = 0; i < h.size(); ++i) {
        if constexpr (IsCx<S>::value) h[i] = Acc{M[i].real(), M[i].imag()};
        else h[i] = M[i].real();
    }
    KFC_HIP(w.ctx, w.small2.ensure(std::max<size_t>(1, h.size() * sizeof(Acc))));
    if (!h.empty())
        KFC_HIP(w.ctx, hipMemcpyAsync(w.small2.p, h.data(), h.size() * sizeof(Acc), hipMemcpyHostToDevice, w.ctx->stream));
    const unsigned gx = unsigned(std::max<int64_t>(1, std::min<int64_t>((N + 255) / 256, 4096)));
    hipLaunchKernelGGL(k_gemm<S>, dim3(gx, (q + 15) / 16), dim3(256), 0, w.ctx->stream, A, lda, p, w.small2.as<Acc>(), q,
                       C, ldc, N, beta);
    KFC_HIP(w.ctx, hipGetLastError());
    // the host staging vector dies here: wait for the copy to finish
    KFC_HIP(w.ctx, hipStreamSynchronize(w.ctx->stream));
    return KFC_OK;
}

// ------------------------------------------------------------------------------------------
// Column reductions: mode 0 |A|^2; mode 1 R = AX - theta X and |R|^2; mode 2 D|X|^2 and |X|^2.

template <class S>
__global__ __launch_bounds__(256) void k_colred(int mode, S* __restrict__ R, int64_t ldr, const S* __restrict__ AX,
                                                int64_t ldax, const S* __restrict__ X, int64_t ldx, int64_t N,
                                                int64_t rows_per, const double* __restrict__ theta,
                                                const double* __restrict__ D, double* __restrict__ part) {
    __shared__ double red[4];
    const int col = blockIdx.y, chunk = blockIdx.x;
    const int64_t r0 = chunk * rows_per, r1 = min(N, r0 + rows_per);
    double s1 = 0, s2 = 0;
    if (mode == 0) {
        const S* a = X + col * ldx;
        for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) s1 += double(s_abs2(a[r]));
    } else if (mode == 1) {
        const S* ax = AX + col * ldax;
        const S* x = X + col * ldx;
        S* rr = R + col * ldr;
        const auto th = typename RealOf<S>::type(theta[col]);
        for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) {
            S v = s_sub(ax[r], s_scale(th, x[r]));
            rr[r] = v;
            s1 += double(s_abs2(v));
        }
    } else {
        const S* x = X + col * ldx;
        for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) {
            const double a = double(s_abs2(x[r]));
            s1 += D[r] * a;
            s2 += a;
        }
    }
    s1 = block_sum(s1, red);
    if (threadIdx.x == 0) part[(col * gridDim.x + chunk) * 2] = s1;
    s2 = block_sum(s2, red);
    if (threadIdx.x == 0) part[(col * gridDim.x + chunk) * 2 + 1] = s2;
}

__global__ void k_colred_reduce(const double* __restrict__ part, int nchunk, int p, double* __restrict__ out) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= p) return;
    double a = 0, b = 0;
    for (int c = 0; c < nchunk; ++c) {
        a += part[(col * nchunk + c) * 2];
        b += part[(col * nchunk + c) * 2 + 1];
    }
    out[2 * col] = a;
    out[2 * col + 1] = b;
}

template <class S>
static int colred(Work& w, int mode, S* R, int64_t ldr, const S* AX, int64_t ldax, const S* X, int64_t ldx, int p,
                  int64_t N, const std::vector<double>* theta, const double* D, std::vector<double>& s1,
                  std::vector<double>* s2) {
    s1.assign(p, 0.0);
    if (s2) s2->assign(p, 0.0);
    if (p == 0) return KFC_OK;
    const int nchunk = nchunks_for(N, 16384, 256);
    const int64_t rows_per = (N + nchunk - 1) / nchunk;
    KFC_HIP(w.ctx, w.part.ensure(sizeof(double) * 2 * size_t(nchunk) * p));
    KFC_HIP(w.ctx, w.small.ensure(sizeof(double) * (2 * size_t(p) + size_t(p))));
    double* th = w.small.as<double>() + 2 * p;
    if (theta)
        KFC_HIP(w.ctx, hipMemcpyAsync(th, theta->data(), sizeof(double) * p, hipMemcpyHostToDevice, w.ctx->stream));
    hipLaunchKernelGGL(k_colred<S>, dim3(nchunk, p), dim3(256), 0, w.ctx->stream, mode, R, ldr, AX, ldax, X, ldx, N,
                       rows_per, th, D, w.part.as<double>());
    hipLaunchKernelGGL(k_colred_reduce, dim3((p + 63) / 64), dim3(64), 0, w.ctx->stream, w.part.as<double>(), nchunk, p,
                       w.small.as<double>());
    KFC_HIP(w.ctx, hipGetLastError());
    std::vector<double> h(2 * size_t(p));
    KFC_HIP(w.ctx, hipMemcpyAsync(h.data(), w.small.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, w.ctx->stream));
    KFC_HIP(w.ctx, hipStreamSynchronize(w.ctx->stream));
    for (int j = 0; j < p; ++j) {
        s1[j] = h[2 * j];
        if (s2) (*s2)[j] = h[2 * j + 1];
    }
    return KFC_OK;
}

template <class S>
int blas_colnorms(Work& w, const S* A, int64_t lda, int p, int64_t N, std::vector<double>& nrm) {
    int rc = colred<S>(w, 0, nullptr, 0, nullptr, 0, A, lda, p, N, nullptr, nullptr, nrm, nullptr);
    for (double& x : nrm) x = std::sqrt(x);
    return rc;
}

template <class S>
int blas_residual(Work& w, S* R, int64_t ldr, const S* AX, int64_t ldax, const S* X, int64_t ldx, int p, int64_t N,
                  const std::vector<double>& theta, std::vector<double>& nrm) {
    int rc = colred<S>(w, 1, R, ldr, AX, ldax, X, ldx, p, N, &theta, nullptr, nrm, nullptr);
    for (double& x : nrm) x = std::sqrt(x);
    return rc;
}

template <class S>
int blas_diag_expect(Work& w, const S* X, int64_t ldx, int p, int64_t N, const double* D, std::vector<double>& theta) {
    std::vector<double> num, den;
    int rc = colred<S>(w, 2, nullptr, 0, nullptr, 0, X, ldx, p, N, nullptr, D, num, &den);
    theta.resize(p);
    for (int j = 0; j < p; ++j) theta[j] = num[j] / den[j];
    return rc;
}

// ------------------------------------------------------------------------------------------
// Elementwise column kernels

template <class S>
__global__ void k_scale_cols(S* A, int64_t lda, int p, int64_t N, const double* s) {
    const int64_t total = N * p;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t col = x / N, r = x - col * N;
        A[r + col * lda] = s_scale(typename RealOf<S>::type(s[col]), A[r + col * lda]);
    }
}

template <class S>
int blas_scale_cols(Work& w, S* A, int64_t lda, int p, int64_t N, const std::vector<double>& s) {
    if (p == 0) return KFC_OK;
    KFC_HIP(w.ctx, w.small.ensure(sizeof(double) * p));
    KFC_HIP(w.ctx, hipMemcpyAsync(w.small.p, s.data(), sizeof(double) * p, hipMemcpyHostToDevice, w.ctx->stream));
    hipLaunchKernelGGL(k_scale_cols<S>, dim3(grid1(N * p)), dim3(256), 0, w.ctx->stream, A, lda, p, N,
                       w.small.as<double>());
    KFC_HIP(w.ctx, hipGetLastError());
    KFC_HIP(w.ctx, hipStreamSynchronize(w.ctx->stream));
    return KFC_OK;
}

template <class S>
__global__ void k_precond(S* R, int64_t ldr, int p, int64_t N, const double* D, const double* theta, double eta) {
    const int64_t total = N * p;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t col = x / N, r = x - col * N;
        double f;
        if (eta > 0) f = 1.0 / hypot(D[r] - theta[col], eta);
        else f = D[r];
        R[r + col * ldr] = s_scale(typename RealOf<S>::type(f), R[r + col * ldr]);
    }
}

template <class S>
int blas_precond(Work& w, S* R, int64_t ldr, int p, int64_t N, const double* D, const std::vector<double>& theta,
                 double eta) {
    if (p == 0) return KFC_OK;
    KFC_HIP(w.ctx, w.small.ensure(sizeof(double) * p));
    if (!theta.empty())
        KFC_HIP(w.ctx, hipMemcpyAsync(w.small.p, theta.data(), sizeof(double) * p, hipMemcpyHostToDevice, w.ctx->stream));
    hipLaunchKernelGGL(k_precond<S>, dim3(grid1(N * p)), dim3(256), 0, w.ctx->stream, R, ldr, p, N, D,
                       w.small.as<double>(), eta);
    KFC_HIP(w.ctx, hipGetLastError());
    KFC_HIP(w.ctx, hipStreamSynchronize(w.ctx->stream));
    return KFC_OK;
}

template <class S>
__global__ void k_axpby(const S* X, int64_t ldx, S* Y, int64_t ldy, int p, int64_t N, double alpha, double beta) {
    using R = typename RealOf<S>::type;
    const int64_t total = N * p;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t col = x / N, r = x - col * N;
        S y = beta == 0.0 ? s_zero(X[0]) : s_scale(R(beta), Y[r + col * ldy]);
        Y[r + col * ldy] = s_add(y, s_scale(R(alpha), X[r + col * ldx]));
    }
}

template <class S>
int blas_axpby(Work& w, const S* X, int64_t ldx, S* Y, int64_t ldy, int p, int64_t N, double alpha, double beta) {
    if (p == 0 || N == 0) return KFC_OK;
    hipLaunchKernelGGL(k_axpby<S>, dim3(grid1(N * p)), dim3(256), 0, w.ctx->stream, X, ldx, Y, ldy, p, N, alpha, beta);
    KFC_HIP(w.ctx, hipGetLastError());
    return KFC_OK;
}

template <class S, class T> __device__ inline T conv(S x);
template <> __device__ inline float conv<double, float>(double x) { return float(x); }
template <> __device__ inline double conv<float, double>(float x) { return double(x); }
template <> __device__ inline double conv<double, double>(double x) { return x; }
template <> __device__ inline float conv<float, float>(float x) { return x; }
template <> __device__ inline Cx<float> conv<Cx<double>, Cx<float>>(Cx<double> x) { return Cx<float>{float(x.re), float(x.im)}; }
template <> __device__ inline Cx<double> conv<Cx<float>, Cx<double>>(Cx<float> x) { return Cx<double>{x.re, x.im}; }
template <> __device__ inline Cx<double> conv<Cx<double>, Cx<double>>(Cx<double> x) { return x; }
template <> __device__ inline Cx<float> conv<Cx<float>, Cx<float>>(Cx<float> x) { return x; }

template <class S, class T>
__global__ void k_convert(const S* X, int64_t ldx, T* Y, int64_t ldy, int p, int64_t N) {
    const int64_t total = N * p;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t col = x / N, r = x - col * N;
        Y[r + col * ldy] = conv<S, T>(X[r + col * ldx]);
    }
}

template <class S, class T>
int blas_convert(Work& w, const S* X, int64_t ldx, T* Y, int64_t ldy, int p, int64_t N) {
    if (p == 0 || N == 0) return KFC_OK;
    hipLaunchKernelGGL((k_convert<S, T>), dim3(grid1(N * p)), dim3(256), 0, w.ctx->stream, X, ldx, Y, ldy, p, N);
    KFC_HIP(w.ctx, hipGetLastError());
    return KFC_OK;
}

// Counter-based normal deviates: splitmix64 of (seed, column, row), Box-Muller.
__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ inline void normal2(uint64_t key, double& a, double& b) {
    const uint64_t h = mix64(key);
    const double u1 = ((h >> 11) + 1.0) * (1.0 / 9007199254740993.0);      // (0, 1]
    const double u2 = (mix64(h) >> 11) * (1.0 / 9007199254740992.0);      // [0, 1)
    const double r = sqrt(-2.0 * log(u1));
    a = r * cos(6.283185307179586 * u2);
    b = r * sin(6.283185307179586 * u2);
}

template <class S>
__global__ void k_randn(S* X, int64_t ldx, int p, int64_t N, uint64_t seed) {
    using R = typename RealOf<S>::type;
    const int64_t total = N * p;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t col = x / N, r = x - col * N;
        double a, b;
        normal2(mix64(seed ^ mix64(uint64_t(col) * 0x100000001B3ull)) + uint64_t(r), a, b);
        if constexpr (IsCx<S>::value) X[r + col * ldx] = S{R(a), R(b)};
        else X[r + col * ldx] = R(a);
    }
}

template <class S>
int blas_randn(Work& w, S* X, int64_t ldx, int p, int64_t N, uint64_t seed) {
    if (p == 0 || N == 0) return KFC_OK;
    hipLaunchKernelGGL(k_randn<S>, dim3(grid1(N * p)), dim3(256), 0, w.ctx->stream, X, ldx, p, N, seed);
    KFC_HIP(w.ctx, hipGetLastError());
    return KFC_OK;
}

// ------------------------------------------------------------------------------------------
// Lanczos orthogonalization passes (one read of the basis each).

template <class A, int KM>
struct LzArgs {
    A hs[KM];
    double sc[KM];
};

template <class S, int KM>
__global__ __launch_bounds__(256) void k_lz_multidot(const S* __restrict__ V, int64_t ldv, int k,
                                                     const LzArgs<typename AccOf<S>::type, KM> arg,
                                                     const S* __restrict__ wv, int64_t N, int64_t rows_per,
                                                     typename AccOf<S>::type* __restrict__ part) {
    using Acc = typename AccOf<S>::type;
    __shared__ Acc red[4][KM + 1];
    const int chunk = blockIdx.x;
    const int64_t r0 = chunk * rows_per, r1 = min(N, r0 + rows_per);
    Acc acc[KM];
    double wn = 0;
#pragma unroll
    for (int i = 0; i < KM; ++i) acc[i] = acc_zero<Acc>();
    for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) {
        const Acc x = to_acc(wv[r]);
        wn += double(s_abs2(x));
#pragma unroll
        for (int i = 0; i < KM; ++i)
            if (i < k) acc[i] = s_fma(acc[i], s_conj(to_acc(V[r + i * ldv])), x);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < KM; ++i) {
        if (i < k) {
            if constexpr (IsCx<S>::value) {
                acc[i].re = wave_sum(acc[i].re);
                acc[i].im = wave_sum(acc[i].im);
            } else {
                acc[i] = wave_sum(acc[i]);
            }
            if (lane == 0) red[w][i] = acc[i];
        }
    }
    wn = wave_sum(wn);
    if (lane == 0) {
        if constexpr (IsCx<S>::value) red[w][KM] = Acc{wn, 0.0};
        else red[w][KM] = wn;
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= KM; i += 256) {
        if (i < k || i == KM) {
            Acc s = s_add(s_add(red[0][i], red[1][i]), s_add(red[2][i], red[3][i]));
            if (i < k) s = s_scale(arg.sc[i], s);
            part[(int64_t)chunk * (KM + 1) + i] = s;
        }
    }
}

template <class S, int KM>
__global__ __launch_bounds__(256) void k_lz_update(const S* __restrict__ V, int64_t ldv, int k,
                                                   const LzArgs<typename AccOf<S>::type, KM> arg, S* __restrict__ wv,
                                                   int64_t N, int64_t rows_per, double* __restrict__ part) {
    using Acc = typename AccOf<S>::type;
    __shared__ double red[4];
    const int chunk = blockIdx.x;
    const int64_t r0 = chunk * rows_per, r1 = min(N, r0 + rows_per);
    double wn = 0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) {
        Acc x = to_acc(wv[r]);
#pragma unroll
        for (int i = 0; i < KM; ++i)
            if (i < k) x = s_sub(x, s_mul(to_acc(V[r + i * ldv]), arg.hs[i]));
        const S y = from_acc<S>(x);
        wv[r] = y;
        wn += double(s_abs2(to_acc(y)));
    }
    wn = block_sum(wn, red);
    if (threadIdx.x == 0) part[chunk] = wn;
}

template <class A>
__global__ void k_sum_parts(const A* __restrict__ part, int nchunk, int stride, int n, A* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    A s = part[i];
    for (int c = 1; c < nchunk; ++c) s = s_add(s, part[(int64_t)c * stride + i]);
    out[i] = s;
}

template <class S, int KM>
static int lz_multidot_t(Work& w, const S* V, int64_t ldv, int k, const double* scale, const S* wv, int64_t N,
                         std::vector<cd>& h, double* wnorm2) {
    using Acc = typename AccOf<S>::type;
    LzArgs<Acc, KM> arg;
    std::memset(&arg, 0, sizeof(arg));
    for (int i = 0; i < k; ++i) arg.sc[i] = scale ? scale[i] : 1.0;
    const int nchunk = nchunks_for(N, 32768, 1024);
    const int64_t rows_per = (N + nchunk - 1) / nchunk;
    KFC_HIP(w.ctx, w.part.ensure(sizeof(Acc) * size_t(nchunk) * (KM + 1)));
    KFC_HIP(w.ctx, w.small.ensure(sizeof(Acc) * (KM + 1)));
    hipLaunchKernelGGL((k_lz_multidot<S, KM>), dim3(nchunk), dim3(256), 0, w.ctx->stream, V, ldv, k, arg, wv, N, rows_per,
                       w.part.as<Acc>());
    hipLaunchKernelGGL(k_sum_parts<Acc>, dim3(1), dim3(128), 0, w.ctx->stream, w.part.as<Acc>(), nchunk, KM + 1, KM + 1,
                       w.small.as<Acc>());
    KFC_HIP(w.ctx, hipGetLastError());
    std::vector<Acc> hh(KM + 1);
    KFC_HIP(w.ctx, hipMemcpyAsync(hh.data(), w.small.p, sizeof(Acc) * (KM + 1), hipMemcpyDeviceToHost, w.ctx->stream));
    KFC_HIP(w.ctx, hipStreamSynchronize(w.ctx->stream));
    h.resize(k);
    for (int i = 0; i < k; ++i) {
        if constexpr (IsCx<S>::value) h[i] = cd(hh[i].re, hh[i].im);
        else h[i] = cd(hh[i], 0.0);
    }
    if (wnorm2) {
        if constexpr (IsCx<S>::value) *wnorm2 = hh[KM].re;
        else *wnorm2 = hh[KM];
    }
    return KFC_OK;
}

template <class S, int KM>
static int lz_update_t(Work& w, const S* V, int64_t ldv, int k, const std::vector<cd>& hs, S* wv, int64_t N,
                       double* wnorm2) {
    using Acc = typename AccOf<S>::type;
    LzArgs<Acc, KM> arg;
    std::memset(&arg, 0, sizeof(arg));
    for (int i = 0; i < k; ++i) {
        if constexpr (IsCx<S>::value) arg.hs[i] = Acc{hs[i].real(), hs[i].imag()};
        else arg.hs[i] = hs[i].real();
    }
    const int nchunk = nchunks_for(N, 32768, 1024);
    const int64_t rows_per = (N + nchunk - 1) / nchunk;
    KFC_HIP(w.ctx, w.part.ensure(sizeof(double) * size_t(nchunk)));
    KFC_HIP(w.ctx, w.small.ensure(sizeof(double)));
    hipLaunchKernelGGL((k_lz_update<S, KM>), dim3(nchunk), dim3(256), 0, w.ctx->stream, V, ldv, k, arg, wv, N, rows_per,
                       w.part.as<double>());
    hipLaunchKernelGGL(k_sum_parts<double>, dim3(1), dim3(64), 0, w.ctx->stream, w.part.as<double>(), nchunk, 1, 1,
                       w.small.as<double>());
    KFC_HIP(w.ctx, hipGetLastError());
    double v = 0;
    KFC_HIP(w.ctx, hipMemcpyAsync(&v, w.small.p, sizeof(double), hipMemcpyDeviceToHost, w.ctx->stream));
    KFC_HIP(w.ctx, hipStreamSynchronize(w.ctx->stream));
    if (wnorm2) *wnorm2 = v;
    return KFC_OK;
}

template <class S>
int lz_multidot(Work& w, const S* V, int64_t ldv, int k, const double* scale, const S* wv, int64_t N,
                std::vector<cd>& h, double* wnorm2) {
    if (k <= 8) return lz_multidot_t<S, 8>(w, V, ldv, k, scale, wv, N, h, wnorm2);
    if (k <= 16) return lz_multidot_t<S, 16>(w, V, ldv, k, scale, wv, N, h, wnorm2);
    if (k <= 32) return lz_multidot_t<S, 32>(w, V, ldv, k, scale, wv, N, h, wnorm2);
    if (k <= 64) return lz_multidot_t<S, 64>(w, V, ldv, k, scale, wv, N, h, wnorm2);
    return kfc_fail(w.ctx, KFC_EUNSUPPORTED, "Lanczos: Krylov dimension above 63 is not supported");
}

template <class S>
int lz_update(Work& w, const S* V, int64_t ldv, int k, const std::vector<cd>& hs, S* wv, int64_t N, double* wnorm2) {
    if (k <= 8) return lz_update_t<S, 8>(w, V, ldv, k, hs, wv, N, wnorm2);
    if (k <= 16) return lz_update_t<S, 16>(w, V, ldv, k, hs, wv, N, wnorm2);
    if (k <= 32) return lz_update_t<S, 32>(w, V, ldv, k, hs, wv, N, wnorm2);
    if (k <= 64) return lz_update_t<S, 64>(w, V, ldv, k, hs, wv, N, wnorm2);
    return kfc_fail(w.ctx, KFC_EUNSUPPORTED, "Lanczos: Krylov dimension above 63 is not supported");
}

// ------------------------------------------------------------------------------------------
// Explicit instantiations

#define KFC_INST(S)                                                                                                 \
    template int blas_gram<S>(Work&, const S*, int64_t, int, const S*, int64_t, int, int64_t, std::vector<cd>&);     \
    template int blas_gemm<S>(Work&, const S*, int64_t, int, const std::vector<cd>&, int, S*, int64_t, int64_t,      \
                              double);                                                                              \
    template int blas_colnorms<S>(Work&, const S*, int64_t, int, int64_t, std::vector<double>&);                    \
    template int blas_scale_cols<S>(Work&, S*, int64_t, int, int64_t, const std::vector<double>&);                  \
    template int blas_residual<S>(Work&, S*, int64_t, const S*, int64_t, const S*, int64_t, int, int64_t,            \
                                  const std::vector<double>&, std::vector<double>&);                                \
    template int blas_precond<S>(Work&, S*, int64_t, int, int64_t, const double*, const std::vector<double>&,       \
                                 double);                                                                           \
    template int blas_diag_expect<S>(Work&, const S*, int64_t, int, int64_t, const double*, std::vector<double>&);  \
    template int blas_axpby<S>(Work&, const S*, int64_t, S*, int64_t, int, int64_t, double, double);                \
    template int blas_randn<S>(Work&, S*, int64_t, int, int64_t, uint64_t);                                         \
    template int lz_multidot<S>(Work&, const S*, int64_t, int, const double*, const S*, int64_t, std::vector<cd>&,   \
                                double*);                                                                           \
    template int lz_update<S>(Work&, const S*, int64_t, int, const std::vector<cd>&, S*, int64_t, double*);

KFC_INST(double)
KFC_INST(float)
KFC_INST(Cx<double>)
KFC_INST(Cx<float>)

template int blas_convert<double, float>(Work&, const double*, int64_t, float*, int64_t, int, int64_t);
template int blas_convert<float, double>(Work&, const float*, int64_t, double*, int64_t, int, int64_t);
template int blas_convert<double, double>(Work&, const double*, int64_t, double*, int64_t, int, int64_t);
template int blas_convert<Cx<double>, Cx<float>>(Work&, const Cx<double>*, int64_t, Cx<float>*, int64_t, int, int64_t);
template int blas_convert<Cx<float>, Cx<double>>(Work&, const Cx<float>*, int64_t, Cx<double>*, int64_t, int, int64_t);
template int blas_convert<Cx<double>, Cx<double>>(Work&, const Cx<double>*, int64_t, Cx<double>*, int64_t, int,
                                                  int64_t);
