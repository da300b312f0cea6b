// The warm-start network: a GELU MLP with BF16 GEMM operands on MFMA
// (gfx950's v_mfma_f32_16x16x32_bf16) and FP32 accumulation, FP32 master weights and
// Adam on the device.  It maps circuit parameters to the coefficients of an
// initial LOBPCG block (no reference counterpart; it feeds the X0 argument of
// the LOBPCG solver, src/lobpcg.jl:217).
#include <algorithm>
#include <cmath>
#include <random>

#include "kfc_internal.h"

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef short s8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

__device__ inline unsigned short f2bf(float x) {
    // round to nearest even
    unsigned int u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);   // inf/nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ inline float gelu(float x) {
    const float c = 0.7978845608028654f;
    return 0.5f * x * (1.f + tanhf(c * (x + 0.044715f * x * x * x)));
}
__device__ inline float gelu_grad(float x) {
    const float c = 0.7978845608028654f;
    const float t = tanhf(c * (x + 0.044715f * x * x * x));
    return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * c * (1.f + 3.f * 0.044715f * x * x);
}

enum Epi { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2, EPI_GELU_GRAD = 3 };

// C[M x N] = op(A) op(B), row-major FP32 in global memory, BF16 operands on MFMA.
// TA: A is stored K x M (use A^T); TB: B is stored N x K (use B^T).
// Epilogue: EPI_BIAS_GELU writes Z = acc + bias to C and gelu(Z) to C2;
// EPI_BIAS writes acc + bias; EPI_GELU_GRAD writes acc * gelu'(Zsaved).
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_bf16(const float* __restrict__ A, const float* __restrict__ B,
                                                   float* __restrict__ C, float* __restrict__ C2, int M, int N, int K,
                                                   const float* __restrict__ bias, const float* __restrict__ Zs, int epi) {
    __shared__ unsigned short sA[64][33];   // [m][k]
    __shared__ unsigned short sB[64][33];   // [n][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;   // this wave's 32 x 32 quarter
    f4v acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f4v{0, 0, 0, 0};
    const int r16 = lane & 15, kq = (lane >> 4) * 4, k8 = (lane >> 4) * 8;
    for (int k0 = 0; k0 < K; k0 += 32) {
        // stage 64 x 32 of A and B as BF16
        for (int e = tid; e < 64 * 32; e += 256) {
            const int mm = e / 32, kk = e % 32;
            const int gm = m0 + mm, gk = k0 + kk;
            float av = 0.f;
            if (gm < M && gk < K) av = TA ? A[(int64_t)gk * M + gm] : A[(int64_t)gm * K + gk];
            sA[mm][kk] = f2bf(av);
            const int gn = n0 + mm;
            float bv = 0.f;
            if (gn < N && gk < K) bv = TB ? B[(int64_t)gn * K + gk] : B[(int64_t)gk * N + gn];
            sB[mm][kk] = f2bf(bv);
        }
        __syncthreads();
        // the whole 32-wide K tile in one 16 x 16 x 32 step: lane group g = lane / 16
        // supplies k = 8 g .. 8 g + 7 of both operands (the same k in the same slot of
        // A and B, so the products pair up whatever order the unit sums them in)
        s8v a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int row = wm + i * 16 + r16;
#pragma unroll
            for (int j = 0; j < 8; ++j) a[i][j] = short(sA[row][k8 + j]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int coln = wn + i * 16 + r16;
#pragma unroll
            for (int j = 0; j < 8; ++j) b[i][j] = short(sB[coln][k8 + j]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a[i]),
                                                                    __builtin_bit_cast(bf8v, b[j]), acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gm = m0 + wm + i * 16 + kq + q;
                const int gn = n0 + wn + j * 16 + r16;
                if (gm >= M || gn >= N) continue;
                float v = acc[i][j][q];
                const int64_t idx = (int64_t)gm * N + gn;
                if (epi == EPI_BIAS_GELU) {
                    v += bias[gn];
                    C[idx] = v;
                    C2[idx] = gelu(v);
                } else if (epi == EPI_BIAS) {
                    C[idx] = v + bias[gn];
                } else if (epi == EPI_GELU_GRAD) {
                    C[idx] = v * gelu_grad(Zs[idx]);
                } else {
                    C[idx] = v;
                }
            }
}

__global__ void k_colsum(const float* __restrict__ X, int M, int N, float* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; ++m) s += X[(int64_t)m * N + n];
    out[n] = s;
}

// dY = 2 (Y - T) / (M N) and the per-row squared error (loss partials)
__global__ void k_mse_grad(const float* __restrict__ Y, const float* __restrict__ T, float* __restrict__ dY,
                           float* __restrict__ sq, int64_t n, float scale) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        dY[i] = 2.f * d * scale;
        sq[i] = d * d;
    }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2,
                       int64_t n, float lr, float b1, float b2, float c1, float c2, float eps) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.f - b1) * gi;
        const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    }
}

__global__ void k_gather_rows(const float* __restrict__ src, const int* __restrict__ idx, int rows, int cols,
                              float* __restrict__ dst) {
    const int64_t n = (int64_t)rows * cols;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = int(i / cols), c = int(i % cols);
        dst[i] = src[(int64_t)idx[r] * cols + c];
    }
}

unsigned g1(int64_t n) { return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384))); }

}  // namespace

struct kfc_mlp {
    kfc_ctx* ctx = nullptr;
    std::vector<int> dims;        // layer widths, dims[0] inputs ... dims[L] outputs
    int64_t nparams = 0;
    std::vector<int64_t> woff, boff;
    DevBuf P, G, M1, M2;          // parameters, gradients, Adam moments (FP32)
    DevBuf act;                   // activations workspace
    int64_t act_batch = 0;
    int64_t adam_t = 0;
    std::vector<float*> Z, A;     // per layer pre-activations and activations (A[0] = input)
    DevBuf dbuf1, dbuf2, lossb, Xd, Td, idxb;
};

static int mlp_workspace(kfc_mlp* m, int64_t batch) {
    if (batch <= m->act_batch) return KFC_OK;
    const int L = int(m->dims.size()) - 1;
    int64_t total = 0;
    for (int l = 0; l <= L; ++l) total += 2 * batch * m->dims[l];
    KFC_HIP(m->ctx, m->act.ensure(sizeof(float) * total));
    int64_t maxw = *std::max_element(m->dims.begin(), m->dims.end());
    KFC_HIP(m->ctx, m->dbuf1.ensure(sizeof(float) * batch * maxw));
    KFC_HIP(m->ctx, m->dbuf2.ensure(sizeof(float) * batch * maxw));
    KFC_HIP(m->ctx, m->lossb.ensure(sizeof(float) * batch * m->dims[L]));
    m->Z.assign(L + 1, nullptr);
    m->A.assign(L + 1, nullptr);
    float* p = m->act.as<float>();
    for (int l = 0; l <= L; ++l) {
        m->A[l] = p;
        p += batch * m->dims[l];
        m->Z[l] = p;
        p += batch * m->dims[l];
    }
    m->act_batch = batch;
    return KFC_OK;
}

template <bool TA, bool TB>
static void gemm(hipStream_t st, const float* A, const float* B, float* C, float* C2, int M, int N, int K,
                 const float* bias, const float* Zs, int epi) {
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    hipLaunchKernelGGL((k_gemm_bf16<TA, TB>), grid, dim3(256), 0, st, A, B, C, C2, M, N, K, bias, Zs, epi);
}

// forward on a device input A[0] (batch x dims[0]); output in Z[L]
static int mlp_fwd(kfc_mlp* m, int64_t batch) {
    const int L = int(m->dims.size()) - 1;
    float* P = m->P.as<float>();
    hipStream_t st = m->ctx->stream;
    for (int l = 1; l <= L; ++l) {
        const float* W = P + m->woff[l - 1];   // dims[l] x dims[l-1], row-major
        const float* b = P + m->boff[l - 1];
        if (l < L)
            gemm<false, true>(st, m->A[l - 1], W, m->Z[l], m->A[l], int(batch), m->dims[l], m->dims[l - 1], b, nullptr,
                              EPI_BIAS_GELU);
        else
            gemm<false, true>(st, m->A[l - 1], W, m->Z[l], nullptr, int(batch), m->dims[l], m->dims[l - 1], b, nullptr,
                              EPI_BIAS);
    }
    KFC_HIP(m->ctx, hipGetLastError());
    return KFC_OK;
}

// backward from dY (batch x dims[L]) on the device into G
static int mlp_bwd(kfc_mlp* m, const float* dY, int64_t batch) {
    const int L = int(m->dims.size()) - 1;
    float* P = m->P.as<float>();
    float* G = m->G.as<float>();
    hipStream_t st = m->ctx->stream;
    const float* dZ = dY;
    float* bufs[2] = {m->dbuf1.as<float>(), m->dbuf2.as<float>()};
    for (int l = L; l >= 1; --l) {
        const int out = m->dims[l], in = m->dims[l - 1];
        // dW = dZ^T A[l-1]  (out x in)
        gemm<true, false>(st, dZ, m->A[l - 1], G + m->woff[l - 1], nullptr, out, in, int(batch), nullptr, nullptr, EPI_STORE);
        hipLaunchKernelGGL(k_colsum, dim3((out + 255) / 256), dim3(256), 0, st, dZ, int(batch), out, G + m->boff[l - 1]);
        if (l > 1) {
            // dA = dZ W (batch x in), then dZ_prev = dA * gelu'(Z[l-1])
            float* dst = bufs[l & 1];
            gemm<false, false>(st, dZ, P + m->woff[l - 1], dst, nullptr, int(batch), in, out, nullptr, m->Z[l - 1],
                               EPI_GELU_GRAD);
            dZ = dst;
        }
    }
    KFC_HIP(m->ctx, hipGetLastError());
    return KFC_OK;
}

extern "C" {

int kfc_mlp_create(kfc_ctx* ctx, const int32_t* layer_dims, int32_t nlayers, uint64_t seed, kfc_mlp** out) {
    if (!ctx || !layer_dims || !out || nlayers < 2) return kfc_fail(ctx, KFC_EARG, "kfc_mlp_create: bad arguments");
    auto m = new kfc_mlp();
    m->ctx = ctx;
    m->dims.assign(layer_dims, layer_dims + nlayers);
    for (int d : m->dims)
        if (d < 1 || d > 8192) {
            delete m;
            return kfc_fail(ctx, KFC_EARG, "kfc_mlp_create: layer widths must lie in 1:8192");
        }
    int64_t off = 0;
    for (int l = 1; l < nlayers; ++l) {
        m->woff.push_back(off);
        off += int64_t(m->dims[l]) * m->dims[l - 1];
        m->boff.push_back(off);
        off += m->dims[l];
    }
    m->nparams = off;
    // He-style init, deterministic in seed
    std::vector<float> h(off, 0.f);
    std::mt19937_64 g(seed);
    for (int l = 1; l < nlayers; ++l) {
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / m->dims[l - 1]));
        for (int64_t i = 0; i < int64_t(m->dims[l]) * m->dims[l - 1]; ++i) h[m->woff[l - 1] + i] = nd(g);
    }
    hipError_t e = m->P.ensure(sizeof(float) * off);
    if (e == hipSuccess) e = m->G.ensure(sizeof(float) * off);
    if (e == hipSuccess) e = m->M1.ensure(sizeof(float) * off);
    if (e == hipSuccess) e = m->M2.ensure(sizeof(float) * off);
    if (e == hipSuccess) e = hipMemcpy(m->P.p, h.data(), sizeof(float) * off, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->M1.p, 0, sizeof(float) * off);
    if (e == hipSuccess) e = hipMemset(m->M2.p, 0, sizeof(float) * off);
    if (e != hipSuccess) {
        delete m;
        return kfc_fail(ctx, KFC_ENOMEM, std::string("kfc_mlp_create: ") + hipGetErrorString(e));
    }
    *out = m;
    return KFC_OK;
}

int kfc_mlp_destroy(kfc_mlp* m) {
    if (m) {
        (void)hipStreamSynchronize(m->ctx->stream);
        delete m;
    }
    return KFC_OK;
}

int64_t kfc_mlp_num_params(const kfc_mlp* m) { return m ? m->nparams : 0; }

int kfc_mlp_set_params(kfc_mlp* m, const float* p) {
    if (!m || !p) return kfc_fail(m ? m->ctx : nullptr, KFC_EARG, "kfc_mlp_set_params: null argument");
    KFC_HIP(m->ctx, hipMemcpy(m->P.p, p, sizeof(float) * m->nparams, hipMemcpyHostToDevice));
    KFC_HIP(m->ctx, hipMemset(m->M1.p, 0, sizeof(float) * m->nparams));
    KFC_HIP(m->ctx, hipMemset(m->M2.p, 0, sizeof(float) * m->nparams));
    m->adam_t = 0;
    return KFC_OK;
}

int kfc_mlp_get_params(kfc_mlp* m, float* p) {
    if (!m || !p) return kfc_fail(m ? m->ctx : nullptr, KFC_EARG, "kfc_mlp_get_params: null argument");
    KFC_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
    KFC_HIP(m->ctx, hipMemcpy(p, m->P.p, sizeof(float) * m->nparams, hipMemcpyDeviceToHost));
    return KFC_OK;
}

int kfc_mlp_forward(kfc_mlp* m, const float* X, int64_t batch, float* Y) {
    if (!m || !X || !Y || batch < 1) return kfc_fail(m ? m->ctx : nullptr, KFC_EARG, "kfc_mlp_forward: bad arguments");
    int rc = mlp_workspace(m, batch);
    if (rc != KFC_OK) return rc;
    const int L = int(m->dims.size()) - 1;
    KFC_HIP(m->ctx, hipMemcpyAsync(m->A[0], X, sizeof(float) * batch * m->dims[0], hipMemcpyHostToDevice, m->ctx->stream));
    if ((rc = mlp_fwd(m, batch)) != KFC_OK) return rc;
    KFC_HIP(m->ctx, hipMemcpyAsync(Y, m->Z[L], sizeof(float) * batch * m->dims[L], hipMemcpyDeviceToHost, m->ctx->stream));
    KFC_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
    return KFC_OK;
}

int kfc_mlp_backward(kfc_mlp* m, const float* X, const float* dY, int64_t batch, float* grads) {
    if (!m || !X || !dY || !grads || batch < 1)
        return kfc_fail(m ? m->ctx : nullptr, KFC_EARG, "kfc_mlp_backward: bad arguments");
    int rc = mlp_workspace(m, batch);
    if (rc != KFC_OK) return rc;
    const int L = int(m->dims.size()) - 1;
    KFC_HIP(m->ctx, hipMemcpyAsync(m->A[0], X, sizeof(float) * batch * m->dims[0], hipMemcpyHostToDevice, m->ctx->stream));
    if ((rc = mlp_fwd(m, batch)) != KFC_OK) return rc;
    KFC_HIP(m->ctx, m->Td.ensure(sizeof(float) * batch * m->dims[L]));
    KFC_HIP(m->ctx, hipMemcpyAsync(m->Td.p, dY, sizeof(float) * batch * m->dims[L], hipMemcpyHostToDevice, m->ctx->stream));
    if ((rc = mlp_bwd(m, m->Td.as<float>(), batch)) != KFC_OK) return rc;
    KFC_HIP(m->ctx, hipMemcpyAsync(grads, m->G.p, sizeof(float) * m->nparams, hipMemcpyDeviceToHost, m->ctx->stream));
    KFC_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
    return KFC_OK;
}

int kfc_mlp_train(kfc_mlp* m, const float* X, const float* T, int64_t n, int64_t batch, int64_t nsteps, float lr,
                  uint64_t seed, float* last_loss) {
    if (!m || !X || !T || n < 1 || batch < 1 || nsteps < 0)
        return kfc_fail(m ? m->ctx : nullptr, KFC_EARG, "kfc_mlp_train: bad arguments");
    batch = std::min(batch, n);
    int rc = mlp_workspace(m, batch);
    if (rc != KFC_OK) return rc;
    const int L = int(m->dims.size()) - 1;
    const int din = m->dims[0], dout = m->dims[L];
    hipStream_t st = m->ctx->stream;
    KFC_HIP(m->ctx, m->Xd.ensure(sizeof(float) * n * din));
    DevBuf Tall;
    KFC_HIP(m->ctx, Tall.ensure(sizeof(float) * n * dout));
    KFC_HIP(m->ctx, m->Td.ensure(sizeof(float) * batch * dout));
    KFC_HIP(m->ctx, m->idxb.ensure(sizeof(int) * batch));
    KFC_HIP(m->ctx, hipMemcpy(m->Xd.p, X, sizeof(float) * n * din, hipMemcpyHostToDevice));
    KFC_HIP(m->ctx, hipMemcpy(Tall.p, T, sizeof(float) * n * dout, hipMemcpyHostToDevice));
    std::mt19937_64 g(seed);
    std::vector<int> perm(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = int(i);
    int64_t pos = n;
    std::vector<int> idx(batch);
    DevBuf dYb;
    KFC_HIP(m->ctx, dYb.ensure(sizeof(float) * batch * dout));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int64_t s = 0; s < nsteps; ++s) {
        for (int64_t i = 0; i < batch; ++i) {
            if (pos >= n) {
                std::shuffle(perm.begin(), perm.end(), g);
                pos = 0;
            }
            idx[i] = perm[pos++];
        }
        KFC_HIP(m->ctx, hipMemcpyAsync(m->idxb.p, idx.data(), sizeof(int) * batch, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_gather_rows, dim3(g1(batch * din)), dim3(256), 0, st, m->Xd.as<float>(), m->idxb.as<int>(),
                           int(batch), din, m->A[0]);
        hipLaunchKernelGGL(k_gather_rows, dim3(g1(batch * dout)), dim3(256), 0, st, Tall.as<float>(), m->idxb.as<int>(),
                           int(batch), dout, m->Td.as<float>());
        if ((rc = mlp_fwd(m, batch)) != KFC_OK) return rc;
        hipLaunchKernelGGL(k_mse_grad, dim3(g1(batch * dout)), dim3(256), 0, st, m->Z[L], m->Td.as<float>(),
                           dYb.as<float>(), m->lossb.as<float>(), batch * dout, 1.f / float(batch * dout));
        if ((rc = mlp_bwd(m, dYb.as<float>(), batch)) != KFC_OK) return rc;
        ++m->adam_t;
        const float c1 = 1.f - std::pow(b1, float(m->adam_t)), c2 = 1.f - std::pow(b2, float(m->adam_t));
        hipLaunchKernelGGL(k_adam, dim3(g1(m->nparams)), dim3(256), 0, st, m->P.as<float>(), m->G.as<float>(),
                           m->M1.as<float>(), m->M2.as<float>(), m->nparams, lr, b1, b2, c1, c2, eps);
        KFC_HIP(m->ctx, hipGetLastError());
        // the host index vector is reused next step: wait for its copy
        KFC_HIP(m->ctx, hipStreamSynchronize(st));
    }
    if (last_loss) {
        std::vector<float> sq(batch * dout);
        KFC_HIP(m->ctx, hipMemcpy(sq.data(), m->lossb.p, sizeof(float) * sq.size(), hipMemcpyDeviceToHost));
        double s = 0;
        for (float v : sq) s += v;
        *last_loss = float(s / double(sq.size()));
    }
    return KFC_OK;
}

}  // extern "C"
