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
    // this chunk's partial of the p x q product, column-major
    A* out = part + (int64_t)chunk * p * q;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gi = pt * 64 + ti * 4 + i, gj = qt * 64 + tj * 4 + j;
            if (gi < p && gj < q) out[gi + (int64_t)gj * p] = acc[i][j];
        }
}

// G[e] = sum over chunks of part[c][e]: one block per entry, threads stride the chunks
template <class A>
__global__ __launch_bounds__(256) void k_gram_reduce(const A* __restrict__ part, int nchunk, int64_t pq,
                                                     A* __restrict__ G) {
    __shared__ A red[256];
    const int64_t e = blockIdx.x;
    A s = acc_zero<A>();
    for (int c = threadIdx.x; c < nchunk; c += 256) s = s_add(s, part[(int64_t)c * pq + e]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] = s_add(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) G[e] = red[0];
}

// Gram of real blocks on the matrix cores: G = A^T B for p, q <= 64.  Each
// 64-row slab of A and B is staged in LDS as FP64 (one coalesced 512-byte load
// per column and wave) and contracted with v_mfma_f64_16x16x4f64 on TP x TQ
// output tiles of 16 x 16; the four waves take alternate 4-row k-steps of the
// slab and their tiles are summed in LDS at the end.  FP64 accumulation for
// FP32 blocks too.
typedef double gd4 __attribute__((ext_vector_type(4)));

// ntq > 0 (TP = TQ = 4): a Gram wider than 64 in 64 x 64 tiles, tile blockIdx.y = (ti,
// tj) with ti = y / ntq, tj = y % ntq; block x of tile y writes its partial tile
// (column-major, leading dimension the tile's rows) at part + (x * gridDim.y + y) * 4096.
template <class S, int TP, int TQ>
__global__ __launch_bounds__(256) void k_gram_mfma(const S* __restrict__ Am, int64_t lda, int p,
                                                   const S* __restrict__ Bm, int64_t ldb, int q, int64_t N,
                                                   int64_t rows_per, double* __restrict__ part, int ntq) {
    if (ntq > 0) {
        const int ti = int(blockIdx.y) / ntq, tj = int(blockIdx.y) % ntq;
        Am += int64_t(64) * ti * lda;
        Bm += int64_t(64) * tj * ldb;
        p = min(64, p - 64 * ti);
        q = min(64, q - 64 * tj);
    }
    constexpr int LDR = 65;   // odd row stride (in doubles) against bank conflicts
    __shared__ double smem[(TP + TQ) * 16 * LDR];
    double* sA = smem;
    double* sB = smem + TP * 16 * LDR;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t r0 = blockIdx.x * rows_per, r1 = min(N, r0 + rows_per);
    gd4 acc[TP][TQ];
#pragma unroll
    for (int i = 0; i < TP; ++i)
#pragma unroll
        for (int j = 0; j < TQ; ++j) acc[i][j] = gd4{0, 0, 0, 0};
    // the next slab is fetched into registers while the current one is contracted
    double pa[TP * 4], pb[TQ * 4];
    auto fetch = [&](int64_t rb) {
        const int64_t row = rb + lane;
        const bool rok = row < r1;
#pragma unroll
        for (int j = 0; j < TP * 4; ++j) {
            const int c = wv + 4 * j;
            pa[j] = (rok && c < p) ? double(Am[row + c * lda]) : 0.0;
        }
#pragma unroll
        for (int j = 0; j < TQ * 4; ++j) {
            const int c = wv + 4 * j;
            pb[j] = (rok && c < q) ? double(Bm[row + c * ldb]) : 0.0;
        }
    };
    if (r0 < r1) fetch(r0);
    for (int64_t rb = r0; rb < r1; rb += 64) {
        __syncthreads();   // the previous slab has been consumed
#pragma unroll
        for (int j = 0; j < TP * 4; ++j) sA[(wv + 4 * j) * LDR + lane] = pa[j];
#pragma unroll
        for (int j = 0; j < TQ * 4; ++j) sB[(wv + 4 * j) * LDR + lane] = pb[j];
        __syncthreads();
        if (rb + 64 < r1) fetch(rb + 64);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int rr = 4 * (wv + 4 * u) + (lane >> 4);   // slab row of this lane's k index
            double a[TP], b[TQ];
#pragma unroll
            for (int i = 0; i < TP; ++i) a[i] = sA[(16 * i + (lane & 15)) * LDR + rr];
#pragma unroll
            for (int j = 0; j < TQ; ++j) b[j] = sB[(16 * j + (lane & 15)) * LDR + rr];
#pragma unroll
            for (int i = 0; i < TP; ++i)
#pragma unroll
                for (int j = 0; j < TQ; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
    // sum the four waves' tiles in LDS: red[tile][row][col]
    double* red = smem;
    __syncthreads();
    for (int w = 0; w < 4; ++w) {
        if (wv == w) {
#pragma unroll
            for (int i = 0; i < TP; ++i)
#pragma unroll
                for (int j = 0; j < TQ; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        // f64 accumulator layout: register r of lane l holds row l/16 + 4r, column l%16
                        const int e = (i * TQ + j) * 256 + ((lane >> 4) + 4 * r) * 16 + (lane & 15);
                        red[e] = (w == 0 ? 0.0 : red[e]) + acc[i][j][r];
                    }
        }
        __syncthreads();
    }
    double* out = ntq > 0 ? part + ((int64_t)blockIdx.x * gridDim.y + blockIdx.y) * 4096
                          : part + (int64_t)blockIdx.x * p * q;
    for (int e = tid; e < TP * TQ * 256; e += 256) {
        const int tile = e >> 8, rc = e & 255, i = tile / TQ, j = tile % TQ;
        const int gi = 16 * i + (rc >> 4), gj = 16 * j + (rc & 15);
        if (gi < p && gj < q) out[gi + (int64_t)gj * p] = red[e];
    }
}

template <class S, int TP, int TQ>
void launch_gram_mfma(int nchunk, hipStream_t st, const S* A, int64_t lda, int p, const S* B, int64_t ldb, int q,
                      int64_t N, int64_t rows_per, double* part) {
    hipLaunchKernelGGL((k_gram_mfma<S, TP, TQ>), dim3(nchunk), dim3(256), 0, st, A, lda, p, B, ldb, q, N, rows_per,
                       part, 0);
}

template <class S, int TP>
void gram_mfma_q(int tq, int nchunk, hipStream_t st, const S* A, int64_t lda, int p, const S* B, int64_t ldb, int q,
                 int64_t N, int64_t rows_per, double* part) {
    switch (tq) {
        case 1: launch_gram_mfma<S, TP, 1>(nchunk, st, A, lda, p, B, ldb, q, N, rows_per, part); break;
        case 2: launch_gram_mfma<S, TP, 2>(nchunk, st, A, lda, p, B, ldb, q, N, rows_per, part); break;
        case 3: launch_gram_mfma<S, TP, 3>(nchunk, st, A, lda, p, B, ldb, q, N, rows_per, part); break;
        default: launch_gram_mfma<S, TP, 4>(nchunk, st, A, lda, p, B, ldb, q, N, rows_per, part); break;
    }
}

template <class S>
void gram_mfma(int nchunk, hipStream_t st, const S* A, int64_t lda, int p, const S* B, int64_t ldb, int q, int64_t N,
               int64_t rows_per, double* part) {
    const int tp = (p + 15) / 16, tq = (q + 15) / 16;
    switch (tp) {
        case 1: gram_mfma_q<S, 1>(tq, nchunk, st, A, lda, p, B, ldb, q, N, rows_per, part); break;
        case 2: gram_mfma_q<S, 2>(tq, nchunk, st, A, lda, p, B, ldb, q, N, rows_per, part); break;
        case 3: gram_mfma_q<S, 3>(tq, nchunk, st, A, lda, p, B, ldb, q, N, rows_per, part); break;
        default: gram_mfma_q<S, 4>(tq, nchunk, st, A, lda, p, B, ldb, q, N, rows_per, part); break;
    }
}

// Tall-skinny product on the matrix cores, with an optional Gram of the result:
// Y[:, 0:q] = X[:, 0:p] M for real blocks, p <= 80, q <= 64.  A block stages a
// 64-row slab of X in LDS (as FP64), each wave forms the 16 x q rows of Y of its
// quarter as Y^T = M^T X^T with v_mfma_f64_16x16x4f64 (so each store writes 16
// consecutive rows of a column), and with GR = 1, 2 or 3 also contracts the slab's
// rows into G = X[:, 0:pz]^T Y, Y^T Y or [X[:, 0:pz] Y]^T Y (q <= 16) while the slab
// is in LDS.  Y may be columns of X: a slab is read completely before any of its
// rows is written.  GR = 3 places the slab's Y rows (as stored) in columns
// pz..pz+q-1 of the staged slab (pz + q <= p): each wave reads and then overwrites
// only its own 16 rows there, so no other wave's product sees them.  GR = 5 is
// GR = 3 without storing Y (only its Gram leaves the kernel).  GR = 6 chains two
// products on the slab: Z = X M (q columns, rounded to S) replaces columns
// pz..pz+q-1 as in GR = 3, then Y = [X with Z] M2 (q2 <= 16 columns) is stored with
// its Gram Y^T Y; Z is the very value GR = 5 forms from the same X and M (same
// products in the same order), so a GR = 5 Gram describes the Z of a later GR = 6.
constexpr int TS_LDR = 65;

template <class S, int TQ, int GR, int TPZ, int NPF>
__global__ __launch_bounds__(256) void k_tsmm(const S* __restrict__ X, int64_t ldx, int p, int ppad,
                                              const double* __restrict__ M, int q, S* Y, int64_t ldy, int64_t N,
                                              int64_t rows_per, int pz, double* __restrict__ part,
                                              const double* __restrict__ M2, int q2) {
    constexpr int QP = 16 * TQ;
    static_assert(GR != 6 || TQ == 1, "the chained product has one 16-column tile per stage");
    extern __shared__ double sm[];
    double* sX = sm;                      // ppad columns x 64 rows (stride TS_LDR)
    double* sM = sX + ppad * TS_LDR;      // ppad x QP, row k at sM[k * QP]
    double* sY = sM + ppad * QP;          // 16 columns x 64 rows (GR != 0)
    double* sM2 = sY + 16 * TS_LDR;       // ppad x 16 (GR = 6)
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, lo = lane & 15, hi = lane >> 4;
    const int qy = GR == 6 ? q2 : q;      // columns of Y
    for (int e = tid; e < ppad * QP; e += 256) {
        const int k = e / QP, j = e - k * QP;
        sM[e] = (k < p && j < q) ? M[k + (int64_t)j * p] : 0.0;
    }
    if (GR == 6) {
        for (int e = tid; e < ppad * 16; e += 256) {
            const int k = e / 16, j = e - k * 16;
            sM2[e] = (k < p && j < q2) ? M2[k + (int64_t)j * p] : 0.0;
        }
    }
    const int64_t r0 = blockIdx.x * rows_per, r1 = min(N, r0 + rows_per);
    gd4 g[TPZ];
#pragma unroll
    for (int i = 0; i < TPZ; ++i) g[i] = gd4{0, 0, 0, 0};
    // the next slab is fetched into registers while the current one is contracted
    // (NPF >= ppad / 4 columns per wave)
    double pf[NPF];
    auto fetch = [&](double* dst, int64_t rb) {
        const int64_t row = rb + lane;
        const bool rok = row < r1;
#pragma unroll
        for (int j = 0; j < NPF; ++j) {
            const int c = wv + 4 * j;
            dst[j] = (rok && c < p) ? double(X[row + c * ldx]) : 0.0;
        }
    };
    // M's operands for this lane held in registers (k step i: row 4 i + hi), the
    // same values in the same order as the LDS reads they replace
    constexpr bool HOIST = TQ * NPF <= 16;
    double mr[HOIST ? TQ : 1][HOIST ? NPF : 1];
    double m2r[GR == 6 ? NPF : 1];
    if (HOIST || GR == 6) {
        __syncthreads();   // sM, sM2 in place
        if constexpr (HOIST) {
#pragma unroll
            for (int j = 0; j < TQ; ++j)
#pragma unroll
                for (int i = 0; i < NPF; ++i) mr[j][i] = 4 * i < ppad ? sM[(4 * i + hi) * QP + 16 * j + lo] : 0.0;
        }
        if constexpr (GR == 6) {
#pragma unroll
            for (int i = 0; i < NPF; ++i) m2r[i] = 4 * i < ppad ? sM2[(4 * i + hi) * 16 + lo] : 0.0;
        }
    }
    // one 64-row slab from the registers cur (refilled with the next slab)
    auto slab = [&](int64_t rb, double* cur) {
        __syncthreads();   // the previous slab's readers are done (and sM is in place)
#pragma unroll
        for (int j = 0; j < NPF; ++j) {
            const int c = wv + 4 * j;
            if (c < ppad) sX[c * TS_LDR + lane] = cur[j];
        }
        __syncthreads();
        if (rb + 64 < r1) fetch(cur, rb + 64);
        gd4 acc[TQ];
#pragma unroll
        for (int j = 0; j < TQ; ++j) acc[j] = gd4{0, 0, 0, 0};
        if constexpr (HOIST) {
#pragma unroll
            for (int i = 0; i < NPF; ++i) {
                if (4 * i < ppad) {
                    const double b = sX[(4 * i + hi) * TS_LDR + 16 * wv + lo];
#pragma unroll
                    for (int j = 0; j < TQ; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(mr[j][i], b, acc[j], 0, 0, 0);
                }
            }
        } else {
            for (int t = 0; t < ppad; t += 4) {
                const int k = t + hi;
                const double b = sX[k * TS_LDR + 16 * wv + lo];
#pragma unroll
                for (int j = 0; j < TQ; ++j)
                    acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(sM[k * QP + 16 * j + lo], b, acc[j], 0, 0, 0);
            }
        }
        // register r of lane l holds Y[16 wv + l % 16][16 j + l / 16 + 4 r]
        const int64_t yrow = rb + 16 * wv + lo;
        if (GR == 6) {
            // Z replaces slab columns pz..pz+q-1 of this wave's rows, then Y = slab M2
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = hi + 4 * r;
                if (col < q) sX[(pz + col) * TS_LDR + 16 * wv + lo] = yrow < r1 ? double(S(acc[0][r])) : 0.0;
            }
            __syncthreads();
            gd4 a2 = gd4{0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < NPF; ++i)
                if (4 * i < ppad)
                    a2 = __builtin_amdgcn_mfma_f64_16x16x4f64(m2r[i], sX[(4 * i + hi) * TS_LDR + 16 * wv + lo], a2, 0, 0, 0);
            acc[0] = a2;
        }
#pragma unroll
        for (int j = 0; j < TQ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = 16 * j + hi + 4 * r;
                if (GR != 5 && yrow < r1 && col < qy) Y[yrow + col * ldy] = S(acc[j][r]);
                if (GR != 0 && j == 0) {
                    const double yv = (yrow < r1 && col < qy)
                                          ? ((GR == 3 || GR == 5 || GR == 6) ? double(S(acc[0][r])) : acc[0][r])
                                          : 0.0;
                    sY[col * TS_LDR + 16 * wv + lo] = yv;
                    if ((GR == 3 || GR == 5) && col < q) sX[(pz + col) * TS_LDR + 16 * wv + lo] = yv;
                }
            }
        if (GR != 0) {
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int rr = 4 * (wv + 4 * u) + hi;
                const double b = sY[lo * TS_LDR + rr];
                if (GR == 1 || GR == 3 || GR == 5) {
#pragma unroll
                    for (int z = 0; z < TPZ; ++z)
                        g[z] = __builtin_amdgcn_mfma_f64_16x16x4f64(sX[(16 * z + lo) * TS_LDR + rr], b, g[z], 0, 0, 0);
                } else {
                    g[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b, b, g[0], 0, 0, 0);
                }
            }
        }
    };
    if (r0 < r1) fetch(pf, r0);
    for (int64_t rb = r0; rb < r1; rb += 64) slab(rb, pf);
    if (GR == 0) return;
    // sum the waves' Gram tiles in LDS
    double* red = sm;
    const int gp = GR == 1 ? pz : ((GR == 3 || GR == 5) ? pz + q : qy);
    __syncthreads();
    for (int w = 0; w < 4; ++w) {
        if (wv == w) {
#pragma unroll
            for (int z = 0; z < TPZ; ++z)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int e = z * 256 + (hi + 4 * r) * 16 + lo;
                    red[e] = (w == 0 ? 0.0 : red[e]) + g[z][r];
                }
        }
        __syncthreads();
    }
    // a block's partial Gram is gp x qy (qy = q2 for the chained product)
    double* out = part + (int64_t)blockIdx.x * gp * qy;
    for (int e = tid; e < TPZ * 256; e += 256) {
        const int z = e >> 8, rc = e & 255;
        const int gi = 16 * z + (rc >> 4), gj = rc & 15;
        if (gi < gp && gj < qy) out[gi + (int64_t)gj * gp] = red[e];
    }
}

// resident (non-null): report the kernel's resident blocks per compute unit instead of
// launching it.
template <class S, int TQ, int GR, int TPZ>
void launch_tsmm(int nchunk, size_t shmem, hipStream_t st, const S* X, int64_t ldx, int p, int ppad, const double* M,
                 int q, S* Y, int64_t ldy, int64_t N, int64_t rows_per, int pz, double* part, const double* M2,
                 int q2, int* resident) {
    auto go = [&](auto kernel) {
        if (resident) {
            int b = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, 256, shmem) != hipSuccess || b < 1) b = 1;
            *resident = b;
            return;
        }
        hipLaunchKernelGGL(kernel, dim3(nchunk), dim3(256), shmem, st, X, ldx, p, ppad, M, q, Y, ldy, N, rows_per, pz,
                           part, M2, q2);
    };
    // prefetch registers sized to the block width: fewer registers, more waves
    if (ppad <= 16) go(k_tsmm<S, TQ, GR, TPZ, 4>);
    else if (ppad <= 32) go(k_tsmm<S, TQ, GR, TPZ, 8>);
    else if (ppad <= 48) go(k_tsmm<S, TQ, GR, TPZ, 12>);
    else go(k_tsmm<S, TQ, GR, TPZ, 20>);
}

template <class S>
void tsmm_dispatch(int gr, int tq, int tpz, int nchunk, size_t shmem, hipStream_t st, const S* X, int64_t ldx, int p,
                   int ppad, const double* M, int q, S* Y, int64_t ldy, int64_t N, int64_t rows_per, int pz,
                   double* part, const double* M2 = nullptr, int q2 = 0, int* resident = nullptr) {
#define KFC_TS(TQv, GRv, TPZv) \
    launch_tsmm<S, TQv, GRv, TPZv>(nchunk, shmem, st, X, ldx, p, ppad, M, q, Y, ldy, N, rows_per, pz, part, M2, q2, \
                                   resident)
    if (gr == 0) {
        switch (tq) {
            case 1: KFC_TS(1, 0, 1); break;
            case 2: KFC_TS(2, 0, 1); break;
            case 3: KFC_TS(3, 0, 1); break;
            default: KFC_TS(4, 0, 1); break;
        }
    } else if (gr == 2) {
        KFC_TS(1, 2, 1);
    } else if (gr == 6) {
        KFC_TS(1, 6, 1);
    } else if (gr == 3) {
        switch (tpz) {
            case 1: KFC_TS(1, 3, 1); break;
            case 2: KFC_TS(1, 3, 2); break;
            case 3: KFC_TS(1, 3, 3); break;
            default: KFC_TS(1, 3, 4); break;
        }
    } else if (gr == 5) {
        switch (tpz) {
            case 1: KFC_TS(1, 5, 1); break;
            case 2: KFC_TS(1, 5, 2); break;
            case 3: KFC_TS(1, 5, 3); break;
            default: KFC_TS(1, 5, 4); break;
        }
    } else {
        switch (tpz) {
            case 1: KFC_TS(1, 1, 1); break;
            case 2: KFC_TS(1, 1, 2); break;
            case 3: KFC_TS(1, 1, 3); break;
            default: KFC_TS(1, 1, 4); break;
        }
    }
#undef KFC_TS
}

// Row chunks of a k_tsmm launch: one block per resident slot (the kernel's occupancy x
// the compute units), each taking whole 64-row slabs, so that every block runs from the
// start and none waits for a second round (2048 blocks at 3 a unit left a third of the
// device idle for the last third of a pass).
template <class S>
int tsmm_chunks(const Work& w, int gr, int tq, int tpz, int ppad, size_t shmem, int64_t N, int64_t& rows_per) {
    int b = 1;
    tsmm_dispatch<S>(gr, tq, tpz, 0, shmem, nullptr, nullptr, 0, 0, ppad, nullptr, 0, nullptr, 0, 0, 0, 0, nullptr,
                     nullptr, 0, &b);
    const int64_t slabs = std::max<int64_t>(1, (N + 63) / 64);
    const int64_t nch = std::max<int64_t>(1, std::min<int64_t>(int64_t(b) * w.ctx->num_cus, slabs));
    rows_per = (slabs + nch - 1) / nch * 64;
    return int((N + rows_per - 1) / rows_per);
}

// A real Gram wider than 64 columns on either side: all its 64 x 64 tiles in one
// k_gram_mfma launch, the chunks' partial tiles summed by one reduction.
template <class S>
int gram_tiled(Work& w, const S* A, int64_t lda, int p, const S* B, int64_t ldb, int q, int64_t N,
               std::vector<cd>& G) {
    if (N <= 0) return KFC_OK;   // G is zero
    const int ntp = (p + 63) / 64, ntq = (q + 63) / 64, nt = ntp * ntq;
    // about a thousand blocks over all tiles, whole 64-row slabs each
    const int64_t slabs = std::max<int64_t>(1, (N + 63) / 64);
    const int64_t nch = std::max<int64_t>(1, std::min<int64_t>(slabs, (1024 + nt - 1) / nt));
    const int64_t rows_per = (slabs + nch - 1) / nch * 64;
    const int nchunk = int((N + rows_per - 1) / rows_per);
    KFC_HIP(w.ctx, w.part.ensure(sizeof(double) * size_t(nchunk) * nt * 4096));
    KFC_HIP(w.ctx, w.small.ensure(sizeof(double) * size_t(nt) * 4096));
    hipLaunchKernelGGL((k_gram_mfma<S, 4, 4>), dim3(nchunk, nt), dim3(256), 0, w.ctx->stream, A, lda, p, B, ldb, q, N,
                       rows_per, w.part.as<double>(), ntq);
    KFC_HIP(w.ctx, hipGetLastError());
    hipLaunchKernelGGL(k_gram_reduce<double>, dim3(unsigned(nt * 4096)), dim3(256), 0, w.ctx->stream,
                       w.part.as<double>(), nchunk, int64_t(nt) * 4096, w.small.as<double>());
    KFC_HIP(w.ctx, hipGetLastError());
    std::vector<double> h(size_t(nt) * 4096);
    KFC_HIP(w.ctx, hipMemcpyAsync(h.data(), w.small.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, w.ctx->stream));
    KFC_HIP(w.ctx, hipStreamSynchronize(w.ctx->stream));
    for (int t = 0; t < nt; ++t) {
        const int ti = t / ntq, tj = t % ntq, pt = std::min(64, p - 64 * ti), qt = std::min(64, q - 64 * tj);
        for (int j = 0; j < qt; ++j)
            for (int i = 0; i < pt; ++i)
                G[(64 * ti + i) + size_t(64 * tj + j) * p] = cd(h[size_t(t) * 4096 + i + size_t(j) * pt], 0.0);
    }
    return KFC_OK;
}

template <class S>
int blas_gram(Work& w, const S* A, int64_t lda, int p, const S* B, int64_t ldb, int q, int64_t N, std::vector<cd>& G) {
    using Acc = typename AccOf<S>::type;
    G.assign(size_t(p) * q, cd(0));
    if (p == 0 || q == 0) return KFC_OK;
    if constexpr (!IsCx<S>::value) {
        if (p > 64 || q > 64) return gram_tiled<S>(w, A, lda, p, B, ldb, q, N, G);
    }
    const int ntp = (p + 63) / 64, ntq = (q + 63) / 64;
    // enough row chunks to fill the chip several times over
    const int nchunk = nchunks_for(N, 4096, std::max(64, 2048 / (ntp * ntq)));
    const int64_t rows_per = (N + nchunk - 1) / nchunk;
    KFC_HIP(w.ctx, w.part.ensure(sizeof(Acc) * size_t(nchunk) * p * q));
    KFC_HIP(w.ctx, w.small.ensure(sizeof(Acc) * size_t(p) * q));
    if constexpr (!IsCx<S>::value) {
        if (p <= 64 && q <= 64)
            gram_mfma<S>(nchunk, w.ctx->stream, A, lda, p, B, ldb, q, N, rows_per, w.part.as<double>());
        else
            hipLaunchKernelGGL((k_gram<S, Acc>), dim3(nchunk, ntp, ntq), dim3(256), 0, w.ctx->stream, A, lda, p, B, ldb,
                               q, N, rows_per, w.part.as<Acc>());
    } else {
        hipLaunchKernelGGL((k_gram<S, Acc>), dim3(nchunk, ntp, ntq), dim3(256), 0, w.ctx->stream, A, lda, p, B, ldb, q,
                           N, rows_per, w.part.as<Acc>());
    }
    hipLaunchKernelGGL(k_gram_reduce<Acc>, dim3(unsigned(int64_t(p) * q)), dim3(256), 0, w.ctx->stream,
                       w.part.as<Acc>(), nchunk, int64_t(p) * q, w.small.as<Acc>());
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

template <class S>
int blas_tsmm(Work& w, const S* X, int64_t ldx, int p, const std::vector<cd>& M, int q, S* Y, int64_t ldy, int64_t N,
              int gram, int pz, std::vector<cd>* G) {
    const int gp = gram == 1 ? pz : ((gram == 3 || gram == 5) ? pz + q : q);   // rows of G
    if (q == 0 || N == 0) {
        if (G) G->assign(size_t(gp) * q, cd(0));
        return KFC_OK;
    }
    const int gz = gram == 1 || gram == 3 || gram == 5 ? gp : 0;   // columns the Gram takes from the slab
    bool mfma = !IsCx<S>::value && p <= 80 && q <= 64 && (gram == 0 || (q <= 16 && gz <= 64 && gz <= p));
    if (gram == 5 && !mfma) return kfc_fail(w.ctx, KFC_EUNSUPPORTED, "tsmm: Gram-only product outside its shapes");
    if (!mfma) {
        // general path: product into scratch, copy, then the Gram
        KFC_HIP(w.ctx, w.tmp.ensure(sizeof(S) * size_t(N) * q));
        int rc = blas_gemm<S>(w, X, ldx, p, M, q, w.tmp.as<S>(), N, N, 0.0);
        if (rc == KFC_OK) rc = blas_axpby<S>(w, w.tmp.as<S>(), N, Y, ldy, q, N, 1.0, 0.0);
        if (rc == KFC_OK && gram == 1) rc = blas_gram<S>(w, X, ldx, pz, Y, ldy, q, N, *G);
        if (rc == KFC_OK && gram == 2) rc = blas_gram<S>(w, Y, ldy, q, Y, ldy, q, N, *G);
        if (rc == KFC_OK && gram == 3) {
            std::vector<cd> G1, G2;
            rc = blas_gram<S>(w, X, ldx, pz, Y, ldy, q, N, G1);
            if (rc == KFC_OK) rc = blas_gram<S>(w, Y, ldy, q, Y, ldy, q, N, G2);
            if (rc == KFC_OK) {
                G->assign(size_t(gp) * q, cd(0));
                for (int j = 0; j < q; ++j) {
                    for (int i = 0; i < pz; ++i) (*G)[i + size_t(j) * gp] = G1[i + size_t(j) * pz];
                    for (int i = 0; i < q; ++i) (*G)[pz + i + size_t(j) * gp] = G2[i + size_t(j) * q];
                }
            }
        }
        return rc;
    }
    if constexpr (!IsCx<S>::value) {
        const int tq = (q + 15) / 16, ppad = (p + 3) / 4 * 4, tpz = (gz + 15) / 16;
        size_t words = size_t(ppad) * TS_LDR + size_t(ppad) * 16 * tq + (gram ? 16 * TS_LDR : 0);
        words = std::max<size_t>(words, size_t(gz > 0 ? tpz : 1) * 256);
        int64_t rows_per = 0;
        const int nchunk = tsmm_chunks<S>(w, gram, tq, tpz, ppad, words * sizeof(double), N, rows_per);
        std::vector<double> h(size_t(p) * q);
        for (size_t i = 0; i < h.size(); ++i) h[i] = M[i].real();
        KFC_HIP(w.ctx, w.small2.ensure(sizeof(double) * h.size()));
        KFC_HIP(w.ctx, hipMemcpyAsync(w.small2.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                                      w.ctx->stream));
        if (gram) {
            KFC_HIP(w.ctx, w.part.ensure(sizeof(double) * size_t(nchunk) * gp * q));
            KFC_HIP(w.ctx, w.small.ensure(sizeof(double) * size_t(gp) * q));
        }
        KFC_HIP(w.ctx, w.tsmm_begin());
        tsmm_dispatch<S>(gram, tq, tpz, nchunk, words * sizeof(double), w.ctx->stream, X, ldx, p, ppad,
                         w.small2.as<double>(), q, Y, ldy, N, rows_per, pz, gram ? w.part.as<double>() : nullptr);
        KFC_HIP(w.ctx, hipGetLastError());
        KFC_HIP(w.ctx, hipEventRecord(w.te1, w.ctx->stream));
        if (gram) {
            hipLaunchKernelGGL(k_gram_reduce<double>, dim3(unsigned(gp * q)), dim3(256), 0, w.ctx->stream,
                               w.part.as<double>(), nchunk, int64_t(gp) * q, w.small.as<double>());
            KFC_HIP(w.ctx, hipGetLastError());
            std::vector<double> hg(size_t(gp) * q);
            KFC_HIP(w.ctx, hipMemcpyAsync(hg.data(), w.small.p, hg.size() * sizeof(double), hipMemcpyDeviceToHost,
                                          w.ctx->stream));
            KFC_HIP(w.ctx, hipStreamSynchronize(w.ctx->stream));
            G->resize(hg.size());
            for (size_t i = 0; i < hg.size(); ++i) (*G)[i] = cd(hg[i], 0.0);
        } else {
            KFC_HIP(w.ctx, hipStreamSynchronize(w.ctx->stream));   // the staged M dies with this call
        }
        KFC_HIP(w.ctx, w.tsmm_end_synced(double(N) * sizeof(S) * (p + (gram == 5 ? 0 : q))));
    }
    return KFC_OK;
}

template <class S>
bool tsmm_chain_ok(int p, int pz, int q1, int q2) {
    return !IsCx<S>::value && p <= 64 && q1 >= 1 && q1 <= 16 && q2 >= 1 && q2 <= 16 && pz >= 0 && pz + q1 <= p;
}

template <class S>
int blas_tsmm_chain(Work& w, const S* X, int64_t ldx, int p, const std::vector<cd>& M1, int q1, int pz,
                    const std::vector<cd>& M2, int q2, S* Y, int64_t ldy, int64_t N, std::vector<cd>& G) {
    if (!tsmm_chain_ok<S>(p, pz, q1, q2)) return kfc_fail(w.ctx, KFC_EUNSUPPORTED, "tsmm: chained product outside its shapes");
    if (N == 0) {
        G.assign(size_t(q2) * q2, cd(0));
        return KFC_OK;
    }
    if constexpr (!IsCx<S>::value) {
        const int ppad = (p + 3) / 4 * 4;
        size_t words = size_t(ppad) * TS_LDR + size_t(ppad) * 16 + 16 * TS_LDR + size_t(ppad) * 16;
        words = std::max<size_t>(words, 256);
        int64_t rows_per = 0;
        const int nchunk = tsmm_chunks<S>(w, 6, 1, 1, ppad, words * sizeof(double), N, rows_per);
        std::vector<double> h(size_t(p) * (q1 + q2));
        for (size_t i = 0; i < size_t(p) * q1; ++i) h[i] = M1[i].real();
        for (size_t i = 0; i < size_t(p) * q2; ++i) h[size_t(p) * q1 + i] = M2[i].real();
        KFC_HIP(w.ctx, w.small2.ensure(sizeof(double) * h.size()));
        KFC_HIP(w.ctx, hipMemcpyAsync(w.small2.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                                      w.ctx->stream));
        KFC_HIP(w.ctx, w.part.ensure(sizeof(double) * size_t(nchunk) * q2 * q2));
        KFC_HIP(w.ctx, w.small.ensure(sizeof(double) * size_t(q2) * q2));
        KFC_HIP(w.ctx, w.tsmm_begin());
        tsmm_dispatch<S>(6, 1, 1, nchunk, words * sizeof(double), w.ctx->stream, X, ldx, p, ppad,
                         w.small2.as<double>(), q1, Y, ldy, N, rows_per, pz, w.part.as<double>(),
                         w.small2.as<double>() + size_t(p) * q1, q2);
        KFC_HIP(w.ctx, hipGetLastError());
        KFC_HIP(w.ctx, hipEventRecord(w.te1, w.ctx->stream));
        hipLaunchKernelGGL(k_gram_reduce<double>, dim3(unsigned(q2 * q2)), dim3(256), 0, w.ctx->stream,
                           w.part.as<double>(), nchunk, int64_t(q2) * q2, w.small.as<double>());
        KFC_HIP(w.ctx, hipGetLastError());
        std::vector<double> hg(size_t(q2) * q2);
        KFC_HIP(w.ctx, hipMemcpyAsync(hg.data(), w.small.p, hg.size() * sizeof(double), hipMemcpyDeviceToHost,
                                      w.ctx->stream));
        KFC_HIP(w.ctx, hipStreamSynchronize(w.ctx->stream));
        G.resize(hg.size());
        for (size_t i = 0; i < hg.size(); ++i) G[i] = cd(hg[i], 0.0);
        KFC_HIP(w.ctx, w.tsmm_end_synced(double(N) * sizeof(S) * (p + q2)));
    }
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

// The seed perturbation of LOBPCG (_perturbed_seed, src/lobpcg.jl:167-177): column j
// of block b plus a random vector r scaled to sqrt(eps) times the column's norm,
// r from a counter-based normal stream (seed, j, row); complex entries have
// independent parts scaled by 1/sqrt(2).  One block per column: the norms in a
// first pass, the update in a second (r regenerated).  Columns j >= ncols[b] are
// left as they are.
template <class S>
__global__ __launch_bounds__(256) void k_perturb(S* X, int64_t ldx, int64_t N, int64_t bstride,
                                                 const int* __restrict__ ncols, uint64_t seed, double rel) {
    using R = typename RealOf<S>::type;
    const int j = blockIdx.x, bi = blockIdx.y;
    if (ncols && j >= ncols[bi]) return;   // the whole block leaves before any barrier
    S* x = X + bi * bstride + (int64_t)j * ldx;
    const uint64_t ks = mix64(seed ^ mix64(uint64_t(j) * 0x100000001B3ull));
    constexpr double sc = IsCx<S>::value ? 0.7071067811865476 : 1.0;
    double nx = 0, nr = 0;
    for (int64_t i = threadIdx.x; i < N; i += 256) {
        double a, b;
        normal2(ks + uint64_t(i), a, b);
        nx += s_abs2(x[i]);
        nr += IsCx<S>::value ? 0.5 * (a * a + b * b) : a * a;
    }
    __shared__ double red[2][256];
    red[0][threadIdx.x] = nx;
    red[1][threadIdx.x] = nr;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            red[0][threadIdx.x] += red[0][threadIdx.x + h];
            red[1][threadIdx.x] += red[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    const double f = red[1][0] > 0 ? rel * sqrt(red[0][0]) / sqrt(red[1][0]) : 0.0;
    for (int64_t i = threadIdx.x; i < N; i += 256) {
        double a, b;
        normal2(ks + uint64_t(i), a, b);
        if constexpr (IsCx<S>::value) x[i] = S{R(x[i].re + f * sc * a), R(x[i].im + f * sc * b)};
        else x[i] = R(x[i] + f * a);
    }
}

template <class S>
int blas_perturb(Work& w, S* X, int64_t ldx, int p, int64_t N, int batch, int64_t bstride, const int* ncols,
                 uint64_t seed, double rel) {
    if (p <= 0 || N == 0 || batch <= 0) return KFC_OK;
    hipLaunchKernelGGL(k_perturb<S>, dim3(unsigned(p), unsigned(batch)), dim3(256), 0, w.ctx->stream, X, ldx, N,
                       bstride, ncols, seed, rel);
    KFC_HIP(w.ctx, hipGetLastError());
    return KFC_OK;
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
    template int blas_tsmm<S>(Work&, const S*, int64_t, int, const std::vector<cd>&, int, S*, int64_t, int64_t, int, \
                              int, std::vector<cd>*);                                                               \
    template bool tsmm_chain_ok<S>(int, int, int, int);                                                             \
    template int blas_tsmm_chain<S>(Work&, const S*, int64_t, int, const std::vector<cd>&, int, int,                \
                                    const std::vector<cd>&, int, S*, int64_t, int64_t, std::vector<cd>&);            \
    template int blas_colnorms<S>(Work&, const S*, int64_t, int, int64_t, std::vector<double>&);                    \
    template int blas_scale_cols<S>(Work&, S*, int64_t, int, int64_t, const std::vector<double>&);                  \
    template int blas_residual<S>(Work&, S*, int64_t, const S*, int64_t, const S*, int64_t, int, int64_t,            \
                                  const std::vector<double>&, std::vector<double>&);                                \
    template int blas_precond<S>(Work&, S*, int64_t, int, int64_t, const double*, const std::vector<double>&,       \
                                 double);                                                                           \
    template int blas_diag_expect<S>(Work&, const S*, int64_t, int, int64_t, const double*, std::vector<double>&);  \
    template int blas_axpby<S>(Work&, const S*, int64_t, S*, int64_t, int, int64_t, double, double);                \
    template int blas_randn<S>(Work&, S*, int64_t, int, int64_t, uint64_t);                                         \
    template int blas_perturb<S>(Work&, S*, int64_t, int, int64_t, int, int64_t, const int*, uint64_t, double);   \
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
