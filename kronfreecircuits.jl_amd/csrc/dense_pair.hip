// Dense kept-basis terms on gfx950 MFMA.
//
// A branch node's kept terms are products of two dense port matrices on two
// axes (p, q): H x = D .* x + sum_t c_t (A_t (x) B_t) x.  Seen as slabs
// X_s = x[l, :, m, :, r] (d_p x d_q), each group of terms on one axis pair is
//   Y_s = [D .* X_s] + sum_t c_t A_t (X_s B_t^T),
// two small GEMMs per term per slab, computed here with
// v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32 (FP32) tiles, the
// slab and the intermediate staged in LDS, the port matrices read through L1
// (they are shared by every slab).  No permutes: the reference moves a middle
// axis to the front and back around its gemm (src/apply.jl:322-337).
// Terms sharing one factor are merged on the host first
// (c1 A(x)B1 + c2 A(x)B2 = A (x) (c1 B1 + c2 B2)).
#include <algorithm>
#include <functional>
#include <map>
#include <mutex>
#include <cstdlib>
#include <string>

#include "kfc_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline f4 mfma(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// Row of the 16x16 output tile held in accumulator register i of a lane in
// quarter k4 (lane / 16); the column is lane % 16.  The f64 instruction
// interleaves rows across the quarters, the f32 one gives each quarter four
// consecutive rows.
__device__ inline int out_row(double, int k4, int i) { return k4 + 4 * i; }
__device__ inline int out_row(float, int k4, int i) { return 4 * k4 + i; }

template <class T> struct V4;
template <> struct V4<double> { using type = d4; };
template <> struct V4<float> { using type = f4; };

constexpr int MAXT = 8;   // terms per group

template <class T>
struct PairArgs {
    const T* X;
    T* Y;
    const T* D;           // full diagonal or null (accumulate into Y)
    int64_t ld;
    int dp, dq, Mp, Nq;   // sizes and their multiples of 16
    int64_t L, M, R;      // slab decomposition
    int64_t sp, sq, sr;   // strides of axis p, axis q, and the slab r index
    int64_t nslab;        // L*M*R*ncols
    int nterm;
    T c[MAXT];
    const T* A[MAXT];     // dp x dp column-major
    const T* B[MAXT];     // dq x dq column-major
    T* Z;                 // N x ncols scratch of the split form (null: the group runs whole)
};

// P: padded extent (multiple of 16) bounding dp and dq.  Slabs s0, s0 + step, ...
template <class T, int P>
__device__ inline void dense_pair_slabs(const PairArgs<T>& a, int64_t s0, int64_t step) {
    using v4 = typename V4<T>::type;
    constexpr int NS = ((P / 16) * (P / 16) + 3) / 4;   // output tiles per wave
    __shared__ T Xs[P * P];   // column-major, leading dimension P
    __shared__ T Ts[P * P];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tm = a.Mp / 16, tn = a.Nq / 16, ntile = tm * tn;
    const int r16 = lane & 15, k4 = lane >> 4;
    const int64_t perc = a.L * a.M * a.R;
    for (int64_t s = s0; s < a.nslab; s += step) {
        const int64_t col = s / perc;
        int64_t rem = s - col * perc;
        const int64_t l = rem % a.L;
        rem /= a.L;
        const int64_t m = rem % a.M;
        const int64_t r = rem / a.M;
        const int64_t base = col * a.ld + l + m * a.L * a.dp + r * a.sr;
        __syncthreads();   // the previous slab's readers of Xs and Ts are done
        for (int e = tid; e < P * P; e += 256) {
            const int i = e % P, j = e / P;
            Xs[e] = (i < a.dp && j < a.dq) ? a.X[base + i * a.sp + j * a.sq] : T(0);
        }
        v4 acc[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) acc[k] = v4{0, 0, 0, 0};
        for (int t = 0; t < a.nterm; ++t) {
            const T* __restrict__ At = a.A[t];
            const T* __restrict__ Bt = a.B[t];
            __syncthreads();   // Xs written (t == 0) / Ts free (t > 0)
            // Ts = Xs B^T : (dp x dq)(dq x dq)
            for (int tile = wave; tile < ntile; tile += 4) {
                const int ti = tile % tm, tj = tile / tm;
                v4 c = v4{0, 0, 0, 0};
                const int row = ti * 16 + r16, bc = tj * 16 + r16;
                for (int kk = 0; kk < a.Nq; kk += 4) {
                    const int k = kk + k4;
                    const T av = Xs[row + k * P];
                    const T bv = (k < a.dq && bc < a.dq) ? Bt[bc + (int64_t)k * a.dq] : T(0);
                    c = mfma(av, bv, c);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) Ts[(ti * 16 + out_row(T(0), k4, i)) + (tj * 16 + r16) * P] = c[i];
            }
            __syncthreads();
            // acc += c_t A Ts : (dp x dp)(dp x dq)
#pragma unroll
            for (int slot = 0; slot < NS; ++slot) {
                const int tile = wave + 4 * slot;
                if (tile >= ntile) break;
                const int ti = tile % tm, tj = tile / tm;
                v4 c = v4{0, 0, 0, 0};
                const int row = ti * 16 + r16, cc = tj * 16 + r16;
                for (int kk = 0; kk < a.Mp; kk += 4) {
                    const int k = kk + k4;
                    const T av = (row < a.dp && k < a.dp) ? At[row + (int64_t)k * a.dp] : T(0);
                    const T bv = Ts[k + cc * P];
                    c = mfma(av, bv, c);
                }
                acc[slot] += a.c[t] * c;
            }
        }
        // epilogue: Y_s = D .* X_s + acc   (or Y_s += acc)
#pragma unroll
        for (int slot = 0; slot < NS; ++slot) {
            const int tile = wave + 4 * slot;
            if (tile >= ntile) break;
            const int ti = tile % tm, tj = tile / tm;
            const int cc = tj * 16 + r16;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = ti * 16 + out_row(T(0), k4, i);
                if (row < a.dp && cc < a.dq) {
                    const int64_t idx = base + row * a.sp + cc * a.sq;
                    const int64_t didx = idx - col * a.ld;
                    const T y0 = a.D ? a.D[didx] * Xs[row + cc * P] : a.Y[idx];
                    a.Y[idx] = y0 + acc[slot][i];
                }
            }
        }
    }
}

// The same slabs with the port matrices staged in LDS (P <= 48): each term's B, then
// its A, is copied once per slab into a shared buffer with coalesced loads, so the
// matrix-core loops read both operands from LDS instead of issuing a global load per
// step.  Xs, Ts and the matrix buffer take 3 P^2 scalars (55 KB at P = 48, FP64).
template <class T, int P>
__device__ inline void dense_pair_slabs_lds(const PairArgs<T>& a, int64_t s0, int64_t step) {
    using v4 = typename V4<T>::type;
    constexpr int NS = ((P / 16) * (P / 16) + 3) / 4;   // output tiles per wave
    __shared__ T Xs[P * P];   // column-major, leading dimension P
    __shared__ T Ts[P * P];
    __shared__ T Ms[P * P];   // B_t, then A_t (zero padded)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tm = a.Mp / 16, tn = a.Nq / 16, ntile = tm * tn;
    const int r16 = lane & 15, k4 = lane >> 4;
    const int64_t perc = a.L * a.M * a.R;
    auto stage = [&](const T* __restrict__ Mg, int d) {   // Ms[i + k P] = M[i + k d], zero outside d x d
        for (int e = tid; e < P * P; e += 256) {
            const int i = e % P, k = e / P;
            Ms[e] = (i < d && k < d) ? Mg[i + (int64_t)k * d] : T(0);
        }
    };
    for (int64_t s = s0; s < a.nslab; s += step) {
        const int64_t col = s / perc;
        int64_t rem = s - col * perc;
        const int64_t l = rem % a.L;
        rem /= a.L;
        const int64_t m = rem % a.M;
        const int64_t r = rem / a.M;
        const int64_t base = col * a.ld + l + m * a.L * a.dp + r * a.sr;
        __syncthreads();   // the previous slab's readers of Xs, Ts and Ms are done
        for (int e = tid; e < P * P; e += 256) {
            const int i = e % P, j = e / P;
            Xs[e] = (i < a.dp && j < a.dq) ? a.X[base + i * a.sp + j * a.sq] : T(0);
        }
        v4 acc[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) acc[k] = v4{0, 0, 0, 0};
        for (int t = 0; t < a.nterm; ++t) {
            if (t > 0) __syncthreads();   // the previous term's readers of Ms and Ts are done
            stage(a.B[t], a.dq);
            __syncthreads();
            // Ts = Xs B^T : (dp x dq)(dq x dq)
            for (int tile = wave; tile < ntile; tile += 4) {
                const int ti = tile % tm, tj = tile / tm;
                v4 c = v4{0, 0, 0, 0};
                const int row = ti * 16 + r16, bc = tj * 16 + r16;
#pragma unroll 4
                for (int kk = 0; kk < a.Nq; kk += 4) {
                    const int k = kk + k4;
                    c = mfma(Xs[row + k * P], Ms[bc + k * P], c);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) Ts[(ti * 16 + out_row(T(0), k4, i)) + (tj * 16 + r16) * P] = c[i];
            }
            __syncthreads();   // Ts complete, B no longer read
            stage(a.A[t], a.dp);
            __syncthreads();
            // acc += c_t A Ts : (dp x dp)(dp x dq)
#pragma unroll
            for (int slot = 0; slot < NS; ++slot) {
                const int tile = wave + 4 * slot;
                if (tile >= ntile) break;
                const int ti = tile % tm, tj = tile / tm;
                v4 c = v4{0, 0, 0, 0};
                const int row = ti * 16 + r16, cc = tj * 16 + r16;
#pragma unroll 4
                for (int kk = 0; kk < a.Mp; kk += 4) {
                    const int k = kk + k4;
                    c = mfma(Ms[row + k * P], Ts[k + cc * P], c);
                }
                acc[slot] += a.c[t] * c;
            }
        }
        // epilogue: Y_s = D .* X_s + acc   (or Y_s += acc)
#pragma unroll
        for (int slot = 0; slot < NS; ++slot) {
            const int tile = wave + 4 * slot;
            if (tile >= ntile) break;
            const int ti = tile % tm, tj = tile / tm;
            const int cc = tj * 16 + r16;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = ti * 16 + out_row(T(0), k4, i);
                if (row < a.dp && cc < a.dq) {
                    const int64_t idx = base + row * a.sp + cc * a.sq;
                    const int64_t didx = idx - col * a.ld;
                    const T y0 = a.D ? a.D[didx] * Xs[row + cc * P] : a.Y[idx];
                    a.Y[idx] = y0 + acc[slot][i];
                }
            }
        }
    }
}

// Register-fragment variant for groups of at most NT terms (P <= 48): a block has
// P / 16 waves; in Ts = Xs B^T wave w forms tile column w, in acc += A Ts tile row
// w, so each wave needs only its own 16 rows of every B_t and A_t, held in registers
// for the whole block (loaded once, not per slab).  The next slab is fetched into
// registers while the current one is contracted, the tiles' accumulation chains are
// interleaved, and the result goes through LDS so that Y is read and written along
// the slab's memory order.
template <class T, int P, int NT>
__device__ inline void dense_pair_slabs_reg(const PairArgs<T>& a, int64_t s0, int64_t step) {
    using v4 = typename V4<T>::type;
    constexpr int NW = P / 16, NTH = 64 * NW, KS = P / 4;
    // the next slab is prefetched into registers only where they are not all taken
    // by the fragments
    constexpr bool PREF = NT == 1 || sizeof(T) == 4;
    constexpr int NPF = PREF ? (P * P + NTH - 1) / NTH : 1;
    __shared__ T Xs[P * P];   // column-major, leading dimension P
    __shared__ T Ts[P * P];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int r16 = lane & 15, k4 = lane >> 4;
    const int tm = a.Mp / 16, tn = a.Nq / 16, ksq = a.Nq / 4, ksp = a.Mp / 4;
    const int nterm = a.nterm;
    T Bf[NT][KS], Af[NT][KS];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int k = 4 * ks + k4, j = w * 16 + r16;
            Bf[t][ks] = (t < nterm && j < a.dq && k < a.dq) ? a.B[t][j + (int64_t)k * a.dq] : T(0);
            Af[t][ks] = (t < nterm && j < a.dp && k < a.dp) ? a.A[t][j + (int64_t)k * a.dp] : T(0);
        }
    const int64_t perc = a.L * a.M * a.R;
    auto base_of = [&](int64_t s, int64_t& col) {
        col = s / perc;
        int64_t rem = s - col * perc;
        const int64_t l = rem % a.L;
        rem /= a.L;
        const int64_t m = rem % a.M;
        const int64_t r = rem / a.M;
        return col * a.ld + l + m * a.L * a.dp + r * a.sr;
    };
    T pf[NPF];
    auto fetch = [&](int64_t s) {
        int64_t col;
        const int64_t base = base_of(s, col);
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const int e = tid + u * NTH, i = e % P, j = e / P;
            pf[u] = (e < P * P && i < a.dp && j < a.dq) ? a.X[base + i * a.sp + j * a.sq] : T(0);
        }
    };
    if (PREF && s0 < a.nslab) fetch(s0);
    for (int64_t s = s0; s < a.nslab; s += step) {
        int64_t col;
        const int64_t base = base_of(s, col);
        __syncthreads();   // the previous slab's readers of Xs and Ts are done
        if constexpr (PREF) {
#pragma unroll
            for (int u = 0; u < NPF; ++u) {
                const int e = tid + u * NTH;
                if (e < P * P) Xs[e] = pf[u];
            }
        } else {
            for (int e = tid; e < P * P; e += NTH) {
                const int i = e % P, j = e / P;
                Xs[e] = (i < a.dp && j < a.dq) ? a.X[base + i * a.sp + j * a.sq] : T(0);
            }
        }
        __syncthreads();
        if (PREF && s + step < a.nslab) fetch(s + step);
        v4 acc[NW];
#pragma unroll
        for (int j = 0; j < NW; ++j) acc[j] = v4{0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            if (t < nterm) {   // uniform across the block
                // Ts[:, tile column w] = Xs B_t^T
                if (w < tn) {
                    v4 c[NW];
#pragma unroll
                    for (int ti = 0; ti < NW; ++ti) c[ti] = v4{0, 0, 0, 0};
#pragma unroll
                    for (int ks = 0; ks < KS; ++ks)
                        if (ks < ksq)
#pragma unroll
                            for (int ti = 0; ti < NW; ++ti)
                                if (ti < tm) c[ti] = mfma(Xs[(ti * 16 + r16) + (4 * ks + k4) * P], Bf[t][ks], c[ti]);
#pragma unroll
                    for (int ti = 0; ti < NW; ++ti)
                        if (ti < tm)
#pragma unroll
                            for (int i = 0; i < 4; ++i)
                                Ts[(ti * 16 + out_row(T(0), k4, i)) + (w * 16 + r16) * P] = c[ti][i];
                }
                __syncthreads();
                // acc[tile row w] += c_t A_t Ts
                if (w < tm) {
                    v4 c[NW];
#pragma unroll
                    for (int tj = 0; tj < NW; ++tj) c[tj] = v4{0, 0, 0, 0};
#pragma unroll
                    for (int ks = 0; ks < KS; ++ks)
                        if (ks < ksp)
#pragma unroll
                            for (int tj = 0; tj < NW; ++tj)
                                if (tj < tn) c[tj] = mfma(Af[t][ks], Ts[(4 * ks + k4) + (tj * 16 + r16) * P], c[tj]);
#pragma unroll
                    for (int tj = 0; tj < NW; ++tj) acc[tj] += a.c[t] * c[tj];
                }
                __syncthreads();   // Ts is free again
            }
        }
        // the result through Ts, then Y along the slab's memory order
        if (w < tm)
#pragma unroll
            for (int tj = 0; tj < NW; ++tj)
                if (tj < tn)
#pragma unroll
                    for (int i = 0; i < 4; ++i) Ts[(w * 16 + out_row(T(0), k4, i)) + (tj * 16 + r16) * P] = acc[tj][i];
        __syncthreads();
        for (int e = tid; e < P * P; e += NTH) {
            const int i = e % P, j = e / P;
            if (i < a.dp && j < a.dq) {
                const int64_t idx = base + i * a.sp + j * a.sq;
                const T y0 = a.D ? a.D[idx - col * a.ld] * Xs[e] : a.Y[idx];
                a.Y[idx] = y0 + Ts[e];
            }
        }
    }
}

// Register-chained variant for groups of at most NT terms (P <= 48).  A block has
// P / 16 waves and wave w owns output tile column w of every slab:
//   T_t[:, w] = X_s (c_t B_t)^T[:, w]      (GEMM1: X_s from LDS, the wave's 16 rows of
//                                           c_t B_t in registers, loaded once)
//   Y_s[:, w] += A_t T_t[:, w]              (GEMM2: A_t from LDS)
// GEMM1's accumulator tiles are GEMM2's B operands as they stand: lane (r, k) of a
// 16 x 16 tile holds rows out_row(k, i), i < 4, of column r, which is the B-operand
// layout of a 16 x 16 x 4 step whose contraction index for that lane is out_row(k, i);
// GEMM2's A operand is read with the same index.  So the intermediate never leaves
// registers, and the waves of a block meet only to stage a slab and to write the
// result back in the slab's memory order.  The contraction loops stop at the real
// extents (ceil(d / 4) steps); only the output tiles are padded to 16.
template <class T, int P, int NT>
__device__ inline void dense_pair_slabs_chain(const PairArgs<T>& a, int64_t s0, int64_t step) {
    using v4 = typename V4<T>::type;
    constexpr int NW = P / 16, NTH = 64 * NW, KS = P / 4;
    constexpr int NPF = (P * P + NTH - 1) / NTH;
    __shared__ T Xs[P * P];       // the slab, column-major (leading dimension P), zero outside dp x dq
    __shared__ T Rs[P * P];       // the result tiles, same layout
    __shared__ T As[NT * P * P];  // A_t, column-major (leading dimension P), zero outside dp x dp
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int r16 = lane & 15, k4 = lane >> 4;
    const int tm = a.Mp / 16, tn = a.Nq / 16, ksq = (a.dq + 3) / 4;
    const int nterm = a.nterm;
    // A_t into LDS: every load issued before the stores
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        T av[NPF];
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const int e = tid + u * NTH, i = e % P, k = e / P;
            av[u] = (e < P * P && t < nterm && i < a.dp && k < a.dp) ? a.A[t][i + (int64_t)k * a.dp] : T(0);
        }
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const int e = tid + u * NTH;
            if (e < P * P) As[t * P * P + e] = av[u];
        }
    }
    // the wave's rows of c_t B_t: Bf[t][ks] = c_t B_t[w * 16 + r, 4 ks + k]
    T Bf[NT][KS];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int k = 4 * ks + k4, j = w * 16 + r16;
            Bf[t][ks] = (t < nterm && j < a.dq && k < a.dq) ? a.c[t] * a.B[t][j + (int64_t)k * a.dq] : T(0);
        }
    const int64_t perc = a.L * a.M * a.R;
    auto base_of = [&](int64_t s, int64_t& col) {
        col = s / perc;
        int64_t rem = s - col * perc;
        const int64_t l = rem % a.L;
        rem /= a.L;
        const int64_t m = rem % a.M;
        const int64_t r = rem / a.M;
        return col * a.ld + l + m * a.L * a.dp + r * a.sr;
    };
    T pf[NPF];
    auto fetch = [&](int64_t s) {
        int64_t col;
        const int64_t base = base_of(s, col);
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const int e = tid + u * NTH, i = e % P, j = e / P;
            pf[u] = (e < P * P && i < a.dp && j < a.dq) ? a.X[base + i * a.sp + j * a.sq] : T(0);
        }
    };
    if (s0 < a.nslab) fetch(s0);
    for (int64_t s = s0; s < a.nslab; s += step) {
        int64_t col;
        const int64_t base = base_of(s, col);
        __syncthreads();   // the previous slab's readers of Xs and Rs are done (and As is staged)
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const int e = tid + u * NTH;
            if (e < P * P) Xs[e] = pf[u];
        }
        __syncthreads();
        if (s + step < a.nslab) fetch(s + step);
        if (w < tn) {
            v4 acc[NW];
#pragma unroll
            for (int tp = 0; tp < NW; ++tp) acc[tp] = v4{0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                if (t < nterm) {   // uniform across the block
                    v4 Tt[NW];
#pragma unroll
                    for (int ti = 0; ti < NW; ++ti) Tt[ti] = v4{0, 0, 0, 0};
#pragma unroll
                    for (int ks = 0; ks < KS; ++ks)
                        if (ks < ksq)
#pragma unroll
                            for (int ti = 0; ti < NW; ++ti)
                                if (ti < tm) Tt[ti] = mfma(Xs[(ti * 16 + r16) + (4 * ks + k4) * P], Bf[t][ks], Tt[ti]);
                    const T* At = As + t * P * P;
#pragma unroll
                    for (int ti = 0; ti < NW; ++ti)
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            // the step's lowest contraction index; rows of T at or beyond dp are zero
                            if (ti < tm && ti * 16 + out_row(T(0), 0, i) < a.dp)
#pragma unroll
                                for (int tp = 0; tp < NW; ++tp)
                                    if (tp < tm)
                                        acc[tp] = mfma(At[(tp * 16 + r16) + (ti * 16 + out_row(T(0), k4, i)) * P],
                                                       Tt[ti][i], acc[tp]);
                        }
                }
            }
#pragma unroll
            for (int tp = 0; tp < NW; ++tp)
                if (tp < tm)
#pragma unroll
                    for (int i = 0; i < 4; ++i) Rs[(tp * 16 + out_row(T(0), k4, i)) + (w * 16 + r16) * P] = acc[tp][i];
        }
        __syncthreads();
        // the epilogue's loads (D, or Y to accumulate into) all issued before any store, so
        // the slab pays one memory round trip, not one per element
        T yv[NPF];
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const int e = tid + u * NTH, i = e % P, j = e / P;
            const int64_t idx = base + i * a.sp + j * a.sq;
            yv[u] = (e < P * P && i < a.dp && j < a.dq) ? (a.D ? a.D[idx - col * a.ld] : a.Y[idx]) : T(0);
        }
#pragma unroll
        for (int u = 0; u < NPF; ++u) {
            const int e = tid + u * NTH, i = e % P, j = e / P;
            if (e < P * P && i < a.dp && j < a.dq) {
                const int64_t idx = base + i * a.sp + j * a.sq;
                a.Y[idx] = (a.D ? yv[u] * Xs[e] : yv[u]) + Rs[e];
            }
        }
    }
}

template <class T, int P, int NT>
__global__ __launch_bounds__(64 * (P / 16)) void k_dense_pair_chain(const PairArgs<T> a) {
    dense_pair_slabs_chain<T, P, NT>(a, blockIdx.x, gridDim.x);
}

template <class T, int P, int NT>
__global__ __launch_bounds__(64 * (P / 16)) void k_dense_pair_chain_batch(const PairArgs<T>* __restrict__ tab) {
    const PairArgs<T> a = tab[blockIdx.y];
    dense_pair_slabs_chain<T, P, NT>(a, blockIdx.x, gridDim.x);
}

// Wave-per-slab variant for groups of at most NT terms with extents up to P <= 48: the
// block stages A_t and c_t B_t in LDS once, then each of its W waves works through whole
// slabs on its own (no block barrier after the staging), so a wave's loads and stores
// overlap the other waves' matrix-core work.  Per slab and per 16-wide tile tc of the
// p axis (the axis of the result's lanes):
//   U_t^T[:, tc] = X_s^T A_t^T[:, tc]        (GEMM1: X_s^T fragments in registers, loaded
//                                             once per slab; A_t from LDS)
//   Y_s^T[:, tc] += (c_t B_t) U_t^T[:, tc]   (GEMM2: c_t B_t from LDS)
// so Y_s = A_t X_s (c_t B_t)^T.  GEMM1's accumulator tiles are GEMM2's B operands as
// they stand (the same chaining as above), and the result tile holds 16 consecutive
// p-axis entries per register: with a unit-strided p axis the epilogue's loads and
// stores are whole 128-byte lines.  P / 4 contraction steps (P = 40: ten, not twelve).
// Buffer access to one slab: the descriptor's range ends where the slab's last column
// ends, so a lane of a padded tile past dq reads zero and its store is dropped, and a
// lane past dp sets bit 31 of its offset (past every range) for the same effect; no lane
// masks.  dq * sq >= (dq - 1) * sq + dp * sp puts every column index >= dq past the range.
typedef unsigned u2 __attribute__((ext_vector_type(2)));
__device__ inline __amdgpu_buffer_rsrc_t slab_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, int(bytes), 0x00020000);
}
__device__ inline double bload(double, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ inline float bload(float, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ inline void bstore(double v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, off, 0, 0);
}
__device__ inline void bstore(float v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}
// bit 31 when index >= extent
__device__ inline uint32_t past(int index, int extent) { return uint32_t(extent - 1 - index) & 0x80000000u; }

// LDS layout of a staged d x d port matrix on a P-wide side: column-major with leading
// dimension ld_of and cols_of columns.  FP64 packs it at P (a 16-row tile past P reads
// the next column's entries, which only reach result entries outside the slab, and the
// pad keeps the last one in the array); FP32 pads to whole tiles, since its contraction
// steps reach 12 rows past their first.
template <class T, int P>
constexpr int lds_ld() { return sizeof(T) == 8 ? P : 16 * ((P + 15) / 16); }
template <class T, int P>
constexpr int lds_size() { return lds_ld<T, P>() * lds_ld<T, P>(); }

// PI, PJ: compile-time extents (16, 32, 40 or 48) bounding dp and dq; every loop bound is
// a constant, so the compiler pipelines the LDS reads under the matrix-core work, and
// rows or columns past dp, dq cost only padded work: their loads read zero and their
// stores are dropped (buffer ranges above).
//
// Per slab and per 16-wide tile tp of the q axis:
//   U[:, tp] = X_s (c_t B_t)^T[:, tp]      (GEMM1: X_s fragments in registers, loaded once
//                                          per slab, 16 consecutive p-axis entries a row;
//                                          c_t B_t from LDS)
//   Y_s^T[tp, :] += U[:, tp]^T A_t^T       (GEMM2: A_t from LDS)
// GEMM1's accumulator tile, lane (r, k) holding rows out_row(k, q) of column r, is
// GEMM2's A operand (row r, contraction index out_row(k, q)) as it stands, and GEMM2's
// result holds 16 consecutive p-axis entries per register: the loads of X and of the
// epilogue and the stores all move whole 128-byte rows.
template <class T, int PI, int PJ, int NT, int W, bool DIAG>
__device__ inline void dense_pair_slabs_wave(const PairArgs<T>& a, int64_t w0, int64_t wstep) {
    using v4 = typename V4<T>::type;
    constexpr int NTI = (PI + 15) / 16, NTJ = (PJ + 15) / 16, KSJ = PJ / 4;
    constexpr int LDA = lds_ld<T, PI>(), LDB = lds_ld<T, PJ>();
    constexpr int MSA = lds_size<T, PI>(), MSB = lds_size<T, PJ>();
    constexpr int PADA = 16 * NTI - LDA + 1, PADB = 16 * NTJ - LDB + 1;
    __shared__ T As[NT * MSA + PADA];   // A_t, zero outside dp x dp
    __shared__ T Bs[NT * MSB + PADB];   // c_t B_t, zero outside dq x dq
    const int tid = threadIdx.x, lane = tid & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
    const int dp = a.dp, dq = a.dq, nterm = a.nterm;
    for (int e = tid; e < NT * MSA + PADA; e += 64 * W) {
        const int t = e / MSA, f = e - t * MSA, i = f % LDA, k = f / LDA;
        As[e] = (t < nterm && i < dp && k < dp) ? a.A[t][i + (int64_t)k * dp] : T(0);
    }
    for (int e = tid; e < NT * MSB + PADB; e += 64 * W) {
        const int t = e / MSB, f = e - t * MSB, i = f % LDB, k = f / LDB;
        Bs[e] = (t < nterm && i < dq && k < dq) ? a.c[t] * a.B[t][i + (int64_t)k * dq] : T(0);
    }
    __syncthreads();
    const int64_t perc = a.L * a.M * a.R;
    // in-slab byte offsets fit 31 bits (the launcher checks it)
    const uint32_t bsp0 = uint32_t(a.sp * sizeof(T)), bsq0 = uint32_t(a.sq * sizeof(T));
    const uint32_t extent = uint32_t((dq - 1) * bsq0 + dp * bsp0);
    const int64_t wave = w0 + __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int64_t s = wave; s < a.nslab; s += wstep) {
        const int64_t col = s / perc;
        int64_t rem = s - col * perc;
        const int64_t l = rem % a.L;
        rem /= a.L;
        const int64_t m = rem % a.M;
        const int64_t r = rem / a.M;
        const int64_t sbase = l + m * a.L * a.dp + r * a.sr;
        const __amdgpu_buffer_rsrc_t xr = slab_rsrc(a.X + col * a.ld + sbase, extent);
        const __amdgpu_buffer_rsrc_t yr = slab_rsrc(a.Y + col * a.ld + sbase, extent);
        const __amdgpu_buffer_rsrc_t dr = slab_rsrc(a.D ? a.D + sbase : a.X, extent);
        // the strides, opaque to the compiler per slab: offsets are formed where they are
        // used rather than held across the loop
        uint32_t bsp = __builtin_amdgcn_readfirstlane(bsp0), bsq = __builtin_amdgcn_readfirstlane(bsq0);
        asm volatile("" : "+s"(bsp), "+s"(bsq));
        // p-axis offset of tile ti's lane (bit 31 past dp; the q-axis part, below 2^31,
        // cannot carry it back into range)
        auto poff = [&](int ti) { return (uint32_t(16 * ti + r16) * bsp) | past(16 * ti + r16, dp); };
        // GEMM1's A operand: xa[ti][ks] = X_s[16 ti + r16, 4 ks + k4]
        T xa[NTI][KSJ];
#pragma unroll
        for (int ti = 0; ti < NTI; ++ti)
#pragma unroll
            for (int ks = 0; ks < KSJ; ++ks) xa[ti][ks] = bload(T(0), xr, poff(ti) + uint32_t(4 * ks + k4) * bsq);
#pragma unroll
        for (int tp = 0; tp < NTJ; ++tp) {
            // c_t B_t and A_t are read again from LDS for every tile: kept in registers
            // across them they would take far more per lane
            asm volatile("" ::: "memory");
            // the epilogue's HBM operand for this tile, in flight during its GEMMs: X (the
            // first group, for D .* X; D is shared by every column and stays in L2, so it
            // is read at the end) or Y (groups accumulating into it), at
            // Y_s[16 tc + r16, 16 tp + out_row(k4, q)]
            // In FP64 the entry X_s[16 tc + r16, 16 tp + k4 + 4 q] of the result's layout is
            // the GEMM1 fragment xa[tc][4 tp + q] this lane already holds, so D .* X needs
            // no second read of X (FP32 interleaves its rows differently).
            constexpr bool XREG = DIAG && sizeof(T) == 8;
            T pv[NTI][4];
            if constexpr (!XREG) {
#pragma unroll
                for (int tc = 0; tc < NTI; ++tc)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        pv[tc][q] = bload(T(0), DIAG ? xr : yr, poff(tc) + uint32_t(16 * tp + out_row(T(0), k4, q)) * bsq);
            }
            v4 acc[NTI];
#pragma unroll
            for (int tc = 0; tc < NTI; ++tc) acc[tc] = v4{0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                if (t < nterm) {   // uniform
                    const T* At = As + t * MSA;
                    const T* Bt = Bs + t * MSB;
                    // one 16-row tile of U at a time: GEMM1 into it, then its share of GEMM2
#pragma unroll
                    for (int ti = 0; ti < NTI; ++ti) {
                        v4 u = v4{0, 0, 0, 0};
#pragma unroll
                        for (int ks = 0; ks < KSJ; ++ks)
                            u = mfma(xa[ti][ks], Bt[(16 * tp + r16) + (4 * ks + k4) * LDB], u);   // c_t B_t[j, j']
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            // steps whose contraction indices all lie past PI are padding
                            if (16 * ti + out_row(T(0), 0, q) < PI)
#pragma unroll
                                for (int tc = 0; tc < NTI; ++tc)
                                    acc[tc] = mfma(u[q], At[(16 * tc + r16) + (16 * ti + out_row(T(0), k4, q)) * LDA],
                                                   acc[tc]);   // A_t[i, i']
                        }
                    }
                }
            }
            if constexpr (XREG) {
#pragma unroll
                for (int tc = 0; tc < NTI; ++tc)
#pragma unroll
                    for (int q = 0; q < 4; ++q) pv[tc][q] = (4 * tp + q < KSJ) ? xa[tc][(4 * tp + q) % KSJ] : T(0);
            }
            T dv[NTI][4];
#pragma unroll
            for (int tc = 0; tc < NTI; ++tc)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    dv[tc][q] = DIAG ? bload(T(0), dr, poff(tc) + uint32_t(16 * tp + out_row(T(0), k4, q)) * bsq) : T(1);
#pragma unroll
            for (int tc = 0; tc < NTI; ++tc)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    bstore(T(dv[tc][q] * pv[tc][q] + acc[tc][q]), yr,
                           poff(tc) + uint32_t(16 * tp + out_row(T(0), k4, q)) * bsq);
        }
    }
}

// Four-wave blocks; the register budget allows two waves per SIMD (at three the compiler
// spills, and FP64 with two terms at 40 x 40 fits only three blocks a unit in LDS).
constexpr int WAVE_W = 4;

// Blocks go to the eight XCDs round-robin, so blocks b, b + 8, ... of a row of the grid share
// one XCD and its L2.  The logical block of b numbers those classes one after another:
// consecutive slabs then run on one XCD, and a cache line shared by neighbouring slabs
// (a strided p axis interleaves them) is fetched and written back by one L2, not four.
__device__ inline int64_t xcd_block(int b, int g) {
    const int r = b % 8, k = b / 8;
    int off = 0;
    for (int c = 0; c < r; ++c) off += (g - c + 7) / 8;
    return int64_t(off) + k;
}

// DIAG: the group writes Y = D .* X + ... (the first group of an operator); otherwise it
// adds to Y
template <class T, int PI, int PJ, int NT, bool DIAG>
__global__ __launch_bounds__(64 * WAVE_W) __attribute__((amdgpu_waves_per_eu(2)))
void k_dense_pair_wave(const PairArgs<T> a) {
    const int64_t lb = xcd_block(int(blockIdx.x), int(gridDim.x));
    dense_pair_slabs_wave<T, PI, PJ, NT, WAVE_W, DIAG>(a, lb * WAVE_W, int64_t(gridDim.x) * WAVE_W);
}

template <class T, int PI, int PJ, int NT, bool DIAG>
__global__ __launch_bounds__(64 * WAVE_W) __attribute__((amdgpu_waves_per_eu(2)))
void k_dense_pair_wave_batch(const PairArgs<T>* __restrict__ tab) {
    const PairArgs<T> a = tab[blockIdx.y];
    // (gridDim.x * blockIdx.y + b) % 8 is constant over a residue class of b, so the same
    // numbering groups one XCD's blocks in every row
    const int64_t lb = xcd_block(int(blockIdx.x), int(gridDim.x));
    dense_pair_slabs_wave<T, PI, PJ, NT, WAVE_W, DIAG>(a, lb * WAVE_W, int64_t(gridDim.x) * WAVE_W);
}

// Groups whose p axis is strided (L > 1: axes before p) interleave neighbouring slabs
// element by element, so a wave that gathers one slab moves 8-byte pieces.  Here a block
// of WS waves takes WS neighbouring slabs l0 .. l0 + WS - 1 of one (m, r, column):
//   load      the block reads X[l0 .. l0 + WS - 1, i, j], WS consecutive entries per
//             (i, j) (32 bytes in FP64, 16 in FP32, with WS = 4), into one LDS tile
//             per slab (rows padded to whole tiles, zero outside dp x dq);
//   compute   wave w applies slab l0 + w with the register-chained GEMMs above, its X
//             fragments read from its tile, and writes the result back over the tile;
//   epilogue  the block forms Y = D .* X (or Y) + result in the same segments.
template <class T, int PI, int PJ, int NT, int WS>
__device__ inline void dense_pair_chunks_staged(const PairArgs<T>& a, int64_t c0, int64_t cstep) {
    using v4 = typename V4<T>::type;
    constexpr int NTI = (PI + 15) / 16, NTJ = (PJ + 15) / 16, KSJ = PJ / 4;
    constexpr int LDA = lds_ld<T, PI>(), LDB = lds_ld<T, PJ>();
    constexpr int MSA = lds_size<T, PI>(), MSB = lds_size<T, PJ>();
    constexpr int PADA = 16 * NTI - LDA + 1, PADB = 16 * NTJ - LDB + 1;
    constexpr int LDX = 16 * NTI, MSX = LDX * PJ;          // a slab's tile: LDX x PJ
    constexpr int NTH = 64 * WS, NPASS = (PI * PJ + 63) / 64;   // 64 (i, j) pairs a pass
    // passes a batch: every load of a phase in flight at once up to 24 (one round trip
    // each way at 40 x 40)
    constexpr int NB = NPASS < 24 ? NPASS : 24;
    __shared__ T As[NT * MSA + PADA];
    __shared__ T Bs[NT * MSB + PADB];
    __shared__ T Xs[WS * MSX];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int r16 = lane & 15, k4 = lane >> 4;
    const int dp = a.dp, dq = a.dq, nterm = a.nterm;
    for (int e = tid; e < NT * MSA + PADA; e += NTH) {
        const int t = e / MSA, f = e - t * MSA, i = f % LDA, k = f / LDA;
        As[e] = (t < nterm && i < dp && k < dp) ? a.A[t][i + (int64_t)k * dp] : T(0);
    }
    for (int e = tid; e < NT * MSB + PADB; e += NTH) {
        const int t = e / MSB, f = e - t * MSB, i = f % LDB, k = f / LDB;
        Bs[e] = (t < nterm && i < dq && k < dq) ? a.c[t] * a.B[t][i + (int64_t)k * dq] : T(0);
    }
    for (int e = tid; e < WS * MSX; e += NTH) Xs[e] = T(0);   // the padding stays zero
    __syncthreads();
    // chunk counts fit 32 bits (the launcher checks N * sizeof(T) < 2^31)
    const uint32_t L = uint32_t(a.L), M = uint32_t(a.M), R = uint32_t(a.R);
    const uint32_t nlc = (L + WS - 1) / WS;                   // chunks of WS slabs along l
    const uint32_t nchunk = nlc * M * R * uint32_t(a.nslab / (a.L * a.M * a.R));
    // the cooperative phases: lane group of WS threads <-> l, pairs (i, j) over the rest
    const int wl = tid % WS, p0 = tid / WS, npair = dp * dq;
    const int di = 64 % dp, dj = 64 / dp;
    for (uint32_t c = uint32_t(c0); c < nchunk; c += uint32_t(cstep)) {
        const uint32_t lc = c % nlc, rest = c / nlc;
        const uint32_t m = rest % M, rr = (rest / M) % R, col = rest / (M * R);
        const uint32_t l0 = lc * WS;
        const int64_t cbase = int64_t(col) * a.ld + l0 + int64_t(m) * L * a.dp + int64_t(rr) * a.sr;   // slab l0's (0, 0)
        const bool lok = l0 + wl < L;
        // offsets below 2^31 (the launcher checks N * sizeof(T)); pass u of a batch covers
        // pair p0 + 64 (ub + u)
        const T* __restrict__ xc = a.X + cbase;
        T* __restrict__ yc = a.Y + cbase;
        const T* __restrict__ dc = a.D ? a.D + (cbase - col * a.ld) : nullptr;
        const uint32_t sp = uint32_t(a.sp), sq = uint32_t(a.sq);
        // load: X into the tiles
        for (int ub = 0; ub < NPASS; ub += NB) {
            T v[NB];
            int i = (p0 + 64 * ub) % dp, j = (p0 + 64 * ub) / dp;
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                v[u] = (lok && p0 + 64 * (ub + u) < npair) ? xc[wl + uint32_t(i) * sp + uint32_t(j) * sq] : T(0);
                i += di;
                j += dj;
                if (i >= dp) {
                    i -= dp;
                    ++j;
                }
            }
            i = (p0 + 64 * ub) % dp;
            j = (p0 + 64 * ub) / dp;
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                if (p0 + 64 * (ub + u) < npair) Xs[wl * MSX + i + j * LDX] = v[u];
                i += di;
                j += dj;
                if (i >= dp) {
                    i -= dp;
                    ++j;
                }
            }
        }
        __syncthreads();
        // compute: wave w, slab l0 + w
        if (l0 + w < L) {
            T* xs = Xs + w * MSX;
            T xa[NTI][KSJ];
#pragma unroll
            for (int ti = 0; ti < NTI; ++ti)
#pragma unroll
                for (int ks = 0; ks < KSJ; ++ks) xa[ti][ks] = xs[(16 * ti + r16) + (4 * ks + k4) * LDX];
            // every wave has its fragments before any result lands on its own tile
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int tp = 0; tp < NTJ; ++tp) {
                v4 acc[NTI];
#pragma unroll
                for (int tc = 0; tc < NTI; ++tc) acc[tc] = v4{0, 0, 0, 0};
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    if (t < nterm) {
                        const T* At = As + t * MSA;
                        const T* Bt = Bs + t * MSB;
#pragma unroll
                        for (int ti = 0; ti < NTI; ++ti) {
                            v4 u = v4{0, 0, 0, 0};
#pragma unroll
                            for (int ks = 0; ks < KSJ; ++ks)
                                u = mfma(xa[ti][ks], Bt[(16 * tp + r16) + (4 * ks + k4) * LDB], u);
#pragma unroll
                            for (int q = 0; q < 4; ++q) {
                                if (16 * ti + out_row(T(0), 0, q) < PI)
#pragma unroll
                                    for (int tc = 0; tc < NTI; ++tc)
                                        acc[tc] = mfma(u[q], At[(16 * tc + r16) + (16 * ti + out_row(T(0), k4, q)) * LDA],
                                                       acc[tc]);
                            }
                        }
                    }
                }
                // the result tile back over the slab's tile (entries inside dp x dq only)
#pragma unroll
                for (int tc = 0; tc < NTI; ++tc)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int i = 16 * tc + r16, j = 16 * tp + out_row(T(0), k4, q);
                        if (i < dp && j < dq) xs[i + j * LDX] = acc[tc][q];
                    }
            }
        }
        __syncthreads();
        // epilogue: Y = D .* X (or Y) + result, a batch's loads before its stores
        for (int ub = 0; ub < NPASS; ub += NB) {
            T v[NB];
            int i = (p0 + 64 * ub) % dp, j = (p0 + 64 * ub) / dp;
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                if (lok && p0 + 64 * (ub + u) < npair) {
                    const uint32_t off = wl + uint32_t(i) * sp + uint32_t(j) * sq;
                    v[u] = dc ? dc[off] * xc[off] : yc[off];
                }
                i += di;
                j += dj;
                if (i >= dp) {
                    i -= dp;
                    ++j;
                }
            }
            i = (p0 + 64 * ub) % dp;
            j = (p0 + 64 * ub) / dp;
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                if (lok && p0 + 64 * (ub + u) < npair)
                    yc[wl + uint32_t(i) * sp + uint32_t(j) * sq] = v[u] + Xs[wl * MSX + i + j * LDX];
                i += di;
                j += dj;
                if (i >= dp) {
                    i -= dp;
                    ++j;
                }
            }
        }
        __syncthreads();   // the tiles are free for the next chunk
    }
}

// slabs (and waves) a block: 4, two blocks a compute unit at 40 x 32 (8 spills FP32
// registers at the four waves per SIMD it would need)
template <class T> constexpr int staged_ws() { return 4; }
template <class T> constexpr int staged_wpe() { return 2; }

template <class T, int PI, int PJ, int NT>
__global__ __launch_bounds__(64 * staged_ws<T>()) __attribute__((amdgpu_waves_per_eu(staged_wpe<T>())))
void k_dense_pair_staged(const PairArgs<T> a) {
    dense_pair_chunks_staged<T, PI, PJ, NT, staged_ws<T>()>(a, xcd_block(int(blockIdx.x), int(gridDim.x)), gridDim.x);
}

template <class T, int PI, int PJ, int NT>
__global__ __launch_bounds__(64 * staged_ws<T>()) __attribute__((amdgpu_waves_per_eu(staged_wpe<T>())))
void k_dense_pair_staged_batch(const PairArgs<T>* __restrict__ tab) {
    const PairArgs<T> a = tab[blockIdx.y];
    dense_pair_chunks_staged<T, PI, PJ, NT, staged_ws<T>()>(a, xcd_block(int(blockIdx.x), int(gridDim.x)), gridDim.x);
}

// The split form of a one-term group whose p axis is strided (L >= 16 entries of the
// axes before it), as two mode products that each read and write whole runs of
// contiguous entries:
//   STAGE 1   Z[.., j, r] = sum_j' (c B)[j, j'] X[.., j', r]     the q axis, on runs of
//                                                              stride(q) = L dp M entries
//   STAGE 2   Y[l, i, ..] (+)= sum_i' A[i, i'] Z[l, i', ..]      the p axis, on runs of
//                                                              L entries
// Each is a product of a d x d matrix G with the d columns of an (rows x d) block (column
// stride rows, blocks rows * d apart, block columns ld apart):  out^T = G in^T on
// 16 x 16 x 4 MFMA tiles, a wave per 16 consecutive rows, G's fragments in registers
// for the whole launch.  The loads and stores of a tile move 16 consecutive entries of
// each column (a 128-byte line in FP64) instead of the slab kernels' gathers at stride
// L.  Traffic: X read and Z written, then Z, Y (or X and the diagonal) read and Y
// written; the slab kernels read X and Y and write Y once, but through 8-byte pieces.
// MODE 0: out = G in (stage 1); 1: out += G in; 2: out = D .* X + G in (the group that
// writes Y first).
// FLAT (family 5): a tile is 16 consecutive rows of the (blocks x rows) sequence of one
// block column, crossing block boundaries (no lanes idle when rows is not a multiple of
// 16), and G's fragments are read from LDS at each step instead of being held in
// registers (about half the registers, twice the waves a SIMD).
template <class T, int KS, int STAGE, int MODE, bool FLAT = false>
__device__ inline void pair_mode(const PairArgs<T>& a, int64_t w0, int64_t wstep) {
    using v4 = typename V4<T>::type;
    constexpr int NTJ = KS / 4;   // 16-wide output tiles; KS contraction steps of 4 bound d
    // tile counts and extents fit 32 bits (split_ok checks them)
    const int d = STAGE == 1 ? a.dq : a.dp;
    const uint32_t rows = uint32_t(STAGE == 1 ? a.sq : a.sp);
    const uint32_t nout = uint32_t(STAGE == 1 ? a.R : a.M * a.dq * a.R);
    const uint32_t ncol = uint32_t(a.nslab / (a.L * a.M * a.R));
    const T* __restrict__ in = STAGE == 1 ? a.X : a.Z;
    T* __restrict__ out = STAGE == 1 ? a.Z : a.Y;
    const T* __restrict__ G = STAGE == 1 ? a.B[0] : a.A[0];
    const T gc = STAGE == 1 ? a.c[0] : T(1);
    const int lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    if constexpr (FLAT) {
        constexpr int LDG = 16 * NTJ;   // G staged column-major, zero-padded to LDG x 4 KS
        __shared__ T Gs[LDG * 4 * KS];
        for (int e = threadIdx.x; e < LDG * 4 * KS; e += blockDim.x) {
            const int j = e % LDG, jp = e / LDG;
            Gs[e] = (j < d && jp < d) ? gc * G[j + (int64_t)jp * d] : T(0);
        }
        __syncthreads();
        const uint32_t nrow = rows * nout, ntile = (nrow + 15) / 16, total = ntile * ncol;
        const uint32_t wave = uint32_t(w0) + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
        for (uint32_t t = wave; t < total; t += uint32_t(wstep)) {
            const uint32_t tr = t % ntile, col = t / ntile;
            const uint32_t gr = 16 * tr + r16;   // this lane's row of the block sequence
            const bool rok = gr < nrow;
            const uint32_t o = gr / rows, row = gr - o * rows;
            const int64_t cb = int64_t(o) * rows * d + row, base = int64_t(col) * a.ld + cb;
            T b[KS];
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const int jp = 4 * ks + k4;
                b[ks] = (rok && jp < d) ? in[base + (int64_t)jp * rows] : T(0);
            }
            // MODE 1: the output's old value, loaded with the operands so that its
            // latency overlaps the products
            T pre[NTJ][4];
#pragma unroll
            for (int tj = 0; tj < NTJ; ++tj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int j = 16 * tj + out_row(T(0), k4, v);
                    const int64_t e = (int64_t)j * rows;
                    pre[tj][v] = (MODE == 1 && rok && j < d) ? out[base + e] : T(0);
                }
            v4 acc[NTJ];
#pragma unroll
            for (int tj = 0; tj < NTJ; ++tj) acc[tj] = v4{0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
                if (4 * ks < d)
#pragma unroll
                    for (int tj = 0; tj < NTJ; ++tj)
                        acc[tj] = mfma(Gs[(16 * tj + r16) + (4 * ks + k4) * LDG], b[ks], acc[tj]);
#pragma unroll
            for (int tj = 0; tj < NTJ; ++tj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int j = 16 * tj + out_row(T(0), k4, v);
                    if (rok && j < d) {
                        const int64_t e = (int64_t)j * rows;
                        T r = acc[tj][v];
                        if constexpr (MODE == 1) r += pre[tj][v];
                        if constexpr (MODE == 2) r += a.D[cb + e] * a.X[base + e];
                        out[base + e] = r;
                    }
                }
        }
        return;
    }
    // A operand of step ks for output tile tj: (c) G[16 tj + r16, 4 ks + k4]
    T g[NTJ][KS];
#pragma unroll
    for (int tj = 0; tj < NTJ; ++tj)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int j = 16 * tj + r16, jp = 4 * ks + k4;
            g[tj][ks] = (j < d && jp < d) ? gc * G[j + (int64_t)jp * d] : T(0);
        }
    const uint32_t ntr = (rows + 15) / 16, total = ntr * nout * ncol;
    const uint32_t wave = uint32_t(w0) + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    for (uint32_t t = wave; t < total; t += uint32_t(wstep)) {
        const uint32_t rt = t % ntr, rest = t / ntr, o = rest % nout, col = rest / nout;
        const int64_t cb = int64_t(o) * rows * d, base = int64_t(col) * a.ld + cb;
        const uint32_t row = 16 * rt + r16;
        const bool rok = row < rows;
        T b[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int jp = 4 * ks + k4;
            b[ks] = (rok && jp < d) ? in[base + row + (int64_t)jp * rows] : T(0);
        }
        T pre[NTJ][4];
#pragma unroll
        for (int tj = 0; tj < NTJ; ++tj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = 16 * tj + out_row(T(0), k4, v);
                const int64_t e = row + (int64_t)j * rows;
                pre[tj][v] = (MODE == 1 && rok && j < d) ? out[base + e] : T(0);
            }
        v4 acc[NTJ];
#pragma unroll
        for (int tj = 0; tj < NTJ; ++tj) acc[tj] = v4{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
            if (4 * ks < d)   // uniform: steps past d are padding
#pragma unroll
                for (int tj = 0; tj < NTJ; ++tj) acc[tj] = mfma(g[tj][ks], b[ks], acc[tj]);
        // lane (r16, k4) holds out[row, 16 tj + out_row(k4, v)]
#pragma unroll
        for (int tj = 0; tj < NTJ; ++tj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = 16 * tj + out_row(T(0), k4, v);
                if (rok && j < d) {
                    const int64_t e = row + (int64_t)j * rows;
                    T r = acc[tj][v];
                    if constexpr (MODE == 1) r += pre[tj][v];
                    if constexpr (MODE == 2) r += a.D[cb + e] * a.X[base + e];
                    out[base + e] = r;
                }
            }
    }
}

template <class T, int KS, int STAGE, int MODE, bool FLAT>
__global__ __launch_bounds__(256) void k_pair_mode(const PairArgs<T> a) {
    pair_mode<T, KS, STAGE, MODE, FLAT>(a, int64_t(blockIdx.x) * 4, int64_t(gridDim.x) * 4);
}

template <class T, int KS, int STAGE, int MODE, bool FLAT>
__global__ __launch_bounds__(256) void k_pair_mode_batch(const PairArgs<T>* __restrict__ tab) {
    const PairArgs<T> a = tab[blockIdx.y];
    pair_mode<T, KS, STAGE, MODE, FLAT>(a, int64_t(blockIdx.x) * 4, int64_t(gridDim.x) * 4);
}

// Blocks of kernel K that fit on the device at once (occupancy x compute units): a
// chain-kernel grid is capped there, so that every block works through many slabs and
// its setup (A_t into LDS, the B_t fragments) is paid once per block, not per slab.
std::mutex g_resident_mu;
std::map<const void*, int> g_resident;   // kernel -> resident blocks per compute unit

template <class K>
int64_t resident_blocks(K kernel, int threads, int cus) {
    const void* key = reinterpret_cast<const void*>(kernel);
    std::lock_guard<std::mutex> g(g_resident_mu);
    auto it = g_resident.find(key);
    if (it == g_resident.end()) {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, threads, 0) != hipSuccess || b < 1) b = 1;
        it = g_resident.emplace(key, b).first;
    }
    return int64_t(it->second) * cus;
}

template <class T, int P, int NT>
__global__ __launch_bounds__(64 * (P / 16)) void k_dense_pair_reg(const PairArgs<T> a) {
    dense_pair_slabs_reg<T, P, NT>(a, blockIdx.x, gridDim.x);
}

template <class T, int P, int NT>
__global__ __launch_bounds__(64 * (P / 16)) void k_dense_pair_reg_batch(const PairArgs<T>* __restrict__ tab) {
    // a private copy: every term index in this kernel is a compile-time constant, so
    // the copy is scalarized (no scratch)
    const PairArgs<T> a = tab[blockIdx.y];
    dense_pair_slabs_reg<T, P, NT>(a, blockIdx.x, gridDim.x);
}

template <class T, int P>
__global__ __launch_bounds__(256) void k_dense_pair(const PairArgs<T> a) {
    if constexpr (P <= 48) dense_pair_slabs_lds<T, P>(a, blockIdx.x, gridDim.x);
    else dense_pair_slabs<T, P>(a, blockIdx.x, gridDim.x);
}

// One group of many operators of one structure: problem blockIdx.y takes its
// arguments from the table.
template <class T, int P>
__global__ __launch_bounds__(256) void k_dense_pair_batch(const PairArgs<T>* __restrict__ tab) {
    // the record is read in place (a private copy indexed by term would live in scratch)
    const PairArgs<T>& a = tab[blockIdx.y];
    if constexpr (P <= 48) dense_pair_slabs_lds<T, P>(a, blockIdx.x, gridDim.x);
    else dense_pair_slabs<T, P>(a, blockIdx.x, gridDim.x);
}

bool same(const std::vector<cd>& x, const std::vector<cd>& y) { return x == y; }

}  // namespace

// Host plan: every kept term is a product of exactly two factors on axes with
// d <= 64 on a real operator; factors of any kind are densified.
void dense_of_factor(const Factor& f, std::vector<cd>& M) {
    const int64_t d = f.d;
    M.assign(size_t(d) * d, cd(0));
    if (f.kind == KFC_DENSE) M = f.matrix;
    else if (f.kind == KFC_DIAG)
        for (int64_t i = 0; i < d; ++i) M[i + i * d] = f.diag[i];
    else if (f.kind == KFC_BANDED) {
        for (size_t b = 0; b < f.offsets.size(); ++b)
            for (int64_t i = 0; i < d; ++i) {
                const int64_t j = i + f.offsets[b];
                if (j >= 0 && j < d) M[i + j * d] += f.weights[i + b * d];
            }
    } else {
        for (int64_t i = 0; i < d; ++i) {
            const int64_t j = i + f.shift;
            if (j >= 0 && j < d) M[j + i * d] = f.factor;
        }
    }
}

int op_build_pairs(kfc_op* op) {
    PairPlan& P = op->pairs;
    P = PairPlan();
    if (op->cplx || op->plan.ok || op->terms.empty()) return KFC_OK;
    for (const Term& t : op->terms) {
        if (t.factors.size() != 2) return KFC_OK;
        for (const Factor& f : t.factors)
            if (f.d > 64) return KFC_OK;
    }
    for (const Term& t : op->terms) {
        const int p = t.factors[0].axis, q = t.factors[1].axis;
        PairGroup* g = nullptr;
        for (PairGroup& x : P.groups)
            if (x.p == p && x.q == q) g = &x;
        if (!g) {
            P.groups.push_back(PairGroup{});
            g = &P.groups.back();
            g->p = p;
            g->q = q;
            g->dp = t.factors[0].d;
            g->dq = t.factors[1].d;
        }
        std::vector<cd> A, B;
        dense_of_factor(t.factors[0], A);
        dense_of_factor(t.factors[1], B);
        bool merged = false;
        for (size_t k = 0; k < g->A.size() && !merged; ++k) {
            if (same(g->A[k], A)) {   // c1 A(x)B1 + c A(x)B = A (x) (c1 B1 + c B)
                for (size_t i = 0; i < B.size(); ++i) g->B[k][i] = g->c[k] * g->B[k][i] + t.coeff * B[i];
                g->c[k] = 1.0;
                merged = true;
            } else if (same(g->B[k], B)) {
                for (size_t i = 0; i < A.size(); ++i) g->A[k][i] = g->c[k] * g->A[k][i] + t.coeff * A[i];
                g->c[k] = 1.0;
                merged = true;
            }
        }
        if (!merged) {
            g->A.push_back(std::move(A));
            g->B.push_back(std::move(B));
            g->c.push_back(t.coeff);
        }
    }
    for (const PairGroup& g : P.groups)
        if (g.A.size() > size_t(MAXT)) return KFC_OK;
    // the first group writes Y = D .* X + ...: a group the split form may take goes after
    // the others, whose slab kernels form D .* X from the X they read anyway (the split
    // form's second stage would read X once more for it)
    {
        std::vector<int64_t> st(op->ndims + 1, 1);
        for (int a = 1; a <= op->ndims; ++a) st[a] = st[a - 1] * op->dims[a - 1];
        std::stable_partition(P.groups.begin(), P.groups.end(), [&](const PairGroup& g) {
            return !(g.A.size() == 1 && st[g.p] >= 16 && g.dp <= 64 && g.dq <= 64);
        });
    }
    P.ok = true;
    return KFC_OK;
}

template <class T>
static const T* up(kfc_op* op, DeviceData& dd, const std::vector<cd>& v) {
    std::vector<T> h(v.size());
    for (size_t i = 0; i < v.size(); ++i) h[i] = T(v[i].real());
    auto b = std::make_unique<DevBuf>();
    if (b->ensure(h.size() * sizeof(T)) != hipSuccess) return nullptr;
    if (hipMemcpy(b->p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    const T* p = b->as<T>();
    dd.blobs.push_back(std::move(b));
    return p;
}

template <class T>
int pairs_upload(kfc_op* op, DeviceData& dd) {
    dd.pair_ptr.clear();
    for (const PairGroup& g : op->pairs.groups) {
        std::vector<const void*> ptrs;
        for (size_t k = 0; k < g.A.size(); ++k) {
            const T* A = up<T>(op, dd, g.A[k]);
            const T* B = up<T>(op, dd, g.B[k]);
            if (!A || !B) return kfc_fail(op->ctx, KFC_ENOMEM, "dense pair upload failed");
            ptrs.push_back(A);
            ptrs.push_back(B);
        }
        dd.pair_ptr.push_back(ptrs);
    }
    return KFC_OK;
}

template int pairs_upload<double>(kfc_op*, DeviceData&);
template int pairs_upload<float>(kfc_op*, DeviceData&);

// The launch arguments of group gi of op on ncols columns of X -> Y.
// Whether some group of op can run in the split form (then its applies want N x ncols
// scalars of scratch).
static bool op_wants_pairz(const kfc_op* op) {
    int64_t stride = 1;
    std::vector<int64_t> st(op->ndims + 1, 1);
    for (int a = 0; a < op->ndims; ++a) {
        st[a] = stride;
        stride *= op->dims[a];
    }
    for (const PairGroup& g : op->pairs.groups)
        if (g.A.size() == 1 && st[g.p] >= 16 && g.dp <= 64 && g.dq <= 64) return true;
    return false;
}

// The context's scratch for the split form, grown to hold `bytes`; null when no group of
// op splits or when the allocation fails (the groups then run whole).
static void* pairz_for(kfc_op* op, size_t bytes) {
    if (!op_wants_pairz(op)) return nullptr;
    if (op->ctx->pairz.ensure(bytes) == hipSuccess) return op->ctx->pairz.p;
    (void)hipGetLastError();   // a failed allocation must not fail the launches after it
    return nullptr;
}

// z: N x ncols scalars of scratch for the split form (null: every group runs whole).
template <class T>
static PairArgs<T> pair_args(const kfc_op* op, size_t gi, const void* X, void* Y, int64_t ncols, int64_t ld,
                             void* z = nullptr) {
    const int prec = std::is_same<T, double>::value ? 0 : 1;
    const DeviceData& dd = op->dd[prec];
    std::vector<int64_t> stride(op->ndims + 1, 1);
    for (int a = 1; a <= op->ndims; ++a) stride[a] = stride[a - 1] * op->dims[a - 1];
    const PairGroup& g = op->pairs.groups[gi];
    PairArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.X = static_cast<const T*>(X);
    a.Y = static_cast<T*>(Y);
    a.D = gi == 0 ? dd.diag.as<T>() : nullptr;
    a.ld = ld;
    a.dp = int(g.dp);
    a.dq = int(g.dq);
    a.Mp = (a.dp + 15) / 16 * 16;
    a.Nq = (a.dq + 15) / 16 * 16;
    a.L = stride[g.p];
    a.M = stride[g.q] / (stride[g.p] * g.dp);
    a.R = op->N / stride[g.q + 1];
    a.sp = stride[g.p];
    a.sq = stride[g.q];
    a.sr = stride[g.q + 1];
    a.nslab = a.L * a.M * a.R * ncols;
    a.nterm = int(g.A.size());
    for (int t = 0; t < a.nterm; ++t) {
        a.c[t] = T(g.c[t].real());
        a.A[t] = static_cast<const T*>(dd.pair_ptr[gi][2 * t]);
        a.B[t] = static_cast<const T*>(dd.pair_ptr[gi][2 * t + 1]);
    }
    a.Z = static_cast<T*>(z);
    return a;
}

bool pairs_same_structure(const kfc_op* a, const kfc_op* b) {
    if (!a->pairs.ok || !b->pairs.ok || a->N != b->N || a->dims != b->dims) return false;
    if (a->pairs.groups.size() != b->pairs.groups.size()) return false;
    for (size_t g = 0; g < a->pairs.groups.size(); ++g) {
        const PairGroup &x = a->pairs.groups[g], &y = b->pairs.groups[g];
        if (x.p != y.p || x.q != y.q || x.dp != y.dp || x.dq != y.dq) return false;
    }
    return true;
}

// Launch shapes of the register-resident kernels (groups of at most two terms, extents
// up to 48): the kernel family (0: register-chained, 1: register fragments with the
// intermediate through LDS, 2: wave per slab, 3: neighbouring slabs of a strided p axis
// staged through LDS by the block, else as 2) and the
// grid (0: one block per slab, at most 4096 blocks in all; k > 0: k times the blocks
// that fit on the device at once, each looping over slabs; the wave-per-slab kernels
// always take the latter, counting a wave per slab).  Which is fastest depends on the
// work in flight (a single operator's columns or a whole lockstep sweep), so the
// launcher times them per group on the first application of each structure and keeps
// the winner.
// Family 4 runs a one-term group with a strided p axis in the split form (pair_mode),
// when the launch record carries its scratch; other groups take family 2 instead.
// Family 5 is family 4 with rows tiled across blocks and G in LDS (pair_mode FLAT).
struct PairShape {
    int family, grid;
};
constexpr PairShape kPairShapes[] = {{0, 0}, {0, 1}, {0, 2}, {0, 4}, {1, 0}, {1, 1}, {1, 2},
                                     {2, 1}, {2, 2}, {3, 1}, {3, 2}, {4, 1}, {4, 2}, {5, 1}, {5, 2}};
constexpr int kNumPairShapes = int(sizeof(kPairShapes) / sizeof(kPairShapes[0]));

template <class T, int P, int NT>
int launch_reg_group(const PairShape& sh, const PairArgs<T>& a, const PairArgs<T>* tab, int n, int cus,
                     hipStream_t st) {
    constexpr int threads = 64 * (P / 16);
    auto grid_for = [&](auto kernel) -> unsigned {
        if (sh.grid == 0) return unsigned(std::max<int64_t>(1, std::min<int64_t>(a.nslab, std::max(1, 4096 / n))));
        const int64_t cap = resident_blocks(kernel, threads, cus) * sh.grid;
        return unsigned(std::max<int64_t>(1, std::min<int64_t>(a.nslab, (cap + n - 1) / n)));
    };
    if (tab) {
        if (sh.family == 0) {
            auto k = k_dense_pair_chain_batch<T, P, NT>;
            hipLaunchKernelGGL(k, dim3(grid_for(k), unsigned(n)), dim3(threads), 0, st, tab);
        } else {
            auto k = k_dense_pair_reg_batch<T, P, NT>;
            hipLaunchKernelGGL(k, dim3(grid_for(k), unsigned(n)), dim3(threads), 0, st, tab);
        }
    } else {
        if (sh.family == 0) {
            auto k = k_dense_pair_chain<T, P, NT>;
            hipLaunchKernelGGL(k, dim3(grid_for(k)), dim3(threads), 0, st, a);
        } else {
            auto k = k_dense_pair_reg<T, P, NT>;
            hipLaunchKernelGGL(k, dim3(grid_for(k)), dim3(threads), 0, st, a);
        }
    }
    return hipGetLastError() == hipSuccess ? KFC_OK : KFC_EDEVICE;
}

template <class T, int PI, int PJ, int NT>
int launch_staged_group(const PairShape& sh, const PairArgs<T>& a, const PairArgs<T>* tab, int n, int cus,
                        hipStream_t st) {
    constexpr int WS = staged_ws<T>(), threads = 64 * WS;
    const int mult = std::max(1, sh.grid);
    const int64_t ncols = a.nslab / (a.L * a.M * a.R);
    const int64_t nchunk = (a.L + WS - 1) / WS * a.M * a.R * ncols;
    auto grid_for = [&](auto kernel) -> unsigned {
        const int64_t cap = resident_blocks(kernel, threads, cus) * mult;
        return unsigned(std::max<int64_t>(1, std::min<int64_t>(nchunk, (cap + n - 1) / n)));
    };
    if (tab) {
        auto k = k_dense_pair_staged_batch<T, PI, PJ, NT>;
        hipLaunchKernelGGL(k, dim3(grid_for(k), unsigned(n)), dim3(threads), 0, st, tab);
    } else {
        auto k = k_dense_pair_staged<T, PI, PJ, NT>;
        hipLaunchKernelGGL(k, dim3(grid_for(k)), dim3(threads), 0, st, a);
    }
    return hipGetLastError() == hipSuccess ? KFC_OK : KFC_EDEVICE;
}

template <class T, int PI, int PJ, int NT>
int launch_wave_group(const PairShape& sh, const PairArgs<T>& a, const PairArgs<T>* tab, int n, int cus,
                      hipStream_t st) {
    // family 3 stages neighbouring slabs of a strided p axis through LDS
    if (sh.family == 3 && a.L > 1) return launch_staged_group<T, PI, PJ, NT>(sh, a, tab, n, cus, st);
    constexpr int W = WAVE_W, threads = 64 * W;
    const int mult = std::max(1, sh.grid);
    auto grid_for = [&](auto kernel) -> unsigned {
        const int64_t cap = resident_blocks(kernel, threads, cus) * mult;
        const int64_t need = (a.nslab + W - 1) / W;   // blocks that give every slab its own wave
        return unsigned(std::max<int64_t>(1, std::min<int64_t>(need, (cap + n - 1) / n)));
    };
    // every problem of a batch is the same group, so a.D decides for all of them
    if (tab) {
        auto k = a.D ? k_dense_pair_wave_batch<T, PI, PJ, NT, true> : k_dense_pair_wave_batch<T, PI, PJ, NT, false>;
        hipLaunchKernelGGL(k, dim3(grid_for(k), unsigned(n)), dim3(threads), 0, st, tab);
    } else {
        auto k = a.D ? k_dense_pair_wave<T, PI, PJ, NT, true> : k_dense_pair_wave<T, PI, PJ, NT, false>;
        hipLaunchKernelGGL(k, dim3(grid_for(k)), dim3(threads), 0, st, a);
    }
    return hipGetLastError() == hipSuccess ? KFC_OK : KFC_EDEVICE;
}

// A group that can run in the split form: one term, at least 16 entries before its p
// axis, extents up to 64.
template <class T>
bool split_ok(const PairArgs<T>& a, int ntmax) {
    const int64_t ncol = a.nslab / (a.L * a.M * a.R);
    const int64_t tiles = std::max((a.sq + 15) / 16 * a.R, (a.sp + 15) / 16 * a.M * a.dq * a.R) * ncol;
    return ntmax == 1 && a.L >= 16 && a.dp <= 64 && a.dq <= 64 && tiles < (int64_t(1) << 31);
}

template <class T, int KS>
int launch_split_sized(const PairShape& sh, const PairArgs<T>& a, const PairArgs<T>* tab, int n, int cus,
                       hipStream_t st, int stage) {
    const int mult = std::max(1, sh.grid);
    const int64_t ncol = a.nslab / (a.L * a.M * a.R);
    const int64_t rows = stage == 1 ? a.sq : a.sp, nout = stage == 1 ? a.R : a.M * a.dq * a.R;
    const bool flat = sh.family == 5;
    const int64_t tiles = (flat ? (rows * nout + 15) / 16 : (rows + 15) / 16 * nout) * ncol;
    auto grid_for = [&](auto kernel) -> unsigned {
        const int64_t cap = resident_blocks(kernel, 256, cus) * mult;
        return unsigned(std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, (cap + n - 1) / n)));
    };
    auto go = [&](auto kb, auto k1) {
        if (tab) hipLaunchKernelGGL(kb, dim3(grid_for(kb), unsigned(n)), dim3(256), 0, st, tab);
        else hipLaunchKernelGGL(k1, dim3(grid_for(k1)), dim3(256), 0, st, a);
    };
    if (flat) {
        if (stage == 1) go(k_pair_mode_batch<T, KS, 1, 0, true>, k_pair_mode<T, KS, 1, 0, true>);
        else if (a.D) go(k_pair_mode_batch<T, KS, 2, 2, true>, k_pair_mode<T, KS, 2, 2, true>);
        else go(k_pair_mode_batch<T, KS, 2, 1, true>, k_pair_mode<T, KS, 2, 1, true>);
    } else {
        if (stage == 1) go(k_pair_mode_batch<T, KS, 1, 0, false>, k_pair_mode<T, KS, 1, 0, false>);
        else if (a.D) go(k_pair_mode_batch<T, KS, 2, 2, false>, k_pair_mode<T, KS, 2, 2, false>);
        else go(k_pair_mode_batch<T, KS, 2, 1, false>, k_pair_mode<T, KS, 2, 1, false>);
    }
    return hipGetLastError() == hipSuccess ? KFC_OK : KFC_EDEVICE;
}

// The split form: stage 1 (the q axis into Z), then stage 2 (the p axis from Z into Y).
template <class T>
int launch_split(const PairShape& sh, const PairArgs<T>& a, const PairArgs<T>* tab, int n, int cus, hipStream_t st) {
    for (int stage = 1; stage <= 2; ++stage) {
        const int d = stage == 1 ? a.dq : a.dp;
        int rc;
        if (d <= 16) rc = launch_split_sized<T, 4>(sh, a, tab, n, cus, st, stage);
        else if (d <= 32) rc = launch_split_sized<T, 8>(sh, a, tab, n, cus, st, stage);
        else if (d <= 48) rc = launch_split_sized<T, 12>(sh, a, tab, n, cus, st, stage);
        else rc = launch_split_sized<T, 16>(sh, a, tab, n, cus, st, stage);
        if (rc != KFC_OK) return rc;
    }
    return KFC_OK;
}

// the smallest of 16, 32, 40, 48 bounding d (0: none)
inline int wave_extent(int d) { return d <= 16 ? 16 : d <= 32 ? 32 : d <= 40 ? 40 : d <= 48 ? 48 : 0; }

template <class T, int NT, int PI>
int launch_wave_j(const PairShape& sh, const PairArgs<T>& a, const PairArgs<T>* tab, int n, int cus, hipStream_t st) {
    switch (wave_extent(a.dq)) {
        case 16: return launch_wave_group<T, PI, 16, NT>(sh, a, tab, n, cus, st);
        case 32: return launch_wave_group<T, PI, 32, NT>(sh, a, tab, n, cus, st);
        case 40: return launch_wave_group<T, PI, 40, NT>(sh, a, tab, n, cus, st);
        default: return launch_wave_group<T, PI, 48, NT>(sh, a, tab, n, cus, st);
    }
}

template <class T, int NT>
int launch_wave_sized(const PairShape& sh, const PairArgs<T>& a, const PairArgs<T>* tab, int n, int cus,
                      hipStream_t st) {
    switch (wave_extent(a.dp)) {
        case 16: return launch_wave_j<T, NT, 16>(sh, a, tab, n, cus, st);
        case 32: return launch_wave_j<T, NT, 32>(sh, a, tab, n, cus, st);
        case 40: return launch_wave_j<T, NT, 40>(sh, a, tab, n, cus, st);
        default: return launch_wave_j<T, NT, 48>(sh, a, tab, n, cus, st);
    }
}

// One group on n problems (tab: their device records, or null with n = 1 and a passed
// by value).  Groups of at most two terms with extents up to 48 take shape sh.
template <class T>
int launch_group(const PairShape& sh0, const PairArgs<T>& a, int ntmax, const PairArgs<T>* tab, int n, int cus,
                 hipStream_t st) {
    PairShape sh = sh0;
    if (sh.family == 4 || sh.family == 5) {
        // every problem of a batch carries scratch or none (pairs_fill_table)
        if (a.Z && split_ok(a, ntmax)) return launch_split<T>(sh, a, tab, n, cus, st);
        sh.family = 2;
    }
    const int Pm = std::max(a.Mp, a.Nq);
    // the wave-per-slab kernels address a slab with byte offsets below 2^31, padded
    // tiles included
    const bool wave_ok = int64_t(Pm) * (a.sp + a.sq) * int64_t(sizeof(T)) < (int64_t(1) << 31);
    if (Pm <= 48 && ntmax <= 2 && sh.family >= 2 && wave_ok) {
        return ntmax <= 1 ? launch_wave_sized<T, 1>(sh, a, tab, n, cus, st)
                          : launch_wave_sized<T, 2>(sh, a, tab, n, cus, st);
    }
    if (Pm <= 48 && ntmax <= 2) {
        if (Pm <= 16) return ntmax <= 1 ? launch_reg_group<T, 16, 1>(sh, a, tab, n, cus, st)
                                        : launch_reg_group<T, 16, 2>(sh, a, tab, n, cus, st);
        if (Pm <= 32) return ntmax <= 1 ? launch_reg_group<T, 32, 1>(sh, a, tab, n, cus, st)
                                        : launch_reg_group<T, 32, 2>(sh, a, tab, n, cus, st);
        return ntmax <= 1 ? launch_reg_group<T, 48, 1>(sh, a, tab, n, cus, st)
                          : launch_reg_group<T, 48, 2>(sh, a, tab, n, cus, st);
    }
    const dim3 grid(unsigned(std::max<int64_t>(1, std::min<int64_t>(a.nslab, std::max(1, 4096 / n)))), unsigned(n));
    if (tab) {
        if (Pm <= 16) hipLaunchKernelGGL((k_dense_pair_batch<T, 16>), grid, dim3(256), 0, st, tab);
        else if (Pm <= 32) hipLaunchKernelGGL((k_dense_pair_batch<T, 32>), grid, dim3(256), 0, st, tab);
        else if (Pm <= 48) hipLaunchKernelGGL((k_dense_pair_batch<T, 48>), grid, dim3(256), 0, st, tab);
        else hipLaunchKernelGGL((k_dense_pair_batch<T, 64>), grid, dim3(256), 0, st, tab);
    } else {
        if (Pm <= 16) hipLaunchKernelGGL((k_dense_pair<T, 16>), dim3(grid.x), dim3(256), 0, st, a);
        else if (Pm <= 32) hipLaunchKernelGGL((k_dense_pair<T, 32>), dim3(grid.x), dim3(256), 0, st, a);
        else if (Pm <= 48) hipLaunchKernelGGL((k_dense_pair<T, 48>), dim3(grid.x), dim3(256), 0, st, a);
        else hipLaunchKernelGGL((k_dense_pair<T, 64>), dim3(grid.x), dim3(256), 0, st, a);
    }
    return hipGetLastError() == hipSuccess ? KFC_OK : KFC_EDEVICE;
}

// The tuner's cache: (dims, each group's axes, extents and term count, scalar size,
// total slabs rounded up to a power of two) -> index into kPairShapes for each group.
struct PairTuneKey {
    std::vector<int64_t> sig;
    bool operator<(const PairTuneKey& o) const { return sig < o.sig; }
};
std::mutex g_pair_tune_mu;
std::map<PairTuneKey, std::vector<int>> g_pair_tune;

PairTuneKey pair_tune_key(const kfc_op* op, size_t scalar, int64_t slabs_total, bool scratch) {
    PairTuneKey k;
    k.sig = op->dims;
    k.sig.push_back(scratch ? 1 : 0);   // the split form is a candidate only with scratch
    for (const PairGroup& g : op->pairs.groups) {
        k.sig.push_back(g.p);
        k.sig.push_back(g.q);
        k.sig.push_back(g.dp);
        k.sig.push_back(g.dq);
        k.sig.push_back(int64_t(g.A.size()));
    }
    int64_t b = 1;
    while (b < slabs_total) b <<= 1;
    k.sig.push_back(int64_t(scalar));
    k.sig.push_back(b);
    return k;
}

// Applies the ng groups in order (the first writes Y, the rest add to it), each with its
// tuned shape.  On the first application of a key every group times every candidate
// (once to warm up, then the better of two runs timed with HIP events on the stream); a
// group run out of order leaves
// Y wrong, so the whole apply runs again with the picks afterwards.
int run_pair_groups(kfc_ctx* ctx, const PairTuneKey& key, int ng, hipStream_t st,
                    const std::function<int(const PairShape&, int)>& one) {
    auto apply = [&](const std::vector<int>& pick) {
        for (int g = 0; g < ng; ++g) {
            const int rc = one(kPairShapes[pick[g]], g);
            if (rc != KFC_OK) return rc;
        }
        return int(KFC_OK);
    };
    // KFC_PAIR_SHAPE=k pins shape k of kPairShapes for every group (read at each call:
    // tests and measurements switch it within one process)
    if (const char* env = std::getenv("KFC_PAIR_SHAPE")) {
        const int k = std::atoi(env);
        if (env[0] != '\0' && k >= 0 && k < kNumPairShapes) return apply(std::vector<int>(ng, k));
    }
    std::vector<int> pick;
    {
        std::lock_guard<std::mutex> g(g_pair_tune_mu);
        auto it = g_pair_tune.find(key);
        if (it != g_pair_tune.end()) pick = it->second;
    }
    if (int(pick.size()) == ng) return apply(pick);
    hipEvent_t e0, e1;
    KFC_HIP(ctx, hipEventCreate(&e0));
    KFC_HIP(ctx, hipEventCreate(&e1));
    pick.assign(ng, 0);
    int rc = KFC_OK;
    for (int g = 0; g < ng && rc == KFC_OK; ++g) {
        float best = 1e30f;
        for (int c = 0; c < kNumPairShapes && rc == KFC_OK; ++c) {
            if ((rc = one(kPairShapes[c], g)) != KFC_OK) break;
            for (int rep = 0; rep < 2 && rc == KFC_OK; ++rep) {   // the better of two timed runs
                if ((rc = hipEventRecord(e0, st) == hipSuccess ? KFC_OK : KFC_EDEVICE) != KFC_OK) break;
                if ((rc = one(kPairShapes[c], g)) != KFC_OK) break;
                if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                    rc = KFC_EDEVICE;
                    break;
                }
                float ms = 0;
                if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    rc = KFC_EDEVICE;
                    break;
                }
                if (ms < best) {
                    best = ms;
                    pick[g] = c;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != KFC_OK) return kfc_fail(ctx, rc, "dense pair launch failed");
    {
        std::lock_guard<std::mutex> g(g_pair_tune_mu);
        g_pair_tune[key] = pick;
    }
    return apply(pick);   // Y from the chosen shapes
}

// Problem q's scratch for the split form is the q-th N x ncols piece of the context's
// buffer, when the problems share one context and the buffer already holds all n;
// this fill allocates nothing (a graph replays its launches after each refill).
template <class T>
void pairs_fill_table(kfc_op* const* ops, int n, const void* const* X, void* const* Y, int64_t ncols, int64_t ld,
                      void* host_tab) {
    const size_t ng = ops[0]->pairs.groups.size();
    PairArgs<T>* h = static_cast<PairArgs<T>*>(host_tab);
    const size_t per = size_t(ops[0]->N) * size_t(ncols);
    const DevBuf& zb = ops[0]->ctx->pairz;
    bool zok = op_wants_pairz(ops[0]) && zb.p && zb.bytes >= per * size_t(n) * sizeof(T);
    for (int q = 1; q < n && zok; ++q) zok = ops[q]->ctx == ops[0]->ctx;
    for (size_t gi = 0; gi < ng; ++gi)
        for (int q = 0; q < n; ++q)
            h[gi * n + q] = pair_args<T>(ops[q], gi, X[q], Y[q], ncols, ld, zok ? zb.as<T>() + per * q : nullptr);
}

int pairs_reserve(kfc_op* op, int n, int64_t ncols, int prec) {
    // when the device cannot hold it, the tables get no scratch and every group runs whole
    (void)pairz_for(op, size_t(op->N) * size_t(ncols) * size_t(n) * (prec == 0 ? sizeof(double) : sizeof(float)));
    return KFC_OK;
}

template <class T>
int launch_pairs_batch(kfc_op* const* ops, int n, const void* const* X, void* const* Y, int64_t ncols, int64_t ld,
                       hipStream_t st, void* host_tab, void* dev_tab) {
    if (n == 0) return KFC_OK;
    // the scratch of the split form, if the context can hold it (else every group runs whole)
    (void)pairz_for(ops[0], size_t(ops[0]->N) * size_t(ncols) * size_t(n) * sizeof(T));
    pairs_fill_table<T>(ops, n, X, Y, ncols, ld, host_tab);
    return launch_pairs_filled<T>(ops, n, st, host_tab, dev_tab);
}

template <class T>
int launch_pairs_filled(kfc_op* const* ops, int n, hipStream_t st, void* host_tab, void* dev_tab) {
    if (n == 0) return KFC_OK;
    const kfc_op* op0 = ops[0];
    const size_t ng = op0->pairs.groups.size();
    PairArgs<T>* h = static_cast<PairArgs<T>*>(host_tab);
    PairArgs<T>* d = static_cast<PairArgs<T>*>(dev_tab);
    KFC_HIP(op0->ctx, hipMemcpyAsync(d, h, sizeof(PairArgs<T>) * ng * n, hipMemcpyHostToDevice, st));
    std::vector<int> ntmax(ng, 0);
    int64_t slabs = 0;
    for (size_t gi = 0; gi < ng; ++gi) {
        for (int q = 0; q < n; ++q) ntmax[gi] = std::max(ntmax[gi], h[gi * n + q].nterm);
        slabs += h[gi * n].nslab * n;
    }
    const int cus = op0->ctx->num_cus;
    auto one = [&](const PairShape& sh, int gi) { return launch_group<T>(sh, h[gi * n], ntmax[gi], d + gi * n, n, cus, st); };
    bool scratch = false;
    for (size_t gi = 0; gi < ng; ++gi) scratch |= h[gi * n].Z != nullptr;
    return run_pair_groups(op0->ctx, pair_tune_key(op0, sizeof(T), slabs, scratch), int(ng), st, one);
}

template void pairs_fill_table<double>(kfc_op* const*, int, const void* const*, void* const*, int64_t, int64_t, void*);
template void pairs_fill_table<float>(kfc_op* const*, int, const void* const*, void* const*, int64_t, int64_t, void*);
template int launch_pairs_filled<double>(kfc_op* const*, int, hipStream_t, void*, void*);
template int launch_pairs_filled<float>(kfc_op* const*, int, hipStream_t, void*, void*);
template int launch_pairs_batch<double>(kfc_op* const*, int, const void* const*, void* const*, int64_t, int64_t,
                                        hipStream_t, void*, void*);
template int launch_pairs_batch<float>(kfc_op* const*, int, const void* const*, void* const*, int64_t, int64_t,
                                       hipStream_t, void*, void*);

size_t pair_args_bytes(int prec) { return prec == 0 ? sizeof(PairArgs<double>) : sizeof(PairArgs<float>); }

template <class T>
int launch_pairs(kfc_op* op, const void* X, void* Y, int64_t ncols, int64_t ld, hipStream_t st, bool scratch) {
    if (op->pairs.groups.empty()) return kfc_fail(op->ctx, KFC_EUNSUPPORTED, "dense pair path without groups");
    const size_t ng = op->pairs.groups.size();
    void* z = scratch ? pairz_for(op, size_t(op->N) * size_t(ncols) * sizeof(T)) : nullptr;
    std::vector<PairArgs<T>> args(ng);
    int64_t slabs = 0;
    for (size_t gi = 0; gi < ng; ++gi) {
        args[gi] = pair_args<T>(op, gi, X, Y, ncols, ld, z);
        slabs += args[gi].nslab;
    }
    const int cus = op->ctx->num_cus;
    auto one = [&](const PairShape& sh, int gi) { return launch_group<T>(sh, args[gi], args[gi].nterm, nullptr, 1, cus, st); };
    return run_pair_groups(op->ctx, pair_tune_key(op, sizeof(T), slabs, z != nullptr), int(ng), st, one);
}

template int launch_pairs<double>(kfc_op*, const void*, void*, int64_t, int64_t, hipStream_t, bool);
template int launch_pairs<float>(kfc_op*, const void*, void*, int64_t, int64_t, hipStream_t, bool);
