// Device forms of the operator's derived quantities, written straight into device
// memory so that nothing of length N passes through the host:
//   * the diagonal: _device_diagonal / _full_diagonal (src/apply.jl:455-478), the
//     fused part alone (the apply's diagonal table) or with the kept terms' diagonals;
//   * the range of the fused diagonal that gershgorin_bound reads (src/apply.jl:586-593);
//   * product_seed(H, k) (src/apply.jl:655-685): the k product states of the per-axis
//     eigenbases with the smallest summed energies, or, when every axis is already an
//     eigenbasis, the unit vectors of the k smallest entries of the full diagonal.
//
// Every entry is formed by the host formula's own sequence of complex operations
// (constant, then each term's coefficient times its factors' entries in factor order),
// with contraction into FMA switched off, so device and host agree bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <queue>
#include <unordered_set>

#include "blas.h"
#include "kfc_internal.h"

namespace {

constexpr int DG_MAXF = 16;   // factors per term (one per axis at most)

struct DTerm {
    double cr, ci;
    int nf;
    int axis[DG_MAXF];
    int64_t off[DG_MAXF];   // first entry of the factor's diagonal in the value table
};

struct DGeom {
    int K = 0;
    int nterm = 0;            // terms [0, nfused) are the fused ones, the rest kept
    int nfused = 0;
    double c0r = 0, c0i = 0;  // the constant
    int64_t stride[16] = {};
    int64_t dims[16] = {};
};

// The term table of the diagonal: fused terms, then (with `kept`) the kept terms
// whose factors all have a nonzero diagonal, in the host's order.
struct DiagTable {
    DGeom g;
    std::vector<DTerm> terms;
    std::vector<double> vals;   // interleaved re, im
};

void push_term(DiagTable& T, cd coeff, const std::vector<int>& axes, const std::vector<std::vector<cd>>& dv) {
    DTerm t{};
    t.cr = coeff.real();
    t.ci = coeff.imag();
    t.nf = int(axes.size());
    for (size_t f = 0; f < axes.size(); ++f) {
        t.axis[f] = axes[f];
        t.off[f] = int64_t(T.vals.size() / 2);
        for (const cd& x : dv[f]) {
            T.vals.push_back(x.real());
            T.vals.push_back(x.imag());
        }
    }
    T.terms.push_back(t);
}

int diag_table(kfc_op* op, bool kept, DiagTable& T) {
    T.g = DGeom();
    T.terms.clear();
    T.vals.clear();
    T.g.K = op->ndims;
    T.g.c0r = op->constant.real();
    T.g.c0i = op->constant.imag();
    int64_t s = 1;
    for (int a = 0; a < op->ndims; ++a) {
        T.g.stride[a] = s;
        T.g.dims[a] = op->dims[a];
        s *= op->dims[a];
    }
    for (const Term& t : op->diagterms) {
        if (t.factors.empty()) continue;   // the constant
        if (t.factors.size() > size_t(DG_MAXF))
            return kfc_fail(op->ctx, KFC_EUNSUPPORTED, "diagonal: a term with more than 16 factors");
        std::vector<int> axes;
        std::vector<std::vector<cd>> dv;
        for (const Factor& f : t.factors) {
            axes.push_back(f.axis);
            dv.push_back(factor_diagonal(f));
        }
        push_term(T, t.coeff, axes, dv);
    }
    T.g.nfused = int(T.terms.size());
    if (kept) {
        for (const Term& t : op->terms) {
            if (t.factors.size() > size_t(DG_MAXF))
                return kfc_fail(op->ctx, KFC_EUNSUPPORTED, "diagonal: a term with more than 16 factors");
            std::vector<int> axes;
            std::vector<std::vector<cd>> dv;
            bool zero = false;
            for (const Factor& f : t.factors) {
                axes.push_back(f.axis);
                dv.push_back(kept_factor_diagonal(f));
                bool allz = true;
                for (const cd& x : dv.back()) allz &= x == cd(0);
                zero |= allz;
            }
            if (!zero) push_term(T, t.coeff, axes, dv);
        }
    }
    T.g.nterm = int(T.terms.size());
    return KFC_OK;
}

// The table on the device, in one allocation: terms, then values.
struct DevTable {
    DevBuf buf;
    const DTerm* terms = nullptr;
    const double2* vals = nullptr;
};

int upload_table(kfc_op* op, const DiagTable& T, DevTable& D) {
    const size_t tb = (sizeof(DTerm) * std::max<size_t>(1, T.terms.size()) + 255) / 256 * 256;
    const size_t vb = sizeof(double) * std::max<size_t>(2, T.vals.size());
    KFC_HIP(op->ctx, D.buf.ensure(tb + vb));
    if (!T.terms.empty())
        KFC_HIP(op->ctx, hipMemcpyAsync(D.buf.p, T.terms.data(), sizeof(DTerm) * T.terms.size(), hipMemcpyHostToDevice,
                                        op->ctx->stream));
    if (!T.vals.empty())
        KFC_HIP(op->ctx, hipMemcpyAsync(static_cast<char*>(D.buf.p) + tb, T.vals.data(), sizeof(double) * T.vals.size(),
                                        hipMemcpyHostToDevice, op->ctx->stream));
    D.terms = static_cast<const DTerm*>(D.buf.p);
    D.vals = reinterpret_cast<const double2*>(static_cast<char*>(D.buf.p) + tb);
    return KFC_OK;
}

template <class S> __device__ inline void store_s(S* p, double r, double i);
template <> __device__ inline void store_s<double>(double* p, double r, double) { *p = r; }
template <> __device__ inline void store_s<float>(float* p, double r, double) { *p = float(r); }
template <> __device__ inline void store_s<Cx<double>>(Cx<double>* p, double r, double i) { *p = Cx<double>{r, i}; }
template <> __device__ inline void store_s<Cx<float>>(Cx<float>* p, double r, double i) {
    *p = Cx<float>{float(r), float(i)};
}

// Sum of terms [t0, t1) at flat index e, onto (ar, ai).
template <class I>
__device__ inline void term_sum(const DGeom& g, const DTerm* __restrict__ T, const double2* __restrict__ V, I e, int t0,
                                int t1, double& ar, double& ai) {
#pragma clang fp contract(off)
    for (int t = t0; t < t1; ++t) {
        double vr = T[t].cr, vi = T[t].ci;
        const int nf = T[t].nf;
        for (int f = 0; f < nf; ++f) {
            const int a = T[t].axis[f];
            const I i = (e / I(g.stride[a])) % I(g.dims[a]);
            const double2 u = V[T[t].off[f] + int64_t(i)];
            const double r = vr * u.x - vi * u.y;
            const double m = vr * u.y + vi * u.x;
            vr = r;
            vi = m;
        }
        ar += vr;
        ai += vi;
    }
}

// MODE 0: full real diagonal into out (double), with per-block min and the count of
//         non-finite entries in part[2 b], part[2 b + 1];
// MODE 1: fused diagonal into out (S);
// MODE 2: per-block max Re, min Re, max |Im| of the fused diagonal into part[3 b ..].
template <class I, class S, int MODE>
__global__ __launch_bounds__(256) void k_diag(DGeom g, const DTerm* __restrict__ T, const double2* __restrict__ V,
                                              int64_t N, S* __restrict__ out, double* __restrict__ part) {
    double r0 = MODE == 0 ? INFINITY : -INFINITY, r1 = MODE == 0 ? 0.0 : INFINITY, r2 = 0.0;
    const int tend = MODE == 0 ? g.nterm : g.nfused;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < N; e += int64_t(gridDim.x) * 256) {
        double ar = g.c0r, ai = g.c0i;
        term_sum<I>(g, T, V, I(e), 0, tend, ar, ai);
        if constexpr (MODE == 0) {
            store_s<S>(out + e, ar, ai);
            r0 = fmin(r0, ar);
            r1 += __builtin_isfinite(ar) ? 0.0 : 1.0;
        } else if constexpr (MODE == 1) {
            store_s<S>(out + e, ar, ai);
        } else {
            r0 = fmax(r0, ar);
            r1 = fmin(r1, ar);
            r2 = fmax(r2, fabs(ai));
        }
    }
    if (MODE == 1) return;
    __shared__ double red[3][256];
    red[0][threadIdx.x] = r0;
    red[1][threadIdx.x] = r1;
    red[2][threadIdx.x] = r2;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) {
            const int o = threadIdx.x + h;
            if (MODE == 0) {
                red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][o]);
                red[1][threadIdx.x] += red[1][o];
            } else {
                red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][o]);
                red[1][threadIdx.x] = fmin(red[1][threadIdx.x], red[1][o]);
                red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][o]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (MODE == 0) {
            part[2 * blockIdx.x] = red[0][0];
            part[2 * blockIdx.x + 1] = red[1][0];
        } else {
            part[3 * blockIdx.x] = red[0][0];
            part[3 * blockIdx.x + 1] = red[1][0];
            part[3 * blockIdx.x + 2] = red[2][0];
        }
    }
}

int diag_grid(int64_t N) { return int(std::max<int64_t>(1, std::min<int64_t>((N + 255) / 256, 2048))); }

template <class S, int MODE>
void launch_diag(const DGeom& g, const DevTable& D, int64_t N, S* out, double* part, hipStream_t st) {
    const int nb = diag_grid(N);
    if (N <= int64_t(UINT32_MAX))
        hipLaunchKernelGGL((k_diag<uint32_t, S, MODE>), dim3(nb), dim3(256), 0, st, g, D.terms, D.vals, N, out, part);
    else
        hipLaunchKernelGGL((k_diag<uint64_t, S, MODE>), dim3(nb), dim3(256), 0, st, g, D.terms, D.vals, N, out, part);
}

// ---- product seed -------------------------------------------------------------------------

// X[e, j] = prod_a U_a[i_a(e), c_a(j)]: the column table holds, per seed column j and
// axis a, the d_a entries of the chosen eigenvector, starting at coff[j K + a].
template <class I, class S>
__global__ __launch_bounds__(256) void k_seed_product(DGeom g, const double2* __restrict__ cols,
                                                      const int64_t* __restrict__ coff, int64_t N, S* __restrict__ X,
                                                      int64_t ldx) {
#pragma clang fp contract(off)
    const int j = blockIdx.y;
    const int64_t* co = coff + int64_t(j) * g.K;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < N; e += int64_t(gridDim.x) * 256) {
        double vr = 1.0, vi = 0.0;
        for (int a = 0; a < g.K; ++a) {
            const I i = (I(e) / I(g.stride[a])) % I(g.dims[a]);
            const double2 u = cols[co[a] + int64_t(i)];
            const double r = vr * u.x - vi * u.y;
            const double m = vr * u.y + vi * u.x;
            vr = r;
            vi = m;
        }
        store_s<S>(X + e + int64_t(j) * ldx, vr, vi);
    }
}

// X[:, j] = unit vector at sel[j]
template <class S>
__global__ __launch_bounds__(256) void k_seed_unit(const int64_t* __restrict__ sel, int64_t N, S* __restrict__ X,
                                                   int64_t ldx) {
    const int j = blockIdx.y;
    const int64_t s = sel[j];
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < N; e += int64_t(gridDim.x) * 256)
        store_s<S>(X + e + int64_t(j) * ldx, e == s ? 1.0 : 0.0, 0.0);
}

// ---- the k smallest entries of a device vector (radix select) ---------------------------
//
// Entries are ordered by (seed_order_key(value), index), a total order, so the k
// smallest are unique.  Each pass counts the entries that match the prefix found so far
// by their next 8-bit digit (the key's 64 bits first, then the index's bits) and
// extends the prefix by the digit where the running count reaches k; once every entry
// matching the prefix is needed, one more pass takes the entries at or below it.

__device__ inline uint64_t order_key_dev(double d) {
    if (d != d) return ~0ull;
    const uint64_t u = uint64_t(__double_as_longlong(d));
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

struct RadixState {
    uint64_t pk = 0, pi = 0;   // known key / index prefixes
    int nk = 0, ni = 0;        // their bit counts
    int ib = 8;                // index width in bits (multiple of 8)
};

// -1: below the prefix, 0: matches it, 1: above it
__device__ inline int radix_cmp(const RadixState& s, uint64_t key, uint64_t idx) {
    if (s.nk > 0) {
        const uint64_t tk = key >> (64 - s.nk);
        if (tk != s.pk) return tk < s.pk ? -1 : 1;
    }
    if (s.ni > 0) {
        const uint64_t ti = idx >> (s.ib - s.ni);
        if (ti != s.pi) return ti < s.pi ? -1 : 1;
    }
    return 0;
}

__global__ __launch_bounds__(256) void k_radix_hist(const double* __restrict__ D, int64_t N, RadixState s,
                                                    unsigned long long* __restrict__ hist) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < N; e += int64_t(gridDim.x) * 256) {
        const uint64_t key = order_key_dev(D[e]);
        if (radix_cmp(s, key, uint64_t(e)) != 0) continue;
        const int digit = s.nk < 64 ? int((key >> (56 - s.nk)) & 255) : int((uint64_t(e) >> (s.ib - s.ni - 8)) & 255);
        atomicAdd(&h[digit], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_radix_take(const double* __restrict__ D, int64_t N, RadixState s, int k,
                                                    unsigned long long* __restrict__ cnt, uint64_t* __restrict__ okey,
                                                    int64_t* __restrict__ oidx) {
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < N; e += int64_t(gridDim.x) * 256) {
        const uint64_t key = order_key_dev(D[e]);
        if (radix_cmp(s, key, uint64_t(e)) > 0) continue;
        const unsigned long long slot = atomicAdd(cnt, 1ull);
        if (slot < (unsigned long long)k) {
            okey[slot] = key;
            oidx[slot] = e;
        }
    }
}

// Indices of the k smallest entries of D (device, N entries), ascending by
// (value, index) under seed_order_key.
int smallest_k_dev(kfc_op* op, const double* D, int64_t N, int64_t k, std::vector<int64_t>& sel) {
    hipStream_t st = op->ctx->stream;
    RadixState s;
    int bits = 1;
    while (bits < 63 && (int64_t(1) << bits) < N) ++bits;
    s.ib = std::max(8, (bits + 7) / 8 * 8);
    DevBuf hb;
    KFC_HIP(op->ctx, hb.ensure(sizeof(unsigned long long) * 257));
    auto* hist = hb.as<unsigned long long>();
    const int nb = diag_grid(N);
    int64_t krem = k;
    std::vector<unsigned long long> h(256);
    for (;;) {
        if (s.nk >= 64 && s.ni >= s.ib) break;   // the prefix names one entry
        KFC_HIP(op->ctx, hipMemsetAsync(hist, 0, sizeof(unsigned long long) * 256, st));
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(256), 0, st, D, N, s, hist);
        KFC_HIP(op->ctx, hipGetLastError());
        KFC_HIP(op->ctx, hipMemcpyAsync(h.data(), hist, sizeof(unsigned long long) * 256, hipMemcpyDeviceToHost, st));
        KFC_HIP(op->ctx, hipStreamSynchronize(st));
        int b = 0;
        int64_t below = 0;
        while (b < 256 && below + int64_t(h[b]) < krem) below += int64_t(h[b++]);
        if (b == 256) return kfc_fail(op->ctx, KFC_EDEVICE, "product_seed: selection lost entries");
        krem -= below;
        if (s.nk < 64) {
            s.pk = (s.pk << 8) | uint64_t(b);
            s.nk += 8;
        } else {
            s.pi = (s.pi << 8) | uint64_t(b);
            s.ni += 8;
        }
        if (int64_t(h[b]) == krem) break;   // every entry matching the prefix is taken
    }
    DevBuf ob;
    const size_t kk = size_t(k);
    KFC_HIP(op->ctx, ob.ensure(sizeof(uint64_t) * kk * 2 + 256));
    auto* okey = ob.as<uint64_t>();
    auto* oidx = reinterpret_cast<int64_t*>(okey + kk);
    KFC_HIP(op->ctx, hipMemsetAsync(hist + 256, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_radix_take, dim3(nb), dim3(256), 0, st, D, N, s, int(k), hist + 256, okey, oidx);
    KFC_HIP(op->ctx, hipGetLastError());
    unsigned long long got = 0;
    std::vector<uint64_t> hk(kk);
    std::vector<int64_t> hi(kk);
    KFC_HIP(op->ctx, hipMemcpyAsync(&got, hist + 256, sizeof(got), hipMemcpyDeviceToHost, st));
    KFC_HIP(op->ctx, hipMemcpyAsync(hk.data(), okey, sizeof(uint64_t) * kk, hipMemcpyDeviceToHost, st));
    KFC_HIP(op->ctx, hipMemcpyAsync(hi.data(), oidx, sizeof(int64_t) * kk, hipMemcpyDeviceToHost, st));
    KFC_HIP(op->ctx, hipStreamSynchronize(st));
    if (got != (unsigned long long)k)
        return kfc_fail(op->ctx, KFC_EDEVICE, "product_seed: selected " + std::to_string(got) + " entries for " +
                                                    std::to_string(k));
    std::vector<std::pair<uint64_t, int64_t>> v(kk);
    for (size_t i = 0; i < kk; ++i) v[i] = {hk[i], hi[i]};
    std::sort(v.begin(), v.end());
    sel.resize(kk);
    for (size_t i = 0; i < kk; ++i) sel[i] = v[i].second;
    return KFC_OK;
}

__global__ __launch_bounds__(256) void k_shift_inverse(const double* __restrict__ D, int64_t N, double sigma,
                                                       double* __restrict__ out) {
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < N; e += int64_t(gridDim.x) * 256)
        out[e] = 1.0 / (D[e] - sigma);
}

}  // namespace

int op_precond_diag_dev(kfc_op* op, bool jacobi, DevBuf& dD) {
    // the :ritz / :jacobi preconditioners' diagonal (src/lobpcg.jl:138-173): D, and for
    // :jacobi 1 / (D - sigma) after it, sigma = minimum(D) - 1
    const int64_t N = op->N;
    KFC_HIP(op->ctx, dD.ensure(sizeof(double) * size_t(N) * (jacobi ? 2 : 1)));
    double dmin = INFINITY;
    int64_t bad = 0;
    int rc = op_full_diag_dev(op, dD.as<double>(), &dmin, &bad);
    if (rc != KFC_OK || !jacobi) return rc;
    if (bad) return kfc_fail(op->ctx, KFC_EARG, "FixedJacobi: the diagonal has non-finite entries");
    const double sigma = dmin - 1;
    if (!(sigma < dmin))
        return kfc_fail(op->ctx, KFC_EARG, "FixedJacobi: the shift minimum(D) - 1 is not representable below minimum(D)");
    return op_shift_inverse_dev(op, dD.as<double>(), sigma, dD.as<double>() + N);
}

int op_shift_inverse_dev(kfc_op* op, const double* D, double sigma, double* out) {
    hipLaunchKernelGGL(k_shift_inverse, dim3(diag_grid(op->N)), dim3(256), 0, op->ctx->stream, D, op->N, sigma, out);
    KFC_HIP(op->ctx, hipGetLastError());
    KFC_HIP(op->ctx, hipStreamSynchronize(op->ctx->stream));
    return KFC_OK;
}

// The k smallest of the Kronecker sum of the per-axis values lam[a] (axis 0 fastest),
// ascending by (seed_order_key(sum), flat index), the sum formed left to right from
// 0 as the host scan does: a best-first walk over the per-axis ascending orders (the
// rounded sum is monotone in each term), continued through the ties of the k-th
// value.  Returns false (nothing decided) when the values are not all finite or the
// ties run past `cap` candidates; the caller then scans all N sums.
bool kron_sum_smallest(const std::vector<std::vector<double>>& lam, int64_t k, size_t cap,
                       std::vector<std::vector<int64_t>>& out) {
    const int K = int(lam.size());
    std::vector<std::vector<int64_t>> ord(K);
    for (int a = 0; a < K; ++a) {
        for (double x : lam[a])
            if (!std::isfinite(x)) return false;
        ord[a].resize(lam[a].size());
        for (size_t i = 0; i < ord[a].size(); ++i) ord[a][i] = int64_t(i);
        std::stable_sort(ord[a].begin(), ord[a].end(), [&](int64_t x, int64_t y) { return lam[a][x] < lam[a][y]; });
    }
    struct Cand {
        double v;
        std::vector<int32_t> r;
    };
    auto value = [&](const std::vector<int32_t>& r) {
        double s = 0;
        for (int a = 0; a < K; ++a) s += lam[a][ord[a][r[a]]];
        return s;
    };
    auto code = [&](const std::vector<int32_t>& r) {
        uint64_t c = 0, m = 1;
        for (int a = 0; a < K; ++a) {
            c += uint64_t(r[a]) * m;
            m *= uint64_t(lam[a].size());
        }
        return c;
    };
    auto cmp = [](const Cand& x, const Cand& y) { return x.v > y.v; };
    std::priority_queue<Cand, std::vector<Cand>, decltype(cmp)> heap(cmp);
    std::unordered_set<uint64_t> seen;
    Cand c0{0, std::vector<int32_t>(K, 0)};
    c0.v = value(c0.r);
    heap.push(c0);
    seen.insert(code(c0.r));
    std::vector<Cand> got;
    double vk = 0;
    while (!heap.empty()) {
        Cand c = heap.top();
        if (int64_t(got.size()) >= k && c.v > vk) break;
        heap.pop();
        got.push_back(c);
        if (int64_t(got.size()) == k) vk = c.v;
        if (got.size() > cap) return false;
        for (int a = 0; a < K; ++a) {
            if (size_t(c.r[a]) + 1 >= lam[a].size()) continue;
            Cand n{0, c.r};
            ++n.r[a];
            if (!seen.insert(code(n.r)).second) continue;
            n.v = value(n.r);
            heap.push(std::move(n));
        }
    }
    if (int64_t(got.size()) < k) return false;
    struct Pick {
        uint64_t key;
        int64_t flat;
        std::vector<int64_t> idx;
    };
    std::vector<Pick> picks;
    picks.reserve(got.size());
    for (const Cand& c : got) {
        Pick p{seed_order_key(c.v), 0, std::vector<int64_t>(K)};
        int64_t st = 1;
        for (int a = 0; a < K; ++a) {
            p.idx[a] = ord[a][c.r[a]];
            p.flat += p.idx[a] * st;
            st *= int64_t(lam[a].size());
        }
        picks.push_back(std::move(p));
    }
    std::sort(picks.begin(), picks.end(),
              [](const Pick& x, const Pick& y) { return x.key != y.key ? x.key < y.key : x.flat < y.flat; });
    out.clear();
    for (int64_t j = 0; j < k; ++j) out.push_back(picks[j].idx);
    return true;
}

int op_full_diag_dev(kfc_op* op, double* D, double* dmin, int64_t* nonfinite) {
    DiagTable T;
    int rc = diag_table(op, true, T);
    if (rc != KFC_OK) return rc;
    DevTable dt;
    rc = upload_table(op, T, dt);
    if (rc != KFC_OK) return rc;
    const int64_t N = op->N;
    const int nb = diag_grid(N);
    DevBuf part;
    KFC_HIP(op->ctx, part.ensure(sizeof(double) * 2 * nb));
    launch_diag<double, 0>(T.g, dt, N, D, part.as<double>(), op->ctx->stream);
    KFC_HIP(op->ctx, hipGetLastError());
    std::vector<double> h(2 * size_t(nb));
    KFC_HIP(op->ctx, hipMemcpyAsync(h.data(), part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, op->ctx->stream));
    KFC_HIP(op->ctx, hipStreamSynchronize(op->ctx->stream));   // the table dies with this call
    double mn = INFINITY, bad = 0;
    for (int b = 0; b < nb; ++b) {
        mn = std::min(mn, h[2 * b]);
        bad += h[2 * b + 1];
    }
    if (dmin) *dmin = mn;
    if (nonfinite) *nonfinite = int64_t(bad);
    return KFC_OK;
}

template <class S>
int op_fused_diag_dev(kfc_op* op, S* out) {
    DiagTable T;
    int rc = diag_table(op, false, T);
    if (rc != KFC_OK) return rc;
    DevTable dt;
    rc = upload_table(op, T, dt);
    if (rc != KFC_OK) return rc;
    launch_diag<S, 1>(T.g, dt, op->N, out, nullptr, op->ctx->stream);
    KFC_HIP(op->ctx, hipGetLastError());
    KFC_HIP(op->ctx, hipStreamSynchronize(op->ctx->stream));
    return KFC_OK;
}
template int op_fused_diag_dev<double>(kfc_op*, double*);
template int op_fused_diag_dev<float>(kfc_op*, float*);
template int op_fused_diag_dev<Cx<double>>(kfc_op*, Cx<double>*);
template int op_fused_diag_dev<Cx<float>>(kfc_op*, Cx<float>*);

int op_fused_diag_range_dev(kfc_op* op, double* dmax, double* dmin, double* imax) {
    DiagTable T;
    int rc = diag_table(op, false, T);
    if (rc != KFC_OK) return rc;
    DevTable dt;
    rc = upload_table(op, T, dt);
    if (rc != KFC_OK) return rc;
    const int nb = diag_grid(op->N);
    DevBuf part;
    KFC_HIP(op->ctx, part.ensure(sizeof(double) * 3 * nb));
    launch_diag<double, 2>(T.g, dt, op->N, nullptr, part.as<double>(), op->ctx->stream);
    KFC_HIP(op->ctx, hipGetLastError());
    std::vector<double> h(3 * size_t(nb));
    KFC_HIP(op->ctx, hipMemcpyAsync(h.data(), part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, op->ctx->stream));
    KFC_HIP(op->ctx, hipStreamSynchronize(op->ctx->stream));
    double mx = -INFINITY, mn = INFINITY, im = 0;
    for (int b = 0; b < nb; ++b) {
        mx = std::max(mx, h[3 * b]);
        mn = std::min(mn, h[3 * b + 1]);
        im = std::max(im, h[3 * b + 2]);
    }
    *dmax = mx;
    *dmin = mn;
    *imax = im;
    return KFC_OK;
}

template <class S>
int op_product_seed_dev(kfc_op* op, int64_t k, S* X, int64_t ldx) {
    const int64_t N = op->N;
    if (k < 1 || k > N) return kfc_fail(op->ctx, KFC_EARG, "product_seed: k must lie in 1:" + std::to_string(N));
    if (k > 65535) return kfc_fail(op->ctx, KFC_EUNSUPPORTED, "product_seed: more than 65535 columns");
    if (ldx < N) return kfc_fail(op->ctx, KFC_EARG, "product_seed: the leading dimension is below the dimension");
    std::vector<std::vector<cd>> U;
    std::vector<std::vector<double>> lam;
    std::vector<bool> ident;
    int rc = op_separable(op, U, lam, ident);
    if (rc != KFC_OK) return rc;
    hipStream_t st = op->ctx->stream;
    const int nbx = int(std::max<int64_t>(1, std::min<int64_t>((N + 255) / 256, 1024)));
    if (std::all_of(ident.begin(), ident.end(), [](bool b) { return b; })) {
        DevBuf dD;
        KFC_HIP(op->ctx, dD.ensure(sizeof(double) * N));
        rc = op_full_diag_dev(op, dD.as<double>(), nullptr, nullptr);
        if (rc != KFC_OK) return rc;
        std::vector<int64_t> sel;
        rc = smallest_k_dev(op, dD.as<double>(), N, k, sel);
        if (rc != KFC_OK) return rc;
        DevBuf ds;
        KFC_HIP(op->ctx, ds.ensure(sizeof(int64_t) * sel.size()));
        KFC_HIP(op->ctx, hipMemcpyAsync(ds.p, sel.data(), sizeof(int64_t) * sel.size(), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL((k_seed_unit<S>), dim3(nbx, unsigned(k)), dim3(256), 0, st, ds.as<int64_t>(), N, X, ldx);
        KFC_HIP(op->ctx, hipGetLastError());
        KFC_HIP(op->ctx, hipStreamSynchronize(st));
        return KFC_OK;
    }
    const int K = op->ndims;
    std::vector<std::vector<int64_t>> picks;
    if (!kron_sum_smallest(lam, k, size_t(1) << 22, picks)) {
        // values that are not all finite, or ties past the walk's bound: scan every sum
        std::vector<int64_t> stride(K, 1);
        for (int a = 1; a < K; ++a) stride[a] = stride[a - 1] * op->dims[a - 1];
        std::priority_queue<std::pair<uint64_t, int64_t>> heap;
        for (int64_t i = 0; i < N; ++i) {
            double s = 0;
            for (int a = 0; a < K; ++a) s += lam[a][(i / stride[a]) % op->dims[a]];
            const std::pair<uint64_t, int64_t> p{seed_order_key(s), i};
            if (int64_t(heap.size()) < k) heap.push(p);
            else if (p < heap.top()) {
                heap.pop();
                heap.push(p);
            }
        }
        std::vector<std::pair<uint64_t, int64_t>> v;
        while (!heap.empty()) {
            v.push_back(heap.top());
            heap.pop();
        }
        std::sort(v.begin(), v.end());
        picks.clear();
        for (const auto& p : v) {
            std::vector<int64_t> idx(K);
            for (int a = 0; a < K; ++a) idx[a] = (p.second / stride[a]) % op->dims[a];
            picks.push_back(idx);
        }
    }
    // the chosen eigenvectors, per column and axis
    std::vector<double> cols;
    std::vector<int64_t> coff(size_t(k) * K);
    for (int64_t j = 0; j < k; ++j)
        for (int a = 0; a < K; ++a) {
            coff[size_t(j) * K + a] = int64_t(cols.size() / 2);
            const int64_t d = op->dims[a];
            const cd* u = &U[a][size_t(picks[j][a]) * d];
            for (int64_t i = 0; i < d; ++i) {
                cols.push_back(u[i].real());
                cols.push_back(u[i].imag());
            }
        }
    DGeom g;
    g.K = K;
    int64_t s = 1;
    for (int a = 0; a < K; ++a) {
        g.stride[a] = s;
        g.dims[a] = op->dims[a];
        s *= op->dims[a];
    }
    DevBuf db;
    const size_t cb = (sizeof(double) * cols.size() + 255) / 256 * 256;
    KFC_HIP(op->ctx, db.ensure(cb + sizeof(int64_t) * coff.size()));
    KFC_HIP(op->ctx, hipMemcpyAsync(db.p, cols.data(), sizeof(double) * cols.size(), hipMemcpyHostToDevice, st));
    auto* dco = reinterpret_cast<int64_t*>(static_cast<char*>(db.p) + cb);
    KFC_HIP(op->ctx, hipMemcpyAsync(dco, coff.data(), sizeof(int64_t) * coff.size(), hipMemcpyHostToDevice, st));
    const auto* dcols = static_cast<const double2*>(db.p);
    if (N <= int64_t(UINT32_MAX))
        hipLaunchKernelGGL((k_seed_product<uint32_t, S>), dim3(nbx, unsigned(k)), dim3(256), 0, st, g, dcols, dco, N, X,
                           ldx);
    else
        hipLaunchKernelGGL((k_seed_product<uint64_t, S>), dim3(nbx, unsigned(k)), dim3(256), 0, st, g, dcols, dco, N, X,
                           ldx);
    KFC_HIP(op->ctx, hipGetLastError());
    KFC_HIP(op->ctx, hipStreamSynchronize(st));   // the tables die with this call
    return KFC_OK;
}
template int op_product_seed_dev<double>(kfc_op*, int64_t, double*, int64_t);
template int op_product_seed_dev<Cx<double>>(kfc_op*, int64_t, Cx<double>*, int64_t);

extern "C" {

int kfc_op_diagonal_device(kfc_op* op, double* D) {
    if (!op || !D) return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_op_diagonal_device: null argument");
    if (op->N == 0) return KFC_OK;
    return op_full_diag_dev(op, D, nullptr, nullptr);
}

int kfc_op_product_seed_device(kfc_op* op, int64_t k, void* X, int64_t ldx) {
    if (!op || !X) return kfc_fail(op ? op->ctx : nullptr, KFC_EARG, "kfc_op_product_seed_device: null argument");
    return op->cplx ? op_product_seed_dev<Cx<double>>(op, k, static_cast<Cx<double>*>(X), ldx)
                    : op_product_seed_dev<double>(op, k, static_cast<double*>(X), ldx);
}

}  // extern "C"
