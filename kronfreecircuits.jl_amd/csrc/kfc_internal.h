// Internal declarations of libkfc: device scalar types, the operator
// structure, and the launch helpers shared by the translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/kfc.h"

// ------------------------------------------------------------------------------------------
// Scalar types.  Complex values are interleaved (re, im) pairs, as the C ABI passes them.

template <class R>
struct Cx {
    R re, im;
};

template <class S> struct RealOf { using type = S; };
template <class R> struct RealOf<Cx<R>> { using type = R; };
template <class S> struct IsCx { static constexpr bool value = false; };
template <class R> struct IsCx<Cx<R>> { static constexpr bool value = true; };

__host__ __device__ inline float s_zero(float) { return 0.f; }
__host__ __device__ inline double s_zero(double) { return 0.0; }
template <class R> __host__ __device__ inline Cx<R> s_zero(Cx<R>) { return Cx<R>{R(0), R(0)}; }

__host__ __device__ inline float s_mul(float a, float b) { return a * b; }
__host__ __device__ inline double s_mul(double a, double b) { return a * b; }
template <class R> __host__ __device__ inline Cx<R> s_mul(Cx<R> a, Cx<R> b) {
    return Cx<R>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline float s_add(float a, float b) { return a + b; }
__host__ __device__ inline double s_add(double a, double b) { return a + b; }
template <class R> __host__ __device__ inline Cx<R> s_add(Cx<R> a, Cx<R> b) { return Cx<R>{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline float s_sub(float a, float b) { return a - b; }
__host__ __device__ inline double s_sub(double a, double b) { return a - b; }
template <class R> __host__ __device__ inline Cx<R> s_sub(Cx<R> a, Cx<R> b) { return Cx<R>{a.re - b.re, a.im - b.im}; }
// a + b*c
__host__ __device__ inline float s_fma(float a, float b, float c) { return fmaf(b, c, a); }
__host__ __device__ inline double s_fma(double a, double b, double c) { return fma(b, c, a); }
__host__ __device__ inline float r_fma(float a, float b, float c) { return fmaf(a, b, c); }
__host__ __device__ inline double r_fma(double a, double b, double c) { return fma(a, b, c); }
template <class R> __host__ __device__ inline Cx<R> s_fma(Cx<R> a, Cx<R> b, Cx<R> c) {
    return Cx<R>{r_fma(b.re, c.re, r_fma(-b.im, c.im, a.re)), r_fma(b.re, c.im, r_fma(b.im, c.re, a.im))};
}
__host__ __device__ inline float s_conj(float a) { return a; }
__host__ __device__ inline double s_conj(double a) { return a; }
template <class R> __host__ __device__ inline Cx<R> s_conj(Cx<R> a) { return Cx<R>{a.re, -a.im}; }
__host__ __device__ inline float s_abs2(float a) { return a * a; }
__host__ __device__ inline double s_abs2(double a) { return a * a; }
template <class R> __host__ __device__ inline R s_abs2(Cx<R> a) { return a.re * a.re + a.im * a.im; }
// real scalar times S
__host__ __device__ inline float s_scale(float r, float a) { return r * a; }
__host__ __device__ inline double s_scale(double r, double a) { return r * a; }
template <class R> __host__ __device__ inline Cx<R> s_scale(R r, Cx<R> a) { return Cx<R>{r * a.re, r * a.im}; }
// conversions from a host complex<double>
template <class S> inline S s_from(std::complex<double> z);
template <> inline float s_from<float>(std::complex<double> z) { return float(z.real()); }
template <> inline double s_from<double>(std::complex<double> z) { return z.real(); }
template <> inline Cx<float> s_from<Cx<float>>(std::complex<double> z) { return Cx<float>{float(z.real()), float(z.imag())}; }
template <> inline Cx<double> s_from<Cx<double>>(std::complex<double> z) { return Cx<double>{z.real(), z.imag()}; }
template <class S> __host__ __device__ inline S s_one();
template <> __host__ __device__ inline float s_one<float>() { return 1.f; }
template <> __host__ __device__ inline double s_one<double>() { return 1.0; }
template <> __host__ __device__ inline Cx<float> s_one<Cx<float>>() { return Cx<float>{1.f, 0.f}; }
template <> __host__ __device__ inline Cx<double> s_one<Cx<double>>() { return Cx<double>{1.0, 0.0}; }

using cd = std::complex<double>;

// ------------------------------------------------------------------------------------------
// Context

// A device allocation owned by the library.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t n) {
        if (n <= bytes) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        else p = nullptr;
        return e;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct kfc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int num_cus = 256;
    kfc_callback cb = nullptr;      // per-iteration solver callback (kfc_ctx_set_callback)
    void* cb_user = nullptr;
    DevBuf pairz;                   // the split dense-pair products' intermediate (grow-only)
};

int kfc_fail(kfc_ctx* ctx, int code, const std::string& msg);
// Hands info to the context's solver callback, if any: KFC_EABORT (with the message set)
// when the callback asks to stop, else KFC_OK.
int ctx_callback(kfc_ctx* ctx, const kfc_iter_info& info, const char* who);

#define KFC_HIP(ctx, call)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return kfc_fail((ctx), KFC_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)


// ------------------------------------------------------------------------------------------
// Operator (host copy of the descriptor plus device data per precision)

struct Factor {
    int axis = 0;
    int kind = KFC_DIAG;
    int64_t d = 0;
    std::vector<int64_t> offsets;
    std::vector<cd> weights;   // d x nbands, column-major
    int64_t shift = 0;
    cd factor = 0;
    std::vector<cd> matrix;    // d x d column-major
    std::vector<cd> diag;      // d
};

struct Term {
    cd coeff = 0;
    std::vector<Factor> factors;
};

// One position of the fused stencil: x at (i0+dx, i1+dy, i2+dz) times
// c * u[i0] * v[i1] * w[i2] (vectors empty = ones).
struct StencilPos {
    bool used = false;
    cd c = 0;
    std::vector<cd> u, v, w;
};

struct StencilPlan {
    bool ok = false;
    int kaxis[3] = {0, -1, -1};   // kernel axis -> operator axis (-1: unit axis)
    int64_t kd[3] = {1, 1, 1};
    StencilPos pos[27];
    std::vector<cd> D0, D1, D2;   // separable diagonal per kernel axis (empty = zero)
    struct DProd { cd c; std::vector<cd> a0, a1, a2; };
    std::vector<DProd> dprod;     // multi-axis diagonal products (<= 4)
    bool need_full_diag = false;
};

// Dense kept-basis terms grouped by axis pair: sum_t c_t A_t (x) B_t on axes (p, q).
struct PairGroup {
    int p = 0, q = 0;
    int64_t dp = 0, dq = 0;
    std::vector<cd> c;
    std::vector<std::vector<cd>> A, B;
};
struct PairPlan {
    bool ok = false;
    std::vector<PairGroup> groups;
};

// Device data of one precision/scalar kind, built on first use.
struct DeviceData {
    bool built = false;
    size_t elem = 0;                // bytes per scalar
    DevBuf diag;                    // N (only when the chain path or a non-separable diagonal needs it)
    std::vector<std::unique_ptr<DevBuf>> blobs;  // factor and plan tables
    // kept terms: per factor, device pointers
    struct FDev { const void* data = nullptr; const int64_t* offs = nullptr; };
    std::vector<std::vector<FDev>> fdev;
    // stencil plan tables
    const void* pu[27] = {}; const void* pv[27] = {}; const void* pw[27] = {};
    const void* D0 = nullptr; const void* D1 = nullptr; const void* D2 = nullptr;
    const void* dp0[4] = {}; const void* dp1[4] = {}; const void* dp2[4] = {};
    DevBuf buf1, buf2;              // chain scratch
    std::vector<std::vector<const void*>> pair_ptr;   // per pair group: A0, B0, A1, B1, ...
};

struct kfc_op {
    kfc_ctx* ctx = nullptr;
    int ndims = 0;
    std::vector<int64_t> dims;
    int64_t N = 0;
    bool cplx = false;
    cd constant = 0;
    std::vector<Term> diagterms;    // fused into the diagonal (the constant as a zero-factor term)
    std::vector<Term> terms;        // kept
    int64_t capacity = 1;
    bool hermitian = true;
    std::vector<cd> hdiag;          // host fused diagonal (built on demand)
    StencilPlan plan;
    int stencil_variant = 1;        // 0: LDS plane tile, 1: register queue (chosen by the tuner)
    int stencil_march = 0;          // march chunk along the slowest axis (0: default)
    bool stencil_pinned = false;    // launch shape set by the caller: no autotuning
    int64_t tuned_cols[3] = {-1, -1, -1};   // column count the stencil variant was tuned for: FP64, FP32,
                                            // and the complex (promoted) apply of a real operator
    int variant_z = -1, march_z = 0;        // the promoted apply's own tuned launch shape
    PairPlan pairs;
    DevBuf dotpart;                 // block partials of the fused apply dots (apply.hip)
    DevBuf evo[4];                  // scratch of the time-evolution applies (evolve.hip)
    DeviceData dd[2];               // [KFC_PREC_F64], [KFC_PREC_F32]
    DeviceData ddz;                 // complex FP64 tables of a real stencil operator (complex kets)
    double gershgorin = 0;
    double gershgorin_lo = 0;       // lower end of the Gershgorin interval
    bool have_gershgorin = false;
    // separable_eigenbases, memoized as the reference memoizes H.separable
    bool have_separable = false;
    std::vector<std::vector<cd>> sep_U;
    std::vector<std::vector<double>> sep_lam;
    std::vector<bool> sep_ident;
};

// Gershgorin interval [lo, hi] holding the spectrum (kfc_api.hip)
int op_spectral_interval(kfc_op* op, double* lo, double* hi);

// Builders and launches (apply.hip)
int op_build_plan(kfc_op* op);
int op_build_pairs(kfc_op* op);
template <class T> int pairs_upload(kfc_op* op, DeviceData& dd);
// scratch: the split form of a group may use the context's scratch (grown here); false
// for callers that must not allocate (every group then runs whole).
template <class T>
int launch_pairs(kfc_op* op, const void* X, void* Y, int64_t ncols, int64_t ld, hipStream_t st, bool scratch = true);
// Grows the context's scratch for n problems of ncols columns at precision prec (0: FP64)
// when a group of op can run in the split form; a graph that captures pair launches
// reserves it first, since filling a table allocates nothing.  If the device cannot hold
// it, the groups run whole (KFC_OK either way).
int pairs_reserve(kfc_op* op, int n, int64_t ncols, int prec);
// Many operators of one pair structure in one launch per group (dense_pair.hip):
// problem q maps X[q] -> Y[q]; host_tab (pinned) and dev_tab hold n x groups
// launch records of pair_args_bytes(prec) bytes each.
bool pairs_same_structure(const kfc_op* a, const kfc_op* b);
size_t pair_args_bytes(int prec);
template <class T>
int launch_pairs_batch(kfc_op* const* ops, int n, const void* const* X, void* const* Y, int64_t ncols, int64_t ld,
                       hipStream_t st, void* host_tab, void* dev_tab);
// The same in two parts: the host table filled, then its upload and the launches (a
// graph captures the second part once and replays it after each refill of host_tab;
// the shapes must have been tuned by an earlier launch of the same key).
template <class T>
void pairs_fill_table(kfc_op* const* ops, int n, const void* const* X, void* const* Y, int64_t ncols, int64_t ld,
                      void* host_tab);
template <class T>
int launch_pairs_filled(kfc_op* const* ops, int n, hipStream_t st, void* host_tab, void* dev_tab);
int op_host_diag(kfc_op* op);

// ---- derived quantities (host: kfc_api.hip, apply.hip; device: seeds.hip) ----------------
// The order in which Julia's isless sorts doubles (-0.0 before 0.0, NaN last) as an
// unsigned key: the order of the product seed's selection.
inline uint64_t seed_order_key(double d) {
    if (d != d) return ~0ull;
    uint64_t u;
    std::memcpy(&u, &d, sizeof u);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
std::vector<cd> factor_diagonal(const Factor& f);        // a fused (diagonal) factor's entries
std::vector<cd> kept_factor_diagonal(const Factor& f);   // the diagonal of a kept factor, dense(op)[i, i]
int op_separable(kfc_op* op, std::vector<std::vector<cd>>& U, std::vector<std::vector<double>>& lam,
                 std::vector<bool>& ident);
int op_full_diagonal(kfc_op* op, std::vector<double>& D);
int op_product_seed(kfc_op* op, int64_t k, std::vector<cd>& X);
// The k smallest of the Kronecker sum of the per-axis values (indices per axis), or
// false when it cannot decide without scanning every sum.
bool kron_sum_smallest(const std::vector<std::vector<double>>& lam, int64_t k, size_t cap,
                       std::vector<std::vector<int64_t>>& out);
// Device: the full real diagonal into D (N doubles) with its minimum and the count of
// non-finite entries; the fused diagonal as S; the fused diagonal's max/min real part and
// max |imaginary part|; the product seed into X (N x k, leading dimension ldx), S the
// operator's FP64 scalar.  Each synchronizes the context's stream before returning.
int op_full_diag_dev(kfc_op* op, double* D, double* dmin, int64_t* nonfinite);
template <class S> int op_fused_diag_dev(kfc_op* op, S* out);
int op_fused_diag_range_dev(kfc_op* op, double* dmax, double* dmin, double* imax);
template <class S> int op_product_seed_dev(kfc_op* op, int64_t k, S* X, int64_t ldx);
// The LOBPCG preconditioner's diagonal D (and 1 / (D - minimum(D) + 1) after it with
// jacobi) in dD, with the :jacobi shift's checks.
int op_precond_diag_dev(kfc_op* op, bool jacobi, DevBuf& dD);
// out = 1 / (D - sigma) over N device entries
int op_shift_inverse_dev(kfc_op* op, const double* D, double sigma, double* out);
int op_ensure_device(kfc_op* op, int prec);
int op_apply(kfc_op* op, const void* X, void* Y, int64_t ncols, int64_t ld, int prec, hipStream_t st);
// Y = alpha H X + beta X (+ gamma A) for complex128 blocks and a real operator on the
// fused stencil path (op->plan.ok), the stencil running on complex tables (op->ddz).  A
// may be Y itself: every stencil variant reads A only at the entry it then writes.
int op_apply_promoted(kfc_op* op, const void* X, void* Y, int64_t ncols, int64_t ld, cd alpha, cd beta,
                      hipStream_t st, const void* A = nullptr,
                      cd gamma = cd(0));
// The argument checks of kfc_apply (precision, sizes, aliasing); a KFC_PREC_Z64 precision
// on a complex operator becomes KFC_PREC_F64, on a real one sets real_on_complex.
int check_apply_args(kfc_op* op, const void* X, const void* Y, int64_t ncols, int64_t ld, int32_t& precision,
                     bool& real_on_complex, const char* who);
// Bytes of chain-path scratch one apply of ncols columns at precision prec needs (0 on the
// fused-stencil and dense-pair paths, and when no kept term has two factors).
size_t op_chain_scratch_bytes(const kfc_op* op, int64_t ncols, int prec);
// Y = H X with no device allocation and no tuning: the tables of precision prec must be on
// the device already, and the chain path's two intermediate blocks are taken from scratch
// (op_chain_scratch_bytes, the second block at the next 256-byte boundary after the first).
int op_apply_fixed(kfc_op* op, const void* X, void* Y, int64_t ncols, int64_t ld, int prec, void* scratch,
                   hipStream_t st);
// Takes the current device's address of kfc_apply_noalloc's scratch array (noalloc.hip).
int noalloc_setup(kfc_ctx* ctx);
// Y = H X for interleaved complex128 blocks X, Y (leading dimension ld) and a real H
// (evolve.hip): the reference's _mul_real_operator!.
int op_apply_real_on_complex(kfc_op* op, const void* X, void* Y, int64_t ncols, int64_t ld);
// Chebyshev step fused into the apply: C = a*(H B) + b*B + g*A  (A may be null)
int op_apply_cheb(kfc_op* op, const void* B, const void* A, void* C, int64_t ncols, int64_t ld, int prec,
                  double a, double b, double g, hipStream_t st);
// One column: op_apply_cheb that also leaves sum B^2 and sum B C (FP64) in dots[0], dots[1]
// (device) when the fused plane sweep runs the apply (fused = true); otherwise dots are
// not written (fused = false).  Real blocks only.
int op_apply_cheb_dots(kfc_op* op, const void* B, const void* A, void* C, int64_t ld, int prec, double a, double b,
                       double g, double* dots, bool& fused, hipStream_t st);

// kfc_lobpcg with the context's callback on or off (solvers.hip); the lockstep sweep
// re-solves collapsed points through it with the callback off.
int lobpcg_solve(kfc_op* op, int64_t m, int64_t guard, double tol, int64_t maxiter, int32_t precond, double eta,
                 int32_t seed_kind, const double* X0, uint64_t seed, int32_t mode, double* E, double* V,
                 kfc_stats* stats, bool use_callback);

// Host linear algebra (linalg.cpp): Hermitian eigendecomposition, ascending.
// A is n x n column-major (overwritten); w[n]; Z[n x n] eigenvectors.
void herm_eig(int n, std::vector<cd>& A, std::vector<double>& w, std::vector<cd>& Z);
void sym_eig(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z);
// The lowest k eigenpairs (w: all n eigenvalues ascending; Z: n x k), by inverse
// iteration on the tridiagonal form: several times faster than sym_eig for k << n.
void sym_eig_lowest(int n, int k, std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z);
