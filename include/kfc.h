/* kfc.h — C ABI of libkfc.so, the MI355X (gfx950) Kronecker-free apply and
 * eigensolver library.
 *
 * Every entry point takes plain pointers and sizes.  Host descriptors are
 * copied to the device at creation; device buffers passed to kfc_apply are
 * owned by the caller; the library never frees caller memory.  Status codes
 * are returned as int (KFC_OK == 0); kfc_last_error(ctx) holds the message of
 * the last failure on that context.
 *
 * Each declaration names the KronFreeCircuits.jl interface it replaces
 * (file:line in the reference repository).
 */
#ifndef KFC_H
#define KFC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the error classes the reference raises) ----------------------- */
enum kfc_status {
    KFC_OK = 0,
    KFC_EARG = 1,        /* ArgumentError (src/lobpcg.jl:224-233, src/apply.jl:225-231) */
    KFC_EDIM = 2,        /* DimensionMismatch (src/apply.jl:10-13, 410-411) */
    KFC_ENOCONV = 3,     /* "did not converge" ErrorException (src/lobpcg.jl:292-294, src/solvers.jl:97-98) */
    KFC_EBREAKDOWN = 4,  /* eigensolver breakdown (src/lobpcg.jl:195-198, src/chebyshev.jl:124-125) */
    KFC_EDEVICE = 5,     /* HIP runtime failure */
    KFC_ENOMEM = 6,      /* device allocation failure */
    KFC_EUNSUPPORTED = 7, /* a combination this build does not serve */
    KFC_EABORT = 8       /* the solver callback asked to stop: the reference's callback threw,
                            and its exception propagates with no retry (src/lobpcg.jl:247-253, 279) */
};

/* ---- element types: the operator's type T (src/apply.jl:232-233) ---------------- */
enum kfc_dtype { KFC_F64 = 0, KFC_C128 = 1 };

/* ---- arithmetic precision of an application (FP64 or the FP32 bulk path) --------- */
/* KFC_PREC_Z64: interleaved complex128 blocks whatever the operator's type; on a real
 * (KFC_F64) operator this is _mul_real_operator! (src/apply.jl:397-407), H applied to
 * the real and imaginary parts, on a complex one the same as KFC_PREC_F64. */
enum kfc_precision { KFC_PREC_F64 = 0, KFC_PREC_F32 = 1, KFC_PREC_Z64 = 2 };

/* ---- solver precision mode ------------------------------------------------------- */
enum kfc_mode {
    KFC_MODE_F64 = 0,    /* every apply and reduction in FP64 (the plain path) */
    KFC_MODE_MIXED = 1   /* FP32 applies for the bulk iteration, FP64 Rayleigh-Ritz,
                            reorthogonalization and a final FP64 refinement */
};

/* ---- local operator kinds (src/localops.jl:18-73) -------------------------------- */
enum kfc_kind { KFC_DIAG = 0, KFC_BANDED = 1, KFC_SHIFT = 2, KFC_DENSE = 3 };

/* One factor (axis, op) of an AxisTerm (src/apply.jl:157-166).  Complex data is
 * interleaved (re, im) doubles; real data is plain doubles. */
typedef struct {
    int32_t axis;            /* 0-based axis, strictly increasing within a term */
    int32_t kind;            /* enum kfc_kind */
    int64_t d;               /* dimension of the op == dims[axis] */
    int32_t is_complex;      /* 1: the data arrays below are interleaved complex */
    int32_t nbands;          /* BANDED: number of bands */
    const int64_t* offsets;  /* BANDED: nbands offsets o_b, Y[i] += w_b[i] X[i+o_b] */
    const double* weights;   /* BANDED: d x nbands, column-major (src/apply.jl:82-88) */
    int64_t shift;           /* SHIFT: s, Y[i+s] = f X[i] */
    double factor_re, factor_im; /* SHIFT: f */
    const double* matrix;    /* DENSE: d x d column-major */
    const double* diag;      /* DIAG: d entries */
} kfc_factor_desc;

/* AxisTerm(coeff, factors) (src/apply.jl:157-166). */
typedef struct {
    double coeff_re, coeff_im;
    int32_t nfactors;
    const kfc_factor_desc* factors;
} kfc_term_desc;

/* KronFreeOperator(dims, terms; constant, block_capacity, hermitian)
 * (src/apply.jl:222-253).  Terms whose factors are all DIAG (or none) are fused
 * with the constant into the diagonal, the rest are kept, as the reference does. */
typedef struct {
    int32_t ndims;
    const int64_t* dims;     /* axis 0 fastest (column-major, src/apply.jl:18-19) */
    int32_t dtype;           /* enum kfc_dtype: KFC_F64 on a real node, KFC_C128 otherwise */
    double constant_re, constant_im;
    int32_t nterms;
    const kfc_term_desc* terms;
    int64_t block_capacity;  /* solver block width reserved for scratch */
    int32_t hermitian;
} kfc_op_desc;

/* Work counters returned by the solvers; they feed node_timings
 * (src/materialize.jl:674-684, 801-809). */
typedef struct {
    int64_t iterations;      /* LOBPCG/Chebyshev iterations, Lanczos restarts */
    int64_t n_matvec;        /* H columns applied (all precisions) */
    int64_t n_matvec_f32;    /* of which in FP32 */
    int64_t n_apply_calls;   /* block apply launches */
    double seconds_total;
    double seconds_apply;     /* host wall time around the applies */
    double seconds_apply_gpu; /* HIP-event time of the applies on the context stream (-1: the
                                 applies ran inside captured step graphs, untimed) */
    double max_residual;     /* worst fresh FP64 true residual of the returned pairs */
    int32_t n_converged;
    int32_t warm;            /* 1 when an X0 was used */
    int32_t degree_last;     /* Chebyshev: degree of the last filter */
    int32_t retried;         /* LOBPCG: 1 when the random-block retry ran */
    double seconds_ortho_gpu; /* Lanczos: HIP-event time of the matrix-core block passes
                                 (orthogonalization, Cholesky scaling, restart) */
    double bytes_ortho;       /* Lanczos: their algorithmic bytes (columns read + written) */
    int64_t n_ortho_calls;    /* Lanczos: their launches */
} kfc_stats;

typedef struct kfc_ctx kfc_ctx;
typedef struct kfc_op kfc_op;

/* ---- context (device + stream) ---------------------------------------------------- */
int kfc_ctx_create(kfc_ctx** ctx, int device);
int kfc_ctx_destroy(kfc_ctx* ctx);
const char* kfc_last_error(const kfc_ctx* ctx);
/* The HIP stream every call of this context runs on (a hipStream_t). */
void* kfc_ctx_stream(kfc_ctx* ctx);

/* ---- per-iteration solver callback ----------------------------------------------------
 * The record a solver hands its callback once per iteration: the fields of the `info`
 * the reference passes to `callback(info)` from LOBPCG (n_iter, n_matvec, n_locked,
 * lambdas of the active columns; LOBPCGEigensolver 0.1.4, src/lobpcg.jl:1-8, 144-157)
 * and from ChebyshevFilter (iteration, lambda, residuals, n_locked, lambda_cut,
 * lambda_max, degree; src/chebyshev.jl:353-363).  The arrays are valid during the call. */
typedef struct {
    int32_t solver;          /* 0 = LOBPCG, 1 = ChebyshevFilter */
    int32_t precision;       /* KFC_PREC_F64, or KFC_PREC_F32 in the mixed path's bulk phase */
    int64_t iteration;       /* LOBPCG: n_iter (1 at the first call); Chebyshev: filters applied so far */
    int64_t n_matvec;        /* H columns applied so far in this solve */
    int32_t width;           /* entries of lambda and residuals (the block width) */
    int32_t n_locked;        /* converged columns (a prefix), counted this iteration */
    int32_t active0;         /* first column active during the iteration (LOBPCG's lambdas = lambda[active0:]) */
    int32_t degree;          /* ChebyshevFilter: degree of the last filter (0 before the first) */
    const double* lambda;    /* Ritz values, ascending within the active columns */
    const double* residuals; /* residual norms ||H x - theta x|| */
    double lambda_cut;       /* ChebyshevFilter: lower end of the damped interval (NaN before the first filter) */
    double lambda_max;       /* ChebyshevFilter: upper end of the filter interval */
} kfc_iter_info;

/* Returns 0 to continue; any other value stops the solve with KFC_EABORT. */
typedef int (*kfc_callback)(const kfc_iter_info* info, void* user);

/* Replaces the `callback` keyword of LOBPCG and ChebyshevFilter (src/tree.jl:66-87,
 * 293-321): kfc_lobpcg and kfc_chebyshev on this context call cb(info, user) once per
 * iteration from the calling host thread (both phases of the mixed mode).  NULL clears
 * it.  The batched and lockstep solvers do not call it. */
int kfc_ctx_set_callback(kfc_ctx* ctx, kfc_callback cb, void* user);

/* ---- operator ------------------------------------------------------------------- */
/* Replaces KronFreeOperator(dims, terms; constant, block_capacity, on)
 * (src/apply.jl:222-253) with on = a device array type, and _on_storage
 * (src/apply.jl:258-267): host data copied to the device once. */
int kfc_op_create(kfc_ctx* ctx, const kfc_op_desc* desc, kfc_op** op);
int kfc_op_destroy(kfc_op* op);
/* prod(dims) */
int64_t kfc_op_size(const kfc_op* op);
/* 1 when every application of op runs as one fused stencil pass */
int kfc_op_fused(const kfc_op* op);
/* Fused-stencil launch shape (no counterpart in the reference: the choice of
 * the project's autotuner).  variant 0 = LDS tile, 1 = register queue,
 * 2 = plane sweep, 3 = plane sweep fed by LDS-DMA loads, 4 = plane sweep
 * loading two points a lane, 5 = as 4 with the stencil positions fixed when the
 * kernel is compiled (19 or all 27) and buffer addressing (2-5: constant weights
 * only, others fall back to 1; 4 and 5 on real, aligned blocks only, else 2; 5 on
 * columns below 2^31 bytes, else 4); march = planes per block along the last axis.  Setting variant -1 returns the choice to the autotuner, which
 * times the candidates on the first application of each column count. */
int kfc_op_stencil_get(const kfc_op* op, int32_t* variant, int32_t* march);
int kfc_op_stencil_set(kfc_op* op, int32_t variant, int32_t march);
/* Fills every scratch block the operator holds (the chain path's two intermediate
 * blocks at both precisions, the fused dots' partials, the time-evolution scratch)
 * with `value`: the reference's device tests fill H.buf1 / H.buf2 with NaN and
 * 1e300 before applying (test/cuda/test_device.jl:163-207); here the scratch is
 * internal, so the tests poison it through this call.  No effect on results. */
int kfc_op_fill_scratch(kfc_op* op, double value);
/* Algorithmic work of one column of kfc_apply on the path the operator takes
 * (no counterpart in the reference; the roofline denominators): *bytes = HBM
 * bytes that must move (read x, write y, plus the stored diagonal when the path
 * reads it), *flops = floating-point operations (2 per multiply-add; the dense
 * kept-basis pairs count 2 N (d_p + d_q) per merged term). */
int kfc_op_work(const kfc_op* op, double* bytes, double* flops);

/* Replaces LinearAlgebra.mul!(y, H, x) (src/apply.jl:416-432):
 * Y = H X for ncols columns with leading dimension ld (>= prod(dims)).
 * X, Y are device pointers of FP64 (or interleaved complex128) when precision is
 * KFC_PREC_F64, FP32 (complex64) when KFC_PREC_F32, complex128 for KFC_PREC_Z64 (a real
 * operator on complex blocks, src/apply.jl:397-407).  Y must not alias X
 * (ArgumentError, src/apply.jl:412).  Enqueued on the context stream; returns
 * after the launch (kfc_sync waits). */
int kfc_apply(kfc_op* op, const void* X, void* Y, int64_t ncols, int64_t ld, int32_t precision);

/* No counterpart in the reference: kfc_apply for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and result as kfc_apply; it calls no hipMalloc, hipFree or
 * allocating routine.  The chain path's two intermediate blocks come from a static
 * device array of 64 MiB that every GPU the library runs on holds until the process
 * ends: 2 x prod(dims) x ncols x element size (plus 256 bytes) must fit, else KFC_EARG;
 * the fused-stencil path needs no scratch, and the dense-pair path runs every group as
 * one kernel (the split form kfc_apply may pick, which takes context scratch, is not
 * used here).  The operator's tables for the
 * precision must be on the device (kfc_op_create places the FP64 ones, the first FP32
 * kfc_apply the FP32 ones), the stencil keeps the launch shape the tuner last chose, and
 * complex blocks on a real operator (KFC_PREC_Z64) are refused (KFC_EUNSUPPORTED).  All
 * calls on one GPU share the array, so calls that overlap on one GPU race. */
int kfc_apply_noalloc(kfc_op* op, const void* X, void* Y, int64_t ncols, int64_t ld, int32_t precision);

/* One Chebyshev step of a real operator with its two inner products:
 * C = a H B + b B + g A (A may be null) on one column, and host dots[0] = B.B,
 * dots[1] = B.C (FP64 sums).  The recurrence is _chebyshev_step!
 * (src/chebyshev.jl:25-30); the dots are the moments q^T T_j(H) q that the
 * s-step Lanczos (kfc_lanczos) takes from its block applies, via
 * T_i T_i = (T_2i + T_0) / 2 and T_i T_(i+1) = (T_2i+1 + T_1) / 2, in place of the
 * reference's separate inner products (KrylovKit's Lanczos reorthogonalization).
 * *fused (may be null) is 1 when the fused plane sweep computed the dots in the
 * apply's own pass, 0 when they came from a separate pass.  C must not alias B.
 * Synchronizes the context stream. */
int kfc_apply_cheb_dots(kfc_op* op, const void* B, const void* A, void* C, double a, double b, double g,
                        int32_t precision, double* dots, int32_t* fused);
int kfc_sync(kfc_ctx* ctx);

/* Replaces diagonal(H) / _full_diagonal(H) (src/apply.jl:455-478): the real
 * diagonal of H written to host out[prod(dims)]. */
int kfc_op_diagonal(kfc_op* op, double* out);
/* Replaces gershgorin_bound(H) (src/apply.jl:586-593). */
int kfc_op_gershgorin(kfc_op* op, double* out);
/* Replaces product_seed(H, k) (src/apply.jl:655-685): k orthonormal product
 * states into host X[prod(dims) x k] (complex interleaved for KFC_C128). */
int kfc_op_product_seed(kfc_op* op, int64_t k, double* X);
/* Replaces _device_diagonal / _full_diagonal(H) on the device (src/apply.jl:455-478):
 * the real diagonal of H into device D[prod(dims)], formed on the device by the
 * same term-by-term arithmetic as kfc_op_diagonal (equal bit for bit).  Returns
 * after the kernel finished. */
int kfc_op_diagonal_device(kfc_op* op, double* D);
/* Replaces product_seed(H, k) (src/apply.jl:655-685) with the block built on the
 * device: device X[prod(dims) x k], leading dimension ldx >= prod(dims), FP64 (or
 * interleaved complex128 for KFC_C128); 1 <= k <= min(prod(dims), 65535).  The k
 * smallest summed energies are found without forming the prod(dims) sums (a
 * best-first walk of the per-axis orders); when every axis is already an eigenbasis,
 * the k smallest diagonal entries are selected on the device (radix select).  Equal
 * to kfc_op_product_seed bit for bit.  Returns after the kernels finished. */
int kfc_op_product_seed_device(kfc_op* op, int64_t k, void* X, int64_t ldx);

/* ---- the linear solve on the complement (the derivatives' Sternheimer step) ------- */
/* Replaces _complement_solve(H, D, basis, B, shifts; tol, maxiter, eta, scale)
 * (src/resolvent.jl:146-256): X[:, j] solves Q (H - shifts[j]) Q x = Q B[:, j] with x
 * orthogonal to the k orthonormal columns of basis (Q = 1 - basis basis^H), by
 * preconditioned CG on all n columns in lockstep (1 <= n <= 64), preconditioner
 * 1 / hypot(D - shifts[j], eta), the reference's stopping rules (each column to
 * tol |Q B_j|, or to the round-off floor 1000 eps scale[j]).  Device: D (prod(dims)
 * doubles, H's real diagonal), basis (prod(dims) x k), B and X (prod(dims) x n), the
 * operator's FP64 scalar, leading dimension prod(dims).  Host: shifts[n], scale[n] (NULL:
 * the column norms of B), and the outputs iterations, residuals[n] (true residual
 * relative to |Q B_j| for a column that met tol, else 0), flags[n] (bit 0 met tol,
 * bit 1 converged from the start with x = 0, bit 2 stopped at the round-off floor) and
 * floor_residuals[n] (relative to scale[j] for a column at the floor, else 0); the
 * output arrays may be NULL.  KFC_EBREAKDOWN: H - shifts[j] is not positive on the
 * complement; KFC_ENOCONV: a column did not converge. */
int kfc_complement_solve(kfc_op* op, const double* D, const void* basis, int64_t k, const void* B, int64_t n,
                         const double* shifts, double tol, int64_t maxiter, double eta, const double* scale, void* X,
                         int64_t* iterations, double* residuals, int32_t* flags, double* floor_residuals);

/* ---- matrix-free time evolution (the port_operator consumers) -------------------- */
/* Replaces KronFreeApply (ext/KronFreeCircuitsQuantumToolboxExt.jl:44-87), the apply
 * behind drive_hamiltonian(...; basis=:product, frame=:lab): W = alpha scale
 * (H - shift) V + beta W for complex128 device blocks V, W (interleaved, prod(dims)
 * x ncols, leading dimension ld, 1 <= ncols <= 4096, W not overlapping V).  A real H
 * on the fused stencil path runs the stencil on complex tables (one pass when beta is
 * zero); any other real H applies once to the real block [Re V | Im V].  beta == 0
 * leaves W unread.  Enqueued on the context stream (kfc_sync waits). */
int kfc_apply_affine(kfc_op* op, const void* V, void* W, int64_t ncols, int64_t ld, double alpha_re, double alpha_im,
                     double scale, double shift, double beta_re, double beta_im);
/* Replaces InteractionApply (ext/KronFreeCircuitsQuantumToolboxExt.jl:211-251), the
 * apply behind drive_hamiltonian(...; basis=:product, frame=:interaction): with
 * ph = exp(-i Omega t), W = alpha conj(ph) .* [2 pi (H0 - E0) - diag Omega +
 * sum_k f_k 2 pi O_k] (ph .* V) + beta W.  Omega is a device vector of prod(dims)
 * doubles (2 pi (diag(H0) - E0) in the reference), f holds nops complex coefficients
 * (re, im), ops the drive operators (port_operator) on H0's context and product
 * basis.  Enqueued on the context stream (kfc_sync waits). */
int kfc_apply_interaction(kfc_op* H0, double E0, const double* Omega, double t, kfc_op* const* ops, const double* f,
                          int32_t nops, const void* V, void* W, int64_t ncols, int64_t ld, double alpha_re,
                          double alpha_im, double beta_re, double beta_im);

/* ---- solvers (the Solver plugin contract, src/solvers.jl:3-31) ------------------- */
/* Each writes E[m] ascending and V[prod(dims) x m] (host, column-major,
 * complex interleaved for KFC_C128) and fills stats (may be NULL). */

/* Replaces lanczos_solve(::Lanczos, H, m) (src/solvers.jl:85-106):
 * thick-restart Lanczos, krylovdim = max(2m+10, krylovdim), residual tol;
 * verbosity > 0 prints a progress line to stderr every `verbosity` restarts. */
int kfc_lanczos(kfc_op* op, int64_t m, int64_t krylovdim, double tol, int64_t maxiter, uint64_t seed,
                int32_t mode, int32_t verbosity, double* E, double* V, kfc_stats* stats);

/* Replaces lobpcg_solve(::LOBPCG, H, m; X0) (src/lobpcg.jl:217-301).
 * precond: 0 = :ritz, 1 = :jacobi, 2 = :none; seed_kind: 0 = :product, 1 = :random.
 * X0: NULL or host prod(dims) x m (the warm start, src/materialize.jl:656-666). */
enum kfc_precond { KFC_PRECOND_RITZ = 0, KFC_PRECOND_JACOBI = 1, KFC_PRECOND_NONE = 2 };
/* seed = :product / :random (src/tree.jl:66-87).  KFC_SEED_WARM_GUARDS (no
 * counterpart in the reference) says X0 holds m + g columns, the guard columns
 * included, instead of m columns with product-state guards: a learned warm
 * start supplies all the columns the solver converges. */
enum kfc_seed { KFC_SEED_PRODUCT = 0, KFC_SEED_RANDOM = 1, KFC_SEED_WARM_GUARDS = 2 };
int kfc_lobpcg(kfc_op* op, int64_t m, int64_t guard, double tol, int64_t maxiter, int32_t precond,
               double eta, int32_t seed_kind, const double* X0, uint64_t seed, int32_t mode,
               double* E, double* V, kfc_stats* stats);

/* A parameter sweep of independent LOBPCG solves (as materialize runs one per
 * sweep point, bench/common.jl:135-163): n operators, each created on one of a few
 * contexts; one host thread per distinct context, the contexts' streams running
 * concurrently.  X0s: NULL or n pointers (each NULL or prod(dims) x m); E[n x m];
 * Vs: n host arrays; stats: NULL or n records. */
int kfc_lobpcg_batch(int32_t n, kfc_op* const* ops, int64_t m, int64_t guard, double tol, int64_t maxiter,
                     int32_t precond, double eta, int32_t seed_kind, const double* const* X0s, uint64_t seed,
                     int32_t mode, double* E, double* const* Vs, kfc_stats* stats);

/* The same sweep in lockstep: n real operators of one size on ONE context (the
 * sweep points of one node: same dims and term structure), every LOBPCG step one
 * launch for all points, the contract of lobpcg_solve (src/lobpcg.jl:217-301) per
 * point.  m + guard <= 21.  Arguments and outputs as kfc_lobpcg_batch; a point
 * whose block collapses is re-solved by kfc_lobpcg. */
int kfc_lobpcg_lockstep(int32_t n, kfc_op* const* ops, int64_t m, int64_t guard, double tol, int64_t maxiter,
                        int32_t precond, double eta, int32_t seed_kind, const double* const* X0s, uint64_t seed,
                        int32_t mode, double* E, double* const* Vs, kfc_stats* stats);

/* The lockstep sweep started from a learned subspace.  No counterpart in the
 * reference: it supplies the X0 of src/lobpcg.jl:217 (src/materialize.jl:656-666)
 * for every sweep point at once.  Point i starts from the lowest m + g Ritz
 * vectors of H_i on span(P_i B C_i): B (DEVICE, nf x nb, column-major) spans a
 * parameter-independent frame of the parent basis, P_i puts frame row f at
 * parent row rowmaps[i][f] (dropped when -1; host arrays of nf entries, no two
 * rows alike), C_i (host, nb x ns) selects ns directions, or C = NULL for all nb
 * (ns is then nb).  The start is perturbed as a warm start; everything else as
 * kfc_lobpcg_lockstep. */
int kfc_lobpcg_lockstep_subspace(int32_t n, kfc_op* const* ops, int64_t m, int64_t guard, double tol,
                                 int64_t maxiter, int32_t precond, double eta, const double* B, int64_t nf,
                                 int32_t nb, const int64_t* const* rowmaps, const double* const* C, int32_t ns,
                                 uint64_t seed, int32_t mode, double* E, double* const* Vs, kfc_stats* stats);

/* Replaces solve(::ChebyshevFilter, H, m) (src/chebyshev.jl:310-406).
 * ortho: 0 = :svqb, 1 = :cholqr; bound: a finite number, or NaN for :gershgorin. */
int kfc_chebyshev(kfc_op* op, int64_t m, int32_t degree, int64_t extra, int64_t maxiter,
                  int32_t rescale_every, int32_t ortho, double bound, int32_t seed_kind, double tol,
                  uint64_t seed, int32_t mode, double* E, double* V, kfc_stats* stats);

/* Replaces solve(::Dense, H, m) (src/solvers.jl:44-50): Matrix(H) by applying H
 * to identity blocks on the device, then a Hermitian eigendecomposition.
 * m == 0 returns all prod(dims) pairs. */
int kfc_dense(kfc_op* op, int64_t m, double* E, double* V, kfc_stats* stats);

/* solve(::Dense, H, m) for each of n operators (the tops of a batched sweep, which
 * the reference solves one after another in materialize): Matrix(H) of each
 * operator of at most 256 states formed on the host from its terms and
 * diagonalized on host threads, larger ones as kfc_dense.  1 <= m <= prod(dims)
 * of every operator; E[n x m]; Vs: n host arrays; stats: NULL or n records. */
int kfc_dense_batch(int32_t n, kfc_op* const* ops, int64_t m, double* E, double* const* Vs, kfc_stats* stats);

/* Fresh FP64 true residuals ||H v_j - E_j v_j|| of host pairs (E, V[N x m]):
 * _worst_true_residual per column (src/solvers.jl:65-69). */
int kfc_residuals(kfc_op* op, int64_t m, const double* E, const double* V, double* resid);

/* G = V^H (op V) for a DEVICE block V (k <= 1024 columns of prod(dims), leading
 * dimension ld, FP64 or complex FP64 as op), G[k x k] on the host (complex:
 * interleaved).  With op a single product term — the node's port operators on
 * their axes — this is the port formation _port_matrix / _joint_port
 * (src/materialize.jl:441-444, 490-493) on the device. */
int kfc_sandwich(kfc_op* op, const void* V, int64_t k, int64_t ld, double* G);

/* ---- batched sweeps of small one-axis nodes (no reference counterpart: a batch
 * of the ChebyshevFilter solves of src/chebyshev.jl:310-406, which the
 * reference runs one node at a time; closest caller bench/common.jl:135-163).
 * n operators on one context, each with a single axis and banded, shift or
 * diagonal terms (the leaves of a sweep of periodic-grid or charge-basis
 * modes), all real or all complex.  Every filter step, Gram and rotation of
 * the whole batch is one launch.  nb = m + extra <= 32 block columns per
 * problem, a fixed filter degree, mode KFC_MODE_MIXED runs FP32 filter steps to
 * the FP32 floor and then FP64 to tol.  E: n x m ascending; V[i]: host N_i x m
 * (interleaved complex for complex operators); stats: n records. */
int kfc_chebyshev_batch(int32_t n, kfc_op* const* ops, int64_t m, int64_t extra, int32_t degree, double tol,
                        int64_t maxiter, uint64_t seed, int32_t mode, double* E, double* const* V,
                        kfc_stats* stats);

/* ---- learned warm start (no reference counterpart; feeds X0 of kfc_lobpcg).
 * A 6 -> 256 -> 256 -> 1024 GELU MLP with BF16 weights and FP32 accumulation. */
typedef struct kfc_mlp kfc_mlp;
int kfc_mlp_create(kfc_ctx* ctx, const int32_t* layer_dims, int32_t nlayers, uint64_t seed, kfc_mlp** mlp);
int kfc_mlp_destroy(kfc_mlp* mlp);
/* Host FP32 parameters in the order W1, b1, W2, b2, ... (W row-major out x in). */
int kfc_mlp_set_params(kfc_mlp* mlp, const float* params);
int kfc_mlp_get_params(kfc_mlp* mlp, float* params);
int64_t kfc_mlp_num_params(const kfc_mlp* mlp);
/* Host X[batch x in] -> host Y[batch x out]. */
int kfc_mlp_forward(kfc_mlp* mlp, const float* X, int64_t batch, float* Y);
/* Gradients of sum(dY * Y) into host grads (same layout as params). */
int kfc_mlp_backward(kfc_mlp* mlp, const float* X, const float* dY, int64_t batch, float* grads);
/* nsteps of Adam on the mean-squared error against host targets T[n x out];
 * returns the last loss. */
int kfc_mlp_train(kfc_mlp* mlp, const float* X, const float* T, int64_t n, int64_t batch, int64_t nsteps,
                  float lr, uint64_t seed, float* last_loss);

#ifdef __cplusplus
}
#endif
#endif /* KFC_H */
