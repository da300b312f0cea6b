/* CPU restatement of KronFreeCircuits.jl's Kronecker-free apply, for the CPU
 * baseline and as an independent check of the numpy oracle.  TEST
 * INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg and tests/).
 *
 * mul!(y, H, x) (src/apply.jl:416-432): y = diag .* x, then for every kept
 * term its factors applied one after another through two scratch buffers
 * (_accumulate_term!, src/apply.jl:370-393) and y .+= c * src.  The factor
 * kernels follow _banded_kernel! (src/apply.jl:66-79), _shift_kernel!
 * (src/apply.jl:111-118), _diagonal_axis! (src/apply.jl:60-63) and
 * _dense_axis! (src/apply.jl:37-53).  Real FP64 data; OpenMP over rows. */
#include <stdint.h>
#include <string.h>

typedef struct {
    int32_t axis, kind;   /* kind: 0 diag, 1 banded, 2 shift, 3 dense */
    int64_t d;
    int32_t nb;
    const int64_t* offs;
    const double* w;      /* d x nb column-major */
    int64_t s;
    double f;
    const double* M;      /* d x d column-major */
    const double* dg;     /* d */
} of_factor;

typedef struct {
    double c;
    int32_t nf;
    const of_factor* f;
} of_term;

static void factor_apply(const of_factor* F, int64_t stride, int64_t N, const double* src, double* dst) {
    const int64_t d = F->d;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < N; ++t) {
        const int64_t i = (t / stride) % d;
        const int64_t base = t - i * stride;
        double acc = 0.0;
        if (F->kind == 0) {
            acc = F->dg[i] * src[t];
        } else if (F->kind == 1) {
            for (int b = 0; b < F->nb; ++b) {
                const int64_t j = i + F->offs[b];
                if (j >= 0 && j < d) acc += F->w[i + b * d] * src[base + j * stride];
            }
        } else if (F->kind == 2) {
            const int64_t j = i - F->s;
            acc = (j >= 0 && j < d) ? F->f * src[base + j * stride] : 0.0;
        } else {
            for (int64_t j = 0; j < d; ++j) acc += F->M[i + j * d] * src[base + j * stride];
        }
        dst[t] = acc;
    }
}

void oracle_apply(int32_t K, const int64_t* dims, const double* diag, int32_t nt, const of_term* terms,
                  const double* x, double* y, double* b1, double* b2) {
    int64_t N = 1, stride[16];
    for (int a = 0; a < K; ++a) {
        stride[a] = N;
        N *= dims[a];
    }
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < N; ++t) y[t] = diag[t] * x[t];
    for (int ti = 0; ti < nt; ++ti) {
        const of_term* T = &terms[ti];
        const double* src = x;
        double* dst = b1;
        double* other = b2;
        for (int k = 0; k < T->nf; ++k) {
            factor_apply(&T->f[k], stride[T->f[k].axis], N, src, dst);
            src = dst;
            double* tmp = dst;
            dst = other;
            other = tmp;
        }
        const double c = T->c;
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < N; ++t) y[t] += c * src[t];
    }
}
