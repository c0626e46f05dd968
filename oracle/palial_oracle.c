/*
 * palial_oracle.c -- CPU oracle (test infrastructure only; see palial_oracle.h).
 *
 * Every routine here is the checker, never the thing shipped: the product
 * library does not link it.  Two BLAS backends share one call sequence:
 *   - "naive": plain C restatements of the BLAS/LAPACK routines the reference
 *     calls (published reference-BLAS / LAPACK algorithms: dgemm, dtrsm, dsyrk,
 *     dtrmm, dgemv, dger, dtrmv, dscal, daxpy, dcopy, dpotf2, dgeqr2, dlarfg,
 *     dlarft, dlarfb, dlamch);
 *   - "openblas": the same entry points bound at run time with dlopen to an
 *     LP64 OpenBLAS that exports cblas_* / LAPACKE_* (optionally scipy_-prefixed).
 */
#define _GNU_SOURCE
#include "palial_oracle.h"

#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* CBLAS / LAPACKE enumerators: the values src/common.h:9-20 also uses. */
enum { ColMajor = 102, NoTrans = 111, Trans = 112, Upper = 121, Lower = 122,
       NonUnit = 131, Unit = 132, Left = 141, Right = 142 };
#define LAPACK_COL 102

/* ------------------------------------------------------------------ */
/* backend function table                                               */
/* ------------------------------------------------------------------ */
typedef void (*f_dgemm)(int, int, int, int, int, int, double, const double *, int,
                        const double *, int, double, double *, int);
typedef void (*f_dtrsm)(int, int, int, int, int, int, int, double, const double *, int,
                        double *, int);
typedef void (*f_dsyrk)(int, int, int, int, int, double, const double *, int, double,
                        double *, int);
typedef void (*f_dgemv)(int, int, int, int, double, const double *, int, const double *,
                        int, double, double *, int);
typedef void (*f_dger)(int, int, int, double, const double *, int, const double *, int,
                       double *, int);
typedef void (*f_dtrmv)(int, int, int, int, int, const double *, int, double *, int);
typedef void (*f_dscal)(int, double, double *, int);
typedef void (*f_daxpy)(int, double, const double *, int, double *, int);
typedef void (*f_dcopy)(int, const double *, int, double *, int);
typedef int (*f_dpotrf)(int, char, int, double *, int);
typedef int (*f_dgeqr2)(int, int, int, double *, int, double *, double *);
typedef int (*f_dlarft)(int, char, char, int, int, const double *, int, const double *,
                        double *, int);
typedef int (*f_dlarfb)(int, char, char, char, char, int, int, int, const double *, int,
                        const double *, int, double *, int, double *, int);
typedef int (*f_dlarfg)(int, double *, double *, int, double *);
typedef double (*f_dlamch)(char);

typedef struct {
    f_dgemm dgemm; f_dtrsm dtrsm; f_dsyrk dsyrk; f_dtrsm dtrmm; f_dgemv dgemv;
    f_dger dger; f_dtrmv dtrmv; f_dscal dscal; f_daxpy daxpy; f_dcopy dcopy;
    f_dpotrf dpotrf; f_dgeqr2 dgeqr2; f_dlarft dlarft; f_dlarfb dlarfb;
    f_dlarfg dlarfg; f_dlamch dlamch;
} blas_table;

/* ------------------------------------------------------------------ */
/* naive BLAS (column major only)                                       */
/* ------------------------------------------------------------------ */
static inline double tri_at(const double *A, int lda, int tr, int i, int j)
{
    return tr ? A[j + (long)i * lda] : A[i + (long)j * lda];
}

static void n_dgemm(int order, int ta, int tb, int M, int N, int K, double alpha,
                    const double *A, int lda, const double *B, int ldb, double beta,
                    double *C, int ldc)
{
    (void)order;
    if (M == 0 || N == 0) return;
    for (int j = 0; j < N; j++) {
        double *c = C + (long)j * ldc;
        if (beta == 0.0) for (int i = 0; i < M; i++) c[i] = 0.0;
        else if (beta != 1.0) for (int i = 0; i < M; i++) c[i] *= beta;
        if (K == 0 || alpha == 0.0) continue;
        if (ta == NoTrans) {
            for (int p = 0; p < K; p++) {
                double bpj = (tb == NoTrans) ? B[p + (long)j * ldb] : B[j + (long)p * ldb];
                double t = alpha * bpj;
                const double *a = A + (long)p * lda;
                for (int i = 0; i < M; i++) c[i] += t * a[i];
            }
        } else {
            for (int i = 0; i < M; i++) {
                const double *a = A + (long)i * lda;
                double s = 0.0;
                if (tb == NoTrans) {
                    const double *bb = B + (long)j * ldb;
                    for (int p = 0; p < K; p++) s += a[p] * bb[p];
                } else {
                    for (int p = 0; p < K; p++) s += a[p] * B[j + (long)p * ldb];
                }
                c[i] += alpha * s;
            }
        }
    }
}

/* op(A) X = alpha B (Left) or X op(A) = alpha B (Right) */
static void n_dtrsm(int order, int side, int uplo, int ta, int diag, int M, int N,
                    double alpha, const double *A, int lda, double *B, int ldb)
{
    (void)order;
    int tr = (ta != NoTrans);
    int lower = ((uplo == Lower) != tr);
    int unit = (diag == Unit);
    if (M == 0 || N == 0) return;
    if (side == Left) {
        for (int j = 0; j < N; j++) {
            double *x = B + (long)j * ldb;
            if (alpha != 1.0) for (int i = 0; i < M; i++) x[i] *= alpha;
            if (lower) {
                for (int i = 0; i < M; i++) {
                    double s = x[i];
                    for (int p = 0; p < i; p++) s -= tri_at(A, lda, tr, i, p) * x[p];
                    x[i] = unit ? s : s / tri_at(A, lda, tr, i, i);
                }
            } else {
                for (int i = M - 1; i >= 0; i--) {
                    double s = x[i];
                    for (int p = i + 1; p < M; p++) s -= tri_at(A, lda, tr, i, p) * x[p];
                    x[i] = unit ? s : s / tri_at(A, lda, tr, i, i);
                }
            }
        }
    } else {
        if (alpha != 1.0)
            for (int j = 0; j < N; j++)
                for (int i = 0; i < M; i++) B[i + (long)j * ldb] *= alpha;
        if (!lower) { /* op(A) upper: forward over columns */
            for (int j = 0; j < N; j++) {
                double *bj = B + (long)j * ldb;
                for (int p = 0; p < j; p++) {
                    double a = tri_at(A, lda, tr, p, j);
                    if (a == 0.0) continue;
                    const double *bp = B + (long)p * ldb;
                    for (int i = 0; i < M; i++) bj[i] -= a * bp[i];
                }
                if (!unit) {
                    double d = tri_at(A, lda, tr, j, j);
                    for (int i = 0; i < M; i++) bj[i] /= d;
                }
            }
        } else {
            for (int j = N - 1; j >= 0; j--) {
                double *bj = B + (long)j * ldb;
                for (int p = j + 1; p < N; p++) {
                    double a = tri_at(A, lda, tr, p, j);
                    if (a == 0.0) continue;
                    const double *bp = B + (long)p * ldb;
                    for (int i = 0; i < M; i++) bj[i] -= a * bp[i];
                }
                if (!unit) {
                    double d = tri_at(A, lda, tr, j, j);
                    for (int i = 0; i < M; i++) bj[i] /= d;
                }
            }
        }
    }
}

/* B = alpha op(A) B (Left) or alpha B op(A) (Right), A triangular */
static void n_dtrmm(int order, int side, int uplo, int ta, int diag, int M, int N,
                    double alpha, const double *A, int lda, double *B, int ldb)
{
    (void)order;
    int tr = (ta != NoTrans);
    int lower = ((uplo == Lower) != tr);
    int unit = (diag == Unit);
    if (M == 0 || N == 0) return;
    if (side == Left) {
        double *t = malloc(sizeof(double) * M);
        for (int j = 0; j < N; j++) {
            double *b = B + (long)j * ldb;
            memcpy(t, b, sizeof(double) * M);
            for (int i = 0; i < M; i++) {
                double s = 0.0;
                int p0 = lower ? 0 : i, p1 = lower ? i : M - 1;
                for (int p = p0; p <= p1; p++) {
                    double a = (p == i && unit) ? 1.0 : tri_at(A, lda, tr, i, p);
                    s += a * t[p];
                }
                b[i] = alpha * s;
            }
        }
        free(t);
    } else {
        double *t = malloc(sizeof(double) * N);
        for (int r = 0; r < M; r++) {
            for (int p = 0; p < N; p++) t[p] = B[r + (long)p * ldb];
            for (int j = 0; j < N; j++) {
                double s = 0.0;
                int p0 = lower ? j : 0, p1 = lower ? N - 1 : j;
                for (int p = p0; p <= p1; p++) {
                    double a = (p == j && unit) ? 1.0 : tri_at(A, lda, tr, p, j);
                    s += t[p] * a;
                }
                B[r + (long)j * ldb] = alpha * s;
            }
        }
        free(t);
    }
}

/* C = alpha A^T A + beta C (Upper, Trans) or alpha A A^T + beta C (Upper, NoTrans) */
static void n_dsyrk(int order, int uplo, int trans, int N, int K, double alpha,
                    const double *A, int lda, double beta, double *C, int ldc)
{
    (void)order;
    for (int j = 0; j < N; j++) {
        int i0 = (uplo == Upper) ? 0 : j, i1 = (uplo == Upper) ? j : N - 1;
        for (int i = i0; i <= i1; i++) {
            double s = 0.0;
            if (trans == NoTrans)
                for (int p = 0; p < K; p++) s += A[i + (long)p * lda] * A[j + (long)p * lda];
            else
                for (int p = 0; p < K; p++) s += A[p + (long)i * lda] * A[p + (long)j * lda];
            double c = (beta == 0.0) ? 0.0 : beta * C[i + (long)j * ldc];
            C[i + (long)j * ldc] = c + alpha * s;
        }
    }
}

static void n_dgemv(int order, int trans, int M, int N, double alpha, const double *A,
                    int lda, const double *x, int incx, double beta, double *y, int incy)
{
    (void)order;
    int leny = (trans == NoTrans) ? M : N, lenx = (trans == NoTrans) ? N : M;
    if (M == 0 || N == 0) return;
    for (int i = 0; i < leny; i++) {
        double v = (beta == 0.0) ? 0.0 : beta * y[(long)i * incy];
        y[(long)i * incy] = v;
    }
    if (alpha == 0.0) return;
    if (trans == NoTrans) {
        for (int j = 0; j < lenx; j++) {
            double t = alpha * x[(long)j * incx];
            for (int i = 0; i < leny; i++) y[(long)i * incy] += t * A[i + (long)j * lda];
        }
    } else {
        for (int j = 0; j < leny; j++) {
            double s = 0.0;
            for (int i = 0; i < lenx; i++) s += A[i + (long)j * lda] * x[(long)i * incx];
            y[(long)j * incy] += alpha * s;
        }
    }
}

static void n_dger(int order, int M, int N, double alpha, const double *x, int incx,
                   const double *y, int incy, double *A, int lda)
{
    (void)order;
    for (int j = 0; j < N; j++) {
        double t = alpha * y[(long)j * incy];
        if (t == 0.0) continue;
        for (int i = 0; i < M; i++) A[i + (long)j * lda] += x[(long)i * incx] * t;
    }
}

static void n_dtrmv(int order, int uplo, int trans, int diag, int N, const double *A,
                    int lda, double *x, int incx)
{
    (void)order;
    if (N == 0) return;
    int tr = (trans != NoTrans);
    int lower = ((uplo == Lower) != tr);
    double *t = malloc(sizeof(double) * N);
    for (int i = 0; i < N; i++) t[i] = x[(long)i * incx];
    for (int i = 0; i < N; i++) {
        double s = 0.0;
        int p0 = lower ? 0 : i, p1 = lower ? i : N - 1;
        for (int p = p0; p <= p1; p++) {
            double a = (p == i && diag == Unit) ? 1.0 : tri_at(A, lda, tr, i, p);
            s += a * t[p];
        }
        x[(long)i * incx] = s;
    }
    free(t);
}

static void n_dscal(int N, double alpha, double *x, int incx)
{
    for (int i = 0; i < N; i++) x[(long)i * incx] *= alpha;
}
static void n_daxpy(int N, double alpha, const double *x, int incx, double *y, int incy)
{
    if (alpha == 0.0) return;
    for (int i = 0; i < N; i++) y[(long)i * incy] += alpha * x[(long)i * incx];
}
static void n_dcopy(int N, const double *x, int incx, double *y, int incy)
{
    for (int i = 0; i < N; i++) y[(long)i * incy] = x[(long)i * incx];
}

static double n_dlamch(char c)
{
    if (c == 'S' || c == 's') return DBL_MIN;         /* safe minimum: 1/huge < tiny */
    if (c == 'E' || c == 'e') return DBL_EPSILON * 0.5; /* eps with rounding */
    if (c == 'P' || c == 'p') return DBL_EPSILON;
    return DBL_MIN;
}

/* LAPACK dpotf2, uplo = 'U' (unblocked): returns info */
static int n_dpotrf(int layout, char uplo, int n, double *A, int lda)
{
    (void)layout; (void)uplo;
    for (int j = 0; j < n; j++) {
        double *aj = A + (long)j * lda;
        double s = 0.0;
        for (int p = 0; p < j; p++) s += aj[p] * aj[p];
        double ajj = aj[j] - s;
        if (ajj <= 0.0 || isnan(ajj)) { aj[j] = ajj; return j + 1; }
        ajj = sqrt(ajj);
        aj[j] = ajj;
        if (j < n - 1) {
            /* A(j, j+1:n) -= A(0:j, j)^T A(0:j, j+1:n);  A(j, j+1:n) /= ajj */
            n_dgemv(ColMajor, Trans, j, n - j - 1, -1.0, A + (long)(j + 1) * lda, lda, aj, 1,
                    1.0, A + j + (long)(j + 1) * lda, lda);
            n_dscal(n - j - 1, 1.0 / ajj, A + j + (long)(j + 1) * lda, lda);
        }
    }
    return 0;
}

static double dnrm2_(int n, const double *x, int incx)
{
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; i++) {
        double v = x[(long)i * incx];
        if (v != 0.0) {
            double a = fabs(v);
            if (scale < a) { ssq = 1.0 + ssq * (scale / a) * (scale / a); scale = a; }
            else ssq += (a / scale) * (a / scale);
        }
    }
    return scale * sqrt(ssq);
}

static double dlapy2_(double x, double y)
{
    double xa = fabs(x), ya = fabs(y);
    double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
    if (z == 0.0 || w > DBL_MAX) return w;
    return w * sqrt(1.0 + (z / w) * (z / w));
}

/* LAPACK dlarfg: H (alpha; x) = (beta; 0), H = I - tau (1; v)(1; v)^T */
static int n_dlarfg(int n, double *alpha, double *x, int incx, double *tau)
{
    if (n <= 1) { *tau = 0.0; return 0; }
    double xnorm = dnrm2_(n - 1, x, incx);
    if (xnorm == 0.0) { *tau = 0.0; return 0; }
    double beta = -copysign(dlapy2_(*alpha, xnorm), *alpha);
    double safmin = n_dlamch('S') / n_dlamch('E');
    int knt = 0;
    if (fabs(beta) < safmin) {
        double rsafmn = 1.0 / safmin;
        do {
            knt++;
            n_dscal(n - 1, rsafmn, x, incx);
            beta *= rsafmn;
            *alpha *= rsafmn;
        } while (fabs(beta) < safmin && knt < 20);
        xnorm = dnrm2_(n - 1, x, incx);
        beta = -copysign(dlapy2_(*alpha, xnorm), *alpha);
    }
    *tau = (beta - *alpha) / beta;
    n_dscal(n - 1, 1.0 / (*alpha - beta), x, incx);
    for (int j = 0; j < knt; j++) beta *= safmin;
    *alpha = beta;
    return 0;
}

/* LAPACK dlarf, side = 'L': C = (I - tau v v^T) C */
static void n_dlarf_left(int m, int n, const double *v, double tau, double *C, int ldc,
                         double *work)
{
    if (tau == 0.0) return;
    n_dgemv(ColMajor, Trans, m, n, 1.0, C, ldc, v, 1, 0.0, work, 1);
    n_dger(ColMajor, m, n, -tau, v, 1, work, 1, C, ldc);
}

/* LAPACK dgeqr2 */
static int n_dgeqr2(int layout, int m, int n, double *A, int lda, double *tau, double *work)
{
    (void)layout;
    int k = m < n ? m : n;
    for (int i = 0; i < k; i++) {
        int r = (i + 1 < m) ? i + 1 : m - 1;
        n_dlarfg(m - i, &A[i + (long)i * lda], &A[r + (long)i * lda], 1, &tau[i]);
        if (i < n - 1) {
            double aii = A[i + (long)i * lda];
            A[i + (long)i * lda] = 1.0;
            n_dlarf_left(m - i, n - i - 1, &A[i + (long)i * lda], tau[i],
                         &A[i + (long)(i + 1) * lda], lda, work);
            A[i + (long)i * lda] = aii;
        }
    }
    return 0;
}

/* LAPACK dlarft, direct = 'F', storev = 'C' */
static int n_dlarft(int layout, char direct, char storev, int n, int k, const double *V,
                    int ldv, const double *tau, double *T, int ldt)
{
    (void)layout; (void)direct; (void)storev;
    for (int i = 0; i < k; i++) {
        if (tau[i] == 0.0) {
            for (int j = 0; j <= i; j++) T[j + (long)i * ldt] = 0.0;
            continue;
        }
        for (int j = 0; j < i; j++) T[j + (long)i * ldt] = -tau[i] * V[i + (long)j * ldv];
        /* T(0:i,i) -= tau V(i+1:n, 0:i)^T V(i+1:n, i) */
        if (n - i - 1 > 0 && i > 0)
            n_dgemv(ColMajor, Trans, n - i - 1, i, -tau[i], &V[i + 1], ldv,
                    &V[i + 1 + (long)i * ldv], 1, 1.0, &T[(long)i * ldt], 1);
        n_dtrmv(ColMajor, Upper, NoTrans, NonUnit, i, T, ldt, &T[(long)i * ldt], 1);
        T[i + (long)i * ldt] = tau[i];
    }
    return 0;
}

/* LAPACK dlarfb, side='L', trans='T', direct='F', storev='C' */
static int n_dlarfb(int layout, char side, char trans, char direct, char storev, int M,
                    int N, int K, const double *V, int ldv, const double *T, int ldt,
                    double *C, int ldc, double *W, int ldw)
{
    (void)layout; (void)direct; (void)storev;
    if (M <= 0 || N <= 0) return 0;
    if (!(side == 'L' && trans == 'T')) return -1;
    /* W = C1^T */
    for (int j = 0; j < K; j++) n_dcopy(N, &C[j], ldc, &W[(long)j * ldw], 1);
    /* W = W V1 (V1 unit lower) */
    n_dtrmm(ColMajor, Right, Lower, NoTrans, Unit, N, K, 1.0, V, ldv, W, ldw);
    if (M > K)
        n_dgemm(ColMajor, Trans, NoTrans, N, K, M - K, 1.0, &C[K], ldc, &V[K], ldv, 1.0, W, ldw);
    /* W = W T   (trans = 'T' -> TRANST = 'N') */
    n_dtrmm(ColMajor, Right, Upper, NoTrans, NonUnit, N, K, 1.0, T, ldt, W, ldw);
    if (M > K)
        n_dgemm(ColMajor, NoTrans, Trans, M - K, N, K, -1.0, &V[K], ldv, W, ldw, 1.0, &C[K], ldc);
    n_dtrmm(ColMajor, Right, Lower, Trans, Unit, N, K, 1.0, V, ldv, W, ldw);
    for (int j = 0; j < K; j++)
        for (int i = 0; i < N; i++) C[j + (long)i * ldc] -= W[i + (long)j * ldw];
    return 0;
}

static const blas_table NAIVE = {
    n_dgemm, n_dtrsm, n_dsyrk, n_dtrmm, n_dgemv, n_dger, n_dtrmv, n_dscal, n_daxpy,
    n_dcopy, n_dpotrf, n_dgeqr2, n_dlarft, n_dlarfb, n_dlarfg, n_dlamch,
};

static blas_table OB;
static int g_backend = 0;
static const blas_table *B = &NAIVE;
static void (*ob_set_threads)(int) = NULL;

static void *sym2(void *h, const char *prefix, const char *name)
{
    char buf[128];
    snprintf(buf, sizeof buf, "%s%s", prefix, name);
    return dlsym(h, buf);
}

int orc_use_openblas(const char *path)
{
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 0;
    const char *pre = sym2(h, "scipy_", "cblas_dgemm") ? "scipy_" : "";
#define BIND(field, name) OB.field = (void *)sym2(h, pre, name); if (!OB.field) return 0;
    BIND(dgemm, "cblas_dgemm") BIND(dtrsm, "cblas_dtrsm") BIND(dsyrk, "cblas_dsyrk")
    BIND(dtrmm, "cblas_dtrmm") BIND(dgemv, "cblas_dgemv") BIND(dger, "cblas_dger")
    BIND(dtrmv, "cblas_dtrmv") BIND(dscal, "cblas_dscal") BIND(daxpy, "cblas_daxpy")
    BIND(dcopy, "cblas_dcopy") BIND(dpotrf, "LAPACKE_dpotrf") BIND(dgeqr2, "LAPACKE_dgeqr2_work")
    BIND(dlarft, "LAPACKE_dlarft_work") BIND(dlarfb, "LAPACKE_dlarfb_work")
    BIND(dlarfg, "LAPACKE_dlarfg_work") BIND(dlamch, "LAPACKE_dlamch_work")
#undef BIND
    ob_set_threads = (void (*)(int))sym2(h, pre, "openblas_set_num_threads");
    B = &OB;
    g_backend = 1;
    return 1;
}
void orc_use_naive(void) { B = &NAIVE; g_backend = 0; }
int orc_backend(void) { return g_backend; }
void orc_openblas_threads(int n) { if (ob_set_threads) ob_set_threads(n); }

double orc_wtime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------ */
/* layout                                                               */
/* ------------------------------------------------------------------ */
long orc_tile_offset(int N, int b, int m, int n)
{
    /* include/tile_address.h:15: tile_nelements * (m + (N/b) * n) */
    return (long)b * b * ((long)m + (long)(N / b) * n);
}
long orc_elem_offset(int N, int b, int i, int j)
{
    return orc_tile_offset(N, b, i / b, j / b) + (i % b) + (long)(j % b) * b;
}
void orc_tiles_to_colmajor(int N, int b, const double *t, double *g)
{
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) g[i + (long)j * N] = t[orc_elem_offset(N, b, i, j)];
}
void orc_colmajor_to_tiles(int N, int b, const double *g, double *t)
{
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) t[orc_elem_offset(N, b, i, j)] = g[i + (long)j * N];
}
#define TILE(D, m, n) ((double *)(D).matrix + orc_tile_offset((D).matrix_size, (D).tile_size, (m), (n)))

/* ------------------------------------------------------------------ */
/* generators                                                           */
/* ------------------------------------------------------------------ */
static inline uint64_t sm64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double orc_uniform(uint64_t seed, uint64_t i, uint64_t j)
{
    uint64_t x = sm64(sm64(seed) + ((i << 32) | (j & 0xffffffffull)));
    return (double)(x >> 11) * 0x1.0p-53 - 0.5;
}
void orc_gen_sym_shift(orc_desc A, uint64_t seed, double shift)
{
    int N = A.matrix_size, b = A.tile_size;
    double *a = A.matrix;
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            double v = (i <= j) ? orc_uniform(seed, i, j) : orc_uniform(seed, j, i);
            if (i == j) v += shift;
            a[orc_elem_offset(N, b, i, j)] = v;
        }
}
void orc_gen_dense_shift(orc_desc A, uint64_t seed, double shift)
{
    int N = A.matrix_size, b = A.tile_size;
    double *a = A.matrix;
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            double v = orc_uniform(seed, i, j);
            if (i == j) v += shift;
            a[orc_elem_offset(N, b, i, j)] = v;
        }
}
/* Standard normal entries by Box-Muller from one counter per entry (the
 * build's PALIAL_GEN_GAUSS): h = sm64(sm64(seed) + (i<<32 | j)),
 * u1 = ((h >> 11) + 1) 2^-53 in (0,1], u2 = (sm64(h) >> 11) 2^-53,
 * z = sqrt(-2 ln u1) cos(2 pi u2). */
double orc_gauss(uint64_t seed, uint64_t i, uint64_t j)
{
    uint64_t h = sm64(sm64(seed) + ((i << 32) | (j & 0xffffffffull)));
    double u1 = (double)((h >> 11) + 1) * 0x1.0p-53;
    double u2 = (double)(sm64(h) >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}
void orc_gen_gauss(orc_desc A, uint64_t seed)
{
    int N = A.matrix_size, b = A.tile_size;
    double *a = A.matrix;
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) a[orc_elem_offset(N, b, i, j)] = orc_gauss(seed, i, j);
}
/* diag(s), s_i = kappa^(-i/(N-1)): a geometric spectrum from 1 to 1/kappa */
void orc_gen_geom_diag(orc_desc A, double kappa)
{
    int N = A.matrix_size, b = A.tile_size;
    double *a = A.matrix;
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++)
            a[orc_elem_offset(N, b, i, j)] = (i != j) ? 0.0 : (N > 1 ? pow(kappa, -(double)i / (double)(N - 1)) : 1.0);
}
void orc_gen_palial(orc_desc A, uint64_t seed)
{
    /* include/populate.h:51-66: diagonal 173*(1+r), r in [0,1]; everything else 0.5 */
    int N = A.matrix_size, b = A.tile_size;
    double *a = A.matrix;
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            double v = 0.5;
            if (i == j) v = (orc_uniform(seed, i, i) + 0.5) * 173.0 + 173.0;
            a[orc_elem_offset(N, b, i, j)] = v;
        }
}
void orc_gen_identity(orc_desc A)
{
    int N = A.matrix_size, b = A.tile_size;
    double *a = A.matrix;
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) a[orc_elem_offset(N, b, i, j)] = (i == j) ? 1.0 : 0.0;
}
void orc_gen_special4x4(orc_desc A)
{
    /* include/populate.h:13-33, tile-by-tile, each tile column-major 2x2 */
    static const double t00[4] = {4.5, -0.095026, -0.095026, 3.719688};
    static const double t10[4] = {0.361857, 0.388551, -0.447549, 0.155058};
    static const double t01[4] = {0.361857, -0.447549, 0.388551, 0.155058};
    static const double t11[4] = {4.484953, 0.342457, 0.342457, 3.982519};
    memcpy(TILE(A, 0, 0), t00, sizeof t00);
    memcpy(TILE(A, 1, 0), t10, sizeof t10);
    memcpy(TILE(A, 0, 1), t01, sizeof t01);
    memcpy(TILE(A, 1, 1), t11, sizeof t11);
}

/* ------------------------------------------------------------------ */
/* tile kernels                                                          */
/* ------------------------------------------------------------------ */
int orc_dpotrf_tile(int b, double *A, int lda)
{
    return B->dpotrf(LAPACK_COL, 'U', b, A, lda);
}

/* src/srclu/dgetrf_no_pivot.h:11-57 */
int orc_dgetrf_no_piv(double *A, int lda, int n, int internal)
{
    if (n == 0 || internal == 0) return 0;
    double sfmin = B->dlamch('S');
    int k = n < internal ? n : internal;
    for (int i = 0; i < k; i++) {
        double alpha = A[(long)i * lda + i];
        if (alpha != 0.0) {
            if (fabs(alpha) > sfmin) {
                B->dscal(n - i - 1, 1.0 / alpha, &A[(long)i * lda + i + 1], 1);
            } else {
                for (int j = i + 1; j < n; j++) A[(long)lda * i + j] = A[(long)lda * i + j] / alpha;
            }
        }
        B->dger(ColMajor, n - i - 1, internal - i - 1, -1.0, &A[(long)lda * i + i + 1], 1,
                &A[(long)lda * (i + 1) + i], lda, &A[(long)lda * (i + 1) + i + 1], lda);
    }
    return 0;
}

/* src/srclu/dgetrf.h:11-59 (the i-loop runs once with gamma = n; its blocked tail is dead) */
int orc_dgetrf(int n, double *A, int lda)
{
    if (n == 0) return 0;
    return orc_dgetrf_no_piv(A, lda, n, n);
}

/* src/srcqr/dgeqrt.h:9-56 with ib = nb (= n in every reference call) */
int orc_dgeqrt(int n, int nb, double *A, int lda, double *T, int ldt, double *tau, double *work)
{
    if (n == 0 || nb == 0) return 0;
    for (int i = 0; i < n; i += nb) {
        int g = nb < n - i ? nb : n - i;
        B->dgeqr2(LAPACK_COL, n - i, g, &A[(long)lda * i + i], lda, &tau[i], work);
        B->dlarft(LAPACK_COL, 'F', 'C', n - i, g, &A[(long)lda * i + i], lda, &tau[i],
                  &T[(long)ldt * i], ldt);
        if (n > i + g)
            B->dlarfb(LAPACK_COL, 'L', 'T', 'F', 'C', n - i, n - i - g, g, &A[(long)lda * i + i],
                      lda, &T[(long)ldt * i], ldt, &A[(long)lda * (i + g) + i], lda, work,
                      n - i - g);
    }
    return 0;
}

/* src/srcqr/dormqr.h:9-53; only (Left, Trans) is a correct application there */
int orc_dormqr(int side, int trans, int n, const double *V, int ldv, const double *T, int ldt,
               double *C, int ldc, double *work, int ldw)
{
    if (n == 0) return 0;
    if (!(side == Left && trans != NoTrans)) return 1;
    B->dlarfb(LAPACK_COL, 'L', 'T', 'F', 'C', n, n, n, V, ldv, T, ldt, C, ldc, work, ldw);
    return 0;
}

/* src/srcqr/dtsqrt.h:11-109 (lambda loop runs once; the trailing dtsmqr call is dead) */
int orc_dtsqrt(int n, double *A1, int lda1, double *A2, int lda2, double *T, int ldt,
               double *tau, double *work)
{
    if (n == 0) return 0;
    for (int i = 0; i < n; i++) {
        B->dlarfg(n + 1, &A1[(long)lda1 * i + i], &A2[(long)lda2 * i], 1, &tau[i]);
        if (i + 1 < n) {
            double alpha = -tau[i];
            B->dcopy(n - i - 1, &A1[(long)lda1 * (i + 1) + i], lda1, work, 1);
            B->dgemv(ColMajor, Trans, n, n - i - 1, 1.0, &A2[(long)lda2 * (i + 1)], lda2,
                     &A2[(long)lda2 * i], 1, 1.0, work, 1);
            B->daxpy(n - i - 1, alpha, work, 1, &A1[(long)lda1 * (i + 1) + i], lda1);
            B->dger(ColMajor, n, n - i - 1, alpha, &A2[(long)lda2 * i], 1, work, 1,
                    &A2[(long)lda2 * (i + 1)], lda2);
        }
        double alpha = -tau[i];
        B->dgemv(ColMajor, Trans, n, i, alpha, A2, lda2, &A2[(long)lda2 * i], 1, 0.0,
                 &T[(long)ldt * i], 1);
        B->dtrmv(ColMajor, Upper, NoTrans, NonUnit, i, T, ldt, &T[(long)ldt * i], 1);
        T[(long)ldt * i + i] = tau[i];
    }
    return 0;
}

/*
 * src/srcqr/dtsmqr.h:11-59 -> dparfb.h:21-74 -> dpamm.h (L = 0).  The reference
 * only calls it with M = N = K = ib and (Left, Trans); the net effect is
 *   W = V^T A2 + A1;  W = T^T W;  A1 -= W;  A2 -= V W.
 */
int orc_dtsmqr(int side, int trans, int M, int N, int K, int ib, double *A1, int lda1,
               double *A2, int lda2, const double *V, int ldv, const double *T, int ldt,
               double *work, int ldw)
{
    if (M == 0 || N == 0 || K == 0 || ib == 0) return 0;
    if (!(side == Left && trans != NoTrans)) return 1;
    if (!(M == N && N == K && K <= ib)) return 1;
    /* dpamm_W: W(K x N) = V(N x K)^T A2(N x N), then W += A1 */
    B->dgemm(ColMajor, Trans, NoTrans, K, N, N, 1.0, V, ldv, A2, lda2, 0.0, work, ldw);
    for (int j = 0; j < N; j++) B->daxpy(K, 1.0, &A1[(long)lda1 * j], 1, &work[(long)ldw * j], 1);
    /* dparfb: W = T^T W; A1 -= W */
    B->dtrmm(ColMajor, Left, Upper, Trans, NonUnit, K, N, 1.0, T, ldt, work, ldw);
    for (int j = 0; j < M; j++) B->daxpy(K, -1.0, &work[(long)ldw * j], 1, &A1[(long)lda1 * j], 1);
    /* dpamm_A2: A2 -= V W */
    B->dgemm(ColMajor, NoTrans, NoTrans, N, N, K, -1.0, V, ldv, work, ldw, 1.0, A2, lda2);
    return 0;
}

void orc_chol_trsm(int b, const double *Akk, double *Akm)
{
    B->dtrsm(ColMajor, Left, Upper, Trans, NonUnit, b, b, 1.0, Akk, b, Akm, b);
}
void orc_chol_syrk(int b, const double *Akm, double *Amm)
{
    B->dsyrk(ColMajor, Upper, Trans, b, b, -1.0, Akm, b, 1.0, Amm, b);
}
void orc_chol_gemm(int b, const double *Akn, const double *Akm, double *Anm)
{
    B->dgemm(ColMajor, Trans, NoTrans, b, b, b, -1.0, Akn, b, Akm, b, 1.0, Anm, b);
}
void orc_lu_trsm_r(int b, const double *Akk, double *Amk)
{
    B->dtrsm(ColMajor, Right, Upper, NoTrans, NonUnit, b, b, 1.0, Akk, b, Amk, b);
}
void orc_lu_trsm_l(int b, const double *Akk, double *Akn)
{
    B->dtrsm(ColMajor, Left, Lower, NoTrans, Unit, b, b, 1.0, Akk, b, Akn, b);
}
void orc_lu_gemm(int b, const double *Amk, const double *Akn, double *Amn)
{
    B->dgemm(ColMajor, NoTrans, NoTrans, b, b, b, -1.0, Amk, b, Akn, b, 1.0, Amn, b);
}

/* ------------------------------------------------------------------ */
/* algorithms, sequential program order                                  */
/* ------------------------------------------------------------------ */
int orc_cholesky(orc_desc A)
{
    /* src/cholesky.h:15-176 */
    int MT = A.matrix_size / A.tile_size, b = A.tile_size, info = 0;
    for (int k = 0; k < MT; k++) {
        int r = orc_dpotrf_tile(b, TILE(A, k, k), b);   /* info is ignored by the reference */
        if (r && !info) info = k * b + r;
        for (int m = k + 1; m < MT; m++) orc_chol_trsm(b, TILE(A, k, k), TILE(A, k, m));
        for (int m = k + 1; m < MT; m++) {
            orc_chol_syrk(b, TILE(A, k, m), TILE(A, m, m));
            for (int n = k + 1; n < m; n++)
                orc_chol_gemm(b, TILE(A, k, n), TILE(A, k, m), TILE(A, n, m));
        }
    }
    return info;
}

int orc_lu(orc_desc A)
{
    /* src/lu.h:14-173 */
    int MT = A.matrix_size / A.tile_size, b = A.tile_size;
    for (int k = 0; k < MT; k++) {
        orc_dgetrf(b, TILE(A, k, k), b);
        for (int m = k + 1; m < MT; m++) orc_lu_trsm_r(b, TILE(A, k, k), TILE(A, m, k));
        for (int n = k + 1; n < MT; n++) {
            orc_lu_trsm_l(b, TILE(A, k, k), TILE(A, k, n));
            for (int m = k + 1; m < MT; m++)
                orc_lu_gemm(b, TILE(A, m, k), TILE(A, k, n), TILE(A, m, n));
        }
    }
    return 0;
}

int orc_qr(orc_desc A, orc_desc S)
{
    /* src/qr.h:18-190; T factors of the flat-tree TS QR go to S */
    int MT = A.matrix_size / A.tile_size, b = A.tile_size;
    double *tau = malloc(sizeof(double) * b);
    double *work = malloc(sizeof(double) * (size_t)b * b);
    for (int k = 0; k < MT; k++) {
        orc_dgeqrt(b, S.tile_size, TILE(A, k, k), b, TILE(S, k, k), S.tile_size, tau, work);
        for (int n = k + 1; n < MT; n++)
            orc_dormqr(Left, Trans, b, TILE(A, k, k), b, TILE(S, k, k), S.tile_size,
                       TILE(A, k, n), b, work, S.tile_size);
        for (int m = k + 1; m < MT; m++) {
            orc_dtsqrt(b, TILE(A, k, k), b, TILE(A, m, k), b, TILE(S, m, k), S.tile_size, tau, work);
            for (int n = k + 1; n < MT; n++)
                orc_dtsmqr(Left, Trans, b, b, b, b, TILE(A, k, n), b, TILE(A, m, n), b,
                           TILE(A, m, k), b, TILE(S, m, k), S.tile_size, work, b);
        }
    }
    free(tau);
    free(work);
    return 0;
}

/* ------------------------------------------------------------------ */
/* algorithms, OpenMP task DAG (correct dependences; src/lu.h:62,104 are */
/* not reproduced)                                                       */
/* ------------------------------------------------------------------ */
int orc_cholesky_omp(orc_desc A, int nthreads)
{
    int MT = A.matrix_size / A.tile_size, b = A.tile_size;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#pragma omp master
    {
        for (int k = 0; k < MT; k++) {
            double *akk = TILE(A, k, k);
#pragma omp task depend(inout : akk[0])
            orc_dpotrf_tile(b, akk, b);
            for (int m = k + 1; m < MT; m++) {
                double *akm = TILE(A, k, m);
#pragma omp task depend(in : akk[0]) depend(inout : akm[0])
                orc_chol_trsm(b, akk, akm);
            }
            for (int m = k + 1; m < MT; m++) {
                double *akm = TILE(A, k, m), *amm = TILE(A, m, m);
#pragma omp task depend(in : akm[0]) depend(inout : amm[0])
                orc_chol_syrk(b, akm, amm);
                for (int n = k + 1; n < m; n++) {
                    double *akn = TILE(A, k, n), *anm = TILE(A, n, m);
#pragma omp task depend(in : akn[0], akm[0]) depend(inout : anm[0])
                    orc_chol_gemm(b, akn, akm, anm);
                }
            }
        }
#pragma omp taskwait
    }
    return 0;
}

int orc_lu_omp(orc_desc A, int nthreads)
{
    int MT = A.matrix_size / A.tile_size, b = A.tile_size;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#pragma omp master
    {
        for (int k = 0; k < MT; k++) {
            double *akk = TILE(A, k, k);
#pragma omp task depend(inout : akk[0])
            orc_dgetrf(b, akk, b);
            for (int m = k + 1; m < MT; m++) {
                double *amk = TILE(A, m, k);
#pragma omp task depend(in : akk[0]) depend(inout : amk[0])
                orc_lu_trsm_r(b, akk, amk);
            }
            for (int n = k + 1; n < MT; n++) {
                double *akn = TILE(A, k, n);
#pragma omp task depend(in : akk[0]) depend(inout : akn[0])
                orc_lu_trsm_l(b, akk, akn);
                for (int m = k + 1; m < MT; m++) {
                    double *amk = TILE(A, m, k), *amn = TILE(A, m, n);
#pragma omp task depend(in : amk[0], akn[0]) depend(inout : amn[0])
                    orc_lu_gemm(b, amk, akn, amn);
                }
            }
        }
#pragma omp taskwait
    }
    return 0;
}

int orc_qr_omp(orc_desc A, orc_desc S, int nthreads)
{
    int MT = A.matrix_size / A.tile_size, b = A.tile_size, sb = S.tile_size;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#pragma omp master
    {
        for (int k = 0; k < MT; k++) {
            double *akk = TILE(A, k, k), *skk = TILE(S, k, k);
#pragma omp task depend(inout : akk[0]) depend(out : skk[0])
            {
                double *tau = malloc(sizeof(double) * b), *work = malloc(sizeof(double) * (size_t)b * b);
                orc_dgeqrt(b, sb, akk, b, skk, sb, tau, work);
                free(tau); free(work);
            }
            for (int n = k + 1; n < MT; n++) {
                double *akn = TILE(A, k, n);
#pragma omp task depend(in : akk[0], skk[0]) depend(inout : akn[0])
                {
                    double *work = malloc(sizeof(double) * (size_t)b * b);
                    orc_dormqr(Left, Trans, b, akk, b, skk, sb, akn, b, work, sb);
                    free(work);
                }
            }
            for (int m = k + 1; m < MT; m++) {
                double *amk = TILE(A, m, k), *smk = TILE(S, m, k);
#pragma omp task depend(inout : akk[0], amk[0]) depend(out : smk[0])
                {
                    double *tau = malloc(sizeof(double) * b), *work = malloc(sizeof(double) * (size_t)b * b);
                    orc_dtsqrt(b, akk, b, amk, b, smk, sb, tau, work);
                    free(tau); free(work);
                }
                for (int n = k + 1; n < MT; n++) {
                    double *akn = TILE(A, k, n), *amn = TILE(A, m, n);
#pragma omp task depend(inout : akn[0], amn[0]) depend(in : amk[0], smk[0])
                    {
                        double *work = malloc(sizeof(double) * (size_t)b * b);
                        orc_dtsmqr(Left, Trans, b, b, b, b, akn, b, amn, b, amk, b, smk, sb, work, b);
                        free(work);
                    }
                }
            }
        }
#pragma omp taskwait
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* solves                                                               */
/* ------------------------------------------------------------------ */
void orc_chol_solve(orc_desc U, double *Bm, int nrhs, int ldb)
{
    int N = U.matrix_size, b = U.tile_size;
    const double *u = U.matrix;
    for (int r = 0; r < nrhs; r++) {
        double *x = Bm + (long)r * ldb;
        for (int i = 0; i < N; i++) {     /* U^T y = x */
            double s = x[i];
            for (int p = 0; p < i; p++) s -= u[orc_elem_offset(N, b, p, i)] * x[p];
            x[i] = s / u[orc_elem_offset(N, b, i, i)];
        }
        for (int i = N - 1; i >= 0; i--) { /* U x = y */
            double s = x[i];
            for (int p = i + 1; p < N; p++) s -= u[orc_elem_offset(N, b, i, p)] * x[p];
            x[i] = s / u[orc_elem_offset(N, b, i, i)];
        }
    }
}

void orc_lu_solve(orc_desc LU, double *Bm, int nrhs, int ldb)
{
    int N = LU.matrix_size, b = LU.tile_size;
    const double *a = LU.matrix;
    for (int r = 0; r < nrhs; r++) {
        double *x = Bm + (long)r * ldb;
        for (int i = 0; i < N; i++) {
            double s = x[i];
            for (int p = 0; p < i; p++) s -= a[orc_elem_offset(N, b, i, p)] * x[p];
            x[i] = s;
        }
        for (int i = N - 1; i >= 0; i--) {
            double s = x[i];
            for (int p = i + 1; p < N; p++) s -= a[orc_elem_offset(N, b, i, p)] * x[p];
            x[i] = s / a[orc_elem_offset(N, b, i, i)];
        }
    }
}

/* ------------------------------------------------------------------ */
/* task names in creation order, with C's stale loop variables           */
/* (src/cholesky.h:19-30; src/lu.h:16-25; src/qr.h:21-30).  The LU and QR */
/* loops read m and n uninitialised on their first task; 0 is used here. */
/* ------------------------------------------------------------------ */
typedef struct { char *buf; size_t len, pos; long count; } namebuf;
static void emit(namebuf *nb, const char *kind, int k, int m, int n)
{
    char tmp[64];
    int w = snprintf(tmp, sizeof tmp, "%s%d%d%d\n", kind, k, m, n);
    if (nb->buf && nb->pos + (size_t)w < nb->len) memcpy(nb->buf + nb->pos, tmp, (size_t)w + 1);
    nb->pos += (size_t)w;
    nb->count++;
}

long orc_task_names(int algo, int MT, char *buf, size_t buflen)
{
    namebuf nb = {buf, buflen, 0, 0};
    int k, m = 0, n = 0;
    if (algo == 0) {
        for (k = 0; k < MT; k++) {
            emit(&nb, "potrf", k, m, n);
            for (m = k + 1; m < MT; m++) emit(&nb, "trsm", k, m, n);
            for (m = k + 1; m < MT; m++) {
                emit(&nb, "syrk", k, m, n);
                for (n = k + 1; n < m; n++) emit(&nb, "gemm", k, m, n);
            }
        }
    } else if (algo == 1) {
        for (k = 0; k < MT; k++) {
            emit(&nb, "getrf", k, m, n);
            for (m = k + 1; m < MT; m++) emit(&nb, "trsm", k, m, n);
            for (n = k + 1; n < MT; n++) {
                emit(&nb, "trsm", k, m, n);
                for (m = k + 1; m < MT; m++) emit(&nb, "gemm", k, m, n);
            }
        }
    } else {
        for (k = 0; k < MT; k++) {
            emit(&nb, "geqrt", k, m, n);
            for (n = k + 1; n < MT; n++) emit(&nb, "ormqr", k, m, n);
            for (m = k + 1; m < MT; m++) {
                emit(&nb, "tsqrt", k, m, n);
                for (n = k + 1; n < MT; n++) emit(&nb, "tsmqr", k, m, n);
            }
        }
    }
    return nb.count;
}
