/*
 * palial_oracle.h -- CPU oracle for the tiled dense factorizations.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * check and bench.py's CPU-baseline leg may load it; the product library
 * (libpalial_hip.so) and the palial driver never link or call it.
 *
 * It restates, in plain C, the reference's tiled algorithms in sequential
 * program order (and as an OpenMP task DAG with correct tile dependences):
 *   cholesky  -> src/cholesky.h:12-177
 *   lu        -> src/lu.h:11-175, src/srclu/dgetrf.h:11-59,
 *                src/srclu/dgetrf_no_pivot.h:11-57
 *   qr        -> src/qr.h:15-192, src/srcqr/dgeqrt.h:9-56,
 *                src/srcqr/dormqr.h:9-53, src/srcqr/dtsqrt.h:11-109,
 *                src/srcqr/dtsmqr.h:11-59, src/srcqr/dparfb.h:14-135,
 *                src/srcqr/dpamm.h:13-396
 * The BLAS/LAPACK routines those files call (OpenBLAS, un-vendored) are
 * restated here from the published LAPACK/BLAS algorithms ("naive" backend),
 * or bound at run time to an LP64 OpenBLAS ("openblas" backend) when one is
 * found (scipy's bundled libscipy_openblas exports them with a scipy_ prefix).
 */
#ifndef PALIAL_ORACLE_H
#define PALIAL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and types as include/descriptor.h:13-20 (sizeof 40 on x86-64). */
typedef struct orc_matrix_desc {
    int      tile_size;
    long int matrix_nelements;
    long int tile_nelements;
    int      matrix_size;
    void    *matrix;
} orc_desc;

/* ---- layout (include/tile_address.h:9-17, include/print.h:27) ---- */
long orc_tile_offset(int N, int b, int m, int n);           /* element offset of tile (m,n) */
long orc_elem_offset(int N, int b, int i, int j);           /* element (i,j) in layout L */
void orc_tiles_to_colmajor(int N, int b, const double *t, double *g);
void orc_colmajor_to_tiles(int N, int b, const double *g, double *t);

/* ---- seeded generators (counter based, written straight into layout L) ---- */
double orc_uniform(uint64_t seed, uint64_t i, uint64_t j);  /* U[-0.5,0.5) */
void orc_gen_sym_shift(orc_desc A, uint64_t seed, double shift);   /* S + shift*I */
void orc_gen_dense_shift(orc_desc A, uint64_t seed, double shift); /* G + shift*I */
void orc_gen_palial(orc_desc A, uint64_t seed);  /* D + 0.5*11^T, include/populate.h:35-66 */
double orc_gauss(uint64_t seed, uint64_t i, uint64_t j);   /* Box-Muller N(0,1) entry */
void orc_gen_gauss(orc_desc A, uint64_t seed);                /* G_ij ~ N(0,1) */
void orc_gen_geom_diag(orc_desc A, double kappa);             /* diag(kappa^(-i/(N-1))) */
void orc_gen_identity(orc_desc A);               /* include/populate.h:48,58 (IDENTITY) */
void orc_gen_special4x4(orc_desc A);             /* include/populate.h:13-33 (N=4, b=2) */

/* ---- BLAS backend ---- */
int  orc_use_openblas(const char *path);   /* 1 bound, 0 not found (naive stays active) */
void orc_use_naive(void);
int  orc_backend(void);                    /* 0 naive, 1 openblas */
void orc_openblas_threads(int n);

/* ---- tile kernels (reference signatures) ---- */
int orc_dpotrf_tile(int b, double *A, int lda);                                   /* cholesky.h:40 */
int orc_dgetrf(int n, double *A, int lda);                                        /* dgetrf.h:11 */
int orc_dgetrf_no_piv(double *A, int lda, int n, int internal);                   /* dgetrf_no_pivot.h:11 */
int orc_dgeqrt(int n, int nb, double *A, int lda, double *T, int ldt,
               double *tau, double *work);                                        /* dgeqrt.h:9 */
int orc_dormqr(int side, int trans, int n, const double *V, int ldv,
               const double *T, int ldt, double *C, int ldc, double *work, int ldw); /* dormqr.h:9 */
int orc_dtsqrt(int n, double *A1, int lda1, double *A2, int lda2, double *T, int ldt,
               double *tau, double *work);                                        /* dtsqrt.h:11 */
int orc_dtsmqr(int side, int trans, int M, int N, int K, int ib,
               double *A1, int lda1, double *A2, int lda2, const double *V, int ldv,
               const double *T, int ldt, double *work, int ldw);                  /* dtsmqr.h:11 */
/* Cholesky / LU tile tasks exactly as the DAG calls them */
void orc_chol_trsm(int b, const double *Akk, double *Akm);      /* cholesky.h:73-84 */
void orc_chol_syrk(int b, const double *Akm, double *Amm);      /* cholesky.h:115-125 */
void orc_chol_gemm(int b, const double *Akn, const double *Akm, double *Anm); /* cholesky.h:156-169 */
void orc_lu_trsm_r(int b, const double *Akk, double *Amk);      /* lu.h:67-78 */
void orc_lu_trsm_l(int b, const double *Akk, double *Akn);      /* lu.h:109-120 */
void orc_lu_gemm(int b, const double *Amk, const double *Akn, double *Amn);   /* lu.h:151-164 */

/* ---- algorithms: sequential program order (the oracle semantics) ---- */
int orc_cholesky(orc_desc A);            /* returns first failing POTRF info as k*b+info, 0 ok */
int orc_lu(orc_desc A);
int orc_qr(orc_desc A, orc_desc S);
/* ---- algorithms: OpenMP task DAG with correct tile dependences ---- */
int orc_cholesky_omp(orc_desc A, int nthreads);
int orc_lu_omp(orc_desc A, int nthreads);
int orc_qr_omp(orc_desc A, orc_desc S, int nthreads);

/* ---- solves on the factored layout-L matrix (column-major RHS, N x nrhs) ---- */
void orc_chol_solve(orc_desc U, double *B, int nrhs, int ldb);   /* U^T U x = b */
void orc_lu_solve(orc_desc LU, double *B, int nrhs, int ldb);    /* L U x = b  */

/* ---- task enumeration (names kind+k+m+n, creation order, cholesky.h:19-30) ---- */
/* algo: 0 cholesky, 1 lu, 2 qr.  Writes "name\n" lines; returns the task count. */
long orc_task_names(int algo, int MT, char *buf, size_t buflen);

double orc_wtime(void);

#ifdef __cplusplus
}
#endif
#endif
