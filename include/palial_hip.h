/*
 * palial_hip.h -- C ABI of libpalial_hip.so, the MI355X (gfx950) tiled dense
 * factorization library.  Plain C types only: a drop-in for the reference's
 * algorithm layer (src/cholesky.h, src/lu.h, src/qr.h) and tile kernels
 * (the headers of src/srclu and src/srcqr), with the reference's tile layout
 * (include/descriptor.h, include/tile_address.h).
 *
 * Tile layout L: an N x N matrix (N % b == 0) is stored as (N/b)^2 contiguous
 * b x b tiles; tile (m,n) starts at element b*b*(m + (N/b)*n) and is
 * column-major inside with leading dimension b.
 *
 * Every entry point is synchronous: it returns after the device work it
 * issued has completed.  One call at a time per process (calls are
 * serialised by an internal mutex).  Return codes: 0 success; > 0 a
 * LAPACK-style info (first failing pivot, 1-based global column); < 0 one of
 * the PALIAL_ERR_* codes below.  The library never calls exit().
 */
#ifndef PALIAL_HIP_H
#define PALIAL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Byte-identical to the reference descriptor (include/descriptor.h:13-20):
 * field offsets 0/8/16/24/32, sizeof 40 on x86-64.  Passed by value.  A
 * translation unit that already has the reference's typedef defines
 * PALIAL_HIP_MATRIX_DESC before including this header. */
#ifndef PALIAL_HIP_MATRIX_DESC
#define PALIAL_HIP_MATRIX_DESC
typedef struct palial_matrix_desc {
    int      tile_size;        /* b */
    long int matrix_nelements; /* N*N */
    long int tile_nelements;   /* b*b */
    int      matrix_size;      /* N */
    void    *matrix;           /* host or device pointer to layout-L storage */
} MATRIX_desc;
#endif

/* error codes */
#define PALIAL_OK            0
#define PALIAL_ERR_ARG      -1   /* bad argument (N % b != 0, NULL pointer, unsupported side/trans) */
#define PALIAL_ERR_HIP      -2   /* a HIP runtime call failed */
#define PALIAL_ERR_NOMEM    -3   /* device allocation failed */
#define PALIAL_ERR_NODEV    -4   /* no usable gfx950 device */
#define PALIAL_ERR_STATE    -5   /* library not initialised */
#define PALIAL_ERR_NOCONV   -6   /* iterative refinement did not reach the requested accuracy */

/* precision tier of the O(n^3) trailing updates */
#define PALIAL_PREC_FP64  0
#define PALIAL_PREC_FP32  1
#define PALIAL_PREC_BF16  2
#define PALIAL_PREC_FP8   3   /* E4M3 panels, one power-of-two scale per panel tile (largest magnitude in
                                 [64, 128)), FP32 accumulation; ~2^-4 relative per operand */
#define PALIAL_PREC_BF16X3 4  /* split BF16: each panel element as hi + lo BF16 terms, three MFMA
                                 products (hi*hi + hi*lo + lo*hi), ~2^-16 relative per product */

/* options; a NULL pointer means all defaults (zero-initialised struct) */
typedef struct palial_opts {
    int   device_matrix;  /* nonzero: MATRIX_desc.matrix (and S, B) are device pointers */
    int   precision;      /* PALIAL_PREC_*; factorization calls honour FP64 only */
    int   no_graph;       /* nonzero: launch stream-ordered instead of replaying a hipGraph */
    int   no_lookahead;   /* nonzero: single stream, no panel look-ahead */
    void *stream;         /* hipStream_t to run on (device_matrix only); NULL = library stream */
    int   refine;         /* solves: 0 = auto (LU-IR, GMRES-IR once LU-IR stalls), 1 = GMRES-IR, 2 = LU-IR only */
    int   max_iter;       /* solves: refinement steps (0 = 30) */
    double tol;           /* solves: target backward error (0 = 1e-14) */
    int   factor_tile;    /* solves: tile size of the internal factorization (0 = automatic) */
    int   no_fallback;    /* solves, mixed tiers: nonzero reports a breakdown or a refinement that
                             does not converge as an error; zero (default) retries a broken-down
                             Cholesky with a diagonal shift (then GMRES-IR) and finally solves in FP64 */
    int   schedule;       /* factorization calls, task priority / look-ahead of the tile DAG (0 = default):
                             Cholesky, LU: 1 diagonal chain on its own stream with the trailing update
                             split in two launches per step, 2 diagonal chain with one update launch,
                             3 row look-ahead, 4 one stream in the reference's order;
                             QR: 1..8 panel chains in flight (steps progressing together) */
    int   batch_grid;     /* palial_hip_batched: workgroups of the persistent launch (0 = as many as
                             the CUs hold at once) */
    int   sweep_block;    /* solves: tiles per step of the triangular correction sweeps, which apply
                             precomputed inverses of the diagonal blocks (0 = the library default,
                             PALIAL_SWEEP_BLOCK or 1; 1 = one tile; blocks of q > 1 tiles need q <= 8
                             dividing N / tile with N / tile >= 2q, <= 4 right-hand sides and
                             q * tile * (4 when nrhs > 1) <= 4096; a mixed tier's FP64 fallback
                             uses the default) */
    int   reserved[3];
} palial_opts;

/* per-call report; any pointer may be NULL */
typedef struct palial_report {
    double factor_ms;     /* device time of the factorization (hipEvents) */
    double solve_ms;      /* device time of solve + refinement */
    double h2d_ms;        /* host->device copies issued by this call */
    double d2h_ms;        /* device->host copies issued by this call */
    double gflop;         /* algorithmic flops of the factorization (N^3/3, 2N^3/3, 4N^3/3) / 1e9 */
    double tflops;        /* gflop / factor_ms */
    double backward_error;/* solves: ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) */
    int    info;          /* same as the return value when > 0 */
    int    iterations;    /* solves: refinement steps taken */
    long   ntasks;        /* tile tasks of the reference DAG this call covers */
    long   nlaunches;     /* kernel launches issued (or graph nodes replayed) */
    int    gmres_iterations; /* solves: GMRES inner iterations over all GMRES-IR steps */
    int    precision_used;   /* solves: tier of the factorization that produced the solution */
    int    shifts;           /* solves: diagonal shifts applied after a low-precision breakdown */
    int    reserved[5];
} palial_report;

/* ---------------------------------------------------------------- lifecycle */
int         palial_hip_init(int device);      /* select device, create streams/workspace */
int         palial_hip_finalize(void);
const char *palial_hip_strerror(int code);
const char *palial_hip_last_error(void);      /* text of the last failure (HIP message etc.) */
int         palial_hip_version(void);         /* 100*major + minor */

/* device memory helpers for C callers that pass device_matrix = 1 */
int palial_hip_malloc(void **dptr, size_t bytes);
int palial_hip_free(void *dptr);
int palial_hip_memcpy_h2d(void *dst, const void *src, size_t bytes);
int palial_hip_memcpy_d2h(void *dst, const void *src, size_t bytes);
int palial_hip_memcpy_d2d(void *dst, const void *src, size_t bytes);
int palial_hip_synchronize(void);             /* hipDeviceSynchronize */
int palial_hip_copy(void *dst, const void *src, size_t bytes);  /* device copy by the library's own streaming kernel */

/* Measurement hooks (no counterpart in the reference): per-kernel-class
 * HIP-event timing of the following algorithm calls, which then launch
 * stream-ordered instead of replaying a graph.  Classes:
 *   0 trailing-update GEMM/SYRK (FP64, or low-precision with FP64 operands),
 *   1 diagonal-tile POTRF/GETRF/GEQRT, 2 panel TRSM, 3 TSQRT,
 *   4 QR block-reflector GEMMs (ORMQR/TSMQR), 5 panel packing (mixed tiers),
 *   6 trailing update from packed low-precision panels.
 * flops / bytes are the algorithmic counts of the timed launches. */
int palial_hip_profile(int enable);   /* 1: per-class totals, 2: also one record per launch; resets */
/* Per-launch records kept under palial_hip_profile(2): 6 doubles each --
 * kernel class, tile size b, precision tier, algorithmic flops, algorithmic
 * bytes, device ms.  Copies at most cap records; returns how many exist. */
long palial_hip_profile_records(double *out, long cap);
int palial_hip_profile_stats(int kclass, long *launches, double *total_ms, double *flops,
                             double *bytes);

/* Per-task trace (the device-side counterpart of the reference's OMPT tracer,
 * palial_trace/palial_trace.c:21-71, log.h:43-50).  palial_hip_trace(1): the
 * following factorizations (plain, or inside a solve) launch stream-ordered
 * and record the device time of every reference tile task: tasks inside a
 * batched launch by device timestamps of their first and last workgroup,
 * the diagonal-tile and QR panel tasks by HIP events.  When a loaded tracer
 * defines the reference's hooks (include/common.h:9-11) they are called once
 * per task after the factorization, in task-creation order, with host times
 * (start of the call + device offsets).  palial_hip_trace_records copies the
 * last factorization's tasks in creation order, 6 doubles each: kind, k, m, n
 * (as palial_hip_task_name numbers them; unused -1), device ms (-1: not
 * measured), start in ms from the start of the factorization; it returns the
 * task count (negative count - 1 if the timestamp buffer overflowed). */
int  palial_hip_trace(int enable);
long palial_hip_trace_records(double *out, long cap);
/* Reference task list (no GPU needed): algo 0 Cholesky, 1 LU, 2 QR with MT
 * tiles per side, in the order the reference's loops create the tasks, with
 * the names they give them (kind followed by k, m, n as the loop variables
 * hold them then, src/cholesky.h:14-173, src/lu.h:13-173, src/qr.h:17-185).
 * Writes the name of task `index` to buf (when buf is not NULL) and returns
 * the task count.  Kinds: Cholesky 0 potrf(k), 1 trsm(k,m), 2 syrk(k,m),
 * 3 gemm(k,m,n); LU 4 getrf(k), 5 trsm(k,m) of tile (m,k), 6 trsm(k,n) of
 * tile (k,n), 7 gemm(k,m,n); QR 8 geqrt(k), 9 ormqr(k,n), 10 tsqrt(k,m),
 * 11 tsmqr(k,m,n). */
long palial_hip_task_name(int algo, int MT, long index, char *buf, int len);

/* Schedule settings that palial_hip_init reads from the environment, changed
 * at run time (no counterpart in the reference; for tests and tuners):
 * "pair" (PALIAL_PAIR: 1 pairs the trailing updates of consecutive steps of
 * the BF16 and split-BF16 Cholesky and LU solves), "sweep_block"
 * (PALIAL_SWEEP_BLOCK, 1..8).  Returns
 * the previous value, or PALIAL_ERR_ARG for an unknown name or a value out of
 * range.  Calls are serialised with the library's other calls. */
int  palial_hip_set_option(const char *name, int value);

/* CU sets per task class (no counterpart in the reference; the user-space GPU
 * analogue of palial_power/server.c:69-84, which caps a core's clock by the
 * type of the task it runs).  palial_hip_init reads PALIAL_CU_POLICY, e.g.
 * "bulk=8-255,aux=8-127,chain=0-7": entries name=ranges separated by ',',
 * name bulk (trailing updates), aux (row TRSMs and packing) or chain (the
 * diagonal chain), ranges "lo-hi" or "cu" joined by '+'.  The mixed tiers'
 * streams take these CU masks in place of PALIAL_RESERVE_CUS and
 * PALIAL_CHAIN_CUS; unnamed classes get all CUs (bulk), the bulk's set (aux)
 * or no mask (chain).  This function parses a policy for ncu CUs into three
 * masks of `words` 32-bit words (bulk, aux, chain) and returns 1 when the
 * chain is masked, 0 when not, PALIAL_ERR_ARG for a malformed policy. */
int  palial_hip_cu_policy(const char *policy, int ncu, unsigned *out, int words);

/* Test hook (no counterpart in the reference): enumerate the tile tasks of
 * one batched trailing-update launch (mode = internal GEMM mode, step k, TS
 * row m for the QR modes) without touching the GPU.  Writes 3 longs per task:
 * element offsets of the A, B and output tiles from the matrix base (-1 none,
 * -2 a workspace or T-factor tile).  Returns the task count, -1 if cap is too
 * small. */
int palial_hip_task_tiles(int mode, int MT, int b, int k, int m, long *out, int cap);

/* Test hooks of the blocked correction sweeps (no counterpart in the
 * reference), host views of what the solve launches, without the GPU:
 *  - palial_hip_sweep_plan: the tile_gemv launches of the two sweeps of a
 *    solve (algo 0 Cholesky, 1 LU; MT tiles per side, q tiles per step) in
 *    issue order, 7 ints each: mode (0 diagonal inverse, 1 Cholesky forward,
 *    2 LU forward, 3 backward), step, transposed, tasks, inverse set (0 U-side,
 *    1 L-side), input and output vector (0 r, 1 y, 2 x);
 *  - palial_hip_gemv_tiles: for each task of one such launch: output tile, the
 *    source range j0, j1, then q source-tile element offsets from the inverse
 *    base (mode 0) or the matrix base (-1 outside [j0, j1)): 3 + q longs;
 *  - palial_hip_gather_map: for each tile (K, i, j) of the q x q blocks
 *    (K-major, then i, then j): kind (0 zero, 1 tile inverse, 2 factor tile),
 *    transposed, source offset, destination offset: 4 longs;
 *  - palial_hip_blkinv_plan: the batched GEMM launches that complete the
 *    block inverses, 5 doubles each: mode, level, term, alpha, beta;
 *  - palial_hip_blkinv_tiles: for each task of one such launch: buffer (0 the
 *    inverses X, 1 the tiles of T, 2 the products W; -1 none) and element
 *    offset of the A, B, C and D operands: 8 longs.
 * Each returns the number of records (steps, tasks, tiles or launches), or
 * PALIAL_ERR_ARG when cap is too small or an argument is out of range. */
int palial_hip_sweep_plan(int algo, int MT, int q, int *out, int cap);
int palial_hip_gemv_tiles(int mode, int trans, int MT, int b, int q, int k, int ntasks, long *out, long cap);
int palial_hip_gather_map(int MT, int b, int q, int upper, int dtrans, long *out, long cap);
int palial_hip_blkinv_plan(int q, double *out, int cap);
int palial_hip_blkinv_tiles(int mode, int MT, int b, int q, int bd, int bt, long *out, long cap);

/* Test hook (no counterpart in the reference): a factorization schedule,
 * recorded on the host instead of launched: the diagonal-chain schedule of a
 * solve or factorization (algo 0 Cholesky, 1 LU; prec 0 the FP64 copy, 1..4
 * the FP32 working copy with packed panels; variant = pair + 2 * schedule,
 * pair 1: paired BF16 / split-BF16 updates, schedule 0 diagonal chain with the split
 * update, 1 chain with one update launch, 2 row look-ahead, 3 one stream, the
 * mixed tiers taking schedule 0 only) or the QR schedule
 * (algo 2, prec 0, variant = panel chains in flight, 1..8); MT tiles of b per
 * side.  Writes a stream of longs, one record per launch, event record or
 * event wait in issue order: type (0 launch, 1 record, 2 wait), stream (1
 * caller, 2 diagonal chain, 3 bulk, 4 row TRSMs and packing; QR: 10 + i panel
 * chain i, 20 + i update stream i), event, kind (1 diagonal factor, 2 TRSM,
 * 3 GEMM, 4 pack, 5 QR panel), mode (the internal GEMM mode; QR panel 0
 * GEQRT, 1 TSQRT), k, m, nr, nw, nx, nr resource ids read, nw written, and nx
 * groups of 4 (GEMM: output tile, step, A and B operand ids; TRSM: tile,
 * step, diagonal tile, inverses; pack: tile, step, packed id, -1; diagonal
 * factor: tile, step, inverses U, L or -1).  Resource ids: matrix tile
 * m + MT n; 2^24 + packed panel tile; 2^25 + diagonal-block inverses (2k U
 * side, 2k + 1 L side); 3 * 2^24 + T-factor tile, 4 * 2^24 and 5 * 2^24 + W
 * and op(T) W workspace tiles, 6 * 2^24 + tau and 7 * 2^24 + Gram workspace of
 * a panel chain.  Returns the number of longs (written when cap holds them),
 * PALIAL_ERR_ARG for arguments the schedule does not take. */
int palial_hip_schedule_ops(int algo, int MT, int b, int prec, int variant, long *out, long cap);

/* Test hook (no counterpart in the reference): average device time of one
 * batched trailing-update launch of step k (algo 0 Cholesky, 1 LU) over reps
 * launches on a device-resident matrix (its contents are overwritten). */
int palial_hip_time_update(int algo, MATRIX_desc A, int k, int reps, double *ms);

/* Test hook (no counterpart in the reference): applies the trailing update of
 * step k (algo 0 Cholesky: A_ij -= A_ki^T A_kj on the upper triangle, 1 LU:
 * A_ij -= A_ik A_kj, all tiles i, j > k) reps times to a device-resident
 * matrix.  prec 0 uses FP64 operands; 1 (FP32), 2 (BF16) and 4 (split BF16)
 * round the panel of step k once into packed K-contiguous copies and
 * accumulate in FP32, as the mixed-precision solves do (b % 128 == 0).
 * prec | 8 (with 1, 2 or 4): the
 * matrix is held as FP32 tiles in the first N*N*4 bytes of the buffer, as the
 * mixed-precision solves' working copy is.  *ms: average time per rep
 * including the packing. */
int palial_hip_update(int algo, MATRIX_desc A, int k, int prec, int reps, double *ms);

/* ---------------------------------------------------------------- inputs */
/* Seeded counter-based generators writing layout L (host or device per opts):
 *   kind 0: S + shift*I, S symmetric, S_ij = U(seed, min(i,j), max(i,j))
 *   kind 1: G + shift*I, G_ij = U(seed, i, j)
 *   kind 2: reference populate.h generator (include/populate.h:35-66) made
 *           reproducible: diagonal 173*(1 + U01), everything else 0.5
 *   kind 3: identity (include/populate.h:48,58)
 *   kind 4, 5: see PALIAL_GEN_GAUSS, PALIAL_GEN_GEOM_DIAG
 * U(seed,i,j) = (splitmix64(splitmix64(seed) + (i<<32 | j)) >> 11) * 2^-53 - 0.5 */
#define PALIAL_GEN_SYM 0
#define PALIAL_GEN_DENSE 1
#define PALIAL_GEN_PALIAL 2
#define PALIAL_GEN_IDENTITY 3
#define PALIAL_GEN_GAUSS 4      /* G_ij ~ N(0,1): h = splitmix64(splitmix64(seed) + (i<<32 | j)),
                                   u1 = ((h >> 11) + 1) 2^-53, u2 = (splitmix64(h) >> 11) 2^-53,
                                   G_ij = sqrt(-2 ln u1) cos(2 pi u2) (shift unused) */
#define PALIAL_GEN_GEOM_DIAG 5  /* diag(s), s_i = kappa^(-i/(N-1)) with kappa = shift >= 1 */
int palial_hip_generate(MATRIX_desc A, int kind, uint64_t seed, double shift,
                        const palial_opts *opts);

/* A = A1 (x) A2 (Kronecker product; A1 n1 x n1, A2 n2 x n2 column-major host
 * arrays, n1*n2 = N) written in layout L.  With A_i = Q_i D_i Q_i^T this is
 * Q diag(s) Q^T for Q = Q1 (x) Q2 and s = diag(D1) (x) diag(D2): a geometric
 * spectrum from 1 to 1/kappa factorises exactly this way. */
int palial_hip_generate_kron(MATRIX_desc A, const double *A1, int n1, const double *A2, int n2,
                             const palial_opts *opts);

/* ---------------------------------------------------------------- algorithms */
/* replaces void cholesky(MATRIX_desc A)   (src/cholesky.h:12)
 * Upper right-looking tiled Cholesky A = U^T U, in place; the strictly lower
 * part of diagonal tiles and the strictly lower tiles keep their input. */
int palial_hip_cholesky(MATRIX_desc A, const palial_opts *opts, palial_report *rep);

/* replaces void lu(MATRIX_desc A)          (src/lu.h:11)
 * Right-looking tiled LU without pivoting (src/srclu/dgetrf.h:20), in place:
 * unit-lower L below the diagonal, U on and above.  ipiv (may be NULL)
 * receives the identity permutation, 1-based, N entries. */
int palial_hip_lu(MATRIX_desc A, int *ipiv, const palial_opts *opts, palial_report *rep);

/* replaces void qr(MATRIX_desc A, MATRIX_desc S)   (src/qr.h:15)
 * Flat-tree TS Householder QR: R and V in A, T factors in the upper triangles
 * of S(k,k) and S(m,k); all other S entries are not written. */
int palial_hip_qr(MATRIX_desc A, MATRIX_desc S, const palial_opts *opts, palial_report *rep);

/* Q of a finished palial_hip_qr (V in A, T factors in S) applied from the
 * left to C, an N x ncols matrix in layout L with the same tile size
 * (N/b row tiles, ncols/b column tiles; tile (m,n) at b*b*(m + (N/b)*n)):
 * trans PALIAL_TRANS gives C <- Q^T C (the transformation qr() applies to
 * its own trailing columns, PALIAL_dormqr / PALIAL_dtsmqr at src/qr.h:43-182),
 * PALIAL_NOTRANS gives C <- Q C, which the reference's tile kernels do not
 * compute correctly (src/srcqr/dormqr.h:19-29, src/srcqr/dtsmqr.h:19-29).
 * side must be PALIAL_LEFT.  Q itself is formed by applying PALIAL_NOTRANS to
 * the identity (palial_hip_generate kind PALIAL_GEN_IDENTITY).
 * rep->solve_ms is the device time, rep->gflop 2 N^2 ncols. */
int palial_hip_qr_apply(int side, int trans, MATRIX_desc A, MATRIX_desc S, double *C, int ncols,
                        const palial_opts *opts, palial_report *rep);

/* Least-squares solution of A X = B from a finished palial_hip_qr (no
 * reference counterpart): X = R^-1 (Q^T B), which for the square nonsingular
 * A the reference factors is the solution of A X = B.  B (N x nrhs,
 * column-major, leading dimension ldb >= N, nrhs <= 16) holds the right-hand
 * sides on entry and X on return.  Requires b % 64 == 0 and b <= 512.  A and S
 * are not modified. */
int palial_hip_qr_solve(MATRIX_desc A, MATRIX_desc S, double *B, int nrhs, int ldb, const palial_opts *opts,
                        palial_report *rep);

/* In-place A <- A^T (average = 0) or A <- (A + A^T)/2 (average = 1) of a
 * layout-L matrix, b % 32 == 0 (no reference counterpart; builds the
 * symmetric test matrices Q diag(s) Q^T). */
int palial_hip_transpose(MATRIX_desc A, int average, const palial_opts *opts);

/* ---------------------------------------------------------------- solve */
/* No counterpart in the reference (which only factors): solve A X = B for
 * nrhs <= 16 right-hand sides with iterative refinement to the backward
 * error opts->tol (default 1e-14), eta = max_j ||b_j - A x_j||_inf /
 * (||A||_inf ||x_j||_inf + ||b_j||_inf).  algo 0: tiled Cholesky (A SPD),
 * 1: tiled LU without pivoting.  A (layout L, FP64) is not modified: a copy
 * is factored with its O(n^3) trailing updates in opts->precision (FP64,
 * FP32, BF16 or split-BF16 operands with FP32 accumulation; the mixed tiers
 * factor an FP32 copy with FP32 panel solves when the factor tile is a
 * multiple of 128); the diagonal-tile factorizations, the residual b - A x
 * and the solution update stay FP64.  PALIAL_PREC_FP8 and PALIAL_PREC_BF16X3
 * need a factor tile multiple of 128, N above it, and the default schedule.
 * B (column-major, leading dimension ldb >= N) holds the right-hand sides on
 * entry and the solution on return.  Requires b % 64 == 0 and b <= 512.
 * A mixed tier whose Cholesky breaks down is refactored with a diagonal
 * shift (then GMRES-IR), and a tier that still fails is solved again in FP64
 * (rep->precision_used, rep->shifts); with opts->no_fallback set it returns
 * the breakdown column (> 0) or PALIAL_ERR_NOCONV (with the reached backward
 * error in rep) when opts->max_iter steps do not reach the tolerance. */
int palial_hip_solve(int algo, MATRIX_desc A, double *B, int nrhs, int ldb, const palial_opts *opts,
                     palial_report *rep);

/* ---------------------------------------------------------------- batched */
/* Many independent small factorizations in one persistent launch (no
 * reference counterpart; for callers with batches of small systems): a
 * workgroup per CU takes matrices, largest first, from a work counter; the
 * launch is captured into a hipGraph once per argument set and replayed
 * (opts->no_graph: launched directly).  Matrix i has
 * order n[i] (1..512), is column-major with leading dimension n[i] and
 * starts at element off[i] of A; kind[i] 0 = Cholesky (A = U^T U, U in the
 * upper triangle, strict lower part untouched, as palial_hip_dpotrf), 1 = LU
 * without pivoting (as palial_hip_dgetrf).  prec PALIAL_PREC_FP64 (A and B
 * hold double) or PALIAL_PREC_FP32 (float; factored with FP64 arithmetic,
 * rounded on output).  B (optional) holds one right-hand side per matrix at
 * element boff[i]; it is overwritten by the solution.  info[i] (optional):
 * 0, or the first non-positive pivot (Cholesky, 1-based).  n, off, kind,
 * boff, info are host arrays; A and B follow opts->device_matrix.
 * rep->factor_ms is the device time of the launch. */
int palial_hip_batched(int count, const int *n, const long *off, const int *kind, void *A, void *B,
                       const long *boff, int *info, int prec, const palial_opts *opts, palial_report *rep);

/* Seeded inputs for a batch: kind 0 -> S + n I (S symmetric), kind 1 ->
 * G + n I, entries U(seed + i, r, c) as palial_hip_generate. */
int palial_hip_batched_generate(int count, const int *n, const long *off, const int *kind, void *A,
                                uint64_t seed, int prec, const palial_opts *opts);

/* ---------------------------------------------------------------- cost model */
/* Task-time cost model (no reference counterpart; for the autotuner,
 * palial_amd/autotune.py): a 3-layer perceptron with 64 hidden units in FP32,
 * y = relu(relu(x W1 + b1) W2 + b2) . w3 + b3, nfeat <= 16 input features.
 * params: W1 [nfeat x 64] row by row, b1 [64], W2 [64 x 64] row by row, b2 [64],
 * w3 [64], b3 [1]; palial_hip_mlp_nparams gives the count.  X is n x nfeat
 * row-major.  Pointers follow opts->device_matrix. */
int palial_hip_mlp_nparams(int nfeat);
int palial_hip_mlp_forward(int n, int nfeat, const float *X, const float *params, float *y,
                           const palial_opts *opts);
/* Adam on the mean squared error, mini-batches of 128 samples in order, fresh
 * optimizer state per call; loss[epoch] = mean squared error of the epoch
 * (before its steps complete).  epochs == 0: grad receives the full-batch
 * gradient of the mean squared error at params (params unchanged). */
int palial_hip_mlp_train(int n, int nfeat, const float *X, const float *y, float *params, int epochs,
                         float lr, float *loss, float *grad, const palial_opts *opts);

/* ---------------------------------------------------------------- tile kernels */
/* side / trans values of the reference (src/common.h:9-24; equal to the CBLAS
 * enum values it passes through) */
#define PALIAL_NOTRANS 111
#define PALIAL_TRANS   112
#define PALIAL_LEFT    141
#define PALIAL_RIGHT   142
/* Single-tile versions of the reference tile kernels; pointers follow
 * opts->device_matrix.  The leading dimensions must equal n (tile storage). */
int palial_hip_dpotrf(int n, double *A, int lda, const palial_opts *opts);      /* LAPACKE_dpotrf('U') at src/cholesky.h:40 */
int palial_hip_dgetrf(int n, double *A, int lda, const palial_opts *opts);      /* PALIAL_dgetrf, src/srclu/dgetrf.h:11 */
int palial_hip_dgeqrt(int n, int nb, double *A, int lda, double *T, int ldt,
                      double *tau, double *work, const palial_opts *opts);       /* PALIAL_dgeqrt, src/srcqr/dgeqrt.h:9 */
int palial_hip_dormqr(int side, int trans, int n, const double *V, int ldv,
                      const double *T, int ldt, double *C, int ldc, double *work,
                      int ldw, const palial_opts *opts);                         /* PALIAL_dormqr, src/srcqr/dormqr.h:9 */
int palial_hip_dtsqrt(int n, double *A1, int lda1, double *A2, int lda2, double *T,
                      int ldt, double *tau, double *work, const palial_opts *opts); /* PALIAL_dtsqrt, src/srcqr/dtsqrt.h:11 */
int palial_hip_dtsmqr(int side, int trans, int M, int N, int K, int ib, double *A1,
                      int lda1, double *A2, int lda2, const double *V, int ldv,
                      const double *T, int ldt, double *work, int ldw,
                      const palial_opts *opts);                                  /* PALIAL_dtsmqr, src/srcqr/dtsmqr.h:11 */
/* No counterpart in the reference: palial_hip_dtsmqr for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, semantics and results
 * (bitwise) as palial_hip_dtsmqr, but it allocates nothing: its two b x b
 * workspace tiles live in a static device array of the library (4 MiB per
 * device, held until the process ends), so M = N = K = ib <= 512 (larger
 * tiles return PALIAL_ERR_ARG), and A1, A2, V, T must be device pointers
 * (opts->device_matrix = 1; host tiles would need staging and return
 * PALIAL_ERR_ARG).  With opts->stream set the launches are enqueued on that
 * stream and the call returns without waiting; with opts->stream NULL it
 * runs on the library stream and waits.  Every call on one device shares
 * the same workspace: calls that overlap on one device (two streams, or two
 * threads with streams of their own) race. */
int palial_hip_dtsmqr_noalloc(int side, int trans, int M, int N, int K, int ib, double *A1,
                              int lda1, double *A2, int lda2, const double *V, int ldv,
                              const double *T, int ldt, double *work, int ldw,
                              const palial_opts *opts);

/* The BLAS tile tasks of the Cholesky and LU DAGs, as the reference calls them:
 *   PALIAL_TASK_CHOL_TRSM  C <- A^-T C, A upper          src/cholesky.h:73-84
 *   PALIAL_TASK_CHOL_SYRK  C <- C - A^T A (upper of C)    src/cholesky.h:115-125
 *   PALIAL_TASK_CHOL_GEMM  C <- C - A^T B                 src/cholesky.h:156-169
 *   PALIAL_TASK_LU_TRSM_R  C <- C A^-1, A upper           src/lu.h:67-78
 *   PALIAL_TASK_LU_TRSM_L  C <- A^-1 C, A unit lower      src/lu.h:109-120
 *   PALIAL_TASK_LU_GEMM    C <- C - A B                   src/lu.h:151-164
 * A, B, C are n x n tiles (ld = n); B is ignored by TRSM/SYRK. */
#define PALIAL_TASK_CHOL_TRSM 1
#define PALIAL_TASK_CHOL_SYRK 2
#define PALIAL_TASK_CHOL_GEMM 3
#define PALIAL_TASK_LU_TRSM_R 4
#define PALIAL_TASK_LU_TRSM_L 5
#define PALIAL_TASK_LU_GEMM   6
int palial_hip_tile_task(int task, int n, const double *A, const double *B, double *C,
                         const palial_opts *opts);

#ifdef __cplusplus
}
#endif
#endif /* PALIAL_HIP_H */
