// internal.hpp -- shared declarations of libpalial_hip.so (host + device).
//
// Tile layout L (reference include/tile_address.h:15): tile (m,n) of an N x N
// matrix with tile size b starts at element b*b*(m + (N/b)*n); inside a tile
// element (r,c) is at r + c*b.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace palial {

__host__ __device__ inline size_t tile_off(int b, int MT, int m, int n)
{
    return (size_t)b * (size_t)b * ((size_t)m + (size_t)MT * (size_t)n);
}

// ------------------------------------------------------------------ GEMM
// Task enumerations of the batched tile GEMM kernel.
enum GemmMode : int {
    GM_SINGLE = 0,      // one task: C = alpha op(A) op(B) + beta D ; E -= alpha op(A) op(B)
    GM_CHOL_TRAIL = 1,  // step k: C(n,m) -= A(k,n)^T A(k,m), k < n <= m < MT (diag: upper only)
    GM_LU_TRAIL = 2,    // step k: C(m,n) -= A(m,k) A(k,n),  m, n > k
    GM_QR_W = 3,        // TSMQR (k,m) over n>k: W_n = V^T A(m,n) + A(k,n), V = A(m,k)
    GM_QR_TW = 4,       // W2_n = T^T W_n (T = S(m,k) upper);  A(k,n) -= W2_n
    GM_QR_VW = 5,       // A(m,n) -= V W2_n
    GM_ORM_W = 6,       // ORMQR k over n>k: W_n = V^T A(k,n), V = A(k,k) unit lower
    GM_ORM_TW = 7,      // W2_n = T^T W_n (T = S(k,k) upper)
    GM_ORM_VW = 8,      // A(k,n) -= V W2_n
    GM_CHOL_ROW = 9,    // step k, only tiles (k+1, m), m >= k+1 (look-ahead part of GM_CHOL_TRAIL)
    GM_CHOL_REST = 10,  // step k, tiles (n, m) with k+2 <= n <= m
    GM_LU_PANEL = 11,   // step k, tiles (m, k+1) and (k+1, n) for m, n >= k+1 (look-ahead part)
    GM_LU_REST = 12,    // step k, tiles (m, n) with m, n >= k+2
    GM_CHOL_FIRST = 13, // GM_CHOL_ROW followed by tile (k+2, k+2) (task T)
    GM_LU_FIRST = 14,   // GM_LU_PANEL followed by tile (k+2, k+2) (task 2T-1)
    // Q of a finished factorization (V in mat, T in smat) applied from the
    // left to another layout-L matrix cmat with the same row tiles; task n =
    // column tile n of cmat (every column tile, not only n > k)
    GM_QA_W = 15,       // TS block (k,m): W_n = V^T C(m,n) + C(k,n), V = A(m,k)
    GM_QA_TW = 16,      // W2_n = op(T) W_n (T = S(m,k) upper);  C(k,n) -= W2_n
    GM_QA_VW = 17,      // C(m,n) -= V W2_n
    GM_OA_W = 18,       // diagonal block k: W_n = V^T C(k,n), V = A(k,k) unit lower
    GM_OA_TW = 19,      // W2_n = op(T) W_n (T = S(k,k) upper)
    GM_OA_VW = 20,      // C(k,n) -= V W2_n
    // inverses of the bq x bq tile blocks on the diagonal of a lower
    // triangular matrix T (blocked correction sweeps).  mat: the inverses,
    // block K's tile (i, j) at ((K bq + i) bq + j) b^2, diagonal tiles already
    // in place; w1: the tiles T(i, m), m < i, at the same offsets.  Level bd
    // (i = j + bd), task (K, j) = (task / (bq - bd), task % (bq - bd)):
    GM_BI_W = 21,       // term bt: W_task (+)= T(i, j + bt) X(j + bt, j)   (w2: W)
    GM_BI_X = 22,       // X(i, j) = -X(i, i) W_task
    // Cholesky steps k and k+1 paired (packed panels): the bulk update of
    // step k is split into GM_CHOL_FIRST, GM_CHOL_ROW2 and the part that
    // GM_CHOL_REST2 applies together with step k+1's (pa2 / pb2: panel k+1)
    GM_CHOL_ROW2 = 23,  // step k: tiles (k+2, m), m >= k+3, then tile (k+3, k+3) (task T-2)
    GM_CHOL_REST2 = 24, // steps k, k+1: tiles (n, m), k+3 <= n <= m, from panels k and k+1
    // the same for the LU (GM_LU_FIRST, GM_LU_ROW2, GM_LU_REST2)
    GM_LU_ROW2 = 25,    // step k: column k+2 (m >= k+3), row k+2 (n >= k+3), then tile (k+3, k+3)
    GM_LU_REST2 = 26,   // steps k, k+1: tiles (m, n), m, n >= k+3, from panels k and k+1
};

enum MaskA : int { MK_NONE = 0, MK_UPPER = 1, MK_UNIT_LOWER = 2 };

struct GemmArgs {
    int mode;            // GemmMode
    // GM_SINGLE operands; GM_ORM_W / GM_ORM_VW: A, when set, is V(k) as a plain
    // unit-lower tile (launch_unit_lower_copy) read instead of A(k,k) masked
    const double *A;
    const double *B;
    double *C;
    const double *D;
    double *E;
    // matrix modes
    double *mat;         // layout-L matrix the tasks address
    const double *smat;  // QR T-factor matrix (layout L)
    double *w1;          // workspace tiles, indexed by task
    double *w2;
    int N, b, MT;
    int k, m;            // step k; TS row m for the QR modes
    int T;               // trailing tile count used by the enumeration
    int ntasks;
    int task0;           // first task of the enumeration this launch covers
    const void *pa;      // packed low-precision panel operands (launch_gemm_packed):
    const void *pb;      // tile t of the panel at t*b*b elements, K-contiguous
    const void *pa2;     // GM_CHOL_REST2: the packed panel of step k+1 (tile t-1 pairs with
    const void *pb2;     // tile t of pa / pb)
    double alpha, beta;
    int maskA;
    int single_diag;     // GM_SINGLE: output is a SYRK tile (upper triangle only)
    int f32c;            // packed updates: C and D are FP32 tiles at the same element offsets from mat
    double *cmat;        // GM_QA_* / GM_OA_*: the matrix Q is applied to (layout L, MT row tiles)
    unsigned long long *ts;   // per-task tracing: timestamp slots (nullptr: not traced)
    long ts_base;             // slot of task 0 of this launch
    int bq, bd, bt;           // GM_BI_*: tiles per block, level, term
};

// ------------------------------------------------------------------ task trace
// Device timestamps of the tile tasks of a batched launch (per-task tracing
// only; ts == nullptr otherwise).  Slot s: ts[2s] = earliest start and
// ts[2s+1] = latest end over the task's workgroups, in ticks of the GPU's
// constant-rate wall clock.  The host fills starts with ~0 and ends with 0
// before the factorization.
__device__ inline void task_begin(unsigned long long *ts, long slot)
{
    if (ts && threadIdx.x == 0) atomicMin(ts + 2 * slot, (unsigned long long)wall_clock64());
}
// every thread of the workgroup calls it (it holds a barrier) after its last store
__device__ inline void task_end(unsigned long long *ts, long slot)
{
    if (!ts) return;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(ts + 2 * slot + 1, (unsigned long long)wall_clock64());
}

// host launchers (k_gemm.hip)
hipError_t launch_gemm(const GemmArgs &a, bool transA, bool transB, hipStream_t s);
// same with 32 x 32 output blocks of one wave each (b % 32 == 0): many
// workgroups for launches of few tiles
hipError_t launch_gemm_fine(const GemmArgs &a, bool transA, bool transB, hipStream_t s);
long gemm_task_count(int mode, int T);

// ------------------------------------------------------------------ panels
void panel_init();   // sets LDS limits of the panel kernels (call once, outside capture)
hipError_t launch_potrf(double *tile, int b, int *info, int info_base, hipStream_t s);
hipError_t launch_getrf(double *tile, int b, hipStream_t s);
// TRSM variants: 0 = Cholesky (U^T X = B, columns), 1 = LU left unit lower
// (L X = B, columns), 2 = LU right upper (X U = B, rows of the tile)
enum TrsmKind : int { TR_CHOL = 0, TR_LUL = 1, TR_LUR = 2 };
// Solves for tiles mat[tile_off(first + i*stride_m...)]; see k_panel.hip
struct TrsmArgs {
    const double *diag;  // the b x b diagonal tile
    double *mat;         // layout-L matrix
    double *single;      // single-tile target (ntiles == 1 && mat == nullptr)
    int N, b, MT, k;
    int first;           // first tile index along the panel (m or n)
    int ntiles;
    // batch mode (trsm2 only): tile t of the launch solves with diagonal tile
    // (t, t) of mat and dinv + t*dinv_stride, in place on single + t*b*b;
    // ident: the right-hand side is the identity (computes the tile inverse)
    int batch, ident;
    long dinv_stride;
    int f32;             // trsm2: the tiles of mat and diag are FP32 (single stays FP64)
    unsigned long long *ts;   // per-task tracing: timestamp slots (nullptr: not traced)
    long ts_base;             // slot of tile 0 of this launch
};
hipError_t launch_trsm(int kind, const TrsmArgs &a, hipStream_t s);

// Task trace, host side (api.hip): while a traced factorization is being
// issued, register a batched launch on the factorization's matrix and give it
// its timestamp slots (sets a.ts / a.ts_base); otherwise leave them as they are.
void trace_gemm(GemmArgs &a);
void trace_trsm(int kind, TrsmArgs &a);

// MFMA panel kernels for b % 32 == 0, b <= 512 (k_panel2.hip).  dinv holds
// the inverses of the 32x32 diagonal blocks of the factored tile (b/32 blocks
// of 32*32 doubles, column-major).
bool panel2_ok(int b);
void panel2_init();
// Batched small factorizations (k_panel2.hip): one persistent launch; matrix
// i (order n[i] <= 512, column-major, ld n[i]) at A + off[i] is copied into
// a 32-padded FP64 workspace (identity padding), factored there by the tile
// kernels' code (kind 0 Cholesky, 1 LU without pivoting), optionally solved
// for one right-hand side at B + boff[i], and copied back (rounded to FP32
// when f32).  All index arrays are device arrays.
struct BatchArgs {
    int count;
    const int *order;    // processing order of the matrices (largest first)
    const int *kind;     // per matrix: 0 Cholesky, 1 LU
    int *queue;          // work counter of the persistent launch (zero before each launch)
    const int *n;
    const long *off;
    const long *woff;    // workspace offsets (doubles)
    const long *boff;
    void *A;
    void *B;             // nullptr: factor only
    double *work;
    int *info;
    int f32;
};
// one persistent launch of `grid` workgroups (batch_grid: what the CUs hold
// at once for matrices of order <= bmax, at most count)
int batch_grid(int bmax, int count);
hipError_t launch_batch(const BatchArgs &a, int bmax, int grid, hipStream_t s);
// seeded batch inputs: kind 0 -> S + n I (S symmetric), 1 -> G + n I; seed of matrix i: seed + i
hipError_t launch_batch_generate(const int *n, const long *off, const int *kind, void *A, int count, int f32,
                                 uint64_t seed, hipStream_t s);

// MLP cost model (k_mlp.hip)
int mlp_nparams(int F);
void mlp_init();
hipError_t launch_mlp_forward(int n, int F, const float *X, const float *P, float *Y, hipStream_t s);
hipError_t launch_mlp_train(int n, int F, const float *X, const float *Y, float *P, float *m, float *v, int epochs,
                            float lr, float *loss, float *grad, hipStream_t s);

hipError_t launch_potrf2(double *tile, int b, int *info, int info_base, double *dinv, hipStream_t s);
hipError_t launch_getrf2(double *tile, int b, double *dinvU, double *dinvL, hipStream_t s);
hipError_t launch_trsm2(int kind, const TrsmArgs &a, const double *dinv, hipStream_t s);

// low-precision trailing updates (k_gemm_lp.hip): prec 1 = FP32, 2 = BF16
// Panel tiles of step k rounded to `prec` (1 FP32, 2 BF16, 3 FP8 with a
// per-tile power-of-two scale, 4 split BF16), K-contiguous: row-panel tiles
// (k, first..first+n-1) keep their layout, column-panel tiles (first.., k)
// are transposed; tile j lands at dst + (j-k-1)*b*b elements (FP8: in a slot
// of b*b + 512 bytes that also holds the scale; split BF16: 2*b*b elements).
// src32: mat holds FP32 tiles (the FP32 working copy; same element offsets)
hipError_t launch_pack(const double *mat, int N, int b, int k, int first, int n, bool colpanel, int prec,
                       void *dst, hipStream_t s, bool src32 = false);
// Trailing update from packed operands (a.pa / a.pb); TA/TB are implied.
hipError_t launch_gemm_packed(const GemmArgs &a, int prec, hipStream_t s);
hipError_t launch_gemm_lp(const GemmArgs &a, bool transA, bool transB, int prec, hipStream_t s);

// solve path (k_solve.hip)
enum GemvMode : int { GV_DIAG = 0, GV_CHOL_FWD = 1, GV_LU_FWD = 2, GV_BWD = 3 };
// A sweep step covers q consecutive tiles (q = 1 or 0: one tile): step k
// solves with the inverse of diagonal block k (tiles kq .. kq+q-1; q = 1:
// tile k at k*b*b; q > 1: lower triangular, tile (i, j) of block k at
// ((kq + i) q + j) b^2, used transposed when trans) and the update launches
// read the q tiles of the block's tile column (forward) or row (backward).
struct GemvArgs {
    int mode;
    const void *mat;     // factored matrix (layout L), double or float (f32)
    const void *inv;     // inverses of the diagonal tiles or blocks
    const double *vin;   // input vector block (N x nrhs, leading dimension ld)
    double *vout;        // output vector block
    int N, b, k, ld, nrhs, ntasks, trans;
    int f32;             // mat / inv hold FP32 copies
    int q;               // tiles per sweep step
};
// Tiles of the lower triangular blocks the blocked sweeps invert, as FP64:
// block k (q tiles) of T = F^T from the upper tiles (trans: Cholesky U and
// LU's U) or T = F from the lower tiles (LU's unit L).  Writes X(k, i, i) =
// op(D_{kq+i}) (D: the tile inverses, tile t at t*b*b; op = transpose for
// LU's U, whose tile inverses are upper), zero tiles X(k, i, j > i), and
// W1(k, i, m) = T(kq+i, kq+m) for m < i, all at ((kq + i) q + j) b^2.
hipError_t launch_blk_gather(const void *F, bool f32, int N, int b, int q, bool upper, bool dtrans,
                             const double *D, double *X, double *W1, hipStream_t s);
// What launch_blk_gather writes to tile (K, i, j) (the kernel and the host
// test hook share it): kind 0 zeros into X, 1 tile src of D into X, 2 tile
// src of F into W1; tr: transposed; dst: the offset in X or W1.
struct GatherTile {
    int kind;
    bool tr;
    size_t src, dst;
};
__host__ __device__ inline GatherTile gather_tile(int MT, int b, int q, bool upper, bool dtrans, int K, int i, int j)
{
    const size_t bb = (size_t)b * b;
    GatherTile g;
    g.dst = ((size_t)(K * q + i) * q + j) * bb;
    g.src = 0;
    g.tr = false;
    if (j > i) {
        g.kind = 0;
    } else if (j == i) {
        g.kind = 1;
        g.src = (size_t)(K * q + i) * bb;
        g.tr = dtrans;
    } else {
        g.kind = 2;
        g.src = upper ? tile_off(b, MT, K * q + j, K * q + i) : tile_off(b, MT, K * q + i, K * q + j);
        g.tr = upper;
    }
    return g;
}
// The batched GEMM launches of the block inverses of q-tile blocks, in issue
// order: level bd (1..q-1), term bt (GM_BI_W, bt < bd) then GM_BI_X.
struct BlkInvLaunch {
    int mode, bd, bt;
    double alpha, beta;
};
// (api.hip: blkinv_plan(q, out, cap) returns the launch count)
hipError_t launch_tile_gemv(const GemvArgs &a, hipStream_t s);

// Tile map of one tile_gemv task (the kernel and the host test hook share it):
// output tile vout and source tiles j0 <= j < j1 of step k (q tiles per step)
struct GemvMap {
    int vout, j0, j1;
};
__host__ __device__ inline GemvMap gemv_map(int mode, int trans, int k, int q, int task)
{
    const int s0 = k * q;
    GemvMap m;
    m.j0 = 0;
    m.j1 = q;
    switch (mode) {
    case GV_DIAG:
        m.vout = s0 + task;
        if (trans) m.j0 = task;      // block inverse read transposed: its tiles (j, task), j >= task
        else m.j1 = task + 1;        // lower block inverse: tiles (task, j), j <= task
        break;
    case GV_CHOL_FWD:
    case GV_LU_FWD: m.vout = s0 + q + task; break;
    default: m.vout = task; break;   // GV_BWD
    }
    return m;
}
// element offset of source tile j of a task: from the inverse base (GV_DIAG:
// block k's tile (vout - kq, j), or (j, vout - kq) when trans) or from the
// matrix base (U tile (kq + j, vout) for GV_CHOL_FWD; (vout, kq + j) else)
__host__ __device__ inline size_t gemv_src(int mode, int trans, int k, int q, int b, int MT, int vout, int j)
{
    const int s0 = k * q, i = vout - s0;
    const size_t bb = (size_t)b * b;
    switch (mode) {
    case GV_DIAG: return ((size_t)(s0 + (trans ? j : i)) * q + (trans ? i : j)) * bb;
    case GV_CHOL_FWD: return tile_off(b, MT, s0 + j, vout);
    default: return tile_off(b, MT, vout, s0 + j);
    }
}
// The launches of the two triangular sweeps (host side), in issue order.
// inv: 0 the U-side inverses (Cholesky (U^T)^-1 tiles or blocks; LU U^-1 tiles
// or (U^T)^-1 blocks), 1 the L-side ones; vin / vout: 0 r, 1 y, 2 x.
struct SweepStep {
    int mode, k, trans, ntasks, inv, vin, vout;
};
// (api.hip: sweep_plan(algo, MT, q, out, cap) returns the step count)
// FP32 copy of a layout-L matrix (N > 0; upper_only: tiles m <= n) or of n doubles (N == 0)
hipError_t launch_to_f32(const double *src, float *dst, int N, int b, bool upper_only, size_t n, hipStream_t s);
hipError_t launch_residual(const double *A, int N, int b, const double *x, const double *rhs, double *r, int ld,
                           int nrhs, double *part, bool sym, double *rowsum, hipStream_t s);
size_t residual_scratch(int N, int b, int nrhs);   // doubles of `part` launch_residual needs
void solve_init();   // LDS limits of the solve kernels (call once, outside capture)
// F(i,i) += sigma |A(i,i)|: F the working copy (tile b, FP32 when f32), A the FP64 matrix (tile bA)
hipError_t launch_add_diag(void *F, bool f32, int N, int b, const double *A, int bA, double sigma, hipStream_t s);
hipError_t launch_colmax(const double *v, int N, int ld, int nrhs, double *out, hipStream_t s);
hipError_t launch_axpy(double *x, const double *d, int N, int ld, int nrhs, hipStream_t s);
// GMRES: out[i] = V(:,i).w (i < k); w = alpha*w - V(:,0:k) h
hipError_t launch_multi_dot(const double *V, size_t ldv, int k, const double *w, int N, double *out, hipStream_t s);
hipError_t launch_vh_update(const double *V, size_t ldv, int k, const double *h, double *w, double alpha, int N,
                            hipStream_t s);

// QR panels.  G: b x b workspace (Gram blocks of the panel reflectors); with
// G != nullptr and b % 32 == 0, 32 <= b <= 512 the blocked kernels run
hipError_t launch_geqrt(double *A, double *T, double *tau, int b, hipStream_t s, double *G = nullptr);
hipError_t launch_tsqrt(double *A1, double *A2, double *T, double *tau, int b, hipStream_t s, double *G = nullptr);
// V <- the unit-lower V of a QR diagonal tile A (its strictly lower part, ones
// on the diagonal, zeros above) as a plain b x b tile; reads only A's strictly
// lower part, so it may run beside the TSQRTs that rewrite A's upper R
hipError_t launch_unit_lower_copy(double *V, const double *A, int b, hipStream_t s);

// misc
hipError_t launch_generate(double *mat, int N, int b, int kind, uint64_t seed, double shift,
                           hipStream_t s);
hipError_t launch_fill_identity_ipiv(int *ipiv, int N, hipStream_t s);
hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s);
hipError_t launch_relayout(double *dst, int b2, const double *src, int b1, int N, hipStream_t s);
// b x b column-major tile <-> its bs-sub-tiles in layout L (to_sub: tile -> sub)
hipError_t launch_tile_split(double *sub, double *tile, int b, int bs, bool to_sub, hipStream_t s,
                             bool tile32 = false);   // tile32: the tile is FP32 (sub stays FP64)
hipError_t launch_kron(double *mat, int N, int b, const double *A1, int n1, const double *A2, int n2, hipStream_t s);
// task trace: fill nslots slot pairs (start ~0, end 0); write the device wall clock to *dst
hipError_t launch_trace_init(unsigned long long *ts, size_t nslots, hipStream_t s);
hipError_t launch_clock_stamp(unsigned long long *dst, hipStream_t s);
// in-place A <- A^T (avg false) or A <- (A + A^T) / 2 (avg true) of a layout-L matrix, b % 32 == 0
hipError_t launch_transpose(double *mat, int N, int b, bool avg, hipStream_t s);
// N x nrhs column-major block (leading dimension ld) <-> N x b layout-L block
// of one tile column (columns nrhs..b-1 zero on the way in), nrhs <= b
hipError_t launch_cols_to_tiles(const double *v, int N, int ld, int nrhs, double *tiles, int b, hipStream_t s);
hipError_t launch_tiles_to_cols(const double *tiles, int N, int b, double *v, int ld, int nrhs, hipStream_t s);
// inverses of the 32x32 diagonal blocks of the upper triangles of the
// diagonal tiles of a layout-L matrix: tile k's block j at dinv + k*stride + j*1024
// (column-major, ld 32), as the factorization kernels write them (b % 32 == 0)
hipError_t launch_upper_block_inverses(const double *mat, int N, int b, double *dinv, long stride, hipStream_t s);

// ------------------------------------------------------------------ schedule recorder
// Host test hook: while sched_rec_on() the launchers of the factorization
// schedules record each launch (its stream and the tiles, packed panel tiles
// and diagonal-block inverses it reads and writes) instead of launching it,
// and the schedules record their event records and waits, so that the happens-
// before order of a whole schedule can be checked without a GPU.
bool sched_rec_on();
void rec_gemm(const GemmArgs &a, bool packed, hipStream_t s);
void rec_trsm(int kind, const TrsmArgs &t, const double *dinv, hipStream_t s);
void rec_pack(const double *mat, int b, int k, int first, int n, bool colpanel, const void *dst, hipStream_t s);
void rec_qr_panel(const double *A1, const double *A2, const double *T, const double *tau, const double *G,
                  hipStream_t s);
void rec_unit_lower_copy(const double *V, const double *A, hipStream_t s);

}  // namespace palial
