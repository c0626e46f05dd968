// k_solve.hip -- kernels of the solve / iterative-refinement path.
//
//   trtri_kernel   : full inverse of every diagonal tile of a factored
//                    matrix (upper U_kk^-1 or unit-lower L_kk^-1), one
//                    workgroup per tile, from the 32x32 block inverses the
//                    panel kernels left in dinv.  Runs once after the
//                    factorization, off its critical path, and turns each
//                    diagonal solve of the triangular sweeps into a GEMV.
//   tile_gemv      : batched out -= op(T) v (or out = op(T) v) over a list
//                    of b x b tiles of layout L, nrhs <= 16 right-hand
//                    sides; the two launches of one sweep step.
//   residual       : r = b - A x over the FP64 layout-L matrix (HBM-bound),
//                    chunked partial sums + a deterministic reduction.
//   rowsum_abs     : row sums of |A| (for ||A||_inf).
//   colmax_abs     : per-column max |v| of an N x nrhs block.
//   axpy_cols      : x += d.
#include "internal.hpp"

namespace palial {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma16s(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// ---------------------------------------------------------------- TRTRI
// X = U^{-1} for an upper-triangular b x b tile, column block by column
// block (32 wide):  X[0:j, j] = -X[0:j, 0:j] (U[0:j, j] X_jj),  X_jj = U_jj^{-1}.
// X = L^{-1} for a unit-lower tile, column blocks from the last:
//   X[j+1:, j] = -X[j+1:, j+1:] (L[j+1:, j] X_jj).
// Everything below (upper) / above (lower) the block diagonal is zero.
constexpr int TT_THREADS = 256;

template <bool UPPER>
__global__ __launch_bounds__(TT_THREADS) void trtri_kernel(const double *mat, int N, int b, const double *dinv,
                                                          double *inv)
{
    constexpr int NB = 32;
    __shared__ double sW[512 * NB];      // W = U[0:j,j] X_jj (upper) or L[j+1:,j] X_jj (lower), b <= 512
    const int k = blockIdx.x;
    const int MT = N / b;
    const double *T = mat + tile_off(b, MT, k, k);
    double *X = inv + (size_t)k * b * b;
    const double *Dk = dinv + (size_t)k * 2 * 32 * b + (UPPER ? 0 : (size_t)32 * b);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int nbk = b / NB;
    // zero + diagonal blocks
    for (size_t e = t; e < (size_t)b * b; e += TT_THREADS) {
        const int r = (int)(e % b), c = (int)(e / b);
        double v = 0.0;
        if (r / NB == c / NB) v = Dk[(size_t)(r / NB) * NB * NB + (r % NB) + (size_t)(c % NB) * NB];
        X[e] = v;
    }
    __syncthreads();
    for (int step = 1; step < nbk; step++) {
        const int j = UPPER ? step : nbk - 1 - step;       // column block being completed
        const int c0 = j * NB;
        const int r_lo = UPPER ? 0 : c0 + NB, r_hi = UPPER ? c0 : b;   // off-diagonal rows of column block j
        const int nr = r_hi - r_lo;
        const double *Xjj = Dk + (size_t)j * NB * NB;
        // W(r, c) = sum_q T(r_lo + r, c0 + q) Xjj(q, c)
        for (int e = t; e < nr * NB; e += TT_THREADS) {
            const int r = e % nr, c = e / nr;
            double s = 0.0;
            if (UPPER) {
                for (int q = 0; q <= c; q++) s += T[(size_t)(r_lo + r) + (size_t)(c0 + q) * b] * Xjj[q + c * NB];
            } else {
                for (int q = c; q < NB; q++) s += T[(size_t)(r_lo + r) + (size_t)(c0 + q) * b] * Xjj[q + c * NB];
            }
            sW[r + c * 512] = s;
        }
        __syncthreads();
        // X[r_lo:r_hi, j] = -X[r_lo:r_hi, r_lo:r_hi] W : 16x16 output tiles, MFMA,
        // K restricted to the triangle of X
        const int ntr = nr / 16;
        for (int e = w; e < ntr * 2; e += TT_THREADS / 64) {
            const int it = e >> 1, ct = e & 1;
            const int i0 = r_lo + it * 16;
            d4 acc = (d4){0.0, 0.0, 0.0, 0.0};
            const int kbeg = UPPER ? i0 : r_lo, kend = UPPER ? r_hi : i0 + 16;
            for (int k4 = kbeg; k4 < kend; k4 += 4) {
                const int q = k4 + (lane >> 4);
                acc = mfma16s(X[(size_t)(i0 + (lane & 15)) + (size_t)q * b], sW[(q - r_lo) + (ct * 16 + (lane & 15)) * 512], acc);
            }
#pragma unroll
            for (int v = 0; v < 4; v++)
                X[(size_t)(i0 + (lane >> 4) + 4 * v) + (size_t)(c0 + ct * 16 + (lane & 15)) * b] = -acc[v];
        }
        __syncthreads();
    }
}

hipError_t launch_trtri(const double *mat, int N, int b, const double *dinv, double *inv, bool upper, hipStream_t s)
{
    if (b > 512 || b % 32) return hipErrorInvalidValue;
    const int MT = N / b;
    if (upper) hipLaunchKernelGGL(trtri_kernel<true>, dim3(MT), dim3(TT_THREADS), 0, s, mat, N, b, dinv, inv);
    else hipLaunchKernelGGL(trtri_kernel<false>, dim3(MT), dim3(TT_THREADS), 0, s, mat, N, b, dinv, inv);
    return hipGetLastError();
}

// ---------------------------------------------------------------- tile GEMV
// For each task t: out_t(0:b, 0:nrhs) = op(T_t) v  (GV_DIAG) or
// out_t -= op(T_t) v (updates).  A workgroup (256 threads) owns 16 output
// rows of one tile; the input segment (b x nrhs) sits in LDS.
//   op = N: out(r) = sum_c T(r, c) v(c): 16 row lanes x 16 column groups,
//           a column of 16 rows is one 128-byte read
//   op = T: out(r) = sum_c T(c, r) v(c): a wave owns 4 rows; its 64 lanes
//           read 64 consecutive c of a row (512 bytes) and reduce by shuffles
constexpr int GV_THREADS = 256;
constexpr int GV_ROWS = 16;          // output rows per workgroup
constexpr int GV_MAXR = 16;          // right-hand sides

template <int NR>
__global__ __launch_bounds__(GV_THREADS) void tile_gemv_kernel(GemvArgs p)
{
    extern __shared__ double sm[];
    double *sv = sm;                          // b x NR input segment: sv[c + j*b]
    double *sred = sm + (size_t)p.b * NR;     // 16 groups x 16 rows x NR partial sums
    const int b = p.b;
    const int rb_per_tile = b / GV_ROWS;
    const int task = blockIdx.x / rb_per_tile, rb = blockIdx.x % rb_per_tile;
    const int t = threadIdx.x;
    const double *T;
    int vin, vout;
    {
        const int MT = p.N / b, k = p.k;
        switch (p.mode) {
        case GV_DIAG: T = p.inv + (size_t)k * b * b; vin = k; vout = k; break;
        case GV_CHOL_FWD: T = p.mat + tile_off(b, MT, k, k + 1 + task); vin = k; vout = k + 1 + task; break;
        case GV_LU_FWD: T = p.mat + tile_off(b, MT, k + 1 + task, k); vin = k; vout = k + 1 + task; break;
        default: T = p.mat + tile_off(b, MT, task, k); vin = k; vout = task; break;
        }
    }
    const double *vi = p.vin + (size_t)vin * b;
    double *vo = p.vout + (size_t)vout * b;
    for (int e = t; e < b * NR; e += GV_THREADS) {
        const int i = e % b, c = e / b;
        sv[e] = (c < p.nrhs) ? vi[(size_t)i + (size_t)c * p.ld] : 0.0;   // padded right-hand sides are zero
    }
    __syncthreads();
    const int r0 = rb * GV_ROWS;
    double acc[NR];
#pragma unroll
    for (int j = 0; j < NR; j++) acc[j] = 0.0;
    if (!p.trans) {
        const int r = t % GV_ROWS, g = t / GV_ROWS;       // g: column group, columns c = g, g+16, ...
        const double *Tr = T + r0 + r;
#pragma unroll 8
        for (int c = g; c < b; c += 16) {
            const double a = Tr[(size_t)c * b];
#pragma unroll
            for (int j = 0; j < NR; j++) acc[j] += a * sv[c + j * b];
        }
#pragma unroll
        for (int j = 0; j < NR; j++) sred[(g * GV_ROWS + r) * NR + j] = acc[j];
        __syncthreads();
        if (t < GV_ROWS * NR && t / GV_ROWS < p.nrhs) {
            const int rr = t % GV_ROWS, j = t / GV_ROWS;
            double s = 0.0;
            for (int q = 0; q < 16; q++) s += sred[(q * GV_ROWS + rr) * NR + j];
            double *o = vo + (size_t)(r0 + rr) + (size_t)j * p.ld;
            if (p.mode == GV_DIAG) *o = s;
            else *o -= s;
        }
    } else {
        const int lane = t & 63, wv = t >> 6;
        double part[4][NR];
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
            for (int j = 0; j < NR; j++) part[x][j] = 0.0;
        for (int c0 = 0; c0 < b; c0 += 64) {
            const int c = c0 + lane;
            double a[4];
#pragma unroll
            for (int x = 0; x < 4; x++) a[x] = T[(size_t)c + (size_t)(r0 + wv * 4 + x) * b];
#pragma unroll
            for (int x = 0; x < 4; x++)
#pragma unroll
                for (int j = 0; j < NR; j++) part[x][j] += a[x] * sv[c + j * b];
        }
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
            for (int j = 0; j < NR; j++) {
                double v = part[x][j];
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
                part[x][j] = v;
            }
        if (lane < 4 * NR && lane % NR < p.nrhs) {
            const int x = lane / NR, j = lane % NR;
            double s = 0.0;
#pragma unroll
            for (int xx = 0; xx < 4; xx++)
#pragma unroll
                for (int jj = 0; jj < NR; jj++)
                    if (xx == x && jj == j) s = part[xx][jj];
            double *o = vo + (size_t)(r0 + wv * 4 + x) + (size_t)j * p.ld;
            if (p.mode == GV_DIAG) *o = s;
            else *o -= s;
        }
    }
}

hipError_t launch_tile_gemv(const GemvArgs &a, hipStream_t s)
{
    if (a.ntasks <= 0) return hipSuccess;
    if (a.b > 512 || a.b % 64 || a.nrhs > GV_MAXR || a.nrhs < 1) return hipErrorInvalidValue;
    dim3 grid((unsigned)(a.ntasks * (a.b / GV_ROWS))), block(GV_THREADS);
    auto lds = [&](int nr) { return ((size_t)a.b * nr + 16 * GV_ROWS * nr) * 8; };
    if (a.nrhs == 1) hipLaunchKernelGGL(tile_gemv_kernel<1>, grid, block, lds(1), s, a);
    else if (a.nrhs <= 4) hipLaunchKernelGGL(tile_gemv_kernel<4>, grid, block, lds(4), s, a);
    else hipLaunchKernelGGL(tile_gemv_kernel<16>, grid, block, lds(16), s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------- residual
// r = b - A x for the FP64 layout-L matrix in two deterministic passes:
// partial sums over chunks of RES_CHUNK tile columns, then the reduction.
constexpr int RES_CHUNK = 8;

template <int NR>
__global__ __launch_bounds__(256) void residual_partial_kernel(const double *A, int N, int bt, const double *x,
                                                               int ld, int nrhs, double *part)
{
    __shared__ double sred[4][64 * NR];
    const int MT = N / bt;
    const int rows_per = bt / 64;
    const int nchunk = (MT + RES_CHUNK - 1) / RES_CHUNK;
    const int rowblk = blockIdx.x / nchunk, chunk = blockIdx.x % nchunk;
    const int m = rowblk / rows_per, r0 = (rowblk % rows_per) * 64;
    const int t = threadIdx.x, rr = t % 64, q = t / 64;
    double acc[NR];
#pragma unroll
    for (int j = 0; j < NR; j++) acc[j] = 0.0;
    const int n1 = min(MT, (chunk + 1) * RES_CHUNK);
    for (int n = chunk * RES_CHUNK; n < n1; n++) {
        const double *T = A + tile_off(bt, MT, m, n) + r0 + rr;
        const double *xs = x + (size_t)n * bt;
#pragma unroll 8
        for (int c = q; c < bt; c += 4) {
            const double a = T[(size_t)c * bt];
#pragma unroll
            for (int j = 0; j < NR; j++)
                if (j < nrhs) acc[j] += a * xs[(size_t)c + (size_t)j * ld];
        }
    }
#pragma unroll
    for (int j = 0; j < NR; j++) sred[q][rr + j * 64] = acc[j];
    __syncthreads();
    for (int e = t; e < 64 * nrhs; e += 256) {
        const int i = e % 64, j = e / 64;
        // part layout: [chunk][j][row]
        part[((size_t)chunk * NR + j) * N + (size_t)m * bt + r0 + i] = sred[0][e] + sred[1][e] + sred[2][e] + sred[3][e];
    }
}

template <int NR>
__global__ void residual_reduce_kernel(const double *part, int nchunk, const double *rhs, double *r, int N, int ld,
                                       int nrhs)
{
    const size_t total = (size_t)N * nrhs;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t i = e % N, j = e / N;
        double s = 0.0;
        for (int c = 0; c < nchunk; c++) s += part[((size_t)c * NR + j) * N + i];
        r[i + j * ld] = rhs[i + j * ld] - s;
    }
}

hipError_t launch_residual(const double *A, int N, int b, const double *x, const double *rhs, double *r, int ld,
                           int nrhs, double *part, hipStream_t s)
{
    if (b % 64 || nrhs > GV_MAXR || nrhs < 1) return hipErrorInvalidValue;
    const int MT = N / b;
    const int nchunk = (MT + RES_CHUNK - 1) / RES_CHUNK;
    dim3 grid((unsigned)(MT * (b / 64) * nchunk)), block(256);
    if (nrhs == 1) {
        hipLaunchKernelGGL(residual_partial_kernel<1>, grid, block, 0, s, A, N, b, x, ld, nrhs, part);
        hipLaunchKernelGGL(residual_reduce_kernel<1>, dim3(1024), dim3(256), 0, s, part, nchunk, rhs, r, N, ld, nrhs);
    } else {
        hipLaunchKernelGGL(residual_partial_kernel<16>, grid, block, 0, s, A, N, b, x, ld, nrhs, part);
        hipLaunchKernelGGL(residual_reduce_kernel<16>, dim3(1024), dim3(256), 0, s, part, nchunk, rhs, r, N, ld, nrhs);
    }
    return hipGetLastError();
}

size_t residual_scratch(int N, int b, int nrhs)
{
    const int MT = N / b;
    const int nchunk = (MT + RES_CHUNK - 1) / RES_CHUNK;
    return (size_t)nchunk * (nrhs == 1 ? 1 : 16) * N;
}

// ||A||_inf: row sums of |A|, then max (two passes)
__global__ __launch_bounds__(256) void rowsum_abs_kernel(const double *A, int N, int bt, double *rs)
{
    __shared__ double sred[4][64];
    const int MT = N / bt;
    const int rows_per = bt / 64;
    const int m = blockIdx.x / rows_per, r0 = (blockIdx.x % rows_per) * 64;
    const int t = threadIdx.x, rr = t % 64, q = t / 64;
    double acc = 0.0;
    for (int n = 0; n < MT; n++) {
        const double *T = A + tile_off(bt, MT, m, n) + r0 + rr;
        for (int c = q; c < bt; c += 4) acc += fabs(T[(size_t)c * bt]);
    }
    sred[q][rr] = acc;
    __syncthreads();
    if (t < 64) rs[(size_t)m * bt + r0 + t] = sred[0][t] + sred[1][t] + sred[2][t] + sred[3][t];
}

// per-column max |v| of an N x nrhs block (one workgroup per column)
__global__ __launch_bounds__(256) void colmax_abs_kernel(const double *v, int N, int ld, double *out)
{
    __shared__ double sred[256];
    const int j = blockIdx.x, t = threadIdx.x;
    double m = 0.0;
    for (int i = t; i < N; i += 256) m = fmax(m, fabs(v[(size_t)i + (size_t)j * ld]));
    sred[t] = m;
    __syncthreads();
    for (int s = 128; s >= 1; s >>= 1) {
        if (t < s) sred[t] = fmax(sred[t], sred[t + s]);
        __syncthreads();
    }
    if (t == 0) out[j] = sred[0];
}

hipError_t launch_norm_inf(const double *A, int N, int b, double *rowsum, double *out, hipStream_t s)
{
    if (b % 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(rowsum_abs_kernel, dim3((unsigned)((N / b) * (b / 64))), dim3(256), 0, s, A, N, b, rowsum);
    hipLaunchKernelGGL(colmax_abs_kernel, dim3(1), dim3(256), 0, s, rowsum, N, N, out);
    return hipGetLastError();
}

hipError_t launch_colmax(const double *v, int N, int ld, int nrhs, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(colmax_abs_kernel, dim3((unsigned)nrhs), dim3(256), 0, s, v, N, ld, out);
    return hipGetLastError();
}

__global__ void axpy_cols_kernel(double *x, const double *d, int N, int ld, int nrhs)
{
    const size_t total = (size_t)N * nrhs;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t i = e % N, j = e / N;
        x[i + j * ld] += d[i + j * ld];
    }
}

hipError_t launch_axpy(double *x, const double *d, int N, int ld, int nrhs, hipStream_t s)
{
    hipLaunchKernelGGL(axpy_cols_kernel, dim3(2048), dim3(256), 0, s, x, d, N, ld, nrhs);
    return hipGetLastError();
}

}  // namespace palial
