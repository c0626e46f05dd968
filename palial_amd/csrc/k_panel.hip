// k_panel.hip -- diagonal-tile factorizations and batched triangular solves.
//
//   potrf_kernel : upper Cholesky of one b x b tile (LAPACKE_dpotrf('U') at
//                  src/cholesky.h:40); strictly-lower part left untouched.
//   getrf_kernel : LU without pivoting of one tile (PALIAL_dgetrf,
//                  src/srclu/dgetrf.h:11 -> dgetrf_no_pivot.h:11), keeping the
//                  reference's pivot handling: multiply by 1/alpha when
//                  |alpha| > sfmin, divide otherwise, skip when alpha == 0.
//   trsm_kernel  : every panel TRSM of one DAG step in a single launch
//                  (src/cholesky.h:73-84, src/lu.h:67-78, src/lu.h:109-120).
//
// Both factorizations are right-looking with NB-column blocks inside the
// tile: the NB x NB diagonal block is factored by one wave holding one column
// per lane (cross-lane data via shuffles), the panel solve is one column (or
// row) per thread, and the rank-NB trailing update runs out of LDS in 4x4
// register blocks.
#include "internal.hpp"
#include <float.h>

namespace palial {

constexpr int PANEL_THREADS = 512;

// ---------------------------------------------------------------- POTRF
template <int NB>
__global__ __launch_bounds__(PANEL_THREADS) void potrf_kernel(double *A, int b, int *info,
                                                              int info_base)
{
    extern __shared__ double sp[];         // NB x b row panel: sp[q*b + c]
    __shared__ double sd[NB * NB];         // factored diagonal block (column-major)
    __shared__ int s_fail;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) s_fail = 0;
    __syncthreads();

    for (int j0 = 0; j0 < b; j0 += NB) {
        // 1. diagonal block, wave 0, lane c holds column c
        if (w == 0) {
            const int c = lane;
            double a[NB];
#pragma unroll
            for (int r = 0; r < NB; r++)
                a[r] = (c < NB && r <= c) ? A[(size_t)(j0 + r) + (size_t)(j0 + c) * b] : 0.0;
            int fail = 0;
#pragma unroll
            for (int j = 0; j < NB; j++) {
                double s = a[j];
#pragma unroll
                for (int p = 0; p < j; p++) s -= __shfl(a[p], j) * a[p];
                const double sj = __shfl(s, j);
                if (!(sj > 0.0)) {
                    if (!fail) fail = j + 1;
                    if (c == j) a[j] = sj;
                } else if (!fail) {
                    const double d = sqrt(sj);
                    if (c == j) a[j] = d;
                    else if (c > j) a[j] = s / d;
                }
            }
            if (c < NB) {
#pragma unroll
                for (int r = 0; r < NB; r++) {
                    if (r <= c) {
                        A[(size_t)(j0 + r) + (size_t)(j0 + c) * b] = a[r];
                        sd[r + c * NB] = a[r];
                    }
                }
            }
            if (lane == 0 && fail) {
                s_fail = 1;
                if (info) *info = info_base + j0 + fail;
            }
        }
        __syncthreads();
        if (s_fail) return;

        // 2. row panel: P <- U_D^{-T} P, one column per thread
        for (int c = j0 + NB + t; c < b; c += PANEL_THREADS) {
            double x[NB];
            const double *col = A + (size_t)c * b + j0;
#pragma unroll
            for (int r = 0; r < NB; r++) x[r] = col[r];
#pragma unroll
            for (int r = 0; r < NB; r++) {
                double s = x[r];
#pragma unroll
                for (int p = 0; p < r; p++) s -= sd[p + r * NB] * x[p];
                x[r] = s / sd[r + r * NB];
            }
#pragma unroll
            for (int r = 0; r < NB; r++) {
                A[(size_t)(j0 + r) + (size_t)c * b] = x[r];
                sp[r * b + c] = x[r];
            }
        }
        __syncthreads();

        // 3. trailing upper update: A(i,j) -= sum_q P(q,i) P(q,j), 4x4 blocks
        const int base = j0 + NB;
        const int nb4 = (b - base) / 4;
        const int ntri = nb4 * (nb4 + 1) / 2;
        for (int e = t; e < ntri; e += PANEL_THREADS) {
            int bj = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
            while ((bj + 1) * (bj + 2) / 2 <= e) bj++;
            while (bj * (bj + 1) / 2 > e) bj--;
            const int bi = e - bj * (bj + 1) / 2;
            const int i = base + 4 * bi, j = base + 4 * bj;
            double acc[4][4];
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int s = 0; s < 4; s++) acc[r][s] = 0.0;
#pragma unroll 4
            for (int q = 0; q < NB; q++) {
                const double *row = sp + q * b;
                double pi[4], pj[4];
#pragma unroll
                for (int r = 0; r < 4; r++) { pi[r] = row[i + r]; pj[r] = row[j + r]; }
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int s = 0; s < 4; s++) acc[r][s] += pi[r] * pj[s];
            }
#pragma unroll
            for (int s = 0; s < 4; s++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (bi < bj || r <= s) A[(size_t)(i + r) + (size_t)(j + s) * b] -= acc[r][s];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- GETRF
__device__ inline double piv_scale(double x, double alpha)
{
    // src/srclu/dgetrf_no_pivot.h:21-39
    if (alpha == 0.0) return x;
    if (fabs(alpha) > DBL_MIN) return x * (1.0 / alpha);
    return x / alpha;
}

template <int NB>
__global__ __launch_bounds__(PANEL_THREADS) void getrf_kernel(double *A, int b)
{
    extern __shared__ double sm[];
    double *sl = sm;            // sl[q*b + i] = L(i, j0+q)
    double *su = sm + NB * b;   // su[q*b + j] = U(j0+q, j)
    __shared__ double sd[NB * NB];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;

    for (int j0 = 0; j0 < b; j0 += NB) {
        if (w == 0) {
            const int c = lane;
            double a[NB];
#pragma unroll
            for (int r = 0; r < NB; r++) a[r] = (c < NB) ? A[(size_t)(j0 + r) + (size_t)(j0 + c) * b] : 0.0;
#pragma unroll
            for (int i = 0; i < NB; i++) {
                const double alpha = __shfl(a[i], i);
                if (c == i) {
#pragma unroll
                    for (int r = 0; r < NB; r++)
                        if (r > i) a[r] = piv_scale(a[r], alpha);
                }
#pragma unroll
                for (int r = 0; r < NB; r++) {
                    if (r > i) {
                        const double l = __shfl(a[r], i);
                        if (c > i) a[r] -= l * a[i];
                    }
                }
            }
            if (c < NB) {
#pragma unroll
                for (int r = 0; r < NB; r++) {
                    A[(size_t)(j0 + r) + (size_t)(j0 + c) * b] = a[r];
                    sd[r + c * NB] = a[r];
                }
            }
        }
        __syncthreads();

        const int base = j0 + NB;
        const int rest = b - base;
        // 2. L panel rows (x U_D = a) and U panel columns (L_D y = a)
        for (int e = t; e < 2 * rest; e += PANEL_THREADS) {
            if (e < rest) {
                const int i = base + e;
                double x[NB];
#pragma unroll
                for (int q = 0; q < NB; q++) x[q] = A[(size_t)i + (size_t)(j0 + q) * b];
#pragma unroll
                for (int q = 0; q < NB; q++) {
                    double s = x[q];
#pragma unroll
                    for (int p = 0; p < q; p++) s -= x[p] * sd[p + q * NB];
                    x[q] = piv_scale(s, sd[q + q * NB]);
                }
#pragma unroll
                for (int q = 0; q < NB; q++) {
                    A[(size_t)i + (size_t)(j0 + q) * b] = x[q];
                    sl[q * b + i] = x[q];
                }
            } else {
                const int j = base + (e - rest);
                double y[NB];
                const double *col = A + (size_t)j * b + j0;
#pragma unroll
                for (int r = 0; r < NB; r++) y[r] = col[r];
#pragma unroll
                for (int r = 0; r < NB; r++) {
                    double s = y[r];
#pragma unroll
                    for (int p = 0; p < r; p++) s -= sd[r + p * NB] * y[p];
                    y[r] = s;
                }
#pragma unroll
                for (int r = 0; r < NB; r++) {
                    A[(size_t)(j0 + r) + (size_t)j * b] = y[r];
                    su[r * b + j] = y[r];
                }
            }
        }
        __syncthreads();

        // 3. trailing update A(i,j) -= sum_q L(i,q) U(q,j)
        const int nb4 = rest / 4;
        const int nblk = nb4 * nb4;
        for (int e = t; e < nblk; e += PANEL_THREADS) {
            const int bi = e % nb4, bj = e / nb4;
            const int i = base + 4 * bi, j = base + 4 * bj;
            double acc[4][4];
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int s = 0; s < 4; s++) acc[r][s] = 0.0;
#pragma unroll 4
            for (int q = 0; q < NB; q++) {
                double li[4], uj[4];
#pragma unroll
                for (int r = 0; r < 4; r++) { li[r] = sl[q * b + i + r]; uj[r] = su[q * b + j + r]; }
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int s = 0; s < 4; s++) acc[r][s] += li[r] * uj[s];
            }
#pragma unroll
            for (int s = 0; s < 4; s++)
#pragma unroll
                for (int r = 0; r < 4; r++) A[(size_t)(i + r) + (size_t)(j + s) * b] -= acc[r][s];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- TRSM
// Forward substitution M X = B with M lower triangular, on a strip of SW
// right-hand sides held in LDS (sx[i*SW + c]).
//   TR_CHOL: M = U^T (M(i,q) = D[q + i*b]), X(i,c) = tile[i + c*b]
//   TR_LUL : M = L unit (M(i,q) = D[i + q*b]), X(i,c) = tile[i + c*b]
//   TR_LUR : M = U^T (X U = B transposed), X(i,c) = tile[c + i*b]
constexpr int TRSM_SW = 16;
constexpr int TRSM_THREADS = 256;

template <int KIND>
__global__ __launch_bounds__(TRSM_THREADS) void trsm_kernel(TrsmArgs p)
{
    extern __shared__ double sx[];
    __shared__ double sdb[16 * 16];          // current 16x16 diagonal block of M: sdb[r*16+q] = M(r0+r, r0+q)
    constexpr int SW = TRSM_SW;
    const int b = p.b;
    const int strips = b / SW;
    const int tix = blockIdx.x / strips, strip = blockIdx.x % strips;
    const int c0 = strip * SW;
    const int t = threadIdx.x;
    double *X;
    if (p.mat) {
        const int idx = p.first + tix;
        X = (KIND == TR_LUR) ? p.mat + tile_off(b, p.MT, idx, p.k) : p.mat + tile_off(b, p.MT, p.k, idx);
    } else {
        X = p.single;
    }
    const double *D = p.diag;
    auto Mat = [&](int i, int q) -> double {
        return (KIND == TR_LUL) ? D[(size_t)i + (size_t)q * b] : D[(size_t)q + (size_t)i * b];
    };
    task_begin(p.ts, p.ts_base + tix);

    if (KIND == TR_LUR) {
        for (int e = t; e < b * SW; e += TRSM_THREADS) {
            const int c = e % SW, i = e / SW;
            sx[i * SW + c] = X[(size_t)(c0 + c) + (size_t)i * b];
        }
    } else {
        for (int e = t; e < b * SW; e += TRSM_THREADS) {
            const int i = e % b, c = e / b;
            sx[i * SW + c] = X[(size_t)i + (size_t)(c0 + c) * b];
        }
    }
    __syncthreads();

    for (int r0 = 0; r0 < b; r0 += 16) {
        {   // stage the 16x16 diagonal block of M
            const int r = t / 16, q = t % 16;
            sdb[r * 16 + q] = (q <= r) ? Mat(r0 + r, r0 + q) : 0.0;
        }
        __syncthreads();
        if (t < SW) {
            const int c = t;
            double x[16];
#pragma unroll
            for (int r = 0; r < 16; r++) x[r] = sx[(r0 + r) * SW + c];
#pragma unroll
            for (int r = 0; r < 16; r++) {
                double s = x[r];
#pragma unroll
                for (int q = 0; q < r; q++) s -= sdb[r * 16 + q] * x[q];
                x[r] = (KIND == TR_LUL) ? s : s / sdb[r * 16 + r];
            }
#pragma unroll
            for (int r = 0; r < 16; r++) sx[(r0 + r) * SW + c] = x[r];
        }
        __syncthreads();
        const int below = b - r0 - 16;
        for (int e = t; e < below * SW; e += TRSM_THREADS) {
            const int i = r0 + 16 + e / SW, c = e % SW;
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 16; q++) s += Mat(i, r0 + q) * sx[(r0 + q) * SW + c];
            sx[i * SW + c] -= s;
        }
        __syncthreads();
    }

    if (KIND == TR_LUR) {
        for (int e = t; e < b * SW; e += TRSM_THREADS) {
            const int c = e % SW, i = e / SW;
            X[(size_t)(c0 + c) + (size_t)i * b] = sx[i * SW + c];
        }
    } else {
        for (int e = t; e < b * SW; e += TRSM_THREADS) {
            const int i = e % b, c = e / b;
            X[(size_t)i + (size_t)(c0 + c) * b] = sx[i * SW + c];
        }
    }
    task_end(p.ts, p.ts_base + tix);
}

// ---------------------------------------------------------------- tiny tiles
// Tile sizes that are not a multiple of 16 (the reference's 4x4 test matrix
// uses b = 2): one thread runs the unblocked algorithm.
__global__ void potrf_tiny_kernel(double *A, int b, int *info, int info_base)
{
    for (int j = 0; j < b; j++) {
        double s = A[j + (size_t)j * b];
        for (int p = 0; p < j; p++) s -= A[p + (size_t)j * b] * A[p + (size_t)j * b];
        if (!(s > 0.0)) {
            A[j + (size_t)j * b] = s;
            if (info) *info = info_base + j + 1;
            return;
        }
        const double d = sqrt(s);
        A[j + (size_t)j * b] = d;
        for (int c = j + 1; c < b; c++) {
            double x = A[j + (size_t)c * b];
            for (int p = 0; p < j; p++) x -= A[p + (size_t)j * b] * A[p + (size_t)c * b];
            A[j + (size_t)c * b] = x / d;
        }
    }
}

__global__ void getrf_tiny_kernel(double *A, int b)
{
    for (int i = 0; i < b; i++) {
        const double alpha = A[i + (size_t)i * b];
        for (int r = i + 1; r < b; r++) A[r + (size_t)i * b] = piv_scale(A[r + (size_t)i * b], alpha);
        for (int c = i + 1; c < b; c++)
            for (int r = i + 1; r < b; r++) A[r + (size_t)c * b] -= A[r + (size_t)i * b] * A[i + (size_t)c * b];
    }
}

__global__ void trsm_tiny_kernel(int kind, TrsmArgs p)
{
    const int b = p.b;
    const int tix = blockIdx.x;
    const int c = threadIdx.x;
    if (c >= b) return;
    double *X;
    if (p.mat) {
        const int idx = p.first + tix;
        X = (kind == TR_LUR) ? p.mat + tile_off(b, p.MT, idx, p.k) : p.mat + tile_off(b, p.MT, p.k, idx);
    } else {
        X = p.single;
    }
    const double *D = p.diag;
    for (int i = 0; i < b; i++) {
        const size_t oi = (kind == TR_LUR) ? (size_t)c + (size_t)i * b : (size_t)i + (size_t)c * b;
        double s = X[oi];
        for (int q = 0; q < i; q++) {
            const size_t oq = (kind == TR_LUR) ? (size_t)c + (size_t)q * b : (size_t)q + (size_t)c * b;
            const double m = (kind == TR_LUL) ? D[(size_t)i + (size_t)q * b] : D[(size_t)q + (size_t)i * b];
            s -= m * X[oq];
        }
        X[oi] = (kind == TR_LUL) ? s : s / D[(size_t)i + (size_t)i * b];
    }
}

// ---------------------------------------------------------------- launchers
static bool g_attr_set = false;

void panel_init()
{
    if (g_attr_set) return;
    const int lim = 150 * 1024;   // largest dynamic request of these kernels; static LDS <= 2 KB
    hipFuncSetAttribute((const void *)potrf_kernel<16>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    hipFuncSetAttribute((const void *)potrf_kernel<8>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    hipFuncSetAttribute((const void *)getrf_kernel<16>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    hipFuncSetAttribute((const void *)getrf_kernel<8>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    hipFuncSetAttribute((const void *)trsm_kernel<TR_CHOL>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    hipFuncSetAttribute((const void *)trsm_kernel<TR_LUL>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    hipFuncSetAttribute((const void *)trsm_kernel<TR_LUR>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    g_attr_set = true;
}

hipError_t launch_potrf(double *tile, int b, int *info, int info_base, hipStream_t s)
{
    panel_init();
    if (b % 16 == 0 && b * 16 * 8 <= 150 * 1024) {
        hipLaunchKernelGGL(potrf_kernel<16>, dim3(1), dim3(PANEL_THREADS), (size_t)16 * b * 8, s, tile, b,
                           info, info_base);
    } else if (b % 8 == 0 && b * 8 * 8 <= 150 * 1024) {
        hipLaunchKernelGGL(potrf_kernel<8>, dim3(1), dim3(PANEL_THREADS), (size_t)8 * b * 8, s, tile, b,
                           info, info_base);
    } else if (b <= 256) {
        hipLaunchKernelGGL(potrf_tiny_kernel, dim3(1), dim3(1), 0, s, tile, b, info, info_base);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_getrf(double *tile, int b, hipStream_t s)
{
    panel_init();
    if (b % 16 == 0 && 2 * b * 16 * 8 <= 150 * 1024) {
        hipLaunchKernelGGL(getrf_kernel<16>, dim3(1), dim3(PANEL_THREADS), (size_t)2 * 16 * b * 8, s, tile, b);
    } else if (b % 8 == 0 && 2 * b * 8 * 8 <= 150 * 1024) {
        hipLaunchKernelGGL(getrf_kernel<8>, dim3(1), dim3(PANEL_THREADS), (size_t)2 * 8 * b * 8, s, tile, b);
    } else if (b <= 256) {
        hipLaunchKernelGGL(getrf_tiny_kernel, dim3(1), dim3(1), 0, s, tile, b);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_trsm(int kind, const TrsmArgs &a, hipStream_t s)
{
    panel_init();
    if (a.ntiles <= 0) return hipSuccess;
    if (a.b % 16 != 0 && a.b <= 256) {
        hipLaunchKernelGGL(trsm_tiny_kernel, dim3((unsigned)a.ntiles), dim3((unsigned)a.b), 0, s, kind, a);
        return hipGetLastError();
    }
    if (a.b % 16 != 0 || (size_t)a.b * TRSM_SW * 8 > 150 * 1024) return hipErrorInvalidValue;
    TrsmArgs t = a;
    trace_trsm(kind, t);
    dim3 grid((unsigned)(t.ntiles * (t.b / TRSM_SW))), block(TRSM_THREADS);
    const size_t lds = (size_t)a.b * TRSM_SW * 8;
    if (kind == TR_CHOL) hipLaunchKernelGGL(trsm_kernel<TR_CHOL>, grid, block, lds, s, t);
    else if (kind == TR_LUL) hipLaunchKernelGGL(trsm_kernel<TR_LUL>, grid, block, lds, s, t);
    else hipLaunchKernelGGL(trsm_kernel<TR_LUR>, grid, block, lds, s, t);
    return hipGetLastError();
}

}  // namespace palial
