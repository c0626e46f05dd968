// k_qr.hip -- Householder panel kernels of the flat-tree tiled QR.
//
//   geqrt_kernel : dgeqr2 + dlarft('F','C') of one b x b tile
//                  (PALIAL_dgeqrt, src/srcqr/dgeqrt.h:18-34, ib = b): R in the
//                  upper triangle, V (unit lower, implicit 1) below, tau, and
//                  the upper-triangular T with H_1..H_b = I - V T V^T.
//   tsqrt_kernel : QR of [R; A2] with R upper (PALIAL_dtsqrt,
//                  src/srcqr/dtsqrt.h:21-86): reflector i is (e_i; v_i) with
//                  v_i stored in column i of A2; T built column by column.
//
// One workgroup walks the columns; for each column the Householder norm and
// the reflector are computed in FP64, then the reflector is applied to the
// remaining columns one wave per column (lanes split the rows, wave
// reduction by shuffles) while the other waves build the T column.
#include "internal.hpp"
#include <float.h>

namespace palial {

constexpr int QR_THREADS = 512;
constexpr int QR_WAVES = QR_THREADS / 64;

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ inline double block_sum(double v, double *red)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < QR_WAVES; q++) s += red[q];
    return s;
}

// LAPACK dlapy2 without overflow: sqrt(x^2 + y^2)
__device__ inline double lapy2(double x, double y)
{
    const double xa = fabs(x), ya = fabs(y);
    const double w = fmax(xa, ya), z = fmin(xa, ya);
    if (z == 0.0) return w;
    return w * sqrt(1.0 + (z / w) * (z / w));
}

// Householder reflector from alpha and ||x||^2 (LAPACK dlarfg): returns
// tau, writes beta (new alpha) and the scale 1/(alpha - beta) for x.
__device__ inline void householder(double alpha, double xnorm2, double &tau, double &beta, double &scal)
{
    if (xnorm2 == 0.0) { tau = 0.0; beta = alpha; scal = 1.0; return; }
    const double xnorm = sqrt(xnorm2);
    beta = -copysign(lapy2(alpha, xnorm), alpha);
    tau = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
}

__global__ __launch_bounds__(QR_THREADS) void geqrt_kernel(double *A, double *T, double *tau_out, int b)
{
    extern __shared__ double sh[];
    double *v = sh;            // v[r - i] for rows r >= i (v[0] = 1)
    double *tcol = sh + b;     // T column under construction
    __shared__ double red[QR_WAVES];
    __shared__ double s_tau, s_scal;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;

    for (int i = 0; i < b; i++) {
        double part = 0.0;
        for (int r = i + 1 + t; r < b; r += QR_THREADS) {
            const double x = A[(size_t)r + (size_t)i * b];
            part += x * x;
        }
        const double xn2 = block_sum(part, red);
        if (t == 0) {
            double tau, beta, scal;
            householder(A[(size_t)i + (size_t)i * b], xn2, tau, beta, scal);
            A[(size_t)i + (size_t)i * b] = beta;
            s_tau = tau;
            s_scal = scal;
            tau_out[i] = tau;
        }
        __syncthreads();
        const double tau = s_tau, scal = s_scal;
        for (int r = i + t; r < b; r += QR_THREADS) {
            double x = 1.0;
            if (r > i) {
                x = A[(size_t)r + (size_t)i * b];
                if (tau != 0.0) { x *= scal; A[(size_t)r + (size_t)i * b] = x; }
            }
            v[r - i] = x;
        }
        __syncthreads();
        // apply H_i to columns j > i, and T(0:i, i) = -tau V(i:b, 0:i)^T v
        for (int job = w; job < b - 1; job += QR_WAVES) {
            if (job < i) {
                const int j = job;   // T gemv for column j < i (V(i,j) stored below diag)
                double s = 0.0;
                for (int r = i + lane; r < b; r += 64) s += A[(size_t)r + (size_t)j * b] * v[r - i];
                s = wave_sum(s);
                if (lane == 0) tcol[j] = -tau * s;
            } else {
                const int j = job + 1;  // columns i+1 .. b-1
                if (tau == 0.0) continue;
                double s = 0.0;
                for (int r = i + lane; r < b; r += 64) s += A[(size_t)r + (size_t)j * b] * v[r - i];
                s = wave_sum(s);
                const double f = tau * s;
                for (int r = i + lane; r < b; r += 64) A[(size_t)r + (size_t)j * b] -= f * v[r - i];
            }
        }
        __syncthreads();
        // T(0:i, i) = T(0:i, 0:i) tcol ; T(i, i) = tau
        for (int j = t; j <= i; j += QR_THREADS) {
            double s;
            if (tau == 0.0) s = 0.0;
            else if (j == i) s = tau;
            else {
                s = 0.0;
                for (int p = j; p < i; p++) s += T[(size_t)j + (size_t)p * b] * tcol[p];
            }
            T[(size_t)j + (size_t)i * b] = s;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(QR_THREADS) void tsqrt_kernel(double *A1, double *A2, double *T,
                                                           double *tau_out, int b)
{
    extern __shared__ double sh[];
    double *tcol = sh;
    __shared__ double red[QR_WAVES];
    __shared__ double s_tau;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;

    for (int i = 0; i < b; i++) {
        double part = 0.0;
        for (int r = t; r < b; r += QR_THREADS) {
            const double x = A2[(size_t)r + (size_t)i * b];
            part += x * x;
        }
        const double xn2 = block_sum(part, red);
        __shared__ double s_scal;
        if (t == 0) {
            double tau, beta, scal;
            householder(A1[(size_t)i + (size_t)i * b], xn2, tau, beta, scal);
            A1[(size_t)i + (size_t)i * b] = beta;
            s_tau = tau;
            s_scal = scal;
            tau_out[i] = tau;
        }
        __syncthreads();
        const double tau = s_tau;
        if (tau != 0.0) {
            const double scal = s_scal;
            for (int r = t; r < b; r += QR_THREADS) A2[(size_t)r + (size_t)i * b] *= scal;
        }
        __syncthreads();
        const double *vi = A2 + (size_t)i * b;
        for (int job = w; job < b - 1; job += QR_WAVES) {
            if (job < i) {
                const int j = job;
                double s = 0.0;
                for (int r = lane; r < b; r += 64) s += A2[(size_t)r + (size_t)j * b] * vi[r];
                s = wave_sum(s);
                if (lane == 0) tcol[j] = -tau * s;
            } else {
                const int j = job + 1;
                if (tau == 0.0) continue;
                double s = 0.0;
                for (int r = lane; r < b; r += 64) s += A2[(size_t)r + (size_t)j * b] * vi[r];
                s = wave_sum(s) + A1[(size_t)i + (size_t)j * b];
                const double f = tau * s;
                if (lane == 0) A1[(size_t)i + (size_t)j * b] -= f;
                for (int r = lane; r < b; r += 64) A2[(size_t)r + (size_t)j * b] -= f * vi[r];
            }
        }
        __syncthreads();
        for (int j = t; j <= i; j += QR_THREADS) {
            double s;
            if (j == i) s = tau;
            else if (tau == 0.0) s = 0.0;
            else {
                s = 0.0;
                for (int p = j; p < i; p++) s += T[(size_t)j + (size_t)p * b] * tcol[p];
            }
            T[(size_t)j + (size_t)i * b] = s;
        }
        __syncthreads();
    }
}

hipError_t launch_geqrt(double *A, double *T, double *tau, int b, hipStream_t s)
{
    hipLaunchKernelGGL(geqrt_kernel, dim3(1), dim3(QR_THREADS), (size_t)2 * b * 8, s, A, T, tau, b);
    return hipGetLastError();
}

hipError_t launch_tsqrt(double *A1, double *A2, double *T, double *tau, int b, hipStream_t s)
{
    hipLaunchKernelGGL(tsqrt_kernel, dim3(1), dim3(QR_THREADS), (size_t)b * 8, s, A1, A2, T, tau, b);
    return hipGetLastError();
}

}  // namespace palial
