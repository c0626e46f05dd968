// k_misc.hip -- input generators and small helpers.
//
// The generators write layout L directly, walking the storage linearly so
// that stores are coalesced; entry (i,j) of the global matrix depends only on
// (seed, i, j), so host and device produce bit-identical matrices for any
// tile size.
#include "internal.hpp"

// Built with -ffp-contract=off (see palial_amd/build.py) so that every
// multiply and add rounds separately, as in a plain C evaluation.

namespace palial {

__device__ inline uint64_t sm64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ inline double uniform(uint64_t seed, uint64_t i, uint64_t j)
{
    const uint64_t x = sm64(sm64(seed) + ((i << 32) | (j & 0xffffffffull)));
    return (double)(x >> 11) * 0x1.0p-53 - 0.5;
}

__global__ void generate_kernel(double *mat, int N, int b, int kind, uint64_t seed, double shift)
{
    const size_t total = (size_t)N * N;
    const size_t bb = (size_t)b * b;
    const int MT = N / b;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const size_t tile = e / bb, within = e % bb;
        const int r = (int)(within % b), c = (int)(within / b);
        const int m = (int)(tile % MT), n = (int)(tile / MT);
        const uint64_t i = (uint64_t)m * b + r, j = (uint64_t)n * b + c;
        double v;
        switch (kind) {
        case 0:
            v = (i <= j) ? uniform(seed, i, j) : uniform(seed, j, i);
            if (i == j) v = v + shift;
            break;
        case 1:
            v = uniform(seed, i, j);
            if (i == j) v = v + shift;
            break;
        case 2:
            v = 0.5;
            if (i == j) v = (uniform(seed, i, i) + 0.5) * 173.0 + 173.0;
            break;
        default:
            v = (i == j) ? 1.0 : 0.0;
        }
        mat[e] = v;
    }
}

// Batch inputs: matrix i (column-major, order n[i]) from seed + i; kind 0:
// S + n I with S_rc = U(seed+i, min(r,c), max(r,c)); kind 1: G + n I with
// G_rc = U(seed+i, r, c).  One workgroup per matrix.
__global__ void batch_generate_kernel(const int *n, const long *off, const int *kind, void *A, int f32,
                                      uint64_t seed)
{
    const int i = blockIdx.x;
    const int m = n[i];
    const uint64_t sd = seed + (uint64_t)i;
    const size_t o = (size_t)off[i];
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
        const int r = e % m, c = e / m;
        double v = (kind[i] == 0) ? ((r <= c) ? uniform(sd, r, c) : uniform(sd, c, r)) : uniform(sd, r, c);
        if (r == c) v += (double)m;
        if (f32) ((float *)A)[o + e] = (float)v;
        else ((double *)A)[o + e] = v;
    }
}

hipError_t launch_batch_generate(const int *n, const long *off, const int *kind, void *A, int count, int f32,
                                 uint64_t seed, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_generate_kernel, dim3((unsigned)count), dim3(256), 0, s, n, off, kind, A, f32, seed);
    return hipGetLastError();
}

hipError_t launch_generate(double *mat, int N, int b, int kind, uint64_t seed, double shift, hipStream_t s)
{
    const size_t total = (size_t)N * N;
    size_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(generate_kernel, dim3((unsigned)blocks), dim3(256), 0, s, mat, N, b, kind, seed, shift);
    return hipGetLastError();
}

// Plain streaming copy (16-byte vectors): the reference point for the HBM
// byte counters (its algorithmic traffic is exactly 2 x bytes).
__global__ void copy_kernel(double2 *__restrict__ dst, const double2 *__restrict__ src, size_t n2)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    const size_t n2 = bytes / 16;
    hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(256), 0, s, (double2 *)dst, (const double2 *)src, n2);
    return hipGetLastError();
}

// A(i,j) = A1(i / n2, j / n2) * A2(i % n2, j % n2): the Kronecker product of
// two small dense column-major matrices, written in layout L.
__global__ void kron_kernel(double *mat, int N, int b, const double *A1, int n1, const double *A2, int n2)
{
    const size_t total = (size_t)N * N;
    const size_t bb = (size_t)b * b;
    const int MT = N / b;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t tile = e / bb, within = e % bb;
        const int r = (int)(within % b), c = (int)(within / b);
        const int m = (int)(tile % MT), n = (int)(tile / MT);
        const int i = m * b + r, j = n * b + c;
        mat[e] = A1[(size_t)(i / n2) + (size_t)(j / n2) * n1] * A2[(size_t)(i % n2) + (size_t)(j % n2) * n2];
    }
}

hipError_t launch_kron(double *mat, int N, int b, const double *A1, int n1, const double *A2, int n2, hipStream_t s)
{
    hipLaunchKernelGGL(kron_kernel, dim3(65536), dim3(256), 0, s, mat, N, b, A1, n1, A2, n2);
    return hipGetLastError();
}

// Copy an N x N matrix from layout L with tile size b1 to layout L with tile
// size b2 (the destination is walked linearly).
__global__ void relayout_kernel(double *__restrict__ dst, int b2, const double *__restrict__ src, int b1, int N)
{
    const size_t total = (size_t)N * N;
    const size_t bb2 = (size_t)b2 * b2;
    const int MT1 = N / b1, MT2 = N / b2;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t tile = e / bb2, within = e % bb2;
        const int i = (int)(tile % MT2) * b2 + (int)(within % b2);
        const int j = (int)(tile / MT2) * b2 + (int)(within / b2);
        dst[e] = src[tile_off(b1, MT1, i / b1, j / b1) + (size_t)(i % b1) + (size_t)(j % b1) * b1];
    }
}

hipError_t launch_relayout(double *dst, int b2, const double *src, int b1, int N, hipStream_t s)
{
    hipLaunchKernelGGL(relayout_kernel, dim3(65536), dim3(256), 0, s, dst, b2, src, b1, N);
    return hipGetLastError();
}

// One b x b column-major tile <-> its (b/bs) x (b/bs) sub-tiles in layout L
// (sub-tile (i,j) at bs*bs*(i + (b/bs)*j), column-major inside).  A thread
// moves runs of 4 rows of one column (b and bs are multiples of 4, so a run
// stays inside one sub-tile), 4 runs in flight; loads are clamped into the
// tile so that only the stores are guarded.  The grid is small: the diagonal
// chain that calls this may own only a few reserved CUs.
template <typename TT> __device__ inline void ld4(const TT *p, double (&v)[4]);
template <> __device__ inline void ld4<float>(const float *p, double (&v)[4])
{
    const float4 f = *(const float4 *)p;
    v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w;
}
template <> __device__ inline void ld4<double>(const double *p, double (&v)[4])
{
    const double2 a = *(const double2 *)p, c = *(const double2 *)(p + 2);
    v[0] = a.x; v[1] = a.y; v[2] = c.x; v[3] = c.y;
}
template <typename TT> __device__ inline void st4(TT *p, const double (&v)[4]);
template <> __device__ inline void st4<float>(float *p, const double (&v)[4])
{
    *(float4 *)p = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
}
template <> __device__ inline void st4<double>(double *p, const double (&v)[4])
{
    *(double2 *)p = make_double2(v[0], v[1]);
    *(double2 *)(p + 2) = make_double2(v[2], v[3]);
}

template <typename TT>
__global__ __launch_bounds__(256) void tile_split_kernel(double *__restrict__ sub, TT *__restrict__ tile, int b, int bs,
                                                         int to_sub)
{
    const int bq = b / 4, nq = bq * b, ms = b / bs;
    const int stride = gridDim.x * blockDim.x;
    for (int q0 = blockIdx.x * blockDim.x + threadIdx.x; q0 < nq; q0 += 4 * stride) {
        double v[4][4];
        size_t so[4], to[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int q = min(q0 + u * stride, nq - 1);
            const int r = 4 * (q % bq), c = q / bq;
            to[u] = (size_t)r + (size_t)c * b;
            so[u] = tile_off(bs, ms, r / bs, c / bs) + (size_t)(r % bs) + (size_t)(c % bs) * bs;
            if (to_sub) ld4(tile + to[u], v[u]);
            else ld4(sub + so[u], v[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (q0 + u * stride >= nq) continue;
            if (to_sub) st4(sub + so[u], v[u]);
            else st4(tile + to[u], v[u]);
        }
    }
}

hipError_t launch_tile_split(double *sub, double *tile, int b, int bs, bool to_sub, hipStream_t s, bool tile32)
{
    if (b <= 0 || bs <= 0 || b % bs != 0 || b % 4 != 0 || bs % 4 != 0) return hipErrorInvalidValue;
    // 4 runs of 4 elements per thread and pass, at most 64 workgroups
    const long nq = (long)b * b / 4;
    int blocks = (int)((nq + 1023) / 1024);
    if (blocks > 64) blocks = 64;
    if (tile32)
        hipLaunchKernelGGL(tile_split_kernel<float>, dim3(blocks), dim3(256), 0, s, sub, (float *)(void *)tile, b, bs,
                           to_sub ? 1 : 0);
    else
        hipLaunchKernelGGL(tile_split_kernel<double>, dim3(blocks), dim3(256), 0, s, sub, tile, b, bs, to_sub ? 1 : 0);
    return hipGetLastError();
}

__global__ void iota_ipiv_kernel(int *ipiv, int N)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) ipiv[i] = i + 1;
}

hipError_t launch_fill_identity_ipiv(int *ipiv, int N, hipStream_t s)
{
    int blocks = (N + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(iota_ipiv_kernel, dim3(blocks), dim3(256), 0, s, ipiv, N);
    return hipGetLastError();
}

}  // namespace palial
