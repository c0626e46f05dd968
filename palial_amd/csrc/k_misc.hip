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

// Standard normal entry (i,j) by the Box-Muller transform of two uniforms
// drawn from the same counter: h = splitmix64(splitmix64(seed) + (i<<32 | j)),
// u1 = ((h >> 11) + 1) 2^-53 in (0, 1], u2 = (splitmix64(h) >> 11) 2^-53,
// z = sqrt(-2 ln u1) cos(2 pi u2).  Host code with another libm agrees to a
// few ulp of log and cos, not bit for bit.
__device__ inline double gaussian(uint64_t seed, uint64_t i, uint64_t j)
{
    const uint64_t h = sm64(sm64(seed) + ((i << 32) | (j & 0xffffffffull)));
    const double u1 = (double)((h >> 11) + 1) * 0x1.0p-53;
    const double u2 = (double)(sm64(h) >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
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
        case 4:
            v = gaussian(seed, i, j);
            break;
        case 5:   // diag(s), s_i = kappa^(-i/(N-1)) from 1 down to 1/kappa (kappa = shift)
            v = (i != j) ? 0.0 : (N > 1 ? pow(shift, -(double)i / (double)(N - 1)) : 1.0);
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

// In-place transpose (avg = 0: A <- A^T) or symmetrization (avg = 1:
// A <- (A + A^T) / 2) of a layout-L matrix.  Workgroup w takes the pair of
// 32x32 blocks (bi, bj) / (bj, bi), bi <= bj, of the global matrix: both are
// read through LDS (32 rows of one column are 256 contiguous bytes of a tile
// column), combined and written back; the diagonal blocks pair with
// themselves.  b % 32 == 0.
__global__ __launch_bounds__(256) void transpose_kernel(double *mat, int N, int b, int avg, long npairs)
{
    __shared__ double s1[32][33], s2[32][33];
    const int MT = N / b, NBK = N / 32;
    const int t = threadIdx.x;
    for (long w = blockIdx.x; w < npairs; w += gridDim.x) {
        // w -> (bi, bj), bi <= bj: w = bj (bj + 1) / 2 + bi
        int bj = (int)((sqrt(8.0 * (double)w + 1.0) - 1.0) * 0.5);
        while ((long)(bj + 1) * (bj + 2) / 2 <= w) bj++;
        while ((long)bj * (bj + 1) / 2 > w) bj--;
        const int bi = (int)(w - (long)bj * (bj + 1) / 2);
        if (bj >= NBK) break;
        auto addr = [&](int gi, int gj) -> size_t {   // global element (gi, gj) in layout L
            return tile_off(b, MT, gi / b, gj / b) + (size_t)(gi % b) + (size_t)(gj % b) * b;
        };
        for (int e = t; e < 1024; e += 256) {
            const int r = e % 32, c = e / 32;
            s1[r][c] = mat[addr(bi * 32 + r, bj * 32 + c)];
            s2[r][c] = mat[addr(bj * 32 + r, bi * 32 + c)];
        }
        __syncthreads();
        for (int e = t; e < 1024; e += 256) {
            const int r = e % 32, c = e / 32;
            // new (bi,bj)(r,c) from old (bj,bi)(c,r) and vice versa
            const double x = s1[r][c], y = s2[c][r];
            const double u = avg ? 0.5 * (x + y) : y;
            const double v = avg ? u : x;
            mat[addr(bi * 32 + r, bj * 32 + c)] = u;
            if (bi != bj) mat[addr(bj * 32 + c, bi * 32 + r)] = v;
        }
        __syncthreads();
    }
}

hipError_t launch_transpose(double *mat, int N, int b, bool avg, hipStream_t s)
{
    if (b % 32 || N % b) return hipErrorInvalidValue;
    const long nbk = N / 32, npairs = nbk * (nbk + 1) / 2;
    const long blocks = npairs < 65536 ? npairs : 65536;
    hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)blocks), dim3(256), 0, s, mat, N, b, avg ? 1 : 0, npairs);
    return hipGetLastError();
}

__global__ void cols_to_tiles_kernel(const double *v, int N, int ld, int nrhs, double *tiles, int b)
{
    // tiles: N x b in layout L with one tile column, i.e. tile m at m*b*b
    const size_t total = (size_t)N * b;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t m = e / ((size_t)b * b), within = e % ((size_t)b * b);
        const int r = (int)(within % b), c = (int)(within / b);
        const size_t i = m * b + r;
        tiles[e] = (c < nrhs) ? v[i + (size_t)c * ld] : 0.0;
    }
}

__global__ void tiles_to_cols_kernel(const double *tiles, int N, int b, double *v, int ld, int nrhs)
{
    const size_t total = (size_t)N * nrhs;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t i = e % N, c = e / N;
        v[i + c * ld] = tiles[(i / b) * b * b + (i % b) + c * b];
    }
}

hipError_t launch_cols_to_tiles(const double *v, int N, int ld, int nrhs, double *tiles, int b, hipStream_t s)
{
    if (nrhs < 1 || nrhs > b || ld < N || N % b) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cols_to_tiles_kernel, dim3(2048), dim3(256), 0, s, v, N, ld, nrhs, tiles, b);
    return hipGetLastError();
}

hipError_t launch_tiles_to_cols(const double *tiles, int N, int b, double *v, int ld, int nrhs, hipStream_t s)
{
    if (nrhs < 1 || nrhs > b || ld < N || N % b) return hipErrorInvalidValue;
    hipLaunchKernelGGL(tiles_to_cols_kernel, dim3(2048), dim3(256), 0, s, tiles, N, b, v, ld, nrhs);
    return hipGetLastError();
}

// task-trace slots: starts ~0 (atomicMin targets), ends 0 (atomicMax targets)
__global__ void trace_init_kernel(unsigned long long *ts, size_t nslots)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += (size_t)gridDim.x * blockDim.x) {
        ts[2 * i] = ~0ull;
        ts[2 * i + 1] = 0ull;
    }
}

__global__ void clock_stamp_kernel(unsigned long long *dst)
{
    if (threadIdx.x == 0) *dst = (unsigned long long)wall_clock64();
}

hipError_t launch_trace_init(unsigned long long *ts, size_t nslots, hipStream_t s)
{
    hipLaunchKernelGGL(trace_init_kernel, dim3(1024), dim3(256), 0, s, ts, nslots);
    return hipGetLastError();
}

hipError_t launch_clock_stamp(unsigned long long *dst, hipStream_t s)
{
    hipLaunchKernelGGL(clock_stamp_kernel, dim3(1), dim3(64), 0, s, dst);
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
