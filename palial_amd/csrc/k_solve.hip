// k_solve.hip -- kernels of the solve / iterative-refinement path.
//
//   tile_gemv      : batched out -= op(T) v (or out = op(T) v) over a list
//                    of b x b tiles of layout L, nrhs <= 16 right-hand
//                    sides; the two launches of one sweep step.
//   residual       : r = b - A x over the FP64 layout-L matrix (HBM-bound),
//                    partial sums in fixed slots + a deterministic reduction;
//                    optionally the row sums of |A| (||A||_inf) in the same
//                    pass; symmetric matrices read only their upper tiles.
//   colmax_abs     : per-column max |v| of an N x nrhs block.
//   axpy_cols      : x += d.
#include "internal.hpp"
#include "gemm_common.hpp"

namespace palial {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------- tile GEMV
// For each task t: out_t(0:b, 0:nrhs) = sum_j op(T_tj) v_j  (GV_DIAG) or
// out_t -= sum_j op(T_tj) v_j (updates), over the source tiles j of the step
// (gemv_map, gemv_src).  A workgroup (256 threads) owns ROWS output rows of
// one tile; the input segment (the step's q tiles x nrhs) sits in LDS.
//   op = N: out(r) = sum_c T(r, c) v(c): ROWS row lanes x 256/ROWS column
//           groups; ROWS = 16 for FP64 and 32 for FP32 tiles, so that a
//           column piece is one full 128-byte line
//   op = T: out(r) = sum_c T(c, r) v(c): a wave owns 4 rows; its 64 lanes
//           read 64 consecutive c of a row (512 bytes) and reduce by shuffles
constexpr int GV_THREADS = 256;
constexpr int GV_ROWS = 16;          // output rows per workgroup
constexpr int GV_MAXR = 16;          // right-hand sides

// E: element type of the tiles (double, or float for the FP32 copies the
// mixed-precision correction solves stream); sums are accumulated in FP64.
// One-tile steps (q = 1, the default): the source tile is found once, before
// the loops.
template <int NR, typename E, int ROWS>
__device__ inline void tile_gemv_one(const GemvArgs &p)
{
    constexpr int G = GV_THREADS / ROWS;      // column groups (op = N)
    extern __shared__ double sm[];
    double *sv = sm;                          // b x NR input segment (layout by op, below)
    double *sred = sm + (size_t)p.b * NR;     // G groups x ROWS rows x NR partial sums
    const int b = p.b;
    const int rb_per_tile = b / ROWS;
    const int task = blockIdx.x / rb_per_tile, rb = blockIdx.x % rb_per_tile;
    const int t = threadIdx.x;
    const E *mat = (const E *)p.mat;
    const E *inv = (const E *)p.inv;
    const E *T;
    int vin, vout;
    {
        const int MT = p.N / b, k = p.k;
        switch (p.mode) {
        case GV_DIAG: T = inv + (size_t)k * b * b; vin = k; vout = k; break;
        case GV_CHOL_FWD: T = mat + tile_off(b, MT, k, k + 1 + task); vin = k; vout = k + 1 + task; break;
        case GV_LU_FWD: T = mat + tile_off(b, MT, k + 1 + task, k); vin = k; vout = k + 1 + task; break;
        default: T = mat + tile_off(b, MT, task, k); vin = k; vout = task; break;
        }
    }
    const double *vi = p.vin + (size_t)vin * b;
    double *vo = p.vout + (size_t)vout * b;
    // op = N reads the NR values of one row together: sv[i * NR + c]; op = T
    // reads one right-hand side across the lanes: sv[i + c * b]
    for (int e = t; e < b * NR; e += GV_THREADS) {
        const int i = e % b, c = e / b;
        const double x = (c < p.nrhs) ? vi[(size_t)i + (size_t)c * p.ld] : 0.0;   // padded right-hand sides are zero
        sv[p.trans ? e : i * NR + c] = x;
    }
    __syncthreads();
    const int r0 = rb * ROWS;
    double acc[NR];
#pragma unroll
    for (int j = 0; j < NR; j++) acc[j] = 0.0;
    if (!p.trans) {
        const int r = t % ROWS, g = t / ROWS;             // g: column group, columns c = g, g+G, ...
        const E *Tr = T + r0 + r;
#pragma unroll 8
        for (int c = g; c < b; c += G) {
            const double a = (double)Tr[(size_t)c * b];
            const double *svc = sv + c * NR;
#pragma unroll
            for (int j = 0; j < NR; j++) acc[j] += a * svc[j];
        }
#pragma unroll
        for (int j = 0; j < NR; j++) sred[(g * ROWS + r) * NR + j] = acc[j];
        __syncthreads();
        for (int e = t; e < ROWS * p.nrhs; e += GV_THREADS) {   // ROWS * NR may exceed the threads
            const int rr = e % ROWS, j = e / ROWS;
            double s = 0.0;
            for (int q = 0; q < G; q++) s += sred[(q * ROWS + rr) * NR + j];
            double *o = vo + (size_t)(r0 + rr) + (size_t)j * p.ld;
            if (p.mode == GV_DIAG) *o = s;
            else *o -= s;
        }
    } else if constexpr (ROWS == 16) {           // op = T: 4 waves x 4 rows (launched with ROWS = 16 only)
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
            for (int x = 0; x < 4; x++) a[x] = (double)T[(size_t)c + (size_t)(r0 + wv * 4 + x) * b];
#pragma unroll
            for (int x = 0; x < 4; x++)
#pragma unroll
                for (int j = 0; j < NR; j++) part[x][j] += a[x] * sv[c + j * b];
        }
        // Sum the 4 NR values over the 64 lanes.  Butterfly steps on the high
        // lane bits first (every lane keeps all values), then halving steps:
        // at offset o a lane keeps one half of its values (the upper half when
        // its bit o is set), sends the other half to lane ^ o and adds what it
        // receives, so that in the end lane L holds the sum of value L mod 4 NR.
        constexpr int NV = 4 * NR;
        constexpr int NH = NV == 64 ? 6 : NV == 16 ? 4 : 2;    // halving steps: log2(NV)
        double v[NV];
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
            for (int j = 0; j < NR; j++) v[x * NR + j] = part[x][j];
#pragma unroll
        for (int o = 32; o > (1 << (NH - 1)); o >>= 1)
#pragma unroll
            for (int i = 0; i < NV; i++) v[i] += __shfl_xor(v[i], o);
        int n = NV;
#pragma unroll
        for (int o = 1 << (NH - 1); o >= 1; o >>= 1) {
            const bool up = lane & o;
            n >>= 1;
#pragma unroll
            for (int i = 0; i < NV / 2; i++) {
                if (i < n) {
                    const double send = up ? v[i] : v[i + n];
                    const double keep = up ? v[i + n] : v[i];
                    v[i] = keep + __shfl_xor(send, o);
                }
            }
        }
        if (lane < NV && lane % NR < p.nrhs) {
            const int x = lane / NR, j = lane % NR;
            const double s = v[0];
            double *o = vo + (size_t)(r0 + wv * 4 + x) + (size_t)j * p.ld;
            if (p.mode == GV_DIAG) *o = s;
            else *o -= s;
        }
    }
}

// QB: steps of p.q tiles (blocked sweeps), the sum running over the step's
// source tiles (gemv_map, gemv_src); without it, tile_gemv_one.
template <int NR, typename E, int ROWS, bool QB>
__global__ __launch_bounds__(GV_THREADS) void tile_gemv_kernel(GemvArgs p)
{
    if constexpr (!QB) {
        tile_gemv_one<NR, E, ROWS>(p);
        return;
    } else {
        constexpr int G = GV_THREADS / ROWS;      // column groups (op = N)
        extern __shared__ double sm[];
        const int b = p.b;
        const int q = QB && p.q > 1 ? p.q : 1;
        const int seg = q * b;                    // input segment: the q tiles of the step
        double *sv = sm;                          // seg x NR input segment (layout by op, below)
        double *sred = sm + (size_t)seg * NR;     // G groups x ROWS rows x NR partial sums
        const int rb_per_tile = b / ROWS;
        const int task = blockIdx.x / rb_per_tile, rb = blockIdx.x % rb_per_tile;
        const int t = threadIdx.x;
        const E *mat = (const E *)p.mat;
        const E *inv = (const E *)p.inv;
        const int MT = p.N / b, s0 = p.k * q;     // first tile of the step
        // output tile and the source tiles j0 <= j < j1 of the step (T_j below)
        const GemvMap mp = gemv_map(p.mode, p.trans, p.k, q, task);
        // one-tile steps read exactly source tile 0 (a diagonal step has one task)
        const int vout = mp.vout, j0 = QB ? mp.j0 : 0, j1 = QB ? mp.j1 : 1;
        auto tileT = [&](int j) -> const E * {
            return (p.mode == GV_DIAG ? inv : mat) + gemv_src(p.mode, p.trans, p.k, q, b, MT, vout, j);
        };
        const E *const T0 = QB ? nullptr : tileT(0);    // one-tile steps: the source tile, once
        const double *vi = p.vin + (size_t)s0 * b;
        double *vo = p.vout + (size_t)vout * b;
        // op = N reads the NR values of one row together: sv[i * NR + c]; op = T
        // reads one right-hand side across the lanes: sv[i + c * seg]
        for (int e = t; e < seg * NR; e += GV_THREADS) {
            const int i = e % seg, c = e / seg;
            const double x = (c < p.nrhs) ? vi[(size_t)i + (size_t)c * p.ld] : 0.0;   // padded right-hand sides are zero
            sv[p.trans ? e : i * NR + c] = x;
        }
        __syncthreads();
        const int r0 = rb * ROWS;
        double acc[NR];
    #pragma unroll
        for (int j = 0; j < NR; j++) acc[j] = 0.0;
        if (!p.trans) {
            const int r = t % ROWS, g = t / ROWS;             // g: column group, columns c = g, g+G, ...
            for (int jt = j0; jt < j1; jt++) {
                const E *Tr = (QB ? tileT(jt) : T0) + r0 + r;
                const double *svt = sv + (size_t)jt * b * NR;
    #pragma unroll 8
                for (int c = g; c < b; c += G) {
                    const double a = (double)Tr[(size_t)c * b];
                    const double *svc = svt + c * NR;
    #pragma unroll
                    for (int j = 0; j < NR; j++) acc[j] += a * svc[j];
                }
            }
    #pragma unroll
            for (int j = 0; j < NR; j++) sred[(g * ROWS + r) * NR + j] = acc[j];
            __syncthreads();
            for (int e = t; e < ROWS * p.nrhs; e += GV_THREADS) {   // ROWS * NR may exceed the threads
                const int rr = e % ROWS, j = e / ROWS;
                double s = 0.0;
                for (int q = 0; q < G; q++) s += sred[(q * ROWS + rr) * NR + j];
                double *o = vo + (size_t)(r0 + rr) + (size_t)j * p.ld;
                if (p.mode == GV_DIAG) *o = s;
                else *o -= s;
            }
        } else if constexpr (ROWS == 16) {           // op = T: 4 waves x 4 rows (launched with ROWS = 16 only)
            const int lane = t & 63, wv = t >> 6;
            double part[4][NR];
    #pragma unroll
            for (int x = 0; x < 4; x++)
    #pragma unroll
                for (int j = 0; j < NR; j++) part[x][j] = 0.0;
            for (int jt = j0; jt < j1; jt++) {
                const E *T = QB ? tileT(jt) : T0;
                const double *svt = sv + (size_t)jt * b;
                for (int c0 = 0; c0 < b; c0 += 64) {
                    const int c = c0 + lane;
                    double a[4];
    #pragma unroll
                    for (int x = 0; x < 4; x++) a[x] = (double)T[(size_t)c + (size_t)(r0 + wv * 4 + x) * b];
    #pragma unroll
                    for (int x = 0; x < 4; x++)
    #pragma unroll
                        for (int j = 0; j < NR; j++) part[x][j] += a[x] * svt[c + j * seg];
                }
            }
            // Sum the 4 NR values over the 64 lanes.  Butterfly steps on the high
            // lane bits first (every lane keeps all values), then halving steps:
            // at offset o a lane keeps one half of its values (the upper half when
            // its bit o is set), sends the other half to lane ^ o and adds what it
            // receives, so that in the end lane L holds the sum of value L mod 4 NR.
            constexpr int NV = 4 * NR;
            constexpr int NH = NV == 64 ? 6 : NV == 16 ? 4 : 2;    // halving steps: log2(NV)
            double v[NV];
    #pragma unroll
            for (int x = 0; x < 4; x++)
    #pragma unroll
                for (int j = 0; j < NR; j++) v[x * NR + j] = part[x][j];
    #pragma unroll
            for (int o = 32; o > (1 << (NH - 1)); o >>= 1)
    #pragma unroll
                for (int i = 0; i < NV; i++) v[i] += __shfl_xor(v[i], o);
            int n = NV;
    #pragma unroll
            for (int o = 1 << (NH - 1); o >= 1; o >>= 1) {
                const bool up = lane & o;
                n >>= 1;
    #pragma unroll
                for (int i = 0; i < NV / 2; i++) {
                    if (i < n) {
                        const double send = up ? v[i] : v[i + n];
                        const double keep = up ? v[i + n] : v[i];
                        v[i] = keep + __shfl_xor(send, o);
                    }
                }
            }
            if (lane < NV && lane % NR < p.nrhs) {
                const int x = lane / NR, j = lane % NR;
                const double s = v[0];
                double *o = vo + (size_t)(r0 + wv * 4 + x) + (size_t)j * p.ld;
                if (p.mode == GV_DIAG) *o = s;
                else *o -= s;
            }
        }
    }
}

// The same products for FP32 tiles and 5..16 right-hand sides on
// v_mfma_f32_16x16x4f32: FP32 products and sums, the input segment rounded to
// FP32 (the mixed-tier correction solves, whose factor carries BF16/FP32-level
// error anyway).  A wave owns 16 output rows and runs over all b columns:
// lane l supplies row i = l % 16 of op(T) at the four consecutive k =
// k0 + 4 (l / 16) + s of each 16-k group, and the matching rows of the input
// segment from LDS; D (16 rows x 16 right-hand sides) lands at rows
// 4 (l / 16) + v, right-hand side l % 16.
typedef float gf4 __attribute__((ext_vector_type(4)));

template <bool TRANS>
__global__ __launch_bounds__(GV_THREADS) void tile_gemm16_kernel(GemvArgs p)
{
    constexpr int SX = 20;                       // LDS row stride in floats: conflict-free operand reads
    extern __shared__ float sxf[];               // b x SX: sxf[c * SX + j] = v(c, j)
    const int b = p.b;
    const int rb_per_tile = b / 64;
    const int task = blockIdx.x / rb_per_tile, rb = blockIdx.x % rb_per_tile;
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const float *mat = (const float *)p.mat;
    const float *inv = (const float *)p.inv;
    const float *T;
    int vin, vout;
    {
        const int MT = p.N / b, k = p.k;
        switch (p.mode) {
        case GV_DIAG: T = inv + (size_t)k * b * b; vin = k; vout = k; break;
        case GV_CHOL_FWD: T = mat + tile_off(b, MT, k, k + 1 + task); vin = k; vout = k + 1 + task; break;
        case GV_LU_FWD: T = mat + tile_off(b, MT, k + 1 + task, k); vin = k; vout = k + 1 + task; break;
        default: T = mat + tile_off(b, MT, task, k); vin = k; vout = task; break;
        }
    }
    const double *vi = p.vin + (size_t)vin * b;
    double *vo = p.vout + (size_t)vout * b;
    for (int e = t; e < b * 16; e += GV_THREADS) {
        const int c = e % b, j = e / b;
        sxf[c * SX + j] = j < p.nrhs ? (float)vi[(size_t)c + (size_t)j * p.ld] : 0.0f;   // padded columns are zero
    }
    __syncthreads();
    const int r0 = rb * 64 + w * 16, i = l & 15, kk = l >> 4;
    gf4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    // op(T) values of 16-k group g for this lane (four loads in flight per group)
    auto load = [&](int g, float *av) {
        const int kb = 16 * g + 4 * kk;
        if (TRANS) {
            const float4 x = *(const float4 *)(T + (size_t)kb + (size_t)(r0 + i) * b);
            av[0] = x.x; av[1] = x.y; av[2] = x.z; av[3] = x.w;
        } else {
#pragma unroll
            for (int s = 0; s < 4; s++) av[s] = T[(size_t)(r0 + i) + (size_t)(kb + s) * b];
        }
    };
    constexpr int D = 4;                         // groups loaded ahead
    const int ng = b / 16;                       // a multiple of 4 (b % 64 == 0)
    float ab[D][4];
#pragma unroll
    for (int u = 0; u < D; u++) load(u, ab[u]);
    for (int g0 = 0; g0 < ng; g0 += D) {
#pragma unroll
        for (int u = 0; u < D; u++) {
            float av[4];
#pragma unroll
            for (int s = 0; s < 4; s++) av[s] = ab[u][s];
            if (g0 + u + D < ng) load(g0 + u + D, ab[u]);
            const int kb = 16 * (g0 + u) + 4 * kk;
#pragma unroll
            for (int s = 0; s < 4; s++)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s], sxf[(kb + s) * SX + i], acc, 0, 0, 0);
        }
    }
    if (i < p.nrhs) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            double *o = vo + (size_t)(r0 + 4 * kk + v) + (size_t)i * p.ld;
            if (p.mode == GV_DIAG) *o = (double)acc[v];
            else *o -= (double)acc[v];
        }
    }
}

// FP64 tiles with 5..16 right-hand sides on v_mfma_f64_16x16x4f64 (FP64
// products and sums): the same wave layout as tile_gemm16_kernel; D lands at
// rows l / 16 + 4 v, right-hand side l % 16.
template <bool TRANS>
__global__ __launch_bounds__(GV_THREADS) void tile_dgemm16_kernel(GemvArgs p)
{
    constexpr int SX = 20;                       // LDS row stride in doubles
    extern __shared__ double sxd[];              // b x SX: sxd[c * SX + j] = v(c, j)
    const int b = p.b;
    const int rb_per_tile = b / 64;
    const int task = blockIdx.x / rb_per_tile, rb = blockIdx.x % rb_per_tile;
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const double *mat = (const double *)p.mat;
    const double *inv = (const double *)p.inv;
    const double *T;
    int vin, vout;
    {
        const int MT = p.N / b, k = p.k;
        switch (p.mode) {
        case GV_DIAG: T = inv + (size_t)k * b * b; vin = k; vout = k; break;
        case GV_CHOL_FWD: T = mat + tile_off(b, MT, k, k + 1 + task); vin = k; vout = k + 1 + task; break;
        case GV_LU_FWD: T = mat + tile_off(b, MT, k + 1 + task, k); vin = k; vout = k + 1 + task; break;
        default: T = mat + tile_off(b, MT, task, k); vin = k; vout = task; break;
        }
    }
    const double *vi = p.vin + (size_t)vin * b;
    double *vo = p.vout + (size_t)vout * b;
    for (int e = t; e < b * 16; e += GV_THREADS) {
        const int c = e % b, j = e / b;
        sxd[c * SX + j] = j < p.nrhs ? vi[(size_t)c + (size_t)j * p.ld] : 0.0;   // padded columns are zero
    }
    __syncthreads();
    const int r0 = rb * 64 + w * 16, i = l & 15, kk = l >> 4;
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    auto load = [&](int g, double *av) {
        const int kb = 16 * g + 4 * kk;
        if (TRANS) {
            const d4 x = *(const d4 *)(T + (size_t)kb + (size_t)(r0 + i) * b);
            av[0] = x[0]; av[1] = x[1]; av[2] = x[2]; av[3] = x[3];
        } else {
#pragma unroll
            for (int s = 0; s < 4; s++) av[s] = T[(size_t)(r0 + i) + (size_t)(kb + s) * b];
        }
    };
    constexpr int D = 4;                         // groups loaded ahead
    const int ng = b / 16;
    double ab[D][4];
#pragma unroll
    for (int u = 0; u < D; u++) load(u, ab[u]);
    for (int g0 = 0; g0 < ng; g0 += D) {
#pragma unroll
        for (int u = 0; u < D; u++) {
            double av[4];
#pragma unroll
            for (int s = 0; s < 4; s++) av[s] = ab[u][s];
            if (g0 + u + D < ng) load(g0 + u + D, ab[u]);
            const int kb = 16 * (g0 + u) + 4 * kk;
#pragma unroll
            for (int s = 0; s < 4; s++)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], sxd[(kb + s) * SX + i], acc, 0, 0, 0);
        }
    }
    if (i < p.nrhs) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            double *o = vo + (size_t)(r0 + kk + 4 * v) + (size_t)i * p.ld;
            if (p.mode == GV_DIAG) *o = acc[v];
            else *o -= acc[v];
        }
    }
}

hipError_t launch_tile_gemv(const GemvArgs &a, hipStream_t s)
{
    if (a.ntasks <= 0) return hipSuccess;
    if (a.b > 512 || a.b % 64 || a.nrhs > GV_MAXR || a.nrhs < 1) return hipErrorInvalidValue;
    const int q = a.q > 1 ? a.q : 1;
    // blocked steps: 1..4 right-hand sides, the segment within the default LDS limit
    if (q > 1 && (a.nrhs > 4 || (size_t)q * a.b * (a.nrhs == 1 ? 1 : 4) > 4096)) return hipErrorInvalidValue;
    // FP32 tiles read by columns (op = N): 32 rows per workgroup (128-byte column pieces)
    if (a.nrhs > 4) {
        dim3 g16((unsigned)(a.ntasks * (a.b / 64))), b16(GV_THREADS);
        if (a.f32) {
            const size_t lds = (size_t)a.b * 20 * 4;
            if (a.trans) hipLaunchKernelGGL(tile_gemm16_kernel<true>, g16, b16, lds, s, a);
            else hipLaunchKernelGGL(tile_gemm16_kernel<false>, g16, b16, lds, s, a);
        } else {
            const size_t lds = (size_t)a.b * 20 * 8;
            if (a.trans) hipLaunchKernelGGL(tile_dgemm16_kernel<true>, g16, b16, lds, s, a);
            else hipLaunchKernelGGL(tile_dgemm16_kernel<false>, g16, b16, lds, s, a);
        }
        return hipGetLastError();
    }
    const bool wide = a.f32 && !a.trans;
    dim3 grid((unsigned)(a.ntasks * (a.b / (wide ? 32 : GV_ROWS)))), block(GV_THREADS);
    auto lds = [&](int nr) { return ((size_t)q * a.b * nr + (size_t)GV_THREADS * nr) * 8; };
    // 1..4 right-hand sides here (5..16 took the MFMA kernels above)
#define GV_LAUNCH(NR, E, ROWS)                                                                           \
    do {                                                                                                 \
        if (q > 1) hipLaunchKernelGGL((tile_gemv_kernel<NR, E, ROWS, true>), grid, block, lds(NR), s, a);  \
        else hipLaunchKernelGGL((tile_gemv_kernel<NR, E, ROWS, false>), grid, block, lds(NR), s, a);       \
    } while (0)
    if (wide) {
        if (a.nrhs == 1) GV_LAUNCH(1, float, 32);
        else GV_LAUNCH(4, float, 32);
    } else if (a.f32) {
        if (a.nrhs == 1) GV_LAUNCH(1, float, 16);
        else GV_LAUNCH(4, float, 16);
    } else {
        if (a.nrhs == 1) GV_LAUNCH(1, double, 16);
        else GV_LAUNCH(4, double, 16);
    }
#undef GV_LAUNCH
    return hipGetLastError();
}

// FP64 -> FP32 copy of the tiles the sweeps read (upper_only: tiles (m, n),
// m <= n, as the Cholesky sweeps use) or of a plain array (N = 0: n elements)
__global__ __launch_bounds__(256) void to_f32_kernel(const double *src, float *dst, int N, int b, int upper_only,
                                                     size_t n)
{
    const size_t total = N ? (size_t)N * N : n;
    const size_t bb = (size_t)b * b;
    const int MT = N ? N / b : 1;
    for (size_t e = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4; e < total;
         e += (size_t)gridDim.x * blockDim.x * 4) {
        if (N && upper_only) {
            // layout L: tile (m, n) occupies [bb*(m + MT*n), +bb)
            const size_t tile = e / bb;
            const int m = (int)(tile % MT), nn = (int)(tile / MT);
            if (m > nn) continue;
        }
        const double2 x = *(const double2 *)(src + e), y = *(const double2 *)(src + e + 2);
        *(float4 *)(dst + e) = make_float4((float)x.x, (float)x.y, (float)y.x, (float)y.y);
    }
}

hipError_t launch_to_f32(const double *src, float *dst, int N, int b, bool upper_only, size_t n, hipStream_t s)
{
    const size_t total = N ? (size_t)N * N : n;
    if (total % 4) return hipErrorInvalidValue;
    hipLaunchKernelGGL(to_f32_kernel, dim3(8192), dim3(256), 0, s, src, dst, N, b, upper_only ? 1 : 0, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------- block inverse inputs
// One workgroup per 64 x 64 sub-block of one tile (K, i, j) of the q x q tile
// blocks: zeros above the block diagonal, op(D) on it, the tiles of T below
// it.  Transposed copies go through LDS, so reads and writes both run down
// columns.
template <typename E>
__global__ __launch_bounds__(256) void blk_gather_kernel(const E *F, int N, int b, int q, int upper, int dtrans,
                                                         const double *D, double *X, double *W1)
{
    __shared__ double tl[64][65];
    const int nsb = b / 64, per_tile = nsb * nsb;
    const long id = blockIdx.x;
    const int sub = (int)(id % per_tile);
    long tt = id / per_tile;
    const int j = (int)(tt % q);
    tt /= q;
    const int i = (int)(tt % q);
    const int K = (int)(tt / q);
    const int r0 = (sub % nsb) * 64, c0 = (sub / nsb) * 64;   // output sub-block origin
    const int MT = N / b;
    const GatherTile gt = gather_tile(MT, b, q, upper != 0, dtrans != 0, K, i, j);
    const int t = threadIdx.x, r = t & 63, cw = t >> 6;
    if (gt.kind == 0) {
        for (int c = cw; c < 64; c += 4) X[gt.dst + (size_t)(r0 + r) + (size_t)(c0 + c) * b] = 0.0;
        return;
    }
    const E *srcE = gt.kind == 2 ? F + gt.src : nullptr;
    const double *srcD = gt.kind == 1 ? D + gt.src : nullptr;
    double *dst = (gt.kind == 1 ? X : W1) + gt.dst;
    const bool tr = gt.tr;
    auto src = [&](int rr, int cc) -> double {
        const size_t e = (size_t)rr + (size_t)cc * b;
        return srcD ? srcD[e] : (double)srcE[e];
    };
    if (!tr) {
        for (int c = cw; c < 64; c += 4) dst[(size_t)(r0 + r) + (size_t)(c0 + c) * b] = src(r0 + r, c0 + c);
    } else {
        // dst(r0 + r, c0 + c) = src(c0 + c, r0 + r): stage tl[c][r] = src(c0 + r, r0 + c)
        for (int c = cw; c < 64; c += 4) tl[c][r] = src(c0 + r, r0 + c);
        __syncthreads();
        for (int c = cw; c < 64; c += 4) dst[(size_t)(r0 + r) + (size_t)(c0 + c) * b] = tl[r][c];
    }
}

hipError_t launch_blk_gather(const void *F, bool f32, int N, int b, int q, bool upper, bool dtrans,
                             const double *D, double *X, double *W1, hipStream_t s)
{
    if (q < 2 || b % 64 || N % b || (N / b) % q) return hipErrorInvalidValue;
    const long KT = (long)(N / b) / q, nsb = b / 64;
    const dim3 grid((unsigned)(KT * q * q * nsb * nsb)), block(256);
    if (f32)
        hipLaunchKernelGGL((blk_gather_kernel<float>), grid, block, 0, s, (const float *)F, N, b, q, (int)upper,
                           (int)dtrans, D, X, W1);
    else
        hipLaunchKernelGGL((blk_gather_kernel<double>), grid, block, 0, s, (const double *)F, N, b, q, (int)upper,
                           (int)dtrans, D, X, W1);
    return hipGetLastError();
}

// ---------------------------------------------------------------- residual
// r = b - A x for the FP64 layout-L matrix, in two deterministic passes:
// partial sums into fixed slots, then a reduction over the slots in order.
// NORM additionally accumulates the row sums of |A| (for ||A||_inf) in the
// same pass over A.
//
// General matrices: wave w of a workgroup owns the 64 rows of row block
// 4 g + w (g = the workgroup's row group) and runs over a chunk of RES_CHUNK
// tile columns; slot = chunk.  The x segment of each tile column is staged in
// LDS once for the four waves, row c of it (the NR right-hand sides) stored
// contiguously, so that a column of A costs one coalesced load per lane and
// NR / 2 broadcast 16-byte LDS reads.  xt: x transposed to [row][NR] (NR > 1).
constexpr int RES_CHUNK = 8;

template <int NR>
__global__ void residual_xt_kernel(const double *x, int N, int ld, int nrhs, double *xt)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < (size_t)N * NR;
         e += (size_t)gridDim.x * blockDim.x) {
        const size_t i = e / NR;
        const int j = (int)(e % NR);
        xt[e] = j < nrhs ? x[i + (size_t)j * ld] : 0.0;   // padded right-hand sides are zero
    }
}

template <int NR, bool NORM>
__global__ __launch_bounds__(256) void residual_partial_kernel(const double *A, int N, int bt, const double *x,
                                                               int nrhs, double *part, double *partn)
{
    constexpr int SX = NR == 1 ? 1 : NR + 2;   // LDS row stride (16-byte aligned, 4-way write conflicts)
    extern __shared__ double sx[];             // bt x SX
    const int MT = N / bt;
    const int nchunk = (MT + RES_CHUNK - 1) / RES_CHUNK;
    const int rg = blockIdx.x / nchunk, chunk = blockIdx.x % nchunk;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int rowblk = rg * 4 + w;
    const bool active = rowblk < N / 64;
    const int rpt = bt / 64;
    const int m = rowblk / rpt, r0 = (rowblk % rpt) * 64;
    double acc[NR], nacc = 0.0;
#pragma unroll
    for (int j = 0; j < NR; j++) acc[j] = 0.0;
    const int n1 = min(MT, (chunk + 1) * RES_CHUNK);
    for (int n = chunk * RES_CHUNK; n < n1; n++) {
        __syncthreads();   // every wave is done with the previous segment
        const double *xs = x + (size_t)n * bt * NR;
        for (int e = t; e < bt * NR; e += 256) sx[(e / NR) * SX + e % NR] = xs[e];
        __syncthreads();
        if (active) {
            const double *T = A + tile_off(bt, MT, m, n) + r0 + lane;
#pragma unroll 8
            for (int c = 0; c < bt; c++) {
                const double a = T[(size_t)c * bt];
                if (NORM) nacc += fabs(a);
                const double *xc = sx + c * SX;
#pragma unroll
                for (int j = 0; j < NR; j++) acc[j] += a * xc[j];
            }
        }
    }
    if (active) {
        const size_t row = (size_t)m * bt + r0 + lane;
#pragma unroll
        for (int j = 0; j < NR; j++)
            if (j < nrhs) part[((size_t)chunk * NR + j) * N + row] = acc[j];
        if (NORM) partn[(size_t)chunk * N + row] = nacc;
    }
}

// Symmetric A given by its upper tiles (the Cholesky solves; one right-hand
// side): a workgroup per tile (i, j), i <= j, reads the tile once and
// produces A_ij x_j (slot j of row block i) and A_ij^T x_i (slot i of row
// block j); the diagonal tile uses its upper triangle for both and adds the
// two into slot i.  Wave w owns the columns c = w, w+4, ...; lane l holds
// rows l + 64q (q < b/64), so a column costs b/64 coalesced loads per lane
// and one cross-lane reduction (its transposed sum, owned by this wave).
// The direct partials of the four waves are combined in a fixed order.
template <bool NORM>
__global__ __launch_bounds__(256) void residual_sym_kernel(const double *A, int N, int bt, const double *x,
                                                           double *part, double *partn)
{
    __shared__ double sxj[512];
    __shared__ double scol[512], sncol[512];
    __shared__ double sdir[4][512], sndir[4][512];
    const int MT = N / bt;
    int ti, tj;
    tri_decode((int)blockIdx.x, ti, tj);                  // ti <= tj
    const double *T = A + tile_off(bt, MT, ti, tj);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int e = t; e < bt; e += 256) sxj[e] = x[(size_t)tj * bt + e];
    const int nq = bt / 64;
    const bool diag = (ti == tj);
    double xr[8], accd[8], nd[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        xr[q] = (q < nq) ? x[(size_t)ti * bt + lane + 64 * q] : 0.0;
        accd[q] = 0.0;
        nd[q] = 0.0;
    }
    __syncthreads();
    for (int c = w; c < bt; c += 4) {
        const double xc = sxj[c];
        const double *col = T + (size_t)c * bt + lane;
        double a[8];
#pragma unroll
        for (int q = 0; q < 8; q++) a[q] = (q < nq) ? col[64 * q] : 0.0;
        double s = 0.0, sn = 0.0;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int r = lane + 64 * q;
            const bool dir = !diag || r <= c, tr = !diag || r < c;
            const double ad = dir ? a[q] : 0.0, at = tr ? a[q] : 0.0;
            accd[q] += ad * xc;
            s += at * xr[q];
            if (NORM) {
                nd[q] += fabs(ad);
                sn += fabs(at);
            }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            s += __shfl_xor(s, o);
            if (NORM) sn += __shfl_xor(sn, o);
        }
        if (lane == 0) {
            scol[c] = s;
            if (NORM) sncol[c] = sn;
        }
    }
#pragma unroll
    for (int q = 0; q < 8; q++)
        if (q < nq) {
            sdir[w][lane + 64 * q] = accd[q];
            if (NORM) sndir[w][lane + 64 * q] = nd[q];
        }
    __syncthreads();
    for (int r = t; r < bt; r += 256) {
        double v = sdir[0][r] + sdir[1][r] + sdir[2][r] + sdir[3][r];
        double vn = NORM ? sndir[0][r] + sndir[1][r] + sndir[2][r] + sndir[3][r] : 0.0;
        if (diag) {
            v += scol[r];
            if (NORM) vn += sncol[r];
        } else {
            part[(size_t)ti * N + (size_t)tj * bt + r] = scol[r];
            if (NORM) partn[(size_t)ti * N + (size_t)tj * bt + r] = sncol[r];
        }
        part[(size_t)tj * N + (size_t)ti * bt + r] = v;
        if (NORM) partn[(size_t)tj * N + (size_t)ti * bt + r] = vn;
    }
}

template <int NR>
__global__ void residual_reduce_kernel(const double *part, const double *partn, int nslot, const double *rhs,
                                       double *r, double *rowsum, int N, int ld, int nrhs)
{
    const size_t total = (size_t)N * nrhs;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t i = e % N, j = e / N;
        double s = 0.0;
        for (int c = 0; c < nslot; c++) s += part[((size_t)c * NR + j) * N + i];
        r[i + j * ld] = rhs[i + j * ld] - s;
        if (rowsum && j == 0) {
            double sn = 0.0;
            for (int c = 0; c < nslot; c++) sn += partn[(size_t)c * N + i];
            rowsum[i] = sn;
        }
    }
}

hipError_t launch_residual(const double *A, int N, int b, const double *x, const double *rhs, double *r, int ld,
                           int nrhs, double *part, bool sym, double *rowsum, hipStream_t s)
{
    if (b % 64 || b > 512 || nrhs > GV_MAXR || nrhs < 1) return hipErrorInvalidValue;
    const int MT = N / b;
    if (sym && nrhs == 1) {
        double *partn = part + (size_t)MT * N;
        dim3 grid((unsigned)(MT * (MT + 1) / 2)), block(256);
        if (rowsum) hipLaunchKernelGGL(residual_sym_kernel<true>, grid, block, 0, s, A, N, b, x, part, partn);
        else hipLaunchKernelGGL(residual_sym_kernel<false>, grid, block, 0, s, A, N, b, x, part, partn);
        hipLaunchKernelGGL(residual_reduce_kernel<1>, dim3(1024), dim3(256), 0, s, part, partn, MT, rhs, r, rowsum,
                           N, ld, nrhs);
        return hipGetLastError();
    }
    if (sym) return hipErrorInvalidValue;   // symmetric residual: one right-hand side
    const int nchunk = (MT + RES_CHUNK - 1) / RES_CHUNK;
    const int nrg = (N / 64 + 3) / 4;
    dim3 grid((unsigned)(nrg * nchunk)), block(256);
    const int NRk = (nrhs == 1) ? 1 : 16;
    double *partn = part + (size_t)nchunk * NRk * N;
    if (nrhs == 1) {
        const size_t lds = (size_t)b * 8;
        if (rowsum) hipLaunchKernelGGL((residual_partial_kernel<1, true>), grid, block, lds, s, A, N, b, x, nrhs, part, partn);
        else hipLaunchKernelGGL((residual_partial_kernel<1, false>), grid, block, lds, s, A, N, b, x, nrhs, part, partn);
        hipLaunchKernelGGL(residual_reduce_kernel<1>, dim3(1024), dim3(256), 0, s, part, partn, nchunk, rhs, r, rowsum,
                           N, ld, nrhs);
    } else {
        double *xt = partn + (size_t)nchunk * N;
        const size_t lds = (size_t)b * 18 * 8;
        hipLaunchKernelGGL(residual_xt_kernel<16>, dim3(512), dim3(256), 0, s, x, N, ld, nrhs, xt);
        if (rowsum) hipLaunchKernelGGL((residual_partial_kernel<16, true>), grid, block, lds, s, A, N, b, xt, nrhs, part, partn);
        else hipLaunchKernelGGL((residual_partial_kernel<16, false>), grid, block, lds, s, A, N, b, xt, nrhs, part, partn);
        hipLaunchKernelGGL(residual_reduce_kernel<16>, dim3(1024), dim3(256), 0, s, part, partn, nchunk, rhs, r,
                           rowsum, N, ld, nrhs);
    }
    return hipGetLastError();
}

// F(i, i) += sigma |A(i, i)| on the working copy F (layout L, tile b, FP64 or
// FP32) from the FP64 matrix A (layout L, tile bA)
__global__ void add_diag_kernel(void *F, int f32, int N, int b, const double *A, int bA, double sigma)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const double d = A[tile_off(bA, N / bA, i / bA, i / bA) + (size_t)(i % bA) * (bA + 1)];
        const size_t o = tile_off(b, N / b, i / b, i / b) + (size_t)(i % b) * (b + 1);
        if (f32) ((float *)F)[o] = (float)((double)((float *)F)[o] + sigma * fabs(d));
        else ((double *)F)[o] += sigma * fabs(d);
    }
}

hipError_t launch_add_diag(void *F, bool f32, int N, int b, const double *A, int bA, double sigma, hipStream_t s)
{
    hipLaunchKernelGGL(add_diag_kernel, dim3((N + 255) / 256), dim3(256), 0, s, F, f32 ? 1 : 0, N, b, A, bA, sigma);
    return hipGetLastError();
}

void solve_init()
{
    // the FP64 16-right-hand-side sweeps stage b x 20 doubles (80 KB at b = 512)
    hipFuncSetAttribute((const void *)tile_dgemm16_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        512 * 20 * 8);
    hipFuncSetAttribute((const void *)tile_dgemm16_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        512 * 20 * 8);
    // the 16-right-hand-side residual stages b x 18 doubles (74 KB at b = 512)
    hipFuncSetAttribute((const void *)residual_partial_kernel<16, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        512 * 18 * 8);
    hipFuncSetAttribute((const void *)residual_partial_kernel<16, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        512 * 18 * 8);
}

// device scratch (doubles) launch_residual needs in `part`
size_t residual_scratch(int N, int b, int nrhs)
{
    const int MT = N / b;
    const int nchunk = (MT + RES_CHUNK - 1) / RES_CHUNK;
    const size_t chunked = (size_t)nchunk * ((nrhs == 1 ? 1 : 16) + 1) * N + (nrhs == 1 ? 0 : (size_t)16 * N);
    const size_t sym = 2 * (size_t)MT * N;
    return chunked > sym ? chunked : sym;
}

// max that keeps a NaN (fmax drops it): a NaN anywhere in the column must
// make its norm NaN, or a broken solution would read as a zero residual
__device__ inline double max_nan(double m, double a) { return (a > m || a != a) ? a : m; }

// per-column max |v| of an N x nrhs block (one workgroup per column); NaN if
// the column holds one
__global__ __launch_bounds__(256) void colmax_abs_kernel(const double *v, int N, int ld, double *out)
{
    __shared__ double sred[256];
    const int j = blockIdx.x, t = threadIdx.x;
    double m = 0.0;
    for (int i = t; i < N; i += 256) m = max_nan(m, fabs(v[(size_t)i + (size_t)j * ld]));
    sred[t] = m;
    __syncthreads();
    for (int s = 128; s >= 1; s >>= 1) {
        if (t < s) sred[t] = max_nan(sred[t], sred[t + s]);
        __syncthreads();
    }
    if (t == 0) out[j] = sred[0];
}

hipError_t launch_colmax(const double *v, int N, int ld, int nrhs, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(colmax_abs_kernel, dim3((unsigned)nrhs), dim3(256), 0, s, v, N, ld, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------- GMRES helpers
// out[i] = V(:, i) . w for i < k (one workgroup per column, fixed-order
// reduction); V has leading dimension ldv
__global__ __launch_bounds__(256) void multi_dot_kernel(const double *V, size_t ldv, const double *w, int N,
                                                        double *out)
{
    __shared__ double sred[256];
    const int i = blockIdx.x, t = threadIdx.x;
    const double *v = V + (size_t)i * ldv;
    double a = 0.0;
    for (int r = t; r < N; r += 256) a += v[r] * w[r];
    sred[t] = a;
    __syncthreads();
    for (int s = 128; s >= 1; s >>= 1) {
        if (t < s) sred[t] += sred[t + s];
        __syncthreads();
    }
    if (t == 0) out[i] = sred[0];
}

// w = alpha * w - V(:, 0:k) h  (h on the device; alpha = 0: w is not read)
__global__ void vh_update_kernel(const double *V, size_t ldv, int k, const double *h, double *w, double alpha,
                                 int N)
{
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < N; r += gridDim.x * blockDim.x) {
        double a = (alpha == 0.0) ? 0.0 : alpha * w[r];   // alpha = 0 ignores w entirely
        for (int i = 0; i < k; i++) a -= V[(size_t)i * ldv + r] * h[i];
        w[r] = a;
    }
}

hipError_t launch_multi_dot(const double *V, size_t ldv, int k, const double *w, int N, double *out, hipStream_t s)
{
    if (k <= 0) return hipSuccess;
    hipLaunchKernelGGL(multi_dot_kernel, dim3((unsigned)k), dim3(256), 0, s, V, ldv, w, N, out);
    return hipGetLastError();
}

hipError_t launch_vh_update(const double *V, size_t ldv, int k, const double *h, double *w, double alpha, int N,
                            hipStream_t s)
{
    hipLaunchKernelGGL(vh_update_kernel, dim3(512), dim3(256), 0, s, V, ldv, k, h, w, alpha, N);
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
