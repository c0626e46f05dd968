// k_gemm_lp.hip -- batched tile GEMM with low-precision operands and FP32
// accumulation: the O(n^3) trailing updates of the mixed-precision
// factorizations (same task enumerations as k_gemm.hip).
//
//   PR_FP32: operands rounded to FP32, v_mfma_f32_16x16x4_f32
//   PR_BF16: operands rounded to BF16 (round to nearest even),
//            v_mfma_f32_16x16x32_bf16
//
// The FP64 output tile receives C = alpha*acc + beta*D in FP64, so only the
// update products carry the low precision; the factor stays stored in FP64.
// Workgroup: 256 threads, 128 x 128 output block, 2 x 2 waves of 64 x 64.
//
// Operand layouts (lane l of 64):
//   f32 16x16x4 : A[row = l%16][k = l/16], B[k = l/16][col = l%16]
//   bf16 16x16x32: A[row = l%16][k = 8*(l/16) .. +7], B[k = 8*(l/16) .. +7][col = l%16]
//   f32 result  : D[row = 4*(l/16) + v][col = l%16]
// (operands are passed swapped, see lp_epilogue)
#include "gemm_common.hpp"

namespace palial {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ inline unsigned short f2bf(float f)
{
    // round to nearest even (inputs are finite)
    const unsigned int u = __float_as_uint(f);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// Operands are fed to the MFMA swapped, so acc[x][c][v] holds
// C(i0 + wm*WM + 16x + l%16, j0 + wn*WM + 16c + 4(l/16) + v): 16 lanes cover
// 16 consecutive rows (128 B) of one column of C.  D (C itself in the
// trailing updates) is read for a whole fragment row before any store, and E
// (never aliased with C or D) is updated in its own pass.
template <int FM, int WM>
__device__ inline void lp_epilogue(const TaskPtrs &tp, const f4 (&acc)[FM][FM], int i0, int j0, int wm, int wn,
                                   int lane, int b, double alpha, double beta)
{
#pragma unroll
    for (int x = 0; x < FM; x++) {
        const int i = i0 + wm * WM + x * 16 + (lane & 15);
        double dv[FM][4];
#pragma unroll
        for (int c = 0; c < FM; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int j = j0 + wn * WM + c * 16 + 4 * (lane >> 4) + v;
                dv[c][v] = (tp.D && !(tp.diag && i > j)) ? tp.D[(size_t)i + (size_t)j * b] : 0.0;
            }
#pragma unroll
        for (int c = 0; c < FM; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int j = j0 + wn * WM + c * 16 + 4 * (lane >> 4) + v;
                if (tp.diag && i > j) continue;
                const double r = alpha * (double)acc[x][c][v];
                tp.C[(size_t)i + (size_t)j * b] = tp.D ? r + beta * dv[c][v] : r;
            }
    }
    if (tp.E) {
#pragma unroll
        for (int x = 0; x < FM; x++)
#pragma unroll
            for (int c = 0; c < FM; c++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = i0 + wm * WM + x * 16 + (lane & 15);
                    const int j = j0 + wn * WM + c * 16 + 4 * (lane >> 4) + v;
                    if (tp.diag && i > j) continue;
                    tp.E[(size_t)i + (size_t)j * b] -= alpha * (double)acc[x][c][v];
                }
    }
}

// ------------------------------------------------------------------ FP32
template <bool TA, bool TB>
__global__ __launch_bounds__(256, 2) void gemm_f32_kernel(GemmArgs p)
{
    constexpr int BM = 128, KB = 16, NT = 256, WM = 64, FM = 4;
    constexpr int LD = BM + 16;          // floats; k-rows 64 B apart modulo 256 B
    constexpr int EPT = BM * KB / NT;    // 8
    constexpr int BUF = KB * LD;
    __shared__ float sA[2 * BUF];
    __shared__ float sB[2 * BUF];

    const int b = p.b;
    const int nbd = b / BM;
    const int blk_per_task = nbd * nbd;
    const int lid = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int task = lid / blk_per_task, blk = lid % blk_per_task;
    const int bi = blk % nbd, bj = blk / nbd;
    if (task >= p.ntasks) return;
    TaskPtrs tp = decode(p, task);
    if (!tp.valid) return;
    if (tp.diag && bi > bj) return;
    task_begin(p.ts, p.ts_base + task);
    const int i0 = bi * BM, j0 = bj * BM;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;

    f4 acc[FM][FM];
#pragma unroll
    for (int a = 0; a < FM; a++)
#pragma unroll
        for (int c = 0; c < FM; c++) acc[a][c] = (f4){0.f, 0.f, 0.f, 0.f};
    float ra[EPT], rb[EPT];

    auto load = [&](int kb) {
        if (TA) {
            const int row = t >> 1, kc = (t & 1) * EPT;
            const double2 *src = (const double2 *)(tp.A + (size_t)(i0 + row) * b + kb + kc);
#pragma unroll
            for (int q = 0; q < EPT / 2; q++) { double2 v = src[q]; ra[2 * q] = (float)v.x; ra[2 * q + 1] = (float)v.y; }
        } else {
            const int kk = t / (BM / 8), ic = (t % (BM / 8)) * 8;
            const double2 *src = (const double2 *)(tp.A + (size_t)(kb + kk) * b + i0 + ic);
#pragma unroll
            for (int q = 0; q < 4; q++) { double2 v = src[q]; ra[2 * q] = (float)v.x; ra[2 * q + 1] = (float)v.y; }
        }
        if (!TB) {
            const int col = t >> 1, kc = (t & 1) * EPT;
            const double2 *src = (const double2 *)(tp.B + (size_t)(j0 + col) * b + kb + kc);
#pragma unroll
            for (int q = 0; q < EPT / 2; q++) { double2 v = src[q]; rb[2 * q] = (float)v.x; rb[2 * q + 1] = (float)v.y; }
        } else {
            const int kk = t / (BM / 8), jc = (t % (BM / 8)) * 8;
            const double2 *src = (const double2 *)(tp.B + (size_t)(kb + kk) * b + j0 + jc);
#pragma unroll
            for (int q = 0; q < 4; q++) { double2 v = src[q]; rb[2 * q] = (float)v.x; rb[2 * q + 1] = (float)v.y; }
        }
    };
    auto store = [&](int buf) {
        float *a = sA + buf * BUF, *bb = sB + buf * BUF;
        if (TA) {
            const int row = t >> 1, kc = (t & 1) * EPT;
#pragma unroll
            for (int q = 0; q < EPT; q++) a[(kc + q) * LD + row] = ra[q];
        } else {
            const int kk = t / (BM / 8), ic = (t % (BM / 8)) * 8;
#pragma unroll
            for (int q = 0; q < 8; q++) a[kk * LD + ic + q] = ra[q];
        }
        if (!TB) {
            const int col = t >> 1, kc = (t & 1) * EPT;
#pragma unroll
            for (int q = 0; q < EPT; q++) bb[(kc + q) * LD + col] = rb[q];
        } else {
            const int kk = t / (BM / 8), jc = (t % (BM / 8)) * 8;
#pragma unroll
            for (int q = 0; q < 8; q++) bb[kk * LD + jc + q] = rb[q];
        }
    };

    load(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int kb = 0; kb < b; kb += KB) {
        const bool more = kb + KB < b;
        if (more) load(kb + KB);
        const float *a = sA + buf * BUF, *bb = sB + buf * BUF;
#pragma unroll
        for (int kk = 0; kk < KB; kk += 4) {
            const int kr = kk + (lane >> 4);
            float av[FM], bv[FM];
#pragma unroll
            for (int x = 0; x < FM; x++) av[x] = a[kr * LD + wm * WM + x * 16 + (lane & 15)];
#pragma unroll
            for (int c = 0; c < FM; c++) bv[c] = bb[kr * LD + wn * WM + c * 16 + (lane & 15)];
#pragma unroll
            for (int x = 0; x < FM; x++)
#pragma unroll
                for (int c = 0; c < FM; c++)
                    acc[x][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[c], av[x], acc[x][c], 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }

    lp_epilogue<FM, WM>(tp, acc, i0, j0, wm, wn, lane, b, p.alpha, p.beta);
    task_end(p.ts, p.ts_base + task);
}

// ------------------------------------------------------------------ BF16
template <bool TA, bool TB>
__global__ __launch_bounds__(256, 2) void gemm_bf16_kernel(GemmArgs p)
{
    constexpr int BM = 128, KB = 32, NT = 256, WM = 64, FM = 4;
    constexpr int LDK = KB + 8;          // bf16 per LDS row ([mn][k] layout, k contiguous)
    constexpr int EPT = BM * KB / NT;    // 16 operand elements per thread per slice
    constexpr int BUF = BM * LDK;
    __shared__ unsigned short sA[2 * BUF];
    __shared__ unsigned short sB[2 * BUF];

    const int b = p.b;
    const int nbd = b / BM;
    const int blk_per_task = nbd * nbd;
    const int lid = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int task = lid / blk_per_task, blk = lid % blk_per_task;
    const int bi = blk % nbd, bj = blk / nbd;
    if (task >= p.ntasks) return;
    TaskPtrs tp = decode(p, task);
    if (!tp.valid) return;
    if (tp.diag && bi > bj) return;
    task_begin(p.ts, p.ts_base + task);
    const int i0 = bi * BM, j0 = bj * BM;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;

    f4 acc[FM][FM];
#pragma unroll
    for (int a = 0; a < FM; a++)
#pragma unroll
        for (int c = 0; c < FM; c++) acc[a][c] = (f4){0.f, 0.f, 0.f, 0.f};
    unsigned short ra[EPT], rb[EPT];

    // K-contiguous operand: thread owns 16 consecutive k of one row.
    // MN-contiguous operand: thread owns 8 consecutive rows at 2 k values.
    auto load = [&](int kb) {
        if (TA) {
            const int row = t >> 1, kc = (t & 1) * EPT;
            const double2 *src = (const double2 *)(tp.A + (size_t)(i0 + row) * b + kb + kc);
#pragma unroll
            for (int q = 0; q < EPT / 2; q++) { double2 v = src[q]; ra[2 * q] = f2bf((float)v.x); ra[2 * q + 1] = f2bf((float)v.y); }
        } else {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int id = t + h * NT;
                const int kk = id / (BM / 8), ic = (id % (BM / 8)) * 8;
                const double2 *src = (const double2 *)(tp.A + (size_t)(kb + kk) * b + i0 + ic);
#pragma unroll
                for (int q = 0; q < 4; q++) { double2 v = src[q]; ra[8 * h + 2 * q] = f2bf((float)v.x); ra[8 * h + 2 * q + 1] = f2bf((float)v.y); }
            }
        }
        if (!TB) {
            const int col = t >> 1, kc = (t & 1) * EPT;
            const double2 *src = (const double2 *)(tp.B + (size_t)(j0 + col) * b + kb + kc);
#pragma unroll
            for (int q = 0; q < EPT / 2; q++) { double2 v = src[q]; rb[2 * q] = f2bf((float)v.x); rb[2 * q + 1] = f2bf((float)v.y); }
        } else {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int id = t + h * NT;
                const int kk = id / (BM / 8), jc = (id % (BM / 8)) * 8;
                const double2 *src = (const double2 *)(tp.B + (size_t)(kb + kk) * b + j0 + jc);
#pragma unroll
                for (int q = 0; q < 4; q++) { double2 v = src[q]; rb[8 * h + 2 * q] = f2bf((float)v.x); rb[8 * h + 2 * q + 1] = f2bf((float)v.y); }
            }
        }
    };
    auto store = [&](int buf) {
        unsigned short *a = sA + buf * BUF, *bb = sB + buf * BUF;
        if (TA) {
            const int row = t >> 1, kc = (t & 1) * EPT;
#pragma unroll
            for (int q = 0; q < EPT; q++) a[row * LDK + kc + q] = ra[q];
        } else {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int id = t + h * NT;
                const int kk = id / (BM / 8), ic = (id % (BM / 8)) * 8;
#pragma unroll
                for (int q = 0; q < 8; q++) a[(ic + q) * LDK + kk] = ra[8 * h + q];
            }
        }
        if (!TB) {
            const int col = t >> 1, kc = (t & 1) * EPT;
#pragma unroll
            for (int q = 0; q < EPT; q++) bb[col * LDK + kc + q] = rb[q];
        } else {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int id = t + h * NT;
                const int kk = id / (BM / 8), jc = (id % (BM / 8)) * 8;
#pragma unroll
                for (int q = 0; q < 8; q++) bb[(jc + q) * LDK + kk] = rb[8 * h + q];
            }
        }
    };

    load(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int kb = 0; kb < b; kb += KB) {
        const bool more = kb + KB < b;
        if (more) load(kb + KB);
        const unsigned short *a = sA + buf * BUF, *bb = sB + buf * BUF;
        bf8 av[FM], bv[FM];
        const int ko = 8 * (lane >> 4);
#pragma unroll
        for (int x = 0; x < FM; x++) av[x] = *(const bf8 *)(a + (wm * WM + x * 16 + (lane & 15)) * LDK + ko);
#pragma unroll
        for (int c = 0; c < FM; c++) bv[c] = *(const bf8 *)(bb + (wn * WM + c * 16 + (lane & 15)) * LDK + ko);
#pragma unroll
        for (int x = 0; x < FM; x++)
#pragma unroll
            for (int c = 0; c < FM; c++)
                acc[x][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[c], av[x], acc[x][c], 0, 0, 0);
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }

    lp_epilogue<FM, WM>(tp, acc, i0, j0, wm, wn, lane, b, p.alpha, p.beta);
    task_end(p.ts, p.ts_base + task);
}

hipError_t launch_gemm_lp(const GemmArgs &a_in, bool ta, bool tb, int prec, hipStream_t s)
{
    if (a_in.ntasks <= 0) return hipSuccess;
    if (sched_rec_on()) { rec_gemm(a_in, false, s); return hipSuccess; }
    GemmArgs a = a_in;
    trace_gemm(a);
    if (a.b % 128 != 0 || a.maskA != MK_NONE) return hipErrorInvalidValue;
    const long nbd = a.b / 128;
    dim3 grid((unsigned)((long)a.ntasks * nbd * nbd)), block(256);
    if (prec == 1) {
        if (ta && !tb) hipLaunchKernelGGL((gemm_f32_kernel<true, false>), grid, block, 0, s, a);
        else if (!ta && !tb) hipLaunchKernelGGL((gemm_f32_kernel<false, false>), grid, block, 0, s, a);
        else return hipErrorInvalidValue;
    } else if (prec == 2) {
        if (ta && !tb) hipLaunchKernelGGL((gemm_bf16_kernel<true, false>), grid, block, 0, s, a);
        else if (!ta && !tb) hipLaunchKernelGGL((gemm_bf16_kernel<false, false>), grid, block, 0, s, a);
        else return hipErrorInvalidValue;
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace palial
