// k_gemm.hip -- batched FP64 tile GEMM on CDNA4 matrix cores.
//
// One launch performs every tile contraction of one DAG stage (e.g. all
// GEMM+SYRK tasks of Cholesky step k, src/cholesky.h:107-173), instead of one
// launch per tile task.  A workgroup owns a BM x BM block of one b x b output
// tile; K (= b) is streamed through LDS in KB = 16 slices with register
// prefetch; each wave accumulates a 64x64 (BM = 128) or 32x32 (BM = 32)
// sub-block with v_mfma_f64_16x16x4_f64.
//
// MFMA 16x16x4 f64 operand layout (lane l of 64):
//   A fragment: A[row = l % 16][k = l / 16]
//   B fragment: B[k = l / 16][col = l % 16]
//   D (4 doubles): D[row = (l/16) + 4*v][col = l % 16], v = 0..3
//   (the f64 variant interleaves rows across the four lane groups)
// The operands are passed swapped (D = B^T A^T), so lane l's accumulators sit
// in row l%16 of C and the epilogue moves 128-byte row runs.
#include "gemm_common.hpp"
#include <cstring>
#include <cstdlib>

namespace palial {

typedef double d4 __attribute__((ext_vector_type(4)));

template <int BM, int KB, int MINB, bool TA, bool TB>
__global__ __launch_bounds__(2 * BM, MINB) void gemm_mfma_kernel(GemmArgs p)
{
    constexpr int NT = 2 * BM;          // threads per workgroup
    constexpr int WM = (BM >= 64) ? 64 : BM;
    constexpr int WAVES_M = BM / WM;
    constexpr int FM = WM / 16;         // MFMA tiles per wave edge
    constexpr int LD = BM + 16;         // LDS row stride (doubles): rows 128 B apart mod 256 B
    constexpr int EPT = BM * KB / NT;   // operand elements per thread per K slice (= KB/2)
    constexpr int BUF = KB * LD;        // one LDS buffer of one operand

    // double-buffered LDS: slice t is computed from buffer t&1 while slice t+1
    // is staged into the other buffer; one barrier per slice
    __shared__ double sA[2 * BUF];
    __shared__ double sB[2 * BUF];

    const int b = p.b;
    const int nbd = b / BM;
    const int blk_per_task = nbd * nbd;
    const int lid = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int task = lid / blk_per_task;
    const int blk = lid % blk_per_task;
    const int bi = blk % nbd, bj = blk / nbd;
    if (task >= p.ntasks) return;
    TaskPtrs tp = decode(p, task);
    if (!tp.valid) return;
    if (tp.diag && bi > bj) return;     // strictly below the diagonal of a SYRK tile
    task_begin(p.ts, p.ts_base + task);

    const int i0 = bi * BM, j0 = bj * BM;
    const int t = (int)threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int wm = w % WAVES_M, wn = w / WAVES_M;
    const int mask = p.maskA;

    // K range: triangular operands contribute only inside their triangle
    int kbeg = 0, kend = b;
    if (mask == MK_UPPER) {
        if (TA) kend = min(b, i0 + BM);       // opA(i,k) = T(k,i), k <= i
        else kbeg = i0;                       // opA(i,k) = T(i,k), i <= k
    } else if (mask == MK_UNIT_LOWER) {
        if (TA) kbeg = i0;                    // opA(i,k) = V(k,i), k >= i
        else kend = min(b, i0 + BM);          // opA(i,k) = V(i,k), k <= i
    }
    kbeg = (kbeg / KB) * KB;

    d4 acc[FM][FM];
#pragma unroll
    for (int a = 0; a < FM; a++)
#pragma unroll
        for (int c = 0; c < FM; c++) acc[a][c] = (d4){0.0, 0.0, 0.0, 0.0};

    double ra[EPT], rb[EPT];

    // ---- global -> registers.  K-contiguous operands: each thread loads EPT
    // consecutive k of one row; MN-contiguous: chunks of 8 along the row.
    auto load_a = [&](int kb) {
        if (TA) {   // opA(i,k) = X[k + i*b]
            const int row = t >> 1, kc = (t & 1) * EPT;
            const double2 *src = (const double2 *)(tp.A + (size_t)(i0 + row) * b + kb + kc);
#pragma unroll
            for (int q = 0; q < EPT / 2; q++) { double2 v = src[q]; ra[2 * q] = v.x; ra[2 * q + 1] = v.y; }
            if (mask) {
#pragma unroll
                for (int q = 0; q < EPT; q++) ra[q] = apply_mask(mask, ra[q], kb + kc + q, i0 + row);
            }
        } else {    // opA(i,k) = X[i + k*b]
#pragma unroll
            for (int c = 0; c < EPT / 8; c++) {
                const int id = t + c * NT;
                const int kk = id / (BM / 8), ic = (id % (BM / 8)) * 8;
                const double2 *src = (const double2 *)(tp.A + (size_t)(kb + kk) * b + i0 + ic);
#pragma unroll
                for (int q = 0; q < 4; q++) { double2 v = src[q]; ra[8 * c + 2 * q] = v.x; ra[8 * c + 2 * q + 1] = v.y; }
                if (mask) {
#pragma unroll
                    for (int q = 0; q < 8; q++) ra[8 * c + q] = apply_mask(mask, ra[8 * c + q], i0 + ic + q, kb + kk);
                }
            }
        }
    };
    auto load_b = [&](int kb) {
        if (!TB) {  // opB(k,j) = Y[k + j*b]
            const int col = t >> 1, kc = (t & 1) * EPT;
            const double2 *src = (const double2 *)(tp.B + (size_t)(j0 + col) * b + kb + kc);
#pragma unroll
            for (int q = 0; q < EPT / 2; q++) { double2 v = src[q]; rb[2 * q] = v.x; rb[2 * q + 1] = v.y; }
        } else {    // opB(k,j) = Y[j + k*b]
#pragma unroll
            for (int c = 0; c < EPT / 8; c++) {
                const int id = t + c * NT;
                const int kk = id / (BM / 8), jc = (id % (BM / 8)) * 8;
                const double2 *src = (const double2 *)(tp.B + (size_t)(kb + kk) * b + j0 + jc);
#pragma unroll
                for (int q = 0; q < 4; q++) { double2 v = src[q]; rb[8 * c + 2 * q] = v.x; rb[8 * c + 2 * q + 1] = v.y; }
            }
        }
    };
    // ---- registers -> LDS ([k][mn] layout)
    auto store_lds = [&](int buf) {
        double *a = sA + buf * BUF, *bb = sB + buf * BUF;
        if (TA) {
            const int row = t >> 1, kc = (t & 1) * EPT;
#pragma unroll
            for (int q = 0; q < EPT; q++) a[(kc + q) * LD + row] = ra[q];
        } else {
#pragma unroll
            for (int c = 0; c < EPT / 8; c++) {
                const int id = t + c * NT;
                const int kk = id / (BM / 8), ic = (id % (BM / 8)) * 8;
#pragma unroll
                for (int q = 0; q < 8; q++) a[kk * LD + ic + q] = ra[8 * c + q];
            }
        }
        if (!TB) {
            const int col = t >> 1, kc = (t & 1) * EPT;
#pragma unroll
            for (int q = 0; q < EPT; q++) bb[(kc + q) * LD + col] = rb[q];
        } else {
#pragma unroll
            for (int c = 0; c < EPT / 8; c++) {
                const int id = t + c * NT;
                const int kk = id / (BM / 8), jc = (id % (BM / 8)) * 8;
#pragma unroll
                for (int q = 0; q < 8; q++) bb[kk * LD + jc + q] = rb[8 * c + q];
            }
        }
    };

    if (kbeg < kend) {
        load_a(kbeg);
        load_b(kbeg);
        store_lds(0);
        __syncthreads();
        int buf = 0;
        for (int kb = kbeg; kb < kend; kb += KB) {
            const bool more = kb + KB < kend;
            if (more) { load_a(kb + KB); load_b(kb + KB); }
            const double *a = sA + buf * BUF, *bb = sB + buf * BUF;
#pragma unroll
            for (int kk = 0; kk < KB; kk += 4) {
                double av[FM], bv[FM];
                const int kr = kk + (lane >> 4);
#pragma unroll
                for (int x = 0; x < FM; x++) av[x] = a[kr * LD + wm * WM + x * 16 + (lane & 15)];
#pragma unroll
                for (int c = 0; c < FM; c++) bv[c] = bb[kr * LD + wn * WM + c * 16 + (lane & 15)];
#pragma unroll
                for (int x = 0; x < FM; x++)
#pragma unroll
                    for (int c = 0; c < FM; c++)   // operands swapped: acc holds C^T fragments
                        acc[x][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[c], av[x], acc[x][c], 0, 0, 0);
            }
            if (more) store_lds(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }

    // ---- epilogue: C = alpha*acc + beta*D ; E -= alpha*acc.
    // acc[x][c][v] = C(i0 + wm*WM + 16x + l%16, j0 + wn*WM + 16c + l/16 + 4v):
    // 16 lanes cover 16 consecutive rows (128 B) of one column.  D (which may
    // be C itself) is read for a whole fragment row before any store, so the
    // loads are in flight together instead of one round trip each.  E (the
    // QR updates only; never aliased with C or D) follows in its own pass.
    const double alpha = p.alpha, beta = p.beta;
    auto ldD = [&](int x, double (&dv)[FM][4]) {
        const int i = i0 + wm * WM + x * 16 + (lane & 15);
#pragma unroll
        for (int c = 0; c < FM; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int j = j0 + wn * WM + c * 16 + (lane >> 4) + 4 * v;
                dv[c][v] = (tp.D && !(tp.diag && i > j)) ? tp.D[(size_t)i + (size_t)j * b] : 0.0;
            }
    };
    auto stC = [&](int x, const double (&dv)[FM][4]) {
        const int i = i0 + wm * WM + x * 16 + (lane & 15);
#pragma unroll
        for (int c = 0; c < FM; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int j = j0 + wn * WM + c * 16 + (lane >> 4) + 4 * v;
                if (tp.diag && i > j) continue;
                const double r = alpha * acc[x][c][v];
                tp.C[(size_t)i + (size_t)j * b] = tp.D ? r + beta * dv[c][v] : r;
            }
    };
    if constexpr (FM == 4) {
        // two fragment rows of D in flight ahead of the stores (D may be C: a
        // row is read completely before any of its elements is written)
        double d0[FM][4], d1[FM][4];
        ldD(0, d0);
        ldD(1, d1);
        stC(0, d0);
        ldD(2, d0);
        stC(1, d1);
        ldD(3, d1);
        stC(2, d0);
        stC(3, d1);
    } else {
#pragma unroll
        for (int x = 0; x < FM; x++) {
            double dv[FM][4];
            ldD(x, dv);
            stC(x, dv);
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
                    const int j = j0 + wn * WM + c * 16 + (lane >> 4) + 4 * v;
                    if (tp.diag && i > j) continue;
                    tp.E[(size_t)i + (size_t)j * b] -= alpha * acc[x][c][v];
                }
    }
    task_end(p.ts, p.ts_base + task);
}

// Scalar fallback for tile sizes that are not a multiple of 32 (tiny tests,
// e.g. the reference's 4x4 matrix with b = 2).  One thread per output element.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_scalar_kernel(GemmArgs p)
{
    const int b = p.b;
    const long per_task = (long)b * b;
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int task = (int)(gid / per_task);
    if (task >= p.ntasks) return;
    const int e = (int)(gid % per_task);
    const int i = e % b, j = e / b;
    TaskPtrs tp = decode(p, task);
    if (!tp.valid) return;
    if (tp.diag && i > j) return;
    double s = 0.0;
    for (int k = 0; k < b; k++) {
        double av = TA ? apply_mask(p.maskA, tp.A[k + (size_t)i * b], k, i)
                       : apply_mask(p.maskA, tp.A[i + (size_t)k * b], i, k);
        double bv = TB ? tp.B[j + (size_t)k * b] : tp.B[k + (size_t)j * b];
        s += av * bv;
    }
    const size_t o = (size_t)i + (size_t)j * b;
    const double r = p.alpha * s;
    double out = r;
    if (tp.D) out += p.beta * tp.D[o];
    tp.C[o] = out;
    if (tp.E) tp.E[o] -= r;
}

}  // namespace palial

// Host-side view of the task enumeration (test hook): for each task of a
// batched launch, the element offsets of its A, B and output tiles relative to
// the matrix base (or -1 when the operand is not a matrix tile).  Lets the CPU
// tests prove every launch stays inside the N x N buffer.
extern "C" int palial_hip_task_tiles(int mode, int MT, int b, int k, int m, long *out, int cap)
{
    using namespace palial;
    GemmArgs a;
    memset(&a, 0, sizeof a);
    a.mode = mode;
    a.b = b;
    a.MT = MT;
    a.N = MT * b;
    a.k = k;
    a.m = m;
    a.T = MT - k - 1;
    a.ntasks = (int)gemm_task_count(mode, a.T);
    // a fake base far from zero so offsets are recoverable
    const size_t base = (size_t)1 << 40;
    a.mat = (double *)(base * 8);
    a.smat = (const double *)(base * 8 * 4);
    a.w1 = (double *)(base * 8 * 8);
    a.w2 = (double *)(base * 8 * 12);
    int n = 0;
    auto rel = [&](const double *p) -> long {
        if (!p) return -1;
        const long d = (long)((const double *)p - (const double *)a.mat);
        return (d >= 0 && d < (long)base) ? d : -2;   // -2: a workspace / S tile
    };
    for (int t = 0; t < a.ntasks; t++) {
        TaskPtrs tp = decode(a, t);
        if (n + 3 > cap) return -1;
        out[n++] = rel(tp.A);
        out[n++] = rel(tp.B);
        out[n++] = rel(tp.C);
    }
    return a.ntasks;
}

namespace palial {

long gemm_task_count(int mode, int T)
{
    switch (mode) {
    case GM_SINGLE: return 1;
    case GM_CHOL_TRAIL: return (long)T * (T + 1) / 2;
    case GM_CHOL_ROW: return T;
    case GM_CHOL_REST: return (long)(T - 1) * T / 2;
    case GM_LU_TRAIL: return (long)T * T;
    case GM_LU_PANEL: return 2L * T - 1;
    case GM_LU_REST: return (long)(T - 1) * (T - 1);
    case GM_CHOL_FIRST: return T + 1;
    case GM_LU_FIRST: return 2L * T;
    case GM_CHOL_ROW2: return T >= 3 ? T - 1 : 0;
    case GM_CHOL_REST2: return T >= 3 ? (long)(T - 2) * (T - 1) / 2 : 0;
    case GM_LU_ROW2: return T >= 3 ? 2L * T - 3 : 0;
    case GM_LU_REST2: return T >= 3 ? (long)(T - 2) * (T - 2) : 0;
    default: return T;  // QR / ORMQR modes: one task per trailing column tile
    }
}

template <int BM, int KB, int MINB>
static hipError_t launch_mfma(const GemmArgs &a, bool ta, bool tb, hipStream_t s)
{
    const long nbd = a.b / BM;
    const long nblk = (long)a.ntasks * nbd * nbd;
    if (nblk <= 0) return hipSuccess;
    dim3 grid((unsigned)nblk), block(2 * BM);
    if (ta && !tb) hipLaunchKernelGGL((gemm_mfma_kernel<BM, KB, MINB, true, false>), grid, block, 0, s, a);
    else if (!ta && !tb) hipLaunchKernelGGL((gemm_mfma_kernel<BM, KB, MINB, false, false>), grid, block, 0, s, a);
    else if (!ta && tb) hipLaunchKernelGGL((gemm_mfma_kernel<BM, KB, MINB, false, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((gemm_mfma_kernel<BM, KB, MINB, true, true>), grid, block, 0, s, a);
    return hipGetLastError();
}

static int gemm_variant()
{
    static int v = -1;
    if (v < 0) {
        const char *e = getenv("PALIAL_GEMM_VARIANT");
        v = e ? atoi(e) : 0;
    }
    return v;
}

hipError_t launch_gemm_fine(const GemmArgs &a_in, bool ta, bool tb, hipStream_t s)
{
    if (a_in.ntasks <= 0) return hipSuccess;
    if (sched_rec_on()) { rec_gemm(a_in, false, s); return hipSuccess; }
    if (a_in.b % 32 != 0) return launch_gemm(a_in, ta, tb, s);
    GemmArgs a = a_in;
    trace_gemm(a);
    return launch_mfma<32, 32, 1>(a, ta, tb, s);      // K slices of 32: half the barriers
}

hipError_t launch_gemm(const GemmArgs &a_in, bool ta, bool tb, hipStream_t s)
{
    if (a_in.ntasks <= 0) return hipSuccess;
    if (sched_rec_on()) { rec_gemm(a_in, false, s); return hipSuccess; }
    GemmArgs a = a_in;
    if (a.b % 32 == 0) trace_gemm(a);     // the scalar fallback below keeps no timestamps
    if (a.b % 128 == 0) {
        // 2 waves per SIMD (<= 256 registers per lane, LDS 2 x 73.7 KB per CU)
        switch (gemm_variant()) {
        case 9: return launch_mfma<128, 16, 1>(a, ta, tb, s);   // 1 wave per SIMD, for comparison
        default: return launch_mfma<128, 16, 2>(a, ta, tb, s);
        }
    }
    if (a.b % 32 == 0) return launch_mfma<32, 16, 1>(a, ta, tb, s);
    const long total = (long)a.ntasks * a.b * a.b;
    dim3 grid((unsigned)((total + 255) / 256)), block(256);
    if (ta && !tb) hipLaunchKernelGGL((gemm_scalar_kernel<true, false>), grid, block, 0, s, a);
    else if (!ta && !tb) hipLaunchKernelGGL((gemm_scalar_kernel<false, false>), grid, block, 0, s, a);
    else if (!ta && tb) hipLaunchKernelGGL((gemm_scalar_kernel<false, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((gemm_scalar_kernel<true, true>), grid, block, 0, s, a);
    return hipGetLastError();
}

}  // namespace palial
