// k_gemm_pk.hip -- trailing updates of the mixed-precision factorizations
// from packed low-precision panels.
//
// The panel of step k (row tiles U_kj of Cholesky, L_ik and U_kj of LU) is
// rounded once to FP32, BF16, split BF16 or FP8 (E4M3, per-tile scale) by
// pack_kernel and stored K-contiguous, so the update GEMM streams 4, 2, 2 + 2
// or 1 bytes per operand element instead of re-reading and converting FP64
// tiles in every workgroup:
//   row-panel tile (k, j) : element (q, c) at q + c*b   (its own layout)
//   col-panel tile (i, k) : element (r, q) at q + r*b   (transposed)
// with q the K index.  The FP64 output tile receives C = alpha*acc + beta*D.
//
// MFMA operands are swapped (D = B^T A^T) so that each lane's accumulator
// column maps to 16 consecutive rows of the column-major C tile: every load
// and store of the FP64 epilogue covers 4 full 128-byte row runs.
//
// Operand layouts (lane l of 64):
//   bf16 16x16x32: A[row = l%16][k = 8*(l/16) .. +7], B[k = 8*(l/16) .. +7][col = l%16]
//   fp8  16x16x32: the same with one byte per k
//   f32  16x16x4 : A[row = l%16][k = l/16],           B[k = l/16][col = l%16]
//   f32 result   : D[row = 4*(l/16) + v][col = l%16]
#include "gemm_common.hpp"
#include <cstdlib>

namespace palial {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef long l2 __attribute__((ext_vector_type(2)));

__device__ inline unsigned short to_bf16(double d)
{
    // FP64 -> FP32 (nearest) -> BF16 (nearest even); inputs are finite
    const unsigned int u = __float_as_uint((float)d);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// BF16 (nearest even) of an FP32 value, and the split x = hi + lo with both
// terms BF16 (lo = the BF16 of the exact FP32 remainder x - hi)
__device__ inline unsigned short bf16_bits(float f)
{
    const unsigned int u = __float_as_uint(f);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline void bf16_split(float x, unsigned short &hi, unsigned short &lo)
{
    hi = bf16_bits(x);
    lo = bf16_bits(x - __uint_as_float((unsigned int)hi << 16));
}

// Output storage: FP64 tiles, or (F32C) FP32 tiles at the same element
// offsets from the matrix base (the FP32 working copy of the mixed-precision
// solves).  Task pointers are formed on the FP64 view of the base, so the
// element offset of a pointer is its distance from p.mat.
template <bool F32C> struct Store {
    const double *base;
    __device__ double ld(const double *ptr, size_t i) const
    {
        if (F32C) return (double)((const float *)(const void *)base)[(size_t)(ptr - base) + i];
        return ptr[i];
    }
    __device__ void st(double *ptr, size_t i, double v) const
    {
        if (F32C) ((float *)(void *)const_cast<double *>(base))[(size_t)(ptr - base) + i] = (float)v;
        else ptr[i] = v;
    }
};

// C = alpha*acc + beta*D with acc[x][c][v] = C(i0 + wm*WM + 16x + l%16,
// j0 + wn*WM + 16c + 4(l/16) + v).  D (C itself in the trailing updates) is
// read for a whole fragment row before any store, so those loads are in
// flight together instead of one dependent round trip each.
template <int FM, int WM, bool F32C>
__device__ inline void epilogue(const TaskPtrs &tp, const f4 (&acc)[FM][FM], int i0, int j0, int wm, int wn,
                                int lane, int b, double alpha, double beta, const double *base)
{
    const Store<F32C> S{base};
#pragma unroll
    for (int x = 0; x < FM; x++) {
        const int i = i0 + wm * WM + x * 16 + (lane & 15);
        double dv[FM][4];
#pragma unroll
        for (int c = 0; c < FM; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int j = j0 + wn * WM + c * 16 + 4 * (lane >> 4) + v;
                dv[c][v] = (tp.D && !(tp.diag && i > j)) ? S.ld(tp.D, (size_t)i + (size_t)j * b) : 0.0;
            }
#pragma unroll
        for (int c = 0; c < FM; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int j = j0 + wn * WM + c * 16 + 4 * (lane >> 4) + v;
                if (tp.diag && i > j) continue;
                const double r = alpha * (double)acc[x][c][v];
                S.st(tp.C, (size_t)i + (size_t)j * b, tp.D ? r + beta * dv[c][v] : r);
            }
    }
}

// FP8 panels (precision 3, E4M3): packed tile t is a slot of fp8_slot(b)
// bytes at t * fp8_slot(b): b*b bytes in the K-contiguous layout of the other
// tiers, then the maxima |FP32(x)| of its 64 x 64 blocks (floats, 256 bytes),
// then the tile's scale exponent e (int).  The stored values are
// FP8(FP32(x) * 2^e) with e = 7 - ex, mx = f * 2^ex, f in [0.5, 1), mx the
// tile's largest magnitude: scaled magnitudes stay below 128, inside the E4M3
// range.  An update divides the product of two tiles by 2^(e_a + e_b), which
// is exact.
__host__ __device__ constexpr size_t fp8_slot(int b) { return (size_t)b * b + 512; }

__device__ inline void ld4f(const float *p, float (&v)[4])
{
    const float4 f = *(const float4 *)p;
    v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w;
}
__device__ inline void ld4f(const double *p, float (&v)[4])
{
    const double2 a = *(const double2 *)p, c = *(const double2 *)(p + 2);
    v[0] = (float)a.x; v[1] = (float)a.y; v[2] = (float)c.x; v[3] = (float)c.y;
}

// One workgroup per 64 x 64 block of one panel tile (the grid of pack_kernel):
// max |FP32(x)| of the block into its tile's slot.  NaNs are skipped.
template <typename ST>
__global__ __launch_bounds__(256) void block_max_kernel(const ST *mat, int b, int MT, int k, int first, int colp,
                                                        unsigned char *dst)
{
    const int nsb = b / 64, per = nsb * nsb;
    const int tix = blockIdx.x / per, sbk = blockIdx.x % per;
    const int r0 = (sbk % nsb) * 64, c0 = (sbk / nsb) * 64;
    const int idx = first + tix;
    const ST *src = colp ? mat + tile_off(b, MT, idx, k) : mat + tile_off(b, MT, k, idx);
    const int t = threadIdx.x;
    float m = 0.f;
#pragma unroll
    for (int pass = 0; pass < 4; pass++) {
        const int e = t + pass * 256;                  // 16 runs of 4 rows per column, 64 columns
        float v[4];
        ld4f(src + (size_t)(r0 + (e % 16) * 4) + (size_t)(c0 + e / 16) * b, v);
#pragma unroll
        for (int q = 0; q < 4; q++) m = fmaxf(m, fabsf(v[q]));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    __shared__ float sm[4];
    if ((t & 63) == 0) sm[t >> 6] = m;
    __syncthreads();
    if (t == 0) {
        float *bm = (float *)(dst + (size_t)(idx - k - 1) * fp8_slot(b) + (size_t)b * b);
        bm[sbk] = fmaxf(fmaxf(sm[0], sm[1]), fmaxf(sm[2], sm[3]));
    }
}

// scale exponent of packed FP8 tile `slot` from its block maxima
__device__ inline int fp8_exponent(const unsigned char *slot, int b)
{
    const float *bm = (const float *)(slot + (size_t)b * b);
    const int per = (b / 64) * (b / 64);
    float mx = 0.f;
    for (int i = 0; i < per; i++) mx = fmaxf(mx, bm[i]);
    int ex = 0;
    if (mx > 0.f && mx <= 3.402823466e38f) frexpf(mx, &ex);
    return mx > 0.f && mx <= 3.402823466e38f ? 7 - ex : 0;
}

__device__ inline int fp8x4(float a, float b, float c, float d)
{
    int v = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    return __builtin_amdgcn_cvt_pk_fp8_f32(c, d, v, true);
}

// One workgroup per 64 x 64 sub-block of one panel tile.
template <int PREC, typename ST>
__global__ __launch_bounds__(256) void pack_kernel(const ST *mat, int b, int MT, int k, int first, int colp,
                                                   void *dst)
{
    __shared__ float sT[64][65];
    const int nsb = b / 64, per = nsb * nsb;
    const int tix = blockIdx.x / per, sbk = blockIdx.x % per;
    const int r0 = (sbk % nsb) * 64, c0 = (sbk / nsb) * 64;
    const int idx = first + tix;
    const ST *src = colp ? mat + tile_off(b, MT, idx, k) : mat + tile_off(b, MT, k, idx);
    // PREC 4 (split BF16): tile t holds hi at t*2b^2 and lo at t*2b^2 + b^2 (BF16 elements);
    // PREC 3 (FP8): dbase counts bytes (slot t at t*fp8_slot(b))
    const size_t dbase = PREC == 3 ? (size_t)(idx - k - 1) * fp8_slot(b) : (size_t)(idx - k - 1) * b * b * (PREC == 4 ? 2 : 1);
    const int t = threadIdx.x;
    int e8 = 0;
    if constexpr (PREC == 3) {
        unsigned char *slot = (unsigned char *)dst + dbase;
        e8 = fp8_exponent(slot, b);
        if (sbk == 0 && t == 0) *(int *)(slot + (size_t)b * b + 256) = e8;   // read by the update
    }
    if (!colp) {
        // same layout: 4 consecutive rows of one column per thread and pass
#pragma unroll
        for (int pass = 0; pass < 4; pass++) {
            const int e = t + pass * 256;
            const int r = (e % 16) * 4, c = e / 16;
            const size_t o = (size_t)(r0 + r) + (size_t)(c0 + c) * b;
            double2 x, y;
            if constexpr (sizeof(ST) == 4) {
                const float4 f = *(const float4 *)(src + o);
                x = make_double2(f.x, f.y);
                y = make_double2(f.z, f.w);
            } else {
                x = *(const double2 *)(src + o);
                y = *(const double2 *)(src + o + 2);
            }
            if (PREC == 2) {
                ushort4 v;
                v.x = to_bf16(x.x); v.y = to_bf16(x.y); v.z = to_bf16(y.x); v.w = to_bf16(y.y);
                *(ushort4 *)((unsigned short *)dst + dbase + o) = v;
            } else if (PREC == 4) {
                ushort4 h, l;
                bf16_split((float)x.x, h.x, l.x); bf16_split((float)x.y, h.y, l.y);
                bf16_split((float)y.x, h.z, l.z); bf16_split((float)y.y, h.w, l.w);
                *(ushort4 *)((unsigned short *)dst + dbase + o) = h;
                *(ushort4 *)((unsigned short *)dst + dbase + (size_t)b * b + o) = l;
            } else if (PREC == 3) {
                *(int *)((unsigned char *)dst + dbase + o) =
                    fp8x4(ldexpf((float)x.x, e8), ldexpf((float)x.y, e8), ldexpf((float)y.x, e8), ldexpf((float)y.y, e8));
            } else {
                *(float4 *)((float *)dst + dbase + o) = make_float4((float)x.x, (float)x.y, (float)y.x, (float)y.y);
            }
        }
    } else {
        // transpose through LDS: read along rows r, write along q
#pragma unroll
        for (int pass = 0; pass < 16; pass++) {
            const int e = t + pass * 256;
            const int r = e % 64, q = e / 64;
            sT[q][r] = (float)src[(size_t)(r0 + r) + (size_t)(c0 + q) * b];
        }
        __syncthreads();
#pragma unroll
        for (int pass = 0; pass < 16; pass++) {
            const int e = t + pass * 256;
            const int q = e % 64, r = e / 64;
            const size_t o = dbase + (size_t)(c0 + q) + (size_t)(r0 + r) * b;
            // BF16 of the FP32 rounding, as to_bf16() does
            if (PREC == 2) {
                ((unsigned short *)dst)[o] = bf16_bits(sT[q][r]);
            } else if (PREC == 4) {
                unsigned short h, l;
                bf16_split(sT[q][r], h, l);
                ((unsigned short *)dst)[o] = h;
                ((unsigned short *)dst)[o + (size_t)b * b] = l;
            } else if (PREC == 3) {
                ((unsigned char *)dst)[o] = (unsigned char)(__builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(sT[q][r], e8), 0.f, 0, false) & 0xff);
            } else {
                ((float *)dst)[o] = sT[q][r];
            }
        }
    }
}

// 256 threads, 128 x 128 output block, 2 x 2 waves of 64 x 64, K slices of
// KB.  The first PF fragment rows of D are loaded before the K loop and the
// rest one row ahead in the epilogue, so the HBM latency of the C tile hides
// under the MFMA work instead of following it.
template <int KB, int MINB, int PF>
__global__ __launch_bounds__(256, MINB) void gemm_bf16p_kernel(GemmArgs p)
{
    constexpr int BM = 128, WM = 64, FM = 4;
    constexpr int LDK = KB + 8;           // bf16 per LDS row ([mn][k], k contiguous)
    constexpr int BUF = BM * LDK;
    constexpr int CH = KB / 16;           // 16-byte chunks (8 bf16) per thread and operand: KB/2 k values
    __shared__ unsigned short sA[2 * BUF];
    __shared__ unsigned short sB[2 * BUF];

    const int b = p.b;
    const int nbd = b / BM;
    const int blk_per_task = nbd * nbd;
    const int lid = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int task = lid / blk_per_task, blk = lid % blk_per_task;
    const int bi = blk % nbd, bj = blk / nbd;
    if (task >= p.ntasks) return;
    const TaskPtrs tp = decode(p, task);
    if (!tp.valid) return;
    if (tp.diag && bi > bj) return;
    task_begin(p.ts, p.ts_base + task);
    const int i0 = bi * BM, j0 = bj * BM;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;
    const unsigned short *PA = (const unsigned short *)p.pa + (size_t)tp.ia * b * b;
    const unsigned short *PB = (const unsigned short *)p.pb + (size_t)tp.ib * b * b;

    // acc[x][c][v] = C(i0 + wm*64 + 16x + l%16, j0 + wn*64 + 16c + 4(l/16) + v)
    auto cidx = [&](int x, int c, int v) -> size_t {
        return (size_t)(i0 + wm * WM + x * 16 + (lane & 15)) + (size_t)(j0 + wn * WM + c * 16 + 4 * (lane >> 4) + v) * b;
    };
    auto live = [&](int x, int c, int v) {
        return !(tp.diag && i0 + wm * WM + x * 16 + (lane & 15) > j0 + wn * WM + c * 16 + 4 * (lane >> 4) + v);
    };
    auto ldD = [&](int x, double (&dv)[FM][4]) {
#pragma unroll
        for (int c = 0; c < FM; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) dv[c][v] = (tp.D && live(x, c, v)) ? tp.D[cidx(x, c, v)] : 0.0;
    };

    f4 acc[FM][FM];
#pragma unroll
    for (int a = 0; a < FM; a++)
#pragma unroll
        for (int c = 0; c < FM; c++) acc[a][c] = (f4){0.f, 0.f, 0.f, 0.f};

    // each thread stages KB/2 consecutive k of one row of each operand
    const int row = t >> 1, kc = (t & 1) * (KB / 2);
    uint4 ra[CH], rb[CH];
    auto load = [&](int kb) {
        const uint4 *a = (const uint4 *)(PA + (size_t)(i0 + row) * b + kb + kc);
        const uint4 *bb = (const uint4 *)(PB + (size_t)(j0 + row) * b + kb + kc);
#pragma unroll
        for (int q = 0; q < CH; q++) { ra[q] = a[q]; rb[q] = bb[q]; }
    };
    auto store = [&](int buf) {
        uint4 *a = (uint4 *)(sA + buf * BUF + row * LDK + kc);
        uint4 *bb = (uint4 *)(sB + buf * BUF + row * LDK + kc);
#pragma unroll
        for (int q = 0; q < CH; q++) { a[q] = ra[q]; bb[q] = rb[q]; }
    };

    load(0);
    double d0[FM][4], d1[FM][4], d2[FM][4];
    if (PF >= 1) ldD(0, d0);
    if (PF >= 2) ldD(1, d1);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int kb = 0; kb < b; kb += KB) {
        const bool more = kb + KB < b;
        if (more) load(kb + KB);
        const unsigned short *a = sA + buf * BUF, *bb = sB + buf * BUF;
#pragma unroll
        for (int ks = 0; ks < KB; ks += 32) {
            const int ko = ks + 8 * (lane >> 4);
            bf8 av[FM], bv[FM];
#pragma unroll
            for (int x = 0; x < FM; x++) av[x] = *(const bf8 *)(a + (wm * WM + x * 16 + (lane & 15)) * LDK + ko);
#pragma unroll
            for (int c = 0; c < FM; c++) bv[c] = *(const bf8 *)(bb + (wn * WM + c * 16 + (lane & 15)) * LDK + ko);
#pragma unroll
            for (int x = 0; x < FM; x++)
#pragma unroll
                for (int c = 0; c < FM; c++)
                    acc[x][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[c], av[x], acc[x][c], 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }

    const double alpha = p.alpha, beta = p.beta;
    auto stC = [&](int x, const double (&dv)[FM][4]) {
#pragma unroll
        for (int c = 0; c < FM; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                if (!live(x, c, v)) continue;
                const double r = alpha * (double)acc[x][c][v];
                tp.C[cidx(x, c, v)] = tp.D ? r + beta * dv[c][v] : r;
            }
    };
    // rows 0..3 with a rolling prefetch of the next row's D (FM == 4)
    if (PF < 1) ldD(0, d0);
    if (PF < 2) ldD(1, d1);
    ldD(2, d2);
    stC(0, d0);
    ldD(3, d0);
    stC(1, d1);
    stC(2, d2);
    stC(3, d0);
    task_end(p.ts, p.ts_base + task);
}

// BF16 update with asynchronous global -> LDS loads (global_load_lds_dwordx4,
// gfx950): NST LDS stages of KB = 32 K values, up to NST-1 slices in flight
// per workgroup, one barrier per slice.  A slice of one operand is 128 rows x
// 64 B; each wave's DMA instruction fills 1 KB of LDS (16 rows, lane l at byte
// 16*l), and lane l fetches row 16j + l/4, 16-byte chunk (l%4) ^ ((row/4)%4)
// of that row: the XOR swizzle makes each MFMA fragment read (16 rows x one
// chunk per 16 lanes) hit 16 distinct 16-byte bank groups.
// Waits are explicit (vmcnt counts this wave's DMA instructions, 4 per
// slice) and the barriers carry no fence, so loads of later slices stay in
// flight across them.
typedef __attribute__((address_space(3))) void lds_void_t;

__device__ inline void wait_vm(int n)
{
    // s_waitcnt vmcnt(n) (expcnt and lgkmcnt left unconstrained); n in {0, 4, 8, 16}
    // (vmcnt bits [3:0] and [15:14])
    if (n >= 16) __builtin_amdgcn_s_waitcnt(0x4F70);
    else if (n >= 8) __builtin_amdgcn_s_waitcnt(0x0F78);
    else if (n >= 4) __builtin_amdgcn_s_waitcnt(0x0F74);
    else __builtin_amdgcn_s_waitcnt(0x0F70);
}

// X3 (split BF16 panels, tile t = hi at 2t*b^2, lo at (2t+1)*b^2 elements):
// each stage also holds the lo slices and acc += B_hi A_hi + B_lo A_hi + B_hi A_lo,
// products good to ~2^-16 instead of 2^-8 at three times the MFMA work.
// NST stages, NST-1 of them in flight; LDS = NST * (2 or 4) * 8 KB.
// F8 (FP8 E4M3 panels in fp8_slot(b) slots): a slice row holds 64 K values
// in the same 64 bytes; of a lane's 16-byte chunk (K values 16 (l/16) .. +15
// of the slice) the low 8 bytes feed one v_mfma_f32_16x16x32_fp8_fp8 and the
// high 8 bytes a second, so A and B pair the same K values.  alpha is scaled
// by 2^-(e_a + e_b).
// TWO: paired steps (pa2 / pb2 set); without it the K loop covers one panel and
// the panel-switch arithmetic is compiled out.
template <int MINB, bool F32C, bool X3 = false, int NST = 4, bool F8 = false, bool TWO = false>
__global__ __launch_bounds__(256, MINB) void gemm_bf16d_kernel(GemmArgs p)
{
    const Store<F32C> S{p.mat};
    constexpr int BM = 128, KB = 32, WM = 64, FM = 4, AHEAD = NST - 1;
    constexpr int NOPS = X3 ? 4 : 2;                    // operand slices per stage (A, B[, A lo, B lo])
    constexpr int DPS = 2 * NOPS;                       // DMA instructions per wave and stage
    constexpr int SLICE = BM * KB * 2;                  // bytes of one operand slice
    __shared__ __attribute__((aligned(1024))) char lds[NST][NOPS][SLICE];

    const int b = p.b;
    const int nbd = b / BM;
    const int blk_per_task = nbd * nbd;
    const int lid = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int task = lid / blk_per_task, blk = lid % blk_per_task;
    const int bi = blk % nbd, bj = blk / nbd;
    if (task >= p.ntasks) return;
    const TaskPtrs tp = decode(p, task);
    if (!tp.valid) return;
    if (tp.diag && bi > bj) return;
    task_begin(p.ts, p.ts_base + task);
    const int i0 = bi * BM, j0 = bj * BM;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;
    // byte addressing: a panel row is rs bytes, a slice 64 bytes of it
    const size_t tstride = F8 ? fp8_slot(b) : (size_t)b * b * (X3 ? 4 : 2);
    const char *PA_ = (const char *)p.pa + (size_t)tp.ia * tstride;
    const char *PB_ = (const char *)p.pb + (size_t)tp.ib * tstride;
    // paired steps (GM_CHOL_REST2, GM_LU_REST2): the K loop runs on through panel k+1, whose
    // tiles sit one position lower
    const char *PA2 = TWO ? (const char *)p.pa2 + (size_t)(tp.ia - 1) * tstride : nullptr;
    const char *PB2 = TWO ? (const char *)p.pb2 + (size_t)(tp.ib - 1) * tstride : nullptr;
    const size_t rs = F8 ? (size_t)b : 2 * (size_t)b;
    const int NS1 = F8 ? b / 64 : b / KB;               // slices of one panel
    const int NS = TWO ? 2 * NS1 : NS1;

    auto issue = [&](int s) {
        const int st = s % NST;
        const bool second = TWO && s >= NS1;
        const char *PA = second ? PA2 : PA_, *PB = second ? PB2 : PB_;
        const size_t kb = (size_t)(second ? s - NS1 : s) * 64;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int j = w * 2 + q;                    // 1 KB chunk of the slice
            const int row = 16 * j + (lane >> 2);
            const int c = (lane & 3) ^ ((row >> 2) & 3);
            __builtin_amdgcn_global_load_lds((const void *)(PA + (size_t)(i0 + row) * rs + kb + c * 16),
                                             (lds_void_t *)(lds[st][0] + j * 1024), 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void *)(PB + (size_t)(j0 + row) * rs + kb + c * 16),
                                             (lds_void_t *)(lds[st][1] + j * 1024), 16, 0, 0);
            if (X3) {
                const size_t lo = 2 * (size_t)b * b;    // bytes from hi to lo
                __builtin_amdgcn_global_load_lds((const void *)(PA + lo + (size_t)(i0 + row) * rs + kb + c * 16),
                                                 (lds_void_t *)(lds[st][NOPS - 2] + j * 1024), 16, 0, 0);
                __builtin_amdgcn_global_load_lds((const void *)(PB + lo + (size_t)(j0 + row) * rs + kb + c * 16),
                                                 (lds_void_t *)(lds[st][NOPS - 1] + j * 1024), 16, 0, 0);
            }
        }
    };
    auto frag = [&](const char *base, int row) -> bf8 {
        const int c = lane >> 4;
        return *(const bf8 *)(base + (row * 4 + (c ^ ((row >> 2) & 3))) * 16);
    };

    f4 acc[FM][FM];
#pragma unroll
    for (int a = 0; a < FM; a++)
#pragma unroll
        for (int c = 0; c < FM; c++) acc[a][c] = (f4){0.f, 0.f, 0.f, 0.f};

#pragma unroll
    for (int q = 0; q < AHEAD; q++)
        if (q < NS) issue(q);
    for (int s = 0; s < NS; s++) {
        // slice s has landed (this wave's part): stages s+1 .. s+AHEAD-1 may still be in flight
        const int later = min(AHEAD - 1, NS - 1 - s);
        wait_vm(later * DPS);
        __builtin_amdgcn_s_barrier();                     // ... and every wave's part
        asm volatile("" ::: "memory");
        if (s + AHEAD < NS) issue(s + AHEAD);             // into the stage read at s-1 (all waves are past it)
        const char *A = lds[s % NST][0], *B = lds[s % NST][1];
        bf8 av[FM], bv[FM];
#pragma unroll
        for (int x = 0; x < FM; x++) av[x] = frag(A, wm * WM + x * 16 + (lane & 15));
#pragma unroll
        for (int c = 0; c < FM; c++) bv[c] = frag(B, wn * WM + c * 16 + (lane & 15));
        if constexpr (X3) {
            const char *Al = lds[s % NST][NOPS - 2], *Bl = lds[s % NST][NOPS - 1];
            bf8 al[FM], bl[FM];
#pragma unroll
            for (int x = 0; x < FM; x++) al[x] = frag(Al, wm * WM + x * 16 + (lane & 15));
#pragma unroll
            for (int c = 0; c < FM; c++) bl[c] = frag(Bl, wn * WM + c * 16 + (lane & 15));
#pragma unroll
            for (int x = 0; x < FM; x++)
#pragma unroll
                for (int c = 0; c < FM; c++) {
                    acc[x][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[c], av[x], acc[x][c], 0, 0, 0);
                    acc[x][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[c], al[x], acc[x][c], 0, 0, 0);
                }
        }
        if constexpr (F8) {
#pragma unroll
            for (int x = 0; x < FM; x++)
#pragma unroll
                for (int c = 0; c < FM; c++) {
                    const l2 a2 = __builtin_bit_cast(l2, av[x]), b2 = __builtin_bit_cast(l2, bv[c]);
                    acc[x][c] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(b2[0], a2[0], acc[x][c], 0, 0, 0);
                    acc[x][c] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(b2[1], a2[1], acc[x][c], 0, 0, 0);
                }
        } else {
#pragma unroll
            for (int x = 0; x < FM; x++)
#pragma unroll
                for (int c = 0; c < FM; c++)
                    acc[x][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[c], av[x], acc[x][c], 0, 0, 0);
        }
        asm volatile("" ::: "memory");
    }
    double alpha_d = p.alpha;
    if constexpr (F8) {
        const int ea = *(const int *)(PA_ + (size_t)b * b + 256), eb = *(const int *)(PB_ + (size_t)b * b + 256);
        alpha_d = ldexp(p.alpha, -(ea + eb));
    }

    if constexpr (F32C) {
        // FP32 tiles: 16 lanes of one accumulator column would cover only 64
        // bytes, so each wave stages half of its 64 x 64 block (32 rows) in
        // LDS, column by column, and moves it as float4 runs: 8 lanes write one
        // full 128-byte line (32 rows of a column), one instruction 8 columns.
        // Lower-triangle elements of a diagonal (SYRK) tile keep their value.
        constexpr int LDC = 36;                          // floats per staged column (banks spread)
        __syncthreads();                                 // all waves are done with the operand stages
        float *sc = (float *)(void *)&lds[0][0][0] + w * 64 * LDC;
        float *Cf = (float *)(void *)const_cast<double *>(p.mat) + (size_t)(tp.C - p.mat);
        const float *Df = tp.D ? (const float *)(const void *)p.mat + (size_t)(tp.D - p.mat) : nullptr;
        const float alpha = (float)alpha_d, beta = (float)p.beta;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int r0 = i0 + wm * WM + 32 * h + 4 * (lane & 7);      // first of this lane's 4 rows
            float4 dv[8];
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int col = j0 + wn * WM + 8 * it + (lane >> 3);
                dv[it] = Df ? *(const float4 *)(Df + (size_t)r0 + (size_t)col * b) : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int c = 0; c < FM; c++)
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        sc[(c * 16 + 4 * (lane >> 4) + v) * LDC + x * 16 + (lane & 15)] = acc[2 * h + x][c][v];
            __syncthreads();
#pragma unroll
            for (int it = 0; it < 8; it++) {
                const int cl = 8 * it + (lane >> 3), col = j0 + wn * WM + cl;
                const float4 a = *(const float4 *)(sc + cl * LDC + 4 * (lane & 7));
                float4 o;
                o.x = alpha * a.x + beta * dv[it].x;
                o.y = alpha * a.y + beta * dv[it].y;
                o.z = alpha * a.z + beta * dv[it].z;
                o.w = alpha * a.w + beta * dv[it].w;
                if (tp.diag) {
                    if (r0 > col) o.x = dv[it].x;
                    if (r0 + 1 > col) o.y = dv[it].y;
                    if (r0 + 2 > col) o.z = dv[it].z;
                    if (r0 + 3 > col) o.w = dv[it].w;
                }
                *(float4 *)(Cf + (size_t)r0 + (size_t)col * b) = o;
            }
            __syncthreads();
        }
        task_end(p.ts, p.ts_base + task);
        return;
    }
    // acc[x][c][v] = C(i0 + wm*64 + 16x + l%16, j0 + wn*64 + 16c + 4(l/16) + v);
    // D rows are loaded one fragment row ahead of the stores
    auto cidx = [&](int x, int c, int v) -> size_t {
        return (size_t)(i0 + wm * WM + x * 16 + (lane & 15)) + (size_t)(j0 + wn * WM + c * 16 + 4 * (lane >> 4) + v) * b;
    };
    auto live = [&](int x, int c, int v) {
        return !(tp.diag && i0 + wm * WM + x * 16 + (lane & 15) > j0 + wn * WM + c * 16 + 4 * (lane >> 4) + v);
    };
    auto ldD = [&](int x, double (&dv)[FM][4]) {
#pragma unroll
        for (int c = 0; c < FM; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) dv[c][v] = (tp.D && live(x, c, v)) ? S.ld(tp.D, cidx(x, c, v)) : 0.0;
    };
    const double alpha = alpha_d, beta = p.beta;
    auto stC = [&](int x, const double (&dv)[FM][4]) {
#pragma unroll
        for (int c = 0; c < FM; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                if (!live(x, c, v)) continue;
                const double r = alpha * (double)acc[x][c][v];
                S.st(tp.C, cidx(x, c, v), tp.D ? r + beta * dv[c][v] : r);
            }
    };
    double d0[FM][4], d1[FM][4];
    ldD(0, d0);
    ldD(1, d1);
    stC(0, d0);
    ldD(2, d0);
    stC(1, d1);
    ldD(3, d1);
    stC(2, d0);
    stC(3, d1);
    task_end(p.ts, p.ts_base + task);
}

// FP32 operands, v_mfma_f32_16x16x4_f32; same blocking with KB = 32.
template <bool F32C>
__global__ __launch_bounds__(256, 2) void gemm_f32p_kernel(GemmArgs p)
{
    constexpr int BM = 128, KB = 32, WM = 64, FM = 4;
    constexpr int LDK = KB + 4;           // floats; 16 rows x 4 k of a fragment hit distinct banks
    constexpr int BUF = BM * LDK;
    __shared__ float sA[2 * BUF];
    __shared__ float sB[2 * BUF];

    const int b = p.b;
    const int nbd = b / BM;
    const int blk_per_task = nbd * nbd;
    const int lid = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int task = lid / blk_per_task, blk = lid % blk_per_task;
    const int bi = blk % nbd, bj = blk / nbd;
    if (task >= p.ntasks) return;
    const TaskPtrs tp = decode(p, task);
    if (!tp.valid) return;
    if (tp.diag && bi > bj) return;
    task_begin(p.ts, p.ts_base + task);
    const int i0 = bi * BM, j0 = bj * BM;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;
    const float *PA = (const float *)p.pa + (size_t)tp.ia * b * b;
    const float *PB = (const float *)p.pb + (size_t)tp.ib * b * b;

    f4 acc[FM][FM];
#pragma unroll
    for (int a = 0; a < FM; a++)
#pragma unroll
        for (int c = 0; c < FM; c++) acc[a][c] = (f4){0.f, 0.f, 0.f, 0.f};

    const int row = t >> 1, kc = (t & 1) * 16;
    float4 ra[4], rb[4];
    auto load = [&](int kb) {
        const float4 *a = (const float4 *)(PA + (size_t)(i0 + row) * b + kb + kc);
        const float4 *bb = (const float4 *)(PB + (size_t)(j0 + row) * b + kb + kc);
#pragma unroll
        for (int q = 0; q < 4; q++) { ra[q] = a[q]; rb[q] = bb[q]; }
    };
    auto store = [&](int buf) {
        float4 *a = (float4 *)(sA + buf * BUF + row * LDK + kc);
        float4 *bb = (float4 *)(sB + buf * BUF + row * LDK + kc);
#pragma unroll
        for (int q = 0; q < 4; q++) { a[q] = ra[q]; bb[q] = rb[q]; }
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
        for (int ks = 0; ks < KB; ks += 4) {
            const int ko = ks + (lane >> 4);
            float av[FM], bv[FM];
#pragma unroll
            for (int x = 0; x < FM; x++) av[x] = a[(wm * WM + x * 16 + (lane & 15)) * LDK + ko];
#pragma unroll
            for (int c = 0; c < FM; c++) bv[c] = bb[(wn * WM + c * 16 + (lane & 15)) * LDK + ko];
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

    epilogue<FM, WM, F32C>(tp, acc, i0, j0, wm, wn, lane, b, p.alpha, p.beta, p.mat);
    task_end(p.ts, p.ts_base + task);
}

}  // namespace

hipError_t launch_pack(const double *mat, int N, int b, int k, int first, int n, bool colpanel, int prec,
                       void *dst, hipStream_t s, bool src32)
{
    if (n <= 0) return hipSuccess;
    if (sched_rec_on()) { rec_pack(mat, b, k, first, n, colpanel, dst, s); return hipSuccess; }
    if (b % 64 != 0 || first <= k || first + n > N / b || (prec != 1 && prec != 2 && prec != 3 && prec != 4))
        return hipErrorInvalidValue;
    const int nsb = b / 64, cp = colpanel ? 1 : 0;
    dim3 grid((unsigned)(n * nsb * nsb)), block(256);
    const float *m32 = (const float *)(const void *)mat;
    if (prec == 3) {
        // block maxima first (same grid), then the scaled FP8 rounding
        unsigned char *d8 = (unsigned char *)dst;
        if (src32) hipLaunchKernelGGL((block_max_kernel<float>), grid, block, 0, s, m32, b, N / b, k, first, cp, d8);
        else hipLaunchKernelGGL((block_max_kernel<double>), grid, block, 0, s, mat, b, N / b, k, first, cp, d8);
        if (src32) hipLaunchKernelGGL((pack_kernel<3, float>), grid, block, 0, s, m32, b, N / b, k, first, cp, dst);
        else hipLaunchKernelGGL((pack_kernel<3, double>), grid, block, 0, s, mat, b, N / b, k, first, cp, dst);
    }
    else if (prec == 4 && src32) hipLaunchKernelGGL((pack_kernel<4, float>), grid, block, 0, s, m32, b, N / b, k, first, cp, dst);
    else if (prec == 4) hipLaunchKernelGGL((pack_kernel<4, double>), grid, block, 0, s, mat, b, N / b, k, first, cp, dst);
    else if (prec == 2 && src32) hipLaunchKernelGGL((pack_kernel<2, float>), grid, block, 0, s, m32, b, N / b, k, first, cp, dst);
    else if (prec == 2) hipLaunchKernelGGL((pack_kernel<2, double>), grid, block, 0, s, mat, b, N / b, k, first, cp, dst);
    else if (src32) hipLaunchKernelGGL((pack_kernel<1, float>), grid, block, 0, s, m32, b, N / b, k, first, cp, dst);
    else hipLaunchKernelGGL((pack_kernel<1, double>), grid, block, 0, s, mat, b, N / b, k, first, cp, dst);
    return hipGetLastError();
}

hipError_t launch_gemm_packed(const GemmArgs &a_in, int prec, hipStream_t s)
{
    if (a_in.ntasks <= 0) return hipSuccess;
    GemmArgs a = a_in;
    if (sched_rec_on()) { rec_gemm(a, true, s); return hipSuccess; }
    trace_gemm(a);
    if (a.b % 128 != 0 || a.maskA != MK_NONE || !a.pa || !a.pb) return hipErrorInvalidValue;
    if (a.mode != GM_CHOL_TRAIL && a.mode != GM_CHOL_REST && a.mode != GM_CHOL_ROW && a.mode != GM_LU_TRAIL &&
        a.mode != GM_LU_REST && a.mode != GM_LU_PANEL && a.mode != GM_CHOL_FIRST && a.mode != GM_LU_FIRST &&
        a.mode != GM_CHOL_ROW2 && a.mode != GM_CHOL_REST2 && a.mode != GM_LU_ROW2 && a.mode != GM_LU_REST2)
        return hipErrorInvalidValue;
    // two panels per launch: BF16 and split BF16 only, and exactly for the *_REST2 modes
    const bool two = a.mode == GM_CHOL_REST2 || a.mode == GM_LU_REST2;
    if (two != (a.pa2 != nullptr && a.pb2 != nullptr) || (a.pa2 && prec != 2 && prec != 4))
        return hipErrorInvalidValue;
    const long nbd = a.b / 128;
    dim3 grid((unsigned)((long)a.ntasks * nbd * nbd)), block(256);
    if (prec == 3) {
        if (a.f32c) hipLaunchKernelGGL((gemm_bf16d_kernel<2, true, false, 4, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((gemm_bf16d_kernel<2, false, false, 4, true>), grid, block, 0, s, a);
    }
    else if (prec == 4) {
        // two stages (64 KB LDS, two workgroups per CU)
        if (two) {
            if (!a.f32c) return hipErrorInvalidValue;        // paired steps run on the FP32 working copy
            hipLaunchKernelGGL((gemm_bf16d_kernel<2, true, true, 2, false, true>), grid, block, 0, s, a);
        }
        else if (a.f32c) hipLaunchKernelGGL((gemm_bf16d_kernel<2, true, true, 2>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((gemm_bf16d_kernel<2, false, true, 2>), grid, block, 0, s, a);
    }
    else if (prec == 2) {
        static const int var = [] {
            const char *e = getenv("PALIAL_PK_VARIANT");
            return e ? atoi(e) : 0;
        }();
        if (two) {
            // paired steps: the default kernel on the FP32 working copy
            if (var != 0 || !a.f32c) return hipErrorInvalidValue;
            hipLaunchKernelGGL((gemm_bf16d_kernel<2, true, false, 4, false, true>), grid, block, 0, s, a);
        }
        else if (a.f32c) {
            if (var != 0) return hipErrorInvalidValue;       // FP32 output only in the default kernel
            hipLaunchKernelGGL((gemm_bf16d_kernel<2, true>), grid, block, 0, s, a);
        }
        else if (var == 0) hipLaunchKernelGGL((gemm_bf16d_kernel<2, false>), grid, block, 0, s, a);
        else if (var == 1) hipLaunchKernelGGL((gemm_bf16p_kernel<32, 3, 0>), grid, block, 0, s, a);
        else if (var == 2) hipLaunchKernelGGL((gemm_bf16p_kernel<64, 2, 0>), grid, block, 0, s, a);
        else if (var == 3) hipLaunchKernelGGL((gemm_bf16p_kernel<64, 2, 2>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((gemm_bf16p_kernel<32, 2, 2>), grid, block, 0, s, a);   // var 4
    }
    else if (prec == 1 && a.f32c) hipLaunchKernelGGL((gemm_f32p_kernel<true>), grid, block, 0, s, a);
    else if (prec == 1) hipLaunchKernelGGL((gemm_f32p_kernel<false>), grid, block, 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace palial
