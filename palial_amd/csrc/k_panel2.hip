// k_panel2.hip -- MFMA panel kernels for tiles with b % 32 == 0 and b <= 512.
//
//   potrf2_kernel : upper Cholesky of one tile (LAPACKE_dpotrf('U'),
//                   src/cholesky.h:40) in 32-column blocks.  For each block:
//                   the 32x32 diagonal block is factored by one wave (one
//                   column per lane, readlane broadcasts), inverted (U_D^-1, kept in LDS
//                   and written to `dinv` for the panel TRSM), the 32-row panel
//                   is solved as U_D^-T P with v_mfma_f64_16x16x4_f64, and the
//                   rank-32 trailing SYRK runs out of the LDS copy of P.
//   getrf2_kernel : LU without pivoting of one tile (PALIAL_dgetrf,
//                   src/srclu/dgetrf.h:11), same blocking; writes U_D^-1 and
//                   L_D^-1 of every diagonal block.  A zero or sub-sfmin pivot
//                   in a block switches that block's L panel to the
//                   reference's per-element scaling rule
//                   (src/srclu/dgetrf_no_pivot.h:21-39).
//   trsm2_kernel  : the panel TRSMs of one DAG step (src/cholesky.h:73-84,
//                   src/lu.h:67-78, src/lu.h:109-120): each workgroup solves a
//                   32-wide strip of one tile held in LDS by right-looking
//                   blocked substitution; diagonal blocks are applied through
//                   the inverses written by potrf2/getrf2, off-diagonal
//                   updates are MFMA.
#include "internal.hpp"
#include <float.h>
#include <type_traits>
#include <cstdlib>

namespace palial {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int P2_THREADS = 512;           // TRSM strips
constexpr int P2_WAVES = P2_THREADS / 64;
constexpr int PF_THREADS = 512;           // diagonal-tile factorizations
constexpr int PF_WAVES = PF_THREADS / 64;
constexpr int PF_G = 4;                   // POTRF trailing-update tiles in flight per wave
constexpr int GF_G = 2;                   // GETRF trailing-update tiles per job (one row tile)
constexpr int NB2 = 32;            // inner block
constexpr int LDD = NB2 + 1;       // LDS stride of 32x32 blocks

__device__ inline d4 mfma16(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ inline d4 zero4() { return (d4){0.0, 0.0, 0.0, 0.0}; }

// value of v held by lane `l` of the wave (broadcast through SGPRs)
__device__ inline double readlane_d(double v, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// triangular index e -> (bi <= bj), e = bj(bj+1)/2 + bi
__device__ inline void tri_ij(int e, int &bi, int &bj)
{
    int j = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
    while ((j + 1) * (j + 2) / 2 <= e) j++;
    while (j * (j + 1) / 2 > e) j--;
    bj = j;
    bi = e - j * (j + 1) / 2;
}

// Parallel inverse of a 32x32 triangular block by recursive doubling, called
// by all threads of the workgroup (contains barriers).  X (ld LDD) receives
// the inverse, sT (32*32) is scratch.  At level s every 2s-block
//   upper: [[X11, X12], [0, X22]],  X12 = -X11 U12 X22
//   unit lower: [[X11, 0], [X21, X22]],  X21 = -X22 L21 X11
// is completed from its two already inverted s-blocks.
template <bool UPPER>
__device__ void tri_inverse32(const double *M, double *X, double *sT, int nthreads)
{
    const int t = threadIdx.x;
    for (int e = t; e < NB2 * NB2; e += nthreads) {
        const int r = e % NB2, c = e / NB2;
        X[r + c * LDD] = (r == c) ? (UPPER ? 1.0 / M[r + r * LDD] : 1.0) : 0.0;
    }
    __syncthreads();
    for (int sz = 1; sz < NB2; sz *= 2) {
        const int nel = (NB2 / (2 * sz)) * sz * sz;
        // T = U12 X22 (upper) or L21 X11 (lower), one element per thread
        for (int e = t; e < nel; e += nthreads) {
            const int pair = e / (sz * sz), ij = e % (sz * sz), i = ij % sz, j = ij / sz;
            const int r0 = pair * 2 * sz;
            double acc = 0.0;
            if (UPPER) {
                for (int p = 0; p <= j; p++) acc += M[(r0 + i) + (r0 + sz + p) * LDD] * X[(r0 + sz + p) + (r0 + sz + j) * LDD];
            } else {
                for (int p = j; p < sz; p++) acc += M[(r0 + sz + i) + (r0 + p) * LDD] * X[(r0 + p) + (r0 + j) * LDD];
            }
            sT[e] = acc;
        }
        __syncthreads();
        for (int e = t; e < nel; e += nthreads) {
            const int pair = e / (sz * sz), ij = e % (sz * sz), i = ij % sz, j = ij / sz;
            const int r0 = pair * 2 * sz;
            const double *T = sT + pair * sz * sz;
            double acc = 0.0;
            if (UPPER) {
                for (int p = i; p < sz; p++) acc += X[(r0 + i) + (r0 + p) * LDD] * T[p + j * sz];
                X[(r0 + i) + (r0 + sz + j) * LDD] = -acc;
            } else {
                for (int p = 0; p <= i; p++) acc += X[(r0 + sz + i) + (r0 + sz + p) * LDD] * T[p + j * sz];
                X[(r0 + sz + i) + (r0 + j) * LDD] = -acc;
            }
        }
        __syncthreads();
    }
}

// Copy the NB x (b - base) row panel A[j0:j0+NB, base:b] into sP, eight
// independent global loads in flight per thread before the LDS stores.
__device__ inline void stage_panel(const double *A, int b, int j0, int base, double *sP, int LDP, int nthreads)
{
    const int total = NB2 * (b - base);
    for (int e0 = threadIdx.x; e0 < total; e0 += nthreads * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = e0 + u * nthreads;
            v[u] = (e < total) ? A[(size_t)(j0 + e % NB2) + (size_t)(base + e / NB2) * b] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = e0 + u * nthreads;
            if (e < total) sP[(e % NB2) * LDP + base + e / NB2] = v[u];
        }
    }
}

// ---------------------------------------------------------------- POTRF
// LDS of the tile factorizations, carved from one dynamic buffer so that a
// kernel running both (the batched one) stays within 160 KB:
//   sP: NB x (b+16) panel rows, then 2 (POTRF) or 3 (GETRF) 32x33 blocks
__host__ __device__ constexpr size_t panel_lds_bytes(int b, int nblk) { return ((size_t)NB2 * (b + 16) + (size_t)nblk * NB2 * LDD) * 8; }

__device__ __forceinline__ void potrf2_tile(double *A, int b, int *info, int info_base, double *dinv, double *smem)
{
    constexpr int NB = NB2;
    const int LDP = b + 16;               // panel rows 128 B apart modulo 256 B
    double *sP = smem;                    // NB x LDP: sP[q*LDP + c], c = absolute column
    double *sU = smem + (size_t)NB * LDP;
    double *sI = sU + NB * LDD;
    __shared__ int s_fail;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) s_fail = 0;
    __syncthreads();

    for (int j0 = 0; j0 < b; j0 += NB) {
        // 1. diagonal block: staged in LDS (upper part, zeros below), factored
        //    by wave 0 in registers, lane c holding column c (dpotf2 order:
        //    sqrt of the pivot, row scaled by its reciprocal, rank-1 update).
        //    Row j of U reaches every lane through v_readlane (uniform SGPR
        //    operands), so the 32 dependent steps need no LDS round trips.
        //    Entries below the diagonal become garbage and are never read.
        for (int e = t; e < NB * NB; e += PF_THREADS) {
            const int r = e % NB, c = e / NB;
            sU[r + c * LDD] = (r <= c) ? A[(size_t)(j0 + r) + (size_t)(j0 + c) * b] : 0.0;
        }
        __syncthreads();
        if (w == 0) {
            const int c = lane & (NB - 1);
            double x[NB];
#pragma unroll
            for (int r = 0; r < NB; r++) x[r] = sU[r + c * LDD];
            int bad = -1;   // wave-uniform; a failed pivot freezes the block (no break:
                            // the loop must unroll fully to keep x[] in registers)
#pragma unroll
            for (int j = 0; j < NB; j++) {
                if (bad < 0) {
                    const double piv = readlane_d(x[j], j);
                    if (!(piv > 0.0)) {
                        bad = j;
                    } else {
                        // 1/sqrt(piv): hardware estimate refined by two Newton steps
                        double rs = __builtin_amdgcn_rsq(piv);
                        rs = rs * (1.5 - 0.5 * piv * rs * rs);
                        rs = rs * (1.5 - 0.5 * piv * rs * rs);
                        x[j] = (c == j) ? piv * rs : x[j] * rs;
                        // row j of U into SGPRs first, then the rank-1 update
                        double u[NB];
#pragma unroll
                        for (int r = j + 1; r < NB; r++) u[r] = readlane_d(x[j], r);
#pragma unroll
                        for (int r = j + 1; r < NB; r++) x[r] -= u[r] * x[j];
                    }
                }
            }
            if (lane < NB) {
#pragma unroll
                for (int r = 0; r < NB; r++)
                    if (r <= c) sU[r + c * LDD] = x[r];
            }
            // the first failure of a factorization is kept (later tiles factor garbage)
            if (bad >= 0 && lane == 0) { s_fail = 1; if (info) atomicCAS(info, 0, info_base + j0 + bad + 1); }
        }
        __syncthreads();
        for (int e = t; e < NB * NB; e += PF_THREADS) {
            const int r = e % NB, c = e / NB;
            if (r <= c) A[(size_t)(j0 + r) + (size_t)(j0 + c) * b] = sU[r + c * LDD];
        }
        if (s_fail) return;
        // 2. U_D^{-1} by recursive doubling (all threads), kept in LDS and in dinv
        tri_inverse32<true>(sU, sI, sP, PF_THREADS);
        if (dinv)
            for (int e = t; e < NB * NB; e += PF_THREADS)
                dinv[(size_t)(j0 / NB) * NB * NB + e] = sI[(e % NB) + (e / NB) * LDD];
        const int base = j0 + NB, rest = b - base;
        // 3. stage the raw row panel A[j0:j0+NB, base:b]
        stage_panel(A, b, j0, base, sP, LDP, PF_THREADS);
        __syncthreads();
        if (rest == 0) break;
        // 4. P <- U_D^{-T} P: a wave owns 16-column tiles (both 16-row halves)
#pragma unroll 1
        for (int ct = w; ct < rest / 16; ct += PF_WAVES) {
            const int col = base + ct * 16 + (lane & 15);
            d4 acc0 = zero4(), acc1 = zero4();
#pragma unroll
            for (int k4 = 0; k4 < NB; k4 += 4) {
                const int q = k4 + (lane >> 4);
                const double bv = sP[q * LDP + col];
                acc0 = mfma16(sI[q + (lane & 15) * LDD], bv, acc0);         // opA(i,q) = Uinv(q,i)
                acc1 = mfma16(sI[q + (16 + (lane & 15)) * LDD], bv, acc1);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = (lane >> 4) + 4 * v;
                sP[i * LDP + col] = acc0[v];
                sP[(i + 16) * LDP + col] = acc1[v];
                A[(size_t)(j0 + i) + (size_t)col * b] = acc0[v];
                A[(size_t)(j0 + i + 16) + (size_t)col * b] = acc1[v];
            }
        }
        __syncthreads();
        // 5. trailing SYRK on the upper triangle of A[base:b, base:b], 16x16 tiles.
        //    Each wave walks groups of PF_G tiles; the C tiles of the next group
        //    are loaded while the MFMAs of the current group run.
        const int nt = rest / 16;
        const int ntri = nt * (nt + 1) / 2;
        const int stride = PF_WAVES * PF_G;
        auto load_group = [&](int e0, double (&cv)[PF_G][4]) {
#pragma unroll
            for (int g = 0; g < PF_G; g++) {
                if (e0 + g < ntri) {
                    int bi, bj;
                    tri_ij(e0 + g, bi, bj);
                    const int i = base + 16 * bi + (lane & 15);
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        cv[g][v] = A[(size_t)i + (size_t)(base + 16 * bj + (lane >> 4) + 4 * v) * b];
                }
            }
        };
        auto run_group = [&](int e0, const double (&cv)[PF_G][4]) {
            int ti[PF_G], tj[PF_G];
            d4 acc[PF_G];
#pragma unroll
            for (int g = 0; g < PF_G; g++) {
                acc[g] = zero4();
                ti[g] = -1;
                tj[g] = 0;
                if (e0 + g < ntri) tri_ij(e0 + g, ti[g], tj[g]);
            }
#pragma unroll
            for (int k4 = 0; k4 < NB; k4 += 4) {
                const int q = k4 + (lane >> 4);
#pragma unroll
                for (int g = 0; g < PF_G; g++)
                    if (ti[g] >= 0)   // operands swapped: D holds the tile transposed
                        acc[g] = mfma16(sP[q * LDP + base + 16 * tj[g] + (lane & 15)],
                                        sP[q * LDP + base + 16 * ti[g] + (lane & 15)], acc[g]);
            }
#pragma unroll
            for (int g = 0; g < PF_G; g++) {
                if (ti[g] < 0) continue;
                const int i = base + 16 * ti[g] + (lane & 15);
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int j = base + 16 * tj[g] + (lane >> 4) + 4 * v;
                    if (ti[g] < tj[g] || i <= j) A[(size_t)i + (size_t)j * b] = cv[g][v] - acc[g][v];
                }
            }
        };
        {
            double cva[PF_G][4], cvb[PF_G][4];
            int e0 = w * PF_G;
            if (e0 < ntri) load_group(e0, cva);
#pragma unroll 1
            while (e0 < ntri) {
                if (e0 + stride < ntri) load_group(e0 + stride, cvb);
                run_group(e0, cva);
                e0 += stride;
                if (e0 >= ntri) break;
                if (e0 + stride < ntri) load_group(e0 + stride, cva);
                run_group(e0, cvb);
                e0 += stride;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- GETRF
__device__ inline double piv_scale2(double x, double alpha)
{
    if (alpha == 0.0) return x;
    if (fabs(alpha) > DBL_MIN) return x * (1.0 / alpha);
    return x / alpha;
}

__device__ __forceinline__ void getrf2_tile(double *A, int b, double *dinvU, double *dinvL, double *smem)
{
    constexpr int NB = NB2;
    const int LDP = b + 16;
    double *sP = smem;                    // U panel rows: sP[q*LDP + c]
    double *sD = smem + (size_t)NB * LDP; // factored diagonal block (L strict lower, U upper)
    double *sUi = sD + NB * LDD;
    double *sLi = sUi + NB * LDD;
    double *sT = sP;                      // inverse scratch: the panel is not live yet
    __shared__ int s_slow;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;

    for (int j0 = 0; j0 < b; j0 += NB) {
        // 1. diagonal block, factored by wave 0 in registers with lane r holding
        //    row r: right-looking no-pivot LU (src/srclu/dgetrf_no_pivot.h:18-55
        //    on the 32x32 block).  Column i's multipliers are scaled in
        //    parallel across lanes (one piv_scale2 per lane); row i of U
        //    reaches the other lanes through v_readlane.
        for (int e = t; e < NB * NB; e += PF_THREADS)
            sD[(e % NB) + (e / NB) * LDD] = A[(size_t)(j0 + e % NB) + (size_t)(j0 + e / NB) * b];
        __syncthreads();
        if (w == 0) {
            const int r = lane & (NB - 1);
            double x[NB];
#pragma unroll
            for (int c = 0; c < NB; c++) x[c] = sD[r + c * LDD];
            bool slow = false;
#pragma unroll
            for (int i = 0; i < NB; i++) {
                const double alpha = readlane_d(x[i], i);
                slow |= !(fabs(alpha) > DBL_MIN);
                if (r > i) x[i] = piv_scale2(x[i], alpha);
                const double l = x[i];
                // row i of U into SGPRs first, then the rank-1 update
                double u[NB];
#pragma unroll
                for (int c = i + 1; c < NB; c++) u[c] = readlane_d(x[c], i);
#pragma unroll
                for (int c = i + 1; c < NB; c++)
                    if (r > i) x[c] -= l * u[c];
            }
            if (lane < NB) {
#pragma unroll
                for (int c = 0; c < NB; c++) sD[r + c * LDD] = x[c];
            }
            if (lane == 0) s_slow = slow ? 1 : 0;
        }
        __syncthreads();
        for (int e = t; e < NB * NB; e += PF_THREADS)
            A[(size_t)(j0 + e % NB) + (size_t)(j0 + e / NB) * b] = sD[(e % NB) + (e / NB) * LDD];
        tri_inverse32<true>(sD, sUi, sT, PF_THREADS);
        tri_inverse32<false>(sD, sLi, sT, PF_THREADS);
        for (int e = t; e < NB * NB; e += PF_THREADS) {
            const int r = e % NB, c = e / NB;
            if (dinvU) dinvU[(size_t)(j0 / NB) * NB * NB + e] = sUi[r + c * LDD];
            if (dinvL) dinvL[(size_t)(j0 / NB) * NB * NB + e] = sLi[r + c * LDD];
        }
        const int base = j0 + NB, rest = b - base;
        stage_panel(A, b, j0, base, sP, LDP, PF_THREADS);
        __syncthreads();
        if (rest == 0) break;
        const int nt = rest / 16;
        // U panel: Y = L_D^{-1} A[j0:j0+NB, base:b]   (tiles 0 .. nt-1)
        // L panel: X = A[base:b, j0:j0+NB] U_D^{-1}   (tiles nt .. 2nt-1, two 16-col halves each)
#pragma unroll 1
        for (int job = w; job < 2 * nt; job += PF_WAVES) {
            if (job < nt) {
                const int col = base + job * 16 + (lane & 15);
                d4 acc0 = zero4(), acc1 = zero4();
#pragma unroll
                for (int k4 = 0; k4 < NB; k4 += 4) {
                    const int q = k4 + (lane >> 4);
                    const double bv = sP[q * LDP + col];
                    acc0 = mfma16(sLi[(lane & 15) + q * LDD], bv, acc0);          // opA(i,q) = Linv(i,q)
                    acc1 = mfma16(sLi[(16 + (lane & 15)) + q * LDD], bv, acc1);
                }
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = (lane >> 4) + 4 * v;
                    sP[i * LDP + col] = acc0[v];
                    sP[(i + 16) * LDP + col] = acc1[v];
                    A[(size_t)(j0 + i) + (size_t)col * b] = acc0[v];
                    A[(size_t)(j0 + i + 16) + (size_t)col * b] = acc1[v];
                }
            } else if (!s_slow) {
                const int i0 = base + (job - nt) * 16;
                d4 acc0 = zero4(), acc1 = zero4();
#pragma unroll
                for (int k4 = 0; k4 < NB; k4 += 4) {
                    const int q = k4 + (lane >> 4);
                    const double av = A[(size_t)(i0 + (lane & 15)) + (size_t)(j0 + q) * b];
                    // D = X^T (rows of X along lanes): opA(c,q) = Uinv(q,c), opB(q,i) = A(i,q)
                    acc0 = mfma16(sUi[q + (lane & 15) * LDD], av, acc0);
                    acc1 = mfma16(sUi[q + (16 + (lane & 15)) * LDD], av, acc1);
                }
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = i0 + (lane & 15), cc = j0 + (lane >> 4) + 4 * v;
                    A[(size_t)i + (size_t)cc * b] = acc0[v];
                    A[(size_t)i + (size_t)(cc + 16) * b] = acc1[v];
                }
            } else if (lane < 16) {
                // reference scaling rule, one row per lane: x U_D = a (rare path:
                // a zero or sub-sfmin pivot in this block)
                const int i = base + (job - nt) * 16 + lane;
                double *row = A + (size_t)i + (size_t)j0 * b;
                for (int q = 0; q < NB; q++) {
                    double s = row[(size_t)q * b];
                    for (int p = 0; p < q; p++) s -= row[(size_t)p * b] * sD[p + q * LDD];
                    row[(size_t)q * b] = piv_scale2(s, sD[q + q * LDD]);
                }
            }
        }
        __syncthreads();
        // trailing: A[base:, base:] -= L_panel U_panel.  A job is one 16-row tile
        // row and GF_G consecutive 16-column tiles: the L-panel fragment (global)
        // is loaded once, the C tiles are loaded before the MFMAs.
        const int ngrp = (nt + GF_G - 1) / GF_G;
#pragma unroll 1
        for (int job = w; job < nt * ngrp; job += PF_WAVES) {
            const int bi = job / ngrp, bj0 = (job % ngrp) * GF_G;
            const int i0 = base + 16 * bi;
            double av[NB / 4];
#pragma unroll
            for (int k4 = 0; k4 < NB; k4 += 4)
                av[k4 / 4] = A[(size_t)(i0 + (lane & 15)) + (size_t)(j0 + k4 + (lane >> 4)) * b];
            double cv[GF_G][4];
            d4 acc[GF_G];
#pragma unroll
            for (int g = 0; g < GF_G; g++) {
                acc[g] = zero4();
                if (bj0 + g < nt) {
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        cv[g][v] = A[(size_t)(i0 + (lane & 15)) +
                                     (size_t)(base + 16 * (bj0 + g) + (lane >> 4) + 4 * v) * b];
                }
            }
            // operands swapped so that D holds the C tile transposed: 16 lanes
            // then cover 16 consecutive rows (128 B) of one column of C
#pragma unroll
            for (int k4 = 0; k4 < NB; k4 += 4) {
                const int q = k4 + (lane >> 4);
#pragma unroll
                for (int g = 0; g < GF_G; g++)
                    if (bj0 + g < nt)
                        acc[g] = mfma16(sP[q * LDP + base + 16 * (bj0 + g) + (lane & 15)], av[k4 / 4], acc[g]);
            }
#pragma unroll
            for (int g = 0; g < GF_G; g++) {
                if (bj0 + g >= nt) continue;
#pragma unroll
                for (int v = 0; v < 4; v++)
                    A[(size_t)(i0 + (lane & 15)) + (size_t)(base + 16 * (bj0 + g) + (lane >> 4) + 4 * v) * b] =
                        cv[g][v] - acc[g][v];
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(PF_THREADS) void potrf2_kernel(double *A, int b, int *info, int info_base,
                                                            double *dinv)
{
    extern __shared__ double smem[];
    potrf2_tile(A, b, info, info_base, dinv, smem);
}

__global__ __launch_bounds__(PF_THREADS) void getrf2_kernel(double *A, int b, double *dinvU, double *dinvL)
{
    extern __shared__ double smem[];
    getrf2_tile(A, b, dinvU, dinvL, smem);
}

// ---------------------------------------------------------------- batched
// 32x32 triangular solves by one wave, lane i (< 32) holding v_i and
// reg[j] = M(i, j): forward for lower M, backward for upper M.
__device__ inline double tri32_fwd(double v, const double (&reg)[NB2], bool unit, int lane)
{
#pragma unroll
    for (int j = 0; j < NB2; j++) {
        double yj = __shfl(v, j);
        if (!unit) yj /= __shfl(reg[j], j);
        if (lane == j) v = yj;
        else if (lane > j) v -= reg[j] * yj;
    }
    return v;
}

__device__ inline double tri32_bwd(double v, const double (&reg)[NB2], int lane)
{
#pragma unroll
    for (int j = NB2 - 1; j >= 0; j--) {
        const double xj = __shfl(v, j) / __shfl(reg[j], j);
        if (lane == j) v = xj;
        else if (lane < j) v -= reg[j] * xj;
    }
    return v;
}

// Matrix i of a batch, by one workgroup: copy into the 32-padded FP64
// workspace (identity padding), factor with the tile kernels' code (chol:
// Cholesky, else LU without pivoting), optionally solve, copy back.
__device__ void batch_one(const BatchArgs &p, int i, bool chol, double *smem)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    {
        const int n = p.n[i], b = (n + NB2 - 1) / NB2 * NB2;
        double *W = p.work + p.woff[i];
        const size_t off = (size_t)p.off[i];
        for (int e = t; e < b * b; e += PF_THREADS) {
            const int r = e % b, c = e / b;
            double v = (r == c) ? 1.0 : 0.0;
            if (r < n && c < n)
                v = p.f32 ? (double)((const float *)p.A)[off + r + (size_t)c * n]
                          : ((const double *)p.A)[off + r + (size_t)c * n];
            W[e] = v;
        }
        if (t == 0) p.info[i] = 0;
        __syncthreads();
        if (chol) potrf2_tile(W, b, p.info + i, 0, nullptr, smem);
        else getrf2_tile(W, b, nullptr, nullptr, smem);
        __syncthreads();
        if (p.B) {
            // x in LDS (padding rows are 0); forward then backward blocked
            // substitution with 32-row blocks
            double *sx = smem;
            const size_t bo = (size_t)p.boff[i];
            for (int e = t; e < b; e += PF_THREADS)
                sx[e] = (e < n) ? (p.f32 ? (double)((const float *)p.B)[bo + e] : ((const double *)p.B)[bo + e]) : 0.0;
            __syncthreads();
            // forward: Cholesky U^T y = r (lower operand U^T), LU L y = r (unit lower)
            for (int j0 = 0; j0 < b; j0 += NB2) {
                if (w == 0) {
                    double reg[NB2];
                    const int li = lane & (NB2 - 1);
#pragma unroll
                    for (int j = 0; j < NB2; j++)
                        reg[j] = chol ? W[(size_t)(j0 + j) + (size_t)(j0 + li) * b]    // U^T(i,j) = U(j,i)
                                      : W[(size_t)(j0 + li) + (size_t)(j0 + j) * b];   // L(i,j)
                    const double v = tri32_fwd(lane < NB2 ? sx[j0 + lane] : 0.0, reg, !chol, lane);
                    if (lane < NB2) sx[j0 + lane] = v;
                }
                __syncthreads();
                for (int r = j0 + NB2 + t; r < b; r += PF_THREADS) {
                    double acc = 0.0;
#pragma unroll 8
                    for (int q = 0; q < NB2; q++)
                        acc += (chol ? W[(size_t)(j0 + q) + (size_t)r * b] : W[(size_t)r + (size_t)(j0 + q) * b]) * sx[j0 + q];
                    sx[r] -= acc;
                }
                __syncthreads();
            }
            // backward: U x = y
            for (int j0 = b - NB2; j0 >= 0; j0 -= NB2) {
                if (w == 0) {
                    double reg[NB2];
                    const int li = lane & (NB2 - 1);
#pragma unroll
                    for (int j = 0; j < NB2; j++) reg[j] = W[(size_t)(j0 + li) + (size_t)(j0 + j) * b];   // U(i,j)
                    const double v = tri32_bwd(lane < NB2 ? sx[j0 + lane] : 0.0, reg, lane);
                    if (lane < NB2) sx[j0 + lane] = v;
                }
                __syncthreads();
                for (int r = t; r < j0; r += PF_THREADS) {
                    double acc = 0.0;
#pragma unroll 8
                    for (int q = 0; q < NB2; q++) acc += W[(size_t)r + (size_t)(j0 + q) * b] * sx[j0 + q];
                    sx[r] -= acc;
                }
                __syncthreads();
            }
            for (int e = t; e < n; e += PF_THREADS) {
                if (p.f32) ((float *)p.B)[bo + e] = (float)sx[e];
                else ((double *)p.B)[bo + e] = sx[e];
            }
        }
        for (int e = t; e < n * n; e += PF_THREADS) {
            const int r = e % n, c = e / n;
            const double v = W[(size_t)r + (size_t)c * b];
            if (p.f32) ((float *)p.A)[off + e] = (float)v;
            else ((double *)p.A)[off + e] = v;
        }
    }
}

// Persistent batch: a fixed grid of workgroups (what the CUs hold at once)
// takes matrices from p.queue[0] in the order list (largest first, both
// kinds: dynamic longest-processing-time scheduling) until the list is
// exhausted; every workgroup leaves through the same test.  The queue counter
// is zeroed before each launch.
__global__ __launch_bounds__(PF_THREADS) void batch_persist_kernel(BatchArgs p)
{
    extern __shared__ double smem[];
    __shared__ int s_idx;
    const int t = threadIdx.x;
    for (;;) {
        if (t == 0) s_idx = atomicAdd(p.queue, 1);
        __syncthreads();
        const int idx = s_idx;
        __syncthreads();                       // s_idx is rewritten by the next round
        if (idx >= p.count) break;
        const int i = p.order[idx];
        batch_one(p, i, p.kind[i] == 0, smem);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- TRSM
// Strip layout in LDS: row i, column c at i*32 + (c ^ 16*(i&1)) (odd rows
// swap their 16-column halves so four consecutive rows read by one MFMA
// fragment load fall on distinct bank halves).
__device__ inline int xidx(int i, int c) { return i * NB2 + (c ^ ((i & 1) << 4)); }

// F32: the tiles of p.mat and p.diag are FP32 (element offsets as for FP64
// tiles); the strip is solved in FP64 in LDS and rounded back on the store.
template <int KIND, bool F32>
__global__ __launch_bounds__(P2_THREADS) void trsm2_kernel(TrsmArgs p, const double *dinv)
{
    using DT = typename std::conditional<F32, float, double>::type;
    constexpr int NB = NB2, SW = NB2;
    extern __shared__ double sX[];
    const int b = p.b;
    const int strips = b / SW;
    const int tix = blockIdx.x / strips, strip = blockIdx.x % strips;
    const int c0 = strip * SW;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const DT *const mat = (const DT *)(const void *)p.mat;
    double *X = nullptr;
    DT *X32 = nullptr;                  // the strip's tile when it lives in an FP32 matrix
    const DT *D = (const DT *)(const void *)p.diag;
    if (p.batch) {
        D = mat + tile_off(b, p.MT, tix, tix);
        X = p.single + (size_t)tix * b * b;
        dinv += (size_t)tix * p.dinv_stride;
    } else if (p.mat) {
        const int idx = p.first + tix;
        const size_t xo = (KIND == TR_LUR) ? tile_off(b, p.MT, idx, p.k) : tile_off(b, p.MT, p.k, idx);
        if (F32) X32 = (DT *)(void *)p.mat + xo;
        else X = p.mat + xo;
    } else {
        X = p.single;
    }
    task_begin(p.ts, p.ts_base + tix);
    auto ldx = [&](size_t i) -> double { return (F32 && X32) ? (double)X32[i] : X[i]; };
    auto stx = [&](size_t i, double v) {
        if (F32 && X32) X32[i] = (DT)v;
        else X[i] = v;
    };
    // M(i,q): CHOL/LUR M = U^T -> D[q + i*b];  LUL M = L -> D[i + q*b]
    for (int e0 = t; e0 < b * SW; e0 += P2_THREADS * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = e0 + u * P2_THREADS;
            if (e < b * SW) {
                // element (row, col) of the tile this strip holds
                const int row = (KIND == TR_LUR) ? c0 + e % SW : e % b;
                const int col = (KIND == TR_LUR) ? e / SW : c0 + e / b;
                v[u] = p.ident ? (row == col ? 1.0 : 0.0) : ldx((size_t)row + (size_t)col * b);
            }
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = e0 + u * P2_THREADS;
            if (e < b * SW) {
                if (KIND == TR_LUR) sX[xidx(e / SW, e % SW)] = v[u];
                else sX[xidx(e % b, e / b)] = v[u];
            }
        }
    }
    __syncthreads();

    // fragments of the inverse of diagonal block r0/NB for phase (a), loaded
    // one block ahead (waves 0..3)
    const int ti = w & 1, tc = w >> 1;
    double mi[NB / 4];
    auto load_mi = [&](int r0) {
        const double *Mi = dinv + (size_t)(r0 / NB) * NB * NB;   // inverse of the diagonal block
#pragma unroll
        for (int k4 = 0; k4 < NB; k4 += 4) {
            const int q = k4 + (lane >> 4);
            const int i = ti * 16 + (lane & 15);
            // CHOL/LUR: M_rr^{-1} = (U_rr^{-1})^T -> Uinv(q,i);  LUL: Linv(i,q)
            mi[k4 / 4] = (KIND == TR_LUL) ? Mi[i + q * NB] : Mi[q + i * NB];
        }
    };
    if (w < 4) load_mi(0);
    for (int r0 = 0; r0 < b; r0 += NB) {
        // the M fragments of phase (b) do not depend on the solve: their loads
        // are issued first and land during phase (a) and its barriers.  A wave
        // owns row tiles it = w, w+8, ... below the block (both 16-column halves).
        const int below = (b - r0 - NB) / 16;
        constexpr int MAXT = 4;                 // row tiles per wave for b <= 512
        double av[MAXT][NB / 4];
#pragma unroll
        for (int m = 0; m < MAXT; m++) {
            const int it = w + m * P2_WAVES;
            if (it < below) {
                const int i = r0 + NB + it * 16 + (lane & 15);
#pragma unroll
                for (int k4 = 0; k4 < NB; k4 += 4) {
                    const int q = k4 + (lane >> 4);
                    av[m][k4 / 4] = (KIND == TR_LUL) ? (double)D[(size_t)i + (size_t)(r0 + q) * b]
                                                     : (double)D[(size_t)(r0 + q) + (size_t)i * b];
                }
            }
        }
        // (a) X_r <- M_rr^{-1} X_r : waves 0..3 own one 16x16 output tile each
        d4 acc = zero4();
        if (w < 4) {
#pragma unroll
            for (int k4 = 0; k4 < NB; k4 += 4) {
                const int q = k4 + (lane >> 4);
                acc = mfma16(mi[k4 / 4], sX[xidx(r0 + q, tc * 16 + (lane & 15))], acc);
            }
            if (r0 + NB < b) load_mi(r0 + NB);
        }
        __syncthreads();
        if (w < 4) {
#pragma unroll
            for (int v = 0; v < 4; v++)
                sX[xidx(r0 + ti * 16 + (lane >> 4) + 4 * v, tc * 16 + (lane & 15))] = acc[v];
        }
        __syncthreads();
        // (b) rows below: X[i] -= M(i, r-block) X_r
#pragma unroll
        for (int m = 0; m < MAXT; m++) {
            const int it = w + m * P2_WAVES;
            if (it >= below) continue;
            const int i0 = r0 + NB + it * 16;
            d4 u0 = zero4(), u1 = zero4();
#pragma unroll
            for (int k4 = 0; k4 < NB; k4 += 4) {
                const int q = k4 + (lane >> 4);
                u0 = mfma16(av[m][k4 / 4], sX[xidx(r0 + q, (lane & 15))], u0);
                u1 = mfma16(av[m][k4 / 4], sX[xidx(r0 + q, 16 + (lane & 15))], u1);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                sX[xidx(i0 + (lane >> 4) + 4 * v, (lane & 15))] -= u0[v];
                sX[xidx(i0 + (lane >> 4) + 4 * v, 16 + (lane & 15))] -= u1[v];
            }
        }
        __syncthreads();
    }

    for (int e = t; e < b * SW; e += P2_THREADS) {
        if (KIND == TR_LUR) stx((size_t)(c0 + e % SW) + (size_t)(e / SW) * b, sX[xidx(e / SW, e % SW)]);
        else stx((size_t)(e % b) + (size_t)(c0 + e / b) * b, sX[xidx(e % b, e / b)]);
    }
    task_end(p.ts, p.ts_base + tix);
}

// FP32 variant for the mixed-precision tiers' FP32 working copy: the strip
// is solved in FP32 in LDS (half the LDS, two workgroups per CU) with
// v_mfma_f32_16x16x4_f32 (twice the FP64 rate).  The block inverses stay FP64
// in memory and are rounded on load; D and the matrix tiles are FP32.  Same
// blocking as trsm2_kernel; the f32 MFMA result D[4(l/16)+v][l%16] differs
// from the f64 one only in its row mapping.  Batch (inverse) mode writes FP64.
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ inline f4v mfma16f(float a, float b, f4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int KIND>
__global__ __launch_bounds__(P2_THREADS, 2) void trsm2f_kernel(TrsmArgs p, const double *dinv)
{
    constexpr int NB = NB2, SW = NB2;
    extern __shared__ double sX_raw[];
    float *sX = (float *)sX_raw;
    const int b = p.b;
    const int strips = b / SW;
    const int tix = blockIdx.x / strips, strip = blockIdx.x % strips;
    const int c0 = strip * SW;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const float *const mat = (const float *)(const void *)p.mat;
    double *Xd = nullptr;               // FP64 output (batch mode / single tile)
    float *X32 = nullptr;               // the strip's tile in the FP32 matrix
    const float *D = (const float *)(const void *)p.diag;
    if (p.batch) {
        D = mat + tile_off(b, p.MT, tix, tix);
        Xd = p.single + (size_t)tix * b * b;
        dinv += (size_t)tix * p.dinv_stride;
    } else if (p.mat) {
        const int idx = p.first + tix;
        X32 = (float *)(void *)p.mat + ((KIND == TR_LUR) ? tile_off(b, p.MT, idx, p.k) : tile_off(b, p.MT, p.k, idx));
    } else {
        Xd = p.single;
    }
    task_begin(p.ts, p.ts_base + tix);
    for (int e0 = t; e0 < b * SW; e0 += P2_THREADS * 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = e0 + u * P2_THREADS;
            if (e < b * SW) {
                const int row = (KIND == TR_LUR) ? c0 + e % SW : e % b;
                const int col = (KIND == TR_LUR) ? e / SW : c0 + e / b;
                const size_t o = (size_t)row + (size_t)col * b;
                v[u] = p.ident ? (row == col ? 1.f : 0.f) : (X32 ? X32[o] : (float)Xd[o]);
            }
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = e0 + u * P2_THREADS;
            if (e < b * SW) {
                if (KIND == TR_LUR) sX[xidx(e / SW, e % SW)] = v[u];
                else sX[xidx(e % b, e / b)] = v[u];
            }
        }
    }
    __syncthreads();

    const int ti = w & 1, tc = w >> 1;
    float mi[NB / 4];
    auto load_mi = [&](int r0) {
        const double *Mi = dinv + (size_t)(r0 / NB) * NB * NB;
#pragma unroll
        for (int k4 = 0; k4 < NB; k4 += 4) {
            const int q = k4 + (lane >> 4);
            const int i = ti * 16 + (lane & 15);
            mi[k4 / 4] = (float)((KIND == TR_LUL) ? Mi[i + q * NB] : Mi[q + i * NB]);
        }
    };
    if (w < 4) load_mi(0);
    for (int r0 = 0; r0 < b; r0 += NB) {
        const int below = (b - r0 - NB) / 16;
        constexpr int MAXT = 4;
        float av[MAXT][NB / 4];
#pragma unroll
        for (int m = 0; m < MAXT; m++) {
            const int it = w + m * P2_WAVES;
            if (it < below) {
                const int i = r0 + NB + it * 16 + (lane & 15);
#pragma unroll
                for (int k4 = 0; k4 < NB; k4 += 4) {
                    const int q = k4 + (lane >> 4);
                    av[m][k4 / 4] = (KIND == TR_LUL) ? D[(size_t)i + (size_t)(r0 + q) * b] : D[(size_t)(r0 + q) + (size_t)i * b];
                }
            }
        }
        // (a) X_r <- M_rr^{-1} X_r : waves 0..3 own one 16x16 output tile each
        f4v acc = (f4v){0.f, 0.f, 0.f, 0.f};
        if (w < 4) {
#pragma unroll
            for (int k4 = 0; k4 < NB; k4 += 4) {
                const int q = k4 + (lane >> 4);
                acc = mfma16f(mi[k4 / 4], sX[xidx(r0 + q, tc * 16 + (lane & 15))], acc);
            }
            if (r0 + NB < b) load_mi(r0 + NB);
        }
        __syncthreads();
        if (w < 4) {
#pragma unroll
            for (int v = 0; v < 4; v++)
                sX[xidx(r0 + ti * 16 + 4 * (lane >> 4) + v, tc * 16 + (lane & 15))] = acc[v];
        }
        __syncthreads();
        // (b) rows below: X[i] -= M(i, r-block) X_r
#pragma unroll
        for (int m = 0; m < MAXT; m++) {
            const int it = w + m * P2_WAVES;
            if (it >= below) continue;
            const int i0 = r0 + NB + it * 16;
            f4v u0 = (f4v){0.f, 0.f, 0.f, 0.f}, u1 = u0;
#pragma unroll
            for (int k4 = 0; k4 < NB; k4 += 4) {
                const int q = k4 + (lane >> 4);
                u0 = mfma16f(av[m][k4 / 4], sX[xidx(r0 + q, (lane & 15))], u0);
                u1 = mfma16f(av[m][k4 / 4], sX[xidx(r0 + q, 16 + (lane & 15))], u1);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                sX[xidx(i0 + 4 * (lane >> 4) + v, (lane & 15))] -= u0[v];
                sX[xidx(i0 + 4 * (lane >> 4) + v, 16 + (lane & 15))] -= u1[v];
            }
        }
        __syncthreads();
    }

    for (int e = t; e < b * SW; e += P2_THREADS) {
        const size_t o = (KIND == TR_LUR) ? (size_t)(c0 + e % SW) + (size_t)(e / SW) * b
                                          : (size_t)(e % b) + (size_t)(c0 + e / b) * b;
        const float v = (KIND == TR_LUR) ? sX[xidx(e / SW, e % SW)] : sX[xidx(e % b, e / b)];
        if (X32) X32[o] = v;
        else Xd[o] = (double)v;
    }
    task_end(p.ts, p.ts_base + tix);
}

// ---------------------------------------------------------------- R blocks
// Inverses of the 32x32 diagonal blocks of the upper triangle of every
// diagonal tile (the R of a tiled QR): workgroup (k, j) inverts block j of
// tile (k,k) with the factorizations' recursive doubling and writes it where
// they write theirs, so that the batched TRSM2 can invert whole tiles.
__global__ __launch_bounds__(256) void upper_block_inv_kernel(const double *mat, int N, int b, double *dinv,
                                                               long stride)
{
    __shared__ double sM[NB2 * LDD], sX[NB2 * LDD], sT[NB2 * NB2];
    const int MT = N / b, nbk = b / NB2;
    const int k = blockIdx.x / nbk, j = blockIdx.x % nbk;
    const double *D = mat + tile_off(b, MT, k, k);
    const int j0 = j * NB2;
    for (int e = threadIdx.x; e < NB2 * NB2; e += 256) {
        const int r = e % NB2, c = e / NB2;
        sM[r + c * LDD] = (r <= c) ? D[(size_t)(j0 + r) + (size_t)(j0 + c) * b] : 0.0;
    }
    __syncthreads();
    tri_inverse32<true>(sM, sX, sT, 256);
    double *out = dinv + (size_t)k * stride + (size_t)j * NB2 * NB2;
    for (int e = threadIdx.x; e < NB2 * NB2; e += 256) out[e] = sX[(e % NB2) + (e / NB2) * LDD];
}

hipError_t launch_upper_block_inverses(const double *mat, int N, int b, double *dinv, long stride, hipStream_t s)
{
    if (b % NB2 || N % b) return hipErrorInvalidValue;
    const int MT = N / b;
    hipLaunchKernelGGL(upper_block_inv_kernel, dim3((unsigned)(MT * (b / NB2))), dim3(256), 0, s, mat, N, b, dinv,
                       stride);
    return hipGetLastError();
}

// ---------------------------------------------------------------- launchers
bool panel2_ok(int b) { return b % NB2 == 0 && b >= NB2 && b <= 512; }

void panel2_init()
{
    // dynamic LDS at the largest supported tile (b = 512)
    const int lds_trsm = 512 * NB2 * 8;           // 131072 B
    hipFuncSetAttribute((const void *)potrf2_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)panel_lds_bytes(512, 2));
    hipFuncSetAttribute((const void *)getrf2_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)panel_lds_bytes(512, 3));
    hipFuncSetAttribute((const void *)batch_persist_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)panel_lds_bytes(512, 3));
    hipFuncSetAttribute((const void *)trsm2_kernel<TR_CHOL, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_trsm);
    hipFuncSetAttribute((const void *)trsm2_kernel<TR_LUL, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_trsm);
    hipFuncSetAttribute((const void *)trsm2_kernel<TR_LUR, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_trsm);
    hipFuncSetAttribute((const void *)trsm2_kernel<TR_CHOL, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_trsm);
    hipFuncSetAttribute((const void *)trsm2_kernel<TR_LUL, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_trsm);
    hipFuncSetAttribute((const void *)trsm2_kernel<TR_LUR, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_trsm);
    hipFuncSetAttribute((const void *)trsm2f_kernel<TR_CHOL>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_trsm / 2);
    hipFuncSetAttribute((const void *)trsm2f_kernel<TR_LUL>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_trsm / 2);
    hipFuncSetAttribute((const void *)trsm2f_kernel<TR_LUR>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_trsm / 2);
}

hipError_t launch_potrf2(double *tile, int b, int *info, int info_base, double *dinv, hipStream_t s)
{
    const size_t lds = panel_lds_bytes(b, 2);
    hipLaunchKernelGGL(potrf2_kernel, dim3(1), dim3(PF_THREADS), lds, s, tile, b, info, info_base, dinv);
    return hipGetLastError();
}

hipError_t launch_getrf2(double *tile, int b, double *dinvU, double *dinvL, hipStream_t s)
{
    const size_t lds = panel_lds_bytes(b, 3);
    hipLaunchKernelGGL(getrf2_kernel, dim3(1), dim3(PF_THREADS), lds, s, tile, b, dinvU, dinvL);
    return hipGetLastError();
}

int batch_grid(int bmax, int count)
{
    int per = 0, dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_persist_kernel, PF_THREADS,
                                                     panel_lds_bytes(bmax, 3)) != hipSuccess || per < 1)
        per = 1;
    const int g = per * prop.multiProcessorCount;
    return count < g ? count : g;
}

hipError_t launch_batch(const BatchArgs &a, int bmax, int grid, hipStream_t s)
{
    if (a.count <= 0) return hipSuccess;
    if (bmax <= 0 || bmax > 512 || bmax % NB2 || grid < 1 || !a.queue || !a.kind) return hipErrorInvalidValue;
    hipLaunchKernelGGL(batch_persist_kernel, dim3((unsigned)grid), dim3(PF_THREADS), panel_lds_bytes(bmax, 3), s, a);
    return hipGetLastError();
}

hipError_t launch_trsm2(int kind, const TrsmArgs &a_in, const double *dinv, hipStream_t s)
{
    if (a_in.ntiles <= 0) return hipSuccess;
    if (sched_rec_on()) { rec_trsm(kind, a_in, dinv, s); return hipSuccess; }
    TrsmArgs a = a_in;
    trace_trsm(kind, a);
    dim3 grid((unsigned)(a.ntiles * (a.b / NB2))), block(P2_THREADS);
    const size_t lds = (size_t)a.b * NB2 * 8;
    if (a.f32) {
        // FP32 tiles: solved in FP32 (PALIAL_TRSM_F64=1 keeps FP64 arithmetic)
        static const bool f64 = [] {
            const char *e = getenv("PALIAL_TRSM_F64");
            return e && atoi(e) != 0;
        }();
        const size_t lds32 = (size_t)a.b * NB2 * 4;
        if (!f64 && kind == TR_CHOL) hipLaunchKernelGGL(trsm2f_kernel<TR_CHOL>, grid, block, lds32, s, a, dinv);
        else if (!f64 && kind == TR_LUL) hipLaunchKernelGGL(trsm2f_kernel<TR_LUL>, grid, block, lds32, s, a, dinv);
        else if (!f64) hipLaunchKernelGGL(trsm2f_kernel<TR_LUR>, grid, block, lds32, s, a, dinv);
        else if (kind == TR_CHOL) hipLaunchKernelGGL((trsm2_kernel<TR_CHOL, true>), grid, block, lds, s, a, dinv);
        else if (kind == TR_LUL) hipLaunchKernelGGL((trsm2_kernel<TR_LUL, true>), grid, block, lds, s, a, dinv);
        else hipLaunchKernelGGL((trsm2_kernel<TR_LUR, true>), grid, block, lds, s, a, dinv);
    } else {
        if (kind == TR_CHOL) hipLaunchKernelGGL((trsm2_kernel<TR_CHOL, false>), grid, block, lds, s, a, dinv);
        else if (kind == TR_LUL) hipLaunchKernelGGL((trsm2_kernel<TR_LUL, false>), grid, block, lds, s, a, dinv);
        else hipLaunchKernelGGL((trsm2_kernel<TR_LUR, false>), grid, block, lds, s, a, dinv);
    }
    return hipGetLastError();
}

}  // namespace palial
