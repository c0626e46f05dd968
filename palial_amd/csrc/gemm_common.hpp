// gemm_common.hpp -- task decode shared by the batched tile GEMM kernels
// (FP64 in k_gemm.hip, FP32/BF16 in k_gemm_lp.hip).
#pragma once
#include "internal.hpp"

namespace palial {

__device__ inline double apply_mask(int mask, double v, int r, int c)
{
    // r, c: coordinates inside the stored tile
    if (mask == MK_UPPER) return (r <= c) ? v : 0.0;
    if (mask == MK_UNIT_LOWER) return (r > c) ? v : ((r == c) ? 1.0 : 0.0);
    return v;
}

struct TaskPtrs {
    const double *A, *B, *D;
    double *C, *E;
    bool diag;   // SYRK diagonal tile: only the upper triangle is written
    bool valid;
    int ia, ib;  // panel positions of the A / B operand tiles (packed operands), -1: none
};

__host__ __device__ inline void tri_decode(int idx, int &n, int &m)
{
    // idx = m(m+1)/2 + n with 0 <= n <= m
    int mm = (int)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
    while ((mm + 1) * (mm + 2) / 2 <= idx) mm++;
    while (mm * (mm + 1) / 2 > idx) mm--;
    m = mm;
    n = idx - mm * (mm + 1) / 2;
}

__host__ __device__ inline TaskPtrs decode(const GemmArgs &p, int task)
{
    task += p.task0;
    TaskPtrs t;
    t.A = t.B = t.D = nullptr;
    t.C = t.E = nullptr;
    t.diag = false;
    t.valid = true;
    t.ia = t.ib = -1;
    const int b = p.b, MT = p.MT, k = p.k;
    const size_t bb = (size_t)b * b;
    double *M = p.mat;
    switch (p.mode) {
    case GM_SINGLE:
        t.A = p.A; t.B = p.B; t.C = p.C; t.D = p.D; t.E = p.E;
        t.diag = p.single_diag != 0;
        break;
    case GM_CHOL_TRAIL:
    case GM_CHOL_REST: {
        int n, m;
        tri_decode(task, n, m);
        int off = (p.mode == GM_CHOL_TRAIL) ? k + 1 : k + 2;
        n += off; m += off;
        t.A = M + tile_off(b, MT, k, n);
        t.B = M + tile_off(b, MT, k, m);
        t.C = M + tile_off(b, MT, n, m);
        t.D = t.C;
        t.diag = (n == m);
        t.ia = n - k - 1;
        t.ib = m - k - 1;
        break;
    }
    case GM_CHOL_ROW:
    case GM_CHOL_FIRST: {
        // GM_CHOL_FIRST: task T is the diagonal tile (k+2, k+2)
        const bool last = p.mode == GM_CHOL_FIRST && task == p.T;
        int n = last ? k + 2 : k + 1, m = last ? k + 2 : k + 1 + task;
        t.A = M + tile_off(b, MT, k, n);
        t.B = M + tile_off(b, MT, k, m);
        t.C = M + tile_off(b, MT, n, m);
        t.D = t.C;
        t.diag = (n == m);
        t.ia = n - k - 1;
        t.ib = m - k - 1;
        break;
    }
    case GM_LU_TRAIL:
    case GM_LU_REST: {
        // GM_LU_TRAIL covers T x T tiles from k+1, GM_LU_REST (T-1) x (T-1) from k+2
        const int off = (p.mode == GM_LU_TRAIL) ? k + 1 : k + 2;
        const int rows = (p.mode == GM_LU_TRAIL) ? p.T : p.T - 1;
        int m = off + task % rows, n = off + task / rows;
        t.A = M + tile_off(b, MT, m, k);
        t.B = M + tile_off(b, MT, k, n);
        t.C = M + tile_off(b, MT, m, n);
        t.D = t.C;
        t.ia = m - k - 1;
        t.ib = n - k - 1;
        break;
    }
    case GM_LU_PANEL:
    case GM_LU_FIRST: {
        // tasks [0, T): column k+1, rows m = k+1..MT-1 ; [T, 2T-1): row k+1, cols n = k+2..
        // GM_LU_FIRST: task 2T-1 is tile (k+2, k+2)
        int m, n;
        if (task < p.T) { m = k + 1 + task; n = k + 1; }
        else if (task < 2 * p.T - 1) { m = k + 1; n = k + 2 + (task - p.T); }
        else { m = k + 2; n = k + 2; }
        t.A = M + tile_off(b, MT, m, k);
        t.B = M + tile_off(b, MT, k, n);
        t.C = M + tile_off(b, MT, m, n);
        t.D = t.C;
        t.ia = m - k - 1;
        t.ib = n - k - 1;
        break;
    }
    case GM_QR_W: {
        int n = k + 1 + task;
        t.A = M + tile_off(b, MT, p.m, k);       // V (transposed)
        t.B = M + tile_off(b, MT, p.m, n);       // A2
        t.C = p.w1 + (size_t)task * bb;          // W
        t.D = M + tile_off(b, MT, k, n);         // + A1
        break;
    }
    case GM_QR_TW: {
        int n = k + 1 + task;
        t.A = p.smat + tile_off(b, MT, p.m, k);  // T (transposed, upper)
        t.B = p.w1 + (size_t)task * bb;
        t.C = p.w2 + (size_t)task * bb;
        t.E = M + tile_off(b, MT, k, n);         // A1 -= T^T W
        break;
    }
    case GM_QR_VW: {
        int n = k + 1 + task;
        t.A = M + tile_off(b, MT, p.m, k);       // V
        t.B = p.w2 + (size_t)task * bb;
        t.C = M + tile_off(b, MT, p.m, n);
        t.D = t.C;
        break;
    }
    case GM_ORM_W: {
        int n = k + 1 + task;
        t.A = p.A ? p.A : M + tile_off(b, MT, k, k);   // V (transposed, unit lower)
        t.B = M + tile_off(b, MT, k, n);
        t.C = p.w1 + (size_t)task * bb;
        break;
    }
    case GM_ORM_TW: {
        t.A = p.smat + tile_off(b, MT, k, k);    // T (transposed, upper)
        t.B = p.w1 + (size_t)task * bb;
        t.C = p.w2 + (size_t)task * bb;
        break;
    }
    case GM_ORM_VW: {
        int n = k + 1 + task;
        t.A = p.A ? p.A : M + tile_off(b, MT, k, k);   // V (unit lower)
        t.B = p.w2 + (size_t)task * bb;
        t.C = M + tile_off(b, MT, k, n);
        t.D = t.C;
        break;
    }
    case GM_QA_W: {
        const int n = task;
        t.A = M + tile_off(b, MT, p.m, k);        // V (transposed)
        t.B = p.cmat + tile_off(b, MT, p.m, n);   // C2
        t.C = p.w1 + (size_t)task * bb;           // W
        t.D = p.cmat + tile_off(b, MT, k, n);     // + C1
        break;
    }
    case GM_QA_TW: {
        const int n = task;
        t.A = p.smat + tile_off(b, MT, p.m, k);   // T (op by the launch), upper
        t.B = p.w1 + (size_t)task * bb;
        t.C = p.w2 + (size_t)task * bb;
        t.E = p.cmat + tile_off(b, MT, k, n);     // C1 -= op(T) W
        break;
    }
    case GM_QA_VW: {
        const int n = task;
        t.A = M + tile_off(b, MT, p.m, k);        // V
        t.B = p.w2 + (size_t)task * bb;
        t.C = p.cmat + tile_off(b, MT, p.m, n);
        t.D = t.C;
        break;
    }
    case GM_OA_W: {
        const int n = task;
        t.A = M + tile_off(b, MT, k, k);          // V (transposed, unit lower)
        t.B = p.cmat + tile_off(b, MT, k, n);
        t.C = p.w1 + (size_t)task * bb;
        break;
    }
    case GM_OA_TW: {
        t.A = p.smat + tile_off(b, MT, k, k);     // T (op by the launch), upper
        t.B = p.w1 + (size_t)task * bb;
        t.C = p.w2 + (size_t)task * bb;
        break;
    }
    case GM_OA_VW: {
        const int n = task;
        t.A = M + tile_off(b, MT, k, k);          // V (unit lower)
        t.B = p.w2 + (size_t)task * bb;
        t.C = p.cmat + tile_off(b, MT, k, n);
        t.D = t.C;
        break;
    }
    case GM_CHOL_ROW2: {
        // tasks 0..T-3: tiles (k+2, k+3+task); task T-2: tile (k+3, k+3)
        const bool last = task == p.T - 2;
        const int n = last ? k + 3 : k + 2, m = last ? k + 3 : k + 3 + task;
        t.A = M + tile_off(b, MT, k, n);
        t.B = M + tile_off(b, MT, k, m);
        t.C = M + tile_off(b, MT, n, m);
        t.D = t.C;
        t.diag = (n == m);
        t.ia = n - k - 1;
        t.ib = m - k - 1;
        break;
    }
    case GM_CHOL_REST2: {
        int n, m;
        tri_decode(task, n, m);
        n += k + 3; m += k + 3;
        t.A = M + tile_off(b, MT, k, n);
        t.B = M + tile_off(b, MT, k, m);
        t.C = M + tile_off(b, MT, n, m);
        t.D = t.C;
        t.diag = (n == m);
        t.ia = n - k - 1;        // panel k; panel k+1 holds the same tiles at ia - 1, ib - 1
        t.ib = m - k - 1;
        break;
    }
    case GM_LU_ROW2:
    case GM_LU_REST2: {
        int m, n;
        if (p.mode == GM_LU_REST2) {
            const int rows = p.T - 2;
            m = k + 3 + task % rows;
            n = k + 3 + task / rows;
        } else if (task < p.T - 2) {
            m = k + 3 + task; n = k + 2;                      // column k+2
        } else if (task < 2 * p.T - 4) {
            m = k + 2; n = k + 3 + (task - (p.T - 2));        // row k+2
        } else {
            m = k + 3; n = k + 3;
        }
        t.A = M + tile_off(b, MT, m, k);
        t.B = M + tile_off(b, MT, k, n);
        t.C = M + tile_off(b, MT, m, n);
        t.D = t.C;
        t.ia = m - k - 1;        // GM_LU_REST2: panel k+1 holds the same tiles at ia - 1, ib - 1
        t.ib = n - k - 1;
        break;
    }
    case GM_BI_W:
    case GM_BI_X: {
        const int q = p.bq, w = q - p.bd;
        const int K = task / w, j = task % w, i = j + p.bd;
        auto tl = [&](int r, int c) { return ((size_t)(K * q + r) * q + c) * bb; };
        if (p.mode == GM_BI_W) {
            const int m = j + p.bt;
            t.A = p.w1 + tl(i, m);
            t.B = M + tl(m, j);
            t.C = p.w2 + (size_t)task * bb;
            t.D = p.bt > 0 ? t.C : nullptr;
        } else {
            t.A = M + tl(i, i);
            t.B = p.w2 + (size_t)task * bb;
            t.C = M + tl(i, j);
        }
        break;
    }
    default:
        t.valid = false;
    }
    return t;
}

// Workgroup index remap so that consecutive logical blocks run on the same
// XCD (hardware dispatches workgroups round-robin over the 8 XCDs, each with
// its own L2): blocks of one output tile then share that XCD's L2 copy of
// their two operand tiles.
__device__ inline int xcd_remap(int pid, int nwg)
{
    constexpr int NX = 8;
    int q = nwg / NX, r = nwg % NX;
    int xcd = pid % NX, local = pid / NX;
    return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}


}  // namespace palial
