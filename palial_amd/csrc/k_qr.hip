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
#include <type_traits>

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

// ------------------------------------------------------------------ blocked
// For b % 32 == 0, b <= 512 the tile is factored in panels of 32 columns:
//   qr_panel_kernel  : dgeqr2 / the dtsqrt column loop over the 32 panel
//                      columns (one workgroup, panel in registers), writing V,
//                      the head rows, tau and the diagonal 32 x 32 block of T;
//   qr_apply_kernel  : the panel's block reflector I - Y Tpp Y^T applied to the
//                      trailing columns in 32-column chunks (FP64 MFMA), the
//                      Gram blocks G(q, p) = Y_q^T Y_p of earlier panels, and
//                      the off-diagonal T blocks of the previous panel by
//                      T(q, pc) = -(sum_r T(q, r) G(r, pc)) T(pc, pc).
// The result is the reference's T (src/srcqr/dtsqrt.h:65-86, dlarft 'F','C'):
// T is fixed by V and tau, T(0:i,i) = -tau_i T(0:i,0:i) Y(:,0:i)^T Y(:,i), and
// the block formula evaluates the same recurrence 32 columns at a time.
//
// Rows of one panel's reflectors ("stacked rows"): 32 head rows, then nt tail rows.
//   GE (geqrt of tile A1): head = tile rows c0..c0+31, tail = rows c0+32..b-1;
//      reflector i has 1 at head row i, v below, 0 above.
//   TS (tsqrt of [A1; A2]): head = rows c0..c0+31 of R (A1), tail = A2 rows
//      0..b-1; reflector i is e_i on the head and column i of A2 on the tail.
namespace {

constexpr int QP = 32;
constexpr int QP_THREADS = 512;    // thread (j, g): panel column j = t / 16, row group g = t % 16
constexpr int QA_THREADS = 512;    // eight waves: a 16 x 16 quarter of a 32 x 32 block and one half of K each

typedef double qd4 __attribute__((ext_vector_type(4)));

__device__ inline qd4 qmfma(double a, double b, qd4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// DPP move of a double inside rows of 16 lanes
template <int CTRL>
__device__ inline double dpp_d(double v)
{
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 16 lanes of a DPP row: pairs, quads (quad_perm), then the
// half-row and row mirrors.  Each step adds two values that the partner lane
// adds in the other order, so every lane ends with the same bits.
__device__ inline double sum16(double v)
{
    v += dpp_d<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_d<0x141>(v);   // row_half_mirror
    v += dpp_d<0x140>(v);   // row_mirror
    return v;
}

__device__ inline double readlane_d(double v, int lane)
{
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)x, lane);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes of a wave, the same bits in every lane.
__device__ inline double wave_sum64(double v)
{
    v = sum16(v);
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

// Thread layout: wave w owns panel columns w, w + 8, w + 16, w + 24; lane g
// holds tail rows g + 64 s (s < 8) of them in registers, so that the
// reflector's tail is read from LDS once per row set (8 values per thread)
// and the branches on the column index are uniform per wave.
// One barrier per column: after the update of step i, the wave holding
// column i+1 builds reflector i+1 (norm, tau, scaled v into the other half of
// the double-buffered sv) while the other waves finish step i.  Column c of
// the diagonal T block is formed in step c+1, 16 lanes per row of T.
template <bool GE>
__device__ __forceinline__ void panel_body(double *A1, double *A2, double *T, double *tau_out, int b, int c0)
{
    constexpr int SV = 10;              // sv stride per lane (80 B: b128 reads without bank conflicts)
    __shared__ double sH[QP][QP + 1];   // head rows: sH[h][c] = (head row h, panel column c)
    __shared__ double sG[QP][QP + 1];   // sG[q][i] = Y_q^T Y_i, q < i
    __shared__ double sT[QP][QP + 1];   // diagonal block of T (upper part)
    __shared__ __align__(16) double sv[2][64 * SV];
    __shared__ double sTau[QP];
    const int t = threadIdx.x, g = t & 63, w = t >> 6;
    const int nt = GE ? b - c0 - QP : b;
    double *tail = GE ? A1 + c0 + QP : A2;
    double *head = A1 + c0;
    double a[4][8];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const double *src = tail + (size_t)(c0 + w + 8 * q) * b;
#pragma unroll
        for (int s = 0; s < 8; s++) a[q][s] = g + 64 * s < nt ? src[g + 64 * s] : 0.0;
    }
    for (int e = t; e < QP * QP; e += QP_THREADS) {
        const int h = e & 31, c = e >> 5;
        sH[h][c] = head[(size_t)(c0 + c) * b + h];
    }
    __syncthreads();

    // reflector of column c = w + 8 qc, by its wave
    auto reflect = [&](int c) {
        const int qc = c >> 3;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (q != qc) continue;
            double p0 = 0.0, p1 = 0.0;
#pragma unroll
            for (int s = 0; s < 8; s += 2) {
                p0 += a[q][s] * a[q][s];
                p1 += a[q][s + 1] * a[q][s + 1];
            }
            double p = p0 + p1;
            if (GE && g < QP && g > c) p += sH[g][c] * sH[g][c];
            p = wave_sum64(p);
            const double alpha = readlane_d(g == c ? sH[c][c] : 0.0, c);
            double tau, beta, scal;
            householder(alpha, p, tau, beta, scal);
            if (g == c) {
                sH[c][c] = beta;
                sTau[c] = tau;
                tau_out[c0 + c] = tau;
            }
            if (tau != 0.0) {
#pragma unroll
                for (int s = 0; s < 8; s++) a[q][s] *= scal;
                if (GE && g < QP && g > c) sH[g][c] *= scal;
            }
            double *dst = sv[c & 1] + g * SV;
#pragma unroll
            for (int s = 0; s < 8; s += 2) *(double2 *)(dst + s) = double2{a[q][s], a[q][s + 1]};
        }
    };
    // T(jr, c) for rows jr <= c (G(:, c) and T(jr, 0:c) complete); 16 lanes per row
    auto tcol = [&](int c) {
        const int jr = 4 * w + (g >> 4), gl = g & 15;
        if (jr < c) {
            double s = 0.0;
            for (int q = jr + gl; q < c; q += 16) s += sT[jr][q] * sG[q][c];
            s = sum16(s);
            const double tc = sTau[c];
            if (gl == 0) sT[jr][c] = tc == 0.0 ? 0.0 : -tc * s;
        } else if (jr == c && gl == 0) {
            sT[c][c] = sTau[c];
        }
    };

    if (w == 0) reflect(0);
    for (int i = 0; i < QP; i++) {
        __syncthreads();
        const double tau = sTau[i];
        const double *vp = sv[i & 1] + g * SV;
        double v[8];
#pragma unroll
        for (int s = 0; s < 8; s += 2) {
            const double2 x = *(const double2 *)(vp + s);
            v[s] = x.x;
            v[s + 1] = x.y;
        }
        const bool hrow = GE && g < QP && g > i;   // head row of reflector i below its 1
        const double vh = hrow ? sH[g][i] : 0.0;
        // the wave holding column i+1 (slot qn) updates that column first and
        // builds reflector i+1 before its other columns: that is the path to
        // the next barrier.  slot<q> keeps the register index static.
        auto slot = [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int jc = w + 8 * q;
            if (jc == i) return;
            double d0 = 0.0, d1 = 0.0;
#pragma unroll
            for (int s = 0; s < 8; s += 2) {
                d0 += v[s] * a[q][s];
                d1 += v[s + 1] * a[q][s + 1];
            }
            double dq = d0 + d1;
            if (hrow) dq += vh * sH[g][jc];
            dq = wave_sum64(dq);
            if (jc > i) {
                const double f = tau * (sH[i][jc] + dq);
#pragma unroll
                for (int s = 0; s < 8; s++) a[q][s] -= f * v[s];
                if (hrow) sH[g][jc] -= f * vh;
                if (g == 0) sH[i][jc] -= f;
            } else if (g == 0) {
                sG[jc][i] = GE ? sH[i][jc] + dq : dq;
            }
        };
        using S0 = std::integral_constant<int, 0>;
        using S1 = std::integral_constant<int, 1>;
        using S2 = std::integral_constant<int, 2>;
        using S3 = std::integral_constant<int, 3>;
        const bool owner = i + 1 < QP && w == ((i + 1) & 7);
        const int qn = owner ? (i + 1) >> 3 : -1;
        if (owner) {
            switch (qn) {
            case 0: slot(S0{}); break;
            case 1: slot(S1{}); break;
            case 2: slot(S2{}); break;
            default: slot(S3{}); break;
            }
            reflect(i + 1);
        }
        if (qn != 0) slot(S0{});
        if (qn != 1) slot(S1{});
        if (qn != 2) slot(S2{});
        if (qn != 3) slot(S3{});
        if (i > 0) tcol(i - 1);
    }
    __syncthreads();
    tcol(QP - 1);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double *dst = tail + (size_t)(c0 + w + 8 * q) * b;
#pragma unroll
        for (int s = 0; s < 8; s++)
            if (g + 64 * s < nt) dst[g + 64 * s] = a[q][s];
    }
    for (int e = t; e < QP * QP; e += QP_THREADS) {
        const int r = e & 31, c = e >> 5;
        if (GE || r <= c) head[(size_t)(c0 + c) * b + r] = sH[r][c];
        if (r <= c) T[(size_t)(c0 + r) + (size_t)(c0 + c) * b] = sT[r][c];
    }
}

template <bool GE>
__global__ __launch_bounds__(QP_THREADS) void qr_panel_kernel(double *A1, double *A2, double *T, double *tau_out,
                                                              int b, int c0)
{
    panel_body<GE>(A1, A2, T, tau_out, b, c0);
}

// acc += Y(0:n, 0:16)^T X(0:n, 0:16) for column-major Y, X (ld b), n % 16 == 0.
// Lane l takes column l % 16 and the four consecutive rows 4 (l / 16) + s of
// each 16-row group, so that both operands are read as 32-byte pieces.
__device__ inline qd4 gram16(const double *Y, const double *X, int n, int b, qd4 acc)
{
    constexpr int D = 4;   // 16-row groups in flight
    const int l = threadIdx.x & 63;
    const size_t o = (size_t)(l & 15) * b + 4 * (l >> 4);
    const double *y = Y + o, *x = X + o;
    const int m = n >> 4;
    qd4 yb[D], xb[D];
#pragma unroll
    for (int u = 0; u < D; u++)
        if (u < m) {
            yb[u] = *(const qd4 *)(y + 16 * u);
            xb[u] = *(const qd4 *)(x + 16 * u);
        }
    for (int r = 0; r < m; r += D) {
#pragma unroll
        for (int u = 0; u < D; u++) {
            if (r + u < m) {
                const qd4 yv = yb[u], xv = xb[u];
                if (r + u + D < m) {
                    yb[u] = *(const qd4 *)(y + 16 * (r + u + D));
                    xb[u] = *(const qd4 *)(x + 16 * (r + u + D));
                }
                acc = qmfma(yv.x, xv.x, acc);
                acc = qmfma(yv.y, xv.y, acc);
                acc = qmfma(yv.z, xv.z, acc);
                acc = qmfma(yv.w, xv.w, acc);
            }
        }
    }
    return acc;
}

// Role A (blocks [0, na)): chunk x = columns cc = c0 + 32 + 32 x:
//   W = Yp^T X;  W = Tpp^T W;  X -= Yp W          (dlarfb / dtsqrt's gemv + ger)
//   with look != 0, block 0 (chunk 0 = the columns of panel p+1) then factors
//   panel p+1 (panel_body), while the other blocks finish this launch's work
// Role B (blocks [na, na + nbk)): q < p: G(32q + j, c0 + i) = Y_q(:, j)^T Y_p(:, i)
// Role C (the rest): row chunk q < pc of T column block pc (cq = 32 pc):
//   T(q, pc) = -(sum_{r=q}^{pc-1} T(q, r) G(r, pc)) T(pc, pc)
// No two blocks of a launch touch the same columns: panel p+1 and the chunks
// are disjoint, role B reads only panels <= p, role C writes T column block pc.
template <bool GE>
__global__ __launch_bounds__(QA_THREADS) void qr_apply_kernel(double *A1, double *A2, double *T, double *G,
                                                              double *tau, int b, int p, int na, int nbk, int pc,
                                                              int look)
{
    __shared__ double sW[QP][QP + 1];
    __shared__ double sW2[QP][QP + 1];
    __shared__ double sTp[QP][QP + 1];
    __shared__ double sY[QP][QP + 1];
    __shared__ qd4 sAcc[4][64];                    // K-half partial products of waves 4..7
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const int blk = w & 3, wi = blk & 1, wj = blk >> 1, kh = w >> 2;
    const int c0 = QP * p;
    const int x = blockIdx.x;
    const int nt = GE ? b - c0 - QP : b;
    double *head = A1 + c0;                        // stacked row 0 (column offsets added)
    double *tail = GE ? A1 + c0 + QP : A2;         // stacked row 32
    const size_t ldb = b;
    if (x < na + nbk) {
        const bool roleA = x < na;
        const int xc = roleA ? c0 + QP + QP * x : QP * (x - na);   // first column of X
        qd4 acc = {0.0, 0.0, 0.0, 0.0};
        // head rows: Yp unit lower (GE) or identity (TS); K half 0 only
        if (GE && kh == 0) {
            const int c = l & 15, kk = l >> 4;
            const int ic = 16 * wi + c;                 // panel column of the A operand
            for (int r = 0; r < QP; r += 16) {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const int h = r + 4 * kk + s;
                    const double yv = h > ic ? head[(size_t)(c0 + ic) * ldb + h] : (h == ic ? 1.0 : 0.0);
                    const double xv = head[(size_t)(xc + 16 * wj + c) * ldb + h];
                    acc = qmfma(yv, xv, acc);
                }
            }
        }
        {
            const int half = (nt >> 5) << 4;            // tail rows [0, half) | [half, nt)
            const int r0 = kh ? half : 0, r1 = kh ? nt : half;
            acc = gram16(tail + (size_t)(c0 + 16 * wi) * ldb + r0, tail + (size_t)(xc + 16 * wj) * ldb + r0, r1 - r0,
                         b, acc);
        }
        if (kh) sAcc[blk][l] = acc;
        __syncthreads();
        if (!kh) {
            const qd4 o = sAcc[blk][l];
            acc[0] += o[0]; acc[1] += o[1]; acc[2] += o[2]; acc[3] += o[3];
        }
        // D layout: lane l, value v -> (i = 16 wi + l / 16 + 4 v, j = 16 wj + l % 16)
        if (!roleA) {
            if (!kh) {
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = 16 * wi + (l >> 4) + 4 * v, jj = 16 * wj + (l & 15);
                    G[(size_t)(xc + jj) + (size_t)(c0 + i) * ldb] = acc[v];
                }
            }
            return;
        }
        if (!kh) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = 16 * wi + (l >> 4) + 4 * v, jj = 16 * wj + (l & 15);
                double wv = acc[v];
                if (!GE) wv += head[(size_t)(xc + jj) * ldb + i];   // identity head of Yp
                sW[i][jj] = wv;
            }
        }
        for (int e = t; e < QP * QP; e += QA_THREADS) {
            const int r = e & 31, c = e >> 5;
            sTp[r][c] = r <= c ? T[(size_t)(c0 + r) + (size_t)(c0 + c) * ldb] : 0.0;
            if (GE) sY[r][c] = r > c ? head[(size_t)(c0 + c) * ldb + r] : (r == c ? 1.0 : 0.0);
        }
        __syncthreads();
        // W2 = Tpp^T W
        for (int e = t; e < QP * QP; e += QA_THREADS) {
            const int i = e >> 5, jj = e & 31;
            double s = 0.0;
            for (int k = 0; k <= i; k++) s += sTp[k][i] * sW[k][jj];
            sW2[i][jj] = s;
        }
        __syncthreads();
        // head rows: X(h, j) -= (Yp_head W2)(h, j)
        for (int e = t; e < QP * QP; e += QA_THREADS) {
            const int h = e & 31, jj = e >> 5;
            double s = sW2[h][jj];
            if (GE)
                for (int k = 0; k < h; k++) s += sY[h][k] * sW2[k][jj];
            head[(size_t)(xc + jj) * ldb + h] -= s;
        }
        // tail rows: X(r, j) -= sum_k Yp(r, k) W2(k, j), as D'[j][r] = sum_k W2(k, j) Yp(r, k).
        // Wave w: columns j0 = 16 (w & 1), row blocks (w >> 1) + 4 m; the next
        // block's operands are loaded before the current block's MFMAs.
        const int nrb = nt >> 4, j0 = 16 * (w & 1);
        double wa[8];
#pragma unroll
        for (int k = 0; k < 8; k++) wa[k] = sW2[4 * k + (l >> 4)][j0 + (l & 15)];
        const double *yp = tail + (size_t)c0 * ldb + (size_t)(l >> 4) * ldb + (l & 15);
        double *xp = tail + (size_t)(xc + j0 + (l >> 4)) * ldb + (l & 15);
        int rb = w >> 1;
        double yb[8], xv[4];
        if (rb < nrb) {
#pragma unroll
            for (int k = 0; k < 8; k++) yb[k] = yp[(size_t)4 * k * ldb + 16 * rb];
#pragma unroll
            for (int v = 0; v < 4; v++) xv[v] = xp[(size_t)4 * v * ldb + 16 * rb];
        }
        for (; rb < nrb; rb += 4) {
            qd4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 8; k++) d = qmfma(wa[k], yb[k], d);
            double xo[4];
#pragma unroll
            for (int v = 0; v < 4; v++) xo[v] = xv[v] - d[v];
            const int rn = rb + 4;
            if (rn < nrb) {
#pragma unroll
                for (int k = 0; k < 8; k++) yb[k] = yp[(size_t)4 * k * ldb + 16 * rn];
#pragma unroll
                for (int v = 0; v < 4; v++) xv[v] = xp[(size_t)4 * v * ldb + 16 * rn];
            }
#pragma unroll
            for (int v = 0; v < 4; v++) xp[(size_t)4 * v * ldb + 16 * rb] = xo[v];
        }
        if (x == 0 && look) {
            // panel p+1 = the chunk this block just updated (visible after the barrier)
            __syncthreads();
            panel_body<GE>(A1, A2, T, tau, b, c0 + QP);
        }
        return;
    }
    // role C (waves 0..3 form the product, all threads finish it)
    const int q = x - na - nbk, cq = QP * pc;
    if (w < 4) {
        qd4 acc = {0.0, 0.0, 0.0, 0.0};
        const int c = l & 15, kk = l >> 4;
        const int ir = 16 * wi + c;                      // row inside chunk q
        const double *trow = T + (size_t)(QP * q + ir);
        const double *gcol = G + (size_t)(cq + 16 * wj + c) * ldb + 4 * kk;
        const int kb = QP * q, m = (cq - kb) >> 4;     // 16-column groups of T's row chunk
        constexpr int D = 2;
        qd4 gb[D];
        double tb[D][4];
#pragma unroll
        for (int u = 0; u < D; u++)
            if (u < m) {
                gb[u] = *(const qd4 *)(gcol + kb + 16 * u);
#pragma unroll
                for (int s = 0; s < 4; s++) tb[u][s] = trow[(size_t)(kb + 16 * u + 4 * kk + s) * ldb];
            }
        for (int r = 0; r < m; r += D) {
#pragma unroll
            for (int u = 0; u < D; u++) {
                if (r + u < m) {
                    const int k0 = kb + 16 * (r + u);
                    const qd4 gv = gb[u];
                    double tv[4];
#pragma unroll
                    for (int s = 0; s < 4; s++) {
                        tv[s] = tb[u][s];
                        if (k0 + 4 * kk + s < kb + ir) tv[s] = 0.0;   // lower part of the diagonal block
                    }
                    if (r + u + D < m) {
                        const int kn = k0 + 16 * D;
                        gb[u] = *(const qd4 *)(gcol + kn);
#pragma unroll
                        for (int s = 0; s < 4; s++) tb[u][s] = trow[(size_t)(kn + 4 * kk + s) * ldb];
                    }
#pragma unroll
                    for (int s = 0; s < 4; s++) acc = qmfma(tv[s], gv[s], acc);
                }
            }
        }
        // D layout: lane l, value v -> Z(16 wi + l / 16 + 4 v, 16 wj + l % 16)
#pragma unroll
        for (int v = 0; v < 4; v++) sW[16 * wi + (l >> 4) + 4 * v][16 * wj + (l & 15)] = acc[v];
    }
    for (int e = t; e < QP * QP; e += QA_THREADS) {
        const int r = e & 31, c = e >> 5;
        sTp[r][c] = r <= c ? T[(size_t)(cq + r) + (size_t)(cq + c) * ldb] : 0.0;
    }
    __syncthreads();
    for (int e = t; e < QP * QP; e += QA_THREADS) {
        const int i = e & 31, jj = e >> 5;
        double s = 0.0;
        for (int k = 0; k <= jj; k++) s += sW[i][k] * sTp[k][jj];
        T[(size_t)(QP * q + i) + (size_t)(cq + jj) * ldb] = -s;
    }
}

// Panel 0 on its own; then launch p (p < P) applies panel p, factors panel
// p+1 in block 0, and forms G(:, p) and the T blocks of panel p-1; the last
// launch forms the T blocks of panel P-1.
hipError_t launch_blocked(bool ge, double *A1, double *A2, double *T, double *tau, double *G, int b, hipStream_t s)
{
    const int P = b / QP;
    if (ge) hipLaunchKernelGGL(qr_panel_kernel<true>, dim3(1), dim3(QP_THREADS), 0, s, A1, A2, T, tau, b, 0);
    else hipLaunchKernelGGL(qr_panel_kernel<false>, dim3(1), dim3(QP_THREADS), 0, s, A1, A2, T, tau, b, 0);
    for (int p = 0; p <= P; p++) {
        const int na = p < P ? P - 1 - p : 0, nbk = p < P ? p : 0, nc = p > 1 ? p - 1 : 0;
        const int look = p + 1 < P ? 1 : 0;      // na >= 1 exactly then: block 0 is chunk 0
        if (na + nbk + nc == 0) continue;
        if (ge)
            hipLaunchKernelGGL(qr_apply_kernel<true>, dim3(na + nbk + nc), dim3(QA_THREADS), 0, s, A1, A2, T, G, tau,
                               b, p, na, nbk, p - 1, look);
        else
            hipLaunchKernelGGL(qr_apply_kernel<false>, dim3(na + nbk + nc), dim3(QA_THREADS), 0, s, A1, A2, T, G,
                               tau, b, p, na, nbk, p - 1, look);
    }
    return hipGetLastError();
}

bool blocked_ok(int b, const double *G) { return G && b % QP == 0 && b >= QP && b <= 512; }

}  // namespace

// V of a QR diagonal tile as a plain operand: ones on the diagonal, zeros
// above it, A's Householder vectors below.  Only the strictly lower part of A
// is read.
__global__ __launch_bounds__(256) void unit_lower_copy_kernel(double *__restrict__ V, const double *__restrict__ A,
                                                              int b)
{
    const size_t bb = (size_t)b * b;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < bb; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % (size_t)b), j = (int)(e / (size_t)b);
        V[e] = i > j ? A[e] : (i == j ? 1.0 : 0.0);
    }
}

hipError_t launch_unit_lower_copy(double *V, const double *A, int b, hipStream_t s)
{
    if (sched_rec_on()) { rec_unit_lower_copy(V, A, s); return hipSuccess; }
    size_t blocks = ((size_t)b * b + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(unit_lower_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, V, A, b);
    return hipGetLastError();
}

hipError_t launch_geqrt(double *A, double *T, double *tau, int b, hipStream_t s, double *G)
{
    if (sched_rec_on()) { rec_qr_panel(A, nullptr, T, tau, G, s); return hipSuccess; }
    if (blocked_ok(b, G)) return launch_blocked(true, A, nullptr, T, tau, G, b, s);
    hipLaunchKernelGGL(geqrt_kernel, dim3(1), dim3(QR_THREADS), (size_t)2 * b * 8, s, A, T, tau, b);
    return hipGetLastError();
}

hipError_t launch_tsqrt(double *A1, double *A2, double *T, double *tau, int b, hipStream_t s, double *G)
{
    if (sched_rec_on()) { rec_qr_panel(A1, A2, T, tau, G, s); return hipSuccess; }
    if (blocked_ok(b, G)) return launch_blocked(false, A1, A2, T, tau, G, b, s);
    hipLaunchKernelGGL(tsqrt_kernel, dim3(1), dim3(QR_THREADS), (size_t)b * 8, s, A1, A2, T, tau, b);
    return hipGetLastError();
}

}  // namespace palial
