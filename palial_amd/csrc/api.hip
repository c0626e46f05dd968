// api.hip -- C ABI of libpalial_hip.so (declared in include/palial_hip.h).
//
// The reference runs each algorithm as an OpenMP task DAG created by one
// master thread (palial_src/main.c:124-159).  Here each DAG step becomes a
// handful of batched launches on two HIP streams:
//   chain stream : the diagonal chain -- factorization of tile (k,k), TRSM of
//                  tile (k,k+1), update of tile (k+1,k+1) by row k, then the
//                  factorization of tile (k+1,k+1);
//   bulk stream  : the TRSMs of the rest of panel k and the trailing update of
//                  every other tile, on CUs masked away from a few reserved
//                  CUs, so the chain's single-workgroup kernels never wait for
//                  a CU to drain.
// Cross-stream dependences are events.  Without a CU reservation the sequence
// is captured once into a hipGraph per (algorithm, N, b, device pointer) and
// replayed (graph replays do not keep stream CU masks, so a reserved schedule
// is launched eagerly).
#include "internal.hpp"
#include "gemm_common.hpp"
#include "../../include/palial_hip.h"

#include <algorithm>
#include <cmath>
#include <array>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <sys/time.h>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

using namespace palial;

// Workspace of palial_hip_dtsmqr_noalloc: the two b x b tiles W = A1 + V^T A2
// and T^T W of one TSMQR, b <= NOALLOC_MAXB, held by the code object itself
// (4 MiB per device, for the life of the process) so that the entry point
// allocates nothing.  Host code never passes `scratch` itself to a kernel:
// palial_hip_init resolves its device address per device (g_scratch below).
constexpr int NOALLOC_MAXB = 512;
static __device__ __align__(256) double scratch[2 * (size_t)NOALLOC_MAXB * NOALLOC_MAXB];

namespace {

// device address of `scratch` on each device (hipGetSymbolAddress with that
// device current, at palial_hip_init); kept until the process ends
constexpr int MAX_DEVICES = 64;
double *g_scratch[MAX_DEVICES] = {};

struct Graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long nodes = 0;
};

struct BatchGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int count = 0, f32 = 0, bmax = 0, grid = 0;
    const void *A = nullptr, *B = nullptr, *idx = nullptr, *work = nullptr;
};

struct Ctx {
    bool init = false;
    int device = -1;
    hipStream_t s_main = nullptr;   // update stream (and the caller-visible stream)
    hipStream_t s_panel = nullptr;  // look-ahead panel / diagonal-chain stream
    hipStream_t s_bulk = nullptr;   // bulk stream of the chain schedule: CU-masked away from
                                    // the reserved CUs (s_main when nothing is reserved)
    hipStream_t s_aux = nullptr;    // chain schedule with the split update: the row/column
                                    // TRSMs and panel packing (CU mask of s_bulk)
    hipStream_t s_chainm = nullptr; // PALIAL_CHAIN_CUS=1: the diagonal chain of the mixed tiers
                                    // masked to the reserved CUs (nullptr: s_panel, all CUs)
    hipStream_t s_qr[8] = {};       // QR panel-chain streams (all CUs)
    hipStream_t s_qrb[8] = {};      // QR update streams: CU-masked away from the first qr_reserve CUs
    int qr_streams = 8;             // QR: panel chains / update streams in use (1..8)
    int qr_reserve = 0;             // QR: CUs kept for the panel chains
    bool qr_up = false;             // s_qr / s_qrb exist (created by the first QR factorization)
    int ncu = 0;
    int reserve = 0;                // CUs kept free of bulk work for the diagonal chain
    int dag = 2;                    // 2: diagonal chain + split update, 1: diagonal chain,
                                    // 0: row look-ahead
    std::vector<hipEvent_t> events;
    int *d_info = nullptr;
    double *d_tau = nullptr;        // QR tau scratch (b doubles per panel-chain stream)
    size_t tau_cap = 0;
    double *d_w1 = nullptr, *d_w2 = nullptr;  // QR/ORMQR workspace tiles (MT tiles per bulk stream)
    double *d_qg = nullptr;                    // b x b Gram workspace of the blocked QR panels, per chain
    size_t qg_cap = 0;
    size_t w_cap = 0;
    double *d_dinv = nullptr;       // inverses of the 32x32 diagonal blocks (U then L)
    size_t dinv_cap = 0;
    double *d_sub = nullptr;        // one diagonal tile as sub-tiles (potrf_tile)
    size_t sub_cap = 0;             // doubles
    int subtile = 1;                // factor diagonal tiles of b >= 256 as 128-sub-tiles
    int f32work = 1;                // mixed-tier solves factor an FP32 working copy
    void *d_pack = nullptr;         // low-precision packed panels of PACK_SETS steps (U then L each)
    size_t pack_cap = 0;            // bytes
    // solve workspace: factor copy F, diagonal-tile inverses, vectors
    double *d_F = nullptr;
    size_t F_cap = 0;
    double *d_inv = nullptr;
    size_t inv_cap = 0;
    double *d_blk = nullptr;        // block inverses of the blocked sweeps: tiles of T and products
    size_t blk_cap = 0;
    int sweep_block = 1;            // default tiles per correction-sweep step
    int pair = 0;                   // BF16 / split-BF16 Cholesky and LU solves: paired trailing updates
    BatchGraph bgraph;              // batched factorizations: the captured persistent launch
    void *d_bidx = nullptr;         // batched factorizations: per-matrix index arrays
    size_t bidx_cap = 0;            // bytes
    double *d_bwork = nullptr;      // batched factorizations: padded FP64 workspace
    size_t bwork_cap = 0;           // doubles
    double *d_gm = nullptr;         // GMRES-IR: Krylov basis and vectors
    size_t gm_cap = 0;
    double *d_F32 = nullptr;        // FP32 copies of the factor and the tile inverses
    size_t F32_cap = 0;             // (in doubles: the buffer is reused as floats)
    double *d_vec = nullptr;
    size_t vec_cap = 0;
    double *d_aw = nullptr;         // Q application: W and op(T) W tiles, one pair per column tile of C
    size_t aw_cap = 0;
    double *d_qs = nullptr;         // QR solve: right-hand-side tiles, vectors, R tile inverses
    size_t qs_cap = 0;
    std::map<std::tuple<int, void *, void *, int, int, int>, Graph> graphs;
};

Ctx g;
std::mutex g_mu;
thread_local std::string g_last;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last = buf;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(PALIAL_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                                \
    } while (0)

palial_opts default_opts()
{
    palial_opts o;
    memset(&o, 0, sizeof o);
    return o;
}

int ensure_init()
{
    if (g.init) return 0;
    return fail(PALIAL_ERR_STATE, "palial_hip_init() has not been called");
}

int ensure_events(size_t n)
{
    while (g.events.size() < n) {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        g.events.push_back(e);
    }
    return 0;
}

// captured factorizations kept at most (each holds its schedule's launches)
constexpr size_t MAX_GRAPHS = 16;

void drop_factor_graphs()
{
    for (auto &kv : g.graphs) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    g.graphs.clear();
}

void drop_graphs()
{
    drop_factor_graphs();
    if (g.bgraph.exec) (void)hipGraphExecDestroy(g.bgraph.exec);
    if (g.bgraph.graph) (void)hipGraphDestroy(g.bgraph.graph);
    g.bgraph = BatchGraph();
}

constexpr int QR_MAXS = 8;     // most panel-chain (and update) streams of the QR schedule

int ensure_ws(int b, int MT)
{
    const int QR_CHAINS = g.qr_streams, QR_BULK = g.qr_streams;
    const size_t need = (size_t)QR_BULK * b * b * (size_t)(MT > 1 ? MT : 1);
    if (g.w_cap < need || g.tau_cap < (size_t)QR_CHAINS * b) drop_graphs();   // captured graphs hold the old pointers
    if (g.w_cap < need) {
        if (g.d_w1) (void)hipFree(g.d_w1);
        if (g.d_w2) (void)hipFree(g.d_w2);
        g.d_w1 = g.d_w2 = nullptr;
        g.w_cap = 0;
        if (hipMalloc(&g.d_w1, need * 8) != hipSuccess || hipMalloc(&g.d_w2, need * 8) != hipSuccess)
            return fail(PALIAL_ERR_NOMEM, "workspace allocation of %zu bytes failed", 2 * need * 8);
        g.w_cap = need;
    }
    if (g.qg_cap < (size_t)QR_CHAINS * b * b) {
        drop_graphs();
        if (g.d_qg) (void)hipFree(g.d_qg);
        g.d_qg = nullptr;
        g.qg_cap = 0;
        if (hipMalloc(&g.d_qg, (size_t)QR_CHAINS * b * b * 8) != hipSuccess)
            return fail(PALIAL_ERR_NOMEM, "QR Gram workspace of %zu bytes failed", (size_t)QR_CHAINS * b * b * 8);
        g.qg_cap = (size_t)QR_CHAINS * b * b;
    }
    if (g.tau_cap < (size_t)QR_CHAINS * b) {
        if (g.d_tau) (void)hipFree(g.d_tau);
        g.d_tau = nullptr;
        g.tau_cap = 0;
        if (hipMalloc(&g.d_tau, (size_t)QR_CHAINS * b * 8) != hipSuccess)
            return fail(PALIAL_ERR_NOMEM, "tau alloc failed");
        g.tau_cap = (size_t)QR_CHAINS * b;
    }
    return 0;
}

// inverses of the 32x32 diagonal blocks of every diagonal tile: tile k holds
// U blocks at k*64*b and L blocks at k*64*b + 32*b (doubles)
int ensure_dinv(int b, int MT)
{
    const size_t need = (size_t)MT * 64 * b;
    if (g.dinv_cap < need) {
        drop_graphs();
        if (g.d_dinv) (void)hipFree(g.d_dinv);
        g.d_dinv = nullptr;
        if (hipMalloc(&g.d_dinv, need * 8) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "dinv alloc failed");
        g.dinv_cap = need;
    }
    const size_t sub = (size_t)b * b;
    if (g.subtile && b >= 256 && b % 128 == 0 && g.sub_cap < sub) {
        drop_graphs();
        if (g.d_sub) (void)hipFree(g.d_sub);
        g.d_sub = nullptr;
        g.sub_cap = 0;
        if (hipMalloc(&g.d_sub, sub * 8) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "sub-tile alloc failed");
        g.sub_cap = sub;
    }
    return 0;
}


// packed low-precision panel operands of one step (a set): row panel at 0,
// column panel at MT*b*b*4 bytes, 4 bytes per element covering FP32, BF16,
// split BF16 and FP8.  The split schedule packs step k into set k % PACK_SETS
// (paired steps read two sets while the next two are packed).
size_t pack_set_bytes(int b, int MT) { return 2 * (size_t)MT * b * b * 4; }
constexpr int PACK_SETS = 4;

int ensure_pack(int b, int MT)
{
    // the step sets, then the one or two packed tiles of the chain's S1 (FP32 working copy)
    const size_t need = PACK_SETS * pack_set_bytes(b, MT) + 2 * (size_t)b * b * 4;
    if (g.pack_cap < need) {
        drop_graphs();
        if (g.d_pack) (void)hipFree(g.d_pack);
        g.d_pack = nullptr;
        g.pack_cap = 0;
        if (hipMalloc(&g.d_pack, need) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "packed panel alloc failed");
        g.pack_cap = need;
    }
    return 0;
}

int check_desc(const MATRIX_desc &A)
{
    if (!A.matrix) return fail(PALIAL_ERR_ARG, "descriptor has a NULL matrix pointer");
    if (A.tile_size <= 0 || A.matrix_size <= 0)
        return fail(PALIAL_ERR_ARG, "bad sizes N=%d b=%d", A.matrix_size, A.tile_size);
    if (A.matrix_size % A.tile_size != 0)
        return fail(PALIAL_ERR_ARG, "Tile size does not divide the matrix size (N=%d b=%d)", A.matrix_size,
                    A.tile_size);
    if (A.tile_size > 1024) return fail(PALIAL_ERR_ARG, "tile size %d above the supported 1024", A.tile_size);
    return 0;
}

// ------------------------------------------------------------------ profiling
// Per-kernel-class HIP-event timing (stream-ordered launches only): each
// launch of a class is bracketed by two events on the stream it runs on.
// KC_PKGEMM: trailing updates from packed low-precision panels; their bytes
// are the algorithmic HBM traffic (FP64 C read + write, the packed panel once)
enum KClass : int { KC_GEMM = 0, KC_DIAG = 1, KC_TRSM = 2, KC_TSQRT = 3, KC_QRUPD = 4, KC_PACK = 5, KC_PKGEMM = 6,
                    KC_N = 7 };
struct ProfRec { int cls; double flops; double bytes; int b, prec; hipEvent_t a, b_ev; };
struct Prof {
    bool on = false;
    bool keep = false;                    // also keep one record per launch
    std::vector<ProfRec> pending;
    std::vector<std::array<double, 6>> records;   // class, b, precision, flops, bytes, ms
    long launches[KC_N] = {0};
    double ms[KC_N] = {0}, flops[KC_N] = {0}, bytes[KC_N] = {0};
};
Prof g_prof;

// ------------------------------------------------------------------ per-task trace
// palial_hip_trace(1): the next factorizations (plain or inside a solve)
// record the device time of every reference tile task.  Batched launches on
// the factorization's matrix get timestamp slots (trace_gemm / trace_trsm,
// filled in by the kernels); the single-task launches of the diagonal chain
// and the QR panels get a pair of HIP events.  trace_collect maps each launch
// task back to its reference task with the kernels' own task decode.
struct TraceLaunch {
    bool gemm;
    GemmArgs g;
    int kind;
    TrsmArgs t;
};
struct TraceSingle {
    int cls;                  // KC_DIAG or KC_TSQRT
    hipEvent_t a, e;
};
struct TaskTrace {
    bool on = false;          // palial_hip_trace(1)
    bool active = false;      // a traced factorization is being issued
    const double *mat = nullptr;
    int N = 0, b = 0, MT = 0, algo = -1;
    unsigned long long *d_ts = nullptr;   // slot pairs; the last pair's start holds the t0 tick
    size_t cap = 0;                       // slots
    long next = 0;
    bool overflow = false;
    std::vector<TraceLaunch> launches;
    std::vector<TraceSingle> singles;
    hipEvent_t t0 = nullptr;
    double host_t0 = 0.0;                 // gettimeofday at the start of the factorization (s)
    std::vector<std::array<double, 6>> records;   // creation order: kind, k, m, n, ms, start ms
};
TaskTrace g_tr;

// ------------------------------------------------------------------ DAGs
struct Launcher {
    long launches = 0;
    hipError_t err = hipSuccess;
    bool capturing = false;
    int b = 0, prec = 0;   // context of the records: tile size and precision tier
    void chk(hipError_t e)
    {
        launches++;
        if (e != hipSuccess && err == hipSuccess) err = e;
    }
    void note(hipError_t e)      // a runtime call that is not a kernel launch
    {
        if (e != hipSuccess && err == hipSuccess) err = e;
    }
    // run one kernel launch of class cls with its algorithmic flops / bytes
    template <class F>
    void run(int cls, double flops, double bytes, hipStream_t s, F &&f)
    {
        const bool prof = g_prof.on && !capturing;
        const bool tr = g_tr.active && !capturing && (cls == KC_DIAG || cls == KC_TSQRT);
        ProfRec r{cls, flops, bytes, b, prec, nullptr, nullptr};
        TraceSingle ts{cls, nullptr, nullptr};
        if (prof) {
            note(hipEventCreate(&r.a));
            note(hipEventCreate(&r.b_ev));
            note(hipEventRecord(r.a, s));
        }
        if (tr) {
            note(hipEventCreate(&ts.a));
            note(hipEventCreate(&ts.e));
            note(hipEventRecord(ts.a, s));
        }
        chk(f());
        if (prof) {
            note(hipEventRecord(r.b_ev, s));
            g_prof.pending.push_back(r);
        }
        if (tr) {
            note(hipEventRecord(ts.e, s));
            g_tr.singles.push_back(ts);
        }
    }
};

void prof_collect()
{
    for (auto &r : g_prof.pending) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, r.a, r.b_ev);
        g_prof.launches[r.cls]++;
        g_prof.ms[r.cls] += ms;
        g_prof.flops[r.cls] += r.flops;
        g_prof.bytes[r.cls] += r.bytes;
        if (g_prof.keep) g_prof.records.push_back({(double)r.cls, (double)r.b, (double)r.prec, r.flops, r.bytes, (double)ms});
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b_ev);
    }
    g_prof.pending.clear();
}

// ------------------------------------------------------------------ per-task trace
// Reference tile tasks in creation order with the names the reference's task
// loops give them: the kind, then k, m and n printed back to back, m and n
// being whatever the loop variables hold when the task is created
// (src/cholesky.h:14-173, src/lu.h:13-173, src/qr.h:17-185; the k = 0 tasks
// of LU and QR print variables the reference leaves uninitialised, here 0).
// Kinds: Cholesky 0 potrf(k), 1 trsm(k,m) [tile (k,m)], 2 syrk(k,m) [tile
// (m,m)], 3 gemm(k,m,n) [tile (n,m)]; LU 4 getrf(k), 5 trsm(k,m) [tile
// (m,k)], 6 trsm(k,n) [tile (k,n)], 7 gemm(k,m,n) [tile (m,n)]; QR 8
// geqrt(k), 9 ormqr(k,n), 10 tsqrt(k,m), 11 tsmqr(k,m,n).  Unused indices -1.
struct RefTask {
    int kind, k, m, n;
    char name[32];
};
enum TaskKind : int { TK_POTRF = 0, TK_CTRSM, TK_SYRK, TK_CGEMM, TK_GETRF, TK_TRSML, TK_TRSMU, TK_LGEMM,
                      TK_GEQRT, TK_ORMQR, TK_TSQRT, TK_TSMQR };
const char *const TK_NAME[] = {"potrf", "trsm", "syrk", "gemm", "getrf", "trsm", "trsm", "gemm",
                               "geqrt", "ormqr", "tsqrt", "tsmqr"};

void ref_tasks(int algo, int MT, std::vector<RefTask> &out)
{
    out.clear();
    int k = 0, m = 0, n = 0;
    auto add = [&](int kind, int kk, int mm, int nn) {
        RefTask t{kind, kk, mm, nn, {0}};
        snprintf(t.name, sizeof t.name, "%s%d%d%d", TK_NAME[kind], k, m, n);
        out.push_back(t);
    };
    for (k = 0; k < MT; k++) {
        if (algo == 0) {
            add(TK_POTRF, k, -1, -1);
            for (m = k + 1; m < MT; m++) add(TK_CTRSM, k, m, -1);
            for (m = k + 1; m < MT; m++) {
                add(TK_SYRK, k, m, -1);
                for (n = k + 1; n < m; n++) add(TK_CGEMM, k, m, n);
            }
        } else if (algo == 1) {
            add(TK_GETRF, k, -1, -1);
            for (m = k + 1; m < MT; m++) add(TK_TRSML, k, m, -1);
            for (n = k + 1; n < MT; n++) {
                add(TK_TRSMU, k, -1, n);
                for (m = k + 1; m < MT; m++) add(TK_LGEMM, k, m, n);
            }
        } else {
            add(TK_GEQRT, k, -1, -1);
            for (n = k + 1; n < MT; n++) add(TK_ORMQR, k, -1, n);
            for (m = k + 1; m < MT; m++) {
                add(TK_TSQRT, k, m, -1);
                for (n = k + 1; n < MT; n++) add(TK_TSMQR, k, m, n);
            }
        }
    }
}

// Start tracing the factorization of mat that is about to be issued on s.
void trace_begin(double *mat, int N, int b, int algo, hipStream_t s, Launcher &L)
{
    if (!g_tr.on) return;
    for (auto &x : g_tr.singles) {
        (void)hipEventDestroy(x.a);
        (void)hipEventDestroy(x.e);
    }
    g_tr.singles.clear();
    g_tr.launches.clear();
    g_tr.records.clear();
    g_tr.mat = mat;
    g_tr.N = N; g_tr.b = b; g_tr.MT = N / b; g_tr.algo = algo;
    g_tr.next = 0;
    g_tr.overflow = false;
    const size_t MT = (size_t)(N / b);
    // every reference task once per launch covering it (QR updates: 3 phases)
    const size_t need = 3 * MT * MT * MT / 2 + 4 * MT * MT + 64;
    if (g_tr.cap < need) {
        if (g_tr.d_ts) (void)hipFree(g_tr.d_ts);
        g_tr.d_ts = nullptr;
        g_tr.cap = 0;
        if (hipMalloc(&g_tr.d_ts, (need + 1) * 16) != hipSuccess) {
            L.note(hipErrorOutOfMemory);
            return;
        }
        g_tr.cap = need;
    }
    if (!g_tr.t0) L.note(hipEventCreate(&g_tr.t0));
    L.note(launch_trace_init(g_tr.d_ts, g_tr.cap, s));
    L.note(hipEventRecord(g_tr.t0, s));
    L.note(launch_clock_stamp(g_tr.d_ts + 2 * g_tr.cap, s));
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    g_tr.host_t0 = tv.tv_sec + 1e-6 * tv.tv_usec;
    g_tr.active = true;
}

void trace_end() { g_tr.active = false; }

}  // namespace

// the reference's trace hooks (include/common.h:9-11), called only when a
// tracer that defines them is loaded
extern "C" {
void palial_set_task_name(const char *name) __attribute__((weak));
void palial_set_task_cpu_node(int cpu, int node, char *name) __attribute__((weak));
void palial_get_task_time(struct timeval start, struct timeval end, char *name) __attribute__((weak));
}

namespace palial {

// ------------------------------------------------------------------ schedule recorder
// Resource ids: matrix tile m + MT n; RES + packed panel tile (set, half,
// index); 2 RES + diagonal-block inverses (2k: U side, 2k+1: L side); then
// the QR buffers by address region: 3 RES + T-factor tile, 4 RES / 5 RES +
// W / op(T) W workspace tile (per update stream), 6 RES + tau and 7 RES + Gram
// workspace of a panel chain.
// Record of one operation (longs): type (0 launch, 1 event record, 2 event
// wait), stream, event, kind (1 diagonal factor, 2 TRSM, 3 GEMM, 4 pack,
// 5 QR panel), mode (GEMM mode; QR panel: 0 GEQRT, 1 TSQRT), k, m, nr, nw,
// nx, the nr read ids, the nw written ids, then nx groups of 4: GEMM (C tile,
// step, A operand id, B operand id), TRSM (tile, step, diagonal tile,
// inverses), pack (tile, step, written id, -1), diagonal factor (tile, step,
// inverses, inverses or -1).
constexpr long REC_RES = 1L << 24;
struct RecRegion {
    const char *base;
    size_t bytes, unit;   // region size and bytes per resource
    long id0;
};
struct SchedRec {
    bool on = false;
    const char *mat = nullptr;    // matrix base and the element size of its tiles
    int es = 8, b = 0, MT = 0;
    const char *pack = nullptr;   // packed panel sets
    size_t psb = 0;
    const char *dinv = nullptr;
    std::vector<RecRegion> regions;   // further address regions (QR buffers)
    std::vector<long> out;
};
SchedRec g_rec;

bool sched_rec_on() { return g_rec.on; }

long rec_tile_elem(long eoff) { return eoff / ((long)g_rec.b * g_rec.b); }
long rec_tile_ptr(const void *p) { return rec_tile_elem((long)((const char *)p - g_rec.mat) / g_rec.es); }
long rec_pack_id(const void *p, int idx)
{
    const size_t off = (size_t)((const char *)p - g_rec.pack);
    const size_t half_bytes = (size_t)g_rec.MT * g_rec.b * g_rec.b * 4;
    const size_t set = off / g_rec.psb, rem = off % g_rec.psb;
    const long half = (long)(rem / half_bytes), first = (long)((rem % half_bytes) / ((size_t)g_rec.b * g_rec.b * 4));
    return REC_RES + (long)set * 2 * g_rec.MT + half * g_rec.MT + first + idx;
}
long rec_dinv_id(const void *p) { return 2 * REC_RES + (long)(((const char *)p - g_rec.dinv) / 8 / (32L * g_rec.b)); }
// any recorded address: a matrix tile, or a resource of one of the regions
long rec_id(const void *p)
{
    const char *c = (const char *)p;
    for (const RecRegion &r : g_rec.regions)
        if (c >= r.base && c < r.base + r.bytes) return r.id0 + (long)((size_t)(c - r.base) / r.unit);
    return rec_tile_ptr(p);
}

void rec_put(int type, hipStream_t s, hipEvent_t e, int kind, const std::vector<long> &rd,
             const std::vector<long> &wr, const std::vector<long> &x, int mode = 0, int k = -1, int m = -1)
{
    std::vector<long> &o = g_rec.out;
    o.push_back(type);
    o.push_back((long)(intptr_t)s);
    o.push_back((long)(intptr_t)e);
    o.push_back(kind);
    o.push_back(mode);
    o.push_back(k);
    o.push_back(m);
    o.push_back((long)rd.size());
    o.push_back((long)wr.size());
    o.push_back((long)x.size() / 4);
    o.insert(o.end(), rd.begin(), rd.end());
    o.insert(o.end(), wr.begin(), wr.end());
    o.insert(o.end(), x.begin(), x.end());
}

void rec_gemm(const GemmArgs &a, bool packed, hipStream_t s)
{
    std::vector<long> rd, wr, x;
    for (int t = 0; t < a.ntasks; t++) {
        const TaskPtrs tp = decode(a, t);
        if (!tp.valid) continue;
        // matrix operands as element offsets (the FP32 working copy is addressed
        // through the FP64 view of its base), the rest by address region
        auto id = [&](const double *ptr) -> long {
            if (!ptr) return -1;
            const bool in_mat = (const char *)ptr >= g_rec.mat && (const char *)ptr < g_rec.mat + ((size_t)1 << 39);
            if (in_mat && a.mode != GM_SINGLE) return rec_tile_elem((long)(ptr - a.mat));
            return rec_id(ptr);
        };
        const long c = id(tp.C);
        if (tp.D) rd.push_back(id(tp.D));
        wr.push_back(c);
        if (tp.E) {
            rd.push_back(id(tp.E));
            wr.push_back(id(tp.E));
        }
        long ta = id(tp.A);
        const long tb = id(tp.B);
        // the unit-lower V of a QR diagonal tile is a resource of its own: the
        // TSQRTs of step k rewrite only the tile's upper R while ORMQR(k) reads V
        if (a.maskA == MK_UNIT_LOWER && ta >= 0 && ta < REC_RES && ta % g_rec.MT == ta / g_rec.MT) ta += 8 * REC_RES;
        // the step of a product: the smaller index of its A tile (Cholesky U(k, n),
        // LU L(m, k)); mode launches carry it as a.k
        const int kk = a.mode == GM_SINGLE ? (int)std::min(ta % g_rec.MT, ta / g_rec.MT) : a.k;
        if (packed) {
            const long pa = rec_pack_id(a.pa, tp.ia), pb = rec_pack_id(a.pb, tp.ib);
            rd.push_back(pa);
            rd.push_back(pb);
            x.insert(x.end(), {c, kk, pa, pb});
            if (a.pa2) {
                const long pa2 = rec_pack_id(a.pa2, tp.ia - 1), pb2 = rec_pack_id(a.pb2, tp.ib - 1);
                rd.push_back(pa2);
                rd.push_back(pb2);
                x.insert(x.end(), {c, (long)kk + 1, pa2, pb2});
            }
        } else {
            rd.push_back(ta);
            rd.push_back(tb);
            x.insert(x.end(), {c, kk, ta, tb});
        }
    }
    rec_put(0, s, nullptr, 3, rd, wr, x, a.mode, a.k, a.m);
}

void rec_trsm(int kind, const TrsmArgs &t, const double *dinv, hipStream_t s)
{
    std::vector<long> rd, wr, x;
    const long d = rec_tile_ptr(t.diag), di = rec_dinv_id(dinv);
    rd.push_back(d);
    rd.push_back(di);
    for (int i = 0; i < t.ntiles; i++) {
        const int idx = t.first + i;
        const long tl = kind == TR_LUR ? (long)idx + (long)t.MT * t.k : (long)t.k + (long)t.MT * idx;
        rd.push_back(tl);
        wr.push_back(tl);
        x.insert(x.end(), {tl, (long)t.k, d, di});
    }
    rec_put(0, s, nullptr, 2, rd, wr, x, kind, t.k, -1);
}

void rec_pack(const double *mat, int b, int k, int first, int n, bool colpanel, const void *dst, hipStream_t s)
{
    (void)mat;
    (void)b;
    std::vector<long> rd, wr, x;
    for (int i = 0; i < n; i++) {
        const long tl = colpanel ? (long)(first + i) + (long)g_rec.MT * k : (long)k + (long)g_rec.MT * (first + i);
        const long pid = rec_pack_id(dst, i);
        rd.push_back(tl);
        wr.push_back(pid);
        x.insert(x.end(), {tl, (long)k, pid, -1L});
    }
    rec_put(0, s, nullptr, 4, rd, wr, x, colpanel ? 1 : 0, k, -1);
}

// QR panels: GEQRT (A2 == nullptr) of tile A1, or TSQRT of [A1; A2]; T
// factor, tau and Gram workspace written
void rec_qr_panel(const double *A1, const double *A2, const double *T, const double *tau, const double *G,
                  hipStream_t s)
{
    std::vector<long> rd, wr;
    const long a1 = rec_id(A1);
    rd.push_back(a1);
    wr.push_back(a1);
    if (!A2) wr.push_back(a1 + 8 * REC_RES);    // GEQRT writes V (below the diagonal) as well
    if (A2) {
        rd.push_back(rec_id(A2));
        wr.push_back(rec_id(A2));
    }
    wr.push_back(rec_id(T));
    wr.push_back(rec_id(tau));
    if (G) wr.push_back(rec_id(G));
    const long t2 = A2 ? rec_id(A2) : a1;
    rec_put(0, s, nullptr, 5, rd, wr, {}, A2 ? 1 : 0, (int)(a1 / g_rec.MT), (int)(t2 % g_rec.MT));
}

// V of diagonal tile A (its own resource, as in rec_gemm) copied to workspace V
void rec_unit_lower_copy(const double *V, const double *A, hipStream_t s)
{
    const long a = rec_id(A);
    rec_put(0, s, nullptr, 6, {a + 8 * REC_RES}, {rec_id(V)}, {}, 0, (int)(a % g_rec.MT), -1);
}

// PALIAL_CU_POLICY: CU sets of the mixed tiers' task classes, entries
// "name=ranges" separated by ',' with name bulk (trailing updates), aux (row
// TRSMs and packing) or chain (diagonal chain), ranges "lo-hi" or "cu" joined
// by '+'.  Unnamed classes: bulk all CUs, aux the bulk's set, chain unmasked
// (given[2] false).  Returns 0, or -1 for a malformed policy.
int parse_cu_policy(const char *pol, int ncu, std::vector<uint32_t> (&m)[3], bool (&given)[3])
{
    const int words = (ncu + 31) / 32;
    static const char *names[3] = {"bulk", "aux", "chain"};
    for (int c = 0; c < 3; c++) {
        m[c].assign((size_t)words, 0u);
        given[c] = false;
    }
    if (!pol || !*pol || ncu <= 0) return -1;
    const char *p = pol;
    while (*p) {
        int c = -1;
        for (int i = 0; i < 3; i++) {
            const size_t n = strlen(names[i]);
            if (!strncmp(p, names[i], n) && p[n] == '=') { c = i; p += n + 1; break; }
        }
        if (c < 0 || given[c]) return -1;
        given[c] = true;
        for (;;) {
            char *e;
            const long lo = strtol(p, &e, 10);
            if (e == p) return -1;
            long hi = lo;
            p = e;
            if (*p == '-') {
                hi = strtol(p + 1, &e, 10);
                if (e == p + 1) return -1;
                p = e;
            }
            if (lo < 0 || hi < lo || hi >= ncu) return -1;
            for (long cu = lo; cu <= hi; cu++) m[c][(size_t)cu / 32] |= 1u << (cu % 32);
            if (*p != '+') break;
            p++;
        }
        if (*p == ',') {
            if (!*++p) return -1;      // a trailing separator
        } else if (*p) {
            return -1;
        }
    }
    if (!given[0])
        for (int cu = 0; cu < ncu; cu++) m[0][(size_t)cu / 32] |= 1u << (cu % 32);
    if (!given[1]) m[1] = m[0];
    return 0;
}

void trace_gemm(GemmArgs &a)
{
    if (!g_tr.active) return;
    const double *base = g_tr.mat;
    const size_t nn = (size_t)g_tr.N * g_tr.N;
    const bool mine = a.mode == GM_SINGLE ? (a.C >= base && a.C < base + nn) : a.mat == base;
    if (!mine || a.ntasks <= 0) return;
    if ((size_t)(g_tr.next + a.ntasks) > g_tr.cap) {
        g_tr.overflow = true;
        return;
    }
    a.ts = g_tr.d_ts;
    a.ts_base = g_tr.next;
    g_tr.next += a.ntasks;
    TraceLaunch r;
    memset(&r, 0, sizeof r);
    r.gemm = true;
    r.g = a;
    g_tr.launches.push_back(r);
}

void trace_trsm(int kind, TrsmArgs &a)
{
    if (!g_tr.active || a.batch || a.mat != g_tr.mat || a.ntiles <= 0) return;
    if ((size_t)(g_tr.next + a.ntiles) > g_tr.cap) {
        g_tr.overflow = true;
        return;
    }
    a.ts = g_tr.d_ts;
    a.ts_base = g_tr.next;
    g_tr.next += a.ntiles;
    TraceLaunch r;
    memset(&r, 0, sizeof r);
    r.gemm = false;
    r.kind = kind;
    r.t = a;
    g_tr.launches.push_back(r);
}

}  // namespace palial

namespace {

// After the traced factorization has completed: per reference task, the sum
// of the device durations of the launch tasks that worked on it (a QR update
// is three launches) and the earliest start, in ms from the start of the
// factorization; -1 for a task nothing measured.  Calls the reference's trace
// hooks when a tracer defines them.
int trace_collect()
{
    if (!g_tr.on || g_tr.algo < 0) return 0;
    const int MT = g_tr.MT, b = g_tr.b;
    const size_t nslot = g_tr.cap + 1;
    std::vector<unsigned long long> ts(2 * nslot);
    HIPCHK(hipMemcpy(ts.data(), g_tr.d_ts, ts.size() * 8, hipMemcpyDeviceToHost));
    const unsigned long long tick0 = ts[2 * g_tr.cap];
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g.device));
    const double tick_ms = khz > 0 ? 1.0 / khz : 1e-5;          // 100 MHz when the query fails
    std::map<std::array<int, 4>, std::array<double, 2>> acc;     // (kind,k,m,n) -> (ms, start)
    auto add = [&](int kind, int k, int m, int n, double ms, double start) {
        auto &v = acc.emplace(std::array<int, 4>{kind, k, m, n}, std::array<double, 2>{0.0, 1e300}).first->second;
        v[0] += ms;
        v[1] = std::min(v[1], start);
    };
    const double *base = g_tr.mat;
    auto tile_of = [&](const double *p, int &r, int &c) {
        const size_t off = (size_t)(p - base) / ((size_t)b * b);
        r = (int)(off % MT);
        c = (int)(off / MT);
    };
    for (const auto &L : g_tr.launches) {
        const long base_slot = L.gemm ? L.g.ts_base : L.t.ts_base;
        const int nt = L.gemm ? L.g.ntasks : L.t.ntiles;
        for (int t = 0; t < nt; t++) {
            const unsigned long long st = ts[2 * (base_slot + t)], en = ts[2 * (base_slot + t) + 1];
            if (st == ~0ull || en < st) continue;
            const double ms = (double)(en - st) * tick_ms, start = ((double)st - (double)tick0) * tick_ms;
            if (!L.gemm) {
                const int k = L.t.k, idx = L.t.first + t;
                if (L.kind == TR_CHOL) add(TK_CTRSM, k, idx, -1, ms, start);
                else if (L.kind == TR_LUR) add(TK_TRSML, k, idx, -1, ms, start);
                else add(TK_TRSMU, k, -1, idx, ms, start);
                continue;
            }
            const GemmArgs &a = L.g;
            const int tg = a.task0 + t;
            switch (a.mode) {
            case GM_QR_W: case GM_QR_TW: case GM_QR_VW:
                add(TK_TSMQR, a.k, a.m, a.k + 1 + tg, ms, start);
                break;
            case GM_ORM_W: case GM_ORM_TW: case GM_ORM_VW:
                add(TK_ORMQR, a.k, -1, a.k + 1 + tg, ms, start);
                break;
            default: {
                const TaskPtrs tp = decode(a, t);
                int r, c, ar, ac;
                tile_of(tp.C, r, c);
                tile_of(tp.A, ar, ac);
                if (g_tr.algo == 0) {
                    const int k = a.mode == GM_SINGLE ? ar : a.k;    // A = tile (k, n)
                    if (r == c) add(TK_SYRK, k, c, -1, ms, start);
                    else add(TK_CGEMM, k, c, r, ms, start);
                } else {
                    const int k = a.mode == GM_SINGLE ? ac : a.k;    // A = tile (m, k)
                    add(TK_LGEMM, k, r, c, ms, start);
                }
            }
            }
        }
    }
    // single-task launches in issue order: diagonal tasks k = 0.. (QR: GEQRT
    // k), TSQRT (k, m) in (k, m) order
    int kd = 0, kq = 0, mq = 1;
    for (const auto &x : g_tr.singles) {
        float ms = -1, st = 0;
        if (hipEventElapsedTime(&ms, x.a, x.e) != hipSuccess) ms = -1;
        if (hipEventElapsedTime(&st, g_tr.t0, x.a) != hipSuccess) st = 0;
        if (x.cls == KC_DIAG) {
            add(g_tr.algo == 0 ? TK_POTRF : g_tr.algo == 1 ? TK_GETRF : TK_GEQRT, kd, -1, -1, ms, st);
            kd++;
        } else {
            if (mq >= MT) { kq++; mq = kq + 1; }
            add(TK_TSQRT, kq, mq, -1, ms, st);
            mq++;
        }
    }
    std::vector<RefTask> tasks;
    ref_tasks(g_tr.algo, MT, tasks);
    g_tr.records.clear();
    for (const auto &t : tasks) {
        auto it = acc.find({t.kind, t.k, t.m, t.n});
        const double ms = it == acc.end() ? -1.0 : it->second[0];
        const double start = it == acc.end() ? -1.0 : it->second[1];
        g_tr.records.push_back({(double)t.kind, (double)t.k, (double)t.m, (double)t.n, ms, start});
    }
    if (palial_set_task_name && palial_set_task_cpu_node && palial_get_task_time) {
        for (size_t i = 0; i < tasks.size(); i++) {
            const auto &r = g_tr.records[i];
            if (r[4] < 0) continue;
            char name[32];
            snprintf(name, sizeof name, "%s", tasks[i].name);
            auto tv = [](double sec) {
                struct timeval v;
                v.tv_sec = (time_t)sec;
                v.tv_usec = (suseconds_t)((sec - (double)v.tv_sec) * 1e6);
                return v;
            };
            const double s0 = g_tr.host_t0 + 1e-3 * r[5];
            palial_set_task_name(name);
            palial_set_task_cpu_node(g.device, 0, name);
            palial_get_task_time(tv(s0), tv(s0 + 1e-3 * r[4]), name);
        }
    }
    return 0;
}

// algorithmic work of the batched launches (leading terms, LAPACK counts)
double b3(int b) { return (double)b * b * b; }
double tile_bytes(int b) { return 8.0 * b * b; }

GemmArgs gargs(int mode, double *mat, int N, int b, int k, int T)
{
    GemmArgs a;
    memset(&a, 0, sizeof a);
    a.mode = mode;
    a.mat = mat;
    a.N = N;
    a.b = b;
    a.MT = N / b;
    a.k = k;
    a.T = T;
    a.ntasks = (int)gemm_task_count(mode, T);
    a.alpha = -1.0;
    a.beta = 1.0;
    return a;
}

TrsmArgs targs(double *mat, int N, int b, int k)
{
    TrsmArgs t;
    memset(&t, 0, sizeof t);
    const int MT = N / b;
    t.diag = mat + tile_off(b, MT, k, k);
    t.mat = mat;
    t.N = N;
    t.b = b;
    t.MT = MT;
    t.k = k;
    t.first = k + 1;
    t.ntiles = MT - k - 1;
    return t;
}

// Cholesky of a diagonal tile of the DAG (LAPACKE_dpotrf('U'),
// src/cholesky.h:40), writing the 32x32 block inverses to dinv.  For b >= 256
// (b % 128 == 0) the tile is copied out as a grid of 128-sub-tiles and
// factored by the tiled algorithm itself: the one-workgroup kernel runs only
// on the 128-diagonal sub-tiles, and the sub-tile TRSMs and SYRK/GEMM updates
// spread over many workgroups instead of one CU.  info receives the first
// failing global column (1-based) as with the one-kernel factorization.
// tile32: the tile is FP32 (the working copy of a mixed-precision solve); it
// is factored in FP64 in the copy and rounded back.
hipError_t potrf_tile(double *tile, int b, int *info, int info_base, double *dinv, hipStream_t s, bool tile32 = false)
{
    if (sched_rec_on()) {
        const long t = rec_tile_ptr(tile), d = rec_dinv_id(dinv);
        rec_put(0, s, nullptr, 1, {t}, {t, d}, {t, (long)(t % g_rec.MT), d, -1L}, 0, (int)(t % g_rec.MT), -1);
        return hipSuccess;
    }
    const bool sub = g.subtile && b >= 256 && b % 128 == 0;
    const bool copy = (sub || tile32) && dinv && g.d_sub && g.sub_cap >= (size_t)b * b;
    if (!copy) {
        if (tile32) return hipErrorInvalidValue;
        return launch_potrf2(tile, b, info, info_base, dinv, s);
    }
    const int bs = sub ? 128 : b, ms = b / bs;
    hipError_t e;
    if ((e = launch_tile_split(g.d_sub, tile, b, bs, true, s, tile32)) != hipSuccess) return e;
    for (int j = 0; j < ms; j++) {
        double *dj = dinv + (size_t)j * bs * 32;    // the bs/32 block inverses of sub-tile j
        if ((e = launch_potrf2(g.d_sub + tile_off(bs, ms, j, j), bs, info, info_base + j * bs, dj, s)) != hipSuccess)
            return e;
        if (j + 1 == ms) break;
        TrsmArgs t = targs(g.d_sub, b, bs, j);
        t.first = j + 1;
        t.ntiles = ms - j - 1;
        if ((e = launch_trsm2(TR_CHOL, t, dj, s)) != hipSuccess) return e;
        if ((e = launch_gemm_fine(gargs(GM_CHOL_TRAIL, g.d_sub, b, bs, j, ms - j - 1), true, false, s)) != hipSuccess)
            return e;
    }
    return launch_tile_split(g.d_sub, tile, b, bs, false, s, tile32);
}

// LU of a diagonal tile (PALIAL_dgetrf, src/srclu/dgetrf.h:11) writing the
// 32x32 block inverses of U (dU) and L (dL); sub-tiles and tile32 as for
// potrf_tile.  Inside a sub-tile the one-workgroup kernel keeps the
// reference's small-pivot scaling rule; the rows of the sub-tiles below are
// solved with the block inverses, as the DAG solves the tiles below a
// diagonal tile.
hipError_t getrf_tile(double *tile, int b, double *dU, double *dL, hipStream_t s, bool tile32 = false)
{
    if (sched_rec_on()) {
        const long t = rec_tile_ptr(tile), u = rec_dinv_id(dU), l = rec_dinv_id(dL);
        rec_put(0, s, nullptr, 1, {t}, {t, u, l}, {t, (long)(t % g_rec.MT), u, l}, 1, (int)(t % g_rec.MT), -1);
        return hipSuccess;
    }
    const bool sub = g.subtile && b >= 256 && b % 128 == 0;
    const bool copy = (sub || tile32) && g.d_sub && g.sub_cap >= (size_t)b * b;
    if (!copy) {
        if (tile32) return hipErrorInvalidValue;
        return launch_getrf2(tile, b, dU, dL, s);
    }
    const int bs = sub ? 128 : b, ms = b / bs;
    hipError_t e;
    if ((e = launch_tile_split(g.d_sub, tile, b, bs, true, s, tile32)) != hipSuccess) return e;
    for (int j = 0; j < ms; j++) {
        double *uj = dU + (size_t)j * bs * 32, *lj = dL + (size_t)j * bs * 32;
        if ((e = launch_getrf2(g.d_sub + tile_off(bs, ms, j, j), bs, uj, lj, s)) != hipSuccess) return e;
        if (j + 1 == ms) break;
        TrsmArgs t = targs(g.d_sub, b, bs, j);
        t.first = j + 1;
        t.ntiles = ms - j - 1;
        if ((e = launch_trsm2(TR_LUL, t, lj, s)) != hipSuccess) return e;    // row j: L_jj X = A_jn
        if ((e = launch_trsm2(TR_LUR, t, uj, s)) != hipSuccess) return e;    // column j: X U_jj = A_mj
        if ((e = launch_gemm_fine(gargs(GM_LU_TRAIL, g.d_sub, b, bs, j, ms - j - 1), false, false, s)) != hipSuccess)
            return e;
    }
    return launch_tile_split(g.d_sub, tile, b, bs, false, s, tile32);
}

// event record / wait of the chain schedules (recorded under the schedule recorder)
hipError_t ev_record(hipEvent_t e, hipStream_t s)
{
    if (sched_rec_on()) {
        rec_put(1, s, e, 0, {}, {}, {});
        return hipSuccess;
    }
    return hipEventRecord(e, s);
}
hipError_t ev_wait(hipStream_t s, hipEvent_t e)
{
    if (sched_rec_on()) {
        rec_put(2, s, e, 0, {}, {}, {});
        return hipSuccess;
    }
    return hipStreamWaitEvent(s, e, 0);
}

// events of one factorization: [0, MT+1) panel k ready, [MT+1, 2MT+2) update
// k done (split update: its first launch), 2MT+2 fork, 2MT+3 graph gate,
// [2MT+4, 3MT+5) chain TRSM k done, [3MT+6, 4MT+7) row/column TRSM and packing
// of step k done, 4MT+7 end of the bulk stream
size_t events_needed(int MT) { return 4 * (size_t)MT + 24; }   // QR: 2 MT + 3 + 2 * 8 at most

GemmArgs single_tile(const double *A, const double *B, double *C, int b, bool syrk)
{
    GemmArgs a;
    memset(&a, 0, sizeof a);
    a.mode = GM_SINGLE;
    a.A = A; a.B = B; a.C = C; a.D = C;
    a.N = b; a.b = b; a.MT = 1; a.T = 1; a.ntasks = 1;
    a.alpha = -1.0; a.beta = 1.0;
    a.single_diag = syrk ? 1 : 0;
    return a;
}

// Diagonal-chain schedule of the upper Cholesky (src/cholesky.h:15-176).
//   chain (sc): P(k) -> T1(k) TRSM of tile (k,k+1) -> S1(k) SYRK of tile
//               (k+1,k+1) by row k -> P(k+1)
//   bulk  (sb): TB(k) TRSM of tiles (k, k+2..) -> R(k) update of every tile
//               (i,j), k < i <= j, except (k+1,k+1)
// Before T1(k) the chain waits for R(k-1), which brings tiles (k,k+1) and
// (k+1,k+1) up to date with rows < k; R(k) waits for T1(k).  Every tile
// therefore receives the updates of rows 0..k in increasing k as in the
// reference, with the same kernels: only the grouping into launches differs.
//
// Split update (sa != nullptr): R(k) is launched as F(k), the rest of tile
// row k+1 and tile (k+2,k+2), then E(k), every other tile.  TB(k) and the
// packing of row k run on a third stream sa after F(k-1) and P(k), i.e.
// beside E(k-1), so they leave the bulk stream's critical path; the chain
// waits for F(k-1) only.  The bulk order F(k-1) E(k-1) F(k) E(k) keeps each
// tile's updates in increasing k, and packed set k % 4 is rewritten by step
// k+4's packing only after F(k+3), which follows every launch that reads
// panel k on the bulk stream (paired steps read panels k and k+1 until E of
// step k+1, so two sets would not do; palial_hip_schedule_ops records the
// schedule for a host-side race check).
//
// fs (mixed tiers with packed panels): dA holds FP32 tiles (element offsets
// of the FP64 layout).  Diagonal tiles are factored in FP64 copies, the TRSMs
// solve in FP64 and round back, the updates read and write FP32 C tiles, and
// S1 packs tile (k,k+1) and runs on the packed kernel like the bulk.
void dag_cholesky_chain(double *dA, int N, int b, hipStream_t s0, hipStream_t sc, hipStream_t sb, hipStream_t sa,
                        int *dinfo, Launcher &L, int prec, bool fs = false)
{
    L.b = b;
    L.prec = prec;
    const int MT = N / b;
    const bool v2 = panel2_ok(b);
    const bool lp = prec > 0 && b % 128 == 0;
    const bool split = sa != nullptr;
    const size_t psb = pack_set_bytes(b, MT);
    const bool pk = lp && g.d_pack && g.pack_cap >= (split ? PACK_SETS : 1) * psb;
    if (fs && !(pk && split && v2 && g.pack_cap >= PACK_SETS * psb + 2 * (size_t)b * b * 4)) {
        L.chk(hipErrorInvalidValue);     // the FP32 working copy needs the packed split schedule
        return;
    }
    const double esz = prec == 2 ? 2.0 : prec == 3 ? 1.0 : 4.0;   // packed bytes per element (split BF16: hi + lo)
    const double sbt = fs ? 4.0 * b * b : tile_bytes(b);      // bytes of one stored tile
    auto pset = [&](int k) { return (char *)g.d_pack + (split ? (size_t)(k % PACK_SETS) * psb : 0); };
    char *pk1 = (char *)g.d_pack + PACK_SETS * psb;            // S1's packed tile (fs)
    auto tile = [&](int m, int n) -> double * {
        return fs ? (double *)(void *)((float *)(void *)dA + tile_off(b, MT, m, n)) : dA + tile_off(b, MT, m, n);
    };
    auto dinvU = [&](int k) { return g.d_dinv + (size_t)k * 64 * b; };
    auto potrf = [&](int k, hipStream_t s) {
        L.run(KC_DIAG, b3(b) / 3, 2 * tile_bytes(b), s, [&] {
            return v2 ? potrf_tile(tile(k, k), b, dinfo, k * b, dinvU(k), s, fs)
                      : launch_potrf(dA + tile_off(b, MT, k, k), b, dinfo, k * b, s);
        });
    };
    auto trsm = [&](int k, int first, int n, hipStream_t s) {
        TrsmArgs t = targs(dA, N, b, k);
        t.diag = tile(k, k);
        t.first = first;
        t.ntiles = n;
        t.f32 = fs ? 1 : 0;
        L.run(KC_TRSM, n * b3(b), 3.0 * n * sbt, s, [&] {
            return v2 ? launch_trsm2(TR_CHOL, t, dinvU(k), s) : launch_trsm(TR_CHOL, t, s);
        });
    };
    auto gemm = [&](const GemmArgs &a, double flops, double tiles, hipStream_t s) {
        L.run(KC_GEMM, flops, 3.0 * tiles * tile_bytes(b), s,
              [&] { return lp ? launch_gemm_lp(a, true, false, prec, s) : launch_gemm(a, true, false, s); });
    };
    auto pack = [&](int k, int n, void *dst, hipStream_t s) {
        L.run(KC_PACK, 0, n * (sbt + esz * b * b), s,
              [&] { return launch_pack(dA, N, b, k, k + 1, n, false, prec, dst, s, fs); });
    };
    // trailing update of step k over tasks [task0, count) of `mode`: ngemm GEMM
    // and nsyrk SYRK tiles, reading npanel tiles of the packed row k
    // panel2: the packed row k+1 as well (GM_CHOL_REST2: twice the products per output tile)
    auto update = [&](int mode, int k, int task0, int ntasks, double ngemm, double nsyrk, double npanel, void *panel,
                      hipStream_t s, void *panel2 = nullptr) {
        const int T = MT - k - 1;
        GemmArgs a = gargs(mode, dA, N, b, k, T);
        a.task0 = task0;
        a.ntasks = ntasks;
        a.f32c = fs ? 1 : 0;
        if (a.ntasks <= 0) return;
        const double np = panel2 ? 2.0 : 1.0;
        const double fl = np * (2 * b3(b) * ngemm + b3(b) * nsyrk);
        if (pk) {
            a.pa = a.pb = panel;
            a.pa2 = a.pb2 = panel2;
            L.run(KC_PKGEMM, fl, 2.0 * (ngemm + nsyrk) * sbt + np * npanel * esz * b * b, s,
                  [&] { return launch_gemm_packed(a, prec, s); });
        } else {
            gemm(a, fl, ngemm + nsyrk, s);
        }
    };
    // Paired steps (BF16 and split BF16 on the FP32 working copy): an even step j with at
    // least 4 trailing tiles leaves the tiles (n, m), n >= j+3 of its update
    // to step j+1, which applies panels j and j+1 to them in one pass, so
    // those C tiles are read and written once for two steps.  Step j updates
    // rows j+1 and j+2 (GM_CHOL_FIRST, GM_CHOL_ROW2): the chain's next two
    // steps never wait for a full trailing update, only every other one.
    const bool pair = fs && (prec == 2 || prec == 4) && split && pk && g.pair && !g_tr.active;
    auto lead = [&](int k) { return pair && k % 2 == 0 && MT - k - 1 >= 4; };
    hipEvent_t *ev = g.events.data();
    hipEvent_t *evP = ev, *evR = ev + MT + 1, *evT = ev + 2 * MT + 4, *evB = ev + 3 * MT + 6;
    hipEvent_t fork = ev[2 * MT + 2], evEnd = ev[4 * MT + 7];
    hipStream_t st = split ? sa : sb;    // stream of TB(k) and the packing
    L.chk(ev_record(fork, s0));
    L.chk(ev_wait(sc, fork));
    if (sb != s0) L.chk(ev_wait(sb, fork));
    if (split) L.chk(ev_wait(sa, fork));
    potrf(0, sc);
    L.chk(ev_record(evP[0], sc));
    for (int k = 0; k < MT; k++) {
        const int T = MT - k - 1;
        if (T > 0) {
            if (k > 0) L.chk(ev_wait(sc, evR[k - 1]));
            trsm(k, k + 1, 1, sc);
            L.chk(ev_record(evT[k], sc));
            if (fs) {
                // GM_CHOL_ROW task 0 is tile (k+1,k+1) from packed tile 0 of row k
                pack(k, 1, pk1, sc);
                update(GM_CHOL_ROW, k, 0, 1, 0, 1, 1, pk1, sc);
            } else {
                gemm(single_tile(dA + tile_off(b, MT, k, k + 1), dA + tile_off(b, MT, k, k + 1),
                                 dA + tile_off(b, MT, k + 1, k + 1), b, true),
                     b3(b), 1, sc);
            }
            potrf(k + 1, sc);
            L.chk(ev_record(evP[k + 1], sc));
        }
        if (split && k > 0) L.chk(ev_wait(sa, evR[k - 1]));
        L.chk(ev_wait(st, evP[k]));
        if (T > 1) trsm(k, k + 2, T - 1, st);
        if (T > 1 && pk) {
            // row k rounded once; the update streams the packed copy
            L.chk(ev_wait(st, evT[k]));
            pack(k, T, pset(k), st);
        }
        if (split) {
            L.chk(ev_record(evB[k], sa));
            L.chk(ev_wait(sb, evB[k]));
        }
        if (T > 0) L.chk(ev_wait(sb, evT[k]));
        // task 0 of each enumeration is tile (k+1,k+1) (the chain's) or (k+2,k+2) (F(k)'s)
        if (!split) {
            if (T > 1) update(GM_CHOL_TRAIL, k, 1, T * (T + 1) / 2 - 1, (double)T * (T - 1) / 2, T - 1, T, pset(k), sb);
            L.chk(ev_record(evR[k], sb));
        } else {
            if (T > 1) update(GM_CHOL_FIRST, k, 1, T, T - 1, 1, T, pset(k), sb);
            L.chk(ev_record(evR[k], sb));
            if (lead(k)) {
                // row k+2 and tile (k+3, k+3); the rest waits for step k+1
                update(GM_CHOL_ROW2, k, 0, T - 1, T - 2, 1, T - 1, pset(k), sb);
            } else if (k > 0 && lead(k - 1)) {
                // steps k-1 and k on the tiles (n, m), k+2 <= n <= m, but (k+2, k+2)
                const int R = T - 1;          // rows k+2 .. MT-1
                update(GM_CHOL_REST2, k - 1, 1, R * (R + 1) / 2 - 1, (double)R * (R - 1) / 2, R - 1, R, pset(k - 1),
                       sb, pset(k));
            } else if (T > 2) {
                update(GM_CHOL_REST, k, 1, (T - 1) * T / 2 - 1, (double)(T - 1) * (T - 2) / 2, T - 2, T - 1, pset(k), sb);
            }
        }
    }
    L.chk(ev_record(evEnd, sb));
    if (sb != s0) L.chk(ev_wait(s0, evEnd));
}

// Diagonal-chain schedule of the no-pivot LU (src/lu.h:14-173), same shape:
// the chain solves tiles (k,k+1) and (k+1,k) and updates (k+1,k+1); the bulk
// solves the rest of row and column k and updates every other trailing tile.
// Split update as for the Cholesky: F(k) is tile row and column k+1 and tile
// (k+2,k+2), E(k) the tiles (m,n), m,n >= k+2, but (k+2,k+2).  fs as for the
// Cholesky.
void dag_lu_chain(double *dA, int N, int b, hipStream_t s0, hipStream_t sc, hipStream_t sb, hipStream_t sa,
                  Launcher &L, int prec, bool fs = false)
{
    L.b = b;
    L.prec = prec;
    const int MT = N / b;
    const bool v2 = panel2_ok(b);
    const bool lp = prec > 0 && b % 128 == 0;
    const bool split = sa != nullptr;
    const size_t psb = pack_set_bytes(b, MT);
    const bool pk = lp && g.d_pack && g.pack_cap >= (split ? PACK_SETS : 1) * psb;
    if (fs && !(pk && split && v2 && g.pack_cap >= PACK_SETS * psb + 2 * (size_t)b * b * 4)) {
        L.chk(hipErrorInvalidValue);
        return;
    }
    const double esz = prec == 2 ? 2.0 : prec == 3 ? 1.0 : 4.0;   // packed bytes per element (split BF16: hi + lo)
    const double sbt = fs ? 4.0 * b * b : tile_bytes(b);
    auto pset = [&](int k) { return (char *)g.d_pack + (split ? (size_t)(k % PACK_SETS) * psb : 0); };
    char *pk1U = (char *)g.d_pack + PACK_SETS * psb, *pk1L = pk1U + (size_t)b * b * 4;   // S1's packed tiles (fs)
    auto tile = [&](int m, int n) -> double * {
        return fs ? (double *)(void *)((float *)(void *)dA + tile_off(b, MT, m, n)) : dA + tile_off(b, MT, m, n);
    };
    auto dU = [&](int k) { return g.d_dinv + (size_t)k * 64 * b; };
    auto dL = [&](int k) { return g.d_dinv + (size_t)k * 64 * b + (size_t)32 * b; };
    auto getrf = [&](int k, hipStream_t s) {
        L.run(KC_DIAG, 2 * b3(b) / 3, 2 * tile_bytes(b), s, [&] {
            return v2 ? getrf_tile(tile(k, k), b, dU(k), dL(k), s, fs) : launch_getrf(dA + tile_off(b, MT, k, k), b, s);
        });
    };
    auto trsm = [&](int kind, int k, int first, int n, hipStream_t s) {
        TrsmArgs t = targs(dA, N, b, k);
        t.diag = tile(k, k);
        t.first = first;
        t.ntiles = n;
        t.f32 = fs ? 1 : 0;
        L.run(KC_TRSM, n * b3(b), 3.0 * n * sbt, s, [&] {
            return v2 ? launch_trsm2(kind, t, kind == TR_LUR ? dU(k) : dL(k), s) : launch_trsm(kind, t, s);
        });
    };
    auto gemm = [&](const GemmArgs &a, double tiles, hipStream_t s) {
        L.run(KC_GEMM, 2 * b3(b) * tiles, 3.0 * tiles * tile_bytes(b), s,
              [&] { return lp ? launch_gemm_lp(a, false, false, prec, s) : launch_gemm(a, false, false, s); });
    };
    // row (U) and column (L, transposed) panels of step k, n tiles from k+1
    auto pack = [&](int k, int n, char *pU, char *pL, hipStream_t s) {
        const double pb = n * (sbt + esz * b * b);
        L.run(KC_PACK, 0, pb, s, [&] { return launch_pack(dA, N, b, k, k + 1, n, false, prec, pU, s, fs); });
        L.run(KC_PACK, 0, pb, s, [&] { return launch_pack(dA, N, b, k, k + 1, n, true, prec, pL, s, fs); });
    };
    // trailing update of step k over tasks [task0, task0 + ntasks) of `mode`,
    // reading npanel tiles of each packed panel
    // pU2 / pL2: the packed panels of step k+1 as well (GM_LU_REST2)
    auto update = [&](int mode, int k, int task0, int ntasks, double npanel, char *pU, char *pL, hipStream_t s,
                      char *pU2 = nullptr, char *pL2 = nullptr) {
        const int T = MT - k - 1;
        GemmArgs a = gargs(mode, dA, N, b, k, T);
        a.task0 = task0;
        a.ntasks = ntasks;
        a.f32c = fs ? 1 : 0;
        if (a.ntasks <= 0) return;
        const double nt = a.ntasks, np = pU2 ? 2.0 : 1.0;
        if (pk) {
            a.pa = pL;
            a.pb = pU;
            a.pa2 = pL2;
            a.pb2 = pU2;
            L.run(KC_PKGEMM, np * 2 * b3(b) * nt, 2.0 * nt * sbt + np * 2.0 * npanel * esz * b * b, s,
                  [&] { return launch_gemm_packed(a, prec, s); });
        } else {
            gemm(a, nt, s);
        }
    };
    // paired steps as in the Cholesky chain schedule
    const bool pair = fs && (prec == 2 || prec == 4) && split && pk && g.pair && !g_tr.active;
    auto lead = [&](int k) { return pair && k % 2 == 0 && MT - k - 1 >= 4; };
    auto psetU = [&](int k) { return pset(k); };
    auto psetL = [&](int k) { return pset(k) + (size_t)MT * b * b * 4; };
    hipEvent_t *ev = g.events.data();
    hipEvent_t *evP = ev, *evR = ev + MT + 1, *evT = ev + 2 * MT + 4, *evB = ev + 3 * MT + 6;
    hipEvent_t fork = ev[2 * MT + 2], evEnd = ev[4 * MT + 7];
    hipStream_t st = split ? sa : sb;
    L.chk(ev_record(fork, s0));
    L.chk(ev_wait(sc, fork));
    if (sb != s0) L.chk(ev_wait(sb, fork));
    if (split) L.chk(ev_wait(sa, fork));
    getrf(0, sc);
    L.chk(ev_record(evP[0], sc));
    for (int k = 0; k < MT; k++) {
        const int T = MT - k - 1;
        if (T > 0) {
            if (k > 0) L.chk(ev_wait(sc, evR[k - 1]));
            trsm(TR_LUL, k, k + 1, 1, sc);
            trsm(TR_LUR, k, k + 1, 1, sc);
            L.chk(ev_record(evT[k], sc));
            if (fs) {
                // GM_LU_PANEL task 0 is tile (k+1,k+1) from packed tiles 0 of both panels
                pack(k, 1, pk1U, pk1L, sc);
                update(GM_LU_PANEL, k, 0, 1, 1, pk1U, pk1L, sc);
            } else {
                gemm(single_tile(dA + tile_off(b, MT, k + 1, k), dA + tile_off(b, MT, k, k + 1),
                                 dA + tile_off(b, MT, k + 1, k + 1), b, false),
                     1, sc);
            }
            getrf(k + 1, sc);
            L.chk(ev_record(evP[k + 1], sc));
        }
        if (split && k > 0) L.chk(ev_wait(sa, evR[k - 1]));
        L.chk(ev_wait(st, evP[k]));
        if (T > 1) {
            trsm(TR_LUL, k, k + 2, T - 1, st);
            trsm(TR_LUR, k, k + 2, T - 1, st);
        }
        if (T > 1 && pk) {
            L.chk(ev_wait(st, evT[k]));
            pack(k, T, psetU(k), psetL(k), st);
        }
        if (split) {
            L.chk(ev_record(evB[k], sa));
            L.chk(ev_wait(sb, evB[k]));
        }
        if (T > 0) L.chk(ev_wait(sb, evT[k]));
        // task 0 of each enumeration is tile (k+1,k+1) (the chain's) or (k+2,k+2) (F(k)'s)
        if (!split) {
            if (T > 1) update(GM_LU_TRAIL, k, 1, T * T - 1, T, psetU(k), psetL(k), sb);
            L.chk(ev_record(evR[k], sb));
        } else {
            if (T > 1) update(GM_LU_FIRST, k, 1, 2 * T - 1, T, psetU(k), psetL(k), sb);
            L.chk(ev_record(evR[k], sb));
            if (lead(k)) {
                // column and row k+2 and tile (k+3, k+3); the rest waits for step k+1
                update(GM_LU_ROW2, k, 0, 2 * T - 3, T - 1, psetU(k), psetL(k), sb);
            } else if (k > 0 && lead(k - 1)) {
                // steps k-1 and k on the tiles (m, n), m, n >= k+2, but (k+2, k+2)
                const int R = T - 1;
                update(GM_LU_REST2, k - 1, 1, R * R - 1, R, psetU(k - 1), psetL(k - 1), sb, psetU(k), psetL(k));
            } else if (T > 2) {
                update(GM_LU_REST, k, 1, (T - 1) * (T - 1) - 1, T - 1, psetU(k), psetL(k), sb);
            }
        }
    }
    L.chk(ev_record(evEnd, sb));
    if (sb != s0) L.chk(ev_wait(s0, evEnd));
}

// Bulk stream of the chain schedule.  The mixed tiers are bound by the
// diagonal chain and keep their trailing updates off the reserved CUs; the
// FP64 updates are MFMA-bound and take every CU (measured at N=32768:
// Cholesky 238 vs 261 ms, LU 445 vs 491 ms without the reservation), which
// also lets the FP64 schedule be captured into a graph.
hipStream_t bulk_stream(int prec, hipStream_t s0)
{
    return prec > 0 ? g.s_bulk : s0;
}
// Diagonal-chain stream: the mixed tiers may confine it to the reserved CUs,
// where its workgroups never share a CU with the bulk update's.
hipStream_t chain_stream(int prec, hipStream_t s1)
{
    return prec > 0 && g.s_chainm ? g.s_chainm : s1;
}

// Cholesky, src/cholesky.h:15-176.  Look-ahead: step k's update of tile row
// k+1 runs on the panel stream ahead of the rest, so POTRF/TRSM of step k+1
// overlap the bulk of step k's trailing update.
void dag_cholesky(double *dA, int N, int b, bool lookahead, hipStream_t s0, hipStream_t s1, int *dinfo,
                  Launcher &L, int prec = 0, bool fs = false)
{
    if (fs && !(lookahead && g.dag == 2)) {
        L.chk(hipErrorInvalidValue);     // the FP32 working copy runs on the split chain schedule only
        return;
    }
    L.b = b;
    L.prec = prec;
    const int MT = N / b;
    const bool v2 = panel2_ok(b);
    const bool lp = prec > 0 && b % 128 == 0;
    auto dinvU = [&](int k) { return g.d_dinv + (size_t)k * 64 * b; };
    auto potrf = [&](int k, hipStream_t s) {
        L.run(KC_DIAG, b3(b) / 3, 2 * tile_bytes(b), s, [&] {
            return v2 ? potrf_tile(dA + tile_off(b, MT, k, k), b, dinfo, k * b, dinvU(k), s)
                      : launch_potrf(dA + tile_off(b, MT, k, k), b, dinfo, k * b, s);
        });
    };
    auto trsm = [&](int k, hipStream_t s) {
        const int T = MT - k - 1;
        L.run(KC_TRSM, T * b3(b), 3.0 * T * tile_bytes(b), s, [&] {
            return v2 ? launch_trsm2(TR_CHOL, targs(dA, N, b, k), dinvU(k), s)
                      : launch_trsm(TR_CHOL, targs(dA, N, b, k), s);
        });
    };
    auto update = [&](int mode, int k, hipStream_t s) {
        const int T = MT - k - 1;
        GemmArgs a = gargs(mode, dA, N, b, k, T);
        long ngemm, nsyrk;
        if (mode == GM_CHOL_TRAIL) { nsyrk = T; ngemm = (long)T * (T - 1) / 2; }
        else if (mode == GM_CHOL_ROW) { nsyrk = 1; ngemm = T - 1; }
        else { nsyrk = T - 1; ngemm = (long)(T - 1) * (T - 2) / 2; }
        L.run(KC_GEMM, 2 * b3(b) * ngemm + b3(b) * nsyrk, 3.0 * (ngemm + nsyrk) * tile_bytes(b), s,
              [&] { return lp ? launch_gemm_lp(a, true, false, prec, s) : launch_gemm(a, true, false, s); });
    };
    if (lookahead && g.dag >= 1) {
        dag_cholesky_chain(dA, N, b, s0, chain_stream(prec, s1), bulk_stream(prec, s0), g.dag == 2 ? g.s_aux : nullptr,
                           dinfo, L, prec, fs);
        return;
    }
    if (!lookahead) {
        for (int k = 0; k < MT; k++) {
            potrf(k, s0);
            trsm(k, s0);
            if (MT - k - 1 > 0) update(GM_CHOL_TRAIL, k, s0);
        }
        return;
    }
    hipEvent_t *evP = g.events.data();          // panel k ready
    hipEvent_t *evU = g.events.data() + MT + 1; // rest-update k done
    hipEvent_t fork = g.events[2 * MT + 2];
    L.chk(ev_record(fork, s0));
    L.chk(ev_wait(s1, fork));
    potrf(0, s1);
    trsm(0, s1);
    L.chk(ev_record(evP[0], s1));
    for (int k = 0; k < MT; k++) {
        const int T = MT - k - 1;
        if (T > 0) {
            if (k > 0) L.chk(ev_wait(s1, evU[k - 1]));
            update(GM_CHOL_ROW, k, s1);
            potrf(k + 1, s1);
            trsm(k + 1, s1);
            L.chk(ev_record(evP[k + 1], s1));
        }
        L.chk(ev_wait(s0, evP[k]));
        if (T > 1) update(GM_CHOL_REST, k, s0);
        L.chk(ev_record(evU[k], s0));
    }
}

// LU without pivoting, src/lu.h:14-173 (with correct tile dependences).
void dag_lu(double *dA, int N, int b, bool lookahead, hipStream_t s0, hipStream_t s1, Launcher &L, int prec = 0,
            bool fs = false)
{
    if (fs && !(lookahead && g.dag == 2)) {
        L.chk(hipErrorInvalidValue);
        return;
    }
    L.b = b;
    L.prec = prec;
    const int MT = N / b;
    const bool v2 = panel2_ok(b);
    const bool lp = prec > 0 && b % 128 == 0;
    auto panel = [&](int k, hipStream_t s) {
        const int T = MT - k - 1;
        double *dU = g.d_dinv + (size_t)k * 64 * b, *dL = dU + (size_t)32 * b;
        L.run(KC_DIAG, 2 * b3(b) / 3, 2 * tile_bytes(b), s, [&] {
            return v2 ? getrf_tile(dA + tile_off(b, MT, k, k), b, dU, dL, s)
                      : launch_getrf(dA + tile_off(b, MT, k, k), b, s);
        });
        L.run(KC_TRSM, T * b3(b), 3.0 * T * tile_bytes(b), s, [&] {
            return v2 ? launch_trsm2(TR_LUR, targs(dA, N, b, k), dU, s) : launch_trsm(TR_LUR, targs(dA, N, b, k), s);
        });
        L.run(KC_TRSM, T * b3(b), 3.0 * T * tile_bytes(b), s, [&] {
            return v2 ? launch_trsm2(TR_LUL, targs(dA, N, b, k), dL, s) : launch_trsm(TR_LUL, targs(dA, N, b, k), s);
        });
    };
    auto update = [&](int mode, int k, hipStream_t s) {
        const int T = MT - k - 1;
        GemmArgs a = gargs(mode, dA, N, b, k, T);
        const double ntile = (double)a.ntasks;
        L.run(KC_GEMM, 2 * b3(b) * ntile, 3.0 * ntile * tile_bytes(b), s,
              [&] { return lp ? launch_gemm_lp(a, false, false, prec, s) : launch_gemm(a, false, false, s); });
    };
    if (lookahead && g.dag >= 1) {
        dag_lu_chain(dA, N, b, s0, chain_stream(prec, s1), bulk_stream(prec, s0), g.dag == 2 ? g.s_aux : nullptr, L,
                     prec, fs);
        return;
    }
    if (!lookahead) {
        for (int k = 0; k < MT; k++) {
            panel(k, s0);
            if (MT - k - 1 > 0) update(GM_LU_TRAIL, k, s0);
        }
        return;
    }
    hipEvent_t *evP = g.events.data();
    hipEvent_t *evU = g.events.data() + MT + 1;
    hipEvent_t fork = g.events[2 * MT + 2];
    L.chk(ev_record(fork, s0));
    L.chk(ev_wait(s1, fork));
    panel(0, s1);
    L.chk(ev_record(evP[0], s1));
    for (int k = 0; k < MT; k++) {
        const int T = MT - k - 1;
        if (T > 0) {
            if (k > 0) L.chk(ev_wait(s1, evU[k - 1]));
            update(GM_LU_PANEL, k, s1);
            panel(k + 1, s1);
            L.chk(ev_record(evP[k + 1], s1));
        }
        L.chk(ev_wait(s0, evP[k]));
        if (T > 1) update(GM_LU_REST, k, s0);
        L.chk(ev_record(evU[k], s0));
    }
}

// The QR schedule's streams are created at the first QR factorization and
// kept until palial_hip_finalize.  All of them carry a CU mask, which gives
// each its own hardware queue (streams without one share the runtime's few
// queues and would run one after the other): the panel chains may use every
// CU, the update streams stay off the first qr_reserve CUs.
int qr_streams_up()
{
    if (g.qr_up) return 0;
    g.qr_up = true;   // qr_streams_down() releases whatever was created, also after a failure
    const int words = (g.ncu + 31) / 32;
    std::vector<uint32_t> all(words, 0), upd(words, 0);
    for (int cu = 0; cu < g.ncu; cu++) all[cu / 32] |= 1u << (cu % 32);
    for (int cu = g.qr_reserve; cu < g.ncu; cu++) upd[cu / 32] |= 1u << (cu % 32);
    for (int i = 0; i < g.qr_streams; i++) {
        HIPCHK(hipExtStreamCreateWithCUMask(&g.s_qr[i], words, all.data()));
        HIPCHK(hipExtStreamCreateWithCUMask(&g.s_qrb[i], words, upd.data()));
    }
    return 0;
}

void qr_streams_down()
{
    if (!g.qr_up) return;
    for (auto &q : g.s_qr)
        if (q) { (void)hipStreamDestroy(q); q = nullptr; }
    for (auto &q : g.s_qrb)
        if (q) { (void)hipStreamDestroy(q); q = nullptr; }
    g.qr_up = false;
}

// Flat-tree tiled QR, src/qr.h:18-190.
// two == false: every task on s0 in the reference's loop order.
// two == true: with ns = PALIAL_QR_STREAMS (default 8), step k's panel chain
// (GEQRT(k), then TSQRT(k, m) in m order) runs on chain stream k % ns and its
// updates (ORMQR(k), then TSMQR(k, m) in m order) on update stream k % ns,
// each stream with its own workspaces:
//   GEQRT(k) and TSQRT(k, m) wait for TSMQR(k-1, k) / TSMQR(k-1, m) (tile final),
//   ORMQR(k) waits for GEQRT(k), TSMQR(k, m) for TSQRT(k, m).
// Tasks that may run at the same time touch disjoint tiles (ORMQR(k) reads
// only the unit-lower V of A(k,k) while TSQRT(k, m) rewrites its upper R;
// TSMQR(k, m) and TSMQR(k+1, m') write rows {k, m} and {k+1, m'}, m' < m),
// so up to ns steps progress together and the factors are the same bits as
// in the reference's order.
void dag_qr(double *dA, double *dS, int N, int b, hipStream_t s0, hipStream_t s1, bool two, Launcher &L, int ns)
{
    L.b = b;
    L.prec = 0;
    const int MT = N / b;
    const size_t wstride = (size_t)b * b * (size_t)MT;   // one bulk stream's W tiles
    auto geqrt = [&](int k, hipStream_t s, int c) {
        L.run(KC_DIAG, 4 * b3(b) / 3, 3 * tile_bytes(b), s, [&] {
            return launch_geqrt(dA + tile_off(b, MT, k, k), dS + tile_off(b, MT, k, k), g.d_tau + (size_t)c * b, b, s,
                                g.d_qg + (size_t)c * b * b);
        });
    };
    auto ormqr = [&](int k, hipStream_t s, int u) {   // PALIAL_dormqr on A(k,n), n > k: C -= V T^T V^T C
        const int T = MT - k - 1;
        // V(k) as a plain tile first: the products then never load A(k,k)'s
        // upper R, which step k's TSQRTs rewrite meanwhile on the chain stream.
        // It takes the stream's last W tile (ORMQR and TSMQR use T <= MT - 1).
        double *vk = g.d_w1 + u * wstride + (size_t)(MT - 1) * b * b;
        L.run(KC_PACK, 0, 2 * tile_bytes(b), s,
              [&] { return launch_unit_lower_copy(vk, dA + tile_off(b, MT, k, k), b, s); });
        GemmArgs a = gargs(GM_ORM_W, dA, N, b, k, T);
        a.A = vk; a.smat = dS; a.w1 = g.d_w1 + u * wstride; a.w2 = g.d_w2 + u * wstride;
        a.alpha = 1.0; a.beta = 0.0; a.maskA = MK_NONE;
        L.run(KC_QRUPD, T * b3(b), 3.0 * T * tile_bytes(b), s, [&] { return launch_gemm(a, true, false, s); });
        a.mode = GM_ORM_TW; a.maskA = MK_UPPER;
        L.run(KC_QRUPD, T * b3(b), 3.0 * T * tile_bytes(b), s, [&] { return launch_gemm(a, true, false, s); });
        a.mode = GM_ORM_VW; a.maskA = MK_NONE; a.alpha = -1.0; a.beta = 1.0;
        L.run(KC_QRUPD, T * b3(b), 3.0 * T * tile_bytes(b), s, [&] { return launch_gemm(a, false, false, s); });
    };
    auto tsqrt = [&](int k, int m, hipStream_t s, int c) {
        L.run(KC_TSQRT, 2 * b3(b), 5 * tile_bytes(b), s, [&] {
            return launch_tsqrt(dA + tile_off(b, MT, k, k), dA + tile_off(b, MT, m, k), dS + tile_off(b, MT, m, k),
                                g.d_tau + (size_t)c * b, b, s, g.d_qg + (size_t)c * b * b);
        });
    };
    auto tsmqr = [&](int k, int m, hipStream_t s, int u) {   // PALIAL_dtsmqr on (A(k,n), A(m,n)), n > k
        const int T = MT - k - 1;
        GemmArgs a = gargs(GM_QR_W, dA, N, b, k, T);
        a.m = m; a.smat = dS; a.w1 = g.d_w1 + u * wstride; a.w2 = g.d_w2 + u * wstride;
        a.alpha = 1.0; a.beta = 1.0;
        L.run(KC_QRUPD, 2 * T * b3(b), 4.0 * T * tile_bytes(b), s, [&] { return launch_gemm(a, true, false, s); });
        a.mode = GM_QR_TW; a.maskA = MK_UPPER; a.beta = 0.0;
        L.run(KC_QRUPD, T * b3(b), 4.0 * T * tile_bytes(b), s, [&] { return launch_gemm(a, true, false, s); });
        a.mode = GM_QR_VW; a.maskA = MK_NONE; a.alpha = -1.0; a.beta = 1.0;
        L.run(KC_QRUPD, 2 * T * b3(b), 3.0 * T * tile_bytes(b), s, [&] { return launch_gemm(a, false, false, s); });
    };
    if (!two) {
        for (int k = 0; k < MT; k++) {
            geqrt(k, s0, 0);
            if (k + 1 < MT) ormqr(k, s0, 0);
            for (int m = k + 1; m < MT; m++) {
                tsqrt(k, m, s0, 0);
                tsmqr(k, m, s0, 0);
            }
        }
        return;
    }
    (void)s1;
    const hipStream_t *chain = g.s_qr, *bulk = g.s_qrb;
    ns = std::max(1, std::min(ns, g.qr_streams));
    hipEvent_t *E = g.events.data();                 // E[(k & 1) MT + m]: TSMQR(k, m) done
    hipEvent_t evG = g.events[2 * MT], evQ = g.events[2 * MT + 1], fork = g.events[2 * MT + 2];
    hipEvent_t *join = g.events.data() + 2 * MT + 3;   // 2 ns events
    L.chk(ev_record(fork, s0));
    for (int i = 0; i < ns; i++) {
        L.chk(ev_wait(chain[i], fork));
        L.chk(ev_wait(bulk[i], fork));
    }
    for (int k = 0; k < MT; k++) {
        const int c = k % ns, u = k % ns;
        hipStream_t sc = chain[c], sb = bulk[u];
        if (k > 0) L.chk(ev_wait(sc, E[((k - 1) & 1) * MT + k]));
        geqrt(k, sc, c);
        if (k + 1 == MT) break;
        L.chk(ev_record(evG, sc));
        L.chk(ev_wait(sb, evG));
        ormqr(k, sb, u);
        for (int m = k + 1; m < MT; m++) {
            if (k > 0) L.chk(ev_wait(sc, E[((k - 1) & 1) * MT + m]));
            tsqrt(k, m, sc, c);
            L.chk(ev_record(evQ, sc));
            L.chk(ev_wait(sb, evQ));
            tsmqr(k, m, sb, u);
            L.chk(ev_record(E[(k & 1) * MT + m], sb));
        }
    }
    for (int i = 0; i < ns; i++) {
        L.chk(ev_record(join[2 * i], chain[i]));
        L.chk(ev_wait(s0, join[2 * i]));
        L.chk(ev_record(join[2 * i + 1], bulk[i]));
        L.chk(ev_wait(s0, join[2 * i + 1]));
    }
}

// Q of a tiled QR (V in dA, T factors in dS) applied from the left to the N x
// NC*b layout-L matrix dC: C <- Q^T C (trans) applies the reflector blocks in
// the order the factorization applied them to its own trailing columns
// (src/qr.h:43-182: ORMQR(k), then TSMQR(k, m) for m = k+1..), C <- Q C the
// same blocks untransposed in reverse order.  A TS block (k, m) of rows k and
// m is I - Y T Y^T with Y = [I; V], V = A(m,k), T = S(m,k), so
//   W = C(k,:) + V^T C(m,:);  W2 = op(T) W;  C(k,:) -= W2;  C(m,:) -= V W2,
// and the diagonal block k is I - V T V^T with V the unit lower part of A(k,k).
// (The reference's own NoTrans paths start at the wrong index,
// src/srcqr/dtsmqr.h:19-29, src/srcqr/dormqr.h:19-29; this follows the
// algebra.)  The column tiles of C are independent, so one launch covers all
// of them; w1 / w2 hold NC tiles each.
void qr_apply_dev(bool trans, double *dA, const double *dS, double *dC, int N, int b, int NC, double *w1, double *w2,
                  hipStream_t s, Launcher &L)
{
    const int MT = N / b;
    L.b = b;
    L.prec = 0;
    // few column tiles: 32x32 output blocks give each launch enough workgroups
    const bool fine = NC <= 2 && b % 32 == 0;
    auto gemm = [&](const GemmArgs &a, bool ta, double flops) {
        L.run(KC_QRUPD, flops, 3.0 * NC * tile_bytes(b), s, [&] {
            return fine ? launch_gemm_fine(a, ta, false, s) : launch_gemm(a, ta, false, s);
        });
    };
    const double bc = b3(b) * NC;
    auto orm = [&](int k) {
        GemmArgs a = gargs(GM_OA_W, dA, N, b, k, NC);
        a.smat = dS; a.cmat = dC; a.w1 = w1; a.w2 = w2;
        a.alpha = 1.0; a.beta = 0.0; a.maskA = MK_UNIT_LOWER;
        gemm(a, true, bc);
        a.mode = GM_OA_TW; a.maskA = MK_UPPER;
        gemm(a, trans, bc);
        a.mode = GM_OA_VW; a.maskA = MK_UNIT_LOWER; a.alpha = -1.0; a.beta = 1.0;
        gemm(a, false, bc);
    };
    auto tsm = [&](int k, int m) {
        GemmArgs a = gargs(GM_QA_W, dA, N, b, k, NC);
        a.m = m; a.smat = dS; a.cmat = dC; a.w1 = w1; a.w2 = w2;
        a.alpha = 1.0; a.beta = 1.0; a.maskA = MK_NONE;
        gemm(a, true, 2 * bc);
        a.mode = GM_QA_TW; a.maskA = MK_UPPER; a.beta = 0.0;
        gemm(a, trans, bc);
        a.mode = GM_QA_VW; a.maskA = MK_NONE; a.alpha = -1.0; a.beta = 1.0;
        gemm(a, false, 2 * bc);
    };
    if (trans) {
        for (int k = 0; k < MT; k++) {
            orm(k);
            for (int m = k + 1; m < MT; m++) tsm(k, m);
        }
    } else {
        for (int k = MT - 1; k >= 0; k--) {
            for (int m = MT - 1; m > k; m--) tsm(k, m);
            orm(k);
        }
    }
}

// ------------------------------------------------------------------ solve
int grow(double **p, size_t *cap, size_t need)
{
    if (*cap >= need) return 0;
    drop_graphs();
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc(p, need * 8) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "solve workspace of %zu bytes", need * 8);
    *cap = need;
    return 0;
}

// Triangular sweeps with the factored tiles F and the inverted diagonal tiles
// (q = 1; Cholesky: Uinv holds U_kk^-T; LU: U_kk^-1 and L_kk^-1) or blocks of
// q tiles (q > 1; lower triangular: Cholesky (U^T)^-1, LU L^-1 and (U^T)^-1):
//   Cholesky: U^T y = r, U x = y;  LU: L y = r, U x = y.
// Each step solves with one inverse (a product) and updates the rest of the
// vector from the step's q tile columns (forward) or rows (backward).
// r is destroyed (used as the running right-hand side of the forward sweep).
// f32: F, Uinv, Linv point at FP32 copies (mixed-precision correction solves)
int sweep_plan(int algo, int MT, int q, SweepStep *out, int cap)
{
    const int KT = MT / q;
    int n = 0;
    auto put = [&](int mode, int k, int trans, int ntasks, int inv, int vin, int vout) {
        if (n < cap && out) out[n] = SweepStep{mode, k, trans, ntasks, inv, vin, vout};
        n++;
    };
    // forward: y_k = inv_k r_k, then r_i -= op(F) y_k below the step
    for (int k = 0; k < KT; k++) {
        put(GV_DIAG, k, 0, q, algo == 0 ? 0 : 1, 0, 1);
        if (k + 1 < KT) put(algo == 0 ? GV_CHOL_FWD : GV_LU_FWD, k, algo == 0, MT - (k + 1) * q, 0, 1, 0);
    }
    // backward: x_k = inv_k y_k (Cholesky, and LU with blocks: (U^T)^-1 read transposed),
    // then y_i -= F x_k above the step
    for (int k = KT - 1; k >= 0; k--) {
        put(GV_DIAG, k, (algo == 0 || q > 1) ? 1 : 0, q, 0, 1, 2);
        if (k > 0) put(GV_BWD, k, 0, k * q, 0, 2, 1);
    }
    return n;
}

void sweeps(int algo, const void *F, const void *Uinv, const void *Linv, double *r, double *y, double *x,
            int N, int b, int nrhs, bool f32, hipStream_t s, Launcher &L, int q = 1)
{
    const int MT = N / b;
    std::vector<SweepStep> plan((size_t)sweep_plan(algo, MT, q, nullptr, 0));
    sweep_plan(algo, MT, q, plan.data(), (int)plan.size());
    double *vec[3] = {r, y, x};
    GemvArgs a;
    memset(&a, 0, sizeof a);
    a.mat = F; a.N = N; a.b = b; a.ld = N; a.nrhs = nrhs; a.f32 = f32 ? 1 : 0; a.q = q;
    for (const SweepStep &st : plan) {
        a.mode = st.mode; a.k = st.k; a.trans = st.trans; a.ntasks = st.ntasks;
        a.inv = st.inv ? Linv : Uinv;
        a.vin = vec[st.vin];
        a.vout = vec[st.vout];
        L.chk(launch_tile_gemv(a, s));
    }
}

// Inverses of the q x q tile blocks on the diagonal of the lower triangular
// T (upper: T = F^T from F's upper tiles; else F's unit lower tiles), into X
// (block k's tile (i, j) at ((kq + i) q + j) b^2) from the tile inverses D
// (dtrans: D holds the inverses of F's upper diagonal tiles, X takes their
// transposes).  Level by level (i = j + d):
//   X(i, j) = -X(i, i) sum_{m=j}^{i-1} T(i, m) X(m, j)
// one batched GEMM launch per term and one for the product with X(i, i).
// W1: q MT tiles (the tiles of T, same offsets as X); W2: (q-1) MT / q tiles.
int blkinv_plan(int q, BlkInvLaunch *out, int cap)
{
    int n = 0;
    for (int d = 1; d < q; d++) {
        for (int t = 0; t < d; t++) {
            if (n < cap && out) out[n] = BlkInvLaunch{GM_BI_W, d, t, 1.0, t > 0 ? 1.0 : 0.0};
            n++;
        }
        if (n < cap && out) out[n] = BlkInvLaunch{GM_BI_X, d, 0, -1.0, 0.0};
        n++;
    }
    return n;
}

void block_inverses(const void *F, bool f32, int N, int b, int q, bool upper, bool dtrans, const double *D,
                    double *X, double *W1, double *W2, hipStream_t s, Launcher &L)
{
    L.chk(launch_blk_gather(F, f32, N, b, q, upper, dtrans, D, X, W1, s));
    std::vector<BlkInvLaunch> plan((size_t)blkinv_plan(q, nullptr, 0));
    blkinv_plan(q, plan.data(), (int)plan.size());
    GemmArgs a;
    memset(&a, 0, sizeof a);
    a.mat = X; a.w1 = W1; a.w2 = W2; a.N = N; a.b = b; a.MT = N / b; a.bq = q; a.maskA = MK_NONE;
    const int KT = N / b / q;
    for (const BlkInvLaunch &l : plan) {
        a.mode = l.mode; a.bd = l.bd; a.bt = l.bt; a.alpha = l.alpha; a.beta = l.beta;
        a.ntasks = KT * (q - l.bd);
        L.chk(launch_gemm(a, false, false, s));
    }
}

int run_solve(int algo, MATRIX_desc A, double *B, int nrhs, int ldb, const palial_opts *o_in, palial_report *rep)
{
    int rc;
    if ((rc = ensure_init()) || (rc = check_desc(A))) return rc;
    palial_opts o = o_in ? *o_in : default_opts();
    const int N = A.matrix_size, bA = A.tile_size;
    // tile size of the internal factorization: the copy of A that is factored
    // is re-laid out, so it need not be the caller's b
    int b = o.factor_tile > 0 ? o.factor_tile : (bA > 256 && N % 256 == 0 ? 256 : bA);
    if (N % b) return fail(PALIAL_ERR_ARG, "factor tile %d does not divide N=%d", b, N);
    const int MT = N / b;
    if (!panel2_ok(b) || b % 64) return fail(PALIAL_ERR_ARG, "solve needs a factor tile with b %% 64 == 0, b <= 512 (b=%d)", b);
    if (bA % 64) return fail(PALIAL_ERR_ARG, "solve needs the matrix tile size %% 64 == 0 (b=%d)", bA);
    if (nrhs < 1 || nrhs > 16) return fail(PALIAL_ERR_ARG, "nrhs=%d outside 1..16", nrhs);
    if (!B || ldb < N) return fail(PALIAL_ERR_ARG, "bad right-hand side (ldb=%d)", ldb);
    if (o.precision < 0 || o.precision > 4)
        return fail(PALIAL_ERR_ARG, "precision %d not available", o.precision);
    const double tol = o.tol > 0 ? o.tol : 1e-14;
    const int max_iter = o.max_iter > 0 ? o.max_iter : 30;
    const size_t nn = (size_t)N * N, nv = (size_t)N * nrhs;
    hipStream_t s = g.s_main;
    if ((rc = ensure_events(events_needed(MT))) || (rc = ensure_dinv(b, MT))) return rc;
    if (o.precision > 0 && b % 128 == 0 && (rc = ensure_pack(b, MT))) return rc;
    // mixed tiers: the correction solves stream FP32 copies (half the bytes);
    // the factor itself carries BF16/FP32-level errors, so this costs nothing
    const bool f32 = o.precision > 0;
    // ... and on the split chain schedule with packed panels the factorization
    // itself runs on the FP32 copy (trailing updates move 8 instead of 16 bytes
    // per element; diagonal tiles are factored in FP64 copies)
    const bool fs = f32 && g.f32work && b % 128 == 0 && MT > 1 && g.dag == 2;
    if ((o.precision == 3 || o.precision == 4) && !fs)
        return fail(PALIAL_ERR_ARG, "%s needs the FP32 working copy on the split chain schedule "
                    "(factor tile %% 128 == 0, N > tile, PALIAL_DAG=2, PALIAL_F32WORK=1)",
                    o.precision == 3 ? "FP8" : "split BF16");
    // tiles per sweep step (the blocked steps' limits: launch_tile_gemv)
    int q = o.sweep_block > 0 ? o.sweep_block : g.sweep_block;
    if (q > 1 && (q > 8 || MT % q || MT < 2 * q || nrhs > 4 || (size_t)q * b * (nrhs == 1 ? 1 : 4) > 4096)) {
        if (o.sweep_block > 1) return fail(PALIAL_ERR_ARG, "sweep block %d: needs q <= 8 dividing N/b=%d with "
                                           "N/b >= 2q, nrhs <= 4 and q*b*(nrhs > 1 ? 4 : 1) <= 4096", q, MT);
        q = 1;
    }
    const int nf = algo == 0 ? 1 : 2;                     // triangular factors swept
    const size_t tinv = (size_t)nf * MT * b * b;          // tile inverses
    const size_t binv = q > 1 ? (size_t)q * tinv : 0;     // block inverses
    if ((!fs || b != bA) && (rc = grow(&g.d_F, &g.F_cap, nn))) return rc;
    if (fs && (rc = grow(&g.d_sub, &g.sub_cap, (size_t)b * b))) return rc;
    if ((rc = grow(&g.d_inv, &g.inv_cap, tinv + binv))) return rc;
    // W1: q MT tiles (the tiles of T at the offsets of X); W2: (q - 1) MT / q tiles
    if (q > 1 && (rc = grow(&g.d_blk, &g.blk_cap, ((size_t)q * MT + (size_t)(q - 1) * (MT / q)) * b * b))) return rc;
    const size_t inv32 = q > 1 ? binv : tinv;             // inverses the FP32 sweeps read
    if (f32 && (rc = grow(&g.d_F32, &g.F32_cap, (nn + inv32) / 2 + 2))) return rc;
    float *F32 = (float *)g.d_F32, *Inv32 = F32 + nn;
    const size_t npart = residual_scratch(N, bA, nrhs);
    if ((rc = grow(&g.d_vec, &g.vec_cap, 6 * nv + 2 * (size_t)N + 64 + npart))) return rc;
    double *X = g.d_vec, *R = X + nv, *Y = R + nv, *D = Y + nv, *Bd = D + nv, *R2 = Bd + nv;
    double *rowsum = R2 + nv, *red = rowsum + N;     // red: [0]=||A||, [1..]=col maxima
    double *part = red + 64;
    double *Uinv = g.d_inv, *Linv = g.d_inv + (size_t)MT * b * b;
    // block inverses (q > 1): (U^T)^-1, then (LU) L^-1
    double *XU = g.d_inv + tinv, *XL = XU + (size_t)q * MT * b * b;
    double *BW1 = g.d_blk, *BW2 = g.d_blk + (size_t)q * MT * b * b;

    // inputs on the device
    double *dA = (double *)A.matrix;
    bool ownA = false;
    if (!o.device_matrix) {
        if (hipMalloc(&dA, nn * 8) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "device copy of A");
        ownA = true;
        HIPCHK(hipMemcpyAsync(dA, A.matrix, nn * 8, hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipMemcpy2DAsync(Bd, (size_t)N * 8, B, (size_t)ldb * 8, (size_t)N * 8, nrhs,
                            o.device_matrix ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    hipEvent_t e0, e1, e2;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventCreate(&e2));
    auto done = [&](int code) {
        (void)hipStreamSynchronize(s);
        if (ownA) (void)hipFree(dA);
        (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(e2);
        return code;
    };
    Launcher L;
    HIPCHK(hipEventRecord(e0, s));
    // factorization of a copy (A stays FP64 and intact for the residuals),
    // with sigma |A(i,i)| added to its diagonal when sigma > 0
    double *Fw = fs ? (double *)(void *)F32 : g.d_F;      // the matrix the DAG factors
    // inverses of the diagonal tiles: one batched strip TRSM per factor with the
    // identity as right-hand side (Cholesky: U^-T; LU: U^-1 and L^-1)
    auto inverses = [&](int kind, double *out, const double *dbase) {
        TrsmArgs t;
        memset(&t, 0, sizeof t);
        t.mat = Fw; t.single = out; t.N = N; t.b = b; t.MT = MT; t.ntiles = MT;
        t.batch = 1; t.ident = 1; t.dinv_stride = 64L * b; t.f32 = fs ? 1 : 0;
        L.chk(launch_trsm2(kind, t, dbase, s));
    };
    auto factor_once = [&](double sigma) {
        if (fs) {
            // FP32 copy (Cholesky: upper tiles only, the ones the factorization reads)
            if (b != bA) L.chk(launch_relayout(g.d_F, b, dA, bA, N, s));
            L.chk(launch_to_f32(b == bA ? dA : g.d_F, F32, N, b, algo == 0, 0, s));
        } else if (b == bA) {
            L.chk(launch_copy(g.d_F, dA, nn * 8, s));
        } else {
            L.chk(launch_relayout(g.d_F, b, dA, bA, N, s));
        }
        if (sigma > 0.0) L.chk(launch_add_diag(Fw, fs, N, b, dA, bA, sigma, s));
        L.note(hipMemsetAsync(g.d_info, 0, sizeof(int), s));
        trace_begin(Fw, N, b, algo, s, L);
        if (algo == 0) dag_cholesky(Fw, N, b, MT > 1, s, g.s_panel, g.d_info, L, o.precision, fs);
        else dag_lu(Fw, N, b, MT > 1, s, g.s_panel, L, o.precision, fs);
        trace_end();
        if (algo == 0) {
            inverses(TR_CHOL, Uinv, g.d_dinv);
        } else {
            inverses(TR_LUR, Uinv, g.d_dinv);
            inverses(TR_LUL, Linv, g.d_dinv + (size_t)32 * b);
        }
        if (q > 1) {
            block_inverses(Fw, fs, N, b, q, true, algo == 1, Uinv, XU, BW1, BW2, s, L);
            if (algo == 1) block_inverses(Fw, fs, N, b, q, false, false, Linv, XL, BW1, BW2, s, L);
        }
        if (f32) {
            if (!fs) L.chk(launch_to_f32(g.d_F, F32, N, b, algo == 0, 0, s));
            L.chk(launch_to_f32(q > 1 ? g.d_inv + tinv : Uinv, Inv32, 0, b, false, inv32, s));
        }
    };
    // FP64 solve in place of this one (a mixed tier that broke down or did not
    // converge); its report adds the time spent here
    const bool fallback_ok = o.precision > 0 && !o.no_fallback;
    auto fallback = [&](int shifts_done) -> int {
        HIPCHK(hipEventRecord(e2, s));
        HIPCHK(hipStreamSynchronize(s));
        float spent = 0;
        (void)hipEventElapsedTime(&spent, e0, e2);
        done(0);
        palial_opts o2 = o;
        o2.precision = 0;
        o2.factor_tile = 0;
        o2.sweep_block = 0;     // the FP64 solve picks its own tile, so its own sweep steps
        palial_report r2;
        memset(&r2, 0, sizeof r2);
        const int rc2 = run_solve(algo, A, B, nrhs, ldb, &o2, &r2);
        r2.factor_ms += spent;
        r2.shifts = shifts_done;
        if (rep) *rep = r2;
        return rc2;
    };
    // unit roundoff of the tier's update operands: the diagonal shift that
    // keeps a low-precision Cholesky from breaking down (Higham and Pranesh,
    // SIAM J. Sci. Comput. 43(1), 2021) starts at 4 u and grows 16x per retry
    const double u_tier = o.precision == 1 ? 0x1p-24 : o.precision == 2 ? 0x1p-8 : o.precision == 3 ? 0x1p-4
                        : o.precision == 4 ? 0x1p-16 : 0.0;
    int shifts = 0;
    double sigma = 0.0;
    int info = 0;
    factor_once(0.0);
    HIPCHK(hipMemcpyAsync(&info, g.d_info, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if ((rc = trace_collect())) return done(rc);
    while (info > 0 && fallback_ok && algo == 0 && shifts < 3) {
        sigma = sigma == 0.0 ? 4.0 * u_tier : 16.0 * sigma;
        shifts++;
        factor_once(sigma);
        HIPCHK(hipMemcpyAsync(&info, g.d_info, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if ((rc = trace_collect())) return done(rc);
    }
    if (info > 0 && fallback_ok) return fallback(shifts);
    const void *sF = f32 ? (const void *)F32 : (const void *)g.d_F;
    const void *sU, *sL;
    if (q > 1) {
        // FP32 copies in the order of the FP64 blocks
        sU = f32 ? (const void *)Inv32 : (const void *)XU;
        sL = f32 ? (const void *)(Inv32 + (size_t)q * MT * b * b) : (const void *)XL;
    } else {
        sU = f32 ? (const void *)Inv32 : (const void *)Uinv;
        sL = f32 ? (const void *)(Inv32 + (size_t)MT * b * b) : (const void *)Linv;
    }
    HIPCHK(hipEventRecord(e1, s));
    if (L.err != hipSuccess) return done(fail(PALIAL_ERR_HIP, "factorization launch: %s", hipGetErrorString(L.err)));
    // ||b||_inf now; ||A||_inf comes with the first residual (same pass over A)
    L.chk(launch_colmax(Bd, N, N, nrhs, red + 1, s));
    // x0 = F^{-1} b
    L.chk(launch_copy(R, Bd, nv * 8, s));
    sweeps(algo, sF, sU, sL, R, Y, X, N, b, nrhs, f32, s, L, q);
    // the Cholesky residual reads only the upper tiles of A (the symmetric
    // matrix the factorization sees)
    const bool sym = (algo == 0 && nrhs == 1);
    double hred[64];
    double normA = 0.0, normB[16];
    double eta = 0.0, eta_prev = 0.0;
    int it = 0, stall = 0, gmres_total = 0, gm_slow = 0;
    // 0 LU-IR, 1 GMRES-IR (always after a shift: F then factors A + sigma |diag A|)
    int mode = (o.refine == 1 || (shifts > 0 && o.refine != 2)) ? 1 : 0;
    // GMRES workspace: Krylov basis V (N x (GM+1)), two vectors, a zero
    // right-hand side, the Hessenberg column on the device
    constexpr int GM = 30;
    const double gm_tol = 1e-8;              // inner relative residual per refinement step
    double *V = nullptr, *Wv = nullptr, *T1 = nullptr, *Z0 = nullptr, *hd = nullptr;
    if (o.refine != 2) {
        if ((rc = grow(&g.d_gm, &g.gm_cap, (size_t)N * (GM + 4) + 2 * (GM + 2)))) return done(rc);
        V = g.d_gm; Wv = V + (size_t)N * (GM + 1); T1 = Wv + N; Z0 = T1 + N; hd = Z0 + N;
        HIPCHK(hipMemsetAsync(Z0, 0, (size_t)N * 8, s));
    }
    auto dots = [&](int k, const double *w, double *out) -> int {
        L.chk(launch_multi_dot(V, (size_t)N, k, w, N, hd, s));
        if (hipMemcpyAsync(out, hd, (size_t)k * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return fail(PALIAL_ERR_HIP, "GMRES reduction");
        return 0;
    };
    auto norm2 = [&](const double *w, double &out) -> int {
        L.chk(launch_multi_dot(w, (size_t)N, 1, w, N, hd, s));
        if (hipMemcpyAsync(&out, hd, 8, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
            return fail(PALIAL_ERR_HIP, "GMRES norm");
        out = sqrt(out);
        return 0;
    };
    // d = solution of A d = r (one column) by GMRES on F^-1 A d = F^-1 r, d0 = 0
    auto gmres = [&](const double *rj, double *dj, int &inner) -> int {
        int err;
        std::vector<double> H((size_t)(GM + 1) * GM, 0.0), cs(GM, 0.0), sn(GM, 0.0), gv(GM + 1, 0.0), h(GM + 1);
        L.chk(launch_copy(T1, rj, (size_t)N * 8, s));
        sweeps(algo, sF, sU, sL, T1, Y, V, N, b, 1, f32, s, L, q);          // v0 = F^-1 r
        double beta;
        if ((err = norm2(V, beta))) return err;
        inner = 0;
        if (!(beta > 0.0)) {
            L.chk(launch_vh_update(V, (size_t)N, 0, hd, dj, 0.0, N, s));     // d = 0
            return 0;
        }
        L.chk(launch_vh_update(V, (size_t)N, 0, hd, V, 1.0 / beta, N, s));
        gv[0] = beta;
        int k = 0;
        for (; k < GM; k++) {
            double *vk = V + (size_t)k * N, *vn = V + (size_t)(k + 1) * N;
            // w = F^-1 A v_k  (the residual kernel with a zero right-hand side gives -A v_k)
            L.chk(launch_residual(dA, N, bA, vk, Z0, T1, N, 1, part, sym, nullptr, s));
            sweeps(algo, sF, sU, sL, T1, Y, Wv, N, b, 1, f32, s, L, q);
            L.chk(launch_vh_update(V, (size_t)N, 0, hd, Wv, -1.0, N, s));
            // classical Gram-Schmidt, twice
            std::vector<double> h2(k + 1);
            if ((err = dots(k + 1, Wv, h.data()))) return err;
            L.chk(launch_vh_update(V, (size_t)N, k + 1, hd, Wv, 1.0, N, s));
            if ((err = dots(k + 1, Wv, h2.data()))) return err;
            L.chk(launch_vh_update(V, (size_t)N, k + 1, hd, Wv, 1.0, N, s));
            for (int i = 0; i <= k; i++) h[i] += h2[i];
            double hn;
            if ((err = norm2(Wv, hn))) return err;
            h[k + 1] = hn;
            if (hn > 0.0) L.chk(launch_vh_update(V, (size_t)N, 0, hd, Wv, 1.0 / hn, N, s));
            L.chk(launch_copy(vn, Wv, (size_t)N * 8, s));
            // Givens rotations on the new Hessenberg column
            for (int i = 0; i < k; i++) {
                const double a = cs[i] * h[i] + sn[i] * h[i + 1];
                h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
                h[i] = a;
            }
            const double r = hypot(h[k], h[k + 1]);
            cs[k] = r > 0 ? h[k] / r : 1.0;
            sn[k] = r > 0 ? h[k + 1] / r : 0.0;
            h[k] = r;
            h[k + 1] = 0.0;
            gv[k + 1] = -sn[k] * gv[k];
            gv[k] = cs[k] * gv[k];
            for (int i = 0; i <= k; i++) H[(size_t)i + (size_t)k * (GM + 1)] = h[i];
            inner++;
            if (fabs(gv[k + 1]) <= gm_tol * beta || !(hn > 0.0)) { k++; break; }
        }
        // y = R^-1 g, d = V y
        std::vector<double> y(k, 0.0);
        for (int i = k - 1; i >= 0; i--) {
            double a = gv[i];
            for (int c = i + 1; c < k; c++) a -= H[(size_t)i + (size_t)c * (GM + 1)] * y[c];
            y[i] = a / H[(size_t)i + (size_t)i * (GM + 1)];
        }
        for (int i = 0; i < k; i++) y[i] = -y[i];
        if (hipMemcpyAsync(hd, y.data(), (size_t)k * 8, hipMemcpyHostToDevice, s) != hipSuccess)
            return fail(PALIAL_ERR_HIP, "GMRES coefficients");
        L.chk(launch_vh_update(V, (size_t)N, k, hd, dj, 0.0, N, s));     // d = 0*d - V(-y)
        return 0;
    };
    for (;; it++) {
        // r = b - A x ; eta = max_j ||r_j|| / (||A|| ||x_j|| + ||b_j||)
        L.chk(launch_residual(dA, N, bA, X, Bd, R, N, nrhs, part, sym, it == 0 ? rowsum : nullptr, s));
        if (it == 0) L.chk(launch_colmax(rowsum, N, N, 1, red, s));
        L.chk(launch_colmax(R, N, N, nrhs, red + 1 + nrhs, s));
        L.chk(launch_colmax(X, N, N, nrhs, red + 1 + 2 * nrhs, s));
        HIPCHK(hipMemcpyAsync(hred, red, (1 + 3 * nrhs) * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (L.err != hipSuccess) return done(fail(PALIAL_ERR_HIP, "refinement launch: %s", hipGetErrorString(L.err)));
        if (it == 0) {
            if (info > 0) return done(fail(info, "factorization failed at column %d", info));
            normA = hred[0];
            for (int j = 0; j < nrhs; j++) normB[j] = hred[1 + j];
        }
        eta = 0.0;
        for (int j = 0; j < nrhs; j++) {
            const double den = normA * hred[1 + 2 * nrhs + j] + normB[j];
            const double rn = hred[1 + nrhs + j];
            // den == 0 only for b = 0 and x = 0 (then r = 0); a NaN or infinite
            // norm makes eta NaN
            const double ej = (den > 0 && std::isfinite(den)) ? rn / den : (den == 0 && rn == 0 ? 0.0 : NAN);
            if (!(ej <= eta)) eta = ej;      // NaN propagates
        }
        // a factor whose solution overflowed or went NaN cannot be refined
        if (eta <= tol || it >= max_iter || !std::isfinite(eta)) break;
        // with the FP64 fallback allowed, GMRES-IR that stalls (two steps in a
        // row that gain less than 4x) or runs long (more than one restart
        // cycle in all) would cost more than the FP64 solve
        if (fallback_ok && mode == 1 && it > 0) {
            gm_slow = (eta <= 0.25 * eta_prev) ? 0 : gm_slow + 1;
            if (gm_slow >= 2 || gmres_total > 2 * GM) break;
        }
        // auto: LU-IR until it fails twice in a row to halve eta, then GMRES-IR
        if (o.refine == 0 && mode == 0 && it > 0 && !(eta <= 0.5 * eta_prev)) {
            if (++stall >= 2) mode = 1;
        } else {
            stall = 0;
        }
        eta_prev = eta;
        if (mode == 1) {
            // GMRES-IR: A d = r by GMRES preconditioned on the left with F^-1
            bool abandon = false;
            for (int j = 0; j < nrhs; j++) {
                int inner = 0;
                if ((rc = gmres(R + (size_t)j * N, D + (size_t)j * N, inner))) return done(rc);
                gmres_total += inner;
                // one GMRES(30) cycle costs about as much as the FP64 solve
                if (fallback_ok && gmres_total > GM && j + 1 < nrhs) { abandon = true; break; }
            }
            if (abandon) break;
        } else {
            // LU-IR: d = F^{-1} r
            L.chk(launch_copy(R2, R, nv * 8, s));
            sweeps(algo, sF, sU, sL, R2, Y, D, N, b, nrhs, f32, s, L, q);
        }
        L.chk(launch_axpy(X, D, N, N, nrhs, s));
    }
    // a mixed tier that did not converge hands over to the FP64 solve before
    // anything is written to B (the fallback reads B as the right-hand side)
    if (!(eta <= tol) && fallback_ok) {
        prof_collect();
        return fallback(shifts);
    }
    HIPCHK(hipEventRecord(e2, s));
    HIPCHK(hipMemcpy2DAsync(B, (size_t)ldb * 8, X, (size_t)N * 8, (size_t)N * 8, nrhs,
                            o.device_matrix ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    float fms = 0, sms = 0;
    (void)hipEventElapsedTime(&fms, e0, e1);
    (void)hipEventElapsedTime(&sms, e1, e2);
    prof_collect();
    if (rep) {
        const double n = N;
        rep->factor_ms = fms;
        rep->solve_ms = sms;
        rep->gflop = (algo == 0 ? n * n * n / 3 : 2 * n * n * n / 3) / 1e9;
        rep->tflops = (algo == 0 ? n * n * n / 3 : 2 * n * n * n / 3) / ((fms + sms) * 1e-3) / 1e12;
        rep->backward_error = eta;
        rep->iterations = it;
        rep->gmres_iterations = gmres_total;
        rep->info = info;
        rep->nlaunches = L.launches;
        rep->precision_used = o.precision;
        rep->shifts = shifts;
    }
    if (!(eta <= tol)) return done(fail(PALIAL_ERR_NOCONV, "refinement stopped at backward error %.3e after %d steps", eta, it));
    return done(0);
}

// ------------------------------------------------------------------ run helper
enum Algo { ALG_CHOL = 0, ALG_LU = 1, ALG_QR = 2 };

int run_algo(int algo, MATRIX_desc A, MATRIX_desc *S, const palial_opts *o_in, palial_report *rep)
{
    int rc;
    if ((rc = ensure_init())) return rc;
    if ((rc = check_desc(A))) return rc;
    if (S && (rc = check_desc(*S))) return rc;
    if (S && (S->matrix_size != A.matrix_size || S->tile_size != A.tile_size))
        return fail(PALIAL_ERR_ARG, "S must have the same N and b as A");
    palial_opts o = o_in ? *o_in : default_opts();
    const int N = A.matrix_size, b = A.tile_size, MT = N / b;
    const size_t bytes = (size_t)N * N * 8;
    hipStream_t s0 = (o.device_matrix && o.stream) ? (hipStream_t)o.stream : g.s_main;
    hipStream_t s1 = g.s_panel;
    // schedule of this call (palial_opts.schedule): Cholesky / LU 1 diagonal
    // chain with the split update, 2 diagonal chain, 3 row look-ahead, 4 one
    // stream; QR: panel chains in flight (1..8); 0: the library default
    if (o.schedule < 0 || o.schedule > (algo == ALG_QR ? QR_MAXS : 4))
        return fail(PALIAL_ERR_ARG, "schedule %d out of range", o.schedule);
    const int dag = (algo != ALG_QR && o.schedule >= 1 && o.schedule <= 3) ? 3 - o.schedule : g.dag;
    const int qr_chains = (algo == ALG_QR && o.schedule > 0) ? std::min(o.schedule, g.qr_streams) : g.qr_streams;
    const bool lookahead = !o.no_lookahead && MT > 1 && algo != ALG_QR && !(algo != ALG_QR && o.schedule == 4);
    // QR's two-stream schedule runs eagerly (a graph replay drops the bulk stream's CU mask)
    const bool qr_two = algo == ALG_QR && !o.no_lookahead && MT > 1;
    // the factorizations here are FP64: their chain schedule uses no CU mask (bulk_stream)
    const bool use_graph = !o.no_graph && !g_prof.on && !g_tr.on && !qr_two;
    if ((rc = ensure_events(events_needed(MT)))) return rc;
    if (algo == ALG_QR && (rc = ensure_ws(b, MT))) return rc;
    if ((rc = ensure_dinv(b, MT))) return rc;

    double *dA = nullptr, *dS = nullptr;
    bool owned = false;
    hipEvent_t e0, e1, e2, e3;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventCreate(&e2));
    HIPCHK(hipEventCreate(&e3));
    auto cleanup = [&]() {
        if (owned) {
            if (dA) (void)hipFree(dA);
            if (dS) (void)hipFree(dS);
        }
        (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(e2); (void)hipEventDestroy(e3);
    };
    if (o.device_matrix) {
        dA = (double *)A.matrix;
        dS = S ? (double *)S->matrix : nullptr;
    } else {
        owned = true;
        if (hipMalloc(&dA, bytes) != hipSuccess) { cleanup(); return fail(PALIAL_ERR_NOMEM, "device alloc %zu B", bytes); }
        if (S && hipMalloc(&dS, bytes) != hipSuccess) { cleanup(); return fail(PALIAL_ERR_NOMEM, "device alloc %zu B", bytes); }
    }
    hipError_t he = hipEventRecord(e0, s0);
    if (!o.device_matrix) {
        if (he == hipSuccess) he = hipMemcpyAsync(dA, A.matrix, bytes, hipMemcpyHostToDevice, s0);
        if (S && he == hipSuccess) he = hipMemcpyAsync(dS, S->matrix, bytes, hipMemcpyHostToDevice, s0);
    }
    if (he == hipSuccess) he = hipMemsetAsync(g.d_info, 0, sizeof(int), s0);
    if (he == hipSuccess) he = hipEventRecord(e1, s0);
    if (he != hipSuccess) { cleanup(); return fail(PALIAL_ERR_HIP, "staging before the factorization: %s", hipGetErrorString(he)); }

    if (qr_two && (rc = qr_streams_up())) {
        qr_streams_down();
        cleanup();
        return rc;
    }
    Launcher L;
    auto issue = [&](hipStream_t st) {
        const int dag_default = g.dag;
        g.dag = dag;        // the dispatchers read the schedule from the context (calls are serialised)
        if (algo == ALG_CHOL) dag_cholesky(dA, N, b, lookahead, st, s1, g.d_info, L);
        else if (algo == ALG_LU) dag_lu(dA, N, b, lookahead, st, s1, L);
        else dag_qr(dA, dS, N, b, st, s1, qr_two, L, qr_chains);
        g.dag = dag_default;
    };
    long nodes = 0;
    if (use_graph) {
        auto key = std::make_tuple(algo, (void *)dA, (void *)dS, N, b, (int)lookahead + 2 * dag);
        auto it = g.graphs.find(key);
        if (it == g.graphs.end() || it->second.exec == nullptr) {
            Graph G;
            hipStream_t cs = g.s_main;   // capture on the library stream
            L.capturing = true;
            hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) { cleanup(); return fail(PALIAL_ERR_HIP, "begin capture: %s", hipGetErrorString(e)); }
            issue(cs);
            e = hipStreamEndCapture(cs, &G.graph);
            if (e != hipSuccess || L.err != hipSuccess) {
                if (G.graph) (void)hipGraphDestroy(G.graph);
                cleanup();
                return fail(PALIAL_ERR_HIP, "graph capture failed: %s / %s", hipGetErrorString(e),
                            hipGetErrorString(L.err));
            }
            e = hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(G.graph);
                cleanup();
                return fail(PALIAL_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
            }
            size_t n = 0;
            if (hipGraphGetNodes(G.graph, nullptr, &n) != hipSuccess) n = 0;
            G.nodes = (long)n;
            // the key holds the device address: calls on host buffers stage A
            // in a fresh allocation each time, so the cache is bounded
            if (g.graphs.size() >= MAX_GRAPHS) drop_factor_graphs();
            g.graphs[key] = G;
            it = g.graphs.find(key);
        }
        // order the replay after the copies on s0
        hipEvent_t gate = g.events[2 * MT + 3];
        hipError_t e = hipEventRecord(gate, s0);
        if (e == hipSuccess) e = hipStreamWaitEvent(g.s_main, gate, 0);
        if (e == hipSuccess) e = hipGraphLaunch(it->second.exec, g.s_main);
        if (e == hipSuccess && s0 != g.s_main) {
            e = hipEventRecord(gate, g.s_main);
            if (e == hipSuccess) e = hipStreamWaitEvent(s0, gate, 0);
        }
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(s0);
            cleanup();
            return fail(PALIAL_ERR_HIP, "graph launch: %s", hipGetErrorString(e));
        }
        nodes = it->second.nodes;
    } else {
        trace_begin(dA, N, b, algo, s0, L);
        issue(s0);
        trace_end();
        if (L.err != hipSuccess) {
            (void)hipStreamSynchronize(s0);
            cleanup();
            return fail(PALIAL_ERR_HIP, "launch failed: %s", hipGetErrorString(L.err));
        }
        nodes = L.launches;
    }
    he = hipEventRecord(e2, s0);
    int info = 0;
    if (he == hipSuccess) he = hipMemcpyAsync(&info, g.d_info, sizeof(int), hipMemcpyDeviceToHost, s0);
    if (!o.device_matrix) {
        if (he == hipSuccess) he = hipMemcpyAsync(A.matrix, dA, bytes, hipMemcpyDeviceToHost, s0);
        if (S && he == hipSuccess) he = hipMemcpyAsync(S->matrix, dS, bytes, hipMemcpyDeviceToHost, s0);
    }
    if (he == hipSuccess) he = hipEventRecord(e3, s0);
    hipError_t e = hipStreamSynchronize(s0);
    if (he != hipSuccess) e = he;
    if (e != hipSuccess) { cleanup(); return fail(PALIAL_ERR_HIP, "synchronize: %s", hipGetErrorString(e)); }
    if ((rc = trace_collect())) { cleanup(); return rc; }
    float h2d = 0, fac = 0, d2h = 0;
    (void)hipEventElapsedTime(&h2d, e0, e1);
    (void)hipEventElapsedTime(&fac, e1, e2);
    (void)hipEventElapsedTime(&d2h, e2, e3);
    prof_collect();
    if (rep) {
        const double n = N;
        const double flops = (algo == ALG_CHOL) ? n * n * n / 3 : (algo == ALG_LU) ? 2 * n * n * n / 3 : 4 * n * n * n / 3;
        rep->factor_ms = fac;
        rep->h2d_ms = h2d;
        rep->d2h_ms = d2h;
        rep->gflop = flops / 1e9;
        rep->tflops = fac > 0 ? flops / (fac * 1e-3) / 1e12 : 0;
        rep->info = info;
        const long mt = MT;
        rep->ntasks = (algo == ALG_CHOL) ? mt + mt * (mt - 1) + mt * (mt - 1) * (mt - 2) / 6
                    : mt + mt * (mt - 1) + (mt - 1) * mt * (2 * mt - 1) / 6;
        rep->nlaunches = nodes;
    }
    cleanup();
    return info;
}

// single-tile staging: copy n x n tiles in/out when the caller passes host pointers
struct Stage {
    std::vector<std::pair<double *, const double *>> back;  // (host dst, device src)
    std::vector<double *> owned;
    size_t bytes;
    bool dev;
    Stage(size_t by, bool d) : bytes(by), dev(d) {}
    ~Stage()
    {
        for (double *p : owned) (void)hipFree(p);
    }
    double *in(const double *h, bool copy_back)
    {
        if (dev || !h) return (double *)h;
        double *d = nullptr;
        if (hipMalloc(&d, bytes) != hipSuccess) return nullptr;
        owned.push_back(d);
        if (hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, g.s_main) != hipSuccess) return nullptr;
        if (copy_back) back.push_back({(double *)h, d});
        return d;
    }
    double *scratch(size_t by)
    {
        double *d = nullptr;
        if (hipMalloc(&d, by) != hipSuccess) return nullptr;
        owned.push_back(d);
        return d;
    }
    int finish()
    {
        for (auto &pr : back) HIPCHK(hipMemcpyAsync(pr.first, pr.second, bytes, hipMemcpyDeviceToHost, g.s_main));
        HIPCHK(hipStreamSynchronize(g.s_main));
        return 0;
    }
};

int check_tile(int n, int lda, const void *A)
{
    if (n < 0) return fail(PALIAL_ERR_ARG, "n < 0");
    if (n > 1024) return fail(PALIAL_ERR_ARG, "tile size %d above 1024", n);
    if (n > 0 && !A) return fail(PALIAL_ERR_ARG, "NULL tile pointer");
    if (lda != n) return fail(PALIAL_ERR_ARG, "leading dimension %d must equal the tile size %d", lda, n);
    return 0;
}

GemmArgs single(const double *A, const double *B, double *C, const double *D, double *E, int n,
                double alpha, double beta, int mask)
{
    GemmArgs a;
    memset(&a, 0, sizeof a);
    a.mode = GM_SINGLE;
    a.A = A; a.B = B; a.C = C; a.D = D; a.E = E;
    a.N = n; a.b = n; a.MT = 1; a.T = 1; a.ntasks = 1;
    a.alpha = alpha; a.beta = beta; a.maskA = mask;
    return a;
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

int palial_hip_version(void) { return 100; }

const char *palial_hip_strerror(int code)
{
    switch (code) {
    case PALIAL_OK: return "success";
    case PALIAL_ERR_ARG: return "invalid argument";
    case PALIAL_ERR_HIP: return "HIP runtime error";
    case PALIAL_ERR_NOMEM: return "device out of memory";
    case PALIAL_ERR_NODEV: return "no usable gfx950 device";
    case PALIAL_ERR_STATE: return "library not initialised";
    case PALIAL_ERR_NOCONV: return "iterative refinement did not converge";
    default: return code > 0 ? "numerical failure (info > 0: non-positive pivot)" : "unknown error";
    }
}

const char *palial_hip_last_error(void) { return g_last.c_str(); }

int palial_hip_init(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g.init) return device == g.device || device < 0 ? 0 : fail(PALIAL_ERR_STATE, "already initialised on device %d", g.device);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(PALIAL_ERR_NODEV, "no HIP device visible");
    if (device < 0) device = 0;
    if (device >= n) return fail(PALIAL_ERR_NODEV, "device %d out of range (%d devices)", device, n);
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(PALIAL_ERR_NODEV, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
    if (device >= MAX_DEVICES) return fail(PALIAL_ERR_NODEV, "device %d above the supported %d", device, MAX_DEVICES);
    if (!g_scratch[device]) {
        void *sp = nullptr;
        HIPCHK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
        if (!sp) return fail(PALIAL_ERR_HIP, "no device address for the no-allocation scratch");
        g_scratch[device] = (double *)sp;
    }
    // PALIAL_RESERVE_CUS=n (default 8): the bulk stream of the chain schedule
    // runs on all CUs but the first n, which stay free for the diagonal chain.
    // PALIAL_DAG=1 selects the chain schedule with one update launch per
    // step, PALIAL_DAG=0 the row look-ahead schedule.
    const char *res = getenv("PALIAL_RESERVE_CUS");
    const char *dag = getenv("PALIAL_DAG");
    g.dag = dag ? atoi(dag) : 2;
    if (g.dag < 0 || g.dag > 2) return fail(PALIAL_ERR_ARG, "PALIAL_DAG=%s: expected 0, 1 or 2", dag);
    g.reserve = res ? atoi(res) : 8;
    // PALIAL_SUBTILE=0 factors diagonal tiles with the one-workgroup kernel only
    const char *sub = getenv("PALIAL_SUBTILE");
    g.subtile = sub ? atoi(sub) : 1;
    // PALIAL_F32WORK=0 keeps the FP64 working copy in the mixed-tier solves
    const char *fw = getenv("PALIAL_F32WORK");
    g.f32work = fw ? atoi(fw) : 1;
    // PALIAL_PAIR=1: BF16 and split-BF16 solves apply two steps' updates per pass over the trailing tiles
    const char *pr = getenv("PALIAL_PAIR");
    g.pair = pr ? atoi(pr) : 0;
    // PALIAL_SWEEP_BLOCK=q: default tiles per step of the correction sweeps
    const char *sb = getenv("PALIAL_SWEEP_BLOCK");
    g.sweep_block = sb ? atoi(sb) : 1;
    if (g.sweep_block < 1 || g.sweep_block > 8) return fail(PALIAL_ERR_ARG, "PALIAL_SWEEP_BLOCK=%s: expected 1..8", sb);
    HIPCHK(hipStreamCreateWithFlags(&g.s_main, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&g.s_panel, hipStreamNonBlocking));
    // PALIAL_CU_POLICY (parse_cu_policy) replaces PALIAL_RESERVE_CUS / PALIAL_CHAIN_CUS
    const char *pol = getenv("PALIAL_CU_POLICY");
    if (pol) {
        const int ncu = prop.multiProcessorCount, words = (ncu + 31) / 32;
        std::vector<uint32_t> m[3];
        bool given[3];
        if (parse_cu_policy(pol, ncu, m, given))
            return fail(PALIAL_ERR_ARG, "PALIAL_CU_POLICY=%s: expected entries bulk|aux|chain=lo-hi[+lo-hi...] "
                        "separated by ',' with CUs below %d", pol, ncu);
        g.reserve = 1;
        HIPCHK(hipExtStreamCreateWithCUMask(&g.s_bulk, words, m[0].data()));
        HIPCHK(hipExtStreamCreateWithCUMask(&g.s_aux, words, m[1].data()));
        if (given[2]) HIPCHK(hipExtStreamCreateWithCUMask(&g.s_chainm, words, m[2].data()));
    } else if (g.reserve > 0 && g.reserve < prop.multiProcessorCount / 2) {
        const int ncu = prop.multiProcessorCount, words = (ncu + 31) / 32;
        std::vector<uint32_t> mu(words, 0);
        for (int cu = g.reserve; cu < ncu; cu++) mu[cu / 32] |= 1u << (cu % 32);
        HIPCHK(hipExtStreamCreateWithCUMask(&g.s_bulk, words, mu.data()));
        HIPCHK(hipExtStreamCreateWithCUMask(&g.s_aux, words, mu.data()));
        const char *cm = getenv("PALIAL_CHAIN_CUS");
        if (cm && atoi(cm) == 1) {
            for (auto &x : mu) x = ~x;
            if (ncu % 32) mu[words - 1] &= (1u << (ncu % 32)) - 1;
            HIPCHK(hipExtStreamCreateWithCUMask(&g.s_chainm, words, mu.data()));
        }
    } else {
        g.reserve = 0;
        g.s_bulk = g.s_main;
        HIPCHK(hipStreamCreateWithFlags(&g.s_aux, hipStreamNonBlocking));
    }
    // QR schedule: PALIAL_QR_STREAMS panel chains and as many update streams;
    // the update streams stay off the first PALIAL_QR_RESERVE CUs
    const char *qs = getenv("PALIAL_QR_STREAMS");
    g.qr_streams = qs ? atoi(qs) : 8;
    if (g.qr_streams < 1) g.qr_streams = 1;
    if (g.qr_streams > QR_MAXS) g.qr_streams = QR_MAXS;
    const char *qres = getenv("PALIAL_QR_RESERVE");
    g.qr_reserve = qres ? atoi(qres) : 32;
    if (g.qr_reserve < 0 || g.qr_reserve >= prop.multiProcessorCount / 2) g.qr_reserve = 0;
    g.ncu = prop.multiProcessorCount;
    HIPCHK(hipMalloc(&g.d_info, sizeof(int)));
    panel_init();
    panel2_init();
    solve_init();
    mlp_init();
    {
        hipError_t e = hipGetLastError();   // attribute calls must not leave a sticky error
        if (e != hipSuccess) return fail(PALIAL_ERR_HIP, "kernel attribute setup: %s", hipGetErrorString(e));
    }
    g.device = device;
    g.init = true;
    return 0;
}

int palial_hip_finalize(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g.init) return 0;
    (void)hipDeviceSynchronize();
    drop_graphs();
    qr_streams_down();
    for (auto e : g.events) (void)hipEventDestroy(e);
    g.events.clear();
    if (g.d_w1) (void)hipFree(g.d_w1);
    if (g.d_w2) (void)hipFree(g.d_w2);
    if (g.d_qg) (void)hipFree(g.d_qg);
    if (g.d_tau) (void)hipFree(g.d_tau);
    if (g.d_dinv) (void)hipFree(g.d_dinv);
    if (g.d_sub) (void)hipFree(g.d_sub);
    if (g.d_pack) (void)hipFree(g.d_pack);
    if (g.d_F) (void)hipFree(g.d_F);
    if (g.d_inv) (void)hipFree(g.d_inv);
    if (g.d_blk) (void)hipFree(g.d_blk);
    if (g.d_F32) (void)hipFree(g.d_F32);
    if (g.d_gm) (void)hipFree(g.d_gm);
    if (g.d_bidx) (void)hipFree(g.d_bidx);
    if (g.d_bwork) (void)hipFree(g.d_bwork);
    if (g.d_vec) (void)hipFree(g.d_vec);
    if (g.d_aw) (void)hipFree(g.d_aw);
    for (auto &x : g_tr.singles) {
        (void)hipEventDestroy(x.a);
        (void)hipEventDestroy(x.e);
    }
    if (g_tr.t0) (void)hipEventDestroy(g_tr.t0);
    if (g_tr.d_ts) (void)hipFree(g_tr.d_ts);
    g_tr = TaskTrace();
    if (g.d_qs) (void)hipFree(g.d_qs);
    if (g.d_info) (void)hipFree(g.d_info);
    if (g.s_bulk && g.s_bulk != g.s_main) (void)hipStreamDestroy(g.s_bulk);
    if (g.s_aux) (void)hipStreamDestroy(g.s_aux);
    if (g.s_chainm) (void)hipStreamDestroy(g.s_chainm);
    (void)hipStreamDestroy(g.s_main);
    (void)hipStreamDestroy(g.s_panel);
    g = Ctx();
    return 0;
}

int palial_hip_sweep_plan(int algo, int MT, int q, int *out, int cap)
{
    if (algo < 0 || algo > 1 || q < 1 || MT < 1 || MT % q) return PALIAL_ERR_ARG;
    const int n = sweep_plan(algo, MT, q, nullptr, 0);
    if (!out || cap < n) return PALIAL_ERR_ARG;
    std::vector<SweepStep> plan((size_t)n);
    sweep_plan(algo, MT, q, plan.data(), n);
    for (int i = 0; i < n; i++) {
        const SweepStep &st = plan[(size_t)i];
        const int v[7] = {st.mode, st.k, st.trans, st.ntasks, st.inv, st.vin, st.vout};
        for (int j = 0; j < 7; j++) out[7 * i + j] = v[j];
    }
    return n;
}

int palial_hip_gemv_tiles(int mode, int trans, int MT, int b, int q, int k, int ntasks, long *out, long cap)
{
    if (mode < GV_DIAG || mode > GV_BWD || q < 1 || b < 1 || MT < 1 || ntasks < 0 || !out ||
        cap < (long)ntasks * (3 + q))
        return PALIAL_ERR_ARG;
    for (int t = 0; t < ntasks; t++) {
        const GemvMap m = gemv_map(mode, trans, k, q, t);
        long *o = out + (long)t * (3 + q);
        o[0] = m.vout; o[1] = m.j0; o[2] = m.j1;
        for (int j = 0; j < q; j++)
            o[3 + j] = (j >= m.j0 && j < m.j1) ? (long)gemv_src(mode, trans, k, q, b, MT, m.vout, j) : -1;
    }
    return ntasks;
}

int palial_hip_gather_map(int MT, int b, int q, int upper, int dtrans, long *out, long cap)
{
    if (q < 2 || MT % q || b < 1 || !out || cap < 4L * MT * q) return PALIAL_ERR_ARG;
    const int KT = MT / q;
    long n = 0;
    for (int K = 0; K < KT; K++)
        for (int i = 0; i < q; i++)
            for (int j = 0; j < q; j++, n++) {
                const GatherTile g2 = gather_tile(MT, b, q, upper != 0, dtrans != 0, K, i, j);
                out[4 * n] = g2.kind; out[4 * n + 1] = g2.tr ? 1 : 0;
                out[4 * n + 2] = (long)g2.src; out[4 * n + 3] = (long)g2.dst;
            }
    return (int)n;
}

int palial_hip_blkinv_plan(int q, double *out, int cap)
{
    if (q < 1) return PALIAL_ERR_ARG;
    const int n = blkinv_plan(q, nullptr, 0);
    if (n > 0 && (!out || cap < n)) return PALIAL_ERR_ARG;
    std::vector<BlkInvLaunch> plan((size_t)n);
    blkinv_plan(q, plan.data(), n);
    for (int i = 0; i < n; i++) {
        const BlkInvLaunch &l = plan[(size_t)i];
        out[5 * i] = l.mode; out[5 * i + 1] = l.bd; out[5 * i + 2] = l.bt;
        out[5 * i + 3] = l.alpha; out[5 * i + 4] = l.beta;
    }
    return n;
}

int palial_hip_blkinv_tiles(int mode, int MT, int b, int q, int bd, int bt, long *out, long cap)
{
    if ((mode != GM_BI_W && mode != GM_BI_X) || q < 2 || MT % q || bd < 1 || bd >= q || bt < 0 || bt >= bd || b < 1)
        return PALIAL_ERR_ARG;
    const int ntasks = MT / q * (q - bd);
    if (!out || cap < 8L * ntasks) return PALIAL_ERR_ARG;
    GemmArgs a;
    memset(&a, 0, sizeof a);
    a.mode = mode; a.b = b; a.MT = MT; a.N = MT * b; a.bq = q; a.bd = bd; a.bt = bt; a.ntasks = ntasks;
    // three fake bases far apart, so that each pointer maps back to one buffer
    const size_t span = (size_t)1 << 40;
    double *base[3] = {(double *)(span * 8), (double *)(span * 8 * 3), (double *)(span * 8 * 5)};
    a.mat = base[0]; a.w1 = base[1]; a.w2 = base[2];
    auto rel = [&](const double *p, long *o) {
        o[0] = -1; o[1] = -1;
        if (!p) return;
        for (int bi = 0; bi < 3; bi++) {
            const long d = (long)(p - base[bi]);
            if (d >= 0 && d < (long)span) { o[0] = bi; o[1] = d; return; }
        }
        o[0] = -2;
    };
    for (int t = 0; t < ntasks; t++) {
        const TaskPtrs tp = decode(a, t);
        long *o = out + 8L * t;
        rel(tp.A, o); rel(tp.B, o + 2); rel(tp.C, o + 4); rel(tp.D, o + 6);
    }
    return ntasks;
}

int palial_hip_schedule_ops(int algo, int MT, int b, int prec, int variant, long *out, long cap)
{
    // the chain schedules of the solves: FP64 (prec 0) on the FP64 copy, the
    // mixed tiers on the FP32 working copy with packed panels (b % 128 == 0);
    // QR (algo 2, FP64) with `variant` panel chains in flight
    // Cholesky / LU: variant = pair + 2 * schedule (0 diagonal chain with the split
    // update, 1 chain with one update launch, 2 row look-ahead, 3 one stream;
    // the mixed tiers run the split chain only)
    const int pair = algo < 2 ? variant & 1 : 0, sched = algo < 2 ? variant >> 1 : 0;
    if (algo < 0 || algo > 2 || MT < 2 || MT > 64 || prec < 0 || prec > 4 || !panel2_ok(b) ||
        (prec > 0 && b % 128) || (algo < 2 && (variant < 0 || sched > 3 || (prec > 0 && sched != 0))) ||
        (algo == 2 && (prec != 0 || variant < 1 || variant > QR_MAXS)))
        return PALIAL_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    Ctx live;
    std::swap(g, live);                  // the schedule runs on fake handles and addresses only
    const bool prof = g_prof.on, tr = g_tr.active;
    g_prof.on = false;
    g_tr.active = false;
    const int N = MT * b;
    const bool fs = prec > 0;
    const size_t psb = pack_set_bytes(b, MT);
    char *const fake = (char *)((size_t)1 << 44);          // never dereferenced
    g.dag = 2;
    g.subtile = 1;
    g.f32work = 1;
    g.pair = pair;
    g.d_pack = fake + ((size_t)1 << 40);
    g.pack_cap = PACK_SETS * psb + 2 * (size_t)b * b * 4;
    g.d_dinv = (double *)(fake + ((size_t)2 << 40));
    g.dinv_cap = (size_t)MT * 64 * b;
    g.d_sub = (double *)(fake + ((size_t)3 << 40));
    g.sub_cap = (size_t)b * b;
    g.events.resize(events_needed(MT));
    for (size_t i = 0; i < g.events.size(); i++) g.events[i] = (hipEvent_t)(intptr_t)(1000 + i);
    g_rec = SchedRec();
    g_rec.on = true;
    g_rec.mat = fake;
    g_rec.es = fs ? 4 : 8;
    g_rec.b = b;
    g_rec.MT = MT;
    g_rec.pack = (const char *)g.d_pack;
    g_rec.psb = psb;
    g_rec.dinv = (const char *)g.d_dinv;
    hipStream_t s0 = (hipStream_t)(intptr_t)1, sc = (hipStream_t)(intptr_t)2, sb = (hipStream_t)(intptr_t)3,
                sa = (hipStream_t)(intptr_t)4;
    Launcher L;
    double *dA = (double *)(void *)fake;
    if (algo == 2) {
        // QR buffers: T factors, the W / op(T) W tiles of every update stream, tau and
        // Gram workspace of every panel chain, each a region of its own
        const size_t bb8 = (size_t)b * b * 8;
        double *dS = (double *)(fake + ((size_t)4 << 40));
        g.d_w1 = (double *)(fake + ((size_t)5 << 40));
        g.d_w2 = (double *)(fake + ((size_t)6 << 40));
        g.d_tau = (double *)(fake + ((size_t)7 << 40));
        g.d_qg = (double *)(fake + ((size_t)8 << 40));
        g_rec.regions = {{(const char *)dS, (size_t)MT * MT * bb8, bb8, 3 * REC_RES},
                         {(const char *)g.d_w1, (size_t)QR_MAXS * MT * bb8, bb8, 4 * REC_RES},
                         {(const char *)g.d_w2, (size_t)QR_MAXS * MT * bb8, bb8, 5 * REC_RES},
                         {(const char *)g.d_tau, (size_t)QR_MAXS * b * 8, (size_t)b * 8, 6 * REC_RES},
                         {(const char *)g.d_qg, (size_t)QR_MAXS * bb8, bb8, 7 * REC_RES}};
        for (int i = 0; i < QR_MAXS; i++) {
            g.s_qr[i] = (hipStream_t)(intptr_t)(10 + i);
            g.s_qrb[i] = (hipStream_t)(intptr_t)(20 + i);
        }
        g.qr_streams = QR_MAXS;
        dag_qr(dA, dS, N, b, s0, sc, true, L, variant);
    } else {
        // the streams the schedules take from the context: bulk and aux (CU-masked), no chain mask
        g.s_bulk = sb;
        g.s_aux = sa;
        g.s_chainm = nullptr;
        g.dag = sched == 0 ? 2 : sched == 1 ? 1 : 0;
        const bool lookahead = sched != 3;
        if (algo == 0) dag_cholesky(dA, N, b, lookahead, s0, sc, (int *)(fake + 8), L, prec, fs);
        else dag_lu(dA, N, b, lookahead, s0, sc, L, prec, fs);
    }
    std::vector<long> rec;
    rec.swap(g_rec.out);
    g_rec = SchedRec();
    std::swap(g, live);
    g_prof.on = prof;
    g_tr.active = tr;
    if (L.err != hipSuccess) return fail(PALIAL_ERR_ARG, "schedule refused its arguments");
    if (out && cap >= (long)rec.size()) std::copy(rec.begin(), rec.end(), out);
    return (int)rec.size();
}

int palial_hip_cu_policy(const char *policy, int ncu, unsigned *out, int words)
{
    if (ncu <= 0 || words < (ncu + 31) / 32 || !out) return PALIAL_ERR_ARG;
    std::vector<uint32_t> m[3];
    bool given[3];
    if (parse_cu_policy(policy, ncu, m, given)) return PALIAL_ERR_ARG;
    for (int c = 0; c < 3; c++)
        for (int w = 0; w < words; w++) out[c * words + w] = (c < 2 || given[2]) && w < (int)m[c].size() ? m[c][(size_t)w] : 0u;
    return given[2] ? 1 : 0;
}

int palial_hip_set_option(const char *name, int value)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init())) return rc;
    if (!name) return fail(PALIAL_ERR_ARG, "option name is NULL");
    int *slot = nullptr;
    int lo = 0, hi = 0;
    if (!strcmp(name, "pair")) { slot = &g.pair; lo = 0; hi = 1; }
    else if (!strcmp(name, "sweep_block")) { slot = &g.sweep_block; lo = 1; hi = 8; }
    if (!slot) return fail(PALIAL_ERR_ARG, "unknown option %s", name);
    if (value < lo || value > hi) return fail(PALIAL_ERR_ARG, "option %s=%d outside %d..%d", name, value, lo, hi);
    const int prev = *slot;
    *slot = value;
    return prev;
}

int palial_hip_trace(int enable)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init())) return rc;
    g_tr.on = enable != 0;
    g_tr.records.clear();
    g_tr.algo = -1;
    return 0;
}

long palial_hip_trace_records(double *out, long cap)
{
    std::lock_guard<std::mutex> lk(g_mu);
    const long n = (long)g_tr.records.size();
    for (long i = 0; i < n && i < cap && out; i++)
        for (int j = 0; j < 6; j++) out[6 * i + j] = g_tr.records[i][j];
    return g_tr.overflow ? -n - 1 : n;
}

long palial_hip_task_name(int algo, int MT, long index, char *buf, int len)
{
    if (algo < 0 || algo > 2 || MT < 1 || MT > 4096) return fail(PALIAL_ERR_ARG, "task list of algo %d, MT %d", algo, MT);
    std::vector<RefTask> tasks;
    ref_tasks(algo, MT, tasks);
    if (buf && len > 0) {
        if (index < 0 || index >= (long)tasks.size()) return fail(PALIAL_ERR_ARG, "task %ld of %zu", index, tasks.size());
        snprintf(buf, (size_t)len, "%s", tasks[(size_t)index].name);
    }
    return (long)tasks.size();
}

int palial_hip_profile(int enable)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_prof.on = enable != 0;
    g_prof.keep = enable == 2;
    g_prof.records.clear();
    for (int c = 0; c < KC_N; c++) { g_prof.launches[c] = 0; g_prof.ms[c] = g_prof.flops[c] = g_prof.bytes[c] = 0; }
    return 0;
}

long palial_hip_profile_records(double *out, long cap)
{
    std::lock_guard<std::mutex> lk(g_mu);
    const long n = (long)g_prof.records.size();
    if (out)
        for (long i = 0; i < n && i < cap; i++)
            for (int f = 0; f < 6; f++) out[i * 6 + f] = g_prof.records[i][f];
    return n;
}

int palial_hip_profile_stats(int cls, long *launches, double *total_ms, double *flops, double *bytes)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (cls < 0 || cls >= KC_N) return fail(PALIAL_ERR_ARG, "kernel class %d out of range", cls);
    if (launches) *launches = g_prof.launches[cls];
    if (total_ms) *total_ms = g_prof.ms[cls];
    if (flops) *flops = g_prof.flops[cls];
    if (bytes) *bytes = g_prof.bytes[cls];
    return 0;
}

int palial_hip_time_update(int algo, MATRIX_desc A, int k, int reps, double *ms)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init()) || (rc = check_desc(A))) return rc;
    const int N = A.matrix_size, b = A.tile_size, MT = N / b, T = MT - k - 1;
    if (k < 0 || T <= 0 || reps <= 0) return fail(PALIAL_ERR_ARG, "bad step %d or reps %d", k, reps);
    GemmArgs a = gargs(algo == ALG_CHOL ? GM_CHOL_TRAIL : GM_LU_TRAIL, (double *)A.matrix, N, b, k, T);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(launch_gemm(a, algo == ALG_CHOL, false, g.s_main));   // warm
    HIPCHK(hipEventRecord(e0, g.s_main));
    for (int r = 0; r < reps; r++) HIPCHK(launch_gemm(a, algo == ALG_CHOL, false, g.s_main));
    HIPCHK(hipEventRecord(e1, g.s_main));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t / reps;
    return 0;
}

int palial_hip_update(int algo, MATRIX_desc A, int k, int prec, int reps, double *ms)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init()) || (rc = check_desc(A))) return rc;
    const int N = A.matrix_size, b = A.tile_size, MT = N / b, T = MT - k - 1;
    if (algo != ALG_CHOL && algo != ALG_LU) return fail(PALIAL_ERR_ARG, "algo %d has no trailing update", algo);
    if (k < 0 || T <= 0 || reps <= 0) return fail(PALIAL_ERR_ARG, "bad step %d or reps %d", k, reps);
    const bool w32 = (prec & 8) != 0;    // FP32 tiles (the first N*N*4 bytes of the buffer)
    prec &= 7;
    if (prec < 0 || prec > 4 || (w32 && prec == 0))
        return fail(PALIAL_ERR_ARG, "precision %d not available", prec);
    if (prec > 0 && (b % 128 != 0)) return fail(PALIAL_ERR_ARG, "packed updates need b %% 128 == 0 (b=%d)", b);
    if (prec > 0 && (rc = ensure_pack(b, MT))) return rc;
    double *dA = (double *)A.matrix;
    GemmArgs a = gargs(algo == ALG_CHOL ? GM_CHOL_TRAIL : GM_LU_TRAIL, dA, N, b, k, T);
    a.f32c = w32 ? 1 : 0;
    char *pU = (char *)g.d_pack, *pL = pU + (size_t)MT * b * b * 4;
    a.pa = algo == ALG_CHOL ? (const void *)pU : (const void *)pL;
    a.pb = pU;
    hipStream_t s = g.s_main;
    auto once = [&]() -> hipError_t {
        if (prec == 0) return launch_gemm(a, algo == ALG_CHOL, false, s);
        hipError_t e = launch_pack(dA, N, b, k, k + 1, T, false, prec, pU, s, w32);
        if (e == hipSuccess && algo == ALG_LU) e = launch_pack(dA, N, b, k, k + 1, T, true, prec, pL, s, w32);
        if (e == hipSuccess) e = launch_gemm_packed(a, prec, s);
        return e;
    };
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++) HIPCHK(once());
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t / reps;
    return 0;
}

// ------------------------------------------------------------------ batched
namespace {

int grow_bytes(void **p, size_t *cap, size_t need)
{
    if (*cap >= need) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc(p, need) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "batch workspace of %zu bytes", need);
    *cap = need;
    return 0;
}

struct BatchPlan {
    std::vector<int> order;          // all matrices, largest first
    std::vector<long> woff;
    size_t work = 0, a_ext = 0, b_ext = 0;
    int bmax = 32;
};

int batch_plan(int count, const int *n, const long *off, const int *kind, const long *boff, BatchPlan &P)
{
    if (count <= 0) return fail(PALIAL_ERR_ARG, "batch count %d", count);
    if (!n || !off || !kind) return fail(PALIAL_ERR_ARG, "NULL batch description array");
    P.woff.resize(count);
    for (int i = 0; i < count; i++) {
        if (n[i] < 1 || n[i] > 512) return fail(PALIAL_ERR_ARG, "matrix %d: order %d outside 1..512", i, n[i]);
        if (kind[i] != 0 && kind[i] != 1) return fail(PALIAL_ERR_ARG, "matrix %d: kind %d", i, kind[i]);
        if (off[i] < 0 || (boff && boff[i] < 0)) return fail(PALIAL_ERR_ARG, "matrix %d: negative offset", i);
        const int b = (n[i] + 31) / 32 * 32;
        P.woff[i] = (long)P.work;
        P.work += (size_t)b * b;
        P.a_ext = std::max(P.a_ext, (size_t)off[i] + (size_t)n[i] * n[i]);
        if (boff) P.b_ext = std::max(P.b_ext, (size_t)boff[i] + (size_t)n[i]);
        P.order.push_back(i);
        P.bmax = std::max(P.bmax, b);
    }
    auto larger = [&](int a, int b) { return n[a] > n[b]; };
    std::stable_sort(P.order.begin(), P.order.end(), larger);
    return 0;
}

}  // namespace

int palial_hip_batched(int count, const int *n, const long *off, const int *kind, void *A, void *B,
                       const long *boff, int *info, int prec, const palial_opts *o_in, palial_report *rep)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init())) return rc;
    palial_opts o = o_in ? *o_in : default_opts();
    if (prec != PALIAL_PREC_FP64 && prec != PALIAL_PREC_FP32) return fail(PALIAL_ERR_ARG, "batched precision %d", prec);
    if (!A) return fail(PALIAL_ERR_ARG, "NULL matrix storage");
    if (B && !boff) return fail(PALIAL_ERR_ARG, "right-hand sides need boff");
    BatchPlan P;
    if ((rc = batch_plan(count, n, off, kind, B ? boff : nullptr, P))) return rc;
    const size_t es = prec == PALIAL_PREC_FP32 ? 4 : 8;
    // device index arrays: off | woff | boff | n | order | kind | info | queue
    const size_t ibytes = (size_t)count * (8 + 8 + 8 + 4 + 4 + 4 + 4) + 64;
    if ((rc = grow_bytes(&g.d_bidx, &g.bidx_cap, ibytes))) return rc;
    if ((rc = grow(&g.d_bwork, &g.bwork_cap, P.work))) return rc;
    char *base = (char *)g.d_bidx;
    long *d_off = (long *)base, *d_woff = d_off + count, *d_boff = d_woff + count;
    int *d_n = (int *)(d_boff + count), *d_ord = d_n + count, *d_kind = d_ord + count, *d_info = d_kind + count;
    int *d_queue = d_info + count;
    hipStream_t s = g.s_main;
    std::vector<long> hb(count, 0);
    if (B) for (int i = 0; i < count; i++) hb[i] = boff[i];
    HIPCHK(hipMemcpyAsync(d_off, off, (size_t)count * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_woff, P.woff.data(), (size_t)count * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_boff, hb.data(), (size_t)count * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_n, n, (size_t)count * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_ord, P.order.data(), (size_t)count * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_kind, kind, (size_t)count * 4, hipMemcpyHostToDevice, s));
    void *dA = A, *dB = B;
    bool own = false;
    if (!o.device_matrix) {
        own = true;
        dA = dB = nullptr;
        if (hipMalloc(&dA, P.a_ext * es) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "batch matrices");
        if (B && hipMalloc(&dB, P.b_ext * es) != hipSuccess) { (void)hipFree(dA); return fail(PALIAL_ERR_NOMEM, "batch vectors"); }
        HIPCHK(hipMemcpyAsync(dA, A, P.a_ext * es, hipMemcpyHostToDevice, s));
        if (B) HIPCHK(hipMemcpyAsync(dB, B, P.b_ext * es, hipMemcpyHostToDevice, s));
    }
    if ((rc = ensure_events(4))) return rc;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    BatchArgs a;
    memset(&a, 0, sizeof a);
    a.n = d_n; a.off = d_off; a.woff = d_woff; a.boff = d_boff; a.A = dA; a.B = dB;
    a.work = g.d_bwork; a.info = d_info; a.f32 = (prec == PALIAL_PREC_FP32);
    a.count = count; a.order = d_ord; a.kind = d_kind; a.queue = d_queue;
    if (o.batch_grid < 0) return fail(PALIAL_ERR_ARG, "batch_grid %d", o.batch_grid);
    // persistent workgroups: what the CUs hold at once, or the caller's count
    const int grid = o.batch_grid > 0 ? std::min(o.batch_grid, count) : batch_grid(P.bmax, count);
    if (grid < 1) return fail(PALIAL_ERR_HIP, "batched launch shape");
    // one persistent launch for both kinds, captured once into a hipGraph
    // (per argument set) and replayed; the queue counter is zeroed before it,
    // outside the graph
    HIPCHK(hipMemsetAsync(d_queue, 0, sizeof(int), s));
    HIPCHK(hipEventRecord(e0, s));
    const bool use_graph = !o.no_graph && !g_prof.on;
    if (use_graph) {
        BatchGraph &bg = g.bgraph;
        const bool same = bg.exec && bg.count == count && bg.A == dA && bg.B == dB && bg.idx == g.d_bidx &&
                          bg.work == g.d_bwork && bg.f32 == a.f32 && bg.bmax == P.bmax && bg.grid == grid;
        if (!same) {
            if (bg.exec) (void)hipGraphExecDestroy(bg.exec);
            if (bg.graph) (void)hipGraphDestroy(bg.graph);
            bg = BatchGraph();
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const hipError_t le = launch_batch(a, P.bmax, grid, s);
            hipGraph_t gr = nullptr;
            const hipError_t ce = hipStreamEndCapture(s, &gr);
            if (le != hipSuccess || ce != hipSuccess || !gr)
                return fail(PALIAL_ERR_HIP, "batched capture: %s / %s", hipGetErrorString(le), hipGetErrorString(ce));
            bg.graph = gr;
            HIPCHK(hipGraphInstantiate(&bg.exec, gr, nullptr, nullptr, 0));
            bg.count = count; bg.A = dA; bg.B = dB; bg.idx = g.d_bidx; bg.work = g.d_bwork;
            bg.f32 = a.f32; bg.bmax = P.bmax; bg.grid = grid;
        }
        HIPCHK(hipGraphLaunch(bg.exec, s));
    } else {
        HIPCHK(launch_batch(a, P.bmax, grid, s));
    }
    HIPCHK(hipEventRecord(e1, s));
    if (info) HIPCHK(hipMemcpyAsync(info, d_info, (size_t)count * 4, hipMemcpyDeviceToHost, s));
    if (own) {
        HIPCHK(hipMemcpyAsync(A, dA, P.a_ext * es, hipMemcpyDeviceToHost, s));
        if (B) HIPCHK(hipMemcpyAsync(B, dB, P.b_ext * es, hipMemcpyDeviceToHost, s));
    }
    const hipError_t e = hipStreamSynchronize(s);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (own) {
        (void)hipFree(dA);
        if (dB) (void)hipFree(dB);
    }
    if (e != hipSuccess) return fail(PALIAL_ERR_HIP, "batched factorization: %s", hipGetErrorString(e));
    if (rep) {
        double fl = 0;
        for (int i = 0; i < count; i++) fl += (kind[i] == 0 ? 1.0 : 2.0) * n[i] * (double)n[i] * n[i] / 3.0;
        rep->factor_ms = ms;
        rep->gflop = fl / 1e9;
        rep->tflops = ms > 0 ? fl / (ms * 1e-3) / 1e12 : 0.0;
        rep->ntasks = count;
        rep->nlaunches = 1;
    }
    return 0;
}

int palial_hip_batched_generate(int count, const int *n, const long *off, const int *kind, void *A, uint64_t seed,
                                int prec, const palial_opts *o_in)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init())) return rc;
    palial_opts o = o_in ? *o_in : default_opts();
    if (prec != PALIAL_PREC_FP64 && prec != PALIAL_PREC_FP32) return fail(PALIAL_ERR_ARG, "batched precision %d", prec);
    if (!A) return fail(PALIAL_ERR_ARG, "NULL matrix storage");
    BatchPlan P;
    if ((rc = batch_plan(count, n, off, kind, nullptr, P))) return rc;
    const size_t es = prec == PALIAL_PREC_FP32 ? 4 : 8;
    const size_t ibytes = (size_t)count * (4 + 8 + 4) + 64;
    if ((rc = grow_bytes(&g.d_bidx, &g.bidx_cap, std::max(ibytes, (size_t)count * 36 + 64)))) return rc;
    long *d_off = (long *)g.d_bidx;
    int *d_n = (int *)(d_off + count), *d_k = d_n + count;
    hipStream_t s = g.s_main;
    HIPCHK(hipMemcpyAsync(d_off, off, (size_t)count * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_n, n, (size_t)count * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_k, kind, (size_t)count * 4, hipMemcpyHostToDevice, s));
    void *dA = A;
    if (!o.device_matrix && hipMalloc(&dA, P.a_ext * es) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "batch matrices");
    HIPCHK(launch_batch_generate(d_n, d_off, d_k, dA, count, prec == PALIAL_PREC_FP32, seed, s));
    if (!o.device_matrix) HIPCHK(hipMemcpyAsync(A, dA, P.a_ext * es, hipMemcpyDeviceToHost, s));
    const hipError_t e = hipStreamSynchronize(s);
    if (!o.device_matrix) (void)hipFree(dA);
    if (e != hipSuccess) return fail(PALIAL_ERR_HIP, "batch generate: %s", hipGetErrorString(e));
    return 0;
}

// ------------------------------------------------------------------ cost model
int palial_hip_mlp_nparams(int nfeat) { return (nfeat >= 1 && nfeat <= 16) ? mlp_nparams(nfeat) : -1; }

int palial_hip_mlp_forward(int n, int nfeat, const float *X, const float *params, float *y, const palial_opts *o_in)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init())) return rc;
    if (n < 0 || nfeat < 1 || nfeat > 16 || !X || !params || !y) return fail(PALIAL_ERR_ARG, "bad cost-model arguments");
    palial_opts o = o_in ? *o_in : default_opts();
    const int np = mlp_nparams(nfeat);
    const float *dX = X, *dP = params;
    float *dY = y, *buf = nullptr;
    hipStream_t s = g.s_main;
    if (!o.device_matrix) {
        if (hipMalloc(&buf, ((size_t)n * nfeat + np + n) * 4 + 64) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "cost-model buffers");
        float *bx = buf, *bp = bx + (size_t)n * nfeat, *by = bp + np;
        HIPCHK(hipMemcpyAsync(bx, X, (size_t)n * nfeat * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(bp, params, (size_t)np * 4, hipMemcpyHostToDevice, s));
        dX = bx; dP = bp; dY = by;
    }
    HIPCHK(launch_mlp_forward(n, nfeat, dX, dP, dY, s));
    if (!o.device_matrix) HIPCHK(hipMemcpyAsync(y, dY, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    const hipError_t e = hipStreamSynchronize(s);
    if (buf) (void)hipFree(buf);
    if (e != hipSuccess) return fail(PALIAL_ERR_HIP, "cost-model forward: %s", hipGetErrorString(e));
    return 0;
}

int palial_hip_mlp_train(int n, int nfeat, const float *X, const float *y, float *params, int epochs, float lr,
                         float *loss, float *grad, const palial_opts *o_in)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init())) return rc;
    if (n < 1 || nfeat < 1 || nfeat > 16 || epochs < 0 || !X || !y || !params || (epochs == 0 && !grad))
        return fail(PALIAL_ERR_ARG, "bad cost-model arguments");
    palial_opts o = o_in ? *o_in : default_opts();
    const int np = mlp_nparams(nfeat);
    const int nl = epochs > 0 ? epochs : 1;
    // device layout: X | y | params | adam m | adam v | loss | grad
    float *buf = nullptr;
    const size_t tot = (size_t)n * nfeat + n + 4 * (size_t)np + nl + 64;
    if (hipMalloc(&buf, tot * 4) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "cost-model buffers");
    float *bx = buf, *by = bx + (size_t)n * nfeat, *bp = by + n, *bm = bp + np, *bv = bm + np, *bl = bv + np, *bg = bl + nl;
    hipStream_t s = g.s_main;
    const hipMemcpyKind in = o.device_matrix ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind out = o.device_matrix ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    auto done = [&](int code) { (void)hipStreamSynchronize(s); (void)hipFree(buf); return code; };
    if (hipMemcpyAsync(bx, X, (size_t)n * nfeat * 4, in, s) != hipSuccess ||
        hipMemcpyAsync(by, y, (size_t)n * 4, in, s) != hipSuccess ||
        hipMemcpyAsync(bp, params, (size_t)np * 4, in, s) != hipSuccess ||
        hipMemsetAsync(bm, 0, 2 * (size_t)np * 4, s) != hipSuccess)
        return done(fail(PALIAL_ERR_HIP, "cost-model copies"));
    if (launch_mlp_train(n, nfeat, bx, by, bp, bm, bv, epochs, lr, bl, bg, s) != hipSuccess)
        return done(fail(PALIAL_ERR_HIP, "cost-model training launch"));
    hipError_t e = hipSuccess;
    if (epochs > 0) e = hipMemcpyAsync(params, bp, (size_t)np * 4, out, s);
    if (grad && e == hipSuccess) e = hipMemcpyAsync(grad, bg, (size_t)np * 4, out, s);
    if (loss && e == hipSuccess) e = hipMemcpyAsync(loss, bl, (size_t)nl * 4, out, s);
    const hipError_t es = hipStreamSynchronize(s);
    if (e == hipSuccess) e = es;
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(PALIAL_ERR_HIP, "cost-model training: %s", hipGetErrorString(e));
    return 0;
}

int palial_hip_malloc(void **p, size_t bytes)
{
    if (!p) return fail(PALIAL_ERR_ARG, "NULL out pointer");
    if (hipMalloc(p, bytes) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
    return 0;
}
int palial_hip_free(void *p)
{
    std::lock_guard<std::mutex> lk(g_mu);
    // graphs captured on this pointer are no longer valid
    for (auto it = g.graphs.begin(); it != g.graphs.end();) {
        if (std::get<1>(it->first) == p || std::get<2>(it->first) == p) {
            if (it->second.exec) (void)hipGraphExecDestroy(it->second.exec);
            if (it->second.graph) (void)hipGraphDestroy(it->second.graph);
            it = g.graphs.erase(it);
        } else {
            ++it;
        }
    }
    HIPCHK(hipFree(p));
    return 0;
}
int palial_hip_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}
int palial_hip_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return 0;
}
int palial_hip_copy(void *dst, const void *src, size_t bytes)
{
    int rc;
    if ((rc = ensure_init())) return rc;
    if (bytes % 16) return fail(PALIAL_ERR_ARG, "copy size must be a multiple of 16 bytes");
    HIPCHK(launch_copy(dst, src, bytes, g.s_main));
    HIPCHK(hipStreamSynchronize(g.s_main));
    return 0;
}
int palial_hip_synchronize(void)
{
    HIPCHK(hipDeviceSynchronize());
    return 0;
}
int palial_hip_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}

int palial_hip_generate(MATRIX_desc A, int kind, uint64_t seed, double shift, const palial_opts *o_in)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init())) return rc;
    if ((rc = check_desc(A))) return rc;
    if (kind < 0 || kind > 5) return fail(PALIAL_ERR_ARG, "unknown generator kind %d", kind);
    if (kind == 5 && !(shift >= 1.0)) return fail(PALIAL_ERR_ARG, "geometric spectrum needs kappa >= 1 (got %g)", shift);
    palial_opts o = o_in ? *o_in : default_opts();
    const int N = A.matrix_size, b = A.tile_size;
    const size_t bytes = (size_t)N * N * 8;
    hipStream_t s = (o.device_matrix && o.stream) ? (hipStream_t)o.stream : g.s_main;
    double *d = (double *)A.matrix;
    if (!o.device_matrix) {
        if (hipMalloc(&d, bytes) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "device alloc");
    }
    hipError_t e = launch_generate(d, N, b, kind, seed, shift, s);
    if (e == hipSuccess && !o.device_matrix) e = hipMemcpyAsync(A.matrix, d, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (!o.device_matrix) (void)hipFree(d);
    if (e != hipSuccess) return fail(PALIAL_ERR_HIP, "generate: %s", hipGetErrorString(e));
    return 0;
}

int palial_hip_generate_kron(MATRIX_desc A, const double *A1, int n1, const double *A2, int n2,
                             const palial_opts *o_in)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init()) || (rc = check_desc(A))) return rc;
    if (!A1 || !A2 || n1 <= 0 || n2 <= 0 || (long)n1 * n2 != A.matrix_size)
        return fail(PALIAL_ERR_ARG, "n1*n2 must equal N");
    palial_opts o = o_in ? *o_in : default_opts();
    const int N = A.matrix_size, b = A.tile_size;
    const size_t bytes = (size_t)N * N * 8;
    double *d = (double *)A.matrix, *d1 = nullptr, *d2 = nullptr;
    if (hipMalloc(&d1, (size_t)n1 * n1 * 8) != hipSuccess || hipMalloc(&d2, (size_t)n2 * n2 * 8) != hipSuccess)
        return fail(PALIAL_ERR_NOMEM, "kron factors");
    if (!o.device_matrix && hipMalloc(&d, bytes) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "device alloc");
    hipError_t e = hipMemcpy(d1, A1, (size_t)n1 * n1 * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d2, A2, (size_t)n2 * n2 * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_kron(d, N, b, d1, n1, d2, n2, g.s_main);
    if (e == hipSuccess && !o.device_matrix) e = hipMemcpyAsync(A.matrix, d, bytes, hipMemcpyDeviceToHost, g.s_main);
    if (e == hipSuccess) e = hipStreamSynchronize(g.s_main);
    (void)hipFree(d1);
    (void)hipFree(d2);
    if (!o.device_matrix) (void)hipFree(d);
    if (e != hipSuccess) return fail(PALIAL_ERR_HIP, "generate_kron: %s", hipGetErrorString(e));
    return 0;
}

int palial_hip_cholesky(MATRIX_desc A, const palial_opts *opts, palial_report *rep)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return run_algo(ALG_CHOL, A, nullptr, opts, rep);
}

int palial_hip_lu(MATRIX_desc A, int *ipiv, const palial_opts *opts, palial_report *rep)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = run_algo(ALG_LU, A, nullptr, opts, rep);
    if (rc >= 0 && ipiv) {
        // no row interchanges (src/srclu/dgetrf.h:20): identity, 1-based
        const int N = A.matrix_size;
        if (opts && opts->device_matrix) {
            hipError_t e = launch_fill_identity_ipiv(ipiv, N, g.s_main);
            if (e == hipSuccess) e = hipStreamSynchronize(g.s_main);
            if (e != hipSuccess) return fail(PALIAL_ERR_HIP, "ipiv: %s", hipGetErrorString(e));
        } else {
            for (int i = 0; i < N; i++) ipiv[i] = i + 1;
        }
    }
    return rc;
}

int palial_hip_solve(int algo, MATRIX_desc A, double *B, int nrhs, int ldb, const palial_opts *opts,
                     palial_report *rep)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (algo != 0 && algo != 1) return fail(PALIAL_ERR_ARG, "algo must be 0 (Cholesky) or 1 (LU)");
    return run_solve(algo, A, B, nrhs, ldb, opts, rep);
}

int palial_hip_qr(MATRIX_desc A, MATRIX_desc S, const palial_opts *opts, palial_report *rep)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return run_algo(ALG_QR, A, &S, opts, rep);
}

// Host tiles of the Q-application entry points are staged through device
// copies; device pointers are used in place.
static int qr_operands(const MATRIX_desc &A, const MATRIX_desc &S, bool dev, std::vector<double *> &owned, double *&dA,
                double *&dS, hipStream_t s)
{
    const size_t bytes = (size_t)A.matrix_size * A.matrix_size * 8;
    dA = (double *)A.matrix;
    dS = (double *)S.matrix;
    if (dev) return 0;
    dA = dS = nullptr;
    if (hipMalloc(&dA, bytes) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "device copy of A");
    owned.push_back(dA);
    if (hipMalloc(&dS, bytes) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "device copy of S");
    owned.push_back(dS);
    HIPCHK(hipMemcpyAsync(dA, A.matrix, bytes, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dS, S.matrix, bytes, hipMemcpyHostToDevice, s));
    return 0;
}

static int check_qr_pair(const MATRIX_desc &A, const MATRIX_desc &S)
{
    int rc;
    if ((rc = check_desc(A)) || (rc = check_desc(S))) return rc;
    if (S.matrix_size != A.matrix_size || S.tile_size != A.tile_size)
        return fail(PALIAL_ERR_ARG, "S must have the same N and b as A");
    return 0;
}

int palial_hip_qr_apply(int side, int trans, MATRIX_desc A, MATRIX_desc S, double *C, int ncols,
                        const palial_opts *o_in, palial_report *rep)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init()) || (rc = check_qr_pair(A, S))) return rc;
    if (side != PALIAL_LEFT) return fail(PALIAL_ERR_ARG, "only side = Left (PALIAL_LEFT) is supported");
    if (trans != PALIAL_TRANS && trans != PALIAL_NOTRANS) return fail(PALIAL_ERR_ARG, "trans %d", trans);
    const int N = A.matrix_size, b = A.tile_size;
    if (!C || ncols <= 0 || ncols % b)
        return fail(PALIAL_ERR_ARG, "C needs a whole number of %d-column tiles (ncols=%d)", b, ncols);
    const int NC = ncols / b;
    palial_opts o = o_in ? *o_in : default_opts();
    const bool dev = o.device_matrix != 0;
    hipStream_t s = (dev && o.stream) ? (hipStream_t)o.stream : g.s_main;
    std::vector<double *> owned;
    auto release = [&](int code) {
        (void)hipStreamSynchronize(s);
        for (double *p : owned) (void)hipFree(p);
        return code;
    };
    double *dA, *dS, *dC = C;
    if ((rc = qr_operands(A, S, dev, owned, dA, dS, s))) return release(rc);
    const size_t cbytes = (size_t)N * ncols * 8;
    if (!dev) {
        if (hipMalloc(&dC, cbytes) != hipSuccess) return release(fail(PALIAL_ERR_NOMEM, "device copy of C"));
        owned.push_back(dC);
        if (hipMemcpyAsync(dC, C, cbytes, hipMemcpyHostToDevice, s) != hipSuccess)
            return release(fail(PALIAL_ERR_HIP, "copy of C"));
    }
    if ((rc = grow(&g.d_aw, &g.aw_cap, 2 * (size_t)NC * b * b))) return release(rc);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    Launcher L;
    hipError_t er = hipEventRecord(e0, s);
    qr_apply_dev(trans == PALIAL_TRANS, dA, dS, dC, N, b, NC, g.d_aw, g.d_aw + (size_t)NC * b * b, s, L);
    if (er == hipSuccess) er = hipEventRecord(e1, s);
    if (er == hipSuccess && !dev) er = hipMemcpyAsync(C, dC, cbytes, hipMemcpyDeviceToHost, s);
    if (er == hipSuccess) er = hipStreamSynchronize(s);
    float ms = 0;
    if (er == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    prof_collect();
    if (L.err != hipSuccess) return release(fail(PALIAL_ERR_HIP, "Q application launch: %s", hipGetErrorString(L.err)));
    if (er != hipSuccess) return release(fail(PALIAL_ERR_HIP, "Q application: %s", hipGetErrorString(er)));
    if (rep) {
        memset(rep, 0, sizeof *rep);
        const double fl = 2.0 * (double)N * N * ncols;   // dormqr with m = k = N, n = ncols: 4mnk - 2nk^2
        rep->solve_ms = ms;
        rep->gflop = fl / 1e9;
        rep->tflops = ms > 0 ? fl / (ms * 1e-3) / 1e12 : 0.0;
        rep->nlaunches = L.launches;
    }
    return release(0);
}

int palial_hip_qr_solve(MATRIX_desc A, MATRIX_desc S, double *B, int nrhs, int ldb, const palial_opts *o_in,
                        palial_report *rep)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init()) || (rc = check_qr_pair(A, S))) return rc;
    const int N = A.matrix_size, b = A.tile_size, MT = N / b;
    if (b % 64 || b > 512) return fail(PALIAL_ERR_ARG, "QR solve needs b %% 64 == 0 and b <= 512 (b=%d)", b);
    if (nrhs < 1 || nrhs > 16) return fail(PALIAL_ERR_ARG, "nrhs=%d outside 1..16", nrhs);
    if (!B || ldb < N) return fail(PALIAL_ERR_ARG, "bad right-hand side (ldb=%d)", ldb);
    palial_opts o = o_in ? *o_in : default_opts();
    const bool dev = o.device_matrix != 0;
    hipStream_t s = (dev && o.stream) ? (hipStream_t)o.stream : g.s_main;
    std::vector<double *> owned;
    auto release = [&](int code) {
        (void)hipStreamSynchronize(s);
        for (double *p : owned) (void)hipFree(p);
        return code;
    };
    double *dA, *dS;
    if ((rc = qr_operands(A, S, dev, owned, dA, dS, s))) return release(rc);
    // workspace: right-hand-side tiles (N x b), Y and X (N x nrhs), the R tile
    // inverses (MT b^2), the 32x32 block inverses (MT 32 b)
    const size_t nv = (size_t)N * nrhs, nt = (size_t)N * b;
    if ((rc = grow(&g.d_qs, &g.qs_cap, nt + 2 * nv + (size_t)MT * b * b + (size_t)MT * 32 * b))) return release(rc);
    if ((rc = grow(&g.d_aw, &g.aw_cap, 2 * (size_t)b * b))) return release(rc);
    double *Ct = g.d_qs, *Y = Ct + nt, *X = Y + nv, *Rinv = X + nv, *dinv = Rinv + (size_t)MT * b * b;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    Launcher L;
    hipError_t er = hipMemcpy2DAsync(Y, (size_t)N * 8, B, (size_t)ldb * 8, (size_t)N * 8, nrhs,
                                     dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s);
    if (er == hipSuccess) er = hipEventRecord(e0, s);
    // Q^T B on one tile column
    L.chk(launch_cols_to_tiles(Y, N, N, nrhs, Ct, b, s));
    qr_apply_dev(true, dA, dS, Ct, N, b, 1, g.d_aw, g.d_aw + (size_t)b * b, s, L);
    L.chk(launch_tiles_to_cols(Ct, N, b, Y, N, nrhs, s));
    // R x = Q^T B by tile back substitution: R lives where LU keeps U (upper
    // triangles of the diagonal tiles and the tiles right of them)
    L.chk(launch_upper_block_inverses(dA, N, b, dinv, 32L * b, s));
    {
        TrsmArgs t;
        memset(&t, 0, sizeof t);
        t.mat = dA; t.single = Rinv; t.N = N; t.b = b; t.MT = MT; t.ntiles = MT;
        t.batch = 1; t.ident = 1; t.dinv_stride = 32L * b;
        L.chk(launch_trsm2(TR_LUR, t, dinv, s));
    }
    GemvArgs a;
    memset(&a, 0, sizeof a);
    a.mat = dA; a.N = N; a.b = b; a.ld = N; a.nrhs = nrhs; a.inv = Rinv;
    for (int k = MT - 1; k >= 0; k--) {
        a.k = k;
        a.mode = GV_DIAG; a.trans = 0; a.vin = Y; a.vout = X; a.ntasks = 1;
        L.chk(launch_tile_gemv(a, s));
        if (k > 0) {
            a.mode = GV_BWD; a.vin = X; a.vout = Y; a.ntasks = k;
            L.chk(launch_tile_gemv(a, s));
        }
    }
    if (er == hipSuccess) er = hipEventRecord(e1, s);
    if (er == hipSuccess)
        er = hipMemcpy2DAsync(B, (size_t)ldb * 8, X, (size_t)N * 8, (size_t)N * 8, nrhs,
                              dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s);
    if (er == hipSuccess) er = hipStreamSynchronize(s);
    float ms = 0;
    if (er == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    prof_collect();
    if (L.err != hipSuccess) return release(fail(PALIAL_ERR_HIP, "QR solve launch: %s", hipGetErrorString(L.err)));
    if (er != hipSuccess) return release(fail(PALIAL_ERR_HIP, "QR solve: %s", hipGetErrorString(er)));
    if (rep) {
        memset(rep, 0, sizeof *rep);
        rep->solve_ms = ms;
        rep->nlaunches = L.launches;
    }
    return release(0);
}

int palial_hip_transpose(MATRIX_desc A, int average, const palial_opts *o_in)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init()) || (rc = check_desc(A))) return rc;
    const int N = A.matrix_size, b = A.tile_size;
    if (b % 32) return fail(PALIAL_ERR_ARG, "transpose needs b %% 32 == 0 (b=%d)", b);
    palial_opts o = o_in ? *o_in : default_opts();
    hipStream_t s = (o.device_matrix && o.stream) ? (hipStream_t)o.stream : g.s_main;
    const size_t bytes = (size_t)N * N * 8;
    double *d = (double *)A.matrix;
    if (!o.device_matrix) {
        if (hipMalloc(&d, bytes) != hipSuccess) return fail(PALIAL_ERR_NOMEM, "device copy");
    }
    hipError_t e = o.device_matrix ? hipSuccess : hipMemcpyAsync(d, A.matrix, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_transpose(d, N, b, average != 0, s);
    if (e == hipSuccess && !o.device_matrix) e = hipMemcpyAsync(A.matrix, d, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    else (void)hipStreamSynchronize(s);
    if (!o.device_matrix) (void)hipFree(d);
    if (e != hipSuccess) return fail(PALIAL_ERR_HIP, "transpose: %s", hipGetErrorString(e));
    return 0;
}

// ---------------------------------------------------------------- tile kernels
int palial_hip_dpotrf(int n, double *A, int lda, const palial_opts *opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init()) || (rc = check_tile(n, lda, A))) return rc;
    if (n == 0) return 0;
    Stage st((size_t)n * n * 8, opts && opts->device_matrix);
    double *d = st.in(A, true);
    if (!d) return fail(PALIAL_ERR_NOMEM, "tile alloc");
    HIPCHK(hipMemsetAsync(g.d_info, 0, sizeof(int), g.s_main));
    if ((rc = ensure_dinv(n, 1))) return rc;
    if (panel2_ok(n)) HIPCHK(launch_potrf2(d, n, g.d_info, 0, g.d_dinv, g.s_main));
    else HIPCHK(launch_potrf(d, n, g.d_info, 0, g.s_main));
    int info = 0;
    HIPCHK(hipMemcpyAsync(&info, g.d_info, sizeof(int), hipMemcpyDeviceToHost, g.s_main));
    if ((rc = st.finish())) return rc;
    return info;
}

int palial_hip_dgetrf(int n, double *A, int lda, const palial_opts *opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init()) || (rc = check_tile(n, lda, A))) return rc;
    if (n == 0) return 0;
    Stage st((size_t)n * n * 8, opts && opts->device_matrix);
    double *d = st.in(A, true);
    if (!d) return fail(PALIAL_ERR_NOMEM, "tile alloc");
    if ((rc = ensure_dinv(n, 1))) return rc;
    if (panel2_ok(n)) HIPCHK(launch_getrf2(d, n, g.d_dinv, g.d_dinv + (size_t)32 * n, g.s_main));
    else HIPCHK(launch_getrf(d, n, g.s_main));
    return st.finish();
}

int palial_hip_dgeqrt(int n, int nb, double *A, int lda, double *T, int ldt, double *tau, double *work,
                      const palial_opts *opts)
{
    (void)work;
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init()) || (rc = check_tile(n, lda, A)) || (rc = check_tile(n, ldt, T))) return rc;
    if (n == 0 || nb == 0) return 0;
    if (nb != n) return fail(PALIAL_ERR_ARG, "inner blocking nb=%d: only nb = n (as every reference call) is supported", nb);
    const bool dev = opts && opts->device_matrix;
    Stage st((size_t)n * n * 8, dev);
    double *dA = st.in(A, true), *dT = st.in(T, true);
    double *dtau = dev ? tau : st.scratch((size_t)n * 8);
    if (!dA || !dT || !dtau) return fail(PALIAL_ERR_NOMEM, "tile alloc");
    double *dG = st.scratch((size_t)n * n * 8);
    if (!dG) return fail(PALIAL_ERR_NOMEM, "tile alloc");
    HIPCHK(launch_geqrt(dA, dT, dtau, n, g.s_main, dG));
    if (!dev && tau) HIPCHK(hipMemcpyAsync(tau, dtau, (size_t)n * 8, hipMemcpyDeviceToHost, g.s_main));
    return st.finish();
}

int palial_hip_dormqr(int side, int trans, int n, const double *V, int ldv, const double *T, int ldt, double *C,
                      int ldc, double *work, int ldw, const palial_opts *opts)
{
    (void)work; (void)ldw;
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init()) || (rc = check_tile(n, ldv, V)) || (rc = check_tile(n, ldt, T)) ||
        (rc = check_tile(n, ldc, C)))
        return rc;
    if (n == 0) return 0;
    // src/srcqr/dormqr.h:18-29: only (Left, Trans) applies Q^T correctly there
    if (!(side == PALIAL_LEFT && trans != PALIAL_NOTRANS)) return fail(PALIAL_ERR_ARG, "only side=Left, trans=Trans is supported");
    const bool dev = opts && opts->device_matrix;
    Stage st((size_t)n * n * 8, dev);
    double *dV = st.in(V, false), *dT = st.in(T, false), *dC = st.in(C, true);
    double *w1 = st.scratch((size_t)n * n * 8), *w2 = st.scratch((size_t)n * n * 8);
    if (!dV || !dT || !dC || !w1 || !w2) return fail(PALIAL_ERR_NOMEM, "tile alloc");
    HIPCHK(launch_gemm(single(dV, dC, w1, nullptr, nullptr, n, 1.0, 0.0, MK_UNIT_LOWER), true, false, g.s_main));
    HIPCHK(launch_gemm(single(dT, w1, w2, nullptr, nullptr, n, 1.0, 0.0, MK_UPPER), true, false, g.s_main));
    HIPCHK(launch_gemm(single(dV, w2, dC, dC, nullptr, n, -1.0, 1.0, MK_UNIT_LOWER), false, false, g.s_main));
    return st.finish();
}

int palial_hip_dtsqrt(int n, double *A1, int lda1, double *A2, int lda2, double *T, int ldt, double *tau,
                      double *work, const palial_opts *opts)
{
    (void)work;
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init()) || (rc = check_tile(n, lda1, A1)) || (rc = check_tile(n, lda2, A2)) ||
        (rc = check_tile(n, ldt, T)))
        return rc;
    if (n == 0) return 0;
    const bool dev = opts && opts->device_matrix;
    Stage st((size_t)n * n * 8, dev);
    double *d1 = st.in(A1, true), *d2 = st.in(A2, true), *dT = st.in(T, true);
    double *dtau = dev ? tau : st.scratch((size_t)n * 8);
    if (!d1 || !d2 || !dT || !dtau) return fail(PALIAL_ERR_NOMEM, "tile alloc");
    double *dG = st.scratch((size_t)n * n * 8);
    if (!dG) return fail(PALIAL_ERR_NOMEM, "tile alloc");
    HIPCHK(launch_tsqrt(d1, d2, dT, dtau, n, g.s_main, dG));
    if (!dev && tau) HIPCHK(hipMemcpyAsync(tau, dtau, (size_t)n * 8, hipMemcpyDeviceToHost, g.s_main));
    return st.finish();
}

int palial_hip_dtsmqr(int side, int trans, int M, int N, int K, int ib, double *A1, int lda1, double *A2,
                      int lda2, const double *V, int ldv, const double *T, int ldt, double *work, int ldw,
                      const palial_opts *opts)
{
    (void)work; (void)ldw;
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init())) return rc;
    if (M == 0 || N == 0 || K == 0 || ib == 0) return 0;
    if (!(side == PALIAL_LEFT && trans != PALIAL_NOTRANS)) return fail(PALIAL_ERR_ARG, "only side=Left, trans=Trans is supported");
    if (!(M == N && N == K && K == ib)) return fail(PALIAL_ERR_ARG, "only M = N = K = ib (every reference call) is supported");
    const int n = M;
    if ((rc = check_tile(n, lda1, A1)) || (rc = check_tile(n, lda2, A2)) || (rc = check_tile(n, ldv, V)) ||
        (rc = check_tile(n, ldt, T)))
        return rc;
    const bool dev = opts && opts->device_matrix;
    Stage st((size_t)n * n * 8, dev);
    double *d1 = st.in(A1, true), *d2 = st.in(A2, true), *dV = st.in(V, false), *dT = st.in(T, false);
    double *w1 = st.scratch((size_t)n * n * 8), *w2 = st.scratch((size_t)n * n * 8);
    if (!d1 || !d2 || !dV || !dT || !w1 || !w2) return fail(PALIAL_ERR_NOMEM, "tile alloc");
    HIPCHK(launch_gemm(single(dV, d2, w1, d1, nullptr, n, 1.0, 1.0, MK_NONE), true, false, g.s_main));
    HIPCHK(launch_gemm(single(dT, w1, w2, nullptr, d1, n, 1.0, 0.0, MK_UPPER), true, false, g.s_main));
    HIPCHK(launch_gemm(single(dV, w2, d2, d2, nullptr, n, -1.0, 1.0, MK_NONE), false, false, g.s_main));
    return st.finish();
}

// palial_hip_dtsmqr with its two workspace tiles in the static `scratch`
// array instead of two hipMalloc'ed tiles: the same three launches of the same
// kernels, so the results are bitwise those of palial_hip_dtsmqr.  Both tiles
// are written in full by the first two launches before they are read, so
// nothing in the scratch has to start at zero.
int palial_hip_dtsmqr_noalloc(int side, int trans, int M, int N, int K, int ib, double *A1, int lda1, double *A2,
                              int lda2, const double *V, int ldv, const double *T, int ldt, double *work, int ldw,
                              const palial_opts *opts)
{
    (void)work; (void)ldw;
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init())) return rc;
    if (M == 0 || N == 0 || K == 0 || ib == 0) return 0;
    if (!(side == PALIAL_LEFT && trans != PALIAL_NOTRANS)) return fail(PALIAL_ERR_ARG, "only side=Left, trans=Trans is supported");
    if (!(M == N && N == K && K == ib)) return fail(PALIAL_ERR_ARG, "only M = N = K = ib (every reference call) is supported");
    const int n = M;
    if ((rc = check_tile(n, lda1, A1)) || (rc = check_tile(n, lda2, A2)) || (rc = check_tile(n, ldv, V)) ||
        (rc = check_tile(n, ldt, T)))
        return rc;
    if (!opts || !opts->device_matrix)
        return fail(PALIAL_ERR_ARG, "palial_hip_dtsmqr_noalloc needs device tiles (opts->device_matrix = 1): "
                    "staging host tiles would allocate");
    if (n > NOALLOC_MAXB)
        return fail(PALIAL_ERR_ARG, "tile size %d above %d, the most the static workspace of "
                    "palial_hip_dtsmqr_noalloc holds", n, NOALLOC_MAXB);
    double *const sp = g_scratch[g.device];
    if (!sp) return fail(PALIAL_ERR_STATE, "no-allocation workspace not set up on device %d", g.device);
    double *w1 = sp, *w2 = sp + (size_t)n * n;
    const hipStream_t s = opts->stream ? (hipStream_t)opts->stream : g.s_main;
    HIPCHK(launch_gemm(single(V, A2, w1, A1, nullptr, n, 1.0, 1.0, MK_NONE), true, false, s));
    HIPCHK(launch_gemm(single(T, w1, w2, nullptr, A1, n, 1.0, 0.0, MK_UPPER), true, false, s));
    HIPCHK(launch_gemm(single(V, w2, A2, A2, nullptr, n, -1.0, 1.0, MK_NONE), false, false, s));
    if (!opts->stream) HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int palial_hip_tile_task(int task, int n, const double *A, const double *B, double *C, const palial_opts *opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = ensure_init()) || (rc = check_tile(n, n, A)) || (rc = check_tile(n, n, C))) return rc;
    const bool needB = (task == PALIAL_TASK_CHOL_GEMM || task == PALIAL_TASK_LU_GEMM);
    if (needB && !B) return fail(PALIAL_ERR_ARG, "task %d needs B", task);
    if (n == 0) return 0;
    const bool dev = opts && opts->device_matrix;
    Stage st((size_t)n * n * 8, dev);
    double *dA = st.in(A, false), *dB = needB ? st.in(B, false) : nullptr, *dC = st.in(C, true);
    if (!dA || !dC || (needB && !dB)) return fail(PALIAL_ERR_NOMEM, "tile alloc");
    TrsmArgs t;
    memset(&t, 0, sizeof t);
    t.diag = dA; t.single = dC; t.N = n; t.b = n; t.MT = 1; t.ntiles = 1;
    switch (task) {
    case PALIAL_TASK_CHOL_TRSM: HIPCHK(launch_trsm(TR_CHOL, t, g.s_main)); break;
    case PALIAL_TASK_LU_TRSM_R: HIPCHK(launch_trsm(TR_LUR, t, g.s_main)); break;
    case PALIAL_TASK_LU_TRSM_L: HIPCHK(launch_trsm(TR_LUL, t, g.s_main)); break;
    case PALIAL_TASK_CHOL_SYRK: {
        GemmArgs a = single(dA, dA, dC, dC, nullptr, n, -1.0, 1.0, MK_NONE);
        a.single_diag = 1;
        HIPCHK(launch_gemm(a, true, false, g.s_main));
        break;
    }
    case PALIAL_TASK_CHOL_GEMM:
        HIPCHK(launch_gemm(single(dA, dB, dC, dC, nullptr, n, -1.0, 1.0, MK_NONE), true, false, g.s_main));
        break;
    case PALIAL_TASK_LU_GEMM:
        HIPCHK(launch_gemm(single(dA, dB, dC, dC, nullptr, n, -1.0, 1.0, MK_NONE), false, false, g.s_main));
        break;
    default: return fail(PALIAL_ERR_ARG, "unknown tile task %d", task);
    }
    return st.finish();
}

}  // extern "C"
