// hipemu.cpp -- runtime of the CPU execution of the library's HIP sources
// (tests only; the header describes the execution model).
#include "hipemu.hpp"

#include <chrono>
#include <cstdarg>
#include <dlfcn.h>
#include <map>
#include <string>
#include <sys/mman.h>
#include <vector>

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define EMU_ASAN 1
#endif
#endif
#ifdef EMU_ASAN
#include <sanitizer/asan_interface.h>
#include <sanitizer/common_interface_defs.h>
#endif

// ------------------------------------------------------------------ context switch
// void emu_switch(void **save_sp, void *new_sp): saves the callee-saved
// registers, MXCSR and the x87 control word on the current stack, stores the
// stack pointer in *save_sp and resumes the context saved at new_sp.
// emu_trampoline: first code of a fresh fiber; calls r13(r12).
asm(R"(
    .text
    .globl emu_switch
    .type emu_switch,@function
emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    subq $8, %rsp
    stmxcsr (%rsp)
    fnstcw 4(%rsp)
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    ldmxcsr (%rsp)
    fldcw 4(%rsp)
    addq $8, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_switch, .-emu_switch

    .globl emu_trampoline
    .type emu_trampoline,@function
emu_trampoline:
    movq %r12, %rdi
    callq *%r13
    ud2
    .size emu_trampoline, .-emu_trampoline
)");
extern "C" void emu_switch(void **save_sp, void *new_sp);
extern "C" void emu_trampoline();

struct emu_stream { int id; };
struct emu_event { double t_ms; };
struct emu_graph { std::vector<std::function<void()>> ops; };

namespace emu {

Lane *g_lane = nullptr;
dim3 g_block, g_grid, g_bdim;
void *g_dyn_lds = nullptr;

namespace {

enum State { RUN = 0, BAR = 1, WAVE = 2, DONE = 3 };

constexpr size_t STACK = 256 * 1024;
constexpr int MAXT = 1024;

struct Fiber {
    Lane lane;
    void *sp = nullptr;
    char *stack = nullptr;           // lowest address (guard page below the usable part)
    int state = DONE;
    int wave = 0;
};

Fiber g_fib[MAXT];
int g_nt = 0;                        // lanes of the current workgroup
Fiber *g_cur = nullptr;              // running fiber (nullptr: the scheduler)
void *g_sched_sp = nullptr;
const std::function<void()> *g_body = nullptr;
alignas(64) unsigned char g_slots[MAXT / 64][64][128];
const void *g_kfn = nullptr;
hipError_t g_last_err = hipSuccess;
std::map<const void *, int> g_dyn_limit;
emu_graph *g_capture = nullptr;
hipStream_t g_capture_stream = nullptr;
unsigned long long g_tick = 0;
auto g_t0 = std::chrono::steady_clock::now();

#ifdef EMU_ASAN
void *g_fake_sched = nullptr;
const void *g_sched_bottom = nullptr;
size_t g_sched_size = 0;
#endif

std::string kname(const void *f)
{
    Dl_info di;
    if (f && dladdr(f, &di) && di.dli_sname) return di.dli_sname;
    return "?";
}

// scheduler -> fiber f
void enter(Fiber *f)
{
    g_cur = f;
    g_lane = &f->lane;
#ifdef EMU_ASAN
    __sanitizer_start_switch_fiber(&g_fake_sched, f->stack + 4096, STACK - 4096);
#endif
    emu_switch(&g_sched_sp, f->sp);
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(g_fake_sched, nullptr, nullptr);
#endif
    g_cur = nullptr;
}

// running fiber -> scheduler (state already set)
void leave()
{
    Fiber *f = g_cur;
#ifdef EMU_ASAN
    void *fake = nullptr;
    __sanitizer_start_switch_fiber(f->state == DONE ? nullptr : &fake, g_sched_bottom, g_sched_size);
#endif
    emu_switch(&f->sp, g_sched_sp);
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
}

extern "C" void emu_fiber_main(Fiber *f)
{
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &g_sched_bottom, &g_sched_size);
#endif
    (*g_body)();
    f->state = DONE;
    leave();
    __builtin_trap();
}

void init_fiber(Fiber *f)
{
    if (!f->stack) {
        void *p = mmap(nullptr, STACK, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
        if (p == MAP_FAILED) fatal("fiber stack allocation failed");
        mprotect(p, 4096, PROT_NONE);            // guard page: a stack overflow faults
        f->stack = (char *)p;
    }
#ifdef EMU_ASAN
    // a fiber that ended left its last frames' redzones in the shadow memory
    __asan_unpoison_memory_region(f->stack + 4096, STACK - 4096);
#endif
    uintptr_t top = ((uintptr_t)(f->stack + STACK)) & ~(uintptr_t)15;
    uint64_t *s = (uint64_t *)top;
    *--s = (uint64_t)(uintptr_t)&emu_trampoline;  // ret target
    *--s = 0;                                     // rbp
    *--s = 0;                                     // rbx
    *--s = (uint64_t)(uintptr_t)f;                // r12: argument
    *--s = (uint64_t)(uintptr_t)&emu_fiber_main;  // r13: entry
    *--s = 0;                                     // r14
    *--s = 0;                                     // r15
    --s;
    ((uint32_t *)s)[0] = 0x1F80;                  // MXCSR default
    ((uint16_t *)s)[2] = 0x037F;                  // x87 control word default
    f->sp = s;
}

void run_block()
{
    for (int i = 0; i < g_nt; i++) {
        init_fiber(&g_fib[i]);
        g_fib[i].state = RUN;
    }
    const int nw = (g_nt + 63) / 64;
    for (;;) {
        // Lanes run one after the other between synchronization points, in
        // descending order: the device executes a wave in lockstep, and code
        // written for that reads a location in every lane before lane 0 (or
        // thread 0) overwrites it further on in program order; lane 0 runs
        // last here, so those reads see the value the device's would.  The
        // other direction (a lane reading what a lower lane wrote before it
        // in program order, with no wave operation between) is not modelled.
        bool ran = false;
        for (int n = 0; n < g_nt; n++) {
            const int i = g_nt - 1 - n;
            if (g_fib[i].state == RUN) {
                enter(&g_fib[i]);
                ran = true;
            }
        }
        if (ran) continue;
        // nobody runnable: the lanes of a wave that wait at a wave operation
        // execute it together (the others of the wave wait at the barrier or
        // have returned: on the device they are inactive for this operation),
        // else the workgroup barrier opens when every live lane waits there
        bool released = false;
        for (int w = 0; w < nw; w++) {
            int waiting = 0;
            for (int i = w * 64; i < std::min(g_nt, w * 64 + 64); i++) waiting += g_fib[i].state == WAVE;
            if (waiting > 0) {
                for (int i = w * 64; i < std::min(g_nt, w * 64 + 64); i++)
                    if (g_fib[i].state == WAVE) g_fib[i].state = RUN;
                released = true;
            }
        }
        if (released) continue;
        int live = 0, bar = 0;
        for (int i = 0; i < g_nt; i++) {
            live += g_fib[i].state != DONE;
            bar += g_fib[i].state == BAR;
        }
        if (live == 0) break;
        if (bar == live) {
            for (int i = 0; i < g_nt; i++)
                if (g_fib[i].state == BAR) g_fib[i].state = RUN;
            continue;
        }
        fatal("deadlock in %s block (%u,%u,%u): %d live lanes, %d at the barrier", kname(g_kfn).c_str(), g_block.x,
              g_block.y, g_block.z, live, bar);
    }
}

void run_grid(const void *kfn, dim3 g, dim3 b, size_t shm, const std::function<void()> &body)
{
    const int nt = (int)(b.x * b.y * b.z);
    g_kfn = kfn;
    g_grid = g;
    g_bdim = b;
    g_nt = nt;
    g_body = &body;
    for (int i = 0; i < nt; i++) {
        Lane &l = g_fib[i].lane;
        l.lin = i;
        l.tid = dim3(i % b.x, (i / b.x) % b.y, i / (b.x * b.y));
        g_fib[i].wave = i / 64;
    }
    for (unsigned z = 0; z < g.z; z++)
        for (unsigned y = 0; y < g.y; y++)
            for (unsigned x = 0; x < g.x; x++) {
                g_block = dim3(x, y, z);
                // fresh dynamic LDS per workgroup, exactly the launch's size,
                // filled with a NaN pattern (LDS is not zeroed on the device)
                void *lds = malloc(shm ? shm : 1);
                memset(lds, 0xff, shm ? shm : 1);
                g_dyn_lds = lds;
                run_block();
                free(lds);
                g_dyn_lds = nullptr;
            }
    g_lane = nullptr;
}

void exec_launch(const void *kfn, dim3 g, dim3 b, size_t shm, const std::function<void()> &body)
{
    const long nt = (long)b.x * b.y * b.z;
    if (g.x == 0 || g.y == 0 || g.z == 0 || nt <= 0 || nt > MAXT || shm > 160 * 1024) {
        fprintf(stderr, "hipemu: invalid launch of %s: grid (%u,%u,%u) block (%u,%u,%u) dynamic LDS %zu\n",
                kname(kfn).c_str(), g.x, g.y, g.z, b.x, b.y, b.z, shm);
        g_last_err = hipErrorInvalidConfiguration;
        return;
    }
    run_grid(kfn, g, b, shm, body);
}

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - g_t0).count();
}

// run a stream operation now, or record it while a capture is open
hipError_t stream_op(std::function<void()> op)
{
    if (g_capture) g_capture->ops.push_back(std::move(op));
    else op();
    return hipSuccess;
}

}  // namespace

void fatal(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "hipemu: ");
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    va_end(ap);
    fflush(stderr);
    abort();
}

unsigned long long clock_tick() { return ++g_tick; }

void barrier()
{
    if (!g_cur) fatal("__syncthreads outside a kernel");
    g_cur->state = BAR;
    leave();
}

void wave_sync()
{
    if (!g_cur) fatal("wave operation outside a kernel");
    g_cur->state = WAVE;
    leave();
}

unsigned char *slot(int lane)
{
    const int w = g_lane->lin / 64;
    return g_slots[w][lane & 63];
}

void submit_launch(hipStream_t, const void *kfn, dim3 g, dim3 b, size_t shm, std::function<void()> lane_fn)
{
    if (g_capture) {
        g_capture->ops.push_back([kfn, g, b, shm, lane_fn]() { exec_launch(kfn, g, b, shm, lane_fn); });
        return;
    }
    exec_launch(kfn, g, b, shm, lane_fn);
}

// ------------------------------------------------------------------ MFMA
// D(i, j) = C(i, j) + sum_k A(i, k) B(k, j), 16 x 16 outputs.  Lane l
// supplies A(l % 16, K block l / 16) and B(K block l / 16, l % 16).  With FP32
// accumulators lane l holds D(4 (l / 16) + v, l % 16), v = 0..3; with FP64
// accumulators D(l / 16 + 4 v, l % 16).
v4d mfma_f64_16x16x4(double a, double b, v4d c)
{
    struct S { double a, b; } s{a, b};
    const int l = lane_id();
    memcpy(slot(l), &s, sizeof s);
    wave_sync();
    v4d d;
    for (int v = 0; v < 4; v++) {
        const int i = l / 16 + 4 * v, j = l % 16;
        double acc = c[v];
        for (int k = 0; k < 4; k++) {
            S sa, sb;
            memcpy(&sa, slot(i + 16 * k), sizeof sa);
            memcpy(&sb, slot(j + 16 * k), sizeof sb);
            acc = std::fma(sa.a, sb.b, acc);
        }
        d[v] = acc;
    }
    wave_sync();
    return d;
}

v4f mfma_f32_16x16x4(float a, float b, v4f c)
{
    struct S { float a, b; } s{a, b};
    const int l = lane_id();
    memcpy(slot(l), &s, sizeof s);
    wave_sync();
    v4f d;
    for (int v = 0; v < 4; v++) {
        const int i = 4 * (l / 16) + v, j = l % 16;
        double acc = 0.0;
        for (int k = 0; k < 4; k++) {
            S sa, sb;
            memcpy(&sa, slot(i + 16 * k), sizeof sa);
            memcpy(&sb, slot(j + 16 * k), sizeof sb);
            acc += (double)sa.a * (double)sb.b;
        }
        d[v] = (float)((double)c[v] + acc);
    }
    wave_sync();
    return d;
}

namespace {
inline float bf16f(unsigned short h) { return __builtin_bit_cast(float, (unsigned)h << 16); }

// OCP E4M3 (bias 7, no infinities, S.1111.111 = NaN)
float fp8_value(unsigned char x)
{
    const int s = x >> 7, e = (x >> 3) & 15, m = x & 7;
    float v;
    if (e == 15 && m == 7) v = NAN;
    else if (e == 0) v = std::ldexp((float)m, -9);
    else v = std::ldexp(1.0f + m / 8.0f, e - 7);
    return s ? -v : v;
}

// nearest-even E4M3; magnitudes beyond 448 saturate
unsigned char fp8_of(float f)
{
    if (std::isnan(f)) return 0x7f;
    const unsigned char sgn = std::signbit(f) ? 0x80 : 0;
    double a = std::fabs((double)f);
    if (a >= 448.0) return sgn | 0x7e;
    // candidates: every finite E4M3 magnitude, ordered by code
    int best = 0;
    double bd = a;
    for (int c = 1; c < 0x7f; c++) {
        const double v = fp8_value((unsigned char)c);
        const double d = std::fabs(v - a);
        if (d < bd || (d == bd && (c & 1) == 0 && (best & 1) == 1)) {
            best = c;
            bd = d;
        }
    }
    return sgn | (unsigned char)best;
}
}  // namespace

v4f mfma_f32_16x16x32_bf16(v8u16 a, v8u16 b, v4f c)
{
    struct S { v8u16 a, b; } s{a, b};
    const int l = lane_id();
    memcpy(slot(l), &s, sizeof s);
    wave_sync();
    v4f d;
    for (int v = 0; v < 4; v++) {
        const int i = 4 * (l / 16) + v, j = l % 16;
        double acc = 0.0;
        for (int kb = 0; kb < 4; kb++) {
            S sa, sb;
            memcpy(&sa, slot(i + 16 * kb), sizeof sa);
            memcpy(&sb, slot(j + 16 * kb), sizeof sb);
            for (int e = 0; e < 8; e++) acc += (double)bf16f(sa.a[e]) * (double)bf16f(sb.b[e]);
        }
        d[v] = (float)((double)c[v] + acc);
    }
    wave_sync();
    return d;
}

v4f mfma_f32_16x16x32_fp8(long a, long b, v4f c)
{
    struct S { long a, b; } s{a, b};
    const int l = lane_id();
    memcpy(slot(l), &s, sizeof s);
    wave_sync();
    v4f d;
    for (int v = 0; v < 4; v++) {
        const int i = 4 * (l / 16) + v, j = l % 16;
        double acc = 0.0;
        for (int kb = 0; kb < 4; kb++) {
            S sa, sb;
            memcpy(&sa, slot(i + 16 * kb), sizeof sa);
            memcpy(&sb, slot(j + 16 * kb), sizeof sb);
            for (int e = 0; e < 8; e++)
                acc += (double)fp8_value((unsigned char)(sa.a >> (8 * e))) *
                       (double)fp8_value((unsigned char)(sb.b >> (8 * e)));
        }
        d[v] = (float)((double)c[v] + acc);
    }
    wave_sync();
    return d;
}

void load_lds(void *dst, const void *src, size_t n)
{
#ifdef EMU_ASAN
    const bool bad_src = __asan_region_is_poisoned(const_cast<void *>(src), n) != nullptr;
    const bool bad_dst = __asan_region_is_poisoned(dst, n) != nullptr;
    if (bad_src || bad_dst)
        fatal("global_load_lds out of bounds in %s block (%u,%u,%u) lane %d: %s %p (%zu bytes)", kname(g_kfn).c_str(),
              g_block.x, g_block.y, g_block.z, g_lane->lin, bad_src ? "source" : "LDS destination",
              bad_src ? src : dst, n);
#endif
    memcpy(dst, src, n);
}

int cvt_pk_fp8_f32(float a, float b, int old, bool hi)
{
    const unsigned v = (unsigned)fp8_of(a) | ((unsigned)fp8_of(b) << 8);
    const unsigned o = (unsigned)old;
    return (int)(hi ? ((o & 0xffffu) | (v << 16)) : ((o & 0xffff0000u) | v));
}

// DPP source lane inside the 16-lane row: quad_perm, row_mirror, row_half_mirror
int update_dpp(int old, int src, int ctrl)
{
    const int l = lane_id();
    int s;
    if (ctrl >= 0 && ctrl <= 0xff) s = (l & ~3) | ((ctrl >> (2 * (l & 3))) & 3);
    else if (ctrl == 0x140) s = (l & ~15) | (15 - (l & 15));
    else if (ctrl == 0x141) s = (l & ~7) | (7 - (l & 7));
    else fatal("DPP control 0x%x not emulated", ctrl);
    (void)old;
    return xchg(src, s);
}

}  // namespace emu

using namespace emu;

// ------------------------------------------------------------------ runtime API
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d)
{
    if (d != 0) return hipErrorInvalidDevice;
    memset(p, 0, sizeof *p);
    snprintf(p->name, sizeof p->name, "hipemu (CPU execution)");
    snprintf(p->gcnArchName, sizeof p->gcnArchName, "gfx950:sramecc+:xnack-");
    p->totalGlobalMem = (size_t)288 << 30;
    p->sharedMemPerBlock = 160 * 1024;
    p->maxSharedMemoryPerMultiProcessor = 160 * 1024;
    p->warpSize = 64;
    p->maxThreadsPerBlock = 1024;
    p->multiProcessorCount = 256;
    p->clockRate = 2400000;
    p->major = 9;
    p->minor = 5;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int)
{
    if (a == hipDeviceAttributeWallClockRate) *v = 100000;
    else if (a == hipDeviceAttributeMultiprocessorCount) *v = 256;
    else *v = 0;
    return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError()
{
    const hipError_t e = g_last_err;
    g_last_err = hipSuccess;
    return e;
}
hipError_t hipPeekAtLastError() { return g_last_err; }
const char *hipGetErrorString(hipError_t e)
{
    switch (e) {
    case hipSuccess: return "hipSuccess";
    case hipErrorInvalidValue: return "hipErrorInvalidValue";
    case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
    case hipErrorInvalidConfiguration: return "hipErrorInvalidConfiguration";
    case hipErrorInvalidDevice: return "hipErrorInvalidDevice";
    default: return "hipemu error";
    }
}

hipError_t hipMalloc(void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return hipSuccess;
    if (posix_memalign(p, 256, bytes) != 0) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipSuccess;
}
hipError_t hipFree(void *p)
{
    free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind)
{
    if (n) memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t)
{
    return stream_op([=]() { if (n) memmove(d, s, n); });
}
hipError_t hipMemcpy2DAsync(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h, hipMemcpyKind,
                            hipStream_t)
{
    return stream_op([=]() {
        for (size_t r = 0; r < h; r++) memmove((char *)d + r * dp, (const char *)s + r * sp, w);
    });
}
hipError_t hipMemset(void *p, int v, size_t n)
{
    if (n) memset(p, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t)
{
    return stream_op([=]() { if (n) memset(p, v, n); });
}

hipError_t hipStreamCreate(hipStream_t *s)
{
    static int next = 1;
    *s = new emu_stream{next++};
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { return hipStreamCreate(s); }
hipError_t hipExtStreamCreateWithCUMask(hipStream_t *s, uint32_t words, const uint32_t *mask)
{
    if (words == 0 || !mask) return hipErrorInvalidValue;
    return hipStreamCreate(s);
}
hipError_t hipStreamDestroy(hipStream_t s)
{
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s)
{
    if (g_capture && s == g_capture_stream) return hipErrorStreamCaptureUnsupported;
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t e, unsigned) { return e ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipEventCreate(hipEvent_t *e)
{
    *e = new emu_event{0.0};
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e)
{
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t)
{
    if (!e) return hipErrorInvalidValue;
    if (g_capture) return hipSuccess;
    e->t_ms = now_ms();
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) { return e ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b)
{
    if (!a || !b) return hipErrorInvalidValue;
    *ms = (float)(b->t_ms - a->t_ms);
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode)
{
    if (g_capture) return hipErrorStreamCaptureUnsupported;
    g_capture = new emu_graph();
    g_capture_stream = s;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g)
{
    if (!g_capture || s != g_capture_stream) return hipErrorStreamCaptureInvalidated;
    *g = g_capture;
    g_capture = nullptr;
    g_capture_stream = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t g, hipGraphNode_t *, char *, size_t)
{
    *x = new emu_graph(*g);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t)
{
    if (!x) return hipErrorInvalidValue;
    for (auto &op : x->ops) op();
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x)
{
    delete x;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete g;
    return hipSuccess;
}
hipError_t hipGraphGetNodes(hipGraph_t g, hipGraphNode_t *, size_t *n)
{
    *n = g->ops.size();
    return hipSuccess;
}
hipError_t hipFuncSetAttribute(const void *f, hipFuncAttribute a, int v)
{
    if (a == hipFuncAttributeMaxDynamicSharedMemorySize) {
        if (v < 0 || v > 160 * 1024) return hipErrorInvalidValue;
        g_dyn_limit[f] = v;
    }
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void **p, const void *sym)
{
    *p = const_cast<void *>(sym);
    return hipSuccess;
}

// ------------------------------------------------------------------ test hooks
extern "C" {
// launches and blocks run so far (a test can check that work was done)
long hipemu_ticks(void) { return (long)g_tick; }
}
