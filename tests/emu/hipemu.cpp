// hipemu.cpp -- runtime of the CPU execution of the library's HIP sources
// (tests only; the header describes the execution model).
#include "hipemu.hpp"

#include <chrono>
#include <cxxabi.h>
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

// A stream's and an event's vector clock (race-checking build): entry i is
// the number of operations of stream slot i that happen before what the
// stream issues next / what the event marks.
struct emu_stream {
    int id;
    bool blocking = true;             // synchronises with the null stream (hipStreamNonBlocking not set)
    std::vector<uint32_t> vc;
};
struct emu_event {
    double t_ms = 0.0;
    std::vector<uint32_t> vc;
};
// a captured node: a stream operation (what it is -- a kernel's address, or a
// copy or fill -- and the code that performs it), or an event record or wait,
// with the stream it was issued on
struct emu_op {
    const void *kfn;
    const char *what;
    std::function<void()> run;
    int type = 0;                     // 0 operation, 1 event record, 2 event wait
    emu_stream *stream = nullptr;
    emu_event *event = nullptr;
};
struct emu_graph { std::vector<emu_op> ops; };

namespace emu {

// ------------------------------------------------------------------ race checker interface
// Race-checking build (HIPEMU_RACE): every stream operation is an "op" with
// a vector clock; the library's loads and stores reach the checker through
// the compiler's thread-sanitizer instrumentation (callbacks at the end of
// this file), and accesses to device memory are checked against the
// happens-before order of the ops.  Elsewhere these are no-ops.
namespace race {
#ifdef HIPEMU_RACE
void begin_op(hipStream_t s, const void *kfn, const char *what);
void end_op();
void set_workgroup(unsigned wg);
void barrier_opened();
void wave_synced(int wave);
void dma_issued(const void *lds, size_t n);
void outside_capture(const void *kfn, const char *what);
bool drop_this_wait();
void access(const void *p, size_t n, int kind);
void *dev_alloc(size_t bytes);
bool dev_free(void *p);
void record(hipEvent_t e, hipStream_t s);
bool wait(hipStream_t s, hipEvent_t e);
void host_sync_stream(hipStream_t s);
void host_sync_event(hipEvent_t e);
void host_sync_device();
void stream_created(hipStream_t s, bool blocking);
void stream_destroyed(hipStream_t s);
#else
inline void begin_op(hipStream_t, const void *, const char *) {}
inline void end_op() {}
inline void set_workgroup(unsigned) {}
inline void barrier_opened() {}
inline void wave_synced(int) {}
inline void dma_issued(const void *, size_t) {}
inline void outside_capture(const void *, const char *) {}
inline bool drop_this_wait() { return false; }
inline void access(const void *, size_t, int) {}
inline void record(hipEvent_t, hipStream_t) {}
inline bool wait(hipStream_t, hipEvent_t) { return true; }
inline void host_sync_stream(hipStream_t) {}
inline void host_sync_event(hipEvent_t) {}
inline void host_sync_device() {}
inline void stream_created(hipStream_t, bool) {}
inline void stream_destroyed(hipStream_t) {}
#endif
enum { READ = 0, WRITE = 1, ATOMIC = 2 };
}  // namespace race

Lane *g_lane = nullptr;
dim3 g_block, g_grid, g_bdim;
void *g_dyn_lds = nullptr;
size_t g_dyn_bytes = 0;                  // size of this launch's dynamic LDS block

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
// streams taking part in the open capture (its origin, and every stream that
// waited for an event recorded in it) and the events recorded in it
std::vector<hipStream_t> g_cap_streams;
std::vector<hipEvent_t> g_cap_events;
bool in_capture(hipStream_t s) { return std::find(g_cap_streams.begin(), g_cap_streams.end(), s) != g_cap_streams.end(); }
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
                race::wave_synced(w);
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
            race::barrier_opened();
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
                race::set_workgroup(x + g.x * (y + g.y * z));
                // fresh dynamic LDS per workgroup, exactly the launch's size,
                // filled with a NaN pattern (LDS is not zeroed on the device)
                void *lds = malloc(shm ? shm : 1);
                memset(lds, 0xff, shm ? shm : 1);
                g_dyn_lds = lds;
                g_dyn_bytes = shm;
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
hipError_t stream_op(hipStream_t s, const void *kfn, const char *what, std::function<void()> op)
{
    if (g_capture) {
        // work on a stream that is not part of the capture would run at once
        // on the device instead of joining the graph
        if (!in_capture(s)) race::outside_capture(kfn, what);
        emu_op n{kfn, what, std::move(op)};
        n.stream = s;
        g_capture->ops.push_back(std::move(n));
        return hipSuccess;
    }
    race::begin_op(s, kfn, what);
    op();
    race::end_op();
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

void submit_launch(hipStream_t s, const void *kfn, const char *name, dim3 g, dim3 b, size_t shm,
                   std::function<void()> lane_fn)
{
    stream_op(s, kfn, name, [kfn, g, b, shm, lane_fn]() { exec_launch(kfn, g, b, shm, lane_fn); });
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
    race::access(src, n, race::READ);
    race::access(dst, n, race::WRITE);       // the lane's store to LDS
    race::dma_issued(dst, n);
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
hipError_t hipDeviceSynchronize()
{
    race::host_sync_device();
    return hipSuccess;
}
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
#ifdef HIPEMU_RACE
    *p = race::dev_alloc(bytes);
    return *p ? hipSuccess : hipErrorOutOfMemory;
#else
    if (posix_memalign(p, 256, bytes) != 0) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipSuccess;
#endif
}
hipError_t hipFree(void *p)
{
#ifdef HIPEMU_RACE
    if (p && !race::dev_free(p)) return hipErrorInvalidValue;
#else
    free(p);
#endif
    return hipSuccess;
}
// the synchronous copies and fills run on the null stream and return when done
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k)
{
    const hipError_t e = hipMemcpyAsync(d, s, n, k, nullptr);
    race::host_sync_stream(nullptr);
    return e;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t st)
{
    return stream_op(st, nullptr, "copy", [=]() {
        if (!n) return;
        race::access(s, n, race::READ);
        race::access(d, n, race::WRITE);
        memmove(d, s, n);
    });
}
hipError_t hipMemcpy2DAsync(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h, hipMemcpyKind,
                            hipStream_t st)
{
    return stream_op(st, nullptr, "2-D copy", [=]() {
        for (size_t r = 0; r < h; r++) {
            race::access((const char *)s + r * sp, w, race::READ);
            race::access((char *)d + r * dp, w, race::WRITE);
            memmove((char *)d + r * dp, (const char *)s + r * sp, w);
        }
    });
}
hipError_t hipMemset(void *p, int v, size_t n)
{
    const hipError_t e = hipMemsetAsync(p, v, n, nullptr);
    race::host_sync_stream(nullptr);
    return e;
}
hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t st)
{
    return stream_op(st, nullptr, "fill", [=]() {
        if (!n) return;
        race::access(p, n, race::WRITE);
        memset(p, v, n);
    });
}

hipError_t hipStreamCreate(hipStream_t *s) { return hipStreamCreateWithFlags(s, hipStreamDefault); }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags)
{
    static int next = 1;
    *s = new emu_stream{next++};
    race::stream_created(*s, !(flags & hipStreamNonBlocking));
    return hipSuccess;
}
hipError_t hipExtStreamCreateWithCUMask(hipStream_t *s, uint32_t words, const uint32_t *mask)
{
    if (words == 0 || !mask) return hipErrorInvalidValue;
    return hipStreamCreateWithFlags(s, hipStreamDefault);
}
hipError_t hipStreamDestroy(hipStream_t s)
{
    race::stream_destroyed(s);
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s)
{
    if (g_capture && s == g_capture_stream) return hipErrorStreamCaptureUnsupported;
    race::host_sync_stream(s);
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned)
{
    if (!e) return hipErrorInvalidValue;
    if (race::drop_this_wait()) return hipSuccess;     // the checker's planted mistake
    if (g_capture && std::find(g_cap_events.begin(), g_cap_events.end(), e) != g_cap_events.end()) {
        // a graph edge; the waiting stream joins the capture
        if (!in_capture(s)) g_cap_streams.push_back(s);
        emu_op n{nullptr, "wait", nullptr};
        n.type = 2;
        n.stream = s;
        n.event = e;
        g_capture->ops.push_back(std::move(n));
        return hipSuccess;
    }
    race::wait(s, e);
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e)
{
    *e = new emu_event();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e)
{
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s)
{
    if (!e) return hipErrorInvalidValue;
    if (g_capture && in_capture(s)) {
        if (std::find(g_cap_events.begin(), g_cap_events.end(), e) == g_cap_events.end()) g_cap_events.push_back(e);
        emu_op n{nullptr, "record", nullptr};
        n.type = 1;
        n.stream = s;
        n.event = e;
        g_capture->ops.push_back(std::move(n));
        return hipSuccess;
    }
    e->t_ms = now_ms();
    race::record(e, s);
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e)
{
    if (!e) return hipErrorInvalidValue;
    race::host_sync_event(e);
    return hipSuccess;
}
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
    g_cap_streams.assign(1, s);
    g_cap_events.clear();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g)
{
    if (!g_capture || s != g_capture_stream) return hipErrorStreamCaptureInvalidated;
    *g = g_capture;
    g_capture = nullptr;
    g_capture_stream = nullptr;
    g_cap_streams.clear();
    g_cap_events.clear();
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t g, hipGraphNode_t *, char *, size_t)
{
    *x = new emu_graph(*g);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s)
{
    if (!x) return hipErrorInvalidValue;
#ifdef HIPEMU_RACE
    // The nodes in capture order, each captured stream replaced by a fresh
    // stream that starts after the launch stream's earlier work and each
    // captured event by a fresh event, so that the graph keeps exactly its
    // captured edges; the launch stream then waits for all of them.
    std::vector<std::pair<hipStream_t, hipStream_t>> vs;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ve;
    hipEvent_t start;
    (void)hipEventCreate(&start);
    race::record(start, s);
    auto vstream = [&](hipStream_t o) {
        for (auto &p : vs)
            if (p.first == o) return p.second;
        hipStream_t v;
        (void)hipStreamCreateWithFlags(&v, hipStreamNonBlocking);
        race::wait(v, start);
        vs.push_back({o, v});
        return v;
    };
    auto vevent = [&](hipEvent_t o) {
        for (auto &p : ve)
            if (p.first == o) return p.second;
        hipEvent_t v;
        (void)hipEventCreate(&v);
        ve.push_back({o, v});
        return v;
    };
    for (auto &op : x->ops) {
        if (op.type == 1) race::record(vevent(op.event), vstream(op.stream));
        else if (op.type == 2) race::wait(vstream(op.stream), vevent(op.event));
        else stream_op(vstream(op.stream), op.kfn, op.what, op.run);
    }
    for (auto &p : vs) {
        hipEvent_t end;
        (void)hipEventCreate(&end);
        race::record(end, p.second);
        race::wait(s, end);
        (void)hipEventDestroy(end);
        (void)hipStreamDestroy(p.second);
    }
    for (auto &p : ve) (void)hipEventDestroy(p.second);
    (void)hipEventDestroy(start);
#else
    // the recorded operations one after the other on the launch stream
    for (auto &op : x->ops)
        if (op.type == 0) stream_op(s, op.kfn, op.what, op.run);
#endif
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

// ------------------------------------------------------------------ race checker
// Built with HIPEMU_RACE: the library's sources are compiled with
// -fsanitize=thread, which makes the compiler call __tsan_read4(p),
// __tsan_write8(p), __tsan_memcpy(...) and so on around every load, store and
// memory intrinsic; this file (compiled without it) defines those callbacks
// itself -- no sanitizer runtime is linked -- and checks each access to device
// memory against the happens-before order of the stream operations:
//   * every copy, fill and kernel launch is an op on its stream; an op issued
//     on stream s happens after the ops s has seen: its own earlier ops, the
//     ops behind every event it waited for (at the time of the wait), the
//     host's synchronisations so far, and for a blocking stream the null
//     stream's ops (and the other way round);
//   * device memory is a reserved address range with a shadow record per
//     4-byte granule: the op and workgroup that last wrote it and up to two
//     ops that read it since;
//   * a read of a granule whose writer op does not happen before the reader,
//     and a write of a granule whose writer or readers do not happen before
//     the writer, is a race between streams; within one launch, a granule
//     written by one workgroup and read or written by another is a race
//     between workgroups (atomics excepted);
//   * host code that loads or stores device memory is reported too (on the
//     device it would fault);
//   * LDS (the __shared__ arrays, gathered in section emu_lds, and a launch's
//     dynamic block): an access that conflicts with another wave's access
//     in the same barrier interval is a race, and a read of a granule this
//     workgroup has not written reads what an earlier workgroup left;
//   * lanes of one wave that pass a value through memory with no wave
//     operation between, in an order whose outcome differs between the
//     lockstep device and the emulator's lane-by-lane execution (lane_check);
//   * an LDS read of a global_load_lds destination whose load the issuing
//     lane has not yet retired with s_waitcnt vmcnt (dma_issued, waitcnt);
//   * a vector access at an address that is not a multiple of its size
//     ends the run (aligned);
//   * a captured graph keeps its edges: its streams and events are replayed
//     as fresh ones forked from the launch stream (hipGraphLaunch), and work
//     issued during a capture on a stream outside it is reported.
// What it does not see: accesses wider than 16 bytes in one instruction (the
// instrumentation skips them).
#ifdef HIPEMU_RACE
namespace emu {
namespace race {
namespace {

constexpr uintptr_t DEV_SPAN = (uintptr_t)1 << 40;  // address space reserved for device memory
constexpr uintptr_t ALIGN = 1 << 16;
constexpr uint32_t ATOMIC_WG = 0xffffffffu;          // granule last written by an atomic
constexpr uint32_t MANY_WG = 0xfffffffeu;            // read by several workgroups of one op

struct Gr { uint32_t w, wwg, r[2], rwg[2]; };        // 24 bytes per 4-byte granule

uintptr_t g_base = 0;
Gr *g_shadow = nullptr;
uintptr_t g_top = 0;
std::map<uintptr_t, size_t> g_blocks;                 // device block -> reserved bytes
std::map<uintptr_t, size_t> g_req;                    // device block -> requested bytes

struct Op {
    int slot;
    uint32_t seq;
    std::vector<uint32_t> vc;
    const void *kfn;
    const char *what;
    int sid;
    // device-memory traffic of the op: bytes its accesses asked for (reads,
    // writes), and its footprint: the distinct 4-byte granules it read
    // (first access a read) and wrote without reading them first
    uint64_t rd = 0, wr = 0, fp_rd = 0, fp_wr = 0;
};
std::vector<Op> g_ops;                                // op id i + 1
uint32_t g_cur = 0;                                   // op running now (0: host code)
uint32_t g_wg = 0;
int g_nslots = 1;                                     // slot 0: the null stream
emu_stream g_null{0};
std::vector<uint32_t> g_host;                         // what the host has waited for
std::vector<emu_stream *> g_streams;
int g_drop_wait = 0;                                  // planted mistake: skip the n-th wait from now

long g_count = 0;
std::string g_report;
std::map<std::tuple<uint32_t, uint32_t, int>, int> g_seen;

// LDS: per 4-byte granule, the workgroup run ("generation") and barrier
// interval of the last write and its wave, and the waves that read it in the
// last interval it was read in
struct Ls {
    uint32_t wgen, wep, rgen, rep;
    uint16_t wwave, rmask;
};
uint32_t g_gen = 0;                                   // workgroups started so far
uint32_t g_epoch = 0;                                 // barriers opened in this workgroup
std::vector<Ls> g_lds_static, g_lds_dyn;

// Lanes of one wave between two of its synchronisation points (a wave
// operation or a barrier): the emulator runs them one after the other (lane
// 63 first), the device in lockstep, instruction by instruction.  When one
// lane writes a location another lane of the wave reads in that interval,
// the two agree if the read instruction comes first in program order and the
// reader runs first here, or the write comes first and the writer runs
// first.  The lanes' own sequences tell the order (control flow uniform
// across the wave): a writer that read the location before writing it ran
// the read instruction first.  So it is reported when
//   * a lane reads what another lane wrote after reading it itself (the read
//     instruction is the earlier one, and the device's reader sees the old
//     value), or
//   * a lane writes what another lane read without having read it itself
//     (the write is the earlier instruction, and the device's reader sees
//     the new value),
// and when two lanes write one location.  Per wave, a table of the granules
// touched in its current interval (tag), with the writing lane, whether it
// had read the granule first, and the reading lanes.
struct LaneRec {                                      // one 4-byte granule, per byte
    uintptr_t key;
    uint64_t tag;
    uint64_t rmask[4];
    int8_t wlane[4];
    bool wread[4];
};
constexpr int LANE_TAB = 1 << 15;
constexpr int MAX_WAVES = 16;
std::vector<LaneRec> g_lane_tab;                      // MAX_WAVES tables of LANE_TAB records
uint64_t g_wave_tag[MAX_WAVES];
uint64_t g_tag_next = 1;
long g_lane_overflow = 0;                             // granules not tracked (table full)

void new_interval(int w) { g_wave_tag[w] = g_tag_next++; }

// global_load_lds stores in flight: per lane, its DMA instructions in issue
// order (LDS address, bytes); per LDS granule, whether a load to it is in
// flight.  s_waitcnt vmcnt(n) of a lane retires its oldest loads until n
// remain (other vector memory instructions are not counted, so a load may
// be held in flight longer than on the device, never shorter); a barrier
// retires nothing.
std::vector<std::vector<std::pair<uintptr_t, size_t>>> g_dma(MAXT);
std::map<uintptr_t, int> g_dma_live;                  // granule -> loads in flight to it
size_t g_dma_lanes = 0;                               // lanes with loads in flight

void dma_reset()
{
    if (!g_dma_lanes) return;
    for (auto &q : g_dma) q.clear();
    g_dma_live.clear();
    g_dma_lanes = 0;
}

__attribute__((constructor)) void reserve()
{
    void *p = mmap(nullptr, DEV_SPAN, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    void *q = mmap(nullptr, DEV_SPAN / 4 * sizeof(Gr), PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (p == MAP_FAILED || q == MAP_FAILED) fatal("race checker: address space reservation failed");
    g_base = ((uintptr_t)p + ALIGN - 1) & ~(ALIGN - 1);
    g_shadow = (Gr *)q;
}

void join(std::vector<uint32_t> &a, const std::vector<uint32_t> &b)
{
    if (a.size() < b.size()) a.resize(b.size(), 0);
    for (size_t i = 0; i < b.size(); i++) a[i] = std::max(a[i], b[i]);
}

int slot_of(hipStream_t s) { return s ? s->id : 0; }
emu_stream *st(hipStream_t s) { return s ? s : &g_null; }

// does op a (1-based) happen before op b?
bool hb(uint32_t a, uint32_t b)
{
    const Op &A = g_ops[a - 1], &B = g_ops[b - 1];
    return (size_t)A.slot < B.vc.size() && B.vc[A.slot] >= A.seq;
}

std::string op_name(uint32_t id)
{
    const Op &o = g_ops[id - 1];
    char buf[64];
    snprintf(buf, sizeof buf, " (op %u, stream %d)", id, o.sid);
    // a kernel by its symbol (the expression at its launch when the symbol is
    // local to its file), a copy or fill by what it is
    std::string k = o.kfn ? kname(o.kfn) : std::string();
    int st = 0;
    if (k == "?") k = o.what;
    else if (char *d = abi::__cxa_demangle(k.c_str(), nullptr, nullptr, &st)) {
        k = d;
        free(d);
    }
    if (k.empty()) k = o.what;
    return k + buf;
}

std::string where(uintptr_t a)
{
    auto it = g_req.upper_bound(a);
    if (it == g_req.begin()) return "?";
    --it;
    char buf[96];
    snprintf(buf, sizeof buf, "byte %zu of a %zu-byte device block", (size_t)(a - it->first), it->second);
    return buf;
}

const char *g_report_what = nullptr;                   // kind 15: the operation

void report(int kind, uint32_t first, uint32_t second, uintptr_t a, uint32_t wg1, uint32_t wg2)
{
    static const char *const names[] = {"read after an unordered write", "write after an unordered write",
                                        "write after an unordered read", "host access to device memory",
                                        "read of another workgroup's write in one launch",
                                        "write over another workgroup's write in one launch",
                                        "write over another workgroup's read in one launch",
                                        "LDS read of another wave's write with no barrier between",
                                        "LDS write over another wave's write with no barrier between",
                                        "LDS write over another wave's read with no barrier between",
                                        "LDS read before the workgroup wrote it",
                                        "read of what another lane of the wave wrote, with no wave operation between",
                                        "write over what another lane of the wave read, with no wave operation between",
                                        "two lanes of one wave write one location, with no wave operation between",
                                        "LDS read of a global_load_lds destination before its s_waitcnt vmcnt",
                                        "operation issued during a stream capture on a stream outside it"};
    g_count++;
    auto key = std::make_tuple(first, second, kind);
    if (g_seen.count(key) || g_seen.size() >= 64) return;
    g_seen[key] = 1;
    const bool dev = a - g_base < DEV_SPAN;
    std::string m = std::string(names[kind]) + " at " +
                    (kind >= 7 && !dev ? std::string("an LDS byte") : where(a)) + ": ";
    if (kind == 15) {
        m = std::string(names[kind]) + ": " + (g_report_what ? g_report_what : "?");
    } else if (kind == 14) {
        char buf[96];
        snprintf(buf, sizeof buf, " workgroup %u, lane %d of the reading wave", g_wg, g_lane->lin);
        m += op_name(second) + buf;
    } else if (kind >= 11) {
        char buf[96];
        snprintf(buf, sizeof buf, " workgroup %u, wave %d, lanes %u and %u", g_wg, g_lane->lin >> 6, wg1, wg2);
        m += op_name(second) + buf;
    } else if (kind >= 7) {
        char buf[96];
        snprintf(buf, sizeof buf, " workgroup %u, waves %u and %u", g_wg, wg1, wg2);
        m += op_name(second) + buf;
    } else if (kind == 3) {
        m += "host code";
    } else if (kind >= 4) {
        char buf[64];
        snprintf(buf, sizeof buf, " workgroups %u and %u", wg1, wg2);
        m += op_name(second) + buf;
    } else {
        m += op_name(first) + " and " + op_name(second);
    }
    g_report += m + "\n";
}

extern "C" {
extern char __start_emu_lds[] __attribute__((weak, visibility("hidden")));
extern char __stop_emu_lds[] __attribute__((weak, visibility("hidden")));
}

// bytes [b0, b1] of the 4-byte granule at ga accessed by the running lane of a kernel
void lane_check(uintptr_t ga, int b0, int b1, int kind)
{
    const int lane = g_lane->lin & 63, w = g_lane->lin >> 6;
    if (w >= MAX_WAVES) return;
    if (g_lane_tab.empty()) g_lane_tab.assign((size_t)MAX_WAVES * LANE_TAB, LaneRec{});
    LaneRec *tab = g_lane_tab.data() + (size_t)w * LANE_TAB;
    const uint64_t tag = g_wave_tag[w];
    size_t h = (size_t)((ga >> 2) * 0x9E3779B97F4A7C15ull >> 49) & (LANE_TAB - 1);
    LaneRec *r = nullptr;
    for (int probe = 0; probe < 64; probe++, h = (h + 1) & (LANE_TAB - 1)) {
        LaneRec &c = tab[h];
        if (c.tag != tag) {                           // free in this interval
            c.key = ga;
            c.tag = tag;
            for (int k = 0; k < 4; k++) {
                c.rmask[k] = 0;
                c.wlane[k] = -1;
                c.wread[k] = false;
            }
            r = &c;
            break;
        }
        if (c.key == ga) {
            r = &c;
            break;
        }
    }
    if (!r) {
        g_lane_overflow++;
        return;
    }
    const uint64_t me = 1ull << lane;
    for (int k = b0; k <= b1; k++) {
        if (kind == READ) {
            if (r->wlane[k] >= 0 && r->wlane[k] != lane && r->wread[k])
                report(11, 0, g_cur, ga + k, (unsigned)r->wlane[k], (unsigned)lane);
            r->rmask[k] |= me;
        } else {
            if ((r->rmask[k] & ~me) && !(r->rmask[k] & me))
                report(12, 0, g_cur, ga + k, (unsigned)__builtin_ctzll(r->rmask[k] & ~me), (unsigned)lane);
            if (r->wlane[k] >= 0 && r->wlane[k] != lane && kind != ATOMIC)
                report(13, 0, g_cur, ga + k, (unsigned)r->wlane[k], (unsigned)lane);
            r->wlane[k] = (int8_t)lane;
            r->wread[k] = (r->rmask[k] & me) != 0;
        }
    }
}

// the bytes of [a, a + n) inside the granule at ga
inline void lane_bytes(uintptr_t a, size_t n, uintptr_t ga, int &b0, int &b1)
{
    b0 = a > ga ? (int)(a - ga) : 0;
    b1 = a + n < ga + 4 ? (int)(a + n - 1 - ga) : 3;
}

// an access of the running lane to LDS (a __shared__ array or the launch's
// dynamic block); anything else is not LDS
void lds_check(uintptr_t a, size_t n, int kind)
{
    Ls *sh;
    const uintptr_t s0 = (uintptr_t)__start_emu_lds, s1 = (uintptr_t)__stop_emu_lds;
    const uintptr_t d0 = (uintptr_t)g_dyn_lds;
    if (s0 && a >= s0 && a < s1) {
        if (g_lds_static.empty()) g_lds_static.assign((s1 - s0) / 4 + 1, Ls{});
        sh = g_lds_static.data() + (a - s0) / 4;
        n = std::min<size_t>(n, s1 - a);
    } else if (d0 && a >= d0 && a < d0 + g_dyn_bytes) {
        if (g_lds_dyn.size() < g_dyn_bytes / 4 + 1) g_lds_dyn.assign(160 * 1024 / 4 + 1, Ls{});
        sh = g_lds_dyn.data() + (a - d0) / 4;
        n = std::min<size_t>(n, d0 + g_dyn_bytes - a);
    } else {
        return;
    }
    const uint16_t w = (uint16_t)(g_lane->lin >> 6);
    const size_t cnt = (a + n - 1) / 4 - a / 4 + 1;
    for (size_t i = 0; i < cnt; i++) {
        Ls &l = sh[i];
        const uintptr_t ga = (a & ~(uintptr_t)3) + 4 * i;
        int b0, b1;
        lane_bytes(a, n, ga, b0, b1);
        lane_check(ga, b0, b1, kind);
        if (kind == READ) {
            if (!g_dma_live.empty() && g_dma_live.count(ga)) report(14, 0, g_cur, ga, 0, 0);
            if (l.wgen != g_gen) report(10, 0, g_cur, ga, w, w);
            else if (l.wep == g_epoch && l.wwave != w) report(7, 0, g_cur, ga, l.wwave, w);
            if (l.rgen == g_gen && l.rep == g_epoch) {
                l.rmask |= (uint16_t)(1u << w);
            } else {
                l.rgen = g_gen;
                l.rep = g_epoch;
                l.rmask = (uint16_t)(1u << w);
            }
        } else {
            if (l.wgen == g_gen && l.wep == g_epoch && l.wwave != w) report(8, 0, g_cur, ga, l.wwave, w);
            if (l.rgen == g_gen && l.rep == g_epoch && (l.rmask & ~(1u << w))) {
                unsigned o = l.rmask & ~(1u << w);
                report(9, 0, g_cur, ga, (unsigned)__builtin_ctz(o), w);
            }
            l.wgen = g_gen;
            l.wep = g_epoch;
            l.wwave = w;
        }
    }
}

void check(uintptr_t a, size_t n, int kind)
{
    if (n == 0) return;
    if (a - g_base >= DEV_SPAN) {
        if (g_cur && g_lane) lds_check(a, n, kind);
        return;
    }
    if (!g_cur) {
        report(3, 0, 0, a, 0, 0);
        return;
    }
    const uint32_t me = g_cur, wg = kind == ATOMIC ? ATOMIC_WG : g_wg;
    Op &op = g_ops[me - 1];
    (kind == READ ? op.rd : op.wr) += n;
    const size_t i0 = (a - g_base) / 4, i1 = (a + n - 1 - g_base) / 4;
    for (size_t i = i0; i <= i1; i++) {
        Gr &g = g_shadow[i];
        const uintptr_t ga = g_base + 4 * i;
        if (g.w != me && g.r[0] != me && g.r[1] != me) (kind == READ ? op.fp_rd : op.fp_wr) += 4;
        if (g_lane) {
            int b0, b1;
            lane_bytes(a, n, ga, b0, b1);
            lane_check(ga, b0, b1, kind);
        }
        if (kind == READ) {
            if (g.w && g.w != me && !hb(g.w, me)) report(0, g.w, me, ga, 0, 0);
            else if (g.w == me && g.wwg != wg && g.wwg != ATOMIC_WG) report(4, me, me, ga, g.wwg, wg);
            int k = -1;
            for (int j = 0; j < 2; j++)
                if (g.r[j] == me) k = j;
            if (k >= 0) {
                if (g.rwg[k] != wg) g.rwg[k] = MANY_WG;
                continue;
            }
            // readers that happen before this one are covered by it
            for (int j = 0; j < 2; j++)
                if (g.r[j] && hb(g.r[j], me)) g.r[j] = 0;
            int f = g.r[0] == 0 ? 0 : g.r[1] == 0 ? 1 : -1;
            if (f < 0) {     // keep the newer of the two
                g.r[0] = g.r[1];
                g.rwg[0] = g.rwg[1];
                f = 1;
            }
            g.r[f] = me;
            g.rwg[f] = wg;
        } else {
            if (g.w && g.w != me && !hb(g.w, me)) report(1, g.w, me, ga, 0, 0);
            else if (g.w == me && g.wwg != wg && g.wwg != ATOMIC_WG && wg != ATOMIC_WG)
                report(5, me, me, ga, g.wwg, wg);
            for (int j = 0; j < 2; j++) {
                if (!g.r[j]) continue;
                if (g.r[j] != me && !hb(g.r[j], me)) report(2, g.r[j], me, ga, 0, 0);
                else if (g.r[j] == me && g.rwg[j] != wg && wg != ATOMIC_WG) report(6, me, me, ga, g.rwg[j], wg);
            }
            g.w = me;
            g.wwg = wg;
            g.r[0] = g.r[1] = 0;
        }
    }
}

}  // namespace

void *dev_alloc(size_t bytes)
{
    const size_t span = (bytes + ALIGN - 1) & ~(ALIGN - 1);
    if (g_top + span + ALIGN > DEV_SPAN) return nullptr;
    const uintptr_t p = g_base + g_top;
    Gr *sh = g_shadow + (p - g_base) / 4;
    if (mprotect((void *)p, span, PROT_READ | PROT_WRITE) != 0 ||
        mprotect((void *)sh, span / 4 * sizeof(Gr), PROT_READ | PROT_WRITE) != 0)
        return nullptr;
    g_top += span + ALIGN;                            // an inaccessible gap after each block
    g_blocks[p] = span;
    g_req[p] = bytes;
    return (void *)p;
}

bool dev_free(void *vp)
{
    const uintptr_t p = (uintptr_t)vp;
    auto it = g_blocks.find(p);
    if (it == g_blocks.end()) return false;
    const size_t span = it->second;
    Gr *sh = g_shadow + (p - g_base) / 4;
    // addresses are never reused: a later access to the block faults
    madvise((void *)p, span, MADV_DONTNEED);
    mprotect((void *)p, span, PROT_NONE);
    madvise((void *)sh, span / 4 * sizeof(Gr), MADV_DONTNEED);
    mprotect((void *)sh, span / 4 * sizeof(Gr), PROT_NONE);
    g_blocks.erase(it);
    g_req.erase(p);
    return true;
}

void stream_created(hipStream_t s, bool blocking)
{
    s->blocking = blocking;
    g_nslots = std::max(g_nslots, s->id + 1);
    s->vc = g_host;                                   // work the host waited for precedes it
    g_streams.push_back(s);
}

void stream_destroyed(hipStream_t s)
{
    g_streams.erase(std::remove(g_streams.begin(), g_streams.end(), s), g_streams.end());
}

void begin_op(hipStream_t s, const void *kfn, const char *what)
{
    emu_stream *t = st(s);
    join(t->vc, g_host);
    if (!s) {
        for (emu_stream *o : g_streams)
            if (o->blocking) join(t->vc, o->vc);
    } else if (s->blocking) {
        join(t->vc, g_null.vc);
    }
    const int sl = slot_of(s);
    if ((int)t->vc.size() <= sl) t->vc.resize(sl + 1, 0);
    t->vc[sl]++;
    g_ops.push_back(Op{sl, t->vc[sl], t->vc, kfn, what, s ? s->id : 0});
    g_cur = (uint32_t)g_ops.size();
    g_wg = 0;
}

void end_op() { g_cur = 0; }

void set_workgroup(unsigned wg)
{
    g_wg = wg;
    g_gen++;
    g_epoch = 0;
    dma_reset();
    for (int w = 0; w < MAX_WAVES; w++) new_interval(w);
}

void barrier_opened()
{
    g_epoch++;
    for (int w = 0; w < MAX_WAVES; w++) new_interval(w);
}

void wave_synced(int wave)
{
    if (wave < MAX_WAVES) new_interval(wave);
}

void outside_capture(const void *kfn, const char *what)
{
    std::string k = kfn ? kname(kfn) : std::string(what);
    if (k == "?") k = what;
    static std::string held;
    held = k;
    g_report_what = held.c_str();
    report(15, 0, 0, 0, 0, 0);
}

void dma_issued(const void *lds, size_t n)
{
    if (!g_lane || !n) return;
    auto &q = g_dma[g_lane->lin];
    if (q.empty()) g_dma_lanes++;
    q.push_back({(uintptr_t)lds, n});
    const uintptr_t a = (uintptr_t)lds & ~(uintptr_t)3;
    for (uintptr_t ga = a; ga < (uintptr_t)lds + n; ga += 4) g_dma_live[ga]++;
}


void access(const void *p, size_t n, int kind) { check((uintptr_t)p, n, kind); }

void record(hipEvent_t e, hipStream_t s)
{
    emu_stream *t = st(s);
    join(t->vc, g_host);
    e->vc = t->vc;
}

bool drop_this_wait() { return g_drop_wait > 0 && --g_drop_wait == 0; }

bool wait(hipStream_t s, hipEvent_t e)
{
    emu_stream *t = st(s);
    join(t->vc, g_host);
    join(t->vc, e->vc);
    return true;
}

void host_sync_stream(hipStream_t s)
{
    join(g_host, st(s)->vc);
}

void host_sync_event(hipEvent_t e) { join(g_host, e->vc); }

void host_sync_device()
{
    join(g_host, g_null.vc);
    for (emu_stream *o : g_streams) join(g_host, o->vc);
}

}  // namespace race

void race_atomic(const void *p, size_t n) { race::check((uintptr_t)p, n, race::ATOMIC); }

namespace race {
// a kernel's access of n >= 4 bytes, taken as aligned by the compiler, to
// device memory or LDS: the address must be a multiple of n.  Such an access
// cannot run here either (an aligned vector instruction faults), so the run
// ends with the report.
void aligned(uintptr_t a, size_t n)
{
    if (n < 4 || !g_cur || !g_lane || (a & (n - 1)) == 0) return;
    const uintptr_t s0 = (uintptr_t)__start_emu_lds, s1 = (uintptr_t)__stop_emu_lds;
    const uintptr_t d0 = (uintptr_t)g_dyn_lds;
    const bool lds = (s0 && a >= s0 && a < s1) || (d0 && a >= d0 && a < d0 + g_dyn_bytes);
    if (a - g_base >= DEV_SPAN && !lds) return;
    fatal("vector access at an address not aligned to its size: %zu bytes at %s %p in %s, workgroup %u, lane %d",
          n, lds ? "LDS" : "device", (void *)a, op_name(g_cur).c_str(), g_wg, g_lane->lin);
}
}  // namespace race

// s_waitcnt of the running lane: vmcnt is bits [3:0] and [15:14] of the
// immediate (gfx9 encoding)
void waitcnt(int imm)
{
    if (!g_lane) return;
    const size_t vm = (size_t)((imm & 0xF) | (((imm >> 14) & 3) << 4));
    auto &q = race::g_dma[g_lane->lin];
    if (q.size() <= vm) return;
    const size_t drop = q.size() - vm;
    for (size_t i = 0; i < drop; i++) {
        const uintptr_t a = q[i].first & ~(uintptr_t)3;
        for (uintptr_t ga = a; ga < q[i].first + q[i].second; ga += 4) {
            auto it = race::g_dma_live.find(ga);
            if (it != race::g_dma_live.end() && --it->second == 0) race::g_dma_live.erase(it);
        }
    }
    q.erase(q.begin(), q.begin() + (long)drop);
    if (q.empty()) race::g_dma_lanes--;
}

}  // namespace emu

// the thread-sanitizer instrumentation's callbacks
extern "C" {
void __tsan_init() {}
void __tsan_func_entry(void *) {}
void __tsan_func_exit() {}
void __tsan_vptr_update(void **, void *) {}
void __tsan_vptr_read(void **) {}
// the aligned forms: the compiler took the address as aligned to the access
// size (at least 8 bytes for the 16-byte forms), so an address that is not
// is a misaligned vector access on the device
#define EMU_TSAN_RW(N)                                                                            \
    void __tsan_read##N(void *p)                                                                  \
    {                                                                                             \
        emu::race::aligned((uintptr_t)p, N);                                                      \
        emu::race::check((uintptr_t)p, N, emu::race::READ);                                       \
    }                                                                                             \
    void __tsan_write##N(void *p)                                                                 \
    {                                                                                             \
        emu::race::aligned((uintptr_t)p, N);                                                      \
        emu::race::check((uintptr_t)p, N, emu::race::WRITE);                                      \
    }                                                                                             \
    void __tsan_unaligned_read##N(void *p) { emu::race::check((uintptr_t)p, N, emu::race::READ); } \
    void __tsan_unaligned_write##N(void *p) { emu::race::check((uintptr_t)p, N, emu::race::WRITE); } \
    void __tsan_volatile_read##N(void *p) { emu::race::check((uintptr_t)p, N, emu::race::READ); }  \
    void __tsan_volatile_write##N(void *p) { emu::race::check((uintptr_t)p, N, emu::race::WRITE); }
EMU_TSAN_RW(1)
EMU_TSAN_RW(2)
EMU_TSAN_RW(4)
EMU_TSAN_RW(8)
EMU_TSAN_RW(16)
#undef EMU_TSAN_RW
void __tsan_read_range(void *p, unsigned long n) { emu::race::check((uintptr_t)p, n, emu::race::READ); }
void __tsan_write_range(void *p, unsigned long n) { emu::race::check((uintptr_t)p, n, emu::race::WRITE); }
void *__tsan_memcpy(void *d, const void *s, size_t n)
{
    emu::race::check((uintptr_t)s, n, emu::race::READ);
    emu::race::check((uintptr_t)d, n, emu::race::WRITE);
    return memcpy(d, s, n);
}
void *__tsan_memmove(void *d, const void *s, size_t n)
{
    emu::race::check((uintptr_t)s, n, emu::race::READ);
    emu::race::check((uintptr_t)d, n, emu::race::WRITE);
    return memmove(d, s, n);
}
void *__tsan_memset(void *d, int v, size_t n)
{
    emu::race::check((uintptr_t)d, n, emu::race::WRITE);
    return memset(d, v, n);
}
// host-side atomics of the library (reference counts, locks): performed as such
#define EMU_TSAN_ATOMIC(B, T)                                                                              \
    T __tsan_atomic##B##_load(const volatile T *a, int) { return __atomic_load_n(a, __ATOMIC_SEQ_CST); }   \
    void __tsan_atomic##B##_store(volatile T *a, T v, int) { __atomic_store_n(a, v, __ATOMIC_SEQ_CST); }   \
    T __tsan_atomic##B##_exchange(volatile T *a, T v, int) { return __atomic_exchange_n(a, v, __ATOMIC_SEQ_CST); } \
    T __tsan_atomic##B##_fetch_add(volatile T *a, T v, int) { return __atomic_fetch_add(a, v, __ATOMIC_SEQ_CST); } \
    T __tsan_atomic##B##_fetch_sub(volatile T *a, T v, int) { return __atomic_fetch_sub(a, v, __ATOMIC_SEQ_CST); } \
    T __tsan_atomic##B##_fetch_and(volatile T *a, T v, int) { return __atomic_fetch_and(a, v, __ATOMIC_SEQ_CST); } \
    T __tsan_atomic##B##_fetch_or(volatile T *a, T v, int) { return __atomic_fetch_or(a, v, __ATOMIC_SEQ_CST); }   \
    T __tsan_atomic##B##_fetch_xor(volatile T *a, T v, int) { return __atomic_fetch_xor(a, v, __ATOMIC_SEQ_CST); } \
    int __tsan_atomic##B##_compare_exchange_strong(volatile T *a, T *c, T v, int, int)                    \
    {                                                                                                      \
        return __atomic_compare_exchange_n(a, c, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);            \
    }                                                                                                      \
    int __tsan_atomic##B##_compare_exchange_weak(volatile T *a, T *c, T v, int, int)                      \
    {                                                                                                      \
        return __atomic_compare_exchange_n(a, c, v, true, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);             \
    }                                                                                                      \
    T __tsan_atomic##B##_compare_exchange_val(volatile T *a, T c, T v, int, int)                          \
    {                                                                                                      \
        __atomic_compare_exchange_n(a, &c, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);                  \
        return c;                                                                                          \
    }
EMU_TSAN_ATOMIC(8, unsigned char)
EMU_TSAN_ATOMIC(16, unsigned short)
EMU_TSAN_ATOMIC(32, unsigned int)
EMU_TSAN_ATOMIC(64, unsigned long long)
#undef EMU_TSAN_ATOMIC
void __tsan_atomic_thread_fence(int) { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
void __tsan_atomic_signal_fence(int) { __atomic_signal_fence(__ATOMIC_SEQ_CST); }

// test hooks: races found so far, their description, and a planted mistake
long hipemu_races(void) { return emu::race::g_count; }
long hipemu_race_report(char *buf, long cap)
{
    const std::string &r = emu::race::g_report;
    if (buf && cap > 0) {
        const size_t n = std::min((size_t)cap - 1, r.size());
        memcpy(buf, r.data(), n);
        buf[n] = 0;
    }
    return (long)r.size();
}
void hipemu_race_reset(void)
{
    emu::race::g_count = 0;
    emu::race::g_report.clear();
    emu::race::g_seen.clear();
}
// the n-th hipStreamWaitEvent from now is skipped (0: none)
void hipemu_drop_wait(int n) { emu::race::g_drop_wait = n; }
long hipemu_ops(void) { return (long)emu::race::g_ops.size(); }
// op i (1-based): its name and device-memory traffic (bytes read, written;
// footprint read, written), -1 for no such op
long hipemu_op_stats(long i, double *out, char *name, long cap)
{
    if (i < 1 || i > (long)emu::race::g_ops.size()) return -1;
    const auto &o = emu::race::g_ops[i - 1];
    out[0] = (double)o.rd;
    out[1] = (double)o.wr;
    out[2] = (double)o.fp_rd;
    out[3] = (double)o.fp_wr;
    if (name && cap > 0) snprintf(name, (size_t)cap, "%s", emu::race::op_name((uint32_t)i).c_str());
    return 0;
}
// granules the lane-order check could not track (a wave's table was full)
long hipemu_lane_overflow(void) { return emu::race::g_lane_overflow; }
}
#endif
