// hipemu.hpp -- CPU execution of the library's HIP sources, for tests only.
//
// The library's .hip files compile unchanged as host C++ against this header
// (reached through the stand-in <hip/hip_runtime.h> next to it).  A kernel
// launch runs every workgroup of the grid one after the other; the lanes of a
// workgroup are fibers on one OS thread, switched at every __syncthreads()
// and at every cross-lane operation, and between those points they run in
// descending lane order (lockstep code reads a location in every lane before
// lane 0 overwrites it; lane 0 running last keeps that order):
//   * __syncthreads / s_barrier: all lanes of the workgroup that have not
//     returned meet;
//   * wave operations (__shfl, __shfl_xor, readlane, DPP, MFMA): all live
//     lanes of the 64-lane wave deposit their operands, then each computes its
//     own result from the deposits;
//   * global_load_lds: the lane's 16 bytes land at the LDS address plus 16x
//     the lane index, at once (the race-checking build tracks them as in
//     flight until the lane's s_waitcnt vmcnt retires them);
//   * MFMA: the CDNA3/CDNA4 operand layouts -- lane l supplies row l % 16 of
//     A (column l % 16 of B) at K block l / 16 and holds column l % 16 of D,
//     rows 4 (l / 16) + v for FP32 accumulators, l / 16 + 4 v for FP64;
//     products summed in FP64 and rounded once to the accumulator type (FP64:
//     one fused multiply-add per K step).
// Streams execute synchronously in host issue order; hipStreamBeginCapture
// records the launches, copies and event edges issued until
// hipStreamEndCapture, and hipGraphLaunch replays them.  Dynamic LDS is a heap block of exactly the
// launch's size and hipMalloc returns blocks of exactly the requested size,
// so an AddressSanitizer build reports any access outside them.  The
// race-checking build (HIPEMU_RACE) instead checks every access to device
// memory against the happens-before order of the stream operations, which
// issue order is only one instance of, and reports LDS races between waves
// and lane-order dependences inside a wave -- the one place where the
// lane-by-lane execution above could compute what the device does not
// (hipemu.cpp, end of file).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <tuple>
#include <type_traits>
#include <utility>

// ------------------------------------------------------------------ qualifiers
#define __global__
#define __device__
#define __host__
#ifdef HIPEMU_RACE
// LDS arrays are gathered in one section, so that the race checker knows them
#define __shared__ static __attribute__((section("emu_lds")))
#else
#define __shared__ static
#endif
#define __constant__ static
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define __forceinline__ inline __attribute__((always_inline))

struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

// HIP's vector types as clang vector types, as on the device: one access of
// the whole vector (and the type's natural alignment), .x .y .z .w members
typedef float float2 __attribute__((ext_vector_type(2)));
typedef float float4 __attribute__((ext_vector_type(4)));
typedef double double2 __attribute__((ext_vector_type(2)));
typedef int int2 __attribute__((ext_vector_type(2)));
typedef int int4 __attribute__((ext_vector_type(4)));
typedef unsigned uint4 __attribute__((ext_vector_type(4)));
// (ushort4 stays a structure: the library binds references to its members)
struct alignas(8) ushort4 { unsigned short x, y, z, w; };
inline float2 make_float2(float x, float y) { return {x, y}; }
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
inline double2 make_double2(double x, double y) { return {x, y}; }
inline int2 make_int2(int x, int y) { return {x, y}; }
inline uint4 make_uint4(unsigned x, unsigned y, unsigned z, unsigned w) { return {x, y, z, w}; }
inline ushort4 make_ushort4(unsigned short x, unsigned short y, unsigned short z, unsigned short w)
{
    return {x, y, z, w};
}

// ------------------------------------------------------------------ runtime types
typedef enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotInitialized = 3,
    hipErrorInvalidConfiguration = 9,
    hipErrorInvalidDevice = 101,
    hipErrorInvalidResourceHandle = 400,
    hipErrorNotReady = 600,
    hipErrorLaunchFailure = 719,
    hipErrorStreamCaptureUnsupported = 900,
    hipErrorStreamCaptureInvalidated = 901,
    hipErrorUnknown = 999,
} hipError_t;

typedef enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4,
} hipMemcpyKind;

typedef enum hipStreamCaptureMode {
    hipStreamCaptureModeGlobal = 0,
    hipStreamCaptureModeThreadLocal = 1,
    hipStreamCaptureModeRelaxed = 2,
} hipStreamCaptureMode;

typedef enum hipFuncAttribute {
    hipFuncAttributeMaxDynamicSharedMemorySize = 8,
    hipFuncAttributePreferredSharedMemoryCarveout = 9,
} hipFuncAttribute;

typedef enum hipDeviceAttribute_t {
    hipDeviceAttributeMultiprocessorCount = 63,
    hipDeviceAttributeWallClockRate = 10015,
} hipDeviceAttribute_t;

#define hipStreamDefault 0x0
#define hipStreamNonBlocking 0x1
#define hipEventDefault 0x0
#define hipEventBlockingSync 0x1
#define hipEventDisableTiming 0x2

struct emu_stream;
struct emu_event;
struct emu_graph;
struct emu_graph_node;
typedef emu_stream *hipStream_t;
typedef emu_event *hipEvent_t;
typedef emu_graph *hipGraph_t;
typedef emu_graph *hipGraphExec_t;
typedef emu_graph_node *hipGraphNode_t;

struct hipDeviceProp_t {
    char name[256];
    size_t totalGlobalMem;
    size_t sharedMemPerBlock;
    int regsPerBlock;
    int warpSize;
    int maxThreadsPerBlock;
    int maxThreadsDim[3];
    int maxGridSize[3];
    int clockRate;
    int memoryClockRate;
    int memoryBusWidth;
    int multiProcessorCount;
    int major, minor;
    size_t maxSharedMemoryPerMultiProcessor;
    int l2CacheSize;
    char gcnArchName[256];
};

// ------------------------------------------------------------------ runtime API
hipError_t hipSetDevice(int);
hipError_t hipGetDevice(int *);
hipError_t hipGetDeviceCount(int *);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *, int);
hipError_t hipDeviceGetAttribute(int *, hipDeviceAttribute_t, int);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
hipError_t hipPeekAtLastError();
const char *hipGetErrorString(hipError_t);
hipError_t hipMalloc(void **, size_t);
template <class T>
inline hipError_t hipMalloc(T **p, size_t bytes) { return hipMalloc((void **)p, bytes); }
hipError_t hipFree(void *);
hipError_t hipMemcpy(void *, const void *, size_t, hipMemcpyKind);
hipError_t hipMemcpyAsync(void *, const void *, size_t, hipMemcpyKind, hipStream_t = nullptr);
hipError_t hipMemcpy2DAsync(void *, size_t, const void *, size_t, size_t, size_t, hipMemcpyKind,
                            hipStream_t = nullptr);
hipError_t hipMemset(void *, int, size_t);
hipError_t hipMemsetAsync(void *, int, size_t, hipStream_t = nullptr);
hipError_t hipStreamCreate(hipStream_t *);
hipError_t hipStreamCreateWithFlags(hipStream_t *, unsigned);
hipError_t hipExtStreamCreateWithCUMask(hipStream_t *, uint32_t, const uint32_t *);
hipError_t hipStreamDestroy(hipStream_t);
hipError_t hipStreamSynchronize(hipStream_t);
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned = 0);
hipError_t hipEventCreate(hipEvent_t *);
hipError_t hipEventCreateWithFlags(hipEvent_t *, unsigned);
hipError_t hipEventDestroy(hipEvent_t);
hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr);
hipError_t hipEventSynchronize(hipEvent_t);
hipError_t hipEventElapsedTime(float *, hipEvent_t, hipEvent_t);
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode);
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *);
hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, hipGraphNode_t *, char *, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t);
hipError_t hipGraphExecDestroy(hipGraphExec_t);
hipError_t hipGraphDestroy(hipGraph_t);
hipError_t hipGraphGetNodes(hipGraph_t, hipGraphNode_t *, size_t *);
hipError_t hipFuncSetAttribute(const void *, hipFuncAttribute, int);
hipError_t hipGetSymbolAddress(void **, const void *);
#define HIP_SYMBOL(x) ((const void *)&(x))
template <class F>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, F, int, size_t)
{
    *n = 1;
    return hipSuccess;
}

// ------------------------------------------------------------------ execution core
namespace emu {

struct Lane {
    dim3 tid;
    int lin;        // linear thread index in the workgroup
};
extern Lane *g_lane;                     // the lane running now
extern dim3 g_block, g_grid, g_bdim;
extern void *g_dyn_lds;                  // this launch's dynamic LDS block

void barrier();                          // __syncthreads
void wave_sync();                        // all live lanes of this lane's wave
unsigned char *slot(int lane);           // 128-byte exchange slot of a lane of this wave
[[noreturn]] void fatal(const char *fmt, ...);
unsigned long long clock_tick();

inline int lane_id() { return g_lane->lin & 63; }

// submit one launch (run now, or record while a capture is open)
// (name: the kernel expression as written at the launch, for reports)
void submit_launch(hipStream_t s, const void *kfn, const char *name, dim3 g, dim3 b, size_t shm,
                   std::function<void()> lane_fn);

// one value per lane in, the value of lane src out
template <class T>
inline T xchg(const T &v, int src)
{
    static_assert(sizeof(T) <= 128, "exchange slot too small");
    const int l = lane_id();
    memcpy(slot(l), &v, sizeof(T));
    wave_sync();
    T r;
    memcpy(&r, slot(src & 63), sizeof(T));
    wave_sync();
    return r;
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned short v8u16 __attribute__((ext_vector_type(8)));

v4d mfma_f64_16x16x4(double a, double b, v4d c);
v4f mfma_f32_16x16x4(float a, float b, v4f c);
v4f mfma_f32_16x16x32_bf16(v8u16 a, v8u16 b, v4f c);
v4f mfma_f32_16x16x32_fp8(long a, long b, v4f c);
int cvt_pk_fp8_f32(float a, float b, int old, bool hi);
// global_load_lds of one lane (under AddressSanitizer a source or destination
// outside its block is reported with the lane, workgroup and kernel)
void load_lds(void *lds_dst, const void *src, size_t n);
int update_dpp(int old, int src, int ctrl);

}  // namespace emu

template <class... P, class... A>
inline void emu_launch(const char *name, void (*k)(P...), dim3 g, dim3 b, size_t shm, hipStream_t s, A &&...a)
{
    auto args = std::make_shared<std::tuple<std::decay_t<P>...>>(static_cast<std::decay_t<P>>(a)...);
    emu::submit_launch(s, (const void *)k, name, g, b, shm, [k, args]() { std::apply(k, *args); });
}
#define hipLaunchKernelGGL(K, G, B, S, STR, ...) \
    emu_launch(#K, (K), dim3(G), dim3(B), (size_t)(S), (hipStream_t)(STR), ##__VA_ARGS__)

#define threadIdx (emu::g_lane->tid)
#define blockIdx (emu::g_block)
#define blockDim (emu::g_bdim)
#define gridDim (emu::g_grid)

// ------------------------------------------------------------------ device functions
inline void __syncthreads() { emu::barrier(); }
#define __builtin_amdgcn_s_barrier() emu::barrier()
#ifdef HIPEMU_RACE
namespace emu { void waitcnt(int imm); }
#define __builtin_amdgcn_s_waitcnt(n) emu::waitcnt(n)
#else
#define __builtin_amdgcn_s_waitcnt(n) ((void)(n))
#endif
#define __builtin_amdgcn_rsq(x) (1.0 / std::sqrt((double)(x)))

template <class T>
inline T __shfl(T v, int src, int width = 64)
{
    const int l = emu::lane_id();
    return emu::xchg(v, (l & ~(width - 1)) | (src & (width - 1)));
}
template <class T>
inline T __shfl_xor(T v, int mask, int width = 64)
{
    const int l = emu::lane_id();
    int s = l ^ mask;
    if ((s & ~(width - 1)) != (l & ~(width - 1))) s = l;
    return emu::xchg(v, s);
}
#define __builtin_amdgcn_readlane(v, l) emu::xchg((v), (l))
#define __builtin_amdgcn_readfirstlane(v) emu::xchg((v), 0)
#define __builtin_amdgcn_update_dpp(old, src, ctrl, rm, bm, bc) emu::update_dpp((old), (src), (ctrl))

#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) emu::mfma_f64_16x16x4((a), (b), (c))
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, x, y, z) emu::mfma_f32_16x16x4((a), (b), (c))
#define __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, x, y, z) \
    emu::mfma_f32_16x16x32_bf16(__builtin_bit_cast(emu::v8u16, (a)), __builtin_bit_cast(emu::v8u16, (b)), (c))
#define __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b, c, x, y, z) emu::mfma_f32_16x16x32_fp8((a), (b), (c))
#define __builtin_amdgcn_cvt_pk_fp8_f32(a, b, old, hi) emu::cvt_pk_fp8_f32((a), (b), (old), (hi))
// each lane's `size` bytes land at the (wave-uniform) LDS address + size x lane
#define __builtin_amdgcn_global_load_lds(src, dst, size, off, aux) \
    emu::load_lds((char *)(void *)(dst) + (off) + (size_t)(size) * emu::lane_id(), (const void *)(src), (size))

inline long long __double_as_longlong(double x) { return __builtin_bit_cast(long long, x); }
inline double __longlong_as_double(long long x) { return __builtin_bit_cast(double, x); }
inline unsigned __float_as_uint(float x) { return __builtin_bit_cast(unsigned, x); }
inline float __uint_as_float(unsigned x) { return __builtin_bit_cast(float, x); }
inline int __float_as_int(float x) { return __builtin_bit_cast(int, x); }
inline float __int_as_float(int x) { return __builtin_bit_cast(float, x); }
inline int __double2hiint(double x) { return (int)(__builtin_bit_cast(unsigned long long, x) >> 32); }
inline int __double2loint(double x) { return (int)(unsigned)__builtin_bit_cast(unsigned long long, x); }
inline double __hiloint2double(int hi, int lo)
{
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
inline unsigned long long wall_clock64() { return emu::clock_tick(); }
inline unsigned long long clock64() { return emu::clock_tick(); }

// one OS thread runs every lane, so plain read-modify-writes are atomic.  In
// the race-checking build (HIPEMU_RACE) each one is reported to the checker
// as one atomic access and its own loads and stores are left uninstrumented.
namespace emu {
#ifdef HIPEMU_RACE
void race_atomic(const void *p, size_t n);
#else
inline void race_atomic(const void *, size_t) {}
#endif
}  // namespace emu
#define EMU_ATOMIC __attribute__((no_sanitize("thread")))
template <class T>
EMU_ATOMIC inline T atomicAdd(T *p, T v) { emu::race_atomic(p, sizeof(T)); T o = *p; *p = o + v; return o; }
template <class T>
EMU_ATOMIC inline T atomicCAS(T *p, T cmp, T v)
{
    emu::race_atomic(p, sizeof(T));
    T o = *p;
    if (o == cmp) *p = v;
    return o;
}
template <class T>
EMU_ATOMIC inline T atomicMin(T *p, T v) { emu::race_atomic(p, sizeof(T)); T o = *p; if (v < o) *p = v; return o; }
template <class T>
EMU_ATOMIC inline T atomicMax(T *p, T v) { emu::race_atomic(p, sizeof(T)); T o = *p; if (v > o) *p = v; return o; }
template <class T>
EMU_ATOMIC inline T atomicExch(T *p, T v) { emu::race_atomic(p, sizeof(T)); T o = *p; *p = v; return o; }

#define EMU_MINMAX(T)                                  \
    inline T min(T a, T b) { return b < a ? b : a; } \
    inline T max(T a, T b) { return a < b ? b : a; }
EMU_MINMAX(int)
EMU_MINMAX(unsigned)
EMU_MINMAX(long)
EMU_MINMAX(unsigned long)
EMU_MINMAX(long long)
EMU_MINMAX(unsigned long long)
EMU_MINMAX(float)
EMU_MINMAX(double)
#undef EMU_MINMAX
inline double rsqrt(double x) { return 1.0 / std::sqrt(x); }
inline float rsqrtf(float x) { return 1.0f / std::sqrt(x); }
using std::fabs;
using std::fma;
using std::ldexp;
using std::sqrt;
