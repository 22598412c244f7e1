// Host emulation of the HIP subset the rasterizer kernels use, for running
// the unmodified kernel sources on the CPU (test infrastructure only).
//
// Every work-item of a workgroup is a fiber on the calling host thread;
// workgroups run one after another.  Fibers switch only at barriers, so
// cross-lane operations (ballot, readlane, ds_bpermute, shfl_xor) exchange
// values through a per-wave slot array between two barrier phases of that
// wave; __builtin_amdgcn_wave_barrier is a barrier of the wave and
// __syncthreads one of the workgroup, so the ordering the kernels rely on (a
// wave's LDS operations in program order) holds wherever the sources mark
// it.  A work-item that returns drops out of both barriers and reads as an
// inactive lane.
//
// __shared__ becomes function-scope static thread-local storage: a host
// thread runs one workgroup at a time (several threads with EMU_THREADS).
// Device pointers are host pointers, so AddressSanitizer sees every
// out-of-bounds access of a kernel.
#pragma once

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <functional>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>
#include <type_traits>
#include <utility>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static thread_local

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_uint3 {
  unsigned x, y, z;
};

struct alignas(8) float2 { float x, y; };
struct float3 { float x, y, z; };
struct alignas(16) float4 { float x, y, z, w; };
struct alignas(8) int2 { int x, y; };
struct alignas(8) uint2 { unsigned x, y; };
struct alignas(16) uint4 { unsigned x, y, z, w; };
struct alignas(16) int4 { int x, y, z, w; };
inline float2 make_float2(float a, float b) { return {a, b}; }
inline float3 make_float3(float a, float b, float c) { return {a, b, c}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline int2 make_int2(int a, int b) { return {a, b}; }
inline int4 make_int4(int a, int b, int c, int d) { return {a, b, c, d}; }
inline uint2 make_uint2(unsigned a, unsigned b) { return {a, b}; }

namespace emu {

constexpr int kMaxThreads = 1024;

struct Ctx {
  emu_uint3 tid, bid;
  dim3 bdim, gdim;
  int lane, wave;
};
inline thread_local Ctx ctx;

// A barrier among the fibers of a workgroup (or of one wave).
struct Bar {
  int expected = 0;
  int arrived = 0;
  uint64_t gen = 0;
};
void bar_wait(Bar* b);  // blocks the calling fiber until every member arrived
void pause();           // lets other host threads run (a polling wave)

struct BlockSync {
  Bar block;
  Bar wave[kMaxThreads / 64];
  uint64_t slot[kMaxThreads / 64][64];
  bool alive[kMaxThreads / 64][64];
};
BlockSync* sync_state();

// Runs fn on every work-item of the grid, workgroup by workgroup.
void run_grid(dim3 grid, dim3 block, const std::function<void()>& fn);

inline uint64_t exchange(uint64_t v, int src_lane) {
  BlockSync& S = *sync_state();
  const int w = ctx.wave;
  S.slot[w][ctx.lane] = v;
  bar_wait(&S.wave[w]);
  const uint64_t r = (src_lane >= 0 && src_lane < 64 && S.alive[w][src_lane]) ? S.slot[w][src_lane]
                                                                               : 0;
  bar_wait(&S.wave[w]);
  return r;
}

}  // namespace emu

#define threadIdx (::emu::ctx.tid)
#define blockIdx (::emu::ctx.bid)
#define blockDim (::emu::ctx.bdim)
#define gridDim (::emu::ctx.gdim)

inline void __syncthreads() { emu::bar_wait(&emu::sync_state()->block); }
inline void __threadfence_block() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __builtin_amdgcn_wave_barrier() {
  emu::bar_wait(&emu::sync_state()->wave[emu::ctx.wave]);
}

inline unsigned long long __ballot(int p) {
  emu::BlockSync& S = *emu::sync_state();
  const int w = emu::ctx.wave;
  S.slot[w][emu::ctx.lane] = p ? 1u : 0u;
  emu::bar_wait(&S.wave[w]);
  unsigned long long m = 0;
  for (int i = 0; i < 64; ++i)
    if (S.alive[w][i] && S.slot[w][i]) m |= 1ull << i;
  emu::bar_wait(&S.wave[w]);
  return m;
}
inline int __builtin_amdgcn_readlane(int v, int l) {
  return (int)(uint32_t)emu::exchange((uint32_t)v, l);
}
inline int __builtin_amdgcn_ds_bpermute(int addr, int v) {
  return (int)(uint32_t)emu::exchange((uint32_t)v, (addr >> 2) & 63);
}
inline int __shfl_xor(int v, int m) {
  return (int)(uint32_t)emu::exchange((uint32_t)v, emu::ctx.lane ^ m);
}
inline unsigned __builtin_amdgcn_mbcnt_lo(unsigned mask, unsigned acc) {
  const int l = emu::ctx.lane < 32 ? emu::ctx.lane : 32;
  const unsigned below = l >= 32 ? 0xffffffffu : ((1u << l) - 1u);
  return acc + (unsigned)__builtin_popcount(mask & below);
}
inline unsigned __builtin_amdgcn_mbcnt_hi(unsigned mask, unsigned acc) {
  const int l = emu::ctx.lane - 32;
  const unsigned below = l <= 0 ? 0u : ((1u << l) - 1u);
  return acc + (unsigned)__builtin_popcount(mask & below);
}
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
inline int __popcll(unsigned long long m) { return __builtin_popcountll(m); }
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }
inline float __builtin_amdgcn_exp2f(float x) { return exp2f(x); }
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }

inline float __uint_as_float(unsigned u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
inline unsigned __float_as_uint(float f) {
  unsigned u;
  memcpy(&u, &f, 4);
  return u;
}
inline float __int_as_float(int u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
inline int __float_as_int(float f) {
  int u;
  memcpy(&u, &f, 4);
  return u;
}

template <typename T>
inline T atomicAdd(T* p, T v) {
  return std::atomic_ref<T>(*p).fetch_add(v);
}
template <typename T>
inline T atomicExch(T* p, T v) {
  return std::atomic_ref<T>(*p).exchange(v);
}
template <typename T>
inline T atomicOr(T* p, T v) {
  return std::atomic_ref<T>(*p).fetch_or(v);
}
template <typename T>
inline T atomicAnd(T* p, T v) {
  return std::atomic_ref<T>(*p).fetch_and(v);
}
template <typename T>
inline T atomicMax(T* p, T v) {
  std::atomic_ref<T> a(*p);
  T cur = a.load();
  while (cur < v && !a.compare_exchange_weak(cur, v)) {
  }
  return cur;
}

// Scoped atomics, sequentially consistent on the host (stronger than the
// orders the kernels ask for).
#define __HIP_MEMORY_SCOPE_AGENT 3
template <typename T>
inline T __hip_atomic_load(const T* p, int, int) {
  return std::atomic_ref<T>(*const_cast<T*>(p)).load();
}
template <typename T, typename V>
inline void __hip_atomic_store(T* p, V v, int, int) {
  std::atomic_ref<T>(*p).store((T)v);
}
inline void __builtin_amdgcn_s_sleep(int) { emu::pause(); }
inline void __builtin_amdgcn_sched_barrier(int) {}

// ---------------------------------------------------------------- runtime
typedef struct ihipStream_t* hipStream_t;
typedef struct ihipEvent_t* hipEvent_t;
typedef struct emu_graph* hipGraph_t;
typedef struct emu_graph* hipGraphExec_t;
typedef void* hipGraphNode_t;
typedef enum {
  hipStreamCaptureModeGlobal = 0,
  hipStreamCaptureModeThreadLocal = 1,
  hipStreamCaptureModeRelaxed = 2
} hipStreamCaptureMode;
enum { hipStreamDefault = 0, hipStreamNonBlocking = 1 };
typedef enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorNotReady = 600 } hipError_t;
typedef enum {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
} hipMemcpyKind;
enum { hipHostMallocDefault = 0, hipEventDisableTiming = 2 };
typedef enum { hipDeviceAttributeMultiprocessorCount = 0 } hipDeviceAttribute_t;

// Graph capture: launches on a capturing stream are recorded (kernel,
// launch shape and a copy of the arguments) instead of run; a launched
// graph runs them in order.  Copies and event records on a capturing
// stream are refused, so a captured path that would record them fails its
// tests here.
struct emu_graph {
  std::vector<std::function<void()>> steps;
};
namespace emu {
inline std::mutex& capture_mu() {
  static std::mutex m;
  return m;
}
inline std::map<hipStream_t, emu_graph*>& captures() {
  static std::map<hipStream_t, emu_graph*> m;
  return m;
}
inline emu_graph* capturing(hipStream_t s) {
  std::lock_guard<std::mutex> g(capture_mu());
  auto it = captures().find(s);
  return it == captures().end() ? nullptr : it->second;
}
}  // namespace emu
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = reinterpret_cast<hipStream_t>(new char[1]);
  return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t s) {
  delete[] reinterpret_cast<char*>(s);
  return hipSuccess;
}
inline hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (s == nullptr || emu::capturing(s)) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> g(emu::capture_mu());
  emu::captures()[s] = new emu_graph();
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  emu_graph* c = emu::capturing(s);
  if (!c) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> l(emu::capture_mu());
  emu::captures().erase(s);
  *g = c;
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*,
                                      size_t) {
  *e = new emu_graph(*g);
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
  if (emu::capturing(s)) return hipErrorInvalidValue;
  for (auto& f : e->steps) f();
  return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t st) {
  if (emu::capturing(st)) return hipErrorInvalidValue;
  memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipHostMalloc(void** p, size_t n, unsigned) {
  *p = malloc(n);
  return *p ? hipSuccess : hipErrorInvalidValue;
}
inline hipError_t hipHostFree(void* p) {
  free(p);
  return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) {
  *e = reinterpret_cast<hipEvent_t>(new char[1]);
  return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e) {
  delete[] reinterpret_cast<char*>(e);
  return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t s) {
  return emu::capturing(s) ? hipErrorInvalidValue : hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }  // launches run to completion
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "emulated"; }
inline hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
// The emulated device has few compute units by default, so persistent grids
// stay small; EMU_CUS and EMU_OCCUPANCY set them to a real device's values
// (256 and the kernel's waves per CU on MI355X) to run the launch shapes the
// GPU runs.
inline int emu_env_int(const char* name, int dflt) {
  const char* e = getenv(name);
  return (e && *e) ? atoi(e) : dflt;
}
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
  *v = emu_env_int("EMU_CUS", 4);
  return hipSuccess;
}
// Device symbols are host statics here: their address is the symbol itself.
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
  *p = const_cast<void*>(symbol);
  return hipSuccess;
}
template <typename K>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, K, int, size_t) {
  *n = emu_env_int("EMU_OCCUPANCY", 2);
  return hipSuccess;
}

template <typename... P, typename... A>
inline void emu_launch(hipStream_t stream, void (*kernel)(P...), dim3 grid, dim3 block,
                       A&&... args) {
  std::tuple<std::decay_t<P>...> t(static_cast<std::decay_t<P>>(args)...);
  if (emu_graph* c = emu::capturing(stream)) {
    c->steps.push_back([kernel, grid, block, t]() {
      emu::run_grid(grid, block, [&]() { std::apply(kernel, t); });
    });
    return;
  }
  emu::run_grid(grid, block, [&]() { std::apply(kernel, t); });
}
#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
  emu_launch(stream, kernel, dim3(grid), dim3(block), ##__VA_ARGS__)
