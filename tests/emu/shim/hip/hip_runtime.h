// Host-execution shim for the library's kernel sources: HIP keywords map to
// plain C++, device intrinsics to the fiber scheduler in emu_core.cpp, and
// hipLaunchKernelGGL to emu::launch (every workgroup run on the host, its
// lanes as fibers).
#pragma once
#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __constant__
#define __align__(n) alignas(n)

typedef enum hipError_t {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorInvalidSymbol = 13,
  hipErrorInvalidDevice = 101,
  hipErrorLaunchFailure = 719
} hipError_t;
typedef enum hipDeviceAttribute_t {
  hipDeviceAttributeMultiprocessorCount = 63,
  hipDeviceAttributeMaxSharedMemoryPerBlock = 74
} hipDeviceAttribute_t;
typedef struct ihipStream_t *hipStream_t;
struct dim3 {
  uint32_t x, y, z;
  constexpr dim3(uint32_t x_ = 1, uint32_t y_ = 1, uint32_t z_ = 1) : x(x_), y(y_), z(z_) {}
};
struct alignas(8) int2 {
  int x, y;
};
inline int2 make_int2(int x, int y) { return int2{x, y}; }

extern "C" {
hipError_t hipGetLastError(void);
hipError_t hipGetDevice(int *);
hipError_t hipDeviceGetAttribute(int *, hipDeviceAttribute_t, int);
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *, const void *, int, size_t);
hipError_t hipMemsetAsync(void *, int, size_t, hipStream_t);
hipError_t hipMalloc(void **, size_t);
}
#define HIP_SYMBOL(x) (&(x))
#define hipGetSymbolAddress(pp, sym) ((*(pp) = (void *)(sym)), hipSuccess)

namespace emu {
struct U3 {
  unsigned x, y, z;
};
enum Op { kShfl = 1, kDown, kXor, kBallot, kAny, kAll, kWaveBarrier };
uint64_t wave_op(int op, uint64_t val, int arg, int width);
void block_barrier();
void *dyn_lds();
unsigned xcc();
// one launch: grid, block, dynamic LDS and the kernel body bound to its
// arguments (copied, as a launch copies them)
struct Job {
  dim3 grid, block;
  size_t shm;
  std::function<void()> body;
  std::string name;
};
void submit(Job *job, hipStream_t s);

template <class... P, class... A>
void launch(const char *name, void (*k)(P...), dim3 g, dim3 b, size_t shm, hipStream_t s,
            A &&...a) {
  static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
  auto args = std::make_shared<std::tuple<std::decay_t<P>...>>(std::forward<A>(a)...);
  Job *j = new Job{g, b, shm, [k, args]() { std::apply(k, *args); }, name};
  submit(j, s);
}
template <class T>
inline uint64_t bits(T v) {
  static_assert(sizeof(T) <= 8, "wave op on a value wider than 8 bytes");
  uint64_t r = 0;
  std::memcpy(&r, &v, sizeof(T));
  return r;
}
template <class T>
inline T unbits(uint64_t r) {
  T v;
  std::memcpy(&v, &r, sizeof(T));
  return v;
}
}  // namespace emu

extern emu::U3 threadIdx, blockIdx, blockDim, gridDim;

#define hipLaunchKernelGGL(k, g, b, shm, s, ...) \
  emu::launch(#k, k, dim3(g), dim3(b), (size_t)(shm), (s), ##__VA_ARGS__)

template <class T>
inline T __shfl(T v, int src, int width = 64) {
  return emu::unbits<T>(emu::wave_op(emu::kShfl, emu::bits(v), src, width));
}
template <class T>
inline T __shfl_down(T v, unsigned delta, int width = 64) {
  return emu::unbits<T>(emu::wave_op(emu::kDown, emu::bits(v), int(delta), width));
}
template <class T>
inline T __shfl_xor(T v, int mask, int width = 64) {
  return emu::unbits<T>(emu::wave_op(emu::kXor, emu::bits(v), mask, width));
}
inline unsigned long long __ballot(int p) { return emu::wave_op(emu::kBallot, p != 0, 0, 64); }
inline int __any(int p) { return int(emu::wave_op(emu::kAny, p != 0, 0, 64)); }
inline int __all(int p) { return int(emu::wave_op(emu::kAll, p != 0, 0, 64)); }
inline void __syncthreads() { emu::block_barrier(); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline double rsqrt(double x) { return 1.0 / std::sqrt(x); }
inline float rsqrt(float x) { return 1.0f / std::sqrt(x); }
#define __builtin_amdgcn_s_getreg(x) (emu::xcc())
#define __builtin_amdgcn_fence(...) ((void)0)
#define __builtin_amdgcn_wave_barrier() ((void)emu::wave_op(emu::kWaveBarrier, 0, 0, 64))

// atomics: lanes are fibers of one host thread, so a plain update is atomic
template <class T>
inline T atomicAdd(T *p, T v) {
  T o = *p;
  *p = o + v;
  return o;
}
inline unsigned atomicAdd(unsigned *p, int v) { return atomicAdd(p, unsigned(v)); }
template <class T>
inline T atomicOr(T *p, T v) {
  T o = *p;
  *p = o | v;
  return o;
}
template <class T>
inline T atomicMax(T *p, T v) {
  T o = *p;
  *p = o > v ? o : v;
  return o;
}
template <class T>
inline T atomicMin(T *p, T v) {
  T o = *p;
  *p = o < v ? o : v;
  return o;
}
template <class T>
inline T atomicExch(T *p, T v) {
  T o = *p;
  *p = v;
  return o;
}
template <class T>
inline T atomicCAS(T *p, T c, T v) {
  T o = *p;
  if (o == c) *p = v;
  return o;
}
