// Host build of the tile kernels for tests/test_kernel_emulation.py: the
// HIP vocabulary the kernel sources use, mapped onto C++ on the CPU. A block
// runs as blockDim.x std::threads; __syncthreads is a std::barrier.
#pragma once
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstring>
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__
#define __align__(n) alignas(n)
struct emu_dim3
{
  unsigned x, y, z;
};
extern thread_local emu_dim3 threadIdx;
extern emu_dim3 blockIdx, blockDim, gridDim;
extern std::barrier<>* emu_bar;
inline void __syncthreads() { emu_bar->arrive_and_wait(); }
// ds_bpermute: lane i of a wave reads `val` of lane (addr / 4) mod 64 of the
// same wave; every thread of the block takes part, as in the kernels that
// call it (block-uniform control flow)
extern int emu_perm[1024];
inline int __builtin_amdgcn_ds_bpermute(int addr, int val)
{
  emu_perm[threadIdx.x] = val;
  __syncthreads();
  const int r = emu_perm[(threadIdx.x & ~63u) + ((unsigned)(addr >> 2) & 63u)];
  __syncthreads();
  return r;
}
// __shfl_down: lane i reads `v` of lane i + off of the same wave (its own
// value where that lane lies outside the wave); every thread of the block
// takes part.
extern unsigned long long emu_shfl[1024];
template <typename T>
inline T __shfl_down(T v, unsigned off)
{
  static_assert(sizeof(T) <= 8, "64-bit lanes");
  std::memcpy(&emu_shfl[threadIdx.x], &v, sizeof(T));
  __syncthreads();
  T r = v;
  if ((threadIdx.x & 63u) + off < 64u)
    std::memcpy(&r, &emu_shfl[threadIdx.x + off], sizeof(T));
  __syncthreads();
  return r;
}
// the hardware transcendentals (log2, sin/cos of revolutions, sqrt)
inline float __builtin_amdgcn_logf(float x) { return std::log2(x); }
inline float __builtin_amdgcn_sinf(float x) { return std::sin(6.283185307179586f * x); }
inline float __builtin_amdgcn_cosf(float x) { return std::cos(6.283185307179586f * x); }
inline float __builtin_amdgcn_sqrtf(float x) { return std::sqrt(x); }
inline void sincospi(double x, double* s, double* c)
{
  *s = std::sin(M_PI * x);
  *c = std::cos(M_PI * x);
}
typedef void* hipStream_t;
