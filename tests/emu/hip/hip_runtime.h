// hip_runtime.h -- CPU emulation shim for the kernel and runtime sources of
// libitsca_hip.so (TEST INFRASTRUCTURE ONLY).
//
// Compiling the product's .hip files for the host with clang++ against this header runs every
// kernel on the CPU: one fiber per GPU thread, a cooperative barrier per
// workgroup for __syncthreads, a per-wavefront exchange for __shfl_up, and
// plain host memory for device buffers.  Workgroups run one after another, or
// with EMU_THREADS=n on n OS threads at once (EMU_ORDER=reverse deals them from
// the last one), so that work of different workgroups interleaves as it does
// on the GPU.  Built with AddressSanitizer, an
// out-of-bounds access in a kernel becomes an ASan report.  Used to check
// kernel logic on machines without a GPU; never a product path.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define amdgpu_waves_per_eu(...)
// LDS: one copy per OS thread, i.e. per workgroup in flight (EMU_THREADS > 1
// runs workgroups concurrently on several OS threads)
#define __shared__ static thread_local
#define __align__(n) __attribute__((aligned(n)))

struct dim3 {
  unsigned x = 1, y = 1, z = 1;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint3e {
  unsigned x = 0, y = 0, z = 0;
};

typedef short emu_bf16x4 __attribute__((ext_vector_type(4)));
typedef float emu_f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 emu_bf16x8 __attribute__((ext_vector_type(8)));

namespace emu {
extern thread_local uint3e tIdx, bIdx;
extern thread_local dim3 bDim, gDim;
void syncthreads();
int shflUp(int v, unsigned d, int width);
float shflXor(float v, int mask, int width);
int shflXorInt(int v, int mask, int width);
int shflInt(int v, int src, int width);
unsigned long long ballot(bool pred);
void waveBarrier();
// global_load_lds (LDS DMA): this lane's `size` bytes at src go to dst + offset + size * lane,
// dst being the wavefront's uniform LDS base
void globalLoadLds(const void* src, void* dst, unsigned size, int offset);
// v_mfma_f32_16x16x16_bf16 as documented for CDNA wave64: lane i holds
// A[i%16][4*(i/16)+r], B[4*(i/16)+r][i%16] and C/D[4*(i/16)+r][i%16].
emu_f32x4 mfma16x16x16bf16(emu_bf16x4 a, emu_bf16x4 b, emu_f32x4 c);
// v_mfma_f32_16x16x32_bf16 (gfx950): lane i holds A[i%16][8*(i/16)+r] and
// B[8*(i/16)+r][i%16], r = 0..7; C/D as for 16x16x16.
emu_f32x4 mfma16x16x32bf16(emu_bf16x8 a, emu_bf16x8 b, emu_f32x4 c);
void launch(dim3 grid, dim3 block, const std::function<void()>& body);
}  // namespace emu

#define threadIdx (emu::tIdx)
#define blockIdx (emu::bIdx)
#define blockDim (emu::bDim)
#define gridDim (emu::gDim)

inline void __syncthreads() { emu::syncthreads(); }
inline void __threadfence_block() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
inline int __shfl_up(int v, unsigned d, int width = 64) { return emu::shflUp(v, d, width); }
inline long long __shfl_up(long long v, unsigned d, int width = 64)
{
  const int lo = emu::shflUp((int)(unsigned)(unsigned long long)v, d, width);
  const int hi = emu::shflUp((int)((unsigned long long)v >> 32), d, width);
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
inline float __shfl_xor(float v, int mask, int width = 64) { return emu::shflXor(v, mask, width); }
inline int __shfl_xor(int v, int mask, int width = 64) { return emu::shflXorInt(v, mask, width); }
inline int __shfl(int v, int src, int width = 64) { return emu::shflInt(v, src, width); }
inline unsigned long long __ballot(int pred) { return emu::ballot(pred != 0); }
inline int __float_as_int(float f)
{
  int i;
  std::memcpy(&i, &f, 4);
  return i;
}
inline float __int_as_float(int i)
{
  float f;
  std::memcpy(&f, &i, 4);
  return f;
}
inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }

// Wave-level builtins used by the MFMA code.
#define __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, cbsz, abid, blgp) emu::mfma16x16x16bf16(a, b, c)
#define __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, cbsz, abid, blgp) emu::mfma16x16x32bf16(a, b, c)
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(__ATOMIC_SEQ_CST)
#define __builtin_amdgcn_wave_barrier() emu::waveBarrier()
#define __builtin_amdgcn_global_load_lds(src, dst, size, offset, aux) emu::globalLoadLds(src, dst, size, offset)
#define __builtin_amdgcn_s_waitcnt(imm) ((void)(imm))  // the emulated loads complete at once
inline float log1pf(float x) { return std::log1p(x); }
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int atomicMin(int* p, int v)
{
  int old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return old;
}
inline int atomicMax(int* p, int v)
{
  int old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (v > old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return old;
}

inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a > b ? a : b; }

struct float2 {
  float x, y;
};
struct float4 {
  float x, y, z, w;
};
struct int2 {
  int x, y;
};
struct int4 {
  int x, y, z, w;
};
struct uint2 {
  unsigned x, y;
};
struct uint4 {
  unsigned x, y, z, w;
};
inline float2 make_float2(float a, float b) { return float2{a, b}; }
inline int2 make_int2(int a, int b) { return int2{a, b}; }
inline int4 make_int4(int a, int b, int c, int d) { return int4{a, b, c, d}; }
inline uint2 make_uint2(unsigned a, unsigned b) { return uint2{a, b}; }
inline uint4 make_uint4(unsigned a, unsigned b, unsigned c, unsigned d) { return uint4{a, b, c, d}; }
inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }

// --- runtime API subset --------------------------------------------------
typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorInvalidDevice = 101 };
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
#define hipStreamNonBlocking 1
struct hipDeviceProp_t {
  int multiProcessorCount;
};
hipError_t hipMalloc(void** p, size_t bytes);
template <typename T>
hipError_t hipMalloc(T** p, size_t bytes)
{
  return hipMalloc(reinterpret_cast<void**>(p), bytes);
}
hipError_t hipFree(void* p);
#define hipHostMallocDefault 0u
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags);  // host memory (no pinning here)
hipError_t hipHostFree(void* p);
// Device symbols are host objects here: the "device address" is the object's own.
#define HIP_SYMBOL(sym) (&(sym))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym)
{
  *p = const_cast<void*>(sym);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t);
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind);
hipError_t hipStreamSynchronize(hipStream_t);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned);
hipError_t hipStreamDestroy(hipStream_t);
hipError_t hipGetDeviceCount(int* n);
hipError_t hipSetDevice(int);
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int);
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);

// Graph capture is not emulated: capture reports an error.
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorInvalidValue; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return hipErrorInvalidValue; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) { return hipErrorInvalidValue; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorInvalidValue; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
  emu::launch(dim3(grid), dim3(block), [&]() { kernel(__VA_ARGS__); })
