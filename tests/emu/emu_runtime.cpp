// emu_runtime.cpp -- the launch machinery and HIP API subset of the CPU
// emulation shim (TEST INFRASTRUCTURE ONLY; see hip/hip_runtime.h).
#include <barrier>
#include <chrono>
#include <cstdlib>
#include <memory>
#include <thread>
#include <vector>

#include "hip/hip_runtime.h"

namespace emu {
thread_local uint3e tIdx, bIdx;
thread_local dim3 bDim, gDim;

namespace {
struct Block {
  std::unique_ptr<std::barrier<>> all;
  std::vector<std::unique_ptr<std::barrier<>>> waves;
  std::vector<int> shfl;
};
thread_local Block* cur = nullptr;
}  // namespace

void syncthreads() { cur->all->arrive_and_wait(); }

int shflUp(int v, unsigned d, int width)
{
  const int lane = (int)(tIdx.x % 64), wave = (int)(tIdx.x / 64);
  std::barrier<>& b = *cur->waves[wave];
  cur->shfl[tIdx.x] = v;
  b.arrive_and_wait();
  const int src = lane - (int)d;
  const int r = (src >= 0 && lane / width == src / width) ? cur->shfl[wave * 64 + src] : v;
  b.arrive_and_wait();
  return r;
}

void launch(dim3 grid, dim3 block, const std::function<void()>& body)
{
  const unsigned nt = block.x;
  for (unsigned b = 0; b < grid.x; ++b) {
    Block blk;
    blk.all = std::make_unique<std::barrier<>>(nt);
    const unsigned nw = (nt + 63) / 64;
    for (unsigned w = 0; w < nw; ++w) blk.waves.push_back(std::make_unique<std::barrier<>>(std::min(64u, nt - 64 * w)));
    blk.shfl.assign(nt, 0);
    std::vector<std::thread> th;
    th.reserve(nt);
    for (unsigned t = 0; t < nt; ++t)
      th.emplace_back([&, t, b]() {
        tIdx.x = t;
        bIdx.x = b;
        bDim = block;
        gDim = grid;
        cur = &blk;
        body();
      });
    for (auto& x : th) x.join();
  }
}
}  // namespace emu

hipError_t hipMalloc(void** p, size_t bytes)
{
  *p = std::malloc(bytes ? bytes : 1);
  return *p ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipFree(void* p)
{
  std::free(p);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t)
{
  std::memcpy(dst, src, bytes);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned)
{
  *s = reinterpret_cast<hipStream_t>(1);
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipGetDeviceCount(int* n)
{
  *n = 1;
  return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int)
{
  p->multiProcessorCount = 1;
  return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "emulated error"; }
hipError_t hipEventCreate(hipEvent_t* e)
{
  *e = new std::chrono::steady_clock::time_point();
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e)
{
  delete static_cast<std::chrono::steady_clock::time_point*>(e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t)
{
  *static_cast<std::chrono::steady_clock::time_point*>(e) = std::chrono::steady_clock::now();
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b)
{
  const auto& ta = *static_cast<std::chrono::steady_clock::time_point*>(a);
  const auto& tb = *static_cast<std::chrono::steady_clock::time_point*>(b);
  *ms = std::chrono::duration<float, std::milli>(tb - ta).count();
  return hipSuccess;
}
