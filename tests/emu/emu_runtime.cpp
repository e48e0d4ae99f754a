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
  std::vector<float> shflF;
  std::vector<emu_bf16x4> ma, mb;
  std::vector<emu_f32x4> mc;
  std::vector<emu_bf16x8> ma8, mb8;
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

float shflXor(float v, int mask, int width)
{
  const int lane = (int)(tIdx.x % 64), wave = (int)(tIdx.x / 64);
  std::barrier<>& b = *cur->waves[wave];
  cur->shflF[tIdx.x] = v;
  b.arrive_and_wait();
  const int src = lane ^ mask;
  const float r = (src < 64 && lane / width == src / width) ? cur->shflF[wave * 64 + src] : v;
  b.arrive_and_wait();
  return r;
}

unsigned long long ballot(bool pred)
{
  const int wave = (int)(tIdx.x / 64);
  std::barrier<>& b = *cur->waves[wave];
  cur->shfl[tIdx.x] = pred ? 1 : 0;
  b.arrive_and_wait();
  unsigned long long m = 0;
  const int lanes = (int)std::min<unsigned>(64u, bDim.x - 64u * wave);
  for (int l = 0; l < lanes; ++l)
    if (cur->shfl[wave * 64 + l]) m |= 1ull << l;
  b.arrive_and_wait();
  return m;
}

int shflInt(int v, int srcLane, int width)
{
  const int lane = (int)(tIdx.x % 64), wave = (int)(tIdx.x / 64);
  std::barrier<>& b = *cur->waves[wave];
  cur->shfl[tIdx.x] = v;
  b.arrive_and_wait();
  const int src = (lane / width) * width + (srcLane % width);
  const int r = cur->shfl[wave * 64 + src];
  b.arrive_and_wait();
  return r;
}

int shflXorInt(int v, int mask, int width)
{
  const int lane = (int)(tIdx.x % 64), wave = (int)(tIdx.x / 64);
  std::barrier<>& b = *cur->waves[wave];
  cur->shfl[tIdx.x] = v;
  b.arrive_and_wait();
  const int src = lane ^ mask;
  const int r = (src < 64 && lane / width == src / width) ? cur->shfl[wave * 64 + src] : v;
  b.arrive_and_wait();
  return r;
}

void waveBarrier() { cur->waves[tIdx.x / 64]->arrive_and_wait(); }

emu_f32x4 mfma16x16x16bf16(emu_bf16x4 a, emu_bf16x4 b, emu_f32x4 c)
{
  const int lane = (int)(tIdx.x % 64), wave = (int)(tIdx.x / 64);
  std::barrier<>& bar = *cur->waves[wave];
  cur->ma[tIdx.x] = a;
  cur->mb[tIdx.x] = b;
  cur->mc[tIdx.x] = c;
  bar.arrive_and_wait();
  auto bf = [](short h) {
    const uint32_t u = (uint32_t)(uint16_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
  };
  // A[m][k]: lane m%16 + 16*(k/4), element k%4; B[k][n]: lane n + 16*(k/4), element k%4
  emu_f32x4 d;
  const int col = lane % 16, grp = lane / 16;
  for (int r = 0; r < 4; ++r) {
    const int row = grp * 4 + r;
    float acc = c[r];
    for (int k = 0; k < 16; ++k)
      acc += bf(cur->ma[wave * 64 + row + 16 * (k / 4)][k % 4]) * bf(cur->mb[wave * 64 + col + 16 * (k / 4)][k % 4]);
    d[r] = acc;
  }
  bar.arrive_and_wait();
  return d;
}

emu_f32x4 mfma16x16x32bf16(emu_bf16x8 a, emu_bf16x8 b, emu_f32x4 c)
{
  const int lane = (int)(tIdx.x % 64), wave = (int)(tIdx.x / 64);
  std::barrier<>& bar = *cur->waves[wave];
  cur->ma8[tIdx.x] = a;
  cur->mb8[tIdx.x] = b;
  bar.arrive_and_wait();
  // A[m][k]: lane m + 16*(k/8), element k%8; B[k][n]: lane n + 16*(k/8), element k%8
  emu_f32x4 d;
  const int col = lane % 16, grp = lane / 16;
  for (int r = 0; r < 4; ++r) {
    const int row = grp * 4 + r;
    float acc = c[r];
    for (int k = 0; k < 32; ++k)
      acc += (float)cur->ma8[wave * 64 + row + 16 * (k / 8)][k % 8] * (float)cur->mb8[wave * 64 + col + 16 * (k / 8)][k % 8];
    d[r] = acc;
  }
  bar.arrive_and_wait();
  return d;
}

void launch(dim3 grid, dim3 block, const std::function<void()>& body)
{
  const unsigned nt = block.x;
  for (unsigned b = 0; b < grid.x * grid.y * grid.z; ++b) {
    Block blk;
    blk.all = std::make_unique<std::barrier<>>(nt);
    const unsigned nw = (nt + 63) / 64;
    for (unsigned w = 0; w < nw; ++w) blk.waves.push_back(std::make_unique<std::barrier<>>(std::min(64u, nt - 64 * w)));
    blk.shfl.assign(nt, 0);
    blk.shflF.assign(nt, 0.0f);
    blk.ma.assign(nt, emu_bf16x4{0, 0, 0, 0});
    blk.mb.assign(nt, emu_bf16x4{0, 0, 0, 0});
    blk.mc.assign(nt, emu_f32x4{0, 0, 0, 0});
    blk.ma8.assign(nt, emu_bf16x8{});
    blk.mb8.assign(nt, emu_bf16x8{});
    std::vector<std::thread> th;
    th.reserve(nt);
    for (unsigned t = 0; t < nt; ++t)
      th.emplace_back([&, t, b]() {
        tIdx.x = t;
        bIdx.x = b % grid.x;
        bIdx.y = (b / grid.x) % grid.y;
        bIdx.z = b / (grid.x * grid.y);
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
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind)
{
  std::memcpy(dst, src, bytes);
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
