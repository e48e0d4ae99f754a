// emu_runtime.cpp -- the launch machinery and HIP API subset of the CPU
// emulation shim (TEST INFRASTRUCTURE ONLY; see hip/hip_runtime.h).
//
// A workgroup runs on one OS thread: every GPU thread is a fiber (ucontext)
// with its own stack, and the fibers are resumed round-robin.  A barrier
// (workgroup or wavefront) is a counter: a fiber that arrives before the last
// one yields until the barrier's phase advances.  Stack switches are announced
// to AddressSanitizer.  With EMU_THREADS=n (n > 1) a launch deals its
// workgroups to a pool of n OS threads, each with its own fibers and its own
// copy of the LDS (thread_local), so workgroups of one launch run at the same
// time and interleave their global-memory accesses; EMU_ORDER=reverse deals
// them from the last workgroup down.
#include <sanitizer/common_interface_defs.h>
#include <sys/mman.h>
#include <ucontext.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "hip/hip_runtime.h"

namespace emu {
thread_local uint3e tIdx, bIdx;
thread_local dim3 bDim, gDim;

namespace {
constexpr size_t kStack = 256 * 1024;  // per fiber
constexpr unsigned kMaxThreads = 1024;

struct Fiber {
  ucontext_t ctx;
  unsigned tid = 0;
  bool done = false;
  void* fake = nullptr;  // ASan fake-stack handle while switched out
};

struct Sched {
  ucontext_t ctx;
  const void* bottom = nullptr;  // the scheduler's stack, as ASan reports it
  size_t size = 0;
  void* fake = nullptr;
  char* stacks = nullptr;  // kMaxThreads * kStack, mapped once
  Fiber fibers[kMaxThreads];
  Fiber* cur = nullptr;
  const std::function<void()>* body = nullptr;
};
Sched* sched()
{
  thread_local Sched* s = [] {
    Sched* p = new Sched();
    void* m = mmap(nullptr, kMaxThreads * kStack, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE,
                   -1, 0);
    if (m == MAP_FAILED) {
      std::perror("emu: mmap");
      std::abort();
    }
    p->stacks = static_cast<char*>(m);
    return p;
  }();
  return s;
}

// Fiber -> scheduler.
void yield()
{
  Sched* s = sched();
  Fiber* f = s->cur;
  __sanitizer_start_switch_fiber(&f->fake, s->bottom, s->size);
  swapcontext(&f->ctx, &s->ctx);
  __sanitizer_finish_switch_fiber(f->fake, &s->bottom, &s->size);
}

void fiberEntry()
{
  Sched* s = sched();
  __sanitizer_finish_switch_fiber(nullptr, &s->bottom, &s->size);
  (*s->body)();
  s->cur->done = true;
  __sanitizer_start_switch_fiber(nullptr, s->bottom, s->size);  // this fiber's stack dies
}

struct Barrier {
  int n = 0, count = 0, phase = 0;
  void wait()
  {
    const int ph = phase;
    if (++count == n) {
      count = 0;
      ++phase;
      return;
    }
    while (phase == ph) yield();
  }
};

struct Block {
  Barrier all;
  std::vector<Barrier> waves;
  std::vector<int> shfl;
  std::vector<float> shflF;
  std::vector<emu_bf16x4> ma, mb;
  std::vector<emu_f32x4> mc;
  std::vector<emu_bf16x8> ma8, mb8;
};
thread_local Block* cur = nullptr;
}  // namespace

void syncthreads() { cur->all.wait(); }

int shflUp(int v, unsigned d, int width)
{
  const int lane = (int)(tIdx.x % 64), wave = (int)(tIdx.x / 64);
  Barrier& b = cur->waves[wave];
  cur->shfl[tIdx.x] = v;
  b.wait();
  const int src = lane - (int)d;
  const int r = (src >= 0 && lane / width == src / width) ? cur->shfl[wave * 64 + src] : v;
  b.wait();
  return r;
}

float shflXor(float v, int mask, int width)
{
  const int lane = (int)(tIdx.x % 64), wave = (int)(tIdx.x / 64);
  Barrier& b = cur->waves[wave];
  cur->shflF[tIdx.x] = v;
  b.wait();
  const int src = lane ^ mask;
  const float r = (src < 64 && lane / width == src / width) ? cur->shflF[wave * 64 + src] : v;
  b.wait();
  return r;
}

unsigned long long ballot(bool pred)
{
  const int wave = (int)(tIdx.x / 64);
  Barrier& b = cur->waves[wave];
  cur->shfl[tIdx.x] = pred ? 1 : 0;
  b.wait();
  unsigned long long m = 0;
  const int lanes = (int)std::min<unsigned>(64u, bDim.x - 64u * wave);
  for (int l = 0; l < lanes; ++l)
    if (cur->shfl[wave * 64 + l]) m |= 1ull << l;
  b.wait();
  return m;
}

int shflInt(int v, int srcLane, int width)
{
  const int lane = (int)(tIdx.x % 64), wave = (int)(tIdx.x / 64);
  Barrier& b = cur->waves[wave];
  cur->shfl[tIdx.x] = v;
  b.wait();
  const int src = (lane / width) * width + (srcLane % width);
  const int r = cur->shfl[wave * 64 + src];
  b.wait();
  return r;
}

int shflXorInt(int v, int mask, int width)
{
  const int lane = (int)(tIdx.x % 64), wave = (int)(tIdx.x / 64);
  Barrier& b = cur->waves[wave];
  cur->shfl[tIdx.x] = v;
  b.wait();
  const int src = lane ^ mask;
  const int r = (src < 64 && lane / width == src / width) ? cur->shfl[wave * 64 + src] : v;
  b.wait();
  return r;
}

void waveBarrier() { cur->waves[tIdx.x / 64].wait(); }

void globalLoadLds(const void* src, void* dst, unsigned size, int offset)
{
  std::memcpy(static_cast<char*>(dst) + offset + (size_t)size * (tIdx.x % 64), src, size);
}

emu_f32x4 mfma16x16x16bf16(emu_bf16x4 a, emu_bf16x4 b, emu_f32x4 c)
{
  const int lane = (int)(tIdx.x % 64), wave = (int)(tIdx.x / 64);
  Barrier& bar = cur->waves[wave];
  cur->ma[tIdx.x] = a;
  cur->mb[tIdx.x] = b;
  cur->mc[tIdx.x] = c;
  bar.wait();
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
  bar.wait();
  return d;
}

emu_f32x4 mfma16x16x32bf16(emu_bf16x8 a, emu_bf16x8 b, emu_f32x4 c)
{
  const int lane = (int)(tIdx.x % 64), wave = (int)(tIdx.x / 64);
  Barrier& bar = cur->waves[wave];
  cur->ma8[tIdx.x] = a;
  cur->mb8[tIdx.x] = b;
  bar.wait();
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
  bar.wait();
  return d;
}

namespace {
int emuThreads()
{
  static const int n = [] {
    const char* e = std::getenv("EMU_THREADS");
    const int v = e ? std::atoi(e) : 1;
    return v < 1 ? 1 : (v > 64 ? 64 : v);
  }();
  return n;
}

bool emuReverse()
{
  static const bool r = [] {
    const char* e = std::getenv("EMU_ORDER");
    return e && std::strcmp(e, "reverse") == 0;
  }();
  return r;
}

// Worker threads that run the workgroups of one launch at a time; they live
// until the process ends.
struct Pool {
  std::mutex m;
  std::condition_variable go, done;
  const std::function<void(unsigned)>* job = nullptr;
  unsigned blocks = 0;
  std::atomic<unsigned> next{0};
  int active = 0;
  unsigned long long gen = 0;
  explicit Pool(int n)
  {
    for (int i = 0; i < n; ++i) std::thread([this] { loop(); }).detach();
  }
  void loop()
  {
    unsigned long long seen = 0;
    for (;;) {
      const std::function<void(unsigned)>* j;
      unsigned nb;
      {
        std::unique_lock<std::mutex> lk(m);
        go.wait(lk, [&] { return gen != seen; });
        seen = gen;
        j = job;
        nb = blocks;
      }
      for (unsigned k = next++; k < nb; k = next++) (*j)(k);
      std::lock_guard<std::mutex> lk(m);
      if (--active == 0) done.notify_all();
    }
  }
  void run(unsigned nb, int workers, const std::function<void(unsigned)>& f)
  {
    std::unique_lock<std::mutex> lk(m);
    job = &f;
    blocks = nb;
    next = 0;
    active = workers;
    ++gen;
    go.notify_all();
    done.wait(lk, [&] { return active == 0; });
  }
};
}  // namespace

void launch(dim3 grid, dim3 block, const std::function<void()>& body)
{
  const unsigned nt = block.x;
  if (nt == 0 || nt > kMaxThreads) {
    std::fprintf(stderr, "emu: workgroup of %u threads\n", nt);
    std::abort();
  }
  const unsigned nb = grid.x * grid.y * grid.z;
  const bool reverse = emuReverse();
  auto runBlock = [&](unsigned k) {
    const unsigned b = reverse ? nb - 1 - k : k;
    Sched* s = sched();
    s->body = &body;
    Block blk;
    blk.all.n = (int)nt;
    const unsigned nw = (nt + 63) / 64;
    blk.waves.resize(nw);
    for (unsigned w = 0; w < nw; ++w) blk.waves[w].n = (int)std::min(64u, nt - 64 * w);
    blk.shfl.assign(nt, 0);
    blk.shflF.assign(nt, 0.0f);
    blk.ma.assign(nt, emu_bf16x4{0, 0, 0, 0});
    blk.mb.assign(nt, emu_bf16x4{0, 0, 0, 0});
    blk.mc.assign(nt, emu_f32x4{0, 0, 0, 0});
    blk.ma8.assign(nt, emu_bf16x8{});
    blk.mb8.assign(nt, emu_bf16x8{});
    cur = &blk;
    bIdx.x = b % grid.x;
    bIdx.y = (b / grid.x) % grid.y;
    bIdx.z = b / (grid.x * grid.y);
    bDim = block;
    gDim = grid;
    for (unsigned t = 0; t < nt; ++t) {
      Fiber& f = s->fibers[t];
      f.tid = t;
      f.done = false;
      f.fake = nullptr;
      getcontext(&f.ctx);
      f.ctx.uc_stack.ss_sp = s->stacks + (size_t)t * kStack;
      f.ctx.uc_stack.ss_size = kStack;
      f.ctx.uc_link = &s->ctx;
      makecontext(&f.ctx, fiberEntry, 0);
    }
    unsigned left = nt;
    while (left > 0) {
      for (unsigned t = 0; t < nt; ++t) {
        Fiber& f = s->fibers[t];
        if (f.done) continue;
        s->cur = &f;
        tIdx.x = t;
        tIdx.y = tIdx.z = 0;
        __sanitizer_start_switch_fiber(&s->fake, s->stacks + (size_t)t * kStack, kStack);
        swapcontext(&s->ctx, &f.ctx);
        __sanitizer_finish_switch_fiber(s->fake, nullptr, nullptr);
        if (f.done) --left;
      }
    }
    cur = nullptr;
  };
  const int workers = emuThreads();
  if (workers <= 1 || nb <= 1) {
    for (unsigned k = 0; k < nb; ++k) runBlock(k);
    return;
  }
  static Pool* pool = new Pool(workers);  // never destroyed: its threads live until the process ends
  pool->run(nb, workers, runBlock);
}
}  // namespace emu

static long gAllocCalls = 0;  // hipMalloc and hipFree calls, for tests of entry points that must not allocate
extern "C" long emu_alloc_calls() { return gAllocCalls; }

hipError_t hipMalloc(void** p, size_t bytes)
{
  ++gAllocCalls;
  *p = std::malloc(bytes ? bytes : 1);
  return *p ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipFree(void* p)
{
  ++gAllocCalls;
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned)
{
  *p = std::malloc(bytes ? bytes : 1);
  return *p ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipHostFree(void* p)
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
