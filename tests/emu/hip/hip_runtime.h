// hip_runtime.h (CPU emulation) - runs the solver's HIP kernels on the host
// for AddressSanitizer / bounds checking of their indexing. TEST
// INFRASTRUCTURE ONLY: the emulated library (tests/emu/libcfdgpu_emu.so) is
// built and loaded by tests/ alone; the product loads libcfdgpu.so and has no
// CPU path.
//
// Model: blocks run one after another; the lanes of a block are worker
// threads; __syncthreads / s_barrier are block-wide std::barrier waits and
// __shfl_down exchanges through a per-wave buffer between two wave barriers.
// Lanes that return early drop out of both barriers, as terminated waves do
// on the hardware. Static __shared__ arrays become function statics (one
// block at a time, so one copy suffices). Memory calls are malloc / memcpy.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <barrier>
#include <condition_variable>
#include <mutex>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
using std::isfinite;
// HIP device code has integer min / max overloads in the global namespace.
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a < b ? b : a; }

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_uint3 {
  unsigned x{0}, y{0}, z{0};
};
struct double2 {
  double x, y;
};
inline double2 make_double2(double a, double b) { return double2{a, b}; }

namespace cfdemu {
struct BlockCtx {
  unsigned threads{0};
  std::unique_ptr<std::barrier<>> block;
  std::vector<std::unique_ptr<std::barrier<>>> waves;
  std::vector<double> xchg;
  std::vector<char> lds;
};
inline thread_local emu_uint3 tIdx;
inline thread_local emu_uint3 bIdx;
inline thread_local dim3 bDim;
inline thread_local BlockCtx* ctx = nullptr;
inline void syncBlock() { ctx->block->arrive_and_wait(); }
inline double shflDown(double v, int offset, int width) {
  const unsigned lane = tIdx.x % static_cast<unsigned>(width);
  const unsigned wave = tIdx.x / 64u;
  auto& wb = *ctx->waves[wave];
  ctx->xchg[tIdx.x] = v;
  wb.arrive_and_wait();
  const unsigned src = lane + static_cast<unsigned>(offset);
  const unsigned waveLanes = std::min(64u, ctx->threads - wave * 64u);
  const double r = (src < static_cast<unsigned>(width) && src < waveLanes) ? ctx->xchg[tIdx.x + offset] : v;
  wb.arrive_and_wait();
  return r;
}
// DPP move with wave_shr:1 (the only control the kernels use): lane l gets
// lane l - 1's value, lane 0 keeps `old`.
inline int dppWaveShr1(int old, int v, int ctrl) {
  if (ctrl != 0x138) std::abort();
  const unsigned lane = tIdx.x % 64u;
  const unsigned wave = tIdx.x / 64u;
  auto& wb = *ctx->waves[wave];
  ctx->xchg[tIdx.x] = static_cast<double>(v);
  wb.arrive_and_wait();
  const int r = (lane > 0) ? static_cast<int>(ctx->xchg[tIdx.x - 1]) : old;
  wb.arrive_and_wait();
  return r;
}
// A persistent pool of lane threads (created once, up to 1024), so long runs
// do not create threads per launch.
class LanePool {
 public:
  static LanePool& get() {
    // never destroyed: the detached lanes wait on its condition variables
    // until the process exits
    static LanePool* pool = new LanePool;
    return *pool;
  }
  void run(unsigned lanes, const std::function<void(unsigned)>& body) {
    std::unique_lock<std::mutex> lock(m_);
    ensure(lanes);
    body_ = &body;
    lanes_ = lanes;
    pending_ = lanes;
    ++generation_;
    go_.notify_all();
    done_.wait(lock, [this] { return pending_ == 0; });
  }

 private:
  void ensure(unsigned lanes) {
    while (threads_.size() < lanes) {
      const unsigned id = static_cast<unsigned>(threads_.size());
      threads_.emplace_back([this, id] { loop(id); });
      threads_.back().detach();
    }
  }
  void loop(unsigned id) {
    unsigned long seen = 0;
    for (;;) {
      const std::function<void(unsigned)>* body;
      {
        std::unique_lock<std::mutex> lock(m_);
        go_.wait(lock, [&] { return generation_ != seen; });
        seen = generation_;
        if (id >= lanes_) continue;
        body = body_;
      }
      (*body)(id);
      std::lock_guard<std::mutex> lock(m_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::mutex m_;
  std::condition_variable go_, done_;
  std::vector<std::thread> threads_;
  const std::function<void(unsigned)>* body_{nullptr};
  unsigned lanes_{0}, pending_{0};
  unsigned long generation_{0};
};

template <class F>
void runBlock(dim3 b, dim3 t, size_t shmem, F&& body) {
  BlockCtx c;
  c.threads = t.x * t.y * t.z;
  c.block = std::make_unique<std::barrier<>>(static_cast<std::ptrdiff_t>(c.threads));
  for (unsigned w = 0; w * 64 < c.threads; ++w)
    c.waves.push_back(std::make_unique<std::barrier<>>(static_cast<std::ptrdiff_t>(std::min(64u, c.threads - w * 64))));
  c.xchg.assign(c.threads, 0.0);
  c.lds.assign(shmem, 0);
  const std::function<void(unsigned)> lane = [&](unsigned q) {
    ctx = &c;
    tIdx.x = q;
    bIdx.x = b.x;
    bIdx.y = b.y;
    bDim = t;
    body();
    c.waves[q / 64]->arrive_and_drop();
    c.block->arrive_and_drop();
  };
  LanePool::get().run(c.threads, lane);
}
}  // namespace cfdemu

#define threadIdx (cfdemu::tIdx)
#define blockIdx (cfdemu::bIdx)
#define blockDim (cfdemu::bDim)
#define __syncthreads() cfdemu::syncBlock()
#define __builtin_amdgcn_s_barrier() cfdemu::syncBlock()
#define __builtin_amdgcn_s_waitcnt(x) ((void)(x))
#define HIP_DYNAMIC_SHARED(type, var) type* var = reinterpret_cast<type*>(cfdemu::ctx->lds.data());
inline double __shfl_down(double v, int offset, int width = 64) { return cfdemu::shflDown(v, offset, width); }
#define __builtin_amdgcn_update_dpp(old, v, ctrl, rm, bm, bc) cfdemu::dppWaveShr1((old), (v), (ctrl))
#define __builtin_amdgcn_s_sleep(x) std::this_thread::yield()
// lanes are threads here: each keeps its own copy of a wave-uniform value
#define __builtin_amdgcn_readfirstlane(x) (x)
#define __HIP_MEMORY_SCOPE_WORKGROUP 2
template <class T>
inline T __hip_atomic_load(const T* p, int order, int) {
  (void)order;
  return std::atomic_ref<T>(*const_cast<T*>(p)).load(std::memory_order_seq_cst);
}
template <class T>
inline void __hip_atomic_store(T* p, T v, int order, int) {
  (void)order;
  std::atomic_ref<T>(*p).store(v, std::memory_order_seq_cst);
}
inline int atomicAdd(int* p, int v) { return std::atomic_ref<int>(*p).fetch_add(v); }
#define __HIP_MEMORY_SCOPE_AGENT 3
template <class T>
inline T __hip_atomic_fetch_add(T* p, T v, int order, int) {
  (void)order;
  return std::atomic_ref<T>(*p).fetch_add(v);
}

// runtime API
using hipError_t = int;
struct emu_stream;
using hipStream_t = emu_stream*;
constexpr hipError_t hipSuccess = 0;
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
constexpr unsigned hipStreamNonBlocking = 1;
inline const char* hipGetErrorString(hipError_t) { return "emulation error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : 1; }
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))
template <class T>
inline hipError_t hipGetSymbolAddress(void** p, T* sym) { *p = static_cast<void*>(sym); return hipSuccess; }
template <class T>
inline hipError_t hipMalloc(T** p, size_t bytes) {
  *p = static_cast<T*>(std::malloc(bytes));
  return *p ? hipSuccess : 2;
}
template <class T>
inline hipError_t hipMallocAsync(T** p, size_t bytes, hipStream_t) { return hipMalloc(p, bytes); }
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
constexpr unsigned hipHostMallocDefault = 0;
inline hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
  *p = std::malloc(bytes);
  return *p ? hipSuccess : 2;
}
inline hipError_t hipHostFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipFreeAsync(void* p, hipStream_t) { return hipFree(p); }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  std::memmove(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  static char token;
  *s = reinterpret_cast<hipStream_t>(&token);
  return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
// Launches run to completion when queued, so an event is a host timestamp.
struct emu_event {
  std::chrono::steady_clock::time_point t;
};
using hipEvent_t = emu_event*;
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = new emu_event(); return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = std::chrono::steady_clock::now(); return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }

// Graphs: a capture records each launch as a closure (arguments by value);
// a launch of the instantiated graph runs them in order.
struct emu_graph {
  std::vector<std::function<void()>> ops;
};
using hipGraph_t = emu_graph*;
using hipGraphExec_t = emu_graph*;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
namespace cfdemu {
inline thread_local emu_graph* capturing = nullptr;
}
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  cfdemu::capturing = new emu_graph();
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  *g = cfdemu::capturing;
  cfdemu::capturing = nullptr;
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, char*, size_t) {
  *e = new emu_graph(*g);
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto& f : e->ops) f();
  return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }

template <class... P, class... A>
void hipLaunchKernelGGL(void (*kernel)(P...), dim3 grid, dim3 block, size_t shmem, hipStream_t, A... args) {
  auto run = [=] {
    for (unsigned by = 0; by < grid.y; ++by)
      for (unsigned bx = 0; bx < grid.x; ++bx)
        cfdemu::runBlock(dim3(bx, by), block, shmem, [&] { kernel(static_cast<P>(args)...); });
  };
  if (cfdemu::capturing) cfdemu::capturing->ops.push_back(run);
  else run();
}

