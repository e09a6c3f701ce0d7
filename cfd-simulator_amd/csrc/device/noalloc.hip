// noalloc.hip - the divergence field without a device allocation, for callers
// that cannot allocate (code under HIP graph capture, tight serving loops).
// No reference counterpart: SimpleSolver::divergence() (SimpleSolver.cpp:
// 796-817) returns a new vector; cfdgpu_divergence stages the field in a
// buffer it allocates per call.
//
// The staging buffer here is a file-scope device array, `scratch`, of
// kNoAllocMaxCells doubles (8 MiB), which every GPU the library runs on holds
// until the process ends. Host code never uses the array's name as an address:
// on each device's first call hipGetSymbolAddress gives that device's copy,
// and the pointer is kept per device. Calls that overlap on one device share
// the array and race.
#include <hip/hip_runtime.h>

#include <array>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../host/engine.hpp"
#include "layout.hpp"

namespace cfdgpu {

namespace {

constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocMaxCells];

std::mutex gScratchMutex;
std::array<double*, kMaxDevices> gScratch{};  // this device's address of `scratch`, once looked up

double* scratchOnCurrentDevice() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    throw std::runtime_error("no current HIP device for the divergence scratch");
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (!gScratch[static_cast<size_t>(dev)]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw std::runtime_error(std::string("hipGetSymbolAddress (divergence scratch): ") + hipGetErrorString(e));
    gScratch[static_cast<size_t>(dev)] = static_cast<double*>(ptr);
  }
  return gScratch[static_cast<size_t>(dev)];
}

}  // namespace

void launchDivergenceNoAlloc(const Inst* d, const LaunchCfg& cfg, int cells, double* out, int* flagStopped,
                             hipStream_t s) {
  if (cells > kNoAllocMaxCells)
    throw std::invalid_argument("cfdgpu_divergence_noalloc supports at most " + std::to_string(kNoAllocMaxCells) +
                                " cells; use cfdgpu_divergence");
  double* buf = scratchOnCurrentDevice();
  int zero = 0;
  if (hipMemcpyAsync(flagStopped, &zero, sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess)
    throw std::runtime_error("divergence: flags");
  launchDivergence(d, cfg, buf, s);
  if (hipGetLastError() != hipSuccess) throw std::runtime_error("divergence: launch");
  if (hipMemcpyAsync(out, buf, sizeof(double) * static_cast<size_t>(cells), hipMemcpyDeviceToHost, s) != hipSuccess)
    throw std::runtime_error("divergence: download");
  if (hipStreamSynchronize(s) != hipSuccess) throw std::runtime_error("divergence: sync");
}

}  // namespace cfdgpu
