// common.hpp - device helpers shared by the SIMPLE kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"

namespace cfdgpu {

// std::max / std::min / std::clamp with the standard's exact NaN and signed
// zero behaviour (the reference uses these, not fmax/fmin).
__device__ __forceinline__ double mx(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double mn(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double clampd(double v, double lo, double hi) {
  return (v < lo) ? lo : ((hi < v) ? hi : v);
}
__device__ __forceinline__ float mxf(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ bool imposes(const FaceBC& bc, double F) {
  // SimpleSolver::imposesVelocity (SimpleSolver.cpp:247-262)
  return bc.kind == kWall || bc.kind == kInlet || (bc.kind == kFarField && F < bc.dead);
}

// Lane l of the wave receives lane l - 1's value and lane 0 receives `fill`
// (a DPP lane with no source keeps the old value): two 32-bit DPP moves with
// the whole-wave shift wave_shr:1.
__device__ __forceinline__ double waveShiftUp(double v, double fill = 0.0) {
  const long long b = __builtin_bit_cast(long long, v);
  const long long f = __builtin_bit_cast(long long, fill);
  const int lo = __builtin_amdgcn_update_dpp(static_cast<int>(f), static_cast<int>(b), 0x138, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(static_cast<int>(f >> 32), static_cast<int>(b >> 32), 0x138, 0xF, 0xF, false);
  return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// Step counters shared by the waves of one workgroup, in LDS.
__device__ __forceinline__ int ldsPoll(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void ldsPost(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// This wave's earlier LDS accesses are complete before its later ones issue
// (global loads in flight are not waited for).
__device__ __forceinline__ void ldsOrder() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt untouched
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ __forceinline__ void spinPause() { __builtin_amdgcn_s_sleep(1); }
// Compiler-only barrier: the values in v are computed before it and no memory
// access is moved across it (emits no instruction).
__device__ __forceinline__ void holdUntilHere(double v) { asm volatile("" : : "v"(v) : "memory"); }
// Compiler-only: v is computed on every lane at this point (the compiler may
// not sink its computation into a branch of the select that uses it).
__device__ __forceinline__ void computedHere(double v) { asm volatile("" : : "v"(v)); }
// Compiler-only: the offset is an opaque 32-bit register value here, so the
// accesses that use it stay (scalar base + 32-bit offset) instead of being
// strength-reduced into 64-bit pointer increments.
__device__ __forceinline__ void opaqueOffset(unsigned& o) { asm volatile("" : "+v"(o)); }
template <int N>
__device__ __forceinline__ void holdUntilHere(const double (&v)[N]) {
#pragma unroll
  for (int q = 0; q < N; ++q) asm volatile("" : : "v"(v[q]) : "memory");
}

// Blocks are handed to the 8 XCDs round-robin; remap so each XCD works on a
// contiguous run of tiles and the j-1 / j+1 halo rows it re-reads stay in its
// own L2.
__device__ __forceinline__ int xcdTile(int b, int nb) {
  constexpr int kXcd = 8;
  if (nb < 2 * kXcd) return b;
  const int per = nb / kXcd;
  const int rem = nb - per * kXcd;
  const int xcd = b % kXcd;
  const int k = b / kXcd;
  // XCD x owns tiles [start(x), start(x) + count(x)); the first `rem` XCDs own one extra.
  const int start = xcd * per + (xcd < rem ? xcd : rem);
  const int count = per + (xcd < rem ? 1 : 0);
  return (k < count) ? start + k : -1;
}

// Deterministic block reduction of NV sums over a 256-thread block (4 waves of
// 64): wave shuffles, then the 4 wave results added in wave order by thread 0.
template <int NV>
__device__ __forceinline__ void blockSum(double (&v)[NV], double* sm) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] += __shfl_down(v[q], o, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) sm[q * 4 + wave] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = ((sm[q * 4 + 0] + sm[q * 4 + 1]) + sm[q * 4 + 2]) + sm[q * 4 + 3];
  }
  __syncthreads();
}

template <int NV>
__device__ __forceinline__ void blockMax(double (&v)[NV], double* sm) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = mx(v[q], __shfl_down(v[q], o, 64));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) sm[q * 4 + wave] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = mx(mx(mx(sm[q * 4 + 0], sm[q * 4 + 1]), sm[q * 4 + 2]), sm[q * 4 + 3]);
  }
  __syncthreads();
}

// Global-address-space views of device pointers. Loads and stores through a
// generic pointer compile to FLAT instructions, which on CDNA count in
// lgkmcnt as well as vmcnt: a wait for LDS traffic (lgkmcnt(0)) then also
// waits for every outstanding HBM access. Kernels that overlap memory traffic
// with LDS hand-offs go through these so their accesses are GLOBAL ones.
#define CFD_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ CFD_GLOBAL T* gmem(T* p) {
  return (CFD_GLOBAL T*)(p);
}
template <class T>
__device__ __forceinline__ const CFD_GLOBAL T* gmem(const T* p) {
  return (const CFD_GLOBAL T*)(p);
}

// A pair of doubles as one 16-byte vector: loads and stores of it through a
// global pointer are single dwordx4 accesses. (HIP's double2 wrapper cannot be
// copied out of an address-space-qualified location.)
typedef double dpair __attribute__((vector_size(16)));
__device__ __forceinline__ const CFD_GLOBAL dpair* gpair(const double2* p) { return (const CFD_GLOBAL dpair*)(p); }
__device__ __forceinline__ CFD_GLOBAL dpair* gpair(double2* p) { return (CFD_GLOBAL dpair*)(p); }
// Element i of a double2 array, and a store to it, as one 16-byte GLOBAL access.
__device__ __forceinline__ double2 ld2(const double2* p, int i) {
  const dpair v = gpair(p)[i];
  return make_double2(v[0], v[1]);
}
__device__ __forceinline__ void st2(double2* p, int i, double x, double y) {
  const dpair v = {x, y};
  gpair(p)[i] = v;
}

// Cross-block scalars: each kernel writes one partial per block and the
// kernel that needs the total sums the partials itself, in a fixed order, at
// its start (the kernel boundary makes them visible). No fences, no atomics:
// deterministic, and nothing forces the per-XCD L2s to write back mid-kernel.
//
// Wave 0 folds p[0..n): lane l combines p[l], p[l + 64], p[l + 128], ... in
// that order, then a shuffle tree combines the lanes. A lane issues up to
// kPartialUnroll loads before its first combine, so up to 64 * kPartialUnroll
// partials cost one memory round trip. `identity` is op's neutral element and
// stands in for the slots past n.
constexpr int kPartialUnroll = 16;

struct OpSum {
  __device__ double operator()(double a, double b) const { return a + b; }
};
struct OpMax {
  __device__ double operator()(double a, double b) const { return mx(a, b); }
};
struct OpMin {
  __device__ double operator()(double a, double b) const { return mn(a, b); }
};

template <int NS, class Op, int UNROLL = kPartialUnroll>
__device__ __forceinline__ void wavePartials(const double* const (&p)[NS], int n, double identity, Op op,
                                             double (&out)[NS]) {
  const int lane = static_cast<int>(threadIdx.x) & 63;
#pragma unroll
  for (int q = 0; q < NS; ++q) out[q] = identity;
  for (int b0 = 0; b0 < n; b0 += 64 * UNROLL) {
    double t[NS][UNROLL];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const CFD_GLOBAL double* g = gmem(p[q]);
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        // every lane loads (clamped to the last partial, n >= 1) and the
        // value is then chosen: no branch around the load, so all loads of
        // the wave are in flight together
        const int b = b0 + u * 64 + lane;
        const double val = g[min(b, n - 1)];
        t[q][u] = (b < n) ? val : identity;
      }
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) {
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) out[q] = op(out[q], t[q][u]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int q = 0; q < NS; ++q) out[q] = op(out[q], __shfl_down(out[q], o, 64));
  }
}

// Single-slot forms for wave 0; the result is in lane 0.
// (UNROLL sets only how many loads are in flight per lane, not the order of the sum)
template <int UNROLL = kPartialUnroll>
__device__ __forceinline__ double waveSumPartials(const double* p, int n) {
  const double* const s[1] = {p};
  double r[1];
  wavePartials<1, OpSum, UNROLL>(s, n, 0.0, OpSum{}, r);
  return r[0];
}
__device__ __forceinline__ double waveMaxPartials(const double* p, int n, double identity = 0.0) {
  const double* const s[1] = {p};
  double r[1];
  wavePartials<1>(s, n, identity, OpMax{}, r);
  return r[0];
}
__device__ __forceinline__ double waveMinPartials(const double* p, int n, double identity) {
  const double* const s[1] = {p};
  double r[1];
  wavePartials<1>(s, n, identity, OpMin{}, r);
  return r[0];
}

// Sums of NS slots' partials, broadcast to every thread of the block. Wave 0
// sums them (up to UNROLL loads per lane in flight per slot) while the other
// waves wait at the barrier; splitting the sum over all four waves measured
// slower.
template <int NS, int UNROLL = kPartialUnroll>
__device__ __forceinline__ void blockSumPartials(const double* const (&p)[NS], int n, double* sm,
                                                 double (&out)[NS]) {
  if (threadIdx.x < 64) {
    wavePartials<NS, OpSum, UNROLL>(p, n, 0.0, OpSum{}, out);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < NS; ++q) sm[q] = out[q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NS; ++q) out[q] = sm[q];
  __syncthreads();
}

template <int UNROLL = kPartialUnroll>
__device__ __forceinline__ double blockSumPartials(const double* p, int n, double* sm) {
  const double* const s[1] = {p};
  double r[1];
  blockSumPartials<1, UNROLL>(s, n, sm, r);
  return r[0];
}

__device__ __forceinline__ double* partialSlot(const Inst& I, int slot) {
  return I.partials + static_cast<size_t>(slot) * kMaxRedBlocks;
}

}  // namespace cfdgpu
