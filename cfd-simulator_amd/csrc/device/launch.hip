// launch.hip - the launch schedule of one SIMPLE outer iteration
// (SimpleSolver.cpp:662-794), for a batch of instances sharing one descriptor
// array (grid.y = instance).
#include <hip/hip_runtime.h>

#include "../host/engine.hpp"
#include "layout.hpp"

namespace cfdgpu {

__global__ void k_begin_iteration(const Inst*, int);
__global__ void k_grad(const Inst*);
__global__ void k_mom(const Inst*);
__global__ void k_gs_wave_mom(const Inst*, int);
__global__ void k_gs_wave_turb(const Inst*, int);
__global__ void k_gs_wave_mom_wide(const Inst*, int);
__global__ void k_gs_wave_turb_wide(const Inst*, int);
__global__ void k_gs_unskew(const Inst*, int, int);
__global__ void k_wall_distance(const double2*, int, const double2*, int, double*);
__global__ void k_gs_color(const Inst*, int, int, int);
__global__ void k_rhiechow(const Inst*);
__global__ void k_pressure(const Inst*, int);
__global__ void k_cg_init(const Inst*, int);
__global__ void k_cg_a(const Inst*, int);
__global__ void k_cg_b(const Inst*, int);
__global__ void k_corr_mean(const Inst*);
__global__ void k_correct(const Inst*);
__global__ void k_divergence(const Inst*, double*);
__global__ void k_monitor(const Inst*);
__global__ void k_sst_blend(const Inst*);
__global__ void k_sst_asm(const Inst*);
__global__ void k_sst_wall(const Inst*);
__global__ void k_sst_mut(const Inst*);
__global__ void k_cg_init_f32(const Inst*, int);
__global__ void k_lcg_init(const Inst*, int);
__global__ void k_lcg_a_c1(const Inst*, int);
__global__ void k_lcg_a_c2(const Inst*, int);
__global__ void k_lcg_a_c4(const Inst*, int);
__global__ void k_lcg_a_tile(const Inst*, int);
__global__ void k_lcg_b_u16(const Inst*, int, int);
__global__ void k_lcg_b_u32(const Inst*, int, int);
__global__ void k_lcg_b_seg(const Inst*, int, int);
__global__ void k_lcg_init_f32(const Inst*, int);
__global__ void k_lcg_a_f32_c1(const Inst*, int);
__global__ void k_lcg_a_f32_c2(const Inst*, int);
__global__ void k_lcg_a_f32_c4(const Inst*, int);
__global__ void k_lcg_a_f32_tile(const Inst*, int);
__global__ void k_lcg_b_f32_u16(const Inst*, int, int);
__global__ void k_lcg_b_f32_u32(const Inst*, int, int);
__global__ void k_lcg_b_f32_seg(const Inst*, int, int);
__global__ void k_lcg_final(const Inst*);
__global__ void k_lcg_final_f32(const Inst*);
__global__ void k_refine_acc(const Inst*, int);
__global__ void k_refine_res(const Inst*);
__global__ void k_refine_final(const Inst*);
__global__ void k_rb32(const Inst*, int, int, int, int);
__global__ void k_cg_a_f32(const Inst*, int);
__global__ void k_cg_b_f32(const Inst*, int);

int gsThreadsFor(int nJ) { return gsLanes(nJ); }

// Dynamic LDS of the lexicographic sweep (gsLexWave): row 0, the [waves][64]
// hand-off ring and a dummy per lane (doubles); the partner map, the step
// counters and a dummy per lane (ints).
size_t gsLdsFor(int nJ, int nI, int threads) {
  (void)nJ;
  const size_t t = static_cast<size_t>(threads), n = static_cast<size_t>(nI), waves = (t + 63) / 64;
  return (n + waves * 64 + t) * sizeof(double) + (n + waves + t) * sizeof(int);
}

void launchIteration(const Inst* d, const LaunchCfg& cfg, const Inst& h, hipStream_t s, StageRecorder* rec) {
  const dim3 grid(static_cast<unsigned>(cfg.maxBlocks), static_cast<unsigned>(cfg.batch));
  const dim3 quadGrid(static_cast<unsigned>(quadBlocks(cfg.maxBlocks)), static_cast<unsigned>(cfg.batch));
  const dim3 block(kBlock);
  const dim3 one((cfg.batch + 63) / 64);
  const bool lex = h.gsOrder == 0;
  const bool mixed = h.precision == 1;
  // the mixed path's two-colour sweeps run on FP32 defect systems
  const bool rb32 = mixed && !lex;
  const bool line = h.pSolver == 1;
  const int refine = h.refinePasses < 1 ? 1 : h.refinePasses;
  const dim3 colGrid(static_cast<unsigned>(cfg.maxColBlocks), static_cast<unsigned>(cfg.batch));
  // n sweeps of one system in two colours; the first colour of the first
  // sweep reads no neighbour, the last colour of the last applies the result
  const dim3 rbGrid(static_cast<unsigned>(cfg.maxRbBlocks), static_cast<unsigned>(cfg.batch));
  auto rbSweeps = [&](int sys, int n) {
    for (int sw = 0; sw < n; ++sw)
      for (int color = 0; color < 2; ++color)
        hipLaunchKernelGGL(k_rb32, rbGrid, block, 0, s, d, sys, color, (sw == 0 && color == 0) ? 1 : 0,
                           (sw == n - 1 && color == 1) ? 1 : 0);
  };
  auto begin = [&](int stage) { if (rec) rec->begin(stage, s); };
  auto end = [&](int stage) { if (rec) rec->end(stage, s); };
  begin(CFDGPU_STAGE_GRADIENTS);
  hipLaunchKernelGGL(k_begin_iteration, one, dim3(64), 0, s, d, cfg.batch);
  hipLaunchKernelGGL(k_grad, grid, block, 0, s, d);
  end(CFDGPU_STAGE_GRADIENTS);
  begin(CFDGPU_STAGE_MOMENTUM_ASSEMBLY);
  hipLaunchKernelGGL(k_mom, grid, block, 0, s, d);
  end(CFDGPU_STAGE_MOMENTUM_ASSEMBLY);
  begin(CFDGPU_STAGE_MOMENTUM_SWEEPS);
  if (lex) {
    // u and v in two workgroups side by side
    const dim3 momGrid(2, cfg.batch);
    if (cfg.gsThreads <= 256) hipLaunchKernelGGL(k_gs_wave_mom, momGrid, dim3(cfg.gsThreads), cfg.gsLds, s, d, h.momSweeps);
    else hipLaunchKernelGGL(k_gs_wave_mom_wide, momGrid, dim3(cfg.gsThreads), cfg.gsLds, s, d, h.momSweeps);
    end(CFDGPU_STAGE_MOMENTUM_SWEEPS);
    begin(CFDGPU_STAGE_SWEEP_COPY);
    hipLaunchKernelGGL(k_gs_unskew, grid, block, 0, s, d, 0, h.momSweeps);
    end(CFDGPU_STAGE_SWEEP_COPY);
  } else if (rb32) {
    rbSweeps(0, h.momSweeps);
    end(CFDGPU_STAGE_MOMENTUM_SWEEPS);
  } else {
    for (int sw = 0; sw < h.momSweeps; ++sw) {
      hipLaunchKernelGGL(k_gs_color, grid, block, 0, s, d, 0, 0, 0);
      hipLaunchKernelGGL(k_gs_color, grid, block, 0, s, d, 0, 1, 0);
    }
    end(CFDGPU_STAGE_MOMENTUM_SWEEPS);
  }
  begin(CFDGPU_STAGE_RHIE_CHOW);
  hipLaunchKernelGGL(k_rhiechow, grid, block, 0, s, d);
  end(CFDGPU_STAGE_RHIE_CHOW);
  // SimpleSolver.cpp:716-748
  const int passes = h.nonOrthPasses < 1 ? 1 : h.nonOrthPasses;
  for (int pass = 0; pass < passes; ++pass) {
    begin(CFDGPU_STAGE_PRESSURE_ASSEMBLY);
    hipLaunchKernelGGL(k_pressure, grid, block, 0, s, d, pass);
    end(CFDGPU_STAGE_PRESSURE_ASSEMBLY);
    begin(CFDGPU_STAGE_PRESSURE_SOLVE);
    const dim3 cgGrid(cfg.maxCgBlocks, cfg.batch);
    const dim3 col(kColThreads);
    for (int rp = 0; rp < refine; ++rp) {
      if (line) {
        // the launch tuning's shapes (cfdgpu_set_launch_tuning)
        using KA = void (*)(const Inst*, int);
        using KB = void (*)(const Inst*, int, int);
        const int cells = cfg.lcgCells;
        const KA ka = mixed ? (cells == 0   ? k_lcg_a_f32_tile
                               : cells == 1 ? k_lcg_a_f32_c1
                               : cells == 4 ? k_lcg_a_f32_c4
                                            : k_lcg_a_f32_c2)
                            : (cells == 0   ? k_lcg_a_tile
                               : cells == 1 ? k_lcg_a_c1
                               : cells == 4 ? k_lcg_a_c4
                                            : k_lcg_a_c2);
        // columns of at most 256 rows: the segmented launch B (1024 threads per
        // 64 columns); taller ones: one lane per column, walking the tuned depth ahead
        const bool seg = cfg.maxNJ <= kLcgSegMaxRows;
        const KB kb = seg ? (mixed ? k_lcg_b_f32_seg : k_lcg_b_seg)
                          : mixed ? (cfg.colUnroll == 16 ? k_lcg_b_f32_u16 : k_lcg_b_f32_u32)
                                  : (cfg.colUnroll == 16 ? k_lcg_b_u16 : k_lcg_b_u32);
        const dim3 colB(seg ? kLcgSegThreads : kColThreads);
        const dim3 aGrid(static_cast<unsigned>(cfg.maxLcgBlocks), static_cast<unsigned>(cfg.batch));
        hipLaunchKernelGGL(mixed ? k_lcg_init_f32 : k_lcg_init, colGrid, col, 0, s, d, rp);
        begin(CFDGPU_STAGE_CG_ITERATIONS);
        for (int it = 0; it < cfg.cgIters; ++it) {
          hipLaunchKernelGGL(ka, aGrid, block, 0, s, d, it);
          hipLaunchKernelGGL(kb, colGrid, colB, 0, s, d, it, cells);
        }
        end(CFDGPU_STAGE_CG_ITERATIONS);
        hipLaunchKernelGGL(mixed ? k_lcg_final_f32 : k_lcg_final, quadGrid, block, 0, s, d);
      } else if (mixed) {
        hipLaunchKernelGGL(k_cg_init_f32, cgGrid, block, 0, s, d, rp);
        begin(CFDGPU_STAGE_CG_ITERATIONS);
        for (int it = 0; it < cfg.cgIters; ++it) {
          hipLaunchKernelGGL(k_cg_a_f32, cgGrid, block, 0, s, d, it);
          hipLaunchKernelGGL(k_cg_b_f32, cgGrid, block, 0, s, d, it);
        }
        end(CFDGPU_STAGE_CG_ITERATIONS);
      } else {
        hipLaunchKernelGGL(k_cg_init, cgGrid, block, 0, s, d, rp);
        begin(CFDGPU_STAGE_CG_ITERATIONS);
        for (int it = 0; it < cfg.cgIters; ++it) {
          hipLaunchKernelGGL(k_cg_a, cgGrid, block, 0, s, d, it);
          hipLaunchKernelGGL(k_cg_b, cgGrid, block, 0, s, d, it);
        }
        end(CFDGPU_STAGE_CG_ITERATIONS);
      }
      if (rp + 1 < refine) {
        // FP64-residual refinement: accumulate the correction, then the
        // next pass's right-hand side b - A x in FP64
        hipLaunchKernelGGL(k_refine_acc, grid, block, 0, s, d, rp);
        hipLaunchKernelGGL(k_refine_res, grid, block, 0, s, d);
      }
    }
    // the last pass's solution plus what the earlier passes accumulated
    if (refine > 1) hipLaunchKernelGGL(k_refine_final, grid, block, 0, s, d);
    end(CFDGPU_STAGE_PRESSURE_SOLVE);
    begin(CFDGPU_STAGE_CORRECTION);
    if (!h.hasPRef) hipLaunchKernelGGL(k_corr_mean, grid, block, 0, s, d);
    hipLaunchKernelGGL(k_correct, grid, block, 0, s, d);
    end(CFDGPU_STAGE_CORRECTION);
  }
  begin(CFDGPU_STAGE_CORRECTION);
  hipLaunchKernelGGL(k_divergence, quadGrid, block, 0, s, d, static_cast<double*>(nullptr));
  end(CFDGPU_STAGE_CORRECTION);
  if (h.model == 2) {
    begin(CFDGPU_STAGE_SST_BLENDING);
    hipLaunchKernelGGL(k_sst_blend, grid, block, 0, s, d);
    end(CFDGPU_STAGE_SST_BLENDING);
    begin(CFDGPU_STAGE_SST_ASSEMBLY);
    hipLaunchKernelGGL(k_sst_asm, grid, block, 0, s, d);
    end(CFDGPU_STAGE_SST_ASSEMBLY);
    begin(CFDGPU_STAGE_SST_SWEEPS);
    if (lex) {
      if (cfg.gsThreads <= 256) hipLaunchKernelGGL(k_gs_wave_turb, dim3(2, cfg.batch), dim3(cfg.gsThreads), cfg.gsLds, s, d, 3);
      else hipLaunchKernelGGL(k_gs_wave_turb_wide, dim3(2, cfg.batch), dim3(cfg.gsThreads), cfg.gsLds, s, d, 3);
      end(CFDGPU_STAGE_SST_SWEEPS);
      begin(CFDGPU_STAGE_SWEEP_COPY);
      hipLaunchKernelGGL(k_gs_unskew, grid, block, 0, s, d, 1, 3);
      end(CFDGPU_STAGE_SWEEP_COPY);
    } else if (rb32) {
      rbSweeps(1, 3);
      end(CFDGPU_STAGE_SST_SWEEPS);
    } else {
      for (int sw = 0; sw < 3; ++sw) {
        const int floor = (sw == 2) ? 1 : 0;
        hipLaunchKernelGGL(k_gs_color, grid, block, 0, s, d, 1, 0, 0);
        hipLaunchKernelGGL(k_gs_color, grid, block, 0, s, d, 1, 1, floor);
        hipLaunchKernelGGL(k_gs_color, grid, block, 0, s, d, 2, 0, 0);
        hipLaunchKernelGGL(k_gs_color, grid, block, 0, s, d, 2, 1, floor);
      }
      end(CFDGPU_STAGE_SST_SWEEPS);
    }
    begin(CFDGPU_STAGE_SST_WALL_VISCOSITY);
    hipLaunchKernelGGL(k_sst_wall, dim3(1, cfg.batch), block, 0, s, d);
    hipLaunchKernelGGL(k_sst_mut, grid, block, 0, s, d);
    end(CFDGPU_STAGE_SST_WALL_VISCOSITY);
  }
  begin(CFDGPU_STAGE_MONITOR);
  hipLaunchKernelGGL(k_monitor, dim3(1, cfg.batch), dim3(64), 0, s, d);
  end(CFDGPU_STAGE_MONITOR);
}

namespace {
__global__ void k_resume(const Inst* insts, int count) {
  const int b = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (b < count) insts[b].flags[fStopped] = 0;
}
}  // namespace

void launchResumeAll(const Inst* d, int count, hipStream_t s) {
  hipLaunchKernelGGL(k_resume, dim3((count + 63) / 64), dim3(64), 0, s, d, count);
}

void launchWallDistance(const double2* centroid, int n, const double2* walls, int nw, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_wall_distance, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, centroid, n, walls, nw,
                     out);
}

void launchDivergence(const Inst* d, const LaunchCfg& cfg, double* out, hipStream_t s) {
  const dim3 grid(static_cast<unsigned>(quadBlocks(cfg.maxBlocks)), static_cast<unsigned>(cfg.batch));
  hipLaunchKernelGGL(k_divergence, grid, dim3(kBlock), 0, s, d, out);
}

}  // namespace cfdgpu
