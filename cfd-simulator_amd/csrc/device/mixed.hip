// mixed.hip - the fast path: FP32 bulk, FP64 state, residuals and reductions.
//
// Sweeps run in defect-correction form: the FP64 assembly kernels hand over
// the FP64 residual r = b - A x of the starting field (rounded to FP32) and
// the FP32 matrix; the sweeps solve A d = r from d = 0 and the FP64 state is
// updated x += d. In exact arithmetic this is the same Gauss-Seidel iteration
// on x (LinearSystem.cpp:77-116), and the converged fixed point is the FP64
// one, because the residual that decides convergence is formed in FP64.
//
// The pressure correction is a CG in FP32 with FP64 inner products and L1
// norms, preconditioned either by the reference's Jacobi scaling
// (LinearSystem.cpp:151-159) or by exact solves along the wall-normal grid
// lines (a block-diagonal, tridiagonal preconditioner: the near-wall cells are
// up to 1e5 times longer than thick, which is the coupling Jacobi cannot
// see). Its solution is checked and refined against the FP64 matrix.
#include "common.hpp"

namespace cfdgpu {

#define MX_PROLOGUE                                           \
  const Inst& I = insts[blockIdx.y];                          \
  if (static_cast<int>(blockIdx.x) >= I.nBlocks) return;     \
  if (I.flags[fStopped]) return;                              \
  const int tile = xcdTile(static_cast<int>(blockIdx.x), I.nBlocks); \
  const int c = tile * kBlock + static_cast<int>(threadIdx.x); \
  const bool live = c < I.N;                                  \
  const int nI = I.nI;                                        \
  const int nJ = I.nJ;                                        \
  const int i = live ? c % nI : 0;                            \
  const int j = live ? c / nI : 0;                            \
  (void)i; (void)j; (void)nJ;

// ---------------------------------------------------------------------------
// Two-colour sweeps on the FP32 defect. sys 0: momentum (two right-hand
// sides, one matrix); sys 1: k; sys 2: omega. On the last launch every lane
// adds its correction to the FP64 state (and applies the k / omega floor).
// ---------------------------------------------------------------------------

struct FSys {
  const float *aP, *aS, *aN, *aW, *aE;
  const float* r[2];
  float* d[2];
};

__device__ __forceinline__ FSys fsys(const Inst& I, int sys) {
  if (sys == 0) return FSys{I.faP, I.faS, I.faN, I.faW, I.faE, {I.fbx, I.fby}, {I.fdx, I.fdy}};
  if (sys == 1) return FSys{I.fkP, I.fkS, I.fkN, I.fkW, I.fkE, {I.fkR, nullptr}, {I.fkD, nullptr}};
  return FSys{I.fwP, I.fwS, I.fwN, I.fwW, I.fwE, {I.fwR, nullptr}, {I.fwD, nullptr}};
}

template <int NC>
__device__ __forceinline__ void fsweepPoint(const Inst& I, const FSys& S, int c, int i, int j) {
  const int nI = I.nI, nJ = I.nJ;
  const float a = S.aP[c];
  if (!(fabsf(a) > 0.0f)) return;
  float off[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) off[q] = 0.0f;
  if (j > 0) {
    const float s = S.aS[c];
#pragma unroll
    for (int q = 0; q < NC; ++q) off[q] = fmaf(s, S.d[q][c - nI], off[q]);
  } else {
    const int pr = I.partner[i];
    if (pr >= 0) {
      const float s = S.aS[c];
#pragma unroll
      for (int q = 0; q < NC; ++q) off[q] = fmaf(s, S.d[q][pr], off[q]);
    }
  }
  if (i < nI - 1) {
    const float e = S.aE[c];
#pragma unroll
    for (int q = 0; q < NC; ++q) off[q] = fmaf(e, S.d[q][c + 1], off[q]);
  }
  if (j < nJ - 1) {
    const float n = S.aN[c];
#pragma unroll
    for (int q = 0; q < NC; ++q) off[q] = fmaf(n, S.d[q][c + nI], off[q]);
  }
  if (i > 0) {
    const float w = S.aW[c];
#pragma unroll
    for (int q = 0; q < NC; ++q) off[q] = fmaf(w, S.d[q][c - 1], off[q]);
  }
  const float inv = 1.0f / a;
#pragma unroll
  for (int q = 0; q < NC; ++q) S.d[q][c] = (S.r[q][c] - off[q]) * inv;
}

__global__ void __launch_bounds__(kBlock) k_fsweep(const Inst* insts, int sys, int color, int last) {
  MX_PROLOGUE
  if (!live) return;
  const FSys S = fsys(I, sys);
  if (((i + j) & 1) == color) {
    if (sys == 0) fsweepPoint<2>(I, S, c, i, j);
    else fsweepPoint<1>(I, S, c, i, j);
  }
  if (!last) return;
  if (sys == 0) {
    I.u[c] += static_cast<double>(S.d[0][c]);
    I.v[c] += static_cast<double>(S.d[1][c]);
  } else if (sys == 1) {
    I.k[c] = mx(I.k[c] + static_cast<double>(S.d[0][c]), 0.0);
  } else {
    I.w[c] = mx(I.w[c] + static_cast<double>(S.d[0][c]), 1.0e-10);
  }
}

// ---------------------------------------------------------------------------
// Pressure correction: FP32 CG, FP64 reductions.
//   init   : b (mean removed when unanchored) -> r, z = M^-1 r, p = z, x = 0
//   A(it)  : finish the test of it-1, beta, p = z + beta p (halo recomputed),
//            Ap, p.Ap
//   B(it)  : alpha, x += alpha p, r -= alpha Ap, |r|_1, and for Jacobi
//            z = r / diag and r.z
//   L(it)  : line preconditioner only: z = M^-1 r column by column, r.z
// The FP32 rhs is the FP64 residual b - A x64 of the current correction, so a
// second pass refines x64 against the FP64 matrix.
// ---------------------------------------------------------------------------

__device__ __forceinline__ float spmvF(const Inst& I, const float* x, int c, int i, int j) {
  const int nI = I.nI, nJ = I.nJ;
  float y = I.fpP[c] * x[c];
  if (j > 0) y = fmaf(I.fpN[c - nI], x[c - nI], y);
  else {
    const int pr = I.partner[i];
    if (pr >= 0) y = fmaf(I.fpWk[c], x[pr], y);
  }
  if (j < nJ - 1) y = fmaf(I.fpN[c], x[c + nI], y);
  if (i > 0) y = fmaf(I.fpE[c - 1], x[c - 1], y);
  if (i < nI - 1) y = fmaf(I.fpE[c], x[c + 1], y);
  return y;
}

// FP64 residual of the current FP64 correction against the FP64 matrix, as
// the FP32 right-hand side of the next pass (pass 0: x64 = 0, rhs = b).
__global__ void __launch_bounds__(kBlock) k_cgf_rhs(const Inst* insts, int pass) {
  MX_PROLOGUE
  __shared__ double sm[16];
  double mean = 0.0;
  if (!I.hasPRef && pass == 0)
    mean = blockSumPartials(partialSlot(I, sSourceSum), I.nBlocks, sm) / static_cast<double>(I.N);
  double part[1] = {0.0};
  if (live) {
    double b = I.pb[c];
    if (!I.hasPRef && pass == 0) {
      b -= mean;
      I.pb[c] = b;
    }
    double r = b;
    if (pass > 0) {
      const double* x = I.x;
      double ax = I.pP[c] * x[c];
      if (j > 0) ax += I.pN[c - nI] * x[c - nI];
      else {
        const int pr = I.partner[i];
        if (pr >= 0) ax += I.pWk[c] * x[pr];
      }
      if (j < nJ - 1) ax += I.pN[c] * x[c + nI];
      if (i > 0) ax += I.pE[c - 1] * x[c - 1];
      if (i < nI - 1) ax += I.pE[c] * x[c + 1];
      r = b - ax;
    } else {
      I.x[c] = 0.0;
    }
    I.fb[c] = static_cast<float>(r);
    part[0] = fabs(r);
  }
  blockSum<1>(part, sm);
  if (threadIdx.x == 0) partialSlot(I, sRefineNorm)[blockIdx.x] = part[0];
}

// Line preconditioner setup: Thomas factors of each wall-normal column's
// tridiagonal block (diagonal, below/above couplings of the pressure matrix).
// One lane per column; stores 1/pivot and the eliminated upper coefficient.
__global__ void __launch_bounds__(64) k_line_factor(const Inst* insts) {
  const Inst& I = insts[blockIdx.y];
  if (I.flags[fStopped]) return;
  const int col = static_cast<int>(blockIdx.x * 64 + threadIdx.x);
  const int nI = I.nI, nJ = I.nJ;
  if (col >= nI) return;
  float cPrev = 0.0f;
  for (int j = 0; j < nJ; ++j) {
    const int c = j * nI + col;
    const float lower = (j > 0) ? I.fpN[c - nI] : 0.0f;
    const float piv = I.fpP[c] - lower * cPrev;
    const float inv = 1.0f / piv;
    const float upper = (j < nJ - 1) ? I.fpN[c] : 0.0f;
    cPrev = upper * inv;
    I.fLinv[c] = inv;
    I.fLc[c] = cPrev;
  }
}

// z = M^-1 r along every column, plus the r.z partial of this block of columns.
__global__ void __launch_bounds__(64) k_line_solve(const Inst* insts, int it) {
  const Inst& I = insts[blockIdx.y];
  if (I.flags[fStopped]) return;
  if (it >= 0 && I.flags[fCgDone]) return;
  __shared__ double sm[4];
  const int col = static_cast<int>(blockIdx.x * 64 + threadIdx.x);
  const int nI = I.nI, nJ = I.nJ;
  double rz = 0.0;
  if (col < nI) {
    float yPrev = 0.0f;
    for (int j = 0; j < nJ; ++j) {
      const int c = j * nI + col;
      const float lower = (j > 0) ? I.fpN[c - nI] : 0.0f;
      yPrev = (I.fr[c] - lower * yPrev) * I.fLinv[c];
      I.fz[c] = yPrev;
    }
    float zNext = 0.0f;
    for (int j = nJ - 1; j >= 0; --j) {
      const int c = j * nI + col;
      zNext = I.fz[c] - I.fLc[c] * zNext;
      I.fz[c] = zNext;
      rz += static_cast<double>(I.fr[c]) * static_cast<double>(zNext);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) rz += __shfl_down(rz, o, 64);
  (void)sm;
  if (threadIdx.x == 0) partialSlot(I, sCgRzLine)[blockIdx.x] = rz;
}

template <int PRE>
__global__ void __launch_bounds__(kBlock) k_cgf_init(const Inst* insts) {
  MX_PROLOGUE
  __shared__ double sm[16];
  double part[2] = {0.0, 0.0};
  if (live) {
    const float r = I.fb[c];
    I.fr[c] = r;
    I.fx[c] = 0.0f;
    if (PRE == 0) {
      const float d = I.fpP[c];
      const float z = (fabsf(d) > 0.0f) ? r / d : r;
      I.fz[c] = z;
      I.fpd0[c] = z;
      part[1] = static_cast<double>(r) * static_cast<double>(z);
    }
    part[0] = fabs(static_cast<double>(r));
  }
  blockSum<2>(part, sm);
  if (threadIdx.x == 0) {
    partialSlot(I, sCgNorm)[blockIdx.x] = part[0];
    if (PRE == 0) partialSlot(I, sCgRz)[blockIdx.x] = part[1];
    if (blockIdx.x == 0) {
      I.flags[fCgDone] = 0;
      I.flags[fCgIters] = 0;
    }
  }
}

// The line preconditioner's p0 = z is copied by the first launch A.
template <int PRE>
__global__ void __launch_bounds__(kBlock) k_cgf_a(const Inst* insts, int it) {
  MX_PROLOGUE
  if (I.flags[fCgDone]) return;
  __shared__ double sm[16];
  const double nrm = blockSumPartials(partialSlot(I, sCgNorm), I.nBlocks, sm);
  const double rzNew = (PRE == 0) ? blockSumPartials(partialSlot(I, sCgRz), I.nBlocks, sm)
                                  : blockSumPartials(partialSlot(I, sCgRzLine), I.lineBlocks, sm);
  double beta = 0.0;
  bool done = false;
  if (it == 0) {
    if (!(nrm > 0.0)) done = true;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
      I.red[sCgNorm0] = nrm;
      I.red[sCgRz0] = rzNew;
    }
  } else {
    const double rzOld = I.red[((it - 1) & 1) ? sCgRz1 : sCgRz0];
    if (nrm <= I.cgTol * I.red[sCgNorm0]) done = true;
    else if (!(fabs(rzOld) > 0.0)) done = true;
    else beta = rzNew / rzOld;
    if (threadIdx.x == 0 && blockIdx.x == 0 && !done) I.red[(it & 1) ? sCgRz1 : sCgRz0] = rzNew;
  }
  if (done) {
    if (threadIdx.x == 0 && blockIdx.x == 0) I.flags[fCgDone] = 1;
    return;
  }
  const float fb = static_cast<float>(beta);
  double part[1] = {0.0};
  if (live) {
    float* pNew = (it & 1) ? I.fpd1 : I.fpd0;
    const float* pOld = (it & 1) ? I.fpd0 : I.fpd1;
    float y, pc;
    if (it == 0) {
      if (PRE == 0) {
        pc = pNew[c];
        y = spmvF(I, pNew, c, i, j);
      } else {
        // p0 = z: the neighbours' z is final (line solve ran before)
        pc = I.fz[c];
        y = spmvF(I, I.fz, c, i, j);
        pNew[c] = pc;
      }
    } else {
      auto pv = [&](int q) { return fmaf(fb, pOld[q], I.fz[q]); };
      pc = pv(c);
      y = I.fpP[c] * pc;
      if (j > 0) y = fmaf(I.fpN[c - nI], pv(c - nI), y);
      else {
        const int pr = I.partner[i];
        if (pr >= 0) y = fmaf(I.fpWk[c], pv(pr), y);
      }
      if (j < nJ - 1) y = fmaf(I.fpN[c], pv(c + nI), y);
      if (i > 0) y = fmaf(I.fpE[c - 1], pv(c - 1), y);
      if (i < nI - 1) y = fmaf(I.fpE[c], pv(c + 1), y);
      pNew[c] = pc;
    }
    I.fAp[c] = y;
    part[0] = static_cast<double>(pc) * static_cast<double>(y);
  }
  blockSum<1>(part, sm);
  if (threadIdx.x == 0) partialSlot(I, sCgPAp)[blockIdx.x] = part[0];
}

template <int PRE>
__global__ void __launch_bounds__(kBlock) k_cgf_b(const Inst* insts, int it) {
  MX_PROLOGUE
  if (I.flags[fCgDone]) return;
  __shared__ double sm[16];
  const double pap = blockSumPartials(partialSlot(I, sCgPAp), I.nBlocks, sm);
  if (!(fabs(pap) > 0.0)) {
    if (threadIdx.x == 0 && blockIdx.x == 0) I.flags[fCgDone] = 1;
    return;
  }
  const float alpha = static_cast<float>(I.red[(it & 1) ? sCgRz1 : sCgRz0] / pap);
  double part[2] = {0.0, 0.0};
  if (live) {
    const float* p = (it & 1) ? I.fpd1 : I.fpd0;
    I.fx[c] = fmaf(alpha, p[c], I.fx[c]);
    const float r = fmaf(-alpha, I.fAp[c], I.fr[c]);
    I.fr[c] = r;
    part[0] = fabs(static_cast<double>(r));
    if (PRE == 0) {
      const float d = I.fpP[c];
      const float z = (fabsf(d) > 0.0f) ? r / d : r;
      I.fz[c] = z;
      part[1] = static_cast<double>(r) * static_cast<double>(z);
    }
  }
  blockSum<2>(part, sm);
  if (threadIdx.x == 0) {
    partialSlot(I, sCgNorm)[blockIdx.x] = part[0];
    if (PRE == 0) partialSlot(I, sCgRz)[blockIdx.x] = part[1];
    if (blockIdx.x == 0) {
      I.flags[fCgIters] = it + 1;
      I.flags[fCgTotal] += 1;
    }
  }
}

// x64 += x32 after a pass.
__global__ void __launch_bounds__(kBlock) k_cgf_accumulate(const Inst* insts) {
  MX_PROLOGUE
  if (!live) return;
  I.x[c] += static_cast<double>(I.fx[c]);
}

template __global__ void k_cgf_init<0>(const Inst*);
template __global__ void k_cgf_init<1>(const Inst*);
template __global__ void k_cgf_a<0>(const Inst*, int);
template __global__ void k_cgf_a<1>(const Inst*, int);
template __global__ void k_cgf_b<0>(const Inst*, int);
template __global__ void k_cgf_b<1>(const Inst*, int);

}  // namespace cfdgpu
