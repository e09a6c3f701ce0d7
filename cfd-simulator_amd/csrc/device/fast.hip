// fast.hip - kernels of the fast path (no reference counterpart; selected by
// cfdgpu_settings, judged against the FP64 path and the oracle):
//
//   * the pressure correction by CG preconditioned with the exact inverse of
//     each wall-normal column's tridiagonal part (line Jacobi), in the storage
//     and arithmetic type T (double, or float on the mixed path), with FP64
//     inner products and norms. The reference's Jacobi preconditioner
//     (LinearSystem.cpp:151-159) leaves the wall-normal coupling, which is up
//     to 1e8 times the streamwise one in the near-wall cells, to the Krylov
//     iteration; the line factor removes it. The CG's stopping test, cap and
//     breakdown guards are the reference's (LinearSystem.cpp:118-210).
//   * FP64-residual iterative refinement of the pressure correction: each
//     further pass solves A d = b - A x with the residual formed in FP64 from
//     the FP64 matrix, and adds d to x in FP64.
//   * the mixed path's two-colour Gauss-Seidel sweeps on FP32 defect systems:
//     the sweeps solve A delta = b - A phi_old for the correction delta in
//     FP32 (the defect and the coefficients rounded once from their FP64
//     assembly), and phi += delta is applied in FP64. In exact arithmetic the
//     iterates are those of the sweeps on phi itself; in FP32 the rounding
//     error scales with delta, which vanishes at convergence, so the fixed
//     point is the FP64 one.
//
// Column-parallel launches give one lane per wall-normal column i (cells
// c = j * nI + i, j = 0 .. nJ - 1), so a wave's accesses at one j are 64
// consecutive cells. The recurrence of the column solve runs along j in each
// lane.
#include "common.hpp"

namespace cfdgpu {

// The line CG's arrays in its storage type T, as global-address-space
// pointers: a generic (FLAT) access would make every wait for LDS traffic
// also wait for all loads in flight.
template <typename T>
struct LineStore {
  const CFD_GLOBAL T *pP, *pN, *pE, *pWk;
  CFD_GLOBAL T *inv, *up, *r, *z, *pd0, *pd1, *Ap, *x;
};

template <typename T>
__device__ __forceinline__ LineStore<T> lineStoreOf(const T* pP, const T* pN, const T* pE, const T* pWk, T* inv, T* up,
                                                  T* r, T* z, T* pd0, T* pd1, T* Ap, T* x) {
  return LineStore<T>{gmem(pP), gmem(pN), gmem(pE), gmem(pWk), gmem(inv), gmem(up), gmem(r), gmem(z), gmem(pd0),
                      gmem(pd1), gmem(Ap), gmem(x)};
}
template <typename T>
__device__ __forceinline__ LineStore<T> lineStore(const Inst& I);
template <>
__device__ __forceinline__ LineStore<double> lineStore<double>(const Inst& I) {
  return lineStoreOf<double>(I.pP, I.pN, I.pE, I.pWk, I.lInv, I.lUp, I.r, I.z, I.pd0, I.pd1, I.Ap, I.x);
}
// FP32 storage: the correction x too (its rounding error scales with the
// correction, which vanishes at convergence); lcgFinal hands it over in FP64
template <>
__device__ __forceinline__ LineStore<float> lineStore<float>(const Inst& I) {
  return lineStoreOf<float>(I.fpP, I.fpN, I.fpE, I.fpWk, I.flInv, I.flUp, I.fr, I.fz, I.fpd0, I.fpd1, I.fAp, I.fx);
}

__device__ __forceinline__ double waveSum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// One lane per column; the block is one wave.
#define LCG_COLUMN_PROLOGUE                                                                  \
  const Inst& I = insts[blockIdx.y];                                                         \
  if (static_cast<int>(blockIdx.x) >= I.colBlocks) return;                                   \
  if (I.flags[fStopped]) return;                                                             \
  const int i = static_cast<int>(blockIdx.x) * kColThreads + static_cast<int>(threadIdx.x); \
  const bool col = i < I.nI;                                                                 \
  const int nI = I.nI;                                                                       \
  const int nJ = I.nJ;

#define FAST_CELL_PROLOGUE                                            \
  const Inst& I = insts[blockIdx.y];                                  \
  if (static_cast<int>(blockIdx.x) >= I.nBlocks) return;             \
  if (I.flags[fStopped]) return;                                      \
  const int tile = xcdTile(static_cast<int>(blockIdx.x), I.nBlocks); \
  const int c = tile * kBlock + static_cast<int>(threadIdx.x);        \
  const bool live = c < I.N;                                          \
  const int nI = I.nI;                                                \
  const int nJ = I.nJ;                                                \
  const int i = live ? c % nI : 0;                                    \
  const int j = live ? c / nI : 0;

// ---------------------------------------------------------------------------
// Line-preconditioned CG
// ---------------------------------------------------------------------------

// A column sweep walks its nJ rows in blocks of kColUnroll. The operands of
// the next block are loaded before the current block is computed (rows past
// the end are clamped to the last row and not used), so the recurrence waits
// for memory once per block instead of once per row.
constexpr int kColUnroll = 16;  // start of a pass (the CG iteration's depth is the launch tuning's)

// Back substitution z_j = y_j - up_j z_{j+1}, j = nJ-1 .. 0, in place over z
// (which holds y), with up_j = pN_j / pivot_j stored by the factorisation.
template <typename T, int U>
__device__ __forceinline__ void lineBack(const LineStore<T>& S, int ic, bool col, int nI, int nJ) {
  T yA[U], uA[U];
  T yB[U] = {}, uB[U] = {};
  auto load = [&](int top, T (&y)[U], T (&up)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int c = max(top - u, 0) * nI + ic;
      y[u] = S.z[c];
      up[u] = S.up[c];
    }
  };
  T zNext = T(0);
  load(nJ - 1, yA, uA);
  for (int top = nJ - 1; top >= 0; top -= U) {
    if (top - U >= 0) load(top - U, yB, uB);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = top - u;
      if (j >= 0) {
        const T z = yA[u] - uA[u] * zNext;
        if (col) S.z[j * nI + ic] = z;
        zNext = z;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      yA[u] = yB[u];
      uA[u] = uB[u];
    }
  }
}

// Start of pass `pass`: r = b (pass 0: the assembled source, its mean removed
// when nothing anchors the pressure level, as the Jacobi solver does; later
// passes: the FP64 residual of the correction so far), x = 0, the factor of
// every column's tridiagonal part (pivot inverses), z = M^-1 r, and the
// partials of |r|_1 and r.z. The column tridiagonal of cell (i, j) has the
// diagonal aP and the wall-normal couplings pN of the faces below and above
// (the pressure matrix is symmetric); the wake cut and the streamwise
// couplings are left to the Krylov iteration.
template <typename T>
__device__ __forceinline__ void lcgInit(const Inst* insts, int pass) {
  LCG_COLUMN_PROLOGUE
  constexpr int U = kColUnroll;
  __shared__ double bc[1];
  const LineStore<T> S = lineStore<T>(I);
  const bool demean = pass == 0 && !I.hasPRef;
  if (demean) {
    const double tot = waveSumPartials(partialSlot(I, sSourceSum), I.nBlocks);
    if (threadIdx.x == 0) bc[0] = tot;
  }
  __syncthreads();
  const double mean = demean ? bc[0] / static_cast<double>(I.N) : 0.0;
  const CFD_GLOBAL double* src = gmem(static_cast<const double*>((pass == 0) ? I.pb : I.rb));
  CFD_GLOBAL double* pb = gmem(I.pb);
  const int ic = col ? i : nI - 1;  // idle lanes read a real column and store nothing
  double bA[U], bB[U] = {};
  T dA[U], nA[U], dB[U] = {}, nB[U] = {};
  auto load = [&](int j0, double (&b)[U], T (&d)[U], T (&n)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int c = min(j0 + u, nJ - 1) * nI + ic;
      b[u] = src[c];
      d[u] = S.pP[c];
      n[u] = S.pN[c];
    }
  };
  // r.z of z = M^-1 r accumulated in the forward substitution: with
  // w_j = r_j - l_j y_{j-1} (the unit lower factor's solve) and y_j = w_j /
  // pivot_j, r^T M^-1 r = sum_j w_j^2 / pivot_j = sum_j y_j w_j
  double l1 = 0.0, rz = 0.0;
  T cPrev = T(0), yPrev = T(0), nPrev = T(0);
  load(0, bA, dA, nA);
  for (int j0 = 0; j0 < nJ; j0 += U) {
    if (j0 + U < nJ) load(j0 + U, bB, dB, nB);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + u;
      if (j < nJ) {
        const int c = j * nI + ic;
        double b = bA[u];
        if (demean) {
          b -= mean;
          if (col) pb[c] = b;
        }
        const T r = static_cast<T>(b);
        const T lo = nPrev;  // the coupling with the row below (zero on row 0)
        const T den = dA[u] - lo * cPrev;
        const T inv = (fabs(static_cast<double>(den)) > 0.0) ? T(1) / den : T(1);
        const T w = r - lo * yPrev;
        const T y = w * inv;
        cPrev = nA[u] * inv;  // zero on the top row
        if (col) {
          S.inv[c] = inv;
          S.up[c] = cPrev;
          S.r[c] = r;
          S.z[c] = y;
          S.x[c] = T(0);
        }
        l1 += fabs(static_cast<double>(r));
        rz += static_cast<double>(y) * static_cast<double>(w);
        yPrev = y;
        nPrev = nA[u];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      bA[u] = bB[u];
      dA[u] = dB[u];
      nA[u] = nB[u];
    }
  }
  lineBack<T, kColUnroll>(S, ic, col, nI, nJ);
  l1 = waveSum(col ? l1 : 0.0);
  rz = waveSum(col ? rz : 0.0);
  if (threadIdx.x == 0) {
    partialSlot(I, sCgNorm)[blockIdx.x] = l1;
    partialSlot(I, sCgRz)[blockIdx.x] = rz;
    if (blockIdx.x == 0) {
      I.flags[fCgDone] = 0;
      I.flags[fCgIters] = 0;
    }
  }
}

// Iteration `it`, launch A (cell-parallel): the test of iteration it - 1
// from the column partials (|r|_1 <= tol |r0|_1, then the r.z guard), beta,
// p = z + beta p_old (neighbours recomputed locally), Ap and p.Ap.
// Each thread takes CELLS cells (c = tile * kBlock * CELLS + q * kBlock +
// threadIdx.x; 1, 2 or 4, the launch tuning's choice). Every operand is read through the global address space,
// without a branch around its load (a missing neighbour is read at the cell
// itself and its coefficient masked to zero), and issued before the stop
// flags are tested and the partials summed, so the stencil's loads and the
// partial sums are one memory round trip; only the lanes of row 0 wait again
// for their wake partner.
template <typename T, int CELLS>
__device__ __forceinline__ void lcgA(const Inst* insts, int it, double* sm) {
  const Inst& I = insts[blockIdx.y];
  const int nb = lineBlocksA(I.N, CELLS);
  if (static_cast<int>(blockIdx.x) >= nb) return;
  if (it >= I.cgMax) return;  // a batch member with a smaller cap than the batch's
  const int tile = xcdTile(static_cast<int>(blockIdx.x), nb);
  const int base = tile * kBlock * CELLS + static_cast<int>(threadIdx.x);
  const int nI = I.nI, nJ = I.nJ;
  const LineStore<T> S = lineStore<T>(I);
  const CFD_GLOBAL int* flagsG = gmem(static_cast<const int*>(I.flags));
  const CFD_GLOBAL double* red = gmem(static_cast<const double*>(I.red));
  const CFD_GLOBAL int* partnerG = gmem(static_cast<const int*>(I.partner));
  const CFD_GLOBAL T* z = S.z;
  const CFD_GLOBAL T* pOld = (it & 1) ? S.pd0 : S.pd1;
  int cq[CELLS], iq[CELLS], jq[CELLS], prq[CELLS];
  bool live[CELLS];
  T zc[CELLS], zS[CELLS], zN[CELLS], zW[CELLS], zE[CELLS];
  T oc[CELLS], oS[CELLS], oN[CELLS], oW[CELLS], oE[CELLS];
  T dP[CELLS], nB[CELLS], nA[CELLS], eL[CELLS], eR[CELLS], wk[CELLS];
#pragma unroll
  for (int q = 0; q < CELLS; ++q) {
    const int c0 = base + q * kBlock;
    live[q] = c0 < I.N;
    const int c = live[q] ? c0 : 0;
    const int i = c % nI, j = c / nI;
    cq[q] = c;
    iq[q] = i;
    jq[q] = j;
    prq[q] = partnerG[i];
    const int cS = (j > 0) ? c - nI : c;
    const int cN = (j < nJ - 1) ? c + nI : c;
    const int cW = (i > 0) ? c - 1 : c;
    const int cE = (i < nI - 1) ? c + 1 : c;
    zc[q] = z[c]; zS[q] = z[cS]; zN[q] = z[cN]; zW[q] = z[cW]; zE[q] = z[cE];
    oc[q] = oS[q] = oN[q] = oW[q] = oE[q] = T(0);
    if (it > 0) { oc[q] = pOld[c]; oS[q] = pOld[cS]; oN[q] = pOld[cN]; oW[q] = pOld[cW]; oE[q] = pOld[cE]; }
    dP[q] = S.pP[c]; nB[q] = S.pN[cS]; nA[q] = S.pN[c]; eL[q] = S.pE[cW]; eR[q] = S.pE[c];
    wk[q] = (j == 0) ? S.pWk[c] : T(0);  // the wake coupling exists on row 0 only
  }
  const int stopped = flagsG[fStopped];
  const int cgDone = flagsG[fCgDone];
  const int itersDone = flagsG[fCgIters];
  const double norm0 = red[sCgNorm0];
  const double rzOld = red[((it - 1) & 1) ? sCgRz1 : sCgRz0];
  const double alphaPrev = red[sCgAlpha];
  T xq[CELLS];
#pragma unroll
  for (int q = 0; q < CELLS; ++q) xq[q] = S.x[cq[q]];
  if (stopped) return;  // the whole block: the flags are uniform
  // x += alpha p of iteration it - 1 (launch B leaves it to this launch, whose
  // cells hold p_{it-1}), whatever this launch's test decides. This comes
  // before the test of cgDone: block 0 of this launch may already have set
  // it for a converged solve, and every block must still add its cells'
  // share of the last step (itersDone == it only when launch B of it - 1
  // made one).
  if (it > 0 && itersDone == it) {
#pragma unroll
    for (int q = 0; q < CELLS; ++q)
      if (live[q]) S.x[cq[q]] = xq[q] + static_cast<T>(alphaPrev) * oc[q];
  }
  if (cgDone) return;
  // row 0: the wake partner's values (a second round trip for those lanes only)
  T zPr[CELLS], oPr[CELLS];
  bool wake[CELLS];
#pragma unroll
  for (int q = 0; q < CELLS; ++q) {
    wake[q] = live[q] && jq[q] == 0 && prq[q] >= 0;
    zPr[q] = oPr[q] = T(0);
    if (wake[q]) {
      zPr[q] = z[prq[q]];
      if (it > 0) oPr[q] = pOld[prq[q]];
    }
  }
  double sums[2];
  {
    const double* const slots[2] = {partialSlot(I, sCgNorm), partialSlot(I, sCgRz)};
    blockSumPartials<2, 1>(slots, I.colBlocks, sm, sums);
  }
  const double nrm = sums[0], rzNew = sums[1];
  double beta = 0.0;
  bool done = false;
  if (it == 0) {
    if (!(nrm > 0.0)) done = true;  // already solved (LinearSystem.cpp:144-149)
    if (threadIdx.x == 0 && blockIdx.x == 0) {
      gmem(I.red)[sCgNorm0] = nrm;
      gmem(I.red)[sCgRz0] = rzNew;
    }
  } else {
    if (nrm <= I.cgTol * norm0) done = true;
    else if (!(fabs(rzOld) > 0.0)) done = true;
    else beta = rzNew / rzOld;
    if (threadIdx.x == 0 && blockIdx.x == 0 && !done) gmem(I.red)[(it & 1) ? sCgRz1 : sCgRz0] = rzNew;
  }
  if (done) {
    if (threadIdx.x == 0 && blockIdx.x == 0) gmem(I.flags)[fCgDone] = 1;
    return;
  }
  const T b = static_cast<T>(beta);
  auto pv = [&](T zv, T ov) -> T { return (it == 0) ? zv : static_cast<T>(zv + b * ov); };
  double part[1] = {0.0};
#pragma unroll
  for (int q = 0; q < CELLS; ++q) {
    const int i = iq[q], j = jq[q];
    // masked couplings: zero where the neighbour is missing
    const T aS = (j > 0) ? nB[q] : (wake[q] ? wk[q] : T(0));
    const T aN = (j < nJ - 1) ? nA[q] : T(0);
    const T aW = (i > 0) ? eL[q] : T(0);
    const T aE = (i < nI - 1) ? eR[q] : T(0);
    const T pc = pv(zc[q], oc[q]);
    const T pS = (j > 0) ? pv(zS[q], oS[q]) : pv(zPr[q], oPr[q]);
    // diagonal, then below (or the wake partner), above, left, right
    T y = dP[q] * pc;
    y += aS * pS;
    y += aN * pv(zN[q], oN[q]);
    y += aW * pv(zW[q], oW[q]);
    y += aE * pv(zE[q], oE[q]);
    if (live[q]) {
      ((it & 1) ? S.pd1 : S.pd0)[cq[q]] = pc;
      S.Ap[cq[q]] = y;
      part[0] += static_cast<double>(pc) * static_cast<double>(y);
    }
  }
  blockSum<1>(part, sm);
  if (threadIdx.x == 0) partialSlot(I, sCgPAp)[blockIdx.x] = part[0];
}

// Launch A on a tile of kTileW columns x kTileH rows per block (4 cells per
// thread; wave w owns rows w, w + 4, w + 8, w + 12 of the tile, a row's 64
// cells contiguous across its lanes): z and p_old of the tile and its
// one-cell halo are loaded once into LDS, p = z + beta p_old is formed there,
// and A p reads its neighbours' p from LDS, so a cell costs its own z, p_old,
// coefficients and x from memory instead of five z and five p_old loads.
// Same per-cell arithmetic as lcgA; the p.Ap partial sums are per tile.
template <typename T>
__device__ __forceinline__ void lcgATile(const Inst* insts, int it, double* sm, T (&sp)[kTileH + 2][kTileW + 2]) {
  const Inst& I = insts[blockIdx.y];
  const int nI = I.nI, nJ = I.nJ;
  const int tilesI = (nI + kTileW - 1) / kTileW;
  const int nb = tilesA(nI, nJ);
  if (static_cast<int>(blockIdx.x) >= nb) return;
  if (it >= I.cgMax) return;  // a batch member with a smaller cap than the batch's
  const int tile = xcdTile(static_cast<int>(blockIdx.x), nb);
  const int i0 = (tile % tilesI) * kTileW, j0 = (tile / tilesI) * kTileH;
  const int lane = static_cast<int>(threadIdx.x) & 63, wv = static_cast<int>(threadIdx.x) >> 6;
  const int i = i0 + lane;
  const LineStore<T> S = lineStore<T>(I);
  const CFD_GLOBAL int* flagsG = gmem(static_cast<const int*>(I.flags));
  const CFD_GLOBAL double* red = gmem(static_cast<const double*>(I.red));
  const CFD_GLOBAL T* z = S.z;
  const CFD_GLOBAL T* pOld = (it & 1) ? S.pd0 : S.pd1;
  constexpr int Q = kTileH / 4;
  int cq[Q], jq[Q];
  bool live[Q];
  T zc[Q], oc[Q], dP[Q], nB[Q], nA[Q], eL[Q], eR[Q], wk[Q], xq[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = j0 + wv + 4 * q;
    live[q] = i < nI && j < nJ;
    const int c = live[q] ? j * nI + i : 0;
    const int jj = live[q] ? j : 0, ii = live[q] ? i : 0;
    cq[q] = c;
    jq[q] = jj;
    const int cS = (jj > 0) ? c - nI : c;
    const int cW = (ii > 0) ? c - 1 : c;
    zc[q] = live[q] ? z[c] : T(0);
    oc[q] = (live[q] && it > 0) ? pOld[c] : T(0);
    dP[q] = S.pP[c]; nB[q] = S.pN[cS]; nA[q] = S.pN[c]; eL[q] = S.pE[cW]; eR[q] = S.pE[c];
    wk[q] = (jj == 0) ? S.pWk[c] : T(0);  // the wake coupling exists on row 0 only
    xq[q] = S.x[c];
  }
  // the halo: the rows below and above the tile (waves 0 and 1, a lane per
  // column) and the columns left and right of it (wave 2, a lane per row);
  // cells outside the mesh hold zero
  T hz = T(0), ho = T(0);
  int hr = -1, hc = -1;
  if (wv < 2) {
    const int j = (wv == 0) ? j0 - 1 : j0 + kTileH;
    hr = (wv == 0) ? 0 : kTileH + 1;
    hc = lane + 1;
    if (i < nI && j >= 0 && j < nJ) {
      hz = z[j * nI + i];
      if (it > 0) ho = pOld[j * nI + i];
    }
  } else if (wv == 2 && lane < 2 * kTileH) {
    const int r = lane % kTileH, side = lane / kTileH;
    const int ii = side ? i0 + kTileW : i0 - 1, j = j0 + r;
    hr = r + 1;
    hc = side ? kTileW + 1 : 0;
    if (ii >= 0 && ii < nI && j < nJ) {
      hz = z[j * nI + ii];
      if (it > 0) ho = pOld[j * nI + ii];
    }
  }
  // row 0's wake partners (tiles on row 0 only)
  const int pr = (j0 == 0 && wv == 0 && i < nI) ? I.partner[i] : -1;
  const int stopped = flagsG[fStopped];
  const int cgDone = flagsG[fCgDone];
  const int itersDone = flagsG[fCgIters];
  const double norm0 = red[sCgNorm0];
  const double rzOld = red[((it - 1) & 1) ? sCgRz1 : sCgRz0];
  const double alphaPrev = red[sCgAlpha];
  if (stopped) return;  // the whole block: the flags are uniform
  // x += alpha p of iteration it - 1 (see lcgA)
  if (it > 0 && itersDone == it) {
#pragma unroll
    for (int q = 0; q < Q; ++q)
      if (live[q]) S.x[cq[q]] = xq[q] + static_cast<T>(alphaPrev) * oc[q];
  }
  if (cgDone) return;
  T zPr = T(0), oPr = T(0);
  if (pr >= 0) {
    zPr = z[pr];
    if (it > 0) oPr = pOld[pr];
  }
  double sums[2];
  {
    const double* const slots[2] = {partialSlot(I, sCgNorm), partialSlot(I, sCgRz)};
    blockSumPartials<2, 1>(slots, I.colBlocks, sm, sums);
  }
  const double nrm = sums[0], rzNew = sums[1];
  double beta = 0.0;
  bool done = false;
  if (it == 0) {
    if (!(nrm > 0.0)) done = true;  // already solved (LinearSystem.cpp:144-149)
    if (threadIdx.x == 0 && blockIdx.x == 0) {
      gmem(I.red)[sCgNorm0] = nrm;
      gmem(I.red)[sCgRz0] = rzNew;
    }
  } else {
    if (nrm <= I.cgTol * norm0) done = true;
    else if (!(fabs(rzOld) > 0.0)) done = true;
    else beta = rzNew / rzOld;
    if (threadIdx.x == 0 && blockIdx.x == 0 && !done) gmem(I.red)[(it & 1) ? sCgRz1 : sCgRz0] = rzNew;
  }
  if (done) {
    if (threadIdx.x == 0 && blockIdx.x == 0) gmem(I.flags)[fCgDone] = 1;
    return;
  }
  const T b = static_cast<T>(beta);
  auto pv = [&](T zv, T ov) -> T { return (it == 0) ? zv : static_cast<T>(zv + b * ov); };
  T pc[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    pc[q] = pv(zc[q], oc[q]);
    sp[1 + wv + 4 * q][1 + lane] = pc[q];
  }
  if (hr >= 0) sp[hr][hc] = pv(hz, ho);
  const T pPr = pv(zPr, oPr);
  __syncthreads();
  double part[1] = {0.0};
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int r = 1 + wv + 4 * q;
    const int jj = jq[q];
    const bool wake = live[q] && jj == 0 && pr >= 0;
    // masked couplings: zero where the neighbour is missing
    const T aS = (jj > 0) ? nB[q] : (wake ? wk[q] : T(0));
    const T aN = (jj < nJ - 1) ? nA[q] : T(0);
    const T aW = (i > 0) ? eL[q] : T(0);
    const T aE = (i < nI - 1) ? eR[q] : T(0);
    const T pS = (jj > 0) ? sp[r - 1][1 + lane] : pPr;
    // diagonal, then below (or the wake partner), above, left, right
    T y = dP[q] * pc[q];
    y += aS * pS;
    y += aN * sp[r + 1][1 + lane];
    y += aW * sp[r][lane];
    y += aE * sp[r][lane + 2];
    if (live[q]) {
      ((it & 1) ? S.pd1 : S.pd0)[cq[q]] = pc[q];
      S.Ap[cq[q]] = y;
      part[0] += static_cast<double>(pc[q]) * static_cast<double>(y);
    }
  }
  blockSum<1>(part, sm);
  if (threadIdx.x == 0) partialSlot(I, sCgPAp)[blockIdx.x] = part[0];
}

// Iteration `it`, launch B (column-parallel): alpha = r.z / p.Ap (kept for
// the next launch A, which adds alpha p to x), r -= alpha Ap, |r|_1,
// z = M^-1 r (forward and back substitution along the column), r.z.
template <typename T, int U>
__device__ __forceinline__ void lcgB(const Inst* insts, int it, int cellsA) {
  LCG_COLUMN_PROLOGUE
  __shared__ double bc[1];
  if (it >= I.cgMax) return;
  if (I.flags[fCgDone]) return;
  const LineStore<T> S = lineStore<T>(I);
  const double papLane = waveSumPartials<4>(partialSlot(I, sCgPAp), blocksA(I.nI, I.nJ, cellsA));
  if (threadIdx.x == 0) bc[0] = papLane;
  __syncthreads();
  const double pap = bc[0];
  if (!(fabs(pap) > 0.0)) {  // the direction has no component the matrix acts on
    if (threadIdx.x == 0 && blockIdx.x == 0) I.flags[fCgDone] = 1;
    return;
  }
  const double alpha = I.red[(it & 1) ? sCgRz1 : sCgRz0] / pap;
  const T alphaT = static_cast<T>(alpha);
  const int ic = col ? i : nI - 1;
  T aA[U], rA[U], iA[U], nA[U];
  T aB[U] = {}, rB[U] = {}, iB[U] = {}, nB[U] = {};
  auto load = [&](int j0, T (&av)[U], T (&rv)[U], T (&iv)[U], T (&nv)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int c = min(j0 + u, nJ - 1) * nI + ic;
      av[u] = S.Ap[c];
      rv[u] = S.r[c];
      iv[u] = S.inv[c];
      nv[u] = S.pN[c];
    }
  };
  double l1 = 0.0, rz = 0.0;  // r.z as in the start of the pass
  T yPrev = T(0), nPrev = T(0);
  load(0, aA, rA, iA, nA);
  for (int j0 = 0; j0 < nJ; j0 += U) {
    if (j0 + U < nJ) load(j0 + U, aB, rB, iB, nB);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + u;
      if (j < nJ) {
        const int c = j * nI + ic;
        const T r = rA[u] - alphaT * aA[u];
        const T w = r - nPrev * yPrev;
        const T y = w * iA[u];
        if (col) {
          S.r[c] = r;
          S.z[c] = y;
        }
        l1 += fabs(static_cast<double>(r));
        rz += static_cast<double>(y) * static_cast<double>(w);
        yPrev = y;
        nPrev = nA[u];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      aA[u] = aB[u];
      rA[u] = rB[u];
      iA[u] = iB[u];
      nA[u] = nB[u];
    }
  }
  lineBack<T, U>(S, ic, col, nI, nJ);
  l1 = waveSum(col ? l1 : 0.0);
  rz = waveSum(col ? rz : 0.0);
  if (threadIdx.x == 0) {
    partialSlot(I, sCgNorm)[blockIdx.x] = l1;
    partialSlot(I, sCgRz)[blockIdx.x] = rz;
    if (blockIdx.x == 0) {
      I.red[sCgAlpha] = alpha;
      I.flags[fCgIters] = it + 1;
      I.flags[fCgTotal] += 1;
    }
  }
}

// Launch B with every column's rows split over the waves of a block: one
// block per 64-column tile, wave w owning rows [w L, w L + L) of its 64
// columns, L = kSegRows, all of them in registers
// (loads along a row are contiguous across the wave's lanes). The forward
// solve y_j = w_j / pivot_j, w_j = r_j - l_j y_{j-1}, is the affine recurrence
// y_j = a_j y_{j-1} + b_j (a_j = -l_j inv_j, b_j = r_j inv_j): each wave
// composes its rows' maps, the value entering each segment is composed from
// the segments below it through LDS, and each wave then runs the recurrence
// over its rows from that value; the back substitution z_j = y_j - up_j
// z_{j+1} is done the same way from the top. A column's 2 nJ dependent steps
// become 2 (L + kSegWaves) short ones, and r and z are written once (y stays
// in registers). Same arithmetic per row as lcgB, with the entering values
// composed in another order.
constexpr int kSegWaves = kLcgSegThreads / 64;            // waves per block
constexpr int kSegRows = kLcgSegMaxRows / kSegWaves;      // rows per wave at most
static_assert(kSegWaves * kSegRows == kLcgSegMaxRows, "segmented launch B shape");

template <typename T>
__device__ __forceinline__ void lcgBSeg(const Inst* insts, int it, int cellsA) {
  const Inst& I = insts[blockIdx.y];
  if (static_cast<int>(blockIdx.x) >= I.colBlocks) return;
  if (I.flags[fStopped]) return;
  if (it >= I.cgMax) return;
  if (I.flags[fCgDone]) return;
  __shared__ double bc[1];
  __shared__ T segA[kSegWaves][64];
  __shared__ T segB[kSegWaves][64];
  __shared__ double sums[2][kSegWaves];
  const int lane = static_cast<int>(threadIdx.x) & 63, wv = static_cast<int>(threadIdx.x) >> 6;
  const int nI = I.nI, nJ = I.nJ;
  const int i = static_cast<int>(blockIdx.x) * 64 + lane;
  const bool col = i < nI;
  const int ic = col ? i : nI - 1;  // idle lanes read a real column and store nothing
  const LineStore<T> S = lineStore<T>(I);
  // segments of kSegRows rows: a column of nJ rows occupies the first
  // ceil(nJ / kSegRows) waves, and the waves above its top row load nothing
  // (their maps are the identity)
  const int j0 = wv * kSegRows;
  const int n = max(0, min(nJ, j0 + kSegRows) - j0);  // this wave's rows
  // every operand of this wave's rows, issued before the partial sums. The
  // back substitution's factor up_j = pN_j inv_j is formed here from the
  // coupling above row j (lo of row j + 1, or pNTop for the wave's last row)
  // as the start of the pass formed it, instead of being read.
  T ap[kSegRows] = {}, rr[kSegRows] = {}, lo[kSegRows] = {}, iv[kSegRows] = {};
  T pNTop = T(0);
  if (n > 0) {
#pragma unroll
    for (int u = 0; u < kSegRows; ++u) {
      const int j = min(j0 + u, nJ - 1);
      const int c = j * nI + ic;
      ap[u] = S.Ap[c];
      rr[u] = S.r[c];
      iv[u] = S.inv[c];
      lo[u] = S.pN[(j > 0 ? j - 1 : 0) * nI + ic];
    }
    pNTop = S.pN[(j0 + n - 1) * nI + ic];
  }
  if (wv == 0) {
    const double papLane = waveSumPartials<4>(partialSlot(I, sCgPAp), blocksA(I.nI, I.nJ, cellsA));
    if (lane == 0) bc[0] = papLane;
  }
  __syncthreads();
  const double pap = bc[0];
  if (!(fabs(pap) > 0.0)) {  // the direction has no component the matrix acts on
    if (threadIdx.x == 0 && blockIdx.x == 0) I.flags[fCgDone] = 1;
    return;
  }
  const double alpha = I.red[(it & 1) ? sCgRz1 : sCgRz0] / pap;
  const T alphaT = static_cast<T>(alpha);
  // r -= alpha Ap; the segment's composed forward map y_out = A y_in + B
  T A = T(1), B = T(0);
#pragma unroll
  for (int u = 0; u < kSegRows; ++u) {
    if (u < n) {
      rr[u] = rr[u] - alphaT * ap[u];
      if (j0 + u == 0) lo[u] = T(0);  // row 0 has no row below
      const T a = -lo[u] * iv[u];
      const T b = rr[u] * iv[u];
      A = a * A;
      B = a * B + b;
    }
  }
  segA[wv][lane] = A;
  segB[wv][lane] = B;
  __syncthreads();
  T y = T(0);  // y of the row below this wave's first row
  for (int q = 0; q < wv; ++q) y = segA[q][lane] * y + segB[q][lane];
  double l1 = 0.0, rz = 0.0;
  T yv[kSegRows];
#pragma unroll
  for (int u = 0; u < kSegRows; ++u) {
    yv[u] = T(0);
    if (u < n) {
      const T w = rr[u] - lo[u] * y;
      y = w * iv[u];
      yv[u] = y;
      if (col) S.r[(j0 + u) * nI + ic] = rr[u];
      l1 += fabs(static_cast<double>(rr[u]));
      rz += static_cast<double>(y) * static_cast<double>(w);
    }
  }
  // back substitution: the segment's composed map z_bottom = C z_above + D
  T up[kSegRows];
#pragma unroll
  for (int u = 0; u < kSegRows; ++u) up[u] = ((u + 1 < n) ? lo[u + 1] : pNTop) * iv[u];
  T Cm = T(1), Dm = T(0);
#pragma unroll
  for (int u = kSegRows - 1; u >= 0; --u) {
    if (u < n) {
      const T cu = -up[u];
      Cm = cu * Cm;
      Dm = cu * Dm + yv[u];
    }
  }
  __syncthreads();  // every wave has read the forward maps
  segA[wv][lane] = Cm;
  segB[wv][lane] = Dm;
  __syncthreads();
  T z = T(0);  // z of the row above this wave's last row
  for (int q = kSegWaves - 1; q > wv; --q) z = segA[q][lane] * z + segB[q][lane];
#pragma unroll
  for (int u = kSegRows - 1; u >= 0; --u) {
    if (u < n) {
      z = yv[u] - up[u] * z;
      if (col) S.z[(j0 + u) * nI + ic] = z;
    }
  }
  l1 = waveSum(col ? l1 : 0.0);
  rz = waveSum(col ? rz : 0.0);
  if (lane == 0) {
    sums[0][wv] = l1;
    sums[1][wv] = rz;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tl = 0.0, tr = 0.0;
    for (int q = 0; q < kSegWaves; ++q) {
      tl += sums[0][q];
      tr += sums[1][q];
    }
    partialSlot(I, sCgNorm)[blockIdx.x] = tl;
    partialSlot(I, sCgRz)[blockIdx.x] = tr;
    if (blockIdx.x == 0) {
      I.red[sCgAlpha] = alpha;
      I.flags[fCgIters] = it + 1;
      I.flags[fCgTotal] += 1;
    }
  }
}

// After the CG iterations of a pass: the x update of the last launch B, when
// no launch A followed it (the cap was reached); with FP32 storage also the
// hand-over of x to the FP64 correction the velocity and flux updates read.
template <typename T>
__device__ __forceinline__ void lcgFinal(const Inst* insts) {
  const Inst& I = insts[blockIdx.y];
  const int nq = quadBlocks(I.nBlocks);
  if (static_cast<int>(blockIdx.x) >= nq) return;
  if (I.flags[fStopped]) return;
  const int tile = xcdTile(static_cast<int>(blockIdx.x), nq);
  const int n = I.flags[fCgIters];
  const bool capped = n == I.cgMax && n != 0;
  const LineStore<T> S = lineStore<T>(I);
  const CFD_GLOBAL T* p = ((n - 1) & 1) ? S.pd1 : S.pd0;
  const double alpha = I.red[sCgAlpha];
  CFD_GLOBAL double* x = gmem(I.x);
  // four cells per thread (quadBlocks), every load issued before the stores
  T pv[kQuad], xv[kQuad];
#pragma unroll
  for (int q = 0; q < kQuad; ++q) {
    const int c = min((tile * kQuad + q) * kBlock + static_cast<int>(threadIdx.x), I.N - 1);
    pv[q] = p[c];
    xv[q] = S.x[c];
  }
#pragma unroll
  for (int q = 0; q < kQuad; ++q) {
    const int c = (tile * kQuad + q) * kBlock + static_cast<int>(threadIdx.x);
    if (c >= I.N) continue;
    if constexpr (sizeof(T) == sizeof(double)) {
      // (S.x is the FP64 correction itself)
      if (capped) x[c] = static_cast<double>(xv[q]) + alpha * static_cast<double>(pv[q]);
    } else {
      // the FP32 accumulation handed over in FP64, with the last step in FP64
      const double last = capped ? alpha * static_cast<double>(pv[q]) : 0.0;
      x[c] = static_cast<double>(xv[q]) + last;
    }
  }
}

// The launches in the storage type T, with the shapes the launch tuning picks
// from: launch A with 1, 2 or 4 cells per thread, launch B walking its columns
// 16 or 32 rows ahead.
#define CFD_LCG_KERNELS(SUFFIX, T)                                                                     \
  __global__ void __launch_bounds__(kColThreads) k_lcg_init##SUFFIX(const Inst* insts, int pass) {      \
    lcgInit<T>(insts, pass);                                                                           \
  }                                                                                                    \
  __global__ void __launch_bounds__(kBlock) k_lcg_a##SUFFIX##_c1(const Inst* insts, int it) {           \
    __shared__ double sm[16];                                                                          \
    lcgA<T, 1>(insts, it, sm);                                                                         \
  }                                                                                                    \
  __global__ void __launch_bounds__(kBlock) k_lcg_a##SUFFIX##_c2(const Inst* insts, int it) {           \
    __shared__ double sm[16];                                                                          \
    lcgA<T, 2>(insts, it, sm);                                                                         \
  }                                                                                                    \
  __global__ void __launch_bounds__(kBlock) k_lcg_a##SUFFIX##_c4(const Inst* insts, int it) {           \
    __shared__ double sm[16];                                                                          \
    lcgA<T, 4>(insts, it, sm);                                                                         \
  }                                                                                                    \
  __global__ void __launch_bounds__(kBlock) k_lcg_a##SUFFIX##_tile(const Inst* insts, int it) {         \
    __shared__ double sm[16];                                                                          \
    __shared__ T sp[kTileH + 2][kTileW + 2];                                                           \
    lcgATile<T>(insts, it, sm, sp);                                                                    \
  }                                                                                                    \
  __global__ void __launch_bounds__(kColThreads) k_lcg_b##SUFFIX##_u16(const Inst* insts, int it, int cellsA) { \
    lcgB<T, 16>(insts, it, cellsA);                                                                    \
  }                                                                                                    \
  __global__ void __launch_bounds__(kColThreads) k_lcg_b##SUFFIX##_u32(const Inst* insts, int it, int cellsA) { \
    lcgB<T, 32>(insts, it, cellsA);                                                                    \
  }                                                                                                    \
  __global__ void __launch_bounds__(kSegWaves * 64) k_lcg_b##SUFFIX##_seg(const Inst* insts, int it, int cellsA) { \
    lcgBSeg<T>(insts, it, cellsA);                                                                     \
  }                                                                                                    \
  __global__ void __launch_bounds__(kBlock) k_lcg_final##SUFFIX(const Inst* insts) { lcgFinal<T>(insts); }

CFD_LCG_KERNELS(, double)
CFD_LCG_KERNELS(_f32, float)

// ---------------------------------------------------------------------------
// FP64-residual refinement of the pressure correction
// ---------------------------------------------------------------------------

// After pass `pass`: the accumulated correction takes the pass's solution.
__global__ void __launch_bounds__(kBlock) k_refine_acc(const Inst* insts, int pass) {
  FAST_CELL_PROLOGUE
  (void)i; (void)j; (void)nI; (void)nJ;
  if (live) I.xacc[c] = (pass == 0) ? I.x[c] : I.xacc[c] + I.x[c];
}

// The next pass's right-hand side b - A x in FP64 (the FP64 pressure matrix,
// the accumulated correction), and its L1 norm's partials.
__global__ void __launch_bounds__(kBlock) k_refine_res(const Inst* insts) {
  FAST_CELL_PROLOGUE
  __shared__ double sm[16];
  double part[1] = {0.0};
  if (live) {
    const double* x = I.xacc;
    double y = I.pP[c] * x[c];
    if (j > 0) {
      y += I.pN[c - nI] * x[c - nI];
    } else {
      const int pr = I.partner[i];
      if (pr >= 0) y += I.pWk[c] * x[pr];
    }
    if (j < nJ - 1) y += I.pN[c] * x[c + nI];
    if (i > 0) y += I.pE[c - 1] * x[c - 1];
    if (i < nI - 1) y += I.pE[c] * x[c + 1];
    const double r = I.pb[c] - y;
    I.rb[c] = r;
    part[0] = fabs(r);
  }
  blockSum<1>(part, sm);
  if (threadIdx.x == 0) partialSlot(I, sRefineNorm)[blockIdx.x] = part[0];
}

// After the last pass: the correction the velocity and flux updates read.
__global__ void __launch_bounds__(kBlock) k_refine_final(const Inst* insts) {
  FAST_CELL_PROLOGUE
  (void)i; (void)j; (void)nI; (void)nJ;
  if (live) I.x[c] = I.xacc[c] + I.x[c];
}

// ---------------------------------------------------------------------------
// Mixed path: two-colour Gauss-Seidel on FP32 defect systems
// ---------------------------------------------------------------------------

// One colour of one sweep. sys 0: u and v (one matrix); sys 1: k and omega.
// `first`: the first colour of the first sweep, whose neighbours' corrections
// are all zero; `last`: the second colour of the last sweep, which also adds
// every cell's correction to its field (k and omega then floored at 0 and
// 1e-10, KOmegaSST.cpp:468-471). Off-diagonal terms in cellFaces order
// (below or the wake partner, right, above, left; Mesh.cpp:206-213); a row
// with a zero diagonal is skipped (LinearSystem.cpp:88-90): its correction is
// zero. The arrays are colour-packed (rbIndex): a launch reads its colour's
// rows contiguously and the other colour's corrections at the neighbours.
// The C-grid has an even nI, so a wake cell and its partner differ in colour.
// Every operand is loaded without a branch around it (a missing neighbour is
// read at the cell's own place and its term dropped by a select; a zero
// diagonal's row computed and discarded), so a lane's loads are one round
// trip. The sum is the one the branches would form: off starts at +0 and a
// sum of floats is -0 only when both terms are, so adding +0 for a dropped
// term leaves it unchanged.
template <int NS>
__device__ __forceinline__ void rbDefect(const float* const (&A)[NS][5], const float* const (&R)[NS],
                                         float* const (&D)[NS], int pk, const int (&nb)[4], const bool (&has)[4],
                                         bool first, float (&out)[NS]) {
  float d[NS], a[NS][4], rhs[NS], dn[NS][4];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    d[s] = A[s][0][pk];
    rhs[s] = R[s][pk];
    // the coefficients in cellFaces order: below, right, above, left
    a[s][0] = A[s][1][pk];
    a[s][1] = A[s][4][pk];
    a[s][2] = A[s][2][pk];
    a[s][3] = A[s][3][pk];
#pragma unroll
    for (int k = 0; k < 4; ++k) dn[s][k] = first ? 0.0f : D[s][nb[k]];
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    float off = 0.0f;
    if (!first) {
#pragma unroll
      for (int k = 0; k < 4; ++k) off += has[k] ? a[s][k] * dn[s][k] : 0.0f;
    }
    out[s] = (fabsf(d[s]) > 0.0f) ? (rhs[s] - off) / d[s] : 0.0f;
  }
}

// One colour of one sweep, one lane per cell of that colour (the colour-packed
// order, so every lane of a wave has a row to solve and its loads are
// contiguous). The launch that ends the last sweep also applies the other
// colour's corrections, lane q taking the other colour's cell q.
template <int NS>
__device__ __forceinline__ void rbSolveCell(const float* const (&A)[NS][5], const float* const (&R)[NS],
                                            float* const (&D)[NS], const Inst& I, int i, int j, int pk, bool first,
                                            float (&out)[NS]) {
  const int nI = I.nI, nJ = I.nJ;
  // the four neighbours in cellFaces order: below (or the wake partner), right, above, left
  int nb[4] = {pk, pk, pk, pk};
  bool has[4] = {false, false, false, false};
  if (!first) {
    if (j > 0) {
      nb[0] = rbIndex(i, j - 1, nI, nJ);
      has[0] = true;
    } else {
      const int pr = I.partner[i];
      if (pr >= 0) {
        nb[0] = rbIndex(pr, 0, nI, nJ);
        has[0] = true;
      }
    }
    has[1] = i < nI - 1;
    has[2] = j < nJ - 1;
    has[3] = i > 0;
    nb[1] = has[1] ? rbIndex(i + 1, j, nI, nJ) : pk;
    nb[2] = has[2] ? rbIndex(i, j + 1, nI, nJ) : pk;
    nb[3] = has[3] ? rbIndex(i - 1, j, nI, nJ) : pk;
  }
  rbDefect<NS>(A, R, D, pk, nb, has, first, out);
#pragma unroll
  for (int s = 0; s < NS; ++s) D[s][pk] = out[s];
}

__device__ __forceinline__ void rbApply(const Inst& I, int sys, int c, const float (&d)[2]) {
  if (sys == 0) {
    I.u[c] += static_cast<double>(d[0]);
    I.v[c] += static_cast<double>(d[1]);
  } else {
    I.k[c] = mx(I.k[c] + static_cast<double>(d[0]), 0.0);
    I.w[c] = mx(I.w[c] + static_cast<double>(d[1]), 1.0e-10);
  }
}

__global__ void __launch_bounds__(kBlock) k_rb32(const Inst* insts, int sys, int color, int first, int last) {
  const Inst& I = insts[blockIdx.y];
  const int nI = I.nI, nJ = I.nJ, N = I.N;
  const int count0 = rbRowBase(nJ, 0, nI);  // cells of colour 0
  const int nb = rbBlocks(nI, nJ);
  if (static_cast<int>(blockIdx.x) >= nb) return;
  if (I.flags[fStopped]) return;
  const int q0 = xcdTile(static_cast<int>(blockIdx.x), nb) * (kBlock * kRbCells) + static_cast<int>(threadIdx.x);
  const int countC = color ? N - count0 : count0;
  const int countO = color ? count0 : N - count0;
  const float* const Am[2][5] = {{I.fMom[0], I.fMom[1], I.fMom[2], I.fMom[3], I.fMom[4]},
                                 {I.fMom[0], I.fMom[1], I.fMom[2], I.fMom[3], I.fMom[4]}};
  const float* const At[2][5] = {{I.fTurb[0][0], I.fTurb[0][1], I.fTurb[0][2], I.fTurb[0][3], I.fTurb[0][4]},
                                 {I.fTurb[1][0], I.fTurb[1][1], I.fTurb[1][2], I.fTurb[1][3], I.fTurb[1][4]}};
  const float* const R[2] = {sys == 0 ? I.fMomRes[0] : I.fTurbRes[0], sys == 0 ? I.fMomRes[1] : I.fTurbRes[1]};
  float* const D[2] = {sys == 0 ? I.fMomDelta[0] : I.fTurbDelta[0], sys == 0 ? I.fMomDelta[1] : I.fTurbDelta[1]};
#pragma unroll
  for (int k = 0; k < kRbCells; ++k) {
    const int q = q0 + k * kBlock;
    if (q < countC) {
      int i, j;
      rbCell(q, color, nI, i, j);
      const int pk = (color ? count0 : 0) + q;
      float d[2];
      // u and v share their coefficients (sys 0); k and omega have their own
      if (sys == 0) rbSolveCell<2>(Am, R, D, I, i, j, pk, first != 0, d);
      else rbSolveCell<2>(At, R, D, I, i, j, pk, first != 0, d);
      if (last) rbApply(I, sys, j * nI + i, d);
    }
    if (last && q < countO) {
      int i, j;
      rbCell(q, 1 - color, nI, i, j);
      const int pk = (color ? 0 : count0) + q;
      const float d[2] = {D[0][pk], D[1][pk]};
      rbApply(I, sys, j * nI + i, d);
    }
  }
}

}  // namespace cfdgpu
