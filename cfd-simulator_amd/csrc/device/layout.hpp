// layout.hpp - device data layout of one SIMPLE instance on a structured C-grid.
//
// Cells are i-major: c = j*nI + i (nI cells streamwise, nJ wall-normal), the
// reference's own numbering (include/cfd/mesh/Mesh.hpp:167-168). Faces come in
// two families, both row-major so a wavefront of 64 lanes walking i reads
// contiguous memory:
//   j-faces (constant j, between rows j-1 and j): index j*nI + i, j in [0, nJ]
//   i-faces (constant i, between columns i-1 and i): index j*(nI+1) + i, i in [0, nI]
// The reference stores the i-family column-major (src/mesh/Mesh.cpp:153-156);
// the host repacks it on upload and back on download.
//
// Boundary data lives on the four edges only: bottom (j = 0, nI faces, which
// holds the wall and both sides of the wake cut), top (j = nJ), left (i = 0)
// and right (i = nI). A wake-cut face in the bottom row couples cell (i, 0) to
// cell (partner[i], 0) (src/mesh/CGrid.cpp:370-381).
#pragma once

#include <hip/hip_runtime.h>

namespace cfdgpu {

// SimpleSolver::FaceGeometry (include/cfd/solver/SimpleSolver.hpp:199-205),
// precomputed on the host in FP64 exactly as SimpleSolver.cpp:82-131 does.
struct FaceGeo {
  double Sx, Sy;  // S = n |S|, out of the owner
  double Tx, Ty;  // T = S - g d, the non-orthogonal remainder
  double g;       // |S|^2 / (S.d)
  double w;       // owner interpolation weight
};

// The descriptor's device arrays: in device code they point into the global
// address space, so every access through them is a GLOBAL instruction. A
// FLAT (generic) access also counts against the scalar-and-LDS wait counter,
// so every wait for a descriptor field or an LDS result would also wait for
// all loads in flight. Host code sees plain pointers of the same size.
#if defined(__HIP_DEVICE_COMPILE__)
#define CFD_G __attribute__((address_space(1)))
#else
#define CFD_G
#endif

// In device memory each FaceGeo field is an array of its own, so a kernel
// reads only the fields it uses (the SST assembly only w, the gradients S and
// w) and every field load is contiguous across a wave.
struct FaceGeoArrays {
  CFD_G const double *Sx, *Sy, *Tx, *Ty, *g, *w;
};
__host__ __device__ __forceinline__ FaceGeo geo(const FaceGeoArrays& a, int f) {
  return FaceGeo{a.Sx[f], a.Sy[f], a.Tx[f], a.Ty[f], a.g[f], a.w[f]};
}

// Per boundary face: the FaceCondition (BoundaryConditions.hpp:101-109) plus
// the frozen far-field dead band (SimpleSolver.cpp:150-166) and the k-omega
// SST geometric factor |S|^2/|S.d| (KOmegaSST.cpp:423-427).
struct FaceBC {
  double u, v, p;
  double dead;
  double gSst;
  int kind;  // flow::BoundaryKind
  int pad;
};

enum BKind { kInternal = 0, kInlet = 1, kOutlet = 2, kFarField = 3, kWall = 4 };

// Reduction slots of one instance (FP64 scalars in device memory).
enum Slot {
  sVelScale = 0,   // max |u|
  sMomResX,        // sum |b - A u|
  sMomResY,
  sMomDiagL1,      // sum |a_P|
  sNetFluxL1,      // sum |netFlux|
  sSourceSum,      // sum b (mean removal when unanchored)
  sCgNorm0,        // CG: initial L1 residual
  sCgNorm,         // CG: current L1 residual
  sCgRz,           // CG: r.z
  sCgAlpha,
  sCgBeta,
  sCgPAp,
  sCorrMean,       // mean of the pressure correction (unanchored)
  sMaxDiv,         // max |div|
  sTurbResK,       // SST residual numerators and denominators
  sTurbDiagK,
  sTurbResW,
  sTurbDiagW,
  sMinK, sMaxK, sMinW, sMaxW, sMinMut, sMaxMut,
  sMaxEddyRatio,
  sMaxYPlus,
  sRefineNorm,     // mixed path: FP64 residual norm of the pressure correction
  sCgRz0, sCgRz1,  // CG: r.z of even / odd iterations (persisted by block 0)
  sCgRzLine,       // CG: r.z partials of the line-preconditioner launch
  kSlots
};

// Integer state of one instance.
enum IFlag {
  fCgDone = 0,     // CG finished (converged or broke down)
  fCgIters,        // iterations of the current CG solve
  fCgTotal,        // iterations over all passes of this outer iteration
  fStopped,        // solve finished: later launches for this instance are no-ops
  fIteration,      // completed outer iterations
  fStatus,         // 0 running/converged, 1 iteration-limit, 2 diverged
  fSweepStall,     // waits of the decoupled sweep that gave up (0 unless broken)
  kIFlags
};

// Monitor written by the device each outer iteration (SolverMonitor,
// include/cfd/solver/SimpleSolver.hpp:135-154).
struct DevMonitor {
  double continuity, momentumX, momentumY;
  double massImbalance, maxDivergence;
  int pressureIterations;
  int pad;
  double turbFirst, turbSecond;
  double minK, maxK, minW, maxW, minMut, maxMut;
  double maxEddyRatio;
};

constexpr int kMaxRedBlocks = 4096;  // partial-sum capacity per slot

// Skewed layout of the lexicographic sweep: cell (i, j) at row i + j (its
// wavefront step) plus kSkewFront, column j, of rows R wide. The sweep reads
// and writes row k + kSkewFront on step k for every k it visits, from D steps
// before its first (the read-ahead fill) to 2 D + 1 past the last, without
// clamping: kSkewFront and kSkewTail rows of padding cover them. Slots of no
// cell (i outside [0, nI), or j >= nJ) are zero when allocated and stay zero:
// their matrix row is zero, so a sweep stores back the old value. Lanes of a
// block wider than R (a batch member narrower than the widest) store into
// kSkewPark parking slots past the last row, which nothing reads.
constexpr int kSkewFront = 16;    // >= the sweeps' read-ahead depth D
constexpr int kSkewTail = 34;     // >= 2 D + 2
constexpr int kSkewPark = 1024;   // >= the widest block's lanes
// Lanes of the sweep for nJ rows: one per row, rounded up to whole waves.
__host__ __device__ constexpr int gsLanes(int nJ) {
  return nJ <= 64 ? 64 : (nJ >= 1024 ? 1024 : ((nJ + 63) / 64) * 64);
}
__host__ __device__ constexpr int skewSteps(int nI, int nJ) { return nI + nJ - 1; }
__host__ __device__ constexpr int skewIndex(int i, int j, int R) { return (i + j + kSkewFront) * R + j; }
__host__ __device__ constexpr int skewParkIndex(int nI, int nJ, int R, int lane) {
  return (kSkewFront + skewSteps(nI, nJ) + kSkewTail) * R + (lane - R);
}
__host__ __device__ constexpr int skewLength(int nI, int nJ, int R) {
  return (kSkewFront + skewSteps(nI, nJ) + kSkewTail) * R + kSkewPark;
}

// The mixed path's two-colour sweep arrays hold each colour contiguously, so
// a launch of one colour reads only that colour's rows: colour C = (i + j) & 1
// (colour 0 first), rows in j order, a row's cells of that colour in i order
// (i = i0, i0 + 2, ... with i0 = (C + j) & 1, so the cell's place in its row is
// i >> 1). rbRowBase counts the colour-C cells of rows 0 .. j - 1 (two
// consecutive rows hold nI of them).
__host__ __device__ inline int rbRowBase(int j, int C, int nI) {
  const int pairs = (j >> 1) * nI;
  if (!(j & 1)) return pairs;
  const int i0 = (C + j - 1) & 1;  // first colour-C column of row j - 1
  return pairs + (nI - i0 + 1) / 2;
}
__host__ __device__ inline int rbIndex(int i, int j, int nI, int nJ) {
  const int C = (i + j) & 1;
  return (C ? rbRowBase(nJ, 0, nI) : 0) + rbRowBase(j, C, nI) + (i >> 1);
}
// The cell (i, j) of colour C at place q of that colour's packed order
// (q < its count; the inverse of rbIndex within one colour): two consecutive
// rows 2r, 2r + 1 hold nI cells of colour C, the first (nI - C + 1) / 2 of
// them in row 2r from column C, the rest in row 2r + 1 from column 1 - C.
__host__ __device__ inline void rbCell(int q, int C, int nI, int& i, int& j) {
  const int r = q / nI;
  const int rem = q - r * nI;
  const int n0 = (nI - C + 1) / 2;
  if (rem < n0) {
    j = 2 * r;
    i = C + 2 * rem;
  } else {
    j = 2 * r + 1;
    i = (1 - C) + 2 * (rem - n0);
  }
}

constexpr int kBlock = 256;          // threads per block for cell-parallel kernels
// The two-colour sweep's lanes each take kRbCells cells of the swept colour
// (kBlock apart; 1 measured faster than 2); blocks of such a launch over the
// larger colour.
constexpr int kRbCells = 1;
__host__ __device__ inline int rbBlocks(int nI, int nJ) {
  const int c0 = rbRowBase(nJ, 0, nI), c1 = nI * nJ - c0;
  const int m = c0 > c1 ? c0 : c1;
  return (m + kBlock * kRbCells - 1) / (kBlock * kRbCells);
}
constexpr int kCgCells = 2;          // cells per thread in the CG launches (1, 2, 4, 8 measured: 2 is fastest)
constexpr int kColThreads = 64;      // columns per block of a column-parallel launch (one wave)
// Launches with four cells per thread (cells (t kQuad + q) kBlock + threadIdx.x
// of tile t, q = 0 .. kQuad - 1): kernels with few loads per cell, whose
// waves would otherwise wait on one memory round trip each
constexpr int kQuad = 4;
__host__ __device__ constexpr int quadBlocks(int nBlocks) { return (nBlocks + kQuad - 1) / kQuad; }
constexpr int kLcgSegThreads = 1024; // the line CG's segmented launch B: 16 waves over one 64-column tile
constexpr int kLcgSegMaxRows = 256;  // ... each wave holding at most 16 rows
constexpr int kNoAllocMaxCells = 1 << 20;  // cells of cfdgpu_divergence_noalloc's staging array (8 MiB)
// blocks of the line CG's launch A for n cells at `cells` cells per thread
__host__ __device__ constexpr int lineBlocksA(int n, int cells) { return (n + kBlock * cells - 1) / (kBlock * cells); }
// the LDS-tiled launch A (cells per thread 0 in the launch tuning): tiles of
// kTileW columns x kTileH rows, four cells per thread
constexpr int kTileW = 64, kTileH = 16;
static_assert(kTileW * kTileH == 4 * kBlock, "launch A tile");
__host__ __device__ constexpr int tilesA(int nI, int nJ) {
  return ((nI + kTileW - 1) / kTileW) * ((nJ + kTileH - 1) / kTileH);
}
// launch A's blocks (and p.Ap partial sums) for an nI x nJ mesh
__host__ __device__ constexpr int blocksA(int nI, int nJ, int cells) {
  return cells == 0 ? tilesA(nI, nJ) : lineBlocksA(nI * nJ, cells);
}

struct Inst {
  int nI, nJ, N;
  int hasPRef;     // a boundary sets the pressure level (SimpleSolver.cpp:71-77)
  int scheme;      // 0 upwind, 1 second-order upwind (deferred correction)
  int model;       // 0 none, 1 laminar, 2 SST
  double alphaU, alphaP, blend, alphaT;
  int momSweeps, cgMax;
  double cgTol;
  int nonOrthPasses;  // max(1, nonOrthogonalCorrectors)
  int gsOrder;        // 0 lexicographic (exact), 1 two-colour
  int precision;      // 0 FP64, 1 mixed (FP32 bulk, FP64 reductions and residuals)
  int pSolver;        // 0 Jacobi-PCG (the reference's), 1 wall-normal line-preconditioned CG
  int refinePasses;   // pressure solves per corrector pass (> 1: FP64-residual refinement)
  double refMassFlow;
  double stopTol;  // device-side stop rule (worst() < stopTol); <= 0 disables
  int stopMaxIter;
  int nBlocks;     // blocks of a cell-parallel launch for this instance
  int cgBlocks;    // blocks of a CG launch (kBlock x kCgCells cells each)
  int colBlocks;   // blocks of a column-parallel launch (kColThreads columns each)

  // geometry (read-only)
  CFD_G const double* V;
  FaceGeoArrays jg;  // j-faces
  FaceGeoArrays ig;  // i-faces
  CFD_G const int* partner;  // nI entries: partner cell in row 0, or -1
  CFD_G const FaceBC* bcB;   // nI
  CFD_G const FaceBC* bcT;   // nI
  CFD_G const FaceBC* bcL;   // nJ
  CFD_G const FaceBC* bcR;   // nJ
  CFD_G const double* wallDist;
  CFD_G const double* jgSst;  // |g| for SST on j-faces (interior and wake)
  CFD_G const double* igSst;

  // flow state (FP64)
  CFD_G double *u, *v, *p, *rho, *mu, *muMol;
  CFD_G double *Fj, *Fi;
  double2 *gU, *gV, *gP;

  // momentum system (cell-centric rows; aP is the relaxed diagonal = momentumDiagonal_)
  CFD_G double *aP, *aS, *aN, *aW, *aE, *bx, *by;
  // pressure system: diagonal, coefficient of the above / right face, wake coefficient
  CFD_G double *pP, *pN, *pE, *pWk, *pb;
  // CG vectors
  CFD_G double *x, *r, *z, *pd0, *pd1, *Ap;
  // ping-pong buffers of the two-colour sweep's callers (unused by the lexicographic one)
  CFD_G double *gsTmp0, *gsTmp1;

  // Lexicographic sweep in the anti-diagonal ("skewed") layout: cell (i, j) is
  // stored at skewIndex(i, j, gsR), so the cells of one wavefront step are
  // contiguous and a wave's loads coalesce. The assembly kernels write the
  // systems here; the sweeps ping-pong between x0 and x1 and the last sweep
  // writes the row-major field.
  int gsR;        // lanes of the sweep (rows rounded up to 64) = stride of the layout
  int gsSkewLen;  // elements of each skewed array: skewLength(nI, nJ, gsR)
  // Matrix rows in pairs, so a lane fetches a row with three 16-byte loads:
  // row[0] = (aP, aS), row[1] = (aN, aW), row[2] = (aE, b); momentum keeps
  // b_y in its own array. Unknowns in pairs too: (u, v) for momentum, (phi, -)
  // for k and omega; x[0] holds the starting values, x[1] the ping-pong copy.
  double2* skMomRow[3];
  CFD_G double* skMomBy;
  double2* skMomX[2];
  double2* skTurbRow[2][3];  // [k, omega][pair]
  double2* skTurbX[2][2];    // [k, omega][copy]

  // k-omega SST state and systems
  CFD_G double *k, *w, *mut, *f1, *f2, *cd, *strain, *prod;
  double2 *gK, *gW;
  CFD_G double *kP, *kS, *kN, *kW, *kE, *kB;
  CFD_G double *wP, *wS, *wN, *wW, *wE, *wB;
  double kIn, wIn;
  double sigK1, sigW1, beta1, sigK2, sigW2, beta2, betaStar, kappa, a1, prodLimit;
  double gamma1, gamma2, betaStarQ;  // gamma1(), gamma2(), betaStar^(1/4) (host-computed)

  // mixed precision: the pressure system and the CG vectors stored in FP32
  // (the Jacobi CG's arithmetic, its reductions and x stay FP64; the line CG
  // accumulates x in FP32, fx, and hands it over in FP64 at the end)
  CFD_G float *fpP, *fpN, *fpE, *fpWk;
  CFD_G float *fr, *fz, *fpd0, *fpd1, *fAp, *fx;

  // Line-preconditioned CG: 1 / pivot of each column's tridiagonal factor
  // (Thomas) and the eliminated upper coefficient (pN / pivot), in the CG's
  // storage type; the refinement's FP64 residual right-hand side and
  // accumulated correction.
  CFD_G double *lInv, *lUp;
  CFD_G float *flInv, *flUp;
  CFD_G double *rb, *xacc;

  // Mixed path with two-colour sweeps: every swept system in FP32 as
  // coefficients (aP, aS, aN, aW, aE), the FP64 defect b - A phi rounded to
  // FP32, and the FP32 correction the sweeps solve for (phi += delta after
  // the last sweep, in FP64). Momentum: u and v share the coefficients.
  // Cell (i, j) of these arrays is at rbIndex(i, j, nI, nJ).
  CFD_G float* fMom[5];
  CFD_G float* fMomRes[2];
  CFD_G float* fMomDelta[2];
  CFD_G float* fTurb[2][5];     // [k, omega][coefficient]
  CFD_G float* fTurbRes[2];
  CFD_G float* fTurbDelta[2];

  // reductions
  CFD_G double* red;        // kSlots scalars
  CFD_G double* partials;   // kSlots * kMaxRedBlocks
  CFD_G unsigned* counter;  // one arrival counter per slot group
  CFD_G int* flags;         // kIFlags
  CFD_G DevMonitor* monitor;
  CFD_G DevMonitor* history;  // optional ring of monitors, one per outer iteration
  int historyCap;
};

}  // namespace cfdgpu
