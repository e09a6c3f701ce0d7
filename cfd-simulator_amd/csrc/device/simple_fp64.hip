// simple_fp64.hip - FP64 kernels of one SIMPLE outer iteration on a structured
// C-grid (reference: src/solver/SimpleSolver.cpp:662-794).
//
// Every kernel is cell-centric: one lane per cell gathers the contributions
// of its four faces (below, above, left, right) in the order the reference's
// face loop scatters them (faces are visited by index: the j-family first,
// then the i-family, so a cell sees below, above, left, right). With the same
// per-face arithmetic and FP contraction disabled, each cell's sums are the
// reference's sums bit for bit; only the global reductions (L1 norms, dot
// products) are re-associated, as tree sums.
//
// Grid: blockIdx.y = instance of a batch, blockIdx.x = tile of 256 cells.
// Every kernel first checks the instance's stop flag, so a batch can keep
// launching while finished instances idle.
#include "common.hpp"

namespace cfdgpu {

#define CFD_PROLOGUE                                          \
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
// Face interpolation + Green-Gauss gradients of u, v, p
// (SimpleSolver.cpp:268-317, Gradient.cpp:20-50) and the velocity scale
// max |u| (SimpleSolver.cpp:677-683).
// ---------------------------------------------------------------------------

struct Face3 {
  double u, v, p;
};

__device__ __forceinline__ Face3 faceValues(const Inst& I, int c, int o, int n, const FaceGeo& f) {
  const double w = f.w;
  (void)c;
  return Face3{w * I.u[o] + (1.0 - w) * I.u[n], w * I.v[o] + (1.0 - w) * I.v[n],
               w * I.p[o] + (1.0 - w) * I.p[n]};
}

__device__ __forceinline__ Face3 boundaryValues(const Inst& I, int c, const FaceBC& bc, double F) {
  if (imposes(bc, F)) return Face3{bc.u, bc.v, I.p[c]};
  return Face3{I.u[c], I.v[c], bc.p};
}

__global__ void __launch_bounds__(kBlock) k_grad(const Inst* insts) {
  CFD_PROLOGUE
  __shared__ double sm[16];
  double vmax[1] = {0.0};
  if (live) {
    double ux = 0.0, uy = 0.0, vx = 0.0, vy = 0.0, px = 0.0, py = 0.0;
    auto add = [&](const FaceGeo& f, const Face3& q, bool plus) {
      if (plus) {
        ux += f.Sx * q.u; uy += f.Sy * q.u;
        vx += f.Sx * q.v; vy += f.Sy * q.v;
        px += f.Sx * q.p; py += f.Sy * q.p;
      } else {
        ux -= f.Sx * q.u; uy -= f.Sy * q.u;
        vx -= f.Sx * q.v; vy -= f.Sy * q.v;
        px -= f.Sx * q.p; py -= f.Sy * q.p;
      }
    };
    // below: j-face (i, j)
    {
      const FaceGeo f = geo(I.jg, j * nI + i);
      if (j > 0) {
        add(f, faceValues(I, c, c - nI, c, f), false);
      } else {
        const int pr = I.partner[i];
        if (pr >= 0) add(f, faceValues(I, c, c, pr, f), true);
        else add(f, boundaryValues(I, c, I.bcB[i], I.Fj[i]), true);
      }
    }
    // above: j-face (i, j+1), owned by this cell
    {
      const int fi = (j + 1) * nI + i;
      const FaceGeo f = geo(I.jg, fi);
      if (j < nJ - 1) add(f, faceValues(I, c, c, c + nI, f), true);
      else add(f, boundaryValues(I, c, I.bcT[i], I.Fj[fi]), true);
    }
    // left: i-face (i, j)
    {
      const int fi = j * (nI + 1) + i;
      const FaceGeo f = geo(I.ig, fi);
      if (i > 0) add(f, faceValues(I, c, c - 1, c, f), false);
      else add(f, boundaryValues(I, c, I.bcL[j], I.Fi[fi]), true);
    }
    // right: i-face (i+1, j), owned by this cell
    {
      const int fi = j * (nI + 1) + i + 1;
      const FaceGeo f = geo(I.ig, fi);
      if (i < nI - 1) add(f, faceValues(I, c, c, c + 1, f), true);
      else add(f, boundaryValues(I, c, I.bcR[j], I.Fi[fi]), true);
    }
    const double vol = I.V[c];
    if (vol > 0.0) {
      const double inv = 1.0 / vol;
      ux *= inv; uy *= inv; vx *= inv; vy *= inv; px *= inv; py *= inv;
    }
    st2(I.gU, c, ux, uy);
    st2(I.gV, c, vx, vy);
    st2(I.gP, c, px, py);
    vmax[0] = mx(0.0, hypot(I.u[c], I.v[c]));
  }
  blockMax<1>(vmax, sm);
  if (threadIdx.x == 0) partialSlot(I, sVelScale)[blockIdx.x] = vmax[0];
}

// ---------------------------------------------------------------------------
// Momentum assembly for both components + implicit relaxation + the L1
// residual measured before the sweeps (SimpleSolver.cpp:319-443, 685-696).
// ---------------------------------------------------------------------------

struct MomRow {
  double d, bx, by;
};

// An interior face seen from one of its two cells.
__device__ __forceinline__ double momInterior(const Inst& I, const FaceGeo& f, double F, int o, int n,
                                              bool cellIsOwner, MomRow& row) {
  const double w = f.w;
  const double visc = w * I.mu[o] + (1.0 - w) * I.mu[n];
  const double diff = visc * f.g;
  const double outward = mx(F, 0.0);
  const double inward = mx(-F, 0.0);
  double off;
  if (cellIsOwner) {
    row.d += outward + diff;
    off = -inward - diff;
  } else {
    row.d += inward + diff;
    off = -outward - diff;
  }
  const double2 guo = ld2(I.gU, o), gun = ld2(I.gU, n), gvo = ld2(I.gV, o), gvn = ld2(I.gV, n);
  const double fux = guo.x * w + gun.x * (1.0 - w), fuy = guo.y * w + gun.y * (1.0 - w);
  const double fvx = gvo.x * w + gvn.x * (1.0 - w), fvy = gvo.y * w + gvn.y * (1.0 - w);
  const double cu = I.blend * visc * (fux * f.Tx + fuy * f.Ty);
  const double cv = I.blend * visc * (fvx * f.Tx + fvy * f.Ty);
  if (cellIsOwner) { row.bx += cu; row.by += cv; }
  else { row.bx -= cu; row.by -= cv; }
  if (I.scheme == 1) {
    const double uo = I.u[o], un = I.u[n], vo = I.v[o], vn = I.v[n];
    const double upU = (F >= 0.0) ? uo : un;
    const double upV = (F >= 0.0) ? vo : vn;
    const double cenU = w * uo + (1.0 - w) * un;
    const double cenV = w * vo + (1.0 - w) * vn;
    const double defU = F * (cenU - upU);
    const double defV = F * (cenV - upV);
    if (cellIsOwner) { row.bx -= defU; row.by -= defV; }
    else { row.bx += defU; row.by += defV; }
  }
  return off;
}

// A wake-cut face, owned by cell c; `a` is the cell across the cut.
__device__ __forceinline__ double momWake(const Inst& I, const FaceGeo& f, double F, int c, int a, MomRow& row) {
  const double w = f.w;
  const double visc = w * I.mu[c] + (1.0 - w) * I.mu[a];
  const double diff = visc * f.g;
  const double outward = mx(F, 0.0);
  const double inward = mx(-F, 0.0);
  row.d += outward + diff;
  const double off = -inward - diff;
  const double2 guo = ld2(I.gU, c), gun = ld2(I.gU, a), gvo = ld2(I.gV, c), gvn = ld2(I.gV, a);
  const double fux = guo.x * w + gun.x * (1.0 - w), fuy = guo.y * w + gun.y * (1.0 - w);
  const double fvx = gvo.x * w + gvn.x * (1.0 - w), fvy = gvo.y * w + gvn.y * (1.0 - w);
  row.bx += I.blend * visc * (fux * f.Tx + fuy * f.Ty);
  row.by += I.blend * visc * (fvx * f.Tx + fvy * f.Ty);
  return off;
}

__device__ __forceinline__ void momBoundary(const Inst& I, const FaceGeo& f, const FaceBC& bc, double F, int c,
                                            MomRow& row) {
  const double visc = I.mu[c];
  const double diff = visc * f.g;
  if (imposes(bc, F)) {
    const double2 gu = ld2(I.gU, c), gv = ld2(I.gV, c);
    row.bx -= F * bc.u;
    row.by -= F * bc.v;
    row.d += diff;
    row.bx += diff * bc.u;
    row.by += diff * bc.v;
    row.bx += I.blend * visc * (gu.x * f.Tx + gu.y * f.Ty);
    row.by += I.blend * visc * (gv.x * f.Tx + gv.y * f.Ty);
  } else {
    row.d += F;
  }
}

__global__ void __launch_bounds__(kBlock) k_mom(const Inst* insts) {
  CFD_PROLOGUE
  __shared__ double sm[16];
  double part[3] = {0.0, 0.0, 0.0};
  if (live) {
    MomRow row{0.0, 0.0, 0.0};
    double aS = 0.0, aN = 0.0, aW = 0.0, aE = 0.0;
    int mirror = -1;
    // below
    {
      const FaceGeo f = geo(I.jg, j * nI + i);
      const double F = I.Fj[j * nI + i];
      if (j > 0) {
        aS = momInterior(I, f, F, c - nI, c, false, row);
      } else {
        mirror = I.partner[i];
        if (mirror >= 0) aS = momWake(I, f, F, c, mirror, row);
        else momBoundary(I, f, I.bcB[i], F, c, row);
      }
    }
    // above
    {
      const int fi = (j + 1) * nI + i;
      const FaceGeo f = geo(I.jg, fi);
      const double F = I.Fj[fi];
      if (j < nJ - 1) aN = momInterior(I, f, F, c, c + nI, true, row);
      else momBoundary(I, f, I.bcT[i], F, c, row);
    }
    // left
    {
      const int fi = j * (nI + 1) + i;
      const FaceGeo f = geo(I.ig, fi);
      const double F = I.Fi[fi];
      if (i > 0) aW = momInterior(I, f, F, c - 1, c, false, row);
      else momBoundary(I, f, I.bcL[j], F, c, row);
    }
    // right
    {
      const int fi = j * (nI + 1) + i + 1;
      const FaceGeo f = geo(I.ig, fi);
      const double F = I.Fi[fi];
      if (i < nI - 1) aE = momInterior(I, f, F, c, c + 1, true, row);
      else momBoundary(I, f, I.bcR[j], F, c, row);
    }
    const double2 gp = ld2(I.gP, c);
    const double vol = I.V[c];
    row.bx -= gp.x * vol;
    row.by -= gp.y * vol;
    // implicit under-relaxation (SimpleSolver.cpp:425-443)
    const double uc = I.u[c], vc = I.v[c];
    double d = row.d;
    if (!(fabs(d) > 0.0)) {
      d = 1.0;
    } else {
      const double relaxed = d / I.alphaU;
      row.bx += (relaxed - d) * uc;
      row.by += (relaxed - d) * vc;
      d = relaxed;
    }
    I.aP[c] = d;  // the momentum diagonal of Rhie-Chow and the correction
    if (I.gsOrder == 1 && I.precision == 0) {
      // the FP64 two-colour sweep's rows
      I.aS[c] = aS;
      I.aN[c] = aN;
      I.aW[c] = aW;
      I.aE[c] = aE;
      I.bx[c] = row.bx;
      I.by[c] = row.by;
    } else if (I.gsOrder == 1) {
      // the mixed path's FP32 rows, colour-packed (the defects follow below)
      const int pk = rbIndex(i, j, nI, nJ);
      I.fMom[0][pk] = static_cast<float>(d);
      I.fMom[1][pk] = static_cast<float>(aS);
      I.fMom[2][pk] = static_cast<float>(aN);
      I.fMom[3][pk] = static_cast<float>(aW);
      I.fMom[4][pk] = static_cast<float>(aE);
    } else {
      // the lexicographic sweep's copy, anti-diagonal layout, in pairs
      const int sk = skewIndex(i, j, I.gsR);
      I.skMomRow[0][sk] = make_double2(d, aS);
      I.skMomRow[1][sk] = make_double2(aN, aW);
      I.skMomRow[2][sk] = make_double2(aE, row.bx);
      I.skMomBy[sk] = row.by;
      I.skMomX[0][sk] = make_double2(uc, vc);
    }
    // residual with the starting field (LinearSystem.cpp:18-54, face order)
    double axu = d * uc, axv = d * vc;
    if (j > 0) { axu += aS * I.u[c - nI]; axv += aS * I.v[c - nI]; }
    else if (mirror >= 0) { axu += aS * I.u[mirror]; axv += aS * I.v[mirror]; }
    if (j < nJ - 1) { axu += aN * I.u[c + nI]; axv += aN * I.v[c + nI]; }
    if (i > 0) { axu += aW * I.u[c - 1]; axv += aW * I.v[c - 1]; }
    if (i < nI - 1) { axu += aE * I.u[c + 1]; axv += aE * I.v[c + 1]; }
    part[0] = fabs(row.bx - axu);
    part[1] = fabs(row.by - axv);
    part[2] = fabs(d);
    if (I.gsOrder == 1 && I.precision == 1) {
      const int pk = rbIndex(i, j, nI, nJ);
      I.fMomRes[0][pk] = static_cast<float>(row.bx - axu);
      I.fMomRes[1][pk] = static_cast<float>(row.by - axv);
    }
  }
  blockSum<3>(part, sm);
  if (threadIdx.x == 0) {
    partialSlot(I, sMomResX)[blockIdx.x] = part[0];
    partialSlot(I, sMomResY)[blockIdx.x] = part[1];
    partialSlot(I, sMomDiagL1)[blockIdx.x] = part[2];
  }
}

// ---------------------------------------------------------------------------
// Lexicographic Gauss-Seidel, exactly (LinearSystem.cpp:77-116).
//
// The reference sweeps cells in index order c = j*nI + i, in place. Cell (i,j)
// then sees new values below and left and old values above and right. All
// cells on one anti-diagonal i + j = k are independent given diagonal k-1, so
// a sweep is nI + nJ - 1 dependent steps with lane j owning row j. The wake
// cut couples (i,0) to (p,0) in the same row, which is walked in i order by
// lane 0 - exactly the reference's order - so the schedule reproduces the
// reference bit for bit for any partner map. Off-diagonal terms are summed in
// the reference's cellFaces order: below, right, above, left (Mesh.cpp:206-213).
//
// One workgroup per component; u and v share one matrix.
// ---------------------------------------------------------------------------

struct GsSys {
  const double *aP, *aS, *aN, *aW, *aE;
  const double* b[2];
  double* x[2];
  double floorValue;  // applied after the sweeps when hasFloor
  int hasFloor;
};

// One system of the lexicographic sweep in the skewed layout (see Inst::gsR):
// matrix rows as three pairs (aP, aS), (aN, aW), (aE, b[0]) plus b[1] for
// momentum, and the two ping-pong copies of the unknowns as pairs (x[0] holds
// the starting values, written by the assembly kernel; after n sweeps the
// result is in x[n & 1], which k_gs_unskew copies to the field). All pointers
// are global-space (see gmem), so a wait on LDS never also waits for HBM
// traffic.
template <int NC>
struct GsSkew {
  const CFD_GLOBAL dpair* row[3];
  const CFD_GLOBAL double* b1;  // second right-hand side (NC == 2)
  CFD_GLOBAL dpair* x[2];
};

// Lexicographic Gauss-Seidel on the anti-diagonal schedule: lane j owns row j
// and handles cell (k - j, j) on step k; row 0 takes its wake partner's
// current value from an LDS copy of the row, updated in place by lane 0, so
// the partner coupling follows the reference's order for any partner map.
// One component per workgroup, and the waves decoupled: inside a wave the
// value below comes from the lane underneath by a whole-wave DPP shift of the
// previous step's results, so no LDS round trip and no block barrier separate
// two steps. Only the row at a wave boundary
// crosses waves: lane 63 of wave w posts its value for step k in an LDS ring
// (kSweepHand slots) and then raises the wave's step counter; wave w + 1
// reads it on step k + 1 once the counter shows k. A wave runs ahead of the
// one above it by at most kSweepHand - 16 steps (it checks the counter above
// every 16 steps before reusing ring slots). Every wait is a bounded spin: if
// the bound is hit the sweep gives up and counts a stall in
// flags[fSweepStall], so no wave can hang the device. The step has no
// divergent branch: writes that only one lane makes go, on the other lanes,
// to a per-lane dummy slot. Off-diagonal terms are summed in the reference's
// cellFaces order (below, right, above, left; Mesh.cpp:206-213). A missing
// neighbour's coefficient is zero in the assembled row, so its term adds a
// +-0 product to a sum that is never -0, which leaves the sum bit for bit as
// the reference's skip does.
constexpr int kSweepHand = 64;  // a power of two
constexpr int kSpinLimit = 1 << 22;

// Waits until *counter >= target; the counter is read as a wave-uniform value.
__device__ __forceinline__ bool waitStep(const int* counter, int target, int* stall) {
  int spins = 0;
#pragma unroll 1
  while (__builtin_amdgcn_readfirstlane(ldsPoll(counter)) < target) {
    if (++spins > kSpinLimit) {
      // global address space, not FLAT: a FLAT access in flight would make
      // every later LDS wait also wait for all outstanding global loads
      __hip_atomic_fetch_add(gmem(stall), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    spinPause();
  }
  // The poll is a relaxed atomic load; the slot it guards is read with a plain
  // load. This fence (no instruction) keeps the compiler from moving that read
  // above the poll.
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return true;
}

// One component of a skewed system: the matrix rows (aP, aS), (aN, aW),
// (aE, b) as pairs, and the unknowns as one half of the pair arrays. With
// SepB the right-hand side comes from its own array (v of momentum) and only
// aE is read from the third pair.
struct GsOne {
  const CFD_GLOBAL dpair* row[3];
  const CFD_GLOBAL double* b;
  CFD_GLOBAL double* x[2];  // the pair arrays as doubles: unknown q at 2 q + half
  int half;
};
struct GsCell1 {
  double aP, aS, aN, aW, aE, b, xC, xN, xE;
};

// A wave-uniform base plus a 32-bit byte offset per lane: the form of a
// global access with the base in scalar registers and one offset register.
template <class T>
__device__ __forceinline__ T ldOff(const CFD_GLOBAL char* base, unsigned off) {
  return *reinterpret_cast<const CFD_GLOBAL T*>(base + off);
}
template <class T>
__device__ __forceinline__ void stOff(CFD_GLOBAL char* base, unsigned off, T v) {
  *reinterpret_cast<CFD_GLOBAL T*>(base + off) = v;
}

// One sweep's arrays as byte bases; the unknowns' bases already point at
// this component's half of each pair, so a slot's byte offset (16 bytes a
// slot) addresses the matrix row and the unknown alike.
struct GsBytes {
  const CFD_GLOBAL char *row0, *row1, *row2, *b;
  const CFD_GLOBAL char* xOld;
  CFD_GLOBAL char* xNew;
};

// Fetch the lane's cell from slot byte offset o, with oE = o + 16 R the same
// lane one row on: the matrix row, the right-hand side, and the old values of
// the cell, the cell to the right (oE) and the cell above (oE + 16). Every
// value is loaded at its own width, none with an unused half: a load whose
// result is partly unused leaves registers the compiler reuses at once, and
// reusing a register a load is still writing forces a wait for every load in
// flight. Every lane loads on every step (the padding rows keep the offsets
// in bounds); values of cells off the wavefront are never used.
template <bool SepB>
__device__ __forceinline__ void gsFetch1(const GsBytes& P, unsigned o, unsigned oE, GsCell1& g) {
  const dpair r0 = ldOff<dpair>(P.row0, o);
  const dpair r1 = ldOff<dpair>(P.row1, o);
  g.aP = r0[0];
  g.aS = r0[1];
  g.aN = r1[0];
  g.aW = r1[1];
  if constexpr (SepB) {
    g.aE = ldOff<double>(P.row2, o);
    g.b = ldOff<double>(P.b, o >> 1);
  } else {
    const dpair r2 = ldOff<dpair>(P.row2, o);
    g.aE = r2[0];
    g.b = r2[1];
  }
  g.xC = ldOff<double>(P.xOld, o);
  g.xE = ldOff<double>(P.xOld, oE);
  g.xN = ldOff<double>(P.xOld, oE + 16u);
}

template <int D, bool SepB>
__device__ void gsLexWave(int nI, int nJ, int R, const int* partnerG, const GsOne& S, int sweeps, int* stall) {
  HIP_DYNAMIC_SHARED(double, lds)
  const int B = static_cast<int>(blockDim.x);
  const int W = B >> 6;
  double* row0 = lds;                                       // [nI]: row j = 0, updated in place
  double* hand = row0 + nI;                                 // [W][kSweepHand]: lane 63's values
  double* dummy = hand + W * kSweepHand;                    // [B]: writes of the lanes that make none
  int* partner = reinterpret_cast<int*>(dummy + B);         // [nI]
  int* posted = partner + nI;                               // [W]: last step each wave finished
  int* dummyInt = posted + W;                               // [B]
  const int lane = static_cast<int>(threadIdx.x);
  // the wave index as a scalar, so that branches on it are uniform
  const int w = __builtin_amdgcn_readfirstlane(lane >> 6), l = lane & 63;
  const int laneG = min(lane, R - 1);
  const bool active = lane < nJ;
  const int steps = skewSteps(nI, nJ);
  // the steps on which some row of this wave can have a cell (0 <= k - j < nI)
  const int kBegin = 64 * w;
  const int kEnd = min(steps, 64 * w + 63 + nI);
  // the loop starts D steps early: those steps only fill the read-ahead ring
  // (every lane is off the wavefront there), so no value is loaded before the
  // loop and the compiler's waits inside it stay partial
  const int kFirst = kBegin - D;
  const int count = max(kEnd - kFirst, 0);
  const int rings = (count + D - 1) / D;
  const CFD_GLOBAL int* partnerSrc = gmem(partnerG);
  for (int t = lane; t < nI; t += B) partner[t] = partnerSrc[t];
  bool ok = true;
  const unsigned R16 = static_cast<unsigned>(R) * 16u;
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    const CFD_GLOBAL double* xOld = (sweep & 1) ? S.x[1] : S.x[0];
    CFD_GLOBAL double* xNew = (sweep & 1) ? S.x[0] : S.x[1];
    for (int t = lane; t < nI; t += B) row0[t] = xOld[2 * skewIndex(t, 0, R) + S.half];
    const GsBytes P{reinterpret_cast<const CFD_GLOBAL char*>(S.row[0]), reinterpret_cast<const CFD_GLOBAL char*>(S.row[1]),
                    reinterpret_cast<const CFD_GLOBAL char*>(S.row[2]), reinterpret_cast<const CFD_GLOBAL char*>(S.b),
                    reinterpret_cast<const CFD_GLOBAL char*>(xOld + S.half), reinterpret_cast<CFD_GLOBAL char*>(xNew + S.half)};
    // byte offsets: the slot fetched (step k + D) and the slot stored (step
    // k), one row on per step; a lane past R parks its stores
    unsigned oF = static_cast<unsigned>((kFirst + D + kSkewFront) * R + laneG) * 16u;
    unsigned oS = static_cast<unsigned>(lane < R ? (kFirst + kSkewFront) * R + lane : skewParkIndex(nI, nJ, R, lane)) * 16u;
    const unsigned incS = lane < R ? R16 : 0u;
    if (l == 0) ldsPost(&posted[w], kBegin - 1);
    __syncthreads();
    double xw = 0.0, up = 0.0;
    GsCell1 ring[D];
#pragma unroll
    for (int q = 0; q < D; ++q) ring[q] = GsCell1{};
    int seenBelow = kBegin - 65;  // the wave underneath has finished this step (its start - 1)
    // lane 0's value below for the coming step and, in wave 0, the wake
    // partners of this lane's next two cells: read one (two) steps ahead so
    // the step does not wait for LDS
    int prCur = -1, prNext = -1;
    double lowNext = 0.0;
    if (w == 0) {
      prCur = partner[min(max(kFirst - lane, 0), nI - 1)];
      prNext = partner[min(max(kFirst + 1 - lane, 0), nI - 1)];
      lowNext = row0[max(prCur, 0)];
    }
    for (int r = 0; r < rings; ++r) {
#pragma unroll
      for (int q = 0; q < D; ++q) {
        // no per-step condition: steps past kEnd touch no cell (they store into
        // slots of no cell), so the body stays branch-free for the read-ahead
        const int k = kFirst + r * D + q;
        const int i = k - lane;
        const bool on = active && i >= 0 && i < nI;
        const GsCell1& g = ring[q];
        // the reads issued on the last step are complete, and so are its
        // writes: post its hand-off value (counters run one step behind)
        ldsOrder();
        ldsPost((k - 1 >= kBegin && l == 63) ? &posted[w] : &dummyInt[lane], k - 1);
        // the value below: lane l - 1's result of the last step, and on lane
        // 0 the value read ahead for this step (both put in place by the shift)
        const double below = up;
        // read ahead for step k + 1
        if (w == 0) {
          prCur = prNext;
          lowNext = row0[max(prCur, 0)];  // forwarded below if this step writes that cell
          prNext = partner[min(max(i + 2, 0), nI - 1)];
        } else {
          const int i1 = k + 1 - 64 * w;  // lane 0's cell on the next step
          if (ok && i1 >= 0 && i1 < nI && seenBelow < k) {
            ok = waitStep(&posted[w - 1], k, stall);
            seenBelow = __builtin_amdgcn_readfirstlane(ldsPoll(&posted[w - 1]));
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
          }
          lowNext = hand[(w - 1) * kSweepHand + (k & (kSweepHand - 1))];
        }
        // cellFaces order: below, right, above, left (Mesh.cpp:206-213). The
        // reference skips a missing neighbour; its coefficient is zero in the
        // assembled row, so here its term adds (+-0) to a sum that is never
        // -0 (it starts at +0, and x + y is -0 only if both are), which
        // leaves the sum as it is. (The neighbour's value is a finite one:
        // a cell of the field, the zeroed padding, or row 0.)
        double off = 0.0;
        off = off + g.aS * below;
        off = off + g.aE * g.xE;
        off = off + g.aN * g.xN;
        off = off + g.aW * xw;
        // a row with a zero diagonal keeps its value (LinearSystem.cpp:86-88);
        // the quotient is formed on every lane so the choice is a select
        const double quotient = (g.b - off) / g.aP;
        computedHere(quotient);
        const double xn = (fabs(g.aP) > 0.0) ? quotient : g.xC;
        // every lane stores on every step: off the wavefront into a slot of
        // no cell (the old value, zero, goes back) or a parking slot
        stOff<double>(P.xNew, oS, xn);
        oS += incS;
        opaqueOffset(oS);
        *((on && lane == 0) ? &row0[max(i, 0)] : &dummy[lane]) = xn;
        // the value read ahead is this cell's old one when the next cell's
        // partner is this cell (the write above comes after that read)
        if (w == 0) lowNext = (on && prCur == i) ? xn : lowNext;
        // lane 63's value for the wave above; every 16 steps make sure the
        // ring slots of the next 16 are free (read by the wave above)
        const bool feeds = (w + 1 < W) && k >= kBegin;
        if (feeds && (k & 15) == 0 && ok) ok = waitStep(&posted[w + 1], k + 16 - kSweepHand, stall);
        *((feeds && l == 63) ? &hand[w * kSweepHand + (k & (kSweepHand - 1))] : &dummy[lane]) = xn;
        up = waveShiftUp(xn, lowNext);
        xw = xn;
        // the refill may not be hoisted above the shift: the slot's old
        // values and this step's results are dead by here, so the loads land
        // in the slot's own registers (no copies of loads in flight, which
        // would force a full wait at the loop latch)
        holdUntilHere(up);
        gsFetch1<SepB>(P, oF, oF + R16, ring[q]);
        oF += R16;
        opaqueOffset(oF);
      }
    }
    ldsOrder();  // the last step's hand-off value is written
    // nobody waits on this wave for the rest of the sweep
    if (l == 63) ldsPost(&posted[w], steps + kSweepHand);
    __syncthreads();  // this sweep's writes are the next sweep's old values
  }
}

// Momentum: u and v share one matrix, one workgroup. Turbulence: blockIdx.x
// 0 sweeps k, 1 sweeps omega (independent systems, run side by side). The
// launch gives the block I.gsR lanes (the layout's stride).
template <int D>
__device__ __forceinline__ void gsLexMomentum(const Inst* insts, int sweeps) {
  const Inst& I = insts[blockIdx.y];
  if (I.flags[fStopped] || sweeps <= 0) return;
  const GsSkew<2> S{{gpair(static_cast<const double2*>(I.skMomRow[0])), gpair(static_cast<const double2*>(I.skMomRow[1])),
                     gpair(static_cast<const double2*>(I.skMomRow[2]))},
                    gmem(static_cast<const double*>(I.skMomBy)),
                    {gpair(I.skMomX[0]), gpair(I.skMomX[1])}};
  // u (blockIdx.x 0) and v (1) in two workgroups, side by side
  const int half = (blockIdx.x == 0) ? 0 : 1;
  const GsOne C{{S.row[0], S.row[1], S.row[2]},
                S.b1,
                {reinterpret_cast<CFD_GLOBAL double*>(S.x[0]), reinterpret_cast<CFD_GLOBAL double*>(S.x[1])},
                half};
  if (half == 0) gsLexWave<D, false>(I.nI, I.nJ, I.gsR, I.partner, C, sweeps, I.flags + fSweepStall);
  else gsLexWave<D, true>(I.nI, I.nJ, I.gsR, I.partner, C, sweeps, I.flags + fSweepStall);
}

template <int D>
__device__ __forceinline__ void gsLexTurbulence(const Inst* insts, int sweeps) {
  const Inst& I = insts[blockIdx.y];
  if (I.flags[fStopped] || sweeps <= 0) return;
  const int sys = (blockIdx.x == 0) ? 0 : 1;  // 0: k, 1: omega
  double2* const* row = I.skTurbRow[sys];
  const GsSkew<1> S{{gpair(static_cast<const double2*>(row[0])), gpair(static_cast<const double2*>(row[1])),
                     gpair(static_cast<const double2*>(row[2]))},
                    nullptr,
                    {gpair(I.skTurbX[sys][0]), gpair(I.skTurbX[sys][1])}};
  const GsOne C{{S.row[0], S.row[1], S.row[2]},
                nullptr,
                {reinterpret_cast<CFD_GLOBAL double*>(S.x[0]), reinterpret_cast<CFD_GLOBAL double*>(S.x[1])},
                0};
  gsLexWave<D, false>(I.nI, I.nJ, I.gsR, I.partner, C, sweeps, I.flags + fSweepStall);
}

// The sweep kernels (gsLexWave): one workgroup per component (u and v side by
// side for momentum, k and omega for turbulence). Up to 256 rows: one wave
// per 64 rows, each lane holding a ring of 8 cells fetched ahead; up to 1024
// rows: a ring of 3.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_gs_wave_mom(const Inst* insts, int sweeps) {
  gsLexMomentum<8>(insts, sweeps);
}
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_gs_wave_turb(const Inst* insts, int sweeps) {
  gsLexTurbulence<8>(insts, sweeps);
}
__global__ void __launch_bounds__(1024) k_gs_wave_mom_wide(const Inst* insts, int sweeps) {
  gsLexMomentum<3>(insts, sweeps);
}
__global__ void __launch_bounds__(1024) k_gs_wave_turb_wide(const Inst* insts, int sweeps) {
  gsLexTurbulence<3>(insts, sweeps);
}

// After `sweeps` sweeps the result is in the skewed copy x[sweeps & 1]:
// copy it to the row-major field. System 0: u and v; system 1: k and omega
// with their floors (KOmegaSST.cpp:468-471).
__global__ void __launch_bounds__(kBlock) k_gs_unskew(const Inst* insts, int system, int sweeps) {
  CFD_PROLOGUE
  if (!live || sweeps <= 0) return;
  const int sk = skewIndex(i, j, I.gsR);
  const int src = sweeps & 1;
  if (system == 0) {
    const dpair v = gpair(static_cast<const double2*>(I.skMomX[src]))[sk];
    I.u[c] = v[0];
    I.v[c] = v[1];
  } else {
    const dpair k = gpair(static_cast<const double2*>(I.skTurbX[0][src]))[sk];
    const dpair w = gpair(static_cast<const double2*>(I.skTurbX[1][src]))[sk];
    I.k[c] = mx(k[0], 0.0);
    I.w[c] = mx(w[0], 1.0e-10);
  }
}

// ---------------------------------------------------------------------------
// Two-colour Gauss-Seidel: colour (i + j) & 1, one launch per colour. Same
// fixed point as the lexicographic sweep, full parallelism. The C-grid has an
// even nI, so a wake cell and its partner always differ in colour.
// ---------------------------------------------------------------------------

template <int NC>
__device__ __forceinline__ void gsPoint(const Inst& I, const GsSys& S, int c, int i, int j) {
  const int nI = I.nI, nJ = I.nJ;
  const double d = S.aP[c];
  if (!(fabs(d) > 0.0)) return;
  double off[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) off[q] = 0.0;
  if (j > 0) {
    const double a = S.aS[c];
#pragma unroll
    for (int q = 0; q < NC; ++q) off[q] += a * S.x[q][c - nI];
  } else {
    const int pr = I.partner[i];
    if (pr >= 0) {
      const double a = S.aS[c];
#pragma unroll
      for (int q = 0; q < NC; ++q) off[q] += a * S.x[q][pr];
    }
  }
  if (i < nI - 1) {
    const double a = S.aE[c];
#pragma unroll
    for (int q = 0; q < NC; ++q) off[q] += a * S.x[q][c + 1];
  }
  if (j < nJ - 1) {
    const double a = S.aN[c];
#pragma unroll
    for (int q = 0; q < NC; ++q) off[q] += a * S.x[q][c + nI];
  }
  if (i > 0) {
    const double a = S.aW[c];
#pragma unroll
    for (int q = 0; q < NC; ++q) off[q] += a * S.x[q][c - 1];
  }
#pragma unroll
  for (int q = 0; q < NC; ++q) S.x[q][c] = (S.b[q][c] - off[q]) / d;
}

__global__ void __launch_bounds__(kBlock) k_gs_color(const Inst* insts, int sys, int color, int applyFloor) {
  CFD_PROLOGUE
  if (!live) return;
  if (sys == 0) {
    GsSys S{I.aP, I.aS, I.aN, I.aW, I.aE, {I.bx, I.by}, {I.u, I.v}, 0.0, 0};
    if (((i + j) & 1) == color) gsPoint<2>(I, S, c, i, j);
  } else if (sys == 1) {
    GsSys S{I.kP, I.kS, I.kN, I.kW, I.kE, {I.kB, nullptr}, {I.k, nullptr}, 0.0, 1};
    if (((i + j) & 1) == color) gsPoint<1>(I, S, c, i, j);
    if (applyFloor) I.k[c] = mx(I.k[c], 0.0);
  } else {
    GsSys S{I.wP, I.wS, I.wN, I.wW, I.wE, {I.wB, nullptr}, {I.w, nullptr}, 1.0e-10, 1};
    if (((i + j) & 1) == color) gsPoint<1>(I, S, c, i, j);
    if (applyFloor) I.w[c] = mx(I.w[c], 1.0e-10);
  }
}

// ---------------------------------------------------------------------------
// Rhie-Chow face fluxes (SimpleSolver.cpp:445-496). Cell (i,j) writes its
// below and left faces; the last row / column also write the top / right.
// ---------------------------------------------------------------------------

__device__ __forceinline__ double rcInterior(const Inst& I, const FaceGeo& f, int o, int n) {
  const double w = f.w;
  const double ux = I.u[o] * w + I.u[n] * (1.0 - w);
  const double uy = I.v[o] * w + I.v[n] * (1.0 - w);
  const double dens = w * I.rho[o] + (1.0 - w) * I.rho[n];
  const double mob = w * (I.V[o] / I.aP[o]) + (1.0 - w) * (I.V[n] / I.aP[n]);
  const double2 go = ld2(I.gP, o), gn = ld2(I.gP, n);
  const double sgx = go.x * w + gn.x * (1.0 - w);
  const double sgy = go.y * w + gn.y * (1.0 - w);
  const double compact = (I.p[n] - I.p[o]) * f.g;
  const double smooth = sgx * f.Sx + sgy * f.Sy;
  return dens * (ux * f.Sx + uy * f.Sy) - dens * mob * (compact - smooth);
}

__device__ __forceinline__ double rcBoundary(const Inst& I, const FaceGeo& f, const FaceBC& bc, double Fold, int o) {
  const double dens = I.rho[o];
  if (imposes(bc, Fold)) return dens * (bc.u * f.Sx + bc.v * f.Sy);
  const double mob = I.V[o] / I.aP[o];
  const double compact = (bc.p - I.p[o]) * f.g;
  const double2 g = ld2(I.gP, o);
  const double smooth = g.x * f.Sx + g.y * f.Sy;
  return dens * (I.u[o] * f.Sx + I.v[o] * f.Sy) - dens * mob * (compact - smooth);
}

__global__ void __launch_bounds__(kBlock) k_rhiechow(const Inst* insts) {
  CFD_PROLOGUE
  if (!live) return;
  // below
  {
    const int fi = j * nI + i;
    const FaceGeo f = geo(I.jg, fi);
    if (j > 0) {
      I.Fj[fi] = rcInterior(I, f, c - nI, c);
    } else {
      const int pr = I.partner[i];
      if (pr >= 0) I.Fj[fi] = rcInterior(I, f, c, pr);
      else I.Fj[fi] = rcBoundary(I, f, I.bcB[i], I.Fj[fi], c);
    }
  }
  if (j == nJ - 1) {
    const int fi = (j + 1) * nI + i;
    I.Fj[fi] = rcBoundary(I, geo(I.jg, fi), I.bcT[i], I.Fj[fi], c);
  }
  // left
  {
    const int fi = j * (nI + 1) + i;
    const FaceGeo f = geo(I.ig, fi);
    if (i > 0) I.Fi[fi] = rcInterior(I, f, c - 1, c);
    else I.Fi[fi] = rcBoundary(I, f, I.bcL[j], I.Fi[fi], c);
  }
  if (i == nI - 1) {
    const int fi = j * (nI + 1) + i + 1;
    I.Fi[fi] = rcBoundary(I, geo(I.ig, fi), I.bcR[j], I.Fi[fi], c);
  }
}

// ---------------------------------------------------------------------------
// Pressure-correction assembly (SimpleSolver.cpp:498-580) and the continuity
// residual of pass 0 (:720-727).
// ---------------------------------------------------------------------------

__device__ __forceinline__ double presCoef(const Inst& I, const FaceGeo& f, int o, int n) {
  const double w = f.w;
  const double dens = w * I.rho[o] + (1.0 - w) * I.rho[n];
  const double mob = w * (I.V[o] / I.aP[o]) + (1.0 - w) * (I.V[n] / I.aP[n]);
  return dens * mob * f.g;
}

__device__ __forceinline__ void presBoundary(const Inst& I, const FaceGeo& f, const FaceBC& bc, double F, int o,
                                             double& diag) {
  if (imposes(bc, F)) return;
  const double dens = I.rho[o];
  const double mob = I.V[o] / I.aP[o];
  diag += dens * mob * f.g;
}

__global__ void __launch_bounds__(kBlock) k_pressure(const Inst* insts, int pass) {
  CFD_PROLOGUE
  __shared__ double sm[16];
  double part[2] = {0.0, 0.0};
  if (live) {
    double net = 0.0, diag = 0.0, cN = 0.0, cE = 0.0, cWk = 0.0;
    // below
    {
      const int fi = j * nI + i;
      const FaceGeo f = geo(I.jg, fi);
      const double F = I.Fj[fi];
      if (j > 0) {
        net -= F;
        diag += presCoef(I, f, c - nI, c);
      } else {
        net += F;
        const int pr = I.partner[i];
        if (pr >= 0) {
          const double cf = presCoef(I, f, c, pr);
          diag += cf;
          cWk = -cf;
        } else {
          presBoundary(I, f, I.bcB[i], F, c, diag);
        }
      }
    }
    // above
    {
      const int fi = (j + 1) * nI + i;
      const FaceGeo f = geo(I.jg, fi);
      const double F = I.Fj[fi];
      net += F;
      if (j < nJ - 1) {
        const double cf = presCoef(I, f, c, c + nI);
        diag += cf;
        cN = -cf;
      } else {
        presBoundary(I, f, I.bcT[i], F, c, diag);
      }
    }
    // left
    {
      const int fi = j * (nI + 1) + i;
      const FaceGeo f = geo(I.ig, fi);
      const double F = I.Fi[fi];
      if (i > 0) {
        net -= F;
        diag += presCoef(I, f, c - 1, c);
      } else {
        net += F;
        presBoundary(I, f, I.bcL[j], F, c, diag);
      }
    }
    // right
    {
      const int fi = j * (nI + 1) + i + 1;
      const FaceGeo f = geo(I.ig, fi);
      const double F = I.Fi[fi];
      net += F;
      if (i < nI - 1) {
        const double cf = presCoef(I, f, c, c + 1);
        diag += cf;
        cE = -cf;
      } else {
        presBoundary(I, f, I.bcR[j], F, c, diag);
      }
    }
    I.pP[c] = diag;
    I.pN[c] = cN;
    I.pE[c] = cE;
    I.pWk[c] = cWk;
    I.pb[c] = -net;
    if (I.precision == 1) {
      I.fpP[c] = static_cast<float>(diag);
      I.fpN[c] = static_cast<float>(cN);
      I.fpE[c] = static_cast<float>(cE);
      I.fpWk[c] = static_cast<float>(cWk);
    }
    part[0] = fabs(net);
    part[1] = -net;
  }
  blockSum<2>(part, sm);
  if (threadIdx.x == 0) {
    if (pass == 0) partialSlot(I, sNetFluxL1)[blockIdx.x] = part[0];
    partialSlot(I, sSourceSum)[blockIdx.x] = part[1];
  }
}

// ---------------------------------------------------------------------------
// Jacobi-preconditioned CG with the reference's L1 stopping test
// (LinearSystem.cpp:118-210), two launches per iteration:
//   A: p_k = z + beta p_{k-1} (halo recomputed locally), Ap, p.Ap
//   B: x += alpha p, r -= alpha Ap, |r|_1, z = r / diag, r.z
// The scalars live in device memory: every block of a launch sums the
// previous launch's per-block partials, so the host never waits inside the
// solve; once the test passes (or the iteration breaks down) every later
// launch is a no-op.
//
// The bodies are written once for the storage type T of the matrix and of
// r, z, p and Ap: double (the plain path, the reference's arithmetic) or
// float (the mixed path, half the bytes per iteration). Either way every
// value is widened on load, the arithmetic and the reductions are FP64, x is
// FP64, and a value is rounded to T once, where it is stored; the rounded
// value is the one used afterwards (p in the stencil, r in |r|_1 and r.z), so
// the FP32 solve is a consistent CG on the FP32-rounded system. With T =
// double every conversion is the identity.
// ---------------------------------------------------------------------------

template <typename T>
struct CgStore {
  const T *pP, *pN, *pE, *pWk;
  T *r, *z, *pd0, *pd1, *Ap;
};

template <typename T>
__device__ __forceinline__ CgStore<T> cgStore(const Inst& I);
template <>
__device__ __forceinline__ CgStore<double> cgStore<double>(const Inst& I) {
  return CgStore<double>{I.pP, I.pN, I.pE, I.pWk, I.r, I.z, I.pd0, I.pd1, I.Ap};
}
template <>
__device__ __forceinline__ CgStore<float> cgStore<float>(const Inst& I) {
  return CgStore<float>{I.fpP, I.fpN, I.fpE, I.fpWk, I.fr, I.fz, I.fpd0, I.fpd1, I.fAp};
}

// Loads in flight per lane when a CG launch sums the previous launch's
// partials: 64 x 8 = 512 partials (a 262,144-cell solve's CG blocks) in one
// round trip, at half the registers of the general sum (launch A holds two
// slots' loads beside its stencil operands).
constexpr int kCgPartialUnroll = 8;

// a value as it is after a store and a load of type T
template <typename T>
__device__ __forceinline__ double asStored(double v) {
  return static_cast<double>(static_cast<T>(v));
}

// The CG launches give each thread kCgCells cells (c = tile * kBlock *
// kCgCells + q * kBlock + threadIdx.x, q in order), so a launch has
// 1 / kCgCells of the blocks of the cell-parallel kernels and every block of
// the next launch sums that many fewer partials. The per-thread sums run over
// q in order before the block sum: the result is fixed run to run.
template <typename T>
__device__ __forceinline__ void cgInit(const Inst* insts, int pass, double* sm) {
  const Inst& I = insts[blockIdx.y];
  if (static_cast<int>(blockIdx.x) >= I.cgBlocks) return;
  const CFD_GLOBAL int* flagsG = gmem(static_cast<const int*>(I.flags));
  if (flagsG[fStopped]) return;
  const int tile = xcdTile(static_cast<int>(blockIdx.x), I.cgBlocks);
  const int base = tile * kBlock * kCgCells + static_cast<int>(threadIdx.x);
  const CgStore<T> S = cgStore<T>(I);
  // pass 0 solves for the assembled source (its mean removed when nothing
  // anchors the pressure level); a refinement pass for the FP64 residual
  const bool demean = pass == 0 && !I.hasPRef;
  const double* src = (pass == 0) ? I.pb : I.rb;
  double mean = 0.0;
  if (demean) mean = blockSumPartials(partialSlot(I, sSourceSum), I.nBlocks, sm) / static_cast<double>(I.N);
  double part[2] = {0.0, 0.0};
  double bq[kCgCells], dq[kCgCells];
#pragma unroll
  for (int q = 0; q < kCgCells; ++q) {
    const int cc = min(base + q * kBlock, I.N - 1);
    bq[q] = gmem(src)[cc];
    dq[q] = static_cast<double>(gmem(S.pP)[cc]);
  }
#pragma unroll
  for (int q = 0; q < kCgCells; ++q) {
    const int c = base + q * kBlock;
    if (c < I.N) {
      double b = bq[q];
      if (demean) {
        b -= mean;
        gmem(I.pb)[c] = b;
      }
      const double r = asStored<T>(b - 0.0);  // x0 = 0, so A x0 = 0
      const double d = dq[q];
      const double z = asStored<T>((fabs(d) > 0.0) ? r / d : r);
      gmem(S.r)[c] = static_cast<T>(r);
      gmem(S.z)[c] = static_cast<T>(z);
      gmem(S.pd0)[c] = static_cast<T>(z);
      gmem(I.x)[c] = 0.0;
      part[0] += fabs(r);
      part[1] += r * z;
    }
  }
  blockSum<2>(part, sm);
  if (threadIdx.x == 0) {
    partialSlot(I, sCgNorm)[blockIdx.x] = part[0];
    partialSlot(I, sCgRz)[blockIdx.x] = part[1];
    if (blockIdx.x == 0) {
      gmem(I.flags)[fCgDone] = 0;
      gmem(I.flags)[fCgIters] = 0;
    }
  }
}

// Iteration `it`, launch A. Finishes the test of iteration it-1 from the
// partials of launch B (|r|_1 <= tol |r0|_1, then the r.z guard), forms beta,
// then p = z + beta p_old (neighbours' p recomputed locally, so no second
// launch is needed before the stencil), Ap and p.Ap. The cells' operands do
// not depend on beta, so they are loaded before the partials are summed and
// the two memory round trips overlap.
template <typename T>
__device__ __forceinline__ void cgA(const Inst* insts, int it, double* sm) {
  constexpr int CELLS = kCgCells;
  constexpr int THREADS = kBlock;
  const Inst& I = insts[blockIdx.y];
  if (static_cast<int>(blockIdx.x) >= I.cgBlocks) return;
  if (it >= I.cgMax) return;  // a batch member with a smaller cap than the batch's
  const int tile = xcdTile(static_cast<int>(blockIdx.x), I.cgBlocks);
  const int base = tile * kBlock * kCgCells + static_cast<int>(threadIdx.x);
  const int nI = I.nI;
  const int nJ = I.nJ;
  const CgStore<T> S = cgStore<T>(I);
  // Every operand is read through the global address space, without a branch
  // around its load, and issued before the stop flags are tested: the flags,
  // the CG scalars, the cells' wake partners and the stencil's operands are
  // one memory round trip (a FLAT or branch-guarded load would make a wait
  // drain every load in flight). Only the lanes of row 0 then wait for their
  // partner's values, and the partials are a second round trip.
  const CFD_GLOBAL int* flagsG = gmem(static_cast<const int*>(I.flags));
  const CFD_GLOBAL double* red = gmem(static_cast<const double*>(I.red));
  const CFD_GLOBAL int* partnerG = gmem(static_cast<const int*>(I.partner));
  int cq[CELLS], iq[CELLS], jq[CELLS], prq[CELLS];
  bool live[CELLS];
#pragma unroll
  for (int q = 0; q < CELLS; ++q) {
    const int c = base + q * THREADS;
    live[q] = c < I.N;
    cq[q] = live[q] ? c : 0;
    iq[q] = cq[q] % nI;
    jq[q] = cq[q] / nI;
    prq[q] = partnerG[iq[q]];
  }
  const int stopped = flagsG[fStopped];
  const int cgDone = flagsG[fCgDone];
  const double norm0 = red[sCgNorm0];
  const double rzOld = red[((it - 1) & 1) ? sCgRz1 : sCgRz0];
  // operands: z and p_old at each cell and its four neighbours, the matrix row
  const CFD_GLOBAL T* z = gmem(static_cast<const T*>(S.z));
  const CFD_GLOBAL T* pOld = gmem(static_cast<const T*>((it & 1) ? S.pd0 : S.pd1));
  const CFD_GLOBAL T* pFirst = gmem(static_cast<const T*>(S.pd0));
  const CFD_GLOBAL T* src = (it == 0) ? pFirst : pOld;  // iteration 0 uses p0 = z (in pd0)
  const CFD_GLOBAL T* pP = gmem(S.pP);
  const CFD_GLOBAL T* pN = gmem(S.pN);
  const CFD_GLOBAL T* pE = gmem(S.pE);
  const CFD_GLOBAL T* pWk = gmem(S.pWk);
  double zc[CELLS], zS[CELLS], zN[CELLS], zW[CELLS], zE[CELLS];
  double oc[CELLS], oS[CELLS], oN[CELLS], oW[CELLS], oE[CELLS];
  double aP[CELLS], aS[CELLS], aN[CELLS], aW[CELLS], aE[CELLS];
  bool hasS[CELLS], hasN[CELLS], hasW[CELLS], hasE[CELLS];
#pragma unroll
  for (int q = 0; q < CELLS; ++q) {
    const int cc = cq[q], i = iq[q], j = jq[q];
    hasN[q] = j < nJ - 1;
    hasW[q] = i > 0;
    hasE[q] = i < nI - 1;
    // the cell below; row 0 takes its wake partner's values further down,
    // so only the lanes of row 0 wait for the partner index
    const int cS = (j > 0) ? cc - nI : cc;
    const int cN = hasN[q] ? cc + nI : cc;
    const int cW = hasW[q] ? cc - 1 : cc;
    const int cE = hasE[q] ? cc + 1 : cc;
    aP[q] = pP[cc];
    aS[q] = pN[(j > 0) ? cc - nI : cc];  // row 0: the wake coefficient, below
    aN[q] = pN[cc];
    aW[q] = pE[cW];  // the left face's coefficient is stored with the cell to its left
    aE[q] = pE[cc];
    oc[q] = src[cc]; oS[q] = src[cS]; oN[q] = src[cN]; oW[q] = src[cW]; oE[q] = src[cE];
    zc[q] = zS[q] = zN[q] = zW[q] = zE[q] = 0.0;
    if (it > 0) { zc[q] = z[cc]; zS[q] = z[cS]; zN[q] = z[cN]; zW[q] = z[cW]; zE[q] = z[cE]; }
    hasS[q] = j > 0;
  }
  if (stopped || cgDone) return;  // the whole block: the flags are uniform
#pragma unroll
  for (int q = 0; q < CELLS; ++q) {
    if (jq[q] == 0 && prq[q] >= 0) {
      hasS[q] = true;
      aS[q] = pWk[cq[q]];
      oS[q] = src[prq[q]];
      if (it > 0) zS[q] = z[prq[q]];
    }
  }
  double sums[2];
  {
    const double* const slots[2] = {partialSlot(I, sCgNorm), partialSlot(I, sCgRz)};
    blockSumPartials<2, kCgPartialUnroll>(slots, I.cgBlocks, sm, sums);
  }
  const double nrm = sums[0];
  const double rzNew = sums[1];
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
  double part[1] = {0.0};
  // p = z + beta p_old (LinearSystem.cpp:199-201), as stored; iteration 0 has p = z
  auto pv = [&](double zq, double oq) { return (it == 0) ? oq : asStored<T>(zq + beta * oq); };
#pragma unroll
  for (int q = 0; q < CELLS; ++q) {
    if (!live[q]) continue;
    const int c = base + q * THREADS;
    const double pc = pv(zc[q], oc[q]);
    // LinearSystem::multiply order: diagonal, then the faces of the cell
    // (below/wake, above, left, right as the structured sweep visits them)
    double y = aP[q] * pc;
    if (hasS[q]) y += aS[q] * pv(zS[q], oS[q]);
    if (hasN[q]) y += aN[q] * pv(zN[q], oN[q]);
    if (hasW[q]) y += aW[q] * pv(zW[q], oW[q]);
    if (hasE[q]) y += aE[q] * pv(zE[q], oE[q]);
    y = asStored<T>(y);
    if (it > 0) gmem((it & 1) ? S.pd1 : S.pd0)[c] = static_cast<T>(pc);
    gmem(S.Ap)[c] = static_cast<T>(y);
    part[0] += pc * y;
  }
  blockSum<1>(part, sm);
  if (threadIdx.x == 0) partialSlot(I, sCgPAp)[blockIdx.x] = part[0];
}

// Iteration `it`, launch B: alpha = r.z / p.Ap, x += alpha p, r -= alpha Ap,
// z = r / a_P, and the partials of |r|_1 and r.z (LinearSystem.cpp:160-198).
template <typename T>
__device__ __forceinline__ void cgB(const Inst* insts, int it, double* sm) {
  const Inst& I = insts[blockIdx.y];
  if (static_cast<int>(blockIdx.x) >= I.cgBlocks) return;
  if (it >= I.cgMax) return;
  const int tile = xcdTile(static_cast<int>(blockIdx.x), I.cgBlocks);
  const int base = tile * kBlock * kCgCells + static_cast<int>(threadIdx.x);
  const CgStore<T> S = cgStore<T>(I);
  // the flags, r.z of this iteration, the cells' operands and the partials in
  // one memory round trip; a stopped or finished solve returns after the sum
  // (the whole block together)
  const CFD_GLOBAL int* flagsG = gmem(static_cast<const int*>(I.flags));
  const int stopped = flagsG[fStopped];
  const int cgDone = flagsG[fCgDone];
  const double rz = gmem(static_cast<const double*>(I.red))[(it & 1) ? sCgRz1 : sCgRz0];
  double xc[kCgCells], pc[kCgCells], rc[kCgCells], apc[kCgCells], d[kCgCells];
#pragma unroll
  for (int q = 0; q < kCgCells; ++q) {
    const int c = base + q * kBlock;
    const int cc = (c < I.N) ? c : 0;
    xc[q] = gmem(static_cast<const double*>(I.x))[cc];
    pc[q] = gmem(static_cast<const T*>((it & 1) ? S.pd1 : S.pd0))[cc];
    rc[q] = gmem(static_cast<const T*>(S.r))[cc];
    apc[q] = gmem(static_cast<const T*>(S.Ap))[cc];
    d[q] = gmem(S.pP)[cc];
  }
  const double pap = blockSumPartials<kCgPartialUnroll>(partialSlot(I, sCgPAp), I.cgBlocks, sm);
  if (stopped || cgDone) return;
  if (!(fabs(pap) > 0.0)) {  // the direction has no component the matrix acts on
    if (threadIdx.x == 0 && blockIdx.x == 0) gmem(I.flags)[fCgDone] = 1;
    return;
  }
  const double alpha = rz / pap;
  double part[2] = {0.0, 0.0};
#pragma unroll
  for (int q = 0; q < kCgCells; ++q) {
    const int c = base + q * kBlock;
    if (c >= I.N) continue;
    const double x = xc[q] + alpha * pc[q];
    const double r = asStored<T>(rc[q] - alpha * apc[q]);
    gmem(I.x)[c] = x;
    gmem(S.r)[c] = static_cast<T>(r);
    const double zq = asStored<T>((fabs(d[q]) > 0.0) ? r / d[q] : r);
    gmem(S.z)[c] = static_cast<T>(zq);
    part[0] += fabs(r);
    part[1] += r * zq;
  }
  blockSum<2>(part, sm);
  if (threadIdx.x == 0) {
    partialSlot(I, sCgNorm)[blockIdx.x] = part[0];
    partialSlot(I, sCgRz)[blockIdx.x] = part[1];
    if (blockIdx.x == 0) {
      gmem(I.flags)[fCgIters] = it + 1;
      gmem(I.flags)[fCgTotal] += 1;
    }
  }
}

// FP64 storage: the plain path
__global__ void __launch_bounds__(kBlock) k_cg_init(const Inst* insts, int pass) {
  __shared__ double sm[16];
  cgInit<double>(insts, pass, sm);
}
// At most 128 VGPRs (4 waves per SIMD; 150 unconstrained): measured faster in
// batches, the same for one solve.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) k_cg_a(const Inst* insts, int it) {
  __shared__ double sm[16];
  cgA<double>(insts, it, sm);
}
__global__ void __launch_bounds__(kBlock) k_cg_b(const Inst* insts, int it) {
  __shared__ double sm[16];
  cgB<double>(insts, it, sm);
}

// FP32 storage: the mixed path
__global__ void __launch_bounds__(kBlock) k_cg_init_f32(const Inst* insts, int pass) {
  __shared__ double sm[16];
  cgInit<float>(insts, pass, sm);
}
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) k_cg_a_f32(const Inst* insts, int it) {
  __shared__ double sm[16];
  cgA<float>(insts, it, sm);
}
__global__ void __launch_bounds__(kBlock) k_cg_b_f32(const Inst* insts, int it) {
  __shared__ double sm[16];
  cgB<float>(insts, it, sm);
}

// Partial sums of the correction for its mean when no boundary anchors the
// pressure level (SimpleSolver.cpp:735-745); the correction kernels finish it.
__global__ void __launch_bounds__(kBlock) k_corr_mean(const Inst* insts) {
  CFD_PROLOGUE
  __shared__ double sm[16];
  double part[1] = {live ? I.x[c] : 0.0};
  blockSum<1>(part, sm);
  if (threadIdx.x == 0) partialSlot(I, sCorrMean)[blockIdx.x] = part[0];
}

// ---------------------------------------------------------------------------
// Correction (SimpleSolver.cpp:582-660): p' face values -> Green-Gauss ->
// p += alpha_p p', u -= grad p' V / a_P. Face fluxes in a second launch, so
// the imposed-velocity test of the first still sees the uncorrected fluxes.
// ---------------------------------------------------------------------------

// (eight partials in flight per lane rather than sixteen: the kernel then
// fits in 60 registers and eight waves per SIMD; the order of the sum is the same)
__device__ __forceinline__ double corrMean(const Inst& I, double* sm) {
  return I.hasPRef ? 0.0 : blockSumPartials<8>(partialSlot(I, sCorrMean), I.nBlocks, sm) / static_cast<double>(I.N);
}

// p, u and v of cell c (SimpleSolver.cpp:582-632).
__device__ __forceinline__ void correctCell(const Inst& I, int c, int i, int j, double mean) {
  const int nI = I.nI, nJ = I.nJ;
  auto corrAt = [&](const Inst& J, int q) { return J.hasPRef ? J.x[q] : J.x[q] - mean; };
  double gx = 0.0, gy = 0.0;
  const double pc = corrAt(I, c);
  auto add = [&](const FaceGeo& f, double val, bool plus) {
    if (plus) { gx += f.Sx * val; gy += f.Sy * val; }
    else { gx -= f.Sx * val; gy -= f.Sy * val; }
  };
  {
    const int fi = j * nI + i;
    const FaceGeo f = geo(I.jg, fi);
    if (j > 0) add(f, f.w * corrAt(I, c - nI) + (1.0 - f.w) * pc, false);
    else {
      const int pr = I.partner[i];
      if (pr >= 0) add(f, f.w * pc + (1.0 - f.w) * corrAt(I, pr), true);
      else add(f, imposes(I.bcB[i], I.Fj[fi]) ? pc : 0.0, true);
    }
  }
  {
    const int fi = (j + 1) * nI + i;
    const FaceGeo f = geo(I.jg, fi);
    if (j < nJ - 1) add(f, f.w * pc + (1.0 - f.w) * corrAt(I, c + nI), true);
    else add(f, imposes(I.bcT[i], I.Fj[fi]) ? pc : 0.0, true);
  }
  {
    const int fi = j * (nI + 1) + i;
    const FaceGeo f = geo(I.ig, fi);
    if (i > 0) add(f, f.w * corrAt(I, c - 1) + (1.0 - f.w) * pc, false);
    else add(f, imposes(I.bcL[j], I.Fi[fi]) ? pc : 0.0, true);
  }
  {
    const int fi = j * (nI + 1) + i + 1;
    const FaceGeo f = geo(I.ig, fi);
    if (i < nI - 1) add(f, f.w * pc + (1.0 - f.w) * corrAt(I, c + 1), true);
    else add(f, imposes(I.bcR[j], I.Fi[fi]) ? pc : 0.0, true);
  }
  const double vol = I.V[c];
  if (vol > 0.0) {
    const double inv = 1.0 / vol;
    gx *= inv;
    gy *= inv;
  }
  I.p[c] += I.alphaP * pc;
  const double mob = vol / I.aP[c];
  I.u[c] -= gx * mob;
  I.v[c] -= gy * mob;
}

__device__ __forceinline__ double corrFlux(const Inst& I, const FaceGeo& f, int o, int n, double mean) {
  const double w = f.w;
  const double dens = w * I.rho[o] + (1.0 - w) * I.rho[n];
  const double mob = w * (I.V[o] / I.aP[o]) + (1.0 - w) * (I.V[n] / I.aP[n]);
  const double pn = I.hasPRef ? I.x[n] : I.x[n] - mean;
  const double po = I.hasPRef ? I.x[o] : I.x[o] - mean;
  return dens * mob * f.g * (pn - po);
}

__device__ __forceinline__ void corrBoundaryFlux(const Inst& I, const FaceGeo& f, const FaceBC& bc, double* F,
                                                 int o, double mean) {
  if (imposes(bc, *F)) return;
  const double dens = I.rho[o];
  const double mob = I.V[o] / I.aP[o];
  const double po = I.hasPRef ? I.x[o] : I.x[o] - mean;
  *F -= dens * mob * f.g * (0.0 - po);
}

// The face fluxes cell c owns: its j-face and i-face, and the top and right
// boundary faces on the last row and column (SimpleSolver.cpp:634-660).
__device__ __forceinline__ void correctFaces(const Inst& I, int c, int i, int j, double mean) {
  const int nI = I.nI, nJ = I.nJ;
  {
    const int fi = j * nI + i;
    const FaceGeo f = geo(I.jg, fi);
    if (j > 0) I.Fj[fi] -= corrFlux(I, f, c - nI, c, mean);
    else {
      const int pr = I.partner[i];
      if (pr >= 0) I.Fj[fi] -= corrFlux(I, f, c, pr, mean);
      else corrBoundaryFlux(I, f, I.bcB[i], &I.Fj[fi], c, mean);
    }
  }
  if (j == nJ - 1) {
    const int fi = (j + 1) * nI + i;
    corrBoundaryFlux(I, geo(I.jg, fi), I.bcT[i], &I.Fj[fi], c, mean);
  }
  {
    const int fi = j * (nI + 1) + i;
    const FaceGeo f = geo(I.ig, fi);
    if (i > 0) I.Fi[fi] -= corrFlux(I, f, c - 1, c, mean);
    else corrBoundaryFlux(I, f, I.bcL[j], &I.Fi[fi], c, mean);
  }
  if (i == nI - 1) {
    const int fi = j * (nI + 1) + i + 1;
    corrBoundaryFlux(I, geo(I.ig, fi), I.bcR[j], &I.Fi[fi], c, mean);
  }
}

// The correction of one pressure pass: every cell's p, u, v, then the face
// fluxes it owns. A cell's own update reads the old flux only on the boundary
// faces it owns itself, before correcting them, so one kernel does both stages
// with the reference's results.
__global__ void __launch_bounds__(kBlock) k_correct(const Inst* insts) {
  CFD_PROLOGUE
  __shared__ double sm[4];
  const double mean = corrMean(I, sm);
  if (!live) return;
  correctCell(I, c, i, j, mean);
  correctFaces(I, c, i, j, mean);
}

// SimpleSolver::divergence (SimpleSolver.cpp:796-817) -> max |div|.
__device__ __forceinline__ double divergenceAt(const Inst& I, int c, int i, int j) {
  const int nI = I.nI, nJ = I.nJ;
  constexpr double kTiny = 1e-300;
  double net = 0.0;
  if (j > 0) net -= I.Fj[j * nI + i] / mx(I.rho[c - nI], kTiny);
  else net += I.Fj[i] / mx(I.rho[c], kTiny);
  net += I.Fj[(j + 1) * nI + i] / mx(I.rho[c], kTiny);
  (void)nJ;
  if (i > 0) net -= I.Fi[j * (nI + 1) + i] / mx(I.rho[c - 1], kTiny);
  else net += I.Fi[j * (nI + 1) + i] / mx(I.rho[c], kTiny);
  net += I.Fi[j * (nI + 1) + i + 1] / mx(I.rho[c], kTiny);
  const double a = I.V[c];
  return (a > 0.0) ? net / a : 0.0;
}

// Four cells per thread (quadBlocks): the block's max goes to the first of
// its tiles' partial slots and zero, the identity of a max of |d|, to the
// others, so the monitor's reduction over nBlocks slots is unchanged.
__global__ void __launch_bounds__(kBlock) k_divergence(const Inst* insts, double* out) {
  const Inst& I = insts[blockIdx.y];
  const int nq = quadBlocks(I.nBlocks);
  if (static_cast<int>(blockIdx.x) >= nq) return;
  if (I.flags[fStopped]) return;
  const int tile = xcdTile(static_cast<int>(blockIdx.x), nq);
  __shared__ double sm[16];
  const int nI = I.nI;
  double d[kQuad];
#pragma unroll
  for (int q = 0; q < kQuad; ++q) {
    const int c = (tile * kQuad + q) * kBlock + static_cast<int>(threadIdx.x);
    d[q] = (c < I.N) ? divergenceAt(I, c, c % nI, c / nI) : 0.0;
  }
  double part[1] = {0.0};
#pragma unroll
  for (int q = 0; q < kQuad; ++q) {
    const int c = (tile * kQuad + q) * kBlock + static_cast<int>(threadIdx.x);
    if (c < I.N) {
      if (out) out[c] = d[q];
      part[0] = mx(part[0], fabs(d[q]));
    }
  }
  blockMax<1>(part, sm);
  if (threadIdx.x == 0) {
    for (int q = 0; q < kQuad; ++q) {
      const int slot = tile * kQuad + q;
      if (slot < I.nBlocks) partialSlot(I, sMaxDiv)[slot] = (q == 0) ? part[0] : 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// Monitor assembly + the caller's stop rule (SolverWorker.cpp:161-192).
// One thread per instance.
// ---------------------------------------------------------------------------

// One 64-lane block per instance: finishes the reductions of the iteration's
// kernels from their per-block partials, writes the monitor and applies the
// stop rule.
__global__ void __launch_bounds__(64) k_monitor(const Inst* insts) {
  const Inst& I = insts[blockIdx.y];
  if (I.flags[fStopped]) return;
  const int n = I.nBlocks;
  double sums[4], maxs[2];
  {
    const double* const s4[4] = {partialSlot(I, sMomResX), partialSlot(I, sMomResY), partialSlot(I, sMomDiagL1),
                                 partialSlot(I, sNetFluxL1)};
    wavePartials<4>(s4, n, 0.0, OpSum{}, sums);
    const double* const m2[2] = {partialSlot(I, sVelScale), partialSlot(I, sMaxDiv)};
    wavePartials<2>(m2, n, 0.0, OpMax{}, maxs);
  }
  const double resX = sums[0], resY = sums[1], diagL1 = sums[2], net = sums[3];
  const double vscale = maxs[0], maxDiv = maxs[1];
  double tk = 0.0, tkd = 0.0, tw = 0.0, twd = 0.0, mxk = 0.0, mxw = 0.0, mxm = 0.0, mxr = 0.0, mnk = 0.0, mnw = 0.0,
         mnm = 0.0;
  if (I.model == 2) {
    double t4[4], hi[3], lo[3], er[1];
    const double* const s4[4] = {partialSlot(I, sTurbResK), partialSlot(I, sTurbDiagK), partialSlot(I, sTurbResW),
                                 partialSlot(I, sTurbDiagW)};
    wavePartials<4>(s4, n, 0.0, OpSum{}, t4);
    const double* const h3[3] = {partialSlot(I, sMaxK), partialSlot(I, sMaxW), partialSlot(I, sMaxMut)};
    wavePartials<3>(h3, n, -INFINITY, OpMax{}, hi);
    const double* const e1[1] = {partialSlot(I, sMaxEddyRatio)};
    wavePartials<1>(e1, n, 0.0, OpMax{}, er);
    const double* const l3[3] = {partialSlot(I, sMinK), partialSlot(I, sMinW), partialSlot(I, sMinMut)};
    wavePartials<3>(l3, n, INFINITY, OpMin{}, lo);
    tk = t4[0]; tkd = t4[1]; tw = t4[2]; twd = t4[3];
    mxk = hi[0]; mxw = hi[1]; mxm = hi[2]; mxr = er[0];
    mnk = lo[0]; mnw = lo[1]; mnm = lo[2];
  }
  if (threadIdx.x != 0) return;
  constexpr double kTiny = 1e-300;
  const double vs = (vscale > 0.0) ? vscale : 1.0;  // SimpleSolver.cpp:681-683
  DevMonitor m{};
  m.continuity = net / I.refMassFlow;
  m.massImbalance = net;
  const double scale = diagL1 * vs;
  m.momentumX = resX / mx(scale, kTiny);
  m.momentumY = resY / mx(scale, kTiny);
  m.maxDivergence = maxDiv;
  m.pressureIterations = I.flags[fCgTotal];
  if (I.model == 2) {
    m.turbFirst = tk / mx(tkd * mx(I.kIn, 1e-30), 1e-30);
    m.turbSecond = tw / mx(twd * mx(I.wIn, 1e-30), 1e-30);
    m.minK = mnk;
    m.maxK = mxk;
    m.minW = mnw;
    m.maxW = mxw;
    m.minMut = mnm;
    m.maxMut = mxm;
    m.maxEddyRatio = mxr;
  }
  *I.monitor = m;
  const int iter = I.flags[fIteration] + 1;
  I.flags[fIteration] = iter;
  if (I.history && I.historyCap > 0) I.history[(iter - 1) % I.historyCap] = m;
  if (I.stopTol > 0.0 || I.stopMaxIter > 0) {
    const bool finite = isfinite(m.continuity) && isfinite(m.momentumX) && isfinite(m.momentumY);
    const double worst = mx(mx(fabs(m.continuity), fabs(m.momentumX)), fabs(m.momentumY));
    if (!finite) {
      I.flags[fStatus] = 2;
      I.flags[fStopped] = 1;
    } else if (worst < I.stopTol) {
      I.flags[fStatus] = 0;
      I.flags[fStopped] = 1;
    } else if (I.stopMaxIter > 0 && iter >= I.stopMaxIter) {
      I.flags[fStatus] = 1;
      I.flags[fStopped] = 1;
    }
  }
}

// Reset the per-iteration CG total before the pressure passes.
__global__ void k_begin_iteration(const Inst* insts, int count) {
  const int b = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (b >= count) return;
  const Inst& I = insts[b];
  if (I.flags[fStopped]) return;
  I.flags[fCgTotal] = 0;
}

}  // namespace cfdgpu
