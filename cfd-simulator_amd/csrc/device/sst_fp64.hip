// sst_fp64.hip - k-omega SST update on the corrected field
// (reference: src/solver/KOmegaSST.cpp:531-548), FP64, cell-centric.
//
//   k_sst_blend : strain (start-of-iteration gradients), Green-Gauss of k and
//                 omega, cross diffusion, F1, F2, limited production
//                 (KOmegaSST.cpp:202-287, 357-368)
//   k_sst_asm   : both transport systems + sources + relaxation + residuals
//                 (:370-466, 475-510)
//   k_gs_lex / k_gs_color (systems 1, 2) : 3 sweeps + floors (:468-471)
//   k_sst_wall  : wall-cell omega and y+ (:305-345)
//   k_sst_mut   : limited eddy viscosity, mu_eff = mu + mu_t, ranges
//                 (:289-303, 513-529; SimpleSolver.cpp:232-245, 780-790)
#include "common.hpp"

namespace cfdgpu {

#define SST_PROLOGUE                                          \
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
  (void)nJ;

constexpr double kMinOmega = 1.0e-10;
constexpr double kTinySst = 1.0e-30;

// KOmegaSST.cpp:26-68: face value of k or omega on a boundary face.
__device__ __forceinline__ double sstBoundaryValue(const FaceBC& bc, double F, double cell, double wallValue,
                                                   double inflow) {
  switch (bc.kind) {
    case kWall: return wallValue;
    case kInlet: return inflow;
    case kFarField: return (F < 0.0) ? inflow : cell;
    default: return cell;
  }
}

__global__ void __launch_bounds__(kBlock) k_sst_blend(const Inst* insts) {
  SST_PROLOGUE
  if (!live) return;
  // strain from the gradients of the start of the iteration (KOmegaSST.cpp:202-216)
  const double2 gu = ld2(I.gU, c), gv = ld2(I.gV, c);
  const double s11 = gu.x, s22 = gv.y, s12 = 0.5 * (gu.y + gv.x);
  const double S = sqrt(2.0 * (s11 * s11 + s22 * s22 + 2.0 * s12 * s12));
  I.strain[c] = S;

  // Green-Gauss of k (wall value 0) and omega (wall faces take the owner value)
  const double kc = I.k[c], wc = I.w[c];
  double kx = 0.0, ky = 0.0, wx = 0.0, wy = 0.0;
  auto add = [&](const FaceGeo& f, double fk, double fw, bool plus) {
    if (plus) { kx += f.Sx * fk; ky += f.Sy * fk; wx += f.Sx * fw; wy += f.Sy * fw; }
    else { kx -= f.Sx * fk; ky -= f.Sy * fk; wx -= f.Sx * fw; wy -= f.Sy * fw; }
  };
  auto interior = [&](const FaceGeo& f, int o, int n, bool plus) {
    const double w = f.w;
    add(f, w * I.k[o] + (1.0 - w) * I.k[n], w * I.w[o] + (1.0 - w) * I.w[n], plus);
  };
  auto boundary = [&](const FaceGeo& f, const FaceBC& bc, double F) {
    const double fk = sstBoundaryValue(bc, F, kc, 0.0, I.kIn);
    const double fw = (bc.kind == kWall) ? wc : sstBoundaryValue(bc, F, wc, 0.0, I.wIn);
    add(f, fk, fw, true);
  };
  {
    const int fi = j * nI + i;
    const FaceGeo f = geo(I.jg, fi);
    if (j > 0) interior(f, c - nI, c, false);
    else {
      const int pr = I.partner[i];
      if (pr >= 0) {
        // the wake face has no neighbour, so ownerWeight() is 1 (Gradient.cpp:9-11)
        add(f, 1.0 * kc + (1.0 - 1.0) * I.k[pr], 1.0 * wc + (1.0 - 1.0) * I.w[pr], true);
      } else {
        boundary(f, I.bcB[i], I.Fj[fi]);
      }
    }
  }
  {
    const int fi = (j + 1) * nI + i;
    const FaceGeo f = geo(I.jg, fi);
    if (j < nJ - 1) interior(f, c, c + nI, true);
    else boundary(f, I.bcT[i], I.Fj[fi]);
  }
  {
    const int fi = j * (nI + 1) + i;
    const FaceGeo f = geo(I.ig, fi);
    if (i > 0) interior(f, c - 1, c, false);
    else boundary(f, I.bcL[j], I.Fi[fi]);
  }
  {
    const int fi = j * (nI + 1) + i + 1;
    const FaceGeo f = geo(I.ig, fi);
    if (i < nI - 1) interior(f, c, c + 1, true);
    else boundary(f, I.bcR[j], I.Fi[fi]);
  }
  const double vol = I.V[c];
  if (vol > 0.0) {
    const double inv = 1.0 / vol;
    kx *= inv; ky *= inv; wx *= inv; wy *= inv;
  }
  st2(I.gK, c, kx, ky);
  st2(I.gW, c, wx, wy);

  // blending functions (KOmegaSST.cpp:247-286)
  const double d = mx(I.wallDist[c], kTinySst);
  const double rho = I.rho[c];
  const double nu = I.muMol[c] / mx(rho, kTinySst);
  const double omega = mx(wc, kMinOmega);
  const double energy = mx(kc, 0.0);
  const double cd = 2.0 * rho * I.sigW2 * (kx * wx + ky * wy) / omega;
  const double cdPlus = mx(cd, 1.0e-10);
  const double term1 = sqrt(mx(energy, 0.0)) / (I.betaStar * omega * d);
  const double term2 = 500.0 * nu / (d * d * omega);
  const double term3 = 4.0 * rho * I.sigW2 * energy / (cdPlus * d * d);
  const double arg1 = mn(mx(term1, term2), term3);
  const double f1 = tanh(clampd(arg1 * arg1 * arg1 * arg1, 0.0, 50.0));
  const double arg2 = mx(2.0 * term1, term2);
  const double f2 = tanh(clampd(arg2 * arg2, 0.0, 50.0));
  I.cd[c] = cd;
  I.f1[c] = f1;
  I.f2[c] = f2;
  // limited production (KOmegaSST.cpp:362-368)
  const double raw = I.mut[c] * S * S;
  const double ceiling = I.prodLimit * I.betaStar * rho * mx(kc, 0.0) * mx(wc, kMinOmega);
  I.prod[c] = mn(raw, ceiling);
}

// One transport equation's row for cell c.
struct SstRow {
  double d, b, aS, aN, aW, aE;
};

// The diffusivity mu + sigma(F1) mu_t of the k (isK) or omega row at a cell
// (KOmegaSST.cpp:384-444).
template <bool isK>
__device__ __forceinline__ double sstGamma(const Inst& I, double muMol, double f1, double mut) {
  return muMol + (isK ? f1 * I.sigK1 + (1.0 - f1) * I.sigK2 : f1 * I.sigW1 + (1.0 - f1) * I.sigW2) * mut;
}

// One face of a row (KOmegaSST.cpp:384-444). An interior face (owner value
// gO, neighbour value gN of the diffusivity, owner weight w) adds its
// outflow, or inflow, and its diffusion to the diagonal and gives the
// off-diagonal coefficient; a boundary face (the cell's diffusivity gC)
// adds to the diagonal and the source, its value imposed on a wall or
// inlet face and on an inflowing far-field face. Both are formed and a
// select keeps one, so no load waits on the face's kind.
__device__ __forceinline__ double sstFace(bool interior, bool owner, double w, double gO, double gN, double gs,
                                          double F, double bcG, int bcKind, double gC, double wallValue,
                                          double inflow, double& d, double& b) {
  const double diff = (w * gO + (1.0 - w) * gN) * gs;
  const double outward = mx(F, 0.0), inward = mx(-F, 0.0);
  const double dI = owner ? outward + diff : inward + diff;
  const double aI = owner ? -inward - diff : -outward - diff;
  const double diffB = gC * bcG;
  const bool imposed = bcKind == kWall || bcKind == kInlet || (bcKind == kFarField && F < 0.0);
  const double value = (bcKind == kWall) ? wallValue : inflow;
  const double addD = imposed ? diffB + mx(F, 0.0) : mx(F, 0.0);
  const double addB = imposed ? (diffB - mn(F, 0.0)) * value : 0.0;
  d += interior ? dI : addD;
  // +0 on an interior face leaves b unchanged: a sum of doubles is -0 only
  // when both terms are, so b, which starts at +0, is never -0
  b += interior ? 0.0 : addB;
  return interior ? aI : 0.0;
}

// The stencil of cell c: its four faces in cellFaces order (below or the
// wake partner, above, left, right; Mesh.cpp:206-213), the neighbour across
// each (the cell itself where there is none) and what the faces carry.
struct SstStencil {
  int nb[4];
  bool interior[4];
  bool sOwner;  // row 0: the cell owns its face below (a wake face)
  double w[4], gs[4], F[4], bcG[4];
  int bcKind[4];
};

__device__ __forceinline__ SstStencil sstStencil(const Inst& I, int c, int i, int j) {
  const int nI = I.nI, nJ = I.nJ;
  SstStencil t;
  // the wake partner of a row-0 cell: the only operand whose load waits on another
  const int pr = (j == 0) ? I.partner[i] : -1;
  const int fS = j * nI + i, fN = (j + 1) * nI + i, fW = j * (nI + 1) + i, fE = fW + 1;
  t.nb[0] = (j > 0) ? c - nI : ((pr >= 0) ? pr : c);
  t.nb[1] = (j < nJ - 1) ? c + nI : c;
  t.nb[2] = (i > 0) ? c - 1 : c;
  t.nb[3] = (i < nI - 1) ? c + 1 : c;
  t.sOwner = j == 0;
  t.interior[0] = j > 0 || pr >= 0;
  t.interior[1] = j < nJ - 1;
  t.interior[2] = i > 0;
  t.interior[3] = i < nI - 1;
  // a wake face weights the cell itself fully (owner c, neighbour the partner)
  t.w[0] = (j > 0) ? I.jg.w[fS] : 1.0;
  t.w[1] = I.jg.w[fN];
  t.w[2] = I.ig.w[fW];
  t.w[3] = I.ig.w[fE];
  t.gs[0] = I.jgSst[fS];
  t.gs[1] = I.jgSst[fN];
  t.gs[2] = I.igSst[fW];
  t.gs[3] = I.igSst[fE];
  t.F[0] = I.Fj[fS];
  t.F[1] = I.Fj[fN];
  t.F[2] = I.Fi[fW];
  t.F[3] = I.Fi[fE];
  const FaceBC* const bcs[4] = {&I.bcB[i], &I.bcT[i], &I.bcL[j], &I.bcR[j]};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    t.bcG[q] = bcs[q]->gSst;
    t.bcKind[q] = bcs[q]->kind;
  }
  return t;
}

// The k (isK) or omega row of cell c from its stencil and the diffusivities
// of the cell (g[4]) and of its neighbours (g[0..3], in the stencil's order).
template <bool isK>
__device__ __forceinline__ SstRow sstRow(const SstStencil& t, const double (&g)[5], double wallValue,
                                         double inflow) {
  SstRow r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const double gC = g[4];
  // below: the face's owner is the cell below (neighbour the cell), or on a
  // wake face the cell itself (neighbour its partner)
  r.aS = sstFace(t.interior[0], t.sOwner, t.w[0], t.sOwner ? gC : g[0], t.sOwner ? g[0] : gC, t.gs[0], t.F[0],
                 t.bcG[0], t.bcKind[0], gC, wallValue, inflow, r.d, r.b);
  r.aN = sstFace(t.interior[1], true, t.w[1], gC, g[1], t.gs[1], t.F[1], t.bcG[1], t.bcKind[1], gC, wallValue,
                 inflow, r.d, r.b);
  r.aW = sstFace(t.interior[2], false, t.w[2], g[2], gC, t.gs[2], t.F[2], t.bcG[2], t.bcKind[2], gC, wallValue,
                 inflow, r.d, r.b);
  r.aE = sstFace(t.interior[3], true, t.w[3], gC, g[3], t.gs[3], t.F[3], t.bcG[3], t.bcKind[3], gC, wallValue,
                 inflow, r.d, r.b);
  return r;
}

// b - A phi of one row (the signed defect; its magnitude is the residual
// term), every neighbour read without a branch (a missing one at the cell
// itself, its term dropped by a select: ax may then differ from the
// branching sum only in the sign of a zero, which b - ax does not see, b
// never being -0)
__device__ __forceinline__ double sstDefect(const SstStencil& t, const SstRow& r, const CFD_G double* phi, int c) {
  const double pc = phi[c], p0 = phi[t.nb[0]], p1 = phi[t.nb[1]], p2 = phi[t.nb[2]], p3 = phi[t.nb[3]];
  double ax = r.d * pc;
  ax += t.interior[0] ? r.aS * p0 : 0.0;
  ax += t.interior[1] ? r.aN * p1 : 0.0;
  ax += t.interior[2] ? r.aW * p2 : 0.0;
  ax += t.interior[3] ? r.aE * p3 : 0.0;
  return r.b - ax;
}

// The lexicographic sweep's copy of one row and the current value, in the
// anti-diagonal layout: pairs (aP, aS), (aN, aW), (aE, b) and (phi, -).
__device__ __forceinline__ void skewRow(const Inst& I, int sys, int at, const SstRow& r, double phi) {
  I.skTurbRow[sys][0][at] = make_double2(r.d, r.aS);
  I.skTurbRow[sys][1][at] = make_double2(r.aN, r.aW);
  I.skTurbRow[sys][2][at] = make_double2(r.aE, r.b);
  I.skTurbX[sys][0][at] = make_double2(phi, 0.0);
}

// The mixed path's FP32 copy of one row (two-colour defect sweeps).
__device__ __forceinline__ void mixedTurbRow(const Inst& I, int sys, int pk, const SstRow& r) {
  I.fTurb[sys][0][pk] = static_cast<float>(r.d);
  I.fTurb[sys][1][pk] = static_cast<float>(r.aS);
  I.fTurb[sys][2][pk] = static_cast<float>(r.aN);
  I.fTurb[sys][3][pk] = static_cast<float>(r.aW);
  I.fTurb[sys][4][pk] = static_cast<float>(r.aE);
}

__global__ void __launch_bounds__(kBlock) k_sst_asm(const Inst* insts) {
  SST_PROLOGUE
  __shared__ double sm[16];
  double part[4] = {0.0, 0.0, 0.0, 0.0};
  if (live) {
    const SstStencil t = sstStencil(I, c, i, j);
    // the diffusivities of the neighbours and of the cell, their operands
    // loaded once for both rows
    double gk[5], gw[5];
    {
      const int q5[5] = {t.nb[0], t.nb[1], t.nb[2], t.nb[3], c};
      double mu[5], fo[5], mt[5];
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        mu[q] = I.muMol[q5[q]];
        fo[q] = I.f1[q5[q]];
        mt[q] = I.mut[q5[q]];
      }
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        gk[q] = sstGamma<true>(I, mu[q], fo[q], mt[q]);
        gw[q] = sstGamma<false>(I, mu[q], fo[q], mt[q]);
      }
    }
    const double vol = I.V[c];
    const double rho = I.rho[c];
    const double f1 = I.f1[c];
    const double S = I.strain[c];
    const double relax = clampd(I.alphaT, 0.05, 1.0);
    // k equation
    {
      SstRow r = sstRow<true>(t, gk, 0.0, I.kIn);
      const double srcE = I.prod[c];
      const double srcI = I.betaStar * rho * mx(I.w[c], kMinOmega);
      r.b += srcE * vol;
      r.d += srcI * vol;
      const double phi = I.k[c];
      if (fabs(r.d) > kTinySst) {
        const double relaxed = r.d / relax;
        r.b += (relaxed - r.d) * phi;
        r.d = relaxed;
      }
      if (I.gsOrder == 0) {
        skewRow(I, 0, skewIndex(i, j, I.gsR), r, phi);
      } else if (I.precision == 0) {
        I.kP[c] = r.d; I.kS[c] = r.aS; I.kN[c] = r.aN; I.kW[c] = r.aW; I.kE[c] = r.aE; I.kB[c] = r.b;
      } else {
        mixedTurbRow(I, 0, rbIndex(i, j, nI, nJ), r);
      }
      const double defect = sstDefect(t, r, I.k, c);
      if (I.gsOrder == 1 && I.precision == 1) I.fTurbRes[0][rbIndex(i, j, nI, nJ)] = static_cast<float>(defect);
      part[0] = fabs(defect);
      part[1] = fabs(r.d);
    }
    // omega equation (wall value is omega_inflow: KOmegaSST.cpp:507-509)
    {
      SstRow r = sstRow<false>(t, gw, I.wIn, I.wIn);
      const double gamma = f1 * I.gamma1 + (1.0 - f1) * I.gamma2;
      const double beta = f1 * I.beta1 + (1.0 - f1) * I.beta2;
      double srcE = gamma * rho * S * S;
      srcE += (1.0 - f1) * I.cd[c];
      const double srcI = beta * rho * mx(I.w[c], kMinOmega);
      r.b += srcE * vol;
      r.d += srcI * vol;
      const double phi = I.w[c];
      if (fabs(r.d) > kTinySst) {
        const double relaxed = r.d / relax;
        r.b += (relaxed - r.d) * phi;
        r.d = relaxed;
      }
      if (I.gsOrder == 0) {
        skewRow(I, 1, skewIndex(i, j, I.gsR), r, phi);
      } else if (I.precision == 0) {
        I.wP[c] = r.d; I.wS[c] = r.aS; I.wN[c] = r.aN; I.wW[c] = r.aW; I.wE[c] = r.aE; I.wB[c] = r.b;
      } else {
        mixedTurbRow(I, 1, rbIndex(i, j, nI, nJ), r);
      }
      const double defect = sstDefect(t, r, I.w, c);
      if (I.gsOrder == 1 && I.precision == 1) I.fTurbRes[1][rbIndex(i, j, nI, nJ)] = static_cast<float>(defect);
      part[2] = fabs(defect);
      part[3] = fabs(r.d);
    }
  }
  blockSum<4>(part, sm);
  if (threadIdx.x == 0) {
    partialSlot(I, sTurbResK)[blockIdx.x] = part[0];
    partialSlot(I, sTurbDiagK)[blockIdx.x] = part[1];
    partialSlot(I, sTurbResW)[blockIdx.x] = part[2];
    partialSlot(I, sTurbDiagW)[blockIdx.x] = part[3];
  }
}

// KOmegaSST.cpp:86-92
__device__ __forceinline__ double wallOmega(const Inst& I, double nu, double distance, double energy) {
  const double d = mx(distance, kTinySst);
  const double viscous = 6.0 * nu / (I.beta1 * d * d);
  const double logarithmic = sqrt(mx(energy, 0.0)) / (I.betaStarQ * I.kappa * d);
  return sqrt(viscous * viscous + logarithmic * logarithmic);
}

// One block per instance walks the four boundary edges (KOmegaSST.cpp:305-345).
__global__ void __launch_bounds__(kBlock) k_sst_wall(const Inst* insts) {
  const Inst& I = insts[blockIdx.y];
  if (I.flags[fStopped]) return;
  __shared__ double sm[16];
  const int nI = I.nI, nJ = I.nJ;
  const int total = 2 * nI + 2 * nJ;
  double ymax[1] = {0.0};
  for (int e = static_cast<int>(threadIdx.x); e < total; e += kBlock) {
    const FaceBC* bc;
    int fi, o;
    bool jFace;
    if (e < nI) { bc = &I.bcB[e]; fi = e; jFace = true; o = e; }
    else if (e < 2 * nI) { const int q = e - nI; bc = &I.bcT[q]; fi = nJ * nI + q; jFace = true; o = (nJ - 1) * nI + q; }
    else if (e < 2 * nI + nJ) { const int q = e - 2 * nI; bc = &I.bcL[q]; fi = q * (nI + 1); jFace = false; o = q * nI; }
    else { const int q = e - 2 * nI - nJ; bc = &I.bcR[q]; fi = q * (nI + 1) + nI; jFace = false; o = q * nI + nI - 1; }
    if (bc->kind != kWall) continue;
    const double fSx = jFace ? I.jg.Sx[fi] : I.ig.Sx[fi];
    const double fSy = jFace ? I.jg.Sy[fi] : I.ig.Sy[fi];
    const double d = mx(I.wallDist[o], kTinySst);
    const double rho = mx(I.rho[o], kTinySst);
    const double nu = I.muMol[o] / rho;
    I.w[o] = wallOmega(I, nu, d, I.k[o]);
    // y+ with the tangential velocity of the wall cell (KOmegaSST.cpp:337-343)
    const double area = hypot(fSx, fSy);
    const double nx = -(fSx / area), ny = -(fSy / area);
    const double uo = I.u[o], vo = I.v[o];
    const double un = uo * nx + vo * ny;
    const double px = uo - nx * un, py = vo - ny * un;
    const double tang = hypot(px, py);
    const double shear = I.muMol[o] * tang / d;
    const double fv = sqrt(shear / rho);
    ymax[0] = mx(ymax[0], fv * d / mx(nu, kTinySst));
  }
  blockMax<1>(ymax, sm);
  if (threadIdx.x == 0) I.red[sMaxYPlus] = ymax[0];
}

__global__ void __launch_bounds__(kBlock) k_sst_mut(const Inst* insts) {
  SST_PROLOGUE
  (void)i; (void)j; (void)nJ;
  __shared__ double sm[32];
  double mxv[4] = {-INFINITY, -INFINITY, -INFINITY, 0.0};  // max k, max w, max mut, max ratio
  double mnv[3] = {INFINITY, INFINITY, INFINITY};
  if (live) {
    const double energy = mx(I.k[c], 0.0);
    const double omega = mx(I.w[c], kMinOmega);
    const double den = mx(I.a1 * omega, I.strain[c] * I.f2[c]);
    const double mut = I.rho[c] * I.a1 * energy / mx(den, kTinySst);
    I.mut[c] = mut;
    I.mu[c] = I.muMol[c] + mut;
    mxv[0] = I.k[c]; mxv[1] = I.w[c]; mxv[2] = mut;
    mnv[0] = I.k[c]; mnv[1] = I.w[c]; mnv[2] = mut;
    const double m = I.muMol[c];
    if (m > 0.0) mxv[3] = mx(0.0, mut / m);
  }
  blockMax<4>(mxv, sm);
  double neg[3] = {-mnv[0], -mnv[1], -mnv[2]};
  blockMax<3>(neg, sm + 16);
  if (threadIdx.x == 0) {
    partialSlot(I, sMaxK)[blockIdx.x] = mxv[0];
    partialSlot(I, sMaxW)[blockIdx.x] = mxv[1];
    partialSlot(I, sMaxMut)[blockIdx.x] = mxv[2];
    partialSlot(I, sMaxEddyRatio)[blockIdx.x] = mxv[3];
    partialSlot(I, sMinK)[blockIdx.x] = -neg[0];
    partialSlot(I, sMinW)[blockIdx.x] = -neg[1];
    partialSlot(I, sMinMut)[blockIdx.x] = -neg[2];
  }
}

// Laminar closure (TurbulenceModel.cpp:67-77): mu_eff = mu + 0.
__global__ void __launch_bounds__(kBlock) k_laminar_mu(const Inst* insts) {
  SST_PROLOGUE
  (void)i; (void)j; (void)nJ;
  if (live) I.mu[c] = I.muMol[c] + 0.0;
}

// Wall distance for the SST set-up (TurbulenceModel.cpp:28-65): from each
// cell centroid to the nearest wall-face centre, by brute force over the wall
// faces, staged through LDS 256 at a time. The minimum is taken over squared
// distances in wall-face order and the square root last, as on the host, so
// the result is bit-identical to it.
__global__ void __launch_bounds__(kBlock) k_wall_distance(const double2* centroid, int n, const double2* walls,
                                                         int nw, double* out) {
  __shared__ double2 tile[kBlock];
  const int c = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
  const double2 p = (c < n) ? centroid[c] : make_double2(0.0, 0.0);
  double best = 1.7976931348623157e308;  // DBL_MAX
  for (int base = 0; base < nw; base += kBlock) {
    __syncthreads();
    if (base + static_cast<int>(threadIdx.x) < nw) tile[threadIdx.x] = walls[base + threadIdx.x];
    __syncthreads();
    const int m = min(kBlock, nw - base);
    for (int q = 0; q < m; ++q) {
      const double dx = p.x - tile[q].x, dy = p.y - tile[q].y;
      best = mn(best, dx * dx + dy * dy);
    }
  }
  if (c < n) out[c] = sqrt(best);
}

}  // namespace cfdgpu
