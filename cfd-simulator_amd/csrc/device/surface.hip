// surface.hip - surface extraction and force integration on the device, FP64
// (SurfaceData.cpp:27-348, Forces.cpp:16-130), for one solver's field. The
// stations are evaluated in parallel; the stagnation point is the first
// maximum of the wall pressure, and the separation scans and the force sums
// run in the reference's station order on one thread, so the result is the
// host routine's (csrc/host/post.cpp) bit for bit.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "surface.hpp"

namespace cfdgpu {

namespace {

constexpr int kSurfaceThreads = 256;

// SurfaceData.cpp:179-204 over rows [first, first + n)
__device__ double separationScan(const double* rows, int first, int n, int* face) {
  for (int s = 1; s < n; ++s) {
    const double* a = rows + 12 * (first + s - 1);
    const double* b = rows + 12 * (first + s);
    if (b[10] != 0.0 || a[10] != 0.0) continue;  // next to the stagnation point
    const double before = a[8], after = b[8];
    if (!(before > 0.0) || !(after < 0.0)) continue;
    const double span = before - after;
    const double fr = (span > 0.0) ? before / span : 0.0;
    *face = static_cast<int>(b[1]);
    return a[4] + fr * (b[4] - a[4]);
  }
  *face = -1;
  return -1.0;
}

}  // namespace

// rows: 12 per station, upper surface (leading edge to the trailing edge)
// then lower; terms: 6 per station (the pressure and friction force and
// moment contributions of Forces.cpp:60-104); fd / fi: the forces record.
__global__ void __launch_bounds__(kSurfaceThreads)
k_surface(const double* u, const double* v, const double* p, const double* mu, const WallStation* st, int nWall,
          int le, double q, double pref, double chord, double ca, double sa, double* rows, double* terms, double* fd,
          int* fi) {
  __shared__ double bestV[kSurfaceThreads];
  __shared__ int bestK[kSurfaceThreads];
  __shared__ int stagShared;
  const int t = static_cast<int>(threadIdx.x);
  // the stagnation point: the first face with the highest pressure
  {
    double bv = -1.7976931348623157e308;
    int bk = 0;
    for (int k = t; k < nWall; k += kSurfaceThreads) {
      const double pv = p[st[k].owner];
      if (pv > bv) {
        bv = pv;
        bk = k;
      }
    }
    bestV[t] = bv;
    bestK[t] = bk;
  }
  __syncthreads();
  if (t == 0) {
    double bv = bestV[0];
    int bk = bestK[0];
    for (int s = 1; s < kSurfaceThreads; ++s) {
      if (bestV[s] > bv || (bestV[s] == bv && bestK[s] < bk)) {
        bv = bestV[s];
        bk = bestK[s];
      }
    }
    stagShared = bk;
  }
  __syncthreads();
  const int stag = stagShared;
  const int nUpper = nWall - le;
  for (int k = t; k < nWall; k += kSurfaceThreads) {
    const WallStation S = st[k];
    // the tangent points with the flow away from the stagnation point
    const bool along = (k == stag) ? (k >= le) : (k > stag);
    const double tx = along ? S.tx : S.tx * -1.0;
    const double ty = along ? S.ty : S.ty * -1.0;
    const double pk = p[S.owner];
    const double cp = (q > 0.0) ? (pk - pref) / q : 0.0;
    double tau = 0.0;
    if (S.wd > 0.0) {
      const double cvx = u[S.owner], cvy = v[S.owner];
      const double dn = cvx * S.nx + cvy * S.ny;
      const double parx = cvx - S.nx * dn, pary = cvy - S.ny * dn;
      double grad = (parx * tx + pary * ty) / S.wd;
      if (S.second >= 0) {
        const double ovx = u[S.second], ovy = v[S.second];
        const double on = ovx * S.nx + ovy * S.ny;
        const double oparx = ovx - S.nx * on, opary = ovy - S.ny * on;
        const double u1 = parx * tx + pary * ty, u2 = oparx * tx + opary * ty;
        const double d1 = S.wd, d2 = S.od;
        grad = (u1 * d2 * d2 - u2 * d1 * d1) / (d1 * d2 * (d2 - d1));
      }
      tau = mu[S.owner] * grad;
    }
    const double cf = (q > 0.0) ? tau / q : 0.0;
    const double xc = (chord > 0.0) ? S.cx / chord : S.cx;
    const int row = (k >= le) ? k - le : nUpper + (le - 1 - k);
    double* r = rows + 12 * row;
    r[0] = (k >= le) ? 1.0 : 0.0;
    r[1] = static_cast<double>(S.face);
    r[2] = S.cx;
    r[3] = S.cy;
    r[4] = xc;
    r[5] = S.arc;
    r[6] = pk;
    r[7] = cp;
    r[8] = tau;
    r[9] = cf;
    r[10] = (k == stag) ? 1.0 : 0.0;
    r[11] = S.len;
    // Forces.cpp: dP = n (-(p - pref) ds), dF = t (tau ds), moments about c/4
    const double ds = S.len;
    const double sP = -(pk - pref) * ds;
    const double dPx = S.nx * sP, dPy = S.ny * sP;
    const double sF = tau * ds;
    const double dFx = tx * sF, dFy = ty * sF;
    const double armx = S.cx - 0.25 * chord, army = S.cy - 0.0;
    double* w = terms + 6 * row;
    w[0] = dPx;
    w[1] = dPy;
    w[2] = dFx;
    w[3] = dFy;
    w[4] = armx * dPy - army * dPx;
    w[5] = armx * dFy - army * dFx;
  }
  __syncthreads();
  if (t != 0) return;
  double pfx = 0.0, pfy = 0.0, ffx = 0.0, ffy = 0.0, pm = 0.0, fm = 0.0;
  int stations = 0;
  for (int r = 0; r < nWall; ++r) {
    if (!(rows[12 * r + 11] > 0.0)) continue;
    const double* w = terms + 6 * r;
    pfx = pfx + w[0];
    pfy = pfy + w[1];
    ffx = ffx + w[2];
    ffy = ffy + w[3];
    pm += w[4];
    fm += w[5];
    ++stations;
  }
  int upperFace = -1, lowerFace = -1;
  const double upperSep = separationScan(rows, 0, nUpper, &upperFace);
  const double lowerSep = separationScan(rows, nUpper, le, &lowerFace);
  const double pd = pfx * ca + pfy * sa, fdr = ffx * ca + ffy * sa;
  const double pl = pfy * ca - pfx * sa, fl = ffy * ca - ffx * sa;
  const double drag = pd + fdr, lift = pl + fl, moment = -pm - fm;
  const double fs = q * chord, ms = fs * chord;
  fd[0] = lift / fs;    // cl
  fd[1] = drag / fs;    // cd
  fd[2] = moment / ms;  // cm
  fd[3] = pd / fs;      // cd_pressure
  fd[4] = fdr / fs;     // cd_friction
  fd[5] = lift;
  fd[6] = drag;
  fd[7] = moment;
  fd[8] = pfx;
  fd[9] = pfy;
  fd[10] = ffx;
  fd[11] = ffy;
  fd[12] = upperSep;
  fd[13] = lowerSep;
  fd[14] = (chord > 0.0) ? st[stag].cx / chord : st[stag].cx;
  fi[0] = nUpper;
  fi[1] = le;
  fi[2] = stations;
  fi[3] = upperFace;
  fi[4] = lowerFace;
}

void launchSurface(const double* u, const double* v, const double* p, const double* mu, const WallStation* st,
                   int nWall, int le, double q, double pref, double chord, double ca, double sa, double* rows,
                   double* terms, double* fd, int* fi, hipStream_t s) {
  hipLaunchKernelGGL(k_surface, dim3(1), dim3(kSurfaceThreads), 0, s, u, v, p, mu, st, nWall, le, q, pref, chord,
                     ca, sa, rows, terms, fd, fi);
}

}  // namespace cfdgpu
