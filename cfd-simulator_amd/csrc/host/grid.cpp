// grid.cpp - NACA sections, spacing laws, C-grid and box generation, and the
// structured metrics. Arithmetic follows the reference statement for
// statement (citations per function) so the meshes are identical bit for bit.
#include "grid.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace cfdgpu {

namespace {
constexpr double kPi = 3.141592653589793238462643383279502884;

Vec2 unit(Vec2 v) {
  const double l = length(v);
  return (l > 0.0) ? v * (1.0 / l) : Vec2{0.0, 0.0};
}
Vec2 turnLeft(Vec2 v) { return Vec2{-v.y, v.x}; }
Vec2 turnRight(Vec2 v) { return Vec2{v.y, -v.x}; }

// --- NACA four-digit (src/geom/Naca4.cpp:103-168) ---
double halfThickness(double x, double t, bool closed) {
  if (!(x > 0.0)) return 0.0;
  const double xc = std::min(x, 1.0);
  const double a4 = closed ? -0.1036 : -0.1015;
  const double x2 = xc * xc, x3 = x2 * xc, x4 = x3 * xc;
  return 5.0 * t * (0.2969 * std::sqrt(xc) + -0.1260 * xc + -0.3516 * x2 + 0.2843 * x3 + a4 * x4);
}
double camberY(double x, double m, double p) {
  if (m == 0.0 || p <= 0.0 || p >= 1.0) return 0.0;
  const double xc = std::clamp(x, 0.0, 1.0);
  if (xc <= p) return (m / (p * p)) * (2.0 * p * xc - xc * xc);
  const double q = 1.0 - p;
  return (m / (q * q)) * ((1.0 - 2.0 * p) + 2.0 * p * xc - xc * xc);
}
double camberDyDx(double x, double m, double p) {
  if (m == 0.0 || p <= 0.0 || p >= 1.0) return 0.0;
  const double xc = std::clamp(x, 0.0, 1.0);
  const double den = (xc <= p) ? (p * p) : ((1.0 - p) * (1.0 - p));
  return (2.0 * m / den) * (p - xc);
}

std::vector<double> uniformSpacing(int intervals) {  // Distribution.cpp:8-19
  if (intervals < 1) return {};
  std::vector<double> t(static_cast<size_t>(intervals) + 1);
  for (int k = 0; k <= intervals; ++k) t[static_cast<size_t>(k)] = static_cast<double>(k) / static_cast<double>(intervals);
  t.back() = 1.0;
  return t;
}

double growthRatio(double first, double total, int n) {  // Distribution.cpp:21-51
  if (n <= 1 || first <= 0.0 || total <= 0.0) return 1.0;
  if (first * static_cast<double>(n) >= total) return 1.0;
  auto span = [first, n](double r) {
    if (std::abs(r - 1.0) < 1e-12) return first * static_cast<double>(n);
    return first * (std::pow(r, static_cast<double>(n)) - 1.0) / (r - 1.0);
  };
  double lo = 1.0, hi = 3.0;
  if (span(hi) < total) return hi;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    (span(mid) < total ? lo : hi) = mid;
  }
  return 0.5 * (lo + hi);
}

std::vector<double> mirroredGeometric(double first, double total, int n) {  // Distribution.cpp:78-109
  if (n < 2) return uniformSpacing(n);
  const int half = n / 2;
  const std::vector<double> lower = geometricDistribution(first, 0.5 * total, half);
  std::vector<double> t(static_cast<size_t>(n) + 1, 0.0);
  for (int k = 0; k <= half; ++k) t[static_cast<size_t>(k)] = 0.5 * lower[static_cast<size_t>(k)];
  for (int k = 0; k <= half; ++k) t[static_cast<size_t>(n - k)] = 1.0 - 0.5 * lower[static_cast<size_t>(k)];
  if (n % 2 != 0) {
    const size_t mid = static_cast<size_t>(half + 1);
    t[mid] = 0.5 * (t[mid - 1] + t[mid + 1]);
  }
  t.front() = 0.0;
  t.back() = 1.0;
  return t;
}

struct Section {
  std::vector<Vec2> contour;
  Vec2 trailingEdge;
};

// src/geom/Airfoil.cpp:50-166 (contour and trailing-edge point only)
Section naca(int mi, int pi, int ti, int points, double c, bool closed) {
  const double m = static_cast<double>(mi) / 100.0, p = static_cast<double>(pi) / 10.0;
  const double t = static_cast<double>(ti) / 100.0;
  const int n = points;
  std::vector<Vec2> up(static_cast<size_t>(n)), lo(static_cast<size_t>(n)), cam(static_cast<size_t>(n));
  const double last = static_cast<double>(n - 1);
  for (int k = 0; k < n; ++k) {
    double x = 0.5 * (1.0 - std::cos(kPi * static_cast<double>(k) / last));  // Naca4.cpp:148-168
    if (k == 0) x = 0.0;
    if (k == n - 1) x = 1.0;
    const double yc = camberY(x, m, p);
    const double yt = halfThickness(x, t, closed);
    const double th = std::atan(camberDyDx(x, m, p));
    const double s = std::sin(th), co = std::cos(th);
    cam[static_cast<size_t>(k)] = Vec2{x * c, yc * c};
    up[static_cast<size_t>(k)] = Vec2{(x - yt * s) * c, (yc + yt * co) * c};
    lo[static_cast<size_t>(k)] = Vec2{(x + yt * s) * c, (yc - yt * co) * c};
  }
  up.front() = lo.front() = cam.front();
  if (closed) up.back() = lo.back() = cam.back();
  Section sec;
  for (int k = n - 1; k >= 0; --k) sec.contour.push_back(up[static_cast<size_t>(k)]);
  const int lowerEnd = closed ? n - 1 : n;
  for (int k = 1; k < lowerEnd; ++k) sec.contour.push_back(lo[static_cast<size_t>(k)]);
  sec.contour.push_back(sec.contour.front());
  sec.trailingEdge = (up.back() + lo.back()) * 0.5;
  return sec;
}

Vec2 rotated(Vec2 v, double a) {
  const double s = std::sin(a), co = std::cos(a);
  return Vec2{v.x * co - v.y * s, v.x * s + v.y * co};
}

double smooth01(double x) {
  const double t = std::clamp(x, 0.0, 1.0);
  return t * t * (3.0 - 2.0 * t);
}

}  // namespace

double length(Vec2 v) { return std::hypot(v.x, v.y); }

std::vector<double> geometricDistribution(double first, double total, int n) {  // Distribution.cpp:53-76
  if (n < 1) return {};
  std::vector<double> t(static_cast<size_t>(n) + 1, 0.0);
  const double r = growthRatio(first, total, n);
  double step = first, acc = 0.0;
  for (int k = 1; k <= n; ++k) {
    acc += step;
    t[static_cast<size_t>(k)] = acc;
    step *= r;
  }
  if (acc > 0.0)
    for (double& v : t) v /= acc;
  t.front() = 0.0;
  t.back() = 1.0;
  return t;
}

std::vector<Vec2> nacaContour(int m, int p, int t, int points, double chord, bool closed) {
  return naca(m, p, t, points, chord, closed).contour;
}

CGridOptions presetOptions(int resolution) {
  CGridOptions o;
  if (resolution == 0) { o.surfacePoints = 80; o.wakePoints = 32; o.normalPoints = 40; o.firstLayer = 1.0e-3; }
  else if (resolution == 1) { o.surfacePoints = 160; o.wakePoints = 56; o.normalPoints = 72; o.firstLayer = 3.0e-4; }
  else { o.surfacePoints = 320; o.wakePoints = 96; o.normalPoints = 128; o.firstLayer = 1.0e-4; }
  return o;
}

int Grid::faceType(int f) const {
  if (f < jFaces()) {
    const int j = f / nI, i = f % nI;
    if (j == 0) return bottomType[static_cast<size_t>(i)];
    if (j == nJ) return topType;
    return kInterior;
  }
  const int q = f - jFaces();
  const int i = q / nJ;
  if (i == 0) return leftType;
  if (i == nI) return rightType;
  return kInterior;
}

// src/mesh/Mesh.cpp:89-339, structured: cell metrics by shoelace moments, face
// metrics per family, then the quality summary.
bool buildGrid(Grid& g, int nodesI, int nodesJ, std::vector<Vec2> nodes, std::vector<int> bottomType,
               std::vector<int> bottomPartner, int top, int left, int right, std::string& error) {
  if (nodesI < 2 || nodesJ < 2) { error = "a structured block needs at least 2x2 nodes"; return false; }
  if (nodes.size() != static_cast<size_t>(nodesI) * static_cast<size_t>(nodesJ)) { error = "node count does not match the block"; return false; }
  if (bottomType.size() != static_cast<size_t>(nodesI - 1)) { error = "j-min boundary entries do not match"; return false; }
  if (!bottomPartner.empty() && bottomPartner.size() != static_cast<size_t>(nodesI - 1)) { error = "j-min partner entries do not match"; return false; }
  g.nodesI = nodesI;
  g.nodesJ = nodesJ;
  g.nI = nodesI - 1;
  g.nJ = nodesJ - 1;
  g.nodes = std::move(nodes);
  g.bottomType = std::move(bottomType);
  g.bottomPartner.assign(static_cast<size_t>(g.nI), -1);
  for (int i = 0; i < g.nI && !bottomPartner.empty(); ++i) {
    const int pi = bottomPartner[static_cast<size_t>(i)];
    // a partner only couples through a wake cut (Mesh.cpp:78-87)
    if (pi >= 0 && pi < g.nI && g.bottomType[static_cast<size_t>(i)] == kWakeCutB) g.bottomPartner[static_cast<size_t>(i)] = pi;
  }
  g.topType = top;
  g.leftType = left;
  g.rightType = right;
  const int nI = g.nI, nJ = g.nJ, NI = nodesI;
  auto node = [&](int i, int j) { return g.nodes[static_cast<size_t>(j) * static_cast<size_t>(NI) + static_cast<size_t>(i)]; };
  const size_t N = static_cast<size_t>(nI) * static_cast<size_t>(nJ);
  g.area.assign(N, 0.0);
  g.centroid.assign(N, Vec2{});
  for (int j = 0; j < nJ; ++j) {
    for (int i = 0; i < nI; ++i) {
      const Vec2 q[4] = {node(i, j), node(i + 1, j), node(i + 1, j + 1), node(i, j + 1)};
      double twice = 0.0, cx = 0.0, cy = 0.0;
      for (int k = 0; k < 4; ++k) {
        const Vec2 a = q[k], b = q[(k + 1) % 4];
        const double w = cross(a, b);
        twice += w;
        cx += (a.x + b.x) * w;
        cy += (a.y + b.y) * w;
      }
      const size_t c = static_cast<size_t>(j) * static_cast<size_t>(nI) + static_cast<size_t>(i);
      g.area[c] = 0.5 * twice;
      g.centroid[c] = (std::abs(twice) > 0.0) ? Vec2{cx / (3.0 * twice), cy / (3.0 * twice)}
                                              : (q[0] + q[1] + q[2] + q[3]) * 0.25;
    }
  }
  auto faceMetrics = [](Vec2 a, Vec2 b, bool rightTurn, Vec2& centre, Vec2& normal, double& len) {
    const Vec2 tng = b - a;
    len = length(tng);
    centre = (a + b) * 0.5;
    normal = Vec2{0.0, 0.0};
    if (len > 0.0) normal = (rightTurn ? turnRight(tng) : turnLeft(tng)) * (1.0 / len);
  };
  g.jCentre.resize(static_cast<size_t>(g.jFaces()));
  g.jNormal.resize(static_cast<size_t>(g.jFaces()));
  g.jLen.resize(static_cast<size_t>(g.jFaces()));
  for (int j = 0; j <= nJ; ++j)
    for (int i = 0; i < nI; ++i) {
      const size_t f = static_cast<size_t>(j * nI + i);
      faceMetrics(node(i, j), node(i + 1, j), j == 0, g.jCentre[f], g.jNormal[f], g.jLen[f]);
    }
  g.iCentre.resize(static_cast<size_t>(g.iFaces()));
  g.iNormal.resize(static_cast<size_t>(g.iFaces()));
  g.iLen.resize(static_cast<size_t>(g.iFaces()));
  for (int j = 0; j < nJ; ++j)
    for (int i = 0; i <= nI; ++i) {
      const size_t f = static_cast<size_t>(j * (nI + 1) + i);
      faceMetrics(node(i, j), node(i, j + 1), i != 0, g.iCentre[f], g.iNormal[f], g.iLen[f]);
    }
  // quality (Mesh.cpp:268-336), in the reference's cell and face order
  double minA = std::numeric_limits<double>::max(), maxA = std::numeric_limits<double>::lowest(), aspect = 0.0;
  int inverted = 0;
  g.totalArea = 0.0;
  for (int j = 0; j < nJ; ++j)
    for (int i = 0; i < nI; ++i) {
      const double a = g.area[static_cast<size_t>(j * nI + i)];
      g.totalArea += a;
      minA = std::min(minA, a);
      maxA = std::max(maxA, a);
      if (!(a > 0.0)) ++inverted;
      const Vec2 p0 = node(i, j), p1 = node(i + 1, j), p2 = node(i + 1, j + 1), p3 = node(i, j + 1);
      const double alongI = 0.5 * (distance(p0, p1) + distance(p3, p2));
      const double alongJ = 0.5 * (distance(p1, p2) + distance(p0, p3));
      const double small = std::min(alongI, alongJ);
      if (small > 0.0) aspect = std::max(aspect, std::max(alongI, alongJ) / small);
    }
  double nonOrtho = 0.0;
  auto link = [&](Vec2 n, int o, int nb) {
    const Vec2 l = g.centroid[static_cast<size_t>(nb)] - g.centroid[static_cast<size_t>(o)];
    const double ll = length(l);
    if (ll <= 0.0) return;
    const double cosine = std::clamp(dot(n, l) / ll, -1.0, 1.0);
    nonOrtho = std::max(nonOrtho, std::acos(cosine));
  };
  for (int j = 1; j < nJ; ++j)
    for (int i = 0; i < nI; ++i) link(g.jNormal[static_cast<size_t>(j * nI + i)], (j - 1) * nI + i, j * nI + i);
  for (int i = 1; i < nI; ++i)
    for (int j = 0; j < nJ; ++j) link(g.iNormal[static_cast<size_t>(j * (nI + 1) + i)], j * nI + i - 1, j * nI + i);
  double wall = std::numeric_limits<double>::max();
  bool sawWall = false;
  for (int i = 0; i < nI; ++i) {
    if (g.bottomType[static_cast<size_t>(i)] != kWallB) continue;
    sawWall = true;
    wall = std::min(wall, distance(node(i, 0), node(i, 1)));
  }
  g.minArea = N > 0 ? minA : 0.0;
  g.maxArea = N > 0 ? maxA : 0.0;
  g.maxAspect = aspect;
  g.maxNonOrthoDeg = nonOrtho * 180.0 / kPi;
  g.inverted = inverted;
  g.minWallSpacing = sawWall ? wall : 0.0;
  return true;
}

// src/mesh/CGrid.cpp:98-399
bool makeCGrid(Grid& g, int mi, int pi, int ti, double c, const CGridOptions& o, std::string& error) {
  if (o.surfacePoints < 8) { error = "surfacePoints must be at least 8"; return false; }
  if (o.wakePoints < 2) { error = "wakePoints must be at least 2"; return false; }
  if (o.normalPoints < 3) { error = "normalPoints must be at least 3"; return false; }
  if (!(o.upstream > 0.5) || !(o.downstream > 0.5) || !(o.vertical > 0.5)) { error = "domain extents must each be more than half a chord"; return false; }
  if (!(o.firstLayer > 0.0) || o.firstLayer > 0.1) { error = "firstLayerHeight must lie in (0, 0.1] chords"; return false; }
  if (!(c > 0.0) || !std::isfinite(c)) { error = "chord must be a positive finite length"; return false; }
  if ((mi > 0 && pi == 0) || (mi == 0 && pi > 0) || ti <= 0 || ti > 40 || mi < 0 || mi > 9 || pi < 0 || pi > 9) {
    error = "not a valid NACA four-digit designation (Naca4.cpp:75-88)";
    return false;
  }
  const Vec2 te = naca(mi, pi, ti, 121, c, true).trailingEdge;
  std::vector<Vec2> wall = naca(mi, pi, ti, o.surfacePoints, c, true).contour;
  std::reverse(wall.begin(), wall.end());
  const int wallNodes = static_cast<int>(wall.size());
  const int wake = o.wakePoints;
  const int ni = 2 * wake + wallNodes, nj = o.normalPoints;
  const int firstWall = wake, lastWall = wake + wallNodes - 1;
  const double xT = te.x;
  const double wakeLen = o.downstream * c;
  const double teSpacing = std::max(distance(wall[0], wall[1]), 1e-9 * c);
  const std::vector<double> innerFrac = geometricDistribution(teSpacing, wakeLen, wake);
  const std::vector<double> outerFrac = geometricDistribution(wakeLen / (2.0 * static_cast<double>(wake)), wakeLen, wake);
  const double halfH = o.vertical * c;
  const double frontR = xT + o.upstream * c;
  std::vector<Vec2> inner(static_cast<size_t>(ni)), outer(static_cast<size_t>(ni));
  for (int i = 0; i < wake; ++i) {
    const double x = xT + innerFrac[static_cast<size_t>(wake - i)] * wakeLen;
    const double ox = xT + outerFrac[static_cast<size_t>(wake - i)] * wakeLen;
    inner[static_cast<size_t>(i)] = inner[static_cast<size_t>(ni - 1 - i)] = Vec2{x, 0.0};
    outer[static_cast<size_t>(i)] = Vec2{ox, -halfH};
    outer[static_cast<size_t>(ni - 1 - i)] = Vec2{ox, halfH};
  }
  for (int k = 0; k < wallNodes; ++k) {
    inner[static_cast<size_t>(firstWall + k)] = wall[static_cast<size_t>(k)];
    const double phi = kPi * static_cast<double>(k) / static_cast<double>(wallNodes - 1);
    outer[static_cast<size_t>(firstWall + k)] = Vec2{xT - frontR * std::sin(phi), -halfH * std::cos(phi)};
  }
  std::vector<Vec2> nrm(static_cast<size_t>(ni));
  for (int i = 0; i < ni; ++i) {
    const size_t u = static_cast<size_t>(i);
    const Vec2 tng = (i == 0) ? inner[1] - inner[0] : (i == ni - 1) ? inner[u] - inner[u - 1] : inner[u + 1] - inner[u - 1];
    nrm[u] = turnLeft(unit(tng));
  }
  for (int pass = 0; pass < o.smoothingPasses; ++pass) {
    std::vector<Vec2> next = nrm;
    for (int i = 1; i < ni - 1; ++i) {
      const size_t u = static_cast<size_t>(i);
      next[u] = unit(nrm[u - 1] + nrm[u] * 2.0 + nrm[u + 1]);
    }
    nrm.swap(next);
  }
  // curvature-limited blend zone (CGrid.cpp:36-60, 253-284)
  const double limit = o.blendLength * c;
  std::vector<double> blend(static_cast<size_t>(ni), limit);
  for (int i = 1; i + 1 < ni; ++i) {
    const size_t u = static_cast<size_t>(i);
    const Vec2 back = inner[u] - inner[u - 1], fwd = inner[u + 1] - inner[u];
    const double lb = length(back), lf = length(fwd);
    if (lb <= 0.0 || lf <= 0.0) continue;
    const Vec2 tb = back * (1.0 / lb), tf = fwd * (1.0 / lf);
    const double kappa = std::atan2(cross(tb, tf), dot(tb, tf)) / (0.5 * (lb + lf));
    if (kappa > 0.0) blend[u] = std::min(limit, 0.3 / kappa);
  }
  for (int pass = 0; pass < 2; ++pass) {
    for (int i = 1; i < ni; ++i) {
      const size_t u = static_cast<size_t>(i);
      blend[u] = std::min(blend[u], blend[u - 1] + 0.25 * distance(inner[u], inner[u - 1]));
    }
    for (int i = ni - 2; i >= 0; --i) {
      const size_t u = static_cast<size_t>(i);
      blend[u] = std::min(blend[u], blend[u + 1] + 0.25 * distance(inner[u], inner[u + 1]));
    }
  }
  const double first = o.firstLayer * c;
  const double ortho = 5.0 * first;
  std::vector<Vec2> nodes(static_cast<size_t>(ni) * static_cast<size_t>(nj));
  for (int i = 0; i < ni; ++i) {
    const size_t u = static_cast<size_t>(i);
    const Vec2 span = outer[u] - inner[u];
    const double reachMax = length(span);
    const Vec2 straight = unit(span);
    const Vec2 wn = nrm[u];
    const double turn = std::atan2(cross(wn, straight), dot(wn, straight));
    const double need = std::max(1.5, 2.5 * std::abs(turn));
    const double outerB = std::max(blend[u], 1e-9 * c);
    const double innerB = std::min(ortho, outerB * std::exp(-need));
    const double logSpan = std::log(outerB / innerB);
    const std::vector<double> t = geometricDistribution(first, reachMax, nj - 1);
    for (int j = 0; j < nj; ++j) {
      const double reach = t[static_cast<size_t>(j)] * reachMax;
      const double w = (reach <= innerB) ? 0.0 : (reach < outerB) ? smooth01(std::log(reach / innerB) / logSpan) : 1.0;
      nodes[static_cast<size_t>(j) * static_cast<size_t>(ni) + u] = inner[u] + rotated(wn, w * turn) * reach;
    }
  }
  std::vector<int> bottom(static_cast<size_t>(ni - 1), kWallB), partner(static_cast<size_t>(ni - 1), -1);
  for (int k = 0; k < ni - 1; ++k) {
    if (k >= firstWall && k < lastWall) continue;
    bottom[static_cast<size_t>(k)] = kWakeCutB;
    partner[static_cast<size_t>(k)] = ni - 2 - k;
  }
  return buildGrid(g, ni, nj, std::move(nodes), std::move(bottom), std::move(partner), kFarfieldB, kOutletB, kOutletB,
                   error);
}

// src/mesh/BoxGrid.cpp:10-66
bool makeBox(Grid& g, const BoxOptions& o, std::string& error) {
  if (o.cellsX < 1 || o.cellsY < 1) { error = "a box needs at least one cell per direction"; return false; }
  if (!std::isfinite(o.length) || o.length <= 0.0 || !std::isfinite(o.height) || o.height <= 0.0) { error = "box extents must be positive and finite"; return false; }
  if (o.grading != 0 && (!(o.firstCellHeight > 0.0) || o.firstCellHeight >= 0.5)) { error = "firstCellHeight must lie in (0, 0.5)"; return false; }
  const std::vector<double> xs = uniformSpacing(o.cellsX);
  const std::vector<double> ys = (o.grading == 0) ? uniformSpacing(o.cellsY)
                                 : (o.grading == 1) ? geometricDistribution(o.firstCellHeight, 1.0, o.cellsY)
                                                    : mirroredGeometric(o.firstCellHeight, 1.0, o.cellsY);
  const int ni = o.cellsX + 1, nj = o.cellsY + 1;
  std::vector<Vec2> nodes;
  nodes.reserve(static_cast<size_t>(ni) * static_cast<size_t>(nj));
  for (int j = 0; j < nj; ++j) {
    const double y = o.originY + o.height * ys[static_cast<size_t>(j)];
    for (int i = 0; i < ni; ++i) nodes.push_back(Vec2{o.originX + o.length * xs[static_cast<size_t>(i)], y});
  }
  return buildGrid(g, ni, nj, std::move(nodes), std::vector<int>(static_cast<size_t>(ni - 1), o.lower), {}, o.upper,
                   o.left, o.right, error);
}

}  // namespace cfdgpu
