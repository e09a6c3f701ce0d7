// oracle.cpp - FP64 CPU restatement of the reference SIMPLE solver.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP product path, never part of
// it. Only tests/, the smoke check and the benchmark's CPU-baseline leg load
// this library.
//
// Each routine follows one reference routine statement by statement, with the
// same operation order (no FMA contraction: built with -ffp-contract=off), so
// that its results are those of the reference on the same inputs. Citations are
// path:line into sidhanth-sridhar/CFD-SIMULATOR.
//
// How it is pinned (tests/test_oracle_*.py): the reference's own closed-form
// tests restated - NACA ordinates, mesh invariants, Green-Gauss exactness on
// linear fields, the LDU hand product, CG known solution, GS residual drop,
// the uniform / Couette / Poiseuille / Blasius / mass-balance validation cases,
// the SST constants and decaying-turbulence closed form, closed-contour force
// identities - plus the two polars the reference ships under examples/.

#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

struct V2 {
  double x{0.0}, y{0.0};
};
inline V2 operator+(V2 a, V2 b) { return V2{a.x + b.x, a.y + b.y}; }
inline V2 operator-(V2 a, V2 b) { return V2{a.x - b.x, a.y - b.y}; }
inline V2 operator*(V2 a, double s) { return V2{a.x * s, a.y * s}; }
inline V2& operator+=(V2& a, V2 b) { a.x += b.x; a.y += b.y; return a; }
inline V2& operator-=(V2& a, V2 b) { a.x -= b.x; a.y -= b.y; return a; }
inline V2& operator*=(V2& a, double s) { a.x *= s; a.y *= s; return a; }
inline double dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline double cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
inline double len(V2 v) { return std::hypot(v.x, v.y); }  // include/cfd/core/Vec2.hpp:64-66
inline double dist(V2 a, V2 b) { return len(b - a); }

// ----------------------------------------------------------------------------
// Geometry: src/geom/Naca4.cpp:103-168, src/geom/Airfoil.cpp:50-166
// ----------------------------------------------------------------------------

double thickness(double x, double t, bool closed) {  // Naca4.cpp:103-118
  if (!(x > 0.0)) return 0.0;
  const double xc = std::min(x, 1.0);
  const double a4 = closed ? -0.1036 : -0.1015;
  const double x2 = xc * xc;
  const double x3 = x2 * xc;
  const double x4 = x3 * xc;
  return 5.0 * t * (0.2969 * std::sqrt(xc) + -0.1260 * xc + -0.3516 * x2 + 0.2843 * x3 + a4 * x4);
}

double camber(double x, double m, double p) {  // Naca4.cpp:120-133
  if (m == 0.0 || p <= 0.0 || p >= 1.0) return 0.0;
  const double xc = std::clamp(x, 0.0, 1.0);
  if (xc <= p) return (m / (p * p)) * (2.0 * p * xc - xc * xc);
  const double q = 1.0 - p;
  return (m / (q * q)) * ((1.0 - 2.0 * p) + 2.0 * p * xc - xc * xc);
}

double camberSlope(double x, double m, double p) {  // Naca4.cpp:135-146
  if (m == 0.0 || p <= 0.0 || p >= 1.0) return 0.0;
  const double xc = std::clamp(x, 0.0, 1.0);
  const double den = (xc <= p) ? (p * p) : ((1.0 - p) * (1.0 - p));
  return (2.0 * m / den) * (p - xc);
}

std::vector<double> cosineSpacing(int count) {  // Naca4.cpp:148-168
  if (count < 2) return {};
  std::vector<double> s(static_cast<std::size_t>(count));
  const double last = static_cast<double>(count - 1);
  for (int i = 0; i < count; ++i) {
    const double beta = kPi * static_cast<double>(i) / last;
    s[static_cast<std::size_t>(i)] = 0.5 * (1.0 - std::cos(beta));
  }
  s.front() = 0.0;
  s.back() = 1.0;
  return s;
}

struct Section {
  std::vector<V2> upper, lower, camberPts, contour;
  V2 te;
};

// Airfoil.cpp:50-166 (only what the mesh needs: surfaces, contour, TE point).
Section naca(int mi, int pi, int ti, int points, double c, bool closed) {
  const double m = static_cast<double>(mi) / 100.0;
  const double p = static_cast<double>(pi) / 10.0;
  const double t = static_cast<double>(ti) / 100.0;
  Section s;
  const std::vector<double> st = cosineSpacing(points);
  const std::size_t n = st.size();
  for (const double x : st) {
    const double yc = camber(x, m, p);
    const double yt = thickness(x, t, closed);
    const double theta = std::atan(camberSlope(x, m, p));
    const double sn = std::sin(theta);
    const double cs = std::cos(theta);
    s.camberPts.push_back(V2{x * c, yc * c});
    s.upper.push_back(V2{(x - yt * sn) * c, (yc + yt * cs) * c});
    s.lower.push_back(V2{(x + yt * sn) * c, (yc - yt * cs) * c});
  }
  s.upper.front() = s.camberPts.front();
  s.lower.front() = s.camberPts.front();
  if (closed) {
    s.upper.back() = s.camberPts.back();
    s.lower.back() = s.camberPts.back();
  }
  for (std::size_t i = n; i-- > 0;) s.contour.push_back(s.upper[i]);
  const std::size_t lowerEnd = closed ? n - 1 : n;
  for (std::size_t i = 1; i < lowerEnd; ++i) s.contour.push_back(s.lower[i]);
  s.contour.push_back(s.contour.front());
  s.te = (s.upper.back() + s.lower.back()) * 0.5;
  return s;
}

// ----------------------------------------------------------------------------
// Distributions: src/mesh/Distribution.cpp:8-109
// ----------------------------------------------------------------------------

std::vector<double> uniformDist(int intervals) {  // Distribution.cpp:8-19
  if (intervals < 1) return {};
  std::vector<double> t(static_cast<std::size_t>(intervals) + 1, 0.0);
  for (int k = 0; k <= intervals; ++k)
    t[static_cast<std::size_t>(k)] = static_cast<double>(k) / static_cast<double>(intervals);
  t.back() = 1.0;
  return t;
}

double geometricRatio(double first, double total, int intervals) {  // Distribution.cpp:21-51
  if (intervals <= 1 || first <= 0.0 || total <= 0.0) return 1.0;
  const double uniform = first * static_cast<double>(intervals);
  if (uniform >= total) return 1.0;
  const auto sum = [first, intervals](double r) {
    if (std::abs(r - 1.0) < 1e-12) return first * static_cast<double>(intervals);
    return first * (std::pow(r, static_cast<double>(intervals)) - 1.0) / (r - 1.0);
  };
  double lo = 1.0, hi = 3.0;
  if (sum(hi) < total) return hi;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (sum(mid) < total) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi);
}

std::vector<double> geometricDist(double first, double total, int intervals) {  // :53-76
  if (intervals < 1) return {};
  std::vector<double> t(static_cast<std::size_t>(intervals) + 1, 0.0);
  const double ratio = geometricRatio(first, total, intervals);
  double step = first, acc = 0.0;
  for (int k = 1; k <= intervals; ++k) {
    acc += step;
    t[static_cast<std::size_t>(k)] = acc;
    step *= ratio;
  }
  if (acc > 0.0)
    for (double& v : t) v /= acc;
  t.front() = 0.0;
  t.back() = 1.0;
  return t;
}

std::vector<double> symmetricGeometricDist(double first, double total, int intervals) {  // :78-109
  if (intervals < 2) return uniformDist(intervals);
  const int half = intervals / 2;
  const std::vector<double> lower = geometricDist(first, 0.5 * total, half);
  std::vector<double> t(static_cast<std::size_t>(intervals) + 1, 0.0);
  for (int k = 0; k <= half; ++k) t[static_cast<std::size_t>(k)] = 0.5 * lower[static_cast<std::size_t>(k)];
  for (int k = 0; k <= half; ++k)
    t[static_cast<std::size_t>(intervals - k)] = 1.0 - 0.5 * lower[static_cast<std::size_t>(k)];
  if (intervals % 2 != 0) {
    const auto mid = static_cast<std::size_t>(half + 1);
    t[mid] = 0.5 * (t[mid - 1] + t[mid + 1]);
  }
  t.front() = 0.0;
  t.back() = 1.0;
  return t;
}

// ----------------------------------------------------------------------------
// Mesh: src/mesh/Mesh.cpp:31-339
// ----------------------------------------------------------------------------

struct Face {
  int n0{-1}, n1{-1};
  int owner{-1}, neighbour{-1};
  int boundary{ORC_INTERIOR};
  int partner{-1};
};

}  // namespace

struct orc_mesh {
  int ni{0}, nj{0};
  std::vector<V2> nodes;
  std::vector<std::array<int, 4>> cellNodes, cellFaces;
  std::vector<Face> faces;
  std::vector<V2> centroids;
  std::vector<double> areas;
  std::vector<V2> fcentre, fnormal;
  std::vector<double> farea;
  double totalArea{0.0};
  double quality[6]{};
  int cellCount() const { return static_cast<int>(cellNodes.size()); }
  int faceCount() const { return static_cast<int>(faces.size()); }
};

namespace {

using Mesh = orc_mesh;

inline V2 rotL(V2 v) { return V2{-v.y, v.x}; }
inline V2 rotR(V2 v) { return V2{v.y, -v.x}; }

// Mesh.cpp:78-87
int opposite(const Mesh& mesh, int f) {
  const Face& face = mesh.faces[static_cast<std::size_t>(f)];
  if (face.neighbour >= 0) return face.neighbour;
  if (face.boundary == ORC_WAKECUT && face.partner >= 0)
    return mesh.faces[static_cast<std::size_t>(face.partner)].owner;
  return -1;
}

// Mesh.cpp:31-57
void quadMoments(const V2 p[4], double& area, V2& centroid) {
  double twice = 0.0, cx = 0.0, cy = 0.0;
  for (int k = 0; k < 4; ++k) {
    const V2 a = p[k];
    const V2 b = p[(k + 1) % 4];
    const double w = cross(a, b);
    twice += w;
    cx += (a.x + b.x) * w;
    cy += (a.y + b.y) * w;
  }
  area = 0.5 * twice;
  if (std::abs(twice) > 0.0) {
    centroid = V2{cx / (3.0 * twice), cy / (3.0 * twice)};
  } else {
    centroid = (p[0] + p[1] + p[2] + p[3]) * 0.25;
  }
}

// Mesh.cpp:89-339
std::unique_ptr<Mesh> buildStructured(int ni, int nj, std::vector<V2> nodes,
                                      const std::vector<int>& jmin, const std::vector<int>& jminPartner,
                                      int jmax, int imin, int imax) {
  if (ni < 2 || nj < 2) return nullptr;
  if (nodes.size() != static_cast<std::size_t>(ni) * static_cast<std::size_t>(nj)) return nullptr;
  if (jmin.size() != static_cast<std::size_t>(ni - 1)) return nullptr;
  if (!jminPartner.empty() && jminPartner.size() != static_cast<std::size_t>(ni - 1)) return nullptr;
  auto mesh = std::make_unique<Mesh>();
  Mesh& m = *mesh;
  m.ni = ni;
  m.nj = nj;
  m.nodes = std::move(nodes);
  const int ci = ni - 1, cj = nj - 1;
  const std::size_t cells = static_cast<std::size_t>(ci) * static_cast<std::size_t>(cj);
  auto nodeAt = [ni](int i, int j) { return j * ni + i; };
  auto cellAt = [ci](int i, int j) { return j * ci + i; };
  m.cellNodes.resize(cells);
  for (int j = 0; j < cj; ++j)
    for (int i = 0; i < ci; ++i)
      m.cellNodes[static_cast<std::size_t>(cellAt(i, j))] = {nodeAt(i, j), nodeAt(i + 1, j),
                                                             nodeAt(i + 1, j + 1), nodeAt(i, j + 1)};
  const std::size_t jCount = static_cast<std::size_t>(ci) * static_cast<std::size_t>(nj);
  const std::size_t iCount = static_cast<std::size_t>(ni) * static_cast<std::size_t>(cj);
  m.faces.resize(jCount + iCount);
  m.cellFaces.resize(cells);
  auto jFace = [ci](int i, int j) { return j * ci + i; };
  auto iFace = [cj, jCount](int i, int j) { return static_cast<int>(jCount) + i * cj + j; };
  for (int j = 0; j < nj; ++j) {
    for (int i = 0; i < ci; ++i) {
      Face& f = m.faces[static_cast<std::size_t>(jFace(i, j))];
      f.n0 = nodeAt(i, j);
      f.n1 = nodeAt(i + 1, j);
      if (j == 0) {
        f.owner = cellAt(i, 0);
        f.boundary = jmin[static_cast<std::size_t>(i)];
        if (!jminPartner.empty()) {
          const int pi = jminPartner[static_cast<std::size_t>(i)];
          f.partner = (pi >= 0 && pi < ci) ? jFace(pi, 0) : -1;
        }
      } else if (j == nj - 1) {
        f.owner = cellAt(i, cj - 1);
        f.boundary = jmax;
      } else {
        f.owner = cellAt(i, j - 1);
        f.neighbour = cellAt(i, j);
        f.boundary = ORC_INTERIOR;
      }
    }
  }
  for (int i = 0; i < ni; ++i) {
    for (int j = 0; j < cj; ++j) {
      Face& f = m.faces[static_cast<std::size_t>(iFace(i, j))];
      f.n0 = nodeAt(i, j);
      f.n1 = nodeAt(i, j + 1);
      if (i == 0) {
        f.owner = cellAt(0, j);
        f.boundary = imin;
      } else if (i == ni - 1) {
        f.owner = cellAt(ci - 1, j);
        f.boundary = imax;
      } else {
        f.owner = cellAt(i - 1, j);
        f.neighbour = cellAt(i, j);
        f.boundary = ORC_INTERIOR;
      }
    }
  }
  for (int j = 0; j < cj; ++j)
    for (int i = 0; i < ci; ++i)
      m.cellFaces[static_cast<std::size_t>(cellAt(i, j))] = {jFace(i, j), iFace(i + 1, j),
                                                             jFace(i, j + 1), iFace(i, j)};
  m.areas.resize(cells);
  m.centroids.resize(cells);
  for (std::size_t c = 0; c < cells; ++c) {
    const auto& n = m.cellNodes[c];
    const V2 p[4] = {m.nodes[static_cast<std::size_t>(n[0])], m.nodes[static_cast<std::size_t>(n[1])],
                     m.nodes[static_cast<std::size_t>(n[2])], m.nodes[static_cast<std::size_t>(n[3])]};
    quadMoments(p, m.areas[c], m.centroids[c]);
  }
  const std::size_t nf = m.faces.size();
  m.fcentre.resize(nf);
  m.fnormal.resize(nf);
  m.farea.resize(nf);
  for (std::size_t f = 0; f < nf; ++f) {
    const Face& face = m.faces[f];
    const V2 a = m.nodes[static_cast<std::size_t>(face.n0)];
    const V2 b = m.nodes[static_cast<std::size_t>(face.n1)];
    const V2 tangent = b - a;
    const double l = len(tangent);
    m.fcentre[f] = (a + b) * 0.5;
    m.farea[f] = l;
    V2 normal{0.0, 0.0};
    if (l > 0.0) {
      if (f < jCount) {
        const int j = static_cast<int>(f / static_cast<std::size_t>(ci));
        normal = (j == 0) ? rotR(tangent) : rotL(tangent);
      } else {
        const int i = static_cast<int>((f - jCount) / static_cast<std::size_t>(cj));
        normal = (i == 0) ? rotL(tangent) : rotR(tangent);
      }
      normal *= 1.0 / l;
    }
    m.fnormal[f] = normal;
  }
  // Mesh.cpp:268-336 quality
  double minA = std::numeric_limits<double>::max(), maxA = std::numeric_limits<double>::lowest();
  double maxAspect = 0.0;
  int inverted = 0;
  m.totalArea = 0.0;
  for (std::size_t c = 0; c < cells; ++c) {
    const double a = m.areas[c];
    m.totalArea += a;
    minA = std::min(minA, a);
    maxA = std::max(maxA, a);
    if (!(a > 0.0)) ++inverted;
    const auto& n = m.cellNodes[c];
    const V2 p0 = m.nodes[static_cast<std::size_t>(n[0])], p1 = m.nodes[static_cast<std::size_t>(n[1])];
    const V2 p2 = m.nodes[static_cast<std::size_t>(n[2])], p3 = m.nodes[static_cast<std::size_t>(n[3])];
    const double alongI = 0.5 * (dist(p0, p1) + dist(p3, p2));
    const double alongJ = 0.5 * (dist(p1, p2) + dist(p0, p3));
    const double small = std::min(alongI, alongJ);
    if (small > 0.0) maxAspect = std::max(maxAspect, std::max(alongI, alongJ) / small);
  }
  double maxNonOrtho = 0.0;
  for (std::size_t f = 0; f < nf; ++f) {
    const Face& face = m.faces[f];
    if (face.boundary != ORC_INTERIOR) continue;
    const V2 link = m.centroids[static_cast<std::size_t>(face.neighbour)] - m.centroids[static_cast<std::size_t>(face.owner)];
    const double ll = len(link);
    if (ll <= 0.0) continue;
    const double cosine = std::clamp(dot(m.fnormal[f], link) / ll, -1.0, 1.0);
    maxNonOrtho = std::max(maxNonOrtho, std::acos(cosine));
  }
  double minWall = std::numeric_limits<double>::max();
  bool sawWall = false;
  for (int i = 0; i < ci; ++i) {
    if (jmin[static_cast<std::size_t>(i)] != ORC_WALL) continue;
    sawWall = true;
    minWall = std::min(minWall, dist(m.nodes[static_cast<std::size_t>(nodeAt(i, 0))],
                                     m.nodes[static_cast<std::size_t>(nodeAt(i, 1))]));
  }
  m.quality[0] = cells > 0 ? minA : 0.0;
  m.quality[1] = cells > 0 ? maxA : 0.0;
  m.quality[2] = maxAspect;
  m.quality[3] = maxNonOrtho * 180.0 / kPi;
  m.quality[4] = static_cast<double>(inverted);
  m.quality[5] = sawWall ? minWall : 0.0;
  return mesh;
}

// ----------------------------------------------------------------------------
// C-grid: src/mesh/CGrid.cpp:17-399
// ----------------------------------------------------------------------------

V2 normalized(V2 v) {
  const double l = len(v);
  return (l > 0.0) ? v * (1.0 / l) : V2{0.0, 0.0};
}

V2 rotate(V2 v, double angle) {
  const double s = std::sin(angle);
  const double co = std::cos(angle);
  return V2{v.x * co - v.y * s, v.x * s + v.y * co};
}

double smoothStep(double x) {
  const double t = std::clamp(x, 0.0, 1.0);
  return t * t * (3.0 - 2.0 * t);
}

std::vector<double> signedCurvature(const std::vector<V2>& curve) {  // CGrid.cpp:44-60
  std::vector<double> kappa(curve.size(), 0.0);
  for (std::size_t i = 1; i + 1 < curve.size(); ++i) {
    const V2 back = curve[i] - curve[i - 1];
    const V2 fwd = curve[i + 1] - curve[i];
    const double lb = len(back), lf = len(fwd);
    if (lb <= 0.0 || lf <= 0.0) continue;
    const V2 tb = back * (1.0 / lb);
    const V2 tf = fwd * (1.0 / lf);
    const double turn = std::atan2(cross(tb, tf), dot(tb, tf));
    kappa[i] = turn / (0.5 * (lb + lf));
  }
  return kappa;
}

std::unique_ptr<Mesh> cgrid(int mi, int pi, int ti, double c, const orc_cgrid_options& o) {
  if (o.surface_points < 8 || o.wake_points < 2 || o.normal_points < 3) return nullptr;
  if (!(o.upstream > 0.5) || !(o.downstream > 0.5) || !(o.vertical > 0.5)) return nullptr;
  if (!(o.first_layer > 0.0) || o.first_layer > 0.1) return nullptr;
  // The airfoil the caller passes is only used for its designation, chord and
  // trailing edge point (CGrid.cpp:129-157); with a closed trailing edge that
  // point is the camber line's end, independent of the caller's resolution.
  const Section passed = naca(mi, pi, ti, 121, c, true);
  const Section wallSection = naca(mi, pi, ti, o.surface_points, c, true);
  std::vector<V2> wall = wallSection.contour;
  std::reverse(wall.begin(), wall.end());
  const int wallNodes = static_cast<int>(wall.size());
  const int wakeNodes = o.wake_points;
  const int ni = 2 * wakeNodes + wallNodes;
  const int nj = o.normal_points;
  const int iTL = wakeNodes;
  const int iTU = wakeNodes + wallNodes - 1;
  const double xT = passed.te.x;
  const double wakeLength = o.downstream * c;
  const double trailingSpacing = std::max(dist(wall[0], wall[1]), 1e-9 * c);
  const std::vector<double> wakeFraction = geometricDist(trailingSpacing, wakeLength, wakeNodes);
  const double outerSpacing = wakeLength / (2.0 * static_cast<double>(wakeNodes));
  const std::vector<double> outerWake = geometricDist(outerSpacing, wakeLength, wakeNodes);
  const double halfHeight = o.vertical * c;
  const double frontRadius = xT + o.upstream * c;
  std::vector<V2> inner(static_cast<std::size_t>(ni)), outer(static_cast<std::size_t>(ni));
  for (int i = 0; i < wakeNodes; ++i) {
    const double s = wakeFraction[static_cast<std::size_t>(wakeNodes - i)] * wakeLength;
    const double x = xT + s;
    const double os = outerWake[static_cast<std::size_t>(wakeNodes - i)] * wakeLength;
    const double ox = xT + os;
    inner[static_cast<std::size_t>(i)] = V2{x, 0.0};
    outer[static_cast<std::size_t>(i)] = V2{ox, -halfHeight};
    const int mirror = ni - 1 - i;
    inner[static_cast<std::size_t>(mirror)] = V2{x, 0.0};
    outer[static_cast<std::size_t>(mirror)] = V2{ox, halfHeight};
  }
  for (int k = 0; k < wallNodes; ++k) {
    const int i = iTL + k;
    inner[static_cast<std::size_t>(i)] = wall[static_cast<std::size_t>(k)];
    const double phi = kPi * static_cast<double>(k) / static_cast<double>(wallNodes - 1);
    outer[static_cast<std::size_t>(i)] = V2{xT - frontRadius * std::sin(phi), -halfHeight * std::cos(phi)};
  }
  std::vector<V2> normals(static_cast<std::size_t>(ni));
  for (int i = 0; i < ni; ++i) {
    const auto u = static_cast<std::size_t>(i);
    V2 tangent;
    if (i == 0) tangent = inner[1] - inner[0];
    else if (i == ni - 1) tangent = inner[u] - inner[u - 1];
    else tangent = inner[u + 1] - inner[u - 1];
    normals[u] = rotL(normalized(tangent));
  }
  for (int pass = 0; pass < o.normal_smoothing_passes; ++pass) {
    std::vector<V2> sm = normals;
    for (int i = 1; i < ni - 1; ++i) {
      const auto u = static_cast<std::size_t>(i);
      sm[u] = normalized(normals[u - 1] + normals[u] * 2.0 + normals[u + 1]);
    }
    normals = std::move(sm);
  }
  const std::vector<double> kappa = signedCurvature(inner);
  const double blendLimit = o.normal_blend_length * c;
  std::vector<double> blend(static_cast<std::size_t>(ni), blendLimit);
  for (int i = 0; i < ni; ++i) {
    const auto u = static_cast<std::size_t>(i);
    if (kappa[u] > 0.0) blend[u] = std::min(blendLimit, 0.3 / kappa[u]);
  }
  constexpr double kGrowth = 0.25;
  for (int pass = 0; pass < 2; ++pass) {
    for (int i = 1; i < ni; ++i) {
      const auto u = static_cast<std::size_t>(i);
      blend[u] = std::min(blend[u], blend[u - 1] + kGrowth * dist(inner[u], inner[u - 1]));
    }
    for (int i = ni - 2; i >= 0; --i) {
      const auto u = static_cast<std::size_t>(i);
      blend[u] = std::min(blend[u], blend[u + 1] + kGrowth * dist(inner[u], inner[u + 1]));
    }
  }
  const double firstLayer = o.first_layer * c;
  const double orthoZone = 5.0 * firstLayer;
  std::vector<V2> nodes(static_cast<std::size_t>(ni) * static_cast<std::size_t>(nj));
  for (int i = 0; i < ni; ++i) {
    const auto u = static_cast<std::size_t>(i);
    const V2 start = inner[u];
    const V2 span = outer[u] - start;
    const double distanceOut = len(span);
    const V2 straight = normalized(span);
    const V2 wn = normals[u];
    const double turn = std::atan2(cross(wn, straight), dot(wn, straight));
    const double requiredSpan = std::max(1.5, 2.5 * std::abs(turn));
    const double outerBlend = std::max(blend[u], 1e-9 * c);
    const double innerBlend = std::min(orthoZone, outerBlend * std::exp(-requiredSpan));
    const double logSpan = std::log(outerBlend / innerBlend);
    const std::vector<double> t = geometricDist(firstLayer, distanceOut, nj - 1);
    for (int j = 0; j < nj; ++j) {
      const double reach = t[static_cast<std::size_t>(j)] * distanceOut;
      double w = 1.0;
      if (reach <= innerBlend) w = 0.0;
      else if (reach < outerBlend) w = smoothStep(std::log(reach / innerBlend) / logSpan);
      const V2 dir = rotate(wn, w * turn);
      nodes[static_cast<std::size_t>(j) * static_cast<std::size_t>(ni) + u] = start + dir * reach;
    }
  }
  std::vector<int> jmin(static_cast<std::size_t>(ni - 1), ORC_WALL), partner(static_cast<std::size_t>(ni - 1), -1);
  for (int k = 0; k < ni - 1; ++k) {
    const bool onWall = (k >= iTL && k < iTU);
    if (!onWall) {
      jmin[static_cast<std::size_t>(k)] = ORC_WAKECUT;
      partner[static_cast<std::size_t>(k)] = ni - 2 - k;
    }
  }
  return buildStructured(ni, nj, std::move(nodes), jmin, partner, ORC_FARFIELD, ORC_OUTLET, ORC_OUTLET);
}

// src/mesh/BoxGrid.cpp:10-66
std::unique_ptr<Mesh> box(const orc_box_options& o) {
  if (o.cells_x < 1 || o.cells_y < 1) return nullptr;
  if (!std::isfinite(o.length) || o.length <= 0.0 || !std::isfinite(o.height) || o.height <= 0.0) return nullptr;
  if (o.grading != 0 && (!(o.first_cell_height > 0.0) || o.first_cell_height >= 0.5)) return nullptr;
  const std::vector<double> xf = uniformDist(o.cells_x);
  std::vector<double> yf;
  if (o.grading == 0) yf = uniformDist(o.cells_y);
  else if (o.grading == 1) yf = geometricDist(o.first_cell_height, 1.0, o.cells_y);
  else yf = symmetricGeometricDist(o.first_cell_height, 1.0, o.cells_y);
  const int ni = o.cells_x + 1, nj = o.cells_y + 1;
  std::vector<V2> nodes;
  nodes.reserve(static_cast<std::size_t>(ni) * static_cast<std::size_t>(nj));
  for (int j = 0; j < nj; ++j) {
    const double y = o.origin_y + o.height * yf[static_cast<std::size_t>(j)];
    for (int i = 0; i < ni; ++i) {
      const double x = o.origin_x + o.length * xf[static_cast<std::size_t>(i)];
      nodes.push_back(V2{x, y});
    }
  }
  std::vector<int> jmin(static_cast<std::size_t>(ni - 1), o.lower);
  return buildStructured(ni, nj, std::move(nodes), jmin, {}, o.upper, o.left, o.right);
}

// ----------------------------------------------------------------------------
// Linear algebra: src/solver/LinearSystem.cpp:8-210, Gradient.cpp:7-50
// ----------------------------------------------------------------------------

struct Ldu {
  std::vector<double> diag, source, upper, lower;
  void resize(std::size_t cells, std::size_t faces) {
    diag.assign(cells, 0.0);
    source.assign(cells, 0.0);
    upper.assign(faces, 0.0);
    lower.assign(faces, 0.0);
  }
  void clear() {
    std::fill(diag.begin(), diag.end(), 0.0);
    std::fill(source.begin(), source.end(), 0.0);
    std::fill(upper.begin(), upper.end(), 0.0);
    std::fill(lower.begin(), lower.end(), 0.0);
  }
};

// LinearSystem.cpp:18-43
void lduMultiply(const Mesh& mesh, const double* diag, const double* upper, const double* lower,
                 const double* x, double* y) {
  const int n = mesh.cellCount();
  for (int c = 0; c < n; ++c) y[c] = diag[c] * x[c];
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const Face& face = mesh.faces[static_cast<std::size_t>(f)];
    if (face.neighbour >= 0) {
      y[face.owner] += upper[f] * x[face.neighbour];
      y[face.neighbour] += lower[f] * x[face.owner];
      continue;
    }
    const int across = opposite(mesh, f);
    if (across >= 0) y[face.owner] += upper[f] * x[across];
  }
}

// Order of the global sums below (L1 norms, CG inner products, means). 0, the
// default, is the reference's: one sequential pass from cell 0. A positive
// value B sums consecutive blocks of B cells sequentially and then the block
// totals in order; -1 sums pairwise (a balanced tree). Both are the same
// terms in another association, used only to measure how far a solve's
// answer moves with the rounding of its global sums (a device sums in blocks
// and trees).
int gSumBlock = 0;

// Evaluation of the SST blending tanh: 0, the default, is std::tanh as in the
// reference; 1 evaluates (1 - e^-2x) / (1 + e^-2x), which differs from it by
// a few units in the last place. Like gSumBlock, only for measuring how far
// a solve's answer moves with the rounding of a device's own tanh.
int gTanhVariant = 0;

double blendTanh(double x) {
  if (gTanhVariant == 0) return std::tanh(x);
  const double e = std::exp(-2.0 * x);
  return (1.0 - e) / (1.0 + e);
}

template <class Term>
double orderedSum(int n, Term term) {
  if (gSumBlock == 0) {
    double total = 0.0;
    for (int c = 0; c < n; ++c) total += term(c);
    return total;
  }
  if (gSumBlock < 0) {
    // pairwise: adjacent terms, then adjacent pair sums, ... (a balanced tree)
    std::vector<double> t(static_cast<std::size_t>(n));
    for (int c = 0; c < n; ++c) t[static_cast<std::size_t>(c)] = term(c);
    std::size_t m = t.size();
    while (m > 1) {
      const std::size_t half = m / 2;
      for (std::size_t q = 0; q < half; ++q) t[q] = t[2 * q] + t[2 * q + 1];
      if (m & 1) t[half] = t[m - 1];
      m = half + (m & 1);
    }
    return n > 0 ? t[0] : 0.0;
  }
  double total = 0.0;
  for (int b = 0; b < n; b += gSumBlock) {
    const int e = std::min(n, b + gSumBlock);
    double part = 0.0;
    for (int c = b; c < e; ++c) part += term(c);
    total += part;
  }
  return total;
}

double residualL1(const Mesh& mesh, const Ldu& s, const std::vector<double>& x) {  // :45-54
  std::vector<double> ax(s.diag.size());
  lduMultiply(mesh, s.diag.data(), s.upper.data(), s.lower.data(), x.data(), ax.data());
  return orderedSum(static_cast<int>(s.diag.size()), [&](int c) {
    return std::abs(s.source[static_cast<std::size_t>(c)] - ax[static_cast<std::size_t>(c)]);
  });
}

double diagonalL1(const Ldu& s) {  // :56-62
  return orderedSum(static_cast<int>(s.diag.size()),
                    [&](int c) { return std::abs(s.diag[static_cast<std::size_t>(c)]); });
}

// LinearSystem.cpp:77-116
void gaussSeidel(const Mesh& mesh, const double* diag, const double* upper, const double* lower,
                 const double* source, double* x, int sweeps) {
  const int n = mesh.cellCount();
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    for (int c = 0; c < n; ++c) {
      const double d = diag[c];
      if (!(std::abs(d) > 0.0)) continue;
      double off = 0.0;
      for (const int f : mesh.cellFaces[static_cast<std::size_t>(c)]) {
        const Face& face = mesh.faces[static_cast<std::size_t>(f)];
        if (face.neighbour >= 0) {
          if (face.owner == c) off += upper[f] * x[face.neighbour];
          else off += lower[f] * x[face.owner];
          continue;
        }
        const int across = opposite(mesh, f);
        if (across >= 0 && face.owner == c) off += upper[f] * x[across];
      }
      x[c] = (source[c] - off) / d;
    }
  }
}

struct CgReport {
  int iterations{0};
  double initial{0.0}, final{0.0};
  bool converged{false};
};

// LinearSystem.cpp:118-210
CgReport conjugateGradient(const Mesh& mesh, const double* diag, const double* upper,
                           const double* lower, const double* source, double* x, double tol,
                           int maxIt) {
  CgReport rep;
  const int n = mesh.cellCount();
  if (n == 0) return rep;
  std::vector<double> r(static_cast<std::size_t>(n)), d(static_cast<std::size_t>(n)),
      z(static_cast<std::size_t>(n)), ad(static_cast<std::size_t>(n));
  lduMultiply(mesh, diag, upper, lower, x, ad.data());
  for (int c = 0; c < n; ++c) r[static_cast<std::size_t>(c)] = source[c] - ad[static_cast<std::size_t>(c)];
  auto norm = [&]() { return orderedSum(n, [&](int c) { return std::abs(r[static_cast<std::size_t>(c)]); }); };
  auto rDotZ = [&]() {
    return orderedSum(n, [&](int c) { return r[static_cast<std::size_t>(c)] * z[static_cast<std::size_t>(c)]; });
  };
  rep.initial = norm();
  rep.final = rep.initial;
  if (rep.initial <= 0.0) {
    rep.converged = true;
    return rep;
  }
  auto precondition = [&]() {
    for (int c = 0; c < n; ++c) {
      const double dd = diag[c];
      z[static_cast<std::size_t>(c)] = (std::abs(dd) > 0.0) ? r[static_cast<std::size_t>(c)] / dd : r[static_cast<std::size_t>(c)];
    }
  };
  precondition();
  d = z;
  double rz = rDotZ();
  for (int it = 0; it < maxIt; ++it) {
    lduMultiply(mesh, diag, upper, lower, d.data(), ad.data());
    const double dAd =
        orderedSum(n, [&](int c) { return d[static_cast<std::size_t>(c)] * ad[static_cast<std::size_t>(c)]; });
    if (!(std::abs(dAd) > 0.0)) break;
    const double step = rz / dAd;
    for (int c = 0; c < n; ++c) {
      x[c] += step * d[static_cast<std::size_t>(c)];
      r[static_cast<std::size_t>(c)] -= step * ad[static_cast<std::size_t>(c)];
    }
    rep.iterations = it + 1;
    rep.final = norm();
    if (rep.final <= tol * rep.initial) {
      rep.converged = true;
      break;
    }
    precondition();
    const double rzNext = rDotZ();
    if (!(std::abs(rz) > 0.0)) break;
    const double beta = rzNext / rz;
    rz = rzNext;
    for (int c = 0; c < n; ++c) d[static_cast<std::size_t>(c)] = z[static_cast<std::size_t>(c)] + beta * d[static_cast<std::size_t>(c)];
  }
  return rep;
}

// Gradient.cpp:7-18
double ownerWeight(const Mesh& mesh, int f) {
  const Face& face = mesh.faces[static_cast<std::size_t>(f)];
  if (face.neighbour < 0) return 1.0;
  const V2 centre = mesh.fcentre[static_cast<std::size_t>(f)];
  const double toO = dist(centre, mesh.centroids[static_cast<std::size_t>(face.owner)]);
  const double toN = dist(centre, mesh.centroids[static_cast<std::size_t>(face.neighbour)]);
  const double total = toO + toN;
  return (total > 0.0) ? toN / total : 0.5;
}

// Gradient.cpp:20-50
void greenGauss(const Mesh& mesh, const double* faceValues, std::vector<V2>& grad) {
  grad.assign(static_cast<std::size_t>(mesh.cellCount()), V2{});
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const Face& face = mesh.faces[static_cast<std::size_t>(f)];
    const V2 area = mesh.fnormal[static_cast<std::size_t>(f)] * mesh.farea[static_cast<std::size_t>(f)];
    const V2 contribution = area * faceValues[f];
    grad[static_cast<std::size_t>(face.owner)] += contribution;
    if (face.neighbour >= 0) grad[static_cast<std::size_t>(face.neighbour)] -= contribution;
  }
  for (int c = 0; c < mesh.cellCount(); ++c) {
    const double vol = mesh.areas[static_cast<std::size_t>(c)];
    if (vol > 0.0) grad[static_cast<std::size_t>(c)] *= 1.0 / vol;
  }
}

// src/solver/TurbulenceModel.cpp:28-65
std::vector<double> wallDistances(const Mesh& mesh, const std::vector<int>& kind) {
  std::vector<double> d(static_cast<std::size_t>(mesh.cellCount()), 0.0);
  std::vector<V2> walls;
  for (int f = 0; f < mesh.faceCount() && f < static_cast<int>(kind.size()); ++f)
    if (kind[static_cast<std::size_t>(f)] == ORC_K_WALL) walls.push_back(mesh.fcentre[static_cast<std::size_t>(f)]);
  if (walls.empty()) {
    const double far = 1.0e6 * std::sqrt(std::abs(mesh.totalArea) + 1.0);
    std::fill(d.begin(), d.end(), far);
    return d;
  }
  for (int c = 0; c < mesh.cellCount(); ++c) {
    const V2 cc = mesh.centroids[static_cast<std::size_t>(c)];
    double best = std::numeric_limits<double>::max();
    for (const V2& w : walls) {
      const double dx = cc.x - w.x;
      const double dy = cc.y - w.y;
      best = std::min(best, dx * dx + dy * dy);
    }
    d[static_cast<std::size_t>(c)] = std::sqrt(best);
  }
  return d;
}

// ----------------------------------------------------------------------------
// k-omega SST: src/solver/KOmegaSST.cpp
// ----------------------------------------------------------------------------

constexpr double kMinOmega = 1.0e-10;  // KOmegaSST.cpp:20
constexpr double kMinEnergy = 0.0;     // :21
constexpr double kTinySst = 1.0e-30;   // :22

struct SstState {
  orc_sst_constants c{};
  std::vector<double> k, omega, mut, wallDistance, strain, f1, f2, cd, cdPlus, production;
  std::vector<V2> gradK, gradOmega;
  double kIn{0.0}, omegaIn{0.0}, maxYPlus{0.0};
  double resFirst{0.0}, resSecond{0.0};
  double ranges[6]{};
  double gamma1() const { return c.beta1 / c.beta_star - c.sigma_w1 * c.kappa * c.kappa / std::sqrt(c.beta_star); }
  double gamma2() const { return c.beta2 / c.beta_star - c.sigma_w2 * c.kappa * c.kappa / std::sqrt(c.beta_star); }
};

// KOmegaSST.cpp:86-92
double wallOmega(double nu, double distance, double energy, const orc_sst_constants& c) {
  const double d = std::max(distance, kTinySst);
  const double viscous = 6.0 * nu / (c.beta1 * d * d);
  const double logarithmic = std::sqrt(std::max(energy, 0.0)) / (std::pow(c.beta_star, 0.25) * c.kappa * d);
  return std::sqrt(viscous * viscous + logarithmic * logarithmic);
}

// KOmegaSST.cpp:26-68
void sstInterpolate(const Mesh& mesh, const std::vector<int>& kind, const std::vector<double>& cell,
                    double wallValue, double inflowValue, const std::vector<double>& flux,
                    std::vector<double>& out) {
  out.assign(static_cast<std::size_t>(mesh.faceCount()), 0.0);
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const Face& face = mesh.faces[static_cast<std::size_t>(f)];
    const auto o = static_cast<std::size_t>(face.owner);
    if (face.neighbour >= 0) {
      const double w = ownerWeight(mesh, f);
      out[static_cast<std::size_t>(f)] = w * cell[o] + (1.0 - w) * cell[static_cast<std::size_t>(face.neighbour)];
      continue;
    }
    const int across = opposite(mesh, f);
    if (across >= 0) {
      const double w = ownerWeight(mesh, f);
      out[static_cast<std::size_t>(f)] = w * cell[o] + (1.0 - w) * cell[static_cast<std::size_t>(across)];
      continue;
    }
    switch (kind[static_cast<std::size_t>(f)]) {
      case ORC_K_WALL: out[static_cast<std::size_t>(f)] = wallValue; break;
      case ORC_K_INLET: out[static_cast<std::size_t>(f)] = inflowValue; break;
      case ORC_K_FARFIELD:
        out[static_cast<std::size_t>(f)] = (flux[static_cast<std::size_t>(f)] < 0.0) ? inflowValue : cell[o];
        break;
      default: out[static_cast<std::size_t>(f)] = cell[o]; break;
    }
  }
}

}  // namespace

// ----------------------------------------------------------------------------
// SIMPLE: src/solver/SimpleSolver.cpp
// ----------------------------------------------------------------------------

struct orc_solver {
  const Mesh* mesh{nullptr};
  std::vector<int> kind;
  std::vector<V2> bcVelocity;
  std::vector<double> bcPressure;
  orc_settings settings{};
  // field
  std::vector<V2> vel;
  std::vector<double> p, rho, mu, muMol;
  std::vector<double> flux;
  // geometry (SimpleSolver.hpp FaceGeometry)
  std::vector<V2> gArea, gDelta, gTang;
  std::vector<double> gDiff, gWeight;
  std::vector<double> faceU, faceV, faceP;
  std::vector<V2> gradU, gradV, gradP;
  Ldu mom, pres;
  std::vector<double> momDiag, netFlux, deadBand, lastCorrection;
  bool hasPressureRef{false};
  double refMassFlow{1.0};
  int model{0};  // 0 none, 1 laminar, 2 SST
  SstState sst;
  std::vector<double> eddyZero;
};

namespace {

using Solver = orc_solver;
constexpr double kTiny = 1e-300;  // SimpleSolver.cpp:16

// SimpleSolver.cpp:82-131
void precomputeGeometry(Solver& s) {
  const Mesh& mesh = *s.mesh;
  const std::size_t nf = static_cast<std::size_t>(mesh.faceCount());
  s.gArea.assign(nf, V2{});
  s.gDelta.assign(nf, V2{});
  s.gTang.assign(nf, V2{});
  s.gDiff.assign(nf, 0.0);
  s.gWeight.assign(nf, 1.0);
  for (std::size_t f = 0; f < nf; ++f) {
    const Face& face = mesh.faces[f];
    s.gArea[f] = mesh.fnormal[f] * mesh.farea[f];
    const V2 oc = mesh.centroids[static_cast<std::size_t>(face.owner)];
    const int across = opposite(mesh, static_cast<int>(f));
    if (face.neighbour >= 0) {
      s.gDelta[f] = mesh.centroids[static_cast<std::size_t>(face.neighbour)] - oc;
      s.gWeight[f] = ownerWeight(mesh, static_cast<int>(f));
    } else if (across >= 0) {
      const V2 ac = mesh.centroids[static_cast<std::size_t>(across)];
      s.gDelta[f] = ac - oc;
      const double toO = dist(mesh.fcentre[f], oc);
      const double toA = dist(mesh.fcentre[f], ac);
      const double span = toO + toA;
      s.gWeight[f] = (span > 0.0) ? toA / span : 0.5;
    } else {
      s.gDelta[f] = mesh.fcentre[f] - oc;
      s.gWeight[f] = 1.0;
    }
    const double ad = dot(s.gArea[f], s.gDelta[f]);
    const double aa = dot(s.gArea[f], s.gArea[f]);
    if (std::abs(ad) > kTiny) {
      s.gDiff[f] = aa / ad;
      s.gTang[f] = s.gArea[f] - s.gDelta[f] * s.gDiff[f];
    } else {
      s.gDiff[f] = 0.0;
      s.gTang[f] = V2{};
    }
  }
}

inline double comp(V2 v, int i) { return i == 0 ? v.x : v.y; }

// SimpleSolver.cpp:247-262
bool imposesVelocity(const Solver& s, int f) {
  switch (s.kind[static_cast<std::size_t>(f)]) {
    case ORC_K_WALL:
    case ORC_K_INLET: return true;
    case ORC_K_FARFIELD: return s.flux[static_cast<std::size_t>(f)] < s.deadBand[static_cast<std::size_t>(f)];
    default: return false;
  }
}

// the cell across a face: neighbour, or the wake partner's owner (SimpleSolver.cpp:277-280)
inline int otherCell(const Mesh& mesh, const Face& face) {
  int other = face.neighbour;
  if (other < 0 && face.boundary == ORC_WAKECUT && face.partner >= 0)
    other = mesh.faces[static_cast<std::size_t>(face.partner)].owner;
  return other;
}

// SimpleSolver.cpp:268-304
void interpolateFaceValues(Solver& s) {
  const Mesh& mesh = *s.mesh;
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const auto fu = static_cast<std::size_t>(f);
    const Face& face = mesh.faces[fu];
    const auto o = static_cast<std::size_t>(face.owner);
    const int other = otherCell(mesh, face);
    if (other >= 0) {
      const auto n = static_cast<std::size_t>(other);
      const double w = s.gWeight[fu];
      s.faceU[fu] = w * s.vel[o].x + (1.0 - w) * s.vel[n].x;
      s.faceV[fu] = w * s.vel[o].y + (1.0 - w) * s.vel[n].y;
      s.faceP[fu] = w * s.p[o] + (1.0 - w) * s.p[n];
      continue;
    }
    if (imposesVelocity(s, f)) {
      s.faceU[fu] = s.bcVelocity[fu].x;
      s.faceV[fu] = s.bcVelocity[fu].y;
      s.faceP[fu] = s.p[o];
    } else {
      s.faceU[fu] = s.vel[o].x;
      s.faceV[fu] = s.vel[o].y;
      s.faceP[fu] = s.bcPressure[fu];
    }
  }
}

// SimpleSolver.cpp:306-317
void computeGradients(Solver& s) {
  greenGauss(*s.mesh, s.faceU.data(), s.gradU);
  greenGauss(*s.mesh, s.faceV.data(), s.gradV);
  greenGauss(*s.mesh, s.faceP.data(), s.gradP);
}

// SimpleSolver.cpp:319-423
void assembleMomentum(Solver& s, int index) {
  const Mesh& mesh = *s.mesh;
  s.mom.clear();
  const std::vector<V2>& gradient = (index == 0) ? s.gradU : s.gradV;
  const double blend = s.settings.non_orthogonal_blend;
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const auto fu = static_cast<std::size_t>(f);
    const Face& face = mesh.faces[fu];
    const auto o = static_cast<std::size_t>(face.owner);
    const int other = otherCell(mesh, face);
    const double F = s.flux[fu];
    if (other >= 0 && face.neighbour >= 0) {
      const auto n = static_cast<std::size_t>(other);
      const double w = s.gWeight[fu];
      const double visc = w * s.mu[o] + (1.0 - w) * s.mu[n];
      const double diff = visc * s.gDiff[fu];
      const double outward = std::max(F, 0.0);
      const double inward = std::max(-F, 0.0);
      s.mom.diag[o] += outward + diff;
      s.mom.upper[fu] += -inward - diff;
      s.mom.diag[n] += inward + diff;
      s.mom.lower[fu] += -outward - diff;
      const V2 fg = gradient[o] * w + gradient[n] * (1.0 - w);
      const double corr = blend * visc * dot(fg, s.gTang[fu]);
      s.mom.source[o] += corr;
      s.mom.source[n] -= corr;
      if (s.settings.scheme == 1) {
        const double up = (F >= 0.0) ? comp(s.vel[o], index) : comp(s.vel[n], index);
        const double central = w * comp(s.vel[o], index) + (1.0 - w) * comp(s.vel[n], index);
        const double deferred = F * (central - up);
        s.mom.source[o] -= deferred;
        s.mom.source[n] += deferred;
      }
      continue;
    }
    if (other >= 0) {
      const auto n = static_cast<std::size_t>(other);
      const double w = s.gWeight[fu];
      const double visc = w * s.mu[o] + (1.0 - w) * s.mu[n];
      const double diff = visc * s.gDiff[fu];
      const double outward = std::max(F, 0.0);
      const double inward = std::max(-F, 0.0);
      s.mom.diag[o] += outward + diff;
      s.mom.upper[fu] += -inward - diff;
      const V2 fg = gradient[o] * w + gradient[n] * (1.0 - w);
      s.mom.source[o] += blend * visc * dot(fg, s.gTang[fu]);
      continue;
    }
    const double visc = s.mu[o];
    const double diff = visc * s.gDiff[fu];
    if (imposesVelocity(s, f)) {
      const double imposed = comp(s.bcVelocity[fu], index);
      s.mom.source[o] -= F * imposed;
      s.mom.diag[o] += diff;
      s.mom.source[o] += diff * imposed;
      s.mom.source[o] += blend * visc * dot(gradient[o], s.gTang[fu]);
    } else {
      s.mom.diag[o] += F;
    }
  }
  for (int c = 0; c < mesh.cellCount(); ++c) {
    const auto cu = static_cast<std::size_t>(c);
    s.mom.source[cu] -= comp(s.gradP[cu], index) * mesh.areas[cu];
  }
}

// SimpleSolver.cpp:425-443
void applyRelaxation(Solver& s, int index) {
  const double alpha = s.settings.velocity_relaxation;
  for (std::size_t c = 0; c < s.mom.diag.size(); ++c) {
    const double un = s.mom.diag[c];
    if (!(std::abs(un) > 0.0)) {
      s.mom.diag[c] = 1.0;
      continue;
    }
    const double relaxed = un / alpha;
    s.mom.diag[c] = relaxed;
    s.mom.source[c] += (relaxed - un) * comp(s.vel[c], index);
  }
}

// SimpleSolver.cpp:445-496
void rhieChow(Solver& s) {
  const Mesh& mesh = *s.mesh;
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const auto fu = static_cast<std::size_t>(f);
    const Face& face = mesh.faces[fu];
    const auto o = static_cast<std::size_t>(face.owner);
    const int other = otherCell(mesh, face);
    if (other >= 0) {
      const auto n = static_cast<std::size_t>(other);
      const double w = s.gWeight[fu];
      const V2 interp = s.vel[o] * w + s.vel[n] * (1.0 - w);
      const double dens = w * s.rho[o] + (1.0 - w) * s.rho[n];
      const double mob = w * (mesh.areas[o] / s.momDiag[o]) + (1.0 - w) * (mesh.areas[n] / s.momDiag[n]);
      const V2 smoothG = s.gradP[o] * w + s.gradP[n] * (1.0 - w);
      const double compact = (s.p[n] - s.p[o]) * s.gDiff[fu];
      const double smooth = dot(smoothG, s.gArea[fu]);
      s.flux[fu] = dens * dot(interp, s.gArea[fu]) - dens * mob * (compact - smooth);
      continue;
    }
    const double dens = s.rho[o];
    if (imposesVelocity(s, f)) {
      s.flux[fu] = dens * dot(s.bcVelocity[fu], s.gArea[fu]);
    } else {
      const double mob = mesh.areas[o] / s.momDiag[o];
      const double compact = (s.bcPressure[fu] - s.p[o]) * s.gDiff[fu];
      const double smooth = dot(s.gradP[o], s.gArea[fu]);
      s.flux[fu] = dens * dot(s.vel[o], s.gArea[fu]) - dens * mob * (compact - smooth);
    }
  }
}

// SimpleSolver.cpp:498-580
void assemblePressure(Solver& s) {
  const Mesh& mesh = *s.mesh;
  s.pres.clear();
  std::fill(s.netFlux.begin(), s.netFlux.end(), 0.0);
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const auto fu = static_cast<std::size_t>(f);
    const Face& face = mesh.faces[fu];
    const auto o = static_cast<std::size_t>(face.owner);
    s.netFlux[o] += s.flux[fu];
    if (face.neighbour >= 0) s.netFlux[static_cast<std::size_t>(face.neighbour)] -= s.flux[fu];
    if (face.neighbour >= 0) {
      const auto n = static_cast<std::size_t>(face.neighbour);
      const double w = s.gWeight[fu];
      const double dens = w * s.rho[o] + (1.0 - w) * s.rho[n];
      const double mob = w * (mesh.areas[o] / s.momDiag[o]) + (1.0 - w) * (mesh.areas[n] / s.momDiag[n]);
      const double coef = dens * mob * s.gDiff[fu];
      s.pres.diag[o] += coef;
      s.pres.upper[fu] += -coef;
      s.pres.diag[n] += coef;
      s.pres.lower[fu] += -coef;
      continue;
    }
    const int across = opposite(mesh, f);
    if (across >= 0) {
      const auto n = static_cast<std::size_t>(across);
      const double w = s.gWeight[fu];
      const double dens = w * s.rho[o] + (1.0 - w) * s.rho[n];
      const double mob = w * (mesh.areas[o] / s.momDiag[o]) + (1.0 - w) * (mesh.areas[n] / s.momDiag[n]);
      const double coef = dens * mob * s.gDiff[fu];
      s.pres.diag[o] += coef;
      s.pres.upper[fu] += -coef;
      continue;
    }
    if (imposesVelocity(s, f)) continue;
    const double dens = s.rho[o];
    const double mob = mesh.areas[o] / s.momDiag[o];
    s.pres.diag[o] += dens * mob * s.gDiff[fu];
  }
  for (std::size_t c = 0; c < s.netFlux.size(); ++c) s.pres.source[c] = -s.netFlux[c];
  if (!s.hasPressureRef) {
    double mean = orderedSum(static_cast<int>(s.pres.source.size()),
                             [&](int c) { return s.pres.source[static_cast<std::size_t>(c)]; });
    mean /= static_cast<double>(s.pres.source.size());
    for (double& v : s.pres.source) v -= mean;
  }
}

// SimpleSolver.cpp:582-660
void correctField(Solver& s, const std::vector<double>& pc) {
  const Mesh& mesh = *s.mesh;
  std::vector<double> fc(static_cast<std::size_t>(mesh.faceCount()), 0.0);
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const auto fu = static_cast<std::size_t>(f);
    const Face& face = mesh.faces[fu];
    const auto o = static_cast<std::size_t>(face.owner);
    const int other = otherCell(mesh, face);
    if (other >= 0) {
      const double w = s.gWeight[fu];
      fc[fu] = w * pc[o] + (1.0 - w) * pc[static_cast<std::size_t>(other)];
    } else if (imposesVelocity(s, f)) {
      fc[fu] = pc[o];
    } else {
      fc[fu] = 0.0;
    }
  }
  std::vector<V2> gc;
  greenGauss(mesh, fc.data(), gc);
  for (int c = 0; c < mesh.cellCount(); ++c) {
    const auto cu = static_cast<std::size_t>(c);
    s.p[cu] += s.settings.pressure_relaxation * pc[cu];
    const double mob = mesh.areas[cu] / s.momDiag[cu];
    s.vel[cu] -= gc[cu] * mob;
  }
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const auto fu = static_cast<std::size_t>(f);
    const Face& face = mesh.faces[fu];
    const auto o = static_cast<std::size_t>(face.owner);
    if (face.neighbour >= 0) {
      const auto n = static_cast<std::size_t>(face.neighbour);
      const double w = s.gWeight[fu];
      const double dens = w * s.rho[o] + (1.0 - w) * s.rho[n];
      const double mob = w * (mesh.areas[o] / s.momDiag[o]) + (1.0 - w) * (mesh.areas[n] / s.momDiag[n]);
      s.flux[fu] -= dens * mob * s.gDiff[fu] * (pc[n] - pc[o]);
      continue;
    }
    const int across = opposite(mesh, f);
    if (across >= 0) {
      const auto n = static_cast<std::size_t>(across);
      const double w = s.gWeight[fu];
      const double dens = w * s.rho[o] + (1.0 - w) * s.rho[n];
      const double mob = w * (mesh.areas[o] / s.momDiag[o]) + (1.0 - w) * (mesh.areas[n] / s.momDiag[n]);
      s.flux[fu] -= dens * mob * s.gDiff[fu] * (pc[n] - pc[o]);
      continue;
    }
    if (imposesVelocity(s, f)) continue;
    const double dens = s.rho[o];
    const double mob = mesh.areas[o] / s.momDiag[o];
    s.flux[fu] -= dens * mob * s.gDiff[fu] * (0.0 - pc[o]);
  }
}

// SimpleSolver.cpp:796-817
std::vector<double> divergence(const Solver& s) {
  const Mesh& mesh = *s.mesh;
  std::vector<double> net(static_cast<std::size_t>(mesh.cellCount()), 0.0);
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const auto fu = static_cast<std::size_t>(f);
    const Face& face = mesh.faces[fu];
    const auto o = static_cast<std::size_t>(face.owner);
    const double dens = std::max(s.rho[o], kTiny);
    const double vf = s.flux[fu] / dens;
    net[o] += vf;
    if (face.neighbour >= 0) net[static_cast<std::size_t>(face.neighbour)] -= vf;
  }
  for (int c = 0; c < mesh.cellCount(); ++c) {
    const auto cu = static_cast<std::size_t>(c);
    const double a = mesh.areas[cu];
    net[cu] = (a > 0.0) ? net[cu] / a : 0.0;
  }
  return net;
}

// ---- SST pieces ----

// KOmegaSST.cpp:104-200
bool sstInitialise(Solver& s) {
  SstState& t = s.sst;
  const Mesh& mesh = *s.mesh;
  if (!std::isfinite(s.settings.intensity) || s.settings.intensity < 0.0) return false;
  if (!std::isfinite(s.settings.viscosity_ratio) || s.settings.viscosity_ratio <= 0.0) return false;
  if (!std::isfinite(s.settings.length_scale) || s.settings.length_scale < 0.0) return false;
  const std::size_t n = static_cast<std::size_t>(mesh.cellCount());
  t.wallDistance = wallDistances(mesh, s.kind);
  double refSpeed = 0.0, refRho = 0.0, refMu = 0.0;
  for (std::size_t c = 0; c < n; ++c) {
    const double sp = len(s.vel[c]);
    if (sp > refSpeed) refSpeed = sp;
    refRho = std::max(refRho, s.rho[c]);
    refMu = std::max(refMu, s.mu[c]);
  }
  if (!(refSpeed > 0.0) || !(refRho > 0.0) || !(refMu > 0.0)) return false;
  const double intensity = std::max(s.settings.intensity, 1.0e-8);
  t.kIn = 1.5 * (intensity * refSpeed) * (intensity * refSpeed);
  double eddy = 0.0;
  if (s.settings.length_scale > 0.0) {
    t.omegaIn = std::max(std::sqrt(t.kIn) / (std::pow(t.c.beta_star, 0.25) * s.settings.length_scale), kMinOmega);
    eddy = refRho * t.kIn / t.omegaIn;
  } else {
    eddy = s.settings.viscosity_ratio * refMu;
    t.omegaIn = std::max(refRho * t.kIn / eddy, kMinOmega);
  }
  t.k.assign(n, t.kIn);
  t.omega.assign(n, t.omegaIn);
  t.mut.assign(n, eddy);
  for (std::size_t c = 0; c < n; ++c) {
    const double d = std::max(t.wallDistance[c], kTinySst);
    const double nu = s.mu[c] / std::max(s.rho[c], kTinySst);
    t.omega[c] = std::max(t.omegaIn, wallOmega(nu, d, t.k[c], t.c));
    t.mut[c] = s.rho[c] * t.k[c] / t.omega[c];
  }
  t.strain.assign(n, 0.0);
  t.f1.assign(n, 0.0);
  t.f2.assign(n, 0.0);
  t.cd.assign(n, 0.0);
  t.cdPlus.assign(n, 0.0);
  t.production.assign(n, 0.0);
  t.gradK.assign(n, V2{});
  t.gradOmega.assign(n, V2{});
  t.maxYPlus = 0.0;
  t.resFirst = t.resSecond = 0.0;
  return true;
}

// KOmegaSST.cpp:202-216
void sstStrain(Solver& s) {
  SstState& t = s.sst;
  for (std::size_t c = 0; c < t.strain.size(); ++c) {
    const double s11 = s.gradU[c].x;
    const double s22 = s.gradV[c].y;
    const double s12 = 0.5 * (s.gradU[c].y + s.gradV[c].x);
    t.strain[c] = std::sqrt(2.0 * (s11 * s11 + s22 * s22 + 2.0 * s12 * s12));
  }
}

// KOmegaSST.cpp:218-287
void sstBlending(Solver& s) {
  SstState& t = s.sst;
  const Mesh& mesh = *s.mesh;
  std::vector<double> fv;
  sstInterpolate(mesh, s.kind, t.k, 0.0, t.kIn, s.flux, fv);
  greenGauss(mesh, fv.data(), t.gradK);
  sstInterpolate(mesh, s.kind, t.omega, t.omega.empty() ? 0.0 : t.omega[0], t.omegaIn, s.flux, fv);
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const Face& face = mesh.faces[static_cast<std::size_t>(f)];
    if (face.neighbour < 0 && opposite(mesh, f) < 0 && s.kind[static_cast<std::size_t>(f)] == ORC_K_WALL)
      fv[static_cast<std::size_t>(f)] = t.omega[static_cast<std::size_t>(face.owner)];
  }
  greenGauss(mesh, fv.data(), t.gradOmega);
  const orc_sst_constants& k = t.c;
  for (std::size_t c = 0; c < t.f1.size(); ++c) {
    const double d = std::max(t.wallDistance[c], kTinySst);
    const double rho = s.rho[c];
    const double nu = s.muMol[c] / std::max(rho, kTinySst);
    const double omega = std::max(t.omega[c], kMinOmega);
    const double energy = std::max(t.k[c], 0.0);
    t.cd[c] = 2.0 * rho * k.sigma_w2 * dot(t.gradK[c], t.gradOmega[c]) / omega;
    t.cdPlus[c] = std::max(t.cd[c], 1.0e-10);
    const double term1 = std::sqrt(std::max(energy, 0.0)) / (k.beta_star * omega * d);
    const double term2 = 500.0 * nu / (d * d * omega);
    const double term3 = 4.0 * rho * k.sigma_w2 * energy / (t.cdPlus[c] * d * d);
    const double arg1 = std::min(std::max(term1, term2), term3);
    t.f1[c] = blendTanh(std::clamp(arg1 * arg1 * arg1 * arg1, 0.0, 50.0));
    const double arg2 = std::max(2.0 * term1, term2);
    t.f2[c] = blendTanh(std::clamp(arg2 * arg2, 0.0, 50.0));
  }
}

// KOmegaSST.cpp:289-303
void sstEddyViscosity(Solver& s) {
  SstState& t = s.sst;
  for (std::size_t i = 0; i < t.mut.size(); ++i) {
    const double energy = std::max(t.k[i], 0.0);
    const double omega = std::max(t.omega[i], kMinOmega);
    const double den = std::max(t.c.a1 * omega, t.strain[i] * t.f2[i]);
    t.mut[i] = s.rho[i] * t.c.a1 * energy / std::max(den, kTinySst);
  }
}

// KOmegaSST.cpp:305-345
void sstWall(Solver& s) {
  SstState& t = s.sst;
  const Mesh& mesh = *s.mesh;
  t.maxYPlus = 0.0;
  for (int f = 0; f < mesh.faceCount(); ++f) {
    if (s.kind[static_cast<std::size_t>(f)] != ORC_K_WALL) continue;
    const auto o = static_cast<std::size_t>(mesh.faces[static_cast<std::size_t>(f)].owner);
    const double d = std::max(t.wallDistance[o], kTinySst);
    const double rho = std::max(s.rho[o], kTinySst);
    const double nu = s.muMol[o] / rho;
    t.omega[o] = wallOmega(nu, d, t.k[o], t.c);
    const V2 normal = mesh.fnormal[static_cast<std::size_t>(f)] * -1.0;
    const V2 vel = s.vel[o];
    const V2 par = vel - normal * dot(vel, normal);
    const double tang = len(par);
    const double shear = s.muMol[o] * tang / d;
    const double fv = std::sqrt(shear / rho);
    t.maxYPlus = std::max(t.maxYPlus, fv * d / std::max(nu, kTinySst));
  }
}

// KOmegaSST.cpp:347-510
void sstTransport(Solver& s) {
  SstState& t = s.sst;
  const Mesh& mesh = *s.mesh;
  const orc_sst_constants& c = t.c;
  const std::size_t n = static_cast<std::size_t>(mesh.cellCount());
  Ldu sys;
  sys.resize(n, static_cast<std::size_t>(mesh.faceCount()));
  for (std::size_t i = 0; i < n; ++i) {
    const double raw = t.mut[i] * t.strain[i] * t.strain[i];
    const double ceiling = c.production_limit * c.beta_star * s.rho[i] * std::max(t.k[i], 0.0) * std::max(t.omega[i], kMinOmega);
    t.production[i] = std::min(raw, ceiling);
  }
  auto blendOf = [&](double nearV, double farV, std::size_t i) { return t.f1[i] * nearV + (1.0 - t.f1[i]) * farV; };
  auto solve = [&](std::vector<double>& phi, double wallValue, double inflowValue, const std::vector<double>& sigma,
                   const std::vector<double>& srcE, const std::vector<double>& srcI, double floorV) {
    sys.clear();
    for (int f = 0; f < mesh.faceCount(); ++f) {
      const auto fu = static_cast<std::size_t>(f);
      const Face& face = mesh.faces[fu];
      const auto o = static_cast<std::size_t>(face.owner);
      const int other = (face.neighbour >= 0) ? face.neighbour : opposite(mesh, f);
      const V2 area = mesh.fnormal[fu] * mesh.farea[fu];
      const double F = s.flux[fu];
      if (other >= 0) {
        const auto nb = static_cast<std::size_t>(other);
        const double w = ownerWeight(mesh, f);
        const double diffusivity = w * (s.muMol[o] + sigma[o] * t.mut[o]) + (1.0 - w) * (s.muMol[nb] + sigma[nb] * t.mut[nb]);
        const V2 delta = mesh.centroids[nb] - mesh.centroids[o];
        const double ad = dot(area, delta);
        const double geo = (std::abs(ad) > kTinySst) ? dot(area, area) / ad : 0.0;
        const double diff = diffusivity * std::abs(geo);
        const double outward = std::max(F, 0.0);
        const double inward = std::max(-F, 0.0);
        sys.diag[o] += outward + diff;
        sys.upper[fu] += -inward - diff;
        if (face.neighbour >= 0) {
          sys.diag[nb] += inward + diff;
          sys.lower[fu] += -outward - diff;
        }
        continue;
      }
      const double diffusivity = s.muMol[o] + sigma[o] * t.mut[o];
      const V2 delta = mesh.fcentre[fu] - mesh.centroids[o];
      const double ad = dot(area, delta);
      const double geo = (std::abs(ad) > kTinySst) ? dot(area, area) / ad : 0.0;
      const double diff = diffusivity * std::abs(geo);
      const int kd = s.kind[fu];
      const bool imposed = kd == ORC_K_WALL || kd == ORC_K_INLET || (kd == ORC_K_FARFIELD && F < 0.0);
      if (imposed) {
        const double value = (kd == ORC_K_WALL) ? wallValue : inflowValue;
        sys.diag[o] += diff + std::max(F, 0.0);
        sys.source[o] += (diff - std::min(F, 0.0)) * value;
      } else {
        sys.diag[o] += std::max(F, 0.0);
      }
    }
    for (std::size_t i = 0; i < n; ++i) {
      const double vol = mesh.areas[i];
      sys.source[i] += srcE[i] * vol;
      sys.diag[i] += srcI[i] * vol;
    }
    const double relax = std::clamp(s.settings.turbulence_relaxation, 0.05, 1.0);
    for (std::size_t i = 0; i < n; ++i) {
      if (std::abs(sys.diag[i]) > kTinySst) {
        const double relaxed = sys.diag[i] / relax;
        sys.source[i] += (relaxed - sys.diag[i]) * phi[i];
        sys.diag[i] = relaxed;
      }
    }
    const double residual = residualL1(mesh, sys, phi) / std::max(diagonalL1(sys) * std::max(inflowValue, kTinySst), kTinySst);
    gaussSeidel(mesh, sys.diag.data(), sys.upper.data(), sys.lower.data(), sys.source.data(), phi.data(), 3);
    for (std::size_t i = 0; i < n; ++i) phi[i] = std::max(phi[i], floorV);
    return residual;
  };
  std::vector<double> sigmaK(n), kP(n), kD(n);
  for (std::size_t i = 0; i < n; ++i) {
    sigmaK[i] = blendOf(c.sigma_k1, c.sigma_k2, i);
    kP[i] = t.production[i];
    kD[i] = c.beta_star * s.rho[i] * std::max(t.omega[i], kMinOmega);
  }
  t.resFirst = solve(t.k, 0.0, t.kIn, sigmaK, kP, kD, kMinEnergy);
  std::vector<double> sigmaW(n), wP(n), wD(n);
  const double g1 = t.gamma1(), g2 = t.gamma2();
  for (std::size_t i = 0; i < n; ++i) {
    sigmaW[i] = blendOf(c.sigma_w1, c.sigma_w2, i);
    const double gamma = blendOf(g1, g2, i);
    const double beta = blendOf(c.beta1, c.beta2, i);
    const double rho = s.rho[i];
    wP[i] = gamma * rho * t.strain[i] * t.strain[i];
    wP[i] += (1.0 - t.f1[i]) * t.cd[i];
    wD[i] = beta * rho * std::max(t.omega[i], kMinOmega);
  }
  t.resSecond = solve(t.omega, t.omegaIn, t.omegaIn, sigmaW, wP, wD, kMinOmega);
}

// KOmegaSST.cpp:513-529
void sstRanges(Solver& s) {
  SstState& t = s.sst;
  for (double& r : t.ranges) r = 0.0;
  if (t.k.empty()) return;
  t.ranges[0] = t.ranges[1] = t.k[0];
  t.ranges[2] = t.ranges[3] = t.omega[0];
  t.ranges[4] = t.ranges[5] = t.mut[0];
  for (std::size_t c = 1; c < t.k.size(); ++c) {
    t.ranges[0] = std::min(t.ranges[0], t.k[c]);
    t.ranges[1] = std::max(t.ranges[1], t.k[c]);
    t.ranges[2] = std::min(t.ranges[2], t.omega[c]);
    t.ranges[3] = std::max(t.ranges[3], t.omega[c]);
    t.ranges[4] = std::min(t.ranges[4], t.mut[c]);
    t.ranges[5] = std::max(t.ranges[5], t.mut[c]);
  }
}

// SimpleSolver.cpp:232-245
void applyEddyViscosity(Solver& s) {
  if (s.model == 0) return;
  const std::vector<double>& eddy = (s.model == 2) ? s.sst.mut : s.eddyZero;
  if (eddy.size() != s.mu.size()) return;
  for (std::size_t c = 0; c < s.mu.size(); ++c) s.mu[c] = s.muMol[c] + eddy[c];
}

bool modelInitialise(Solver& s) {
  if (s.model == 1) {
    s.eddyZero.assign(static_cast<std::size_t>(s.mesh->cellCount()), 0.0);
    return true;
  }
  if (s.model == 2) return sstInitialise(s);
  return true;
}

// SimpleSolver.cpp:662-794
void iterateOnce(Solver& s, orc_monitor& mon) {
  const Mesh& mesh = *s.mesh;
  std::memset(&mon, 0, sizeof(mon));
  interpolateFaceValues(s);
  computeGradients(s);
  const std::size_t n = static_cast<std::size_t>(mesh.cellCount());
  std::vector<double> sol(n, 0.0);
  double momRes[2] = {0.0, 0.0};
  double vscale = 0.0;
  for (std::size_t c = 0; c < n; ++c) vscale = std::max(vscale, len(s.vel[c]));
  if (!(vscale > 0.0)) vscale = 1.0;
  for (int index = 0; index < 2; ++index) {
    assembleMomentum(s, index);
    applyRelaxation(s, index);
    for (std::size_t c = 0; c < n; ++c) sol[c] = comp(s.vel[c], index);
    const double residual = residualL1(mesh, s.mom, sol);
    const double scale = diagonalL1(s.mom) * vscale;
    momRes[index] = residual / std::max(scale, kTiny);
    gaussSeidel(mesh, s.mom.diag.data(), s.mom.upper.data(), s.mom.lower.data(), s.mom.source.data(), sol.data(),
                s.settings.momentum_sweeps);
    for (std::size_t c = 0; c < n; ++c) {
      if (index == 0) s.vel[c].x = sol[c]; else s.vel[c].y = sol[c];
    }
    if (index == 0) s.momDiag = s.mom.diag;
  }
  rhieChow(s);
  std::vector<double> corr(n, 0.0);
  for (int pass = 0; pass < std::max(1, s.settings.non_orthogonal_correctors); ++pass) {
    assemblePressure(s);
    if (pass == 0) {
      const double imb = orderedSum(static_cast<int>(s.netFlux.size()),
                                    [&](int c) { return std::abs(s.netFlux[static_cast<std::size_t>(c)]); });
      mon.mass_imbalance = imb;
      mon.continuity = imb / s.refMassFlow;
    }
    std::fill(corr.begin(), corr.end(), 0.0);
    const CgReport rep = conjugateGradient(mesh, s.pres.diag.data(), s.pres.upper.data(), s.pres.lower.data(),
                                           s.pres.source.data(), corr.data(), s.settings.pressure_tolerance,
                                           s.settings.pressure_iterations);
    mon.pressure_iterations += rep.iterations;
    if (!s.hasPressureRef) {
      double mean = orderedSum(static_cast<int>(corr.size()),
                               [&](int q) { return corr[static_cast<std::size_t>(q)]; });
      mean /= static_cast<double>(corr.size());
      for (double& v : corr) v -= mean;
    }
    correctField(s, corr);
    s.lastCorrection = corr;
  }
  mon.momentum_x = momRes[0];
  mon.momentum_y = momRes[1];
  const std::vector<double> dv = divergence(s);
  for (const double v : dv) mon.max_divergence = std::max(mon.max_divergence, std::abs(v));
  if (s.model != 0) {
    if (s.model == 2) {
      // KOmegaSST.cpp:531-548
      sstStrain(s);
      sstBlending(s);
      sstTransport(s);
      sstWall(s);
      sstEddyViscosity(s);
      sstRanges(s);
    }
    applyEddyViscosity(s);
    if (s.model == 2) {
      mon.turb_first = s.sst.resFirst;
      mon.turb_second = s.sst.resSecond;
      mon.min_k = s.sst.ranges[0];
      mon.max_k = s.sst.ranges[1];
      mon.min_w = s.sst.ranges[2];
      mon.max_w = s.sst.ranges[3];
      mon.min_mut = s.sst.ranges[4];
      mon.max_mut = s.sst.ranges[5];
    }
    const std::vector<double>& eddy = (s.model == 2) ? s.sst.mut : s.eddyZero;
    mon.max_eddy_ratio = 0.0;
    for (std::size_t c = 0; c < eddy.size(); ++c) {
      const double m = s.muMol[c];
      if (m > 0.0) mon.max_eddy_ratio = std::max(mon.max_eddy_ratio, eddy[c] / m);
    }
  }
}

// ---- post-processing: src/post/SurfaceData.cpp, src/post/Forces.cpp ----

struct WallFace {
  int face{0}, owner{0};
  V2 centre, tangent, normal;
  double length{0.0};
};

struct Station {
  int face{0};
  V2 position, tangent, normal;
  double segment{0.0}, xc{0.0}, pressure{0.0}, cp{0.0}, shear{0.0}, cf{0.0}, speed{0.0}, arc{0.0};
  bool reversed{false}, nearStag{false};
};

// SurfaceData.cpp:67-98
int secondCell(const Mesh& mesh, const WallFace& wall) {
  int best = -1;
  double bestAlign = 0.5;
  for (const int f : mesh.cellFaces[static_cast<std::size_t>(wall.owner)]) {
    if (f < 0 || f == wall.face) continue;
    const int across = opposite(mesh, f);
    if (across < 0 || across == wall.owner) continue;
    V2 outward = mesh.fnormal[static_cast<std::size_t>(f)];
    if (mesh.faces[static_cast<std::size_t>(f)].owner != wall.owner) outward = outward * -1.0;
    const double a = dot(outward, wall.normal);
    if (a > bestAlign) {
      bestAlign = a;
      best = across;
    }
  }
  return best;
}

// SurfaceData.cpp:101-172
Station evaluateStation(const Mesh& mesh, const std::vector<V2>& vel, const double* p, const double* mu,
                        const WallFace& wall, V2 tangent, double q, double pref, double chord) {
  Station st;
  st.face = wall.face;
  st.position = wall.centre;
  st.tangent = tangent;
  st.normal = wall.normal;
  st.segment = wall.length;
  st.xc = (chord > 0.0) ? wall.centre.x / chord : wall.centre.x;
  st.pressure = p[wall.owner];
  st.cp = (q > 0.0) ? (st.pressure - pref) / q : 0.0;
  const V2 offset = mesh.centroids[static_cast<std::size_t>(wall.owner)] - wall.centre;
  const double wd = std::abs(dot(offset, wall.normal));
  const V2 cv = vel[static_cast<std::size_t>(wall.owner)];
  const V2 par = cv - wall.normal * dot(cv, wall.normal);
  st.speed = len(par);
  if (wd > 0.0) {
    double g = dot(par, tangent) / wd;
    const int second = secondCell(mesh, wall);
    if (second >= 0) {
      const V2 oo = mesh.centroids[static_cast<std::size_t>(second)] - wall.centre;
      const double od = std::abs(dot(oo, wall.normal));
      if (od > wd * 1.05) {
        const V2 ov = vel[static_cast<std::size_t>(second)];
        const V2 opar = ov - wall.normal * dot(ov, wall.normal);
        const double u1 = dot(par, tangent);
        const double u2 = dot(opar, tangent);
        const double d1 = wd, d2 = od;
        g = (u1 * d2 * d2 - u2 * d1 * d1) / (d1 * d2 * (d2 - d1));
      }
    }
    st.shear = mu[wall.owner] * g;
  }
  st.cf = (q > 0.0) ? st.shear / q : 0.0;
  st.reversed = st.shear < 0.0;
  return st;
}

// SurfaceData.cpp:179-204
double findSeparation(const std::vector<Station>& s) {
  for (std::size_t i = 1; i < s.size(); ++i) {
    if (s[i].nearStag || s[i - 1].nearStag) continue;
    const double before = s[i - 1].shear, after = s[i].shear;
    if (!(before > 0.0) || !(after < 0.0)) continue;
    const double span = before - after;
    const double fr = (span > 0.0) ? before / span : 0.0;
    return s[i - 1].xc + fr * (s[i].xc - s[i - 1].xc);
  }
  return -1.0;
}

}  // namespace

// ============================================================================
// C entry points
// ============================================================================

extern "C" {

int orc_naca_contour(int m, int p, int t, int points, double chord, int closed, double* xy, int capacity) {
  const Section s = naca(m, p, t, points, chord, closed != 0);
  const int n = static_cast<int>(s.contour.size());
  for (int i = 0; i < n && i < capacity; ++i) {
    xy[2 * i] = s.contour[static_cast<std::size_t>(i)].x;
    xy[2 * i + 1] = s.contour[static_cast<std::size_t>(i)].y;
  }
  return n;
}
double orc_thickness(double x, double t, int closed) { return thickness(x, t, closed != 0); }
double orc_camber(double x, double m, double p) { return camber(x, m, p); }
double orc_camber_slope(double x, double m, double p) { return camberSlope(x, m, p); }

int orc_geometric_distribution(double first, double total, int intervals, double* out) {
  const std::vector<double> t = geometricDist(first, total, intervals);
  for (std::size_t i = 0; i < t.size(); ++i) out[i] = t[i];
  return static_cast<int>(t.size());
}

// CGrid.cpp:73-96 presets, CGrid.hpp defaults for the rest
int orc_cgrid_defaults(int resolution, orc_cgrid_options* o) {
  o->upstream = 12.0;
  o->downstream = 25.0;
  o->vertical = 12.0;
  o->normal_blend_length = 0.06;
  o->normal_smoothing_passes = 6;
  switch (resolution) {
    case 0: o->surface_points = 80; o->wake_points = 32; o->normal_points = 40; o->first_layer = 1.0e-3; break;
    case 1: o->surface_points = 160; o->wake_points = 56; o->normal_points = 72; o->first_layer = 3.0e-4; break;
    case 2: o->surface_points = 320; o->wake_points = 96; o->normal_points = 128; o->first_layer = 1.0e-4; break;
    default: return -1;
  }
  return 0;
}

orc_mesh* orc_cgrid(int m, int p, int t, double chord, const orc_cgrid_options* options) {
  return cgrid(m, p, t, chord, *options).release();
}
orc_mesh* orc_box(const orc_box_options* options) { return box(*options).release(); }
orc_mesh* orc_structured(int ni, int nj, const double* nodes_xy, const int* jmin_type, const int* jmin_partner,
                         int jmax_type, int imin_type, int imax_type) {
  std::vector<V2> nodes(static_cast<std::size_t>(ni) * static_cast<std::size_t>(nj));
  for (std::size_t i = 0; i < nodes.size(); ++i) nodes[i] = V2{nodes_xy[2 * i], nodes_xy[2 * i + 1]};
  std::vector<int> jmin(jmin_type, jmin_type + (ni - 1));
  std::vector<int> part;
  if (jmin_partner) part.assign(jmin_partner, jmin_partner + (ni - 1));
  return buildStructured(ni, nj, std::move(nodes), jmin, part, jmax_type, imin_type, imax_type).release();
}
void orc_mesh_free(orc_mesh* mesh) { delete mesh; }
void orc_mesh_dims(const orc_mesh* m, int* ni, int* nj, int* cells, int* faces) {
  *ni = m->ni;
  *nj = m->nj;
  *cells = m->cellCount();
  *faces = m->faceCount();
}
void orc_mesh_quality(const orc_mesh* m, double* out7) {
  for (int i = 0; i < 6; ++i) out7[i] = m->quality[i];
  out7[6] = m->totalArea;
}
void orc_mesh_nodes(const orc_mesh* m, double* xy) {
  for (std::size_t i = 0; i < m->nodes.size(); ++i) {
    xy[2 * i] = m->nodes[i].x;
    xy[2 * i + 1] = m->nodes[i].y;
  }
}
void orc_mesh_cells(const orc_mesh* m, double* cxy, double* area, int* cf4) {
  for (std::size_t c = 0; c < m->areas.size(); ++c) {
    if (cxy) { cxy[2 * c] = m->centroids[c].x; cxy[2 * c + 1] = m->centroids[c].y; }
    if (area) area[c] = m->areas[c];
    if (cf4) for (int k = 0; k < 4; ++k) cf4[4 * c + static_cast<std::size_t>(k)] = m->cellFaces[c][static_cast<std::size_t>(k)];
  }
}
void orc_mesh_faces(const orc_mesh* m, double* centre, double* normal, double* area, int* owner, int* neighbour,
                    int* boundary, int* partner) {
  for (std::size_t f = 0; f < m->faces.size(); ++f) {
    if (centre) { centre[2 * f] = m->fcentre[f].x; centre[2 * f + 1] = m->fcentre[f].y; }
    if (normal) { normal[2 * f] = m->fnormal[f].x; normal[2 * f + 1] = m->fnormal[f].y; }
    if (area) area[f] = m->farea[f];
    if (owner) owner[f] = m->faces[f].owner;
    if (neighbour) neighbour[f] = m->faces[f].neighbour;
    if (boundary) boundary[f] = m->faces[f].boundary;
    if (partner) partner[f] = m->faces[f].partner;
  }
}

// src/flow/BoundaryConditions.cpp:40-97 with the default BoundaryConditions
void orc_face_conditions(const orc_mesh* m, double speed, double alpha_deg, double pref, int* kind, double* vxy,
                         double* pressure) {
  const double a = alpha_deg * kPi / 180.0;  // Freestream.cpp:9-17
  const V2 stream{speed * std::cos(a), speed * std::sin(a)};
  for (std::size_t f = 0; f < m->faces.size(); ++f) {
    int k = ORC_K_INTERNAL;
    switch (m->faces[f].boundary) {
      case ORC_WALL: k = ORC_K_WALL; break;
      case ORC_FARFIELD: k = ORC_K_FARFIELD; break;
      case ORC_OUTLET: k = ORC_K_OUTLET; break;
      default: k = ORC_K_INTERNAL; break;
    }
    kind[f] = k;
    pressure[f] = pref;
    const V2 v = (k == ORC_K_WALL) ? V2{0.0, 0.0} : stream;
    vxy[2 * f] = v.x;
    vxy[2 * f + 1] = v.y;
  }
}

void orc_set_sum_block(int block) { gSumBlock = block > 0 ? block : (block < 0 ? -1 : 0); }
void orc_set_tanh_variant(int variant) { gTanhVariant = variant == 1 ? 1 : 0; }

// SimpleSolver.hpp:86-132 defaults
void orc_default_settings(orc_settings* s) {
  s->velocity_relaxation = 0.5;
  s->pressure_relaxation = 0.2;
  s->momentum_sweeps = 3;
  s->pressure_iterations = 200;
  s->pressure_tolerance = 0.01;
  s->scheme = 0;
  s->turbulence_relaxation = 0.5;
  s->intensity = 0.001;
  s->viscosity_ratio = 1.0;
  s->length_scale = 0.0;
  s->non_orthogonal_correctors = 1;
  s->non_orthogonal_blend = 1.0;
}

// KOmegaSST.hpp:147-171
void orc_default_sst(orc_sst_constants* c) {
  c->sigma_k1 = 0.85; c->sigma_w1 = 0.5; c->beta1 = 0.075;
  c->sigma_k2 = 1.0; c->sigma_w2 = 0.856; c->beta2 = 0.0828;
  c->beta_star = 0.09; c->kappa = 0.41; c->a1 = 0.31; c->production_limit = 10.0;
}

// SimpleSolver.cpp:30-80
orc_solver* orc_solver_create(const orc_mesh* mesh, const int* kind, const double* vxy, const double* pressure,
                              const orc_settings* settings) {
  if (mesh->cellCount() == 0) return nullptr;
  if (!(settings->velocity_relaxation > 0.0) || settings->velocity_relaxation > 1.0 ||
      !(settings->pressure_relaxation > 0.0) || settings->pressure_relaxation > 1.0)
    return nullptr;
  auto* s = new orc_solver();
  s->mesh = mesh;
  const std::size_t nf = static_cast<std::size_t>(mesh->faceCount());
  const std::size_t nc = static_cast<std::size_t>(mesh->cellCount());
  s->kind.assign(kind, kind + nf);
  s->bcVelocity.resize(nf);
  s->bcPressure.assign(pressure, pressure + nf);
  for (std::size_t f = 0; f < nf; ++f) s->bcVelocity[f] = V2{vxy[2 * f], vxy[2 * f + 1]};
  s->settings = *settings;
  s->flux.assign(nf, 0.0);
  s->deadBand.assign(nf, 0.0);
  s->faceU.assign(nf, 0.0);
  s->faceV.assign(nf, 0.0);
  s->faceP.assign(nf, 0.0);
  s->gradU.assign(nc, V2{});
  s->gradV.assign(nc, V2{});
  s->gradP.assign(nc, V2{});
  s->mom.resize(nc, nf);
  s->pres.resize(nc, nf);
  s->momDiag.assign(nc, 1.0);
  s->netFlux.assign(nc, 0.0);
  precomputeGeometry(*s);
  for (std::size_t f = 0; f < nf; ++f) {
    if (s->kind[f] == ORC_K_OUTLET || s->kind[f] == ORC_K_FARFIELD) {
      s->hasPressureRef = true;
      break;
    }
  }
  orc_default_sst(&s->sst.c);
  return s;
}
void orc_solver_free(orc_solver* s) { delete s; }

// SimpleSolver.cpp:212-230
int orc_solver_set_model(orc_solver* s, int model, const orc_sst_constants* constants) {
  s->model = model;
  if (constants) s->sst.c = *constants; else orc_default_sst(&s->sst.c);
  if (model == 0) {
    if (!s->muMol.empty()) s->mu = s->muMol;
    return 0;
  }
  if (s->muMol.empty()) return 0;
  if (!modelInitialise(*s)) return -1;
  applyEddyViscosity(*s);
  return 0;
}

// SimpleSolver.cpp:133-210
int orc_solver_initialise(orc_solver* s, const double* u, const double* v, const double* p, const double* rho,
                          const double* mu) {
  const Mesh& mesh = *s->mesh;
  const std::size_t nc = static_cast<std::size_t>(mesh.cellCount());
  s->vel.resize(nc);
  s->p.assign(p, p + nc);
  s->rho.assign(rho, rho + nc);
  s->mu.assign(mu, mu + nc);
  for (std::size_t c = 0; c < nc; ++c) s->vel[c] = V2{u[c], v[c]};
  s->muMol = s->mu;
  double refSpeed = 0.0;
  for (std::size_t c = 0; c < nc; ++c) refSpeed = std::max(refSpeed, len(s->vel[c]));
  std::fill(s->deadBand.begin(), s->deadBand.end(), 0.0);
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const auto fu = static_cast<std::size_t>(f);
    if (s->kind[fu] != ORC_K_FARFIELD) continue;
    const auto o = static_cast<std::size_t>(mesh.faces[fu].owner);
    s->deadBand[fu] = 1.0e-3 * s->rho[o] * refSpeed * len(s->gArea[fu]);
  }
  interpolateFaceValues(*s);
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const auto fu = static_cast<std::size_t>(f);
    const auto o = static_cast<std::size_t>(mesh.faces[fu].owner);
    const double w = s->gWeight[fu];
    double dens = s->rho[o];
    if (mesh.faces[fu].neighbour >= 0) {
      const auto n = static_cast<std::size_t>(mesh.faces[fu].neighbour);
      dens = w * s->rho[o] + (1.0 - w) * s->rho[n];
    }
    s->flux[fu] = dens * (s->faceU[fu] * s->gArea[fu].x + s->faceV[fu] * s->gArea[fu].y);
  }
  s->refMassFlow = 0.0;
  for (int f = 0; f < mesh.faceCount(); ++f) {
    const auto fu = static_cast<std::size_t>(f);
    if (mesh.faces[fu].boundary == ORC_INTERIOR || s->kind[fu] == ORC_K_INTERNAL) continue;
    s->refMassFlow += std::max(-s->flux[fu], 0.0);
  }
  if (!(s->refMassFlow > 0.0)) s->refMassFlow = 1.0;
  if (s->model != 0) {
    if (!modelInitialise(*s)) return -1;
    applyEddyViscosity(*s);
  }
  return 0;
}

void orc_solver_set_settings(orc_solver* s, const orc_settings* settings) { s->settings = *settings; }

int orc_solver_iterate(orc_solver* s, int count, orc_monitor* last) {
  orc_monitor mon{};
  for (int i = 0; i < count; ++i) iterateOnce(*s, mon);
  if (last) *last = mon;
  return 0;
}

void orc_solver_field(const orc_solver* s, double* u, double* v, double* p, double* mu) {
  for (std::size_t c = 0; c < s->vel.size(); ++c) {
    if (u) u[c] = s->vel[c].x;
    if (v) v[c] = s->vel[c].y;
    if (p) p[c] = s->p[c];
    if (mu) mu[c] = s->mu[c];
  }
}
void orc_solver_set_field(orc_solver* s, const double* u, const double* v, const double* p) {
  for (std::size_t c = 0; c < s->vel.size(); ++c) {
    if (u) s->vel[c].x = u[c];
    if (v) s->vel[c].y = v[c];
    if (p) s->p[c] = p[c];
  }
}
void orc_solver_flux(const orc_solver* s, double* flux) { std::copy(s->flux.begin(), s->flux.end(), flux); }
void orc_solver_set_flux(orc_solver* s, const double* flux) { std::copy(flux, flux + s->flux.size(), s->flux.begin()); }
int orc_solver_turbulence(const orc_solver* s, double* k, double* w, double* mut, double* wd, double* f1, double* cd) {
  if (s->model != 2) return -1;
  const SstState& t = s->sst;
  for (std::size_t c = 0; c < t.k.size(); ++c) {
    if (k) k[c] = t.k[c];
    if (w) w[c] = t.omega[c];
    if (mut) mut[c] = t.mut[c];
    if (wd) wd[c] = t.wallDistance[c];
    if (f1) f1[c] = t.f1[c];
    if (cd) cd[c] = t.cd[c];
  }
  return 0;
}
void orc_solver_set_turbulence(orc_solver* s, const double* k, const double* w, const double* mut) {
  SstState& t = s->sst;
  for (std::size_t c = 0; c < t.k.size(); ++c) {
    if (k) t.k[c] = k[c];
    if (w) t.omega[c] = w[c];
    if (mut) t.mut[c] = mut[c];
  }
  applyEddyViscosity(*s);
}
void orc_solver_divergence(const orc_solver* s, double* out) {
  const std::vector<double> d = divergence(*s);
  std::copy(d.begin(), d.end(), out);
}
double orc_solver_freestream_k(const orc_solver* s) { return s->sst.kIn; }
double orc_solver_freestream_omega(const orc_solver* s) { return s->sst.omegaIn; }
double orc_solver_max_yplus(const orc_solver* s) { return s->sst.maxYPlus; }

int orc_solver_array(const orc_solver* s, const char* name, double* out, int capacity) {
  const std::string n(name);
  auto put = [&](const std::vector<double>& v) {
    const int sz = static_cast<int>(v.size());
    for (int i = 0; i < sz && i < capacity; ++i) out[i] = v[static_cast<std::size_t>(i)];
    return sz;
  };
  auto put2 = [&](const std::vector<V2>& v) {
    const int sz = static_cast<int>(v.size());
    for (int i = 0; i < sz && 2 * i + 1 < capacity; ++i) {
      out[2 * i] = v[static_cast<std::size_t>(i)].x;
      out[2 * i + 1] = v[static_cast<std::size_t>(i)].y;
    }
    return 2 * sz;
  };
  if (n == "faceU") return put(s->faceU);
  if (n == "faceV") return put(s->faceV);
  if (n == "faceP") return put(s->faceP);
  if (n == "gradU") return put2(s->gradU);
  if (n == "gradV") return put2(s->gradV);
  if (n == "gradP") return put2(s->gradP);
  if (n == "momDiag") return put(s->mom.diag);
  if (n == "momSource") return put(s->mom.source);
  if (n == "momUpper") return put(s->mom.upper);
  if (n == "momLower") return put(s->mom.lower);
  if (n == "momentumDiagonal") return put(s->momDiag);
  if (n == "pDiag") return put(s->pres.diag);
  if (n == "pSource") return put(s->pres.source);
  if (n == "pUpper") return put(s->pres.upper);
  if (n == "pLower") return put(s->pres.lower);
  if (n == "netFlux") return put(s->netFlux);
  if (n == "correction") return put(s->lastCorrection);
  if (n == "deadBand") return put(s->deadBand);
  if (n == "geomDiffusion") return put(s->gDiff);
  if (n == "geomWeight") return put(s->gWeight);
  if (n == "strain") return put(s->sst.strain);
  if (n == "f2") return put(s->sst.f2);
  return -1;
}

void orc_stage_interpolate(orc_solver* s) { interpolateFaceValues(*s); }
void orc_stage_gradients(orc_solver* s) { computeGradients(*s); }
void orc_stage_momentum(orc_solver* s, int component) {
  assembleMomentum(*s, component);
  applyRelaxation(*s, component);
}
void orc_stage_rhie_chow(orc_solver* s) { rhieChow(*s); }
void orc_stage_pressure(orc_solver* s) { assemblePressure(*s); }
int orc_stage_cg(orc_solver* s, double* x) {
  const std::size_t n = s->pres.diag.size();
  std::vector<double> corr(n, 0.0);
  const CgReport rep = conjugateGradient(*s->mesh, s->pres.diag.data(), s->pres.upper.data(), s->pres.lower.data(),
                                         s->pres.source.data(), corr.data(), s->settings.pressure_tolerance,
                                         s->settings.pressure_iterations);
  std::copy(corr.begin(), corr.end(), x);
  return rep.iterations;
}
void orc_stage_correct(orc_solver* s, const double* correction) {
  std::vector<double> pc(correction, correction + s->vel.size());
  correctField(*s, pc);
}
void orc_set_momentum_diagonal(orc_solver* s, const double* diag) {
  std::copy(diag, diag + s->momDiag.size(), s->momDiag.begin());
}

void orc_ldu_multiply(const orc_mesh* mesh, const double* diag, const double* upper, const double* lower,
                      const double* x, double* y) {
  lduMultiply(*mesh, diag, upper, lower, x, y);
}
void orc_ldu_gauss_seidel(const orc_mesh* mesh, const double* diag, const double* upper, const double* lower,
                          const double* source, double* x, int sweeps) {
  gaussSeidel(*mesh, diag, upper, lower, source, x, sweeps);
}
int orc_ldu_cg(const orc_mesh* mesh, const double* diag, const double* upper, const double* lower,
               const double* source, double* x, double tol, int maxIt, double* r0, double* r1) {
  const CgReport rep = conjugateGradient(*mesh, diag, upper, lower, source, x, tol, maxIt);
  if (r0) *r0 = rep.initial;
  if (r1) *r1 = rep.final;
  return rep.converged ? rep.iterations : -1 - rep.iterations;
}
int orc_green_gauss(const orc_mesh* mesh, const double* fv, double* gxy) {
  std::vector<V2> g;
  greenGauss(*mesh, fv, g);
  for (std::size_t c = 0; c < g.size(); ++c) {
    gxy[2 * c] = g[c].x;
    gxy[2 * c + 1] = g[c].y;
  }
  return static_cast<int>(g.size());
}

// SurfaceData.cpp:208-348 + Forces.cpp:60-130. Station rows (12 doubles):
// side(1 upper, 0 lower), face, x, y, chordFraction, arcLength, pressure, cp,
// wallShear, cf, nearStagnation, segmentLength.
int orc_surface(const orc_mesh* mp, const double* u, const double* v, const double* p, const double* mu, double speed,
                double alpha_deg, double density, double pref, double chord, double* rows, int capacity,
                orc_forces* out) {
  const Mesh& mesh = *mp;
  std::vector<V2> vel(static_cast<std::size_t>(mesh.cellCount()));
  for (std::size_t c = 0; c < vel.size(); ++c) vel[c] = V2{u[c], v[c]};
  // collectWallFaces :27-59
  std::vector<WallFace> wall;
  const int ci = mesh.ni - 1;
  for (int i = 0; i < ci; ++i) {
    const Face& face = mesh.faces[static_cast<std::size_t>(i)];
    if (face.boundary != ORC_WALL) continue;
    const V2 a = mesh.nodes[static_cast<std::size_t>(face.n0)];
    const V2 b = mesh.nodes[static_cast<std::size_t>(face.n1)];
    const V2 along = b - a;
    const double l = len(along);
    if (!(l > 0.0)) continue;
    WallFace w;
    w.face = i;
    w.owner = face.owner;
    w.centre = mesh.fcentre[static_cast<std::size_t>(i)];
    w.tangent = along * (1.0 / l);
    w.normal = mesh.fnormal[static_cast<std::size_t>(i)] * -1.0;
    w.length = l;
    wall.push_back(w);
  }
  if (wall.size() < 4) return -1;
  std::size_t le = 0;
  double fwd = std::numeric_limits<double>::max();
  for (std::size_t i = 0; i < wall.size(); ++i) {
    const V2 node = mesh.nodes[static_cast<std::size_t>(mesh.faces[static_cast<std::size_t>(wall[i].face)].n0)];
    if (node.x < fwd) {
      fwd = node.x;
      le = i;
    }
  }
  const double q = 0.5 * density * speed * speed;  // Freestream.cpp:19-21
  std::size_t stag = 0;
  double hp = -std::numeric_limits<double>::max();
  for (std::size_t i = 0; i < wall.size(); ++i) {
    const double pr = p[wall[i].owner];
    if (pr > hp) {
      hp = pr;
      stag = i;
    }
  }
  auto flowTangent = [&](std::size_t i) {
    if (i == stag) return (i >= le) ? wall[i].tangent : wall[i].tangent * -1.0;
    return (i > stag) ? wall[i].tangent : wall[i].tangent * -1.0;
  };
  std::vector<Station> lower, upper;
  for (std::size_t i = le; i-- > 0;) {
    lower.push_back(evaluateStation(mesh, vel, p, mu, wall[i], flowTangent(i), q, pref, chord));
    lower.back().nearStag = (i == stag);
  }
  for (std::size_t i = le; i < wall.size(); ++i) {
    upper.push_back(evaluateStation(mesh, vel, p, mu, wall[i], flowTangent(i), q, pref, chord));
    upper.back().nearStag = (i == stag);
  }
  for (auto* side : {&upper, &lower})
    for (Station& st : *side)
      if (st.nearStag) st.reversed = false;
  auto accumulate = [](std::vector<Station>& s) {
    double t = 0.0;
    for (std::size_t i = 0; i < s.size(); ++i) {
      if (i > 0) t += dist(s[i - 1].position, s[i].position);
      s[i].arc = t;
    }
  };
  accumulate(lower);
  accumulate(upper);
  orc_forces f{};
  f.upper_separation = findSeparation(upper);
  f.lower_separation = findSeparation(lower);
  f.stagnation_xc = (chord > 0.0) ? wall[stag].centre.x / chord : wall[stag].centre.x;
  f.n_upper = static_cast<int>(upper.size());
  f.n_lower = static_cast<int>(lower.size());
  // Forces.cpp:16-46, 60-130
  V2 pf{}, ff{};
  double pm = 0.0, fm = 0.0;
  int stations = 0;
  const V2 ref{0.25 * chord, 0.0};
  for (const auto* side : {&upper, &lower}) {
    for (const Station& st : *side) {
      const double ds = st.segment;
      if (!(ds > 0.0)) continue;
      const V2 dP = st.normal * (-(st.pressure - pref) * ds);
      const V2 dF = st.tangent * (st.shear * ds);
      pf = pf + dP;
      ff = ff + dF;
      const V2 arm = st.position - ref;
      pm += arm.x * dP.y - arm.y * dP.x;
      fm += arm.x * dF.y - arm.y * dF.x;
      ++stations;
    }
  }
  const double a = alpha_deg * kPi / 180.0;
  const double ca = std::cos(a), sa = std::sin(a);
  auto dragOf = [&](V2 v) { return v.x * ca + v.y * sa; };
  auto liftOf = [&](V2 v) { return v.y * ca - v.x * sa; };
  const double pd = dragOf(pf), fd = dragOf(ff), pl = liftOf(pf), fl = liftOf(ff);
  f.drag = pd + fd;
  f.lift = pl + fl;
  f.pitching_moment = -pm - fm;
  const double fs = q * chord;
  const double ms = fs * chord;
  f.cl = f.lift / fs;
  f.cd = f.drag / fs;
  f.cd_pressure = pd / fs;
  f.cd_friction = fd / fs;
  f.cm = f.pitching_moment / ms;
  f.fx_pressure = pf.x;
  f.fy_pressure = pf.y;
  f.fx_friction = ff.x;
  f.fy_friction = ff.y;
  f.stations = stations;
  if (out) *out = f;
  int row = 0;
  for (int sideIdx = 0; sideIdx < 2; ++sideIdx) {
    const std::vector<Station>& side = sideIdx == 0 ? upper : lower;
    for (const Station& st : side) {
      if (rows && (row + 1) * 12 <= capacity) {
        double* r = rows + 12 * row;
        r[0] = sideIdx == 0 ? 1.0 : 0.0;
        r[1] = st.face;
        r[2] = st.position.x;
        r[3] = st.position.y;
        r[4] = st.xc;
        r[5] = st.arc;
        r[6] = st.pressure;
        r[7] = st.cp;
        r[8] = st.shear;
        r[9] = st.cf;
        r[10] = st.nearStag ? 1.0 : 0.0;
        r[11] = st.segment;
      }
      ++row;
    }
  }
  return row;
}

}  // extern "C"
