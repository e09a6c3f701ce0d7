// post.cpp - surface distribution and force integration on a converged field,
// host FP64 (reference: src/post/SurfaceData.cpp:27-348, src/post/Forces.cpp:16-130,
// src/post/Polar.cpp:36-77).
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "post.hpp"

namespace cfdgpu {

namespace {

struct WallFace {
  int i{0};  // j-face (i, 0)
  int owner{0};
  Vec2 centre, tangent, normal;
  double len{0.0};
};

struct Station {
  int face{0};
  Vec2 pos, tangent, normal;
  double seg{0.0}, xc{0.0}, p{0.0}, cp{0.0}, tau{0.0}, cf{0.0}, arc{0.0};
  bool nearStag{false};
};

// SurfaceData.cpp:67-98: of the owner's faces, the one pointing most nearly
// away from the wall leads to the second cell. In the structured layout the
// candidate faces are the owner's other three (right, above, left).
int secondCell(const Grid& g, const WallFace& w) {
  const int nI = g.nI, nJ = g.nJ;
  const int i = w.i;
  int best = -1;
  double bestAlign = 0.5;
  // cellFaces order: below (the wall itself), right, above, left (Mesh.cpp:208-213)
  auto consider = [&](Vec2 outward, int across) {
    if (across < 0 || across == w.owner) return;
    const double a = dot(outward, w.normal);
    if (a > bestAlign) {
      bestAlign = a;
      best = across;
    }
  };
  // right: i-face (i+1, 0), owned by this cell
  if (i + 1 < nI) consider(g.iNormal[static_cast<size_t>(i + 1)], i + 1);
  // above: j-face (i, 1), owned by this cell
  if (nJ > 1) consider(g.jNormal[static_cast<size_t>(nI + i)], nI + i);
  // left: i-face (i, 0). The reference asks oppositeCell() for the cell across
  // it, which returns the face's neighbour - the wall cell itself - so the
  // left face is never a candidate (SurfaceData.cpp:80-83).
  return best;
}

Station evaluate(const Grid& g, const double* u, const double* v, const double* p, const double* mu, const WallFace& w,
                 Vec2 tangent, double q, double pref, double chord) {
  Station s;
  s.face = w.i;
  s.pos = w.centre;
  s.tangent = tangent;
  s.normal = w.normal;
  s.seg = w.len;
  s.xc = (chord > 0.0) ? w.centre.x / chord : w.centre.x;
  s.p = p[w.owner];
  s.cp = (q > 0.0) ? (s.p - pref) / q : 0.0;
  const Vec2 off = g.centroid[static_cast<size_t>(w.owner)] - w.centre;
  const double wd = std::abs(dot(off, w.normal));
  const Vec2 cv{u[w.owner], v[w.owner]};
  const Vec2 par = cv - w.normal * dot(cv, w.normal);
  if (wd > 0.0) {
    double grad = dot(par, tangent) / wd;
    const int second = secondCell(g, w);
    if (second >= 0) {
      const Vec2 oo = g.centroid[static_cast<size_t>(second)] - w.centre;
      const double od = std::abs(dot(oo, w.normal));
      if (od > wd * 1.05) {
        const Vec2 ov{u[second], v[second]};
        const Vec2 opar = ov - w.normal * dot(ov, w.normal);
        const double u1 = dot(par, tangent), u2 = dot(opar, tangent);
        const double d1 = wd, d2 = od;
        grad = (u1 * d2 * d2 - u2 * d1 * d1) / (d1 * d2 * (d2 - d1));
      }
    }
    s.tau = mu[w.owner] * grad;
  }
  s.cf = (q > 0.0) ? s.tau / q : 0.0;
  return s;
}

// SurfaceData.cpp:179-204
double separation(const std::vector<Station>& s, int* face) {
  for (size_t i = 1; i < s.size(); ++i) {
    if (s[i].nearStag || s[i - 1].nearStag) continue;
    const double before = s[i - 1].tau, after = s[i].tau;
    if (!(before > 0.0) || !(after < 0.0)) continue;
    const double span = before - after;
    const double fr = (span > 0.0) ? before / span : 0.0;
    if (face) *face = s[i].face;
    return s[i - 1].xc + fr * (s[i].xc - s[i - 1].xc);
  }
  if (face) *face = -1;
  return -1.0;
}

}  // namespace

namespace {

// collectWallFaces (SurfaceData.cpp:27-59): the j = 0 wall faces in i order
std::vector<WallFace> collectWall(const Grid& g) {
  std::vector<WallFace> wall;
  for (int i = 0; i < g.nI; ++i) {
    if (g.bottomType[static_cast<size_t>(i)] != kWallB) continue;
    const Vec2 a = g.nodes[static_cast<size_t>(i)], b = g.nodes[static_cast<size_t>(i + 1)];
    const Vec2 along = b - a;
    const double l = length(along);
    if (!(l > 0.0)) continue;
    WallFace w;
    w.i = i;
    w.owner = i;
    w.centre = g.jCentre[static_cast<size_t>(i)];
    w.tangent = along * (1.0 / l);
    w.normal = g.jNormal[static_cast<size_t>(i)] * -1.0;
    w.len = l;
    wall.push_back(w);
  }
  return wall;
}

// the wall face whose first node lies furthest forward (SurfaceData.cpp:231-249)
size_t leadingEdge(const Grid& g, const std::vector<WallFace>& wall) {
  size_t le = 0;
  double fwd = std::numeric_limits<double>::max();
  for (size_t k = 0; k < wall.size(); ++k) {
    const Vec2 node = g.nodes[static_cast<size_t>(wall[k].i)];
    if (node.x < fwd) {
      fwd = node.x;
      le = k;
    }
  }
  return le;
}

}  // namespace

bool wallTable(const Grid& g, std::vector<WallStation>& out, int& leOut) {
  const std::vector<WallFace> wall = collectWall(g);
  out.clear();
  if (wall.size() < 4) return false;
  const size_t le = leadingEdge(g, wall);
  out.resize(wall.size());
  for (size_t k = 0; k < wall.size(); ++k) {
    const WallFace& w = wall[k];
    WallStation& s = out[k];
    s.cx = w.centre.x;
    s.cy = w.centre.y;
    s.nx = w.normal.x;
    s.ny = w.normal.y;
    s.tx = w.tangent.x;
    s.ty = w.tangent.y;
    s.len = w.len;
    s.face = w.i;
    s.owner = w.owner;
    s.pad = 0;
    // the static part of evaluate(): wall distance and the second cell
    const Vec2 off = g.centroid[static_cast<size_t>(w.owner)] - w.centre;
    s.wd = std::abs(dot(off, w.normal));
    s.second = -1;
    s.od = 0.0;
    if (s.wd > 0.0) {
      const int second = secondCell(g, w);
      if (second >= 0) {
        const Vec2 oo = g.centroid[static_cast<size_t>(second)] - w.centre;
        const double od = std::abs(dot(oo, w.normal));
        if (od > s.wd * 1.05) {
          s.second = second;
          s.od = od;
        }
      }
    }
  }
  // arc lengths in station order, as the arc lambda of surfaceAndForces sums them
  double t = 0.0;
  for (size_t k = le; k < wall.size(); ++k) {
    if (k > le) t += distance(wall[k - 1].centre, wall[k].centre);
    out[k].arc = t;
  }
  t = 0.0;
  for (size_t k = le; k-- > 0;) {
    if (k + 1 < le) t += distance(wall[k + 1].centre, wall[k].centre);
    out[k].arc = t;
  }
  leOut = static_cast<int>(le);
  return true;
}

int surfaceAndForces(const Grid& g, const double* u, const double* v, const double* p, const double* mu, double speed,
                     double alphaDeg, double density, double pref, double chord, double* rows, int capacity,
                     int* count, cfdgpu_forces* out) {
  const std::vector<WallFace> wall = collectWall(g);
  if (wall.size() < 4) return 3;  // NotFound
  const size_t le = leadingEdge(g, wall);
  const double q = 0.5 * density * speed * speed;
  size_t stag = 0;
  double hp = -std::numeric_limits<double>::max();
  for (size_t k = 0; k < wall.size(); ++k) {
    const double pr = p[wall[k].owner];
    if (pr > hp) {
      hp = pr;
      stag = k;
    }
  }
  auto flowTangent = [&](size_t k) {
    if (k == stag) return (k >= le) ? wall[k].tangent : wall[k].tangent * -1.0;
    return (k > stag) ? wall[k].tangent : wall[k].tangent * -1.0;
  };
  std::vector<Station> lower, upper;
  for (size_t k = le; k-- > 0;) {
    lower.push_back(evaluate(g, u, v, p, mu, wall[k], flowTangent(k), q, pref, chord));
    lower.back().nearStag = (k == stag);
  }
  for (size_t k = le; k < wall.size(); ++k) {
    upper.push_back(evaluate(g, u, v, p, mu, wall[k], flowTangent(k), q, pref, chord));
    upper.back().nearStag = (k == stag);
  }
  auto arc = [](std::vector<Station>& s) {
    double t = 0.0;
    for (size_t k = 0; k < s.size(); ++k) {
      if (k > 0) t += distance(s[k - 1].pos, s[k].pos);
      s[k].arc = t;
    }
  };
  arc(lower);
  arc(upper);
  cfdgpu_forces f{};
  f.upper_separation = separation(upper, &f.upper_separation_face);
  f.lower_separation = separation(lower, &f.lower_separation_face);
  f.stagnation_xc = (chord > 0.0) ? wall[stag].centre.x / chord : wall[stag].centre.x;
  f.n_upper = static_cast<int>(upper.size());
  f.n_lower = static_cast<int>(lower.size());
  // Forces.cpp:16-46, 60-130 (midpoint rule, gauge pressure, wind axes)
  Vec2 pf{}, ff{};
  double pm = 0.0, fm = 0.0;
  int stations = 0;
  const Vec2 ref{0.25 * chord, 0.0};
  for (const auto* side : {&upper, &lower}) {
    for (const Station& s : *side) {
      const double ds = s.seg;
      if (!(ds > 0.0)) continue;
      const Vec2 dP = s.normal * (-(s.p - pref) * ds);
      const Vec2 dF = s.tangent * (s.tau * ds);
      pf = pf + dP;
      ff = ff + dF;
      const Vec2 arm = s.pos - ref;
      pm += arm.x * dP.y - arm.y * dP.x;
      fm += arm.x * dF.y - arm.y * dF.x;
      ++stations;
    }
  }
  const double a = alphaDeg * 3.141592653589793238462643383279502884 / 180.0;
  const double ca = std::cos(a), sa = std::sin(a);
  const double pd = pf.x * ca + pf.y * sa, fd = ff.x * ca + ff.y * sa;
  const double pl = pf.y * ca - pf.x * sa, fl = ff.y * ca - ff.x * sa;
  f.drag = pd + fd;
  f.lift = pl + fl;
  f.pitching_moment = -pm - fm;
  const double fs = q * chord, ms = fs * chord;
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
  for (int side = 0; side < 2; ++side) {
    for (const Station& s : (side == 0 ? upper : lower)) {
      if (rows && (row + 1) * 12 <= capacity) {
        double* r = rows + 12 * row;
        r[0] = side == 0 ? 1.0 : 0.0;
        r[1] = s.face;
        r[2] = s.pos.x;
        r[3] = s.pos.y;
        r[4] = s.xc;
        r[5] = s.arc;
        r[6] = s.p;
        r[7] = s.cp;
        r[8] = s.tau;
        r[9] = s.cf;
        r[10] = s.nearStag ? 1.0 : 0.0;
        r[11] = s.seg;
      }
      ++row;
    }
  }
  if (count) *count = row;
  return 0;
}

// Polar.cpp:36-77
int sweepAngles(double start, double end, double step, std::vector<double>& out) {
  if (!std::isfinite(start) || !std::isfinite(end) || !std::isfinite(step)) return 1;
  if (!(step > 0.0)) return 1;
  const double signedStep = (end < start) ? -step : step;
  const double span = (end - start) / signedStep;
  const double rounded = std::round(span);
  const double count = (std::abs(span - rounded) < 1e-9) ? rounded : std::floor(span);
  if (count < 0.0) return 1;
  const double points = count + 1.0;
  if (points > 400.0) return 1;
  out.clear();
  const auto total = static_cast<size_t>(points);
  for (size_t i = 0; i < total; ++i) out.push_back(start + static_cast<double>(i) * signedStep);
  return 0;
}

}  // namespace cfdgpu
