// polar.cpp - the reference's polar sweep (the GUI sequence of Panels.cpp,
// restated as a loop) and its CSV output, with every solve on the device.
#include "polar.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <stdexcept>

#include "engine.hpp"
#include "post.hpp"

namespace cfdgpu {

namespace {
constexpr double kPi = 3.141592653589793238462643383279502884;
}

void defaultFaceConditions(const Grid& g, double speed, double alphaDeg, double pref, int* kind, double* vxy,
                           double* pressure) {
  const double a = alphaDeg * kPi / 180.0;
  const double sx = speed * std::cos(a), sy = speed * std::sin(a);
  for (int f = 0; f < g.faces(); ++f) {
    int k = CFDGPU_K_INTERNAL;
    switch (g.faceType(f)) {
      case kWallB: k = CFDGPU_K_WALL; break;
      case kFarfieldB: k = CFDGPU_K_FARFIELD; break;
      case kOutletB: k = CFDGPU_K_OUTLET; break;
      default: k = CFDGPU_K_INTERNAL; break;
    }
    kind[f] = k;
    pressure[f] = pref;
    vxy[2 * f] = (k == CFDGPU_K_WALL) ? 0.0 : sx;
    vxy[2 * f + 1] = (k == CFDGPU_K_WALL) ? 0.0 : sy;
  }
}

void uniformField(int cells, double speed, double alphaDeg, double density, double reynolds, double chord,
                  double pref, double* u, double* v, double* p, double* rho, double* mu) {
  const double a = alphaDeg * kPi / 180.0;
  const double ux = speed * std::cos(a), uy = speed * std::sin(a);
  const double visc = density * speed * chord / reynolds;
  for (int c = 0; c < cells; ++c) {
    u[c] = ux;
    v[c] = uy;
    p[c] = pref;
    rho[c] = density;
    mu[c] = visc;
  }
}

// One point: requestAngle (Panels.cpp:1581-1597) -> updateFlow (:1137-1157:
// the uniform field of the new freestream, with velocity and pressure carried
// over when continuing) -> updateSolver (:1244-1318: conditions, create, the
// model attached and initialised cold) -> the worker's stop rule -> surface
// and forces -> the recorded point (:1737-1762).
void polarSweep(const PolarRequest& rq, std::vector<cfdgpu_polar_point>& out) {
  if (!rq.grid) throw std::invalid_argument("polar sweep without a mesh");
  const Grid& g = *rq.grid;
  const int N = g.cells(), F = g.faces();
  std::vector<int> kind(F);
  std::vector<double> vxy(2 * static_cast<size_t>(F)), pres(F);
  std::vector<double> u(N), v(N), p(N), rho(N), mu(N), muOut(N);
  std::vector<double> prevU, prevV, prevP;
  std::vector<double> rows(12 * static_cast<size_t>(g.nI));
  out.clear();
  bool nextCold = false;
  for (double alpha : rq.alphas) {
    bool retried = false;
    bool cold = !rq.continued || nextCold || prevU.empty();
    cfdgpu_solve_report rep{};
    for (;;) {
      defaultFaceConditions(g, rq.speed, alpha, rq.referencePressure, kind.data(), vxy.data(), pres.data());
      uniformField(N, rq.speed, alpha, rq.density, rq.reynolds, rq.chord, rq.referencePressure, u.data(), v.data(),
                   p.data(), rho.data(), mu.data());
      if (!cold) {
        u = prevU;
        v = prevV;
        p = prevP;
      }
      std::vector<Vec2> vel(static_cast<size_t>(F));
      for (int f = 0; f < F; ++f) vel[f] = Vec2{vxy[2 * f], vxy[2 * f + 1]};
      Engine e(g, kind, vel, pres, rq.settings);
      if (rq.model != CFDGPU_MODEL_NONE && e.setModel(rq.model, rq.sst) != CFDGPU_OK)
        throw std::invalid_argument("unknown turbulence model");
      if (e.initialise(u.data(), v.data(), p.data(), rho.data(), mu.data()) != CFDGPU_OK)
        throw std::runtime_error("the solver could not start for this angle");
      if (e.solve(rq.rule, &rep) != CFDGPU_OK) throw std::runtime_error("solve failed");
      e.getField(u.data(), v.data(), p.data(), muOut.data());
      // a blow-up gets one retry from the undisturbed stream (Polar.cpp:112-117)
      if (rep.status == 2 && !retried) {
        retried = true;
        cold = true;
        continue;
      }
      break;
    }
    cfdgpu_polar_point pt{};
    pt.alpha_deg = alpha;
    pt.status = rep.status;
    pt.iterations = rep.iterations;
    pt.continuity = rep.last.continuity;
    pt.retried_cold = retried ? 1 : 0;
    pt.upper_separation = -1.0;
    pt.lower_separation = -1.0;
    cfdgpu_forces f{};
    int count = 0;
    if (surfaceAndForces(g, u.data(), v.data(), p.data(), muOut.data(), rq.speed, alpha, rq.density,
                         rq.referencePressure, rq.chord, rows.data(), g.nI, &count, &f) == 0) {
      pt.cl = f.cl;
      pt.cd = f.cd;
      pt.cd_pressure = f.cd_pressure;
      pt.cd_friction = f.cd_friction;
      pt.cm = f.cm;
      pt.l_over_d = (std::isfinite(f.cd) && f.cd > 0.0) ? f.cl / f.cd : 0.0;  // Forces.cpp:50-56
      pt.upper_separation = f.upper_separation;
      pt.lower_separation = f.lower_separation;
    }
    out.push_back(pt);
    prevU = u;
    prevV = v;
    prevP = p;
    // after a retry that diverged again, the next angle starts cold (Panels.cpp:1797-1799)
    nextCold = retried && rep.status == 2;
  }
}

namespace {

void append(std::string& s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void append(std::string& s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  const int n = std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (n > 0) s.append(buf, static_cast<size_t>(n < static_cast<int>(sizeof(buf)) ? n : sizeof(buf) - 1));
}

// {:.6g} of std::format: shortest of %e / %f at six significant digits, with
// trailing zeros removed (the general format without '#').
std::string general6(double x) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.6g", x);
  return buf;
}

}  // namespace

std::string polarCsv(const cfdgpu_polar_point* pts, int count, const cfdgpu_polar_header& h) {
  std::string out;
  out += "# cfd_simulator aerodynamic polar\n";
  append(out, "# section,%s\n", (h.section && *h.section) ? h.section : "unknown");
  append(out, "# mesh,%s\n", (h.mesh && *h.mesh) ? h.mesh : "unknown");
  out += "# reynolds," + general6(h.reynolds) + "\n";
  out += "# freestream_speed_mps," + general6(h.speed) + "\n";
  out += "# chord_m," + general6(h.chord) + "\n";
  out += "# moment_reference_xc," + general6(h.moment_reference) + "\n";
  append(out, "# continued_between_points,%s\n", h.continued ? "yes" : "no");
  out += "# every row is a separate Navier-Stokes solve; nothing is interpolated\n";
  out += "# status is converged, iteration-limit or diverged\n";
  out += "alpha_deg,cl,cd,cd_pressure,cd_friction,cm,l_over_d,"
         "separation_upper_xc,separation_lower_xc,status,iterations,continuity_residual\n";
  for (int q = 0; q < count; ++q) {
    const cfdgpu_polar_point& p = pts[q];
    char up[64] = "", lo[64] = "";
    if (p.upper_separation >= 0.0) std::snprintf(up, sizeof(up), "%.6f", p.upper_separation);
    if (p.lower_separation >= 0.0) std::snprintf(lo, sizeof(lo), "%.6f", p.lower_separation);
    const char* status = p.status == 2 ? "diverged" : (p.status == 0 ? "converged" : "iteration-limit");
    append(out, "%.6f,%.8f,%.8f,%.8f,%.8f,%.8f,%.6f,%s,%s,%s,%d,%.6e\n", p.alpha_deg, p.cl, p.cd, p.cd_pressure,
           p.cd_friction, p.cm, p.l_over_d, up, lo, status, p.iterations, p.continuity);
  }
  return out;
}

}  // namespace cfdgpu
