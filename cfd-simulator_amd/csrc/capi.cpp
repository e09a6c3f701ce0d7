// capi.cpp - the extern "C" boundary (include/cfdgpu.h). No exception crosses
// it: every entry point catches and maps to the reference's ErrorCode values
// (include/cfd/core/Error.hpp:37-46).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cfdgpu.h"
#include "host/engine.hpp"
#include "host/polar.hpp"
#include "host/post.hpp"
#include "host/serve.hpp"

namespace {

thread_local std::string gLastError;

int fail(int code, const std::string& msg) {
  gLastError = msg;
  return code;
}

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::invalid_argument& e) {
    return fail(CFDGPU_INVALID_ARGUMENT, e.what());
  } catch (const std::bad_alloc& e) {
    return fail(CFDGPU_INITIALIZATION_FAILURE, "out of memory");
  } catch (const std::exception& e) {
    return fail(CFDGPU_INTERNAL, e.what());
  } catch (...) {
    return fail(CFDGPU_INTERNAL, "unknown failure");
  }
}

// Device-path options the library runs.
int checkPathOptions(const cfdgpu_settings& s) {
  if (s.precision != CFDGPU_PRECISION_FP64 && s.precision != CFDGPU_PRECISION_MIXED)
    return fail(CFDGPU_INVALID_ARGUMENT, "unknown precision");
  if (s.gs_ordering != CFDGPU_GS_LEXICOGRAPHIC && s.gs_ordering != CFDGPU_GS_REDBLACK)
    return fail(CFDGPU_INVALID_ARGUMENT, "unknown Gauss-Seidel ordering");
  if (s.pressure_solver != CFDGPU_PSOLVE_JACOBI_CG && s.pressure_solver != CFDGPU_PSOLVE_LINE_CG)
    return fail(CFDGPU_INVALID_ARGUMENT, "unknown pressure solver");
  if (s.refinement_passes < 1 || s.refinement_passes > 8)
    return fail(CFDGPU_INVALID_ARGUMENT, "refinement_passes must be 1..8");
  return CFDGPU_OK;
}

cfdgpu::CGridOptions toOptions(const cfdgpu_cgrid_options& o) {
  cfdgpu::CGridOptions r;
  r.upstream = o.upstream_chords;
  r.downstream = o.downstream_chords;
  r.vertical = o.vertical_chords;
  r.surfacePoints = o.surface_points;
  r.wakePoints = o.wake_points;
  r.normalPoints = o.normal_points;
  r.firstLayer = o.first_layer_height;
  r.blendLength = o.normal_blend_length;
  r.smoothingPasses = o.normal_smoothing_passes;
  return r;
}

}  // namespace

extern "C" {

const char* cfdgpu_version(void) { return "cfdgpu 0.1.0 (gfx950)"; }
const char* cfdgpu_last_error(void) { return gLastError.c_str(); }

int cfdgpu_device_count(int* count) {
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (count) *count = (e == hipSuccess) ? n : 0;
  return (e == hipSuccess) ? CFDGPU_OK : fail(CFDGPU_INITIALIZATION_FAILURE, hipGetErrorString(e));
}

int cfdgpu_cgrid_defaults(int resolution, cfdgpu_cgrid_options* out) {
  if (!out || resolution < 0 || resolution > 2) return fail(CFDGPU_INVALID_ARGUMENT, "resolution must be 0, 1 or 2");
  const cfdgpu::CGridOptions o = cfdgpu::presetOptions(resolution);
  out->upstream_chords = 12.0;
  out->downstream_chords = 25.0;
  out->vertical_chords = 12.0;
  out->surface_points = o.surfacePoints;
  out->wake_points = o.wakePoints;
  out->normal_points = o.normalPoints;
  out->first_layer_height = o.firstLayer;
  out->normal_blend_length = 0.06;
  out->normal_smoothing_passes = 6;
  return CFDGPU_OK;
}

int cfdgpu_mesh_cgrid(int m, int p, int t, double chord, const cfdgpu_cgrid_options* options, cfdgpu_mesh** out) {
  if (!options || !out) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  return guarded([&]() -> int {
    auto* mesh = new cfdgpu_mesh();
    std::string err;
    if (!cfdgpu::makeCGrid(mesh->grid, m, p, t, chord, toOptions(*options), err)) {
      delete mesh;
      return fail(CFDGPU_INVALID_ARGUMENT, err);
    }
    *out = mesh;
    return CFDGPU_OK;
  });
}

int cfdgpu_mesh_box(const cfdgpu_box_options* o, cfdgpu_mesh** out) {
  if (!o || !out) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  return guarded([&]() -> int {
    cfdgpu::BoxOptions b;
    b.originX = o->origin_x;
    b.originY = o->origin_y;
    b.length = o->length;
    b.height = o->height;
    b.cellsX = o->cells_x;
    b.cellsY = o->cells_y;
    b.grading = o->grading;
    b.firstCellHeight = o->first_cell_height;
    b.left = o->left;
    b.right = o->right;
    b.lower = o->lower;
    b.upper = o->upper;
    auto* mesh = new cfdgpu_mesh();
    std::string err;
    if (!cfdgpu::makeBox(mesh->grid, b, err)) {
      delete mesh;
      return fail(CFDGPU_INVALID_ARGUMENT, err);
    }
    *out = mesh;
    return CFDGPU_OK;
  });
}

int cfdgpu_mesh_structured(int ni, int nj, const double* xy, const int* jmin, const int* partner, int jmax, int imin,
                           int imax, cfdgpu_mesh** out) {
  if (!xy || !jmin || !out) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  return guarded([&]() -> int {
    if (ni < 2 || nj < 2) return fail(CFDGPU_INVALID_ARGUMENT, "a structured block needs at least 2x2 nodes");
    std::vector<cfdgpu::Vec2> nodes(static_cast<size_t>(ni) * static_cast<size_t>(nj));
    for (size_t q = 0; q < nodes.size(); ++q) nodes[q] = cfdgpu::Vec2{xy[2 * q], xy[2 * q + 1]};
    std::vector<int> bottom(jmin, jmin + (ni - 1));
    std::vector<int> part;
    if (partner) part.assign(partner, partner + (ni - 1));
    auto* mesh = new cfdgpu_mesh();
    std::string err;
    if (!cfdgpu::buildGrid(mesh->grid, ni, nj, std::move(nodes), std::move(bottom), std::move(part), jmax, imin, imax,
                           err)) {
      delete mesh;
      return fail(CFDGPU_INVALID_ARGUMENT, err);
    }
    *out = mesh;
    return CFDGPU_OK;
  });
}

void cfdgpu_mesh_destroy(cfdgpu_mesh* mesh) { delete mesh; }

int cfdgpu_mesh_dims(const cfdgpu_mesh* m, int* ni, int* nj, int* cells, int* faces) {
  if (!m) return fail(CFDGPU_INVALID_ARGUMENT, "null mesh");
  if (ni) *ni = m->grid.nodesI;
  if (nj) *nj = m->grid.nodesJ;
  if (cells) *cells = m->grid.cells();
  if (faces) *faces = m->grid.faces();
  return CFDGPU_OK;
}

int cfdgpu_mesh_nodes(const cfdgpu_mesh* m, double* xy) {
  if (!m || !xy) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  for (size_t q = 0; q < m->grid.nodes.size(); ++q) {
    xy[2 * q] = m->grid.nodes[q].x;
    xy[2 * q + 1] = m->grid.nodes[q].y;
  }
  return CFDGPU_OK;
}

int cfdgpu_mesh_cells(const cfdgpu_mesh* m, double* cxy, double* area) {
  if (!m) return fail(CFDGPU_INVALID_ARGUMENT, "null mesh");
  for (size_t c = 0; c < m->grid.area.size(); ++c) {
    if (cxy) {
      cxy[2 * c] = m->grid.centroid[c].x;
      cxy[2 * c + 1] = m->grid.centroid[c].y;
    }
    if (area) area[c] = m->grid.area[c];
  }
  return CFDGPU_OK;
}

// faces in the reference's numbering (Mesh.cpp:138-204)
int cfdgpu_mesh_faces(const cfdgpu_mesh* m, double* centre, double* normal, double* area, int* owner, int* neighbour,
                      int* boundary, int* partner) {
  if (!m) return fail(CFDGPU_INVALID_ARGUMENT, "null mesh");
  const cfdgpu::Grid& g = m->grid;
  const int nI = g.nI, nJ = g.nJ;
  auto put = [&](int f, cfdgpu::Vec2 c, cfdgpu::Vec2 n, double a, int o, int nb, int b, int pt) {
    if (centre) { centre[2 * f] = c.x; centre[2 * f + 1] = c.y; }
    if (normal) { normal[2 * f] = n.x; normal[2 * f + 1] = n.y; }
    if (area) area[f] = a;
    if (owner) owner[f] = o;
    if (neighbour) neighbour[f] = nb;
    if (boundary) boundary[f] = b;
    if (partner) partner[f] = pt;
  };
  for (int j = 0; j <= nJ; ++j)
    for (int i = 0; i < nI; ++i) {
      const size_t f = static_cast<size_t>(j * nI + i);
      int o, nb = -1, b = cfdgpu::kInterior, pt = -1;
      if (j == 0) {
        o = i;
        b = g.bottomType[static_cast<size_t>(i)];
        pt = g.bottomPartner[static_cast<size_t>(i)];
      } else if (j == nJ) {
        o = (nJ - 1) * nI + i;
        b = g.topType;
      } else {
        o = (j - 1) * nI + i;
        nb = j * nI + i;
      }
      put(g.refJFace(i, j), g.jCentre[f], g.jNormal[f], g.jLen[f], o, nb, b, pt);
    }
  for (int j = 0; j < nJ; ++j)
    for (int i = 0; i <= nI; ++i) {
      const size_t f = static_cast<size_t>(j * (nI + 1) + i);
      int o, nb = -1, b = cfdgpu::kInterior;
      if (i == 0) {
        o = j * nI;
        b = g.leftType;
      } else if (i == nI) {
        o = j * nI + nI - 1;
        b = g.rightType;
      } else {
        o = j * nI + i - 1;
        nb = j * nI + i;
      }
      put(g.refIFace(i, j), g.iCentre[f], g.iNormal[f], g.iLen[f], o, nb, b, -1);
    }
  return CFDGPU_OK;
}

int cfdgpu_mesh_quality(const cfdgpu_mesh* m, double* out7) {
  if (!m || !out7) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  const cfdgpu::Grid& g = m->grid;
  out7[0] = g.minArea;
  out7[1] = g.maxArea;
  out7[2] = g.maxAspect;
  out7[3] = g.maxNonOrthoDeg;
  out7[4] = g.inverted;
  out7[5] = g.minWallSpacing;
  out7[6] = g.totalArea;
  return CFDGPU_OK;
}

// buildFaceConditions with the default BoundaryConditions (BoundaryConditions.cpp:40-97)
int cfdgpu_face_conditions(const cfdgpu_mesh* m, double speed, double alpha, double pref, int* kind, double* vxy,
                           double* pressure) {
  if (!m || !kind || !vxy || !pressure) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  if (!std::isfinite(speed) || speed <= 0.0) return fail(CFDGPU_INVALID_ARGUMENT, "freestream speed must be positive");
  if (!std::isfinite(alpha) || std::abs(alpha) > 90.0) return fail(CFDGPU_INVALID_ARGUMENT, "angle of attack must lie within +/-90 degrees");
  cfdgpu::defaultFaceConditions(m->grid, speed, alpha, pref, kind, vxy, pressure);
  return CFDGPU_OK;
}

// FlowField::uniform (FlowField.cpp:21-47) with Freestream.cpp:9-40
int cfdgpu_uniform_field(int cells, double speed, double alpha, double density, double reynolds, double chord,
                         double pref, double* u, double* v, double* p, double* rho, double* mu) {
  if (cells <= 0) return fail(CFDGPU_INVALID_ARGUMENT, "cannot initialise a field with no cells");
  if (!std::isfinite(speed) || speed <= 0.0 || !std::isfinite(density) || density <= 0.0 ||
      !std::isfinite(reynolds) || reynolds <= 0.0 || !std::isfinite(chord) || chord <= 0.0)
    return fail(CFDGPU_INVALID_ARGUMENT, "freestream must be positive and finite");
  const double a = alpha * 3.141592653589793238462643383279502884 / 180.0;
  const double ux = speed * std::cos(a), uy = speed * std::sin(a);
  const double visc = density * speed * chord / reynolds;
  for (int c = 0; c < cells; ++c) {
    u[c] = ux;
    v[c] = uy;
    p[c] = pref;
    rho[c] = density;
    mu[c] = visc;
  }
  return CFDGPU_OK;
}

// SimpleSettings defaults (SimpleSolver.hpp:96-131), plain FP64 path
void cfdgpu_default_settings(cfdgpu_settings* s) {
  if (!s) return;
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
  s->precision = CFDGPU_PRECISION_FP64;
  s->gs_ordering = CFDGPU_GS_LEXICOGRAPHIC;
  s->pressure_solver = CFDGPU_PSOLVE_JACOBI_CG;
  s->refinement_passes = 1;
}

// SSTConstants (KOmegaSST.hpp:147-171)
void cfdgpu_default_sst(cfdgpu_sst_constants* c) {
  if (!c) return;
  c->sigma_k1 = 0.85; c->sigma_w1 = 0.5; c->beta1 = 0.075;
  c->sigma_k2 = 1.0; c->sigma_w2 = 0.856; c->beta2 = 0.0828;
  c->beta_star = 0.09; c->kappa = 0.41; c->a1 = 0.31; c->production_limit = 10.0;
}

int cfdgpu_create(cfdgpu_solver** out, const cfdgpu_mesh* mesh, const int* kind, const double* vxy,
                  const double* pressure, const cfdgpu_settings* settings) {
  if (!out || !mesh || !kind || !vxy || !pressure || !settings) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  if (!(settings->velocity_relaxation > 0.0) || settings->velocity_relaxation > 1.0 ||
      !(settings->pressure_relaxation > 0.0) || settings->pressure_relaxation > 1.0)
    return fail(CFDGPU_INVALID_ARGUMENT, "relaxation factors must lie in (0, 1]");
  if (const int rc = checkPathOptions(*settings); rc != CFDGPU_OK) return rc;
  return guarded([&]() -> int {
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0)
      return fail(CFDGPU_INITIALIZATION_FAILURE, "no HIP device: the MI355X path needs a GPU");
    const int nf = mesh->grid.faces();
    std::vector<int> k(kind, kind + nf);
    std::vector<cfdgpu::Vec2> vel(static_cast<size_t>(nf));
    for (int f = 0; f < nf; ++f) vel[static_cast<size_t>(f)] = cfdgpu::Vec2{vxy[2 * f], vxy[2 * f + 1]};
    std::vector<double> pr(pressure, pressure + nf);
    auto* s = new cfdgpu_solver();
    s->engine = std::make_unique<cfdgpu::Engine>(mesh->grid, std::move(k), std::move(vel), std::move(pr), *settings);
    *out = s;
    return CFDGPU_OK;
  });
}

void cfdgpu_destroy(cfdgpu_solver* s) { delete s; }

#define SOLVER_GUARD(s) \
  if (!(s) || !(s)->engine) return fail(CFDGPU_INVALID_ARGUMENT, "null solver");

int cfdgpu_set_turbulence(cfdgpu_solver* s, int model, const cfdgpu_sst_constants* c) {
  SOLVER_GUARD(s)
  return guarded([&]() -> int { return s->engine->setModel(model, c); });
}

int cfdgpu_initialise(cfdgpu_solver* s, const double* u, const double* v, const double* p, const double* rho,
                      const double* mu) {
  SOLVER_GUARD(s)
  if (!u || !v || !p || !rho || !mu) return fail(CFDGPU_INVALID_ARGUMENT, "null field");
  return guarded([&]() -> int {
    const int rc = s->engine->initialise(u, v, p, rho, mu);
    if (rc != CFDGPU_OK) return fail(rc, "initialise failed (turbulence inflow or starting field invalid)");
    return rc;
  });
}

int cfdgpu_iterate(cfdgpu_solver* s, int count, cfdgpu_monitor* out) {
  SOLVER_GUARD(s)
  if (count < 0) return fail(CFDGPU_INVALID_ARGUMENT, "negative iteration count");
  return guarded([&]() -> int { return s->engine->iterate(count, out); });
}

int cfdgpu_iterate_timed(cfdgpu_solver* s, int count, cfdgpu_monitor* out, double* stage_ms, int* stage_calls) {
  SOLVER_GUARD(s)
  if (count < 0) return fail(CFDGPU_INVALID_ARGUMENT, "negative iteration count");
  return guarded([&]() -> int { return s->engine->iterateTimed(count, out, stage_ms, stage_calls); });
}

int cfdgpu_solve(cfdgpu_solver* s, const cfdgpu_stop_rule* rule, cfdgpu_solve_report* out) {
  SOLVER_GUARD(s)
  if (!rule) return fail(CFDGPU_INVALID_ARGUMENT, "null stop rule");
  return guarded([&]() -> int { return s->engine->solve(*rule, out); });
}

int cfdgpu_batch_create(cfdgpu_batch** out, cfdgpu_solver* const* solvers, int count) {
  if (!out || !solvers || count <= 0) return fail(CFDGPU_INVALID_ARGUMENT, "null argument or empty batch");
  return guarded([&]() -> int {
    std::vector<cfdgpu::Engine*> members;
    members.reserve(static_cast<size_t>(count));
    for (int q = 0; q < count; ++q) {
      if (!solvers[q] || !solvers[q]->engine) return fail(CFDGPU_INVALID_ARGUMENT, "null solver in batch");
      members.push_back(solvers[q]->engine.get());
    }
    auto* b = new cfdgpu_batch();
    try {
      b->batch = std::make_unique<cfdgpu::Batch>(std::move(members));
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
    return CFDGPU_OK;
  });
}

void cfdgpu_batch_destroy(cfdgpu_batch* b) { delete b; }

int cfdgpu_batch_iterate(cfdgpu_batch* b, int count, cfdgpu_monitor* out) {
  if (!b || !b->batch) return fail(CFDGPU_INVALID_ARGUMENT, "null batch");
  if (count < 0) return fail(CFDGPU_INVALID_ARGUMENT, "negative iteration count");
  return guarded([&]() -> int { return b->batch->iterate(count, out); });
}

int cfdgpu_batch_iterate_timed(cfdgpu_batch* b, int count, cfdgpu_monitor* out, double* stage_ms,
                               int* stage_calls) {
  if (!b || !b->batch) return fail(CFDGPU_INVALID_ARGUMENT, "null batch");
  if (count < 0) return fail(CFDGPU_INVALID_ARGUMENT, "negative iteration count");
  return guarded([&]() -> int { return b->batch->iterateTimed(count, out, stage_ms, stage_calls); });
}

int cfdgpu_batch_solve(cfdgpu_batch* b, const cfdgpu_stop_rule* rule, cfdgpu_solve_report* out) {
  if (!b || !b->batch) return fail(CFDGPU_INVALID_ARGUMENT, "null batch");
  if (!rule) return fail(CFDGPU_INVALID_ARGUMENT, "null stop rule");
  return guarded([&]() -> int { return b->batch->solve(*rule, out); });
}

int cfdgpu_get_settings(const cfdgpu_solver* s, cfdgpu_settings* out) {
  SOLVER_GUARD(s)
  if (!out) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  *out = s->engine->settings();
  return CFDGPU_OK;
}

int cfdgpu_set_settings(cfdgpu_solver* s, const cfdgpu_settings* settings) {
  SOLVER_GUARD(s)
  if (!settings) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  if (const int rc = checkPathOptions(*settings); rc != CFDGPU_OK) return rc;
  return guarded([&]() -> int {
    s->engine->setSettings(*settings);
    return CFDGPU_OK;
  });
}

int cfdgpu_get_field(const cfdgpu_solver* s, double* u, double* v, double* p, double* mu) {
  SOLVER_GUARD(s)
  return guarded([&]() -> int { return s->engine->getField(u, v, p, mu); });
}
int cfdgpu_set_field(cfdgpu_solver* s, const double* u, const double* v, const double* p) {
  SOLVER_GUARD(s)
  return guarded([&]() -> int { return s->engine->setField(u, v, p); });
}
int cfdgpu_get_massflux(const cfdgpu_solver* s, double* flux) {
  SOLVER_GUARD(s)
  if (!flux) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  return guarded([&]() -> int { return s->engine->getMassFlux(flux); });
}
int cfdgpu_set_massflux(cfdgpu_solver* s, const double* flux) {
  SOLVER_GUARD(s)
  if (!flux) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  return guarded([&]() -> int { return s->engine->setMassFlux(flux); });
}
int cfdgpu_get_turbulence(const cfdgpu_solver* s, double* k, double* w, double* mut, double* wd) {
  SOLVER_GUARD(s)
  return guarded([&]() -> int {
    const int rc = s->engine->getTurbulence(k, w, mut, wd);
    return rc == CFDGPU_OK ? rc : fail(rc, "no k-omega SST model attached");
  });
}
int cfdgpu_set_turbulence_fields(cfdgpu_solver* s, const double* k, const double* w, const double* mut) {
  SOLVER_GUARD(s)
  return guarded([&]() -> int {
    const int rc = s->engine->setTurbulence(k, w, mut);
    return rc == CFDGPU_OK ? rc : fail(rc, "no k-omega SST model attached");
  });
}
int cfdgpu_turbulence_info(const cfdgpu_solver* s, double* kIn, double* wIn, double* yplus) {
  SOLVER_GUARD(s)
  return guarded([&]() -> int { return s->engine->turbulenceInfo(kIn, wIn, yplus); });
}
int cfdgpu_divergence(const cfdgpu_solver* s, double* out) {
  SOLVER_GUARD(s)
  if (!out) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  return guarded([&]() -> int { return s->engine->divergence(out); });
}

int cfdgpu_divergence_noalloc(const cfdgpu_solver* s, double* out) {
  SOLVER_GUARD(s)
  if (!out) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  return guarded([&]() -> int { return s->engine->divergenceNoAlloc(out); });
}

int cfdgpu_set_launch_tuning(int cells, int rows) {
  if (cells != 0 && cells != 1 && cells != 2 && cells != 4)
    return fail(CFDGPU_INVALID_ARGUMENT, "cells per thread must be 0 (LDS tiles), 1, 2 or 4");
  if (rows != 16 && rows != 32) return fail(CFDGPU_INVALID_ARGUMENT, "column rows must be 16 or 32");
  cfdgpu::LaunchTuning t;
  t.lcgCells = cells;
  t.colUnroll = rows;
  cfdgpu::setLaunchTuning(t);
  return CFDGPU_OK;
}

int cfdgpu_get_launch_tuning(int* cells, int* rows) {
  if (!cells || !rows) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  const cfdgpu::LaunchTuning t = cfdgpu::launchTuning();
  *cells = t.lcgCells;
  *rows = t.colUnroll;
  return CFDGPU_OK;
}

int cfdgpu_surface(const cfdgpu_mesh* m, const double* u, const double* v, const double* p, const double* mu,
                   double speed, double alpha, double density, double pref, double chord, double* rows, int capacity,
                   int* count, cfdgpu_forces* forces) {
  if (!m || !u || !v || !p || !mu) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  const double q = 0.5 * density * speed * speed;
  if (!(q > 0.0)) return fail(CFDGPU_INVALID_ARGUMENT, "the freestream has no dynamic pressure to form coefficients with");
  return guarded([&]() -> int {
    const int rc = cfdgpu::surfaceAndForces(m->grid, u, v, p, mu, speed, alpha, density, pref, chord, rows, capacity,
                                            count, forces);
    return rc == 0 ? CFDGPU_OK : fail(rc, "the mesh has no aerofoil surface to extract");
  });
}

int cfdgpu_solver_surface(cfdgpu_solver* s, double speed, double alpha_deg, double density, double pref, double chord,
                          double* rows, int capacity, int* count, cfdgpu_forces* forces) {
  if (!s) return fail(CFDGPU_INVALID_ARGUMENT, "null solver");
  if (capacity < 0 || (capacity > 0 && !rows)) return fail(CFDGPU_INVALID_ARGUMENT, "bad row buffer");
  return guarded([&]() -> int {
    const int rc = s->engine->surface(speed, alpha_deg, density, pref, chord, rows, capacity, count, forces);
    if (rc == CFDGPU_NOT_FOUND) return fail(rc, "the mesh has fewer than four wall faces");
    if (rc == CFDGPU_INITIALIZATION_FAILURE) return fail(rc, "the solver is not initialised");
    return rc;
  });
}

int cfdgpu_polar_sweep(const cfdgpu_mesh* m, const cfdgpu_settings* settings, int model,
                       const cfdgpu_sst_constants* sst, double speed, double density, double reynolds, double chord,
                       double pref, const double* alphas, int count, int continued, const cfdgpu_stop_rule* rule,
                       cfdgpu_polar_point* out) {
  if (!m || !settings || !alphas || !rule || !out || count < 0) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  if (!(speed > 0.0) || !(density > 0.0) || !(reynolds > 0.0) || !(chord > 0.0))
    return fail(CFDGPU_INVALID_ARGUMENT, "freestream must be positive and finite");
  if (model < CFDGPU_MODEL_NONE || model > CFDGPU_MODEL_SST) return fail(CFDGPU_INVALID_ARGUMENT, "unknown model");
  if (const int rc = checkPathOptions(*settings); rc != CFDGPU_OK) return rc;
  return guarded([&]() -> int {
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0)
      return fail(CFDGPU_INITIALIZATION_FAILURE, "no HIP device: the MI355X path needs a GPU");
    cfdgpu::PolarRequest rq;
    rq.grid = &m->grid;
    rq.settings = *settings;
    rq.model = model;
    rq.sst = sst;
    rq.speed = speed;
    rq.density = density;
    rq.reynolds = reynolds;
    rq.chord = chord;
    rq.referencePressure = pref;
    rq.alphas.assign(alphas, alphas + count);
    rq.continued = continued != 0;
    rq.rule = *rule;
    std::vector<cfdgpu_polar_point> pts;
    cfdgpu::polarSweep(rq, pts);
    for (size_t q = 0; q < pts.size(); ++q) out[q] = pts[q];
    return CFDGPU_OK;
  });
}

void cfdgpu_server_defaults(cfdgpu_server_options* o) {
  if (!o) return;
  *o = cfdgpu_server_options{};
  o->hbm_budget_bytes = 200e9;
  o->max_slots = 64;
  o->chunk_iterations = 20;
  o->rule = cfdgpu_stop_rule{1e-6, 10000, 1};
  cfdgpu_default_settings(&o->settings);
  o->model = CFDGPU_MODEL_SST;
  o->warm_start = 1;
  o->keep_fields = 0;
  o->settle_window = 0;
  o->settle_cl = 5e-4;
  o->settle_cd = 5e-5;
}

int cfdgpu_server_create(const cfdgpu_server_options* o, cfdgpu_server** out) {
  if (!o || !out) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  if (const int rc = checkPathOptions(o->settings); rc != CFDGPU_OK) return rc;
  return guarded([&]() -> int {
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0)
      return fail(CFDGPU_INITIALIZATION_FAILURE, "no HIP device: the MI355X path needs a GPU");
    cfdgpu::ServerOptions so;
    so.budgetBytes = o->hbm_budget_bytes;
    so.maxSlots = o->max_slots;
    so.chunk = o->chunk_iterations;
    so.rule = o->rule;
    so.settings = o->settings;
    so.model = o->model;
    so.warmStart = o->warm_start != 0;
    so.keepFields = o->keep_fields != 0;
    if (o->settle_window < 0 || !(o->settle_cl >= 0.0) || !(o->settle_cd >= 0.0))
      return fail(CFDGPU_INVALID_ARGUMENT, "settle_window, settle_cl and settle_cd must not be negative");
    so.settleWindow = o->settle_window;
    so.settleCl = o->settle_cl;
    so.settleCd = o->settle_cd;
    auto* s = new cfdgpu_server();
    try {
      s->server = std::make_unique<cfdgpu::Server>(so);
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
    return CFDGPU_OK;
  });
}

void cfdgpu_server_destroy(cfdgpu_server* s) { delete s; }

int cfdgpu_server_submit(cfdgpu_server* s, const cfdgpu_serve_request* r, int* ticket) {
  if (!s || !r) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  return guarded([&]() -> int {
    cfdgpu::ServeRequest q;
    q.m = r->m;
    q.p = r->p;
    q.t = r->t;
    q.chord = r->chord;
    q.mesh = toOptions(r->mesh);
    q.alphaDeg = r->alpha_deg;
    q.reynolds = r->reynolds;
    q.speed = r->speed;
    q.density = r->density;
    q.referencePressure = r->reference_pressure;
    q.arrival = r->arrival_s;
    if (r->settings) {
      if (const int rc = checkPathOptions(*r->settings); rc != CFDGPU_OK) return rc;
      q.ownSettings = true;
      q.settings = *r->settings;
    }
    const int t = s->server->submit(q);
    if (ticket) *ticket = t;
    return CFDGPU_OK;
  });
}

int cfdgpu_server_step(cfdgpu_server* s, cfdgpu_server_stats* out) {
  if (!s) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  return guarded([&]() -> int {
    s->server->step();
    if (out) s->server->stats(out);
    return CFDGPU_OK;
  });
}

int cfdgpu_server_run(cfdgpu_server* s, double wall_limit_s, cfdgpu_server_stats* out) {
  if (!s) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  return guarded([&]() -> int {
    s->server->run(wall_limit_s);
    if (out) s->server->stats(out);
    return CFDGPU_OK;
  });
}

int cfdgpu_server_collect(cfdgpu_server* s, cfdgpu_serve_result* out, int capacity, int* count) {
  if (!s || !count || (capacity > 0 && !out) || capacity < 0) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  return guarded([&]() -> int {
    *count = s->server->collect(out, capacity);
    return CFDGPU_OK;
  });
}

int cfdgpu_server_fields(const cfdgpu_server* s, int ticket, double* u, double* v, double* p, double* mu) {
  if (!s) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  return guarded([&]() -> int {
    if (!s->server->fields(ticket, u, v, p, mu))
      return fail(CFDGPU_NOT_FOUND, "no kept fields for this ticket (keep_fields off, unfinished or unknown)");
    return CFDGPU_OK;
  });
}

int cfdgpu_polar_csv(const cfdgpu_polar_point* points, int count, const cfdgpu_polar_header* header, char* buf,
                     int capacity) {
  if ((!points && count > 0) || !header || count < 0) return -fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  const std::string text = cfdgpu::polarCsv(points, count, *header);
  if (buf && capacity > static_cast<int>(text.size())) std::memcpy(buf, text.c_str(), text.size() + 1);
  return static_cast<int>(text.size());
}

int cfdgpu_sweep_angles(double start, double end, double step, double* out, int capacity) {
  std::vector<double> a;
  if (cfdgpu::sweepAngles(start, end, step, a) != 0) return -fail(CFDGPU_INVALID_ARGUMENT, "invalid sweep");
  for (size_t q = 0; q < a.size() && static_cast<int>(q) < capacity; ++q) out[q] = a[q];
  return static_cast<int>(a.size());
}

int cfdgpu_debug_array(const cfdgpu_solver* s, const char* name, double* out, int capacity) {
  SOLVER_GUARD(s)
  if (!name || !out) return fail(CFDGPU_INVALID_ARGUMENT, "null argument");
  return guarded([&]() -> int { return s->engine->debugArray(name, out, capacity); });
}

}  // extern "C"
