// Compiles the reference-style C++ API (include/cfdgpu_facade.hpp) and, with
// "solve" as argv[1], runs a small laminar C-grid solve through it on the GPU.
// Without a GPU it exercises the host-side parts: meshes, conditions, fields,
// sweepAngles and the error path of SimpleSolver::create.
#include <cstdio>
#include <cstring>
#include <memory>

#include "cfdgpu_facade.hpp"

using namespace cfd;

int main(int argc, char** argv) {
  const bool solve = argc > 1 && std::strcmp(argv[1], "solve") == 0;
  auto foil = geom::makeNaca4Digit("0012", geom::AirfoilOptions{.chord = 1.0, .trailingEdge = geom::TrailingEdge::Closed});
  if (!foil) { std::printf("FAIL naca\n"); return 1; }
  mesh::CGridOptions opts = mesh::optionsFor(mesh::MeshResolution::Coarse);
  opts.surfacePoints = 40;
  opts.wakePoints = 20;
  opts.normalPoints = 28;
  auto mesh = mesh::generateCGrid(foil.value(), opts);
  if (!mesh) { std::printf("FAIL mesh %s\n", mesh.error().format().c_str()); return 1; }
  const mesh::Mesh& m = mesh.value();
  if (m.cellCount() != 118u * 27u || m.countFaces(mesh::BoundaryType::Wall) != 78u) {
    std::printf("FAIL dims %zu %zu\n", m.cellCount(), m.countFaces(mesh::BoundaryType::Wall));
    return 1;
  }
  flow::FreestreamConditions fs{.speed = 1.0, .angleOfAttackDeg = 4.0, .density = 1.0, .reynoldsNumber = 500.0};
  auto conditions = flow::buildFaceConditions(m, flow::BoundaryConditions{}, fs);
  auto field = flow::FlowField::uniform(m.cellCount(), fs, 1.0);
  if (!conditions || !field) { std::printf("FAIL inputs\n"); return 1; }
  auto angles = post::sweepAngles(-4.0, 16.0, 1.0);
  if (!angles || angles.value().size() != 21) { std::printf("FAIL sweep\n"); return 1; }
  auto bad = solver::SimpleSolver::create(m, flow::FaceConditions(3), solver::SimpleSettings{});
  if (!bad.hasError() || bad.error().code() != ErrorCode::InvalidArgument) { std::printf("FAIL create check\n"); return 1; }
  if (!solve) { std::printf("ok\n"); return 0; }

  auto created = solver::SimpleSolver::create(m, conditions.value(), solver::SimpleSettings{});
  if (!created) { std::printf("FAIL create %s\n", created.error().format().c_str()); return 1; }
  solver::SimpleSolver& s = created.value();
  if (!s.initialise(field.value())) { std::printf("FAIL initialise\n"); return 1; }
  solver::SolverMonitor first = s.iterate(), last{};
  for (int it = 1; it < 200; ++it) last = s.iterate();
  auto surface = post::extractSurface(m, s.field(), fs, 1.0);
  if (!surface) { std::printf("FAIL surface\n"); return 1; }
  auto forces = post::integrateForces(surface.value(), fs);
  std::printf("continuity %.3e -> %.3e cl %.6f cd %.6f\n", first.residuals.continuity, last.residuals.continuity,
              forces.value().liftCoefficient, forces.value().dragCoefficient);
  if (!(last.residuals.continuity < first.residuals.continuity)) { std::printf("FAIL residual\n"); return 1; }
  // A warm start written through field() reaches the device.
  s.field().velocity[0] = Vec2{123.0, 0.0};
  (void)s.iterate();
  std::printf("ok\n");
  return 0;
}
