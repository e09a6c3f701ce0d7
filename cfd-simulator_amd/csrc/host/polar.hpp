// polar.hpp - the reference's polar sweep and its CSV, on top of the device
// solver (Panels.cpp:1137-1157, 1581-1800; Polar.cpp:79-163).
#pragma once

#include <string>
#include <vector>

#include "../../../include/cfdgpu.h"
#include "grid.hpp"

namespace cfdgpu {

struct PolarRequest {
  const Grid* grid{nullptr};
  cfdgpu_settings settings{};
  int model{CFDGPU_MODEL_NONE};
  const cfdgpu_sst_constants* sst{nullptr};
  double speed{50.0}, density{1.225}, reynolds{1e6}, chord{1.0}, referencePressure{0.0};
  std::vector<double> alphas;
  bool continued{true};
  cfdgpu_stop_rule rule{1e-6, 10000, 25};
};

// Runs the sweep; out gets one point per angle.
void polarSweep(const PolarRequest& request, std::vector<cfdgpu_polar_point>& out);

// Polar.cpp:126-163 (toCsv).
std::string polarCsv(const cfdgpu_polar_point* points, int count, const cfdgpu_polar_header& header);

// BoundaryConditions.cpp:67-97 with the default kinds, and FlowField::uniform
// (FlowField.cpp:21-47, Freestream.cpp:9-40).
void defaultFaceConditions(const Grid& g, double speed, double alphaDeg, double pref, int* kind, double* vxy,
                           double* pressure);
void uniformField(int cells, double speed, double alphaDeg, double density, double reynolds, double chord,
                  double pref, double* u, double* v, double* p, double* rho, double* mu);

}  // namespace cfdgpu
