// post.hpp - FP64 surface data, forces and sweep angles on the host.
#pragma once

#include <vector>

#include "../../../include/cfdgpu.h"
#include "grid.hpp"

namespace cfdgpu {

// extractSurface + integrateForces (SurfaceData.cpp:208-348, Forces.cpp:60-130).
// Returns 0 or an ErrorCode value.
int surfaceAndForces(const Grid& g, const double* u, const double* v, const double* p, const double* mu, double speed,
                     double alphaDeg, double density, double pref, double chord, double* rows, int capacity,
                     int* count, cfdgpu_forces* out);

// sweepAngles (Polar.cpp:36-77). Returns 0 or an ErrorCode value.
int sweepAngles(double start, double end, double step, std::vector<double>& out);

}  // namespace cfdgpu
