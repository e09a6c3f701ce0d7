// post.hpp - FP64 surface data, forces and sweep angles on the host.
#pragma once

#include <vector>

#include "../../../include/cfdgpu.h"
#include "../device/surface.hpp"
#include "grid.hpp"

namespace cfdgpu {

// extractSurface + integrateForces (SurfaceData.cpp:208-348, Forces.cpp:60-130).
// Returns 0 or an ErrorCode value.
int surfaceAndForces(const Grid& g, const double* u, const double* v, const double* p, const double* mu, double speed,
                     double alphaDeg, double density, double pref, double chord, double* rows, int capacity,
                     int* count, cfdgpu_forces* out);

// The mesh-static part of surfaceAndForces for the device (k_surface): the
// wall faces in i order with their geometry, second cells and arc lengths,
// and the leading-edge index. False when the mesh has fewer than four wall
// faces (surfaceAndForces returns NotFound then).
bool wallTable(const Grid& g, std::vector<WallStation>& out, int& le);

// sweepAngles (Polar.cpp:36-77). Returns 0 or an ErrorCode value.
int sweepAngles(double start, double end, double step, std::vector<double>& out);

}  // namespace cfdgpu
