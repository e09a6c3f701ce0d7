// surface.hpp - the wall table of a C-grid for surface extraction and force
// integration on the device (SurfaceData.cpp:27-348, Forces.cpp:60-130). Its
// entries depend only on the mesh (every length and distance is formed once
// on the host, with std::hypot as the reference does); what depends on the
// flow is evaluated by k_surface.
#pragma once

namespace cfdgpu {

// One wall face, in i order (collectWallFaces, SurfaceData.cpp:27-59).
struct WallStation {
  double cx, cy;    // face centre
  double nx, ny;    // normal into the fluid
  double tx, ty;    // unit tangent along increasing i
  double len;       // face length (the integration segment)
  double wd;        // |(owner centroid - centre) . n|
  double od;        // the same for the second cell (0 when it is not used)
  double arc;       // arc length from the leading edge along its side
  int face, owner;  // j-face index (i, 0) and its cell
  int second;       // the cell above used for the two-point gradient, or -1
  int pad;
};

// The forces record written by the device: the cfdgpu_forces doubles in
// their declaration order, then its ints.
constexpr int kForceDoubles = 15;
constexpr int kForceInts = 5;

}  // namespace cfdgpu
