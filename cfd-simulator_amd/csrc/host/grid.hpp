// grid.hpp - structured C-grid / box meshes and their metrics, host FP64.
//
// The mesh is the layout contract of every kernel, so it is generated here
// with the reference's arithmetic (sidhanth-sridhar/CFD-SIMULATOR
// src/geom/Naca4.cpp, src/geom/Airfoil.cpp, src/mesh/Distribution.cpp,
// src/mesh/CGrid.cpp, src/mesh/BoxGrid.cpp, src/mesh/Mesh.cpp) but stored the
// way the device reads it: cell-major arrays plus two row-major face families
// instead of a generic face list.
#pragma once

#include <array>
#include <string>
#include <vector>

namespace cfdgpu {

struct Vec2 {
  double x{0.0}, y{0.0};
};
inline Vec2 operator+(Vec2 a, Vec2 b) { return Vec2{a.x + b.x, a.y + b.y}; }
inline Vec2 operator-(Vec2 a, Vec2 b) { return Vec2{a.x - b.x, a.y - b.y}; }
inline Vec2 operator*(Vec2 a, double s) { return Vec2{a.x * s, a.y * s}; }
double length(Vec2 v);  // std::hypot, as include/cfd/core/Vec2.hpp:64-66
inline double dot(Vec2 a, Vec2 b) { return a.x * b.x + a.y * b.y; }
inline double cross(Vec2 a, Vec2 b) { return a.x * b.y - a.y * b.x; }
inline double distance(Vec2 a, Vec2 b) { return length(b - a); }

enum BoundaryType { kInterior = 0, kWallB = 1, kFarfieldB = 2, kOutletB = 3, kWakeCutB = 4 };

struct CGridOptions {
  double upstream{12.0}, downstream{25.0}, vertical{12.0};
  int surfacePoints{160}, wakePoints{56}, normalPoints{72};
  double firstLayer{3e-4};
  double blendLength{0.06};
  int smoothingPasses{6};
};

struct BoxOptions {
  double originX{0.0}, originY{0.0}, length{1.0}, height{1.0};
  int cellsX{40}, cellsY{20};
  int grading{0};
  double firstCellHeight{0.01};
  int left{kFarfieldB}, right{kOutletB}, lower{kWallB}, upper{kWallB};
};

// A structured block of nI x nJ cells.
struct Grid {
  int nodesI{0}, nodesJ{0};
  int nI{0}, nJ{0};
  std::vector<Vec2> nodes;          // j*nodesI + i
  std::vector<int> bottomType;      // nI: type of j-face (i, 0)
  std::vector<int> bottomPartner;   // nI: partner cell i' across the wake cut, or -1
  int topType{kFarfieldB}, leftType{kOutletB}, rightType{kOutletB};

  std::vector<double> area;         // N, signed shoelace area
  std::vector<Vec2> centroid;       // N
  // j-faces: index j*nI + i (j = 0..nJ); i-faces: index j*(nI+1) + i (i = 0..nI)
  std::vector<Vec2> jCentre, jNormal, iCentre, iNormal;
  std::vector<double> jLen, iLen;
  double minArea{0}, maxArea{0}, maxAspect{0}, maxNonOrthoDeg{0}, minWallSpacing{0}, totalArea{0};
  int inverted{0};

  int cells() const { return nI * nJ; }
  int jFaces() const { return nI * (nJ + 1); }
  int iFaces() const { return (nI + 1) * nJ; }
  int faces() const { return jFaces() + iFaces(); }
  // the reference's face numbering (src/mesh/Mesh.cpp:149-156)
  int refJFace(int i, int j) const { return j * nI + i; }
  int refIFace(int i, int j) const { return jFaces() + i * nJ + j; }
  // face type in the reference's numbering
  int faceType(int refFace) const;
};

bool buildGrid(Grid& g, int nodesI, int nodesJ, std::vector<Vec2> nodes, std::vector<int> bottomType,
               std::vector<int> bottomPartner, int top, int left, int right, std::string& error);
bool makeCGrid(Grid& g, int m, int p, int t, double chord, const CGridOptions& o, std::string& error);
bool makeBox(Grid& g, const BoxOptions& o, std::string& error);
CGridOptions presetOptions(int resolution);  // CGrid.cpp:73-96

// NACA four-digit closed contour (Airfoil.cpp:110-125), exposed for tests.
std::vector<Vec2> nacaContour(int m, int p, int t, int points, double chord, bool closed);
std::vector<double> geometricDistribution(double first, double total, int intervals);

}  // namespace cfdgpu
