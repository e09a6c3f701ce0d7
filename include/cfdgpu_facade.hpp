// cfdgpu_facade.hpp - the reference's C++ solver API, served by the MI355X
// library (libcfdgpu.so) through its C ABI (cfdgpu.h).
//
// Header-only. The names, members and signatures follow the reference
// (sidhanth-sridhar/CFD-SIMULATOR, include/cfd/...) for the hot path and its
// inputs and outputs, so code written against
//   cfd::mesh::generateCGrid / generateBox, cfd::flow::buildFaceConditions,
//   cfd::flow::FlowField::uniform, cfd::solver::SimpleSolver (+ KOmegaSST,
//   LaminarModel), cfd::post::extractSurface / integrateForces / sweepAngles
// compiles against this header and runs on the GPU. Differences from the
// reference, all forced by the device residency of the solver state:
//   * SimpleSolver::field() / massFlux() return host copies refreshed from the
//     device on each call; writes through the non-const field() are uploaded
//     (velocity and pressure) before the next iterate().
//   * setTurbulenceModel accepts the closures the device implements
//     (KOmegaSST, LaminarModel, nullptr); any other TurbulenceModel returns
//     ErrorCode::NotImplemented, since its update() would run on the host.
//   * Meshes are reference-counted handles to the library's mesh; copies share
//     it.
// Requires C++17.
#pragma once

#include <array>
#include <cmath>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>
#include <variant>
#include <vector>

#include "cfdgpu.h"

namespace cfd {

// ---- core (include/cfd/core/Vec2.hpp, Error.hpp) ----------------------------

struct Vec2 {
  double x{0.0};
  double y{0.0};
  friend Vec2 operator+(Vec2 a, Vec2 b) { return {a.x + b.x, a.y + b.y}; }
  friend Vec2 operator-(Vec2 a, Vec2 b) { return {a.x - b.x, a.y - b.y}; }
  friend Vec2 operator*(double s, Vec2 a) { return {s * a.x, s * a.y}; }
  friend Vec2 operator*(Vec2 a, double s) { return {s * a.x, s * a.y}; }
  double length() const { return std::hypot(x, y); }
};
inline double dot(Vec2 a, Vec2 b) { return a.x * b.x + a.y * b.y; }
inline double distance(Vec2 a, Vec2 b) { return (a - b).length(); }

enum class ErrorCode {
  InvalidArgument,
  OutOfRange,
  NotFound,
  IoFailure,
  NotImplemented,
  InitializationFailure,
  NumericalFailure,
  Internal,
};

class Error {
 public:
  Error(ErrorCode code, std::string message) : code_(code), message_(std::move(message)) {}
  ErrorCode code() const noexcept { return code_; }
  const std::string& message() const noexcept { return message_; }
  std::string format() const { return message_; }

 private:
  ErrorCode code_;
  std::string message_;
};

class BadResultAccess : public std::logic_error {
 public:
  using std::logic_error::logic_error;
};

template <typename T>
class Result {
 public:
  using value_type = T;
  Result(T value) : storage_(std::in_place_index<0>, std::move(value)) {}
  Result(Error error) : storage_(std::in_place_index<1>, std::move(error)) {}
  bool hasValue() const noexcept { return storage_.index() == 0; }
  bool hasError() const noexcept { return !hasValue(); }
  explicit operator bool() const noexcept { return hasValue(); }
  const T& value() const& { need(true); return std::get<0>(storage_); }
  T& value() & { need(true); return std::get<0>(storage_); }
  T&& value() && { need(true); return std::get<0>(std::move(storage_)); }
  const Error& error() const& { need(false); return std::get<1>(storage_); }
  template <typename U>
  T valueOr(U&& fallback) const& {
    return hasValue() ? std::get<0>(storage_) : static_cast<T>(std::forward<U>(fallback));
  }

 private:
  void need(bool value) const {
    if (value != hasValue()) throw BadResultAccess(value ? "Result holds an error" : "Result holds a value");
  }
  std::variant<T, Error> storage_;
};

class Status {
 public:
  Status(Error error) : error_(std::make_shared<Error>(std::move(error))) {}
  static Status ok() { return Status(); }
  bool hasError() const noexcept { return static_cast<bool>(error_); }
  bool hasValue() const noexcept { return !hasError(); }
  explicit operator bool() const noexcept { return !hasError(); }
  const Error& error() const {
    if (!error_) throw BadResultAccess("Status holds no error");
    return *error_;
  }

 private:
  Status() = default;
  std::shared_ptr<Error> error_;
};

namespace detail {

inline ErrorCode codeOf(int status) {
  switch (status) {
    case CFDGPU_INVALID_ARGUMENT: return ErrorCode::InvalidArgument;
    case CFDGPU_OUT_OF_RANGE: return ErrorCode::OutOfRange;
    case CFDGPU_NOT_FOUND: return ErrorCode::NotFound;
    case CFDGPU_IO_FAILURE: return ErrorCode::IoFailure;
    case CFDGPU_NOT_IMPLEMENTED: return ErrorCode::NotImplemented;
    case CFDGPU_INITIALIZATION_FAILURE: return ErrorCode::InitializationFailure;
    case CFDGPU_NUMERICAL_FAILURE: return ErrorCode::NumericalFailure;
    default: return ErrorCode::Internal;
  }
}
inline Error lastError(int status) { return Error(codeOf(status), cfdgpu_last_error()); }
inline Status statusOf(int status) { return status == CFDGPU_OK ? Status::ok() : Status(lastError(status)); }
inline void require(int status) {
  if (status != CFDGPU_OK) throw std::runtime_error(cfdgpu_last_error());
}

}  // namespace detail

// ---- geometry (include/cfd/geom/Airfoil.hpp, Naca4.hpp) ---------------------

namespace geom {

enum class TrailingEdge { Open, Closed };

struct AirfoilOptions {
  int pointsPerSurface{121};
  double chord{1.0};
  TrailingEdge trailingEdge{TrailingEdge::Open};
};

// A NACA four-digit section as the mesh generator needs it (the device path
// meshes closed trailing edges only, as the reference's C-grid does:
// CGrid.cpp:100-104).
class Airfoil {
 public:
  Airfoil(int m, int p, int t, AirfoilOptions options) : m_(m), p_(p), t_(t), options_(options) {}
  double chord() const noexcept { return options_.chord; }
  TrailingEdge trailingEdgeStyle() const noexcept { return options_.trailingEdge; }
  Vec2 trailingEdge() const noexcept { return {options_.chord, 0.0}; }
  int maxCamberDigit() const noexcept { return m_; }
  int camberPositionDigit() const noexcept { return p_; }
  int thicknessDigits() const noexcept { return t_; }

 private:
  int m_, p_, t_;
  AirfoilOptions options_;
};

// makeNaca4Digit("2412", options) (Naca4.hpp).
inline Result<Airfoil> makeNaca4Digit(std::string_view designation, const AirfoilOptions& options = {}) {
  if (designation.size() != 4) return Error(ErrorCode::InvalidArgument, "a NACA four-digit designation has four digits");
  for (char ch : designation)
    if (ch < '0' || ch > '9') return Error(ErrorCode::InvalidArgument, "a NACA designation is digits only");
  const int m = designation[0] - '0', p = designation[1] - '0';
  const int t = (designation[2] - '0') * 10 + (designation[3] - '0');
  if (t <= 0) return Error(ErrorCode::InvalidArgument, "zero thickness");
  if ((m == 0) != (p == 0)) return Error(ErrorCode::InvalidArgument, "camber and its position must both be zero or both set");
  if (!(options.chord > 0.0)) return Error(ErrorCode::InvalidArgument, "chord must be positive");
  return Airfoil(m, p, t, options);
}

}  // namespace geom

// ---- mesh (include/cfd/mesh/Mesh.hpp, CGrid.hpp, BoxGrid.hpp) -----------------

namespace mesh {

enum class BoundaryType { Interior = CFDGPU_INTERIOR, Wall = CFDGPU_WALL, Farfield = CFDGPU_FARFIELD,
                          Outlet = CFDGPU_OUTLET, WakeCut = CFDGPU_WAKECUT };

struct Face {
  int owner{-1};
  int neighbour{-1};
  BoundaryType boundary{BoundaryType::Interior};
  int partner{-1};
  bool isInterior() const noexcept { return neighbour >= 0; }
};

struct MeshQuality {
  double minCellArea{0.0};
  double maxCellArea{0.0};
  double maxAspectRatio{0.0};
  double maxNonOrthogonalityDeg{0.0};
  std::size_t invertedCells{0};
  double minWallSpacing{0.0};
};

class Mesh {
 public:
  explicit Mesh(cfdgpu_mesh* handle) : handle_(handle, cfdgpu_mesh_destroy) { load(); }
  const cfdgpu_mesh* handle() const noexcept { return handle_.get(); }

  const std::vector<Vec2>& nodes() const noexcept { return nodes_; }
  const std::vector<Face>& faces() const noexcept { return faces_; }
  const std::vector<Vec2>& cellCentroids() const noexcept { return centroids_; }
  const std::vector<double>& cellAreas() const noexcept { return areas_; }
  const std::vector<Vec2>& faceCentres() const noexcept { return faceCentres_; }
  const std::vector<Vec2>& faceNormals() const noexcept { return faceNormals_; }
  const std::vector<double>& faceAreas() const noexcept { return faceAreas_; }
  std::size_t nodeCount() const noexcept { return nodes_.size(); }
  std::size_t cellCount() const noexcept { return areas_.size(); }
  std::size_t faceCount() const noexcept { return faces_.size(); }
  std::size_t countFaces(BoundaryType type) const {
    std::size_t n = 0;
    for (const Face& f : faces_) n += (f.boundary == type) ? 1 : 0;
    return n;
  }
  double totalArea() const noexcept { return totalArea_; }
  const MeshQuality& quality() const noexcept { return quality_; }
  int nodesI() const noexcept { return nodesI_; }
  int nodesJ() const noexcept { return nodesJ_; }
  bool isStructured() const noexcept { return nodesI_ > 0 && nodesJ_ > 0; }
  int nodeIndex(int i, int j) const noexcept { return j * nodesI_ + i; }
  int cellIndex(int i, int j) const noexcept { return j * (nodesI_ - 1) + i; }

 private:
  void load() {
    int ni = 0, nj = 0, nc = 0, nf = 0;
    detail::require(cfdgpu_mesh_dims(handle_.get(), &ni, &nj, &nc, &nf));
    nodesI_ = ni;
    nodesJ_ = nj;
    std::vector<double> xy(2 * static_cast<size_t>(ni) * nj), cxy(2 * static_cast<size_t>(nc)), area(nc);
    detail::require(cfdgpu_mesh_nodes(handle_.get(), xy.data()));
    detail::require(cfdgpu_mesh_cells(handle_.get(), cxy.data(), area.data()));
    std::vector<double> fc(2 * static_cast<size_t>(nf)), fn(2 * static_cast<size_t>(nf)), fa(nf);
    std::vector<int> owner(nf), neighbour(nf), boundary(nf), partner(nf);
    detail::require(cfdgpu_mesh_faces(handle_.get(), fc.data(), fn.data(), fa.data(), owner.data(), neighbour.data(),
                                      boundary.data(), partner.data()));
    nodes_.resize(xy.size() / 2);
    for (size_t q = 0; q < nodes_.size(); ++q) nodes_[q] = {xy[2 * q], xy[2 * q + 1]};
    centroids_.resize(nc);
    for (int c = 0; c < nc; ++c) centroids_[c] = {cxy[2 * c], cxy[2 * c + 1]};
    areas_ = std::move(area);
    faces_.resize(nf);
    faceCentres_.resize(nf);
    faceNormals_.resize(nf);
    for (int f = 0; f < nf; ++f) {
      faces_[f] = Face{owner[f], neighbour[f], static_cast<BoundaryType>(boundary[f]), partner[f]};
      faceCentres_[f] = {fc[2 * f], fc[2 * f + 1]};
      faceNormals_[f] = {fn[2 * f], fn[2 * f + 1]};
    }
    faceAreas_ = std::move(fa);
    double q[7] = {};
    detail::require(cfdgpu_mesh_quality(handle_.get(), q));
    quality_ = MeshQuality{q[0], q[1], q[2], q[3], static_cast<std::size_t>(q[4]), q[5]};
    totalArea_ = q[6];
  }

  std::shared_ptr<cfdgpu_mesh> handle_;
  std::vector<Vec2> nodes_, centroids_, faceCentres_, faceNormals_;
  std::vector<double> areas_, faceAreas_;
  std::vector<Face> faces_;
  MeshQuality quality_{};
  double totalArea_{0.0};
  int nodesI_{0}, nodesJ_{0};
};

enum class MeshResolution { Coarse, Medium, Fine };

struct CGridOptions {
  double upstreamChords{12.0};
  double downstreamChords{25.0};
  double verticalChords{12.0};
  int surfacePoints{160};
  int wakePoints{56};
  int normalPoints{72};
  double firstLayerHeight{3e-4};
  double normalBlendLength{0.06};
  int normalSmoothingPasses{6};
};

inline CGridOptions optionsFor(MeshResolution resolution) {
  cfdgpu_cgrid_options o{};
  detail::require(cfdgpu_cgrid_defaults(static_cast<int>(resolution), &o));
  return CGridOptions{o.upstream_chords, o.downstream_chords, o.vertical_chords, o.surface_points, o.wake_points,
                      o.normal_points, o.first_layer_height, o.normal_blend_length, o.normal_smoothing_passes};
}

inline Result<Mesh> generateCGrid(const geom::Airfoil& airfoil, const CGridOptions& options) {
  if (airfoil.trailingEdgeStyle() != geom::TrailingEdge::Closed)
    return Error(ErrorCode::InvalidArgument, "the C-grid needs a closed trailing edge");
  const cfdgpu_cgrid_options o{options.upstreamChords, options.downstreamChords, options.verticalChords,
                               options.surfacePoints, options.wakePoints, options.normalPoints,
                               options.firstLayerHeight, options.normalBlendLength, options.normalSmoothingPasses};
  cfdgpu_mesh* h = nullptr;
  const int rc = cfdgpu_mesh_cgrid(airfoil.maxCamberDigit(), airfoil.camberPositionDigit(), airfoil.thicknessDigits(),
                                   airfoil.chord(), &o, &h);
  if (rc != CFDGPU_OK) return detail::lastError(rc);
  return Mesh(h);
}

enum class Grading { Uniform, TowardLower, TowardBoth };

struct BoxOptions {
  double originX{0.0};
  double originY{0.0};
  double length{1.0};
  double height{1.0};
  int cellsX{40};
  int cellsY{20};
  Grading grading{Grading::Uniform};
  double firstCellHeight{0.01};
  BoundaryType left{BoundaryType::Farfield};
  BoundaryType right{BoundaryType::Outlet};
  BoundaryType lower{BoundaryType::Wall};
  BoundaryType upper{BoundaryType::Wall};
};

inline Result<Mesh> generateBox(const BoxOptions& options) {
  const cfdgpu_box_options o{options.originX, options.originY, options.length, options.height, options.cellsX,
                             options.cellsY, static_cast<int>(options.grading), options.firstCellHeight,
                             static_cast<int>(options.left), static_cast<int>(options.right),
                             static_cast<int>(options.lower), static_cast<int>(options.upper)};
  cfdgpu_mesh* h = nullptr;
  const int rc = cfdgpu_mesh_box(&o, &h);
  if (rc != CFDGPU_OK) return detail::lastError(rc);
  return Mesh(h);
}

}  // namespace mesh

// ---- flow (include/cfd/flow/...) ----------------------------------------------

namespace flow {

struct FreestreamConditions {
  double speed{50.0};
  double angleOfAttackDeg{0.0};
  double density{1.225};
  double referencePressure{0.0};
  double reynoldsNumber{1.0e6};
  double dynamicViscosityOverride{0.0};

  double angleOfAttackRad() const noexcept { return angleOfAttackDeg * 3.141592653589793238462643383279502884 / 180.0; }
  Vec2 velocity() const noexcept {
    const double a = angleOfAttackRad();
    return {speed * std::cos(a), speed * std::sin(a)};
  }
  double dynamicPressure() const noexcept { return 0.5 * density * speed * speed; }
  double dynamicViscosity(double chord) const noexcept {
    return dynamicViscosityOverride > 0.0 ? dynamicViscosityOverride : density * speed * chord / reynoldsNumber;
  }
  Status validate() const {
    if (!(speed > 0.0) || !std::isfinite(speed)) return Error(ErrorCode::InvalidArgument, "freestream speed must be positive");
    if (!(density > 0.0)) return Error(ErrorCode::InvalidArgument, "density must be positive");
    if (!(reynoldsNumber > 0.0)) return Error(ErrorCode::InvalidArgument, "Reynolds number must be positive");
    if (!(std::abs(angleOfAttackDeg) <= 90.0)) return Error(ErrorCode::InvalidArgument, "angle of attack must lie within +/-90 degrees");
    return Status::ok();
  }
};

struct FlowField {
  std::vector<Vec2> velocity;
  std::vector<double> pressure;
  std::vector<double> density;
  std::vector<double> viscosity;

  void resize(std::size_t cells) {
    velocity.assign(cells, Vec2{});
    pressure.assign(cells, 0.0);
    density.assign(cells, 0.0);
    viscosity.assign(cells, 0.0);
  }
  std::size_t size() const noexcept { return velocity.size(); }
  bool empty() const noexcept { return velocity.empty(); }
  bool isConsistent() const noexcept {
    return pressure.size() == velocity.size() && density.size() == velocity.size() &&
           viscosity.size() == velocity.size();
  }
  static Result<FlowField> uniform(std::size_t cells, const FreestreamConditions& freestream, double chord) {
    if (cells == 0) return Error(ErrorCode::InvalidArgument, "cannot initialise a field with no cells");
    if (const Status s = freestream.validate(); !s) return s.error();
    if (!(chord > 0.0)) return Error(ErrorCode::InvalidArgument, "chord must be positive");
    FlowField f;
    f.velocity.assign(cells, freestream.velocity());
    f.pressure.assign(cells, freestream.referencePressure);
    f.density.assign(cells, freestream.density);
    f.viscosity.assign(cells, freestream.dynamicViscosity(chord));
    return f;
  }
};

struct ResidualSet {
  double continuity{0.0};
  double momentumX{0.0};
  double momentumY{0.0};
  double worst() const noexcept {
    double w = std::abs(continuity);
    if (std::abs(momentumX) > w) w = std::abs(momentumX);
    if (std::abs(momentumY) > w) w = std::abs(momentumY);
    return w;
  }
};

enum class BoundaryKind { Internal = CFDGPU_K_INTERNAL, Inlet = CFDGPU_K_INLET, Outlet = CFDGPU_K_OUTLET,
                          FarField = CFDGPU_K_FARFIELD, NoSlipWall = CFDGPU_K_WALL };

struct BoundaryConditions {
  BoundaryKind wall{BoundaryKind::NoSlipWall};
  BoundaryKind farField{BoundaryKind::FarField};
  BoundaryKind outlet{BoundaryKind::Outlet};

  BoundaryKind kindFor(mesh::BoundaryType type) const noexcept {
    switch (type) {
      case mesh::BoundaryType::Wall: return wall;
      case mesh::BoundaryType::Farfield: return farField;
      case mesh::BoundaryType::Outlet: return outlet;
      default: return BoundaryKind::Internal;
    }
  }
  Status validate() const {
    const bool anchors = outlet == BoundaryKind::Outlet || outlet == BoundaryKind::FarField ||
                         farField == BoundaryKind::Outlet || farField == BoundaryKind::FarField;
    if (!anchors) return Error(ErrorCode::InvalidArgument, "no boundary sets a pressure");
    if (wall == BoundaryKind::Internal) return Error(ErrorCode::InvalidArgument, "the airfoil surface cannot be an internal face");
    return Status::ok();
  }
};

struct FaceCondition {
  BoundaryKind kind{BoundaryKind::Internal};
  Vec2 velocity{};
  double pressure{0.0};
};
using FaceConditions = std::vector<FaceCondition>;

// BoundaryConditions.cpp:67-97: each face's kind from its boundary type, the
// stream velocity except on walls, the reference pressure everywhere.
inline Result<FaceConditions> buildFaceConditions(const mesh::Mesh& mesh, const BoundaryConditions& conditions,
                                                  const FreestreamConditions& freestream) {
  if (const Status s = conditions.validate(); !s) return s.error();
  if (const Status s = freestream.validate(); !s) return s.error();
  const Vec2 stream = freestream.velocity();
  FaceConditions out(mesh.faceCount());
  for (std::size_t f = 0; f < mesh.faceCount(); ++f) {
    const BoundaryKind kind = conditions.kindFor(mesh.faces()[f].boundary);
    out[f].kind = kind;
    out[f].pressure = freestream.referencePressure;
    out[f].velocity = (kind == BoundaryKind::NoSlipWall) ? Vec2{0.0, 0.0} : stream;
  }
  return out;
}

}  // namespace flow

// ---- solver (include/cfd/solver/...) --------------------------------------------

namespace solver {

enum class ConvectionScheme { Upwind, SecondOrderUpwind };

struct TurbulenceInflow {
  double intensity{0.001};
  double viscosityRatio{1.0};
  double lengthScale{0.0};
};

struct TurbulenceRanges {
  double minEnergy{0.0}, maxEnergy{0.0};
  double minDissipation{0.0}, maxDissipation{0.0};
  double minEddyViscosity{0.0}, maxEddyViscosity{0.0};
};

struct TurbulenceResiduals {
  double first{0.0};
  double second{0.0};
  double worst() const noexcept { return first > second ? first : second; }
};

struct SSTConstants {
  double sigmaK1{0.85};
  double sigmaOmega1{0.5};
  double beta1{0.075};
  double sigmaK2{1.0};
  double sigmaOmega2{0.856};
  double beta2{0.0828};
  double betaStar{0.09};
  double kappa{0.41};
  double a1{0.31};
  double productionLimit{10.0};
  double gamma1() const noexcept { return beta1 / betaStar - sigmaOmega1 * kappa * kappa / std::sqrt(betaStar); }
  double gamma2() const noexcept { return beta2 / betaStar - sigmaOmega2 * kappa * kappa / std::sqrt(betaStar); }
};

class SimpleSolver;

// The plug-in interface (TurbulenceModel.hpp:155-206). On this path the
// closures run on the device; a model object is the caller's view of one.
class TurbulenceModel {
 public:
  virtual ~TurbulenceModel() = default;
  virtual std::string_view name() const noexcept = 0;
  virtual const std::vector<double>& eddyViscosity() const noexcept = 0;
  virtual TurbulenceResiduals residuals() const noexcept = 0;
  virtual TurbulenceRanges ranges() const noexcept { return {}; }
  virtual const std::vector<double>* turbulentEnergyField() const noexcept { return nullptr; }
  virtual const std::vector<double>* dissipationField() const noexcept { return nullptr; }
  virtual double maxWallYPlus() const noexcept { return 0.0; }

 protected:
  friend class SimpleSolver;
  virtual int deviceModel() const noexcept { return -1; }  // CFDGPU_MODEL_*, -1: not on the device
  virtual void attach(cfdgpu_solver*, std::size_t /*cells*/) {}
  virtual void refresh(const cfdgpu_monitor&) {}
};

class LaminarModel final : public TurbulenceModel {
 public:
  std::string_view name() const noexcept override { return "laminar"; }
  const std::vector<double>& eddyViscosity() const noexcept override { return zero_; }
  TurbulenceResiduals residuals() const noexcept override { return {}; }

 protected:
  int deviceModel() const noexcept override { return CFDGPU_MODEL_LAMINAR; }

 private:
  std::vector<double> zero_;
};

class KOmegaSST final : public TurbulenceModel {
 public:
  explicit KOmegaSST(SSTConstants constants = {}) : constants_(constants) {}
  std::string_view name() const noexcept override { return "k-omega SST"; }
  const SSTConstants& constants() const noexcept { return constants_; }
  const std::vector<double>& eddyViscosity() const noexcept override { sync(); return mut_; }
  TurbulenceResiduals residuals() const noexcept override { return residuals_; }
  TurbulenceRanges ranges() const noexcept override { return ranges_; }
  const std::vector<double>* turbulentEnergyField() const noexcept override { sync(); return &k_; }
  const std::vector<double>* dissipationField() const noexcept override { sync(); return &omega_; }
  double maxWallYPlus() const noexcept override { info(); return yplus_; }
  const std::vector<double>& turbulentEnergy() const noexcept { sync(); return k_; }
  const std::vector<double>& specificDissipation() const noexcept { sync(); return omega_; }
  const std::vector<double>& wallDistance() const noexcept { sync(); return wd_; }
  double freestreamEnergy() const noexcept { info(); return kIn_; }
  double freestreamDissipation() const noexcept { info(); return wIn_; }

 protected:
  int deviceModel() const noexcept override { return CFDGPU_MODEL_SST; }
  void attach(cfdgpu_solver* s, std::size_t cells) override {
    solver_ = s;
    cells_ = cells;
  }
  void refresh(const cfdgpu_monitor& m) override {
    residuals_ = {m.turb_first, m.turb_second};
    ranges_ = {m.min_k, m.max_k, m.min_w, m.max_w, m.min_mut, m.max_mut};
  }

 private:
  void sync() const noexcept {
    if (!solver_) return;
    k_.resize(cells_); omega_.resize(cells_); mut_.resize(cells_); wd_.resize(cells_);
    (void)cfdgpu_get_turbulence(solver_, k_.data(), omega_.data(), mut_.data(), wd_.data());
  }
  void info() const noexcept {
    if (solver_) (void)cfdgpu_turbulence_info(solver_, &kIn_, &wIn_, &yplus_);
  }
  SSTConstants constants_;
  cfdgpu_solver* solver_{nullptr};
  std::size_t cells_{0};
  mutable std::vector<double> k_, omega_, mut_, wd_;
  mutable double kIn_{0.0}, wIn_{0.0}, yplus_{0.0};
  TurbulenceResiduals residuals_{};
  TurbulenceRanges ranges_{};
};

struct SimpleSettings {
  double velocityRelaxation{0.5};
  double pressureRelaxation{0.2};
  int momentumSweeps{3};
  int pressureIterations{200};
  double pressureTolerance{0.01};
  ConvectionScheme scheme{ConvectionScheme::Upwind};
  double turbulenceRelaxation{0.5};
  TurbulenceInflow inflow{};
  int nonOrthogonalCorrectors{1};
  double nonOrthogonalBlend{1.0};
};

struct SolverMonitor {
  flow::ResidualSet residuals;
  double massImbalance{0.0};
  double maxDivergence{0.0};
  int pressureIterations{0};
  TurbulenceResiduals turbulence{};
  TurbulenceRanges turbulenceRanges{};
  double maxEddyViscosityRatio{0.0};
};

// solver::SimpleSolver (SimpleSolver.hpp:159-197) on the device.
class SimpleSolver {
 public:
  static Result<SimpleSolver> create(const mesh::Mesh& mesh, flow::FaceConditions conditions,
                                     const SimpleSettings& settings) {
    const size_t nf = mesh.faceCount();
    if (conditions.size() != nf) return Error(ErrorCode::InvalidArgument, "expected one condition per face");
    std::vector<int> kind(nf);
    std::vector<double> vxy(2 * nf), pressure(nf);
    for (size_t f = 0; f < nf; ++f) {
      kind[f] = static_cast<int>(conditions[f].kind);
      vxy[2 * f] = conditions[f].velocity.x;
      vxy[2 * f + 1] = conditions[f].velocity.y;
      pressure[f] = conditions[f].pressure;
    }
    const cfdgpu_settings s = toC(settings);
    cfdgpu_solver* h = nullptr;
    const int rc = cfdgpu_create(&h, mesh.handle(), kind.data(), vxy.data(), pressure.data(), &s);
    if (rc != CFDGPU_OK) return detail::lastError(rc);
    return SimpleSolver(mesh, h, settings);
  }

  Status initialise(const flow::FlowField& initial) {
    const size_t n = mesh_->cellCount();
    if (!initial.isConsistent() || initial.size() != n)
      return Error(ErrorCode::InvalidArgument, "the starting field does not match the mesh");
    std::vector<double> u(n), v(n);
    for (size_t c = 0; c < n; ++c) { u[c] = initial.velocity[c].x; v[c] = initial.velocity[c].y; }
    const int rc = cfdgpu_initialise(handle_.get(), u.data(), v.data(), initial.pressure.data(),
                                     initial.density.data(), initial.viscosity.data());
    if (rc == CFDGPU_OK) { field_ = initial; dirty_ = false; }
    return detail::statusOf(rc);
  }

  Status setTurbulenceModel(std::unique_ptr<TurbulenceModel> model) {
    const int kind = model ? model->deviceModel() : CFDGPU_MODEL_NONE;
    if (kind < 0) return Error(ErrorCode::NotImplemented, "only the device closures (k-omega SST, laminar) run on this path");
    cfdgpu_sst_constants c{};
    const cfdgpu_sst_constants* pc = nullptr;
    if (auto* sst = dynamic_cast<KOmegaSST*>(model.get())) {
      const SSTConstants& k = sst->constants();
      c = cfdgpu_sst_constants{k.sigmaK1, k.sigmaOmega1, k.beta1, k.sigmaK2, k.sigmaOmega2, k.beta2,
                               k.betaStar, k.kappa, k.a1, k.productionLimit};
      pc = &c;
    }
    const int rc = cfdgpu_set_turbulence(handle_.get(), kind, pc);
    if (rc != CFDGPU_OK) return detail::lastError(rc);
    if (model) model->attach(handle_.get(), mesh_->cellCount());
    turbulence_ = std::move(model);
    return Status::ok();
  }
  const TurbulenceModel* turbulenceModel() const noexcept { return turbulence_.get(); }

  SolverMonitor iterate() {
    if (dirty_) upload();
    cfdgpu_monitor m{};
    detail::require(cfdgpu_iterate(handle_.get(), 1, &m));
    if (turbulence_) turbulence_->refresh(m);
    SolverMonitor out;
    out.residuals = {m.continuity, m.momentum_x, m.momentum_y};
    out.massImbalance = m.mass_imbalance;
    out.maxDivergence = m.max_divergence;
    out.pressureIterations = m.pressure_iterations;
    out.turbulence = {m.turb_first, m.turb_second};
    out.turbulenceRanges = {m.min_k, m.max_k, m.min_w, m.max_w, m.min_mut, m.max_mut};
    out.maxEddyViscosityRatio = m.max_eddy_ratio;
    return out;
  }

  const flow::FlowField& field() const { download(); return field_; }
  flow::FlowField& field() {
    download();
    dirty_ = true;  // the caller may write; velocity and pressure go back before the next iterate()
    return field_;
  }
  const std::vector<double>& massFlux() const {
    flux_.resize(mesh_->faceCount());
    detail::require(cfdgpu_get_massflux(handle_.get(), flux_.data()));
    return flux_;
  }
  const SimpleSettings& settings() const noexcept { return settings_; }
  void setSettings(const SimpleSettings& settings) {
    const cfdgpu_settings s = toC(settings);
    detail::require(cfdgpu_set_settings(handle_.get(), &s));
    settings_ = settings;
  }
  std::vector<double> divergence() const {
    std::vector<double> d(mesh_->cellCount());
    detail::require(cfdgpu_divergence(handle_.get(), d.data()));
    return d;
  }
  cfdgpu_solver* handle() const noexcept { return handle_.get(); }

 private:
  SimpleSolver(const mesh::Mesh& mesh, cfdgpu_solver* h, const SimpleSettings& settings)
      : mesh_(&mesh), handle_(h, cfdgpu_destroy), settings_(settings) {}

  static cfdgpu_settings toC(const SimpleSettings& s) {
    cfdgpu_settings c{};
    cfdgpu_default_settings(&c);  // device-path fields (precision, ordering) at their defaults
    c.velocity_relaxation = s.velocityRelaxation;
    c.pressure_relaxation = s.pressureRelaxation;
    c.momentum_sweeps = s.momentumSweeps;
    c.pressure_iterations = s.pressureIterations;
    c.pressure_tolerance = s.pressureTolerance;
    c.scheme = static_cast<int>(s.scheme);
    c.turbulence_relaxation = s.turbulenceRelaxation;
    c.intensity = s.inflow.intensity;
    c.viscosity_ratio = s.inflow.viscosityRatio;
    c.length_scale = s.inflow.lengthScale;
    c.non_orthogonal_correctors = s.nonOrthogonalCorrectors;
    c.non_orthogonal_blend = s.nonOrthogonalBlend;
    return c;
  }
  void download() const {
    const size_t n = mesh_->cellCount();
    std::vector<double> u(n), v(n), p(n), mu(n);
    detail::require(cfdgpu_get_field(handle_.get(), u.data(), v.data(), p.data(), mu.data()));
    if (field_.size() != n) field_.resize(n);
    for (size_t c = 0; c < n; ++c) field_.velocity[c] = {u[c], v[c]};
    field_.pressure = std::move(p);
    field_.viscosity = std::move(mu);
  }
  void upload() {
    const size_t n = mesh_->cellCount();
    std::vector<double> u(n), v(n);
    for (size_t c = 0; c < n; ++c) { u[c] = field_.velocity[c].x; v[c] = field_.velocity[c].y; }
    detail::require(cfdgpu_set_field(handle_.get(), u.data(), v.data(), field_.pressure.data()));
    dirty_ = false;
  }

  const mesh::Mesh* mesh_;
  std::unique_ptr<cfdgpu_solver, void (*)(cfdgpu_solver*)> handle_;
  SimpleSettings settings_;
  std::unique_ptr<TurbulenceModel> turbulence_;
  mutable flow::FlowField field_;
  mutable std::vector<double> flux_;
  bool dirty_{false};
};

}  // namespace solver

// ---- post-processing (include/cfd/post/SurfaceData.hpp, Forces.hpp, Polar.hpp) ----

namespace post {

enum class SurfaceSide { Upper, Lower };

// One wall face's station (SurfaceData.hpp).
struct SurfaceStation {
  SurfaceSide side{SurfaceSide::Upper};
  int face{-1};
  Vec2 position{};
  double xOverChord{0.0};
  double arcLength{0.0};
  double pressure{0.0};
  double pressureCoefficient{0.0};
  double wallShearStress{0.0};
  double skinFrictionCoefficient{0.0};
  bool nearStagnation{false};
  double panelLength{0.0};
};

struct AerodynamicForces {
  double liftCoefficient{0.0};
  double dragCoefficient{0.0};
  double momentCoefficient{0.0};
  double pressureDragCoefficient{0.0};
  double frictionDragCoefficient{0.0};
  double lift{0.0};
  double drag{0.0};
  double pitchingMoment{0.0};
  Vec2 pressureForce{};
  Vec2 frictionForce{};
  double upperSeparationXc{-1.0};
  double lowerSeparationXc{-1.0};
  double stagnationXc{0.0};
};

// The wall stations of a solved field together with the integrated forces:
// extractSurface followed by integrateForces in one call to the library.
struct SurfaceDistribution {
  std::vector<SurfaceStation> stations;
  flow::FreestreamConditions freestream{};
  double chord{1.0};
  AerodynamicForces forces{};
};

inline Result<SurfaceDistribution> extractSurface(const mesh::Mesh& mesh, const flow::FlowField& field,
                                                  const flow::FreestreamConditions& freestream, double chord) {
  const size_t n = mesh.cellCount();
  if (!field.isConsistent() || field.size() != n) return Error(ErrorCode::InvalidArgument, "field does not match the mesh");
  std::vector<double> u(n), v(n);
  for (size_t c = 0; c < n; ++c) { u[c] = field.velocity[c].x; v[c] = field.velocity[c].y; }
  const int cap = 12 * static_cast<int>(mesh.faceCount());
  std::vector<double> rows(static_cast<size_t>(cap));
  int count = 0;
  cfdgpu_forces f{};
  const int rc = cfdgpu_surface(mesh.handle(), u.data(), v.data(), field.pressure.data(), field.viscosity.data(),
                                freestream.speed, freestream.angleOfAttackDeg, freestream.density,
                                freestream.referencePressure, chord, rows.data(), cap, &count, &f);
  if (rc != CFDGPU_OK) return detail::lastError(rc);
  SurfaceDistribution out;
  out.freestream = freestream;
  out.chord = chord;
  out.stations.resize(static_cast<size_t>(count));
  for (int s = 0; s < count; ++s) {
    const double* r = &rows[12 * static_cast<size_t>(s)];
    out.stations[s] = SurfaceStation{r[0] > 0.5 ? SurfaceSide::Lower : SurfaceSide::Upper, static_cast<int>(r[1]),
                                     {r[2], r[3]}, r[4], r[5], r[6], r[7], r[8], r[9], r[10] > 0.5, r[11]};
  }
  out.forces = AerodynamicForces{f.cl, f.cd, f.cm, f.cd_pressure, f.cd_friction, f.lift, f.drag,
                                 f.pitching_moment, {f.fx_pressure, f.fy_pressure}, {f.fx_friction, f.fy_friction},
                                 f.upper_separation, f.lower_separation, f.stagnation_xc};
  return out;
}

inline Result<AerodynamicForces> integrateForces(const SurfaceDistribution& surface,
                                                 const flow::FreestreamConditions& /*freestream*/) {
  if (surface.stations.empty()) return Error(ErrorCode::InvalidArgument, "no surface stations");
  return surface.forces;
}

inline Result<std::vector<double>> sweepAngles(double start, double end, double step) {
  std::vector<double> out(4096);
  const int n = cfdgpu_sweep_angles(start, end, step, out.data(), static_cast<int>(out.size()));
  if (n < 0) return detail::lastError(-n);
  out.resize(static_cast<size_t>(n));
  return out;
}

}  // namespace post

}  // namespace cfd
