// engine.cpp - host setup and device residency of one SIMPLE instance.
#include "engine.hpp"
#include "post.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <mutex>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace cfdgpu {

namespace {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

constexpr double kTiny = 1e-300;     // SimpleSolver.cpp:16
constexpr double kTinySst = 1.0e-30; // KOmegaSST.cpp:22
constexpr double kMinOmega = 1.0e-10;

struct Carver {
  size_t offset{0};
  char* base{nullptr};
  template <class T>
  T* take(size_t count) {
    offset = (offset + 255) & ~static_cast<size_t>(255);
    T* p = base ? reinterpret_cast<T*>(base + offset) : nullptr;
    offset += count * sizeof(T);
    return p;
  }
};

}  // namespace

// Event pairs around each stage occurrence, drawn from a pool that grows to
// the largest timed call and is reused after that.
class StageClock final : public StageRecorder {
 public:
  ~StageClock() {
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
  }
  void reset() {
    used_ = 0;
    marks_.clear();
    open_.clear();
  }
  // stages may nest (the CG iterations inside the pressure solve)
  void begin(int stage, hipStream_t s) override {
    open_.push_back(marks_.size());
    marks_.push_back({stage, used_, used_});
    record(s);
  }
  void end(int, hipStream_t s) override {
    marks_[open_.back()].end = used_;
    open_.pop_back();
    record(s);
  }
  // after the stream has drained
  void collect(double* ms, int* calls) const {
    for (int q = 0; q < CFDGPU_STAGE_COUNT; ++q) {
      if (ms) ms[q] = 0.0;
      if (calls) calls[q] = 0;
    }
    for (const Mark& m : marks_) {
      float t = 0.0f;
      check(hipEventElapsedTime(&t, pool_[m.begin], pool_[m.end]), "stage timing");
      if (ms) ms[m.stage] += static_cast<double>(t);
      if (calls) calls[m.stage] += 1;
    }
  }

 private:
  void record(hipStream_t s) {
    if (used_ == pool_.size()) {
      hipEvent_t e{};
      check(hipEventCreate(&e), "hipEventCreate");
      pool_.push_back(e);
    }
    check(hipEventRecord(pool_[used_], s), "hipEventRecord");
    ++used_;
  }
  struct Mark {
    int stage;
    size_t begin, end;  // event indices
  };
  std::vector<hipEvent_t> pool_;
  size_t used_{0};
  std::vector<Mark> marks_;
  std::vector<size_t> open_;  // marks begun and not yet ended
};

DeviceArena::~DeviceArena() {
  if (base) (void)hipFree(base);
}

Engine::Engine(const Grid& grid, std::vector<int> kind, std::vector<Vec2> bcVelocity, std::vector<double> bcPressure,
               const cfdgpu_settings& settings)
    : grid_(grid), kind_(std::move(kind)), bcVel_(std::move(bcVelocity)), bcP_(std::move(bcPressure)),
      settings_(settings) {
  const int nf = grid_.faces();
  if (static_cast<int>(kind_.size()) != nf || static_cast<int>(bcVel_.size()) != nf ||
      static_cast<int>(bcP_.size()) != nf)
    throw std::invalid_argument("face conditions do not match the mesh");
  if (grid_.cells() == 0) throw std::invalid_argument("cannot solve on an empty mesh");
  if ((grid_.cells() + kBlock - 1) / kBlock > kMaxRedBlocks)
    throw std::invalid_argument("mesh exceeds the per-instance cell capacity (1,048,576 cells)");
  if (grid_.nJ > 1024) throw std::invalid_argument("more than 1024 cells wall-normal is not supported");
  // SimpleSolver.cpp:71-77
  for (int f = 0; f < nf; ++f) {
    if (kind_[static_cast<size_t>(f)] == CFDGPU_K_OUTLET || kind_[static_cast<size_t>(f)] == CFDGPU_K_FARFIELD) {
      hasPRef_ = true;
      break;
    }
  }
  cfdgpu_default_sst(&sst_);
  dead_.assign(static_cast<size_t>(nf), 0.0);
  precomputeGeometry();
  check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  allocate();
  checkOrdering(settings_);
  uploadStatic();
}

Engine::~Engine() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (dWall_) (void)hipFree(dWall_);
  if (dSurf_) (void)hipFree(dSurf_);
  if (dInst_) (void)hipFree(dInst_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// SimpleSolver.cpp:82-131 (+ the SST geometric factor of KOmegaSST.cpp:399-427),
// in the structured layout.
void Engine::precomputeGeometry() {
  const Grid& g = grid_;
  const int nI = g.nI, nJ = g.nJ;
  auto centroid = [&](int i, int j) { return g.centroid[static_cast<size_t>(j * nI + i)]; };
  auto make = [&](Vec2 n, double len, Vec2 centre, Vec2 co, int kindOf, Vec2 cn, double& gSst) {
    // kindOf: 0 interior (cn = neighbour centroid), 1 wake (cn = across), 2 boundary
    FaceGeo f{};
    const Vec2 S = n * len;
    f.Sx = S.x;
    f.Sy = S.y;
    Vec2 d;
    if (kindOf == 0) {
      d = cn - co;
      const double toO = distance(centre, co), toN = distance(centre, cn);
      const double total = toO + toN;
      f.w = (total > 0.0) ? toN / total : 0.5;
    } else if (kindOf == 1) {
      d = cn - co;
      const double toO = distance(centre, co), toA = distance(centre, cn);
      const double span = toO + toA;
      f.w = (span > 0.0) ? toA / span : 0.5;
    } else {
      d = centre - co;
      f.w = 1.0;
    }
    const double ad = dot(S, d), aa = dot(S, S);
    if (std::abs(ad) > kTiny) {
      f.g = aa / ad;
      const Vec2 T = S - d * f.g;
      f.Tx = T.x;
      f.Ty = T.y;
    } else {
      f.g = 0.0;
      f.Tx = f.Ty = 0.0;
    }
    gSst = (std::abs(ad) > kTinySst) ? std::abs(aa / ad) : 0.0;
    return f;
  };
  jg_.resize(static_cast<size_t>(g.jFaces()));
  jgSst_.assign(static_cast<size_t>(g.jFaces()), 0.0);
  ig_.resize(static_cast<size_t>(g.iFaces()));
  igSst_.assign(static_cast<size_t>(g.iFaces()), 0.0);
  bcB_.assign(static_cast<size_t>(nI), FaceBC{});
  bcT_.assign(static_cast<size_t>(nI), FaceBC{});
  bcL_.assign(static_cast<size_t>(nJ), FaceBC{});
  bcR_.assign(static_cast<size_t>(nJ), FaceBC{});
  auto fillBC = [&](FaceBC& bc, int refFace, double gs) {
    bc.kind = kind_[static_cast<size_t>(refFace)];
    bc.u = bcVel_[static_cast<size_t>(refFace)].x;
    bc.v = bcVel_[static_cast<size_t>(refFace)].y;
    bc.p = bcP_[static_cast<size_t>(refFace)];
    bc.dead = 0.0;
    bc.gSst = gs;
  };
  for (int j = 0; j <= nJ; ++j) {
    for (int i = 0; i < nI; ++i) {
      const size_t f = static_cast<size_t>(j * nI + i);
      double gs = 0.0;
      if (j > 0 && j < nJ) {
        jg_[f] = make(g.jNormal[f], g.jLen[f], g.jCentre[f], centroid(i, j - 1), 0, centroid(i, j), gs);
        jgSst_[f] = gs;
      } else if (j == 0) {
        const int pr = g.bottomPartner[static_cast<size_t>(i)];
        if (pr >= 0) {
          jg_[f] = make(g.jNormal[f], g.jLen[f], g.jCentre[f], centroid(i, 0), 1, centroid(pr, 0), gs);
          jgSst_[f] = gs;
          fillBC(bcB_[static_cast<size_t>(i)], g.refJFace(i, 0), 0.0);
        } else {
          jg_[f] = make(g.jNormal[f], g.jLen[f], g.jCentre[f], centroid(i, 0), 2, Vec2{}, gs);
          fillBC(bcB_[static_cast<size_t>(i)], g.refJFace(i, 0), gs);
        }
      } else {
        jg_[f] = make(g.jNormal[f], g.jLen[f], g.jCentre[f], centroid(i, nJ - 1), 2, Vec2{}, gs);
        fillBC(bcT_[static_cast<size_t>(i)], g.refJFace(i, nJ), gs);
      }
    }
  }
  for (int j = 0; j < nJ; ++j) {
    for (int i = 0; i <= nI; ++i) {
      const size_t f = static_cast<size_t>(j * (nI + 1) + i);
      double gs = 0.0;
      if (i > 0 && i < nI) {
        ig_[f] = make(g.iNormal[f], g.iLen[f], g.iCentre[f], centroid(i - 1, j), 0, centroid(i, j), gs);
        igSst_[f] = gs;
      } else if (i == 0) {
        ig_[f] = make(g.iNormal[f], g.iLen[f], g.iCentre[f], centroid(0, j), 2, Vec2{}, gs);
        fillBC(bcL_[static_cast<size_t>(j)], g.refIFace(0, j), gs);
      } else {
        ig_[f] = make(g.iNormal[f], g.iLen[f], g.iCentre[f], centroid(nI - 1, j), 2, Vec2{}, gs);
        fillBC(bcR_[static_cast<size_t>(j)], g.refIFace(nI, j), gs);
      }
    }
  }
}

namespace {

// Every device array of one instance, carved in a fixed order; a sizing pass
// (base == nullptr) gives the arena's byte count.
void carveInst(Carver& cv, Inst& I, const Grid& grid) {
  const size_t N = static_cast<size_t>(grid.cells());
  const size_t nJf = static_cast<size_t>(grid.jFaces()), nIf = static_cast<size_t>(grid.iFaces());
  const size_t nI = static_cast<size_t>(grid.nI), nJ = static_cast<size_t>(grid.nJ);
  {
    I.V = cv.take<double>(N);
    for (FaceGeoArrays* a : {&I.jg, &I.ig}) {
      const size_t n = (a == &I.jg) ? nJf : nIf;
      a->Sx = cv.take<double>(n);
      a->Sy = cv.take<double>(n);
      a->Tx = cv.take<double>(n);
      a->Ty = cv.take<double>(n);
      a->g = cv.take<double>(n);
      a->w = cv.take<double>(n);
    }
    I.partner = cv.take<int>(nI);
    I.bcB = cv.take<FaceBC>(nI);
    I.bcT = cv.take<FaceBC>(nI);
    I.bcL = cv.take<FaceBC>(nJ);
    I.bcR = cv.take<FaceBC>(nJ);
    I.wallDist = cv.take<double>(N);
    I.jgSst = cv.take<double>(nJf);
    I.igSst = cv.take<double>(nIf);
    double** cellArrays[] = {&I.u, &I.v, &I.p, &I.rho, &I.mu, &I.muMol, &I.aP, &I.aS, &I.aN, &I.aW, &I.aE,
                             &I.bx, &I.by, &I.pP, &I.pN, &I.pE, &I.pWk, &I.pb, &I.x, &I.r, &I.z, &I.pd0,
                             &I.pd1, &I.Ap, &I.gsTmp0, &I.gsTmp1, &I.k, &I.w, &I.mut, &I.f1, &I.f2, &I.cd, &I.strain, &I.prod,
                             &I.kP, &I.kS, &I.kN, &I.kW, &I.kE, &I.kB, &I.wP, &I.wS, &I.wN, &I.wW, &I.wE, &I.wB};
    for (double** a : cellArrays) *a = cv.take<double>(N);
    double2** vecArrays[] = {&I.gU, &I.gV, &I.gP, &I.gK, &I.gW};
    for (double2** a : vecArrays) *a = cv.take<double2>(N);
    I.Fj = cv.take<double>(nJf);
    I.Fi = cv.take<double>(nIf);
    float** fArrays[] = {&I.fpP, &I.fpN, &I.fpE, &I.fpWk, &I.fr, &I.fz, &I.fpd0, &I.fpd1, &I.fAp, &I.flInv, &I.flUp,
                         &I.fx};
    for (float** a : fArrays) *a = cv.take<float>(N);
    double** lineArrays[] = {&I.lInv, &I.lUp, &I.rb, &I.xacc};
    for (double** a : lineArrays) *a = cv.take<double>(N);
    for (float*& a : I.fMom) a = cv.take<float>(N);
    for (float*& a : I.fMomRes) a = cv.take<float>(N);
    for (float*& a : I.fMomDelta) a = cv.take<float>(N);
    for (auto& sys : I.fTurb)
      for (float*& a : sys) a = cv.take<float>(N);
    for (float*& a : I.fTurbRes) a = cv.take<float>(N);
    for (float*& a : I.fTurbDelta) a = cv.take<float>(N);
    const size_t skew = static_cast<size_t>(skewLength(grid.nI, grid.nJ, gsThreadsFor(grid.nJ)));
    for (double2*& a : I.skMomRow) a = cv.take<double2>(skew);
    I.skMomBy = cv.take<double>(skew);
    for (double2*& a : I.skMomX) a = cv.take<double2>(skew);
    for (auto& sys : I.skTurbRow)
      for (double2*& a : sys) a = cv.take<double2>(skew);
    for (auto& sys : I.skTurbX)
      for (double2*& a : sys) a = cv.take<double2>(skew);
    I.red = cv.take<double>(kSlots);
    I.partials = cv.take<double>(static_cast<size_t>(kSlots) * kMaxRedBlocks);
    I.counter = cv.take<unsigned>(16);
    I.flags = cv.take<int>(kIFlags);
    I.monitor = cv.take<DevMonitor>(1);
    I.historyCap = 0;
    I.history = nullptr;
  }
}

}  // namespace

size_t Engine::deviceBytesFor(const Grid& grid) {
  Carver sizing;
  Inst dummy{};
  carveInst(sizing, dummy, grid);
  return sizing.offset + sizeof(Inst);
}

void Engine::allocate() {
  Carver sizing;
  Inst dummy{};
  carveInst(sizing, dummy, grid_);
  arena_.bytes = sizing.offset;
  check(hipMalloc(&arena_.base, arena_.bytes), "hipMalloc (solver state)");
  check(hipMemsetAsync(arena_.base, 0, arena_.bytes, stream_), "hipMemset");
  Carver real;
  real.base = static_cast<char*>(arena_.base);
  carveInst(real, h_, grid_);
  h_.nI = grid_.nI;
  h_.nJ = grid_.nJ;
  h_.N = grid_.cells();
  h_.nBlocks = (h_.N + kBlock - 1) / kBlock;
  h_.cgBlocks = (h_.N + kBlock * kCgCells - 1) / (kBlock * kCgCells);
  h_.colBlocks = (h_.nI + kColThreads - 1) / kColThreads;
  h_.gsR = gsThreadsFor(grid_.nJ);
  h_.gsSkewLen = skewLength(grid_.nI, grid_.nJ, h_.gsR);
  check(hipMalloc(&dInst_, sizeof(Inst)), "hipMalloc (descriptor)");
  cfg_.maxBlocks = h_.nBlocks;
  cfg_.maxCgBlocks = h_.cgBlocks;
  cfg_.maxColBlocks = h_.colBlocks;
  cfg_.maxRbBlocks = rbBlocks(h_.nI, h_.nJ);
  cfg_.maxNJ = h_.nJ;
  cfg_.maxTilesA = tilesA(h_.nI, h_.nJ);
  cfg_.batch = 1;
  cfg_.gsThreads = gsThreadsFor(grid_.nJ);
  cfg_.gsLds = gsLdsFor(grid_.nJ, grid_.nI, cfg_.gsThreads);
  if (cfg_.gsThreads != h_.gsR) throw std::logic_error("sweep lanes differ from the skewed layout's stride");
}

void Engine::uploadStatic() {
  const size_t N = static_cast<size_t>(grid_.cells());
  auto up = [&](void* dst, const void* src, size_t bytes) {
    check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_), "upload");
  };
  up(const_cast<double*>(h_.V), grid_.area.data(), N * sizeof(double));
  // each field of the face geometry as an array of its own (staged in a host
  // buffer that lives until the stream is synchronised below)
  std::vector<double> fields;
  for (const std::vector<FaceGeo>* src : {&jg_, &ig_}) {
    const size_t n = src->size();
    const size_t at = fields.size();
    fields.resize(at + 6 * n);
    for (size_t f = 0; f < n; ++f) {
      const FaceGeo& g = (*src)[f];
      fields[at + f] = g.Sx;
      fields[at + n + f] = g.Sy;
      fields[at + 2 * n + f] = g.Tx;
      fields[at + 3 * n + f] = g.Ty;
      fields[at + 4 * n + f] = g.g;
      fields[at + 5 * n + f] = g.w;
    }
  }
  {
    size_t at = 0;
    for (const FaceGeoArrays* a : {&h_.jg, &h_.ig}) {
      const size_t n = (a == &h_.jg) ? jg_.size() : ig_.size();
      const double* const dst[6] = {a->Sx, a->Sy, a->Tx, a->Ty, a->g, a->w};
      for (int k = 0; k < 6; ++k) up(const_cast<double*>(dst[k]), fields.data() + at + k * n, n * sizeof(double));
      at += 6 * n;
    }
  }
  up(const_cast<int*>(h_.partner), grid_.bottomPartner.data(), grid_.bottomPartner.size() * sizeof(int));
  up(const_cast<FaceBC*>(h_.bcB), bcB_.data(), bcB_.size() * sizeof(FaceBC));
  up(const_cast<FaceBC*>(h_.bcT), bcT_.data(), bcT_.size() * sizeof(FaceBC));
  up(const_cast<FaceBC*>(h_.bcL), bcL_.data(), bcL_.size() * sizeof(FaceBC));
  up(const_cast<FaceBC*>(h_.bcR), bcR_.data(), bcR_.size() * sizeof(FaceBC));
  up(const_cast<double*>(h_.jgSst), jgSst_.data(), jgSst_.size() * sizeof(double));
  up(const_cast<double*>(h_.igSst), igSst_.data(), igSst_.size() * sizeof(double));
  check(hipStreamSynchronize(stream_), "upload sync");
}

void Engine::fillInstScalars() {
  h_.hasPRef = hasPRef_ ? 1 : 0;
  h_.scheme = settings_.scheme;
  h_.model = model_;
  h_.alphaU = settings_.velocity_relaxation;
  h_.alphaP = settings_.pressure_relaxation;
  h_.blend = settings_.non_orthogonal_blend;
  h_.alphaT = settings_.turbulence_relaxation;
  h_.momSweeps = settings_.momentum_sweeps;
  h_.cgMax = settings_.pressure_iterations;
  cfg_.cgIters = h_.cgMax;
  applyTuning(cfg_, grid_.cells());
  h_.cgTol = settings_.pressure_tolerance;
  h_.nonOrthPasses = std::max(1, settings_.non_orthogonal_correctors);
  h_.gsOrder = settings_.gs_ordering;
  h_.precision = settings_.precision;
  h_.pSolver = settings_.pressure_solver;
  h_.refinePasses = std::max(1, settings_.refinement_passes);
  h_.refMassFlow = refMassFlow_;
  h_.kIn = kIn_;
  h_.wIn = wIn_;
  h_.sigK1 = sst_.sigma_k1; h_.sigW1 = sst_.sigma_w1; h_.beta1 = sst_.beta1;
  h_.sigK2 = sst_.sigma_k2; h_.sigW2 = sst_.sigma_w2; h_.beta2 = sst_.beta2;
  h_.betaStar = sst_.beta_star; h_.kappa = sst_.kappa; h_.a1 = sst_.a1; h_.prodLimit = sst_.production_limit;
  // SSTConstants::gamma1/gamma2 (KOmegaSST.cpp:96-102) and beta*^(1/4) (:90)
  h_.gamma1 = sst_.beta1 / sst_.beta_star - sst_.sigma_w1 * sst_.kappa * sst_.kappa / std::sqrt(sst_.beta_star);
  h_.gamma2 = sst_.beta2 / sst_.beta_star - sst_.sigma_w2 * sst_.kappa * sst_.kappa / std::sqrt(sst_.beta_star);
  h_.betaStarQ = std::pow(sst_.beta_star, 0.25);
}

void Engine::uploadInst() {
  fillInstScalars();
  check(hipMemcpyAsync(dInst_, &h_, sizeof(Inst), hipMemcpyHostToDevice, stream_), "upload descriptor");
  check(hipStreamSynchronize(stream_), "descriptor sync");
}

void Engine::checkOrdering(const cfdgpu_settings& s) const {
  constexpr size_t kLdsLimit = 64 * 1024;
  if (s.gs_ordering == CFDGPU_GS_LEXICOGRAPHIC && cfg_.gsLds > kLdsLimit)
    throw std::invalid_argument("mesh too wide for the lexicographic sweep (row 0 must fit in 64 KiB of LDS); "
                                "use CFDGPU_GS_REDBLACK");
}

void Engine::setSettings(const cfdgpu_settings& s) {
  checkOrdering(s);
  settings_ = s;
  if (initialised_) uploadInst();
}

// SimpleSolver.cpp:212-230
int Engine::setModel(int model, const cfdgpu_sst_constants* constants) {
  if (model < CFDGPU_MODEL_NONE || model > CFDGPU_MODEL_SST) return CFDGPU_INVALID_ARGUMENT;
  model_ = model;
  if (constants) sst_ = *constants;
  else cfdgpu_default_sst(&sst_);
  if (!initialised_) return CFDGPU_OK;
  // attached after initialise: bring the model up on the current field
  const size_t N = static_cast<size_t>(grid_.cells());
  std::vector<double> u(N), v(N), p(N);
  getField(u.data(), v.data(), p.data(), nullptr);
  u_ = u;
  v_ = v;
  p_ = p;
  if (model_ == CFDGPU_MODEL_NONE) {
    mu_ = muMol_;
  } else if (model_ == CFDGPU_MODEL_SST) {
    if (!sstInitialise()) return CFDGPU_INVALID_ARGUMENT;
    applyEddyViscosity();
  }
  check(hipMemcpyAsync(h_.mu, mu_.data(), N * sizeof(double), hipMemcpyHostToDevice, stream_), "upload mu");
  if (model_ == CFDGPU_MODEL_SST) {
    check(hipMemcpyAsync(h_.k, k_.data(), N * sizeof(double), hipMemcpyHostToDevice, stream_), "upload k");
    check(hipMemcpyAsync(h_.w, w_.data(), N * sizeof(double), hipMemcpyHostToDevice, stream_), "upload w");
    check(hipMemcpyAsync(h_.mut, mut_.data(), N * sizeof(double), hipMemcpyHostToDevice, stream_), "upload mut");
    check(hipMemcpyAsync(const_cast<double*>(h_.wallDist), wallDist_.data(), N * sizeof(double),
                         hipMemcpyHostToDevice, stream_), "upload d");
  }
  uploadInst();
  return CFDGPU_OK;
}

// KOmegaSST.cpp:104-200 with wallDistances (TurbulenceModel.cpp:28-65)
bool Engine::sstInitialise() {
  if (!std::isfinite(settings_.intensity) || settings_.intensity < 0.0) return false;
  if (!std::isfinite(settings_.viscosity_ratio) || settings_.viscosity_ratio <= 0.0) return false;
  if (!std::isfinite(settings_.length_scale) || settings_.length_scale < 0.0) return false;
  const Grid& g = grid_;
  const size_t N = static_cast<size_t>(g.cells());
  wallDist_.assign(N, 0.0);
  std::vector<Vec2> walls;
  auto gather = [&](int refFace, Vec2 centre) {
    if (kind_[static_cast<size_t>(refFace)] == CFDGPU_K_WALL) walls.push_back(centre);
  };
  // reference face order: j-family (row-major), then i-family (column-major)
  for (int j = 0; j <= g.nJ; ++j)
    for (int i = 0; i < g.nI; ++i) gather(g.refJFace(i, j), g.jCentre[static_cast<size_t>(j * g.nI + i)]);
  for (int i = 0; i <= g.nI; ++i)
    for (int j = 0; j < g.nJ; ++j) gather(g.refIFace(i, j), g.iCentre[static_cast<size_t>(j * (g.nI + 1) + i)]);
  if (walls.empty()) {
    const double far = 1.0e6 * std::sqrt(std::abs(g.totalArea) + 1.0);
    std::fill(wallDist_.begin(), wallDist_.end(), far);
  } else {
    // brute force on the device (O(cells x wall faces)); set-up only
    std::vector<double2> cxy(N), wxy(walls.size());
    for (size_t c = 0; c < N; ++c) cxy[c] = make_double2(g.centroid[c].x, g.centroid[c].y);
    for (size_t w = 0; w < walls.size(); ++w) wxy[w] = make_double2(walls[w].x, walls[w].y);
    double2 *dC = nullptr, *dW = nullptr;
    double* dOut = nullptr;
    check(hipMalloc(&dC, N * sizeof(double2)), "hipMalloc (wall distance)");
    check(hipMalloc(&dW, wxy.size() * sizeof(double2)), "hipMalloc (wall distance)");
    check(hipMalloc(&dOut, N * sizeof(double)), "hipMalloc (wall distance)");
    check(hipMemcpyAsync(dC, cxy.data(), N * sizeof(double2), hipMemcpyHostToDevice, stream_), "upload");
    check(hipMemcpyAsync(dW, wxy.data(), wxy.size() * sizeof(double2), hipMemcpyHostToDevice, stream_), "upload");
    launchWallDistance(dC, static_cast<int>(N), dW, static_cast<int>(wxy.size()), dOut, stream_);
    check(hipGetLastError(), "launch (wall distance)");
    check(hipMemcpyAsync(wallDist_.data(), dOut, N * sizeof(double), hipMemcpyDeviceToHost, stream_), "download");
    check(hipStreamSynchronize(stream_), "wall distance sync");
    (void)hipFree(dC);
    (void)hipFree(dW);
    (void)hipFree(dOut);
  }
  double refSpeed = 0.0, refRho = 0.0, refMu = 0.0;
  for (size_t c = 0; c < N; ++c) {
    const double sp = std::hypot(u_[c], v_[c]);
    if (sp > refSpeed) refSpeed = sp;
    refRho = std::max(refRho, rho_[c]);
    refMu = std::max(refMu, mu_[c]);
  }
  if (!(refSpeed > 0.0) || !(refRho > 0.0) || !(refMu > 0.0)) return false;
  const double intensity = std::max(settings_.intensity, 1.0e-8);
  kIn_ = 1.5 * (intensity * refSpeed) * (intensity * refSpeed);
  double eddy = 0.0;
  if (settings_.length_scale > 0.0) {
    wIn_ = std::max(std::sqrt(kIn_) / (std::pow(sst_.beta_star, 0.25) * settings_.length_scale), kMinOmega);
    eddy = refRho * kIn_ / wIn_;
  } else {
    eddy = settings_.viscosity_ratio * refMu;
    wIn_ = std::max(refRho * kIn_ / eddy, kMinOmega);
  }
  k_.assign(N, kIn_);
  w_.assign(N, wIn_);
  mut_.assign(N, eddy);
  const double bq = std::pow(sst_.beta_star, 0.25);
  for (size_t c = 0; c < N; ++c) {
    const double d = std::max(wallDist_[c], kTinySst);
    const double nu = mu_[c] / std::max(rho_[c], kTinySst);
    // wallOmega (KOmegaSST.cpp:86-92)
    const double dd = std::max(d, kTinySst);
    const double viscous = 6.0 * nu / (sst_.beta1 * dd * dd);
    const double logarithmic = std::sqrt(std::max(k_[c], 0.0)) / (bq * sst_.kappa * dd);
    const double wall = std::sqrt(viscous * viscous + logarithmic * logarithmic);
    w_[c] = std::max(wIn_, wall);
    mut_[c] = rho_[c] * k_[c] / w_[c];
  }
  return true;
}

// SimpleSolver.cpp:232-245
void Engine::applyEddyViscosity() {
  if (model_ == CFDGPU_MODEL_SST) {
    for (size_t c = 0; c < mu_.size(); ++c) mu_[c] = muMol_[c] + mut_[c];
  } else if (model_ == CFDGPU_MODEL_LAMINAR) {
    for (size_t c = 0; c < mu_.size(); ++c) mu_[c] = muMol_[c] + 0.0;
  }
}

// SimpleSolver.cpp:133-210: everything here is per-solve setup in FP64 on the host.
int Engine::initialise(const double* u, const double* v, const double* p, const double* rho, const double* mu) {
  const Grid& g = grid_;
  const int nI = g.nI, nJ = g.nJ;
  const size_t N = static_cast<size_t>(g.cells());
  u_.assign(u, u + N);
  v_.assign(v, v + N);
  p_.assign(p, p + N);
  rho_.assign(rho, rho + N);
  mu_.assign(mu, mu + N);
  muMol_ = mu_;
  // far-field dead band from the starting field (:150-166)
  double refSpeed = 0.0;
  for (size_t c = 0; c < N; ++c) refSpeed = std::max(refSpeed, std::hypot(u_[c], v_[c]));
  auto deadFor = [&](FaceBC& bc, int owner, const FaceGeo& f) {
    bc.dead = 0.0;
    if (bc.kind != CFDGPU_K_FARFIELD) return;
    bc.dead = 1.0e-3 * rho_[static_cast<size_t>(owner)] * refSpeed * std::hypot(f.Sx, f.Sy);
  };
  for (int i = 0; i < nI; ++i) {
    deadFor(bcB_[static_cast<size_t>(i)], i, jg_[static_cast<size_t>(i)]);
    deadFor(bcT_[static_cast<size_t>(i)], (nJ - 1) * nI + i, jg_[static_cast<size_t>(nJ * nI + i)]);
  }
  for (int j = 0; j < nJ; ++j) {
    deadFor(bcL_[static_cast<size_t>(j)], j * nI, ig_[static_cast<size_t>(j * (nI + 1))]);
    deadFor(bcR_[static_cast<size_t>(j)], j * nI + nI - 1, ig_[static_cast<size_t>(j * (nI + 1) + nI)]);
  }
  // face values with zero flux, then the seeded flux F = rho_f (u_f . S) (:168-183)
  Fj_.assign(static_cast<size_t>(g.jFaces()), 0.0);
  Fi_.assign(static_cast<size_t>(g.iFaces()), 0.0);
  auto imposesHost = [](const FaceBC& bc, double F) {
    return bc.kind == CFDGPU_K_WALL || bc.kind == CFDGPU_K_INLET || (bc.kind == CFDGPU_K_FARFIELD && F < bc.dead);
  };
  auto seedInterior = [&](const FaceGeo& f, size_t o, size_t n, bool trueNeighbour) {
    const double w = f.w;
    const double fu = w * u_[o] + (1.0 - w) * u_[n];
    const double fv = w * v_[o] + (1.0 - w) * v_[n];
    double dens = rho_[o];
    if (trueNeighbour) dens = w * rho_[o] + (1.0 - w) * rho_[n];
    return dens * (fu * f.Sx + fv * f.Sy);
  };
  auto seedBoundary = [&](const FaceGeo& f, const FaceBC& bc, size_t o) {
    const bool imp = imposesHost(bc, 0.0);
    const double fu = imp ? bc.u : u_[o];
    const double fv = imp ? bc.v : v_[o];
    return rho_[o] * (fu * f.Sx + fv * f.Sy);
  };
  for (int j = 0; j <= nJ; ++j) {
    for (int i = 0; i < nI; ++i) {
      const size_t f = static_cast<size_t>(j * nI + i);
      if (j > 0 && j < nJ) Fj_[f] = seedInterior(jg_[f], static_cast<size_t>((j - 1) * nI + i), static_cast<size_t>(j * nI + i), true);
      else if (j == 0) {
        const int pr = g.bottomPartner[static_cast<size_t>(i)];
        if (pr >= 0) Fj_[f] = seedInterior(jg_[f], static_cast<size_t>(i), static_cast<size_t>(pr), false);
        else Fj_[f] = seedBoundary(jg_[f], bcB_[static_cast<size_t>(i)], static_cast<size_t>(i));
      } else {
        Fj_[f] = seedBoundary(jg_[f], bcT_[static_cast<size_t>(i)], static_cast<size_t>((nJ - 1) * nI + i));
      }
    }
  }
  for (int j = 0; j < nJ; ++j) {
    for (int i = 0; i <= nI; ++i) {
      const size_t f = static_cast<size_t>(j * (nI + 1) + i);
      if (i > 0 && i < nI) Fi_[f] = seedInterior(ig_[f], static_cast<size_t>(j * nI + i - 1), static_cast<size_t>(j * nI + i), true);
      else if (i == 0) Fi_[f] = seedBoundary(ig_[f], bcL_[static_cast<size_t>(j)], static_cast<size_t>(j * nI));
      else Fi_[f] = seedBoundary(ig_[f], bcR_[static_cast<size_t>(j)], static_cast<size_t>(j * nI + nI - 1));
    }
  }
  // inflow through the boundary, for normalising continuity (:188-198), reference face order
  refMassFlow_ = 0.0;
  auto addInflow = [&](int refFace, double F) {
    const int type = g.faceType(refFace);
    if (type == kInterior || kind_[static_cast<size_t>(refFace)] == CFDGPU_K_INTERNAL) return;
    refMassFlow_ += std::max(-F, 0.0);
  };
  for (int j = 0; j <= nJ; ++j)
    for (int i = 0; i < nI; ++i) addInflow(g.refJFace(i, j), Fj_[static_cast<size_t>(j * nI + i)]);
  for (int i = 0; i <= nI; ++i)
    for (int j = 0; j < nJ; ++j) addInflow(g.refIFace(i, j), Fi_[static_cast<size_t>(j * (nI + 1) + i)]);
  if (!(refMassFlow_ > 0.0)) refMassFlow_ = 1.0;
  if (model_ == CFDGPU_MODEL_SST) {
    if (!sstInitialise()) return CFDGPU_INVALID_ARGUMENT;
  }
  applyEddyViscosity();
  // upload state
  auto up = [&](void* dst, const void* src, size_t bytes) {
    check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_), "upload state");
  };
  up(h_.u, u_.data(), N * sizeof(double));
  up(h_.v, v_.data(), N * sizeof(double));
  up(h_.p, p_.data(), N * sizeof(double));
  up(h_.rho, rho_.data(), N * sizeof(double));
  up(h_.mu, mu_.data(), N * sizeof(double));
  up(h_.muMol, muMol_.data(), N * sizeof(double));
  up(h_.Fj, Fj_.data(), Fj_.size() * sizeof(double));
  up(h_.Fi, Fi_.data(), Fi_.size() * sizeof(double));
  up(const_cast<FaceBC*>(h_.bcB), bcB_.data(), bcB_.size() * sizeof(FaceBC));
  up(const_cast<FaceBC*>(h_.bcT), bcT_.data(), bcT_.size() * sizeof(FaceBC));
  up(const_cast<FaceBC*>(h_.bcL), bcL_.data(), bcL_.size() * sizeof(FaceBC));
  up(const_cast<FaceBC*>(h_.bcR), bcR_.data(), bcR_.size() * sizeof(FaceBC));
  if (model_ == CFDGPU_MODEL_SST) {
    up(h_.k, k_.data(), N * sizeof(double));
    up(h_.w, w_.data(), N * sizeof(double));
    up(h_.mut, mut_.data(), N * sizeof(double));
    up(const_cast<double*>(h_.wallDist), wallDist_.data(), N * sizeof(double));
  }
  // momentumDiagonal_ starts at 1 (SimpleSolver.cpp:64)
  std::vector<double> ones(N, 1.0);
  up(h_.aP, ones.data(), N * sizeof(double));
  std::vector<int> flags(kIFlags, 0);
  up(h_.flags, flags.data(), flags.size() * sizeof(int));
  check(hipStreamSynchronize(stream_), "initialise sync");
  initialised_ = true;
  h_.stopTol = 0.0;
  h_.stopMaxIter = 0;
  uploadInst();
  return CFDGPU_OK;
}

void Engine::readMonitor(cfdgpu_monitor* out) const { readMonitorOn(stream_, out); }

void Engine::readMonitorOn(hipStream_t s, cfdgpu_monitor* out) const {
  DevMonitor m{};
  check(hipMemcpyAsync(&m, h_.monitor, sizeof(DevMonitor), hipMemcpyDeviceToHost, s), "read monitor");
  check(hipStreamSynchronize(s), "monitor sync");
  out->continuity = m.continuity;
  out->momentum_x = m.momentumX;
  out->momentum_y = m.momentumY;
  out->mass_imbalance = m.massImbalance;
  out->max_divergence = m.maxDivergence;
  out->pressure_iterations = m.pressureIterations;
  out->turb_first = m.turbFirst;
  out->turb_second = m.turbSecond;
  out->min_k = m.minK;
  out->max_k = m.maxK;
  out->min_w = m.minW;
  out->max_w = m.maxW;
  out->min_mut = m.minMut;
  out->max_mut = m.maxMut;
  out->max_eddy_ratio = m.maxEddyRatio;
}

// The lexicographic sweep's waves wait on each other with bounded spins; a
// wait that gave up left wrong values behind, so the call fails.
void checkSweepStall(const int* flags) {
  if (flags[fSweepStall] != 0)
    throw std::runtime_error("the lexicographic sweep stalled (a wave's wait for its neighbour timed out)");
}

namespace {
bool graphsEnabled() {
  static const bool on = [] {
    const char* e = std::getenv("CFDGPU_GRAPHS");
    return !(e && e[0] == '0');
  }();
  return on;
}
}  // namespace

bool IterationGraph::Key::operator==(const Key& o) const {
  return d == o.d && maxBlocks == o.maxBlocks && maxCgBlocks == o.maxCgBlocks && maxColBlocks == o.maxColBlocks &&
         lcgCells == o.lcgCells && colUnroll == o.colUnroll && maxLcgBlocks == o.maxLcgBlocks &&
         maxRbBlocks == o.maxRbBlocks && maxNJ == o.maxNJ && maxTilesA == o.maxTilesA && batch == o.batch && gsThreads == o.gsThreads && gsLds == o.gsLds && gsOrder == o.gsOrder &&
         precision == o.precision && pSolver == o.pSolver && refinePasses == o.refinePasses &&
         momSweeps == o.momSweeps && nonOrthPasses == o.nonOrthPasses && cgIters == o.cgIters && hasPRef == o.hasPRef &&
         model == o.model;
}

IterationGraph::Key IterationGraph::keyOf(const Inst* d, const LaunchCfg& cfg, const Inst& h) {
  return Key{d,           cfg.maxBlocks,   cfg.maxCgBlocks, cfg.maxColBlocks, cfg.lcgCells, cfg.colUnroll,
             cfg.maxLcgBlocks, cfg.maxRbBlocks, cfg.maxNJ, cfg.maxTilesA, cfg.batch, cfg.gsThreads,
             cfg.gsLds,   h.gsOrder,       h.precision,     h.pSolver,        h.refinePasses, h.momSweeps,
             h.nonOrthPasses, cfg.cgIters, h.hasPRef, h.model};
}

namespace {
std::mutex gTuningMutex;
LaunchTuning gTuning;
}  // namespace

LaunchTuning launchTuning() {
  std::lock_guard<std::mutex> lock(gTuningMutex);
  return gTuning;
}

void setLaunchTuning(const LaunchTuning& t) {
  std::lock_guard<std::mutex> lock(gTuningMutex);
  gTuning = t;
}

void applyTuning(LaunchCfg& cfg, int maxCells) {
  const LaunchTuning t = launchTuning();
  cfg.lcgCells = t.lcgCells;
  cfg.colUnroll = t.colUnroll;
  cfg.maxLcgBlocks = (t.lcgCells == 0) ? cfg.maxTilesA : lineBlocksA(maxCells, t.lcgCells);
}

LaunchCfg launchCfgFor(const Inst* insts, size_t count) {
  int lanes = 64, blocks = 1, cgBlocks = 1, colBlocks = 1, rbBlk = 1, maxNI = 0, maxNJ = 1, tiles = 1, cgIters = 0,
      maxN = 1;
  for (size_t q = 0; q < count; ++q) {
    const Inst& b = insts[q];
    cgIters = std::max(cgIters, b.cgMax);
    rbBlk = std::max(rbBlk, rbBlocks(b.nI, b.nJ));
    maxN = std::max(maxN, b.N);
    lanes = std::max(lanes, b.gsR);
    blocks = std::max(blocks, b.nBlocks);
    cgBlocks = std::max(cgBlocks, b.cgBlocks);
    colBlocks = std::max(colBlocks, b.colBlocks);
    maxNI = std::max(maxNI, b.nI);
    maxNJ = std::max(maxNJ, b.nJ);
    tiles = std::max(tiles, tilesA(b.nI, b.nJ));
  }
  LaunchCfg cfg{blocks, cgBlocks, colBlocks, cgIters, 2, 32, 1, rbBlk, maxNJ, tiles, static_cast<int>(count), lanes,
                gsLdsFor(0, maxNI, lanes)};
  applyTuning(cfg, maxN);
  return cfg;
}

std::vector<Inst> largestFirst(const std::vector<Inst>& members) {
  std::vector<Inst> out(members);
  std::stable_sort(out.begin(), out.end(), [](const Inst& a, const Inst& b) { return a.N > b.N; });
  return out;
}

bool sameSchedule(const Inst& b, const Inst& a) {
  return b.model == a.model && b.momSweeps == a.momSweeps &&
         b.nonOrthPasses == a.nonOrthPasses && b.gsOrder == a.gsOrder && b.precision == a.precision &&
         b.pSolver == a.pSolver && b.refinePasses == a.refinePasses && b.hasPRef == a.hasPRef;
}

void IterationGraph::reset() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
}

void IterationGraph::run(const Inst* d, const LaunchCfg& cfg, const Inst& h, hipStream_t s, int count) {
  if (count <= 0) return;
  if (!graphsEnabled()) {
    for (int n = 0; n < count; ++n) launchIteration(d, cfg, h, s);
    return;
  }
  const Key k = keyOf(d, cfg, h);
  if (!exec_ || !(k == key_)) {
    reset();
    check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    launchIteration(d, cfg, h, s);
    const hipError_t launched = hipGetLastError();
    const hipError_t ended = hipStreamEndCapture(s, &graph_);
    check(launched, "launch (capture)");
    check(ended, "hipStreamEndCapture");
    check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
    key_ = k;
  }
  for (int n = 0; n < count; ++n) check(hipGraphLaunch(exec_, s), "hipGraphLaunch");
}

void Engine::checkStall() const {
  int flags[kIFlags] = {};
  check(hipMemcpyAsync(flags, h_.flags, sizeof(flags), hipMemcpyDeviceToHost, stream_), "flags");
  check(hipStreamSynchronize(stream_), "flags sync");
  checkSweepStall(flags);
}

int Engine::iterate(int count, cfdgpu_monitor* out) {
  if (!initialised_) return CFDGPU_INITIALIZATION_FAILURE;
  applyTuning(cfg_, grid_.cells());  // (the launch tuning may have changed)
  // a plain iterate() never stops by itself (SimpleSolver::iterate never fails)
  int zero[2] = {0, 0};
  check(hipMemcpyAsync(h_.flags + fStopped, zero, sizeof(int), hipMemcpyHostToDevice, stream_), "flags");
  graph_.run(dInst_, cfg_, h_, stream_, count);
  check(hipGetLastError(), "launch");
  if (out) readMonitor(out);
  else check(hipStreamSynchronize(stream_), "iterate sync");
  checkStall();
  return CFDGPU_OK;
}

int Engine::iterateTimed(int count, cfdgpu_monitor* out, double* stageMs, int* stageCalls) {
  if (!initialised_) return CFDGPU_INITIALIZATION_FAILURE;
  if (!clock_) clock_ = std::make_unique<StageClock>();
  StageClock& clock = *clock_;
  clock.reset();
  applyTuning(cfg_, grid_.cells());
  int zero[2] = {0, 0};
  check(hipMemcpyAsync(h_.flags + fStopped, zero, sizeof(int), hipMemcpyHostToDevice, stream_), "flags");
  for (int n = 0; n < count; ++n) launchIteration(dInst_, cfg_, h_, stream_, &clock);
  check(hipGetLastError(), "launch");
  check(hipStreamSynchronize(stream_), "iterate sync");
  if (out) readMonitor(out);
  clock.collect(stageMs, stageCalls);
  checkStall();
  return CFDGPU_OK;
}

// SolverWorker::runLoop stop rules (SolverWorker.cpp:161-192), evaluated on
// the device after every outer iteration; the host only polls.
int Engine::solve(const cfdgpu_stop_rule& rule, cfdgpu_solve_report* out) {
  if (!initialised_) return CFDGPU_INITIALIZATION_FAILURE;
  const auto t0 = std::chrono::steady_clock::now();
  int startIter = 0;
  check(hipMemcpyAsync(&startIter, h_.flags + fIteration, sizeof(int), hipMemcpyDeviceToHost, stream_), "flags");
  check(hipStreamSynchronize(stream_), "sync");
  h_.stopTol = rule.tolerance;
  h_.stopMaxIter = (rule.max_iterations > 0) ? startIter + rule.max_iterations : 0;
  uploadInst();
  applyTuning(cfg_, grid_.cells());
  int zero = 0;
  check(hipMemcpyAsync(h_.flags + fStopped, &zero, sizeof(int), hipMemcpyHostToDevice, stream_), "flags");
  const int chunk = std::max(1, rule.check_every);
  int launched = 0;
  int flags[kIFlags] = {};
  for (;;) {
    const int n = (rule.max_iterations > 0) ? std::min(chunk, rule.max_iterations - launched) : chunk;
    graph_.run(dInst_, cfg_, h_, stream_, n);
    launched += n;
    check(hipGetLastError(), "launch");
    check(hipMemcpyAsync(flags, h_.flags, sizeof(flags), hipMemcpyDeviceToHost, stream_), "flags");
    check(hipStreamSynchronize(stream_), "solve sync");
    checkSweepStall(flags);
    if (flags[fStopped]) break;
    if (rule.max_iterations > 0 && launched >= rule.max_iterations) break;
  }
  if (out) {
    out->iterations = flags[fIteration] - startIter;
    out->status = flags[fStopped] ? flags[fStatus] : 1;
    readMonitor(&out->last);
    out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  h_.stopTol = 0.0;
  h_.stopMaxIter = 0;
  uploadInst();
  return CFDGPU_OK;
}

int Engine::surface(double speed, double alphaDeg, double density, double pref, double chord, double* rows,
                    int capacity, int* count, cfdgpu_forces* out) {
  if (!initialised_) return CFDGPU_INITIALIZATION_FAILURE;
  if (!wallBuilt_) {
    std::vector<WallStation> table;
    hasWall_ = wallTable(grid_, table, le_);
    wallBuilt_ = true;
    if (hasWall_) {
      nWall_ = static_cast<int>(table.size());
      check(hipMalloc(&dWall_, sizeof(WallStation) * table.size()), "hipMalloc (wall table)");
      // rows (12 per station), force terms (6 per station), the forces record
      check(hipMalloc(&dSurf_, sizeof(double) * (18 * table.size() + kForceDoubles) + sizeof(int) * kForceInts),
            "hipMalloc (surface)");
      check(hipMemcpyAsync(dWall_, table.data(), sizeof(WallStation) * table.size(), hipMemcpyHostToDevice, stream_),
            "upload wall table");
      check(hipStreamSynchronize(stream_), "wall table sync");
    }
  }
  if (!hasWall_) return CFDGPU_NOT_FOUND;  // fewer than four wall faces (SurfaceData.cpp:27-59)
  const size_t n = static_cast<size_t>(nWall_);
  double* dRows = dSurf_;
  double* dTerms = dRows + 12 * n;
  double* dFd = dTerms + 6 * n;
  int* dFi = reinterpret_cast<int*>(dFd + kForceDoubles);
  const double q = 0.5 * density * speed * speed;
  const double a = alphaDeg * 3.141592653589793238462643383279502884 / 180.0;
  launchSurface(h_.u, h_.v, h_.p, h_.mu, static_cast<const WallStation*>(dWall_), nWall_, le_, q, pref, chord,
                std::cos(a), std::sin(a), dRows, dTerms, dFd, dFi, stream_);
  check(hipGetLastError(), "launch surface");
  std::vector<double> hostRows(12 * n);
  double fd[kForceDoubles];
  int fi[kForceInts];
  check(hipMemcpyAsync(hostRows.data(), dRows, sizeof(double) * hostRows.size(), hipMemcpyDeviceToHost, stream_),
        "download stations");
  check(hipMemcpyAsync(fd, dFd, sizeof(fd), hipMemcpyDeviceToHost, stream_), "download forces");
  check(hipMemcpyAsync(fi, dFi, sizeof(fi), hipMemcpyDeviceToHost, stream_), "download forces");
  check(hipStreamSynchronize(stream_), "surface sync");
  for (size_t r = 0; r < n; ++r)
    if (rows && static_cast<long long>(r + 1) * 12 <= capacity) std::copy_n(hostRows.data() + 12 * r, 12, rows + 12 * r);
  if (count) *count = nWall_;
  if (out) {
    cfdgpu_forces f{};
    f.cl = fd[0];
    f.cd = fd[1];
    f.cm = fd[2];
    f.cd_pressure = fd[3];
    f.cd_friction = fd[4];
    f.lift = fd[5];
    f.drag = fd[6];
    f.pitching_moment = fd[7];
    f.fx_pressure = fd[8];
    f.fy_pressure = fd[9];
    f.fx_friction = fd[10];
    f.fy_friction = fd[11];
    f.upper_separation = fd[12];
    f.lower_separation = fd[13];
    f.stagnation_xc = fd[14];
    f.n_upper = fi[0];
    f.n_lower = fi[1];
    f.stations = fi[2];
    f.upper_separation_face = fi[3];
    f.lower_separation_face = fi[4];
    *out = f;
  }
  return CFDGPU_OK;
}

int Engine::getField(double* u, double* v, double* p, double* mu) const {
  const size_t N = static_cast<size_t>(grid_.cells());
  auto down = [&](double* dst, const double* src) {
    if (dst) check(hipMemcpyAsync(dst, src, N * sizeof(double), hipMemcpyDeviceToHost, stream_), "download");
  };
  down(u, h_.u);
  down(v, h_.v);
  down(p, h_.p);
  down(mu, h_.mu);
  check(hipStreamSynchronize(stream_), "download sync");
  return CFDGPU_OK;
}

int Engine::setField(const double* u, const double* v, const double* p) {
  const size_t N = static_cast<size_t>(grid_.cells());
  auto up = [&](double* dst, const double* src) {
    if (src) check(hipMemcpyAsync(dst, src, N * sizeof(double), hipMemcpyHostToDevice, stream_), "upload");
  };
  up(h_.u, u);
  up(h_.v, v);
  up(h_.p, p);
  check(hipStreamSynchronize(stream_), "upload sync");
  return CFDGPU_OK;
}

// device faces -> reference numbering
int Engine::getMassFlux(double* flux) const {
  const Grid& g = grid_;
  std::vector<double> fj(static_cast<size_t>(g.jFaces())), fi(static_cast<size_t>(g.iFaces()));
  check(hipMemcpyAsync(fj.data(), h_.Fj, fj.size() * sizeof(double), hipMemcpyDeviceToHost, stream_), "flux");
  check(hipMemcpyAsync(fi.data(), h_.Fi, fi.size() * sizeof(double), hipMemcpyDeviceToHost, stream_), "flux");
  check(hipStreamSynchronize(stream_), "flux sync");
  std::copy(fj.begin(), fj.end(), flux);
  for (int i = 0; i <= g.nI; ++i)
    for (int j = 0; j < g.nJ; ++j) flux[g.refIFace(i, j)] = fi[static_cast<size_t>(j * (g.nI + 1) + i)];
  return CFDGPU_OK;
}

int Engine::setMassFlux(const double* flux) {
  const Grid& g = grid_;
  std::vector<double> fi(static_cast<size_t>(g.iFaces()));
  for (int i = 0; i <= g.nI; ++i)
    for (int j = 0; j < g.nJ; ++j) fi[static_cast<size_t>(j * (g.nI + 1) + i)] = flux[g.refIFace(i, j)];
  check(hipMemcpyAsync(h_.Fj, flux, static_cast<size_t>(g.jFaces()) * sizeof(double), hipMemcpyHostToDevice, stream_), "flux");
  check(hipMemcpyAsync(h_.Fi, fi.data(), fi.size() * sizeof(double), hipMemcpyHostToDevice, stream_), "flux");
  check(hipStreamSynchronize(stream_), "flux sync");
  return CFDGPU_OK;
}

int Engine::getTurbulence(double* k, double* omega, double* mut, double* wd) const {
  if (model_ != CFDGPU_MODEL_SST) return CFDGPU_NOT_FOUND;
  const size_t N = static_cast<size_t>(grid_.cells());
  auto down = [&](double* dst, const double* src) {
    if (dst) check(hipMemcpyAsync(dst, src, N * sizeof(double), hipMemcpyDeviceToHost, stream_), "download");
  };
  down(k, h_.k);
  down(omega, h_.w);
  down(mut, h_.mut);
  down(wd, h_.wallDist);
  check(hipStreamSynchronize(stream_), "download sync");
  return CFDGPU_OK;
}

int Engine::setTurbulence(const double* k, const double* omega, const double* mut) {
  if (model_ != CFDGPU_MODEL_SST) return CFDGPU_NOT_FOUND;
  const size_t N = static_cast<size_t>(grid_.cells());
  auto up = [&](double* dst, const double* src) {
    if (src) check(hipMemcpyAsync(dst, src, N * sizeof(double), hipMemcpyHostToDevice, stream_), "upload");
  };
  up(h_.k, k);
  up(h_.w, omega);
  up(h_.mut, mut);
  if (mut) {
    std::vector<double> mu(N);
    for (size_t c = 0; c < N; ++c) mu[c] = muMol_[c] + mut[c];
    check(hipMemcpyAsync(h_.mu, mu.data(), N * sizeof(double), hipMemcpyHostToDevice, stream_), "upload");
  }
  check(hipStreamSynchronize(stream_), "upload sync");
  return CFDGPU_OK;
}

int Engine::turbulenceInfo(double* kIn, double* wIn, double* yplus) const {
  if (kIn) *kIn = kIn_;
  if (wIn) *wIn = wIn_;
  if (yplus) {
    check(hipMemcpyAsync(yplus, h_.red + sMaxYPlus, sizeof(double), hipMemcpyDeviceToHost, stream_), "y+");
    check(hipStreamSynchronize(stream_), "sync");
  }
  return CFDGPU_OK;
}

int Engine::divergence(double* out) const {
  double* d = nullptr;
  const size_t N = static_cast<size_t>(grid_.cells());
  check(hipMallocAsync(reinterpret_cast<void**>(&d), N * sizeof(double), stream_), "hipMalloc");
  int zero = 0;
  check(hipMemcpyAsync(h_.flags + fStopped, &zero, sizeof(int), hipMemcpyHostToDevice, stream_), "flags");
  launchDivergence(dInst_, cfg_, d, stream_);
  check(hipMemcpyAsync(out, d, N * sizeof(double), hipMemcpyDeviceToHost, stream_), "download");
  check(hipFreeAsync(d, stream_), "hipFree");
  check(hipStreamSynchronize(stream_), "sync");
  return CFDGPU_OK;
}

int Engine::divergenceNoAlloc(double* out) const {
  launchDivergenceNoAlloc(dInst_, cfg_, grid_.cells(), out, h_.flags + fStopped, stream_);
  return CFDGPU_OK;
}

int Engine::debugArray(const std::string& name, double* out, int capacity) const {
  const size_t N = static_cast<size_t>(grid_.cells());
  const double* src = nullptr;
  size_t count = N;
  const double2* vec = nullptr;
  // The momentum rows other than the diagonal live where the ordering's sweep
  // reads them: the skewed pairs (lexicographic), the FP32 rows (mixed path,
  // two colours) or the FP64 rows (two colours).
  static const char* const momNames[] = {"momS", "momN", "momW", "momE", "momBx", "momBy"};
  int momIdx = -1;
  for (int q = 0; q < 6; ++q)
    if (name == momNames[q]) momIdx = q;
  if (momIdx >= 0 && settings_.gs_ordering == CFDGPU_GS_LEXICOGRAPHIC) {
    if (static_cast<int>(N) > capacity) return -CFDGPU_OUT_OF_RANGE;
    const size_t L = static_cast<size_t>(h_.gsSkewLen);
    std::vector<double> sk(2 * L);
    // pair (aP, aS) | (aN, aW) | (aE, bx); by on its own
    static const int pairOf[] = {0, 1, 1, 2, 2, -1};
    static const int halfOf[] = {1, 0, 1, 0, 1, 0};
    const int pr = pairOf[momIdx];
    if (pr >= 0) check(hipMemcpyAsync(sk.data(), h_.skMomRow[pr], L * sizeof(double2), hipMemcpyDeviceToHost, stream_),
                       "debug");
    else check(hipMemcpyAsync(sk.data(), h_.skMomBy, L * sizeof(double), hipMemcpyDeviceToHost, stream_), "debug");
    check(hipStreamSynchronize(stream_), "debug sync");
    for (int j = 0; j < grid_.nJ; ++j)
      for (int i = 0; i < grid_.nI; ++i) {
        const size_t at = static_cast<size_t>(skewIndex(i, j, h_.gsR));
        out[static_cast<size_t>(j) * grid_.nI + i] = (pr >= 0) ? sk[2 * at + halfOf[momIdx]] : sk[at];
      }
    return static_cast<int>(N);
  }
  if (momIdx >= 0 && momIdx < 4 && settings_.precision == CFDGPU_PRECISION_MIXED) {
    if (static_cast<int>(N) > capacity) return -CFDGPU_OUT_OF_RANGE;
    std::vector<float> f(N);
    check(hipMemcpyAsync(f.data(), h_.fMom[momIdx + 1], N * sizeof(float), hipMemcpyDeviceToHost, stream_), "debug");
    check(hipStreamSynchronize(stream_), "debug sync");
    for (int j = 0; j < grid_.nJ; ++j)
      for (int i = 0; i < grid_.nI; ++i)
        out[static_cast<size_t>(j) * grid_.nI + i] = f[static_cast<size_t>(rbIndex(i, j, grid_.nI, grid_.nJ))];
    return static_cast<int>(N);
  }
  if (momIdx >= 4 && settings_.precision == CFDGPU_PRECISION_MIXED) return -CFDGPU_NOT_FOUND;
  if (name == "gradU") vec = h_.gU;
  else if (name == "gradV") vec = h_.gV;
  else if (name == "gradP") vec = h_.gP;
  else if (name == "momDiag") src = h_.aP;
  else if (name == "momS") src = h_.aS;
  else if (name == "momN") src = h_.aN;
  else if (name == "momW") src = h_.aW;
  else if (name == "momE") src = h_.aE;
  else if (name == "momBx") src = h_.bx;
  else if (name == "momBy") src = h_.by;
  else if (name == "pDiag") src = h_.pP;
  else if (name == "pN") src = h_.pN;
  else if (name == "pE") src = h_.pE;
  else if (name == "pWake") src = h_.pWk;
  else if (name == "pSource") src = h_.pb;
  else if (name == "correction") src = h_.x;
  else if (name == "strain") src = h_.strain;
  else if (name == "f1") src = h_.f1;
  else if (name == "f2") src = h_.f2;
  else if (name == "crossDiffusion") src = h_.cd;
  else return -CFDGPU_NOT_FOUND;
  if (vec) {
    count = 2 * N;
    if (static_cast<int>(count) > capacity) return -CFDGPU_OUT_OF_RANGE;
    check(hipMemcpyAsync(out, vec, N * sizeof(double2), hipMemcpyDeviceToHost, stream_), "debug");
  } else {
    if (static_cast<int>(count) > capacity) return -CFDGPU_OUT_OF_RANGE;
    check(hipMemcpyAsync(out, src, N * sizeof(double), hipMemcpyDeviceToHost, stream_), "debug");
  }
  check(hipStreamSynchronize(stream_), "debug sync");
  return static_cast<int>(count);
}

}  // namespace cfdgpu

namespace cfdgpu {

Batch::Batch(std::vector<Engine*> members) : members_(std::move(members)) {
  if (members_.empty()) throw std::invalid_argument("a batch needs at least one solver");
  if (members_.size() > 65535) throw std::invalid_argument("too many solvers for one batch");
  for (Engine* e : members_)
    if (!e || !e->ready()) throw std::invalid_argument("every solver of a batch must be initialised");
  std::vector<Inst> all;
  for (Engine* e : members_) all.push_back(e->refreshInst());
  for (const Inst& b : all)
    if (!sameSchedule(b, all.front()))
      throw std::invalid_argument("batched solvers must share closure, sweeps, CG cap, correctors, ordering, "
                                  "precision, pressure solver and pressure anchoring");
  cfg_ = launchCfgFor(all.data(), all.size());
  if (all.front().gsOrder == CFDGPU_GS_LEXICOGRAPHIC && cfg_.gsLds > 64 * 1024)
    throw std::invalid_argument("batch too wide for the lexicographic sweep's LDS; use CFDGPU_GS_REDBLACK");
  check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  check(hipMalloc(&dInsts_, sizeof(Inst) * members_.size()), "hipMalloc (batch descriptors)");
  host_.resize(members_.size());
}

Batch::~Batch() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (dInsts_) (void)hipFree(dInsts_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// Members' own streams drain first; then their current descriptors go up.
void Batch::gather() {
  for (size_t q = 0; q < members_.size(); ++q) {
    check(hipStreamSynchronize(members_[q]->stream()), "member sync");
    host_[q] = members_[q]->refreshInst();
    // every launch takes its schedule from the first member: a member whose
    // settings changed since the batch was made must not run under it
    if (!sameSchedule(host_[q], host_.front()))
      throw std::invalid_argument("a batched solver's closure, sweeps, CG cap, correctors, ordering, precision, "
                                  "pressure solver or anchoring changed after the batch was made");
  }
  devOrder_ = largestFirst(host_);
  check(hipMemcpyAsync(dInsts_, devOrder_.data(), sizeof(Inst) * devOrder_.size(), hipMemcpyHostToDevice, stream_),
        "upload batch descriptors");
  cfg_ = launchCfgFor(host_.data(), host_.size());  // (the launch tuning may have changed)
}

int Batch::iterate(int count, cfdgpu_monitor* out) {
  gather();
  launchResumeAll(dInsts_, size(), stream_);
  graph_.run(dInsts_, cfg_, host_.front(), stream_, count);
  check(hipGetLastError(), "launch");
  check(hipStreamSynchronize(stream_), "batch sync");
  for (Engine* e : members_) e->checkStall();
  if (out)
    for (size_t q = 0; q < members_.size(); ++q) members_[q]->readMonitorOn(stream_, &out[q]);
  return CFDGPU_OK;
}

int Batch::iterateTimed(int count, cfdgpu_monitor* out, double* stageMs, int* stageCalls) {
  gather();
  if (!clock_) clock_ = std::make_unique<StageClock>();
  StageClock& clock = *clock_;
  clock.reset();
  launchResumeAll(dInsts_, size(), stream_);
  for (int n = 0; n < count; ++n) launchIteration(dInsts_, cfg_, host_.front(), stream_, &clock);
  check(hipGetLastError(), "launch");
  check(hipStreamSynchronize(stream_), "batch sync");
  for (Engine* e : members_) e->checkStall();
  clock.collect(stageMs, stageCalls);
  if (out)
    for (size_t q = 0; q < members_.size(); ++q) members_[q]->readMonitorOn(stream_, &out[q]);
  return CFDGPU_OK;
}

// The stop rule of each member, evaluated on the device; finished members
// return at once from every later launch while the others continue.
int Batch::solve(const cfdgpu_stop_rule& rule, cfdgpu_solve_report* out) {
  const auto t0 = std::chrono::steady_clock::now();
  const size_t n = members_.size();
  std::vector<int> start(n, 0);
  std::vector<std::array<int, kIFlags>> flags(n);
  for (size_t q = 0; q < n; ++q) {
    check(hipMemcpyAsync(&start[q], members_[q]->inst().flags + fIteration, sizeof(int), hipMemcpyDeviceToHost,
                         members_[q]->stream()),
          "flags");
    check(hipStreamSynchronize(members_[q]->stream()), "member sync");
    Inst& h = members_[q]->hostInst();
    h.stopTol = rule.tolerance;
    h.stopMaxIter = (rule.max_iterations > 0) ? start[q] + rule.max_iterations : 0;
  }
  // refreshInst() keeps stopTol / stopMaxIter (they are not settings)
  gather();
  launchResumeAll(dInsts_, size(), stream_);
  const int chunk = std::max(1, rule.check_every);
  int launched = 0;
  for (;;) {
    const int m = (rule.max_iterations > 0) ? std::min(chunk, rule.max_iterations - launched) : chunk;
    graph_.run(dInsts_, cfg_, host_.front(), stream_, m);
    launched += m;
    check(hipGetLastError(), "launch");
    for (size_t q = 0; q < n; ++q)
      check(hipMemcpyAsync(flags[q].data(), host_[q].flags, sizeof(int) * kIFlags, hipMemcpyDeviceToHost, stream_),
            "flags");
    check(hipStreamSynchronize(stream_), "batch solve sync");
    bool all = true;
    for (size_t q = 0; q < n; ++q) {
      checkSweepStall(flags[q].data());
      all = all && flags[q][fStopped];
    }
    if (all) break;
    if (rule.max_iterations > 0 && launched >= rule.max_iterations) break;
  }
  const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (size_t q = 0; q < n; ++q) {
    if (out) {
      out[q].iterations = flags[q][fIteration] - start[q];
      out[q].status = flags[q][fStopped] ? flags[q][fStatus] : 1;
      members_[q]->readMonitorOn(stream_, &out[q].last);
      out[q].seconds = seconds;
    }
    Inst& h = members_[q]->hostInst();
    h.stopTol = 0.0;
    h.stopMaxIter = 0;
    members_[q]->uploadInst();
  }
  return CFDGPU_OK;
}

}  // namespace cfdgpu
