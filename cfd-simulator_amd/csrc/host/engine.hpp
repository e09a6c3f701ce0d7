// engine.hpp - host side of the MI355X SIMPLE solver: setup (FP64, on the
// host, once per solve), device residency, and the launch schedule of one
// outer iteration. The C ABI (capi.cpp) and the C++ facade (SimpleSolver.hpp)
// are thin layers over this.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../../include/cfdgpu.h"
#include "../device/layout.hpp"
#include "grid.hpp"

struct cfdgpu_mesh {
  cfdgpu::Grid grid;
};

namespace cfdgpu {

// Kernel launchers (defined in launch.hip).
struct LaunchCfg {
  int maxBlocks;   // grid.x for cell-parallel kernels
  int maxCgBlocks; // grid.x for the CG kernels
  int maxColBlocks; // grid.x for the column-parallel kernels (line-preconditioned CG)
  int cgIters;      // CG iterations launched per solve: the largest cap of the batch (a member
                    // with a smaller cap stops its CG at its own)
  int lcgCells;     // line CG launch A: cells per thread (1, 2, 4), or 0: the LDS-tiled launch
  int colUnroll;    // line CG launch B: rows of a column loaded ahead (16, 32)
  int maxLcgBlocks; // grid.x of the line CG's launch A
  int maxRbBlocks;  // grid.x of the two-colour FP32 sweep (one lane per cell of a colour)
  int maxNJ;        // the tallest member's rows (the line CG's segmented launch B holds up to 256)
  int maxTilesA;    // the most 64 x 16 tiles of a member (the LDS-tiled launch A)
  int batch;       // grid.y
  int gsThreads;   // block size of the lexicographic sweep
  size_t gsLds;    // dynamic LDS of the lexicographic sweep
};
// Optional hook around each stage of an outer iteration (CFDGPU_STAGE_*),
// called on the host as the launches are queued.
struct StageRecorder {
  virtual void begin(int stage, hipStream_t s) = 0;
  virtual void end(int stage, hipStream_t s) = 0;
 protected:
  ~StageRecorder() = default;
};
void launchIteration(const Inst* dInst, const LaunchCfg& cfg, const Inst& h, hipStream_t s,
                     StageRecorder* rec = nullptr);
// The launch shape of a batch of instances (the widest member's grid of each kind).
LaunchCfg launchCfgFor(const Inst* insts, size_t count);
// The order a batch's descriptors are uploaded in (blockIdx.y): the largest
// meshes first, ties in member order, so the hardware dispatches their blocks
// first and the small members' blocks fill the end of each launch.
std::vector<Inst> largestFirst(const std::vector<Inst>& members);
// Process-wide launch tuning of the line CG (cfdgpu_set_launch_tuning); read
// whenever a launch shape is formed, so a change takes effect at the next call
// (a captured iteration graph is captured again).
struct LaunchTuning {
  int lcgCells{4};
  int colUnroll{32};
};
LaunchTuning launchTuning();
void setLaunchTuning(const LaunchTuning& t);
void applyTuning(LaunchCfg& cfg, int maxCells);
// Members of one batch share everything that shapes the launch schedule.
bool sameSchedule(const Inst& a, const Inst& b);
void launchDivergence(const Inst* dInst, const LaunchCfg& cfg, double* out, hipStream_t s);
// launchDivergence staged through the library's static device array (noalloc.hip): no allocation;
// clears the instance's stop flag first, downloads `cells` values into the host buffer `out`
void launchDivergenceNoAlloc(const Inst* dInst, const LaunchCfg& cfg, int cells, double* out, int* flagStopped,
                             hipStream_t s);
// distance from each of n centroids to the nearest of nw wall-face centres
void launchWallDistance(const double2* centroid, int n, const double2* walls, int nw, double* out, hipStream_t s);
struct WallStation;
// surface stations and forces of one field (surface.hip)
void launchSurface(const double* u, const double* v, const double* p, const double* mu, const WallStation* st,
                   int nWall, int le, double q, double pref, double chord, double ca, double sa, double* rows,
                   double* terms, double* fd, int* fi, hipStream_t s);
// clears the stop flag of `count` instances (a plain iterate never stops by itself)
void launchResumeAll(const Inst* dInst, int count, hipStream_t s);
int gsThreadsFor(int nJ);
size_t gsLdsFor(int nJ, int nI, int threads);

class StageClock;
// throws when flags (kIFlags ints of one instance) record a stalled sweep
void checkSweepStall(const int* flags);


// The launches of one outer iteration captured once into a HIP graph and
// replayed: the same kernels with the same arguments in the same order (the
// iteration's scalars and stop flags live on the device, so one capture
// serves every iteration). It is captured again whenever anything the launch
// schedule reads changes: the descriptor array, the launch shape, or the
// settings that pick kernels and loop counts. Set CFDGPU_GRAPHS=0 in the
// environment to queue the launches one by one instead.
class IterationGraph {
 public:
  IterationGraph() = default;
  ~IterationGraph() { reset(); }
  IterationGraph(const IterationGraph&) = delete;
  IterationGraph& operator=(const IterationGraph&) = delete;
  // queue `count` outer iterations on s; h is the host copy of the first
  // instance's descriptor (it drives the launch schedule)
  void run(const Inst* d, const LaunchCfg& cfg, const Inst& h, hipStream_t s, int count);
  void reset();

 private:
  struct Key {
    const Inst* d;
    int maxBlocks, maxCgBlocks, maxColBlocks, lcgCells, colUnroll, maxLcgBlocks, maxRbBlocks, maxNJ, maxTilesA,
        batch, gsThreads;
    size_t gsLds;
    int gsOrder, precision, pSolver, refinePasses, momSweeps, nonOrthPasses, cgIters, hasPRef, model;
    bool operator==(const Key& o) const;
  };
  static Key keyOf(const Inst* d, const LaunchCfg& cfg, const Inst& h);
  Key key_{};
  hipGraph_t graph_{nullptr};
  hipGraphExec_t exec_{nullptr};
};

// Device arena for one instance: every array of Inst carved out of one allocation.
struct DeviceArena {
  void* base{nullptr};
  size_t bytes{0};
  ~DeviceArena();
};

class Engine {
 public:
  Engine(const Grid& grid, std::vector<int> kind, std::vector<Vec2> bcVelocity, std::vector<double> bcPressure,
         const cfdgpu_settings& settings);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  int setModel(int model, const cfdgpu_sst_constants* constants);
  int initialise(const double* u, const double* v, const double* p, const double* rho, const double* mu);
  int iterate(int count, cfdgpu_monitor* out);
  // iterate() with HIP events around every stage; per-stage device time summed
  // over the iterations (stageMs) and how often each stage ran (stageCalls)
  int iterateTimed(int count, cfdgpu_monitor* out, double* stageMs, int* stageCalls);
  int solve(const cfdgpu_stop_rule& rule, cfdgpu_solve_report* out);
  void setSettings(const cfdgpu_settings& s);
  const cfdgpu_settings& settings() const { return settings_; }

  int getField(double* u, double* v, double* p, double* mu) const;
  int setField(const double* u, const double* v, const double* p);
  int getMassFlux(double* flux) const;
  int setMassFlux(const double* flux);
  int getTurbulence(double* k, double* omega, double* mut, double* wd) const;
  int setTurbulence(const double* k, const double* omega, const double* mut);
  int turbulenceInfo(double* kIn, double* wIn, double* yplus) const;
  int divergence(double* out) const;
  int divergenceNoAlloc(double* out) const;  // the same values, staged without allocating
  // surface distribution and forces of the current field, on the device; the
  // same numbers as surfaceAndForces on the downloaded field
  int surface(double speed, double alphaDeg, double density, double pref, double chord, double* rows, int capacity,
              int* count, cfdgpu_forces* out);
  int debugArray(const std::string& name, double* out, int capacity) const;

  const Grid& grid() const { return grid_; }
  bool ready() const { return initialised_; }
  // device memory one solver on this mesh holds (state arena + descriptor)
  static size_t deviceBytesFor(const Grid& grid);
  size_t deviceBytes() const { return arena_.bytes + sizeof(Inst); }
  hipStream_t stream() const { return stream_; }
  // for batching: the instance descriptor with device pointers, refreshed from
  // the current settings, and the monitor read on another stream
  const Inst& inst() const { return h_; }
  const Inst& refreshInst() {
    fillInstScalars();
    return h_;
  }
  Inst& hostInst() { return h_; }
  void readMonitorOn(hipStream_t s, cfdgpu_monitor* out) const;
  void uploadInst();
  // throws when a lexicographic sweep of this solver stalled (see checkSweepStall)
  void checkStall() const;

 private:
  void precomputeGeometry();
  void allocate();
  void uploadStatic();
  bool sstInitialise();
  void applyEddyViscosity();
  void fillInstScalars();
  void checkOrdering(const cfdgpu_settings& s) const;
  void readMonitor(cfdgpu_monitor* out) const;

  const Grid& grid_;
  std::vector<int> kind_;
  std::vector<Vec2> bcVel_;
  std::vector<double> bcP_;
  cfdgpu_settings settings_;
  int model_{CFDGPU_MODEL_NONE};
  cfdgpu_sst_constants sst_{};
  bool initialised_{false};
  bool hasPRef_{false};
  double refMassFlow_{1.0};

  // host-side FP64 setup (SimpleSolver.cpp:82-210)
  std::vector<FaceGeo> jg_, ig_;
  std::vector<double> jgSst_, igSst_;
  std::vector<FaceBC> bcB_, bcT_, bcL_, bcR_;
  std::vector<double> dead_;  // per reference face
  // host mirrors used during initialise
  std::vector<double> u_, v_, p_, rho_, mu_, muMol_, Fj_, Fi_;
  std::vector<double> k_, w_, mut_, wallDist_;
  double kIn_{0.0}, wIn_{0.0};

  DeviceArena arena_;
  Inst h_{};
  Inst* dInst_{nullptr};
  hipStream_t stream_{nullptr};
  LaunchCfg cfg_{};
  IterationGraph graph_;
  std::unique_ptr<StageClock> clock_;
  // device surface integration: the wall table and its scratch, made on first use
  bool wallBuilt_{false}, hasWall_{false};
  int nWall_{0}, le_{0};
  void* dWall_{nullptr};
  double* dSurf_{nullptr};
};

// Several solves advanced together: one launch per stage for all of them
// (grid.y = instance), each instance's state in its own arena. Members must
// share the launch-shaping settings (closure, sweeps, CG cap, corrector
// passes, ordering, pressure anchoring); meshes may differ.
class Batch {
 public:
  explicit Batch(std::vector<Engine*> members);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;
  int size() const { return static_cast<int>(members_.size()); }
  int iterate(int count, cfdgpu_monitor* out);
  // iterate() launched directly with HIP events on the batch's stream around
  // every stage (per-stage device time summed over the iterations)
  int iterateTimed(int count, cfdgpu_monitor* out, double* stageMs, int* stageCalls);
  int solve(const cfdgpu_stop_rule& rule, cfdgpu_solve_report* out);

 private:
  void gather();
  std::unique_ptr<StageClock> clock_;
  std::vector<Engine*> members_;
  std::vector<Inst> host_;     // in member order
  std::vector<Inst> devOrder_;  // as uploaded: largest mesh first
  Inst* dInsts_{nullptr};
  hipStream_t stream_{nullptr};
  LaunchCfg cfg_{};
  IterationGraph graph_;
};

}  // namespace cfdgpu

struct cfdgpu_batch {
  std::unique_ptr<cfdgpu::Batch> batch;
};

struct cfdgpu_solver {
  std::unique_ptr<cfdgpu::Engine> engine;
};
