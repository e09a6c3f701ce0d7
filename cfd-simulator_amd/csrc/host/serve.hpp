// serve.hpp - a request loop for polar-point solves with continuous batching.
//
// Requests (a NACA section, its C-grid, alpha, Re) arrive over time. A request
// is admitted when its arrival time has passed, a slot is free and its device
// state fits the memory budget; admitted solves stay resident in HBM and are
// advanced together, one launch per SIMPLE stage for all of them, in chunks of
// outer iterations. Between chunks finished solves leave (their forces are
// integrated and their memory is returned) and new arrivals join. The
// per-solve semantics are the reference polar's: the worker's stop rule
// (SolverWorker.cpp:161-192), one cold retry after a divergence
// (Polar.cpp:112-117), the point status of Polar.cpp:79-84, and the warm start
// that carries velocity and pressure over from a converged neighbour
// (Panels.cpp:1137-1157), here the nearest converged alpha of the same section
// on a mesh of the same dimensions.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "../../../include/cfdgpu.h"
#include "engine.hpp"
#include "grid.hpp"

namespace cfdgpu {

struct ServeRequest {
  int m{0}, p{0}, t{12};
  double chord{1.0};
  CGridOptions mesh{};
  double alphaDeg{0.0}, reynolds{1e6}, speed{1.0}, density{1.0}, referencePressure{0.0};
  double arrival{0.0};
  bool ownSettings{false};  // settings below instead of the server's
  cfdgpu_settings settings{};
};

struct ServerOptions {
  double budgetBytes{200e9};
  int maxSlots{64};
  int chunk{20};
  cfdgpu_stop_rule rule{1e-6, 10000, 1};
  cfdgpu_settings settings{};
  int model{CFDGPU_MODEL_SST};
  bool warmStart{true};
  bool keepFields{false};
  int settleWindow{0};  // > 0: converged only once Cl and Cd have settled (see cfdgpu_server_options)
  double settleCl{5e-4}, settleCd{5e-5};
};

// The change still to come of a converging sequence from its last three
// samples spaced evenly in iterations (Aitken's geometric tail): |d2| rho /
// (1 - rho) with rho = d2 / d1 for a monotone, shrinking change; |d2| for an
// oscillation; infinity for a change that does not shrink.
double remainingChange(double c0, double c1, double c2);

class Server {
 public:
  explicit Server(const ServerOptions& options);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  int submit(const ServeRequest& r);
  // One scheduling round: queue a chunk for the resident solves, admit what
  // has arrived while it runs, then retire the solves that stopped.
  void step();
  // Rounds until nothing is queued or resident, or until wallLimit seconds
  // of the server clock have passed (<= 0: no limit).
  void run(double wallLimit);
  int collect(cfdgpu_serve_result* out, int capacity);
  bool fields(int ticket, double* u, double* v, double* p, double* mu) const;
  int cellsOf(int ticket) const;  // cells of a kept result, -1 if none
  void stats(cfdgpu_server_stats* out) const;
  double clock() const;

 private:
  struct Slot;
  struct Donor {
    double alpha, reynolds;
    int ticket;
    std::vector<double> u, v, p;
  };
  using MeshKey = std::tuple<int, int, int, double, double, double, double, int, int, int, double, double, int>;
  using DonorKey = std::tuple<int, int, int, double, int, int, double, double>;

  static MeshKey meshKeyOf(const ServeRequest& r);
  std::shared_ptr<const Grid> cachedMesh(const MeshKey& key);
  void keepMesh(const MeshKey& key, std::shared_ptr<const Grid> g);
  void admitReady();
  void reject(int ticket, const std::shared_ptr<const Grid>& grid);
  void startAttempt(Slot& s, bool cold);
  void finish(Slot& s, int status, int iterations, double continuity);
  bool settled(Slot& s, int iteration);
  void launchChunk();
  void retire();

  ServerOptions opt_;
  int device_{0};        // the GPU the server's solves run on
  int admitThreads_{1};  // host threads that build admitted requests' meshes and solvers
  std::chrono::steady_clock::time_point t0_;
  std::vector<ServeRequest> requests_;
  std::deque<int> pending_;  // tickets in arrival order
  std::vector<std::unique_ptr<Slot>> active_;   // the roster of the running chunk
  std::vector<std::unique_ptr<Slot>> joining_;  // admitted during it, join the next
  std::list<std::pair<MeshKey, std::shared_ptr<const Grid>>> meshes_;  // most recent first
  std::map<DonorKey, std::vector<Donor>> donors_;
  std::deque<std::pair<DonorKey, int>> donorOrder_;  // oldest first
  size_t donorBytes_{0};
  std::vector<cfdgpu_serve_result> done_;  // finished, not yet collected
  std::map<int, std::vector<double>> kept_;   // ticket -> u, v, p, mu

  // the resident roster's launch state
  Inst* dInsts_{nullptr};
  std::vector<Inst> host_;
  int* hFlags_{nullptr};  // pinned, kIFlags per slot
  hipStream_t stream_{nullptr};
  LaunchCfg cfg_{};
  IterationGraph graph_;
  bool launched_{false};
  bool rosterDirty_{true};

  double resident_{0.0}, peakResident_{0.0};
  int peakActive_{0};
  long long memberIterations_{0};
  int rounds_{0}, finished_{0}, converged_{0}, admitted_{0};
};

}  // namespace cfdgpu

struct cfdgpu_server {
  std::unique_ptr<cfdgpu::Server> server;
};
