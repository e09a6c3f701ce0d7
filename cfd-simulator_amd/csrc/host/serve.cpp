// serve.cpp - the request loop: admission by memory budget, device-resident
// solves advanced together in chunks, retirement with the reference polar's
// per-point semantics (Polar.cpp:79-124, Panels.cpp:1137-1157, 1723-1730).
#include "serve.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <stdexcept>
#include <string>
#include <thread>

#include "polar.hpp"
#include "post.hpp"

namespace cfdgpu {

namespace {

constexpr size_t kMeshCache = 32;                 // meshes kept for later requests
constexpr size_t kDonorsPerKey = 16;              // converged fields kept per section and mesh size
constexpr size_t kDonorBytes = size_t(2) << 30;   // host bytes of kept warm-start fields

void hipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Runs task(0 .. n-1) on up to `threads` host threads (the calling one
// included), each with `device` current. Exceptions stay with the tasks.
void parallelFor(int n, int threads, int device, const std::function<void(int)>& task) {
  std::atomic<int> next{0};
  auto work = [&](bool setDevice) {
    if (setDevice) (void)hipSetDevice(device);
    for (int q = next++; q < n; q = next++) task(q);
  };
  const int extra = std::max(0, std::min(threads, n) - 1);
  std::vector<std::thread> pool;
  pool.reserve(static_cast<size_t>(extra));
  for (int t = 0; t < extra; ++t) pool.emplace_back(work, true);
  work(false);
  for (std::thread& t : pool) t.join();
}

}  // namespace

struct Server::Slot {
  int ticket{-1};
  std::shared_ptr<const Grid> grid;
  std::unique_ptr<Engine> engine;
  bool retried{false};
  int warmFrom{-1};
  double admitted{0.0};
  size_t bytes{0};
  int lastIteration{0};
  // coefficient settling: Cl and Cd every settle window, the last three kept
  int nextSample{0};
  std::deque<std::pair<double, double>> samples;
  int settledChecks{0};  // consecutive checks that passed
};

double remainingChange(double c0, double c1, double c2) {
  const double d1 = c1 - c0, d2 = c2 - c1;
  if (!std::isfinite(d1) || !std::isfinite(d2)) return INFINITY;
  if (d2 == 0.0) return 0.0;
  if (d1 * d2 < 0.0) return std::fabs(d2);
  if (std::fabs(d2) < std::fabs(d1)) {
    const double rho = d2 / d1;
    return std::fabs(d2) * rho / (1.0 - rho);
  }
  return INFINITY;
}

Server::Server(const ServerOptions& options) : opt_(options), t0_(std::chrono::steady_clock::now()) {
  if (opt_.maxSlots < 1 || opt_.maxSlots > 65535) throw std::invalid_argument("max_slots must be in 1..65535");
  if (opt_.chunk < 1) throw std::invalid_argument("chunk_iterations must be at least 1");
  if (!(opt_.budgetBytes > 0.0)) throw std::invalid_argument("the memory budget must be positive");
  if (opt_.model != CFDGPU_MODEL_NONE && opt_.model != CFDGPU_MODEL_LAMINAR && opt_.model != CFDGPU_MODEL_SST)
    throw std::invalid_argument("unknown turbulence model");
  hipCheck(hipGetDevice(&device_), "hipGetDevice");
  // the host set-up of admitted requests runs on up to 16 threads
  // (CFDGPU_ADMIT_THREADS overrides, 1 builds them one after another)
  const unsigned hw = std::thread::hardware_concurrency();
  admitThreads_ = std::max(1, std::min(16, static_cast<int>(hw ? hw : 1)));
  if (const char* e = std::getenv("CFDGPU_ADMIT_THREADS")) admitThreads_ = std::max(1, std::min(64, std::atoi(e)));
  hipCheck(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  hipCheck(hipMalloc(&dInsts_, sizeof(Inst) * static_cast<size_t>(opt_.maxSlots)), "hipMalloc (roster)");
  hipCheck(hipHostMalloc(reinterpret_cast<void**>(&hFlags_), sizeof(int) * kIFlags * static_cast<size_t>(opt_.maxSlots),
                         hipHostMallocDefault),
           "hipHostMalloc (flags)");
}

Server::~Server() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  active_.clear();
  joining_.clear();
  if (hFlags_) (void)hipHostFree(hFlags_);
  if (dInsts_) (void)hipFree(dInsts_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

double Server::clock() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

int Server::submit(const ServeRequest& r) {
  if (r.m < 0 || r.m > 9 || r.p < 0 || r.p > 9 || r.t < 1 || r.t > 40)
    throw std::invalid_argument("NACA digits out of range");
  if (!(r.chord > 0.0) || !(r.reynolds > 0.0) || !(r.speed > 0.0) || !(r.density > 0.0))
    throw std::invalid_argument("chord, Reynolds number, speed and density must be positive");
  if (!std::isfinite(r.alphaDeg) || !std::isfinite(r.arrival)) throw std::invalid_argument("non-finite request");
  if (r.ownSettings) {
    const cfdgpu_settings& a = r.settings;
    const cfdgpu_settings& b = opt_.settings;
    if (a.momentum_sweeps != b.momentum_sweeps ||
        a.non_orthogonal_correctors != b.non_orthogonal_correctors || a.gs_ordering != b.gs_ordering ||
        a.precision != b.precision || a.pressure_solver != b.pressure_solver ||
        a.refinement_passes != b.refinement_passes)
      throw std::invalid_argument("a request's sweeps, correctors, ordering, precision, pressure solver and "
                                  "refinement passes must be the server's (they shape the shared launches)");
  }
  const int ticket = static_cast<int>(requests_.size());
  requests_.push_back(r);
  // arrival order; ties keep submission order
  auto at = std::upper_bound(pending_.begin(), pending_.end(), r.arrival,
                             [&](double a, int q) { return a < requests_[static_cast<size_t>(q)].arrival; });
  pending_.insert(at, ticket);
  return ticket;
}

// One mesh per (section, C-grid options), built on first use and kept while
// any resident solve refers to it or it is among the most recent.
Server::MeshKey Server::meshKeyOf(const ServeRequest& r) {
  const CGridOptions& o = r.mesh;
  return MeshKey{r.m, r.p, r.t, r.chord, o.upstream, o.downstream, o.vertical, o.surfacePoints,
                 o.wakePoints, o.normalPoints, o.firstLayer, o.blendLength, o.smoothingPasses};
}

std::shared_ptr<const Grid> Server::cachedMesh(const MeshKey& key) {
  for (auto it = meshes_.begin(); it != meshes_.end(); ++it) {
    if (it->first == key) {
      meshes_.splice(meshes_.begin(), meshes_, it);
      return meshes_.front().second;
    }
  }
  return nullptr;
}

void Server::keepMesh(const MeshKey& key, std::shared_ptr<const Grid> g) {
  meshes_.emplace_front(key, std::move(g));
  // drop the oldest meshes no solve holds
  for (auto it = meshes_.end(); meshes_.size() > kMeshCache && it != meshes_.begin();) {
    --it;
    if (it->second.use_count() == 1) it = meshes_.erase(it);
  }
}

// A new solver for the slot's request: the uniform field of its freestream
// (FlowField.cpp:21-47), with velocity and pressure taken from the nearest
// converged alpha of the same section and mesh size unless `cold`; the closure
// is attached and initialised cold either way (Panels.cpp:1274-1306).
void Server::startAttempt(Slot& s, bool cold) {
  const ServeRequest& r = requests_[static_cast<size_t>(s.ticket)];
  const Grid& g = *s.grid;
  const int N = g.cells(), F = g.faces();
  std::vector<int> kind(static_cast<size_t>(F));
  std::vector<double> vxy(2 * static_cast<size_t>(F)), pres(static_cast<size_t>(F));
  defaultFaceConditions(g, r.speed, r.alphaDeg, r.referencePressure, kind.data(), vxy.data(), pres.data());
  std::vector<Vec2> vel(static_cast<size_t>(F));
  for (int f = 0; f < F; ++f) vel[static_cast<size_t>(f)] = Vec2{vxy[2 * static_cast<size_t>(f)], vxy[2 * static_cast<size_t>(f) + 1]};
  std::vector<double> u(static_cast<size_t>(N)), v(u.size()), p(u.size()), rho(u.size()), mu(u.size());
  uniformField(N, r.speed, r.alphaDeg, r.density, r.reynolds, r.chord, r.referencePressure, u.data(), v.data(),
               p.data(), rho.data(), mu.data());
  s.warmFrom = -1;
  if (!cold && opt_.warmStart) {
    const DonorKey key{r.m, r.p, r.t, r.chord, g.nI, g.nJ, r.speed, r.density};
    auto it = donors_.find(key);
    if (it != donors_.end() && !it->second.empty()) {
      const Donor* best = nullptr;
      double bestA = 0.0, bestR = 0.0;
      for (const Donor& d : it->second) {
        const double da = std::fabs(d.alpha - r.alphaDeg);
        const double dr = std::fabs(std::log(d.reynolds / r.reynolds));
        if (!best || da < bestA || (da == bestA && dr < bestR)) {
          best = &d;
          bestA = da;
          bestR = dr;
        }
      }
      u = best->u;
      v = best->v;
      p = best->p;
      s.warmFrom = best->ticket;
    }
  }
  // a retry replaces the diverged solver once the new one is ready
  auto e = std::make_unique<Engine>(g, std::move(kind), std::move(vel), std::move(pres),
                                    r.ownSettings ? r.settings : opt_.settings);
  if (opt_.model != CFDGPU_MODEL_NONE && e->setModel(opt_.model, nullptr) != CFDGPU_OK)
    throw std::invalid_argument("unknown turbulence model");
  if (e->initialise(u.data(), v.data(), p.data(), rho.data(), mu.data()) != CFDGPU_OK)
    throw std::runtime_error("the solver could not start");
  Inst& h = e->hostInst();
  // with coefficient settling the device stops only on divergence and at the
  // limit; the residual test and the settling are checked between chunks
  h.stopTol = opt_.settleWindow > 0 ? 0.0 : opt_.rule.tolerance;
  h.stopMaxIter = opt_.rule.max_iterations > 0 ? opt_.rule.max_iterations : 0;
  hipCheck(hipStreamSynchronize(e->stream()), "solver set-up");
  s.bytes = e->deviceBytes();
  s.lastIteration = 0;
  s.nextSample = opt_.settleWindow;
  s.samples.clear();
  s.settledChecks = 0;
  s.engine = std::move(e);
}

// A request that cannot be served (its device state exceeds the budget, or
// its mesh or solver could not be built) leaves with status 3.
void Server::reject(int ticket, const std::shared_ptr<const Grid>& grid) {
  const ServeRequest& r = requests_[static_cast<size_t>(ticket)];
  cfdgpu_serve_result res{};
  res.ticket = ticket;
  res.status = 3;
  res.warm_from = -1;
  res.upper_separation = res.lower_separation = -1.0;
  res.arrival_s = r.arrival;
  res.admitted_s = res.finished_s = clock();
  res.cells = grid ? grid->cells() : 0;
  done_.push_back(res);
  ++finished_;
}

// Admits, in arrival order, the queued requests that have arrived while slots
// are free and their device state fits the budget (the first that does not
// fit waits for memory, and the ones behind it wait with it). Their meshes
// (the ones not cached) and their solvers (host geometry, device arena,
// uploads, wall distance) are built on host threads side by side: the set-up
// of a 1024 x 256 request takes tens of milliseconds on the host, and a burst
// of arrivals would otherwise hold the next chunk back for the sum of them.
void Server::admitReady() {
  const size_t busy = active_.size() + joining_.size();
  const size_t free = busy < static_cast<size_t>(opt_.maxSlots) ? static_cast<size_t>(opt_.maxSlots) - busy : 0;
  std::vector<int> cand;
  for (const int t : pending_) {
    if (cand.size() >= free || requests_[static_cast<size_t>(t)].arrival > clock()) break;
    cand.push_back(t);
  }
  if (cand.empty()) return;
  // meshes: the cached ones, then the missing distinct keys built in parallel
  const size_t n = cand.size();
  std::vector<MeshKey> keys(n);
  std::vector<std::shared_ptr<const Grid>> grids(n);
  std::vector<int> build;  // index of the first candidate of each missing key
  for (size_t k = 0; k < n; ++k) {
    keys[k] = meshKeyOf(requests_[static_cast<size_t>(cand[k])]);
    grids[k] = cachedMesh(keys[k]);
    if (grids[k]) continue;
    bool dup = false;
    for (const int b : build) dup = dup || keys[static_cast<size_t>(b)] == keys[k];
    if (!dup) build.push_back(static_cast<int>(k));
  }
  std::vector<std::shared_ptr<Grid>> built(build.size());
  parallelFor(static_cast<int>(build.size()), admitThreads_, device_, [&](int b) {
    const ServeRequest& r = requests_[static_cast<size_t>(cand[static_cast<size_t>(build[static_cast<size_t>(b)])])];
    try {
      auto g = std::make_shared<Grid>();
      std::string err;
      if (makeCGrid(*g, r.m, r.p, r.t, r.chord, r.mesh, err)) built[static_cast<size_t>(b)] = g;
    } catch (const std::exception&) {
    }
  });
  for (size_t b = 0; b < build.size(); ++b)
    if (built[b]) keepMesh(keys[static_cast<size_t>(build[b])], built[b]);
  for (size_t k = 0; k < n; ++k)
    if (!grids[k]) grids[k] = cachedMesh(keys[k]);
  // the budget, in arrival order
  std::vector<std::unique_ptr<Slot>> admit;
  double planned = 0.0;
  size_t taken = 0;
  for (size_t k = 0; k < n; ++k) {
    const int ticket = cand[k];
    if (!grids[k]) {  // the mesh could not be built
      reject(ticket, nullptr);
      ++taken;
      continue;
    }
    const double need = static_cast<double>(Engine::deviceBytesFor(*grids[k]));
    if (need > opt_.budgetBytes) {
      reject(ticket, grids[k]);
      ++taken;
      continue;
    }
    if (resident_ + planned + need > opt_.budgetBytes) break;  // waits for memory
    planned += need;
    auto s = std::make_unique<Slot>();
    s->ticket = ticket;
    s->grid = grids[k];
    s->admitted = clock();
    admit.push_back(std::move(s));
    ++taken;
  }
  for (size_t k = 0; k < taken; ++k) pending_.pop_front();
  // the solvers, in parallel
  std::vector<char> ok(admit.size(), 0);
  parallelFor(static_cast<int>(admit.size()), admitThreads_, device_, [&](int q) {
    try {
      startAttempt(*admit[static_cast<size_t>(q)], false);
      ok[static_cast<size_t>(q)] = 1;
    } catch (const std::exception&) {
    }
  });
  for (size_t q = 0; q < admit.size(); ++q) {
    if (!ok[q]) {
      reject(admit[q]->ticket, admit[q]->grid);
      continue;
    }
    resident_ += static_cast<double>(admit[q]->bytes);
    peakResident_ = std::max(peakResident_, resident_);
    ++admitted_;
    joining_.push_back(std::move(admit[q]));
  }
}

void Server::finish(Slot& s, int status, int iterations, double continuity) {
  const ServeRequest& r = requests_[static_cast<size_t>(s.ticket)];
  const Grid& g = *s.grid;
  const size_t N = static_cast<size_t>(g.cells());
  // the fields come to the host only when they are kept or donated
  const bool donate = status == 0 && opt_.warmStart;
  std::vector<double> u, v, p, mu;
  if (opt_.keepFields || donate) {
    u.resize(N);
    v.resize(N);
    p.resize(N);
    mu.resize(N);
    s.engine->getField(u.data(), v.data(), p.data(), mu.data());
  }
  cfdgpu_serve_result res{};
  res.ticket = s.ticket;
  res.status = status;
  res.iterations = iterations;
  res.retried_cold = s.retried ? 1 : 0;
  res.warm_from = s.warmFrom;
  res.cells = g.cells();
  res.continuity = continuity;
  res.upper_separation = res.lower_separation = -1.0;
  // surface and forces on the device (the host routine's numbers, without the fields)
  cfdgpu_forces f{};
  int count = 0;
  if (s.engine->surface(r.speed, r.alphaDeg, r.density, r.referencePressure, r.chord, nullptr, 0, &count, &f) ==
      CFDGPU_OK) {
    res.cl = f.cl;
    res.cd = f.cd;
    res.cd_pressure = f.cd_pressure;
    res.cd_friction = f.cd_friction;
    res.cm = f.cm;
    res.l_over_d = (std::isfinite(f.cd) && f.cd > 0.0) ? f.cl / f.cd : 0.0;  // Forces.cpp:50-56
    res.upper_separation = f.upper_separation;
    res.lower_separation = f.lower_separation;
  }
  res.arrival_s = r.arrival;
  res.admitted_s = s.admitted;
  res.finished_s = clock();
  res.device_bytes = static_cast<double>(s.bytes);
  done_.push_back(res);
  ++finished_;
  if (status == 0) ++converged_;
  if (opt_.keepFields) {
    std::vector<double> all;
    all.reserve(4 * N);
    for (const auto* a : {&u, &v, &p, &mu}) all.insert(all.end(), a->begin(), a->end());
    kept_[s.ticket] = std::move(all);
  }
  if (donate) {
    const DonorKey key{r.m, r.p, r.t, r.chord, g.nI, g.nJ, r.speed, r.density};
    auto& list = donors_[key];
    list.push_back(Donor{r.alphaDeg, r.reynolds, s.ticket, std::move(u), std::move(v), std::move(p)});
    donorBytes_ += 3 * N * sizeof(double);
    donorOrder_.emplace_back(key, s.ticket);
    if (list.size() > kDonorsPerKey) {
      donorBytes_ -= 3 * list.front().u.size() * sizeof(double);
      list.erase(list.begin());
    }
    while (donorBytes_ > kDonorBytes && !donorOrder_.empty()) {
      auto [k, t] = donorOrder_.front();
      donorOrder_.pop_front();
      auto it = donors_.find(k);
      if (it == donors_.end()) continue;
      auto& l = it->second;
      auto d = std::find_if(l.begin(), l.end(), [&](const Donor& x) { return x.ticket == t; });
      if (d == l.end()) continue;
      donorBytes_ -= 3 * d->u.size() * sizeof(double);
      l.erase(d);
    }
  }
  resident_ -= static_cast<double>(s.bytes);
  s.engine.reset();
}

// Queues `chunk` outer iterations for every resident solve, one launch per
// stage for all of them, then the read-back of each one's stop flags.
void Server::launchChunk() {
  if (active_.empty()) return;
  if (rosterDirty_) {
    // the largest meshes first, as a batch uploads them (largestFirst): the
    // slots themselves are reordered, so host_[q] stays slot q's descriptor
    std::stable_sort(active_.begin(), active_.end(), [](const std::unique_ptr<Slot>& a, const std::unique_ptr<Slot>& b) {
      return a->engine->inst().N > b->engine->inst().N;
    });
    host_.resize(active_.size());
    for (size_t q = 0; q < active_.size(); ++q) host_[q] = active_[q]->engine->refreshInst();
    cfg_ = launchCfgFor(host_.data(), host_.size());
    if (opt_.settings.gs_ordering == CFDGPU_GS_LEXICOGRAPHIC && cfg_.gsLds > 64 * 1024)
      throw std::invalid_argument("resident meshes too wide for the lexicographic sweep's LDS");
    hipCheck(hipMemcpyAsync(dInsts_, host_.data(), sizeof(Inst) * host_.size(), hipMemcpyHostToDevice, stream_),
             "upload roster");
    rosterDirty_ = false;
  } else {
    cfg_ = launchCfgFor(host_.data(), host_.size());  // (the launch tuning may have changed)
  }
  graph_.run(dInsts_, cfg_, host_.front(), stream_, opt_.chunk);
  hipCheck(hipGetLastError(), "launch");
  for (size_t q = 0; q < host_.size(); ++q)
    hipCheck(hipMemcpyAsync(hFlags_ + q * kIFlags, host_[q].flags, sizeof(int) * kIFlags, hipMemcpyDeviceToHost,
                            stream_),
             "read flags");
  launched_ = true;
}

// Waits for the chunk; solves that stopped leave the roster (a first
// divergence is retried from the undisturbed stream in the same slot).
void Server::retire() {
  hipCheck(hipStreamSynchronize(stream_), "chunk");
  launched_ = false;
  ++rounds_;
  std::vector<std::unique_ptr<Slot>> keep;
  keep.reserve(active_.size());
  for (size_t q = 0; q < active_.size(); ++q) {
    Slot& s = *active_[q];
    const int* fl = hFlags_ + q * kIFlags;
    checkSweepStall(fl);
    memberIterations_ += fl[fIteration] - s.lastIteration;
    s.lastIteration = fl[fIteration];
    if (!fl[fStopped]) {
      if (opt_.settleWindow > 0 && settled(s, fl[fIteration])) {
        rosterDirty_ = true;
        cfdgpu_monitor mon{};
        s.engine->readMonitorOn(s.engine->stream(), &mon);
        finish(s, 0, fl[fIteration], mon.continuity);
        continue;
      }
      keep.push_back(std::move(active_[q]));
      continue;
    }
    rosterDirty_ = true;
    const int status = fl[fStatus];
    cfdgpu_monitor mon{};
    s.engine->readMonitorOn(s.engine->stream(), &mon);
    if (status == 2 && !s.retried) {
      // Polar.cpp:112-117: one retry from the undisturbed stream
      s.retried = true;
      const size_t before = s.bytes;
      bool restarted = true;
      try {
        startAttempt(s, true);
      } catch (const std::exception&) {
        restarted = false;  // reported as diverged below
      }
      if (restarted) {
        resident_ += static_cast<double>(s.bytes) - static_cast<double>(before);
        peakResident_ = std::max(peakResident_, resident_);
        keep.push_back(std::move(active_[q]));
        continue;
      }
    }
    finish(s, status, fl[fIteration], mon.continuity);
  }
  active_ = std::move(keep);
}

// Coefficient settling, checked when a solve has run another settle window:
// the residual test of the worker's rule on the last monitor, then Cl and Cd
// of the current field (integrated on the device) against the last samples.
bool Server::settled(Slot& s, int iteration) {
  if (iteration < s.nextSample) return false;
  while (s.nextSample <= iteration) s.nextSample += opt_.settleWindow;
  const ServeRequest& r = requests_[static_cast<size_t>(s.ticket)];
  cfdgpu_forces f{};
  int count = 0;
  if (s.engine->surface(r.speed, r.alphaDeg, r.density, r.referencePressure, r.chord, nullptr, 0, &count, &f) !=
      CFDGPU_OK)
    return false;
  s.samples.emplace_back(f.cl, f.cd);
  if (s.samples.size() > 3) s.samples.pop_front();
  if (s.samples.size() < 3) return false;
  cfdgpu_monitor mon{};
  s.engine->readMonitorOn(s.engine->stream(), &mon);
  const double worst = std::max(std::max(std::fabs(mon.continuity), std::fabs(mon.momentum_x)),
                                std::fabs(mon.momentum_y));
  const double rl = remainingChange(s.samples[0].first, s.samples[1].first, s.samples[2].first);
  const double rd = remainingChange(s.samples[0].second, s.samples[1].second, s.samples[2].second);
  const bool pass = worst < opt_.rule.tolerance && rl <= opt_.settleCl && rd <= opt_.settleCd;
  s.settledChecks = pass ? s.settledChecks + 1 : 0;
  // two consecutive checks, so an early transient that happens to look
  // geometric for three samples does not end the solve
  return s.settledChecks >= 2;
}

void Server::step() {
  launchChunk();
  // admission overlaps the chunk running on the device
  admitReady();
  if (launched_) retire();
  if (!joining_.empty()) {
    for (auto& s : joining_) active_.push_back(std::move(s));
    joining_.clear();
    rosterDirty_ = true;
  }
  peakActive_ = std::max(peakActive_, static_cast<int>(active_.size()));
  if (active_.empty() && !pending_.empty()) {
    // idle until the next arrival
    const double wait = requests_[static_cast<size_t>(pending_.front())].arrival - clock();
    if (wait > 0.0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.02)));
  }
}

void Server::run(double wallLimit) {
  while (!pending_.empty() || !active_.empty() || !joining_.empty()) {
    if (wallLimit > 0.0 && clock() >= wallLimit) break;
    step();
  }
}

int Server::collect(cfdgpu_serve_result* out, int capacity) {
  const int n = std::min(capacity, static_cast<int>(done_.size()));
  for (int q = 0; q < n; ++q) out[q] = done_[static_cast<size_t>(q)];
  done_.erase(done_.begin(), done_.begin() + n);
  return n;
}

bool Server::fields(int ticket, double* u, double* v, double* p, double* mu) const {
  auto it = kept_.find(ticket);
  if (it == kept_.end()) return false;
  const size_t N = it->second.size() / 4;
  double* dst[4] = {u, v, p, mu};
  for (int a = 0; a < 4; ++a)
    if (dst[a]) std::copy_n(it->second.begin() + static_cast<std::ptrdiff_t>(a * N), N, dst[a]);
  return true;
}

int Server::cellsOf(int ticket) const {
  auto it = kept_.find(ticket);
  return it == kept_.end() ? -1 : static_cast<int>(it->second.size() / 4);
}

void Server::stats(cfdgpu_server_stats* out) const {
  out->clock_s = clock();
  out->queued = static_cast<int>(pending_.size());
  out->active = static_cast<int>(active_.size() + joining_.size());
  out->admitted = admitted_;
  out->finished = finished_;
  out->converged = converged_;
  out->peak_active = peakActive_;
  out->rounds = rounds_;
  out->resident_bytes = resident_;
  out->peak_resident_bytes = peakResident_;
  out->member_iterations = static_cast<double>(memberIterations_);
  out->meshes_cached = static_cast<int>(meshes_.size());
}

}  // namespace cfdgpu
