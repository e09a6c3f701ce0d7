/* cfdgpu.h - C ABI of the MI355X (gfx950) SIMPLE solver.
 *
 * Drop-in boundary for the reference's cfd::solver::SimpleSolver class API and
 * its TurbulenceModel plugin (sidhanth-sridhar/CFD-SIMULATOR,
 * include/cfd/solver/SimpleSolver.hpp:159-197, TurbulenceModel.hpp:155-206),
 * plus the inputs it is built from (mesh, face conditions) and the outputs it
 * feeds (surface data, forces, polar). Every entry point returns an int status
 * whose values map 1:1 to the reference's ErrorCode
 * (include/cfd/core/Error.hpp:37-46); 0 is success. No exception crosses the
 * ABI. Arrays are caller-owned host buffers in the reference's numbering:
 *   cell  c = j*(nodesI-1) + i
 *   faces [0, nJfaces)            j-family, index j*(nodesI-1) + i
 *         [nJfaces, nJfaces+nIf)  i-family, index nJfaces + i*(nodesJ-1) + j
 * Device copies are owned by the solver handle. One handle per host thread.
 */
#ifndef CFDGPU_H
#define CFDGPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: include/cfd/core/Error.hpp:37-46 ---- */
enum {
  CFDGPU_OK = 0,
  CFDGPU_INVALID_ARGUMENT = 1,
  CFDGPU_OUT_OF_RANGE = 2,
  CFDGPU_NOT_FOUND = 3,
  CFDGPU_IO_FAILURE = 4,
  CFDGPU_NOT_IMPLEMENTED = 5,
  CFDGPU_INITIALIZATION_FAILURE = 6,
  CFDGPU_NUMERICAL_FAILURE = 7,
  CFDGPU_INTERNAL = 8
};

/* mesh::BoundaryType (include/cfd/mesh/Mesh.hpp:42-48) */
enum { CFDGPU_INTERIOR = 0, CFDGPU_WALL = 1, CFDGPU_FARFIELD = 2, CFDGPU_OUTLET = 3, CFDGPU_WAKECUT = 4 };
/* flow::BoundaryKind (include/cfd/flow/BoundaryConditions.hpp:34-60) */
enum { CFDGPU_K_INTERNAL = 0, CFDGPU_K_INLET = 1, CFDGPU_K_OUTLET = 2, CFDGPU_K_FARFIELD = 3,
       CFDGPU_K_WALL = 4 };
/* turbulence closures */
enum { CFDGPU_MODEL_NONE = 0, CFDGPU_MODEL_LAMINAR = 1, CFDGPU_MODEL_SST = 2 };
/* arithmetic of the hot path */
enum {
  CFDGPU_PRECISION_FP64 = 0, /* plain FP64 path: the drop-in and the judge */
  CFDGPU_PRECISION_MIXED = 1 /* FP32 bulk, FP64 reductions. Pressure solve: the matrix and
                                r, z, p, Ap in FP32 (line CG: FP32 arithmetic too; Jacobi CG:
                                FP64 arithmetic), inner products, norms and the correction x in
                                FP64. With CFDGPU_GS_REDBLACK the momentum and k-omega sweeps
                                run in FP32 on the defect b - A phi formed in FP64 and add
                                their correction to the FP64 field. Assembly, fluxes, pressure,
                                residual norms, the stop test and the force integrals: FP64 */
};
/* Gauss-Seidel ordering of the momentum and turbulence sweeps */
enum {
  CFDGPU_GS_LEXICOGRAPHIC = 0, /* the reference's cell order, exactly (wavefront schedule) */
  CFDGPU_GS_REDBLACK = 1       /* two-colour order: fully parallel, same fixed point */
};
/* pressure-correction solver */
enum {
  CFDGPU_PSOLVE_JACOBI_CG = 0, /* reference: Jacobi-PCG, L1 stop (LinearSystem.cpp:118-210) */
  CFDGPU_PSOLVE_LINE_CG = 1    /* the fast path's CG preconditioned by the exact inverse of each
                                  wall-normal column's tridiagonal part; the reference's L1 stop,
                                  cap and breakdown guards; FP32 storage and arithmetic with FP64
                                  inner products under CFDGPU_PRECISION_MIXED */
};

/* stages of one outer iteration, for cfdgpu_iterate_timed */
enum {
  CFDGPU_STAGE_GRADIENTS = 0,          /* face values + Green-Gauss of u, v, p (SimpleSolver.cpp:268-317) */
  CFDGPU_STAGE_MOMENTUM_ASSEMBLY = 1,  /* assembleMomentum + applyRelaxation + residual (:319-443) */
  CFDGPU_STAGE_MOMENTUM_SWEEPS = 2,    /* 3 Gauss-Seidel sweeps of u and v (LinearSystem.cpp:77-116) */
  CFDGPU_STAGE_RHIE_CHOW = 3,          /* computeRhieChowFluxes (:445-496) */
  CFDGPU_STAGE_PRESSURE_ASSEMBLY = 4,  /* assemblePressureCorrection (:498-580) */
  CFDGPU_STAGE_PRESSURE_SOLVE = 5,     /* conjugateGradient (LinearSystem.cpp:118-210) */
  CFDGPU_STAGE_CORRECTION = 6,         /* correct (:582-660) + divergence monitor (:796-817) */
  CFDGPU_STAGE_SST_BLENDING = 7,       /* strain, gradients of k and omega, F1, F2, CD, production */
  CFDGPU_STAGE_SST_ASSEMBLY = 8,       /* k and omega rows, relaxation, residuals */
  CFDGPU_STAGE_SST_SWEEPS = 9,         /* 3 Gauss-Seidel sweeps of k and of omega */
  CFDGPU_STAGE_SST_WALL_VISCOSITY = 10,/* wall omega + eddy viscosity + effective viscosity */
  CFDGPU_STAGE_MONITOR = 11,           /* reductions, SolverMonitor, stop rule */
  CFDGPU_STAGE_SWEEP_COPY = 12,        /* skewed sweep results copied to the fields (+ k, omega floors) */
  CFDGPU_STAGE_CG_ITERATIONS = 13,     /* the CG iterations alone: part of PRESSURE_SOLVE, not added to it */
  CFDGPU_STAGE_COUNT = 14
};

typedef struct cfdgpu_mesh cfdgpu_mesh;
typedef struct cfdgpu_solver cfdgpu_solver;
typedef struct cfdgpu_batch cfdgpu_batch;

/* mesh::CGridOptions (include/cfd/mesh/CGrid.hpp:53-84) */
typedef struct {
  double upstream_chords, downstream_chords, vertical_chords;
  int surface_points, wake_points, normal_points;
  double first_layer_height;
  double normal_blend_length;
  int normal_smoothing_passes;
} cfdgpu_cgrid_options;

/* mesh::BoxOptions (include/cfd/mesh/BoxGrid.hpp:29-53) */
typedef struct {
  double origin_x, origin_y, length, height;
  int cells_x, cells_y;
  int grading; /* 0 uniform, 1 toward lower, 2 toward both */
  double first_cell_height;
  int left, right, lower, upper; /* CFDGPU_* boundary types */
} cfdgpu_box_options;

/* solver::SimpleSettings (include/cfd/solver/SimpleSolver.hpp:86-132) + TurbulenceInflow
 * (TurbulenceModel.hpp:78-101) + the MI355X path selection. */
typedef struct {
  double velocity_relaxation, pressure_relaxation;
  int momentum_sweeps, pressure_iterations;
  double pressure_tolerance;
  int scheme; /* 0 upwind, 1 second-order upwind */
  double turbulence_relaxation;
  double intensity, viscosity_ratio, length_scale;
  int non_orthogonal_correctors;
  double non_orthogonal_blend;
  /* MI355X path selection (no reference counterpart) */
  int precision;      /* CFDGPU_PRECISION_* */
  int gs_ordering;    /* CFDGPU_GS_* */
  int pressure_solver;/* CFDGPU_PSOLVE_* */
  int refinement_passes; /* 1..8 pressure solves per corrector: pass k > 1 solves for the
                            FP64 residual b - A x of the correction so far and adds its
                            solution in FP64 (iterative refinement); 1 = the reference */
} cfdgpu_settings;

/* solver::SSTConstants (include/cfd/solver/KOmegaSST.hpp:147-171) */
typedef struct {
  double sigma_k1, sigma_w1, beta1, sigma_k2, sigma_w2, beta2, beta_star, kappa, a1,
      production_limit;
} cfdgpu_sst_constants;

/* solver::SolverMonitor (include/cfd/solver/SimpleSolver.hpp:135-154) */
typedef struct {
  double continuity, momentum_x, momentum_y;
  double mass_imbalance, max_divergence;
  int pressure_iterations;
  double turb_first, turb_second;
  double min_k, max_k, min_w, max_w, min_mut, max_mut;
  double max_eddy_ratio;
} cfdgpu_monitor;

/* Stop rule of the caller loop (src/app/SolverWorker.cpp:161-192). */
typedef struct {
  double tolerance;   /* converged when residuals.worst() < tolerance */
  int max_iterations; /* iteration limit */
  int check_every;    /* monitor read-back interval (1 = every iteration, like the reference) */
} cfdgpu_stop_rule;

typedef struct {
  int status;         /* 0 converged, 1 iteration-limit, 2 diverged (post/Polar.cpp:79-84) */
  int iterations;
  cfdgpu_monitor last;
  double seconds;
} cfdgpu_solve_report;

/* post::AerodynamicForces (include/cfd/post/Forces.hpp) + separation summary */
typedef struct {
  double cl, cd, cm, cd_pressure, cd_friction, lift, drag, pitching_moment;
  double fx_pressure, fy_pressure, fx_friction, fy_friction;
  double upper_separation, lower_separation; /* chord fraction, -1 if attached */
  double stagnation_xc;
  int n_upper, n_lower;
  int stations;
  int upper_separation_face, lower_separation_face; /* wall face index (after), -1 if none */
} cfdgpu_forces;

/* post::PolarPoint (include/cfd/post/Polar.hpp:32-60) */
typedef struct {
  double alpha_deg, cl, cd, cd_pressure, cd_friction, cm, l_over_d;
  double upper_separation, lower_separation; /* chord fraction, -1 where attached */
  int status;     /* 0 converged, 1 iteration-limit, 2 diverged (Polar.cpp:79-84) */
  int iterations;
  double continuity; /* last continuity residual */
  int retried_cold;
} cfdgpu_polar_point;

/* The header fields of post::Polar that toCsv prints (Polar.cpp:126-140) */
typedef struct {
  const char* section;
  const char* mesh;
  double reynolds, speed, chord, moment_reference;
  int continued;
} cfdgpu_polar_header;

/* ---- library ---- */
const char* cfdgpu_version(void);
const char* cfdgpu_last_error(void);
int cfdgpu_device_count(int* count);

/* ---- mesh (host side; replaces mesh::generateCGrid CGrid.hpp:92-96, generateBox
 *      BoxGrid.hpp:53, buildStructured Mesh.hpp:209) ---- */
int cfdgpu_cgrid_defaults(int resolution, cfdgpu_cgrid_options* out); /* CGrid.cpp:73-96 */
int cfdgpu_mesh_cgrid(int m, int p, int t, double chord, const cfdgpu_cgrid_options* options,
                      cfdgpu_mesh** out);
int cfdgpu_mesh_box(const cfdgpu_box_options* options, cfdgpu_mesh** out);
int cfdgpu_mesh_structured(int nodes_i, int nodes_j, const double* nodes_xy,
                           const int* jmin_type, const int* jmin_partner, int jmax_type,
                           int imin_type, int imax_type, cfdgpu_mesh** out);
void cfdgpu_mesh_destroy(cfdgpu_mesh* mesh);
int cfdgpu_mesh_dims(const cfdgpu_mesh* mesh, int* nodes_i, int* nodes_j, int* cells, int* faces);
int cfdgpu_mesh_nodes(const cfdgpu_mesh* mesh, double* xy);
int cfdgpu_mesh_cells(const cfdgpu_mesh* mesh, double* centroid_xy, double* area);
int cfdgpu_mesh_faces(const cfdgpu_mesh* mesh, double* centre_xy, double* normal_xy,
                      double* area, int* owner, int* neighbour, int* boundary, int* partner);
/* min area, max area, max aspect, max non-orthogonality (deg), inverted, min wall spacing,
 * total area (Mesh.hpp:80-94) */
int cfdgpu_mesh_quality(const cfdgpu_mesh* mesh, double* out7);

/* ---- flow inputs (flow::buildFaceConditions BoundaryConditions.hpp:111-113 with the default
 *      BoundaryConditions; FlowField::uniform FlowField.hpp:55-57) ---- */
int cfdgpu_face_conditions(const cfdgpu_mesh* mesh, double speed, double alpha_deg,
                           double reference_pressure, int* kind, double* velocity_xy,
                           double* pressure);
int cfdgpu_uniform_field(int cells, double speed, double alpha_deg, double density,
                         double reynolds, double chord, double reference_pressure, double* u,
                         double* v, double* p, double* rho, double* mu);

/* ---- the solver (SimpleSolver.hpp:159-197) ---- */
void cfdgpu_default_settings(cfdgpu_settings* out);
void cfdgpu_default_sst(cfdgpu_sst_constants* out);
/* SimpleSolver::create: the mesh must outlive the solver (SimpleSolver.hpp:156-158) */
int cfdgpu_create(cfdgpu_solver** out, const cfdgpu_mesh* mesh, const int* kind,
                  const double* velocity_xy, const double* pressure,
                  const cfdgpu_settings* settings);
void cfdgpu_destroy(cfdgpu_solver* solver);
/* SimpleSolver::setTurbulenceModel (SimpleSolver.cpp:212-230); constants NULL = defaults */
int cfdgpu_set_turbulence(cfdgpu_solver* solver, int model, const cfdgpu_sst_constants* constants);
/* SimpleSolver::initialise (SimpleSolver.cpp:133-210) */
int cfdgpu_initialise(cfdgpu_solver* solver, const double* u, const double* v, const double* p,
                      const double* rho, const double* mu);
/* SimpleSolver::iterate (SimpleSolver.cpp:662-794), count times; monitor of the last */
int cfdgpu_iterate(cfdgpu_solver* solver, int count, cfdgpu_monitor* out);
/* cfdgpu_iterate with HIP events recorded on the solver's stream around every
   stage: stage_ms[CFDGPU_STAGE_COUNT] receives each stage's device time summed
   over the iterations, stage_calls[CFDGPU_STAGE_COUNT] how often it ran (either
   may be NULL). No reference counterpart: the measurement hook of this path. */
int cfdgpu_iterate_timed(cfdgpu_solver* solver, int count, cfdgpu_monitor* out, double* stage_ms,
                         int* stage_calls);
/* the SolverWorker loop (SolverWorker.cpp:134-277) run device-resident */
int cfdgpu_solve(cfdgpu_solver* solver, const cfdgpu_stop_rule* rule, cfdgpu_solve_report* out);
/* ---- batches: several solves advanced together, one launch per stage for all
 *      of them. No reference counterpart (the reference runs one solve per
 *      thread); it serves polar sweeps and mixed-size batches. The batch refers
 *      to the solvers, which must be initialised, outlive it, and not be used
 *      on their own while a batch call runs. Members must share the closure,
 *      sweep count, corrector passes, ordering, precision, pressure solver,
 *      refinement passes and pressure anchoring; meshes and CG caps may differ
 *      (each member stops its CG at its own cap). ---- */
int cfdgpu_batch_create(cfdgpu_batch** out, cfdgpu_solver* const* solvers, int count);
void cfdgpu_batch_destroy(cfdgpu_batch* batch);
/* count outer iterations of every member; out[i] = member i's last monitor (may be NULL) */
int cfdgpu_batch_iterate(cfdgpu_batch* batch, int count, cfdgpu_monitor* out);
/* cfdgpu_batch_iterate launched directly, with HIP events on the batch's stream around every
   stage: stage_ms / stage_calls as in cfdgpu_iterate_timed, summed over all members (either may
   be NULL). No reference counterpart: the measurement hook of the batched path. */
int cfdgpu_batch_iterate_timed(cfdgpu_batch* batch, int count, cfdgpu_monitor* out, double* stage_ms,
                               int* stage_calls);
/* every member's stop rule on the device until all have stopped; out[i] per member (may be NULL) */
int cfdgpu_batch_solve(cfdgpu_batch* batch, const cfdgpu_stop_rule* rule, cfdgpu_solve_report* out);
int cfdgpu_get_settings(const cfdgpu_solver* solver, cfdgpu_settings* out);
int cfdgpu_set_settings(cfdgpu_solver* solver, const cfdgpu_settings* settings);
/* SimpleSolver::field() */
int cfdgpu_get_field(const cfdgpu_solver* solver, double* u, double* v, double* p, double* mu);
int cfdgpu_set_field(cfdgpu_solver* solver, const double* u, const double* v, const double* p);
/* SimpleSolver::massFlux() */
int cfdgpu_get_massflux(const cfdgpu_solver* solver, double* flux);
int cfdgpu_set_massflux(cfdgpu_solver* solver, const double* flux);
/* TurbulenceModel::turbulentEnergyField/dissipationField/eddyViscosity (+ wall distance) */
int cfdgpu_get_turbulence(const cfdgpu_solver* solver, double* k, double* omega, double* mut,
                          double* wall_distance);
int cfdgpu_set_turbulence_fields(cfdgpu_solver* solver, const double* k, const double* omega,
                                 const double* mut);
int cfdgpu_turbulence_info(const cfdgpu_solver* solver, double* k_inflow, double* omega_inflow,
                           double* max_wall_yplus);
/* SimpleSolver::divergence() (SimpleSolver.cpp:796-817) */
int cfdgpu_divergence(const cfdgpu_solver* solver, double* out);
/* cfdgpu_divergence without allocating device memory: no counterpart in the reference; for callers
   that cannot allocate (code under HIP graph capture, tight serving loops). The field is staged in
   a static device array of 1,048,576 doubles (8 MiB) that each GPU the library runs on holds until
   the process ends; larger meshes are rejected with CFDGPU_INVALID_ARGUMENT. Calls that overlap on
   one GPU share that array and race. Same values as cfdgpu_divergence. */
int cfdgpu_divergence_noalloc(const cfdgpu_solver* solver, double* out);

/* Launch shapes of the line-preconditioned CG (CFDGPU_PSOLVE_LINE_CG), process-wide: cells per
   thread of the cell-parallel launch (1, 2 or 4, or 0 for tiles of 64 columns x 16 rows staged
   in LDS) and rows of a column loaded ahead by the column sweep (16 or 32; batches whose
   meshes have at most 256 rows use the segmented column launch instead). Every shape computes the same iterates; the block partial sums of
   the cell-parallel launch split differently, so results agree to rounding (each shape is
   deterministic). The next iterate, solve, batch call or server chunk uses the new shapes. No
   reference counterpart: the hook of the launch tuner (cfd-simulator_amd/tune.py).
   INVALID_ARGUMENT for other values. Defaults 4 and 32. */
int cfdgpu_set_launch_tuning(int cells_per_thread, int column_rows);
int cfdgpu_get_launch_tuning(int* cells_per_thread, int* column_rows);

/* ---- post-processing, FP64 (SurfaceData.hpp:159, Forces.hpp:127-133) ----
 * rows (12 doubles per station, upper surface then lower): side(1 upper/0 lower), face, x, y,
 * x/c, arc length, p, Cp, tau_w, Cf, nearStagnation, segment length. */
int cfdgpu_surface(const cfdgpu_mesh* mesh, const double* u, const double* v, const double* p,
                   const double* mu, double speed, double alpha_deg, double density,
                   double reference_pressure, double chord, double* rows, int capacity,
                   int* station_count, cfdgpu_forces* forces);
/* cfdgpu_surface of the solver's current field, computed on the device (the wall
 * stations evaluated in parallel, the stagnation point, separation scans and force sums in
 * the reference's order): the same rows and forces as cfdgpu_surface on the downloaded
 * field, without downloading it. NOT_FOUND when the mesh has fewer than four wall faces. */
int cfdgpu_solver_surface(cfdgpu_solver* solver, double speed, double alpha_deg, double density,
                          double reference_pressure, double chord, double* rows, int capacity,
                          int* station_count, cfdgpu_forces* forces);
/* The reference's polar sweep (app/Panels.cpp:1137-1157, 1581-1800 with
 * post::nextSweepAction, Polar.cpp:86-124): for each angle a new solve from the
 * uniform field of that freestream, with velocity and pressure carried over from
 * the previous angle when continue_between_points (the first angle starts from
 * the uniform field); the closure is attached and initialised cold at every
 * point; a diverged point is retried once from cold, and if the retry diverges
 * too the next angle starts cold. Every solve runs on the device under `rule`.
 * model: CFDGPU_MODEL_*; sst NULL = default constants. out[count]. */
int cfdgpu_polar_sweep(const cfdgpu_mesh* mesh, const cfdgpu_settings* settings, int model,
                       const cfdgpu_sst_constants* sst, double speed, double density, double reynolds,
                       double chord, double reference_pressure, const double* alphas, int count,
                       int continue_between_points, const cfdgpu_stop_rule* rule, cfdgpu_polar_point* out);
/* post::toCsv (Polar.cpp:126-163). Returns the text length (without the NUL);
 * the text is written to buf when capacity exceeds it. Negative: -status. */
int cfdgpu_polar_csv(const cfdgpu_polar_point* points, int count, const cfdgpu_polar_header* header, char* buf,
                     int capacity);
/* post::sweepAngles (Polar.cpp:36-77): returns count, or -status on error */
int cfdgpu_sweep_angles(double start, double end, double step, double* out, int capacity);

/* ---- serving: a request loop for polar-point solves with continuous batching.
 *      No reference counterpart as a service; the per-solve semantics are the
 *      reference polar's (app/Panels.cpp:1137-1157, 1581-1800; post/Polar.cpp:79-124).
 *      A request is admitted once its arrival time has passed on the server
 *      clock (seconds since cfdgpu_server_create), a slot is free and its device
 *      state fits the memory budget, in arrival order. Resident solves are
 *      advanced together, one launch per SIMPLE stage for all of them, in chunks
 *      of outer iterations; between chunks stopped solves leave and arrivals
 *      join. Each solve runs the worker's stop rule (SolverWorker.cpp:161-192);
 *      a first divergence is retried once from the undisturbed stream
 *      (Polar.cpp:112-117). With warm_start, a new request starts from the
 *      velocity and pressure of the converged solve of the same section, chord,
 *      mesh size, speed and density nearest in alpha (then in log Re); the
 *      closure starts cold. A result is bit-identical to the same request served
 *      alone when it starts from the same field. One server per host thread. ---- */
typedef struct cfdgpu_server cfdgpu_server;

typedef struct {
  double hbm_budget_bytes; /* device bytes the resident solves may hold */
  int max_slots;           /* most resident solves (1..65535) */
  int chunk_iterations;    /* outer iterations between two scheduling rounds */
  cfdgpu_stop_rule rule;   /* tolerance and iteration limit of every solve (check_every unused) */
  cfdgpu_settings settings;
  int model;               /* CFDGPU_MODEL_*, default constants */
  int warm_start;          /* 1: warm start from converged neighbours in alpha */
  int keep_fields;         /* 1: keep each result's u, v, p, mu for cfdgpu_server_fields */
  /* Coefficient settling (0 = off, the worker's residual rule alone). With settle_window > 0 a
     solve converges when its residuals are below rule.tolerance and its lift and drag have
     settled: every settle_window outer iterations Cl and Cd are integrated on the device, and
     from the last three samples the change still to come is estimated (geometric tail, Aitken:
     |d2| rho / (1 - rho), rho = d2 / d1; an oscillation counts |d2|; a growing change never
     settles). Both estimates must be at most settle_cl and settle_cd, with the residual test
     passed, at two consecutive samples. No reference
     counterpart: the residual rule alone leaves the lift of fine near-wall grids far from its
     converged value. */
  int settle_window;
  double settle_cl, settle_cd;
} cfdgpu_server_options;

typedef struct {
  int m, p, t;                 /* NACA four-digit section (closed trailing edge) */
  double chord;
  cfdgpu_cgrid_options mesh;   /* its C-grid */
  double alpha_deg, reynolds, speed, density, reference_pressure;
  double arrival_s;            /* server-clock arrival time; 0 = at once */
  const cfdgpu_settings* settings; /* NULL: the server's. Relaxation, scheme, tolerances, CG cap and
                                      inflow turbulence may differ; sweeps, correctors, ordering,
                                      precision, pressure solver and refinement passes must be the
                                      server's */
} cfdgpu_serve_request;

typedef struct {
  int ticket;
  int status;       /* 0 converged, 1 iteration-limit, 2 diverged (Polar.cpp:79-84),
                       3 rejected (mesh not buildable or larger than the whole budget) */
  int iterations;   /* of the last attempt */
  int retried_cold;
  int warm_from;    /* ticket whose converged field started the solve, -1 cold */
  int cells;
  double cl, cd, cd_pressure, cd_friction, cm, l_over_d;
  double upper_separation, lower_separation; /* chord fraction, -1 where attached */
  double continuity;
  double arrival_s, admitted_s, finished_s;  /* server clock */
  double device_bytes;
} cfdgpu_serve_result;

typedef struct {
  double clock_s;
  int queued, active, admitted, finished, converged;
  int peak_active, rounds, meshes_cached;
  double resident_bytes, peak_resident_bytes;
  double member_iterations; /* outer iterations advanced, summed over solves */
} cfdgpu_server_stats;

void cfdgpu_server_defaults(cfdgpu_server_options* out);
int cfdgpu_server_create(const cfdgpu_server_options* options, cfdgpu_server** out);
void cfdgpu_server_destroy(cfdgpu_server* server);
int cfdgpu_server_submit(cfdgpu_server* server, const cfdgpu_serve_request* request, int* ticket);
/* one scheduling round: a chunk for the resident solves, admission while it runs, retirement */
int cfdgpu_server_step(cfdgpu_server* server, cfdgpu_server_stats* out);
/* rounds until nothing is queued or resident, or the server clock reaches wall_limit_s (<= 0: none) */
int cfdgpu_server_run(cfdgpu_server* server, double wall_limit_s, cfdgpu_server_stats* out);
/* moves up to capacity finished results into out, oldest first; *count = how many */
int cfdgpu_server_collect(cfdgpu_server* server, cfdgpu_serve_result* out, int capacity, int* count);
/* the final fields of a finished ticket (keep_fields); cells from cfdgpu_serve_result */
int cfdgpu_server_fields(const cfdgpu_server* server, int ticket, double* u, double* v, double* p,
                         double* mu);

/* ---- staged access for parity tests (one SIMPLE stage on the current device state) ----
 * names: "gradU" "gradV" "gradP" (cells, xy interleaved), "momDiag" "momS" "momN" "momW"
 * "momE" "momBx" "momBy", "pDiag" "pN" "pE" "pWake" "pSource", "correction", "strain",
 * "f1" "f2" "crossDiffusion" */
int cfdgpu_debug_array(const cfdgpu_solver* solver, const char* name, double* out, int capacity);

#ifdef __cplusplus
}
#endif
#endif
