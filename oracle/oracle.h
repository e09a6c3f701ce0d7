/* oracle.h - C entry points of the FP64 CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY. This library is a line-by-line CPU restatement of
 * the reference solver (sidhanth-sridhar/CFD-SIMULATOR) used to check the HIP
 * product path. Only the test suite, the smoke check and the CPU-baseline leg
 * of the benchmark may load it. The product library never links or calls it.
 *
 * Every function mirrors one reference routine; the citation is given at the
 * definition in oracle.cpp. Arrays are plain double/int buffers in the
 * reference's own numbering: cells c = j*(nodesI-1) + i; faces are the j-family
 * [j*(nodesI-1) + i] followed by the i-family [offset + i*(nodesJ-1) + j].
 */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mesh orc_mesh;
typedef struct orc_solver orc_solver;

/* Boundary types of a face (mesh level) - mirrors mesh::BoundaryType. */
enum { ORC_INTERIOR = 0, ORC_WALL = 1, ORC_FARFIELD = 2, ORC_OUTLET = 3, ORC_WAKECUT = 4 };
/* Boundary kinds of a face condition - mirrors flow::BoundaryKind. */
enum { ORC_K_INTERNAL = 0, ORC_K_INLET = 1, ORC_K_OUTLET = 2, ORC_K_FARFIELD = 3, ORC_K_WALL = 4 };

typedef struct {
  double upstream, downstream, vertical;
  int surface_points, wake_points, normal_points;
  double first_layer;
  double normal_blend_length;
  int normal_smoothing_passes;
} orc_cgrid_options;

typedef struct {
  double origin_x, origin_y, length, height;
  int cells_x, cells_y;
  int grading; /* 0 uniform, 1 toward lower, 2 toward both */
  double first_cell_height;
  int left, right, lower, upper; /* ORC_* boundary types */
} orc_box_options;

typedef struct {
  double velocity_relaxation, pressure_relaxation;
  int momentum_sweeps, pressure_iterations;
  double pressure_tolerance;
  int scheme; /* 0 upwind, 1 second-order upwind */
  double turbulence_relaxation;
  double intensity, viscosity_ratio, length_scale;
  int non_orthogonal_correctors;
  double non_orthogonal_blend;
} orc_settings;

typedef struct {
  double sigma_k1, sigma_w1, beta1, sigma_k2, sigma_w2, beta2, beta_star, kappa, a1,
      production_limit;
} orc_sst_constants;

typedef struct {
  double continuity, momentum_x, momentum_y;
  double mass_imbalance, max_divergence;
  int pressure_iterations;
  double turb_first, turb_second;
  double min_k, max_k, min_w, max_w, min_mut, max_mut;
  double max_eddy_ratio;
} orc_monitor;

typedef struct {
  double cl, cd, cm, cd_pressure, cd_friction, lift, drag, pitching_moment;
  double fx_pressure, fy_pressure, fx_friction, fy_friction;
  double upper_separation, lower_separation; /* chord fraction, -1 if none */
  double stagnation_xc;
  int n_upper, n_lower;
  int stations;
} orc_forces;

/* geometry / mesh */
int orc_naca_contour(int m, int p, int t, int points, double chord, int closed, double* xy,
                     int capacity);
double orc_thickness(double x, double t, int closed);
double orc_camber(double x, double m, double p);
double orc_camber_slope(double x, double m, double p);
int orc_geometric_distribution(double first, double total, int intervals, double* out);
int orc_cgrid_defaults(int resolution, orc_cgrid_options* out);
orc_mesh* orc_cgrid(int m, int p, int t, double chord, const orc_cgrid_options* options);
orc_mesh* orc_box(const orc_box_options* options);
orc_mesh* orc_structured(int ni, int nj, const double* nodes_xy, const int* jmin_type,
                         const int* jmin_partner, int jmax_type, int imin_type, int imax_type);
void orc_mesh_free(orc_mesh* mesh);
void orc_mesh_dims(const orc_mesh* mesh, int* nodes_i, int* nodes_j, int* cells, int* faces);
/* quality: min area, max area, max aspect, max non-orth deg, inverted, min wall spacing, total */
void orc_mesh_quality(const orc_mesh* mesh, double* out7);
void orc_mesh_nodes(const orc_mesh* mesh, double* xy);
void orc_mesh_cells(const orc_mesh* mesh, double* centroid_xy, double* area, int* cell_faces4);
void orc_mesh_faces(const orc_mesh* mesh, double* centre_xy, double* normal_xy, double* area,
                    int* owner, int* neighbour, int* boundary, int* partner);

/* boundary conditions: reference defaults wall->NoSlipWall, farfield->FarField, outlet->Outlet */
void orc_face_conditions(const orc_mesh* mesh, double speed, double alpha_deg,
                         double reference_pressure, int* kind, double* velocity_xy,
                         double* pressure);

/* solver */
/* Order of the global sums (L1 norms, CG inner products, means): 0, the
   default, sums sequentially from cell 0 as the reference does; B > 0 sums
   blocks of B cells and then the block totals; -1 sums pairwise (process-wide;
   only for measuring a solve's sensitivity to the rounding of its global
   sums). */
void orc_set_sum_block(int block);
/* SST blending tanh: 0 std::tanh (the default, the reference's); 1 (1 - e^-2x)
   / (1 + e^-2x), a few ulps away (process-wide; sensitivity measurement only). */
void orc_set_tanh_variant(int variant);
void orc_default_settings(orc_settings* out);
void orc_default_sst(orc_sst_constants* out);
orc_solver* orc_solver_create(const orc_mesh* mesh, const int* kind, const double* velocity_xy,
                              const double* pressure, const orc_settings* settings);
void orc_solver_free(orc_solver* solver);
/* model: 0 none, 1 laminar closure, 2 k-omega SST (constants may be NULL for defaults) */
int orc_solver_set_model(orc_solver* solver, int model, const orc_sst_constants* constants);
int orc_solver_initialise(orc_solver* solver, const double* u, const double* v, const double* p,
                          const double* rho, const double* mu);
void orc_solver_set_settings(orc_solver* solver, const orc_settings* settings);
int orc_solver_iterate(orc_solver* solver, int count, orc_monitor* last);
void orc_solver_field(const orc_solver* solver, double* u, double* v, double* p, double* mu);
void orc_solver_set_field(orc_solver* solver, const double* u, const double* v, const double* p);
void orc_solver_flux(const orc_solver* solver, double* flux);
void orc_solver_set_flux(orc_solver* solver, const double* flux);
int orc_solver_turbulence(const orc_solver* solver, double* k, double* omega, double* mut,
                          double* wall_distance, double* f1, double* cross_diffusion);
void orc_solver_set_turbulence(orc_solver* solver, const double* k, const double* omega,
                               const double* mut);
void orc_solver_divergence(const orc_solver* solver, double* out);
double orc_solver_freestream_k(const orc_solver* solver);
double orc_solver_freestream_omega(const orc_solver* solver);
double orc_solver_max_yplus(const orc_solver* solver);
/* Stage-level access for per-stage parity tests. Names: "faceU" "faceV" "faceP" (faces),
 * "gradU" "gradV" "gradP" (cells, x then y interleaved), "momDiag" "momSource" "momUpper"
 * "momLower" (last assembled component), "momentumDiagonal", "pDiag" "pSource" "pUpper"
 * "pLower", "netFlux", "correction", "deadBand", "geomDiffusion" "geomWeight" (faces). */
int orc_solver_array(const orc_solver* solver, const char* name, double* out, int capacity);
/* Individual stages, each run on the current solver state. */
void orc_stage_interpolate(orc_solver* solver);
void orc_stage_gradients(orc_solver* solver);
void orc_stage_momentum(orc_solver* solver, int component); /* assemble + relax */
void orc_stage_rhie_chow(orc_solver* solver);
void orc_stage_pressure(orc_solver* solver);
int orc_stage_cg(orc_solver* solver, double* x_out);
void orc_stage_correct(orc_solver* solver, const double* correction);
void orc_set_momentum_diagonal(orc_solver* solver, const double* diag);

/* LDU primitives on a mesh (reference LinearSystem): system arrays in face numbering */
void orc_ldu_multiply(const orc_mesh* mesh, const double* diag, const double* upper,
                      const double* lower, const double* x, double* y);
void orc_ldu_gauss_seidel(const orc_mesh* mesh, const double* diag, const double* upper,
                          const double* lower, const double* source, double* x, int sweeps);
int orc_ldu_cg(const orc_mesh* mesh, const double* diag, const double* upper,
               const double* lower, const double* source, double* x, double tolerance,
               int max_iterations, double* initial_residual, double* final_residual);
int orc_green_gauss(const orc_mesh* mesh, const double* face_values, double* grad_xy);

/* post-processing */
int orc_surface(const orc_mesh* mesh, const double* u, const double* v, const double* p,
                const double* mu, double speed, double alpha_deg, double density,
                double reference_pressure, double chord, double* stations, int capacity,
                orc_forces* forces);

#ifdef __cplusplus
}
#endif
#endif
