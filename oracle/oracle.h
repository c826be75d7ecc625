/*
 * CPU oracle for the Mallard explicit-residual hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a clean-room FP64 restatement of the
 * reference algorithm (reference numbering, AoS state, atomics in the
 * OpenMP mode) that the tests compare the HIP path with, and the timed CPU
 * baseline. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it. The product library never links or calls it.
 *
 * Every routine cites the reference file:line it restates. Parity status per
 * routine is listed in oracle.cpp's header comment.
 */
#ifndef MALLARD_ORACLE_H
#define MALLARD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mesh orc_mesh;
typedef struct orc_solver orc_solver;

/* mesh type codes follow mesh.h:23-28 (file=0 is rejected) */
enum { ORC_MESH_CARTESIAN = 1, ORC_MESH_CARTESIAN_TRI = 2, ORC_MESH_WEDGE = 3 };
enum { ORC_RECON_FO = 0, ORC_RECON_TENO = 1 };
enum { ORC_BASIS_MONOMIAL = 0, ORC_BASIS_LEGENDRE = 1 };
enum { ORC_RS_RUSANOV = 0, ORC_RS_HLL = 1, ORC_RS_HLLC = 2, ORC_RS_ROE = 3 };
enum { ORC_TI_FE = 0, ORC_TI_RK4 = 1, ORC_TI_SSPRK3 = 2 };
/* TENO variants. REFERENCE restates face_reconstruction.cpp:866-1039 as
   written: in the ENO branch the central stencil keeps its raw weight and
   CT2-failing directional stencils keep raw weights (:964-981), and face
   values use psi_k + ipt_k / ta_0 (:1025-1030). CORRECTED is the documented
   deviation used where the REFERENCE form diverges: ENO selection among the
   CT2-passing directional stencils only, zero-mean face evaluation
   psi_k - ipt_k, Zhang-Shu positivity scaling of the face states, and a
   vacuum guard in the two-rarefaction pressure estimate. */
enum { ORC_TENO_REFERENCE = 0, ORC_TENO_CORRECTED = 1 };
enum { ORC_BC_SYMMETRY = 0, ORC_BC_EXTRAPOLATION = 1, ORC_BC_WALL_ADIABATIC = 2,
       ORC_BC_UPT = 3, ORC_BC_P_OUT = 4 };

typedef struct {
    int32_t zone;      /* index into the mesh's zone list (1..4; 0 is interior) */
    int32_t type;      /* ORC_BC_* */
    double u[2], p, T; /* UPT: u, p, T; P_OUT: p */
} orc_bc;

typedef struct {
    int32_t recon, basis, order;
    double stencil_factor;
    int32_t quad_order_cell, quad_order_face; /* <=0 -> reference defaults */
    int32_t riemann, integrator;
    double gamma, p_ref, T_ref, rho_ref, p_min, p_max;
    int32_t n_bc;
    orc_bc bc[8];
    int32_t teno_variant; /* ORC_TENO_* */
    int32_t positivity;   /* 1: positivity-preserving flux limiter (FE, SSP-RK3) */
} orc_config;

/* ---- mesh ---- */
orc_mesh* orc_mesh_create(int type, uint32_t nx, uint32_t ny, double Lx, double Ly,
                          double jitter, uint64_t seed);
/* A caller's mesh in the reference's array layout: 2 nodes per face,
   faces_of_cell sharing offsets_nodes_of_cell, zone 0 named "interior". */
orc_mesh* orc_mesh_create_arrays(uint32_t n_nodes, const double* node_coords, uint32_t n_cells,
                                 const uint32_t* offsets_nodes_of_cell, const uint32_t* nodes_of_cell,
                                 const uint32_t* faces_of_cell, uint32_t n_faces, const uint32_t* nodes_of_face,
                                 const int32_t* cells_of_face, uint32_t n_zones, const char* const* zone_names,
                                 const uint32_t* zone_offsets, const uint32_t* zone_faces);
void orc_mesh_destroy(orc_mesh* m);
/* counts[0..5] = n_cells, n_nodes, n_faces (allocated), n_faces_real,
   nodes_of_cell size, n_zones */
void orc_mesh_counts(const orc_mesh* m, int64_t* counts);
/* Copy-out of the reference arrays. Any pointer may be NULL. */
void orc_mesh_arrays(const orc_mesh* m, double* node_coords, double* cell_coords,
                     double* cell_volume, double* face_area, double* face_normals,
                     uint32_t* nodes_of_cell, uint32_t* offsets_nodes_of_cell,
                     uint32_t* faces_of_cell, uint32_t* offsets_faces_of_cell,
                     uint32_t* nodes_of_face, uint32_t* offsets_nodes_of_face,
                     int32_t* cells_of_face);
int64_t orc_mesh_zone(const orc_mesh* m, int zone, uint32_t* faces_out, char* name_out);

/* ---- solver ---- */
orc_solver* orc_solver_create(orc_mesh* m, const orc_config* cfg, int n_threads);
/* faces whose flux the positivity limiter changed, summed over the steps so far */
int64_t orc_pp_limited(const orc_solver* s);
/* the limiter's pieces, for its tests: the largest theta in [0, theta] that
 * keeps U + sc (Flo + theta (F - Flo)) admissible, and the low-order flux of
 * two means across an interior face (b < 0) or of one mean at boundary b */
double orc_pp_theta(const orc_solver* s, const double* U, double sc, const double* F, const double* Flo,
                    double theta);
int orc_pp_low_flux(const orc_solver* s, int b, const double* nu, const double* U0, const double* U1, double* F);
void orc_solver_destroy(orc_solver* s);
/* on != 0: the residual's face contributions are gathered per cell in the
   one-thread order instead of added with atomics, so every thread count gives
   the one-thread result bit for bit (off by default). */
void orc_set_canonical(orc_solver* s, int on);
const char* orc_last_error(void);
/* TENO sizes: out[0..4] = n_dof, max_cells_per_stencil, n stencils entries,
   n offsets_stencils, n reconstruction-matrix entries */
void orc_teno_sizes(const orc_solver* s, int64_t* out);
void orc_teno_arrays(const orc_solver* s, uint32_t* offsets_stencil_groups,
                     uint32_t* offsets_stencils, uint32_t* stencils,
                     uint64_t* offsets_recon, double* recon, double* transformed_areas,
                     double* integral_psi_target, double* oscillation_indicator,
                     uint8_t* poly_indices);
/* quadrant initial state, solver_initialize.cpp:101-186 semantics with the
   state chosen by the centroid's quadrant. states[4][4] = (rho,u,v,p) for
   UR, UL, LL, LR. */
void orc_init_quadrant(orc_solver* s, double xs, double ys, const double* states);
void orc_init_constant(orc_solver* s, const double* u, double p, double T);
void orc_set_state(orc_solver* s, const double* U, const double* prim);
void orc_get_state(const orc_solver* s, double* U, double* prim);
/* One residual evaluation R(U) into rhs (N_c x 4). */
int orc_calc_rhs(orc_solver* s, const double* U, double* rhs);
/* The residual over the interior faces (parts 1), the boundary zones (2) or
   both (3), divided by the volume (solver_rhs.cpp:44-115). */
int orc_calc_rhs_parts(orc_solver* s, const double* U, int parts, double* rhs);
/* Face states of the last residual evaluation, [n_faces alloc][Q][2][4]. */
void orc_get_face_values(const orc_solver* s, double* out);
int orc_calc_dt(orc_solver* s, double cfl, double* dt_out, double* cfl_local);
int orc_take_step(orc_solver* s, double dt);
/* n steps with dt from cfl each step (the Solver::run loop body without
   checks/output). dts_out may be NULL. */
int orc_run(orc_solver* s, int n_steps, double cfl, double* dts_out);
double orc_time(const orc_solver* s);
/* cut-off decision of the TENO kernel for every cell and variable in the last
   residual evaluation: 1 if the central stencil was selected. */
void orc_teno_decisions(const orc_solver* s, uint8_t* out);

/* ---- pointwise routines pinned by the reference's unit tests ---- */
/* RiemannSolver::calc_flux as the reference writes it (no vacuum guard) */
void orc_riemann_flux(int type, double* flux, const double* n_unit, double rho_l,
                      const double* u_l, double p_l, double gamma_l, double h_l,
                      double rho_r, const double* u_r, double p_r, double gamma_r,
                      double h_r);
/* the same with the corrected variant's vacuum guard on (guard != 0) or off */
void orc_riemann_flux_guarded(int type, int guard, double* flux, const double* n_unit, double rho_l,
                              const double* u_l, double p_l, double gamma_l, double h_l,
                              double rho_r, const double* u_r, double p_r, double gamma_r,
                              double h_r);
void orc_physics_constants(double gamma, double p_ref, double T_ref, double rho_ref,
                           double* R_cp_cv);
void orc_primitives(const double* consts /* gamma,R,cp,cv,p_min,p_max */,
                    const double* cons, double* prim);
void orc_qr_noq(const double* A, double* R, int m, int n);
void orc_forward_substitution(const double* L, const double* B, double* X, int m, int n,
                              int p, int tL, int tB);
void orc_back_substitution(const double* U, const double* B, double* X, int m, int n,
                           int p, int tU, int tB);
void orc_gemm(const double* A, const double* B, double* C, int m, int n, int p, int q,
              int tA, int tB);
void orc_invert2(const double* A, double* Ainv);
void orc_invert3(const double* A, double* Ainv);
double orc_triangle_area(const double* v0, const double* v1, const double* v2);
void orc_triangle_J(const double* v0, const double* v1, const double* v2, double* J);
double orc_basis_1d(int basis, int p, double x);
double orc_basis_deriv_1d(int basis, int n, int p, double x);
/* in-place integrator step with a constant RHS callback value (time_integrator_test) */
void orc_integrator_constant_rhs(int type, double dt, double* U, int n, const double* rhs);
/* the solver's integrator sequence on dU/dt = lam U (closed-form amplification test) */
int orc_integrator_linear_rhs(int type, double dt, double lam, double* U, int n);

#ifdef __cplusplus
}
#endif
#endif
