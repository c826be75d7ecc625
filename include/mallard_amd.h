/*
 * mallard_amd: C ABI of the MI355X-native explicit residual engine for
 * Mallard's compressible Euler path on unstructured 2D meshes.
 *
 * Plain C types only. Every call returns 0 on success and a nonzero status
 * otherwise; exceptions and HIP errors never cross this boundary, the text of
 * the last failure is returned by mlr_last_error(ctx) (or mlr_last_error(NULL)
 * for failures that have no context). A context owns all device memory and one
 * HIP stream; calls on one context must come from one thread at a time.
 *
 * Arrays passed in or out are host arrays in the reference numbering of
 * cells and faces (mesh.cpp generators); the library renumbers internally.
 * Conserved state is AoS double[N_cells][4] = (rho, rho u, rho v, rho E);
 * primitives are double[N_cells][5] = (u, v, p, T, h), as Kokkos::View<rtype
 * *[N]> LayoutRight in the reference.
 */
#ifndef MALLARD_AMD_H
#define MALLARD_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlr_ctx mlr_ctx;

/* Mesh type codes (mesh.h:23-35). */
enum { MLR_MESH_FILE = 0, MLR_MESH_CARTESIAN = 1, MLR_MESH_CARTESIAN_TRI = 2, MLR_MESH_WEDGE = 3 };
/* Face reconstruction (face_reconstruction.h:30-33). */
enum { MLR_RECON_FO = 0, MLR_RECON_TENO = 1 };
/* Basis (basis.h:24-32). */
enum { MLR_BASIS_MONOMIAL = 0, MLR_BASIS_LEGENDRE = 1 };
/* TENO weighting: 0 = as the reference computes it
 * (face_reconstruction.cpp:948-1031); 1 = corrected variant (ENO selection
 * among CT2-passing directional stencils, zero-mean face evaluation,
 * positivity scaling, vacuum-guarded two-rarefaction estimate). */
enum { MLR_TENO_REFERENCE = 0, MLR_TENO_CORRECTED = 1 };
/* Riemann solvers (riemann_solver.h:21-31; Roe has no reference counterpart). */
enum { MLR_RS_RUSANOV = 0, MLR_RS_HLL = 1, MLR_RS_HLLC = 2, MLR_RS_ROE = 3 };
/* Time integrators (time_integrator.h:29-33). */
enum { MLR_TI_FE = 0, MLR_TI_RK4 = 1, MLR_TI_SSPRK3 = 2 };
/* Boundary conditions (boundary.h:22-37). */
enum { MLR_BC_SYMMETRY = 0, MLR_BC_EXTRAPOLATION = 1, MLR_BC_WALL_ADIABATIC = 2, MLR_BC_UPT = 3,
       MLR_BC_P_OUT = 4, MLR_BC_WALL_NOSLIP = 5 /* adiabatic no-slip wall; no reference counterpart */ };

/* One [[boundaries]] entry (solver.cpp:188-237): zone name, type and the
 * parameters UPT (u, p, T) and P_OUT (p) read. */
typedef struct {
    char zone[32];
    int32_t type;
    double u[2];
    double p, T;
} mlr_bc_desc;

/* [numerics], [physics] and [run] settings (solver.cpp:110-273,
 * physics.cpp:27-53, face_reconstruction.cpp:109-152). quad orders <= 0 take
 * the reference defaults (Dunavant p+1, Gauss-Legendre (p+1)/2). cfl != 0
 * selects CFL time steps (a negative CFL number is refused with "dt
 * negative", as solver.cpp:580-590 refuses the negative dt it gives),
 * cfl = 0 the fixed dt, which must be positive. */
typedef struct {
    int32_t recon, basis, order;
    double stencil_factor;
    int32_t quad_order_cell, quad_order_face, teno_variant;
    int32_t riemann, integrator;
    double gamma, p_ref, T_ref, rho_ref, p_min, p_max;
    double cfl, dt;
    int32_t n_bc;
    mlr_bc_desc bc[14];
    int32_t n_threads; /* host precompute threads, 0 = default */
    /* Navier-Stokes (no reference counterpart; the reference solves Euler):
     * viscous != 0 adds the viscous fluxes with Sutherland viscosity
     * mu = mu_ref (T/T_mu_ref)^1.5 (T_mu_ref + S)/(T + S) (constant mu_ref when
     * sutherland_S < 0) and heat conductivity mu cp / prandtl. */
    int32_t viscous;
    double mu_ref, T_mu_ref, sutherland_S, prandtl;
    /* positivity != 0: the positivity-preserving flux limiter of Hu, Adams &
     * Shu (J. Comput. Phys. 242 (2013) 169) on every face of every stage
     * update (no reference counterpart). Each face's flux is blended toward
     * the Rusanov flux of the two cell means just enough that both cells'
     * one-face forward-Euler brackets keep rho, p > 0; faces it leaves alone
     * give bit for bit the unlimited result. FE and SSP-RK3, Euler only. */
    int32_t positivity;
} mlr_numerics_desc;

/* Structured generators of the reference (mesh.cpp:305-848). jitter moves
 * interior nodes by jitter*(dx,dy)*(2u-1), u from a splitmix64 stream keyed by
 * (seed, node id); jitter = 0 is the reference generator. */
typedef struct {
    int32_t type;
    uint32_t nx, ny;
    double Lx, Ly, jitter;
    uint64_t seed;
} mlr_mesh_gen;

/* Unstructured mesh in reference numbering (the arrays of mesh.h:228-253):
 * 2 nodes per face, faces_of_cell CSR sharing offsets_nodes_of_cell, zone 0
 * named "interior". */
typedef struct {
    uint32_t n_nodes;
    const double* node_coords;
    uint32_t n_cells;
    const uint32_t* offsets_nodes_of_cell;
    const uint32_t* nodes_of_cell;
    const uint32_t* faces_of_cell;
    uint32_t n_faces;
    const uint32_t* nodes_of_face;
    const int32_t* cells_of_face;
    uint32_t n_zones;
    const char* const* zone_names;
    const uint32_t* zone_offsets;
    const uint32_t* zone_faces;
} mlr_mesh_desc;

/* Solver::init(input_file) (solver.cpp:39-80): reads a Mallard TOML input
 * (mesh, physics, numerics, boundaries, run, initialize) and builds the
 * device state. */
int mlr_ctx_create_toml(const char* input_toml, mlr_ctx** out);
/* Same with a generated mesh and explicit settings (Mesh::init, mesh.cpp:32-66). */
int mlr_ctx_create_generated(const mlr_mesh_gen* gen, const mlr_numerics_desc* num, mlr_ctx** out);
/* Same with a caller mesh (the host arrays copy_host_to_device uploads, mesh.cpp:263-282). */
int mlr_ctx_create(const mlr_mesh_desc* mesh, const mlr_numerics_desc* num, mlr_ctx** out);
int mlr_ctx_destroy(mlr_ctx* ctx);
const char* mlr_last_error(const mlr_ctx* ctx);

/* sizes[0..14] = n_cells, n_faces (allocated), n_faces_real, face slots per
 * cell, face quadrature points, n_dof, cells per stencil, stencils per cell,
 * device bytes allocated by the context, step, recon, teno_variant, non-empty
 * TENO stencils over all cells, TENO tile-list entries (all 64-cell blocks),
 * bytes per in-tile member position */
int mlr_get_sizes(const mlr_ctx* ctx, int64_t* sizes);

/* Initial states (solver_initialize.cpp:35-186). states[16] = (rho,u,v,p)
 * for the upper-right, upper-left, lower-left and lower-right quadrants of
 * the split point (xs, ys), chosen by the cell centroid. */
int mlr_init_quadrant(mlr_ctx* ctx, double xs, double ys, const double* states);
int mlr_init_constant(mlr_ctx* ctx, const double* u, double p, double T);
/* Host <-> device state copies (Solver::copy_host_to_device / _to_host,
 * solver.cpp:322-334). prim may be NULL on input (computed from U). */
int mlr_set_state(mlr_ctx* ctx, const double* U, const double* prim);
int mlr_get_state(mlr_ctx* ctx, double* U, double* prim);

/* Solver::calc_rhs (solver_rhs.cpp:44-55): residual R(U) of the current
 * state, double[N][4]. */
int mlr_calc_rhs(mlr_ctx* ctx, double* rhs);
/* Solver::calc_dt (solver.cpp:580-590) with the configured cfl or dt. */
int mlr_compute_dt(mlr_ctx* ctx, double* dt_out);
/* Solver::take_step (solver.cpp:521-531) with the dt of the last
 * mlr_compute_dt. */
int mlr_step(mlr_ctx* ctx);
/* n iterations of the run loop body (calc_dt + take_step, solver.cpp:358-364),
 * captured once in a hipGraph and replayed; dt values written to dts if not
 * NULL. */
int mlr_run_steps(mlr_ctx* ctx, int64_t n, double* dts);
/* The same n steps launched kernel by kernel (no graph), with a HIP event
 * recorded on the context's stream after every launch. No reference
 * counterpart: it is the measurement hook behind the throughput benchmark.
 * stats[9] = { total event ms,
 *              TENO face-state ms, launches,  flux+stage-update ms, launches,
 *              spectral-radius+dt ms, launches, copy ms, launches }. */
int mlr_run_steps_profiled(mlr_ctx* ctx, int64_t n, double* stats);
/* Compact TENO tables (corrected weighting, p = 2 with the default stencil
 * size; no reference counterpart). The member areas are folded into the
 * pseudo-inverses, a = (A+ diag(ta)) (u_i - u_c), where the reference forms
 * b_i = ta_i (u_i - u_c) first (face_reconstruction.cpp:900-925), and the
 * face-point basis values are those of the reference triangle's edges
 * instead of each cell's face points mapped through its inverse Jacobian
 * (:985-1031). Both are the same quantities in exact arithmetic; the TENO
 * pass reads 472 fewer bytes per cell. Results differ from the exact tables'
 * by rounding in one residual, which the TENO cut-off can amplify over many
 * steps (4.4e-8 relative after 100 steps of the 8.4M-cell four-quadrant
 * problem), so this is an optional fast path, not the parity path. One way:
 * the exact tables are released. */
int mlr_compact_tables(mlr_ctx* ctx);
/* Launch shapes; no reference counterpart (the reference leaves them to
 * Kokkos' defaults). tune[5] = { TENO occupancy target in waves per SIMD
 * (0: the compiler's choice; 5 or 6 at p = 2 and 4 or 5 for plain TENO at
 * p = 3, each with the default stencil size), cells per flux/stage-update block (64, 128, 256), cells per
 * spectral-radius block (64, 128, 256), the TENO occupancy target of gated
 * passes that reuse recorded decisions, the plain TENO pass's face-basis
 * source: 0 reads the stored face-point basis values, 1 derives them in
 * the kernel from the cell's node coordinates with the precompute's
 * operations (p = 2, Legendre basis, default stencil size; occupancy target
 * 0, 4 or 5): 72 fewer bytes per cell to read }. Every choice gives
 * bit-identical results. mlr_autotune times the candidates on the current state and
 * returns the fastest (also run by itself before the first steps, cached per
 * mesh size and numerics); mlr_set_tune pins a choice. */
int mlr_autotune(mlr_ctx* ctx, int32_t* tune);
int mlr_set_tune(mlr_ctx* ctx, const int32_t* tune);
int mlr_get_tune(mlr_ctx* ctx, int32_t* tune);
/* Solver::run (solver.cpp:352-373) with the configured stop conditions. */
int mlr_run(mlr_ctx* ctx);
int mlr_set_cfl(mlr_ctx* ctx, double cfl);
/* Length of the device ring that holds each step's dt until mlr_run_steps /
 * mlr_run read it back (1 .. 2^20 steps, default 2^20; no reference
 * counterpart: the reference reads dt on the host every step,
 * solver.cpp:358-364). Longer runs are issued in chunks of at most this many
 * steps. */
int mlr_set_dt_history(mlr_ctx* ctx, int64_t steps);
int mlr_get_time(const mlr_ctx* ctx, double* t, int64_t* step);
/* Face states of the last residual, face_solution[n_faces][Q][2][4] as in
 * FaceReconstruction::calc_face_values (face_reconstruction.cpp:95-99,
 * :1081-1106); entries no cell writes are NaN. */
int mlr_get_face_values(mlr_ctx* ctx, double* out);
/* TENO cut-off outcome of the last residual per cell: bit v set when the
 * central stencil was selected for variable v (face_reconstruction.cpp:958). */
int mlr_get_decisions(mlr_ctx* ctx, uint8_t* out);
/* cfl_local after the last dt (solver.cpp:584, :705). */
int mlr_get_cfl_local(mlr_ctx* ctx, double* out);

/* ---- The strategy interfaces on their own, for callers that keep their own
 * Solver loop and swap in one GPU piece at a time. */

/* FaceReconstruction::calc_face_values(solution, face_solution)
 * (face_reconstruction.h:43-90; TENO face_reconstruction.cpp:1081-1106, FO
 * :59-99) on a caller's state: U[N_cells][4] in reference numbering,
 * face_solution[n_faces][Q][2][4] as mlr_get_face_values. The context's own
 * state, time and dt are left as they were. */
int mlr_calc_face_values(mlr_ctx* ctx, const double* U, double* face_solution);

/* Face sums of the residual (solver_rhs.cpp:44-115): MLR_RHS_INTERIOR is
 * calc_rhs_interior (the interior flux functor, flux_functor.h:106-178),
 * MLR_RHS_BOUNDARY is calc_rhs_boundaries, i.e. Boundary::apply of every zone
 * (boundary.h:47-126, boundary_*.cpp), MLR_RHS_ALL both. Each result is divided
 * by the cell volume (DivideVolumeFunctor, solver_rhs.cpp:18-42), so the
 * interior and boundary parts add up to the residual (to rounding: the full
 * residual sums each cell's faces in one pass). */
enum { MLR_RHS_INTERIOR = 1, MLR_RHS_BOUNDARY = 2, MLR_RHS_ALL = 3 };
/* Solver::calc_rhs(solution, face_solution, rhs) on a caller's state:
 * U[N][4] and rhs[N][4] host arrays in reference numbering. */
int mlr_calc_rhs_of(mlr_ctx* ctx, const double* U, int parts, double* rhs);
/* The same on DEVICE arrays U_dev[N][4], rhs_dev[N][4] in the context's
 * numbering (cell j of these arrays is reference cell perm[j],
 * mlr_get_permutation), queued on the context's stream without a host copy or
 * a synchronisation: the form a device rhs function for
 * mlr_integrator_take_step takes. */
int mlr_calc_rhs_device(mlr_ctx* ctx, const double* U_dev, int parts, double* rhs_dev);

/* RiemannSolver::calc_flux(flux, n_unit, rho_l, u_l, p_l, gamma_l, h_l, rho_r,
 * u_r, p_r, gamma_r, h_r) (riemann_solver.h:42-94; HLLC :441-519, HLL
 * :377-439, Rusanov :331-375; Roe has no reference counterpart) for n face
 * points at once, host arrays: n_unit[n][2], left[n][6] and right[n][6] =
 * (rho, u_x, u_y, p, gamma, h), flux[n][4]. type = MLR_RS_*. guard != 0 floors
 * the two-rarefaction pressure terms at zero (the corrected TENO variant's
 * vacuum guard); guard = 0 is the reference's formula. The points are staged
 * through device memory allocated for the call. */
int mlr_riemann_flux(int type, int guard, int64_t n, const double* n_unit, const double* left, const double* right,
                     double* flux);

/* mlr_riemann_flux without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The points are staged through one fixed device array of
 * mlr_riemann_flux_noalloc_capacity() points (65,536 points, 9.4 MB), which
 * every GPU the library runs on holds until the process ends; larger batches
 * are refused with an error. Calls on one GPU share that array, so calls that
 * overlap on one GPU (from several host threads) race. */
int mlr_riemann_flux_noalloc(int type, int guard, int64_t n, const double* n_unit, const double* left,
                             const double* right, double* flux);
int64_t mlr_riemann_flux_noalloc_capacity(void);

/* TimeIntegrator::take_step(dt, solution_vec, face_solution, rhs_vec, calc_rhs)
 * (time_integrator.h:41-99, time_integrator.cpp:57-163) on DEVICE vectors of n
 * doubles: solution_vec[0] is the state, advanced in place; solution_vec[1]
 * the stage state (RK4, SSPRK3); rhs_vec[0..] the stage residuals (FE 1, RK4 4,
 * SSPRK3 3: get_n_solution_vectors / get_n_rhs_vectors). rhs(U, rhs, user) is
 * called once per stage with device pointers and returns 0 on success. The
 * copies and KokkosBlas axpy / axpby updates run as kernels in the reference's
 * order on the context's stream (ctx may be NULL: the null stream), so an rhs
 * function that queues its work on that stream, such as mlr_calc_rhs_device
 * on the same context, needs no synchronisation; any other must finish its
 * work before returning. The call returns after the step has finished. */
typedef int (*mlr_rhs_fn)(const double* U, double* rhs, void* user);
int mlr_integrator_take_step(mlr_ctx* ctx, int type, double dt, int64_t n, double* const* solution_vec,
                             double* const* rhs_vec, mlr_rhs_fn rhs, void* user);

/* Index parity hooks: new -> reference cell permutation, and one cell's
 * stencils / pseudo-inverses recomputed in reference numbering
 * (face_reconstruction.cpp:406-433, :477-741). ids[G*M], sizes[G],
 * aplus[G*K*M] (row-major K x M), ta[G*M], ipt[K], oi[K*K]. */
int mlr_get_permutation(const mlr_ctx* ctx, uint32_t* perm);
int mlr_get_stencils(const mlr_ctx* ctx, uint32_t cell, uint32_t* ids, uint32_t* sizes);
int mlr_get_teno_matrices(const mlr_ctx* ctx, uint32_t cell, double* aplus, double* ta, double* ipt,
                          double* oi);
/* Device-side stencil ids of one cell mapped back to reference numbering
 * (the uploaded layout, target cell first), ids[G*M]. */
int mlr_get_device_stencils(const mlr_ctx* ctx, uint32_t cell, uint32_t* ids);
/* The uploaded pseudo-inverses of one cell (aplus[G*K*M], K x M per stencil
 * with the zero first row and column) and member areas (ta[G*M], entry 0 of
 * each stencil left 0): for checking the GPU precompute against the host's. */
int mlr_get_device_matrices(const mlr_ctx* ctx, uint32_t cell, double* aplus, double* ta);

/* Host-only precompute (no GPU needed): the mesh generator, the SFC
 * renumbering and the packed TENO layout exactly as a context would upload
 * them, for the index and precompute parity tests. The query functions map
 * the packed arrays back to reference numbering:
 *   mesh_arrays: node_coords[2Nn], cell_coords[2Nc], cell_volume[Nc],
 *     face_area[Nf], face_normals[2Nf], nodes_of_cell[], faces_of_cell[],
 *     nodes_of_face[2Nf], cells_of_face[2Nf] (any pointer may be NULL)
 *   cell_stencils: ids[G*M] (target first, 0xFFFFFFFF for empty stencils)
 *   cell_matrices: aplus[G*K*M] row-major K x M with the zero row/column,
 *     ta[G*M], fbasis[S*Q*(K-1)]
 *   teno_tables: ipt[K], oi[K*K]
 *   slots: per face slot of the cell (S entries each): neighbour (reference
 *     id, or the cell itself on a boundary), info word, unit normal x/y,
 *     area, spectral-radius length. */
typedef struct mlr_layout mlr_layout;
int mlr_layout_create(const mlr_mesh_gen* gen, const mlr_numerics_desc* num, mlr_layout** out);
int mlr_layout_destroy(mlr_layout* lay);
/* sizes[0..13] = n_cells, n_faces (allocated), n_faces_real, face slots per
 * cell, face quadrature points, n_dof, cells per stencil, stencils per cell,
 * device bytes of the packed arrays, n_nodes, nodes_of_cell entries, zones,
 * longest 64-cell tile list, bytes per in-tile position (1 or 2) */
int mlr_layout_sizes(const mlr_layout* lay, int64_t* sizes);
int mlr_layout_permutation(const mlr_layout* lay, uint32_t* perm);
int mlr_layout_mesh_arrays(const mlr_layout* lay, double* node_coords, double* cell_coords,
                           double* cell_volume, double* face_area, double* face_normals,
                           uint32_t* nodes_of_cell, uint32_t* faces_of_cell, uint32_t* nodes_of_face,
                           int32_t* cells_of_face);
int mlr_layout_cell_stencils(const mlr_layout* lay, uint32_t cell, uint32_t* ids);
int mlr_layout_cell_matrices(const mlr_layout* lay, uint32_t cell, double* aplus, double* ta,
                             double* fbasis);
int mlr_layout_teno_tables(const mlr_layout* lay, double* ipt, double* oi);
int mlr_layout_slots(const mlr_layout* lay, uint32_t cell, uint32_t* nb, uint32_t* info, double* nx,
                     double* ny, double* area, double* dxn);

/* ---- batched ensembles (no reference counterpart): n independent cases with
 * their own meshes, states and CFL time steps, every kernel phase of a step
 * one launch over all cases driven by a device table of per-case
 * descriptors and block offsets, the step captured in a hipGraph. Each case's
 * results equal its standalone run bit for bit. The cases share the numerics
 * (num) and differ in mesh (gens[n_cases]); FE is not supported. */
typedef struct mlr_ensemble mlr_ensemble;
int mlr_ensemble_create(const mlr_mesh_gen* gens, int n_cases, const mlr_numerics_desc* num, mlr_ensemble** out);
int mlr_ensemble_destroy(mlr_ensemble* ens);
int mlr_ensemble_sizes(const mlr_ensemble* ens, int64_t* n_cells);  /* [n_cases] */
int mlr_ensemble_init_quadrant(mlr_ensemble* ens, int k, double xs, double ys, const double* states);
int mlr_ensemble_get_state(mlr_ensemble* ens, int k, double* U, double* prim);
int mlr_ensemble_get_time(mlr_ensemble* ens, int k, double* t, int64_t* step);
/* cell volumes of case k, reference numbering */
int mlr_ensemble_get_volumes(mlr_ensemble* ens, int k, double* vol);
int mlr_ensemble_run_steps(mlr_ensemble* ens, int64_t n);
/* kernel-by-kernel with events (the stats layout of mlr_run_steps_profiled) */
int mlr_ensemble_run_steps_profiled(mlr_ensemble* ens, int64_t n, double* stats);
/* launch shapes of the batched kernels and the plain pass's face-basis
 * source, as mlr_autotune / mlr_set_tune (the flux and spectral block sizes
 * are one value here: tune[2] = tune[1]), tune[5] */
int mlr_ensemble_autotune(mlr_ensemble* ens, int32_t* tune);
int mlr_ensemble_set_tune(mlr_ensemble* ens, const int32_t* tune);
int mlr_ensemble_set_indicator(mlr_ensemble* ens, const float* params, double threshold);
int mlr_ensemble_gated(mlr_ensemble* ens, int reset, int64_t* gated);
/* 1: mlr_ensemble_run_steps issues each step as ONE persistent cooperative
 * launch over every case (the spectral radii and dt, then per stage the TENO
 * and flux + update phases, grid-wide barriers between phases) instead of
 * the graph of per-phase launches; results are identical. For ensembles of
 * many small cases, whose per-phase launches are too short to fill the GPU.
 * Plain TENO at p = 2 with the default stencil size and tile width, or first
 * order; fixes the flux and spectral blocks at 256 cells. 0: back to the
 * graph. mlr_ensemble_run_steps_profiled then times each phase from device
 * clock stamps at the barriers. Like mlr_ensemble_run_steps, it has no
 * counterpart in the reference, which steps one case per process
 * (solver/solver.cpp:352-373). */
int mlr_ensemble_set_persistent(mlr_ensemble* ens, int on);
/* blocks of the last persistent launch (0 before the first) */
int mlr_ensemble_persistent_grid(const mlr_ensemble* ens, int32_t* blocks);
/* Watchdog of the persistent step: the longest a block waits at a grid
 * barrier (device wall clock) before it gives up; the launch then ends and
 * mlr_ensemble_run_steps fails with the cases' states invalid. Default
 * 2000 ms, within [0, 10000]. */
int mlr_ensemble_persistent_wait_limit(mlr_ensemble* ens, double ms);

/* Host-only input parsing (no GPU), for the input-compatibility tests.
 * Solver::init's TOML reading (solver.cpp:82-273, mesh.cpp:32-66,
 * face_reconstruction.cpp:109-152): mesh and numerics as the descriptors
 * above; run[7] = cfl (0 when dt is fixed), dt, n_steps (-1: none), t_stop,
 * t_wall_stop, check_interval, check_nan; init_type 0 constant,
 * 1 analytical, 2 quadrant. */
int mlr_config_parse(const char* toml_path, mlr_mesh_gen* mesh, mlr_numerics_desc* num, double* run,
                     int32_t* init_type);
/* The analytical-initialisation expression evaluator (the exprtk subset of
 * solver_initialize.cpp:101-186) at n points (x, y). */
int mlr_expr_eval(const char* expr, int64_t n, const double* x, const double* y, double* out);

/* DataWriter::write_vtu (data_writer.cpp:93-244): the context's state as one
 * VTU file (Float64 cell arrays named RHO RHOU_X RHOU_Y RHOE U_X U_Y P T H
 * CFL, in the order given). The [[write_data]] tables of a TOML input drive
 * the same writer from mlr_run. */
int mlr_write_vtu(mlr_ctx* ctx, const char* path, const char* const* variables, int n_vars);
/* Host-only variant on a layout's mesh with caller arrays (reference
 * numbering: U[N][4], prim[N][5], cfl[N]). */
int mlr_layout_write_vtu(const mlr_layout* lay, const char* path, const char* const* variables, int n_vars,
                         const double* U, const double* prim, const double* cfl);

/* ---- learned troubled-cell indicator (no reference counterpart; the
 * BF16-MFMA MLP that lets cells judged smooth skip the directional TENO
 * stencils). Parameters are the FP32 master vector of 7361 values:
 * W1[48][64], b1[64], W2[64][64], b2[64], w3[64], b3. Features are 48 bf16
 * per cell (12 central-stencil members x 4), labels 1 where plain TENO
 * selected the central stencil for all four variables. */
typedef struct mlr_dataset mlr_dataset;
/* device-resident rows of (features, label) */
int mlr_dataset_create(int64_t capacity, mlr_dataset** out);
int mlr_dataset_destroy(mlr_dataset* ds);
int mlr_dataset_size(const mlr_dataset* ds, int64_t* n);
int mlr_dataset_get(const mlr_dataset* ds, int64_t start, int64_t n, uint16_t* X, float* y);
/* plain TENO decisions of the context's current state, then every cell's
 * features and label appended (new-numbering order) */
int mlr_collect_features(mlr_ctx* ctx, mlr_dataset* ds);
/* the same for the current state, in reference numbering: X[N][48], y[N] */
int mlr_indicator_features(mlr_ctx* ctx, uint16_t* X, float* y);
/* logits of n host rows on the GPU (the inference kernel) */
int mlr_mlp_forward(const float* params, const uint16_t* X, int64_t n, float* z);
/* gradient and loss of the weighted mean BCE over dataset rows idx[0..B):
 * label-1 rows weigh pos_w, label-0 rows neg_w (the training kernels) */
int mlr_mlp_grad(const float* params, const mlr_dataset* ds, const uint32_t* idx, int64_t B, float pos_w,
                 float neg_w, float* grad, float* loss);
/* Adam on the GPU: `steps` batches of `batch` rows drawn from pool by a hash
 * of (seed, step, i); params updated in place; loss_hist[steps] may be NULL */
int mlr_mlp_train(float* params, const mlr_dataset* ds, const uint32_t* pool, int64_t n_pool, int steps, int batch,
                  float lr, uint64_t seed, float pos_w, float neg_w, float* loss_hist);
/* gate the context's TENO with the network (params NULL: off); cells with
 * sigmoid(z) >= threshold take the central stencil alone */
int mlr_set_indicator(mlr_ctx* ctx, const float* params, double threshold);
/* every_stage = 0 (default): the network judges the cells once per step, on
 * the step's first stage, and the later stages reuse its decisions;
 * 1: every stage evaluates it */
int mlr_indicator_cadence(mlr_ctx* ctx, int every_stage);
/* 1 (default): a gated TENO pass runs as two kernels, one for the cells the
 * network gated (no full-TENO code, so it keeps few registers) and one for
 * the other cells, in the 64-cell blocks that hold any; 0: one kernel takes
 * both kinds. The results are identical. */
int mlr_indicator_split(mlr_ctx* ctx, int split);
/* the decisions of the network's last evaluation inside the TENO pass, per
 * cell in reference numbering (1: gated, the central stencil alone) */
int mlr_indicator_decisions(mlr_ctx* ctx, uint8_t* out);
int mlr_ensemble_indicator_cadence(mlr_ensemble* ens, int every_stage);
int mlr_ensemble_indicator_split(mlr_ensemble* ens, int split);
/* cells judged smooth (summed over every gated TENO pass) since the last reset */
int mlr_indicator_gated(mlr_ctx* ctx, int reset, int64_t* gated);
/* cell volumes in reference numbering */
int mlr_get_volumes(const mlr_ctx* ctx, double* vol);

/* Device math probes for the parity tests: 0 sqrt, 1 divide, 2 the
 * double-double pow of the two-rarefaction estimate, 3 x^6, 4 the platform
 * pow. */
int mlr_probe_math(int which, const double* x, const double* y, double* out, int n);
/* Streaming bandwidth probe (no reference counterpart; context for the
 * roofline figures): kind 0 copies `bytes`, kind 1 reads them, with 16-byte
 * loads per lane over a grid-stride loop; kind 2 is hipMemcpyAsync device to
 * device; kind 3 reads 64 contiguous bytes per lane per iteration; kind 4
 * reads with 16-byte non-temporal loads. ms_avg = mean duration of `reps`
 * launches after one warm-up, by HIP events. */
int mlr_probe_stream(int kind, int64_t bytes, int reps, double* ms_avg);
/* Host builds of the same double-double pow and correctly rounded x^6. */
int mlr_host_pow(const double* x, const double* y, double* out, int n);
int mlr_host_pow6(const double* x, double* out, int n);

#ifdef __cplusplus
}
#endif
#endif
