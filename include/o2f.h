/*
 * o2f.h -- C ABI of the MI355X (gfx950) residual / SSP-RK3 / adjoint path.
 *
 * A drop-in for the OpenMP hot loop of ThibaultMaurelOujia/Optimisation2Forme
 * (namespace navier_stokes).  Plain C types only: pointers, sizes, int status.
 * Every array that crosses this boundary is float64 SoA in the reference's
 * own order: cells in Gmsh element order, edges in the order of the
 * reference's std::map keyed by (min vertex, max vertex), vertices in node
 * order (src/mesh.cpp:14-229).  Cell state is [rho | rho_u | rho_v | E],
 * each block n_cells long.  Calls are synchronous (they return with results
 * in host memory) unless documented otherwise.  A context is not
 * thread-safe: one context per host thread.
 *
 * Errors: every function returns O2F_OK (0) or a negative code; the message
 * is in o2f_last_error(ctx) (or o2f_global_error() when no context exists).
 * The reference throws C++ exceptions with the same messages (cited).
 */
#ifndef O2F_H
#define O2F_H

#ifdef __cplusplus
extern "C" {
#endif

#define O2F_OK 0
#define O2F_ERR_ARG -1      /* bad argument / size mismatch               */
#define O2F_ERR_BC -2       /* std::invalid_argument, src/boundary.cpp:90 */
#define O2F_ERR_HIP -3      /* HIP runtime failure / no GPU               */
#define O2F_ERR_IO -4       /* std::runtime_error on file I/O             */
#define O2F_ERR_DIVERGED -5 /* "BiCGSTAB diverged", shape_optimization.cpp:725        */
#define O2F_ERR_NOCONV -6   /* "BiCGSTAB did not converge", shape_optimization.cpp:730 */

#define O2F_LIMITER_VENKAT 0 /* computeVenkatAlpha, src/advection.cpp:209-230 */
#define O2F_LIMITER_BJ 1     /* computeBarthJespersenAlpha, :195-207 */

#define O2F_INTERP_REF 0 /* src/advection.cpp:26-131 exactly: valence >= 5 never written (stays 0) */
#define O2F_INTERP_MVC 1 /* same, with the valence-4 mean-value weights generalised to any valence */

/* How the derived geometry follows a vertex displacement (geometric
 * derivatives below). */
#define O2F_GEOM_REFERENCE 0 /* updateDerivedGeometry (src/shape_optimization.cpp:748-805): only edges touching
                                a wall vertex and their left cells are refreshed */
#define O2F_GEOM_FULL 1      /* every cell and edge follows its vertices (a re-built mesh) */

typedef struct o2f_ctx o2f_ctx;

/* Heat flux of the viscous term.  REF: q = -(mu c_p / Pr) (gamma-1) grad e
 * as computeDiffusionTerm evaluates it (src/diffusion.cpp:233-240), i.e. R
 * times Fourier's law; with the reference's dt limit 0.5 min_h^2 / mu this
 * is unstable once viscosity matters (Re ~ 5000 at chord 1 m).  FOURIER:
 * q = -(mu c_p / Pr) grad T with T = (gamma-1) e / R, R = c_p - c_p / gamma
 * (src/params.cpp:5-8); an extension, not reference behaviour. */
#define O2F_HEAT_FLUX_REF 0
#define O2F_HEAT_FLUX_FOURIER 1

/* Physical and numerical parameters: the members of SimulationParams that
 * the hot path reads (include/params.hpp:15-73), plus the limiter and
 * interpolation choices the reference hard-codes. */
typedef struct o2f_params {
    double gamma;           /* SimulationParams::gamma            (1.4)      */
    double viscosity;       /* ::viscosity, dynamic mu            (1e-6)     */
    double Pr;              /* ::Pr                               (0.71)     */
    double cp;              /* ::cp                               (1005)     */
    double cfl_number;      /* ::cfl_number                       (0.1)      */
    double rho_ref;         /* ::rho_ref                          (1.225)    */
    double p_ref;           /* ::p_ref                            (101325)   */
    double inflow_velocity; /* ::inflow_velocity (a Mach number)  (1.0)      */
    double venkat_k;        /* k_eps of computeVenkatAlpha        (1e-2)     */
    int limiter;            /* O2F_LIMITER_*                      (VENKAT)   */
    int interp_mode;        /* O2F_INTERP_*                       (REF)      */
    int bc_inflow_outflow_x;/* 1 iff boundary_conditions == "inflow_outflow_x"; 0 makes
                               every operator fail like apply_boundary_conditions */
    int heat_flux;          /* O2F_HEAT_FLUX_*                    (REF)      */
} o2f_params;

/* Fills *p with the defaults of SimulationParams (include/params.hpp). */
void o2f_params_default(o2f_params *p);

/* Mesh as the Gmsh API hands it to Mesh::Mesh (src/mesh.cpp:7-179):
 *  xy[2*n_vertices]        node coordinates in node order;
 *  cell_nv[n_cells]        3 (triangle) or 4 (quad), in element order;
 *  cell_vtx[4*n_cells]     vertex indices (0-based), unused slots -1;
 *  bnd_vtx[2*n_bnd]        boundary line elements, in physical-group order;
 *  bnd_tag[n_bnd]          'i','o','f','w' (Inlet, Outlet, FreeStream,
 *                          Wall/Airfoil); the last line touching a vertex
 *                          or an edge sets its tag (src/mesh.cpp:76-110). */
typedef struct o2f_mesh_desc {
    int n_vertices;
    const double *xy;
    int n_cells;
    const int *cell_nv;
    const int *cell_vtx;
    int n_bnd;
    const int *bnd_vtx;
    const char *bnd_tag;
} o2f_mesh_desc;

/* ---- mesh files ---------------------------------------------------- */

/* Reads a Gmsh MSH 2.2 ASCII file into library-owned arrays with the
 * element order of the Gmsh API as used by Mesh::Mesh (physical groups
 * ascending, entities ascending, file order).  Replaces gmsh::open +
 * getNodes/getElements (src/mesh.cpp:9-111).  Free with o2f_msh2_free. */
int o2f_msh2_read(const char *path, o2f_mesh_desc *out);
void o2f_msh2_free(o2f_mesh_desc *desc);

/* Topology and geometry the library derives from a mesh description,
 * exactly as o2f_ctx_create does (Mesh::Mesh, src/mesh.cpp:112-229), computed
 * on the host only (no GPU needed).  Call once with every array NULL to get
 * *n_edges, then again with arrays of that size; any array may be NULL.
 *   edge_vtx[2*ne]    (min, max) vertex of each edge, edges in std::map order
 *   edge_cells[2*ne]  left cell (first recorder), right cell or -1
 *   edge_bc[ne]       '\0', 'i', 'o', 'f' or 'w'
 *   edge_geom[5*ne]   length | n_x | n_y | mid_x | mid_y (n into the left cell)
 *   cell_geom[3*nc]   area | centre_x | centre_y
 *   cell_edges[4*nc]  the cell's edges in ascending id, -1 pad */
int o2f_mesh_topology(const o2f_mesh_desc *mesh, int *n_edges, int *edge_vtx, int *edge_cells,
                      char *edge_bc, double *edge_geom, double *cell_geom, int *cell_edges);

/* A renumbered copy of a mesh for the device's gathers, and the two
 * permutations (new -> old): cell_perm[nc], vertex_perm[nv] (either may be
 * NULL).  No counterpart in the reference, which keeps the mesher's order
 * (Mesh::Mesh, src/mesh.cpp:9-111); its callers are those who load a mesh
 * numbered without locality (a state of the renumbered mesh is
 * U_new[q][i] = U_old[q][cell_perm[i]]).
 *   O2F_REORDER_RCM      cells by reverse Cuthill-McKee over cells sharing a
 *                        vertex, vertices in order of first use
 *   O2F_REORDER_RCM_XCD  the same, then blocks of 64 cells from eight
 *                        contiguous chunks interleaved (every eighth
 *                        workgroup, one XCD, covers one contiguous part)
 * Cell vertex order (orientation) and boundary elements are kept.  The
 * renumbered mesh has other edge ids, so results on it differ from the
 * original's in the last bits (the sums over a cell's edges run in
 * ascending edge id).  Free *out with o2f_msh2_free. */
#define O2F_REORDER_RCM 0
#define O2F_REORDER_RCM_XCD 1
int o2f_mesh_reorder(const o2f_mesh_desc *mesh, int method, o2f_mesh_desc *out, int *cell_perm,
                     int *vertex_perm);

/* ---- context -------------------------------------------------------- */

/* Builds topology and geometry (Mesh::Mesh, src/mesh.cpp:112-248) on the
 * host, uploads it to the current HIP device, allocates state.  Replaces the
 * function-local static buffers of src/advection.cpp:9-23,422-430 and
 * src/diffusion.cpp:9-25,212-221 with per-context device buffers. */
int o2f_ctx_create(const o2f_mesh_desc *mesh, const o2f_params *params, o2f_ctx **out);
void o2f_ctx_destroy(o2f_ctx *ctx);
const char *o2f_last_error(const o2f_ctx *ctx);
const char *o2f_global_error(void);
int o2f_counts(const o2f_ctx *ctx, int *n_vertices, int *n_edges, int *n_cells);
/* Hot reload of parameters (src/simulation.cpp:118-119). */
int o2f_set_params(o2f_ctx *ctx, const o2f_params *params);

/* ---- state ---------------------------------------------------------- */

int o2f_set_state(o2f_ctx *ctx, const double *U);   /* 4*n_cells */
int o2f_get_state(o2f_ctx *ctx, double *U);         /* 4*n_cells */
/* init_one_x: uniform freestream cold start (src/init.cpp:59-81). */
int o2f_init_freestream(o2f_ctx *ctx);
/* loadCellStates / saveCellStates text format (src/utils.cpp:29-107). */
int o2f_load_state_file(o2f_ctx *ctx, const char *path);
int o2f_save_state_file(o2f_ctx *ctx, const char *path);

/* ---- residual (computeAdvectionTerm, computeDiffusionTerm) ---------- */

/* computeAdvectionTerm (src/advection.cpp:415-466): cell gradients, limited
 * face states with boundary ghosts, HLLC, R = -(1/V) sum s F L.  Overwrites
 * the device residual.  R_out (4*n_cells) may be NULL. */
int o2f_advection(o2f_ctx *ctx, double *R_out);
/* computeDiffusionTerm (src/diffusion.cpp:206-243): adds the viscous
 * divergence to the device residual, using the gradients of the last
 * o2f_advection.  R_out may be NULL. */
int o2f_diffusion(o2f_ctx *ctx, double *R_out);
/* computeResidual (src/shape_optimization.cpp:171-189) = both of the above. */
int o2f_residual(o2f_ctx *ctx, double *R_out);
/* Side effects the reference leaves in Cell / Edge after the residual:
 * cell gradients [rho_x|rho_u_x|rho_v_x|E_x|rho_y|rho_u_y|rho_v_y|E_y]
 * (8*n_cells), face states L and R (4*n_edges each, src/advection.cpp:302-317). */
int o2f_get_gradients(o2f_ctx *ctx, double *G);
int o2f_get_face_states(o2f_ctx *ctx, double *face_L, double *face_R);
int o2f_set_face_states(o2f_ctx *ctx, const double *face_L, const double *face_R);

/* compute_hllc_flux (src/advection.cpp:335-412) on a batch of n faces:
 * qL, qR are 4*n SoA conservative states, normal is 2*n SoA (the stored
 * edge normal, pointing into the left cell), F is 4*n.  Runs on the GPU. */
int o2f_hllc_flux_batch(const double *qL, const double *qR, const double *normal, double gamma,
                        double *F, int n);

/* o2f_hllc_flux_batch for callers that must not allocate device memory
 * (serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same arguments and bit-identical results; the faces are
 * staged through a fixed device array of the library (one per GPU, held
 * until the process ends) instead of a hipMalloc'd buffer, so at most
 * O2F_NOALLOC_MAX_FACES faces per call (O2F_ERR_ARG beyond).  Calls on one
 * GPU share that array and are serialised. */
#define O2F_NOALLOC_MAX_FACES 16384
int o2f_hllc_flux_batch_noalloc(const double *qL, const double *qR, const double *normal, double gamma,
                                double *F, int n);

/* ---- time integration (Simulator::time_loop / step_SSP_RK3) --------- */

/* dt = min(CFL min_h / max_wave, 0.5 min_h^2 / mu) (src/simulation.cpp:77-110). */
int o2f_compute_dt(o2f_ctx *ctx, double *dt);
/* step_SSP_RK3 with a given dt (src/simulation.cpp:130-171). */
int o2f_rk3_step(o2f_ctx *ctx, double dt);
/* The decomposition of step_SSP_RK3: the U_old <- U copy that opens it
 * (src/simulation.cpp:130-133), then Simulator::rk3_substep(c0, c1, dt)
 * (:145-171): computeAdvectionTerm + computeDiffusionTerm on the current U
 * (gradients, face states and R left as those operators leave them), then
 * U = c0 U_old + c1 (U + dt R).  o2f_rk3_save_old followed by substeps
 * (0, 1), (3/4, 1/4), (1/3, 2/3) is bitwise o2f_rk3_step. */
int o2f_rk3_save_old(o2f_ctx *ctx);
int o2f_rk3_substep(o2f_ctx *ctx, double c0, double c1, double dt);
/* nsteps x { dt reduction; SSP-RK3 step; t += dt }, device resident, no host
 * synchronisation inside.  dts (nsteps) may be NULL; t_inout (may be NULL)
 * carries the simulated time. */
int o2f_run_steps(o2f_ctx *ctx, int nsteps, double *dts, double *t_inout);
/* while (t < final_time) { ... } of src/simulation.cpp:75-124 without the
 * output / reload side branches; returns steps taken and final t. */
int o2f_time_loop(o2f_ctx *ctx, double final_time, int max_steps, int *steps, double *t);
/* The same loop, device resident: steps are launched in batches and a
 * device-side flag stops them once t >= final_time, so the host synchronises
 * once per batch, not once per step.  *t_inout carries t in and out;
 * *t_last_start (may be NULL) is t at the start of the last step and *dt_last
 * (may be NULL) its dt (the values output_step prints, src/simulation.cpp:115-116). */
int o2f_run_until(o2f_ctx *ctx, double final_time, int max_steps, int *steps, double *t_inout,
                  double *t_last_start, double *dt_last);
/* sqrt(mean_c R_rho^2) / (rho_inf U_inf / c), c = 1: the density residual
 * norm of the last residual evaluation (convergence criterion, not in the
 * reference, which has none). */
int o2f_residual_norm(o2f_ctx *ctx, double *rnorm);

/* ---- forces (computeLiftDragCoefficients, computeDirectObjectiveTerm) - */

/* src/compute.cpp:92-132: wall pressure + viscous forces from the face L
 * states of the last residual. */
int o2f_lift_drag(o2f_ctx *ctx, double *C_L, double *C_D);
/* src/shape_optimization.cpp:893-939: J = -C_L + beta C_D over wall edges
 * whose centre has x <= 1.3. */
int o2f_objective(o2f_ctx *ctx, double beta, double *J);

/* ---- discrete adjoint ----------------------------------------------- */

/* assembleAdjointRHS (src/shape_optimization.cpp:491-554); b is 4*n_cells. */
int o2f_adjoint_rhs(o2f_ctx *ctx, double beta, double *b);
/* applyJacobianTranspose (src/shape_optimization.cpp:237-484): y = J^T x
 * with central finite-difference (absolute step eps) HLLC blocks at the
 * current face states.  x, y are 4*n_cells. */
int o2f_apply_jt(o2f_ctx *ctx, const double *x, double *y, double eps);
/* Exact discrete adjoint: y = (dR/dU)^T x at the current state, R the full
 * residual of o2f_residual (computeResidual, src/shape_optimization.cpp:
 * 171-189).  Unlike applyJacobianTranspose (:237-484), which freezes the
 * gradients and limiters and finite-differences HLLC, every dependence of R
 * on U is differentiated exactly (interpolation, Green-Gauss, limiter
 * through its active branch, HLLC, ghosts, stresses, heat flux); limiter
 * state is recomputed per cell, not stored.  x, y are 4*n_cells SoA. */
int o2f_apply_jt_exact(o2f_ctx *ctx, const double *x, double *y);
/* The exact (dJ/dU)^T of J = -C_L + beta C_D as computeDirectObjectiveTerm
 * evaluates it (src/shape_optimization.cpp:893-939: pressure of the wall
 * face L states, wall edges with centre x > 1.3 skipped), differentiated
 * through the face-state reconstruction (limited gradients, interpolation,
 * boundary states) by the exact adjoint chain.  assembleAdjointRHS
 * (:491-554) keeps only the face-state derivative, taken as the cell
 * state's, on every wall edge.  b (4*n_cells SoA) out. */
int o2f_adjoint_rhs_exact(o2f_ctx *ctx, double beta, double *b);
/* Options of o2f_bicgstab_exact (bitwise or). */
#define O2F_ADJ_EXACT_RHS 1    /* right-hand side o2f_adjoint_rhs_exact (else assembleAdjointRHS's) */
#define O2F_ADJ_BLOCK_JACOBI 2 /* right preconditioner: the inverse of each cell's 4x4 diagonal block of the
                                  operator (else the identity, as solveAdjointSystem's
                                  IdentityPreconditioner) */
/* solveAdjointSystem's BiCGSTAB (chunks, true residual, same error codes as
 * o2f_bicgstab) on the exact operator of o2f_apply_jt_exact; options:
 * O2F_ADJ_* flags (0: the reference's right-hand side, no
 * preconditioner). */
int o2f_bicgstab_exact(o2f_ctx *ctx, double beta, double tol, int chunk, int max_total, double *lambda,
                       int *iters, double *relres, int options);
/* The 4x4 diagonal blocks D_c = dR_c/dU_c (16 per cell, row q = residual
 * variable, column r = state variable) of the exact operator (exact = 1)
 * or of applyJacobianTranspose's (exact = 0), obtained by probing J^T with
 * cells coloured two vertex rings apart (the block-Jacobi preconditioner's
 * input).  No counterpart in the reference. */
int o2f_adjoint_block_diagonal(o2f_ctx *ctx, int exact, double *blocks);
/* The epsilon of the orthogonality restart test |r0.r| < eps^2 |r0|^2 of
 * every BiCGSTAB of this context (Eigen's internal::bicgstab, called by
 * solveAdjointSystem, src/shape_optimization.cpp:699-717, uses
 * NumTraits<double>::epsilon(), the default here).  A larger value makes
 * the restart branch reachable on small test systems.  O2F_ERR_ARG for a
 * negative or NaN value. */
int o2f_bicgstab_set_restart_eps(o2f_ctx *ctx, double restart_eps);
/* solveAdjointSystem (src/shape_optimization.cpp:658-740): BiCGSTAB with an
 * identity preconditioner on J^T lambda = b, restarted every `chunk`
 * iterations up to `max_total`, true residual checked after each chunk.
 * Reference values: eps 1e-6, tol 1e-6, chunk 500, max_total 100000.
 * lambda (4*n_cells) out; iters, relres out.  Returns O2F_ERR_DIVERGED or
 * O2F_ERR_NOCONV like the reference's throws (lambda still written). */
int o2f_bicgstab(o2f_ctx *ctx, double beta, double eps, double tol, int chunk, int max_total,
                 double *lambda, int *iters, double *relres);
/* o2f_bicgstab with options: O2F_ADJ_BLOCK_JACOBI right-preconditions the
 * reference operator by the inverses of its 4x4 diagonal blocks (probed
 * exactly, o2f_adjoint_block_diagonal).  solveAdjointSystem
 * (src/shape_optimization.cpp:658-740) uses the identity; on this operator
 * the blocks map the spectrum into the right half-plane and cut the
 * iterations by an order of magnitude.  Other flags: O2F_ERR_ARG. */
int o2f_bicgstab_options(o2f_ctx *ctx, double beta, double eps, double tol, int chunk, int max_total, int options,
                         double *lambda, int *iters, double *relres);
/* computeShapeGradient (src/shape_optimization.cpp:942-1016) with the
 * NACA 4-digit wall perturbation of applyShapePerturbationNACA12 (:828-890).
 * names_all/values_all: every key of the shape-parameter file (pos_x, c, m,
 * p, t, beta ...); opt_names: the keys marked "*=", in file order.
 * grad[n_opt] out.  Leaves the face states at the current state. */
int o2f_shape_gradient(o2f_ctx *ctx, const char *const *names_all, const double *values_all,
                       int n_all, const char *const *opt_names, int n_opt, const double *lambda,
                       double beta, double eps, double *grad);

/* ---- geometric derivatives (forward mode) ----------------------------- */

/* Wall-vertex displacement d(x, y)/d(param) of applyShapePerturbationNACA12
 * (src/shape_optimization.cpp:807-890) for param "m" / "cambrure", "p" /
 * "position_cambrure" or "t" (thickness: an extension, the reference
 * perturbs only m and p): y = yc(m, p, x_r) + side * yt(t, x_r) with x
 * fixed, zero for x_r >= 0.9 and off the wall.  names_all / values_all as
 * for o2f_shape_gradient (pos_x, c, m, p, t defaults :830-834).  dxy
 * (2*n_vertices, x then y per vertex) out.  O2F_ERR_ARG for another name. */
int o2f_naca_displacement(o2f_ctx *ctx, const char *name, const char *const *names_all,
                          const double *values_all, int n_all, double *dxy);
/* Directional derivatives at the current state U (held fixed) along the
 * vertex displacement dxy (2*n_vertices): dR = (dR/dX) dxy (4*n_cells SoA)
 * and dJ = (dJ/dX) dxy for J = -C_L + beta C_D (computeDirectObjectiveTerm,
 * src/shape_optimization.cpp:893-939), by forward-mode (dual number)
 * differentiation of the geometry (host) and of the whole residual
 * (device: interpolation weights, Green-Gauss, limiter, face states, HLLC
 * with its moving normal, stresses, viscous and convective sums).  This is
 * the eps -> 0 limit of computeShapeGradient's central differences
 * (:942-1016) without the mesh copies.  geometry: O2F_GEOM_*.  R (may be
 * NULL) receives the residual's value, bit-identical to o2f_residual's.
 * Any of R, dR, dJ may be NULL.  Leaves the face states at the current
 * state. */
int o2f_residual_tangent(o2f_ctx *ctx, const double *dxy, int geometry, double beta, double *R, double *dR,
                         double *dJ);
/* Jacobian-vector product y = (dR/dU) dU at the current state by the same
 * forward-mode (dual number) residual, with the geometry fixed: the
 * forward counterpart of o2f_apply_jt_exact, so <x, J y> = <J^T x, y> holds
 * to rounding.  dU, dR: 4*n_cells SoA; R (may be NULL) receives the
 * residual, bit-identical to o2f_residual's.  No counterpart in the
 * reference (src/shape_optimization.cpp:192-227 finite-differences it). */
int o2f_residual_jvp(o2f_ctx *ctx, const double *dU, double *R, double *dR);
/* computeShapeGradient's g_k = dJ/dk - lambda . dR/dk with both
 * derivatives from o2f_residual_tangent along o2f_naca_displacement (one
 * tangent evaluation per parameter instead of two perturbed residuals and
 * mesh copies).  Arguments as o2f_shape_gradient; "t" is differentiable;
 * other names give 0. */
int o2f_shape_gradient_tangent(o2f_ctx *ctx, const char *const *names_all, const double *values_all, int n_all,
                               const char *const *opt_names, int n_opt, const double *lambda, double beta,
                               int geometry, double *grad);

/* ---- run to convergence --------------------------------------------- */

/* The time loop of src/simulation.cpp:75-124 driven to a steady state
 * instead of a final time (the reference stops only at final_time): every
 * check_every steps (steps 0, k, 2k, ...) the density residual
 * r = sqrt(mean_c R_rho(U^n)^2) / (rho_inf U_inf / c) of the step's first
 * stage is reduced on the device; a step with r <= tol is completed and the
 * loop stops after it.  *steps = steps taken, *res = last r evaluated.
 * tol <= 0 runs max_steps steps.  use_graph: replay check_every steps as one
 * captured hipGraph.  Device resident, one host synchronisation per group of
 * steps. */
int o2f_run_to_residual(o2f_ctx *ctx, double tol, int max_steps, int check_every, int use_graph, int *steps,
                        double *res);

/* ---- batches of shape candidates -------------------------------------- */

/* A batch advances several contexts (meshes of any sizes, flow conditions
 * of their own, all on the current device) together: each stage kernel is
 * ONE launch over all members' workgroup tiles, optionally persistent and
 * graph-captured.  Every member's results are bit-identical to its own
 * single-mesh run (same per-entity arithmetic and sum orders); launch shapes
 * change only speed.  Members must outlive the batch and must not be used
 * on their own while a batch call runs.  No counterpart in the reference,
 * which solves one shape per process (run_opti.sh:211-254). */
typedef struct o2f_batch o2f_batch;
int o2f_batch_create(o2f_ctx *const *members, int n_members, o2f_batch **out);
void o2f_batch_destroy(o2f_batch *batch);
const char *o2f_batch_last_error(const o2f_batch *batch);
/* Launch shapes of the four batched stage kernels (vertex gather, cell
 * reconstruction, viscous gather + edge flux, flux + update): shapes[2k] =
 * threads per workgroup (64, 128 or 256), shapes[2k+1] = persistent grid
 * size (0: one workgroup per tile). */
int o2f_batch_set_shapes(o2f_batch *batch, const int *shapes);
int o2f_batch_get_shapes(const o2f_batch *batch, int *shapes);
/* Times every shape for each kernel on residual-only passes (the state is
 * not changed) and keeps the fastest; shapes (8, may be NULL) out.  On the
 * automatic tile stages (o2f_batch_set_tile_cells -1) it also times the
 * rule's tile size and 90 % and 80 % of it on first-stage launches from the
 * members' states (into scratch buffers) and keeps the fastest for the
 * batch's runs (o2f_batch_tile_info reports it). */
int o2f_batch_autotune(o2f_batch *batch, int reps, int *shapes);
/* o2f_run_to_residual for every member at once: steps[i], res[i] per member
 * (may be NULL). */
int o2f_batch_run(o2f_batch *batch, double tol, int max_steps, int check_every, int use_graph, int *steps,
                  double *res);
/* Simulator::time_loop's while (t < final_time) (src/simulation.cpp:75-124)
 * for every member at once, from t = 0, each member stopping at its own
 * step (its own dt); at most max_steps steps.  steps[i], t[i] per member
 * (may be NULL). */
int o2f_batch_run_until(o2f_batch *batch, double final_time, int max_steps, int use_graph, int *steps, double *t);
/* Elapsed milliseconds of nsteps batched steps (tol <= 0, no early stop). */
int o2f_batch_time_steps(o2f_batch *batch, int nsteps, int use_graph, double *total_ms);
/* The batch's steps on fused tile stages (no counterpart in the reference):
 * one launch per SSP-RK3 stage over the cell tiles of every member, the
 * members' tile tables concatenated (a tile as in o2f_set_tile_cells).
 * cells: -1 (the default) the largest tiles of which one CU holds as many
 * as the kernel's registers allow, 0 the four-kernel batched stages (whose
 * launch shapes o2f_batch_set_shapes sets), n > 0 at most n cells per tile.
 * A batch with a member that cannot be tiled runs the four-kernel stages.
 * Every member's results are bitwise the same either way; after a run the
 * members hold the intermediates of their last stage as on the four-kernel
 * stages. */
int o2f_batch_set_tile_cells(o2f_batch *batch, int cells);
/* The adjoint solve of every member (no counterpart in the reference, which
 * solves one shape per process, run_opti.sh:211-254): each member's
 * o2f_bicgstab (reference-form operator, eps as there) or, with exact = 1,
 * its o2f_bicgstab_exact (no options), the members' solves driven from
 * several host threads at once, each on the member's own stream, so that
 * the solves of meshes too small to fill the GPU overlap on it.  Each
 * member's lambda, iterations and relative residual are bitwise those of its
 * own call.  lambdas[i] (4 x nc of member i, host; the array or an entry
 * may be NULL), iters, relres and status (per member: O2F_OK,
 * O2F_ERR_NOCONV, O2F_ERR_DIVERGED; may be NULL).  Returns O2F_OK unless
 * an argument or a device call failed. */
int o2f_batch_bicgstab(o2f_batch *batch, double beta, double eps, double tol, int chunk, int max_total, int exact,
                       double *const *lambdas, int *iters, double *relres, int *status);
/* The tiling in use (built now if needed): *n_tiles 0 on the four-kernel
 * stages (o2f_batch_last_error says why), else the tiles over all members,
 * the cell cap, the LDS bytes of one tile, the tiles one CU holds at once
 * and the kernel's VGPRs. */
int o2f_batch_tile_info(o2f_batch *batch, int *n_tiles, int *cells, int *lds_bytes, int *tiles_per_cu, int *vgprs);

/* ---- learned warm start ---------------------------------------------- */

/* A per-cell MLP mapping (x, y, wall distance, m, p, t, Mach, AoA in
 * degrees) plus 16 Fourier features of (x, y) to the conservative state
 * (normalised: rho/rho_inf - 1, rho u/(rho_inf a_inf) - M,
 * rho v/(rho_inf a_inf), (E - E_inf)/p_inf), 4 hidden layers of 256 with
 * SiLU, trained with Adam on fields the solver produced.  BF16 operands on
 * the matrix cores with FP32 accumulation and FP32 master weights.  Its
 * prediction replaces the freestream cold start (init_one_x,
 * src/init.cpp:59-81) as the initial state; the solver still iterates to
 * the same residual tolerance, so only the step count changes.  No
 * counterpart in the reference.  Weights: 5 layers W_l (K_l x N_l,
 * row-major) then the 5 biases, fp32 (o2f_mlp_num_params gives the
 * layout: K, N, weight offset, bias offset per layer). */
typedef struct o2f_mlp o2f_mlp;
int o2f_mlp_create(unsigned long long seed, o2f_mlp **out);
void o2f_mlp_destroy(o2f_mlp *net);
const char *o2f_mlp_last_error(const o2f_mlp *net);
int o2f_mlp_num_params(const o2f_mlp *net, long *n_params, int *layout /* 20, may be NULL */);
/* fp32 weights (n_params) and the Fourier frequencies (2 x 16: x row then y row). */
int o2f_mlp_get(o2f_mlp *net, float *weights, float *fourier);
int o2f_mlp_set(o2f_mlp *net, const float *weights);
/* Training set: raw (n x 8 inputs above), targets (n x 4 normalised state). */
int o2f_mlp_set_data(o2f_mlp *net, const float *raw, const float *targets, long n);
/* steps Adam updates (lr, beta 0.9 / 0.999, eps 1e-8) on minibatches of
 * `batch` samples (a multiple of 64) drawn from the training set by a
 * seeded hash; losses[steps] (may be NULL) receives each step's MSE. */
int o2f_mlp_train(o2f_mlp *net, int steps, int batch, float lr, unsigned long long seed, float *losses);
/* out (n x 4) = network(raw (n x 8)). */
int o2f_mlp_forward(o2f_mlp *net, const float *raw, long n, float *out);
/* Mean squared error over the 4 outputs of n samples and its gradient with
 * respect to every parameter (n_params, fp32). */
int o2f_mlp_loss_grad(o2f_mlp *net, const float *raw, const float *targets, long n, float *loss, float *grad);
/* Sets the context's state to the network's prediction for its cells
 * (shape m, p, t and angle of attack as inputs; Mach from the context). */
int o2f_mlp_warm_start(o2f_mlp *net, o2f_ctx *ctx, double m, double p, double t, double aoa_deg);
/* The GEMM kernel choice: -1 (default) times the admissible kernels once
 * per shape and keeps the fastest; 0-6 forces one wherever its tile divides
 * the shape, the 64x64 kernel 0 elsewhere (0-2: single-buffered 64x64 /
 * 128x64 / 128x128 workgroup tiles with K stages of 32; 3-6:
 * double-buffered 128x128 / 128x64 / 64x128 / 64x64 with K stages of 64).
 * Every kernel accumulates each output in the same k order, so the choice
 * changes speed only. */
int o2f_mlp_set_gemm_variant(o2f_mlp *net, int variant);

/* ---- measurement ----------------------------------------------------- */

/* Runs nsteps device-resident SSP-RK3 steps (as o2f_run_steps) between HIP
 * events on the context's stream and returns the elapsed milliseconds.
 * kernel_ms[4] (may be NULL) receives the summed durations of the four
 * per-stage kernels (vertex gather, cell reconstruction, vertex viscous
 * gather, flux + update), each bracketed by its own events. */
int o2f_time_steps(o2f_ctx *ctx, int nsteps, double *total_ms, double *kernel_ms);
/* nprod device-resident products y = J^T x at the current state between HIP
 * events: exact = 0 the reference-form operator (FD blocks, built once
 * before timing), 1 the exact adjoint.  Returns the elapsed milliseconds. */
int o2f_time_adjoint(o2f_ctx *ctx, int nprod, int exact, double *total_ms);
/* `launches` back-to-back launches of one per-stage kernel at the current
 * state between a single pair of HIP events (no event between launches, so
 * the figure is the kernel's own duration plus the dispatch gap, as a
 * kernel trace reports it): kernel 0 the vertex gather, 1 the cell
 * reconstruction, 2 the vertex viscous gather + edge flux, 3 the flux sum
 * and viscous divergence in residual form (no stage update), 4 the whole
 * residual (those four launches in order, per "launch"), 5 one fused tile
 * stage (residual and stage update of every cell in one launch, written to
 * a spare state buffer; O2F_ERR_ARG when the context does not tile).  Every
 * launch recomputes the same outputs from the same inputs.  *avg_us
 * receives the microseconds per launch.  A measurement entry point of this
 * library (no counterpart in the reference). */
int o2f_time_kernel(o2f_ctx *ctx, int kernel, int launches, double *avg_us);

/* ---- fused tile stages (no counterpart in the reference) ------------ */

/* The time-stepping entry points (o2f_run_steps, o2f_run_until,
 * o2f_time_loop, o2f_time_steps) evaluate each SSP-RK3 stage in one launch
 * over compact tiles of cells, one workgroup per tile, with the stage's
 * intermediates in LDS (the mesh's rim entities formed by every tile that
 * needs them).  Results are bit-identical to the four-kernel stages.
 * cells: -1 (the default) sizes the tiles from the device's CU count and
 * the kernel's occupancy, 0 uses the four-kernel stages, n > 0 caps a tile
 * at n cells.  The tables are built on the next step. */
int o2f_set_tile_cells(o2f_ctx *ctx, int cells);
/* The tiling in use (built now if needed): *n_tiles 0 when the stages are
 * the four-kernel ones (o2f_last_error then says why, unless cells was 0);
 * the cell cap, the LDS bytes of one tile, the tiles one CU holds at once
 * and the kernel's VGPRs.  Any pointer may be NULL. */
int o2f_tile_info(o2f_ctx *ctx, int *n_tiles, int *cells, int *lds_bytes, int *tiles_per_cu, int *vgprs);

#ifdef __cplusplus
}
#endif

#endif /* O2F_H */
