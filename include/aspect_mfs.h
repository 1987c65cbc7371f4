/*
 * aspect_mfs.h — C-ABI of the MI355X (gfx950) matrix-free Stokes solver.
 *
 * This is the drop-in boundary for ASPECT's StokesMatrixFreeHandler
 * (include/aspect/simulator/solver/stokes_matrix_free.h:47-98): host C++ on the
 * ASPECT side (a StokesMatrixFreeHandler subclass returned by
 * create_matrix_free_solver, source/simulator/solver/stokes_matrix_free.cc:29-62)
 * calls these functions with plain pointers and sizes.  No torch or HIP types
 * appear in the signatures.
 *
 * Conventions
 *   - Every function returns an int status: MFS_OK, MFS_ERROR (message in
 *     mfs_last_error()), or MFS_NO_CONVERGENCE (solve only; result filled).
 *   - Host buffers are caller-owned and copied in/out; device buffers are
 *     handle-owned.  *_device entry points take device pointers that the caller
 *     owns (allocated with hipMalloc on the handle's device).
 *   - One handle per host thread; calls on one handle are not reentrant.
 *   - Vector layout: velocity block first, component-major, nodes
 *     lexicographic with x fastest ((k*n_x+1) x (k*n_y+1) [x (k*n_z+1)] nodes per
 *     component); the pressure block follows ((k-1)*n+1 nodes per direction).
 *     Per-cell data: cells lexicographic with x fastest; per-quadrature-point
 *     data: [cell][q] with q lexicographic over the QGauss(k+1) points, x fastest.
 */
#ifndef ASPECT_MFS_H
#define ASPECT_MFS_H

#ifdef __cplusplus
extern "C" {
#endif

#define MFS_OK 0
#define MFS_ERROR 1
#define MFS_NO_CONVERGENCE 2

/* Krylov method of the cheap solver steps (Parameters::StokesKrylovType). */
#define MFS_KRYLOV_GMRES 0
#define MFS_KRYLOV_IDRS 1

/* GMGNumberType of the V-cycle (stokes_matrix_free.h:40 fixes it to double). */
#define MFS_GMG_FP64 0
#define MFS_GMG_FP32 1

/* Material averaging applied to per-quadrature-point viscosities
 * (source/material_model/interface.cc:556-680). */
#define MFS_AVERAGE_NONE 0
#define MFS_AVERAGE_ARITHMETIC 1
#define MFS_AVERAGE_HARMONIC 2
#define MFS_AVERAGE_GEOMETRIC 3
/* "project to Q1 only viscosity": cellwise L2 projection onto Q1 (DGQ1), values
 * at the quadrature points clamped to the evaluated range; level tables per
 * quadrature point (matrix_free_operators.cc:200-326, local_smoothing.cc:211-294). */
#define MFS_AVERAGE_PROJECT_Q1 4

/* Boundary type per face (x-, x+, y-, y+, z-, z+). */
#define MFS_BC_OPEN 0
#define MFS_BC_TANGENTIAL 1 /* "Tangential velocity boundary indicators" (no normal flux) */
#define MFS_BC_ZERO 2       /* "Zero velocity boundary indicators" */

typedef struct mfs_handle mfs_handle;

/* A box geometry refined globally (GeometryModel::Box with repetitions, plus
 * "Initial global refinement"); the level hierarchy is level 0 = coarse mesh. */
typedef struct {
  int dim;                 /* 2 or 3 */
  int degree;              /* "Stokes velocity polynomial degree": 2 or 3 */
  int coarse[3];           /* cells per direction on level 0 */
  int n_levels;            /* levels; finest has coarse << (n_levels-1) cells */
  double extent[3];        /* box extents, origin at 0 */
  int bc[6];               /* MFS_BC_* per face */
  int compressible;        /* subtract (2/3) eta div(u) I (matrix_free_operators.cc:466-469) */
  double pressure_scaling; /* Simulator::get_pressure_scaling() */
} mfs_box;

/* "Linear solver failure strategy" (source/simulator/parameters.cc:287-292),
 * applied when the expensive steps fail (local_smoothing.cc:1353-1395). */
#define MFS_FAILURE_ABORT 0    /* status MFS_NO_CONVERGENCE, message of Utilities::throw_linear_solver_failure_exception */
#define MFS_FAILURE_CONTINUE 1 /* "continue with nonlinear solver": status MFS_OK, result.converged = 0 */

typedef struct {
  double tolerance;         /* "Linear solver tolerance" */
  int n_cheap_steps;        /* "Number of cheap Stokes solver steps" */
  int n_expensive_steps;    /* "Maximum number of expensive Stokes solver steps" */
  int restart;              /* "GMRES solver restart length" */
  double a_block_tolerance; /* "Linear solver A block tolerance" */
  double s_block_tolerance; /* "Linear solver S block tolerance" */
  int cheap_krylov;         /* "Krylov method for cheap solver steps": MFS_KRYLOV_GMRES or MFS_KRYLOV_IDRS
                               (stokes_matrix_free_local_smoothing.cc:1276-1301) */
  int idr_s;                /* "IDR(s) parameter", 1..8 (used with MFS_KRYLOV_IDRS) */
  int solve_newton_system;  /* solve()'s argument (interface.h:89-94): zero initial guess, tolerance relative
                               to ||rhs||, initial_nonlinear_residual = ||rhs||, no pressure normalization
                               (local_smoothing.cc:950-969, 1028-1041, 1443-1444) */
  int failure_strategy;     /* MFS_FAILURE_ABORT or MFS_FAILURE_CONTINUE */
  const char *history_file; /* MFS_FAILURE_ABORT: residual history of every solver phase that ran
                               ("step residual" lines, a blank line between phases), as ASPECT writes
                               solver_history.txt (source/utilities.cc:2854-2876); NULL: none */
} mfs_solve_params;

/* StokesSolver::SolverOutputs (include/aspect/simulator/solver/interface.h:36-58)
 * plus the counters ASPECT's post_stokes_solver signal reports. */
typedef struct {
  int cheap_iterations;
  int expensive_iterations;
  int a_preconditioner_applications;
  int s_preconditioner_applications;
  int converged;
  double initial_nonlinear_residual;
  double final_linear_residual;
  double solver_tolerance;
  double r0_norm;
  double pressure_normalization_adjustment;
  double solve_seconds; /* device time of the Krylov solve (hipEvent) */
  double cheap_final_residual; /* residual norm at the end of the cheap phase */
  double cheap_seconds;        /* device time of the cheap phase (hipEvent) */
} mfs_solve_result;

/* Error message of the last failing call on this thread. */
const char *mfs_last_error(void);

/* Operator codes of mfs_validate_operator_arguments. */
#define MFS_OP_STOKES 0
#define MFS_OP_ABLOCK 1
#define MFS_OP_BT 2
#define MFS_OP_MASS 3
#define MFS_OP_ADIAG 4
#define MFS_OP_MDIAG 5
#define MFS_OP_RHS 6
#define MFS_OP_LAPLACE 7 /* vector Laplace of the mesh-deformation solve */
#define MFS_OP_LDIAG 8   /* its diagonal */
/* Operator modes: 0 dst = op(src), 1 dst = rhs - op(src), 2 dst += op(src). */
#define MFS_MODE_SET 0
#define MFS_MODE_RESID 1
#define MFS_MODE_ADD 2

/* Host-side check that every pointer the (op, mode) kernel reads or writes is
 * non-null; the library runs it before every operator launch and the *_device
 * entry points return MFS_ERROR instead of launching on a null pointer.
 * Needs no device.  No counterpart in the reference (deal.II's FEEvaluation
 * takes its vectors by reference). */
int mfs_validate_operator_arguments(int op, int mode, const void *src_u, const void *src_p, const void *eta,
                                    const void *rhs_u, const void *rhs_p, const void *dst_u, const void *dst_p);

/* The DoF map between a deal.II DoF numbering and this library's vectors
 * (the reference's setup_dofs distributes and renumbers the velocity and
 * pressure DoFs hierarchically, local_smoothing.cc:1487-1491 and :1557-1561):
 * for n DoFs with support points points[3 i + d] (DoFTools::
 * map_dofs_to_support_points; z ignored in 2D) and vector component
 * component[i] (0..dim-1 velocity, dim pressure), index[i] = position of the
 * DoF in the block vector of the box's finest level (velocity component-major,
 * nodes lexicographic with x fastest, then pressure).  FE_Q(k) support points
 * are Gauss-Lobatto per cell.  Returns MFS_ERROR if a point is not a node of
 * the box or two DoFs map to one position.  Host only, needs no device.
 * No counterpart in the reference. */
int mfs_box_dof_index(const mfs_box *box, long long n, const double *points, const int *component,
                      long long *index);

/* setup_dofs(): stokes_matrix_free_local_smoothing.cc:1450-1851 — level meshes,
 * DoF layout, level constraints and device storage. */
int mfs_create(const mfs_box *box, int gmg_precision, mfs_handle **out);
int mfs_destroy(mfs_handle *h);

/* name(): local_smoothing.cc:146-151 ("GMG"). Writes at most len bytes. */
int mfs_name(const mfs_handle *h, char *buf, int len);

/* Sizes of a level: out[0] cells, out[1] velocity DoFs, out[2] pressure DoFs,
 * out[3] velocity nodes. */
int mfs_sizes(const mfs_handle *h, int level, long long out[4]);

/* assemble() / evaluate_material_model(): local_smoothing.cc:173-294 and
 * fill_active_cell_data (matrix_free_operators.cc:137-330).  eta holds the
 * material model's viscosity at every quadrature point of the active level
 * ([cell][q]); averaging selects the active table (per quadrature point for
 * MFS_AVERAGE_NONE, otherwise one value per cell).  Level tables are the DGQ0
 * projection restricted by averaging children. */
int mfs_set_viscosity(mfs_handle *h, const double *eta_qp, int averaging);
int mfs_set_pressure_scaling(mfs_handle *h, double pressure_scaling);

/* General geometry of one level: the MappingQ data that MatrixFree stores per
 * quadrature point (update_gradients | update_JxW_values; active level
 * stokes_matrix_free_local_smoothing.cc:1643-1656, multigrid levels with the
 * level mapping :1699-1760): geo
 * holds, for every cell of the level (lexicographic) and every QGauss(k+1)
 * point (lexicographic), dim*dim + 1 values: J^{-1} row-major (entry
 * e*dim + d is d xi_e / d x_d, xi the unit cell coordinates) and JxW (det J
 * times the quadrature weight).  Every operator, diagonal, right-hand side and
 * level viscosity projection of that level then uses them (the uniform-box
 * fast paths are bypassed).  Set every level, then call mfs_set_viscosity;
 * faces of type MFS_BC_TANGENTIAL and the free-surface term need the box
 * geometry and are refused. */
int mfs_set_geometry(mfs_handle *h, int level, const double *geo);

/* Free-surface stabilization of the active-level Stokes operator,
 * StokesOperator::local_apply_boundary_face (matrix_free_operators.cc:563-602):
 * on every face with faces[f] = 1 (order x-, x+, y-, y+, z-, z+) the operator
 * adds -(t . u)(phi . n) integrated over the face (QGauss(k+1) on the face).
 * stab holds t: for each marked face in that order, for each face cell
 * (lexicographic over the two tangential directions, the lower direction
 * fastest), for each face quadrature point (same order), dim values (ASPECT
 * fills rho dt theta |g| g_hat, local_smoothing.cc:527-608).  It enters
 * mfs_stokes_vmult, the solver's Stokes operator and mfs_correct_stokes_rhs,
 * not the multigrid level operators (as in the reference).  All faces 0
 * switches it off; stab is copied. */
int mfs_set_free_surface(mfs_handle *h, const int faces[6], const double *stab);

/* Mesh-deformation displacement solve on the handle's mesh,
 * MeshDeformationHandler::compute_mesh_displacements_gmg_for_degree
 * (source/mesh_deformation/interface.cc:1076-1360): the vector Laplace
 * operator (dim components of FE_Q(k), k the Stokes velocity degree,
 * coefficient 1) with geometric multigrid (Chebyshev degree 5 over
 * [1.2 lambda_max / 15, 1.2 lambda_max] from 10 eigenvalue iterations; on
 * level 0 Chebyshev as a solver, 50 eigenvalue iterations, range 1e-3, degree
 * from the Chebyshev error bound) inside CG to tolerance * ||rhs|| within
 * 5 N iterations.  bc per face (x-, x+, y-, y+, z-, z+): 0 none, 1 no normal
 * displacement, 2 zero displacement, 3 prescribed displacement (values taken
 * from u0 on those DoFs; the inhomogeneous constraints are moved to the
 * right-hand side, rhs = -A u0, as the reference does at :1116-1156).  u0 and
 * sol hold dim * velocity-nodes values (the velocity layout); only the
 * constrained entries of u0 are read.  iterations: CG steps; residuals[0..1]:
 * final residual and required residual.  Returns MFS_NO_CONVERGENCE (with the
 * result filled) when CG stops at the iteration limit. */
int mfs_mesh_deformation_solve(mfs_handle *h, const int bc[6], const double *u0, double tolerance, double *sol,
                               int *iterations, double *residuals);
/* Its pieces for parity checks: the level Laplace operator (Base::vmult
 * semantics: constrained rows copy src) and one V-cycle of the mesh multigrid;
 * smoother_info (or NULL): per level lambda_min, lambda_max, theta, delta,
 * Chebyshev degree, eigenvalue-CG iterations. */
int mfs_mesh_laplace_vmult(mfs_handle *h, const int bc[6], int level, const double *src, double *dst);
int mfs_mesh_vcycle(mfs_handle *h, const int bc[6], const double *src, double *dst, double *smoother_info);

/* build_preconditioner(): local_smoothing.cc:1855-1867 (level diagonals) and the
 * Chebyshev eigenvalue estimates of solve() (:827-849). */
int mfs_build_preconditioner(mfs_handle *h);

/* StokesOperator::vmult (matrix_free_operators.cc:376-547, 606-627), FP64,
 * active level.  src/dst hold velocity then pressure. */
int mfs_stokes_vmult(mfs_handle *h, const double *src, double *dst);
/* ABlockOperator::vmult (:974-1064).  level >= 0: multigrid level operator in
 * the GMG precision with the level viscosity; level == -1: the active-level
 * operator (FP64, active viscosity table) used by the expensive solver. */
int mfs_ablock_vmult(mfs_handle *h, int level, const double *src, double *dst);
/* MassMatrixOperator::vmult (:787-857), same level convention. */
int mfs_mass_vmult(mfs_handle *h, int level, const double *src, double *dst);
/* BTBlockOperator::vmult (:668-725), active level, FP64: dst_u = B^T src_p. */
int mfs_bt_vmult(mfs_handle *h, const double *src_p, double *dst_u);
/* compute_diagonal of the A block / mass (:1068-1098, :861-908), level >= 0,
 * the diagonal itself (ones on constrained DoFs), not its inverse. */
int mfs_ablock_diagonal(mfs_handle *h, int level, double *diag);
int mfs_mass_diagonal(mfs_handle *h, int level, double *diag);
/* MGTransferMatrixFree prolongate / restrict_and_add (set up at
 * local_smoothing.cc:1843-1850); kind 0 velocity, 1 pressure. */
int mfs_prolongate(mfs_handle *h, int fine_level, int kind, const double *src_coarse, double *dst_fine);
int mfs_restrict(mfs_handle *h, int fine_level, int kind, const double *src_fine, double *dst_coarse);
/* Level viscosity table (cells lexicographic), local_smoothing.cc:211-294:
 * one value per cell, or cells x nq values with MFS_AVERAGE_PROJECT_Q1. */
int mfs_level_viscosity(mfs_handle *h, int level, double *out);
/* Chebyshev set-up per level: out[(level*2+kind)*5 + {0..4}] =
 * lambda_min, lambda_max, theta, delta, eigenvalue-CG iterations. */
int mfs_smoother_info(mfs_handle *h, double *out);
/* One PreconditionChebyshev application with zero initial guess. */
int mfs_chebyshev(mfs_handle *h, int kind, int level, const double *src, double *dst);
/* One PreconditionMG::vmult V-cycle (kind 0: A block, 1: Schur mass). */
int mfs_vcycle(mfs_handle *h, int kind, const double *src, double *dst);
/* BlockSchurPreconditioner::vmult, cheap variant
 * (block_stokes_preconditioner.h:220-255). */
int mfs_block_preconditioner(mfs_handle *h, const double *src, double *dst);

/* solve(): local_smoothing.cc:758-1446.  rhs is the Stokes right-hand side,
 * x0 the current linearization point (NULL: zero); sol receives velocity and
 * pressure (pressure rescaled and volume-normalized). */
int mfs_solve(mfs_handle *h, const double *rhs, const double *x0, double *sol, const mfs_solve_params *prm,
              mfs_solve_result *res);

/* Buoyancy right-hand side of the incompressible Stokes system,
 * rhs_u = int rho g . phi_u, rhs_p = 0, constrained rows zero
 * (source/simulator/assemblers/stokes.cc:393-396).  rho_qp: density at every
 * quadrature point of the active level ([cell][q]); gravity: dim components. */
int mfs_assemble_buoyancy_rhs(mfs_handle *h, const double *rho_qp, const double *gravity, double *rhs);

/* StokesMatrixFreeHandler::correct_stokes_rhs (stokes_matrix_free_local_smoothing.cc:617-754):
 * rhs += -A(u0) where u0 holds the boundary values on constrained velocity
 * DoFs (zero elsewhere), read as stored; constrained rows of rhs unchanged.
 * rhs and u0: N_u + N_p values. */
int mfs_correct_stokes_rhs(mfs_handle *h, const double *u0, double *rhs);

/* Device-resident variants: all pointers are device pointers. */
int mfs_stokes_vmult_device(mfs_handle *h, const double *d_src, double *d_dst);
/* Level A-block apply in the GMG precision (float or double device arrays). */
int mfs_ablock_vmult_device(mfs_handle *h, int level, const void *d_src, void *d_dst);
/* Allocation-free variant of mfs_ablock_vmult_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The degree-2 split apply's work array (in-tile sums of the tile-boundary
 * nodes) lives in a fixed device array of the library,
 * MFS_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends), instead of
 * the handle's buffer; the call allocates and frees nothing.  Levels whose
 * work array (3 x 386 values per 4^3-cell tile in 3D, 2 x 289 per 8^2-cell
 * tile in 2D) exceeds it return MFS_ERROR: at most 8149 tiles in FP64 and
 * 16299 in FP32 (80^3 and 100^3 cells) in 3D.  Degree-3 levels use no work
 * array.  Every call on one GPU
 * shares that array (from any handle): calls that overlap on one GPU race. */
#define MFS_NOALLOC_SCRATCH_BYTES (72LL << 20)
int mfs_ablock_vmult_device_noalloc(mfs_handle *h, int level, const void *d_src, void *d_dst);
int mfs_solve_device(mfs_handle *h, const double *d_rhs, const double *d_x0, double *d_sol,
                     const mfs_solve_params *prm, mfs_solve_result *res);
int mfs_assemble_buoyancy_rhs_device(mfs_handle *h, const double *d_rho_qp, const double *gravity, double *d_rhs);
int mfs_synchronize(mfs_handle *h);
/* V-cycles are captured once into a hipGraph and replayed (default on); 0
 * launches them kernel by kernel.  No counterpart in the reference (deal.II
 * runs the V-cycle on the host). */
int mfs_set_graph_capture(mfs_handle *h, int enable);

/* "Execute solver timings" harness (local_smoothing.cc:1119-1258): average
 * device time of reps applications of an operator, bracketed by hipEvents on
 * the handle's stream.  which: 0 Stokes (FP64, active level; d_src/d_dst FP64
 * block vectors), 1 level A block (GMG precision), 2 level mass (GMG
 * precision), 3 A-block V-cycle, 4 Schur V-cycle (FP64 vectors),
 * 5 cheap block preconditioner (FP64 block vectors). */
int mfs_time_operator(mfs_handle *h, int which, int level, const void *d_src, void *d_dst, int reps,
                      double *avg_seconds);

/* Memory consumption getters of StokesMatrixFreeHandler (stokes_matrix_free.h:81-97),
 * bytes of device memory: out[0] vectors/levels, out[1] cell data, out[2] Krylov space. */
int mfs_memory_consumption(const mfs_handle *h, long long out[3]);

/* Ensembles of independent problems (uneven sizes) solved together: one
 * handle and HIP stream per problem, host worker threads drive the solves so
 * that the kernels of many small problems overlap on the device.  No
 * counterpart in the reference (one ASPECT run solves one problem). */
typedef struct mfs_ensemble mfs_ensemble;
int mfs_ensemble_create(int n, const mfs_box *boxes, int gmg_precision, mfs_ensemble **out);
int mfs_ensemble_destroy(mfs_ensemble *e);
/* Handle of member i (owned by the ensemble): any per-handle call applies. */
int mfs_ensemble_member(mfs_ensemble *e, int i, mfs_handle **h);
/* mfs_set_viscosity + mfs_build_preconditioner on every member. */
int mfs_ensemble_setup(mfs_ensemble *e, const double *const *eta_qp, int averaging, int n_threads);
/* mfs_solve on every member (zero initial guess); results[i] per member;
 * *seconds = wall time of the whole ensemble solve.  Returns MFS_NO_CONVERGENCE
 * if any member did not converge. */
int mfs_ensemble_solve(mfs_ensemble *e, const double *const *rhs, double *const *sol, const mfs_solve_params *prm,
                       mfs_solve_result *results, int n_threads, double *seconds);

/* General level meshes (degree 2): the level operators of the global-coarsening
 * handler on a mesh of any topology with hanging nodes
 * (source/simulator/solver/stokes_matrix_free_global_coarsening.cc:1455-1625:
 * one MatrixFree object per level with the velocity and pressure
 * AffineConstraints of :1481-1598; operators of matrix_free_operators.cc:376-547,
 * :668-701, :808-845, :974-1009).  The caller describes the mesh the way a
 * DoFHandler and AffineConstraints do: node ids per cell, the geometry at every
 * quadrature point, and closed constraint lines. */
typedef struct mfs_tria mfs_tria;
typedef struct {
  int dim;                  /* 2 or 3 */
  int degree;               /* velocity degree: 2 */
  long long n_cells;
  long long n_nodes_u;      /* velocity nodes; velocity DoF = component * n_nodes_u + node */
  long long n_nodes_p;      /* pressure nodes = pressure DoFs */
  const int *cell_nodes_u;  /* n_cells x 3^dim node ids, lexicographic in the cell (x fastest) */
  const int *cell_nodes_p;  /* n_cells x 2^dim node ids, lexicographic */
  const double *geometry;   /* n_cells x 3^dim QGauss(3) points (lexicographic) x (dim*dim + 1):
                               J^{-1} row-major (entry e*dim + d = d xi_e / d x_d), then JxW */
  int compressible;         /* subtract (2/3) eta div(u) I (matrix_free_operators.cc:466-469) */
  double pressure_scaling;
} mfs_tria_desc;
/* Closed, homogeneous constraint lines of one block (AffineConstraints after
 * close(): no master is itself constrained): DoF dofs[i] = sum over
 * k in [offsets[i], offsets[i+1]) of weights[k] * DoF masters[k]; a line
 * without masters fixes the DoF to zero (Dirichlet). */
typedef struct {
  long long n_lines;
  const long long *dofs;
  const long long *offsets; /* n_lines + 1 */
  const long long *masters;
  const double *weights;
} mfs_constraints;
/* velocity / pressure may be NULL (no constraints). */
int mfs_tria_create(const mfs_tria_desc *desc, const mfs_constraints *velocity, const mfs_constraints *pressure,
                    mfs_tria **out);
int mfs_tria_destroy(mfs_tria *t);
/* out[0] cells, out[1] velocity DoFs, out[2] pressure DoFs. */
int mfs_tria_sizes(const mfs_tria *t, long long out[3]);
/* Viscosity at every quadrature point, n_cells x 3^dim (OperatorCellData with
 * per-point tables, matrix_free_operators.h:91-208). */
int mfs_tria_set_viscosity(mfs_tria *t, const double *eta_qp);
/* dst = op(src) with the MatrixFree constraint semantics: constrained source
 * entries are taken from their masters, contributions to constrained DoFs go
 * to their masters, constrained rows are identity rows (zero for B^T).
 * op: MFS_OP_STOKES ([u|p] -> [u|p]), MFS_OP_ABLOCK (u -> u), MFS_OP_BT
 * (p -> u), MFS_OP_MASS (s^2/eta pressure mass, p -> p).  Host vectors. */
int mfs_tria_vmult(mfs_tria *t, int op, const double *src, double *dst);
/* The same on device vectors (hipMalloc'd by the caller). */
int mfs_tria_vmult_device(mfs_tria *t, int op, const double *d_src, double *d_dst);
/* correct_stokes_rhs on a general level (local_smoothing.cc:617-754): rhs -= K u0
 * with u0's entries read as stored (read_dof_values_plain: u0 carries the
 * prescribed values of the constrained DoFs, e.g. after distribute() with the
 * inhomogeneous constraints) and the contributions to constrained DoFs
 * distributed to their masters or dropped (constrained rows receive nothing).
 * Host vectors [u|p]. */
int mfs_tria_correct_stokes_rhs(mfs_tria *t, const double *u0, double *rhs);
/* The viscosity table of the object, [cell][q] (host). */
int mfs_tria_get_viscosity(mfs_tria *t, double *eta_qp);
/* Average device time of reps applies on device vectors, bracketed by hipEvents
 * on the object's stream ("Execute solver timings", local_smoothing.cc:1119-1258). */
int mfs_tria_time(mfs_tria *t, int op, const double *d_src, double *d_dst, int reps, double *avg_seconds);

/* Global-coarsening multigrid over general levels (coarse first) and the
 * cheap Stokes solve on the finest one (stokes_matrix_free_global_coarsening.cc:
 * smoothers :800-846, coarse solver :877-883, transfers :1688-1728, cheap GMRES
 * :1259-1313).  Prolongation l (l = 0 .. n_levels-2) maps level l to level
 * l+1 as a CSR matrix over the DoFs of the two levels: the fine interpolant of
 * the coarse function, read with the coarse constraints, zero on constrained
 * fine DoFs (MGTwoLevelTransfer); restriction is its transpose.  The levels
 * stay owned by the caller and must outlive the multigrid object. */
typedef struct mfs_tria_gmg mfs_tria_gmg;
typedef struct {
  long long rows, cols;
  const long long *row_ptr; /* rows + 1 */
  const long long *col_idx;
  const double *values;
} mfs_csr;
/* parents (may be NULL): n_levels-1 arrays, parents[l][c] = the cell of level l
 * that contains cell c of level l+1 (its parent, or the same cell when it was
 * not refined); needed by mfs_tria_gmg_set_viscosity. */
int mfs_tria_gmg_create(int n_levels, mfs_tria *const *levels, const mfs_csr *prolong_u, const mfs_csr *prolong_p,
                        const long long *const *parents, mfs_tria_gmg **out);
int mfs_tria_gmg_destroy(mfs_tria_gmg *g);
/* Viscosity of every level from the finest level's values at its quadrature
 * points ([cell][q], host), as the global-coarsening handler's
 * evaluate_material_model does (stokes_matrix_free_global_coarsening.cc:173-330):
 * the finest level takes them with the material averaging (MFS_AVERAGE_NONE,
 * ARITHMETIC, HARMONIC or GEOMETRIC), every coarser level one value per cell,
 * the mean of its cells on the next finer level of their DGQ0 values (JxW-weighted
 * cell means).  Replaces mfs_tria_set_viscosity on every level. */
int mfs_tria_gmg_set_viscosity(mfs_tria_gmg *g, const double *eta_qp, int averaging);
/* Diagonals of every level operator (exact, constraints included: probing by
 * DoF colours; 1 on constrained DoFs) and the Chebyshev set-up: levels > 0
 * degree 4 over [lambda_max / 15, lambda_max] from 10 eigenvalue iterations,
 * level 0 degree 8, range 1e-3, 100 iterations.  The viscosity must be set on
 * every level (mfs_tria_set_viscosity). */
int mfs_tria_gmg_build(mfs_tria_gmg *g);
/* V-cycles captured once into hipGraphs and replayed (default on); 0 launches
 * them kernel by kernel.  No counterpart in the reference. */
int mfs_tria_gmg_set_graph_capture(mfs_tria_gmg *g, int enable);
/* out = {lambda_min, lambda_max, eigenvalue iterations, colours} of level /
 * kind (0 A block, 1 Schur complement); diagonal (may be NULL): the level
 * operator's diagonal (host, velocity or pressure DoFs of the level). */
int mfs_tria_gmg_info(mfs_tria_gmg *g, int level, int kind, double out[4], double *diagonal);
/* One V-cycle (kind 0: A block, 1 smoothing step; 1: Schur complement, 4 steps)
 * on host vectors of the finest level. */
int mfs_tria_gmg_vcycle(mfs_tria_gmg *g, int kind, const double *src, double *dst);
/* Solve on the finest level (local_smoothing.cc:944-1445 sequence).  Initial
 * guess x0 (host, may be NULL: zero; ignored for a Newton update): its
 * constrained velocity entries are zeroed and its pressure divided by the
 * pressure scaling; tolerance prm->tolerance * r0 with r0 the norm of
 * (rhs_u - [K (0, p0)]_u, rhs_p) (|rhs| without x0 or for a Newton update);
 * the solution's pressure is multiplied by the pressure scaling (no pressure
 * normalization: the caller's).  Right-preconditioned GMRES
 * (restart prm->restart) with the cheap block preconditioner (one A and one
 * Schur V-cycle) for at most prm->n_cheap_steps iterations, then, if not
 * converged, flexible GMRES with the expensive one (CG inner solves to
 * prm->a_block_tolerance / prm->s_block_tolerance, V-cycle preconditioned)
 * for at most prm->n_expensive_steps.  Host vectors [u|p]; the IDR(s)
 * fields of prm are not used.  Not converged: MFS_NO_CONVERGENCE with a
 * message (MFS_FAILURE_ABORT) or MFS_OK with res->converged = 0
 * (MFS_FAILURE_CONTINUE). */
int mfs_tria_solve(mfs_tria_gmg *g, const double *rhs, const double *x0, double *sol, const mfs_solve_params *prm,
                   mfs_solve_result *res);

/* Measurement of the matrix-core alternative for the sum-factorization
 * contractions (no counterpart in the reference): FP32 3-point basis
 * contractions y = S x of n_lines lines (a positive multiple of 80), each
 * repeated reps times, variant 0 as VALU FMAs (one line per thread), variant 1
 * on the matrix cores (v_mfma_f32_16x16x4f32, block-diagonal S packed into
 * 16 x 16 tiles: 4 products of 80 lines per repetition and wave).  *seconds:
 * device time of one launch; out (may be NULL, 3 n_lines floats as doubles):
 * the result of one repetition (variant 0: S x + x, variant 1: 4 S x). */
int mfs_probe_contraction(int variant, long long n_lines, int reps, double *seconds, double *out);

#ifdef __cplusplus
}
#endif
#endif /* ASPECT_MFS_H */
