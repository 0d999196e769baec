/*
 * platoanalyze_amd.h — C ABI of the MI355X-native Tet4 assembly library.
 *
 * Drop-in boundary for Plato Analyze's workset residual / Jacobian evaluation
 * and assembly on linear tetrahedra (SIMP elastostatics, steady thermal and
 * J2 plasticity).
 * Every entry point returns an int status (PA_OK = 0) and never throws; the
 * message of the last failure on the calling thread is pa_last_error().
 * Array arguments named *_dev are device pointers (hipMalloc'd, e.g. through
 * pa_malloc); *_host are host pointers.  Work is enqueued on the given
 * hipStream_t (passed as void*, 0 = default stream); the call returns before
 * the work finishes unless it says otherwise.
 *
 * Layouts are the reference's (src/PlatoStaticsTypes.hpp, src/alg/CrsMatrix.hpp):
 *   state  u[node*dpn + d]      control z[node]      coords x[node*3 + d]
 *   connectivity conn[cell*4 + a] (int32)
 *   block CRS: rowmap[N+1], colidx[nnzb] (sorted, diagonal included), values
 *   [nnzb * bi * bj] with the entry of block e at e*bi*bj + i*bj + j.
 */
#ifndef PLATOANALYZE_AMD_H
#define PLATOANALYZE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PA_OK = 0, PA_ERR_ARG = 1, PA_ERR_HIP = 2, PA_ERR_LIMIT = 3, PA_ERR_STATE = 4 };

/* Arithmetic of the assembled values. */
enum { PA_FP64 = 0, PA_FP32 = 1 };

/* Physics of a vector function (the reference's PhysicsT template argument). */
enum { PA_MECHANICS = 0, PA_THERMAL = 1 };

/* Scalar function types ("Scalar Function Type" in the reference input). */
enum { PA_INTERNAL_ELASTIC_ENERGY = 0, PA_INTERNAL_THERMAL_ENERGY = 1 };

typedef struct pa_mesh pa_mesh;
typedef struct pa_vector_function pa_vector_function;
typedef struct pa_scalar_function pa_scalar_function;

/* Message of the last failed call on this thread ("" if none). */
const char* pa_last_error(void);
/* Library version string. */
const char* pa_version(void);

/* ---------------------------------------------------------------- device */
/* Thin wrappers so callers without their own HIP runtime can move data.    */
int pa_device_count(int* count);
int pa_malloc(void** ptr_dev, size_t bytes);
int pa_free(void* ptr_dev);
int pa_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes);
int pa_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes);
int pa_memcpy_d2d(void* dst_dev, const void* src_dev, size_t bytes, void* stream);
int pa_memset(void* ptr_dev, int value, size_t bytes, void* stream);
/* Device-to-device copy by a plain streaming kernel (16-byte vectors); bytes
 * must be a multiple of 16.  Used to calibrate HBM byte counters. */
int pa_copy_kernel(void* dst_dev, const void* src_dev, size_t bytes, void* stream);
int pa_stream_create(void** stream);
int pa_stream_destroy(void* stream);
int pa_stream_synchronize(void* stream);
int pa_event_create(void** event);
int pa_event_destroy(void* event);
int pa_event_record(void* event, void* stream);
int pa_event_synchronize(void* event);
int pa_event_elapsed_ms(float* ms, void* start, void* stop);

/* ---------------------------------------------------------------- mesh */
/*
 * Synthetic structured box: nx*ny*nz hexes on [0,lx]x[0,ly]x[0,lz], each split
 * into 6 tets around the v000->v111 diagonal; node = ix*(ny+1)*(nz+1) +
 * iy*(nz+1) + iz, hex = (ix*ny + iy)*nz + iz, 6 consecutive tets per hex.
 * Replaces the BamG test mesh of unit_tests/PlatoTestHelpers.hpp:31-84.
 * coords_host: (nx+1)(ny+1)(nz+1)*3 doubles; conn_host: 24*nx*ny*nz ints.
 */
int pa_box_mesh(int nx, int ny, int nz, double lx, double ly, double lz, double* coords_host, int* conn_host);

/*
 * Mesh + assembly graph.  Replaces Omega_h's mesh / NodeNodeGraph
 * (src/OmegaHMesh.cpp:131-149) and the graph half of CreateBlockMatrix
 * (src/ImplicitFunctors.hpp:932-1005), which the reference rebuilds on every
 * gradient call; here it is built once.  Copies coords/conn to the device.
 */
int pa_mesh_create(const double* coords_host, int num_nodes, const int* conn_host, int num_cells, pa_mesh** mesh);
int pa_mesh_destroy(pa_mesh* mesh);
/* Sizes: nodes, cells, CRS blocks. */
int pa_mesh_sizes(const pa_mesh* mesh, int* num_nodes, int* num_cells, int64_t* num_blocks);
/*
 * Assembly schedule of the mesh (diagnostic; no reference counterpart):
 * number of cell tiles and distinct tile schedules (0 when the node
 * numbering is not a lexicographic grid, or PA_TILES=0 was set at creation),
 * and of distinct row patterns (0 when the generic row kernels are used).
 */
int pa_mesh_assembly_info(const pa_mesh* mesh, int* num_tiles, int* num_tile_patterns, int* num_row_patterns);
/*
 * Node grid of the streaming grid kernels (diagnostic; no reference
 * counterpart): dims = (x, y, z) node counts when the nodes are a
 * lexicographic grid (z fastest) and the cells are the six tets of every hex
 * around its v0 -> v111 diagonal, else zeros (also when PA_GRID=0 was set at
 * creation).  On such meshes value() and the thermal gradient_u run the
 * streaming kernels.
 */
int pa_mesh_grid_info(const pa_mesh* mesh, int* dims);
/* Copy the block-CRS graph to the host (rowmap: N+1, colidx: nnzb). */
int pa_mesh_graph(const pa_mesh* mesh, int* rowmap_host, int* colidx_host);
/* Device pointers of the graph (owned by the mesh). */
int pa_mesh_graph_dev(const pa_mesh* mesh, const int** rowmap_dev, const int** colidx_dev);
/*
 * Register a side set (Omega_h side set: for each face, the owning cell and
 * the three local node ordinals of the face in that cell;
 * src/SurfaceLoadIntegral.hpp:106-107).
 */
int pa_mesh_add_side_set(pa_mesh* mesh, const char* name, int num_faces, const int* face_cells_host,
                         const int* face_local_nodes_host);
/* Register a node set (src/AbstractPlatoMesh.hpp node sets). */
int pa_mesh_add_node_set(pa_mesh* mesh, const char* name, int num_nodes, const int* nodes_host);
/* Number of nodes and a host copy of a node set. */
int pa_mesh_node_set(const pa_mesh* mesh, const char* name, int* num_nodes, int* nodes_host);

/* ------------------------------------------------------- vector function */
/*
 * Elliptic::VectorFunction<Mechanics<3>|Thermal<3>>
 * (src/elliptic/VectorFunction.hpp:115-234) with the evaluator of one domain:
 * ElastostaticResidual (src/elliptic/ElastostaticResidual.hpp:165-262) or
 * ThermostaticResidual (src/elliptic/ThermostaticResidual.hpp:128-223),
 * isotropic material, MSIMP penalty (src/Simp.hpp:41-52,130-134).
 * material[0..1] = {Youngs Modulus, Poissons Ratio} for PA_MECHANICS,
 * material[0] = {Thermal Conductivity} for PA_THERMAL;
 * penalty = {Exponent, Minimum Value}.
 */
int pa_vector_function_create(pa_mesh* mesh, int physics, const double* material, const double* penalty,
                              pa_vector_function** vf);
int pa_vector_function_destroy(pa_vector_function* vf);
/* Degrees of freedom per node (3 or 1). */
int pa_vector_function_dofs_per_node(const pa_vector_function* vf, int* dpn);
/*
 * Uniform natural BC ("Natural Boundary Conditions/<n>/Type=Uniform") on a side
 * set: src/NaturalBCs.hpp:150-224, src/SurfaceLoadIntegral.hpp:97-149.
 * values: dpn doubles (traction or flux).
 */
int pa_vector_function_add_uniform_natural_bc(pa_vector_function* vf, const char* side_set, const double* values);
/*
 * Constant body load ("Body Loads/<n>/{Index, Function}"), scaled by the
 * unpenalised mean nodal density: src/BodyLoads.hpp:54-122.
 */
int pa_vector_function_add_body_load(pa_vector_function* vf, int dof, double value);

/* A body load given by a 'Function' string of x, y and z on any dof of any
 * physics: BodyLoad::get (src/BodyLoads.hpp:54-122) with the function
 * evaluated at each cell's cubature point, the centroid (src/PlatoMeshExpr.hpp:
 * 16-50); R[cell node, dof] += -(1/6) |det J| f (1/4) mean cell density.  The
 * function is evaluated once, here, on the host.  gradient_z, gradient_x and the
 * problem's criterion gradient include the term; like the reference,
 * gradient_x takes f as constant per cell (its point is mapped in plain
 * doubles).  Syntax: numbers, x y z t, pi, + - * / ^, parentheses, sin cos tan
 * asin acos atan sinh cosh tanh exp log log10 sqrt abs fabs pow atan2 min max,
 * leading 'name = expr;' assignments. */
int pa_vector_function_add_body_load_function(pa_vector_function* vf, int dof, const char* function);

/* Evaluates a 'Function' string at n points (host arrays; a null coordinate
 * array reads as zeros) and time t.  No GPU work. */
int pa_expression_eval(const char* function, int n, const double* x, const double* y, const double* z, double t,
                       double* out);

/* VectorFunction::value (src/elliptic/VectorFunction.hpp:250-334): R[N*dpn]. */
int pa_vector_function_value(const pa_vector_function* vf, const double* u_dev, const double* z_dev, double* r_dev,
                             void* stream);
/*
 * VectorFunction::gradient_u (src/elliptic/VectorFunction.hpp:524-615):
 * Jacobian values [nnzb*dpn*dpn] on the mesh graph, in FP64 (double*) or FP32
 * (float*) per `precision`.
 */
int pa_vector_function_gradient_u(const pa_vector_function* vf, const double* u_dev, const double* z_dev,
                                  void* values_dev, int precision, void* stream);
/* VectorFunction::gradient_u_T (src/elliptic/VectorFunction.hpp:431-523). */
int pa_vector_function_gradient_u_T(const pa_vector_function* vf, const double* u_dev, const double* z_dev,
                                    void* values_dev, int precision, void* stream);
/*
 * value and gradient_u in one call, as a Newton step needs them (no
 * reference counterpart; src/elliptic/Problem.hpp:250-262 calls the two):
 * the two separate calls' results, up to rounding.  On elastic Kuhn grids
 * one kernel pass forms the blocks and R = K u + f from them; elsewhere the
 * residual and Jacobian kernels run back to back, or one fused pass where no
 * grid or cell tiles exist.  r_dev: N*dpn doubles, always FP64.
 */
int pa_vector_function_value_gradient_u(const pa_vector_function* vf, const double* u_dev, const double* z_dev,
                                        double* r_dev, void* values_dev, int precision, void* stream);
/*
 * VectorFunction::gradient_z (src/elliptic/VectorFunction.hpp:617-710): the
 * transposed dR/dz as a block matrix on the node graph with 1 x dpn blocks
 * (row = control node, col = state node): values [nnzb*dpn] doubles.
 */
int pa_vector_function_gradient_z(const pa_vector_function* vf, const double* u_dev, const double* z_dev,
                                  double* values_dev, void* stream);
/*
 * VectorFunction::gradient_x (src/elliptic/VectorFunction.hpp:337-420): the
 * transposed dR/dx (node coordinates) as a block matrix on the node graph with
 * 3 x dpn blocks (row = node whose position moves, col = state node); entry
 * [i*dpn + j] of block (a, b) is dR_{b,j}/dx_{a,i}.  Includes the natural
 * BCs' face-area terms.  values [nnzb*3*dpn] doubles.
 */
int pa_vector_function_gradient_x(const pa_vector_function* vf, const double* u_dev, const double* z_dev,
                                  double* values_dev, void* stream);

/* ------------------------------------------------------- scalar function */
/*
 * Elliptic::PhysicsScalarFunction with InternalElasticEnergy
 * (src/elliptic/InternalElasticEnergy.hpp:132-199) or InternalThermalEnergy
 * (src/elliptic/InternalThermalEnergy.hpp:96-150).
 */
int pa_scalar_function_create(pa_mesh* mesh, int type, const double* material, const double* penalty,
                              pa_scalar_function** sf);
int pa_scalar_function_destroy(pa_scalar_function* sf);
/* PhysicsScalarFunction::value (src/elliptic/PhysicsScalarFunction.hpp:225-284): one double at value_dev. */
int pa_scalar_function_value(const pa_scalar_function* sf, const double* u_dev, const double* z_dev,
                             double* value_dev, void* stream);
/*
 * pa_scalar_function_value without allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory in the call
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Same kernels and result as pa_scalar_function_value; its
 * per-workgroup partial sums (at most 1024 doubles, enough for any mesh: the
 * cell kernel strides over the cells) live in a static device array each GPU
 * holds until the process ends, shared by every call on that GPU, so calls
 * that overlap in time on one GPU race.  The first call on a device looks up
 * the array's address (no allocation).
 */
int pa_scalar_function_value_noalloc(const pa_scalar_function* sf, const double* u_dev, const double* z_dev,
                                     double* value_dev, void* stream);
/* PhysicsScalarFunction::gradient_u: N*dpn doubles. */
int pa_scalar_function_gradient_u(const pa_scalar_function* sf, const double* u_dev, const double* z_dev,
                                  double* grad_dev, void* stream);
/* PhysicsScalarFunction::gradient_z (src/elliptic/PhysicsScalarFunction.hpp:431-494): N doubles. */
int pa_scalar_function_gradient_z(const pa_scalar_function* sf, const double* u_dev, const double* z_dev,
                                  double* grad_dev, void* stream);
/* PhysicsScalarFunction::gradient_x (src/elliptic/PhysicsScalarFunction.hpp:293-355): N*3 doubles. */
int pa_scalar_function_gradient_x(const pa_scalar_function* sf, const double* u_dev, const double* z_dev,
                                  double* grad_dev, void* stream);

/* ---------------------------------------------------------- constraints */
/*
 * applyBlockConstraints<dpn> (src/ApplyConstraints.hpp:40-112): symmetric
 * Dirichlet elimination on a block-CRS matrix of the mesh graph (FP64 or FP32
 * values per `precision`) and its right-hand side (FP64).
 */
int pa_apply_block_constraints(const pa_mesh* mesh, int dpn, void* values_dev, int precision, double* rhs_dev,
                               int num_bcs, const int* bc_dofs_dev, const double* bc_values_dev, double scale,
                               void* stream);

/*
 * Block matrix-vector product y = A x + beta*y on the mesh graph
 * (src/PlatoMathHelpers.hpp:79-159, MatrixTimesVectorPlusVector).
 * bi x bj blocks; values FP64 or FP32 per `precision`; x, y FP64.
 */
int pa_block_spmv(const pa_mesh* mesh, int bi, int bj, const void* values_dev, int precision, const double* x_dev,
                  double beta, double* y_dev, void* stream);

/* -------------------------------------------------------------- solver */
/*
 * Preconditioned conjugate gradients on a block-CRS matrix of the mesh graph
 * (square dpn x dpn blocks).  Replaces the reference's linear-solver stacks
 * for this path (AztecOO CG, src/alg/EpetraLinearSolver.cpp:192-288; AmgX
 * PCG, src/alg/AmgXConfigs.cpp:110-134).  Vectors and all reductions are
 * FP64; the matrix is FP64 or FP32 per `precision`.
 */
enum { PA_PRECOND_NONE = 0, PA_PRECOND_JACOBI = 1, PA_PRECOND_BLOCK_JACOBI = 2, PA_PRECOND_MULTIGRID = 3 };

/*
 * Geometric multigrid preconditioner for meshes built by pa_box_mesh (nx x ny
 * x nz hexes, 6 tets each): levels halve the grid while the cell counts are
 * even; trilinear transfer; Galerkin coarse operators P^T A P stored as
 * 27-point block stencils; damped block-Jacobi smoothing; dense inverse on the
 * coarsest level.  FP32 throughout (a preconditioner).  No counterpart in the
 * reference (its AmgX path, src/alg/AmgXConfigs.cpp, is algebraic multigrid).
 */
/* Smoothers: damped block Jacobi with `omega`, or the Chebyshev polynomial of
   degree `pre_sweeps` / `post_sweeps` in D^-1 A (for symmetric positive definite
   operators) on [b/12, b], b = 1.2 x the largest eigenvalue of D^-1 A that power
   iteration finds on each level at setup; the post-smoothing steps run in reverse
   order, so Chebyshev needs pre_sweeps == post_sweeps.  DEFAULT: Chebyshev for
   the elliptic problems' multigrid when the sweep counts agree, Jacobi
   otherwise. */
enum { PA_SMOOTHER_DEFAULT = 0, PA_SMOOTHER_JACOBI = 1, PA_SMOOTHER_CHEBYSHEV = 2 };
typedef struct {
    int max_levels;       /* 0: coarsen while possible */
    int pre_sweeps;       /* block-Jacobi sweeps before the coarse correction */
    int post_sweeps;      /* ... and after */
    double omega;         /* Jacobi damping (0: 0.6) */
    int max_coarse_nodes; /* limit of the dense coarsest level (0: 2000) */
    int symmetric;        /* 1: the matrices given to setup are symmetric, block (m, n) the transpose of
                             (n, m); the fine level then keeps half of its block stencil (the elliptic
                             problems set this themselves) */
    int smoother;         /* PA_SMOOTHER_* */
    int coarse_sweeps;    /* pre- and post-smoothing sweeps on the coarse levels of at most coarse_max_nodes
                             nodes, whose sweeps cost little more than a launch (0: as pre/post_sweeps) */
    int coarse_max_nodes; /* 0: 16384 */
} pa_multigrid_options;

typedef struct pa_multigrid pa_multigrid;
int pa_multigrid_create(const pa_mesh* mesh, int nx, int ny, int nz, int dpn, const pa_multigrid_options* options,
                        pa_multigrid** mg);
int pa_multigrid_destroy(pa_multigrid* mg);
int pa_multigrid_levels(const pa_multigrid* mg, int* levels, int* coarse_dofs);
/* Build the hierarchy from the (constrained) fine matrix; constrained dofs are masked out of the transfer. */
int pa_multigrid_setup(pa_multigrid* mg, const void* values_dev, int precision, int num_bcs, const int* bc_dofs_dev,
                       void* stream);
/* z = M^-1 r by one V-cycle. */
int pa_multigrid_apply(pa_multigrid* mg, const double* r_dev, double* z_dev, void* stream);
/*
 * Copy the Galerkin operator of `level` (>= 1) to the host as 27-point block
 * stencils: nodes x 27 x dpn*dpn floats, offset index (dx+1)*9 + (dy+1)*3 +
 * (dz+1).  Synchronizes.  (Diagnostic: lets tests check P^T A P.)
 */
int pa_multigrid_level_operator(const pa_multigrid* mg, int level, float* stencil_host, int* nx, int* ny, int* nz);

typedef struct pa_solver pa_solver;
typedef struct {
    int iterations;      /* CG iterations performed (summed over refinement sweeps) */
    int sweeps;          /* refinement sweeps (1 for a plain solve) */
    double rel_residual; /* final ||b - A x|| / ||b|| (FP64 residual) */
    int converged;       /* 1 if rel_residual <= tolerance */
} pa_solve_stats;

int pa_solver_create(const pa_mesh* mesh, int dpn, int precond, pa_solver** solver);
int pa_solver_destroy(pa_solver* solver);
/* Use `mg` (set up by the caller) as the preconditioner; the solver does not own it. */
int pa_solver_set_multigrid(pa_solver* solver, pa_multigrid* mg);
/*
 * Solve A x = b starting from the x given.  Stops when the recursive residual
 * satisfies ||r|| <= tol ||b|| (checked every `check_every` iterations; the
 * host reads one scalar then) or after max_iter iterations; the returned
 * rel_residual is recomputed in FP64 from A.  Synchronizes `stream`.
 */
int pa_solver_pcg(pa_solver* solver, const void* values_dev, int precision, const double* b_dev, double* x_dev,
                  double tol, int max_iter, int check_every, pa_solve_stats* stats, void* stream);

/* -------------------------------------------------------------- problem */
/*
 * Elliptic::Problem (src/elliptic/Problem.hpp:237-450): solution (one Newton
 * step: R = value(0, z), K = gradient_u, applyBlockConstraints, solve K du =
 * -R), criterion value and the self-adjoint criterion gradient
 * dF/dz = dF/dz|_u + (dR/dz)^T lambda with lambda = -u.
 *
 * mode PA_FP64: FP64 assembly and FP64-matrix PCG to `tol`.
 * mode PA_FP32 (mixed): FP32 assembly and PCG on the FP32 matrix with
 * reliable updates: every `replace_every` iterations and at apparent
 * convergence the recursive residual is replaced by the exact FP64 residual
 * b - K(z) x, computed matrix-free by the assembly kernels; the solve ends
 * when that FP64 residual satisfies ||r|| <= tol ||b||.
 */
typedef struct {
    int mode;          /* PA_FP64 or PA_FP32 */
    int precond;       /* PA_PRECOND_* */
    double tol;        /* relative residual target (FP64 residual) */
    int max_iter;      /* CG iteration cap */
    int replace_every; /* mixed mode: iterations between exact-residual replacements (0: only at convergence) */
    int check_every;   /* convergence check period (iterations) */
} pa_problem_options;

typedef struct pa_problem pa_problem;
int pa_problem_create(pa_vector_function* vf, pa_scalar_function* criterion, int num_bcs, const int* bc_dofs_host,
                      const double* bc_values_host, const pa_problem_options* options, pa_problem** problem);
int pa_problem_destroy(pa_problem* problem);
int pa_problem_set_options(pa_problem* problem, const pa_problem_options* options);
/*
 * Precondition the solves with geometric multigrid on the nx x ny x nz box
 * (rebuilt from each new Jacobian inside pa_problem_solution); selects
 * PA_PRECOND_MULTIGRID.
 */
int pa_problem_set_multigrid(pa_problem* problem, int nx, int ny, int nz, const pa_multigrid_options* options);
/* State u (N*dpn) for control z; statistics of the linear solve. Synchronizes. */
int pa_problem_solution(pa_problem* problem, const double* z_dev, double* u_dev, pa_solve_stats* stats, void* stream);
/* Criterion value at (u, z), one double at value_dev. */
int pa_problem_criterion_value(pa_problem* problem, const double* z_dev, const double* u_dev, double* value_dev,
                               void* stream);
/*
 * Total criterion gradient dF/dz (N doubles) with the self-adjoint adjoint
 * lambda = -u ("Self-Adjoint" = true, src/elliptic/Problem.hpp:423-427), in one
 * fused node-ordered kernel: per incident cell, dF_e/dz_a + lambda_e.dR_e/dz_a,
 * without assembling the N x N*dpn dR/dz matrix.
 */
int pa_problem_criterion_gradient(pa_problem* problem, const double* z_dev, const double* u_dev, double* grad_dev,
                                  void* stream);
/*
 * Criterion gradient wrt the node coordinates (N*3 doubles) of a self-adjoint
 * problem: the reference's criterionGradientX returns -1 times the criterion's
 * partial derivative there (src/elliptic/Problem.hpp:492-515).
 */
int pa_problem_criterion_gradient_x(pa_problem* problem, const double* z_dev, const double* u_dev, double* grad_dev,
                                    void* stream);

/* -------------------------------------------------------- J2 plasticity */
/*
 * Infinitesimal-strain J2 plasticity in the reference's stabilized mixed u-p
 * form (Physics "Plasticity", src/SimplexPlasticity.hpp:25-31): 4 global dofs
 * per node (ux, uy, uz, p) in u[node*4 + d]; 14 local states per cell in
 * c[cell*14 + k] (accumulated plastic strain, plastic multiplier increment,
 * plastic strain xx yy zz yz xz xy, backstress xx yy zz yz xz xy;
 * src/J2PlasticityLocalResidual.hpp:99-119); the nodal projected pressure
 * gradient pi[node*3 + d].  Cell worksets are [cell][row][column] with the
 * local global dof node*4 + d, as the reference's ScalarArray3D worksets.
 */
typedef struct {
    double youngs, poisson;                                  /* Isotropic Linear Elastic */
    double hardening_isotropic, hardening_kinematic;         /* J2 Plasticity */
    double initial_yield_stress;
    double elastic_penalty_exponent, elastic_min_ersatz;     /* defaults 1, 1e-9 */
    double plastic_penalty_exponent, plastic_min_ersatz;     /* defaults 0.5, 1e-4 */
    double global_penalty_exponent, global_min_ersatz;       /* Elliptic / Penalty Function: 1, 1e-9 */
    double pressure_scaling;                                 /* Material Models / Pressure Scaling: 1 */
    int projection_penalty;                                  /* State Gradient Projection penalty: 0 none, 1 SIMP */
    double projection_penalty_exponent, projection_min_ersatz;
    double work_penalty_exponent, work_min_ersatz;           /* plastic work criterion: 3, 1e-9 */
    int formulation;                                         /* PA_J2_MIXED (default) or PA_J2_DISPLACEMENT */
} pa_j2_params;
/*
 * PA_J2_MIXED: the reference's stabilized mixed u-p form above (4 dofs per node).
 * PA_J2_DISPLACEMENT: pure-displacement J2 with no reference counterpart,
 * sigma = Dev(mu~)(eps - eps_p) + kappa~ tr(eps - eps_p) I with the global
 * penalty on both moduli, 3 dofs per node (u[node*3 + d], R, lambda, Dirichlet
 * dofs and 3x3 blocks of the condensed operator), no pressure, stabilization or
 * projection (pi is read and ignored); the plastic work uses kappa~_w tr(eps_e)
 * in place of the pressure.  Cell worksets keep the 16 rows/columns node*4 + d,
 * with the pressure slots zero.
 */
enum { PA_J2_MIXED = 0, PA_J2_DISPLACEMENT = 1 };
/* The reference's defaults (src/J2PlasticityLocalResidual.hpp:195-220,
 * src/InfinitesimalStrainPlasticityResidual.hpp:142-160,214,
 * src/PlasticWorkCriterion.hpp:295-303); moduli and yield stress zero. */
void pa_j2_params_default(pa_j2_params* params);

/* Which partial derivative a workset Jacobian holds. */
enum { PA_WRT_U = 0, PA_WRT_C = 1, PA_WRT_Z = 2, PA_WRT_C_PREV = 3, PA_WRT_PI = 4 };

typedef struct pa_plasticity pa_plasticity;
int pa_plasticity_create(const pa_mesh* mesh, const pa_j2_params* params, pa_plasticity** plasticity);
int pa_plasticity_destroy(pa_plasticity* plasticity);

/* LocalVectorFunctionInc::updateLocalState, closed-form radial return
 * (src/LocalVectorFunctionInc.hpp:236-300, src/J2PlasticityLocalResidual.hpp:440-543). */
int pa_j2_update_local_state(pa_plasticity* p, const double* u_dev, const double* c_prev_dev, const double* z_dev,
                             double* c_dev, void* stream);
/* LocalVectorFunctionInc::valueWorkSet: H[cell*14 + k]
 * (src/LocalVectorFunctionInc.hpp:302-370, src/J2PlasticityLocalResidual.hpp:310-428). */
int pa_j2_local_value(pa_plasticity* p, const double* u_dev, const double* c_dev, const double* c_prev_dev,
                      const double* z_dev, double* H_dev, void* stream);
/* LocalVectorFunctionInc::gradient_{u,c,z,cp} worksets [cell][14][W], W = 16 (U), 14 (C, C_PREV), 4 (Z)
 * (src/LocalVectorFunctionInc.hpp:372-874), hand-derived in place of the FAD evaluation types. */
int pa_j2_local_gradient(pa_plasticity* p, int wrt, const double* u_dev, const double* c_dev,
                         const double* c_prev_dev, const double* z_dev, double* out_dev, void* stream);
/* GlobalVectorFunctionInc::value: assembled R[node*dpn + d] (dpn 4 mixed, 3 displacement)
 * (src/GlobalVectorFunctionInc.hpp:919-960, src/InfinitesimalStrainPlasticityResidual.hpp:393-521). */
int pa_j2_global_value(pa_plasticity* p, const double* u_dev, const double* c_dev, const double* pi_dev,
                       const double* z_dev, double* R_dev, void* stream);
/* GlobalVectorFunctionInc::gradient_{u,c,z,n} worksets [cell][16][W], W = 16 (U), 14 (C), 4 (Z), 12 (PI)
 * (src/GlobalVectorFunctionInc.hpp:962-1155). */
int pa_j2_global_gradient(pa_plasticity* p, int wrt, const double* u_dev, const double* c_dev, const double* pi_dev,
                          const double* z_dev, double* out_dev, void* stream);
/*
 * One Newton linearization of the condensed system
 * (NewtonRaphsonSolver::assembleTangentMatrix / assembleResidual,
 * src/NewtonRaphsonSolver.hpp:204-330): block-CRS dpn x dpn values on the mesh graph
 *   K_t = dR/du - dR/dc (dH/dc)^-1 dH/du,   R_t = R - dR/dc (dH/dc)^-1 H.
 * transpose != 0 assembles K_t^T (the adjoint's operator,
 * src/PathDependentAdjointSolver.hpp:232-275); R_dev may then be null.
 */
int pa_j2_condensed_system(pa_plasticity* p, const double* u_dev, const double* c_dev, const double* c_prev_dev,
                           const double* pi_dev, const double* z_dev, double* values_dev, double* R_dev,
                           int transpose, void* stream);
/* Projected pressure gradient from the pressure in u (PlasticityProblem::updateProjectedPressureGradient,
 * src/PlasticityProblem.hpp:872-891; lumped solve src/Plato_Solve.hpp:22-45). */
int pa_j2_projection(pa_plasticity* p, const double* u_dev, const double* z_dev, double* pi_dev, void* stream);
/* PlasticWorkCriterion::evaluate of one step summed over cells, one double at value_dev
 * (src/PlasticWorkCriterion.hpp:163-265). */
int pa_j2_plastic_work(pa_plasticity* p, const double* u_dev, const double* u_prev_dev, const double* c_dev,
                       const double* c_prev_dev, const double* z_dev, double* value_dev, void* stream);

/*
 * PlasticityProblem (src/PlasticityProblem.hpp:784-868, 1035-1068): pseudo-time
 * steps with load factor (k+1)/K on the Dirichlet values, the condensed Newton
 * solve per step (elastic trial state, Dirichlet increment on the first
 * iteration, absolute-residual stopping test, src/NewtonRaphsonSolver.hpp:555-610)
 * with flexible GMRES (or BiCGStab) linear solves preconditioned by dpn x dpn block
 * Jacobi or geometric multigrid, the projection update (mixed form) after
 * each step, every step's states kept on the device (the checkpoints of the
 * adjoint), the plastic-work criterion summed over steps and its total
 * gradient by the path-dependent adjoint (src/PathDependentAdjointSolver.hpp).
 */
enum { PA_KRYLOV_GMRES = 0, PA_KRYLOV_BICGSTAB = 1 };
typedef struct {
    int num_steps;                  /* Initial Num. Pseudo Time Steps (default 20) */
    int max_newton;                 /* Maximum Number Iterations (default 10) */
    double newton_stop_tol;         /* Stopping Tolerance (default 1e-8) */
    double newton_current_norm_tol; /* Current Residual Norm Stopping Tolerance (default 1e-8) */
    int linear_max_iter;            /* Krylov iteration cap (default 20000) */
    double linear_tol;              /* Krylov relative residual (default 1e-12) */
    int krylov;                     /* PA_KRYLOV_GMRES (flexible, restarted; default) or PA_KRYLOV_BICGSTAB */
    int restart;                    /* GMRES restart length (default 50) */
    /* Checkpoint stride m of the path-dependent adjoint (no counterpart in the
     * reference, which keeps every step, src/PlasticityProblem.hpp:288-292).
     * 1 (default): every step's (u, c, pi) stays on the device.  m > 1: only
     * steps m-1, 2m-1, ... and the last are kept; the criterion value and the
     * backward sweep recompute the others segment by segment from the
     * preceding kept step (bitwise the same states: every reduction is
     * fixed-order), holding ceil(K/m) + m - 1 step states instead of K
     * (m = ceil(sqrt(K)) minimizes it).  The criterion calls must then get the
     * control of the last solution. */
    int checkpoint_stride;
} pa_plasticity_options;
void pa_plasticity_options_default(pa_plasticity_options* options);

typedef struct pa_plasticity_problem pa_plasticity_problem;
int pa_plasticity_problem_create(pa_plasticity* p, int num_bcs, const int* bc_dofs_host,
                                 const double* bc_values_host, const pa_plasticity_options* options,
                                 pa_plasticity_problem** problem);
int pa_plasticity_problem_destroy(pa_plasticity_problem* problem);
/* Precondition the BiCGStab solves with geometric multigrid on the nx x ny x nz
 * box mesh (4x4 nodal blocks: damped block-Jacobi smoothing, Galerkin coarse
 * operators, pivoted dense coarsest solve), rebuilt from every new tangent.
 * No counterpart in the reference (its solver stacks are external). */
int pa_plasticity_problem_set_multigrid(pa_plasticity_problem* problem, int nx, int ny, int nz,
                                        const pa_multigrid_options* options);
/* Forward solve for control z; returns PA_OK with *converged = 0 if a step's
 * Newton solve stopped without converging.  Synchronizes. */
int pa_plasticity_problem_solution(pa_plasticity_problem* problem, const double* z_dev, int* converged,
                                   void* stream);
/* Device bytes the problem holds for step states (u, c, pi). */
int pa_plasticity_problem_state_bytes(const pa_plasticity_problem* problem, size_t* bytes);
/* Device pointers to the states of step k (valid until the next solution call);
 * with checkpoint_stride m > 1 only the kept steps (m-1, 2m-1, ..., the last),
 * else PA_ERR_STATE. */
int pa_plasticity_problem_state(const pa_plasticity_problem* problem, int step, const double** u_dev,
                                const double** c_dev, const double** pi_dev);
/* Newton iterations, summed linear iterations and final residual norm of step k. */
int pa_plasticity_problem_stats(const pa_plasticity_problem* problem, int step, int* newton_iterations,
                                int* linear_iterations, double* residual_norm);
/* multiplier * sum over steps of the plastic work (host value).  Synchronizes. */
int pa_plasticity_problem_criterion_value(pa_plasticity_problem* problem, const double* z_dev, double multiplier,
                                          double* value_host, void* stream);
/* Total derivative of the criterion with respect to z (N doubles) by the
 * backward recursion over the steps.  Returns PA_ERR_STATE (the gradient is
 * then not to be used) when a step's adjoint linear solve stops short of
 * linear_tol (iteration limit or a non-finite residual).  Synchronizes. */
int pa_plasticity_problem_criterion_gradient(pa_plasticity_problem* problem, const double* z_dev,
                                             double multiplier, double* grad_dev, void* stream);

/* ------------------------------------------------- warm-start network */
/*
 * Kernels of the learned warm start (no counterpart in the reference): a small
 * 3D U-Net on the voxelised density field predicts the displacement field that
 * starts the linear solve.  BF16 activations/weights (raw uint16 bits),
 * FP32 accumulation, FP32 master weights.  A tensor is a channels-last view
 * (pointer, stride): element (voxel v, channel c) at ptr[v * stride + c], voxels
 * ordered (d, h, w) with w fastest.  Channel counts of convolutions are
 * multiples of 16.
 */
/* 3x3x3 convolution, zero padding: y = act(bias + conv(x, W)), bf16 and/or FP32
 * output (y_f32 may be null, y may be null when y_f32 is given); accumulate adds
 * the existing bf16 y before the activation.  wfrag: weights packed by
 * pa_cnn_pack_weights; bias may be null; relu 0/1.  D*H*W must be a multiple of 16. */
int pa_cnn_conv3d(int D, int H, int W, const void* x_dev, int x_stride, int cin, const void* wfrag_dev,
                  const float* bias_dev, int cout, void* y_dev, int y_stride, int relu, float* y_f32_dev,
                  int yf_stride, int accumulate, void* stream);
/* Pack FP32 weights W[27][cin][cout] (tap t = (dz+1)*9 + (dy+1)*3 + dx+1) into
 * the convolution's bf16 fragments; transposed != 0 packs the data-gradient
 * weights W'[t][cout][cin] = W[26 - t][cin][cout] (a convolution with cin and
 * cout exchanged).  wfrag: 27*cin*cout bf16. */
int pa_cnn_pack_weights(const float* w_dev, int cin, int cout, int transposed, void* wfrag_dev, void* stream);
/* Weight-gradient partials parts[chunk][27][cin][cout] = sum over the chunk's
 * voxels of x[nb(v, t)][ci] dy[v][co]; sum them with pa_cnn_sum_parts. */
int pa_cnn_conv3d_wgrad_parts(int D, int H, int W, const void* x_dev, int x_stride, int cin, const void* dy_dev,
                              int dy_stride, int cout, long long chunk, float* parts_dev, void* stream);
/* dst[k] = sum_p parts[p*n + k] (fixed order). */
int pa_cnn_sum_parts(long long n, int nparts, const float* parts_dev, float* dst_dev, void* stream);
/* Bias-gradient partials parts[block][c] over voxel chunks. */
int pa_cnn_bias_parts(long long V, const void* dy_dev, int dy_stride, int c, long long chunk, float* parts_dev,
                      void* stream);
/* 2x2x2 average pooling of a (D, H, W) tensor and its gradient (accumulate adds). */
int pa_cnn_pool(int D, int H, int W, const void* x_dev, int x_stride, int c, void* y_dev, int y_stride, void* stream);
int pa_cnn_pool_bwd(int D, int H, int W, const void* dy_dev, int dy_stride, int c, void* dx_dev, int dx_stride,
                    int accumulate, void* stream);
/* Nearest 2x upsampling to the fine grid (D, H, W), and its gradient onto the coarse grid (D, H, W). */
int pa_cnn_upsample(int D, int H, int W, const void* x_dev, int x_stride, int c, void* y_dev, int y_stride,
                    void* stream);
int pa_cnn_upsample_bwd(int D, int H, int W, const void* dy_dev, int dy_stride, int c, void* dx_dev, int dx_stride,
                        void* stream);
/* dy = y > 0 ? dy : 0 (the ReLU gradient, y the activation). */
int pa_cnn_relu_bwd(long long V, int c, const void* y_dev, int y_stride, void* dy_dev, int dy_stride, void* stream);
/* Weighted mean squared error of the first nc of c FP32 output channels against
 * target (V x nc), weight rho^3 + wmin with rho channel 0 of the bf16 input x
 * (the voxel density; x null: weight 1): gradient (bf16, c channels) and
 * per-256-voxel partial sums of the weighted squared error. */
int pa_cnn_mse(long long V, int nc, int c, const float* out_dev, int out_stride, const float* target_dev,
               const void* x_dev, int x_stride, float wmin, void* dout_dev, int dout_stride, float* part_dev,
               void* stream);
/* Energy-norm proxy loss of e = out - target over the first nc channels:
 * sum_v w_v (l2 |e_v|^2 + sum over the +x/+y/+z voxel neighbours u of |e_u - e_v|^2) / (V nc),
 * w = rho^3 + wmin from the input's density channel; gradient (bf16) and per-256-voxel partial sums. */
int pa_cnn_energy_loss(int D, int H, int W, int nc, int c, const float* out_dev, int out_stride,
                       const float* target_dev, const void* x_dev, int x_stride, float wmin, float l2, void* dout_dev,
                       int dout_stride, float* part_dev, void* stream);
/* One Adam step on n FP32 parameters (lr_t includes the bias correction). */
int pa_cnn_adam(long long n, float lr_t, float beta1, float beta2, float eps, const float* grad_dev, float* w_dev,
                float* m1_dev, float* m2_dev, void* stream);
/* FP32 view -> bf16 view; bf16 y += bf16 x. */
int pa_cnn_to_bf16(long long V, int c, const float* x_dev, int x_stride, void* y_dev, int y_stride, void* stream);
int pa_cnn_add_bf16(long long V, int c, const void* x_dev, int x_stride, void* y_dev, int y_stride, void* stream);
/* Box-mesh node field (N x nc FP64, N = (D+1)(H+1)(W+1), the pa_box_mesh node
 * order) to the hex voxels (mean of 8 nodes) times scale, clipped to
 * [-clip, clip] when clip > 0, FP32, written at channel offset oc of a view
 * with stride os; and voxels back to nodes (mean of the adjacent voxels). */
int pa_cnn_nodes_to_voxels(int D, int H, int W, int nc, const double* nodes_dev, double scale, double clip,
                           float* out_dev, int out_stride, int channel_offset, void* stream);
int pa_cnn_voxels_to_nodes(int D, int H, int W, int nc, const float* vox_dev, int vox_stride, double scale,
                           double* nodes_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif
