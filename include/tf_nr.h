/*
 * tf_nr.h -- C ABI of the MI355X implicit Newton-Raphson step for growing prism tissues.
 *
 * One handle owns one or more independent tissues ("segments") resident in GPU memory.
 * Every vector crossing this boundary uses the reference's interleaved DoF order
 * [x0, y0, z0, x1, y1, z1, ...]; element-local arrays use the reference's node-major
 * order 3a+i.  Host pointers are read or written synchronously inside the call and
 * never retained.  Every entry point returns TF_OK (0) or a tf_status code and never
 * exits the process; tf_last_error() describes the last failure.
 *
 * Reference call sites each entry point stands in for are cited with paths relative
 * to TissueFolding/TissueFolding/SourceCode/ of the reference tree.
 */
#ifndef TF_NR_H
#define TF_NR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tf_handle tf_handle;

typedef enum tf_status {
    TF_OK = 0,
    TF_NOT_CONVERGED = 1, /* NR hit its iteration cap (the reference prints and continues) */
    TF_KRYLOV_STALL = 2,  /* Krylov breakdown or iteration cap without reaching tolerance    */
    TF_NEG_JACOBIAN = 3,  /* det Fe <= 0 in some element (inverted prism)                    */
    TF_NAN = 4,           /* NaN/Inf in det Fe (positions, growth state), residual or update */
    TF_OOM = 5,           /* device allocation failed                                        */
    TF_HIP_ERROR = 6,     /* any other HIP runtime error                                     */
    TF_INVALID = 7        /* invalid argument                                                */
} tf_status;

/* Mesh and material description, copied at tf_create.
 * Replaces the state the reference keeps in Node / Prism / ReferenceShapeBase objects:
 * Node::Position, FixedPos, externalViscosity (Node.h:19-38), ShapeBase::NodeIds,
 * ReferenceShape->Positions, lambda/mu/internalViscosity (ShapeBase.h:178-187),
 * exposed apical/basal faces (Prism.cpp:2674-2711). */
typedef struct tf_mesh_desc {
    int32_t n_nodes;
    int32_t n_elements;
    const double* node_pos;       /* [3*n_nodes]  initial positions                            */
    const int32_t* conn;          /* [6*n_elements] node ids, 0-2 basal, 3-5 apical            */
    const double* ref_pos;        /* [18*n_elements] element reference positions, or NULL =     */
                                  /*   gathered from node_pos                                  */
    const double* lambda;         /* [n_elements] Lame lambda (stiffness multiplier applied)    */
    const double* mu;             /* [n_elements] Lame mu                                       */
    const double* eta_int;        /* [n_elements] internal viscosity, or NULL = 0               */
    const double* eta_ext;        /* [3*n_nodes] external viscosity per DoF, or NULL = 0        */
    const uint8_t* fixed;         /* [3*n_nodes] 1 = fixed DoF, or NULL = none                  */
    const uint8_t* exposed;       /* [n_elements] bit0 basal face, bit1 apical face, or NULL    */
    const double* growth_rate;    /* [3*n_elements] growth rate per second (x,y,z), or NULL     */
    int32_t n_segments;           /* independent tissues in this handle (>= 1)                  */
    const int32_t* node_seg_ptr;  /* [n_segments+1] contiguous node ranges, NULL if 1 segment   */
    const int32_t* elem_seg_ptr;  /* [n_segments+1] contiguous element ranges, NULL if 1 seg.   */
    const int8_t* node_placement; /* [n_nodes] Node::tissuePlacement: 0 basal, 1 apical, 2 mid,  */
                                  /*   3 lateral; NULL = corners of exposed caps (implicit packing) */
} tf_mesh_desc;

/* TF_KRYLOV_CG1: CG with one fused kernel per iteration (single-reduction recurrences of
 * Chronopoulos and Gear); TF_KRYLOV_CG: textbook PCG, two kernels per iteration;
 * TF_KRYLOV_PIPECG: pipelined CG (Ghysels and Vanroose), one kernel per iteration whose SpMV
 * operand was completed by the previous kernel; TF_KRYLOV_BICGSTAB for the non-symmetric
 * Jacobian of internal viscosity (value 2 is not used). */
enum { TF_KRYLOV_BICGSTAB = 0, TF_KRYLOV_CG = 1, TF_KRYLOV_CG1 = 3, TF_KRYLOV_PIPECG = 4 };
enum { TF_PREC_FP64 = 0, TF_PREC_FP32 = 1, TF_PREC_BF16 = 2 };
/* TF_PRECOND_AUTO (the default): two-level for a handle with 4096 nodes or more, else block
 * Jacobi.  TF_PRECOND_BLOCK_JACOBI: the inverse 3x3 diagonal blocks of M.
 * TF_PRECOND_TWO_LEVEL (CG solvers): block Jacobi over the node columns (the diagonal blocks of
 * M over the nodes that prisms' vertical edges join; 3x3 nodal blocks when the columns are not
 * uniform or column_blocks = 0) plus an additive coarse correction P (P^T M P)^-1 P^T r, P the
 * rigid-body modes of aggregates of whole columns (square cells of the tissue's x-y extent).
 * Every tissue of a handle has its own aggregates and its own block of the coarse operator (at
 * most 4096 coarse unknowns per tissue).  A tissue's block is rebuilt at the first Newton
 * iteration of a timestep once its Newton iteration 1 Krylov count exceeds 1.25x its count right
 * after its last rebuild; every block coarse_refresh timesteps after the last full rebuild, and
 * for every solve outside the Newton loop.  The column-block inverses are held in FP32 and the
 * coarse inverse in FP16 scaled by its largest diagonal entry (they only precondition:
 * residuals, matrix and Krylov vectors stay in the solver's precision; precond_fp64 holds both
 * in FP64). */
enum { TF_PRECOND_AUTO = -1, TF_PRECOND_BLOCK_JACOBI = 0, TF_PRECOND_TWO_LEVEL = 1 };

typedef struct tf_solver_opts {
    int32_t krylov;            /* TF_KRYLOV_*                                                  */
    int32_t krylov_max_iter;   /* per linear solve                                             */
    double krylov_rtol;        /* ||r|| <= max(rtol*||g||, atol)                               */
    double krylov_atol;
    int32_t iters_per_sync;    /* Krylov iterations between host convergence checks            */
    double nr_tol;             /* ||dU||_2 <= nr_tol (absolute, per tissue) -- reference 1e-8  */
    int32_t nr_max_iter;       /* reference: 20 (stops when the count exceeds it)              */
    int32_t tangent_precision; /* TF_PREC_*: arithmetic of the element tangent contraction     */
    int32_t matrix_precision;  /* TF_PREC_FP64 or TF_PREC_FP32 storage of the Jacobian values  */
    int32_t growth_rotation;   /* 1: rotate growth by the element's mean rotation (1 deg rule) */
    int32_t use_graph;         /* 1: replay Krylov iteration chunks from a captured hipGraph   */
    int32_t krylov_forcing;    /* 1: inside the NR loop each linear solve stops at the accuracy
                                  the next Newton update needs (inexact Newton, per tissue);
                                  0: every solve to krylov_rtol                                 */
    int32_t xcd_tile_map;      /* 1: each XCD runs a contiguous run of Krylov row tiles         */
    int32_t krylov_warm_start; /* 1: the solves of NR iterations 0 and 1 start from the same
                                  iteration's update of the previous timestep                  */
    int32_t implicit_packing;  /* 1: tf_timestep detects packing pairs after the growth update,
                                  as Simulation::runOneStep does (Simulation.cpp:37, :5911-5920);
                                  0: tf_timestep clears the pair list                          */
    int32_t preconditioner;    /* TF_PRECOND_*: the preconditioner of the CG solvers             */
    int32_t coarse_aggregates; /* TF_PRECOND_TWO_LEVEL: target number of node aggregates of the
                                  coarse space (0: one per 44 nodes, at most 600, and at most two
                                  aggregates per compute unit)                                  */
    int32_t coarse_refresh;    /* TF_PRECOND_TWO_LEVEL: most timesteps between rebuilds of the
                                  coarse operator (0 or 1: every timestep; default 100)         */
    int32_t column_blocks;     /* TF_PRECOND_TWO_LEVEL smoother: 1 the blocks of the node columns
                                  (nodes joined by prisms' vertical edges) when every column has
                                  the same 1-4 nodes, else 3x3 nodal blocks (default); 0 always
                                  3x3 nodal blocks                                              */
    int32_t precond_fp64;      /* TF_PRECOND_TWO_LEVEL: 1 holds the column-block inverses and the
                                  coarse inverse in FP64 (every array of the solve in FP64);
                                  0 (default): FP32 column blocks, FP16 coarse inverse scaled by
                                  its largest diagonal entry                                    */
    int32_t mixed_fp64_from;   /* with tangent_precision or matrix_precision below FP64: Newton
                                  iterations from this index on assemble their tangents and solve
                                  with FP64 Jacobian values (the early, large updates run in the
                                  reduced precision, the last ones converge as in FP64); default 1,
                                  < 0: the reduced precision in every iteration                  */
    int32_t nr_graph;          /* 1 (default, with use_graph): a Newton iteration whose solve has a
                                  prediction from the previous timestep runs assembly, warm start,
                                  solve, update and convergence test on the device and the host
                                  reads one report per iteration (a solve that is not finished by
                                  then continues with host checks); 2: the same, the iteration's
                                  launches replayed from one captured hipGraph; 0: the host checks
                                  the assembly status and every Krylov round                     */
    int32_t tl_persistent;     /* TF_PRECOND_TWO_LEVEL: 1 (default) runs a solve of one tissue whose
                                  aggregates fit on the device at once, one workgroup each, as one
                                  launch that keeps the aggregate's coarse rows in LDS, with a grid
                                  barrier after each half of an iteration (same iterates as the two
                                  launches per iteration it replaces); a barrier wait that gives up
                                  (workgroups not co-resident) falls back to BiCGSTAB; 2: the same
                                  as a cooperative launch (co-residency guaranteed by the runtime,
                                  slower to launch); 0: two launches per iteration, replayed from
                                  graphs                                                         */
    int32_t tl_barrier_polls;  /* polls of one grid-barrier wait of the persistent solve before it
                                  gives up (0, the default: 2^22); a small value forces the
                                  fallback that a grid without co-residency takes              */
} tf_solver_opts;

typedef struct tf_step_stats {
    int32_t nr_iterations;       /* max over tissues                                           */
    int32_t krylov_iterations;   /* summed over NR iterations (max over tissues)               */
    int32_t converged;           /* 1 when every tissue met nr_tol                             */
    int32_t rotated_elements;    /* growth increments rotated this step                        */
    int32_t krylov_fallbacks;    /* linear solves that fell back from CG to BiCGSTAB           */
    int32_t packing_pairs;       /* implicit-packing pairs in force during the step            */
    double last_du_norm;         /* max over tissues of the last ||dU||_2                      */
    double ms_grow, ms_assemble, ms_solve, ms_total;
    int32_t krylov_by_newton[4]; /* Krylov iterations of Newton iterations 0-3 (max over tissues) */
    int32_t two_level;           /* 1 when the step's CG solves used the two-level preconditioner */
    int32_t krylov_stalls;       /* linear solves that stopped short of their tolerance (CG and its
                                    BiCGSTAB fallback) with a residual at most 1e-2 of the right-hand
                                    side: taken as inexact Newton steps.  A solve that ends above
                                    that ends its tissue's step: tf_nr_step returns TF_KRYLOV_STALL */
    int32_t host_waits;          /* times the host waited for the device during the NR loop       */
    int32_t nr_graph_builds;     /* Newton iterations captured into a new graph (nr_graph = 2)    */
    int32_t krylov_persistent;   /* linear solves run as one persistent launch (tl_persistent)     */
} tf_step_stats;

/* ---- lifetime --------------------------------------------------------------------------- */
/* Set-up: NewtonRaphsonSolver::NewtonRaphsonSolver (NewtonRaphsonSolver.cpp:16-36) plus the
 * element construction of Simulation::initiateMesh (Simulation.cpp:2947-3009, Prism.cpp:9-269)
 * and the shape-function set-up (Prism.cpp:705-728). */
int tf_create(const tf_mesh_desc* desc, tf_handle** out);
/* NewtonRaphsonSolver::~NewtonRaphsonSolver (NewtonRaphsonSolver.cpp:38-70). */
void tf_destroy(tf_handle* h);
/* Last error text of h (or of the last failed tf_create when h is NULL). */
const char* tf_last_error(const tf_handle* h);
void tf_default_options(tf_solver_opts* opts);
int tf_set_options(tf_handle* h, const tf_solver_opts* opts);
int tf_get_options(const tf_handle* h, tf_solver_opts* opts);
int tf_get_sizes(const tf_handle* h, int32_t* n_nodes, int32_t* n_elements, int32_t* n_blocks,
                 int32_t* n_segments);

/* ---- state ------------------------------------------------------------------------------- */
/* Node positions (Node::Position / NewtonRaphsonSolver::uk). */
int tf_set_positions(tf_handle* h, const double* x);
int tf_get_positions(tf_handle* h, double* x);
/* Growth and material state (ShapeBase::Fg / InvFg, ShapeBase.cpp:1215-1251; lambda, mu). */
int tf_set_growth_rates(tf_handle* h, const double* rate);
int tf_set_fg(tf_handle* h, const double* Fg);
/* Per-element growth increments for the following growth updates (tf_grow, tf_timestep): the
 * product of the step's growth-function increments, row-major 3x3 per element, as
 * Simulation::calculateGrowth builds them (Simulation.cpp:9651-9670, ShapeBase::updateGrowthIncrement
 * ShapeBase.cpp:567-586) -- angled uniform and ring functions (calculateFgFromRates,
 * ShapeBase.cpp:396-421) and grid functions (calculateFgFromGridCorners, :432-565).  They replace
 * the diagonal exp(rate dt) increments; the growth-rotation rule still applies.  NULL returns to
 * the rates. */
int tf_set_growth_increments(tf_handle* h, const double* dFg);
/* Growth functions evaluated on the device at every growth update (tf_grow, tf_timestep), in
 * place of host-given increments: Simulation::calculateGrowth (Simulation.cpp:9651-9670) over the
 * GrowthFunctions list -- uniform (type 1, calculateGrowthUniform :9757-9799), ring (type 2,
 * calculateGrowthRing :9801-9898) and grid (type 3, calculateGrowthGridBased :10065-10117, with
 * ShapeBase::calculateFgFromGridCorners, ShapeBase.cpp:432-565) -- columnar layer, tissue
 * weight 1.  A function is active while t0 <= time < t1 at the step's start time (tf_set_time;
 * tf_timestep advances the time by dt); active functions multiply from the left in list order. */
enum { TF_GROWTH_UNIFORM = 1, TF_GROWTH_RING = 2, TF_GROWTH_GRID = 3 };
typedef struct tf_growth_function {
    int32_t kind;               /* TF_GROWTH_*                                                  */
    int32_t nx, ny;             /* grid points along x and y (grid, >= 2 each)                  */
    double t0, t1;              /* active while t0 <= time < t1 (seconds)                       */
    double rate[3];             /* 1/s along the rotated x, y and z (uniform; ring: the rate at
                                   the inner radius, falling linearly to 0 at the outer one)    */
    double angle;               /* xy angle of the rates, radians (uniform, ring)               */
    double centre[2];           /* ring centre (read as float, as the reference stores it)      */
    double inner, outer;        /* ring radii (float)                                           */
    double zmin, zmax;          /* grid: range of the element's initial relative height         */
    const double* grid_rates;   /* grid: [nx][ny][3] rates, 1/s                                 */
    const double* grid_angles;  /* grid: [nx][ny] angles, radians                               */
} tf_growth_function;
/* n functions (n = 0 removes them: back to the rates).  initial_z[E]: every element's initial
 * relative height (ShapeBase::setInitialZPosition); rel_src[E]: the element whose relative
 * position an element takes (its column's apical element for the first layers - 2 elements of
 * the column, else itself).  pinned = 1: grids read the relative positions pinned on the mesh at
 * this call (GridGrowthsPinnedOnInitialMesh) and re-pinned by tf_pin_growth_positions; 0: the
 * current positions every step.  interpolation: 0 nearest corner, 1 bilinear
 * (GridGrowthsInterpolationType).  Replaces host-given increments. */
int tf_set_growth_functions(tf_handle* h, int32_t n, const tf_growth_function* f, const double* initial_z,
                            const int32_t* rel_src, int32_t pinned, int32_t interpolation);
/* Pins the grids' relative positions on the current positions (the reference's pinning update at
 * each PinningUpdateTime, Simulation::checkForPinningPositionsUpdate, Simulation.cpp:10425-10443). */
int tf_pin_growth_positions(tf_handle* h);
/* Simulation time (seconds) at the start of the next timestep. */
int tf_set_time(tf_handle* h, double t);
int tf_get_time(const tf_handle* h, double* t);
/* The growth increments of the last growth update, row-major 3x3 per element (device growth
 * functions or host-given increments; TF_INVALID under rate-based growth). */
int tf_get_growth_increments(tf_handle* h, double* dFg);
int tf_get_fg(tf_handle* h, double* Fg, double* InvFg);
int tf_set_material(tf_handle* h, const double* lambda, const double* mu, const double* eta_int);

/* ---- DoF constraints beyond FixedPos ----------------------------------------------------------- */
/* Slave/master DoF binding: NewtonRaphsonSolver::slaveMasterList with
 * calculateBoundKWithSlavesMasterDoF and equateSlaveDisplacementsToMasters
 * (NewtonRaphsonSolver.cpp:154-235, :397-399).  pairs[2k] = slave DoF, pairs[2k+1] = master DoF
 * (DoF = 3*node + dim).  Every later assembly applies M <- N^T M N + I_slaves and g <- N^T g after
 * the fixed DoFs, and every solve sets dU[slave] = dU[master].  Replaces the previous list
 * (n_pairs = 0 removes all binding) and widens the Jacobian pattern (tf_get_sizes reports the new
 * block count).  A slave is bound at most once, a slave is never a master, and both DoFs belong to
 * one tissue; otherwise TF_INVALID. */
int tf_set_bound_dofs(tf_handle* h, int32_t n_pairs, const int32_t* pairs);
/* Ablated nodes (Simulation::AblatedNodes): a zero Jacobian diagonal entry of theirs is set to 1
 * before the fixed DoFs are applied (NewtonRaphsonSolver::checkJacobianForAblatedNodes,
 * NewtonRaphsonSolver.cpp:350-361).  Replaces the previous list; n = 0 clears it. */
int tf_set_ablated_nodes(tf_handle* h, int32_t n, const int32_t* nodes);

/* ---- implicit packing ------------------------------------------------------------------------ */
/* Node masses (Node::mass) from the current Fg: every prism adds det(Fg) V_ref / 6 to its nodes
 * (Simulation::updateNodeMasses, Simulation.cpp:7552-7560; ShapeBase.cpp:1243-1245, :3938-3942;
 * V_ref by Prism::calculateReferenceVolume, Prism.cpp:876-908).  Ablated nodes get 0. */
int tf_get_node_masses(tf_handle* h, double* mass);
/* Simulation::updateNodeMasses then Simulation::detectPacingNodes (Simulation.cpp:8167-8440, with
 * getAverageSideLength :3205-3265 and calculateBoundingBox :9148-9189) on the current positions,
 * for every tissue.  The pairs and their weights hold for every following assembly until the
 * next detection or tf_set_packing_pairs: each assembly adds the sigmoid packing forces to g and
 * their Jacobian to M (calculatePackingForcesImplicit3D, calculatePackingJacobian3D,
 * Simulation.cpp:7149-7295, addPackingForces :8869-8906).  New pairs widen the Jacobian pattern.
 * n_pairs may be NULL. */
int tf_detect_packing(tf_handle* h, int32_t* n_pairs);
/* Replaces the pair list (pacingNodeCouples0/1, initialWeightPointx/y/z): pairs[2k], pairs[2k+1]
 * are nodes of one tissue, weights[3k..3k+2] the pair's unit weights, threshold[s] the packing
 * threshold T of tissue s (packingThreshold), > 0 where a tissue has pairs.  Node masses are
 * updated from the current Fg.  Every unordered pair may appear once.  n_pairs = 0 clears. */
int tf_set_packing_pairs(tf_handle* h, int32_t n_pairs, const int32_t* pairs, const double* weights,
                         const double* threshold);
/* The pair list in force: n_pairs always; pairs [2*n], weights [3*n], threshold [n_segments] and
 * detect_grid [50*n_segments] (the 10 x 5 detection thresholds of each tissue, row-major) when not
 * NULL.  Thresholds and grid are those of the last detection (threshold: or of
 * tf_set_packing_pairs). */
int tf_get_packing_pairs(tf_handle* h, int32_t* n_pairs, int32_t* pairs, double* weights, double* threshold,
                         double* detect_grid);

/* ---- per-step set-up ----------------------------------------------------------------------- */
/* Simulation::updateGrowthRotationMatrices + calculateGrowth + ShapeBase::growShapeByFg
 * (Simulation.cpp:5876-5909, :9651-9799; ShapeBase.cpp:396-421, :1215-1251, :1656-1811). */
int tf_grow(tf_handle* h, double dt, int32_t* n_rotated);
/* Simulation::updateNodeViscositySurfaces (Simulation.cpp:7563-7581). */
int tf_update_viscosity_surfaces(tf_handle* h);
int tf_get_viscosity_surfaces(tf_handle* h, double* Sa);

/* ---- Newton-Raphson pieces (Simulation::updateStepNR, Simulation.cpp:6449-6590) ---------------- */
/* constructUnMatrix + initialteUkMatrix (NewtonRaphsonSolver.cpp:135-145). */
int tf_begin_step(tf_handle* h);
/* calculateDisplacementMatrix, calculateForcesAndJacobianMatrixNR, writeForcesTogeAndgvInternal,
 * writeImplicitElementalKToJacobian, calculateExternalViscousForcesForNR,
 * addImplicitKViscousExternalToJacobian, calculateSumOfInternalForces, calcutateFixedK
 * (NewtonRaphsonSolver.cpp:148-152, :237-348, :363-367): residual g and Jacobian M on device. */
int tf_assemble(tf_handle* h, double dt);
/* gSum (NewtonRaphsonSolver.h:35). */
int tf_get_residual(tf_handle* h, double* g);
/* Jacobian in natural orientation M = d(internal force)/du, which is the transpose of the
 * reference's K (the reference solves K^T dU = g, NewtonRaphsonSolver.cpp:636-643).
 * 3x3 block CSR: row_ptr [n_nodes+1], col [n_blocks], val [9*n_blocks] row-major in-block. */
int tf_get_jacobian_bsr(tf_handle* h, int32_t* row_ptr, int32_t* col, double* val);
/* Element-local arrays of one prism at the current uk/un, in the reference's layout:
 * ElementalElasticSystemForces, ElementalInternalViscousSystemForces (ShapeBase.cpp:2198-2226),
 * TriPointKe and TriPointKv as stored by the reference (ShapeBase.cpp:2619-2705). */
int tf_element_dump(tf_handle* h, int32_t e, double dt, double* fe, double* fv, double* Ke, double* Kv);
/* tf_element_dump for the n prisms e0 .. e0 + n - 1 without allocating: no hipMalloc, hipFree or
 * allocating library call.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under stream or graph capture, tight inspection loops, harnesses
 * that forbid allocation).  The kernel writes into a static device array of the library
 * (room for 512 prisms, 2.8 MB, one copy per GPU, held until the process ends); tf_create resolves
 * its address.  fe and fv receive [n][18], Ke and Kv [n][324], each prism as tf_element_dump gives
 * it; any of them may be NULL.  n > 512 returns TF_INVALID.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (several host threads, any handles) race. */
int tf_element_dump_range_noalloc(tf_handle* h, int32_t e0, int32_t n, double dt, double* fe, double* fv, double* Ke,
                                  double* Kv);
/* solveForDeltaU (NewtonRaphsonSolver.cpp:382-404): M dU = g by preconditioned Krylov. */
int tf_solve(tf_handle* h, double* dU, int32_t* krylov_iters, double* rel_residual);
/* checkConvergenceViaDeltaU + updateUkInIteration (NewtonRaphsonSolver.cpp:727-761). */
int tf_update_uk(tf_handle* h, double* du_norm, int32_t* converged);
/* The whole NR loop of one step (Simulation::updateStepNR). */
int tf_nr_step(tf_handle* h, double dt, tf_step_stats* stats);
/* Growth + viscosity surfaces + node masses and packing detection (when implicit_packing) + NR
 * loop (the NR-adjacent part of Simulation::runOneStep, Simulation.cpp:5820-5974). */
int tf_timestep(tf_handle* h, double dt, tf_step_stats* stats);
/* Per-tissue results of the last NR step: iterations, last ||dU||, converged flag. */
int tf_get_segment_stats(tf_handle* h, int32_t* nr_iters, double* du_norm, int32_t* converged);

/* ---- linear algebra on the assembled Jacobian and the discrete adjoint ---------------------------
 * No reference counterpart: the reference has no sensitivities.  They serve gradient-based
 * fitting of growth rates and stiffnesses through the implicit timestep. */
/* Fg and InvFg exactly as given (row-major 3x3 per element), e.g. to restore a checkpoint. */
int tf_set_growth_state(tf_handle* h, const double* Fg, const double* InvFg);
/* Growth-rotation angle about z applied to each element's increment by the last tf_grow
 * (0 where the 1-degree rule did not fire). */
int tf_get_growth_rotation(tf_handle* h, double* theta_z);
/* y = M x (transpose = 0) or y = M^T x (transpose = 1) with the current FP64 Jacobian. */
int tf_apply_jacobian(tf_handle* h, int32_t transpose, const double* x, double* y);
/* Solves M x = b or M^T x = b with the Krylov solver and the current Jacobian (FP64 values). */
int tf_solve_system(tf_handle* h, int32_t transpose, const double* b, double* x, int32_t* krylov_iters,
                    double* rel_residual);
/* Vector-Jacobian product of the residual g at the current state with y: y^T dg/dInvFg
 * ([E][9]), y^T dg/dlambda and y^T dg/dmu ([E]). */
int tf_adjoint_elements(tf_handle* h, const double* y, double* invfg_bar, double* lambda_bar, double* mu_bar);
/* Zeroes the accumulated parameter adjoints and the growth-state adjoint. */
int tf_adjoint_reset(tf_handle* h);
/* One reverse timestep.  State on entry: positions = x_{n+1} (converged), un = x_n, surfaces at
 * x_n, Fg/InvFg = Fg_{n+1}.  Inputs: xbar_next = dL/dx_{n+1}, fg_prev = Fg_n, theta_z = the
 * step's growth-rotation angles (held fixed).  Output: xbar_prev = dL/dx_n.  Accumulates dL/drate
 * ([E][3]), dL/dlambda, dL/dmu and carries dL/dFg to step n.  Needs eta_int = 0 and no bound or
 * ablated DoFs (TF_INVALID otherwise). */
int tf_adjoint_step(tf_handle* h, double dt, const double* xbar_next, const double* fg_prev, const double* theta_z,
                    double* xbar_prev, int32_t* krylov_iters);
int tf_adjoint_get(tf_handle* h, double* rate_bar, double* lambda_bar, double* mu_bar, double* fg_bar);

/* The two-level preconditioner's coarse operator Ac = P^T M P of the current Jacobian and the FP64
 * inverse of the operator its last rebuild factored ([nc][nc] row-major each; either may be NULL;
 * call with both NULL to learn nc).  No reference counterpart: diagnostics of this implementation's
 * preconditioner.  TF_INVALID before a two-level solve has built one. */
int tf_get_coarse_operator(tf_handle* h, int32_t* nc, double* Ac, double* Ac_inv);

/* ---- tuning (no reference counterpart: schedule choices of this implementation) ----------------- */
/* Row tiling of the Krylov kernels: workgroup size wg in {256, 512, 1024} and tile capacity
 * 1 <= tile_cap <= min(4 wg, 2048) blocks (a tile never holds more than wg/3 block rows). */
int tf_set_tiling(tf_handle* h, int32_t wg, int32_t tile_cap);
int tf_get_tiling(const tf_handle* h, int32_t* wg, int32_t* tile_cap, int32_t* n_tiles);
/* Times n_iters Krylov iterations (the two-level PCG when it is active, else the single-kernel
 * CG) on the Jacobian at the current positions for every
 * candidate tiling (power-of-two capacities, and capacities that cut the rows into 1-4 tiles per
 * compute unit), keeps the fastest; per-variant results (up to 32) are written to the arrays. */
int tf_autotune(tf_handle* h, double dt, int32_t n_iters, int32_t* n_variants, int32_t* wg, int32_t* tile_cap,
                double* ms_per_iter);

/* ---- measurement ------------------------------------------------------------------------------ */
/* Average device time (ms) per launch of the named kernel over the calls since the last reset,
 * recorded with HIP events on the stream the kernel runs on.  name: "assemble_elements",
 * "assemble_gather", "spmv". */
int tf_kernel_timing(tf_handle* h, const char* name, int32_t enable, double* avg_ms, int64_t* launches);
/* Algorithmic bytes of one launch of the named kernel (unique minimum traffic): "spmv",
 * "coarse", "assemble_elements", "assemble_gather"; "persistent": one iteration of the persistent
 * two-level solve. */
int tf_kernel_bytes(tf_handle* h, const char* name, double* bytes);
/* Average device time (ms) of one launch of the kernel priced under "spmv" (the SpMV launch of
 * the active solver: k_tp_q for the two-level PCG, k_cg_q for the textbook CG, k_cg1, k_pipe),
 * n_launches replayed back to back in one graph on the current Jacobian and timed with HIP events
 * around the replay.  Leaves the Krylov work vectors undefined.  0 for BiCGSTAB. */
int tf_kernel_probe(tf_handle* h, int32_t n_launches, double* avg_ms);
/* The same probe for a named kernel: "spmv" (as tf_kernel_probe), "coarse" (k_tp_z, the
 * two-level PCG's coarse-correction launch: 0 when the two-level preconditioner is not active) or
 * "persistent" (ms per iteration of one persistent launch of n_launches iterations with tolerance 0:
 * 0 when the solve does not run persistently). */
int tf_kernel_probe_named(tf_handle* h, const char* name, int32_t n_launches, double* avg_ms);
/* Krylov chunk replays timed with HIP events on the handle's stream since the last reset:
 * total device time of the replays (ms), the iterations they launched (no-op launches after a
 * solve's end included) and the iterations the solves took.  enable >= 0 resets the sums and
 * switches the timing on (1) or off (0); -1 reads only. */
int tf_krylov_timing(tf_handle* h, int32_t enable, double* replay_ms, int64_t* launched, int64_t* iterations);
/* Device-to-device copy of `bytes` (1 MiB .. 16 GiB) into a second buffer, `reps` times on the
 * handle's stream: the best read + write bandwidth (GB/s), the achievable HBM rate that the
 * kernels' roofline fractions are read against. */
int tf_copy_bandwidth(tf_handle* h, int64_t bytes, int32_t reps, double* best_GBps);

#ifdef __cplusplus
}
#endif
#endif /* TF_NR_H */
