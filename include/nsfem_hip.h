/*
 * nsfem_hip.h - C ABI of the MI355X (gfx950) Chorin-projection library.
 *
 * Every entry point returns an int status (NSFEM_OK on success) and records a
 * thread-local message readable with nsfem_last_error().  Status codes follow
 * the reference's exception split: std::invalid_argument -> NSFEM_EINVAL,
 * std::runtime_error -> NSFEM_ERUNTIME, failed HIP calls -> NSFEM_EDEVICE, a
 * non-finite solver residual -> NSFEM_ENONFINITE.
 *
 * Device pointers are plain HIP device addresses (hipMalloc'd by
 * nsfem_malloc or owned by a handle).  Blocking behaviour matches the
 * reference: every call that the reference runs synchronously (cudaMemcpy,
 * host-pointer-mode cuBLAS reductions, solver calls) has completed on return.
 * Vectors with several channels are planar: channel c of a vector with n rows
 * per channel starts at element c * n.
 *
 * Reference citations are relative to kosyonikolov/NsFemOptim.
 */
#ifndef NSFEM_HIP_H
#define NSFEM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NSFEM_OK 0
#define NSFEM_EINVAL 1
#define NSFEM_ERUNTIME 2
#define NSFEM_EDEVICE 3
#define NSFEM_ENONFINITE 4

/* Value types; ids match the linalg .bin dtypeId (cpp/libs/LinAlg/src/io.cpp:11-21). */
#define NSFEM_F32 0
#define NSFEM_F64 1

const char * nsfem_last_error(void);
int nsfem_version(int * major, int * minor);
int nsfem_synchronize(void);

/* ------------------------------------------------------------------------
 * Device buffers.  Replace cu::vec<T> (cpp/libs/CudaUtils/include/cu/vec.h:16-387):
 * allocation (:73-86, size 0 rejected), upload/download (:248-274),
 * copyTo (:167-178), memsetZero (:180-191).
 * ---------------------------------------------------------------------- */
int nsfem_malloc(size_t bytes, void ** dptr);
int nsfem_free(void * dptr);
int nsfem_upload(void * dst, const void * hostSrc, size_t bytes);
int nsfem_download(void * hostDst, const void * src, size_t bytes);
int nsfem_copy(void * dst, const void * src, size_t bytes);
int nsfem_memset_zero(void * dptr, size_t bytes);

/* ------------------------------------------------------------------------
 * Device CSR.  Replaces cu::csrF (cpp/libs/CudaUtils/src/csrF.cu:5-40):
 * int32, 0-based, deep copy of host arrays, values mutable in place.
 * ---------------------------------------------------------------------- */
typedef struct nsfem_csr nsfem_csr;
int nsfem_csr_create(int rows, int cols, int nnz, const int * rowStart, const int * column, const void * values, int dtype,
                     nsfem_csr ** out);
int nsfem_csr_destroy(nsfem_csr * m);
int nsfem_csr_values(nsfem_csr * m, void ** dvalues);
int nsfem_csr_shape(const nsfem_csr * m, int * rows, int * cols, int * nnz, int * dtype);

/* y = A x.  Replaces cu::spmv::compute (cpp/libs/CudaUtils/src/spmv.cu:45-63). */
int nsfem_spmv(nsfem_csr * m, const void * x, void * y);
/* Y = A X, X and Y column-major with ld = cols / rows.
 * Replaces cu::spmm::compute (cpp/libs/CudaUtils/src/spmm.cu:45-70). */
int nsfem_spmm(nsfem_csr * m, int numCh, const void * x, void * y);
/* General form: y = alpha A x + beta y, explicit leading dimensions, lanes per
 * row (0 = automatic; 4/8/16/32/64). */
int nsfem_csr_mul(nsfem_csr * m, int numCh, const void * x, long ldx, void * y, long ldy, double alpha, double beta, int lanes);

/* ------------------------------------------------------------------------
 * Vector operations.  Replace cu::saxpy / cu::scale (cpp/libs/CudaUtils/src/blas.cu:51-67),
 * cublasSgeam used as a vector add (execs/NavierStokes/src/chorinCuda.cu:246-250),
 * cublasSdot (src/conjGradF.cu:57), cublasSnrm2 (src/gaussSeidel.cu:278),
 * cusparseScatter / cusparseGather (chorinCuda.cu:73, :134, :146).
 * Reductions accumulate in FP64 and are deterministic.
 * ---------------------------------------------------------------------- */
int nsfem_axpy(int dtype, long n, double alpha, const void * src, void * dst); /* dst += alpha src */
int nsfem_scale(int dtype, long n, double alpha, void * dst);
/* Achievable streaming bandwidth of the device: a device-to-device copy of
 * `bytes` (rounded down to 16; use more than the 256 MB Infinity Cache for
 * the HBM figure) repeated `reps` times back to back, timed with HIP events
 * on the library stream; *bytesPerSecond = read + write bytes per second.
 * Allocates and frees its two buffers.  No reference counterpart
 * (measurement hook). */
int nsfem_copy_bandwidth(size_t bytes, int reps, double * bytesPerSecond);
int nsfem_geam(int dtype, long n, double alpha, const void * a, double beta, const void * b, void * c);
int nsfem_dot(int dtype, long n, const void * x, const void * y, double * out);
int nsfem_nrm2(int dtype, long n, const void * x, double * out);
/* nsfem_dot without allocating: no counterpart in the reference; for callers
 * that cannot allocate device memory (tight serving loops, harnesses that
 * forbid hipMalloc).  Same kernels and result as nsfem_dot, bit for bit; its
 * scratch (2 * 2048 + 2 doubles: one partial sum per workgroup, at most 2048
 * workgroups, so any n fits; n < 0 is rejected) is a fixed device array the
 * library holds on each GPU until the process ends.  Calls from several host
 * threads are serialised by a lock; blocking, like nsfem_dot. */
int nsfem_dot_noalloc(int dtype, long n, const void * x, const void * y, double * out);
int nsfem_scatter(int dtype, int nnz, const int * idx, const void * values, void * dense);
int nsfem_gather(int dtype, int nnz, const void * dense, const int * idx, void * values);
int nsfem_convert(int srcDtype, int dstDtype, long n, const void * src, void * dst);

/* ------------------------------------------------------------------------
 * Linear solvers.  Replace cu::createSolver (cpp/libs/CudaUtils/src/solvers/solverFactory.cu:209-237)
 * and cu::AbstractSolver (include/cu/solvers/abstractSolver.h:8-32).
 *   name "gs"        : multicolour Gauss-Seidel with the reference's semantics
 *                      (gaussSeidel.cu:121-404): RMS residual |b-Ax|/sqrt(n)
 *                      checked every mseCheckInterval sweeps, stop below targetMse.
 *   name "pcg"       : CG preconditioned by a symmetric multicolour GS sweep;
 *                      targetMse is a relative residual |b-Ax| / |b|.
 *   name "pcg_mixed" : FP64 PCG (vectors, products, reductions) with the
 *                      symmetric GS preconditioner in FP32 (values must be
 *                      NSFEM_F64); option "mixedRestart" switches to FP64
 *                      iterative refinement around FP32 inner PCG.
 *   name "bpcg"      : block PCG for a tile of numCh = 4, 8, 16 or 32
 *                      right-hand sides of one system: one Krylov space for
 *                      the tile with numCh x numCh step matrices (FP64 small
 *                      Cholesky solves that drop a direction whose pivot falls
 *                      below 1e-13 of the largest), the symmetric GS
 *                      preconditioner; stop when every column has
 *                      |b_c - A x_c| <= targetMse |b_c|.  FP64 values.
 *   name "bpcg_mixed": bpcg with the preconditioner in FP32.
 *                      No reference counterpart for either.
 *   name "amg_pcg"   : FP64 CG preconditioned by a smoothed-aggregation
 *                      V-cycle (host set-up: strength 0.08 halved per level,
 *                      greedy aggregation, Jacobi-smoothed prolongation,
 *                      Galerkin coarse operators down to <= 512 rows, inverted
 *                      densely) smoothed by one multicolour GS sweep per
 *                      direction on the finest level (forward down, backward
 *                      up) and two Chebyshev steps on D^-1 A on the coarser
 *                      ones (options "amgChebyshev*" below); one system,
 *                      numCh = 1, 2, 4, 8 or 16 right-hand sides (each with its
 *                      own scalars and stopping test); the stopping rule of
 *                      "pcg".  The matrix must be SPD, or positive
 *                      semidefinite with zero row sums (a pure-Neumann
 *                      Laplacian: the coarsest level is completed by a rank-one
 *                      term, right-hand sides must have zero mean; option
 *                      "projectMean" returns the zero-mean solution).
 *   name "amg_pcg_mixed": amg_pcg with the V-cycle in FP32.  No reference
 *                      counterpart (the reference's pressure solve is "gs" or
 *                      the cuDSS "dss").
 *   name "dss"       : rejected with NSFEM_EINVAL (no direct solver here).
 * The compute precision is valueDtype ("gs"/"pcg").  Colouring is the
 * reference's smallest-last greedy colouring (graphs.cpp:64-176).
 * rhs/sol are device buffers with stable addresses, numCh * rows elements of
 * the io dtype (NSFEM_F32 for an F32 solver, NSFEM_F64 otherwise).
 * numCh: 1 or 2 (the reference's solvers), or 4, 8, 16, 32 right-hand sides
 * sharing the matrix (no reference counterpart; one lane per channel).
 * ---------------------------------------------------------------------- */
typedef struct nsfem_solver nsfem_solver;
int nsfem_solver_create(const char * name, int numCh, int rows, int nnz, const int * rowStart, const int * column, const void * values,
                        int valueDtype, int maxIters, double targetMse, int mseCheckInterval, nsfem_solver ** out);
/* Block-diagonal batch of independent systems, each coloured on its own,
 * solved together with per-system convergence (values FP64 host arrays). */
int nsfem_solver_create_batch(const char * name, int numCh, int numSystems, const int * rows, const int * nnz,
                              const int * const * rowStart, const int * const * column, const double * const * values, int dtype,
                              int maxIters, double targetMse, int mseCheckInterval, nsfem_solver ** out);
int nsfem_solver_destroy(nsfem_solver * s);
int nsfem_solver_num_ch(const nsfem_solver * s, int * out);
int nsfem_solver_rows(const nsfem_solver * s, int * out);
int nsfem_solver_io_dtype(const nsfem_solver * s, int * out);
int nsfem_solver_rhs(nsfem_solver * s, void ** dptr);
int nsfem_solver_sol(nsfem_solver * s, void ** dptr);
int nsfem_solver_solve(nsfem_solver * s);
/* ch = -1: combined value, sqrt(m0^2 + m1^2) for two channels (gaussSeidel.cu:397). */
int nsfem_solver_last_mse(const nsfem_solver * s, int ch, double * out);
int nsfem_solver_last_relres(const nsfem_solver * s, int system, int ch, double * out);
int nsfem_solver_last_iterations(const nsfem_solver * s, int * out);
int nsfem_solver_last_iterations_system(const nsfem_solver * s, int system, int * out);
/* pcg_mixed with mixedRestart: refinement steps of the last solve; bpcg:
 * restart and verification cycles after the first (each restarts from the
 * true residual); 0 otherwise.  No reference counterpart. */
int nsfem_solver_last_outer_iterations(const nsfem_solver * s, int * out);
int nsfem_solver_set_max_iters(nsfem_solver * s, int n);
int nsfem_solver_set_target_mse(nsfem_solver * s, double target);
int nsfem_solver_set_mse_check_interval(nsfem_solver * s, int interval);
/* Options (no reference counterpart; the reference solvers have none):
 *   "projectMean"  1: pure-Neumann system with a zero-mean right-hand side;
 *                  the solution is returned with zero mean (and the FP32
 *                  inner right-hand sides of mixedRestart are projected).
 *   "mixedRestart" 1: pcg_mixed as iterative refinement (FP32 PCG inner
 *                  solves from zero on the FP64 residual) instead of the
 *                  default FP64 PCG with an FP32 symmetric GS preconditioner.
 *   "innerTol"     pcg_mixed with mixedRestart: inner relative tolerance.
 *   "graphs"       0: launch every iteration from the host instead of
 *                  replaying captured graphs of iteration chunks (default 1).
 *   "fusedReductions" pcg: 1 or -1 (default) lets the direction kernel
 *                  finish the previous iteration's rho and convergence test
 *                  and the update kernel the step length, every workgroup
 *                  from its system's partial sums, instead of two
 *                  one-workgroup finalize kernels per iteration; 0 keeps the
 *                  finalize kernels.
 *   "rowsDirection", "rowsUpdate", "rowsRho"
 *                  fused reductions on one system: rows per workgroup of
 *                  those kernels (0: the main shape's); autotune sets them.
 *   "lanesDirection" lanes per row of the direction kernel alone (0: the
 *                  main shape's; 1 and 2 channels); autotune sets it.
 *   "regionDirection" fused reductions on one system: 1 (default) orders
 *                  the direction kernel's workgroups region-major, so the
 *                  workgroups of one XCD take the same stretch (one eighth
 *                  of the rows' locality order) of every colour and gather
 *                  from one eighth of the vector; 0: rows in colour order.
 *   "fuseUpdateSweep" fused reductions on one system, 1-2 channels: 1
 *                  (default) runs the update kernel and the colour-1 forward
 *                  sweep as one launch (the sweep computes its colour-0
 *                  values from r and q; bit-identical results) while the
 *                  update workgroups hold at most two sweep slabs of rows
 *                  (rows per update workgroup x sweep lanes <= 512); 0: two
 *                  launches.
 *   "preconditioner" pcg / pcg_mixed (without mixedRestart; any
 *                  channel count, batches too): 0 (default) one symmetric
 *                  multicolour GS sweep, the reference's; 1 the factorized
 *                  sparse approximate inverse z = G^T G r (Kolotilina and
 *                  Yeremin), G lower triangular on each row's lower-colour
 *                  neighbours and itself, G_i* from M[P_i, P_i] g = e_i
 *                  scaled to (G M G^T)_ii = 1, built on the host when first
 *                  chosen: two parallel launches per application instead of
 *                  2 colours - 2 dependent sweeps (the P2 mass matrix: 12
 *                  iterations where GS takes 13).
 *   "rowsPreconditioner" fused reductions on one system, preconditioner
 *                  1: rows per workgroup of the approximate inverse's two
 *                  launches on a region-major map of their own (a positive
 *                  multiple of 256); 0 (default) shares the direction
 *                  kernel's map.
 *   "amgChebyshevFrom", "amgChebyshevDegree", "amgChebyshevFraction"
 *                  amg_pcg: the V-cycle levels from this one on (but the
 *                  dense coarsest) are smoothed by this many steps of the
 *                  Chebyshev iteration on D^-1 A over [fraction, 1] x the
 *                  level's lambda_max bound (1.1 x its power-iteration
 *                  estimate, at most the Gershgorin bound), the same
 *                  polynomial on the way down and up, instead of one
 *                  multicolour GS sweep (degree 2 to 8; 2 deg + 2 launches
 *                  per level instead of 2 colours + 3).  Defaults 1, 2, 0.2
 *                  (GS on the finest level only); a first level past the
 *                  last smoothing level keeps GS everywhere. */
int nsfem_solver_set_option(nsfem_solver * s, const char * key, double value);
/* Kernel timing (HIP events on the library stream; reps launches queued
 * behind a 20 ms delay kernel so they run back to back and the events
 * bracket GPU execution only), replaying the launches of one PCG iteration: "direction" (A z fused with the direction update and
 * p.q), "sweep" (the colour sweeps of one preconditioner application:
 * forward colours 1..C-1 on the lower part, from 2 with fuseUpdateSweep,
 * backward C-2..1), "rho" (backward colour 0 fused with r.z), "update" (x,
 * r, colour-0 z, r.r; with fuseUpdateSweep also the colour-1 sweep),
 * "finalize" (a one-block reduction finalize).  avgMs and bytes
 * (algorithmic, x read once) are per kernel launch, averaged over the
 * launches; launchesPerIteration (optional) is how many launches of that
 * kind one iteration makes.  pcg solvers only; overwrites the solver's
 * internal vectors.  No reference counterpart (measurement hook). */
int nsfem_solver_time_kernel(nsfem_solver * s, const char * which, int reps, double * avgMs, double * bytes,
                             double * launchesPerIteration);
int nsfem_solver_num_colors(const nsfem_solver * s, int system, int * out);
/* Launch shape of the solver kernels: lanes per matrix row (1, 2, 4, 8, 16; 0
 * picks from the mean row length) and rows per workgroup (0: default).
 * reshape rebuilds the workgroup map only (colouring and matrices kept) and
 * gives the colour sweeps the same shape; reshape_sweeps then sets the
 * sweeps' own (they write no partial sums).  autotune times one PCG
 * iteration replayed from a captured graph (pcg solvers, reps iterations)
 * for lanes {1, 2, 4, 8, 16} x {1, 2, 4, 8} row slabs, keeps the fastest, then
 * tries the sweeps on lanes x {1, 2, 4} slabs of their own, then (fused
 * reductions) the direction / update / rho rows and the direction lanes, and
 * last, unless fusedReductions was set, the finalize kernels against the
 * fused reductions; it overwrites the internal work vectors, not rhs/sol.
 * Solvers of 4 to 32 channels run one
 * lane per channel (lanes = numCh) and keep at most about 16k workgroups.
 * No reference counterpart (the reference's launch shapes are fixed). */
int nsfem_solver_reshape(nsfem_solver * s, int lanes, int rowsPerBlock);
int nsfem_solver_shape(const nsfem_solver * s, int * lanes, int * rowsPerBlock);
int nsfem_solver_reshape_sweeps(nsfem_solver * s, int lanes, int rowsPerBlock);
int nsfem_solver_sweep_shape(const nsfem_solver * s, int * lanes, int * rowsPerBlock);
/* Rows per workgroup of the direction, update and rho kernels (the main
 * shape's unless fused reductions give them their own). */
int nsfem_solver_kind_rows(const nsfem_solver * s, int * direction, int * update, int * rho);
int nsfem_solver_autotune(nsfem_solver * s, int reps, int * lanes, int * rowsPerBlock);
int nsfem_solver_coloring(const nsfem_solver * s, int system, int * coloring, int * partitionStart);

/* Smallest-last greedy colouring of a square CSR pattern, host only
 * (buildSmallestLastOrdering + partitionGraphGreedy + per-colour sort,
 * graphs.cpp:64-176, gaussSeidel.cu:140-159).  partitionStart needs rows + 1 slots. */
int nsfem_mcgs_coloring(int rows, const int * rowStart, const int * column, int * coloring, int * partitionStart, int * numColors);

/* ------------------------------------------------------------------------
 * Meshes.  Replace mesh::parseTriangleGmsh (cpp/libs/Mesh/src/io.cpp:65-214)
 * and mesh::createMesh (src/concreteMesh.cpp:78-293).
 * ---------------------------------------------------------------------- */
typedef struct nsfem_mesh nsfem_mesh;
int nsfem_mesh_load_gmsh(const char * path, nsfem_mesh ** out);
/* borderEdges: numBorderEdges x (from, to, group), oriented like gmsh line elements. */
int nsfem_mesh_from_arrays(int numNodes, const double * xy, int numTriangles, const int * triangles, int numBorderEdges,
                           const int * borderEdges, int numGroups, const char * const * groupNames, nsfem_mesh ** out);
int nsfem_mesh_destroy(nsfem_mesh * m);
int nsfem_mesh_info(const nsfem_mesh * m, int * numNodes, int * numTriangles, int * numBorderElements, int * numGroups);
/* xy: 2*numNodes, triangles: 3*numTriangles, border: 3*numBorderElements (triangle, side, group). */
int nsfem_mesh_arrays(const nsfem_mesh * m, double * xy, int * triangles, int * border);
int nsfem_mesh_group_name(const nsfem_mesh * m, int group, char * buf, int bufSize);

/* ------------------------------------------------------------------------
 * Chorin context: Dirichlet lists and all FP64 matrices of one mesh.
 * Replaces buildChorinContext (cpp/execs/NavierStokes/src/buildContext.cpp:17-142)
 * with fem::buildChorinCsrMatrices (cpp/libs/Fem/src/chorinCsr.cpp:437-502).
 * ---------------------------------------------------------------------- */
typedef struct nsfem_dfg_conditions
{
    double viscosity;
    double peakVelocity;
    double channelHeight; /* 0.41 in the reference (dfgCondtions.h:9-12) */
    double profileScale;  /* 1 in the reference; inflow = scale*peak*y(H-y)/H^2 */
} nsfem_dfg_conditions;

typedef struct nsfem_context nsfem_context;
int nsfem_context_build(const nsfem_mesh * mesh, const nsfem_dfg_conditions * cond, int numThreads, nsfem_context ** out);
/* flags: NSFEM_CONTEXT_GPU_ASSEMBLY computes the matrix values on the GPU
 * (patterns and per-entry slot lists on the host): the values part of
 * buildChorinCsrMatrices (cpp/libs/Fem/src/chorinCsr.cpp:296-502), bit-identical
 * to the host assembly. */
#define NSFEM_CONTEXT_GPU_ASSEMBLY 1
int nsfem_context_build_ex(const nsfem_mesh * mesh, const nsfem_dfg_conditions * cond, int numThreads, int flags,
                           nsfem_context ** out);
/* A batch of independent instances (different meshes and conditions) as one
 * block-diagonal context: velocity [ux of all; uy of all], pressure and
 * internal-pressure vectors concatenated in instance order; the Chorin solvers
 * keep one system per instance (own colouring, own convergence).  No
 * counterpart in the reference (one mesh per run there). */
int nsfem_context_build_batch(const nsfem_mesh * const * meshes, const nsfem_dfg_conditions * conds, int count, int numThreads,
                              int flags, nsfem_context ** out);
int nsfem_context_num_systems(const nsfem_context * c, int * out);
/* Per instance: offset into one velocity component and into the pressure vector. */
int nsfem_context_system_offsets(const nsfem_context * c, int * velocityOffsets, int * pressureOffsets);
int nsfem_context_destroy(nsfem_context * c);
int nsfem_context_sizes(const nsfem_context * c, int * numVelocityNodes, int * numPressureNodes, int * numDirichletVelocity,
                        int * numInternalPressure, int * numElements);

#define NSFEM_MAT_VELOCITY_MASS 0
#define NSFEM_MAT_VELOCITY_STIFFNESS 1
#define NSFEM_MAT_PRESSURE_STIFFNESS 2
#define NSFEM_MAT_PRESSURE_STIFFNESS_INTERNAL 3
#define NSFEM_MAT_VELOCITY_PRESSURE_DIV 4
#define NSFEM_MAT_PRESSURE_VELOCITY_DIV 5
/* Borrowed host pointers, valid while the context lives. */
int nsfem_context_matrix(const nsfem_context * c, int which, int * rows, int * cols, int * nnz, const int ** rowStart,
                         const int ** column, const double ** values);
/* which: 0 = velocity x, 1 = velocity y, 2 = pressure; arrays of the matching count. */
int nsfem_context_dirichlet(const nsfem_context * c, int which, int * count, int * ids, double * values);
int nsfem_context_internal_pressure(const nsfem_context * c, int * ids);
/* P2 (velocity) or P1 (pressure) numbering: xy 2*nodes, elements 6*E or 3*E. */
int nsfem_context_velocity_mesh(const nsfem_context * c, double * xy, int * elements);
int nsfem_context_pressure_mesh(const nsfem_context * c, double * xy, int * elements);
/* FastConvection integration map (cpp/libs/Fem/src/fastConvection.cpp:9-100),
 * FP64, rows = nnz of the velocity pattern, cols = 2N.  Arrays are malloc'd;
 * release with nsfem_host_free. */
int nsfem_fast_convection_map(const nsfem_context * c, int numThreads, int * rows, int * cols, int * nnz, int ** rowStart, int ** column,
                              double ** values);
void nsfem_host_free(void * p);

/* ------------------------------------------------------------------------
 * Chorin projection time stepping on the GPU.  Replaces the loop of
 * solveNsChorinCuda (cpp/execs/NavierStokes/src/chorinCuda.cu:154-379) and
 * its config (include/NavierStokes/chorinCudaConfig.h:6-12).
 * ---------------------------------------------------------------------- */
typedef struct nsfem_solver_config
{
    const char * method; /* "gs", "pcg", "pcg_mixed" */
    int maxIterations;
    double targetMse;
    int mseCheckInterval;
    int lanes;        /* lanes per matrix row: 1, 2, 4, 8, 16 (0: from the mean row length) */
    int rowsPerBlock; /* rows per workgroup of the solver kernels (0: default) */
    int sweepLanes;   /* the colour sweeps' own shape (0, 0: as lanes, rowsPerBlock) */
    int sweepRowsPerBlock;
} nsfem_solver_config;

typedef struct nsfem_chorin_config
{
    nsfem_solver_config velocitySolver;
    nsfem_solver_config pressureSolver;
    int dtype;       /* NSFEM_F64 (plain / mixed) or NSFEM_F32 (reference precision) */
    double innerTol; /* pcg_mixed with mixedRestart: inner relative tolerance */
    int eagerLaunches; /* 1: launch every solver iteration from the host instead
                        * of replaying captured graphs (for tools that cannot
                        * trace graph launches); 0: graphs */
    int timeStepPolicy; /* NSFEM_TAU_REJECT (0, default), NSFEM_TAU_CLAMP or
                         * NSFEM_TAU_UNCHECKED; see nsfem_chorin_set_time_step */
    int stepGraph;      /* 1: each step (after a first regular one) is one captured
                         * graph: a device snapshot of the state, convection, the
                         * three solves with iteration budgets from the last
                         * counts (their kernels idle once converged, so
                         * convergence is decided on the device), the updates and
                         * impositions; a solve out of budget rolls the step back
                         * to the snapshot and reruns it regularly.  pcg solvers
                         * without mixedRestart, or amg_pcg / amg_pcg_mixed for the
                         * pressure.  No reference counterpart. */
    int pressureStart;  /* the pressure solve's initial guess: NSFEM_PRESSURE_START_LAST
                         * (0, the last pressure, as the reference's in-place
                         * solve) or NSFEM_PRESSURE_START_EXTRAPOLATE (2 p_n - p_n-1;
                         * p_n on the first step after a reset) */
    int velocityStart;  /* the velocity solves' initial guess, each of the two solves
                         * from its own history: NSFEM_VELOCITY_START_ZERO (0, as
                         * the reference's accel.memsetZero(), chorinCuda.cu:265,
                         * 343), NSFEM_VELOCITY_START_LAST (its solution one step
                         * before) or NSFEM_VELOCITY_START_EXTRAPOLATE (2 a_n - a_n-1);
                         * zero after a reset */
} nsfem_chorin_config;

#define NSFEM_PRESSURE_START_LAST 0
#define NSFEM_PRESSURE_START_EXTRAPOLATE 1
#define NSFEM_VELOCITY_START_ZERO 0
#define NSFEM_VELOCITY_START_LAST 1
#define NSFEM_VELOCITY_START_EXTRAPOLATE 2

/* Time-step policies.  The step is explicit (forward Euler) in diffusion and
 * convection (chorinCuda.cu:263-278) and stays bounded only while
 * tau nu lambda_max(M^-1 K) <= 2; the limit enforced is half of that,
 * tau_max = 1 / lambda_max(M^-1 nu K), estimated once per Chorin object by
 * power iteration on the GPU (the largest over the instances of a batch).
 * The reference checks nothing (NSFEM_TAU_UNCHECKED). */
#define NSFEM_TAU_REJECT 0
#define NSFEM_TAU_CLAMP 1
#define NSFEM_TAU_UNCHECKED 2

typedef struct nsfem_step_stats
{
    double mseTentX, mseTentY, msePressure, mseFinalX, mseFinalY;
    double relTent, relPressure, relFinal;
    double msConvection, msTentative, msPressure, msFinal, msTotal;
    int itersTent, itersPressure, itersFinal;
    int outerTent, outerPressure, outerFinal;
    double tauRequested, tauUsed; /* the step asked for and the step taken */
    int capturedStep;             /* 1: ran as one captured graph (phase times 0) */
} nsfem_step_stats;

typedef struct nsfem_chorin nsfem_chorin;
int nsfem_chorin_create(const nsfem_context * c, const nsfem_chorin_config * cfg, nsfem_chorin ** out);
/* Ensemble: `instances` (1, 2, 4, 8 or 16) independent flows on the one mesh
 * of a single-mesh context, instance b's Dirichlet velocity (the inflow
 * profile) scaled by inflowScale[b] (NULL: all 1).  They share every matrix:
 * each velocity solve is one solve of 2 x instances right-hand sides, the
 * pressure solve one of `instances` (amg_pcg_mixed, or a block solver such
 * as bpcg_mixed for 4 or more); the convection is each instance's own.  The state functions
 * then hold the instances one after another: velocity [u_0 .. u_B-1] (2N
 * each), pressure [p_0 .. p_B-1] (V each); the step statistics are the worst
 * over the instances.  Parameter sweeps over one geometry (inflow speeds,
 * perturbed starts) in place of one chorinCuda.cu run per case.  No reference
 * counterpart. */
int nsfem_chorin_create_ensemble(const nsfem_context * c, const nsfem_chorin_config * cfg, int instances, const double * inflowScale,
                                 nsfem_chorin ** out);
int nsfem_chorin_instances(const nsfem_chorin * ch, int * instances);
int nsfem_chorin_destroy(nsfem_chorin * ch);
/* Replaces the fixed tau = maxT / ceil(maxT / tau0) of chorinCuda.cu:221-222.
 * Under NSFEM_TAU_REJECT a tau above tau_max fails with NSFEM_EINVAL (the
 * state is kept; the previous step stays in force); NSFEM_TAU_CLAMP uses
 * tau_max instead; NSFEM_TAU_UNCHECKED takes any tau > 0.  A step taken
 * without a set_time_step call uses 1e-3 under the same policy. */
int nsfem_chorin_set_time_step(nsfem_chorin * ch, double tau);
/* The last requested and the used time step, lambda_max(M^-1 nu K) and
 * tau_max = 1 / lambda_max (any pointer may be NULL; asking for lambdaNu or
 * tauMax runs the estimate if it has not run).  No reference counterpart. */
int nsfem_chorin_time_step_info(nsfem_chorin * ch, double * tauRequested, double * tauUsed, double * lambdaNu, double * tauMax);
/* Captured-step bookkeeping: graphs built (and the host seconds spent
 * capturing and instantiating them) and steps rolled back so far, and the
 * current iteration budgets (tentative, pressure, final; 0 before the first
 * step).  Any pointer may be NULL.  No reference counterpart. */
int nsfem_chorin_step_graph_info(const nsfem_chorin * ch, int * builds, double * buildSeconds, int * rollbacks, int * budgets);
/* Sets the iteration budgets (tentative, pressure, final; each > 0) of the
 * next captured step, e.g. from counts the caller knows; later steps adapt
 * them from their own counts again.  No reference counterpart. */
int nsfem_chorin_set_step_graph_budgets(nsfem_chorin * ch, int tentative, int pressure, int final);
/* One projection step; stats may be NULL.  NSFEM_ENONFINITE on a bad
 * velocity solve ("Tentative/Final acceleration is bad", chorinCuda.cu:270-273, 349-352). */
int nsfem_chorin_step(nsfem_chorin * ch, nsfem_step_stats * stats);
/* Device state: velocity 2N ([ux; uy]), pressure V, in the chorin dtype
 * (an ensemble: B times each, instance after instance). */
int nsfem_chorin_state(nsfem_chorin * ch, void ** velocity, void ** pressure);
int nsfem_chorin_download(nsfem_chorin * ch, void * velocity, void * pressure);
/* Output path, replacing the two blocking per-step downloads
 * (chorinCuda.cu:327, 363) for frame-strided output: download_async
 * snapshots (u, p) after the last step on the device and copies the
 * snapshot to library-owned pinned host memory on a stream of its own while
 * later steps run; download_wait waits for that copy and writes it to the
 * caller's buffers (either may be NULL).  One download at a time
 * (NSFEM_EINVAL otherwise). */
int nsfem_chorin_download_async(nsfem_chorin * ch);
int nsfem_chorin_download_wait(nsfem_chorin * ch, void * velocity, void * pressure);
int nsfem_chorin_upload_velocity(nsfem_chorin * ch, const void * velocity);
/* u = 0 + Dirichlet values, p = 0, pressure warm start cleared. */
int nsfem_chorin_reset(nsfem_chorin * ch);
/* Launch shapes of the step's solvers, shape[18] = {velocity lanes, velocity
 * rows per workgroup, pressure lanes, pressure rows per workgroup, the same
 * four for the colour sweeps, velocity direction / update / rho rows,
 * pressure direction / update / rho rows, the velocity and pressure
 * direction kernels' lanes, then whether each solver fuses its reductions};
 * autotune as nsfem_solver_autotune for each pcg / pcg_mixed solver of
 * the step (other methods keep their shapes; the state u, p is kept). */
int nsfem_chorin_shape(const nsfem_chorin * ch, int * shape);
int nsfem_chorin_autotune(nsfem_chorin * ch, int reps, int * shape);
/* nsfem_solver_set_option for the step's solvers: which = 0 velocity,
 * 1 pressure, 2 both.  No reference counterpart. */
int nsfem_chorin_set_solver_option(nsfem_chorin * ch, int which, const char * key, double value);
/* Kernel timing inside a Chorin context: "convection", "spmm",
 * "velocity_<k>", "pressure_<k>" (k as in nsfem_solver_time_kernel). */
int nsfem_chorin_time_kernel(nsfem_chorin * ch, const char * which, int reps, double * avgMs, double * bytes,
                             double * launchesPerIteration);
/* A = nu K + C(u) in the velocity pattern (C optional): device arrays. */
int nsfem_chorin_assemble_convection(nsfem_chorin * ch, const void * velocity, void * aValues, void * cValues);

#ifdef __cplusplus
}
#endif

#endif /* NSFEM_HIP_H */
