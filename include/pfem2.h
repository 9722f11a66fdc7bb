/*
 * pfem2.h -- C ABI of libpfem2_hip.so, the MI355X (gfx950) PFEM-2 hot path.
 *
 * Every entry point returns 0 on success or a negative PFEM_E* code; pfem_last_error() gives the
 * message. Nothing in the library calls exit(). All device work is queued on the context's stream;
 * the calls that hand a value back to the host (counts, solver statistics, downloads) synchronise
 * that stream. Pointers named d_* are device pointers; every other pointer is host memory that is
 * only borrowed for the duration of the call.
 *
 * Each declaration names the reference interface it replaces (paths relative to the reference
 * repository andreyypopov/gpuPfem2).
 */
#ifndef PFEM2_H
#define PFEM2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFEM_OK 0
#define PFEM_EINVAL (-1)   /* bad argument */
#define PFEM_EHIP (-2)     /* HIP runtime error */
#define PFEM_EIO (-3)      /* file could not be opened or parsed */
#define PFEM_ENUMERIC (-4) /* zero/negative pivot, NaN */
#define PFEM_ECAPACITY (-5)/* a fixed-size buffer would overflow */

typedef struct pfem_context* pfem_ctx;
typedef struct pfem_mesh2d_t* pfem_mesh2d;
typedef struct pfem_mesh3d_t* pfem_mesh3d;
typedef struct pfem_csr_t* pfem_csr;
typedef struct pfem_dirichlet_t* pfem_dirichlet;
typedef struct pfem_precond_t* pfem_precond;
typedef struct pfem_solver_t* pfem_solver;
typedef struct pfem_particles_t* pfem_particles;
typedef struct pfem_pfem2_t* pfem_pfem2;

/* ------------------------------------------------------------------ context and device memory */
/* replaces the implicit CUDA context + legacy default stream (src/common/cuda_memory.cuh:17-148).
 * A context, and every handle made on it, is used from one host thread at a time: its calls are ordered
 * on its one stream and share its scratch buffers (e.g. the reduction partials of pfem_dot / pfem_nrm2).
 * Work from several threads at once takes one context per thread (contexts on one device run
 * concurrently). */
int pfem_ctx_create(int device, pfem_ctx* out);
int pfem_ctx_destroy(pfem_ctx ctx);
const char* pfem_last_error(pfem_ctx ctx);
int pfem_ctx_synchronize(pfem_ctx ctx);
/* the context stream, as a void* (hipStream_t) for callers that order their own work on it */
void* pfem_ctx_stream(pfem_ctx ctx);
/* deviceVector<T> substrate: allocate_device / free_device / copy_h2d / copy_d2h / copy_d2d /
 * zero_value_device (src/common/cuda_memory.cuh:17-148, src/common/device_vector.cuh:19-146) */
int pfem_malloc(pfem_ctx ctx, size_t bytes, void** d_out);
int pfem_free(pfem_ctx ctx, void* d_ptr);
int pfem_memcpy_h2d(pfem_ctx ctx, void* d_dst, const void* src, size_t bytes);
int pfem_memcpy_d2h(pfem_ctx ctx, void* dst, const void* d_src, size_t bytes);
int pfem_memcpy_d2d(pfem_ctx ctx, void* d_dst, const void* d_src, size_t bytes);
int pfem_memset_zero(pfem_ctx ctx, void* d_ptr, size_t bytes);

/* Node and cell order for memory locality, host only (no counterpart in the reference, which keeps the
 * DAT file's numbering, src/mesh_2d.cu:36-96 and src/mesh_3d.cu:84-147): for nt cells of nvc = 3
 * (triangles) or 4 (tetrahedra) vertices, node_order[new] = old in reverse Cuthill-McKee order of the
 * vertex adjacency graph, and cell_order[new] = old with the cells sorted by the smallest new index of
 * their vertices (ties: the next smallest, then the old index). A mesh renumbered by these orders gives
 * the same solutions, permuted; its rows, cells and cell-sorted particles lie close in memory. */
int pfem_mesh_locality_order(int nv, int nt, int nvc, const int* cells, int* node_order, int* cell_order);

/* ------------------------------------------------------------------ 2D mesh (src/mesh_2d.cuh:11-69) */
/* Mesh2D::loadMeshFromFile(filename, fillNeighborLists, scale) (src/mesh_2d.cu:36-96), including the
 * reader's quirk that a trailing newline drops the last cell from the device mesh. */
int pfem_mesh2d_load_dat(pfem_ctx ctx, const char* path, int fill_neighbors, double scale, pfem_mesh2d* out);
/* Mesh2D from host arrays (vertices xy[2*nv], triangles tri[3*nt], 0-based) -- generated meshes */
int pfem_mesh2d_create(pfem_ctx ctx, int nv, const double* xy, int nt, const int* tri, int fill_neighbors,
                       pfem_mesh2d* out);
int pfem_mesh2d_destroy(pfem_mesh2d m);
/* sizes: vertices, device cells, host cells, neighbour entries */
int pfem_mesh2d_info(pfem_mesh2d m, int* nv, int* nt, int* nt_host, long* n_neighbors);
/* getters of src/mesh_2d.cuh:18-53, as host copies. what: 0 vertices (2 dbl/vtx), 1 cells (3 int),
 * 2 cell area (1 dbl), 3 inverse Jacobian (4 dbl, row-major), 4 neighbour offsets (nt+1 int),
 * 5 neighbour indices (int), 6 edge boundary IDs (3 int per cell), 7 host cells (3 int per host cell) */
int pfem_mesh2d_download(pfem_mesh2d m, int what, void* dst);
/* the same arrays in device memory (what 0..6), owned by the mesh: the deviceVector members behind the
 * getters (getVertices().data, getCells().data, ...) */
int pfem_mesh2d_device_ptr(pfem_mesh2d m, int what, void** d_ptr);
/* kSetEdgeBoundaryIDs (cases/Cylinder2D/main.cu:35-70, cases/PoiseuilleFlow2D/main.cu:48-79), with the
 * geometry as parameters: edge midpoints within 2e-6 of x0 / x1 / y0 / y1 get id_x0 / id_x1 / id_y0 /
 * id_y1 (checked in that order); when circle_r > 0, midpoints with | |m - c| - r | < circle_tol get
 * id_circle. Other edges get -1. */
int pfem_mesh2d_classify_edges(pfem_mesh2d m, double x0, double x1, double y0, double y1, int id_x0, int id_x1,
                               int id_y0, int id_y1, double circle_cx, double circle_cy, double circle_r,
                               double circle_tol, int id_circle);
/* explicit edge IDs (3 per device cell) */
int pfem_mesh2d_set_edge_ids(pfem_mesh2d m, const int* ids);

/* ------------------------------------------------------------------ 3D mesh (src/mesh_3d.cuh) */
/* Mesh3D::loadMeshFromFile (src/mesh_3d.cu:84-147) without the neighbour lists (pfem_mesh3d_fill_neighbors
 * builds them). zero_small_det = 1 keeps the reference's zero inverse Jacobian for |det| < 2e-6
 * (src/common/matrix3x3.cuh:139-140); 0 keeps the true inverse. */
int pfem_mesh3d_load_dat(pfem_ctx ctx, const char* path, double scale, int zero_small_det, pfem_mesh3d* out);
int pfem_mesh3d_create(pfem_ctx ctx, int nv, const double* xyz, int nt, const int* tet, int zero_small_det,
                       pfem_mesh3d* out);
int pfem_mesh3d_destroy(pfem_mesh3d m);
int pfem_mesh3d_info(pfem_mesh3d m, int* nv, int* nt);
/* Mesh3D::fillCellNeighborIndices (src/mesh_3d.cu:37-82, 198-227; run by loadMeshFromFile when
 * fillNeighborLists is set): per cell, the cells sharing a face (exactly 3 equal vertex pairs),
 * ascending, as offsets (nt + 1) and indices. Built on the GPU from the vertex -> cell incidence in O(T)
 * instead of the reference's all-pairs O(T^2) test; a second call returns the lists already built.
 * n_neighbors (may be null) receives the total. */
int pfem_mesh3d_fill_neighbors(pfem_mesh3d m, long* n_neighbors);
/* what: 0 vertices (3 dbl), 1 cells (4 int), 2 volume (1 dbl), 3 inverse Jacobian (9 dbl, row-major),
 * 4 face-neighbour offsets (nt + 1 int), 5 face-neighbour indices (int; after pfem_mesh3d_fill_neighbors) */
int pfem_mesh3d_download(pfem_mesh3d m, int what, void* dst);
/* the same arrays in device memory (what 0..5; 4 and 5 null until the neighbours are built), owned by the
 * mesh: the deviceVector members behind Mesh3D's getters (getVertices().data, getCellNeighborIndices().data) */
int pfem_mesh3d_device_ptr(pfem_mesh3d m, int what, void** d_ptr);

/* ------------------------------------------------------------------ CSR (src/linear_algebra/sparse_matrix.cuh) */
/* SparseMatrixCSR(const Mesh2D&) / (const Mesh3D&) (sparse_matrix.cu:8-99): pattern {i} U {vertices
 * sharing a cell with i}, columns ascending, values zero. Built in O(nnz) instead of std::set. */
int pfem_csr_create_from_mesh2d(pfem_mesh2d m, pfem_csr* out);
int pfem_csr_create_from_mesh3d(pfem_mesh3d m, pfem_csr* out);
/* a matrix from host CSR arrays (values may be NULL: zero) */
int pfem_csr_create(pfem_ctx ctx, int n, int nnz, const int* rowptr, const int* col, const double* val,
                    pfem_csr* out);
/* a second value array on the same (shared) pattern */
int pfem_csr_create_like(pfem_csr a, pfem_csr* out);
int pfem_csr_destroy(pfem_csr a);
int pfem_csr_info(pfem_csr a, int* n, int* nnz);
/* getRowOffset / getColIndices / getMatrixValues (sparse_matrix.cuh:24-34): device pointers */
int pfem_csr_pointers(pfem_csr a, int** d_rowptr, int** d_col, double** d_val);
/* clearValues (sparse_matrix.cuh:46-48) */
int pfem_csr_clear_values(pfem_csr a);
int pfem_csr_upload_values(pfem_csr a, const double* val);
int pfem_csr_download(pfem_csr a, int* rowptr, int* col, double* val);
/* y = A x: LinearAlgebra::sparseMV (src/linear_algebra/linear_algebra.h:75-80, cusparseSpMV CSR, FP64).
 * Per-row sums run left to right, so results are deterministic. */
int pfem_spmv(pfem_csr a, const double* d_x, double* d_y);
/* FP32 copy of the values (mixed-precision inner solves) and its product y32 = A32 x32 */
int pfem_csr_refresh_fp32(pfem_csr a);
int pfem_spmv_fp32(pfem_csr a, const float* d_x, float* d_y);

/* Dense BLAS-1/2 of LinearAlgebra (src/linear_algebra/linear_algebra.h:27-54), FP64, on the context's stream;
 * vectors, matrices and the scalars res, alpha, beta are device pointers (the reference's cuBLAS pointer
 * mode is DEVICE). Reductions sum in a fixed order, so results are deterministic. */
/* *d_res = x . y: LinearAlgebra::dot (linear_algebra.h:35-38, cublasDdot); n = 0 gives 0 */
int pfem_dot(pfem_ctx ctx, int n, const double* d_x, const double* d_y, double* d_res);
/* *d_res = ||x||_2 (not squared, whatever the name): LinearAlgebra::normSquared (linear_algebra.h:40-43,
 * cublasDnrm2). Vectors whose squares overflow or underflow are summed again scaled by max |x_i|. */
int pfem_nrm2(pfem_ctx ctx, int n, const double* d_x, double* d_res);
/* x <- U^-1 x, U the upper triangle (non-unit diagonal) of the column-major n x n matrix at d_a with leading
 * dimension lda >= n: LinearAlgebra::solveGeneralTriangularSystem (linear_algebra.h:45-49, cublasDtrsv).
 * One workgroup, for small systems (the GMRES Hessenberg triangle). */
int pfem_trsv_upper(pfem_ctx ctx, int n, const double* d_a, int lda, double* d_x);
/* y = alpha A x + beta y, A m x n column-major with lda >= m: LinearAlgebra::generalMV
 * (linear_algebra.h:51-54, cublasDgemv, no transpose). As BLAS: m = 0 or n = 0 leaves y unchanged, beta = 0
 * does not read y. */
int pfem_gemv(pfem_ctx ctx, int m, int n, const double* d_alpha, const double* d_a, int lda, const double* d_x,
              const double* d_beta, double* d_y);

/* ------------------------------------------------------------------ Dirichlet BCs (src/Dirichlet_bcs.cuh) */
#define PFEM_DIRICHLET_PARITY 0  /* reference heuristic: skip row c when |first off-diagonal| < 2e-6 */
#define PFEM_DIRICHLET_MASK 1    /* skip exactly the Dirichlet rows (symmetric result) */
/* DirichletBCs::setupDirichletBCs + setupNodeMap (Dirichlet_bcs.cu:53-71). n = matrix rows. */
int pfem_dirichlet_create(pfem_ctx ctx, int n, int nbc, const int* node, const double* value, pfem_dirichlet* out);
int pfem_dirichlet_destroy(pfem_dirichlet bc);
int pfem_dirichlet_set_values(pfem_dirichlet bc, const double* value);            /* host values */
int pfem_dirichlet_values_device(pfem_dirichlet bc, double** d_value);            /* device value array */
/* DirichletBCs::applyBCs(matrix, rhs) (Dirichlet_bcs.cu:73-86) */
int pfem_dirichlet_apply(pfem_dirichlet bc, pfem_csr a, double* d_rhs, int mode);
/* two right-hand sides with their own values on one matrix (both velocity components share their BC
 * node set, cases/Cylinder2D/main.cu:675-689): same result as applying bc0 and bc1 to two copies */
int pfem_dirichlet_apply2(pfem_dirichlet bc0, pfem_dirichlet bc1, pfem_csr a, double* d_rhs0, double* d_rhs1,
                          int mode);

/* ------------------------------------------------------------------ preconditioners (preconditioners.cuh) */
#define PFEM_PRECOND_NONE 0
#define PFEM_PRECOND_JACOBI 1
#define PFEM_PRECOND_IC0 2
#define PFEM_PRECOND_ILU0 3
#define PFEM_PRECOND_CUSTOM 4   /* host callbacks (pfem_precond_create_custom) */
/* ordering for IC0/ILU0: natural (exact reference factor, level-scheduled), multicolour (the
 * factor of the symmetrically permuted matrix; few levels) or blocks (blocks of up to 8 neighbouring
 * nodes, coloured as blocks: the factor keeps the couplings inside each block, which takes fewer Krylov
 * iterations than multicolour; each block is solved by one lane). No counterpart in the reference,
 * whose factor is the natural-order one. */
#define PFEM_ORDER_NATURAL 0
#define PFEM_ORDER_MULTICOLOR 1
#define PFEM_ORDER_BLOCKS 2
/* PreconditionerJacobi(n, LA) / PreconditionerIC(matrix, LA) / PreconditionerILU(matrix, LA)
 * (preconditioners.cuh:9-90); the pattern of `a` fixes the structure */
int pfem_precond_create(pfem_csr a, int kind, int ordering, pfem_precond* out);
/* The Preconditioner plugin point (preconditioners.cuh:9-26): a preconditioner defined by the caller.
 * init(user, a) runs where Preconditioner::initialize would (the start of every solve,
 * linear_solver.cu:306-307), apply(user, dest, src) where applyPreconditioner would; dest may equal src.
 * Both run on the host and must enqueue their device work on the context stream (pfem_ctx_stream). FP64
 * solves only. */
typedef void (*pfem_precond_init_fn)(void* user, pfem_csr a);
typedef void (*pfem_precond_apply_fn)(void* user, double* d_dest, const double* d_src);
int pfem_precond_create_custom(pfem_ctx ctx, int n, pfem_precond_init_fn init, pfem_precond_apply_fn apply,
                               void* user, pfem_precond* out);
int pfem_precond_destroy(pfem_precond pc);
/* Preconditioner::initialize(const SparseMatrixCSR&) */
int pfem_precond_initialize(pfem_precond pc, pfem_csr a);
/* Preconditioner::applyPreconditioner(dest, src); dest may equal src */
int pfem_precond_apply(pfem_precond pc, double* d_dest, const double* d_src);
/* factor values (nnz doubles, in the matrix's own CSR layout and row order), levels, permutation */
int pfem_precond_download(pfem_precond pc, double* factor, int* n_levels, int* perm);
/* IC0/ILU0 apply as one launch per triangular solve (1, default) or one launch per level (0); both give
 * bit-identical results. Replaces the analysis + cusparseSpSV_solve pair of applyPreconditioner
 * (preconditioners.cu:116-133, :195-212). */
int pfem_precond_set_sweep(pfem_precond pc, int single_pass);
/* Polls a single-pass wait makes before it is declared failed (no counterpart in the reference; default
 * 65536, about 0.1 s). A failed wait does not hang or fault: the launch completes, and the solve or apply
 * reports PFEM_ENUMERIC and re-arms the sweep buffers. 0 fails every wait that is not satisfied at once. */
int pfem_precond_set_spin_limit(pfem_precond pc, int polls);
/* Storage of the reduced-precision factor copy that PFEM_PRECISION_MIXED_PRECOND applies to FP64 vectors
 * (no counterpart in the reference, whose factor is FP64): 32 (default) or 16 bits per value (IEEE half;
 * multicolour factors with the single-pass sweeps only). Call before the solver's precision is set. */
int pfem_precond_set_factor_bits(pfem_precond pc, int bits);

/* ------------------------------------------------------------------ Krylov solvers (linear_solver.cuh) */
#define PFEM_SOLVER_CG 0
#define PFEM_SOLVER_GMRES 1
#define PFEM_TOL_ABSOLUTE 0  /* reference: CG (r, M^-1 r) < tol^2; GMRES |beta| < tol */
#define PFEM_TOL_RELATIVE 1  /* CG (r, M^-1 r) < tol^2 (b, M^-1 b); GMRES |beta| < tol ||M^-1 b|| */
#define PFEM_PRECISION_FP64 0
#define PFEM_PRECISION_MIXED 1 /* FP32 inner Krylov + FP64 iterative refinement */
/* FP64 Krylov whose IC0/ILU0 factor is stored in FP32 (the triangular solves compute in FP64). The
 * refinement of MIXED contracts by about cond(A) * 2^-24 per pass and stalls on the pressure Poisson
 * matrices, whose condition number is of that order. */
#define PFEM_PRECISION_MIXED_PRECOND 2
typedef struct {
    int converged;
    int iterations;        /* Krylov iterations (for MIXED: inner iterations summed) */
    double final_residual; /* CG: sqrt((r, M^-1 r)); GMRES: |beta| */
    int nan_detected;
    int refinements;       /* MIXED: outer FP64 refinement steps */
} pfem_solve_stats;
/* SolverCG(tol, maxIt, LA, precond) / SolverGMRES(...) (linear_solver.cuh:60-137) + init(matrix).
 * restart: GMRES(m) restart length; restart >= max_iterations reproduces the reference's
 * non-restarted GMRES (x untouched unless converged). precond may be NULL. */
int pfem_solver_create(pfem_csr a, int kind, double tol, int max_iterations, int restart, int tol_mode,
                       pfem_precond precond, pfem_solver* out);
int pfem_solver_set_precision(pfem_solver s, int precision, double inner_tol, int max_refinements);
/* Solver options (no counterpart in the reference). PFEM_SOLVER_OPT_FACTOR_ORDER (default 1): with a
 * multicolour IC0/ILU0, run the FP64 Krylov solve on P A P^T, P b, P x (the factor's ordering; results
 * differ from the natural-order solve only by reduction order). PFEM_SOLVER_OPT_POLL (default 8 for CG,
 * 4 for GMRES): iterations launched between two host reads of the convergence flag. */
#define PFEM_SOLVER_OPT_FACTOR_ORDER 1
#define PFEM_SOLVER_OPT_POLL 2
#define PFEM_SOLVER_OPT_REPLACE 4 /* multi-RHS / batched PFEM_PRECISION_MIXED, default 2: a residual replacement
                                    follows every fall of (r, M^-1 r) by 10^-value (value 1..8) */
#define PFEM_SOLVER_OPT_GS_PASSES 6 /* GMRES Arnoldi: classical Gram-Schmidt passes, 2 (default: one
                                       re-orthogonalisation) or 1 (for well-conditioned systems; the
                                       reference's modified Gram-Schmidt lies between the two) */
#define PFEM_SOLVER_OPT_GRAPH 5  /* fused PCG (fusion level 2 or more), default 0: 1 makes each solve capture one
                                    batch of iterations (the poll interval) as a HIP graph and replay it; 0
                                    launches the kernels one by one (the same bits either way) */
#define PFEM_SOLVER_OPT_FUSE 3   /* PCG kernel fusion in factor order, default 3: 0 none; 1 the x/r update runs
                                    inside the forward sweep; 2 also alpha and beta are finished by the last
                                    workgroups of the SpMV and of the backward sweep; 3 also the two sweeps
                                    share one buffer whose values carry a 1-bit tag in the lowest mantissa
                                    bit instead of re-armed sentinels (y and z change by at most one unit in
                                    the last place; levels 0 to 2 give the same bits as each other), and
                                    with IC0 the forward sweep ends the iteration with gamma = (y, y),
                                    y = L^-1 r, equal to (r, z) in exact arithmetic, and the backward
                                    sweep applies the x and p updates (three launches per iteration) */
int pfem_solver_set_option(pfem_solver s, int option, int value);
int pfem_solver_destroy(pfem_solver s);
/* LinearSolver::solve(A, x, b) -> bool (linear_solver.cu:146-154, 301-376, 421-521): warm start from x,
 * the preconditioner is re-initialised from A first. stats may be NULL. */
int pfem_solver_solve(pfem_solver s, pfem_csr a, double* d_x, const double* d_b, pfem_solve_stats* stats);
/* Two systems on one pattern solved together (the two velocity-prediction systems, which the reference
 * solves one after the other with one solver, cases/Cylinder2D/main.cu:820-824): s0 solves a0 x0 = b0 and
 * s1 solves a1 x1 = b1, each with its own Krylov state, scalars and stopping test, so each x is that of a
 * single pfem_solver_solve. For GMRES with a Jacobi or no preconditioner, in FP64 or PFEM_PRECISION_MIXED,
 * and matrices sharing one pattern, the Arnoldi steps run in lockstep with one SpMV launch for both
 * (the pattern streamed once) and one host poll for both; other combinations solve one after the other.
 * stats may be NULL (else two entries). */
int pfem_solver_solve_pair(pfem_solver s0, pfem_csr a0, double* d_x0, const double* d_b0, pfem_solver s1,
                           pfem_csr a1, double* d_x1, const double* d_b1, pfem_solve_stats stats[2]);
/* Multi-right-hand-side PCG (no counterpart in the reference, whose SolverCG::solve, linear_solver.cu:301-376,
 * takes one right-hand side per call): A X = B for nrhs (1..8) columns sharing the matrix, column j of X and
 * B at d_X + j*n and d_B + j*n (X is also the initial guess). Each column runs the solver's PCG recurrence
 * and stopping test with its own scalars and stops on its own; the matrix is streamed once per iteration
 * for all columns. The solver must be PCG with a Jacobi preconditioner or none, in PFEM_PRECISION_FP64 or
 * PFEM_PRECISION_MIXED (PCG with reliable updates: FP32 vectors and matrix copy in the iterations, the solution
 * in FP64, FP64 residual replacement after every tenfold drop and for the stopping test; stats.refinements
 * counts the replacements). stats: nrhs entries (or NULL). */
int pfem_solver_solve_multi(pfem_solver s, pfem_csr a, int nrhs, double* d_X, const double* d_B,
                            pfem_solve_stats* stats);
/* Tile limits of pfem_solver_solve_batched per system size (no counterpart in the reference): a system of
 * n rows takes the first bucket k with n <= max_rows[k] and is cut into SpMV tiles of at most tile_rows[k]
 * rows (<= 256) and tile_nnz[k] non-zeros (<= 2048); larger systems, and every system when nbuckets = 0
 * (the default), take 256 rows and 2048 non-zeros. Set by the offline sweep of gpupfem2_amd/tune.py. */
int pfem_solver_set_batch_tiles(pfem_solver s, int nbuckets, const int* max_rows, const int* tile_rows,
                                const int* tile_nnz);
/* Batched independent PCG solves (no counterpart in the reference, which solves one system per
 * SolverCG::solve call, linear_solver.cu:301-376): a is block diagonal, system k owning rows and columns
 * [h_sys_rows[k], h_sys_rows[k+1]) (h_sys_rows: nsys + 1 host ints from 0 to n; checked). Each system runs
 * the solver's PCG recurrence and stopping test with its own scalars and stops on its own; all systems
 * share every launch. Jacobi preconditioner or none; PFEM_PRECISION_FP64 or PFEM_PRECISION_MIXED (reliable
 * updates per system, as pfem_solver_solve_multi). d_x is the initial guess (FP64 path) and the solution.
 * stats: nsys entries (or NULL); h_time_ms (or NULL): per-system time to solution on the device clock,
 * from the start of the call's first Krylov or refinement pass to the pass where the system stopped. */
int pfem_solver_solve_batched(pfem_solver s, pfem_csr a, int nsys, const int* h_sys_rows, double* d_x,
                              const double* d_b, pfem_solve_stats* stats, double* h_time_ms);

/* ------------------------------------------------------------------ quadrature */
/* createEdgeQuadratureFormula / createFaceQuadratureFormula / createCellQuadratureFormula
 * (src/integration/quadrature_formula_{1d,2d,3d}.cu). kind 1: edge rule i in 0..4 -> (coord, weight);
 * kind 2: face rule i in 0..7 -> (l1, l2, l3, weight); kind 3: tet rule i in 0..1 -> (l1..l4, weight).
 * Returns the number of points (or a negative error code); out needs room for 13*5 doubles. */
int pfem_quadrature(int kind, int index, double* out);

/* ------------------------------------------------------------------ assembly (case integrands) */
/* cases/Poisson2D/main.cu:28-68 (kIntegrateOverCell) as a library operator: coeff replaces the
 * hard-coded 0.5; rhs_kind 0 exp(-(x^2+10y^2)), 1 rhs_param*sin(pi x)sin(pi y), 2 constant rhs_param.
 * Overwrites the matrix values and d_rhs (no clearValues needed). quad_index: face rule. */
int pfem_assemble_poisson2d(pfem_mesh2d m, pfem_csr a, double* d_rhs, double coeff, int rhs_kind,
                            double rhs_param, int quad_index);
/* cases/Poisson3D/main.cu:26-67; rhs_kind as above with z; quad_index: tet rule (0 = 4 pt, 1 = 6 pt) */
int pfem_assemble_poisson3d(pfem_mesh3d m, pfem_csr a, double* d_rhs, double coeff, int rhs_kind,
                            double rhs_param, int quad_index);

/* SimulationParameters (src/parameters.cuh:6-51), field for field, plus the build's extras */
typedef struct {
    double rho, mu, tFinal, dt, tolerance;
    int maxIterations, simulationScheme;
    const char* meshFileName;
    int outputFrequency, particleAdvectionSubsteps, exportParticles, calculateLoads, bodyBoundaryID;
    double thickness, meanVelocity;
} pfem_simulation_parameters;
void pfem_simulation_parameters_default(pfem_simulation_parameters* p); /* setDefaultParameters() */

typedef struct {
    double rho, mu, dt;
    int scheme;            /* 0 Scheme A, 1 Scheme B (reference: pressureOld != nullptr) */
    int face_quad, edge_quad;
    int traction_mask;     /* bit b: boundary ID b carries the viscous traction edge term */
    double normal[8][2];   /* outward unit normal per boundary ID */
} pfem_pfem2_params;
/* kIntegrateVelocityPrediction (cases/Cylinder2D/main.cu:72-167, PoiseuilleFlow2D/main.cu:81-176):
 * two matrices on one pattern + two rhs; d_p_old NULL for Scheme A */
int pfem_assemble_prediction(pfem_mesh2d m, pfem_csr a0, pfem_csr a1, double* d_rhs0, double* d_rhs1,
                             const double* d_u0, const double* d_u1, const double* d_uo0, const double* d_uo1,
                             const double* d_p_old, const pfem_pfem2_params* prm);
/* kIntegratePressureEquation (cases/Cylinder2D/main.cu:169-211) */
int pfem_assemble_pressure(pfem_mesh2d m, pfem_csr a, double* d_rhs, const double* d_vp0, const double* d_vp1,
                           const double* d_p_old, const pfem_pfem2_params* prm);
/* kIntegrateVelocityCorrection (cases/Cylinder2D/main.cu:213-259): the two matrices are identical, one
 * is written */
int pfem_assemble_correction(pfem_mesh2d m, pfem_csr a, double* d_rhs0, double* d_rhs1, const double* d_vp0,
                             const double* d_vp1, const double* d_p, const double* d_p_old,
                             const pfem_pfem2_params* prm);
/* VelocityDirichletBCs::setDirichletValues (cases/Cylinder2D/main.cu:261-297, 467-479): writes
 * g*_b = g_b + dt/rho * avg_cells(d_component(p - p_old)) into target's device values */
int pfem_prediction_bc_values(pfem_mesh2d m, pfem_dirichlet target, pfem_dirichlet source, const double* d_p,
                              const double* d_p_old, int component, const pfem_pfem2_params* prm);
/* boundaryLoadsCalculator (cases/Cylinder2D/main.cu:299-445): forces {Fp_x, Fp_y, Fv_x, Fv_y} on the
 * edges with boundary ID body_id, outward normal taken from the centre (cx, cy) */
int pfem_body_forces(pfem_mesh2d m, int body_id, double cx, double cy, const double* d_u0, const double* d_u1,
                     const double* d_p, const pfem_pfem2_params* prm, double out[4]);
/* No counterpart in the reference: pfem_body_forces for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc). Same inputs and the same bits
 * as pfem_body_forces; it calls no hipMalloc/hipFree: the body's edges are found on the device and the
 * per-edge loads go to one static device array per GPU (about 2.4 MB, held until the process ends).
 * Limits: at most 16777216 cells and 65536 body edges (PFEM_EINVAL beyond). Calls that overlap on one
 * GPU share that array and race. Synchronises the context stream to return the totals. */
int pfem_body_forces_noalloc(pfem_mesh2d m, int body_id, double cx, double cy, const double* d_u0,
                             const double* d_u1, const double* d_p, const pfem_pfem2_params* prm, double out[4]);

/* ------------------------------------------------------------------ particles (particle_handler_2d.cuh) */
/* ParticleHandler2D(const Mesh2D*, int cellDivisionLevel) (particle_handler_2d.cu:238-294) */
int pfem_particles_create(pfem_mesh2d m, int level, pfem_particles* out);
int pfem_particles_destroy(pfem_particles ph);
/* seedParticles (particle_handler_2d.cu:304-320). mode 0: L^2 particles at the subcell centroids
 * (reference); mode 1: ppc particles per cell at uniform random barycentric positions from
 * Philox4x32-10(key = seed, counter = (cell, k, 0, 0)). */
int pfem_particles_seed(pfem_particles ph, int mode, int ppc, uint64_t seed);
/* initParticleVelocity (:322-326) */
int pfem_particles_init_velocity(pfem_particles ph, const double* d_u0, const double* d_u1);
/* advectParticles(velocity, dt, substeps) (:328-342): advection + relocation per substep, deletion of
 * particles that leave the 1-ring, subcell refill; the particles are then ordered by cell */
int pfem_particles_advect(pfem_particles ph, const double* d_u0, const double* d_u1, double dt, int substeps);
/* projectVelocityOntoGrid (:350-361) */
int pfem_particles_project(pfem_particles ph, double* d_u0, double* d_u1);
/* correctParticleVelocity (:344-348) */
int pfem_particles_correct(pfem_particles ph, const double* d_u0, const double* d_u1, const double* d_uo0,
                           const double* d_uo1);
/* getParticleCount (:30-32) */
int pfem_particles_count(pfem_particles ph, long* count);
/* host copy of the particle state: cell, lambda (3 per), position (2), velocity (2), ID; NULL skips */
int pfem_particles_download(pfem_particles ph, int* cell, double* lam, double* pos, double* vel, uint32_t* id);
int pfem_particles_upload(pfem_particles ph, long count, const int* cell, const double* lam, const double* pos,
                          const double* vel, const uint32_t* id, uint32_t next_id);
/* particles deleted / added by the last advect */
int pfem_particles_stats(pfem_particles ph, long* deleted_last, long* added_last);

/* ------------------------------------------------------------------ PFEM-2 time step (cases: Cylinder2D, PoiseuilleFlow2D main.cu) */
typedef struct {
    pfem_pfem2_params prm;
    int outer_iterations;      /* Cylinder2D: 2 (main.cu:810); PoiseuilleFlow2D: 1 (main.cu:669) */
    int substeps;              /* particleAdvectionSubsteps */
    int pred_solver, pred_precond, pres_solver, pres_precond, corr_solver, corr_precond;
    double tol;
    int tol_mode, maxit, restart;
    int dirichlet_mode;
    int body_id;               /* -1: no loads */
    double body_cx, body_cy, force_coeff;
    int precision;             /* PFEM_PRECISION_FP64, or PFEM_PRECISION_MIXED: solves with Jacobi/no
                                  preconditioner run FP32-inner/FP64-refinement, IC0/ILU0 solves run
                                  PFEM_PRECISION_MIXED_PRECOND */
    double inner_tol;          /* MIXED: relative tolerance of each FP32 inner solve */
    int pres_ordering;         /* PFEM_ORDER_* for an IC0/ILU0 pressure preconditioner */
    int pres_factor_bits;      /* MIXED with a multicolour IC0 pressure factor: bits of its reduced copy, 16 or
                                  32 (0 means 32; pfem_precond_set_factor_bits) */
    int pred_gs_passes;        /* GMRES prediction solves: classical Gram-Schmidt passes per Arnoldi step, 1 or
                                  2 (PFEM_SOLVER_OPT_GS_PASSES; 0 means 1, enough for these mass-dominated
                                  systems; 2 orthogonalises like the reference's modified Gram-Schmidt) */
} pfem_step_config;
/* The time loop body of cases/Cylinder2D/main.cu:793-880 on a mesh with edge IDs set. Velocity BCs
 * share one node list with a value per component; pressure BCs have their own list. */
int pfem_pfem2_create(pfem_mesh2d m, pfem_particles ph, const pfem_step_config* cfg, int nbcu, const int* bcu_node,
                      const double* bcu_val0, const double* bcu_val1, int nbcp, const int* bcp_node,
                      const double* bcp_val, pfem_pfem2* out);
int pfem_pfem2_destroy(pfem_pfem2 s);
/* one time step; forces (may be NULL) receives {Cx, Cy} = 2 F / (rho U^2 D) */
int pfem_pfem2_step(pfem_pfem2 s, double forces[2]);
/* device pointers of the fields: 0 u_x, 1 u_y, 2 p, 3 u*_x, 4 u*_y, 5 u_old_x, 6 u_old_y, 7 p_old */
int pfem_pfem2_field(pfem_pfem2 s, int which, double** d_ptr);
/* Krylov iterations of the last step, 5 per outer iteration (pred x, pred y, pressure, corr x, corr y) */
int pfem_pfem2_last_iterations(pfem_pfem2 s, int* its, int cap);
/* accumulated device time per phase since the last reset, in ms: 0 particles (advect+refill+sort),
 * 1 projection, 2 assembly+BCs, 3 prediction solves, 4 pressure solves, 5 correction solves,
 * 6 particle correction + loads. Timing is off unless enabled (it synchronises at phase ends). */
int pfem_pfem2_enable_timing(pfem_pfem2 s, int on);
int pfem_pfem2_phase_times(pfem_pfem2 s, double* ms, int cap);
/* Whole-step orchestration (replaces the kernel-by-kernel launches of the time loop,
 * cases/Cylinder2D/main.cu:793-880): with on != 0, the work between the solves (particle step,
 * projection, the three assemblies with their Dirichlet passes, particle correction and loads) is
 * captured as HIP graphs at the step after the call (a first step, eager, makes the one-time set-ups)
 * and replayed at every later step; the Krylov solves stay kernel by kernel. The results are the same
 * bits as the eager step. Graphs are re-captured after the particle buffers grow; 0 destroys them. */
int pfem_pfem2_set_graphs(pfem_pfem2 s, int on);
/* Solves re-run since creation because a single-pass triangular sweep's wait exceeded its spin limit (no
 * counterpart in the reference, whose SpSV has no such wait): such a solve is repeated from the same start
 * with per-level sweeps and the FP64 factor instead of failing the step, and a line goes to stderr. */
int pfem_pfem2_sweep_fallbacks(pfem_pfem2 s);
/* the stepper's matrices (borrowed, owned by s): 0 prediction x, 1 prediction y, 2 pressure,
 * 3 correction; and its preconditioners: 0 prediction, 1 pressure, 2 correction */
int pfem_pfem2_matrix(pfem_pfem2 s, int which, pfem_csr* out);
int pfem_pfem2_precond(pfem_pfem2 s, int which, pfem_precond* out);

/* ------------------------------------------------------------------ measurement */
/* Average device time of one launch, measured with HIP events around `launches` back-to-back
 * launches on the context stream (no counterpart in the reference, whose GpuTimer times phases,
 * src/common/gpu_timer.cuh:14-98). op 0: FP64 SpMV of a; op 1: FP32 SpMV of a (values converted
 * first); op 2: FP64 apply of pc (initialised from a first); op 3: the FP32 apply of pc; op 4: the apply
 * to FP64 vectors with the FP32 factor (PFEM_PRECISION_MIXED_PRECOND); ops 5 and 6: ops 2 and 4 on
 * vectors already in a multicolour factor's ordering (as the Krylov solvers run them); op 7: FP64 SpMV
 * of the matrix in that ordering (P a P^T, as the Krylov solvers run it); ops 8 and 9: ops 6 and 5 with
 * the tagged sweeps of PCG fusion level 3 (sliced layout); ops 10 and 11: one iteration of the fused PCG
 * as pfem_solver_solve runs it with pc (op 10 FP64 factor, op 11 its PFEM_PRECISION_MIXED_PRECOND copy),
 * the device time of a solve of 2 launches iterations minus one of `launches` iterations, over launches. */
int pfem_time_kernel(int op, pfem_csr a, pfem_precond pc, int launches, double* avg_ms);
/* The device's streaming ceiling (no counterpart in the reference; the roofline's measured reference
 * point): a copy kernel (16-byte loads, 4 or 8 in flight per lane, grid-stride) over two buffers of `bytes`
 * bytes, timed with HIP events over `launches` launches for each of the 14 forms (seven grid sizes, 256 to
 * 16,384 workgroups of 256 lanes, times the two load depths); returns the best average time of one copy
 * (it reads and writes `bytes` each) and, if best_grid is not NULL, the form that gave it as
 * depth * 100000 + workgroups. bytes: a positive multiple of 16. */
int pfem_stream_copy(pfem_ctx ctx, size_t bytes, int launches, double* avg_ms, int* best_grid);

/* ------------------------------------------------------------------ export (src/data_export.cuh) */
/* DataExport::exportToVTK (src/data_export.cu:17-95): ASCII VTU with nodal scalars (device arrays) */
int pfem_export_vtk(pfem_mesh2d m, const char* path, int nfields, const char* const* names,
                    const double* const* d_fields);
/* DataExport3D::exportToVTK (src/data_export_3d.cu:15-93): ASCII VTU of a tetrahedral mesh (cell type 10)
 * with nodal scalars (device arrays) */
int pfem_export_vtk3d(pfem_mesh3d m, const char* path, int nfields, const char* const* names,
                      const double* const* d_fields);

#ifdef __cplusplus
}
#endif
#endif
