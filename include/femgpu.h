/*
 * femgpu.h — C ABI of the MI355X (gfx950) sparse solve path.
 *
 * This is the drop-in boundary for the sparse solve path of the C++17
 * library cpp-FEM-solver. The reference exposes it as the C++ classes
 * sparse::SMatrix (ode-solver/include/math/array.hpp:264-373) and
 * block::BlockMatrix (ode-solver/include/math/blockMatrix.hpp:121-451); the
 * C++ facade shipped with this library (cxx/include/math/array.hpp and
 * cxx/include/math/blockMatrix.hpp) keeps those declarations and calls the
 * functions below. Every entry point says which reference interface it
 * replaces.
 *
 * Conventions
 *  - Plain pointers and sizes only. Unless a name ends in _device, vector
 *    arguments are host pointers and the call is synchronous, like the
 *    reference's std::vector calls.
 *  - Indices are uint32 and values float64, as in SMatrix (array.hpp:269-273).
 *  - Every call returns a status code; fem_last_error() gives the message of
 *    the last failure on the calling thread. The codes map onto the
 *    reference's exception types (array.cpp): FEM_ERR_INVALID_ARGUMENT ->
 *    std::invalid_argument, FEM_ERR_OUT_OF_RANGE -> std::out_of_range,
 *    FEM_ERR_RUNTIME -> std::runtime_error. FEM_ERR_DEVICE means no usable
 *    gfx950 device or a HIP failure; there is no CPU fallback.
 *  - Non-convergence is not an error (array.cpp:3597, 3841, 4400): solvers
 *    return FEM_OK with the reference's iteration semantics.
 *  - A matrix handle owns one HIP stream; calls on one handle are not
 *    reentrant (the reference's preconditioner closures are not either).
 *    Preconditioners and solvers run on their matrix's stream and keep
 *    pointers into it: destroy them before the matrix.
 */
#ifndef FEMGPU_H
#define FEMGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fem_status {
    FEM_OK = 0,
    FEM_ERR_INVALID_ARGUMENT = 1,
    FEM_ERR_OUT_OF_RANGE = 2,
    FEM_ERR_RUNTIME = 3,
    FEM_ERR_DEVICE = 4
};

typedef struct fem_csr_s* fem_csr;
typedef struct fem_bsr_s* fem_bsr;
typedef struct fem_pc_s* fem_pc;
typedef struct fem_solver_s* fem_solver;

/* Preconditioner kinds. The reference's string names map onto these through
 * per-solver alias tables (array.cpp:3484-3511, 3641-3669, 4215-4238). */
enum fem_pc_kind {
    FEM_PC_NONE = 0,    /* "none"/"identity"/"noprecond"(/"cg")            */
    FEM_PC_DIAG = 1,    /* "diag"/"diagonal"/"jacobi"   array.cpp:2450     */
    FEM_PC_IC0 = 2,     /* "ic"/"ic0"/"ichol"/...        array.cpp:2509     */
    FEM_PC_ILU0 = 3,    /* "ilu"/"ilu0"                   array.cpp:2733     */
    FEM_PC_SSOR = 4,    /* "sor"/"ssor"                   array.cpp:3068     */
    FEM_PC_AMG = 5,     /* "amg" (classical C/F)          array.cpp:3390     */
    FEM_PC_SAAMG = 6,   /* smoothed aggregation on a BS=1 condensation,
                           blockMatrix.hpp:1051                              */
    FEM_PC_LEARNED = 7, /* learned sparse approximate inverse (no reference
                           counterpart; see fem_spai_*)                      */
    FEM_PC_BJACOBI = 8, /* block-Jacobi, blockMatrix.hpp:714-735 and
                           test_BlockMatrix_mpi.cpp:132-155                   */
    FEM_PC_BILU0 = 9,   /* block ILU(0), blockMatrix.hpp:454-623             */
    FEM_PC_BSAAMG = 10, /* block SA-AMG on a BSR matrix, blockMatrix.hpp:1051 */
    FEM_PC_SCHUR = 11,  /* Schur-complement preconditioner of a saddle-point
                           matrix, blockMatrix.hpp:625-685 (see
                           fem_precond_create_schur)                         */
    FEM_PC_SPAI = 12    /* a given sparse approximate inverse M, or G with
                           M = G^T G (no reference counterpart; see
                           fem_precond_create_spai)                          */
};

enum fem_solver_kind {
    FEM_SOLVER_PCG = 0,      /* SMatrix::pcg       array.cpp:3445-3601 */
    FEM_SOLVER_BICGSTAB = 1, /* SMatrix::bicgstab  array.cpp:3602-3842 */
    FEM_SOLVER_GMRES = 2,    /* SMatrix::gmres     array.cpp:4190-4401 */
    FEM_SOLVER_MINICG = 3    /* mini_cg of test_BlockMatrix_mpi.cpp:157-182 */
};

enum fem_precision {
    FEM_PREC_FP64 = 0,    /* every operation in float64 (the reference)       */
    FEM_PREC_FP32_IR = 1  /* float32 Krylov + preconditioner inside float64
                             iterative refinement (true residual, update and
                             AMG coarsest solve in float64)                   */
};

enum fem_smoother {
    FEM_SMOOTH_JACOBI = 0, /* damped (block-)Jacobi, array.cpp:3355-3369 */
    FEM_SMOOTH_SGS = 1     /* symmetric multicolour Gauss-Seidel/SSOR      */
};

typedef struct fem_pc_params {
    double omega;       /* SSOR relaxation (1.2, array.hpp:312); AMG smoother damping;
                           SA-AMG: <= 0 selects 4/(3 rho(D^{-1}A)) per level (default) */
    double theta;       /* AMG strength threshold (0.25) */
    int npre, npost;    /* AMG pre/post smoothing sweeps */
    int max_levels;     /* AMG level cap (25 C/F, 20 SA) */
    int coarse_size;    /* AMG coarsest size (300 rows C/F, 50 block rows SA) */
    int max_agg;        /* SA aggregate size cap (4) */
    int smoother;       /* enum fem_smoother */
    double p_omega;     /* SA prolongator smoothing weight; <= 0 (default) selects
                           the robust variant: 4/(3 rho(D^{-1}A)) per level,
                           aggregation by a distance-2 maximal independent set
                           of the strength graph, and a stop when a level keeps
                           > 85% of its rows; > 0 (2/3 = the reference) keeps
                           the reference's greedy aggregation */
    int coarse_exact;   /* SA: 1 = pivoted LU coarse solve; 0 = the reference's
                           elimination sequence (blockMatrix.hpp:958-999) */
    int fp32;           /* also build a float32 copy for FEM_PREC_FP32_IR */
    int setup_device;   /* SA robust variant: 1 (default) builds the hierarchy on
                           the device (same result as the host build), 0 on the
                           host */
} fem_pc_params;

typedef struct fem_solve_opts {
    int solver;        /* enum fem_solver_kind */
    double tol;        /* relative residual tolerance (reference default 1e-10) */
    int max_iter;      /* iteration cap (reference default 2000) */
    int restart;       /* GMRES restart length */
    int precision;     /* enum fem_precision */
    double inner_tol;  /* FP32_IR: relative reduction asked of each inner solve */
    int max_refine;    /* FP32_IR: refinement step cap */
    int check_every;   /* iterations captured per hipGraph replay (0 = auto) */
    int device_ptrs;   /* 1: b and x are device pointers on this device */
    int use_graph;     /* 1: replay captured hipGraphs (default), 0: plain launches */
} fem_solve_opts;

typedef struct fem_solve_stats {
    int iterations;        /* reference iteration semantics (array.cpp:3597 etc.) */
    int inner_iterations;  /* Krylov iterations actually run */
    int refinements;       /* FP32_IR outer steps */
    int converged;         /* 1 if the stopping test passed */
    double true_relres;    /* ||b - A x||_2 / ||b||_2 in float64 at exit */
    double setup_ms;       /* preconditioner setup time of this call (0 if prebuilt) */
    double solve_ms;       /* wall time of the iterations */
    long long spmv_count;  /* operator applications */
    int fallback;          /* 1 if a learned preconditioner's budget ran out and
                              the system was solved with its fallback */
} fem_solve_stats;

/* ---- library ---------------------------------------------------------- */
const char* fem_last_error(void);
/* Name and compute-unit count of the current device; FEM_ERR_DEVICE if none. */
int fem_device_info(char* name, int name_len, int* cu_count);
void fem_pc_params_default(int kind, fem_pc_params* p);
void fem_solve_opts_default(int solver, fem_solve_opts* o);

/* ---- CSR matrix (sparse::SMatrix) -------------------------------------- */
/* CSR constructor without checks: SMatrix(rows, cols, row_ptr, col_idx,
 * values), array.cpp:2183-2189. Arrays are copied to the device. */
int fem_csr_create(uint32_t rows, uint32_t cols, const uint32_t* row_ptr, const uint32_t* col_idx,
                   const double* values, fem_csr* out);
/* COO constructor: SMatrix(rows, cols, values, rowlist, collist),
 * array.cpp:2072-2182 (validate, drop zeros, sort, merge duplicates by sum,
 * drop merged zeros), done on the device. */
int fem_csr_create_coo(uint32_t rows, uint32_t cols, uint64_t nnz, const double* values,
                       const uint32_t* rowlist, const uint32_t* collist, fem_csr* out);
/* rows()/cols()/nnz(), array.hpp:296-298 */
int fem_csr_info(fem_csr A, uint32_t* rows, uint32_t* cols, uint64_t* nnz);
/* row_ptr()/col_idx()/values(), array.hpp:300-302 (copies to host) */
int fem_csr_export(fem_csr A, uint32_t* row_ptr, uint32_t* col_idx, double* values);
/* SMatrix::transpose, array.cpp:2202-2235 (on the device) */
int fem_csr_transpose(fem_csr A, fem_csr* out);
/* SMatrix::operator*, array.cpp:2336-2447 (mathematically exact product, on
 * the device) */
int fem_csr_multiply(fem_csr A, fem_csr B, fem_csr* out);
int fem_csr_destroy(fem_csr A);
/* SpMV variant of A (no counterpart in the reference, which has one loop):
 * lanes = 0 selects the CSR-stream kernel, whose row sums run sequentially in
 * stored order and match SMatrix::spmv bit for bit; 4, 8, 16, 32 or 64 give
 * each row that many lanes and a butterfly sum (last bits differ). Default 0. */
int fem_csr_set_row_lanes(fem_csr A, int lanes);
/* Column-index storage of A's CSR-stream SpMV (no counterpart in the
 * reference): with on = 1 the columns are also kept as 16-bit offsets from a
 * base per 64 entries, read instead of the 32-bit indices (2 bytes less per
 * nonzero), when every such chunk spans fewer than 65536 columns and nnz >=
 * 65536; on = 0 drops that copy; on < 0 changes nothing. *packed (may be
 * NULL) receives 1 if the packed form is in use. Results are identical either
 * way. Matrices are packed by default where possible. Change it only before
 * creating preconditioners or solvers on A (captured solves keep the form
 * they were captured with). */
int fem_csr_pack_columns(fem_csr A, int on, int* packed);

/* ---- primitives --------------------------------------------------------- */
/* SMatrix::spmv, array.cpp:3418-3442 */
int fem_spmv(fem_csr A, const double* x, double* y);
/* Same on device buffers of this device, enqueued on A's stream; fp32 != 0
 * uses the float32 copy of A with float32 x and y. */
int fem_spmv_device(fem_csr A, const void* x, void* y, int fp32);
/* SMatrix::dot, array.cpp:3401-3416 */
int fem_dot(uint64_t n, const double* a, const double* b, double* out);

/* No counterpart in the reference: fem_dot for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops). a, b and
 * out are device pointers on the calling thread's current device; a[0..n) .
 * b[0..n) is written to *out, enqueued on `stream` (a hipStream_t; NULL = the
 * null stream), and the call returns without waiting. It allocates nothing:
 * its partial sums live in a static device array of 2048 doubles (16 KB) per
 * device, held until the process ends. Kernels and summation order are
 * fem_dot's, so the result is bit-identical. n must be below 2^32
 * (FEM_ERR_INVALID_ARGUMENT otherwise). Calls that overlap on one device
 * share that array and race: order them (one stream, or events). */
int fem_dot_device_noalloc(uint64_t n, const double* a, const double* b, double* out, void* stream);

/* ---- preconditioners (SMatrix::make_*_preconditioner, array.hpp:309-313) -- */
int fem_precond_create(fem_csr A, int kind, const fem_pc_params* params, fem_pc* out);
/* make_schur_preconditioner(sys, A_precond, S_precond), blockMatrix.hpp:
 * 625-685, for the saddle-point matrix [[A, B^T], [B, 0]]: B is n_pres x
 * n_vel, Bt its transpose (n_vel x n_pres), A_pc a preconditioner of size
 * n_vel, S_pc one of size n_pres or NULL (identity); S_pc's output is
 * multiplied by s_scale. z = M^{-1} r is
 *   w_u = A_pc(r_u), z_p = s_scale S_pc(B w_u - r_p), z_u = A_pc(r_u - B^T z_p),
 * composed on the device (no host round trip; usable inside captured
 * solves). The result keeps pointers to B, Bt, A_pc and S_pc: destroy it
 * first. */
int fem_precond_create_schur(fem_csr B, fem_csr Bt, fem_pc A_pc, fem_pc S_pc, double s_scale, fem_pc* out);
/* z = M^{-1} r: the std::function the reference factories return */
int fem_precond_apply(fem_pc M, const double* r, double* z);
int fem_precond_destroy(fem_pc M);
/* Setup time and summary of a preconditioner: levels = triangular-solve levels
 * (IC0/ILU0/block ILU0) or AMG levels; fills up to cap entries of rows/nnz
 * per AMG level. */
int fem_precond_info(fem_pc M, int* levels, uint32_t* rows, uint64_t* nnz, int cap, double* setup_ms);
/* AMG hierarchy inspection (no counterpart in the reference, whose AMG data
 * stays inside the std::function closure, array.cpp:3390-3398): level l's
 * operator (which = 0), prolongator (1) or restriction (2) as CSR. dims
 * receives rows, columns and nnz; with row_ptr NULL only dims is written.
 * FEM_ERR_OUT_OF_RANGE if the level has no such matrix. */
int fem_amg_level_export(fem_pc M, int level, int which, uint32_t* dims, uint32_t* row_ptr, uint32_t* col_idx,
                         double* values);
/* C/F AMG: the C/F split of coarsened level l (n bytes: 1 C point, 2 F point) */
int fem_amg_split(fem_pc M, int level, uint8_t* cf);
/* Algorithmic bytes of one application of M in float64 (fp32 = 0) or float32:
 * the sum over the kernels it launches, each reading its matrices (uint32
 * indices) and vectors once and writing its output once. 0 for the kinds
 * not tallied (tallied: diag, IC0, ILU0, AMG, SA-AMG). A measurement aid with no
 * counterpart in the reference. */
int fem_precond_bytes(fem_pc M, int fp32, double* bytes);

/* ---- Krylov solvers ----------------------------------------------------- */
/* General entry: solver, precision, graph capture and pointer space from opts.
 * M may be NULL (no preconditioner). */
int fem_solve(fem_csr A, fem_pc M, const double* b, double* x, const fem_solve_opts* opts,
              fem_solve_stats* stats);
/* SMatrix::pcg(A, b, precond, tol, max_iter), array.cpp:3445 */
int fem_pcg(fem_csr A, fem_pc M, const double* b, double tol, int max_iter, double* x, int* iterations);
/* SMatrix::bicgstab, array.cpp:3602 */
int fem_bicgstab(fem_csr A, fem_pc M, const double* b, double tol, int max_iter, double* x, int* iterations);
/* SMatrix::gmres(A, b, precond, restart, tol, max_iter), array.cpp:4190 */
int fem_gmres(fem_csr A, fem_pc M, const double* b, int restart, double tol, int max_iter, double* x,
              int* iterations);
/* SMatrix::gmres(A, b, std::function solveM, ...), array.cpp:4402: the
 * preconditioner is a host callback, called with host vectors of length n. */
typedef void (*fem_host_precond_fn)(const double* r, double* z, uint64_t n, void* user);
int fem_gmres_callback(fem_csr A, fem_host_precond_fn fn, void* user, const double* b, int restart, double tol,
                       int max_iter, double* x, int* iterations);

/* Persistent solver: keeps workspace and captured graphs across solves. */
int fem_solver_create(fem_csr A, fem_pc M, const fem_solve_opts* opts, fem_solver* out);
int fem_solver_solve(fem_solver S, const double* b, double* x, fem_solve_stats* stats);
int fem_solver_destroy(fem_solver S);

/* ---- block CSR (block::BlockMatrix<double,BS>) -------------------------- */
/* BlockMatrix(nbr, nbc, row_ptr, col_idx, values), blockMatrix.hpp:170-181;
 * bs in {1,2,3,4}; values row-major bs*bs per block. */
int fem_bsr_create(int bs, uint32_t nbr, uint32_t nbc, const uint32_t* row_ptr, const uint32_t* col_idx,
                   const double* values, fem_bsr* out);
/* BlockMatrix::condense(SMatrix, nbr, nbc), blockMatrix.hpp:269-315 */
int fem_bsr_condense(fem_csr A, int bs, uint32_t nbr, uint32_t nbc, fem_bsr* out);
int fem_bsr_info(fem_bsr A, int* bs, uint32_t* nbr, uint32_t* nbc, uint64_t* nnzb);
int fem_bsr_export(fem_bsr A, uint32_t* row_ptr, uint32_t* col_idx, double* values);
/* BlockMatrix::transpose, blockMatrix.hpp:237-267 */
int fem_bsr_transpose(fem_bsr A, fem_bsr* out);
/* BlockMatrix::matvec, blockMatrix.hpp:208-234 */
int fem_bsr_matvec(fem_bsr A, const double* x, double* y);
int fem_bsr_matvec_device(fem_bsr A, const void* x, void* y, int fp32);
/* FEM_PC_BJACOBI / FEM_PC_BILU0 (make_block_ilu0_preconditioner,
 * blockMatrix.hpp:455) / FEM_PC_BSAAMG (make_block_amg_preconditioner,
 * blockMatrix.hpp:1051-1053) */
int fem_bsr_precond_create(fem_bsr A, int kind, const fem_pc_params* params, fem_pc* out);
/* PCG / mini_cg / GMRES on a block matrix; BlockMatrix::gmres(A, b, precond,
 * tol, max_iter, restart), blockMatrix.hpp:444-451 when opts->solver is GMRES */
int fem_bsr_solve(fem_bsr A, fem_pc M, const double* b, double* x, const fem_solve_opts* opts,
                  fem_solve_stats* stats);
int fem_bsr_destroy(fem_bsr A);

/* ---- learned sparse approximate inverse ------------------------------------
 * No counterpart in the reference: an MLP (6 -> 64 -> 64 -> 7, ReLU) reads
 * each row's scaled stencil couplings f_i[d] = -6 a_ij / sqrt(a_ii a_jj) of a
 * 7-point grid matrix (rows (z*ny + y)*nx + x, directions d = 0..5 for the
 * column offsets -nx*ny, -nx, -1, +1, +nx, +nx*ny) and emits the row's
 * entries of a lower-triangular G on A's lower pattern:
 *   G_ii = softplus(o_i[0]) / sqrt(a_ii), G_ij = o_i[1+d] / sqrt(a_jj) for the
 *   lower neighbours (d = 0, 1, 2; outputs 4..6 unused).
 * The preconditioner is M = G^T G, SPD for any output, applied as two sparse
 * products. Parameters are float32, flat, FEM_MLP_NPARAM of them: W1[6][64],
 * b1[64], W2[64][64], b2[64], W3[64][7], b3[7] (W[in][out]). The device
 * kernels run the layers on the matrix cores with bfloat16 operands and
 * float32 accumulation. Training minimises L = 1/(n k) sum_c ||v_c - M A v_c||^2
 * over k probe vectors with Adam. grid[3] = {nx, ny, nz}. */
#define FEM_MLP_NPARAM 5063
typedef struct fem_mlp_s* fem_mlp;
typedef struct fem_adam_opts {
    double lr, beta1, beta2, eps;
} fem_adam_opts;
/* params: FEM_MLP_NPARAM floats, or NULL for He-initialised weights from seed */
int fem_mlp_create(const float* params, uint64_t seed, fem_mlp* out);
int fem_mlp_destroy(fem_mlp net);
int fem_mlp_get_params(fem_mlp net, float* params);
int fem_mlp_set_params(fem_mlp net, const float* params);
/* network outputs (n x 7, host) for host features (n x 6) */
int fem_mlp_forward(fem_mlp net, const float* feat, uint32_t n, float* out);
/* the features of A (n x 6, host); FEM_ERR_INVALID_ARGUMENT if A is not a
 * symmetric-pattern 7-point matrix on the grid with a positive diagonal */
int fem_spai_features(fem_csr A, const uint32_t* grid, float* feat);
/* loss and gradient (FEM_MLP_NPARAM floats, may be NULL) for k probe vectors
 * V (k x n, probe c at V + c*n; device pointers if device_ptrs) */
int fem_spai_loss_grad(fem_mlp net, fem_csr A, const uint32_t* grid, const double* V, int k, int device_ptrs,
                       double* loss, float* grad);
/* one Adam step on that loss; *loss is the loss before the step */
int fem_spai_train_step(fem_mlp net, fem_csr A, const uint32_t* grid, const double* V, int k, int device_ptrs,
                        const fem_adam_opts* opt, double* loss);
/* M = G^T G from the network (two sparse products per apply). A solve with
 * this preconditioner stops after `budget` iterations (0: no explicit limit)
 * or, once the fallback's iteration count on this matrix is known, after
 * that many (a learned solve needing more is a regression); it is then
 * solved again with the fallback preconditioner (FEM_PC_IC0, FEM_PC_ILU0,
 * FEM_PC_AMG, FEM_PC_SAAMG or FEM_PC_DIAG, built on first use) and
 * fem_solve_stats.fallback is set. The fallback's count is recorded whenever
 * it runs, or given by fem_precond_learned_reference. */
int fem_precond_create_learned(fem_csr A, const uint32_t* grid, fem_mlp net, int fallback_kind, int budget,
                               fem_pc* out);
/* the fallback's iteration count on this matrix (0: unknown), e.g. from one
 * solve with the fallback preconditioner */
int fem_precond_learned_reference(fem_pc M, int iterations);
/* A given sparse approximate inverse (no counterpart in the reference; the
 * non-learned baselines of the learned one use it): factorized = 0 applies
 * z = Mat r, factorized = 1 takes Mat as G and applies z = G^T (G r). Mat
 * must be n x n for A n x n; the result keeps a pointer to Mat. */
int fem_precond_create_spai(fem_csr A, fem_csr Mat, int factorized, fem_pc* out);

/* ---- batched solves ----------------------------------------------------------
 * No counterpart in the reference: many independent SPD systems (block size
 * 1 or 3, up to 4 right-hand sides each) solved in one persistent launch by
 * block-Jacobi PCG (the reference's block-Jacobi, blockMatrix.hpp:714-735,
 * with cached inverse blocks as in test_BlockMatrix_mpi.cpp:132-155) in
 * float32 inside float64 iterative refinement. A workgroup solves one
 * system at a time, all its right-hand sides together, and takes the next
 * system from a work queue. Systems are BSR with node-major vectors. */
typedef struct fem_batch_s* fem_batch;
typedef struct fem_batch_system {
    int bs;                   /* 1 or 3 */
    uint32_t nb;              /* block rows (= block columns) */
    const uint32_t* row_ptr;  /* nb + 1 */
    const uint32_t* col_idx;  /* row_ptr[nb] block columns */
    const double* values;     /* row_ptr[nb] row-major bs x bs blocks */
} fem_batch_system;
typedef struct fem_batch_opts {
    double tol;               /* ||b - A x|| < tol ||b|| per right-hand side */
    double inner_tol;         /* float32 inner solves: relative reduction */
    int max_refine;           /* float64 refinement steps */
    int max_inner;            /* inner iterations per refinement step */
} fem_batch_opts;
typedef struct fem_batch_stats {
    int iterations;           /* inner iterations (all right-hand sides together; for a
                                 split system the largest of its right-hand sides') */
    int refinements;
    int converged;            /* right-hand sides that reached tol */
    double max_relres;        /* largest final ||b - A x|| / ||b|| (float64) */
} fem_batch_stats;
typedef struct fem_batch_schedule {
    int block;                /* workgroup size: 256, 512 or 1024 */
    int wg_per_cu;            /* resident workgroups per CU (0: occupancy limit) */
    int order;                /* queue: 0 largest estimated work first, 1 input order */
    int split;                /* the `split` systems of largest estimated work are
                                 queued as one work item per right-hand side */
} fem_batch_schedule;
/* copies the systems to the device and inverts the diagonal blocks */
int fem_batch_create(int nsys, const fem_batch_system* systems, int nrhs, fem_batch* out);
/* right-hand sides: for each system in turn, nrhs vectors of its nb*bs rows */
int fem_batch_set_rhs(fem_batch b, const double* B);
/* solves every system from x0 = 0; per_system (nsys entries) may be NULL;
 * *ms = device time of the persistent launch */
int fem_batch_solve(fem_batch b, const fem_batch_opts* opts, fem_batch_stats* per_system, double* ms);
/* solutions, same layout as the right-hand sides */
int fem_batch_get_solution(fem_batch b, double* X);
/* schedule autotuner (profile-guided): solves once unsplit to record every
 * system's iterations (they become the work estimates), times each of the 18
 * unsplit schedules on the whole batch with every system capped at
 * 1/divisor of its recorded iterations, then times the fastest of each
 * workgroup size on full solves, unsplit and with the systems whose work
 * exceeds an even share per resident workgroup split, and returns the
 * fastest; for a largest-first order it then runs that schedule once more to
 * record every work item's device time, and orders the queue by those times
 * when that is faster (trial_ms receives up to cap timings: the 18 capped
 * runs, the full runs, then that run). The batch is left unsplit with its
 * previous schedule; apply
 * the result with fem_batch_set_schedule. Solutions are left as the last
 * trial computed them. */
int fem_batch_tune(fem_batch b, const fem_batch_opts* opts, int divisor, fem_batch_schedule* best, double* trial_ms,
                   int cap);
int fem_batch_set_schedule(fem_batch b, const fem_batch_schedule* schedule);
int fem_batch_destroy(fem_batch b);

/* ---- measurement --------------------------------------------------------- */
/* Average device time of `reps` back-to-back SpMV launches on A's stream,
 * timed with HIP events (inputs resident in HBM). */
int fem_bench_spmv(fem_csr A, int fp32, int reps, double* avg_ms);
/* The same, each launch timed alone after a 1 GiB write has evicted the L2
 * and the 256 MB Infinity Cache, so the matrix streams from HBM. */
int fem_bench_spmv_cold(fem_csr A, int fp32, int reps, double* avg_ms);
/* Average device time of `reps` back-to-back applications of M (float32
 * sweeps when fp32 != 0), timed with HIP events on M's stream, input and
 * output resident in HBM. */
int fem_bench_precond(fem_pc M, int fp32, int reps, double* avg_ms);
/* The device's streaming-read ceiling: average time of a read of `bytes`
 * (16-byte loads), back to back (cold = 0) or each after a 1 GiB read has
 * evicted the L2 and the Infinity Cache (cold != 0). No counterpart in the
 * reference: it calibrates the roofline fractions. */
int fem_bench_stream(uint64_t bytes, int cold, int reps, double* avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* FEMGPU_H */
