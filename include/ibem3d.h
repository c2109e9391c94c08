/* ibem3d -- MI355X engine for the iBEM elastic-3D spherical-inclusion solve.
 *
 * C ABI. Plain pointers and sizes; every call returns 0 on success and a negative code on error
 * (message from ibem3d_last_error). No exception crosses the ABI. Caller-owned inputs are copied to
 * the device at create; outputs are caller-allocated. A context owns its device memory and one HIP
 * stream; calls on one context are serialised by the caller; several contexts may coexist.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   configElastic3D::Readpartical_info / ReadBEM_info / ReadPostProcess_info
 *       configElastic3D.cpp:32-197, configElastic3D.h:76-129      -> ibem3d_create
 *   BEMBuilderElastic3D::addBEM             BEMBuilderElastic3D.cpp:1355  -> ibem3d_add_bem
 *   integratorElastic3D::addFieldToBEM      integratorElastic3D.cpp:135   -> ibem3d_add_field
 *   integratorElastic3D::addFluxEquivalent{,FirstOrder,SecondOrder}Inclusion
 *                                           integratorElastic3D.cpp:341, 455, 568 -> ibem3d_add_incl
 *   BEMBuilderElastic3D::addFluxEquivalent{,FirstOrder,SecondOrder}BEM
 *                                           BEMBuilderElastic3D.cpp:1511, 1629, 1743 -> ibem3d_add_bem_rows
 *   BEMBuilderElastic3D::ApplyBC            BEMBuilderElastic3D.cpp:1920  -> ibem3d_apply_bc
 *   systemSolverElastic3D::solve            systemSolverElastic3D.h:44-60 -> ibem3d_solve
 *   BEMBuilderElastic3D::getUTVector        BEMBuilderElastic3D.cpp:2053  -> ibem3d_get_ut
 *   postProcessorElastic3D::getFlux         postProcessorElastic3D.cpp:525 -> ibem3d_stress
 *
 * Unknown layout. The reference keeps 78 unknowns per sphere (6 uniform, 18 linear, 54 quadratic
 * eigenstrain coefficients, configElastic3D.cpp:308-343). The quadratic block E_ij,kl is symmetric in
 * (k,l), so the device system keeps 60 = 6 + 18 + 36 (k <= l) per sphere; n = 3 NN + 60 num.
 * ibem3d_get_ut returns the reference's 78-layout vectors.
 *
 * Boundary conditions are folded while assembling: the dense system A (n x n, column-major) is built
 * directly in the column order ApplyBC produces (displacement-unknown node dofs, then traction-unknown
 * (node, dof) keys in first-encounter order, then eigenstrain columns), and b is accumulated at the
 * same time; the reference's HMAT and GMAT are never materialised.
 */
#ifndef IBEM3D_H
#define IBEM3D_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ibem3d_ctx ibem3d_ctx;

/* One RVE in the reference's input formats (position.txt, BEM.txt, postprocess.txt). */
typedef struct {
  int NN, NE, num, nump;   /* nodes, elements, spheres, stress points */
  double E0, nu0;          /* matrix (BEM.txt lines 2-3) */
  double E1, nu1;          /* particles (position.txt lines 2-3) */
  const double* nodes;     /* NN x 3, row-major */
  const int* conn;         /* NE x 4, row-major, one-based node ids */
  const double* bc;        /* 3NN x 2, row-major: (flag 1 = displacement unknown / 0 = known, value) */
  const double* tbc;       /* NE x 24, row-major: 12 flags (1 = traction known, 0 = unknown), 12 values */
  const double* centres;   /* num x 3, row-major */
  const double* radii;     /* num */
  const double* points;    /* nump x 3, row-major (may be NULL when nump == 0) */
} ibem3d_problem;

enum {
  IBEM3D_OK = 0,
  IBEM3D_EINVAL = -1,   /* bad argument or inconsistent boundary conditions */
  IBEM3D_EHIP = -2,     /* HIP runtime error (device, memory) */
  IBEM3D_ESTATE = -3,   /* call out of order */
  IBEM3D_ESOLVE = -4    /* singular system or refinement did not converge */
};

/* solver modes for ibem3d_solve */
enum {
  IBEM3D_SOLVE_FP64_LU = 0,     /* FP64 blocked LU with row pivoting (FP64 MFMA update) */
  IBEM3D_SOLVE_MIXED_FP32 = 1,  /* FP32 LU (FP32 MFMA update) + FP64 iterative refinement */
  IBEM3D_SOLVE_MIXED_FP16 = 2,  /* FP32 LU whose trailing updates run on FP16 MFMA with scaled hi/lo split
                                   operands (FP32-level accuracy) + FP64 refinement */
  IBEM3D_SOLVE_MIXED_FP16_1 = 3,/* FP32 LU whose trailing updates are single FP16 MFMA products of
                                   operands scaled per row of L21 / column of U12 by powers of two
                                   (FP32 accumulation) + FP64 refinement */
  IBEM3D_SOLVE_MIXED_BF16 = 4,  /* the same with single BF16 products */
  IBEM3D_SOLVE_BJ_GMRES = 5,    /* no n x n factor: FP64 GMRES right-preconditioned by block Jacobi
                                   (the LU of the equilibrated boundary block, 3 NN x 3 NN, and the
                                   inverse of every sphere's 60 x 60 diagonal block), from the initial
                                   guess of ibem3d_set_initial_guess when one is set; when GMRES does
                                   not reach tol in 300 iterations (three restarts of 100), the
                                   MIXED_FP16_1 LU solves the system instead
                                   (ibem3d_solve_fell_back reports it) */
  IBEM3D_SOLVE_BJ_GMRES16 = 6,  /* BJ_GMRES whose outer FP64 GMRES (FP64 residuals and products with
                                   A) is preconditioned by a few inner block-Jacobi GMRES iterations
                                   on a 16-bit copy of A (FP16, every 256 x 32 tile scaled by a power
                                   of two; 2 n^2 bytes beside A, read at a quarter of A's bytes per
                                   product). Without room for the copy it runs as BJ_GMRES. Same
                                   fallback as BJ_GMRES. */
  IBEM3D_SOLVE_MIXED_BF16X3 = 7,/* MIXED_FP16's three-product split with BF16 hi/lo operands (16
                                   significant bits each) on BF16 MFMA + FP64 refinement */
  IBEM3D_SOLVE_GMRES = 0x100,   /* flag for the mixed modes: refine by GMRES-IR (FP64 GMRES, right-
                                   preconditioned by the LU factors) from the start. Without it the
                                   refinement is classic and switches to GMRES-IR when an iteration
                                   gains less than a factor of 10. */
  IBEM3D_SOLVE_REUSE = 0x200,   /* flag: solve with the factors of the last factorisation in the same
                                   mode (no new factorisation). Valid while the system matrix is the
                                   one factored: same context, only known values changed
                                   (ibem3d_set_known_values). */
  IBEM3D_SOLVE_COMPACT = 0x400  /* flag: the factorisation takes the place of A instead of a second
                                   n x n copy (chosen without the flag when free device memory cannot
                                   hold that copy). FP64_LU: A <- R A C and its LU in place; the
                                   residual is not computed (ibem3d_residual assembles A again for
                                   it). Mixed modes: FP32 R A C converted in place, factored, then
                                   stored as an FP16 copy (every 128 x 128 tile scaled by a power
                                   of two), A assembled again over
                                   the FP32 factor, and FP64 GMRES-IR against A preconditioned by
                                   the 16-bit factor. Peak memory 8 n^2 + 2 n^2 bytes instead of
                                   8 n^2 + 4 n^2 (mixed) or 16 n^2 (FP64). */
};

/* Create a context on HIP device `device`: copies the problem to the device, builds the BC column
 * map, allocates the n x n system. Returns NULL on failure (reason in ibem3d_last_error(NULL)). */
ibem3d_ctx* ibem3d_create(const ibem3d_problem* prob, int device);
void ibem3d_destroy(ibem3d_ctx* ctx);
const char* ibem3d_last_error(const ibem3d_ctx* ctx);

/* Runtime tunables of the solver, process-wide (no counterpart in the reference); they shape the
 * launches enqueued or captured after the call. Set by the batch autotuner. */
enum {
  IBEM3D_TUNE_BIG_MIN_TILES = 1,     /* 256 x 256 tiles from which a trailing update takes the persistent
                                        kernel (default 512; below it the 128 x 128 kernel) */
  IBEM3D_TUNE_PANEL_CU_PERCENT = 2,  /* percent of the default schedule of CUs the bulk trailing update
                                        leaves to the concurrent panel (default 100, 0..400) */
  IBEM3D_TUNE_INNER_ITERS = 3,       /* IBEM3D_SOLVE_BJ_GMRES16: inner GMRES iterations on the 16-bit A
                                        per outer direction (default 6, 1..16) */
  IBEM3D_TUNE_INNER_TOL_EXP = 4,     /* IBEM3D_SOLVE_BJ_GMRES16: the inner GMRES stops at relative
                                        residual 10^-value (default 3, 1..8) */
  IBEM3D_TUNE_ROWS_POINTWISE = 5,    /* 1: the boundary->inclusion rows contract every Gauss point's
                                        Kelvin table (the reference's order of sums) even where every
                                        element is planar; 0 (default): there the tables are integrated
                                        over the element first (4 contractions per element and row) */
  IBEM3D_TUNE_BJ_NOPIV = 6,          /* the block-Jacobi boundary block: 1 (default) factored without
                                        row interchanges in its natural column order (each boundary
                                        dof's own unknown on the diagonal) when the column map gives
                                        one, falling back to partial pivoting on a pivot below 1e-6;
                                        0 always partial pivoting; 2 a threshold no pivot meets, so
                                        the fallback runs every time */
  IBEM3D_TUNE_BJ_OVERLAP = 7,        /* IBEM3D_SOLVE_BJ_GMRES16: 1 (default) the boundary block's
                                        triangular solves of each inner iteration run on a second
                                        stream while the FP16 product covers the columns past the
                                        boundary block; 0: before the product */
  IBEM3D_TUNE_STRESS_BESIDE = 8,     /* 1 (default): each GMRES cycle computes the stresses of its new
                                        solution on a second stream beside its FP64 residual check;
                                        ibem3d_stress returns them when that residual ended the solve
                                        (bit-identical to computing them afterwards); 0: ibem3d_stress
                                        computes them */
  IBEM3D_TUNE_BJ_FP32 = 9,           /* 1 (default): the block-Jacobi boundary block, when factored
                                        without row interchanges, is factored and solved in FP32 (it
                                        only shapes search directions; residuals stay FP64); 0: FP64 */
  IBEM3D_TUNE_RESIDUAL_UPDATE = 10   /* 1 (default): a GMRES cycle checks its new x with
                                        r - (A Z) y, r the cycle's starting residual and A Z its FP64
                                        products (the same b - A x without another pass over A), and
                                        computes b - A x by a product only when that misses the
                                        tolerance; 0: always by a product */
};
int ibem3d_set_tuning(int key, long value);
long ibem3d_get_tuning(int key);  /* -1 for an unknown key */

/* System size n = 3 NN + 60 num. */
long ibem3d_size(const ibem3d_ctx* ctx);

/* ---- assembly phases (same call order as iBEMSystemService::run, iBEMSystemService.h:58-67) ---- */
int ibem3d_add_bem(ibem3d_ctx* ctx);                 /* boundary rows x boundary columns, b rows */
int ibem3d_add_field(ibem3d_ctx* ctx);               /* boundary rows x eigenstrain columns */
int ibem3d_add_incl(ibem3d_ctx* ctx, int order);     /* inclusion rows of order 0/1/2 (or -1: all) */
int ibem3d_add_bem_rows(ibem3d_ctx* ctx, int order); /* inclusion rows x boundary columns, b rows */
int ibem3d_apply_bc(ibem3d_ctx* ctx);                /* checks that every phase ran; system ready */
/* All of the above in order, the inclusion and boundary->inclusion phases once each with order -1
 * (the same entries as the three per-order calls, with the per-pair tables built once). */
int ibem3d_assemble(ibem3d_ctx* ctx);
/* Clear the assembled system so the phases can run again on the same inputs (no counterpart in
 * the reference, which re-reads its inputs for every run). */
int ibem3d_reset(ibem3d_ctx* ctx);

/* Solve A x = b. mode: IBEM3D_SOLVE_*; tol: relative residual target ||b - A x|| / ||b|| for the
 * refined modes (ignored by FP64_LU). */
int ibem3d_solve(ibem3d_ctx* ctx, int mode, double tol);

/* Solution: U (3NN + 78 num, known displacements reinserted, eigenstrains in the 78 layout),
 * T (12 NE, known tractions reinserted) -- BEMBuilderElastic3D::getUTVector. X60 (n) may be NULL. */
int ibem3d_get_ut(ibem3d_ctx* ctx, double* U, double* T, double* X60);

/* Stresses at the problem's points: nump x 6 row-major (s00 s01 s02 s11 s12 s22). */
int ibem3d_stress(ibem3d_ctx* ctx, double* s);

/* Replace the known boundary values (no counterpart in the reference, which re-reads its inputs
 * for every load): bc_values (3 NN) are the prescribed displacements where BEM.txt marks a
 * displacement known, tbc_values (12 NE) the prescribed tractions where a traction is known; the
 * flags of which values are known stay those of ibem3d_create, so the system matrix is unchanged
 * and its factors stay valid (IBEM3D_SOLVE_REUSE). Clears the assembled system: run the assembly
 * phases again, then solve. Entries at positions that are not known are ignored. */
int ibem3d_set_known_values(ibem3d_ctx* ctx, const double* bc_values, const double* tbc_values);

/* The same new known values without a new matrix (no counterpart in the reference): b is assembled
 * again (the known-value folds of ApplyBC, BEMBuilderElastic3D.cpp:1973-1990, from the boundary and
 * boundary->inclusion phases, in the full assembly's order, so b equals the full assembly's bit for
 * bit) and A -- or the factor that occupies its memory after a compact FP64 solve -- is left as it
 * is. Then solve with IBEM3D_SOLVE_REUSE. Requires an assembled system. */
int ibem3d_update_known_values(ibem3d_ctx* ctx, const double* bc_values, const double* tbc_values);

/* Volume average of the equivalent eigenstrain (no counterpart in the reference):
 * <e*_ij> = (1/V) sum_spheres V_s (E_ij + (a_s^2 / 5) sum_k E_ij,kk), the integral of the quadratic
 * eigenstrain polynomial over each sphere; V as in ibem3d_average_fields. With it the average
 * stress follows from the inclusions, <s> = C0 : (<e> - <e*>), without the boundary tractions.
 * eps_star: 6 values in the order 00 01 02 11 12 22. */
int ibem3d_average_eigenstrain(ibem3d_ctx* ctx, double* eps_star);

/* Volume averages over the RVE from the boundary solution (no counterpart in the reference: the
 * effective-property post-step). With U, T as ibem3d_get_ut returns them, on every element's 8 x 8
 * Gauss rule: stress <s_ij> = (1/V) int_S t_i x_j dS (symmetrised), strain
 * <e_ij> = (1/V) int_S (u_i n_j + u_j n_i) / 2 dS, V = (1/3) int_S x . n dS. sigma, eps: 6 values
 * each in the order 00 01 02 11 12 22; any pointer may be NULL. Host arithmetic in a fixed order. */
int ibem3d_average_fields(ibem3d_ctx* ctx, double* sigma, double* eps, double* volume);

/* Copy the assembled system (A column-major n x n, b n) to the host; either may be NULL. Every
 * assembly phase must have run since the last reset (IBEM3D_ESTATE otherwise: the phases write
 * their blocks of A rather than accumulating into a cleared matrix). */
int ibem3d_get_system(ibem3d_ctx* ctx, double* A, double* b);

/* Initial guess of the eigenstrain unknowns (60 num, the device's 60 layout: 6 uniform, 18 linear,
 * 36 quadratic coefficients with k <= l, per sphere) for the IBEM3D_SOLVE_BJ_GMRES solves that follow
 * -- for instance the learned per-sphere prediction; the boundary unknowns start from the boundary
 * block's solve against it. NULL clears it. (No counterpart in the reference.) */
int ibem3d_set_initial_guess(ibem3d_ctx* ctx, const double* x_eig);
/* 1 if the last IBEM3D_SOLVE_BJ_GMRES solve fell back to the mixed LU, else 0. */
int ibem3d_solve_fell_back(const ibem3d_ctx* ctx);
/* Inner GMRES iterations (on the 16-bit A) of the last IBEM3D_SOLVE_BJ_GMRES16 solve, beside the outer
 * iterations ibem3d_solve reports; -1 when that solve ran as BJ_GMRES (no room for the copy). No
 * counterpart in the reference. */
long ibem3d_inner_iterations(const ibem3d_ctx* ctx);
/* 1 when the boundary->inclusion rows (ibem3d_add_bem_rows, and the right-hand side of
 * ibem3d_update_known_values) are integrated in moment form: every boundary element is planar (its
 * Gauss-point normals agree to 1e-14), so the Kelvin tables are summed over the element's Gauss
 * points before the row contraction, one contraction per shape function instead of one per point;
 * 0 for the per-point contraction (a non-planar element, or IBEM3D_TUNE_ROWS_POINTWISE = 1). The
 * entries agree to rounding. No counterpart in the reference. */
int ibem3d_rows_moment_form(const ibem3d_ctx* ctx);
/* The block-Jacobi preconditioner's current boundary factor: 2 computed without row interchanges in
 * FP32 (IBEM3D_TUNE_BJ_NOPIV, IBEM3D_TUNE_BJ_FP32), 1 without interchanges in FP64, 0 with partial
 * pivoting (FP64) or before any block-Jacobi setup. No counterpart in the reference. */
int ibem3d_bj_nopiv(const ibem3d_ctx* ctx);
/* The last block-Jacobi solve's start from the initial guess: relative residual of the scaled guess
 * alpha x0 (alpha in [0, 1] minimises ||b - alpha A x0||, so the start is never worse than zero);
 * rel0 = -1 when no guess was used. */
int ibem3d_warm_start_info(const ibem3d_ctx* ctx, double* rel0, double* alpha);

/* Relative residual ||b - A x|| / ||b|| of the current solution with the FP64 system matrix. After a
 * compact FP64 solve (A overwritten by its factor) the system is assembled again first, which ends
 * the factor's validity (no counterpart in the reference, which reports no residual). */
int ibem3d_residual(ibem3d_ctx* ctx, double* rel_residual);

/* Refinement statistics of the last solve: relative residual (-1 when not computed: compact FP64
 * solve), iterations (0 for FP64_LU). */
int ibem3d_solve_info(const ibem3d_ctx* ctx, double* rel_residual, int* iterations);

/* Per-kernel timings of the calls since the last reset, measured with HIP events on the context's
 * stream: names (NUL-separated, up to cap bytes) and milliseconds; returns the entry count. */
int ibem3d_timings(const ibem3d_ctx* ctx, char* names, int cap, double* ms, int max_entries);
void ibem3d_reset_timings(ibem3d_ctx* ctx);

/* ---- batched RVE stiffness (no counterpart in the reference, which solves one RVE per process,
 * configElastic3D.cpp:27-30) ---- */
typedef struct ibem3d_batch ibem3d_batch;
typedef struct {
  int mode;             /* IBEM3D_SOLVE_MIXED_* */
  int refine_steps;     /* classic refinement steps per load inside the graph (>= 1) */
  int wave_max;         /* RVEs per wave at most */
  double mem_budget_gb; /* device memory a wave may use (GB); a wave holds at least one RVE */
  int use_graph;        /* 1: each wave captured as one graph and run by one launch; 0: direct launches */
  int host_threads;     /* direct launches: host threads enqueuing the wave's RVEs (<= 1: one) */
} ibem3d_batch_options;
/* count problems whose boundary displacements are all prescribed (every BEM.txt displacement flag 0,
 * every traction flag 0): the affine loads of the apparent stiffness. The problems are copied. */
ibem3d_batch* ibem3d_batch_create(const ibem3d_problem* probs, int count, int device);
void ibem3d_batch_destroy(ibem3d_batch* b);
const char* ibem3d_batch_last_error(void);
/* Apparent stiffness of every RVE from the six unit macro strains u = e_t . x (t in the order 00 01 02
 * 11 12 22, e_ij = e_ji = 1), one factorisation per RVE: C (count x 36, row-major 6 x 6 per RVE,
 * entry [c][t] = stress component c under load t) from the inclusions, <s> = C0 : (<e> - <e*>);
 * C_bnd (may be NULL) from the boundary tractions, <s> = (1/V) int t x dS; rel (count x 6, may be
 * NULL): each load's final relative FP64 residual; host_refined (count x 6, may be NULL): 1 where the
 * in-graph refinement had not reached tol and the load was refined further on the host; times (6,
 * may be NULL): seconds of setup, capture, instantiation, device run, host post-processing, and the
 * number of waves; eig (may be NULL): every RVE's eigenstrain solution, RVE after RVE, 6 loads x
 * 60 num values each (the device's 60 layout per sphere). */
int ibem3d_batch_stiffness(ibem3d_batch* b, const ibem3d_batch_options* opt, double tol, double* C,
                           double* C_bnd, double* rel, int* host_refined, double* times, double* eig);

/* ---- learned per-sphere eigenstrain predictor (no counterpart in the reference): a message-passing
 * MLP over each sphere's 16 nearest neighbours whose 360 outputs are the sphere's 60 eigenstrain
 * coefficients (coefficient c of order o times r^o) under each of the six unit macro strains; trained
 * and evaluated by its own HIP kernels (FP32 training, BF16-MFMA or FP32 inference). The prediction
 * contracted with the macro strain is an initial guess for IBEM3D_SOLVE_BJ_GMRES. ---- */
typedef struct ibem3d_gnn ibem3d_gnn;
const char* ibem3d_gnn_last_error(void);
long ibem3d_gnn_param_count(void);
/* Graph of one RVE on the device: nbr (num x 16; the 16 nearest other spheres by exact integer
 * distance of centres quantised to 2^-28, ties to the smaller index, -1 padding), edge features
 * (num x 16 x 8 floats) and node features (num x 4 floats). */
int ibem3d_gnn_graph(const double* centres, const double* radii, int num, double E1_over_E0, double nu1,
                     int device, int* nbr, float* edge, float* node);
ibem3d_gnn* ibem3d_gnn_create(int device, unsigned long long seed);
void ibem3d_gnn_destroy(ibem3d_gnn* g);
int ibem3d_gnn_get_params(ibem3d_gnn* g, float* p);
int ibem3d_gnn_set_params(ibem3d_gnn* g, const float* p);
/* Training set on the device: S spheres' edges (S x 16 x 8), nodes (S x 4), targets (S x 360). */
int ibem3d_gnn_set_data(ibem3d_gnn* g, const float* edge, const float* node, const float* target, long S);
/* Mean squared error and its gradient (param_count floats, may be NULL) over dataset rows idx[0..B). */
int ibem3d_gnn_loss_grad(ibem3d_gnn* g, const int* idx, long B, float* loss, float* grad);
/* Adam (lr) over epochs passes, each a permutation of the S rows (perms: epochs x S), minibatches of
 * B rows; loss_hist (epochs, may be NULL): mean minibatch loss per epoch. */
int ibem3d_gnn_train(ibem3d_gnn* g, const int* perms, int epochs, long B, float lr, float* loss_hist);
/* y (S x 360) for S spheres; bf16 = 1: dense layers on BF16 MFMA with FP32 accumulation. */
int ibem3d_gnn_predict(ibem3d_gnn* g, const float* edge, const float* node, long S, int bf16, float* y);
/* No counterpart in the reference: ibem3d_gnn_predict for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops). d_edge (S x 16 x 8), d_node (S x 4) and d_y
 * (S x 360) are device pointers on g's device; the call allocates nothing: the activations live in a
 * 16 MiB file-scope device array (one per device, held until the process ends) that the spheres pass
 * through in chunks of 1664, and the kernels are enqueued on `stream` (a hipStream_t; NULL: g's own
 * stream) without waiting. Every call on a device shares that array, so calls that overlap on one
 * device race; no part of it needs clearing (every element is written before it is read). Results
 * equal ibem3d_gnn_predict's bit for bit. */
int ibem3d_gnn_predict_noalloc(ibem3d_gnn* g, const float* d_edge, const float* d_node, long S, int bf16,
                               float* d_y, void* stream);

/* ---- the reference's text formats (host only; no GPU call) ---- */
typedef struct ibem3d_inputs ibem3d_inputs;
/* Read position.txt, BEM.txt and (if present) postprocess.txt from directory `dir`:
 * configElastic3D::Readparticle_num / Readpartical_info (configElastic3D.cpp:105-125),
 * ReadBEM_num / ReadBEM_info (configElastic3D.h:85-110, configElastic3D.cpp:32-103),
 * ReadPostProcess_info (configElastic3D.h:113-126) -- without the dense HMAT / GMAT / A the
 * reference allocates there (configElastic3D.cpp:95-101). Returns NULL on error (ibem3d_io_error). */
ibem3d_inputs* ibem3d_read_inputs(const char* dir);
/* The parsed problem; its pointers stay valid until ibem3d_free_inputs. */
const ibem3d_problem* ibem3d_inputs_problem(const ibem3d_inputs* in);
const char* ibem3d_inputs_title(const ibem3d_inputs* in);
void ibem3d_free_inputs(ibem3d_inputs* in);
const char* ibem3d_io_error(void);
/* Write nump x 6 stresses as stress.txt rows: full = 0 is the reference's format (std::scientific,
 * 6 digits; postProcessorElastic3D.cpp:578-585), full = 1 writes %.17g. */
int ibem3d_write_stress(const char* path, const double* s, int nump, int full);

#ifdef __cplusplus
}
#endif
#endif
