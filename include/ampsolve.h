/*
 * ampsolve: supernodal multifrontal Cholesky / LDL^T numeric factorization and supernodal triangular solves on one
 * AMD Instinct MI355X (gfx950), behind a plain C ABI.
 *
 * This is the boundary the reference's JavaScript host would bind through a thin Node-API addon.
 * Each entry point names the reference interface it replaces (jwt625/applied-multiphysics, packages/core/src/la/).
 * All pointers are plain host pointers unless the name ends in _device, in which case they are HIP device pointers
 * (HBM). The library copies what it needs during a call and never keeps a caller's host pointer.
 * Every function returns AMP_OK (0) or a negative AMP_ERR_* code; amp_last_error() describes the last failure of the
 * calling thread. A non-positive Cholesky pivot or a non-finite LDL^T pivot is NOT an error: the factor call returns
 * AMP_OK with info.ok = 0, as the reference's factor returns false / {ok: false}.
 * Handles are not reentrant: calls that share an amp_num must not overlap.
 */
#ifndef AMPSOLVE_H
#define AMPSOLVE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMP_OK 0
#define AMP_ERR_ARG -1         /* invalid argument (null pointer, negative size, unknown mode) */
#define AMP_ERR_HIP -2         /* a HIP runtime call failed (no device, out of device memory, launch failure) */
#define AMP_ERR_STATE -3       /* e.g. solve without a successful numeric factorization */
#define AMP_ERR_PATTERN -4     /* value count or sparsity pattern does not match the symbolic analysis */
#define AMP_ERR_INTERNAL -5    /* symbolic analysis failed internally */

#define AMP_CHOLESKY 0
#define AMP_LDLT 1

#define AMP_F64 0 /* factor, update matrices and solves in FP64 (MFMA f64) */
#define AMP_F32 1 /* factor and update matrices in FP32 (MFMA f32); solves in FP32 */
#define AMP_F32_BF16X3 2 /* as AMP_F32, with the bulk trailing updates of large fronts as three bf16 products per FP32
                          * product (bf16 matrix cores, FP32 accumulation; relative error about 2^-16 per product) */

typedef struct amp_sym amp_sym; /* symbolic analysis + execution plan (host), device copy of the plan */
typedef struct amp_num amp_num; /* numeric factor resident in HBM */
typedef struct amp_graph amp_graph; /* captured factor + solve + refinement sequence of one amp_num */
typedef struct amp_mlp amp_mlp;     /* warm-start network of the parametric sweep */

typedef struct {
  int32_t n, nsuper, maxFront, nlevels;
  int64_t nnzL, nnzA, nnzLStored, nnzPattern;
  double flops, fill;
  double gemmFlops;   /* algorithmic flops of the matrix-core tile kernel per FP32 factorization (lower parts only;
                       * FP64 factors also run the panel steps on it) */
  double bulkBytes;   /* algorithmic HBM bytes per FP32 factorization of its wide bulk trailing-update launches (the
                       * bf16 products of AMP_F32_BF16X3): lower output entries read and written, 4 bytes each way
                       * (written only by a front's first such update), and each split panel entry read once */
  int32_t bulkLaunches; /* those launches per factorization */
  char patternKey[64];
} amp_sym_info;

/* Read-only views of the analysis arrays (valid while the amp_sym lives); names and meaning as analyze() returns
 * them (sparse.mjs:873-901). Lxp holds int64 offsets (the reference keeps them in a Float64Array). */
typedef struct {
  const int32_t *perm, *pinv, *parent, *colCount;
  const int32_t *lowerP, *lowerI, *lowerSrc;
  const int32_t *superStart, *colSuper, *superParent, *childHead, *childNext;
  const int32_t *Sp, *Si;
  const int64_t *Lxp;
} amp_sym_arrays;

/* A symbolic analysis made by the caller's own symbolic phase (the reference's analyze(), sparse.mjs:666-902),
 * handed over as the arrays it returned. Everything else (superParent, children, Lxp, maxFront) is derived. */
typedef struct {
  int32_t n, nsuper;
  int64_t nnzPattern;           /* number of values of the CSR pattern the analysis was made from */
  const int32_t *perm;          /* n, new -> old */
  const int32_t *parent;        /* n, elimination tree in new indices (-1 at roots) */
  const int32_t *colCount;      /* n, column counts of L (may be NULL: supernodal counts are used) */
  const int32_t *superStart;    /* nsuper + 1 */
  const int32_t *Sp;            /* nsuper + 1 */
  const int32_t *Si;            /* Sp[nsuper] */
  const int32_t *lowerP;        /* n + 1 */
  const int32_t *lowerI;        /* lowerP[n] */
  const int32_t *lowerSrc;      /* lowerP[n] */
  const char *patternKey;       /* the analysis' patternKey (may be NULL) */
} amp_sym_desc;

typedef struct {
  int32_t ok;          /* 1: factor complete; 0: a pivot failed (see failColumn) */
  int32_t nNeg;        /* LDL^T: negative pivots before the first failure (Sylvester inertia when ok) */
  int32_t nPerturbed;  /* LDL^T: pivots replaced by +-tau before the first failure */
  int32_t failColumn;  /* first failing column (new index), -1 when ok */
  double ms;           /* device time of the factorization (HIP events on the handle's stream) */
} amp_factor_info;

/* ---- symbolic phase (host) ---------------------------------------------------------------------------------- */

/* amdOrder (sparse.mjs:297-592): approximate minimum degree, perm new -> old. denseFactor: 10 in the reference. */
int amp_amd_order(int32_t n, const int32_t *rowptr, const int32_t *colidx, double denseFactor, int32_t *perm);

/* ndOrder (sparse.mjs:230-283): nested dissection by recursive coordinate bisection with node coordinates x, y
 * (ordering 'nd' of analyze, sparse.mjs:678-679), perm new -> old; ranges of at most `leaf` nodes (48 in the reference)
 * keep their order. AMP_ERR_ARG for leaf < 1 or a malformed pattern. */
int amp_nd_order(int32_t n, const int32_t *rowptr, const int32_t *colidx, const double *x, const double *y, int32_t leaf,
                 int32_t *perm);

/* analyze (sparse.mjs:666-902): ordering (AMD when perm is NULL), etree, postorder, column counts, relaxed supernodes
 * (relax != 0), row structures and block offsets, plus the device execution plan. AMP_ERR_ARG when the plan's index
 * ranges cannot hold the system (a tiled front with rows x super-panel width >= 2^31, or more than 2^31 extend-add
 * ranges). */
int amp_analyze(int32_t n, const int32_t *rowptr, const int32_t *colidx, const int32_t *perm, int32_t relax,
                amp_sym **out);

/* Take over an analysis from the caller's symbolic phase (the drop-in path for the reference's own analyze()). */
int amp_sym_import(const amp_sym_desc *desc, amp_sym **out);

/* patternFingerprint (sparse.mjs:908-925) into out (>= 48 bytes); checkSymbolic compares it (:927-930). */
int amp_pattern_fingerprint(int32_t n, const int32_t *rowptr, const int32_t *colidx, char *out);

int amp_sym_get_info(const amp_sym *sym, amp_sym_info *info);
int amp_sym_get_arrays(const amp_sym *sym, amp_sym_arrays *arrays);
void amp_sym_free(amp_sym *sym);

/* ---- numeric phase (device) --------------------------------------------------------------------------------- */

/* Device pointers (the *_device calls, amp_graph_*, amp_pcg_device, amp_mlp_forward / amp_mlp_gradient): each call
 * orders its work after everything queued on the null (legacy default) stream before it; inputs produced on any
 * other stream must be complete before the call. The calls return with their outputs complete (amp_graph_launch:
 * after amp_graph_sync). */

/* Storage of one numeric factor in HBM (the NumericFactor / SupernodalLDLT storage of sparse.mjs:946-949 and
 * ldlt.mjs:428-448): mode AMP_CHOLESKY or AMP_LDLT, precision AMP_F64 or AMP_F32. Builds the device plan on first
 * use of the analysis. */
int amp_num_create(amp_sym *sym, int32_t mode, int32_t precision, amp_num **out);

/* factorNumeric (sparse.mjs:941-1125) / SupernodalLDLT.factor (ldlt.mjs:450-562): values aligned with the CSR
 * pattern (nvalues must equal the pattern's entry count, else AMP_ERR_PATTERN). pivotTol (LDL^T only, default 1e-13
 * in the reference) scales max |a_jj|. */
int amp_factor(amp_num *num, const double *values, int64_t nvalues, double pivotTol, amp_factor_info *info);
/* Same with the values already in HBM. */
int amp_factor_device(amp_num *num, const double *d_values, int64_t nvalues, double pivotTol, amp_factor_info *info);

/* solveNumeric / solveNumericMulti (sparse.mjs:1261-1310), SupernodalLDLT.solve / solveMany (ldlt.mjs:565-643):
 * nrhs vectors stored one after another, original ordering. AMP_ERR_STATE without a successful factor. */
int amp_solve(amp_num *num, const double *b, double *x, int32_t nrhs);
int amp_solve_device(amp_num *num, const double *d_b, double *d_x, int32_t nrhs);

/* Allocation-free variant of amp_solve_device (no counterpart in the reference): for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses that forbid hipMalloc). Same kernels,
 * arguments and results as amp_solve_device; the solve's scratch comes from a fixed array of
 * AMP_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) instead of the handle's growable buffers, and the
 * call allocates nothing. Requests needing more scratch are refused with AMP_ERR_ARG. Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (from several threads or handles) race. */
#define AMP_NOALLOC_SCRATCH_BYTES (64 << 20)
int amp_solve_device_noalloc(amp_num *num, const double *d_b, double *d_x, int32_t nrhs);

/* ---- refinement (device, FP64 residual and update) ---------------------------------------------------------- */

#define AMP_REFINE_IR 0    /* solveRefined (ldlt.mjs:253-275): x = solve(b); x += solve(b - A x) */
#define AMP_REFINE_GMRES 1 /* gmres (gmres-real.mjs:31-104) right-preconditioned by the factor, x0 = solve(b) */

typedef struct {
  int32_t method;   /* AMP_REFINE_IR or AMP_REFINE_GMRES */
  int32_t maxSteps; /* IR: refinement steps (3 in the reference); GMRES: maximum iterations (400) */
  int32_t restart;  /* GMRES restart length, 1..64 (40 in the reference) */
  double tol;       /* relative residual |b - A x| / |b| to stop at (IR 1e-14, GMRES 1e-10 in the reference) */
} amp_refine_opts;

typedef struct {
  int32_t steps;      /* IR: refinement steps taken; GMRES: iterations */
  int32_t converged;  /* IR: relResidual < tol; GMRES: true residual <= max(10 tol, 1e-14) */
  double relResidual; /* |b - A x| / |b| of the returned x (FP64) */
  double ms;          /* device time of the whole refinement */
} amp_refine_info;

/* Refine the factor's solution against the unperturbed values on the analysed CSR pattern (one right-hand side).
 * The residual r = b - A x, the update and the Krylov vectors stay in FP64 in HBM whatever the factor precision. */
int amp_refine(amp_num *num, const double *values, const double *b, double *x, const amp_refine_opts *opts,
               amp_refine_info *info);
int amp_refine_device(amp_num *num, const double *d_values, const double *d_b, double *d_x,
                      const amp_refine_opts *opts, amp_refine_info *info);
/* The CSR pattern (rowptr n + 1, colidx) for the residual of an analysis taken over with amp_sym_import. */
int amp_sym_set_pattern(amp_sym *sym, const int32_t *rowptr, const int32_t *colidx);

/* The factor in the reference's layout: Lx (nnzLStored, m x nc column-major blocks at Lxp; the upper part of each
 * diagonal block is unspecified, as in the reference) and D (n, LDL^T pivots; ignored for Cholesky; may be NULL). */
int amp_download_factor(amp_num *num, double *Lx, double *D);

/* Device time (ms) of the last factor and of the last solve ("solve_total", always recorded), and the total of each
 * kernel class of the last factor and solve when profiling is enabled (amp_num_set_profiling). `kernel` is a class
 * name ("gemm_tile", "gemm_bf16x3" (the bf16 bulk products of AMP_F32_BF16X3), "split_panel", "panel_step",
 * "solve_fwd_flow", ...) or a class name and an elimination-tree level ("gemm_tile@23") for that level's launches. */
int amp_num_set_profiling(amp_num *num, int32_t enable);
int amp_num_get_kernel_ms(amp_num *num, const char *kernel, double *total_ms, int32_t *launches);

/* FP32 factors (no counterpart in the reference): a 64 x 64 diagonal block of a tiled front whose largest elimination
 * multiplier (|l_ik / l_kk| for Cholesky, |l_ik| for LDL^T) exceeds `limit` is factored again in FP64 arithmetic, with
 * FP64 pivots, and stored rounded to FP32 (0: never; default 1e4 for LDL^T, whose static pivots can make the
 * multipliers grow, and 0 for Cholesky, whose factor of an SPD matrix cannot; ignored by FP64 factors). The number of blocks redone
 * in the last factor is reported as the launch count of the kernel class "diag_fp64_redo" (amp_num_get_kernel_ms). */
int amp_num_set_growth_limit(amp_num *num, double limit);

void amp_num_free(amp_num *num);
/* Free the factor's update-matrix workspace (needed only while factoring; the next factor allocates it again), so
 * that many factors can stay resident for solves only, as the parametric sweep's factor cache keeps them.
 * AMP_ERR_STATE while captured graphs of the handle exist. */
int amp_num_release_workspace(amp_num *num);

/* Shape thresholds of the device plans (no counterpart in the reference, whose kernel choice is per front,
 * front-real.mjs:189-244): fronts with height m <= smallMax are factored and solved whole in LDS (64 threads when
 * m <= smallNarrow, else 256); product launches use 128 x 128 tiles when they have at least wideLaunch of them (else
 * 64 x 64); tiled fronts are factored in super-panels of superPanel columns; a forward solve launch of the tiled
 * fronts of a level gives each workgroup several 64-row blocks (passed on inside the workgroup) when it has at least
 * flowWideMin such tasks, else one block each; whole fronts of at most 8 (16) rows are packed 8 (4) to a wave when a
 * level has at least packMin of them; with fusedPanel each 64-column step of a tiled front (diagonal block, panel,
 * the super-panel's later columns) is one launch, for super-panels of at most 256 columns; with fusedSolve (bit 0: forward, bit 1:
 * backward; default 3) each run of consecutive levels of tiled fronts of at most 16 blocks of 64 columns whose solve
 * tasks are one block wide is solved by one launch per direction (its gather or initialisation tasks among the flow
 * tasks, handed over through counters) instead of two launches per level and direction. amp_sym_autotune times the factor and one solve of d_values (device, pattern order) on the device for a set of candidate thresholds, starting from the
 * defaults and changing one threshold at a time, and keeps the fastest; amp_sym_set_params applies given thresholds.
 * Both rebuild the plans, so they refuse (AMP_ERR_STATE) while numeric handles of the analysis exist. */
typedef struct {
  int32_t smallMax, smallNarrow, wideLaunch, superPanel, flowWideMin, packMin, fusedPanel, fusedSolve;
  int32_t candidates;        /* out: plans timed */
  double defaultMs, bestMs;  /* out: factor + solve device time with the default and with the chosen thresholds */
} amp_tune_info;
int amp_sym_autotune(amp_sym *sym, const double *d_values, int32_t mode, int32_t precision, amp_tune_info *info);
int amp_sym_set_params(amp_sym *sym, const amp_tune_info *params);

/* Whole-sequence capture (no counterpart in the reference, whose factor and solve are separate synchronous calls):
 * the factor of the values at d_values, the solve of the nrhs right-hand sides at d_b (device, n doubles each, column
 * after column) into d_x, then irSteps steps of iterative refinement x += solve(b - A x) (FP64 residual on the CSR
 * pattern), captured once into a HIP graph and replayed by amp_graph_launch with no host synchronisation inside --
 * for callers that refactor and resolve the same pattern many times, such as a batch of independent systems
 * assembled into one block-diagonal pattern. The device buffers are read at every replay (new values / right-hand
 * sides in place). Cholesky handles only (AMP_ERR_ARG otherwise); the pattern must be known for irSteps > 0.
 * amp_graph_sync waits for the last replay and reports its factor status (info->ms is the replay's device time).
 * A graph holds the addresses of its factor's device buffers: while graphs of an amp_num exist, calls that would move
 * those buffers are refused with AMP_ERR_STATE (amp_num_release_workspace, and solves with more right-hand sides than
 * the buffers hold), and every graph must be freed before its amp_num. */
int amp_graph_create(amp_num *num, const double *d_values, const double *d_b, double *d_x, int32_t nrhs,
                     int32_t irSteps, amp_graph **out);
int amp_graph_launch(amp_graph *g);
int amp_graph_sync(amp_graph *g, amp_factor_info *info);
void amp_graph_free(amp_graph *g);

/* Shift-invert operator of the eigen driver (shiftInvertEigen's `op`, eigen.ts:156-166, applied by blockArnoldi,
 * arnoldi.mjs:39-157): Y = A^-1 (M X) for nvec vectors stored one after another, with M given by its values on the
 * analysed CSR pattern and A = K - sigma M factored in num (Cholesky or LDL^T, FP64 or FP32). The products with M are
 * FP64; the nvec solves are one multi-right-hand-side pass. The pattern must be known (amp_sym_set_pattern for an
 * imported analysis). */
int amp_shift_invert_apply(amp_num *num, const double *M, const double *X, double *Y, int32_t nvec);
int amp_shift_invert_apply_device(amp_num *num, const double *d_M, const double *d_X, double *d_Y, int32_t nvec);

/* Preconditioned conjugate gradients on the device (no counterpart in the reference; the parametric sweep's solver):
 * A x = b with A given by d_values on the analysed CSR pattern, the preconditioner being the factor held by `num`
 * (typically of a nearby matrix), starting from the guess in d_x; stops when |b - A x| / |b| <= tol or after maxIter
 * iterations. FP64 vectors in HBM. */
typedef struct {
  int32_t maxIter;
  double tol;
} amp_pcg_opts;
typedef struct {
  int32_t iters, converged;
  double relResidual, initialResidual, ms;
} amp_pcg_info;
int amp_pcg_device(amp_num *num, const double *d_values, const double *d_b, double *d_x, const amp_pcg_opts *opts,
                   amp_pcg_info *info);

/* The same for nb <= 64 independent systems in lockstep, sharing the preconditioner: system j has the matrix values
 * d_values + j nnzPattern, the right-hand side d_b + j bStride (bStride 0: one shared right-hand side) and its guess and
 * solution at d_x + j n. Each iteration applies the preconditioner to all systems in one multi-right-hand-side solve;
 * a system stops updating once it meets tol. info[j] reports system j (ms: the whole call). */
int amp_pcg_batch_device(amp_num *num, const double *d_values, const double *d_b, int64_t bStride, double *d_x,
                         int32_t nb, const amp_pcg_opts *opts, amp_pcg_info *info);

/* Warm-start network (no counterpart in the reference): an FP32 multilayer perceptron nIn -> nHidden -> nHidden -> nOut
 * (tanh hidden layers, linear output) with its training on the device. Parameters (flat, amp_mlp_param_count):
 * W1 (nHidden x nIn, row-major), b1, W2 (nHidden x nHidden), b2, W3 (nOut x nHidden), b3. Activations are batch-major
 * (batch x features, device pointers, batch <= maxBatch). amp_mlp_gradient forms the gradient of the mean squared
 * error against the targets d_t into the handle (amp_mlp_get_grads copies it out) and returns the error;
 * amp_mlp_adam takes one Adam step with it. */
int amp_mlp_create(int32_t nIn, int32_t nHidden, int32_t nOut, int32_t maxBatch, amp_mlp **out);
int64_t amp_mlp_param_count(const amp_mlp *mlp);
int amp_mlp_set_params(amp_mlp *mlp, const float *params, int64_t count);
int amp_mlp_get_params(amp_mlp *mlp, float *params, int64_t count);
int amp_mlp_get_grads(amp_mlp *mlp, float *grads, int64_t count);
int amp_mlp_forward(amp_mlp *mlp, const float *d_x, int32_t batch, float *d_y);
int amp_mlp_gradient(amp_mlp *mlp, const float *d_x, const float *d_t, int32_t batch, double *loss);
int amp_mlp_adam(amp_mlp *mlp, float lr, float beta1, float beta2, float eps);
void amp_mlp_free(amp_mlp *mlp);

/* ---- device assembly of P1 diffusion matrices (FEM assembly feeding the factor) --------------------------------- */

typedef struct amp_asm amp_asm;

/* assembleCells (packages/physics/src/common/scalar.ts:96-144) for P1 simplices (dim 2: tri3, dim 3: tet4) with a
 * per-element coefficient, set up once on the host: the element pattern (buildElementPattern, sparse.mjs:86-110), the
 * element slots (elementSlots, sparse.mjs:113-125), the restriction to the free nodes (restrict, scalar.ts:210-233;
 * freeMask NULL keeps every node) and the geometric element matrices |T| grad N_a . grad N_b. nodes: nn x dim,
 * conn: ne x (dim + 1). */
int amp_asm_create(int32_t dim, int32_t nn, const double *nodes, int32_t ne, const int32_t *conn,
                   const uint8_t *freeMask, amp_asm **out);
/* The restricted pattern (valid while the handle lives): n free DOFs, full symmetric CSR with sorted rows. */
int amp_asm_get_pattern(const amp_asm *a, int32_t *n, int64_t *nnz, const int32_t **rowptr, const int32_t **colidx);
/* Values of nb matrices on the restricted pattern (d_values: nb x nnz, device) from nb coefficient sets (d_k: nb x ne,
 * device): every entry sums k_e Kgeo_e over its element contributions in element order, as the reference's scatter
 * adds them. */
int amp_assemble_device(amp_asm *a, const double *d_k, double *d_values, int32_t nb);
/* Parametric coefficients of the learned sweep: k_e = exp(sum_m theta_m Phi_em) for nb parameter vectors (d_theta:
 * nb x nm, device; d_k: nb x ne, device), with the mode values Phi (ne x nm, host) set once. No counterpart in the
 * reference (its sweeps assemble on the host per point). */
int amp_asm_set_modes(amp_asm *a, int32_t nm, const double *Phi);
int amp_param_coefficients_device(amp_asm *a, const double *d_theta, double *d_k, int32_t nb);
void amp_asm_free(amp_asm *a);

/* Message of the last failure on the calling thread ("" when none). */
const char *amp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
