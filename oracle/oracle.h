/*
 * CPU oracle for the supernodal Cholesky / LDL^T path. TEST INFRASTRUCTURE ONLY: tests, the smoke check and the CPU
 * baseline leg of bench.py call it as the checker; the shipped library never links or calls it.
 *
 * It restates, in plain C, the algorithms of jwt625/applied-multiphysics packages/core/src/la:
 *   - symbolic analysis: amdOrder (sparse.mjs:297-592), postorder (:597-625), etreeOf (:628-645), analyze (:666-902),
 *     patternFingerprint (:908-925);
 *   - numeric multifrontal Cholesky: factorNumeric + panelUpdate (sparse.mjs:941-1158) with the per-entry operation
 *     order of wasm/front-real.mjs:9-16;
 *   - Cholesky solves: belowAccumulate, dotBelow, solveNumeric, supernodalSolve (sparse.mjs:1161-1310);
 *   - SupernodalLDLT factor / solve / solveMany / solveRefined (ldlt.mjs:253-275, 417-643);
 *   - UpLookingCholesky (sparse.mjs:1379-1544), the reference's own test oracle for the structure;
 *   - restarted right-preconditioned GMRES (gmres-real.mjs:31-104).
 * Built with -ffp-contract=off and no fast-math, so every product and sum is rounded exactly where the reference
 * rounds it (JavaScript and WebAssembly never fuse a multiply-add).
 *
 * Parity pins (each function is checked by tests/ against something it could get wrong):
 *   amd / analyze: valid permutation; colCount equal to the up-looking structure; supernode rows own-first and sorted;
 *     nnzLStored identity without relaxation; brute-force symbolic fill on tiny patterns; the exact AMD permutation of
 *     the reference's JS run is NOT pinned here (the reference needs Node >= 23.6): "parity unpinned" for perm bits.
 *   factor / solve: closed-form tridiagonal Cholesky, dense LAPACK factor on small matrices, residual < 1e-12,
 *     agreement with the up-looking oracle, 2A -> x/2, Sylvester inertia, failure semantics.
 *   gmres: nonsymmetric chain to 1e-11 (core-solvers.test.mjs:102-115).
 */
#ifndef AMP_ORACLE_H
#define AMP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Symbolic analysis, field for field the object analyze() returns (sparse.mjs:873-901). Lxp is int64 here (the
 * reference keeps it in a Float64Array; all offsets are integers below 2^53). */
typedef struct {
  int32_t n, nsuper, maxFront;
  int64_t nnzL, nnzA, nnzLStored;
  double flops, fill;
  int32_t *perm, *pinv, *parent, *colCount;
  int32_t *lowerP, *lowerI, *lowerSrc;
  int32_t *superStart, *colSuper, *superParent, *childHead, *childNext;
  int32_t *Sp, *Si;
  int64_t *Lxp;
  char patternKey[64];
} orc_sym;

/* amdOrder (sparse.mjs:297-592). perm: new -> old, length n. Returns 0, or -1 on internal error. */
int orc_amd(int32_t n, const int32_t *rowptr, const int32_t *colidx, double denseFactor, int32_t *perm);

/* ndOrder (sparse.mjs:230-283): recursive coordinate bisection, perm new -> old. Pinned by a hand-derived order of a
 * 100-node path (split at half the length, separator from the first half, first half / second half / separator
 * order, leaves of `leaf` nodes), the axis rule on a vertical path and equal extents, the stability of ties and the
 * fill of grids against the natural order (tests/test_oracle.py). Returns 0, or -1 on allocation failure. */
int orc_nd_order(int32_t n, const int32_t *rowptr, const int32_t *colidx, const double *x, const double *y,
                 int32_t leaf, int32_t *perm);

/* analyze (sparse.mjs:666-902). perm0 may be NULL (AMD). relax: relaxed amalgamation. NULL on allocation failure. */
orc_sym *orc_analyze(int32_t n, const int32_t *rowptr, const int32_t *colidx, const int32_t *perm0, int relax);
void orc_sym_free(orc_sym *s);

/* patternFingerprint (sparse.mjs:908-925): writes "n:nnz:hex" into out (>= 48 bytes). */
void orc_fingerprint(int32_t n, const int32_t *rowptr, const int32_t *colidx, char *out);

/* Threads of the numeric factorization (OpenMP; rows and target columns of a front are split among them without
 * changing any entry's operation order, so the factor is bitwise the same for every thread count). */
void orc_set_threads(int32_t k);
int32_t orc_get_threads(void);

/* factorNumeric (sparse.mjs:941-1125): Lx has nnzLStored entries. Returns 1 when positive definite, else 0. */
int orc_chol_factor(const orc_sym *s, const double *values, double *Lx);
/* factorNumeric restricted to the complete subtree of supernode `root` (a contiguous postorder range); flops of its
 * columns (sum colCount^2) to *flops. Used to time a bounded sample of a large factorization. */
int orc_chol_factor_subtree(const orc_sym *s, const double *values, int32_t root, double *Lx, double *flops);
/* solveNumeric (sparse.mjs:1261-1305); b and x in the original ordering; x may alias nothing. */
void orc_chol_solve(const orc_sym *s, const double *Lx, const double *b, double *x);
/* solveNumericMulti (sparse.mjs:1308): q vectors stored one after another (same per-vector arithmetic). */
void orc_chol_solve_multi(const orc_sym *s, const double *Lx, const double *b, int32_t q, double *x);

/* SupernodalLDLT.factor (ldlt.mjs:450-562). Returns 1 ok, 0 on a non-finite pivot; counts up to the failure. */
int orc_ldlt_factor(const orc_sym *s, const double *values, double pivotTol, double *Lx, double *D, int32_t *nNeg,
                    int32_t *nPert);
/* SupernodalLDLT.solve (ldlt.mjs:613-643), one vector. */
void orc_ldlt_solve(const orc_sym *s, const double *Lx, const double *D, const double *b, double *x);
/* SupernodalLDLT.solveMany (ldlt.mjs:565-611), q vectors one after another (note the 1/D scaling for q > 1). */
void orc_ldlt_solve_many(const orc_sym *s, const double *Lx, const double *D, const double *b, int32_t q, double *x);
/* solveRefined (ldlt.mjs:253-275) on the LDL^T factor: CSR values (unperturbed), b -> x. Returns steps taken. */
int32_t orc_ldlt_solve_refined(const orc_sym *s, const double *Lx, const double *D, const int32_t *rowptr,
                               const int32_t *colidx, const double *values, const double *b, int32_t maxSteps,
                               double tol, double *x, double *relResidual);

/* UpLookingCholesky (sparse.mjs:1379-1544): structure (colCount via ereach) and numeric factor / solve. */
typedef struct {
  int32_t n;
  int32_t *perm, *pinv, *Cp, *Ci, *srcIndex, *parent, *Lp, *Li;
  double *Lx;
  int64_t nnzL;
} orc_uplook;
orc_uplook *orc_uplook_create(int32_t n, const int32_t *rowptr, const int32_t *colidx, const int32_t *perm);
int orc_uplook_factor(orc_uplook *u, const double *values);
void orc_uplook_solve(const orc_uplook *u, const double *b, double *x);
void orc_uplook_free(orc_uplook *u);

/* gmres (gmres-real.mjs:31-104): restarted right-preconditioned GMRES with MGS and Givens rotations. */
typedef void (*orc_apply_fn)(void *ctx, const double *v, double *out);
typedef struct {
  int32_t iterations, converged;
  double relres;
} orc_gmres_info;
void orc_gmres(int32_t n, orc_apply_fn matvec, void *mctx, orc_apply_fn precond, void *pctx, const double *b,
               double tol, int32_t restart, int32_t maxIter, const double *x0, double *x, orc_gmres_info *info);

/* CSR y = A x in row order (sparse.mjs matVec :220-229). */
void orc_csr_matvec(int32_t n, const int32_t *rowptr, const int32_t *colidx, const double *val, const double *x,
                    double *y);

/* P1 diffusion assembly (buildElementPattern sparse.mjs:86-110, elementSlots :113-125, assembleCells scalar.ts:96-144,
 * restrict scalar.ts:210-233): nodes nn x dim, conn ne x (dim + 1), kel per element, freeMask NULL = all free. Writes the
 * restricted pattern and values (rowptr: free + 1, colidx / values: capacity ne (dim + 1)^2 + nn) and returns nnz. */
int64_t orc_assemble_p1(int32_t dim, int32_t nn, const double *nodes, int32_t ne, const int32_t *conn,
                        const double *kel, const uint8_t *freeMask, int32_t *rowptr, int32_t *colidx,
                        double *values);

/* ComplexSupernodalLDLT (ldlt.mjs:648-866): factor (Lr, Li: nnzLStored; Dr, Di: n; returns 1 ok, 0 on a non-finite
 * pivot), solve (:780-829) and complex iterative refinement (:838-862, returns the steps taken). */
int orc_cldlt_factor(const orc_sym *s, const double *re, const double *im, double pivotTol, double *Lr, double *Li,
                     double *Dr, double *Di, int32_t *nPert);
void orc_cldlt_solve(const orc_sym *s, const double *Lr, const double *Li, const double *Dr, const double *Di,
                     const double *br, const double *bi, double *xr, double *xi);
int32_t orc_cldlt_solve_refined(const orc_sym *s, const double *Lr, const double *Li, const double *Dr,
                                const double *Di, const int32_t *rowptr, const int32_t *colidx, const double *re,
                                const double *im, const double *br, const double *bi, int32_t maxSteps, double tol,
                                double *xr, double *xi, double *rel);

#ifdef __cplusplus
}
#endif

#endif
