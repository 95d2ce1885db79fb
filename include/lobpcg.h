/*
 * LOBPCG entry points, tolerance / LAPACK / small-dense interpreter structs
 * and exit codes of the BLOPEX C API (blopex_abstract/krylov/lobpcg.h:13-168).
 *
 * lobpcg_solve_double on a block built by HipSetupInterpreter
 * (hip_multivector.h) runs the device-resident solver: all block work and
 * the FP64 Cholesky / Rayleigh-Ritz stay on the GPU, and blap_fn may be
 * zero-filled.  On any other backend it runs the generic host algorithm
 * through the interpreter table and the caller's dpotrf/dsygv, as the
 * reference does.  lobpcg_solve always runs the generic algorithm with the
 * caller's small-dense interpreters.
 */
#ifndef BLOPEX_AMD_LOBPCG_H
#define BLOPEX_AMD_LOBPCG_H

#include "fortran_interpreter.h"
#include "fortran_matrix.h"
#include "multivector.h"

#define PROBLEM_SIZE_TOO_SMALL 1
#define WRONG_BLOCK_SIZE 2
#define WRONG_CONSTRAINTS 3
#define REQUESTED_ACCURACY_NOT_ACHIEVED -1

typedef struct {
    double absolute;
    double relative;
} lobpcg_Tolerance;

/* caller-supplied LAPACK (Fortran calling convention) */
typedef struct {
    BlopexInt (*dpotrf)(char *uplo, BlopexInt *n, double *a, BlopexInt *lda, BlopexInt *info);
    BlopexInt (*dsygv)(BlopexInt *itype, char *jobz, char *uplo, BlopexInt *n, double *a,
                       BlopexInt *lda, double *b, BlopexInt *ldb, double *w, double *work,
                       BlopexInt *lwork, BlopexInt *info);
    BlopexInt (*zpotrf)(char *uplo, BlopexInt *n, komplex *a, BlopexInt *lda, BlopexInt *info);
    BlopexInt (*zhegv)(BlopexInt *itype, char *jobz, char *uplo, BlopexInt *n, komplex *a,
                       BlopexInt *lda, komplex *b, BlopexInt *ldb, double *w, komplex *work,
                       BlopexInt *lwork, double *rwork, BlopexInt *info);
} lobpcg_BLASLAPACKFunctions;

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    BlopexInt (*chol)(utilities_FortranMatrix *a, lobpcg_BLASLAPACKFunctions blap_fn,
                      utilities_FortranInterpreter *util);
    BlopexInt (*solveGEVP)(utilities_FortranMatrix *mtxA, utilities_FortranMatrix *mtxB,
                           utilities_FortranMatrix *eigVal, lobpcg_BLASLAPACKFunctions blap_fn,
                           utilities_FortranInterpreter *util);
} lobpcg_Interpreter;

/*
 * Smallest eigenpairs of A x = lambda B x by LOBPCG.  blockVectorX (n x k) is
 * the initial block on entry and the eigenvectors on return; operatorA/B/T
 * are op(data, X, Y) callbacks on the backend blocks (B NULL: identity,
 * T NULL: no preconditioner); blockVectorY holds constraints (NULL: none).
 * lambda_values and residualNorms_values hold k doubles; the histories are
 * optional (NULL) column-major k x (maxIterations+1) with leading dimension
 * *_gh.  Returns 0, WRONG_BLOCK_SIZE, WRONG_CONSTRAINTS or
 * REQUESTED_ACCURACY_NOT_ACHIEVED.
 */
int lobpcg_solve_double(mv_MultiVectorPtr blockVectorX, void *operatorAData,
                        void (*operatorA)(void *, void *, void *), void *operatorBData,
                        void (*operatorB)(void *, void *, void *), void *operatorTData,
                        void (*operatorT)(void *, void *, void *), mv_MultiVectorPtr blockVectorY,
                        lobpcg_BLASLAPACKFunctions blap_fn, lobpcg_Tolerance tolerance,
                        int maxIterations, int verbosityLevel, int *iterationNumber,
                        double *lambda_values, double *lambdaHistory_values,
                        BlopexInt lambdaHistory_gh, double *residualNorms_values,
                        double *residualNormsHistory_values, BlopexInt residualNormsHistory_gh);

/* Complex Hermitian problems (lobpcg.c:417-498).  A block wrapped with the
   HIP interpreter must be HIPMV_C64 (WRONG_BLOCK_SIZE otherwise); up to 32
   columns it runs on the device solver through the real symmetric form of
   the pencil (order 2n, each eigenvalue twice; the operators still see
   complex blocks, HipDenseMatMultiVec being the dense one), and blap_fn is
   not used.  Wider blocks and other backends run the generic algorithm with
   the complex FortranMatrix routines (zutilities_*) and blap_fn.zpotrf /
   zhegv (required there: -1 without them).  lambda and its history are
   komplex with zero imaginary parts. */
int lobpcg_solve_complex(mv_MultiVectorPtr blockVectorX, void *operatorAData,
                         void (*operatorA)(void *, void *, void *), void *operatorBData,
                         void (*operatorB)(void *, void *, void *), void *operatorTData,
                         void (*operatorT)(void *, void *, void *), mv_MultiVectorPtr blockVectorY,
                         lobpcg_BLASLAPACKFunctions blap_fn, lobpcg_Tolerance tolerance,
                         int maxIterations, int verbosityLevel, int *iterationNumber,
                         komplex *lambda_values, komplex *lambdaHistory_values,
                         BlopexInt lambdaHistory_gh, double *residualNorms_values,
                         double *residualNormsHistory_values, BlopexInt residualNormsHistory_gh);

/* Generic algorithm with caller-supplied small-dense interpreters. */
int lobpcg_solve(mv_MultiVectorPtr blockVectorX, void *operatorAData,
                 void (*operatorA)(void *, void *, void *), void *operatorBData,
                 void (*operatorB)(void *, void *, void *), void *operatorTData,
                 void (*operatorT)(void *, void *, void *), mv_MultiVectorPtr blockVectorY,
                 lobpcg_BLASLAPACKFunctions blap_fn, lobpcg_Tolerance tolerance, int maxIterations,
                 int verbosityLevel, int *iterationNumber, void *lambda_values,
                 void *lambdaHistory_values, BlopexInt lambdaHistory_gh,
                 double *residualNorms_values, double *residualNormsHistory_values,
                 BlopexInt residualNormsHistory_gh, lobpcg_Interpreter *lobpcg,
                 utilities_FortranInterpreter *util);

#ifdef __cplusplus
}
#endif

#endif
