// lobpcg_solve_double / lobpcg_solve_complex entry points
// (krylov/lobpcg.c:332-498): a block wrapped with the HIP interpreter runs
// the device-resident solver (complex blocks through the real form of their
// pencil); other backends, and blocks wider than the device solver takes, run
// the generic algorithm with the FortranMatrix routines of this library and
// the caller's LAPACK.
#include <cstdio>

#include "common.h"
#include "lobpcg.h"

extern "C" int blopex_amd_is_hip_interpreter(const mv_InterfaceInterpreter *i);
extern "C" mv_InterfaceInterpreter *blopex_amd_mv_interpreter(mv_MultiVectorPtr x);
extern "C" void blopex_amd_default_interpreters(lobpcg_Interpreter *lp,
                                                utilities_FortranInterpreter *u);
extern "C" void blopex_amd_complex_interpreters(lobpcg_Interpreter *lp,
                                                utilities_FortranInterpreter *u);

namespace bx {
int device_lobpcg(mv_MultiVectorPtr X, void *Ad, void (*opA)(void *, void *, void *), void *Bd,
                  void (*opB)(void *, void *, void *), void *Td,
                  void (*opT)(void *, void *, void *), mv_MultiVectorPtr Y, lobpcg_Tolerance tol,
                  int maxit, int verb, int *iters, double *lam, double *lamh, BlopexInt lamh_gh,
                  double *res, double *resh, BlopexInt resh_gh);
bool device_lobpcg_fits(int k, int guard);
int device_lobpcg_complex(mv_MultiVectorPtr X, void *Ad, void (*opA)(void *, void *, void *),
                          void *Bd, void (*opB)(void *, void *, void *), void *Td,
                          void (*opT)(void *, void *, void *), mv_MultiVectorPtr Y,
                          lobpcg_Tolerance tol, int maxit, int verb, int *iters, komplex *lam,
                          komplex *lamh, BlopexInt lamh_gh, double *res, double *resh,
                          BlopexInt resh_gh);
}

extern "C" {

int lobpcg_solve_double(mv_MultiVectorPtr X, void *Ad, void (*opA)(void *, void *, void *),
                        void *Bd, void (*opB)(void *, void *, void *), void *Td,
                        void (*opT)(void *, void *, void *), mv_MultiVectorPtr Y,
                        lobpcg_BLASLAPACKFunctions blap, lobpcg_Tolerance tol, int maxit, int verb,
                        int *iters, double *lam, double *lamh, BlopexInt lamh_gh, double *res,
                        double *resh, BlopexInt resh_gh) {
    if (blopex_amd_is_hip_interpreter(blopex_amd_mv_interpreter(X))) {
        if (((hip_Multi_Vector *)mv_MultiVectorGetData(X))->precision == HIPMV_C64) {
            bx::set_error("lobpcg_solve_double: complex block (use lobpcg_solve_complex)");
            if (verb) fprintf(stderr, "Error in LOBPCG:\n%s\n", blopex_amd_last_error());
            *iters = 0;
            return WRONG_BLOCK_SIZE;
        }
        // blocks wider than the device solver's Rayleigh-Ritz takes run the
        // generic algorithm over the same device blocks when the caller
        // passes dpotrf/dsygv, as the reference accepts any block size
        const int k = mv_MultiVectorWidth(X);
        if (k > bx::kMaxCols) {
            bx::set_error("lobpcg_solve_double: %d columns (device blocks take at most %d)", k,
                          bx::kMaxCols);
            if (verb) fprintf(stderr, "Error in LOBPCG:\n%s\n", blopex_amd_last_error());
            *iters = 0;
            return WRONG_BLOCK_SIZE;
        }
        const int guard = ((hip_Multi_Vector *)mv_MultiVectorGetData(X))->settings.guard_columns;
        if (bx::device_lobpcg_fits(k, guard) || !blap.dpotrf || !blap.dsygv)
            return bx::device_lobpcg(X, Ad, opA, Bd, opB, Td, opT, Y, tol, maxit, verb, iters,
                                     lam, lamh, lamh_gh, res, resh, resh_gh);
    }
    if (!blap.dpotrf || !blap.dsygv) {
        bx::set_error("lobpcg_solve_double: dpotrf/dsygv are required for a host backend");
        if (verb) fprintf(stderr, "Error in LOBPCG:\n%s\n", blopex_amd_last_error());
        *iters = 0;
        return REQUESTED_ACCURACY_NOT_ACHIEVED;
    }
    lobpcg_Interpreter lp;
    utilities_FortranInterpreter u;
    blopex_amd_default_interpreters(&lp, &u);
    return lobpcg_solve(X, Ad, opA, Bd, opB, Td, opT, Y, blap, tol, maxit, verb, iters, lam, lamh,
                        lamh_gh, res, resh, resh_gh, &lp, &u);
}

int lobpcg_solve_complex(mv_MultiVectorPtr X, void *Ad, void (*opA)(void *, void *, void *),
                         void *Bd, void (*opB)(void *, void *, void *), void *Td,
                         void (*opT)(void *, void *, void *), mv_MultiVectorPtr Y,
                         lobpcg_BLASLAPACKFunctions blap, lobpcg_Tolerance tol, int maxit,
                         int verb, int *iters, komplex *lam, komplex *lamh, BlopexInt lamh_gh,
                         double *res, double *resh, BlopexInt resh_gh) {
    // a complex HIP block runs on the device solver through the real form of
    // the pencil (lobpcg_complex.cpp) when twice its width fits; wider blocks
    // and other backends run the generic algorithm (block work through X's
    // interpreter, small dense work through the complex FortranMatrix
    // routines and the caller's zpotrf/zhegv)
    if (blopex_amd_is_hip_interpreter(blopex_amd_mv_interpreter(X))) {
        const hip_Multi_Vector *x = (const hip_Multi_Vector *)mv_MultiVectorGetData(X);
        if (x->precision != HIPMV_C64) {
            bx::set_error("lobpcg_solve_complex: the block is not complex (HIPMV_C64)");
            if (verb) fprintf(stderr, "Error in LOBPCG:\n%s\n", blopex_amd_last_error());
            if (iters) *iters = 0;
            return WRONG_BLOCK_SIZE;
        }
        const int k = (int)x->num_vectors;
        if (bx::device_lobpcg_fits(2 * k, 0) || !blap.zpotrf || !blap.zhegv)
            return bx::device_lobpcg_complex(X, Ad, opA, Bd, opB, Td, opT, Y, tol, maxit, verb,
                                             iters, lam, lamh, lamh_gh, res, resh, resh_gh);
    }
    if (!blap.zpotrf || !blap.zhegv) {
        bx::set_error("lobpcg_solve_complex: zpotrf/zhegv are required");
        if (verb) fprintf(stderr, "Error in LOBPCG:\n%s\n", blopex_amd_last_error());
        if (iters) *iters = 0;
        return REQUESTED_ACCURACY_NOT_ACHIEVED;
    }
    lobpcg_Interpreter lp;
    utilities_FortranInterpreter u;
    blopex_amd_complex_interpreters(&lp, &u);
    return lobpcg_solve(X, Ad, opA, Bd, opB, Td, opT, Y, blap, tol, maxit, verb, iters, lam, lamh,
                        lamh_gh, res, resh, resh_gh, &lp, &u);
}

}  // extern "C"
