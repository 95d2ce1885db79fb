// Generic LOBPCG over the interpreter tables (the algorithm of
// blopex_abstract/krylov/lobpcg.c:502-1388): block work goes through the
// mv_* dispatch of whichever backend wraps X, small dense work through the
// caller's utilities_FortranInterpreter and lobpcg_Interpreter.  The
// device-resident solver (lobpcg_device.cpp) runs the same steps on the GPU
// when X is a hip_Multi_Vector.
//
// Step order, soft locking, symmetrisation points, histories and exit codes
// follow the reference; the one divergence is a rank-deficient P (failed
// Cholesky of P'BP), which here drops P from that iteration's basis and
// rebuilds it from R (the reference keeps stale P-block bounds, lobpcg.c:
// 1056-1063 and 1200).
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "lobpcg.h"

namespace {

using FM = utilities_FortranMatrix;

struct Generic {
    // arguments
    mv_MultiVectorPtr X, Y;
    void *Ad, *Bd, *Td;
    void (*opA)(void *, void *, void *);
    void (*opB)(void *, void *, void *);
    void (*opT)(void *, void *, void *);
    lobpcg_BLASLAPACKFunctions blap;
    lobpcg_Tolerance tol;
    int maxit, verb;
    lobpcg_Interpreter *lp;
    utilities_FortranInterpreter *u;

    int k = 0, ny = 0;
    bool noB = true, noT = true, noY = true;

    FM *mk() { return u->FortranMatrixCreate(); }
    void sel(FM *m, long i0, long i1, long j0, long j1, FM *b) {
        u->FortranMatrixSelectBlock(m, i0, i1, j0, j1, b);
    }
    void gram(mv_MultiVectorPtr a, mv_MultiVectorPtr b, FM *g) {
        mv_MultiVectorByMultiVector(a, b, u->FortranMatrixGlobalHeight(g), u->FortranMatrixHeight(g),
                                    u->FortranMatrixWidth(g), u->FortranMatrixValues(g));
    }
    void vecmat(mv_MultiVectorPtr a, FM *c, mv_MultiVectorPtr b) {
        mv_MultiVectorByMatrix(a, u->FortranMatrixGlobalHeight(c), u->FortranMatrixHeight(c),
                               u->FortranMatrixWidth(c), u->FortranMatrixValues(c), b);
    }
    void apply(void (*op)(void *, void *, void *), void *d, mv_MultiVectorPtr a, mv_MultiVectorPtr b) {
        op(d, mv_MultiVectorGetData(a), mv_MultiVectorGetData(b));
    }
    // B-orthonormalise x (bx = Bx) with the Cholesky factor of x'bx kept in r
    int orthonormalize(mv_MultiVectorPtr x, mv_MultiVectorPtr bx, FM *r, mv_MultiVectorPtr z) {
        gram(x, bx, r);
        int info = lp->chol(r, blap, u);
        if (info) return info;
        u->FortranMatrixUpperInv(r);
        u->FortranMatrixClearL(r);
        mv_MultiVectorCopy(x, z);
        vecmat(z, r, x);
        return 0;
    }
    void error(const char *msg) {
        if (verb) fprintf(stderr, "Error in LOBPCG:\n%s", msg);
    }
    void dump(FM *res, FM *lam, int it) {
        printf("\nEigenvalue lambda       Residual              \n");
        for (int i = 1; i <= k; i++)
            printf("%22.14e  %22.14e\n", *(double *)u->FortranMatrixValuePtr(lam, i, 1),
                   utilities_FortranMatrixAbs(res, i, 1));
        printf("\n%d iterations\n", it);
    }

    int run(int *iters, void *lam_v, void *lamh_v, BlopexInt lamh_gh, double *res_v,
            double *resh_v, BlopexInt resh_gh);
};

int Generic::run(int *iters, void *lam_v, void *lamh_v, BlopexInt lamh_gh, double *res_v,
                 double *resh_v, BlopexInt resh_gh) {
    *iters = 0;
    noT = opT == NULL;
    noB = opB == NULL;
    ny = mv_MultiVectorWidth(Y);
    noY = ny == 0;
    k = mv_MultiVectorWidth(X);
    if (k < 1) {
        error("The bloc size is wrong.\n");
        return WRONG_BLOCK_SIZE;
    }

    FM *lam = mk(), *lamh = NULL, *res = utilities_FortranMatrixCreate(), *resh = NULL;
    u->FortranMatrixWrap(lam_v, k, k, 1, lam);
    if (lamh_v) {
        lamh = mk();
        u->FortranMatrixWrap(lamh_v, lamh_gh, k, maxit + 1, lamh);
    }
    utilities_FortranMatrixWrap(res_v, k, k, 1, res);
    if (resh_v) {
        resh = utilities_FortranMatrixCreate();
        utilities_FortranMatrixWrap(resh_v, resh_gh, k, maxit + 1, resh);
    }
    FM *rdiag = mk();
    u->FortranMatrixAllocateData(k, 1, rdiag);

    mv_MultiVectorPtr W = mv_MultiVectorCreateCopy(X, 0), BY = Y;
    FM *yby = mk(), *ybx = mk(), *tybx = mk(), *ybr = mk(), *tybr = mk();
    int flag = 0;

    auto release_small = [&]() {
        u->FortranMatrixDestroy(yby);
        u->FortranMatrixDestroy(ybx);
        u->FortranMatrixDestroy(tybx);
        u->FortranMatrixDestroy(ybr);
        u->FortranMatrixDestroy(tybr);
        u->FortranMatrixDestroy(lam);
        u->FortranMatrixDestroy(lamh);
        utilities_FortranMatrixDestroy(res);
        utilities_FortranMatrixDestroy(resh);
        u->FortranMatrixDestroy(rdiag);
    };

    if (!noY) {
        u->FortranMatrixAllocateData(ny, ny, yby);
        u->FortranMatrixAllocateData(ny, k, ybx);
        u->FortranMatrixAllocateData(ny, k, tybx);
        if (!noB) {
            BY = mv_MultiVectorCreateCopy(Y, 0);
            apply(opB, Bd, Y, BY);
        }
        gram(BY, Y, yby);
        if (lp->chol(yby, blap, u)) {
            if (verb) printf("Cannot handle linear dependent constraints\n");
            if (!noB) mv_MultiVectorDestroy(BY);
            mv_MultiVectorDestroy(W);
            release_small();
            return WRONG_CONSTRAINTS;
        }
        u->FortranMatrixUpperInv(yby);
        u->FortranMatrixClearL(yby);
        // X -= Y (Y'BY)^{-1} (BY)'X
        gram(BY, X, ybx);
        u->FortranMatrixMultiply(yby, 1, ybx, 0, tybx);
        u->FortranMatrixMultiply(yby, 0, tybx, 0, ybx);
        vecmat(Y, ybx, W);
        mv_MultiVectorAxpy(-1.0, W, X);
    }

    if (verb) {
        printf("\nSolving %s eigenvalue problem with%s preconditioning\n\n",
               noB ? "standard" : "generalized", noT ? "out" : "");
        printf("block size %d\n\n", k);
        if (noY)
            printf("No constraints\n\n");
        else
            printf("%d constraint%s\n\n", ny, ny > 1 ? "s" : "");
    }

    FM *gA = mk(), *gB = mk(), *lamAB = mk(), *lamX = mk();
    FM *xax = mk(), *rax = mk(), *pax = mk(), *rar = mk(), *par = mk(), *pap = mk();
    FM *xbx = mk(), *rbx = mk(), *pbx = mk(), *rbr = mk(), *pbr = mk(), *pbp = mk();
    FM *cx = mk(), *cxx = mk(), *crx = mk(), *cpx = mk(), *upper = mk();
    FM *lamcol = utilities_FortranMatrixCreate(), *rescol = utilities_FortranMatrixCreate();

    BlopexInt *mask = (BlopexInt *)calloc(k, sizeof(BlopexInt));
    for (int i = 0; i < k; i++) mask[i] = 1;
    const int k3 = 3 * k;
    u->FortranMatrixAllocateData(k3, k3, gA);
    u->FortranMatrixAllocateData(k3, k3, gB);
    u->FortranMatrixAllocateData(k3, 1, lamAB);

    mv_MultiVectorPtr R = mv_MultiVectorCreateCopy(X, 0), P = mv_MultiVectorCreateCopy(X, 0);
    mv_MultiVectorPtr AX = mv_MultiVectorCreateCopy(X, 0), AR = mv_MultiVectorCreateCopy(X, 0),
                      AP = mv_MultiVectorCreateCopy(X, 0);
    mv_MultiVectorPtr BX = X, BR = R, BP = P;
    if (!noB) {
        BX = mv_MultiVectorCreateCopy(X, 0);
        BR = mv_MultiVectorCreateCopy(X, 0);
        BP = mv_MultiVectorCreateCopy(X, 0);
    }
    auto mask_iter_blocks = [&](BlopexInt *m) {
        mv_MultiVectorSetMask(R, m);
        mv_MultiVectorSetMask(P, m);
        mv_MultiVectorSetMask(AR, m);
        mv_MultiVectorSetMask(AP, m);
        if (!noB) {
            mv_MultiVectorSetMask(BR, m);
            mv_MultiVectorSetMask(BP, m);
        }
        mv_MultiVectorSetMask(W, m);
    };
    mask_iter_blocks(mask);

    auto store_hist = [&](int col) {
        if (lamh) {
            sel(lamh, 1, k, col, col, lamcol);
            u->FortranMatrixCopy(lam, 0, lamcol);
        }
        if (resh) {
            utilities_FortranMatrixSelectBlock(resh, 1, k, col, col, rescol);
            utilities_FortranMatrixCopy(res, 0, rescol);
        }
    };
    auto norms = [&](BlopexInt *m) {
        mv_MultiVectorByMultiVectorDiag(R, R, m, k, u->FortranMatrixValues(rdiag));
        double *rp = (double *)res->value;
        for (int i = 0; i < k; i++)
            if (!m || m[i]) rp[i] = sqrt(u->FortranMatrixAbs(rdiag, i + 1, 1));
    };

    // ---- initial Rayleigh-Ritz in span(X) ----
    sel(gB, 1, k, 1, k, upper);
    if (!noB) apply(opB, Bd, X, BX);
    flag = orthonormalize(X, BX, upper, W);
    if (flag) {
        error("Bad initial vectors: orthonormalization failed\n");
        if (verb) printf("DPOTRF INFO = %d\n", flag);
    } else {
        if (!noB) {
            vecmat(BX, upper, W);
            mv_MultiVectorCopy(W, BX);
        }
        apply(opA, Ad, X, AX);
        sel(gA, 1, k, 1, k, xax);
        gram(X, AX, xax);
        u->FortranMatrixSymmetrize(xax);
        sel(gB, 1, k, 1, k, xbx);
        gram(X, BX, xbx);
        u->FortranMatrixSymmetrize(xbx);
        if ((flag = lp->solveGEVP(xax, xbx, lam, blap, u)) != 0) {
            error("Bad problem: Rayleigh-Ritz in the initial subspace failed\n");
            if (verb) printf("DSYGV INFO = %d\n", flag);
        } else {
            sel(xax, 1, k, 1, k, cx);
            vecmat(X, cx, W);
            mv_MultiVectorCopy(W, X);
            vecmat(AX, cx, W);
            mv_MultiVectorCopy(W, AX);
            if (!noB) {
                vecmat(BX, cx, W);
                mv_MultiVectorCopy(W, BX);
            }
            mv_MultiVectorByDiagonal(BX, NULL, k, u->FortranMatrixValues(lam), R);
            mv_MultiVectorAxpy(-1.0, AX, R);
            norms(NULL);
            store_hist(1);
            if (verb == 2)
                dump(res, lam, *iters);
            else if (verb == 1)
                printf("\nInitial Max. Residual %22.14e\n", utilities_FortranMatrixMaxValue(res));
        }
    }

    // ---- main loop ----
    int it;
    for (it = 1; flag == 0 && it <= maxit; it++) {
        int nr = 0;
        for (int i = 0; i < k; i++) {
            mask[i] = utilities_FortranMatrixAbs(res, i + 1, 1) >
                      u->FortranMatrixAbs(lam, i + 1, 1) * tol.relative + tol.absolute + DBL_EPSILON;
            nr += mask[i];
        }
        if (nr < 1) break;
        mask_iter_blocks(mask);

        if (!noT) {
            apply(opT, Td, R, W);
            mv_MultiVectorCopy(W, R);
        }
        if (!noY) {
            sel(ybx, 1, ny, 1, nr, ybr);
            sel(tybx, 1, ny, 1, nr, tybr);
            gram(BY, R, ybr);
            u->FortranMatrixMultiply(yby, 1, ybr, 0, tybr);
            u->FortranMatrixMultiply(yby, 0, tybr, 0, ybr);
            vecmat(Y, ybr, W);
            mv_MultiVectorAxpy(-1.0, W, R);
        }
        const long r0 = k + 1, r1 = k + nr, p0 = r1 + 1;
        sel(gB, r0, r1, r0, r1, upper);
        if (!noB) apply(opB, Bd, R, BR);
        flag = orthonormalize(R, BR, upper, W);
        if (flag) {
            error("Orthonormalization of residuals failed\n");
            if (verb) printf("DPOTRF INFO = %d\n", flag);
            break;
        }
        if (!noB) {
            vecmat(BR, upper, W);
            mv_MultiVectorCopy(W, BR);
        }
        apply(opA, Ad, R, AR);

        long np = 0;
        if (it > 1) {
            np = nr;
            sel(gB, p0, r1 + np, p0, r1 + np, upper);
            if (orthonormalize(P, BP, upper, W)) {
                np = 0;
            } else {
                if (!noB) {
                    vecmat(BP, upper, W);
                    mv_MultiVectorCopy(W, BP);
                }
                vecmat(AP, upper, W);
                mv_MultiVectorCopy(W, AP);
            }
        }
        const long p1 = r1 + np, na = r1 + np;

        sel(gA, 1, k, 1, k, xax);
        sel(gA, r0, r1, 1, k, rax);
        sel(gA, r0, r1, r0, r1, rar);
        sel(gB, 1, k, 1, k, xbx);
        sel(gB, r0, r1, 1, k, rbx);
        sel(gB, r0, r1, r0, r1, rbr);
        u->FortranMatrixClear(xax);
        u->FortranMatrixSetDiagonal(xax, lam);
        gram(R, AX, rax);
        gram(R, AR, rar);
        u->FortranMatrixSymmetrize(rar);
        u->FortranMatrixSetToIdentity(xbx);
        gram(R, BX, rbx);
        u->FortranMatrixSetToIdentity(rbr);
        if (np > 0) {
            sel(gA, p0, p1, 1, k, pax);
            sel(gA, p0, p1, r0, r1, par);
            sel(gA, p0, p1, p0, p1, pap);
            sel(gB, p0, p1, 1, k, pbx);
            sel(gB, p0, p1, r0, r1, pbr);
            sel(gB, p0, p1, p0, p1, pbp);
            gram(P, AX, pax);
            gram(P, AR, par);
            gram(P, AP, pap);
            u->FortranMatrixSymmetrize(pap);
            gram(P, BX, pbx);
            gram(P, BR, pbr);
            u->FortranMatrixSetToIdentity(pbp);
        }
        sel(gA, 1, na, 1, na, xax);
        sel(gB, 1, na, 1, na, xbx);
        if ((flag = lp->solveGEVP(xax, xbx, lamAB, blap, u)) != 0) {
            error("GEVP solver failure\n");
            if (verb) printf("INFO = %d\n", flag);
            it--;
            break;
        }
        sel(lamAB, 1, k, 1, 1, lamX);
        u->FortranMatrixCopy(lamX, 0, lam);
        sel(gA, 1, na, 1, k, cx);
        sel(cx, 1, k, 1, k, cxx);
        sel(cx, r0, r1, 1, k, crx);

        if (np > 0) {
            sel(cx, p0, p1, 1, k, cpx);
            mv_MultiVectorSetMask(W, NULL);
            vecmat(P, cpx, W);
            mv_MultiVectorSetMask(P, NULL);
            mv_MultiVectorCopy(W, P);
            vecmat(AP, cpx, W);
            mv_MultiVectorSetMask(AP, NULL);
            mv_MultiVectorCopy(W, AP);
            if (!noB) {
                vecmat(BP, cpx, W);
                mv_MultiVectorSetMask(BP, NULL);
                mv_MultiVectorCopy(W, BP);
            }
            vecmat(R, crx, W);
            mv_MultiVectorAxpy(1.0, W, P);
            vecmat(AR, crx, W);
            mv_MultiVectorAxpy(1.0, W, AP);
            if (!noB) {
                vecmat(BR, crx, W);
                mv_MultiVectorAxpy(1.0, W, BP);
            }
        } else {
            mv_MultiVectorSetMask(P, NULL);
            vecmat(R, crx, P);
            mv_MultiVectorSetMask(AP, NULL);
            vecmat(AR, crx, AP);
            if (!noB) {
                mv_MultiVectorSetMask(BP, NULL);
                vecmat(BR, crx, BP);
            }
        }
        mv_MultiVectorSetMask(W, NULL);
        mv_MultiVectorCopy(X, W);
        vecmat(W, cxx, X);
        mv_MultiVectorAxpy(1.0, P, X);
        mv_MultiVectorCopy(AX, W);
        vecmat(W, cxx, AX);
        mv_MultiVectorAxpy(1.0, AP, AX);
        if (!noB) {
            mv_MultiVectorCopy(BX, W);
            vecmat(W, cxx, BX);
            mv_MultiVectorAxpy(1.0, BP, BX);
        }
        mv_MultiVectorSetMask(AX, mask);
        mv_MultiVectorSetMask(BX, mask);
        mv_MultiVectorByDiagonal(BX, mask, k, u->FortranMatrixValues(lam), R);
        mv_MultiVectorAxpy(-1.0, AX, R);
        norms(mask);
        store_hist(it + 1);
        if (verb == 2) {
            printf("Iteration %d \tbsize %d\n", it, nr);
            dump(res, lam, it);
        } else if (verb == 1) {
            printf("Iteration %d \tbsize %d \tmaxres %22.14e\n", it, nr,
                   utilities_FortranMatrixMaxValue(res));
        }
        mv_MultiVectorSetMask(AX, NULL);
        mv_MultiVectorSetMask(BX, NULL);
        mv_MultiVectorSetMask(AP, mask);
        mv_MultiVectorSetMask(BP, mask);
        mv_MultiVectorSetMask(P, mask);
        mv_MultiVectorSetMask(W, mask);
    }
    if (flag != 0 || it > maxit) flag = REQUESTED_ACCURACY_NOT_ACHIEVED;
    *iters = it - 1;
    if (verb == 1) dump(res, lam, *iters);

    mv_MultiVectorDestroy(R);
    mv_MultiVectorDestroy(P);
    mv_MultiVectorDestroy(AX);
    mv_MultiVectorDestroy(AR);
    mv_MultiVectorDestroy(AP);
    if (!noB) {
        mv_MultiVectorDestroy(BX);
        mv_MultiVectorDestroy(BR);
        mv_MultiVectorDestroy(BP);
        if (!noY) mv_MultiVectorDestroy(BY);
    }
    mv_MultiVectorDestroy(W);
    FM *all[] = {gA, gB, lamAB, lamX, xax, rax, pax, rar, par, pap, xbx, rbx, pbx, rbr, pbr, pbp,
                 cx, cxx, crx, cpx, upper, lamcol, rescol};
    for (FM *m : all) u->FortranMatrixDestroy(m);
    release_small();
    free(mask);
    return flag;
}

// ---- the small-dense routines lobpcg_solve_double binds ----
BlopexInt chol_lapack(utilities_FortranMatrix *a, lobpcg_BLASLAPACKFunctions blap,
                      utilities_FortranInterpreter *u) {
    char uplo = 'U';
    BlopexInt n = (BlopexInt)u->FortranMatrixHeight(a), lda = (BlopexInt)u->FortranMatrixGlobalHeight(a);
    BlopexInt info = 0;
    blap.dpotrf(&uplo, &n, (double *)u->FortranMatrixValues(a), &lda, &info);
    return info;
}

BlopexInt gevp_lapack(utilities_FortranMatrix *A, utilities_FortranMatrix *B,
                      utilities_FortranMatrix *lam, lobpcg_BLASLAPACKFunctions blap,
                      utilities_FortranInterpreter *u) {
    BlopexInt itype = 1, n = (BlopexInt)u->FortranMatrixHeight(A);
    BlopexInt lda = (BlopexInt)u->FortranMatrixGlobalHeight(A), ldb = (BlopexInt)u->FortranMatrixGlobalHeight(B);
    BlopexInt lwork = 10 * n, info = 0;
    char jobz = 'V', uplo = 'L';
    double *work = (double *)calloc(lwork, sizeof(double));
    blap.dsygv(&itype, &jobz, &uplo, &n, (double *)u->FortranMatrixValues(A), &lda,
               (double *)u->FortranMatrixValues(B), &ldb, (double *)u->FortranMatrixValues(lam),
               work, &lwork, &info);
    free(work);
    return info;
}

}  // namespace

extern "C" {

// Fills the FP64 small-dense interpreters lobpcg_solve_double uses
// (lobpcg.c:357-387): the FortranMatrix routines of this library and
// chol/solveGEVP calling blap_fn.dpotrf/dsygv.
void blopex_amd_default_interpreters(lobpcg_Interpreter *lp, utilities_FortranInterpreter *u) {
    u->FortranMatrixCreate = utilities_FortranMatrixCreate;
    u->FortranMatrixAllocateData = utilities_FortranMatrixAllocateData;
    u->FortranMatrixWrap = utilities_FortranMatrixWrap;
    u->FortranMatrixDestroy = utilities_FortranMatrixDestroy;
    u->FortranMatrixGlobalHeight = utilities_FortranMatrixGlobalHeight;
    u->FortranMatrixHeight = utilities_FortranMatrixHeight;
    u->FortranMatrixWidth = utilities_FortranMatrixWidth;
    u->FortranMatrixValues = utilities_FortranMatrixValues;
    u->FortranMatrixClear = utilities_FortranMatrixClear;
    u->FortranMatrixClearL = utilities_FortranMatrixClearL;
    u->FortranMatrixSetToIdentity = utilities_FortranMatrixSetToIdentity;
    u->FortranMatrixTransposeSquare = utilities_FortranMatrixTransposeSquare;
    u->FortranMatrixSymmetrize = utilities_FortranMatrixSymmetrize;
    u->FortranMatrixCopy = utilities_FortranMatrixCopy;
    u->FortranMatrixIndexCopy = utilities_FortranMatrixIndexCopy;
    u->FortranMatrixSetDiagonal = utilities_FortranMatrixSetDiagonal;
    u->FortranMatrixGetDiagonal = utilities_FortranMatrixGetDiagonal;
    u->FortranMatrixAdd = utilities_FortranMatrixAdd;
    u->FortranMatrixDMultiply = utilities_FortranMatrixDMultiply;
    u->FortranMatrixMultiplyD = utilities_FortranMatrixMultiplyD;
    u->FortranMatrixMultiply = utilities_FortranMatrixMultiply;
    u->FortranMatrixFNorm = utilities_FortranMatrixFNorm;
    u->FortranMatrixAbs = utilities_FortranMatrixAbs;
    u->FortranMatrixValuePtr = utilities_FortranMatrixValuePtr;
    u->FortranMatrixMaxValue = utilities_FortranMatrixMaxValue;
    u->FortranMatrixSelectBlock = utilities_FortranMatrixSelectBlock;
    u->FortranMatrixUpperInv = utilities_FortranMatrixUpperInv;
    u->FortranMatrixPrint = utilities_FortranMatrixPrint;
    lp->chol = chol_lapack;
    lp->solveGEVP = gevp_lapack;
}

int lobpcg_solve(mv_MultiVectorPtr X, void *Ad, void (*opA)(void *, void *, void *), void *Bd,
                 void (*opB)(void *, void *, void *), void *Td, void (*opT)(void *, void *, void *),
                 mv_MultiVectorPtr Y, lobpcg_BLASLAPACKFunctions blap, lobpcg_Tolerance tol,
                 int maxit, int verb, int *iters, void *lam, void *lamh, BlopexInt lamh_gh,
                 double *res, double *resh, BlopexInt resh_gh, lobpcg_Interpreter *lp,
                 utilities_FortranInterpreter *u) {
    Generic g{X, Y, Ad, Bd, Td, opA, opB, opT, blap, tol, maxit, verb, lp, u};
    return g.run(iters, lam, lamh, lamh_gh, res, resh, resh_gh);
}

}  // extern "C"
