/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Drives the reference's own lobpcg_solve_double (compiled unmodified from
 * /root/reference by oracle/Makefile into oracle/_ref/) with the reference's
 * serial multivector backend (blopex_serial_double/pcg_multi/pcg_multi.c:
 * 203-242, SerialSetupInterpreter).  The reference has no CSR operator, so
 * this file supplies CSR-A, diagonal-B and diagonal-T callbacks with the
 * operator signature void op(void*, void*, void*) that lobpcg_solve takes
 * (krylov/lobpcg.h:68-81) and that honour the active-column list
 * (multivector/multi_vector.h:22-32), the way MatMultiVec does
 * (pcg_multi.c:167-193).  LAPACK comes in as function pointers, as in
 * lobpcg_BLASLAPACKFunctions (krylov/lobpcg.h:25-46).
 */
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "interpreter.h"
#include "lobpcg.h"
#include "multi_vector.h"
#include "multivector.h"
#include "pcg_multi.h"

typedef struct {
    long n;
    const int *rowptr;
    const int *colind;
    const double *val;
} rh_csr;

typedef struct {
    long n;
    const double *d;
} rh_diag;

/* CLOCK_MONOTONIC at the start of each operator-A call of the last solve:
   one call for the initial AX, then one per iteration (on R), so the
   difference of consecutive stamps after the first is the time of one
   iteration */
#define RH_MAX_STAMPS 64
static double rh_stamps[RH_MAX_STAMPS];
static int rh_nstamps = 0;

int ref_operator_stamps(double *out, int max)
{
    int m = rh_nstamps < max ? rh_nstamps : max;
    for (int i = 0; i < m; i++) out[i] = rh_stamps[i];
    return m;
}

static void rh_csr_apply(void *A_, void *X_, void *Y_)
{
    rh_csr *A = (rh_csr *)A_;
    if (rh_nstamps < RH_MAX_STAMPS) {
        struct timespec ts;
        clock_gettime(CLOCK_MONOTONIC, &ts);
        rh_stamps[rh_nstamps++] = (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
    }
    serial_Multi_Vector *X = (serial_Multi_Vector *)X_, *Y = (serial_Multi_Vector *)Y_;
    for (int j = 0; j < X->num_active_vectors; j++) {
        const double *x = X->data + (size_t)X->active_indices[j] * X->size;
        double *y = Y->data + (size_t)Y->active_indices[j] * Y->size;
        for (long r = 0; r < A->n; r++) {
            double s = 0.0;
            for (int e = A->rowptr[r]; e < A->rowptr[r + 1]; e++) s += A->val[e] * x[A->colind[e]];
            y[r] = s;
        }
    }
}

static void rh_diag_apply(void *D_, void *X_, void *Y_)
{
    rh_diag *D = (rh_diag *)D_;
    serial_Multi_Vector *X = (serial_Multi_Vector *)X_, *Y = (serial_Multi_Vector *)Y_;
    for (int j = 0; j < X->num_active_vectors; j++) {
        const double *x = X->data + (size_t)X->active_indices[j] * X->size;
        double *y = Y->data + (size_t)Y->active_indices[j] * Y->size;
        for (long r = 0; r < D->n; r++) y[r] = D->d[r] * x[r];
    }
}

/* Same argument list as oracle_lobpcg_solve (oracle/lobpcg_oracle.c). */
int ref_lobpcg_solve(long n, int k, double *Xdata, const int *a_rowptr, const int *a_colind,
                     const double *a_val, const double *bdiag, const double *tdiag, int ny,
                     const double *Ydata, double atol, double rtol, int maxit, int verbosity,
                     void *dpotrf_ptr, void *dsygv_ptr, double *lambda, double *lambda_hist,
                     int lambda_hist_gh, double *res, double *res_hist, int res_hist_gh,
                     int *iterations)
{
    mv_InterfaceInterpreter ii;
    lobpcg_BLASLAPACKFunctions blap;
    lobpcg_Tolerance tol;
    rh_csr A = {n, a_rowptr, a_colind, a_val};
    rh_diag B = {n, bdiag}, T = {n, tdiag};

    SerialSetupInterpreter(&ii);
    memset(&blap, 0, sizeof(blap));
    blap.dpotrf = (BlopexInt(*)(char *, BlopexInt *, double *, BlopexInt *, BlopexInt *))dpotrf_ptr;
    blap.dsygv = (BlopexInt(*)(BlopexInt *, char *, char *, BlopexInt *, double *, BlopexInt *,
                               double *, BlopexInt *, double *, double *, BlopexInt *,
                               BlopexInt *))dsygv_ptr;
    tol.absolute = atol;
    tol.relative = rtol;
    rh_nstamps = 0;

    /* wrap caller memory: the serial multivector does not own it */
    serial_Multi_Vector *x = serial_Multi_VectorCreate((BlopexInt)n, k);
    x->data = Xdata;
    serial_Multi_VectorSetDataOwner(x, 0);
    serial_Multi_VectorInitialize(x);
    mv_MultiVectorPtr xx = mv_MultiVectorWrap(&ii, x, 0);

    serial_Multi_Vector *y = NULL;
    mv_MultiVectorPtr yy = NULL;
    if (ny > 0) {
        y = serial_Multi_VectorCreate((BlopexInt)n, ny);
        y->data = (double *)Ydata;
        serial_Multi_VectorSetDataOwner(y, 0);
        serial_Multi_VectorInitialize(y);
        yy = mv_MultiVectorWrap(&ii, y, 0);
    }

    int flag = lobpcg_solve_double(xx, &A, rh_csr_apply, bdiag ? &B : NULL,
                                   bdiag ? rh_diag_apply : NULL, tdiag ? &T : NULL,
                                   tdiag ? rh_diag_apply : NULL, yy, blap, tol, maxit, verbosity,
                                   iterations, lambda, lambda_hist, lambda_hist_gh, res, res_hist,
                                   res_hist_gh);

    mv_MultiVectorDestroy(xx);
    serial_Multi_VectorDestroy(x);
    if (yy) {
        mv_MultiVectorDestroy(yy);
        serial_Multi_VectorDestroy(y);
    }
    return flag;
}

/* X0 through the reference's own generator (multi_vector.c:166-184). */
void ref_drand48_fill(double *xdata, long n, int k, int seed)
{
    serial_Multi_Vector *x = serial_Multi_VectorCreate((BlopexInt)n, k);
    x->data = xdata;
    serial_Multi_VectorSetDataOwner(x, 0);
    serial_Multi_VectorInitialize(x);
    serial_Multi_VectorSetRandomValues(x, seed);
    serial_Multi_VectorDestroy(x);
}
