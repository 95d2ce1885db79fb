/*
 * ref_harness_complex.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Drives the reference's own lobpcg_solve_complex (krylov/lobpcg.c:417-498,
 * compiled unmodified from /root/reference by oracle/Makefile into
 * oracle/_ref/libblopex_ref_complex.so) with the reference's complex serial
 * backend, the way its complex driver does (blopex_serial_complex/driver/
 * serial_driver.c:40-140): a dense complex A applied by MatMultiVec
 * (pcg_multi.c:175-200), B = T = Y = NULL, zpotrf / zhegv passed in as
 * function pointers (krylov/lobpcg.h:25-46).
 */
#include <stdlib.h>
#include <string.h>

#include "interpreter.h"
#include "lobpcg.h"
#include "multi_vector.h"
#include "multivector.h"
#include "pcg_multi.h"

/* A: n x n, X: n x k (in: X0, out: eigenvectors), both column-major komplex
   (interleaved real, imaginary).  lam, lamh: komplex k and k x (maxit + 1);
   res, resh: double k and k x (maxit + 1). */
int ref_lobpcg_solve_complex(int n, int k, const double *A, double *X, double atol, double rtol,
                             int maxit, void *zpotrf_ptr, void *zhegv_ptr, int *iters,
                             double *lam, double *lamh, double *res, double *resh)
{
    serial_Multi_Vector *a = serial_Multi_VectorCreate(n, n);
    serial_Multi_VectorInitialize(a);
    memcpy(a->data, A, sizeof(double) * 2 * (size_t)n * n);
    serial_Multi_Vector *x = serial_Multi_VectorCreate(n, k);
    serial_Multi_VectorInitialize(x);
    memcpy(x->data, X, sizeof(double) * 2 * (size_t)n * k);

    mv_InterfaceInterpreter ii;
    SerialSetupInterpreter(&ii);
    mv_MultiVectorPtr xx = mv_MultiVectorWrap(&ii, x, 0);

    lobpcg_BLASLAPACKFunctions blap;
    memset(&blap, 0, sizeof blap);
    blap.zpotrf = (BlopexInt(*)(char *, BlopexInt *, komplex *, BlopexInt *, BlopexInt *))zpotrf_ptr;
    blap.zhegv = (BlopexInt(*)(BlopexInt *, char *, char *, BlopexInt *, komplex *, BlopexInt *,
                               komplex *, BlopexInt *, double *, komplex *, BlopexInt *, double *,
                               BlopexInt *))zhegv_ptr;
    lobpcg_Tolerance tol;
    tol.absolute = atol;
    tol.relative = rtol;
    int flag = lobpcg_solve_complex(xx, a, MatMultiVec, NULL, NULL, NULL, NULL, NULL, blap, tol,
                                    maxit, 0, iters, (komplex *)lam, (komplex *)lamh, k, res,
                                    resh, k);
    memcpy(X, x->data, sizeof(double) * 2 * (size_t)n * k);
    mv_MultiVectorDestroy(xx);
    serial_Multi_VectorDestroy(x);
    serial_Multi_VectorDestroy(a);
    return flag;
}

/* The reference's complex serial backend for a solver other than its own
   (the library's generic lobpcg_solve_complex, on the CPU): an n x k block
   holding a copy of data, its contents, the interpreter table and the dense
   operator callback. */
void *ref_complex_block(int n, int k, const double *data)
{
    serial_Multi_Vector *x = serial_Multi_VectorCreate(n, k);
    serial_Multi_VectorInitialize(x);
    memcpy(x->data, data, sizeof(double) * 2 * (size_t)n * k);
    return x;
}

void ref_complex_block_get(void *x_, double *out)
{
    serial_Multi_Vector *x = (serial_Multi_Vector *)x_;
    memcpy(out, x->data, sizeof(double) * 2 * (size_t)x->size * x->num_vectors);
}

void ref_complex_block_destroy(void *x) { serial_Multi_VectorDestroy((serial_Multi_Vector *)x); }

void ref_complex_setup_interpreter(mv_InterfaceInterpreter *ii) { SerialSetupInterpreter(ii); }

void *ref_complex_matmultivec(void) { return (void *)MatMultiVec; }
