/*
 * hip_driver.c -- a C caller of libblopex_amd.so in the pattern of the
 * reference's serial driver (blopex_serial_double/driver/serial_driver.c:
 * 67-175): build the operator, create a block, fill it with
 * serial_Multi_VectorSetRandomValues' stream, wrap it with the interpreter,
 * call lobpcg_solve_double.  The only differences are the objects: a device
 * CSR matrix and hip_Multi_Vector instead of serial multivectors, and no
 * LAPACK (device solves need none).
 *
 *   hip_driver nx ny nz k [maxit] [tol] [fp64|mixed] [jacobi]
 *
 * The operator is the 7-point Dirichlet Laplacian on an nx x ny x nz grid
 * (diagonal 6, off-diagonals -1: the stencil ij.c:3853-3873 drives
 * GenerateLaplacian with).  Prints "flag <f> iterations <it>", then one line
 * "<lambda> <residual>" per eigenpair.
 */
#include <stdio.h>
#include <stdlib.h>

#include "hip_multivector.h"
#include "lobpcg.h"

static void laplacian_7pt(int nx, int ny, int nz, int **rowptr, int **colind, double **val,
                          long *nnz)
{
    const long n = (long)nx * ny * nz;
    *rowptr = (int *)malloc(sizeof(int) * (n + 1));
    *colind = (int *)malloc(sizeof(int) * n * 7);
    *val = (double *)malloc(sizeof(double) * n * 7);
    long e = 0;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const long r = x + (long)nx * (y + (long)ny * z);
                (*rowptr)[r] = (int)e;
                /* columns in ascending order */
                const long nb[7] = {z > 0 ? r - (long)nx * ny : -1, y > 0 ? r - nx : -1,
                                    x > 0 ? r - 1 : -1, r, x < nx - 1 ? r + 1 : -1,
                                    y < ny - 1 ? r + nx : -1,
                                    z < nz - 1 ? r + (long)nx * ny : -1};
                for (int t = 0; t < 7; t++)
                    if (nb[t] >= 0) {
                        (*colind)[e] = (int)nb[t];
                        (*val)[e] = nb[t] == r ? 6.0 : -1.0;
                        e++;
                    }
            }
    (*rowptr)[n] = (int)e;
    *nnz = e;
}

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s nx ny nz k [maxit] [tol] [fp64|mixed] [jacobi]\n", argv[0]);
        return 2;
    }
    const int nx = atoi(argv[1]), ny = atoi(argv[2]), nz = atoi(argv[3]), k = atoi(argv[4]);
    const int maxit = argc > 5 ? atoi(argv[5]) : 200;
    const double tol = argc > 6 ? atof(argv[6]) : 1e-6;
    const int mixed = argc > 7 && argv[7][0] == 'm';
    const int jacobi = argc > 8 && argv[8][0] == 'j';

    int *rowptr, *colind;
    double *val;
    long nnz;
    laplacian_7pt(nx, ny, nz, &rowptr, &colind, &val, &nnz);
    const long n = (long)nx * ny * nz;

    hip_CSRMatrix *A = hip_CSRMatrixCreate(n, nnz, rowptr, colind, val);
    hip_DiagMatrix *T = jacobi ? hip_DiagMatrixCreateJacobi(A) : NULL;
    hip_Multi_Vector *x = hip_Multi_VectorCreate(n, k, HIPMV_FP64);
    if (!A || !x || (jacobi && !T)) {
        fprintf(stderr, "setup failed: %s\n", blopex_amd_last_error());
        return 1;
    }
    /* the stream of serial_Multi_VectorSetRandomValues(x, 1) */
    hip_Multi_VectorSetRandomValues(x, 1);

    hip_LobpcgSettings s;
    hip_LobpcgSettingsDefault(&s);
    s.mode = mixed ? HIPLOBPCG_MIXED : HIPLOBPCG_FP64;
    hip_Multi_VectorSetSettings(x, &s);

    mv_InterfaceInterpreter ii;
    HipSetupInterpreter(&ii); /* replaces SerialSetupInterpreter(&ii) */
    mv_MultiVectorPtr xx = mv_MultiVectorWrap(&ii, x, 0);

    lobpcg_BLASLAPACKFunctions blap_fn = {0}; /* device solves need no LAPACK */
    lobpcg_Tolerance lobpcg_tol = {.absolute = tol, .relative = 1e-50};
    double *eigs = (double *)malloc(sizeof(double) * k);
    double *resid = (double *)malloc(sizeof(double) * k);
    int iterations = 0;

    const int flag = lobpcg_solve_double(xx, A, HipCSRMatMultiVec, NULL, NULL, T,
                                         T ? HipDiagMatMultiVec : NULL, NULL, blap_fn, lobpcg_tol,
                                         maxit, 0, &iterations, eigs, NULL, 0, resid, NULL, 0);

    printf("flag %d iterations %d\n", flag, iterations);
    for (int i = 0; i < k; i++) printf("%.17g %.17g\n", eigs[i], resid[i]);

    mv_MultiVectorDestroy(xx);
    hip_Multi_VectorDestroy(x);
    if (T) hip_DiagMatrixDestroy(T);
    hip_CSRMatrixDestroy(A);
    free(eigs);
    free(resid);
    free(rowptr);
    free(colind);
    free(val);
    return flag == 0 ? 0 : 1;
}
