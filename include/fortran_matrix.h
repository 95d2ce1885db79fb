/*
 * Small column-major host matrices used by lobpcg_solve for its Gram blocks
 * and Ritz coordinates.  Same struct layout and double-precision entry points
 * as blopex_abstract/utilities/fortran_matrix.h:11-112 (implementation
 * fortran_matrix.c), and the complex (z*) twins of fortran_matrix.h:111-166
 * that lobpcg_solve_complex binds: entries are komplex, transposes are
 * conjugate transposes, Symmetrize makes the matrix Hermitian.
 */
#ifndef BLOPEX_AMD_FORTRAN_MATRIX_H
#define BLOPEX_AMD_FORTRAN_MATRIX_H

#include "fortran_options.h"

typedef struct {
    long globalHeight; /* leading dimension */
    long height;
    long width;
    void *value;
    BlopexInt ownsValues;
} utilities_FortranMatrix;

typedef struct {
    double real, imag;
} komplex;

#ifdef __cplusplus
extern "C" {
#endif

utilities_FortranMatrix *utilities_FortranMatrixCreate(void);
void utilities_FortranMatrixWrap(void *, long gh, long h, long w, utilities_FortranMatrix *mtx);
void utilities_FortranMatrixDestroy(utilities_FortranMatrix *mtx);
long utilities_FortranMatrixGlobalHeight(utilities_FortranMatrix *mtx);
long utilities_FortranMatrixHeight(utilities_FortranMatrix *mtx);
long utilities_FortranMatrixWidth(utilities_FortranMatrix *mtx);
void *utilities_FortranMatrixValues(utilities_FortranMatrix *mtx);
double utilities_FortranMatrixFNorm(utilities_FortranMatrix *mtx);

void utilities_FortranMatrixAllocateData(long h, long w, utilities_FortranMatrix *mtx);
void utilities_FortranMatrixClear(utilities_FortranMatrix *mtx);
void utilities_FortranMatrixClearL(utilities_FortranMatrix *mtx);
void utilities_FortranMatrixSetToIdentity(utilities_FortranMatrix *mtx);
void utilities_FortranMatrixTransposeSquare(utilities_FortranMatrix *mtx);
void utilities_FortranMatrixSymmetrize(utilities_FortranMatrix *mtx);
void utilities_FortranMatrixCopy(utilities_FortranMatrix *src, BlopexInt t,
                                 utilities_FortranMatrix *dest);
void utilities_FortranMatrixIndexCopy(BlopexInt *index, utilities_FortranMatrix *src, BlopexInt t,
                                      utilities_FortranMatrix *dest);
void utilities_FortranMatrixSetDiagonal(utilities_FortranMatrix *mtx, utilities_FortranMatrix *d);
void utilities_FortranMatrixGetDiagonal(utilities_FortranMatrix *mtx, utilities_FortranMatrix *d);
void utilities_FortranMatrixAdd(double a, utilities_FortranMatrix *mtxA,
                                utilities_FortranMatrix *mtxB, utilities_FortranMatrix *mtxC);
void utilities_FortranMatrixDMultiply(utilities_FortranMatrix *d, utilities_FortranMatrix *mtx);
void utilities_FortranMatrixMultiplyD(utilities_FortranMatrix *mtx, utilities_FortranMatrix *d);
void utilities_FortranMatrixMultiply(utilities_FortranMatrix *mtxA, BlopexInt tA,
                                     utilities_FortranMatrix *mtxB, BlopexInt tB,
                                     utilities_FortranMatrix *mtxC);
double utilities_FortranMatrixAbs(utilities_FortranMatrix *mtx, long i, long j);
void *utilities_FortranMatrixValuePtr(utilities_FortranMatrix *mtx, long i, long j);
double utilities_FortranMatrixMaxValue(utilities_FortranMatrix *mtx);
void utilities_FortranMatrixSelectBlock(utilities_FortranMatrix *mtx, long iFrom, long iTo,
                                        long jFrom, long jTo, utilities_FortranMatrix *block);
void utilities_FortranMatrixUpperInv(utilities_FortranMatrix *u);
BlopexInt utilities_FortranMatrixPrint(utilities_FortranMatrix *mtx, char fileName[]);

/* complex routines (fortran_matrix.h:111-166) */
void zutilities_FortranMatrixAllocateData(long h, long w, utilities_FortranMatrix *mtx);
void zutilities_FortranMatrixClear(utilities_FortranMatrix *mtx);
void zutilities_FortranMatrixClearL(utilities_FortranMatrix *mtx);
void zutilities_FortranMatrixSetToIdentity(utilities_FortranMatrix *mtx);
void zutilities_FortranMatrixTransposeSquare(utilities_FortranMatrix *mtx);
void zutilities_FortranMatrixSymmetrize(utilities_FortranMatrix *mtx);
void zutilities_FortranMatrixCopy(utilities_FortranMatrix *src, BlopexInt t,
                                  utilities_FortranMatrix *dest);
void zutilities_FortranMatrixIndexCopy(BlopexInt *index, utilities_FortranMatrix *src,
                                       BlopexInt t, utilities_FortranMatrix *dest);
void zutilities_FortranMatrixSetDiagonal(utilities_FortranMatrix *mtx, utilities_FortranMatrix *d);
void zutilities_FortranMatrixGetDiagonal(utilities_FortranMatrix *mtx, utilities_FortranMatrix *d);
void zutilities_FortranMatrixAdd(double a, utilities_FortranMatrix *mtxA,
                                 utilities_FortranMatrix *mtxB, utilities_FortranMatrix *mtxC);
void zutilities_FortranMatrixDMultiply(utilities_FortranMatrix *d, utilities_FortranMatrix *mtx);
void zutilities_FortranMatrixMultiplyD(utilities_FortranMatrix *mtx, utilities_FortranMatrix *d);
void zutilities_FortranMatrixMultiply(utilities_FortranMatrix *mtxA, BlopexInt tA,
                                      utilities_FortranMatrix *mtxB, BlopexInt tB,
                                      utilities_FortranMatrix *mtxC);
double zutilities_FortranMatrixAbs(utilities_FortranMatrix *mtx, long i, long j);
void *zutilities_FortranMatrixValuePtr(utilities_FortranMatrix *mtx, long i, long j);
void zutilities_FortranMatrixSelectBlock(utilities_FortranMatrix *mtx, long iFrom, long iTo,
                                         long jFrom, long jTo, utilities_FortranMatrix *block);
void zutilities_FortranMatrixUpperInv(utilities_FortranMatrix *u);
BlopexInt zutilities_FortranMatrixPrint(utilities_FortranMatrix *mtx, char fileName[]);

#ifdef __cplusplus
}
#endif

#endif
