/*
 * Abstract multivector: an opaque {data, ownsData, interpreter} handle whose
 * operations dispatch through mv_InterfaceInterpreter.  Same entry points as
 * blopex_abstract/multivector/multivector.h:23-103 (implementation:
 * multivector.c:15-254).
 */
#ifndef BLOPEX_AMD_MULTIVECTOR_H
#define BLOPEX_AMD_MULTIVECTOR_H

#include "fortran_options.h"
#include "interpreter.h"

typedef struct mv_MultiVector *mv_MultiVectorPtr;

#ifdef __cplusplus
extern "C" {
#endif

/* backend block behind x (what operator callbacks receive) */
void *mv_MultiVectorGetData(mv_MultiVectorPtr x);

/* wrap a backend block; ownsData != 0 destroys it with the handle */
mv_MultiVectorPtr mv_MultiVectorWrap(mv_InterfaceInterpreter *ii, void *data, BlopexInt ownsData);

/* n-wide block created by the backend from a sample vector */
mv_MultiVectorPtr mv_MultiVectorCreateFromSampleVector(void *ii, BlopexInt n, void *sample);

/* block of x's shape; copies values when copyValues != 0 */
mv_MultiVectorPtr mv_MultiVectorCreateCopy(mv_MultiVectorPtr x, BlopexInt copyValues);

void mv_MultiVectorDestroy(mv_MultiVectorPtr);
BlopexInt mv_MultiVectorWidth(mv_MultiVectorPtr v);
BlopexInt mv_MultiVectorHeight(mv_MultiVectorPtr v);

/* later operations touch only the columns with mask[i] != 0 (NULL: all) */
void mv_MultiVectorSetMask(mv_MultiVectorPtr v, BlopexInt *mask);

void mv_MultiVectorClear(mv_MultiVectorPtr);
void mv_MultiVectorSetRandom(mv_MultiVectorPtr v, BlopexInt seed);
void mv_MultiVectorCopy(mv_MultiVectorPtr src, mv_MultiVectorPtr dest);

/* y = a*x + y */
void mv_MultiVectorAxpy(double a, mv_MultiVectorPtr x, mv_MultiVectorPtr y);

/* v = x'*y, column-major h x w with leading dimension gh */
void mv_MultiVectorByMultiVector(mv_MultiVectorPtr x, mv_MultiVectorPtr y, BlopexInt gh,
                                 BlopexInt h, BlopexInt w, void *v);

/* diag(mask) = diag(x'*y) */
void mv_MultiVectorByMultiVectorDiag(mv_MultiVectorPtr, mv_MultiVectorPtr, BlopexInt *mask,
                                     BlopexInt n, void *diag);

/* y = x*v, v column-major h x w with leading dimension gh */
void mv_MultiVectorByMatrix(mv_MultiVectorPtr x, BlopexInt gh, BlopexInt h, BlopexInt w, void *v,
                            mv_MultiVectorPtr y);

/* y = x*v + y */
void mv_MultiVectorXapy(mv_MultiVectorPtr x, BlopexInt gh, BlopexInt h, BlopexInt w, void *v,
                        mv_MultiVectorPtr y);

/* y = x*diag(mask) */
void mv_MultiVectorByDiagonal(mv_MultiVectorPtr x, BlopexInt *mask, BlopexInt n, void *diag,
                              mv_MultiVectorPtr y);

/* y = f(x) column by column */
void mv_MultiVectorEval(void (*f)(void *, void *, void *), void *par, mv_MultiVectorPtr x,
                        mv_MultiVectorPtr y);

void mv_MultiVectorPrint(mv_MultiVectorPtr x, char *tag, BlopexInt limit);

#ifdef __cplusplus
}
#endif

#endif
