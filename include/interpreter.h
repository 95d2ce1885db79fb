/*
 * mv_InterfaceInterpreter: the callback table through which lobpcg_solve
 * reaches a multivector backend.  Same 26 slots, in the same order and with
 * the same signatures, as blopex_abstract/multivector/interpreter.h:11-50, so
 * a table filled by HipSetupInterpreter (hip_multivector.h) can be handed to
 * code compiled against the reference's header.
 */
#ifndef BLOPEX_AMD_INTERPRETER_H
#define BLOPEX_AMD_INTERPRETER_H

#include "fortran_options.h"

typedef struct {
    /* single-vector slots (left NULL by the block backends) */
    void *(*CreateVector)(void *vector);
    BlopexInt (*DestroyVector)(void *vector);
    BlopexInt (*InnerProd)(void *x, void *y, void *result);
    BlopexInt (*CopyVector)(void *x, void *y);
    BlopexInt (*ClearVector)(void *x);
    BlopexInt (*SetRandomValues)(void *x, BlopexInt seed);
    BlopexInt (*ScaleVector)(double alpha, void *x);
    BlopexInt (*Axpy)(void *alpha, void *x, void *y);
    BlopexInt (*VectorSize)(void *vector);

    /* block slots */
    void *(*CreateMultiVector)(void *, BlopexInt n, void *vector);
    void *(*CopyCreateMultiVector)(void *x, BlopexInt copyValues);
    void (*DestroyMultiVector)(void *x);
    BlopexInt (*Width)(void *x);
    BlopexInt (*Height)(void *x);
    void (*SetMask)(void *x, BlopexInt *mask);
    void (*CopyMultiVector)(void *x, void *y);
    void (*ClearMultiVector)(void *x);
    void (*SetRandomVectors)(void *x, BlopexInt seed);
    void (*MultiInnerProd)(void *x, void *y, BlopexInt gh, BlopexInt h, BlopexInt w, void *v);
    void (*MultiInnerProdDiag)(void *x, void *y, BlopexInt *mask, BlopexInt n, void *diag);
    void (*MultiVecMat)(void *x, BlopexInt gh, BlopexInt h, BlopexInt w, void *v, void *y);
    void (*MultiVecMatDiag)(void *x, BlopexInt *mask, BlopexInt n, void *diag, void *y);
    void (*MultiAxpy)(double alpha, void *x, void *y);
    void (*MultiPrint)(void *x, char *tag, BlopexInt limit);
    void (*MultiXapy)(void *x, BlopexInt gh, BlopexInt h, BlopexInt w, void *v, void *y);
    void (*Eval)(void (*f)(void *, void *, void *), void *par, void *x, void *y);
} mv_InterfaceInterpreter;

#endif
