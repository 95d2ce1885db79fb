// Abstract multivector: {data, ownsData, interpreter}; every operation
// forwards to the interpreter slot of x (multivector.c:15-254).
#include <cstdlib>

#include "multivector.h"

struct mv_MultiVector {
    void *data;
    BlopexInt ownsData;
    mv_InterfaceInterpreter *interpreter;
};

extern "C" {

void *mv_MultiVectorGetData(mv_MultiVectorPtr x) {
    BlopexAssert(x != NULL);
    return x->data;
}

mv_MultiVectorPtr mv_MultiVectorWrap(mv_InterfaceInterpreter *ii, void *data, BlopexInt ownsData) {
    mv_MultiVectorPtr x = (mv_MultiVectorPtr)malloc(sizeof(mv_MultiVector));
    BlopexAssert(x != NULL);
    x->data = data;
    x->ownsData = ownsData;
    x->interpreter = ii;
    return x;
}

mv_MultiVectorPtr mv_MultiVectorCreateFromSampleVector(void *ii_, BlopexInt n, void *sample) {
    mv_InterfaceInterpreter *ii = (mv_InterfaceInterpreter *)ii_;
    BlopexAssert(ii != NULL && ii->CreateMultiVector != NULL);
    return mv_MultiVectorWrap(ii, ii->CreateMultiVector(ii, n, sample), 1);
}

mv_MultiVectorPtr mv_MultiVectorCreateCopy(mv_MultiVectorPtr x, BlopexInt copyValues) {
    BlopexAssert(x != NULL);
    mv_InterfaceInterpreter *ii = x->interpreter;
    return mv_MultiVectorWrap(ii, ii->CopyCreateMultiVector(x->data, copyValues), 1);
}

void mv_MultiVectorDestroy(mv_MultiVectorPtr v) {
    if (!v) return;
    if (v->ownsData) v->interpreter->DestroyMultiVector(v->data);
    free(v);
}

BlopexInt mv_MultiVectorWidth(mv_MultiVectorPtr v) {
    return v ? v->interpreter->Width(v->data) : 0;
}

BlopexInt mv_MultiVectorHeight(mv_MultiVectorPtr v) {
    if (!v || !v->interpreter->Height) return 0;
    return v->interpreter->Height(v->data);
}

void mv_MultiVectorSetMask(mv_MultiVectorPtr v, BlopexInt *mask) {
    BlopexAssert(v != NULL);
    v->interpreter->SetMask(v->data, mask);
}

void mv_MultiVectorClear(mv_MultiVectorPtr v) {
    BlopexAssert(v != NULL);
    v->interpreter->ClearMultiVector(v->data);
}

void mv_MultiVectorSetRandom(mv_MultiVectorPtr v, BlopexInt seed) {
    BlopexAssert(v != NULL);
    v->interpreter->SetRandomVectors(v->data, seed);
}

void mv_MultiVectorCopy(mv_MultiVectorPtr src, mv_MultiVectorPtr dest) {
    BlopexAssert(src != NULL && dest != NULL);
    src->interpreter->CopyMultiVector(src->data, dest->data);
}

void mv_MultiVectorAxpy(double a, mv_MultiVectorPtr x, mv_MultiVectorPtr y) {
    BlopexAssert(x != NULL && y != NULL);
    x->interpreter->MultiAxpy(a, x->data, y->data);
}

void mv_MultiVectorByMultiVector(mv_MultiVectorPtr x, mv_MultiVectorPtr y, BlopexInt gh,
                                 BlopexInt h, BlopexInt w, void *v) {
    BlopexAssert(x != NULL && y != NULL);
    x->interpreter->MultiInnerProd(x->data, y->data, gh, h, w, v);
}

void mv_MultiVectorByMultiVectorDiag(mv_MultiVectorPtr x, mv_MultiVectorPtr y, BlopexInt *mask,
                                     BlopexInt n, void *d) {
    BlopexAssert(x != NULL && y != NULL);
    x->interpreter->MultiInnerProdDiag(x->data, y->data, mask, n, d);
}

void mv_MultiVectorByMatrix(mv_MultiVectorPtr x, BlopexInt gh, BlopexInt h, BlopexInt w, void *v,
                            mv_MultiVectorPtr y) {
    BlopexAssert(x != NULL && y != NULL);
    x->interpreter->MultiVecMat(x->data, gh, h, w, v, y->data);
}

void mv_MultiVectorXapy(mv_MultiVectorPtr x, BlopexInt gh, BlopexInt h, BlopexInt w, void *v,
                        mv_MultiVectorPtr y) {
    BlopexAssert(x != NULL && y != NULL);
    x->interpreter->MultiXapy(x->data, gh, h, w, v, y->data);
}

void mv_MultiVectorByDiagonal(mv_MultiVectorPtr x, BlopexInt *mask, BlopexInt n, void *d,
                              mv_MultiVectorPtr y) {
    BlopexAssert(x != NULL && y != NULL);
    x->interpreter->MultiVecMatDiag(x->data, mask, n, d, y->data);
}

void mv_MultiVectorEval(void (*f)(void *, void *, void *), void *par, mv_MultiVectorPtr x,
                        mv_MultiVectorPtr y) {
    BlopexAssert(x != NULL && y != NULL);
    x->interpreter->Eval(f, par, x->data, y->data);
}

void mv_MultiVectorPrint(mv_MultiVectorPtr x, char *tag, BlopexInt limit) {
    BlopexAssert(x != NULL);
    x->interpreter->MultiPrint(x->data, tag, limit);
}

/* the interpreter table behind x (used by the solver to pick its path) */
mv_InterfaceInterpreter *blopex_amd_mv_interpreter(mv_MultiVectorPtr x) {
    return x ? x->interpreter : NULL;
}

}  // extern "C"
