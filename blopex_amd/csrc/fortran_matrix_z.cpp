// Complex twins of the small host matrices (blopex_abstract/utilities/
// fortran_matrix.c, the zutilities_* routines): komplex entries, conjugate
// transposes, Hermitian symmetrisation.  With the double routines that do not
// depend on the entry type (create, wrap, destroy, sizes, values) they make
// the utilities_FortranInterpreter lobpcg_solve_complex binds
// (krylov/lobpcg.c:440-468), plus chol / solveGEVP over the caller's
// zpotrf / zhegv (lobpcg.c:40-62, 104-158).
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "lobpcg.h"

namespace {

using cplx = std::complex<double>;
static_assert(sizeof(cplx) == sizeof(komplex), "komplex layout");

inline cplx &at(utilities_FortranMatrix *m, long i, long j) {
    return ((cplx *)m->value)[i + j * m->globalHeight];
}
inline cplx &el(utilities_FortranMatrix *v, long i) { return ((cplx *)v->value)[i]; }

BlopexInt zchol_lapack(utilities_FortranMatrix *a, lobpcg_BLASLAPACKFunctions blap,
                       utilities_FortranInterpreter *u) {
    char uplo = 'U';
    BlopexInt n = (BlopexInt)u->FortranMatrixHeight(a);
    BlopexInt lda = (BlopexInt)u->FortranMatrixGlobalHeight(a), info = 0;
    blap.zpotrf(&uplo, &n, (komplex *)u->FortranMatrixValues(a), &lda, &info);
    return info;
}

// dsygv's complex counterpart: eigenvalues come back real from zhegv and are
// stored as komplex with a zero imaginary part
BlopexInt zgevp_lapack(utilities_FortranMatrix *A, utilities_FortranMatrix *B,
                       utilities_FortranMatrix *lam, lobpcg_BLASLAPACKFunctions blap,
                       utilities_FortranInterpreter *u) {
    BlopexInt itype = 1, n = (BlopexInt)u->FortranMatrixHeight(A);
    BlopexInt lda = (BlopexInt)u->FortranMatrixGlobalHeight(A);
    BlopexInt ldb = (BlopexInt)u->FortranMatrixGlobalHeight(B);
    BlopexInt lwork = 10 * n, info = 0;
    char jobz = 'V', uplo = 'L';
    std::vector<komplex> work(lwork > 0 ? lwork : 1);
    std::vector<double> w(n > 0 ? n : 1), rwork(3 * (n > 0 ? n : 1));
    blap.zhegv(&itype, &jobz, &uplo, &n, (komplex *)u->FortranMatrixValues(A), &lda,
               (komplex *)u->FortranMatrixValues(B), &ldb, w.data(), work.data(), &lwork,
               rwork.data(), &info);
    komplex *l = (komplex *)u->FortranMatrixValues(lam);
    for (BlopexInt i = 0; i < n; i++) {
        l[i].real = w[i];
        l[i].imag = 0.0;
    }
    return info;
}

}  // namespace

extern "C" {

void zutilities_FortranMatrixAllocateData(long h, long w, utilities_FortranMatrix *m) {
    BlopexAssert(h > 0 && w > 0);
    if (m->value && m->ownsValues) free(m->value);
    m->value = calloc((size_t)h * w, sizeof(komplex));
    BlopexAssert(m->value != NULL);
    m->globalHeight = m->height = h;
    m->width = w;
    m->ownsValues = 1;
}

void zutilities_FortranMatrixClear(utilities_FortranMatrix *m) {
    for (long j = 0; j < m->width; j++)
        for (long i = 0; i < m->height; i++) at(m, i, j) = 0.0;
}

void zutilities_FortranMatrixClearL(utilities_FortranMatrix *m) {
    long w = m->width < m->height ? m->width : m->height;
    for (long j = 0; j + 1 < w; j++)
        for (long i = j + 1; i < m->height; i++) at(m, i, j) = 0.0;
}

void zutilities_FortranMatrixSetToIdentity(utilities_FortranMatrix *m) {
    zutilities_FortranMatrixClear(m);
    for (long j = 0; j < m->width && j < m->height; j++) at(m, j, j) = 1.0;
}

// in place m <- m^H
void zutilities_FortranMatrixTransposeSquare(utilities_FortranMatrix *m) {
    BlopexAssert(m->height == m->width);
    for (long j = 0; j < m->width; j++) {
        at(m, j, j) = std::conj(at(m, j, j));
        for (long i = j + 1; i < m->height; i++) {
            const cplx t = at(m, i, j);
            at(m, i, j) = std::conj(at(m, j, i));
            at(m, j, i) = std::conj(t);
        }
    }
}

// m <- (m + m^H) / 2: a real diagonal, the upper triangle the conjugate of the lower
void zutilities_FortranMatrixSymmetrize(utilities_FortranMatrix *m) {
    BlopexAssert(m->height == m->width);
    for (long j = 0; j < m->width; j++) {
        at(m, j, j) = at(m, j, j).real();
        for (long i = j + 1; i < m->height; i++) {
            const cplx a = (at(m, i, j) + std::conj(at(m, j, i))) * 0.5;
            at(m, i, j) = a;
            at(m, j, i) = std::conj(a);
        }
    }
}

// t == 0: dst = src; t != 0: dst = src^H
void zutilities_FortranMatrixCopy(utilities_FortranMatrix *src, BlopexInt t,
                                  utilities_FortranMatrix *dst) {
    if (t == 0) {
        BlopexAssert(src->height == dst->height && src->width == dst->width);
        for (long j = 0; j < dst->width; j++)
            for (long i = 0; i < dst->height; i++) at(dst, i, j) = at(src, i, j);
    } else {
        BlopexAssert(src->height == dst->width && src->width == dst->height);
        for (long j = 0; j < dst->width; j++)
            for (long i = 0; i < dst->height; i++) at(dst, i, j) = std::conj(at(src, j, i));
    }
}

// column j of dst <- column index[j]-1 of src (the conjugate of row index[j]-1 when t != 0)
void zutilities_FortranMatrixIndexCopy(BlopexInt *index, utilities_FortranMatrix *src,
                                       BlopexInt t, utilities_FortranMatrix *dst) {
    for (long j = 0; j < dst->width; j++) {
        const long c = index[j] - 1;
        for (long i = 0; i < dst->height; i++)
            at(dst, i, j) = t == 0 ? at(src, i, c) : std::conj(at(src, c, i));
    }
}

void zutilities_FortranMatrixSetDiagonal(utilities_FortranMatrix *m, utilities_FortranMatrix *d) {
    BlopexAssert(d->height >= m->height);
    for (long j = 0; j < m->width && j < m->height; j++) at(m, j, j) = el(d, j);
}

void zutilities_FortranMatrixGetDiagonal(utilities_FortranMatrix *m, utilities_FortranMatrix *d) {
    BlopexAssert(d->height >= m->height);
    for (long j = 0; j < m->width && j < m->height; j++) el(d, j) = at(m, j, j);
}

// C = a A + B, a real
void zutilities_FortranMatrixAdd(double a, utilities_FortranMatrix *A, utilities_FortranMatrix *B,
                                 utilities_FortranMatrix *C) {
    BlopexAssert(A->height == C->height && B->height == C->height && A->width == C->width &&
                 B->width == C->width);
    for (long j = 0; j < C->width; j++)
        for (long i = 0; i < C->height; i++) at(C, i, j) = a * at(A, i, j) + at(B, i, j);
}

// row i of m times d(i)
void zutilities_FortranMatrixDMultiply(utilities_FortranMatrix *d, utilities_FortranMatrix *m) {
    BlopexAssert(d->height == m->height);
    for (long j = 0; j < m->width; j++)
        for (long i = 0; i < m->height; i++) at(m, i, j) *= el(d, i);
}

// column j of m times d(j)
void zutilities_FortranMatrixMultiplyD(utilities_FortranMatrix *m, utilities_FortranMatrix *d) {
    BlopexAssert(d->height == m->width);
    for (long j = 0; j < m->width; j++)
        for (long i = 0; i < m->height; i++) at(m, i, j) *= el(d, j);
}

// C = op(A) op(B), op(M) = M^H when its flag is set
void zutilities_FortranMatrixMultiply(utilities_FortranMatrix *A, BlopexInt tA,
                                      utilities_FortranMatrix *B, BlopexInt tB,
                                      utilities_FortranMatrix *C) {
    const long l = tA ? A->height : A->width;
    BlopexAssert((tA ? A->width : A->height) == C->height);
    BlopexAssert((tB ? B->width : B->height) == l && (tB ? B->height : B->width) == C->width);
    for (long j = 0; j < C->width; j++)
        for (long i = 0; i < C->height; i++) {
            cplx s = 0.0;
            for (long q = 0; q < l; q++)
                s += (tA ? std::conj(at(A, q, i)) : at(A, i, q)) *
                     (tB ? std::conj(at(B, j, q)) : at(B, q, j));
            at(C, i, j) = s;
        }
}

double zutilities_FortranMatrixAbs(utilities_FortranMatrix *m, long i, long j) {
    BlopexAssert(i > 0 && i <= m->height && j > 0 && j <= m->width);
    return std::abs(at(m, i - 1, j - 1));
}

void *zutilities_FortranMatrixValuePtr(utilities_FortranMatrix *m, long i, long j) {
    BlopexAssert(i > 0 && i <= m->height && j > 0 && j <= m->width);
    return &at(m, i - 1, j - 1);
}

void zutilities_FortranMatrixSelectBlock(utilities_FortranMatrix *m, long iFrom, long iTo,
                                         long jFrom, long jTo, utilities_FortranMatrix *b) {
    if (b->value && b->ownsValues) free(b->value);
    b->globalHeight = m->globalHeight;
    b->ownsValues = 0;
    if (iTo < iFrom || jTo < jFrom) {
        b->height = b->width = 0;
        b->value = NULL;
        return;
    }
    b->height = iTo - iFrom + 1;
    b->width = jTo - jFrom + 1;
    b->value = (cplx *)m->value + (iFrom - 1) + (jFrom - 1) * m->globalHeight;
}

// inverse of an upper triangle in place, rows from the bottom (as the double routine)
void zutilities_FortranMatrixUpperInv(utilities_FortranMatrix *u) {
    const long n = u->height;
    BlopexAssert(u->width == n);
    std::vector<cplx> d(n > 0 ? n : 1);
    for (long i = 0; i < n; i++) {
        d[i] = at(u, i, i);
        at(u, i, i) = 1.0 / d[i];
    }
    for (long i = n - 2; i >= 0; i--)
        for (long j = n - 1; j > i; j--) {
            cplx v = 0.0;
            for (long m = i + 1; m <= j; m++) v -= at(u, i, m) * at(u, m, j);
            at(u, i, j) = v / d[i];
        }
}

BlopexInt zutilities_FortranMatrixPrint(utilities_FortranMatrix *m, char fileName[]) {
    FILE *f = fopen(fileName, "w");
    if (!f) return 1;
    fprintf(f, "%ld\n%ld\n", m->height, m->width);
    for (long j = 0; j < m->width; j++)
        for (long i = 0; i < m->height; i++)
            fprintf(f, "%.14e %.14e\n", at(m, i, j).real(), at(m, i, j).imag());
    fclose(f);
    return 0;
}

// Fills the complex small-dense interpreters lobpcg_solve_complex uses
// (lobpcg.c:440-468): chol / solveGEVP call blap_fn.zpotrf / zhegv.
void blopex_amd_complex_interpreters(lobpcg_Interpreter *lp, utilities_FortranInterpreter *u) {
    u->FortranMatrixCreate = utilities_FortranMatrixCreate;
    u->FortranMatrixAllocateData = zutilities_FortranMatrixAllocateData;
    u->FortranMatrixWrap = utilities_FortranMatrixWrap;
    u->FortranMatrixDestroy = utilities_FortranMatrixDestroy;
    u->FortranMatrixGlobalHeight = utilities_FortranMatrixGlobalHeight;
    u->FortranMatrixHeight = utilities_FortranMatrixHeight;
    u->FortranMatrixWidth = utilities_FortranMatrixWidth;
    u->FortranMatrixValues = utilities_FortranMatrixValues;
    u->FortranMatrixClear = zutilities_FortranMatrixClear;
    u->FortranMatrixClearL = zutilities_FortranMatrixClearL;
    u->FortranMatrixSetToIdentity = zutilities_FortranMatrixSetToIdentity;
    u->FortranMatrixTransposeSquare = zutilities_FortranMatrixTransposeSquare;
    u->FortranMatrixSymmetrize = zutilities_FortranMatrixSymmetrize;
    u->FortranMatrixCopy = zutilities_FortranMatrixCopy;
    u->FortranMatrixIndexCopy = zutilities_FortranMatrixIndexCopy;
    u->FortranMatrixSetDiagonal = zutilities_FortranMatrixSetDiagonal;
    u->FortranMatrixGetDiagonal = zutilities_FortranMatrixGetDiagonal;
    u->FortranMatrixAdd = zutilities_FortranMatrixAdd;
    u->FortranMatrixDMultiply = zutilities_FortranMatrixDMultiply;
    u->FortranMatrixMultiplyD = zutilities_FortranMatrixMultiplyD;
    u->FortranMatrixMultiply = zutilities_FortranMatrixMultiply;
    u->FortranMatrixFNorm = utilities_FortranMatrixFNorm;
    u->FortranMatrixAbs = zutilities_FortranMatrixAbs;
    u->FortranMatrixValuePtr = zutilities_FortranMatrixValuePtr;
    u->FortranMatrixMaxValue = utilities_FortranMatrixMaxValue;
    u->FortranMatrixSelectBlock = zutilities_FortranMatrixSelectBlock;
    u->FortranMatrixUpperInv = zutilities_FortranMatrixUpperInv;
    u->FortranMatrixPrint = zutilities_FortranMatrixPrint;
    lp->chol = zchol_lapack;
    lp->solveGEVP = zgevp_lapack;
}

}  // extern "C"
