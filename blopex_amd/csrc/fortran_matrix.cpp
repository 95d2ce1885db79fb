// Small column-major host matrices (double precision subset of
// blopex_abstract/utilities/fortran_matrix.c).  Entries are addressed
// (i, j) -> value[i + j*globalHeight]; blocks are views that share storage.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fortran_matrix.h"

namespace {
inline double &at(utilities_FortranMatrix *m, long i, long j) {
    return ((double *)m->value)[i + j * m->globalHeight];
}
}  // namespace

extern "C" {

utilities_FortranMatrix *utilities_FortranMatrixCreate(void) {
    utilities_FortranMatrix *m = (utilities_FortranMatrix *)malloc(sizeof(utilities_FortranMatrix));
    BlopexAssert(m != NULL);
    m->globalHeight = m->height = m->width = 0;
    m->value = NULL;
    m->ownsValues = 0;
    return m;
}

void utilities_FortranMatrixAllocateData(long h, long w, utilities_FortranMatrix *m) {
    BlopexAssert(h > 0 && w > 0);
    if (m->value && m->ownsValues) free(m->value);
    m->value = calloc((size_t)h * w, sizeof(double));
    BlopexAssert(m->value != NULL);
    m->globalHeight = m->height = h;
    m->width = w;
    m->ownsValues = 1;
}

void utilities_FortranMatrixWrap(void *v, long gh, long h, long w, utilities_FortranMatrix *m) {
    BlopexAssert(h > 0 && w > 0);
    if (m->value && m->ownsValues) free(m->value);
    m->value = v;
    BlopexAssert(m->value != NULL);
    m->globalHeight = gh;
    m->height = h;
    m->width = w;
    m->ownsValues = 0;
}

void utilities_FortranMatrixDestroy(utilities_FortranMatrix *m) {
    if (!m) return;
    if (m->ownsValues && m->value) free(m->value);
    free(m);
}

long utilities_FortranMatrixGlobalHeight(utilities_FortranMatrix *m) { return m->globalHeight; }
long utilities_FortranMatrixHeight(utilities_FortranMatrix *m) { return m->height; }
long utilities_FortranMatrixWidth(utilities_FortranMatrix *m) { return m->width; }
void *utilities_FortranMatrixValues(utilities_FortranMatrix *m) { return m->value; }

void utilities_FortranMatrixClear(utilities_FortranMatrix *m) {
    for (long j = 0; j < m->width; j++)
        for (long i = 0; i < m->height; i++) at(m, i, j) = 0.0;
}

void utilities_FortranMatrixClearL(utilities_FortranMatrix *m) {
    long w = m->width < m->height ? m->width : m->height;
    for (long j = 0; j + 1 < w; j++)
        for (long i = j + 1; i < m->height; i++) at(m, i, j) = 0.0;
}

void utilities_FortranMatrixSetToIdentity(utilities_FortranMatrix *m) {
    utilities_FortranMatrixClear(m);
    for (long j = 0; j < m->width && j < m->height; j++) at(m, j, j) = 1.0;
}

void utilities_FortranMatrixTransposeSquare(utilities_FortranMatrix *m) {
    BlopexAssert(m->height == m->width);
    for (long j = 0; j < m->width; j++)
        for (long i = j + 1; i < m->height; i++) {
            double t = at(m, i, j);
            at(m, i, j) = at(m, j, i);
            at(m, j, i) = t;
        }
}

void utilities_FortranMatrixSymmetrize(utilities_FortranMatrix *m) {
    BlopexAssert(m->height == m->width);
    for (long j = 0; j < m->width; j++)
        for (long i = j + 1; i < m->height; i++) {
            double a = (at(m, i, j) + at(m, j, i)) * 0.5;
            at(m, i, j) = a;
            at(m, j, i) = a;
        }
}

void utilities_FortranMatrixCopy(utilities_FortranMatrix *src, BlopexInt t,
                                 utilities_FortranMatrix *dst) {
    if (t == 0) {
        BlopexAssert(src->height == dst->height && src->width == dst->width);
        for (long j = 0; j < dst->width; j++)
            for (long i = 0; i < dst->height; i++) at(dst, i, j) = at(src, i, j);
    } else {
        BlopexAssert(src->height == dst->width && src->width == dst->height);
        for (long j = 0; j < dst->width; j++)
            for (long i = 0; i < dst->height; i++) at(dst, i, j) = at(src, j, i);
    }
}

void utilities_FortranMatrixIndexCopy(BlopexInt *index, utilities_FortranMatrix *src, BlopexInt t,
                                      utilities_FortranMatrix *dst) {
    // column j of dst <- column index[j]-1 of src (row index[j]-1 when t != 0)
    for (long j = 0; j < dst->width; j++) {
        long c = index[j] - 1;
        for (long i = 0; i < dst->height; i++) at(dst, i, j) = t == 0 ? at(src, i, c) : at(src, c, i);
    }
}

void utilities_FortranMatrixSetDiagonal(utilities_FortranMatrix *m, utilities_FortranMatrix *d) {
    BlopexAssert(d->height >= m->height);
    for (long j = 0; j < m->width && j < m->height; j++) at(m, j, j) = ((double *)d->value)[j];
}

void utilities_FortranMatrixGetDiagonal(utilities_FortranMatrix *m, utilities_FortranMatrix *d) {
    BlopexAssert(d->height >= m->height);
    for (long j = 0; j < m->width && j < m->height; j++) ((double *)d->value)[j] = at(m, j, j);
}

void utilities_FortranMatrixAdd(double a, utilities_FortranMatrix *A, utilities_FortranMatrix *B,
                                utilities_FortranMatrix *C) {
    BlopexAssert(A->height == C->height && B->height == C->height && A->width == C->width &&
                 B->width == C->width);
    for (long j = 0; j < C->width; j++)
        for (long i = 0; i < C->height; i++) at(C, i, j) = a * at(A, i, j) + at(B, i, j);
}

void utilities_FortranMatrixDMultiply(utilities_FortranMatrix *d, utilities_FortranMatrix *m) {
    BlopexAssert(d->height == m->height);
    for (long j = 0; j < m->width; j++)
        for (long i = 0; i < m->height; i++) at(m, i, j) *= ((double *)d->value)[i];
}

void utilities_FortranMatrixMultiplyD(utilities_FortranMatrix *m, utilities_FortranMatrix *d) {
    BlopexAssert(d->height == m->width);
    for (long j = 0; j < m->width; j++)
        for (long i = 0; i < m->height; i++) at(m, i, j) *= ((double *)d->value)[j];
}

void utilities_FortranMatrixMultiply(utilities_FortranMatrix *A, BlopexInt tA,
                                     utilities_FortranMatrix *B, BlopexInt tB,
                                     utilities_FortranMatrix *C) {
    long l = tA ? A->height : A->width;
    BlopexAssert((tA ? A->width : A->height) == C->height);
    BlopexAssert((tB ? B->width : B->height) == l && (tB ? B->height : B->width) == C->width);
    for (long j = 0; j < C->width; j++)
        for (long i = 0; i < C->height; i++) {
            double s = 0.0;
            for (long q = 0; q < l; q++)
                s += (tA ? at(A, q, i) : at(A, i, q)) * (tB ? at(B, j, q) : at(B, q, j));
            at(C, i, j) = s;
        }
}

double utilities_FortranMatrixFNorm(utilities_FortranMatrix *m) {
    double s = 0.0;
    for (long j = 0; j < m->width; j++)
        for (long i = 0; i < m->height; i++) s += at(m, i, j) * at(m, i, j);
    return sqrt(s);
}

double utilities_FortranMatrixAbs(utilities_FortranMatrix *m, long i, long j) {
    BlopexAssert(i > 0 && i <= m->height && j > 0 && j <= m->width);
    return fabs(at(m, i - 1, j - 1));
}

void *utilities_FortranMatrixValuePtr(utilities_FortranMatrix *m, long i, long j) {
    BlopexAssert(i > 0 && i <= m->height && j > 0 && j <= m->width);
    return &at(m, i - 1, j - 1);
}

double utilities_FortranMatrixMaxValue(utilities_FortranMatrix *m) {
    double mx = at(m, 0, 0);
    for (long j = 0; j < m->width; j++)
        for (long i = 0; i < m->height; i++)
            if (at(m, i, j) > mx) mx = at(m, i, j);
    return mx;
}

void utilities_FortranMatrixSelectBlock(utilities_FortranMatrix *m, long iFrom, long iTo, long jFrom,
                                        long jTo, utilities_FortranMatrix *b) {
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
    b->value = (double *)m->value + (iFrom - 1) + (jFrom - 1) * m->globalHeight;
}

// Inverse of an upper triangle in place.  Row i of the inverse (from the
// bottom) is -sum_{m>i} U(i,m) * Uinv(m,j) / U(i,i), summed with m ascending.
void utilities_FortranMatrixUpperInv(utilities_FortranMatrix *u) {
    long n = u->height;
    BlopexAssert(u->width == n);
    std::vector<double> d(n > 0 ? n : 1);
    for (long i = 0; i < n; i++) {
        d[i] = at(u, i, i);
        at(u, i, i) = 1.0 / d[i];
    }
    for (long i = n - 2; i >= 0; i--)
        for (long j = n - 1; j > i; j--) {
            double v = 0.0;
            for (long m = i + 1; m <= j; m++) v -= at(u, i, m) * at(u, m, j);
            at(u, i, j) = v / d[i];
        }
}

BlopexInt utilities_FortranMatrixPrint(utilities_FortranMatrix *m, char fileName[]) {
    FILE *f = fopen(fileName, "w");
    if (!f) return 1;
    fprintf(f, "%ld\n%ld\n", m->height, m->width);
    for (long j = 0; j < m->width; j++)
        for (long i = 0; i < m->height; i++) fprintf(f, "%.14e\n", at(m, i, j));
    fclose(f);
    return 0;
}

}  // extern "C"
