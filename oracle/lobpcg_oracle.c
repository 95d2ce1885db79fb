/*
 * lobpcg_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64, single-threaded restatement of the BLOPEX LOBPCG solver
 * (blopex_abstract/krylov/lobpcg.c:502-1388) over serial column-major block
 * vectors with an active-column list (blopex_serial_double/multivector/
 * multi_vector.c:22-32, 110-135).  It is the checker the tests, smoke() and
 * bench.py's cpu_baseline leg compare the HIP library against; nothing in the
 * shipped library links, loads or calls it.
 *
 * Operators: A is CSR (there is no CSR operator in the reference; the
 * per-row sum order here is CSR order, one active column at a time, the
 * same loop the oracle/_ref harness uses), B is an optional diagonal, T an
 * optional diagonal preconditioner (T = diag(t), Jacobi when t = 1/diag(A)).
 * LAPACK dpotrf/dsygv come in as function pointers, as in
 * lobpcg_BLASLAPACKFunctions (krylov/lobpcg.h:25-46).
 *
 * Pinned by: the compiled reference (oracle/_ref, same inputs, eigenvalue
 * and residual histories), the closed-form Dirichlet Laplacian spectrum,
 * diag(1..40) -> 1..5 and diag(1..n) with 3 constraints -> 4,5,6
 * (tests/test_oracle.py).
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef void (*orc_dpotrf_fn)(char *uplo, int *n, double *a, int *lda, int *info);
typedef void (*orc_dsygv_fn)(int *itype, char *jobz, char *uplo, int *n, double *a,
                             int *lda, double *b, int *ldb, double *w, double *work,
                             int *lwork, int *info);

/* ---------------------------------------------------------------------------
 * Block of column vectors with an active list (multi_vector.h:22-32).
 * ------------------------------------------------------------------------- */
typedef struct {
    double *v;
    long n;
    int k;
    int na;
    int *act;
} blk;

static blk blk_new(long n, int k)
{
    blk b;
    b.v = (double *)calloc((size_t)n * k, sizeof(double));
    b.n = n;
    b.k = k;
    b.act = (int *)malloc(sizeof(int) * (k > 0 ? k : 1));
    b.na = k;
    for (int i = 0; i < k; i++) b.act[i] = i;
    return b;
}

static blk blk_wrap(double *v, long n, int k)
{
    blk b = blk_new(0, k);
    free(b.v);
    b.v = v;
    b.n = n;
    return b;
}

static void blk_free(blk *b, int owns)
{
    if (owns) free(b->v);
    free(b->act);
}

/* mask semantics of multi_vector.c:110-135: NULL = all columns */
static void blk_mask(blk *b, const int *mask)
{
    b->na = 0;
    for (int i = 0; i < b->k; i++)
        if (!mask || mask[i]) b->act[b->na++] = i;
}

static double *bcol(const blk *b, int j) { return b->v + (size_t)b->act[j] * b->n; }

/* g(0:h,0:w) with leading dim gh = x(act)' * y(act)   (multi_vector.c:354-406) */
static void mv_gram(const blk *x, const blk *y, int gh, double *g)
{
    for (int j = 0; j < y->na; j++) {
        const double *yc = bcol(y, j);
        for (int i = 0; i < x->na; i++) {
            const double *xc = bcol(x, i);
            double s = 0.0;
            for (long r = 0; r < x->n; r++) s += xc[r] * yc[r];
            g[i + (size_t)j * gh] = s;
        }
    }
}

/* y(act) = x(act) * c, c is h x w with leading dim gh, overwrite
   (multi_vector.c:475-524) */
static void mv_vecmat(const blk *x, int gh, const double *c, blk *y)
{
    for (int j = 0; j < y->na; j++) {
        double *yc = bcol(y, j);
        const double *cj = c + (size_t)j * gh;
        const double *x0 = bcol(x, 0);
        for (long r = 0; r < x->n; r++) yc[r] = cj[0] * x0[r];
        for (int i = 1; i < x->na; i++) {
            const double *xi = bcol(x, i);
            double cc = cj[i];
            for (long r = 0; r < x->n; r++) yc[r] += cc * xi[r];
        }
    }
}

static void mv_axpy(double a, const blk *x, blk *y)
{
    for (int j = 0; j < x->na; j++) {
        const double *xc = bcol(x, j);
        double *yc = bcol(y, j);
        for (long r = 0; r < x->n; r++) yc[r] += a * xc[r];
    }
}

static void mv_copy(const blk *x, blk *y)
{
    for (int j = 0; j < x->na; j++)
        memcpy(bcol(y, j), bcol(x, j), sizeof(double) * x->n);
}

/* y(act) = d(mask) .* x(act)   (multi_vector.c:292-349) */
static void mv_bydiag(const blk *x, const int *mask, int kk, const double *d, blk *y)
{
    int t = 0;
    for (int i = 0; i < kk; i++) {
        if (mask && !mask[i]) continue;
        const double *xc = bcol(x, t);
        double *yc = bcol(y, t);
        for (long r = 0; r < x->n; r++) yc[r] = d[i] * xc[r];
        t++;
    }
}

/* d(mask) = diag(x(act)' y(act))   (multi_vector.c:411-469) */
static void mv_diagdot(const blk *x, const blk *y, const int *mask, int kk, double *d)
{
    int t = 0;
    for (int i = 0; i < kk; i++) {
        if (mask && !mask[i]) continue;
        const double *xc = bcol(x, t), *yc = bcol(y, t);
        double s = 0.0;
        for (long r = 0; r < x->n; r++) s += xc[r] * yc[r];
        d[i] = s;
        t++;
    }
}

/* ---------------------------------------------------------------------------
 * Operators.
 * ------------------------------------------------------------------------- */
typedef struct {
    long n;
    const int *rowptr;
    const int *colind;
    const double *val;
} orc_csr;

static void op_csr(const orc_csr *A, const blk *x, blk *y)
{
    for (int j = 0; j < x->na; j++) {
        const double *xc = bcol(x, j);
        double *yc = bcol(y, j);
        for (long r = 0; r < A->n; r++) {
            double s = 0.0;
            for (int e = A->rowptr[r]; e < A->rowptr[r + 1]; e++) s += A->val[e] * xc[A->colind[e]];
            yc[r] = s;
        }
    }
}

static void op_diag(const double *d, const blk *x, blk *y)
{
    for (int j = 0; j < x->na; j++) {
        const double *xc = bcol(x, j);
        double *yc = bcol(y, j);
        for (long r = 0; r < x->n; r++) yc[r] = d[r] * xc[r];
    }
}

/* ---------------------------------------------------------------------------
 * Small column-major matrices (utilities/fortran_matrix.c subset).
 * ------------------------------------------------------------------------- */
typedef struct {
    double *v;
    int gh, h, w;
} sm;

/* 1-based inclusive block, as FortranMatrixSelectBlock (fortran_matrix.c:1312-1334) */
static sm sm_sel(sm m, int i0, int i1, int j0, int j1)
{
    sm b;
    b.gh = m.gh;
    if (i1 < i0 || j1 < j0) {
        b.h = b.w = 0;
        b.v = NULL;
        return b;
    }
    b.h = i1 - i0 + 1;
    b.w = j1 - j0 + 1;
    b.v = m.v + (i0 - 1) + (size_t)(j0 - 1) * m.gh;
    return b;
}

#define SM(m, i, j) ((m).v[(i) + (size_t)(j) * (m).gh])

static void sm_clear(sm m)
{
    for (int j = 0; j < m.w; j++)
        for (int i = 0; i < m.h; i++) SM(m, i, j) = 0.0;
}

static void sm_eye(sm m)
{
    sm_clear(m);
    for (int j = 0; j < m.w && j < m.h; j++) SM(m, j, j) = 1.0;
}

/* (m + m')/2 in place, fortran_matrix.c:428-452 */
static void sm_sym(sm m)
{
    for (int j = 0; j < m.w; j++)
        for (int i = j + 1; i < m.h; i++) {
            double a = (SM(m, i, j) + SM(m, j, i)) * 0.5;
            SM(m, i, j) = a;
            SM(m, j, i) = a;
        }
}

static void sm_clear_lower(sm m)
{
    int w = m.w < m.h ? m.w : m.h;
    for (int j = 0; j + 1 < w; j++)
        for (int i = j + 1; i < m.h; i++) SM(m, i, j) = 0.0;
}

/* inverse of an upper triangle in place: row-oriented back substitution with
   the k-ascending inner sum of fortran_matrix.c:1368-1416 */
static void sm_upper_inv(sm u)
{
    int n = u.h;
    double *d = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        d[i] = SM(u, i, i);
        SM(u, i, i) = 1.0 / d[i];
    }
    for (int i = n - 2; i >= 0; i--)
        for (int j = n - 1; j > i; j--) {
            double v = 0.0;
            for (int m = i + 1; m <= j; m++) v -= SM(u, i, m) * SM(u, m, j);
            SM(u, i, j) = v / d[i];
        }
    free(d);
}

/* c = op(a) * op(b)   (fortran_matrix.c:1016-1077) */
static void sm_mult(sm a, int ta, sm b, int tb, sm c)
{
    int l = ta ? a.h : a.w;
    for (int j = 0; j < c.w; j++)
        for (int i = 0; i < c.h; i++) {
            double s = 0.0;
            for (int q = 0; q < l; q++) {
                double av = ta ? SM(a, q, i) : SM(a, i, q);
                double bv = tb ? SM(b, j, q) : SM(b, q, j);
                s += av * bv;
            }
            SM(c, i, j) = s;
        }
}

static void sm_copy(sm src, sm dst)
{
    for (int j = 0; j < dst.w; j++)
        for (int i = 0; i < dst.h; i++) SM(dst, i, j) = SM(src, i, j);
}

static void sm_set_diag(sm m, const double *d)
{
    for (int j = 0; j < m.w && j < m.h; j++) SM(m, j, j) = d[j];
}

static int sm_chol(sm a, orc_dpotrf_fn potrf)
{
    char uplo = 'U';
    int n = a.h, lda = a.gh, info = 0;
    potrf(&uplo, &n, a.v, &lda, &info);
    return info;
}

/* dsygv itype=1, jobz='V', uplo='L', lwork=10n (krylov/lobpcg.c:66-103) */
static int sm_gevp(sm a, sm b, double *lam, orc_dsygv_fn sygv)
{
    int itype = 1, n = a.h, lda = a.gh, ldb = b.gh, lwork = 10 * n, info = 0;
    char jobz = 'V', uplo = 'L';
    double *work = (double *)calloc((size_t)lwork, sizeof(double));
    sygv(&itype, &jobz, &uplo, &n, a.v, &lda, b.v, &ldb, lam, work, &lwork, &info);
    free(work);
    return info;
}

/* ---------------------------------------------------------------------------
 * Solver state.
 * ------------------------------------------------------------------------- */
typedef struct {
    const orc_csr *A;
    const double *bdiag;
    const double *tdiag;
    orc_dpotrf_fn potrf;
    orc_dsygv_fn sygv;
} orc_ops;

static void applyA(const orc_ops *o, const blk *x, blk *y) { op_csr(o->A, x, y); }
static void applyB(const orc_ops *o, const blk *x, blk *y) { op_diag(o->bdiag, x, y); }

/* B-orthonormalise x given bx = Bx: r = x'bx -> chol U -> U^{-1}; x = (copy of x)*U^{-1}
   (krylov/lobpcg.c:197-232) */
static int implicit_qr(const orc_ops *o, blk *x, blk *bx, sm r, blk *z)
{
    mv_gram(x, bx, r.gh, r.v);
    int info = sm_chol(r, o->potrf);
    if (info) return info;
    sm_upper_inv(r);
    sm_clear_lower(r);
    mv_copy(x, z);
    mv_vecmat(z, r.gh, r.v, x);
    return 0;
}

static int check_residuals(int k, const double *res, const double *lam, double atol,
                           double rtol, int *mask)
{
    int na = 0;
    for (int i = 0; i < k; i++) {
        mask[i] = res[i] > fabs(lam[i]) * rtol + atol + DBL_EPSILON;
        na += mask[i];
    }
    return na;
}

static void residual_norms(const blk *r, const int *mask, int k, double *diag, double *res)
{
    mv_diagdot(r, r, mask, k, diag);
    for (int i = 0; i < k; i++)
        if (!mask || mask[i]) res[i] = sqrt(fabs(diag[i]));
}

static void hist_store(double *h, int gh, int col, const double *v, int k)
{
    if (!h) return;
    for (int i = 0; i < k; i++) h[i + (size_t)col * gh] = v[i];
}

/*
 * Returns the exit flag of lobpcg_solve (lobpcg.h:13-16): 0, 2 (block size),
 * 3 (dependent constraints) or -1 (not converged / failure).
 */
int oracle_lobpcg_solve(long n, int k, double *Xdata, const int *a_rowptr, const int *a_colind,
                        const double *a_val, const double *bdiag, const double *tdiag, int ny,
                        const double *Ydata, double atol, double rtol, int maxit, int verbosity,
                        void *dpotrf_ptr, void *dsygv_ptr, double *lambda, double *lambda_hist,
                        int lambda_hist_gh, double *res, double *res_hist, int res_hist_gh,
                        int *iterations)
{
    orc_csr A = {n, a_rowptr, a_colind, a_val};
    orc_ops o = {&A, bdiag, tdiag, (orc_dpotrf_fn)dpotrf_ptr, (orc_dsygv_fn)dsygv_ptr};
    int noB = bdiag == NULL, noT = tdiag == NULL, noY = ny == 0;
    int exit_flag = 0;
    *iterations = 0;
    if (k < 1) return 2;

    blk X = blk_wrap(Xdata, n, k);
    blk W = blk_new(n, k);
    blk Y = blk_wrap((double *)Ydata, n, ny), BY = Y;
    double *ybx_v = NULL, *tybx_v = NULL, *yby_v = NULL;
    sm yby, ybx, tybx;

    double *rdiag = (double *)calloc((size_t)k, sizeof(double));

    if (!noY) {
        yby_v = (double *)calloc((size_t)ny * ny, sizeof(double));
        ybx_v = (double *)calloc((size_t)ny * k, sizeof(double));
        tybx_v = (double *)calloc((size_t)ny * k, sizeof(double));
        yby = (sm){yby_v, ny, ny, ny};
        ybx = (sm){ybx_v, ny, ny, k};
        tybx = (sm){tybx_v, ny, ny, k};
        if (!noB) {
            BY = blk_new(n, ny);
            applyB(&o, &Y, &BY);
        }
        mv_gram(&BY, &Y, ny, yby_v);
        if (sm_chol(yby, o.potrf)) {
            if (verbosity) printf("Cannot handle linear dependent constraints\n");
            if (!noB) blk_free(&BY, 1);
            free(yby_v); free(ybx_v); free(tybx_v); free(rdiag);
            blk_free(&W, 1);
            blk_free(&X, 0);
            blk_free(&Y, 0);
            return 3;
        }
        sm_upper_inv(yby);
        sm_clear_lower(yby);
        mv_gram(&BY, &X, ny, ybx_v);
        sm_mult(yby, 1, ybx, 0, tybx);
        sm_mult(yby, 0, tybx, 0, ybx);
        mv_vecmat(&Y, ny, ybx_v, &W);
        mv_axpy(-1.0, &W, &X);
    }

    int k3 = 3 * k;
    double *gA_v = (double *)calloc((size_t)k3 * k3, sizeof(double));
    double *gB_v = (double *)calloc((size_t)k3 * k3, sizeof(double));
    double *lamAB = (double *)calloc((size_t)k3, sizeof(double));
    sm gA = {gA_v, k3, k3, k3}, gB = {gB_v, k3, k3, k3};
    int *mask = (int *)malloc(sizeof(int) * k);
    for (int i = 0; i < k; i++) mask[i] = 1;

    blk R = blk_new(n, k), P = blk_new(n, k), AX = blk_new(n, k), AR = blk_new(n, k),
        AP = blk_new(n, k);
    blk BX = X, BR = R, BP = P;
    if (!noB) {
        BX = blk_new(n, k);
        BR = blk_new(n, k);
        BP = blk_new(n, k);
    }
    /* mask bookkeeping is per block; with B = I the B-blocks alias their partners */
    blk *Rm = &R, *Pm = &P, *BRm = noB ? &R : &BR, *BPm = noB ? &P : &BP, *BXm = noB ? &X : &BX;

    blk_mask(&R, mask); blk_mask(&P, mask); blk_mask(&AR, mask); blk_mask(&AP, mask);
    if (!noB) { blk_mask(&BR, mask); blk_mask(&BP, mask); }
    blk_mask(&W, mask);

    sm upper = sm_sel(gB, 1, k, 1, k);
    if (!noB) applyB(&o, &X, BXm);
    exit_flag = implicit_qr(&o, &X, BXm, upper, &W);
    if (exit_flag) {
        if (verbosity) fprintf(stderr, "Bad initial vectors: orthonormalization failed\n");
    } else {
        if (!noB) {
            mv_vecmat(BXm, upper.gh, upper.v, &W);
            mv_copy(&W, BXm);
        }
        applyA(&o, &X, &AX);
        sm xax = sm_sel(gA, 1, k, 1, k);
        mv_gram(&X, &AX, k3, xax.v);
        sm_sym(xax);
        sm xbx = sm_sel(gB, 1, k, 1, k);
        mv_gram(&X, BXm, k3, xbx.v);
        sm_sym(xbx);
        if ((exit_flag = sm_gevp(xax, xbx, lambda, o.sygv)) != 0) {
            if (verbosity) fprintf(stderr, "Bad problem: Rayleigh-Ritz in the initial subspace failed\n");
        } else {
            sm cx = sm_sel(xax, 1, k, 1, k);
            mv_vecmat(&X, cx.gh, cx.v, &W);
            mv_copy(&W, &X);
            mv_vecmat(&AX, cx.gh, cx.v, &W);
            mv_copy(&W, &AX);
            if (!noB) {
                mv_vecmat(BXm, cx.gh, cx.v, &W);
                mv_copy(&W, BXm);
            }
            mv_bydiag(BXm, NULL, k, lambda, &R);
            mv_axpy(-1.0, &AX, &R);
            residual_norms(&R, NULL, k, rdiag, res);
            hist_store(lambda_hist, lambda_hist_gh, 0, lambda, k);
            hist_store(res_hist, res_hist_gh, 0, res, k);
        }
    }

    int it;
    for (it = 1; exit_flag == 0 && it <= maxit; it++) {
        int nr = check_residuals(k, res, lambda, atol, rtol, mask);
        if (nr < 1) break;
        blk_mask(&R, mask); blk_mask(&P, mask); blk_mask(&AR, mask); blk_mask(&AP, mask);
        if (!noB) { blk_mask(&BR, mask); blk_mask(&BP, mask); }
        blk_mask(&W, mask);

        if (!noT) {
            op_diag(o.tdiag, Rm, &W);
            mv_copy(&W, Rm);
        }
        if (!noY) {
            sm ybr = sm_sel(ybx, 1, ny, 1, nr), tybr = sm_sel(tybx, 1, ny, 1, nr);
            mv_gram(&BY, Rm, ybr.gh, ybr.v);
            sm_mult(yby, 1, ybr, 0, tybr);
            sm_mult(yby, 0, tybr, 0, ybr);
            mv_vecmat(&Y, ybr.gh, ybr.v, &W);
            mv_axpy(-1.0, &W, Rm);
        }
        int r0 = k + 1, r1 = k + nr, p0 = r1 + 1, p1, np;
        upper = sm_sel(gB, r0, r1, r0, r1);
        if (!noB) applyB(&o, Rm, BRm);
        exit_flag = implicit_qr(&o, Rm, BRm, upper, &W);
        if (exit_flag) {
            if (verbosity) fprintf(stderr, "Orthonormalization of residuals failed\n");
            break;
        }
        if (!noB) {
            mv_vecmat(BRm, upper.gh, upper.v, &W);
            mv_copy(&W, BRm);
        }
        applyA(&o, Rm, &AR);

        if (it > 1) {
            np = nr;
            p1 = r1 + np;
            upper = sm_sel(gB, p0, p1, p0, p1);
            if (implicit_qr(&o, Pm, BPm, upper, &W)) {
                /* a rank-deficient P drops out of this iteration's basis and P is
                   rebuilt from R alone (steepest-descent restart); the reference
                   keeps the stale P-block bounds here (krylov/lobpcg.c:1056-1063,
                   1200) and reads past the Ritz block */
                np = 0;
            } else {
                if (!noB) {
                    mv_vecmat(BPm, upper.gh, upper.v, &W);
                    mv_copy(&W, BPm);
                }
                mv_vecmat(&AP, upper.gh, upper.v, &W);
                mv_copy(&W, &AP);
            }
        } else {
            np = 0;
        }
        p1 = r1 + np;
        int na = r1 + np;

        sm xax = sm_sel(gA, 1, k, 1, k), rax = sm_sel(gA, r0, r1, 1, k),
           rar = sm_sel(gA, r0, r1, r0, r1), xbx = sm_sel(gB, 1, k, 1, k),
           rbx = sm_sel(gB, r0, r1, 1, k), rbr = sm_sel(gB, r0, r1, r0, r1);
        sm_clear(xax);
        sm_set_diag(xax, lambda);
        mv_gram(Rm, &AX, rax.gh, rax.v);
        mv_gram(Rm, &AR, rar.gh, rar.v);
        sm_sym(rar);
        sm_eye(xbx);
        mv_gram(Rm, BXm, rbx.gh, rbx.v);
        sm_eye(rbr);
        if (np > 0) {
            sm pax = sm_sel(gA, p0, p1, 1, k), par = sm_sel(gA, p0, p1, r0, r1),
               pap = sm_sel(gA, p0, p1, p0, p1), pbx = sm_sel(gB, p0, p1, 1, k),
               pbr = sm_sel(gB, p0, p1, r0, r1), pbp = sm_sel(gB, p0, p1, p0, p1);
            mv_gram(Pm, &AX, pax.gh, pax.v);
            mv_gram(Pm, &AR, par.gh, par.v);
            mv_gram(Pm, &AP, pap.gh, pap.v);
            sm_sym(pap);
            mv_gram(Pm, BXm, pbx.gh, pbx.v);
            mv_gram(Pm, BRm, pbr.gh, pbr.v);
            sm_eye(pbp);
        }
        sm ga = sm_sel(gA, 1, na, 1, na), gb = sm_sel(gB, 1, na, 1, na);
        if ((exit_flag = sm_gevp(ga, gb, lamAB, o.sygv)) != 0) {
            if (verbosity) fprintf(stderr, "GEVP solver failure\n");
            it--;
            break;
        }
        memcpy(lambda, lamAB, sizeof(double) * k);
        sm cx = sm_sel(gA, 1, na, 1, k);
        sm cxx = sm_sel(cx, 1, k, 1, k), crx = sm_sel(cx, r0, r1, 1, k);

        if (np > 0) {
            sm cpx = sm_sel(cx, p0, p1, 1, k);
            blk_mask(&W, NULL);
            mv_vecmat(Pm, cpx.gh, cpx.v, &W);
            blk_mask(Pm, NULL);
            mv_copy(&W, Pm);
            mv_vecmat(&AP, cpx.gh, cpx.v, &W);
            blk_mask(&AP, NULL);
            mv_copy(&W, &AP);
            if (!noB) {
                mv_vecmat(BPm, cpx.gh, cpx.v, &W);
                blk_mask(BPm, NULL);
                mv_copy(&W, BPm);
            }
            mv_vecmat(Rm, crx.gh, crx.v, &W);
            mv_axpy(1.0, &W, Pm);
            mv_vecmat(&AR, crx.gh, crx.v, &W);
            mv_axpy(1.0, &W, &AP);
            if (!noB) {
                mv_vecmat(BRm, crx.gh, crx.v, &W);
                mv_axpy(1.0, &W, BPm);
            }
        } else {
            blk_mask(Pm, NULL);
            mv_vecmat(Rm, crx.gh, crx.v, Pm);
            blk_mask(&AP, NULL);
            mv_vecmat(&AR, crx.gh, crx.v, &AP);
            if (!noB) {
                blk_mask(BPm, NULL);
                mv_vecmat(BRm, crx.gh, crx.v, BPm);
            }
        }
        blk_mask(&W, NULL);
        mv_copy(&X, &W);
        mv_vecmat(&W, cxx.gh, cxx.v, &X);
        mv_axpy(1.0, Pm, &X);
        mv_copy(&AX, &W);
        mv_vecmat(&W, cxx.gh, cxx.v, &AX);
        mv_axpy(1.0, &AP, &AX);
        if (!noB) {
            mv_copy(BXm, &W);
            mv_vecmat(&W, cxx.gh, cxx.v, BXm);
            mv_axpy(1.0, BPm, BXm);
        }
        blk_mask(&AX, mask);
        blk_mask(BXm, mask);
        mv_bydiag(BXm, mask, k, lambda, Rm);
        mv_axpy(-1.0, &AX, Rm);
        residual_norms(Rm, mask, k, rdiag, res);
        hist_store(lambda_hist, lambda_hist_gh, it, lambda, k);
        hist_store(res_hist, res_hist_gh, it, res, k);
        if (verbosity == 1) {
            double mx = res[0];
            for (int i = 1; i < k; i++) mx = res[i] > mx ? res[i] : mx;
            printf("Iteration %d \tbsize %d \tmaxres %22.14e\n", it, nr, mx);
        }
        blk_mask(&AX, NULL);
        blk_mask(BXm, NULL);
        blk_mask(&AP, mask);
        if (!noB) blk_mask(&BP, mask);
        blk_mask(Pm, mask);
        blk_mask(&W, mask);
    }
    if (exit_flag != 0 || it > maxit) exit_flag = -1;
    *iterations = it - 1;

    blk_free(&R, 1); blk_free(&P, 1); blk_free(&AX, 1); blk_free(&AR, 1); blk_free(&AP, 1);
    if (!noB) {
        blk_free(&BX, 1); blk_free(&BR, 1); blk_free(&BP, 1);
        if (!noY) blk_free(&BY, 1);
    }
    blk_free(&W, 1);
    blk_free(&X, 0);
    blk_free(&Y, 0);
    free(gA_v); free(gB_v); free(lamAB); free(mask); free(rdiag);
    free(yby_v); free(ybx_v); free(tybx_v);
    return exit_flag;
}

/* ---------------------------------------------------------------------------
 * Helpers the tests call directly.
 * ------------------------------------------------------------------------- */

/* X0 exactly as serial_Multi_VectorSetRandomValues (multi_vector.c:166-184):
   srand48(seed), then 2*drand48()-1 column by column */
void oracle_drand48_fill(double *x, long n, int k, int seed)
{
    srand48(seed);
    for (size_t i = 0; i < (size_t)n * k; i++) x[i] = 2.0 * drand48() - 1.0;
}

/* Y = A X for every column, CSR order */
void oracle_spmm(long n, int k, const int *rowptr, const int *colind, const double *val,
                 const double *x, double *y)
{
    orc_csr A = {n, rowptr, colind, val};
    blk xb = blk_wrap((double *)x, n, k), yb = blk_wrap(y, n, k);
    op_csr(&A, &xb, &yb);
    blk_free(&xb, 0);
    blk_free(&yb, 0);
}

/* the oracle's inverse of an upper triangle (n x n, leading dim ld), in place */
void oracle_upper_inv(int n, int ld, double *u)
{
    sm m = {u, ld, n, n};
    sm_upper_inv(m);
}
