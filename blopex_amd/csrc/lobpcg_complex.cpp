// lobpcg_solve_complex on the device (the reference's complex LOBPCG,
// krylov/lobpcg.c:417-498 with the complex multivector of
// blopex_serial_complex): the Hermitian pencil (A, B) of order n runs on the
// real device solver through the real symmetric pencil of order 2n that
// represents it,
//     A~ = [Re A  -Im A]      x~ = [Re x]
//          [Im A   Re A],           [Im x],
// A~ x~ being the real form of A x.  Its eigenvalues are those of (A, B),
// each twice: x~ and J x~ = [-Im x; Re x] (the real form of i x) belong to
// the same one.  The complex block X (n x k) becomes the real block
// [x~_1 .. x~_k, J x~_1 .. J x~_k] (2n x 2k), whose span is closed under J;
// R, P and every Rayleigh-Ritz basis of the real iteration then stay closed
// under J, so each real iteration spans the realified subspace of the
// complex one: the same Ritz values (each twice), residual norms and
// convergence.  Every Gram product, block update and Rayleigh-Ritz step runs
// on the real MFMA kernels; the price is twice the flops and bytes of
// complex kernels on n x k blocks.  The caller's complex operators see
// complex blocks: each application goes through [Re; Im] <-> complex copies.
//
// At the end, Ritz vector pairs (columns 2i, 2i + 1, ascending values) give
// complex candidates z = a + i b for [a; b]; one per pair is kept, chosen
// in order by a pivoted Cholesky factorisation of their B-Gram matrix (a
// candidate whose B-norm lies mostly in the span of those kept, as the
// partner of a kept one does, is skipped), and the kept ones are
// B-orthonormalised: X = Z_S L^{-H}.  For simple eigenvalues L = I and the
// returned vectors are the candidates themselves.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dense.h"
#include "kernels.h"
#include "lobpcg.h"

extern "C" mv_InterfaceInterpreter *blopex_amd_mv_interpreter(mv_MultiVectorPtr x);
extern "C" int blopex_amd_is_hip_interpreter(const mv_InterfaceInterpreter *i);

namespace bx {

int device_lobpcg(mv_MultiVectorPtr X, void *Ad, void (*opA)(void *, void *, void *), void *Bd,
                  void (*opB)(void *, void *, void *), void *Td,
                  void (*opT)(void *, void *, void *), mv_MultiVectorPtr Y, lobpcg_Tolerance tol,
                  int maxit, int verb, int *iters, double *lam, double *lamh, BlopexInt lamh_gh,
                  double *res, double *resh, BlopexInt resh_gh);
bool device_lobpcg_fits(int k, int guard);

namespace {

using Op = void (*)(void *, void *, void *);

BlopexInt *all_indices() {
    static BlopexInt ids[kMaxCols];
    static bool init = false;
    if (!init) {
        for (int i = 0; i < kMaxCols; i++) ids[i] = i;
        init = true;
    }
    return ids;
}

// a complex view of w contiguous columns of n rows, all active
hip_Multi_Vector cview(void *p, long n, int w) {
    hip_Multi_Vector v;
    memset(&v, 0, sizeof(v));
    v.data = p;
    v.size = n;
    v.ld = n;
    v.num_vectors = w;
    v.num_active_vectors = w;
    v.active_indices = all_indices();
    v.precision = HIPMV_C64;
    return v;
}

// The caller's complex operator, applied to realified blocks: the solver's
// views (2n rows, all columns active) are copied to complex staging, the
// operator runs there, and the result comes back in real form.
struct ZOp {
    void *d;
    Op op;
    long n;
    int wmax;
    void *zx, *zy;  // staging, n x wmax komplex each
};

void zop_apply(void *ctx, void *vx, void *vy) {
    ZOp *z = (ZOp *)ctx;
    const hip_Multi_Vector *x = (const hip_Multi_Vector *)vx;
    hip_Multi_Vector *y = (hip_Multi_Vector *)vy;
    const int w = (int)x->num_vectors;
    BX_REQUIRE(x->size == 2 * z->n && y->size == 2 * z->n && w <= z->wmax &&
                   x->num_active_vectors == w && y->num_vectors == w,
               "complex operator: a %ld x %d real block for %ld complex rows", x->size, w, z->n);
    real_to_z(z->n, w, x->data, x->ld, x->precision, z->zx, z->n);
    hip_Multi_Vector cx = cview(z->zx, z->n, w), cy = cview(z->zy, z->n, w);
    z->op(z->d, &cx, &cy);
    z_to_real(z->n, w, z->zy, z->n, y->data, y->ld, y->precision, false);
}

struct DevBuf {
    void *p = nullptr;
    explicit DevBuf(size_t bytes) { BX_HIP(hipMalloc(&p, bytes > 0 ? bytes : 16)); }
    ~DevBuf() {
        (void)hipStreamSynchronize(stream());
        (void)hipFree(p);
    }
};

struct Realified {
    hip_Multi_Vector *blk = nullptr;
    mv_MultiVectorPtr mv = nullptr;
    ~Realified() {
        if (mv) mv_MultiVectorDestroy(mv);
        if (blk) hip_Multi_VectorDestroy(blk);
    }
};

inline komplex kmul_conj(komplex a, komplex b) {  // conj(a) * b
    return {a.real * b.real + a.imag * b.imag, a.real * b.imag - a.imag * b.real};
}

// Picks k of the m = 2k candidates (in order) by a pivoted Cholesky factor
// of their B-Gram matrix M (m x m, column-major): candidate j is kept when
// more than half of its squared B-norm lies outside the span of those kept
// before it.  Returns the m x k coefficients C with Z C B-orthonormal
// (rows of the skipped candidates zero), or false when fewer than k are kept.
bool select_pairs(const std::vector<komplex> &M, int m, int k, std::vector<komplex> &C) {
    std::vector<int> sel;
    std::vector<komplex> L((size_t)k * k, komplex{0.0, 0.0});  // lower, row s = kept s
    for (int j = 0; j < m && (int)sel.size() < k; j++) {
        const int s = (int)sel.size();
        std::vector<komplex> l(s);
        double nl = 0.0;
        for (int a = 0; a < s; a++) {  // L l = M[sel, j]
            komplex v = M[sel[a] + (size_t)j * m];
            for (int b = 0; b < a; b++) {
                const komplex t = L[a + (size_t)b * k], u = l[b];
                v.real -= t.real * u.real - t.imag * u.imag;
                v.imag -= t.real * u.imag + t.imag * u.real;
            }
            const double d = L[a + (size_t)a * k].real;
            l[a] = {v.real / d, v.imag / d};
            nl += l[a].real * l[a].real + l[a].imag * l[a].imag;
        }
        const double g = M[j + (size_t)j * m].real, d2 = g - nl;
        if (!(g > 0.0) || !(d2 > 0.5 * g)) continue;
        for (int b = 0; b < s; b++) L[s + (size_t)b * k] = {l[b].real, -l[b].imag};  // conj(l)
        L[s + (size_t)s * k] = {std::sqrt(d2), 0.0};
        sel.push_back(j);
    }
    if ((int)sel.size() < k) return false;
    // C_S = L^{-H}: solve L^H W = I (upper triangular) column by column
    std::vector<komplex> W((size_t)k * k, komplex{0.0, 0.0});
    for (int c = 0; c < k; c++)
        for (int r = c; r >= 0; r--) {
            komplex v = {r == c ? 1.0 : 0.0, 0.0};
            for (int q = r + 1; q <= c; q++) {
                // (L^H)[r][q] = conj(L[q][r])
                const komplex t = kmul_conj(L[q + (size_t)r * k], W[q + (size_t)c * k]);
                v.real -= t.real;
                v.imag -= t.imag;
            }
            const double d = L[r + (size_t)r * k].real;
            W[r + (size_t)c * k] = {v.real / d, v.imag / d};
        }
    C.assign((size_t)m * k, komplex{0.0, 0.0});
    for (int r = 0; r < k; r++)
        for (int c = 0; c < k; c++) C[sel[r] + (size_t)c * m] = W[r + (size_t)c * k];
    return true;
}

}  // namespace

int device_lobpcg_complex(mv_MultiVectorPtr X, void *Ad, Op opA, void *Bd, Op opB, void *Td,
                          Op opT, mv_MultiVectorPtr Y, lobpcg_Tolerance tol, int maxit, int verb,
                          int *iters, komplex *lam, komplex *lamh, BlopexInt lamh_gh,
                          double *res, double *resh, BlopexInt resh_gh) {
    auto t0 = std::chrono::steady_clock::now();
    *iters = 0;
    hip_Multi_Vector *x = (hip_Multi_Vector *)mv_MultiVectorGetData(X);
    const long n = x->size;
    const int k = (int)x->num_vectors, m = 2 * k;
    if (k < 1 || !device_lobpcg_fits(m, 0)) {
        set_error("lobpcg_solve_complex: %d columns (the device solver takes 1 to %d)", k,
                  kMaxRR / 6);
        if (verb) fprintf(stderr, "Error in LOBPCG:\n%s\n", blopex_amd_last_error());
        return WRONG_BLOCK_SIZE;
    }
    if (x->settings.guard_columns != 0) {
        set_error("lobpcg_solve_complex: guard columns are not supported for complex blocks");
        if (verb) fprintf(stderr, "Error in LOBPCG:\n%s\n", blopex_amd_last_error());
        return WRONG_BLOCK_SIZE;
    }
    hip_Multi_Vector *y = nullptr;
    int ny = 0;
    if (Y) {
        y = blopex_amd_is_hip_interpreter(blopex_amd_mv_interpreter(Y))
                ? (hip_Multi_Vector *)mv_MultiVectorGetData(Y)
                : nullptr;
        ny = y ? (int)y->num_vectors : 0;
        if (!y || y->precision != HIPMV_C64 || y->size != n || 2 * ny > kMaxCols) {
            set_error("lobpcg_solve_complex: constraints must be a complex HIP block of %ld rows "
                      "and at most %d columns",
                      n, kMaxCols / 2);
            if (verb) fprintf(stderr, "Error in LOBPCG:\n%s\n", blopex_amd_last_error());
            return WRONG_CONSTRAINTS;
        }
    }
    mv_InterfaceInterpreter *ii = blopex_amd_mv_interpreter(X);
    // the realified blocks: [x~ | J x~] from X's columns, likewise Y
    Realified xr, yr;
    xr.blk = hip_Multi_VectorCreate(2 * n, m, HIPMV_FP64);
    BX_REQUIRE(xr.blk, "lobpcg_solve_complex: allocation of a %ld x %d block", 2 * n, m);
    xr.blk->settings = x->settings;
    // every column of X, as lobpcg_solve starts from (its mask is its own business)
    z_to_real(n, k, x->data, x->ld, xr.blk->data, xr.blk->ld, HIPMV_FP64, true);
    xr.mv = mv_MultiVectorWrap(ii, xr.blk, 0);
    if (y && ny > 0) {
        yr.blk = hip_Multi_VectorCreate(2 * n, 2 * ny, HIPMV_FP64);
        BX_REQUIRE(yr.blk, "lobpcg_solve_complex: allocation of a %ld x %d block", 2 * n, 2 * ny);
        z_to_real(n, ny, y->data, y->ld, yr.blk->data, yr.blk->ld, HIPMV_FP64, true);
        yr.mv = mv_MultiVectorWrap(ii, yr.blk, 0);
    }
    // staging of the operator adapters (shared: one application at a time)
    const int wmax = m > 2 * ny ? m : 2 * ny;
    DevBuf zx((size_t)n * wmax * sizeof(komplex)), zy((size_t)n * wmax * sizeof(komplex));
    ZOp za{Ad, opA, n, wmax, zx.p, zy.p}, zb{Bd, opB, n, wmax, zx.p, zy.p},
        zt{Td, opT, n, wmax, zx.p, zy.p};
    if (verb) {
        printf("\nSolving %s eigenvalue problem with%s preconditioning\n\n",
               opB ? "generalized" : "standard", opT ? "" : "out");
        printf("block size %d\n\n", k);
        if (ny == 0)
            printf("No constraints\n\n");
        else
            printf("%d constraint%s\n\n", ny, ny > 1 ? "s" : "");
    }
    std::vector<double> lr(m), rr(m);
    const size_t hist = (size_t)m * (size_t)(maxit > 0 ? maxit + 1 : 1);
    std::vector<double> lhr(lamh ? hist : 0), rhr(resh ? hist : 0);
    const int flag = device_lobpcg(xr.mv, &za, zop_apply, opB ? &zb : nullptr,
                                   opB ? zop_apply : nullptr, opT ? &zt : nullptr,
                                   opT ? zop_apply : nullptr, yr.mv, tol, maxit, 0, iters,
                                   lr.data(), lamh ? lhr.data() : nullptr, m, rr.data(),
                                   resh ? rhr.data() : nullptr, m);
    x->stats = xr.blk->stats;
    if (flag == WRONG_BLOCK_SIZE || flag == WRONG_CONSTRAINTS || flag == PROBLEM_SIZE_TOO_SMALL)
        return flag;
    // pair i: real columns 2i and 2i + 1 (ascending values, each twice)
    for (int i = 0; i < k; i++) {
        lam[i] = {lr[2 * i], 0.0};
        res[i] = rr[2 * i];
    }
    for (int t = 0; t <= *iters && t <= maxit; t++)
        for (int i = 0; i < k; i++) {
            if (lamh) lamh[i + (size_t)t * lamh_gh] = {lhr[2 * i + (size_t)t * m], 0.0};
            if (resh) resh[i + (size_t)t * resh_gh] = rhr[2 * i + (size_t)t * m];
        }
    // complex candidates from the real Ritz vectors, their B-Gram matrix
    real_to_z(n, m, xr.blk->data, xr.blk->ld, HIPMV_FP64, zx.p, n);
    const void *bz = zx.p;
    if (opB) {
        hip_Multi_Vector cz = cview(zx.p, n, m), cbz = cview(zy.p, n, m);
        opB(Bd, &cz, &cbz);
        bz = zy.p;
    }
    DevBuf dm((size_t)m * m * sizeof(komplex));
    zgram_cols(n, zx.p, n, colmap_identity(m), bz, n, colmap_identity(m), dm.p, m);
    std::vector<komplex> M((size_t)m * m), C;
    download_sync(M.data(), dm.p, M.size() * sizeof(komplex));
    if (!select_pairs(M, m, k, C)) {
        set_error("lobpcg_solve_complex: the Ritz vectors did not give %d independent complex "
                  "vectors",
                  k);
        if (verb) fprintf(stderr, "Error in LOBPCG:\n%s\n", blopex_amd_last_error());
        return flag != 0 ? flag : REQUESTED_ACCURACY_NOT_ACHIEVED;
    }
    upload_async(dm.p, C.data(), C.size() * sizeof(komplex));
    zvecmat_cols(n, zx.p, n, colmap_identity(m), dm.p, m, x->data, x->ld, colmap_identity(k),
                 false);
    if (verb == 1) {
        printf("\nEigenvalue lambda       Residual              \n");
        for (int i = 0; i < k; i++) printf("%22.14e  %22.14e\n", lam[i].real, res[i]);
        printf("\n%d iterations\n", *iters);
    }
    BX_HIP(hipStreamSynchronize(stream()));
    x->stats.exit_flag = flag;
    x->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return flag;
}

}  // namespace bx
