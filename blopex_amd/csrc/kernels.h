// Launch wrappers of the HIP kernels (all enqueue on bx::stream()).
#pragma once

#include "common.h"

namespace bx {

// Column-mapped block operations on n-row column-major blocks with leading
// dimension ld; `prec` is HIPMV_FP64 or HIPMV_FP32 for both operands.
void copy_cols(int prec, long n, const void *x, long ldx, const ColMap &xm, void *y, long ldy,
               const ColMap &ym);
void axpy_cols(int prec, long n, double alpha, const void *x, long ldx, const ColMap &xm, void *y,
               long ldy, const ColMap &ym);
// y(:, ym[j]) = coef[j] * x(:, xm[j]); coef has xm.n doubles (host)
void scale_cols(int prec, long n, const double *coef, const void *x, long ldx, const ColMap &xm,
                void *y, long ldy, const ColMap &ym);
void clear_cols(int prec, long n, void *x, long ldx, const ColMap &xm);
// column j of x (in map order) gets values 2*u-1 of the drand48 stream after
// srand48(seed), stream positions j*n .. j*n+n-1
void drand48_cols(int prec, long n, int seed, void *x, long ldx, const ColMap &xm);
// dst (prec_d) <- src (prec_s), n x k contiguous blocks
void convert_block(long n, int k, const void *src, long lds, int prec_s, void *dst, long ldd,
                   int prec_d);

// g (device FP64, h x w, leading dim ldg) = x(:, xm)' * y(:, ym), FP64 accumulation
void gram_cols(int prec, long n, const void *x, long ldx, const ColMap &xm, const void *y,
               long ldy, const ColMap &ym, double *g, long ldg);
// d (device FP64, xm.n) = diag(x(:, xm)' y(:, ym))
void diagdot_cols(int prec, long n, const void *x, long ldx, const ColMap &xm, const void *y,
                  long ldy, const ColMap &ym, double *d);
// y(:, ym) = x(:, xm) * c (+ y(:, ym) if accumulate); c device FP64 h x w, leading dim ldc
void vecmat_cols(int prec, long n, const void *x, long ldx, const ColMap &xm, const double *c,
                 long ldc, void *y, long ldy, const ColMap &ym, bool accumulate);

// r(:, t) = lam[act[t]] * bx(:, act[t]) - ax(:, act[t]), t < act.n (lam on the device)
void residual_cols(int prec, long n, const void *bx, const void *ax, long ld, const ColMap &act,
                   const double *lam, void *r, long ldr);
// c (m x w) = alpha * op(a) * op(b), all device FP64 column-major, inner size l
void small_mm(double alpha, const double *a, long lda, bool ta, const double *b, long ldb, bool tb,
              double *c, long ldc, int m, int w, int l);

// CSR SpMM Y(:, ym) = A X(:, xm)
void spmm_csr(int prec, const hip_CSRMatrix *A, const void *x, long ldx, const ColMap &xm, void *y,
              long ldy, const ColMap &ym);
// Y(:, ym) = diag(d) X(:, xm)
void diag_apply(int prec, const hip_DiagMatrix *D, const void *x, long ldx, const ColMap &xm,
                void *y, long ldy, const ColMap &ym);

}  // namespace bx
