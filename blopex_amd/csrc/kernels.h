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

// ---- MFMA kernels of the device solver (blocks with leading dimension ld,
// a multiple of 64 whose padding rows are zero) ----

// One group of the block update: tail = in1 C1 + in2 C2, full = tail + in0 C0,
// C_b = C[crow_b : crow_b + h_b, 0:w].  Absent inputs / outputs are nullptr.
struct UpdGroup {
    const void *in[3];
    int h[3];
    int crow[3];
    void *out_tail;
    void *out_full;
};
// Up to 3 groups in one launch; C device FP64 (crows x w, leading dim ldc);
// a no-op when st->abort is set (st may be nullptr).
void block_update(int prec, long ld, const UpdGroup *groups, int ngroups, const double *C, int ldc,
                  int crows, int w, const DevStatus *st);

// Fused Gram products G = L' Q over 16-column tiles, FP64 accumulation
// (FP32 inputs: FP32 MFMA over 64 rows, then FP64).
constexpr int kGramBlocks = 12;  // distinct n-row blocks
constexpr int kGramJobs = 8;     // left tiles
constexpr int kGramQ = 14;       // right tiles per left tile
constexpr int kGramOut = 96;     // output tiles
struct GramPlan {
    const void *blk[kGramBlocks];
    int blkw[kGramBlocks];
    int nblk;
    int njobs;
    int lblk[kGramJobs], ltile[kGramJobs], nq[kGramJobs];
    int qblk[kGramJobs][kGramQ], qtile[kGramJobs][kGramQ], qout[kGramJobs][kGramQ];
    int ntiles;
    double *dst[kGramOut];  // output tile origin (FP64, leading dim ldg)
    int rows[kGramOut], cols[kGramOut];
};
// Adds the product L' Q (L: lw columns, Q: qw columns) landing at g (leading
// dim ldg); returns false, leaving the plan unchanged, when it does not fit.
bool gram_plan_product(GramPlan &p, const void *L, int lw, const void *Q, int qw, double *g,
                       long ldg);
void gram_run(int prec, long ld, GramPlan &p, long ldg);

// Accumulates products and runs the plan whenever it is full.
class GramBatch {
  public:
    GramBatch(int prec, long ld, long ldg) : prec_(prec), ld_(ld), ldg_(ldg), p_() {}
    void add(const void *L, int lw, const void *Q, int qw, double *g);
    void flush();

  private:
    int prec_;
    long ld_, ldg_;
    GramPlan p_;
};

// CSR SpMM Y(:, ym) = A X(:, xm)
void spmm_csr(int prec, const hip_CSRMatrix *A, const void *x, long ldx, const ColMap &xm, void *y,
              long ldy, const ColMap &ym);
// Y(:, ym) = diag(d) X(:, xm)
void diag_apply(int prec, const hip_DiagMatrix *D, const void *x, long ldx, const ColMap &xm,
                void *y, long ldy, const ColMap &ym);

}  // namespace bx
