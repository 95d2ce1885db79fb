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
// The same with the caller's workspace for the per-workgroup partial tiles
// (part_words doubles; fewer partial sets when it holds fewer than
// gram_cols would use).  Allocates nothing.
void gram_cols_into(int prec, long n, const void *x, long ldx, const ColMap &xm, const void *y,
                    long ldy, const ColMap &ym, double *g, long ldg, double *part,
                    size_t part_words);
// d (device FP64, xm.n) = diag(x(:, xm)' y(:, ym))
void diagdot_cols(int prec, long n, const void *x, long ldx, const ColMap &xm, const void *y,
                  long ldy, const ColMap &ym, double *d);
// y(:, ym) = x(:, xm) * c (+ y(:, ym) if accumulate); c device FP64 h x w, leading dim ldc
void vecmat_cols(int prec, long n, const void *x, long ldx, const ColMap &xm, const double *c,
                 long ldc, void *y, long ldy, const ColMap &ym, bool accumulate);

// r(:, t) = lam[act[t]] * bx(:, act[t]) - ax(:, act[t]), t < act.n (lam on the device)
void residual_cols(int prec, long n, const void *bx, const void *ax, long ld, const ColMap &act,
                   const double *lam, void *r, long ldr);
// The same residual columns, their squared norms d[t] (device FP64, summed
// exactly as diagdot_cols sums them) and, when t is not null, r(:, t) scaled
// by the diagonal t (the Jacobi preconditioner) after the norms are taken:
// one pass instead of residual_cols + diagdot_cols + the diagonal operator.
void residual_norm_cols(int prec, long n, const void *bx, const void *ax, long ld,
                        const ColMap &act, const double *lam, void *r, long ldr, const void *t,
                        double *d);
// in-place compaction: x(:, i) = x(:, src[i]) for i < src.n, src ascending with src[i] >= i
void compact_cols(int prec, long n, void *x, long ld, const ColMap &src);
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
// a no-op when st->abort is set (st may be nullptr).  tail_pos (host, w
// entries, nullptr = identity): column of out_tail that receives output
// column c, -1 to drop it (the tail can be compacted to the active columns
// as it is written).
void block_update(int prec, long ld, const UpdGroup *groups, int ngroups, const double *C, int ldc,
                  int crows, int w, const DevStatus *st, const int *tail_pos = nullptr);
// FP32 or FP64 blocks, B = I, at most 32 output columns: the X group
// (groups[0]) and AX group (groups[1]) of block_update, then
// R = T (X' diag(lam) - AX') into the ntail tail columns of r (tail_pos as
// above; tdiag and r in the blocks' precision) and their squared norms
// (before T) into rn2[0..ntail) -- the update and the residual pass in one.
void block_update_resid(int prec, long n, long ld, const UpdGroup *groups, const double *C,
                        int ldc, int crows, int w, const DevStatus *st, const int *tail_pos,
                        int ntail, const double *lam, const void *tdiag, void *r, double *rn2);
// out[j] = sum of part[j * nparts + 0 .. nparts) in a fixed order, j < ncols
void row_sums(const double *part, int nparts, int ncols, double *out);

// Fused Gram products G = L' Q: every block of a plan is staged through LDS
// a chunk of rows at a time, and each wave accumulates a register block of
// up to 2 x 4 output tiles (16 x 16) on MFMA in FP64 (FP32 inputs: FP32 MFMA
// over each chunk, then FP64).
constexpr int kGramBlocks = 10;    // distinct n-row blocks per plan
constexpr int kGramCols = 288;     // staged columns per plan (sum of block widths)
constexpr int kGramWaves = 8;      // waves per workgroup = register blocks per plan
constexpr int kGramBL = 2, kGramBQ = 4;                    // largest register block (tiles)
constexpr int kGramTiles = kGramWaves * kGramBL * kGramBQ;  // tile slots per plan
struct GramPlan {
    const void *blk[kGramBlocks];
    int bw[kGramBlocks];   // block widths
    int bc0[kGramBlocks];  // first staged column of each block
    int nblk, ncols;
    int nwb;                                  // register blocks in use
    short wl[kGramWaves][kGramBL];            // staged column of each L tile
    short wq[kGramWaves][kGramBQ];            // staged column of each Q tile
    unsigned char wnl[kGramWaves], wnq[kGramWaves];
    // tile (w, i, j) = slot (w * kGramBL + i) * kGramBQ + j
    unsigned char rows[kGramTiles], cols[kGramTiles];  // valid part of the tile
    double *dst[kGramTiles];   // tile origin (FP64, leading dim ldg); null: not stored
    double *mdst[kGramTiles];  // origin of the transposed copy (symmetric products) or null
    // products queued for the plan (blocks are formed when it runs)
    int nprod;
    short pl[16], pq[16];      // block index of L and Q
    double *pg[16];            // output origin
    unsigned char pt[16];      // 1: store (L'Q)' = Q'L at pg
};
// Adds the product L' Q (L: lw columns, Q: qw columns) landing at g (leading
// dim ldg); a product with Q == L (same block and width) is symmetric: only
// its upper tiles are stored, each also as its transpose.  Returns false,
// leaving the plan unchanged, when it does not fit.
bool gram_plan_product(GramPlan &p, const void *L, int lw, const void *Q, int qw, double *g,
                       long ldg, bool transposed = false);
void gram_run(int prec, long ld, GramPlan &p, long ldg);

// Accumulates products and runs the plan whenever it is full.
class GramBatch {
  public:
    GramBatch(int prec, long ld, long ldg) : prec_(prec), ld_(ld), ldg_(ldg), p_() {}
    void add(const void *L, int lw, const void *Q, int qw, double *g);
    // g (qw x lw) receives Q'L, computed as the transpose of L'Q: lets a
    // product join the register blocks of L
    void add_transposed(const void *L, int lw, const void *Q, int qw, double *g);
    void flush();

  private:
    int prec_;
    long ld_, ldg_;
    GramPlan p_;
};

// ---- complex (komplex) blocks: column-major n x k, interleaved FP64 pairs ----
// g (device komplex, h x w, leading dim ldg) = x(:, xm)^H y(:, ym)
void zgram_cols(long n, const void *x, long ldx, const ColMap &xm, const void *y, long ldy,
                const ColMap &ym, void *g, long ldg);
// d (device komplex, xm.n) = diag(x(:, xm)^H y(:, ym))
void zdiagdot_cols(long n, const void *x, long ldx, const ColMap &xm, const void *y, long ldy,
                   const ColMap &ym, void *d);
// y(:, ym) = x(:, xm) c (+ y); c device komplex xm.n x ym.n, leading dim ldc
void zvecmat_cols(long n, const void *x, long ldx, const ColMap &xm, const void *c, long ldc,
                  void *y, long ldy, const ColMap &ym, bool accumulate);
// y(:, ym[j]) = d[j] x(:, xm[j]), d device komplex
void zscale_cols(long n, const void *d, const void *x, long ldx, const ColMap &xm, void *y,
                 long ldy, const ColMap &ym);
// Realified blocks of the complex solve: column j of r (2n rows, precision
// rprec) = [Re z_j; Im z_j] for z (n x w komplex, leading dim ldz); with
// jpair also column w + j of r = [-Im z_j; Re z_j] (the column of i z_j).
void z_to_real(long n, int w, const void *z, long ldz, void *r, long ldr, int rprec, bool jpair);
// z_j = r_j[0:n] + i r_j[n:2n] for the w columns of r
void real_to_z(long n, int w, const void *r, long ldr, int rprec, void *z, long ldz);
// y(:, ym) = A x(:, xm), A dense FP64 n x n (leading dim lda), x and y FP64 or
// FP32 (prec), FP64 arithmetic
void dense_apply(int prec, long n, const double *A, long lda, const void *x, long ldx,
                 const ColMap &xm, void *y, long ldy, const ColMap &ym);
// y(:, ym) = L x(:, xm), L the n x n 1D Laplacian [-1 2 -1] with Dirichlet ends
void lapl1d_apply(int prec, long n, const void *x, long ldx, const ColMap &xm, void *y, long ldy,
                  const ColMap &ym);
// y(:, ym) = A x(:, xm), A dense komplex n x n
void zdense_apply(long n, const void *A, long lda, const void *x, long ldx, const ColMap &xm,
                  void *y, long ldy, const ColMap &ym);

// ---- eigenpair adjoint ----
// y(:, c) = a[j] x(:, c) + b[j] y(:, c) for c = cm.idx[j]; a, b device FP64 (cm.n)
void axpby_cols(int prec, long n, const double *a, const void *x, long ldx, const double *b,
                void *y, long ldy, const ColMap &cm);
// CG scalars on the device (m columns): a = rz / pq, na = -a for running
// columns (done[j] == 0), *bad = j + 1 for a running column with pq <= 0
void cg_alpha(int m, const double *rz, const double *pq, const int *done, double *a, double *na,
              int *bad);
// done[j] |= ||r_j|| <= target[j] (rr = ||r||^2); beta = rzn / rz for running
// columns, 0 for finished; rz <- rzn
void cg_beta(int m, const double *rr, const double *rzn, const double *target, int *done,
             double *rz, double *beta);
// gval[e] = sum_i gl[i] x_i[r] x_i[c] - (w_i[r] x_i[c] + x_i[r] w_i[c]) / 2 over the
// stored entries e = (r, c) of A; x, w n x k FP64 (leading dim ld), gl device
void sampled_outer(const hip_CSRMatrix *A, const double *x, const double *w, long ld, int k,
                   const double *gl, double *gval);
// gb[r] = sum_i x_i[r] (a[i] x_i[r] + lam[i] w_i[r]); a, lam device
void diag_outer(long n, const double *x, const double *w, long ld, int k, const double *a,
                const double *lam, double *gb);

// kernel spmm_csr uses for ncols contiguous columns with leading dimension
// ld: 0 CSR, 1 diagonal copy, 2 z-marching stencil kernel
int spmm_kernel(int prec, const hip_CSRMatrix *A, long ld, int ncols);
// CSR SpMM Y(:, ym) = A X(:, xm)
void spmm_csr(int prec, const hip_CSRMatrix *A, const void *x, long ldx, const ColMap &xm, void *y,
              long ldy, const ColMap &ym);
// Y(:, ym) = diag(d) X(:, xm)
void diag_apply(int prec, const hip_DiagMatrix *D, const void *x, long ldx, const ColMap &xm,
                void *y, long ldy, const ColMap &ym);

}  // namespace bx
