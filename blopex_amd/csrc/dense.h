// FP64 small-dense kernels of the device-resident solver (one workgroup each).
#pragma once

#include "common.h"

namespace bx {

constexpr int kCholX = 0, kCholR = 1, kCholP = 2;
constexpr int kMaxRR = 192;              // largest Rayleigh-Ritz pencil (3 x 64)
constexpr size_t kRRLdsBytes = 147456;   // 2 x 96 x 96 doubles

// g (m x m, leading dim ld, device) <- U^{-1} of its Cholesky factor
void chol_inv(double *g, int ld, int m, DevStatus *st, int which);
// Rayleigh-Ritz on the Gram buffers (see dense.hip)
void rr_solve(const double *ga, const double *gb, int lda, int k, int a, int p, int initial,
              const double *lam_in, double *lam_out, double *coef, int ldc, DevStatus *st);

struct RR2Args {
    const double *ga, *gb;  // raw Gram buffers, leading dim lda
    int lda;
    int k, a, p;            // block widths (p = 0: no P this iteration)
    const double *lam_in;   // Ritz values of X (XAX = diag)
    double *lam_out;
    double *coef;           // (k+a+p) x k, leading dim ldc
    int ldc;
    double *work;           // global scratch
    DevStatus *st;
    int explicit_x;         // 1: X'AX, X'BX from the Gram buffers (FP32 iterates)
    int retry_without_p;    // 1: a pencil that is not positive definite is solved again
                            // without P (p_info = -1); the tridiagonal solver only
    double cond_limit;      // > 0: a pencil whose B-part condition estimate (log10 of the
                            // Cholesky pivot spread + 1, into st->condest) exceeds it is
                            // solved again without P (p_info = -2); tridiagonal solver only
};

// pencils the one-workgroup tridiagonal solver takes (rr_fast.hip): both
// pencil matrices in LDS up to 96 (k <= 32); the large variant up to 192
// (k, a, p <= 64) keeps A_n in global memory and B_n in LDS while it fits
constexpr int kFastRRN = 96;
constexpr int kFastRRK = 32;
constexpr int kBigRRN = 192;
constexpr int kBigRRK = 64;
bool rr_fast_fits(int k, int a, int p);
void rr_fast(const RR2Args &r);
bool rr_big_fits(int k, int a, int p);
void rr_big(const RR2Args &r);
// work words of one pencil of order up to n
size_t rr_fast_work_words(int n);
// The same solver over a batch: workgroup i takes problem act[i] (i < *nact)
// with k = a = kvec[p] (and p = k when with_p), its buffers at p times the
// strides (Gram buffers s_gram, Ritz values s_lam, coefficients s_coef, work
// s_work, one DevStatus each); grid nprob.  Problems with run[p] == 0 are
// skipped (run may be null).
void rr_fast_batched(const RR2Args &base, const int *act, const int *nact, const int *kvec,
                     const int *run, int nprob, int kmax, bool with_p, long s_gram, long s_lam,
                     long s_coef, long s_work);

// Rayleigh-Ritz on raw Gram blocks of [X R P], folding in the Cholesky
// normalisation of R and P (see dense.hip); coef gets (k+a+p) x k coordinates
// for the raw blocks.  explicit_x: take X'AX and X'BX from the Gram buffers
// instead of diag(lambda) and I.  retry_without_p: see RR2Args.  Sets
// st->r_info / abort, st->p_info (P dropped) and st->gevp_info.
void rr_solve_implicit(const double *ga, const double *gb, int lda, int k, int a, int p,
                       const double *lam_in, double *lam_out, double *coef, int ldc,
                       DevStatus *st, int explicit_x, int retry_without_p = 0,
                       double cond_limit = 0.0);

}  // namespace bx
