/*
 * Kernel-level entry points of libblopex_amd.so: the MFMA building blocks of
 * the device-resident solver, callable on host arrays for tests and tuning.
 * They have no counterpart in the reference (whose block operations are the
 * per-column loops of blopex_serial_double/multivector/multi_vector.c:
 * 354-524); each allocates, runs the kernel on the current device and
 * copies the result back.  All arrays are column-major FP64 on the host;
 * precision selects the device arithmetic (HIPMV_FP64 / HIPMV_FP32).
 */
#ifndef BLOPEX_AMD_KERNELS_H
#define BLOPEX_AMD_KERNELS_H

#ifdef __cplusplus
extern "C" {
#endif

/* tail = R*C[k:k+a,:] + P*C[k+a:k+a+p,:], full = tail + X*C[0:k,:]
   (X n x k, R n x a, P n x p, C (k+a+p) x k); a or p may be 0. */
int blopex_amd_kernel_update(int precision, long n, int k, int a, int p, const double *X,
                             const double *R, const double *P, const double *C, double *full,
                             double *tail);

/* The update above for the X group (X, R, P) and the AX group (AX, AR, AP)
   together, with the residual of the new block formed from the registers
   that hold X' and AX' (B = I, k <= 32): Xo = X', AXo = AX'; output column
   c with tail_pos[c] = q >= 0 (q < ntail; -1: none) also gives Po(:, q) and
   APo(:, q) (its P' and AP' columns), Ro(:, q) = t .* (X'(:, c) lam[c] -
   AX'(:, c)) with t = tdiag (n entries) or 1 when tdiag is NULL, and
   rn2[q] = ||X'(:, c) lam[c] - AX'(:, c)||^2.  Po, APo, Ro are n x ntail. */
int blopex_amd_kernel_update_resid(int precision, long n, int k, int a, int p, const double *X,
                                   const double *AX, const double *R, const double *AR,
                                   const double *P, const double *AP, const double *C,
                                   const int *tail_pos, int ntail, const double *lam,
                                   const double *tdiag, double *Xo, double *AXo, double *Po,
                                   double *APo, double *Ro, double *rn2);

/* G (lw x qw) = L' * Q for L n x lw, Q n x qw. */
int blopex_amd_kernel_gram(int precision, long n, int lw, int qw, const double *L, const double *Q,
                           double *G);

/* Rayleigh-Ritz step of the device solver on host Gram buffers ga, gb
   (lda x (k+a+p), raw blocks of [X R P] as the solver stores them, only the
   lower block triangle and the R-R / P-P blocks read): coefficients
   coef ((k+a+p) x k) for the raw blocks and the k smallest Ritz values.
   explicit_x: X'AX, X'BX from the buffers instead of diag(lam_in), I.
   retry_without_p: a pencil that is not positive definite is solved again
   without P (p_info = -1; tridiagonal solver only).
   jacobi != 0 forces the cyclic-Jacobi eigensolver instead of the
   tridiagonal one.  flags = {abort, r_info, p_info, gevp_info}. */
int blopex_amd_kernel_rr(int k, int a, int p, const double *ga, const double *gb, int lda,
                         const double *lam_in, int explicit_x, int retry_without_p, int jacobi,
                         double *lam_out, double *coef, int *flags);
/* The same step with the condition-estimate restart of lobpcg.m:846-863:
   when log10 of the spread of B_n's Cholesky pivots + 1 (written to
   *condest) exceeds cond_limit > 0, the step is solved again without P
   (flags[2] = p_info = -2). */
int blopex_amd_kernel_rr_cond(int k, int a, int p, const double *ga, const double *gb, int lda,
                              const double *lam_in, int explicit_x, double cond_limit,
                              double *lam_out, double *coef, int *flags, double *condest);

/* For each lane l (0..63) and accumulator register r (0..3) of the 16x16x4
   MFMA used by the kernels (FP64 or FP32), the row and column of D it holds:
   rows[4*l+r], cols[4*l+r].  Checks the layout the kernels assume. */
int blopex_amd_kernel_mfma_probe(int precision, int *rows, int *cols);

#ifdef __cplusplus
}
#endif

#endif
