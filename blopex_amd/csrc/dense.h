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

// Rayleigh-Ritz on raw Gram blocks of [X R P], folding in the Cholesky
// normalisation of R and P (see dense.hip); coef gets (k+a+p) x k coordinates
// for the raw blocks.  explicit_x: take X'AX and X'BX from the Gram buffers
// instead of diag(lambda) and I.  Sets st->r_info / abort, st->p_info (P
// dropped) and st->gevp_info.
void rr_solve_implicit(const double *ga, const double *gb, int lda, int k, int a, int p,
                       const double *lam_in, double *lam_out, double *coef, int ldc,
                       DevStatus *st, int explicit_x);

}  // namespace bx
