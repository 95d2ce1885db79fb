// FP64 small-dense kernels of the device-resident solver (one workgroup each).
#pragma once

#include "common.h"

namespace bx {

// Device-side status of one solve; kernels after a fatal failure are no-ops.
struct DevStatus {
    int abort;      // set on a failure that ends the solve
    int x_info;     // Cholesky info of X'BX (initial orthonormalisation)
    int r_info;     // Cholesky info of R'BR
    int p_info;     // Cholesky info of P'BP (nonzero: P dropped this iteration)
    int gevp_info;  // Rayleigh-Ritz failure (dsygv-style info)
    int pad[3];
};

constexpr int kCholX = 0, kCholR = 1, kCholP = 2;
constexpr int kMaxRR = 192;              // largest Rayleigh-Ritz pencil (3 x 64)
constexpr size_t kRRLdsBytes = 147456;   // 2 x 96 x 96 doubles

// g (m x m, leading dim ld, device) <- U^{-1} of its Cholesky factor
void chol_inv(double *g, int ld, int m, DevStatus *st, int which);
// Rayleigh-Ritz on the Gram buffers (see dense.hip)
void rr_solve(const double *ga, const double *gb, int lda, int k, int a, int p, int initial,
              const double *lam_in, double *lam_out, double *coef, int ldc, DevStatus *st);

}  // namespace bx
