// Kernels of the eigenpair adjoint: per-column axpby for the deflated
// conjugate gradients, and the sampled outer products that turn the
// eigenvectors X and the adjoint vectors W into gradients with respect to the
// stored values of A and the diagonal of B.
#include "kernels.h"

namespace bx {

namespace {

constexpr int kThreads = 256;

unsigned rows_grid(long n) {
    size_t blocks = ((size_t)n + kThreads - 1) / kThreads;
    size_t cap = (size_t)num_cus() * 8;
    return (unsigned)(blocks < 1 ? 1 : (blocks > cap ? cap : blocks));
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_axpby(long n, const double *a, const T *x, long ldx,
                                                    const double *b, T *y, long ldy, ColMap cm) {
    const int j = blockIdx.y;
    const int c = cm.idx[j];
    const T aj = (T)a[j], bj = (T)b[j];
    const T *xc = x + (size_t)c * ldx;
    T *yc = y + (size_t)c * ldy;
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads)
        yc[r] = aj * xc[r] + bj * yc[r];
}

// One thread per row r; for each stored entry (r, c):
//   sum_i gl[i] x_i[r] x_i[c] - (w_i[r] x_i[c] + x_i[r] w_i[c]) / 2
__global__ void __launch_bounds__(kThreads) k_sampled_outer(long n, const int *rowptr,
                                                            const int *colind, const double *x,
                                                            const double *w, long ld, int k,
                                                            const double *gl, double *gval) {
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads) {
        const int e0 = rowptr[r], e1 = rowptr[r + 1];
        for (int e = e0; e < e1; e++) {
            const long c = colind[e];
            double s = 0.0;
            for (int i = 0; i < k; i++) {
                const double *xi = x + (size_t)i * ld, *wi = w + (size_t)i * ld;
                const double xr = xi[r], xc = xi[c];
                s = fma(gl[i] * xr, xc, s);
                s = fma(-0.5, fma(wi[r], xc, xr * wi[c]), s);
            }
            gval[e] = s;
        }
    }
}

// gb[r] = sum_i x_i[r] (a[i] x_i[r] + lam[i] w_i[r])
__global__ void __launch_bounds__(kThreads) k_diag_outer(long n, const double *x, const double *w,
                                                         long ld, int k, const double *a,
                                                         const double *lam, double *gb) {
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads) {
        double s = 0.0;
        for (int i = 0; i < k; i++) {
            const double xr = x[(size_t)i * ld + r];
            s = fma(xr, fma(a[i], xr, lam[i] * w[(size_t)i * ld + r]), s);
        }
        gb[r] = s;
    }
}

// CG step lengths on the device: a[j] = rz[j] / pq[j] and na[j] = -a[j] for
// the columns still running (0 for the finished ones); a non-positive pq of a
// running column sets *bad = j + 1
__global__ void k_cg_alpha(int m, const double *rz, const double *pq, const int *done, double *a,
                           double *na, int *bad) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    double v = 0.0;
    if (!done[j]) {
        if (!(pq[j] > 0.0) || !isfinite(rz[j]))
            *bad = j + 1;
        else
            v = rz[j] / pq[j];
    }
    a[j] = v;
    na[j] = -v;
}

// after the residual update: a column finishes once ||r|| <= target; beta =
// rz_new / rz for the running ones (0 for the finished), rz <- rz_new
// done[m + 1] counts the iterations in which some column was still running
// (one workgroup: m <= kMaxCols <= blockDim.x)
__global__ void k_cg_beta(int m, const double *rr, const double *rzn, const double *target,
                          int *done, double *rz, double *beta) {
    const int j = threadIdx.x;
    const bool running = j < m && !done[j];
    if (__syncthreads_or(running) && j == 0) done[m + 1]++;
    if (j >= m) return;
    if (running && sqrt(fmax(rr[j], 0.0)) <= target[j]) done[j] = 1;
    beta[j] = done[j] ? 0.0 : rzn[j] / rz[j];
    rz[j] = rzn[j];
}

}  // namespace

void cg_alpha(int m, const double *rz, const double *pq, const int *done, double *a, double *na,
              int *bad) {
    if (m == 0) return;
    hipLaunchKernelGGL(k_cg_alpha, dim3((m + 63) / 64), dim3(64), 0, stream(), m, rz, pq, done, a,
                       na, bad);
    BX_HIP(hipGetLastError());
}

void cg_beta(int m, const double *rr, const double *rzn, const double *target, int *done,
             double *rz, double *beta) {
    if (m == 0) return;
    BX_REQUIRE(m <= kMaxCols, "cg_beta: %d columns", m);
    hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(kMaxCols), 0, stream(), m, rr, rzn, target, done,
                       rz, beta);
    BX_HIP(hipGetLastError());
}

void axpby_cols(int prec, long n, const double *a, const void *x, long ldx, const double *b,
                void *y, long ldy, const ColMap &cm) {
    if (cm.n == 0 || n == 0) return;
    dim3 g(rows_grid(n), cm.n);
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_axpby<float>, g, dim3(kThreads), 0, stream(), n, a, (const float *)x,
                           ldx, b, (float *)y, ldy, cm);
    else
        hipLaunchKernelGGL(k_axpby<double>, g, dim3(kThreads), 0, stream(), n, a,
                           (const double *)x, ldx, b, (double *)y, ldy, cm);
    BX_HIP(hipGetLastError());
}

void sampled_outer(const hip_CSRMatrix *A, const double *x, const double *w, long ld, int k,
                   const double *gl, double *gval) {
    if (A->n == 0) return;
    hipLaunchKernelGGL(k_sampled_outer, dim3(rows_grid(A->n)), dim3(kThreads), 0, stream(), A->n,
                       A->rowptr, A->colind, x, w, ld, k, gl, gval);
    BX_HIP(hipGetLastError());
}

void diag_outer(long n, const double *x, const double *w, long ld, int k, const double *a,
                const double *lam, double *gb) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_diag_outer, dim3(rows_grid(n)), dim3(kThreads), 0, stream(), n, x, w, ld,
                       k, a, lam, gb);
    BX_HIP(hipGetLastError());
}

}  // namespace bx
