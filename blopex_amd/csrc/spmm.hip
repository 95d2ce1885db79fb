// CSR SpMM Y(:, ym) = A X(:, xm) over a column-major block, and the diagonal
// operator Y = diag(d) X.  These replace the per-column operator loops of the
// reference backend (pcg_multi.c:167-193, matmultivec.c:11-48).
//
// One thread per row.  A row's (column, value) pairs are loaded once into
// registers and reused for every block column, so the matrix streams from
// HBM once per product; for the stencil matrices the X reads of neighbouring
// lanes hit consecutive addresses of the same column.  Blocks are renumbered
// so that each of the 8 XCDs walks one contiguous range of rows (workgroups
// are dealt to XCDs round-robin): the X rows a workgroup needs from the
// previous/next grid plane were then fetched into that XCD's L2 a short time
// before.
#include "kernels.h"

namespace bx {

namespace {

constexpr int kSpThreads = 256;
constexpr int kXcds = 8;

__device__ __forceinline__ long xcd_block(unsigned b, unsigned nb) {
    // nb is a multiple of kXcds
    return (long)(b % kXcds) * (nb / kXcds) + b / kXcds;
}

template <typename T, int MAXE>
__global__ void __launch_bounds__(kSpThreads) k_spmm(long n, const int *__restrict__ rowptr,
                                                     const int *__restrict__ colind,
                                                     const T *__restrict__ val, const T *x, long ldx,
                                                     ColMap xm, T *y, long ldy, ColMap ym) {
    const long r = xcd_block(blockIdx.x, gridDim.x) * kSpThreads + threadIdx.x;
    if (r >= n) return;
    const int e0 = rowptr[r], e1 = rowptr[r + 1];
    const int len = e1 - e0;
    if (len <= MAXE) {
        int ci[MAXE];
        T v[MAXE];
#pragma unroll
        for (int e = 0; e < MAXE; e++) {
            const bool ok = e < len;
            ci[e] = ok ? colind[e0 + e] : (int)r;
            v[e] = ok ? val[e0 + e] : T(0);
        }
        int j = 0;
        for (; j + 1 < xm.n; j += 2) {
            const T *xa = x + (size_t)xm.idx[j] * ldx;
            const T *xb = x + (size_t)xm.idx[j + 1] * ldx;
            T sa = T(0), sb = T(0);
#pragma unroll
            for (int e = 0; e < MAXE; e++) {
                if (e < len) {
                    sa = fma(v[e], xa[ci[e]], sa);
                    sb = fma(v[e], xb[ci[e]], sb);
                }
            }
            y[(size_t)ym.idx[j] * ldy + r] = sa;
            y[(size_t)ym.idx[j + 1] * ldy + r] = sb;
        }
        if (j < xm.n) {
            const T *xa = x + (size_t)xm.idx[j] * ldx;
            T sa = T(0);
#pragma unroll
            for (int e = 0; e < MAXE; e++)
                if (e < len) sa = fma(v[e], xa[ci[e]], sa);
            y[(size_t)ym.idx[j] * ldy + r] = sa;
        }
    } else {
        for (int j = 0; j < xm.n; j++) {
            const T *xa = x + (size_t)xm.idx[j] * ldx;
            T s = T(0);
            for (int e = e0; e < e1; e++) s = fma(val[e], xa[colind[e]], s);
            y[(size_t)ym.idx[j] * ldy + r] = s;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kSpThreads) k_diag(long n, const T *__restrict__ d, const T *x,
                                                     long ldx, ColMap xm, T *y, long ldy, ColMap ym) {
    const long r = xcd_block(blockIdx.x, gridDim.x) * kSpThreads + threadIdx.x;
    if (r >= n) return;
    const T dr = d[r];
    for (int j = 0; j < xm.n; j++)
        y[(size_t)ym.idx[j] * ldy + r] = dr * x[(size_t)xm.idx[j] * ldx + r];
}

unsigned xcd_grid(long n) {
    unsigned nb = (unsigned)((n + kSpThreads - 1) / kSpThreads);
    return (nb + kXcds - 1) / kXcds * kXcds;
}

}  // namespace

void spmm_csr(int prec, const hip_CSRMatrix *A, const void *x, long ldx, const ColMap &xm, void *y,
              long ldy, const ColMap &ym) {
    BX_REQUIRE(xm.n == ym.n, "spmm: active column counts differ (%d vs %d)", xm.n, ym.n);
    if (xm.n == 0 || A->n == 0) return;
    dim3 g(xcd_grid(A->n));
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL((k_spmm<float, 8>), g, dim3(kSpThreads), 0, stream(), A->n, A->rowptr,
                           A->colind, A->val32, (const float *)x, ldx, xm, (float *)y, ldy, ym);
    else
        hipLaunchKernelGGL((k_spmm<double, 8>), g, dim3(kSpThreads), 0, stream(), A->n, A->rowptr,
                           A->colind, A->val, (const double *)x, ldx, xm, (double *)y, ldy, ym);
    BX_HIP(hipGetLastError());
}

void diag_apply(int prec, const hip_DiagMatrix *D, const void *x, long ldx, const ColMap &xm,
                void *y, long ldy, const ColMap &ym) {
    BX_REQUIRE(xm.n == ym.n, "diag: active column counts differ (%d vs %d)", xm.n, ym.n);
    if (xm.n == 0 || D->n == 0) return;
    dim3 g(xcd_grid(D->n));
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_diag<float>, g, dim3(kSpThreads), 0, stream(), D->n, D->d32,
                           (const float *)x, ldx, xm, (float *)y, ldy, ym);
    else
        hipLaunchKernelGGL(k_diag<double>, g, dim3(kSpThreads), 0, stream(), D->n, D->d,
                           (const double *)x, ldx, xm, (double *)y, ldy, ym);
    BX_HIP(hipGetLastError());
}

}  // namespace bx
