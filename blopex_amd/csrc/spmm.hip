// CSR SpMM Y(:, ym) = A X(:, xm) over a column-major block, and the diagonal
// operator Y = diag(d) X.  These replace the per-column operator loops of the
// reference backend (pcg_multi.c:167-193, matmultivec.c:11-48).
//
// Matrices whose nonzeros lie on a few diagonals (stencils; see
// hip_CSRMatrix) use their diagonal copy: a thread owns 4 consecutive rows,
// keeps their diagonal values in registers for all block columns, and reads
// X with one 16-byte load per diagonal and column -- no index stream and no
// 4-byte gathers.  The diagonals are summed in ascending offset order, which
// is the column order the CSR kernel sums in, and absent entries are zeros,
// so both kernels give the same bits.
//
// General CSR: one thread per row.  A row's (column, value) pairs are loaded once into
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
        // short rows padded with (r, 0): fma(0, x, s) == s, so the padding
        // changes no bit and the inner loop has no branch
        int ci[MAXE];
        T v[MAXE];
#pragma unroll
        for (int e = 0; e < MAXE; e++) {
            const bool ok = e < len;
            ci[e] = ok ? colind[e0 + e] : (int)r;
            v[e] = ok ? val[e0 + e] : T(0);
        }
        for (int j = 0; j < xm.n; j++) {
            const T *xa = x + (size_t)xm.idx[j] * ldx;
            T sa = T(0);
#pragma unroll
            for (int e = 0; e < MAXE; e++) sa = fma(v[e], xa[ci[e]], sa);
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

struct DiaOffsets {
    int o[HIP_CSR_MAX_DIAGONALS];
};

// 4 consecutive values at p (4-byte aligned for float, 8-byte for double)
__device__ __forceinline__ void ld4(const float *p, float v[4]) {
    float4 q;
    __builtin_memcpy(&q, p, sizeof(q));
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
}
__device__ __forceinline__ void ld4(const double *p, double v[4]) {
    double2 a, b;
    __builtin_memcpy(&a, p, sizeof(a));
    __builtin_memcpy(&b, p + 2, sizeof(b));
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
}

template <typename T, int ND>
__global__ void __launch_bounds__(kSpThreads) k_dia(long n, DiaOffsets off,
                                                    const T *__restrict__ dval, long dld,
                                                    const T *x, long ldx, ColMap xm, T *y, long ldy,
                                                    ColMap ym) {
    const long r0 = 4 * (xcd_block(blockIdx.x, gridDim.x) * kSpThreads + threadIdx.x);
    if (r0 >= n) return;
    T dv[ND][4];
#pragma unroll
    for (int d = 0; d < ND; d++) ld4(dval + (size_t)d * dld + r0, dv[d]);
    // rows whose every neighbour block lies inside [0, n) take 16-byte loads
    bool inner = r0 + 4 <= n;
#pragma unroll
    for (int d = 0; d < ND; d++) inner = inner && r0 + off.o[d] >= 0 && r0 + off.o[d] + 4 <= n;
    for (int j = 0; j < xm.n; j++) {
        const T *xa = x + (size_t)xm.idx[j] * ldx;
        T s[4] = {T(0), T(0), T(0), T(0)};
        if (inner) {
#pragma unroll
            for (int d = 0; d < ND; d++) {
                T xv[4];
                ld4(xa + r0 + off.o[d], xv);
#pragma unroll
                for (int i = 0; i < 4; i++) s[i] = fma(dv[d][i], xv[i], s[i]);
            }
        } else {
#pragma unroll
            for (int d = 0; d < ND; d++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    long c = r0 + i + off.o[d];
                    c = c < 0 ? 0 : (c >= n ? n - 1 : c);  // the value is 0 there
                    s[i] = fma(dv[d][i], xa[c], s[i]);
                }
        }
        T *ya = y + (size_t)ym.idx[j] * ldy + r0;
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (r0 + i < n) ya[i] = s[i];
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
    if (A->ndiag > 0) {
        DiaOffsets off;
        for (int d = 0; d < HIP_CSR_MAX_DIAGONALS; d++) off.o[d] = d < A->ndiag ? A->diag_offset[d] : 0;
        dim3 gd(xcd_grid((A->n + 3) / 4));
        const void *dv = prec == HIPMV_FP32 ? (const void *)A->diag_val32 : (const void *)A->diag_val;
#define BX_DIA(ND)                                                                              \
    case ND:                                                                                    \
        if (prec == HIPMV_FP32)                                                                 \
            hipLaunchKernelGGL((k_dia<float, ND>), gd, dim3(kSpThreads), 0, stream(), A->n, off, \
                               (const float *)dv, A->diag_ld, (const float *)x, ldx, xm,        \
                               (float *)y, ldy, ym);                                             \
        else                                                                                    \
            hipLaunchKernelGGL((k_dia<double, ND>), gd, dim3(kSpThreads), 0, stream(), A->n,    \
                               off, (const double *)dv, A->diag_ld, (const double *)x, ldx, xm, \
                               (double *)y, ldy, ym);                                            \
        break;
        switch (A->ndiag) {
            BX_DIA(1) BX_DIA(2) BX_DIA(3) BX_DIA(4) BX_DIA(5) BX_DIA(6) BX_DIA(7) BX_DIA(8)
            BX_DIA(9) BX_DIA(10) BX_DIA(11) BX_DIA(12) BX_DIA(13) BX_DIA(14) BX_DIA(15) BX_DIA(16)
            default: BX_REQUIRE(false, "spmm: %d diagonals", A->ndiag);
        }
#undef BX_DIA
        BX_HIP(hipGetLastError());
        return;
    }
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
