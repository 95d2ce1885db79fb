// Complex (komplex: interleaved FP64 real, imaginary) block kernels for the
// hip_Multi_Vector slots that lobpcg_solve_complex drives: the Gram product
// X^H Y and its diagonal, block times small complex matrix, column scaling by
// complex diagonals, and the reference's MatMultiVec operators: a dense
// operator on FP64 MFMA for complex and real blocks
// (blopex_serial_complex/multivector/multi_vector.c:567-610,
// blopex_serial_double/pcg_multi/pcg_multi.c:167-193) and the 1D Laplacian it
// applies when given no matrix (blopex_serial_double/matmultivec/
// matmultivec.c:11-48).  Copies, clears, real axpy and the drand48 fill treat a
// complex column as 2n doubles and use the FP64 kernels.
// Reductions are two-stage with a fixed summation order, as in blas.hip.
#include "kernels.h"
#include "mfma.h"

namespace bx {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, a.y * b.y), fma(a.x, b.y, -a.y * b.x));
}

// Realified columns: complex z (n rows) <-> real r (2n rows: the real parts,
// then the imaginary parts).  With jpair, column w + j of r also receives
// J z_j = i z_j, that is [-Im z_j; Re z_j].
template <typename T>
__global__ void __launch_bounds__(kThreads) k_z_to_real(long n, const double2 *z, long ldz, T *r,
                                                        long ldr, int w, int jpair) {
    const int j = blockIdx.y;
    const double2 *zc = z + (size_t)j * ldz;
    T *rc = r + (size_t)j * ldr, *rp = r + (size_t)(j + w) * ldr;
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads) {
        const double2 v = zc[i];
        rc[i] = (T)v.x;
        rc[n + i] = (T)v.y;
        if (jpair) {
            rp[i] = (T)-v.y;
            rp[n + i] = (T)v.x;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_real_to_z(long n, const T *r, long ldr, double2 *z,
                                                        long ldz) {
    const int j = blockIdx.y;
    const T *rc = r + (size_t)j * ldr;
    double2 *zc = z + (size_t)j * ldz;
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads)
        zc[i] = make_double2((double)rc[i], (double)rc[n + i]);
}

unsigned zrow_grid(long n) {
    size_t blocks = ((size_t)n + kThreads - 1) / kThreads;
    size_t cap = (size_t)num_cus() * 4;
    return (unsigned)(blocks < 1 ? 1 : (blocks > cap ? cap : blocks));
}

// part[(pair) * nparts + blockIdx.x] = sum over this block's rows of conj(x_i) y_j,
// pair = blockIdx.y: (i, j) = (pair % h, pair / h), or (pair, pair) when diag
__global__ void __launch_bounds__(kThreads) k_zgram(long n, const double2 *x, long ldx, ColMap xm,
                                                    const double2 *y, long ldy, ColMap ym, int h,
                                                    int diag, double2 *part) {
    __shared__ double2 red[kThreads];
    const int pr = blockIdx.y;
    const int i = diag ? pr : pr % h, j = diag ? pr : pr / h;
    const double2 *xc = x + (size_t)xm.idx[i] * ldx;
    const double2 *yc = y + (size_t)ym.idx[j] * ldy;
    double2 s = make_double2(0.0, 0.0);
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads) {
        const double2 p = cmulc(xc[r], yc[r]);
        s.x += p.x;
        s.y += p.y;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[threadIdx.x].x += red[threadIdx.x + w].x;
            red[threadIdx.x].y += red[threadIdx.x + w].y;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)pr * gridDim.x + blockIdx.x] = red[0];
}

// g[(pair % h) + (pair / h) * ldg] (or g[pair] when diag) = sum of the pair's partials
__global__ void k_zgram_sum(const double2 *part, int nparts, int npairs, int h, int diag, long ldg,
                            double2 *g) {
    const int pr = blockIdx.x * blockDim.x + threadIdx.x;
    if (pr >= npairs) return;
    double2 s = make_double2(0.0, 0.0);
    for (int p = 0; p < nparts; p++) {
        s.x += part[(size_t)pr * nparts + p].x;
        s.y += part[(size_t)pr * nparts + p].y;
    }
    if (diag)
        g[pr] = s;
    else
        g[(pr % h) + (size_t)(pr / h) * ldg] = s;
}

// y(:, ym[j]) = sum_i c[i + j ldc] x(:, xm[i]) (+ y when accumulate)
__global__ void __launch_bounds__(kThreads) k_zvecmat(long n, const double2 *x, long ldx,
                                                      ColMap xm, const double2 *c, long ldc,
                                                      double2 *y, long ldy, ColMap ym,
                                                      int accumulate) {
    const int j = blockIdx.y;
    double2 *yc = y + (size_t)ym.idx[j] * ldy;
    const double2 *cj = c + (size_t)j * ldc;
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads) {
        double2 s = accumulate ? yc[r] : make_double2(0.0, 0.0);
        for (int i = 0; i < xm.n; i++) {
            const double2 p = cmul(cj[i], x[(size_t)xm.idx[i] * ldx + r]);
            s.x += p.x;
            s.y += p.y;
        }
        yc[r] = s;
    }
}

// y(:, ym[j]) = d[j] x(:, xm[j])
__global__ void __launch_bounds__(kThreads) k_zscale(long n, const double2 *d, const double2 *x,
                                                     long ldx, ColMap xm, double2 *y, long ldy,
                                                     ColMap ym) {
    const int j = blockIdx.y;
    const double2 a = d[j];
    const double2 *xc = x + (size_t)xm.idx[j] * ldx;
    double2 *yc = y + (size_t)ym.idx[j] * ldy;
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads)
        yc[r] = cmul(a, xc[r]);
}

// y(:, ym[j]) = A x(:, xm[j]), A dense n x n (leading dim lda)
// Y(:, ym) = A X(:, xm), A dense n x n, on FP64 MFMA 16x16x4: a wave owns 16
// rows of Y and a slice of the inner index, with Y_r = A_r X_r - A_i X_i and
// Y_i = A_r X_i + A_i X_r as four real products per step of 4, and CT tiles of
// 16 X columns (columns c0 .. c0 + w of the maps).  Lane (li, lk) supplies A's
// row r0 + li and X's column 16 c + li at inner index k + lk; entries past n
// or w are read from a clamped place and zeroed.  One slice writes Y itself;
// several write partials part[s][col][row] that k_zdense_sum adds in order.
template <int CT>
__global__ void __launch_bounds__(64) k_zdense_mfma(long n, const double2 *__restrict__ A,
                                                    long lda, const double2 *__restrict__ x,
                                                    long ldx, ColMap xm, int c0, int w,
                                                    long kslice, double2 *__restrict__ y,
                                                    long ldy, ColMap ym, double2 *part) {
    const int lane = threadIdx.x, li = lane & 15, lk = lane >> 4;
    const long r0 = (long)blockIdx.x * 16, row = r0 + li;
    const long kb = (long)blockIdx.y * kslice, ke = kb + kslice < n ? kb + kslice : n;
    const bool rok = row < n;
    const double2 *ar = A + (rok ? row : n - 1);
    const double2 *xc[CT];
    bool cok[CT];
#pragma unroll
    for (int c = 0; c < CT; c++) {
        const int col = 16 * c + li;
        cok[c] = col < w;
        xc[c] = x + (size_t)xm.idx[c0 + (cok[c] ? col : w - 1)] * ldx;
    }
    f64x4 yr[CT], yi[CT];
#pragma unroll
    for (int c = 0; c < CT; c++) yr[c] = yi[c] = f64x4{0.0, 0.0, 0.0, 0.0};
    auto load = [&](long k, double2 &a, double2 (&b)[CT]) {
        const long kk = k + lk;
        const bool kok = kk < ke;
        const long kc = kok ? kk : ke - 1;
        const double2 av = ar[(size_t)kc * lda];
        a = (kok && rok) ? av : make_double2(0.0, 0.0);
#pragma unroll
        for (int c = 0; c < CT; c++) {
            const double2 bv = xc[c][kc];
            b[c] = (kok && cok[c]) ? bv : make_double2(0.0, 0.0);
        }
    };
    // kD steps of loads in flight: a ring refilled as each step is used
    // (steps past the slice read clamped places and come out zero)
    constexpr int kD = 4;
    double2 a[kD], b[kD][CT];
#pragma unroll
    for (int d = 0; d < kD; d++) load(kb + 4 * d, a[d], b[d]);
    for (long k = kb; k < ke; k += 4 * kD) {
#pragma unroll
        for (int d = 0; d < kD; d++) {
#pragma unroll
            for (int c = 0; c < CT; c++) {
                yr[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[d].x, b[d][c].x, yr[c], 0, 0, 0);
                yi[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[d].x, b[d][c].y, yi[c], 0, 0, 0);
            }
#pragma unroll
            for (int c = 0; c < CT; c++) {
                yr[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[d].y, b[d][c].y, yr[c], 0, 0, 0);
                yi[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[d].y, b[d][c].x, yi[c], 0, 0, 0);
            }
            load(k + 4 * (d + kD), a[d], b[d]);
        }
    }
    // register q of lane (li, lk): row r0 + lk + 4 q, column 16 c + li
#pragma unroll
    for (int c = 0; c < CT; c++) {
        const int col = 16 * c + li;
        if (col >= w) continue;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const long r = r0 + lk + 4 * q;
            if (r >= n) continue;
            const double2 v = make_double2(yr[c][q], yi[c][q]);
            if (gridDim.y == 1)
                y[(size_t)ym.idx[c0 + col] * ldy + r] = v;
            else
                part[((size_t)blockIdx.y * w + col) * n + r] = v;
        }
    }
}

// y(:, ym[c0 + col]) = sum over the slices of part[s][col][:], in slice order
__global__ void __launch_bounds__(kThreads) k_zdense_sum(long n, int w, int nslice,
                                                         const double2 *part, double2 *y,
                                                         long ldy, ColMap ym, int c0) {
    const int col = blockIdx.y;
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads) {
        double2 s = make_double2(0.0, 0.0);
        for (int q = 0; q < nslice; q++) {
            const double2 v = part[((size_t)q * w + col) * n + r];
            s.x += v.x;
            s.y += v.y;
        }
        y[(size_t)ym.idx[c0 + col] * ldy + r] = s;
    }
}

// The real form of the same product, Y(:, ym) = A X(:, xm) with A dense FP64
// and X, Y in FP64 or FP32 (FP64 arithmetic).  A wave owns RT tiles of 16
// rows (rows r0 + 16 t + li), so each X operand it loads feeds RT products;
// per step of 4 it issues RT x CT 16x16x4 products, with KD steps of loads in
// flight.  Loads read clamped places (rows past n are never stored); the main
// loop runs whole rounds of KD steps inside the slice, and a last partial
// round zeroes the steps past it.  Slices and the partial layout are those of
// the complex kernel.
template <typename T, int CT, int RT, int KD, int WV>
__global__ void __launch_bounds__(64 * WV) k_dense_mfma(long n, const double *__restrict__ A,
                                                        long lda, const T *__restrict__ x, long ldx,
                                                        ColMap xm, int c0, int w, long kslice,
                                                        T *__restrict__ y, long ldy, ColMap ym,
                                                        double *part) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const long r0 = ((long)blockIdx.x * WV + (threadIdx.x >> 6)) * 16 * RT;
    const long kb = (long)blockIdx.y * kslice, ke = kb + kslice < n ? kb + kslice : n;
    const double *ar[RT];
#pragma unroll
    for (int t = 0; t < RT; t++) {
        const long row = r0 + 16 * t + li;
        ar[t] = A + (row < n ? row : n - 1);
    }
    const T *xc[CT];
#pragma unroll
    for (int c = 0; c < CT; c++) {
        const int col = 16 * c + li;
        xc[c] = x + (size_t)xm.idx[c0 + (col < w ? col : w - 1)] * ldx;
    }
    f64x4 acc[RT][CT];
#pragma unroll
    for (int t = 0; t < RT; t++)
#pragma unroll
        for (int c = 0; c < CT; c++) acc[t][c] = f64x4{0.0, 0.0, 0.0, 0.0};
    auto load = [&](long k, double (&a)[RT], double (&b)[CT]) {
        const long kc = k + lk < n ? k + lk : n - 1;
#pragma unroll
        for (int t = 0; t < RT; t++) a[t] = ar[t][(size_t)kc * lda];
#pragma unroll
        for (int c = 0; c < CT; c++) b[c] = (double)xc[c][kc];
    };
    auto mma = [&](const double (&a)[RT], const double (&b)[CT]) {
#pragma unroll
        for (int t = 0; t < RT; t++)
#pragma unroll
            for (int c = 0; c < CT; c++)
                acc[t][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[c], acc[t][c], 0, 0, 0);
    };
    double a[KD][RT], b[KD][CT];
#pragma unroll
    for (int d = 0; d < KD; d++) load(kb + 4 * d, a[d], b[d]);
    long k = kb;
    for (; k + 4 * KD <= ke; k += 4 * KD) {
#pragma unroll
        for (int d = 0; d < KD; d++) {
            mma(a[d], b[d]);
            load(k + 4 * (d + KD), a[d], b[d]);
        }
    }
    if (k < ke) {
#pragma unroll
        for (int d = 0; d < KD; d++) {
            const bool kok = k + 4 * d + lk < ke;
#pragma unroll
            for (int t = 0; t < RT; t++) a[d][t] = kok ? a[d][t] : 0.0;
#pragma unroll
            for (int c = 0; c < CT; c++) b[d][c] = kok ? b[d][c] : 0.0;
            mma(a[d], b[d]);
        }
    }
#pragma unroll
    for (int c = 0; c < CT; c++) {
        const int col = 16 * c + li;
        if (col >= w) continue;
#pragma unroll
        for (int t = 0; t < RT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const long r = r0 + 16 * t + lk + 4 * q;
                if (r >= n) continue;
                if (gridDim.y == 1)
                    y[(size_t)ym.idx[c0 + col] * ldy + r] = (T)acc[t][c][q];
                else
                    part[((size_t)blockIdx.y * w + col) * n + r] = acc[t][c][q];
            }
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_dense_sum(long n, int w, int nslice,
                                                        const double *part, T *y, long ldy,
                                                        ColMap ym, int c0) {
    const int col = blockIdx.y;
    const size_t qs = (size_t)w * n;
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads) {
        // four running sums over the slices q = j (mod 4), then added in a fixed order
        const double *p = part + (size_t)col * n + r;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        int q = 0;
        for (; q + 4 <= nslice; q += 4)
#pragma unroll
            for (int j = 0; j < 4; j++) s[j] += p[(q + j) * qs];
        for (; q < nslice; q++) s[q & 3] += p[q * qs];
        y[(size_t)ym.idx[c0 + col] * ldy + r] = (T)((s[0] + s[1]) + (s[2] + s[3]));
    }
}

// 1D Laplacian with Dirichlet ends, y_j = -x_{j-1} + 2 x_j - x_{j+1} on each
// active column (the stencil of matmultivec.c:36-45), FP64 arithmetic
template <typename T>
__global__ void __launch_bounds__(kThreads) k_lapl1d(long n, const T *x, long ldx, ColMap xm,
                                                     T *y, long ldy, ColMap ym) {
    const T *xc = x + (size_t)xm.idx[blockIdx.y] * ldx;
    T *yc = y + (size_t)ym.idx[blockIdx.y] * ldy;
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads) {
        double v = 2.0 * (double)xc[i];
        if (i > 0) v -= (double)xc[i - 1];
        if (i + 1 < n) v -= (double)xc[i + 1];
        yc[i] = (T)v;
    }
}

// Shape of the real dense product: 4 row tiles of 16 per wave, 4 steps of
// loads in flight, 4 waves per workgroup (256 rows), one wave per SIMD.
// Measured at n = 16000, 32 columns: 532 us against 366 us with the loads
// cut out and 412 us with the products cut out; 2 tiles per wave, 8 steps in
// flight, 8 waves per workgroup or 2 waves per SIMD were no faster.
constexpr int kDenseRT = 4, kDenseKD = 4, kDenseWV = 4, kDenseWaves = 1024;

// slices of the inner index (lengths a multiple of kmult) so that about
// `units` workgroups run, each owning tile_rows rows
void dense_slices(long n, long tile_rows, long units, long kmult, long &tiles, long &ns,
                  long &kslice) {
    tiles = (n + tile_rows - 1) / tile_rows;
    ns = (units + tiles - 1) / tiles;
    if (ns > (n + 3) / 4) ns = (n + 3) / 4;
    kslice = ((n + ns - 1) / ns + kmult - 1) / kmult * kmult;
    ns = (n + kslice - 1) / kslice;
    BX_REQUIRE(tiles <= 0x7fffffffL && ns <= 65535, "dense operator: %ld rows", n);
}

template <typename T>
void dense_apply_t(long n, const double *A, long lda, const T *x, long ldx, const ColMap &xm,
                   T *y, long ldy, const ColMap &ym) {
    constexpr int RT = kDenseRT, WV = kDenseWV;
    long tiles, ns, kslice;
    // a whole number of workgroups per CU where it can (kDenseWaves waves on
    // 256 CUs), so that the SIMDs hold equal numbers of waves, and slices of
    // at least 128 inner indices (32 steps per wave)
    long wgs = (long)kDenseWaves * num_cus() / (256L * WV);
    tiles = (n + 16 * RT * WV - 1) / (16 * RT * WV);
    long per = wgs / tiles;
    if (per > n / 128) per = n / 128;
    if (per < 1) per = 1;
    dense_slices(n, 16 * RT * WV, per * tiles, 4 * kDenseKD, tiles, ns, kslice);
    for (int c0 = 0; c0 < xm.n; c0 += 64) {
        const int w = xm.n - c0 < 64 ? xm.n - c0 : 64;
        double *part = ns > 1 ? (double *)scratch(15, (size_t)ns * w * n * sizeof(double)) : nullptr;
        const dim3 g((unsigned)tiles, (unsigned)ns);
        const int ct = (w + 15) / 16;
#define BX_DD(CT)                                                                             \
    hipLaunchKernelGGL((k_dense_mfma<T, CT, RT, kDenseKD, WV>), g, dim3(64 * WV), 0, stream(), n,    \
                       A, lda, x, ldx, xm, c0, w, kslice, y, ldy, ym, part)
        if (ct == 1) BX_DD(1);
        else if (ct == 2) BX_DD(2);
        else if (ct == 3) BX_DD(3);
        else BX_DD(4);
#undef BX_DD
        BX_HIP(hipGetLastError());
        if (ns > 1) {
            hipLaunchKernelGGL(k_dense_sum<T>, dim3(zrow_grid(n), w), dim3(kThreads), 0, stream(), n,
                               w, (int)ns, part, y, ldy, ym, c0);
            BX_HIP(hipGetLastError());
        }
    }
}

void zgram_launch(long n, const void *x, long ldx, const ColMap &xm, const void *y, long ldy,
                  const ColMap &ym, int diag, void *g, long ldg) {
    const int h = xm.n, npairs = diag ? xm.n : xm.n * ym.n;
    if (npairs == 0) return;
    unsigned nparts = zrow_grid(n);
    // at most ~64K workgroups in all
    while (nparts > 1 && (size_t)nparts * npairs > 65536) nparts = (nparts + 1) / 2;
    double2 *part = (double2 *)scratch(0, (size_t)nparts * npairs * sizeof(double2));
    hipLaunchKernelGGL(k_zgram, dim3(nparts, npairs), dim3(kThreads), 0, stream(), n,
                       (const double2 *)x, ldx, xm, (const double2 *)y, ldy, ym, h, diag, part);
    BX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_zgram_sum, dim3((npairs + 255) / 256), dim3(256), 0, stream(), part,
                       (int)nparts, npairs, h, diag, ldg, (double2 *)g);
    BX_HIP(hipGetLastError());
}

}  // namespace

void zgram_cols(long n, const void *x, long ldx, const ColMap &xm, const void *y, long ldy,
                const ColMap &ym, void *g, long ldg) {
    zgram_launch(n, x, ldx, xm, y, ldy, ym, 0, g, ldg);
}

void zdiagdot_cols(long n, const void *x, long ldx, const ColMap &xm, const void *y, long ldy,
                   const ColMap &ym, void *d) {
    BX_REQUIRE(xm.n == ym.n, "zdiagdot: active column counts differ (%d vs %d)", xm.n, ym.n);
    zgram_launch(n, x, ldx, xm, y, ldy, ym, 1, d, 1);
}

void zvecmat_cols(long n, const void *x, long ldx, const ColMap &xm, const void *c, long ldc,
                  void *y, long ldy, const ColMap &ym, bool accumulate) {
    if (ym.n == 0 || n == 0) return;
    hipLaunchKernelGGL(k_zvecmat, dim3(zrow_grid(n), ym.n), dim3(kThreads), 0, stream(), n,
                       (const double2 *)x, ldx, xm, (const double2 *)c, ldc, (double2 *)y, ldy, ym,
                       (int)accumulate);
    BX_HIP(hipGetLastError());
}

void zscale_cols(long n, const void *d, const void *x, long ldx, const ColMap &xm, void *y,
                 long ldy, const ColMap &ym) {
    BX_REQUIRE(xm.n == ym.n, "zscale: active column counts differ (%d vs %d)", xm.n, ym.n);
    if (xm.n == 0 || n == 0) return;
    hipLaunchKernelGGL(k_zscale, dim3(zrow_grid(n), xm.n), dim3(kThreads), 0, stream(), n,
                       (const double2 *)d, (const double2 *)x, ldx, xm, (double2 *)y, ldy, ym);
    BX_HIP(hipGetLastError());
}

void zdense_apply(long n, const void *A, long lda, const void *x, long ldx, const ColMap &xm,
                  void *y, long ldy, const ColMap &ym) {
    BX_REQUIRE(xm.n == ym.n, "dense operator: active column counts differ (%d vs %d)", xm.n, ym.n);
    if (xm.n == 0 || n == 0) return;
    long tiles, ns, kslice;
    dense_slices(n, 16, 2048, 4, tiles, ns, kslice);
    for (int c0 = 0; c0 < xm.n; c0 += 64) {
        const int w = xm.n - c0 < 64 ? xm.n - c0 : 64;
        double2 *part = ns > 1 ? (double2 *)scratch(15, (size_t)ns * w * n * sizeof(double2))
                               : nullptr;
        const dim3 g((unsigned)tiles, (unsigned)ns);
        const int ct = (w + 15) / 16;
#define BX_ZD(CT)                                                                            \
    hipLaunchKernelGGL(k_zdense_mfma<CT>, g, dim3(64), 0, stream(), n, (const double2 *)A, lda, \
                       (const double2 *)x, ldx, xm, c0, w, kslice, (double2 *)y, ldy, ym, part)
        if (ct == 1) BX_ZD(1);
        else if (ct == 2) BX_ZD(2);
        else if (ct == 3) BX_ZD(3);
        else BX_ZD(4);
#undef BX_ZD
        BX_HIP(hipGetLastError());
        if (ns > 1) {
            hipLaunchKernelGGL(k_zdense_sum, dim3(zrow_grid(n), w), dim3(kThreads), 0, stream(), n,
                               w, (int)ns, part, (double2 *)y, ldy, ym, c0);
            BX_HIP(hipGetLastError());
        }
    }
}

void dense_apply(int prec, long n, const double *A, long lda, const void *x, long ldx,
                 const ColMap &xm, void *y, long ldy, const ColMap &ym) {
    BX_REQUIRE(xm.n == ym.n, "dense operator: active column counts differ (%d vs %d)", xm.n, ym.n);
    if (xm.n == 0 || n == 0) return;
    if (prec == HIPMV_FP64)
        dense_apply_t<double>(n, A, lda, (const double *)x, ldx, xm, (double *)y, ldy, ym);
    else
        dense_apply_t<float>(n, A, lda, (const float *)x, ldx, xm, (float *)y, ldy, ym);
}

void lapl1d_apply(int prec, long n, const void *x, long ldx, const ColMap &xm, void *y, long ldy,
                  const ColMap &ym) {
    BX_REQUIRE(xm.n == ym.n, "1D Laplacian: active column counts differ (%d vs %d)", xm.n, ym.n);
    BX_REQUIRE(n > 1, "1D Laplacian: %ld rows (at least 2 needed)", n);
    if (xm.n == 0) return;
    const dim3 g(zrow_grid(n), xm.n);
    if (prec == HIPMV_FP64)
        hipLaunchKernelGGL(k_lapl1d<double>, g, dim3(kThreads), 0, stream(), n, (const double *)x,
                           ldx, xm, (double *)y, ldy, ym);
    else
        hipLaunchKernelGGL(k_lapl1d<float>, g, dim3(kThreads), 0, stream(), n, (const float *)x,
                           ldx, xm, (float *)y, ldy, ym);
    BX_HIP(hipGetLastError());
}

void z_to_real(long n, int w, const void *z, long ldz, void *r, long ldr, int rprec, bool jpair) {
    if (n == 0 || w == 0) return;
    BX_REQUIRE(rprec == HIPMV_FP64 || rprec == HIPMV_FP32, "z_to_real: precision %d", rprec);
    dim3 g(zrow_grid(n), w);
    if (rprec == HIPMV_FP64)
        hipLaunchKernelGGL(k_z_to_real<double>, g, dim3(kThreads), 0, stream(), n,
                           (const double2 *)z, ldz, (double *)r, ldr, w, (int)jpair);
    else
        hipLaunchKernelGGL(k_z_to_real<float>, g, dim3(kThreads), 0, stream(), n,
                           (const double2 *)z, ldz, (float *)r, ldr, w, (int)jpair);
    BX_HIP(hipGetLastError());
}

void real_to_z(long n, int w, const void *r, long ldr, int rprec, void *z, long ldz) {
    if (n == 0 || w == 0) return;
    BX_REQUIRE(rprec == HIPMV_FP64 || rprec == HIPMV_FP32, "real_to_z: precision %d", rprec);
    dim3 g(zrow_grid(n), w);
    if (rprec == HIPMV_FP64)
        hipLaunchKernelGGL(k_real_to_z<double>, g, dim3(kThreads), 0, stream(), n,
                           (const double *)r, ldr, (double2 *)z, ldz);
    else
        hipLaunchKernelGGL(k_real_to_z<float>, g, dim3(kThreads), 0, stream(), n,
                           (const float *)r, ldr, (double2 *)z, ldz);
    BX_HIP(hipGetLastError());
}

}  // namespace bx
