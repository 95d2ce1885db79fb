// Column-mapped block kernels: the bodies of the hip_Multi_Vector
// interpreter slots (copy, axpy, diagonal scaling, clear, random fill, Gram,
// diagonal of the Gram, block times small matrix) and precision conversion.
// Reductions are two-stage (per-block FP64 partials, then a fixed-order sum),
// so every result is reproducible run to run.
#include "dpp.h"
#include "kernels.h"

namespace bx {

namespace {

constexpr int kThreads = 256;

template <typename T>
__device__ __forceinline__ const T *colp(const T *base, long ld, int c) {
    return base + (size_t)c * (size_t)ld;
}
template <typename T>
__device__ __forceinline__ T *colp(T *base, long ld, int c) {
    return base + (size_t)c * (size_t)ld;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_copy(long n, const T *x, long ldx, ColMap xm, T *y,
                                                   long ldy, ColMap ym) {
    const int j = blockIdx.y;
    const T *xc = colp(x, ldx, xm.idx[j]);
    T *yc = colp(y, ldy, ym.idx[j]);
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads)
        yc[r] = xc[r];
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_axpy(long n, double alpha, const T *x, long ldx,
                                                   ColMap xm, T *y, long ldy, ColMap ym) {
    const int j = blockIdx.y;
    const T *xc = colp(x, ldx, xm.idx[j]);
    T *yc = colp(y, ldy, ym.idx[j]);
    const T a = (T)alpha;
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads)
        yc[r] = yc[r] + a * xc[r];
}

struct Coef {
    double v[kMaxCols];
};

template <typename T>
__global__ void __launch_bounds__(kThreads) k_scale(long n, Coef coef, const T *x, long ldx,
                                                    ColMap xm, T *y, long ldy, ColMap ym) {
    const int j = blockIdx.y;
    const T *xc = colp(x, ldx, xm.idx[j]);
    T *yc = colp(y, ldy, ym.idx[j]);
    const T a = (T)coef.v[j];
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads)
        yc[r] = a * xc[r];
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_clear(long n, T *x, long ldx, ColMap xm) {
    T *xc = colp(x, ldx, xm.idx[blockIdx.y]);
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads)
        xc[r] = T(0);
}

// drand48: x' = (a x + c) mod 2^48.  A thread jumps to its first stream
// position with the binary powers of the affine map, then steps.
constexpr uint64_t kM48 = (1ull << 48) - 1;
constexpr uint64_t kA48 = 0x5DEECE66Dull;
constexpr uint64_t kC48 = 0xBull;
constexpr int kRandRun = 16;

__device__ __forceinline__ uint64_t lcg_jump(uint64_t x, uint64_t steps) {
    uint64_t a = kA48, c = kC48;  // map for 2^i steps
    while (steps) {
        if (steps & 1) x = (a * x + c) & kM48;
        c = (a * c + c) & kM48;
        a = (a * a) & kM48;
        steps >>= 1;
    }
    return x;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_drand48(long n, uint64_t x0, T *x, long ldx,
                                                      ColMap xm) {
    const int j = blockIdx.y;
    T *xc = colp(x, ldx, xm.idx[j]);
    for (long r0 = ((long)blockIdx.x * kThreads + threadIdx.x) * kRandRun; r0 < n;
         r0 += (long)gridDim.x * kThreads * kRandRun) {
        uint64_t s = lcg_jump(x0, (uint64_t)j * (uint64_t)n + (uint64_t)r0);
        const long r1 = r0 + kRandRun < n ? r0 + kRandRun : n;
        for (long r = r0; r < r1; r++) {
            s = (kA48 * s + kC48) & kM48;
            const double u = (double)s * (1.0 / 281474976710656.0);
            xc[r] = (T)(2.0 * u - 1.0);
        }
    }
}

template <typename S, typename D>
__global__ void __launch_bounds__(kThreads) k_convert(long n, const S *x, long ldx, D *y, long ldy) {
    const int j = blockIdx.y;
    const S *xc = x + (size_t)j * ldx;
    D *yc = y + (size_t)j * ldy;
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads)
        yc[r] = (D)xc[r];
}

// ---- Gram of mapped columns: 16 x 16 output tile per block, LDS-staged ----
constexpr int kGT = 16;     // tile edge
constexpr int kGRows = 64;  // rows staged per step

template <typename T>
__global__ void __launch_bounds__(kThreads) k_gram(long n, const T *x, long ldx, ColMap xm,
                                                   const T *y, long ldy, ColMap ym, int ntx,
                                                   double *part) {
    __shared__ double xs[kGRows][kGT + 1];
    __shared__ double ys[kGRows][kGT + 1];
    const int t = blockIdx.y;
    const int i0 = (t % ntx) * kGT, j0 = (t / ntx) * kGT;
    const int ti = threadIdx.x % kGT, tj = threadIdx.x / kGT;
    double acc = 0.0;
    for (long r0 = (long)blockIdx.x * kGRows; r0 < n; r0 += (long)gridDim.x * kGRows) {
        for (int q = 0; q < (kGRows * kGT) / kThreads; q++) {
            const int e = threadIdx.x + q * kThreads;
            const int rr = e % kGRows, cc = e / kGRows;
            const long r = r0 + rr;
            double xv = 0.0, yv = 0.0;
            if (r < n && i0 + cc < xm.n) xv = (double)colp(x, ldx, xm.idx[i0 + cc])[r];
            if (r < n && j0 + cc < ym.n) yv = (double)colp(y, ldy, ym.idx[j0 + cc])[r];
            xs[rr][cc] = xv;
            ys[rr][cc] = yv;
        }
        __syncthreads();
#pragma unroll 8
        for (int rr = 0; rr < kGRows; rr++) acc = fma(xs[rr][ti], ys[rr][tj], acc);
        __syncthreads();
    }
    part[((size_t)blockIdx.x * gridDim.y + t) * (kGT * kGT) + threadIdx.x] = acc;
}

// Sums the nparts partial tiles: kRE consecutive entries per workgroup, one
// per lane (coalesced loads), each entry's partials split over
// kThreads / kRE lanes and the slices summed in LDS.  (One thread per entry
// would sum a thousand partials serially when the output is a single tile.)
constexpr int kRE = 16;
static_assert(kThreads % kRE == 0, "reduction slices");

__global__ void __launch_bounds__(kThreads) k_gram_reduce(const double *part, int nparts,
                                                          int ntiles, int ntx, int h, int w,
                                                          double *g, long ldg) {
    __shared__ double red[kThreads / kRE][kRE + 1];
    const int le = threadIdx.x % kRE, sl = threadIdx.x / kRE;
    const int total = ntiles * kGT * kGT;
    const int e = blockIdx.x * kRE + le;
    double s = 0.0;
    if (e < total)
        for (int p = sl; p < nparts; p += kThreads / kRE) s += part[(size_t)p * total + e];
    red[sl][le] = s;
    __syncthreads();
    if (sl != 0 || e >= total) return;
    for (int q = 1; q < kThreads / kRE; q++) s += red[q][le];
    const int t = e / (kGT * kGT), l = e % (kGT * kGT);
    const int i = (t % ntx) * kGT + l % kGT, j = (t / ntx) * kGT + l / kGT;
    if (i < h && j < w) g[i + (size_t)j * ldg] = s;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_diagdot(long n, const T *x, long ldx, ColMap xm,
                                                      const T *y, long ldy, ColMap ym,
                                                      double *part) {
    __shared__ double red[kThreads / 64];
    const int j = blockIdx.y;
    const T *xc = colp(x, ldx, xm.idx[j]);
    const T *yc = colp(y, ldy, ym.idx[j]);
    double s = 0.0;
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads)
        s = fma((double)xc[r], (double)yc[r], s);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kThreads / 64; w++) t += red[w];
        part[(size_t)j * gridDim.x + blockIdx.x] = t;
    }
}

// out[j] = sum of the nparts partials of column j: one wave per column, lane
// l adds partials l, l + 64, ... and the lanes are summed in a fixed order
__global__ void __launch_bounds__(64) k_rowsum(const double *part, int nparts, double *out) {
    const int j = blockIdx.x, l = threadIdx.x;
    const double *pj = part + (size_t)j * nparts;
    double s = 0.0;
    for (int p = l; p < nparts; p += 64) s += pj[p];
    s = wave_sum(s);
    if (l == 0) out[j] = s;
}

// residual, its squared norm partials (the k_diagdot partition and order)
// and the optional diagonal scaling, in one pass
template <typename T>
__global__ void __launch_bounds__(kThreads) k_resid_norm(long n, const T *bx, const T *ax, long ld,
                                                         ColMap act, const double *lam, T *r,
                                                         long ldr, const T *tdiag, double *part) {
    __shared__ double red[kThreads / 64];
    // blockIdx.x: column (fastest, so the workgroups of one row part run
    // together and share the tdiag lines in L2); blockIdx.y: row part
    const int t = blockIdx.x, np = gridDim.y;
    const int c = act.idx[t];
    const T l = (T)lam[c];
    const T *bc = colp(bx, ld, c), *ac = colp(ax, ld, c);
    T *rc = colp(r, ldr, t);
    double s = 0.0;
    const long stride = (long)np * kThreads;
    long i = (long)blockIdx.y * kThreads + threadIdx.x;
    // four rows of the part in flight at a time, summed in row order
    for (; i + 3 * stride < n; i += 4 * stride) {
        T b[4], a4[4], d[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            b[u] = bc[i + u * stride];
            a4[u] = ac[i + u * stride];
            d[u] = tdiag ? tdiag[i + u * stride] : T(1);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const T v = l * b[u] - a4[u];
            s = fma((double)v, (double)v, s);
            rc[i + u * stride] = tdiag ? d[u] * v : v;
        }
    }
    for (; i < n; i += stride) {
        const T v = l * bc[i] - ac[i];
        s = fma((double)v, (double)v, s);
        rc[i] = tdiag ? tdiag[i] * v : v;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double u = 0.0;
        for (int w = 0; w < kThreads / 64; w++) u += red[w];
        part[(size_t)t * np + blockIdx.y] = u;
    }
}

// one thread per row walks the destinations in ascending order, so every
// source is read before a later destination could overwrite it
template <typename T>
__global__ void __launch_bounds__(kThreads) k_compact(long n, T *x, long ld, ColMap src) {
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads)
        for (int t = 0; t < src.n; t++)
            if (src.idx[t] != t) x[(size_t)t * ld + i] = x[(size_t)src.idx[t] * ld + i];
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_vecmat(long n, const T *x, long ldx, ColMap xm,
                                                     const double *c, long ldc, T *y, long ldy,
                                                     ColMap ym, int accumulate) {
    const int j = blockIdx.y;
    T *yc = colp(y, ldy, ym.idx[j]);
    const double *cj = c + (size_t)j * ldc;
    for (long r = (long)blockIdx.x * kThreads + threadIdx.x; r < n; r += (long)gridDim.x * kThreads) {
        T s = accumulate ? yc[r] : T(0);
        for (int i = 0; i < xm.n; i++) s = fma((T)cj[i], colp(x, ldx, xm.idx[i])[r], s);
        yc[r] = s;
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_residual(long n, const T *bx, const T *ax, long ld,
                                                       ColMap act, const double *lam, T *r,
                                                       long ldr) {
    const int t = blockIdx.y;
    const int c = act.idx[t];
    const T l = (T)lam[c];
    const T *bc = colp(bx, ld, c), *ac = colp(ax, ld, c);
    T *rc = colp(r, ldr, t);
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads)
        rc[i] = l * bc[i] - ac[i];
}

__global__ void k_small_mm(double alpha, const double *a, long lda, int ta, const double *b,
                           long ldb, int tb, double *c, long ldc, int m, int w, int l) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m * w; e += gridDim.x * blockDim.x) {
        const int i = e % m, j = e / m;
        double s = 0.0;
        for (int q = 0; q < l; q++) {
            const double av = ta ? a[q + (size_t)i * lda] : a[i + (size_t)q * lda];
            const double bv = tb ? b[j + (size_t)q * ldb] : b[q + (size_t)j * ldb];
            s += av * bv;
        }
        c[i + (size_t)j * ldc] = alpha * s;
    }
}

unsigned row_grid(long n, int per_thread = 1) {
    size_t blocks = ((size_t)n + (size_t)kThreads * per_thread - 1) / ((size_t)kThreads * per_thread);
    size_t cap = (size_t)num_cus() * 8;
    return (unsigned)(blocks < 1 ? 1 : (blocks > cap ? cap : blocks));
}

}  // namespace

#define BX_DISPATCH(prec, KERNEL, GRID, ...)                                             \
    do {                                                                                 \
        if ((prec) == HIPMV_FP32)                                                        \
            hipLaunchKernelGGL(KERNEL<float>, GRID, dim3(kThreads), 0, stream(), __VA_ARGS__); \
        else                                                                             \
            hipLaunchKernelGGL(KERNEL<double>, GRID, dim3(kThreads), 0, stream(), __VA_ARGS__); \
        BX_HIP(hipGetLastError());                                                       \
    } while (0)

template <typename T>
static inline T *tp(void *p) {
    return (T *)p;
}

void copy_cols(int prec, long n, const void *x, long ldx, const ColMap &xm, void *y, long ldy,
               const ColMap &ym) {
    BX_REQUIRE(xm.n == ym.n, "copy: active column counts differ (%d vs %d)", xm.n, ym.n);
    if (xm.n == 0 || n == 0) return;
    dim3 g(row_grid(n), xm.n);
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_copy<float>, g, dim3(kThreads), 0, stream(), n, (const float *)x, ldx,
                           xm, (float *)y, ldy, ym);
    else
        hipLaunchKernelGGL(k_copy<double>, g, dim3(kThreads), 0, stream(), n, (const double *)x,
                           ldx, xm, (double *)y, ldy, ym);
    BX_HIP(hipGetLastError());
}

void axpy_cols(int prec, long n, double alpha, const void *x, long ldx, const ColMap &xm, void *y,
               long ldy, const ColMap &ym) {
    BX_REQUIRE(xm.n == ym.n, "axpy: active column counts differ (%d vs %d)", xm.n, ym.n);
    if (xm.n == 0 || n == 0) return;
    dim3 g(row_grid(n), xm.n);
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_axpy<float>, g, dim3(kThreads), 0, stream(), n, alpha,
                           (const float *)x, ldx, xm, (float *)y, ldy, ym);
    else
        hipLaunchKernelGGL(k_axpy<double>, g, dim3(kThreads), 0, stream(), n, alpha,
                           (const double *)x, ldx, xm, (double *)y, ldy, ym);
    BX_HIP(hipGetLastError());
}

void scale_cols(int prec, long n, const double *coef, const void *x, long ldx, const ColMap &xm,
                void *y, long ldy, const ColMap &ym) {
    BX_REQUIRE(xm.n == ym.n, "scale: active column counts differ (%d vs %d)", xm.n, ym.n);
    if (xm.n == 0 || n == 0) return;
    Coef c;
    for (int j = 0; j < xm.n; j++) c.v[j] = coef[j];
    dim3 g(row_grid(n), xm.n);
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_scale<float>, g, dim3(kThreads), 0, stream(), n, c, (const float *)x,
                           ldx, xm, (float *)y, ldy, ym);
    else
        hipLaunchKernelGGL(k_scale<double>, g, dim3(kThreads), 0, stream(), n, c,
                           (const double *)x, ldx, xm, (double *)y, ldy, ym);
    BX_HIP(hipGetLastError());
}

void clear_cols(int prec, long n, void *x, long ldx, const ColMap &xm) {
    if (xm.n == 0 || n == 0) return;
    dim3 g(row_grid(n), xm.n);
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_clear<float>, g, dim3(kThreads), 0, stream(), n, (float *)x, ldx, xm);
    else
        hipLaunchKernelGGL(k_clear<double>, g, dim3(kThreads), 0, stream(), n, (double *)x, ldx, xm);
    BX_HIP(hipGetLastError());
}

void drand48_cols(int prec, long n, int seed, void *x, long ldx, const ColMap &xm) {
    if (xm.n == 0 || n == 0) return;
    // srand48: high 32 bits of the state from the seed, low 16 bits 0x330E
    const uint64_t x0 = (((uint64_t)(uint32_t)seed) << 16) | 0x330Eull;
    dim3 g(row_grid(n, kRandRun), xm.n);
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_drand48<float>, g, dim3(kThreads), 0, stream(), n, x0, (float *)x,
                           ldx, xm);
    else
        hipLaunchKernelGGL(k_drand48<double>, g, dim3(kThreads), 0, stream(), n, x0, (double *)x,
                           ldx, xm);
    BX_HIP(hipGetLastError());
}

void convert_block(long n, int k, const void *src, long lds, int prec_s, void *dst, long ldd,
                   int prec_d) {
    if (k == 0 || n == 0) return;
    dim3 g(row_grid(n), k);
    if (prec_s == HIPMV_FP64 && prec_d == HIPMV_FP32)
        hipLaunchKernelGGL((k_convert<double, float>), g, dim3(kThreads), 0, stream(), n,
                           (const double *)src, lds, (float *)dst, ldd);
    else if (prec_s == HIPMV_FP32 && prec_d == HIPMV_FP64)
        hipLaunchKernelGGL((k_convert<float, double>), g, dim3(kThreads), 0, stream(), n,
                           (const float *)src, lds, (double *)dst, ldd);
    else if (prec_s == HIPMV_FP64)
        hipLaunchKernelGGL((k_convert<double, double>), g, dim3(kThreads), 0, stream(), n,
                           (const double *)src, lds, (double *)dst, ldd);
    else
        hipLaunchKernelGGL((k_convert<float, float>), g, dim3(kThreads), 0, stream(), n,
                           (const float *)src, lds, (float *)dst, ldd);
    BX_HIP(hipGetLastError());
}

// partial tiles gram_cols' grid writes for these shapes
static unsigned gram_parts(long n, int ntiles) {
    size_t blocks = ((size_t)n + kGRows - 1) / kGRows;
    size_t cap = (size_t)num_cus() * 4 / (size_t)ntiles + 1;
    return (unsigned)(blocks < 1 ? 1 : (blocks > cap ? cap : blocks));
}

void gram_cols(int prec, long n, const void *x, long ldx, const ColMap &xm, const void *y,
               long ldy, const ColMap &ym, double *g, long ldg) {
    if (xm.n == 0 || ym.n == 0) return;
    const int ntiles = ((xm.n + kGT - 1) / kGT) * ((ym.n + kGT - 1) / kGT);
    const size_t words = (size_t)gram_parts(n, ntiles) * ntiles * kGT * kGT;
    gram_cols_into(prec, n, x, ldx, xm, y, ldy, ym, g, ldg,
                   (double *)scratch(0, words * sizeof(double)), words);
}

void gram_cols_into(int prec, long n, const void *x, long ldx, const ColMap &xm, const void *y,
                    long ldy, const ColMap &ym, double *g, long ldg, double *part,
                    size_t part_words) {
    if (xm.n == 0 || ym.n == 0) return;
    const int ntx = (xm.n + kGT - 1) / kGT, nty = (ym.n + kGT - 1) / kGT;
    const int ntiles = ntx * nty;
    unsigned nparts = gram_parts(n, ntiles);
    const size_t fit = part_words / ((size_t)ntiles * kGT * kGT);
    BX_REQUIRE(fit >= 1, "gram_cols_into: workspace of %zu words below one partial set",
               part_words);
    if (nparts > fit) nparts = (unsigned)fit;
    dim3 grid(nparts, ntiles);
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_gram<float>, grid, dim3(kThreads), 0, stream(), n, (const float *)x,
                           ldx, xm, (const float *)y, ldy, ym, ntx, part);
    else
        hipLaunchKernelGGL(k_gram<double>, grid, dim3(kThreads), 0, stream(), n, (const double *)x,
                           ldx, xm, (const double *)y, ldy, ym, ntx, part);
    BX_HIP(hipGetLastError());
    const int total = ntiles * kGT * kGT;
    hipLaunchKernelGGL(k_gram_reduce, dim3((total + kRE - 1) / kRE), dim3(kThreads), 0, stream(),
                       part, (int)nparts, ntiles, ntx, xm.n, ym.n, g, ldg);
    BX_HIP(hipGetLastError());
}

void diagdot_cols(int prec, long n, const void *x, long ldx, const ColMap &xm, const void *y,
                  long ldy, const ColMap &ym, double *d) {
    BX_REQUIRE(xm.n == ym.n, "diagdot: active column counts differ (%d vs %d)", xm.n, ym.n);
    if (xm.n == 0) return;
    unsigned nparts = row_grid(n);
    if (nparts > 512) nparts = 512;
    double *part = (double *)scratch(1, (size_t)nparts * xm.n * sizeof(double));
    dim3 grid(nparts, xm.n);
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_diagdot<float>, grid, dim3(kThreads), 0, stream(), n, (const float *)x,
                           ldx, xm, (const float *)y, ldy, ym, part);
    else
        hipLaunchKernelGGL(k_diagdot<double>, grid, dim3(kThreads), 0, stream(), n,
                           (const double *)x, ldx, xm, (const double *)y, ldy, ym, part);
    BX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rowsum, dim3(xm.n), dim3(64), 0, stream(), part, (int)nparts, d);
    BX_HIP(hipGetLastError());
}

void row_sums(const double *part, int nparts, int ncols, double *out) {
    if (ncols == 0) return;
    hipLaunchKernelGGL(k_rowsum, dim3(ncols), dim3(64), 0, stream(), part, nparts, out);
    BX_HIP(hipGetLastError());
}

void vecmat_cols(int prec, long n, const void *x, long ldx, const ColMap &xm, const double *c,
                 long ldc, void *y, long ldy, const ColMap &ym, bool accumulate) {
    if (ym.n == 0 || n == 0) return;
    dim3 g(row_grid(n), ym.n);
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_vecmat<float>, g, dim3(kThreads), 0, stream(), n, (const float *)x,
                           ldx, xm, c, ldc, (float *)y, ldy, ym, (int)accumulate);
    else
        hipLaunchKernelGGL(k_vecmat<double>, g, dim3(kThreads), 0, stream(), n,
                           (const double *)x, ldx, xm, c, ldc, (double *)y, ldy, ym,
                           (int)accumulate);
    BX_HIP(hipGetLastError());
}

void residual_norm_cols(int prec, long n, const void *bx, const void *ax, long ld,
                        const ColMap &act, const double *lam, void *r, long ldr, const void *t,
                        double *d) {
    if (act.n == 0 || n == 0) return;
    unsigned nparts = row_grid(n);
    if (nparts > 512) nparts = 512;  // the diagdot_cols partition
    double *part = (double *)scratch(1, (size_t)nparts * act.n * sizeof(double));
    dim3 g(act.n, nparts);
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_resid_norm<float>, g, dim3(kThreads), 0, stream(), n,
                           (const float *)bx, (const float *)ax, ld, act, lam, (float *)r, ldr,
                           (const float *)t, part);
    else
        hipLaunchKernelGGL(k_resid_norm<double>, g, dim3(kThreads), 0, stream(), n,
                           (const double *)bx, (const double *)ax, ld, act, lam, (double *)r, ldr,
                           (const double *)t, part);
    BX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rowsum, dim3(act.n), dim3(64), 0, stream(), part, (int)nparts, d);
    BX_HIP(hipGetLastError());
}

void compact_cols(int prec, long n, void *x, long ld, const ColMap &src) {
    bool moves = false;
    for (int t = 0; t < src.n; t++) {
        BX_REQUIRE(src.idx[t] >= t && (t == 0 || src.idx[t] > src.idx[t - 1]),
                   "compact_cols: sources must ascend from their destinations");
        moves = moves || src.idx[t] != t;
    }
    if (!moves || n == 0) return;
    dim3 g(row_grid(n));
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_compact<float>, g, dim3(kThreads), 0, stream(), n, (float *)x, ld, src);
    else
        hipLaunchKernelGGL(k_compact<double>, g, dim3(kThreads), 0, stream(), n, (double *)x, ld,
                           src);
    BX_HIP(hipGetLastError());
}

void residual_cols(int prec, long n, const void *bx, const void *ax, long ld, const ColMap &act,
                   const double *lam, void *r, long ldr) {
    if (act.n == 0 || n == 0) return;
    dim3 g(row_grid(n), act.n);
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_residual<float>, g, dim3(kThreads), 0, stream(), n, (const float *)bx,
                           (const float *)ax, ld, act, lam, (float *)r, ldr);
    else
        hipLaunchKernelGGL(k_residual<double>, g, dim3(kThreads), 0, stream(), n,
                           (const double *)bx, (const double *)ax, ld, act, lam, (double *)r, ldr);
    BX_HIP(hipGetLastError());
}

void small_mm(double alpha, const double *a, long lda, bool ta, const double *b, long ldb, bool tb,
              double *c, long ldc, int m, int w, int l) {
    if (m == 0 || w == 0) return;
    hipLaunchKernelGGL(k_small_mm, dim3((m * w + 255) / 256), dim3(256), 0, stream(), alpha, a,
                       lda, (int)ta, b, ldb, (int)tb, c, ldc, m, w, l);
    BX_HIP(hipGetLastError());
}

}  // namespace bx
