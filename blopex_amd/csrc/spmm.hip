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
// 7-point stencils whose diagonal copy is laid out on an nx x ny x nz grid
// (hip_CSRMatrix::grid, nx = 64 or 128) march along z instead: a
// 512-thread workgroup owns 512 / nx full x-lines, one row per thread, of a
// chunk of consecutive planes.  X of planes z-1, z, z+1 at the thread's own
// row stays in registers, so each X value comes from HBM once per chunk
// instead of once for each of the three planes that use it; the x and y
// neighbours come from the plane-z slab of the tile in LDS, whose two halo
// lines are fetched together with X(z+1), when the workgroup of the
// neighbouring tile -- on the same XCD, at the same plane -- reads them as
// its own rows.  The sum runs over the same offsets in the same order, and
// neighbours outside the grid (zero coefficients) enter as zeros, so the
// bits equal those of the diagonal and CSR kernels.
//
// General CSR: one thread per row.  A row group's (column, value) pairs are
// staged through LDS with coalesced loads, then each row's pairs sit in
// registers for every block column, so the matrix streams from HBM once per
// product; for the stencil matrices the X reads of neighbouring lanes hit
// consecutive addresses of the same column.  Blocks are renumbered
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

// General CSR: a workgroup owns kSpThreads consecutive rows.  Their (column,
// value) pairs are one contiguous range of the CSR arrays, staged into LDS
// with coalesced non-temporal loads (they stream past L2, which is left to
// X); each thread then takes its row's pairs from LDS into registers and
// reuses them for every block column.  Y streams out non-temporally.
constexpr int kSpSeg = kSpThreads * 10;  // staged pairs per row group (rows of <= 10 on average)

template <typename T, int MAXE>
__global__ void __launch_bounds__(kSpThreads) k_spmm(long n, const int *__restrict__ rowptr,
                                                     const int *__restrict__ colind,
                                                     const T *__restrict__ val, const T *x, long ldx,
                                                     ColMap xm, T *y, long ldy, ColMap ym) {
    __shared__ int sc[kSpSeg];
    __shared__ T sv[kSpSeg];
    const long rb = xcd_block(blockIdx.x, gridDim.x) * kSpThreads;
    if (rb >= n) return;  // uniform over the workgroup
    const long r = rb + threadIdx.x;
    const long rend = rb + kSpThreads < n ? rb + kSpThreads : n;
    const int s0 = rowptr[rb], s1 = rowptr[rend];
    const bool staged = s1 - s0 <= kSpSeg;  // uniform
    if (staged)
        for (int e = threadIdx.x; e < s1 - s0; e += kSpThreads) {
            sc[e] = __builtin_nontemporal_load(colind + s0 + e);
            sv[e] = __builtin_nontemporal_load(val + s0 + e);
        }
    __syncthreads();
    if (r >= n) return;
    const int e0 = rowptr[r], e1 = rowptr[r + 1];
    const int len = e1 - e0;
    auto col = [&](int e) { return staged ? sc[e - s0] : colind[e]; };
    auto value = [&](int e) { return staged ? sv[e - s0] : val[e]; };
    if (len <= MAXE) {
        // short rows padded with (r, 0): fma(0, x, s) == s, so the padding
        // changes no bit and the inner loop has no branch
        int ci[MAXE];
        T v[MAXE];
#pragma unroll
        for (int e = 0; e < MAXE; e++) {
            const bool ok = e < len;
            ci[e] = ok ? col(e0 + e) : (int)r;
            v[e] = ok ? value(e0 + e) : T(0);
        }
        // one block column at a time: the gathers of a wave then touch one
        // column's lines, which the neighbouring rows reuse from L1
        for (int j = 0; j < xm.n; j++) {
            const T *xa = x + (size_t)xm.idx[j] * ldx;
            T sa = T(0);
#pragma unroll
            for (int e = 0; e < MAXE; e++) sa = fma(v[e], xa[ci[e]], sa);
            __builtin_nontemporal_store(sa, y + (size_t)ym.idx[j] * ldy + r);
        }
    } else {
        for (int j = 0; j < xm.n; j++) {
            const T *xa = x + (size_t)xm.idx[j] * ldx;
            T s = T(0);
            for (int e = e0; e < e1; e++) s = fma(value(e), xa[col(e)], s);
            __builtin_nontemporal_store(s, y + (size_t)ym.idx[j] * ldy + r);
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

constexpr int kZT = 512;  // threads of a z-marching workgroup

template <typename T, int KC>
struct ZmArgs {
    const T *dval;  // 7 x dld: offsets -P, -nx, -1, 0, 1, nx, P
    const T *x;     // first column of the chunk
    T *y;
    unsigned dld, ld;  // leading dimensions of dval and of x and y (elements)
    int ny, nz, zc, ntiles, nc, xcd;
};

// Element offsets are 32-bit (the host checks that they fit), so loads take
// one scalar base and a 32-bit vector offset.
template <typename T, int KC, int NXT>
__global__ void __launch_bounds__(kZT) k_zm(const ZmArgs<T, KC> a) {
    constexpr int TY = kZT / NXT;            // x-lines per workgroup
    constexpr int HPT = 2 * NXT * KC / kZT;  // halo values per thread
    constexpr int LINE = KC * NXT;           // one x-line of all columns
    __shared__ T slab[TY * LINE];            // plane z: [line][column][x]
    __shared__ T hbuf[2][2 * LINE];          // halo lines y0-1, y0+TY of planes z, z+1
    const int t = threadIdx.x, xq = t % NXT, ly = t / NXT;
    const long lin = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : (long)blockIdx.x;
    const int zc = (int)(lin / a.ntiles), tile = (int)(lin % a.ntiles);
    const int y0 = tile * TY, z0 = zc * a.zc, yy = y0 + ly;
    const unsigned plane = (unsigned)NXT * (unsigned)a.ny, ld = a.ld;
    auto row = [&](int z, int yv, int xv) {
        return (unsigned)z * plane + (unsigned)yv * NXT + (unsigned)xv;
    };
    // X(z-1), X(z), X(z+1) in use and X(z+2) in flight, the coefficients one
    // plane ahead: every load has a whole step to land.  The halo lines of
    // plane z+2 are fetched in the step in which the neighbouring workgroups
    // fetch them as their own X(z+2), so that one of the two reads finds them
    // in L2; they land in registers and go to LDS one step later.
    // Coefficients and Y stream past L2 (non-temporal).
    T xm1[KC], x0v[KC], xp1[KC], xp2[KC], hal[HPT], dv[7], dvn[7];
    auto load_plane = [&](int z, T (&v)[KC]) {
#pragma unroll
        for (int j = 0; j < KC; j++)
            v[j] = (j < a.nc && z >= 0 && z < a.nz) ? a.x[j * ld + row(z, yy, xq)] : T(0);
    };
    auto load_dv = [&](int z, T (&v)[7]) {
        const unsigned r = row(z < a.nz ? z : a.nz - 1, yy, xq);
#pragma unroll
        for (int d = 0; d < 7; d++) v[d] = __builtin_nontemporal_load(a.dval + d * a.dld + r);
    };
    // halo element e = t + kZT i: x = e % NXT, column (e / NXT) % KC, line
    // y0 - 1 for e < LINE, y0 + TY after; stored at hbuf[.][e]
    auto load_halo = [&](int z) {
#pragma unroll
        for (int i = 0; i < HPT; i++) {
            const int e = t + kZT * i, xv = e % NXT, j = (e / NXT) % KC;
            const int yv = e < LINE ? y0 - 1 : y0 + TY;
            const bool ok = j < a.nc && yv >= 0 && yv < a.ny;
            hal[i] = ok ? a.x[(unsigned)j * ld + row(z, ok ? yv : 0, xv)] : T(0);
        }
    };
    auto store_halo = [&](T *h) {
#pragma unroll
        for (int i = 0; i < HPT; i++) h[t + kZT * i] = hal[i];
    };
    load_plane(z0 - 1, xm1);
    load_plane(z0, x0v);
    load_plane(z0 + 1, xp1);
    load_dv(z0, dvn);
    load_halo(z0);
    store_halo(hbuf[z0 & 1]);
    if (z0 + 1 < z0 + a.zc) load_halo(z0 + 1);
    for (int z = z0; z < z0 + a.zc; z++) {
#pragma unroll
        for (int j = 0; j < KC; j++) slab[ly * LINE + j * NXT + xq] = x0v[j];
        if (z + 1 < z0 + a.zc) store_halo(hbuf[(z + 1) & 1]);
#pragma unroll
        for (int d = 0; d < 7; d++) dv[d] = dvn[d];
        if (z + 1 < z0 + a.zc) {  // the next step of this chunk
            load_plane(z + 2, xp2);
            load_dv(z + 1, dvn);
        }
        if (z + 2 < z0 + a.zc) load_halo(z + 2);
        const unsigned r = row(z, yy, xq);
        // lines y - 1 and y + 1 of this wave (ly is uniform in a wave)
        const T *hb = hbuf[z & 1];
        const T *lm = ly > 0 ? slab + (ly - 1) * LINE : hb;
        const T *lp = ly + 1 < TY ? slab + (ly + 1) * LINE : hb + LINE;
        const T *lc = slab + ly * LINE;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < KC; j++) {
            if (j >= a.nc) continue;
            const T ym = lm[j * NXT + xq];
            const T xmn = xq > 0 ? lc[j * NXT + xq - 1] : T(0);
            const T xpn = xq + 1 < NXT ? lc[j * NXT + xq + 1] : T(0);
            const T yp = lp[j * NXT + xq];
            T s = fma(dv[0], xm1[j], T(0));
            s = fma(dv[1], ym, s);
            s = fma(dv[2], xmn, s);
            s = fma(dv[3], x0v[j], s);
            s = fma(dv[4], xpn, s);
            s = fma(dv[5], yp, s);
            s = fma(dv[6], xp1[j], s);
            __builtin_nontemporal_store(s, a.y + j * ld + r);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < KC; j++) {
            xm1[j] = x0v[j];
            x0v[j] = xp1[j];
            xp1[j] = xp2[j];
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

// planes per z-marching chunk for this grid, or 0 when the diagonal kernel
// is used: about one workgroup per CU, chunks of at least min_zc planes
// (BLOPEX_AMD_ZM_MIN_ZC, default 4)
int zm_chunk(const hip_CSRMatrix *A) {
    const int nx = A->grid[0], ny = A->grid[1], nz = A->grid[2];
    if (A->ndiag != 7 || (nx != 64 && nx != 128)) return 0;
    const int ty = kZT / nx;
    if (ny % ty) return 0;
    const long ntiles = ny / ty;
    int min_zc = 4;
    if (const char *e = getenv("BLOPEX_AMD_ZM_MIN_ZC")) min_zc = atoi(e);
    int zc = (int)(ntiles * nz / num_cus());
    if (zc < 1) zc = 1;
    while (nz % zc) zc--;
    return zc >= min_zc ? zc : 0;
}

// the z-marching kernel takes contiguous column ranges
bool contiguous(const ColMap &m) {
    for (int j = 1; j < m.n; j++)
        if (m.idx[j] != m.idx[0] + j) return false;
    return true;
}

template <typename T, int KC, int NXT>
void launch_zm(const hip_CSRMatrix *A, int zc, const T *dv, const T *x, long ldx, const ColMap &xm,
               T *y, long ldy, const ColMap &ym) {
    ZmArgs<T, KC> a{};
    a.dval = dv;
    a.dld = (unsigned)A->diag_ld;
    a.ld = (unsigned)ldx;
    a.ny = A->grid[1];
    a.nz = A->grid[2];
    a.zc = zc;
    a.ntiles = A->grid[1] / (kZT / NXT);
    const long nb = (long)a.ntiles * (a.nz / zc);
    a.xcd = nb % kXcds == 0;
    for (int c0 = 0; c0 < xm.n; c0 += KC) {
        a.nc = xm.n - c0 < KC ? xm.n - c0 : KC;
        a.x = x + (size_t)(xm.idx[0] + c0) * ldx;
        a.y = y + (size_t)(ym.idx[0] + c0) * ldy;
        hipLaunchKernelGGL((k_zm<T, KC, NXT>), dim3((unsigned)nb), dim3(kZT), 0, stream(), a);
    }
}

template <typename T, int KC>
void spmm_zm(const hip_CSRMatrix *A, int zc, const T *dv, const T *x, long ldx, const ColMap &xm,
             T *y, long ldy, const ColMap &ym) {
    switch (A->grid[0]) {
        case 64: launch_zm<T, KC, 64>(A, zc, dv, x, ldx, xm, y, ldy, ym); break;
        default: launch_zm<T, KC, 128>(A, zc, dv, x, ldx, xm, y, ldy, ym); break;
    }
}

}  // namespace

int spmm_zm_chunk(int prec, const hip_CSRMatrix *A, long ld, int ncols) {
    // one leading dimension and 32-bit element offsets within a chunk of KC
    // columns (32 FP32, 16 FP64)
    const long kc = prec == HIPMV_FP32 ? 32 : 16;
    if (7 * A->diag_ld >= (1L << 31) || (ncols < kc ? ncols : kc) * ld >= (1L << 31)) return 0;
    return zm_chunk(A);
}

int spmm_kernel(int prec, const hip_CSRMatrix *A, long ld, int ncols) {
    if (spmm_zm_chunk(prec, A, ld, ncols)) return 2;
    return A->ndiag > 0 ? 1 : 0;
}

void spmm_csr(int prec, const hip_CSRMatrix *A, const void *x, long ldx, const ColMap &xm, void *y,
              long ldy, const ColMap &ym) {
    BX_REQUIRE(xm.n == ym.n, "spmm: active column counts differ (%d vs %d)", xm.n, ym.n);
    if (xm.n == 0 || A->n == 0) return;
    const int zc = ldx == ldy && contiguous(xm) && contiguous(ym)
                       ? spmm_zm_chunk(prec, A, ldx, xm.n)
                       : 0;
    if (zc) {
        if (prec == HIPMV_FP32)
            spmm_zm<float, 32>(A, zc, A->diag_val32, (const float *)x, ldx, xm, (float *)y, ldy, ym);
        else
            spmm_zm<double, 16>(A, zc, A->diag_val, (const double *)x, ldx, xm, (double *)y, ldy,
                                ym);
        BX_HIP(hipGetLastError());
        return;
    }
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
