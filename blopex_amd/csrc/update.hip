// Block update on MFMA: for each group g,
//     tail = in1 * C[crow1 : crow1+h1, :] + in2 * C[crow2 : crow2+h2, :]
//     full = tail + in0 * C[crow0 : crow0+h0, :]
// written to out_tail / out_full (either may be absent).  With in0 = X,
// in1 = R, in2 = P this is LOBPCG's P' = R C_R + P C_P and X' = X C_X + P'
// in one pass over [X R P] (krylov/lobpcg.c:1200-1272, there 9-14
// MultiVecMat/Axpy/Copy calls); with one input it is the B-orthonormalising
// X <- X U^{-1} and the Ritz rotation X <- X C.
//
// Layout: each wave owns 64 consecutive rows r0..r0+63 and computes them for
// every output column, as four interleaved 16-row MFMA tiles (tile t holds
// rows r0 + 4j + t), so a lane reads and writes 4 consecutive rows of one
// column with one vector access.  The transposed product is formed,
//     D'(16 cols x 16 rows) += C'(16 x 4) * In'(4 x 16 rows),
// so C comes in as the A operand (from LDS) and the block rows as B.
// Rows a wave owns are read completely before they are written, so every
// output may alias the input of the same group it replaces.
#include "dense.h"
#include "dpp.h"
#include "kernels.h"
#include "mfma.h"

namespace bx {

namespace {

constexpr int kUT = 256;  // threads per workgroup (4 waves)

struct UpdKArgs {
    UpdGroup g[3];
    const double *C;
    int ldc;
    int crows;
    int w;
    long ld;
    const DevStatus *st;
    signed char tail_pos[64];  // out_tail column of each output column (-1: not stored)
};

// U groups of 4 input columns per step: U loads in flight per lane before
// their MFMAs wait on the first (one at a time leaves the kernel bound by
// load latency, not bandwidth)
template <typename T, int NOT>
__device__ __forceinline__ void upd_block(const T *in, int h, int crow, const T *Cl, long ld,
                                          long r0, int li, int lk,
                                          typename Mfma<T>::acc (&acc)[NOT][4]) {
    constexpr int WT = NOT * 16;
    constexpr int U = 4;
    for (int s = 0; s < h; s += 4 * U) {
        T v[U][4];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int kc = s + 4 * u + lk;
#pragma unroll
            for (int t = 0; t < 4; t++) v[u][t] = T(0);
            if (kc < h) load4(in + (size_t)kc * ld + r0, v[u]);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int kc = s + 4 * u + lk;
#pragma unroll
            for (int o = 0; o < NOT; o++) {
                const T cf = (kc < h) ? Cl[(crow + kc) * WT + o * 16 + li] : T(0);
#pragma unroll
                for (int t = 0; t < 4; t++) acc[o][t] = Mfma<T>::mma(cf, v[u][t], acc[o][t]);
            }
        }
    }
}

// output column c of the lane's accumulators goes to column pos[o][r] of out
// (skipped when negative)
template <typename T, int NOT>
__device__ __forceinline__ void upd_store(T *out, long ld, long r0, const int (&pos)[NOT][4],
                                          const typename Mfma<T>::acc (&acc)[NOT][4]) {
#pragma unroll
    for (int o = 0; o < NOT; o++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int c = pos[o][r];
            if (c >= 0) store4(out + (size_t)c * ld + r0, acc[o][0][r], acc[o][1][r], acc[o][2][r],
                               acc[o][3][r]);
        }
}

template <typename T, int NOT>
__global__ void __launch_bounds__(kUT) k_update(UpdKArgs a) {
    extern __shared__ unsigned char smem_raw[];
    T *Cl = reinterpret_cast<T *>(smem_raw);
    constexpr int WT = NOT * 16;
    if (a.st && a.st->abort) return;
    const UpdGroup &g = a.g[blockIdx.y];
    for (int e = threadIdx.x; e < a.crows * WT; e += kUT) {
        const int r = e / WT, c = e % WT;
        Cl[e] = (c < a.w) ? (T)a.C[r + (size_t)c * a.ldc] : T(0);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const long nchunks = a.ld / 64;
    const T *in0 = (const T *)g.in[0], *in1 = (const T *)g.in[1], *in2 = (const T *)g.in[2];
    int pfull[NOT][4], ptail[NOT][4];
#pragma unroll
    for (int o = 0; o < NOT; o++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int c = o * 16 + Mfma<T>::drow(lk, r);
            pfull[o][r] = c < a.w ? c : -1;
            ptail[o][r] = c < a.w ? (int)a.tail_pos[c] : -1;
        }
    for (long ch = (long)blockIdx.x * 4 + wave; ch < nchunks; ch += (long)gridDim.x * 4) {
        const long r0 = ch * 64 + 4 * li;
        typename Mfma<T>::acc acc[NOT][4];
#pragma unroll
        for (int o = 0; o < NOT; o++)
#pragma unroll
            for (int t = 0; t < 4; t++) acc[o][t] = typename Mfma<T>::acc{};
        if (in1) upd_block<T, NOT>(in1, g.h[1], g.crow[1], Cl, a.ld, r0, li, lk, acc);
        if (in2) upd_block<T, NOT>(in2, g.h[2], g.crow[2], Cl, a.ld, r0, li, lk, acc);
        if (g.out_tail) upd_store<T, NOT>((T *)g.out_tail, a.ld, r0, ptail, acc);
        if (in0) upd_block<T, NOT>(in0, g.h[0], g.crow[0], Cl, a.ld, r0, li, lk, acc);
        if (g.out_full) upd_store<T, NOT>((T *)g.out_full, a.ld, r0, pfull, acc);
    }
}

template <typename T, int NOT>
void launch_update(const UpdKArgs &a, int ngroups) {
    const size_t lds = (size_t)a.crows * NOT * 16 * sizeof(T);
    static bool attr = false;
    if (!attr) {
        BX_HIP(hipFuncSetAttribute((const void *)k_update<T, NOT>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024));
        attr = true;
    }
    // as many workgroups (over all groups) as are resident at once: the loop
    // over row chunks is persistent, and a second round would run as a tail
    static size_t occ_lds[8];
    static int occ_val[8], occ_n = 0;
    int per_cu = 0;
    for (int i = 0; i < occ_n; i++)
        if (occ_lds[i] == lds) per_cu = occ_val[i];
    if (per_cu == 0) {
        BX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_update<T, NOT>,
                                                            kUT, lds));
        if (per_cu < 1) per_cu = 1;
        if (occ_n < 8) {
            occ_lds[occ_n] = lds;
            occ_val[occ_n++] = per_cu;
        }
    }
    const long nchunks = a.ld / 64;
    long blocks = (nchunks + 3) / 4;
    const long cap = (long)num_cus() * per_cu / ngroups;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_update<T, NOT>), dim3((unsigned)blocks, ngroups), dim3(kUT), lds, stream(),
                       a);
    BX_HIP(hipGetLastError());
}

// The block update of the X and AX groups (B = I) with the residual of the
// new block folded in: one wave computes AX' and then X' for its 64 rows
// (both groups' accumulators live at once), stores them and P', AP' as
// k_update does, and forms R = T (X' diag(lam) - AX') for the columns that
// keep a tail position, with the squared norms of X' lam - AX' summed per
// workgroup (FP64) into part[tail column][workgroup].  Saves the residual
// pass's reads of X' and AX'.  T is the block's precision (FP32 or FP64).
struct UpdRArgs {
    UpdKArgs u;         // g[0] = X group, g[1] = AX group
    const double *lam;  // new Ritz values, by output column
    const void *tdiag;  // diagonal preconditioner (n entries, precision T) or null
    long n;             // rows of the problem (the rows from n to u.ld are padding)
    void *r;            // residual block (tail columns, precision T), leading dimension u.ld
    double *part;       // [tail column][nparts]
    int nparts;         // workgroups in the grid
};

// MINW: waves per SIMD the register allocation must allow (FP64 takes 2,
// else its registers allow one; FP32 is faster at its own choice of 2 than
// held to the 3 that a bound of 2 gives it)
template <typename T, int MINW>
__global__ void __launch_bounds__(kUT, MINW) k_update_resid(UpdRArgs a) {
    constexpr int NOT = 2;
    extern __shared__ unsigned char smem_raw[];
    T *Cl = reinterpret_cast<T *>(smem_raw);
    constexpr int WT = NOT * 16;
    const UpdKArgs &u = a.u;
    if (u.st && u.st->abort) return;
    for (int e = threadIdx.x; e < u.crows * WT; e += kUT) {
        const int r = e / WT, c = e % WT;
        Cl[e] = (c < u.w) ? (T)u.C[r + (size_t)c * u.ldc] : T(0);
    }
    __syncthreads();
    // tail positions and Ritz values of the output columns, read where used
    __shared__ int tposl[WT];
    __shared__ T laml[WT];
    for (int c = threadIdx.x; c < WT; c += kUT) {
        tposl[c] = c < u.w ? (int)u.tail_pos[c] : -1;
        laml[c] = c < u.w ? (T)a.lam[c] : T(0);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const long nchunks = u.ld / 64;
    const T *tdiag = (const T *)a.tdiag;
    T *rout = (T *)a.r;
    int pfull[NOT][4], ptail[NOT][4];
    double ss[NOT][4];
#pragma unroll
    for (int o = 0; o < NOT; o++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int c = o * 16 + Mfma<T>::drow(lk, r);
            pfull[o][r] = c < u.w ? c : -1;
            ptail[o][r] = tposl[c];
            ss[o][r] = 0.0;
        }
    const UpdGroup &gx = u.g[0], &ga = u.g[1];
    for (long ch = (long)blockIdx.x * 4 + wave; ch < nchunks; ch += (long)gridDim.x * 4) {
        const long r0 = ch * 64 + 4 * li;
        typename Mfma<T>::acc ax[NOT][4], xx[NOT][4];
#pragma unroll
        for (int o = 0; o < NOT; o++)
#pragma unroll
            for (int t = 0; t < 4; t++) ax[o][t] = xx[o][t] = typename Mfma<T>::acc{};
        // AX' (with AP' on the way), then X' (with P')
        if (ga.in[1]) upd_block<T, NOT>((const T *)ga.in[1], ga.h[1], ga.crow[1], Cl, u.ld, r0, li, lk, ax);
        if (ga.in[2]) upd_block<T, NOT>((const T *)ga.in[2], ga.h[2], ga.crow[2], Cl, u.ld, r0, li, lk, ax);
        if (ga.out_tail) upd_store<T, NOT>((T *)ga.out_tail, u.ld, r0, ptail, ax);
        upd_block<T, NOT>((const T *)ga.in[0], ga.h[0], ga.crow[0], Cl, u.ld, r0, li, lk, ax);
        upd_store<T, NOT>((T *)ga.out_full, u.ld, r0, pfull, ax);
        if (gx.in[1]) upd_block<T, NOT>((const T *)gx.in[1], gx.h[1], gx.crow[1], Cl, u.ld, r0, li, lk, xx);
        if (gx.in[2]) upd_block<T, NOT>((const T *)gx.in[2], gx.h[2], gx.crow[2], Cl, u.ld, r0, li, lk, xx);
        if (gx.out_tail) upd_store<T, NOT>((T *)gx.out_tail, u.ld, r0, ptail, xx);
        upd_block<T, NOT>((const T *)gx.in[0], gx.h[0], gx.crow[0], Cl, u.ld, r0, li, lk, xx);
        upd_store<T, NOT>((T *)gx.out_full, u.ld, r0, pfull, xx);
        // R = T (X' lam - AX') for the tail columns, as the residual kernel forms it
        // (the diagonal has n entries: padding rows, whose X' and AX' are
        // zero, keep the factor 1)
        T td[4] = {T(1), T(1), T(1), T(1)};
        if (tdiag) {
            if (r0 + 3 < a.n) {
                load4(tdiag + r0, td);
            } else {
#pragma unroll
                for (int t = 0; t < 4; t++)
                    if (r0 + t < a.n) td[t] = tdiag[r0 + t];
            }
        }
#pragma unroll
        for (int o = 0; o < NOT; o++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int q = ptail[o][r];
                if (q < 0) continue;
                const T lc = laml[o * 16 + Mfma<T>::drow(lk, r)];
                T v[4];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    v[t] = lc * xx[o][t][r] - ax[o][t][r];
                    ss[o][r] = fma((double)v[t], (double)v[t], ss[o][r]);
                }
                if (tdiag)
                    store4(rout + (size_t)q * u.ld + r0, td[0] * v[0], td[1] * v[1], td[2] * v[2], td[3] * v[3]);
                else
                    store4(rout + (size_t)q * u.ld + r0, v[0], v[1], v[2], v[3]);
            }
    }
    // norms: sum over the 16 lanes that share the output columns, then over
    // the four waves in order; one partial per workgroup
    __shared__ double wsum[4][WT];
#pragma unroll
    for (int o = 0; o < NOT; o++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const double t = group_sum16(ss[o][r]);
            if (li == 0) wsum[wave][o * 16 + Mfma<T>::drow(lk, r)] = t;
        }
    __syncthreads();
    for (int c = threadIdx.x; c < WT; c += kUT) {
        const int q = tposl[c];
        if (q >= 0)
            a.part[(size_t)q * a.nparts + blockIdx.x] = (wsum[0][c] + wsum[1][c]) + (wsum[2][c] + wsum[3][c]);
    }
}

// resident workgroups of k_update_resid<T> with coefficient rows for the
// largest pencil (LDS never limits it); sets the dynamic LDS limit once
template <typename T, int MINW>
int update_resid_occupancy() {
    static int per_cu = 0;
    if (per_cu == 0) {
        const size_t lds = (size_t)kMaxRR * 32 * sizeof(T);
        BX_HIP(hipFuncSetAttribute((const void *)k_update_resid<T, MINW>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        BX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_update_resid<T, MINW>,
                                                            kUT, lds));
        if (per_cu < 1) per_cu = 1;
    }
    return per_cu;
}

// D = A B with A[i][0] = i (other k-slots 0), B[0][j] = 1: D[i][j] = i;
// then A[i][0] = 1, B[0][j] = j: D[i][j] = j.
template <typename T>
__global__ void k_mfma_probe(int *rows, int *cols) {
    const int l = threadIdx.x, li = l & 15, lk = l >> 4;
    typename Mfma<T>::acc d1{}, d2{};
    d1 = Mfma<T>::mma(lk == 0 ? T(li) : T(0), lk == 0 ? T(1) : T(0), d1);
    d2 = Mfma<T>::mma(lk == 0 ? T(1) : T(0), lk == 0 ? T(li) : T(0), d2);
    for (int r = 0; r < 4; r++) {
        rows[4 * l + r] = (int)d1[r];
        cols[4 * l + r] = (int)d2[r];
    }
}

}  // namespace

void mfma_layout_probe(int prec, int *rows, int *cols) {
    int *d;
    BX_HIP(hipMalloc(&d, 2 * 256 * sizeof(int)));
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_mfma_probe<float>, dim3(1), dim3(64), 0, stream(), d, d + 256);
    else
        hipLaunchKernelGGL(k_mfma_probe<double>, dim3(1), dim3(64), 0, stream(), d, d + 256);
    BX_HIP(hipGetLastError());
    download_sync(rows, d, 256 * sizeof(int));
    download_sync(cols, d + 256, 256 * sizeof(int));
    BX_HIP(hipFree(d));
}

void block_update_resid(int prec, long n, long ld, const UpdGroup *groups, const double *C,
                        int ldc, int crows, int w, const DevStatus *st, const int *tail_pos,
                        int ntail, const double *lam, const void *tdiag, void *r, double *rn2) {
    BX_REQUIRE(prec == HIPMV_FP32 || prec == HIPMV_FP64, "block_update_resid: precision %d", prec);
    BX_REQUIRE(ld % 64 == 0, "block_update_resid: leading dimension %ld not a multiple of 64", ld);
    BX_REQUIRE(n >= 1 && n <= ld, "block_update_resid: %ld rows with leading dimension %ld", n, ld);
    BX_REQUIRE(w >= 1 && w <= 32, "block_update_resid: %d output columns (max 32)", w);
    UpdRArgs a;
    for (int i = 0; i < 2; i++) {
        a.u.g[i] = groups[i];
        BX_REQUIRE(groups[i].in[0] && groups[i].out_full, "block_update_resid: group %d", i);
        for (int b = 0; b < 3; b++) {
            if (!a.u.g[i].in[b]) continue;
            BX_REQUIRE(a.u.g[i].crow[b] >= 0 && a.u.g[i].crow[b] + a.u.g[i].h[b] <= crows,
                       "block_update_resid: coefficient rows %d..%d outside %d", a.u.g[i].crow[b],
                       a.u.g[i].crow[b] + a.u.g[i].h[b], crows);
        }
    }
    a.u.C = C;
    a.u.ldc = ldc;
    a.u.crows = crows;
    a.u.w = w;
    a.u.ld = ld;
    a.u.st = st;
    for (int c = 0; c < 64; c++) {
        const int q = c < w ? tail_pos[c] : -1;
        BX_REQUIRE(q >= -1 && q < ntail, "block_update_resid: tail position %d of column %d", q, c);
        a.u.tail_pos[c] = (signed char)q;
    }
    a.lam = lam;
    a.tdiag = tdiag;
    a.n = n;
    a.r = r;
    BX_REQUIRE(crows <= kMaxRR, "block_update_resid: %d coefficient rows", crows);
    const bool f32 = prec == HIPMV_FP32;
    const size_t lds = (size_t)crows * 32 * (f32 ? sizeof(float) : sizeof(double));
    const int per_cu = f32 ? update_resid_occupancy<float, 1>() : update_resid_occupancy<double, 2>();
    const long nchunks = ld / 64;
    long blocks = (nchunks + 3) / 4;
    const long cap = (long)num_cus() * per_cu;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    a.nparts = (int)blocks;
    a.part = (double *)scratch(1, (size_t)a.nparts * ntail * sizeof(double));
    if (f32)
        hipLaunchKernelGGL((k_update_resid<float, 1>), dim3((unsigned)blocks), dim3(kUT), lds, stream(), a);
    else
        hipLaunchKernelGGL((k_update_resid<double, 2>), dim3((unsigned)blocks), dim3(kUT), lds, stream(), a);
    BX_HIP(hipGetLastError());
    row_sums(a.part, a.nparts, ntail, rn2);
}

void block_update(int prec, long ld, const UpdGroup *groups, int ngroups, const double *C, int ldc,
                  int crows, int w, const DevStatus *st, const int *tail_pos) {
    BX_REQUIRE(ld % 64 == 0, "block_update: leading dimension %ld not a multiple of 64", ld);
    BX_REQUIRE(ngroups >= 1 && ngroups <= 3, "block_update: %d groups", ngroups);
    BX_REQUIRE(w >= 1 && w <= 64, "block_update: %d output columns (max 64)", w);
    UpdKArgs a;
    for (int i = 0; i < ngroups; i++) {
        a.g[i] = groups[i];
        for (int b = 0; b < 3; b++) {
            if (!a.g[i].in[b]) continue;
            BX_REQUIRE(a.g[i].crow[b] >= 0 && a.g[i].crow[b] + a.g[i].h[b] <= crows,
                       "block_update: coefficient rows %d..%d outside %d", a.g[i].crow[b],
                       a.g[i].crow[b] + a.g[i].h[b], crows);
        }
    }
    a.C = C;
    a.ldc = ldc;
    a.crows = crows;
    a.w = w;
    a.ld = ld;
    a.st = st;
    for (int c = 0; c < 64; c++) {
        const int q = c < w ? (tail_pos ? tail_pos[c] : c) : -1;
        BX_REQUIRE(q >= -1 && q < w, "block_update: tail position %d of column %d", q, c);
        a.tail_pos[c] = (signed char)q;
    }
    const int NOT = (w + 15) / 16;
    if (prec == HIPMV_FP32) {
        if (NOT == 1) launch_update<float, 1>(a, ngroups);
        else if (NOT == 2) launch_update<float, 2>(a, ngroups);
        else launch_update<float, 4>(a, ngroups);
    } else {
        if (NOT == 1) launch_update<double, 1>(a, ngroups);
        else if (NOT == 2) launch_update<double, 2>(a, ngroups);
        else launch_update<double, 4>(a, ngroups);
    }
}

}  // namespace bx
