// hip_LobpcgSolveBatched (include/lobpcg_batched.h): LOBPCG on many
// independent problems at once.
//
// Layout: the problems are stacked row-wise into one block-diagonal CSR
// matrix of N = sum n_p rows, and every block (X, AX, R, AR, P, AP) is one
// N x 24 array (FP64; FP32 copies for the mixed phase; column-major, leading
// dimension N); problem p owns rows off_p .. off_p + n_p and its first k_p
// columns (k wanted plus guard columns).  Rows are cut into chunks (2048
// rows, or the autotuner's choice) that never straddle two problems; every
// grouped kernel is launched over the chunks of the unconverged problems, so
// one launch advances the whole batch:
//   k_bresid     R = T (AX - X diag(lambda)) and per-chunk partial norms
//   k_bspmm      AR = A R over the active chunks of the stacked CSR matrix
//   k_bgram      per-chunk partial Gram blocks [X R P]' [AX AR AP] and
//                [X R P]' [X R P] (upper triangles, the problem's columns only)
//                of FP64 blocks; FP32 blocks take k_bgram_mfma (FP32 MFMA tiles)
//   k_bgram_sum  per-problem fixed-order sums of the partials
//   k_rr_tri_batched  one workgroup per problem: the single-problem path's
//                tridiagonal Rayleigh-Ritz solver (rr_fast.hip) on the
//                problem's pencil (FP64); the initial Rayleigh-Ritz in span(X)
//                uses k_brr (diagonal scaling, Cholesky, cyclic Jacobi)
//   k_bupd       X' = [X R P] C_X, P' = [R P] C_RP and the same for AX, AP
// The host reads the residual norms once per iteration.  The launches of an
// iteration are captured once into a hipGraph and replayed: their grids cover
// every chunk and problem, and workgroups past the active counts (device
// memory, rewritten when problems converge) exit at once.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "dense.h"
#include "dpp.h"
#include "kernels.h"
#include "lobpcg_batched.h"
#include "mfma.h"

namespace bx {
namespace {

constexpr int kK = HIP_BATCH_MAX_K + 8;  // largest block width: k wanted plus guard columns
constexpr int kM = 3 * kK;               // largest pencil
constexpr int kT = 256;                  // threads per workgroup
constexpr int kChunk = 2048;             // rows per chunk
constexpr int kRefresh = 20;             // FP32 iterations between recomputations of AX
constexpr int kHostEvery = 4;            // graph-driven iterations between host reads
constexpr int kSub = 32;                 // rows staged per Gram step
constexpr int kPairs = kM * (kM + 1) / 2;
constexpr int kPairsPerThread = (kPairs + kT - 1) / kT;

struct Chunks {
    const int *prob;   // problem of each active chunk
    const long *r0;    // first (global) row
    const int *nr;     // rows
    const int *count;  // active chunks (device): workgroups past it exit, so a captured
                       // launch stays valid while the batch shrinks
};

struct Probs {
    const long *off;
    const int *n;
    const int *k;
    const int *run;  // 0: the problem is skipped (it left after the last list rebuild); or null
};

template <typename T>
__device__ __forceinline__ double blk_get(const T *b, long ld, long row, int col) {
    return (double)b[(size_t)col * ld + row];
}

// ---- residual, Jacobi, per-chunk squared norms ----
// The problem's columns in groups of eight (KB: kp rounded up to 8, 16 or
// 24): a thread owns rows t, t + kT, ... of the chunk, issues the loads of a
// row's eight columns together and keeps their squared norms in registers;
// each group ends with DPP sums in every wave, and the four waves' sums are
// added in order once per chunk.
template <int KB, typename T>
__device__ __forceinline__ void bresid_rows(int kp, long ld, long r0, int nr, const T *X,
                                            const T *AX, const double *lj, const double *dinv,
                                            T *R, double (&red)[kT / 64][kK], double *out) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int g = 0; g < KB; g += 8) {
        if (g < kp) {  // uniform
            double acc[8];
#pragma unroll
            for (int u = 0; u < 8; u++) acc[u] = 0.0;
            for (int t = threadIdx.x; t < nr; t += kT) {
                const long r = r0 + t;
                T ax[8], xv[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const bool ok = g + u < kp;
                    ax[u] = ok ? AX[(size_t)(g + u) * ld + r] : T(0);
                    xv[u] = ok ? X[(size_t)(g + u) * ld + r] : T(0);
                }
                const double di = dinv ? dinv[r] : 1.0;
#pragma unroll
                for (int u = 0; u < 8; u++)
                    if (g + u < kp) {
                        double v = (double)ax[u] - lj[g + u] * (double)xv[u];
                        acc[u] = fma(v, v, acc[u]);
                        if (dinv) v *= di;
                        R[(size_t)(g + u) * ld + r] = (T)v;
                    }
            }
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (g + u < kp) {
                    const double sj = wave_sum(acc[u]);
                    if (lane == 0) red[wave][g + u] = sj;
                }
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < kp)
        out[threadIdx.x] = (red[0][threadIdx.x] + red[1][threadIdx.x]) +
                           (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

template <typename T>
__global__ void __launch_bounds__(kT) k_bresid(Chunks ch, Probs pr, long ld, const T *X,
                                               const T *AX, const double *lam,
                                               const double *dinv, T *R, double *part) {
    static_assert(kT == 256, "four waves");
    __shared__ double red[kT / 64][kK];
    __shared__ double lj[kK];
    const int c = blockIdx.x;
    if (c >= *ch.count) return;
    const int p = ch.prob[c];
    if (pr.run && !pr.run[p]) return;
    const int kp = pr.k[p];
    const long r0 = ch.r0[c];
    const int nr = ch.nr[c];
    if ((int)threadIdx.x < kK) lj[threadIdx.x] = lam[p * kK + threadIdx.x];
    __syncthreads();
    // the problem's columns only: R's other columns are never written and stay zero
    double *out = part + (size_t)c * kK;
    if (kp <= 8)
        bresid_rows<8, T>(kp, ld, r0, nr, X, AX, lj, dinv, R, red, out);
    else if (kp <= 16)
        bresid_rows<16, T>(kp, ld, r0, nr, X, AX, lj, dinv, R, red, out);
    else
        bresid_rows<kK, T>(kp, ld, r0, nr, X, AX, lj, dinv, R, red, out);
}

// res2[p K + j] = sum of problem p's chunk partials (fixed order)
__global__ void k_bres_sum(const int *act_prob, const int *nact, const int *pc_begin,
                           const int *pc_count, const int *run, const double *part, double *res2) {
    if ((int)blockIdx.x >= *nact) return;
    const int p = act_prob[blockIdx.x];
    if (run && !run[p]) return;
    const int j = threadIdx.x;
    if (j >= kK) return;
    // chunk by chunk in order; eight chunks' loads in flight at a time
    double s = 0.0;
    int c = pc_begin[p];
    const int e = pc_begin[p] + pc_count[p];
    for (; c + 8 <= e; c += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = part[(size_t)(c + u) * kK + j];
#pragma unroll
        for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; c < e; c++) s += part[(size_t)c * kK + j];
    res2[p * kK + j] = s;
}

// ---- grouped Gram blocks ----
// pencil index q = b * kp + col (block b of {X, R, P}); partials of the upper
// triangles of GA = L' A L and GB = L' L over the chunk's rows.  Each thread
// owns 2 x 2 blocks of the (pair-padded) upper triangle: six LDS reads feed
// eight FMAs.
constexpr int kHM = kM / 2;                               // block rows / columns
constexpr int kBlk = kHM * (kHM + 1) / 2;                 // upper 2 x 2 blocks
constexpr int kBlkPerThread = (kBlk + kT - 1) / kT;
// FP64 blocks (FP32 blocks take k_bgram_mfma).
__global__ void __launch_bounds__(kT) k_bgram(Chunks ch, Probs pr, long ld, int nb,
                                              const double *X, const double *R, const double *P,
                                              const double *AX, const double *AR,
                                              const double *AP, double *part) {
    using T = double;
    __shared__ T Ls[kSub][kM + 2];
    __shared__ T As[kSub][kM + 2];
    const int c = blockIdx.x;
    if (c >= *ch.count) return;
    const int p = ch.prob[c];
    if (pr.run && !pr.run[p]) return;
    const int kp = pr.k[p];
    const int m = nb * kp;
    const int hm = (m + 1) / 2;
    const int nblk = hm * (hm + 1) / 2;
    const long r0 = ch.r0[c];
    const int nr = ch.nr[c];
    const T *Lb[3] = {X, R, P};
    const T *Ab[3] = {AX, AR, AP};
    int bi[kBlkPerThread], bj[kBlkPerThread];
    double ga[kBlkPerThread][4], gb[kBlkPerThread][4];
#pragma unroll
    for (int s = 0; s < kBlkPerThread; s++) {
        const int o = threadIdx.x + s * kT;
        int j = (int)((sqrt(8.0 * o + 1.0) - 1.0) * 0.5);
        while ((j + 1) * (j + 2) / 2 <= o) j++;
        while (j * (j + 1) / 2 > o) j--;
        bi[s] = 2 * (o - j * (j + 1) / 2);
        bj[s] = 2 * j;
#pragma unroll
        for (int q = 0; q < 4; q++) ga[s][q] = gb[s][q] = 0.0;
    }
    // staged columns 0 .. 2 hm (an odd m gets a zero column); element
    // e = threadIdx.x + kT u of a step is row e % kSub of column e / kSub.
    // The next step's elements are loaded into registers while this step's
    // products run.
    constexpr int kPer = (kSub * (kM + 2) + kT - 1) / kT;
    const int nst = kSub * 2 * hm;
    T lreg[kPer], areg[kPer];
    auto fetch = [&](int s0) {
        const int rows = min(kSub, nr - s0);
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int e = threadIdx.x + kT * u, rr = e % kSub, q = e / kSub;
            lreg[u] = areg[u] = T(0);
            if (e < nst && rr < rows && q < m) {
                const int b = q / kp, col = q % kp;
                const size_t o = (size_t)col * ld + (size_t)(r0 + s0 + rr);
                lreg[u] = Lb[b][o];
                areg[u] = Ab[b][o];
            }
        }
    };
    fetch(0);
    for (int s0 = 0; s0 < nr; s0 += kSub) {
        const int rows = min(kSub, nr - s0);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int e = threadIdx.x + kT * u, rr = e % kSub, q = e / kSub;
            if (e < nst) {
                Ls[rr][q] = lreg[u];
                As[rr][q] = areg[u];
            }
        }
        __syncthreads();
        if (s0 + kSub < nr) fetch(s0 + kSub);
#pragma unroll
        for (int s = 0; s < kBlkPerThread; s++) {
            const int o = threadIdx.x + s * kT;
            if (o < nblk) {
                const int i = bi[s], j = bj[s];
                T a0 = ga[s][0], a1 = ga[s][1], a2 = ga[s][2], a3 = ga[s][3];
                T g0 = gb[s][0], g1 = gb[s][1], g2 = gb[s][2], g3 = gb[s][3];
                for (int rr = 0; rr < rows; rr++) {
                    const T l0 = Ls[rr][i], l1 = Ls[rr][i + 1];
                    const T x0 = As[rr][j], x1 = As[rr][j + 1];
                    const T y0 = Ls[rr][j], y1 = Ls[rr][j + 1];
                    a0 = fma(l0, x0, a0);
                    a1 = fma(l1, x0, a1);
                    a2 = fma(l0, x1, a2);
                    a3 = fma(l1, x1, a3);
                    g0 = fma(l0, y0, g0);
                    g1 = fma(l1, y0, g1);
                    g2 = fma(l0, y1, g2);
                    g3 = fma(l1, y1, g3);
                }
                ga[s][0] = a0; ga[s][1] = a1; ga[s][2] = a2; ga[s][3] = a3;
                gb[s][0] = g0; gb[s][1] = g1; gb[s][2] = g2; gb[s][3] = g3;
            }
        }
    }
    // entries (i + di, j + dj) with i + di <= j + dj < m, at pair index
    // (j + dj)(j + dj + 1)/2 + i + di
#pragma unroll
    for (int s = 0; s < kBlkPerThread; s++) {
        const int o = threadIdx.x + s * kT;
        if (o < nblk) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int ii = bi[s] + (q & 1), jj = bj[s] + (q >> 1);
                if (ii <= jj && jj < m) {
                    const int pair = jj * (jj + 1) / 2 + ii;
                    part[((size_t)c * 2 + 0) * kPairs + pair] = ga[s][q];
                    part[((size_t)c * 2 + 1) * kPairs + pair] = gb[s][q];
                }
            }
        }
    }
}

// FP32 blocks on MFMA: the chunk's rows in 16-, 32- or 64-row steps staged in LDS (the
// problem's m = nb kp columns of L = [X R P] and of [AX AR AP], m padded to
// whole 16-column tiles; the next step is loaded into registers while this
// one is multiplied), every upper tile product (ti <= tj) of GA = L' A L and
// GB = L' L dealt over the four waves, v_mfma_f32_16x16x4 over the chunk's
// rows in FP32 (the FP32 phase stops at switch_factor * tol, far above what
// this rounding affects); the chunk partials are summed in FP64 by
// k_bgram_sum.  Partials in k_bgram's layout.
constexpr int kMT = (kM + 15) / 16;                    // tiles per side
constexpr int kMP = kMT * 16;                          // padded columns
constexpr int kTilePairs = kMT * (kMT + 1) / 2;        // upper tile pairs
constexpr int kMProd = 2 * kTilePairs;                 // GA and GB
constexpr int kMPerWave = (kMProd + kT / 64 - 1) / (kT / 64);
constexpr int kMS = kMP + 4;                           // LDS row stride (floats)
// kMRows: rows per staged step (16, 32 or 64; the autotuner picks one)
template <int kMRows>
__global__ void __launch_bounds__(kT) k_bgram_mfma(Chunks ch, Probs pr, long ld, int nb,
                                                   const float *X, const float *R,
                                                   const float *P, const float *AX,
                                                   const float *AR, const float *AP,
                                                   double *part) {
    __shared__ float Ls[kMRows * kMS];
    __shared__ float As[kMRows * kMS];
    const int c = blockIdx.x;
    if (c >= *ch.count) return;
    const int p = ch.prob[c];
    if (pr.run && !pr.run[p]) return;
    const int kp = pr.k[p];
    const int m = nb * kp;
    const int tm = (m + 15) / 16;
    const int npair = tm * (tm + 1) / 2;
    const long r0 = ch.r0[c];
    const int nr = ch.nr[c];
    const float *Lb[3] = {X, R, P};
    const float *Ab[3] = {AX, AR, AP};
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    // product s of this wave: index w = wave + 4 s; w < npair: GA tile pair
    // w, else GB tile pair w - npair; pair index -> (ti, tj), ti <= tj
    int pti[kMPerWave], ptj[kMPerWave], pga[kMPerWave];
    f32x4 acc[kMPerWave];
#pragma unroll
    for (int s = 0; s < kMPerWave; s++) {
        int w = wave + 4 * s;
        pga[s] = w < npair ? 1 : (w < 2 * npair ? 0 : -1);
        if (w >= npair) w -= npair;
        int tj = 0;
        while ((tj + 1) * (tj + 2) / 2 <= w) tj++;
        pti[s] = w - tj * (tj + 1) / 2;
        ptj[s] = tj;
        acc[s] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    const int mp = tm * 16;
    // staging: element e = threadIdx.x + kT u (row fastest, so each column's
    // rows are one coalesced read; padding rows and columns are zeros).  The
    // next step's elements are loaded into registers while this step's
    // MFMAs run.
    constexpr int kPer = (kMRows * kMP + kT - 1) / kT;
    float lreg[kPer], areg[kPer];
    auto fetch = [&](int s0) {
        const int rows = min(kMRows, nr - s0);
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int e = threadIdx.x + kT * u, rr = e % kMRows, q = e / kMRows;
            lreg[u] = areg[u] = 0.f;
            if (q < m && rr < rows) {
                const int b = q / kp, col = q % kp;
                const size_t o = (size_t)col * ld + (size_t)(r0 + s0 + rr);
                lreg[u] = Lb[b][o];
                areg[u] = Ab[b][o];
            }
        }
    };
    fetch(0);
    for (int s0 = 0; s0 < nr; s0 += kMRows) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int e = threadIdx.x + kT * u, rr = e % kMRows, q = e / kMRows;
            if (q < mp) {
                Ls[rr * kMS + q] = lreg[u];
                As[rr * kMS + q] = areg[u];
            }
        }
        __syncthreads();
        if (s0 + kMRows < nr) fetch(s0 + kMRows);
#pragma unroll
        for (int s = 0; s < kMPerWave; s++) {
            if (pga[s] < 0) continue;  // uniform in the wave
            const int i0 = 16 * pti[s], j0 = 16 * ptj[s];
            const float *Q = pga[s] ? As : Ls;
            f32x4 d = acc[s];
#pragma unroll
            for (int k0 = 0; k0 < kMRows; k0 += 4) {
                const float a = Ls[(k0 + lk) * kMS + i0 + li];
                const float b = Q[(k0 + lk) * kMS + j0 + li];
                d = Mfma<float>::mma(a, b, d);
            }
            acc[s] = d;
        }
    }
    // lane (li, lk), register q: entry (i0 + drow(lk, q), j0 + li)
#pragma unroll
    for (int s = 0; s < kMPerWave; s++) {
        if (pga[s] < 0) continue;
        const int i0 = 16 * pti[s], j0 = 16 * ptj[s];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int ii = i0 + Mfma<float>::drow(lk, q), jj = j0 + li;
            if (ii <= jj && jj < m) {
                const int pair = jj * (jj + 1) / 2 + ii;
                part[((size_t)c * 2 + (pga[s] ? 0 : 1)) * kPairs + pair] = (double)acc[s][q];
            }
        }
    }
}

// full symmetric GA, GB (kM x kM, pencil indexing) of each active problem
__global__ void __launch_bounds__(kT) k_bgram_sum(const int *act_prob, const int *nact,
                                                  const int *pc_begin, const int *pc_count,
                                                  Probs pr, int nb, const double *part,
                                                  double *GA, double *GB) {
    if ((int)blockIdx.x >= *nact) return;
    const int p = act_prob[blockIdx.x];
    if (pr.run && !pr.run[p]) return;
    const int m = nb * pr.k[p];
    const int npairs = m * (m + 1) / 2;
    double *ga = GA + (size_t)p * kM * kM, *gb = GB + (size_t)p * kM * kM;
    for (int o = threadIdx.x; o < npairs; o += kT) {
        int j = (int)((sqrt(8.0 * o + 1.0) - 1.0) * 0.5);
        while ((j + 1) * (j + 2) / 2 <= o) j++;
        while (j * (j + 1) / 2 > o) j--;
        const int i = o - j * (j + 1) / 2;
        // chunk by chunk in order; eight chunks' loads in flight at a time
        double a = 0.0, g = 0.0;
        int c = pc_begin[p];
        const int e = pc_begin[p] + pc_count[p];
        for (; c + 8 <= e; c += 8) {
            double va[8], vg[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                va[u] = part[((size_t)(c + u) * 2 + 0) * kPairs + o];
                vg[u] = part[((size_t)(c + u) * 2 + 1) * kPairs + o];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                a += va[u];
                g += vg[u];
            }
        }
        for (; c < e; c++) {
            a += part[((size_t)c * 2 + 0) * kPairs + o];
            g += part[((size_t)c * 2 + 1) * kPairs + o];
        }
        ga[i + j * kM] = a;
        ga[j + i * kM] = a;
        gb[i + j * kM] = g;
        gb[j + i * kM] = g;
    }
}

// ---- the initial Rayleigh-Ritz in span(X), one workgroup per problem ----
constexpr int kRT = 256;
// Solves GA v = theta GB v on the leading m = nb kp rows / columns (launched
// with nb = 1); writes the kp smallest theta to lam and the GB-orthonormal v
// to C (pencil rows, kp columns, leading dim kM), nbe[p] = blocks used.  A GB
// that is not positive definite is retried with one block fewer; if no
// pencil is, status[p] = 1 and X is left as it is (C = [I; 0]).  The sweeps
// of its cyclic Jacobi run to completion (off-diagonal mass below 1e-32 of
// the diagonal's, or 30 sweeps).
__global__ void __launch_bounds__(kRT) k_brr(const int *act_prob, const int *nact, Probs pr, int nb,
                                             const double *GA, const double *GB, double *lam,
                                             double *C, int *nbe, int *status) {
    if ((int)blockIdx.x >= *nact) return;
    // the initial Rayleigh-Ritz only: pencils of one block (m = k <= kK)
    __shared__ double A[kK][kK + 1], B[kK][kK + 1], V[kK][kK + 1];
    __shared__ double sc[kM], ev[kM], cs[kM / 2], sn[kM / 2];
    __shared__ int order[kM], top[kM / 2], bot[kM / 2];
    __shared__ int fail, mm;
    const int p = act_prob[blockIdx.x];
    const int kp = pr.k[p];
    const int tid = threadIdx.x;
    const double *ga = GA + (size_t)p * kM * kM, *gb = GB + (size_t)p * kM * kM;
    for (int nbt = nb; nbt >= 1; nbt--) {
        const int m = nbt * kp;
        // load and scale to a unit diagonal of B
        for (int e = tid; e < m * m; e += kRT) {
            const int i = e % m, j = e / m;
            A[i][j] = ga[i + j * kM];
            B[i][j] = gb[i + j * kM];
        }
        if (tid == 0) fail = 0;
        __syncthreads();
        if (tid < m) {
            const double d = B[tid][tid];
            if (!(d > 0.0)) fail = 1;
            sc[tid] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < m * m; e += kRT) {
            const int i = e % m, j = e / m;
            A[i][j] *= sc[i] * sc[j];
            B[i][j] *= sc[i] * sc[j];
        }
        __syncthreads();
        // Cholesky B = L L' (lower, in place), right-looking
        for (int j = 0; j < m && !fail; j++) {
            if (tid == 0) {
                const double d = B[j][j];
                if (!(d > 1e-14)) fail = 1;
                B[j][j] = d > 0.0 ? sqrt(d) : 1.0;
            }
            __syncthreads();
            if (fail) break;
            for (int i = j + 1 + tid; i < m; i += kRT) B[i][j] /= B[j][j];
            __syncthreads();
            for (int e = tid; e < (m - j - 1) * (m - j - 1); e += kRT) {
                const int i = j + 1 + e % (m - j - 1), q = j + 1 + e / (m - j - 1);
                if (q <= i) B[i][q] -= B[i][j] * B[q][j];
            }
            __syncthreads();
        }
        __syncthreads();
        if (fail) {
            __syncthreads();
            continue;
        }
        // A <- L^-1 A L^-T: columns of L^-1 A (thread per column), then rows
        for (int col = tid; col < m; col += kRT)
            for (int i = 0; i < m; i++) {
                double s = A[i][col];
                for (int q = 0; q < i; q++) s -= B[i][q] * A[q][col];
                A[i][col] = s / B[i][i];
            }
        __syncthreads();
        for (int row = tid; row < m; row += kRT)
            for (int i = 0; i < m; i++) {
                double s = A[row][i];
                for (int q = 0; q < i; q++) s -= B[i][q] * A[row][q];
                A[row][i] = s / B[i][i];
            }
        __syncthreads();
        for (int e = tid; e < m * m; e += kRT) {
            const int i = e % m, j = e / m;
            if (i < j) {
                const double s = 0.5 * (A[i][j] + A[j][i]);
                A[i][j] = s;
                A[j][i] = s;
            }
            V[i][j] = i == j ? 1.0 : 0.0;
        }
        // cyclic Jacobi, round-robin pairs (m padded to even with an identity row)
        const int me = m + (m & 1);
        if (tid == 0) mm = me;
        if (m & 1) {
            for (int e = tid; e < me; e += kRT) {
                A[m][e] = 0.0;
                A[e][m] = 0.0;
                V[m][e] = 0.0;
                V[e][m] = 0.0;
            }
        }
        __syncthreads();
        if (tid == 0 && (m & 1)) {
            A[m][m] = 1e300;  // decoupled: never rotates
            V[m][m] = 1.0;
        }
        const int h = me / 2;
        if (tid < h) {
            top[tid] = 2 * tid;
            bot[tid] = 2 * tid + 1;
        }
        __syncthreads();
        // up to 30 sweeps, every nonzero a_pq rotated: stopping at the
        // rounding level of the off-diagonal mass (or skipping rotations
        // below 1e-15 sqrt(|a_pp a_qq|)) left Ritz vectors that slowed LOBPCG
        // near tight tolerances several-fold on small pencils, so the sweeps
        // run until the off-diagonal mass is below 1e-32 of the diagonal's
        __shared__ int rotated;
        __shared__ double offs, dia, ro[kRT], rd[kRT];
        for (int sweep = 0; sweep < 30; sweep++) {
            {
                double o = 0.0, d = 0.0;
                for (int e = tid; e < m * m; e += kRT) {
                    const int i = e % m, j = e / m;
                    const double v = A[i][j];
                    if (i != j)
                        o = fma(v, v, o);
                    else
                        d = fma(v, v, d);
                }
                ro[tid] = o;
                rd[tid] = d;
            }
            __syncthreads();
            if (tid == 0) {
                double o = 0.0, d = 0.0;
                for (int q = 0; q < kRT; q++) {
                    o += ro[q];
                    d += rd[q];
                }
                offs = o;
                dia = d;
                rotated = 0;
            }
            __syncthreads();
            if (offs <= 1e-32 * dia) break;
            for (int step = 0; step < me - 1; step++) {
                if (tid < h) {
                    const int a = min(top[tid], bot[tid]), b = max(top[tid], bot[tid]);
                    const double apq = A[a][b];
                    double c = 1.0, s = 0.0;
                    if (apq != 0.0) {
                        const double tau = (A[b][b] - A[a][a]) / (2.0 * apq);
                        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                        rotated = 1;
                    }
                    cs[tid] = c;
                    sn[tid] = s;
                }
                __syncthreads();
                // rows a, b of A
                for (int e = tid; e < h * me; e += kRT) {
                    const int pp = e / me, col = e % me;
                    const int a = min(top[pp], bot[pp]), b = max(top[pp], bot[pp]);
                    const double c = cs[pp], s = sn[pp];
                    const double xa = A[a][col], xb = A[b][col];
                    A[a][col] = c * xa - s * xb;
                    A[b][col] = s * xa + c * xb;
                }
                __syncthreads();
                // columns a, b of A and V
                for (int e = tid; e < h * me; e += kRT) {
                    const int pp = e / me, row = e % me;
                    const int a = min(top[pp], bot[pp]), b = max(top[pp], bot[pp]);
                    const double c = cs[pp], s = sn[pp];
                    const double xa = A[row][a], xb = A[row][b];
                    A[row][a] = c * xa - s * xb;
                    A[row][b] = s * xa + c * xb;
                    const double va = V[row][a], vb = V[row][b];
                    V[row][a] = c * va - s * vb;
                    V[row][b] = s * va + c * vb;
                }
                __syncthreads();
                // next round-robin pairing (position 0 of top stays; a single
                // pair never changes)
                if (tid == 0 && h > 1) {
                    const int tl = top[h - 1], b0 = bot[0];
                    for (int q = h - 1; q > 1; q--) top[q] = top[q - 1];
                    top[1] = b0;
                    for (int q = 0; q < h - 1; q++) bot[q] = bot[q + 1];
                    bot[h - 1] = tl;
                }
                __syncthreads();
            }
            if (!rotated) break;
            __syncthreads();
        }
        // ascending order of the m eigenvalues (selection by thread 0)
        if (tid == 0) {
            for (int i = 0; i < m; i++) {
                ev[i] = A[i][i];
                order[i] = i;
            }
            for (int i = 0; i < kp; i++) {
                int b = i;
                for (int j = i + 1; j < m; j++)
                    if (ev[order[j]] < ev[order[b]]) b = j;
                const int t = order[i];
                order[i] = order[b];
                order[b] = t;
            }
        }
        __syncthreads();
        // C = S L^-T V(:, order[0..kp)): back substitution per output column,
        // into A (its diagonal is saved in ev)
        double *cp = C + (size_t)p * kM * kK;
        for (int j = tid; j < kp; j += kRT) {
            const int src = order[j];
            for (int i = m - 1; i >= 0; i--) {
                double s = V[i][src];
                for (int q = i + 1; q < m; q++) s -= B[q][i] * A[q][j];
                A[i][j] = s / B[i][i];
            }
        }
        __syncthreads();
        for (int e = tid; e < kM * kp; e += kRT) {
            const int i = e % kM, j = e / kM;
            cp[i + j * kM] = i < m ? A[i][j] * sc[i] : 0.0;
        }
        for (int j = tid; j < kp; j += kRT) lam[p * kK + j] = ev[order[j]];
        if (tid == 0) {
            nbe[p] = nbt;
            status[p] = 0;
        }
        return;
    }
    // no positive definite pencil: keep X (C = [I; 0])
    double *cp = C + (size_t)p * kM * kK;
    for (int e = tid; e < kM * kK; e += kRT) cp[e] = (e % kM) == (e / kM) ? 1.0 : 0.0;
    if (tid == 0) {
        nbe[p] = 1;
        status[p] = 1;
    }
}

// ---- block update, row by row ----
// X' = sum_q C[q, j] L_q (q over the nbe blocks' columns), P' = the part with
// q >= kp; the same for the A-side blocks.  A thread owns one row; the
// output columns accumulate in registers in the blocks' precision (KB of
// them: the problem's column count rounded up to 8, 16 or 24, chosen per
// workgroup), the coefficients sit in LDS in that precision (each read, the
// same address across the wave, feeds the FMAs of a whole wave), and the
// loads of all of a block's columns are issued before their FMAs.
template <int KB, typename T>
__device__ __forceinline__ void bupd_rows(const T *cs, int kp, int nb, long ld, long r0, int nr,
                                          T *b0, const T *b1, T *b2) {
    const T *bl[3] = {b0, b1, b2};
    for (int t = threadIdx.x; t < nr; t += kT) {
        const long r = r0 + t;
        T xa[KB], pa[KB];
#pragma unroll
        for (int j = 0; j < KB; j++) xa[j] = pa[j] = T(0);
        for (int b = 0; b < nb; b++) {
            const T *src = bl[b] + r;
            T u[KB];
#pragma unroll
            for (int d = 0; d < KB; d++) u[d] = d < kp ? src[(size_t)d * ld] : T(0);
#pragma unroll
            for (int d = 0; d < KB; d++) {
                if (d < kp) {  // uniform
                    const T *cq = cs + (b * kp + d);
                    if (b == 0) {
#pragma unroll
                        for (int j = 0; j < KB; j++)
                            if (j < kp) xa[j] = fma(cq[j * kM], u[d], xa[j]);
                    } else {
#pragma unroll
                        for (int j = 0; j < KB; j++)
                            if (j < kp) pa[j] = fma(cq[j * kM], u[d], pa[j]);
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < KB; j++)
            if (j < kp) {
                if (nb > 1) b2[(size_t)j * ld + r] = pa[j];
                b0[(size_t)j * ld + r] = xa[j] + pa[j];
            }
    }
}

template <typename T>
__global__ void __launch_bounds__(kT) k_bupd(Chunks ch, Probs pr, long ld, const double *C,
                                             const int *nbe, int nb_all, T *X, T *R, T *P, T *AX,
                                             T *AR, T *AP) {
    __shared__ T cs[kM * kK];
    const int c = blockIdx.x;
    if (c >= *ch.count) return;
    const int p = ch.prob[c];
    if (pr.run && !pr.run[p]) return;
    const int kp = pr.k[p];
    const int nb = nbe ? nbe[p] : nb_all;
    const long r0 = ch.r0[c];
    const int nr = ch.nr[c];
    for (int e = threadIdx.x; e < kM * kK; e += kT) cs[e] = (T)C[(size_t)p * kM * kK + e];
    __syncthreads();
    for (int side = 0; side < 2; side++) {
        T *b0 = side ? AX : X, *b1 = side ? AR : R, *b2 = side ? AP : P;
        if (kp <= 8)
            bupd_rows<8, T>(cs, kp, nb, ld, r0, nr, b0, b1, b2);
        else if (kp <= 16)
            bupd_rows<16, T>(cs, kp, nb, ld, r0, nr, b0, b1, b2);
        else
            bupd_rows<kK, T>(cs, kp, nb, ld, r0, nr, b0, b1, b2);
    }
}

// ---- AR = A R over the active chunks (CSR rows of the stacked matrix) ----
template <int KB, typename T>
__device__ __forceinline__ void bspmm_rows(int kp, long ld, long r0, int nr, const int *rowptr,
                                           const int *colind, const T *val, const T *Rb, T *ARb) {
    for (int t = threadIdx.x; t < nr; t += kT) {
        const long r = r0 + t;
        T acc[KB];
#pragma unroll
        for (int j = 0; j < KB; j++) acc[j] = T(0);
        for (int e = rowptr[r], e1 = rowptr[r + 1]; e < e1; e++) {
            const T v = val[e];
            const long col = colind[e];
#pragma unroll
            for (int j = 0; j < KB; j++)
                if (j < kp) acc[j] = fma(v, Rb[(size_t)j * ld + col], acc[j]);
        }
#pragma unroll
        for (int j = 0; j < KB; j++)
            if (j < kp) ARb[(size_t)j * ld + r] = acc[j];
    }
}

template <typename T>
__global__ void __launch_bounds__(kT) k_bspmm(Chunks ch, Probs pr, long ld, const int *rowptr,
                                              const int *colind, const T *val, const T *Rb,
                                              T *ARb) {
    const int c = blockIdx.x;
    if (c >= *ch.count) return;
    const int p = ch.prob[c];
    if (pr.run && !pr.run[p]) return;
    const int kp = pr.k[p];
    const long r0 = ch.r0[c];
    const int nr = ch.nr[c];
    if (kp <= 8)
        bspmm_rows<8, T>(kp, ld, r0, nr, rowptr, colind, val, Rb, ARb);
    else if (kp <= 16)
        bspmm_rows<16, T>(kp, ld, r0, nr, rowptr, colind, val, Rb, ARb);
    else
        bspmm_rows<kK, T>(kp, ld, r0, nr, rowptr, colind, val, Rb, ARb);
}

// ---- X0: drand48 stream, problem by problem, column by column ----
constexpr uint64_t kM48 = (1ull << 48) - 1, kA48 = 0x5DEECE66Dull, kC48 = 0xBull;
__device__ uint64_t jump48(uint64_t x, uint64_t steps) {
    uint64_t a = kA48, c = kC48;
    while (steps) {
        if (steps & 1) x = (a * x + c) & kM48;
        c = (a * c + c) & kM48;
        a = (a * a) & kM48;
        steps >>= 1;
    }
    return x;
}
template <typename T>
__global__ void __launch_bounds__(kT) k_bdrand(Chunks ch, Probs pr, long ld, const long *sbase,
                                               uint64_t x0, T *X) {
    const int c = blockIdx.x;
    const int p = ch.prob[c];
    const int kp = pr.k[p], np = pr.n[p];
    const long r0 = ch.r0[c], off = pr.off[p];
    for (int j = 0; j < kp; j++)
        for (int t = threadIdx.x; t < ch.nr[c]; t += kT) {
            const long r = r0 + t;
            const uint64_t pos = (uint64_t)sbase[p] + (uint64_t)j * np + (uint64_t)(r - off);
            const uint64_t s = jump48(x0, pos + 1);
            X[(size_t)j * ld + r] = (T)(2.0 * ((double)s * (1.0 / 281474976710656.0)) - 1.0);
        }
}

template <typename T>
struct Blocks {
    using type = T;
    T *X, *AX, *R, *AR, *P, *AP;
};

// End of an iteration of a graph-driven phase (one workgroup): problem p
// leaves when all its wanted residuals are at most thr or its pencil broke
// down (the host loop's test, on the same values), recording the iteration
// and the outcome; when none, or fewer than min_active, are left the phase
// stops and every iteration launched after that does nothing.  ctl[0]:
// completed iterations of the phase, ctl[1]: stopped, ctl[2]: iteration of
// the stop.
__global__ void __launch_bounds__(1024) k_bcheck(int nprob, const int *kw, double thr,
                                                 int min_active, const double *res2,
                                                 const int *status, const DevStatus *dstat,
                                                 int *active, int *run, int *leave_it,
                                                 int *leave_conv, int *ctl) {
    __shared__ int left_s;
    const int tid = threadIdx.x;
    if (tid == 0) left_s = 0;
    __syncthreads();
    if (ctl[1]) return;  // uniform
    const int pit = ctl[0];
    int left = 0;
    for (int p = tid; p < nprob; p += blockDim.x) {
        if (!active[p]) continue;
        bool conv = true;
        for (int j = 0; j < kw[p]; j++) conv = conv && sqrt(fabs(res2[(size_t)p * kK + j])) <= thr;
        const bool broke = status[p] != 0 || dstat[p].abort != 0;
        if (conv || broke) {
            active[p] = 0;
            run[p] = 0;
            leave_it[p] = pit;
            leave_conv[p] = conv ? 1 : 0;
        } else {
            left++;
        }
    }
    if (left) atomicAdd(&left_s, left);
    __syncthreads();
    const int lt = left_s;
    const bool stop = lt == 0 || lt < min_active;
    if (stop)
        for (int p = tid; p < nprob; p += blockDim.x) run[p] = 0;
    if (tid == 0) {
        if (stop) {
            ctl[1] = 1;
            ctl[2] = pit;
        } else {
            ctl[0] = pit + 1;
        }
    }
}

template <typename T>
T *dev_alloc(size_t count) {
    T *p = nullptr;
    BX_HIP(hipMalloc((void **)&p, (count ? count : 1) * sizeof(T)));
    return p;
}

}  // namespace
}  // namespace bx

using namespace bx;

extern "C" {

void hip_BatchSettingsDefault(hip_BatchSettings *s) {
    if (!s) return;
    s->tol = 1e-6;
    s->maxit = 500;
    s->jacobi = 1;
    s->seed = 1;
    s->graph = 1;
    s->guard = 2;
    s->mode = HIPLOBPCG_FP64;
    s->switch_factor = 100.0;
    s->autotune = 1;
}

int hip_LobpcgSolveBatched(int nprob, const hip_BatchProblem *probs, const hip_BatchSettings *set,
                           double *lambda, double *resnorm, int *iterations, int *flags,
                           double *Xout, hip_BatchStats *st) {
    hip_BatchSettings s;
    hip_BatchSettingsDefault(&s);
    if (set) s = *set;
    if (nprob < 1 || !probs || !lambda || !resnorm || !iterations || !flags) {
        set_error("hip_LobpcgSolveBatched: need nprob >= 1 and the output arrays");
        return -1;
    }
    long N = 0, nnz = 0;
    int K = 1;
    for (int p = 0; p < nprob; p++) {
        const hip_BatchProblem &q = probs[p];
        if (q.k < 1 || q.k > HIP_BATCH_MAX_K || q.n < 3L * q.k || !q.rowptr || !q.colind || !q.val ||
            q.nnz != q.rowptr[q.n]) {
            set_error("hip_LobpcgSolveBatched: problem %d: need 1 <= k <= %d, n >= 3k and a "
                      "consistent CSR (n %ld, k %d, nnz %ld)",
                      p, HIP_BATCH_MAX_K, q.n, q.k, q.nnz);
            return -1;
        }
        N += q.n;
        nnz += q.nnz;
    }
    if (N > 0x7fffffffL || nnz > 0x7fffffffL) {
        set_error("hip_LobpcgSolveBatched: %ld rows / %ld entries exceed int32 indices", N, nnz);
        return -1;
    }
    // stacked block-diagonal CSR, Jacobi, per-problem data, chunks
    std::vector<int> rowptr(N + 1), colind(nnz);
    std::vector<double> val(nnz), dinv(N);
    std::vector<long> off(nprob), sbase(nprob);
    // pk: columns iterated (k plus guard columns), kw: columns wanted
    std::vector<int> pn(nprob), pk(nprob), kw(nprob);
    std::vector<int> chunk_prob;
    std::vector<long> chunk_r0;
    std::vector<int> chunk_nr, chunk_first(nprob), chunk_cnt(nprob);
    long r = 0, e = 0, sb = 0;
    for (int p = 0; p < nprob; p++) {
        const hip_BatchProblem &q = probs[p];
        off[p] = r;
        sbase[p] = sb;
        pn[p] = (int)q.n;
        kw[p] = q.k;
        pk[p] = (int)std::min<long>(std::min(q.k + std::max(s.guard, 0), kK), q.n / 3);
        K = std::max(K, pk[p]);
        sb += q.n * pk[p];
        for (long i = 0; i < q.n; i++) {
            rowptr[r + i] = (int)(e + q.rowptr[i]);
            double d = 0.0;
            for (int t = q.rowptr[i]; t < q.rowptr[i + 1]; t++) {
                if (q.colind[t] < 0 || q.colind[t] >= q.n) {
                    set_error("hip_LobpcgSolveBatched: problem %d row %ld: column %d outside "
                              "0..%ld",
                              p, i, q.colind[t], q.n - 1);
                    return -1;
                }
                colind[e + t] = (int)(r + q.colind[t]);
                val[e + t] = q.val[t];
                if (q.colind[t] == i) d = q.val[t];
            }
            dinv[r + i] = d != 0.0 ? 1.0 / d : 1.0;
        }
        r += q.n;
        e += q.nnz;
    }
    rowptr[N] = (int)nnz;
    // chunks of at most `rows` rows, problem by problem
    auto make_chunks = [&](int rows) {
        chunk_prob.clear();
        chunk_r0.clear();
        chunk_nr.clear();
        for (int p = 0; p < nprob; p++) {
            chunk_first[p] = (int)chunk_prob.size();
            for (long c0 = 0; c0 < pn[p]; c0 += rows) {
                chunk_prob.push_back(p);
                chunk_r0.push_back(off[p] + c0);
                chunk_nr.push_back((int)std::min<long>(rows, pn[p] - c0));
            }
            chunk_cnt[p] = (int)chunk_prob.size() - chunk_first[p];
        }
        return (int)chunk_prob.size();
    };
    const int kChunkCand[] = {512, 1024, 2048, 4096};
    const int nch_max = make_chunks(s.autotune ? kChunkCand[0] : kChunk);
    const auto t0 = std::chrono::steady_clock::now();
    hip_CSRMatrix *A = hip_CSRMatrixCreate(N, nnz, rowptr.data(), colind.data(), val.data());
    if (!A) return -1;
    int nch = make_chunks(kChunk), chunk_rows = kChunk;
    const size_t nk = (size_t)N * kK;
    double *X = dev_alloc<double>(nk), *AX = dev_alloc<double>(nk), *R = dev_alloc<double>(nk),
           *AR = dev_alloc<double>(nk), *P = dev_alloc<double>(nk), *AP = dev_alloc<double>(nk);
    for (double *b : {X, AX, R, AR, P, AP}) BX_HIP(hipMemsetAsync(b, 0, nk * 8, stream()));
    double *dv = dev_alloc<double>(N);
    upload_async(dv, dinv.data(), N * 8);
    long *d_off = dev_alloc<long>(nprob), *d_sbase = dev_alloc<long>(nprob);
    int *d_n = dev_alloc<int>(nprob), *d_k = dev_alloc<int>(nprob);
    upload_async(d_off, off.data(), nprob * 8);
    upload_async(d_sbase, sbase.data(), nprob * 8);
    upload_async(d_n, pn.data(), nprob * 4);
    upload_async(d_k, pk.data(), nprob * 4);
    int *c_prob = dev_alloc<int>(nch_max), *c_nr = dev_alloc<int>(nch_max);
    long *c_r0 = dev_alloc<long>(nch_max);
    int *d_actp = dev_alloc<int>(nprob), *d_pcb = dev_alloc<int>(nprob),
        *d_pcc = dev_alloc<int>(nprob), *d_counts = dev_alloc<int>(2);
    double *lam = dev_alloc<double>((size_t)nprob * kK);
    // what the host reads after every iteration, in one allocation so that one
    // copy brings it: the residual norms, the small pencils' status, the
    // Rayleigh-Ritz status
    const size_t rb_res = (size_t)nprob * kK * sizeof(double), rb_ds = (size_t)nprob * sizeof(DevStatus),
                 rb_st = (size_t)nprob * sizeof(int);
    char *rb = dev_alloc<char>(rb_res + rb_ds + rb_st);
    double *res2 = (double *)rb;
    DevStatus *dstat = (DevStatus *)(rb + rb_res);
    int *status = (int *)(rb + rb_res + rb_ds);
    double *GA = dev_alloc<double>((size_t)nprob * kM * kM), *GB = dev_alloc<double>((size_t)nprob * kM * kM);
    double *Cm = dev_alloc<double>((size_t)nprob * kM * kK);
    int *nbe = dev_alloc<int>(nprob);
    // the one-workgroup tridiagonal Rayleigh-Ritz solver of the single-problem
    // path, launched over the batch (explicit X blocks, P dropped when the
    // pencil is not positive definite, status per problem)
    const size_t rr_words = rr_fast_work_words(kM);
    double *rr_work = dev_alloc<double>((size_t)nprob * rr_words);
    BX_HIP(hipMemsetAsync(dstat, 0, sizeof(DevStatus) * nprob, stream()));
    double *part_r = dev_alloc<double>((size_t)nch_max * kK);
    double *part_g = dev_alloc<double>((size_t)nch_max * 2 * kPairs);
    BX_HIP(hipMemsetAsync(lam, 0, (size_t)nprob * kK * 8, stream()));
    BX_HIP(hipMemsetAsync(status, 0, nprob * 4, stream()));

    // device-side control of the graph-driven phases: active and run flags,
    // the iteration at which each problem left and how, the wanted widths,
    // and three control words (see k_bcheck)
    int *dflags = dev_alloc<int>((size_t)5 * nprob + 4);
    int *d_active = dflags, *d_run = dflags + nprob, *d_leave_it = dflags + 2 * nprob,
        *d_leave_conv = dflags + 3 * nprob, *d_kw = dflags + 4 * nprob, *d_ctl = dflags + 5 * nprob;
    upload_async(d_kw, kw.data(), nprob * 4);
    Probs pr{d_off, d_n, d_k, d_run};
    Chunks ch{c_prob, c_r0, c_nr, d_counts};
    const int *d_nact = d_counts + 1;
    std::vector<char> active(nprob, 1);
    std::vector<int> flag_h(nprob, -1), iters_h(nprob, 0);
    int nact_ch = 0, nact = 0;
    // active chunk list (chunks of a problem contiguous) and per-problem ranges
    auto rebuild = [&]() {
        std::vector<int> cp, cn, actp, pcb(nprob, 0), pcc(nprob, 0);
        std::vector<long> cr;
        for (int p = 0; p < nprob; p++) {
            if (!active[p]) continue;
            actp.push_back(p);
            pcb[p] = (int)cp.size();
            for (int c = chunk_first[p]; c < chunk_first[p] + chunk_cnt[p]; c++) {
                cp.push_back(chunk_prob[c]);
                cr.push_back(chunk_r0[c]);
                cn.push_back(chunk_nr[c]);
            }
            pcc[p] = (int)cp.size() - pcb[p];
        }
        nact_ch = (int)cp.size();
        nact = (int)actp.size();
        if (nact_ch) {
            upload_async(c_prob, cp.data(), cp.size() * 4);
            upload_async(c_r0, cr.data(), cr.size() * 8);
            upload_async(c_nr, cn.data(), cn.size() * 4);
        }
        if (nact) upload_async(d_actp, actp.data(), actp.size() * 4);
        upload_async(d_pcb, pcb.data(), nprob * 4);
        upload_async(d_pcc, pcc.data(), nprob * 4);
        const int counts[2] = {nact_ch, nact};
        upload_async(d_counts, counts, sizeof counts);
        // every problem in the lists runs
        std::vector<int> run(nprob);
        for (int p = 0; p < nprob; p++) run[p] = active[p] ? 1 : 0;
        upload_async(d_run, run.data(), nprob * 4);
    };
    rebuild();
    const ColMap all = colmap_identity(kK);
    const bool mixed = s.mode == HIPLOBPCG_MIXED;
    Blocks<double> b64{X, AX, R, AR, P, AP};
    Blocks<float> b32{};
    if (mixed) {
        float **f[6] = {&b32.X, &b32.AX, &b32.R, &b32.AR, &b32.P, &b32.AP};
        for (float **q : f) {
            *q = dev_alloc<float>(nk);
            BX_HIP(hipMemsetAsync(*q, 0, nk * 4, stream()));
        }
    }
    // the Gram partials: FP32 blocks on MFMA (gram_rows rows per staged step,
    // the autotuner's choice), FP64 blocks on FP64 FMA
    int gram_rows = 32;
    auto launch_gram = [&](auto &bk, int nb) {
        using T = typename std::remove_reference<decltype(bk)>::type::type;
        if constexpr (std::is_same<T, float>::value)
            hipLaunchKernelGGL(gram_rows == 16   ? k_bgram_mfma<16>
                               : gram_rows == 64 ? k_bgram_mfma<64>
                                                 : k_bgram_mfma<32>,
                               dim3(nch), dim3(kT), 0, stream(), ch, pr, N, nb, bk.X, bk.R, bk.P,
                               bk.AX, bk.AR, bk.AP, part_g);
        else
            hipLaunchKernelGGL(k_bgram, dim3(nch), dim3(kT), 0, stream(), ch, pr, N, nb,
                               bk.X, bk.R, bk.P, bk.AX, bk.AR, bk.AP, part_g);
    };
    // threshold and smallest batch of the device check at the end of each
    // step (0: no check; set by the graph-driven phase)
    double check_thr = 0.0;
    int check_min = 1;
    // grids cover every chunk / problem; workgroups past the device counts
    // exit at once, so one captured graph serves a whole phase
    auto step = [&](auto &bk, int nb, bool with_spmm) {
        using T = typename std::remove_reference<decltype(bk)>::type::type;
        const T *vals = std::is_same<T, float>::value ? (const T *)A->val32 : (const T *)A->val;
        if (with_spmm)
            hipLaunchKernelGGL(k_bspmm<T>, dim3(nch), dim3(kT), 0, stream(), ch, pr, N,
                               A->rowptr, A->colind, vals, (const T *)bk.R, bk.AR);
        launch_gram(bk, nb);
        hipLaunchKernelGGL(k_bgram_sum, dim3(nprob), dim3(kT), 0, stream(), d_actp, d_nact, d_pcb,
                           d_pcc, pr, nb, part_g, GA, GB);
        if (nb == 1) {
            hipLaunchKernelGGL(k_brr, dim3(nprob), dim3(kRT), 0, stream(), d_actp, d_nact, pr, nb,
                               GA, GB, lam, Cm, nbe, status);
        } else {
            RR2Args base{GA, GB, kM, 0, 0, 0, lam, lam, Cm, kM, rr_work, dstat, 1, 1};
            rr_fast_batched(base, d_actp, d_nact, d_k, d_run, nprob, K, nb == 3, (long)kM * kM, kK,
                            (long)kM * kK, (long)rr_words);
        }
        hipLaunchKernelGGL(k_bupd<T>, dim3(nch), dim3(kT), 0, stream(), ch, pr, N, Cm,
                           nb == 1 ? nbe : nullptr, nb, bk.X, bk.R, bk.P, bk.AX, bk.AR, bk.AP);
        hipLaunchKernelGGL(k_bresid<T>, dim3(nch), dim3(kT), 0, stream(), ch, pr, N,
                           (const T *)bk.X, (const T *)bk.AX, lam, s.jacobi ? dv : nullptr, bk.R,
                           part_r);
        hipLaunchKernelGGL(k_bres_sum, dim3(nprob), dim3(64), 0, stream(), d_actp, d_nact, d_pcb,
                           d_pcc, d_run, part_r, res2);
        BX_HIP(hipGetLastError());
        if (check_thr > 0.0) {
            hipLaunchKernelGGL(k_bcheck, dim3(1), dim3(1024), 0, stream(), nprob, d_kw, check_thr,
                               check_min, res2, status, dstat, d_active, d_run, d_leave_it,
                               d_leave_conv, d_ctl);
            BX_HIP(hipGetLastError());
        }
    };

    std::vector<double> res_h((size_t)nprob * kK), lam_h((size_t)nprob * kK);
    std::vector<int> st_h(nprob);
    std::vector<DevStatus> ds_h(nprob);
    std::vector<char> rb_h(rb_res + rb_ds + rb_st);
    int captures = 0, replays = 0, it = 0, it32 = 0;
    // The graph-driven phase: every step ends with k_bcheck, which takes the
    // problems that converged or broke down out on the device (their
    // workgroups exit from the next iteration on) and stops the phase when
    // too few are left.  The host reads the outcome every kHostEvery
    // iterations, calls done for the problems that left with the iteration
    // the device recorded, and rebuilds the lists without them.  The tests
    // are the host loop's below on the same values, so both give the same
    // iterations and the same bits.
    auto phase_graph = [&](auto &bk, double thr, int it_cap, int min_active, auto done) {
        using T = typename std::remove_reference<decltype(bk)>::type::type;
        {
            std::vector<int> a(nprob);
            for (int p = 0; p < nprob; p++) a[p] = active[p] ? 1 : 0;
            upload_async(d_active, a.data(), nprob * 4);
            const int ctl0[3] = {0, 0, 0};
            upload_async(d_ctl, ctl0, sizeof ctl0);
        }
        check_thr = thr;
        check_min = min_active;
        step(bk, 1, false);
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int pit = 0;
        std::vector<int> fl((size_t)5 * nprob + 4);
        for (;;) {
            download_sync(fl.data(), dflags, fl.size() * sizeof(int));
            const int *a = fl.data(), *lit = a + 2 * nprob, *lcv = a + 3 * nprob,
                      *ctl = a + 5 * nprob;
            bool changed = false;
            for (int p = 0; p < nprob; p++)
                if (active[p] && !a[p]) {
                    active[p] = 0;
                    changed = true;
                    done(p, lcv[p] != 0, lit[p]);
                }
            if (ctl[1]) {  // stopped: the iterations launched after the stop did nothing
                it -= pit - ctl[2];
                pit = ctl[2];
                break;
            }
            if (pit >= it_cap) break;
            if (changed) rebuild();
            const int batch = std::min(kHostEvery, it_cap - pit);
            for (int b = 0; b < batch; b++) {
                pit++;
                it++;
                if (std::is_same<T, float>::value && pit % kRefresh == 0)
                    hipLaunchKernelGGL(k_bspmm<T>, dim3(nch), dim3(kT), 0, stream(), ch, pr, N,
                                       A->rowptr, A->colind, (const T *)A->val32,
                                       (const T *)bk.X, bk.AX);
                if (pit == 1) {
                    step(bk, 2, true);
                    continue;
                }
                if (!exec) {
                    BX_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeRelaxed));
                    step(bk, 3, true);
                    BX_HIP(hipStreamEndCapture(stream(), &graph));
                    BX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                    captures++;
                } else {
                    replays++;
                }
                BX_HIP(hipGraphLaunch(exec, stream()));
            }
        }
        check_thr = 0.0;
        if (exec) {
            BX_HIP(hipGraphExecDestroy(exec));
            BX_HIP(hipGraphDestroy(graph));
        }
        return pit;
    };
    // One phase: from the Rayleigh-Ritz step in span(X) (AX current), iterate
    // until every active problem's wanted residuals are below thr, its pencil
    // broke down, or the phase has run it_cap iterations; done(p, conv) is
    // called as each problem leaves.  With graphs, phase_graph.
    auto phase = [&](auto &bk, double thr, int it_cap, int min_active, auto done) {
        if (s.graph) return phase_graph(bk, thr, it_cap, min_active, done);
        step(bk, 1, false);
        int pit = 0;
        for (;;) {
            download_sync(rb_h.data(), rb, rb_h.size());
            memcpy(res_h.data(), rb_h.data(), rb_res);
            memcpy(ds_h.data(), rb_h.data() + rb_res, rb_ds);
            memcpy(st_h.data(), rb_h.data() + rb_res + rb_ds, rb_st);
            bool changed = false;
            for (int p = 0; p < nprob; p++) {
                if (!active[p]) continue;
                bool conv = true;
                for (int j = 0; j < kw[p]; j++) conv = conv && sqrt(fabs(res_h[p * kK + j])) <= thr;
                const bool broke = st_h[p] || ds_h[p].abort;
                if (conv || broke) {
                    active[p] = 0;
                    changed = true;
                    done(p, conv, pit);
                }
            }
            const int left = (int)std::count(active.begin(), active.end(), 1);
            if (left == 0 || left < min_active || pit >= it_cap) break;
            if (changed) rebuild();
            pit++;
            it++;
            const int nb = pit == 1 ? 2 : 3;
            // FP32 blocks: AX follows X by the update recurrence, whose rounding
            // accumulates; every kRefresh iterations it is recomputed from X
            // (residual replacement) so that the FP32 residuals stay those of X
            if (std::is_same<typename std::remove_reference<decltype(bk)>::type::type,
                             float>::value &&
                pit % kRefresh == 0) {
                using T = typename std::remove_reference<decltype(bk)>::type::type;
                hipLaunchKernelGGL(k_bspmm<T>, dim3(nch), dim3(kT), 0, stream(), ch, pr, N,
                                   A->rowptr, A->colind, (const T *)A->val32, (const T *)bk.X,
                                   bk.AX);
            }
            step(bk, nb, true);
        }
        return pit;
    };
    // Autotuner: the launch shapes that make the SpMM and the three-block
    // Gram pass fastest on this batch, timed with HIP events on the blocks of
    // the first phase (still zero: the kernels' cost does not depend on the
    // values).  Coordinate search: the rows per chunk (the work of one
    // workgroup in the grouped kernels), then for FP32 blocks the Gram's
    // staged rows per step (16, 32, 64).
    auto tune = [&](auto &bk) {
        using T = typename std::remove_reference<decltype(bk)>::type::type;
        const T *vals = std::is_same<T, float>::value ? (const T *)A->val32 : (const T *)A->val;
        hipEvent_t e0, e1;
        BX_HIP(hipEventCreate(&e0));
        BX_HIP(hipEventCreate(&e1));
        auto time_pass = [&]() {
            float ms = 0.f;
            for (int rep = 0; rep < 3; rep++) {
                BX_HIP(hipEventRecord(e0, stream()));
                hipLaunchKernelGGL(k_bspmm<T>, dim3(nch), dim3(kT), 0, stream(), ch, pr, N,
                                   A->rowptr, A->colind, vals, (const T *)bk.R, bk.AR);
                launch_gram(bk, 3);
                BX_HIP(hipEventRecord(e1, stream()));
                BX_HIP(hipEventSynchronize(e1));
                float t = 0.f;
                BX_HIP(hipEventElapsedTime(&t, e0, e1));
                if (rep > 0) ms += t;  // the first launch warms up
            }
            return ms;
        };
        float best = 1e30f;
        for (int rows : kChunkCand) {
            nch = make_chunks(rows);
            rebuild();
            const float ms = time_pass();
            if (ms < best) {
                best = ms;
                chunk_rows = rows;
            }
        }
        nch = make_chunks(chunk_rows);
        rebuild();
        if (std::is_same<T, float>::value) {
            int bestg = gram_rows;
            for (int gr : {16, 64}) {
                gram_rows = gr;
                const float ms = time_pass();
                if (ms < best) {
                    best = ms;
                    bestg = gr;
                }
            }
            gram_rows = bestg;
        }
        BX_HIP(hipEventDestroy(e0));
        BX_HIP(hipEventDestroy(e1));
    };
    if (s.autotune) {
        if (mixed)
            tune(b32);
        else
            tune(b64);
    }
    const uint64_t x0 = (((uint64_t)(uint32_t)s.seed) << 16) | 0x330Eull;
    std::vector<int> it_fp32(nprob, 0);
    if (mixed) {
        // FP32 phase to switch_factor * tol (FP64 Gram sums and Rayleigh-Ritz)
        hipLaunchKernelGGL(k_bdrand<float>, dim3(nch), dim3(kT), 0, stream(), ch, pr, N, d_sbase,
                           x0, b32.X);
        BX_HIP(hipGetLastError());
        spmm_csr(HIPMV_FP32, A, b32.X, N, all, b32.AX, N, all);
        // the phase ends when fewer than 5% of the problems are left in it (the
        // last ones would keep the whole batch iterating in FP32)
        it32 = phase(b32, s.switch_factor * s.tol, s.maxit / 2, std::max(1, nprob / 20),
                     [&](int p, bool, int pit) { it_fp32[p] = pit; });
        for (int p = 0; p < nprob; p++)
            if (active[p]) it_fp32[p] = it32;
        // every problem continues in FP64 from its FP32 block; AX is recomputed
        // from X, P starts empty
        convert_block(N, kK, b32.X, N, HIPMV_FP32, X, N, HIPMV_FP64);
        BX_HIP(hipMemsetAsync(dstat, 0, sizeof(DevStatus) * nprob, stream()));
        BX_HIP(hipMemsetAsync(status, 0, nprob * 4, stream()));
        std::fill(active.begin(), active.end(), 1);
        rebuild();
        // a block the FP32 phase left without full rank (its Rayleigh-Ritz
        // in FP64 fails, or its residuals are not finite) starts again from
        // X0 in FP64
        spmm_csr(HIPMV_FP64, A, X, N, all, AX, N, all);
        step(b64, 1, false);
        download_sync(res_h.data(), res2, res_h.size() * 8);
        download_sync(st_h.data(), status, nprob * 4);
        bool any = false;
        for (int p = 0; p < nprob; p++) {
            bool bad = st_h[p] != 0;
            for (int j = 0; j < pk[p]; j++) bad = bad || !std::isfinite(res_h[p * kK + j]);
            active[p] = bad;
            any = any || bad;
        }
        if (any) {
            rebuild();
            hipLaunchKernelGGL(k_bdrand<double>, dim3(nact_ch), dim3(kT), 0, stream(), ch, pr, N,
                               d_sbase, x0, X);
            BX_HIP(hipGetLastError());
            BX_HIP(hipMemsetAsync(status, 0, nprob * 4, stream()));
        }
        std::fill(active.begin(), active.end(), 1);
        rebuild();
    } else {
        hipLaunchKernelGGL(k_bdrand<double>, dim3(nch), dim3(kT), 0, stream(), ch, pr, N,
                           d_sbase, x0, X);
        BX_HIP(hipGetLastError());
    }
    spmm_csr(HIPMV_FP64, A, X, N, all, AX, N, all);
    phase(b64, s.tol, s.maxit - it32, 1, [&](int p, bool conv, int pit) {
        flag_h[p] = conv ? 0 : 3;
        iters_h[p] = it_fp32[p] + pit;
    });
    for (int p = 0; p < nprob; p++)
        if (active[p]) iters_h[p] = it;
    if (mixed)
        for (float *q : {b32.X, b32.AX, b32.R, b32.AR, b32.P, b32.AP}) {
            BX_HIP(hipStreamSynchronize(stream()));
            BX_HIP(hipFree(q));
        }
    download_sync(lam_h.data(), lam, lam_h.size() * 8);
    download_sync(res_h.data(), res2, res_h.size() * 8);
    size_t o = 0, xo = 0;
    std::vector<double> col;
    for (int p = 0; p < nprob; p++) {
        for (int j = 0; j < kw[p]; j++) {
            lambda[o + j] = lam_h[p * kK + j];
            resnorm[o + j] = sqrt(fabs(res_h[p * kK + j]));
        }
        o += kw[p];
        iterations[p] = iters_h[p];
        flags[p] = flag_h[p];
        if (Xout) {
            for (int j = 0; j < kw[p]; j++)
                download_sync(Xout + xo + (size_t)j * pn[p], X + (size_t)j * N + off[p],
                              (size_t)pn[p] * 8);
            xo += (size_t)pn[p] * kw[p];
        }
    }
    BX_HIP(hipStreamSynchronize(stream()));
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (void *b : {(void *)X, (void *)AX, (void *)R, (void *)AR, (void *)P, (void *)AP,
                    (void *)dv, (void *)d_off, (void *)d_sbase, (void *)d_n, (void *)d_k,
                    (void *)c_prob, (void *)c_nr, (void *)c_r0, (void *)d_actp, (void *)d_pcb,
                    (void *)d_pcc, (void *)lam, (void *)rb, (void *)GA, (void *)GB, (void *)Cm,
                    (void *)nbe, (void *)part_r, (void *)part_g, (void *)rr_work, (void *)dflags})
        BX_HIP(hipFree(b));
    hip_CSRMatrixDestroy(A);
    BX_HIP(hipFree(d_counts));
    if (st) {
        st->seconds = secs;
        st->iterations = it;
        st->iterations_fp32 = it32;
        st->chunk_rows = chunk_rows;
        st->gram_rows = gram_rows;
        st->graph_captures = captures;
        st->graph_replays = replays;
        st->total_rows = N;
    }
    bool ok = true;
    for (int p = 0; p < nprob; p++) ok = ok && flag_h[p] == 0;
    return ok ? 0 : -1;
}

}  // extern "C"
