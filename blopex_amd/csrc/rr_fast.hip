// Rayleigh-Ritz step of LOBPCG on one 1024-thread workgroup: pencils up to
// 96 x 96 (k <= 32) with both pencil matrices in LDS, and pencils up to
// 192 x 192 (k <= 64: guard columns, wide blocks) with A_n in global memory
// (L2-resident; it is touched only by the tile products, which are parallel)
// and B_n in LDS while it fits (the sequential Cholesky and triangular
// inverse run on B_n), else in global memory too.
//
// Input: the raw Gram blocks of [X R P] (lower block triangle and diagonal
// blocks, as the Gram kernel fills them).  The reference B-orthonormalises R
// and P explicitly before forming the pencil (krylov/lobpcg.c:1027-1063,
// Cholesky 'U' + FortranMatrixUpperInv + MultiVecMat); here the same factors
// are applied to the Gram blocks instead, A_n = S' A S, B_n = S' B S with
// S = diag(I, U_R^{-1}, U_P^{-1}), and the blocks lobpcg_solve assumes
// (XAX = diag(lambda), XBX = RBR = PBP = I, RAR and PAP symmetrised,
// lobpcg.c:1107-1162) are set.  The k smallest eigenpairs of (A_n, B_n)
// replace dsygv itype 1 'V' 'L' (lobpcg.c:66-103):
//   B_n = U'U (block Cholesky: the identity X block is skipped and its
//   rank-k update done on MFMA), W = U^{-1}, C = W' A_n W (MFMA),
//   C = Q T Q' (Householder with C in registers, two barriers per column),
//   eigenvalues of T by multisection on the three-term Sturm recurrence,
//   eigenvectors by inverse iteration (stored L D L' pivots), re-orthogonalised
//   inside eigenvalue clusters (relative gap 1e-3, as LAPACK dstein; 1e-6
//   for the explicit-X pencils of FP32 iterates),
//   back-transformed through Q and W; coef = S Y.
// Failure codes follow the reference: R not positive definite -> r_info and
// abort (lobpcg.c:1031-1040), P not positive definite -> P dropped for this
// iteration (p_info), B_n not positive definite -> gevp_info = n + info.
//
// Every step keeps O(1) work per thread between barriers, reductions use DPP
// lane moves, and the LDS leading dimension (3 mod 32) keeps row, column and
// block sweeps free of bank conflicts.
#include <cmath>

#include "dense.h"
#include "dpp.h"
#include "mfma.h"

namespace bx {

namespace {

constexpr int kT = 1024;     // threads
constexpr int kW = kT / 64;  // waves
constexpr int kN = kFastRRN;
constexpr int kK = kFastRRK;
constexpr int kBigN = kBigRRN;
constexpr int kBigK = kBigRRK;
constexpr double kEps = 2.220446049250313e-16;

// where the two pencil matrices live
constexpr int kBothLds = 0, kALdsB = 1, kBothGlobal = 2;  // kALdsB: A global, B in LDS

// static LDS of the body (the small arrays: eleven of NM doubles, two of KM,
// sc[4] and cpiv[1]), for sizing the dynamic part; rounded up to 16 bytes
constexpr size_t static_lds_bytes(int nm, int km) {
    return ((size_t)(11 * nm + 2 * km + 5) * 8 + 15) / 16 * 16;
}
// leading dimension of the pencil arrays: 3 (mod 32), which spreads rows,
// columns and 8 x 8 blocks of doubles over the LDS banks
__host__ __device__ inline int pencil_ld(int n0) { return n0 + ((3 - n0) & 31); }

__device__ __forceinline__ double wmin(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wmax(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// 1/q: hardware estimate refined by two Newton steps (full precision for
// normal q; callers keep |q| away from zero)
__device__ __forceinline__ double recip(double q) {
    double r = __builtin_amdgcn_rcp(q);
    r = fma(r, fma(-q, r, 1.0), r);
    r = fma(r, fma(-q, r, 1.0), r);
    return r;
}

// D(16x16) = sum_{l0 <= l < l1} A(r0 + i, l) B(l, c0 + j) on FP64 MFMA.  The
// operand functors mask their own out-of-range rows and columns.  Lane
// (li = lane % 16, lk = lane / 16) ends with rows r0 + lk + 4q, column c0 + li.
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
// v if ok else 0; callers load v unconditionally from a clamped index so
// that the loads of a step are issued together
__device__ __forceinline__ double msk(bool ok, double v) { return ok ? v : 0.0; }

template <class FA, class FB>
__device__ __forceinline__ f64x4 tile_mm(FA fa, FB fb, int r0, int c0, int l0, int l1) {
    const int li = threadIdx.x & 15, lk = (threadIdx.x >> 4) & 3;
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int l = l0; l < l1; l += 16) {  // four MFMA steps per batch of loads
        double x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int ll = l + 4 * u + lk;
            const bool ok = ll < l1;
            const int lc = ok ? ll : l1 - 1;
            x[u] = msk(ok, fa(r0 + li, lc));
            y[u] = msk(ok, fb(lc, c0 + li));
        }
#pragma unroll
        for (int u = 0; u < 4; u++) acc = Mfma<double>::mma(x[u], y[u], acc);
    }
    return acc;
}

// number of eigenvalues below x of the symmetric tridiagonal with diagonal d
// and squared off-diagonal e2 (entries of order one): sign changes of the
// Sturm sequence p_i = (d_i - x) p_{i-1} - e2_{i-1} p_{i-2}, rescaled every
// four steps; an exact zero counts as -pivmin p_{i-1} (LAPACK's guard on
// the pivots q_i = p_i / p_{i-1})
__device__ int sturm_count(const double *d, const double *e2, int n, double x) {
    constexpr double kPivmin = 1e-290;
    double pp = 1.0, p = d[0] - x;
    if (p == 0.0) p = -kPivmin;
    int cnt = p < 0.0;
    auto step = [&](double di, double ei) {
        double pn = fma(di - x, p, -ei * pp);
        if (pn == 0.0) pn = -kPivmin * p;
        cnt += (pn < 0.0) != (p < 0.0);
        pp = p;
        p = pn;
    };
    int i = 1;
    for (; i + 3 < n; i += 4) {  // four steps per batch of loads, then rescale
        const double d0 = d[i], d1 = d[i + 1], d2 = d[i + 2], d3 = d[i + 3];
        const double f0 = e2[i - 1], f1 = e2[i], f2 = e2[i + 1], f3 = e2[i + 2];
        step(d0, f0);
        step(d1, f1);
        step(d2, f2);
        step(d3, f3);
        int ex;
        (void)frexp(fabs(p) > fabs(pp) ? p : pp, &ex);
        p = ldexp(p, -ex);
        pp = ldexp(pp, -ex);
    }
    for (; i < n; i++) step(d[i], e2[i - 1]);
    return cnt;
}

// Householder tridiagonalisation of the n x n matrix C (LDS, leading
// dimension L), held in registers: waves 0..HW-1 keep C, lane (hc = lane % 16,
// hr = lane / 16) of wave w owning rows w + HW (hr + 4 q), q < HQ, and
// columns hc + 16 b, b < HB (4 HW HQ >= n, 16 HB >= n), so a row's columns lie
// in 16 lanes of one wave.  Step j: x = C[j+1:, j], kept in LDS by absolute
// row with zeros above (so every lane reads its rows and columns at fixed
// addresses); H_j = I - tau v v' with v = x - alpha e_{j+1}; p = tau C v (row
// sums in the wave), K = tau v'p / 2, w = p - K v, C -= v w' + w v'.  v goes
// to column j of C in LDS, the next column and the next diagonal entry to
// LDS.  Rows and columns up to j have v = w = 0 and do not change.  Every
// thread of the workgroup calls this (two barriers per column).
template <int NM, int HW, int HQ, int HB>
__device__ __forceinline__ void householder(int n, int L, double *Am, double *col0, double *col1,
                                            double *pvec, double *dT, double *eT, double *tauT) {
    constexpr int XR = (NM + 63) / 64;  // entries of a column vector per lane
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool hw = wave < HW;
    const int hc = lane & 15, hr = lane >> 4;
    int hrow[HQ], hcol[HB];
#pragma unroll
    for (int q = 0; q < HQ; q++) hrow[q] = wave + HW * (hr + 4 * q);
#pragma unroll
    for (int b = 0; b < HB; b++) hcol[b] = hc + 16 * b;
    double M[HQ][HB];
#pragma unroll
    for (int q = 0; q < HQ; q++)
#pragma unroll
        for (int b = 0; b < HB; b++)
            M[q][b] = (hw && hrow[q] < n && hcol[b] < n) ? Am[hrow[q] + hcol[b] * L] : 0.0;
    // the owner lanes of column c (uniform) pick its entries out of their registers
    auto column = [&](int c, double (&cv)[HQ]) {
        const int bs = __builtin_amdgcn_readfirstlane(c >> 4);
#pragma unroll
        for (int q = 0; q < HQ; q++) cv[q] = M[q][HB - 1];
#pragma unroll
        for (int b = 0; b + 1 < HB; b++)
            if (bs == b) {
#pragma unroll
                for (int q = 0; q < HQ; q++) cv[q] = M[q][b];
            }
    };
    for (int i = tid; i < NM; i += kT) {
        col0[i] = i >= 1 && i < n ? Am[i] : 0.0;
        col1[i] = 0.0;
    }
    if (tid == 0) dT[0] = Am[0];
    __syncthreads();
    for (int j = 0; j + 2 < n; j++) {
        const double *cur = (j & 1) ? col1 : col0;
        double *nxt = (j & 1) ? col0 : col1;
        const int o = j + 1;
        double vr[HQ], vc[HB], pr[HQ], tau = 0.0, alpha = 0.0, xs[XR];
#pragma unroll
        for (int u = 0; u < XR; u++) xs[u] = 0.0;
        if (hw) {
            double ss = 0.0;
#pragma unroll
            for (int u = 0; u < XR; u++) {
                xs[u] = lane + 64 * u < NM ? cur[lane + 64 * u] : 0.0;
                ss = u == 0 ? xs[0] * xs[0] : ss + xs[u] * xs[u];
            }
            const double s = wave_sum(ss);
            const double x0 = cur[o];
            alpha = s > 0.0 ? -copysign(sqrt(s), x0) : 0.0;
            const double v0 = x0 - alpha;
            const double vtv = s - x0 * x0 + v0 * v0;
            tau = (s > 0.0 && vtv > 0.0) ? 2.0 / vtv : 0.0;
#pragma unroll
            for (int q = 0; q < HQ; q++) vr[q] = cur[hrow[q]] - (hrow[q] == o ? alpha : 0.0);
#pragma unroll
            for (int b = 0; b < HB; b++) vc[b] = cur[hcol[b]] - (hcol[b] == o ? alpha : 0.0);
#pragma unroll
            for (int q = 0; q < HQ; q++) {
                double acc = 0.0;
#pragma unroll
                for (int b = 0; b < HB; b++) acc = fma(M[q][b], vc[b], acc);
                pr[q] = hrow[q] < o ? 0.0 : tau * group_sum16(acc);
                if (hc == 0) pvec[hrow[q]] = pr[q];
            }
            if (wave == 0) {
                if (lane == 0) {
                    eT[j] = tau != 0.0 ? alpha : x0;
                    tauT[j] = tau;
                    nxt[j] = 0.0;  // the next column is zero down to row j+1
                    nxt[o] = 0.0;
                }
                // v into column j of Am (rows j+1..n-1; zeros above are ignored)
#pragma unroll
                for (int u = 0; u < XR; u++)
                    if (lane + 64 * u < n)
                        Am[lane + 64 * u + (size_t)j * L] =
                            xs[u] - (lane + 64 * u == o ? alpha : 0.0);
            }
        }
        __syncthreads();
        if (hw) {
            double vp = 0.0;
#pragma unroll
            for (int u = 0; u < XR; u++) {
                const double pu = lane + 64 * u < NM ? pvec[lane + 64 * u] : 0.0;
                const double t = (xs[u] - (lane + 64 * u == o ? alpha : 0.0)) * pu;
                vp = u == 0 ? t : vp + t;
            }
            const double K = 0.5 * tau * wave_sum(vp);
            double wr[HQ], wc[HB];
#pragma unroll
            for (int q = 0; q < HQ; q++) wr[q] = pr[q] - K * vr[q];
#pragma unroll
            for (int b = 0; b < HB; b++) wc[b] = pvec[hcol[b]] - K * vc[b];
#pragma unroll
            for (int q = 0; q < HQ; q++)
#pragma unroll
                for (int b = 0; b < HB; b++)
                    M[q][b] = fma(-vr[q], wc[b], fma(-wr[q], vc[b], M[q][b]));
            // column o (rows > o) for the next step, and its diagonal entry
            double cv[HQ];
            column(o, cv);
            if (hc == (o & 15)) {
#pragma unroll
                for (int q = 0; q < HQ; q++) {
                    if (hrow[q] > o && hrow[q] < n) nxt[hrow[q]] = cv[q];
                    if (hrow[q] == o) dT[o] = cv[q];
                }
            }
        }
        __syncthreads();
    }
    // the last diagonal entry and subdiagonal entry
    if (hw && n >= 2) {
        double cv[HQ];
        column(n - 1, cv);
        if (hc == ((n - 1) & 15)) {
#pragma unroll
            for (int q = 0; q < HQ; q++)
                if (hrow[q] == n - 1) dT[n - 1] = cv[q];
        }
    }
    if (tid == 0 && n >= 2) {
        eT[n - 2] = (n & 1 ? col1 : col0)[n - 1];
        tauT[n - 2] = 0.0;
    }
    __syncthreads();
}

// NM, KM: largest pencil order and block width; MODE: where A_n and B_n live
// (gA, gB: global arrays of pencil_ld(n0) x max(n0, 2k) doubles when not in LDS)
template <int NM, int KM, int MODE>
__device__ void rr_tri_body(RR2Args r, double *gA, double *gB) {
    extern __shared__ double lds[];
    __shared__ double dT[NM], eT[NM], dS[NM], e2S[NM], tauT[NM], invu[NM], sdg[NM];
    __shared__ double pvec[NM], pv[NM], col0[NM], col1[NM];
    __shared__ double lamk[KM], proj[KM];
    __shared__ double sc[4], cpiv[1];
    // tile slots per wave of the MFMA phases (the largest tile counts at NM, KM)
    constexpr int tNm = (NM + 15) / 16, tKm = (KM + 15) / 16, tXRm = (2 * KM + 15) / 16;
    constexpr int S_RF = (tXRm * tKm + tNm * tKm + 3 * tKm * tKm + kW - 1) / kW;
    constexpr int S_LF = (4 * tKm * tKm + kW - 1) / kW;
    constexpr int S_SC = (4 * tKm * tKm + kW - 1) / kW;
    constexpr int S_WX = (2 * tKm * tKm + kW - 1) / kW;
    constexpr int S_C1 = (tNm * tNm + kW - 1) / kW;
    constexpr int S_C2 = (tNm * (tNm + 1) / 2 + kW - 1) / kW;
    constexpr int S_Y = (tNm * tKm + kW - 1) / kW;
    constexpr int S_CO = (2 * tKm * tKm + kW - 1) / kW;
    if (r.st->abort) return;
    const int tid = threadIdx.x, tx = tid & 31, ty = tid >> 5;
    const int wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int k = r.k, a = r.a, pq = r.p, n0 = k + a + pq;
    const int L = pencil_ld(n0);
    double *Am = MODE == kBothLds ? lds : gA;
    double *Bm = MODE == kBothLds ? lds + (size_t)L * n0 : (MODE == kALdsB ? lds : gB);
    double *WRg = r.work, *WPg = WRg + (size_t)a * a, *Wg = WPg + (size_t)pq * pq;
    const int oR = k, oP = k + a;
    auto blk = [&](int i) { return i < k ? 0 : (i < oP ? 1 : 2); };

    // ---- raw pencil into LDS (lower block triangle mirrored) ----
    {
        const double *ga = r.ga, *gb = r.gb;
        const size_t lda = (size_t)r.lda;
        for (int j = ty; j < n0; j += 32)
            for (int i = tx; i < n0; i += 32) {
                const int bi = blk(i), bj = blk(j);
                double va, vb;
                if (bi == 0 && bj == 0) {
                    if (r.explicit_x) {
                        va = 0.5 * (ga[i + j * lda] + ga[j + i * lda]);
                        vb = 0.5 * (gb[i + j * lda] + gb[j + i * lda]);
                    } else {
                        va = i == j ? r.lam_in[i] : 0.0;
                        vb = i == j ? 1.0 : 0.0;
                    }
                } else if (bi < bj) {
                    va = ga[j + i * lda];
                    vb = gb[j + i * lda];
                } else {
                    va = ga[i + j * lda];
                    vb = gb[i + j * lda];
                }
                Am[i + j * L] = va;
                Bm[i + j * L] = vb;
            }
    }
    __syncthreads();

    // ---- Cholesky of the R and P Gram blocks (in place, side by side) ----
    // right-looking, pivots divided out of the trailing update and the rows
    // scaled afterwards; a pivot that is not positive fails as in dpotrf
    int infoR = 0, infoP = 0;
    const int mRP = a > pq ? a : pq;
    for (int j = 0; j < mRP; j++) {
        if (j < a && !infoR) {
            const double d = Bm[(oR + j) * (L + 1)];
            if (!(d > 0.0) || !isfinite(d)) {
                infoR = j + 1;
            } else {
                for (int c = j + 1 + ty; c < a; c += 32)
                    for (int i = j + 1 + tx; i <= c; i += 32)
                        Bm[(oR + i) + (oR + c) * L] -=
                            Bm[(oR + j) + (oR + i) * L] * (Bm[(oR + j) + (oR + c) * L] * (1.0 / d));
            }
        }
        if (j < pq && !infoP) {
            const double d = Bm[(oP + j) * (L + 1)];
            if (!(d > 0.0) || !isfinite(d)) {
                infoP = j + 1;
            } else {
                for (int c = j + 1 + ty; c < pq; c += 32)
                    for (int i = j + 1 + tx; i <= c; i += 32)
                        Bm[(oP + i) + (oP + c) * L] -=
                            Bm[(oP + j) + (oP + i) * L] * (Bm[(oP + j) + (oP + c) * L] * (1.0 / d));
            }
        }
        __syncthreads();
    }
    if (infoR) {
        if (tid == 0) {
            r.st->r_info = infoR;
            r.st->abort = 1;
        }
        return;
    }
    int p = infoP ? 0 : pq;
    int n = k + a + p;
    if (tid == 0) r.st->p_info = infoP;
    for (int j = tid; j < a + p; j += kT) sdg[oR + j] = sqrt(Bm[(oR + j) * (L + 1)]);
    __syncthreads();
    // U rows scaled (strict upper); the lower triangle with the diagonal
    // becomes the accumulator of the inverse
    for (int c = ty; c < mRP; c += 32)
        for (int i = tx; i < mRP; i += 32) {
            if (c < a && i < a) {
                double &e = Bm[(oR + i) + (oR + c) * L];
                e = i < c ? e / sdg[oR + i] : 0.0;
            }
            if (c < p && i < p) {
                double &e = Bm[(oP + i) + (oP + c) * L];
                e = i < c ? e / sdg[oP + i] : 0.0;
            }
        }
    for (int j = tid; j < a + p; j += kT) invu[oR + j] = 1.0 / sdg[oR + j];
    __syncthreads();

    // ---- W_R = U_R^{-1}, W_P = U_P^{-1}: row-wise back substitution ----
    // acc(i, c) (i <= c, stored at (c, i)) collects -sum_{l>i} U(i,l) W(l,c);
    // when row j is final it is fanned out to the rows above it
    for (int j = mRP - 1; j >= 0; j--) {
        for (int c = j + ty; c < mRP; c += 32)
            for (int i = tx; i < j; i += 32) {
                if (j < a && c < a)
                    Bm[(oR + c) + (oR + i) * L] -=
                        Bm[(oR + i) + (oR + j) * L] *
                        ((c == j ? 1.0 : 0.0) + Bm[(oR + c) + (oR + j) * L]) * invu[oR + j];
                if (j < p && c < p)
                    Bm[(oP + c) + (oP + i) * L] -=
                        Bm[(oP + i) + (oP + j) * L] *
                        ((c == j ? 1.0 : 0.0) + Bm[(oP + c) + (oP + j) * L]) * invu[oP + j];
            }
        __syncthreads();
    }
    for (int c = ty; c < mRP; c += 32)
        for (int j = tx; j < mRP; j += 32) {
            if (c < a && j < a) {
                double w = 0.0;
                if (j <= c) {
                    w = ((j == c ? 1.0 : 0.0) + Bm[(oR + c) + (oR + j) * L]) * invu[oR + j];
                    Bm[(oR + j) + (oR + c) * L] = w;
                }
                WRg[j + c * a] = w;
            }
            if (c < p && j < p) {
                double w = 0.0;
                if (j <= c) {
                    w = ((j == c ? 1.0 : 0.0) + Bm[(oP + c) + (oP + j) * L]) * invu[oP + j];
                    Bm[(oP + j) + (oP + c) * L] = w;
                }
                WPg[j + c * p] = w;
            }
        }
    __syncthreads();

    const int tk = (k + 15) / 16, ta = (a + 15) / 16;
    int tp = (p + 15) / 16;
    // ---- right factor: T[:, R] = M[:, R] W_R, T[:, P] = M[:, P] W_P ----
    // A: rows X,R for the R columns, all rows for the P columns;
    // B: X rows for R and P columns, R rows for P columns
    {
        const int tXR = (oP + 15) / 16, tN = (n + 15) / 16;
        const int j1 = tXR * ta, j2 = j1 + tN * tp, j3 = j2 + tk * ta, j4 = j3 + tk * tp,
                  j5 = j4 + ta * tp;
        f64x4 acc[S_RF];
        int jr0[S_RF], jrend[S_RF], jc0[S_RF], jcend[S_RF], jmat[S_RF];
#pragma unroll
        for (int s = 0; s < S_RF; s++) {
            const int t = wave + s * kW;
            jmat[s] = -1;
            if (t >= j5) continue;
            int mat, r0, rend, c0, cend, wo;
            if (t < j1) {
                mat = 0; r0 = 16 * (t / ta); rend = oP; wo = oR; c0 = oR + 16 * (t % ta); cend = oP;
            } else if (t < j2) {
                const int u = t - j1;
                mat = 0; r0 = 16 * (u / tp); rend = n; wo = oP; c0 = oP + 16 * (u % tp); cend = n;
            } else if (t < j3) {
                const int u = t - j2;
                mat = 1; r0 = 16 * (u / ta); rend = k; wo = oR; c0 = oR + 16 * (u % ta); cend = oP;
            } else if (t < j4) {
                const int u = t - j3;
                mat = 1; r0 = 16 * (u / tp); rend = k; wo = oP; c0 = oP + 16 * (u % tp); cend = n;
            } else {
                const int u = t - j4;
                mat = 1; r0 = oR + 16 * (u / tp); rend = oP; wo = oP; c0 = oP + 16 * (u % tp); cend = n;
            }
            const double *M = mat ? Bm : Am;
            const int lend = c0 + 16 < cend ? c0 + 16 : cend;
            acc[s] = tile_mm(
                [&](int row, int l) { return msk(row < rend, M[imin(row, rend - 1) + l * L]); },
                [&](int l, int col) { return msk(col < cend && l <= col, Bm[l + imin(col, cend - 1) * L]); },
                r0, c0, wo, lend);
            jmat[s] = mat; jr0[s] = r0; jrend[s] = rend; jc0[s] = c0; jcend[s] = cend;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < S_RF; s++) {
            if (jmat[s] < 0) continue;
            double *M = jmat[s] ? Bm : Am;
            const int col = jc0[s] + li;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = jr0[s] + lk + 4 * q;
                if (row < jrend[s] && col < jcend[s]) M[row + col * L] = acc[s][q];
            }
        }
        __syncthreads();
    }
    // ---- left factor: A_n[R, R u P] = W_R' T, A_n[P, P] = W_P' T, B_n[R, P] = W_R' T_B ----
    {
        const int j1 = ta * (ta + tp), j2 = j1 + tp * tp, j3 = j2 + ta * tp;
        f64x4 accs[S_LF];
        int mats[S_LF], r0s[S_LF], rends[S_LF], c0s[S_LF], cends[S_LF];
#pragma unroll
        for (int s = 0; s < S_LF; s++) {
        const int t = wave + s * kW;
        f64x4 acc;
        int mat = -1, r0 = 0, rend = 0, c0 = 0, cend = 0;
        if (t < j3) {
            int wo;
            if (t < j1) {
                const int rt = t / (ta + tp), ct = t % (ta + tp);
                mat = 0; wo = oR; r0 = oR + 16 * rt; rend = oP;
                if (ct < ta) { c0 = oR + 16 * ct; cend = oP; } else { c0 = oP + 16 * (ct - ta); cend = n; }
            } else if (t < j2) {
                const int u = t - j1;
                mat = 0; wo = oP; r0 = oP + 16 * (u / tp); rend = n; c0 = oP + 16 * (u % tp); cend = n;
            } else {
                const int u = t - j2;
                mat = 1; wo = oR; r0 = oR + 16 * (u / tp); rend = oP; c0 = oP + 16 * (u % tp); cend = n;
            }
            const double *M = mat ? Bm : Am;
            const int lend = r0 + 16 < rend ? r0 + 16 : rend;
            acc = tile_mm(
                [&](int row, int l) { return msk(row < rend && l <= row, Bm[l + imin(row, rend - 1) * L]); },
                [&](int l, int col) { return msk(col < cend, M[l + imin(col, cend - 1) * L]); },
                r0, c0, wo, lend);
        }
        accs[s] = acc;
        mats[s] = mat;
        r0s[s] = r0;
        rends[s] = rend;
        c0s[s] = c0;
        cends[s] = cend;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < S_LF; s++) {
            if (mats[s] < 0) continue;
            double *M = mats[s] ? Bm : Am;
            const int col = c0s[s] + li;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = r0s[s] + lk + 4 * q;
                if (row < rends[s] && col < cends[s]) M[row + col * L] = accs[s][q];
            }
        }
        __syncthreads();
    }
    // ---- pencil blocks the reference sets: symmetric A_n, B_n ----
    for (int j = ty; j < n; j += 32)
        for (int i = tx; i <= j; i += 32) {
            const int bi = blk(i), bj = blk(j);
            if (i == j) {
                if (bi > 0 || !r.explicit_x) Bm[i + j * L] = 1.0;
                continue;
            }
            double va, vb;
            if (bi == bj) {
                va = bi > 0 ? 0.5 * (Am[i + j * L] + Am[j + i * L]) : Am[i + j * L];
                vb = bi > 0 || !r.explicit_x ? 0.0 : Bm[i + j * L];
            } else {
                va = Am[i + j * L];
                vb = Bm[i + j * L];
            }
            Am[i + j * L] = va;
            Am[j + i * L] = va;
            Bm[i + j * L] = vb;
            Bm[j + i * L] = vb;
        }
    __syncthreads();

    // ---- B_n = U'U ----
    // With retry_without_p, a pencil that is not positive definite is factored
    // again without P (the [X R] block is kept in global scratch for that)
    const int j0 = r.explicit_x ? 0 : k;  // the identity X block needs no steps
    double *Sg = Wg + (size_t)n0 * n0;
    const bool cond_on = r.cond_limit > 0.0;
    const bool may_retry = (r.retry_without_p || cond_on) && p > 0;
    if (may_retry) {
        for (int c = ty; c < oP; c += 32)
            for (int i = tx; i < oP; i += 32) Sg[i + (size_t)c * oP] = Bm[i + c * L];
    }
    int infoB = 0;
    for (int attempt = 0; attempt < 2; attempt++) {
        const int m2 = n - j0, t2 = (m2 + 15) / 16;
        if (j0 > 0 && m2 > 0) {
            // Schur complement B[RP, RP] - E'E, E = B[X, RP]: tiles dealt over the waves
            f64x4 acc[S_SC];
#pragma unroll
            for (int s = 0; s < S_SC; s++) {
                const int t = wave + s * kW;
                const bool act = t < t2 * t2;
                const int r0 = j0 + 16 * (act ? t / t2 : 0), c0 = j0 + 16 * (act ? t % t2 : 0);
                if (act)
                    acc[s] = tile_mm([&](int row, int l) { return msk(row < n, Bm[l + imin(row, n - 1) * L]); },
                                     [&](int l, int col) { return msk(col < n, Bm[l + imin(col, n - 1) * L]); },
                                     r0, c0, 0, k);
            }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < S_SC; s++) {
                const int t = wave + s * kW;
                if (t < t2 * t2) {
                    const int r0 = j0 + 16 * (t / t2), c0 = j0 + 16 * (t % t2);
                    const int col = c0 + li;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int row = r0 + lk + 4 * q;
                        if (row < n && col < n) Bm[row + col * L] -= acc[s][q];
                    }
                }
            }
            __syncthreads();
        }
        infoB = 0;
        for (int j = j0; j < n; j++) {
            const double d = Bm[j * (L + 1)];
            if (!(d > 0.0) || !isfinite(d)) {
                infoB = j + 1;
                break;
            }
            const double rd = 1.0 / d;
            for (int c = j + 1 + ty; c < n; c += 32) {
                const double bjc = Bm[j + c * L] * rd;
                for (int i = j + 1 + tx; i <= c; i += 32) Bm[i + c * L] -= Bm[j + i * L] * bjc;
            }
            __syncthreads();
        }
        // condition estimate of B_n (lobpcg.m:846): log10 of the spread of the
        // Cholesky pivots (the identity X block has pivots 1) + 1
        bool cond_trip = false;
        if (!infoB && cond_on) {
            if (wave == 0) {
                double lo = j0 > 0 ? 1.0 : 1e308, hi = j0 > 0 ? 1.0 : 0.0;
                for (int j = j0 + lane; j < n; j += 64) {
                    const double d = Bm[j * (L + 1)];
                    lo = fmin(lo, d);
                    hi = fmax(hi, d);
                }
                lo = wmin(lo);
                hi = wmax(hi);
                if (lane == 0) cpiv[0] = log10(hi / lo) + 1.0;
            }
            __syncthreads();
            const double cest = cpiv[0];
            if (tid == 0) r.st->condest = cest;
            cond_trip = attempt == 0 && p > 0 && cest > r.cond_limit;
        }
        const bool retry_pd = infoB && r.retry_without_p && p > 0 && attempt == 0;
        if (!retry_pd && !cond_trip) break;
        // drop P: restore the [X R] block and factor again
        __syncthreads();
        for (int c = ty; c < oP; c += 32)
            for (int i = tx; i < oP; i += 32) Bm[i + c * L] = Sg[i + (size_t)c * oP];
        p = 0;
        n = oP;
        tp = 0;
        if (tid == 0) r.st->p_info = cond_trip ? -2 : -1;
        __syncthreads();
    }
    if (infoB) {
        if (tid == 0) {
            r.st->gevp_info = n + infoB;
            r.st->abort = 1;
        }
        return;
    }
    const int m2 = n - j0, t2 = (m2 + 15) / 16;
    for (int j = tid; j < n; j += kT) sdg[j] = j < j0 ? 1.0 : sqrt(Bm[j * (L + 1)]);
    __syncthreads();
    for (int c = ty; c < n; c += 32)
        for (int i = tx; i < n; i += 32) {
            if (i < c) {
                if (i >= j0) Bm[i + c * L] /= sdg[i];
            } else if (c >= j0) {
                Bm[i + c * L] = 0.0;  // accumulator of the inverse
            }
        }
    for (int j = tid; j < n; j += kT) invu[j] = 1.0 / sdg[j];
    __syncthreads();

    // ---- W = U^{-1}: trailing block by back substitution, X rows by MFMA ----
    for (int j = n - 1; j >= j0; j--) {
        const double iu = invu[j];
        for (int c = j + ty; c < n; c += 32) {
            const double f = ((c == j ? 1.0 : 0.0) + Bm[c + j * L]) * iu;
            for (int i = j0 + tx; i < j; i += 32) Bm[c + i * L] -= Bm[i + j * L] * f;
        }
        __syncthreads();
    }
    for (int c = j0 + ty; c < n; c += 32)
        for (int j = j0 + tx; j <= c; j += 32)
            Bm[j + c * L] = ((j == c ? 1.0 : 0.0) + Bm[c + j * L]) * invu[j];
    __syncthreads();
    if (j0 > 0 && m2 > 0) {
        // W[X, RP] = -E W2
        f64x4 acc[S_WX];
#pragma unroll
        for (int s = 0; s < S_WX; s++) {
            const int t = wave + s * kW;
            const bool act = t < tk * t2;
            const int r0 = 16 * (act ? t / t2 : 0), c0 = j0 + 16 * (act ? t % t2 : 0);
            if (act) {
                const int lend = c0 + 16 < n ? c0 + 16 : n;
                acc[s] = tile_mm([&](int row, int l) { return msk(row < k, Bm[imin(row, k - 1) + l * L]); },
                                 [&](int l, int col) { return msk(col < n && l <= col, Bm[l + imin(col, n - 1) * L]); },
                                 r0, c0, j0, lend);
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < S_WX; s++) {
            const int t = wave + s * kW;
            if (t < tk * t2) {
                const int r0 = 16 * (t / t2), c0 = j0 + 16 * (t % t2);
                const int col = c0 + li;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int row = r0 + lk + 4 * q;
                    if (row < k && col < n) Bm[row + col * L] = -acc[s][q];
                }
            }
        }
        __syncthreads();
    }

    // ---- C = W' A_n W (lower triangle kept); W saved for the back-transform ----
    const int tN = (n + 15) / 16;
    {
        f64x4 acc[S_C1];
#pragma unroll
        for (int s = 0; s < S_C1; s++) {
            const int t = wave + s * kW;
            if (t < tN * tN) {
                const int r0 = 16 * (t / tN), c0 = 16 * (t % tN);
                const int lend = c0 + 16 < n ? c0 + 16 : n;
                acc[s] = tile_mm(
                    [&](int row, int l) { return msk(row < n, Am[imin(row, n - 1) + l * L]); },
                    [&](int l, int col) { return msk(col < n && l <= col, Bm[l + imin(col, n - 1) * L]); },
                    r0, c0, 0, lend);
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < S_C1; s++) {
            const int t = wave + s * kW;
            if (t < tN * tN) {
                const int r0 = 16 * (t / tN), col = 16 * (t % tN) + li;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int row = r0 + lk + 4 * q;
                    if (row < n && col < n) Am[row + col * L] = acc[s][q];
                }
            }
        }
        for (int c = ty; c < n; c += 32)
            for (int i = tx; i < n; i += 32) Wg[i + (size_t)c * n] = i <= c ? Bm[i + c * L] : 0.0;
        __syncthreads();
    }
    {
        f64x4 acc[S_C2];
        int rr0[S_C2], cc0[S_C2];
#pragma unroll
        for (int s = 0; s < S_C2; s++) {
            int t = wave + s * kW;
            rr0[s] = -1;
            if (t < tN * (tN + 1) / 2) {
                int rt = 0;
                while (t > rt) {
                    t -= rt + 1;
                    rt++;
                }
                const int r0 = 16 * rt, c0 = 16 * t;
                const int lend = r0 + 16 < n ? r0 + 16 : n;
                acc[s] = tile_mm(
                    [&](int row, int l) { return msk(row < n && l <= row, Bm[l + imin(row, n - 1) * L]); },
                    [&](int l, int col) { return msk(col < n, Am[l + imin(col, n - 1) * L]); },
                    r0, c0, 0, lend);
                rr0[s] = r0;
                cc0[s] = c0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < S_C2; s++) {
            if (rr0[s] < 0) continue;
            const int col = cc0[s] + li;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = rr0[s] + lk + 4 * q;
                if (row < n && col < n && row >= col) {
                    Am[row + col * L] = acc[s][q];
                    Am[col + row * L] = acc[s][q];
                }
            }
        }
        __syncthreads();
    }

    // ---- Householder tridiagonalisation of C, held in registers ----
    // pencils up to 48: 4 waves (one per SIMD) with 9 entries per lane; up
    // to 96: 8 waves with 18; up to 128: 16 waves with 16; up to 192: 16
    // waves with 36
    if (n <= 48)
        householder<NM, 4, 3, 3>(n, L, Am, col0, col1, pvec, dT, eT, tauT);
    else if (NM <= 96 || n <= 96)
        householder<NM, 8, 3, 6>(n, L, Am, col0, col1, pvec, dT, eT, tauT);
    else if (NM <= 128 || n <= 128)
        householder<NM, 16, 2, 8>(n, L, Am, col0, col1, pvec, dT, eT, tauT);
    else
        householder<NM, 16, 3, 12>(n, L, Am, col0, col1, pvec, dT, eT, tauT);

    // ---- Gershgorin interval; T scaled by a power of two near its norm ----
    if (wave == 0) {
        double lo = 1e308, hi = -1e308, tn = 0.0;
        for (int i = lane; i < n; i += 64) {
            const double rad = (i > 0 ? fabs(eT[i - 1]) : 0.0) + (i + 1 < n ? fabs(eT[i]) : 0.0);
            lo = fmin(lo, dT[i] - rad);
            hi = fmax(hi, dT[i] + rad);
            tn = fmax(tn, fabs(dT[i]) + rad);
        }
        lo = wmin(lo);
        hi = wmax(hi);
        tn = wmax(tn);
        if (lane == 0) {
            int ex = 0;
            if (tn > 0.0) (void)frexp(tn, &ex);
            sc[2] = tn;
            sc[3] = ldexp(1.0, -ex);  // 1 / scale
            sc[0] = ldexp(lo, -ex) - 4.0 * kEps;
            sc[1] = ldexp(hi, -ex) + 4.0 * kEps;
        }
    }
    __syncthreads();
    {
        const double is = sc[3];
        for (int i = tid; i < n; i += kT) {
            dS[i] = dT[i] * is;
            e2S[i] = i + 1 < n ? (eT[i] * is) * (eT[i] * is) : 0.0;
        }
    }
    __syncthreads();

    // ---- k smallest eigenvalues: multisection, G lanes per index ----
    {
        const int G = k <= 8 ? 32 : (k <= 16 ? 16 : 8);
        const unsigned gmask = G == 32 ? 0xFFFFFFFFu : ((1u << G) - 1u);
        const int idx = tid / G, lg = tid & (G - 1);
        if (idx < k) {
            double lo = sc[0], hi = sc[1];
            // Bracket from the previous Ritz value e: the pencil contains the
            // previous Ritz vectors, so its idx-th eigenvalue is about e or
            // below.  Lane 0 tries e (plus a margin) as an upper bound, lane
            // g > 0 tries e - |e| 2^(2 - 6g) as a lower bound; a bound is kept
            // only where its Sturm count confirms it.
            const double e = r.lam_in[idx] * sc[3];
            if (e > lo && e < hi) {  // false for NaN
                auto cand = [&](int g) {
                    return g == 0 ? e + 16.0 * kEps * fabs(e) + 1e-3 * kEps
                                  : e - fabs(e) * ldexp(1.0, 2 - 6 * g);
                };
                const bool above = sturm_count(dS, e2S, n, cand(lg)) > idx;
                const unsigned long long bal = __ballot(above);
                const unsigned seg = (unsigned)(bal >> (lane & ~(G - 1))) & gmask;
                if (seg & 1u) hi = fmin(hi, cand(0));
                const unsigned low = seg >> 1;  // lanes 1..G-1, points ascending
                const int best = low ? __builtin_ctz(low) : G - 1;  // last lane not above
                if (best >= 1) lo = fmax(lo, cand(best));
                if (!(lo < hi)) {  // cannot happen with exact counts; start over
                    lo = sc[0];
                    hi = sc[1];
                }
            }
            for (int round = 0; round < 48; round++) {
                const double w = hi - lo;
                if (!(w > 2.0 * kEps * fmax(fabs(lo), fabs(hi)) + 1e-3 * kEps)) break;
                const double step = w / (double)(G + 1);
                const bool above = sturm_count(dS, e2S, n, lo + step * (lg + 1)) > idx;
                const unsigned long long bal = __ballot(above);
                const unsigned seg = (unsigned)(bal >> (lane & ~(G - 1))) & gmask;
                if (seg) {
                    const int t = __builtin_ctz(seg);
                    const double nhi = lo + step * (t + 1);
                    if (t > 0) lo = lo + step * t;
                    hi = nhi;
                } else {
                    lo = lo + step * G;
                }
            }
            if (lg == 0) lamk[idx] = 0.5 * (lo + hi) / sc[3];
        }
    }
    __syncthreads();

    // ---- inverse iteration, one thread per eigenvalue (Z, pivots in Bm) ----
    double *Z = Bm, *Dq = Bm + (size_t)L * k;
    const double tnorm = sc[2];
    if (tid < k) {
        // loads go in batches of four ahead of the recurrences that use them
        const int c = tid;
        double *z = Z + (size_t)c * L, *rq = Dq + (size_t)c * L;
        const double lam = lamk[c];
        const double tiny = kEps * tnorm + 1e-280;
        auto guard = [&](double q) { return fabs(q) < tiny ? (q < 0.0 ? -tiny : tiny) : q; };
        auto start = [&](int j) {  // pseudo-random start vector in [0.5, 1.5)
            const unsigned h = (unsigned)(j + 1) * 2654435761u ^ (unsigned)(c + 1) * 40503u;
            return 0.5 + (double)(h % 1024u) / 1024.0;
        };
        // T - lam I = L D L' (reciprocal pivots kept), forward solve of the start
        double q = guard(dT[0] - lam), rqp = recip(q), y = start(0);
        rq[0] = rqp;
        z[0] = y;
        auto fstep = [&](int j, double dj, double ej) {
            const double l = ej * rqp;
            q = guard(dj - lam - l * ej);
            rqp = recip(q);
            rq[j] = rqp;
            y = start(j) - l * y;
            z[j] = y;
        };
        int j = 1;
        for (; j + 3 < n; j += 4) {
            const double d0 = dT[j], d1 = dT[j + 1], d2 = dT[j + 2], d3 = dT[j + 3];
            const double e0 = eT[j - 1], e1 = eT[j], e2 = eT[j + 1], e3 = eT[j + 2];
            fstep(j, d0, e0);
            fstep(j + 1, d1, e1);
            fstep(j + 2, d2, e2);
            fstep(j + 3, d3, e3);
        }
        for (; j < n; j++) fstep(j, dT[j], eT[j - 1]);
        for (int it = 0; it < 2; it++) {
            if (it == 1) {  // forward solve with the stored pivots
                double yy = z[0];
                int jj = 1;
                for (; jj + 3 < n; jj += 4) {
                    double zz[4], ee[4], rr[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        zz[u] = z[jj + u];
                        ee[u] = eT[jj + u - 1];
                        rr[u] = rq[jj + u - 1];
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        yy = zz[u] - ee[u] * rr[u] * yy;
                        z[jj + u] = yy;
                    }
                }
                for (; jj < n; jj++) {
                    yy = z[jj] - eT[jj - 1] * rq[jj - 1] * yy;
                    z[jj] = yy;
                }
            }
            double x = z[n - 1] * rq[n - 1], mx = fabs(x);
            z[n - 1] = x;
            int jb = n - 2;
            for (; jb >= 3; jb -= 4) {
                double zz[4], ee[4], rr[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    zz[u] = z[jb - u];
                    ee[u] = eT[jb - u];
                    rr[u] = rq[jb - u];
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    x = (zz[u] - ee[u] * x) * rr[u];
                    z[jb - u] = x;
                    mx = fmax(mx, fabs(x));
                }
            }
            for (; jb >= 0; jb--) {
                x = (z[jb] - eT[jb] * x) * rq[jb];
                z[jb] = x;
                mx = fmax(mx, fabs(x));
            }
            const double inv = mx > 0.0 ? 1.0 / mx : 1.0;
            double s2 = 0.0;
            for (int i = 0; i < n; i++) {
                const double v = z[i] * inv;
                z[i] = v;
                s2 += v * v;
            }
            const double f = s2 > 0.0 ? 1.0 / sqrt(s2) : 0.0;
            for (int i = 0; i < n; i++) z[i] *= f;
        }
    }
    __syncthreads();

    // ---- Gram-Schmidt inside clusters of close eigenvalues ----
    // Each vector against the earlier (unit) vectors of its cluster; the
    // dot products include its own, so |z - P z|^2 = |z|^2 - sum proj^2 is
    // known without another reduction: a second pass runs only when the
    // first removed more than three quarters of |z|^2 (twice is enough),
    // and the last pass also normalises.
    {
        // With explicit X blocks (FP32 iterates) the next Gram measures X'BX
        // anyway, so the Ritz vectors need orthogonality well below FP32
        // rounding only: eps |T| / gap <= 2e-10 for the clusters below.
        const double ortol = (r.explicit_x ? 1e-6 : 1e-3) * tnorm;
        for (int c = 1; c < k; c++) {
            int c0 = c;
            while (c0 > 0 && lamk[c0] - lamk[c0 - 1] <= ortol) c0--;
            if (c0 == c) continue;
            double *zc = Z + (size_t)c * L;
            for (int pass = 0; pass < 2; pass++) {
                {
                    const int cp = c0 + (tid >> 4), part = tid & 15;  // cp == c: |z|^2
                    double s = 0.0;
                    if (cp <= c)
                        for (int j = part; j < n; j += 16) s += Z[(size_t)cp * L + j] * zc[j];
                    s = group_sum16(s);
                    if (cp <= c && part == 0) proj[cp] = s;
                }
                __syncthreads();
                double rem = proj[c];
                for (int cp = c0; cp < c; cp++) rem -= proj[cp] * proj[cp];
                const bool again = pass == 0 && !(rem > 0.25 * proj[c]);  // uniform
                const double f = again ? 1.0 : (rem > 0.0 ? 1.0 / sqrt(rem) : 0.0);
                for (int j = tid; j < n; j += kT) {
                    double v = zc[j];
                    for (int cp = c0; cp < c; cp++) v -= proj[cp] * Z[(size_t)cp * L + j];
                    zc[j] = v * f;
                }
                __syncthreads();
                if (!again) break;
            }
        }
    }

    // ---- Z <- Q Z: reflections n-3 .. 0, LPC lanes per column ----
    // the lane keeps rows l + LPC q of its column in registers; v of
    // reflection j is column j of Am below the diagonal (zero above)
    {
        constexpr int LPC = KM <= 32 ? 32 : 16, RPL = (NM + LPC - 1) / LPC;
        if (tid / LPC < k) {
            const int c = tid / LPC, l = tid % LPC;
            double *z = Z + (size_t)c * L;
            double zr[RPL];
#pragma unroll
            for (int q = 0; q < RPL; q++) {
                const int i = l + LPC * q;
                zr[q] = i < n ? z[i] : 0.0;
            }
#pragma unroll 2
            for (int j = n - 3; j >= 0; j--) {
                const double tau = tauT[j];  // tau = 0 leaves z unchanged
                double vv[RPL], s = 0.0;
#pragma unroll
                for (int q = 0; q < RPL; q++) {
                    const int i = l + LPC * q, ic = i < n ? i : n - 1;
                    const double v = Am[ic + (size_t)j * L];
                    vv[q] = i > j && i < n ? v : 0.0;
                    s = fma(vv[q], zr[q], s);
                }
                s = tau * (LPC == 32 ? group_sum32(s) : group_sum16(s));
#pragma unroll
                for (int q = 0; q < RPL; q++) zr[q] = fma(-s, vv[q], zr[q]);
            }
#pragma unroll
            for (int q = 0; q < RPL; q++) {
                const int i = l + LPC * q;
                if (i < n) z[i] = zr[q];
            }
        }
    }
    __syncthreads();

    // ---- Y = W Z (into Am), then coef = S Y ----
    {
        const int tKc = (k + 15) / 16;
        f64x4 acc[S_Y];
#pragma unroll
        for (int s = 0; s < S_Y; s++) {
            const int t = wave + s * kW;
            const bool act = t < tN * tKc;
            const int r0 = 16 * (act ? t / tKc : 0), c0 = 16 * (act ? t % tKc : 0);
            if (act)
                acc[s] = tile_mm(
                    [&](int row, int l) { return msk(row < n && l >= row, Wg[imin(row, n - 1) + (size_t)l * n]); },
                    [&](int l, int col) { return msk(col < k, Z[(size_t)imin(col, k - 1) * L + l]); },
                    r0, c0, r0, n);
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < S_Y; s++) {
            const int t = wave + s * kW;
            if (t < tN * tKc) {
                const int r0 = 16 * (t / tKc), c0 = 16 * (t % tKc);
                const int col = c0 + li;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int row = r0 + lk + 4 * q;
                    if (row < n && col < k) Am[row + col * L] = acc[s][q];
                }
            }
        }
        __syncthreads();
        // R and P rows through W_R, W_P (tiles dealt over the waves), X rows copied
        const double *Y = Am;
        const int jr = ta * tKc, jp = jr + tp * tKc;
#pragma unroll
        for (int s = 0; s < S_CO; s++) {
        const int t = wave + s * kW;
        if (t < jp) {
            const bool isR = t < jr;
            const int u = isR ? t : t - jr;
            const int o = isR ? oR : oP, m = isR ? a : p;
            const double *Wb = isR ? WRg : WPg;
            const int rr = 16 * (u / tKc), cc = 16 * (u % tKc);
            const f64x4 ac = tile_mm(
                [&](int row, int l) { return msk(row < m, Wb[imin(row, m - 1) + (size_t)l * m]); },
                [&](int l, int col) { return msk(col < k, Y[(o + l) + imin(col, k - 1) * L]); },
                rr, cc, rr, m);
            const int col = cc + li;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = rr + lk + 4 * q;
                if (row < m && col < k) r.coef[(o + row) + (size_t)col * r.ldc] = ac[q];
            }
        }
        }
        for (int c = ty; c < k; c += 32) {
            for (int i = tx; i < k; i += 32) r.coef[i + (size_t)c * r.ldc] = Y[i + c * L];
            for (int i = n + tx; i < n0; i += 32) r.coef[i + (size_t)c * r.ldc] = 0.0;  // dropped P
        }
        for (int c = tid; c < k; c += kT) r.lam_out[c] = lamk[c];
        if (tid == 0) r.st->gevp_info = 0;
    }
}

__global__ void __launch_bounds__(kT) k_rr_tri(RR2Args r) {
    rr_tri_body<kN, kK, kBothLds>(r, nullptr, nullptr);
}

// pencils up to 192 (k, a, p <= 64): A_n in global memory, B_n in LDS
// (MODE kALdsB) or global memory (kBothGlobal)
template <int NM, int MODE>
__global__ void __launch_bounds__(kT) k_rr_tri_big(RR2Args r, double *gA, double *gB) {
    rr_tri_body<NM, kBigK, MODE>(r, gA, gB);
}

// one workgroup per problem of a batch: problem act[blockIdx.x] has blocks of
// width kvec[p] (X and R, and P when with_p); its Gram buffers, Ritz values,
// coefficients, work space and status sit at p times the given strides
__global__ void __launch_bounds__(kT)
    k_rr_tri_batched(RR2Args base, const int *act, const int *nact, const int *kvec,
                     const int *run, int with_p, long s_gram, long s_lam, long s_coef,
                     long s_work) {
    if ((int)blockIdx.x >= *nact) return;
    const int p = act[blockIdx.x];
    if (run && !run[p]) return;
    RR2Args r = base;
    r.ga += (size_t)p * s_gram;
    r.gb += (size_t)p * s_gram;
    r.k = r.a = kvec[p];
    r.p = with_p ? kvec[p] : 0;
    r.lam_in += (size_t)p * s_lam;
    r.lam_out += (size_t)p * s_lam;
    r.coef += (size_t)p * s_coef;
    r.work += (size_t)p * s_work;
    r.st += p;
    rr_tri_body<kN, kK, kBothLds>(r, nullptr, nullptr);
}

}  // namespace

size_t rr_fast_work_words(int n) {
    const int nE = (n + 1) / 2 * 2;
    return (size_t)4 * nE * nE + (size_t)2 * n * n + (size_t)n * n + 64;
}

void rr_fast_batched(const RR2Args &base, const int *act, const int *nact, const int *kvec,
                     const int *run, int nprob, int kmax, bool with_p, long s_gram, long s_lam,
                     long s_coef, long s_work) {
    const int n0 = (with_p ? 3 : 2) * kmax;
    BX_REQUIRE(rr_fast_fits(kmax, kmax, with_p ? kmax : 0), "rr_fast_batched: k %d too large",
               kmax);
    constexpr size_t kMaxLds = ((size_t)kFastRRN + 3) * 2 * kFastRRN * sizeof(double);
    const int L = n0 + ((3 - n0) & 31);
    const int second = n0 > 2 * kmax ? n0 : 2 * kmax;
    // the largest pencil of the batch sets the dynamic LDS of every workgroup
    size_t lds = 0;
    for (int n = 1; n <= n0; n++) {
        const int Ln = n + ((3 - n) & 31);
        const size_t b = (size_t)Ln * (n + (n > second ? n : second)) * sizeof(double);
        if (b > lds) lds = b;
    }
    (void)L;
    BX_REQUIRE(lds <= kMaxLds, "rr_fast_batched: %zu bytes of LDS", lds);
    static bool attr_set = false;
    if (!attr_set) {
        BX_HIP(hipFuncSetAttribute((const void *)k_rr_tri_batched,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds));
        attr_set = true;
    }
    hipLaunchKernelGGL(k_rr_tri_batched, dim3(nprob), dim3(kT), lds, stream(), base, act, nact,
                       kvec, run, (int)with_p, s_gram, s_lam, s_coef, s_work);
    BX_HIP(hipGetLastError());
}

bool rr_big_fits(int k, int a, int p) {
    return k + a + p <= kBigRRN && k <= kBigRRK && a <= kBigRRK && p <= kBigRRK && k >= 1;
}

void rr_big(const RR2Args &r) {
    BX_REQUIRE(rr_big_fits(r.k, r.a, r.p), "rr_big: pencil %d + %d + %d does not fit", r.k, r.a,
               r.p);
    const int n0 = r.k + r.a + r.p, L = pencil_ld(n0);
    const int second = n0 > 2 * r.k ? n0 : 2 * r.k;
    const size_t a_words = (size_t)L * n0, b_words = (size_t)L * second;
    constexpr size_t kLds = 163840;  // LDS of one workgroup on gfx950
    // pencils up to 128 take the instantiation with fewer registers per lane
    const bool mid = n0 <= 128;
    const size_t dyn_max = kLds - static_lds_bytes(mid ? 128 : kBigN, kBigK);
    const bool b_lds = b_words * sizeof(double) <= dyn_max;
    double *g = (double *)scratch(10, (a_words + (b_lds ? 0 : b_words)) * sizeof(double) + 256);
    static bool attr_set = false;
    if (!attr_set) {
        BX_HIP(hipFuncSetAttribute((const void *)k_rr_tri_big<128, kALdsB>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(kLds - static_lds_bytes(128, kBigK))));
        BX_HIP(hipFuncSetAttribute((const void *)k_rr_tri_big<kBigN, kALdsB>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(kLds - static_lds_bytes(kBigN, kBigK))));
        attr_set = true;
    }
    auto kern = mid ? (b_lds ? k_rr_tri_big<128, kALdsB> : k_rr_tri_big<128, kBothGlobal>)
                    : (b_lds ? k_rr_tri_big<kBigN, kALdsB> : k_rr_tri_big<kBigN, kBothGlobal>);
    if (b_lds)
        hipLaunchKernelGGL(kern, dim3(1), dim3(kT), b_words * sizeof(double), stream(), r, g,
                           nullptr);
    else
        hipLaunchKernelGGL(kern, dim3(1), dim3(kT), 0, stream(), r, g, g + a_words);
    BX_HIP(hipGetLastError());
}

bool rr_fast_fits(int k, int a, int p) {
    return k + a + p <= kFastRRN && k <= kFastRRK && a <= kFastRRK && p <= kFastRRK && k >= 1;
}

void rr_fast(const RR2Args &r) {
    BX_REQUIRE(rr_fast_fits(r.k, r.a, r.p), "rr_fast: pencil %d + %d + %d does not fit", r.k, r.a,
               r.p);
    const int n0 = r.k + r.a + r.p, L = n0 + ((3 - n0) & 31);
    const int second = n0 > 2 * r.k ? n0 : 2 * r.k;
    const size_t lds = (size_t)L * (n0 + second) * sizeof(double);
    constexpr size_t kMaxLds = ((size_t)kFastRRN + 3) * 2 * kFastRRN * sizeof(double);
    BX_REQUIRE(lds <= kMaxLds, "rr_fast: %zu bytes of LDS (max %zu)", lds, kMaxLds);
    static bool attr_set = false;
    if (!attr_set) {
        BX_HIP(hipFuncSetAttribute((const void *)k_rr_tri, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kMaxLds));
        attr_set = true;
    }
    hipLaunchKernelGGL(k_rr_tri, dim3(1), dim3(kT), lds, stream(), r);
    BX_HIP(hipGetLastError());
}

}  // namespace bx
