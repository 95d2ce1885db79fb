// FP64 small-dense kernels of the device-resident solver, one workgroup each:
//
//  * chol_inv: Cholesky G = U'U of an m x m Gram block (LAPACK dpotrf 'U',
//    krylov/lobpcg.c:20-39), then U^{-1} in place and the strict lower part
//    cleared (FortranMatrixUpperInv / ClearL, fortran_matrix.c:257-279,
//    1368-1416).  The failing pivot (LAPACK info) goes to a status word.
//  * rr_solve: the generalized Rayleigh-Ritz A u = lambda B u on the nA x nA
//    Gram pencil (dsygv itype 1, 'V', 'L', krylov/lobpcg.c:66-103):
//    B = U'U, C = U^{-T} A U^{-1}, C = Q diag(lambda) Q' by cyclic Jacobi in
//    parallel (round-robin) order, eigenvectors U^{-1} Q, ascending
//    eigenvalues; the first k pairs are kept.  Before solving it assembles
//    the blocks lobpcg_solve sets instead of computing (XAX = diag(lambda),
//    XBX = RBR = PBP = I, RAR and PAP symmetrised; lobpcg.c:1107-1162).
//
// The eigenvector signs and the order inside exactly repeated eigenvalues
// may differ from LAPACK's; the LOBPCG iteration is invariant under both.
#include <cmath>
#include <cstdlib>

#include "dense.h"

namespace bx {

bool g_force_jacobi = false;

namespace {

constexpr int kDT = 1024;  // threads per small-dense workgroup

__device__ __forceinline__ int tid() { return threadIdx.x; }

// block-wide sum (all threads get the result)
__device__ double block_sum(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();
    if ((tid() & 63) == 0) red[tid() >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < kDT / 64; w++) s += red[w];
    __syncthreads();
    return s;
}

// In-place upper Cholesky of the m x m matrix a (leading dim ld); returns the
// LAPACK-style info (0, or the 1-based column whose pivot fails).  A pivot
// fails when it is not positive (dpotrf's test, rel = 0) or, with rel > 0,
// when it keeps no more than rel of the column's own diagonal entry (the
// column is numerically dependent on the ones before it); orig (>= m doubles
// of shared memory) holds those entries.
__device__ int chol_upper(double *a, int ld, int m, int *shared_info, double *orig,
                          double rel = 0.0) {
    if (tid() == 0) *shared_info = 0;
    for (int j = tid(); j < m; j += kDT) orig[j] = a[j + j * ld];
    __syncthreads();
    for (int j = 0; j < m; j++) {
        // a(j,j) already holds a_jj - sum_{i<j} u_ij^2 (right-looking updates)
        const double d = a[j + j * ld];
        if (!(d > rel * orig[j]) || !isfinite(d)) {
            if (tid() == 0) *shared_info = j + 1;
            __syncthreads();
            return *shared_info;
        }
        const double piv = sqrt(d);
        __syncthreads();
        // row j of U
        for (int c = j + 1 + tid(); c < m; c += kDT) a[j + c * ld] /= piv;
        __syncthreads();
        if (tid() == 0) a[j + j * ld] = piv;
        // trailing update of the upper triangle: a(i,c) -= u(j,i) u(j,c), j < i <= c
        const int t = m - j - 1;
        for (int e = tid(); e < t * t; e += kDT) {
            const int i = j + 1 + e % t, c = j + 1 + e / t;
            if (i <= c) a[i + c * ld] -= a[j + i * ld] * a[j + c * ld];
        }
        __syncthreads();
    }
    return 0;
}

// In-place inverse of the upper triangle, strict lower part cleared.  Each
// thread solves one column j of the inverse bottom-up,
//   x_jj = 1/u_jj,  x_ij = -(sum_{l=i+1..j} u_il x_lj) / u_ii,
// and parks x_ij (i < j) in the free strict lower triangle at (j, i) so the
// original strict upper part stays readable until every column is done.
__device__ void upper_inverse(double *u, int ld, int m, double *diag) {
    for (int i = tid(); i < m; i += kDT) diag[i] = u[i + i * ld];
    __syncthreads();
    for (int j = tid(); j < m; j += kDT) {
        const double xjj = 1.0 / diag[j];
        for (int i = j - 1; i >= 0; i--) {
            double v = 0.0;
            for (int l = i + 1; l <= j; l++) {
                const double xl = (l == j) ? xjj : u[j + l * ld];
                v -= u[i + l * ld] * xl;
            }
            u[j + i * ld] = v / diag[i];
        }
    }
    __syncthreads();
    for (int e = tid(); e < m * m; e += kDT) {
        const int i = e % m, j = e / m;
        if (i < j) u[i + j * ld] = u[j + i * ld];
    }
    __syncthreads();
    for (int e = tid(); e < m * m; e += kDT) {
        const int i = e % m, j = e / m;
        if (i > j) u[i + j * ld] = 0.0;
        if (i == j) u[i + j * ld] = 1.0 / diag[i];
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kDT) k_chol_inv(double *g, int ld, int m, DevStatus *st,
                                                   int which) {
    __shared__ int info;
    __shared__ double diag[kMaxRR];
    if (st->abort) return;
    // the initial block and the constraints are checked for numerical rank
    // (exactly dependent columns leave a pivot of a few ulps); R and P use
    // dpotrf's test, as the reference does
    const double rel = which == kCholX ? 4.0 * 2.220446049250313e-16 : 0.0;
    int r = chol_upper(g, ld, m, &info, diag, rel);
    if (r) {
        // a failed factor becomes the identity, so the block products that
        // follow leave their operand unchanged (P then drops out of the basis)
        for (int e = tid(); e < m * m; e += kDT) g[e % m + (e / m) * ld] = (e % m == e / m) ? 1.0 : 0.0;
        if (tid() == 0) {
            if (which == kCholR) {
                st->r_info = r;
                st->abort = 1;
            } else if (which == kCholP) {
                st->p_info = r;
            } else {
                st->x_info = r;
                st->abort = 1;
            }
        }
        return;
    }
    if (tid() == 0 && which == kCholP) st->p_info = 0;
    upper_inverse(g, ld, m, diag);
}

// ---- Rayleigh-Ritz -----------------------------------------------------

__device__ __forceinline__ int rr_player(int pos, int round, int n) {
    return pos == 0 ? 0 : ((pos - 1 + round) % (n - 1)) + 1;
}

// cyclic Jacobi on the symmetric n x n matrix c (ld n, n even), eigenvectors
// accumulated into v (ld n, starts as identity); returns sweeps used
__device__ int jacobi(double *c, double *v, int n, double *cs, double *sn, int *rot,
                      double *red) {
    const int half = n / 2;
    int sweep;
    for (sweep = 0; sweep < 40; sweep++) {
        if (tid() == 0) *rot = 0;
        __syncthreads();
        for (int round = 0; round < n - 1; round++) {
            for (int a = tid(); a < half; a += kDT) {
                int p = rr_player(a, round, n), q = rr_player(n - 1 - a, round, n);
                if (p > q) {
                    int t = p;
                    p = q;
                    q = t;
                }
                const double apq = c[p + q * n], app = c[p + p * n], aqq = c[q + q * n];
                double cc = 1.0, ss = 0.0;
                // rotate unless a_pq is negligible relative to the diagonal
                if (fabs(apq) > 1e-300 && fabs(apq) > 2.2e-16 * 0.5 * sqrt(fabs(app * aqq))) {
                    const double theta = (aqq - app) / (2.0 * apq);
                    const double t = (theta >= 0.0 ? 1.0 : -1.0) /
                                     (fabs(theta) + sqrt(theta * theta + 1.0));
                    cc = 1.0 / sqrt(t * t + 1.0);
                    ss = t * cc;
                    *rot = 1;
                }
                cs[a] = cc;
                sn[a] = ss;
            }
            __syncthreads();
            // A <- J' A J on 2x2 blocks (pair a rows, pair b columns)
            for (int e = tid(); e < half * half; e += kDT) {
                const int a = e % half, b = e / half;
                int p = rr_player(a, round, n), q = rr_player(n - 1 - a, round, n);
                if (p > q) {
                    int t = p;
                    p = q;
                    q = t;
                }
                int r = rr_player(b, round, n), s = rr_player(n - 1 - b, round, n);
                if (r > s) {
                    int t = r;
                    r = s;
                    s = t;
                }
                const double ca = cs[a], sa = sn[a], cb = cs[b], sb = sn[b];
                const double m00 = c[p + r * n], m01 = c[p + s * n], m10 = c[q + r * n],
                             m11 = c[q + s * n];
                // J = [[c, s], [-s, c]]: J' M
                const double t00 = ca * m00 - sa * m10, t01 = ca * m01 - sa * m11;
                const double t10 = sa * m00 + ca * m10, t11 = sa * m01 + ca * m11;
                // (J' M) J
                double o00 = t00 * cb - t01 * sb, o01 = t00 * sb + t01 * cb;
                double o10 = t10 * cb - t11 * sb, o11 = t10 * sb + t11 * cb;
                if (a == b) {
                    o01 = 0.0;
                    o10 = 0.0;
                }
                c[p + r * n] = o00;
                c[p + s * n] = o01;
                c[q + r * n] = o10;
                c[q + s * n] = o11;
            }
            // V <- V J
            for (int e = tid(); e < n * half; e += kDT) {
                const int row = e % n, a = e / n;
                int p = rr_player(a, round, n), q = rr_player(n - 1 - a, round, n);
                if (p > q) {
                    int t = p;
                    p = q;
                    q = t;
                }
                const double ca = cs[a], sa = sn[a];
                const double vp = v[row + p * n], vq = v[row + q * n];
                v[row + p * n] = ca * vp - sa * vq;
                v[row + q * n] = sa * vp + ca * vq;
            }
            __syncthreads();
        }
        if (*rot == 0) break;
        __syncthreads();
    }
    (void)red;
    return sweep;
}

struct RRArgs {
    const double *ga; // A-Gram, ld lda
    const double *gb; // B-Gram, ld lda
    int lda;
    int k;            // sizeX
    int a;            // sizeR (active residuals)
    int p;            // sizeP requested (0 on the first iteration)
    int initial;      // 1: plain pencil (X'AX, X'BX) of the initial subspace, symmetrised
    const double *lam_in;  // previous Ritz values (XAX = diag)
    double *lam_out;       // k new Ritz values
    double *coef;          // nA x k Ritz coordinates, ld ldc (rows of P zero when P dropped)
    int ldc;
    double *work;          // global scratch: 3 * lda * lda doubles
    DevStatus *st;
};

__global__ void __launch_bounds__(kDT) k_rr(RRArgs r) {
    extern __shared__ double lds[];
    __shared__ double cs[kMaxRR / 2 + 1], sn[kMaxRR / 2 + 1], red[kDT / 64], diag[kMaxRR];
    __shared__ int rot, info;
    __shared__ int rank_of[kMaxRR];
    if (r.st->abort) return;
    const int k = r.k, a = r.a;
    const int p = (r.p > 0 && r.st->p_info == 0) ? r.p : 0;
    const int nA = k + a + p;
    const int nE = (nA + 1) / 2 * 2;  // even size for the round-robin pairing
    // storage: C and V in LDS when they fit, else in the global work area
    double *A = r.work;                       // nA x nA, ld nA (assembled A, then T = A U^-1)
    double *B = r.work + (size_t)nA * nA;     // nA x nA (B, then U, then U^-1)
    const bool in_lds = (size_t)2 * nE * nE * sizeof(double) <= kRRLdsBytes;
    double *C = in_lds ? lds : r.work + (size_t)2 * nA * nA;
    double *V = in_lds ? lds + (size_t)nE * nE : C + (size_t)nE * nE;

    // ---- assemble the pencil (lower triangles of the Gram buffers) ----
    const int pb = k + a;  // first row of the P block in the basis ordering
    for (int e = tid(); e < nA * nA; e += kDT) {
        const int i = e % nA, j = e / nA;
        double va, vb;
        if (r.initial) {
            // symmetrised X'AX and X'BX (lobpcg.c:888-895)
            va = 0.5 * (r.ga[i + j * r.lda] + r.ga[j + i * r.lda]);
            vb = 0.5 * (r.gb[i + j * r.lda] + r.gb[j + i * r.lda]);
        } else {
            const int lo = i >= j ? i : j, hi = i >= j ? j : i;  // lower-triangle entry
            const bool xx = lo < k;
            const bool rr = lo >= k && lo < pb && hi >= k;
            const bool pp = lo >= pb && hi >= pb;
            if (xx) {
                va = (lo == hi) ? r.lam_in[lo] : 0.0;
                vb = (lo == hi) ? 1.0 : 0.0;
            } else if (rr || pp) {
                va = 0.5 * (r.ga[lo + hi * r.lda] + r.ga[hi + lo * r.lda]);
                vb = (lo == hi) ? 1.0 : 0.0;
            } else {
                va = r.ga[lo + hi * r.lda];
                vb = r.gb[lo + hi * r.lda];
            }
        }
        A[i + j * nA] = va;
        B[i + j * nA] = vb;
    }
    __syncthreads();
    // ---- B = U'U ----
    int bi = chol_upper(B, nA, nA, &info, diag);
    if (bi) {
        if (tid() == 0) {
            r.st->gevp_info = nA + bi;  // LAPACK dsygv reports N + i for a B failure
            r.st->abort = 1;
        }
        return;
    }
    for (int e = tid(); e < nA * nA; e += kDT) {
        const int i = e % nA, j = e / nA;
        if (i > j) B[i + j * nA] = 0.0;
    }
    __syncthreads();
    upper_inverse(B, nA, nA, diag);  // B := U^{-1}
    // T = A U^{-1}  (into C temporarily, nA x nA with ld nE)
    for (int e = tid(); e < nA * nA; e += kDT) {
        const int i = e % nA, j = e / nA;
        double s = 0.0;
        for (int l = 0; l <= j; l++) s += A[i + l * nA] * B[l + j * nA];
        C[i + j * nE] = s;
    }
    __syncthreads();
    for (int e = tid(); e < nA * nA; e += kDT) A[e % nA + (e / nA) * nA] = C[e % nA + (e / nA) * nE];
    __syncthreads();
    // C = U^{-T} T, symmetrised; padded to nE with a large isolated diagonal
    for (int e = tid(); e < nE * nE; e += kDT) {
        const int i = e % nE, j = e / nE;
        double s = 0.0;
        if (i < nA && j < nA) {
            for (int l = 0; l <= i; l++) s += B[l + i * nA] * A[l + j * nA];
        }
        C[i + j * nE] = s;
        V[i + j * nE] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int e = tid(); e < nA * nA; e += kDT) {
        const int i = e % nA, j = e / nA;
        if (i < j) {
            const double s = 0.5 * (C[i + j * nE] + C[j + i * nE]);
            C[i + j * nE] = s;
            C[j + i * nE] = s;
        }
    }
    __syncthreads();
    if (nE > nA && tid() == 0) {
        // the padding index gets an isolated eigenvalue above every real one
        // (twice the largest Gershgorin row sum)
        double mx = 0.0;
        for (int i = 0; i < nA; i++) {
            double s = 0.0;
            for (int j = 0; j < nA; j++) s += fabs(C[i + j * nE]);
            mx = s > mx ? s : mx;
        }
        C[nA + nA * nE] = 2.0 * mx + 1.0;
    }
    __syncthreads();
    jacobi(C, V, nE, cs, sn, &rot, red);
    // ---- ascending order of the nA real eigenvalues ----
    for (int i = tid(); i < nE; i += kDT) {
        const double li = C[i + i * nE];
        int rk = 0;
        for (int j = 0; j < nE; j++) {
            const double lj = C[j + j * nE];
            rk += (lj < li) || (lj == li && j < i);
        }
        rank_of[i] = rk;
    }
    __syncthreads();
    for (int i = tid(); i < nE; i += kDT) {
        const int rk = rank_of[i];
        if (rk < k) r.lam_out[rk] = C[i + i * nE];
        // column i of Q goes to position rk: stash the map in diag[]
        diag[rk] = (double)i;
    }
    __syncthreads();
    // ---- Ritz coordinates: coef(:, c) = U^{-1} Q(:, col of rank c) ----
    for (int e = tid(); e < nA * k; e += kDT) {
        const int i = e % nA, c = e / nA;
        const int col = (int)diag[c];
        double s = 0.0;
        for (int l = i; l < nA; l++) s += B[i + l * nA] * V[l + col * nE];
        // rows i < k + a keep their place; P rows follow R rows
        r.coef[i + c * r.ldc] = s;
    }
    // rows of a dropped P block are zero
    if (r.p > 0 && p == 0)
        for (int e = tid(); e < r.p * k; e += kDT)
            r.coef[(k + a + e % r.p) + (e / r.p) * r.ldc] = 0.0;
    if (tid() == 0) r.st->gevp_info = 0;
}


// ---- Rayleigh-Ritz with the R/P orthonormalisation folded in ----------
//
// The Gram buffers hold raw products of [X R P] with [AX AR AP] and
// [BX BR BP] (R, P not yet B-orthonormal).  With U_R = chol(R'BR),
// U_P = chol(P'BP) and S = diag(I, U_R^{-1}, U_P^{-1}), the pencil of the
// B-orthonormalised basis is S' G S: exactly what lobpcg_solve assembles from
// the explicitly normalised blocks (lobpcg.c:1027-1075 then 1092-1165).  The
// Ritz coordinates returned are S Y, so the caller combines the raw blocks.

__global__ void __launch_bounds__(kDT) k_rr_implicit(RR2Args r) {
    extern __shared__ double lds[];
    __shared__ double cs[kMaxRR / 2 + 1], sn[kMaxRR / 2 + 1], red[kDT / 64], diag[kMaxRR];
    __shared__ int rot, info, rank_of[kMaxRR];
    if (r.st->abort) return;
    const int k = r.k, a = r.a, pq = r.p;
    const int nmax = k + a + pq, nEmax = (nmax + 1) / 2 * 2;
    const bool in_lds = (size_t)2 * nEmax * nEmax * sizeof(double) <= kRRLdsBytes;
    double *G1 = r.work, *G2 = G1 + (size_t)nEmax * nEmax;
    double *WR = G2 + (size_t)nEmax * nEmax, *WP = WR + (size_t)a * a;
    double *Y = WP + (size_t)(pq > 0 ? pq : 1) * (pq > 0 ? pq : 1);
    double *M1 = in_lds ? lds : Y + (size_t)nmax * k;
    double *M2 = in_lds ? lds + (size_t)nEmax * nEmax : M1 + (size_t)nEmax * nEmax;

    // ---- U_R^{-1}, U_P^{-1} (small Cholesky factors in the M2 area) ----
    for (int e = tid(); e < a * a; e += kDT) M2[e] = r.gb[(k + e % a) + (size_t)(k + e / a) * r.lda];
    __syncthreads();
    int ci = chol_upper(M2, a, a, &info, diag);
    if (ci) {
        if (tid() == 0) {
            r.st->r_info = ci;
            r.st->abort = 1;
        }
        return;
    }
    upper_inverse(M2, a, a, diag);
    for (int e = tid(); e < a * a; e += kDT) WR[e] = M2[e];
    __syncthreads();
    int p = 0;
    if (pq > 0) {
        for (int e = tid(); e < pq * pq; e += kDT)
            M2[e] = r.gb[(k + a + e % pq) + (size_t)(k + a + e / pq) * r.lda];
        __syncthreads();
        ci = chol_upper(M2, pq, pq, &info, diag);
        if (ci == 0) {
            upper_inverse(M2, pq, pq, diag);
            for (int e = tid(); e < pq * pq; e += kDT) WP[e] = M2[e];
            p = pq;
        }
        if (tid() == 0) r.st->p_info = ci;
        __syncthreads();
    }
    const int nA = k + a + p, nE = (nA + 1) / 2 * 2;
    // block of a basis index: 0 = X, 1 = R, 2 = P
    auto blk = [&](int i) { return i < k ? 0 : (i < k + a ? 1 : 2); };

    // ---- raw symmetric pencil, then M <- S' M S ----
    for (int pass = 0; pass < 2; pass++) {
        double *M = pass == 0 ? M1 : M2;
        const double *g = pass == 0 ? r.ga : r.gb;
        for (int e = tid(); e < nA * nA; e += kDT) {
            const int i = e % nA, j = e / nA, bi = blk(i), bj = blk(j);
            double v;
            if (bi == 0 && bj == 0 && r.explicit_x)
                v = 0.5 * (g[i + (size_t)j * r.lda] + g[j + (size_t)i * r.lda]);
            else if (bi == 0 && bj == 0)
                v = (i == j) ? (pass == 0 ? r.lam_in[i] : 1.0) : 0.0;
            else if (bi > bj)
                v = g[i + (size_t)j * r.lda];
            else if (bi < bj)
                v = g[j + (size_t)i * r.lda];
            else
                v = g[i + (size_t)j * r.lda];
            M[i + j * nE] = v;
        }
        __syncthreads();
        // G1 = M S
        for (int e = tid(); e < nA * nA; e += kDT) {
            const int i = e % nA, j = e / nA, bj = blk(j);
            double s;
            if (bj == 0) {
                s = M[i + j * nE];
            } else {
                const int o = bj == 1 ? k : k + a, m = bj == 1 ? a : p, jj = j - o;
                const double *W = bj == 1 ? WR : WP;
                s = 0.0;
                for (int l = 0; l <= jj; l++) s += M[i + (o + l) * nE] * W[l + jj * m];
            }
            G1[i + (size_t)j * nE] = s;
        }
        __syncthreads();
        // M = S' G1
        for (int e = tid(); e < nA * nA; e += kDT) {
            const int i = e % nA, j = e / nA, bi = blk(i);
            double s;
            if (bi == 0) {
                s = G1[i + (size_t)j * nE];
            } else {
                const int o = bi == 1 ? k : k + a, m = bi == 1 ? a : p, ii = i - o;
                const double *W = bi == 1 ? WR : WP;
                s = 0.0;
                for (int l = 0; l <= ii; l++) s += W[l + ii * m] * G1[(o + l) + (size_t)j * nE];
            }
            M[i + j * nE] = s;
        }
        __syncthreads();
        // lower triangle wins off the diagonal blocks; RAR/PAP symmetrised;
        // RBR = PBP = I (lobpcg.c:1113-1162)
        for (int e = tid(); e < nA * nA; e += kDT) {
            const int i = e % nA, j = e / nA;
            if (i <= j) continue;
            const int bi = blk(i), bj = blk(j);
            double v;
            if (bi == bj && bi > 0)
                v = pass == 0 ? 0.5 * (M[i + j * nE] + M[j + i * nE]) : 0.0;
            else
                v = M[i + j * nE];
            M[i + j * nE] = v;
            M[j + i * nE] = v;
        }
        __syncthreads();
        if (pass == 1)
            for (int i = (r.explicit_x ? k : 0) + tid(); i < nA; i += kDT) M[i + i * nE] = 1.0;
        __syncthreads();
    }

    // ---- generalized eigenproblem M1 u = lambda M2 u ----
    int bi = chol_upper(M2, nE, nA, &info, diag);
    if (bi) {
        if (tid() == 0) {
            r.st->gevp_info = nA + bi;
            r.st->abort = 1;
        }
        return;
    }
    for (int e = tid(); e < nA * nA; e += kDT) {
        const int i = e % nA, j = e / nA;
        if (i > j) M2[i + j * nE] = 0.0;
    }
    __syncthreads();
    upper_inverse(M2, nE, nA, diag);  // M2 := U^{-1}
    // G1 = M1 U^{-1}
    for (int e = tid(); e < nA * nA; e += kDT) {
        const int i = e % nA, j = e / nA;
        double s = 0.0;
        for (int l = 0; l <= j; l++) s += M1[i + l * nE] * M2[l + j * nE];
        G1[i + (size_t)j * nE] = s;
    }
    // G2 = U^{-1} (kept for the eigenvectors)
    for (int e = tid(); e < nA * nA; e += kDT) G2[e % nA + (size_t)(e / nA) * nE] = M2[e % nA + (e / nA) * nE];
    __syncthreads();
    // M1 = U^{-T} G1 (padded to nE), M2 = I
    for (int e = tid(); e < nE * nE; e += kDT) {
        const int i = e % nE, j = e / nE;
        double s = 0.0;
        if (i < nA && j < nA)
            for (int l = 0; l <= i; l++) s += G2[l + (size_t)i * nE] * G1[l + (size_t)j * nE];
        M1[i + j * nE] = s;
        M2[i + j * nE] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int e = tid(); e < nA * nA; e += kDT) {
        const int i = e % nA, j = e / nA;
        if (i < j) {
            const double s = 0.5 * (M1[i + j * nE] + M1[j + i * nE]);
            M1[i + j * nE] = s;
            M1[j + i * nE] = s;
        }
    }
    __syncthreads();
    if (nE > nA && tid() == 0) {
        double mx = 0.0;
        for (int i = 0; i < nA; i++) {
            double s = 0.0;
            for (int j = 0; j < nA; j++) s += fabs(M1[i + j * nE]);
            mx = s > mx ? s : mx;
        }
        M1[nA + nA * nE] = 2.0 * mx + 1.0;
    }
    __syncthreads();
    jacobi(M1, M2, nE, cs, sn, &rot, red);
    for (int i = tid(); i < nE; i += kDT) {
        const double li = M1[i + i * nE];
        int rk = 0;
        for (int j = 0; j < nE; j++) {
            const double lj = M1[j + j * nE];
            rk += (lj < li) || (lj == li && j < i);
        }
        rank_of[i] = rk;
    }
    __syncthreads();
    for (int i = tid(); i < nE; i += kDT) {
        const int rk = rank_of[i];
        if (rk < k) r.lam_out[rk] = M1[i + i * nE];
        diag[rk] = (double)i;
    }
    __syncthreads();
    // Y = U^{-1} Q(:, first k), then coef = S Y
    for (int e = tid(); e < nA * k; e += kDT) {
        const int i = e % nA, c = e / nA, col = (int)diag[c];
        double s = 0.0;
        for (int l = i; l < nA; l++) s += G2[i + (size_t)l * nE] * M2[l + col * nE];
        Y[i + (size_t)c * nA] = s;
    }
    __syncthreads();
    for (int e = tid(); e < (k + a + pq) * k; e += kDT) {
        const int i = e % (k + a + pq), c = e / (k + a + pq), bi2 = blk(i);
        double s;
        if (i >= nA) {
            s = 0.0;  // dropped P block
        } else if (bi2 == 0) {
            s = Y[i + (size_t)c * nA];
        } else {
            const int o = bi2 == 1 ? k : k + a, m = bi2 == 1 ? a : p, ii = i - o;
            const double *W = bi2 == 1 ? WR : WP;
            s = 0.0;
            for (int l = ii; l < m; l++) s += W[ii + l * m] * Y[(o + l) + (size_t)c * nA];
        }
        r.coef[i + (size_t)c * r.ldc] = s;
    }
    if (tid() == 0) r.st->gevp_info = 0;
}

}  // namespace

void chol_inv(double *g, int ld, int m, DevStatus *st, int which) {
    BX_REQUIRE(m <= kMaxRR, "chol_inv: block of %d columns (max %d)", m, kMaxRR);
    if (m == 0) return;
    hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(kDT), 0, stream(), g, ld, m, st, which);
    BX_HIP(hipGetLastError());
}

void rr_solve(const double *ga, const double *gb, int lda, int k, int a, int p, int initial,
              const double *lam_in, double *lam_out, double *coef, int ldc, DevStatus *st) {
    const int nA = k + a + p;
    BX_REQUIRE(nA <= kMaxRR, "rr_solve: pencil of order %d (max %d)", nA, kMaxRR);
    const int nE = (nA + 1) / 2 * 2;
    double *work = (double *)scratch(6, (size_t)(2 * nA * nA + 2 * nE * nE) * sizeof(double));
    RRArgs r{ga, gb, lda, k, a, p, initial, lam_in, lam_out, coef, ldc, work, st};
    static bool attr_set = false;
    if (!attr_set) {
        BX_HIP(hipFuncSetAttribute((const void *)k_rr, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   kRRLdsBytes));
        attr_set = true;
    }
    const bool in_lds = (size_t)2 * nE * nE * sizeof(double) <= kRRLdsBytes;
    hipLaunchKernelGGL(k_rr, dim3(1), dim3(kDT), in_lds ? kRRLdsBytes : 0, stream(), r);
    BX_HIP(hipGetLastError());
}

void rr_solve_implicit(const double *ga, const double *gb, int lda, int k, int a, int p,
                       const double *lam_in, double *lam_out, double *coef, int ldc,
                       DevStatus *st, int explicit_x, int retry_without_p,
                       double cond_limit) {
    const int nmax = k + a + p;
    BX_REQUIRE(nmax <= kMaxRR, "rr_solve_implicit: pencil of order %d (max %d)", nmax, kMaxRR);
    const int nE = (nmax + 1) / 2 * 2;
    const size_t words = (size_t)4 * nE * nE + (size_t)a * a + (size_t)(p > 0 ? p * p : 1) +
                         (size_t)nmax * k + 64;
    double *work = (double *)scratch(7, words * sizeof(double));
    RR2Args r{ga, gb, lda, k, a, p, lam_in, lam_out, coef, ldc, work, st, explicit_x,
              retry_without_p, cond_limit};
    static const bool force_jacobi = [] {
        const char *e = getenv("BLOPEX_AMD_RR");
        return e && e[0] == 'j';
    }();
    if (rr_fast_fits(k, a, p) && !force_jacobi && !g_force_jacobi) {
        rr_fast(r);
        return;
    }
    if (rr_big_fits(k, a, p) && !force_jacobi && !g_force_jacobi) {
        rr_big(r);
        return;
    }
    static bool attr_set = false;
    if (!attr_set) {
        BX_HIP(hipFuncSetAttribute((const void *)k_rr_implicit,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, kRRLdsBytes));
        attr_set = true;
    }
    const bool in_lds = (size_t)2 * nE * nE * sizeof(double) <= kRRLdsBytes;
    hipLaunchKernelGGL(k_rr_implicit, dim3(1), dim3(kDT), in_lds ? kRRLdsBytes : 0, stream(), r);
    BX_HIP(hipGetLastError());
}

}  // namespace bx
