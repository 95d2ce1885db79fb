// hip_LobpcgAdjoint (include/lobpcg_adjoint.h): gradients of a loss of the
// eigenpairs with respect to the stored values of A and the diagonal of B.
//
// Everything runs on the library stream with the forward solve's kernels:
// CSR SpMM and the diagonal operators for A - lambda_i B and T, the Gram and
// block-times-small-matrix kernels for the projections against X, and
// per-column axpby for the conjugate-gradient steps.  The projected systems
//
//   (A - lambda_i B) u_i = g_i - B X X' g_i,   X' B u_i = 0,
//
// of all columns with a nonzero g_i are solved together: one block CG whose
// columns carry their own shift lambda_i and their own step lengths, the
// preconditioned residual projected by Q = I - X X' B (so the iterates stay
// in the B-complement of X) and A p projected by Q' (so the residual stays
// orthogonal to X).  MIXED runs the CG sweeps on FP32 copies and corrects
// them with residuals recomputed in FP64 (iterative refinement).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "kernels.h"
#include "lobpcg_adjoint.h"

using namespace bx;

namespace {

// one device allocation for every block of a call, cleared once; the stream
// is drained before it is freed
struct Arena {
    char *base = nullptr;
    size_t used = 0, size = 0;
    explicit Arena(size_t bytes) : size(bytes) {
        BX_HIP(hipMalloc((void **)&base, size ? size : 256));
        BX_HIP(hipMemsetAsync(base, 0, size ? size : 256, stream()));
    }
    ~Arena() {
        hipStreamSynchronize(stream());
        hipFree(base);
    }
    void *take(size_t bytes) {
        bytes = (bytes + 255) / 256 * 256;
        void *p = base + used;
        used += bytes;
        BX_REQUIRE(used <= size, "hip_LobpcgAdjoint: workspace overflow");
        return p;
    }
};

constexpr int kNumBlocks64 = 13;
constexpr int kNumBlocks32 = 10;
constexpr int kSmall = 16;  // device arrays of kMaxCols doubles

// blocks of one precision for the CG sweeps
struct Set {
    int prec;
    const void *X, *NX, *BX, *NBX;  // X, -X, B X, -B X (B X = X without B)
    void *R, *Z, *P, *Q, *D, *T;
};

struct Ctx {
    long n;
    int k;
    const hip_CSRMatrix *A;
    const hip_DiagMatrix *B, *T;
    ColMap all, cm;
    std::vector<double> lam;
    double *G;       // k x k (device)
    double *dots;    // 3 kMaxCols (device)
    double *coef_a, *coef_b, *neglam, *ones;  // kMaxCols each (device)
    double *rz, *target, *beta;                // CG scalars (device)
    int *done;                                 // kMaxCols + 1 flags (device)
};

void upload(double *dst, const std::vector<double> &v) {
    upload_async(dst, v.data(), v.size() * sizeof(double));
}

// Z(:, cm) = Q T R(:, cm),  Q = I - X X' B
void precond_project(Ctx &c, Set &s, const void *R, void *Z) {
    if (c.T)
        diag_apply(s.prec, c.T, R, c.n, c.cm, Z, c.n, c.cm);
    else
        copy_cols(s.prec, c.n, R, c.n, c.cm, Z, c.n, c.cm);
    gram_cols(s.prec, c.n, s.BX, c.n, c.all, Z, c.n, c.cm, c.G, c.k);
    vecmat_cols(s.prec, c.n, s.NX, c.n, c.all, c.G, c.k, Z, c.n, c.cm, true);
}

// Y(:, cm) = Q' (A - lambda B) P(:, cm),  Q' = I - B X X'
void apply_shifted(Ctx &c, Set &s, const void *P, void *Y) {
    spmm_csr(s.prec, c.A, P, c.n, c.cm, Y, c.n, c.cm);
    const void *bp = P;
    if (c.B) {
        diag_apply(s.prec, c.B, P, c.n, c.cm, s.T, c.n, c.cm);
        bp = s.T;
    }
    axpby_cols(s.prec, c.n, c.neglam, bp, c.n, c.ones, Y, c.n, c.cm);
    gram_cols(s.prec, c.n, s.X, c.n, c.all, Y, c.n, c.cm, c.G, c.k);
    vecmat_cols(s.prec, c.n, s.NBX, c.n, c.all, c.G, c.k, Y, c.n, c.cm, true);
}

// Block CG on the projected systems, D(:, cm) from zero, R(:, cm) the right
// sides (overwritten by the recursive residuals).  Column j stops once its
// residual norm is at most target[j].  The step lengths and the convergence
// flags stay on the device; the host reads the flags (and the residual
// norms) at most kCheck iterations apart, sooner when the residuals' rate
// of decrease says the last column finishes earlier.  A finished column
// takes zero steps meanwhile, and the count returned is that of the
// iterations in which some column still ran (kept on the device).  Returns
// -1 on a breakdown (p' Q' (A - lambda B) p <= 0: the shifted operator is
// not positive definite on the complement of X).
constexpr int kCheck = 8;
int block_cg(Ctx &c, Set &s, const std::vector<double> &target, int maxit) {
    const int m = c.cm.n;
    clear_cols(s.prec, c.n, s.D, c.n, c.cm);
    precond_project(c, s, s.R, s.Z);
    diagdot_cols(s.prec, c.n, s.R, c.n, c.cm, s.Z, c.n, c.cm, c.rz);
    diagdot_cols(s.prec, c.n, s.R, c.n, c.cm, s.R, c.n, c.cm, c.dots);
    copy_cols(s.prec, c.n, s.Z, c.n, c.cm, s.P, c.n, c.cm);
    std::vector<double> h(m);
    download_sync(h.data(), c.dots, m * sizeof(double));
    std::vector<int> done(m + 1, 0);
    bool all = true;
    for (int j = 0; j < m; j++) {
        done[j] = std::sqrt(std::max(h[j], 0.0)) <= target[j];
        all = all && done[j];
    }
    if (all) return 0;
    upload(c.target, target);
    // done[m] is the breakdown flag, done[m + 1] the count of running iterations
    done.push_back(0);
    upload_async(c.done, done.data(), (m + 2) * sizeof(int));
    int *bad = c.done + m;
    // worst log(||r_j|| / target_j) over the running columns at the last check
    auto worst_log = [&](const std::vector<double> &rr) {
        double w = 0.0;
        for (int j = 0; j < m; j++)
            if (!done[j])
                w = std::max(w, 0.5 * std::log(std::max(rr[j], 1e-300)) -
                                    std::log(std::max(target[j], 1e-300)));
        return w;
    };
    double last_worst = worst_log(h);
    int next_check = kCheck, last_check = 0;
    auto body = [&]() {
        apply_shifted(c, s, s.P, s.Q);
        diagdot_cols(s.prec, c.n, s.P, c.n, c.cm, s.Q, c.n, c.cm, c.dots);
        cg_alpha(m, c.rz, c.dots, c.done, c.coef_a, c.coef_b, bad);
        axpby_cols(s.prec, c.n, c.coef_a, s.P, c.n, c.ones, s.D, c.n, c.cm);
        axpby_cols(s.prec, c.n, c.coef_b, s.Q, c.n, c.ones, s.R, c.n, c.cm);
        diagdot_cols(s.prec, c.n, s.R, c.n, c.cm, s.R, c.n, c.cm, c.dots);
        precond_project(c, s, s.R, s.Z);
        diagdot_cols(s.prec, c.n, s.R, c.n, c.cm, s.Z, c.n, c.cm, c.dots + kMaxCols);
        cg_beta(m, c.dots, c.dots + kMaxCols, c.target, c.done, c.rz, c.beta);
        axpby_cols(s.prec, c.n, c.ones, s.Z, c.n, c.beta, s.P, c.n, c.cm);
    };
    int it = 0;
    while (it < maxit) {
        body();
        it++;
        if (it >= next_check || it >= maxit) {
            download_sync(done.data(), c.done, (m + 2) * sizeof(int));
            download_sync(h.data(), c.dots, m * sizeof(double));
            if (done[m]) {
                const int j = done[m] - 1;
                set_error("hip_LobpcgAdjoint: breakdown in column %d by CG iteration %d: "
                          "p'(A - lambda B)p <= 0, lambda_%d is not below the spectrum left "
                          "after deflating X",
                          c.cm.idx[j], it, c.cm.idx[j]);
                return -1;
            }
            all = true;
            for (int j = 0; j < m; j++) all = all && done[j];
            if (all) break;
            // iterations the worst column still needs at the rate since the last check
            const double worst = worst_log(h), rate = (last_worst - worst) / (it - last_check);
            int step = kCheck;
            if (rate > 0.0) step = std::max(1, std::min(kCheck, (int)std::ceil(worst / rate)));
            last_worst = worst;
            last_check = it;
            next_check = it + step;
        }
    }
    download_sync(done.data(), c.done, (m + 2) * sizeof(int));
    return done[m + 1];
}

bool check_block(const hip_Multi_Vector *v, long n, int k, const char *name) {
    if (!v) return true;
    if (v->precision != HIPMV_FP64 || v->size != n || v->num_vectors != k) {
        set_error("hip_LobpcgAdjoint: %s must be an FP64 block of %ld x %d (got %ld x %d, %s)",
                  name, n, k, v->size, (int)v->num_vectors,
                  v->precision == HIPMV_FP64 ? "FP64" : "FP32");
        return false;
    }
    return true;
}

}  // namespace

extern "C" {

void hip_AdjointSettingsDefault(hip_AdjointSettings *s) {
    if (!s) return;
    s->mode = HIPLOBPCG_FP64;
    s->tol = 1e-10;
    s->maxit = 2000;
}

int hip_LobpcgAdjoint(const hip_CSRMatrix *A, const hip_DiagMatrix *B, const hip_DiagMatrix *T,
                      const hip_Multi_Vector *X, const double *lambda, const double *glambda,
                      const hip_Multi_Vector *GX, double *gval, double *gb, hip_Multi_Vector *W,
                      const hip_AdjointSettings *settings, hip_AdjointStats *st) {
    const auto t0 = std::chrono::steady_clock::now();
    hip_AdjointSettings s;
    hip_AdjointSettingsDefault(&s);
    if (settings) s = *settings;
    if (!A || !X || !lambda || !gval) {
        set_error("hip_LobpcgAdjoint: A, X, lambda and gval are required");
        return -1;
    }
    const long n = A->n;
    const int k = X->num_vectors;
    if (k < 1 || k > kMaxCols) {
        set_error("hip_LobpcgAdjoint: X has %d columns (1..%d supported)", k, kMaxCols);
        return -1;
    }
    if (!check_block(X, n, k, "X") || !check_block(GX, n, k, "GX") || !check_block(W, n, k, "W"))
        return -1;
    if ((B && B->n != n) || (T && T->n != n)) {
        set_error("hip_LobpcgAdjoint: B or T does not have %ld rows", n);
        return -1;
    }
    if (s.mode != HIPLOBPCG_FP64 && s.mode != HIPLOBPCG_MIXED) {
        set_error("hip_LobpcgAdjoint: unknown mode %d", s.mode);
        return -1;
    }
    if (!(s.tol > 0.0) || s.maxit < 0) {
        set_error("hip_LobpcgAdjoint: tol must be positive and maxit non-negative");
        return -1;
    }
    const bool mixed = s.mode == HIPLOBPCG_MIXED;
    const size_t nk = (size_t)n * k;
    Arena ar(kNumBlocks64 * (nk * 8 + 256) + (mixed ? kNumBlocks32 * (nk * 4 + 256) : 0) +
             (size_t)k * k * 8 + 256 + kSmall * (kMaxCols * 8 + 256));
    auto b64 = [&]() { return (double *)ar.take(nk * 8); };
    auto b32 = [&]() { return (float *)ar.take(nk * 4); };

    Ctx c;
    c.n = n;
    c.k = k;
    c.A = A;
    c.B = B;
    c.T = T;
    c.all = colmap_identity(k);
    c.lam.assign(lambda, lambda + k);
    c.G = (double *)ar.take((size_t)k * k * 8);
    c.dots = (double *)ar.take(3 * kMaxCols * 8);
    c.coef_a = (double *)ar.take(kMaxCols * 8);
    c.coef_b = (double *)ar.take(kMaxCols * 8);
    c.neglam = (double *)ar.take(kMaxCols * 8);
    c.ones = (double *)ar.take(kMaxCols * 8);
    c.rz = (double *)ar.take(kMaxCols * 8);
    c.target = (double *)ar.take(kMaxCols * 8);
    c.beta = (double *)ar.take(kMaxCols * 8);
    c.done = (int *)ar.take((kMaxCols + 1) * 8);
    double *gl_d = (double *)ar.take(kMaxCols * 8);
    double *lam_d = (double *)ar.take(kMaxCols * 8);

    double *X64 = b64(), *NX64 = b64(), *W64 = b64(), *U64 = b64(), *RHS = b64(), *R64 = b64(),
           *T64 = b64(), *Z64 = b64(), *P64 = b64(), *Q64 = b64(), *D64 = b64();
    double *BX64 = X64, *NBX64 = NX64;
    double *bx_store = b64(), *nbx_store = b64();
    std::vector<double> minus_one(k, -1.0);
    copy_cols(HIPMV_FP64, n, X->data, X->ld, c.all, X64, n, c.all);
    scale_cols(HIPMV_FP64, n, minus_one.data(), X64, n, c.all, NX64, n, c.all);
    if (B) {
        BX64 = bx_store;
        NBX64 = nbx_store;
        diag_apply(HIPMV_FP64, B, X64, n, c.all, BX64, n, c.all);
        scale_cols(HIPMV_FP64, n, minus_one.data(), BX64, n, c.all, NBX64, n, c.all);
    }

    // C = X' G, the in-block part of w, and the columns that need a solve
    std::vector<double> C((size_t)k * k, 0.0), gg(k, 0.0);
    if (GX) {
        gram_cols(HIPMV_FP64, n, X64, n, c.all, GX->data, GX->ld, c.all, c.G, k);
        download_sync(C.data(), c.G, C.size() * sizeof(double));
        diagdot_cols(HIPMV_FP64, n, GX->data, GX->ld, c.all, GX->data, GX->ld, c.all, c.dots);
        download_sync(gg.data(), c.dots, k * sizeof(double));
        std::vector<double> cin((size_t)k * k, 0.0);
        for (int i = 0; i < k; i++)
            for (int j = 0; j < k; j++)
                if (j != i) {
                    const double gap = c.lam[j] - c.lam[i];
                    if (gap == 0.0 && C[j + (size_t)i * k] != 0.0) {
                        set_error("hip_LobpcgAdjoint: lambda_%d == lambda_%d and x_%d' g_%d != 0: "
                                  "the loss has no derivative there",
                                  j, i, j, i);
                        return -1;
                    }
                    cin[j + (size_t)i * k] = gap == 0.0 ? 0.0 : C[j + (size_t)i * k] / gap;
                }
        upload(c.G, cin);
        vecmat_cols(HIPMV_FP64, n, X64, n, c.all, c.G, k, W64, n, c.all, false);
    }
    c.cm.n = 0;
    for (int i = 0; i < k; i++)
        if (gg[i] > 0.0) c.cm.idx[c.cm.n++] = i;
    const int m0 = c.cm.n;

    // right sides g_i - B X X' g_i and their norms
    std::vector<double> bnorm;
    if (m0 > 0) {
        copy_cols(HIPMV_FP64, n, GX->data, GX->ld, c.cm, RHS, n, c.cm);
        std::vector<double> negc((size_t)k * m0);
        for (int j = 0; j < m0; j++)
            for (int i = 0; i < k; i++) negc[i + (size_t)j * k] = -C[i + (size_t)c.cm.idx[j] * k];
        upload(c.G, negc);
        vecmat_cols(HIPMV_FP64, n, BX64, n, c.all, c.G, k, RHS, n, c.cm, true);
        diagdot_cols(HIPMV_FP64, n, RHS, n, c.cm, RHS, n, c.cm, c.dots);
        std::vector<double> h(m0);
        download_sync(h.data(), c.dots, m0 * sizeof(double));
        ColMap keep;
        keep.n = 0;
        for (int j = 0; j < m0; j++)
            if (h[j] > 0.0) {
                keep.idx[keep.n++] = c.cm.idx[j];
                bnorm.push_back(std::sqrt(h[j]));
            }
        c.cm = keep;
    }
    const int m = c.cm.n;

    int iterations = 0, corrections = 0, flag = 0;
    double maxrel = 0.0;
    if (m > 0) {
        std::vector<double> nl(m), one(m, 1.0);
        for (int j = 0; j < m; j++) nl[j] = -c.lam[c.cm.idx[j]];
        upload(c.neglam, nl);
        upload(c.ones, one);
        Set s64{HIPMV_FP64, X64, NX64, BX64, NBX64, R64, Z64, P64, Q64, D64, T64};
        Set s32{};
        if (mixed) {
            float *X32 = b32(), *NX32 = b32(), *BX32 = X32, *NBX32 = NX32;
            float *bx32 = b32(), *nbx32 = b32();
            convert_block(n, k, X64, n, HIPMV_FP64, X32, n, HIPMV_FP32);
            convert_block(n, k, NX64, n, HIPMV_FP64, NX32, n, HIPMV_FP32);
            if (B) {
                BX32 = bx32;
                NBX32 = nbx32;
                convert_block(n, k, BX64, n, HIPMV_FP64, BX32, n, HIPMV_FP32);
                convert_block(n, k, NBX64, n, HIPMV_FP64, NBX32, n, HIPMV_FP32);
            }
            s32 = Set{HIPMV_FP32, X32, NX32, BX32, NBX32, b32(), b32(), b32(), b32(), b32(), b32()};
        }
        copy_cols(HIPMV_FP64, n, RHS, n, c.cm, R64, n, c.cm);
        std::vector<double> h(m), target(m);
        // FP32 sweeps stop at this relative reduction of their own right sides
        const double kSweep = 1e-5;
        for (;;) {
            diagdot_cols(HIPMV_FP64, n, R64, n, c.cm, R64, n, c.cm, c.dots);
            download_sync(h.data(), c.dots, m * sizeof(double));
            maxrel = 0.0;
            for (int j = 0; j < m; j++)
                maxrel = std::max(maxrel, std::sqrt(std::max(h[j], 0.0)) / bnorm[j]);
            if (maxrel <= s.tol || iterations >= s.maxit || corrections >= 64) break;
            int got;
            if (mixed) {
                for (int j = 0; j < m; j++)
                    target[j] = std::max(s.tol * bnorm[j], kSweep * std::sqrt(std::max(h[j], 0.0)));
                convert_block(n, k, R64, n, HIPMV_FP64, s32.R, n, HIPMV_FP32);
                got = block_cg(c, s32, target, s.maxit - iterations);
                if (got > 0) {
                    convert_block(n, k, s32.D, n, HIPMV_FP32, T64, n, HIPMV_FP64);
                    axpy_cols(HIPMV_FP64, n, 1.0, T64, n, c.cm, U64, n, c.cm);
                }
            } else {
                for (int j = 0; j < m; j++) target[j] = s.tol * bnorm[j];
                got = block_cg(c, s64, target, s.maxit - iterations);
                if (got > 0) axpy_cols(HIPMV_FP64, n, 1.0, D64, n, c.cm, U64, n, c.cm);
            }
            if (got < 0) return -1;
            iterations += got;
            corrections++;
            if (got == 0) break;
            // true residual of U in FP64
            apply_shifted(c, s64, U64, Q64);
            copy_cols(HIPMV_FP64, n, RHS, n, c.cm, R64, n, c.cm);
            axpy_cols(HIPMV_FP64, n, -1.0, Q64, n, c.cm, R64, n, c.cm);
        }
        flag = maxrel <= s.tol ? 0 : 1;
        // U back into the B-complement of X (rounding leaves components along
        // X that the projected residual does not see), then w = in-block + U
        gram_cols(HIPMV_FP64, n, BX64, n, c.all, U64, n, c.cm, c.G, k);
        vecmat_cols(HIPMV_FP64, n, NX64, n, c.all, c.G, k, U64, n, c.cm, true);
        axpy_cols(HIPMV_FP64, n, 1.0, U64, n, c.cm, W64, n, c.cm);
    }

    std::vector<double> glv(k, 0.0);
    if (glambda) glv.assign(glambda, glambda + k);
    upload(gl_d, glv);
    sampled_outer(A, X64, W64, n, k, gl_d, gval);
    if (gb) {
        std::vector<double> av(k);
        for (int i = 0; i < k; i++) av[i] = -glv[i] * c.lam[i] - 0.5 * C[i + (size_t)i * k];
        upload(c.coef_a, av);
        upload(lam_d, c.lam);
        diag_outer(n, X64, W64, n, k, c.coef_a, lam_d, gb);
    }
    if (W) copy_cols(HIPMV_FP64, n, W64, n, c.all, W->data, W->ld, c.all);
    BX_HIP(hipStreamSynchronize(stream()));
    if (st) {
        st->iterations = iterations;
        st->corrections = corrections;
        st->residual = maxrel;
        st->seconds =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (flag) set_error("hip_LobpcgAdjoint: relative residual %g above tol %g after %d iterations",
                        maxrel, s.tol, iterations);
    return flag;
}

}  // extern "C"
