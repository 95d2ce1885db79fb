// Host-array wrappers of the MFMA building blocks (include/blopex_amd_kernels.h).
#include <vector>

#include "blopex_amd_kernels.h"
#include "dense.h"
#include "kernels.h"

namespace bx {
void mfma_layout_probe(int prec, int *rows, int *cols);
extern bool g_force_jacobi;
}

using namespace bx;

namespace {

struct DevBlock {
    void *p = nullptr;
    long ld = 0;
    int w = 0;
    DevBlock(long n, int w_, int prec, const double *host) : w(w_) {
        ld = (n + 63) / 64 * 64;
        const size_t bytes = (size_t)ld * (w > 0 ? w : 1) * elem_size(prec);
        BX_HIP(hipMalloc(&p, bytes));
        BX_HIP(hipMemsetAsync(p, 0, bytes, stream()));
        if (host && w > 0) {
            void *tmp = scratch(5, (size_t)n * w * sizeof(double));
            BX_HIP(hipMemcpyAsync(tmp, host, (size_t)n * w * 8, hipMemcpyHostToDevice, stream()));
            convert_block(n, w, tmp, n, HIPMV_FP64, p, ld, prec);
        }
    }
    void download(long n, int prec, double *host) {
        void *tmp = scratch(5, (size_t)n * w * sizeof(double));
        convert_block(n, w, p, ld, prec, tmp, n, HIPMV_FP64);
        download_sync(host, tmp, (size_t)n * w * sizeof(double));
    }
    ~DevBlock() {
        hipStreamSynchronize(stream());
        hipFree(p);
    }
};

}  // namespace

extern "C" {

int blopex_amd_kernel_update(int prec, long n, int k, int a, int p, const double *X,
                             const double *R, const double *P, const double *C, double *full,
                             double *tail) {
    DevBlock dX(n, k, prec, X), dR(n, a, prec, R), dP(n, p, prec, P), dT(n, k, prec, nullptr);
    const int crows = k + a + p;
    double *dC = (double *)scratch(8, (size_t)crows * k * sizeof(double));
    upload_async(dC, C, (size_t)crows * k * sizeof(double));
    UpdGroup g{};
    g.in[0] = dX.p;
    g.h[0] = k;
    g.crow[0] = 0;
    if (a) {
        g.in[1] = dR.p;
        g.h[1] = a;
        g.crow[1] = k;
    }
    if (p) {
        g.in[2] = dP.p;
        g.h[2] = p;
        g.crow[2] = k + a;
    }
    g.out_tail = dT.p;
    g.out_full = dX.p;  // in place, as the solver uses it
    block_update(prec, dX.ld, &g, 1, dC, crows, crows, k, nullptr);
    dX.download(n, prec, full);
    if (tail) dT.download(n, prec, tail);
    return 0;
}

int blopex_amd_kernel_update_resid(int prec, long n, int k, int a, int p, const double *X,
                                   const double *AX, const double *R, const double *AR,
                                   const double *P, const double *AP, const double *C,
                                   const int *tail_pos, int ntail, const double *lam,
                                   const double *tdiag, double *Xo, double *AXo, double *Po,
                                   double *APo, double *Ro, double *rn2) {
    BX_REQUIRE(ntail >= 1 && ntail <= k, "kernel_update_resid: %d tail columns for k = %d", ntail,
               k);
    DevBlock dX(n, k, prec, X), dAX(n, k, prec, AX), dR(n, a, prec, R), dAR(n, a, prec, AR),
        dP(n, p, prec, P), dAP(n, p, prec, AP), dPo(n, ntail, prec, nullptr),
        dAPo(n, ntail, prec, nullptr), dRo(n, ntail, prec, nullptr),
        dT(n, tdiag ? 1 : 0, prec, tdiag);
    const int crows = k + a + p;
    double *dC = (double *)scratch(8, (size_t)crows * k * sizeof(double));
    upload_async(dC, C, (size_t)crows * k * sizeof(double));
    double *dlam = (double *)scratch(9, (size_t)(k + ntail) * sizeof(double)), *drn2 = dlam + k;
    upload_async(dlam, lam, (size_t)k * sizeof(double));
    UpdGroup g[2] = {};
    void *in[2][3] = {{dX.p, dR.p, dP.p}, {dAX.p, dAR.p, dAP.p}};
    void *tails[2] = {dPo.p, dAPo.p};
    for (int q = 0; q < 2; q++) {
        g[q].in[0] = in[q][0];
        g[q].h[0] = k;
        g[q].crow[0] = 0;
        if (a) {
            g[q].in[1] = in[q][1];
            g[q].h[1] = a;
            g[q].crow[1] = k;
        }
        if (p) {
            g[q].in[2] = in[q][2];
            g[q].h[2] = p;
            g[q].crow[2] = k + a;
        }
        g[q].out_tail = tails[q];
        g[q].out_full = in[q][0];  // in place, as the solver uses it
    }
    block_update_resid(prec, n, dX.ld, g, dC, crows, crows, k, nullptr, tail_pos, ntail, dlam,
                       tdiag ? dT.p : nullptr, dRo.p, drn2);
    dX.download(n, prec, Xo);
    dAX.download(n, prec, AXo);
    dPo.download(n, prec, Po);
    dAPo.download(n, prec, APo);
    dRo.download(n, prec, Ro);
    download_sync(rn2, drn2, (size_t)ntail * sizeof(double));
    return 0;
}

int blopex_amd_kernel_gram(int prec, long n, int lw, int qw, const double *L, const double *Q,
                           double *G) {
    DevBlock dL(n, lw, prec, L), dQ(n, qw, prec, Q);
    double *dG = (double *)scratch(8, (size_t)lw * qw * sizeof(double));
    GramBatch gb(prec, dL.ld, lw);
    gb.add(dL.p, lw, dQ.p, qw, dG);
    gb.flush();
    download_sync(G, dG, (size_t)lw * qw * sizeof(double));
    return 0;
}

}  // extern "C"

namespace {
int kernel_rr(int k, int a, int p, const double *ga, const double *gb, int lda,
              const double *lam_in, int explicit_x, int retry_without_p, int jacobi,
              double cond_limit, double *lam_out, double *coef, int *flags, double *condest) {
    const int n = k + a + p;
    const size_t gbytes = (size_t)lda * n * sizeof(double);
    char *d = (char *)scratch(9, 2 * gbytes + (size_t)(2 * k + n * k) * 8 + sizeof(DevStatus) + 64);
    double *dga = (double *)d, *dgb = (double *)(d + gbytes);
    double *dlam = (double *)(d + 2 * gbytes), *dlo = dlam + k, *dc = dlo + k;
    DevStatus *st = (DevStatus *)(dc + (size_t)n * k);
    upload_async(dga, ga, gbytes);
    upload_async(dgb, gb, gbytes);
    upload_async(dlam, lam_in, (size_t)k * 8);
    BX_HIP(hipMemsetAsync(st, 0, sizeof(DevStatus), stream()));
    g_force_jacobi = jacobi != 0;
    rr_solve_implicit(dga, dgb, lda, k, a, p, dlam, dlo, dc, n, st, explicit_x, retry_without_p,
                      cond_limit);
    g_force_jacobi = false;
    DevStatus h;
    download_sync(&h, st, sizeof(h));
    download_sync(lam_out, dlo, (size_t)k * 8);
    download_sync(coef, dc, (size_t)n * k * 8);
    flags[0] = h.abort;
    flags[1] = h.r_info;
    flags[2] = h.p_info;
    flags[3] = h.gevp_info;
    if (condest) *condest = h.condest;
    return 0;
}
}  // namespace

extern "C" {

int blopex_amd_kernel_rr(int k, int a, int p, const double *ga, const double *gb, int lda,
                         const double *lam_in, int explicit_x, int retry_without_p, int jacobi,
                         double *lam_out, double *coef, int *flags) {
    return kernel_rr(k, a, p, ga, gb, lda, lam_in, explicit_x, retry_without_p, jacobi, 0.0,
                     lam_out, coef, flags, nullptr);
}

int blopex_amd_kernel_rr_cond(int k, int a, int p, const double *ga, const double *gb, int lda,
                              const double *lam_in, int explicit_x, double cond_limit,
                              double *lam_out, double *coef, int *flags, double *condest) {
    return kernel_rr(k, a, p, ga, gb, lda, lam_in, explicit_x, 0, 0, cond_limit, lam_out, coef,
                     flags, condest);
}

int blopex_amd_kernel_mfma_probe(int prec, int *rows, int *cols) {
    mfma_layout_probe(prec, rows, cols);
    return 0;
}

}  // extern "C"
