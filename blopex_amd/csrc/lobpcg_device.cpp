// Device-resident LOBPCG: the steps of lobpcg_solve (krylov/lobpcg.c:
// 502-1388) with every block in HBM and the FP64 small-dense work (Cholesky
// of the Gram blocks, triangular inverse, Rayleigh-Ritz) on the GPU.  The host
// keeps only what the reference keeps on the host anyway: the soft-locking
// mask, the eigenvalue / residual histories and the exit bookkeeping; per
// iteration it reads back k Ritz values, the active residual norms and a
// status word (with the next iteration's A R already enqueued behind that
// read, used when the active set is unchanged).
//
// Block layout: X, AX, BX and P, AP, BP are n x k column-major; the residual
// blocks R, AR, BR hold only the active columns, compacted in ascending
// order, so every kernel runs over contiguous column ranges.  When the mask
// shrinks, the surviving columns of R and P/AP/BP are gathered down.  Block
// results are written to a spare buffer and the pointers swapped instead of
// copying back (the reference's W/copy pairs).
//
// Precision: FP64 runs everything in FP64.  MIXED iterates with FP32 blocks
// (FP64 Gram accumulation, FP64 small dense; AX and BX recomputed from X
// every refresh_interval iterations) until every wanted residual is below
// switch_factor times its threshold (|lambda| rtol + atol; with eig_rtol,
// every eigenvalue bound within switch_factor^2 of its own), the FP32
// residuals stop improving, or an FP32 Rayleigh-Ritz step breaks down; then
// it restarts from the last FP32 X in FP64 (re-orthonormalisation
// and Rayleigh-Ritz in FP64) and finishes there, so the final residuals and
// Ritz values are FP64 ones.  refine_iterations adds FP64 iterations with
// every column active after convergence.
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "dense.h"
#include "kernels.h"
#include "lobpcg.h"

extern "C" int blopex_amd_is_hip_interpreter(const mv_InterfaceInterpreter *i);
extern "C" mv_InterfaceInterpreter *blopex_amd_mv_interpreter(mv_MultiVectorPtr x);

namespace bx {

namespace {

using Op = void (*)(void *, void *, void *);

BlopexInt *identity_indices() {
    static BlopexInt ids[kMaxCols];
    static bool init = false;
    if (!init) {
        for (int i = 0; i < kMaxCols; i++) ids[i] = i;
        init = true;
    }
    return ids;
}

hip_Multi_Vector view(void *p, long n, int w, int prec) {
    hip_Multi_Vector v;
    memset(&v, 0, sizeof(v));
    v.data = p;
    v.size = n;
    v.num_vectors = w;
    v.num_active_vectors = w;
    v.active_indices = identity_indices();
    v.precision = prec;
    return v;
}

// Optional device timing of kernel classes: an event pair around each group
// of launches, resolved after the stream has been synchronised.
enum KClass { kSpmm, kGram, kUpdate, kDense, kOther, kNumClasses };

// library workspace slots of the solver (see scratch() in common.h)
constexpr int kSlotPool = 11, kSlotSmall = 12, kSlotY = 13, kSlotXw = 14;

// residual level below which FP64 iterations measure X'AX and X'BX.
// lobpcg.m switches at eps^0.6 (4e-10); here the implicit blocks already broke
// a 10^3 Jacobi solve down near 1e-10, so the switch comes at eps^0.5 (1.5e-8).
// Active columns only: at the usual tolerances (above the switch) columns
// lock before they reach it, and the iterations stay those of lobpcg.c
const double kExplicitGramRes = sqrt(DBL_EPSILON);

class KTimer {
  public:
    explicit KTimer(bool on) : on_(on) {}
    bool on() const { return on_; }
    ~KTimer() {
        for (hipEvent_t e : pool_) hipEventDestroy(e);
    }
    void begin(int cls) {
        if (!on_) return;
        cls_ = cls;
        a_ = ev();
        BX_HIP(hipEventRecord(a_, stream()));
    }
    void end(double bytes = 0.0) {
        if (!on_) return;
        hipEvent_t b = ev();
        BX_HIP(hipEventRecord(b, stream()));
        recs_.push_back({cls_, a_, b, bytes});
    }
    // call after a stream synchronisation
    void resolve() {
        if (!on_) return;
        for (const Rec &r : recs_) {
            float ms = 0.f;
            BX_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            acc_[r.cls] += ms * 1e-3;
            if (r.cls == kSpmm) {
                spmm_bytes_ += r.bytes;
                spmm_calls_++;
            }
        }
        recs_.clear();
        used_ = 0;
    }
    void add_to(hip_LobpcgStats *st) {
        if (!on_) return;
        st->seconds_spmm += acc_[kSpmm];
        st->seconds_gram += acc_[kGram];
        st->seconds_update += acc_[kUpdate];
        st->seconds_small_dense += acc_[kDense];
        st->seconds_other += acc_[kOther];
        st->spmm_bytes += spmm_bytes_;
        st->spmm_calls += spmm_calls_;
    }

  private:
    struct Rec {
        int cls;
        hipEvent_t a, b;
        double bytes;
    };
    hipEvent_t ev() {
        if (used_ == pool_.size()) {
            hipEvent_t e;
            BX_HIP(hipEventCreate(&e));
            pool_.push_back(e);
        }
        return pool_[used_++];
    }
    bool on_;
    int cls_ = 0;
    hipEvent_t a_ = nullptr;
    std::vector<hipEvent_t> pool_;
    size_t used_ = 0;
    std::vector<Rec> recs_;
    double acc_[kNumClasses] = {};
    double spmm_bytes_ = 0.0;
    int spmm_calls_ = 0;
};

struct HostOut {
    double *lam, *lamh, *res, *resh;
    int lamh_gh, resh_gh;
};

class Phase {
  public:
    int fp32_exit_ = 0;  // why the FP32 phase ended (hip_LobpcgStats.fp32_exit)
    int refine_done_ = 0;  // refinement iterations run (hip_LobpcgStats.refine_iterations)
    int final_rr_done_ = 0;  // 1 once the final Rayleigh-Ritz ran
    int cond_restarts_ = 0;  // steps solved again without P on a condition-estimate jump
    // k: columns iterated; kw <= k: the wanted ones (the rest are guard columns)
    Phase(int prec, long n, int k, int kw, int ny, Op opA, void *Ad, Op opB, void *Bd, Op opT,
          void *Td, const hip_Multi_Vector *Yuser, lobpcg_Tolerance tol, int verb, HostOut out,
          KTimer *tm)
        : prec_(prec), n_(n), k_(k), kw_(kw), ny_(ny), opA_(opA), Ad_(Ad), opB_(opB), Bd_(Bd),
          opT_(opT), Td_(Td), Yu_(Yuser), tol_(tol), verb_(verb), out_(out), tm_(tm) {
        noB_ = opB == nullptr;
        noT_ = opT == nullptr;
        fuseT_ = opT == HipDiagMatMultiVec;
        es_ = elem_size(prec);
        ld3_ = 3 * k;
        // padded leading dimension (padding rows stay zero): a multiple of 64
        // rows, and for large blocks a multiple of 256 plus 256, so that the
        // columns of a block do not all start at one large power-of-two
        // alignment (at n = 2^21 the FP64 stencil product took 282 us with
        // columns 16 MiB apart and 247 us with 2 KiB more; the solves 1-2%
        // less)
        ld_ = n >= (1L << 16) ? (n + 255) / 256 * 256 + 256 : (n + 63) / 64 * 64;
        const size_t blk = (size_t)ld_ * k * es_;
        const int nblk = noB_ ? 7 : 10;
        // the block pool, the small dense buffers and the constraint blocks
        // live in library workspace kept between solves (the FP32 and FP64
        // phases of one solve never hold theirs at the same time)
        pool_ = scratch(kSlotPool, blk * nblk + 256);
        BX_HIP(hipMemsetAsync(pool_, 0, blk * nblk + 256, stream()));
        char *p = (char *)pool_;
        for (int i = 0; i < nblk; i++) bufs_.push_back(p + i * blk);
        X_ = bufs_[0];
        AX_ = bufs_[1];
        R_ = bufs_[2];
        AR_ = bufs_[3];
        P_ = bufs_[4];
        AP_ = bufs_[5];
        W_ = bufs_[6];
        if (!noB_) {
            BX_ = bufs_[7];
            BR_ = bufs_[8];
            BP_ = bufs_[9];
        }
        const size_t small = (size_t)ld3_ * ld3_ * 2 + (size_t)ld3_ * k + 6 * (size_t)k + 64;
        dsmall_ = (double *)scratch(kSlotSmall, small * sizeof(double) + sizeof(DevStatus));
        GA_ = dsmall_;
        GB_ = GA_ + (size_t)ld3_ * ld3_;
        coef_ = GB_ + (size_t)ld3_ * ld3_;
        lam_ = coef_ + (size_t)ld3_ * k;  // lam_, lam_new_, rn2_ (2k), status: one readback
        lam_new_ = lam_ + k;
        rn2_ = lam_new_ + k;
        st_ = (DevStatus *)(rn2_ + 2 * k);
        tail_ = lam_;
        tail_bytes_ = 4 * (size_t)k * sizeof(double) + sizeof(DevStatus);
        BX_HIP(hipMemsetAsync(dsmall_, 0, small * sizeof(double) + sizeof(DevStatus), stream()));
        if (ny_ > 0) {
            const size_t yb = (size_t)ld_ * ny_ * es_;
            ypool_ = scratch(kSlotY, yb * 2 + ((size_t)ny_ * ny_ + 2 * (size_t)ny_ * k) * 8);
            BX_HIP(hipMemsetAsync(ypool_, 0, yb * 2, stream()));
            Y_ = ypool_;
            BY_ = noB_ ? Y_ : (char *)ypool_ + yb;
            yby_ = (double *)((char *)ypool_ + yb * 2);
            ybx_ = yby_ + (size_t)ny_ * ny_;
            tybx_ = ybx_ + (size_t)ny_ * k;
        }
    }
    ~Phase() { hipStreamSynchronize(stream()); }

    // Returns the exit flag; *it_io is the iteration count so far (in) and
    // after this phase (out).  fp32_phase: stop when FP32 has done its part.
    int run(void *Xuser, int *it_io, int maxit, bool fp32_phase, const hip_LobpcgSettings &set,
            bool *switched);

  private:
    hip_Multi_Vector V(void *p, int w) {
        hip_Multi_Vector v = view(p, n_, w, prec_);
        v.ld = ld_;
        return v;
    }
    void apply(Op op, void *d, void *x, void *y, int w) {
        hip_Multi_Vector vx = V(x, w), vy = V(y, w);
        const bool isA = op == opA_ && d == Ad_;
        double bytes = 0.0;
        if (isA && op == HipCSRMatMultiVec) {
            const hip_CSRMatrix *A = (const hip_CSRMatrix *)d;
            bytes = A->ndiag > 0 ? (double)A->ndiag * A->diag_ld * es_ + 2.0 * n_ * w * es_
                                 : (double)A->nnz * (es_ + 4) + (double)(A->n + 1) * 4 +
                                       2.0 * n_ * w * es_;
        } else if (isA && op == HipDenseMatMultiVec) {
            bytes = d ? (double)n_ * n_ * 8 + 2.0 * n_ * w * es_ : 2.0 * n_ * w * es_;
        }
        tm_->begin(isA ? kSpmm : kOther);
        op(d, &vx, &vy);
        tm_->end(bytes);
    }
    GramBatch gram_batch() { return GramBatch(prec_, ld_, ld3_); }
    void run_gram(GramBatch &gb) {
        tm_->begin(kGram);
        gb.flush();
        tm_->end();
    }
    void update(const UpdGroup *g, int ng, const double *C, int crows, int w,
                const int *tail_pos = nullptr) {
        tm_->begin(kUpdate);
        block_update(prec_, ld_, g, ng, C, ld3_, crows, w, st_, tail_pos);
        tm_->end();
    }
    // x <- x * c (in place), c a w x w block of the Gram buffer
    void rotate(void *x, const double *c, int w) {
        UpdGroup g{};
        g.in[0] = x;
        g.h[0] = w;
        g.out_full = x;
        update(&g, 1, c, w, w);
    }
    void read_tail(std::vector<double> &buf, DevStatus *hs) {
        buf.resize(tail_bytes_ / sizeof(double) + 1);
        download_sync(buf.data(), tail_, tail_bytes_);
        memcpy(hs, (const char *)buf.data() + 4 * (size_t)k_ * sizeof(double), sizeof(DevStatus));
        tm_->resolve();
    }
    // read_tail with the products A R (and B R) of the w columns of R
    // enqueued behind the copy, so that the device computes them while the
    // host reads the tail and decides the next iteration (which uses them if
    // its active set is unchanged).  Not with kernel timing, whose events
    // are resolved after the copy
    void read_tail_ahead(std::vector<double> &buf, DevStatus *hs, int w) {
        buf.resize(tail_bytes_ / sizeof(double) + 1);
        download_start(tail_, tail_bytes_);
        if (!noB_) apply(opB_, Bd_, R_, BR_, w);
        apply(opA_, Ad_, R_, AR_, w);
        download_finish(buf.data(), tail_bytes_);
        memcpy(hs, (const char *)buf.data() + 4 * (size_t)k_ * sizeof(double), sizeof(DevStatus));
    }
    void read_status(DevStatus *h) {
        download_sync(h, st_, sizeof(DevStatus));
        tm_->resolve();
    }
    void gather(void *buf, const std::vector<int> &src) {
        // in place (src ascending, src[t] >= t), one launch
        ColMap m = colmap_identity((int)src.size());
        for (size_t t = 0; t < src.size(); t++) m.idx[t] = src[t];
        compact_cols(prec_, n_, buf, ld_, m);
    }
    // residuals R = BX diag(lam) - AX of the columns in m and their squared
    // norms; with the library's diagonal preconditioner R is also scaled by it
    // here (after the norms), instead of by a separate pass next iteration
    void residuals(const void *BXp, const ColMap &m, const double *lam) {
        const void *t = nullptr;
        if (fuseT_) {
            const hip_DiagMatrix *T = (const hip_DiagMatrix *)Td_;
            t = prec_ == HIPMV_FP32 ? (const void *)T->d32 : (const void *)T->d;
        }
        residual_norm_cols(prec_, n_, BXp, AX_, ld_, m, lam, R_, ld_, t, rn2_);
    }
    void project_constraints(void *blk, int w);
    void store_hist(int col) {
        if (out_.lamh)
            for (int i = 0; i < kw_; i++) out_.lamh[i + (size_t)col * out_.lamh_gh] = out_.lam[i];
        if (out_.resh)
            for (int i = 0; i < kw_; i++) out_.resh[i + (size_t)col * out_.resh_gh] = out_.res[i];
    }
    void dump(int it) {
        printf("\nEigenvalue lambda       Residual              \n");
        for (int i = 0; i < kw_; i++) printf("%22.14e  %22.14e\n", out_.lam[i], out_.res[i]);
        printf("\n%d iterations\n", it);
    }
    double maxres() {
        double m = out_.res[0];
        for (int i = 1; i < kw_; i++) m = out_.res[i] > m ? out_.res[i] : m;
        return m;
    }
    // residual threshold of column i (lobpcg_checkResiduals, lobpcg.c:265-294)
    double threshold(int i) const {
        return fabs(out_.lam[i]) * tol_.relative + tol_.absolute + DBL_EPSILON;
    }

  public:
    // Eigenvalue error bounds of the wanted Ritz values.  For a window [a, b]
    // of consecutive Ritz pairs, sum_{j in [a, b]} (theta_j - lambda_j) <=
    // sum_{j in [a, b]} ||r_j||^2 / delta, delta the distance from the window
    // to the nearest Ritz value outside it (the cluster form of the quadratic
    // residual bound, exact for B = I; Ritz values stand in for the
    // eigenvalues outside the window).  Column i takes the smallest bound over
    // the windows that contain it, at most kWin pairs either side, with b <
    // k - 1: the top pair has no neighbour above, and without guard columns
    // the top wanted pair falls back to the gap below it.  Only columns with
    // need[i] set are evaluated (the others get 0).
    void eig_bounds(const std::vector<char> &need, std::vector<double> &est) const {
        constexpr int kWin = 24;
        const int k = k_;
        const double *lam = out_.lam, *res = out_.res;
        std::vector<double> S(k + 1, 0.0);  // prefix sums of ||r_j||^2
        for (int j = 0; j < k; j++) S[j + 1] = S[j] + res[j] * res[j];
        est.assign(kw_, 0.0);
        for (int i = 0; i < kw_; i++) {
            if (!need[i]) continue;
            double best = HUGE_VAL;
            for (int a = i; a >= 0 && a >= i - kWin; a--) {
                const double below = a > 0 ? lam[a] - lam[a - 1] : HUGE_VAL;
                for (int b = i; b <= k - 2 && b <= i + kWin; b++) {
                    const double gap = fmin(below, lam[b + 1] - lam[b]);
                    if (gap > 0.0) best = fmin(best, (S[b + 1] - S[a]) / gap);
                }
            }
            if (best == HUGE_VAL && i == k - 1 && i > 0 && lam[i] > lam[i - 1])
                best = res[i] * res[i] / (lam[i] - lam[i - 1]);
            est[i] = best;
        }
    }
    double eig_error_bound() const {
        std::vector<char> need(kw_, 1);
        std::vector<double> est;
        eig_bounds(need, est);
        double m = 0.0;
        for (int i = 0; i < kw_; i++)
            m = fmax(m, out_.lam[i] != 0.0 ? est[i] / fabs(out_.lam[i]) : est[i]);
        return m;
    }

  private:

    int prec_;
    long n_;
    int k_, kw_, ny_;
    Op opA_;
    void *Ad_;
    Op opB_;
    void *Bd_;
    Op opT_;
    void *Td_;
    const hip_Multi_Vector *Yu_;
    lobpcg_Tolerance tol_;
    int verb_;
    HostOut out_;
    KTimer *tm_;
    bool noB_, noT_, fuseT_;
    size_t es_;
    int ld3_;
    long ld_;
    void *pool_ = nullptr, *ypool_ = nullptr;
    std::vector<void *> bufs_;
    void *X_, *AX_, *BX_ = nullptr, *R_, *AR_, *BR_ = nullptr, *P_, *AP_, *BP_ = nullptr, *W_;
    void *Y_ = nullptr, *BY_ = nullptr;
    double *dsmall_ = nullptr, *GA_, *GB_, *coef_, *lam_, *lam_new_, *rn2_;
    double *tail_;
    size_t tail_bytes_;
    double *yby_ = nullptr, *ybx_ = nullptr, *tybx_ = nullptr;
    DevStatus *st_;
};

// blk(:, 0:w) -= Y (Y'BY)^{-1} (BY)' blk   (krylov/lobpcg.c:741-750, 996-1012)
void Phase::project_constraints(void *blk, int w) {
    tm_->begin(kOther);
    gram_cols(prec_, n_, BY_, ld_, colmap_identity(ny_), blk, ld_, colmap_identity(w), ybx_, ny_);
    small_mm(1.0, yby_, ny_, true, ybx_, ny_, false, tybx_, ny_, ny_, w, ny_);
    small_mm(-1.0, yby_, ny_, false, tybx_, ny_, false, ybx_, ny_, ny_, w, ny_);
    vecmat_cols(prec_, n_, Y_, ld_, colmap_identity(ny_), ybx_, ny_, blk, ld_, colmap_identity(w),
                true);
    tm_->end();
}

int Phase::run(void *Xuser, int *it_io, int maxit, bool fp32_phase, const hip_LobpcgSettings &set,
               bool *switched) {
    *switched = false;
    const int k = k_;
    DevStatus hs;
    std::vector<double> tail;
    // ---- working copy of X (and Y) in the padded layout ----
    convert_block(n_, k, Xuser, n_, HIPMV_FP64, X_, ld_, prec_);
    if (ny_ > 0) {
        convert_block(n_, ny_, Yu_->data, Yu_->ld, Yu_->precision, Y_, ld_, prec_);
        if (!noB_) apply(opB_, Bd_, Y_, BY_, ny_);
        gram_cols(prec_, n_, BY_, ld_, colmap_identity(ny_), Y_, ld_, colmap_identity(ny_), yby_,
                  ny_);
        chol_inv(yby_, ny_, ny_, st_, kCholX);
        read_status(&hs);
        if (hs.abort) {
            if (verb_) printf("Cannot handle linear dependent constraints\n");
            return WRONG_CONSTRAINTS;
        }
        project_constraints(X_, k);
    }
    int flag = 0;
    // ---- initial Rayleigh-Ritz in span(X) (lobpcg.c:840-960) ----
    if (!noB_) apply(opB_, Bd_, X_, BX_, k);
    void *BXp = noB_ ? X_ : BX_;
    {
        GramBatch gb = gram_batch();
        gb.add(X_, k, BXp, k, GB_);
        run_gram(gb);
    }
    tm_->begin(kDense);
    chol_inv(GB_, ld3_, k, st_, kCholX);
    tm_->end();
    read_status(&hs);
    if (hs.abort) {
        flag = hs.x_info;
        if (verb_) {
            fprintf(stderr, "Error in LOBPCG:\nBad initial vectors: orthonormalization failed\n");
            printf("DPOTRF INFO = %d\n", flag);
        }
    } else {
        rotate(X_, GB_, k);
        if (!noB_) rotate(BX_, GB_, k);
        apply(opA_, Ad_, X_, AX_, k);
        {
            GramBatch gb = gram_batch();
            gb.add(X_, k, AX_, k, GA_);
            gb.add(X_, k, BXp, k, GB_);
            run_gram(gb);
        }
        tm_->begin(kDense);
        rr_solve(GA_, GB_, ld3_, k, 0, 0, 1, lam_, lam_new_, coef_, ld3_, st_);
        tm_->end();
        UpdGroup g[3] = {};
        int ng = 0;
        for (void *b : {X_, AX_, noB_ ? nullptr : BX_}) {
            if (!b) continue;
            g[ng].in[0] = b;
            g[ng].h[0] = k;
            g[ng].out_full = b;
            ng++;
        }
        update(g, ng, coef_, k, k);
        ColMap all = colmap_identity(k);
        tm_->begin(kOther);
        residuals(BXp, all, lam_new_);
        tm_->end();
        read_tail(tail, &hs);
        if (hs.abort) {
            flag = hs.gevp_info;
            if (verb_) {
                fprintf(stderr,
                        "Error in LOBPCG:\nBad problem: Rayleigh-Ritz in the initial subspace failed\n");
                printf("DSYGV INFO = %d\n", flag);
            }
        } else {
            std::swap(lam_, lam_new_);
            const double *lam_h = tail.data() + (lam_ - tail_);
            const double *rn_h = tail.data() + 2 * k;
            for (int i = 0; i < k; i++) {
                out_.lam[i] = lam_h[i];
                out_.res[i] = sqrt(fabs(rn_h[i]));
            }
            store_hist(*it_io);
            if (verb_ == 2)
                dump(*it_io);
            else if (verb_ == 1)
                printf("\nInitial Max. Residual %22.14e\n", maxres());
        }
    }

    // ---- main loop (lobpcg.c:965-1318) ----
    std::vector<int> mask(k, 1), act, prev_act(k);
    for (int i = 0; i < k; i++) prev_act[i] = i;
    int local_it = 0;
    // FP64: refine_iterations more iterations with every column active once
    // all have converged (the first of them recomputes every residual and
    // starts without P)
    int refine_left = fp32_phase ? 0 : set.refine_iterations;
    bool refining = false;
    std::vector<double> gm_hist;  // geometric mean of the k residual norms, per FP32 iteration
    int it = *it_io + 1;
    std::vector<char> need(kw_);
    std::vector<double> est;
    // condition estimates of the pencils (lobpcg.m:488, 518, 912); the first
    // entry is lobpcg.m's start value, large enough to avoid early restarts
    std::vector<double> condh(1, -log10(DBL_EPSILON) / 2.0);
    // A R and B R of the columns of R as the last iteration left them are
    // already enqueued (read_tail_ahead); valid while R is not rewritten
    bool ahead = false;
    const bool may_ahead = !tm_->on() && (noT_ || fuseT_) && ny_ == 0;
    for (; flag == 0 && it <= maxit; it++) {
        act.clear();
        // wanted columns: residual above the threshold (lobpcg_checkResiduals)
        // or, with eig_rtol, the eigenvalue error bound above eig_rtol |lambda|;
        // a column never returns once it has left.  Guard columns are
        // iterated while any wanted column is.
        bool any = false;
        for (int i = 0; i < kw_; i++) need[i] = mask[i] && !(out_.res[i] > threshold(i));
        if (set.eig_rtol > 0.0) eig_bounds(need, est);
        for (int i = 0; i < kw_; i++) {
            bool m = out_.res[i] > threshold(i);
            if (!m && set.eig_rtol > 0.0) m = est[i] > set.eig_rtol * fabs(out_.lam[i]);
            mask[i] = mask[i] && m;
            any = any || mask[i];
        }
        for (int i = kw_; i < k; i++) mask[i] = any;
        for (int i = 0; i < k; i++)
            if (mask[i]) act.push_back(i);
        int nr = (int)act.size();
        if (fp32_phase) {
            // leave FP32 on convergence, near the tolerance, or on a stall
            // (LOBPCG residuals are not monotone: early on they rise for tens
            // of iterations, so progress is judged window against window);
            // only the wanted columns count
            const double mr = maxres();
            double lg = 0.0;
            for (int i = 0; i < kw_; i++) lg += log(out_.res[i] > 1e-300 ? out_.res[i] : 1e-300);
            gm_hist.push_back(exp(lg / kw_));
            const int W = set.stall_iterations, h = (int)gm_hist.size();
            bool stalled = false;
            if (W > 0 && h >= 2 * W) {
                double m1 = HUGE_VAL, m2 = HUGE_VAL;
                for (int t = h - 2 * W; t < h - W; t++) m1 = fmin(m1, gm_hist[t]);
                for (int t = h - W; t < h; t++) m2 = fmin(m2, gm_hist[t]);
                stalled = m2 > 0.98 * m1;
            }
            // near the tolerance: every residual below switch_factor times
            // its own threshold (the mask's |lambda_i| rtol + atol)
            bool near = true;
            for (int i = 0; i < kw_ && near; i++)
                near = out_.res[i] < set.switch_factor *
                                         (fabs(out_.lam[i]) * tol_.relative + tol_.absolute);
            // with eig_rtol, the bounds (quadratic in the residuals) within
            // switch_factor^2 of theirs too
            if (near && set.eig_rtol > 0.0)
                for (int i = 0; i < kw_ && near; i++)
                    near = !need[i] || est[i] <= set.switch_factor * set.switch_factor *
                                                     set.eig_rtol * fabs(out_.lam[i]);
            (void)mr;
            if (nr == 0 || near || stalled) {
                *switched = true;
                fp32_exit_ = stalled ? 1 : 0;
                break;
            }
        }
        if (refining || (nr < 1 && flag == 0 && refine_left > 0)) {
            if (refine_left == 0) break;
            act.resize(k);
            for (int i = 0; i < k; i++) act[i] = i;
            if (!refining) {
                refining = true;
                tm_->begin(kOther);
                residuals(noB_ ? X_ : BX_, colmap_identity(k), lam_);
                tm_->end();
                prev_act = act;
                local_it = 0;
                ahead = false;
            }
            refine_left--;
            refine_done_++;
            nr = k;
        }
        if (nr < 1) break;
        // R, P, AP and BP are laid out by prev_act (the update writes P' only
        // for the columns active when it ran); compact them when columns
        // have converged since (the active set only shrinks)
        if (act != prev_act) {
            ahead = false;
            tm_->begin(kOther);
            std::vector<int> src;
            size_t q = 0;
            for (int c : act) {
                while (prev_act[q] != c) q++;
                src.push_back((int)q);
            }
            gather(R_, src);
            if (local_it > 0) {
                gather(P_, src);
                gather(AP_, src);
                if (!noB_) gather(BP_, src);
            }
            tm_->end();
        }
        // FP32: AX (and BX) follow X by the update recurrence, whose rounding
        // accumulates; recomputing them from X every refresh_interval
        // iterations keeps the computed residuals those of X ("residual
        // replacement"), so FP32 gets close enough to the tolerance that the
        // FP64 phase needs only a few iterations
        if (fp32_phase && set.refresh_interval > 0 && local_it > 0 &&
            local_it % set.refresh_interval == 0) {
            apply(opA_, Ad_, X_, AX_, k);
            if (!noB_) apply(opB_, Bd_, X_, BX_, k);
        }
        if (!noT_ && !fuseT_) {
            apply(opT_, Td_, R_, W_, nr);
            std::swap(R_, W_);
        }
        if (ny_ > 0) project_constraints(R_, nr);
        const int r0 = k, p0 = k + nr;
        if (!ahead) {
            if (!noB_) apply(opB_, Bd_, R_, BR_, nr);
            apply(opA_, Ad_, R_, AR_, nr);
        }
        ahead = false;
        const int preq = local_it > 0 ? nr : 0;
        // explicit X'AX, X'BX: always for FP32 iterates; in FP64 once an
        // active column's residual is below kExplicitGramRes, where the
        // implicit blocks' rounding would dominate (lobpcg.m:692-696; the
        // columns that converge first are the ones whose R block loses
        // accuracy, so the smallest active residual decides).  With them, a
        // pencil that is not positive definite is solved again without P (the
        // steepest-descent restart of lobpcg.m:845-860)
        double min_active = HUGE_VAL;
        for (int t = 0; t < nr; t++) min_active = fmin(min_active, out_.res[act[t]]);
        // likewise once every wanted residual meets its threshold and only
        // the eigenvalue error bound (eig_rtol) keeps the iteration going:
        // the residuals must then fall well below the tolerance, which the
        // implicit blocks do not sustain
        bool bound_only = set.eig_rtol > 0.0;
        for (int i = 0; i < kw_ && bound_only; i++) bound_only = !(out_.res[i] > threshold(i));
        const int explicit_x =
            prec_ == HIPMV_FP32 || min_active <= kExplicitGramRes || bound_only ? 1 : 0;
        void *BRp = noB_ ? R_ : BR_, *BPp = noB_ ? P_ : BP_;
        BXp = noB_ ? X_ : BX_;
        {
            // every Gram block of the iteration in one pass over the rows
            GramBatch gb = gram_batch();
            if (explicit_x) {
                // FP32 iterates drift from X'BX = I, X'AX = diag(lambda): use the
                // computed X blocks (the explicit-Gram variant of lobpcg.m:692-696)
                gb.add(X_, k, AX_, k, GA_);
                gb.add(X_, k, BXp, k, GB_);
            }
            gb.add(R_, nr, AX_, k, GA_ + r0);
            gb.add(R_, nr, AR_, nr, GA_ + r0 + (size_t)r0 * ld3_);
            // with the explicit X blocks, R'BX as the transpose of (BX)'R: for
            // B = I, X then leads 6 right tiles, R 6 and P 12, which makes 8
            // register blocks of 2 x 3 tiles, two on each SIMD
            if (explicit_x)
                gb.add_transposed(BXp, k, R_, nr, GB_ + r0);
            else
                gb.add(R_, nr, BXp, k, GB_ + r0);
            gb.add(R_, nr, BRp, nr, GB_ + r0 + (size_t)r0 * ld3_);
            if (preq) {
                gb.add(P_, nr, AX_, k, GA_ + p0);
                gb.add(P_, nr, AR_, nr, GA_ + p0 + (size_t)r0 * ld3_);
                gb.add(P_, nr, AP_, nr, GA_ + p0 + (size_t)p0 * ld3_);
                gb.add(P_, nr, BXp, k, GB_ + p0);
                gb.add(P_, nr, BRp, nr, GB_ + p0 + (size_t)r0 * ld3_);
                gb.add(P_, nr, BPp, nr, GB_ + p0 + (size_t)p0 * ld3_);
            }
            run_gram(gb);
        }
        // lobpcg.m:846-863: restart (solve without P) when the condition
        // estimate exceeds min(8, max(2, twice its mean over the last
        // 10 + log(block size) iterations))
        double cond_limit = 0.0;
        if (set.cond_restart) {
            const int w = 10 + (int)lround(log((double)nr));
            const int hi = (int)condh.size() - 1, lo = hi - w > 0 ? hi - w : 0;
            double mean = 0.0;
            for (int t = lo; t <= hi; t++) mean += condh[t];
            mean /= (hi - lo + 1);
            cond_limit = fmin(8.0, fmax(2.0 * mean, 2.0));
        }
        tm_->begin(kDense);
        // FP32 iterates: explicit X blocks, and P is dropped rather than the
        // phase abandoned when the basis has lost rank at FP32 accuracy
        // with cond_restart (lobpcg.m's remedies) a pencil that is not
        // positive definite is also solved again without P, in FP64 too;
        // otherwise FP64 stops there as lobpcg.c does (dsygv failure)
        rr_solve_implicit(GA_, GB_, ld3_, k, nr, preq, lam_, lam_new_, coef_, ld3_, st_,
                          explicit_x, explicit_x || set.cond_restart, cond_limit);
        tm_->end();
        // P' = R C_R + P C_P and X' = X C_X + P' (likewise A-, B-blocks), in place
        UpdGroup g[3] = {};
        int ng = 0;
        void *trip[3][3] = {{X_, R_, P_}, {AX_, AR_, AP_}, {BX_, BR_, BP_}};
        for (int q = 0; q < (noB_ ? 2 : 3); q++) {
            g[ng].in[0] = trip[q][0];
            g[ng].h[0] = k;
            g[ng].crow[0] = 0;
            g[ng].in[1] = trip[q][1];
            g[ng].h[1] = nr;
            g[ng].crow[1] = r0;
            if (preq) {
                g[ng].in[2] = trip[q][2];
                g[ng].h[2] = nr;
                g[ng].crow[2] = p0;
            }
            g[ng].out_tail = trip[q][2];
            g[ng].out_full = trip[q][0];
            ng++;
        }
        std::vector<int> tail_pos(k, -1);  // P' kept for the active columns only
        for (int t = 0; t < nr; t++) tail_pos[act[t]] = t;
        if (noB_ && k <= 32) {
            // B = I: the residual of the new block is formed as it is updated
            // (R = T (X' lam - AX') and its norms, for the active columns)
            const void *t = nullptr;
            if (fuseT_) {
                const hip_DiagMatrix *T = (const hip_DiagMatrix *)Td_;
                t = prec_ == HIPMV_FP32 ? (const void *)T->d32 : (const void *)T->d;
            }
            tm_->begin(kUpdate);
            block_update_resid(prec_, n_, ld_, g, coef_, ld3_, k + nr + preq, k, st_,
                               tail_pos.data(), nr, lam_new_, t, R_, rn2_);
            tm_->end();
        } else {
            update(g, ng, coef_, k + nr + preq, k, tail_pos.data());
            BXp = noB_ ? X_ : BX_;
            ColMap am = colmap_identity(nr);
            for (int t = 0; t < nr; t++) am.idx[t] = act[t];
            tm_->begin(kOther);
            residuals(BXp, am, lam_new_);
            tm_->end();
        }
        BXp = noB_ ? X_ : BX_;
        if (may_ahead && it < maxit) {
            read_tail_ahead(tail, &hs, nr);
            ahead = true;
        } else {
            read_tail(tail, &hs);
        }
        if (hs.abort && fp32_phase) {
            // an FP32 breakdown ends the FP32 phase; X is still the last update
            *switched = true;
            fp32_exit_ = 2;
            break;
        }
        if (hs.abort) {
            if (hs.r_info) {
                flag = hs.r_info;
                if (verb_) {
                    fprintf(stderr, "Error in LOBPCG:\nOrthonormalization of residuals failed\n");
                    printf("DPOTRF INFO = %d\n", flag);
                }
            } else {
                flag = hs.gevp_info ? hs.gevp_info : 1;
                if (verb_) {
                    fprintf(stderr, "Error in LOBPCG:\nGEVP solver failure\n");
                    printf("INFO = %d\n", flag);
                }
                it--;
            }
            break;
        }
        if (set.cond_restart) {
            condh.push_back(hs.condest);
            if (hs.p_info == -2) {
                cond_restarts_++;
                if (verb_) printf("Restart on step %d as condestG %5.4e\n", it, hs.condest);
            }
        }
        std::swap(lam_, lam_new_);
        const double *lam_h = tail.data() + (lam_ - tail_);
        const double *rn_h = tail.data() + 2 * k;
        for (int i = 0; i < k; i++) out_.lam[i] = lam_h[i];
        for (int t = 0; t < nr; t++) out_.res[act[t]] = sqrt(fabs(rn_h[t]));
        store_hist(it);
        if (verb_ == 2) {
            printf("Iteration %d \tbsize %d\n", it, nr);
            dump(it);
        } else if (verb_ == 1) {
            printf("Iteration %d \tbsize %d \tmaxres %22.14e\n", it, nr, maxres());
        }
        prev_act = act;
        local_it++;
    }
    *it_io = it - 1;
    // ---- final Rayleigh-Ritz on X (lobpcg.m:922-974): AX and BX from X, the
    // explicit, symmetrised X'AX and X'BX, X <- X C, residuals of every column
    if (set.final_rr && !*switched && flag == 0) {
        if (!noB_) apply(opB_, Bd_, X_, BX_, k);
        BXp = noB_ ? X_ : BX_;
        apply(opA_, Ad_, X_, AX_, k);
        {
            GramBatch gb = gram_batch();
            gb.add(X_, k, AX_, k, GA_);
            gb.add(X_, k, BXp, k, GB_);
            run_gram(gb);
        }
        tm_->begin(kDense);
        rr_solve(GA_, GB_, ld3_, k, 0, 0, 1, lam_, lam_new_, coef_, ld3_, st_);
        tm_->end();
        UpdGroup g[3] = {};
        int ng = 0;
        for (void *b : {X_, AX_, noB_ ? nullptr : BX_}) {
            if (!b) continue;
            g[ng].in[0] = b;
            g[ng].h[0] = k;
            g[ng].out_full = b;
            ng++;
        }
        update(g, ng, coef_, k, k);
        tm_->begin(kOther);
        residuals(BXp, colmap_identity(k), lam_new_);
        tm_->end();
        read_tail(tail, &hs);
        if (hs.abort) {
            flag = hs.gevp_info ? hs.gevp_info : 1;
            if (verb_) fprintf(stderr, "Error in LOBPCG:\nfinal Rayleigh-Ritz failed\n");
        } else {
            std::swap(lam_, lam_new_);
            const double *lam_h = tail.data() + (lam_ - tail_);
            const double *rn_h = tail.data() + 2 * k;
            for (int i = 0; i < k; i++) {
                out_.lam[i] = lam_h[i];
                out_.res[i] = sqrt(fabs(rn_h[i]));
            }
            store_hist(*it_io);  // replaces the last column, as lobpcg.m does
            final_rr_done_ = 1;
        }
    }
    // ---- copy the block back ----
    convert_block(n_, k, X_, ld_, prec_, Xuser, n_, HIPMV_FP64);
    BX_HIP(hipStreamSynchronize(stream()));
    if (*switched) return 0;
    if (flag != 0 || it > maxit) return REQUESTED_ACCURACY_NOT_ACHIEVED;
    return 0;
}

}  // namespace

// the block (with the guard columns its settings ask for) fits the device solver
bool device_lobpcg_fits(int k, int guard) { return k >= 1 && guard >= 0 && 3 * (k + guard) <= kMaxRR; }

int device_lobpcg(mv_MultiVectorPtr X, void *Ad, Op opA, void *Bd, Op opB, void *Td, Op opT,
                  mv_MultiVectorPtr Y, lobpcg_Tolerance tol, int maxit, int verb, int *iters,
                  double *lam, double *lamh, BlopexInt lamh_gh, double *res, double *resh,
                  BlopexInt resh_gh) {
    auto t0 = std::chrono::steady_clock::now();
    hip_Multi_Vector *x = (hip_Multi_Vector *)mv_MultiVectorGetData(X);
    *iters = 0;
    const int k = x->num_vectors;
    if (k < 1) {
        if (verb) fprintf(stderr, "Error in LOBPCG:\nThe bloc size is wrong.\n");
        return WRONG_BLOCK_SIZE;
    }
    if (x->precision != HIPMV_FP64) {
        set_error("lobpcg_solve_double: X must be an FP64 block");
        if (verb) fprintf(stderr, "Error in LOBPCG:\n%s\n", blopex_amd_last_error());
        return WRONG_BLOCK_SIZE;
    }
    const int g = x->settings.guard_columns, kb = k + g;
    if (g < 0 || 3 * kb > kMaxRR) {
        set_error("device solver: block size %d with %d guard columns above %d", k, g,
                  kMaxRR / 3);
        if (verb) fprintf(stderr, "Error in LOBPCG:\n%s\n", blopex_amd_last_error());
        return WRONG_BLOCK_SIZE;
    }
    hip_Multi_Vector *y = nullptr;
    int ny = 0;
    if (Y) {
        if (!blopex_amd_is_hip_interpreter(blopex_amd_mv_interpreter(Y))) {
            set_error("device solver: constraints must be a hip_Multi_Vector");
            if (verb) fprintf(stderr, "Error in LOBPCG:\n%s\n", blopex_amd_last_error());
            return WRONG_CONSTRAINTS;
        }
        y = (hip_Multi_Vector *)mv_MultiVectorGetData(Y);
        ny = y->num_vectors;
        if (y->size != x->size || y->precision != HIPMV_FP64 || ny > kMaxCols) {
            set_error("device solver: constraints must be an FP64 block of %ld rows and at most "
                      "%d columns (got %ld x %d)",
                      x->size, kMaxCols, y->size, ny);
            if (verb) fprintf(stderr, "Error in LOBPCG:\n%s\n", blopex_amd_last_error());
            return WRONG_CONSTRAINTS;
        }
    }
    const hip_LobpcgSettings set = x->settings;
    memset(&x->stats, 0, sizeof(x->stats));
    if (set.size_guard && x->size - ny < 5L * kb) {
        // lobpcg.m:475-484 (disabled in lobpcg.c:672-687)
        set_error("problem size %ld (minus %d constraints) below 5 x %d columns", x->size, ny, kb);
        if (verb) fprintf(stderr, "Error in LOBPCG:\nProblem size too small compared to block size\n");
        return PROBLEM_SIZE_TOO_SMALL;
    }
    // the phases iterate kb columns: the caller's k, then the guard columns
    // (2 drand48() - 1 after srand48(guard_seed), column by column); the
    // caller's arrays receive the first k
    std::vector<double> lam_w(kb), res_w(kb);
    HostOut out{lam_w.data(), lamh, res_w.data(), resh, (int)lamh_gh, (int)resh_gh};
    void *Xw = x->data;
    if (g > 0 || x->ld != x->size) {
        Xw = scratch(kSlotXw, (size_t)x->size * kb * sizeof(double));
        convert_block(x->size, k, x->data, x->ld, HIPMV_FP64, Xw, x->size, HIPMV_FP64);
        if (g > 0) {
            ColMap gm = colmap_identity(g);
            for (int j = 0; j < g; j++) gm.idx[j] = k + j;
            drand48_cols(HIPMV_FP64, x->size, set.guard_seed, Xw, x->size, gm);
        }
    }
    if (verb) {
        printf("\nSolving %s eigenvalue problem with%s preconditioning\n\n",
               opB ? "generalized" : "standard", opT ? "" : "out");
        printf("block size %d\n\n", k);
        if (ny == 0)
            printf("No constraints\n\n");
        else
            printf("%d constraint%s\n\n", ny, ny > 1 ? "s" : "");
    }
    int it = 0, flag = 0;
    bool switched = false;
    KTimer tm(set.profile != 0);
    if (set.mode == HIPLOBPCG_MIXED) {
        Phase ph(HIPMV_FP32, x->size, kb, k, ny, opA, Ad, opB, Bd, opT, Td, y, tol, verb, out,
                 &tm);
        flag = ph.run(Xw, &it, maxit, true, set, &switched);
        x->stats.iterations_fp32 = it;
        x->stats.fp32_exit = ph.fp32_exit_;
        x->stats.cond_restarts += ph.cond_restarts_;
    }
    if (set.mode != HIPLOBPCG_MIXED || (switched && flag == 0)) {
        const int it0 = it;
        Phase ph(HIPMV_FP64, x->size, kb, k, ny, opA, Ad, opB, Bd, opT, Td, y, tol, verb, out,
                 &tm);
        bool dummy;
        flag = ph.run(Xw, &it, maxit, false, set, &dummy);
        x->stats.iterations_fp64 = it - it0;
        x->stats.refine_iterations = ph.refine_done_;
        x->stats.final_rr = ph.final_rr_done_;
        x->stats.cond_restarts += ph.cond_restarts_;
        if (set.eig_rtol > 0.0) x->stats.eig_error_bound = ph.eig_error_bound();
    }
    for (int i = 0; i < k; i++) {
        lam[i] = lam_w[i];
        res[i] = res_w[i];
    }
    if (Xw != x->data)
        convert_block(x->size, k, Xw, x->size, HIPMV_FP64, x->data, x->ld, HIPMV_FP64);
    x->stats.guard_columns = g;
    *iters = it;
    if (verb == 1) {
        printf("\nEigenvalue lambda       Residual              \n");
        for (int i = 0; i < k; i++) printf("%22.14e  %22.14e\n", lam[i], res[i]);
        printf("\n%d iterations\n", it);
    }
    BX_HIP(hipStreamSynchronize(stream()));
    tm.resolve();
    tm.add_to(&x->stats);
    x->stats.exit_flag = flag;
    x->stats.seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return flag;
}

}  // namespace bx
