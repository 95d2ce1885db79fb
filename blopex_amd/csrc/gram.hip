// Fused tall-skinny Gram products on MFMA: one pass over the n rows computes
// every requested L' Q block of a LOBPCG iteration (the 8-10
// MultiInnerProd calls per iteration of krylov/lobpcg.c:1092-1165, each a
// column-by-column triple loop in multi_vector.c:354-406).
//
// A job is one 16-column tile of a left block L and up to kGramQ tiles of
// right blocks; a wave runs one job over 64-row chunks:
//   * the lane holding (kslot = l/16, column l%16) loads 16 consecutive rows
//     r0 + 16*kslot .. +15 of its column with vector loads -- the MFMA sums
//     over the 4 k-slots of each step, and any row order that L and Q share
//     gives the same sum, so the 64 rows are assigned as 4 slots x 16 steps;
//   * 16 MFMA steps per (L tile, Q tile) pair accumulate D = L' Q (16x16).
// FP64 inputs accumulate on FP64 MFMA throughout.  FP32 inputs use FP32
// MFMA for one 64-row chunk (products exact in the FP32 accumulator up to its
// rounding over 64 terms) and add each chunk's partial into FP64 registers,
// so FP32 rounding never accumulates over more than 64 rows.
// Each workgroup writes one FP64 partial per output tile; a second kernel
// sums the partials in a fixed order (bitwise reproducible results).
#include "kernels.h"
#include "mfma.h"

namespace bx {

namespace {

constexpr int kGTh = 256;

struct GramKArgs {
    GramPlan p;
    long ld;
    double *part;  // [gridDim.x][ntiles][256]
};

template <typename T>
__device__ __forceinline__ void load16(const T *p, bool ok, T v[16]) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (ok) {
            load4(p + 4 * q, v + 4 * q);
        } else {
            v[4 * q] = v[4 * q + 1] = v[4 * q + 2] = v[4 * q + 3] = T(0);
        }
    }
}

template <typename T>
struct GramAcc;

template <>
struct GramAcc<double> {
    // accumulate straight into the FP64 accumulator
    static __device__ __forceinline__ void chunk(const double *lv, const double *qv, f64x4 &acc) {
#pragma unroll
        for (int s = 0; s < 16; s++) acc = Mfma<double>::mma(lv[s], qv[s], acc);
    }
};

template <>
struct GramAcc<float> {
    static __device__ __forceinline__ void chunk(const float *lv, const float *qv, f64x4 &acc) {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 16; s++) c = Mfma<float>::mma(lv[s], qv[s], c);
        acc[0] += (double)c[0];
        acc[1] += (double)c[1];
        acc[2] += (double)c[2];
        acc[3] += (double)c[3];
    }
};

template <typename T>
__global__ void __launch_bounds__(kGTh) k_gram_jobs(GramKArgs a) {
    const GramPlan &p = a.p;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const long nchunks = a.ld / 64;
    for (int j = wave; j < p.njobs; j += kGTh / 64) {
        const int lb = p.lblk[j];
        const int lcol = p.ltile[j] * 16 + li;
        const bool lok = lcol < p.blkw[lb];
        const T *L = (const T *)p.blk[lb] + (size_t)(lok ? lcol : 0) * a.ld;
        const int nq = p.nq[j];
        f64x4 acc[kGramQ];
#pragma unroll
        for (int q = 0; q < kGramQ; q++) acc[q] = f64x4{0.0, 0.0, 0.0, 0.0};
        for (long ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
            const long rb = ch * 64 + 16 * lk;
            T lv[16];
            load16(L + rb, lok, lv);
#pragma unroll
            for (int q = 0; q < kGramQ; q++) {
                if (q < nq) {
                    const int qb = p.qblk[j][q];
                    const int qcol = p.qtile[j][q] * 16 + li;
                    const bool qok = qcol < p.blkw[qb];
                    const T *Q = (const T *)p.blk[qb] + (size_t)(qok ? qcol : 0) * a.ld;
                    T qv[16];
                    load16(Q + rb, qok, qv);
                    GramAcc<T>::chunk(lv, qv, acc[q]);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kGramQ; q++) {
            if (q < nq) {
                double *dst = a.part + ((size_t)blockIdx.x * p.ntiles + p.qout[j][q]) * 256;
#pragma unroll
                for (int r = 0; r < 4; r++) dst[Mfma<T>::drow(lk, r) * 16 + li] = acc[q][r];
            }
        }
    }
}

struct ReduceArgs {
    double *dst[kGramOut];
    int rows[kGramOut], cols[kGramOut];
    int ntiles;
    int nparts;
    long ldg;
    const double *part;
};

__global__ void __launch_bounds__(256) k_gram_sum(ReduceArgs a) {
    const int t = blockIdx.x, e = threadIdx.x;
    const int i = e / 16, jj = e % 16;
    const double *src = a.part + (size_t)t * 256 + e;
    const size_t stride = (size_t)a.ntiles * 256;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int p = 0;
    for (; p + 4 <= a.nparts; p += 4) {
        s0 += src[(size_t)p * stride];
        s1 += src[(size_t)(p + 1) * stride];
        s2 += src[(size_t)(p + 2) * stride];
        s3 += src[(size_t)(p + 3) * stride];
    }
    for (; p < a.nparts; p++) s0 += src[(size_t)p * stride];
    const double s = (s0 + s1) + (s2 + s3);
    if (i < a.rows[t] && jj < a.cols[t]) a.dst[t][i + (size_t)jj * a.ldg] = s;
}

}  // namespace

namespace {
int plan_block(GramPlan &p, const void *blk, int width) {
    for (int b = 0; b < p.nblk; b++)
        if (p.blk[b] == blk && p.blkw[b] == width) return b;
    if (p.nblk >= kGramBlocks) return -1;
    p.blk[p.nblk] = blk;
    p.blkw[p.nblk] = width;
    return p.nblk++;
}
}  // namespace

bool gram_plan_product(GramPlan &p0, const void *L, int lw, const void *Q, int qw, double *g,
                       long ldg) {
    GramPlan p = p0;  // work on a copy: commit only if everything fits
    const int lb = plan_block(p, L, lw), qb = plan_block(p, Q, qw);
    if (lb < 0 || qb < 0) return false;
    const int lt = (lw + 15) / 16, qt = (qw + 15) / 16;
    for (int a = 0; a < lt; a++) {
        for (int b = 0; b < qt; b++) {
            int j = -1;
            for (int c = 0; c < p.njobs; c++)
                if (p.lblk[c] == lb && p.ltile[c] == a && p.nq[c] < kGramQ) j = c;
            if (j < 0) {
                if (p.njobs >= kGramJobs) return false;
                j = p.njobs++;
                p.lblk[j] = lb;
                p.ltile[j] = a;
                p.nq[j] = 0;
            }
            if (p.ntiles >= kGramOut) return false;
            const int t = p.ntiles++;
            p.qblk[j][p.nq[j]] = qb;
            p.qtile[j][p.nq[j]] = b;
            p.qout[j][p.nq[j]] = t;
            p.nq[j]++;
            p.dst[t] = g + 16 * a + (size_t)16 * b * ldg;
            p.rows[t] = lw - 16 * a < 16 ? lw - 16 * a : 16;
            p.cols[t] = qw - 16 * b < 16 ? qw - 16 * b : 16;
        }
    }
    p0 = p;
    return true;
}

void GramBatch::add(const void *L, int lw, const void *Q, int qw, double *g) {
    if (lw == 0 || qw == 0) return;
    if (gram_plan_product(p_, L, lw, Q, qw, g, ldg_)) return;
    flush();
    BX_REQUIRE(gram_plan_product(p_, L, lw, Q, qw, g, ldg_),
               "Gram product of %d x %d columns does not fit one plan", lw, qw);
}

void GramBatch::flush() {
    if (p_.njobs) gram_run(prec_, ld_, p_, ldg_);
    p_ = GramPlan();
}

void gram_run(int prec, long ld, GramPlan &p, long ldg) {
    BX_REQUIRE(ld % 64 == 0, "gram_run: leading dimension %ld not a multiple of 64", ld);
    if (p.njobs == 0) return;
    const long nchunks = ld / 64;
    long parts = (long)num_cus() * 2;
    if (parts > nchunks) parts = nchunks;
    if (parts < 1) parts = 1;
    double *part = (double *)scratch(0, (size_t)parts * p.ntiles * 256 * sizeof(double));
    GramKArgs a;
    a.p = p;
    a.ld = ld;
    a.part = part;
    if (prec == HIPMV_FP32)
        hipLaunchKernelGGL(k_gram_jobs<float>, dim3((unsigned)parts), dim3(kGTh), 0, stream(), a);
    else
        hipLaunchKernelGGL(k_gram_jobs<double>, dim3((unsigned)parts), dim3(kGTh), 0, stream(), a);
    BX_HIP(hipGetLastError());
    ReduceArgs r;
    for (int t = 0; t < p.ntiles; t++) {
        r.dst[t] = p.dst[t];
        r.rows[t] = p.rows[t];
        r.cols[t] = p.cols[t];
    }
    r.ntiles = p.ntiles;
    r.nparts = (int)parts;
    r.ldg = ldg;
    r.part = part;
    hipLaunchKernelGGL(k_gram_sum, dim3(p.ntiles), dim3(256), 0, stream(), r);
    BX_HIP(hipGetLastError());
}

}  // namespace bx
