// Fused tall-skinny Gram products on MFMA: one pass over the n rows computes
// every requested L' Q block of a LOBPCG iteration (the 8-10
// MultiInnerProd calls per iteration of krylov/lobpcg.c:1092-1165, each a
// column-by-column triple loop in multi_vector.c:354-406).
//
// Workgroups (8 waves) take chunks of rows round-robin (32 rows FP64, 64
// FP32), so at any time they stream one contiguous window of every column.
// A chunk of each distinct block of the plan is copied into LDS with 16-byte
// coalesced loads (column-major, stride an odd number of 16-byte pieces);
// while one chunk is multiplied, the next is written into a second LDS stage
// and the one after is in flight in registers.  Each wave owns a
// register block of up to 2 L tiles x 4 Q tiles (16 columns each): per half
// chunk a lane reads its rows of those tiles with 16-byte LDS reads and runs
// the tile products on MFMA 16x16x4.  The k-index of the MFMA runs over the
// rows of the chunk lane-group-major (lane group g = lane / 16 holds rows
// g*RC/4 .. g*RC/4 + RC/4 - 1), the same for both operands.
// FP64 inputs accumulate in FP64 throughout; FP32 inputs use FP32 MFMA
// within a chunk and add each chunk's tile into FP64 registers, so FP32
// rounding never spans more than 64 rows.  Each workgroup writes one FP64
// partial per tile; a second kernel sums the partials in a fixed order
// (bitwise reproducible results).
#include "kernels.h"
#include "mfma.h"

namespace bx {

namespace {

constexpr int kGTh = kGramWaves * 64;  // threads per workgroup
constexpr int kSlots = kGramBL * kGramBQ;  // partial-buffer slots per wave

template <typename T>
struct GramCfg;
template <>
struct GramCfg<double> {
    static constexpr int RC = 32;  // rows per chunk
};
template <>
struct GramCfg<float> {
    static constexpr int RC = 64;
};

struct GramKArgs {
    GramPlan p;
    long ld;
    double *part;  // [gridDim.x][kGramTiles][256]
};

// operand vector of one lane for one tile and half a chunk: RC/8
// consecutive rows (two 16-byte LDS reads)
template <typename T>
struct Rows;
template <>
struct Rows<double> {
    static constexpr int N = 4;
    double v[N];
    __device__ __forceinline__ void load(const double *p) {
#pragma unroll
        for (int q = 0; q < N / 2; q++) {
            const double2 x = reinterpret_cast<const double2 *>(p)[q];
            v[2 * q] = x.x;
            v[2 * q + 1] = x.y;
        }
    }
};
template <>
struct Rows<float> {
    static constexpr int N = 8;
    float v[N];
    __device__ __forceinline__ void load(const float *p) {
#pragma unroll
        for (int q = 0; q < N / 4; q++) {
            const float4 x = reinterpret_cast<const float4 *>(p)[q];
            v[4 * q] = x.x;
            v[4 * q + 1] = x.y;
            v[4 * q + 2] = x.z;
            v[4 * q + 3] = x.w;
        }
    }
};

// The products of one staged chunk for an NL x NQ register block.
template <typename T, int NL, int NQ>
struct GramBlock;

template <int NL, int NQ>
struct GramBlock<double, NL, NQ> {
    // NH: steps of Rows::N rows per lane group in the chunk (RC = 4 N NH)
    template <int BQ, int NH>
    static __device__ __forceinline__ void run(const double *stage, const int *lo, const int *qo,
                                               f64x4 (&acc)[kGramBL * BQ]) {
#pragma unroll
        for (int h = 0; h < NH; h++) {
            Rows<double> x[NL], y[NQ];
#pragma unroll
            for (int i = 0; i < NL; i++) x[i].load(stage + lo[i] + h * Rows<double>::N);
#pragma unroll
            for (int j = 0; j < NQ; j++) y[j].load(stage + qo[j] + h * Rows<double>::N);
#pragma unroll
            for (int st = 0; st < Rows<double>::N; st++)
#pragma unroll
                for (int i = 0; i < NL; i++)
#pragma unroll
                    for (int j = 0; j < NQ; j++)
                        acc[i * BQ + j] = Mfma<double>::mma(x[i].v[st], y[j].v[st], acc[i * BQ + j]);
        }
    }
};

template <int NL, int NQ>
struct GramBlock<float, NL, NQ> {
    // NH: steps of Rows::N rows per lane group in the chunk (RC = 4 N NH); the
    // FP32 tiles are added into the FP64 ones after every two steps (64 rows)
    template <int BQ, int NH>
    static __device__ __forceinline__ void run(const float *stage, const int *lo, const int *qo,
                                               f64x4 (&acc)[kGramBL * BQ]) {
      static_assert(NH % 2 == 0, "whole 64-row groups");
#pragma unroll
      for (int h0 = 0; h0 < NH; h0 += 2) {
        f32x4 c[NL * NQ];
#pragma unroll
        for (int t = 0; t < NL * NQ; t++) c[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int h = h0; h < h0 + 2; h++) {
            Rows<float> x[NL], y[NQ];
#pragma unroll
            for (int i = 0; i < NL; i++) x[i].load(stage + lo[i] + h * Rows<float>::N);
#pragma unroll
            for (int j = 0; j < NQ; j++) y[j].load(stage + qo[j] + h * Rows<float>::N);
#pragma unroll
            for (int st = 0; st < Rows<float>::N; st++)
#pragma unroll
                for (int i = 0; i < NL; i++)
#pragma unroll
                    for (int j = 0; j < NQ; j++)
                        c[i * NQ + j] = Mfma<float>::mma(x[i].v[st], y[j].v[st], c[i * NQ + j]);
        }
#pragma unroll
        for (int i = 0; i < NL; i++)
#pragma unroll
            for (int j = 0; j < NQ; j++)
#pragma unroll
                for (int r = 0; r < 4; r++) acc[i * BQ + j][r] += (double)c[i * NQ + j][r];
      }
    }
};

template <typename T, int BQ, int NH>
__device__ __forceinline__ void gram_block(int nl, int nq, const T *stage, const int *lo,
                                           const int *qo, f64x4 (&acc)[kGramBL * BQ]) {
    switch (nl * 8 + nq) {  // wave-uniform
        case 9: GramBlock<T, 1, 1>::template run<BQ, NH>(stage, lo, qo, acc); break;
        case 10: GramBlock<T, 1, 2>::template run<BQ, NH>(stage, lo, qo, acc); break;
        case 11: GramBlock<T, 1, 3>::template run<BQ, NH>(stage, lo, qo, acc); break;
        case 17: GramBlock<T, 2, 1>::template run<BQ, NH>(stage, lo, qo, acc); break;
        case 18: GramBlock<T, 2, 2>::template run<BQ, NH>(stage, lo, qo, acc); break;
        case 19: GramBlock<T, 2, 3>::template run<BQ, NH>(stage, lo, qo, acc); break;
        default:
            if constexpr (BQ == 4) {
                if (nl == 1) GramBlock<T, 1, 4>::template run<BQ, NH>(stage, lo, qo, acc);
                if (nl == 2) GramBlock<T, 2, 4>::template run<BQ, NH>(stage, lo, qo, acc);
            }
            break;
    }
}

// PRE: 16-byte pieces per thread and chunk (6: plans up to 192 columns at
// the standard chunk, two LDS stages; 9: up to 288 columns, one); BQ:
// right tiles per register block (3 or 4); RC: rows per chunk (the standard
// one, or twice that for narrow plans, so that each chunk round trip moves
// more bytes)
template <typename T, int PRE, int BQ, int RC>
__global__ void __launch_bounds__(kGTh) k_gram(GramKArgs a) {
    constexpr int NH = RC / (4 * Rows<T>::N);
    // PRE <= 6: two LDS stages; a chunk is multiplied out of one while the
    // next is written into the other, one barrier per chunk.  PRE 9: one
    // stage (LDS would not hold two), store - barrier - multiply - barrier.
    constexpr bool kDB = PRE <= 6;
    constexpr int EV = 16 / sizeof(T);  // elements per 16-byte piece
    constexpr int S = RC + EV;          // column stride: an odd number of 16-byte pieces
    constexpr int VPC = RC / EV;                     // pieces per column per chunk
    static_assert(PRE <= 9, "Gram staging: too many pieces per thread");
    extern __shared__ __align__(16) unsigned char smem[];
    const GramPlan &p = a.p;
    T *stage = reinterpret_cast<T *>(smem);
    const int sbuf = (p.ncols + 16) * S;  // elements of one stage
    const T **colp =
        reinterpret_cast<const T **>(smem + (size_t)(kDB ? 2 : 1) * sbuf * sizeof(T));
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar branches
    const int li = lane & 15, lk = lane >> 4;
    for (int c = tid; c < p.ncols; c += kGTh) {
        int b = 0;
        while (b + 1 < p.nblk && c >= p.bc0[b + 1]) b++;
        colp[c] = (const T *)p.blk[b] + (size_t)(c - p.bc0[b]) * a.ld;
    }
    __syncthreads();

    const int nvec = p.ncols * VPC;
    const long nch = a.ld / RC;
    const long gs = gridDim.x;
    const bool act = wave < p.nwb;
    const int nl = act ? p.wnl[wave] : 0, nq = act ? p.wnq[wave] : 0;
    int lo[kGramBL], qo[BQ];
#pragma unroll
    for (int i = 0; i < kGramBL; i++) lo[i] = ((act ? p.wl[wave][i] : 0) + li) * S + lk * (RC / 4);
#pragma unroll
    for (int j = 0; j < BQ; j++) qo[j] = ((act ? p.wq[wave][j] : 0) + li) * S + lk * (RC / 4);
    f64x4 acc[kGramBL * BQ];
#pragma unroll
    for (int s = 0; s < kGramBL * BQ; s++) acc[s] = f64x4{0.0, 0.0, 0.0, 0.0};

    // the (up to 9) 16-byte pieces this thread copies per chunk, in named
    // registers (an array of them would live in scratch memory); pieces past
    // the plan load piece 0 and land in the spare columns
    uint4 p0, p1, p2, p3, p4, p5, p6, p7, p8;
#define BX_GRAM_P(X) X(0, p0) X(1, p1) X(2, p2) X(3, p3) X(4, p4) X(5, p5) X(6, p6) X(7, p7) X(8, p8)
#define BX_GRAM_LOAD(q, r)                                                                    \
    if (q < PRE) {                                                                            \
        const int e = tid + q * kGTh < nvec ? tid + q * kGTh : 0;                             \
        r = *reinterpret_cast<const uint4 *>(colp[e / VPC] + chl * RC + (e % VPC) * EV);      \
    }
#define BX_GRAM_STORE(q, r)                                                                   \
    if (q < PRE) {                                                                            \
        const int e = tid + q * kGTh;                                                         \
        const int o = e < nvec ? (e / VPC) * S + (e % VPC) * EV : (p.ncols + (tid & 15)) * S; \
        *reinterpret_cast<uint4 *>(dstg + o) = r;                                             \
    }
    const long c0 = blockIdx.x;
    if (c0 < nch) {
        const long chl = c0;
        BX_GRAM_P(BX_GRAM_LOAD)
    }
    if (kDB) {
        // chunk ch is in stage cur; the registers hold chunk ch + gs
        if (c0 < nch) {
            T *dstg = stage;
            BX_GRAM_P(BX_GRAM_STORE)
        }
        if (c0 + gs < nch) {
            const long chl = c0 + gs;
            BX_GRAM_P(BX_GRAM_LOAD)
        }
        __syncthreads();
        int cur = 0;
        for (long ch = c0; ch < nch; ch += gs) {
            if (ch + gs < nch) {
                T *dstg = stage + (cur ^ 1) * sbuf;
                BX_GRAM_P(BX_GRAM_STORE)
            }
            if (ch + 2 * gs < nch) {
                const long chl = ch + 2 * gs;
                BX_GRAM_P(BX_GRAM_LOAD)
            }
            gram_block<T, BQ, NH>(nl, nq, stage + cur * sbuf, lo, qo, acc);
            __syncthreads();
            cur ^= 1;
        }
    } else {
        T *dstg = stage;
        for (long ch = c0; ch < nch; ch += gs) {
            BX_GRAM_P(BX_GRAM_STORE)
            __syncthreads();
            if (ch + gs < nch) {
                const long chl = ch + gs;
                BX_GRAM_P(BX_GRAM_LOAD)
            }
            gram_block<T, BQ, NH>(nl, nq, stage, lo, qo, acc);
            __syncthreads();
        }
    }
#undef BX_GRAM_P
#undef BX_GRAM_STORE
#undef BX_GRAM_LOAD
    if (!act) return;
    double *base = a.part + ((size_t)blockIdx.x * kGramTiles + (size_t)wave * kSlots) * 256;
#pragma unroll
    for (int i = 0; i < kGramBL; i++)
#pragma unroll
        for (int j = 0; j < BQ; j++) {
            if (i >= nl || j >= nq) continue;
            double *t = base + (i * kGramBQ + j) * 256;
#pragma unroll
            for (int r = 0; r < 4; r++) t[Mfma<T>::drow(lk, r) * 16 + li] = acc[i * BQ + j][r];
        }
}

struct ReduceArgs {
    double *dst[kGramTiles], *mdst[kGramTiles];
    unsigned char rows[kGramTiles], cols[kGramTiles];
    unsigned char slot[kGramTiles];
    int nparts;
    long ldg;
    const double *part;
};

// one workgroup per stored tile: sum of the per-workgroup partials in a fixed order
__global__ void __launch_bounds__(256) k_gram_sum(ReduceArgs a) {
    const int t = blockIdx.x, e = threadIdx.x;
    const int i = e / 16, jj = e % 16;
    const double *src = a.part + (size_t)a.slot[t] * 256 + e;
    const size_t stride = (size_t)kGramTiles * 256;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int q = 0;
    // sixteen loads in flight, added in the same order as four at a time
    for (; q + 16 <= a.nparts; q += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = src[(size_t)(q + u) * stride];
#pragma unroll
        for (int u = 0; u < 16; u += 4) {
            s0 += v[u];
            s1 += v[u + 1];
            s2 += v[u + 2];
            s3 += v[u + 3];
        }
    }
    for (; q + 4 <= a.nparts; q += 4) {
        s0 += src[(size_t)q * stride];
        s1 += src[(size_t)(q + 1) * stride];
        s2 += src[(size_t)(q + 2) * stride];
        s3 += src[(size_t)(q + 3) * stride];
    }
    for (; q < a.nparts; q++) s0 += src[(size_t)q * stride];
    const double s = (s0 + s1) + (s2 + s3);
    if (i < a.rows[t] && jj < a.cols[t]) {
        if (a.dst[t]) a.dst[t][i + (size_t)jj * a.ldg] = s;
        if (a.mdst[t]) a.mdst[t][jj + (size_t)i * a.ldg] = s;
    }
}

int plan_block(GramPlan &p, const void *blk, int width) {
    for (int b = 0; b < p.nblk; b++)
        if (p.blk[b] == blk && p.bw[b] == width) return b;
    if (p.nblk >= kGramBlocks || p.ncols + width > kGramCols) return -1;
    p.blk[p.nblk] = blk;
    p.bw[p.nblk] = width;
    p.bc0[p.nblk] = p.ncols;
    p.ncols += width;
    return p.nblk++;
}

// Forms the register blocks from the queued products: for each left block,
// its 16-column tiles in pairs, against all right tiles of its products in
// groups of bq.  Returns the number of register blocks (may exceed
// kGramWaves).  When filling, blocks are split while fewer than kGramWaves
// (the largest first: a pair of left tiles into two single ones, else its
// right tiles into halves), then dealt to the waves so that the two waves of
// each SIMD (waves w and w + 4) carry about equal numbers of tiles: the
// kernel is bound by the busiest SIMD's MFMA work.
int form_blocks(GramPlan &p, long ldg, bool fill, int bq) {
    struct Blk {
        int lb, x0, nl, nq;
        int q[kGramBQ], y[kGramBQ];  // product and right tile of each column of tiles
    };
    Blk blks[2 * kGramWaves];
    int nb = 0, nw = 0;
    bool done[kGramBlocks] = {};
    for (int q0 = 0; q0 < p.nprod; q0++) {
        const int lb = p.pl[q0];
        if (done[lb]) continue;
        done[lb] = true;
        // right tiles of every product with this left block
        int qtp[64], qty[64], nqt = 0;
        for (int q = 0; q < p.nprod; q++) {
            if (p.pl[q] != lb) continue;
            const int qt = (p.bw[p.pq[q]] + 15) / 16;
            for (int y = 0; y < qt && nqt < 64; y++) {
                qtp[nqt] = q;
                qty[nqt] = y;
                nqt++;
            }
        }
        const int lt = (p.bw[lb] + 15) / 16;
        for (int x0 = 0; x0 < lt; x0 += kGramBL) {
            const int nl = lt - x0 < kGramBL ? lt - x0 : kGramBL;
            for (int j0 = 0; j0 < nqt; j0 += bq) {
                const int nq = nqt - j0 < bq ? nqt - j0 : bq;
                if (fill && nb < kGramWaves) {
                    Blk &b = blks[nb++];
                    b.lb = lb;
                    b.x0 = x0;
                    b.nl = nl;
                    b.nq = nq;
                    for (int j = 0; j < nq; j++) {
                        b.q[j] = qtp[j0 + j];
                        b.y[j] = qty[j0 + j];
                    }
                }
                nw++;
            }
        }
    }
    if (!fill) return nw;
    BX_REQUIRE(nw <= kGramWaves, "form_blocks: %d register blocks", nw);
    // split the largest blocks into the spare waves
    while (nb < kGramWaves) {
        int big = -1;
        for (int i = 0; i < nb; i++)
            if (blks[i].nl * blks[i].nq > 1 &&
                (big < 0 || blks[i].nl * blks[i].nq > blks[big].nl * blks[big].nq))
                big = i;
        if (big < 0) break;
        Blk &a = blks[big], &b = blks[nb++];
        b = a;
        if (a.nl == 2) {
            a.nl = b.nl = 1;
            b.x0 = a.x0 + 1;
        } else {
            const int h = (a.nq + 1) / 2;
            for (int j = h; j < a.nq; j++) {
                b.q[j - h] = a.q[j];
                b.y[j - h] = a.y[j];
            }
            b.nq = a.nq - h;
            a.nq = h;
        }
    }
    // largest first, each onto the SIMD with the least work and a free wave
    int order[kGramWaves], load[4] = {}, used[4] = {}, wave_of[kGramWaves];
    for (int i = 0; i < nb; i++) order[i] = i;
    for (int i = 1; i < nb; i++)
        for (int j = i; j > 0 && blks[order[j]].nl * blks[order[j]].nq >
                                     blks[order[j - 1]].nl * blks[order[j - 1]].nq; j--)
            std::swap(order[j], order[j - 1]);
    for (int t = 0; t < nb; t++) {
        int best = -1;
        for (int m = 0; m < 4; m++)
            if (used[m] < kGramWaves / 4 && (best < 0 || load[m] < load[best])) best = m;
        const Blk &b = blks[order[t]];
        wave_of[order[t]] = best + 4 * used[best];
        used[best]++;
        load[best] += b.nl * b.nq;
    }
    // waves without a block sit idle (nl = nq = 0)
    int wmax = 0;
    for (int i = 0; i < nb; i++) wmax = std::max(wmax, wave_of[i] + 1);
    for (int w = 0; w < wmax; w++) {
        p.wnl[w] = p.wnq[w] = 0;
        for (int i = 0; i < kGramBL; i++) p.wl[w][i] = 0;
        for (int j = 0; j < kGramBQ; j++) p.wq[w][j] = 0;
        for (int t = 0; t < kSlots; t++) {
            const int s = w * kSlots + t;
            p.dst[s] = p.mdst[s] = nullptr;
            p.rows[s] = p.cols[s] = 0;
        }
    }
    for (int bi = 0; bi < nb; bi++) {
        const Blk &b = blks[bi];
        const int w = wave_of[bi], lb = b.lb, lw = p.bw[lb];
        p.wnl[w] = (unsigned char)b.nl;
        p.wnq[w] = (unsigned char)b.nq;
        for (int i = 0; i < kGramBL; i++)
            p.wl[w][i] = (short)(p.bc0[lb] + 16 * (b.x0 + (i < b.nl ? i : 0)));
        for (int j = 0; j < kGramBQ; j++) {
            const int jj = j < b.nq ? j : 0;
            p.wq[w][j] = (short)(p.bc0[p.pq[b.q[jj]]] + 16 * b.y[jj]);
        }
        for (int i = 0; i < b.nl; i++)
            for (int j = 0; j < b.nq; j++) {
                const int s = (w * kGramBL + i) * kGramBQ + j;
                const int q = b.q[j], x = b.x0 + i, y = b.y[j];
                const int qw = p.bw[p.pq[q]];
                const bool sym = p.pq[q] == lb;
                if (sym && y < x) continue;  // stored as the transpose of (y, x)
                p.rows[s] = (unsigned char)(lw - 16 * x < 16 ? lw - 16 * x : 16);
                p.cols[s] = (unsigned char)(qw - 16 * y < 16 ? qw - 16 * y : 16);
                if (p.pt[q]) {  // only the transpose is stored
                    p.mdst[s] = p.pg[q] + 16 * y + (size_t)16 * x * ldg;
                    continue;
                }
                p.dst[s] = p.pg[q] + 16 * x + (size_t)16 * y * ldg;
                if (sym && x < y) p.mdst[s] = p.pg[q] + 16 * y + (size_t)16 * x * ldg;
            }
    }
    p.nwb = wmax;
    return nw;
}

}  // namespace

bool gram_plan_product(GramPlan &p0, const void *L, int lw, const void *Q, int qw, double *g,
                       long ldg, bool transposed) {
    GramPlan p = p0;  // work on a copy: commit only if everything fits
    if (p.nprod >= 16) return false;
    const int lb = plan_block(p, L, lw), qb = plan_block(p, Q, qw);
    if (lb < 0 || qb < 0) return false;
    BX_REQUIRE(!(transposed && lb == qb), "gram_plan_product: a symmetric product stored transposed");
    p.pl[p.nprod] = (short)lb;
    p.pq[p.nprod] = (short)qb;
    p.pg[p.nprod] = g;
    p.pt[p.nprod] = transposed ? 1 : 0;
    p.nprod++;
    if (form_blocks(p, ldg, false, kGramBQ) > kGramWaves) return false;
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

void GramBatch::add_transposed(const void *L, int lw, const void *Q, int qw, double *g) {
    if (lw == 0 || qw == 0) return;
    if (gram_plan_product(p_, L, lw, Q, qw, g, ldg_, true)) return;
    flush();
    BX_REQUIRE(gram_plan_product(p_, L, lw, Q, qw, g, ldg_, true),
               "Gram product of %d x %d columns does not fit one plan", lw, qw);
}

void GramBatch::flush() {
    if (p_.nprod) gram_run(prec_, ld_, p_, ldg_);
    p_ = GramPlan();
}

void gram_run(int prec, long ld, GramPlan &p, long ldg) {
    BX_REQUIRE(ld % 64 == 0, "gram_run: leading dimension %ld not a multiple of 64", ld);
    if (p.nprod == 0) return;
    // groups of three right tiles spread the work over more waves; four when
    // three would need more register blocks than there are waves
    const int bq = form_blocks(p, ldg, false, 3) <= kGramWaves ? 3 : kGramBQ;
    form_blocks(p, ldg, true, bq);
    const bool f32 = prec == HIPMV_FP32;
    const size_t es = f32 ? sizeof(float) : sizeof(double);
    // narrow plans (at most 96 columns, three right tiles) take chunks of
    // twice the standard rows: their chunk loop is bound by the round trip
    const int rc0 = f32 ? GramCfg<float>::RC : GramCfg<double>::RC;
    // (four times for at most 72 columns, one chunk in flight)
    const bool narrow4 = bq == 3 && p.ncols <= 72 && ld % (4 * rc0) == 0;
    const bool narrow = !narrow4 && bq == 3 && p.ncols <= 96 && ld % (2 * rc0) == 0;
    const int rc = narrow4 ? 4 * rc0 : (narrow ? 2 * rc0 : rc0);
    const int S = rc + (int)(16 / es);
    const long nchunks = ld / rc;
    const int vpc = rc / (int)(16 / es);
    const bool small = p.ncols * vpc <= 6 * kGTh;  // 6 pieces per thread suffice
    // the 6-piece kernels keep two stages in LDS
    const size_t lds = (size_t)(!narrow4 && small ? 2 : 1) * (p.ncols + 16) * S * es +
                       (size_t)p.ncols * sizeof(void *);
    long parts = (long)num_cus();  // one 8-wave workgroup per CU (register bound)
    if (parts > nchunks) parts = nchunks;
    if (parts < 1) parts = 1;
    double *part = (double *)scratch(0, (size_t)parts * kGramTiles * 256 * sizeof(double));
    GramKArgs a;
    a.p = p;
    a.ld = ld;
    a.part = part;
    BX_REQUIRE(lds <= 160 * 1024, "gram_run: %zu bytes of LDS", lds);
    BX_REQUIRE(!narrow || small, "gram_run: narrow plan of %d columns", p.ncols);
    BX_REQUIRE(!narrow4 || p.ncols * vpc <= 9 * kGTh, "gram_run: narrow plan of %d columns",
               p.ncols);
    const void *kern;
    if (f32)
        kern = narrow4  ? (const void *)k_gram<float, 9, 3, 256>
               : narrow ? (const void *)k_gram<float, 6, 3, 128>
               : small ? (bq == 3 ? (const void *)k_gram<float, 6, 3, 64> : (const void *)k_gram<float, 6, 4, 64>)
                       : (bq == 3 ? (const void *)k_gram<float, 9, 3, 64> : (const void *)k_gram<float, 9, 4, 64>);
    else
        kern = narrow4  ? (const void *)k_gram<double, 9, 3, 128>
               : narrow ? (const void *)k_gram<double, 6, 3, 64>
               : small ? (bq == 3 ? (const void *)k_gram<double, 6, 3, 32> : (const void *)k_gram<double, 6, 4, 32>)
                       : (bq == 3 ? (const void *)k_gram<double, 9, 3, 32> : (const void *)k_gram<double, 9, 4, 32>);
    static const void *attr_done[12];
    for (int i = 0; i < 12; i++) {
        if (attr_done[i] == kern) break;
        if (!attr_done[i]) {
            BX_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            attr_done[i] = kern;
            break;
        }
    }
    void *args[] = {&a};
    BX_HIP(hipLaunchKernel(kern, dim3((unsigned)parts), dim3(kGTh), args, lds, stream()));
    BX_HIP(hipGetLastError());
    ReduceArgs r;
    int nt = 0;
    for (int s = 0; s < p.nwb * kSlots; s++) {
        if (!p.dst[s] && !p.mdst[s]) continue;
        r.dst[nt] = p.dst[s];
        r.mdst[nt] = p.mdst[s];
        r.rows[nt] = p.rows[s];
        r.cols[nt] = p.cols[s];
        r.slot[nt] = (unsigned char)s;
        nt++;
    }
    r.nparts = (int)parts;
    r.ldg = ldg;
    r.part = part;
    if (nt) hipLaunchKernelGGL(k_gram_sum, dim3(nt), dim3(256), 0, stream(), r);
    BX_HIP(hipGetLastError());
}

}  // namespace bx
