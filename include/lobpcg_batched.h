/* Many independent eigenproblems in one solve: every iteration advances all
 * unconverged problems together, one grouped launch per step (residual and
 * Jacobi preconditioner, SpMM, Gram blocks, Rayleigh-Ritz, block update), and
 * the launches of an iteration are replayed from one captured hipGraph.
 *
 * No counterpart in the reference, which solves one problem per call
 * (lobpcg_solve, krylov/lobpcg.c:502-1388).  Per problem the iteration is
 * LOBPCG's: Rayleigh-Ritz on [X, T R, P] (lobpcg.c:1085-1190) with the
 * initial Rayleigh-Ritz in span(X) (lobpcg.c:842-905).  Columns are not
 * locked one by one: a problem keeps all k columns in its basis until every
 * residual norm ||A x_i - lambda_i x_i|| is below the tolerance, then it
 * leaves the batch (the convergence test of lobpcg_checkResiduals,
 * lobpcg.c:265-294, applied to the whole block). */
#ifndef BLOPEX_AMD_LOBPCG_BATCHED_H
#define BLOPEX_AMD_LOBPCG_BATCHED_H

#include "hip_multivector.h" /* HIPLOBPCG_FP64, HIPLOBPCG_MIXED */

#ifdef __cplusplus
extern "C" {
#endif

#define HIP_BATCH_MAX_K 16

typedef struct hip_BatchProblem {
    long n;             /* rows */
    long nnz;
    const int *rowptr;  /* host CSR, symmetric positive definite A */
    const int *colind;
    const double *val;
    int k;              /* eigenpairs wanted, 1..HIP_BATCH_MAX_K */
} hip_BatchProblem;

typedef struct hip_BatchSettings {
    double tol;         /* absolute residual tolerance (default 1e-6) */
    int maxit;          /* iterations (default 500) */
    int jacobi;         /* 1: T = diag(A)^-1 (default 1) */
    int seed;           /* X0: drand48 after srand48(seed), problem by problem, column by
                           column (default 1) */
    int graph;          /* 1: replay each iteration's launches from a hipGraph, the
                           convergence test on the device and read by the host every 4
                           iterations (default 1); 0: launches and test from the host every
                           iteration (the same results) */
    int guard;          /* extra columns iterated beside the k wanted, up to
                           HIP_BATCH_MAX_K + 8 columns in all and n/3; only the k wanted
                           are tested for convergence and returned (default 2) */
    int mode;           /* HIPLOBPCG_FP64 (default), or HIPLOBPCG_MIXED: the blocks in FP32
                           (FP32 block updates, Gram chunks on FP32 MFMA summed in FP64,
                           FP64 Rayleigh-Ritz, AX recomputed from X every 20 iterations)
                           until all but 5% of the problems have their wanted residuals
                           below switch_factor * tol (or maxit / 2 iterations have run),
                           then every problem continues in FP64 from its block */
    double switch_factor; /* MIXED: see mode (default 100: on the 2D 512-problem workload
                             30, 100, 300, 1000 and 3000 took 3.15, 3.07, 3.08, 3.21 and
                             3.36 s) */
    int autotune;       /* 1: pick the launch shapes by timing the SpMM and Gram kernels on
                           this batch before the solve: the rows per chunk (the work of one
                           workgroup: 512, 1024, 2048 or 4096), then for MIXED the FP32
                           Gram's staged rows per step (16, 32 or 64) (default 1);
                           0: 2048 rows per chunk, 32 rows per step */
} hip_BatchSettings;

typedef struct hip_BatchStats {
    double seconds;     /* wall time of the solve (device-resident inputs excluded) */
    int iterations;     /* batch iterations (the slowest problem's count) */
    int graph_captures; /* graphs captured (one per phase: the active lists and counts
                           live in device memory) */
    int graph_replays;  /* iterations replayed from the captured graphs */
    long total_rows;    /* sum of n */
    int iterations_fp32; /* MIXED: batch iterations of the FP32 phase */
    int chunk_rows;     /* rows per chunk used */
    int gram_rows;      /* rows per staged step of the FP32 Gram pass used */
} hip_BatchStats;

void hip_BatchSettingsDefault(hip_BatchSettings *s);

/* Solves all nprob problems.  Outputs (host): lambda and resnorm, sum of k
   doubles (problem by problem, ascending lambda); iterations and flags,
   nprob ints (flag 0 converged, -1 not within maxit, 3 Rayleigh-Ritz
   breakdown); X (NULL: not returned) sum of n*k doubles, each problem's n x k
   block column-major.  Returns 0 when every problem converged, -1 otherwise
   (blopex_amd_last_error for argument errors). */
int hip_LobpcgSolveBatched(int nprob, const hip_BatchProblem *probs,
                           const hip_BatchSettings *s, double *lambda, double *resnorm,
                           int *iterations, int *flags, double *X, hip_BatchStats *st);

#ifdef __cplusplus
}
#endif

#endif
