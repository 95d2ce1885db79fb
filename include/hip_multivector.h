/*
 * Device multivector backend for the BLOPEX C API on MI355X (gfx950).
 *
 * The reference's serial backend (blopex_serial_double/multivector/
 * multi_vector.h:22-32, pcg_multi/pcg_multi.h:17-51) stores a block as one
 * column-major n x k host array and loops over its columns one at a time.
 * hip_Multi_Vector keeps the same layout and mask semantics in device memory
 * (HBM) and every block operation is a HIP kernel over the whole block.
 *
 *   hip_Multi_Vector      <- serial_Multi_Vector       (multi_vector.h:22-32)
 *   HipSetupInterpreter   <- SerialSetupInterpreter    (pcg_multi.c:203-242)
 *   HipCSRMatMultiVec     <- MatMultiVec               (pcg_multi.c:167-193)
 *   HipDiagMatMultiVec    <- operatorB / operatorT slots (krylov/lobpcg.c:505-509)
 *
 * lobpcg_solve_double (lobpcg.h) on a block wrapped with this interpreter
 * runs the device-resident solver; its precision and statistics hang off
 * the block (hip_Multi_VectorSetSettings / hip_Multi_VectorGetStats).
 *
 * Operator callbacks receive hip_Multi_Vector pointers.  They act on the
 * active columns of X and Y, enqueue their work on blopex_amd_stream() and
 * may return before it completes: everything the solver does afterwards is
 * ordered behind it on that stream.
 *
 * All work of the library runs on one HIP stream per process (default
 * priority), created on the device current at the first call.
 */
#ifndef BLOPEX_AMD_HIP_MULTIVECTOR_H
#define BLOPEX_AMD_HIP_MULTIVECTOR_H

#include "interpreter.h"
#include "multivector.h"

#define HIPMV_FP64 0
#define HIPMV_FP32 1
#define HIPMV_C64 2 /* komplex entries (interleaved FP64 real, imaginary): blocks of
                       lobpcg_solve_complex, which runs the generic algorithm */

#define HIPLOBPCG_FP64 0  /* every block operation in FP64 */
#define HIPLOBPCG_MIXED 1 /* FP32 blocks, FP64 Gram accumulation and small dense, FP64 finish */

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hip_LobpcgSettings {
    int mode;                 /* HIPLOBPCG_FP64 (default) or HIPLOBPCG_MIXED */
    double switch_factor;     /* MIXED: leave FP32 once every active residual is below
                                 switch_factor * atol (and, with eig_rtol, every eigenvalue
                                 bound below switch_factor^2 * eig_rtol |lambda|), or FP32 has
                                 stalled (default 8: on the 128^3 k = 32 workload the means over
                                 six FP32 trajectories were 1.83, 1.64, 1.65, 1.67 and 1.66 s
                                 at 1.5, 2.5, 4, 6 and 8, and only 8 kept every trajectory
                                 off the stall exit) */
    int stall_iterations;     /* MIXED: window W; FP32 has stalled when the smallest geometric
                                 mean of the k residual norms over the last W iterations is
                                 not 2% below that of the W before (default 50) */
    int refine_iterations;    /* FP64 iterations with all columns active after convergence
                                 (default 0) */
    int profile;              /* 1: time kernel classes with HIP events (adds event records
                                 to the stream, and the next iteration's A R is no longer
                                 enqueued ahead of the host's read of the residual norms;
                                 the results are the same bits; default 0) */
    int refresh_interval;     /* MIXED: FP32 iterations between recomputations of AX (and BX)
                                 from X, which keep the FP32 residuals those of X; 0: never
                                 (default 20) */
    int final_rr;             /* 1: after the last iteration, recompute AX (and BX) from X and
                                 do a Rayleigh-Ritz step on X with explicit X'AX, X'BX, so
                                 that X is B-orthonormal and the residuals are those of the
                                 returned X (lobpcg.m:922-974; default 0) */
    int guard_columns;        /* g >= 0: the solver iterates k + g columns (k + g <= 64) and
                                 tests and returns the first k, the k smallest Ritz pairs.
                                 The guard columns widen the gap that sets the convergence
                                 of the k-th pair; they start from 2*drand48()-1 after
                                 srand48(guard_seed), column by column (default 0) */
    double eig_rtol;          /* > 0: a wanted column also stays active until the error
                                 bound of its eigenvalue, min over windows [a, b] of Ritz
                                 pairs around it of sum_j ||r_j||^2 / gap(a, b), is at most
                                 eig_rtol * |lambda_i| (default 0: residuals only) */
    int cond_restart;         /* 1: when the condition estimate of the Rayleigh-Ritz pencil's
                                 B part jumps (log10 of its Cholesky pivot spread + 1 above
                                 min(8, max(2, twice its recent mean))), or the pencil is not
                                 positive definite, that step is solved again without P, a
                                 steepest-descent restart (lobpcg.m:846-863); without it FP64
                                 stops with -1 on such a pencil, as lobpcg.c does (default 0) */
    int size_guard;           /* 1: return PROBLEM_SIZE_TOO_SMALL when n - sizeY < 5 (k + g)
                                 (lobpcg.m:475-484; lobpcg.c:672-687 has it disabled;
                                 default 0) */
    int guard_seed;           /* srand48 seed of the guard columns (default 7) */
} hip_LobpcgSettings;

typedef struct hip_LobpcgStats {
    double seconds;             /* wall time of the last device solve */
    int iterations_fp32;        /* iterations done in FP32 (MIXED) */
    int iterations_fp64;        /* iterations done in FP64 */
    int refine_iterations;      /* of which refinement */
    int exit_flag;
    /* device time per kernel class, filled when settings.profile = 1 */
    double seconds_small_dense; /* FP64 Cholesky / inverse / Rayleigh-Ritz */
    double seconds_spmm;        /* operator A applications */
    double seconds_gram;        /* Gram products */
    double seconds_update;      /* block updates (block x small matrix, axpy) */
    double seconds_other;       /* residuals, norms, preconditioner, B, gathers, copies */
    double spmm_bytes;          /* algorithmic bytes of the A products timed above (value,
                                   index and block streams; the diagonal copy has no index
                                   stream) */
    int spmm_calls;
    int fp32_exit;              /* MIXED: why the FP32 phase ended: 0 converged or within
                                   switch_factor * atol, 1 stall, 2 breakdown (R or pencil
                                   not positive definite in FP32) */
    int final_rr;               /* 1 when the final Rayleigh-Ritz step ran */
    int guard_columns;          /* guard columns iterated */
    int cond_restarts;          /* steps solved again without P on a condition-estimate jump */
    double eig_error_bound;     /* at exit: max over the wanted columns of the eigenvalue
                                   error bound (see eig_rtol) divided by |lambda_i|; 0 when
                                   eig_rtol is 0 */
} hip_LobpcgStats;

typedef struct hip_Multi_Vector {
    void *data;                    /* device memory, column-major */
    long size;                     /* rows */
    long ld;                       /* leading dimension (>= size); callbacks touch rows < size */
    BlopexInt num_vectors;         /* columns */
    BlopexInt num_active_vectors;
    BlopexInt *active_indices;     /* host, 0-based, ascending */
    BlopexInt owns_data;
    int precision;                 /* HIPMV_FP64 or HIPMV_FP32 */
    hip_LobpcgSettings settings;
    hip_LobpcgStats stats;
} hip_Multi_Vector;

/* ---- library ---------------------------------------------------------- */
int blopex_amd_init(void);              /* optional: 0 on success */
void *blopex_amd_stream(void);          /* the hipStream_t all work runs on */
int blopex_amd_synchronize(void);       /* waits for the stream; 0 on success */
const char *blopex_amd_last_error(void);/* last error message ("" if none) */
/* frees the device workspace the library keeps between calls (the solver's
   block pools, Gram partials, Rayleigh-Ritz buffers); 0 on success */
int blopex_amd_release_workspace(void);

/* ---- blocks ------------------------------------------------------------ */
hip_Multi_Vector *hip_Multi_VectorCreate(long size, BlopexInt num_vectors, int precision);
/* wraps existing device memory (not freed by Destroy) */
hip_Multi_Vector *hip_Multi_VectorWrapDevice(void *data, long size, BlopexInt num_vectors,
                                             int precision);
void hip_Multi_VectorDestroy(hip_Multi_Vector *x);
/* host <-> device, all columns, host array column-major n x k FP64 */
int hip_Multi_VectorUpload(hip_Multi_Vector *x, const double *host);
int hip_Multi_VectorDownload(const hip_Multi_Vector *x, double *host);
/* 2*drand48()-1 over the active columns, column by column, after srand48(seed):
   the stream of serial_Multi_VectorSetRandomValues (multi_vector.c:166-184),
   generated on the device by jumping the LCG ahead per thread */
int hip_Multi_VectorSetRandomValues(hip_Multi_Vector *x, BlopexInt seed);
int hip_Multi_VectorSetSettings(hip_Multi_Vector *x, const hip_LobpcgSettings *s);
/* the defaults every new block carries (FP64 mode, switch_factor 8, stall window 50,
   refresh interval 20, guard seed 7, everything else 0) */
void hip_LobpcgSettingsDefault(hip_LobpcgSettings *s);
int hip_Multi_VectorGetStats(const hip_Multi_Vector *x, hip_LobpcgStats *s);

/* v = X(active)' Y(active), h x w column-major with leading dimension gh in
   host memory: the interpreter's MultiInnerProd slot (mv_MultiVectorByMultiVector
   on HIP blocks) without any allocation.  No counterpart in the reference: it
   is for callers that cannot allocate (tight serving loops, harnesses that
   forbid allocation).  Its workspace is a fixed device array of 2.2 MB held
   by each device for the life of the process; h and w are at most 64 (larger
   results return 1 with a message).  Calls that overlap on one device share
   the array and race.  Real blocks (FP64 or FP32) only; returns 0 on success. */
int hip_Multi_VectorInnerProdNoAlloc(const hip_Multi_Vector *x, const hip_Multi_Vector *y,
                                     BlopexInt gh, BlopexInt h, BlopexInt w, double *v);

/* fills every multivector slot of the table the solver uses */
BlopexInt HipSetupInterpreter(mv_InterfaceInterpreter *ii);

/* ---- operators ----------------------------------------------------------- */
#define HIP_CSR_MAX_DIAGONALS 16
typedef struct hip_CSRMatrix {
    long n;
    long nnz;
    int *rowptr; /* device */
    int *colind; /* device */
    double *val; /* device FP64 */
    float *val32;/* device FP32 copy (for FP32 blocks) */
    /* diagonal copy, built when the nonzeros lie on at most
       HIP_CSR_MAX_DIAGONALS diagonals filled to at least one half and the
       rows have sorted, distinct columns (products then sum in the same
       order as the CSR kernel); ndiag = 0 otherwise */
    int ndiag;
    int diag_offset[HIP_CSR_MAX_DIAGONALS]; /* column - row, ascending */
    long diag_ld;                           /* >= n, multiple of 4 */
    double *diag_val;                       /* device, ndiag x diag_ld, 0 off the matrix */
    float *diag_val32;
    /* nx, ny, nz when the diagonals are a 7-point stencil on that grid (x
       fastest; offsets -nx*ny, -nx, -1, 0, 1, nx, nx*ny, zero coefficients
       for neighbours outside the grid); 0 otherwise */
    int grid[3];
} hip_CSRMatrix;

typedef struct hip_DiagMatrix {
    long n;
    double *d;   /* device FP64 */
    float *d32;  /* device FP32 copy */
} hip_DiagMatrix;

/* square CSR from host arrays (int32 row pointers / columns, FP64 values) */
hip_CSRMatrix *hip_CSRMatrixCreate(long n, long nnz, const int *rowptr, const int *colind,
                                   const double *val);
void hip_CSRMatrixDestroy(hip_CSRMatrix *A);
/* operatorA / operatorB callback: Y(active) = A * X(active) */
void HipCSRMatMultiVec(void *A, void *X, void *Y);
/* the kernel HipCSRMatMultiVec runs for ncols contiguous active columns of
   blocks with leading dimension ld: 0 CSR, 1 diagonal copy, 2 z-marching
   7-point stencil kernel (no counterpart in the reference) */
int hip_CSRMatrixProductKernel(const hip_CSRMatrix *A, int precision, BlopexInt ncols, long ld);

hip_DiagMatrix *hip_DiagMatrixCreate(long n, const double *d);
/* Jacobi preconditioner T = diag(A)^{-1} */
hip_DiagMatrix *hip_DiagMatrixCreateJacobi(const hip_CSRMatrix *A);
void hip_DiagMatrixDestroy(hip_DiagMatrix *D);
/* operatorB / operatorT callback: Y(active) = diag(d) * X(active) */
void HipDiagMatMultiVec(void *D, void *X, void *Y);

/* operatorA callback MatMultiVec (blopex_serial_double/pcg_multi/
   pcg_multi.c:167-193, blopex_serial_complex/pcg_multi/pcg_multi.c,
   multi_vector.c:567-610): A is a hip_Multi_Vector of n x n entries (all
   columns), Y(active) = A * X(active).  A complex (HIPMV_C64) A takes complex
   blocks; a real (HIPMV_FP64) A takes FP64 or FP32 blocks (FP64 arithmetic).
   A == NULL applies the 1D Laplacian [-1 2 -1] with Dirichlet ends to real
   blocks, as serial_Lapl1DMatMultiVec does (matmultivec.c:11-48). */
void HipDenseMatMultiVec(void *A, void *X, void *Y);

#ifdef __cplusplus
}
#endif

#endif
