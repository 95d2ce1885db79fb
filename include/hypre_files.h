/* Reading the HYPRE file formats the reference's workflows use for A, B and X0.
 *
 * IJ matrix files: prefix.00000, prefix.00001, ... (one per process), each
 * "ilower iupper jlower jupper" (0-based row and column range of the part)
 * followed by "i j value" lines with 0-based global indices, as written by
 * blopex_tools/matlab/matlab2hypre/matlab2hypreIJ.m:73-99 and read back by
 * hypreIJ2matlab.m:44-65; the hypre driver loads the same files with
 * `ij -fromfile prefix` (blopex_hypre/.../ij.c:2582-2588).
 *
 * ParVector files: prefix.<vector>.<part>, each the part's row count then one
 * value per line (matlab2hypreParVectors.m:55-75, hypreParVectors2matlab.m).
 *
 * The C reference has no reader of its own (HYPRE does the reading there);
 * these give its drivers' inputs to the device solver. */
#ifndef BLOPEX_AMD_HYPRE_FILES_H
#define BLOPEX_AMD_HYPRE_FILES_H

#include "hip_multivector.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host CSR of an IJ file set: n, nnz and malloc'ed rowptr (n + 1), colind,
   val (nnz); entries of a row sorted by column, repeated (i, j) entries summed
   (hypreIJ2matlab sums them too).  Returns 0, or -1 with blopex_amd_last_error
   set (missing files, a malformed line, an index outside the part's rows or
   the matrix).  Free the arrays with hip_IJFree. */
int hip_IJReadCSR(const char *prefix, long *n, long *nnz, int **rowptr, int **colind,
                  double **val);
/* Host n x m block (column-major, malloc'ed) of a ParVector file set; m is the
   number of vectors found.  Returns 0 or -1 as above.  Free with hip_IJFree. */
int hip_ParVectorsRead(const char *prefix, long *n, int *m, double **v);
void hip_IJFree(void *p);

/* Device objects straight from the files (NULL on error) */
hip_CSRMatrix *hip_CSRMatrixReadIJ(const char *prefix);
hip_Multi_Vector *hip_Multi_VectorReadParVectors(const char *prefix); /* FP64 */

#ifdef __cplusplus
}
#endif

#endif
