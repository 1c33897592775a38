/*
 * ell.h -- MI355X-native drop-in for the reference's fixed-width ELL matrix
 * API (inc/ell.h:1-42, src/ell.c).
 *
 * Host-side functions keep the reference's struct layout, signatures, return
 * codes and semantics.  The compute entry points (ell_mvp, ell_solve_jacobi,
 * ell_solve_pcg) run on the GPU through HIP kernels; they upload the matrix
 * and vectors for the call and copy the result back.  Callers that keep a
 * matrix resident use the ell_device_* mirror, which stores the ELL as
 * column-major slabs in HBM.
 *
 * Error codes (as src/ell.c): 0 ok, 1 NULL argument or index >= dimension,
 * 2 negative index or allocation failure, 3 row full / not found.
 * GPU-side failures return 6 (MM_ERR_HIP) after printing the HIP error.
 */
#ifndef MADMAX_ELL_H_
#define MADMAX_ELL_H_

#include <stdlib.h>
#include <stdio.h>
#include <math.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NRM "\x1B[0m"
#define RED "\x1B[31m"
#define GRN "\x1B[32m"

/* inc/ell.h:12-13 (the reference writes them with a trailing ';') */
#define MAX_ITS 100000
#define MIN_ERR 1.0e-10

/* inc/ell.h:15-21: row-major [nrow x nnz], cols == -1 marks an empty slot */
typedef struct ell_matrix_ {
  int nrow;
  int ncol;
  int nnz;
  int *cols;
  double *vals;
} ell_matrix;

/* inc/ell.h:23-28 */
typedef struct ell_solver_ {
  int max_its;
  int its;
  double min_err;
  double err;
} ell_solver;

/* src/ell.c:3-15 -- allocates cols = -1, vals = 0 */
int ell_init(ell_matrix *m, int nrow, int ncol, int nnz);
/* src/ell.c:17-44 -- overwrite or insert at the first free slot */
int ell_set_val(ell_matrix *m, int row, int col, double val);
/* src/ell.c:46-73 -- accumulate or insert at the first free slot */
int ell_add_val(ell_matrix *m, int row, int col, double val);
/* src/ell.c:75-84 -- vals indexed [i*nx + j] like the reference */
int ell_add_vals(ell_matrix *m, int *ix, int nx, int *iy, int ny, double *vals);
/* src/ell.c:127-141 -- y = A x on the GPU, bit-exact with the reference order */
int ell_mvp(ell_matrix *m, double *x, double *y);
/* src/ell.c:186-208 */
int ell_get_val(ell_matrix *m, int row, int col, double *val);
/* src/ell.c:143-184 -- stationary Jacobi on the GPU; like the reference it
 * uses MAX_ITS / MIN_ERR (absolute ||Ax-b||_2) and writes solver->its, err */
int ell_solve_jacobi(ell_solver *solver, ell_matrix *m, double *b, double *x);
/* src/ell.c:86-101 */
int ell_set_zero_row(ell_matrix *m, int row, double diag_val);
/* src/ell.c:103-118 -- returns at the first empty slot of any row, like the
 * reference, so only the leading rows are touched */
int ell_set_zero_col(ell_matrix *m, int col, double diag_val);
/* src/ell.c:120-124 */
int ell_set_zero_mat(ell_matrix *m);
/* src/ell.c:210-221 */
int ell_print_full(ell_matrix *m);

/* ---- extensions (no counterpart in inc/ell.h) ---------------------------- */

/* Frees cols/vals (the reference never frees, micro/src/finalize.c:3-39). */
int ell_free(ell_matrix *m);
/* Jacobi-preconditioned CG on the GPU in FP64, the algorithm the reference
 * reaches through PETSc -ksp_type cg -pc_type jacobi
 * (micro/src/homogenize.c:325-326).  x is the initial guess on entry.
 * Stops at ||b - A x||_2 <= min_err * ||b||_2 (recursive residual) or after
 * max_its iterations; max_its <= 0 / min_err <= 0 select MAX_ITS / MIN_ERR.
 * Returns 0 when converged, 5 when max_its was reached. */
int ell_solve_pcg(ell_solver *solver, ell_matrix *m, double *b, double *x);

/* Device-resident mirror: slabs of slab_rows rows (a multiple of 64, 0 picks
 * the default 256), column-major inside a slab.  Every column id must be -1
 * or in [0, ncol) (what ell_add_val stores); create and update return 1
 * otherwise, since the GPU would read x at that index. */
typedef struct ell_device_ ell_device;
int ell_device_create(const ell_matrix *m, int slab_rows, ell_device **d);
/* Re-uploads cols/vals of a matrix with the same shape. */
int ell_device_update(ell_device *d, const ell_matrix *m);
/* Copies the device matrix back in the reference's row-major layout. */
int ell_device_download(const ell_device *d, int *cols, double *vals);
/* y = A x with host vectors (x, y of length nrow / ncol). */
int ell_device_mvp(ell_device *d, const double *x, double *y);
/* Average duration in ms of `reps` back-to-back device SpMV kernels on the
 * resident vectors (HIP events on the mirror's stream; no host transfer).
 * No counterpart in the reference: a measurement hook. */
int ell_device_time_mvp(ell_device *d, int reps, double *ms);
int ell_device_solve_pcg(ell_device *d, ell_solver *solver, const double *b, double *x);
int ell_device_solve_jacobi(ell_device *d, ell_solver *solver, const double *b, double *x);
int ell_device_destroy(ell_device *d);
/* How ell_device_solve_pcg (and the _noalloc variant) runs its iterations on
 * this mirror.  0 (the default): two launches per iteration, q = A p and the
 * vector update, each of whose workgroups first sums the other's slab
 * partials itself.  n > 0: one cooperative persistent kernel that runs up to
 * n whole iterations per launch, its workgroups meeting at two grid barriers
 * per iteration.  In both, the host checks for convergence between launches.
 * With 256-row slabs both give the same bits.  Returns 1 for n < 0, or for
 * n > 0 on a mirror whose slabs are not 256 rows or on a device without
 * cooperative launch.  No counterpart in the reference: it selects how the
 * PCG loop is mapped onto kernels. */
int ell_device_set_pcg_fused(ell_device *d, int its_per_launch);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the Jacobi-PCG of ell_device_solve_pcg (same kernels,
 * same bits) on a mirror created beforehand, with b and x (initial guess in,
 * solution out) in device memory the caller owns.  The PCG work vectors,
 * partial sums and iteration state live in one library-wide device array
 * sized for ELL_NOALLOC_MAX_ROWS rows (about 48.5 MiB per GPU, held until the
 * process ends); the call allocates nothing.  Returns 1 for a matrix with more
 * rows, 0 when converged, 5 at max_its.  Every call on one GPU uses the same
 * array: calls that overlap on one GPU race. */
#define ELL_NOALLOC_MAX_ROWS (1 << 20)
int ell_device_solve_pcg_noalloc(ell_device *d, ell_solver *solver, const double *b, double *x);

#ifdef __cplusplus
}
#endif
#endif
