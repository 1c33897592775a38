/*
 * ell_host.c -- host half of the ELL API (include/ell.h): construction,
 * insertion and Dirichlet edits on the caller-owned row-major ELL, with the
 * reference's semantics, return codes and messages (src/ell.c).  The compute
 * entry points (ell_mvp, ell_solve_jacobi, ell_solve_pcg) live in
 * ell_device.hip and run on the GPU.
 */
#include "ell.h"

#include <stdint.h>
#include <string.h>

/* src/ell.c:3-15 */
int ell_init(ell_matrix *m, int nrow, int ncol, int nnz)
{
  if (m == NULL) return 1;
  m->nrow = nrow;
  m->ncol = ncol;
  m->nnz = nnz;
  if (nrow < 0 || nnz < 0) return 2;
  size_t n = (size_t)nrow * (size_t)nnz;
  m->cols = (int *)malloc((n ? n : 1) * sizeof(int));
  m->vals = (double *)malloc((n ? n : 1) * sizeof(double));
  if (m->cols == NULL || m->vals == NULL) return 2;
  for (size_t i = 0; i < n; i++) m->cols[i] = -1;
  memset(m->vals, 0, n * sizeof(double));
  return 0;
}

int ell_free(ell_matrix *m)
{
  if (m == NULL) return 1;
  free(m->cols);
  free(m->vals);
  m->cols = NULL;
  m->vals = NULL;
  return 0;
}

/* Walks the row: the first slot holding `col` or the first empty slot.
 * Returns the slot or -1 when the row is full without `col`. */
static int find_slot(const ell_matrix *m, int row, int col)
{
  const int *c = m->cols + (size_t)row * (size_t)m->nnz;
  for (int j = 0; j < m->nnz; j++)
    if (c[j] == -1 || c[j] == col) return j;
  return -1;
}

/* src/ell.c:17-44 */
int ell_set_val(ell_matrix *m, int row, int col, double val)
{
  if (row >= m->nrow || col >= m->ncol) {
    printf(RED "ell error: row %d or col %d greater than the dimension of the matrix\n" NRM, row, col);
    return 1;
  }
  if (row < 0 || col < 0) {
    printf(RED "ell error: negative values in row %d or col %d\n" NRM, row, col);
    return 2;
  }
  int j = find_slot(m, row, col);
  if (j < 0) {
    printf(RED "ell error: not enought space to store value in row %d and col %d\n" NRM, row, col);
    return 3;
  }
  size_t k = (size_t)row * (size_t)m->nnz + (size_t)j;
  m->cols[k] = col;
  m->vals[k] = val;
  return 0;
}

/* src/ell.c:46-73: the first touch stores, later touches accumulate */
int ell_add_val(ell_matrix *m, int row, int col, double val)
{
  if (row >= m->nrow || col >= m->ncol) {
    printf(RED "ell error: row %d or col %d greater than the dimension of the matrix\n" NRM, row, col);
    return 1;
  }
  if (row < 0 || col < 0) {
    printf(RED "ell error: negative values in row %d or col %d\n" NRM, row, col);
    return 2;
  }
  int j = find_slot(m, row, col);
  if (j < 0) {
    printf(RED "ell error: not enought space to add value in row %d and col %d\n" NRM, row, col);
    return 3;
  }
  size_t k = (size_t)row * (size_t)m->nnz + (size_t)j;
  if (m->cols[k] == -1) {
    m->cols[k] = col;
    m->vals[k] = val;
  } else {
    m->vals[k] += val;
  }
  return 0;
}

/* src/ell.c:75-84 -- like the reference, vals is read with row stride nx */
int ell_add_vals(ell_matrix *m, int *ix, int nx, int *iy, int ny, double *vals)
{
  if (m == NULL || ix == NULL || iy == NULL || vals == NULL) return 1;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) ell_add_val(m, ix[i], iy[j], vals[i * nx + j]);
  return 0;
}

/* src/ell.c:86-101 */
int ell_set_zero_row(ell_matrix *m, int row, double diag_val)
{
  if (m == NULL) return 1;
  int *c = m->cols + (size_t)row * (size_t)m->nnz;
  double *v = m->vals + (size_t)row * (size_t)m->nnz;
  for (int j = 0; j < m->nnz && c[j] != -1; j++) v[j] = (c[j] == row) ? diag_val : 0.0;
  return 0;
}

/* src/ell.c:103-118 -- the scan of every row ends the whole call at the first
 * empty slot it meets (src/ell.c:110); kept so results match the reference. */
int ell_set_zero_col(ell_matrix *m, int col, double diag_val)
{
  if (m == NULL) return 1;
  for (int i = 0; i < m->nrow; i++) {
    int *c = m->cols + (size_t)i * (size_t)m->nnz;
    double *v = m->vals + (size_t)i * (size_t)m->nnz;
    for (int j = 0; j < m->nnz; j++) {
      if (c[j] == -1) return 0;
      if (c[j] == col) v[j] = (i == col) ? diag_val : 0.0;
    }
  }
  return 0;
}

/* src/ell.c:120-124 */
int ell_set_zero_mat(ell_matrix *m)
{
  for (int i = 0; i < m->nrow; i++) ell_set_zero_row(m, i, 0.0);
  return 0;
}

/* src/ell.c:186-208 */
int ell_get_val(ell_matrix *m, int row, int col, double *val)
{
  if (row >= m->nrow || col >= m->ncol) {
    printf(RED "ell_get_val: row %d or col %d greater than the dimension of the matrix\n" NRM, row, col);
    return 1;
  }
  if (row < 0 || col < 0) {
    printf(RED "ell_get_val: negative values in row %d or col %d\n" NRM, row, col);
    return 2;
  }
  int j = find_slot(m, row, col);
  if (j < 0) return 3;
  size_t k = (size_t)row * (size_t)m->nnz + (size_t)j;
  *val = (m->cols[k] == -1) ? 0.0 : m->vals[k];
  return 0;
}

/* src/ell.c:210-221 */
int ell_print_full(ell_matrix *m)
{
  if (m == NULL) return 1;
  if (m->vals == NULL || m->cols == NULL) return 2;
  double val;
  for (int i = 0; i < m->nrow; i++)
    for (int j = 0; j < m->ncol; j++)
      printf("%lf%s", (ell_get_val(m, i, j, &val) == 0) ? val : 0.0, (j == m->ncol - 1) ? "\n" : " ");
  return 0;
}
