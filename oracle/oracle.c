/*
 * oracle.c -- FP64 CPU restatement of the sputnik micro (RVE) homogenization
 * path, written from the reference's behaviour, not copied from it.
 *
 * TEST INFRASTRUCTURE ONLY.  Loaded by tests/, by __graft_entry__.smoke() and
 * by the CPU-baseline leg of bench.py, where its results are what the GPU
 * results are compared with; the product library never uses it.
 *
 * Reference behaviours reproduced on purpose (they fix the data layout):
 *   - ELL first-touch column order and tail-only -1 padding (src/ell.c:46-84);
 *   - ell_set_zero_col returns at the first padded slot of ANY row
 *     (src/ell.c:110), so Dirichlet columns stay in the interior rows;
 *   - ell_add_vals indexes vals[i*nx + j] (src/ell.c:80).
 * Reference bugs deliberately NOT reproduced:
 *   - get_elem_disp on the ELL path reads x_ell[i] instead of
 *     x_ell[elem_index[i]] (micro/src/homogenize.c:409): the oracle gathers
 *     through the connectivity like the PETSc branch (:407);
 *   - x_ell / dx_ell are uninitialised malloc (micro/src/alloc.c:51-52): the
 *     oracle starts every RVE solve from zero interior displacement and a zero
 *     correction;
 *   - ell_solve_jacobi ignores solver->max_its / min_err (src/ell.c:157-159):
 *     or_ell_solve_jacobi takes them as arguments (callers pass 100000 and
 *     1e-10 to get the reference's constants).
 *
 * Parity pins (each function is checked by tests/ against something other
 * than a retyped formula):
 *   ELL ops              -> the reference's own src/ell.c compiled into
 *                           oracle/_ref (bitwise cols/vals/trajectories) and
 *                           the committed fixtures in tests/golden/;
 *   hex8 tables          -> src/fem.c compiled into oracle/_ref (bitwise);
 *   phase map            -> micro/src/micro_struct.c compiled into oracle/_ref;
 *   Ke, C, B             -> symmetry, rigid-body null space, patch test,
 *                           closed-form lambda/mu;
 *   PCG                  -> dense LAPACK solves (numpy) on small systems;
 *   homogenization (2D)  -> published tangent exs/README.md:154-157;
 *   homogenization (3D)  -> homogeneous RVE gives C exactly, major symmetry,
 *                           <eps> = eps_mac, Hill-Mandel, Voigt/Reuss bounds.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* =================================================================== ELL */

int or_ell_init(or_ell *m, int nrow, int ncol, int nnz)
{
  if (m == NULL) return 1;
  m->nrow = nrow;
  m->ncol = ncol;
  m->nnz = nnz;
  size_t n = (size_t)nrow * (size_t)nnz;
  m->cols = (int *)malloc((n ? n : 1) * sizeof(int));
  m->vals = (double *)malloc((n ? n : 1) * sizeof(double));
  if (m->cols == NULL || m->vals == NULL) return 2;
  for (size_t i = 0; i < n; i++) m->cols[i] = -1;
  for (size_t i = 0; i < n; i++) m->vals[i] = 0.0;
  return 0;
}

void or_ell_free(or_ell *m)
{
  if (m == NULL) return;
  free(m->cols);
  free(m->vals);
  m->cols = NULL;
  m->vals = NULL;
}

static int ell_check_index(const or_ell *m, int row, int col)
{
  if (row >= m->nrow || col >= m->ncol) return 1;
  if (row < 0 || col < 0) return 2;
  return 0;
}

/* Insert or overwrite / accumulate: first matching or first free slot. */
static int ell_put(or_ell *m, int row, int col, double val, int accumulate)
{
  int ec = ell_check_index(m, row, col);
  if (ec) return ec;
  size_t base = (size_t)row * (size_t)m->nnz;
  for (int j = 0; j < m->nnz; j++) {
    int c = m->cols[base + j];
    if (c == -1) {
      m->cols[base + j] = col;
      m->vals[base + j] = val;
      return 0;
    }
    if (c == col) {
      if (accumulate) m->vals[base + j] += val;
      else m->vals[base + j] = val;
      return 0;
    }
  }
  return 3;
}

int or_ell_set_val(or_ell *m, int row, int col, double val) { return ell_put(m, row, col, val, 0); }
int or_ell_add_val(or_ell *m, int row, int col, double val) { return ell_put(m, row, col, val, 1); }

int or_ell_add_vals(or_ell *m, const int *ix, int nx, const int *iy, int ny, const double *vals)
{
  if (m == NULL || ix == NULL || iy == NULL || vals == NULL) return 1;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      or_ell_add_val(m, ix[i], iy[j], vals[i * nx + j]);
  return 0;
}

int or_ell_set_zero_row(or_ell *m, int row, double diag_val)
{
  if (m == NULL) return 1;
  size_t base = (size_t)row * (size_t)m->nnz;
  for (int j = 0; j < m->nnz; j++) {
    int c = m->cols[base + j];
    if (c == -1) return 0;
    m->vals[base + j] = (c == row) ? diag_val : 0.0;
  }
  return 0;
}

int or_ell_set_zero_col(or_ell *m, int col, double diag_val)
{
  if (m == NULL) return 1;
  for (int i = 0; i < m->nrow; i++) {
    size_t base = (size_t)i * (size_t)m->nnz;
    for (int j = 0; j < m->nnz; j++) {
      int c = m->cols[base + j];
      if (c == -1) return 0; /* the reference leaves the whole scan here */
      if (c == col) m->vals[base + j] = (i == col) ? diag_val : 0.0;
    }
  }
  return 0;
}

int or_ell_set_zero_mat(or_ell *m)
{
  for (int i = 0; i < m->nrow; i++) or_ell_set_zero_row(m, i, 0.0);
  return 0;
}

int or_ell_mvp(const or_ell *m, const double *x, double *y)
{
  if (m == NULL || x == NULL || y == NULL) return 1;
  for (int i = 0; i < m->nrow; i++) {
    size_t base = (size_t)i * (size_t)m->nnz;
    double acc = 0.0;
    for (int j = 0; j < m->nnz; j++) {
      int c = m->cols[base + j];
      if (c == -1) break;
      acc += m->vals[base + j] * x[c];
    }
    y[i] = acc;
  }
  return 0;
}

int or_ell_get_val(const or_ell *m, int row, int col, double *val)
{
  int ec = ell_check_index(m, row, col);
  if (ec) return ec;
  size_t base = (size_t)row * (size_t)m->nnz;
  for (int j = 0; j < m->nnz; j++) {
    int c = m->cols[base + j];
    if (c == -1) { *val = 0.0; return 0; }
    if (c == col) { *val = m->vals[base + j]; return 0; }
  }
  return 3;
}

int or_ell_solve_jacobi(const or_ell *m, const double *b, double *x, int max_its,
                        double min_err, int *its_out, double *err_out)
{
  if (m == NULL || b == NULL || x == NULL) return 1;
  int n = m->nrow;
  double *r = (double *)malloc((size_t)n * sizeof(double));
  double *dinv = (double *)malloc((size_t)n * sizeof(double));
  for (int i = 0; i < n; i++) {
    or_ell_get_val(m, i, i, &dinv[i]);
    dinv[i] = 1 / dinv[i];
  }
  int its = 0;
  double err = 0.0;
  while (its < max_its) {
    err = 0;
    or_ell_mvp(m, x, r);
    for (int i = 0; i < n; i++) {
      r[i] -= b[i];
      err += r[i] * r[i];
    }
    err = sqrt(err);
    if (err < min_err) break;
    for (int i = 0; i < n; i++) x[i] -= dinv[i] * r[i];
    its++;
  }
  if (its_out) *its_out = its;
  if (err_out) *err_out = err;
  free(r);
  free(dinv);
  return 0;
}

static double dot(int n, const double *a, const double *b)
{
  double s = 0.0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

int or_pcg(const or_ell *m, const double *b, double *x, double rtol, int max_its, int *its_out,
           double *relres_out)
{
  if (m == NULL || b == NULL || x == NULL) return 1;
  int n = m->nrow;
  double *r = (double *)malloc((size_t)n * sizeof(double));
  double *z = (double *)malloc((size_t)n * sizeof(double));
  double *p = (double *)malloc((size_t)n * sizeof(double));
  double *q = (double *)malloc((size_t)n * sizeof(double));
  double *dinv = (double *)malloc((size_t)n * sizeof(double));
  for (int i = 0; i < n; i++) {
    double d;
    or_ell_get_val(m, i, i, &d);
    dinv[i] = 1.0 / d;
  }
  double bnorm = sqrt(dot(n, b, b));
  or_ell_mvp(m, x, q);
  for (int i = 0; i < n; i++) r[i] = b[i] - q[i];
  double rnorm = sqrt(dot(n, r, r));
  int its = 0;
  if (bnorm == 0.0) {
    for (int i = 0; i < n; i++) x[i] = 0.0;
    rnorm = 0.0;
    bnorm = 1.0;
  } else if (rnorm > rtol * bnorm) {
    for (int i = 0; i < n; i++) z[i] = dinv[i] * r[i];
    for (int i = 0; i < n; i++) p[i] = z[i];
    double rz = dot(n, r, z);
    while (its < max_its) {
      or_ell_mvp(m, p, q);
      double alpha = rz / dot(n, p, q);
      for (int i = 0; i < n; i++) x[i] += alpha * p[i];
      for (int i = 0; i < n; i++) r[i] -= alpha * q[i];
      its++;
      rnorm = sqrt(dot(n, r, r));
      if (rnorm <= rtol * bnorm) break;
      for (int i = 0; i < n; i++) z[i] = dinv[i] * r[i];
      double rz_new = dot(n, r, z);
      double beta = rz_new / rz;
      rz = rz_new;
      for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
  }
  if (its_out) *its_out = its;
  if (relres_out) *relres_out = rnorm / bnorm;
  free(r); free(z); free(p); free(q); free(dinv);
  return (rnorm <= rtol * bnorm) ? 0 : 5;
}

/* ============================================================ FEM tables */

/* Reference-element node / gauss point sign pattern shared by src/fem.c:324-361
 * (gauss points) and :364-405 (nodes): 0(-,-,-) 1(+,-,-) 2(+,+,-) 3(-,+,-)
 * 4(-,-,+) 5(+,-,+) 6(+,+,+) 7(-,+,+). */
static const double HEX_SIGN[8][3] = {
  {-1, -1, -1}, {+1, -1, -1}, {+1, +1, -1}, {-1, +1, -1},
  {-1, -1, +1}, {+1, -1, +1}, {+1, +1, +1}, {-1, +1, +1}};
static const double GAUSS_A = 0.577350269189626; /* src/fem.c:324 */

void or_hex8_tables(double xp[8][3], double wp[8], double sh[8][8], double ds[8][3][8])
{
  for (int g = 0; g < 8; g++) {
    for (int d = 0; d < 3; d++) xp[g][d] = HEX_SIGN[g][d] * GAUSS_A;
    wp[g] = 1.0;
  }
  for (int g = 0; g < 8; g++) {
    for (int a = 0; a < 8; a++) {
      /* (1 +- xi) written as 1.0 + s*xi: bitwise equal to 1.0 -+ xi */
      double fx = 1.0 + HEX_SIGN[a][0] * xp[g][0];
      double fy = 1.0 + HEX_SIGN[a][1] * xp[g][1];
      double fz = 1.0 + HEX_SIGN[a][2] * xp[g][2];
      sh[a][g] = fx * fy * fz * 0.125;
      ds[a][0][g] = HEX_SIGN[a][0] * fy * fz * 0.125;
      ds[a][1][g] = HEX_SIGN[a][1] * fx * fz * 0.125;
      ds[a][2][g] = HEX_SIGN[a][2] * fx * fy * 0.125;
    }
  }
}

int or_struct_dsh(int dim, const double *h, double *dsh, double *wp)
{
  if (dim == 2) {
    /* src/fem.c:9-14 gauss points, :41-54 physical derivatives */
    const double sg[4][2] = {{-1, -1}, {+1, -1}, {+1, +1}, {-1, +1}};
    double xp[4][2];
    for (int g = 0; g < 4; g++)
      for (int d = 0; d < 2; d++) xp[g][d] = sg[g][d] * GAUSS_A;
    for (int g = 0; g < 4; g++) {
      for (int a = 0; a < 4; a++) {
        double fx = 1 + sg[a][0] * xp[g][0];
        double fy = 1 + sg[a][1] * xp[g][1];
        dsh[(a * 2 + 0) * 4 + g] = sg[a][0] * fy / 4 * 2 / h[0];
        dsh[(a * 2 + 1) * 4 + g] = sg[a][1] * fx / 4 * 2 / h[1];
      }
    }
    double vol = h[0] * h[1];
    for (int g = 0; g < 4; g++) wp[g] = vol / 4;
    return 0;
  }
  if (dim == 3) {
    double xp[8][3], w[8], sh[8][8], ds[8][3][8];
    or_hex8_tables(xp, w, sh, ds);
    for (int g = 0; g < 8; g++)
      for (int a = 0; a < 8; a++)
        for (int d = 0; d < 3; d++) dsh[(a * 3 + d) * 8 + g] = ds[a][d][g] * 2 / h[d];
    double vol = h[0] * h[1] * h[2];
    for (int g = 0; g < 8; g++) wp[g] = vol / 8;
    return 0;
  }
  return 1;
}

int or_struct_bmat(int dim, const double *h, double *bmat, double *wp)
{
  int npe = (dim == 2) ? 4 : 8, ngp = npe, nvoi = (dim == 2) ? 3 : 6, nd = npe * dim;
  double dsh[8 * 3 * 8];
  if (or_struct_dsh(dim, h, dsh, wp)) return 1;
  memset(bmat, 0, sizeof(double) * (size_t)(nvoi * nd * ngp));
#define B(v, i, g) bmat[((v) * nd + (i)) * ngp + (g)]
#define D(a, d, g) dsh[((a) * dim + (d)) * ngp + (g)]
  for (int g = 0; g < ngp; g++) {
    for (int a = 0; a < npe; a++) {
      if (dim == 2) { /* micro/src/init.c:43-48 */
        B(0, a * 2 + 0, g) = D(a, 0, g);
        B(1, a * 2 + 1, g) = D(a, 1, g);
        B(2, a * 2 + 0, g) = D(a, 1, g);
        B(2, a * 2 + 1, g) = D(a, 0, g);
      } else { /* Voigt xx, yy, zz, xy, yz, xz; engineering shears */
        B(0, a * 3 + 0, g) = D(a, 0, g);
        B(1, a * 3 + 1, g) = D(a, 1, g);
        B(2, a * 3 + 2, g) = D(a, 2, g);
        B(3, a * 3 + 0, g) = D(a, 1, g);
        B(3, a * 3 + 1, g) = D(a, 0, g);
        B(4, a * 3 + 1, g) = D(a, 2, g);
        B(4, a * 3 + 2, g) = D(a, 1, g);
        B(5, a * 3 + 0, g) = D(a, 2, g);
        B(5, a * 3 + 2, g) = D(a, 0, g);
      }
    }
  }
#undef B
#undef D
  return 0;
}

int or_material_c(int dim, double young, double poisson, double *c)
{
  int nvoi = (dim == 2) ? 3 : 6;
  int nnor = (dim == 2) ? 2 : 3;
  for (int i = 0; i < nvoi * nvoi; i++) c[i] = 0.0;
  for (int i = 0; i < nnor; i++)
    for (int j = 0; j < nnor; j++) c[i * nvoi + j] = (i == j) ? 1.0 - poisson : poisson;
  for (int i = nnor; i < nvoi; i++) c[i * nvoi + i] = (1 - 2 * poisson) / 2;
  for (int i = 0; i < nvoi * nvoi; i++) c[i] *= young / ((1 + poisson) * (1 - 2 * poisson));
  return 0;
}

int or_ke(int dim, const double *bmat, const double *wp, const double *c, double *ke)
{
  int npe = (dim == 2) ? 4 : 8, ngp = npe, nvoi = (dim == 2) ? 3 : 6, nd = npe * dim;
  for (int i = 0; i < nd * nd; i++) ke[i] = 0.0;
  for (int g = 0; g < ngp; g++)
    for (int i = 0; i < nd; i++)
      for (int j = 0; j < nd; j++)
        for (int k = 0; k < nvoi; k++)
          for (int h = 0; h < nvoi; h++)
            ke[i * nd + j] += bmat[(h * nd + i) * ngp + g] * c[h * nvoi + k] *
                              bmat[(k * nd + j) * ngp + g] * wp[g];
  return 0;
}

/* ================================================================= mesh */

int or_mesh_init(or_mesh *m, int dim, int nex, int ney, int nez, double lx, double ly,
                 double lz)
{
  if (dim != 2 && dim != 3) return 1;
  if (nex < 1 || ney < 1 || (dim == 3 && nez < 1)) return 1;
  memset(m, 0, sizeof(*m));
  m->dim = dim;
  m->lx = lx;
  m->ly = ly;
  m->lz = (dim == 3) ? lz : 1;
  m->nex = nex;
  m->ney = ney;
  m->nez = (dim == 3) ? nez : 1;
  m->nx = nex + 1;
  m->ny = ney + 1;
  m->nz = (dim == 3) ? nez + 1 : 1;
  m->nn = m->nx * m->ny * m->nz;
  m->nelm = m->nex * m->ney * m->nez;
  m->hx = m->lx / m->nex;
  m->hy = m->ly / m->ney;
  m->hz = (dim == 2) ? 1 : m->lz / m->nez;
  m->vol = m->lx * m->ly * m->lz;
  m->vol_elm = m->hx * m->hy * m->hz;
  m->npe = (dim == 2) ? 4 : 8;
  m->ngp = m->npe;
  m->nvoi = (dim == 2) ? 3 : 6;
  return 0;
}

int or_elem_nodes(const or_mesh *m, int e, int *nods)
{
  int ex = e % m->nex;
  int ey = (e / m->nex) % m->ney;
  int ez = e / (m->nex * m->ney);
  int n0 = ex + ey * m->nx + ez * m->nx * m->ny;
  nods[0] = n0;
  nods[1] = n0 + 1;
  nods[2] = n0 + m->nx + 1;
  nods[3] = n0 + m->nx;
  if (m->dim == 3) {
    int s = m->nx * m->ny;
    for (int k = 0; k < 4; k++) nods[4 + k] = nods[k] + s;
  }
  return 0;
}

int or_elem_index(const or_mesh *m, int e, int *idx)
{
  int nods[8];
  or_elem_nodes(m, e, nods);
  for (int k = 0; k < m->npe; k++)
    for (int d = 0; d < m->dim; d++) idx[k * m->dim + d] = nods[k] * m->dim + d;
  return 0;
}

int or_node_coord(const or_mesh *m, int n, double *x)
{
  x[0] = (n % m->nx) * m->hx;
  x[1] = ((n / m->nx) % m->ny) * m->hy;
  if (m->dim == 3) x[2] = (n / (m->nx * m->ny)) * m->hz;
  return 0;
}

int or_elem_centroid(const or_mesh *m, int e, double *c)
{
  c[0] = (e % m->nex + 0.5) * m->hx;
  c[1] = ((e / m->nex) % m->ney + 0.5) * m->hy;
  if (m->dim == 3) c[2] = (e / (m->nex * m->ney) + 0.5) * m->hz;
  return 0;
}

int or_boundary_nodes(const or_mesh *m, int *nods)
{
  int cnt = 0;
  for (int n = 0; n < m->nn; n++) {
    int ix = n % m->nx, iy = (n / m->nx) % m->ny, iz = n / (m->nx * m->ny);
    int b = (ix == 0 || ix == m->nx - 1 || iy == 0 || iy == m->ny - 1);
    if (m->dim == 3) b = b || iz == 0 || iz == m->nz - 1;
    if (b) {
      if (nods) nods[cnt] = n;
      cnt++;
    }
  }
  return cnt;
}

int or_phase_map(const or_mesh *m, int nfib, const double *fx, const double *fy,
                 const double *fr, int *elem_type)
{
  for (int e = 0; e < m->nelm; e++) {
    double c[3];
    or_elem_centroid(m, e, c);
    elem_type[e] = 0;
    for (int f = 0; f < nfib; f++) {
      double l = 0.0;
      l = l + pow(c[0] - fx[f], 2);
      l = l + pow(c[1] - fy[f], 2);
      l = sqrt(l);
      if (l <= fr[f]) {
        elem_type[e] = 1;
        break;
      }
    }
  }
  return 0;
}

/* ================================================================== RVE */

#define OR_MAXMAT 8

struct or_rve {
  or_mesh mesh;
  int nd;                 /* npe*dim */
  int nmat;
  int *elem_type;
  double *bmat;           /* [nvoi][nd][ngp] */
  double wp[8];
  double c[OR_MAXMAT][36];
  double ke[OR_MAXMAT][24 * 24];
  int nbc;                /* boundary nodes */
  int *bc_nodes;
  double *bc_coord;       /* [nbc][dim] */
  or_ell K;
  int assembled;
};

or_rve *or_rve_create(int dim, int nex, int ney, int nez, double lx, double ly, double lz,
                      const int *elem_type, int nmat, const double *young,
                      const double *poisson)
{
  if (nmat < 1 || nmat > OR_MAXMAT) return NULL;
  or_rve *r = (or_rve *)calloc(1, sizeof(or_rve));
  if (or_mesh_init(&r->mesh, dim, nex, ney, nez, lx, ly, lz)) { free(r); return NULL; }
  or_mesh *m = &r->mesh;
  r->nd = m->npe * dim;
  r->nmat = nmat;
  r->elem_type = (int *)malloc((size_t)m->nelm * sizeof(int));
  for (int e = 0; e < m->nelm; e++) {
    int t = elem_type ? elem_type[e] : 0;
    if (t < 0 || t >= nmat) { or_rve_free(r); return NULL; }
    r->elem_type[e] = t;
  }
  double h[3] = {m->hx, m->hy, m->hz};
  r->bmat = (double *)malloc(sizeof(double) * (size_t)(m->nvoi * r->nd * m->ngp));
  or_struct_bmat(dim, h, r->bmat, r->wp);
  for (int k = 0; k < nmat; k++) {
    or_material_c(dim, young[k], poisson[k], r->c[k]);
    or_ke(dim, r->bmat, r->wp, r->c[k], r->ke[k]);
  }
  r->nbc = or_boundary_nodes(m, NULL);
  r->bc_nodes = (int *)malloc((size_t)r->nbc * sizeof(int));
  r->bc_coord = (double *)malloc((size_t)r->nbc * (size_t)dim * sizeof(double));
  or_boundary_nodes(m, r->bc_nodes);
  for (int i = 0; i < r->nbc; i++) {
    double x[3];
    or_node_coord(m, r->bc_nodes[i], x);
    for (int d = 0; d < dim; d++) r->bc_coord[i * dim + d] = x[d];
  }
  int nnz = (dim == 2) ? 18 : 81; /* micro/src/alloc.c:6 */
  if (or_ell_init(&r->K, m->nn * dim, m->nn * dim, nnz)) { or_rve_free(r); return NULL; }
  return r;
}

void or_rve_free(or_rve *r)
{
  if (!r) return;
  free(r->elem_type);
  free(r->bmat);
  free(r->bc_nodes);
  free(r->bc_coord);
  or_ell_free(&r->K);
  free(r);
}

const or_mesh *or_rve_mesh(const or_rve *r) { return &r->mesh; }
const or_ell *or_rve_ell(const or_rve *r) { return &r->K; }

int or_rve_assemble_jac(or_rve *r)
{
  const or_mesh *m = &r->mesh;
  int dim = m->dim, nd = r->nd;
  int idx[24];
  or_ell_set_zero_mat(&r->K);
  for (int e = 0; e < m->nelm; e++) {
    or_elem_index(m, e, idx);
    or_ell_add_vals(&r->K, idx, nd, idx, nd, r->ke[r->elem_type[e]]);
  }
  for (int d = 0; d < dim; d++) {
    for (int i = 0; i < r->nbc; i++) {
      int dof = r->bc_nodes[i] * dim + d;
      or_ell_set_zero_row(&r->K, dof, 1.0);
      or_ell_set_zero_col(&r->K, dof, 1.0);
    }
  }
  r->assembled = 1;
  return 0;
}

/* strain / stress at one gauss point of one element (micro/src/homogenize.c:416-461) */
static void gp_strain_stress(const or_rve *r, int e, int g, const double *ue, double *strain,
                             double *stress)
{
  const or_mesh *m = &r->mesh;
  int nvoi = m->nvoi, nd = r->nd, ngp = m->ngp;
  for (int v = 0; v < nvoi; v++) {
    strain[v] = 0.0;
    for (int i = 0; i < nd; i++) strain[v] += r->bmat[(v * nd + i) * ngp + g] * ue[i];
  }
  const double *c = r->c[r->elem_type[e]];
  for (int i = 0; i < nvoi; i++) {
    stress[i] = 0.0;
    for (int j = 0; j < nvoi; j++) stress[i] += c[i * nvoi + j] * strain[j];
  }
}

int or_rve_residual(or_rve *r, const double *x, double *res, double *norm)
{
  const or_mesh *m = &r->mesh;
  int nd = r->nd, nvoi = m->nvoi, ngp = m->ngp, dim = m->dim;
  int n = m->nn * dim;
  int idx[24];
  double ue[24], rese[24], strain[6], stress[6];
  for (int i = 0; i < n; i++) res[i] = 0.0;
  for (int e = 0; e < m->nelm; e++) {
    for (int i = 0; i < nd; i++) rese[i] = 0.0;
    or_elem_index(m, e, idx);
    for (int i = 0; i < nd; i++) ue[i] = x[idx[i]];
    for (int g = 0; g < ngp; g++) {
      gp_strain_stress(r, e, g, ue, strain, stress);
      for (int i = 0; i < nd; i++)
        for (int j = 0; j < nvoi; j++)
          rese[i] += r->bmat[(j * nd + i) * ngp + g] * stress[j] * r->wp[g];
    }
    for (int i = 0; i < nd; i++) res[idx[i]] += rese[i];
  }
  for (int d = 0; d < dim; d++)
    for (int i = 0; i < r->nbc; i++) res[r->bc_nodes[i] * dim + d] = 0.0;
  double s = 0;
  for (int i = 0; i < n; i++) s += res[i] * res[i];
  if (norm) *norm = sqrt(s);
  return 0;
}

int or_rve_averages(or_rve *r, const double *x, double *strain_ave, double *stress_ave)
{
  const or_mesh *m = &r->mesh;
  int nd = r->nd, nvoi = m->nvoi, ngp = m->ngp;
  int idx[24];
  double ue[24], strain[6], stress[6];
  for (int i = 0; i < nvoi; i++) { strain_ave[i] = 0.0; stress_ave[i] = 0.0; }
  for (int e = 0; e < m->nelm; e++) {
    or_elem_index(m, e, idx);
    for (int i = 0; i < nd; i++) ue[i] = x[idx[i]];
    for (int g = 0; g < ngp; g++) {
      gp_strain_stress(r, e, g, ue, strain, stress);
      for (int i = 0; i < nvoi; i++) {
        stress_ave[i] += stress[i] * r->wp[g];
        strain_ave[i] += strain[i] * r->wp[g];
      }
    }
  }
  for (int i = 0; i < nvoi; i++) {
    stress_ave[i] /= m->vol;
    strain_ave[i] /= m->vol;
  }
  return 0;
}

/* u = E.X with E_ii = eps_i and E_ij = gamma_ij/2 (micro/src/homogenize.c:335-342) */
static void strain_x_coord(int dim, const double *s, const double *x, double *u)
{
  if (dim == 2) {
    u[0] = s[0] * x[0] + s[2] / 2 * x[1];
    u[1] = s[2] / 2 * x[0] + s[1] * x[1];
  } else {
    u[0] = s[0] * x[0] + s[3] / 2 * x[1] + s[5] / 2 * x[2];
    u[1] = s[3] / 2 * x[0] + s[1] * x[1] + s[4] / 2 * x[2];
    u[2] = s[5] / 2 * x[0] + s[4] / 2 * x[1] + s[2] * x[2];
  }
}

int or_rve_set_disp_0(or_rve *r, const double *strain, double *x)
{
  int dim = r->mesh.dim;
  for (int i = 0; i < r->nbc; i++) {
    double u[3];
    strain_x_coord(dim, strain, &r->bc_coord[i * dim], u);
    for (int d = 0; d < dim; d++) x[r->bc_nodes[i] * dim + d] = u[d];
  }
  return 0;
}

int or_rve_homog_fe2(or_rve *r, const double *strain_mac, int solver, double rtol,
                     int nl_max_its, double nl_min_norm, double *strain_ave,
                     double *stress_ave, double *x_out, int *pcg_its_total, int *nl_its_out)
{
  const or_mesh *m = &r->mesh;
  int n = m->nn * m->dim;
  double *x = (double *)calloc((size_t)n, sizeof(double));
  double *res = (double *)malloc((size_t)n * sizeof(double));
  double *dx = (double *)malloc((size_t)n * sizeof(double));
  int ierr = 0, its_total = 0;
  or_rve_set_disp_0(r, strain_mac, x);
  int nl_its = 0;
  double res_norm = nl_min_norm * 10;
  while (nl_its < nl_max_its && res_norm > nl_min_norm) {
    or_rve_residual(r, x, res, &res_norm);
    if (res_norm < nl_min_norm) break;
    for (int i = 0; i < n; i++) res[i] *= -1.0;
    if (!r->assembled) or_rve_assemble_jac(r);
    for (int i = 0; i < n; i++) dx[i] = 0.0;
    int its = 0;
    if (solver == 0) {
      double rr;
      if (or_pcg(&r->K, res, dx, rtol, 100000, &its, &rr)) ierr = 5;
    } else {
      double err;
      or_ell_solve_jacobi(&r->K, res, dx, 100000, 1.0e-10, &its, &err);
    }
    its_total += its;
    for (int i = 0; i < n; i++) x[i] += dx[i];
    nl_its++;
  }
  or_rve_averages(r, x, strain_ave, stress_ave);
  if (x_out) memcpy(x_out, x, (size_t)n * sizeof(double));
  if (pcg_its_total) *pcg_its_total = its_total;
  if (nl_its_out) *nl_its_out = nl_its;
  free(x); free(res); free(dx);
  return ierr;
}

int or_rve_c_tangent_fd(or_rve *r, const double *strain_mac, double delta, double rtol, double *c)
{
  int nvoi = r->mesh.nvoi;
  double s1[6] = {0}, s2[6] = {0}, e1[6] = {0}, e2[6] = {0}, aux[6] = {0};
  int ierr = 0;
  for (int i = 0; i < nvoi; i++) e1[i] = strain_mac[i];
  ierr |= or_rve_homog_fe2(r, e1, 0, rtol, 2, 1.0e-4, aux, s1, NULL, NULL, NULL);
  for (int i = 0; i < nvoi; i++) {
    for (int k = 0; k < nvoi; k++) e2[k] = strain_mac[k];
    e2[i] = e2[i] + delta;
    ierr |= or_rve_homog_fe2(r, e2, 0, rtol, 2, 1.0e-4, aux, s2, NULL, NULL, NULL);
    for (int j = 0; j < nvoi; j++) c[j * nvoi + i] = (s2[j] - s1[j]) / (e2[i] - e1[i]);
  }
  return ierr;
}

int or_rve_c_tangent_unit(or_rve *r, double amp, double rtol, double *c, int *its_max)
{
  int nvoi = r->mesh.nvoi, ierr = 0, imax = 0;
  double e[6], aux[6], s[6];
  for (int i = 0; i < nvoi; i++) {
    for (int k = 0; k < nvoi; k++) e[k] = (k == i) ? amp : 0.0;
    int its = 0;
    ierr |= or_rve_homog_fe2(r, e, 0, rtol, 2, 1.0e-4, aux, s, NULL, &its, NULL);
    if (its > imax) imax = its;
    for (int j = 0; j < nvoi; j++) c[j * nvoi + i] = s[j] / amp;
  }
  if (its_max) *its_max = imax;
  return ierr;
}

/* Gauss-Jordan inverse with partial pivoting (n <= 6). */
static int inv_small(int n, const double *a, double *ainv)
{
  double w[6][12];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 2 * n; j++) w[i][j] = (j < n) ? a[i * n + j] : (j - n == i ? 1.0 : 0.0);
  for (int col = 0; col < n; col++) {
    int piv = col;
    for (int i = col + 1; i < n; i++)
      if (fabs(w[i][col]) > fabs(w[piv][col])) piv = i;
    if (w[piv][col] == 0.0) return 1;
    if (piv != col)
      for (int j = 0; j < 2 * n; j++) { double t = w[col][j]; w[col][j] = w[piv][j]; w[piv][j] = t; }
    double d = w[col][col];
    for (int j = 0; j < 2 * n; j++) w[col][j] /= d;
    for (int i = 0; i < n; i++) {
      if (i == col) continue;
      double f = w[i][col];
      for (int j = 0; j < 2 * n; j++) w[i][j] -= f * w[col][j];
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) ainv[i * n + j] = w[i][n + j];
  return 0;
}

int or_rve_taylor(or_rve *r, int mode, double *c)
{
  const or_mesh *m = &r->mesh;
  int nvoi = m->nvoi, nn = nvoi * nvoi;
  double vol[OR_MAXMAT] = {0};
  for (int e = 0; e < m->nelm; e++) vol[r->elem_type[e]] += m->vol_elm;
  double acc[36] = {0}, inv[36];
  for (int k = 0; k < r->nmat; k++) {
    double v = vol[k] / m->vol;
    if (mode == 0) {
      for (int i = 0; i < nn; i++) acc[i] += v * r->c[k][i];
    } else {
      if (inv_small(nvoi, r->c[k], inv)) return 1;
      for (int i = 0; i < nn; i++) acc[i] += v * inv[i];
    }
  }
  if (mode == 0) {
    for (int i = 0; i < nn; i++) c[i] = acc[i];
    return 0;
  }
  return inv_small(nvoi, acc, c);
}

/* ========================================================= periodic BC --
 * BC_PER_LM (micro/src/homogenize.c:244-260, micro/src/assembly.c:175-220,
 * 272-354, micro/src/alloc.c:15-16).  The reference writes it for 2D only:
 * corners get u = E.X, and every non-corner node of the high edge (x1 / y1)
 * is tied to its partner on the low edge by one Lagrange multiplier per dof,
 * u+ - u- = E.(X+ - X-).  The 3D extension used here: the 8 corners get
 * u = E.X; every other node with a coordinate on a high face (ix = nex,
 * iy = ney or iz = nez) is tied to the node at its wrapped position (each
 * index taken modulo the element count, which is never on a high face) by
 * u_s - u_m = E.(X_s - X_m), the multiplier entering the force balance as
 * +mu on the slave and -mu on the master rows (assembly.c:183-196).
 * The linear saddle-point system [K G^T; G 0] is solved densely (Gaussian
 * elimination with partial pivoting, constraint rows scaled by the largest
 * stiffness diagonal): small RVEs only.  x (length 3 nn, may be NULL)
 * receives the displacement. */
int or_rve_homog_per_lm(or_rve *r, const double *strain, double *strain_ave, double *stress_ave,
                        double *x_out)
{
  const or_mesh *m = &r->mesh;
  if (m->dim != 3) return 1;
  const int nx = m->nx, ny = m->ny, nz = m->nz, nn = m->nn, nd = r->nd;
  int *slave_of = (int *)malloc(sizeof(int) * (size_t)nn); /* master node, -1 if none, -2 corner */
  int nlm = 0;
  for (int n = 0; n < nn; n++) {
    const int ix = n % nx, iy = (n / nx) % ny, iz = n / (nx * ny);
    const int cx = (ix == 0 || ix == nx - 1), cy = (iy == 0 || iy == ny - 1), cz = (iz == 0 || iz == nz - 1);
    if (cx && cy && cz) {
      slave_of[n] = -2;
    } else if (ix == nx - 1 || iy == ny - 1 || iz == nz - 1) {
      slave_of[n] = (ix % m->nex) + (iy % m->ney) * nx + (iz % m->nez) * nx * ny;
      nlm++;
    } else {
      slave_of[n] = -1;
    }
  }
  const int nu = 3 * nn, N = nu + 3 * nlm;
  double *A = (double *)calloc((size_t)N * (size_t)(N + 1), sizeof(double));
  if (!A) { free(slave_of); return 2; }
#define AA(i, j) A[(size_t)(i) * (N + 1) + (j)]
  int idx[24];
  for (int e = 0; e < m->nelm; e++) {
    or_elem_index(m, e, idx);
    const double *ke = r->ke[r->elem_type[e]];
    for (int i = 0; i < nd; i++)
      for (int j = 0; j < nd; j++) AA(idx[i], idx[j]) += ke[i * nd + j];
  }
  double smax = 0.0;
  for (int i = 0; i < nu; i++) smax = fmax(smax, fabs(AA(i, i)));
  /* corners: identity rows with u = E.X */
  for (int n = 0; n < nn; n++) {
    if (slave_of[n] != -2) continue;
    double X[3], u[3];
    or_node_coord(m, n, X);
    strain_x_coord(3, strain, X, u);
    for (int d = 0; d < 3; d++) {
      for (int j = 0; j <= N; j++) AA(3 * n + d, j) = 0.0;
      AA(3 * n + d, 3 * n + d) = smax;
      AA(3 * n + d, N) = smax * u[d];
    }
  }
  /* multiplier rows and columns */
  int l = nu;
  for (int n = 0; n < nn; n++) {
    const int mm = slave_of[n];
    if (mm < 0) continue;
    double Xs[3], Xm[3], dX[3], u[3];
    or_node_coord(m, n, Xs);
    or_node_coord(m, mm, Xm);
    for (int d = 0; d < 3; d++) dX[d] = Xs[d] - Xm[d];
    strain_x_coord(3, strain, dX, u);
    for (int d = 0; d < 3; d++, l++) {
      AA(l, 3 * n + d) = smax;
      AA(l, 3 * mm + d) = -smax;
      AA(l, N) = smax * u[d];
      if (slave_of[n] >= 0) AA(3 * n + d, l) += smax;
      if (slave_of[mm] != -2) AA(3 * mm + d, l) -= smax;
    }
  }
#undef AA
  /* Gaussian elimination with partial pivoting */
  int ierr = 0;
  for (int c = 0; c < N && !ierr; c++) {
    int p = c;
    for (int i = c + 1; i < N; i++)
      if (fabs(A[(size_t)i * (N + 1) + c]) > fabs(A[(size_t)p * (N + 1) + c])) p = i;
    if (A[(size_t)p * (N + 1) + c] == 0.0) { ierr = 3; break; }
    if (p != c)
      for (int j = c; j <= N; j++) {
        double t = A[(size_t)c * (N + 1) + j];
        A[(size_t)c * (N + 1) + j] = A[(size_t)p * (N + 1) + j];
        A[(size_t)p * (N + 1) + j] = t;
      }
    const double piv = A[(size_t)c * (N + 1) + c];
    for (int i = c + 1; i < N; i++) {
      const double f = A[(size_t)i * (N + 1) + c] / piv;
      if (f == 0.0) continue;
      for (int j = c; j <= N; j++) A[(size_t)i * (N + 1) + j] -= f * A[(size_t)c * (N + 1) + j];
    }
  }
  double *sol = (double *)malloc(sizeof(double) * (size_t)N);
  for (int i = N - 1; i >= 0 && !ierr; i--) {
    double t = A[(size_t)i * (N + 1) + N];
    for (int j = i + 1; j < N; j++) t -= A[(size_t)i * (N + 1) + j] * sol[j];
    sol[i] = t / A[(size_t)i * (N + 1) + i];
  }
  if (!ierr) {
    or_rve_averages(r, sol, strain_ave, stress_ave);
    if (x_out) memcpy(x_out, sol, sizeof(double) * (size_t)nu);
  }
  free(sol);
  free(A);
  free(slave_of);
  return ierr;
}

/* ============================================================= macro hex8 --
 * Isoparametric hex8 of the macro assembly (macro/src/assembly.c:105-130 for
 * the stiffness, :20-41 for the residual, :227-244 for dsh), with the
 * Jacobian, its cofactor inverse and the derivative transform of
 * src/fem.c:411-471 and the hex8 tables of src/fem.c:309-405 (weights 1).
 * coor[8][3] in the hex8 node order; B is the 3D Voigt extension of
 * assembly_get_bmat (xx, yy, zz, xy, yz, xz).  wdet[gp] = |det J|. */
int or_macro_bmat(const double *coor, double *bmat, double *detj)
{
  double xp[8][3], wp[8], sh[8][8], ds[8][3][8];
  or_hex8_tables(xp, wp, sh, ds);
  for (int i = 0; i < 6 * 24 * 8; i++) bmat[i] = 0.0;
  for (int g = 0; g < 8; g++) {
    double J[3][3], c[3][3], iJ[3][3], det;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        J[i][j] = 0.0;
        for (int n = 0; n < 8; n++) J[i][j] += ds[n][i][g] * coor[n * 3 + j];
      }
    c[0][0] = +J[1][1] * J[2][2] - J[2][1] * J[1][2];
    c[0][1] = -J[1][0] * J[2][2] + J[2][0] * J[1][2];
    c[0][2] = +J[1][0] * J[2][1] - J[2][0] * J[1][1];
    c[1][0] = -J[0][1] * J[2][2] + J[2][1] * J[0][2];
    c[1][1] = +J[0][0] * J[2][2] - J[2][0] * J[0][2];
    c[1][2] = -J[0][0] * J[2][1] + J[2][0] * J[0][1];
    c[2][0] = +J[0][1] * J[1][2] - J[1][1] * J[0][2];
    c[2][1] = -J[0][0] * J[1][2] + J[1][0] * J[0][2];
    c[2][2] = +J[0][0] * J[1][1] - J[1][0] * J[0][1];
    det = J[0][0] * c[0][0] + J[0][1] * c[0][1] + J[0][2] * c[0][2];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) iJ[i][j] = c[j][i] / det;
    detj[g] = fabs(det);
    for (int a = 0; a < 8; a++) {
      double d[3];
      for (int i = 0; i < 3; i++) {
        d[i] = 0.0;
        for (int j = 0; j < 3; j++) d[i] += iJ[i][j] * ds[a][j][g];
      }
#define BM(v, i) bmat[((v) * 24 + (i)) * 8 + g]
      BM(0, a * 3 + 0) = d[0];
      BM(1, a * 3 + 1) = d[1];
      BM(2, a * 3 + 2) = d[2];
      BM(3, a * 3 + 0) = d[1];
      BM(3, a * 3 + 1) = d[0];
      BM(4, a * 3 + 1) = d[2];
      BM(4, a * 3 + 2) = d[1];
      BM(5, a * 3 + 0) = d[2];
      BM(5, a * 3 + 2) = d[0];
#undef BM
    }
  }
  return 0;
}

/* k_elem += B[h][i] C[h][k] B[k][j] wp |detJ| per gauss point, loops gp, i,
 * j, k, h (macro/src/assembly.c:173-181); c_gp[gp][36] row-major. */
int or_macro_ke(const double *coor, const double *c_gp, double *ke)
{
  double bmat[6 * 24 * 8], detj[8];
  or_macro_bmat(coor, bmat, detj);
  for (int i = 0; i < 24 * 24; i++) ke[i] = 0.0;
  for (int g = 0; g < 8; g++)
    for (int i = 0; i < 24; i++)
      for (int j = 0; j < 24; j++)
        for (int k = 0; k < 6; k++)
          for (int h = 0; h < 6; h++)
            ke[i * 24 + j] += bmat[(h * 24 + i) * 8 + g] * c_gp[g * 36 + h * 6 + k] * bmat[(k * 24 + j) * 8 + g] *
                              1.0 * detj[g];
  return 0;
}

/* Strain at the gauss points from the element displacement; components
 * below 1e-6 in magnitude are set to zero (macro/src/assembly.c:307-313). */
int or_macro_strain(const double *coor, const double *ue, double *strain)
{
  double bmat[6 * 24 * 8], detj[8];
  or_macro_bmat(coor, bmat, detj);
  for (int g = 0; g < 8; g++)
    for (int v = 0; v < 6; v++) {
      double s = 0.0;
      for (int i = 0; i < 24; i++) s += bmat[(v * 24 + i) * 8 + g] * ue[i];
      strain[g * 6 + v] = (fabs(s) < 1.0e-6) ? 0.0 : s;
    }
  return 0;
}

/* res_elem[i] += B[j][i] stress_gp[j] wp |detJ| (macro/src/assembly.c:37-40) */
int or_macro_res(const double *coor, const double *stress, double *res)
{
  double bmat[6 * 24 * 8], detj[8];
  or_macro_bmat(coor, bmat, detj);
  for (int i = 0; i < 24; i++) res[i] = 0.0;
  for (int g = 0; g < 8; g++)
    for (int i = 0; i < 24; i++)
      for (int j = 0; j < 6; j++) res[i] += bmat[(j * 24 + i) * 8 + g] * stress[g * 6 + j] * 1.0 * detj[g];
  return 0;
}

/* MatZeroRowsColumns(A, n, rows, diag) on an ELL (macro/src/assembly.c:143-148):
 * every listed row keeps its columns with value 0 except diag on the
 * diagonal, and every listed column is zeroed in every other row. */
int or_ell_zero_rows_cols(or_ell *m, int n, const int *rows, double diag)
{
  char *mark = (char *)calloc((size_t)m->ncol, 1);
  for (int k = 0; k < n; k++) {
    if (rows[k] < 0 || rows[k] >= m->nrow || rows[k] >= m->ncol) { free(mark); return 1; }
    mark[rows[k]] = 1;
  }
  for (int i = 0; i < m->nrow; i++)
    for (int j = 0; j < m->nnz; j++) {
      const size_t s = (size_t)i * m->nnz + j;
      const int c = m->cols[s];
      if (c < 0) break;
      if (i < m->ncol && mark[i]) m->vals[s] = (c == i) ? diag : 0.0;
      else if (mark[c]) m->vals[s] = 0.0;
    }
  free(mark);
  return 0;
}
