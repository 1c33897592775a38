// macro.cpp -- C-ABI of the macro assembly (include/macro_api.h) over the
// element assembly of fe.hip in its isoparametric mode.
#include <cstring>
#include <vector>

#include "fe.h"
#include "macro_api.h"

struct mm_macro {
  mm::Fe *fe = nullptr;
  int nrow = 0, nelm = 0;
};

extern "C" int mm_macro_create(int nex, int ney, int nez, double lx, double ly, double lz, const double *coords,
                               int device, mm_macro **out)
{
  if (!out || nex < 1 || ney < 1 || nez < 1 || !(lx > 0) || !(ly > 0) || !(lz > 0)) return MM_ERR_ARG;
  const int nx = nex + 1, ny = ney + 1, nz = nez + 1;
  std::vector<double> grid;
  if (!coords) {  // the regular grid, still integrated isoparametrically
    grid.resize((size_t)nx * ny * nz * 3);
    for (int n = 0; n < nx * ny * nz; n++) {
      grid[3 * (size_t)n] = (n % nx) * (lx / nex);
      grid[3 * (size_t)n + 1] = ((n / nx) % ny) * (ly / ney);
      grid[3 * (size_t)n + 2] = (n / (nx * ny)) * (lz / nez);
    }
    coords = grid.data();
  }
  double cdummy[36];
  for (int i = 0; i < 36; i++) cdummy[i] = (i % 7 == 0) ? 1.0 : 0.0;  // replaced by the gauss-point tangents
  mm::FeMeshDesc d;
  memset(&d, 0, sizeof(d));
  d.nex = nex;
  d.ney = ney;
  d.nez = nez;
  d.lx = lx;
  d.ly = ly;
  d.lz = lz;
  d.coords = coords;
  d.nmat = 1;
  d.cmat = cdummy;
  d.flush_strain = 1;
  d.device = device;
  mm_macro *m = new mm_macro();
  const int rc = mm::fe_create(d, &m->fe);
  if (rc) {
    delete m;
    return rc;
  }
  mm::fe_sizes(m->fe, &m->nrow, &m->nelm, nullptr);
  *out = m;
  return 0;
}

extern "C" int mm_macro_destroy(mm_macro *m)
{
  if (!m) return MM_ERR_ARG;
  mm::fe_destroy(m->fe);
  delete m;
  return 0;
}

extern "C" int mm_macro_sizes(const mm_macro *m, int *nrow, int *nelm)
{
  if (!m) return MM_ERR_ARG;
  if (nrow) *nrow = m->nrow;
  if (nelm) *nelm = m->nelm;
  return 0;
}

extern "C" int mm_macro_set_dirichlet(mm_macro *m, int ndof, const int *dofs)
{
  if (!m) return MM_ERR_ARG;
  return mm::fe_set_dirichlet(m->fe, ndof, dofs, 1);
}

extern "C" int mm_macro_set_u(mm_macro *m, const double *u)
{
  if (!m || !u) return MM_ERR_ARG;
  return mm::fe_set_x(m->fe, u);
}

extern "C" int mm_macro_strain(mm_macro *m, double *strain_gp)
{
  if (!m || !strain_gp) return MM_ERR_ARG;
  return mm::fe_gp_strain(m->fe, nullptr, strain_gp);
}

extern "C" int mm_macro_assemble_b(mm_macro *m, const double *stress_gp, double *b, double *norm)
{
  if (!m || !stress_gp || !b) return MM_ERR_ARG;
  double nrm = 0.0;
  int rc = mm::fe_residual(m->fe, nullptr, stress_gp, &nrm);
  if (rc) return rc;
  if ((rc = mm::fe_get_residual(m->fe, b))) return rc;
  for (int i = 0; i < m->nrow; i++) b[i] = -b[i];  // VecScale(b, -1.0)
  if (norm) *norm = nrm;
  return 0;
}

extern "C" int mm_macro_assemble_a(mm_macro *m, const double *c_gp)
{
  if (!m || !c_gp) return MM_ERR_ARG;
  int rc = mm::fe_set_c_gp(m->fe, c_gp);
  if (rc) return rc;
  return mm::fe_assemble(m->fe);
}

extern "C" int mm_macro_get_ell(mm_macro *m, int *cols, double *vals)
{
  if (!m) return MM_ERR_ARG;
  return mm::fe_get_ell(m->fe, cols, vals);
}

extern "C" int mm_macro_solve(mm_macro *m, const double *b, double *du, double rtol, int *its)
{
  if (!m || !b || !du || !(rtol > 0)) return MM_ERR_ARG;
  return mm::fe_solve(m->fe, b, du, rtol, 100000, its);
}
