// fe.h -- per-element finite-element assembly on the GPU (internal; the
// C-ABI that uses it is in micro.cpp and macro.cpp).
//
// A structured hex8 mesh of nex x ney x nez elements.  Unknowns live on the
// structured nodes, or, with `periodic`, on the nex x ney x nez nodes of the
// torus (a node's unknown is the one at its index modulo the element count);
// then the displacement is u = E X + u~ with u~ the unknown fluctuation.
// Element stiffness matrices are formed per element from a tangent per
// gauss point (B^T C B), B either from the structured hex8 tables (the
// micro, micro/src/init.c:37-62) or isoparametrically from the element's
// nodal coordinates (the macro, macro/src/assembly.c:227-244).  They are
// summed into the reference's fixed-width ELL (first-touch column order of
// src/ell.c:46-84 over the element loop) by one lane per node, in ascending
// element order, so the assembled values equal the reference's bit for bit.
#pragma once

#include <cstdint>

#include "ell.h"
#include "mm_internal.h"

namespace mm {

struct FeMeshDesc {
  int nex, ney, nez;
  double lx, ly, lz;
  int periodic;         // 1: periodic fluctuation with node (0,0,0) pinned
  const double *coords; // host [structured node][3], or null (regular grid)
  const int *elem_mat;  // host [elem] material index, or null (all 0)
  int nmat;
  const double *cmat;   // host [mat][36] tangent per material (row-major)
  int flush_strain;     // 1: gauss-point strain components below 1e-6 set to 0
                        //    (macro/src/assembly.c:312)
  int device;
};

struct Fe;  // opaque

int fe_create(const FeMeshDesc &d, Fe **out);
void fe_destroy(Fe *f);
int fe_sizes(const Fe *f, int *nrow, int *nelm, int *nnode);
// Dirichlet dofs (identity rows); zero_cols also zeroes their columns in the
// other rows (PETSc MatZeroRowsColumns) instead of keeping them (the ELL
// path's ell_set_zero_col, which returns at the first padded slot).
int fe_set_dirichlet(Fe *f, int ndof, const int *dofs, int zero_cols);
// USTRAIN: every dof of the six faces (micro/src/assembly.c:112-135).
int fe_set_dirichlet_boundary(Fe *f);
// Displacement (device copy, length nrow) from / to host.
int fe_set_x(Fe *f, const double *x);
int fe_get_x(Fe *f, double *x);
// x on the Dirichlet dofs = E X (set_disp_0, micro/src/homogenize.c:227-242);
// the other dofs are left as they are.
int fe_set_disp_0(Fe *f, const double *strain);
// Tangent per gauss point [elem][8][36] (null: the element's material).
int fe_set_c_gp(Fe *f, const double *c_gp);
// Residual r = sum_e B^T sigma w (sigma from the gauss-point tangent and the
// strain of x, or stress_gp [elem][8][6] when given), Dirichlet dofs zeroed;
// its 2-norm.  `strain` is the macro strain of the affine part (periodic).
int fe_residual(Fe *f, const double *strain, const double *stress_gp, double *norm);
int fe_get_residual(Fe *f, double *r);
// Gauss-point strains of x [elem][8][6] (with the macro flush if set).
int fe_gp_strain(Fe *f, const double *strain, double *out);
// Element stiffness matrices and their ELL sum (with the Dirichlet rows).
int fe_assemble(Fe *f);
int fe_get_ell(Fe *f, int *cols, double *vals);
// dx solving K dx = -r (Jacobi-PCG, relative tolerance rtol), x += dx.
int fe_solve_update(Fe *f, double rtol, int max_its, int *its);
// K du = b for a host right-hand side (Jacobi-PCG from zero).
int fe_solve(Fe *f, const double *b, double *du, double rtol, int max_its, int *its);
// Volume averages of strain and stress (micro/src/homogenize.c:344-368);
// elem_strain / elem_stress ([elem][6], may be null) receive each element's
// average over its gauss points.
int fe_averages(Fe *f, const double *strain, double *strain_ave, double *stress_ave, double *elem_strain,
                double *elem_stress);
int fe_get_phase(Fe *f, int *elem_type);

// Device ELL of the reference layout (ell_device.hip): allocation without a
// host matrix, its arrays, and the Jacobi-PCG on device vectors.
int ell_device_alloc(int nrow, int nnz, int S, ell_device **out);
void ell_device_arrays(ell_device *d, int **cols, double **vals, double **dinv, int *S, hipStream_t *s);
int ell_device_pcg_device(ell_device *d, ell_solver *solver, const double *b, double *x);

}  // namespace mm
