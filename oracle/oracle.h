/*
 * oracle.h -- FP64 CPU restatement of the sputnik micro (RVE) homogenization
 * path.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and the
 * CPU-baseline leg of the benchmark load it as the reference they compare
 * with.  The product library (madmax_amd) never links, loads or calls
 * anything in this folder.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the reference root).  The reference implements only 2D for the structured
 * micro; the 3D conventions below are straight extensions of the 2D code and
 * are documented next to each function.
 */
#ifndef MADMAX_ORACLE_H
#define MADMAX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ ELL --
 * Same struct shape as inc/ell.h:15-28. */
typedef struct {
  int nrow, ncol, nnz;
  int *cols;
  double *vals;
} or_ell;

int or_ell_init(or_ell *m, int nrow, int ncol, int nnz);          /* src/ell.c:3-15   */
void or_ell_free(or_ell *m);
int or_ell_set_val(or_ell *m, int row, int col, double val);       /* src/ell.c:17-44  */
int or_ell_add_val(or_ell *m, int row, int col, double val);       /* src/ell.c:46-73  */
int or_ell_add_vals(or_ell *m, const int *ix, int nx, const int *iy, int ny,
                    const double *vals);                            /* src/ell.c:75-84  */
int or_ell_set_zero_row(or_ell *m, int row, double diag_val);      /* src/ell.c:86-101 */
int or_ell_set_zero_col(or_ell *m, int col, double diag_val);      /* src/ell.c:103-118*/
int or_ell_set_zero_mat(or_ell *m);                                /* src/ell.c:120-124*/
int or_ell_mvp(const or_ell *m, const double *x, double *y);       /* src/ell.c:127-141*/
int or_ell_get_val(const or_ell *m, int row, int col, double *val);/* src/ell.c:186-208*/
int or_ell_solve_jacobi(const or_ell *m, const double *b, double *x,
                        int max_its, double min_err, int *its, double *err); /* src/ell.c:143-184 */
/* Jacobi-preconditioned CG (Saad, Iterative Methods, Alg. 9.1); the
 * reference reaches it through PETSc -ksp_type cg -pc_type jacobi
 * (micro/src/homogenize.c:325-326).  Stops at ||r_k||_2 <= rtol*||b||_2 with
 * the recursively updated residual.  x holds the initial guess on entry. */
int or_pcg(const or_ell *m, const double *b, double *x, double rtol, int max_its,
           int *its, double *relres);

/* ------------------------------------------------------------ FEM tables --
 * hexa8: src/fem.c:309-405 (gauss points, weights, sh[node][gp],
 * ds[node][dim][gp]).  quad4 structured: src/fem.c:3-59. */
void or_hex8_tables(double xp[8][3], double wp[8], double sh[8][8], double ds[8][3][8]);
/* Structured physical derivative tables and weights for dim 2 or 3:
 * dsh[(node*dim + d)*ngp + gp] and wp[gp] = hx*hy(*hz)/ngp
 * (src/fem.c:42-54 in 2D; 3D uses ds_hexa_8 * 2 / h_d). */
int or_struct_dsh(int dim, const double *h, double *dsh, double *wp);
/* B[(v*npe*dim + i)*ngp + gp] (micro/src/init.c:40-50 in 2D; 3D Voigt order
 * xx, yy, zz, xy, yz, xz with engineering shears). */
int or_struct_bmat(int dim, const double *h, double *bmat, double *wp);
/* Linear elastic tangent (src/material.c:16-22, 47-53 in 2D plane strain;
 * 3D isotropic in the same form).  c is nvoi x nvoi row-major. */
int or_material_c(int dim, double young, double poisson, double *c);
/* Element stiffness in the reference loop order (micro/src/assembly.c:97-107):
 * ke[i*nd+j] += ((B[h][i]*C[h][k])*B[k][j])*wp, loops gp, i, j, k, h. */
int or_ke(int dim, const double *bmat, const double *wp, const double *c, double *ke);

/* ------------------------------------------------------------ mesh ------ */
typedef struct {
  int dim;
  int nex, ney, nez, nx, ny, nz;
  int nn, nelm, npe, nvoi, ngp;
  double lx, ly, lz, hx, hy, hz, vol, vol_elm;
} or_mesh;

/* src/mesh.c:762-786 with element counts (not node counts) as input. */
int or_mesh_init(or_mesh *m, int dim, int nex, int ney, int nez,
                 double lx, double ly, double lz);
int or_elem_nodes(const or_mesh *m, int e, int *nods);      /* src/mesh.c:878-888 */
int or_elem_index(const or_mesh *m, int e, int *idx);       /* src/mesh.c:890-901 */
int or_node_coord(const or_mesh *m, int n, double *x);      /* src/mesh.c:869-876 */
int or_elem_centroid(const or_mesh *m, int e, double *c);   /* micro/src/homogenize.c:487-494 */
/* All nodes on the boundary faces, ascending node id (src/mesh.c:815-864). */
int or_boundary_nodes(const or_mesh *m, int *nods);

/* Phase map: z-aligned cylinders, centroid distance <= r is fiber
 * (micro/src/micro_struct.c:98-129).  elem_type 0 = matrix, 1 = fiber. */
int or_phase_map(const or_mesh *m, int nfib, const double *fx, const double *fy,
                 const double *fr, int *elem_type);

/* ------------------------------------------------------------ RVE ------- */
typedef struct or_rve or_rve;

/* nmat materials (0 = MATRIX, 1 = FIBER), elem_type per element. */
or_rve *or_rve_create(int dim, int nex, int ney, int nez, double lx, double ly, double lz,
                      const int *elem_type, int nmat, const double *young,
                      const double *poisson);
void or_rve_free(or_rve *r);
const or_mesh *or_rve_mesh(const or_rve *r);
/* Jacobian assembly into ELL with the USTRAIN boundary rows/cols
 * (micro/src/assembly.c:84-137; W = 18 in 2D, 81 in 3D, micro/src/alloc.c:6). */
int or_rve_assemble_jac(or_rve *r);
const or_ell *or_rve_ell(const or_rve *r);
/* Residual sum_e sum_gp B^T sigma w with boundary dofs zeroed, and its
 * 2-norm (micro/src/assembly.c:33-82).  Displacement gather through the
 * element connectivity (the PETSc branch, micro/src/homogenize.c:407). */
int or_rve_residual(or_rve *r, const double *x, double *res, double *norm);
/* Volume averages (micro/src/homogenize.c:344-368). */
int or_rve_averages(or_rve *r, const double *x, double *strain_ave, double *stress_ave);
/* Affine boundary displacements x[bc] = E.X (micro/src/homogenize.c:227-242, 335-342). */
int or_rve_set_disp_0(or_rve *r, const double *strain, double *x);
/* homog_fe2 (micro/src/homogenize.c:175-225) with the ELL solver replaced by
 * Jacobi-PCG to relative tolerance rtol (solver = 0) or the reference's
 * stationary Jacobi (solver = 1).  x (length nrow) is the output displacement
 * (may be NULL).  Newton limits nl_max_its / nl_min_norm as in
 * micro/src/init.c:7-8. */
int or_rve_homog_fe2(or_rve *r, const double *strain_mac, int solver, double rtol,
                     int nl_max_its, double nl_min_norm, double *strain_ave,
                     double *stress_ave, double *x, int *pcg_its_total, int *nl_its);
/* homog_calculate_c_tangent (micro/src/homogenize.c:59-94): forward
 * differences with step delta (HOMOGENIZE_DELTA_STRAIN = 0.005,
 * micro/inc/micro.h:29); c[j*nvoi + i] = d sigma_j / d eps_i. */
int or_rve_c_tangent_fd(or_rve *r, const double *strain_mac, double delta, double rtol,
                        double *c);
/* Tangent from nvoi unit load cases eps_i = amp * e_i (linear materials):
 * c[j*nvoi + i] = <sigma_j>(amp e_i) / amp. */
int or_rve_c_tangent_unit(or_rve *r, double amp, double rtol, double *c, int *its_max);
/* Voigt (mode 0, MIXP) / Reuss (mode 1, MIXS) mixture bounds
 * (micro/src/homogenize.c:96-163); LU inverse restated with Gauss-Jordan. */
int or_rve_taylor(or_rve *r, int mode, double *c);

/* BC_PER_LM homogenization (micro/src/assembly.c:175-220, 272-354, 2D in the
 * reference; 3D extension in oracle.c), solved densely: small RVEs only. */
int or_rve_homog_per_lm(or_rve *r, const double *strain, double *strain_ave, double *stress_ave,
                        double *x);

/* ------------------------------------------------------------ macro hex8 --
 * Isoparametric hex8 (macro/src/assembly.c, src/fem.c:411-471): B and
 * |det J| at the 8 gauss points from nodal coordinates coor[8][3],
 * k_elem with a tangent per gauss point, the gauss-point strains (with the
 * 1e-6 flush of assembly.c:312) and the element residual from gauss-point
 * stresses. */
int or_macro_bmat(const double *coor, double *bmat, double *detj);
int or_macro_ke(const double *coor, const double *c_gp, double *ke);
int or_macro_strain(const double *coor, const double *ue, double *strain);
int or_macro_res(const double *coor, const double *stress, double *res);
/* MatZeroRowsColumns semantics on an ELL (macro/src/assembly.c:143-148). */
int or_ell_zero_rows_cols(or_ell *m, int n, const int *rows, double diag);

#ifdef __cplusplus
}
#endif
#endif
