/*
 * macro_api.h -- the macro side of the FE^2 coupling on the GPU: assembly of
 * the macro Jacobian and residual of a hex8 mesh from the tangents and
 * stresses the micro returns per gauss point (macro/src/assembly.c:5-222),
 * in the reference's fixed-width ELL (81 slots per row, first-touch column
 * order of src/ell.c:46-84 over the element loop), and its Jacobi-PCG solve.
 *
 * The mesh is a structured nex x ney x nez hex8 block (node n = ix + iy nx +
 * iz nx ny, element nodes in the hex8 order of src/fem.c:324-361) whose node
 * coordinates may be moved (distorted elements): B and |det J| come from the
 * element's nodal coordinates at each gauss point (isoparametric,
 * src/fem.c:411-471, macro/src/assembly.c:227-244).  Dirichlet dofs are
 * eliminated like PETSc's MatZeroRowsColumns (identity row, zero column,
 * macro/src/assembly.c:143-148).  Gauss-point arrays are [elem][8][...].
 */
#ifndef MADMAX_MACRO_API_H_
#define MADMAX_MACRO_API_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mm_macro mm_macro;

/* coords: [node][3] (nx ny nz nodes) or NULL for the regular grid over
 * [0,lx] x [0,ly] x [0,lz]. */
int mm_macro_create(int nex, int ney, int nez, double lx, double ly, double lz, const double *coords, int device,
                    mm_macro **out);
int mm_macro_destroy(mm_macro *m);
int mm_macro_sizes(const mm_macro *m, int *nrow, int *nelm);
int mm_macro_set_dirichlet(mm_macro *m, int ndof, const int *dofs);
/* the macro displacement (length nrow) */
int mm_macro_set_u(mm_macro *m, const double *u);
/* assembly_get_strain at every gauss point: strain_gp [elem][8][6], with
 * components below 1e-6 set to zero like macro/src/assembly.c:312 */
int mm_macro_strain(mm_macro *m, double *strain_gp);
/* assembly_b (macro/src/assembly.c:5-85): b = -(sum_e B^T sigma_gp w |det J|)
 * with the Dirichlet entries zeroed before the sign change, and |b|_2 */
int mm_macro_assemble_b(mm_macro *m, const double *stress_gp, double *b, double *norm);
/* assembly_A (macro/src/assembly.c:164-222) with c_gp [elem][8][36] */
int mm_macro_assemble_a(mm_macro *m, const double *c_gp);
int mm_macro_get_ell(mm_macro *m, int *cols, double *vals);
/* du solving A du = b (Jacobi-PCG from zero, relative tolerance rtol) */
int mm_macro_solve(mm_macro *m, const double *b, double *du, double rtol, int *its);

#ifdef __cplusplus
}
#endif
#endif
