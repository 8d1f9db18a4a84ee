/*
 * oracle.h -- CPU restatement of PyLith's finite-element residual / Jacobian
 * hot path, used ONLY as a checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  It is never linked into the product library.
 *
 * What it restates (reference file:line):
 *   - pointwise kernels with the PetscPointFn / PetscPointJacFn / PetscBdPointFn
 *     signature (libsrc/pylith/fekernels/ headers, cited per function in oracle.c): elastic and
 *     generalized Maxwell f1 / Jf3, body force and gravity f0, state projections, fault rows;
 *   - PETSc's DMPlexComputeResidualByKey / DMPlexComputeJacobianActionByKey cell
 *     loop as PyLith drives it (libsrc/pylith/feassemble/IntegratorDomain.cc:448-449,
 *     :540-541): closure gather, per-quadrature-point geometry, pointwise kernel
 *     call through a function pointer, sum_q w_q |J_q| (f0 phi + f1 : grad phi),
 *     closure scatter-add (ADD_VALUES) into a local vector;
 *   - the hybrid (cohesive-cell) residual of IntegratorInterface.cc:603-660;
 *   - DMLocalToGlobal with ADD_VALUES dropping constrained dofs (topology/Field.cc:394-404).
 *
 * PetscInt is int32 and PetscScalar/PetscReal are double, as in the reference's
 * PETSc configuration (docker/pylith-testenv:51).
 */
#ifndef PLX_ORACLE_H
#define PLX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*oracle_point_fn)(int dim, int numS, int numA,
                                const int sOff[], const int sOff_x[],
                                const double s[], const double s_t[], const double s_x[],
                                const int aOff[], const int aOff_x[],
                                const double a[], const double a_t[], const double a_x[],
                                double t, const double x[],
                                int numConstants, const double constants[],
                                double f[]);

typedef void (*oracle_point_jac_fn)(int dim, int numS, int numA,
                                    const int sOff[], const int sOff_x[],
                                    const double s[], const double s_t[], const double s_x[],
                                    const int aOff[], const int aOff_x[],
                                    const double a[], const double a_t[], const double a_x[],
                                    double t, double s_tshift, const double x[],
                                    int numConstants, const double constants[],
                                    double J[]);

typedef void (*oracle_bd_point_fn)(int dim, int numS, int numA,
                                   const int sOff[], const int sOff_x[],
                                   const double s[], const double s_t[], const double s_x[],
                                   const int aOff[], const int aOff_x[],
                                   const double a[], const double a_t[], const double a_x[],
                                   double t, const double x[], const double n[],
                                   int numConstants, const double constants[],
                                   double f[]);

/* Cell types (same numbering as include/pylith_amd.h). */
enum { ORACLE_HEX_Q1 = 1, ORACLE_TET_P2 = 2 };

/* Kernel ids (same numbering as include/pylith_amd.h plx_pointfn). */
enum {
    ORACLE_FN_NONE = 0,
    ORACLE_F1V_ISOLINEAR_ELASTICITY_3D = 1,
    ORACLE_JF3VU_ISOLINEAR_ELASTICITY_3D = 2,
    ORACLE_F1V_GENMAXWELL_3D = 3,
    ORACLE_JF3VU_GENMAXWELL_3D = 4,
    ORACLE_VISCOUS_STRAIN_GENMAXWELL_3D = 5,
    ORACLE_TOTAL_STRAIN_3D = 6,
    ORACLE_F0V_BODYFORCE = 7,
    ORACLE_F0V_GRAV = 8,
    ORACLE_F0V_GRAVBODYFORCE = 9,
    ORACLE_F1V_ISOLINEAR_ELASTICITY_REFSTATE_3D = 10,
    /* test-only kernels for the device path's PetscPointFn-form plugin launchers (oracle.c) */
    ORACLE_TEST_F0 = 90,
    ORACLE_TEST_F1 = 91,
    ORACLE_TEST_JF0 = 92,
    ORACLE_TEST_JF3 = 93
};

/* Mesh + discretization + auxiliary field of one integrator key. */
typedef struct {
    int cell_type;            /* ORACLE_HEX_Q1 or ORACLE_TET_P2 */
    int num_cells;
    const int32_t *closure;   /* [num_cells][nodes_per_cell]: node ids; dof = 3*node+c */
    const int32_t *cone;      /* [num_cells][verts_per_cell]: vertex ids into coords */
    const double *coords;     /* [num_vertices][3] */
    int num_aux;              /* numA */
    const int32_t *aux_sizes; /* [num_aux] components per aux subfield */
    int aux_per_qp;           /* 0: aux[cell][naux]; 1: aux[cell][nq][naux] */
    const double *aux;
    int num_constants;
    const double *constants;
} oracle_mesh;

/* Test hook: plant a known mistake in the viscous-strain update (0 resets; see oracle.c). */
void oracle_plant_maxwell_error(int kind);

int oracle_nodes_per_cell(int cell_type);
int oracle_verts_per_cell(int cell_type);
int oracle_num_quad(int cell_type);
/* Reference quadrature: points [nq][3], weights [nq]. */
int oracle_quadrature(int cell_type, double *points, double *weights);
/* Tabulation at reference point xi: basis [nb], reference gradients [nb][3]. */
int oracle_tabulate(int cell_type, const double xi[3], double *phi, double *dphi);

oracle_point_fn oracle_point_kernel(int id);
oracle_point_jac_fn oracle_point_jac_kernel(int id);

/* r_local += residual over `cells` (all cells when cells == NULL). Returns 0 or an error code
 * (PETSc style): 1 = bad argument, 2 = non-positive Jacobian determinant. */
int oracle_compute_residual(const oracle_mesh *m, const int32_t *cells, int ncells,
                            int f0_id, int f1_id, double t,
                            const double *locX, const double *locX_t, double *locF);

/* f_local += J(locX) * locY, with J from the Jf0..Jf3 kernels (ids; 0 = NULL). */
int oracle_compute_jacobian_action(const oracle_mesh *m, const int32_t *cells, int ncells,
                                   int jf0_id, int jf1_id, int jf2_id, int jf3_id,
                                   double t, double s_tshift,
                                   const double *locX, const double *locX_t,
                                   const double *locY, double *locF);

/* Point-block (3x3) diagonal of J: blocks[node][3][3] += ... */
int oracle_compute_jacobian_pbdiag(const oracle_mesh *m, const int32_t *cells, int ncells,
                                   int jf3_id, double t, double s_tshift,
                                   const double *locX, double *blocks);

/* Dense element matrix of one cell (nd x nd, nd = 3*nodes_per_cell). */
int oracle_element_matrix(const oracle_mesh *m, int cell, int jf3_id, double t, double s_tshift,
                          const double *locX, double *Ke);

/* State-variable projection at quadrature points (IntegratorDomain.cc:555-590):
 * aux_out[cell][q][...] = aux_in with subfield `sub` replaced by kernel(id) values. */
int oracle_update_state(const oracle_mesh *m, int kernel_id, int subfield, double t,
                        const double *locX, double *aux_out);

/* Cohesive cells (hybrid integration, IntegratorInterface.cc:603-660), prescribed slip.
 *   closure_neg/pos/lag: [ncoh][4] node ids; dofs: u at 3*node, lagrange at lag_offset + 3*lag_node.
 *   face_cone: [ncoh][4] vertex ids (negative side) for geometry; slip: [num_lag_nodes][3]
 *   (opening, left-lateral, reverse) at the Lagrange nodes, interpolated with the face basis. */
typedef struct {
    int num_cells;
    const int32_t *closure_neg;
    const int32_t *closure_pos;
    const int32_t *closure_lag;
    const int32_t *face_cone;
    const double *coords;
    int lag_offset;
    const double *slip;          /* [num_lag_nodes][3] */
    double ref_dir_1[3];
    double ref_dir_2[3];
} oracle_cohesive;

int oracle_cohesive_residual(const oracle_cohesive *c, double t, const double *locX, double *locF);
int oracle_cohesive_jacobian_action(const oracle_cohesive *c, const double *locY, double *locF);

/* Tangent directions of BoundaryDirections::tangential_directions (BoundaryDirections.hh:50-82). */
void oracle_tangential_directions(double t1[3], double t2[3], const double r1[3],
                                  const double r2[3], const double n[3]);

/* DMLocalToGlobal(ADD_VALUES): g[l2g[i]] += loc[i] for l2g[i] >= 0. */
void oracle_local_to_global(int nlocal, const int32_t *l2g, const double *loc, double *g);
/* DMGlobalToLocal(INSERT_VALUES) on unconstrained dofs. */
void oracle_global_to_local(int nlocal, const int32_t *l2g, const double *g, double *loc);

/* Colored multi-threaded residual (the timed CPU baseline): colors[cell] in [0,ncolors). */
int oracle_compute_residual_colored(const oracle_mesh *m, const int32_t *color_ptr,
                                    const int32_t *color_cells, int ncolors,
                                    int f1_id, double t, const double *locX, double *locF,
                                    int nthreads);

#ifdef __cplusplus
}
#endif
#endif
