/*
 * pylith_amd.h -- C ABI of the MI355X (gfx950) finite-element residual / Jacobian-action path.
 *
 * The entry points replace, one for one, the PETSc calls PyLith's assembly layer makes
 * (libsrc/pylith/feassemble/IntegratorDomain.cc, IntegratorInterface.cc) and the field
 * scatters of libsrc/pylith/topology/Field.cc.  Every function returns an int error code
 * (0 = success, as PetscErrorCode); plx_dm_last_error() gives the message of the last
 * failure on a handle.  No C++ exception crosses this boundary.
 *
 * Vectors (locX, locF, ...) are DEVICE pointers owned by the caller, laid out like a PETSc
 * local vector of PyLith's solution section: the displacement field interleaved per mesh node
 * (dof = 3*node + component), followed by the Lagrange-multiplier field when a fault is present
 * (dof = plx_dm lagrange offset + 3*fault_node + component).  Global vectors drop constrained
 * dofs in ascending local order (PETSc's global section of a single process).
 *
 * All work is enqueued on the HIP stream bound to the handle (plx_dm_set_stream); calls return
 * without waiting for the device unless documented otherwise.  One handle per host thread.
 */
#ifndef PYLITH_AMD_H
#define PYLITH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct plx_dm_s *plx_dm;

/* Cell types (PyLith mesh files: hex vertex order 0-3 one face cyclic, 4-7 opposite face). */
enum plx_cell_type { PLX_HEX_Q1 = 1, PLX_TET_P2 = 2 };

/* Scalar type of the vectors and of the arithmetic of one call. */
enum plx_scalar_type { PLX_FP64 = 0, PLX_FP32 = 1 };

/* ADD_VALUES accumulates (PETSc semantics); INSERT_VALUES overwrites the dofs in the closure
 * of the key's cells, i.e. fuses the caller's zeroing of the local residual
 * (problems/TimeDependent.cc:657) with the assembly. */
enum plx_insert_mode { PLX_ADD_VALUES = 0, PLX_INSERT_VALUES = 1 };

/* pylith::feassemble::Integrator::EquationPart (feassemble/Integrator.hh:26-31). */
enum plx_equation_part { PLX_LHS = 0, PLX_RHS = 1, PLX_LHS_LUMPED_INV = 2, PLX_LHS_WEIGHTED = 3 };

/* Device twins of PyLith pointwise kernels.  A host integration maps the PetscPointFn /
 * PetscPointJacFn address it registers to one of these ids; kernels outside this list are
 * compiled for the device against pylith_amd_pointfn.hpp and registered by address. */
enum plx_pointfn {
    PLX_FN_NONE = 0,
    PLX_F1V_ISOLINEAR_ELASTICITY_3D = 1,      /* IsotropicLinearElasticity3D::f1v_infinitesimalStrain  (IsotropicLinearElasticity.hh:803) */
    PLX_JF3VU_ISOLINEAR_ELASTICITY_3D = 2,    /* IsotropicLinearElasticity3D::Jf3vu_infinitesimalStrain (IsotropicLinearElasticity.hh:890) */
    PLX_F1V_GENMAXWELL_3D = 3,                /* IsotropicLinearGenMaxwell3D::f1v_infinitesimalStrain  (IsotropicLinearGenMaxwell.hh:1054) */
    PLX_JF3VU_GENMAXWELL_3D = 4,              /* IsotropicLinearGenMaxwell3D::Jf3vu_infinitesimalStrain (IsotropicLinearGenMaxwell.hh:1143) */
    PLX_VISCOUS_STRAIN_GENMAXWELL_3D = 5,     /* IsotropicLinearGenMaxwell3D::viscousStrain_infinitesimalStrain_asVector (:1507) */
    PLX_TOTAL_STRAIN_3D = 6,                  /* Elasticity3D::infinitesimalStrain_asVector (Elasticity.hh:586) */
    /* f0 of the quasi-static displacement equation (materials/Elasticity.cc:455-475); aux
     * subfields density(1), then body_force(3) and/or gravitational_acceleration(3), then the
     * rheology's subfields (materials/Elasticity.cc:234-241). */
    PLX_F0V_BODYFORCE = 7,                    /* Elasticity::g0v_bodyforce     (Elasticity.hh:300) */
    PLX_F0V_GRAV = 8,                         /* Elasticity::g0v_grav          (Elasticity.hh:249) */
    PLX_F0V_GRAVBODYFORCE = 9,                /* Elasticity::g0v_gravbodyforce (Elasticity.hh:346) */
    /* f1 with reference stress and strain (IsotropicLinearElasticity::useReferenceState); aux
     * [..., reference_stress(6), reference_strain(6), shear_modulus, bulk_modulus]; its Jacobian
     * is PLX_JF3VU_ISOLINEAR_ELASTICITY_3D (materials/IsotropicLinearElasticity.cc:120-131). */
    PLX_F1V_ISOLINEAR_ELASTICITY_REFSTATE_3D = 10 /* IsotropicLinearElasticity3D::f1v_infinitesimalStrain_refState (IsotropicLinearElasticity.hh:848) */
};

/* ---------------------------------------------------------------------------------------------
 * Mesh, section and closure index.
 * Replaces DMPlexCreateFromCellListPetsc (meshio/MeshBuilder.cc:155) + the solution section
 * (topology/Field.cc:277-283) + DMPlexCreateClosureIndex (topology/CoordsVisitor.icc:145-157).
 *   closure [num_cells][nodes_per_cell] node ids (hex: 8 vertices; tet P2: 4 vertices then the
 *           edge nodes of edges (0,1),(0,2),(0,3),(1,2),(1,3),(2,3));
 *   cone    [num_cells][verts_per_cell] vertex ids into coords (NULL: the first verts_per_cell
 *           closure entries);
 *   coords  [num_vertices][3];
 *   labels  [num_cells] material label values (NULL: all 0).
 * All of these are HOST arrays; they are copied.  Fails with an error for a cell whose
 * Jacobian determinant is not positive at some quadrature point. */
int plx_dm_create(plx_dm *dm, int cell_type, int num_cells, int num_vertices, int num_nodes,
                  const int32_t *closure, const int32_t *cone, const double *coords,
                  const int32_t *labels);
int plx_dm_destroy(plx_dm dm);
/* plx_dm_create with the caller's vertex order of the hexahedra (Q1): our vertex i of every cell
 * is the caller's vertex vertex_perm[i] of that cell (NULL: the order above).  A DMPlex cone made
 * by DMPlexInvertCell from the mesh file's order (meshio/MeshBuilder.cc:151; PETSc swaps
 * hexahedron vertices 1 and 3) is mapped back by {0, 3, 2, 1, 4, 5, 6, 7}; cells are then checked
 * for positive Jacobians as in plx_dm_create, so a wrong order is refused. */
int plx_dm_create_ordered(plx_dm *dm, int cell_type, int num_cells, int num_vertices, int num_nodes,
                          const int32_t *closure, const int32_t *cone, const double *coords,
                          const int32_t *labels, const int32_t *vertex_perm);
/* The caller's quadrature-point order of per-point auxiliary fields (call before registering
 * kernels): our point q is the caller's point quad_perm[q]; per-point aux given to
 * plx_dm_set_auxiliary and returned by plx_dm_get_auxiliary is in the caller's order (NULL: ours).
 * Ours: Q1 q = i + 2 j + 4 k at reference coordinates (-1/sqrt3, +1/sqrt3)[i, j, k] (xi fastest),
 * P2 the 14-point rule's order; a Q1 rule ordered xi slowest (q' = 4 i + 2 j + k) maps by
 * {0, 4, 2, 6, 1, 5, 3, 7}. */
int plx_dm_set_point_order(plx_dm dm, int num_quad, const int32_t *quad_perm);
int plx_dm_set_stream(plx_dm dm, void *hip_stream);
const char *plx_dm_last_error(plx_dm dm);
/* Local and global vector lengths (dofs). */
int plx_dm_get_sizes(plx_dm dm, int64_t *local_size, int64_t *global_size);

/* Element algorithm and in-workgroup accumulation of the cell kernels.
 *   PLX_ALGO_QUADRATURE_POINTS: the device twin of the pointwise kernel is evaluated at every
 *     quadrature point (Q1: sum-factorized 2x2x2 Gauss; P2: 14-point rule);
 *   PLX_ALGO_FACTORED: for pointwise kernels linear in the solution gradient with per-cell
 *     coefficients, on affine cells, the same quadrature sum is evaluated through the polynomial
 *     structure of the gradient (Q1: the 8 Walsh-Hadamard modes with the Gauss orthogonality
 *     weights; P2: the 4 vertex gradients with the rule's P1 moment matrix); other cells fall
 *     back to quadrature points.  Identical results up to rounding;
 *   PLX_ALGO_POINTFN: the generic quadrature-point path of include/pylith_amd_pointfn.hpp,
 *     which calls the device kernels in the PetscPointFn / PetscPointJacFn parameter list (the
 *     library's own twins of the reference kernels, written in that form; LDS-staged tiles with
 *     the accumulation below, or one thread per cell colour by colour under
 *     PLX_ACCUMULATE_COLORED, bitwise reproducible).  Elastic, reference-state and body-force /
 *     gravity keys; generalized Maxwell keys keep the tile kernels;
 *   PLX_ALGO_AUTO (default): FACTORED where it applies.
 *   PLX_ACCUMULATE_COLORED: element vectors enter the tile's LDS node buffer colour by colour, a
 *     fixed order: results are bitwise reproducible;
 *   PLX_ACCUMULATE_LDS_ATOMIC: LDS atomic adds into an FP64 accumulator (FP32 calls round once
 *     when the tile writes out), no colour phases; the last bits may vary run to run;
 *   PLX_ACCUMULATE_AUTO (default): LDS atomics.
 * min_waves_per_simd (0 = default 2): occupancy the cell kernels' register allocation must
 * allow (2: <= 256 VGPRs, 3: <= 168, 4: <= 128); the rest spills. */
enum plx_algorithm { PLX_ALGO_AUTO = 0, PLX_ALGO_QUADRATURE_POINTS = 1, PLX_ALGO_FACTORED = 2, PLX_ALGO_POINTFN = 3 };
enum plx_accumulate { PLX_ACCUMULATE_AUTO = 0, PLX_ACCUMULATE_COLORED = 1, PLX_ACCUMULATE_LDS_ATOMIC = 2 };
int plx_dm_set_kernel_options(plx_dm dm, int algorithm, int accumulate, int min_waves_per_simd);

/* Constrained dofs (Dirichlet, ConstraintSimple/ConstraintSpatialDB; PetscDSAddBoundary at
 * feassemble/ConstraintSimple.cc:122): `dofs` are local dof indices (HOST array). */
int plx_dm_set_constraints(plx_dm dm, int num_constrained, const int32_t *dofs);

/* ---------------------------------------------------------------------------------------------
 * Weak form: PetscWeakFormAddResidual / PetscWeakFormAddJacobian as called by
 * IntegratorDomain::setKernelsResidual / setKernelsJacobian (IntegratorDomain.cc:145-146,183-184).
 * The key is (label_value, part) of the cells' material label.  f0 is PLX_FN_NONE or one of the
 * PLX_F0V_* body force / gravity kernels; f1 an elastic, reference-state or generalized Maxwell
 * f1.  Terms that do not depend on the solution (the f0 load, sigma_ref - C : eps_ref) are
 * integrated once on the host when the kernels or the aux field change and added after the cell
 * pass of every residual; the Jacobian action does not see them. */
int plx_dm_add_residual(plx_dm dm, int label_value, int part, int f0, int f1);
int plx_dm_add_jacobian(plx_dm dm, int label_value, int part, int jf0, int jf1, int jf2, int jf3);

/* DMSetAuxiliaryVec (IntegratorDomain.cc:447).  aux is a HOST array laid out like PyLith's
 * auxiliary local vector: per cell (aux_per_qp = 0) or per cell and quadrature point
 * (aux_per_qp = 1), subfields concatenated in order with the given component counts.
 * The values are captured (copied to the device) at this call.
 * Isotropic linear elasticity: [..., shear_modulus, bulk_modulus] (IsotropicLinearElasticity.hh:
 * 116-117).  Generalized Maxwell (Q1, per quadrature point): [..., shear_modulus, bulk_modulus,
 * maxwell_time(3), shear_modulus_ratio(3), viscous_strain(18), total_strain(6)]
 * (IsotropicLinearGenMaxwell.hh:109-165); moduli, times and ratios uniform within a cell. */
int plx_dm_set_auxiliary(plx_dm dm, int label_value, int part, int num_aux,
                         const int32_t *aux_sizes, int aux_per_qp, const double *aux);

/* PetscDSSetConstants via Integrator::_setKernelConstants (Integrator.cc:264-288).
 * Generalized Maxwell: constants[0] = dt (IsotropicLinearGenMaxwell.cc:211-212); setting it
 * recomputes the Jacobian's effective shear modulus and the state stress of the step. */
int plx_dm_set_constants(plx_dm dm, int label_value, int part, int num_constants,
                         const double *constants);

/* State variables: IntegratorDomain::updateStateVars -> _updateStateVars
 * (IntegratorDomain.cc:555-590, DMProjectFieldLocal of the viscous and total strain kernels,
 * materials/IsotropicLinearGenMaxwell.cc:219-244) at the solution locX (FP64, device) of the
 * step just solved.  No-op for keys without state variables. */
int plx_update_state_vars(plx_dm dm, int label_value, int part, double t, const double *locX);
/* Copy of the key's auxiliary field in the layout given to plx_dm_set_auxiliary, with the
 * current state variables (HOST array; synchronises the stream). */
int plx_dm_get_auxiliary(plx_dm dm, int label_value, int part, double *aux);

/* ---------------------------------------------------------------------------------------------
 * DMPlexComputeResidualByKey(dm, key, cellsIS, t, locX, locX_t, t, locF)
 * (IntegratorDomain.cc:448-449): locF (+)= sum over the key's cells of
 * w|J| (f0 . phi + f1 : grad phi).  locX_t may be NULL (quasi-static).  scalar_type selects
 * FP64 or FP32 vectors and arithmetic. */
int plx_compute_residual(plx_dm dm, int label_value, int part, int scalar_type, double t,
                         const void *locX, const void *locX_t, void *locF, int mode);

/* Allocation-free variant of plx_compute_residual.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory during the call (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments, semantics and results
 * as plx_compute_residual, but the shared-node partial sums of the apply live in a file-scope
 * device array of the library (64 MiB per device, held until the process ends) instead of the
 * key's own buffer, and the call allocates nothing and records no event timers.  Fails with an
 * error when the key needs more than 64 MiB of partial sums (halo slots x 3 x scalar size;
 * about a 140^3 Q1 mesh in FP64).  All calls on one device share the array: calls that overlap on
 * one device (several streams or threads) race and must be serialised by the caller. */
int plx_compute_residual_noalloc(plx_dm dm, int label_value, int part, int scalar_type, double t,
                                 const void *locX, const void *locX_t, void *locF, int mode);

/* DMPlexComputeJacobianActionByKey(dm, key, cellsIS, t, s_tshift, locX, locX_t, locY, locF)
 * (IntegratorDomain.cc:540-541): locF (+)= J(locX) locY without assembling J. */
int plx_compute_jacobian_action(plx_dm dm, int label_value, int part, int scalar_type,
                                double t, double s_tshift, const void *locX, const void *locX_t,
                                const void *locY, void *locF, int mode);

/* Point-block (3x3 per node) diagonal of J, FP64, blocks[node][3][3] (ADD).  The blocks of
 * PETSc's vpbjacobi preconditioner that PyLith selects (materials/Elasticity.cc:307-318). */
int plx_compute_jacobian_pbdiag(plx_dm dm, int label_value, int part, double t,
                                double s_tshift, const double *locX, double *blocks);

/* ---------------------------------------------------------------------------------------------
 * Assembled Jacobian: DMPlexComputeJacobianByKey(dm, key, cellsIS, t, s_tshift, locX, locX_t,
 * J, P) as IntegratorDomain::computeLHSJacobian calls it (IntegratorDomain.cc:458-504, call at
 * :489-490; Integrator.hh:189-191).
 * The matrix is the displacement block in 3x3 blocks over the mesh nodes: block (i, j) exists
 * when nodes i and j share a cell of the mesh (the blocked pattern DMCreateMatrix gives the
 * displacement field, block size 3).  plx_jacobian_pattern returns it as CSR over the nodes:
 * row_ptr [num_nodes + 1] and col [nnzb], columns ascending in each row (HOST arrays; pass NULL
 * to query nnzb only).  plx_compute_jacobian adds (ADD_VALUES) or writes (INSERT_VALUES: all
 * blocks zeroed first) the key's element matrices
 *   K_ab[f][g] = sum_q w|J| (phi_a Jf0[f][g] phi_b + phi_a Jf1[f][g][d] dphi_b[d]
 *                + dphi_a[d] Jf2[f][g][d] phi_b + dphi_a[d] Jf3[f][g][d][e] dphi_b[e])
 * into values [nnzb][3][3] (DEVICE array, FP64, row-major blocks in pattern order), cell colour
 * after cell colour (deterministic).  Rows and columns of constrained dofs are included; a
 * caller inserting through PETSc's local-to-global map drops them (their global index is
 * negative), as DMPlexComputeJacobian does.  Several keys (materials) add into one matrix. */
int plx_jacobian_pattern(plx_dm dm, int64_t *nnzb, int32_t *row_ptr, int32_t *col);
int plx_compute_jacobian(plx_dm dm, int label_value, int part, double t, double s_tshift,
                         const double *locX, const double *locX_t, double *values, int mode);

/* ---------------------------------------------------------------------------------------------
 * Pointwise kernels supplied by the caller in the PetscPointFn / PetscPointJacFn parameter list
 * (docs/developer/implementation/petsc-fem.md:447-490; e.g. IsotropicLinearElasticity.hh:
 * 804-821).  The caller writes them as __device__ functions and instantiates the launchers of
 * include/pylith_amd_pointfn.hpp in its own hipcc translation unit; the launchers are
 * registered on a key here, replacing PetscWeakFormAddResidual / PetscWeakFormAddJacobian
 * (IntegratorDomain.cc:145-146,183-184) for kernels without a built-in twin.  The DM hands a
 * launcher everything the kernels see: */
typedef struct {
    int cell_type, nodes_per_cell, verts_per_cell, num_quad;
    int num_cells;                   /* cells of the key */
    const int32_t *cells;            /* DEVICE [num_cells]: the key's cells, colour after colour */
    int num_colors;
    const int32_t *color_ptr;        /* HOST [num_colors + 1]: cells of colour k share no node */
    const int32_t *closure;          /* DEVICE [num_mesh_cells][nodes_per_cell] */
    const int32_t *cone;             /* DEVICE [num_mesh_cells][verts_per_cell] */
    const double *coords;            /* DEVICE [num_vertices][3] */
    int num_aux;                     /* auxiliary subfields (numA) */
    int aux_off[16];                 /* aOff: offsets of the subfields in a record */
    int aux_stride;                  /* doubles per record */
    int aux_per_qp;                  /* records per cell (0) or per cell and quadrature point (1) */
    const double *aux;               /* DEVICE, the layout given to plx_dm_set_auxiliary */
    int num_constants;
    const double *constants;         /* DEVICE [num_constants] */
    double t, s_tshift;
    const double *u, *u_t;           /* DEVICE local vectors (u_t may be NULL: zero) */
    const double *v;                 /* DEVICE direction of a Jacobian action */
    double *f;                       /* DEVICE output local vector (residual / action) */
    int mode;                        /* ADD_VALUES / INSERT_VALUES */
    const int32_t *row_ptr, *col;    /* DEVICE Jacobian pattern (assembly launchers) */
    double *values;                  /* DEVICE Jacobian values [nnzb][3][3] */
    /* The key's tile plan (residual and action launchers): each tile's cells, its nodes (owned
     * ones first, then the shared ones whose partial sums leave through the halo slots) and its
     * vertices; DEVICE arrays.  deterministic = 1 (PLX_ACCUMULATE_COLORED) asks for the colour
     * passes, whose summation order is fixed. */
    int deterministic;
    int num_tiles, tile_max_nodes, tile_max_verts;
    const int32_t *tile_cell_ptr, *tile_node_ptr, *tile_vert_ptr;  /* [num_tiles + 1] */
    const int32_t *tile_nown, *tile_halo_off;                      /* [num_tiles] */
    const int32_t *tile_nodes, *tile_verts;  /* mesh node / vertex of each tile slot */
    const int32_t *tile_cells;               /* [num_cells] the key's cells in tile order */
    const uint16_t *lclosure, *lcone;        /* [nodes_per_cell | verts_per_cell][num_cells] tile slots */
    int num_shared;
    const int32_t *shared_node, *shared_ptr, *shared_slot;  /* shared nodes and their halo slots */
    double *halo;                            /* [halo slots][3] partial sums */
} plx_pointfn_args;
typedef int (*plx_pointfn_launcher)(const plx_pointfn_args *args, void *hip_stream);
/* residual: f (+)= sum w|J| (f0 . phi + f1 : grad phi).  action: f (+)= J(u) v.  assemble:
 * values (+)= element matrices.  NULL launchers leave that operation unregistered. */
int plx_dm_add_residual_pointfn(plx_dm dm, int label_value, int part, plx_pointfn_launcher residual);
int plx_dm_add_jacobian_pointfn(plx_dm dm, int label_value, int part, plx_pointfn_launcher action,
                                plx_pointfn_launcher assemble);

/* ---------------------------------------------------------------------------------------------
 * Field scatters (topology/Field.cc:394-422) and Dirichlet insertion
 * (DMPlexInsertBoundaryValuesEssential, feassemble/ConstraintSimple.cc:161-162). */
/* DMLocalToGlobal: glob (+)= loc over unconstrained dofs. */
int plx_local_to_global(plx_dm dm, int scalar_type, const void *loc, void *glob, int mode);
/* DMGlobalToLocal(INSERT): loc[unconstrained] = glob; constrained entries untouched. */
int plx_global_to_local(plx_dm dm, int scalar_type, const void *glob, void *loc);
/* loc[constrained dof i] = values[i] (device array, order of plx_dm_set_constraints). */
int plx_insert_boundary_values(plx_dm dm, int scalar_type, const void *values, void *loc);

/* ---------------------------------------------------------------------------------------------
 * Fault with prescribed slip: the cohesive cells of FaultCohesiveKin integrated by
 * IntegratorInterface (faults/FaultCohesiveKin.cc:330-461, feassemble/IntegratorInterface.cc:
 * 603-722).  Q1 hex meshes.  Per cohesive cell [num_cohesive][4], in the face's cyclic vertex
 * order: closure_neg / closure_pos the split displacement nodes of the negative / positive side,
 * closure_lag the fault (Lagrange-multiplier) node, face_cone the vertex ids of the face
 * geometry (into the mesh coords).  Adds 3*num_lagrange dofs after the displacement field;
 * constraints set before are kept.  ref_dir_1/2 are the reference directions of the fault
 * frame (FaultCohesive::setRefDir1/2, BoundaryDirections.hh:50-82).  HOST arrays, copied. */
int plx_dm_add_fault(plx_dm dm, int num_cohesive, const int32_t *closure_neg,
                     const int32_t *closure_pos, const int32_t *closure_lag,
                     const int32_t *face_cone, int num_lagrange, const double ref_dir_1[3],
                     const double ref_dir_2[3]);
/* Slip per fault node [num_lagrange][3] in the fault frame (opening = normal, left-lateral =
 * t1, reverse = t2), the "slip" auxiliary subfield that FaultCohesiveKin::_updateSlip writes
 * (FaultCohesiveKin.cc:264-324).  _device takes a device array (stream-ordered copy). */
int plx_dm_set_fault_slip(plx_dm dm, const double *slip);
int plx_dm_set_fault_slip_device(plx_dm dm, const double *slip);
int plx_dm_get_fault_info(plx_dm dm, int64_t *lagrange_offset, int *num_lagrange);
/* DMPlexComputeResidualHybridByKey (IntegratorInterface.cc:653) with the kernels f0u_neg,
 * f0u_pos, f0l_slip: locF (+)= the cohesive cells' residual, FP64.  INSERT_VALUES zeroes the
 * Lagrange-multiplier block first and adds into the displacement block (the domain residual
 * owns those entries). */
int plx_compute_residual_cohesive(plx_dm dm, double t, const double *locX, double *locF,
                                  int mode);
/* DMPlexComputeJacobianHybridByKey's blocks Jf0ul_neg, Jf0ul_pos, Jf0lu
 * (IntegratorInterface.cc:714) applied to locY, FP64 (same mode rule). */
int plx_compute_jacobian_action_cohesive(plx_dm dm, const double *locY, double *locF, int mode);

/* ---------------------------------------------------------------------------------------------
 * Reductions and BLAS-1 on device vectors, FP64 accumulation (deterministic two-pass tree).
 * Results of dot/norm are written to a device scalar and copied to *result (synchronises the
 * handle's stream). */
int plx_vec_dot(plx_dm dm, int scalar_type, int64_t n, const void *x, const void *y,
                double *result);
int plx_vec_norm2(plx_dm dm, int scalar_type, int64_t n, const void *x, double *result);

/* ---------------------------------------------------------------------------------------------
 * Linear solves around the matrix-free operator A = -J of one key (SPD for elasticity with the
 * DM's constrained dofs held fixed), point-block Jacobi preconditioned.  Replaces the KSP solve
 * of SNESSolve in TimeDependent::solve / GreensFns::solve (problems/TimeDependent.cc:475-484,
 * problems/GreensFns.cc:326).
 *   PLX_KSP_CG        FP64 preconditioned CG (the plain FP64 path);
 *   PLX_KSP_FGMRES_IR FP64 FGMRES on the true residual, each preconditioner application an FP32
 *                     PCG solve (FP32 apply inside FP64 iterative refinement);
 *   PLX_KSP_IR        Richardson iterative refinement with the FP32 PCG correction.
 * x (local FP64, device) holds the initial guess with the boundary values in the constrained dofs
 * and returns the solution; b (local FP64, device) may be NULL (zero).  Convergence:
 * ||b - A x|| <= rtol ||b - A x0|| on the unconstrained dofs, the residual always recomputed in
 * FP64 at the end. */
typedef struct plx_solver_s *plx_solver;
enum plx_ksp_type { PLX_KSP_CG = 0, PLX_KSP_FGMRES_IR = 1, PLX_KSP_IR = 2 };
typedef struct {
    int outer_iterations;   /* FGMRES / IR / CG iterations */
    int inner_iterations;   /* FP32 PCG iterations, all preconditioner applications */
    int applies_fp64;       /* FP64 operator applications */
    int converged;
    double initial_residual;
    double final_residual;  /* true FP64 residual norm */
} plx_solve_stats;
int plx_solver_create(plx_solver *solver, plx_dm dm, int label_value, int part);
int plx_solver_destroy(plx_solver solver);
/* (Re)compute the point-block preconditioner (after the aux field or constraints change). */
int plx_solver_setup(plx_solver solver);
/* Inner FP32 PCG relative tolerance / iteration cap and the FGMRES restart length. */
int plx_solver_set_inner(plx_solver solver, double rtol, int maxit, int restart);
/* Preconditioner of every PCG the solver runs (the FP64 CG and the FP32 inner PCG): the inverse
 * 3x3 point blocks (default; PyLith's vpbjacobi fine smoother) or one FP32 V-cycle of
 * smoothed-aggregation multigrid on the assembled operator (the GAMG role of
 * materials/Elasticity.cc:307-318; keys covering every cell, Q1 or P2).  Without a fault the
 * finest level of the cycle smooths with the solver's matrix-free FP32 operator.
 * plx_solver_setup rebuilds the hierarchy from the current operator. */
enum { PLX_PC_PBJACOBI = 0, PLX_PC_AMG = 1 };
int plx_solver_set_preconditioner(plx_solver solver, int kind);
int plx_solver_solve(plx_solver solver, int method, const double *b, double *x, double rtol,
                     int maxit, plx_solve_stats *stats);

/* ---------------------------------------------------------------------------------------------
 * Batched Green's functions: GreensFns (problems/GreensFns.cc:236-331) with the impulses of
 * FaultCohesiveImpulses (faults/FaultCohesiveImpulses.cc:263-300) -- one linear solve per
 * fault-slip impulse, all sharing the operator.  The operator of the key (the key covering
 * every cell) is assembled once at create as 3x3-block CSR over the nodes, with the DM's
 * constrained dofs and, when the DM has a fault, the split nodes merged (the constraint-
 * eliminated operator); Q1 from its element matrices, P2 (no fault) from the assembled
 * Jacobian (plx_compute_jacobian); plx_gf_solve needs a planar fault (Q1).  Impulses are solved k at a time
 * (k in {8, 16, 32, 64}, autotuned when 0) by FP64 iterative refinement around an FP32 block
 * PCG, each block solved to ||b - A x|| <= rtol ||b|| per impulse in FP64.
 *   slip [num_impulses][num_lagrange][3] (device, fault frame), u [num_impulses][local size]
 *   (device): the displacement of each impulse (Lagrange-multiplier entries set to zero). */
typedef struct plx_gf_s *plx_gf;
/* label_value of plx_solver_create / plx_gf_create selecting every key of the part that has
 * Jacobian kernels and an auxiliary field (one key per material): the operator is their sum. */
#define PLX_ALL_KEYS (-1)
typedef struct {
    int block;                     /* block width used */
    int num_blocks;
    int outer_iterations;          /* FP64 refinement steps, summed over blocks */
    int inner_iterations;          /* FP32 block PCG iterations, summed over blocks */
    int converged;
    double max_relative_residual;  /* worst ||b - A x|| / ||b|| over the impulses, FP64 */
} plx_gf_stats;
int plx_gf_create(plx_gf *gf, plx_dm dm, int label_value, int part);
int plx_gf_destroy(plx_gf gf);
int plx_gf_set_options(plx_gf gf, int block, double inner_rtol, int inner_maxit, int use_graph);
/* Preconditioner of the FP32 block PCG: point-block Jacobi (default) or smoothed-aggregation
 * algebraic multigrid, the role of the GAMG preconditioner PyLith configures for elasticity
 * (materials/Elasticity.cc:307-318): aggregates of the node graph, rigid-body near-null space,
 * smoothed prolongator, Galerkin coarse operators, one V-cycle with Chebyshev smoothing per
 * iteration.  The hierarchy is built on the first switch to AMG (host set-up, FP32 on the device). */
enum { PLX_GF_PC_JACOBI = 0, PLX_GF_PC_AMG = 1 };
int plx_gf_set_preconditioner(plx_gf gf, int kind);
/* Time probe iterations for k in {8, 16, 32, 64}; keeps the fastest per impulse. */
int plx_gf_autotune(plx_gf gf, int *best_block);
/* Microseconds per impulse per FP32 iteration measured by the last autotune, for k = 8..64. */
int plx_gf_probe_times(plx_gf gf, double *us_per_rhs_iteration);
/* Y = A X for a block of k columns ([dof][k] layout, displacement dofs; synchronous). */
int plx_gf_apply(plx_gf gf, int k, int scalar_type, const void *X, void *Y);
int plx_gf_solve(plx_gf gf, int num_impulses, const double *slip, double *u, double rtol, int maxit,
                 plx_gf_stats *stats);

/* Continuous batching on one handle (the serving loop): k slots (k in {8, 16, 32, 64}) solved as
 * independent columns of one block.  A slot is loaded with an impulse while the others keep
 * iterating; each step replays one chunk of the FP32 block PCG for all columns, and the slots
 * whose inner solve converged take an FP64 refinement step: they retire CONVERGED when
 * ||b - A x|| <= rtol ||b|| (FP64), FAILED after `maxit` refinement steps (or a non-finite
 * residual), or restart their inner PCG.  A slot's iterates do not depend on the other columns.
 *   plx_gf_slots_open: (re)initialise k empty slots;
 *   plx_gf_slots_load: load n empty slots; slip [n][num_lag][3] (device, fault frame);
 *   plx_gf_slots_step: one step; state[k] (host, may be NULL) receives the slot states;
 *   plx_gf_slot_take:  the displacement of a finished slot into u (device local vector, may be
 *                      NULL) and its statistics; the slot becomes EMPTY.
 * All synchronise the handle's stream. */
enum { PLX_SLOT_EMPTY = 0, PLX_SLOT_RUNNING = 1, PLX_SLOT_CONVERGED = 2, PLX_SLOT_FAILED = 3 };
typedef struct {
    int converged;
    int outer_iterations;   /* FP64 refinement steps */
    int inner_iterations;   /* FP32 PCG iterations */
    double relative_residual;
} plx_gf_slot_stats;
int plx_gf_slots_open(plx_gf gf, int k, double rtol, int maxit);
int plx_gf_slots_load(plx_gf gf, int n, const int *slot_ids, const double *slip);
int plx_gf_slots_step(plx_gf gf, int *state);
int plx_gf_slot_take(plx_gf gf, int slot, double *u, plx_gf_slot_stats *stats);

/* ---------------------------------------------------------------------------------------------
 * Event timers (the device counterpart of pylith::petsc::EventLogger, petsc/EventLogger.icc:23-47):
 * when on, every kernel launch of the handle is bracketed by hipEvents on its stream.
 * Events: the kernels "plx:apply_tiles", "plx:reduce_halo", and the reference's integrator
 * events around whole entry points (IntegratorDomain.cc:77-85): "PL:IntegratorDomain:
 * computeLHSResidual" / "computeRHSResidual" (plx_compute_residual by part),
 * "PL:IntegratorDomain:computeLHSJacobian" (plx_compute_jacobian) and
 * "PL:IntegratorDomain:updateStateVars" (plx_update_state_vars).  get_event_stats synchronises
 * the stream. */
int plx_dm_set_event_logging(plx_dm dm, int on);
int plx_dm_get_event_stats(plx_dm dm, const char *name, int *count, double *total_ms);
int plx_dm_reset_event_stats(plx_dm dm);

#ifdef __cplusplus
}
#endif
#endif /* PYLITH_AMD_H */
