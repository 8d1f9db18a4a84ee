/*
 * oracle.c -- CPU restatement (plain C, FP64) of PyLith's FE residual / Jacobian hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker or the timed CPU
 * baseline.  The product library (pylith_amd/csrc) never links or calls it.
 *
 * Every pointwise kernel keeps PETSc's pointwise signature and is called through a
 * function pointer at every quadrature point, the way PETSc's PetscFEIntegrateResidual
 * (PETSCFEBASIC, topology/FieldOps.cc:92-93) calls PyLith's fekernels.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------ */
/* Pointwise kernels                                                                           */
/* ------------------------------------------------------------------------------------------ */

/* Infinitesimal strain (xx,yy,zz,xy,yz,xz) from displacement gradient, Elasticity.hh:559-575. */
static void strain_from_grad(const double *g, double e[6]) {
    e[0] = g[0];
    e[1] = g[4];
    e[2] = g[8];
    e[3] = 0.5 * (g[1] + g[3]);
    e[4] = 0.5 * (g[5] + g[7]);
    e[5] = 0.5 * (g[2] + g[6]);
}

/* Symmetric 6-vector -> 3x3 row-major tensor, TensorOps::_toTensor3D (Tensor.hh:235-247). */
static void sym_to_tensor(const double v[6], double T[9]) {
    T[0] = v[0]; T[1] = v[3]; T[2] = v[5];
    T[3] = v[3]; T[4] = v[1]; T[5] = v[4];
    T[6] = v[5]; T[7] = v[4]; T[8] = v[2];
}

/* Deviatoric part, Elasticity::deviatoric (Elasticity.hh:397-408). */
static void deviatoric(const double e[6], double d[6]) {
    const double mean = (e[0] + e[1] + e[2]) / 3.0;
    d[0] = e[0] - mean; d[1] = e[1] - mean; d[2] = e[2] - mean;
    d[3] = e[3]; d[4] = e[4]; d[5] = e[5];
}

/* IsotropicLinearElasticity3D::f1v_infinitesimalStrain (IsotropicLinearElasticity.hh:803-838)
 * -> Elasticity::f1v (Elasticity.hh:171-191) with cauchyStress = meanStress + deviatoricStress
 * (IsotropicLinearElasticity.hh:182-193,221-255).  Aux: [..., shear_modulus, bulk_modulus]. */
static void f1v_isolinear_elasticity_3d(int dim, int numS, int numA, const int sOff[],
                                        const int sOff_x[], const double s[], const double s_t[],
                                        const double s_x[], const int aOff[], const int aOff_x[],
                                        const double a[], const double a_t[], const double a_x[],
                                        double t, const double x[], int numConstants,
                                        const double constants[], double f1[]) {
    (void)dim; (void)numS; (void)sOff; (void)s; (void)s_t; (void)aOff_x; (void)a_t; (void)a_x;
    (void)t; (void)x; (void)numConstants; (void)constants;
    const double mu = a[aOff[numA - 2]];
    const double K = a[aOff[numA - 1]];
    double e[6];
    strain_from_grad(&s_x[sOff_x[0]], e);
    const double tr = e[0] + e[1] + e[2];
    double sig[6] = {0, 0, 0, 0, 0, 0};
    const double mean = K * tr;
    sig[0] += mean; sig[1] += mean; sig[2] += mean;
    const double traceTerm = -2.0 / 3.0 * mu * tr;
    sig[0] += 2.0 * mu * e[0] + traceTerm;
    sig[1] += 2.0 * mu * e[1] + traceTerm;
    sig[2] += 2.0 * mu * e[2] + traceTerm;
    sig[3] += 2.0 * mu * e[3];
    sig[4] += 2.0 * mu * e[4];
    sig[5] += 2.0 * mu * e[5];
    double T[9];
    sym_to_tensor(sig, T);
    for (int i = 0; i < 9; ++i) f1[i] -= T[i];
}

/* IsotropicLinearElasticity3D::f1v_infinitesimalStrain_refState (IsotropicLinearElasticity.hh:848-880)
 * with cauchyStress_refState = meanStress_refState + deviatoricStress_refState (:204-216, :261-304).
 * Aux: [..., reference_stress(6), reference_strain(6), shear_modulus, bulk_modulus], vectors in the
 * order xx, yy, zz, xy, yz, xz (setContext_refState :131-169, TensorOps::_fromVector3D,
 * Tensor.hh:197-207). */
static void f1v_isolinear_elasticity_refstate_3d(int dim, int numS, int numA, const int sOff[],
                                                 const int sOff_x[], const double s[],
                                                 const double s_t[], const double s_x[],
                                                 const int aOff[], const int aOff_x[],
                                                 const double a[], const double a_t[],
                                                 const double a_x[], double t, const double x[],
                                                 int numConstants, const double constants[],
                                                 double f1[]) {
    (void)dim; (void)numS; (void)sOff; (void)s; (void)s_t; (void)aOff_x; (void)a_t; (void)a_x;
    (void)t; (void)x; (void)numConstants; (void)constants;
    const double *refStress = &a[aOff[numA - 4]];
    const double *refStrain = &a[aOff[numA - 3]];
    const double mu = a[aOff[numA - 2]];
    const double K = a[aOff[numA - 1]];
    double e[6];
    strain_from_grad(&s_x[sOff_x[0]], e);
    const double tr = e[0] + e[1] + e[2];
    const double trRef = refStrain[0] + refStrain[1] + refStrain[2];
    const double meanRef = (refStress[0] + refStress[1] + refStress[2]) / 3.0;
    double sig[6] = {0, 0, 0, 0, 0, 0};
    const double mean = meanRef + K * (tr - trRef);
    sig[0] += mean; sig[1] += mean; sig[2] += mean;
    const double traceTerm = -2.0 / 3.0 * mu * (tr - trRef);
    for (int i = 0; i < 3; ++i) sig[i] += refStress[i] - meanRef + 2.0 * mu * (e[i] - refStrain[i]) + traceTerm;
    for (int i = 3; i < 6; ++i) sig[i] += refStress[i] + 2.0 * mu * (e[i] - refStrain[i]);
    double T[9];
    sym_to_tensor(sig, T);
    for (int i = 0; i < 9; ++i) f1[i] -= T[i];
}

/* Fill the 21 nonzeros of -C (isotropic) at index ((f*3+g)*3+df)*3+dg,
 * IsotropicLinearElasticity.hh:1014-1034 / IsotropicLinearGenMaxwell.hh:1284-1304. */
static void jf3_isotropic(double C1111, double C1122, double C1212, double J[]) {
    J[0] -= C1111;  J[4] -= C1212;  J[8] -= C1212;  J[10] -= C1122; J[12] -= C1212;
    J[20] -= C1122; J[24] -= C1212; J[28] -= C1212; J[30] -= C1122; J[36] -= C1212;
    J[40] -= C1111; J[44] -= C1212; J[50] -= C1122; J[52] -= C1212; J[56] -= C1212;
    J[60] -= C1122; J[68] -= C1212; J[70] -= C1122; J[72] -= C1212; J[76] -= C1212;
    J[80] -= C1111;
}

/* IsotropicLinearElasticity3D::Jf3vu_infinitesimalStrain (IsotropicLinearElasticity.hh:890-1035). */
static void jf3vu_isolinear_elasticity_3d(int dim, int numS, int numA, const int sOff[],
                                          const int sOff_x[], const double s[], const double s_t[],
                                          const double s_x[], const int aOff[], const int aOff_x[],
                                          const double a[], const double a_t[], const double a_x[],
                                          double t, double s_tshift, const double x[],
                                          int numConstants, const double constants[], double J[]) {
    (void)dim; (void)numS; (void)sOff; (void)sOff_x; (void)s; (void)s_t; (void)s_x; (void)aOff_x;
    (void)a_t; (void)a_x; (void)t; (void)s_tshift; (void)x; (void)numConstants; (void)constants;
    const double mu = a[aOff[numA - 2]];
    const double K = a[aOff[numA - 1]];
    const double lambda = K - 2.0 / 3.0 * mu;
    const double lambda2mu = lambda + 2.0 * mu;
    jf3_isotropic(lambda2mu, lambda, mu, J);
}

/* Test hook: a known mistake planted in the viscous-strain update, so that the tests can show
 * their pins catch it (0: none; 1: the coefficient without its tau/dt factor; 2: dev e in place
 * of dev e - dev e_t).  Only tests set it, and they reset it to 0. */
static int g_plant_maxwell = 0;
void oracle_plant_maxwell_error(int kind) { g_plant_maxwell = kind; }

/* IsotropicLinearMaxwell::viscousStrainCoeff (IsotropicLinearMaxwell.hh:220-223). */
static double viscous_strain_coeff(double dt, double tau) {
    if (g_plant_maxwell == 1) return 1.0 - exp(-dt / tau);
    return tau * (1.0 - exp(-dt / tau)) / dt;
}

/* IsotropicLinearMaxwell::viscousStrain (IsotropicLinearMaxwell.hh:256-278). */
static void maxwell_viscous_strain(double tau, const double prev[6], const double total[6],
                                   const double e[6], double dt, double out[6]) {
    double de[6], dtot[6];
    deviatoric(e, de);
    deviatoric(total, dtot);
    if (g_plant_maxwell == 2)
        for (int i = 0; i < 6; ++i) dtot[i] = 0.0;
    const double dq = viscous_strain_coeff(dt, tau);
    const double expFac = exp(-dt / tau);
    for (int i = 0; i < 6; ++i) out[i] = expFac * prev[i] + dq * (de[i] - dtot[i]);
}

/* GenMaxwell aux layout, IsotropicLinearGenMaxwell::setContext (IsotropicLinearGenMaxwell.hh:109-165):
 * [..., shear_modulus, bulk_modulus, maxwell_time(3), shear_modulus_ratio(3),
 *  viscous_strain(3x6), total_strain(6)], constants[0] = dt. */
static void f1v_genmaxwell_3d(int dim, int numS, int numA, const int sOff[], const int sOff_x[],
                              const double s[], const double s_t[], const double s_x[],
                              const int aOff[], const int aOff_x[], const double a[],
                              const double a_t[], const double a_x[], double t, const double x[],
                              int numConstants, const double constants[], double f1[]) {
    (void)dim; (void)numS; (void)sOff; (void)s; (void)s_t; (void)aOff_x; (void)a_t; (void)a_x;
    (void)t; (void)x; (void)numConstants;
    const double mu = a[aOff[numA - 6]];
    const double K = a[aOff[numA - 5]];
    const double *tau = &a[aOff[numA - 4]];
    const double *ratio = &a[aOff[numA - 3]];
    const double *visc = &a[aOff[numA - 2]];
    const double *total = &a[aOff[numA - 1]];
    const double dt = constants[0];
    double e[6];
    strain_from_grad(&s_x[sOff_x[0]], e);
    double sig[6] = {0, 0, 0, 0, 0, 0};
    /* meanStress (IsotropicLinearElasticity.hh:205-215) */
    const double mean = K * (e[0] + e[1] + e[2]);
    sig[0] += mean; sig[1] += mean; sig[2] += mean;
    /* cauchyStress (IsotropicLinearGenMaxwell.hh:262-286) */
    double vs[3][6];
    for (int i = 0; i < 3; ++i) maxwell_viscous_strain(tau[i], &visc[6 * i], total, e, dt, vs[i]);
    double de[6];
    deviatoric(e, de);
    /* _deviatoricStress (IsotropicLinearGenMaxwell.hh:404-450) */
    double ratio0 = 1.0;
    for (int i = 0; i < 3; ++i) ratio0 -= ratio[i];
    const double muE = mu * ratio0;
    for (int k = 0; k < 6; ++k) {
        sig[k] += 2.0 * muE * de[k];
        for (int i = 0; i < 3; ++i) sig[k] += 2.0 * mu * ratio[i] * vs[i][k];
    }
    double T[9];
    sym_to_tensor(sig, T);
    for (int i = 0; i < 9; ++i) f1[i] -= T[i];
}

/* IsotropicLinearGenMaxwell3D::Jf3vu_infinitesimalStrain (IsotropicLinearGenMaxwell.hh:1143-1310). */
static void jf3vu_genmaxwell_3d(int dim, int numS, int numA, const int sOff[], const int sOff_x[],
                                const double s[], const double s_t[], const double s_x[],
                                const int aOff[], const int aOff_x[], const double a[],
                                const double a_t[], const double a_x[], double t, double s_tshift,
                                const double x[], int numConstants, const double constants[],
                                double J[]) {
    (void)dim; (void)numS; (void)sOff; (void)sOff_x; (void)s; (void)s_t; (void)s_x; (void)aOff_x;
    (void)a_t; (void)a_x; (void)t; (void)s_tshift; (void)x; (void)numConstants;
    const double mu = a[aOff[numA - 6]];
    const double K = a[aOff[numA - 5]];
    const double *tau = &a[aOff[numA - 4]];
    const double *ratio = &a[aOff[numA - 3]];
    const double dt = constants[0];
    const double ratio0 = 1.0 - ratio[0] - ratio[1] - ratio[2];
    double sf = mu * ratio0;
    for (int i = 0; i < 3; ++i) sf += mu * viscous_strain_coeff(dt, tau[i]) * ratio[i];
    jf3_isotropic(K + 4.0 * sf / 3.0, K - 2.0 * sf / 3.0, sf, J);
}

/* IsotropicLinearGenMaxwell3D::viscousStrain_infinitesimalStrain_asVector
 * (IsotropicLinearGenMaxwell.hh:1507-1535): 3 x 6 new viscous strains. */
static void viscous_strain_genmaxwell_3d(int dim, int numS, int numA, const int sOff[],
                                         const int sOff_x[], const double s[], const double s_t[],
                                         const double s_x[], const int aOff[], const int aOff_x[],
                                         const double a[], const double a_t[], const double a_x[],
                                         double t, const double x[], int numConstants,
                                         const double constants[], double out[]) {
    (void)dim; (void)numS; (void)sOff; (void)s; (void)s_t; (void)aOff_x; (void)a_t; (void)a_x;
    (void)t; (void)x; (void)numConstants;
    const double *tau = &a[aOff[numA - 4]];
    const double *visc = &a[aOff[numA - 2]];
    const double *total = &a[aOff[numA - 1]];
    const double dt = constants[0];
    double e[6];
    strain_from_grad(&s_x[sOff_x[0]], e);
    for (int i = 0; i < 3; ++i) maxwell_viscous_strain(tau[i], &visc[6 * i], total, e, dt, &out[6 * i]);
}

/* Elasticity3D::infinitesimalStrain_asVector (Elasticity.hh:586-610). */
static void total_strain_3d(int dim, int numS, int numA, const int sOff[], const int sOff_x[],
                            const double s[], const double s_t[], const double s_x[],
                            const int aOff[], const int aOff_x[], const double a[],
                            const double a_t[], const double a_x[], double t, const double x[],
                            int numConstants, const double constants[], double out[]) {
    (void)dim; (void)numS; (void)numA; (void)sOff; (void)s; (void)s_t; (void)aOff; (void)aOff_x;
    (void)a; (void)a_t; (void)a_x; (void)t; (void)x; (void)numConstants; (void)constants;
    strain_from_grad(&s_x[sOff_x[0]], out);
}

/* Body force and gravity, the f0 of the quasi-static displacement equation (Elasticity.hh:249-385,
 * chosen in materials/Elasticity.cc:455-475).  Aux subfields: 0 density, then body_force and/or
 * gravitational_acceleration (3 each); f0 += density * g (g0v_grav, Elasticity.hh:249-284),
 * f0 += b (g0v_bodyforce, :300-329), or both, gravity first (g0v_gravbodyforce, :346-373). */
static void f0v_grav(int dim, int numS, int numA, const int sOff[], const int sOff_x[],
                     const double s[], const double s_t[], const double s_x[], const int aOff[],
                     const int aOff_x[], const double a[], const double a_t[], const double a_x[],
                     double t, const double x[], int numConstants, const double constants[],
                     double f0[]) {
    (void)numS; (void)numA; (void)sOff; (void)sOff_x; (void)s; (void)s_t; (void)s_x; (void)aOff_x;
    (void)a_t; (void)a_x; (void)t; (void)x; (void)numConstants; (void)constants;
    const double rho = a[aOff[0]];
    const double *g = &a[aOff[1]];
    for (int i = 0; i < dim; ++i) f0[i] += rho * g[i];
}

static void f0v_bodyforce(int dim, int numS, int numA, const int sOff[], const int sOff_x[],
                          const double s[], const double s_t[], const double s_x[], const int aOff[],
                          const int aOff_x[], const double a[], const double a_t[], const double a_x[],
                          double t, const double x[], int numConstants, const double constants[],
                          double f0[]) {
    (void)numS; (void)numA; (void)sOff; (void)sOff_x; (void)s; (void)s_t; (void)s_x; (void)aOff_x;
    (void)a_t; (void)a_x; (void)t; (void)x; (void)numConstants; (void)constants;
    const double *b = &a[aOff[1]];
    for (int i = 0; i < dim; ++i) f0[i] += b[i];
}

static void f0v_gravbodyforce(int dim, int numS, int numA, const int sOff[], const int sOff_x[],
                              const double s[], const double s_t[], const double s_x[],
                              const int aOff[], const int aOff_x[], const double a[],
                              const double a_t[], const double a_x[], double t, const double x[],
                              int numConstants, const double constants[], double f0[]) {
    const int offGrav[2] = {aOff[0], aOff[2]};
    const int offBody[2] = {aOff[0], aOff[1]};
    f0v_grav(dim, numS, numA, sOff, sOff_x, s, s_t, s_x, offGrav, aOff_x, a, a_t, a_x, t, x,
             numConstants, constants, f0);
    f0v_bodyforce(dim, numS, numA, sOff, sOff_x, s, s_t, s_x, offBody, aOff_x, a, a_t, a_x, t, x,
                  numConstants, constants, f0);
}

/* Test-only pointwise kernels (no counterpart in the reference), for the device path's
 * PetscPointFn-form plugin launchers: they read every input a kernel can see.  Aux [density,
 * shear_modulus, bulk_modulus], constants [c0, c1, c2]:
 *   f0 = c1 s + density (c0 x + t);  f1 = -sigma(mu(x), K), mu(x) = mu (1 + c2 sin(x/1000 + z/2000));
 *   Jf0 = c1 I;  Jf3 = -C(mu(x), K). */
static double test_shear(const double *a, const int aOff[], int numA, const double x[], const double constants[]) {
    return a[aOff[numA - 2]] * (1.0 + constants[2] * sin(x[0] / 1000.0 + x[2] / 2000.0));
}

static void f0_test(int dim, int numS, int numA, const int sOff[], const int sOff_x[], const double s[],
                    const double s_t[], const double s_x[], const int aOff[], const int aOff_x[],
                    const double a[], const double a_t[], const double a_x[], double t, const double x[],
                    int numConstants, const double constants[], double f0[]) {
    (void)numS; (void)numA; (void)sOff_x; (void)s_x; (void)aOff_x; (void)a_t; (void)a_x;
    (void)numConstants;
    for (int i = 0; i < dim; ++i)
        f0[i] += constants[1] * s[sOff[0] + i] + a[aOff[0]] * (constants[0] * x[i] + t + s_t[sOff[0] + i]);
}

static void f1_test(int dim, int numS, int numA, const int sOff[], const int sOff_x[], const double s[],
                    const double s_t[], const double s_x[], const int aOff[], const int aOff_x[],
                    const double a[], const double a_t[], const double a_x[], double t, const double x[],
                    int numConstants, const double constants[], double f1[]) {
    (void)dim; (void)numS; (void)sOff; (void)s; (void)s_t; (void)aOff_x; (void)a_t; (void)a_x; (void)t;
    (void)numConstants;
    const double mu = test_shear(a, aOff, numA, x, constants), K = a[aOff[numA - 1]];
    double e[6], sig[6], T[9];
    strain_from_grad(&s_x[sOff_x[0]], e);
    const double tr = e[0] + e[1] + e[2];
    for (int i = 0; i < 3; ++i) sig[i] = K * tr + 2.0 * mu * (e[i] - tr / 3.0);
    for (int i = 3; i < 6; ++i) sig[i] = 2.0 * mu * e[i];
    sym_to_tensor(sig, T);
    for (int i = 0; i < 9; ++i) f1[i] -= T[i];
}

static void jf0_test(int dim, int numS, int numA, const int sOff[], const int sOff_x[], const double s[],
                     const double s_t[], const double s_x[], const int aOff[], const int aOff_x[],
                     const double a[], const double a_t[], const double a_x[], double t, double s_tshift,
                     const double x[], int numConstants, const double constants[], double J[]) {
    (void)numS; (void)numA; (void)sOff; (void)sOff_x; (void)s; (void)s_t; (void)s_x; (void)aOff_x;
    (void)a_t; (void)a_x; (void)t; (void)x; (void)numConstants;
    for (int i = 0; i < dim; ++i) J[i * dim + i] += constants[1] + s_tshift * a[aOff[0]];
}

static void jf3_test(int dim, int numS, int numA, const int sOff[], const int sOff_x[], const double s[],
                     const double s_t[], const double s_x[], const int aOff[], const int aOff_x[],
                     const double a[], const double a_t[], const double a_x[], double t, double s_tshift,
                     const double x[], int numConstants, const double constants[], double J[]) {
    (void)dim; (void)numS; (void)sOff; (void)sOff_x; (void)s; (void)s_t; (void)s_x; (void)aOff_x; (void)a_t;
    (void)a_x; (void)t; (void)s_tshift; (void)numConstants;
    const double mu = test_shear(a, aOff, numA, x, constants), K = a[aOff[numA - 1]];
    const double lambda = K - 2.0 / 3.0 * mu;
    jf3_isotropic(lambda + 2.0 * mu, lambda, mu, J);
}

oracle_point_fn oracle_point_kernel(int id) {
    switch (id) {
    case ORACLE_TEST_F0: return f0_test;
    case ORACLE_TEST_F1: return f1_test;
    case ORACLE_F1V_ISOLINEAR_ELASTICITY_REFSTATE_3D: return f1v_isolinear_elasticity_refstate_3d;
    case ORACLE_F0V_BODYFORCE: return f0v_bodyforce;
    case ORACLE_F0V_GRAV: return f0v_grav;
    case ORACLE_F0V_GRAVBODYFORCE: return f0v_gravbodyforce;
    case ORACLE_F1V_ISOLINEAR_ELASTICITY_3D: return f1v_isolinear_elasticity_3d;
    case ORACLE_F1V_GENMAXWELL_3D: return f1v_genmaxwell_3d;
    case ORACLE_VISCOUS_STRAIN_GENMAXWELL_3D: return viscous_strain_genmaxwell_3d;
    case ORACLE_TOTAL_STRAIN_3D: return total_strain_3d;
    default: return NULL;
    }
}

oracle_point_jac_fn oracle_point_jac_kernel(int id) {
    switch (id) {
    case ORACLE_TEST_JF0: return jf0_test;
    case ORACLE_TEST_JF3: return jf3_test;
    case ORACLE_JF3VU_ISOLINEAR_ELASTICITY_3D: return jf3vu_isolinear_elasticity_3d;
    case ORACLE_JF3VU_GENMAXWELL_3D: return jf3vu_genmaxwell_3d;
    default: return NULL;
    }
}

/* Fault kernels, FaultCohesiveKin.hh:73-224,298-417 (PetscBdPointFn signature). */
static void f0u_neg(int dim, int numS, int numA, const int sOff[], const int sOff_x[],
                    const double s[], const double s_t[], const double s_x[], const int aOff[],
                    const int aOff_x[], const double a[], const double a_t[], const double a_x[],
                    double t, const double x[], const double n[], int numConstants,
                    const double constants[], double f0[]) {
    (void)numA; (void)sOff_x; (void)s_t; (void)s_x; (void)aOff; (void)aOff_x; (void)a; (void)a_t;
    (void)a_x; (void)t; (void)x; (void)n; (void)numConstants; (void)constants;
    const double *lagrange = &s[sOff[numS - 1]];
    for (int i = 0; i < dim; ++i) f0[i] += -lagrange[i];
}

static void f0u_pos(int dim, int numS, int numA, const int sOff[], const int sOff_x[],
                    const double s[], const double s_t[], const double s_x[], const int aOff[],
                    const int aOff_x[], const double a[], const double a_t[], const double a_x[],
                    double t, const double x[], const double n[], int numConstants,
                    const double constants[], double f0[]) {
    (void)numA; (void)sOff_x; (void)s_t; (void)s_x; (void)aOff; (void)aOff_x; (void)a; (void)a_t;
    (void)a_x; (void)t; (void)x; (void)n; (void)numConstants; (void)constants;
    const double *lagrange = &s[sOff[numS - 1]];
    for (int i = 0; i < dim; ++i) f0[i] += +lagrange[i];
}

static void unit_cross(double r[3], const double u[3], const double v[3]) {
    r[0] = +u[1] * v[2] - u[2] * v[1];
    r[1] = +u[2] * v[0] - u[0] * v[2];
    r[2] = +u[0] * v[1] - u[1] * v[0];
    const double mag = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (int i = 0; i < 3; ++i) r[i] /= mag;
}

/* BoundaryDirections::tangential_directions (BoundaryDirections.hh:50-82).  The sign test
 * at :76 reads refDir2[1]*n[1] twice and never refDir2[2]*n[2]; restated as written. */
void oracle_tangential_directions(double t1[3], double t2[3], const double r1[3],
                                  const double r2[3], const double n[3]) {
    double r[3] = {r1[0], r1[1], r1[2]};
    if (fabs(r[0] * n[0] + r[1] * n[1] + r[2] * n[2]) > 0.999) {
        r[0] = r2[0]; r[1] = r2[1]; r[2] = r2[2];
    }
    unit_cross(t1, r, n);
    unit_cross(t2, n, t1);
    if (r2[0] * n[0] + r2[1] * n[1] + r2[1] * n[1] < 0.0) {
        for (int i = 0; i < 3; ++i) t2[i] *= -1.0;
    }
}

static void f0l_slip(int dim, int numS, int numA, const int sOff[], const int sOff_x[],
                     const double s[], const double s_t[], const double s_x[], const int aOff[],
                     const int aOff_x[], const double a[], const double a_t[], const double a_x[],
                     double t, const double x[], const double n[], int numConstants,
                     const double constants[], double f0[]) {
    (void)numS; (void)numA; (void)sOff_x; (void)s_t; (void)s_x; (void)aOff_x; (void)a_t;
    (void)a_x; (void)t; (void)x; (void)numConstants;
    const double *slip = &a[aOff[0]];
    const double *dispN = &s[sOff[0]];
    const double *dispP = &s[sOff[0] + dim];
    double t1[3], t2[3];
    oracle_tangential_directions(t1, t2, &constants[0], &constants[3], n);
    for (int i = 0; i < 3; ++i) {
        const double slipXYZ = n[i] * slip[0] + t1[i] * slip[1] + t2[i] * slip[2];
        f0[i] += dispP[i] - dispN[i] - slipXYZ;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Reference elements                                                                           */
/* ------------------------------------------------------------------------------------------ */

/* Q1 hexahedron on [-1,1]^3 with the vertex order of PyLith mesh files (tests/libtests/
 * materials/data/hex.mesh): 0-3 one face in cyclic order, 4-7 the opposite face, i+4 over i. */
static const double hex_ref[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                     {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

/* P2 tetrahedron on (0,0,0),(1,0,0),(0,1,0),(0,0,1): 4 vertices, then edge nodes. */
static const int tet_edges[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

/* 14-point, degree-5 symmetric rule for the tetrahedron (two S31 orbits, one S22 orbit;
 * weights sum to the reference volume 1/6).  Parameters solved to 25 digits; exactness for
 * every monomial of degree <= 5 is checked by tests/test_oracle_quadrature.py. */
static const double TQ_A1 = 0.09273525031089125323737967, TQ_W1 = 0.01224884051939366598892406;
static const double TQ_A2 = 0.3108859192633006188837448, TQ_W2 = 0.0187813209530026585737831;
static const double TQ_B = 0.4544962958743504682934192, TQ_W3 = 0.007091003462846894735973002;

int oracle_nodes_per_cell(int ct) { return ct == ORACLE_HEX_Q1 ? 8 : ct == ORACLE_TET_P2 ? 10 : 0; }
int oracle_verts_per_cell(int ct) { return ct == ORACLE_HEX_Q1 ? 8 : ct == ORACLE_TET_P2 ? 4 : 0; }
int oracle_num_quad(int ct) { return ct == ORACLE_HEX_Q1 ? 8 : ct == ORACLE_TET_P2 ? 14 : 0; }

int oracle_quadrature(int ct, double *p, double *w) {
    if (ct == ORACLE_HEX_Q1) {
        /* PetscDTCreateDefaultQuadrature(HEX, 1): 2-point Gauss-Legendre per direction
         * (topology/FieldOps.cc:113). */
        const double g = 1.0 / sqrt(3.0);
        int q = 0;
        for (int k = 0; k < 2; ++k)
            for (int j = 0; j < 2; ++j)
                for (int i = 0; i < 2; ++i, ++q) {
                    p[3 * q + 0] = i ? g : -g;
                    p[3 * q + 1] = j ? g : -g;
                    p[3 * q + 2] = k ? g : -g;
                    w[q] = 1.0;
                }
        return 0;
    }
    if (ct == ORACLE_TET_P2) {
        int q = 0;
        const double A[2] = {TQ_A1, TQ_A2}, W[2] = {TQ_W1, TQ_W2};
        for (int o = 0; o < 2; ++o)
            for (int i = 0; i < 4; ++i, ++q) {
                double l[4] = {A[o], A[o], A[o], A[o]};
                l[i] = 1.0 - 3.0 * A[o];
                p[3 * q + 0] = l[1]; p[3 * q + 1] = l[2]; p[3 * q + 2] = l[3];
                w[q] = W[o];
            }
        for (int e = 0; e < 6; ++e, ++q) {
            double l[4];
            for (int i = 0; i < 4; ++i) l[i] = 0.5 - TQ_B;
            l[tet_edges[e][0]] = TQ_B;
            l[tet_edges[e][1]] = TQ_B;
            p[3 * q + 0] = l[1]; p[3 * q + 1] = l[2]; p[3 * q + 2] = l[3];
            w[q] = TQ_W3;
        }
        return 0;
    }
    return 1;
}

int oracle_tabulate(int ct, const double xi[3], double *phi, double *dphi) {
    if (ct == ORACLE_HEX_Q1) {
        for (int i = 0; i < 8; ++i) {
            const double a = 1.0 + hex_ref[i][0] * xi[0];
            const double b = 1.0 + hex_ref[i][1] * xi[1];
            const double c = 1.0 + hex_ref[i][2] * xi[2];
            phi[i] = 0.125 * a * b * c;
            dphi[3 * i + 0] = 0.125 * hex_ref[i][0] * b * c;
            dphi[3 * i + 1] = 0.125 * a * hex_ref[i][1] * c;
            dphi[3 * i + 2] = 0.125 * a * b * hex_ref[i][2];
        }
        return 0;
    }
    if (ct == ORACLE_TET_P2) {
        const double l[4] = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
        const double dl[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int v = 0; v < 4; ++v) {
            phi[v] = l[v] * (2.0 * l[v] - 1.0);
            for (int d = 0; d < 3; ++d) dphi[3 * v + d] = (4.0 * l[v] - 1.0) * dl[v][d];
        }
        for (int e = 0; e < 6; ++e) {
            const int a = tet_edges[e][0], b = tet_edges[e][1];
            phi[4 + e] = 4.0 * l[a] * l[b];
            for (int d = 0; d < 3; ++d) dphi[3 * (4 + e) + d] = 4.0 * (l[a] * dl[b][d] + l[b] * dl[a][d]);
        }
        return 0;
    }
    return 1;
}

/* Coordinate (geometry) basis: Q1 for the hexahedron, P1 for the tetrahedron. */
static void tabulate_coord(int ct, const double xi[3], double *psi, double *dpsi) {
    if (ct == ORACLE_HEX_Q1) {
        oracle_tabulate(ct, xi, psi, dpsi);
        return;
    }
    const double dl[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    psi[0] = 1.0 - xi[0] - xi[1] - xi[2];
    psi[1] = xi[0]; psi[2] = xi[1]; psi[3] = xi[2];
    for (int v = 0; v < 4; ++v)
        for (int d = 0; d < 3; ++d) dpsi[3 * v + d] = dl[v][d];
}

/* ------------------------------------------------------------------------------------------ */
/* Cell loop                                                                                    */
/* ------------------------------------------------------------------------------------------ */

#define MAXNB 10
#define MAXQ 14
#define MAXAUX 64

typedef struct {
    int ct, nb, nv, nq;
    double qp[MAXQ][3], qw[MAXQ];
    double phi[MAXQ][MAXNB], dphi[MAXQ][MAXNB][3];
    double psi[MAXQ][8], dpsi[MAXQ][8][3];
} tab_t;

static int make_tab(int ct, tab_t *T) {
    T->ct = ct;
    T->nb = oracle_nodes_per_cell(ct);
    T->nv = oracle_verts_per_cell(ct);
    T->nq = oracle_num_quad(ct);
    if (!T->nb) return 1;
    oracle_quadrature(ct, &T->qp[0][0], T->qw);
    for (int q = 0; q < T->nq; ++q) {
        oracle_tabulate(ct, T->qp[q], T->phi[q], &T->dphi[q][0][0]);
        tabulate_coord(ct, T->qp[q], T->psi[q], &T->dpsi[q][0][0]);
    }
    return 0;
}

/* Cell geometry at quadrature point q (DMPlexComputeCellGeometryFEM semantics). */
static int geometry(const tab_t *T, const double X[][3], int q, double x[3], double Jinv[9],
                    double *detJ) {
    double J[9] = {0};
    x[0] = x[1] = x[2] = 0.0;
    for (int v = 0; v < T->nv; ++v) {
        for (int d = 0; d < 3; ++d) {
            x[d] += X[v][d] * T->psi[q][v];
            for (int k = 0; k < 3; ++k) J[3 * d + k] += X[v][d] * T->dpsi[q][v][k];
        }
    }
    const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                       J[2] * (J[3] * J[7] - J[4] * J[6]);
    if (!(det > 0.0)) return 2;
    const double id = 1.0 / det;
    Jinv[0] = (J[4] * J[8] - J[5] * J[7]) * id;
    Jinv[1] = (J[2] * J[7] - J[1] * J[8]) * id;
    Jinv[2] = (J[1] * J[5] - J[2] * J[4]) * id;
    Jinv[3] = (J[5] * J[6] - J[3] * J[8]) * id;
    Jinv[4] = (J[0] * J[8] - J[2] * J[6]) * id;
    Jinv[5] = (J[2] * J[3] - J[0] * J[5]) * id;
    Jinv[6] = (J[3] * J[7] - J[4] * J[6]) * id;
    Jinv[7] = (J[1] * J[6] - J[0] * J[7]) * id;
    Jinv[8] = (J[0] * J[4] - J[1] * J[3]) * id;
    *detJ = det;
    return 0;
}

typedef struct {
    int naux_tot;
    int aOff[16];
    int aOff_x[16];
} auxinfo_t;

static int aux_info(const oracle_mesh *m, auxinfo_t *A) {
    A->naux_tot = 0;
    if (m->num_aux > 16) return 1;
    for (int i = 0; i < m->num_aux; ++i) {
        A->aOff[i] = A->naux_tot;
        A->aOff_x[i] = 0;
        A->naux_tot += m->aux_sizes[i];
    }
    return A->naux_tot > MAXAUX ? 1 : 0;
}

static const double *aux_at(const oracle_mesh *m, const auxinfo_t *A, const tab_t *T, int cell,
                            int q) {
    if (!m->aux) return NULL;
    if (m->aux_per_qp) return m->aux + ((size_t)cell * T->nq + q) * A->naux_tot;
    return m->aux + (size_t)cell * A->naux_tot;
}

/* Per-cell gather (DMPlexVecGetClosure). */
static void gather(const oracle_mesh *m, const tab_t *T, int c, const double *vec,
                   double out[][3]) {
    const int32_t *cl = m->closure + (size_t)c * T->nb;
    for (int n = 0; n < T->nb; ++n)
        for (int k = 0; k < 3; ++k) out[n][k] = vec ? vec[3 * (size_t)cl[n] + k] : 0.0;
}

static void gather_coords(const oracle_mesh *m, const tab_t *T, int c, double X[][3]) {
    const int32_t *cn = m->cone + (size_t)c * T->nv;
    for (int v = 0; v < T->nv; ++v)
        for (int k = 0; k < 3; ++k) X[v][k] = m->coords[3 * (size_t)cn[v] + k];
}

/* Residual of one cell: re[n][c] = sum_q w|J| (f0 phi + f1 : grad phi). */
static int cell_residual(const oracle_mesh *m, const tab_t *T, const auxinfo_t *A, int c,
                         oracle_point_fn f0, oracle_point_fn f1, double t, const double *locX,
                         const double *locX_t, double re[][3]) {
    double ue[MAXNB][3], ute[MAXNB][3], X[8][3];
    gather(m, T, c, locX, ue);
    gather(m, T, c, locX_t, ute);
    gather_coords(m, T, c, X);
    memset(re, 0, sizeof(double) * 3 * T->nb);
    const int sOff[1] = {0}, sOff_x[1] = {0};
    for (int q = 0; q < T->nq; ++q) {
        double x[3], Ji[9], detJ;
        if (geometry(T, (const double(*)[3])X, q, x, Ji, &detJ)) return 2;
        double gphi[MAXNB][3];
        for (int n = 0; n < T->nb; ++n)
            for (int d = 0; d < 3; ++d)
                gphi[n][d] = T->dphi[q][n][0] * Ji[0 * 3 + d] + T->dphi[q][n][1] * Ji[1 * 3 + d] +
                             T->dphi[q][n][2] * Ji[2 * 3 + d];
        double s[3] = {0}, st[3] = {0}, sx[9] = {0};
        for (int n = 0; n < T->nb; ++n)
            for (int k = 0; k < 3; ++k) {
                s[k] += ue[n][k] * T->phi[q][n];
                st[k] += ute[n][k] * T->phi[q][n];
                for (int d = 0; d < 3; ++d) sx[3 * k + d] += ue[n][k] * gphi[n][d];
            }
        const double *a = aux_at(m, A, T, c, q);
        double f0v[3] = {0}, f1v[9] = {0};
        if (f0) f0(3, 1, m->num_aux, sOff, sOff_x, s, st, sx, A->aOff, A->aOff_x, a, NULL, NULL, t, x,
                   m->num_constants, m->constants, f0v);
        if (f1) f1(3, 1, m->num_aux, sOff, sOff_x, s, st, sx, A->aOff, A->aOff_x, a, NULL, NULL, t, x,
                   m->num_constants, m->constants, f1v);
        const double wd = T->qw[q] * detJ;
        for (int n = 0; n < T->nb; ++n)
            for (int k = 0; k < 3; ++k)
                re[n][k] += wd * (f0v[k] * T->phi[q][n] + f1v[3 * k + 0] * gphi[n][0] +
                                  f1v[3 * k + 1] * gphi[n][1] + f1v[3 * k + 2] * gphi[n][2]);
    }
    return 0;
}

static void scatter_add(const oracle_mesh *m, const tab_t *T, int c, const double re[][3],
                        double *locF) {
    const int32_t *cl = m->closure + (size_t)c * T->nb;
    for (int n = 0; n < T->nb; ++n)
        for (int k = 0; k < 3; ++k) locF[3 * (size_t)cl[n] + k] += re[n][k];
}

int oracle_compute_residual(const oracle_mesh *m, const int32_t *cells, int ncells, int f0_id,
                            int f1_id, double t, const double *locX, const double *locX_t,
                            double *locF) {
    tab_t T;
    auxinfo_t A;
    if (!m || !locF || make_tab(m->cell_type, &T) || aux_info(m, &A)) return 1;
    oracle_point_fn f0 = oracle_point_kernel(f0_id), f1 = oracle_point_kernel(f1_id);
    if ((f0_id && !f0) || (f1_id && !f1)) return 1;
    const int n = cells ? ncells : m->num_cells;
    for (int i = 0; i < n; ++i) {
        const int c = cells ? cells[i] : i;
        double re[MAXNB][3];
        const int err = cell_residual(m, &T, &A, c, f0, f1, t, locX, locX_t, re);
        if (err) return err;
        scatter_add(m, &T, c, re, locF);
    }
    return 0;
}

int oracle_compute_residual_colored(const oracle_mesh *m, const int32_t *color_ptr,
                                    const int32_t *color_cells, int ncolors, int f1_id, double t,
                                    const double *locX, double *locF, int nthreads) {
    tab_t T;
    auxinfo_t A;
    if (!m || !locF || make_tab(m->cell_type, &T) || aux_info(m, &A)) return 1;
    oracle_point_fn f1 = oracle_point_kernel(f1_id);
    if (!f1) return 1;
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    for (int col = 0; col < ncolors; ++col) {
        const int b = color_ptr[col], e = color_ptr[col + 1];
#pragma omp parallel for schedule(static) reduction(| : err)
        for (int i = b; i < e; ++i) {
            double re[MAXNB][3];
            const int c = color_cells[i];
            const int ce = cell_residual(m, &T, &A, c, NULL, f1, t, locX, NULL, re);
            if (ce) err |= ce;
            else scatter_add(m, &T, c, re, locF);
        }
    }
    return err;
}

/* Per-quadrature-point Jacobian kernels contracted with a trial field: PETSc's
 * PetscFEIntegrateJacobian with the Jf0..Jf3 layouts [f][g], [f][g][dg], [f][g][df],
 * [f][g][df][dg] (IsotropicLinearElasticity.hh:946-1012). */
int oracle_compute_jacobian_action(const oracle_mesh *m, const int32_t *cells, int ncells,
                                   int jf0_id, int jf1_id, int jf2_id, int jf3_id, double t,
                                   double s_tshift, const double *locX, const double *locX_t,
                                   const double *locY, double *locF) {
    tab_t T;
    auxinfo_t A;
    if (!m || !locF || !locY || make_tab(m->cell_type, &T) || aux_info(m, &A)) return 1;
    oracle_point_jac_fn J0 = oracle_point_jac_kernel(jf0_id), J1 = oracle_point_jac_kernel(jf1_id);
    oracle_point_jac_fn J2 = oracle_point_jac_kernel(jf2_id), J3 = oracle_point_jac_kernel(jf3_id);
    if ((jf0_id && !J0) || (jf1_id && !J1) || (jf2_id && !J2) || (jf3_id && !J3)) return 1;
    const int n = cells ? ncells : m->num_cells;
    const int sOff[1] = {0}, sOff_x[1] = {0};
    for (int i = 0; i < n; ++i) {
        const int c = cells ? cells[i] : i;
        double ue[MAXNB][3], ute[MAXNB][3], ve[MAXNB][3], X[8][3], re[MAXNB][3];
        gather(m, &T, c, locX, ue);
        gather(m, &T, c, locX_t, ute);
        gather(m, &T, c, locY, ve);
        gather_coords(m, &T, c, X);
        memset(re, 0, sizeof(re));
        for (int q = 0; q < T.nq; ++q) {
            double x[3], Ji[9], detJ;
            if (geometry(&T, (const double(*)[3])X, q, x, Ji, &detJ)) return 2;
            double gphi[MAXNB][3];
            for (int b = 0; b < T.nb; ++b)
                for (int d = 0; d < 3; ++d)
                    gphi[b][d] = T.dphi[q][b][0] * Ji[d] + T.dphi[q][b][1] * Ji[3 + d] +
                                 T.dphi[q][b][2] * Ji[6 + d];
            double s[3] = {0}, st[3] = {0}, sx[9] = {0}, v[3] = {0}, vx[9] = {0};
            for (int b = 0; b < T.nb; ++b)
                for (int k = 0; k < 3; ++k) {
                    s[k] += ue[b][k] * T.phi[q][b];
                    st[k] += ute[b][k] * T.phi[q][b];
                    v[k] += ve[b][k] * T.phi[q][b];
                    for (int d = 0; d < 3; ++d) {
                        sx[3 * k + d] += ue[b][k] * gphi[b][d];
                        vx[3 * k + d] += ve[b][k] * gphi[b][d];
                    }
                }
            const double *a = aux_at(m, &A, &T, c, q);
            double j0[9] = {0}, j1[27] = {0}, j2[27] = {0}, j3[81] = {0};
#define CALLJ(F, OUT)                                                                              \
    if (F) F(3, 1, m->num_aux, sOff, sOff_x, s, st, sx, A.aOff, A.aOff_x, a, NULL, NULL, t, s_tshift, \
             x, m->num_constants, m->constants, OUT)
            CALLJ(J0, j0);
            CALLJ(J1, j1);
            CALLJ(J2, j2);
            CALLJ(J3, j3);
#undef CALLJ
            double g0[3] = {0}, g1[9] = {0};
            for (int f = 0; f < 3; ++f)
                for (int g = 0; g < 3; ++g) {
                    g0[f] += j0[3 * f + g] * v[g];
                    for (int d = 0; d < 3; ++d) {
                        g0[f] += j1[(3 * f + g) * 3 + d] * vx[3 * g + d];
                        g1[3 * f + d] += j2[(3 * f + g) * 3 + d] * v[g];
                        for (int e = 0; e < 3; ++e)
                            g1[3 * f + d] += j3[((3 * f + g) * 3 + d) * 3 + e] * vx[3 * g + e];
                    }
                }
            const double wd = T.qw[q] * detJ;
            for (int b = 0; b < T.nb; ++b)
                for (int f = 0; f < 3; ++f)
                    re[b][f] += wd * (g0[f] * T.phi[q][b] + g1[3 * f] * gphi[b][0] +
                                      g1[3 * f + 1] * gphi[b][1] + g1[3 * f + 2] * gphi[b][2]);
        }
        scatter_add(m, &T, c, (const double(*)[3])re, locF);
    }
    return 0;
}

/* Element matrix K[(b*3+f)][(b2*3+g)] = sum_q w|J| Jf3[f][g][df][dg] dphi_b[df] dphi_b2[dg]. */
static int element_matrix(const oracle_mesh *m, const tab_t *T, const auxinfo_t *A, int c,
                          oracle_point_jac_fn J3, double t, double s_tshift, const double *locX,
                          double *Ke) {
    const int nd = 3 * T->nb;
    double ue[MAXNB][3], X[8][3];
    gather(m, T, c, locX, ue);
    gather_coords(m, T, c, X);
    memset(Ke, 0, sizeof(double) * nd * nd);
    const int sOff[1] = {0}, sOff_x[1] = {0};
    for (int q = 0; q < T->nq; ++q) {
        double x[3], Ji[9], detJ;
        if (geometry(T, (const double(*)[3])X, q, x, Ji, &detJ)) return 2;
        double gphi[MAXNB][3];
        for (int b = 0; b < T->nb; ++b)
            for (int d = 0; d < 3; ++d)
                gphi[b][d] = T->dphi[q][b][0] * Ji[d] + T->dphi[q][b][1] * Ji[3 + d] +
                             T->dphi[q][b][2] * Ji[6 + d];
        double s[3] = {0}, sx[9] = {0};
        for (int b = 0; b < T->nb; ++b)
            for (int k = 0; k < 3; ++k) {
                s[k] += ue[b][k] * T->phi[q][b];
                for (int d = 0; d < 3; ++d) sx[3 * k + d] += ue[b][k] * gphi[b][d];
            }
        const double *a = aux_at(m, A, T, c, q);
        double j3[81] = {0};
        J3(3, 1, m->num_aux, sOff, sOff_x, s, NULL, sx, A->aOff, A->aOff_x, a, NULL, NULL, t,
           s_tshift, x, m->num_constants, m->constants, j3);
        const double wd = T->qw[q] * detJ;
        for (int b = 0; b < T->nb; ++b)
            for (int f = 0; f < 3; ++f)
                for (int b2 = 0; b2 < T->nb; ++b2)
                    for (int g = 0; g < 3; ++g) {
                        double acc = 0.0;
                        for (int d = 0; d < 3; ++d)
                            for (int e = 0; e < 3; ++e)
                                acc += j3[((3 * f + g) * 3 + d) * 3 + e] * gphi[b][d] * gphi[b2][e];
                        Ke[(3 * b + f) * nd + 3 * b2 + g] += wd * acc;
                    }
    }
    return 0;
}

int oracle_element_matrix(const oracle_mesh *m, int cell, int jf3_id, double t, double s_tshift,
                          const double *locX, double *Ke) {
    tab_t T;
    auxinfo_t A;
    if (!m || !Ke || make_tab(m->cell_type, &T) || aux_info(m, &A)) return 1;
    oracle_point_jac_fn J3 = oracle_point_jac_kernel(jf3_id);
    if (!J3) return 1;
    return element_matrix(m, &T, &A, cell, J3, t, s_tshift, locX, Ke);
}

int oracle_compute_jacobian_pbdiag(const oracle_mesh *m, const int32_t *cells, int ncells,
                                   int jf3_id, double t, double s_tshift, const double *locX,
                                   double *blocks) {
    tab_t T;
    auxinfo_t A;
    if (!m || !blocks || make_tab(m->cell_type, &T) || aux_info(m, &A)) return 1;
    oracle_point_jac_fn J3 = oracle_point_jac_kernel(jf3_id);
    if (!J3) return 1;
    const int nd = 3 * T.nb;
    double *Ke = (double *)malloc(sizeof(double) * nd * nd);
    const int n = cells ? ncells : m->num_cells;
    for (int i = 0; i < n; ++i) {
        const int c = cells ? cells[i] : i;
        const int err = element_matrix(m, &T, &A, c, J3, t, s_tshift, locX, Ke);
        if (err) { free(Ke); return err; }
        const int32_t *cl = m->closure + (size_t)c * T.nb;
        for (int b = 0; b < T.nb; ++b)
            for (int f = 0; f < 3; ++f)
                for (int g = 0; g < 3; ++g)
                    blocks[9 * (size_t)cl[b] + 3 * f + g] += Ke[(3 * b + f) * nd + 3 * b + g];
    }
    free(Ke);
    return 0;
}

int oracle_update_state(const oracle_mesh *m, int kernel_id, int subfield, double t,
                        const double *locX, double *aux_out) {
    tab_t T;
    auxinfo_t A;
    if (!m || !aux_out || !m->aux || !m->aux_per_qp || make_tab(m->cell_type, &T) ||
        aux_info(m, &A) || subfield < 0 || subfield >= m->num_aux)
        return 1;
    oracle_point_fn F = oracle_point_kernel(kernel_id);
    if (!F) return 1;
    const int sOff[1] = {0}, sOff_x[1] = {0};
    const int na = A.naux_tot;
    for (int c = 0; c < m->num_cells; ++c) {
        double ue[MAXNB][3], X[8][3];
        gather(m, &T, c, locX, ue);
        gather_coords(m, &T, c, X);
        for (int q = 0; q < T.nq; ++q) {
            double x[3], Ji[9], detJ;
            if (geometry(&T, (const double(*)[3])X, q, x, Ji, &detJ)) return 2;
            double s[3] = {0}, sx[9] = {0};
            for (int b = 0; b < T.nb; ++b)
                for (int k = 0; k < 3; ++k) {
                    s[k] += ue[b][k] * T.phi[q][b];
                    for (int d = 0; d < 3; ++d)
                        sx[3 * k + d] += ue[b][k] *
                                         (T.dphi[q][b][0] * Ji[d] + T.dphi[q][b][1] * Ji[3 + d] +
                                          T.dphi[q][b][2] * Ji[6 + d]);
                }
            const double *a = aux_at(m, &A, &T, c, q);
            double out[MAXAUX] = {0};
            F(3, 1, m->num_aux, sOff, sOff_x, s, NULL, sx, A.aOff, A.aOff_x, a, NULL, NULL, t, x,
              m->num_constants, m->constants, out);
            double *dst = aux_out + ((size_t)c * T.nq + q) * na;
            for (int k = 0; k < m->aux_sizes[subfield]; ++k) dst[A.aOff[subfield] + k] = out[k];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Cohesive cells                                                                               */
/* ------------------------------------------------------------------------------------------ */

static const double quad_ref[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
#define FACE_GAUSS 0.57735026918962576451 /* 1/sqrt(3): 2x2 Gauss points on [-1,1]^2, weight 1 */

/* Face geometry at a 2x2 Gauss point: basis psi[4], area element, unit normal
 * (d x/d xi  cross  d x/d eta, which the mesh orients from the negative to the positive side). */
static void face_geometry(const double X[4][3], const double xi[2], double psi[4], double *area,
                          double nrm[3]) {
    double dxi[3] = {0}, deta[3] = {0};
    for (int v = 0; v < 4; ++v) {
        const double a = 1.0 + quad_ref[v][0] * xi[0], b = 1.0 + quad_ref[v][1] * xi[1];
        psi[v] = 0.25 * a * b;
        for (int d = 0; d < 3; ++d) {
            dxi[d] += X[v][d] * 0.25 * quad_ref[v][0] * b;
            deta[d] += X[v][d] * 0.25 * a * quad_ref[v][1];
        }
    }
    double cr[3] = {dxi[1] * deta[2] - dxi[2] * deta[1], dxi[2] * deta[0] - dxi[0] * deta[2],
                    dxi[0] * deta[1] - dxi[1] * deta[0]};
    const double mag = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    *area = mag;
    for (int d = 0; d < 3; ++d) nrm[d] = cr[d] / mag;
}

int oracle_cohesive_residual(const oracle_cohesive *C, double t, const double *locX,
                             double *locF) {
    if (!C || !locX || !locF) return 1;
    const double g = FACE_GAUSS;
    const double constants[6] = {C->ref_dir_1[0], C->ref_dir_1[1], C->ref_dir_1[2],
                                 C->ref_dir_2[0], C->ref_dir_2[1], C->ref_dir_2[2]};
    const int sOff[2] = {0, 6}, sOff_x[2] = {0, 0}, aOff[1] = {0}, aOff_x[1] = {0};
    for (int c = 0; c < C->num_cells; ++c) {
        const int32_t *neg = C->closure_neg + 4 * (size_t)c, *pos = C->closure_pos + 4 * (size_t)c;
        const int32_t *lag = C->closure_lag + 4 * (size_t)c, *fc = C->face_cone + 4 * (size_t)c;
        double X[4][3];
        for (int v = 0; v < 4; ++v)
            for (int d = 0; d < 3; ++d) X[v][d] = C->coords[3 * (size_t)fc[v] + d];
        double rn[4][3] = {{0}}, rp[4][3] = {{0}}, rl[4][3] = {{0}};
        for (int q = 0; q < 4; ++q) {
            const double xi[2] = {(q & 1) ? g : -g, (q & 2) ? g : -g};
            double psi[4], area, n[3];
            face_geometry((const double(*)[3])X, xi, psi, &area, n);
            double s[9] = {0}, slip[3] = {0}, x[3] = {0};
            for (int v = 0; v < 4; ++v)
                for (int k = 0; k < 3; ++k) {
                    s[k] += psi[v] * locX[3 * (size_t)neg[v] + k];
                    s[3 + k] += psi[v] * locX[3 * (size_t)pos[v] + k];
                    s[6 + k] += psi[v] * locX[C->lag_offset + 3 * (size_t)lag[v] + k];
                    slip[k] += psi[v] * C->slip[3 * (size_t)lag[v] + k];
                    x[k] += psi[v] * X[v][k];
                }
            double f0n[3] = {0}, f0p[3] = {0}, f0l[3] = {0};
            f0u_neg(3, 2, 1, sOff, sOff_x, s, NULL, NULL, aOff, aOff_x, slip, NULL, NULL, t, x, n, 6,
                    constants, f0n);
            f0u_pos(3, 2, 1, sOff, sOff_x, s, NULL, NULL, aOff, aOff_x, slip, NULL, NULL, t, x, n, 6,
                    constants, f0p);
            f0l_slip(3, 2, 1, sOff, sOff_x, s, NULL, NULL, aOff, aOff_x, slip, NULL, NULL, t, x, n, 6,
                     constants, f0l);
            for (int v = 0; v < 4; ++v)
                for (int k = 0; k < 3; ++k) {
                    rn[v][k] += area * psi[v] * f0n[k];
                    rp[v][k] += area * psi[v] * f0p[k];
                    rl[v][k] += area * psi[v] * f0l[k];
                }
        }
        for (int v = 0; v < 4; ++v)
            for (int k = 0; k < 3; ++k) {
                locF[3 * (size_t)neg[v] + k] += rn[v][k];
                locF[3 * (size_t)pos[v] + k] += rp[v][k];
                locF[C->lag_offset + 3 * (size_t)lag[v] + k] += rl[v][k];
            }
    }
    return 0;
}

/* Jacobian action of the cohesive rows: Jf0ul_neg = -I, Jf0ul_pos = +I, Jf0lu = [-I, +I]
 * (FaultCohesiveKin.hh:298-417), integrated with the face basis. */
int oracle_cohesive_jacobian_action(const oracle_cohesive *C, const double *locY, double *locF) {
    if (!C || !locY || !locF) return 1;
    const double g = FACE_GAUSS;
    for (int c = 0; c < C->num_cells; ++c) {
        const int32_t *neg = C->closure_neg + 4 * (size_t)c, *pos = C->closure_pos + 4 * (size_t)c;
        const int32_t *lag = C->closure_lag + 4 * (size_t)c, *fc = C->face_cone + 4 * (size_t)c;
        double X[4][3];
        for (int v = 0; v < 4; ++v)
            for (int d = 0; d < 3; ++d) X[v][d] = C->coords[3 * (size_t)fc[v] + d];
        double rn[4][3] = {{0}}, rp[4][3] = {{0}}, rl[4][3] = {{0}};
        for (int q = 0; q < 4; ++q) {
            const double xi[2] = {(q & 1) ? g : -g, (q & 2) ? g : -g};
            double psi[4], area, n[3];
            face_geometry((const double(*)[3])X, xi, psi, &area, n);
            double vn[3] = {0}, vp[3] = {0}, vl[3] = {0};
            for (int v = 0; v < 4; ++v)
                for (int k = 0; k < 3; ++k) {
                    vn[k] += psi[v] * locY[3 * (size_t)neg[v] + k];
                    vp[k] += psi[v] * locY[3 * (size_t)pos[v] + k];
                    vl[k] += psi[v] * locY[C->lag_offset + 3 * (size_t)lag[v] + k];
                }
            for (int v = 0; v < 4; ++v)
                for (int k = 0; k < 3; ++k) {
                    rn[v][k] += area * psi[v] * (-vl[k]);
                    rp[v][k] += area * psi[v] * (+vl[k]);
                    rl[v][k] += area * psi[v] * (vp[k] - vn[k]);
                }
        }
        for (int v = 0; v < 4; ++v)
            for (int k = 0; k < 3; ++k) {
                locF[3 * (size_t)neg[v] + k] += rn[v][k];
                locF[3 * (size_t)pos[v] + k] += rp[v][k];
                locF[C->lag_offset + 3 * (size_t)lag[v] + k] += rl[v][k];
            }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Local <-> global                                                                             */
/* ------------------------------------------------------------------------------------------ */

void oracle_local_to_global(int nlocal, const int32_t *l2g, const double *loc, double *g) {
    for (int i = 0; i < nlocal; ++i)
        if (l2g[i] >= 0) g[l2g[i]] += loc[i];
}

void oracle_global_to_local(int nlocal, const int32_t *l2g, const double *g, double *loc) {
    for (int i = 0; i < nlocal; ++i)
        if (l2g[i] >= 0) loc[i] = g[l2g[i]];
}
