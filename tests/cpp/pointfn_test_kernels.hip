// Test-only pointwise kernels written in the PetscPointFn / PetscPointJacFn parameter list, in a
// translation unit of their own, integrated by the launchers of include/pylith_amd_pointfn.hpp:
// how a caller plugs a kernel without a built-in twin into the device path.  Built into
// tests/cpp/libpointfn_test.so; the test hands the launchers to plx_dm_add_*_pointfn.
//
// Aux [density, shear_modulus, bulk_modulus], constants [c0, c1, c2] (the same kernels as the
// oracle's test kernels, oracle/oracle.c):
//   f0 = c1 s + density (c0 x + t);  f1 = -sigma(mu(x), K), mu(x) = mu (1 + c2 sin(x/1000 + z/2000));
//   Jf0 = c1 I;  Jf3 = -C(mu(x), K).
#include "pylith_amd_pointfn.hpp"

using namespace pylith_amd::pointfn;

__device__ static double shear_at(const PylithScalar a[], const PylithInt aOff[], PylithInt numA,
                                  const PylithScalar x[], const PylithScalar constants[]) {
    return a[aOff[numA - 2]] * (1.0 + constants[2] * sin(x[0] / 1000.0 + x[2] / 2000.0));
}

__device__ void f0_test(const PylithInt dim, const PylithInt numS, const PylithInt numA, const PylithInt sOff[],
                        const PylithInt sOff_x[], const PylithScalar s[], const PylithScalar s_t[],
                        const PylithScalar s_x[], const PylithInt aOff[], const PylithInt aOff_x[],
                        const PylithScalar a[], const PylithScalar a_t[], const PylithScalar a_x[],
                        const PylithReal t, const PylithScalar x[], const PylithInt numConstants,
                        const PylithScalar constants[], PylithScalar f0[]) {
    for (PylithInt i = 0; i < dim; ++i)
        f0[i] += constants[1] * s[sOff[0] + i] + a[aOff[0]] * (constants[0] * x[i] + t + s_t[sOff[0] + i]);
}

__device__ void f1_test(const PylithInt dim, const PylithInt numS, const PylithInt numA, const PylithInt sOff[],
                        const PylithInt sOff_x[], const PylithScalar s[], const PylithScalar s_t[],
                        const PylithScalar s_x[], const PylithInt aOff[], const PylithInt aOff_x[],
                        const PylithScalar a[], const PylithScalar a_t[], const PylithScalar a_x[],
                        const PylithReal t, const PylithScalar x[], const PylithInt numConstants,
                        const PylithScalar constants[], PylithScalar f1[]) {
    const double mu = shear_at(a, aOff, numA, x, constants), K = a[aOff[numA - 1]];
    const PylithScalar *g = &s_x[sOff_x[0]];
    const double tr = g[0] + g[4] + g[8];
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) {
            const double eps = 0.5 * (g[3 * c + d] + g[3 * d + c]);
            f1[3 * c + d] -= 2.0 * mu * eps + (c == d ? (K - 2.0 / 3.0 * mu) * tr : 0.0);
        }
}

__device__ void jf0_test(const PylithInt dim, const PylithInt numS, const PylithInt numA, const PylithInt sOff[],
                         const PylithInt sOff_x[], const PylithScalar s[], const PylithScalar s_t[],
                         const PylithScalar s_x[], const PylithInt aOff[], const PylithInt aOff_x[],
                         const PylithScalar a[], const PylithScalar a_t[], const PylithScalar a_x[],
                         const PylithReal t, const PylithReal s_tshift, const PylithScalar x[],
                         const PylithInt numConstants, const PylithScalar constants[], PylithScalar Jf0[]) {
    for (PylithInt i = 0; i < dim; ++i) Jf0[i * dim + i] += constants[1] + s_tshift * a[aOff[0]];
}

__device__ void jf3_test(const PylithInt dim, const PylithInt numS, const PylithInt numA, const PylithInt sOff[],
                         const PylithInt sOff_x[], const PylithScalar s[], const PylithScalar s_t[],
                         const PylithScalar s_x[], const PylithInt aOff[], const PylithInt aOff_x[],
                         const PylithScalar a[], const PylithScalar a_t[], const PylithScalar a_x[],
                         const PylithReal t, const PylithReal s_tshift, const PylithScalar x[],
                         const PylithInt numConstants, const PylithScalar constants[], PylithScalar Jf3[]) {
    const double mu = shear_at(a, aOff, numA, x, constants), K = a[aOff[numA - 1]];
    const double lambda = K - 2.0 / 3.0 * mu;
    // -C_{f d g e} = -(lambda delta_fd delta_ge + mu (delta_fg delta_de + delta_fe delta_dg))
    for (int f = 0; f < 3; ++f)
        for (int g = 0; g < 3; ++g)
            for (int d = 0; d < 3; ++d)
                for (int e = 0; e < 3; ++e)
                    Jf3[((f * 3 + g) * 3 + d) * 3 + e] -= lambda * (f == d) * (g == e) +
                                                          mu * ((f == g) * (d == e) + (f == e) * (d == g));
}

PLX_POINTFN(TestF0, f0_test);
PLX_POINTFN(TestF1, f1_test);
PLX_POINTJAC(TestJf0, jf0_test);
PLX_POINTJAC(TestJf3, jf3_test);

// The launchers: residual, Jacobian action, Jacobian assembly.
extern "C" int plx_test_pointfn_launchers(plx_pointfn_launcher out[3]) {
    out[0] = launch_residual<TestF0, TestF1>;
    out[1] = launch_jacobian_action<TestJf0, None, None, TestJf3>;
    out[2] = launch_jacobian_assembly<TestJf0, None, None, TestJf3>;
    return 0;
}
