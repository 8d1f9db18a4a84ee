// Drives the C++ IntegratorDomain mirror (include/pylith_amd_feassemble.hpp) the way PyLith's
// problem layer does: register the elasticity kernels, bind the aux field, zero the residual,
// computeLHSResidual, then the Jacobian action.  Input / output are raw binary files written
// and read by tests/test_gpu_integrator_cpp.py:
//   in:  int32 cell_type, num_cells, num_vertices, num_nodes, num_aux; int32 closure[num_cells*nb];
//        double coords[num_vertices*3]; int32 aux_sizes[num_aux]; double aux[num_cells*sum];
//        double u[3*num_nodes]
//   out: double residual[3*num_nodes], double jacobian_action[3*num_nodes],
//        int64 nnzb, int32 row_ptr[num_nodes+1], int32 col[nnzb], double jacobian[nnzb*9]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <numeric>
#include <vector>

#include "pylith_amd_feassemble.hpp"

using namespace pylith_amd::feassemble;

template <typename T>
static void rd(std::ifstream &f, std::vector<T> &v, size_t n) {
    v.resize(n);
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(T)));
}

int main(int argc, char **argv) {
    if (argc != 3) {
        std::cerr << "usage: integrator_driver <in> <out>\n";
        return 2;
    }
    std::ifstream f(argv[1], std::ios::binary);
    int32_t hdr[5];
    f.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
    const int ct = hdr[0], nc = hdr[1], nv = hdr[2], nn = hdr[3], na = hdr[4];
    const int nb = ct == PLX_HEX_Q1 ? 8 : 10;
    std::vector<int32_t> closure, sizes;
    std::vector<double> coords, aux, u;
    rd(f, closure, (size_t)nc * nb);
    rd(f, coords, (size_t)nv * 3);
    rd(f, sizes, (size_t)na);
    const int tot = std::accumulate(sizes.begin(), sizes.end(), 0);
    rd(f, aux, (size_t)nc * tot);
    rd(f, u, (size_t)3 * nn);
    if (!f) {
        std::cerr << "short input\n";
        return 2;
    }
    plx_dm dm = nullptr;
    if (plx_dm_create(&dm, ct, nc, nv, nn, closure.data(), ct == PLX_HEX_Q1 ? closure.data() : nullptr, coords.data(),
                      nullptr)) {
        std::cerr << "plx_dm_create failed\n";
        return 1;
    }
    if (ct == PLX_TET_P2) {
        std::cerr << "the driver covers Q1 meshes\n";
        return 2;
    }
    int rc = 0;
    try {
        Field solution{"solution", nullptr, 3 * (int64_t)nn}, residual{"residual", nullptr, 3 * (int64_t)nn};
        Field jv{"jacobian_action", nullptr, 3 * (int64_t)nn};
        for (Field *fld : {&solution, &residual, &jv})
            if (hipMalloc(&fld->local, fld->size * sizeof(double)) != hipSuccess) throw Error("hipMalloc");
        (void)hipMemcpy(solution.local, u.data(), u.size() * sizeof(double), hipMemcpyHostToDevice);
        IntegratorDomain integrator(dm, 0);
        // IsotropicLinearElasticity::_setKernelsResidual / _setKernelsJacobian (quasi-static,
        // no body force): f0 = NULL, f1 = f1v_infinitesimalStrain; Jf3 = Jf3vu_infinitesimalStrain
        integrator.setKernelsResidual({ResidualKernels("displacement", LHS, PLX_FN_NONE,
                                                       PLX_F1V_ISOLINEAR_ELASTICITY_3D)},
                                      solution);
        integrator.setKernelsJacobian({JacobianKernels("displacement", "displacement", LHS, PLX_FN_NONE,
                                                       PLX_FN_NONE, PLX_FN_NONE, PLX_JF3VU_ISOLINEAR_ELASTICITY_3D)},
                                      solution);
        integrator.setAuxiliaryField(sizes, false, aux);
        IntegrationData data;
        data.setScalar(IntegrationData::time, 0.0);
        data.setScalar(IntegrationData::s_tshift, 0.0);
        data.setField(IntegrationData::solution, &solution);
        (void)hipMemset(residual.local, 0, residual.size * sizeof(double));  // TimeDependent.cc:657
        (void)hipMemset(jv.local, 0, jv.size * sizeof(double));
        integrator.computeRHSResidual(&residual, data);  // no RHS kernels: adds nothing
        integrator.computeLHSResidual(&residual, data);
        integrator.computeLHSJacobianAction(&jv, solution, data);
        // the assembled Jacobian (DMCreateMatrix + IntegratorDomain::computeLHSJacobian), the
        // preconditioner matrix being the Jacobian itself
        Matrix jac = createMatrix(dm);
        integrator.computeLHSJacobian(&jac, &jac, data);
        (void)hipDeviceSynchronize();
        std::vector<double> jvals((size_t)jac.nnzb * 9);
        (void)hipMemcpy(jvals.data(), jac.values, jvals.size() * sizeof(double), hipMemcpyDeviceToHost);
        destroyMatrix(&jac);
        std::vector<double> r(residual.size), j(jv.size);
        (void)hipMemcpy(r.data(), residual.local, r.size() * sizeof(double), hipMemcpyDeviceToHost);
        (void)hipMemcpy(j.data(), jv.local, j.size() * sizeof(double), hipMemcpyDeviceToHost);
        std::ofstream o(argv[2], std::ios::binary);
        o.write(reinterpret_cast<const char *>(r.data()), (std::streamsize)(r.size() * sizeof(double)));
        o.write(reinterpret_cast<const char *>(j.data()), (std::streamsize)(j.size() * sizeof(double)));
        o.write(reinterpret_cast<const char *>(&jac.nnzb), sizeof(int64_t));
        o.write(reinterpret_cast<const char *>(jac.rowPtr.data()), (std::streamsize)(jac.rowPtr.size() * sizeof(int32_t)));
        o.write(reinterpret_cast<const char *>(jac.col.data()), (std::streamsize)(jac.col.size() * sizeof(int32_t)));
        o.write(reinterpret_cast<const char *>(jvals.data()), (std::streamsize)(jvals.size() * sizeof(double)));
        // a kernel without a device twin is refused with an exception (PylithCallPetsc throws)
        try {
            integrator.setKernelsResidual({ResidualKernels("displacement", LHS, PLX_F1V_ISOLINEAR_ELASTICITY_3D,
                                                           PLX_F1V_ISOLINEAR_ELASTICITY_3D)},
                                          solution);
            std::cerr << "expected an exception\n";
            rc = 1;
        } catch (const Error &e) {
            std::cout << "refused: " << e.what() << "\n";
        }
        for (Field *fld : {&solution, &residual, &jv}) (void)hipFree(fld->local);
    } catch (const Error &e) {
        std::cerr << "error: " << e.what() << "\n";
        rc = 1;
    }
    plx_dm_destroy(dm);
    return rc;
}
