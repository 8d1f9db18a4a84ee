// pylith_amd_feassemble.hpp -- C++ host-side mirror of pylith::feassemble::IntegratorDomain
// over the C ABI of libpylith_amd.so (header-only).
//
// Same names, argument meaning and error behaviour as the reference integrator
// (libsrc/pylith/feassemble/IntegratorDomain.hh:25-130, Integrator.hh:26-31,169-200,
// IntegrationData.hh:26-38): kernels are registered per subfield and equation part, the
// residual is *added* into the caller's residual field (the caller zeroes it,
// problems/TimeDependent.cc:657), and every failure throws (as PylithCallPetsc does,
// exceptions/error.hh:33-44).  Fields are PETSc-layout local vectors in device memory.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "pylith_amd.h"

namespace pylith_amd {
namespace feassemble {

// pylith::feassemble::Integrator::EquationPart (Integrator.hh:26-31)
enum EquationPart { LHS = PLX_LHS, RHS = PLX_RHS, LHS_LUMPED_INV = PLX_LHS_LUMPED_INV, LHS_WEIGHTED = PLX_LHS_WEIGHTED };

// A pointwise kernel: the device twin of a PetscPointFn / PetscPointJacFn (plx_pointfn id).
using PointFn = int;

class Error : public std::runtime_error {
  public:
    explicit Error(const std::string &what) : std::runtime_error(what) {}
};

// IntegratorDomain::ResidualKernels (IntegratorDomain.hh:29-47)
struct ResidualKernels {
    std::string subfield;
    EquationPart part = LHS;
    PointFn r0 = PLX_FN_NONE;
    PointFn r1 = PLX_FN_NONE;
    ResidualKernels() = default;
    ResidualKernels(const char *subfieldValue, EquationPart partValue, PointFn r0Value, PointFn r1Value)
        : subfield(subfieldValue), part(partValue), r0(r0Value), r1(r1Value) {}
};

// IntegratorDomain::JacobianKernels (IntegratorDomain.hh:49-80)
struct JacobianKernels {
    std::string subfieldTrial, subfieldBasis;
    EquationPart part = LHS;
    PointFn j0 = PLX_FN_NONE, j1 = PLX_FN_NONE, j2 = PLX_FN_NONE, j3 = PLX_FN_NONE;
    JacobianKernels() = default;
    JacobianKernels(const char *trial, const char *basis, EquationPart partValue, PointFn j0Value, PointFn j1Value,
                    PointFn j2Value, PointFn j3Value)
        : subfieldTrial(trial), subfieldBasis(basis), part(partValue), j0(j0Value), j1(j1Value), j2(j2Value),
          j3(j3Value) {}
};

// feassemble::ProjectKernels (state-variable update kernels of a subfield)
struct ProjectKernels {
    std::string subfield;
    PointFn f = PLX_FN_NONE;
    ProjectKernels() = default;
    ProjectKernels(const char *subfieldValue, PointFn fValue) : subfield(subfieldValue), f(fValue) {}
};

// A local vector of a field on the device (pylith::topology::Field's local vector).
struct Field {
    std::string label;
    double *local = nullptr;  // device, PETSc local layout
    int64_t size = 0;
};

// The assembled Jacobian of the displacement field (the Mat PyLith's problem creates with
// DMCreateMatrix): 3x3 blocks over the mesh nodes, the pattern on the host, the values on the
// device.  createMatrix gives it the DM's pattern and zeroed values.
struct Matrix {
    std::vector<int32_t> rowPtr, col;  // CSR over the nodes, columns ascending per row
    double *values = nullptr;          // device [nnzb][3][3]
    int64_t nnzb = 0;
};

inline Matrix createMatrix(plx_dm dm) {
    Matrix m;
    if (plx_jacobian_pattern(dm, &m.nnzb, nullptr, nullptr)) throw Error(std::string("createMatrix: ") + plx_dm_last_error(dm));
    int64_t nrows = 0, nglob = 0;
    plx_dm_get_sizes(dm, &nrows, &nglob);
    (void)nglob;
    m.col.resize((size_t)m.nnzb);
    // row_ptr holds num_nodes + 1 entries: the displacement dofs are 3 per node
    m.rowPtr.resize((size_t)(nrows / 3) + 1);
    if (plx_jacobian_pattern(dm, &m.nnzb, m.rowPtr.data(), m.col.data()))
        throw Error(std::string("createMatrix: ") + plx_dm_last_error(dm));
    if (hipMalloc(&m.values, (size_t)m.nnzb * 9 * sizeof(double)) != hipSuccess) throw Error("createMatrix: hipMalloc");
    if (hipMemset(m.values, 0, (size_t)m.nnzb * 9 * sizeof(double)) != hipSuccess) throw Error("createMatrix: hipMemset");
    return m;
}

inline void destroyMatrix(Matrix *m) {
    if (m && m->values) (void)hipFree(m->values);
    if (m) m->values = nullptr;
}

// pylith::feassemble::IntegrationData (IntegrationData.hh:26-38): scalars and fields by key.
class IntegrationData {
  public:
    static constexpr const char *time = "time";
    static constexpr const char *time_step = "time_step";
    static constexpr const char *s_tshift = "s_tshift";
    static constexpr const char *solution = "solution";
    static constexpr const char *solution_dot = "solution_dot";
    static constexpr const char *residual = "residual";

    void setScalar(const std::string &key, double value) { scalars_[key] = value; }
    double getScalar(const std::string &key) const {
        auto it = scalars_.find(key);
        if (it == scalars_.end()) throw Error("IntegrationData: no scalar '" + key + "'");
        return it->second;
    }
    bool hasScalar(const std::string &key) const { return scalars_.count(key) != 0; }
    void setField(const std::string &key, const Field *field) { fields_[key] = field; }
    const Field *getField(const std::string &key) const {
        auto it = fields_.find(key);
        if (it == fields_.end()) throw Error("IntegrationData: no field '" + key + "'");
        return it->second;
    }
    bool hasField(const std::string &key) const { return fields_.count(key) != 0; }

  private:
    std::map<std::string, double> scalars_;
    std::map<std::string, const Field *> fields_;
};

// pylith::feassemble::IntegratorDomain for the cells of one material label value.
class IntegratorDomain {
  public:
    IntegratorDomain(plx_dm dm, int labelValue) : dm_(dm), label_(labelValue) {
        if (!dm) throw Error("IntegratorDomain: null DM");
    }

    // IntegratorDomain::setKernelsResidual (IntegratorDomain.cc:134-169)
    void setKernelsResidual(const std::vector<ResidualKernels> &kernels, const Field &solution) {
        (void)solution;
        for (const ResidualKernels &k : kernels) {
            check(plx_dm_add_residual(dm_, label_, k.part, k.r0, k.r1), "setKernelsResidual");
            if (k.part == LHS) hasLHSResidual_ = true;
            if (k.part == RHS) hasRHSResidual_ = true;
        }
    }

    // IntegratorDomain::setKernelsJacobian (IntegratorDomain.cc:171-208)
    void setKernelsJacobian(const std::vector<JacobianKernels> &kernels, const Field &solution) {
        (void)solution;
        for (const JacobianKernels &k : kernels) {
            check(plx_dm_add_jacobian(dm_, label_, k.part, k.j0, k.j1, k.j2, k.j3), "setKernelsJacobian");
            if (k.part == LHS) hasLHSJacobian_ = true;
        }
    }

    // IntegratorDomain::setKernelsUpdateStateVars (IntegratorDomain.cc:220-230): the state
    // update of the material's rheology is a property of its registered kernels here.
    void setKernelsUpdateStateVars(const std::vector<ProjectKernels> &kernels) { updateState_ = !kernels.empty(); }

    // DMSetAuxiliaryVec (IntegratorDomain.cc:447): the aux field of the material, host layout.
    void setAuxiliaryField(const std::vector<int32_t> &subfieldSizes, bool perQuadraturePoint,
                           const std::vector<double> &aux, EquationPart part = LHS) {
        check(plx_dm_set_auxiliary(dm_, label_, part, (int)subfieldSizes.size(), subfieldSizes.data(),
                                   perQuadraturePoint ? 1 : 0, aux.data()),
              "setAuxiliaryField");
    }

    // Integrator::_setKernelConstants (Integrator.cc:264-288)
    void setKernelConstants(const std::vector<double> &constants, EquationPart part = LHS) {
        check(plx_dm_set_constants(dm_, label_, part, (int)constants.size(), constants.data()), "setKernelConstants");
    }

    // IntegratorDomain::computeRHSResidual (IntegratorDomain.cc:385-416)
    void computeRHSResidual(Field *residual, const IntegrationData &data) {
        if (hasRHSResidual_) compute(residual, data, RHS);
    }

    // IntegratorDomain::computeLHSResidual (IntegratorDomain.cc:421-453)
    void computeLHSResidual(Field *residual, const IntegrationData &data) {
        if (hasLHSResidual_) compute(residual, data, LHS);
    }

    // IntegratorDomain::computeLHSJacobian (IntegratorDomain.cc:458-504, Integrator.hh:189-191):
    // adds the material's element matrices into the Jacobian and, when it is another matrix, into
    // the preconditioner matrix (DMPlexComputeJacobianByKey, :489-490; the caller zeroes both,
    // as MatZeroEntries before SNES's Jacobian evaluation).
    void computeLHSJacobian(Matrix *jacobianMat, Matrix *precondMat, const IntegrationData &data) {
        if (!hasLHSJacobian_) return;
        if (!jacobianMat) throw Error("computeLHSJacobian: null Jacobian matrix");
        const Field *solution = data.getField(IntegrationData::solution);
        const Field *dot = data.hasField(IntegrationData::solution_dot) ? data.getField(IntegrationData::solution_dot) : nullptr;
        const double t = data.getScalar(IntegrationData::time);
        const double shift = data.hasScalar(IntegrationData::s_tshift) ? data.getScalar(IntegrationData::s_tshift) : 0.0;
        Matrix *targets[2] = {jacobianMat, precondMat != jacobianMat ? precondMat : nullptr};
        for (Matrix *mat : targets) {
            if (!mat) continue;
            check(plx_compute_jacobian(dm_, label_, LHS, t, shift, solution->local, dot ? dot->local : nullptr,
                                       mat->values, PLX_ADD_VALUES),
                  "computeLHSJacobian");
        }
    }

    // DMPlexComputeJacobianActionByKey (IntegratorDomain.cc:540-541): result += J(solution) v.
    void computeLHSJacobianAction(Field *result, const Field &direction, const IntegrationData &data) {
        if (!hasLHSJacobian_) return;
        const Field *solution = data.getField(IntegrationData::solution);
        const Field *dot = data.hasField(IntegrationData::solution_dot) ? data.getField(IntegrationData::solution_dot) : nullptr;
        const double t = data.getScalar(IntegrationData::time);
        const double shift = data.hasScalar(IntegrationData::s_tshift) ? data.getScalar(IntegrationData::s_tshift) : 0.0;
        check(plx_compute_jacobian_action(dm_, label_, LHS, PLX_FP64, t, shift, solution->local, dot ? dot->local : nullptr,
                                          direction.local, result->local, PLX_ADD_VALUES),
              "computeLHSJacobianAction");
    }

    // The point-block diagonal of the LHS Jacobian (the vpbjacobi preconditioner blocks).
    void computeLHSJacobianPointBlockDiagonal(double *blocks, const IntegrationData &data) {
        const double t = data.getScalar(IntegrationData::time);
        const double shift = data.hasScalar(IntegrationData::s_tshift) ? data.getScalar(IntegrationData::s_tshift) : 0.0;
        check(plx_compute_jacobian_pbdiag(dm_, label_, LHS, t, shift, nullptr, blocks), "computeLHSJacobianPointBlockDiagonal");
    }

    // IntegratorDomain::updateState -> _updateStateVars (IntegratorDomain.cc:555-590)
    void updateState(const IntegrationData &data) {
        if (!updateState_) return;
        const Field *solution = data.getField(IntegrationData::solution);
        check(plx_update_state_vars(dm_, label_, LHS, data.getScalar(IntegrationData::time), solution->local),
              "updateState");
    }

  private:
    void compute(Field *residual, const IntegrationData &data, EquationPart part) {
        if (!residual) throw Error("computeResidual: null residual field");
        const Field *solution = data.getField(IntegrationData::solution);
        const Field *dot = data.hasField(IntegrationData::solution_dot) ? data.getField(IntegrationData::solution_dot) : nullptr;
        check(plx_compute_residual(dm_, label_, part, PLX_FP64, data.getScalar(IntegrationData::time), solution->local,
                                   dot ? dot->local : nullptr, residual->local, PLX_ADD_VALUES),
              "computeResidual");
    }

    void check(int err, const char *where) const {
        if (err) throw Error(std::string(where) + ": " + plx_dm_last_error(dm_) + " (error " + std::to_string(err) + ")");
    }

    plx_dm dm_;
    int label_;
    bool hasLHSResidual_ = false, hasRHSResidual_ = false, hasLHSJacobian_ = false, updateState_ = false;
};

}  // namespace feassemble
}  // namespace pylith_amd
