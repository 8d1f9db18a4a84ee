// plx_pointfn.hip -- the PetscPointFn-form path: the library's twins of the reference kernels
// written with the PetscPointFn / PetscPointJacFn parameter list and integrated by the generic
// launchers of include/pylith_amd_pointfn.hpp, caller-registered launchers, and the assembled
// Jacobian (DMPlexComputeJacobianByKey, IntegratorDomain.cc:489-490).
#include <algorithm>
#include <cstring>
#include <thread>

#include "../../include/pylith_amd_pointfn.hpp"
#include "plx_internal.hpp"

namespace plx {

using pylith_amd::pointfn::None;
using pylith_amd::pointfn::PylithInt;
using pylith_amd::pointfn::PylithReal;
using pylith_amd::pointfn::PylithScalar;

// ---------------------------------------------------------------------------------------------
// Pointwise kernels in the reference's form (device).
// ---------------------------------------------------------------------------------------------
// Infinitesimal strain from the displacement gradient (Elasticity3D::infinitesimalStrain,
// Elasticity.hh:559-575), vector order xx, yy, zz, xy, yz, xz (Tensor.hh:21-249).
__device__ inline void pf_strain(const PylithScalar *g, double e[6]) {
    e[0] = g[0];
    e[1] = g[4];
    e[2] = g[8];
    e[3] = 0.5 * (g[1] + g[3]);
    e[4] = 0.5 * (g[5] + g[7]);
    e[5] = 0.5 * (g[2] + g[6]);
}

// f1 -= sigma for a stress vector (TensorOps::toTensor3D, Tensor.hh:235-247, then Elasticity::f1v).
__device__ inline void pf_sub_stress(const double sig[6], PylithScalar f1[]) {
    f1[0] -= sig[0];
    f1[1] -= sig[3];
    f1[2] -= sig[5];
    f1[3] -= sig[3];
    f1[4] -= sig[1];
    f1[5] -= sig[4];
    f1[6] -= sig[5];
    f1[7] -= sig[4];
    f1[8] -= sig[2];
}

// IsotropicLinearElasticity3D::f1v_infinitesimalStrain (IsotropicLinearElasticity.hh:803-838):
// sigma = meanStress + deviatoricStress (:182-193, :221-255); aux [..., shear_modulus, bulk_modulus].
__device__ void f1v_infinitesimalStrain(const PylithInt dim, const PylithInt numS, const PylithInt numA,
                                        const PylithInt sOff[], const PylithInt sOff_x[], const PylithScalar s[],
                                        const PylithScalar s_t[], const PylithScalar s_x[], const PylithInt aOff[],
                                        const PylithInt aOff_x[], const PylithScalar a[], const PylithScalar a_t[],
                                        const PylithScalar a_x[], const PylithReal t, const PylithScalar x[],
                                        const PylithInt numConstants, const PylithScalar constants[],
                                        PylithScalar f1[]) {
    const double shearModulus = a[aOff[numA - 2]];
    const double bulkModulus = a[aOff[numA - 1]];
    double e[6];
    pf_strain(&s_x[sOff_x[0]], e);
    const double trace = e[0] + e[1] + e[2];
    const double mean = bulkModulus * trace;
    const double traceTerm = -2.0 / 3.0 * shearModulus * trace;
    double sig[6];
    for (int i = 0; i < 3; ++i) sig[i] = mean + 2.0 * shearModulus * e[i] + traceTerm;
    for (int i = 3; i < 6; ++i) sig[i] = 2.0 * shearModulus * e[i];
    pf_sub_stress(sig, f1);
}

// IsotropicLinearElasticity3D::f1v_infinitesimalStrain_refState (IsotropicLinearElasticity.hh:848-880):
// aux [..., reference_stress(6), reference_strain(6), shear_modulus, bulk_modulus]
// (meanStress_refState :204-216, deviatoricStress_refState :261-304).
__device__ void f1v_infinitesimalStrain_refState(const PylithInt dim, const PylithInt numS, const PylithInt numA,
                                                 const PylithInt sOff[], const PylithInt sOff_x[],
                                                 const PylithScalar s[], const PylithScalar s_t[],
                                                 const PylithScalar s_x[], const PylithInt aOff[],
                                                 const PylithInt aOff_x[], const PylithScalar a[],
                                                 const PylithScalar a_t[], const PylithScalar a_x[],
                                                 const PylithReal t, const PylithScalar x[],
                                                 const PylithInt numConstants, const PylithScalar constants[],
                                                 PylithScalar f1[]) {
    const PylithScalar *refStress = &a[aOff[numA - 4]];
    const PylithScalar *refStrain = &a[aOff[numA - 3]];
    const double shearModulus = a[aOff[numA - 2]];
    const double bulkModulus = a[aOff[numA - 1]];
    double e[6];
    pf_strain(&s_x[sOff_x[0]], e);
    const double trace = e[0] + e[1] + e[2];
    const double refTrace = refStrain[0] + refStrain[1] + refStrain[2];
    const double refMean = (refStress[0] + refStress[1] + refStress[2]) / 3.0;
    const double mean = refMean + bulkModulus * (trace - refTrace);
    const double traceTerm = -2.0 / 3.0 * shearModulus * (trace - refTrace);
    double sig[6];
    for (int i = 0; i < 3; ++i)
        sig[i] = mean + refStress[i] - refMean + 2.0 * shearModulus * (e[i] - refStrain[i]) + traceTerm;
    for (int i = 3; i < 6; ++i) sig[i] = refStress[i] + 2.0 * shearModulus * (e[i] - refStrain[i]);
    pf_sub_stress(sig, f1);
}

// Elasticity::g0v_grav (Elasticity.hh:249-284): f0 += density * gravitational_acceleration;
// aux [density, gravitational_acceleration(3), ...].
__device__ void g0v_grav(const PylithInt dim, const PylithInt numS, const PylithInt numA, const PylithInt sOff[],
                         const PylithInt sOff_x[], const PylithScalar s[], const PylithScalar s_t[],
                         const PylithScalar s_x[], const PylithInt aOff[], const PylithInt aOff_x[],
                         const PylithScalar a[], const PylithScalar a_t[], const PylithScalar a_x[],
                         const PylithReal t, const PylithScalar x[], const PylithInt numConstants,
                         const PylithScalar constants[], PylithScalar f0[]) {
    const double density = a[aOff[0]];
    const PylithScalar *g = &a[aOff[1]];
    for (PylithInt i = 0; i < dim; ++i) f0[i] += density * g[i];
}

// Elasticity::g0v_bodyforce (Elasticity.hh:300-329): f0 += body_force; aux [density, body_force(3), ...].
__device__ void g0v_bodyforce(const PylithInt dim, const PylithInt numS, const PylithInt numA,
                              const PylithInt sOff[], const PylithInt sOff_x[], const PylithScalar s[],
                              const PylithScalar s_t[], const PylithScalar s_x[], const PylithInt aOff[],
                              const PylithInt aOff_x[], const PylithScalar a[], const PylithScalar a_t[],
                              const PylithScalar a_x[], const PylithReal t, const PylithScalar x[],
                              const PylithInt numConstants, const PylithScalar constants[], PylithScalar f0[]) {
    const PylithScalar *b = &a[aOff[1]];
    for (PylithInt i = 0; i < dim; ++i) f0[i] += b[i];
}

// Elasticity::g0v_gravbodyforce (Elasticity.hh:346-373): gravity then body force;
// aux [density, body_force(3), gravitational_acceleration(3), ...].
__device__ void g0v_gravbodyforce(const PylithInt dim, const PylithInt numS, const PylithInt numA,
                                  const PylithInt sOff[], const PylithInt sOff_x[], const PylithScalar s[],
                                  const PylithScalar s_t[], const PylithScalar s_x[], const PylithInt aOff[],
                                  const PylithInt aOff_x[], const PylithScalar a[], const PylithScalar a_t[],
                                  const PylithScalar a_x[], const PylithReal t, const PylithScalar x[],
                                  const PylithInt numConstants, const PylithScalar constants[], PylithScalar f0[]) {
    const PylithInt offGrav[2] = {aOff[0], aOff[2]};
    const PylithInt offBody[2] = {aOff[0], aOff[1]};
    g0v_grav(dim, numS, numA, sOff, sOff_x, s, s_t, s_x, offGrav, aOff_x, a, a_t, a_x, t, x, numConstants,
             constants, f0);
    g0v_bodyforce(dim, numS, numA, sOff, sOff_x, s, s_t, s_x, offBody, aOff_x, a, a_t, a_x, t, x, numConstants,
                  constants, f0);
}

// The 21 nonzeros of -C for isotropic C at ((f*3+g)*3+df)*3+dg (IsotropicLinearElasticity.hh:
// 1014-1034, IsotropicLinearGenMaxwell.hh:1284-1304).
__device__ inline void pf_jf3_isotropic(double C1111, double C1122, double C1212, PylithScalar J[]) {
    const int d1111[3] = {0, 40, 80};
    const int d1122[6] = {10, 20, 30, 50, 60, 70};
    const int d1212[12] = {4, 8, 12, 24, 28, 36, 44, 52, 56, 68, 72, 76};
    for (int i = 0; i < 3; ++i) J[d1111[i]] -= C1111;
    for (int i = 0; i < 6; ++i) J[d1122[i]] -= C1122;
    for (int i = 0; i < 12; ++i) J[d1212[i]] -= C1212;
}

// IsotropicLinearElasticity3D::Jf3vu_infinitesimalStrain (IsotropicLinearElasticity.hh:890-1035).
__device__ void Jf3vu_infinitesimalStrain(const PylithInt dim, const PylithInt numS, const PylithInt numA,
                                          const PylithInt sOff[], const PylithInt sOff_x[], const PylithScalar s[],
                                          const PylithScalar s_t[], const PylithScalar s_x[],
                                          const PylithInt aOff[], const PylithInt aOff_x[], const PylithScalar a[],
                                          const PylithScalar a_t[], const PylithScalar a_x[], const PylithReal t,
                                          const PylithReal s_tshift, const PylithScalar x[],
                                          const PylithInt numConstants, const PylithScalar constants[],
                                          PylithScalar Jf3[]) {
    const double shearModulus = a[aOff[numA - 2]];
    const double bulkModulus = a[aOff[numA - 1]];
    const double lambda = bulkModulus - 2.0 / 3.0 * shearModulus;
    pf_jf3_isotropic(lambda + 2.0 * shearModulus, lambda, shearModulus, Jf3);
}

// IsotropicLinearGenMaxwell3D::Jf3vu_infinitesimalStrain (IsotropicLinearGenMaxwell.hh:1143-1310):
// shear factor mu (r0 + sum_i r_i dq_i), dq_i = tau_i (1 - exp(-dt / tau_i)) / dt
// (IsotropicLinearMaxwell::viscousStrainCoeff, IsotropicLinearMaxwell.hh:220-223); aux
// [..., shear_modulus, bulk_modulus, maxwell_time(3), shear_modulus_ratio(3), viscous_strain(18),
// total_strain(6)], constants[0] = dt.
__device__ void Jf3vu_genMaxwell(const PylithInt dim, const PylithInt numS, const PylithInt numA,
                                 const PylithInt sOff[], const PylithInt sOff_x[], const PylithScalar s[],
                                 const PylithScalar s_t[], const PylithScalar s_x[], const PylithInt aOff[],
                                 const PylithInt aOff_x[], const PylithScalar a[], const PylithScalar a_t[],
                                 const PylithScalar a_x[], const PylithReal t, const PylithReal s_tshift,
                                 const PylithScalar x[], const PylithInt numConstants, const PylithScalar constants[],
                                 PylithScalar Jf3[]) {
    const double shearModulus = a[aOff[numA - 6]];
    const double bulkModulus = a[aOff[numA - 5]];
    const PylithScalar *maxwellTime = &a[aOff[numA - 4]];
    const PylithScalar *ratio = &a[aOff[numA - 3]];
    const double dt = constants[0];
    double shearFactor = shearModulus * (1.0 - ratio[0] - ratio[1] - ratio[2]);
    for (int i = 0; i < 3; ++i)
        shearFactor += shearModulus * ratio[i] * maxwellTime[i] * (1.0 - exp(-dt / maxwellTime[i])) / dt;
    pf_jf3_isotropic(bulkModulus + 4.0 * shearFactor / 3.0, bulkModulus - 2.0 * shearFactor / 3.0, shearFactor, Jf3);
}

PLX_POINTFN(F1Iso, f1v_infinitesimalStrain);
PLX_POINTFN(F1IsoRef, f1v_infinitesimalStrain_refState);
PLX_POINTFN(F0Grav, g0v_grav);
PLX_POINTFN(F0Body, g0v_bodyforce);
PLX_POINTFN(F0GravBody, g0v_gravbodyforce);
PLX_POINTJAC(Jf3Iso, Jf3vu_infinitesimalStrain);
PLX_POINTJAC(Jf3GenMaxwell, Jf3vu_genMaxwell);

namespace pf = pylith_amd::pointfn;

template <class F1>
static plx_pointfn_launcher residual_launcher_f0(int f0) {
    switch (f0) {
    case PLX_FN_NONE: return pf::launch_residual<None, F1>;
    case PLX_F0V_BODYFORCE: return pf::launch_residual<F0Body, F1>;
    case PLX_F0V_GRAV: return pf::launch_residual<F0Grav, F1>;
    case PLX_F0V_GRAVBODYFORCE: return pf::launch_residual<F0GravBody, F1>;
    default: return nullptr;
    }
}

static plx_pointfn_launcher builtin_residual(int f0, int f1) {
    if (f1 == PLX_F1V_ISOLINEAR_ELASTICITY_3D) return residual_launcher_f0<F1Iso>(f0);
    if (f1 == PLX_F1V_ISOLINEAR_ELASTICITY_REFSTATE_3D) return residual_launcher_f0<F1IsoRef>(f0);
    return nullptr;
}

static plx_pointfn_launcher builtin_action(int jf3) {
    if (jf3 == PLX_JF3VU_ISOLINEAR_ELASTICITY_3D) return pf::launch_jacobian_action<None, None, None, Jf3Iso>;
    return nullptr;
}

static plx_pointfn_launcher builtin_assembly(int jf3) {
    if (jf3 == PLX_JF3VU_ISOLINEAR_ELASTICITY_3D) return pf::launch_jacobian_assembly<None, None, None, Jf3Iso>;
    if (jf3 == PLX_JF3VU_GENMAXWELL_3D) return pf::launch_jacobian_assembly<None, None, None, Jf3GenMaxwell>;
    return nullptr;
}

bool pointfn_builtin_residual(int f0, int f1) { return builtin_residual(f0, f1) != nullptr; }

// ---------------------------------------------------------------------------------------------
// Set-up: the key's cells by global colour, the aux field and constants on the device.
// ---------------------------------------------------------------------------------------------
static int ensure_colors(plx_dm dm, KeyPlan *kp) {
    if (kp->d_pf_cells) return 0;
    const int nb = dm->nb;
    const size_t n = kp->cells.size();
    std::vector<uint64_t> used(2 * (size_t)dm->num_nodes, 0);
    std::vector<int> color(n);
    int ncolor = 0;
    for (size_t i = 0; i < n; ++i) {
        const int32_t *cl = dm->closure.data() + (size_t)kp->cells[i] * nb;
        uint64_t m0 = 0, m1 = 0;
        for (int b = 0; b < nb; ++b) {
            m0 |= used[2 * (size_t)cl[b]];
            m1 |= used[2 * (size_t)cl[b] + 1];
        }
        int c = 0;
        while (c < 128 && ((c < 64 ? (m0 >> c) : (m1 >> (c - 64))) & 1)) ++c;
        if (c >= 128) return set_error(dm, 1, "cell colouring needs more than 128 colours");
        for (int b = 0; b < nb; ++b) used[2 * (size_t)cl[b] + (c >> 6)] |= (uint64_t)1 << (c & 63);
        color[i] = c;
        ncolor = std::max(ncolor, c + 1);
    }
    kp->pf_color_ptr.assign(ncolor + 1, 0);
    for (size_t i = 0; i < n; ++i) kp->pf_color_ptr[color[i] + 1]++;
    for (int c = 0; c < ncolor; ++c) kp->pf_color_ptr[c + 1] += kp->pf_color_ptr[c];
    std::vector<int32_t> sorted(n);
    std::vector<int32_t> pos(kp->pf_color_ptr.begin(), kp->pf_color_ptr.end() - 1);
    for (size_t i = 0; i < n; ++i) sorted[pos[color[i]]++] = kp->cells[i];
    PLX_HIP(dm, hipMalloc(&kp->d_pf_cells, n * sizeof(int32_t)));
    PLX_HIP(dm, hipMemcpy(kp->d_pf_cells, sorted.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    return 0;
}

static int ensure_aux(plx_dm dm, KeyPlan *kp) {
    if (!kp->aux_raw_valid) {
        if (kp->d_aux_raw) (void)hipFree(kp->d_aux_raw);
        kp->d_aux_raw = nullptr;
        if (!kp->aux_host.empty()) {
            PLX_HIP(dm, hipMalloc(&kp->d_aux_raw, kp->aux_host.size() * sizeof(double)));
            PLX_HIP(dm, hipMemcpy(kp->d_aux_raw, kp->aux_host.data(), kp->aux_host.size() * sizeof(double),
                                  hipMemcpyHostToDevice));
        }
        kp->aux_raw_valid = true;
    }
    if (kp->num_constants_dev != (int)kp->constants.size()) {
        if (kp->d_constants) (void)hipFree(kp->d_constants);
        kp->d_constants = nullptr;
        PLX_HIP(dm, hipMalloc(&kp->d_constants, std::max<size_t>(1, kp->constants.size()) * sizeof(double)));
        kp->num_constants_dev = (int)kp->constants.size();
    }
    if (!kp->constants.empty())
        PLX_HIP(dm, hipMemcpyAsync(kp->d_constants, kp->constants.data(), kp->constants.size() * sizeof(double),
                                   hipMemcpyHostToDevice, dm->stream));
    return 0;
}

static int make_args(plx_dm dm, KeyPlan *kp, plx_pointfn_args *A) {
    int err = ensure_colors(dm, kp);
    if (err || (err = ensure_aux(dm, kp)) || (err = ensure_mesh_indices(dm))) return err;
    int stride = 0;
    for (int s : kp->aux_sizes) stride += s;
    if (kp->num_aux > 16 || stride > pylith_amd::pointfn::kMaxAuxValues)
        return set_error(dm, 1, "the PetscPointFn path takes at most 16 aux subfields and 64 values per record");
    std::memset(A, 0, sizeof(*A));
    A->cell_type = dm->cell_type;
    A->nodes_per_cell = dm->nb;
    A->verts_per_cell = dm->nv;
    A->num_quad = dm->nq;
    A->num_cells = (int)kp->cells.size();
    A->cells = kp->d_pf_cells;
    A->num_colors = (int)kp->pf_color_ptr.size() - 1;
    A->color_ptr = kp->pf_color_ptr.data();
    A->closure = dm->d_closure;
    A->cone = dm->d_cone;
    A->coords = dm->d_coords;
    A->num_aux = kp->num_aux;
    for (int k = 0, off = 0; k < kp->num_aux; ++k) {
        A->aux_off[k] = off;
        off += kp->aux_sizes[k];
    }
    A->aux_stride = stride;
    A->aux_per_qp = kp->aux_per_qp;
    A->aux = kp->d_aux_raw;
    A->num_constants = (int)kp->constants.size();
    A->constants = kp->d_constants;
    A->deterministic = dm->accumulate == PLX_ACCUMULATE_COLORED;
    A->num_tiles = kp->num_tiles;
    A->tile_max_nodes = kp->max_nodes;
    A->tile_max_verts = kp->max_verts;
    A->tile_cell_ptr = kp->d_tile_cell_ptr;
    A->tile_node_ptr = kp->d_tile_node_ptr;
    A->tile_vert_ptr = kp->d_tile_vert_ptr;
    A->tile_nown = kp->d_tile_nown;
    A->tile_halo_off = kp->d_tile_halo_off;
    A->tile_nodes = kp->d_tile_nodes;
    A->tile_verts = kp->d_tile_verts;
    A->tile_cells = kp->d_cells;
    A->lclosure = kp->d_lclosure;
    A->lcone = kp->d_lcone;
    A->num_shared = kp->num_shared;
    A->shared_node = kp->d_shared_node;
    A->shared_ptr = kp->d_shared_ptr;
    A->shared_slot = kp->d_shared_slot;
    A->halo = static_cast<double *>(kp->d_halo);
    return 0;
}

void pointfn_release(plx_dm dm, KeyPlan *kp) {
    (void)dm;
    for (void *p : {(void *)kp->d_pf_cells, (void *)kp->d_aux_raw, (void *)kp->d_constants})
        if (p) (void)hipFree(p);
    kp->d_pf_cells = nullptr;
    kp->d_aux_raw = nullptr;
    kp->d_constants = nullptr;
    kp->aux_raw_valid = false;
    kp->num_constants_dev = -1;
    kp->pf_color_ptr.clear();
}

int pointfn_residual(plx_dm dm, KeyPlan *kp, double t, const void *u, const void *u_t, void *f, int mode) {
    plx_pointfn_launcher fn = kp->pf_residual ? kp->pf_residual : builtin_residual(kp->f0, kp->f1);
    if (!fn) return set_error(dm, 1, "no PetscPointFn-form residual for this key");
    plx_pointfn_args A;
    int err = make_args(dm, kp, &A);
    if (err) return err;
    A.t = t;
    A.u = static_cast<const double *>(u);
    A.u_t = static_cast<const double *>(u_t);
    A.f = static_cast<double *>(f);
    A.mode = mode;
    if ((err = fn(&A, dm->stream))) return set_error(dm, err, "PetscPointFn-form residual launch failed");
    return 0;
}

int pointfn_action(plx_dm dm, KeyPlan *kp, double t, double s_tshift, const void *u, const void *u_t,
                   const void *v, void *f, int mode) {
    plx_pointfn_launcher fn = kp->pf_action ? kp->pf_action : builtin_action(kp->jf3);
    if (!fn) return set_error(dm, 1, "no PetscPointJacFn-form Jacobian action for this key");
    plx_pointfn_args A;
    int err = make_args(dm, kp, &A);
    if (err) return err;
    A.t = t;
    A.s_tshift = s_tshift;
    A.u = static_cast<const double *>(u);
    A.u_t = static_cast<const double *>(u_t);
    A.v = static_cast<const double *>(v);
    A.f = static_cast<double *>(f);
    A.mode = mode;
    if ((err = fn(&A, dm->stream))) return set_error(dm, err, "PetscPointJacFn-form action launch failed");
    return 0;
}

// The blocked node pattern of the mesh: nodes i and j share a cell.
static int ensure_pattern(plx_dm dm) {
    if (dm->d_jac_row_ptr) return 0;
    const int nn = dm->num_nodes, nb = dm->nb;
    std::vector<int32_t> inc_ptr(nn + 1, 0), inc;
    for (int32_t v : dm->closure) inc_ptr[v + 1]++;
    for (int i = 0; i < nn; ++i) inc_ptr[i + 1] += inc_ptr[i];
    inc.resize(inc_ptr[nn]);
    {
        std::vector<int32_t> pos(inc_ptr.begin(), inc_ptr.end() - 1);
        for (int c = 0; c < dm->num_cells; ++c)
            for (int b = 0; b < nb; ++b) inc[pos[dm->closure[(size_t)c * nb + b]]++] = c;
    }
    // rows are independent: each host thread sorts a contiguous range of rows into its own
    // list, and the lists are joined in row order
    const int T = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::vector<int32_t>> part(T);
    std::vector<int64_t> len(nn + 1, 0);
    {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                const int lo = (int)((int64_t)nn * t / T), hi = (int)((int64_t)nn * (t + 1) / T);
                std::vector<int32_t> row;
                for (int i = lo; i < hi; ++i) {
                    row.clear();
                    for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k)
                        for (int b = 0; b < nb; ++b) row.push_back(dm->closure[(size_t)inc[k] * nb + b]);
                    std::sort(row.begin(), row.end());
                    row.erase(std::unique(row.begin(), row.end()), row.end());
                    part[t].insert(part[t].end(), row.begin(), row.end());
                    len[i + 1] = (int64_t)row.size();
                }
            });
        for (auto &x : th) x.join();
    }
    for (int i = 0; i < nn; ++i) len[i + 1] += len[i];
    if (len[nn] > (int64_t)INT32_MAX) return set_error(dm, 1, "Jacobian pattern too large");
    dm->jac_row_ptr.resize(nn + 1);
    for (int i = 0; i <= nn; ++i) dm->jac_row_ptr[i] = (int32_t)len[i];
    dm->jac_col.resize((size_t)len[nn]);
    {
        size_t at = 0;
        for (int t = 0; t < T; ++t) {
            std::copy(part[t].begin(), part[t].end(), dm->jac_col.begin() + at);
            at += part[t].size();
            std::vector<int32_t>().swap(part[t]);
        }
    }
    PLX_HIP(dm, hipMalloc(&dm->d_jac_row_ptr, dm->jac_row_ptr.size() * sizeof(int32_t)));
    PLX_HIP(dm, hipMalloc(&dm->d_jac_col, std::max<size_t>(1, dm->jac_col.size()) * sizeof(int32_t)));
    PLX_HIP(dm, hipMemcpy(dm->d_jac_row_ptr, dm->jac_row_ptr.data(), dm->jac_row_ptr.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice));
    PLX_HIP(dm, hipMemcpy(dm->d_jac_col, dm->jac_col.data(), dm->jac_col.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice));
    return 0;
}

}  // namespace plx

using namespace plx;

extern "C" {

int plx_jacobian_pattern(plx_dm dm, int64_t *nnzb, int32_t *row_ptr, int32_t *col) {
    if (!dm || !nnzb) return set_error(dm, 1, "null argument");
    const int err = ensure_pattern(dm);
    if (err) return err;
    *nnzb = (int64_t)dm->jac_col.size();
    if (row_ptr) std::memcpy(row_ptr, dm->jac_row_ptr.data(), dm->jac_row_ptr.size() * sizeof(int32_t));
    if (col) std::memcpy(col, dm->jac_col.data(), dm->jac_col.size() * sizeof(int32_t));
    return 0;
}

int plx_compute_jacobian(plx_dm dm, int label_value, int part, double t, double s_tshift, const double *locX,
                         const double *locX_t, double *values, int mode) {
    if (!dm || !values) return set_error(dm, 1, "null values");
    if (mode != PLX_ADD_VALUES && mode != PLX_INSERT_VALUES) return set_error(dm, 1, "bad insert mode");
    KeyPlan *kp = find_key(dm, label_value, part);
    if (!kp || (!kp->jf3 && !kp->pf_assemble)) return set_error(dm, 1, "no Jacobian kernels registered for key");
    if (!kp->has_aux) return set_error(dm, 1, "auxiliary field not set");
    plx_pointfn_launcher fn = kp->pf_assemble ? kp->pf_assemble : builtin_assembly(kp->jf3);
    if (!fn) return set_error(dm, 1, "no assembled form of this Jacobian kernel");
    int err = ensure_pattern(dm);
    if (err) return err;
    plx_pointfn_args A;
    if ((err = make_args(dm, kp, &A))) return err;
    ScopedEvent ev(dm, "PL:IntegratorDomain:computeLHSJacobian");
    if (mode == PLX_INSERT_VALUES) {
        const int64_t n = 9 * (int64_t)dm->jac_col.size();
        hipLaunchKernelGGL(pylith_amd::pointfn::fill_values, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 8192)),
                           dim3(256), 0, dm->stream, n, values);
        PLX_HIP(dm, hipGetLastError());
    }
    A.t = t;
    A.s_tshift = s_tshift;
    A.u = locX;
    A.u_t = locX_t;
    A.mode = mode;
    A.row_ptr = dm->d_jac_row_ptr;
    A.col = dm->d_jac_col;
    A.values = values;
    if ((err = fn(&A, dm->stream))) return set_error(dm, err, "Jacobian assembly launch failed");
    return 0;
}

}  // extern "C"
