// plx_genmax.hip -- generalized Maxwell viscoelasticity (three Maxwell elements in parallel with
// a spring), Q1 hexes, auxiliary field per quadrature point.
//
// The residual kernel of the reference, IsotropicLinearGenMaxwell3D::f1v_infinitesimalStrain
// (IsotropicLinearGenMaxwell.hh:1054-1097, cauchyStress :262-286, _deviatoricStress :404-450,
// IsotropicLinearMaxwell::viscousStrain :256-278), evaluates at each point
//   sigma = K tr(e) I + 2 mu r0 dev(e) + sum_i 2 mu r_i [x_i ev_i + dq_i (dev(e) - dev(et))]
// with x_i = exp(-dt/tau_i), dq_i = tau_i (1 - x_i) / dt, r0 = 1 - sum r_i, ev_i the viscous
// strains and et the total strain of the previous step.  This is linear in e:
//   sigma = lambda_eff tr(e) I + 2 mu_eff e + s,   mu_eff = mu (r0 + sum r_i dq_i)
//   (the shear factor of Jf3vu_infinitesimalStrain, :1143-1310), lambda_eff = K - 2/3 mu_eff,
//   s = sum_i 2 mu r_i (x_i ev_i - dq_i dev(et))   (fixed during a time step).
// So the residual is the elastic tile operator at (lambda_eff, mu_eff) plus the per-point stress
// s, which this file computes once per time step (6 values per point instead of the 32-value
// auxiliary record), and the Jacobian action is the elastic operator alone.  The state update
// (IntegratorDomain::_updateStateVars, IntegratorDomain.cc:555-590, projecting
// viscousStrain_infinitesimalStrain_asVector and infinitesimalStrain_asVector,
// IsotropicLinearGenMaxwell.hh:1507-1535, Elasticity.hh:586-610) writes the new state and the
// next step's s in the same pass.
#include <cmath>
#include <vector>

#include "plx_element.hpp"
#include "plx_internal.hpp"

namespace plx {

// Strain (xx, yy, zz, xy, yz, xz) at Gauss point q (q = i + 2j + 4k, coordinate -1/sqrt3 for
// index 0) of the Q1 cell with closure cl (PyLith vertex order).
__device__ __forceinline__ void q1_point_strain(const int32_t *__restrict__ cl,
                                                const double *__restrict__ coords,
                                                const double *__restrict__ u, int q, double e[6]) {
    const double g = 0.57735026918962576451;
    const double xi[3] = {(q & 1) ? g : -g, (q & 2) ? g : -g, (q & 4) ? g : -g};
    const double sgn[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                              {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    double dphi[8][3], J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const double a = 1.0 + sgn[v][0] * xi[0], b = 1.0 + sgn[v][1] * xi[1], c = 1.0 + sgn[v][2] * xi[2];
        dphi[v][0] = 0.125 * sgn[v][0] * b * c;
        dphi[v][1] = 0.125 * a * sgn[v][1] * c;
        dphi[v][2] = 0.125 * a * b * sgn[v][2];
        const size_t n = 3 * (size_t)cl[v];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int r = 0; r < 3; ++r) J[3 * d + r] += coords[n + d] * dphi[v][r];
    }
    double Ji[9];
    inv3(J, Ji);
    double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const size_t n = 3 * (size_t)cl[v];
        double gp[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) gp[d] = dphi[v][0] * Ji[d] + dphi[v][1] * Ji[3 + d] + dphi[v][2] * Ji[6 + d];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) G[3 * c + d] += u[n + c] * gp[d];
    }
    e[0] = G[0];
    e[1] = G[4];
    e[2] = G[8];
    e[3] = 0.5 * (G[1] + G[3]);
    e[4] = 0.5 * (G[5] + G[7]);
    e[5] = 0.5 * (G[2] + G[6]);
}

__device__ __forceinline__ void dev6(const double e[6], double d[6]) {
    const double mean = (e[0] + e[1] + e[2]) / 3.0;
    d[0] = e[0] - mean;
    d[1] = e[1] - mean;
    d[2] = e[2] - mean;
    d[3] = e[3];
    d[4] = e[4];
    d[5] = e[5];
}

struct GmArgs {
    int n;                       // key cells
    const int32_t *cells;        // key cell -> mesh cell
    const int32_t *closure;
    const double *coords;
    const double *mat;           // [8][n]
    double dt;
    double *state;               // [24][8][n]
    double *pre;                 // [6][8][n]
};

// s = sum_i 2 mu r_i (x_i ev_i - dq_i dev(et)) from the stored state (after a change of dt or aux).
__global__ void __launch_bounds__(256) genmax_prestress_kernel(GmArgs A) {
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (pos >= A.n) return;
    const size_t n = (size_t)A.n;
    auto at = [&](int k) { return (size_t)(8 * k + q) * n + pos; };
    const double mu = A.mat[pos];
    double et[6], det[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) et[k] = A.state[at(18 + k)];
    dev6(et, det);
    double s[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double tau = A.mat[(2 + i) * n + pos], ri = A.mat[(5 + i) * n + pos];
        const double x = exp(-A.dt / tau), dq = tau * (1.0 - exp(-A.dt / tau)) / A.dt;
        const double w = 2.0 * mu * ri;
#pragma unroll
        for (int k = 0; k < 6; ++k) s[k] += w * (x * A.state[at(6 * i + k)] - dq * det[k]);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) A.pre[(size_t)(8 * k + q) * n + pos] = s[k];
}

// State update at the new displacement u, fused with the next step's state stress.
__global__ void __launch_bounds__(256) genmax_update_kernel(GmArgs A, const double *__restrict__ u) {
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (pos >= A.n) return;
    const size_t n = (size_t)A.n;
    auto at = [&](int k) { return (size_t)(8 * k + q) * n + pos; };
    double e[6], de[6], et[6], det[6];
    q1_point_strain(A.closure + 8 * (size_t)A.cells[pos], A.coords, u, q, e);
    dev6(e, de);
#pragma unroll
    for (int k = 0; k < 6; ++k) et[k] = A.state[at(18 + k)];
    dev6(et, det);
    const double mu = A.mat[pos];
    double s[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double tau = A.mat[(2 + i) * n + pos], ri = A.mat[(5 + i) * n + pos];
        const double x = exp(-A.dt / tau), dq = tau * (1.0 - exp(-A.dt / tau)) / A.dt;
        const double w = 2.0 * mu * ri;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double ev = x * A.state[at(6 * i + k)] + dq * (de[k] - det[k]);
            A.state[at(6 * i + k)] = ev;
            s[k] += w * (x * ev - dq * de[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        A.state[at(18 + k)] = e[k];
        A.pre[(size_t)(8 * k + q) * n + pos] = s[k];
    }
}

static GmArgs gm_args(plx_dm dm, KeyPlan *kp) {
    GmArgs A;
    A.n = (int)kp->cells.size();
    A.cells = kp->d_cells;
    A.closure = dm->d_closure;
    A.coords = dm->d_coords;
    A.mat = kp->d_gm_mat;
    A.dt = kp->constants.empty() ? 0.0 : kp->constants[0];
    A.state = kp->d_state;
    A.pre = kp->d_prestress;
    return A;
}

int launch_genmax_prestress(plx_dm dm, KeyPlan *kp) {
    const GmArgs A = gm_args(dm, kp);
    if (A.n == 0) return 0;
    ScopedEvent ev(dm, "plx:genmax_prestress");
    hipLaunchKernelGGL(genmax_prestress_kernel, dim3((A.n + 255) / 256, 8), dim3(256), 0, dm->stream, A);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

int launch_genmax_update(plx_dm dm, KeyPlan *kp, const double *u) {
    int err = ensure_mesh_indices(dm);
    if (err) return err;
    const GmArgs A = gm_args(dm, kp);
    if (A.n == 0) return 0;
    ScopedEvent ev(dm, "plx:genmax_update");
    hipLaunchKernelGGL(genmax_update_kernel, dim3((A.n + 255) / 256, 8), dim3(256), 0, dm->stream, A, u);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

// Capture of a generalized Maxwell aux field (reference layout per cell and point:
// [..., shear_modulus, bulk_modulus, maxwell_time(3), shear_modulus_ratio(3), viscous_strain(18),
// total_strain(6)], IsotropicLinearGenMaxwell.hh:109-165) and, once dt is known (constants[0],
// IsotropicLinearGenMaxwell.cc:211-212), the effective moduli and the state stress.
int genmax_capture(plx_dm dm, KeyPlan *kp) {
    const int num_aux = kp->num_aux;
    if (dm->cell_type != PLX_HEX_Q1) return set_error(dm, 1, "generalized Maxwell is implemented for Q1 hexes");
    if (!kp->aux_per_qp) return set_error(dm, 1, "generalized Maxwell needs the aux field per quadrature point");
    if (num_aux < 6) return set_error(dm, 1, "generalized Maxwell needs 6 rheology aux subfields");
    const int want[6] = {1, 1, 3, 3, 18, 6};
    for (int i = 0; i < 6; ++i)
        if (kp->aux_sizes[num_aux - 6 + i] != want[i])
            return set_error(dm, 1, "generalized Maxwell aux subfield sizes must end with 1,1,3,3,18,6");
    std::vector<int> off(num_aux);
    int tot = 0;
    for (int i = 0; i < num_aux; ++i) {
        off[i] = tot;
        tot += kp->aux_sizes[i];
    }
    const int n = (int)kp->cells.size(), nq = dm->nq;
    const int o_mu = off[num_aux - 6];  // mu, K, tau(3), ratio(3) are contiguous
    const int o_vs = off[num_aux - 2], o_et = off[num_aux - 1];
    std::vector<double> mat(8 * (size_t)n), state(24 * (size_t)nq * n);
    for (int p = 0; p < n; ++p) {
        const double *a = kp->aux_host.data() + (size_t)kp->cells[p] * nq * tot;
        for (int k = 0; k < 8; ++k) {
            const double v = a[o_mu + k];
            for (int q = 1; q < nq; ++q)
                if (a[(size_t)q * tot + o_mu + k] != v)
                    return set_error(dm, 1, "shear/bulk modulus, Maxwell times and ratios must be uniform within a cell");
            mat[(size_t)k * n + p] = v;
        }
        if (!(mat[p] > 0.0) || !(mat[(size_t)n + p] > 0.0)) return set_error(dm, 1, "non-positive shear or bulk modulus");
        for (int i = 0; i < 3; ++i)
            if (!(mat[(size_t)(2 + i) * n + p] > 0.0)) return set_error(dm, 1, "non-positive Maxwell time");
        for (int q = 0; q < nq; ++q)
            for (int k = 0; k < 24; ++k) {
                const double v = k < 18 ? a[(size_t)q * tot + o_vs + k] : a[(size_t)q * tot + o_et + k - 18];
                state[(size_t)(8 * k + q) * n + p] = v;
            }
    }
    for (double *p : {kp->d_gm_mat, kp->d_state, kp->d_prestress})
        if (p) (void)hipFree(p);
    kp->d_gm_mat = kp->d_state = kp->d_prestress = nullptr;
    PLX_HIP(dm, hipMalloc(&kp->d_gm_mat, mat.size() * sizeof(double)));
    PLX_HIP(dm, hipMemcpy(kp->d_gm_mat, mat.data(), mat.size() * sizeof(double), hipMemcpyHostToDevice));
    PLX_HIP(dm, hipMalloc(&kp->d_state, state.size() * sizeof(double)));
    PLX_HIP(dm, hipMemcpy(kp->d_state, state.data(), state.size() * sizeof(double), hipMemcpyHostToDevice));
    PLX_HIP(dm, hipMalloc(&kp->d_prestress, 6 * (size_t)nq * n * sizeof(double)));
    kp->genmax = true;
    return genmax_refresh(dm, kp);
}

// Effective moduli (host, per cell) and the state stress (device) for the current dt.
int genmax_refresh(plx_dm dm, KeyPlan *kp) {
    if (!kp->genmax) return 0;
    if (kp->constants.empty() || !(kp->constants[0] > 0.0)) {
        kp->has_aux = false;  // wait for the time step
        return 0;
    }
    const double dt = kp->constants[0];
    const int n = (int)kp->cells.size();
    std::vector<double> mat(8 * (size_t)n);
    PLX_HIP(dm, hipMemcpy(mat.data(), kp->d_gm_mat, mat.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double> m64(2 * (size_t)n);
    std::vector<float> m32(2 * (size_t)n);
    for (int p = 0; p < n; ++p) {
        const double mu = mat[p], K = mat[(size_t)n + p];
        double r0 = 1.0, sf = 0.0;
        for (int i = 0; i < 3; ++i) r0 -= mat[(size_t)(5 + i) * n + p];
        sf = mu * r0;
        for (int i = 0; i < 3; ++i) {
            const double tau = mat[(size_t)(2 + i) * n + p];
            sf += mu * (tau * (1.0 - std::exp(-dt / tau)) / dt) * mat[(size_t)(5 + i) * n + p];
        }
        m64[p] = K - 2.0 / 3.0 * sf;
        m64[(size_t)n + p] = sf;
        m32[p] = (float)m64[p];
        m32[(size_t)n + p] = (float)sf;
    }
    if (kp->d_mat64) (void)hipFree(kp->d_mat64);
    if (kp->d_mat32) (void)hipFree(kp->d_mat32);
    kp->d_mat64 = nullptr;
    kp->d_mat32 = nullptr;
    PLX_HIP(dm, hipMalloc(&kp->d_mat64, m64.size() * sizeof(double)));
    PLX_HIP(dm, hipMemcpy(kp->d_mat64, m64.data(), m64.size() * sizeof(double), hipMemcpyHostToDevice));
    PLX_HIP(dm, hipMalloc(&kp->d_mat32, m32.size() * sizeof(float)));
    PLX_HIP(dm, hipMemcpy(kp->d_mat32, m32.data(), m32.size() * sizeof(float), hipMemcpyHostToDevice));
    int err = launch_genmax_prestress(dm, kp);
    if (err) return err;
    kp->has_aux = true;
    return 0;
}

}  // namespace plx

using namespace plx;

extern "C" {

int plx_update_state_vars(plx_dm dm, int label_value, int part, double t, const double *locX) {
    (void)t;
    if (!dm || !locX) return set_error(dm, 1, "null vector");
    KeyPlan *kp = find_key(dm, label_value, part);
    if (!kp) return set_error(dm, 1, "unknown key");
    if (!kp->genmax) return 0;  // no state variables
    if (!kp->has_aux) return set_error(dm, 1, "auxiliary field or time step not set");
    ScopedEvent ev(dm, "PL:IntegratorDomain:updateStateVars");
    return launch_genmax_update(dm, kp, locX);
}

int plx_dm_get_auxiliary(plx_dm dm, int label_value, int part, double *aux) {
    if (!dm || !aux) return set_error(dm, 1, "null output");
    KeyPlan *kp = find_key(dm, label_value, part);
    if (!kp || kp->num_aux == 0) return set_error(dm, 1, "no auxiliary field for key");
    const int num_aux = kp->num_aux, nq = dm->nq, n = (int)kp->cells.size();
    int tot = 0;
    for (int i = 0; i < num_aux; ++i) tot += kp->aux_sizes[i];
    // per-point records back in the caller's point order
    const bool permute = kp->aux_per_qp && !dm->quad_perm.empty();
    auto out_point = [&](int c, int q) { return aux + ((size_t)c * nq + (permute ? dm->quad_perm[q] : q)) * tot; };
    if (permute) {
        for (int c = 0; c < dm->num_cells; ++c)
            for (int q = 0; q < nq; ++q)
                std::copy(kp->aux_host.begin() + ((size_t)c * nq + q) * tot,
                          kp->aux_host.begin() + ((size_t)c * nq + q + 1) * tot, out_point(c, q));
    } else {
        std::copy(kp->aux_host.begin(), kp->aux_host.end(), aux);
    }
    if (!kp->genmax) return 0;
    const int o_et = tot - 6, o_vs = tot - 24;
    std::vector<double> state(24 * (size_t)nq * n);
    PLX_HIP(dm, hipStreamSynchronize(dm->stream));
    PLX_HIP(dm, hipMemcpy(state.data(), kp->d_state, state.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < nq; ++q) {
            double *a = out_point(kp->cells[p], q);
            for (int k = 0; k < 24; ++k) a[k < 18 ? o_vs + k : o_et + k - 18] = state[(size_t)(8 * k + q) * n + p];
        }
    return 0;
}

}  // extern "C"
