// plx_dm.cpp -- host side of the C ABI: mesh upload, tile plans, weak-form keys, aux capture.
//
// Mirrors the set-up half of PyLith's assembly path: DMPlex + section + closure index
// (topology/Field.cc:277-283, topology/CoordsVisitor.icc:145-157), kernel registration
// (feassemble/IntegratorDomain.cc:134-190) and DMSetAuxiliaryVec (IntegratorDomain.cc:447).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "plx_element.hpp"
#include "plx_internal.hpp"

namespace plx {

int set_error(plx_dm dm, int code, const std::string &msg) {
    if (dm) dm->last_error = msg;
    return code;
}

KeyPlan *find_key(plx_dm dm, int label_value, int part) {
    for (KeyPlan *k : dm->keys)
        if (k->label_value == label_value && k->part == part) return k;
    return nullptr;
}

template <typename T>
static int upload(plx_dm dm, T **dst, const T *src, size_t n) {
    *dst = nullptr;
    if (n == 0) return 0;
    PLX_HIP(dm, hipMalloc((void **)dst, n * sizeof(T)));
    PLX_HIP(dm, hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

static void free_plan_arrays(KeyPlan *k) {
    pointfn_release(nullptr, k);
    void *ptrs[] = {k->d_cells,       k->d_tile_cell_ptr, k->d_tile_node_ptr, k->d_tile_nown,
                    k->d_tile_nodes,  k->d_tile_vert_ptr, k->d_tile_verts,    k->d_tile_halo_off,
                    k->d_lclosure,    k->d_lcone,         k->d_shared_node,   k->d_color,
                    k->d_tile_ncolor,   k->d_records,       k->d_shared_ptr,    k->d_tiles_affine,  k->d_tiles_general,
                    k->d_shared_slot, k->d_halo,          k->d_mat64,         k->d_mat32,
                    k->d_state,       k->d_gm_mat,        k->d_prestress,     k->d_load_node,
                    k->d_load64,      k->d_load32};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    KeyPlan fresh;
    fresh.label_value = k->label_value;
    fresh.part = k->part;
    fresh.f0 = k->f0;
    fresh.f1 = k->f1;
    fresh.jf0 = k->jf0;
    fresh.jf1 = k->jf1;
    fresh.jf2 = k->jf2;
    fresh.jf3 = k->jf3;
    fresh.num_aux = k->num_aux;
    fresh.aux_sizes = std::move(k->aux_sizes);
    fresh.aux_per_qp = k->aux_per_qp;
    fresh.aux_host = std::move(k->aux_host);
    fresh.constants = std::move(k->constants);
    fresh.pf_residual = k->pf_residual;
    fresh.pf_action = k->pf_action;
    fresh.pf_assemble = k->pf_assemble;
    *k = std::move(fresh);
}

static void free_key(KeyPlan *k) {
    free_plan_arrays(k);
    delete k;
}

// Build the key's tile plan on the host (plx_plan.cpp) and upload it.
static int build_plan(plx_dm dm, KeyPlan *kp) {
    MeshView mv{dm->nb, dm->nv, dm->num_cells, dm->num_nodes, dm->num_vertices,
                dm->closure.data(), dm->cone.data(), dm->labels.data(),
                dm->cell_type == PLX_HEX_Q1, dm->coords.data()};
    HostPlan hp;
    std::string msg;
    // P2 tiles keep colour order (the tets of a cube share nodes: colour order keeps a wave's
    // atomic targets distinct); Q1 tiles keep mesh order (structured LDS access).
    kp->sorted_by_color = dm->cell_type == PLX_TET_P2;
    // Small keys get smaller tiles while there would be fewer tiles than half the CUs (not below
    // 64 cells): idle CUs cost more than the extra halo nodes there; larger keys keep the full
    // tile (fewer halo nodes).  Measured on Q1 32^3: 10.5 us per apply at 256-cell tiles against
    // 13.1 us at 512; at 48^3 smaller tiles no longer pay.
    TileLimits lim = tile_limits(dm->cell_type);
    {
        int64_t nkey = 0;
        for (int c = 0; c < dm->num_cells; ++c)
            if ((dm->labels.empty() ? 0 : dm->labels[c]) == kp->label_value) ++nkey;
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || ncu <= 0) ncu = 256;
        int cells = lim.cells;
        while (cells > 64 && nkey < (int64_t)cells * (ncu / 2)) cells >>= 1;
        lim.cells = cells;
    }
    if (build_host_plan(mv, kp->label_value, lim, &hp, &msg, kp->sorted_by_color))
        return set_error(dm, 1, msg);
    kp->cells = hp.cells;
    kp->num_tiles = hp.num_tiles;
    kp->max_nodes = hp.max_nodes;
    kp->max_verts = hp.max_verts;
    kp->max_colors = hp.max_colors;
    kp->halo_slots = hp.halo_slots;
    kp->num_shared = (int)hp.shared_node.size();
    kp->reordered = hp.reordered;
    int err = 0;
    if ((err = upload(dm, &kp->d_cells, hp.cells.data(), hp.cells.size()))) return err;
    if ((err = upload(dm, &kp->d_tile_cell_ptr, hp.tile_cell_ptr.data(), hp.tile_cell_ptr.size()))) return err;
    if ((err = upload(dm, &kp->d_tile_node_ptr, hp.tile_node_ptr.data(), hp.tile_node_ptr.size()))) return err;
    if ((err = upload(dm, &kp->d_tile_nown, hp.tile_nown.data(), hp.tile_nown.size()))) return err;
    if ((err = upload(dm, &kp->d_tile_nodes, hp.tile_nodes.data(), hp.tile_nodes.size()))) return err;
    if ((err = upload(dm, &kp->d_tile_vert_ptr, hp.tile_vert_ptr.data(), hp.tile_vert_ptr.size()))) return err;
    if ((err = upload(dm, &kp->d_tile_verts, hp.tile_verts.data(), hp.tile_verts.size()))) return err;
    if ((err = upload(dm, &kp->d_tile_halo_off, hp.tile_halo_off.data(), hp.tile_halo_off.size()))) return err;
    if ((err = upload(dm, &kp->d_lclosure, hp.lclosure.data(), hp.lclosure.size()))) return err;
    if ((err = upload(dm, &kp->d_lcone, hp.lcone.data(), hp.lcone.size()))) return err;
    if ((err = upload(dm, &kp->d_color, hp.color.data(), hp.color.size()))) return err;
    {
        // packed per-cell records: Q1 8 x uint16 closure; P2 10 closure + 4 cone + 2 pad
        const int recw = dm->cell_type == PLX_HEX_Q1 ? 8 : 16;
        const size_t n = hp.cells.size();
        std::vector<uint16_t> rec(n * recw, 0);
        for (size_t i = 0; i < n; ++i) {
            for (int b = 0; b < dm->nb; ++b) rec[i * recw + b] = hp.lclosure[(size_t)b * n + i];
            if (dm->cell_type != PLX_HEX_Q1)
                for (int v = 0; v < dm->nv; ++v) rec[i * recw + dm->nb + v] = hp.lcone[(size_t)v * n + i];
        }
        if ((err = upload(dm, &kp->d_records, rec.data(), rec.size()))) return err;
    }
    if ((err = upload(dm, &kp->d_tile_ncolor, hp.tile_ncolor.data(), hp.tile_ncolor.size()))) return err;
    if ((err = upload(dm, &kp->d_shared_node, hp.shared_node.data(), hp.shared_node.size()))) return err;
    if ((err = upload(dm, &kp->d_shared_ptr, hp.shared_ptr.data(), hp.shared_ptr.size()))) return err;
    if ((err = upload(dm, &kp->d_shared_slot, hp.shared_slot.data(), hp.shared_slot.size()))) return err;
    if (hp.halo_slots > 0) PLX_HIP(dm, hipMalloc(&kp->d_halo, (size_t)hp.halo_slots * 3 * sizeof(double)));
    // Q1 tiles whose cells are all parallelepipeds take the modal kernel (geometry is static).
    if (dm->cell_type == PLX_HEX_Q1) {
        std::vector<int32_t> aff, gen;
        for (int t = 0; t < hp.num_tiles; ++t) {
            bool all = true;
            for (int i = hp.tile_cell_ptr[t]; i < hp.tile_cell_ptr[t + 1] && all; ++i) {
                double X[3][8], TX[3][8];
                const int32_t *cn = &dm->cone[(size_t)hp.cells[i] * 8];
                for (int l = 0; l < 8; ++l)
                    for (int d = 0; d < 3; ++d) X[d][l] = dm->coords[3 * (size_t)cn[q1_lex(l)] + d];
                for (int d = 0; d < 3; ++d) wht8<double>(X[d], TX[d]);
                all = q1_affine(TX);
            }
            (all ? aff : gen).push_back(t);
        }
        kp->num_tiles_affine = (int)aff.size();
        kp->num_tiles_general = (int)gen.size();
        if ((err = upload(dm, &kp->d_tiles_affine, aff.data(), aff.size()))) return err;
        if ((err = upload(dm, &kp->d_tiles_general, gen.data(), gen.size()))) return err;
    }
    return 0;
}

// Moduli per key cell in tile order (SoA: lambda then mu), FP64 and FP32, from the captured aux.
// Elastic moduli are the last two subfields (IsotropicLinearElasticity::setContext,
// IsotropicLinearElasticity.hh:116-117) or, for generalized Maxwell, subfields numA-6 and
// numA-5 (IsotropicLinearGenMaxwell.hh:133-134); per-qp aux must be uniform in a cell there.
// The residual terms that do not depend on the solution, as one fixed per-node vector:
//  * body force / gravity (Elasticity::g0v_bodyforce, g0v_grav, g0v_gravbodyforce,
//    fekernels/Elasticity.hh:249-385): sum_q w_q |J_q| phi_b(q) f0(q);
//  * the reference state of IsotropicLinearElasticity3D::f1v_infinitesimalStrain_refState
//    (IsotropicLinearElasticity.hh:848-880, :261-304): sigma = C : eps + (sigma_ref - C : eps_ref),
//    so f1 = -C : eps (the cell pass) - S with S = sigma_ref - C : eps_ref, whose part is
//    -sum_q w_q |J_q| S(q) grad phi_b(q).
// Computed here (host, FP64, in the key's cell order) whenever the kernels or the aux field
// change, and added after the cell pass of every residual (not of the Jacobian action, which
// these terms do not enter).
static int build_load(plx_dm dm, KeyPlan *kp) {
    if (kp->d_load_node) (void)hipFree(kp->d_load_node);
    if (kp->d_load64) (void)hipFree(kp->d_load64);
    if (kp->d_load32) (void)hipFree(kp->d_load32);
    kp->d_load_node = nullptr;
    kp->d_load64 = nullptr;
    kp->d_load32 = nullptr;
    kp->num_load = 0;
    const bool refstate = kp->f1 == PLX_F1V_ISOLINEAR_ELASTICITY_REFSTATE_3D;
    if ((kp->f0 == PLX_FN_NONE && !refstate) || !kp->has_aux) return 0;
    // aux subfields the kernel reads (Elasticity.cc:234-241 order)
    const bool grav = kp->f0 == PLX_F0V_GRAV || kp->f0 == PLX_F0V_GRAVBODYFORCE;
    const bool body = kp->f0 == PLX_F0V_BODYFORCE || kp->f0 == PLX_F0V_GRAVBODYFORCE;
    const int i_body = 1, i_grav = kp->f0 == PLX_F0V_GRAVBODYFORCE ? 2 : 1;
    const int nsub = kp->f0 == PLX_F0V_GRAVBODYFORCE ? 3 : 2;
    if (kp->f0 != PLX_FN_NONE &&
        (kp->num_aux < nsub + (refstate ? 4 : 2) || kp->aux_sizes[0] != 1 || kp->aux_sizes[1] != 3 ||
         (nsub == 3 && kp->aux_sizes[2] != 3)))
        return set_error(dm, 1, "body force / gravity need aux subfields density(1), then 3-vectors");
    const int i_rs = kp->num_aux - 4, i_re = kp->num_aux - 3;
    if (refstate && (kp->num_aux < 5 || kp->aux_sizes[i_rs] != 6 || kp->aux_sizes[i_re] != 6))
        return set_error(dm, 1, "reference state needs aux subfields reference_stress(6), reference_strain(6) before the moduli");
    std::vector<int> off(kp->num_aux);
    int naux_tot = 0;
    for (int i = 0; i < kp->num_aux; ++i) {
        off[i] = naux_tot;
        naux_tot += kp->aux_sizes[i];
    }
    const int per = kp->aux_per_qp ? dm->nq : 1;
    // quadrature weights and basis values (Q1: 2x2x2 Gauss on [-1,1]^3, PyLith vertex order;
    // P2: the 14-point rule, vertices then edges (0,1),(0,2),(0,3),(1,2),(1,3),(2,3))
    const int nb = dm->nb, nq = dm->nq;
    std::vector<double> w(nq), phi((size_t)nq * nb), dq((size_t)nq * nb * 3);
    const double ref[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                              {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    if (dm->cell_type == PLX_HEX_Q1) {
        const double g = 0.57735026918962576451;
        for (int q = 0; q < 8; ++q) {
            const double xi[3] = {(q & 1) ? g : -g, (q & 2) ? g : -g, (q & 4) ? g : -g};
            w[q] = 1.0;
            for (int b = 0; b < 8; ++b) {
                const double a0 = 1 + ref[b][0] * xi[0], a1 = 1 + ref[b][1] * xi[1], a2 = 1 + ref[b][2] * xi[2];
                phi[(size_t)q * 8 + b] = 0.125 * a0 * a1 * a2;
                dq[((size_t)q * 8 + b) * 3 + 0] = 0.125 * ref[b][0] * a1 * a2;
                dq[((size_t)q * 8 + b) * 3 + 1] = 0.125 * a0 * ref[b][1] * a2;
                dq[((size_t)q * 8 + b) * 3 + 2] = 0.125 * a0 * a1 * ref[b][2];
            }
        }
    } else {
        P2Tables t;
        p2_fill_tables(&t);
        const int e[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
        for (int q = 0; q < nq; ++q) {
            const double *l = t.lam[q];
            w[q] = t.w[q];
            for (int b = 0; b < 10; ++b)
                for (int k = 0; k < 3; ++k) dq[((size_t)q * 10 + b) * 3 + k] = t.dphi[q][b][k];
            for (int v = 0; v < 4; ++v) phi[(size_t)q * 10 + v] = l[v] * (2.0 * l[v] - 1.0);
            for (int k = 0; k < 6; ++k) phi[(size_t)q * 10 + 4 + k] = 4.0 * l[e[k][0]] * l[e[k][1]];
        }
    }
    std::vector<double> load(3 * (size_t)dm->num_nodes, 0.0);
    std::vector<char> touched(dm->num_nodes, 0);
    for (int32_t cell : kp->cells) {
        const int32_t *cl = dm->closure.data() + (size_t)cell * nb;
        const int32_t *cn = dm->cone.data() + (size_t)cell * dm->nv;
        double X[8][3];
        for (int v = 0; v < dm->nv; ++v)
            for (int d = 0; d < 3; ++d) X[v][d] = dm->coords[3 * (size_t)cn[v] + d];
        const double *a = kp->aux_host.data() + (size_t)cell * per * naux_tot;
        for (int q = 0; q < nq; ++q) {
            double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (dm->cell_type == PLX_HEX_Q1) {
                for (int v = 0; v < 8; ++v)
                    for (int d = 0; d < 3; ++d)
                        for (int k = 0; k < 3; ++k) J[3 * d + k] += X[v][d] * dq[((size_t)q * 8 + v) * 3 + k];
            } else {
                for (int d = 0; d < 3; ++d)
                    for (int k = 0; k < 3; ++k) J[3 * d + k] = X[k + 1][d] - X[0][d];
            }
            double Ji[9];
            const double det = inv3(J, Ji);
            const double *aq = a + (per > 1 ? (size_t)q * naux_tot : 0);
            if (refstate) {
                // S = sigma_ref - C : eps_ref (vectors xx, yy, zz, xy, yz, xz)
                const double *rs = aq + off[i_rs], *re = aq + off[i_re];
                const double mu = aq[off[kp->num_aux - 2]], K = aq[off[kp->num_aux - 1]];
                const double tr = re[0] + re[1] + re[2];
                double S6[6];
                for (int i = 0; i < 3; ++i) S6[i] = rs[i] - K * tr - 2.0 * mu * (re[i] - tr / 3.0);
                for (int i = 3; i < 6; ++i) S6[i] = rs[i] - 2.0 * mu * re[i];
                const double S[3][3] = {{S6[0], S6[3], S6[5]}, {S6[3], S6[1], S6[4]}, {S6[5], S6[4], S6[2]}};
                const double wd = w[q] * det;
                for (int b = 0; b < nb; ++b) {
                    const double *dr = &dq[((size_t)q * nb + b) * 3];
                    double gp[3];
                    for (int d = 0; d < 3; ++d) gp[d] = dr[0] * Ji[0 * 3 + d] + dr[1] * Ji[1 * 3 + d] + dr[2] * Ji[2 * 3 + d];
                    for (int c = 0; c < 3; ++c)
                        load[3 * (size_t)cl[b] + c] -= wd * (S[c][0] * gp[0] + S[c][1] * gp[1] + S[c][2] * gp[2]);
                }
            }
            if (kp->f0 == PLX_FN_NONE) continue;
            double f[3] = {0.0, 0.0, 0.0};
            if (grav)
                for (int c = 0; c < 3; ++c) f[c] += aq[off[0]] * aq[off[i_grav] + c];
            if (body)
                for (int c = 0; c < 3; ++c) f[c] += aq[off[i_body] + c];
            const double wd = w[q] * det;
            for (int b = 0; b < nb; ++b)
                for (int c = 0; c < 3; ++c) load[3 * (size_t)cl[b] + c] += wd * phi[(size_t)q * nb + b] * f[c];
        }
        for (int b = 0; b < nb; ++b) touched[cl[b]] = 1;
    }
    std::vector<int32_t> nodes;
    for (int n = 0; n < dm->num_nodes; ++n)
        if (touched[n]) nodes.push_back(n);
    std::vector<double> v64(3 * nodes.size());
    std::vector<float> v32(3 * nodes.size());
    for (size_t i = 0; i < nodes.size(); ++i)
        for (int c = 0; c < 3; ++c) {
            v64[3 * i + c] = load[3 * (size_t)nodes[i] + c];
            v32[3 * i + c] = (float)v64[3 * i + c];
        }
    int err = 0;
    if ((err = upload(dm, &kp->d_load_node, nodes.data(), nodes.size()))) return err;
    if ((err = upload(dm, &kp->d_load64, v64.data(), v64.size()))) return err;
    if ((err = upload(dm, &kp->d_load32, v32.data(), v32.size()))) return err;
    kp->num_load = (int)nodes.size();
    return 0;
}

static int repack_aux(plx_dm dm, KeyPlan *kp) {
    const int num_aux = kp->num_aux;
    std::vector<int> off(num_aux);
    int naux_tot = 0;
    for (int i = 0; i < num_aux; ++i) {
        off[i] = naux_tot;
        naux_tot += kp->aux_sizes[i];
    }
    const int per = kp->aux_per_qp ? dm->nq : 1;
    const int n = (int)kp->cells.size();
    kp->aux_raw_valid = false;  // the PetscPointFn-form path re-uploads the original layout
    if (kp->f1 == PLX_FN_NONE && kp->jf3 == PLX_FN_NONE) {
        // kernels registered only as PetscPointFn-form launchers read the aux field as given
        kp->genmax = false;
        kp->has_aux = true;
        return 0;
    }
    const bool genmax = kp->f1 == PLX_F1V_GENMAXWELL_3D || kp->jf3 == PLX_JF3VU_GENMAXWELL_3D;
    if (genmax) {
        const int err = genmax_capture(dm, kp);
        return err ? err : build_load(dm, kp);
    }
    kp->genmax = false;
    const int i_mu = num_aux - 2;
    const int i_k = num_aux - 1;
    std::vector<double> m64(2 * (size_t)n);
    std::vector<float> m32(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        const double *a = kp->aux_host.data() + (size_t)kp->cells[i] * per * naux_tot;
        const double mu = a[off[i_mu]], K = a[off[i_k]];
        if (!(mu > 0.0) || !(K > 0.0)) return set_error(dm, 1, "non-positive shear or bulk modulus");
        for (int q = 1; q < per; ++q)
            if (a[q * naux_tot + off[i_mu]] != mu || a[q * naux_tot + off[i_k]] != K)
                return set_error(dm, 1, "moduli must be uniform within a cell");
        m64[i] = K - 2.0 / 3.0 * mu;  // lambda
        m64[n + i] = mu;
        m32[i] = (float)m64[i];
        m32[n + i] = (float)mu;
    }
    if (kp->d_mat64) (void)hipFree(kp->d_mat64);
    if (kp->d_mat32) (void)hipFree(kp->d_mat32);
    kp->d_mat64 = nullptr;
    kp->d_mat32 = nullptr;
    int err = 0;
    if ((err = upload(dm, &kp->d_mat64, m64.data(), m64.size()))) return err;
    if ((err = upload(dm, &kp->d_mat32, m32.data(), m32.size()))) return err;
    kp->has_aux = true;
    return build_load(dm, kp);
}

static int get_or_create_key(plx_dm dm, int label_value, int part, KeyPlan **out) {
    KeyPlan *kp = find_key(dm, label_value, part);
    if (!kp) {
        kp = new KeyPlan();
        kp->label_value = label_value;
        kp->part = part;
        const int err = build_plan(dm, kp);
        if (err) {
            free_key(kp);
            return err;
        }
        dm->keys.push_back(kp);
    }
    *out = kp;
    return 0;
}

int residual_key(plx_dm dm, int label_value, int part, const void *locX, void *locF, KeyPlan **out) {
    if (!dm || !locX || !locF) return set_error(dm, 1, "null vector");
    KeyPlan *kp = find_key(dm, label_value, part);
    if (!kp || (!kp->f1 && !kp->pf_residual)) return set_error(dm, 1, "no residual kernels registered for key");
    if (!kp->has_aux) return set_error(dm, 1, "auxiliary field not set");
    if (kp->pf_residual) {
        *out = kp;
        return 0;
    }
    if (kp->f1 != PLX_F1V_ISOLINEAR_ELASTICITY_3D && kp->f1 != PLX_F1V_ISOLINEAR_ELASTICITY_REFSTATE_3D &&
        kp->f1 != PLX_F1V_GENMAXWELL_3D)
        return set_error(dm, 1, "residual kernel not supported by this entry point");
    if ((kp->f1 == PLX_F1V_GENMAXWELL_3D) != kp->genmax)
        return set_error(dm, 1, "residual kernel and auxiliary layout disagree");
    if ((kp->f0 != PLX_FN_NONE || kp->f1 == PLX_F1V_ISOLINEAR_ELASTICITY_REFSTATE_3D) && kp->num_load == 0)
        return set_error(dm, 1, "body force / gravity / reference state: the auxiliary layout does not match the kernels");
    *out = kp;
    return 0;
}

static bool known_residual_pair(int f0, int f1) {
    return (f0 == PLX_FN_NONE || f0 == PLX_F0V_BODYFORCE || f0 == PLX_F0V_GRAV || f0 == PLX_F0V_GRAVBODYFORCE) &&
           (f1 == PLX_F1V_ISOLINEAR_ELASTICITY_3D || f1 == PLX_F1V_ISOLINEAR_ELASTICITY_REFSTATE_3D ||
            f1 == PLX_F1V_GENMAXWELL_3D);
}

}  // namespace plx

using namespace plx;

extern "C" {

int plx_dm_create(plx_dm *out, int cell_type, int num_cells, int num_vertices, int num_nodes,
                  const int32_t *closure, const int32_t *cone, const double *coords,
                  const int32_t *labels) {
    if (!out) return 1;
    *out = nullptr;
    if ((cell_type != PLX_HEX_Q1 && cell_type != PLX_TET_P2) || num_cells <= 0 ||
        num_vertices <= 0 || num_nodes <= 0 || !closure || !coords)
        return 1;
    plx_dm dm = new plx_dm_s();
    dm->cell_type = cell_type;
    dm->nb = cell_type == PLX_HEX_Q1 ? 8 : 10;
    dm->nv = cell_type == PLX_HEX_Q1 ? 8 : 4;
    dm->nq = cell_type == PLX_HEX_Q1 ? 8 : 14;
    dm->num_cells = num_cells;
    dm->num_vertices = num_vertices;
    dm->num_nodes = num_nodes;
    dm->local_size = 3 * (int64_t)num_nodes;
    dm->closure.assign(closure, closure + (size_t)num_cells * dm->nb);
    if (cone) {
        dm->cone.assign(cone, cone + (size_t)num_cells * dm->nv);
    } else {
        dm->cone.resize((size_t)num_cells * dm->nv);
        for (int c = 0; c < num_cells; ++c)
            for (int b = 0; b < dm->nv; ++b) dm->cone[(size_t)c * dm->nv + b] = closure[(size_t)c * dm->nb + b];
    }
    if (cell_type == PLX_HEX_Q1 && dm->cone != dm->closure) {
        delete dm;  // Q1 is isoparametric: the vertices are the nodes
        return 1;
    }
    for (int32_t n : dm->closure)
        if (n < 0 || n >= num_nodes) {
            delete dm;
            return 1;
        }
    for (int32_t v : dm->cone)
        if (v < 0 || v >= num_vertices) {
            delete dm;
            return 1;
        }
    dm->coords.assign(coords, coords + (size_t)num_vertices * 3);
    if (labels)
        dm->labels.assign(labels, labels + num_cells);
    else
        dm->labels.assign(num_cells, 0);
    if (hipMalloc(&dm->d_coords, dm->coords.size() * sizeof(double)) != hipSuccess ||
        hipMemcpy(dm->d_coords, dm->coords.data(), dm->coords.size() * sizeof(double),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&dm->d_scalar, 8 * sizeof(double)) != hipSuccess) {
        plx_dm_destroy(dm);
        return 76;
    }
    // element tables and the scratch address on the DM's device
    if (device_setup(dm)) {
        plx_dm_destroy(dm);
        return 76;
    }
    // no constrained dofs yet: the global vector is the local one (identity g2l map)
    if (plx_dm_set_constraints(dm, 0, nullptr)) {
        plx_dm_destroy(dm);
        return 76;
    }
    int bad = 0;
    const int verr = validate_geometry(dm, &bad);
    if (verr || bad) {
        plx_dm_destroy(dm);
        return verr ? verr : 2;
    }
    *out = dm;
    return 0;
}

int plx_dm_destroy(plx_dm dm) {
    if (!dm) return 0;
    events_release(dm);
    for (KeyPlan *k : dm->keys) free_key(k);
    if (dm->fault) free_fault(dm->fault);
    if (dm->d_coords) (void)hipFree(dm->d_coords);
    if (dm->d_closure) (void)hipFree(dm->d_closure);
    if (dm->d_cone) (void)hipFree(dm->d_cone);
    if (dm->d_scalar) (void)hipFree(dm->d_scalar);
    if (dm->d_jac_row_ptr) (void)hipFree(dm->d_jac_row_ptr);
    if (dm->d_jac_col) (void)hipFree(dm->d_jac_col);
    if (dm->d_partials) (void)hipFree(dm->d_partials);
    if (dm->con.d_dofs) (void)hipFree(dm->con.d_dofs);
    if (dm->con.d_g2l) (void)hipFree(dm->con.d_g2l);
    delete dm;
    return 0;
}

int plx_dm_set_stream(plx_dm dm, void *stream) {
    if (dm) dm->epoch++;
    if (!dm) return 1;
    dm->stream = (hipStream_t)stream;
    return 0;
}

const char *plx_dm_last_error(plx_dm dm) { return dm ? dm->last_error.c_str() : "null handle"; }

int plx_dm_set_kernel_options(plx_dm dm, int algorithm, int accumulate, int min_waves_per_simd) {
    if (dm) dm->epoch++;
    if (!dm) return 1;
    if (algorithm != PLX_ALGO_AUTO && algorithm != PLX_ALGO_QUADRATURE_POINTS &&
        algorithm != PLX_ALGO_FACTORED && algorithm != PLX_ALGO_POINTFN)
        return set_error(dm, 1, "unknown algorithm");
    if (accumulate != PLX_ACCUMULATE_AUTO && accumulate != PLX_ACCUMULATE_COLORED &&
        accumulate != PLX_ACCUMULATE_LDS_ATOMIC)
        return set_error(dm, 1, "unknown accumulation mode");
    if (min_waves_per_simd < 0 || min_waves_per_simd > 4) return set_error(dm, 1, "min waves must be 0..4");
    dm->algorithm = algorithm;
    dm->accumulate = accumulate;
    dm->min_waves = min_waves_per_simd <= 2 ? 2 : min_waves_per_simd;
    return 0;
}

int plx_dm_get_sizes(plx_dm dm, int64_t *local_size, int64_t *global_size) {
    if (!dm) return 1;
    if (local_size) *local_size = dm->local_size;
    if (global_size) *global_size = dm->con.global_size;
    return 0;
}

int plx_dm_set_constraints(plx_dm dm, int num_constrained, const int32_t *dofs) {
    if (dm) dm->epoch++, dm->op_epoch++;
    if (!dm || num_constrained < 0 || (num_constrained > 0 && !dofs)) return set_error(dm, 1, "bad constraint arguments");
    std::vector<char> is_con(dm->local_size, 0);
    for (int i = 0; i < num_constrained; ++i) {
        if (dofs[i] < 0 || dofs[i] >= dm->local_size) return set_error(dm, 1, "constrained dof out of range");
        if (is_con[dofs[i]]) return set_error(dm, 1, "constrained dof listed twice");
        is_con[dofs[i]] = 1;
    }
    std::vector<int32_t> g2l;
    g2l.reserve(dm->local_size - num_constrained);
    for (int64_t i = 0; i < dm->local_size; ++i)
        if (!is_con[i]) g2l.push_back((int32_t)i);
    if (dm->con.d_dofs) (void)hipFree(dm->con.d_dofs);
    if (dm->con.d_g2l) (void)hipFree(dm->con.d_g2l);
    dm->con.d_dofs = nullptr;
    dm->con.d_g2l = nullptr;
    dm->con.dofs.assign(dofs, dofs + num_constrained);
    dm->con.global_size = (int64_t)g2l.size();
    int err = 0;
    if ((err = upload(dm, &dm->con.d_dofs, dm->con.dofs.data(), dm->con.dofs.size()))) return err;
    if ((err = upload(dm, &dm->con.d_g2l, g2l.data(), g2l.size()))) return err;
    return 0;
}

int plx_dm_add_residual(plx_dm dm, int label_value, int part, int f0, int f1) {
    if (dm) dm->epoch++, dm->op_epoch++;
    if (!dm) return 1;
    if (!known_residual_pair(f0, f1)) return set_error(dm, 1, "no device twin for this residual kernel pair");
    KeyPlan *kp = nullptr;
    int err = get_or_create_key(dm, label_value, part, &kp);
    if (err) return err;
    kp->f0 = f0;
    kp->f1 = f1;
    return build_load(dm, kp);
}

int plx_dm_add_jacobian(plx_dm dm, int label_value, int part, int jf0, int jf1, int jf2, int jf3) {
    if (dm) dm->epoch++, dm->op_epoch++;
    if (!dm) return 1;
    if (jf0 || jf1 || jf2 ||
        (jf3 != PLX_JF3VU_ISOLINEAR_ELASTICITY_3D && jf3 != PLX_JF3VU_GENMAXWELL_3D))
        return set_error(dm, 1, "no device twin for this Jacobian kernel set");
    KeyPlan *kp = nullptr;
    int err = get_or_create_key(dm, label_value, part, &kp);
    if (err) return err;
    kp->jf0 = jf0;
    kp->jf1 = jf1;
    kp->jf2 = jf2;
    kp->jf3 = jf3;
    return 0;
}

int plx_dm_create_ordered(plx_dm *out, int cell_type, int num_cells, int num_vertices, int num_nodes,
                          const int32_t *closure, const int32_t *cone, const double *coords, const int32_t *labels,
                          const int32_t *vertex_perm) {
    if (!vertex_perm) return plx_dm_create(out, cell_type, num_cells, num_vertices, num_nodes, closure, cone, coords, labels);
    if (!out || cell_type != PLX_HEX_Q1 || num_cells <= 0 || !closure) return 1;
    std::vector<char> seen(8, 0);
    for (int i = 0; i < 8; ++i) {
        if (vertex_perm[i] < 0 || vertex_perm[i] > 7 || seen[vertex_perm[i]]) return 1;
        seen[vertex_perm[i]] = 1;
    }
    // our vertex i of every cell is the caller's vertex vertex_perm[i] of that cell
    std::vector<int32_t> cl((size_t)num_cells * 8);
    for (int c = 0; c < num_cells; ++c)
        for (int i = 0; i < 8; ++i) cl[(size_t)c * 8 + i] = closure[(size_t)c * 8 + vertex_perm[i]];
    return plx_dm_create(out, cell_type, num_cells, num_vertices, num_nodes, cl.data(), cl.data(), coords, labels);
}

int plx_dm_set_point_order(plx_dm dm, int num_quad, const int32_t *quad_perm) {
    if (!dm) return 1;
    if (!dm->keys.empty()) return set_error(dm, 1, "set the point order before registering kernels");
    if (quad_perm) {
        std::vector<char> seen(dm->nq, 0);
        if (num_quad != dm->nq) return set_error(dm, 1, "bad quadrature-point permutation");
        for (int i = 0; i < num_quad; ++i) {
            if (quad_perm[i] < 0 || quad_perm[i] >= num_quad || seen[quad_perm[i]])
                return set_error(dm, 1, "bad quadrature-point permutation");
            seen[quad_perm[i]] = 1;
        }
        dm->quad_perm.assign(quad_perm, quad_perm + num_quad);
    } else {
        dm->quad_perm.clear();
    }
    dm->epoch++, dm->op_epoch++;
    return 0;
}

int plx_dm_add_residual_pointfn(plx_dm dm, int label_value, int part, plx_pointfn_launcher residual) {
    if (dm) dm->epoch++, dm->op_epoch++;
    if (!dm || !residual) return set_error(dm, 1, "null launcher");
    KeyPlan *kp = nullptr;
    const int err = get_or_create_key(dm, label_value, part, &kp);
    if (err) return err;
    kp->pf_residual = residual;
    return 0;
}

int plx_dm_add_jacobian_pointfn(plx_dm dm, int label_value, int part, plx_pointfn_launcher action,
                                plx_pointfn_launcher assemble) {
    if (dm) dm->epoch++, dm->op_epoch++;
    if (!dm || (!action && !assemble)) return set_error(dm, 1, "null launchers");
    KeyPlan *kp = nullptr;
    const int err = get_or_create_key(dm, label_value, part, &kp);
    if (err) return err;
    kp->pf_action = action;
    kp->pf_assemble = assemble;
    return 0;
}

int plx_dm_set_constants(plx_dm dm, int label_value, int part, int n, const double *constants) {
    if (dm) dm->epoch++, dm->op_epoch++;
    if (!dm || n < 0 || (n > 0 && !constants)) return 1;
    KeyPlan *kp = find_key(dm, label_value, part);
    if (!kp) return set_error(dm, 1, "unknown key");
    kp->constants.assign(constants, constants + n);
    return genmax_refresh(dm, kp);
}

int plx_dm_set_auxiliary(plx_dm dm, int label_value, int part, int num_aux, const int32_t *aux_sizes,
                         int aux_per_qp, const double *aux) {
    if (!dm || num_aux < 2 || !aux_sizes || !aux) return set_error(dm, 1, "bad auxiliary arguments");
    dm->epoch++, dm->op_epoch++;
    KeyPlan *kp = find_key(dm, label_value, part);
    if (!kp) return set_error(dm, 1, "unknown key");
    int naux_tot = 0;
    for (int i = 0; i < num_aux; ++i) {
        if (aux_sizes[i] <= 0) return set_error(dm, 1, "bad aux subfield size");
        naux_tot += aux_sizes[i];
    }
    const int per = aux_per_qp ? dm->nq : 1;
    kp->num_aux = num_aux;
    kp->aux_sizes.assign(aux_sizes, aux_sizes + num_aux);
    kp->aux_per_qp = aux_per_qp;
    kp->aux_host.assign(aux, aux + (size_t)dm->num_cells * per * naux_tot);
    if (aux_per_qp && !dm->quad_perm.empty())  // caller's point order -> ours
        for (int c = 0; c < dm->num_cells; ++c)
            for (int q = 0; q < per; ++q)
                std::copy(aux + ((size_t)c * per + dm->quad_perm[q]) * naux_tot,
                          aux + ((size_t)c * per + dm->quad_perm[q] + 1) * naux_tot,
                          kp->aux_host.begin() + ((size_t)c * per + q) * naux_tot);
    return repack_aux(dm, kp);
}

int plx_compute_residual(plx_dm dm, int label_value, int part, int scalar_type, double t,
                         const void *locX, const void *locX_t, void *locF, int mode) {
    (void)t;
    (void)locX_t;
    KeyPlan *kp = nullptr;
    if (const int e = residual_key(dm, label_value, part, locX, locF, &kp)) return e;
    if (kp->pf_residual || (dm->algorithm == PLX_ALGO_POINTFN && pointfn_builtin_residual(kp->f0, kp->f1))) {
        if (scalar_type != PLX_FP64) return set_error(dm, 1, "the PetscPointFn-form path is FP64");
        ScopedEvent ev(dm, part == PLX_RHS ? "PL:IntegratorDomain:computeRHSResidual" : "PL:IntegratorDomain:computeLHSResidual");
        return pointfn_residual(dm, kp, t, locX, locX_t, locF, mode);
    }
    // the event of the reference's integrator around the same call (IntegratorDomain.cc:80-81)
    ScopedEvent ev(dm, part == PLX_RHS ? "PL:IntegratorDomain:computeRHSResidual" : "PL:IntegratorDomain:computeLHSResidual");
    const int err = launch_apply(dm, kp, scalar_type, false, locX, locF, mode);
    if (err || kp->num_load == 0) return err;
    return launch_add_load(dm, kp, scalar_type, locF);
}

int plx_compute_jacobian_action(plx_dm dm, int label_value, int part, int scalar_type, double t,
                                double s_tshift, const void *locX, const void *locX_t,
                                const void *locY, void *locF, int mode) {
    (void)t;
    (void)s_tshift;
    (void)locX;
    (void)locX_t;
    if (!dm || !locY || !locF) return set_error(dm, 1, "null vector");
    KeyPlan *kp = find_key(dm, label_value, part);
    if (!kp || (!kp->jf3 && !kp->pf_action)) return set_error(dm, 1, "no Jacobian kernels registered for key");
    if (!kp->has_aux) return set_error(dm, 1, "auxiliary field not set");
    if (kp->pf_action || (dm->algorithm == PLX_ALGO_POINTFN && kp->jf3 == PLX_JF3VU_ISOLINEAR_ELASTICITY_3D)) {
        if (scalar_type != PLX_FP64) return set_error(dm, 1, "the PetscPointFn-form path is FP64");
        return pointfn_action(dm, kp, t, s_tshift, locX, locX_t, locY, locF, mode);
    }
    if ((kp->jf3 == PLX_JF3VU_GENMAXWELL_3D) != kp->genmax)
        return set_error(dm, 1, "Jacobian kernel and auxiliary layout disagree");
    return launch_apply(dm, kp, scalar_type, true, locY, locF, mode);
}

int plx_compute_jacobian_pbdiag(plx_dm dm, int label_value, int part, double t, double s_tshift,
                                const double *locX, double *blocks) {
    (void)t;
    (void)s_tshift;
    (void)locX;
    if (!dm || !blocks) return set_error(dm, 1, "null vector");
    KeyPlan *kp = find_key(dm, label_value, part);
    if (!kp || !kp->jf3) return set_error(dm, 1, "no Jacobian kernels registered for key");
    if (!kp->has_aux) return set_error(dm, 1, "auxiliary field not set");
    return launch_pbdiag(dm, kp, blocks);
}

int plx_local_to_global(plx_dm dm, int scalar_type, const void *loc, void *glob, int mode) {
    if (!dm || !loc || !glob) return set_error(dm, 1, "null vector");
    return launch_l2g(dm, scalar_type, loc, glob, mode);
}

int plx_global_to_local(plx_dm dm, int scalar_type, const void *glob, void *loc) {
    if (!dm || !loc || !glob) return set_error(dm, 1, "null vector");
    return launch_g2l(dm, scalar_type, glob, loc);
}

int plx_insert_boundary_values(plx_dm dm, int scalar_type, const void *values, void *loc) {
    if (!dm || !loc || (!values && !dm->con.dofs.empty())) return set_error(dm, 1, "null vector");
    return launch_insert_bc(dm, scalar_type, values, loc);
}

int plx_vec_dot(plx_dm dm, int scalar_type, int64_t n, const void *x, const void *y, double *result) {
    if (!dm || !x || !y || !result || n < 0) return set_error(dm, 1, "bad dot arguments");
    return launch_dot(dm, scalar_type, n, x, y, result);
}

int plx_vec_norm2(plx_dm dm, int scalar_type, int64_t n, const void *x, double *result) {
    if (!dm || !x || !result || n < 0) return set_error(dm, 1, "bad norm arguments");
    double d = 0.0;
    const int err = launch_dot(dm, scalar_type, n, x, x, &d);
    if (err) return err;
    *result = std::sqrt(d);
    return 0;
}

int plx_debug_plan_info(plx_dm dm, int label_value, int part, int *num_tiles, int64_t *halo_slots,
                        int *reordered) {
    KeyPlan *kp = dm ? find_key(dm, label_value, part) : nullptr;
    if (!kp) return set_error(dm, 1, "unknown key");
    if (num_tiles) *num_tiles = kp->num_tiles;
    if (halo_slots) *halo_slots = kp->halo_slots;
    if (reordered) *reordered = kp->reordered ? 1 : 0;
    return 0;
}

}  // extern "C"
