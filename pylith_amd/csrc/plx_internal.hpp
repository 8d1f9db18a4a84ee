// plx_internal.hpp -- device data layout of the MI355X assembly path (private to the library).
//
// A DM holds the mesh once on the device.  Each weak-form key (label value, equation part)
// owns a tile plan: the key's cells cut into tiles of spatially adjacent cells, one tile per
// workgroup.  A tile lists the nodes its cells touch; nodes touched by no other tile are
// "owned" and written straight to the output vector, the rest are "shared" and leave the
// tile as partial sums in a halo buffer that a second, node-ordered pass reduces in a fixed
// order.  Local closures are 16-bit indices into the tile's node list so the cell kernels
// gather from and scatter into LDS only.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <cstdint>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/pylith_amd.h"
#include "plx_plan.hpp"

namespace plx {

constexpr int kMaxQuad = 14;

// Device-side arrays of one key's tile plan (passed by value to kernels).
struct TilePlanDev {
    int num_tiles;
    int stride;                      // SoA stride of per-cell arrays (= number of key cells)
    const int32_t *tile_cell_ptr;    // [num_tiles+1] cell positions
    const int32_t *tile_node_ptr;    // [num_tiles+1]
    const int32_t *tile_nown;        // [num_tiles] owned nodes lead each tile's node list
    const int32_t *tile_nodes;       // global node ids
    const int32_t *tile_vert_ptr;    // [num_tiles+1]
    const int32_t *tile_verts;       // global vertex ids
    const int32_t *tile_halo_off;    // [num_tiles] first halo slot of the tile's shared nodes
    const uint16_t *lclosure;        // [nb][stride] tile-local node index
    const uint16_t *lcone;           // [nv][stride] tile-local vertex index
    const double *coords;            // [num_vertices][3]
};

struct KeyPlan {
    int label_value = 0;
    int part = 0;
    int f0 = 0, f1 = 0;
    int jf0 = 0, jf1 = 0, jf2 = 0, jf3 = 0;
    std::vector<int32_t> cells;      // key cells, tile order
    int num_tiles = 0;
    int max_nodes = 0, max_verts = 0;
    int64_t halo_slots = 0;
    int num_shared = 0;
    bool reordered = false;          // the planner took the cells in Morton order
    // device arrays
    int32_t *d_cells = nullptr;
    int32_t *d_tile_cell_ptr = nullptr, *d_tile_node_ptr = nullptr, *d_tile_nown = nullptr;
    int32_t *d_tile_nodes = nullptr, *d_tile_vert_ptr = nullptr, *d_tile_verts = nullptr;
    int32_t *d_tile_halo_off = nullptr;
    uint16_t *d_lclosure = nullptr, *d_lcone = nullptr;
    uint16_t *d_records = nullptr;   // [ncells][8 or 16] packed local closure (+ cone) per cell
    uint8_t *d_color = nullptr;      // [ncells] colour within the tile (cells sorted by colour)
    int32_t *d_tile_ncolor = nullptr;
    int max_colors = 0;
    bool sorted_by_color = true;
    int32_t *d_tiles_affine = nullptr, *d_tiles_general = nullptr;  // Q1: tiles of affine cells / others
    int num_tiles_affine = 0, num_tiles_general = 0;
    int32_t *d_shared_node = nullptr, *d_shared_ptr = nullptr, *d_shared_slot = nullptr;
    void *d_halo = nullptr;          // halo partial sums, sized for FP64
    // auxiliary field, repacked per key cell (SoA, tile order)
    int num_aux = 0;
    std::vector<int32_t> aux_sizes;
    int aux_per_qp = 0;
    bool has_aux = false;
    double *d_mat64 = nullptr;       // [2][stride]: lambda, mu (elastic part)
    float *d_mat32 = nullptr;        // same in FP32
    // generalized Maxwell (Q1, aux per quadrature point): per-cell material, per-point state and
    // the per-point state stress of the current time step, SoA over key cells (tile order)
    bool genmax = false;
    double *d_gm_mat = nullptr;      // [8][stride]: mu, K, maxwell_time[3], shear_modulus_ratio[3]
    double *d_state = nullptr;       // [24][8][stride]: viscous_strain[3][6], total_strain[6]
    double *d_prestress = nullptr;   // [6][8][stride]
    std::vector<double> aux_host;    // captured aux (original layout) for state kernels
    std::vector<double> constants;
    // f0 load term (body force / gravity: independent of the solution), added to the residual
    // after the cell pass: the key's closure nodes and their values (FP64 and FP32 copies)
    int num_load = 0;
    int32_t *d_load_node = nullptr;
    double *d_load64 = nullptr;
    float *d_load32 = nullptr;
    // PetscPointFn-form path: launchers registered by the caller (plx_dm_add_*_pointfn), the
    // key's cells by global colour (cells of one colour share no node), the aux field in its
    // original layout and the constants on the device
    plx_pointfn_launcher pf_residual = nullptr, pf_action = nullptr, pf_assemble = nullptr;
    std::vector<int32_t> pf_color_ptr;
    int32_t *d_pf_cells = nullptr;
    double *d_aux_raw = nullptr;
    bool aux_raw_valid = false;
    double *d_constants = nullptr;
    int num_constants_dev = -1;
};

struct EventLog {
    struct Pending {
        std::string name;
        hipEvent_t begin, end;
    };
    bool on = false;
    std::vector<Pending> pending;
    std::map<std::string, std::pair<int, double>> stats;  // name -> (count, total ms)
};

struct Constraints {
    std::vector<int32_t> dofs;       // constrained local dofs (caller order)
    int32_t *d_dofs = nullptr;
    int32_t *d_g2l = nullptr;        // [global] local index
    int64_t global_size = 0;
};

// Cohesive cells of one fault with prescribed slip (FaultCohesiveKin).  Each cohesive cell is a
// quadrilateral face carrying a negative-side, a positive-side and a Lagrange-multiplier node at
// each of its four vertices.  The Lagrange-multiplier field follows the displacement field in
// the local vector: dof lag_offset + 3*l + k.
struct FaultData {
    int num_cohesive = 0, num_lag = 0;
    int64_t lag_offset = 0;
    double ref_dir_1[3] = {0, 0, 1}, ref_dir_2[3] = {0, 1, 0};
    std::vector<int32_t> color_ptr;        // cells of colour k: color_cells[color_ptr[k]..]
    std::vector<int32_t> neg_node, pos_node;  // [num_lag] split-node pair of each fault vertex
    int32_t *d_neg = nullptr, *d_pos = nullptr, *d_lag = nullptr, *d_cone = nullptr;  // [ncoh][4]
    int32_t *d_color_cells = nullptr;
    int32_t *d_neg_node = nullptr, *d_pos_node = nullptr;
    double *d_slip = nullptr;              // [num_lag][3] in the fault frame (normal, t1, t2)
    bool planar = false;                   // one fault frame everywhere: jump = R slip exactly
    double R[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // columns n, t1, t2 (row-major 3x3)
    // scalar fault-face mass matrix M_ab = integral psi_a psi_b over the fault (CSR over Lagrange
    // nodes, the same 2x2 Gauss rule as the cohesive kernels) and its inverse diagonal
    int32_t *d_mass_ptr = nullptr, *d_mass_col = nullptr;
    double *d_mass_val = nullptr, *d_mass_dinv = nullptr;
};

}  // namespace plx

struct plx_dm_s {
    int cell_type = 0;
    int nb = 0, nv = 0, nq = 0;
    int num_cells = 0, num_vertices = 0, num_nodes = 0;
    int64_t local_size = 0;
    std::vector<int32_t> closure, cone, labels;
    // caller's quadrature-point order: our point q is the caller's point quad_perm[q] (empty:
    // the same order); applied to per-point aux on the way in and out
    std::vector<int32_t> quad_perm;
    std::vector<double> coords;
    double *d_coords = nullptr;
    int32_t *d_closure = nullptr, *d_cone = nullptr;  // uploaded on first use (set-up kernels)
    hipStream_t stream = nullptr;
    std::vector<plx::KeyPlan *> keys;
    plx::Constraints con;
    plx::FaultData *fault = nullptr;
    // assembled-Jacobian pattern (3x3 blocks over the mesh nodes), built on first use
    std::vector<int32_t> jac_row_ptr, jac_col;
    int32_t *d_jac_row_ptr = nullptr, *d_jac_col = nullptr;
    double *d_scalar = nullptr;      // reduction scratch
    double *d_partials = nullptr;
    int partials_len = 0;
    plx::EventLog events;
    int algorithm = PLX_ALGO_AUTO;
    int accumulate = PLX_ACCUMULATE_AUTO;
    int min_waves = 2;
    int wg_per_cu = 2;               // resident workgroups per CU of the persistent cell kernels
    int num_cu = 0;                  // compute units of the device (queried once)
    int device = 0;                  // HIP device current at plx_dm_create (the DM's buffers live there)
    // bumped by every call that changes device state or launch choices (buffers, kernels,
    // options, stream): captured graphs recorded at another epoch are stale
    uint64_t epoch = 0;
    // bumped by every call that changes the operator itself (kernels, aux field, constants,
    // constraints, fault): an assembled copy of the operator made at another value is stale
    uint64_t op_epoch = 0;
    std::string last_error;
};

namespace plx {

// Error plumbing: every C-ABI function funnels failures through set_error.
int set_error(plx_dm dm, int code, const std::string &msg);

#define PLX_HIP(dm, call)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return plx::set_error(dm, 76, std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

KeyPlan *find_key(plx_dm dm, int label_value, int part);
// The key of a residual call after the argument checks of plx_compute_residual.
int residual_key(plx_dm dm, int label_value, int part, const void *locX, void *locF, KeyPlan **out);

// Records a named hipEvent pair around a launch when the handle's event logging is on.
class ScopedEvent {
  public:
    ScopedEvent(plx_dm dm, const char *name);
    ~ScopedEvent();

  private:
    plx_dm dm_;
    const char *name_;
    hipEvent_t a_ = nullptr;
    bool active_ = false;
};
void events_release(plx_dm dm);

// PetscPointFn-form path (plx_pointfn.hip): residual / Jacobian action of a key through the
// generic quadrature-point launchers, FP64.
int pointfn_residual(plx_dm dm, KeyPlan *kp, double t, const void *u, const void *u_t, void *f, int mode);
int pointfn_action(plx_dm dm, KeyPlan *kp, double t, double s_tshift, const void *u, const void *u_t,
                   const void *v, void *f, int mode);
bool pointfn_builtin_residual(int f0, int f1);
void pointfn_release(plx_dm dm, KeyPlan *kp);

// Kernel launchers (plx_assembly.hip).
// halo: the shared-node partial-sum buffer (nullptr: the key's own d_halo); log: record the
// kernels' event timers when the handle's logging is on
int launch_apply(plx_dm dm, KeyPlan *kp, int scalar_type, bool jacobian, const void *u,
                 void *r, int mode, double scale = 1.0, void *halo = nullptr, bool log = true);
// Per-device set-up at plx_dm_create with the DM's device current: the element tables in
// constant memory and the device address of the allocation-free entry point's scratch array.
int device_setup(plx_dm dm);
int launch_pbdiag(plx_dm dm, KeyPlan *kp, double *blocks);
int launch_add_load(plx_dm dm, KeyPlan *kp, int scalar_type, void *r);
int ensure_mesh_indices(plx_dm dm);  // uploads d_closure and d_cone (each if absent)
// Generalized Maxwell state (plx_genmax.hip).
int launch_genmax_prestress(plx_dm dm, KeyPlan *kp);
int launch_genmax_update(plx_dm dm, KeyPlan *kp, const double *u);
int genmax_capture(plx_dm dm, KeyPlan *kp);
int genmax_refresh(plx_dm dm, KeyPlan *kp);
int validate_geometry(plx_dm dm, int *num_bad);

// Cohesive cells and split-node maps (plx_fault.hip).
int launch_cohesive(plx_dm dm, bool jacobian, const double *x, double *f, int mode);
enum { kFold = 0, kUnfold = 1, kClearPos = 2 };
int launch_fold(plx_dm dm, int scalar_type, void *v, int op);
int launch_gather_pos(plx_dm dm, const double *v, double *out);
int launch_sync_pos(plx_dm dm, const double *jump, double *v);
int launch_planar_jump(plx_dm dm, int64_t count, const double *slip, double *jump);
int launch_fault_mass(plx_dm dm, const double *x, double *y);
int launch_fault_jacobi(plx_dm dm, const double *r, double *z);
int launch_fault_slip_rhs(plx_dm dm, double *lag_out);
void free_fault(FaultData *f);

// Multigrid V-cycle of a Green's-function handle on one column (plx_gf.hip), r and z in DM
// node order.  gf_set_fine_operator gives it a matrix-free fine-level operator y = A x (FP32,
// DM order, queued on the V-cycle's stream, constrained rows zero): the fine level then smooths
// with it instead of the assembled operator (null: back to the assembled one).
int gf_vcycle1(plx_gf g, hipStream_t stream, const float *r, float *z);
using GfFineOp = int (*)(void *ctx, const float *x, float *y);
int gf_set_fine_operator(plx_gf g, GfFineOp op, void *ctx);

// Vector kernels (plx_vector.hip).
int launch_l2g(plx_dm dm, int scalar_type, const void *loc, void *glob, int mode);
int launch_g2l(plx_dm dm, int scalar_type, const void *glob, void *loc);
int launch_insert_bc(plx_dm dm, int scalar_type, const void *values, void *loc);
int launch_dot(plx_dm dm, int scalar_type, int64_t n, const void *x, const void *y,
               double *result);

}  // namespace plx
