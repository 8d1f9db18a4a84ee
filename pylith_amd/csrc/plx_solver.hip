// plx_solver.hip -- Krylov solves around the matrix-free operator, on the device.
//
// The reference solves J du = -F with PETSc KSP (GMRES + GAMG, vpbjacobi fine smoother;
// materials/Elasticity.cc:307-318, problems/GreensFns.cc:236-256).  Here the operator is the
// matrix-free Jacobian action A = -J (symmetric positive definite for elasticity once the
// constrained dofs are fixed), preconditioned by the inverses of its 3x3 point blocks.
//
//   * PLX_KSP_CG:        plain PCG in FP64 (the FP64 reference path);
//   * PLX_KSP_FGMRES_IR: FGMRES in FP64 on the true residual, each preconditioner application
//                        an FP32 PCG solve (iterative refinement in a flexible Krylov loop);
//   * PLX_KSP_IR:        plain iterative refinement (Richardson) with the FP32 PCG correction.
// Vectors are local vectors; constrained dofs are held fixed (the operator is masked there).
// Scalars of the PCG recurrences stay on the device, so an iteration needs no host round trip.
//
// A fault with prescribed slip (cohesive cells) turns the system into a saddle point
// [K  B^T; B  0] [u; lambda] = [b_u; g - b_lambda].  It is solved by eliminating the constraint:
// B u = M (u+ - u-) = g fixes the jump u+ - u- = M^-1 (g - b_lambda) across every split node pair
// (M the fault-face mass matrix, one small PCG), so u = U w + u_jump with U copying each negative
// node's value to its positive twin.  U^T K U is symmetric positive definite and is what the
// Krylov methods above see (corrections w carry zeros at the positive nodes; "fold" is U^T).
// Finally the positive-side rows give M lambda = (A u)_+ - b_+.
#include <array>
#include <map>
#include <vector>

#include "plx_element.hpp"
#include "plx_internal.hpp"

namespace plx {

constexpr int kRedBlocks = 512;
enum { S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_RZNEW, S_COUNT };

template <typename Real>
__global__ void __launch_bounds__(256) mask_kernel(int n, const int32_t *__restrict__ dofs, Real *v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[dofs[i]] = Real(0);
}

// Block-reduce helper (256 threads): returns the block sum in thread 0.
__device__ __forceinline__ double block_sum(double v, double *s) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    return s[0];
}

template <typename Real>
__global__ void __launch_bounds__(256) dot_partials(int64_t n, const Real *__restrict__ a,
                                                    const Real *__restrict__ b, double *partials) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        acc += (double)a[i] * (double)b[i];
    const double t = block_sum(acc, s);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// scal[S_PQ] = sum(partials); scal[S_ALPHA] = scal[S_RZ] / scal[S_PQ]
__global__ void __launch_bounds__(256) finish_alpha(const double *__restrict__ partials, double *scal) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += 256) acc += partials[i];
    const double t = block_sum(acc, s);
    if (threadIdx.x == 0) {
        scal[S_PQ] = t;
        scal[S_ALPHA] = t != 0.0 ? scal[S_RZ] / t : 0.0;
    }
}

// x += alpha p; r -= alpha q; z = Binv r (per node); partial sums of r.z and r.r
template <typename Real>
__global__ void __launch_bounds__(256) cg_update(int nnodes, const Real *__restrict__ Binv, Real *x,
                                                 Real *r, const Real *__restrict__ p,
                                                 const Real *__restrict__ q, Real *z,
                                                 const double *__restrict__ scal, double *part_rz,
                                                 double *part_rr) {
    __shared__ double s[256];
    const Real alpha = (Real)scal[S_ALPHA];
    double rz = 0.0, rr = 0.0;
    for (int node = blockIdx.x * 256 + threadIdx.x; node < nnodes; node += gridDim.x * 256) {
        const size_t o = 3 * (size_t)node;
        Real rv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x[o + c] += alpha * p[o + c];
            rv[c] = r[o + c] - alpha * q[o + c];
            r[o + c] = rv[c];
        }
        const Real *B = Binv + 9 * (size_t)node;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const Real zc = B[3 * c] * rv[0] + B[3 * c + 1] * rv[1] + B[3 * c + 2] * rv[2];
            z[o + c] = zc;
            rz += (double)rv[c] * (double)zc;
            rr += (double)rv[c] * (double)rv[c];
        }
    }
    const double a = block_sum(rz, s);
    __syncthreads();
    const double b = block_sum(rr, s);
    if (threadIdx.x == 0) {
        part_rz[blockIdx.x] = a;
        part_rr[blockIdx.x] = b;
    }
}

// beta = rz_new / rz; rz = rz_new; rr
__global__ void __launch_bounds__(256) finish_beta(const double *__restrict__ part_rz,
                                                   const double *__restrict__ part_rr, double *scal) {
    __shared__ double s[256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += 256) {
        a += part_rz[i];
        b += part_rr[i];
    }
    const double rz = block_sum(a, s);
    __syncthreads();
    const double rr = block_sum(b, s);
    if (threadIdx.x == 0) {
        const double old = scal[S_RZ];
        scal[S_BETA] = old != 0.0 ? rz / old : 0.0;
        scal[S_RZ] = rz;
        scal[S_RR] = rr;
    }
}

template <typename Real>
__global__ void __launch_bounds__(256) p_update(int64_t n, const Real *__restrict__ z, Real *p,
                                                const double *__restrict__ scal) {
    const Real beta = (Real)scal[S_BETA];
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        p[i] = z[i] + beta * p[i];
}

// Multigrid-preconditioned PCG pieces (the preconditioner is applied between kernels):
// x += alpha p; r -= alpha q; partials of r.r
template <typename Real>
__global__ void __launch_bounds__(256) cg_update_xr(int64_t n, Real *x, Real *r, const Real *__restrict__ p,
                                                    const Real *__restrict__ q, const double *__restrict__ scal,
                                                    double *part_rr) {
    __shared__ double s[256];
    const Real alpha = (Real)scal[S_ALPHA];
    double rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        x[i] += alpha * p[i];
        const Real rv = r[i] - alpha * q[i];
        r[i] = rv;
        rr += (double)rv * (double)rv;
    }
    const double b = block_sum(rr, s);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = b;
}

// partials of r.z and (when part_rr is set) r.r; p = z when p is set
template <typename Real>
__global__ void __launch_bounds__(256) rz_partials(int64_t n, const Real *__restrict__ r, const Real *__restrict__ z,
                                                   Real *p, double *part_rz, double *part_rr) {
    __shared__ double s[256];
    double rz = 0.0, rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double rv = (double)r[i], zv = (double)z[i];
        rz += rv * zv;
        rr += rv * rv;
        if (p) p[i] = z[i];
    }
    const double a = block_sum(rz, s);
    __syncthreads();
    const double b = block_sum(rr, s);
    if (threadIdx.x == 0) {
        part_rz[blockIdx.x] = a;
        if (part_rr) part_rr[blockIdx.x] = b;
    }
}

// z = Binv r, p = z, partials of r.z and r.r (start of PCG)
template <typename Real>
__global__ void __launch_bounds__(256) pc_start(int nnodes, const Real *__restrict__ Binv,
                                                const Real *__restrict__ r, Real *z, Real *p,
                                                double *part_rz, double *part_rr) {
    __shared__ double s[256];
    double rz = 0.0, rr = 0.0;
    for (int node = blockIdx.x * 256 + threadIdx.x; node < nnodes; node += gridDim.x * 256) {
        const size_t o = 3 * (size_t)node;
        const Real *B = Binv + 9 * (size_t)node;
        const Real r0 = r[o], r1 = r[o + 1], r2 = r[o + 2];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const Real zc = B[3 * c] * r0 + B[3 * c + 1] * r1 + B[3 * c + 2] * r2;
            z[o + c] = zc;
            p[o + c] = zc;
            const Real rc = c == 0 ? r0 : c == 1 ? r1 : r2;
            rz += (double)rc * (double)zc;
            rr += (double)rc * (double)rc;
        }
    }
    const double a = block_sum(rz, s);
    __syncthreads();
    const double b = block_sum(rr, s);
    if (threadIdx.x == 0) {
        part_rz[blockIdx.x] = a;
        part_rr[blockIdx.x] = b;
    }
}

__global__ void __launch_bounds__(256) finish_start(const double *__restrict__ part_rz,
                                                    const double *__restrict__ part_rr, double *scal) {
    __shared__ double s[256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += 256) {
        a += part_rz[i];
        b += part_rr[i];
    }
    const double rz = block_sum(a, s);
    __syncthreads();
    const double rr = block_sum(b, s);
    if (threadIdx.x == 0) {
        scal[S_RZ] = rz;
        scal[S_RR] = rr;
        scal[S_BETA] = 0.0;
    }
}

template <typename Src, typename Dst>
__global__ void __launch_bounds__(256) convert_kernel(int64_t n, const Src *__restrict__ a, Dst *b,
                                                      double scale) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        b[i] = (Dst)((double)a[i] * scale);
}

template <typename Real>
__global__ void __launch_bounds__(256) fill_kernel(int64_t n, Real *a, Real v) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        a[i] = v;
}

// y = a*x + b*y (FP64)
// y = a x + b y; b = 0 assigns (y is not read: it may be uninitialised memory, where 0 * NaN would
// propagate), a = 0 scales y alone
__global__ void __launch_bounds__(256) axpby_kernel(int64_t n, double a, const double *__restrict__ x,
                                                    double b, double *y) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        y[i] = b == 0.0 ? a * x[i] : (a == 0.0 ? b * y[i] : a * x[i] + b * y[i]);
}

// Invert the 3x3 point blocks of A = -J; rows/columns of constrained dofs become identity.
__global__ void __launch_bounds__(256) invert_blocks(int nnodes, const double *__restrict__ Jblocks,
                                                     const uint8_t *__restrict__ conmask,
                                                     double *Binv64, float *Binv32) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nnodes) return;
    double A[9];
    for (int i = 0; i < 9; ++i) A[i] = -Jblocks[9 * (size_t)node + i];
    for (int c = 0; c < 3; ++c)
        if (conmask[3 * (size_t)node + c]) {
            for (int k = 0; k < 3; ++k) A[3 * c + k] = A[3 * k + c] = 0.0;
            A[4 * c] = 1.0;
        }
    double Ai[9];
    const double det = inv3(A, Ai);
    for (int i = 0; i < 9; ++i) {
        const double v = det != 0.0 ? Ai[i] : (i % 4 == 0 ? 1.0 : 0.0);
        Binv64[9 * (size_t)node + i] = v;
        Binv32[9 * (size_t)node + i] = (float)v;
    }
}

// blocks[neg] += blocks[pos] for every split node pair (3x3 blocks)
__global__ void __launch_bounds__(256) merge_blocks(int n, const int32_t *__restrict__ negn,
                                                   const int32_t *__restrict__ posn, double *blocks) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 9; ++k) blocks[9 * (size_t)negn[i] + k] += blocks[9 * (size_t)posn[i] + k];
}

}  // namespace plx

using namespace plx;

struct plx_solver_s {
    plx_dm dm = nullptr;
    KeyPlan *kp = nullptr;           // the first key
    std::vector<KeyPlan *> keys;     // every key of the operator (one per material)
    int label_value = 0, part = 0;
    int64_t n = 0;
    int nnodes = 0;
    double *Binv64 = nullptr;
    float *Binv32 = nullptr;
    // FP32 inner PCG work vectors
    float *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *z32 = nullptr;
    // FP64 PCG work vectors
    double *r64 = nullptr, *p64 = nullptr, *q64 = nullptr, *z64 = nullptr, *t64 = nullptr;
    // FGMRES bases
    int restart = 30;
    std::vector<double *> V, Z;
    double *scal = nullptr, *part_a = nullptr, *part_b = nullptr;
    double inner_rtol = 1e-4;
    int inner_maxit = 2000;
    plx_solve_stats last{};
    // multigrid preconditioner: the assembled merged operator and its hierarchy (Q1 keys)
    plx_gf amg = nullptr;
    // solves run on this stream; chunks of Krylov iterations are captured as hipGraphs, keyed by
    // (kind, buffers, preconditioner) and valid for one DM epoch
    hipStream_t stream = nullptr;
    bool use_graph = true;
    uint64_t graph_epoch = ~0ull;
    std::map<std::array<uintptr_t, 4>, hipGraphExec_t> graphs;
    double *lsol = nullptr;  // fault-mass CG iterate
    // fault (null-space method): jump across split nodes and Lagrange-space work vectors
    bool fault = false;
    double *jump = nullptr, *lr = nullptr, *lp = nullptr, *lq = nullptr, *lz = nullptr, *lx = nullptr;
};

namespace plx {

static inline dim3 red_grid() { return dim3(kRedBlocks); }

// y = scale * J x summed over the solver's keys (the dofs of their closures are written: one key
// inserts; several keys zero y first and accumulate)
static int apply_keys(plx_solver s, int st, const void *x, void *y, double scale) {
    plx_dm dm = s->dm;
    if (s->keys.size() == 1) return launch_apply(dm, s->keys[0], st, true, x, y, PLX_INSERT_VALUES, scale);
    if (st == PLX_FP64)
        hipLaunchKernelGGL(fill_kernel<double>, red_grid(), dim3(256), 0, dm->stream, s->n, (double *)y, 0.0);
    else
        hipLaunchKernelGGL(fill_kernel<float>, red_grid(), dim3(256), 0, dm->stream, s->n, (float *)y, 0.0f);
    PLX_HIP(dm, hipGetLastError());
    for (KeyPlan *kp : s->keys)
        if (int err = launch_apply(dm, kp, st, true, x, y, PLX_ADD_VALUES, scale)) return err;
    return 0;
}

// y = A x on the unconstrained dofs.  With a fault x is a merged vector (zeros at positive split
// nodes) and y = U^T A U x; x's positive entries are filled for the apply and cleared again.
template <typename Real>
static int masked_apply(plx_solver s, Real *x, Real *y) {
    plx_dm dm = s->dm;
    const int st = std::is_same<Real, double>::value ? PLX_FP64 : PLX_FP32;
    int err = 0;
    if (s->fault && (err = launch_fold(dm, st, x, kUnfold))) return err;
    if ((err = apply_keys(s, st, x, y, -1.0))) return err;
    if (s->fault && ((err = launch_fold(dm, st, x, kClearPos)) || (err = launch_fold(dm, st, y, kFold))))
        return err;
    const int nc = (int)dm->con.dofs.size();
    if (nc > 0) {
        hipLaunchKernelGGL(mask_kernel<Real>, dim3((nc + 255) / 256), dim3(256), 0, dm->stream, nc,
                           dm->con.d_dofs, y);
        PLX_HIP(dm, hipGetLastError());
    }
    return 0;
}

static int read_scalar(plx_solver s, int idx, double *v) {
    PLX_HIP(s->dm, hipMemcpyAsync(v, s->scal + idx, sizeof(double), hipMemcpyDeviceToHost, s->dm->stream));
    PLX_HIP(s->dm, hipStreamSynchronize(s->dm->stream));
    return 0;
}

static void clear_graphs(plx_solver s) {
    for (auto &g : s->graphs) (void)hipGraphExecDestroy(g.second);
    s->graphs.clear();
}

// Run body() (a fixed chunk of iterations on fixed buffers, no host synchronisation) through a
// captured graph; eager when graphs are off or the handle records events.
template <typename F>
static int run_chunk(plx_solver s, uintptr_t kind, const void *a, const void *b, F &&body) {
    plx_dm dm = s->dm;
    if (!s->use_graph || dm->events.on || !s->stream || dm->stream != s->stream) return body();
    if (s->graph_epoch != dm->epoch) {
        clear_graphs(s);
        s->graph_epoch = dm->epoch;
    }
    const std::array<uintptr_t, 4> key{kind, (uintptr_t)a, (uintptr_t)b, (uintptr_t)s->amg};
    auto it = s->graphs.find(key);
    if (it == s->graphs.end()) {
        hipGraph_t graph = nullptr;
        PLX_HIP(dm, hipStreamBeginCapture(dm->stream, hipStreamCaptureModeThreadLocal));
        const int err = body();
        const hipError_t e_end = hipStreamEndCapture(dm->stream, &graph);
        if (err || e_end != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return err ? err : set_error(dm, 76, std::string("hipStreamEndCapture: ") + hipGetErrorString(e_end));
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t e_inst = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        PLX_HIP(dm, e_inst);
        it = s->graphs.emplace(key, exec).first;
    }
    PLX_HIP(dm, hipGraphLaunch(it->second, dm->stream));
    return 0;
}

static int dot64n(plx_solver s, int64_t n, const double *a, const double *b, double *out) {
    plx_dm dm = s->dm;
    hipLaunchKernelGGL(dot_partials<double>, red_grid(), dim3(256), 0, dm->stream, n, a, b, s->part_a);
    hipLaunchKernelGGL(finish_start, dim3(1), dim3(256), 0, dm->stream, s->part_a, s->part_a, s->scal);
    PLX_HIP(dm, hipGetLastError());
    return read_scalar(s, S_RZ, out);
}

static int dot64(plx_solver s, const double *a, const double *b, double *out) {
    plx_dm dm = s->dm;
    hipLaunchKernelGGL(dot_partials<double>, red_grid(), dim3(256), 0, dm->stream, s->n, a, b, s->part_a);
    hipLaunchKernelGGL(finish_start, dim3(1), dim3(256), 0, dm->stream, s->part_a, s->part_a, s->scal);
    PLX_HIP(dm, hipGetLastError());
    return read_scalar(s, S_RZ, out);
}

// The V-cycle's fine-level operator: the masked matrix-free FP32 apply (no fault: x is not
// touched; with a fault the V-cycle keeps the assembled merged operator).
static int amg_fine_apply(void *ctx, const float *x, float *y) {
    return masked_apply<float>((plx_solver)ctx, const_cast<float *>(x), y);
}

// z = M r with the multigrid V-cycle (FP32; an FP64 pair goes through the FP32 work vectors,
// which no FP64 iteration uses).
template <typename Real>
static int amg_precondition(plx_solver s, const Real *r, Real *z) {
    plx_dm dm = s->dm;
    if (std::is_same<Real, float>::value) return gf_vcycle1(s->amg, dm->stream, (const float *)r, (float *)z);
    hipLaunchKernelGGL((convert_kernel<double, float>), red_grid(), dim3(256), 0, dm->stream, s->n,
                       (const double *)r, s->r32, 1.0);
    int err = gf_vcycle1(s->amg, dm->stream, s->r32, s->z32);
    if (err) return err;
    hipLaunchKernelGGL((convert_kernel<float, double>), red_grid(), dim3(256), 0, dm->stream, s->n, s->z32,
                       (double *)z, 1.0);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

// PCG with the multigrid preconditioner (same contract as pcg below).
template <typename Real>
static int pcg_amg(plx_solver s, Real *x, Real *r, Real *p, Real *q, Real *z, double rtol, int maxit, int *its,
                   double *rnorm_out) {
    plx_dm dm = s->dm;
    int err = amg_precondition<Real>(s, r, z);
    if (err) return err;
    hipLaunchKernelGGL(rz_partials<Real>, red_grid(), dim3(256), 0, dm->stream, s->n, r, z, p, s->part_a, s->part_b);
    hipLaunchKernelGGL(finish_start, dim3(1), dim3(256), 0, dm->stream, s->part_a, s->part_b, s->scal);
    PLX_HIP(dm, hipGetLastError());
    double rr0 = 0.0;
    if ((err = read_scalar(s, S_RR, &rr0))) return err;
    const double target = rtol * rtol * rr0;
    int it = 0;
    double rr = rr0;
    const int check = 2;  // a V-cycle iteration is expensive next to a convergence check
    auto iteration = [&]() -> int {
        int e = masked_apply<Real>(s, p, q);
        if (e) return e;
        hipLaunchKernelGGL(dot_partials<Real>, red_grid(), dim3(256), 0, dm->stream, s->n, p, q, s->part_a);
        hipLaunchKernelGGL(finish_alpha, dim3(1), dim3(256), 0, dm->stream, s->part_a, s->scal);
        hipLaunchKernelGGL(cg_update_xr<Real>, red_grid(), dim3(256), 0, dm->stream, s->n, x, r, p, q, s->scal,
                           s->part_b);
        if ((e = amg_precondition<Real>(s, r, z))) return e;
        hipLaunchKernelGGL(rz_partials<Real>, red_grid(), dim3(256), 0, dm->stream, s->n, r, z, (Real *)nullptr,
                           s->part_a, (double *)nullptr);
        hipLaunchKernelGGL(finish_beta, dim3(1), dim3(256), 0, dm->stream, s->part_a, s->part_b, s->scal);
        hipLaunchKernelGGL(p_update<Real>, red_grid(), dim3(256), 0, dm->stream, s->n, z, p, s->scal);
        PLX_HIP(dm, hipGetLastError());
        return 0;
    };
    while (it < maxit && rr > target && rr0 > 0.0) {
        if (it + check <= maxit) {
            err = run_chunk(s, 10 * sizeof(Real) + 2, x, r, [&]() -> int {
                for (int k = 0; k < check; ++k)
                    if (int e = iteration()) return e;
                return 0;
            });
            if (err) return err;
            it += check;
        } else {
            for (; it < maxit; ++it)
                if ((err = iteration())) return err;
        }
        if ((err = read_scalar(s, S_RR, &rr))) return err;
    }
    *its = it;
    *rnorm_out = rr0 > 0.0 ? std::sqrt(rr / rr0) : 0.0;
    return 0;
}

// PCG on A x = b (x, b, work vectors of type Real; x must hold the initial guess, masked).
// Stops when ||r|| <= rtol ||b|| or maxit; checks convergence every `check` iterations.
template <typename Real>
static int pcg(plx_solver s, Real *x, Real *r, Real *p, Real *q, Real *z, const Real *Binv,
               double rtol, int maxit, int *its, double *rnorm_out) {
    if (s->amg) return pcg_amg<Real>(s, x, r, p, q, z, rtol, maxit, its, rnorm_out);
    plx_dm dm = s->dm;
    // r = b - A x  (caller stored b in r; x == 0 guess skips the apply)
    hipLaunchKernelGGL(pc_start<Real>, red_grid(), dim3(256), 0, dm->stream, s->nnodes, Binv, r, z, p,
                       s->part_a, s->part_b);
    hipLaunchKernelGGL(finish_start, dim3(1), dim3(256), 0, dm->stream, s->part_a, s->part_b, s->scal);
    PLX_HIP(dm, hipGetLastError());
    double rr0 = 0.0;
    int err = read_scalar(s, S_RR, &rr0);
    if (err) return err;
    const double target = rtol * rtol * rr0;
    int it = 0;
    double rr = rr0;
    const int check = 10;
    auto iteration = [&]() -> int {
        int e = masked_apply<Real>(s, p, q);
        if (e) return e;
        hipLaunchKernelGGL(dot_partials<Real>, red_grid(), dim3(256), 0, dm->stream, s->n, p, q, s->part_a);
        hipLaunchKernelGGL(finish_alpha, dim3(1), dim3(256), 0, dm->stream, s->part_a, s->scal);
        hipLaunchKernelGGL(cg_update<Real>, red_grid(), dim3(256), 0, dm->stream, s->nnodes, Binv, x, r, p, q, z,
                           s->scal, s->part_a, s->part_b);
        hipLaunchKernelGGL(finish_beta, dim3(1), dim3(256), 0, dm->stream, s->part_a, s->part_b, s->scal);
        hipLaunchKernelGGL(p_update<Real>, red_grid(), dim3(256), 0, dm->stream, s->n, z, p, s->scal);
        PLX_HIP(dm, hipGetLastError());
        return 0;
    };
    while (it < maxit && rr > target && rr0 > 0.0) {
        if (it + check <= maxit) {
            err = run_chunk(s, 10 * sizeof(Real) + 1, x, r, [&]() -> int {
                for (int k = 0; k < check; ++k)
                    if (int e = iteration()) return e;
                return 0;
            });
            if (err) return err;
            it += check;
        } else {
            for (; it < maxit; ++it)
                if ((err = iteration())) return err;
        }
        if ((err = read_scalar(s, S_RR, &rr))) return err;
    }
    *its = it;
    *rnorm_out = rr0 > 0.0 ? std::sqrt(rr / rr0) : 0.0;
    return 0;
}

// True FP64 residual r = b - A x on unconstrained dofs (x carries the boundary values; with a
// fault x is the full displacement, positive nodes included, and r is folded).
static int true_residual(plx_solver s, const double *b, const double *x, double *r) {
    plx_dm dm = s->dm;
    int err = apply_keys(s, PLX_FP64, x, r, -1.0);
    if (err) return err;
    if (b)
        hipLaunchKernelGGL(axpby_kernel, red_grid(), dim3(256), 0, dm->stream, s->n, 1.0, b, -1.0, r);
    else
        hipLaunchKernelGGL(axpby_kernel, red_grid(), dim3(256), 0, dm->stream, s->n, 0.0, r, -1.0, r);
    PLX_HIP(dm, hipGetLastError());
    if (s->fault && (err = launch_fold(dm, PLX_FP64, r, kFold))) return err;
    const int nc = (int)dm->con.dofs.size();
    if (nc > 0) {
        hipLaunchKernelGGL(mask_kernel<double>, dim3((nc + 255) / 256), dim3(256), 0, dm->stream, nc,
                           dm->con.d_dofs, r);
        PLX_HIP(dm, hipGetLastError());
    }
    return 0;
}

// out = M^-1 rhs on the fault (Lagrange space, three components), Jacobi-preconditioned CG in
// FP64 to round-off (M is the well-conditioned fault-face mass matrix).
static int fault_mass_solve(plx_solver s, const double *rhs, double *out) {
    // Jacobi-preconditioned CG with its scalars on the device (finish_* kernels); the host
    // reads the residual norm every 8 iterations only.
    plx_dm dm = s->dm;
    const int64_t n = 3 * (int64_t)dm->fault->num_lag;
    int err = 0;
    double *xs = s->lsol;  // fixed iterate buffer (the chunks are captured)
    hipLaunchKernelGGL(fill_kernel<double>, red_grid(), dim3(256), 0, dm->stream, n, xs, 0.0);
    hipLaunchKernelGGL(axpby_kernel, red_grid(), dim3(256), 0, dm->stream, n, 1.0, rhs, 0.0, s->lr);
    if ((err = launch_fault_jacobi(dm, s->lr, s->lz))) return err;
    // p = z; r.z and r.r (= b.b)
    hipLaunchKernelGGL(rz_partials<double>, red_grid(), dim3(256), 0, dm->stream, n, s->lr, s->lz, s->lp,
                       s->part_a, s->part_b);
    hipLaunchKernelGGL(finish_start, dim3(1), dim3(256), 0, dm->stream, s->part_a, s->part_b, s->scal);
    PLX_HIP(dm, hipGetLastError());
    double bb = 0.0;
    if ((err = read_scalar(s, S_RR, &bb))) return err;
    const int check = 8;
    double rr = bb;
    for (int it = 0; it < 500 && bb > 0.0 && rr > 1e-30 * bb; it += check) {
        err = run_chunk(s, 3, xs, s->lr, [&]() -> int {
            for (int k = 0; k < check; ++k) {
                int e = launch_fault_mass(dm, s->lp, s->lq);
                if (e) return e;
                hipLaunchKernelGGL(dot_partials<double>, red_grid(), dim3(256), 0, dm->stream, n, s->lp, s->lq,
                                   s->part_a);
                hipLaunchKernelGGL(finish_alpha, dim3(1), dim3(256), 0, dm->stream, s->part_a, s->scal);
                hipLaunchKernelGGL(cg_update_xr<double>, red_grid(), dim3(256), 0, dm->stream, n, xs, s->lr, s->lp,
                                   s->lq, s->scal, s->part_b);
                if ((e = launch_fault_jacobi(dm, s->lr, s->lz))) return e;
                hipLaunchKernelGGL(rz_partials<double>, red_grid(), dim3(256), 0, dm->stream, n, s->lr, s->lz,
                                   (double *)nullptr, s->part_a, (double *)nullptr);
                hipLaunchKernelGGL(finish_beta, dim3(1), dim3(256), 0, dm->stream, s->part_a, s->part_b, s->scal);
                hipLaunchKernelGGL(p_update<double>, red_grid(), dim3(256), 0, dm->stream, n, s->lz, s->lp, s->scal);
            }
            PLX_HIP(dm, hipGetLastError());
            return 0;
        });
        if (err) return err;
        if ((err = read_scalar(s, S_RR, &rr))) return err;
    }
    hipLaunchKernelGGL(axpby_kernel, red_grid(), dim3(256), 0, dm->stream, n, 1.0, xs, 0.0, out);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

// Jump across the split nodes: M jump = g - b_lambda, g = integral psi d (the slip forcing).
static int fault_jump(plx_solver s, const double *b) {
    plx_dm dm = s->dm;
    const int64_t n = 3 * (int64_t)dm->fault->num_lag;
    if (dm->fault->planar && !b) return launch_planar_jump(dm, dm->fault->num_lag, dm->fault->d_slip, s->jump);
    int err = launch_fault_slip_rhs(dm, s->lx);  // lx = -g
    if (err) return err;
    if (b)
        hipLaunchKernelGGL(axpby_kernel, red_grid(), dim3(256), 0, dm->stream, n, -1.0,
                           b + dm->fault->lag_offset, -1.0, s->lx);
    else
        hipLaunchKernelGGL(axpby_kernel, red_grid(), dim3(256), 0, dm->stream, n, 0.0, s->lx, -1.0, s->lx);
    PLX_HIP(dm, hipGetLastError());
    return fault_mass_solve(s, s->lx, s->jump);
}

// Lagrange multipliers from the positive-side rows: M lambda = (A u)_+ - b_+.
static int fault_lambda(plx_solver s, const double *b, double *x) {
    plx_dm dm = s->dm;
    const int64_t n = 3 * (int64_t)dm->fault->num_lag;
    int err = apply_keys(s, PLX_FP64, x, s->t64, -1.0);
    if (err || (err = launch_gather_pos(dm, s->t64, s->lx))) return err;
    if (b) {
        if ((err = launch_gather_pos(dm, b, s->lq))) return err;
        hipLaunchKernelGGL(axpby_kernel, red_grid(), dim3(256), 0, dm->stream, n, -1.0, s->lq, 1.0, s->lx);
        PLX_HIP(dm, hipGetLastError());
    }
    return fault_mass_solve(s, s->lx, x + dm->fault->lag_offset);
}

// Re-impose u+ = u- + jump after an update of the negative-side values.
static int sync_x(plx_solver s, double *x) {
    return s->fault ? launch_sync_pos(s->dm, s->jump, x) : 0;
}

// Inner preconditioner: out (FP64) ~= A^{-1} in (FP64) by FP32 PCG from a zero guess.
static int inner_solve(plx_solver s, const double *in, double *out, int *its) {
    plx_dm dm = s->dm;
    double nrm = 0.0;
    int err = dot64(s, in, in, &nrm);
    if (err) return err;
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) {
        hipLaunchKernelGGL(fill_kernel<double>, red_grid(), dim3(256), 0, dm->stream, s->n, out, 0.0);
        *its = 0;
        return 0;
    }
    hipLaunchKernelGGL((convert_kernel<double, float>), red_grid(), dim3(256), 0, dm->stream, s->n, in,
                       s->r32, 1.0 / nrm);
    hipLaunchKernelGGL(fill_kernel<float>, red_grid(), dim3(256), 0, dm->stream, s->n, s->x32, 0.0f);
    PLX_HIP(dm, hipGetLastError());
    double rn = 0.0;
    if ((err = pcg<float>(s, s->x32, s->r32, s->p32, s->q32, s->z32, s->Binv32, s->inner_rtol,
                          s->inner_maxit, its, &rn)))
        return err;
    hipLaunchKernelGGL((convert_kernel<float, double>), red_grid(), dim3(256), 0, dm->stream, s->n, s->x32,
                       out, nrm);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

}  // namespace plx

extern "C" {

int plx_solver_create(plx_solver *out, plx_dm dm, int label_value, int part) {
    if (!out || !dm) return 1;
    *out = nullptr;
    std::vector<KeyPlan *> keys;
    if (label_value == PLX_ALL_KEYS) {
        for (KeyPlan *k : dm->keys)
            if (k->part == part && k->jf3 && k->has_aux) keys.push_back(k);
        if (keys.empty()) return set_error(dm, 1, "no key of this part has Jacobian kernels and aux");
    } else {
        KeyPlan *kp = find_key(dm, label_value, part);
        if (!kp || !kp->jf3 || !kp->has_aux)
            return set_error(dm, 1, "solver needs a key with Jacobian kernels and aux");
        keys.push_back(kp);
    }
    plx_solver s = new plx_solver_s();
    s->dm = dm;
    s->kp = keys[0];
    s->keys = keys;
    s->label_value = label_value;
    s->part = part;
    s->n = 3 * (int64_t)dm->num_nodes;  // the displacement block; Lagrange dofs are eliminated
    s->nnodes = dm->num_nodes;
    const size_t n = (size_t)s->n;
    for (int32_t d : dm->con.dofs)
        if (d >= s->n) {
            delete s;
            return set_error(dm, 1, "constrained Lagrange-multiplier dofs are not supported");
        }
    if (dm->fault) {
        s->fault = true;
        std::vector<char> is_pos(dm->num_nodes, 0);
        for (int32_t p : dm->fault->pos_node) is_pos[p] = 1;
        for (int32_t d : dm->con.dofs)
            if (is_pos[d / 3]) {
                delete s;
                return set_error(dm, 1, "constrain the negative-side twin of a split fault node, not the positive one");
            }
    }
    bool ok = hipMalloc(&s->Binv64, 9 * (size_t)s->nnodes * sizeof(double)) == hipSuccess &&
              hipMalloc(&s->Binv32, 9 * (size_t)s->nnodes * sizeof(float)) == hipSuccess &&
              hipMalloc(&s->scal, 16 * sizeof(double)) == hipSuccess &&
              hipMalloc(&s->part_a, kRedBlocks * sizeof(double)) == hipSuccess &&
              hipMalloc(&s->part_b, kRedBlocks * sizeof(double)) == hipSuccess;
    for (float **v : {&s->x32, &s->r32, &s->p32, &s->q32, &s->z32})
        ok = ok && hipMalloc(v, n * sizeof(float)) == hipSuccess;
    for (double **v : {&s->r64, &s->p64, &s->q64, &s->z64, &s->t64})
        ok = ok && hipMalloc(v, n * sizeof(double)) == hipSuccess;
    if (s->fault) {
        const size_t nl = 3 * (size_t)dm->fault->num_lag;
        for (double **v : {&s->jump, &s->lr, &s->lp, &s->lq, &s->lz, &s->lx, &s->lsol})
            ok = ok && hipMalloc(v, nl * sizeof(double)) == hipSuccess;
    }
    ok = ok && hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) == hipSuccess;
    if (!ok) {
        plx_solver_destroy(s);
        return set_error(dm, 76, "solver allocation failed");
    }
    *out = s;
    return plx_solver_setup(s);
}

int plx_solver_destroy(plx_solver s) {
    if (!s) return 0;
    void *ptrs[] = {s->Binv64, s->Binv32, s->scal, s->part_a, s->part_b, s->x32, s->r32, s->p32,
                    s->q32,    s->z32,    s->r64,  s->p64,    s->q64,    s->z64, s->t64};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (double *p : {s->jump, s->lr, s->lp, s->lq, s->lz, s->lx, s->lsol})
        if (p) (void)hipFree(p);
    clear_graphs(s);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    for (double *p : s->V) (void)hipFree(p);
    for (double *p : s->Z) (void)hipFree(p);
    if (s->amg) plx_gf_destroy(s->amg);
    delete s;
    return 0;
}

int plx_solver_setup(plx_solver s) {
    if (!s) return 1;
    plx_dm dm = s->dm;
    // set-up buffers, freed on every exit path
    struct Scratch {
        double *blocks = nullptr;
        uint8_t *mask = nullptr;
        ~Scratch() {
            if (blocks) (void)hipFree(blocks);
            if (mask) (void)hipFree(mask);
        }
    } tmp;
    PLX_HIP(dm, hipMalloc(&tmp.blocks, 9 * (size_t)s->nnodes * sizeof(double)));
    PLX_HIP(dm, hipMemsetAsync(tmp.blocks, 0, 9 * (size_t)s->nnodes * sizeof(double), dm->stream));
    int err = 0;
    for (KeyPlan *kp : s->keys)
        if ((err = launch_pbdiag(dm, kp, tmp.blocks))) {
            (void)hipStreamSynchronize(dm->stream);
            return err;
        }
    std::vector<uint8_t> hmask(s->n, 0);
    for (int32_t d : dm->con.dofs) hmask[d] = 1;
    if (s->fault) {
        // merged point blocks: the negative node carries both twins' diagonal blocks (the two
        // sides share no cell); positive nodes are outside the merged space (identity)
        for (int32_t p : dm->fault->pos_node)
            for (int c = 0; c < 3; ++c) hmask[3 * (size_t)p + c] = 1;
        const int nl = dm->fault->num_lag;
        hipLaunchKernelGGL(merge_blocks, dim3((nl + 255) / 256), dim3(256), 0, dm->stream, nl,
                           dm->fault->d_neg_node, dm->fault->d_pos_node, tmp.blocks);
        PLX_HIP(dm, hipGetLastError());
    }
    PLX_HIP(dm, hipMalloc(&tmp.mask, hmask.size()));
    PLX_HIP(dm, hipMemcpyAsync(tmp.mask, hmask.data(), hmask.size(), hipMemcpyHostToDevice, dm->stream));
    hipLaunchKernelGGL(invert_blocks, dim3((s->nnodes + 255) / 256), dim3(256), 0, dm->stream, s->nnodes,
                       tmp.blocks, tmp.mask, s->Binv64, s->Binv32);
    PLX_HIP(dm, hipGetLastError());
    // the host mask and the buffers must outlive the queued copy and kernels
    PLX_HIP(dm, hipStreamSynchronize(dm->stream));
    clear_graphs(s);
    if (s->amg) {  // rebuild the hierarchy from the current operator
        plx_gf_destroy(s->amg);
        s->amg = nullptr;
        return plx_solver_set_preconditioner(s, PLX_PC_AMG);
    }
    return 0;
}

int plx_solver_set_preconditioner(plx_solver s, int kind) {
    if (!s) return 1;
    plx_dm dm = s->dm;
    clear_graphs(s);
    if (kind == PLX_PC_PBJACOBI) {
        if (s->amg) plx_gf_destroy(s->amg);
        s->amg = nullptr;
        return 0;
    }
    if (kind != PLX_PC_AMG) return set_error(dm, 1, "preconditioner must be PLX_PC_PBJACOBI or PLX_PC_AMG");
    if (s->amg) return 0;
    plx_gf g = nullptr;
    int err = plx_gf_create(&g, dm, s->label_value, s->part);
    if (!err) err = plx_gf_set_preconditioner(g, PLX_GF_PC_AMG);
    if (!err && !s->fault) err = gf_set_fine_operator(g, amg_fine_apply, s);
    if (err) {
        if (g) plx_gf_destroy(g);
        return err;
    }
    // the hierarchy is read on this handle's stream from now on
    PLX_HIP(dm, hipDeviceSynchronize());
    s->amg = g;
    return 0;
}

int plx_solver_set_inner(plx_solver s, double rtol, int maxit, int restart) {
    if (!s || rtol <= 0.0 || maxit <= 0 || restart <= 0 || restart > 200) return 1;
    s->inner_rtol = rtol;
    s->inner_maxit = maxit;
    s->restart = restart;
    return 0;
}

}  // extern "C"

namespace plx {

static int solve_impl(plx_solver s, int method, const double *b, double *x, double rtol, int maxit,
                      plx_solve_stats *stats) {
    plx_dm dm = s->dm;
    plx_solve_stats st{};
    int err = 0;
    const int nc = (int)dm->con.dofs.size();
    // fault: fix the jump across the split nodes, then x = U w + u_jump throughout
    if (s->fault && ((err = fault_jump(s, b)) || (err = sync_x(s, x)))) return err;
    // r = b - A x (true FP64 residual)
    if ((err = true_residual(s, b, x, s->r64))) return err;
    double rr0 = 0.0;
    if ((err = dot64(s, s->r64, s->r64, &rr0))) return err;
    const double r0 = std::sqrt(rr0);
    st.initial_residual = r0;
    if (method == PLX_KSP_CG) {
        // PCG in FP64 on the correction: A d = r, x += d
        double *d = s->t64;
        hipLaunchKernelGGL(fill_kernel<double>, red_grid(), dim3(256), 0, dm->stream, s->n, d, 0.0);
        int its = 0;
        double rn = 0.0;
        if ((err = pcg<double>(s, d, s->r64, s->p64, s->q64, s->z64, s->Binv64, rtol, maxit, &its, &rn)))
            return err;
        hipLaunchKernelGGL(axpby_kernel, red_grid(), dim3(256), 0, dm->stream, s->n, 1.0, d, 1.0, x);
        PLX_HIP(dm, hipGetLastError());
        if ((err = sync_x(s, x))) return err;
        st.outer_iterations = its;
        st.applies_fp64 = its + 1;
    } else if (method == PLX_KSP_IR) {
        double rr = rr0;
        while (st.outer_iterations < maxit && r0 > 0.0 && std::sqrt(rr) > rtol * r0) {
            int its = 0;
            if ((err = inner_solve(s, s->r64, s->t64, &its))) return err;
            st.inner_iterations += its;
            hipLaunchKernelGGL(axpby_kernel, red_grid(), dim3(256), 0, dm->stream, s->n, 1.0, s->t64, 1.0, x);
            if ((err = sync_x(s, x))) return err;
            if ((err = true_residual(s, b, x, s->r64))) return err;
            if ((err = dot64(s, s->r64, s->r64, &rr))) return err;
            st.outer_iterations++;
            st.applies_fp64++;
        }
        st.applies_fp64++;
    } else if (method == PLX_KSP_FGMRES_IR) {
        const int m = s->restart;
        while ((int)s->V.size() < m + 1) {
            double *v = nullptr;
            PLX_HIP(dm, hipMalloc(&v, (size_t)s->n * sizeof(double)));
            s->V.push_back(v);
        }
        while ((int)s->Z.size() < m) {
            double *v = nullptr;
            PLX_HIP(dm, hipMalloc(&v, (size_t)s->n * sizeof(double)));
            s->Z.push_back(v);
        }
        double beta = r0;
        std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1);
        bool done = r0 == 0.0;
        while (!done && st.outer_iterations < maxit) {
            // V0 = r / beta
            hipLaunchKernelGGL(axpby_kernel, red_grid(), dim3(256), 0, dm->stream, s->n, 1.0 / beta, s->r64,
                               0.0, s->V[0]);
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            int j = 0;
            for (; j < m && st.outer_iterations < maxit; ++j) {
                int its = 0;
                if ((err = inner_solve(s, s->V[j], s->Z[j], &its))) return err;
                st.inner_iterations += its;
                if (nc > 0)
                    hipLaunchKernelGGL(mask_kernel<double>, dim3((nc + 255) / 256), dim3(256), 0, dm->stream,
                                       nc, dm->con.d_dofs, s->Z[j]);
                if ((err = masked_apply<double>(s, s->Z[j], s->V[j + 1]))) return err;
                st.applies_fp64++;
                // modified Gram-Schmidt
                for (int i = 0; i <= j; ++i) {
                    double h = 0.0;
                    if ((err = dot64(s, s->V[i], s->V[j + 1], &h))) return err;
                    H[(size_t)i * m + j] = h;
                    hipLaunchKernelGGL(axpby_kernel, red_grid(), dim3(256), 0, dm->stream, s->n, -h, s->V[i],
                                       1.0, s->V[j + 1]);
                }
                double hn = 0.0;
                if ((err = dot64(s, s->V[j + 1], s->V[j + 1], &hn))) return err;
                hn = std::sqrt(hn);
                H[(size_t)(j + 1) * m + j] = hn;
                if (hn > 0.0)
                    hipLaunchKernelGGL(axpby_kernel, red_grid(), dim3(256), 0, dm->stream, s->n, 1.0 / hn,
                                       s->V[j + 1], 0.0, s->V[j + 1]);
                // Givens rotations
                for (int i = 0; i < j; ++i) {
                    const double a = H[(size_t)i * m + j], c = H[(size_t)(i + 1) * m + j];
                    H[(size_t)i * m + j] = cs[i] * a + sn[i] * c;
                    H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
                }
                const double a = H[(size_t)j * m + j], c = H[(size_t)(j + 1) * m + j];
                const double rho = std::sqrt(a * a + c * c);
                cs[j] = rho > 0 ? a / rho : 1.0;
                sn[j] = rho > 0 ? c / rho : 0.0;
                H[(size_t)j * m + j] = rho;
                H[(size_t)(j + 1) * m + j] = 0.0;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                st.outer_iterations++;
                if (std::fabs(g[j + 1]) <= rtol * r0 || hn == 0.0) {
                    ++j;
                    break;
                }
            }
            // x += Z y, H y = g (upper triangular j x j)
            std::vector<double> y(j, 0.0);
            for (int i = j - 1; i >= 0; --i) {
                double acc = g[i];
                for (int k = i + 1; k < j; ++k) acc -= H[(size_t)i * m + k] * y[k];
                y[i] = acc / H[(size_t)i * m + i];
            }
            for (int i = 0; i < j; ++i)
                hipLaunchKernelGGL(axpby_kernel, red_grid(), dim3(256), 0, dm->stream, s->n, y[i], s->Z[i], 1.0, x);
            PLX_HIP(dm, hipGetLastError());
            if ((err = sync_x(s, x))) return err;
            // restart from the true residual
            if ((err = true_residual(s, b, x, s->r64))) return err;
            double rr = 0.0;
            if ((err = dot64(s, s->r64, s->r64, &rr))) return err;
            beta = std::sqrt(rr);
            st.applies_fp64++;
            done = beta <= rtol * r0;
        }
    } else {
        return set_error(dm, 1, "unknown solver method");
    }
    if ((err = true_residual(s, b, x, s->r64))) return err;
    double rr = 0.0;
    if ((err = dot64(s, s->r64, s->r64, &rr))) return err;
    st.final_residual = std::sqrt(rr);
    if (s->fault && (err = fault_lambda(s, b, x))) return err;
    st.converged = st.final_residual <= rtol * r0 * (method == PLX_KSP_CG ? 10.0 : 1.0) ? 1 : 0;
    s->last = st;
    if (stats) *stats = st;
    return 0;
}

}  // namespace plx

extern "C" {

// The solve runs on the solver's own stream (ordered after the caller's queued work and before
// anything the caller queues afterwards), where its iteration chunks can be graph-captured.
int plx_solver_solve(plx_solver s, int method, const double *b, double *x, double rtol, int maxit,
                     plx_solve_stats *stats) {
    if (!s || !x || rtol <= 0.0 || maxit <= 0) return 1;
    plx_dm dm = s->dm;
    hipStream_t caller = dm->stream;
    hipEvent_t ev = nullptr;
    PLX_HIP(dm, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    if (hipEventRecord(ev, caller) != hipSuccess || hipStreamWaitEvent(s->stream, ev, 0) != hipSuccess) {
        (void)hipEventDestroy(ev);
        return set_error(dm, 76, "solver stream ordering failed");
    }
    dm->stream = s->stream;
    const int err = solve_impl(s, method, b, x, rtol, maxit, stats);
    dm->stream = caller;
    const bool ordered = hipEventRecord(ev, s->stream) == hipSuccess && hipStreamWaitEvent(caller, ev, 0) == hipSuccess;
    (void)hipEventDestroy(ev);
    if (err) return err;
    if (!ordered) return set_error(dm, 76, "solver stream ordering failed");
    return 0;
}

}  // extern "C"
