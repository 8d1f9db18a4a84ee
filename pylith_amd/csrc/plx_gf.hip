// plx_gf.hip -- batched Green's functions: many fault-slip impulses solved together.
//
// PyLith's GreensFns problem (problems/GreensFns.cc:236-331) assembles the Jacobian once and
// solves one linear system per impulse of FaultCohesiveImpulses (FaultCohesiveImpulses.cc:
// 263-300), each impulse a unit slip on part of the fault.  All impulses share the operator.
// Here:
//   * the operator is assembled once into a 3x3-block CSR matrix (BSR) of the merged system:
//     every positive-side split node is identified with its negative twin, so the matrix is the
//     stiffness U^T K U of the constraint-eliminated fault problem (the fault disappears from
//     it), with constrained dofs as identity rows/columns; element matrices are summed in a
//     fixed order per block (deterministic);
//   * an impulse's right-hand side is -U^T K u_jump, with u_jump the slip jump R s on the
//     positive side (matrix-free FP64 apply);
//   * k impulses form a block: vectors interleave the k columns per dof ([dof][k]), so one
//     sparse-matrix x block-vector product reads each matrix block once for k columns;
//   * the block is solved by FP64 iterative refinement around an FP32 block PCG (point-block
//     Jacobi, per-column scalars on the device), ten iterations per captured hipGraph replay;
//   * the block width k is chosen by timing probe iterations for k in {8, 16, 32, 64}.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "plx_amg.hpp"
#include "plx_element.hpp"
#include "plx_internal.hpp"

namespace plx {

constexpr int kGfRed = 2048;  // blocks of the grid-stride block kernels (= per-column partials)
enum { G_RZ = 0, G_PQ, G_ALPHA, G_BETA, G_RR, G_COUNT };

// ---------------------------------------------------------------------------------------------
// Assembly.
// ---------------------------------------------------------------------------------------------
// Element stiffness blocks K_ab (3x3, a, b Q1 vertices in PyLith order) of one key cell, the
// matrix of the operator -J for isotropic C:  K_ab[f][g] = sum_q w |J| (lambda da_f db_g +
// mu (delta_fg da.db + da_g db_f)).  Thread (cell position p, row vertex a).
__global__ void __launch_bounds__(256) gf_element_kernel(int ncell, const int32_t *__restrict__ cells,
                                                         const int32_t *__restrict__ closure,
                                                         const double *__restrict__ coords,
                                                         const double *__restrict__ mat, double *Ke) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 8 * ncell) return;
    const int p = t >> 3, a = t & 7;
    const int32_t *cl = closure + 8 * (size_t)cells[p];
    const double lam = mat[p], mu = mat[ncell + p];
    const double sgn[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                              {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    double X[8][3];
    for (int v = 0; v < 8; ++v)
        for (int d = 0; d < 3; ++d) X[v][d] = coords[3 * (size_t)cl[v] + d];
    double K[8][9];
    for (int b = 0; b < 8; ++b)
        for (int i = 0; i < 9; ++i) K[b][i] = 0.0;
    const double g = 0.57735026918962576451;
    for (int q = 0; q < 8; ++q) {
        const double xi[3] = {(q & 1) ? g : -g, (q & 2) ? g : -g, (q & 4) ? g : -g};
        double dphi[8][3], J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int v = 0; v < 8; ++v) {
            const double e0 = 1.0 + sgn[v][0] * xi[0], e1 = 1.0 + sgn[v][1] * xi[1], e2 = 1.0 + sgn[v][2] * xi[2];
            dphi[v][0] = 0.125 * sgn[v][0] * e1 * e2;
            dphi[v][1] = 0.125 * e0 * sgn[v][1] * e2;
            dphi[v][2] = 0.125 * e0 * e1 * sgn[v][2];
            for (int d = 0; d < 3; ++d)
                for (int r = 0; r < 3; ++r) J[3 * d + r] += X[v][d] * dphi[v][r];
        }
        double Ji[9];
        const double det = inv3(J, Ji);
        double gr[8][3];
        for (int v = 0; v < 8; ++v)
            for (int d = 0; d < 3; ++d)
                gr[v][d] = dphi[v][0] * Ji[d] + dphi[v][1] * Ji[3 + d] + dphi[v][2] * Ji[6 + d];
        const double *ga = gr[a];
        for (int b = 0; b < 8; ++b) {
            const double *gb = gr[b];
            const double dd = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
            for (int f = 0; f < 3; ++f)
                for (int h = 0; h < 3; ++h)
                    K[b][3 * f + h] += det * (lam * ga[f] * gb[h] + mu * ga[h] * gb[f] + (f == h ? mu * dd : 0.0));
        }
    }
    double *out = Ke + ((size_t)p * 8 + a) * 8 * 9;
    for (int b = 0; b < 8; ++b)
        for (int i = 0; i < 9; ++i) out[9 * b + i] = K[b][i];
}

// One BSR block (row node i, slot s): sum of the element blocks in fixed (cell, vertex) order,
// then identity rows / columns for constrained dofs and for positive split nodes.
__global__ void __launch_bounds__(256) gf_assemble_kernel(int nrows, const int32_t *__restrict__ row_ptr,
                                                          const int32_t *__restrict__ col,
                                                          const int32_t *__restrict__ inc_ptr,
                                                          const int32_t *__restrict__ inc,
                                                          const int32_t *__restrict__ mclosure,
                                                          const double *__restrict__ Ke,
                                                          const uint8_t *__restrict__ conmask,
                                                          const uint8_t *__restrict__ is_pos, double *val) {
    const int i = blockIdx.x;
    const int s = threadIdx.x;
    if (i >= nrows) return;
    const int e = row_ptr[i] + s;
    if (e >= row_ptr[i + 1]) return;
    const int j = col[e];
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!is_pos[i]) {
        for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k) {
            const int p = inc[k] >> 3, a = inc[k] & 7;
            for (int b = 0; b < 8; ++b)
                if (mclosure[8 * (size_t)p + b] == j) {
                    const double *src = Ke + (((size_t)p * 8 + a) * 8 + b) * 9;
                    for (int m = 0; m < 9; ++m) acc[m] += src[m];
                }
        }
    }
    for (int f = 0; f < 3; ++f)
        for (int h = 0; h < 3; ++h) {
            const bool fixed = is_pos[i] || conmask[3 * (size_t)i + f] || conmask[3 * (size_t)j + h];
            if (fixed) acc[3 * f + h] = (i == j && f == h) ? 1.0 : 0.0;
        }
    for (int m = 0; m < 9; ++m) val[9 * (size_t)e + m] = acc[m];
}

__global__ void __launch_bounds__(256) gf_diag_kernel(int nrows, const int32_t *__restrict__ diag_pos,
                                                      const double *__restrict__ val, double *Binv64,
                                                      float *Binv32, float *val32, int64_t nnzb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrows) {
        double Ai[9];
        const double det = inv3(val + 9 * (size_t)diag_pos[i], Ai);
        for (int m = 0; m < 9; ++m) {
            const double v = det != 0.0 ? Ai[m] : (m % 4 == 0 ? 1.0 : 0.0);
            Binv64[9 * (size_t)i + m] = v;
            Binv32[9 * (size_t)i + m] = (float)v;
        }
    }
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < 9 * nnzb;
         t += (int64_t)gridDim.x * blockDim.x)
        val32[t] = (float)val[t];
}

// ---------------------------------------------------------------------------------------------
// Block kernels ([dof][k] layout, k divides 256).
// ---------------------------------------------------------------------------------------------
// Y = A X: thread (row node i, column r) computes the 3 rows of node i for column r.
template <typename Real>
__global__ void __launch_bounds__(256) gf_spmm(int nrows, int k, const int32_t *__restrict__ row_ptr,
                                               const int32_t *__restrict__ col,
                                               const Real *__restrict__ val, const Real *__restrict__ X,
                                               Real *Y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nrows * k) return;
    const int i = (int)(t / k), r = (int)(t - (int64_t)i * k);
    Real y0 = 0, y1 = 0, y2 = 0;
    for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
        const size_t j = (size_t)col[e];
        const Real *A = val + 9 * (size_t)e;
        const Real x0 = X[(3 * j) * k + r], x1 = X[(3 * j + 1) * k + r], x2 = X[(3 * j + 2) * k + r];
        y0 += A[0] * x0 + A[1] * x1 + A[2] * x2;
        y1 += A[3] * x0 + A[4] * x1 + A[5] * x2;
        y2 += A[6] * x0 + A[7] * x1 + A[8] * x2;
    }
    Y[(3 * (size_t)i) * k + r] = y0;
    Y[(3 * (size_t)i + 1) * k + r] = y1;
    Y[(3 * (size_t)i + 2) * k + r] = y2;
}

// One step of the lambda_max power iteration of D^-1 A: v = Binv y * inv_vn (y = A v before), and
// per-block partial sums of |Binv y|^2 (thread = row node).
__global__ void __launch_bounds__(256) gf_binv_step(int nrows, const double *__restrict__ Binv,
                                                    const double *__restrict__ y, double inv_vn, double *v,
                                                    double *part) {
    __shared__ double s[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (i < nrows) {
        const double *B = Binv + 9 * (size_t)i;
        const double *yi = y + 3 * (size_t)i;
        for (int r = 0; r < 3; ++r) {
            const double z = B[3 * r] * yi[0] + B[3 * r + 1] * yi[1] + B[3 * r + 2] * yi[2];
            v[3 * (size_t)i + r] = z * inv_vn;
            acc += z * z;
        }
    }
    s[threadIdx.x] = acc;
    for (int w = 128; w > 0; w >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Per-column block reduction (threads with equal t % k share a column; k divides 256).
__device__ __forceinline__ void col_reduce(double *s, int k) {
    for (int w = 128; w >= k; w >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    }
    __syncthreads();
}

// partials[b][r] = sum over this block's (dof, r) of a*b
template <typename Real>
__global__ void __launch_bounds__(256) gf_dot(int64_t ndof, int k, const Real *__restrict__ a,
                                              const Real *__restrict__ b, double *partials) {
    __shared__ double s[256];
    double acc = 0.0;
    const int64_t n = ndof * k;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
        acc += (double)a[t] * (double)b[t];
    s[threadIdx.x] = acc;
    col_reduce(s, k);
    if ((int)threadIdx.x < k) partials[(size_t)blockIdx.x * k + threadIdx.x] = s[threadIdx.x];
}

// Column r of the per-block partials summed by block r (256 threads, fixed-order tree).
__device__ __forceinline__ double col_sum(const double *partials, int k, int r) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int b = threadIdx.x; b < kGfRed; b += 256) acc += partials[(size_t)b * k + r];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    return s[0];
}

// grid = k blocks of 256 threads: block r finishes column r
__global__ void __launch_bounds__(256) gf_finish_norm(int k, const double *__restrict__ partials, double *out) {
    const int r = blockIdx.x;
    const double v = col_sum(partials, k, r);
    if (threadIdx.x == 0) out[r] = sqrt(v);
}

__global__ void __launch_bounds__(256) gf_finish_alpha(int k, const double *__restrict__ partials, double *scal) {
    const int r = blockIdx.x;
    const double pq = col_sum(partials, k, r);
    if (threadIdx.x == 0) {
        scal[G_PQ * 64 + r] = pq;
        scal[G_ALPHA * 64 + r] = pq != 0.0 ? scal[G_RZ * 64 + r] / pq : 0.0;
    }
}

__global__ void __launch_bounds__(256) gf_finish_beta(int k, const double *__restrict__ prz,
                                                      const double *__restrict__ prr, double *scal, int start) {
    const int r = blockIdx.x;
    const double rz = col_sum(prz, k, r);
    __syncthreads();
    const double rr = col_sum(prr, k, r);
    if (threadIdx.x == 0) {
        const double old = scal[G_RZ * 64 + r];
        scal[G_BETA * 64 + r] = (!start && old != 0.0) ? rz / old : 0.0;
        scal[G_RZ * 64 + r] = rz;
        scal[G_RR * 64 + r] = rr;
    }
}

// Row schedule of the SpMM: XCD x (workgroups are dispatched round-robin over the 8 XCDs)
// owns the slab [lo, hi) of rows; its nbx workgroups walk the slab together, workgroup j taking
// chunk j of every window of nbx chunks.  At any time an XCD's resident workgroups work on one
// compact window of rows, so the neighbour vectors those rows share stay in that XCD's L2.
__device__ __forceinline__ void gf_xcd_slab(int n, int &lo, int &hi, int &j, int &nbx) {
    const int x = blockIdx.x & 7;
    j = blockIdx.x >> 3;
    nbx = gridDim.x >> 3;
    const int per = (n + 7) / 8;
    lo = min(n, x * per);
    hi = min(n, lo + per);
}

// ---- FP32 block PCG kernels with a compile-time block width K (thread = (node, column)) ----
// q = A p and the per-column partial sums of p.q
template <int K>
__global__ void __launch_bounds__(256) gfk_spmm_dot(int nrows, const int32_t *__restrict__ row_ptr,
                                                    const int32_t *__restrict__ col, const float *__restrict__ val,
                                                    const float *__restrict__ P, float *Q, double *ppq) {
    __shared__ double s[256];
    double acc = 0.0;
    int lo, hi, jb, nbx;
    gf_xcd_slab(nrows, lo, hi, jb, nbx);
    constexpr int RW = 256 / K;  // rows per workgroup chunk
    const int r = (int)(threadIdx.x % K);
    for (int base = lo + jb * RW; base < hi; base += nbx * RW) {
        const int64_t i = base + (int)(threadIdx.x / K);
        if (i >= hi) continue;
        float y0 = 0.f, y1 = 0.f, y2 = 0.f;
        const int e1 = row_ptr[i + 1];
        for (int e = row_ptr[i]; e < e1; ++e) {
            const int64_t j = col[e];
            const float *A = val + 9 * (size_t)e;
            const float x0 = P[(3 * j) * K + r], x1 = P[(3 * j + 1) * K + r], x2 = P[(3 * j + 2) * K + r];
            y0 += A[0] * x0 + A[1] * x1 + A[2] * x2;
            y1 += A[3] * x0 + A[4] * x1 + A[5] * x2;
            y2 += A[6] * x0 + A[7] * x1 + A[8] * x2;
        }
        const size_t o = (3 * (size_t)i) * K + r;
        Q[o] = y0;
        Q[o + K] = y1;
        Q[o + 2 * K] = y2;
        acc += (double)P[o] * y0 + (double)P[o + K] * y1 + (double)P[o + 2 * K] * y2;
    }
    s[threadIdx.x] = acc;
    col_reduce(s, K);
    if ((int)threadIdx.x < K) ppq[(size_t)blockIdx.x * K + threadIdx.x] = s[threadIdx.x];
}

// K = 64: one wavefront per block row, lane = column.  The row index is wave-uniform, so the
// row's column indices and 3x3 blocks come through scalar loads and each neighbour's three
// values are one coalesced 256-byte vector load per component.
__global__ void __launch_bounds__(256) gfk_spmm_dot64(int nrows, const int32_t *__restrict__ row_ptr,
                                                      const int32_t *__restrict__ col,
                                                      const float *__restrict__ val, const float *__restrict__ P,
                                                      float *Q, double *ppq) {
    __shared__ double s[256];
    const int lane = threadIdx.x & 63;
    double acc = 0.0;
    int lo, hi, jb, nbx;
    gf_xcd_slab(nrows, lo, hi, jb, nbx);
    for (int row0 = lo + jb * 4 + (int)(threadIdx.x >> 6); row0 < hi; row0 += nbx * 4) {
        const int row = __builtin_amdgcn_readfirstlane(row0);
        const int e0 = row_ptr[row], e1 = row_ptr[row + 1];
        float y0 = 0.f, y1 = 0.f, y2 = 0.f;
        int e = e0;
        // groups of 9 neighbours: all 27 loads of a group are issued before the arithmetic
        for (; e + 9 <= e1; e += 9) {
            float x[9][3];
#pragma unroll
            for (int m = 0; m < 9; ++m) {
                const size_t j = (size_t)col[e + m];
#pragma unroll
                for (int c = 0; c < 3; ++c) x[m][c] = P[(3 * j + c) * 64 + lane];
            }
#pragma unroll
            for (int m = 0; m < 9; ++m) {
                const float *A = val + 9 * (size_t)(e + m);
                y0 += A[0] * x[m][0] + A[1] * x[m][1] + A[2] * x[m][2];
                y1 += A[3] * x[m][0] + A[4] * x[m][1] + A[5] * x[m][2];
                y2 += A[6] * x[m][0] + A[7] * x[m][1] + A[8] * x[m][2];
            }
        }
        for (; e < e1; ++e) {
            const size_t j = (size_t)col[e];
            const float *A = val + 9 * (size_t)e;
            const float x0 = P[(3 * j) * 64 + lane], x1 = P[(3 * j + 1) * 64 + lane], x2 = P[(3 * j + 2) * 64 + lane];
            y0 += A[0] * x0 + A[1] * x1 + A[2] * x2;
            y1 += A[3] * x0 + A[4] * x1 + A[5] * x2;
            y2 += A[6] * x0 + A[7] * x1 + A[8] * x2;
        }
        const size_t o = (3 * (size_t)row) * 64 + lane;
        Q[o] = y0;
        Q[o + 64] = y1;
        Q[o + 128] = y2;
        acc += (double)P[o] * y0 + (double)P[o + 64] * y1 + (double)P[o + 128] * y2;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < 64)
        ppq[(size_t)blockIdx.x * 64 + threadIdx.x] = (s[threadIdx.x] + s[threadIdx.x + 64]) +
                                                     (s[threadIdx.x + 128] + s[threadIdx.x + 192]);
}

// x += alpha p; r -= alpha q; z = Binv r; partial sums of r.z and r.r (start: z = Binv r, p = z)
template <int K, bool START>
__global__ void __launch_bounds__(256) gfk_update(int nrows, const float *__restrict__ Binv,
                                                  const double *__restrict__ scal, float *X, float *R,
                                                  float *P, const float *__restrict__ Q, float *Z, double *prz,
                                                  double *prr) {
    __shared__ double s1[256], s2[256];
    double rz = 0.0, rr = 0.0;
    const int64_t n = (int64_t)nrows * K;
    const int rc = (int)(threadIdx.x % K);  // this thread's column (fixed: the stride is a multiple of K)
    const float a = START ? 0.f : (float)scal[G_ALPHA * 64 + rc];
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const int64_t i = t / K;
        const size_t o = (3 * (size_t)i) * K + rc;
        float rv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (START) {
                rv[c] = R[o + c * K];
            } else {
                X[o + c * K] += a * P[o + c * K];
                rv[c] = R[o + c * K] - a * Q[o + c * K];
                R[o + c * K] = rv[c];
            }
        }
        const float *B = Binv + 9 * i;
        float zc[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            zc[c] = B[3 * c] * rv[0] + B[3 * c + 1] * rv[1] + B[3 * c + 2] * rv[2];
            Z[o + c * K] = zc[c];
            if (START) P[o + c * K] = zc[c];
            rz += (double)rv[c] * zc[c];
            rr += (double)rv[c] * rv[c];
        }
    }
    s1[threadIdx.x] = rz;
    col_reduce(s1, K);
    s2[threadIdx.x] = rr;
    col_reduce(s2, K);
    if ((int)threadIdx.x < K) {
        prz[(size_t)blockIdx.x * K + threadIdx.x] = s1[threadIdx.x];
        prr[(size_t)blockIdx.x * K + threadIdx.x] = s2[threadIdx.x];
    }
}

// p = z + beta p
template <int K>
__global__ void __launch_bounds__(256) gfk_pupdate(int64_t n, const double *__restrict__ scal,
                                                   const float *__restrict__ Z, float *P) {
    const int rc = (int)(threadIdx.x % K);
    const float b = (float)scal[G_BETA * 64 + rc];
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
        P[t] = Z[t] + b * P[t];
}

// ---- multigrid preconditioner kernels ([dof][K] layout) ----
// Block-sparse products of the V-cycle, one wave per (64 / K) block rows, lane = (row, column);
// with K = 64 the row is wave-uniform and the matrix words are scalar loads.  Rows follow the
// per-XCD slab schedule of the fine-level product.
//   kBsrPlain: Y = T X (or Y += T X with accumulate)          transfers (P, R = P^T)
//   kBsrResid: Y = B - A X                                    pre-smoother residual
//   kBsrPost:  Y = X + s Dinv (B - A X)                       post-smoother (damped Jacobi)
enum { kBsrPlain = 0, kBsrResid = 1, kBsrPost = 2 };
template <int BR, int BC, int K, int MODE>
__global__ void __launch_bounds__(256) amg_bsr(int nrows, const int32_t *__restrict__ ptr,
                                               const int32_t *__restrict__ col, const float *__restrict__ val,
                                               const float *__restrict__ X, const float *__restrict__ Bv,
                                               const float *__restrict__ Dinv, float *Y, float s, int accumulate) {
    constexpr int RPW = 64 / K;             // block rows per wave
    constexpr int U = BC == 3 ? 9 : 4;      // entries whose operands are loaded together
    const int lane = threadIdx.x & 63;
    const int c = lane % K;
    int lo, hi, jb, nbx;
    gf_xcd_slab(nrows, lo, hi, jb, nbx);
    for (int row0 = lo + (jb * 4 + (int)(threadIdx.x >> 6)) * RPW; row0 < hi; row0 += nbx * 4 * RPW) {
        int row = row0 + lane / K;
        if (K == 64) row = __builtin_amdgcn_readfirstlane(row);
        if (row >= hi) continue;
        const int e0 = ptr[row], e1 = ptr[row + 1];
        float y[BR];
#pragma unroll
        for (int a = 0; a < BR; ++a) y[a] = 0.f;
        int e = e0;
        for (; e + U <= e1; e += U) {
            float x[U][BC];
#pragma unroll
            for (int m = 0; m < U; ++m) {
                const size_t j = (size_t)col[e + m];
#pragma unroll
                for (int q = 0; q < BC; ++q) x[m][q] = X[(BC * j + q) * K + c];
            }
#pragma unroll
            for (int m = 0; m < U; ++m) {
                const float *A = val + (size_t)BR * BC * (e + m);
#pragma unroll
                for (int a = 0; a < BR; ++a)
#pragma unroll
                    for (int q = 0; q < BC; ++q) y[a] += A[a * BC + q] * x[m][q];
            }
        }
        for (; e < e1; ++e) {
            const size_t j = (size_t)col[e];
            const float *A = val + (size_t)BR * BC * e;
            float x[BC];
#pragma unroll
            for (int q = 0; q < BC; ++q) x[q] = X[(BC * j + q) * K + c];
#pragma unroll
            for (int a = 0; a < BR; ++a)
#pragma unroll
                for (int q = 0; q < BC; ++q) y[a] += A[a * BC + q] * x[q];
        }
        const size_t o = (size_t)BR * row * K + c;
        if (MODE == kBsrPlain) {
#pragma unroll
            for (int a = 0; a < BR; ++a) Y[o + a * K] = accumulate ? Y[o + a * K] + y[a] : y[a];
        } else if (MODE == kBsrResid) {
#pragma unroll
            for (int a = 0; a < BR; ++a) Y[o + a * K] = Bv[o + a * K] - y[a];
        } else {
            float rv[BR];
#pragma unroll
            for (int a = 0; a < BR; ++a) rv[a] = Bv[o + a * K] - y[a];
            const float *Di = Dinv + (size_t)BR * BR * row;
#pragma unroll
            for (int a = 0; a < BR; ++a) {
                float z = 0.f;
#pragma unroll
                for (int q = 0; q < BR; ++q) z += Di[a * BR + q] * rv[q];
                Y[o + a * K] = X[o + a * K] + s * z;
            }
        }
    }
}

// Single-column (K = 1) form of amg_bsr: a group of G lanes per block row walks the row's
// blocks (consecutive lanes read consecutive blocks), then the group's partial sums are combined
// with cross-lane shuffles and the group's first lane writes the row.
template <int BR, int BC, int G, int MODE>
__global__ void __launch_bounds__(256) amg_bsr1(int nrows, const int32_t *__restrict__ ptr,
                                                const int32_t *__restrict__ col, const float *__restrict__ val,
                                                const float *__restrict__ X, const float *__restrict__ Bv,
                                                const float *__restrict__ Dinv, float *Y, float s, int accumulate) {
    constexpr int RPW = 64 / G;
    const int lane = threadIdx.x & 63;
    const int j = lane % G;
    int lo, hi, jb, nbx;
    gf_xcd_slab(nrows, lo, hi, jb, nbx);
    for (int row0 = lo + (jb * 4 + (int)(threadIdx.x >> 6)) * RPW; row0 < hi; row0 += nbx * 4 * RPW) {
        const int row = row0 + lane / G;
        float y[BR];
#pragma unroll
        for (int a = 0; a < BR; ++a) y[a] = 0.f;
        if (row < hi) {
            for (int e = ptr[row] + j; e < ptr[row + 1]; e += G) {
                const size_t jc = (size_t)col[e];
                const float *A = val + (size_t)BR * BC * e;
                float x[BC];
#pragma unroll
                for (int q = 0; q < BC; ++q) x[q] = X[BC * jc + q];
#pragma unroll
                for (int a = 0; a < BR; ++a)
#pragma unroll
                    for (int q = 0; q < BC; ++q) y[a] += A[a * BC + q] * x[q];
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
            for (int a = 0; a < BR; ++a) y[a] += __shfl_xor(y[a], off, 64);
        if (row < hi && j == 0) {
            const size_t o = (size_t)BR * row;
            if (MODE == kBsrPlain) {
#pragma unroll
                for (int a = 0; a < BR; ++a) Y[o + a] = accumulate ? Y[o + a] + y[a] : y[a];
            } else if (MODE == kBsrResid) {
#pragma unroll
                for (int a = 0; a < BR; ++a) Y[o + a] = Bv[o + a] - y[a];
            } else {
                float rv[BR];
#pragma unroll
                for (int a = 0; a < BR; ++a) rv[a] = Bv[o + a] - y[a];
                const float *Di = Dinv + (size_t)BR * BR * row;
#pragma unroll
                for (int a = 0; a < BR; ++a) {
                    float z = 0.f;
#pragma unroll
                    for (int q = 0; q < BR; ++q) z += Di[a * BR + q] * rv[q];
                    Y[o + a] = X[o + a] + s * z;
                }
            }
        }
    }
}

// r = b - t (the fine residual over a matrix-free product t = A x)
__global__ void __launch_bounds__(256) amg_sub(int64_t n, const float *__restrict__ b, const float *__restrict__ t,
                                               float *r) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) r[i] = b[i] - t[i];
}

// z = x + s Dinv (b - t), 3 x 3 blocks (the fine post-smoothing over a matrix-free t = A x)
__global__ void __launch_bounds__(256) amg_post0(int n, const float *__restrict__ Dinv, const float *__restrict__ b,
                                                 const float *__restrict__ t, const float *__restrict__ x, float *z,
                                                 float s) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float *Di = Dinv + 9 * (size_t)i;
        float rv[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) rv[q] = b[3 * (size_t)i + q] - t[3 * (size_t)i + q];
#pragma unroll
        for (int a = 0; a < 3; ++a)
            z[3 * (size_t)i + a] = x[3 * (size_t)i + a] + s * (Di[3 * a] * rv[0] + Di[3 * a + 1] * rv[1] + Di[3 * a + 2] * rv[2]);
    }
}

// coarsest level, one column: a wave per row of the dense inverse
__global__ void __launch_bounds__(256) amg_dense1(int m, const float *__restrict__ Ainv, const float *__restrict__ Bv,
                                                  float *X) {
    const int lane = threadIdx.x & 63;
    for (int i = blockIdx.x * 4 + (int)(threadIdx.x >> 6); i < m; i += gridDim.x * 4) {
        const float *a = Ainv + (size_t)m * i;
        float acc = 0.f;
        for (int jj = lane; jj < m; jj += 64) acc += a[jj] * Bv[jj];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) X[i] = acc;
    }
}

// pre-smoother from x = 0: x = s Dinv b (thread = (block row, column))
template <int B, int K>
__global__ void __launch_bounds__(256) amg_jacobi0(int n, const float *__restrict__ Dinv, const float *__restrict__ Bv,
                                                   float *X, float s) {
    const int64_t nt = (int64_t)n * K;
    const int r = (int)(threadIdx.x % K);
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nt; t += (int64_t)gridDim.x * 256) {
        const int64_t i = t / K;
        const float *Di = Dinv + (size_t)B * B * i;
        float bv[B];
#pragma unroll
        for (int q = 0; q < B; ++q) bv[q] = Bv[(B * i + q) * K + r];
#pragma unroll
        for (int a = 0; a < B; ++a) {
            float z = 0.f;
#pragma unroll
            for (int q = 0; q < B; ++q) z += Di[a * B + q] * bv[q];
            X[(B * i + a) * K + r] = s * z;
        }
    }
}

// coarsest level: x = Ainv b (dense, m x m)
template <int K>
__global__ void __launch_bounds__(256) amg_dense(int m, const float *__restrict__ Ainv, const float *__restrict__ Bv,
                                                 float *X) {
    const int64_t nt = (int64_t)m * K;
    const int r = (int)(threadIdx.x % K);
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nt; t += (int64_t)gridDim.x * 256) {
        const int64_t i = t / K;
        const float *a = Ainv + (size_t)m * i;
        float acc = 0.f;
        for (int j = 0; j < m; ++j) acc += a[j] * Bv[(size_t)j * K + r];
        X[i * K + r] = acc;
    }
}

// x += alpha p; r -= alpha q; partial sums of r.r (preconditioner applied separately)
template <int K>
__global__ void __launch_bounds__(256) gfk_update_xr(int64_t ndof, const double *__restrict__ scal, float *X,
                                                     float *R, const float *__restrict__ P,
                                                     const float *__restrict__ Q, double *prr) {
    __shared__ double s[256];
    double rr = 0.0;
    const int64_t n = ndof * K;
    const int rc = (int)(threadIdx.x % K);
    const float a = (float)scal[G_ALPHA * 64 + rc];
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        X[t] += a * P[t];
        const float rv = R[t] - a * Q[t];
        R[t] = rv;
        rr += (double)rv * rv;
    }
    s[threadIdx.x] = rr;
    col_reduce(s, K);
    if ((int)threadIdx.x < K) prr[(size_t)blockIdx.x * K + threadIdx.x] = s[threadIdx.x];
}

// per-column partials of a.b (and, when bb is set, of b.b) -- fixed column per thread
template <int K>
__global__ void __launch_bounds__(256) gfk_dot2(int64_t ndof, const float *__restrict__ a, const float *__restrict__ b,
                                                double *pab, double *pbb) {
    __shared__ double s1[256], s2[256];
    double ab = 0.0, bb = 0.0;
    const int64_t n = ndof * K;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        ab += (double)a[t] * b[t];
        bb += (double)b[t] * b[t];
    }
    s1[threadIdx.x] = ab;
    col_reduce(s1, K);
    s2[threadIdx.x] = bb;
    col_reduce(s2, K);
    if ((int)threadIdx.x < K) {
        pab[(size_t)blockIdx.x * K + threadIdx.x] = s1[threadIdx.x];
        if (pbb) pbb[(size_t)blockIdx.x * K + threadIdx.x] = s2[threadIdx.x];
    }
}

// p = z (PCG start)
__global__ void __launch_bounds__(256) gf_copyf(int64_t n, const float *__restrict__ A, float *B) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) B[t] = A[t];
}

// R32 = R64 / nrm[col] ; X32 = 0
__global__ void __launch_bounds__(256) gf_to32(int64_t ndof, int k, const double *__restrict__ R64,
                                               const double *__restrict__ nrm, float *R32, float *X32) {
    const int64_t n = ndof * k;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const double s = nrm[t % k];
        R32[t] = s > 0.0 ? (float)(R64[t] / s) : 0.0f;
        X32[t] = 0.0f;
    }
}

// X64 += nrm[col] * X32
__global__ void __launch_bounds__(256) gf_correct(int64_t ndof, int k, const float *__restrict__ X32,
                                                  const double *__restrict__ nrm, double *X64) {
    const int64_t n = ndof * k;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
        X64[t] += nrm[t % k] * (double)X32[t];
}

// R = B - R (R holds A X)
__global__ void __launch_bounds__(256) gf_resid(int64_t n, const double *__restrict__ B, double *R) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
        R[t] = B[t] - R[t];
}

// column c of a block (handle row order) <-> a local vector (DM order, first ndof entries);
// perm[row] is the DM node of a handle row
__global__ void __launch_bounds__(256) gf_pack(int64_t ndof, int k, int c, const int32_t *__restrict__ perm,
                                               const double *__restrict__ v, double *blk) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < ndof; t += (int64_t)gridDim.x * 256) {
        const int64_t i = t / 3;
        blk[t * k + c] = v[3 * (int64_t)perm[i] + (t - 3 * i)];
    }
}

__global__ void __launch_bounds__(256) gf_unpack(int64_t ndof, int k, int c, const int32_t *__restrict__ perm,
                                                 const double *__restrict__ blk, double *v) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < ndof; t += (int64_t)gridDim.x * 256) {
        const int64_t i = t / 3;
        v[3 * (int64_t)perm[i] + (t - 3 * i)] = blk[t * k + c];
    }
}

// [dof][k] blocks between DM order and handle row order (to_rows: DM -> handle)
template <typename T>
__global__ void __launch_bounds__(256) gf_permute_block(int64_t ndof, int k, const int32_t *__restrict__ perm,
                                                        bool to_rows, const T *__restrict__ A, T *B) {
    const int64_t n = ndof * k;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const int64_t d = t / k, i = d / 3;
        const int64_t o = (3 * (int64_t)perm[i] + (d - 3 * i)) * k + (t - d * k);
        if (to_rows) B[t] = A[o];
        else B[o] = A[t];
    }
}

// v[pos node of i] = jump[i] (v zero elsewhere beforehand)
__global__ void __launch_bounds__(256) gf_scatter_jump(int n, const int32_t *__restrict__ posn,
                                                       const double *__restrict__ jump, double *v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 3; ++c) v[3 * (size_t)posn[i] + c] = jump[3 * (size_t)i + c];
}

__global__ void __launch_bounds__(256) gf_fill(int64_t n, double *a, double v) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) a[t] = v;
}

__global__ void __launch_bounds__(256) gf_mask(int n, const int32_t *__restrict__ dofs, double *v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[dofs[i]] = 0.0;
}

}  // namespace plx

using namespace plx;

struct plx_gf_s {
    plx_dm dm = nullptr;
    KeyPlan *kp = nullptr;        // the first key
    std::vector<KeyPlan *> keys;  // every key of the operator (one per material)
    hipStream_t stream = nullptr;
    int nrows = 0;
    int64_t ndof = 0, nnzb = 0;
    int32_t *d_row_ptr = nullptr, *d_col = nullptr;
    // rows are the merged nodes in Morton order of their coordinates (neighbouring rows then share
    // cache lines of the neighbour gathers): perm[row] is the DM node of a row
    std::vector<int32_t> perm;
    int32_t *d_perm = nullptr;
    std::vector<double> xyz;  // [nrows][3] node coordinates in row order (multigrid rigid-body modes)
    // host copy of the row-ordered operator and its identity-row dofs, kept by the assembled
    // (P2) path for the multigrid set-up (which frees it) instead of a download
    std::vector<int32_t> h_ptr, h_col;
    AmgValF h_val;  // FP32 copy for the multigrid set-up (default-initialised: filled by a download)
    std::vector<uint8_t> h_fixed;
    // P2 handles: the P1 interpolation onto the vertices that the multigrid's first level takes
    // (parents of each row in vertex numbering, weights 1 or 1/2; vertex coordinates)
    int p2_nv = 0;
    std::vector<int32_t> p2_par_ptr, p2_par;
    std::vector<double> p2_parw, p2_vxyz;
    double *d_val64 = nullptr;
    float *d_val32 = nullptr;
    double *d_Binv64 = nullptr;
    float *d_Binv32 = nullptr;
    // block work vectors, sized for kmax columns
    int kmax = 0;
    double *B64 = nullptr, *X64 = nullptr, *R64 = nullptr;
    float *X32 = nullptr, *R32 = nullptr, *P32 = nullptr, *Q32 = nullptr, *Z32 = nullptr;
    double *v1 = nullptr, *v2 = nullptr, *jump = nullptr;  // single local vectors
    double *scal = nullptr, *prz = nullptr, *prr = nullptr, *nrm = nullptr;
    // options
    int block = 0;  // 0: autotune
    double inner_rtol = 1e-4;
    int inner_maxit = 4000;
    bool use_graph = true;
    // captured 10-iteration inner PCG chunk per block width
    hipGraphExec_t graph[65] = {};
    std::vector<double> probe_us;  // per-RHS per-iteration time of each probed width
    // smoothed-aggregation multigrid preconditioner (PLX_GF_PC_AMG)
    int pc = 0;
    // strength threshold of the aggregation: Q1 aggregates over the whole node graph; the P2
    // graph (vertex and edge nodes, up to 65 neighbours a row) over its strong couplings
    double amg_theta = 0.0;
    struct Level {
        int n = 0, b = 0, nc = 0;
        int bc = 6;  // block size of the next level (3 below the P2 -> P1 level)
        int32_t *Aptr = nullptr, *Acol = nullptr, *Pptr = nullptr, *Pcol = nullptr, *Rptr = nullptr, *Rcol = nullptr;
        float *Aval = nullptr, *Pval = nullptr, *Rval = nullptr, *Dinv = nullptr;
        float lmax = 0.f;
        float *r = nullptr, *d = nullptr, *bv = nullptr, *xv = nullptr;  // workspace, 64 columns
    };
    std::vector<Level> lv;
    int coarse_m = 0;
    float *coarse_inv = nullptr, *cb = nullptr, *cx = nullptr;
    float *vr1 = nullptr, *vz1 = nullptr;  // single-column V-cycle operands in row order
    // the finest level's transfer operators and inverse diagonal blocks with DM-ordered fine
    // indices, for the single-column V-cycle over a matrix-free fine operator (fine_op)
    int32_t *Rcol_dm = nullptr, *Pptr_dm = nullptr, *Pcol_dm = nullptr;
    float *Pval_dm = nullptr, *Dinv_dm = nullptr;
    GfFineOp fine_op = nullptr;
    void *fine_ctx = nullptr;
    int amg_kmax = 0;
    uint64_t op_epoch = 0;  // the DM's operator epoch the assembled copies were made at
    // continuous batching (plx_gf_slots_*): per-column solves that start and retire on their own
    struct Slot {
        int state = 0;        // PLX_SLOT_EMPTY / RUNNING / CONVERGED / FAILED
        int outer = 0, inner = 0;
        int ready_since = -1; // step at which its inner solve converged (-1: not yet)
        double b0 = 0.0, scale = 0.0, relres = 0.0;
    };
    std::vector<Slot> slots;
    int slots_k = 0, slots_maxit = 0, slots_step = 0;
    double slots_rtol = 0.0;
    double *d_rr0 = nullptr;       // [64] r.r of each column at its inner PCG start
    double *d_slot_slip = nullptr; // [k][num_lag][3] each slot's slip
};

namespace plx {

static inline dim3 gf_grid(int64_t n) { return dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65535)); }

static int gf_alloc(plx_gf g, int k) {
    if (k <= g->kmax) return 0;
    plx_dm dm = g->dm;
    for (double *p : {g->B64, g->X64, g->R64})
        if (p) (void)hipFree(p);
    for (float *p : {g->X32, g->R32, g->P32, g->Q32, g->Z32})
        if (p) (void)hipFree(p);
    for (int i = 0; i <= 64; ++i)
        if (g->graph[i]) {
            (void)hipGraphExecDestroy(g->graph[i]);
            g->graph[i] = nullptr;
        }
    const size_t n = (size_t)g->ndof * k;
    for (double **p : {&g->B64, &g->X64, &g->R64}) PLX_HIP(dm, hipMalloc(p, n * sizeof(double)));
    for (float **p : {&g->X32, &g->R32, &g->P32, &g->Q32, &g->Z32}) PLX_HIP(dm, hipMalloc(p, n * sizeof(float)));
    g->kmax = k;
    return 0;
}

template <typename Real>
static void gf_spmm_launch(plx_gf g, int k, const Real *val, const Real *X, Real *Y) {
    const int64_t n = (int64_t)g->nrows * k;
    hipLaunchKernelGGL(gf_spmm<Real>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g->stream, g->nrows, k,
                       g->d_row_ptr, g->d_col, val, X, Y);
}

// q = A p and the per-column partials of p.q for the k-column block
template <int K>
static void gf_spmm_dot_launch(plx_gf g) {
    if (K == 64)
        hipLaunchKernelGGL(gfk_spmm_dot64, dim3(kGfRed), dim3(256), 0, g->stream, g->nrows, g->d_row_ptr, g->d_col,
                           g->d_val32, g->P32, g->Q32, g->prz);
    else
        hipLaunchKernelGGL(gfk_spmm_dot<K>, dim3(kGfRed), dim3(256), 0, g->stream, g->nrows, g->d_row_ptr, g->d_col,
                           g->d_val32, g->P32, g->Q32, g->prz);
}

// Ten iterations of the FP32 block PCG (p, z, scalars carried on the device): per iteration
// q = A p with p.q fused, the x/r/z update with r.z and r.r fused, and p = z + beta p.
template <int K>
static void gf_pcg_chunk_k(plx_gf g) {
    for (int it = 0; it < 10; ++it) {
        gf_spmm_dot_launch<K>(g);
        hipLaunchKernelGGL(gf_finish_alpha, dim3(K), dim3(256), 0, g->stream, K, g->prz, g->scal);
        hipLaunchKernelGGL((gfk_update<K, false>), dim3(kGfRed), dim3(256), 0, g->stream, g->nrows, g->d_Binv32,
                           g->scal, g->X32, g->R32, g->P32, g->Q32, g->Z32, g->prz, g->prr);
        hipLaunchKernelGGL(gf_finish_beta, dim3(K), dim3(256), 0, g->stream, K, g->prz, g->prr, g->scal, 0);
        hipLaunchKernelGGL(gfk_pupdate<K>, dim3(kGfRed), dim3(256), 0, g->stream, g->ndof * K, g->scal, g->Z32,
                           g->P32);
    }
}

template <int K>
static void gf_pcg_chunk_amg(plx_gf g);
template <int K>
static void gf_pcg_start_amg(plx_gf g);

static void gf_pcg_chunk(plx_gf g, int k) {
    if (g->pc == PLX_GF_PC_AMG) {
        switch (k) {
        case 8: gf_pcg_chunk_amg<8>(g); break;
        case 16: gf_pcg_chunk_amg<16>(g); break;
        case 32: gf_pcg_chunk_amg<32>(g); break;
        default: gf_pcg_chunk_amg<64>(g); break;
        }
        return;
    }
    switch (k) {
    case 8: gf_pcg_chunk_k<8>(g); break;
    case 16: gf_pcg_chunk_k<16>(g); break;
    case 32: gf_pcg_chunk_k<32>(g); break;
    default: gf_pcg_chunk_k<64>(g); break;
    }
}

// PCG start: z = Binv r, p = z, r.z and r.r (X32 already zero)
static void gf_pcg_start(plx_gf g, int k) {
    if (g->pc == PLX_GF_PC_AMG) {
        switch (k) {
        case 8: gf_pcg_start_amg<8>(g); break;
        case 16: gf_pcg_start_amg<16>(g); break;
        case 32: gf_pcg_start_amg<32>(g); break;
        default: gf_pcg_start_amg<64>(g); break;
        }
        return;
    }
#define PLX_GF_START(K)                                                                                      \
    hipLaunchKernelGGL((gfk_update<K, true>), dim3(kGfRed), dim3(256), 0, g->stream, g->nrows, g->d_Binv32, \
                       g->scal, g->X32, g->R32, g->P32, g->Q32, g->Z32, g->prz, g->prr)
    switch (k) {
    case 8: PLX_GF_START(8); break;
    case 16: PLX_GF_START(16); break;
    case 32: PLX_GF_START(32); break;
    default: PLX_GF_START(64); break;
    }
#undef PLX_GF_START
    hipLaunchKernelGGL(gf_finish_beta, dim3(k), dim3(256), 0, g->stream, k, g->prz, g->prr, g->scal, 1);
}

// ---------------------------------------------------------------------------------------------
// Multigrid V-cycle: z = M r on the k-column block (all launches on the handle's stream, no host
// synchronisation: capturable).  Smoother: Chebyshev of degree 1 on [lmax/30, lmax] of D^-1 A,
// i.e. block-Jacobi damped by s = 2 / (lmax + lmax/30), once before and once after the coarse
// correction (a symmetric V-cycle, as PCG needs):
//   x = s D^-1 b;  r = b - A x;  x += P V(R r);  z = x + s D^-1 (b - A x)
// ---------------------------------------------------------------------------------------------
static dim3 amg_grid(int nrows, int rows_per_wave) {
    const int64_t waves = ((int64_t)nrows + rows_per_wave - 1) / rows_per_wave;
    int64_t blocks = (waves + 3) / 4;
    blocks = std::min<int64_t>(kGfRed, std::max<int64_t>(8, (blocks + 7) / 8 * 8));
    return dim3((unsigned)blocks);
}

// G: lanes per block row of the single-column kernel (operator rows ~25-50 blocks, restriction
// rows up to a few hundred, prolongation rows ~4-8)
template <int BR, int BC, int K, int MODE, int G = 16>
static void amg_bsr_launch(plx_gf g, int nrows, const int32_t *ptr, const int32_t *col, const float *val,
                           const float *X, const float *Bv, const float *Dinv, float *Y, float s, int accumulate) {
    if (K == 1)
        hipLaunchKernelGGL((amg_bsr1<BR, BC, G, MODE>), amg_grid(nrows, 64 / G), dim3(256), 0, g->stream, nrows, ptr,
                           col, val, X, Bv, Dinv, Y, s, accumulate);
    else
        hipLaunchKernelGGL((amg_bsr<BR, BC, K, MODE>), amg_grid(nrows, 64 / K), dim3(256), 0, g->stream, nrows, ptr,
                           col, val, X, Bv, Dinv, Y, s, accumulate);
}

template <int B, int K>
static void amg_level(plx_gf g, int l, const float *bv, float *xv);

template <int K>
static void amg_vcycle(plx_gf g, int l, const float *bv, float *xv) {
    if (l == (int)g->lv.size()) {
        if (K == 1)
            hipLaunchKernelGGL(amg_dense1, dim3((g->coarse_m + 3) / 4), dim3(256), 0, g->stream, g->coarse_m,
                               g->coarse_inv, bv, xv);
        else
            hipLaunchKernelGGL(amg_dense<K>, dim3(std::max(1, std::min(kGfRed, (g->coarse_m * K + 255) / 256))),
                               dim3(256), 0, g->stream, g->coarse_m, g->coarse_inv, bv, xv);
        return;
    }
    if (g->lv[l].b == 3) amg_level<3, K>(g, l, bv, xv);
    else amg_level<6, K>(g, l, bv, xv);
}

template <int B, int K>
static void amg_level(plx_gf g, int l, const float *bv, float *xv) {
    const plx_gf_s::Level &L = g->lv[l];
    const int nl = (int)g->lv.size();
    const int32_t *Ap = l == 0 ? g->d_row_ptr : L.Aptr;
    const int32_t *Ac = l == 0 ? g->d_col : L.Acol;
    const float *Av = l == 0 ? g->d_val32 : L.Aval;
    const float s = (float)(2.0 / (L.lmax + L.lmax / 30.0));
    float *x = L.d, *r = L.r;
    float *nb = l + 1 < nl ? g->lv[l + 1].bv : g->cb;
    float *nx = l + 1 < nl ? g->lv[l + 1].xv : g->cx;
    hipLaunchKernelGGL((amg_jacobi0<B, K>), gf_grid((int64_t)L.n * K), dim3(256), 0, g->stream, L.n, L.Dinv, bv, x, s);
    amg_bsr_launch<B, B, K, kBsrResid>(g, L.n, Ap, Ac, Av, x, bv, nullptr, r, 0.f, 0);
    if (B == 3 && L.bc == 3)  // the P2 -> P1 level: 3 x 3 transfer blocks (~15 children a vertex)
        amg_bsr_launch<3, 3, K, kBsrPlain, 16>(g, L.nc, L.Rptr, L.Rcol, L.Rval, r, nullptr, nullptr, nb, 0.f, 0);
    else
        amg_bsr_launch<6, B, K, kBsrPlain, 64>(g, L.nc, L.Rptr, L.Rcol, L.Rval, r, nullptr, nullptr, nb, 0.f, 0);
    amg_vcycle<K>(g, l + 1, nb, nx);
    if (B == 3 && L.bc == 3)  // one or two parents a node
        amg_bsr_launch<3, 3, K, kBsrPlain, 2>(g, L.n, L.Pptr, L.Pcol, L.Pval, nx, nullptr, nullptr, x, 0.f, 1);
    else
        amg_bsr_launch<B, 6, K, kBsrPlain, 8>(g, L.n, L.Pptr, L.Pcol, L.Pval, nx, nullptr, nullptr, x, 0.f, 1);
    amg_bsr_launch<B, B, K, kBsrPost>(g, L.n, Ap, Ac, Av, x, bv, L.Dinv, xv, s, 0);
}

// kAmgChunk iterations of the multigrid-preconditioned FP32 block PCG (few: the V-cycle
// converges in a handful of iterations, and a convergence check costs one small copy).
constexpr int kAmgChunk = 2;
template <int K>
static void gf_pcg_chunk_amg(plx_gf g) {
    for (int it = 0; it < kAmgChunk; ++it) {
        gf_spmm_dot_launch<K>(g);
        hipLaunchKernelGGL(gf_finish_alpha, dim3(K), dim3(256), 0, g->stream, K, g->prz, g->scal);
        hipLaunchKernelGGL(gfk_update_xr<K>, dim3(kGfRed), dim3(256), 0, g->stream, g->ndof, g->scal, g->X32, g->R32,
                           g->P32, g->Q32, g->prr);
        amg_vcycle<K>(g, 0, g->R32, g->Z32);
        hipLaunchKernelGGL(gfk_dot2<K>, dim3(kGfRed), dim3(256), 0, g->stream, g->ndof, g->Z32, g->R32, g->prz,
                           (double *)nullptr);
        hipLaunchKernelGGL(gf_finish_beta, dim3(K), dim3(256), 0, g->stream, K, g->prz, g->prr, g->scal, 0);
        hipLaunchKernelGGL(gfk_pupdate<K>, dim3(kGfRed), dim3(256), 0, g->stream, g->ndof * K, g->scal, g->Z32,
                           g->P32);
    }
}

// start: z = M r, p = z, r.z and r.r
template <int K>
static void gf_pcg_start_amg(plx_gf g) {
    amg_vcycle<K>(g, 0, g->R32, g->Z32);
    hipLaunchKernelGGL(gf_copyf, gf_grid(g->ndof * K), dim3(256), 0, g->stream, g->ndof * K, g->Z32, g->P32);
    hipLaunchKernelGGL(gfk_dot2<K>, dim3(kGfRed), dim3(256), 0, g->stream, g->ndof, g->Z32, g->R32, g->prz, g->prr);
    hipLaunchKernelGGL(gf_finish_beta, dim3(K), dim3(256), 0, g->stream, K, g->prz, g->prr, g->scal, 1);
}

static int gf_run_chunk(plx_gf g, int k) {
    plx_dm dm = g->dm;
    if (!g->use_graph) {
        gf_pcg_chunk(g, k);
        PLX_HIP(dm, hipGetLastError());
        return 0;
    }
    if (!g->graph[k]) {
        hipGraph_t graph = nullptr;
        PLX_HIP(dm, hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
        gf_pcg_chunk(g, k);
        PLX_HIP(dm, hipStreamEndCapture(g->stream, &graph));
        PLX_HIP(dm, hipGraphInstantiate(&g->graph[k], graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
    }
    PLX_HIP(dm, hipGraphLaunch(g->graph[k], g->stream));
    return 0;
}

// FP32 block PCG on A X32 = R32 from X32 = 0; all columns to inner_rtol (normalized rhs).
static int gf_inner(plx_gf g, int k, int *its) {
    plx_dm dm = g->dm;
    gf_pcg_start(g, k);
    PLX_HIP(dm, hipGetLastError());
    std::vector<double> rr(k), rr0(k);
    PLX_HIP(dm, hipMemcpyAsync(rr0.data(), g->scal + G_RR * 64, k * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    PLX_HIP(dm, hipStreamSynchronize(g->stream));
    int it = 0;
    while (it < g->inner_maxit) {
        bool done = true;
        for (int r = 0; r < k; ++r)
            if (rr0[r] > 0.0 && (it == 0 ? rr0[r] : rr[r]) > g->inner_rtol * g->inner_rtol * rr0[r]) done = false;
        if (done) break;
        int err = gf_run_chunk(g, k);
        if (err) return err;
        it += g->pc == PLX_GF_PC_AMG ? kAmgChunk : 10;
        PLX_HIP(dm, hipMemcpyAsync(rr.data(), g->scal + G_RR * 64, k * sizeof(double), hipMemcpyDeviceToHost,
                                   g->stream));
        PLX_HIP(dm, hipStreamSynchronize(g->stream));
    }
    *its = it;
    return 0;
}

// Column norms of an FP64 block -> g->nrm (device) and host copy.
static int gf_norms(plx_gf g, int k, const double *V, std::vector<double> &out) {
    plx_dm dm = g->dm;
    hipLaunchKernelGGL(gf_dot<double>, dim3(kGfRed), dim3(256), 0, g->stream, g->ndof, k, V, V, g->prz);
    hipLaunchKernelGGL(gf_finish_norm, dim3(k), dim3(256), 0, g->stream, k, g->prz, g->nrm);
    PLX_HIP(dm, hipGetLastError());
    out.resize(k);
    PLX_HIP(dm, hipMemcpyAsync(out.data(), g->nrm, k * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    PLX_HIP(dm, hipStreamSynchronize(g->stream));
    return 0;
}

// Solve A_m X64 = B64 (k columns) by FP64 iterative refinement around the FP32 block PCG.
static int gf_block_solve(plx_gf g, int k, double rtol, int maxit, plx_gf_stats *st) {
    plx_dm dm = g->dm;
    const int64_t n = g->ndof * k;
    hipLaunchKernelGGL(gf_fill, gf_grid(n), dim3(256), 0, g->stream, n, g->X64, 0.0);
    std::vector<double> b0, rn;
    int err = gf_norms(g, k, g->B64, b0);
    if (err) return err;
    // R = B (X = 0)
    PLX_HIP(dm, hipMemcpyAsync(g->R64, g->B64, n * sizeof(double), hipMemcpyDeviceToDevice, g->stream));
    rn = b0;
    int outer = 0;
    for (;;) {
        bool done = true;
        for (int r = 0; r < k; ++r)
            if (rn[r] > rtol * b0[r]) done = false;
        if (done || outer >= maxit) break;
        if ((err = gf_norms(g, k, g->R64, rn))) return err;
        hipLaunchKernelGGL(gf_to32, gf_grid(n), dim3(256), 0, g->stream, g->ndof, k, g->R64, g->nrm, g->R32, g->X32);
        int its = 0;
        if ((err = gf_inner(g, k, &its))) return err;
        st->inner_iterations += its;
        hipLaunchKernelGGL(gf_correct, gf_grid(n), dim3(256), 0, g->stream, g->ndof, k, g->X32, g->nrm, g->X64);
        gf_spmm_launch<double>(g, k, g->d_val64, g->X64, g->R64);
        hipLaunchKernelGGL(gf_resid, gf_grid(n), dim3(256), 0, g->stream, n, g->B64, g->R64);
        PLX_HIP(dm, hipGetLastError());
        if ((err = gf_norms(g, k, g->R64, rn))) return err;
        ++outer;
    }
    st->outer_iterations += outer;
    double worst = 0.0;
    for (int r = 0; r < k; ++r)
        if (b0[r] > 0.0) worst = std::max(worst, rn[r] / b0[r]);
    st->max_relative_residual = std::max(st->max_relative_residual, worst);
    return 0;
}

static void gf_amg_free(plx_gf g) {
    for (auto &L : g->lv)
        for (void *p : {(void *)L.Aptr, (void *)L.Acol, (void *)L.Aval, (void *)L.Pptr, (void *)L.Pcol, (void *)L.Pval,
                        (void *)L.Rptr, (void *)L.Rcol, (void *)L.Rval, (void *)L.Dinv, (void *)L.r, (void *)L.d,
                        (void *)L.bv, (void *)L.xv})
            if (p) (void)hipFree(p);
    g->lv.clear();
    for (void *p : {(void *)g->coarse_inv, (void *)g->cb, (void *)g->cx, (void *)g->vr1, (void *)g->vz1,
                    (void *)g->Rcol_dm, (void *)g->Pptr_dm, (void *)g->Pcol_dm, (void *)g->Pval_dm,
                    (void *)g->Dinv_dm})
        if (p) (void)hipFree(p);
    g->coarse_inv = g->cb = g->cx = g->vr1 = g->vz1 = g->Pval_dm = g->Dinv_dm = nullptr;
    g->Rcol_dm = g->Pptr_dm = g->Pcol_dm = nullptr;
    g->coarse_m = 0;
}

// Build the multigrid hierarchy of the merged FP64 operator on the host and upload it in FP32
// with workspace for 64 columns.
// lambda_max estimate of D^-1 A on the device (A: the handle's FP64 operator, D^-1: its inverse
// diagonal blocks): twelve power iterations from the host set-up's start vector, times 1.1 (the
// host's level_lmax on the device).
static int gf_lmax_device(plx_gf g, double *out) {
    plx_dm dm = g->dm;
    const int n = g->nrows;
    const int nblk = (n + 255) / 256;
    const size_t nd = 3 * (size_t)n;
    std::vector<double> hv = lmax_start_vector(nd), part(nblk);
    double vn = 0.0;
    for (double x : hv) vn += x * x;
    vn = std::sqrt(vn);
    double *v = nullptr, *y = nullptr, *dpart = nullptr;
    int err = 0;
    if (hipMalloc(&v, nd * sizeof(double)) != hipSuccess || hipMalloc(&y, nd * sizeof(double)) != hipSuccess ||
        hipMalloc(&dpart, nblk * sizeof(double)) != hipSuccess ||
        hipMemcpy(v, hv.data(), nd * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        err = set_error(dm, 76, "lambda_max estimate: device buffers");
    double lam = 0.0;
    for (int it = 0; it < 12 && !err; ++it) {
        hipLaunchKernelGGL(gf_spmm<double>, dim3(nblk), dim3(256), 0, g->stream, n, 1, g->d_row_ptr, g->d_col,
                           g->d_val64, v, y);
        hipLaunchKernelGGL(gf_binv_step, dim3(nblk), dim3(256), 0, g->stream, n, g->d_Binv64, y, 1.0 / vn, v, dpart);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(part.data(), dpart, nblk * sizeof(double), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
            hipStreamSynchronize(g->stream) != hipSuccess) {
            err = set_error(dm, 76, "lambda_max estimate failed");
            break;
        }
        double acc = 0.0;
        for (double x : part) acc += x;
        const double wn = std::sqrt(acc) / vn;  // || D^-1 A v || / || v ||
        lam = wn;
        vn = wn;
    }
    (void)hipFree(v);
    (void)hipFree(y);
    (void)hipFree(dpart);
    if (!err) *out = 1.1 * lam;
    return err;
}

static int gf_amg_setup(plx_gf g) {
    plx_dm dm = g->dm;
    const int nn = g->nrows;
    std::vector<int32_t> ptr, col;
    AmgVal val;    // FP64 operator (downloaded here: Q1 handles)
    AmgValF valf;  // FP32 operator (the assembled P2 handle's host copy)
    std::vector<uint8_t> fixed;
    if (!g->h_val.empty()) {  // the assembled path's host copy
        if (g->p2_nv <= 0) return set_error(dm, 1, "assembled operator copy without a P2 hierarchy");
        ptr.swap(g->h_ptr);
        col.swap(g->h_col);
        valf.swap(g->h_val);
        fixed.swap(g->h_fixed);
    } else {
        ptr.resize(nn + 1);
        col.resize(g->nnzb);
        val.resize((size_t)g->nnzb * 9);
        PLX_HIP(dm, hipMemcpy(ptr.data(), g->d_row_ptr, ptr.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        PLX_HIP(dm, hipMemcpy(col.data(), g->d_col, col.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        PLX_HIP(dm, hipMemcpy(val.data(), g->d_val64, val.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    // fixed dofs: identity rows of the merged operator (constrained dofs, positive split nodes)
    if (fixed.empty()) {
        fixed.assign(3 * (size_t)nn, 0);
        for (int i = 0; i < nn; ++i)
            for (int c = 0; c < 3; ++c) {
                bool ident = true;
                for (int e = ptr[i]; e < ptr[i + 1] && ident; ++e)
                    for (int d = 0; d < 3; ++d) {
                        const double v = val[(size_t)e * 9 + c * 3 + d];
                        if (v != ((col[e] == i && d == c) ? 1.0 : 0.0)) {
                            ident = false;
                            break;
                        }
                    }
                fixed[3 * (size_t)i + c] = ident;
            }
    }
    AmgHierarchy h;
    std::string msg;
    if (g->xyz.size() != 3 * (size_t)nn) return set_error(dm, 1, "node coordinates missing");
    double lmax0 = 0.0;  // the P2 level's estimate, from the device operator
    if (g->p2_nv > 0) {
        const int lerr = gf_lmax_device(g, &lmax0);
        if (lerr) return lerr;
    }
    const int aerr = g->p2_nv > 0
                         ? amg_setup_p2(nn, ptr.data(), col.data(), valf.data(), fixed.data(), g->p2_nv,
                                        g->p2_par_ptr.data(), g->p2_par.data(), g->p2_parw.data(),
                                        g->p2_vxyz.data(), 768, 8, &h, &msg, g->amg_theta, lmax0)
                         : amg_setup(nn, ptr.data(), col.data(), val.data(), g->xyz.data(), fixed.data(), 768, 8,
                                     &h, &msg, g->amg_theta);
    if (aerr) return set_error(dm, 1, msg.c_str());
    gf_amg_free(g);
    auto upf = [&](float **d, const auto &v) -> int {
        std::vector<float> f(v.begin(), v.end());
        PLX_HIP(dm, hipMalloc((void **)d, std::max<size_t>(1, f.size()) * sizeof(float)));
        PLX_HIP(dm, hipMemcpy(*d, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
        return 0;
    };
    auto upi = [&](int32_t **d, const auto &v) -> int {
        PLX_HIP(dm, hipMalloc((void **)d, std::max<size_t>(1, v.size()) * sizeof(int32_t)));
        PLX_HIP(dm, hipMemcpy(*d, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        return 0;
    };
    auto ws = [&](float **d, size_t n) -> int {
        PLX_HIP(dm, hipMalloc((void **)d, std::max<size_t>(1, n) * 64 * sizeof(float)));
        return 0;
    };
    g->lv.resize(h.levels.size());
    int err = 0;
    for (size_t l = 0; l < h.levels.size(); ++l) {
        const AmgLevel &H = h.levels[l];
        plx_gf_s::Level &L = g->lv[l];
        L.n = H.n;
        L.b = H.b;
        L.nc = H.nc;
        L.bc = H.bc;
        L.lmax = (float)H.lmax;
        if (l > 0 && ((err = upi(&L.Aptr, H.Aptr)) || (err = upi(&L.Acol, H.Acol)) || (err = upf(&L.Aval, H.Aval))))
            return err;
        if ((err = upi(&L.Pptr, H.Pptr)) || (err = upi(&L.Pcol, H.Pcol)) || (err = upf(&L.Pval, H.Pval)) ||
            (err = upi(&L.Rptr, H.Rptr)) || (err = upi(&L.Rcol, H.Rcol)) || (err = upf(&L.Rval, H.Rval)) ||
            (err = upf(&L.Dinv, H.Dinv)))
            return err;
        const size_t nd = (size_t)H.n * H.b;
        if ((err = ws(&L.r, nd)) || (err = ws(&L.d, nd))) return err;
        if (l > 0 && ((err = ws(&L.bv, nd)) || (err = ws(&L.xv, nd)))) return err;
    }
    g->coarse_m = h.coarse_n;
    if ((err = upf(&g->coarse_inv, h.coarse_inv)) || (err = ws(&g->cb, h.coarse_n)) || (err = ws(&g->cx, h.coarse_n)))
        return err;
    PLX_HIP(dm, hipMalloc((void **)&g->vr1, (size_t)g->ndof * sizeof(float)));
    PLX_HIP(dm, hipMalloc((void **)&g->vz1, (size_t)g->ndof * sizeof(float)));
    {  // the finest level in DM order: Dinv and P rows moved to their nodes, R columns renamed
        const AmgLevel &H = h.levels[0];
        std::vector<int32_t> iperm(nn);
        for (int r = 0; r < nn; ++r) iperm[g->perm[r]] = r;
        std::vector<double> dinv(H.Dinv.size());
        for (int v = 0; v < nn; ++v)
            std::copy(&H.Dinv[9 * (size_t)iperm[v]], &H.Dinv[9 * (size_t)iperm[v]] + 9, &dinv[9 * (size_t)v]);
        std::vector<int32_t> rcol(H.Rcol.size());
        for (size_t k = 0; k < rcol.size(); ++k) rcol[k] = g->perm[H.Rcol[k]];
        std::vector<int32_t> pptr(nn + 1, 0), pcol(H.Pcol.size());
        std::vector<double> pval(H.Pval.size());
        const size_t bsz = (size_t)H.b * H.bc;  // values per prolongator block
        for (int v = 0; v < nn; ++v) {
            const int r = iperm[v], e0 = H.Pptr[r], e1 = H.Pptr[r + 1];
            pptr[v + 1] = pptr[v] + (e1 - e0);
            std::copy(H.Pcol.begin() + e0, H.Pcol.begin() + e1, pcol.begin() + pptr[v]);
            std::copy(H.Pval.begin() + bsz * e0, H.Pval.begin() + bsz * e1, pval.begin() + bsz * pptr[v]);
        }
        if ((err = upf(&g->Dinv_dm, dinv)) || (err = upi(&g->Rcol_dm, rcol)) || (err = upi(&g->Pptr_dm, pptr)) ||
            (err = upi(&g->Pcol_dm, pcol)) || (err = upf(&g->Pval_dm, pval)))
            return err;
    }
    return 0;
}

int gf_setup(plx_gf g, int nc, const std::vector<int32_t> &row_ptr, const std::vector<int32_t> &col,
             const std::vector<int32_t> &inc_ptr, const std::vector<int32_t> &inc,
             const std::vector<int32_t> &mcl, const std::vector<int32_t> &diag_pos,
             const std::vector<uint8_t> &conmask, const std::vector<uint8_t> &is_pos);
int gf_alloc_work(plx_gf g);

// Rows of the assembled operator in Morton order: block o of row r is block src[o] of the DM-order
// operator val0, negated into -J; a constrained dof's row and column become identity / zero
// (con: 1 per constrained dof, DM order).  One wavefront per row, lanes over its 9 x blocks values.
__global__ void __launch_bounds__(256) gf_permute_blocks(int nn, const int32_t *__restrict__ rp,
                                                         const int32_t *__restrict__ col,
                                                         const int32_t *__restrict__ perm,
                                                         const int32_t *__restrict__ src,
                                                         const uint8_t *__restrict__ con,
                                                         const double *__restrict__ val0, double *val) {
    const int r = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (r >= nn) return;
    const int lane = threadIdx.x & 63;
    const int i = perm[r];
    const int o0 = rp[r], nv = 9 * (rp[r + 1] - o0);
    for (int v = lane; v < nv; v += 64) {
        const int q = v / 9, fh = v - 9 * q, f = fh / 3, h = fh - 3 * f;
        const size_t o = (size_t)o0 + q;
        const int c = col[o];
        const int j = perm[c];
        double x = -val0[9 * (size_t)src[o] + fh];
        if (con[3 * (size_t)i + f] || con[3 * (size_t)j + h]) x = (c == r && f == h) ? 1.0 : 0.0;
        val[9 * o + fh] = x;
    }
}

// Handle over an operator whose Morton-order rows are already on the device (d_row_ptr, d_col,
// d_perm, d_val64; diagonal block of row i at d_diag[i]): inverse diagonal blocks, FP32 copies,
// workspace.
static int gf_setup_assembled(plx_gf g, int32_t *d_diag) {
    plx_dm dm = g->dm;
    const int nn = g->nrows;
    PLX_HIP(dm, hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    PLX_HIP(dm, hipMalloc(&g->d_val32, (size_t)g->nnzb * 9 * sizeof(float)));
    PLX_HIP(dm, hipMalloc(&g->d_Binv64, 9 * (size_t)nn * sizeof(double)));
    PLX_HIP(dm, hipMalloc(&g->d_Binv32, 9 * (size_t)nn * sizeof(float)));
    hipLaunchKernelGGL(gf_diag_kernel, dim3((nn + 255) / 256), dim3(256), 0, g->stream, nn, d_diag, g->d_val64,
                       g->d_Binv64, g->d_Binv32, g->d_val32, g->nnzb);
    PLX_HIP(dm, hipGetLastError());
    PLX_HIP(dm, hipStreamSynchronize(g->stream));
    return gf_alloc_work(g);
}

// Operator handle from the assembled Jacobian of the keys (plx_compute_jacobian, linearised at
// zero displacement), negated into the positive operator -J of the Q1 path: the path for P2
// tetrahedra, whose element matrices the Q1 kernels above do not form.  Rows are the mesh
// nodes in Morton order of their coordinates (an edge node sits at its edge's midpoint); a
// constrained dof gets an identity row and a zero column.  The handle also records each row's
// P1 parents (its vertex, or its edge's two ends) for the multigrid's P2 -> P1 first level.  It
// serves products and the multigrid preconditioner; fault impulses need the Q1 handle.
static int gf_create_assembled(plx_gf *out, plx_dm dm, const std::vector<KeyPlan *> &keys) {
    if (dm->fault) return set_error(dm, 1, "Green's functions with a fault are implemented for Q1 hexes");
    const int nn = dm->num_nodes, nb = dm->nb, nv = dm->nv;
    int64_t nnzb = 0;
    int err = plx_jacobian_pattern(dm, &nnzb, nullptr, nullptr);  // builds the DM's pattern once
    if (err) return err;
    const int32_t *rp0 = dm->jac_row_ptr.data(), *col0 = dm->jac_col.data();
    // device buffers of the set-up, freed on every exit
    struct DevBuf {
        void *p = nullptr;
        ~DevBuf() {
            if (p) (void)hipFree(p);
        }
    } val0_buf, src_buf, con_buf, diag_buf;
    {  // the operator in DM order, assembled on the device and kept there
        double *d_x = nullptr;
        if (hipMalloc(&val0_buf.p, (size_t)nnzb * 9 * sizeof(double)) != hipSuccess ||
            hipMalloc(&d_x, std::max<int64_t>(1, dm->local_size) * sizeof(double)) != hipSuccess)
            return set_error(dm, 1, "out of device memory for the assembled operator");
        if (hipMemsetAsync(d_x, 0, dm->local_size * sizeof(double), dm->stream) != hipSuccess)
            err = set_error(dm, 76, "hipMemsetAsync failed");
        for (size_t k = 0; k < keys.size() && !err; ++k)
            err = plx_compute_jacobian(dm, keys[k]->label_value, keys[k]->part, 0.0, 0.0, d_x, nullptr,
                                       static_cast<double *>(val0_buf.p), k == 0 ? PLX_INSERT_VALUES : PLX_ADD_VALUES);
        if (!err && hipStreamSynchronize(dm->stream) != hipSuccess)
            err = set_error(dm, 76, "assembling the operator failed");
        (void)hipFree(d_x);
        if (err) return err;
    }
    // node coordinates: the vertex of a vertex node, the midpoint of an edge node's edge (P2
    // closure order: 4 vertices, then edges 01 02 03 12 13 23)
    static const int kEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    std::vector<double> nx(3 * (size_t)nn, 0.0);
    for (int c = 0; c < dm->num_cells; ++c) {
        const int32_t *cl = &dm->closure[(size_t)c * nb];
        const int32_t *cv = &dm->cone[(size_t)c * nv];
        for (int b = 0; b < nb; ++b) {
            const double *p = nullptr, *q = nullptr;
            if (b < nv) {
                p = q = &dm->coords[3 * (size_t)cv[b]];
            } else if (dm->cell_type == PLX_TET_P2) {
                p = &dm->coords[3 * (size_t)cv[kEdge[b - 4][0]]];
                q = &dm->coords[3 * (size_t)cv[kEdge[b - 4][1]]];
            } else {
                return set_error(dm, 1, "no node coordinates for this cell type");
            }
            for (int d = 0; d < 3; ++d) nx[3 * (size_t)cl[b] + d] = 0.5 * (p[d] + q[d]);
        }
    }
    std::vector<int32_t> perm = morton_permutation(nx.data(), nn), iperm(nn);
    for (int i = 0; i < nn; ++i) iperm[perm[i]] = i;
    std::vector<uint8_t> conmask(3 * (size_t)nn, 0);
    for (int32_t d : dm->con.dofs) {
        if (d < 0 || d >= 3 * nn) return set_error(dm, 1, "constrained dof out of range");
        conmask[d] = 1;
    }
    // rows in Morton order, columns renumbered and ascending; src: each block's position in the
    // DM-order operator (the values move on the device)
    std::vector<int32_t> row_ptr(nn + 1, 0), col((size_t)nnzb), diag_pos(nn, -1), src((size_t)nnzb);
    for (int r = 0; r < nn; ++r) row_ptr[r + 1] = row_ptr[r] + (rp0[perm[r] + 1] - rp0[perm[r]]);
    {  // rows are independent: chunks of them on host threads
        const int T = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                std::vector<std::pair<int32_t, int32_t>> rowv;
                const int r0 = (int)((int64_t)nn * t / T), r1 = (int)((int64_t)nn * (t + 1) / T);
                for (int r = r0; r < r1; ++r) {
                    const int i = perm[r];
                    rowv.clear();
                    for (int e = rp0[i]; e < rp0[i + 1]; ++e) rowv.push_back({iperm[col0[e]], e});
                    std::sort(rowv.begin(), rowv.end());
                    int64_t o = row_ptr[r];
                    for (const auto &ce : rowv) {
                        col[o] = ce.first;
                        src[o] = ce.second;
                        if (ce.first == r) diag_pos[r] = (int32_t)o;
                        ++o;
                    }
                }
            });
        for (auto &x : th) x.join();
    }
    for (int r = 0; r < nn; ++r)
        if (diag_pos[r] < 0) return set_error(dm, 1, "assembled operator without a diagonal block");
    plx_gf g = new plx_gf_s();
    g->dm = dm;
    g->kp = keys[0];
    g->keys = keys;
    g->nrows = nn;
    g->ndof = 3 * (int64_t)nn;
    g->nnzb = nnzb;
    g->op_epoch = dm->op_epoch;
    g->amg_theta = 0.05;  // of the P1 operator below the P2 -> P1 level
    if (dm->cell_type == PLX_TET_P2) {
        // vertices in row order; a vertex node's parent is its vertex, an edge node's the two ends
        std::vector<int32_t> vidx(nn, -1), pa(2 * (size_t)nn, -1);
        for (int c = 0; c < dm->num_cells; ++c) {
            const int32_t *cl = &dm->closure[(size_t)c * nb];
            for (int v = 0; v < 4; ++v) vidx[iperm[cl[v]]] = 0;
            for (int e = 0; e < 6; ++e) {
                const size_t r = (size_t)iperm[cl[4 + e]];
                pa[2 * r] = iperm[cl[kEdge[e][0]]];
                pa[2 * r + 1] = iperm[cl[kEdge[e][1]]];
            }
        }
        int nv1 = 0;
        for (int r = 0; r < nn; ++r)
            if (vidx[r] == 0) vidx[r] = nv1++;
        g->p2_nv = nv1;
        g->p2_par_ptr.assign(nn + 1, 0);
        g->p2_vxyz.resize(3 * (size_t)nv1);
        for (int r = 0; r < nn; ++r) {
            if (vidx[r] >= 0) {
                g->p2_par.push_back(vidx[r]);
                g->p2_parw.push_back(1.0);
                for (int d = 0; d < 3; ++d) g->p2_vxyz[3 * (size_t)vidx[r] + d] = nx[3 * (size_t)perm[r] + d];
            } else {
                const int a = pa[2 * (size_t)r], b = pa[2 * (size_t)r + 1];
                if (a < 0 || b < 0 || vidx[a] < 0 || vidx[b] < 0) {
                    plx_gf_destroy(g);
                    return set_error(dm, 1, "a P2 node that is neither a vertex nor an edge midpoint");
                }
                g->p2_par.push_back(std::min(vidx[a], vidx[b]));
                g->p2_par.push_back(std::max(vidx[a], vidx[b]));
                g->p2_parw.push_back(0.5);
                g->p2_parw.push_back(0.5);
            }
            g->p2_par_ptr[r + 1] = (int32_t)g->p2_par.size();
        }
    }
    g->xyz.resize(3 * (size_t)nn);
    for (int r = 0; r < nn; ++r)
        for (int d = 0; d < 3; ++d) g->xyz[3 * (size_t)r + d] = nx[3 * (size_t)perm[r] + d];
    g->perm = std::move(perm);
    // upload the pattern, move the blocks into Morton order on the device, then keep a host copy
    // of the result for the multigrid set-up
    auto up = [&](void **d, const void *h, size_t bytes) -> int {
        PLX_HIP(dm, hipMalloc(d, std::max<size_t>(1, bytes)));
        PLX_HIP(dm, hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
        return 0;
    };
    auto fail = [&](int e) {
        plx_gf_destroy(g);
        return e;
    };
    if ((err = up((void **)&g->d_row_ptr, row_ptr.data(), row_ptr.size() * sizeof(int32_t))) ||
        (err = up((void **)&g->d_col, col.data(), col.size() * sizeof(int32_t))) ||
        (err = up((void **)&g->d_perm, g->perm.data(), g->perm.size() * sizeof(int32_t))) ||
        (err = up(&diag_buf.p, diag_pos.data(), diag_pos.size() * sizeof(int32_t))) ||
        (err = up(&src_buf.p, src.data(), src.size() * sizeof(int32_t))) ||
        (err = up(&con_buf.p, conmask.data(), conmask.size())))
        return fail(err);
    std::vector<int32_t>().swap(src);
    if (hipMalloc(&g->d_val64, (size_t)nnzb * 9 * sizeof(double)) != hipSuccess)
        return fail(set_error(dm, 1, "out of device memory for the assembled operator"));
    hipLaunchKernelGGL(gf_permute_blocks, dim3((unsigned)((nn + 3) / 4)), dim3(256), 0, dm->stream, nn, g->d_row_ptr,
                       g->d_col, g->d_perm, static_cast<const int32_t *>(src_buf.p),
                       static_cast<const uint8_t *>(con_buf.p), static_cast<const double *>(val0_buf.p), g->d_val64);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(dm->stream) != hipSuccess)
        return fail(set_error(dm, 76, "reordering the assembled operator failed"));
    (void)hipFree(val0_buf.p);
    val0_buf.p = nullptr;
    if ((err = gf_setup_assembled(g, static_cast<int32_t *>(diag_buf.p)))) return fail(err);
    // the multigrid levels are FP32 on the device: their set-up reads the FP32 copy (half the bytes)
    g->h_val.resize((size_t)nnzb * 9);
    if (hipMemcpy(g->h_val.data(), g->d_val32, g->h_val.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(set_error(dm, 76, "downloading the assembled operator failed"));
    g->h_fixed.resize(3 * (size_t)nn);
    for (int r = 0; r < nn; ++r)
        for (int c = 0; c < 3; ++c) g->h_fixed[3 * (size_t)r + c] = conmask[3 * (size_t)g->perm[r] + c];
    g->h_ptr = std::move(row_ptr);
    g->h_col = std::move(col);
    *out = g;
    return 0;
}

// One V-cycle on a single column, z = M r (3 x num_nodes FP32 vectors in DM node order), queued
// on `stream`; used by the matrix-free solver as its multigrid preconditioner.
int gf_set_fine_operator(plx_gf g, GfFineOp op, void *ctx) {
    if (!g) return 1;
    g->fine_op = op;
    g->fine_ctx = ctx;
    return 0;
}

// The V-cycle with the finest level in DM order over the matrix-free operator:
//   x = s D^-1 r;  t = A x;  x += P V(R (r - t));  t = A x;  z = x + s D^-1 (r - t)
// (the same cycle as amg_level on level 0; constrained dofs stay zero: their rows of D^-1 are
// identity rows, their rows of P zero, and r is zero there)
static int gf_vcycle1_matfree(plx_gf g, const float *r, float *z) {
    const plx_gf_s::Level &L = g->lv[0];
    const int nl = (int)g->lv.size();
    const float s = (float)(2.0 / (L.lmax + L.lmax / 30.0));
    float *x = L.d, *r1 = L.r, *t = g->vr1;
    float *nb = nl > 1 ? g->lv[1].bv : g->cb;
    float *nx = nl > 1 ? g->lv[1].xv : g->cx;
    hipLaunchKernelGGL((amg_jacobi0<3, 1>), gf_grid((int64_t)L.n), dim3(256), 0, g->stream, L.n, g->Dinv_dm, r, x, s);
    int err = g->fine_op(g->fine_ctx, x, t);
    if (err) return err;
    hipLaunchKernelGGL(amg_sub, gf_grid(g->ndof), dim3(256), 0, g->stream, g->ndof, r, (const float *)t, r1);
    if (L.bc == 3) {  // P2 -> P1 transfer
        amg_bsr_launch<3, 3, 1, kBsrPlain, 16>(g, L.nc, L.Rptr, g->Rcol_dm, L.Rval, r1, nullptr, nullptr, nb, 0.f, 0);
        amg_vcycle<1>(g, 1, nb, nx);
        amg_bsr_launch<3, 3, 1, kBsrPlain, 2>(g, L.n, g->Pptr_dm, g->Pcol_dm, g->Pval_dm, nx, nullptr, nullptr, x, 0.f,
                                              1);
    } else {
        amg_bsr_launch<6, 3, 1, kBsrPlain, 64>(g, L.nc, L.Rptr, g->Rcol_dm, L.Rval, r1, nullptr, nullptr, nb, 0.f, 0);
        amg_vcycle<1>(g, 1, nb, nx);
        amg_bsr_launch<3, 6, 1, kBsrPlain, 8>(g, L.n, g->Pptr_dm, g->Pcol_dm, g->Pval_dm, nx, nullptr, nullptr, x, 0.f,
                                              1);
    }
    if ((err = g->fine_op(g->fine_ctx, x, t))) return err;
    hipLaunchKernelGGL(amg_post0, gf_grid(L.n), dim3(256), 0, g->stream, L.n, (const float *)g->Dinv_dm, r,
                       (const float *)t, (const float *)x, z, s);
    PLX_HIP(g->dm, hipGetLastError());
    return 0;
}

int gf_vcycle1(plx_gf g, hipStream_t stream, const float *r, float *z) {
    if (g->lv.empty()) return set_error(g->dm, 1, "multigrid hierarchy not built");
    if (g->fine_op) {
        hipStream_t saved = g->stream;
        g->stream = stream;
        const int err = gf_vcycle1_matfree(g, r, z);
        g->stream = saved;
        return err;
    }
    hipStream_t saved = g->stream;
    g->stream = stream;
    hipLaunchKernelGGL(gf_permute_block<float>, gf_grid(g->ndof), dim3(256), 0, stream, g->ndof, 1, g->d_perm, true,
                       r, g->vr1);
    amg_vcycle<1>(g, 0, g->vr1, g->vz1);
    hipLaunchKernelGGL(gf_permute_block<float>, gf_grid(g->ndof), dim3(256), 0, stream, g->ndof, 1, g->d_perm, false,
                       (const float *)g->vz1, z);
    g->stream = saved;
    PLX_HIP(g->dm, hipGetLastError());
    return 0;
}

}  // namespace plx

extern "C" {

int plx_gf_create(plx_gf *out, plx_dm dm, int label_value, int part) {
    if (!out || !dm) return 1;
    *out = nullptr;
    std::vector<KeyPlan *> keys;
    if (label_value == PLX_ALL_KEYS) {
        for (KeyPlan *k : dm->keys)
            if (k->part == part && k->jf3 && k->has_aux) keys.push_back(k);
    } else if (KeyPlan *k = find_key(dm, label_value, part)) {
        keys.push_back(k);
    }
    if (keys.empty()) return set_error(dm, 1, "no key with Jacobian kernels and aux");
    std::vector<int32_t> allcells;  // the keys' cells, key after key
    for (KeyPlan *k : keys) {
        if (!k->jf3 || !k->has_aux) return set_error(dm, 1, "the key needs Jacobian kernels and aux");
        allcells.insert(allcells.end(), k->cells.begin(), k->cells.end());
    }
    if (allcells.size() != (size_t)dm->num_cells) return set_error(dm, 1, "the keys must cover every cell");
    if (dm->cell_type != PLX_HEX_Q1) return gf_create_assembled(out, dm, keys);
    const int nn = dm->num_nodes, nc = (int)allcells.size();
    // handle rows: the DM's nodes in Morton order of their coordinates (Q1: the nodes are the
    // vertices), perm[row] = node, iperm[node] = row
    std::vector<int32_t> perm, iperm(nn);
    if (dm->coords.size() == 3 * (size_t)nn) {
        perm = morton_permutation(dm->coords.data(), nn);
    } else {
        perm.resize(nn);
        for (int i = 0; i < nn; ++i) perm[i] = i;
    }
    for (int i = 0; i < nn; ++i) iperm[perm[i]] = i;
    // merged[row]: the row a node's equations go to (a positive split node's go to its negative
    // twin's row; its own row keeps only the identity)
    std::vector<int32_t> merged(nn);
    std::vector<uint8_t> is_pos(nn, 0);
    for (int i = 0; i < nn; ++i) merged[i] = i;
    if (dm->fault) {  // without a fault the merged operator is the masked stiffness itself
        const FaultData &F = *dm->fault;
        for (int l = 0; l < F.num_lag; ++l) {
            merged[iperm[F.pos_node[l]]] = iperm[F.neg_node[l]];
            is_pos[iperm[F.pos_node[l]]] = 1;
        }
    }
    std::vector<uint8_t> conmask(3 * (size_t)nn, 0);
    for (int32_t d : dm->con.dofs) {
        if (d >= 3 * nn) return set_error(dm, 1, "constrained Lagrange-multiplier dofs are not supported");
        const int row = iperm[d / 3];
        if (is_pos[row]) return set_error(dm, 1, "constrain the negative-side twin of a split fault node");
        conmask[3 * (size_t)row + d % 3] = 1;
    }
    // merged closures of the key cells (key order, in rows) and row -> (cell, vertex) incidence
    std::vector<int32_t> mcl(8 * (size_t)nc);
    std::vector<int32_t> inc_ptr(nn + 1, 0), inc;
    for (int p = 0; p < nc; ++p)
        for (int a = 0; a < 8; ++a) {
            mcl[8 * (size_t)p + a] = merged[iperm[dm->closure[8 * (size_t)allcells[p] + a]]];
            inc_ptr[mcl[8 * (size_t)p + a] + 1]++;
        }
    for (int i = 0; i < nn; ++i) inc_ptr[i + 1] += inc_ptr[i];
    inc.resize(inc_ptr[nn]);
    {
        std::vector<int32_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
        for (int p = 0; p < nc; ++p)
            for (int a = 0; a < 8; ++a) inc[fill[mcl[8 * (size_t)p + a]]++] = 8 * p + a;
    }
    // BSR pattern: the merged node graph (positive nodes: diagonal only)
    std::vector<int32_t> row_ptr(nn + 1, 0), col, diag_pos(nn);
    std::vector<int32_t> nb;
    for (int i = 0; i < nn; ++i) {
        nb.clear();
        nb.push_back(i);
        for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k) {
            const int p = inc[k] >> 3;
            for (int b = 0; b < 8; ++b) nb.push_back(mcl[8 * (size_t)p + b]);
        }
        std::sort(nb.begin(), nb.end());
        nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
        if (nb.size() > 256) return set_error(dm, 1, "node degree above 256");
        for (int32_t j : nb) {
            if (j == i) diag_pos[i] = (int32_t)col.size();
            col.push_back(j);
        }
        row_ptr[i + 1] = (int32_t)col.size();
    }
    plx_gf g = new plx_gf_s();
    g->dm = dm;
    g->kp = keys[0];
    g->keys = keys;
    g->nrows = nn;
    g->ndof = 3 * (int64_t)nn;
    g->nnzb = (int64_t)col.size();
    g->op_epoch = dm->op_epoch;
    if (dm->coords.size() == 3 * (size_t)nn) {  // else no multigrid (it needs the coordinates)
        g->xyz.resize(3 * (size_t)nn);
        for (int r = 0; r < nn; ++r)
            for (int d = 0; d < 3; ++d) g->xyz[3 * (size_t)r + d] = dm->coords[3 * (size_t)perm[r] + d];
    }
    g->perm = std::move(perm);
    const int err = gf_setup(g, nc, row_ptr, col, inc_ptr, inc, mcl, diag_pos, conmask, is_pos);
    if (err) {
        plx_gf_destroy(g);
        return err;
    }
    *out = g;
    return 0;
}

}  // extern "C"

namespace plx {

int gf_setup(plx_gf g, int nc, const std::vector<int32_t> &row_ptr, const std::vector<int32_t> &col,
             const std::vector<int32_t> &inc_ptr, const std::vector<int32_t> &inc,
             const std::vector<int32_t> &mcl, const std::vector<int32_t> &diag_pos,
             const std::vector<uint8_t> &conmask, const std::vector<uint8_t> &is_pos) {
    plx_dm dm = g->dm;
    const int nn = g->nrows;
    auto up = [&](auto **d, const auto &h) -> int {
        using T = typename std::decay<decltype(h)>::type::value_type;
        PLX_HIP(dm, hipMalloc((void **)d, h.size() * sizeof(T)));
        PLX_HIP(dm, hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return 0;
    };
    int32_t *d_inc_ptr = nullptr, *d_inc = nullptr, *d_mcl = nullptr, *d_diag = nullptr;
    uint8_t *d_con = nullptr, *d_pos = nullptr;
    double *d_Ke = nullptr;
    int err = 0;
    if ((err = up(&g->d_row_ptr, row_ptr)) || (err = up(&g->d_col, col)) || (err = up(&g->d_perm, g->perm)) ||
        (err = up(&d_inc_ptr, inc_ptr)) ||
        (err = up(&d_inc, inc)) || (err = up(&d_mcl, mcl)) || (err = up(&d_diag, diag_pos)) ||
        (err = up(&d_con, conmask)) || (err = up(&d_pos, is_pos)))
        return err;
    if ((err = ensure_mesh_indices(dm))) return err;
    PLX_HIP(dm, hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    PLX_HIP(dm, hipMalloc(&d_Ke, (size_t)nc * 64 * 9 * sizeof(double)));
    PLX_HIP(dm, hipMalloc(&g->d_val64, (size_t)g->nnzb * 9 * sizeof(double)));
    PLX_HIP(dm, hipMalloc(&g->d_val32, (size_t)g->nnzb * 9 * sizeof(float)));
    PLX_HIP(dm, hipMalloc(&g->d_Binv64, 9 * (size_t)nn * sizeof(double)));
    PLX_HIP(dm, hipMalloc(&g->d_Binv32, 9 * (size_t)nn * sizeof(float)));
    {
        size_t off = 0;  // element matrices of each key's cells with that key's moduli, key after key
        for (KeyPlan *kp : g->keys) {
            const int nk = (int)kp->cells.size();
            hipLaunchKernelGGL(gf_element_kernel, dim3((8 * nk + 255) / 256), dim3(256), 0, g->stream, nk, kp->d_cells,
                               dm->d_closure, dm->d_coords, kp->d_mat64, d_Ke + off * 64 * 9);
            off += (size_t)nk;
        }
    }
    hipLaunchKernelGGL(gf_assemble_kernel, dim3(nn), dim3(256), 0, g->stream, nn, g->d_row_ptr, g->d_col, d_inc_ptr,
                       d_inc, d_mcl, d_Ke, d_con, d_pos, g->d_val64);
    hipLaunchKernelGGL(gf_diag_kernel, dim3((nn + 255) / 256), dim3(256), 0, g->stream, nn, d_diag, g->d_val64,
                       g->d_Binv64, g->d_Binv32, g->d_val32, g->nnzb);
    PLX_HIP(dm, hipGetLastError());
    PLX_HIP(dm, hipStreamSynchronize(g->stream));
    for (void *p : {(void *)d_inc_ptr, (void *)d_inc, (void *)d_mcl, (void *)d_diag, (void *)d_con, (void *)d_pos,
                    (void *)d_Ke})
        (void)hipFree(p);
    return gf_alloc_work(g);
}

// Reduction scalars and single-vector workspace of a handle whose operator is in place.
int gf_alloc_work(plx_gf g) {
    plx_dm dm = g->dm;
    PLX_HIP(dm, hipMalloc(&g->scal, G_COUNT * 64 * sizeof(double)));
    PLX_HIP(dm, hipMalloc(&g->prz, (size_t)kGfRed * 64 * sizeof(double)));
    PLX_HIP(dm, hipMalloc(&g->prr, (size_t)kGfRed * 64 * sizeof(double)));
    PLX_HIP(dm, hipMalloc(&g->nrm, 64 * sizeof(double)));
    PLX_HIP(dm, hipMalloc(&g->v1, dm->local_size * sizeof(double)));
    PLX_HIP(dm, hipMalloc(&g->v2, dm->local_size * sizeof(double)));
    if (dm->fault) PLX_HIP(dm, hipMalloc(&g->jump, 3 * (size_t)dm->fault->num_lag * sizeof(double)));
    return 0;
}

}  // namespace plx

extern "C" {

int plx_gf_destroy(plx_gf g) {
    if (!g) return 0;
    for (int i = 0; i <= 64; ++i)
        if (g->graph[i]) (void)hipGraphExecDestroy(g->graph[i]);
    void *ptrs[] = {g->d_perm, g->d_row_ptr, g->d_col, g->d_val64, g->d_val32, g->d_Binv64, g->d_Binv32, g->B64, g->X64,
                    g->R64,       g->X32,   g->R32,     g->P32,     g->Q32,      g->Z32,      g->v1,  g->v2,
                    g->jump,      g->scal,  g->prz,     g->prr,     g->nrm,      g->d_rr0,    g->d_slot_slip};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    gf_amg_free(g);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return 0;
}

int plx_gf_set_options(plx_gf g, int block, double inner_rtol, int inner_maxit, int use_graph) {
    if (!g) return 1;
    if (block != 0 && block != 8 && block != 16 && block != 32 && block != 64)
        return set_error(g->dm, 1, "block width must be 0 (autotune), 8, 16, 32 or 64");
    if (!(inner_rtol > 0.0) || inner_maxit <= 0) return set_error(g->dm, 1, "bad inner solver options");
    g->block = block;
    g->inner_rtol = inner_rtol;
    g->inner_maxit = inner_maxit;
    g->use_graph = use_graph != 0;
    return 0;
}

int plx_gf_set_preconditioner(plx_gf g, int kind) {
    if (!g) return 1;
    if (kind != PLX_GF_PC_JACOBI && kind != PLX_GF_PC_AMG)
        return set_error(g->dm, 1, "preconditioner must be PLX_GF_PC_JACOBI or PLX_GF_PC_AMG");
    if (kind == g->pc) return 0;
    if (kind == PLX_GF_PC_AMG && g->lv.empty()) {
        const int err = gf_amg_setup(g);
        if (err) {
            gf_amg_free(g);
            return err;
        }
    }
    for (int i = 0; i <= 64; ++i)  // captured chunks belong to the previous preconditioner
        if (g->graph[i]) {
            (void)hipGraphExecDestroy(g->graph[i]);
            g->graph[i] = nullptr;
        }
    g->pc = kind;
    g->probe_us.assign(4, 0.0);
    return 0;
}

int plx_gf_apply(plx_gf g, int k, int scalar_type, const void *X, void *Y) {
    if (!g || !X || !Y) return 1;
    if (k <= 0 || 256 % k != 0 || k > 64) return set_error(g->dm, 1, "k must divide 256 and be <= 64");
    plx_dm dm = g->dm;
    if (g->op_epoch != dm->op_epoch)
        return set_error(dm, 1, "the DM's operator changed after plx_gf_create: destroy and recreate the handle");
    // X and Y are in DM order: through row-ordered copies (a verification path, not a hot one)
    const size_t es = scalar_type == PLX_FP64 ? sizeof(double) : sizeof(float);
    void *Xr = nullptr, *Yr = nullptr;
    PLX_HIP(dm, hipMalloc(&Xr, (size_t)g->ndof * k * es));
    if (hipMalloc(&Yr, (size_t)g->ndof * k * es) != hipSuccess) {
        (void)hipFree(Xr);
        return set_error(dm, 1, "out of device memory");
    }
    if (scalar_type == PLX_FP64) {
        hipLaunchKernelGGL(gf_permute_block<double>, gf_grid(g->ndof * k), dim3(256), 0, g->stream, g->ndof, k,
                           g->d_perm, true, (const double *)X, (double *)Xr);
        gf_spmm_launch<double>(g, k, g->d_val64, (const double *)Xr, (double *)Yr);
        hipLaunchKernelGGL(gf_permute_block<double>, gf_grid(g->ndof * k), dim3(256), 0, g->stream, g->ndof, k,
                           g->d_perm, false, (const double *)Yr, (double *)Y);
    } else {
        hipLaunchKernelGGL(gf_permute_block<float>, gf_grid(g->ndof * k), dim3(256), 0, g->stream, g->ndof, k,
                           g->d_perm, true, (const float *)X, (float *)Xr);
        gf_spmm_launch<float>(g, k, g->d_val32, (const float *)Xr, (float *)Yr);
        hipLaunchKernelGGL(gf_permute_block<float>, gf_grid(g->ndof * k), dim3(256), 0, g->stream, g->ndof, k,
                           g->d_perm, false, (const float *)Yr, (float *)Y);
    }
    const hipError_t e1 = hipGetLastError(), e2 = hipStreamSynchronize(g->stream);
    (void)hipFree(Xr);
    (void)hipFree(Yr);
    PLX_HIP(dm, e1);
    PLX_HIP(dm, e2);
    return 0;
}

int plx_gf_autotune(plx_gf g, int *best) {
    if (!g) return 1;
    plx_dm dm = g->dm;
    const int cand[4] = {8, 16, 32, 64};
    int err = gf_alloc(g, 64);
    if (err) return err;
    g->probe_us.assign(4, 0.0);
    double best_t = 1e300;
    int best_k = 8;
    for (int c = 0; c < 4; ++c) {
        const int k = cand[c];
        const int64_t n = g->ndof * k;
        hipLaunchKernelGGL(gf_fill, gf_grid(n), dim3(256), 0, g->stream, n, g->R64, 1.0);
        hipLaunchKernelGGL(gf_fill, dim3(1), dim3(64), 0, g->stream, 64, g->nrm, std::sqrt((double)g->ndof));
        hipLaunchKernelGGL(gf_to32, gf_grid(n), dim3(256), 0, g->stream, g->ndof, k, g->R64, g->nrm, g->R32, g->X32);
        gf_pcg_start(g, k);
        if ((err = gf_run_chunk(g, k))) return err;  // warm-up (and graph capture)
        PLX_HIP(dm, hipStreamSynchronize(g->stream));
        const auto t0 = std::chrono::steady_clock::now();
        for (int rep = 0; rep < 3; ++rep)
            if ((err = gf_run_chunk(g, k))) return err;
        PLX_HIP(dm, hipStreamSynchronize(g->stream));
        const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
        const double per = us / (3.0 * (g->pc == PLX_GF_PC_AMG ? kAmgChunk : 10) * k);  // per RHS per iteration
        g->probe_us[c] = per;
        if (per < best_t) {
            best_t = per;
            best_k = k;
        }
    }
    g->block = best_k;
    if (best) *best = best_k;
    return 0;
}

int plx_gf_probe_times(plx_gf g, double *us_per_rhs_iteration) {
    if (!g || !us_per_rhs_iteration) return 1;
    for (size_t i = 0; i < g->probe_us.size(); ++i) us_per_rhs_iteration[i] = g->probe_us[i];
    return 0;
}

int plx_gf_solve(plx_gf g, int num_impulses, const double *slip, double *u, double rtol, int maxit,
                 plx_gf_stats *stats) {
    if (!g || num_impulses <= 0 || !slip || !u || !(rtol > 0.0) || maxit <= 0) return 1;
    plx_dm dm = g->dm;
    if (!dm->fault) return set_error(dm, 1, "Green's-function impulses need a fault");
    // the right-hand sides use the live matrix-free operator, the block PCG the copy assembled at
    // create: both must be the same operator
    if (g->op_epoch != dm->op_epoch)
        return set_error(dm, 1, "the DM's operator changed after plx_gf_create: destroy and recreate the handle");
    FaultData &F = *dm->fault;
    int err = 0;
    if (g->block == 0 && (err = plx_gf_autotune(g, nullptr))) return err;
    const int k = g->block;
    if ((err = gf_alloc(g, k))) return err;
    plx_gf_stats st{};
    st.block = k;
    // run everything (including the matrix-free applies) on the solver's own stream, ordered
    // after the work already queued on the handle's stream
    hipEvent_t ev;
    PLX_HIP(dm, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    if (hipEventRecord(ev, dm->stream) != hipSuccess || hipStreamWaitEvent(g->stream, ev, 0) != hipSuccess) {
        (void)hipEventDestroy(ev);
        return set_error(dm, 76, "Green's-function stream ordering failed");
    }
    hipStream_t saved = dm->stream;
    dm->stream = g->stream;
    const int64_t nloc = dm->local_size, ndof = g->ndof, nl3 = 3 * (int64_t)F.num_lag;
    const int nc = (int)dm->con.dofs.size();
    for (int j0 = 0; j0 < num_impulses && !err; j0 += k) {
        const int kb = std::min(k, num_impulses - j0);
        // right-hand sides: B_j = fold(J u_jump_j), masked; columns beyond kb stay zero
        hipLaunchKernelGGL(gf_fill, gf_grid(ndof * k), dim3(256), 0, g->stream, ndof * k, g->B64, 0.0);
        for (int c = 0; c < kb && !err; ++c) {
            const double *s = slip + (size_t)(j0 + c) * nl3;
            if (F.planar) {
                err = launch_planar_jump(dm, F.num_lag, s, g->jump);
            } else {
                err = set_error(dm, 1, "Green's functions need a planar fault");
            }
            if (err) break;
            hipLaunchKernelGGL(gf_fill, gf_grid(nloc), dim3(256), 0, g->stream, nloc, g->v1, 0.0);
            hipLaunchKernelGGL(gf_scatter_jump, dim3((F.num_lag + 255) / 256), dim3(256), 0, g->stream, F.num_lag,
                               F.d_pos_node, g->jump, g->v1);
            if (g->keys.size() == 1) {
                if ((err = launch_apply(dm, g->kp, PLX_FP64, true, g->v1, g->v2, PLX_INSERT_VALUES, 1.0))) break;
            } else {
                hipLaunchKernelGGL(gf_fill, gf_grid(nloc), dim3(256), 0, g->stream, nloc, g->v2, 0.0);
                for (KeyPlan *kp : g->keys)
                    if ((err = launch_apply(dm, kp, PLX_FP64, true, g->v1, g->v2, PLX_ADD_VALUES, 1.0))) break;
                if (err) break;
            }
            if ((err = launch_fold(dm, PLX_FP64, g->v2, kFold))) break;
            if (nc > 0)
                hipLaunchKernelGGL(gf_mask, dim3((nc + 255) / 256), dim3(256), 0, g->stream, nc, dm->con.d_dofs, g->v2);
            hipLaunchKernelGGL(gf_pack, gf_grid(ndof), dim3(256), 0, g->stream, ndof, k, c, g->d_perm, g->v2, g->B64);
        }
        if (err) break;
        if (hipGetLastError() != hipSuccess) {
            err = set_error(dm, 76, "Green's-function right-hand side launch failed");
            break;
        }
        if ((err = gf_block_solve(g, k, rtol, maxit, &st))) break;
        // solutions: u = U w + u_jump (Lagrange multipliers left at zero)
        for (int c = 0; c < kb && !err; ++c) {
            double *uj = u + (size_t)(j0 + c) * nloc;
            hipLaunchKernelGGL(gf_fill, gf_grid(nloc), dim3(256), 0, g->stream, nloc, uj, 0.0);
            hipLaunchKernelGGL(gf_unpack, gf_grid(ndof), dim3(256), 0, g->stream, ndof, k, c, g->d_perm, g->X64, uj);
            if ((err = launch_planar_jump(dm, F.num_lag, slip + (size_t)(j0 + c) * nl3, g->jump))) break;
            err = launch_sync_pos(dm, g->jump, uj);
        }
        st.num_blocks++;
    }
    dm->stream = saved;
    if (!err && (hipEventRecord(ev, g->stream) != hipSuccess || hipStreamWaitEvent(dm->stream, ev, 0) != hipSuccess))
        err = set_error(dm, 76, "Green's-function stream ordering failed");
    (void)hipEventDestroy(ev);
    if (err) return err;
    st.converged = st.max_relative_residual <= rtol ? 1 : 0;
    if (stats) *stats = st;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Continuous batching: k columns solved as independent slots.  A slot is loaded with a new
// right-hand side while the other columns keep iterating; every step replays one captured chunk
// of the FP32 block PCG for all columns, then the slots whose inner solve has converged take
// their FP64 refinement step (X += scale X32, R = B - A X in FP64) and either retire (relative
// FP64 residual <= rtol), fail (outer iterations exhausted) or restart their inner PCG.  Columns
// share no arithmetic: each has its own PCG scalars, so a slot's iterates do not depend on the
// other columns.
// ---------------------------------------------------------------------------------------------
namespace plx {

// Column selections and per-column scales travel as kernel arguments (k <= 64): bit c of `mask`
// selects column c.
struct ColumnScale {
    double v[64];
};

__device__ __forceinline__ bool col_on(uint64_t mask, int c) { return (mask >> c) & 1ull; }

__global__ void __launch_bounds__(256) gf_to32_mask(int64_t ndof, int k, uint64_t mask,
                                                    const double *__restrict__ R64, ColumnScale scale,
                                                    float *R32, float *X32) {
    const int64_t n = ndof * k;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const int c = (int)(t % k);
        if (!col_on(mask, c)) continue;
        const double sc = scale.v[c];
        R32[t] = sc > 0.0 ? (float)(R64[t] / sc) : 0.0f;
        X32[t] = 0.0f;
    }
}

// X64 += scale[c] X32 on the selected columns
__global__ void __launch_bounds__(256) gf_correct_mask(int64_t ndof, int k, uint64_t mask, ColumnScale scale,
                                                       const float *__restrict__ X32, double *X64) {
    const int64_t n = ndof * k;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const int c = (int)(t % k);
        if (col_on(mask, c)) X64[t] += scale.v[c] * (double)X32[t];
    }
}

template <typename T>
__global__ void __launch_bounds__(256) gf_zero_cols(int64_t ndof, int k, uint64_t mask, T *A) {
    const int64_t n = ndof * k;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
        if (col_on(mask, (int)(t % k))) A[t] = T(0);
}

template <typename T>
__global__ void __launch_bounds__(256) gf_copy_cols(int64_t ndof, int k, uint64_t mask, const T *__restrict__ A,
                                                    T *B) {
    const int64_t n = ndof * k;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
        if (col_on(mask, (int)(t % k))) B[t] = A[t];
}

// z = Binv r on every column (the point-block Jacobi preconditioner)
template <int K>
__global__ void __launch_bounds__(256) gfk_zjacobi(int nrows, const float *__restrict__ Binv,
                                                   const float *__restrict__ R, float *Z) {
    const int64_t n = (int64_t)nrows * K;
    const int rc = (int)(threadIdx.x % K);
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const int64_t i = t / K;
        const size_t o = (3 * (size_t)i) * K + rc;
        const float *B = Binv + 9 * i;
        const float r0 = R[o], r1 = R[o + K], r2 = R[o + 2 * K];
        for (int c = 0; c < 3; ++c) Z[o + c * K] = B[3 * c] * r0 + B[3 * c + 1] * r1 + B[3 * c + 2] * r2;
    }
}

// PCG start of the selected columns: beta = 0, r.z and r.r from the partials; r.r also kept as the
// column's starting residual (rr0) for its convergence test
__global__ void __launch_bounds__(256) gf_finish_start_mask(int k, const double *__restrict__ prz,
                                                            const double *__restrict__ prr, double *scal,
                                                            uint64_t mask, double *rr0) {
    const int r = blockIdx.x;
    const double rz = col_sum(prz, k, r);
    __syncthreads();
    const double rr = col_sum(prr, k, r);
    if (threadIdx.x == 0 && col_on(mask, r)) {
        scal[G_BETA * 64 + r] = 0.0;
        scal[G_RZ * 64 + r] = rz;
        scal[G_RR * 64 + r] = rr;
        rr0[r] = rr;
    }
}

// Launches of the handle run on its own stream; the DM's launches inside are redirected there,
// ordered after the work already queued on the DM's stream, and the DM's stream waits for them.
class GfStreamScope {
  public:
    explicit GfStreamScope(plx_gf g) : g_(g), saved_(g->dm->stream) {
        if (hipEventCreateWithFlags(&ev_, hipEventDisableTiming) != hipSuccess) return;
        if (hipEventRecord(ev_, saved_) != hipSuccess || hipStreamWaitEvent(g->stream, ev_, 0) != hipSuccess) return;
        g->dm->stream = g->stream;
        ok_ = true;
    }
    ~GfStreamScope() {
        g_->dm->stream = saved_;
        if (ok_ && hipEventRecord(ev_, g_->stream) == hipSuccess) (void)hipStreamWaitEvent(saved_, ev_, 0);
        if (ev_) (void)hipEventDestroy(ev_);
    }
    bool ok() const { return ok_; }

  private:
    plx_gf g_;
    hipStream_t saved_;
    hipEvent_t ev_ = nullptr;
    bool ok_ = false;
};

// Column c of B64 = the impulse right-hand side -U^T J u_jump of one slip (as plx_gf_solve).
static int gf_rhs_column(plx_gf g, const double *slip, int c, int k) {
    plx_dm dm = g->dm;
    FaultData &F = *dm->fault;
    const int64_t nloc = dm->local_size, ndof = g->ndof;
    const int nc = (int)dm->con.dofs.size();
    int err = launch_planar_jump(dm, F.num_lag, slip, g->jump);
    if (err) return err;
    hipLaunchKernelGGL(gf_fill, gf_grid(nloc), dim3(256), 0, g->stream, nloc, g->v1, 0.0);
    hipLaunchKernelGGL(gf_scatter_jump, dim3((F.num_lag + 255) / 256), dim3(256), 0, g->stream, F.num_lag,
                       F.d_pos_node, g->jump, g->v1);
    if (g->keys.size() == 1) {
        if ((err = launch_apply(dm, g->kp, PLX_FP64, true, g->v1, g->v2, PLX_INSERT_VALUES, 1.0))) return err;
    } else {
        hipLaunchKernelGGL(gf_fill, gf_grid(nloc), dim3(256), 0, g->stream, nloc, g->v2, 0.0);
        for (KeyPlan *kp : g->keys)
            if ((err = launch_apply(dm, kp, PLX_FP64, true, g->v1, g->v2, PLX_ADD_VALUES, 1.0))) return err;
    }
    if ((err = launch_fold(dm, PLX_FP64, g->v2, kFold))) return err;
    if (nc > 0) hipLaunchKernelGGL(gf_mask, dim3((nc + 255) / 256), dim3(256), 0, g->stream, nc, dm->con.d_dofs, g->v2);
    hipLaunchKernelGGL(gf_pack, gf_grid(ndof), dim3(256), 0, g->stream, ndof, k, c, g->d_perm, g->v2, g->B64);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

static uint64_t mask_bits(const std::vector<int> &mask) {
    uint64_t m = 0;
    for (size_t c = 0; c < mask.size() && c < 64; ++c)
        if (mask[c]) m |= 1ull << c;
    return m;
}

static ColumnScale slot_scales(plx_gf g) {
    ColumnScale sc;
    for (int c = 0; c < 64; ++c) sc.v[c] = c < g->slots_k ? g->slots[c].scale : 0.0;
    return sc;
}

template <int K>
static void gf_start_mask_k(plx_gf g, uint64_t mask) {
    if (g->pc == PLX_GF_PC_AMG)
        amg_vcycle<K>(g, 0, g->R32, g->Z32);
    else
        hipLaunchKernelGGL(gfk_zjacobi<K>, dim3(kGfRed), dim3(256), 0, g->stream, g->nrows, g->d_Binv32, g->R32, g->Z32);
    hipLaunchKernelGGL(gf_copy_cols<float>, gf_grid(g->ndof * K), dim3(256), 0, g->stream, g->ndof, K, mask,
                       g->Z32, g->P32);
    hipLaunchKernelGGL(gfk_dot2<K>, dim3(kGfRed), dim3(256), 0, g->stream, g->ndof, g->Z32, g->R32, g->prz, g->prr);
    hipLaunchKernelGGL(gf_finish_start_mask, dim3(K), dim3(256), 0, g->stream, K, g->prz, g->prr, g->scal, mask,
                       g->d_rr0);
}

// (Re)start the inner PCG of the selected slots from R32 = R64 / scale, X32 = 0 (no host wait:
// the starting residuals stay on the device for the next step's convergence test).
static int gf_restart(plx_gf g, uint64_t mask) {
    plx_dm dm = g->dm;
    const int k = g->slots_k;
    hipLaunchKernelGGL(gf_to32_mask, gf_grid(g->ndof * k), dim3(256), 0, g->stream, g->ndof, k, mask, g->R64,
                       slot_scales(g), g->R32, g->X32);
    switch (k) {
    case 8: gf_start_mask_k<8>(g, mask); break;
    case 16: gf_start_mask_k<16>(g, mask); break;
    case 32: gf_start_mask_k<32>(g, mask); break;
    default: gf_start_mask_k<64>(g, mask); break;
    }
    PLX_HIP(dm, hipGetLastError());
    for (int c = 0; c < k; ++c)
        if ((mask >> c) & 1ull) {
            g->slots[c].inner = 0;
            g->slots[c].ready_since = -1;
        }
    return 0;
}

// Idle columns: zero vectors (the PCG then computes alpha = beta = 0 for them).
static int gf_idle(plx_gf g, uint64_t mask, bool clear_solution) {
    plx_dm dm = g->dm;
    const int k = g->slots_k;
    for (float *v : {g->R32, g->X32, g->P32, g->Z32})
        hipLaunchKernelGGL(gf_zero_cols<float>, gf_grid(g->ndof * k), dim3(256), 0, g->stream, g->ndof, k, mask, v);
    if (clear_solution)
        for (double *v : {g->B64, g->X64, g->R64})
            hipLaunchKernelGGL(gf_zero_cols<double>, gf_grid(g->ndof * k), dim3(256), 0, g->stream, g->ndof, k, mask, v);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

}  // namespace plx

extern "C" {

int plx_gf_slots_open(plx_gf g, int k, double rtol, int maxit) {
    if (!g) return 1;
    plx_dm dm = g->dm;
    if (k != 8 && k != 16 && k != 32 && k != 64) return set_error(dm, 1, "slot count must be 8, 16, 32 or 64");
    if (!(rtol > 0.0) || maxit <= 0) return set_error(dm, 1, "bad slot tolerance");
    if (!dm->fault) return set_error(dm, 1, "Green's-function slots need a fault");
    if (g->op_epoch != dm->op_epoch)
        return set_error(dm, 1, "the DM's operator changed after plx_gf_create: destroy and recreate the handle");
    int err = gf_alloc(g, k);
    if (err) return err;
    if (!g->d_rr0) PLX_HIP(dm, hipMalloc(&g->d_rr0, 64 * sizeof(double)));
    if (g->d_slot_slip) (void)hipFree(g->d_slot_slip);
    g->d_slot_slip = nullptr;
    PLX_HIP(dm, hipMalloc(&g->d_slot_slip, (size_t)k * 3 * dm->fault->num_lag * sizeof(double)));
    g->slots.assign(k, plx_gf_s::Slot());
    g->slots_k = k;
    g->slots_rtol = rtol;
    g->slots_maxit = maxit;
    g->slots_step = 0;
    GfStreamScope scope(g);
    if (!scope.ok()) return set_error(dm, 76, "stream ordering failed");
    hipLaunchKernelGGL(gf_fill, dim3(1), dim3(64), 0, g->stream, 64, g->d_rr0, 0.0);
    const uint64_t all = k == 64 ? ~0ull : ((1ull << k) - 1);
    if ((err = gf_idle(g, all, true))) return err;
    PLX_HIP(dm, hipStreamSynchronize(g->stream));
    return 0;
}

int plx_gf_slots_load(plx_gf g, int n, const int *slot_ids, const double *slip) {
    if (!g || n < 0 || (n > 0 && (!slot_ids || !slip))) return 1;
    plx_dm dm = g->dm;
    if (!g->slots_k) return set_error(dm, 1, "plx_gf_slots_open first");
    if (g->op_epoch != dm->op_epoch)
        return set_error(dm, 1, "the DM's operator changed after plx_gf_create: destroy and recreate the handle");
    const int k = g->slots_k;
    const size_t nl3 = 3 * (size_t)dm->fault->num_lag;
    uint64_t mask = 0;
    for (int i = 0; i < n; ++i) {
        const int c = slot_ids[i];
        if (c < 0 || c >= k) return set_error(dm, 1, "slot index out of range");
        if (g->slots[c].state != PLX_SLOT_EMPTY) return set_error(dm, 1, "slot is not empty");
        if ((mask >> c) & 1ull) return set_error(dm, 1, "slot listed twice");
        mask |= 1ull << c;
    }
    if (n == 0) return 0;
    GfStreamScope scope(g);
    if (!scope.ok()) return set_error(dm, 76, "stream ordering failed");
    int err = 0;
    for (int i = 0; i < n && !err; ++i) {
        const int c = slot_ids[i];
        PLX_HIP(dm, hipMemcpyAsync(g->d_slot_slip + c * nl3, slip + i * nl3, nl3 * sizeof(double),
                                   hipMemcpyDeviceToDevice, g->stream));
        err = gf_rhs_column(g, g->d_slot_slip + c * nl3, c, k);
    }
    if (err) return err;
    // X = 0, R = B on the loaded columns
    hipLaunchKernelGGL(gf_zero_cols<double>, gf_grid(g->ndof * k), dim3(256), 0, g->stream, g->ndof, k, mask, g->X64);
    hipLaunchKernelGGL(gf_copy_cols<double>, gf_grid(g->ndof * k), dim3(256), 0, g->stream, g->ndof, k, mask, g->B64,
                       g->R64);
    std::vector<double> b0;
    if ((err = gf_norms(g, k, g->B64, b0))) return err;
    for (int c = 0; c < k; ++c)
        if ((mask >> c) & 1ull) {
            plx_gf_s::Slot &S = g->slots[c];
            S = plx_gf_s::Slot();
            S.state = PLX_SLOT_RUNNING;
            S.b0 = b0[c];
            S.scale = b0[c];
            S.relres = 1.0;
        }
    return gf_restart(g, mask);
}

int plx_gf_slots_step(plx_gf g, int *state) {
    if (!g) return 1;
    plx_dm dm = g->dm;
    const int k = g->slots_k;
    if (!k) return set_error(dm, 1, "plx_gf_slots_open first");
    if (g->op_epoch != dm->op_epoch)
        return set_error(dm, 1, "the DM's operator changed after plx_gf_create: destroy and recreate the handle");
    int running = 0;
    for (const auto &S : g->slots) running += S.state == PLX_SLOT_RUNNING;
    int err = 0;
    if (running) {
        GfStreamScope scope(g);
        if (!scope.ok()) return set_error(dm, 76, "stream ordering failed");
        if ((err = gf_run_chunk(g, k))) return err;
        const int its = g->pc == PLX_GF_PC_AMG ? kAmgChunk : 10;
        const int step = ++g->slots_step;
        double rr[64], rr0[64];
        PLX_HIP(dm, hipMemcpyAsync(rr, g->scal + G_RR * 64, k * sizeof(double), hipMemcpyDeviceToHost, g->stream));
        PLX_HIP(dm, hipMemcpyAsync(rr0, g->d_rr0, k * sizeof(double), hipMemcpyDeviceToHost, g->stream));
        PLX_HIP(dm, hipStreamSynchronize(g->stream));
        int ready = 0, oldest = 0;
        for (int c = 0; c < k; ++c) {
            plx_gf_s::Slot &S = g->slots[c];
            if (S.state != PLX_SLOT_RUNNING) continue;
            S.inner += its;
            const bool done = rr0[c] <= 0.0 || rr[c] <= g->inner_rtol * g->inner_rtol * rr0[c] || S.inner >= g->inner_maxit;
            if (done && S.ready_since < 0) S.ready_since = step;
            if (S.ready_since >= 0) {
                ++ready;
                oldest = std::max(oldest, step - S.ready_since);
            }
        }
        // refinement rounds: when a quarter of the running slots (at least one) are ready, or a
        // ready slot has waited 3 steps; a waiting slot's inner PCG keeps refining its correction
        if (ready > 0 && (4 * ready >= running || oldest >= 3)) {
            uint64_t sel = 0;
            for (int c = 0; c < k; ++c)
                if (g->slots[c].state == PLX_SLOT_RUNNING && g->slots[c].ready_since >= 0) sel |= 1ull << c;
            // X64 += scale X32 on the selected columns, then the FP64 residual of every column
            hipLaunchKernelGGL(gf_correct_mask, gf_grid(g->ndof * k), dim3(256), 0, g->stream, g->ndof, k, sel,
                               slot_scales(g), g->X32, g->X64);
            gf_spmm_launch<double>(g, k, g->d_val64, g->X64, g->R64);
            hipLaunchKernelGGL(gf_resid, gf_grid(g->ndof * k), dim3(256), 0, g->stream, g->ndof * k, g->B64, g->R64);
            PLX_HIP(dm, hipGetLastError());
            std::vector<double> rn;
            if ((err = gf_norms(g, k, g->R64, rn))) return err;
            uint64_t restart = 0, finished = 0;
            for (int c = 0; c < k; ++c) {
                if (!((sel >> c) & 1ull)) continue;
                plx_gf_s::Slot &S = g->slots[c];
                S.outer++;
                S.relres = S.b0 > 0.0 ? rn[c] / S.b0 : 0.0;
                if (S.relres <= g->slots_rtol) {
                    S.state = PLX_SLOT_CONVERGED;
                    finished |= 1ull << c;
                } else if (S.outer >= g->slots_maxit || !std::isfinite(S.relres)) {
                    S.state = PLX_SLOT_FAILED;
                    finished |= 1ull << c;
                } else {
                    S.scale = rn[c];
                    restart |= 1ull << c;
                }
            }
            if (finished && (err = gf_idle(g, finished, false))) return err;
            if (restart && (err = gf_restart(g, restart))) return err;
        }
    }
    if (state)
        for (int c = 0; c < k; ++c) state[c] = g->slots[c].state;
    return 0;
}

int plx_gf_slot_take(plx_gf g, int slot, double *u, plx_gf_slot_stats *stats) {
    if (!g) return 1;
    plx_dm dm = g->dm;
    if (!g->slots_k || slot < 0 || slot >= g->slots_k) return set_error(dm, 1, "slot index out of range");
    plx_gf_s::Slot &S = g->slots[slot];
    if (S.state != PLX_SLOT_CONVERGED && S.state != PLX_SLOT_FAILED) return set_error(dm, 1, "slot has no result");
    if (stats) {
        stats->converged = S.state == PLX_SLOT_CONVERGED;
        stats->outer_iterations = S.outer;
        stats->inner_iterations = S.inner;
        stats->relative_residual = S.relres;
    }
    const int k = g->slots_k;
    const size_t nl3 = 3 * (size_t)dm->fault->num_lag;
    GfStreamScope scope(g);
    if (!scope.ok()) return set_error(dm, 76, "stream ordering failed");
    int err = 0;
    if (u) {
        // u = U w + u_jump (Lagrange multipliers left at zero)
        hipLaunchKernelGGL(gf_fill, gf_grid(dm->local_size), dim3(256), 0, g->stream, dm->local_size, u, 0.0);
        hipLaunchKernelGGL(gf_unpack, gf_grid(g->ndof), dim3(256), 0, g->stream, g->ndof, k, slot, g->d_perm, g->X64, u);
        if ((err = launch_planar_jump(dm, dm->fault->num_lag, g->d_slot_slip + slot * nl3, g->jump))) return err;
        if ((err = launch_sync_pos(dm, g->jump, u))) return err;
    }
    if ((err = gf_idle(g, 1ull << slot, true))) return err;
    S = plx_gf_s::Slot();
    return 0;
}

}  // extern "C"
