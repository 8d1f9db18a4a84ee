// plx_vector.hip -- field scatters and FP64 reductions.
//
// DMLocalToGlobal / DMGlobalToLocal of a single process (topology/Field.cc:394-422) become a
// gather over the global -> local index map, so every output element is written once by
// one thread (no atomics).  Dot products accumulate in FP64 whatever the vector type, in a
// fixed two-pass tree, so they are reproducible run to run.
#include "plx_internal.hpp"

namespace plx {

template <typename Real>
__global__ void __launch_bounds__(256) l2g_kernel(int64_t n, const int32_t *__restrict__ g2l,
                                                  const Real *__restrict__ loc, Real *glob, int mode) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Real v = loc[g2l ? g2l[i] : i];
    glob[i] = mode == PLX_INSERT_VALUES ? v : glob[i] + v;
}

template <typename Real>
__global__ void __launch_bounds__(256) g2l_kernel(int64_t n, const int32_t *__restrict__ g2l,
                                                  const Real *__restrict__ glob, Real *loc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    loc[g2l ? g2l[i] : i] = glob[i];
}

template <typename Real>
__global__ void __launch_bounds__(256) bc_kernel(int n, const int32_t *__restrict__ dofs,
                                                 const Real *__restrict__ values, Real *loc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    loc[dofs[i]] = values[i];
}

constexpr int kDotBlocks = 1024;

template <typename Real>
__global__ void __launch_bounds__(256) dot_partial(int64_t n, const Real *__restrict__ x,
                                                   const Real *__restrict__ y, double *partials) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        acc += (double)x[i] * (double)y[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(256) dot_final(int n, const double *__restrict__ partials,
                                                 double *out) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += partials[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

static inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

int launch_l2g(plx_dm dm, int st, const void *loc, void *glob, int mode) {
    const int64_t n = dm->con.global_size;
    if (n == 0) return 0;
    const int32_t *g2l = dm->con.d_g2l;
    if (st == PLX_FP64)
        hipLaunchKernelGGL(l2g_kernel<double>, grid_for(n), dim3(256), 0, dm->stream, n, g2l,
                           (const double *)loc, (double *)glob, mode);
    else if (st == PLX_FP32)
        hipLaunchKernelGGL(l2g_kernel<float>, grid_for(n), dim3(256), 0, dm->stream, n, g2l,
                           (const float *)loc, (float *)glob, mode);
    else
        return set_error(dm, 1, "bad scalar type");
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

int launch_g2l(plx_dm dm, int st, const void *glob, void *loc) {
    const int64_t n = dm->con.global_size;
    if (n == 0) return 0;
    const int32_t *g2l = dm->con.d_g2l;
    if (st == PLX_FP64)
        hipLaunchKernelGGL(g2l_kernel<double>, grid_for(n), dim3(256), 0, dm->stream, n, g2l,
                           (const double *)glob, (double *)loc);
    else if (st == PLX_FP32)
        hipLaunchKernelGGL(g2l_kernel<float>, grid_for(n), dim3(256), 0, dm->stream, n, g2l,
                           (const float *)glob, (float *)loc);
    else
        return set_error(dm, 1, "bad scalar type");
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

int launch_insert_bc(plx_dm dm, int st, const void *values, void *loc) {
    const int n = (int)dm->con.dofs.size();
    if (n == 0) return 0;
    if (st == PLX_FP64)
        hipLaunchKernelGGL(bc_kernel<double>, grid_for(n), dim3(256), 0, dm->stream, n,
                           dm->con.d_dofs, (const double *)values, (double *)loc);
    else if (st == PLX_FP32)
        hipLaunchKernelGGL(bc_kernel<float>, grid_for(n), dim3(256), 0, dm->stream, n,
                           dm->con.d_dofs, (const float *)values, (float *)loc);
    else
        return set_error(dm, 1, "bad scalar type");
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

int launch_dot(plx_dm dm, int st, int64_t n, const void *x, const void *y, double *result) {
    if (!dm->d_partials) PLX_HIP(dm, hipMalloc(&dm->d_partials, kDotBlocks * sizeof(double)));
    if (st == PLX_FP64)
        hipLaunchKernelGGL(dot_partial<double>, dim3(kDotBlocks), dim3(256), 0, dm->stream, n,
                           (const double *)x, (const double *)y, dm->d_partials);
    else if (st == PLX_FP32)
        hipLaunchKernelGGL(dot_partial<float>, dim3(kDotBlocks), dim3(256), 0, dm->stream, n,
                           (const float *)x, (const float *)y, dm->d_partials);
    else
        return set_error(dm, 1, "bad scalar type");
    hipLaunchKernelGGL(dot_final, dim3(1), dim3(256), 0, dm->stream, kDotBlocks, dm->d_partials,
                       dm->d_scalar);
    PLX_HIP(dm, hipGetLastError());
    PLX_HIP(dm, hipMemcpyAsync(result, dm->d_scalar, sizeof(double), hipMemcpyDeviceToHost, dm->stream));
    PLX_HIP(dm, hipStreamSynchronize(dm->stream));
    return 0;
}

}  // namespace plx
