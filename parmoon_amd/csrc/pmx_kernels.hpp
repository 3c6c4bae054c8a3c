// pmx_kernels.hpp -- HIP kernels of libpmx (gfx950).
//
// All operator kernels work on lattice-ordered vectors and the colour-sliced
// stencil (CSS) value layout described in pmx_internal.hpp.  One thread owns
// one lattice node; a wavefront covers 64 consecutive nodes of one residue
// class, so stencil values are read fully coalesced and every wavefront
// follows the same stencil shape (no divergence).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pmx_internal.hpp"

namespace pmx {


// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs.  Remap
// so that each XCD works on one contiguous slice of the block table: the
// neighbouring nodes whose x values a block gathers then stay in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int G) {
  const int X = 8;
  const int q = G / X, r = G % X, xcd = bid % X, k = bid / X;
  return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + k;
}

// first value of row idx of a class (slot stride kSlice)
__device__ __forceinline__ int64_t vbase(const DevSegment &s, int idx) {
  return s.voff + (int64_t)(idx / kSlice) * ((int64_t)s.ns * kSlice) + (idx % kSlice);
}
// The row-scaled FP16 smoother values are stored with the slots of a row in
// groups of eight: inside a slice of kSlice rows, group o holds, row by row,
// the row's slots 8o .. 8o+7 (16 contiguous bytes per row, a last group of
// ns mod 8 slots likewise), so a row's values arrive in 16-byte loads.  The
// slice sizes are those of the FP32 values.
__device__ __forceinline__ int64_t v16slice(const DevSegment &s, int idx) {
  return s.voff + (int64_t)(idx / kSlice) * ((int64_t)s.ns * kSlice);
}
__host__ __device__ constexpr int v16off(int ns, int r, int slot) {
  return (slot >> 3) * 8 * kSlice + r * ((slot >> 3) < (ns >> 3) ? 8 : (ns & 7)) + (slot & 7);
}

// ---------------------------------------------------------------- BLAS-1
template <typename T>
__global__ void k_fill(T *__restrict__ x, T v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}
template <typename T, typename U>
__global__ void k_convert(T *__restrict__ y, const U *__restrict__ x, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (T)x[i];
}
// y += a*x with a = sgn * (*alpha) (device scalar) or a host constant
template <typename T>
__global__ void k_axpy(T *__restrict__ y, const T *__restrict__ x, const double *alpha, double sgn, double ac,
                       int64_t n) {
  const double a = alpha ? sgn * alpha[0] : ac;
  const T at = (T)a;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += at * x[i];
}
template <typename T>
__global__ void k_copy(T *__restrict__ y, const T *__restrict__ x, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i];
}
// y += (T) x  (precision-changing update)
template <typename T, typename U>
__global__ void k_add_convert(T *__restrict__ y, const U *__restrict__ x, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += (T)x[i];
}
// y = x * (1 / *nrm)   (device scalar)
template <typename T>
__global__ void k_scale_inv(T *__restrict__ y, const T *__restrict__ x, const double *nrm, int64_t n) {
  const T s = (T)(1.0 / nrm[0]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i] * s;
}

// deterministic dot product: fixed grid, fixed per-thread order, fixed tree
template <typename T>
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBlock / 64; i++) s += sh[i];
  __syncthreads();
  return s;
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot_partial(const T *__restrict__ x, const T *__restrict__ y,
                                                        double *__restrict__ part, int64_t n) {
  double s = 0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    s += (double)x[i] * (double)y[i];
  s = block_sum<T>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// out[0] = sum(part) (or sqrt of it)
__global__ __launch_bounds__(kBlock) void k_reduce_final(const double *__restrict__ part, int np,
                                                         double *__restrict__ out, int do_sqrt);

// ------------------------------------------------------------- boundary rows
// Dirichlet rows are identity: y_b = x_b (spmv), r_b = f_b - x_b (defect), x_b = f_b (smoother)
template <typename T, int MODE>
__global__ void k_boundary(const int *__restrict__ bnd, int nb, T *__restrict__ y, const T *__restrict__ x,
                           const T *__restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int p = bnd[i];
  if (MODE == 0) y[p] = x[p];
  else if (MODE == 1) y[p] = f[p] - x[p];
  else if (MODE == 2) y[p] = f[p];
  else y[p] = 0;
}

}  // namespace pmx
