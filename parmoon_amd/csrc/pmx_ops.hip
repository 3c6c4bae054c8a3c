// pmx_ops.hip -- operator, transfer, assembly and Krylov kernels (gfx950) and
// their stream-ordered launchers.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <stdexcept>
#include <type_traits>
#include <utility>

#include "pmx_functors.hpp"
#include "pmx_kernels.hpp"
#include "pmx_ops.hpp"

namespace pmx {

__global__ __launch_bounds__(kBlock) void k_reduce_final(const double *__restrict__ part, int np,
                                                         double *__restrict__ out, int do_sqrt) {
  double s = 0;
  for (int i = threadIdx.x; i < np; i += kBlock) s += part[i];
  s = block_sum<double>(s);
  if (threadIdx.x == 0) out[0] = do_sqrt ? sqrt(s) : s;
}

// ============================================================ stencil rows
// One matrix row is handled by R lanes (R = row_lanes of the level): lane
// `part` accumulates groups of slots part, part+R, ... and a shuffle
// reduction combines them.  Modes:
// 0: y = A x                          (MatVect, Blas2.C:86-118)
// 1: y = f - A x                      (ScalarDefect, Blas2.C:37-83)
// 2: y = (1-w) x + w (f - A_off x)/d  (TMGLevel3D::Jacobi, MGLevel3D.C:469-563), x = old iterate
// 3: x = w ((f - A_off x)/d - x) + x  (TMGLevel3D::SOR, MGLevel3D.C:280-361) in place, one colour
// 4: mode 3 on row-scaled values: val holds a_ij / a_ii (the diagonal slot
//    unused), dinv the lattice vector 1 / a_ii: x = w (f dinv - W_off x - x) + x
template <typename T> struct Acc { using type = T; };

// W consecutive lattice values starting at q (the x-run of one (dy, dz) group
// of a stencil row) in as few loads as the element alignment allows:
// one 16-byte load plus one element for W = 5 floats, 8 + 4 bytes for W = 3.
template <typename T, int W>
__device__ __forceinline__ void load_run(const T *q, T (&r)[W]) {
  if constexpr (sizeof(T) == 4 && W == 5) {
    float4 a;
    __builtin_memcpy(&a, q, 16);
    r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w; r[4] = q[4];
  } else if constexpr (sizeof(T) == 4 && W == 3) {
    float2 a;
    __builtin_memcpy(&a, q, 8);
    r[0] = a.x; r[1] = a.y; r[2] = q[2];
  } else if constexpr (sizeof(T) == 8 && W == 5) {
    double2 a, b;
    __builtin_memcpy(&a, q, 16);
    __builtin_memcpy(&b, q + 2, 16);
    r[0] = a.x; r[1] = a.y; r[2] = b.x; r[3] = b.y; r[4] = q[4];
  } else if constexpr (sizeof(T) == 8 && W == 3) {
    double2 a;
    __builtin_memcpy(&a, q, 16);
    r[0] = a.x; r[1] = a.y; r[2] = q[2];
  } else {
#pragma unroll
    for (int k = 0; k < W; k++) r[k] = q[k];
  }
}

// The slots of a row are grouped by (dy, dz); the WX slots of a group read
// WX consecutive x values, fetched as one run (XL: x lives in LDS and is read
// element by element).  Lane `part` of the R lanes takes groups part,
// part+R, ...  The diagonal, f and the old x of the row are loaded up front
// by all R lanes, so their latency overlaps the stencil loads.
template <typename T, int EX, int EY, int EZ, int R, int MODE, bool XL = false, typename V = T>
__device__ __forceinline__ void row_split(const DevSegment &s, int idx, int part, int M,
                                          const V *__restrict__ val, const T *x, const T *__restrict__ f, T *y,
                                          double omega, const float *__restrict__ dinv = nullptr) {
  using A = typename Acc<T>::type;
  constexpr int WX = 2 * EX + 1, WY = 2 * EY + 1, WZ = 2 * EZ + 1, NG = WY * WZ;
  constexpr int DS = EX + WX * (EY + WY * EZ), DG = DS / WX;
  const int i = idx % s.cx;
  const int t = idx / s.cx;
  const int j = t % s.cy, k = t / s.cy;
  const int p = (s.x0 + s.stride * i) + M * ((s.y0 + s.stride * j) + M * (s.z0 + s.stride * k));
  constexpr int NS = WX * WY * WZ;
  // MODE 4: the FP16 values' slice and the row within it (v16off layout)
  const V *v = MODE == 4 ? val + v16slice(s, idx) : val + vbase(s, idx);
  const int rs = idx % kSlice;
  A fp = 0, diag = 1, told = 0;
  if constexpr (MODE >= 1) fp = (A)f[p];
  if constexpr (MODE == 4) {
    fp *= (A)dinv[p];
    told = (A)x[p];
  } else if constexpr (MODE >= 2) {
    diag = (A)v[DS * kSlice];
    told = (A)x[p];
  }
  // Loads and FMAs are unconditional (a lane past the last group re-reads the
  // last group with zero weights, the diagonal gets a zero weight in the
  // smoothers), so that a wavefront issues all loads before its first FMA
  // instead of one round trip per group.
  A acc = 0;
#pragma unroll
  for (int m = 0; m < (NG + R - 1) / R; m++) {
    const int g0 = part + R * m;
    const bool ok = g0 < NG;
    const int g = ok ? g0 : NG - 1;
    const int dy = g % WY - EY, dz = g / WY - EZ;
    const T *q = x + (p - EX + M * (dy + M * dz));
    T xr[WX];
    V vr[WX];
    if constexpr (XL) {
#pragma unroll
      for (int e = 0; e < WX; e++) xr[e] = q[e];
    } else {
      load_run<T, WX>(q, xr);
    }
#pragma unroll
    for (int e = 0; e < WX; e++) {
      if constexpr (MODE == 4) vr[e] = v[v16off(NS, rs, g * WX + e)];
      else vr[e] = v[(g * WX + e) * kSlice];
    }
#pragma unroll
    for (int e = 0; e < WX; e++) {
      const bool use = ok && !(MODE >= 2 && e == EX && g == DG);
      acc += (use ? (A)vr[e] : (A)0) * (A)xr[e];
    }
  }
#pragma unroll
  for (int o = R / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, R);
  if (part != 0) return;
  if (MODE == 0) y[p] = (T)acc;
  else if (MODE == 1) y[p] = (T)(fp - acc);
  else if (MODE == 2) y[p] = (T)((1 - (A)omega) * told + (A)omega * (fp - acc) / diag);
  // SOR: the Gauss-Seidel value itself when omega = 1, so that a colour
  // applied twice in a row gives the same bits (the neighbours it reads are
  // unchanged) and sweeps may skip such a repeat (diag = 1 in mode 4)
  else if (omega == 1.0) y[p] = (T)((fp - acc) / diag);
  else y[p] = (T)((A)omega * ((fp - acc) / diag - told) + told);
}

// R lanes per row, kRowsPerBlock rows per workgroup; with one lane per row
// the compiler is held to 128 registers so that 4 waves share a SIMD.
template <typename T, int MODE, int R>
__global__ __launch_bounds__(kRowsPerBlock * R, R == 1 ? 4 : 1) void k_rows(const RowBlock *__restrict__ rows, int M,
                                                 const T *__restrict__ val, const T *x, const T *__restrict__ f,
                                                 T *y, double omega) {
  const RowBlock &rb = rows[xcd_remap(blockIdx.x, gridDim.x)];
  const DevSegment s = rb.s;
  const int idx = rb.base + threadIdx.x / R;
  const int part = threadIdx.x % R;
  if (idx >= s.count) return;  // the R lanes of a row leave together
#define PMX_ROW(EX, EY, EZ) row_split<T, EX, EY, EZ, R, MODE>(s, idx, part, M, val, x, f, y, omega)
  switch (s.ex * 100 + s.ey * 10 + s.ez) {
    case 111: PMX_ROW(1, 1, 1); break;
    case 211: PMX_ROW(2, 1, 1); break;
    case 121: PMX_ROW(1, 2, 1); break;
    case 221: PMX_ROW(2, 2, 1); break;
    case 112: PMX_ROW(1, 1, 2); break;
    case 212: PMX_ROW(2, 1, 2); break;
    case 122: PMX_ROW(1, 2, 2); break;
    default: PMX_ROW(2, 2, 2); break;
  }
#undef PMX_ROW
}

// Mode 4 (SOR of one colour on row-scaled FP16 values, FP32 vectors), laid
// out as k_rows.
template <int R>
__global__ __launch_bounds__(kRowsPerBlock * R, R == 1 ? 4 : 1) void k_sor_scaled(
    const RowBlock *__restrict__ rows, int M, const _Float16 *__restrict__ val, const float *__restrict__ dinv,
    float *x, const float *__restrict__ f, double omega) {
  const RowBlock &rb = rows[xcd_remap(blockIdx.x, gridDim.x)];
  const DevSegment s = rb.s;
  const int idx = rb.base + threadIdx.x / R;
  const int part = threadIdx.x % R;
  if (idx >= s.count) return;
#define PMX_ROW(EX, EY, EZ) row_split<float, EX, EY, EZ, R, 4, false, _Float16>(s, idx, part, M, val, x, f, x, omega, dinv)
  switch (s.ex * 100 + s.ey * 10 + s.ez) {
    case 111: PMX_ROW(1, 1, 1); break;
    case 211: PMX_ROW(2, 1, 1); break;
    case 121: PMX_ROW(1, 2, 1); break;
    case 221: PMX_ROW(2, 2, 1); break;
    case 112: PMX_ROW(1, 1, 2); break;
    case 212: PMX_ROW(2, 1, 2); break;
    case 122: PMX_ROW(1, 2, 2); break;
    default: PMX_ROW(2, 2, 2); break;
  }
#undef PMX_ROW
}

// Row-scaled FP16 copy of the FP32 stencil values for mode 4: slot k of an
// active row holds a_k / a_diag rounded to FP16 (saturated at +-65504, the
// largest finite FP16 value), and dinv[p] = 1 / a_diag in FP32.  One thread
// per row.
__global__ __launch_bounds__(kRowsPerBlock) void k_scaled_values(const BlockInfo *__restrict__ blocks,
                                                                 const DevSegment *__restrict__ segs, int M,
                                                                 const float *__restrict__ val,
                                                                 _Float16 *__restrict__ val16,
                                                                 float *__restrict__ dinv) {
  const BlockInfo bi = blocks[blockIdx.x];
  const DevSegment s = segs[bi.seg];
  const int idx = bi.base + threadIdx.x;
  if (idx >= s.count) return;
  const int i = idx % s.cx, t = idx / s.cx;
  const int j = t % s.cy, k = t / s.cy;
  const int p = (s.x0 + s.stride * i) + M * ((s.y0 + s.stride * j) + M * (s.z0 + s.stride * k));
  const int64_t vb = vbase(s, idx);
  const float d = val[vb + (int64_t)s.dslot * kSlice];
  const float r = 1.0f / d;
  dinv[p] = r;
  const int64_t sb = v16slice(s, idx);
  for (int q = 0; q < s.ns; q++) {
    const float w = val[vb + (int64_t)q * kSlice] * r;
    val16[sb + v16off(s.ns, idx % kSlice, q)] = (_Float16)fminf(fmaxf(w, -65504.0f), 65504.0f);
  }
}

// The rows of colour c of a level staged in LDS (segment table lseg, colour
// ranges lcs, each segment's first row in its colour lrow0, rows per colour
// lrows), R lanes per row, computed by row_split in mode MODE with x in LDS.
template <typename T, int R, int MODE, int NT>
__device__ __forceinline__ void wg_colour_rows(const DevSegment *lseg, const int *lcs, const int *lrow0,
                                               const int *lrows, int c, int M, const T *__restrict__ val,
                                               const T *lx, const T *__restrict__ f, T *y, double omega) {
  const int part = threadIdx.x % R;
  const int s0 = lcs[c], s1 = lcs[c + 1], n = lrows[c];
  // the R lanes of a row share r, so they run the same trip count
  for (int r = threadIdx.x / R; r < n; r += NT / R) {
    int si = s0;
    while (si + 1 < s1 && r >= lrow0[si + 1]) si++;
    const DevSegment s = lseg[si];
    const int idx = r - lrow0[si];
#define PMX_ROW(EX, EY, EZ) row_split<T, EX, EY, EZ, R, MODE, true>(s, idx, part, M, val, lx, f, y, omega)
    switch (s.ex * 100 + s.ey * 10 + s.ez) {
      case 111: PMX_ROW(1, 1, 1); break;
      case 211: PMX_ROW(2, 1, 1); break;
      case 121: PMX_ROW(1, 2, 1); break;
      case 221: PMX_ROW(2, 2, 1); break;
      case 112: PMX_ROW(1, 1, 2); break;
      case 212: PMX_ROW(2, 1, 2); break;
      case 122: PMX_ROW(1, 2, 2); break;
      default: PMX_ROW(2, 2, 2); break;
    }
#undef PMX_ROW
  }
}
// stage the segment table, colour ranges and row offsets, and x with
// Dirichlet rows x = f, in LDS
template <typename T, int NT>
__device__ __forceinline__ void wg_stage(const DevSegment *__restrict__ segs, int nseg,
                                         const int *__restrict__ colour_seg, int ncol, int N, const T *x,
                                         const T *__restrict__ f, const int *__restrict__ bnd, int nb,
                                         DevSegment *lseg, int *lcs, int *lrow0, int *lrows, T *lx) {
  for (int i = threadIdx.x; i < nseg; i += NT) lseg[i] = segs[i];
  for (int i = threadIdx.x; i <= ncol; i += NT) lcs[i] = colour_seg[i];
  for (int i = threadIdx.x; i < N; i += NT) lx[i] = x[i];
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += NT) lx[bnd[i]] = f[bnd[i]];
  for (int c = threadIdx.x; c < ncol; c += NT) {
    int n = 0;
    for (int si = lcs[c]; si < lcs[c + 1]; si++) {
      lrow0[si] = n;
      n += (int)lseg[si].count;
    }
    lrows[c] = n;
  }
  __syncthreads();
}

// A whole SOR (ssor = 0) or SSOR (ssor = 1) sweep of a small level in one
// workgroup: the level's x is copied into LDS (Dirichlet rows x = f), the
// colours run one after another (backwards again for SSOR) with a workgroup
// barrier between them, and x is copied back.  The rows of a colour are
// enumerated across its residue classes, R lanes per row as in k_rows, and
// each row is computed by row_split, so the iterate is bit-identical to the
// launch-per-colour sweep.  On a level whose colours have a few hundred rows
// this replaces 2 x 27 launch latencies per SSOR sweep by barriers, and the
// x round trips through memory by LDS accesses.
template <typename T, int R>
__global__ __launch_bounds__(sweep_threads<T>()) void k_sweep_wg(const DevSegment *__restrict__ segs, int nseg,
                                                            const int *__restrict__ colour_seg, int ncol, int M,
                                                            const T *__restrict__ val, T *x,
                                                            const T *__restrict__ f, double omega, int ssor,
                                                            const int *__restrict__ bnd, int nb) {
  constexpr int NT = sweep_threads<T>();
  extern __shared__ __align__(16) unsigned char sweep_lds[];
  T *lx = reinterpret_cast<T *>(sweep_lds);
  __shared__ DevSegment lseg[kMaxSegs];
  __shared__ int lcs[kMaxColours + 1];
  __shared__ int lrow0[kMaxSegs + 1];
  __shared__ int lrows[kMaxColours];
  const int N = M * M * M;
  wg_stage<T, NT>(segs, nseg, colour_seg, ncol, N, x, f, bnd, nb, lseg, lcs, lrow0, lrows, lx);
  const int nsteps = ssor ? 2 * ncol : ncol;
  for (int step = 0; step < nsteps; step++) {
    const int c = step < ncol ? step : 2 * ncol - 1 - step;
    wg_colour_rows<T, R, 3, NT>(lseg, lcs, lrow0, lrows, c, M, val, lx, f, lx, omega);
    __syncthreads();
  }
  for (int i = threadIdx.x; i < N; i += NT) x[i] = lx[i];
}

// ||f - A x|| of the LDS iterate over all rows (Dirichlet rows: 0), the
// same value in every thread.  d: global scratch of N entries whose
// Dirichlet entries hold 0.
template <typename T, int R, int NT>
__device__ double wg_defect_norm(const DevSegment *lseg, const int *lcs, const int *lrow0, const int *lrows,
                                 int ncol, int M, const T *__restrict__ val, const T *lx, const T *__restrict__ f,
                                 T *d, double *lred) {
  const int N = M * M * M;
  for (int c = 0; c < ncol; c++) wg_colour_rows<T, R, 1, NT>(lseg, lcs, lrow0, lrows, c, M, val, lx, f, d, 0.0);
  __syncthreads();
  double s = 0;
  for (int i = threadIdx.x; i < N; i += NT) s += (double)d[i] * (double)d[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if (threadIdx.x % 64 == 0) lred[threadIdx.x / 64] = s;
  __syncthreads();
  double t = 0;
  for (int w = 0; w < NT / 64; w++) t += lred[w];
  __syncthreads();
  return sqrt(t);
}
// The coarse-level loop of TMultiGrid3D::Smooth(CoarseSmooth)
// (MultiGrid3D.C:225-337) in one workgroup: res = ||f - A x||, then while
// res > red * res_0 and fewer than maxit sweeps: one Jacobi (kind 1, omega
// damped), SOR (2) or SSOR (3) sweep and res again.  The sweeps compute each
// row as the launched sweeps do (bit-identical iterates); d and xj are
// global scratch of N entries (defect, Jacobi output).
template <typename T, int R>
__global__ __launch_bounds__(sweep_threads<T>()) void k_coarse_wg(const DevSegment *__restrict__ segs, int nseg,
                                                             const int *__restrict__ colour_seg, int ncol, int M,
                                                             const T *__restrict__ val, T *x,
                                                             const T *__restrict__ f, double omega, int kind,
                                                             int maxit, double red, const int *__restrict__ bnd,
                                                             int nb, T *d, T *xj) {
  constexpr int NT = sweep_threads<T>();
  extern __shared__ __align__(16) unsigned char sweep_lds[];
  T *lx = reinterpret_cast<T *>(sweep_lds);
  __shared__ DevSegment lseg[kMaxSegs];
  __shared__ int lcs[kMaxColours + 1];
  __shared__ int lrow0[kMaxSegs + 1];
  __shared__ int lrows[kMaxColours];
  __shared__ double lred[NT / 64];
  const int N = M * M * M;
  for (int i = threadIdx.x; i < nb; i += NT) {
    d[bnd[i]] = T(0);
    xj[bnd[i]] = f[bnd[i]];
  }
  wg_stage<T, NT>(segs, nseg, colour_seg, ncol, N, x, f, bnd, nb, lseg, lcs, lrow0, lrows, lx);
  double res = wg_defect_norm<T, R, NT>(lseg, lcs, lrow0, lrows, ncol, M, val, lx, f, d, lred);
  const double reduction = red * res;
  for (int it = 0; res > reduction && it < maxit; it++) {
    if (kind == 1) {
      for (int c = 0; c < ncol; c++) wg_colour_rows<T, R, 2, NT>(lseg, lcs, lrow0, lrows, c, M, val, lx, f, xj, omega);
      __syncthreads();
      for (int i = threadIdx.x; i < N; i += NT) lx[i] = xj[i];
      __syncthreads();
    } else {
      const int nsteps = kind == 3 ? 2 * ncol : ncol;
      for (int step = 0; step < nsteps; step++) {
        const int c = step < ncol ? step : 2 * ncol - 1 - step;
        wg_colour_rows<T, R, 3, NT>(lseg, lcs, lrow0, lrows, c, M, val, lx, f, lx, omega);
        __syncthreads();
      }
    }
    res = wg_defect_norm<T, R, NT>(lseg, lcs, lrow0, lrows, ncol, M, val, lx, f, d, lred);
  }
  for (int i = threadIdx.x; i < N; i += NT) x[i] = lx[i];
}

// flag = 1 if a and b differ in any bit (all writers store the same value)
__global__ void k_bits_differ(const double *__restrict__ a, const double *__restrict__ b, int64_t n,
                              int *__restrict__ flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (__double_as_longlong(a[i]) != __double_as_longlong(b[i])) *flag = 1;
}
// ============================================================ transfers
// 1D taps of the nodal interpolation from the coarse to the fine lattice
// (the prolongation matrices of FEDatabase3D.C:1656 for Q1/Q2 Lagrange,
// averaged over the cells sharing a fine DOF as MGComponents3D.C:208-488
// does, which for a continuous interpolant is the interpolant itself).
// Taps are held in fixed slots (Q1: 2 prolongation / 3 restriction slots per
// axis, Q2: 3 / 5) in increasing lattice order; a slot that a node does not
// use gets weight 0 and a clamped index, and the kernels add its term as an
// exact 0.  Every load of a node is then issued before the first product
// (no data-dependent trip counts), and the sum runs over the same terms in
// the same order as over the used taps alone, so the result is the same bit
// for bit.
template <int ORDER> struct TransferTaps {
  static constexpr int P = ORDER == 1 ? 2 : 3;  // prolongation taps per axis
  static constexpr int R = ORDER == 1 ? 3 : 5;  // restriction taps per axis
};
template <int ORDER>
__device__ __forceinline__ void prolong_taps(int x, int nc, int *ci, double *w) {
  if (ORDER == 1) {
    const int c = x >> 1;
    const bool odd = x & 1;
    ci[0] = c; w[0] = odd ? 0.5 : 1.0;
    ci[1] = odd ? c + 1 : c; w[1] = odd ? 0.5 : 0.0;
    return;
  }
  int e = x >> 2;
  if (e > nc - 1) e = nc - 1;
  const int r = x - 4 * e;
  ci[0] = 2 * e; ci[1] = 2 * e + 1; ci[2] = 2 * e + 2;
  w[0] = r == 0 ? 1.0 : r == 1 ? 0.375 : r == 3 ? -0.125 : 0.0;
  w[1] = r == 2 ? 1.0 : (r == 1 || r == 3) ? 0.75 : 0.0;
  w[2] = r == 4 ? 1.0 : r == 1 ? -0.125 : r == 3 ? 0.375 : 0.0;
}
// transposed taps: fine nodes feeding coarse node y (DefectRestriction = P^T)
template <int ORDER>
__device__ __forceinline__ void restrict_taps(int y, int Mf, int *fi, double *w) {
  constexpr int R = TransferTaps<ORDER>::R;
  int off[R];
  double ww[R];
  if (ORDER == 1) {
    off[0] = -1; off[1] = 0; off[2] = 1;
    ww[0] = 0.5; ww[1] = 1.0; ww[2] = 0.5;
  } else {
    const bool even = (y & 1) == 0;
    off[0] = -3; off[1] = -1; off[2] = 0; off[3] = 1; off[4] = 3;
    ww[0] = even ? -0.125 : 0.0; ww[1] = even ? 0.375 : 0.75; ww[2] = 1.0; ww[3] = even ? 0.375 : 0.75;
    ww[4] = even ? -0.125 : 0.0;
  }
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int x = 2 * y + off[k];
    const bool in = x >= 0 && x < Mf;
    fi[k] = in ? x : 0;
    w[k] = in ? ww[k] : 0.0;
  }
}

// ADD = 0: xf = P xc.  ADD = 1: xf += damp (P xc with 0 on the lattice
// surface), i.e. prolongation, CorrectNodes and the damped update in one
// pass (the Dirichlet nodes of every layout are its lattice surface)
template <typename T, int ORDER, int ADD>
__global__ __launch_bounds__(kBlock) void k_prolong(const T *__restrict__ xc, T *__restrict__ xf, int Mc, int Mf,
                                                    double damp) {
  constexpr int P = TransferTaps<ORDER>::P;
  // 32-bit index arithmetic (the launcher checks that the lattice fits)
  const unsigned id = blockIdx.x * kBlock + threadIdx.x;
  const unsigned Nf = (unsigned)Mf * Mf * Mf;
  if (id >= Nf) return;
  const unsigned yz = id / (unsigned)Mf;
  const int x = id - yz * Mf, y = yz % (unsigned)Mf, z = yz / (unsigned)Mf;
  const int nc = (Mc - 1) / ORDER;
  int ix[P], iy[P], iz[P];
  double wx[P], wy[P], wz[P];
  prolong_taps<ORDER>(x, nc, ix, wx);
  prolong_taps<ORDER>(y, nc, iy, wy);
  prolong_taps<ORDER>(z, nc, iz, wz);
  T v[P][P][P];
#pragma unroll
  for (int c = 0; c < P; c++)
#pragma unroll
    for (int b = 0; b < P; b++) {
      const T *row = xc + (unsigned)Mc * (iy[b] + (unsigned)Mc * iz[c]);
#pragma unroll
      for (int a = 0; a < P; a++) v[c][b][a] = row[ix[a]];
    }
  // arithmetic in T: the tap weights and their products are exact in FP32
  T s = 0;
#pragma unroll
  for (int c = 0; c < P; c++)
#pragma unroll
    for (int b = 0; b < P; b++) {
      const T wzy = (T)wz[c] * (T)wy[b];
#pragma unroll
      for (int a = 0; a < P; a++) {
        const T t = wzy * (T)wx[a] * v[c][b][a];
        s += (wz[c] != 0.0 && wy[b] != 0.0 && wx[a] != 0.0) ? t : T(0);
      }
    }
  if (ADD) {
    const bool surface = x == 0 || y == 0 || z == 0 || x == Mf - 1 || y == Mf - 1 || z == Mf - 1;
    xf[id] += (T)damp * (surface ? T(0) : s);
  } else {
    xf[id] = s;
  }
}
// zero_surface: 0 on the coarse lattice surface (CorrectDefect fused)
template <typename T, int ORDER>
__global__ __launch_bounds__(kBlock) void k_restrict(T *__restrict__ rc, const T *__restrict__ rf, int Mc, int Mf,
                                                     int zero_surface) {
  constexpr int R = TransferTaps<ORDER>::R;
  const unsigned id = blockIdx.x * kBlock + threadIdx.x;
  const unsigned Nc = (unsigned)Mc * Mc * Mc;
  if (id >= Nc) return;
  const unsigned yz = id / (unsigned)Mc;
  const int x = id - yz * Mc, y = yz % (unsigned)Mc, z = yz / (unsigned)Mc;
  int ix[R], iy[R], iz[R];
  double wx[R], wy[R], wz[R];
  restrict_taps<ORDER>(x, Mf, ix, wx);
  restrict_taps<ORDER>(y, Mf, iy, wy);
  restrict_taps<ORDER>(z, Mf, iz, wz);
  T s = 0;
#pragma unroll
  for (int c = 0; c < R; c++)
#pragma unroll
    for (int b = 0; b < R; b++) {
      const T wzy = (T)wz[c] * (T)wy[b];
      const T *row = rf + (unsigned)Mf * (iy[b] + (unsigned)Mf * iz[c]);
#pragma unroll
      for (int a = 0; a < R; a++) {
        const T t = wzy * (T)wx[a] * row[ix[a]];
        s += (wz[c] != 0.0 && wy[b] != 0.0 && wx[a] != 0.0) ? t : T(0);
      }
    }
  if (zero_surface && (x == 0 || y == 0 || z == 0 || x == Mc - 1 || y == Mc - 1 || z == Mc - 1)) s = 0;
  rc[id] = s;
}

// ============================================================ assembly
// Coefficients at quadrature point q of cell (cx, cy, cz), in the canonical
// (axis-aligned) frame of the lattice: {w|detJ|, eps, b0, b1, b2, c, f, delta}.
// with_delta = false: the record without its stabilisation parameter
// (o[7] = 0), for callers that read it from the per-point buffer of
// k_qp_delta
template <bool CHEAP = false>
__device__ __forceinline__ void qp_coeffs(const double *__restrict__ prm, const double *__restrict__ qref, int n,
                                          int cx, int cy, int cz, int q, int form, double *o,
                                          const uint8_t *__restrict__ corner = nullptr, bool with_delta = true) {
  const double h = 1.0 / n;
  const double xi = qref[4 * q], eta = qref[4 * q + 1], zeta = qref[4 * q + 2], w = qref[4 * q + 3];
  // THexaAffin::GetOrigFromRef: X = xc0 + xc1 xi + ... with xc0 the cell centre
  const double X = (cx * h + 0.5 * h) + 0.5 * h * xi;
  const double Y = (cy * h + 0.5 * h) + 0.5 * h * eta;
  const double Z = (cz * h + 0.5 * h) + 0.5 * h * zeta;
  const QpCoeffs c = example_coeffs<CHEAP>(prm, X, Y, Z);
  if (!with_delta) {
    o[0] = w * (h * h * h / 8.0);
    o[1] = c.eps; o[2] = c.b0; o[3] = c.b1; o[4] = c.b2; o[5] = c.c; o[6] = c.f; o[7] = 0.0;
    return;
  }
  if (corner && (int)prm[PMX_PRM_CELL_MEASURE] == 4) {
    // the coefficients at the reference's first point of the cell, (a, a, a)
    // with a = qref[0] in the cell's own frame: centre + (h/2) a (e1 + e2 + e3)
    const int code = corner[cx + (int64_t)n * (cy + (int64_t)n * cz)];
    const double a = 0.5 * h * qref[0];
    const QpCoeffs c0 = example_coeffs(prm, (cx * h + 0.5 * h) + ((code & 1) ? a : -a),
                                       (cy * h + 0.5 * h) + ((code & 2) ? a : -a),
                                       (cz * h + 0.5 * h) + ((code & 4) ? a : -a));
    qp_record(prm, form, h, w, c, o, &c0);
  } else {
    qp_record(prm, form, h, w, c, o);
  }
}
// Stage 1 (coefficient functions that are expensive to evaluate): the
// coefficients of every quadrature point of every cell, once.
__global__ __launch_bounds__(kBlock) void k_qp_data(const double *__restrict__ prm, const double *__restrict__ qref,
                                                    int nq, int n, double *__restrict__ out,
                                                    const uint8_t *__restrict__ corner) {
  const int64_t id = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  const int64_t ncell = (int64_t)n * n * n;
  if (id >= ncell * nq) return;
  const int q = id % nq;
  const int64_t cell = id / nq;
  const int cx = cell % n, cy = (cell / n) % n, cz = cell / ((int64_t)n * n);
  qp_coeffs(prm, qref, n, cx, cy, cz, q, (int)prm[PMX_PRM_FORM], out + id * 8, corner);
}
// The stabilisation parameter alone (record entry 7) of every quadrature
// point of every cell, once: the per-row evaluation of the inline
// coefficient functions would otherwise compute it (a square root, a
// hyperbolic tangent and divisions) once for each of the up to 8 rows that
// share a cell.  Point-major, dl[q * n^3 + cell], so that the rows of a
// wavefront (cells two apart along x) read nearby entries.  Grid: x over the
// cells of one z-plane, y = q * n + cz.
__global__ __launch_bounds__(kBlock) void k_qp_delta(const double *__restrict__ prm, const double *__restrict__ qref,
                                                     int n, double *__restrict__ dl) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= n * n) return;
  const int cx = t % n, cy = t / n;
  const int cz = blockIdx.y % n, q = blockIdx.y / n;
  double o[8];
  qp_coeffs<true>(prm, qref, n, cx, cy, cz, q, (int)prm[PMX_PRM_FORM], o);
  dl[(int64_t)q * n * n * n + t + (int64_t)n * n * cz] = o[7];
}
// forms whose records carry a stabilisation parameter
__host__ __device__ inline bool form_has_delta(int form) {
  return form == PMX_FORM_SDFEM || form == PMX_FORM_MASS_SUPG || form == PMX_FORM_RHS_SUPG ||
         form == PMX_FORM_TCD_SUPG_AK;
}
// The curl-Fourier convection (fourier_b) without a cosine per mode and
// point.  The quadrature rules are tensor products (point q = ix + np (iy +
// np iz) at (xi_ix, xi_iy, xi_iz)), so the phase of mode m at point q of cell
// (cx, cy, cz), 2 pi k . X + phi, is a_x(cx, ix) + a_y(cy, iy) + a_z(cz, iz)
// with a_d = 2 pi k_d X_d along one axis (phi in a_x), and
// cos(phase) = Re(e^{i a_x} e^{i a_y} e^{i a_z}).  The per-axis factors are
// one table per level (mode, axis, cell index, 1D point); a point then costs
// two complex products per mode instead of a cosine, and the three loads of
// a mode touch one cache line per wavefront.  The values differ from the
// direct evaluation by rounding only.
__device__ __forceinline__ int rule_points_1d(int nq) { return nq <= 8 ? 2 : (nq <= 27 ? 3 : 4); }
__global__ __launch_bounds__(kBlock) void k_fourier_tables(const double *__restrict__ prm,
                                                           const double *__restrict__ qref, int nq, int n,
                                                           double *__restrict__ tab) {
  const int np = rule_points_1d(nq);
  const int id = blockIdx.x * kBlock + threadIdx.x;
  if (id >= kFourierModes * 3 * n * np) return;
  const int ci = id % (n * np), d = (id / (n * np)) % 3, m = id / (3 * n * np);
  const int c = ci / np, i = ci % np;
  const double twopi = 6.283185307179586476925286766559;
  const double h = 1.0 / n;
  const double *t = prm + PMX_PRM_FOURIER + m * 5;
  // point i of the 1D rule is the first coordinate of point q = i
  const double X = (c * h + 0.5 * h) + 0.5 * h * qref[4 * i];
  const double a = twopi * t[1 + d] * X + (d == 0 ? t[4] : 0.0);
  tab[2 * id] = cos(a);
  tab[2 * id + 1] = sin(a);
}
// The curl-Fourier coefficients (eps, b, c = 0, f = 1 as example_coeffs) at
// the 1D points (ix, iy, iz) of cell (cx, cy, cz) from the table above.
__device__ __forceinline__ QpCoeffs fourier_coeffs_tab(const double *__restrict__ prm,
                                                       const double *__restrict__ tab, int n, int np, int cx, int cy,
                                                       int cz, int ix, int iy, int iz) {
  const double twopi = 6.283185307179586476925286766559;
  const int64_t L = (int64_t)n * np;  // entries per (mode, axis)
  double dA[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll 4
  for (int m = 0; m < kFourierModes; m++) {
    const double *t = prm + PMX_PRM_FOURIER + m * 5;
    const double *fx = tab + 2 * ((m * 3 + 0) * L + cx * np + ix);
    const double *fy = tab + 2 * ((m * 3 + 1) * L + cy * np + iy);
    const double *fz = tab + 2 * ((m * 3 + 2) * L + cz * np + iz);
    const double r1 = fx[0] * fy[0] - fx[1] * fy[1], i1 = fx[0] * fy[1] + fx[1] * fy[0];
    const double cs = t[0] * (r1 * fz[0] - i1 * fz[1]) * twopi;
    const int c = m / 8;
    dA[c][0] += cs * t[1];
    dA[c][1] += cs * t[2];
    dA[c][2] += cs * t[3];
  }
  const double sc = prm[PMX_PRM_BSCALE];
  QpCoeffs k;
  k.eps = prm[PMX_PRM_EPS];
  k.b0 = sc * (dA[2][1] - dA[1][2]);
  k.b1 = sc * (dA[0][2] - dA[2][0]);
  k.b2 = sc * (dA[1][0] - dA[0][1]);
  k.c = 0;
  k.f = 1.0;
  return k;
}
// Stage 1 for the curl-Fourier example: the records of k_qp_data with the
// coefficients from the table above.  A thread takes the np points of one
// x-line (iy, iz) of a cell: per mode it reads the y and z factors once and
// the np x factors, so the table reads per point (the bound of this kernel:
// 24 modes x 3 complex factors per point) drop from 3 to 1 + 2/np complex
// values.  Per point and mode the products and sums are those of
// fourier_coeffs_tab, in the same order.
// Grid: x over the (cx, iy, iz) lines of one row of cells, y = cy + n cz;
// the rule size is a template parameter so that no index needs a runtime
// division.
template <int NQ>
__global__ __launch_bounds__(kBlock) void k_qp_data_fourier(const double *__restrict__ prm,
                                                            const double *__restrict__ qref, int n,
                                                            const double *__restrict__ tab, double *__restrict__ out,
                                                            const uint8_t *__restrict__ corner) {
  constexpr int np = NQ == 8 ? 2 : (NQ == 27 ? 3 : 4);
  static_assert(np * np * np == NQ, "tensor rule");
  constexpr int NL = np * np;  // x-lines per cell
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= n * NL) return;
  const int cx = t / NL, r = t % NL, iy = r % np, iz = r / np;
  const int cy = blockIdx.y % n, cz = blockIdx.y / n;
  const double twopi = 6.283185307179586476925286766559;
  const int64_t L = (int64_t)n * np;  // entries per (mode, axis)
  double dA[np][3][3];
#pragma unroll
  for (int i = 0; i < np; i++)
#pragma unroll
    for (int c = 0; c < 3; c++) dA[i][c][0] = dA[i][c][1] = dA[i][c][2] = 0.0;
  // modes 8 c .. 8 c + 7 belong to component c of the potential (the
  // component loop is unrolled so that every accumulator index is a constant)
#pragma unroll
  for (int c = 0; c < 3; c++) {
#pragma unroll 2
    for (int mm = 0; mm < kFourierModes / 3; mm++) {
      const int m = (kFourierModes / 3) * c + mm;
      const double *tm = prm + PMX_PRM_FOURIER + m * 5;
      const double *fx = tab + 2 * ((m * 3 + 0) * L + cx * np);
      const double *fy = tab + 2 * ((m * 3 + 1) * L + cy * np + iy);
      const double *fz = tab + 2 * ((m * 3 + 2) * L + cz * np + iz);
      const double fy0 = fy[0], fy1 = fy[1], fz0 = fz[0], fz1 = fz[1];
      const double t0 = tm[0], t1 = tm[1], t2 = tm[2], t3 = tm[3];
#pragma unroll
      for (int ix = 0; ix < np; ix++) {
        const double fx0 = fx[2 * ix], fx1 = fx[2 * ix + 1];
        const double r1 = fx0 * fy0 - fx1 * fy1, i1 = fx0 * fy1 + fx1 * fy0;
        const double cs = t0 * (r1 * fz0 - i1 * fz1) * twopi;
        dA[ix][c][0] += cs * t1;
        dA[ix][c][1] += cs * t2;
        dA[ix][c][2] += cs * t3;
      }
    }
  }
  const double sc = prm[PMX_PRM_BSCALE];
  const bool cm4 = corner && (int)prm[PMX_PRM_CELL_MEASURE] == 4;
  QpCoeffs k0;
  if (cm4) {
    // the coefficients at the reference's first point of the cell, (a, a, a)
    // in its own frame: 1D point 0 (a) along an axis where the frame points
    // forward, np - 1 (-a; the rules are symmetric) where it points back
    const int code = corner[cx + (int64_t)n * (cy + (int64_t)n * cz)];
    k0 = fourier_coeffs_tab(prm, tab, n, np, cx, cy, cz, (code & 1) ? 0 : np - 1, (code & 2) ? 0 : np - 1,
                            (code & 4) ? 0 : np - 1);
  }
  const int64_t cell = cx + (int64_t)n * (blockIdx.y);
#pragma unroll
  for (int ix = 0; ix < np; ix++) {
    QpCoeffs k;
    k.eps = prm[PMX_PRM_EPS];
    k.b0 = sc * (dA[ix][2][1] - dA[ix][1][2]);
    k.b1 = sc * (dA[ix][0][2] - dA[ix][2][0]);
    k.b2 = sc * (dA[ix][1][0] - dA[ix][0][1]);
    k.c = 0;
    k.f = 1.0;
    const int q = ix + np * r;
    double *o = out + (cell * NQ + q) * 8;
    if (cm4) qp_record(prm, (int)prm[PMX_PRM_FORM], 1.0 / n, qref[4 * q + 3], k, o, &k0);
    else qp_record(prm, (int)prm[PMX_PRM_FORM], 1.0 / n, qref[4 * q + 3], k, o);
  }
}
// examples whose coefficients are cheaper to evaluate again than to read back
__host__ __device__ inline bool coeffs_inline(int example) {
  return example == PMX_EX_LAPLACE || example == PMX_EX_SINE_CD || example == PMX_EX_HEAT_AMC ||
         example == PMX_EX_ROTATING;
}

// Stage 2: row gather.  The thread of active node p visits the cells that
// contain p, forms row a of each local matrix
//   K_ab = sum_q w [ u0 phi_b + u . grad phi_b ]
// with s_a = phi_a + delta b.grad phi_a, u0 = c s_a, u = eps grad phi_a + b s_a
// (BilinearAssemble / _SD, ConvDiff3D.C:118-145,182-213; mass and TCD forms,
// TCD3D.C:37-322), and adds it into its own CSS row.  No two threads write
// the same row, so no atomics and no colouring of cells are needed, and the
// writes of a wavefront are coalesced.
// The incident cells per direction are compile-time (a Q2 node with an odd
// coordinate lies in one cell along it, every other node in two), so the
// loops are uniform across the wavefront and the basis tables are read with
// scalar loads; TWO selects the second matrix of the SUPG TCD form.
template <int ORDER, int EX, int EY, int EZ, bool TWO, bool INL, bool RONLY>
__device__ __forceinline__ void assemble_row(const DevSegment &s, int idx, int M, int n, int nq, int form,
                                             const double *__restrict__ prm, const double *__restrict__ qref,
                                             const double *__restrict__ qp, const double *__restrict__ tab,
                                             double *__restrict__ val, double *__restrict__ val2,
                                             double *__restrict__ rhs) {
  constexpr int NCX = (ORDER == 2 && EX == 1) ? 1 : 2;
  constexpr int NCY = (ORDER == 2 && EY == 1) ? 1 : 2;
  constexpr int NCZ = (ORDER == 2 && EZ == 1) ? 1 : 2;
  constexpr int K = ORDER + 1, NB = K * K * K;
  const int i = idx % s.cx;
  const int t = idx / s.cx;
  const int j = t % s.cy, k = t / s.cy;
  const int px = s.x0 + s.stride * i, py = s.y0 + s.stride * j, pz = s.z0 + s.stride * k;
  const int p = px + M * (py + M * pz);
  const double g = 2.0 * n;  // reference -> physical derivative factor 2/h
  double *v = val + vbase(s, idx);
  double *v2 = TWO ? val2 + vbase(s, idx) : nullptr;
  constexpr int cnt = kSlice;
  double frhs = 0;
  // Q1 rows of the one-matrix forms (27 slots) are accumulated in registers
  // over the 8 cells (loops unrolled, every slot a compile-time index) and
  // stored once; Q2 rows and the two-matrix form (whose unrolled loops need
  // too many registers) add each cell's contribution into the values (Q2:
  // cleared beforehand; Q1: the first cell of a slot stores)
  constexpr bool REG = ORDER == 1 && !TWO;
  constexpr int UNR = REG ? 2 : 1;
  double racc[REG ? 27 : 1];
  if constexpr (REG) {
#pragma unroll
    for (int q = 0; q < 27; q++) racc[q] = 0.0;
  }
  // incident cells per direction: lattice coordinate x lies in cells
  // floor((x-1)/ORDER) .. floor(x/ORDER) (two cells at a vertex plane)
  const int cx0 = (px - 1) / ORDER, cy0 = (py - 1) / ORDER, cz0 = (pz - 1) / ORDER;
#pragma unroll UNR
  for (int iz = 0; iz < NCZ; iz++)
#pragma unroll UNR
    for (int iy = 0; iy < NCY; iy++)
#pragma unroll UNR
      for (int ix = 0; ix < NCX; ix++) {
        const int cx = cx0 + ix, cy = cy0 + iy, cz = cz0 + iz;
        const int ax = px - ORDER * cx, ay = py - ORDER * cy, az = pz - ORDER * cz;
        const int a = ax + K * (ay + K * az);
        const int64_t cell = cx + (int64_t)n * (cy + (int64_t)n * cz);
        const double *Q = qp + cell * nq * 8;
        double acc[NB], acc2[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) acc[b] = acc2[b] = 0;
        // inline coefficients: qp holds the per-point stabilisation
        // parameters of k_qp_delta (nullptr for the forms without one), read
        // one point ahead
        const int64_t ncell = (int64_t)n * n * n;
        double dl_next = (INL && qp) ? qp[cell] : 0.0;
        for (int q = 0; q < nq; q++) {
          double Qq[8];
          if (INL) {
            qp_coeffs<true>(prm, qref, n, cx, cy, cz, q, form, Qq, nullptr, false);
            Qq[7] = dl_next;
            if (qp) dl_next = qp[(q + 1 < nq ? q + 1 : q) * ncell + cell];
          } else {
#pragma unroll
            for (int e = 0; e < 8; e++) Qq[e] = Q[8 * q + e];
          }
          const double w = Qq[0], eps = Qq[1], b0 = Qq[2], b1 = Qq[3], b2 = Qq[4];
          const double cc = Qq[5], ff = Qq[6], dl = Qq[7];
          const double *Ta = tab + (q * NB + a) * 4;
          const double fa = Ta[0], gax = g * Ta[1], gay = g * Ta[2], gaz = g * Ta[3];
          const double bga = b0 * gax + b1 * gay + b2 * gaz;
          double u0, ux, uy, uz, k0 = 0;
          if (RONLY) {
            // RhsAssemble / RhsAssemble_SUPG: f (phi_a + delta b.grad phi_a)
            // (delta = 0 for the Galerkin rhs)
            frhs += w * ff * (fa + dl * bga);
            continue;
          }
          if (form == PMX_FORM_MASS || form == PMX_FORM_MASS_SUPG) {
            // the mass matrix; the rhs is f (phi_a + delta b.grad phi_a) with
            // delta = 0 for MatrixMRhsAssemble (TCD3D.C:64, 117)
            u0 = fa; ux = uy = uz = 0;
            frhs += w * ff * (fa + dl * bga);
          } else {
            const double sa = fa + dl * bga;
            u0 = cc * sa;
            ux = eps * gax + b0 * sa;
            uy = eps * gay + b1 * sa;
            uz = eps * gaz + b2 * sa;
            frhs += w * ff * sa;
            k0 = dl * bga;
          }
          u0 *= w; ux *= w * g; uy *= w * g; uz *= w * g; k0 *= w;
          const double *Tq = tab + q * NB * 4;
#pragma unroll
          for (int b = 0; b < NB; b++) {
            const double *Tb = Tq + 4 * b;
            acc[b] += u0 * Tb[0] + ux * Tb[1] + uy * Tb[2] + uz * Tb[3];
            if (TWO) acc2[b] += k0 * Tb[0];
          }
        }
        if (RONLY) continue;
        // add the row into the CSS slots of node p
#pragma unroll
        for (int b = 0; b < NB; b++) {
          const int bx = b % K, by = (b / K) % K, bz = b / (K * K);
          const int dx = ORDER * cx + bx - px, dy = ORDER * cy + by - py, dz = ORDER * cz + bz - pz;
          const int slot = (dx + EX) + (2 * EX + 1) * ((dy + EY) + (2 * EY + 1) * (dz + EZ));
          if constexpr (REG) {
            racc[slot] += acc[b];
          } else if (ORDER == 1 && (ix == 0 || bx == 1) && (iy == 0 || by == 1) && (iz == 0 || bz == 1)) {
            // Q1: the first of the row's cells (in loop order) that meets
            // the slot, (max(0, dx), max(0, dy), max(0, dz)), stores; the
            // later ones add, so the values need no clearing beforehand
            v[slot * cnt] = acc[b];
            if (TWO) v2[slot * cnt] = acc2[b];
          } else {
            v[slot * cnt] += acc[b];
            if (TWO) v2[slot * cnt] += acc2[b];
          }
        }
      }
  if constexpr (REG && !RONLY) {
#pragma unroll
    for (int q = 0; q < 27; q++) v[q * cnt] = racc[q];
  }
  rhs[p] = frhs;
}

template <int ORDER, bool TWO, bool INL, bool RONLY>
__global__ __launch_bounds__(kRowsPerBlock) void k_assemble(const BlockInfo *__restrict__ blocks,
                                                     const DevSegment *__restrict__ segs, int M, int n, int nq,
                                                     const double *__restrict__ prm, const double *__restrict__ qref,
                                                     const double *__restrict__ qp, const double *__restrict__ tab,
                                                     double *__restrict__ val, double *__restrict__ val2,
                                                     double *__restrict__ rhs) {
  const BlockInfo bi = blocks[xcd_remap(blockIdx.x, gridDim.x)];
  const DevSegment s = segs[bi.seg];
  const int idx = bi.base + threadIdx.x;
  if (idx >= s.count) return;
  const int form = (int)prm[PMX_PRM_FORM];
#define PMX_AROW(O, EX, EY, EZ) \
  assemble_row<O, EX, EY, EZ, TWO, INL, RONLY>(s, idx, M, n, nq, form, prm, qref, qp, tab, val, val2, rhs)
  if (ORDER == 1) {
    PMX_AROW(1, 1, 1, 1);
    return;
  }
  const int code = s.ex * 100 + s.ey * 10 + s.ez;
  switch (code) {
    case 111: PMX_AROW(2, 1, 1, 1); break;
    case 211: PMX_AROW(2, 2, 1, 1); break;
    case 121: PMX_AROW(2, 1, 2, 1); break;
    case 221: PMX_AROW(2, 2, 2, 1); break;
    case 112: PMX_AROW(2, 1, 1, 2); break;
    case 212: PMX_AROW(2, 2, 1, 2); break;
    case 122: PMX_AROW(2, 1, 2, 2); break;
    default: PMX_AROW(2, 2, 2, 2); break;
  }
#undef PMX_AROW
}
__global__ void k_scatter(double *__restrict__ y, const int *__restrict__ idx, const double *__restrict__ v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[idx[i]] = v[i];
}
void scatter(double *y, const int *idx, const double *v, int n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_scatter, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, y, idx, v, n);
}
// Dirichlet rhs: RHS[dof] = g(node) (AssembleMat3D.C:1031-1096)
__global__ void k_boundary_rhs(const int *__restrict__ bnd, int nbd, int M, const double *__restrict__ prm,
                               double *__restrict__ rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbd) return;
  const int p = bnd[i];
  const double h = 1.0 / (M - 1);
  const int x = p % M, y = (p / M) % M, z = p / (M * M);
  rhs[p] = example_bound_value(prm, x * h, y * h, z * h);
}

// ============================================================ CSR export
// entries in the reference CSR (sorted rows, Dirichlet rows = 1) from CSS values
template <typename T>
__global__ void k_css_to_csr(const int *__restrict__ rowptr, const int *__restrict__ kcol,
                             const int *__restrict__ dof2lat, int64_t N, int64_t active, int M, int S,
                             const int *__restrict__ segidx, const DevSegment *__restrict__ segs,
                             const T *__restrict__ val, double *__restrict__ out) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= N) return;
  const int b = rowptr[r], e = rowptr[r + 1];
  if (r >= active) { out[b] = 1.0; return; }
  const int p = dof2lat[r];
  const int x = p % M, y = (p / M) % M, z = p / (M * M);
  const DevSegment s = segs[segidx[(x % S) + S * ((y % S) + S * (z % S))]];
  const int idx = ((x - s.x0) / s.stride) + s.cx * (((y - s.y0) / s.stride) + s.cy * ((z - s.z0) / s.stride));
  for (int n = b; n < e; n++) {
    const int q = dof2lat[kcol[n]];
    const int dx = q % M - x, dy = (q / M) % M - y, dz = q / (M * M) - z;
    const int slot = (dx + s.ex) + (2 * s.ex + 1) * ((dy + s.ey) + (2 * s.ey + 1) * (dz + s.ez));
    if constexpr (std::is_same<T, _Float16>::value)
      out[n] = (double)val[v16slice(s, idx) + v16off(s.ns, idx % kSlice, slot)];
    else
      out[n] = (double)val[vbase(s, idx) + (int64_t)slot * kSlice];
  }
}

// The inverse of k_css_to_csr for the active rows: the entries of a CSR
// matrix with the level's pattern (reference numbering) into the CSS values.
__global__ void k_csr_to_css(const int *__restrict__ rowptr, const int *__restrict__ kcol,
                             const int *__restrict__ dof2lat, int64_t active, int M, int S,
                             const int *__restrict__ segidx, const DevSegment *__restrict__ segs,
                             const double *__restrict__ in, double *__restrict__ val) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= active) return;
  const int b = rowptr[r], e = rowptr[r + 1];
  const int p = dof2lat[r];
  const int x = p % M, y = (p / M) % M, z = p / (M * M);
  const DevSegment s = segs[segidx[(x % S) + S * ((y % S) + S * (z % S))]];
  const int idx = ((x - s.x0) / s.stride) + s.cx * (((y - s.y0) / s.stride) + s.cy * ((z - s.z0) / s.stride));
  for (int n = b; n < e; n++) {
    const int q = dof2lat[kcol[n]];
    const int dx = q % M - x, dy = (q / M) % M - y, dz = q / (M * M) - z;
    const int slot = (dx + s.ex) + (2 * s.ex + 1) * ((dy + s.ey) + (2 * s.ey + 1) * (dz + s.ez));
    val[vbase(s, idx) + (int64_t)slot * kSlice] = in[n];
  }
}

// ============================================================ permutations
template <typename T, typename U>
__global__ void k_gather(T *__restrict__ out, const U *__restrict__ in, const int *__restrict__ map, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = (T)in[map[i]];
}

// ============================================================ generic CSR
// y = A x on a caller CSR matrix (MatVect on any TSquareMatrix3D): 16 lanes per row
template <typename T>
__global__ __launch_bounds__(kBlock) void k_csr_spmv(int n, const int *__restrict__ rowptr,
                                                     const int *__restrict__ kcol, const T *__restrict__ val,
                                                     const T *__restrict__ x, T *__restrict__ y,
                                                     const T *__restrict__ b) {
  const int lane = threadIdx.x & 15;
  const int row = (blockIdx.x * kBlock + threadIdx.x) >> 4;
  if (row >= n) return;
  double s = 0;
  for (int j = rowptr[row] + lane; j < rowptr[row + 1]; j += 16) s += (double)val[j] * (double)x[kcol[j]];
  for (int o = 8; o > 0; o >>= 1) s += __shfl_down(s, o, 16);
  if (lane == 0) y[row] = b ? (T)((double)b[row] - s) : (T)s;
}

// ============================================================ dense coarse solve
// The coarsest operator is A = [A_II A_IB; 0 I] (interior rows I, Dirichlet
// rows B identity), so A^-1 f = [A_II^-1 (f_I - A_IB f_B); f_B].  The device
// keeps W = A_II^-1 [I_I, -A_IB] as an n_I x N matrix in lattice column order:
// y_I = W f (one wavefront per row, FP64), y_B = f_B.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gemv_coarse(int nI, int N, const double *__restrict__ W,
                                                        const int *__restrict__ iidx, const T *__restrict__ x,
                                                        T *__restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int row = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  if (row >= nI) return;
  double s = 0;
  const double *a = W + (int64_t)row * N;
  for (int j = lane; j < N; j += 64) s += a[j] * (double)x[j];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if (lane == 0) y[iidx[row]] = (T)s;
}

// ============================================================ coarse inverse
// Augmented [A_II | R] (n_I x (n_I + N), row-major), R[i][p] = 1 for the
// interior node p of row i, -a(i, p) for Dirichlet columns p.
__global__ void k_css_to_aug(const BlockInfo *__restrict__ blocks, const DevSegment *__restrict__ segs, int M,
                             const int *__restrict__ cidx, int nI, const double *__restrict__ val,
                             double *__restrict__ aug, int64_t ld) {
  const BlockInfo bi = blocks[blockIdx.x];
  const DevSegment s = segs[bi.seg];
  const int idx = bi.base + threadIdx.x;
  if (idx >= s.count) return;
  const int i = idx % s.cx, t = idx / s.cx, j = t % s.cy, k = t / s.cy;
  const int x = s.x0 + s.stride * i, y = s.y0 + s.stride * j, z = s.z0 + s.stride * k;
  const int p = x + M * (y + M * z);
  const int r = cidx[p];
  double *row = aug + (int64_t)r * ld;
  row[nI + p] = 1.0;
  int slot = 0;
  for (int dz = -s.ez; dz <= s.ez; dz++)
    for (int dy = -s.ey; dy <= s.ey; dy++)
      for (int dx = -s.ex; dx <= s.ex; dx++, slot++) {
        const int q = (x + dx) + M * ((y + dy) + M * (z + dz));
        const double v = val[vbase(s, idx) + (int64_t)slot * kSlice];
        const int cq = cidx[q];
        if (cq >= 0) row[cq] = v;
        else row[nI + q] = -v;
      }
}
// Gauss-Jordan step c with partial pivoting, in one launch: every block finds
// the same pivot row (largest |a(r, c)| among rows not yet used as pivots,
// lowest index on ties), then eliminates column c from the other rows of its
// share.  Rows are not swapped or scaled (the pivot row and column c are
// read-only during the step); piv[c] records the pivot row.
__global__ __launch_bounds__(256) void k_gj_step(double *__restrict__ aug, int nI, int64_t ld, int c,
                                                 int *__restrict__ piv) {
  __shared__ unsigned char used[8192];
  __shared__ double bv[256];
  __shared__ int bix[256];
  for (int r = threadIdx.x; r < nI; r += 256) used[r] = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < c; k += 256) used[piv[k]] = 1;
  __syncthreads();
  double best = -1.0;
  int arg = nI;
  for (int r = threadIdx.x; r < nI; r += 256) {
    if (used[r]) continue;
    const double a = fabs(aug[(int64_t)r * ld + c]);
    if (a > best) { best = a; arg = r; }
  }
  bv[threadIdx.x] = best;
  bix[threadIdx.x] = arg;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const double b2 = bv[threadIdx.x + o];
      const int i2 = bix[threadIdx.x + o];
      if (b2 > bv[threadIdx.x] || (b2 == bv[threadIdx.x] && i2 < bix[threadIdx.x])) {
        bv[threadIdx.x] = b2;
        bix[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  const int pr = bix[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) piv[c] = pr;
  const double *prow = aug + (int64_t)pr * ld;
  const double pinv = 1.0 / prow[c];
  // this block's rows: r = blockIdx.x * 4 + 0..3, columns c+1 .. ld-1
  const int64_t w = ld - (c + 1);
  for (int q = 0; q < 4; q++) {
    const int r = blockIdx.x * 4 + q;
    if (r >= nI || r == pr) continue;
    double *row = aug + (int64_t)r * ld;
    const double f = row[c] * pinv;
    if (f == 0.0) continue;
    for (int64_t j = threadIdx.x; j < w; j += 256) row[c + 1 + j] -= f * prow[c + 1 + j];
  }
}
// W[i][p] = aug[piv[i]][nI + p] / aug[piv[i]][i]
__global__ void k_gj_finish(const double *__restrict__ aug, int nI, int N, int64_t ld, const int *__restrict__ piv,
                            double *__restrict__ W) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)nI * N) return;
  const int i = (int)(id / N), p = (int)(id % N);
  const double *row = aug + (int64_t)piv[i] * ld;
  W[id] = row[nI + p] / row[i];
}

// ============================================================ FGMRES scalars
// Givens update of column i of H and of s (FgmresIte.C:194-222, 447-459),
// H stored column-major with leading dimension m+1; writes |s[i+1]| to res.
__global__ void k_givens(double *H, double *cs, double *sn, double *s, int i, int ld, double *res) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double *h = H + (int64_t)i * ld;
  for (int k = 0; k < i; k++) {
    const double t = cs[k] * h[k] + sn[k] * h[k + 1];
    h[k + 1] = -sn[k] * h[k] + cs[k] * h[k + 1];
    h[k] = t;
  }
  const double dx = h[i], dy = h[i + 1];
  double c, sg;
  if (dy == 0.0) { c = 1.0; sg = 0.0; }
  else if (fabs(dy) > fabs(dx)) { const double t = dx / dy; sg = 1.0 / sqrt(1.0 + t * t); c = t * sg; }
  else { const double t = dy / dx; c = 1.0 / sqrt(1.0 + t * t); sg = t * c; }
  cs[i] = c; sn[i] = sg;
  {
    const double t = c * h[i] + sg * h[i + 1];
    h[i + 1] = -sg * h[i] + c * h[i + 1];
    h[i] = t;
  }
  {
    const double t = c * s[i] + sg * s[i + 1];
    s[i + 1] = -sg * s[i] + c * s[i + 1];
    s[i] = t;
  }
  res[0] = fabs(s[i + 1]);
}
// back substitution of UpdateGmresIterate (FgmresIte.C:224-243): y = H^{-1} s
__global__ void k_backsolve(const double *H, const double *s, double *y, int k, int ld) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = 0; i <= k; i++) y[i] = s[i];
  for (int i = k; i >= 0; i--) {
    y[i] /= H[(int64_t)i * ld + i];
    for (int j = i - 1; j >= 0; j--) y[j] -= H[(int64_t)i * ld + j] * y[i];
  }
}
// x += sum_j y[j] z_j over k+1 vectors (z stored contiguously, stride n)
template <typename T, typename X>
__global__ __launch_bounds__(kBlock) void k_multi_axpy(X *__restrict__ x, const T *__restrict__ z,
                                                       const double *__restrict__ y, int k1, int64_t n,
                                                       int64_t stride) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    double s = 0;
    for (int j = 0; j < k1; j++) s += y[j] * (double)z[j * stride + i];
    x[i] += (X)s;
  }
}
// write a scalar into H(i,j) from a reduction result
__global__ void k_set_scalar(double *dst, const double *src, double scale) {
  if (threadIdx.x == 0 && blockIdx.x == 0) dst[0] = scale * src[0];
}

// ================================================================ launchers
static inline int grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65535 * 8) g = 65535 * 8;
  return (int)(g < 1 ? 1 : g);
}

void check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
void fill(T *x, T v, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n)), dim3(kBlock), 0, st, x, v, n);
}
bool bits_equal(const double *a, const double *b, int64_t n, int *flag, hipStream_t st) {
  fill<int>(flag, 0, 1, st);
  hipLaunchKernelGGL(k_bits_differ, dim3(grid_for(n)), dim3(kBlock), 0, st, a, b, n, flag);
  int h = 1;
  if (hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    throw std::runtime_error("bits_equal: copy failed");
  return h == 0;
}
template <typename T>
void copy(T *y, const T *x, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_copy<T>, dim3(grid_for(n)), dim3(kBlock), 0, st, y, x, n);
}
template <typename T>
void axpy(T *y, const T *x, const double *alpha_dev, double sgn, double ac, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy<T>, dim3(grid_for(n)), dim3(kBlock), 0, st, y, x, alpha_dev, sgn, ac, n);
}
template <typename T, typename U>
void convert(T *y, const U *x, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL((k_convert<T, U>), dim3(grid_for(n)), dim3(kBlock), 0, st, y, x, n);
}
template <typename T, typename U>
void add_convert(T *y, const U *x, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL((k_add_convert<T, U>), dim3(grid_for(n)), dim3(kBlock), 0, st, y, x, n);
}
template <typename T>
void scale_inv(T *y, const T *x, const double *nrm, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_scale_inv<T>, dim3(grid_for(n)), dim3(kBlock), 0, st, y, x, nrm, n);
}
// dot product with one workgroup per kDotChunk elements (no grid-stride
// loop): the number of partial sums grows with n, ceil(n / kDotChunk)
__global__ __launch_bounds__(kBlock) void k_dot_chunk(const double *__restrict__ x, const double *__restrict__ y,
                                                      double *__restrict__ part, int64_t n) {
  const int64_t base = (int64_t)blockIdx.x * kDotChunk;
  double s = 0;
#pragma unroll
  for (int k = 0; k < kDotChunk / kBlock; k++) {
    const int64_t i = base + k * kBlock + threadIdx.x;
    if (i < n) s += x[i] * y[i];
  }
  s = block_sum<double>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
void dot_chunked(const double *x, const double *y, int64_t n, double *part, double *out, int do_sqrt,
                 hipStream_t st) {
  const int64_t nb = (n + kDotChunk - 1) / kDotChunk;
  hipLaunchKernelGGL(k_dot_chunk, dim3((unsigned)nb), dim3(kBlock), 0, st, x, y, part, n);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, st, part, (int)nb, out, do_sqrt);
}
template <typename T>
void dot(const T *x, const T *y, int64_t n, double *part, double *out, int do_sqrt, hipStream_t st) {
  int g = (int)((n + kBlock - 1) / kBlock);
  if (g > kRedBlocks) g = kRedBlocks;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_dot_partial<T>, dim3(g), dim3(kBlock), 0, st, x, y, part, n);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, st, part, g, out, do_sqrt);
}

// ------------------------------------------------------- step length control
template <typename T>
__global__ void k_diff(T *__restrict__ e, const T *__restrict__ u, const T *__restrict__ uold, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    e[i] = u[i] - uold[i];
}
// omega = num / nom (param0 if nom <= 0 or |omega| < 0.1), u = uold + omega (u - uold)
template <typename T>
__global__ void k_step_length(T *__restrict__ u, const T *__restrict__ uold, const double *__restrict__ num_nom,
                              double param0, int64_t n) {
  const double num = num_nom[0], nom = num_nom[1];
  double omega = nom > 0 ? num / nom : param0;
  if (fabs(omega) < 0.1) omega = param0;
  const T a = (T)omega;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    u[i] = uold[i] + a * (u[i] - uold[i]);
}
template <typename T>
void axpy_diff(T *e, const T *u, const T *uold, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_diff<T>, dim3(grid_for(n)), dim3(kBlock), 0, st, e, u, uold, n);
}
template <typename T>
void step_length_update(T *u, const T *uold, const double *num_nom, double param0, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_step_length<T>, dim3(grid_for(n)), dim3(kBlock), 0, st, u, uold, num_nom, param0, n);
}

// ------------------------------------------------------- Gram-Schmidt
// sum of the np partials in k_reduce_final's order, in every workgroup (the
// same bits everywhere), broadcast to the workgroup
__device__ __forceinline__ double sum_partials(const double *__restrict__ part, int np) {
  __shared__ double sh_total;
  double s = 0;
  for (int i = threadIdx.x; i < np; i += kBlock) s += part[i];
  s = block_sum<double>(s);
  if (threadIdx.x == 0) sh_total = s;
  __syncthreads();
  return sh_total;
}
// one modified Gram-Schmidt step: h = (w, v) from the partials of the
// previous launch (stored to *h by workgroup 0), w -= h v, and the partials
// of (w, next) for the following step (next = w: the squared norm).  The
// thread's first kPrefetch elements are loaded before the partials are
// summed, so their latency overlaps that reduction; elements and their
// order per thread are those of the plain grid-stride loop.
constexpr int kPrefetch = 8;
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mgs_step(T *w, const T *__restrict__ v, const T *next,
                                                     const double *__restrict__ part_in, int np,
                                                     double *__restrict__ part_out, double *__restrict__ h,
                                                     int64_t n) {
  const int64_t i0 = blockIdx.x * (int64_t)kBlock + threadIdx.x, stride = (int64_t)gridDim.x * kBlock;
  const bool self = next == w;
  T wv[kPrefetch], vv[kPrefetch], nv[kPrefetch];
#pragma unroll
  for (int k = 0; k < kPrefetch; k++) {
    const int64_t i = i0 + k * stride;
    wv[k] = vv[k] = nv[k] = T(0);
    if (i < n) {
      wv[k] = w[i];
      vv[k] = v[i];
      if (!self) nv[k] = next[i];
    }
  }
  const double hv = sum_partials(part_in, np);
  if (blockIdx.x == 0 && threadIdx.x == 0) h[0] = hv;
  const T at = (T)(-hv);
  double s = 0;
#pragma unroll
  for (int k = 0; k < kPrefetch; k++) {
    const int64_t i = i0 + k * stride;
    if (i < n) {
      T wi = wv[k];
      wi += at * vv[k];
      w[i] = wi;
      s += (double)wi * (double)(self ? wi : nv[k]);
    }
  }
  for (int64_t i = i0 + kPrefetch * stride; i < n; i += stride) {
    T wi = w[i];
    wi += at * v[i];
    w[i] = wi;
    s += (double)wi * (double)(self ? wi : next[i]);
  }
  s = block_sum<T>(s);
  if (threadIdx.x == 0) part_out[blockIdx.x] = s;
}
// *h = ||w|| from the partials of (w, w), y = w / ||w|| (prefetch as above)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_norm_scale(const T *__restrict__ w, const double *__restrict__ part,
                                                       int np, double *__restrict__ h, T *__restrict__ y, int64_t n) {
  const int64_t i0 = blockIdx.x * (int64_t)kBlock + threadIdx.x, stride = (int64_t)gridDim.x * kBlock;
  T wv[kPrefetch];
#pragma unroll
  for (int k = 0; k < kPrefetch; k++) {
    const int64_t i = i0 + k * stride;
    wv[k] = i < n ? w[i] : T(0);
  }
  const double nv = sqrt(sum_partials(part, np));
  if (blockIdx.x == 0 && threadIdx.x == 0) h[0] = nv;
  const T sc = (T)(1.0 / nv);
#pragma unroll
  for (int k = 0; k < kPrefetch; k++) {
    const int64_t i = i0 + k * stride;
    if (i < n) y[i] = wv[k] * sc;
  }
  for (int64_t i = i0 + kPrefetch * stride; i < n; i += stride) y[i] = w[i] * sc;
}
template <typename T>
void mgs_arnoldi(T *w, const T *V, int64_t n, int k1, double *h, T *vnext, double *part, hipStream_t st) {
  int g = (int)((n + kBlock - 1) / kBlock);
  if (g > kRedBlocks) g = kRedBlocks;
  if (g < 1) g = 1;
  double *pa = part, *pb = part + kRedBlocks;
  hipLaunchKernelGGL(k_dot_partial<T>, dim3(g), dim3(kBlock), 0, st, w, V, pa, n);
  for (int k = 0; k < k1; k++) {
    const T *next = k + 1 < k1 ? V + (size_t)(k + 1) * n : w;
    hipLaunchKernelGGL(k_mgs_step<T>, dim3(g), dim3(kBlock), 0, st, w, V + (size_t)k * n, next, pa, g, pb, h + k, n);
    std::swap(pa, pb);
  }
  hipLaunchKernelGGL(k_norm_scale<T>, dim3(g), dim3(kBlock), 0, st, w, pa, g, h + k1, vnext, n);
}

// Per-dispatch timing of the row kernels: while a timer is installed on the
// calling thread, css_apply launches through hipExtLaunchKernelGGL with a
// start and a stop event of the timer around each dispatch (the events take
// the dispatch's own begin and end, not the gaps between dispatches).
static thread_local DispatchTimer *g_timer = nullptr;
void set_dispatch_timer(DispatchTimer *t) { g_timer = t; }
double DispatchTimer::total_ms() const {
  double s = 0;
  for (int k = 0; k < used; k++) {
    float ms = 0;
    check(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]), "dispatch timer");
    s += ms;
  }
  return s;
}
DispatchTimer::~DispatchTimer() {
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
}
static void timer_events(hipEvent_t *e0, hipEvent_t *e1) {
  DispatchTimer &t = *g_timer;
  if (2 * t.used + 1 >= (int)t.ev.size()) {
    for (int k = 0; k < 2; k++) {
      hipEvent_t e;
      check(hipEventCreate(&e), "event");
      t.ev.push_back(e);
    }
  }
  *e0 = t.ev[2 * t.used];
  *e1 = t.ev[2 * t.used + 1];
  t.used++;
}

template <typename T>
void css_apply(const DevLayout &D, int mode, const T *val, const T *x, const T *f, T *y, double omega, int colour,
               hipStream_t st) {
  const RowBlock *blk = colour < 0 ? D.rows_all : D.rows_colour + D.colour_off[colour];
  const int nblk = colour < 0 ? D.nblocks_all : D.colour_off[colour + 1] - D.colour_off[colour];
  if (nblk == 0) return;
#define PMX_ROWS(MODE, R)                                                                                      \
  if (g_timer) {                                                                                               \
    hipEvent_t e0, e1;                                                                                         \
    timer_events(&e0, &e1);                                                                                    \
    hipExtLaunchKernelGGL((k_rows<T, MODE, R>), dim3(nblk), dim3(kRowsPerBlock * R), 0, st, e0, e1, 0, blk, D.M,  \
                          val, x, f, y, omega);                                                                \
  } else                                                                                                       \
    hipLaunchKernelGGL((k_rows<T, MODE, R>), dim3(nblk), dim3(kRowsPerBlock * R), 0, st, blk, D.M, val, x, f, y,  \
                       omega)
#define PMX_ROWS_R(MODE)                                                                                       \
  switch (row_lanes(D)) {                                                                                      \
    case 1: PMX_ROWS(MODE, 1); break;                                                                          \
    case 2: PMX_ROWS(MODE, 2); break;                                                                          \
    default: PMX_ROWS(MODE, 4); break;                                                                         \
  }
  switch (mode) {
    case 0: PMX_ROWS_R(0); break;
    case 1: PMX_ROWS_R(1); break;
    case 2: PMX_ROWS_R(2); break;
    default: PMX_ROWS_R(3); break;
  }
#undef PMX_ROWS_R
#undef PMX_ROWS
}
void css_sor_scaled(const DevLayout &D, const _Float16 *val16, const float *dinv, float *x, const float *f,
                    double omega, int colour, hipStream_t st) {
  const RowBlock *blk = D.rows_colour + D.colour_off[colour];
  const int nblk = D.colour_off[colour + 1] - D.colour_off[colour];
  if (nblk == 0) return;
#define PMX_ROWS(R)                                                                                            \
  if (g_timer) {                                                                                               \
    hipEvent_t e0, e1;                                                                                         \
    timer_events(&e0, &e1);                                                                                    \
    hipExtLaunchKernelGGL((k_sor_scaled<R>), dim3(nblk), dim3(kRowsPerBlock * R), 0, st, e0, e1, 0, blk, D.M,     \
                          val16, dinv, x, f, omega);                                                           \
  } else                                                                                                       \
    hipLaunchKernelGGL((k_sor_scaled<R>), dim3(nblk), dim3(kRowsPerBlock * R), 0, st, blk, D.M, val16, dinv, x,  \
                       f, omega)
#ifdef PMX_SCALED_LANES
  const int lanes = PMX_SCALED_LANES;  // a fixed lane count (tuning builds)
#else
  const int lanes = row_lanes(D);
#endif
  switch (lanes) {
    case 1: PMX_ROWS(1); break;
    case 2: PMX_ROWS(2); break;
    default: PMX_ROWS(4); break;
  }
#undef PMX_ROWS
}
// Calibration kernels of the dispatch measurements: a dispatch that does
// nothing, and a plain streaming read (16 bytes per load, two loads per lane)
// whose sum reaches memory only in a case that never occurs.
__global__ void k_empty_dispatch() {}
__global__ __launch_bounds__(256) void k_stream_read(const float4 *__restrict__ a, int64_t n4,
                                                     float *__restrict__ sink) {
  const int64_t i = (int64_t)blockIdx.x * 512 + threadIdx.x;
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    if (i + 256 * k < n4) {
      const float4 v = a[i + 256 * k];
      s += v.x + v.y + v.z + v.w;
    }
  }
  if (s == -1234567.0f) sink[0] = s;
}
void empty_dispatch(int blocks, hipStream_t st) {
  if (blocks < 1) return;
  if (g_timer) {
    hipEvent_t e0, e1;
    timer_events(&e0, &e1);
    hipExtLaunchKernelGGL(k_empty_dispatch, dim3(blocks), dim3(kRowsPerBlock), 0, st, e0, e1, 0);
  } else {
    hipLaunchKernelGGL(k_empty_dispatch, dim3(blocks), dim3(kRowsPerBlock), 0, st);
  }
}
void stream_read(const void *p, int64_t bytes, float *sink, hipStream_t st) {
  const int64_t n4 = bytes / 16;
  if (n4 < 1) return;
  if ((reinterpret_cast<uintptr_t>(p) & 15) != 0) throw std::invalid_argument("stream_read: 16-byte alignment");
  const unsigned g = (unsigned)((n4 + 511) / 512);
  const float4 *a = static_cast<const float4 *>(p);
  if (g_timer) {
    hipEvent_t e0, e1;
    timer_events(&e0, &e1);
    hipExtLaunchKernelGGL(k_stream_read, dim3(g), dim3(256), 0, st, e0, e1, 0, a, n4, sink);
  } else {
    hipLaunchKernelGGL(k_stream_read, dim3(g), dim3(256), 0, st, a, n4, sink);
  }
}
void scaled_values(const DevLayout &D, const float *val, _Float16 *val16, float *dinv, hipStream_t st) {
  if (D.nblocks_all == 0) return;
  hipLaunchKernelGGL(k_scaled_values, dim3(D.nblocks_all), dim3(kRowsPerBlock), 0, st, D.blocks_all, D.segs, D.M,
                     val, val16, dinv);
}
template <typename T>
void sweep_wg(const DevLayout &D, const T *val, T *x, const T *f, double omega, int ssor, hipStream_t st) {
  if (D.nsegs > kMaxSegs || D.ncolours > kMaxColours) throw std::logic_error("sweep_wg: segment table too large");
  if (!sweep_wg_fits(D, sizeof(T))) throw std::logic_error("sweep_wg: level does not fit in LDS");
  const size_t lds = (size_t)D.N * sizeof(T);
#define PMX_SWEEP(R)                                                                                            \
  hipLaunchKernelGGL((k_sweep_wg<T, R>), dim3(1), dim3(sweep_threads<T>()), lds, st, D.segs, D.nsegs,             \
                     D.colour_seg_dev, D.ncolours, D.M, val, x, f, omega, ssor, D.boundary, D.nboundary)
  switch (row_lanes(D)) {
    case 1: PMX_SWEEP(1); break;
    case 2: PMX_SWEEP(2); break;
    default: PMX_SWEEP(4); break;
  }
#undef PMX_SWEEP
}
template <typename T>
void coarse_wg(const DevLayout &D, const T *val, T *x, const T *f, double omega, int kind, int maxit, double red,
               T *d, T *xj, hipStream_t st) {
  if (D.nsegs > kMaxSegs || D.ncolours > kMaxColours) throw std::logic_error("coarse_wg: segment table too large");
  if (!sweep_wg_fits(D, sizeof(T))) throw std::logic_error("coarse_wg: level does not fit in LDS");
  if (kind < 1 || kind > 3) throw std::invalid_argument("coarse_wg: smoother must be 1, 2 or 3");
  const size_t lds = (size_t)D.N * sizeof(T);
#define PMX_COARSE(R)                                                                                           \
  hipLaunchKernelGGL((k_coarse_wg<T, R>), dim3(1), dim3(sweep_threads<T>()), lds, st, D.segs, D.nsegs,           \
                     D.colour_seg_dev, D.ncolours, D.M, val, x, f, omega, kind, maxit, red, D.boundary,          \
                     D.nboundary, d, xj)
  switch (row_lanes(D)) {
    case 1: PMX_COARSE(1); break;
    case 2: PMX_COARSE(2); break;
    default: PMX_COARSE(4); break;
  }
#undef PMX_COARSE
}
template <typename T>
void boundary(const DevLayout &D, int mode, T *y, const T *x, const T *f, hipStream_t st) {
  const int g = (D.nboundary + kBlock - 1) / kBlock;
  switch (mode) {
    case 0: hipLaunchKernelGGL((k_boundary<T, 0>), dim3(g), dim3(kBlock), 0, st, D.boundary, D.nboundary, y, x, f); break;
    case 1: hipLaunchKernelGGL((k_boundary<T, 1>), dim3(g), dim3(kBlock), 0, st, D.boundary, D.nboundary, y, x, f); break;
    case 2: hipLaunchKernelGGL((k_boundary<T, 2>), dim3(g), dim3(kBlock), 0, st, D.boundary, D.nboundary, y, x, f); break;
    default: hipLaunchKernelGGL((k_boundary<T, 3>), dim3(g), dim3(kBlock), 0, st, D.boundary, D.nboundary, y, x, f); break;
  }
}
template <typename T>
void prolongate(const DevLayout &C, const DevLayout &F, const T *xc, T *xf, hipStream_t st) {
  if (F.N >= (int64_t(1) << 31)) throw std::runtime_error("prolongate: lattice too large for 32-bit indices");
  if (F.order == 2)
    hipLaunchKernelGGL((k_prolong<T, 2, 0>), dim3(grid_for(F.N)), dim3(kBlock), 0, st, xc, xf, C.M, F.M, 1.0);
  else
    hipLaunchKernelGGL((k_prolong<T, 1, 0>), dim3(grid_for(F.N)), dim3(kBlock), 0, st, xc, xf, C.M, F.M, 1.0);
}
template <typename T>
void prolongate_add(const DevLayout &C, const DevLayout &F, const T *xc, T *xf, double damp, hipStream_t st) {
  if (F.N >= (int64_t(1) << 31)) throw std::runtime_error("prolongate: lattice too large for 32-bit indices");
  if (F.order == 2)
    hipLaunchKernelGGL((k_prolong<T, 2, 1>), dim3(grid_for(F.N)), dim3(kBlock), 0, st, xc, xf, C.M, F.M, damp);
  else
    hipLaunchKernelGGL((k_prolong<T, 1, 1>), dim3(grid_for(F.N)), dim3(kBlock), 0, st, xc, xf, C.M, F.M, damp);
}
template <typename T>
void restrict_(const DevLayout &C, const DevLayout &F, T *rc, const T *rf, hipStream_t st, bool zero_surface) {
  if (F.N >= (int64_t(1) << 31)) throw std::runtime_error("restrict: lattice too large for 32-bit indices");
  if (F.order == 2)
    hipLaunchKernelGGL((k_restrict<T, 2>), dim3(grid_for(C.N)), dim3(kBlock), 0, st, rc, rf, C.M, F.M,
                       (int)zero_surface);
  else
    hipLaunchKernelGGL((k_restrict<T, 1>), dim3(grid_for(C.N)), dim3(kBlock), 0, st, rc, rf, C.M, F.M,
                       (int)zero_surface);
}
template <typename T>
void gemv_coarse(int nI, int N, const double *W, const int *iidx, const T *x, T *y, hipStream_t st) {
  const unsigned g = (unsigned)((nI * 64 + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_gemv_coarse<T>, dim3(g), dim3(kBlock), 0, st, nI, N, W, iidx, x, y);
}
void coarse_inverse_dense(const DevLayout &D, const double *val, const int *cidx, int nI, double *aug, int *piv,
                          double *W, hipStream_t st) {
  if (nI > 8192) throw std::invalid_argument("coarse inverse: too many interior nodes");
  const int N = (int)D.N;
  const int64_t ld = nI + (int64_t)N;
  fill<double>(aug, 0.0, nI * ld, st);
  hipLaunchKernelGGL(k_css_to_aug, dim3(D.nblocks_all), dim3(kRowsPerBlock), 0, st, D.blocks_all, D.segs, D.M, cidx,
                     nI, val, aug, ld);
  const unsigned g = (unsigned)((nI + 3) / 4);
  for (int c = 0; c < nI; c++) hipLaunchKernelGGL(k_gj_step, dim3(g), dim3(256), 0, st, aug, nI, ld, c, piv);
  hipLaunchKernelGGL(k_gj_finish, dim3((unsigned)(((int64_t)nI * N + 255) / 256)), dim3(256), 0, st, aug, nI, N, ld,
                     piv, W);
}
template <typename T, typename U>
void gather(T *out, const U *in, const int *map, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL((k_gather<T, U>), dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, out, in,
                     map, n);
}
template <typename T>
void csr_spmv(int n, const int *rp, const int *kc, const T *val, const T *x, T *y, hipStream_t st) {
  const int g = (int)(((int64_t)n * 16 + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_csr_spmv<T>, dim3(g), dim3(kBlock), 0, st, n, rp, kc, val, x, y, (const T *)nullptr);
}
template <typename T>
void csr_defect(int n, const int *rp, const int *kc, const T *val, const T *x, const T *b, T *r, hipStream_t st) {
  const int g = (int)(((int64_t)n * 16 + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_csr_spmv<T>, dim3(g), dim3(kBlock), 0, st, n, rp, kc, val, x, r, b);
}
void csr_to_css(const DevLayout &D, const int *rowptr, const int *kcol, const int *dof2lat, int64_t active,
                const double *in, double *val, hipStream_t st) {
  if (active == 0) return;
  hipLaunchKernelGGL(k_csr_to_css, dim3((unsigned)((active + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, rowptr,
                     kcol, dof2lat, active, D.M, D.S, D.segidx, D.segs, in, val);
}
template <typename T>
void css_to_csr(const DevLayout &D, const int *rowptr, const int *kcol, const int *dof2lat, int64_t active,
                const T *val, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_css_to_csr<T>, dim3((unsigned)((D.N + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, rowptr,
                     kcol, dof2lat, D.N, active, D.M, D.S, D.segidx, D.segs, val, out);
}

// ---- Q2 one-matrix forms on MFMA: local matrices per cell, then a row gather.
// With Phi_q the 4 x 27 table of (phi, dphi/dxi, dphi/deta, dphi/dzeta) at
// point q, every form above is K = sum_q Phi_q^T D_q Phi_q (rows: test
// functions a, columns: trial functions b) with, for the convection-diffusion
// forms (g = 2/h folds the reference derivatives into D, s = delta):
//   D[0][0] = w c          D[0][d] = w g b_d
//   D[e][0] = w g c s b_e  D[e][d] = w g^2 (eps [e == d] + s b_e b_d)
// and the right-hand side f_a = sum_q r_q . Phi_q[:, a] with r = w f (1, g s b);
// the mass form has D[0][0] = w and r = (w f, 0, 0, 0).  One wavefront
// computes a cell as a 32 x 32 product (27 x 27 padded, column 27 carrying the
// right-hand side) over K = 4 nq, one v_mfma_f64_16x16x4f64 per 16 x 16 tile
// and quadrature point: lane l holds A = Phi_q[l/16][16R + l%16] and
// B = (D_q Phi_q)[l/16][16C + l%16]; the result lane l, register r is entry
// (16R + l/16 + 4r, 16C + l%16).  The tables sit in LDS ([q][comp][32]).
typedef double doublex4 __attribute__((ext_vector_type(4)));
// 8 wavefronts share one copy of the tables (64 KB for the 64-point rule)
constexpr int kCellWaves = 8, kCellThreads = 64 * kCellWaves;
constexpr int kKcRow = 28;  // doubles per local row in the cell buffer: 27 entries + rhs
#ifndef PMX_CELLS_PER_PASS
#define PMX_CELLS_PER_PASS 2
#endif
constexpr int kCellsPerPass = PMX_CELLS_PER_PASS;  // cells per wavefront pass of k_cell_mfma
// CPC cells per wavefront pass: every quadrature point's table values are
// read from LDS once for the CPC cells (10 LDS reads per lane and point
// serve CPC x 4 MFMAs), with CPC independent accumulation chains per
// wavefront.  Each cell's products and sums are those of one cell alone.
template <bool INL, bool MASS, int CPC>
__global__ __launch_bounds__(kCellThreads) void k_cell_mfma(int n, int nq, int form, const double *__restrict__ prm,
                                                  const double *__restrict__ qref, const double *__restrict__ qp,
                                                  const double *__restrict__ tab, double *__restrict__ kc) {
  extern __shared__ __align__(16) unsigned char cell_lds[];
  double *lt = reinterpret_cast<double *>(cell_lds);  // [nq][4][32]
  const int tid = threadIdx.x;
  for (int i = tid; i < nq * 128; i += kCellThreads) {
    const int q = i >> 7, comp = (i >> 5) & 3, b = i & 31;
    lt[i] = b < 27 ? tab[((int64_t)q * 27 + b) * 4 + comp] : 0.0;
  }
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int l = tid & 63, comp = l >> 4, col = l & 15;
  const double g = 2.0 * n;
  // per-lane selectors of row `comp` of D (branch-free loop body keeps the
  // accumulators in place across quadrature points)
  const bool r0 = comp == 0;
  const double s1 = comp == 1 ? 1.0 : 0.0, s2 = comp == 2 ? 1.0 : 0.0, s3 = comp == 3 ? 1.0 : 0.0;
  const bool rhs_col = col == 11;  // column 27 = 16 + 11 of the second column tile
  const int64_t ncells = (int64_t)n * n * n;
  const int64_t ngroups = (ncells + CPC - 1) / CPC;
  for (int64_t grp = (int64_t)blockIdx.x * kCellWaves + wave; grp < ngroups;
       grp += (int64_t)gridDim.x * kCellWaves) {
    // the cells of this pass (a pass past the last cell repeats the last
    // one and does not store it)
    int64_t cell[CPC];
    int cx[CPC], cy[CPC], cz[CPC];
#pragma unroll
    for (int k = 0; k < CPC; k++) {
      cell[k] = grp * CPC + k < ncells ? grp * CPC + k : ncells - 1;
      cx[k] = (int)(cell[k] % n);
      cy[k] = (int)((cell[k] / n) % n);
      cz[k] = (int)(cell[k] / ((int64_t)n * n));
    }
    doublex4 a00[CPC], a01[CPC], a10[CPC], a11[CPC];
#pragma unroll
    for (int k = 0; k < CPC; k++) {
      a00[k] = doublex4{0.0, 0.0, 0.0, 0.0};
      a01[k] = a00[k];
      a10[k] = a00[k];
      a11[k] = a00[k];
    }
    // coefficients of the next quadrature point are loaded one step ahead
    double Qn[CPC][8];
    if (!INL) {
#pragma unroll
      for (int k = 0; k < CPC; k++)
#pragma unroll
        for (int e = 0; e < 8; e++) Qn[k][e] = qp[cell[k] * nq * 8 + e];
    }
    for (int q = 0; q < nq; q++) {
      const double *T = lt + q * 128;
      const double t0 = T[col], t1 = T[32 + col], t2 = T[64 + col], t3 = T[96 + col];
      const double u0 = T[16 + col], u1 = T[48 + col], u2 = T[80 + col], u3 = T[112 + col];
      const double av0 = T[comp * 32 + col], av1 = T[comp * 32 + 16 + col];
      const int qn = q + 1 < nq ? q + 1 : q;
#pragma unroll
      for (int k = 0; k < CPC; k++) {
        double Qq[8];
        if (INL) {
          qp_coeffs<true>(prm, qref, n, cx[k], cy[k], cz[k], q, form, Qq);
        } else {
#pragma unroll
          for (int e = 0; e < 8; e++) Qq[e] = Qn[k][e];
#pragma unroll
          for (int e = 0; e < 8; e++) Qn[k][e] = qp[(cell[k] * nq + qn) * 8 + e];
        }
        const double w = Qq[0], eps = Qq[1], cc = Qq[5], ff = Qq[6], dl = Qq[7];
        const double b0 = Qq[2], b1 = Qq[3], b2 = Qq[4];
        // row comp of D = s v + t e_comp with v = (c, g b): s = w for row 0 and
        // w g delta b_comp below it, t = w g^2 eps on the diagonal of rows 1-3;
        // the right-hand-side weight is f s
        double d0, d1, d2, d3, rr;
        if (MASS) {
          d0 = r0 ? w : 0.0;
          d1 = d2 = d3 = 0.0;
          // right-hand-side weights w f (1, g delta b): delta = 0 for
          // MatrixMRhsAssemble, Compute_SDFEM_delta for MatrixMRhsAssemble_SUPG
          const double be = s1 * b0 + s2 * b1 + s3 * b2;
          rr = ff * (r0 ? w : w * g * dl * be);
        } else {
          const double be = s1 * b0 + s2 * b1 + s3 * b2;  // b_comp (0 for comp 0)
          const double sc = r0 ? w : w * g * dl * be;
          const double t = r0 ? 0.0 : w * g * g * eps;
          const double gb0 = g * b0, gb1 = g * b1, gb2 = g * b2;
          d0 = sc * cc;
          d1 = sc * gb0 + s1 * t;
          d2 = sc * gb1 + s2 * t;
          d3 = sc * gb2 + s3 * t;
          rr = ff * sc;
        }
        const double bv0 = d0 * t0 + d1 * t1 + d2 * t2 + d3 * t3;
        double bv1 = d0 * u0 + d1 * u1 + d2 * u2 + d3 * u3;
        bv1 = rhs_col ? rr : bv1;
        a00[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(av0, bv0, a00[k], 0, 0, 0);
        a01[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(av0, bv1, a01[k], 0, 0, 0);
        a10[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(av1, bv0, a10[k], 0, 0, 0);
        a11[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(av1, bv1, a11[k], 0, 0, 0);
      }
    }
#pragma unroll
    for (int k = 0; k < CPC; k++) {
      if (grp * CPC + k >= ncells) continue;
      double *K = kc + cell[k] * (27 * kKcRow);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int a0 = comp + 4 * r, a1 = 16 + comp + 4 * r, b0 = col, b1 = 16 + col;
        K[a0 * kKcRow + b0] = a00[k][r];
        if (b1 < kKcRow) K[a0 * kKcRow + b1] = a01[k][r];
        if (a1 < 27) {
          K[a1 * kKcRow + b0] = a10[k][r];
          if (b1 < kKcRow) K[a1 * kKcRow + b1] = a11[k][r];
        }
      }
    }
  }
}
// Row gather from the cell buffer: the thread of active node p adds row a of
// each incident cell's local matrix into its CSS row and sums the rhs column.
// The row's stencil slots are accumulated in registers over its incident
// cells (in the order (iz, iy, ix)) and every slot is stored once, after the
// last cell that touches it (a slot (dx, dy, dz) gets its last contribution
// from the cell on the positive side of each offset that is >= 0), so the
// values are written without being read and without a prior clear.  The local node of the row in cell (ix, iy, iz) and
// the slot of every cell entry are compile-time constants: along an axis
// with extent 2 (even lattice coordinate, a vertex plane) the node is local
// coordinate 2 - 2 i of cell i, along an axis with extent 1 local coordinate 1.
template <int EX, int EY, int EZ>
__device__ __forceinline__ void assemble_row_kc(const DevSegment &s, int idx, int M, int n,
                                                const double *__restrict__ kc, double *__restrict__ val,
                                                double *__restrict__ rhs) {
  constexpr int NCX = EX == 1 ? 1 : 2, NCY = EY == 1 ? 1 : 2, NCZ = EZ == 1 ? 1 : 2;
  constexpr int WX = 2 * EX + 1, WY = 2 * EY + 1, WZ = 2 * EZ + 1, NS = WX * WY * WZ;
  const int i = idx % s.cx, t = idx / s.cx;
  const int j = t % s.cy, k = t / s.cy;
  const int px = s.x0 + s.stride * i, py = s.y0 + s.stride * j, pz = s.z0 + s.stride * k;
  double *v = val + vbase(s, idx);
  double frhs = 0;
  double acc[NS];
#pragma unroll
  for (int q = 0; q < NS; q++) acc[q] = 0.0;
  const int cx0 = (px - 1) / 2, cy0 = (py - 1) / 2, cz0 = (pz - 1) / 2;
#pragma unroll
  for (int iz = 0; iz < NCZ; iz++)
#pragma unroll
    for (int iy = 0; iy < NCY; iy++)
#pragma unroll
      for (int ix = 0; ix < NCX; ix++) {
        const int ax = EX == 1 ? 1 : 2 - 2 * ix, ay = EY == 1 ? 1 : 2 - 2 * iy, az = EZ == 1 ? 1 : 2 - 2 * iz;
        const int a = ax + 3 * (ay + 3 * az);
        const int64_t cell = (cx0 + ix) + (int64_t)n * ((cy0 + iy) + (int64_t)n * (cz0 + iz));
        const double2 *Kr = reinterpret_cast<const double2 *>(kc + (cell * 27 + a) * kKcRow);
        double row[kKcRow];
#pragma unroll
        for (int e = 0; e < kKcRow / 2; e++) {
          const double2 u = Kr[e];
          row[2 * e] = u.x;
          row[2 * e + 1] = u.y;
        }
#pragma unroll
        for (int b = 0; b < 27; b++) {
          const int dx = b % 3 - ax, dy = (b / 3) % 3 - ay, dz = b / 9 - az;
          acc[(dx + EX) + WX * ((dy + EY) + WY * (dz + EZ))] += row[b];
        }
        frhs += row[27];
#pragma unroll
        for (int q = 0; q < NS; q++) {
          const int dx = q % WX - EX, dy = (q / WX) % WY - EY, dz = q / (WX * WY) - EZ;
          const int lx = (NCX == 2 && dx >= 0) ? 1 : 0, ly = (NCY == 2 && dy >= 0) ? 1 : 0,
                    lz = (NCZ == 2 && dz >= 0) ? 1 : 0;
          if (lx == ix && ly == iy && lz == iz) v[q * kSlice] = acc[q];
        }
      }
  rhs[px + M * (py + M * pz)] = frhs;
}
__global__ __launch_bounds__(kRowsPerBlock) void k_assemble_kc(const BlockInfo *__restrict__ blocks,
                                                               const DevSegment *__restrict__ segs, int M, int n,
                                                               const double *__restrict__ kc,
                                                               double *__restrict__ val, double *__restrict__ rhs) {
  const BlockInfo bi = blocks[xcd_remap(blockIdx.x, gridDim.x)];
  const DevSegment s = segs[bi.seg];
  const int idx = bi.base + threadIdx.x;
  if (idx >= s.count) return;
#define PMX_KROW(EX, EY, EZ) assemble_row_kc<EX, EY, EZ>(s, idx, M, n, kc, val, rhs)
  switch (s.ex * 100 + s.ey * 10 + s.ez) {
    case 111: PMX_KROW(1, 1, 1); break;
    case 211: PMX_KROW(2, 1, 1); break;
    case 121: PMX_KROW(1, 2, 1); break;
    case 221: PMX_KROW(2, 2, 1); break;
    case 112: PMX_KROW(1, 1, 2); break;
    case 212: PMX_KROW(2, 1, 2); break;
    case 122: PMX_KROW(1, 2, 2); break;
    default: PMX_KROW(2, 2, 2); break;
  }
#undef PMX_KROW
}

void assemble(const DevLayout &D, const double *prm_dev, int example, int form, const double *qref, int nq,
              const double *tab, double *qpbuf, double *kc, double *val, double *val2, double *rhs, hipStream_t st,
              bool qp_given, double *ftab, const uint8_t *corner) {
  const int64_t ncells = (int64_t)D.n * D.n * D.n;
  const int64_t ncq = ncells * nq;
  // with the per-cell frame codes (CELL_MEASURE 4) the coefficients go
  // through the qp records, which carry the cell's first-point size
  const bool inl = !qp_given && coeffs_inline(example) && !corner;
  if (!inl && !qp_given) {
    // (grid y = n^2 lines of cells: up to 255 cells per direction)
    if (example == PMX_EX_CURL_FOURIER && ftab && (int64_t)D.n * D.n <= 65535) {
      const int64_t nt = fourier_tab_len(D.n, nq) / 2;  // complex entries
      hipLaunchKernelGGL(k_fourier_tables, dim3((unsigned)((nt + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                         prm_dev, qref, nq, D.n, ftab);
      const int nl = nq == 8 ? 4 : (nq == 27 ? 9 : 16);  // x-lines of points per cell
      const dim3 g((unsigned)((D.n * nl + kBlock - 1) / kBlock), (unsigned)(D.n * D.n));
      if (nq == 8)
        hipLaunchKernelGGL(k_qp_data_fourier<8>, g, dim3(kBlock), 0, st, prm_dev, qref, D.n, ftab, qpbuf, corner);
      else if (nq == 27)
        hipLaunchKernelGGL(k_qp_data_fourier<27>, g, dim3(kBlock), 0, st, prm_dev, qref, D.n, ftab, qpbuf, corner);
      else if (nq == 64)
        hipLaunchKernelGGL(k_qp_data_fourier<64>, g, dim3(kBlock), 0, st, prm_dev, qref, D.n, ftab, qpbuf, corner);
      else throw std::logic_error("assemble: rule without the curl-Fourier kernel");
    } else {
      hipLaunchKernelGGL(k_qp_data, dim3((unsigned)((ncq + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, prm_dev,
                         qref, nq, D.n, qpbuf, corner);
    }
  }
  // the per-row evaluation with inline coefficients reads the stabilisation
  // parameters from a per-point buffer (the start of qpbuf), filled once
  const double *qsrc = qpbuf;
  if (inl) {
    qsrc = nullptr;
    if (form_has_delta(form) && !(kc && D.order == 2 && !val2 && form != PMX_FORM_RHS_SUPG)) {
      if ((int64_t)D.n * nq > 65535) throw std::invalid_argument("assemble: level too large for k_qp_delta");
      hipLaunchKernelGGL(k_qp_delta, dim3((unsigned)((D.n * D.n + kBlock - 1) / kBlock), (unsigned)(D.n * nq)),
                         dim3(kBlock), 0, st, prm_dev, qref, D.n, qpbuf);
      qsrc = qpbuf;
    }
  }
  if (form == PMX_FORM_RHS || form == PMX_FORM_RHS_SUPG) {
    // right-hand-side forms (no local matrices): the matrices are left as they are
#define PMX_RHS(O, IN)                                                                                     \
  hipLaunchKernelGGL((k_assemble<O, false, IN, true>), dim3(D.nblocks_all), dim3(kRowsPerBlock), 0, st,        \
                     D.blocks_all, D.segs, D.M, D.n, nq, prm_dev, qref, qsrc, tab, nullptr, nullptr, rhs)
    if (D.order == 1) {
      if (inl) PMX_RHS(1, true);
      else PMX_RHS(1, false);
    } else {
      if (inl) PMX_RHS(2, true);
      else PMX_RHS(2, false);
    }
#undef PMX_RHS
    hipLaunchKernelGGL(k_boundary_rhs, dim3((D.nboundary + kBlock - 1) / kBlock), dim3(kBlock), 0, st, D.boundary,
                       D.nboundary, D.M, prm_dev, rhs);
    return;
  }
  if (kc && D.order == 2 && !val2) {
    // local matrices on MFMA, then the row gather from the cell buffer
    const int64_t ngroups = (ncells + kCellsPerPass - 1) / kCellsPerPass;
    const unsigned grid = (unsigned)std::min<int64_t>((ngroups + kCellWaves - 1) / kCellWaves, 4096);
    const size_t lds = (size_t)nq * 128 * sizeof(double);
#define PMX_CELL(IN, MS)                                                                                   \
  hipLaunchKernelGGL((k_cell_mfma<IN, MS, kCellsPerPass>), dim3(grid), dim3(kCellThreads), lds, st, D.n, nq, form,   \
                     prm_dev, qref, qpbuf, tab, kc)
    const bool ms = form == PMX_FORM_MASS || form == PMX_FORM_MASS_SUPG;
    if (inl) {
      if (ms) PMX_CELL(true, true);
      else PMX_CELL(true, false);
    } else {
      if (ms) PMX_CELL(false, true);
      else PMX_CELL(false, false);
    }
#undef PMX_CELL
    hipLaunchKernelGGL(k_assemble_kc, dim3(D.nblocks_all), dim3(kRowsPerBlock), 0, st, D.blocks_all, D.segs, D.M, D.n,
                       kc, val, rhs);
    hipLaunchKernelGGL(k_boundary_rhs, dim3((D.nboundary + kBlock - 1) / kBlock), dim3(kBlock), 0, st, D.boundary,
                       D.nboundary, D.M, prm_dev, rhs);
    return;
  }
  // the per-row evaluation of Q2 rows adds every cell's contribution into
  // the cleared values (Q1 rows store each slot first; the padding rows of
  // the slices are never written and keep the zeros of the allocation)
  if (D.order == 2) {
    check(hipMemsetAsync(val, 0, sizeof(double) * D.nvals, st), "memset val");
    if (val2) check(hipMemsetAsync(val2, 0, sizeof(double) * D.nvals, st), "memset val2");
  }
#define PMX_ASM(O, TW, IN)                                                                                   \
  hipLaunchKernelGGL((k_assemble<O, TW, IN, false>), dim3(D.nblocks_all), dim3(kRowsPerBlock), 0, st,             \
                     D.blocks_all,                                                                                \
                     D.segs, D.M, D.n, nq, prm_dev, qref, qsrc, tab, val, val2, rhs)
#define PMX_ASM2(O, TW) \
  do {                   \
    if (inl) PMX_ASM(O, TW, true); \
    else PMX_ASM(O, TW, false);    \
  } while (0)
  if (D.order == 1) {
    if (val2) PMX_ASM2(1, true);
    else PMX_ASM2(1, false);
  } else {
    if (val2) PMX_ASM2(2, true);
    else PMX_ASM2(2, false);
  }
#undef PMX_ASM2
#undef PMX_ASM
  hipLaunchKernelGGL(k_boundary_rhs, dim3((D.nboundary + kBlock - 1) / kBlock), dim3(kBlock), 0, st, D.boundary,
                     D.nboundary, D.M, prm_dev, rhs);
}

// ============================================================ error norms
// ||u - u_h||_L2 and |u - u_h|_H1 of a lattice FE function against the
// example's exact solution (TFEFunction3D::GetErrors with L2H1Errors,
// FEFunction3D.C:83-300, MainUtilities.C:1649): one thread per cell gathers
// its (order + 1)^3 nodal values, evaluates u_h and grad u_h at the rule's
// points from the reference tables (brick of side h: d/dx = (2/h) d/dxi,
// |detJ| = h^3/8) and adds w |detJ| (u - u_h)^2 and the gradient term;
// fixed-order block partials, then k_reduce_errors.
template <int ORDER>
__global__ __launch_bounds__(kBlock) void k_errors(const double *__restrict__ prm, int n, int M,
                                                   const double *__restrict__ qref, const double *__restrict__ tab,
                                                   int nq, const double *__restrict__ u,
                                                   double *__restrict__ part) {
  constexpr int K = ORDER + 1, NB = K * K * K;
  const int64_t ncells = (int64_t)n * n * n;
  const double h = 1.0 / n, adet = h * h * h / 8.0, g = 2.0 / h;
  double e0 = 0, e1 = 0;
  for (int64_t c = blockIdx.x * (int64_t)kBlock + threadIdx.x; c < ncells; c += (int64_t)gridDim.x * kBlock) {
    const int ci = (int)(c % n), cj = (int)((c / n) % n), ck = (int)(c / ((int64_t)n * n));
    double uc[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int a = b % K, bb = (b / K) % K, cc = b / (K * K);
      uc[b] = u[(ORDER * ci + a) + (int64_t)M * ((ORDER * cj + bb) + (int64_t)M * (ORDER * ck + cc))];
    }
    for (int q = 0; q < nq; q++) {
      const double *r = qref + 4 * q;
      double uh[4] = {0, 0, 0, 0};
#pragma unroll
      for (int b = 0; b < NB; b++) {
        const double *t = tab + ((int64_t)q * NB + b) * 4;
        uh[0] += uc[b] * t[0];
        uh[1] += uc[b] * t[1];
        uh[2] += uc[b] * t[2];
        uh[3] += uc[b] * t[3];
      }
      double ex[4];
      example_exact_grad(prm, (ci + 0.5 * (r[0] + 1.0)) * h, (cj + 0.5 * (r[1] + 1.0)) * h,
                         (ck + 0.5 * (r[2] + 1.0)) * h, ex);
      const double w = r[3] * adet;
      const double d0 = uh[0] - ex[0], d1 = g * uh[1] - ex[1], d2 = g * uh[2] - ex[2], d3 = g * uh[3] - ex[3];
      e0 += w * d0 * d0;
      e1 += w * (d1 * d1 + d2 * d2 + d3 * d3);
    }
  }
  e0 = block_sum<double>(e0);
  e1 = block_sum<double>(e1);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = e0;
    part[kRedBlocks + blockIdx.x] = e1;
  }
}
__global__ __launch_bounds__(kBlock) void k_reduce_errors(const double *__restrict__ part, int np,
                                                          double *__restrict__ out) {
  double s = 0, t = 0;
  for (int i = threadIdx.x; i < np; i += kBlock) {
    s += part[i];
    t += part[kRedBlocks + i];
  }
  s = block_sum<double>(s);
  t = block_sum<double>(t);
  if (threadIdx.x == 0) {
    out[0] = sqrt(s);
    out[1] = sqrt(t);
  }
}
void fe_errors(const DevLayout &D, const double *prm_dev, const double *qref, const double *tab, int nq,
               const double *u, double *part, double *out, hipStream_t st) {
  const int64_t ncells = (int64_t)D.n * D.n * D.n;
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>((ncells + kBlock - 1) / kBlock, kRedBlocks));
  if (D.order == 2)
    hipLaunchKernelGGL(k_errors<2>, dim3(g), dim3(kBlock), 0, st, prm_dev, D.n, D.M, qref, tab, nq, u, part);
  else
    hipLaunchKernelGGL(k_errors<1>, dim3(g), dim3(kBlock), 0, st, prm_dev, D.n, D.M, qref, tab, nq, u, part);
  hipLaunchKernelGGL(k_reduce_errors, dim3(1), dim3(kBlock), 0, st, part, g, out);
}

// ============================================================ time stepping
// nodal interpolation of the example's exact (initial) function on the lattice
// (TFEFunction3D::Interpolate for Lagrange nodes)
__global__ void k_interpolate(const double *__restrict__ prm, int M, double *__restrict__ x) {
  const int64_t N = (int64_t)M * M * M;
  const double h = 1.0 / (M - 1);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(i % M), b = (int)((i / M) % M), c = (int)(i / ((int64_t)M * M));
    x[i] = example_exact(prm, a * h, b * h, c * h);
  }
}
// right-hand side of TSystemTCD3D::AssembleSystMat (SystemTCD3D.C:255-283):
// active rows B = t3 oldrhs + t4 rhs + (M (+ K) - t2 A) oldsol, Dirichlet rows
// B = rhs (k_old = K oldsol of the SUPG form, SystemTCD2D.C:294-297, or nullptr)
__global__ void k_tcd_rhs(int M, const double *__restrict__ oldrhs, const double *__restrict__ rhs,
                          const double *__restrict__ m_old, const double *__restrict__ a_old,
                          const double *__restrict__ k_old, double t3, double t4, double t2,
                          double *__restrict__ B) {
  const int64_t N = (int64_t)M * M * M;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(i % M), b = (int)((i / M) % M), c = (int)(i / ((int64_t)M * M));
    const bool interior = a > 0 && b > 0 && c > 0 && a < M - 1 && b < M - 1 && c < M - 1;
    const double mk = k_old ? m_old[i] + k_old[i] : m_old[i];
    B[i] = interior ? (t3 * oldrhs[i] + t4 * rhs[i]) + (mk - t2 * a_old[i]) : rhs[i];
  }
}
void interpolate(const DevLayout &D, const double *prm_dev, double *x, hipStream_t st) {
  hipLaunchKernelGGL(k_interpolate, dim3(grid_for(D.N)), dim3(kBlock), 0, st, prm_dev, D.M, x);
}
// val = mass (+ k) + c a (MatAdd of Blas2.C:128-150 on the stored stencil values)
__global__ void k_system_matrix(double *__restrict__ val, float *__restrict__ val32, const double *__restrict__ mass,
                                const double *__restrict__ k, const double *__restrict__ a, double c, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = (k ? mass[i] + k[i] : mass[i]) + c * a[i];
    val[i] = v;
    if (val32) val32[i] = (float)v;
  }
}
void system_matrix(double *val, float *val32, const double *mass, const double *k, const double *a, double c,
                   int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_system_matrix, dim3(grid_for(n)), dim3(kBlock), 0, st, val, val32, mass, k, a, c, n);
}
void tcd_rhs(const DevLayout &D, const double *oldrhs, const double *rhs, const double *m_old, const double *a_old,
             const double *k_old, double t3, double t4, double t2, double *B, hipStream_t st) {
  hipLaunchKernelGGL(k_tcd_rhs, dim3(grid_for(D.N)), dim3(kBlock), 0, st, D.M, oldrhs, rhs, m_old, a_old, k_old,
                     t3, t4, t2, B);
}

void givens(double *H, double *cs, double *sn, double *s, int i, int ld, double *res, hipStream_t st) {
  hipLaunchKernelGGL(k_givens, dim3(1), dim3(64), 0, st, H, cs, sn, s, i, ld, res);
}
void backsolve(const double *H, const double *s, double *y, int k, int ld, hipStream_t st) {
  hipLaunchKernelGGL(k_backsolve, dim3(1), dim3(64), 0, st, H, s, y, k, ld);
}
template <typename T, typename X>
void multi_axpy(X *x, const T *z, const double *y, int k1, int64_t n, int64_t stride, hipStream_t st) {
  hipLaunchKernelGGL((k_multi_axpy<T, X>), dim3(grid_for(n)), dim3(kBlock), 0, st, x, z, y, k1, n, stride);
}
void set_scalar(double *dst, const double *src, double scale, hipStream_t st) {
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(64), 0, st, dst, src, scale);
}

// explicit instantiations
#define PMX_INST(T)                                                                                         \
  template void fill<T>(T *, T, int64_t, hipStream_t);                                                      \
  template void copy<T>(T *, const T *, int64_t, hipStream_t);                                              \
  template void axpy<T>(T *, const T *, const double *, double, double, int64_t, hipStream_t);              \
  template void scale_inv<T>(T *, const T *, const double *, int64_t, hipStream_t);                         \
  template void dot<T>(const T *, const T *, int64_t, double *, double *, int, hipStream_t);                \
  template void mgs_arnoldi<T>(T *, const T *, int64_t, int, double *, T *, double *, hipStream_t);         \
  template void axpy_diff<T>(T *, const T *, const T *, int64_t, hipStream_t);                              \
  template void step_length_update<T>(T *, const T *, const double *, double, int64_t, hipStream_t);        \
  template void css_apply<T>(const DevLayout &, int, const T *, const T *, const T *, T *, double, int,     \
                             hipStream_t);                                                                  \
  template void boundary<T>(const DevLayout &, int, T *, const T *, const T *, hipStream_t);                \
  template void sweep_wg<T>(const DevLayout &, const T *, T *, const T *, double, int, hipStream_t);          \
  template void coarse_wg<T>(const DevLayout &, const T *, T *, const T *, double, int, int, double, T *, T *, \
                             hipStream_t);                                                                  \
  template void prolongate<T>(const DevLayout &, const DevLayout &, const T *, T *, hipStream_t);           \
  template void prolongate_add<T>(const DevLayout &, const DevLayout &, const T *, T *, double, hipStream_t); \
  template void restrict_<T>(const DevLayout &, const DevLayout &, T *, const T *, hipStream_t, bool);      \
  template void gemv_coarse<T>(int, int, const double *, const int *, const T *, T *, hipStream_t);         \
  template void csr_spmv<T>(int, const int *, const int *, const T *, const T *, T *, hipStream_t);         \
  template void csr_defect<T>(int, const int *, const int *, const T *, const T *, const T *, T *, hipStream_t); \
  template void css_to_csr<T>(const DevLayout &, const int *, const int *, const int *, int64_t, const T *, \
                              double *, hipStream_t);
PMX_INST(double)
PMX_INST(float)
template void css_to_csr<_Float16>(const DevLayout &, const int *, const int *, const int *, int64_t,
                                   const _Float16 *, double *, hipStream_t);
template void fill<uint16_t>(uint16_t *, uint16_t, int64_t, hipStream_t);
template void fill<int>(int *, int, int64_t, hipStream_t);
template void convert<float, double>(float *, const double *, int64_t, hipStream_t);
template void convert<double, float>(double *, const float *, int64_t, hipStream_t);
template void add_convert<double, float>(double *, const float *, int64_t, hipStream_t);
template void add_convert<double, double>(double *, const double *, int64_t, hipStream_t);
template void gather<double, double>(double *, const double *, const int *, int64_t, hipStream_t);
template void gather<float, float>(float *, const float *, const int *, int64_t, hipStream_t);
template void multi_axpy<double, double>(double *, const double *, const double *, int, int64_t, int64_t, hipStream_t);
template void multi_axpy<float, float>(float *, const float *, const double *, int, int64_t, int64_t, hipStream_t);

}  // namespace pmx
