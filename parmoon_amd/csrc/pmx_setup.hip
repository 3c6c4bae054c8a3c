// pmx_setup.hip -- the DOF numbering of a level built on the device.
//
// The same integer results as build_numbering (pmx_mesh.cpp; the reference's
// ConstructSpace, src/FE/FESpace3D.C:1739-1830 over the cell order of
// It_Finest and the regular refinement, src/Refinement/RefHexaRegDesc.C:
// 55-60), computed in parallel:
//  1. one thread per cell (tree order): the cell's frame follows from the
//     base-8 digits of its index (the child taken at each depth), and gives
//     the lattice node of each local DOF and the cell's frame code;
//  2. one thread per position (cell, local DOF): the smallest position that
//     holds each node (atomicMin): the node's first encounter;
//  3. the first encounters of inner and of Dirichlet nodes are counted in
//     position order (two exclusive scans), which numbers inner nodes first
//     and Dirichlet nodes after them, each class in first-encounter order;
//  4. GlobalNumbers and the inverse permutation by gathers.
// and the CSR structure of a level on the device (device_csr).
#include <hip/hip_runtime.h>

#include <climits>
#include <hipcub/hipcub.hpp>
#include <stdexcept>

#include "pmx_internal.hpp"
#include "pmx_ops.hpp"

namespace pmx {

// child j of the regular refinement: local axis k is parent axis |code|-1
// with the sign of code (the table of pmx_mesh.cpp)
__constant__ int kChildAxesDev[8][3] = {{1, 2, 3},  {2, -1, 3},  {-1, -2, 3}, {-2, 1, 3},
                                        {2, 1, -3}, {-1, 2, -3}, {-2, -1, -3}, {1, -2, -3}};
// reference position (0/1 per axis) of hexahedron vertex j
__constant__ int kVertexPosDev[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                        {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// Stage 1: frame of cell i (tree order), its local DOFs' lattice nodes and
// its frame code (signs of e1 + e2 + e3, see Numbering::corner)
__global__ __launch_bounds__(kBlock) void k_cell_nodes(int nref, int order, int M, int64_t ncells,
                                                       int *__restrict__ nodes, uint8_t *__restrict__ corner) {
  const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (i >= ncells) return;
  int o[3] = {0, 0, 0};
  int e[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  int size = 1 << nref;
  for (int l = 0; l < nref; l++) {
    const int j = (int)((i >> (3 * (nref - 1 - l))) & 7);
    int c[3], ce[3][3];
    for (int d = 0; d < 3; d++)
      c[d] = o[d] + size * (kVertexPosDev[j][0] * e[0][d] + kVertexPosDev[j][1] * e[1][d] +
                            kVertexPosDev[j][2] * e[2][d]);
    for (int a = 0; a < 3; a++) {
      const int code = kChildAxesDev[j][a];
      const int src = (code > 0 ? code : -code) - 1, sg = code > 0 ? 1 : -1;
      for (int d = 0; d < 3; d++) ce[a][d] = sg * e[src][d];
    }
    for (int d = 0; d < 3; d++) {
      o[d] = c[d];
      for (int a = 0; a < 3; a++) e[a][d] = ce[a][d];
    }
    size >>= 1;
  }
  const int k = order + 1, nb = k * k * k, n = 1 << nref;
  int q = 0;
  for (int cc = 0; cc < k; cc++)
    for (int b = 0; b < k; b++)
      for (int a = 0; a < k; a++, q++) {
        int X[3];
        for (int d = 0; d < 3; d++) X[d] = order * o[d] + a * e[0][d] + b * e[1][d] + cc * e[2][d];
        nodes[i * nb + q] = X[0] + M * (X[1] + M * X[2]);
      }
  int lo[3];
  uint8_t code = 0;
  for (int d = 0; d < 3; d++) {
    lo[d] = o[d] + min(e[0][d], 0) + min(e[1][d], 0) + min(e[2][d], 0);
    if (e[0][d] + e[1][d] + e[2][d] > 0) code |= (uint8_t)(1 << d);
  }
  corner[lo[0] + (int64_t)n * (lo[1] + (int64_t)n * lo[2])] = code;
}
// Stage 2: first encounter of every node
__global__ __launch_bounds__(kBlock) void k_first_pos(const int *__restrict__ nodes, int npos, int *__restrict__ first) {
  const int pos = blockIdx.x * kBlock + threadIdx.x;
  if (pos < npos) atomicMin(&first[nodes[pos]], pos);
}
__device__ __forceinline__ bool lattice_surface(int l, int M) {
  const int x = l % M, y = (l / M) % M, z = l / (M * M);
  return x == 0 || y == 0 || z == 0 || x == M - 1 || y == M - 1 || z == M - 1;
}
// Stage 3a: flags of the first encounters, inner and Dirichlet
__global__ __launch_bounds__(kBlock) void k_first_flags(const int *__restrict__ nodes, const int *__restrict__ first,
                                                        int npos, int M, int *__restrict__ fi, int *__restrict__ fd) {
  const int pos = blockIdx.x * kBlock + threadIdx.x;
  if (pos >= npos) return;
  const int l = nodes[pos];
  const bool is_first = first[l] == pos;
  const bool bd = lattice_surface(l, M);
  fi[pos] = is_first && !bd ? 1 : 0;
  fd[pos] = is_first && bd ? 1 : 0;
}
// Stage 3b: the numbers of the first encounters
__global__ __launch_bounds__(kBlock) void k_number(const int *__restrict__ nodes, const int *__restrict__ first,
                                                   const int *__restrict__ si, const int *__restrict__ sd, int npos,
                                                   int M, int nactive, int *__restrict__ lat2dof) {
  const int pos = blockIdx.x * kBlock + threadIdx.x;
  if (pos >= npos) return;
  const int l = nodes[pos];
  if (first[l] != pos) return;
  lat2dof[l] = lattice_surface(l, M) ? nactive + sd[pos] : si[pos];
}
// Stage 4: GlobalNumbers (cell-major) and dof2lat
__global__ __launch_bounds__(kBlock) void k_global_numbers(const int *__restrict__ nodes,
                                                           const int *__restrict__ lat2dof, int npos,
                                                           int *__restrict__ gn) {
  const int pos = blockIdx.x * kBlock + threadIdx.x;
  if (pos < npos) gn[pos] = lat2dof[nodes[pos]];
}
__global__ __launch_bounds__(kBlock) void k_inverse_perm(const int *__restrict__ lat2dof, int n,
                                                         int *__restrict__ dof2lat) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  if (l < n) dof2lat[lat2dof[l]] = l;
}

// ---- the CSR structure (TSquareStructure3D + Sort) in the reference numbering
// (build_csr's result): row r of an interior node holds the DOFs of the
// node's stencil box, ascending; a Dirichlet row only its diagonal.
__device__ __forceinline__ int stencil_ext(int order, int c) { return (order == 2 && (c & 1) == 0) ? 2 : 1; }
// row lengths into len[r + 1] (len[0] = 0): the inclusive scan gives RowPtr
__global__ __launch_bounds__(kBlock) void k_csr_len(int M, int order, const int *__restrict__ dof2lat, int64_t nact,
                                                    int64_t N, int *__restrict__ len) {
  const int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (r == 0) len[0] = 0;
  if (r >= N) return;
  int n = 1;
  if (r < nact) {
    const int l = dof2lat[r];
    const int x = l % M, y = (l / M) % M, z = l / (M * M);
    n = (2 * stencil_ext(order, x) + 1) * (2 * stencil_ext(order, y) + 1) * (2 * stencil_ext(order, z) + 1);
  }
  len[r + 1] = n;
}
// one wavefront per row: the row's column numbers go to LDS, then every
// entry's position is its rank among them (the numbers of a row are
// distinct), which sorts the row without a sorting pass
constexpr int kCsrRowsPerBlock = kBlock / 64;
__global__ __launch_bounds__(kBlock) void k_csr_cols(int M, int order, const int *__restrict__ dof2lat,
                                                     const int *__restrict__ lat2dof,
                                                     const int *__restrict__ rowptr, int64_t nact, int64_t N,
                                                     int *__restrict__ kcol) {
  __shared__ int sh[kCsrRowsPerBlock][128];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t r = blockIdx.x * (int64_t)kCsrRowsPerBlock + w;
  if (r >= N) return;  // the whole wavefront
  const int b = rowptr[r];
  if (r >= nact) {
    if (lane == 0) kcol[b] = (int)r;
    return;
  }
  const int l = dof2lat[r];
  const int x = l % M, y = (l / M) % M, z = l / (M * M);
  const int ex = stencil_ext(order, x), ey = stencil_ext(order, y), ez = stencil_ext(order, z);
  const int wx = 2 * ex + 1, wy = 2 * ey + 1, n = wx * wy * (2 * ez + 1);
  int *row = sh[w];
  for (int t = lane; t < n; t += 64) {
    const int dx = t % wx - ex, dy = (t / wx) % wy - ey, dz = t / (wx * wy) - ez;
    row[t] = lat2dof[(x + dx) + M * ((y + dy) + M * (z + dz))];
  }
  // the row's LDS entries are written and read by this wavefront only
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int t = lane; t < n; t += 64) {
    const int v = row[t];
    int rank = 0;
    for (int u = 0; u < n; u++) rank += row[u] < v ? 1 : 0;
    kcol[b + rank] = v;
  }
}
int64_t csr_nnz(const Layout &L) {
  int64_t nnz = L.N - L.nactive;
  for (const Segment &g : L.segs) nnz += g.count * g.ns;
  return nnz;
}
void device_csr(const Layout &L, const int *lat2dof, const int *dof2lat, int *rowptr, int *kcol, hipStream_t st) {
  const int64_t N = L.N;
  if (csr_nnz(L) > INT_MAX) throw std::invalid_argument("device_csr: more than 2^31 - 1 entries");
  // a row's columns are staged in 128 LDS entries: at most 5^3 (Q2)
  if (L.order != 1 && L.order != 2) throw std::invalid_argument("device_csr: order must be 1 or 2");
  const unsigned g = (unsigned)((N + kBlock - 1) / kBlock);
  int *len = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
  auto release = [&]() {
    if (len) (void)hipFree(len);
    if (tmp) (void)hipFree(tmp);
  };
  try {
    check(hipMalloc(&len, sizeof(int) * (N + 1)), "csr: lengths");
    hipLaunchKernelGGL(k_csr_len, dim3(g), dim3(kBlock), 0, st, L.M, L.order, dof2lat, L.nactive, N, len);
    check(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, len, rowptr, (int)(N + 1), st), "csr: scan size");
    check(hipMalloc(&tmp, tmp_bytes), "csr: scan storage");
    check(hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, len, rowptr, (int)(N + 1), st), "csr: scan");
    const unsigned gr = (unsigned)((N + kCsrRowsPerBlock - 1) / kCsrRowsPerBlock);
    hipLaunchKernelGGL(k_csr_cols, dim3(gr), dim3(kBlock), 0, st, L.M, L.order, dof2lat, lat2dof, rowptr,
                       L.nactive, N, kcol);
    check(hipStreamSynchronize(st), "csr");
  } catch (...) {
    release();
    throw;
  }
  release();
}

void device_numbering(const Layout &L, int *lat2dof, int *dof2lat, int *gn, uint8_t *corner, hipStream_t st) {
  const int64_t ncells = (int64_t)L.n * L.n * L.n;
  const int nb = (L.order + 1) * (L.order + 1) * (L.order + 1);
  const int64_t npos64 = ncells * nb;
  if (npos64 > INT_MAX / 2 || L.N > INT_MAX / 2) throw std::invalid_argument("device_numbering: level too large");
  const int npos = (int)npos64, N = (int)L.N;
  int *nodes = nullptr, *first = nullptr, *fi = nullptr, *fd = nullptr, *si = nullptr, *sd = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
  auto release = [&]() {
    for (void *p : {(void *)nodes, (void *)first, (void *)fi, (void *)fd, (void *)si, (void *)sd, tmp})
      if (p) (void)hipFree(p);
  };
  try {
    check(hipMalloc(&nodes, sizeof(int) * npos), "numbering: nodes");
    check(hipMalloc(&first, sizeof(int) * N), "numbering: first");
    check(hipMalloc(&fi, sizeof(int) * npos), "numbering: flags");
    check(hipMalloc(&fd, sizeof(int) * npos), "numbering: flags");
    check(hipMalloc(&si, sizeof(int) * npos), "numbering: scan");
    check(hipMalloc(&sd, sizeof(int) * npos), "numbering: scan");
    const unsigned gc = (unsigned)((ncells + kBlock - 1) / kBlock), gp = (unsigned)((npos + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_cell_nodes, dim3(gc), dim3(kBlock), 0, st, L.nref, L.order, L.M, ncells, nodes, corner);
    fill<int>(first, INT_MAX, N, st);
    hipLaunchKernelGGL(k_first_pos, dim3(gp), dim3(kBlock), 0, st, nodes, npos, first);
    hipLaunchKernelGGL(k_first_flags, dim3(gp), dim3(kBlock), 0, st, nodes, first, npos, L.M, fi, fd);
    check(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, fi, si, npos, st), "numbering: scan size");
    check(hipMalloc(&tmp, tmp_bytes), "numbering: scan storage");
    check(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, fi, si, npos, st), "numbering: scan");
    check(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, fd, sd, npos, st), "numbering: scan");
    // the counts of first encounters must be the level's inner and Dirichlet nodes
    int last[4];
    check(hipMemcpyAsync(&last[0], si + npos - 1, sizeof(int), hipMemcpyDeviceToHost, st), "numbering: count");
    check(hipMemcpyAsync(&last[1], fi + npos - 1, sizeof(int), hipMemcpyDeviceToHost, st), "numbering: count");
    check(hipMemcpyAsync(&last[2], sd + npos - 1, sizeof(int), hipMemcpyDeviceToHost, st), "numbering: count");
    check(hipMemcpyAsync(&last[3], fd + npos - 1, sizeof(int), hipMemcpyDeviceToHost, st), "numbering: count");
    check(hipStreamSynchronize(st), "numbering: count");
    if (last[0] + last[1] != L.nactive || last[2] + last[3] != L.N - L.nactive)
      throw std::runtime_error("device_numbering: node count mismatch");
    hipLaunchKernelGGL(k_number, dim3(gp), dim3(kBlock), 0, st, nodes, first, si, sd, npos, L.M, (int)L.nactive,
                       lat2dof);
    if (gn) hipLaunchKernelGGL(k_global_numbers, dim3(gp), dim3(kBlock), 0, st, nodes, lat2dof, npos, gn);
    hipLaunchKernelGGL(k_inverse_perm, dim3((unsigned)((N + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, lat2dof, N,
                       dof2lat);
    check(hipStreamSynchronize(st), "numbering");
  } catch (...) {
    release();
    throw;
  }
  release();
}

}  // namespace pmx
