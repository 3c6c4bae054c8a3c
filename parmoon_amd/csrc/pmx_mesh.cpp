// pmx_mesh.cpp -- host set-up: cell order, DOF numbering, CSR pattern and the
// colour-sliced stencil layout of a uniformly refined unit cube.
//
// The reference builds these through its refinement tree and mappers
// (src/Refinement/RefHexaRegDesc.C, src/Geometry/It_Finest.C,
// src/FE/FESpace3D.C:485-2080, src/FE/SquareStructure3D.C:59-200).  Here the
// same integer results come from closed forms on the lattice:
//  * a cell of the tree is an origin plus a signed permutation of the axes;
//    each child of a regular refinement starts at a parent vertex and turns
//    its frame by a fixed signed permutation (the child vertex lists of
//    RefHexaRegDesc.C:55-60 written as frames);
//  * ConstructSpace keeps, for every node, the mark at the smallest
//    (cell, local index) position that holds it (FE3DMapper.C:173-296) and
//    numbers marks in position order, inner nodes first and Dirichlet nodes
//    last (FESpace3D.C:1739-1830) -- i.e. first-encounter order per class.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "pmx_internal.hpp"

namespace pmx {

// Child frames of the regular hexahedron refinement: child j starts at parent
// vertex j; its local axis k is parent axis |code|-1 with the sign of code.
static const int kChildAxes[8][3] = {{1, 2, 3},  {2, -1, 3},  {-1, -2, 3}, {-2, 1, 3},
                                     {2, 1, -3}, {-1, 2, -3}, {-2, -1, -3}, {1, -2, -3}};
// reference position (0/1 per axis) of hexahedron vertex j
static const int kVertexPos[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                     {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

struct Frame {
  int o[3];     // origin in units of the finest cell
  int e[3][3];  // local axes as unit vectors
  int size;     // edge length in units of the finest cell
};

int node_colour(const Layout &L, int x, int y, int z) {
  if (L.order == 2) return (x % 3) + 3 * (y % 3) + 9 * (z % 3);
  return (x % 2) + 2 * (y % 2) + 4 * (z % 2);
}

void build_layout(Layout &L, int nref, int order, int block) {
  if (order != 1 && order != 2) throw std::invalid_argument("order must be 1 or 2");
  if (nref < 1 || nref > 10) throw std::invalid_argument("nref out of range");
  L.nref = nref;
  L.order = order;
  L.n = 1 << nref;
  L.M = order * L.n + 1;
  L.S = order == 2 ? 6 : 2;
  L.N = (int64_t)L.M * L.M * L.M;
  L.nactive = (int64_t)(L.M - 2) * (L.M - 2) * (L.M - 2);
  L.ncolours = order == 2 ? 27 : 8;
  L.segs.clear();
  L.nvals = 0;
  const int S = L.S, M = L.M;
  auto first = [&](int r) { return r >= 1 ? r : S; };
  auto cnt = [&](int x0) { return x0 <= M - 2 ? (M - 2 - x0) / S + 1 : 0; };
  // classes ordered by colour, then by residue
  for (int col = 0; col < L.ncolours; col++)
    for (int rz = 0; rz < S; rz++)
      for (int ry = 0; ry < S; ry++)
        for (int rx = 0; rx < S; rx++) {
          if (node_colour(L, rx, ry, rz) != col) continue;
          Segment s{};
          s.rx = rx; s.ry = ry; s.rz = rz;
          s.x0 = first(rx); s.y0 = first(ry); s.z0 = first(rz);
          s.cx = cnt(s.x0); s.cy = cnt(s.y0); s.cz = cnt(s.z0);
          s.count = (int64_t)s.cx * s.cy * s.cz;
          if (s.count == 0) continue;
          if (order == 2) {
            s.ex = (s.x0 % 2 == 0) ? 2 : 1;
            s.ey = (s.y0 % 2 == 0) ? 2 : 1;
            s.ez = (s.z0 % 2 == 0) ? 2 : 1;
          } else {
            s.ex = s.ey = s.ez = 1;
          }
          s.colour = col;
          s.ns = (2 * s.ex + 1) * (2 * s.ey + 1) * (2 * s.ez + 1);
          s.dslot = s.ex + (2 * s.ex + 1) * (s.ey + (2 * s.ey + 1) * s.ez);
          s.voff = L.nvals;
          L.nvals += ((s.count + kSlice - 1) / kSlice) * kSlice * s.ns;
          L.segs.push_back(s);
        }
  L.blocks_all.clear();
  L.blocks_colour.assign(L.ncolours, {});
  for (int si = 0; si < (int)L.segs.size(); si++) {
    const Segment &s = L.segs[si];
    for (int64_t b = 0; b < s.count; b += block) {
      L.blocks_all.push_back({si, (int)b});
      L.blocks_colour[s.colour].push_back({si, (int)b});
    }
  }
  L.boundary.clear();
  for (int z = 0; z < M; z++)
    for (int y = 0; y < M; y++)
      for (int x = 0; x < M; x++)
        if (x == 0 || y == 0 || z == 0 || x == M - 1 || y == M - 1 || z == M - 1)
          L.boundary.push_back(x + M * (y + M * z));
}

// Calls leaf(frame) for the cells of the unit cube refined until the cells
// have size 1, in the reference's cell order: depth first through the
// refinement tree, children 0..7 (It_Finest over the regular refinement).
template <class F>
static void walk_leaves(int n, F &&leaf) {
  std::vector<Frame> stack;
  Frame root{};
  root.e[0][0] = root.e[1][1] = root.e[2][2] = 1;
  root.size = n;
  stack.push_back(root);
  while (!stack.empty()) {
    Frame f = stack.back();
    stack.pop_back();
    if (f.size == 1) {
      leaf(f);
      continue;
    }
    // push children in reverse so that child 0 is visited first
    const int h = f.size / 2;
    for (int j = 7; j >= 0; j--) {
      Frame c{};
      for (int d = 0; d < 3; d++)
        c.o[d] = f.o[d] + f.size * (kVertexPos[j][0] * f.e[0][d] + kVertexPos[j][1] * f.e[1][d] +
                                    kVertexPos[j][2] * f.e[2][d]);
      for (int a = 0; a < 3; a++) {
        const int code = kChildAxes[j][a];
        const int src = (code > 0 ? code : -code) - 1, sg = code > 0 ? 1 : -1;
        for (int d = 0; d < 3; d++) c.e[a][d] = sg * f.e[src][d];
      }
      c.size = h;
      stack.push_back(c);
    }
  }
}

void build_numbering(Numbering &Nb, const Layout &L) {
  const int p = L.order, M = L.M, k = p + 1;
  Nb.nb = k * k * k;
  Nb.ncells = (int64_t)L.n * L.n * L.n;
  Nb.N = L.N;
  Nb.active = L.nactive;
  Nb.begin_index.resize(Nb.ncells + 1);
  for (int64_t i = 0; i <= Nb.ncells; i++) Nb.begin_index[i] = (int)(i * Nb.nb);
  Nb.global_numbers.assign(Nb.ncells * Nb.nb, -1);
  Nb.lat2dof.assign(L.N, -1);
  int64_t inner = 0, dir = 0;
  std::vector<uint8_t> isdir(L.N, 0);
  int64_t cell = 0;
  int loc[27];
  Nb.corner.assign((size_t)L.n * L.n * L.n, 0);
  walk_leaves(L.n, [&](const Frame &f) {
    {
      int lo[3];
      uint8_t code = 0;
      for (int d = 0; d < 3; d++) {
        const int s = f.e[0][d] + f.e[1][d] + f.e[2][d];
        lo[d] = f.o[d] + std::min(f.e[0][d], 0) + std::min(f.e[1][d], 0) + std::min(f.e[2][d], 0);
        if (s > 0) code |= (uint8_t)(1 << d);
      }
      Nb.corner[lo[0] + (size_t)L.n * (lo[1] + (size_t)L.n * lo[2])] = code;
    }
    // local DOF (a,b,c), x fastest, sits at p*origin + a e1 + b e2 + c e3
    int j = 0;
    for (int c = 0; c < k; c++)
      for (int b = 0; b < k; b++)
        for (int a = 0; a < k; a++, j++) {
          int X[3];
          for (int d = 0; d < 3; d++) X[d] = p * f.o[d] + a * f.e[0][d] + b * f.e[1][d] + c * f.e[2][d];
          loc[j] = X[0] + M * (X[1] + M * X[2]);
          const int64_t l = loc[j];
          if (Nb.lat2dof[l] < 0) {
            const bool bd = X[0] == 0 || X[1] == 0 || X[2] == 0 || X[0] == M - 1 || X[1] == M - 1 ||
                            X[2] == M - 1;
            if (bd) { Nb.lat2dof[l] = (int)dir++; isdir[l] = 1; }
            else Nb.lat2dof[l] = (int)inner++;
          }
        }
    for (int q = 0; q < Nb.nb; q++) Nb.global_numbers[cell * Nb.nb + q] = loc[q];  // lattice for now
    cell++;
  });
  if (inner != L.nactive || inner + dir != L.N) throw std::runtime_error("numbering count mismatch");
  for (int64_t l = 0; l < L.N; l++)
    if (isdir[l]) Nb.lat2dof[l] += (int)inner;
  Nb.dof2lat.assign(L.N, -1);
  for (int64_t l = 0; l < L.N; l++) Nb.dof2lat[Nb.lat2dof[l]] = (int)l;
  for (auto &g : Nb.global_numbers) g = Nb.lat2dof[g];
  Nb.have_csr = false;
}

void build_csr(Numbering &Nb, const Layout &L) {
  if (Nb.have_csr) return;
  const int M = L.M;
  const int64_t N = L.N;
  Nb.rowptr.assign(N + 1, 0);
  // row length from the node's stencil box (interior rows) or 1 (Dirichlet)
  auto ext = [&](int x) { return (L.order == 2 && x % 2 == 0) ? 2 : 1; };
  for (int64_t r = 0; r < N; r++) {
    int64_t len = 1;
    if (r < Nb.active) {
      const int l = Nb.dof2lat[r];
      const int x = l % M, y = (l / M) % M, z = l / (M * M);
      len = (int64_t)(2 * ext(x) + 1) * (2 * ext(y) + 1) * (2 * ext(z) + 1);
    }
    Nb.rowptr[r + 1] = (int)len;
  }
  for (int64_t r = 0; r < N; r++) Nb.rowptr[r + 1] += Nb.rowptr[r];
  Nb.kcol.assign((size_t)Nb.rowptr[N], 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t r = 0; r < N; r++) {
    int *c = &Nb.kcol[Nb.rowptr[r]];
    if (r >= Nb.active) { c[0] = (int)r; continue; }
    const int l = Nb.dof2lat[r];
    const int x = l % M, y = (l / M) % M, z = l / (M * M);
    const int ex = ext(x), ey = ext(y), ez = ext(z);
    int n = 0;
    for (int dz = -ez; dz <= ez; dz++)
      for (int dy = -ey; dy <= ey; dy++)
        for (int dx = -ex; dx <= ex; dx++) c[n++] = Nb.lat2dof[(x + dx) + M * ((y + dy) + M * (z + dz))];
    std::sort(c, c + n);
  }
  Nb.have_csr = true;
}

void write_vtk(const Layout &L, const Numbering &Nb, int nvar, const char *const *names,
               const double *const *values, const char *path) {
  const int n = L.n, V = n + 1, p = L.order, M = L.M;
  const int64_t nv = (int64_t)V * V * V, nc = (int64_t)n * n * n;
  FILE *fp = std::fopen(path, "w");
  if (!fp) throw std::runtime_error(std::string("cannot open ") + path + " for writing");
  std::fprintf(fp, "# vtk DataFile Version 4.2\nfile created by ParMooN\nASCII\nDATASET UNSTRUCTURED_GRID\n\n");
  std::fprintf(fp, "POINTS %lld double\n", (long long)nv);
  const double h = 1.0 / n;
  for (int z = 0; z < V; z++)
    for (int y = 0; y < V; y++)
      for (int x = 0; x < V; x++) std::fprintf(fp, "%.6e %.6e %.6e\n", x * h, y * h, z * h);
  std::fprintf(fp, "\nCELLS %lld %lld\n", (long long)nc, (long long)(nc * 9));
  walk_leaves(n, [&](const Frame &f) {
    // vertex j of the cell at the reference position kVertexPos[j] of its frame
    std::fprintf(fp, "8 ");
    for (int j = 0; j < 8; j++) {
      int X[3];
      for (int d = 0; d < 3; d++)
        X[d] = f.o[d] + kVertexPos[j][0] * f.e[0][d] + kVertexPos[j][1] * f.e[1][d] + kVertexPos[j][2] * f.e[2][d];
      std::fprintf(fp, "%d ", X[0] + V * (X[1] + V * X[2]));
    }
    std::fprintf(fp, "\n");
  });
  std::fprintf(fp, "\nCELL_TYPES %lld\n", (long long)nc);
  for (int64_t c = 0; c < nc; c++) std::fputs("12 ", fp);
  std::fprintf(fp, "\n\nPOINT_DATA %lld\n", (long long)nv);
  for (int k = 0; k < nvar; k++) {
    std::fprintf(fp, "SCALARS %s double\nLOOKUP_TABLE default\n", names[k]);
    const double *u = values[k];
    // a nodal Q1/Q2 function's value at a vertex is its coefficient there
    for (int z = 0; z < V; z++)
      for (int y = 0; y < V; y++)
        for (int x = 0; x < V; x++) std::fprintf(fp, "%.6e\n", u[Nb.lat2dof[p * x + M * (p * y + M * p * z)]]);
    std::fprintf(fp, "\n\n");
  }
  std::fprintf(fp, "\nCELL_DATA %lld\nSCALARS RegionID int  1\nLOOKUP_TABLE default  \n", (long long)nc);
  for (int64_t c = 0; c < nc; c++) std::fputs("0\n", fp);
  const bool bad = std::ferror(fp) != 0;
  if (std::fclose(fp) != 0 || bad) throw std::runtime_error(std::string("error writing ") + path);
}

}  // namespace pmx
