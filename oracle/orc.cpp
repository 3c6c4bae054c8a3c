// orc.cpp -- CPU oracle for the CD3D/TCD3D hot path.
//
// TEST INFRASTRUCTURE ONLY: a single-threaded FP64 restatement of ParMooN's
// algorithms, used as the parity checker of the HIP implementation and as the
// timed reference CPU path.  The product (parmoon_amd/) never calls it.
//
// Every function names the reference file:line whose behaviour it restates.
// Parity pins for each part are listed in tests/ (golden log values from
// ReadIn/tcd3d_GMG.out, closed forms, brute force).
#include "orc.h"
#include "../include/pmx_prm.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <unordered_map>
#include <vector>

namespace orc {

static const double kPi = 3.14159265358979323846;

// ===========================================================================
// Geometry: macro cell and regular refinement tree
// ===========================================================================

// data/Wuerfel.GEO lists one hexahedron whose 8 vertices are boundary
// parameters (comp, T, S); ReadGeo maps them through the BdPlane components of
// data/Wuerfel.PRM (src/Geometry/ReadGeo.C:1755-1768).  A plane is
// X = P + T*A + S*B with B = A x N (src/Geometry/BdPlane.C:47-74).
static void macro_vertices(double v[8][3]) {
  // plane components 0 and 3 of Wuerfel.PRM: {P, A, N}
  const double plane[6][3][3] = {
      {{0, 0, 0}, {1, 0, 0}, {0, 0, -1}}, {{0, 0, 0}, {0, 1, 0}, {-1, 0, 0}},
      {{0, 0, 0}, {1, 0, 0}, {0, -1, 0}}, {{0, 0, 1}, {1, 0, 0}, {0, 0, 1}},
      {{1, 0, 0}, {0, 1, 0}, {1, 0, 0}},  {{1, 1, 1}, {-1, 0, 0}, {0, 1, 0}}};
  // DCORVG rows of Wuerfel.GEO: (comp, T, S)
  const double par[8][3] = {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1},
                            {3, 0, 0}, {3, 1, 0}, {3, 1, -1}, {3, 0, -1}};
  for (int i = 0; i < 8; i++) {
    const int c = (int)par[i][0];
    const double *P = plane[c][0], *A = plane[c][1], *N = plane[c][2];
    const double B[3] = {A[1] * N[2] - A[2] * N[1], A[2] * N[0] - A[0] * N[2],
                         A[0] * N[1] - A[1] * N[0]};
    for (int d = 0; d < 3; d++) v[i][d] = P[d] + par[i][1] * A[d] + par[i][2] * B[d];
  }
}

// New vertices of a regularly refined hexahedron as averages of the parent
// vertices: 8..25 edge midpoints and face centres, 26 the cell centre
// (src/Refinement/RefHexaRegDesc.C:163-173 DatEdgeVertex / :202-210 DatFaceVertex).
static const int kNewVertexFrom[27][9] = {
    {1, 0}, {1, 1}, {1, 2}, {1, 3}, {1, 4}, {1, 5}, {1, 6}, {1, 7},
    {2, 0, 1}, {2, 1, 2}, {2, 2, 3}, {2, 3, 0}, {4, 0, 1, 2, 3},
    {2, 1, 5}, {2, 5, 4}, {2, 4, 0}, {4, 0, 1, 5, 4},
    {2, 2, 6}, {2, 6, 5}, {4, 1, 2, 6, 5},
    {2, 3, 7}, {2, 7, 6}, {4, 2, 3, 7, 6},
    {2, 4, 7}, {4, 3, 0, 4, 7}, {4, 4, 5, 6, 7},
    {8, 0, 1, 2, 3, 4, 5, 6, 7}};

// Child vertex lists of the regular refinement (RefHexaRegDesc.C:55-60).
static const int kChildVertex[8][8] = {
    {0, 8, 12, 11, 15, 16, 26, 24}, {1, 9, 12, 8, 13, 19, 26, 16},
    {2, 10, 12, 9, 17, 22, 26, 19}, {3, 11, 12, 10, 20, 24, 26, 22},
    {4, 23, 25, 14, 15, 24, 26, 16}, {5, 14, 25, 18, 13, 16, 26, 19},
    {6, 18, 25, 21, 17, 19, 26, 22}, {7, 21, 25, 23, 20, 22, 26, 24}};

// Hexahedron faces (src/Geometry/Hexahedron.C:42-44).
static const int kFaceVertex[6][4] = {{0, 1, 2, 3}, {0, 4, 5, 1}, {1, 5, 6, 2},
                                      {2, 6, 7, 3}, {0, 3, 7, 4}, {4, 7, 6, 5}};

struct Tree {
  std::vector<std::array<double, 3>> X;  // vertex coordinates
  struct Cell {
    int v[8];
    int child0 = -1;  // index of child 0, children are contiguous
    int parent = -1;
    int depth = 0;
  };
  std::vector<Cell> cells;
  std::map<std::array<double, 3>, int> vmap;

  int vertex(const std::array<double, 3> &p) {
    auto it = vmap.find(p);
    if (it != vmap.end()) return it->second;
    int id = (int)X.size();
    X.push_back(p);
    vmap.emplace(p, id);
    return id;
  }
  Tree() {
    double v[8][3];
    macro_vertices(v);
    Cell c;
    for (int i = 0; i < 8; i++) c.v[i] = vertex({v[i][0], v[i][1], v[i][2]});
    cells.push_back(c);
  }
  // TDomain::RegRefineAll (src/Geometry/Domain.C:1064) over the current leaves
  void refine_leaves() {
    const size_t n = cells.size();
    for (size_t ci = 0; ci < n; ci++) {
      if (cells[ci].child0 >= 0) continue;
      int nv[27];
      for (int k = 0; k < 27; k++) {
        const int m = kNewVertexFrom[k][0];
        std::array<double, 3> p{0, 0, 0};
        for (int q = 0; q < m; q++)
          for (int d = 0; d < 3; d++) p[d] += X[cells[ci].v[kNewVertexFrom[k][1 + q]]][d];
        for (int d = 0; d < 3; d++) p[d] /= m;
        nv[k] = vertex(p);
      }
      const int first = (int)cells.size();
      for (int j = 0; j < 8; j++) {
        Cell ch;
        for (int q = 0; q < 8; q++) ch.v[q] = nv[kChildVertex[j][q]];
        ch.parent = (int)ci;
        ch.depth = cells[ci].depth + 1;
        cells.push_back(ch);
      }
      cells[ci].child0 = first;
    }
  }
  // GetCollection(It_Finest) at a given depth: depth-first, children 0..7
  // (src/Geometry/It_Finest.C, It_LE.C:42-78).
  void collect(int cell, int depth, std::vector<int> &out) const {
    if (cells[cell].depth == depth) {
      out.push_back(cell);
      return;
    }
    for (int j = 0; j < 8; j++) collect(cells[cell].child0 + j, depth, out);
  }
};

// ===========================================================================
// Elements: Q1/Q2 conforming Lagrange on the reference cube, x fastest
// (include/FE3D/BF_C_H_Q1_3D.h:31-52, NF_C_H_Q2_3D.h:39-47).
// ===========================================================================
static void lagrange1d(int order, double t, double *v, double *d) {
  if (order == 1) {
    v[0] = 0.5 * (1 - t); v[1] = 0.5 * (1 + t);
    d[0] = -0.5; d[1] = 0.5;
  } else {
    v[0] = 0.5 * t * (t - 1); v[1] = 1 - t * t; v[2] = 0.5 * t * (t + 1);
    d[0] = t - 0.5; d[1] = -2 * t; d[2] = t + 0.5;
  }
}
// values / d/dxi / d/deta / d/dzeta of all basis functions at one point
static void basis(int order, const double r[3], double *f, double *fx, double *fy,
                  double *fz) {
  const int k = order + 1;
  double v[3][3], d[3][3];
  for (int a = 0; a < 3; a++) lagrange1d(order, r[a], v[a], d[a]);
  for (int c = 0; c < k; c++)
    for (int b = 0; b < k; b++)
      for (int a = 0; a < k; a++) {
        const int j = a + k * (b + k * c);
        f[j] = v[0][a] * v[1][b] * v[2][c];
        fx[j] = d[0][a] * v[1][b] * v[2][c];
        fy[j] = v[0][a] * d[1][b] * v[2][c];
        fz[j] = v[0][a] * v[1][b] * d[2][c];
      }
}
static void node_ref(int order, int j, double r[3]) {
  const int k = order + 1;
  const int a = j % k, b = (j / k) % k, c = j / (k * k);
  const double h = 2.0 / order;
  r[0] = -1 + h * a; r[1] = -1 + h * b; r[2] = -1 + h * c;
}

// ===========================================================================
// Quadrature (src/QuadFormulas/QuadFormulaHexa.C:70-146), index l = k + n*(j + n*i)
// with xi = node[k], eta = node[j], zeta = node[i].
// ===========================================================================
struct Quad {
  std::vector<double> w, xi, eta, zeta;
};
static Quad gauss_hexa(int npt) {
  double w1[4], x1[4];
  if (npt == 2) {
    const double a = 0.5773502691896257645091489;
    x1[0] = -a; x1[1] = a; w1[0] = w1[1] = 1.0;
  } else if (npt == 3) {
    x1[0] = -0.7745966692414833770358530; x1[1] = 0; x1[2] = 0.7745966692414833770358530;
    w1[0] = 0.5555555555555555555555558; w1[1] = 0.8888888888888888888888888;
    w1[2] = 0.5555555555555555555555558;
  } else {
    x1[0] = 0.33998104358485626480266576; x1[1] = 0.86113631159405257522394649;
    x1[2] = -0.86113631159405257522394649; x1[3] = -0.33998104358485626480266576;
    w1[0] = 0.652145154862546142626936051; w1[1] = 0.347854845137453857373063949;
    w1[2] = 0.347854845137453857373063949; w1[3] = 0.652145154862546142626936051;
  }
  Quad q;
  for (int i = 0; i < npt; i++)
    for (int j = 0; j < npt; j++)
      for (int k = 0; k < npt; k++) {
        q.w.push_back(w1[i] * w1[j] * w1[k]);
        q.xi.push_back(x1[k]); q.eta.push_back(x1[j]); q.zeta.push_back(x1[i]);
      }
  return q;
}
// rule from 3*degree (2*degree if INTERNAL_PROBLEM_LINEAR), FEDatabase3D.C:1203-1213,1035-1053
static int rule_points(int order, bool linear) {
  const int deg = (linear ? 2 : 3) * order;
  if (deg <= 2) return 2;
  if (deg <= 5) return 3;
  return 4;
}

// ===========================================================================
// Coefficient functors: restatements of the example headers' callbacks.
// ===========================================================================
struct Coeffs {
  double eps, b[3], c, f;
};
static void eval_b_fourier(const double *prm, double x, double y, double z, double b[3]) {
  // b = curl(Apot), Apot_c = sum_m a sin(2 pi k.x + phi)
  double dA[3][3] = {{0}};  // dA[c][j] = d Apot_c / d x_j
  const double *F = prm + PMX_PRM_FOURIER;
  for (int c = 0; c < 3; c++)
    for (int m = 0; m < 8; m++) {
      const double *t = F + (c * 8 + m) * 5;
      const double arg = 2 * kPi * (t[1] * x + t[2] * y + t[3] * z) + t[4];
      const double cs = t[0] * std::cos(arg) * 2 * kPi;
      dA[c][0] += cs * t[1]; dA[c][1] += cs * t[2]; dA[c][2] += cs * t[3];
    }
  const double s = prm[PMX_PRM_BSCALE];
  b[0] = s * (dA[2][1] - dA[1][2]);
  b[1] = s * (dA[0][2] - dA[2][0]);
  b[2] = s * (dA[1][0] - dA[0][1]);
}
static double grf(const double *prm, double x, double y, double z) {
  const double *G = prm + PMX_PRM_GRF;
  double s = 0;
  for (int k = 0; k < 64; k++) {
    const double *t = G + 4 * k;
    s += std::cos(t[0] * x + t[1] * y + t[2] * z + t[3]);
  }
  return s * std::sqrt(2.0 / 64.0);
}
static Coeffs coeffs_at(const double *prm, double x, double y, double z) {
  Coeffs k;
  const int ex = (int)prm[PMX_PRM_EXAMPLE];
  const double eps = prm[PMX_PRM_EPS];
  k.eps = eps; k.c = prm[PMX_PRM_C];
  k.b[0] = prm[PMX_PRM_B0]; k.b[1] = prm[PMX_PRM_B1]; k.b[2] = prm[PMX_PRM_B2];
  const double t = prm[PMX_PRM_TIME];
  switch (ex) {
    case PMX_EX_LAPLACE:  // Examples/CD_3D/Laplace.h:34-53
      k.b[0] = k.b[1] = k.b[2] = 0; k.c = 0;
      k.f = 3 * kPi * kPi * eps * std::sin(kPi * x) * std::sin(kPi * y) * std::sin(kPi * z);
      break;
    case PMX_EX_SINE_CD: {
      const double sx = std::sin(kPi * x), sy = std::sin(kPi * y), sz = std::sin(kPi * z);
      const double cx = std::cos(kPi * x), cy = std::cos(kPi * y), cz = std::cos(kPi * z);
      const double u = sx * sy * sz;
      k.f = 3 * kPi * kPi * eps * u +
            kPi * (k.b[0] * cx * sy * sz + k.b[1] * sx * cy * sz + k.b[2] * sx * sy * cz) +
            k.c * u;
      break;
    }
    case PMX_EX_CURL_FOURIER:
      eval_b_fourier(prm, x, y, z, k.b); k.c = 0; k.f = 1.0;
      break;
    case PMX_EX_HEAT_AMC:  // Examples/TCD_3D/amc.h:60-100
      k.eps = 1.0; k.b[0] = k.b[1] = k.b[2] = 0; k.c = 0;
      k.f = (3. * 1.0 * kPi * kPi - 0.1) * std::exp(-0.1 * t) * std::sin(kPi * x) *
            std::cos(kPi * y) * std::cos(kPi * z);
      break;
    case PMX_EX_LOGNORMAL:
      k.eps = eps * std::exp(prm[PMX_PRM_GRF_SIGMA] * grf(prm, x, y, z)); k.f = 1.0;
      break;
    case PMX_EX_ROTATING:
      k.b[0] = -(y - 0.5); k.b[1] = x - 0.5; k.b[2] = 0.2; k.c = 0; k.f = 0.0;
      break;
    default:
      k.f = 0;
  }
  return k;
}
// BoundValue / Exact of the examples
static double bound_value(const double *prm, double x, double y, double z) {
  const int ex = (int)prm[PMX_PRM_EXAMPLE];
  const double t = prm[PMX_PRM_TIME];
  switch (ex) {
    case PMX_EX_LAPLACE:
    case PMX_EX_SINE_CD:
      return std::sin(kPi * x) * std::sin(kPi * y) * std::sin(kPi * z);
    case PMX_EX_HEAT_AMC:
      return std::exp(-0.1 * t) * std::sin(kPi * x) * std::cos(kPi * y) * std::cos(kPi * z);
    default:
      return 0.0;
  }
}
static void exact(const double *prm, double x, double y, double z, double v[4]) {
  const int ex = (int)prm[PMX_PRM_EXAMPLE];
  const double t = prm[PMX_PRM_TIME];
  if (ex == PMX_EX_HEAT_AMC) {  // amc.h:16-26
    const double e = std::exp(-0.1 * t);
    v[0] = e * std::sin(kPi * x) * std::cos(kPi * y) * std::cos(kPi * z);
    v[1] = kPi * e * std::cos(kPi * x) * std::cos(kPi * y) * std::cos(kPi * z);
    v[2] = -kPi * e * std::sin(kPi * x) * std::sin(kPi * y) * std::cos(kPi * z);
    v[3] = -kPi * e * std::sin(kPi * x) * std::cos(kPi * y) * std::sin(kPi * z);
  } else if (ex == PMX_EX_ROTATING) {
    const double *B = prm + PMX_PRM_BLOB;
    const double r2 = (x - B[0]) * (x - B[0]) + (y - B[1]) * (y - B[1]) + (z - B[2]) * (z - B[2]);
    v[0] = std::exp(-r2 / (2 * B[3] * B[3]));
    v[1] = v[2] = v[3] = 0;
  } else {  // Laplace.h:12-19
    v[0] = std::sin(kPi * x) * std::sin(kPi * y) * std::sin(kPi * z);
    v[1] = kPi * std::cos(kPi * x) * std::sin(kPi * y) * std::sin(kPi * z);
    v[2] = kPi * std::sin(kPi * x) * std::cos(kPi * y) * std::sin(kPi * z);
    v[3] = kPi * std::sin(kPi * x) * std::sin(kPi * y) * std::cos(kPi * z);
  }
}

// Compute_SDFEM_delta, types 0/1/2 with CELL_MEASURE 0 (src/FE/ConvDiff.C:520-611)
static double sdfem_delta(const double *prm, double hK, double eps, double b1, double b2,
                          double b3) {
  const double delta0 = prm[PMX_PRM_DELTA0], delta1 = prm[PMX_PRM_DELTA1];
  if (std::fabs(eps) < 1e-20) eps = 1e-20;
  const double h = hK;
  const double nb = std::sqrt(b1 * b1 + b2 * b2 + b3 * b3);
  const int type = (int)prm[PMX_PRM_SDFEM_TYPE];
  double delta = 0;
  if (type == 0) {
    delta = (eps < h * nb) ? delta0 * h / nb : delta1 * h * h / eps;
  } else {
    const double pd = prm[PMX_PRM_POLYDEG] > 0 ? prm[PMX_PRM_POLYDEG] : 1.0;
    const double nu = 1.0 / pd;
    if (nb > 0) {
      const double alpha = nu * nb * h / (2 * eps);
      delta = nu * h * (1 / std::tanh(alpha) - 1 / alpha) / (2 * nb);
    }
    delta *= delta0;
  }
  return delta;
}
// RhsAssemble_SUPG's own parameter (src/FE/TCD3D.C:396-408); b = 0 (0/0 in
// the reference) gives 0
static double rhs_supg_tau(double hK, double eps, const double b[3]) {
  const double bb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  if (!(bb > 0)) return 0.0;
  const double pe = hK * bb / (2 * eps);
  double tau;
  if (pe < 1e-2) tau = pe / (6 * bb);
  else tau = (1 / std::tanh(pe) - 1 / pe) / (2 * bb);
  return tau * hK;
}

// ===========================================================================
// FE space: ConstructSpace restatement (src/FE/FESpace3D.C:485-2080)
// ===========================================================================
static const int kFirstMark = -10;  // include/FE/FE3DMapper.h:43

struct Space {
  int order = 1, nref = 0, nb = 8;
  std::vector<int> cells;                     // tree cell ids in collection order
  std::vector<std::array<double, 24>> verts;  // cell vertex coordinates
  std::vector<int> begin, global;             // BeginIndex, GlobalNumbers
  int N = 0, active = 0;
  std::vector<std::array<double, 3>> dofx;    // DOF coordinates
  std::vector<int> rowptr, kcol;              // sorted CSR (TSquareStructure3D + Sort)
  const struct Tree *tree = nullptr;          // refinement tree the collection comes from
  std::shared_ptr<struct Tree> owned;         // set for standalone spaces
};

// THexaAffin::SetCell + GetOrigFromRef (src/FE/HexaAffin.C:891-923, 61-78)
struct Affine {
  double xc[4], yc[4], zc[4], det;
  explicit Affine(const double *v) {
    const double *p0 = v, *p1 = v + 3, *p3 = v + 9, *p4 = v + 12;
    xc[0] = (p1[0] + p3[0] + p4[0] - p0[0]) * 0.5; xc[1] = (p1[0] - p0[0]) * 0.5;
    xc[2] = (p3[0] - p0[0]) * 0.5; xc[3] = (p4[0] - p0[0]) * 0.5;
    yc[0] = (p1[1] + p3[1] + p4[1] - p0[1]) * 0.5; yc[1] = (p1[1] - p0[1]) * 0.5;
    yc[2] = (p3[1] - p0[1]) * 0.5; yc[3] = (p4[1] - p0[1]) * 0.5;
    zc[0] = (p1[2] + p3[2] + p4[2] - p0[2]) * 0.5; zc[1] = (p1[2] - p0[2]) * 0.5;
    zc[2] = (p3[2] - p0[2]) * 0.5; zc[3] = (p4[2] - p0[2]) * 0.5;
    det = xc[1] * yc[2] * zc[3] + xc[2] * yc[3] * zc[1] + xc[3] * yc[1] * zc[2] -
          xc[3] * yc[2] * zc[1] - xc[2] * yc[1] * zc[3] - xc[1] * yc[3] * zc[2];
  }
  void map(const double r[3], double X[3]) const {
    X[0] = xc[0] + xc[1] * r[0] + xc[2] * r[1] + xc[3] * r[2];
    X[1] = yc[0] + yc[1] * r[0] + yc[2] * r[1] + yc[3] * r[2];
    X[2] = zc[0] + zc[1] * r[0] + zc[2] * r[1] + zc[3] * r[2];
  }
  // inverse Jacobian rows i11..i33 (HexaAffin.C:604-625)
  void inv(double I[9]) const {
    const double r = 1 / det;
    I[0] = (yc[2] * zc[3] - yc[3] * zc[2]) * r; I[1] = (yc[3] * zc[1] - yc[1] * zc[3]) * r;
    I[2] = (yc[1] * zc[2] - yc[2] * zc[1]) * r; I[3] = (xc[3] * zc[2] - xc[2] * zc[3]) * r;
    I[4] = (xc[1] * zc[3] - xc[3] * zc[1]) * r; I[5] = (xc[2] * zc[1] - xc[1] * zc[2]) * r;
    I[6] = (xc[2] * yc[3] - xc[3] * yc[2]) * r; I[7] = (xc[3] * yc[1] - xc[1] * yc[3]) * r;
    I[8] = (xc[1] * yc[2] - xc[2] * yc[1]) * r;
  }
};
// THexahedron::GetDiameter (src/Geometry/Hexahedron.C:104-130)
static double diameter(const double *v) {
  const int pr[4][2] = {{0, 6}, {1, 7}, {2, 4}, {3, 5}};
  double d = 0;
  for (auto &p : pr) {
    double s = 0;
    for (int k = 0; k < 3; k++) s += (v[3 * p[0] + k] - v[3 * p[1] + k]) * (v[3 * p[0] + k] - v[3 * p[1] + k]);
    d = std::max(d, s);
  }
  return std::sqrt(d);
}

// hK of a cell for CELL_MEASURE (Assemble3D.C:314-336): 0 and every id
// without its own case the diameter, 2 the shortest edge, 1 and 3 the cube
// root of the measure (adet: |det J| of the affine map from [-1,1]^3)
static double cell_hk(const double *prm, const double *v, double adet) {
  const int m = (int)prm[PMX_PRM_CELL_MEASURE];
  if (m == 2) {
    const int e[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                          {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    double d = HUGE_VAL;
    for (auto &p : e) {
      double s2 = 0;
      for (int k = 0; k < 3; k++) s2 += (v[3 * p[0] + k] - v[3 * p[1] + k]) * (v[3 * p[0] + k] - v[3 * p[1] + k]);
      d = std::min(d, s2);
    }
    return std::sqrt(d);
  }
  if (m == 1 || m == 3) return std::cbrt(8.0 * adet);
  return diameter(v);
}
// Mesh_size_in_convection_direction_without_storing, hexahedron branch
// (src/FE/ConvDiff.C:430-485): the coefficients of the 8 trilinear vertex
// basis functions in (1, x, y, z, xy, xz, yz, xyz) from the vertex values
// (a dense 8 x 8 solve), their gradients at the barycentre, and
// 2 |b| / sum_i |b . grad phi_i| (hK when the sum is below 1e-10)
static double mesh_size_convection(const double *v, double hK, double b1, double b2, double b3) {
  double sx = 0, sy = 0, sz = 0;
  for (int i = 0; i < 8; i++) { sx += v[3 * i]; sy += v[3 * i + 1]; sz += v[3 * i + 2]; }
  sx /= 8; sy /= 8; sz /= 8;
  // A[k][j] = m_j(vertex k); solve A C = I (C column i: coefficients of phi_i)
  double A[8][16];
  for (int k = 0; k < 8; k++) {
    const double x = v[3 * k], y = v[3 * k + 1], z = v[3 * k + 2];
    const double m[8] = {1, x, y, z, x * y, x * z, y * z, x * y * z};
    for (int j = 0; j < 8; j++) { A[k][j] = m[j]; A[k][8 + j] = k == j ? 1.0 : 0.0; }
  }
  for (int c = 0; c < 8; c++) {  // Gauss-Jordan with partial pivoting
    int piv = c;
    for (int r = c + 1; r < 8; r++)
      if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
    for (int j = 0; j < 16; j++) std::swap(A[c][j], A[piv][j]);
    const double d = A[c][c];
    for (int j = 0; j < 16; j++) A[c][j] /= d;
    for (int r = 0; r < 8; r++)
      if (r != c) {
        const double f = A[r][c];
        for (int j = 0; j < 16; j++) A[r][j] -= f * A[c][j];
      }
  }
  double den = 0;
  for (int i = 0; i < 8; i++) {
    double C[8];
    for (int j = 0; j < 8; j++) C[j] = A[j][8 + i];
    double val = b1 * (C[1] + C[4] * sy + C[5] * sz + C[7] * sy * sz);
    val += b2 * (C[2] + C[4] * sx + C[6] * sz + C[7] * sx * sz);
    val += b3 * (C[3] + C[5] * sx + C[6] * sy + C[7] * sx * sy);
    den += std::fabs(val);
  }
  if (den < 1e-10) return hK;
  return 2 * std::sqrt(b1 * b1 + b2 * b2 + b3 * b3) / den;
}

// TFE3DMapper::MapDOF (src/FE/FE3DMapper.C:173-296): union of two local
// positions, the smaller position keeps the mark, larger marks win on merge.
static void map_dof(std::vector<int> &G, int dof0, int dof1, int &counter) {
  int w0 = dof0, w1 = dof1, v;
  while ((v = G[w0]) > -1) w0 = v;
  while ((v = G[w1]) > -1) w1 = v;
  if (G[dof0] == -1 && G[dof1] == -1) {
    counter--;
    if (dof0 < dof1) { G[dof0] = counter; G[dof1] = dof0; }
    else { G[dof1] = counter; G[dof0] = dof1; }
  } else if (G[dof0] == -1) {
    if (w1 > dof0) { G[dof0] = G[w1]; G[dof1] = dof0; G[w1] = dof0; }
    else G[dof0] = w1;
  } else if (G[dof1] == -1) {
    if (w0 > dof1) { G[dof1] = G[w0]; G[dof0] = dof1; G[w0] = dof1; }
    else G[dof1] = w0;
  } else if (G[w0] != G[w1]) {
    const int e = std::max(G[w0], G[w1]);
    const int w = std::min(w0, w1);
    G[w0] = w; G[w1] = w; G[w] = e;
  }
}
// TFE3DMapper::MapBound (FE3DMapper.C:103-171), no hanging nodes
static void map_bound(std::vector<int> &G, int dof, int &bcounter) {
  int w = dof, v;
  while ((v = G[w]) > -1) w = v;
  if (G[w] == -1) {
    bcounter--;
    G[dof] = bcounter;
  } else if (G[w] < bcounter) {
    bcounter--;
    G[w] = bcounter;
    if (dof != w) G[dof] = w;
  }
}

static void build_space(Space &S, const Tree &T, int depth, int order) {
  S.order = order;
  S.nref = depth;
  S.tree = &T;
  S.nb = (order + 1) * (order + 1) * (order + 1);
  T.collect(0, depth, S.cells);
  const int nc = (int)S.cells.size();
  S.verts.resize(nc);
  for (int i = 0; i < nc; i++)
    for (int q = 0; q < 8; q++)
      for (int d = 0; d < 3; d++) S.verts[i][3 * q + d] = T.X[T.cells[S.cells[i]].v[q]][d];

  // local DOF coordinates, used to pair the DOFs of a shared face (the
  // reference pairs them with the mapper tables of FE3DMapper for the joint's
  // map type; for a conforming brick mesh both give geometric coincidence)
  std::vector<std::array<double, 3>> pos((size_t)nc * S.nb);
  for (int i = 0; i < nc; i++) {
    Affine A(S.verts[i].data());
    for (int j = 0; j < S.nb; j++) {
      double r[3], X[3];
      node_ref(order, j, r);
      A.map(r, X);
      pos[(size_t)i * S.nb + j] = {X[0], X[1], X[2]};
    }
  }
  // face DOF lists: local DOFs with the face's reference coordinate fixed
  // face 0: zeta=-1, 1: eta=-1, 2: xi=+1, 3: eta=+1, 4: xi=-1, 5: zeta=+1
  const int fdir[6] = {2, 1, 0, 1, 0, 2};
  const double fval[6] = {-1, -1, 1, 1, -1, 1};
  std::vector<int> facedofs[6];
  for (int f = 0; f < 6; f++)
    for (int j = 0; j < S.nb; j++) {
      double r[3];
      node_ref(order, j, r);
      if (r[fdir[f]] == fval[f]) facedofs[f].push_back(j);
    }
  // face neighbours through the face centre
  std::map<std::array<double, 3>, std::vector<std::pair<int, int>>> faces;
  for (int i = 0; i < nc; i++)
    for (int f = 0; f < 6; f++) {
      std::array<double, 3> c{0, 0, 0};
      for (int q = 0; q < 4; q++)
        for (int d = 0; d < 3; d++) c[d] += S.verts[i][3 * kFaceVertex[f][q] + d] * 0.25;
      faces[c].push_back({i, f});
    }
  std::vector<std::array<int, 6>> neigh(nc);
  std::vector<std::array<int, 6>> neighface(nc);
  for (auto &kv : faces) {
    auto &l = kv.second;
    if (l.size() == 1) { neigh[l[0].first][l[0].second] = -1; }
    else {
      neigh[l[0].first][l[0].second] = l[1].first; neighface[l[0].first][l[0].second] = l[1].second;
      neigh[l[1].first][l[1].second] = l[0].first; neighface[l[1].first][l[1].second] = l[0].second;
    }
  }
  S.begin.resize(nc + 1);
  for (int i = 0; i <= nc; i++) S.begin[i] = i * S.nb;
  const int sum = nc * S.nb;
  // upper bound for Dirichlet marks (FESpace3D.C:640-690): joint DOFs of boundary faces
  int dub = 0;
  for (int i = 0; i < nc; i++)
    for (int f = 0; f < 6; f++)
      if (neigh[i][f] < 0) dub += (int)facedofs[f].size();
  std::vector<int> &G = S.global;
  G.assign(sum, -1);
  int dcounter = kFirstMark + 1;
  int counter = -dub + kFirstMark + 1;  // no non-Dirichlet boundary classes
  for (int i = 0; i < nc; i++) {
    for (int f = 0; f < 6; f++) {
      const int n = neigh[i][f];
      if (n < 0) {  // boundary face, every example has BoundCondition == DIRICHLET
        for (int j : facedofs[f]) map_bound(G, S.begin[i] + j, dcounter);
      } else if (n > i) {
        const int g = neighface[i][f];
        for (int j : facedofs[f]) {
          const auto &p = pos[(size_t)i * S.nb + j];
          for (int l : facedofs[g])
            if (pos[(size_t)n * S.nb + l] == p) { map_dof(G, S.begin[i] + j, S.begin[n] + l, counter); break; }
        }
      }
    }
    // inner DOFs of the element (FESpace3D.C:1467-1474): Q2 cell centre
    if (order == 2) { counter--; G[S.begin[i] + 13] = counter; }
  }
  // final numbering (FESpace3D.C:1739-1830): positions in increasing order,
  // inner marks -> 0.., Dirichlet marks -> N_Inner + ..
  const int inner_mark = kFirstMark - dub;
  int n_inner = 0, n_dir = 0;
  std::vector<int> cls(sum, -1);
  for (int p = 0; p < sum; p++) {
    const int l = G[p];
    if (l < -1) {
      if (l <= inner_mark) { G[p] = n_inner++; cls[p] = 0; }
      else { G[p] = n_dir++; cls[p] = 1; }
    } else {
      G[p] = G[l];
      cls[p] = cls[l];
    }
  }
  for (int p = 0; p < sum; p++)
    if (cls[p] == 1) G[p] += n_inner;
  S.N = n_inner + n_dir;
  S.active = n_inner;
  S.dofx.resize(S.N);
  for (int p = 0; p < sum; p++) S.dofx[G[p]] = pos[p];

  // TSquareStructure3D(space) + Sort (src/FE/SquareStructure3D.C:59-200, SquareStructure.C:91-110)
  std::vector<std::vector<int>> rows(S.active);
  for (int i = 0; i < nc; i++) {
    const int *dof = &G[S.begin[i]];
    for (int a = 0; a < S.nb; a++) {
      if (dof[a] >= S.active) continue;
      auto &r = rows[dof[a]];
      for (int b = 0; b < S.nb; b++) r.push_back(dof[b]);
    }
  }
  S.rowptr.assign(S.N + 1, 0);
  S.kcol.clear();
  for (int r = 0; r < S.N; r++) {
    if (r < S.active) {
      auto &v = rows[r];
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
      S.kcol.insert(S.kcol.end(), v.begin(), v.end());
    } else {
      S.kcol.push_back(r);  // Dirichlet row: diagonal only
    }
    S.rowptr[r + 1] = (int)S.kcol.size();
  }
}

// ===========================================================================
// Assembly: TAssembleMat3D::Assemble3D cell loop (src/System/AssembleMat3D.C:415-521)
// ===========================================================================
// Forms without local matrices (RhsAssemble, RhsAssemble_SUPG) leave E and
// E2 as they are.
static int assemble(const Space &S, const double *prm, double *E, double *rhs, double *E2) {
  const int form = (int)prm[PMX_PRM_FORM];
  const bool linear = prm[PMX_PRM_LINEAR] != 0;
  const bool rhs_only = form == PMX_FORM_RHS || form == PMX_FORM_RHS_SUPG;
  const int nnz = S.rowptr[S.N];
  if (!rhs_only) std::memset(E, 0, sizeof(double) * nnz);
  if (E2 && !rhs_only) std::memset(E2, 0, sizeof(double) * nnz);
  std::memset(rhs, 0, sizeof(double) * S.N);
  const Quad Q = gauss_hexa(rule_points(S.order, linear));
  const int nq = (int)Q.w.size(), nb = S.nb;
  // reference tables of the rule
  std::vector<double> rf(nq * nb), rx(nq * nb), ry(nq * nb), rz(nq * nb);
  for (int q = 0; q < nq; q++) {
    const double r[3] = {Q.xi[q], Q.eta[q], Q.zeta[q]};
    basis(S.order, r, &rf[q * nb], &rx[q * nb], &ry[q * nb], &rz[q * nb]);
  }
  std::vector<double> L(nb * nb), L2(nb * nb), Lr(nb);
  std::vector<double> gx(nb), gy(nb), gz(nb);
  const double tau = prm[PMX_PRM_TAU], th1 = prm[PMX_PRM_THETA1];
  for (int i = 0; i < (int)S.cells.size(); i++) {
    const double *v = S.verts[i].data();
    Affine A(v);
    double I[9];
    A.inv(I);
    const double adet = std::fabs(A.det);
    const double hK = cell_hk(prm, v, adet);
    // the h_K of Compute_SDFEM_delta: with CELL_MEASURE 4 the mesh size in the
    // convection direction of the cell's first call, kept for the cell
    // (INTERNAL_HK_CONVECTION, ConvDiff.C:373-383, Assemble3D.C:389-405)
    double hk_conv = -1;
    auto hd = [&](double b1, double b2, double b3) {
      if ((int)prm[PMX_PRM_CELL_MEASURE] != 4) return hK;
      if (hk_conv < 0) hk_conv = mesh_size_convection(v, hK, b1, b2, b3);
      return hk_conv;
    };
    std::fill(L.begin(), L.end(), 0.0);
    std::fill(L2.begin(), L2.end(), 0.0);
    std::fill(Lr.begin(), Lr.end(), 0.0);
    for (int q = 0; q < nq; q++) {
      const double r[3] = {Q.xi[q], Q.eta[q], Q.zeta[q]};
      double X[3];
      A.map(r, X);
      const Coeffs c = coeffs_at(prm, X[0], X[1], X[2]);
      const double mult = Q.w[q] * adet;
      const double *f = &rf[q * nb];
      for (int j = 0; j < nb; j++) {
        const double a = rx[q * nb + j], b = ry[q * nb + j], d = rz[q * nb + j];
        gx[j] = I[0] * a + I[1] * b + I[2] * d;
        gy[j] = I[3] * a + I[4] * b + I[5] * d;
        gz[j] = I[6] * a + I[7] * b + I[8] * d;
      }
      if (form == PMX_FORM_GALERKIN || form == PMX_FORM_TCD_A) {
        // BilinearAssemble (ConvDiff3D.C:87-146), MatrixARhsAssemble (TCD3D.C:268-322)
        for (int a = 0; a < nb; a++) {
          Lr[a] += mult * f[a] * c.f;
          for (int b = 0; b < nb; b++) {
            double val = c.eps * (gx[a] * gx[b] + gy[a] * gy[b] + gz[a] * gz[b]);
            val += (c.b[0] * gx[b] + c.b[1] * gy[b] + c.b[2] * gz[b]) * f[a];
            val += c.c * f[b] * f[a];
            L[a * nb + b] += mult * val;
          }
        }
      } else if (form == PMX_FORM_SDFEM) {
        // BilinearAssemble_SD (ConvDiff3D.C:150-214): no second derivatives
        const double delta = sdfem_delta(prm, hd(c.b[0], c.b[1], c.b[2]), c.eps, c.b[0], c.b[1], c.b[2]);
        for (int a = 0; a < nb; a++) {
          const double bgv = c.b[0] * gx[a] + c.b[1] * gy[a] + c.b[2] * gz[a];
          Lr[a] += mult * (f[a] + delta * bgv) * c.f;
          for (int b = 0; b < nb; b++) {
            const double bgu = c.b[0] * gx[b] + c.b[1] * gy[b] + c.b[2] * gz[b];
            double val = c.eps * (gx[a] * gx[b] + gy[a] * gy[b] + gz[a] * gz[b]);
            val += bgu * f[a];
            val += c.c * f[b] * f[a];
            val += delta * (bgu + c.c * f[b]) * bgv;
            L[a * nb + b] += mult * val;
          }
        }
      } else if (form == PMX_FORM_MASS) {
        // MatrixMRhsAssemble (TCD3D.C:37-72)
        for (int a = 0; a < nb; a++) {
          Lr[a] += mult * f[a] * c.f;
          for (int b = 0; b < nb; b++) L[a * nb + b] += mult * f[b] * f[a];
        }
      } else if (form == PMX_FORM_MASS_SUPG) {
        // MatrixMRhsAssemble_SUPG (TCD3D.C:74-126): the Galerkin mass matrix and
        // the rhs f (phi + tau b.grad phi), tau from the unscaled coefficients
        const double dl = sdfem_delta(prm, hd(c.b[0], c.b[1], c.b[2]), c.eps, c.b[0], c.b[1], c.b[2]);
        for (int a = 0; a < nb; a++) {
          const double bgv = c.b[0] * gx[a] + c.b[1] * gy[a] + c.b[2] * gz[a];
          Lr[a] += mult * (f[a] + dl * bgv) * c.f;
          for (int b = 0; b < nb; b++) L[a * nb + b] += mult * f[b] * f[a];
        }
      } else if (form == PMX_FORM_RHS) {
        // RhsAssemble (TCD3D.C:423-446)
        for (int a = 0; a < nb; a++) Lr[a] += mult * f[a] * c.f;
      } else if (form == PMX_FORM_RHS_SUPG) {
        // RhsAssemble_SUPG (TCD3D.C:374-421)
        const double tau_r = rhs_supg_tau(hK, c.eps, c.b);
        for (int a = 0; a < nb; a++) {
          const double bgv = c.b[0] * gx[a] + c.b[1] * gy[a] + c.b[2] * gz[a];
          Lr[a] += mult * (f[a] + tau_r * bgv) * c.f;
        }
      } else if (form == PMX_FORM_TCD_SUPG_AK) {
        // MatricesAKRhsAssemble_SUPG (TCD3D.C:128-266), delta from tau*theta1-scaled
        // coefficients and the max-norm of b
        const double val0 = th1 * tau;
        const double c00 = val0 * c.eps, c11 = val0 * c.b[0], c22 = val0 * c.b[1], c33 = val0 * c.b[2];
        double dl = sdfem_delta(prm, hd(c11, c22, c33), c00, c11, c22, c33);
        dl *= val0;
        for (int a = 0; a < nb; a++) {
          const double bgv = dl * (c.b[0] * gx[a] + c.b[1] * gy[a] + c.b[2] * gz[a]);
          const double test = f[a] + bgv;
          Lr[a] += mult * test * c.f;
          for (int b = 0; b < nb; b++) {
            double v1 = c.b[0] * gx[b] + c.b[1] * gy[b] + c.b[2] * gz[b] + c.c * f[b];
            double val = c.eps * (gx[a] * gx[b] + gy[a] * gy[b] + gz[a] * gz[b]) + v1 * test;
            L[a * nb + b] += mult * val;
            L2[a * nb + b] += mult * f[b] * bgv;
          }
        }
      } else {
        return -1;
      }
    }
    // AddLocalSqMatToGlobal: linear search in the sorted row (AssembleMat3D.C:751-857)
    const int *dof = &S.global[S.begin[i]];
    for (int a = 0; a < nb; a++) {
      const int l = dof[a];
      if (rhs_only) {
        if (l < S.active) rhs[l] += Lr[a];
        continue;
      }
      if (l < S.active) {
        for (int b = 0; b < nb; b++) {
          int n = S.rowptr[l];
          const int end = S.rowptr[l + 1];
          for (; n < end; n++)
            if (S.kcol[n] == dof[b]) break;
          if (n == end) return -2;
          E[n] += L[a * nb + b];
          if (E2) E2[n] += L2[a * nb + b];
        }
        rhs[l] += Lr[a];  // AddLocalRhsToGlobal (AssembleMat3D.C:960-976)
      } else {
        const int n = S.rowptr[l];
        if (S.kcol[n] == l) { E[n] = 1.0; if (E2) E2[n] = 1.0; }
      }
    }
    // Dirichlet faces: RHS[dof] = g(node) (AssembleMat3D.C:1031-1096)
    for (int a = 0; a < nb; a++)
      if (dof[a] >= S.active) {
        const auto &p = S.dofx[dof[a]];
        rhs[dof[a]] = bound_value(prm, p[0], p[1], p[2]);
      }
  }
  return 0;
}

// ===========================================================================
// BLAS and level operations (src/FE/Blas1.C, Blas2.C, MGLevel3D.C)
// ===========================================================================
static double ddot(int n, const double *x, const double *y) {
  double s = 0;
  for (int i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}
static void daxpy(int n, double a, const double *x, double *y) {
  for (int i = 0; i < n; i++) y[i] += a * x[i];
}
static void matvec(const Space &S, const double *E, const double *x, double *y) {
  for (int i = 0; i < S.N; i++) {
    double s = 0;
    for (int j = S.rowptr[i]; j < S.rowptr[i + 1]; j++) s += E[j] * x[S.kcol[j]];
    y[i] = s;
  }
}
// ScalarDefect (Blas2.C:37-83): d = f - A x, ||d|| over all rows
static double defect(const Space &S, const double *E, const double *x, const double *f, double *d) {
  double res = 0;
  for (int i = 0; i < S.N; i++) {
    double s = f[i];
    for (int j = S.rowptr[i]; j < S.rowptr[i + 1]; j++) s -= E[j] * x[S.kcol[j]];
    d[i] = s;
    res += s * s;
  }
  return std::sqrt(res);
}
// TMGLevel3D::SOR / SSOR / Jacobi (MGLevel3D.C:280-563); no hanging nodes
static void row_sweep(const Space &S, const double *E, int i, double omega, double *sol,
                      const double *f, const double *src) {
  double s = f[i], diag = 1;
  for (int j = S.rowptr[i]; j < S.rowptr[i + 1]; j++) {
    const int c = S.kcol[j];
    if (c == i) diag = E[j];
    else s -= E[j] * src[c];
  }
  const double t = src[i];
  if (src == sol) sol[i] = omega * (s / diag - t) + t;
  else sol[i] = (1 - omega) * t + omega * s / diag;
}
static void smooth(const Space &S, const double *E, int kind, double omega, double *sol,
                   const double *f, double *aux) {
  const int N = S.N, NA = S.active;
  if (kind == 1) {
    std::memcpy(aux, sol, sizeof(double) * N);
    std::memcpy(sol + NA, f + NA, sizeof(double) * (N - NA));
    for (int i = 0; i < NA; i++) row_sweep(S, E, i, omega, sol, f, aux);
    return;
  }
  std::memcpy(sol + NA, f + NA, sizeof(double) * (N - NA));
  for (int i = 0; i < NA; i++) row_sweep(S, E, i, omega, sol, f, sol);
  if (kind == 3)
    for (int i = NA - 1; i >= 0; i--) row_sweep(S, E, i, omega, sol, f, sol);
}

// SOR / SSOR visiting the active rows in a caller-given order (the same
// update as TMGLevel3D::SOR/SSOR, MGLevel3D.C:280-468, with the row loop
// permuted; used to check multicolour sweeps row for row)
static void smooth_ordered(const Space &S, const double *E, int kind, double omega, double *sol,
                           const double *f, const int *order, int n) {
  std::memcpy(sol + S.active, f + S.active, sizeof(double) * (S.N - S.active));
  for (int k = 0; k < n; k++) row_sweep(S, E, order[k], omega, sol, f, sol);
  if (kind == 3)
    for (int k = n - 1; k >= 0; k--) row_sweep(S, E, order[k], omega, sol, f, sol);
}

// ===========================================================================
// Grid transfers: cellwise Prolongate / DefectRestriction with fine-DOF
// multiplicity (src/FE/MGComponents3D.C:208-488, 732-1014).  QQ of child j is
// the coarse basis evaluated at the child's nodes
// (TFEDatabase3D::GetProlongationMatrix3D, FEDatabase3D.C:1656).
// ===========================================================================
struct Transfer {
  std::vector<int> parent, child;            // per fine cell
  std::vector<std::vector<double>> QQ;       // per fine cell [nb_f][nb_c]
  std::vector<double> mult;                  // per fine DOF
};
static Transfer make_transfer(const Space &C, const Space &F, const Tree &T) {
  Transfer tr;
  std::unordered_map<int, int> cidx;
  for (int i = 0; i < (int)C.cells.size(); i++) cidx[C.cells[i]] = i;
  const int nf = (int)F.cells.size();
  tr.parent.resize(nf); tr.child.resize(nf); tr.QQ.resize(nf);
  for (int i = 0; i < nf; i++) {
    const int tc = F.cells[i];
    const int tp = T.cells[tc].parent;
    tr.parent[i] = cidx.at(tp);
    tr.child[i] = tc - T.cells[tp].child0;
    // fine node -> physical -> parent reference coordinates
    Affine Ap(C.verts[tr.parent[i]].data());
    double I[9];
    Ap.inv(I);
    Affine Af(F.verts[i].data());
    auto &Q = tr.QQ[i];
    Q.resize((size_t)F.nb * C.nb);
    for (int k = 0; k < F.nb; k++) {
      double r[3], X[3], rp[3];
      node_ref(F.order, k, r);
      Af.map(r, X);
      const double dx[3] = {X[0] - Ap.xc[0], X[1] - Ap.yc[0], X[2] - Ap.zc[0]};
      // ref = J^{-1} (X - c0); I[3a+k] = d ref_k / d X_a (HexaAffin.C:612-620)
      for (int k = 0; k < 3; k++) rp[k] = I[k] * dx[0] + I[3 + k] * dx[1] + I[6 + k] * dx[2];
      std::vector<double> v(C.nb), d1(C.nb), d2(C.nb), d3(C.nb);
      basis(C.order, rp, v.data(), d1.data(), d2.data(), d3.data());
      for (int l = 0; l < C.nb; l++) Q[(size_t)k * C.nb + l] = v[l];
    }
  }
  tr.mult.assign(F.N, 0.0);
  for (int i = 0; i < nf; i++)
    for (int j = 0; j < F.nb; j++) tr.mult[F.global[F.begin[i] + j]] += 1;
  return tr;
}
static void prolongate(const Space &C, const Space &F, const Transfer &tr, const double *xc,
                       double *xf) {
  std::memset(xf, 0, sizeof(double) * F.N);
  std::vector<double> val(C.nb);
  for (int i = 0; i < (int)F.cells.size(); i++) {
    const int *cd = &C.global[C.begin[tr.parent[i]]];
    for (int l = 0; l < C.nb; l++) val[l] = xc[cd[l]];
    const int *fd = &F.global[F.begin[i]];
    for (int k = 0; k < F.nb; k++) {
      double s = 0;
      for (int l = 0; l < C.nb; l++) s += tr.QQ[i][(size_t)k * C.nb + l] * val[l];
      xf[fd[k]] += s;
    }
  }
  for (int i = 0; i < F.N; i++)
    if (tr.mult[i]) xf[i] /= tr.mult[i];
}
static void restrict_defect(const Space &C, const Space &F, const Transfer &tr, double *rc,
                            const double *rf) {
  std::memset(rc, 0, sizeof(double) * C.N);
  std::vector<double> fr(F.N);
  for (int i = 0; i < F.N; i++) fr[i] = rf[i] / tr.mult[i];
  for (int i = 0; i < (int)F.cells.size(); i++) {
    const int *cd = &C.global[C.begin[tr.parent[i]]];
    const int *fd = &F.global[F.begin[i]];
    for (int k = 0; k < C.nb; k++) {
      double s = 0;
      for (int l = 0; l < F.nb; l++) s += tr.QQ[i][(size_t)l * C.nb + k] * fr[fd[l]];
      rc[cd[k]] += s;
    }
  }
}

// dense LU with partial pivoting, column-major (restates dgetrf_/dgetrs_ used
// by SolveLinearSystemLapack, src/FE/LinAlg.C:106-160)
static int dense_lu_solve(int n, double *a, double *b) {
  std::vector<int> piv(n);
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (std::fabs(a[k * n + i]) > std::fabs(a[k * n + p])) p = i;
    if (a[k * n + p] == 0) return -1;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; j++) std::swap(a[j * n + k], a[j * n + p]);
    const double inv = 1.0 / a[k * n + k];
    for (int i = k + 1; i < n; i++) a[k * n + i] *= inv;
    for (int j = k + 1; j < n; j++) {
      const double t = a[j * n + k];
      if (t != 0)
        for (int i = k + 1; i < n; i++) a[j * n + i] -= a[k * n + i] * t;
    }
  }
  for (int k = 0; k < n; k++)
    if (piv[k] != k) std::swap(b[k], b[piv[k]]);
  for (int j = 0; j < n; j++)
    for (int i = j + 1; i < n; i++) b[i] -= a[j * n + i] * b[j];
  for (int j = n - 1; j >= 0; j--) {
    b[j] /= a[j * n + j];
    for (int i = 0; i < j; i++) b[i] -= a[j * n + i] * b[j];
  }
  return 0;
}

// ===========================================================================
// Multilevel system: TSystemCD3D levels + TMultiGrid3D + Krylov drivers
// ===========================================================================
struct System {
  Tree tree;
  std::vector<Space> lev;
  std::vector<Transfer> tr;  // tr[l] between lev[l-1] and lev[l]
  std::vector<std::vector<double>> E, rhs, E2;
  // MG level work vectors
  std::vector<std::vector<double>> sol, f, aux0, aux1;
  std::vector<int> rec;
  double opt[16];
  // SC_STEP_LENGTH_CONTROL_FINE_SCALAR (1) / _ALL_SCALAR (2) / off (0)
  int slc = 0;
  std::vector<std::vector<double>> old_sol, old_def;
  // optional row order of the SOR/SSOR sweeps per level (empty: lexicographic,
  // as TMGLevel3D; a multicolour order reproduces a device sweep row for row)
  std::vector<std::vector<int>> order;
};

static void mg_smooth(System &Y, int l, int kind, int n, double omega) {
  const bool ordered = kind != 1 && l < (int)Y.order.size() && !Y.order[l].empty();
  for (int k = 0; k < n; k++) {
    if (ordered)
      smooth_ordered(Y.lev[l], Y.E[l].data(), kind, omega, Y.sol[l].data(), Y.f[l].data(), Y.order[l].data(),
                     (int)Y.order[l].size());
    else
      smooth(Y.lev[l], Y.E[l].data(), kind, omega, Y.sol[l].data(), Y.f[l].data(), Y.aux1[l].data());
  }
}
// TMGLevel3D::StepLengthControl (src/FE/MGLevel3D.C:908-960): with
// x = u - uold and y = A x, omega = (def, y) / (y, y); omega = Parameters[0]
// (the smoother damping, as the system passes SC_SMOOTH_DAMP_FACTOR_SCALAR)
// when (y, y) <= 0 or |omega| < 0.1
static double step_length(const Space &S, const double *E, const double *u, const double *uold, const double *def,
                          double param0) {
  std::vector<double> x(S.N), y(S.N, 0.0);
  for (int i = 0; i < S.N; i++) x[i] = u[i] - uold[i];
  matvec(S, E, x.data(), y.data());
  const double num = ddot(S.N, def, y.data()), nom = ddot(S.N, y.data(), y.data());
  double omega = nom > 0 ? num / nom : param0;
  if (std::fabs(omega) < 0.1) omega = param0;
  return omega;
}
// TMultiGrid3D::Cycle (src/FE/MultiGrid3D.C:449-695); step length control
// (slc, MultiGrid3D.C:465-471, 534-538, 562-566, 671-679) on the finest level
// (1) or every level above the coarsest (2)
static void mg_cycle(System &Y, int i) {
  const double *o = Y.opt;
  const Space &S = Y.lev[i];
  if (i == 0) {
    // CoarseSmooth (MultiGrid3D.C:225-337); ids other than 1, 2, 17 take the
    // switch's default branch, SSOR (as 6 does in the sequential build)
    const int cs = ((int)o[9] == 1 || (int)o[9] == 2 || (int)o[9] == 17) ? (int)o[9] : 3;
    double res = defect(S, Y.E[0].data(), Y.sol[0].data(), Y.f[0].data(), Y.aux0[0].data());
    const double red = o[11] * res;
    int it = 0;
    while (res > red && it < (int)o[10]) {
      if (cs == 17) {
        // TMGLevel3D::SolveExact (MGLevel3D.C:865-905): dense copy + LU
        const int n = S.N;
        std::vector<double> a((size_t)n * n, 0.0);
        for (int r = 0; r < n; r++)
          for (int j = S.rowptr[r]; j < S.rowptr[r + 1]; j++) a[(size_t)S.kcol[j] * n + r] = Y.E[0][j];
        std::vector<double> b(Y.f[0]);
        dense_lu_solve(n, a.data(), b.data());
        Y.sol[0] = b;
      } else {
        mg_smooth(Y, 0, cs, 1, o[8]);
      }
      res = defect(S, Y.E[0].data(), Y.sol[0].data(), Y.f[0].data(), Y.aux0[0].data());
      it++;
    }
    return;
  }
  const Space &C = Y.lev[i - 1];
  const int kind = (int)o[5];
  const bool slc = Y.slc == 2 || (Y.slc == 1 && i == (int)Y.lev.size() - 1);
  if (slc) {
    Y.old_sol[i] = Y.sol[i];
    Y.old_def[i].assign(S.N, 0.0);
    defect(S, Y.E[i].data(), Y.sol[i].data(), Y.f[i].data(), Y.old_def[i].data());
  }
  mg_smooth(Y, i, kind, (int)o[6], o[8]);
  defect(S, Y.E[i].data(), Y.sol[i].data(), Y.f[i].data(), Y.aux0[i].data());
  restrict_defect(C, S, Y.tr[i], Y.f[i - 1].data(), Y.aux0[i].data());
  // CorrectDefect: non-active part of the coarse rhs set to 0; Reset coarse sol
  std::fill(Y.f[i - 1].begin() + C.active, Y.f[i - 1].end(), 0.0);
  std::fill(Y.sol[i - 1].begin(), Y.sol[i - 1].end(), 0.0);
  for (int j = 0; j < Y.rec[i]; j++) mg_cycle(Y, i - 1);
  if ((int)o[12] < 1) Y.rec[i] = 1;  // F-cycle
  prolongate(C, S, Y.tr[i], Y.sol[i - 1].data(), Y.aux0[i].data());
  std::fill(Y.aux0[i].begin() + S.active, Y.aux0[i].end(), 0.0);  // CorrectNodes
  daxpy(S.N, o[13], Y.aux0[i].data(), Y.sol[i].data());           // Update
  mg_smooth(Y, i, kind, (int)o[7], o[8]);
  if (slc) {
    const double a = step_length(S, Y.E[i].data(), Y.sol[i].data(), Y.old_sol[i].data(), Y.old_def[i].data(), o[8]);
    for (int j = 0; j < S.N; j++) Y.sol[i][j] = Y.old_sol[i][j] + a * (Y.sol[i][j] - Y.old_sol[i][j]);
  }
}
// TMultiGridScaIte::Iterate (src/FE/MultiGridScaIte.C:89-130), zero start
static void mg_prec(System &Y, double *z, const double *v) {
  const int L = (int)Y.lev.size() - 1;
  const Space &S = Y.lev[L];
  std::fill(Y.sol[L].begin(), Y.sol[L].end(), 0.0);
  std::memcpy(Y.f[L].data(), v, sizeof(double) * S.N);
  std::memcpy(Y.sol[L].data() + S.active, Y.f[L].data() + S.active, sizeof(double) * (S.N - S.active));
  // SetRecursion (MultiGrid3D.C:717-725)
  const int gam = (int)Y.opt[12];
  Y.rec.assign(Y.lev.size() + 1, 1);
  for (int k = 2; k <= L; k++) Y.rec[k] = gam > 0 ? gam : 2;
  mg_cycle(Y, L);
  std::memcpy(z, Y.sol[L].data(), sizeof(double) * S.N);
}

// Givens helpers and back substitution (src/FE/FgmresIte.C:194-243)
static void gen_rot(double dx, double dy, double &cs, double &sn) {
  if (dy == 0.0) { cs = 1.0; sn = 0.0; }
  else if (std::fabs(dy) > std::fabs(dx)) {
    const double t = dx / dy; sn = 1.0 / std::sqrt(1.0 + t * t); cs = t * sn;
  } else {
    const double t = dy / dx; cs = 1.0 / std::sqrt(1.0 + t * t); sn = t * cs;
  }
}
static void app_rot(double &dx, double &dy, double cs, double sn) {
  const double t = cs * dx + sn * dy;
  dy = -sn * dx + cs * dy;
  dx = t;
}

// TFgmresIte::Iterate, flexible branch (FgmresIte.C:262-625)
static int fgmres(System &Y, const double *rhs, double *sol, double *info) {
  const int L = (int)Y.lev.size() - 1;
  const Space &S = Y.lev[L];
  const double *E = Y.E[L].data();
  const int N = S.N, m = (int)Y.opt[1], maxit = (int)Y.opt[2];
  const double rmin = Y.opt[3], red = Y.opt[4];
  std::vector<std::vector<double>> v(m + 1, std::vector<double>(N)), z(m + 1, std::vector<double>(N));
  std::vector<std::vector<double>> H(m + 1, std::vector<double>(m + 1, 0.0));
  std::vector<double> s(m + 1), cs(m + 1), sn(m + 1), w(N);
  double res = defect(S, E, sol, rhs, v[0].data());
  double res0 = res, beta = res;
  info[2] = res;
  if (beta <= rmin) { info[0] = 0; info[1] = res; return 0; }
  int j = 1;
  while (j <= maxit) {
    for (int k = 0; k < N; k++) v[0][k] *= 1.0 / beta;
    std::fill(s.begin(), s.end(), 0.0);
    s[0] = beta;
    int i;
    for (i = 0; i < m && j <= maxit; i++, j++) {
      mg_prec(Y, z[i].data(), v[i].data());
      matvec(S, E, z[i].data(), w.data());
      for (int k = 0; k <= i; k++) {
        H[k][i] = ddot(N, w.data(), v[k].data());
        daxpy(N, -H[k][i], v[k].data(), w.data());
      }
      H[i + 1][i] = std::sqrt(ddot(N, w.data(), w.data()));
      const double hinv = 1 / H[i + 1][i];
      for (int k = 0; k < N; k++) v[i + 1][k] = w[k] * hinv;
      for (int k = 0; k < i; k++) app_rot(H[k][i], H[k + 1][i], cs[k], sn[k]);
      gen_rot(H[i][i], H[i + 1][i], cs[i], sn[i]);
      app_rot(H[i][i], H[i + 1][i], cs[i], sn[i]);
      app_rot(s[i], s[i + 1], cs[i], sn[i]);
      res = std::fabs(s[i + 1]);
      if (res < rmin || res < red * res0) {
        // UpdateGmresIterate with k = i
        for (int a = i; a >= 0; a--) {
          s[a] /= H[a][a];
          for (int b = a - 1; b >= 0; b--) s[b] -= H[b][a] * s[a];
        }
        for (int a = 0; a <= i; a++) daxpy(N, s[a], z[a].data(), sol);
        info[0] = j; info[1] = res;
        return j;
      }
    }
    // UpdateGmresIterate over the columns this cycle filled (all restart
    // columns in FgmresIte.C:563, which reads unfilled ones when maxit ends
    // the cycle early)
    for (int a = i - 1; a >= 0; a--) {
      s[a] /= H[a][a];
      for (int b = a - 1; b >= 0; b--) s[b] -= H[b][a] * s[a];
    }
    for (int a = 0; a < i; a++) daxpy(N, s[a], z[a].data(), sol);
    res = defect(S, E, sol, rhs, v[0].data());
    res0 = beta = res;
  }
  info[0] = maxit; info[1] = res;
  return maxit;
}
// TFixedPointIte::Iterate (src/FE/FixedPointIte.C:176-300)
static int richardson(System &Y, const double *rhs, double *sol, double *info) {
  const int L = (int)Y.lev.size() - 1;
  const Space &S = Y.lev[L];
  const double *E = Y.E[L].data();
  const int N = S.N, maxit = (int)Y.opt[2];
  std::vector<double> d(N), z(N);
  double res = defect(S, E, sol, rhs, d.data());
  const double res0 = res;
  info[2] = res;
  int it = 0;
  if (res > Y.opt[3])
    for (it = 1; it <= maxit; it++) {
      mg_prec(Y, z.data(), d.data());
      daxpy(N, 1.0, z.data(), sol);
      res = defect(S, E, sol, rhs, d.data());
      if (res <= res0 * Y.opt[4] || res <= Y.opt[3]) break;
    }
  info[0] = std::min(it, maxit); info[1] = res;
  return it;
}
// TBcgs::Iterate (src/FE/Bcgs.C:111-304).  opt[14] selects the flexible branch
// (v = A y); the reference hardcodes flexible = 0 (v = y, Bcgs.C:120,198-206).
static int bicgstab(System &Y, const double *rhs_in, double *sol, double *info) {
  const int L = (int)Y.lev.size() - 1;
  const Space &S = Y.lev[L];
  const double *E = Y.E[L].data();
  const int N = S.N, maxit = (int)Y.opt[2];
  const bool flexible = Y.opt[14] != 0;
  std::vector<double> r(N), rt(N), p(N), y(N), v(N), t(N);
  defect(S, E, sol, rhs_in, r.data());
  rt = r;
  double dnorm0 = std::sqrt(ddot(N, r.data(), rt.data()));
  info[2] = dnorm0;
  double rho_last = 1, alpha = 0, omega = 0, dnorm = dnorm0;
  int i;
  if (dnorm0 <= Y.opt[3]) { info[0] = 0; info[1] = dnorm0; return 0; }
  for (i = 0; i < maxit; i++) {
    const double rho = ddot(N, r.data(), rt.data());
    if (std::fabs(rho) < 1e-50) { info[0] = i; info[1] = dnorm; return -1; }
    if (i == 0) p = r;
    else {
      const double beta = (rho / rho_last) * (alpha / omega);
      for (int k = 0; k < N; k++) p[k] = beta * p[k] + r[k] - beta * omega * v[k];
    }
    rho_last = rho;
    mg_prec(Y, y.data(), p.data());
    if (flexible) matvec(S, E, y.data(), v.data()); else v = y;
    double delta = ddot(N, rt.data(), v.data());
    if (std::fabs(delta) < 1e-50) delta = 1e-50;
    alpha = rho / delta;
    daxpy(N, -alpha, v.data(), r.data());
    daxpy(N, alpha, flexible ? y.data() : p.data(), sol);
    dnorm = std::sqrt(ddot(N, r.data(), r.data()));
    if (dnorm < dnorm0 * Y.opt[4] || dnorm < Y.opt[3]) break;
    mg_prec(Y, y.data(), r.data());
    if (flexible) matvec(S, E, y.data(), t.data()); else t = y;
    delta = ddot(N, t.data(), t.data());
    if (std::fabs(delta) < 1e-50) delta = 1e-50;
    omega = ddot(N, t.data(), r.data()) / delta;
    daxpy(N, omega, flexible ? y.data() : r.data(), sol);
    daxpy(N, -omega, t.data(), r.data());
    dnorm = std::sqrt(ddot(N, r.data(), r.data()));
    if (dnorm < dnorm0 * Y.opt[4] || dnorm < Y.opt[3]) break;
  }
  info[0] = i + 1; info[1] = dnorm;
  return i + 1;
}

static int solve(System &Y, const double *opt, const double *rhs, double *sol, double *info) {
  std::memcpy(Y.opt, opt, sizeof(double) * 16);
  const int L = (int)Y.lev.size() - 1;
  for (int l = 0; l <= L; l++) {
    const int n = Y.lev[l].N;
    Y.sol[l].assign(n, 0.0); Y.f[l].assign(n, 0.0); Y.aux0[l].assign(n, 0.0); Y.aux1[l].assign(n, 0.0);
  }
  const int solver = (int)opt[0];
  if (solver == 16) return fgmres(Y, rhs, sol, info);
  if (solver == 11) return richardson(Y, rhs, sol, info);
  if (solver == 13) return bicgstab(Y, rhs, sol, info);
  // single preconditioner application: sol = MG(rhs)
  mg_prec(Y, sol, rhs);
  info[0] = 1; info[1] = 0; info[2] = 0;
  return 1;
}

// TFEFunction3D::GetErrors + L2H1Errors (src/FE/FEFunction3D.C:83-300, MainUtilities.C:1649)
static void errors(const Space &S, const double *prm, const double *u, double *err) {
  const Quad Q = gauss_hexa(rule_points(S.order, prm[PMX_PRM_LINEAR] != 0));
  const int nq = (int)Q.w.size(), nb = S.nb;
  std::vector<double> f(nb), fx(nb), fy(nb), fz(nb);
  double e0 = 0, e1 = 0;
  for (int i = 0; i < (int)S.cells.size(); i++) {
    Affine A(S.verts[i].data());
    double I[9];
    A.inv(I);
    const double adet = std::fabs(A.det);
    const int *dof = &S.global[S.begin[i]];
    double l0 = 0, l1 = 0;
    for (int q = 0; q < nq; q++) {
      const double r[3] = {Q.xi[q], Q.eta[q], Q.zeta[q]};
      basis(S.order, r, f.data(), fx.data(), fy.data(), fz.data());
      double X[3];
      A.map(r, X);
      double uh[4] = {0, 0, 0, 0};
      for (int j = 0; j < nb; j++) {
        const double c = u[dof[j]];
        uh[0] += c * f[j];
        uh[1] += c * (I[0] * fx[j] + I[1] * fy[j] + I[2] * fz[j]);
        uh[2] += c * (I[3] * fx[j] + I[4] * fy[j] + I[5] * fz[j]);
        uh[3] += c * (I[6] * fx[j] + I[7] * fy[j] + I[8] * fz[j]);
      }
      double ex[4];
      exact(prm, X[0], X[1], X[2], ex);
      const double w = Q.w[q] * adet;
      l0 += w * (uh[0] - ex[0]) * (uh[0] - ex[0]);
      l1 += w * ((uh[1] - ex[1]) * (uh[1] - ex[1]) + (uh[2] - ex[2]) * (uh[2] - ex[2]) +
                 (uh[3] - ex[3]) * (uh[3] - ex[3]));
    }
    e0 += l0; e1 += l1;
  }
  err[0] = std::sqrt(e0);
  err[1] = std::sqrt(e1);
}

}  // namespace orc

using namespace orc;

// ===========================================================================
// C interface
// ===========================================================================
extern "C" {

void *orc_space_new(int nref, int order) {
  auto tree = std::make_shared<Tree>();
  for (int i = 0; i < nref; i++) tree->refine_leaves();
  auto *S = new Space();
  build_space(*S, *tree, nref, order);
  S->owned = tree;
  return S;
}
static Space &SP(void *h) { return *static_cast<Space *>(h); }
void orc_space_free(void *sp) { delete static_cast<Space *>(sp); }
void orc_space_dims(void *sp, long long *d) {
  Space &S = SP(sp);
  d[0] = S.N; d[1] = S.active; d[2] = (long long)S.cells.size(); d[3] = S.rowptr[S.N]; d[4] = S.nb;
}
void orc_space_arrays(void *sp, int *bi, int *gn, int *rp, int *kc) {
  Space &S = SP(sp);
  if (bi) std::memcpy(bi, S.begin.data(), sizeof(int) * S.begin.size());
  if (gn) std::memcpy(gn, S.global.data(), sizeof(int) * S.global.size());
  if (rp) std::memcpy(rp, S.rowptr.data(), sizeof(int) * S.rowptr.size());
  if (kc) std::memcpy(kc, S.kcol.data(), sizeof(int) * S.kcol.size());
}
void orc_dof_coords(void *sp, double *xyz) {
  Space &S = SP(sp);
  for (int i = 0; i < S.N; i++)
    for (int d = 0; d < 3; d++) xyz[3 * i + d] = S.dofx[i][d];
}
void orc_cell_vertices(void *sp, double *xyz) {
  Space &S = SP(sp);
  for (size_t i = 0; i < S.cells.size(); i++) std::memcpy(xyz + 24 * i, S.verts[i].data(), 24 * sizeof(double));
}
int orc_assemble(void *sp, const double *prm, double *E, double *rhs, double *E2) {
  return assemble(SP(sp), prm, E, rhs, E2);
}
void orc_matvec(void *sp, const double *E, const double *x, double *y) { matvec(SP(sp), E, x, y); }
double orc_defect(void *sp, const double *E, const double *x, const double *b, double *r) {
  return defect(SP(sp), E, x, b, r);
}
void orc_smooth(void *sp, const double *E, int kind, double omega, double *sol, const double *f,
                double *aux) {
  smooth(SP(sp), E, kind, omega, sol, f, aux);
}
void orc_smooth_ordered(void *sp, const double *E, int kind, double omega, double *sol, const double *f,
                        const int *order, int n) {
  smooth_ordered(SP(sp), E, kind, omega, sol, f, order, n);
}

// Transfers between two spaces.  Spaces of one system share a tree; two
// standalone spaces are rebuilt inside one tree (same collections, same
// numbering, since both follow from the refinement depth alone).
struct PairCache {
  const Space *c = nullptr, *f = nullptr;
  std::shared_ptr<Tree> tree;
  Space coarse, fine;
  Transfer tr;
};
static PairCache g_pair;
static void pair_setup(void *c, void *f) {
  const Space &C = SP(c), &F = SP(f);
  if (g_pair.c == &C && g_pair.f == &F) return;
  g_pair.tree = std::make_shared<Tree>();
  for (int i = 0; i < C.nref + 1; i++) g_pair.tree->refine_leaves();
  g_pair.coarse = Space(); g_pair.fine = Space();
  build_space(g_pair.coarse, *g_pair.tree, C.nref, C.order);
  build_space(g_pair.fine, *g_pair.tree, C.nref + 1, F.order);
  g_pair.tr = make_transfer(g_pair.coarse, g_pair.fine, *g_pair.tree);
  g_pair.c = &C; g_pair.f = &F;
}
void orc_prolongate(void *c, void *f, const double *xc, double *xf) {
  pair_setup(c, f);
  prolongate(g_pair.coarse, g_pair.fine, g_pair.tr, xc, xf);
}
void orc_restrict(void *c, void *f, double *rc, const double *rf) {
  pair_setup(c, f);
  restrict_defect(g_pair.coarse, g_pair.fine, g_pair.tr, rc, rf);
}

void *orc_system_new(int nref0, int nlev, int order) {
  auto *Y = new System();
  for (int i = 0; i < nref0 + nlev - 1; i++) Y->tree.refine_leaves();
  Y->lev.resize(nlev);
  for (int l = 0; l < nlev; l++) build_space(Y->lev[l], Y->tree, nref0 + l, order);
  Y->tr.resize(nlev);
  for (int l = 1; l < nlev; l++) Y->tr[l] = make_transfer(Y->lev[l - 1], Y->lev[l], Y->tree);
  Y->E.resize(nlev); Y->E2.resize(nlev); Y->rhs.resize(nlev);
  Y->sol.resize(nlev); Y->f.resize(nlev); Y->aux0.resize(nlev); Y->aux1.resize(nlev);
  for (int l = 0; l < nlev; l++) {
    Y->E[l].assign(Y->lev[l].rowptr.back(), 0.0);
    Y->E2[l].assign(Y->lev[l].rowptr.back(), 0.0);
    Y->rhs[l].assign(Y->lev[l].N, 0.0);
  }
  return Y;
}
void orc_system_free(void *sys) { delete static_cast<System *>(sys); }
void *orc_system_space(void *sys, int level) { return &static_cast<System *>(sys)->lev[level]; }
int orc_system_assemble(void *sys, const double *prm) {
  auto *Y = static_cast<System *>(sys);
  for (size_t l = 0; l < Y->lev.size(); l++) {
    int rc = assemble(Y->lev[l], prm, Y->E[l].data(), Y->rhs[l].data(), Y->E2[l].data());
    if (rc) return rc;
  }
  return 0;
}
double *orc_system_entries(void *sys, int level) { return static_cast<System *>(sys)->E[level].data(); }
double *orc_system_rhs(void *sys, int level) { return static_cast<System *>(sys)->rhs[level].data(); }
int orc_system_solve(void *sys, const double *opt, const double *rhs, double *sol, double *info) {
  return solve(*static_cast<System *>(sys), opt, rhs, sol, info);
}
void orc_system_set_slc(void *sys, int mode) {
  System &Y = *static_cast<System *>(sys);
  Y.slc = mode;
  Y.old_sol.resize(Y.lev.size());
  Y.old_def.resize(Y.lev.size());
}
// row order of the SOR/SSOR sweeps of a level (n = 0: lexicographic again)
void orc_system_set_order(void *sys, int level, const int *order, int n) {
  System &Y = *static_cast<System *>(sys);
  Y.order.resize(Y.lev.size());
  Y.order[level].assign(order, order + n);
}
void orc_errors(void *sp, const double *prm, const double *sol, double *err) { errors(SP(sp), prm, sol, err); }
// which = 0: exact solution values at the nodes (TFEFunction3D::Interpolate)
void orc_interpolate(void *sp, const double *prm, int which, double *sol) {
  Space &S = SP(sp);
  for (int i = 0; i < S.N; i++) {
    double v[4];
    exact(prm, S.dofx[i][0], S.dofx[i][1], S.dofx[i][2], v);
    sol[i] = v[which];
  }
}

// 3DPrograms/TCD3D_ParMooN.C:565-753 with TSystemTCD3D (src/System/SystemTCD3D.C:208-428).
// err_out: per step {time, L2 error, H1-semi error, L2(0,t;L2), L2(0,t;H1), solver iterations}
int orc_tcd3d_run(void *sys, const double *prm_in, const double *td, const double *opt,
                  double *sol, double *err_out, int max_steps) {
  auto *Y = static_cast<System *>(sys);
  const int L = (int)Y->lev.size() - 1;
  std::vector<double> prm(prm_in, prm_in + PMX_PRM_LEN);
  const double tau = td[0], tend = td[1], th1 = td[2], th2 = td[3], th3 = td[4], th4 = td[5];
  // TimeDB's CURRENTTIMESTEPLENGTH and THETA1, read by the SUPG forms
  prm[PMX_PRM_TAU] = tau;
  prm[PMX_PRM_THETA1] = th1;
  const int nlev = L + 1;
  // stiffness form: prm[FORM] (TCD_A when it names none); the SUPG form
  // MatricesAKRhsAssemble_SUPG starts from MatrixMRhsAssemble_SUPG and adds its
  // time-derivative matrix K to M, as TSystemTCD2D::AssembleSystMat does with
  // MatAdd(M, S, 1) (src/System/SystemTCD2D.C:294-297) and RestoreMassMat
  // takes it out again (:336-341)
  int sform = (int)prm_in[PMX_PRM_FORM];
  if (sform != PMX_FORM_GALERKIN && sform != PMX_FORM_SDFEM && sform != PMX_FORM_TCD_SUPG_AK) sform = PMX_FORM_TCD_A;
  const bool supg = sform == PMX_FORM_TCD_SUPG_AK;
  std::vector<std::vector<double>> M(nlev), A(nlev), K(nlev), rhsA(nlev);
  // AssembleMRhs at the start time
  prm[PMX_PRM_FORM] = supg ? PMX_FORM_MASS_SUPG : PMX_FORM_MASS;
  for (int l = 0; l < nlev; l++) {
    M[l].assign(Y->lev[l].rowptr.back(), 0.0);
    A[l].assign(Y->lev[l].rowptr.back(), 0.0);
    if (supg) K[l].assign(Y->lev[l].rowptr.back(), 0.0);
    rhsA[l].assign(Y->lev[l].N, 0.0);
    assemble(Y->lev[l], prm.data(), M[l].data(), rhsA[l].data(), nullptr);
  }
  const Space &S = Y->lev[L];
  const int N = S.N, NA = S.active;
  std::memcpy(sol + NA, rhsA[L].data() + NA, sizeof(double) * (N - NA));
  std::vector<double> oldrhs(rhsA[L]), oldsol(N), B(N), d(N);
  double t = prm[PMX_PRM_TIME];
  double e_old[2] = {0, 0}, e[2], acc[2] = {0, 0};
  orc::errors(S, prm.data(), sol, e_old);
  int m = 0;
  while (t < tend && m < max_steps) {
    m++;
    std::memcpy(oldsol.data(), sol, sizeof(double) * N);
    t += tau;
    prm[PMX_PRM_TIME] = t;
    // AssembleARhs
    prm[PMX_PRM_FORM] = sform;
    for (int l = 0; l < nlev; l++)
      assemble(Y->lev[l], prm.data(), A[l].data(), rhsA[l].data(), supg ? K[l].data() : nullptr);
    const double *rhs = rhsA[L].data();
    std::memcpy(sol + NA, rhs + NA, sizeof(double) * (N - NA));
    // AssembleSystMat
    std::fill(B.begin(), B.end(), 0.0);
    daxpy(NA, tau * th3, oldrhs.data(), B.data());
    daxpy(NA, tau * th4, rhs, B.data());
    const int nact = S.rowptr[NA];
    daxpy(nact, -tau * th2, A[L].data(), M[L].data());
    if (supg) daxpy(nact, 1.0, K[L].data(), M[L].data());
    // MatVectActive
    for (int i = 0; i < NA; i++) {
      double s = 0;
      for (int j = S.rowptr[i]; j < S.rowptr[i + 1]; j++) s += M[L][j] * oldsol[S.kcol[j]];
      d[i] = s;
    }
    daxpy(NA, 1.0, d.data(), B.data());
    std::memcpy(B.data() + NA, rhs + NA, sizeof(double) * (N - NA));
    std::memcpy(sol + NA, rhs + NA, sizeof(double) * (N - NA));
    for (int l = 0; l < nlev; l++) {
      const Space &Sl = Y->lev[l];
      const double fac = (l == L) ? (tau * th2 + tau * th1) : tau * th1;
      daxpy(Sl.rowptr[Sl.active], fac, A[l].data(), M[l].data());
      if (supg && l < L) daxpy(Sl.rowptr[Sl.active], 1.0, K[l].data(), M[l].data());
      Y->E[l] = M[l];
    }
    std::memcpy(oldrhs.data(), rhs, sizeof(double) * N);
    double info[3];
    solve(*Y, opt, B.data(), sol, info);
    // RestoreMassMat
    for (int l = 0; l < nlev; l++) {
      const Space &Sl = Y->lev[l];
      daxpy(Sl.rowptr[Sl.active], -tau * th1, A[l].data(), M[l].data());
      if (supg) daxpy(Sl.rowptr[Sl.active], -1.0, K[l].data(), M[l].data());
    }
    orc::errors(S, prm.data(), sol, e);
    if (m > 1) {
      acc[0] += (e[0] * e[0] + e_old[0] * e_old[0]) * tau / 2.0;
      acc[1] += (e[1] * e[1] + e_old[1] * e_old[1]) * tau / 2.0;
    }
    double *o = err_out + 6 * (m - 1);
    o[0] = t; o[1] = e[0]; o[2] = e[1]; o[3] = std::sqrt(acc[0]); o[4] = std::sqrt(acc[1]); o[5] = info[0];
    e_old[0] = e[0]; e_old[1] = e[1];
  }
  return m;
}

int orc_dense_lu_solve(int n, double *a, double *b) { return dense_lu_solve(n, a, b); }

}  // extern "C"
