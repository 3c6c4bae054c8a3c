// Driver of the reference's own assembly (Assemble3D, src/FE/Assemble3D.C,
// with the local forms of src/FE/ConvDiff3D.C and src/FE/TCD3D.C,
// Compute_SDFEM_delta of src/FE/ConvDiff.C), compiled from the reference
// sources by oracle/ref.mk into oracle/_ref/ref_assemble (test
// infrastructure: it pins the oracle's assembled entries and right-hand sides
// against ParMooN itself).  Mesh and space as oracle/ref_numbering.cpp.
//
//   ref_assemble PRM GEO NREF ORDER PRMBLOCK OUT
// PRMBLOCK: the parameter block of include/pmx_prm.h (PMX_PRM_LEN doubles):
// the form, the ParamDB / TimeDB values (SDFEM_TYPE, DELTA0, DELTA1,
// INTERNAL_PROBLEM_LINEAR, INTERNAL_POLYNOMIAL_DEGREE, THETA1,
// CURRENTTIMESTEPLENGTH) and the example.  The examples' coefficient and
// boundary callbacks are written here, as an example header of the
// reference would: PMX_EX_SINE_CD (constant eps, b, c; manufactured
// u = sin sin sin), PMX_EX_ROTATING (b = (-(y-1/2), x-1/2, 0.2), f = 0, g = 0)
// and PMX_EX_CURL_FOURIER (b = scale * curl of the Fourier potential, f = 1,
// g = 0).
// OUT (binary, doubles): N, nnz, then Entries[nnz], Rhs[N], and Entries of
// the second matrix [nnz] for the two-matrix SUPG form.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include <Assemble3D.h>
#include <AuxParam3D.h>
#include <Database.h>
#include <DiscreteForm3D.h>
#include <Domain.h>
#include <FEDatabase3D.h>
#include <FESpace3D.h>
#include <SquareMatrix3D.h>
#include <SquareStructure3D.h>

#include "../include/pmx_prm.h"

// the local forms and their term tables (src/FE/ConvDiff3D.C, TCD3D.C; the
// TCD3D tables are defined in DiscreteForm3D.C, which includes TCD3D.h)
void BilinearAssemble(double, double *, double, double **, int *, double ***, double **);
void BilinearAssemble_SD(double, double *, double, double **, int *, double ***, double **);
void MatrixMRhsAssemble(double, double *, double *, double, double **, int *, double ***, double **);
void MatrixARhsAssemble(double, double *, double *, double, double **, int *, double ***, double **);
void MatrixMRhsAssemble_SUPG(double, double *, double *, double, double **, int *, double ***, double **);
void MatricesAKRhsAssemble_SUPG(double, double *, double *, double, double **, int *, double ***, double **);
void RhsAssemble(double, double *, double *, double, double **, int *, double ***, double **);
void RhsAssemble_SUPG(double, double *, double *, double, double **, int *, double ***, double **);
#define PMX_TABLE(S)                                                                                       \
  extern int N_Terms_##S;                                                                                 \
  extern MultiIndex3D Derivatives_##S[];                                                                  \
  extern int SpacesNumbers_##S[];                                                                         \
  extern int N_Matrices_##S;                                                                              \
  extern int N_Rhs_##S;                                                                                   \
  extern int RhsSpace_##S[];
PMX_TABLE(MatrixMRhs)
PMX_TABLE(MatrixARhs)
PMX_TABLE(MatrixMRhs_SUPG)
PMX_TABLE(MatricesAKRhs_SUPG)
PMX_TABLE(Rhs)
PMX_TABLE(Rhs_SUPG)
extern int RowSpace_MatrixMRhs[], ColumnSpace_MatrixMRhs[], RowSpace_MatrixARhs[], ColumnSpace_MatrixARhs[];
extern int RowSpace_MatrixMRhs_SUPG[], ColumnSpace_MatrixMRhs_SUPG[];
extern int RowSpace_MatricesAKRhs_SUPG[], ColumnSpace_MatricesAKRhs_SUPG[];
extern int *RowSpace_Rhs, *ColumnSpace_Rhs, *RowSpace_Rhs_SUPG, *ColumnSpace_Rhs_SUPG;

static std::vector<double> g_prm;
static const double kPi = 3.14159265358979323846;

static void curl_b(double x, double y, double z, double b[3]) {
  const double *F = g_prm.data() + PMX_PRM_FOURIER;
  double dA[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int c = 0; c < 3; c++)
    for (int m = 0; m < 8; m++) {
      const double *t = F + (c * 8 + m) * 5;
      const double w = t[0] * std::cos(2 * kPi * (t[1] * x + t[2] * y + t[3] * z) + t[4]) * 2 * kPi;
      for (int j = 0; j < 3; j++) dA[c][j] += w * t[1 + j];
    }
  const double s = g_prm[PMX_PRM_BSCALE];
  b[0] = s * (dA[2][1] - dA[1][2]);
  b[1] = s * (dA[0][2] - dA[2][0]);
  b[2] = s * (dA[1][0] - dA[0][1]);
}

static void Coeffs(int n, double *X, double *Y, double *Z, double **, double **coeffs) {
  const int ex = (int)g_prm[PMX_PRM_EXAMPLE];
  const double eps = g_prm[PMX_PRM_EPS];
  for (int i = 0; i < n; i++) {
    double *c = coeffs[i];
    const double x = X[i], y = Y[i], z = Z[i];
    c[0] = eps;
    if (ex == PMX_EX_SINE_CD) {
      const double b0 = g_prm[PMX_PRM_B0], b1 = g_prm[PMX_PRM_B1], b2 = g_prm[PMX_PRM_B2], r = g_prm[PMX_PRM_C];
      const double u = std::sin(kPi * x) * std::sin(kPi * y) * std::sin(kPi * z);
      c[1] = b0; c[2] = b1; c[3] = b2; c[4] = r;
      c[5] = 3 * kPi * kPi * eps * u +
             kPi * (b0 * std::cos(kPi * x) * std::sin(kPi * y) * std::sin(kPi * z) +
                    b1 * std::sin(kPi * x) * std::cos(kPi * y) * std::sin(kPi * z) +
                    b2 * std::sin(kPi * x) * std::sin(kPi * y) * std::cos(kPi * z)) +
             r * u;
    } else if (ex == PMX_EX_ROTATING) {
      c[1] = -(y - 0.5); c[2] = x - 0.5; c[3] = 0.2; c[4] = 0; c[5] = 0;
    } else {  // PMX_EX_CURL_FOURIER
      double b[3];
      curl_b(x, y, z, b);
      c[1] = b[0]; c[2] = b[1]; c[3] = b[2]; c[4] = 0; c[5] = 1.0;
    }
  }
}
static void AllDirichlet(int, double, double, double, BoundCond &cond) { cond = DIRICHLET; }
static void BoundValue(int, double x, double y, double z, double &value) {
  value = (int)g_prm[PMX_PRM_EXAMPLE] == PMX_EX_SINE_CD ? std::sin(kPi * x) * std::sin(kPi * y) * std::sin(kPi * z)
                                                        : 0.0;
}

int main(int argc, char **argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: ref_assemble PRM GEO NREF ORDER PRMBLOCK OUT\n");
    return 2;
  }
  const int nref = std::atoi(argv[3]), order = std::atoi(argv[4]);
  g_prm.assign(PMX_PRM_LEN, 0.0);
  FILE *fp = std::fopen(argv[5], "rb");
  if (!fp || std::fread(g_prm.data(), sizeof(double), PMX_PRM_LEN, fp) != (size_t)PMX_PRM_LEN) return 2;
  std::fclose(fp);
  new TDatabase();
  new TFEDatabase3D();
  TDatabase::ParamDB->SDFEM_TYPE = (int)g_prm[PMX_PRM_SDFEM_TYPE];
  TDatabase::ParamDB->DELTA0 = g_prm[PMX_PRM_DELTA0];
  TDatabase::ParamDB->DELTA1 = g_prm[PMX_PRM_DELTA1];
  TDatabase::ParamDB->INTERNAL_PROBLEM_LINEAR = (int)g_prm[PMX_PRM_LINEAR];
  TDatabase::ParamDB->INTERNAL_POLYNOMIAL_DEGREE = (int)g_prm[PMX_PRM_POLYDEG];
  TDatabase::TimeDB->THETA1 = g_prm[PMX_PRM_THETA1];
  TDatabase::TimeDB->CURRENTTIMESTEPLENGTH = g_prm[PMX_PRM_TAU];
  TDomain *domain = new TDomain();
  domain->Init(argv[1], argv[2]);
  for (int i = 0; i < nref; i++) domain->RegRefineAll();
  TCollection *coll = domain->GetCollection(It_Finest, 0);
  char name[] = "name", desc[] = "description";
  TFESpace3D *space = new TFESpace3D(coll, name, desc, AllDirichlet, order);
  TSquareStructure3D *st = new TSquareStructure3D(space);
  st->Sort();
  const int N = space->GetN_DegreesOfFreedom();
  TSquareMatrix3D *A = new TSquareMatrix3D(st), *K = new TSquareMatrix3D(st);
  std::vector<double> rhs(N, 0.0);
  const int form = (int)g_prm[PMX_PRM_FORM];
  TDiscreteForm3D *df = nullptr;
  int nsq = 1;
#define PMX_DF(S, F)                                                                                          \
  new TDiscreteForm3D(name, desc, N_Terms_##S, Derivatives_##S, SpacesNumbers_##S, N_Matrices_##S, N_Rhs_##S, \
                      RowSpace_##S, ColumnSpace_##S, RhsSpace_##S, F, Coeffs, nullptr)
  static MultiIndex3D der4[4] = {D100, D010, D001, D000};
  static int sp4[4] = {0, 0, 0, 0}, row1[1] = {0}, col1[1] = {0}, rhs1[1] = {0};
  switch (form) {
    case PMX_FORM_GALERKIN:
      df = new TDiscreteForm3D(name, desc, 4, der4, sp4, 1, 1, row1, col1, rhs1, BilinearAssemble, Coeffs, nullptr);
      break;
    case PMX_FORM_SDFEM:
      df = new TDiscreteForm3D(name, desc, 4, der4, sp4, 1, 1, row1, col1, rhs1, BilinearAssemble_SD, Coeffs,
                               nullptr);
      break;
    case PMX_FORM_MASS: df = PMX_DF(MatrixMRhs, MatrixMRhsAssemble); break;
    case PMX_FORM_TCD_A: df = PMX_DF(MatrixARhs, MatrixARhsAssemble); break;
    case PMX_FORM_MASS_SUPG: df = PMX_DF(MatrixMRhs_SUPG, MatrixMRhsAssemble_SUPG); break;
    case PMX_FORM_TCD_SUPG_AK: df = PMX_DF(MatricesAKRhs_SUPG, MatricesAKRhsAssemble_SUPG); nsq = 2; break;
    case PMX_FORM_RHS: df = PMX_DF(Rhs, RhsAssemble); nsq = 0; break;
    case PMX_FORM_RHS_SUPG: df = PMX_DF(Rhs_SUPG, RhsAssemble_SUPG); nsq = 0; break;
    default: return 3;
  }
  TFESpace3D *fesp[1] = {space}, *ferhs[1] = {space};
  TSquareMatrix3D *sq[2] = {A, K};
  double *rhsp[1] = {rhs.data()};
  BoundCondFunct3D *bc[1] = {AllDirichlet};
  BoundValueFunct3D *bv[1] = {BoundValue};
  TAuxParam3D *aux = new TAuxParam3D(1, 0, 0, 0, fesp, nullptr, nullptr, nullptr, nullptr, 0, nullptr);
  A->Reset();
  K->Reset();
  Assemble3D(1, fesp, nsq, nsq ? sq : nullptr, 0, nullptr, 1, rhsp, ferhs, df, bc, bv, aux);
  const int nnz = st->GetRowPtr()[N];
  fp = std::fopen(argv[6], "wb");
  if (!fp) return 1;
  const double head[2] = {(double)N, (double)nnz};
  std::fwrite(head, sizeof(double), 2, fp);
  std::fwrite(A->GetEntries(), sizeof(double), nnz, fp);
  std::fwrite(rhs.data(), sizeof(double), N, fp);
  if (nsq == 2) std::fwrite(K->GetEntries(), sizeof(double), nnz, fp);
  std::fclose(fp);
  return 0;
}
