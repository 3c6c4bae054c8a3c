// Shared part of the drivers of the reference's own code (oracle/ref_assemble.cpp,
// oracle/ref_solve.cpp; test infrastructure): the example callbacks written
// as an example header of the reference would (PMX_EX_SINE_CD: constant eps,
// b, c and the manufactured u = sin sin sin; PMX_EX_ROTATING:
// b = (-(y-1/2), x-1/2, 0.2), f = 0, g = 0; PMX_EX_CURL_FOURIER: b = scale *
// curl of the Fourier potential, f = 1, g = 0), the ParamDB / TimeDB fields
// of the parameter block, and the TDiscreteForm3D of a form id.
#pragma once
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

static std::vector<double> g_prm;  // the parameter block (PMX_PRM_LEN doubles)
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


// reads the parameter block file into g_prm; false on a short file
static bool read_prm_block(const char *path) {
  g_prm.assign(PMX_PRM_LEN, 0.0);
  FILE *fp = std::fopen(path, "rb");
  if (!fp) return false;
  const bool ok = std::fread(g_prm.data(), sizeof(double), PMX_PRM_LEN, fp) == (size_t)PMX_PRM_LEN;
  std::fclose(fp);
  return ok;
}

// the ParamDB / TimeDB fields the assembly reads
static void set_assembly_params() {
  TDatabase::ParamDB->SDFEM_TYPE = (int)g_prm[PMX_PRM_SDFEM_TYPE];
  TDatabase::ParamDB->DELTA0 = g_prm[PMX_PRM_DELTA0];
  TDatabase::ParamDB->DELTA1 = g_prm[PMX_PRM_DELTA1];
  TDatabase::ParamDB->INTERNAL_PROBLEM_LINEAR = (int)g_prm[PMX_PRM_LINEAR];
  TDatabase::ParamDB->INTERNAL_POLYNOMIAL_DEGREE = (int)g_prm[PMX_PRM_POLYDEG];
  TDatabase::TimeDB->THETA1 = g_prm[PMX_PRM_THETA1];
  TDatabase::TimeDB->CURRENTTIMESTEPLENGTH = g_prm[PMX_PRM_TAU];
  TDatabase::ParamDB->CELL_MEASURE = (int)g_prm[PMX_PRM_CELL_MEASURE];
}

// the discrete form of the block's form id; nsq = its number of matrices
static TDiscreteForm3D *make_form(int form, int &nsq) {
  static char name[] = "name", desc[] = "description";
#define PMX_DF(S, F)                                                                                          \
  new TDiscreteForm3D(name, desc, N_Terms_##S, Derivatives_##S, SpacesNumbers_##S, N_Matrices_##S, N_Rhs_##S, \
                      RowSpace_##S, ColumnSpace_##S, RhsSpace_##S, F, Coeffs, nullptr)
  static MultiIndex3D der4[4] = {D100, D010, D001, D000};
  static int sp4[4] = {0, 0, 0, 0}, row1[1] = {0}, col1[1] = {0}, rhs1[1] = {0};
  nsq = 1;
  switch (form) {
    case PMX_FORM_GALERKIN:
      return new TDiscreteForm3D(name, desc, 4, der4, sp4, 1, 1, row1, col1, rhs1, BilinearAssemble, Coeffs, nullptr);
    case PMX_FORM_SDFEM:
      return new TDiscreteForm3D(name, desc, 4, der4, sp4, 1, 1, row1, col1, rhs1, BilinearAssemble_SD, Coeffs,
                                 nullptr);
    case PMX_FORM_MASS: return PMX_DF(MatrixMRhs, MatrixMRhsAssemble);
    case PMX_FORM_TCD_A: return PMX_DF(MatrixARhs, MatrixARhsAssemble);
    case PMX_FORM_MASS_SUPG: return PMX_DF(MatrixMRhs_SUPG, MatrixMRhsAssemble_SUPG);
    case PMX_FORM_TCD_SUPG_AK: nsq = 2; return PMX_DF(MatricesAKRhs_SUPG, MatricesAKRhsAssemble_SUPG);
    case PMX_FORM_RHS: nsq = 0; return PMX_DF(Rhs, RhsAssemble);
    case PMX_FORM_RHS_SUPG: nsq = 0; return PMX_DF(Rhs_SUPG, RhsAssemble_SUPG);
    default: return nullptr;
  }
#undef PMX_DF
}
