// GPU run of the C++ classes in include/pmx_parmoon.hpp, written like the
// reference's drivers (3DPrograms/CD3D_ParMooN.C, TCD3D_ParMooN.C):
//   shim_solve <dir>
// reads <dir>/prm.bin, <dir>/opt.bin (CD3D) and <dir>/prm_tcd.bin,
// <dir>/opt_tcd.bin, <dir>/td.bin, <dir>/u0.bin (TCD3D), and writes
// <dir>/sol.bin, <dir>/rhs.bin, <dir>/entries.bin, <dir>/sol_tcd.bin, the
// entries / rhs of Laplace.h's callbacks through TAssembleMat3D and of the
// built-in functor (<dir>/laplace_{cb,fn}_{entries,rhs}.bin), and the
// solution with the caller's MG level matrices (reads
// <dir>/mg_entries_{0,1,2}.bin and <dir>/mg_rhs.bin, writes <dir>/mg_sol.bin).
#include <cstdio>
#include <string>
#include <vector>

#include <cmath>

#include "pmx_parmoon.hpp"

// Examples/CD_3D/Laplace.h's callbacks with the reference's signatures (the
// BoundCond& / double& of BoundCondFunct3D / BoundValueFunct3D as pointers);
// eps = 1 / RE_NR with RE_NR = 50 here
static const double kPi = 3.14159265358979323846;
static void BoundCondition(int, double, double, double, int *cond) { *cond = 0; /* DIRICHLET */ }
static void BoundValue(int, double x, double y, double z, double *value) {
  *value = std::sin(kPi * x) * std::sin(kPi * y) * std::sin(kPi * z);
}
static void BilinearCoeffs(int n_points, double *x, double *y, double *z, double **, double **coeffs) {
  const double eps = 1.0 / 50.0;
  for (int i = 0; i < n_points; i++) {
    double *coeff = coeffs[i];
    coeff[0] = eps;
    coeff[1] = 0;
    coeff[2] = 0;
    coeff[3] = 0;
    coeff[4] = 0;
    coeff[5] = (3 * kPi * kPi * eps) * std::sin(kPi * x[i]) * std::sin(kPi * y[i]) * std::sin(kPi * z[i]);
  }
}

static std::vector<double> load(const std::string &f, size_t n) {
  std::vector<double> v(n);
  FILE *fp = std::fopen(f.c_str(), "rb");
  if (!fp || std::fread(v.data(), sizeof(double), n, fp) != n) throw std::runtime_error("cannot read " + f);
  std::fclose(fp);
  return v;
}
static void save(const std::string &f, const double *v, size_t n) {
  FILE *fp = std::fopen(f.c_str(), "wb");
  if (!fp || std::fwrite(v, sizeof(double), n, fp) != n) throw std::runtime_error("cannot write " + f);
  std::fclose(fp);
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  try {
    pmx::Context ctx(0);
    // CD3D: 3 levels 2^3..8^3 Q1
    {
      pmx::TFESpace3D fine(3, 1);
      pmx::TSquareStructure3D structure(fine);
      const int N = fine.GetN_DegreesOfFreedom();
      std::vector<double> sol(N, 0.0), rhs(N, 0.0);
      pmx::TSystemCD3D system(ctx, 3, 1, 1, sol.data(), rhs.data());
      const std::vector<double> prm = load(dir + "/prm.bin", PMX_PRM_LEN);
      const std::vector<double> opt = load(dir + "/opt.bin", 24);
      system.Init(prm.data());
      system.Assemble();
      double info[8];
      const int its = system.Solve(opt.data(), info);
      std::vector<double> entries(structure.GetN_Entries());
      system.GetEntries(-1, entries.data());
      save(dir + "/sol.bin", sol.data(), N);
      save(dir + "/rhs.bin", rhs.data(), N);
      save(dir + "/entries.bin", entries.data(), entries.size());
      std::printf("cd3d iterations %d residual %.3e\n", its, info[1]);
    }
    // TAssembleMat3D with Laplace.h's callbacks against the built-in Laplace
    // functor (3 levels 2^2..2^4 Q2, Galerkin): both sets of entries and rhs
    {
      pmx::TFESpace3D fine(4, 2);
      pmx::TSquareStructure3D structure(fine);
      const int N = fine.GetN_DegreesOfFreedom();
      std::vector<double> sol(N, 0.0), rhs_cb(N), rhs_fn(N), e_cb(structure.GetN_Entries()), e_fn(e_cb.size());
      pmx::TSystemCD3D system(ctx, 3, 2, 2, sol.data(), rhs_fn.data());
      std::vector<double> prm(PMX_PRM_LEN, 0.0);
      prm[PMX_PRM_EXAMPLE] = PMX_EX_LAPLACE;
      prm[PMX_PRM_EPS] = 1.0 / 50.0;
      prm[PMX_PRM_FORM] = PMX_FORM_GALERKIN;
      prm[PMX_PRM_POLYDEG] = 1;
      system.Init(prm.data());
      system.Assemble();
      system.GetEntries(-1, e_fn.data());
      pmx::TAssembleMat3D am(system.handle(), -1, prm.data(), BilinearCoeffs, BoundCondition, BoundValue,
                             e_cb.data(), rhs_cb.data());
      am.Init();
      am.Reset();
      am.Assemble3D();
      am.DeAllocate();
      save(dir + "/laplace_cb_entries.bin", e_cb.data(), e_cb.size());
      save(dir + "/laplace_fn_entries.bin", e_fn.data(), e_fn.size());
      save(dir + "/laplace_cb_rhs.bin", rhs_cb.data(), N);
      save(dir + "/laplace_fn_rhs.bin", rhs_fn.data(), N);
      std::printf("laplace callbacks assembled\n");
    }
    // the caller's matrices as the MG levels (TMGLevel3D + TMultiGrid3D::
    // AddLevel) and Iterate on the caller's finest matrix: 3 levels
    // 2^1..2^3 Q1 whose system was assembled with another form; the
    // entries of every level and the rhs come from <dir>/mg_entries_<l>.bin,
    // <dir>/mg_rhs.bin
    {
      pmx::TFESpace3D fine(3, 1);
      const int N = fine.GetN_DegreesOfFreedom();
      std::vector<double> sol(N, 0.0), rhs0(N, 0.0);
      pmx::TSystemCD3D system(ctx, 3, 1, 1, sol.data(), rhs0.data());
      std::vector<double> prm(PMX_PRM_LEN, 0.0);
      prm[PMX_PRM_EXAMPLE] = PMX_EX_LAPLACE;
      prm[PMX_PRM_EPS] = 1.0;
      prm[PMX_PRM_FORM] = PMX_FORM_GALERKIN;
      prm[PMX_PRM_POLYDEG] = 1;
      system.Init(prm.data());
      system.Assemble();
      const std::vector<double> opt = load(dir + "/opt.bin", 24);
      std::vector<double> mgp = {opt[8], opt[8]};
      pmx::TMultiGrid3D mg(system.handle(), 1, 2, mgp.data());
      std::vector<pmx::TFESpace3D *> spaces;
      std::vector<pmx::TSquareStructure3D *> structures;
      std::vector<pmx::TSquareMatrix3D *> matrices;
      std::vector<pmx::TMGLevel3D *> levels;
      for (int l = 0; l < 3; l++) {
        spaces.push_back(new pmx::TFESpace3D(1 + l, 1));
        structures.push_back(new pmx::TSquareStructure3D(*spaces.back()));
        matrices.push_back(new pmx::TSquareMatrix3D(*structures.back()));
        const std::vector<double> e = load(dir + "/mg_entries_" + std::to_string(l) + ".bin",
                                           (size_t)matrices.back()->GetN_Entries());
        std::copy(e.begin(), e.end(), matrices.back()->GetEntries());
        levels.push_back(new pmx::TMGLevel3D(l, matrices.back(), nullptr, nullptr, 2, nullptr));
        if (mg.AddLevel(levels.back()) != l + 1) throw pmx::Error(PMX_ERR_STATE, "AddLevel count");
      }
      std::vector<double> rhs = load(dir + "/mg_rhs.bin", N);
      for (int i = fine.GetActiveBound(); i < N; i++) sol[i] = rhs[i];
      pmx::TItMethod itm(mg, opt.data());
      pmx::TSquareMatrix3D *sq[1] = {matrices.back()};
      const int its = itm.Iterate(sq, nullptr, sol.data(), rhs.data());
      save(dir + "/mg_sol.bin", sol.data(), N);
      std::printf("caller mg levels: %d iterations\n", its);
      // a matrix of another pattern is rejected
      bool rejected = false;
      try {
        pmx::TMGLevel3D wrong(2, matrices[1], nullptr, nullptr, 2, nullptr);
        mg.AddLevel(&wrong);
      } catch (const pmx::Error &e) {
        rejected = e.code() == PMX_ERR_ARG;
      }
      if (!rejected) throw pmx::Error(PMX_ERR_STATE, "a matrix of another level's pattern was accepted");
      for (auto *p : levels) delete p;
      for (auto *p : matrices) delete p;
      for (auto *p : structures) delete p;
      for (auto *p : spaces) delete p;
    }
    // TCD3D: 4 levels 2^1..2^4 Q1, Crank-Nicolson
    {
      pmx::TFESpace3D fine(4, 1);
      const int N = fine.GetN_DegreesOfFreedom();
      const std::vector<double> td = load(dir + "/td.bin", 6);  // tau, steps, theta1..4
      std::vector<double> sol = load(dir + "/u0.bin", N);
      pmx::TSystemTCD3D system(ctx, 4, 1, 1, sol.data());
      const std::vector<double> prm = load(dir + "/prm_tcd.bin", PMX_PRM_LEN);
      const std::vector<double> opt = load(dir + "/opt_tcd.bin", 24);
      system.Init(prm.data(), opt.data());
      system.AssembleMRhs();
      double t = prm[PMX_PRM_TIME];
      for (int m = 0; m < (int)td[1]; m++) {
        t += td[0];
        system.AssembleARhs(t, td[0], td[2], td[3], td[4], td[5]);
        system.AssembleSystMat();
        system.Solve();
        system.RestoreMassMat();
      }
      save(dir + "/sol_tcd.bin", sol.data(), N);
      std::printf("tcd3d steps %d\n", (int)td[1]);
    }
  } catch (const pmx::Error &e) {
    std::printf("pmx error %d: %s\n", e.code(), e.what());
    return 1;
  }
  std::printf("OK\n");
  return 0;
}
