// GPU run of the C++ classes in include/pmx_parmoon.hpp, written like the
// reference's drivers (3DPrograms/CD3D_ParMooN.C, TCD3D_ParMooN.C):
//   shim_solve <dir>
// reads <dir>/prm.bin, <dir>/opt.bin (CD3D) and <dir>/prm_tcd.bin,
// <dir>/opt_tcd.bin, <dir>/td.bin, <dir>/u0.bin (TCD3D), and writes
// <dir>/sol.bin, <dir>/rhs.bin, <dir>/entries.bin and <dir>/sol_tcd.bin.
#include <cstdio>
#include <string>
#include <vector>

#include "pmx_parmoon.hpp"

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
