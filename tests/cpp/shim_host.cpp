// Host-only check of the C++ classes in include/pmx_parmoon.hpp: the
// TFESpace3D / TSquareStructure3D arrays equal the oracle's (bit-exact), and a
// failing call surfaces as pmx::Error instead of exit().  With an argument,
// also writes a VTK file there through TFEFunction3D / TOutput3D, as the
// reference's drivers do (3DPrograms/TCD3D_ParMooN.C:465-489): the 4^3 Q2
// space, u = x + 2y + 3z and v = xyz at the DOFs.  Prints "OK".
//
//   shim_host [VTK_PATH]
#include <cstdio>
#include <cstring>
#include <vector>

#include "pmx_parmoon.hpp"

extern "C" {
void *orc_space_new(int nref, int order);
void orc_space_free(void *sp);
void orc_space_dims(void *sp, long long *dims);
void orc_space_arrays(void *sp, int *begin_index, int *global_numbers, int *rowptr, int *kcol);
void orc_dof_coords(void *sp, double *xyz);
}

static int fail(const char *what, int nref, int order) {
  std::printf("FAIL %s nref=%d order=%d\n", what, nref, order);
  return 1;
}

int main(int argc, char **argv) {
  const int cases[][2] = {{2, 1}, {3, 1}, {4, 1}, {1, 2}, {2, 2}, {3, 2}};
  for (const auto &c : cases) {
    const int nref = c[0], order = c[1];
    pmx::TFESpace3D space(nref, order);
    pmx::TSquareStructure3D structure(space);
    void *o = orc_space_new(nref, order);
    long long d[8];
    orc_space_dims(o, d);
    // oracle dims: [0] N [1] active [2] cells [3] nnz [4] n_b
    if (space.GetN_DegreesOfFreedom() != d[0] || space.GetActiveBound() != d[1] || space.GetN_Cells() != d[2] ||
        structure.GetN_Entries() != d[3] || space.GetN_BaseFunct() != d[4])
      return fail("dims", nref, order);
    std::vector<int> bi(d[2] + 1), gn(d[2] * d[4]), rp(d[0] + 1), kc(d[3]);
    orc_space_arrays(o, bi.data(), gn.data(), rp.data(), kc.data());
    orc_space_free(o);
    if (std::memcmp(bi.data(), space.GetBeginIndex(), sizeof(int) * bi.size()) ||
        std::memcmp(gn.data(), space.GetGlobalNumbers(), sizeof(int) * gn.size()))
      return fail("numbering", nref, order);
    if (std::memcmp(rp.data(), structure.GetRowPtr(), sizeof(int) * rp.size()) ||
        std::memcmp(kc.data(), structure.GetKCol(), sizeof(int) * kc.size()))
      return fail("structure", nref, order);
  }
  // error path: an out-of-range space is reported, not fatal
  try {
    pmx::TFESpace3D bad(0, 3);
    return fail("no error for order 3", 0, 3);
  } catch (const pmx::Error &e) {
    if (e.code() != PMX_ERR_ARG) return fail("error code", 0, 3);
  }
  if (argc > 1) {
    pmx::TFESpace3D space(2, 2);
    const int N = space.GetN_DegreesOfFreedom();
    // the DOF coordinates from the oracle's space
    void *o = orc_space_new(2, 2);
    std::vector<double> xyz(3 * (size_t)N);
    orc_dof_coords(o, xyz.data());
    orc_space_free(o);
    std::vector<double> u(N), v(N);
    for (int i = 0; i < N; i++) {
      const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
      u[i] = x + 2 * y + 3 * z;
      v[i] = x * y * z;
    }
    pmx::TFEFunction3D fu(&space, "u", "u", u.data(), N), fv(&space, "v", "v", v.data(), N);
    pmx::TOutput3D out(2, 2, 1, 1, nullptr);
    if (out.AddFEFunction(&fu) != 1 || out.AddFEFunction(&fv) != 2) return fail("AddFEFunction", 2, 2);
    if (out.WriteVtk(argv[1]) != 0) return fail("WriteVtk", 2, 2);
  }
  std::printf("OK\n");
  return 0;
}
