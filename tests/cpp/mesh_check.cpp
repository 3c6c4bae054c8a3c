// Host set-up (parmoon_amd/csrc/pmx_mesh.cpp) under the address and
// undefined-behaviour sanitizers: numbering, CSR pattern and stencil layout of
// the unit cube refined 1..5 times (Q1) and 1..4 times (Q2), with the
// invariants every caller relies on.  Built with g++ (no device code) by
// tests/test_sanitizers.py; exits non-zero on the first violation.
#include <cstdio>
#include <set>
#include <vector>

#include "../../parmoon_amd/csrc/pmx_internal.hpp"

using namespace pmx;

static int fail(const char *what, int nref, int order) {
  std::printf("FAIL %s (nref %d, order %d)\n", what, nref, order);
  return 1;
}

int main() {
  int checked = 0;
  for (int order = 1; order <= 2; order++)
    for (int nref = 1; nref <= (order == 1 ? 5 : 4); nref++) {
      Layout L;
      Numbering Nb;
      build_layout(L, nref, order, kRowsPerBlock);
      build_numbering(Nb, L);
      build_csr(Nb, L);
      const int64_t N = L.N;
      if ((int64_t)Nb.lat2dof.size() != N || (int64_t)Nb.dof2lat.size() != N) return fail("permutation size", nref, order);
      std::vector<char> seen(N, 0);
      for (int64_t p = 0; p < N; p++) {
        const int r = Nb.lat2dof[p];
        if (r < 0 || r >= N || seen[r]) return fail("lat2dof not a permutation", nref, order);
        seen[r] = 1;
        if (Nb.dof2lat[r] != p) return fail("dof2lat not the inverse", nref, order);
      }
      if ((int64_t)Nb.rowptr.size() != N + 1 || Nb.rowptr[0] != 0) return fail("rowptr", nref, order);
      for (int64_t r = 0; r < N; r++) {
        const int b = Nb.rowptr[r], e = Nb.rowptr[r + 1];
        if (e <= b) return fail("empty row", nref, order);
        if (r >= L.nactive && (e - b != 1 || Nb.kcol[b] != r)) return fail("Dirichlet row", nref, order);
        for (int k = b; k < e; k++) {
          if (Nb.kcol[k] < 0 || Nb.kcol[k] >= N) return fail("column out of range", nref, order);
          if (k > b && Nb.kcol[k] <= Nb.kcol[k - 1]) return fail("columns not sorted", nref, order);
        }
      }
      const int nb = order == 1 ? 8 : 27;
      if ((int64_t)Nb.global_numbers.size() != Nb.ncells * nb) return fail("GlobalNumbers size", nref, order);
      std::set<int> used(Nb.global_numbers.begin(), Nb.global_numbers.end());
      if ((int64_t)used.size() != N) return fail("GlobalNumbers do not cover the DOFs", nref, order);
      int64_t rows = 0;
      for (const Segment &s : L.segs) rows += s.count;
      if (rows != L.nactive) return fail("segments do not cover the active rows", nref, order);
      checked++;
    }
  std::printf("mesh_check: %d spaces checked\n", checked);
  return 0;
}
