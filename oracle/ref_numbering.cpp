// Driver of the reference's own mesh / FE-space / structure code, compiled
// from the reference sources by oracle/ref.mk into oracle/_ref/ref_numbering
// (test infrastructure: it pins the DOF numbering and CSR pattern of the
// oracle and of libpmx against ParMooN itself).  It follows the set-up of
// 3DPrograms/CD3D_ParMooN.C:85-300: TDatabase, TFEDatabase3D, TDomain::Init
// with data/Wuerfel.PRM / .GEO, RegRefineAll, GetCollection(It_Finest, 0),
// TFESpace3D(coll, ..., BoundCondition, ORDER) with every boundary Dirichlet
// (Examples/CD_3D/Laplace.h:21-25), TSquareStructure3D(space) + Sort()
// (SystemCD3D.C:113-119).
//
//   ref_numbering PRM GEO NREF ORDER OUT
// writes to OUT (binary, int32): N, active, n_cells, n_base, nnz, then
// BeginIndex[n_cells + 1], GlobalNumbers[n_cells * n_base], RowPtr[N + 1],
// KCol[nnz].
#include <cstdio>
#include <cstring>
#include <vector>

#include <Database.h>
#include <Domain.h>
#include <FEDatabase3D.h>
#include <FESpace3D.h>
#include <SquareStructure3D.h>

static void AllDirichlet(int, double, double, double, BoundCond &cond) { cond = DIRICHLET; }

int main(int argc, char **argv) {
  if (argc < 6) {
    std::fprintf(stderr, "usage: ref_numbering PRM GEO NREF ORDER OUT\n");
    return 2;
  }
  const int nref = std::atoi(argv[3]), order = std::atoi(argv[4]);
  TDatabase *db = new TDatabase();
  TFEDatabase3D *fedb = new TFEDatabase3D();
  (void)db;
  (void)fedb;
  TDomain *domain = new TDomain();
  domain->Init(argv[1], argv[2]);
  for (int i = 0; i < nref; i++) domain->RegRefineAll();
  TCollection *coll = domain->GetCollection(It_Finest, 0);
  char name[] = "name", desc[] = "description";
  TFESpace3D *space = new TFESpace3D(coll, name, desc, AllDirichlet, order);
  TSquareStructure3D *st = new TSquareStructure3D(space);
  st->Sort();
  const int N = space->GetN_DegreesOfFreedom(), active = space->GetActiveBound();
  const int ncells = coll->GetN_Cells();
  const int *bi = space->GetBeginIndex(), *gn = space->GetGlobalNumbers();
  const int nbase = bi[1] - bi[0];
  const int *rp = st->GetRowPtr(), *kc = st->GetKCol();
  const int nnz = rp[N];
  FILE *fp = std::fopen(argv[5], "wb");
  if (!fp) return 1;
  const int head[5] = {N, active, ncells, nbase, nnz};
  std::fwrite(head, sizeof(int), 5, fp);
  std::fwrite(bi, sizeof(int), ncells + 1, fp);
  std::fwrite(gn, sizeof(int), (size_t)ncells * nbase, fp);
  std::fwrite(rp, sizeof(int), N + 1, fp);
  std::fwrite(kc, sizeof(int), nnz, fp);
  std::fclose(fp);
  std::printf("N %d active %d cells %d nnz %d\n", N, active, ncells, nnz);
  return 0;
}
