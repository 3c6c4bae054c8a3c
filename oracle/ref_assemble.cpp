// Driver of the reference's own assembly (Assemble3D, src/FE/Assemble3D.C,
// with the local forms of src/FE/ConvDiff3D.C and src/FE/TCD3D.C,
// Compute_SDFEM_delta of src/FE/ConvDiff.C), compiled from the reference
// sources by oracle/ref.mk into oracle/_ref/ref_assemble (test
// infrastructure: it pins the oracle's assembled entries and right-hand sides
// against ParMooN itself).  Mesh and space as oracle/ref_numbering.cpp; the
// example callbacks are in oracle/ref_examples.hpp.
//
//   ref_assemble PRM GEO NREF ORDER PRMBLOCK OUT
// PRMBLOCK: the parameter block of include/pmx_prm.h (PMX_PRM_LEN doubles):
// the form, the ParamDB / TimeDB values (SDFEM_TYPE, DELTA0, DELTA1,
// INTERNAL_PROBLEM_LINEAR, INTERNAL_POLYNOMIAL_DEGREE, THETA1,
// CURRENTTIMESTEPLENGTH) and the example.
// OUT (binary, doubles): N, nnz, then Entries[nnz], Rhs[N], and Entries of
// the second matrix [nnz] for the two-matrix SUPG form.
#include "ref_examples.hpp"

int main(int argc, char **argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: ref_assemble PRM GEO NREF ORDER PRMBLOCK OUT\n");
    return 2;
  }
  const int nref = std::atoi(argv[3]), order = std::atoi(argv[4]);
  if (!read_prm_block(argv[5])) return 2;
  new TDatabase();
  new TFEDatabase3D();
  set_assembly_params();
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
  int nsq = 1;
  TDiscreteForm3D *df = make_form((int)g_prm[PMX_PRM_FORM], nsq);
  if (!df) return 3;
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
  FILE *fp = std::fopen(argv[6], "wb");
  if (!fp) return 1;
  const double head[2] = {(double)N, (double)nnz};
  std::fwrite(head, sizeof(double), 2, fp);
  std::fwrite(A->GetEntries(), sizeof(double), nnz, fp);
  std::fwrite(rhs.data(), sizeof(double), N, fp);
  if (nsq == 2) std::fwrite(K->GetEntries(), sizeof(double), nnz, fp);
  std::fclose(fp);
  return 0;
}
