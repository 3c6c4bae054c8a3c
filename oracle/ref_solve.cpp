// Driver of the reference's own GMG-preconditioned solve, compiled from the
// reference sources by oracle/ref.mk into oracle/_ref/ref_solve (test
// infrastructure: it pins the oracle's multigrid cycle, smoothers, grid
// transfers, coarse solve and Krylov methods against ParMooN itself).
//
// The levels, matrices and solver objects are set up as TSystemCD3D does
// (src/System/SystemCD3D.C:173-225 the solver objects, :290-335 Init with
// TMGLevel3D + AddLevel, :340-360 Assemble, :366-395 Solve): the unit cube
// refined nref0 .. nref0+nlev-1 times (3DPrograms/CD3D_ParMooN.C:285-300),
// one space, structure, matrix and rhs per level assembled with the form of
// the parameter block (Assemble3D, the example callbacks of
// oracle/ref_examples.hpp), the Dirichlet values copied into the solution,
// then TMultiGrid3D with one TMGLevel3D per level, TMultiGridScaIte as the
// preconditioner and TFgmresIte / TFixedPointIte as the outer method, and
// Iterate on the finest level from a zero start.
//
//   ref_solve PRM GEO NREF0 NLEV ORDER PRMBLOCK OPTBLOCK OUT
// OPTBLOCK: the solver option vector of pmx.h (16 or more doubles): entries
// 0-13 (solver, restart, maxit, res_norm_min, red_factor, smoother, pre,
// post, omega, coarse smoother, coarse maxit, coarse reduction, cycle, GMG
// damping) are written into the ParamDB fields they stand for; a positive
// entry 16 (relative target) sets SC_LIN_RES_NORM_MIN_SCALAR to it times the
// initial defect norm.
// OUT (binary, doubles): N, the finest solution [N], the finest rhs [N] (the
// Krylov method's copy; the MG levels overwrite their own rhs arrays), then
// iterations, and the seconds of the set-up (refinement, spaces,
// structures), the assembly of all levels and the solve.
#include <chrono>

#include <FgmresIte.h>
#include <FixedPointIte.h>
#include <LinAlg.h>
#include <MGLevel3D.h>
#include <MultiGrid3D.h>
#include <MultiGridScaIte.h>

#include "ref_examples.hpp"

int main(int argc, char **argv) {
  if (argc < 9) {
    std::fprintf(stderr, "usage: ref_solve PRM GEO NREF0 NLEV ORDER PRMBLOCK OPTBLOCK OUT\n");
    return 2;
  }
  const int nref0 = std::atoi(argv[3]), nlev = std::atoi(argv[4]), order = std::atoi(argv[5]);
  if (!read_prm_block(argv[6])) return 2;
  double opt[24] = {0};
  FILE *fp = std::fopen(argv[7], "rb");
  if (!fp || std::fread(opt, sizeof(double), 24, fp) < 16) return 2;
  std::fclose(fp);
  using clock = std::chrono::steady_clock;
  auto secs = [](clock::time_point a, clock::time_point b) { return std::chrono::duration<double>(b - a).count(); };
  const auto t0 = clock::now();
  double t_asm = 0;
  new TDatabase();
  new TFEDatabase3D();
  set_assembly_params();
  TParamDB *P = TDatabase::ParamDB;
  P->SC_SOLVER_SCALAR = (int)opt[0];
  P->SC_PRECONDITIONER_SCALAR = 5;
  P->SC_GMRES_RESTART = (int)opt[1];
  P->SC_LIN_MAXIT_SCALAR = (int)opt[2];
  P->SC_LIN_RES_NORM_MIN_SCALAR = opt[3];
  P->SC_LIN_RED_FACTOR_SCALAR = opt[4];
  P->SC_SMOOTHER_SCALAR = (int)opt[5];
  P->SC_PRE_SMOOTH_SCALAR = (int)opt[6];
  P->SC_POST_SMOOTH_SCALAR = (int)opt[7];
  P->SC_SMOOTH_DAMP_FACTOR_SCALAR = opt[8];
  P->SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR = opt[8];
  P->SC_SMOOTH_DAMP_FACTOR_COARSE_SCALAR = opt[8];
  P->SC_COARSE_SMOOTHER_SCALAR = (int)opt[9];
  P->SC_COARSE_MAXIT_SCALAR = (int)opt[10];
  P->SC_COARSE_RED_FACTOR_SCALAR = opt[11];
  P->SC_MG_CYCLE_SCALAR = (int)opt[12];
  P->SC_GMG_DAMP_FACTOR_SCALAR = opt[13];
  P->SC_GMG_DAMP_FACTOR_FINE_SCALAR = opt[13];
  P->SC_STEP_LENGTH_CONTROL_ALL_SCALAR = 0;
  P->SC_STEP_LENGTH_CONTROL_FINE_SCALAR = 0;
  P->SC_FLEXIBLE_KRYLOV_SPACE_SOLVER = 1;
  P->SC_MINIT = 0;
  P->SC_DIV_FACTOR = 1e10;
  P->SC_VERBOSE = 2;
  P->SC_MG_TYPE_SCALAR = 0;

  TDomain *domain = new TDomain();
  domain->Init(argv[1], argv[2]);
  for (int i = 0; i < nref0; i++) domain->RegRefineAll();
  char name[] = "name", desc[] = "description";
  int nsq = 1;
  TDiscreteForm3D *df = make_form((int)g_prm[PMX_PRM_FORM], nsq);
  if (!df || nsq != 1) return 3;  // one-matrix forms
  double params[2] = {P->SC_SMOOTH_DAMP_FACTOR_SCALAR, P->SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR};
  TMultiGrid3D *mg = new TMultiGrid3D(1, 2, params);
  std::vector<TSquareMatrix3D *> A(nlev);
  std::vector<std::vector<double>> sol(nlev), rhs(nlev);
  int N = 0;
  for (int l = 0; l < nlev; l++) {
    if (l) domain->RegRefineAll();
    TCollection *coll = domain->GetCollection(It_Finest, 0);
    TFESpace3D *space = new TFESpace3D(coll, name, desc, AllDirichlet, order);
    TSquareStructure3D *st = new TSquareStructure3D(space);
    st->Sort();
    N = space->GetN_DegreesOfFreedom();
    const int nact = space->GetActiveBound();
    A[l] = new TSquareMatrix3D(st);
    sol[l].assign(N, 0.0);
    rhs[l].assign(N, 0.0);
    TFESpace3D *fesp[1] = {space}, *ferhs[1] = {space};
    TSquareMatrix3D *sq[1] = {A[l]};
    double *rhsp[1] = {rhs[l].data()};
    BoundCondFunct3D *bc[1] = {AllDirichlet};
    BoundValueFunct3D *bv[1] = {BoundValue};
    TAuxParam3D *aux = new TAuxParam3D(1, 0, 0, 0, fesp, nullptr, nullptr, nullptr, nullptr, 0, nullptr);
    const auto ta = clock::now();
    A[l]->Reset();
    Assemble3D(1, fesp, 1, sq, 0, nullptr, 1, rhsp, ferhs, df, bc, bv, aux);
    t_asm += secs(ta, clock::now());
    std::memcpy(sol[l].data() + nact, rhs[l].data() + nact, sizeof(double) * (N - nact));
    mg->AddLevel(new TMGLevel3D(l, A[l], rhs[l].data(), sol[l].data(), 2, nullptr));
  }
  const double t_setup = secs(t0, clock::now()) - t_asm;
  // Solve: the Krylov method works on copies (SystemCD3D.C:378-395)
  std::vector<double> x(sol[nlev - 1]), b(rhs[nlev - 1]);
  const auto ts = clock::now();
  if (opt[16] > 0) {
    std::vector<double> r(N);
    TSquareMatrix *sqd[1] = {A[nlev - 1]};
    Defect_Scalar(sqd, nullptr, x.data(), b.data(), r.data());
    double rr = 0;
    for (double v : r) rr += v * v;
    P->SC_LIN_RES_NORM_MIN_SCALAR = opt[16] * std::sqrt(rr);
  }
  TItMethod *prec = new TMultiGridScaIte(MatVect_Scalar, Defect_Scalar, nullptr, 0, N, mg, 1);
  TItMethod *it = nullptr;
  if ((int)opt[0] == 16) it = new TFgmresIte(MatVect_Scalar, Defect_Scalar, prec, 0, N, 1);
  else if ((int)opt[0] == 11) it = new TFixedPointIte(MatVect_Scalar, Defect_Scalar, prec, 0, N, 1);
  else return 3;
  TSquareMatrix *sqm[1] = {A[nlev - 1]};
  const int iters = it->Iterate(sqm, nullptr, x.data(), b.data());
  const double t_solve = secs(ts, clock::now());
  fp = std::fopen(argv[8], "wb");
  if (!fp) return 1;
  const double head[1] = {(double)N};
  std::fwrite(head, sizeof(double), 1, fp);
  std::fwrite(x.data(), sizeof(double), N, fp);
  std::fwrite(b.data(), sizeof(double), N, fp);
  const double tail[4] = {(double)iters, t_setup, t_asm, t_solve};
  std::fwrite(tail, sizeof(double), 4, fp);
  std::fclose(fp);
  return 0;
}
