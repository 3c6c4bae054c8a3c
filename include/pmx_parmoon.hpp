// pmx_parmoon.hpp -- C++ classes over the libpmx C-ABI with the names and
// call order of ParMooN's CD3D/TCD3D surface, for host code written against
// that surface.  Header-only; link with libpmx.so.
//
//   TFESpace3D         GetN_DegreesOfFreedom / GetActiveBound / GetN_Cells /
//                      GetBeginIndex / GetGlobalNumbers   (include/FE/FESpace3D.h)
//   TSquareStructure3D GetN_Rows / GetN_Entries / GetActiveBound / GetRowPtr /
//                      GetKCol                            (include/FE/SquareStructure3D.h)
//   TSystemCD3D        ctor(N_levels, ...), Init, Assemble, Solve
//                                                         (include/System/SystemCD3D.h:129-140)
//   TSystemTCD3D       AssembleMRhs, AssembleARhs, AssembleSystMat, Solve,
//                      RestoreMassMat                     (include/System/SystemTCD3D.h)
//   TAssembleMat3D     Init, Reset, Assemble3D, DeAllocate (include/System/AssembleMat3D.h)
//   TFEFunction3D      GetName / GetValues / GetLength    (include/FE/FEFunction3D.h)
//   TOutput3D          AddFEFunction, WriteVtk            (include/FE/Output3D.h)
//   TSquareMatrix3D    GetEntries / Reset                 (include/FE/SquareMatrix3D.h)
//   TMGLevel3D, TMultiGrid3D::AddLevel, TItMethod::Iterate (include/FE/MGLevel3D.h,
//                      MultiGrid3D.h, ItMethod.h)
//
// Differences: parameters come from an explicit block (pmx_prm.h) and an
// option vector (pmx.h) instead of TDatabase; errors throw pmx::Error
// instead of calling exit(); the spaces are the unit cube refined
// nref0 .. nref0 + N_levels - 1 times.
#ifndef PMX_PARMOON_HPP
#define PMX_PARMOON_HPP

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "pmx.h"

namespace pmx {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc, const char *what) {
  if (rc != PMX_OK) throw Error(rc, std::string(what) + ": " + pmx_last_error());
}

class Context {
 public:
  explicit Context(int device = 0) { check(pmx_ctx_create(device, &h_), "pmx_ctx_create"); }
  ~Context() { pmx_ctx_destroy(h_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  pmx_ctx *handle() const { return h_; }

 private:
  pmx_ctx *h_ = nullptr;
};

// Scalar Q1/Q2 space of the unit cube refined nref times (host only).
class TFESpace3D {
 public:
  TFESpace3D(int nref, int order) : nref_(nref), order_(order) {
    check(pmx_host_space_dims(nref, order, dims_), "pmx_host_space_dims");
  }
  int GetN_DegreesOfFreedom() const { return (int)dims_[0]; }
  int GetActiveBound() const { return (int)dims_[1]; }
  int GetN_Cells() const { return (int)dims_[2]; }
  int GetN_BaseFunct() const { return (int)dims_[4]; }
  int GetOrder() const { return order_; }
  int GetRefinements() const { return nref_; }
  const int *GetBeginIndex() { load(); return bi_.data(); }
  const int *GetGlobalNumbers() { load(); return gn_.data(); }
  const int *RowPtr() { load(); return rp_.data(); }
  const int *KCol() { load(); return kc_.data(); }
  long long N_Entries() const { return dims_[3]; }

 private:
  void load() {
    if (!bi_.empty()) return;
    bi_.resize(dims_[2] + 1);
    gn_.resize(dims_[2] * dims_[4]);
    rp_.resize(dims_[0] + 1);
    kc_.resize(dims_[3]);
    check(pmx_host_space_arrays(nref_, order_, bi_.data(), gn_.data(), rp_.data(), kc_.data()),
          "pmx_host_space_arrays");
  }
  int nref_, order_;
  long long dims_[8] = {0};
  std::vector<int> bi_, gn_, rp_, kc_;
};

// CSR pattern of a space (TSquareStructure3D(space) followed by Sort()).
class TSquareStructure3D {
 public:
  explicit TSquareStructure3D(TFESpace3D &space) : space_(space) {}
  int GetN_Rows() const { return space_.GetN_DegreesOfFreedom(); }
  int GetN_Columns() const { return space_.GetN_DegreesOfFreedom(); }
  int GetN_Entries() const { return (int)space_.N_Entries(); }
  int GetActiveBound() const { return space_.GetActiveBound(); }
  const int *GetRowPtr() { return space_.RowPtr(); }
  const int *GetKCol() { return space_.KCol(); }

 private:
  TFESpace3D &space_;
};

// A scalar finite element function: the space and the caller's coefficient
// array (reference numbering), as TFEFunction3D(fespace, name, description,
// values, length) (include/FE/FEFunction3D.h:71).
class TFEFunction3D {
 public:
  TFEFunction3D(TFESpace3D *space, const char *name, const char *description, double *values, int length)
      : space_(space), name_(name ? name : ""), description_(description ? description : ""), values_(values),
        length_(length) {}
  TFESpace3D *GetFESpace3D() const { return space_; }
  const char *GetName() const { return name_.c_str(); }
  const char *GetDescription() const { return description_.c_str(); }
  double *GetValues() const { return values_; }
  int GetLength() const { return length_; }

 private:
  TFESpace3D *space_;
  std::string name_, description_;
  double *values_;
  int length_;
};

// Output of scalar functions of one space: AddFEFunction, WriteVtk
// (include/FE/Output3D.h:128-136, src/FE/Output3D.C:141-162,1773-2305).  The
// capacities of the constructor grow as needed, as in the reference; the
// domain argument is not used (the space carries the mesh).
class TOutput3D {
 public:
  TOutput3D(int maxn_fespaces, int maxn_scalar, int maxn_vect, int maxn_parameters, const void *domain = nullptr) {
    (void)maxn_fespaces; (void)maxn_vect; (void)maxn_parameters; (void)domain;
    funcs_.reserve(maxn_scalar > 0 ? maxn_scalar : 1);
  }
  // returns the number of scalar functions stored (as the reference)
  int AddFEFunction(TFEFunction3D *f) {
    if (!funcs_.empty() && (f->GetFESpace3D()->GetRefinements() != funcs_[0]->GetFESpace3D()->GetRefinements() ||
                            f->GetFESpace3D()->GetOrder() != funcs_[0]->GetFESpace3D()->GetOrder()))
      throw Error(PMX_ERR_ARG, "TOutput3D: functions of different spaces");
    if (f->GetLength() != f->GetFESpace3D()->GetN_DegreesOfFreedom())
      throw Error(PMX_ERR_ARG, "TOutput3D: function length differs from the space's DOFs");
    funcs_.push_back(f);
    return (int)funcs_.size();
  }
  int WriteVtk(const char *name) {
    if (funcs_.empty()) throw Error(PMX_ERR_ARG, "TOutput3D: no function added");
    std::vector<const char *> names;
    std::vector<const double *> vals;
    for (auto *f : funcs_) {
      names.push_back(f->GetName());
      vals.push_back(f->GetValues());
    }
    const TFESpace3D *s = funcs_[0]->GetFESpace3D();
    check(pmx_host_write_vtk(s->GetRefinements(), s->GetOrder(), (int)funcs_.size(), names.data(), vals.data(),
                             name),
          "pmx_host_write_vtk");
    return 0;
  }

 private:
  std::vector<TFEFunction3D *> funcs_;
};

// TSystemCD3D: levels 0 .. N_levels-1 (coarsest first), solution and rhs
// arrays of the finest level owned by the caller (SolArray / RhsArray).
class TSystemCD3D {
 public:
  TSystemCD3D(Context &ctx, int N_levels, int nref_coarsest, int order, double *sol, double *rhs)
      : sol_(sol), rhs_(rhs), nlev_(N_levels) {
    check(pmx_system_create(ctx.handle(), nref_coarsest, N_levels, order, &h_), "pmx_system_create");
    long long d[8];
    check(pmx_system_dims(h_, -1, d), "pmx_system_dims");
    n_dof_ = (int)d[0];
    n_active_ = (int)d[1];
  }
  ~TSystemCD3D() { pmx_system_destroy(h_); }
  TSystemCD3D(const TSystemCD3D &) = delete;
  TSystemCD3D &operator=(const TSystemCD3D &) = delete;

  // Init(BilinearCoeffs, BoundCondition, BoundValue, ...): the coefficient
  // functions, the discretisation and the boundary data are the example and
  // form fields of the parameter block.
  void Init(const double *prm) { prm_.assign(prm, prm + PMX_PRM_LEN); }
  // TSystemCD3D::Assemble (SystemCD3D.C:340-365): all levels on the device;
  // the finest rhs goes to RhsArray and its Dirichlet values to SolArray.
  void Assemble() {
    if (prm_.empty()) throw Error(PMX_ERR_STATE, "Init first");
    check(pmx_system_assemble(h_, prm_.data()), "pmx_system_assemble");
    check(pmx_system_rhs(h_, -1, rhs_), "pmx_system_rhs");
    for (int i = n_active_; i < n_dof_; i++) sol_[i] = rhs_[i];
  }
  // TSystemCD3D::Solve(sol, rhs) with the solver / MG options of opt (pmx.h);
  // returns the number of iterations like TItMethod::Iterate.
  int Solve(const double *opt, double *info = nullptr) {
    double inf[8];
    const int rc = pmx_system_solve(h_, opt, rhs_, sol_, inf);
    if (info)
      for (int i = 0; i < 8; i++) info[i] = inf[i];
    check(rc, "pmx_system_solve");
    return (int)inf[0];
  }
  // TSquareMatrix3D::GetEntries() of a level (CSR order of its structure)
  void GetEntries(int level, double *entries) { check(pmx_matrix_entries(h_, level, 0, entries), "entries"); }
  int GetN_DegreesOfFreedom() const { return n_dof_; }
  int GetActiveBound() const { return n_active_; }
  int GetN_Levels() const { return nlev_; }
  pmx_system *handle() const { return h_; }

 private:
  pmx_system *h_ = nullptr;
  double *sol_, *rhs_;
  int nlev_, n_dof_ = 0, n_active_ = 0;
  std::vector<double> prm_;
};

// A square matrix on a structure with caller-owned entries
// (TSquareMatrix3D(structure), include/FE/SquareMatrix3D.h; GetEntries).
class TSquareMatrix3D {
 public:
  explicit TSquareMatrix3D(TSquareStructure3D &structure)
      : st_(structure), entries_((size_t)structure.GetN_Entries(), 0.0) {}
  TSquareStructure3D &GetStructure() { return st_; }
  int GetN_Rows() const { return st_.GetN_Rows(); }
  int GetActiveBound() const { return st_.GetActiveBound(); }
  int GetN_Entries() const { return st_.GetN_Entries(); }
  const int *GetRowPtr() { return st_.GetRowPtr(); }
  const int *GetKCol() { return st_.GetKCol(); }
  double *GetEntries() { return entries_.data(); }
  void Reset() { std::fill(entries_.begin(), entries_.end(), 0.0); }

 private:
  TSquareStructure3D &st_;
  std::vector<double> entries_;
};

// A multigrid level with the caller's matrix (TMGLevel3D(level, A, rhs, sol,
// n_aux, permutation), include/FE/MGLevel3D.h:142).  rhs and sol are the
// level's work arrays in the reference; the device keeps its own.
class TMGLevel3D {
 public:
  TMGLevel3D(int level, TSquareMatrix3D *A, double *rhs, double *sol, int n_aux, int *permutation)
      : level_(level), A_(A) {
    (void)rhs; (void)sol; (void)n_aux;
    if (permutation) throw Error(PMX_ERR_ARG, "TMGLevel3D: permuted levels are not supported");
  }
  int GetLevel() const { return level_; }
  TSquareMatrix3D *GetMatrix() const { return A_; }
  int GetN_DOF() const { return A_->GetN_Rows(); }

 private:
  int level_;
  TSquareMatrix3D *A_;
};

// TMultiGrid3D (include/FE/MultiGrid3D.h:87-110): AddLevel hands the level's
// matrix to the device hierarchy of a system (pmx_level_set_entries; the
// reference's constructor takes no system, here it is the first argument).
// The cycle's options (smoother, sweeps, damping, cycle type) are the
// solver option vector of the solve, as the reference reads them from
// ParamDB in Cycle; Parameters are the smoother's damping factors.
class TMultiGrid3D {
 public:
  TMultiGrid3D(pmx_system *system, int n_problems, int n_parameters, const double *parameters)
      : h_(system), params_(parameters, parameters + (parameters ? n_parameters : 0)) {
    (void)n_problems;
  }
  // returns the number of levels added (level numbers are the system's,
  // coarsest 0)
  int AddLevel(TMGLevel3D *lv) {
    TSquareMatrix3D *A = lv->GetMatrix();
    check(pmx_level_set_entries(h_, lv->GetLevel(), A->GetN_Rows(), A->GetRowPtr(), A->GetKCol(), A->GetEntries()),
          "pmx_level_set_entries");
    levels_.push_back(lv);
    return (int)levels_.size();
  }
  int GetN_Levels() const { return (int)levels_.size(); }
  TMGLevel3D *GetLevel(int i) const { return levels_.at(i); }
  const std::vector<double> &GetParameters() const { return params_; }
  pmx_system *handle() const { return h_; }

 private:
  pmx_system *h_;
  std::vector<double> params_;
  std::vector<TMGLevel3D *> levels_;
};

// TFgmresIte / TFixedPointIte / TBcgs preconditioned by TMultiGridScaIte on
// a TMultiGrid3D (include/FE/ItMethod.h:105): Iterate(A, B, sol, rhs) on the
// caller's finest matrix (pmx_iterate); the solver id and its parameters are
// opt[0..4] (16 FGMRES, 11 Richardson, 13 BiCGStab).  Returns the iterations.
class TItMethod {
 public:
  TItMethod(TMultiGrid3D &mg, const double *opt) : mg_(mg), opt_(opt, opt + 24) {}
  int Iterate(TSquareMatrix3D **sqmat, void *mat, double *sol, double *rhs, double *info = nullptr) {
    (void)mat;
    TSquareMatrix3D *A = sqmat[0];
    double inf[8];
    const int rc = pmx_iterate(mg_.handle(), A->GetN_Rows(), A->GetRowPtr(), A->GetKCol(), A->GetEntries(),
                               opt_.data(), rhs, sol, inf);
    if (info)
      for (int i = 0; i < 8; i++) info[i] = inf[i];
    check(rc, "pmx_iterate");
    return (int)inf[0];
  }

 private:
  TMultiGrid3D &mg_;
  std::vector<double> opt_;
};

// TAssembleMat3D (include/System/AssembleMat3D.h:53-91, Init / Reset /
// Assemble3D / DeAllocate): the local form of a TDiscreteForm3D (a pmx_form id
// in prm[PMX_PRM_FORM]) with the caller's CoeffFct3D, BoundCondFunct3D and
// BoundValueFunct3D (pmx_system_assemble_cb), assembled on the device for all
// levels of a system; Assemble3D fills the caller's entries of one level (the
// matrix of TSquareStructure3D's RowPtr / KCol, and the second matrix of a
// two-matrix form) and its rhs in place, as the reference accumulates into
// caller-owned TSquareMatrix3D and rhs arrays.
class TAssembleMat3D {
 public:
  TAssembleMat3D(pmx_system *system, int level, const double *prm, pmx_coeff_fct *coeffs,
                 pmx_bound_cond_fct *cond, pmx_bound_value_fct *value, double *entries, double *rhs,
                 double *entries2 = nullptr)
      : h_(system), level_(level), prm_(prm, prm + PMX_PRM_LEN), coeffs_(coeffs), cond_(cond), value_(value),
        entries_(entries), entries2_(entries2), rhs_(rhs) {
    long long d[8];
    check(pmx_system_dims(h_, level, d), "pmx_system_dims");
    n_dof_ = d[0];
    n_entries_ = d[3];
  }
  void Init() {}
  // zero the caller's matrices and rhs (TAssembleMat3D::Reset)
  void Reset() {
    for (long long i = 0; i < n_entries_; i++) {
      entries_[i] = 0.0;
      if (entries2_) entries2_[i] = 0.0;
    }
    for (long long i = 0; i < n_dof_; i++) rhs_[i] = 0.0;
  }
  void Assemble3D() {
    check(pmx_system_assemble_cb(h_, prm_.data(), coeffs_, cond_, value_), "pmx_system_assemble_cb");
    check(pmx_matrix_entries(h_, level_, 0, entries_), "pmx_matrix_entries");
    if (entries2_) check(pmx_matrix_entries(h_, level_, 1, entries2_), "pmx_matrix_entries");
    check(pmx_system_rhs(h_, level_, rhs_), "pmx_system_rhs");
  }
  void DeAllocate() {}

 private:
  pmx_system *h_;
  int level_;
  std::vector<double> prm_;
  pmx_coeff_fct *coeffs_;
  pmx_bound_cond_fct *cond_;
  pmx_bound_value_fct *value_;
  double *entries_, *entries2_, *rhs_;
  long long n_dof_ = 0, n_entries_ = 0;
};

// TSystemTCD3D with the time loop of 3DPrograms/TCD3D_ParMooN.C:565-663:
//   AssembleMRhs();                       once
//   per step: AssembleARhs(); AssembleSystMat(); Solve(sol); RestoreMassMat();
// each call runs its phase on the device (pmx_tcd_assemble_arhs,
// pmx_tcd_assemble_systmat, pmx_tcd_solve, pmx_tcd_restore_massmat), so the
// rhs (after AssembleARhs) and the matrices (GetEntries) can be read between
// them; RestoreMassMat subtracts nothing because M, K and A are kept apart.
class TSystemTCD3D {
 public:
  TSystemTCD3D(Context &ctx, int N_levels, int nref_coarsest, int order, double *sol)
      : sol_(sol) {
    check(pmx_system_create(ctx.handle(), nref_coarsest, N_levels, order, &h_), "pmx_system_create");
  }
  ~TSystemTCD3D() { pmx_system_destroy(h_); }
  TSystemTCD3D(const TSystemTCD3D &) = delete;
  TSystemTCD3D &operator=(const TSystemTCD3D &) = delete;

  void Init(const double *prm, const double *opt) {
    prm_.assign(prm, prm + PMX_PRM_LEN);
    opt_.assign(opt, opt + 24);
  }
  // mass matrix and rhs at the current time; sol (the initial solution) is
  // taken as it is, with the Dirichlet values of the rhs
  void AssembleMRhs() { check(pmx_tcd_init(h_, prm_.data(), sol_), "pmx_tcd_init"); }
  // the stiffness matrix and rhs at `time` with the step's tau and theta
  // (TDatabase::TimeDB CURRENTTIME, CURRENTTIMESTEPLENGTH, THETA1..4 after
  // SetTimeDiscParameters); rhs (optional): the assembled rhs of the finest
  // level in the reference numbering
  void AssembleARhs(double time, double tau, double th1, double th2, double th3, double th4,
                    double *rhs = nullptr) {
    prm_[PMX_PRM_TIME] = time;
    const double td[5] = {tau, th1, th2, th3, th4};
    check(pmx_tcd_assemble_arhs(h_, prm_.data(), td), "pmx_tcd_assemble_arhs");
    if (rhs) check(pmx_system_rhs(h_, -1, rhs), "pmx_system_rhs");
  }
  void AssembleSystMat() { check(pmx_tcd_assemble_systmat(h_), "pmx_tcd_assemble_systmat"); }
  int Solve(double *info = nullptr) {
    double inf[8];
    const int rc = pmx_tcd_solve(h_, opt_.data(), inf);
    if (info)
      for (int i = 0; i < 8; i++) info[i] = inf[i];
    check(rc, "pmx_tcd_solve");
    check(pmx_system_solution(h_, sol_), "pmx_system_solution");
    return (int)inf[0];
  }
  void RestoreMassMat() { check(pmx_tcd_restore_massmat(h_), "pmx_tcd_restore_massmat"); }
  // entries of a level's matrix: 0 the operator (the system matrix between
  // AssembleSystMat and RestoreMassMat), 1 K, 2 M, 3 A
  void GetEntries(int level, int which, double *entries) {
    check(pmx_matrix_entries(h_, level, which, entries), "pmx_matrix_entries");
  }
  pmx_system *handle() const { return h_; }

 private:
  pmx_system *h_ = nullptr;
  double *sol_;
  std::vector<double> prm_, opt_;
};

}  // namespace pmx

#endif
