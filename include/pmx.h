/*
 * pmx.h -- C-ABI of libpmx, the MI355X implementation of ParMooN's CD3D/TCD3D
 * hot path: element assembly with SUPG, geometric multigrid and the
 * FGMRES/Richardson/BiCGStab solvers on one GPU.
 *
 * Conventions
 *  - Every function returns 0 (PMX_OK) or a pmx_status; pmx_last_error()
 *    returns the message of the last failure on the calling thread.  Nothing
 *    calls exit() (the reference aborts with exit(4711) / exit(0)).
 *  - Host arrays are borrowed for the duration of the call.  Vectors passed
 *    as host arrays are in the reference DOF numbering (GlobalNumbers of
 *    TFESpace3D); device storage is internal and lattice ordered.
 *  - Parameters come from an explicit `const double prm[PMX_PRM_LEN]` block
 *    (pmx_prm.h) instead of the global TDatabase::ParamDB.
 *  - All work is ordered on the context's single HIP stream.
 */
#ifndef PMX_H
#define PMX_H

#include "pmx_prm.h"

#ifdef __cplusplus
extern "C" {
#endif

enum pmx_status {
  PMX_OK = 0,
  PMX_ERR_ARG = 1,       /* invalid argument                                  */
  PMX_ERR_HIP = 2,       /* HIP runtime failure                               */
  PMX_ERR_STATE = 3,     /* call out of order (e.g. solve before assemble)    */
  PMX_ERR_NOCONV = 4,    /* iteration limit reached (cf. FgmresIte.C:604-612) */
  PMX_ERR_DIVERGED = 5   /* residual > SC_DIV_FACTOR * initial (FgmresIte.C:535, exit(4711) there) */
};

typedef struct pmx_ctx pmx_ctx;
typedef struct pmx_system pmx_system;

/* ---- context ----------------------------------------------------------- */
int pmx_ctx_create(int device, pmx_ctx **out);
/* fails with PMX_ERR_STATE while systems created on the context are alive */
int pmx_ctx_destroy(pmx_ctx *ctx);
int pmx_ctx_sync(pmx_ctx *ctx);
const char *pmx_last_error(void);

/* ---- host-only set-up (no device) ----------------------------------------
 * The same numbering / structure as pmx_space_numbering / pmx_space_structure
 * for the unit cube refined nref times, computed on the host only. */
int pmx_host_space_dims(int nref, int order, long long *dims);
int pmx_host_space_arrays(int nref, int order, int *begin_index, int *global_numbers, int *rowptr, int *kcol);
/* Legacy-VTK output of nvar scalar functions of that space (values[k]: N
 * doubles in the reference numbering, names[k]: the SCALARS name, no blanks),
 * replacing TOutput3D::WriteVtk (src/FE/Output3D.C:1773-2305) for scalar
 * Q1/Q2 functions: POINTS are the mesh vertices, CELLS the hexahedra
 * (type 12) in the reference's cell order with each cell's vertices in its
 * local order, POINT_DATA the function values at the vertices, CELL_DATA the
 * region id (0), numbers as %.6e.  The reference lists the points in the
 * order of their addresses in memory; here they are in lattice order (x
 * fastest), so the file equals the reference's up to that renumbering. */
int pmx_host_write_vtk(int nref, int order, int nvar, const char *const *names, const double *const *values,
                       const char *path);

/* ---- multilevel system --------------------------------------------------
 * Replaces the per-level set-up of 3DPrograms/CD3D_ParMooN.C:264-351
 * (RegRefineAll + new TFESpace3D(coll, ..., BoundCondition, ORDER)) and the
 * constructor TSystemCD3D(N_levels, fespaces, ...) (src/System/SystemCD3D.C:62-225):
 * levels are the unit cube (data/Wuerfel.GEO) refined nref0 .. nref0+nlev-1
 * times, all with Q1 (order 1) or Q2 (order 2) elements, all-Dirichlet. */
int pmx_system_create(pmx_ctx *ctx, int nref0, int nlev, int order, pmx_system **out);
int pmx_system_destroy(pmx_system *sys);
/* dims[0]=N_DegreesOfFreedom [1]=ActiveBound [2]=N_Cells [3]=N_Entries of the
 * CSR structure [4]=N_BaseFunct [5]=lattice points per direction
 * [6]=stored stencil values [7]=smoother colours */
int pmx_system_dims(pmx_system *sys, int level, long long *dims);

/* TFESpace3D::GetBeginIndex / GetGlobalNumbers (include/FE/FESpace3D.h),
 * identical to ConstructSpace (src/FE/FESpace3D.C:485-2080). */
int pmx_space_numbering(pmx_system *sys, int level, int *begin_index, int *global_numbers);
/* TSquareStructure3D(space) + Sort(): RowPtr / KCol
 * (src/FE/SquareStructure3D.C:59-200, src/FE/SquareStructure.C:91-110). */
int pmx_space_structure(pmx_system *sys, int level, int *rowptr, int *kcol);
/* multicolour smoother class of every DOF (no reference counterpart for the
 * sequential build; the hybrid build colours DOFs in ParFEMapper3D.C:1995) */
int pmx_space_colours(pmx_system *sys, int level, int *colour);

/* ---- assembly -----------------------------------------------------------
 * TSystemCD3D::Assemble (SystemCD3D.C:340-365): for every level
 * TAssembleMat3D::Reset + Assemble3D (src/System/AssembleMat3D.C:235,261-535)
 * with the local form prm[PMX_PRM_FORM] (the AssembleFct3D callbacks of
 * src/FE/ConvDiff3D.C and src/FE/TCD3D.C) and the coefficient functor
 * prm[PMX_PRM_EXAMPLE] (CoeffFct3D, include/General/Constants.h:148).
 * The result stays on the device. */
int pmx_system_assemble(pmx_system *sys, const double *prm);
/* The same with the caller's coefficient and boundary callbacks instead of a
 * built-in functor: the host-function form of CoeffFct3D, BoundCondFunct3D and
 * BoundValueFunct3D (include/General/Constants.h:128-129,148; BoundCond& and
 * double& become int* and double*, which ParMooN's example functions match in
 * the ABI).  coeffs is called once per cell with the cell's n quadrature
 * points (X, Y, Z) and fills coeffs[q][0..5] = eps, b1, b2, b3, c, f
 * (param[q] are zeros), as TDiscreteForm3D::GetLocalForms calls it
 * (DiscreteForm3D.C:212-301); the records go to the device and the forms run
 * there.  cond (optional) must return DIRICHLET (0) at every boundary node;
 * value (optional, else 0) gives the Dirichlet rhs at each boundary node.
 * bdcomp numbers the cube's planes as data/Wuerfel.PRM: 0 z = 0, 1 x = 0,
 * 2 y = 0, 3 z = 1, 4 x = 1, 5 y = 1 (the first that holds the node).  The
 * host evaluation costs n_q x n_cells callback points per level. */
typedef void pmx_coeff_fct(int n_points, double *X, double *Y, double *Z, double **param, double **coeffs);
typedef void pmx_bound_cond_fct(int bdcomp, double x, double y, double z, int *cond);
typedef void pmx_bound_value_fct(int bdcomp, double x, double y, double z, double *value);
int pmx_system_assemble_cb(pmx_system *sys, const double *prm, pmx_coeff_fct *coeffs, pmx_bound_cond_fct *cond,
                           pmx_bound_value_fct *value);
/* TSquareMatrix3D::GetEntries() of a level in CSR order; which = 0 the main
 * matrix (the operator of the solves), 1 the second matrix of two-matrix
 * forms (SUPG K of TCD3D.C:128-266), 2 the mass matrix and 3 the stiffness
 * matrix A of the time stepping; 4 the FP32 copy of the operator and 5 the
 * row-scaled FP16 smoother values of opt[15] = 3 / 4 (a_ij / a_ii, Dirichlet
 * rows 1), both as doubles. */
int pmx_matrix_entries(pmx_system *sys, int level, int which, double *entries);
/* TMGLevel3D(level, A, ...) with the caller's matrix (include/FE/MGLevel3D.h:
 * 142, added by TMultiGrid3D::AddLevel, include/FE/MultiGrid3D.h:94): the
 * operator of a level is replaced by the caller's entries (CSR of the
 * level's own pattern, reference numbering, n = the level's DOFs; Dirichlet
 * rows must be identity rows).  The system must be assembled; the MG cycles
 * of the following solves use these values (the FP32 copies and the coarse
 * inverse follow).  A pattern other than the level's is rejected. */
int pmx_level_set_entries(pmx_system *sys, int level, int n, const int *rowptr, const int *kcol,
                          const double *entries);
/* right-hand side of a level (Rhs_array[level]) */
int pmx_system_rhs(pmx_system *sys, int level, double *rhs);

/* ---- level operations on host vectors (reference numbering) --------------
 * MatVect_Scalar / Defect_Scalar (src/FE/Blas2.C:37-126); precision 0 = FP64
 * entries, 1 = FP32 copy of the entries. */
int pmx_matvec(pmx_system *sys, int level, int precision, const double *x, double *y);
int pmx_defect(pmx_system *sys, int level, const double *x, const double *b, double *r, double *res);
/* No counterpart in the reference: the allocation-free variant of pmx_defect
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  x, b, r are
 * device arrays of the level's N nodes in lattice order (node (i, j, k) of the
 * M^3 lattice at i + M (j + M k), M = pmx_system_dims[5]), res one device
 * double: r = b - A x (Dirichlet rows r = b - x), *res = ||r||_2.  Ordered on
 * the context's stream, no synchronisation, no allocation: the norm's partial
 * sums live in a static device array of 4096 doubles per GPU (one per 2048
 * rows), so levels of at most 8,388,608 rows are accepted (PMX_ERR_ARG
 * beyond).  Calls that overlap on one GPU share that array and race. */
int pmx_defect_noalloc(pmx_system *sys, int level, const double *x, const double *b, double *r, double *res);
/* TMGLevel3D::Jacobi (kind 1) / SOR (2) / SSOR (3) (src/FE/MGLevel3D.C:280-563),
 * SOR/SSOR in multicolour order; `sweeps` sweeps, sol updated in place. */
int pmx_smooth(pmx_system *sys, int level, int kind, double omega, int sweeps, double *sol, const double *f);
/* Prolongate / DefectRestriction between fine_level-1 and fine_level
 * (src/FE/MGComponents3D.C:208-488, 732-1014). */
int pmx_prolongate(pmx_system *sys, int fine_level, const double *xc, double *xf);
int pmx_restrict(pmx_system *sys, int fine_level, double *rc, const double *rf);
/* MatVect on a caller CSR matrix (any TSquareMatrix3D: RowPtr, KCol, Entries). */
int pmx_csr_matvec(pmx_ctx *ctx, int n, const int *rowptr, const int *kcol, const double *entries,
                   const double *x, double *y);

/* ---- solvers -------------------------------------------------------------
 * TSystemCD3D::Solve -> TItMethod::Iterate (include/FE/ItMethod.h) with the
 * TMultiGridScaIte preconditioner (src/FE/MultiGridScaIte.C:89-130).
 * opt (24 doubles; entries 0-15 have the oracle's layout):
 *  [0] solver: 16 FGMRES (FgmresIte.C:245), 11 Richardson (FixedPointIte.C:176),
 *      13 BiCGStab (Bcgs.C:111), 0 = one MG cycle
 *  [1] restart  [2] maxit  [3] res_norm_min (absolute)  [4] red_factor
 *  [5] smoother 1 Jacobi 2 SOR 3 SSOR  [6] pre  [7] post  [8] omega
 *  [9] coarse: 17 exact (FP64 dense inverse), 1/2 = Jacobi/SOR, other = SSOR sweeps
 *      until the coarse defect is reduced by [11] or [10] sweeps were made
 *      (MultiGrid3D.C:225-337; one workgroup where the level fits its LDS)
 *  [10] coarse maxit (SC_COARSE_MAXIT_SCALAR)
 *  [11] coarse reduction factor (SC_COARSE_RED_FACTOR_SCALAR)
 *  [12] cycle 1 V 2 W 0 F  [13] GMG damp
 *  [14] BiCGStab flexible branch (1 = v = A y)
 *  [15] precision: 0 FP64 throughout, 1 FP32 operator/MG/Krylov vectors
 *       inside an FP64 defect correction, 2 = 1 plus the learned post-smoother;
 *       3 / 4 = 1 / 2 with the SOR/SSOR sweeps that run launch per colour
 *       reading row-scaled FP16 values (a_ij / a_ii, saturated to the finite
 *       FP16 range) and 1 / a_ii in FP32: half the value bytes of the sweeps,
 *       which only precondition (defects, transfers and Krylov vectors keep
 *       the FP32 values, the outer residual FP64)
 *  [16] stop when ||r|| <= [16] * ||r0|| (r0 = initial FP64 defect; 0 = off)
 *  [17] FP32 inner solve: residual reduction per outer correction (default 1e-4)
 *  [18] 1 = launch the MG cycle kernel by kernel instead of replaying its HIP graph
 *  [19] precision 2: number of finest levels with the learned post-smoother
 *       (0 = the levels it was trained for)
 *  [20] precision 2: learned steps per post-smoothing (default 1)
 *  [21] SOR/SSOR sweeps of levels whose largest colour has at most this many
 *       rows run in one workgroup (same iterate, fewer launches); 0 = 512 in
 *       FP32 and never in FP64, < 0 = never
 *  [22] precision 2: the finest [22] of the [19] levels keep the plain smoother
 *  [23] precision 2: 1 = the learned steps also replace the pre-smoothing
 * info (out, 8 doubles): [0] iterations (Krylov steps summed) [1] final FP64
 * residual [2] initial residual [3] outer corrections [4] seconds (device
 * events) [5] learned-smoother fallbacks [6..7] reserved. */
int pmx_system_solve(pmx_system *sys, const double *opt, const double *rhs, double *sol, double *info);
/* TItMethod::Iterate(TSquareMatrix **A, TMatrix **B, double *sol, double *rhs)
 * (include/FE/ItMethod.h:105) on a caller's matrix: the Krylov method of
 * opt[0] works with the caller's CSR matrix of the finest level (n = N,
 * RowPtr / KCol / Entries in the reference numbering, FP64, uploaded and
 * permuted to the device order) and is preconditioned by the system's own
 * assembled MG levels (TMultiGridScaIte).  With opt[15] >= 1 the FP64 defect
 * correction uses the caller's matrix and the inner FP32 solves the system's
 * operator.  sol, rhs, info as in pmx_system_solve. */
int pmx_iterate(pmx_system *sys, int n, const int *rowptr, const int *kcol, const double *entries,
                const double *opt, const double *rhs, double *sol, double *info);
/* Step length control of the MG cycle (TMGLevel3D::StepLengthControl,
 * src/FE/MGLevel3D.C:908-960, applied as in TMultiGrid3D::Cycle,
 * src/FE/MultiGrid3D.C:465-471, 562-566, 671-679): on the levels it covers,
 * u and d = f - A u are kept before the pre-smoothing, and after the
 * post-smoothing u = uold + omega (u - uold) with
 * omega = (d, A(u - uold)) / ||A(u - uold)||^2 (opt[8] when that is undefined
 * or |omega| < 0.1).  mode: 0 off (the CD3D/TCD3D input files' setting),
 * 1 finest level (SC_STEP_LENGTH_CONTROL_FINE_SCALAR), 2 every level above
 * the coarsest (SC_STEP_LENGTH_CONTROL_ALL_SCALAR, Database.C:688's default).
 * Applies to the following solves of the system. */
int pmx_system_set_step_length_control(pmx_system *sys, int mode);
/* Device-resident variant: solves with the assembled right-hand side already
 * in HBM, zero initial guess with the Dirichlet values of the rhs
 * (SystemCD3D.C:356).  The solution stays on the device. */
int pmx_system_solve_resident(pmx_system *sys, const double *opt, double *info);
int pmx_system_solution(pmx_system *sys, double *sol);
/* Errors of the solution held on the device (last solve or time step) against
 * the example's exact solution at prm[PMX_PRM_TIME], with the quadrature rule
 * of prm (TFEFunction3D::GetErrors with L2H1Errors, src/FE/FEFunction3D.C:83-300,
 * 3DPrograms/TCD3D_ParMooN.C:640-655).  err: [0] L2 error [1] H1 semi-norm error. */
int pmx_system_errors(pmx_system *sys, const double *prm, double *err);

/* ---- time stepping --------------------------------------------------------
 * TSystemTCD3D (src/System/SystemTCD3D.C:208-428) as driven by the time loop
 * of 3DPrograms/TCD3D_ParMooN.C:565-663.
 * init: AssembleMRhs at prm[PMX_PRM_TIME] (mass matrix on every level and the
 * rhs, which becomes oldrhs); the solution is sol0 (reference numbering) or,
 * if sol0 is NULL, the nodal interpolant of the example's initial function;
 * its Dirichlet values are those of the rhs. */
int pmx_tcd_init(pmx_system *sys, const double *prm, const double *sol0);
/* One step to the time prm[PMX_PRM_TIME] (the caller advances it by tau):
 * AssembleARhs with the stiffness form prm[PMX_PRM_FORM], AssembleSystMat
 * (B = tau th3 oldrhs + tau th4 rhs + (M - tau th2 A) oldsol on active rows,
 * rhs on Dirichlet rows; system matrix M + tau th1 A on every level), Solve
 * with opt (as pmx_system_solve; info likewise), oldrhs = rhs.
 * Stiffness forms: PMX_FORM_TCD_A, _GALERKIN, _SDFEM and the SUPG form
 * PMX_FORM_TCD_SUPG_AK (MatricesAKRhsAssemble_SUPG, src/FE/TCD3D.C:128-266),
 * whose time-derivative matrix K joins M in both places (M + K - tau th2 A,
 * M + K + tau th1 A), as TSystemTCD2D wires its SUPG case
 * (src/System/SystemTCD2D.C:294-297, 336-341); pmx_tcd_init starts that case
 * from MatrixMRhsAssemble_SUPG (TCD3D.C:74-126) when prm[PMX_PRM_FORM] is
 * PMX_FORM_TCD_SUPG_AK or PMX_FORM_MASS_SUPG.
 * td: [0] tau [1] theta1 [2] theta2 [3] theta3 [4] theta4
 * (SetTimeDiscParameters, src/FE/TimeDiscRout.C:59-140); the step sets
 * prm[PMX_PRM_TAU] = td[0] and prm[PMX_PRM_THETA1] = td[1] for the forms
 * (TimeDB->CURRENTTIMESTEPLENGTH / THETA1).  The solution stays on the device
 * (pmx_system_solution reads it). */
int pmx_tcd_step(pmx_system *sys, const double *prm, const double *td, const double *opt, double *info);
/* The same step in TSystemTCD3D's phases (SystemTCD3D.C:229-330), for hosts
 * that drive the time loop of 3DPrograms/TCD3D_ParMooN.C:565-640 call by call:
 *  - AssembleARhs (:229-246): oldsol = sol, the stiffness form (and K of the
 *    SUPG form) and the rhs at prm[PMX_PRM_TIME] on every level, sol's
 *    Dirichlet values from the rhs; td as in pmx_tcd_step;
 *  - AssembleSystMat (:248-300): the right-hand side B and the system matrix
 *    M (+ K) + tau th1 A on every level, then oldrhs = rhs (the loop's copy);
 *    PMX_ERR_STATE if the system is still assembled (the reference exits);
 *  - Solve with opt (as pmx_system_solve; the solution stays on the device);
 *  - RestoreMassMat (:312-330): M, K and A are kept apart, so nothing is
 *    subtracted; the state returns to "not assembled" (PMX_ERR_STATE if it was
 *    not assembled, where the reference exits).
 * Between the phases pmx_matrix_entries reads the matrices: which 0 the
 * operator (the system matrix once assembled), 1 K, 2 M, 3 A. */
int pmx_tcd_assemble_arhs(pmx_system *sys, const double *prm, const double *td);
int pmx_tcd_assemble_systmat(pmx_system *sys);
int pmx_tcd_solve(pmx_system *sys, const double *opt, double *info);
int pmx_tcd_restore_massmat(pmx_system *sys);

/* ---- measurement ---------------------------------------------------------
 * Times `reps` back-to-back launches of one kernel on the context stream with
 * HIP events: 0 stencil SpMV, 1 defect, 2 Jacobi sweep, 3 SSOR sweep,
 * 4 assembly (all kernels of one level), 5 prolongation, 6 restriction,
 * 7 one MG V-cycle, 8 one learned post-smoothing step (FP32), 9 the network's
 * forward convolutions alone on the level's lattice (FP32/BF16); these are
 * timed with two events around the `reps` repetitions (ms per repetition).
 * 10 and 11 give the average duration of a single dispatch, from events
 * recorded by the dispatch itself (hipExtLaunchKernelGGL), without the gaps
 * between dispatches: 10 one colour launch of an SSOR sweep (launch per
 * colour), 11 one SpMV; calibration of 10 on the same level, also per
 * dispatch: 12 a dispatch that does nothing with the workgroup count of each
 * colour launch, 13 a plain streaming read of each colour's stored values in
 * the precision's storage.  precision 0 FP64, 1 FP32, 3 FP32 with the
 * SOR/SSOR colour launches on row-scaled FP16 values (opt[15] = 3).  out: ms. */
int pmx_time_kernel(pmx_system *sys, int kernel, int level, int precision, int reps, double *ms);

/* ---- learned post-smoother (no counterpart in the reference; it replaces
 * the post-sweeps of TMGLevel3D::SOR/SSOR, src/FE/MGLevel3D.C:280-467, on the
 * finest levels of FP32 cycles).  Network: 4 conv layers 3x3x3 with
 * 16 -> 16 -> 16 -> 16 -> 1 channels and GELU, BF16 operands with FP32
 * accumulation.  Parameters are FP32, per layer W[cout][27][cin] then b[cout],
 * tap t = (dx+1) + 3 (dy+1) + 9 (dz+1).  Lattice arrays are x-fastest M^3.
 * Network input per node (16 channels): the scaled defect, then
 * a(p,p+e)/a(p,p) for e = +x, -x, +y, -y, +z, -z, then a(p,p+2e)/a(p,p) for
 * the same six e (0 where the stencil has no such entry), then three zeros. */
int pmx_learned_nparams(int *n);
int pmx_learned_init(pmx_system *sys, unsigned seed);
/* Each level above the coarsest has its own network.  init: fresh draws for
 * all of them.  params: set != 0 uploads, set == 0 downloads (finest level's
 * network; _level for any level).  forward / gradient / adam act on the
 * finest level's network. */
int pmx_learned_params(pmx_system *sys, int set, float *params);
int pmx_learned_params_level(pmx_system *sys, int level, int set, float *params);
/* network output on an M^3 lattice; in: [M^3][16] input channels */
int pmx_learned_forward(pmx_system *sys, int M, const float *in, float *out);
/* loss = sum over interior nodes (out - target)^2 / (M-2)^3 and its gradient
 * with respect to the parameters (fresh, not accumulated); in: [M^3][16] */
int pmx_learned_gradient(pmx_system *sys, int M, const float *in, const float *target, float *grads,
                         double *loss);
/* residual loss on a level: c = network output on the level's lattice (0 on
 * its boundary), loss = sum over interior nodes (in - D^-1 A c)^2 / (M-2)^3,
 * and its parameter gradient (fresh).  in: the scaled defect (channel 0),
 * lattice order; the other channels are the level's operator features. */
int pmx_learned_gradient_level(pmx_system *sys, int level, const float *in, float *grads, double *loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) with the last gradient */
int pmx_learned_adam(pmx_system *sys, double lr);
/* train on pairs from plain cycles (opt: the cycle options of pmx_system_solve).
 * topt (16 doubles, unused entries 0): [0] samples per level, [1] Adam steps per level (levels interleaved),
 * [2] learning rate, [3] seed, [4] finest levels that use the network (0: all
 * but the coarsest), [5] loss: 0 squared error against the cycle's error,
 * 1 the scaled defect left after the correction (as pmx_learned_gradient_level),
 * [6] samples: 0 a random start of the homogeneous problem, 1 a zero start
 * with a random right-hand side, 2 half each, [7] the finest levels among [4]
 * left untrained (for solves with opt[22] of that value), [8] samples per Adam
 * step (batch, default 1; the step uses the mean of their gradients).
 * tinfo (16 doubles): [0] first / [1] last mean loss of the finest trained network,
 * [2] s, [3] total steps, [4 + 2j] / [5 + 2j] first / last mean loss of the
 * network of level L - j. */
int pmx_learned_train(pmx_system *sys, const double *opt, const double *topt, double *tinfo);
/* one learned post-smoothing step on a level (reference numbering, sol in place) */
int pmx_learned_post(pmx_system *sys, int level, double *sol, const double *f);

#ifdef __cplusplus
}
#endif
#endif
