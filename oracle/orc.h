/*
 * orc.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a single-threaded FP64 C++
 * restatement of ParMooN's CD3D/TCD3D hot path (mesh refinement, DOF
 * numbering, CSR structure, local assembling, GMG smoothers and transfers,
 * FGMRES/Richardson/BiCGStab, error norms, the TCD3D time loop).  It is the
 * parity checker for the HIP implementation in parmoon_amd/ and the timed
 * "reference CPU path".  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links or calls it.
 */
#ifndef ORC_H
#define ORC_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- FE space on the unit cube refined `nref` times, Q1 (order 1) or Q2 (order 2) ---- */
void *orc_space_new(int nref, int order);
void orc_space_free(void *sp);
/* dims[0]=N_DOF, [1]=ActiveBound, [2]=N_Cells, [3]=N_Entries (CSR), [4]=N_BaseFunct */
void orc_space_dims(void *sp, long long *dims);
void orc_space_arrays(void *sp, int *begin_index, int *global_numbers,
                      int *rowptr, int *kcol);
/* physical coordinates of every DOF (3*N, xyz interleaved) */
void orc_dof_coords(void *sp, double *xyz);
/* cell vertex coordinates (N_Cells*8*3) in the cell's own local vertex order */
void orc_cell_vertices(void *sp, double *xyz);

/* Assemble one form on one space: entries (N_Entries) and rhs (N) are
 * zeroed first (TAssembleMat3D::Reset) then accumulated cell by cell.
 * entries2 receives the second matrix of two-matrix forms (SUPG K), may be NULL. */
int orc_assemble(void *sp, const double *prm, double *entries, double *rhs,
                 double *entries2);

/* ---- BLAS2 on the space's CSR ---- */
void orc_matvec(void *sp, const double *entries, const double *x, double *y);
double orc_defect(void *sp, const double *entries, const double *x,
                  const double *b, double *r);

/* ---- smoothers (one sweep, TMGLevel3D semantics) : kind 1 Jacobi, 2 SOR, 3 SSOR ---- */
void orc_smooth(void *sp, const double *entries, int kind, double omega,
                double *sol, const double *f, double *aux);

/* SOR (2) / SSOR (3) with the active rows visited in `order` (n rows) */
void orc_smooth_ordered(void *sp, const double *entries, int kind, double omega, double *sol,
                        const double *f, const int *order, int n);

/* ---- grid transfers between space `coarse` (nref) and `fine` (nref+1) ---- */
void orc_prolongate(void *coarse, void *fine, const double *xc, double *xf);
void orc_restrict(void *coarse, void *fine, double *rc, const double *rf);

/* ---- multilevel system: levels nref0 .. nref0+nlev-1, all order `order` ---- */
void *orc_system_new(int nref0, int nlev, int order);
void orc_system_free(void *sys);
void *orc_system_space(void *sys, int level);
/* assemble every level with the same parameter block; returns 0 on success */
int orc_system_assemble(void *sys, const double *prm);
/* the finest level's matrix / rhs (pointers into oracle storage) */
double *orc_system_entries(void *sys, int level);
double *orc_system_rhs(void *sys, int level);

/* solver options, doubles so a numpy array can carry them:
 * [0] solver: 11 Richardson, 16 FGMRES, 13 BiCGStab, 0 = single MG cycle
 * [1] restart (FGMRES)         [2] maxit
 * [3] res_norm_min (absolute)  [4] red_factor
 * [5] smoother 1 Jacobi 2 SOR 3 SSOR   [6] pre  [7] post  [8] omega
 * [9] coarse smoother (1,2,3 or 17=exact LU) [10] coarse maxit [11] coarse red
 * [12] cycle 1=V 2=W 0=F   [13] gmg damp   [14] bicgstab flexible (0 = reference default)
 * info out: [0] iterations [1] final residual [2] initial residual */
int orc_system_solve(void *sys, const double *opt, const double *rhs,
                     double *sol, double *info);
/* step length control of the MG cycle (TMGLevel3D::StepLengthControl,
 * MGLevel3D.C:908): 0 off, 1 finest level (SC_STEP_LENGTH_CONTROL_FINE_SCALAR),
 * 2 every level above the coarsest (SC_STEP_LENGTH_CONTROL_ALL_SCALAR) */
void orc_system_set_slc(void *sys, int mode);

/* ---- errors: L2 and H1-semi against the example's exact solution at prm time ---- */
void orc_errors(void *sp, const double *prm, const double *sol, double *err);
void orc_interpolate(void *sp, const double *prm, int which, double *sol);

/* ---- TCD3D time loop (3DPrograms/TCD3D_ParMooN.C:565-753), Galerkin M and A ----
 * tdisc: [0] tau, [1] end time, [2] theta1 [3] theta2 [4] theta3 [5] theta4
 * sol in/out (initial value interpolated by the caller if desired);
 * err_out: per step {time, L2, H1, L2(0,T;L2), L2(0,T;H1)} (5 * max_steps doubles);
 * returns the number of steps taken. */
int orc_tcd3d_run(void *sys, const double *prm, const double *tdisc,
                  const double *opt, double *sol, double *err_out, int max_steps);

/* ---- dense LU (coarse solve, restates dgetrf/dgetrs semantics) ---- */
int orc_dense_lu_solve(int n, double *a_colmajor, double *b);

#ifdef __cplusplus
}
#endif
#endif
