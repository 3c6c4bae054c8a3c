// pmx_ops.hpp -- stream-ordered launchers of the libpmx kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "pmx_internal.hpp"

namespace pmx {

// device mirror of a Layout
struct DevLayout {
  int order = 1, n = 0, M = 0, S = 2;
  int64_t N = 0, nvals = 0, nactive = 0;
  int nboundary = 0;
  int *boundary = nullptr;
  DevSegment *segs = nullptr;
  int *segidx = nullptr;  // residue class (rx + S(ry + S rz)) -> segment, -1 if empty
  BlockInfo *blocks_all = nullptr;  // all rows, spatial order (z, y of the first row)
  int nblocks_all = 0;
  BlockInfo *blocks_colour = nullptr;
  const RowBlock *rows_all = nullptr;     // blocks_all with the class records inline (row kernels)
  const RowBlock *rows_colour = nullptr;  // blocks_colour likewise
  std::vector<int> colour_off;  // host: block offsets per colour
  int *colour_off_dev = nullptr;  // the same on the device
  int *colour_seg_dev = nullptr;  // first segment of each colour (segments are colour-major), ncolours + 1
  int64_t max_colour_rows = 0;    // rows of the largest colour
  int nsegs = 0;                  // residue classes with rows
  int ncolours = 0;
};

void check(hipError_t e, const char *what);

// the numbering of build_numbering on the device (pmx_setup.hip): lat2dof,
// dof2lat (N each), GlobalNumbers (cells x local DOFs, optional) and the
// per-cell frame codes (n^3), all device arrays; synchronises the stream
void device_numbering(const Layout &L, int *lat2dof, int *dof2lat, int *gn, uint8_t *corner, hipStream_t st);
// the CSR structure of build_csr on the device (pmx_setup.hip): rowptr
// (N + 1) and kcol (csr_nnz(L)) from the device numbering; synchronises st
void device_csr(const Layout &L, const int *lat2dof, const int *dof2lat, int *rowptr, int *kcol, hipStream_t st);
int64_t csr_nnz(const Layout &L);

// event pairs of timed dispatches (see set_dispatch_timer)
struct DispatchTimer {
  std::vector<hipEvent_t> ev;
  int used = 0;
  double total_ms() const;  // sum of the dispatch durations (synchronise first)
  DispatchTimer() = default;
  DispatchTimer(const DispatchTimer &) = delete;
  DispatchTimer &operator=(const DispatchTimer &) = delete;
  ~DispatchTimer();
};
// while installed (non-null), the row kernels of css_apply record each
// dispatch between two events of the timer; nullptr uninstalls
void set_dispatch_timer(DispatchTimer *t);

// calibration of the dispatch measurements (both honour the dispatch timer):
// a dispatch of `blocks` workgroups that does nothing, and a plain streaming
// read of `bytes` bytes at p (16-byte aligned; sink: one float, written never
// in practice)
void empty_dispatch(int blocks, hipStream_t st);
void stream_read(const void *p, int64_t bytes, float *sink, hipStream_t st);

template <typename T> void fill(T *x, T v, int64_t n, hipStream_t st);
template <typename T> void copy(T *y, const T *x, int64_t n, hipStream_t st);
template <typename T>
void axpy(T *y, const T *x, const double *alpha_dev, double sgn, double ac, int64_t n, hipStream_t st);
template <typename T, typename U> void convert(T *y, const U *x, int64_t n, hipStream_t st);
template <typename T, typename U> void add_convert(T *y, const U *x, int64_t n, hipStream_t st);
template <typename T> void scale_inv(T *y, const T *x, const double *nrm, int64_t n, hipStream_t st);
template <typename T>
void dot(const T *x, const T *y, int64_t n, double *part, double *out, int do_sqrt, hipStream_t st);
// out = (x, y) (sqrt with do_sqrt) with one partial sum per kDotChunk
// elements in part (ceil(n / kDotChunk) doubles), summed in a fixed order
constexpr int kDotChunk = 2048;
void dot_chunked(const double *x, const double *y, int64_t n, double *part, double *out, int do_sqrt,
                 hipStream_t st);
// e = u - uold
template <typename T> void axpy_diff(T *e, const T *u, const T *uold, int64_t n, hipStream_t st);
// TMGLevel3D::StepLengthControl's update: omega = num_nom[0] / num_nom[1]
// (param0 when num_nom[1] <= 0 or |omega| < 0.1), u = uold + omega (u - uold)
template <typename T>
void step_length_update(T *u, const T *uold, const double *num_nom, double param0, int64_t n, hipStream_t st);
// modified Gram-Schmidt of w against the k1 vectors V[0..k1) (stride n) and
// normalisation: h[k] = (w, v_k) with w -= h[k] v_k in turn, h[k1] = ||w||,
// vnext = w / ||w||, in k1 + 2 launches (each step reduces the previous
// step's partial sums itself; the values are those of dot + axpy in the same
// order).  part: 2 kRedBlocks doubles.
template <typename T>
void mgs_arnoldi(T *w, const T *V, int64_t n, int k1, double *h, T *vnext, double *part, hipStream_t st);

// mode 0 spmv, 1 defect, 2 jacobi (x = old iterate, y = new), 3 sor (in place on y); colour -1 = all classes
template <typename T>
void css_apply(const DevLayout &D, int mode, const T *val, const T *x, const T *f, T *y, double omega, int colour,
               hipStream_t st);
// out = {||u - u_h||_L2, |u - u_h|_H1} of the lattice function u (FP64) against
// the example's exact solution with the rule of qref/tab (part: 2 kRedBlocks)
void fe_errors(const DevLayout &D, const double *prm_dev, const double *qref, const double *tab, int nq,
               const double *u, double *part, double *out, hipStream_t st);
// a == b bit for bit (n doubles; flag: one device int of scratch; synchronises st)
bool bits_equal(const double *a, const double *b, int64_t n, int *flag, hipStream_t st);
// whole SOR (ssor 0) / SSOR (ssor 1) sweep with Dirichlet rows x = f, one
// workgroup with the level's x in LDS (levels with M^3 * s <= kSweepLdsBytes)
inline bool sweep_wg_fits(const DevLayout &D, size_t s) { return (int64_t)D.N * (int64_t)s <= kSweepLdsBytes; }
// Lanes per row of the row kernels on a level (the same for every kernel and
// precision of the level, so the single-workgroup and the launch-per-colour
// sweeps compute each row alike): 4 where a colour has at most kSweepWgRows
// rows (few rows: more lanes per row shorten the load chain of each), 1 from
// 2^20 active rows on (one lane per row touches the fewest cache lines per
// gathered x run), 2 between.
inline int row_lanes(const DevLayout &D) {
  if (D.max_colour_rows <= kSweepWgRows) return 4;
  return D.nactive >= (int64_t(1) << 20) ? 1 : 2;
}
// one SOR colour (mode 4 of k_rows) on the row-scaled FP16 values and the
// FP32 1/diag of scaled_values; x in place
void css_sor_scaled(const DevLayout &D, const _Float16 *val16, const float *dinv, float *x, const float *f,
                    double omega, int colour, hipStream_t st);
// the row-scaled FP16 copy of the FP32 stencil values (a_ij / a_ii,
// saturated to the finite FP16 range) and dinv = 1 / a_ii on the active rows
void scaled_values(const DevLayout &D, const float *val, _Float16 *val16, float *dinv, hipStream_t st);
template <typename T>
void sweep_wg(const DevLayout &D, const T *val, T *x, const T *f, double omega, int ssor, hipStream_t st);
// the coarse-level smoothing loop of TMultiGrid3D (MultiGrid3D.C:225-337) in
// one workgroup (levels with sweep_wg_fits): sweeps of kind 1 Jacobi, 2 SOR,
// 3 SSOR while ||f - A x|| > red ||f - A x_0|| and fewer than maxit were
// made; d, xj: scratch of D.N entries
template <typename T>
void coarse_wg(const DevLayout &D, const T *val, T *x, const T *f, double omega, int kind, int maxit, double red,
               T *d, T *xj, hipStream_t st);
// Dirichlet rows: mode 0 y=x, 1 y=f-x, 2 y=f, 3 y=0
template <typename T> void boundary(const DevLayout &D, int mode, T *y, const T *x, const T *f, hipStream_t st);
template <typename T>
void prolongate(const DevLayout &C, const DevLayout &F, const T *xc, T *xf, hipStream_t st);
// xf += damp * (P xc with the fine Dirichlet nodes set to 0): prolongation,
// CorrectNodes and the damped update of TMultiGrid3D::Cycle in one launch
// (the Dirichlet nodes of a layout are its lattice surface)
template <typename T>
void prolongate_add(const DevLayout &C, const DevLayout &F, const T *xc, T *xf, double damp, hipStream_t st);
// zero_surface: the coarse Dirichlet entries are set to 0 (CorrectDefect)
template <typename T>
void restrict_(const DevLayout &C, const DevLayout &F, T *rc, const T *rf, hipStream_t st, bool zero_surface = false);
// coarsest-level solve: y_I = W x with W = A_II^-1 [I, -A_IB] (n_I x N, lattice
// columns; iidx: interior row -> lattice node); y_B = x_B is set separately
template <typename T>
void gemv_coarse(int nI, int N, const double *W, const int *iidx, const T *x, T *y, hipStream_t st);
// W of the level by Gauss-Jordan with partial pivoting (cidx: lattice node ->
// interior row or -1; aug: n_I x (n_I + N) workspace; piv: n_I ints)
void coarse_inverse_dense(const DevLayout &D, const double *val, const int *cidx, int nI, double *aug, int *piv,
                          double *W, hipStream_t st);
template <typename T, typename U> void gather(T *out, const U *in, const int *map, int64_t n, hipStream_t st);
template <typename T>
void csr_spmv(int n, const int *rp, const int *kc, const T *val, const T *x, T *y, hipStream_t st);
// r = b - A x on a CSR matrix
template <typename T>
void csr_defect(int n, const int *rp, const int *kc, const T *val, const T *x, const T *b, T *r, hipStream_t st);
// CSR entries with the level's pattern (reference numbering) into the CSS
// values of the active rows (the inverse of css_to_csr there)
void csr_to_css(const DevLayout &D, const int *rowptr, const int *kcol, const int *dof2lat, int64_t active,
                const double *in, double *val, hipStream_t st);
template <typename T>
void css_to_csr(const DevLayout &D, const int *rowptr, const int *kcol, const int *dof2lat, int64_t active,
                const T *val, double *out, hipStream_t st);
// example: prm[PMX_PRM_EXAMPLE] (selects evaluating the coefficients inline or
// through the quadrature-point buffer); form: prm[PMX_PRM_FORM]; kc: cell
// buffer of n^3 x 27 x 28 doubles for the Q2 one-matrix forms on MFMA
// (nullptr: the per-row evaluation); qp_given: qpbuf already holds the
// quadrature-point records (coefficients from host callbacks)
// ftab (optional): room for fourier_tab_len(n, nq) doubles; with it the
// curl-Fourier example's coefficients come from per-level phase tables.
// corner (required for CELL_MEASURE 4, else unused): the per-cell frame codes
// of Numbering::corner on the device.
void assemble(const DevLayout &D, const double *prm_dev, int example, int form, const double *qref, int nq,
              const double *tab, double *qpbuf, double *kc, double *val, double *val2, double *rhs, hipStream_t st,
              bool qp_given = false, double *ftab = nullptr, const uint8_t *corner = nullptr);
// doubles of the curl-Fourier phase table of a level with n cells per
// direction and an nq-point tensor rule (complex factors of the 3 x 8 modes
// of PMX_PRM_FOURIER per axis, cell index and 1D point)
constexpr int kFourierModes = 24;
inline int64_t fourier_tab_len(int n, int nq) {
  const int np = nq <= 8 ? 2 : (nq <= 27 ? 3 : 4);
  return 2 * (int64_t)kFourierModes * 3 * n * np;
}
// y[idx[i]] = v[i]
void scatter(double *y, const int *idx, const double *v, int n, hipStream_t st);
constexpr int64_t kCellMatrixDoubles = 27 * 28;  // cell buffer entries per Q2 cell
// time stepping (TSystemTCD3D): nodal interpolation of the example's initial
// function, the system matrix val = mass (+ k) + c a, and the Crank-Nicolson
// / theta right-hand side (k, k_old: the SUPG time-derivative matrix K and
// K oldsol, nullptr for the Galerkin forms)
void interpolate(const DevLayout &D, const double *prm_dev, double *x, hipStream_t st);
// val = mass (+ k) + c a; val32 (if not null) receives its FP32 rounding
void system_matrix(double *val, float *val32, const double *mass, const double *k, const double *a, double c,
                   int64_t n, hipStream_t st);
void tcd_rhs(const DevLayout &D, const double *oldrhs, const double *rhs, const double *m_old, const double *a_old,
             const double *k_old, double t3, double t4, double t2, double *B, hipStream_t st);
void givens(double *H, double *cs, double *sn, double *s, int i, int ld, double *res, hipStream_t st);
void backsolve(const double *H, const double *s, double *y, int k, int ld, hipStream_t st);
template <typename T, typename X>
void multi_axpy(X *x, const T *z, const double *y, int k1, int64_t n, int64_t stride, hipStream_t st);
void set_scalar(double *dst, const double *src, double scale, hipStream_t st);

}  // namespace pmx
