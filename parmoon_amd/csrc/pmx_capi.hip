// pmx_capi.hip -- the extern "C" boundary of libpmx and the host orchestration
// of assembly, multigrid and Krylov solvers on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pmx.h"
#include "pmx_internal.hpp"
#include "pmx_functors.hpp"
#include "pmx_learned.hpp"
#include "pmx_ops.hpp"

namespace pmx {

static thread_local std::string g_err;
// V(2,2) SSOR with the exact coarse solve, used when a single kernel is timed
// FGMRES residual ratio above which a learned cycle counts as failed
static const double kLearnStall = 0.99;
static const double PMX_DEFAULT_VCYCLE[24] = {0, 1, 1, 0, 0, 3, 2, 2, 1.0, 17, 1, 0, 1, 1.0, 0, 0};

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t m) {
    if (m == n && p) return;
    release();
    if (m == 0) return;
    check(hipMalloc(&p, sizeof(T) * m), "hipMalloc");
    n = m;
  }
  // alloc, and a fresh allocation starts zeroed (stream-ordered): the
  // padding rows of the last value slice of a class are never written by
  // the assembly and stay zero
  void alloc_zeroed(size_t m, hipStream_t st) {
    if (m == n && p) return;
    alloc(m);
    if (p) check(hipMemsetAsync(p, 0, sizeof(T) * m, st), "clear");
  }
  void upload(const T *h, size_t m, hipStream_t st) {
    alloc(m);
    check(hipMemcpyAsync(p, h, sizeof(T) * m, hipMemcpyHostToDevice, st), "upload");
  }
};

struct Ctx {
  int device = 0;
  int systems = 0;  // live systems on this context
  hipStream_t st = nullptr;
  DevBuf<double> part;  // reduction partials: 3 x kRedBlocks, then scalars
  double *pinned = nullptr;
};

// per-precision work vectors of one level
template <typename T>
struct Vecs {
  DevBuf<T> val;                      // CSS values in this precision
  DevBuf<T> x, f, d, aux;             // MG level vectors
  DevBuf<T> sold, sdef;               // step length control: u and f - A u before the pre-smoothing
};

struct Level {
  Layout L;
  Numbering Nb;
  DevLayout D;
  DevBuf<int> boundary, segidx;
  DevBuf<DevSegment> segs;
  DevBuf<BlockInfo> blocks_all, blocks_colour;
  DevBuf<RowBlock> rows_all, rows_colour;  // the same with the class records inline
  DevBuf<int> colour_off, colour_seg;
  DevBuf<int> lat2dof, dof2lat, rowptr, kcol;
  DevBuf<double> val2;                // second matrix (SUPG K)
  DevBuf<double> mass;                // time stepping: mass matrix M
  DevBuf<double> stiff;               // time stepping: stiffness matrix A (the system matrix is in v64.val)
  DevBuf<double> rhs;                 // assembled rhs (lattice order)
  Vecs<double> v64;
  Vecs<float> v32;
  bool have_val32 = false;
  DevBuf<double> ainv;                // coarsest level: A_II^-1 [I, -A_IB] (n_I x N)
  DevBuf<double> ainv_src;            // the values ainv was built from
  DevBuf<int> cidx, iidx;             // lattice node <-> interior row of ainv
  DevBuf<double> aug;                 // Gauss-Jordan workspace
  DevBuf<int> piv, flag;
  int ainv_n = 0;
  // learned post-smoother: 1/diag, correction and its image under A (FP32)
  DevBuf<float> dinv, lc, lac;
  DevBuf<float> valT;  // transposed FP32 stencil values (residual-loss training)
  DevBuf<uint16_t> feat;              // network input features [N][8] (BF16)
  bool have_dinv = false;
  // FP16 smoother values: a_ij / a_ii of the FP32 values and 1 / a_ii
  DevBuf<_Float16> val16;
  DevBuf<float> sdinv;
  bool have_val16 = false;
  DevBuf<uint8_t> corner;  // per-cell frame codes (Numbering::corner), CELL_MEASURE 4
};

struct System {
  Ctx *ctx = nullptr;
  int order = 1;
  std::vector<std::unique_ptr<Level>> lev;
  bool assembled = false;
  bool two_matrices = false;
  DevBuf<double> prm, qref, tab, qpbuf;
  int64_t qp_tab_off = 0;  // offset of the curl-Fourier phase tables in qpbuf
  int cell_measure = 0;    // prm[PMX_PRM_CELL_MEASURE] of the last assembly
  int nq = 0;
  int tab_key = -1;                   // rule of the uploaded tables
  int example = 0;                    // prm[PMX_PRM_EXAMPLE] of the last assembly
  int form = 0;                       // prm[PMX_PRM_FORM] of the last assembly
  bool mfma_assembly = true;          // Q2 one-matrix forms through the cell buffer
  DevBuf<double> kc;                  // Q2 cell matrices (finest level's cells x 27 x 28)
  // Krylov workspace
  DevBuf<double> sol, b, r64;         // finest level, FP64 (lattice)
  DevBuf<double> sol_try, r_try;      // FP32 form: the corrected iterate and its defect before acceptance
  // a caller's finest-level CSR matrix (lattice-permuted) as the FP64 Krylov
  // operator of pmx_iterate; the MG cycle keeps the system's own levels
  DevBuf<int> ucsr_rp, ucsr_kc;
  DevBuf<double> ucsr_val;
  bool use_ucsr = false;
  DevBuf<double> oldrhs, oldsol, tmp1, tmp2, tmp3;  // time stepping (finest level)
  bool tcd_ready = false;
  // phase of the time step (TSystemTCD3D's call order): 0 before
  // AssembleARhs, 1 after it, 2 system matrix assembled (SystMatAssembled)
  int tcd_phase = 0;
  double td[5] = {0};
  DevBuf<double> krylov64;
  DevBuf<float> krylov32;
  DevBuf<double> scal;                // H, cs, sn, s, y, beta, res
  std::vector<int> rec;
  double opt[24] = {0};
  // captured MG cycles (finest f -> finest x), keyed by precision and cycle options
  struct Graph {
    bool valid = false;
    int prec = 0;
    double key[16] = {0};
    hipGraphExec_t exec = nullptr;
  };
  Graph graphs[2];
  bool use_graphs = true;
  // learned post-smoother on the `learned_levels` finest levels (FP32 cycles),
  // one network per level
  static constexpr int kMaxLevels = 10;
  LearnedNet nets[kMaxLevels];
  int learned_levels = 1;
  int learned_skip = 0;     // finest levels inside learned_levels that keep the plain smoother
  bool learned_pre = false; // learned steps also replace the pre-smoothing
  int trained_levels = 1;  // levels of the last training
  // step length control of the MG cycle: 0 off, 1 finest level
  // (SC_STEP_LENGTH_CONTROL_FINE_SCALAR), 2 every level above the coarsest
  // (SC_STEP_LENGTH_CONTROL_ALL_SCALAR)
  int slc = 0;
  bool learn_active = false;
  bool x_zero_shortcut = true;  // see mg_cycle
  int fallbacks = 0;
  // FP32 forms: SOR/SSOR sweeps launched per colour read row-scaled FP16
  // values (opt[15] = 3 / 4)
  bool half_smoother = false;
  DevBuf<double> lscal;               // norm, (d, Ac), (Ac, Ac)
  ~System() {
    for (auto &n : nets) n.release();
    for (auto &g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
  }
  void invalidate_graphs() {
    for (auto &g : graphs) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g.exec = nullptr;
      g.valid = false;
    }
  }
};

// ------------------------------------------------------------- set-up
static void setup_level(System &Y, Level &lv, int nref) {
  hipStream_t st = Y.ctx->st;
  build_layout(lv.L, nref, Y.order, kRowsPerBlock);
  const Layout &L = lv.L;
  {
    // the DOF numbering on the device, with host copies for the host-side
    // users (exports, CSR on demand, permutations of caller data)
    Numbering &Nb = lv.Nb;
    const int k = L.order + 1;
    Nb.nb = k * k * k;
    Nb.ncells = (int64_t)L.n * L.n * L.n;
    Nb.N = L.N;
    Nb.active = L.nactive;
    Nb.begin_index.resize(Nb.ncells + 1);
    for (int64_t i = 0; i <= Nb.ncells; i++) Nb.begin_index[i] = (int)(i * Nb.nb);
    lv.lat2dof.alloc(L.N);
    lv.dof2lat.alloc(L.N);
    lv.corner.alloc(Nb.ncells);
    DevBuf<int> gn;
    gn.alloc(Nb.ncells * Nb.nb);
    device_numbering(L, lv.lat2dof.p, lv.dof2lat.p, gn.p, lv.corner.p, st);
    Nb.global_numbers.resize(Nb.ncells * Nb.nb);
    Nb.lat2dof.resize(L.N);
    Nb.dof2lat.resize(L.N);
    Nb.corner.resize(Nb.ncells);
    check(hipMemcpyAsync(Nb.global_numbers.data(), gn.p, sizeof(int) * gn.n, hipMemcpyDeviceToHost, st), "d2h");
    check(hipMemcpyAsync(Nb.lat2dof.data(), lv.lat2dof.p, sizeof(int) * L.N, hipMemcpyDeviceToHost, st), "d2h");
    check(hipMemcpyAsync(Nb.dof2lat.data(), lv.dof2lat.p, sizeof(int) * L.N, hipMemcpyDeviceToHost, st), "d2h");
    check(hipMemcpyAsync(Nb.corner.data(), lv.corner.p, Nb.ncells, hipMemcpyDeviceToHost, st), "d2h");
    check(hipStreamSynchronize(st), "numbering copies");
    Nb.have_csr = false;
  }
  DevLayout &D = lv.D;
  D.order = L.order; D.n = L.n; D.M = L.M; D.S = L.S;
  D.N = L.N; D.nvals = L.nvals; D.nactive = L.nactive;
  D.ncolours = L.ncolours;
  std::vector<DevSegment> ds(L.segs.size());
  for (size_t i = 0; i < L.segs.size(); i++) {
    const Segment &s = L.segs[i];
    ds[i] = DevSegment{s.x0, s.y0, s.z0, s.cx, s.cy, s.cz, s.ex, s.ey, s.ez, s.ns, s.dslot, L.S, s.count, s.voff};
  }
  lv.segs.upload(ds.data(), ds.size(), st);
  std::vector<int> sidx(L.S * L.S * L.S, -1);
  for (size_t i = 0; i < L.segs.size(); i++) {
    const Segment &s = L.segs[i];
    sidx[s.rx + L.S * (s.ry + L.S * s.rz)] = (int)i;
  }
  lv.segidx.upload(sidx.data(), sidx.size(), st);
  // Launch order of the blocks: sorted by the lattice position of their first
  // row (z, then y) across residue classes, in every list.  A launch's blocks
  // are split into 8 contiguous chunks, one per XCD (xcd_remap), so each XCD
  // works on one z-slab: the x values its rows gather (the slab plus two
  // planes either side) stay in that XCD's L2 instead of every XCD streaming
  // the whole vector once per class, and the rows sharing a cell are
  // assembled close in time.  Rows of one list are independent, so the order
  // changes no result.
  auto key = [&](const BlockInfo &b) {
    const Segment &sg = L.segs[b.seg];
    const int64_t t = b.base / sg.cx;
    const int64_t y = sg.y0 + (int64_t)L.S * (t % sg.cy), z = sg.z0 + (int64_t)L.S * (t / sg.cy);
    return std::make_tuple(z, y, b.seg, b.base);
  };
  auto spatial = [&](std::vector<BlockInfo> v) {
    std::stable_sort(v.begin(), v.end(), [&](const BlockInfo &a, const BlockInfo &b) { return key(a) < key(b); });
    return v;
  };
  auto inline_records = [&](const std::vector<BlockInfo> &v) {
    std::vector<RowBlock> r(v.size());
    for (size_t i = 0; i < v.size(); i++) r[i] = RowBlock{ds[v[i].seg], v[i].base, 0};
    return r;
  };
  {
    const std::vector<BlockInfo> ba = spatial(L.blocks_all);
    lv.blocks_all.upload(ba.data(), ba.size(), st);
    const std::vector<RowBlock> ra = inline_records(ba);
    lv.rows_all.upload(ra.data(), ra.size(), st);
  }
  std::vector<BlockInfo> bc;
  D.colour_off.assign(L.ncolours + 1, 0);
  for (int c = 0; c < L.ncolours; c++) {
    D.colour_off[c] = (int)bc.size();
    const std::vector<BlockInfo> bs = spatial(L.blocks_colour[c]);
    bc.insert(bc.end(), bs.begin(), bs.end());
  }
  D.colour_off[L.ncolours] = (int)bc.size();
  lv.blocks_colour.upload(bc.data(), bc.size(), st);
  {
    const std::vector<RowBlock> rc = inline_records(bc);
    lv.rows_colour.upload(rc.data(), rc.size(), st);
  }
  lv.colour_off.upload(D.colour_off.data(), D.colour_off.size(), st);
  {
    // segments are stored colour-major: colour c owns segments cs[c] .. cs[c+1]-1
    std::vector<int> cs(L.ncolours + 1, 0);
    std::vector<int64_t> rows(L.ncolours, 0);
    for (size_t i = 0; i < L.segs.size(); i++) {
      const int c = L.segs[i].colour;
      if (i > 0 && c < L.segs[i - 1].colour) throw std::logic_error("segments not colour-major");
      cs[c + 1]++;
      rows[c] += L.segs[i].count;
    }
    for (int c = 0; c < L.ncolours; c++) cs[c + 1] += cs[c];
    lv.colour_seg.upload(cs.data(), cs.size(), st);
    D.colour_seg_dev = lv.colour_seg.p;
    D.max_colour_rows = *std::max_element(rows.begin(), rows.end());
    D.nsegs = (int)L.segs.size();
  }
  lv.boundary.upload(L.boundary.data(), L.boundary.size(), st);
  D.boundary = lv.boundary.p; D.nboundary = (int)L.boundary.size();
  D.segs = lv.segs.p; D.segidx = lv.segidx.p;
  D.blocks_all = lv.blocks_all.p; D.nblocks_all = (int)L.blocks_all.size();
  D.blocks_colour = lv.blocks_colour.p;
  D.rows_all = lv.rows_all.p;
  D.rows_colour = lv.rows_colour.p;
  D.colour_off_dev = lv.colour_off.p;

  lv.rhs.alloc(L.N);
  lv.v64.val.alloc_zeroed(L.nvals, st);
  for (auto *b : {&lv.v64.x, &lv.v64.f, &lv.v64.d, &lv.v64.aux}) b->alloc(L.N);
}

// Gauss rules of src/QuadFormulas/QuadFormulaHexa.C:70-146 and Q1/Q2 Lagrange
// tables in the canonical frame: qref[q] = {xi, eta, zeta, w},
// tab[q][b] = {phi, dphi/dxi, dphi/deta, dphi/dzeta}
static void make_tables(int order, bool linear, std::vector<double> &qref, std::vector<double> &tab, int &nq) {
  const int deg = (linear ? 2 : 3) * order;
  const int np = deg <= 2 ? 2 : (deg <= 5 ? 3 : 4);
  double x1[4], w1[4];
  if (np == 2) {
    const double a = 0.5773502691896257645091489;
    x1[0] = -a; x1[1] = a; w1[0] = w1[1] = 1.0;
  } else if (np == 3) {
    x1[0] = -0.7745966692414833770358530; x1[1] = 0; x1[2] = 0.7745966692414833770358530;
    w1[0] = 0.5555555555555555555555558; w1[1] = 0.8888888888888888888888888; w1[2] = 0.5555555555555555555555558;
  } else {
    x1[0] = 0.33998104358485626480266576; x1[1] = 0.86113631159405257522394649;
    x1[2] = -0.86113631159405257522394649; x1[3] = -0.33998104358485626480266576;
    w1[0] = 0.652145154862546142626936051; w1[1] = 0.347854845137453857373063949;
    w1[2] = 0.347854845137453857373063949; w1[3] = 0.652145154862546142626936051;
  }
  nq = np * np * np;
  const int K = order + 1, NB = K * K * K;
  qref.assign(4 * nq, 0.0);
  tab.assign((size_t)nq * NB * 4, 0.0);
  auto lag = [&](double t, double *v, double *d) {
    if (order == 1) { v[0] = 0.5 * (1 - t); v[1] = 0.5 * (1 + t); d[0] = -0.5; d[1] = 0.5; }
    else { v[0] = 0.5 * t * (t - 1); v[1] = 1 - t * t; v[2] = 0.5 * t * (t + 1); d[0] = t - 0.5; d[1] = -2 * t; d[2] = t + 0.5; }
  };
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++) {
        const int q = k + np * (j + np * i);
        const double r[3] = {x1[k], x1[j], x1[i]};
        qref[4 * q] = r[0]; qref[4 * q + 1] = r[1]; qref[4 * q + 2] = r[2];
        qref[4 * q + 3] = w1[i] * w1[j] * w1[k];
        double v[3][3], d[3][3];
        for (int a = 0; a < 3; a++) lag(r[a], v[a], d[a]);
        for (int c = 0; c < K; c++)
          for (int bb = 0; bb < K; bb++)
            for (int a = 0; a < K; a++) {
              const int b = a + K * (bb + K * c);
              double *t = &tab[((size_t)q * NB + b) * 4];
              t[0] = v[0][a] * v[1][bb] * v[2][c];
              t[1] = d[0][a] * v[1][bb] * v[2][c];
              t[2] = v[0][a] * d[1][bb] * v[2][c];
              t[3] = v[0][a] * v[1][bb] * d[2][c];
            }
      }
}

// explicit FP64 inverse of the coarsest operator (device Gauss-Jordan); used
// as the exact coarse solve (SC_COARSE_SMOOTHER_SCALAR 17, MGLevel3D.C:865-905,
// which refactors a dense copy on every call; here it is formed once per assembly)
// reference tables and quadrature of the element for the rule in use
// (uploaded again only when the rule changes)
static void ensure_tables(System &Y, bool linear) {
  const int key = 2 * Y.order + (linear ? 1 : 0);
  if (Y.tab_key == key) return;
  std::vector<double> qref, tab;
  int nq;
  make_tables(Y.order, linear, qref, tab, nq);
  Y.qref.upload(qref.data(), qref.size(), Y.ctx->st);
  Y.tab.upload(tab.data(), tab.size(), Y.ctx->st);
  Y.nq = nq;
  // the qp records of the finest level, then room for the curl-Fourier
  // phase tables of any level
  Y.qp_tab_off = (int64_t)Y.lev.back()->Nb.ncells * nq * 8;
  Y.qpbuf.alloc((size_t)(Y.qp_tab_off + fourier_tab_len(Y.lev.back()->L.n, nq)));
  Y.tab_key = key;
}

// the cell buffer of the MFMA assembly (Q2 only; sized for the finest level)
static double *cell_buffer(System &Y) {
  if (Y.order != 2 || !Y.mfma_assembly) return nullptr;
  const int64_t n = Y.lev.back()->L.n;
  Y.kc.alloc((size_t)(n * n * n * kCellMatrixDoubles));
  return Y.kc.p;
}

static void coarse_inverse(System &Y, Level &lv) {
  const int N = (int)lv.L.N;
  const int nI = (int)lv.L.nactive;
  if (nI > 6000) { lv.ainv_n = 0; return; }
  hipStream_t st = Y.ctx->st;
  if (!lv.cidx.p) {
    const int M = lv.L.M;
    std::vector<int> c(N, -1), ii;
    for (int z = 1; z < M - 1; z++)
      for (int y = 1; y < M - 1; y++)
        for (int x = 1; x < M - 1; x++) {
          const int p = x + M * (y + M * z);
          c[p] = (int)ii.size();
          ii.push_back(p);
        }
    lv.cidx.upload(c.data(), c.size(), st);
    lv.iidx.upload(ii.data(), ii.size(), st);
  }
  // the inverse of the values it was built from is still valid when the
  // coarsest operator has not changed bit for bit (time steps with
  // time-independent coefficients re-assemble the same matrix)
  const int64_t nv = lv.L.nvals;
  lv.flag.alloc(1);
  if (lv.ainv_n == nI && lv.ainv_src.p && bits_equal(lv.v64.val.p, lv.ainv_src.p, nv, lv.flag.p, st)) return;
  lv.aug.alloc((size_t)nI * (nI + N));
  lv.piv.alloc(nI);
  lv.ainv.alloc((size_t)nI * N);
  lv.ainv_n = 0;
  coarse_inverse_dense(lv.D, lv.v64.val.p, lv.cidx.p, nI, lv.aug.p, lv.piv.p, lv.ainv.p, st);
  lv.ainv_src.alloc(nv);
  copy<double>(lv.ainv_src.p, lv.v64.val.p, nv, st);
  check(hipStreamSynchronize(st), "coarse inverse");
  lv.ainv_n = nI;
}

template <typename T> Vecs<T> &vecs(Level &lv);
template <typename T> DevBuf<T> &kry_buf(System &Y);
template <> DevBuf<double> &kry_buf<double>(System &Y) { return Y.krylov64; }
template <> DevBuf<float> &kry_buf<float>(System &Y) { return Y.krylov32; }
template <> Vecs<double> &vecs<double>(Level &lv) { return lv.v64; }
template <> Vecs<float> &vecs<float>(Level &lv) { return lv.v32; }

static void ensure_val32(System &Y) {
  for (auto &lp : Y.lev) {
    Level &lv = *lp;
    if (lv.have_val32) continue;
    if (!lv.v32.val.p) {
      Y.invalidate_graphs();
      lv.v32.val.alloc(lv.L.nvals);
      for (auto *b : {&lv.v32.x, &lv.v32.f, &lv.v32.d, &lv.v32.aux}) b->alloc(lv.L.N);
      // the defect's Dirichlet entries are read by the learned step with zero
      // weights only (dinv and A c vanish there), so they just have to be finite
      fill<float>(lv.v32.d.p, 0.0f, lv.L.N, Y.ctx->st);
    }
    convert<float, double>(lv.v32.val.p, lv.v64.val.p, lv.L.nvals, Y.ctx->st);
    lv.have_val32 = true;
  }
}

// row-scaled FP16 smoother values of every level (from the FP32 values)
static void ensure_val16(System &Y) {
  ensure_val32(Y);
  for (auto &lp : Y.lev) {
    Level &lv = *lp;
    if (lv.have_val16) continue;
    if (!lv.val16.p) {
      Y.invalidate_graphs();
      lv.val16.alloc(lv.L.nvals);
      lv.sdinv.alloc(lv.L.N);
      fill<float>(lv.sdinv.p, 0.0f, lv.L.N, Y.ctx->st);
    }
    scaled_values(lv.D, lv.v32.val.p, lv.val16.p, lv.sdinv.p, Y.ctx->st);
    lv.have_val16 = true;
  }
}

// buffers of the learned post-smoother on the levels that use it
static void ensure_learned(System &Y) {
  ensure_val32(Y);
  const int nl = (int)Y.lev.size();
  Y.lscal.alloc(8);
  for (int i = std::max(0, nl - Y.learned_levels); i < nl; i++) {
    Level &lv = *Y.lev[i];
    if (!lv.dinv.p) {
      lv.dinv.alloc(lv.L.N);
      lv.lc.alloc(lv.L.N);
      lv.lac.alloc(lv.L.N);
      // A c on the Dirichlet rows is c there, which the network's output
      // mask makes zero; the row kernels write interior rows only, so these
      // entries are set once here
      fill<float>(lv.lac.p, 0.0f, lv.L.N, Y.ctx->st);
      lv.feat.alloc((size_t)lv.L.N * kNetIn);
    }
    if (!lv.have_dinv) {
      learned_diag_inv(lv.D, lv.v32.val.p, lv.dinv.p, Y.ctx->st);
      learned_features(lv.D, lv.v32.val.p, lv.feat.p, Y.ctx->st);
      lv.have_dinv = true;
    }
    if (net_reserve(Y.nets[i], lv.L.N, false)) Y.invalidate_graphs();
  }
}

// ------------------------------------------------------------- multigrid
template <typename T>
static void level_smooth(System &Y, Level &lv, int kind, double omega, int sweeps) {
  hipStream_t st = Y.ctx->st;
  Vecs<T> &v = vecs<T>(lv);
  // levels whose largest colour has at most this many rows sweep in one
  // workgroup; by default FP32 only: an FP64 sweep moves twice the bytes
  // through the one CU and is slower there than launch per colour
  const int64_t wg_default = sizeof(T) == 4 ? kSweepWgRows : 0;
  const int64_t wg_limit = Y.opt[21] == 0 ? wg_default : (int64_t)Y.opt[21];
  for (int s = 0; s < sweeps; s++) {
    if (kind == 1) {
      copy<T>(v.aux.p, v.x.p, lv.L.N, st);
      boundary<T>(lv.D, 2, v.x.p, nullptr, v.f.p, st);
      css_apply<T>(lv.D, 2, v.val.p, v.aux.p, v.f.p, v.x.p, omega, -1, st);
    } else if (lv.D.max_colour_rows <= wg_limit && sweep_wg_fits(lv.D, sizeof(T))) {
      sweep_wg<T>(lv.D, v.val.p, v.x.p, v.f.p, omega, kind == 3 ? 1 : 0, st);
    } else {
      // with omega = 1 a colour right after itself (the turn of an SSOR sweep,
      // the start of the next one) reproduces its rows bit for bit and is skipped
      const bool skip = kind == 3 && omega == 1.0;
      // Dirichlet rows x = f (no sweep changes them afterwards)
      if (s == 0) boundary<T>(lv.D, 2, v.x.p, nullptr, v.f.p, st);
      auto colour = [&](int c) {
        if constexpr (std::is_same<T, float>::value) {
          if (Y.half_smoother) {
            if (!lv.have_val16) throw std::logic_error("level_smooth: FP16 smoother values missing");
            css_sor_scaled(lv.D, lv.val16.p, lv.sdinv.p, v.x.p, v.f.p, omega, c, st);
            return;
          }
        }
        css_apply<T>(lv.D, 3, v.val.p, v.x.p, v.f.p, v.x.p, omega, c, st);
      };
      for (int c = (skip && s > 0) ? 1 : 0; c < lv.L.ncolours; c++) colour(c);
      if (kind == 3)
        for (int c = lv.L.ncolours - (skip ? 2 : 1); c >= 0; c--) colour(c);
    }
  }
}
// Learned post-smoothing of one level (replaces the post-sweeps of
// TMGLevel3D; no counterpart in the reference): with d = f - A x and
// u = D^-1 d, the network maps u / rms(u) to a correction c (rescaled by
// rms(u)); x += alpha c with the minimal-residual step
// alpha = max(0, (d, Ac) / (Ac, Ac)), so the level defect never grows in the
// 2-norm (the step-length idea of TMGLevel3D::StepLengthControl,
// MGLevel3D.C:908-960, applied to the network's correction).
static void learned_post(System &Y, Level &lv, LearnedNet &net, bool fresh = true, bool carry = false,
                         bool x_zero = false) {
  hipStream_t st = Y.ctx->st;
  Vecs<float> &v = lv.v32;
  const int64_t N = lv.L.N;
  double *sc = Y.lscal.p, *part = Y.ctx->part.p;
  const double sq = std::sqrt((double)lv.L.nactive);
  // x_zero (a fresh step on x = 0): the defect is f itself, bit for bit
  const float *d0 = fresh && x_zero ? v.f.p : v.d.p;
  if (fresh && !x_zero) {
    // defect of the interior rows; its Dirichlet entries meet zero weights
    // below (dinv = 0 there, A c = 0 there) and are finite (zero or f - x)
    css_apply<float>(lv.D, 1, v.val.p, v.x.p, v.f.p, v.d.p, 0.0, -1, st);
  }
  // the network's first layer builds its input from the features, u = D^-1 d
  // and the partials of ||u||^2
  LearnedInput in{lv.feat.p, lv.lc.p, nullptr, learned_input_parts(N)};
  if (fresh) {
    learned_input_sq(lv.dinv.p, d0, lv.lc.p, N, part, st);
    in.part = part;
  } else {
    // d, u = D^-1 d and the partials of ||u||^2 were left by the previous step
    in.part = part + 2 * kRedBlocks;
  }
  net_forward_correction(net, lv.L.M, sc, sq, lv.lc.p, st, &in);
  css_apply<float>(lv.D, 0, v.val.p, lv.lc.p, nullptr, lv.lac.p, 0.0, -1, st);  // Dirichlet entries stay 0
  if (carry)
    learned_update_carry(d0, v.d.p, lv.lac.p, lv.lc.p, v.x.p, lv.dinv.p, lv.lc.p, N, part, part + 2 * kRedBlocks,
                         st);
  else
    learned_update(d0, lv.lac.p, lv.lc.p, v.x.p, N, part, st);
}
// `steps` learned steps on a level: the first from the level's defect, the
// next ones from the defect the previous step updates (d - alpha A c: the
// same defect to rounding, without a product with A and a pass over d);
// carry_last: the last step leaves the updated defect too (interior rows)
// x_zero: the level's x is 0 (a coarse level entered after its Reset)
static void learned_smooth(System &Y, Level &lv, LearnedNet &net, int steps, bool carry_last = false,
                           bool x_zero = false) {
  for (int k = 0; k < steps; k++) learned_post(Y, lv, net, k == 0, carry_last || k + 1 < steps, k == 0 && x_zero);
}

// level i smooths with the network in this cycle
static bool learned_level(const System &Y, int i) {
  const int n = (int)Y.lev.size();
  return Y.learn_active && i >= n - Y.learned_levels && i < n - Y.learned_skip;
}
template <typename T>
// x_zero: x of level i is 0 on entry (the first visit after the Reset)
static void mg_cycle(System &Y, int i, bool post = true, bool x_zero = false) {
  hipStream_t st = Y.ctx->st;
  const double *o = Y.opt;
  Level &lv = *Y.lev[i];
  Vecs<T> &v = vecs<T>(lv);
  if (i == 0) {
    // coarse smoother ids other than 1, 2, 17 take the default branch of
    // MultiGrid3D.C's CoarseSmooth switch, SSOR (as 6 does in the
    // sequential build)
    const int cs = ((int)o[9] == 1 || (int)o[9] == 2 || (int)o[9] == 17) ? (int)o[9] : 3;
    if (cs == 17 && lv.ainv_n > 0) {
      gemv_coarse<T>(lv.ainv_n, (int)lv.L.N, lv.ainv.p, lv.iidx.p, v.f.p, v.x.p, st);
      boundary<T>(lv.D, 2, v.x.p, nullptr, v.f.p, st);
    } else if (cs >= 1 && cs <= 3 && sweep_wg_fits(lv.D, sizeof(T))) {
      // sweeps until the defect is reduced by SC_COARSE_RED_FACTOR_SCALAR
      // (opt[11]) or SC_COARSE_MAXIT_SCALAR (opt[10]) sweeps were made
      coarse_wg<T>(lv.D, v.val.p, v.x.p, v.f.p, o[8], cs, std::max(1, (int)o[10]), o[11], v.d.p, v.aux.p, st);
    } else {
      level_smooth<T>(Y, lv, cs == 17 ? 3 : cs, o[8], std::max(1, (int)o[10]));
    }
    return;
  }
  Level &cl = *Y.lev[i - 1];
  Vecs<T> &cv = vecs<T>(cl);
  const int kind = (int)o[5];
  // step length control (MultiGrid3D.C:465-471, 534-538, 562-566): keep u
  // and its defect before the pre-smoothing
  const bool slc = Y.slc == 2 || (Y.slc == 1 && i == (int)Y.lev.size() - 1);
  if (slc) {
    if (!v.sold.p || !v.sdef.p) throw std::logic_error("step length control buffers missing");
    copy<T>(v.sold.p, v.x.p, lv.L.N, st);
    css_apply<T>(lv.D, 1, v.val.p, v.x.p, v.f.p, v.sdef.p, 0.0, -1, st);
    boundary<T>(lv.D, 1, v.sdef.p, v.x.p, v.f.p, st);
  }
  bool pre_done = false;
  if constexpr (std::is_same<T, float>::value) {
    if (Y.learned_pre && learned_level(Y, i)) {
      // the steps leave the defect of their last update in v.d
      learned_smooth(Y, lv, Y.nets[i], std::max(1, (int)o[20]), true, x_zero);
      pre_done = true;
    }
  }
  if (!pre_done) {
    level_smooth<T>(Y, lv, kind, o[8], (int)o[6]);
    css_apply<T>(lv.D, 1, v.val.p, v.x.p, v.f.p, v.d.p, 0.0, -1, st);
  }
  // the defect's Dirichlet entries are not set: the restriction to an
  // interior coarse node reads interior fine nodes only (taps 2y-3..2y+3 and
  // 2y-1..2y+1 with 1 <= y <= Mc-2), and the coarse surface is set to 0
  // (CorrectDefect)
  restrict_<T>(cl.D, lv.D, cv.f.p, v.d.p, st, true);
  fill<T>(cv.x.p, (T)0, cl.L.N, st);                   // Reset
  // the first visit of the coarse level starts from its Reset (x = 0); the
  // learned pre-smoothing there takes d = f (restrict_ with zero_surface
  // leaves the coarse f zero on the Dirichlet surface, so the Dirichlet
  // entries of that defect are the ones A x = 0 would give)
  for (int j = 0; j < Y.rec[i]; j++) mg_cycle<T>(Y, i - 1, true, j == 0 && Y.x_zero_shortcut);
  if ((int)o[12] < 1) Y.rec[i] = 1;
  // prolongation, CorrectNodes and x += SC_GMG_DAMP e
  prolongate_add<T>(cl.D, lv.D, cv.x.p, v.x.p, o[13], st);
  if (post) {
    bool post_done = false;
    if constexpr (std::is_same<T, float>::value) {
      if (learned_level(Y, i)) {
        learned_smooth(Y, lv, Y.nets[i], std::max(1, (int)o[20]));
        post_done = true;
      }
    }
    if (!post_done) level_smooth<T>(Y, lv, kind, o[8], (int)o[7]);
  }
  if (slc) {
    // TMGLevel3D::StepLengthControl (MGLevel3D.C:908-960), MultiGrid3D.C:671-679:
    // e = u - uold, y = A e, omega = (def, y) / (y, y), u = uold + omega e
    double *num = Y.ctx->part.p + 3 * kRedBlocks + 16;
    axpy_diff<T>(v.aux.p, v.x.p, v.sold.p, lv.L.N, st);
    css_apply<T>(lv.D, 0, v.val.p, v.aux.p, nullptr, v.d.p, 0.0, -1, st);
    boundary<T>(lv.D, 0, v.d.p, v.aux.p, nullptr, st);
    dot<T>(v.sdef.p, v.d.p, lv.L.N, Y.ctx->part.p, num, 0, st);
    dot<T>(v.d.p, v.d.p, lv.L.N, Y.ctx->part.p, num + 1, 0, st);
    step_length_update<T>(v.x.p, v.sold.p, num, o[8], lv.L.N, st);
  }
}
// TMultiGrid3D::SetRecursion (MultiGrid3D.C:697-725): gamma cycles per level
static void set_recursion(System &Y) {
  const int L = (int)Y.lev.size() - 1;
  const int gam = (int)Y.opt[12];
  Y.rec.assign(Y.lev.size() + 1, 1);
  for (int k = 2; k <= L; k++) Y.rec[k] = gam > 0 ? gam : 2;
}
// one MG cycle on the finest level's f -> x (zero start, SetDirichletNodes)
template <typename T>
static void mg_cycle_from_zero(System &Y) {
  hipStream_t st = Y.ctx->st;
  const int L = (int)Y.lev.size() - 1;
  Level &lv = *Y.lev[L];
  Vecs<T> &v = vecs<T>(lv);
  fill<T>(v.x.p, (T)0, lv.L.N, st);
  boundary<T>(lv.D, 2, v.x.p, nullptr, v.f.p, st);  // SetDirichletNodes
  set_recursion(Y);
  mg_cycle<T>(Y, L);
}
// z = MG(v): TMultiGridScaIte::Iterate with zero start (MultiGridScaIte.C:89-130).
// The cycle (about a thousand kernels for 5 Q2 levels) is captured once into a
// HIP graph and replayed, so the GPU is not left waiting on launches.
template <typename T>
static void mg_prec(System &Y, T *z, const T *vin) {
  hipStream_t st = Y.ctx->st;
  Level &lv = *Y.lev.back();
  Vecs<T> &v = vecs<T>(lv);
  copy<T>(v.f.p, vin, lv.L.N, st);
  const int pi = std::is_same<T, double>::value ? 0 : 1;
  System::Graph &g = Y.graphs[pi];
  double key[16];
  for (int k = 0; k < 9; k++) key[k] = Y.opt[5 + k];
  const bool learned = pi == 1 && Y.learn_active;
  key[9] = learned ? Y.learned_levels : 0;
  key[10] = learned ? std::max(1.0, Y.opt[20]) : 0;
  key[11] = Y.opt[21];
  key[12] = learned ? Y.learned_skip : 0;
  key[13] = learned ? Y.learned_pre : 0;
  key[14] = Y.slc;
  key[15] = pi == 1 && Y.half_smoother ? 1 : 0;
  if (Y.use_graphs) {
    if (!g.valid || std::memcmp(key, g.key, sizeof(key)) != 0) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g.exec = nullptr;
      hipGraph_t graph;
      check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
      mg_cycle_from_zero<T>(Y);
      check(hipStreamEndCapture(st, &graph), "end capture");
      check(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0), "graph instantiate");
      (void)hipGraphDestroy(graph);
      std::memcpy(g.key, key, sizeof(key));
      g.valid = true;
    }
    check(hipGraphLaunch(g.exec, st), "graph launch");
  } else {
    mg_cycle_from_zero<T>(Y);
  }
  copy<T>(z, v.x.p, lv.L.N, st);
}

// y = A x on the finest level (stencil + Dirichlet identity rows); in FP64
// the caller's CSR matrix instead while pmx_iterate runs
template <typename T>
static void fine_matvec(System &Y, const T *x, T *y) {
  Level &lv = *Y.lev.back();
  if constexpr (std::is_same<T, double>::value) {
    if (Y.use_ucsr) {
      csr_spmv<double>((int)lv.L.N, Y.ucsr_rp.p, Y.ucsr_kc.p, Y.ucsr_val.p, x, y, Y.ctx->st);
      return;
    }
  }
  css_apply<T>(lv.D, 0, vecs<T>(lv).val.p, x, nullptr, y, 0.0, -1, Y.ctx->st);
  boundary<T>(lv.D, 0, y, x, nullptr, Y.ctx->st);
}
template <typename T>
static void fine_defect(System &Y, const T *x, const T *b, T *r) {
  Level &lv = *Y.lev.back();
  if constexpr (std::is_same<T, double>::value) {
    if (Y.use_ucsr) {
      csr_defect<double>((int)lv.L.N, Y.ucsr_rp.p, Y.ucsr_kc.p, Y.ucsr_val.p, x, b, r, Y.ctx->st);
      return;
    }
  }
  css_apply<T>(lv.D, 1, vecs<T>(lv).val.p, x, b, r, 0.0, -1, Y.ctx->st);
  boundary<T>(lv.D, 1, r, x, b, Y.ctx->st);
}

static double read_scalar(System &Y, const double *dev) {
  check(hipMemcpyAsync(Y.ctx->pinned, dev, sizeof(double), hipMemcpyDeviceToHost, Y.ctx->st), "d2h scalar");
  check(hipStreamSynchronize(Y.ctx->st), "sync");
  return Y.ctx->pinned[0];
}

// TFgmresIte::Iterate, flexible branch with modified Gram-Schmidt
// (src/FE/FgmresIte.C:262-625).  Vectors in T, Hessenberg/Givens in FP64.
// Returns iterations; sets *res_out to the last residual estimate.
template <typename T>
static int fgmres(System &Y, T *x, const T *b, int m, int maxit, double rmin, double red, double *res_out,
                  double *res0_out, int *status) {
  hipStream_t st = Y.ctx->st;
  Level &lv = *Y.lev.back();
  const int64_t N = lv.L.N;
  DevBuf<T> &kry = kry_buf<T>(Y);
  kry.alloc((size_t)(2 * (m + 1) + 1) * N);
  T *V = kry.p, *Z = kry.p + (size_t)(m + 1) * N, *w = kry.p + (size_t)2 * (m + 1) * N;
  const int ld = m + 1;
  Y.scal.alloc((size_t)ld * ld + 5 * ld + 8);
  double *H = Y.scal.p, *cs = H + ld * ld, *sn = cs + ld, *s = sn + ld, *yv = s + ld, *beta = yv + ld,
         *res = beta + 1, *tmp = res + 1;
  double *part = Y.ctx->part.p;
  fine_defect<T>(Y, x, b, V);
  dot<T>(V, V, N, part, beta, 1, st);
  double bet = read_scalar(Y, beta);
  double res0 = bet, r = bet;
  *res0_out = bet;
  *status = PMX_OK;
  if (bet <= rmin) { *res_out = bet; return 0; }
  int j = 1;
  while (j <= maxit) {
    scale_inv<T>(V, V, beta, N, st);
    fill<double>(s, 0.0, ld, st);
    set_scalar(s, beta, 1.0, st);
    fill<double>(H, 0.0, (int64_t)ld * ld, st);
    int i;
    for (i = 0; i < m && j <= maxit; i++, j++) {
      T *vi = V + (size_t)i * N, *zi = Z + (size_t)i * N;
      mg_prec<T>(Y, zi, vi);
      fine_matvec<T>(Y, zi, w);
      mgs_arnoldi<T>(w, V, N, i + 1, H + (size_t)i * ld, V + (size_t)(i + 1) * N, part, st);
      givens(H, cs, sn, s, i, ld, res, st);
      const double r_prev = r;
      r = read_scalar(Y, res);
      if (!(r == r)) { *status = PMX_ERR_DIVERGED; *res_out = r; return j; }
      if (Y.learn_active && r > kLearnStall * r_prev) {
        // a cycle with the learned post-smoother failed to reduce the residual:
        // plain post-smoothing from here on (the flexible method allows it)
        Y.learn_active = false;
        Y.fallbacks++;
      }
      if (r < rmin || r < red * res0) {
        backsolve(H, s, yv, i, ld, st);
        multi_axpy<T, T>(x, Z, yv, i + 1, N, N, st);
        *res_out = r;
        return j;
      }
      if (r > 1e10 * res0) { *status = PMX_ERR_DIVERGED; *res_out = r; return j; }
    }
    // the columns of this cycle (i < m when the iteration limit ended it; the
    // reference back-substitutes all restart columns there, FgmresIte.C:563,
    // and reads Hessenberg columns the cycle did not fill)
    backsolve(H, s, yv, i - 1, ld, st);
    multi_axpy<T, T>(x, Z, yv, i, N, N, st);
    fine_defect<T>(Y, x, b, V);
    dot<T>(V, V, N, part, beta, 1, st);
    bet = read_scalar(Y, beta);
    res0 = r = bet;
    if (bet <= rmin) break;
  }
  *res_out = r;
  if (j > maxit) *status = PMX_ERR_NOCONV;
  (void)tmp;
  return std::min(j, maxit);
}

// TFixedPointIte::Iterate (src/FE/FixedPointIte.C:176-300)
template <typename T>
static int richardson(System &Y, T *x, const T *b, int maxit, double rmin, double red, double *res_out,
                      double *res0_out, int *status) {
  hipStream_t st = Y.ctx->st;
  Level &lv = *Y.lev.back();
  const int64_t N = lv.L.N;
  DevBuf<T> &kry = kry_buf<T>(Y);
  kry.alloc(2 * N);
  T *d = kry.p, *z = kry.p + N;
  Y.scal.alloc(64);
  double *res = Y.scal.p;
  fine_defect<T>(Y, x, b, d);
  dot<T>(d, d, N, Y.ctx->part.p, res, 1, st);
  double r = read_scalar(Y, res);
  const double r0 = r;
  *res0_out = r0;
  *status = PMX_OK;
  int it = 0;
  if (r > rmin)
    for (it = 1; it <= maxit; it++) {
      mg_prec<T>(Y, z, d);
      axpy<T>(x, z, nullptr, 1.0, 1.0, N, st);
      fine_defect<T>(Y, x, b, d);
      dot<T>(d, d, N, Y.ctx->part.p, res, 1, st);
      r = read_scalar(Y, res);
      if (r <= r0 * red || r <= rmin) break;
      if (!(r == r) || r > 1e10 * r0) { *status = PMX_ERR_DIVERGED; break; }
    }
  if (it > maxit) { it = maxit; *status = PMX_ERR_NOCONV; }
  *res_out = r;
  return it;
}

// TBcgs::Iterate (src/FE/Bcgs.C:111-304); flexible selects v = A y
template <typename T>
static int bicgstab(System &Y, T *x, const T *b, int maxit, double rmin, double red, bool flexible,
                    double *res_out, double *res0_out, int *status) {
  hipStream_t st = Y.ctx->st;
  Level &lv = *Y.lev.back();
  const int64_t N = lv.L.N;
  DevBuf<T> &kry = kry_buf<T>(Y);
  kry.alloc(6 * N);
  T *r = kry.p, *rt = r + N, *p = rt + N, *yv = p + N, *v = yv + N, *t = v + N;
  Y.scal.alloc(64);
  double *sc = Y.scal.p;  // 0 rho, 1 delta, 2 nrm, 3 tt, 4 tr
  double *part = Y.ctx->part.p;
  fine_defect<T>(Y, x, b, r);
  check(hipMemcpyAsync(rt, r, sizeof(T) * N, hipMemcpyDeviceToDevice, st), "copy");
  dot<T>(r, rt, N, part, sc + 2, 1, st);
  const double dn0 = read_scalar(Y, sc + 2);
  *res0_out = dn0;
  *status = PMX_OK;
  double dn = dn0, rho_last = 1, alpha = 0, omega = 0;
  if (dn0 <= rmin) { *res_out = dn0; return 0; }
  int i;
  for (i = 0; i < maxit; i++) {
    dot<T>(r, rt, N, part, sc, 0, st);
    const double rho = read_scalar(Y, sc);
    if (std::fabs(rho) < 1e-50) { *status = PMX_ERR_NOCONV; break; }
    if (i == 0) check(hipMemcpyAsync(p, r, sizeof(T) * N, hipMemcpyDeviceToDevice, st), "copy");
    else {
      const double beta = (rho / rho_last) * (alpha / omega);
      // p = beta p + r - beta omega v
      fill<T>(t, (T)0, N, st);
      axpy<T>(t, p, nullptr, 1.0, beta, N, st);
      axpy<T>(t, r, nullptr, 1.0, 1.0, N, st);
      axpy<T>(t, v, nullptr, 1.0, -beta * omega, N, st);
      check(hipMemcpyAsync(p, t, sizeof(T) * N, hipMemcpyDeviceToDevice, st), "copy");
    }
    rho_last = rho;
    mg_prec<T>(Y, yv, p);
    if (flexible) fine_matvec<T>(Y, yv, v);
    else check(hipMemcpyAsync(v, yv, sizeof(T) * N, hipMemcpyDeviceToDevice, st), "copy");
    dot<T>(rt, v, N, part, sc + 1, 0, st);
    double delta = read_scalar(Y, sc + 1);
    if (std::fabs(delta) < 1e-50) delta = 1e-50;
    alpha = rho / delta;
    axpy<T>(r, v, nullptr, 1.0, -alpha, N, st);
    axpy<T>(x, flexible ? yv : p, nullptr, 1.0, alpha, N, st);
    dot<T>(r, r, N, part, sc + 2, 1, st);
    dn = read_scalar(Y, sc + 2);
    if (dn < dn0 * red || dn < rmin) break;
    mg_prec<T>(Y, yv, r);
    if (flexible) fine_matvec<T>(Y, yv, t);
    else check(hipMemcpyAsync(t, yv, sizeof(T) * N, hipMemcpyDeviceToDevice, st), "copy");
    dot<T>(t, t, N, part, sc + 3, 0, st);
    double tt = read_scalar(Y, sc + 3);
    if (std::fabs(tt) < 1e-50) tt = 1e-50;
    dot<T>(t, r, N, part, sc + 4, 0, st);
    omega = read_scalar(Y, sc + 4) / tt;
    axpy<T>(x, flexible ? yv : r, nullptr, 1.0, omega, N, st);
    axpy<T>(r, t, nullptr, 1.0, -omega, N, st);
    dot<T>(r, r, N, part, sc + 2, 1, st);
    dn = read_scalar(Y, sc + 2);
    if (dn < dn0 * red || dn < rmin) break;
    if (!(dn == dn) || dn > 1e10 * dn0) { *status = PMX_ERR_DIVERGED; break; }
  }
  if (i >= maxit) *status = PMX_ERR_NOCONV;
  *res_out = dn;
  return std::min(i + 1, maxit);
}

template <typename T>
static int krylov(System &Y, T *x, const T *b, int maxit, double rmin, double red, double *res, double *res0,
                  int *status) {
  const int solver = (int)Y.opt[0];
  if (solver == 16) return fgmres<T>(Y, x, b, std::max(1, (int)Y.opt[1]), maxit, rmin, red, res, res0, status);
  if (solver == 11) return richardson<T>(Y, x, b, maxit, rmin, red, res, res0, status);
  if (solver == 13) return bicgstab<T>(Y, x, b, maxit, rmin, red, Y.opt[14] != 0, res, res0, status);
  mg_prec<T>(Y, x, b);
  *res = *res0 = 0;
  *status = PMX_OK;
  return 1;
}

// One residual-loss gradient of the network on level lv for the scaled input
// u (FP32 lattice): c = mask(net(u)), loss = mean over interior nodes of
// (u - D^-1 A c)^2 -- the part of the scaled defect that the correction
// leaves.  Accumulates into net.grads; the loss goes to *loss_dev.
// the level's transposed FP32 values for learned_res_step (from the current
// FP32 values; call after they change)
static void learned_transpose(System &Y, Level &lv) {
  lv.valT.alloc(lv.L.nvals);
  css_transpose_values(lv.D, lv.v32.val.p, lv.valT.p, Y.ctx->st);
}
static void learned_res_step(System &Y, Level &lv, LearnedNet &net, const float *u, float *loss_dev) {
  hipStream_t st = Y.ctx->st;
  if (!lv.valT.p) throw std::logic_error("learned_res_step: transposed values missing");
  const int M = lv.L.M;
  Vecs<float> &v = lv.v32;
  learned_build_input(net, u, nullptr, 1.0, lv.feat.p, lv.L.N, st);
  net_forward_act(net, M, true, st);
  learned_mask(net.out, M, lv.lc.p, st);
  css_apply<float>(lv.D, 0, v.val.p, lv.lc.p, nullptr, lv.lac.p, 0.0, -1, st);  // Dirichlet entries stay 0
  learned_res_loss(net, M, u, lv.dinv.p, lv.lac.p, v.aux.p, st);
  css_transpose(lv.D, lv.valT.p, v.aux.p, v.d.p, st);
  net_backward_from(net, M, v.d.p, st);
  learned_loss_sum(net, loss_dev, st);
}

// Training of the learned post-smoother on (input, target) pairs from plain
// cycles of the homogeneous problem (f = 0, exact solution 0): a random
// interior start x0, then pre-smoothing and the coarse-grid correction of one
// cycle on the finest level give x; the ideal post-correction is -x, the
// network input the scaled D^-1 (f - A x).  Adam on the mean squared error.
// topt: [0] samples per level, [1] Adam steps per level (levels interleaved),
// [2] learning rate, [3] seed, [4] finest levels using the network (default:
// all but the coarsest), [5] loss: 0 squared error against -x, 1 the scaled
// defect left after the correction (learned_res_step), [6] samples: 0 random
// start and f = 0, 1 zero start and a random f, 2 half each, [7] the finest
// levels among [4] that are not trained (they keep the plain smoother), [8]
// samples per Adam step (the batch, default 1: the step takes the mean of
// their gradients; sample k of level j's list of [0] is used at step k / B,
// cyclically).
// tinfo: [0] mean loss of the first 10 steps of the finest trained level, [1]
// of its last ones, [2] seconds, [3] steps, [4 + 2j] / [5 + 2j] the same for
// level L - j.
static void learned_train(System &Y, const double *to, double *ti) {
  hipStream_t st = Y.ctx->st;
  const int ns = to[0] > 0 ? (int)to[0] : 16;
  const int steps = to[1] > 0 ? (int)to[1] : 400;
  const double lr = to[2] > 0 ? to[2] : 2e-3;
  const uint32_t seed = (uint32_t)to[3];
  const int L = (int)Y.lev.size() - 1;
  Y.learned_levels = to[4] > 0 ? std::min((int)to[4], L) : L;
  const bool residual_loss = to[5] != 0.0;
  const int data_mode = (int)to[6];  // 0 random error, f = 0; 1 x = 0, random f; 2 both
  const int nl = Y.learned_levels;
  const int j0 = std::max(0, std::min((int)to[7], nl - 1));  // finest levels left untrained
  const int batch = to[8] > 0 ? (int)to[8] : 1;
  if (batch > 4096) throw std::invalid_argument("learned_train: batch above 4096");
  Y.trained_levels = nl;
  Y.learn_active = false;
  ensure_learned(Y);
  for (int j = j0; j < nl; j++) {
    LearnedNet &net = Y.nets[L - j];
    if (!net.ready) net_init(net, (seed ^ 0x5eedu) + 31u * (uint32_t)j, st);
    if (net_reserve(net, Y.lev[L - j]->L.N, true)) Y.invalidate_graphs();
  }
  // samples of level L - j at off[j] + k N_j
  std::vector<size_t> off(nl + 1, 0);
  for (int j = 0; j < nl; j++) off[j + 1] = off[j] + (j >= j0 ? (size_t)ns * Y.lev[L - j]->L.N : 0);
  DevBuf<float> du, dt;
  du.alloc(off[nl]);
  if (!residual_loss) dt.alloc(off[nl]);
  const int nt = nl - j0;  // trained levels
  const int total = steps * nt;
  DevBuf<float> losses;
  losses.alloc((size_t)total * batch);
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "event");
  check(hipEventCreate(&e1), "event");
  check(hipEventRecord(e0, st), "event");
  set_recursion(Y);
  for (int j = j0; j < nl; j++) {
    Level &lv = *Y.lev[L - j];
    Vecs<float> &v = lv.v32;
    const int64_t N = lv.L.N;
    const int M = lv.L.M;
    const double sq = std::sqrt((double)lv.L.nactive);
    for (int k = 0; k < ns; k++) {
      const uint32_t sk = seed * 7919u + 104729u * (uint32_t)j + (uint32_t)k;
      const bool zero_start = data_mode == 1 || (data_mode == 2 && (k & 1));
      if (zero_start) {
        // as inside a preconditioner application: x = 0, a random right-hand side
        fill<float>(v.x.p, 0.0f, N, st);
        learned_random(v.f.p, M, sk, st);
      } else {
        // a random error of the homogeneous problem
        learned_random(v.x.p, M, sk, st);
        fill<float>(v.f.p, 0.0f, N, st);
      }
      mg_cycle<float>(Y, L - j, false);
      css_apply<float>(lv.D, 1, v.val.p, v.x.p, v.f.p, v.d.p, 0.0, -1, st);
      boundary<float>(lv.D, 1, v.d.p, v.x.p, v.f.p, st);
      learned_input(lv.dinv.p, v.d.p, lv.lc.p, N, st);
      dot<float>(lv.lc.p, lv.lc.p, N, Y.ctx->part.p, Y.lscal.p, 1, st);
      learned_target(lv.lc.p, M, Y.lscal.p, sq, 1.0f, du.p + off[j] + (size_t)k * N, st);
      if (!residual_loss) learned_target(v.x.p, M, Y.lscal.p, sq, -1.0f, dt.p + off[j] + (size_t)k * N, st);
    }
  }
  if (residual_loss)
    for (int j = j0; j < nl; j++) learned_transpose(Y, *Y.lev[L - j]);
  for (int k = 0; k < total; k++) {
    const int j = j0 + k % nt, kk = k / nt;
    Level &lv = *Y.lev[L - j];
    LearnedNet &net = Y.nets[L - j];
    const double lrk = lr * (0.1 + 0.45 * (1.0 + std::cos(M_PI * kk / std::max(1, steps - 1))));
    net_zero_grads(net, st);
    for (int b = 0; b < batch; b++) {
      // the gradients of the batch accumulate in net.grads
      const int smp = (int)(((int64_t)kk * batch + b) % ns);
      const size_t o = off[j] + (size_t)smp * lv.L.N;
      float *lk = losses.p + (size_t)k * batch + b;
      if (residual_loss) {
        learned_res_step(Y, lv, net, du.p + o, lk);
      } else {
        learned_build_input(net, du.p + o, nullptr, 1.0, lv.feat.p, lv.L.N, st);
        net_forward_act(net, lv.L.M, true, st);
        net_backward(net, lv.L.M, dt.p + o, lk, st);
      }
    }
    net_adam(net, (float)lrk, st, 1.0f / (float)batch);
  }
  check(hipEventRecord(e1, st), "event");
  check(hipEventSynchronize(e1), "event");
  float ms = 0;
  check(hipEventElapsedTime(&ms, e0, e1), "event");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::vector<float> hb((size_t)total * batch), h(total, 0.0f);
  check(hipMemcpy(hb.data(), losses.p, sizeof(float) * hb.size(), hipMemcpyDeviceToHost), "losses");
  for (int k = 0; k < total; k++) {
    double a = 0;
    for (int b = 0; b < batch; b++) a += hb[(size_t)k * batch + b];
    h[k] = (float)(a / batch);
  }
  // mean loss of the first / last 10 steps of each level's network
  if (ti) {
    const int w = std::min(10, steps);
    for (int j = j0; j < nl && 4 + 2 * j + 1 < 16; j++) {
      double a = 0, b = 0;
      for (int q = 0; q < w; q++) {
        a += h[(size_t)q * nt + (j - j0)];
        b += h[(size_t)(steps - 1 - q) * nt + (j - j0)];
      }
      if (j == j0) { ti[0] = a / w; ti[1] = b / w; }
      ti[4 + 2 * j] = a / w;
      ti[5 + 2 * j] = b / w;
    }
    ti[2] = ms * 1e-3;
    ti[3] = total;
  }
  Y.invalidate_graphs();
}

// Solve A sol = b on the finest level; sol/b are FP64 lattice vectors on the device.
static int solve_device(System &Y, double *info) {
  hipStream_t st = Y.ctx->st;
  // PMX_EAGER=1 launches kernel by kernel (profilers that cannot follow graph capture)
  const char *eager = std::getenv("PMX_EAGER");
  Y.use_graphs = Y.opt[18] == 0 && !(eager && eager[0] == '1');
  Level &lv = *Y.lev.back();
  const int64_t N = lv.L.N;
  // opt[15]: 0 FP64, 1 FP32, 2 FP32 + learned; 3 / 4 = 1 / 2 with the
  // per-colour SOR/SSOR sweeps on row-scaled FP16 values
  const int prec_opt = (int)Y.opt[15];
  if (prec_opt < 0 || prec_opt > 4) throw std::invalid_argument("opt[15] (precision) must be 0..4");
  const int prec = prec_opt >= 3 ? prec_opt - 2 : prec_opt;
  Y.half_smoother = prec_opt >= 3;
  // ILU (id 4 of MultiGrid3D.C's switches, TMGLevel3D::ILU) is not built;
  // every other id follows the switches' branches (default: SOR / SSOR)
  if ((int)Y.opt[5] == 4 || (int)Y.opt[9] == 4)
    throw std::invalid_argument("smoother id 4 (ILU) is not implemented");
  // PMX_NO_XZERO=1: a coarse level's first learned step forms its defect with
  // the operator instead of taking d = f on x = 0 (the two agree bit for bit)
  const char *nxz = std::getenv("PMX_NO_XZERO");
  const bool xz = !(nxz && nxz[0] == '1');
  if (xz != Y.x_zero_shortcut) {
    Y.x_zero_shortcut = xz;
    Y.invalidate_graphs();
  }
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "event");
  check(hipEventCreate(&e1), "event");
  check(hipEventRecord(e0, st), "event");
  int status = PMX_OK, its = 0, outer = 0, fp64_fallback = 0;
  double res = 0, res0 = 0;
  const int maxit = (int)Y.opt[2];
  const double rel_target = Y.opt[16];
  double *nr = Y.ctx->part.p + 3 * kRedBlocks;
  if (prec == 0) {
    double rmin = Y.opt[3];
    if (rel_target > 0) {
      // absolute SC_LIN_RES_NORM_MIN = rel_target * ||b - A x0|| over all rows
      Y.r64.alloc(N);
      fine_defect<double>(Y, Y.sol.p, Y.b.p, Y.r64.p);
      dot<double>(Y.r64.p, Y.r64.p, N, Y.ctx->part.p, nr, 1, st);
      rmin = std::max(rmin, rel_target * read_scalar(Y, nr));
    }
    its = krylov<double>(Y, Y.sol.p, Y.b.p, maxit, rmin, Y.opt[4], &res, &res0, &status);
    outer = 1;
  } else {
    // FP64 defect correction around an FP32 Krylov/MG solve; precision 2
    // replaces the post-smoother of the finest levels by the learned one
    ensure_val32(Y);
    if (Y.half_smoother) ensure_val16(Y);
    Y.fallbacks = 0;
    if (prec == 2) {
      Y.learned_levels = std::min(Y.opt[19] > 0 ? (int)Y.opt[19] : Y.trained_levels, (int)Y.lev.size() - 1);
      Y.learned_skip = std::max(0, std::min((int)Y.opt[22], Y.learned_levels));
      Y.learned_pre = Y.opt[23] != 0;
      for (int i = (int)Y.lev.size() - Y.learned_levels; i < (int)Y.lev.size() - Y.learned_skip; i++)
        if (!Y.nets[i].ready) throw std::logic_error("learned smoother: train or load parameters first");
      ensure_learned(Y);
      Y.learn_active = true;
    }
    Y.r64.alloc(N);
    Y.sol_try.alloc(N);
    Y.r_try.alloc(N);
    DevBuf<float> r32, e32;
    r32.alloc(N);
    e32.alloc(N);
    const double rmin = Y.opt[3], red = Y.opt[4];
    const double inner_red = Y.opt[17] > 0 ? Y.opt[17] : 1e-4;
    fine_defect<double>(Y, Y.sol.p, Y.b.p, Y.r64.p);
    dot<double>(Y.r64.p, Y.r64.p, N, Y.ctx->part.p, nr, 1, st);
    res0 = res = read_scalar(Y, nr);
    const double target = std::max(std::max(rmin, red * res0), rel_target * res0);
    while (res > target && its < maxit) {
      convert<float, double>(r32.p, Y.r64.p, N, st);
      fill<float>(e32.p, 0.0f, N, st);
      double ri, ri0;
      int st_in;
      const double want = std::max(inner_red, 0.5 * target / res);
      its += krylov<float>(Y, e32.p, r32.p, maxit - its, 0.0, want, &ri, &ri0, &st_in);
      outer++;
      // the correction is accepted only if the inner solve ended with a
      // finite estimate and the FP64 residual of the corrected iterate is
      // finite and smaller; otherwise it is discarded
      double rn = HUGE_VAL;
      const bool inner_ok = st_in != PMX_ERR_DIVERGED && std::isfinite(ri);
      if (inner_ok) {
        copy<double>(Y.sol_try.p, Y.sol.p, N, st);
        add_convert<double, float>(Y.sol_try.p, e32.p, N, st);
        fine_defect<double>(Y, Y.sol_try.p, Y.b.p, Y.r_try.p);
        dot<double>(Y.r_try.p, Y.r_try.p, N, Y.ctx->part.p, nr, 1, st);
        rn = read_scalar(Y, nr);
      }
      if (inner_ok && std::isfinite(rn) && rn < res) {
        std::swap(Y.sol.p, Y.sol_try.p);
        std::swap(Y.r64.p, Y.r_try.p);
        res = rn;
        continue;
      }
      if (Y.learn_active) {
        // the cycles with the learned smoother did not reduce the FP64
        // residual: the same correction again with the plain smoother
        Y.learn_active = false;
        Y.fallbacks++;
        continue;
      }
      if (its >= maxit) break;
      // the FP32 solve itself failed: the rest of the solve in FP64 from the
      // last accepted iterate (FP64 operator, cycle and Krylov vectors)
      fp64_fallback = 1;
      double r64res, r64res0;
      int st64;
      its += krylov<double>(Y, Y.sol.p, Y.b.p, std::max(1, maxit - its), target, 0.0, &r64res, &r64res0, &st64);
      fine_defect<double>(Y, Y.sol.p, Y.b.p, Y.r64.p);
      dot<double>(Y.r64.p, Y.r64.p, N, Y.ctx->part.p, nr, 1, st);
      res = read_scalar(Y, nr);
      outer++;
      if (st64 == PMX_ERR_DIVERGED || !std::isfinite(res)) status = PMX_ERR_DIVERGED;
      break;
    }
    if (res > target && status == PMX_OK) status = PMX_ERR_NOCONV;
  }
  check(hipEventRecord(e1, st), "event");
  check(hipEventSynchronize(e1), "event");
  float ms = 0;
  check(hipEventElapsedTime(&ms, e0, e1), "event");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (info) {
    info[0] = its; info[1] = res; info[2] = res0; info[3] = outer; info[4] = ms * 1e-3;
    info[5] = prec == 2 ? Y.fallbacks : 0;
    info[6] = fp64_fallback;
  }
  Y.learn_active = false;
  return status;
}

static Level &level_of(System &Y, int level) {
  if (level < 0) level += (int)Y.lev.size();
  if (level < 0 || level >= (int)Y.lev.size()) throw std::invalid_argument("level out of range");
  return *Y.lev[level];
}
// the level's CSR structure, built on the device (device_csr); the host copy
// (Numbering::rowptr / kcol) only where a caller needs host arrays
static void ensure_csr_dev(System &Y, Level &lv) {
  if (!lv.rowptr.p) {
    lv.rowptr.alloc(lv.L.N + 1);
    lv.kcol.alloc(csr_nnz(lv.L));
    device_csr(lv.L, lv.lat2dof.p, lv.dof2lat.p, lv.rowptr.p, lv.kcol.p, Y.ctx->st);
  }
}
static void ensure_csr_host(System &Y, Level &lv) {
  if (lv.Nb.have_csr) return;
  ensure_csr_dev(Y, lv);
  lv.Nb.rowptr.resize(lv.L.N + 1);
  lv.Nb.kcol.resize(csr_nnz(lv.L));
  check(hipMemcpyAsync(lv.Nb.rowptr.data(), lv.rowptr.p, sizeof(int) * lv.Nb.rowptr.size(), hipMemcpyDeviceToHost,
                       Y.ctx->st),
        "csr download");
  check(hipMemcpyAsync(lv.Nb.kcol.data(), lv.kcol.p, sizeof(int) * lv.Nb.kcol.size(), hipMemcpyDeviceToHost,
                       Y.ctx->st),
        "csr download");
  check(hipStreamSynchronize(Y.ctx->st), "csr download");
  lv.Nb.have_csr = true;
}
// host (reference numbering) <-> device lattice vector
template <typename T>
static void to_lattice(System &Y, Level &lv, const double *host, T *dev) {
  DevBuf<double> tmp;
  tmp.upload(host, lv.L.N, Y.ctx->st);
  DevBuf<double> lat;
  lat.alloc(lv.L.N);
  gather<double, double>(lat.p, tmp.p, lv.lat2dof.p, lv.L.N, Y.ctx->st);
  if (std::is_same<T, double>::value)
    check(hipMemcpyAsync(dev, lat.p, sizeof(double) * lv.L.N, hipMemcpyDeviceToDevice, Y.ctx->st), "copy");
  else
    convert<float, double>((float *)dev, lat.p, lv.L.N, Y.ctx->st);
  check(hipStreamSynchronize(Y.ctx->st), "sync");
}
template <typename T>
static void from_lattice(System &Y, Level &lv, const T *dev, double *host) {
  DevBuf<double> lat, tmp;
  lat.alloc(lv.L.N);
  tmp.alloc(lv.L.N);
  if (std::is_same<T, double>::value)
    check(hipMemcpyAsync(lat.p, dev, sizeof(double) * lv.L.N, hipMemcpyDeviceToDevice, Y.ctx->st), "copy");
  else
    convert<double, float>(lat.p, (const float *)dev, lv.L.N, Y.ctx->st);
  gather<double, double>(tmp.p, lat.p, lv.dof2lat.p, lv.L.N, Y.ctx->st);
  check(hipMemcpyAsync(host, tmp.p, sizeof(double) * lv.L.N, hipMemcpyDeviceToHost, Y.ctx->st), "d2h");
  check(hipStreamSynchronize(Y.ctx->st), "sync");
}

}  // namespace pmx

using namespace pmx;

struct pmx_ctx { Ctx c; };
struct pmx_system { System s; };

// ---- scratch of the allocation-free entry point (pmx_defect_noalloc) ------
// The partial sums of the defect norm live in this file-scope device array
// (one copy per GPU, held until the process ends) instead of memory the
// library allocates: one partial per kDotChunk rows, so levels up to
// kNoAllocMaxRows rows fit.  The array's name is a host-side stand-in: host
// code uses the address hipGetSymbolAddress returns for the current device,
// kept per device, never the name itself.
namespace pmx {
constexpr int kNoAllocParts = 4096;
constexpr int64_t kNoAllocMaxRows = (int64_t)kNoAllocParts * kDotChunk;  // 8,388,608
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocParts];
static double *g_scratch_ptr[kNoAllocMaxDevices] = {nullptr};
static std::mutex g_scratch_mutex;
// the device address of `scratch` on the context's device (looked up once per
// device, with that device current)
static double *noalloc_scratch(const Ctx &c) {
  if (c.device < 0 || c.device >= kNoAllocMaxDevices) throw std::invalid_argument("device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[c.device]) {
    int cur = 0;
    check(hipGetDevice(&cur), "hipGetDevice");
    if (cur != c.device) check(hipSetDevice(c.device), "hipSetDevice");
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != c.device) (void)hipSetDevice(cur);
    check(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[c.device] = static_cast<double *>(ptr);
  }
  return g_scratch_ptr[c.device];
}
}  // namespace pmx

#define PMX_TRY(...)                                         \
  try {                                                      \
    __VA_ARGS__;                                                    \
    return PMX_OK;                                           \
  } catch (const std::invalid_argument &e) {                 \
    g_err = e.what();                                        \
    return PMX_ERR_ARG;                                      \
  } catch (const std::logic_error &e) {                      \
    g_err = e.what();                                        \
    return PMX_ERR_STATE;                                    \
  } catch (const std::exception &e) {                        \
    g_err = e.what();                                        \
    return PMX_ERR_HIP;                                      \
  }

extern "C" {

const char *pmx_last_error(void) { return g_err.c_str(); }

int pmx_ctx_create(int device, pmx_ctx **out) {
  PMX_TRY({
    if (!out) throw std::invalid_argument("out is NULL");
    int n = 0;
    check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) throw std::invalid_argument("no such device");
    check(hipSetDevice(device), "hipSetDevice");
    auto *c = new pmx_ctx();
    c->c.device = device;
    check(hipStreamCreateWithFlags(&c->c.st, hipStreamNonBlocking), "stream");
    c->c.part.alloc(3 * kRedBlocks + 64);
    check(hipHostMalloc((void **)&c->c.pinned, 64 * sizeof(double), hipHostMallocDefault), "pinned");
    *out = c;
  })
}
int pmx_ctx_destroy(pmx_ctx *ctx) {
  PMX_TRY({
    if (!ctx) return PMX_OK;
    if (ctx->c.systems > 0) throw std::logic_error("pmx_ctx_destroy: destroy the context's systems first");
    (void)hipStreamSynchronize(ctx->c.st);
    ctx->c.part.release();
    if (ctx->c.pinned) (void)hipHostFree(ctx->c.pinned);
    (void)hipStreamDestroy(ctx->c.st);
    delete ctx;
  })
}
int pmx_ctx_sync(pmx_ctx *ctx) { PMX_TRY({ check(hipStreamSynchronize(ctx->c.st), "sync"); }) }

// host-only numbering / structure (no device needed)
int pmx_host_space_dims(int nref, int order, long long *d) {
  PMX_TRY({
    Layout L;
    Numbering Nb;
    build_layout(L, nref, order, kBlock);
    build_numbering(Nb, L);
    build_csr(Nb, L);
    d[0] = L.N; d[1] = L.nactive; d[2] = Nb.ncells; d[3] = Nb.rowptr[L.N]; d[4] = Nb.nb; d[5] = L.M;
    d[6] = L.nvals; d[7] = L.ncolours;
  })
}
int pmx_host_space_arrays(int nref, int order, int *bi, int *gn, int *rp, int *kc) {
  PMX_TRY({
    Layout L;
    Numbering Nb;
    build_layout(L, nref, order, kBlock);
    build_numbering(Nb, L);
    if (bi) std::memcpy(bi, Nb.begin_index.data(), sizeof(int) * Nb.begin_index.size());
    if (gn) std::memcpy(gn, Nb.global_numbers.data(), sizeof(int) * Nb.global_numbers.size());
    if (rp || kc) {
      build_csr(Nb, L);
      if (rp) std::memcpy(rp, Nb.rowptr.data(), sizeof(int) * Nb.rowptr.size());
      if (kc) std::memcpy(kc, Nb.kcol.data(), sizeof(int) * Nb.kcol.size());
    }
  })
}

int pmx_host_write_vtk(int nref, int order, int nvar, const char *const *names, const double *const *values,
                       const char *path) {
  PMX_TRY({
    if (!path || nvar < 0 || (nvar > 0 && (!names || !values)))
      throw std::invalid_argument("NULL argument or negative nvar");
    for (int k = 0; k < nvar; k++) {
      if (!names[k] || !values[k]) throw std::invalid_argument("NULL name or values");
      for (const char *c = names[k]; *c; c++)
        if (*c == ' ' || *c == '\n' || *c == '\t') throw std::invalid_argument("VTK names carry no blanks");
      if (!*names[k]) throw std::invalid_argument("empty name");
    }
    Layout L;
    Numbering Nb;
    build_layout(L, nref, order, kBlock);
    build_numbering(Nb, L);
    write_vtk(L, Nb, nvar, names, values, path);
  })
}

int pmx_system_create(pmx_ctx *ctx, int nref0, int nlev, int order, pmx_system **out) {
  PMX_TRY({
    if (!ctx || !out) throw std::invalid_argument("NULL argument");
    if (order != 1 && order != 2) throw std::invalid_argument("order must be 1 or 2");
    if (nref0 < 1 || nlev < 1 || nref0 + nlev - 1 > (order == 2 ? 7 : 8))
      throw std::invalid_argument("refinement levels out of range");
    auto *h = new pmx_system();
    System &Y = h->s;
    Y.ctx = &ctx->c;
    Y.order = order;
    try {
      for (int l = 0; l < nlev; l++) {
        Y.lev.emplace_back(new Level());
        setup_level(Y, *Y.lev.back(), nref0 + l);
      }
      const Level &f = *Y.lev.back();
      Y.sol.alloc(f.L.N);
      Y.b.alloc(f.L.N);
      check(hipStreamSynchronize(ctx->c.st), "sync");
    } catch (...) {
      delete h;
      throw;
    }
    ctx->c.systems++;
    *out = h;
  })
}
int pmx_system_destroy(pmx_system *sys) {
  PMX_TRY({
    if (!sys) return PMX_OK;
    (void)hipStreamSynchronize(sys->s.ctx->st);
    sys->s.ctx->systems--;
    delete sys;
  })
}
int pmx_system_dims(pmx_system *sys, int level, long long *d) {
  PMX_TRY({
    Level &lv = level_of(sys->s, level);
    d[0] = lv.L.N; d[1] = lv.L.nactive; d[2] = lv.Nb.ncells; d[3] = csr_nnz(lv.L);
    d[4] = lv.Nb.nb; d[5] = lv.L.M; d[6] = lv.L.nvals; d[7] = lv.L.ncolours;
  })
}
int pmx_space_numbering(pmx_system *sys, int level, int *bi, int *gn) {
  PMX_TRY({
    Level &lv = level_of(sys->s, level);
    if (bi) std::memcpy(bi, lv.Nb.begin_index.data(), sizeof(int) * lv.Nb.begin_index.size());
    if (gn) std::memcpy(gn, lv.Nb.global_numbers.data(), sizeof(int) * lv.Nb.global_numbers.size());
  })
}
int pmx_space_structure(pmx_system *sys, int level, int *rp, int *kc) {
  PMX_TRY({
    Level &lv = level_of(sys->s, level);
    if (rp || kc) ensure_csr_host(sys->s, lv);
    if (rp) std::memcpy(rp, lv.Nb.rowptr.data(), sizeof(int) * lv.Nb.rowptr.size());
    if (kc) std::memcpy(kc, lv.Nb.kcol.data(), sizeof(int) * lv.Nb.kcol.size());
  })
}
int pmx_space_colours(pmx_system *sys, int level, int *colour) {
  PMX_TRY({
    Level &lv = level_of(sys->s, level);
    const int M = lv.L.M;
    for (int64_t r = 0; r < lv.L.N; r++) {
      const int l = lv.Nb.dof2lat[r];
      colour[r] = r < lv.L.nactive ? node_colour(lv.L, l % M, (l / M) % M, l / (M * M)) : -1;
    }
  })
}

namespace pmx {
// The coefficient callbacks of a host caller (CoeffFct3D, BoundCondFunct3D,
// BoundValueFunct3D of include/General/Constants.h:128-129,148).
struct HostCallbacks {
  pmx_coeff_fct *coeffs = nullptr;
  pmx_bound_cond_fct *cond = nullptr;
  pmx_bound_value_fct *value = nullptr;
};
// boundary component of a unit-cube node as data/Wuerfel.PRM numbers the
// planes: 0 z = 0, 1 x = 0, 2 y = 0, 3 z = 1, 4 x = 1, 5 y = 1 (the first
// one that holds the node)
static int cube_component(double x, double y, double z) {
  if (z == 0.0) return 0;
  if (x == 0.0) return 1;
  if (y == 0.0) return 2;
  if (z == 1.0) return 3;
  if (x == 1.0) return 4;
  return 5;
}
// The quadrature-point records of every cell of a level from the caller's
// CoeffFct3D, called once per cell with the cell's points as
// TDiscreteForm3D::GetLocalForms does (DiscreteForm3D.C:212-301); coeffs[q]
// holds 20 doubles (eps, b1, b2, b3, c, f at 0..5), param[q] 20 zeros.
static void host_qp_records(const Level &lv, const double *prm, int form, bool linear, pmx_coeff_fct *cf,
                            std::vector<double> &out) {
  std::vector<double> qref, tab;
  int nq;
  make_tables(lv.L.order, linear, qref, tab, nq);
  const int n = lv.L.n;
  const double h = 1.0 / n;
  const int np = nq <= 8 ? 2 : (nq <= 27 ? 3 : 4);  // points of the 1D rule
  out.assign((size_t)n * n * n * nq * 8, 0.0);
  std::vector<double> X(nq), Yc(nq), Z(nq), cbuf((size_t)nq * 20), pbuf((size_t)nq * 20, 0.0);
  std::vector<double *> cptr(nq), pptr(nq);
  for (int q = 0; q < nq; q++) {
    cptr[q] = &cbuf[(size_t)q * 20];
    pptr[q] = &pbuf[(size_t)q * 20];
  }
  for (int cz = 0; cz < n; cz++)
    for (int cy = 0; cy < n; cy++)
      for (int cx = 0; cx < n; cx++) {
        for (int q = 0; q < nq; q++) {
          X[q] = (cx * h + 0.5 * h) + 0.5 * h * qref[4 * q];
          Yc[q] = (cy * h + 0.5 * h) + 0.5 * h * qref[4 * q + 1];
          Z[q] = (cz * h + 0.5 * h) + 0.5 * h * qref[4 * q + 2];
        }
        std::fill(cbuf.begin(), cbuf.end(), 0.0);
        cf(nq, X.data(), Yc.data(), Z.data(), pptr.data(), cptr.data());
        const int64_t cell = cx + (int64_t)n * (cy + (int64_t)n * cz);
        // the reference's first point of the cell, (a, a, a) in its own frame:
        // 1D point 0 (a) along an axis where the frame points forward, np - 1
        // (-a) where it points back (CELL_MEASURE 4)
        const int code = lv.Nb.corner[cell];
        const int q0 = ((code & 1) ? 0 : np - 1) + np * (((code & 2) ? 0 : np - 1) + np * ((code & 4) ? 0 : np - 1));
        QpCoeffs k0;
        {
          const double *c = cptr[q0];
          k0.eps = c[0]; k0.b0 = c[1]; k0.b1 = c[2]; k0.b2 = c[3]; k0.c = c[4]; k0.f = c[5];
        }
        for (int q = 0; q < nq; q++) {
          const double *c = cptr[q];
          QpCoeffs k;
          k.eps = c[0]; k.b0 = c[1]; k.b1 = c[2]; k.b2 = c[3]; k.c = c[4]; k.f = c[5];
          qp_record(prm, form, h, qref[4 * q + 3], k, &out[((size_t)cell * nq + q) * 8], &k0);
        }
      }
}
// the device copy of a level's per-cell frame codes when the assembly needs
// them (CELL_MEASURE 4: the cell's first quadrature point), else nullptr
static const uint8_t *cell_corners(System &Y, Level &lv) {
  if (Y.cell_measure != 4) return nullptr;
  if (!lv.corner.p) lv.corner.upload(lv.Nb.corner.data(), lv.Nb.corner.size(), Y.ctx->st);
  return lv.corner.p;
}
// TSystemCD3D::Assemble of every level with the device form of prm and the
// built-in coefficient functor of prm, or the caller's callbacks (cb)
static void system_assemble(System &Y, const double *prm, const HostCallbacks *cb) {
  hipStream_t st = Y.ctx->st;
  const int form = (int)prm[PMX_PRM_FORM];
  if (form < PMX_FORM_GALERKIN || form > PMX_FORM_RHS) throw std::invalid_argument("unknown form");
  const bool linear = prm[PMX_PRM_LINEAR] != 0;
  Y.prm.upload(prm, PMX_PRM_LEN, st);
  Y.example = (int)prm[PMX_PRM_EXAMPLE];
  Y.cell_measure = (int)prm[PMX_PRM_CELL_MEASURE];
  ensure_tables(Y, linear);
  const int nq = Y.nq;
  const bool rhs_only = form == PMX_FORM_RHS || form == PMX_FORM_RHS_SUPG;
  if (!rhs_only) {
    Y.form = form;
    Y.two_matrices = form == PMX_FORM_TCD_SUPG_AK;
  }
  std::vector<double> rec, bval;
  DevBuf<double> dval;
  for (auto &lp : Y.lev) {
    Level &lv = *lp;
    if (cb) {
      // the records of the level's quadrature points from the caller's
      // coefficients (the buffer is sized for the finest level)
      host_qp_records(lv, prm, form, linear, cb->coeffs, rec);
      check(hipMemcpyAsync(Y.qpbuf.p, rec.data(), sizeof(double) * rec.size(), hipMemcpyHostToDevice, st),
            "upload qp records");
    }
    if (rhs_only) {
      // Assemble3D with no matrices: only the right-hand side
      assemble(lv.D, Y.prm.p, Y.example, form, Y.qref.p, nq, Y.tab.p, Y.qpbuf.p, nullptr, nullptr, nullptr,
               lv.rhs.p, st, cb != nullptr, Y.qpbuf.p + Y.qp_tab_off, cell_corners(Y, lv));
    } else {
      if (Y.two_matrices) lv.val2.alloc_zeroed(lv.L.nvals, st);
      assemble(lv.D, Y.prm.p, Y.example, form, Y.qref.p, nq, Y.tab.p, Y.qpbuf.p, cell_buffer(Y), lv.v64.val.p,
               Y.two_matrices ? lv.val2.p : nullptr, lv.rhs.p, st, cb != nullptr, Y.qpbuf.p + Y.qp_tab_off,
               cell_corners(Y, lv));
      lv.have_val32 = false;
      lv.have_val16 = false;
      lv.have_dinv = false;
    }
    if (cb) {
      // Dirichlet rows: the caller's BoundValue at each boundary node (the
      // nodal functionals of AssembleMat3D.C:1031-1096); BoundCondition
      // must report DIRICHLET there, the only condition the cube's
      // structure holds
      const int M = lv.L.M;
      const double hn = 1.0 / (M - 1);
      bval.resize(lv.L.boundary.size());
      for (size_t b = 0; b < lv.L.boundary.size(); b++) {
        const int p = lv.L.boundary[b];
        const double x = (p % M) * hn, y = ((p / M) % M) * hn, z = (p / (M * M)) * hn;
        const int comp = cube_component(x, y, z);
        if (cb->cond) {
          int c = -1;
          cb->cond(comp, x, y, z, &c);
          if (c != 0) throw std::invalid_argument("BoundCondition: only DIRICHLET (0) boundaries are supported");
        }
        double v = 0.0;
        if (cb->value) cb->value(comp, x, y, z, &v);
        bval[b] = v;
      }
      dval.upload(bval.data(), bval.size(), st);
      scatter(lv.rhs.p, lv.boundary.p, dval.p, (int)bval.size(), st);
      check(hipStreamSynchronize(st), "boundary values");
    }
  }
  check(hipStreamSynchronize(st), "assemble");
  if (!rhs_only) {
    coarse_inverse(Y, *Y.lev.front());
    Y.invalidate_graphs();
    Y.assembled = true;
  }
}
}  // namespace pmx

int pmx_system_assemble(pmx_system *sys, const double *prm) {
  PMX_TRY({
    if (!sys || !prm) throw std::invalid_argument("NULL argument");
    system_assemble(sys->s, prm, nullptr);
  })
}
int pmx_system_assemble_cb(pmx_system *sys, const double *prm, pmx_coeff_fct *coeffs, pmx_bound_cond_fct *cond,
                           pmx_bound_value_fct *value) {
  PMX_TRY({
    if (!sys || !prm || !coeffs) throw std::invalid_argument("NULL argument (prm and coeffs are required)");
    HostCallbacks cb;
    cb.coeffs = coeffs;
    cb.cond = cond;
    cb.value = value;
    system_assemble(sys->s, prm, &cb);
  })
}
// ---- time stepping: TSystemTCD3D (src/System/SystemTCD3D.C:208-428) driven as
// 3DPrograms/TCD3D_ParMooN.C:565-663.  The mass matrix is kept apart, so the
// system matrix M + tau theta1 A is rebuilt each step and no RestoreMassMat
// pass is needed.
// mass: the mass form into lv.mass; otherwise the stiffness form into
// lv.stiff (and K of the SUPG form into lv.val2)
static void assemble_into(System &Y, const double *prm, bool mass) {
  hipStream_t st = Y.ctx->st;
  Y.prm.upload(prm, PMX_PRM_LEN, st);
  Y.example = (int)prm[PMX_PRM_EXAMPLE];
  Y.cell_measure = (int)prm[PMX_PRM_CELL_MEASURE];
  Y.form = (int)prm[PMX_PRM_FORM];
  Y.two_matrices = !mass && Y.form == PMX_FORM_TCD_SUPG_AK;
  ensure_tables(Y, prm[PMX_PRM_LINEAR] != 0);
  const int nq = Y.nq;
  for (auto &lp : Y.lev) {
    Level &lv = *lp;
    if (mass) lv.mass.alloc_zeroed(lv.L.nvals, st);
    else lv.stiff.alloc_zeroed(lv.L.nvals, st);
    if (Y.two_matrices) lv.val2.alloc_zeroed(lv.L.nvals, st);
    assemble(lv.D, Y.prm.p, Y.example, Y.form, Y.qref.p, nq, Y.tab.p, Y.qpbuf.p, cell_buffer(Y),
             mass ? lv.mass.p : lv.stiff.p, Y.two_matrices ? lv.val2.p : nullptr, lv.rhs.p, st, false,
             Y.qpbuf.p + Y.qp_tab_off, cell_corners(Y, lv));
  }
}

int pmx_tcd_init(pmx_system *sys, const double *prm, const double *sol0) {
  PMX_TRY({
    System &Y = sys->s;
    if (!prm) throw std::invalid_argument("prm is NULL");
    hipStream_t st = Y.ctx->st;
    std::vector<double> p(prm, prm + PMX_PRM_LEN);
    // the SUPG time discretisation starts from MatrixMRhsAssemble_SUPG (its
    // rhs carries the streamline term), the others from MatrixMRhsAssemble
    const int f = (int)prm[PMX_PRM_FORM];
    p[PMX_PRM_FORM] = (f == PMX_FORM_TCD_SUPG_AK || f == PMX_FORM_MASS_SUPG) ? PMX_FORM_MASS_SUPG : PMX_FORM_MASS;
    assemble_into(Y, p.data(), true);  // AssembleMRhs (SystemTCD3D.C:208-227)
    Level &lv = *Y.lev.back();
    const int64_t N = lv.L.N;
    for (auto *b : {&Y.oldrhs, &Y.oldsol, &Y.tmp1, &Y.tmp2, &Y.tmp3}) b->alloc(N);
    check(hipMemcpyAsync(Y.oldrhs.p, lv.rhs.p, sizeof(double) * N, hipMemcpyDeviceToDevice, st), "copy");
    if (sol0) to_lattice<double>(Y, lv, sol0, Y.sol.p);
    else interpolate(lv.D, Y.prm.p, Y.sol.p, st);
    boundary<double>(lv.D, 2, Y.sol.p, nullptr, lv.rhs.p, st);  // Dirichlet values of the rhs
    check(hipStreamSynchronize(st), "tcd init");
    Y.tcd_ready = true;
    Y.tcd_phase = 0;
  })
}
// ---- the phases of a time step (TSystemTCD3D, SystemTCD3D.C:229-330, in
// the order of the time loop of 3DPrograms/TCD3D_ParMooN.C:565-640)
namespace pmx {
// AssembleARhs at prm[PMX_PRM_TIME]: oldsol = sol (the loop's memcpy), the
// stiffness form (and K of the SUPG form) and the rhs on every level, the
// Dirichlet values of sol from the rhs
static void tcd_arhs(System &Y, const double *prm, const double *td) {
  if (!Y.tcd_ready) throw std::logic_error("pmx_tcd_init first");
  if (Y.tcd_phase == 2) throw std::logic_error("the system matrix has to be restored before AssembleARhs");
  const int form = (int)prm[PMX_PRM_FORM];
  if (form != PMX_FORM_TCD_A && form != PMX_FORM_GALERKIN && form != PMX_FORM_SDFEM &&
      form != PMX_FORM_TCD_SUPG_AK)
    throw std::invalid_argument("time stepping takes a stiffness form (TCD_A, GALERKIN, SDFEM, TCD_SUPG_AK)");
  hipStream_t st = Y.ctx->st;
  for (int k = 0; k < 5; k++) Y.td[k] = td[k];
  Level &lv = *Y.lev.back();
  const int64_t N = lv.L.N;
  check(hipMemcpyAsync(Y.oldsol.p, Y.sol.p, sizeof(double) * N, hipMemcpyDeviceToDevice, st), "copy");
  // the SUPG form reads the step length and theta1 (TimeDB's
  // CURRENTTIMESTEPLENGTH and THETA1, TCD3D.C:145-147) from td
  std::vector<double> pa(prm, prm + PMX_PRM_LEN);
  pa[PMX_PRM_TAU] = td[0];
  pa[PMX_PRM_THETA1] = td[1];
  assemble_into(Y, pa.data(), false);
  boundary<double>(lv.D, 2, Y.sol.p, nullptr, lv.rhs.p, st);  // sol Dirichlet = rhs
  Y.tcd_phase = 1;
}
// AssembleSystMat: B = tau th3 oldrhs + tau th4 rhs + (M - tau th2 A) oldsol
// on active rows, rhs on Dirichlet rows, system matrix M + tau th1 A on every
// level; the SUPG form adds its time-derivative part K to M in both, as the
// 2D system does (SystemTCD2D.C:294-297).  Then oldrhs = rhs (the loop's
// memcpy after AssembleSystMat) and the coarse inverse of the new operator.
static void tcd_systmat(System &Y) {
  if (Y.tcd_phase == 2) throw std::logic_error("the system is assembled: RestoreMassMat first");
  if (Y.tcd_phase != 1) throw std::logic_error("AssembleARhs first");
  hipStream_t st = Y.ctx->st;
  const double tau = Y.td[0], th1 = Y.td[1], th2 = Y.td[2], th3 = Y.td[3], th4 = Y.td[4];
  Level &lv = *Y.lev.back();
  const int64_t N = lv.L.N;
  const bool supg = Y.two_matrices;
  css_apply<double>(lv.D, 0, lv.mass.p, Y.oldsol.p, nullptr, Y.tmp1.p, 0.0, -1, st);
  css_apply<double>(lv.D, 0, lv.stiff.p, Y.oldsol.p, nullptr, Y.tmp2.p, 0.0, -1, st);
  if (supg) css_apply<double>(lv.D, 0, lv.val2.p, Y.oldsol.p, nullptr, Y.tmp3.p, 0.0, -1, st);
  tcd_rhs(lv.D, Y.oldrhs.p, lv.rhs.p, Y.tmp1.p, Y.tmp2.p, supg ? Y.tmp3.p : nullptr, tau * th3, tau * th4,
          tau * th2, Y.b.p, st);
  for (auto &lp : Y.lev) {
    Level &l = *lp;
    // the FP32 copy (once the FP32 forms have allocated it) in the same pass
    system_matrix(l.v64.val.p, l.v32.val.p, l.mass.p, supg ? l.val2.p : nullptr, l.stiff.p, tau * th1, l.L.nvals,
                  st);
    l.have_val32 = l.v32.val.p != nullptr;
    l.have_val16 = false;
    l.have_dinv = false;
  }
  check(hipMemcpyAsync(Y.oldrhs.p, lv.rhs.p, sizeof(double) * N, hipMemcpyDeviceToDevice, st), "copy");
  coarse_inverse(Y, *Y.lev.front());
  Y.assembled = true;
  Y.tcd_phase = 2;
}
}  // namespace pmx

int pmx_tcd_assemble_arhs(pmx_system *sys, const double *prm, const double *td) {
  PMX_TRY({
    if (!sys || !prm || !td) throw std::invalid_argument("NULL argument");
    tcd_arhs(sys->s, prm, td);
    check(hipStreamSynchronize(sys->s.ctx->st), "AssembleARhs");
  })
}
int pmx_tcd_assemble_systmat(pmx_system *sys) {
  PMX_TRY({
    if (!sys) throw std::invalid_argument("NULL argument");
    tcd_systmat(sys->s);
    check(hipStreamSynchronize(sys->s.ctx->st), "AssembleSystMat");
  })
}
int pmx_tcd_solve(pmx_system *sys, const double *opt, double *info) {
  PMX_TRY({
    if (!sys || !opt) throw std::invalid_argument("NULL argument");
    System &Y = sys->s;
    if (Y.tcd_phase != 2) throw std::logic_error("AssembleSystMat first");
    std::memcpy(Y.opt, opt, sizeof(Y.opt));
    const int rc = solve_device(Y, info);
    if (rc) { g_err = rc == PMX_ERR_NOCONV ? "iteration limit reached" : "diverged"; return rc; }
  })
}
int pmx_tcd_restore_massmat(pmx_system *sys) {
  PMX_TRY({
    if (!sys) throw std::invalid_argument("NULL argument");
    System &Y = sys->s;
    // M, K and A are kept apart, so nothing is subtracted; the system matrix
    // is no longer assembled (SystemTCD3D.C:312-330)
    if (Y.tcd_phase != 2) throw std::logic_error("the system is not assembled: nothing to restore");
    Y.tcd_phase = 0;
  })
}
int pmx_tcd_step(pmx_system *sys, const double *prm, const double *td, const double *opt, double *info) {
  PMX_TRY({
    System &Y = sys->s;
    if (!prm || !td || !opt) throw std::invalid_argument("NULL argument");
    if (Y.tcd_phase == 2) Y.tcd_phase = 0;  // a step after a bare solve restores first
    tcd_arhs(Y, prm, td);
    tcd_systmat(Y);
    std::memcpy(Y.opt, opt, sizeof(Y.opt));
    const int rc = solve_device(Y, info);
    Y.tcd_phase = 0;
    if (rc) { g_err = rc == PMX_ERR_NOCONV ? "iteration limit reached" : "diverged"; return rc; }
  })
}

int pmx_matrix_entries(pmx_system *sys, int level, int which, double *entries) {
  PMX_TRY({
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first");
    Level &lv = level_of(Y, level);
    if (which < 0 || which > 5)
      throw std::invalid_argument("which: 0 operator, 1 K, 2 mass, 3 stiffness, 4 FP32 operator, 5 FP16 smoother");
    ensure_csr_dev(Y, lv);
    DevBuf<double> out;
    out.alloc(csr_nnz(lv.L));
    if (which == 4) {
      ensure_val32(Y);
      css_to_csr<float>(lv.D, lv.rowptr.p, lv.kcol.p, lv.dof2lat.p, lv.L.nactive, lv.v32.val.p, out.p, Y.ctx->st);
    } else if (which == 5) {
      ensure_val16(Y);
      css_to_csr<_Float16>(lv.D, lv.rowptr.p, lv.kcol.p, lv.dof2lat.p, lv.L.nactive, lv.val16.p, out.p,
                           Y.ctx->st);
    } else {
      const double *src = which == 0 ? lv.v64.val.p : which == 1 ? lv.val2.p : which == 2 ? lv.mass.p : lv.stiff.p;
      if (!src) throw std::logic_error("that matrix has not been assembled");
      css_to_csr<double>(lv.D, lv.rowptr.p, lv.kcol.p, lv.dof2lat.p, lv.L.nactive, src, out.p, Y.ctx->st);
    }
    check(hipMemcpyAsync(entries, out.p, sizeof(double) * out.n, hipMemcpyDeviceToHost, Y.ctx->st), "d2h");
    check(hipStreamSynchronize(Y.ctx->st), "sync");
  })
}
int pmx_level_set_entries(pmx_system *sys, int level, int n, const int *rowptr, const int *kcol,
                          const double *entries) {
  PMX_TRY({
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first");
    if (!rowptr || !kcol || !entries) throw std::invalid_argument("NULL argument");
    Level &lv = level_of(Y, level);
    if (n != (int)lv.L.N) throw std::invalid_argument("n differs from the level's DOFs");
    ensure_csr_host(Y, lv);
    // the pattern must be the level's own (TSquareStructure3D of its space)
    if (std::memcmp(rowptr, lv.Nb.rowptr.data(), sizeof(int) * (n + 1)) != 0 ||
        std::memcmp(kcol, lv.Nb.kcol.data(), sizeof(int) * lv.Nb.kcol.size()) != 0)
      throw std::invalid_argument("the matrix pattern differs from the level's structure");
    // Dirichlet rows (the last N - active) are identity rows
    for (int64_t r = lv.L.nactive; r < n; r++)
      if (entries[rowptr[r]] != 1.0) throw std::invalid_argument("a Dirichlet row is not the identity row");
    ensure_csr_dev(Y, lv);
    hipStream_t st = Y.ctx->st;
    DevBuf<double> in;
    in.upload(entries, lv.Nb.kcol.size(), st);
    csr_to_css(lv.D, lv.rowptr.p, lv.kcol.p, lv.dof2lat.p, lv.L.nactive, in.p, lv.v64.val.p, st);
    lv.have_val32 = false;
    lv.have_val16 = false;
    lv.have_dinv = false;
    if (&lv == Y.lev.front().get()) coarse_inverse(Y, lv);
    check(hipStreamSynchronize(st), "set entries");
  })
}
int pmx_system_rhs(pmx_system *sys, int level, double *rhs) {
  PMX_TRY({
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first");
    Level &lv = level_of(Y, level);
    from_lattice<double>(Y, lv, lv.rhs.p, rhs);
  })
}
int pmx_matvec(pmx_system *sys, int level, int precision, const double *x, double *y) {
  PMX_TRY({
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first");
    Level &lv = level_of(Y, level);
    hipStream_t st = Y.ctx->st;
    if (precision == 0) {
      to_lattice<double>(Y, lv, x, lv.v64.aux.p);
      css_apply<double>(lv.D, 0, lv.v64.val.p, lv.v64.aux.p, nullptr, lv.v64.d.p, 0.0, -1, st);
      boundary<double>(lv.D, 0, lv.v64.d.p, lv.v64.aux.p, nullptr, st);
      from_lattice<double>(Y, lv, lv.v64.d.p, y);
    } else {
      ensure_val32(Y);
      to_lattice<float>(Y, lv, x, lv.v32.aux.p);
      css_apply<float>(lv.D, 0, lv.v32.val.p, lv.v32.aux.p, nullptr, lv.v32.d.p, 0.0, -1, st);
      boundary<float>(lv.D, 0, lv.v32.d.p, lv.v32.aux.p, nullptr, st);
      from_lattice<float>(Y, lv, lv.v32.d.p, y);
    }
  })
}
int pmx_defect(pmx_system *sys, int level, const double *x, const double *b, double *r, double *res) {
  PMX_TRY({
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first");
    Level &lv = level_of(Y, level);
    hipStream_t st = Y.ctx->st;
    to_lattice<double>(Y, lv, x, lv.v64.aux.p);
    to_lattice<double>(Y, lv, b, lv.v64.f.p);
    css_apply<double>(lv.D, 1, lv.v64.val.p, lv.v64.aux.p, lv.v64.f.p, lv.v64.d.p, 0.0, -1, st);
    boundary<double>(lv.D, 1, lv.v64.d.p, lv.v64.aux.p, lv.v64.f.p, st);
    double *nr = Y.ctx->part.p + 3 * kRedBlocks;
    dot<double>(lv.v64.d.p, lv.v64.d.p, lv.L.N, Y.ctx->part.p, nr, 1, st);
    if (res) *res = read_scalar(Y, nr);
    from_lattice<double>(Y, lv, lv.v64.d.p, r);
  })
}
int pmx_defect_noalloc(pmx_system *sys, int level, const double *x, const double *b, double *r, double *res) {
  PMX_TRY({
    if (!sys || !x || !b || !r || !res) throw std::invalid_argument("NULL argument");
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first");
    Level &lv = level_of(Y, level);
    if (lv.L.N > kNoAllocMaxRows)
      throw std::invalid_argument("pmx_defect_noalloc: the level has more than 8,388,608 rows (the scratch limit)");
    double *part = noalloc_scratch(*Y.ctx);
    hipStream_t st = Y.ctx->st;
    css_apply<double>(lv.D, 1, lv.v64.val.p, x, b, r, 0.0, -1, st);
    boundary<double>(lv.D, 1, r, x, b, st);
    dot_chunked(r, r, lv.L.N, part, res, 1, st);
    check(hipGetLastError(), "pmx_defect_noalloc launch");
  })
}
int pmx_smooth(pmx_system *sys, int level, int kind, double omega, int sweeps, double *sol, const double *f) {
  PMX_TRY({
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first");
    if (kind < 1 || kind > 3) throw std::invalid_argument("kind must be 1, 2 or 3");
    Level &lv = level_of(Y, level);
    to_lattice<double>(Y, lv, sol, lv.v64.x.p);
    to_lattice<double>(Y, lv, f, lv.v64.f.p);
    level_smooth<double>(Y, lv, kind, omega, sweeps);
    from_lattice<double>(Y, lv, lv.v64.x.p, sol);
  })
}
int pmx_prolongate(pmx_system *sys, int fine_level, const double *xc, double *xf) {
  PMX_TRY({
    System &Y = sys->s;
    if (fine_level < 0) fine_level += (int)Y.lev.size();
    if (fine_level < 1) throw std::invalid_argument("fine_level must be >= 1");
    Level &f = level_of(Y, fine_level), &c = level_of(Y, fine_level - 1);
    to_lattice<double>(Y, c, xc, c.v64.aux.p);
    prolongate<double>(c.D, f.D, c.v64.aux.p, f.v64.aux.p, Y.ctx->st);
    from_lattice<double>(Y, f, f.v64.aux.p, xf);
  })
}
int pmx_restrict(pmx_system *sys, int fine_level, double *rc, const double *rf) {
  PMX_TRY({
    System &Y = sys->s;
    if (fine_level < 0) fine_level += (int)Y.lev.size();
    if (fine_level < 1) throw std::invalid_argument("fine_level must be >= 1");
    Level &f = level_of(Y, fine_level), &c = level_of(Y, fine_level - 1);
    to_lattice<double>(Y, f, rf, f.v64.aux.p);
    restrict_<double>(c.D, f.D, c.v64.aux.p, f.v64.aux.p, Y.ctx->st);
    from_lattice<double>(Y, c, c.v64.aux.p, rc);
  })
}
int pmx_csr_matvec(pmx_ctx *ctx, int n, const int *rowptr, const int *kcol, const double *entries, const double *x,
                   double *y) {
  PMX_TRY({
    if (!ctx || n < 0) throw std::invalid_argument("bad argument");
    hipStream_t st = ctx->c.st;
    const int nnz = rowptr[n];
    DevBuf<int> rp, kc;
    DevBuf<double> v, dx, dy;
    rp.upload(rowptr, n + 1, st);
    kc.upload(kcol, nnz, st);
    v.upload(entries, nnz, st);
    dx.upload(x, n, st);
    dy.alloc(n);
    csr_spmv<double>(n, rp.p, kc.p, v.p, dx.p, dy.p, st);
    check(hipMemcpyAsync(y, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost, st), "d2h");
    check(hipStreamSynchronize(st), "sync");
  })
}

int pmx_system_solve(pmx_system *sys, const double *opt, const double *rhs, double *sol, double *info) {
  PMX_TRY({
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first");
    std::memcpy(Y.opt, opt, sizeof(Y.opt));
    Level &lv = *Y.lev.back();
    to_lattice<double>(Y, lv, rhs, Y.b.p);
    to_lattice<double>(Y, lv, sol, Y.sol.p);
    const int rc = solve_device(Y, info);
    from_lattice<double>(Y, lv, Y.sol.p, sol);
    if (rc) { g_err = rc == PMX_ERR_NOCONV ? "iteration limit reached" : "diverged"; return rc; }
  })
}
int pmx_iterate(pmx_system *sys, int n, const int *rowptr, const int *kcol, const double *entries,
                const double *opt, const double *rhs, double *sol, double *info) {
  PMX_TRY({
    if (!sys || !rowptr || !kcol || !entries || !opt || !rhs || !sol) throw std::invalid_argument("NULL argument");
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first: the MG levels precondition the caller's matrix");
    Level &lv = *Y.lev.back();
    const int64_t N = lv.L.N;
    if (n != N) throw std::invalid_argument("the matrix does not have the finest level's size");
    // the caller's rows and columns in lattice order
    const int64_t nnz = rowptr[n];
    if (rowptr[0] != 0 || nnz < n) throw std::invalid_argument("malformed RowPtr");
    std::vector<int> rp(N + 1), kc(nnz);
    std::vector<double> v(nnz);
    rp[0] = 0;
    for (int64_t p = 0; p < N; p++) {
      const int r = lv.Nb.lat2dof[p];
      int64_t o = rp[p];
      for (int k = rowptr[r]; k < rowptr[r + 1]; k++, o++) {
        if (kcol[k] < 0 || kcol[k] >= n) throw std::invalid_argument("column index out of range");
        kc[o] = lv.Nb.dof2lat[kcol[k]];
        v[o] = entries[k];
      }
      rp[p + 1] = (int)o;
    }
    hipStream_t st = Y.ctx->st;
    Y.ucsr_rp.upload(rp.data(), rp.size(), st);
    Y.ucsr_kc.upload(kc.data(), kc.size(), st);
    Y.ucsr_val.upload(v.data(), v.size(), st);
    std::memcpy(Y.opt, opt, sizeof(Y.opt));
    to_lattice<double>(Y, lv, rhs, Y.b.p);
    to_lattice<double>(Y, lv, sol, Y.sol.p);
    Y.use_ucsr = true;
    int rc;
    try {
      rc = solve_device(Y, info);
    } catch (...) {
      Y.use_ucsr = false;
      throw;
    }
    Y.use_ucsr = false;
    from_lattice<double>(Y, lv, Y.sol.p, sol);
    if (rc) { g_err = rc == PMX_ERR_NOCONV ? "iteration limit reached" : "diverged"; return rc; }
  })
}
int pmx_system_solve_resident(pmx_system *sys, const double *opt, double *info) {
  PMX_TRY({
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first");
    std::memcpy(Y.opt, opt, sizeof(Y.opt));
    Level &lv = *Y.lev.back();
    hipStream_t st = Y.ctx->st;
    check(hipMemcpyAsync(Y.b.p, lv.rhs.p, sizeof(double) * lv.L.N, hipMemcpyDeviceToDevice, st), "copy");
    fill<double>(Y.sol.p, 0.0, lv.L.N, st);
    boundary<double>(lv.D, 2, Y.sol.p, nullptr, Y.b.p, st);
    const int rc = solve_device(Y, info);
    if (rc) { g_err = rc == PMX_ERR_NOCONV ? "iteration limit reached" : "diverged"; return rc; }
  })
}
int pmx_system_set_step_length_control(pmx_system *sys, int mode) {
  PMX_TRY({
    if (mode < 0 || mode > 2) throw std::invalid_argument("step length control: mode must be 0, 1 or 2");
    System &Y = sys->s;
    // the kept iterate and defect of every level, both precisions (allocated
    // here: the cycle may be captured into a graph, where nothing allocates)
    if (mode > 0)
      for (auto &l : Y.lev)
        for (DevBuf<double> *b : {&l->v64.sold, &l->v64.sdef}) b->alloc(l->L.N);
    if (mode > 0)
      for (auto &l : Y.lev)
        for (DevBuf<float> *b : {&l->v32.sold, &l->v32.sdef}) b->alloc(l->L.N);
    Y.slc = mode;
  })
}
int pmx_system_solution(pmx_system *sys, double *sol) {
  PMX_TRY({
    System &Y = sys->s;
    from_lattice<double>(Y, *Y.lev.back(), Y.sol.p, sol);
  })
}

int pmx_system_errors(pmx_system *sys, const double *prm, double *err) {
  PMX_TRY({
    System &Y = sys->s;
    if (!prm || !err) throw std::invalid_argument("null argument");
    if (!Y.sol.p) throw std::logic_error("no solution on the device: solve first");
    hipStream_t st = Y.ctx->st;
    ensure_tables(Y, prm[PMX_PRM_LINEAR] != 0);
    Y.prm.upload(prm, PMX_PRM_LEN, st);
    double *out = Y.ctx->part.p + 3 * kRedBlocks + 8;
    fe_errors(Y.lev.back()->D, Y.prm.p, Y.qref.p, Y.tab.p, Y.nq, Y.sol.p, Y.ctx->part.p, out, st);
    check(hipMemcpyAsync(err, out, 2 * sizeof(double), hipMemcpyDeviceToHost, st), "errors");
    check(hipStreamSynchronize(st), "errors");
  })
}

int pmx_time_kernel(pmx_system *sys, int kernel, int level, int precision, int reps, double *ms) {
  PMX_TRY({
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first");
    if (reps < 1) throw std::invalid_argument("reps < 1");
    Level &lv = level_of(Y, level);
    const int li = level < 0 ? level + (int)Y.lev.size() : level;
    if (kernel == 7 && li != (int)Y.lev.size() - 1) throw std::invalid_argument("V-cycle timing is on the finest level");
    std::memcpy(Y.opt, PMX_DEFAULT_VCYCLE, sizeof(Y.opt));
    hipStream_t st = Y.ctx->st;
    if (precision < 0 || precision > 3 || precision == 2) throw std::invalid_argument("precision must be 0, 1 or 3");
    // 3: FP32 with the SOR/SSOR colour launches on row-scaled FP16 values
    Y.half_smoother = precision == 3;
    if (precision == 3) precision = 1;
    if (precision) ensure_val32(Y);
    if (Y.half_smoother) ensure_val16(Y);
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    auto one = [&]() {
      if (precision == 0) {
        Vecs<double> &v = lv.v64;
        switch (kernel) {
          case 0: css_apply<double>(lv.D, 0, v.val.p, v.x.p, nullptr, v.d.p, 0.0, -1, st); break;
          case 1: css_apply<double>(lv.D, 1, v.val.p, v.x.p, v.f.p, v.d.p, 0.0, -1, st); break;
          case 2: level_smooth<double>(Y, lv, 1, 0.8, 1); break;
          case 3: level_smooth<double>(Y, lv, 3, 1.0, 1); break;
          case 4: assemble(lv.D, Y.prm.p, Y.example, Y.form, Y.qref.p, Y.nq, Y.tab.p, Y.qpbuf.p, cell_buffer(Y), v.val.p, Y.two_matrices ? lv.val2.p : nullptr, lv.rhs.p, st, false, Y.qpbuf.p + Y.qp_tab_off, cell_corners(Y, lv)); break;
          case 5: if (li > 0) prolongate<double>(Y.lev[li - 1]->D, lv.D, Y.lev[li - 1]->v64.x.p, v.aux.p, st); break;
          case 6: if (li > 0) restrict_<double>(Y.lev[li - 1]->D, lv.D, Y.lev[li - 1]->v64.f.p, v.d.p, st); break;
          case 7: mg_prec<double>(Y, v.aux.p, v.d.p); break;
          default: break;
        }
      } else {
        Vecs<float> &v = lv.v32;
        switch (kernel) {
          case 0: css_apply<float>(lv.D, 0, v.val.p, v.x.p, nullptr, v.d.p, 0.0, -1, st); break;
          case 1: css_apply<float>(lv.D, 1, v.val.p, v.x.p, v.f.p, v.d.p, 0.0, -1, st); break;
          case 2: level_smooth<float>(Y, lv, 1, 0.8, 1); break;
          case 3: level_smooth<float>(Y, lv, 3, 1.0, 1); break;
          case 5: if (li > 0) prolongate<float>(Y.lev[li - 1]->D, lv.D, Y.lev[li - 1]->v32.x.p, v.aux.p, st); break;
          case 6: if (li > 0) restrict_<float>(Y.lev[li - 1]->D, lv.D, Y.lev[li - 1]->v32.f.p, v.d.p, st); break;
          case 7: mg_prec<float>(Y, v.aux.p, v.d.p); break;
          case 8: learned_post(Y, lv, Y.nets[li]); break;
          case 9: net_forward_act(Y.nets[li], lv.L.M, false, st); break;
          default: break;
        }
      }
    };
    if (kernel == 10 || kernel == 11) {
      // durations of the individual dispatches (not the gaps between them):
      // 10 the colour launches of SSOR sweeps (launch per colour forced), 11 SpMV
      Y.opt[21] = -1;
      auto run = [&]() {
        if (kernel == 10) {
          if (precision) level_smooth<float>(Y, lv, 3, 1.0, 1);
          else level_smooth<double>(Y, lv, 3, 1.0, 1);
        } else {
          if (precision) css_apply<float>(lv.D, 0, lv.v32.val.p, lv.v32.x.p, nullptr, lv.v32.d.p, 0.0, -1, st);
          else css_apply<double>(lv.D, 0, lv.v64.val.p, lv.v64.x.p, nullptr, lv.v64.d.p, 0.0, -1, st);
        }
      };
      run();
      DispatchTimer tm;
      set_dispatch_timer(&tm);
      try {
        for (int r = 0; r < reps; r++) run();
      } catch (...) {
        set_dispatch_timer(nullptr);
        throw;
      }
      set_dispatch_timer(nullptr);
      check(hipStreamSynchronize(st), "sync");
      *ms = tm.used ? tm.total_ms() / tm.used : 0.0;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return PMX_OK;
    }
    if (kernel == 12 || kernel == 13) {
      // calibration of kind 10 on the same level: 12 a dispatch that does
      // nothing with the workgroup count of each colour launch; 13 a plain
      // streaming read of each colour's stored values (FP64, FP32 or the FP16
      // smoother values), one dispatch per colour; average per dispatch
      std::vector<int64_t> lo(lv.L.ncolours, -1), hi(lv.L.ncolours, 0);
      for (const Segment &g : lv.L.segs) {
        const int64_t len = (g.count + kSlice - 1) / kSlice * kSlice * g.ns;
        if (lo[g.colour] < 0 || g.voff < lo[g.colour]) lo[g.colour] = g.voff;
        hi[g.colour] = std::max(hi[g.colour], g.voff + len);
      }
      const size_t es = Y.half_smoother ? 2 : (precision ? 4 : 8);
      const char *base = Y.half_smoother ? (const char *)lv.val16.p
                                         : (precision ? (const char *)lv.v32.val.p : (const char *)lv.v64.val.p);
      float *sink = lv.v32.aux.p ? lv.v32.aux.p : (float *)lv.v64.aux.p;
      auto run = [&]() {
        for (int c = 0; c < lv.L.ncolours; c++) {
          if (lo[c] < 0) continue;
          if (kernel == 12) empty_dispatch(lv.D.colour_off[c + 1] - lv.D.colour_off[c], st);
          else stream_read(base + lo[c] * es, (hi[c] - lo[c]) * (int64_t)es, sink, st);
        }
      };
      run();
      DispatchTimer tm;
      set_dispatch_timer(&tm);
      try {
        for (int r = 0; r < reps; r++) run();
      } catch (...) {
        set_dispatch_timer(nullptr);
        throw;
      }
      set_dispatch_timer(nullptr);
      check(hipStreamSynchronize(st), "sync");
      *ms = tm.used ? tm.total_ms() / tm.used : 0.0;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return PMX_OK;
    }
    if (kernel >= 8) {
      if (!precision) throw std::invalid_argument("the learned smoother runs in FP32");
      if (!Y.nets[li].ready) throw std::logic_error("learned smoother: train or load parameters first");
      const int keep = Y.learned_levels;
      Y.learned_levels = (int)Y.lev.size() - li;
      ensure_learned(Y);
      Y.learned_levels = keep;
    }
    one();
    check(hipEventRecord(e0, st), "event");
    for (int r = 0; r < reps; r++) one();
    check(hipEventRecord(e1, st), "event");
    check(hipEventSynchronize(e1), "event");
    float t = 0;
    check(hipEventElapsedTime(&t, e0, e1), "event");
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  })
}

int pmx_learned_nparams(int *n) {
  PMX_TRY({
    if (!n) throw std::invalid_argument("NULL argument");
    *n = net_total_params();
  })
}
int pmx_learned_init(pmx_system *sys, unsigned seed) {
  PMX_TRY({
    System &Y = sys->s;
    for (int i = 1; i < (int)Y.lev.size(); i++) net_init(Y.nets[i], seed + 31u * (uint32_t)i, Y.ctx->st);
    check(hipStreamSynchronize(Y.ctx->st), "sync");
    Y.invalidate_graphs();
  })
}
int pmx_learned_params_level(pmx_system *sys, int level, int set, float *params) {
  PMX_TRY({
    System &Y = sys->s;
    if (!params) throw std::invalid_argument("NULL argument");
    (void)level_of(Y, level);
    const int li = level < 0 ? level + (int)Y.lev.size() : level;
    LearnedNet &net = Y.nets[li];
    hipStream_t st = Y.ctx->st;
    const size_t bytes = sizeof(float) * net_total_params();
    if (set) {
      if (!net.ready) net_init(net, 0, st);
      check(hipMemcpyAsync(net.params, params, bytes, hipMemcpyHostToDevice, st), "h2d params");
      net_sync_bf16(net, st);
      Y.invalidate_graphs();
    } else {
      if (!net.ready) throw std::logic_error("learned smoother: no parameters");
      check(hipMemcpyAsync(params, net.params, bytes, hipMemcpyDeviceToHost, st), "d2h params");
    }
    check(hipStreamSynchronize(st), "sync");
  })
}
int pmx_learned_params(pmx_system *sys, int set, float *params) {
  return pmx_learned_params_level(sys, -1, set, params);
}
int pmx_learned_forward(pmx_system *sys, int M, const float *in, float *out) {
  PMX_TRY({
    System &Y = sys->s;
    LearnedNet &net = Y.nets[Y.lev.size() - 1];
    if (!in || !out || M < 3 || M > 1025) throw std::invalid_argument("bad lattice or NULL argument");
    if (!net.ready) throw std::logic_error("learned smoother: no parameters");
    hipStream_t st = Y.ctx->st;
    const int64_t N = (int64_t)M * M * M;
    if (net_reserve(net, N, false)) Y.invalidate_graphs();
    DevBuf<float> din;
    din.upload(in, (size_t)N * kNetIn, st);
    net_forward(net, M, din.p, false, st);
    check(hipMemcpyAsync(out, net.out, sizeof(float) * N, hipMemcpyDeviceToHost, st), "d2h");
    check(hipStreamSynchronize(st), "sync");
  })
}
int pmx_learned_gradient(pmx_system *sys, int M, const float *in, const float *target, float *grads,
                         double *loss) {
  PMX_TRY({
    System &Y = sys->s;
    LearnedNet &net = Y.nets[Y.lev.size() - 1];
    if (!in || !target || M < 3 || M > 1025) throw std::invalid_argument("bad lattice or NULL argument");
    if (!net.ready) throw std::logic_error("learned smoother: no parameters");
    hipStream_t st = Y.ctx->st;
    const int64_t N = (int64_t)M * M * M;
    if (net_reserve(net, N, true)) Y.invalidate_graphs();
    DevBuf<float> din, dtg, dl;
    din.upload(in, (size_t)N * kNetIn, st);
    dtg.upload(target, N, st);
    dl.alloc(1);
    net_zero_grads(net, st);
    net_forward(net, M, din.p, true, st);
    net_backward(net, M, dtg.p, dl.p, st);
    if (grads)
      check(hipMemcpyAsync(grads, net.grads, sizeof(float) * net_total_params(), hipMemcpyDeviceToHost, st),
            "d2h grads");
    float l = 0;
    check(hipMemcpyAsync(&l, dl.p, sizeof(float), hipMemcpyDeviceToHost, st), "d2h loss");
    check(hipStreamSynchronize(st), "sync");
    if (loss) *loss = l;
  })
}
int pmx_learned_gradient_level(pmx_system *sys, int level, const float *in, float *grads, double *loss) {
  PMX_TRY({
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first");
    if (!in) throw std::invalid_argument("NULL argument");
    Level &lv = level_of(Y, level);
    const int li = level < 0 ? level + (int)Y.lev.size() : level;
    LearnedNet &net = Y.nets[li];
    if (!net.ready) throw std::logic_error("learned smoother: no parameters");
    hipStream_t st = Y.ctx->st;
    const int keep = Y.learned_levels;
    Y.learned_levels = (int)Y.lev.size() - li;
    ensure_learned(Y);
    Y.learned_levels = keep;
    if (net_reserve(net, lv.L.N, true)) Y.invalidate_graphs();
    DevBuf<float> din, dl;
    din.upload(in, lv.L.N, st);
    dl.alloc(1);
    net_zero_grads(net, st);
    learned_transpose(Y, lv);
    learned_res_step(Y, lv, net, din.p, dl.p);
    if (grads)
      check(hipMemcpyAsync(grads, net.grads, sizeof(float) * net_total_params(), hipMemcpyDeviceToHost, st),
            "d2h grads");
    float l = 0;
    check(hipMemcpyAsync(&l, dl.p, sizeof(float), hipMemcpyDeviceToHost, st), "d2h loss");
    check(hipStreamSynchronize(st), "sync");
    if (loss) *loss = l;
  })
}
int pmx_learned_adam(pmx_system *sys, double lr) {
  PMX_TRY({
    System &Y = sys->s;
    LearnedNet &net = Y.nets[Y.lev.size() - 1];
    if (!net.ready) throw std::logic_error("learned smoother: no parameters");
    net_adam(net, (float)lr, Y.ctx->st);
    check(hipStreamSynchronize(Y.ctx->st), "sync");
    Y.invalidate_graphs();
  })
}
int pmx_learned_train(pmx_system *sys, const double *opt, const double *topt, double *tinfo) {
  PMX_TRY({
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first");
    if (!opt || !topt) throw std::invalid_argument("NULL argument");
    if (Y.lev.size() < 2) throw std::invalid_argument("the learned smoother needs at least two levels");
    std::memcpy(Y.opt, opt, sizeof(Y.opt));
    learned_train(Y, topt, tinfo);
  })
}
int pmx_learned_post(pmx_system *sys, int level, double *sol, const double *f) {
  PMX_TRY({
    System &Y = sys->s;
    if (!Y.assembled) throw std::logic_error("assemble first");
    Level &lv = level_of(Y, level);
    const int li = level < 0 ? level + (int)Y.lev.size() : level;
    if (!Y.nets[li].ready) throw std::logic_error("learned smoother: no parameters");
    const int keep = Y.learned_levels;
    Y.learned_levels = (int)Y.lev.size() - li;
    ensure_learned(Y);
    Y.learned_levels = keep;
    to_lattice<float>(Y, lv, sol, lv.v32.x.p);
    to_lattice<float>(Y, lv, f, lv.v32.f.p);
    learned_post(Y, lv, Y.nets[li]);
    from_lattice<float>(Y, lv, lv.v32.x.p, sol);
  })
}

}  // extern "C"
