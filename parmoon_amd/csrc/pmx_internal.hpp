// pmx_internal.hpp -- internal data model of libpmx (not part of the C-ABI).
//
// Layout in HBM.  Every level of the hierarchy is a uniformly refined unit
// cube with n = 2^nref cells per direction.  Its Q1/Q2 nodes form a lattice of
// M = order*n + 1 points per direction, and every device vector is stored in
// lattice order (x fastest), length M^3.  The reference DOF numbering
// (ConstructSpace) is only used at the boundary: `lat2dof`/`dof2lat` permute
// between the two.
//
// Operators are stored "colour-sliced stencil" (CSS): the active (interior)
// lattice nodes are split into residue classes modulo S (S = 2 for Q1, 6 for
// Q2).  A class has one smoother colour (coordinates mod 2 resp. mod 3) and
// one stencil shape (Q2: extent 2 along an even coordinate, 1 along an odd
// one), so a class is a strided sub-lattice whose rows all have the same
// column offsets.  Inside a class the rows are cut into slices of 16 and a
// slice stores its values slot-major:
//   val[seg.voff + (idx/16)*(16*ns) + slot*16 + idx%16]
// so the wavefront that owns 16 rows with 4 lanes per row reads 4
// consecutive slots of 16 rows = 256 contiguous bytes (FP32) per load, and no
// column indices are stored.  Dirichlet rows (lattice boundary) are identity
// and not stored.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pmx {

constexpr int kBlock = 256;       // threads per workgroup of the operator kernels
constexpr int kRowsPerBlock = 64; // rows per entry of the block tables (and per row-kernel workgroup)
// workgroup of the single-workgroup sweep: FP64 rows need twice the
// registers for their loads in flight, so half the lanes
template <typename T> constexpr int sweep_threads() { return sizeof(T) == 8 ? 512 : 1024; }
constexpr int kMaxSegs = 216;        // residue classes of a level: at most 6^3
constexpr int kMaxColours = 27;
constexpr int64_t kSweepLdsBytes = 48 * 1024;  // largest x staged in LDS by the single-workgroup sweep
constexpr int kSweepWgRows = 512;    // default: single-workgroup sweep up to this many rows per colour
#ifndef PMX_SLICE
#define PMX_SLICE 16
#endif
constexpr int kSlice = PMX_SLICE;  // rows per value slice (see Layout)
constexpr int kRedBlocks = 1024;  // partial sums of the deterministic reductions

struct Segment {
  int rx, ry, rz;     // residues mod S
  int x0, y0, z0;     // first interior lattice coordinate of the class
  int cx, cy, cz;     // node counts per direction
  int ex, ey, ez;     // stencil half extents
  int colour;
  int ns;             // slots = (2ex+1)(2ey+1)(2ez+1)
  int dslot;          // slot of offset (0,0,0)
  int64_t count;      // cx*cy*cz
  int64_t voff;       // offset of the class in the value array
};

// device-side copy of a segment (plain struct, kernel argument via array)
struct DevSegment {
  int x0, y0, z0, cx, cy, cz;
  int ex, ey, ez, ns, dslot, stride;
  int64_t count, voff;
};

// one entry per launched block: which segment and the first node index
struct BlockInfo {
  int seg;
  int base;
};

// a block entry of the row kernels with its class record inline, so that a
// workgroup finds its rows with one load instead of two dependent ones
struct RowBlock {
  DevSegment s;
  int base, pad;
};

struct Layout {
  int nref = 0, order = 1, n = 0, M = 0, S = 2;
  int64_t N = 0;            // lattice nodes = DOFs
  int64_t nactive = 0;      // interior nodes
  int ncolours = 8;
  std::vector<Segment> segs;
  int64_t nvals = 0;        // total stored stencil values
  // block tables: all segments (spmv) and per colour (smoothers)
  std::vector<BlockInfo> blocks_all;
  std::vector<std::vector<BlockInfo>> blocks_colour;
  std::vector<int> boundary;  // lattice indices of Dirichlet nodes
};

// host mesh / numbering products (the TFESpace3D + TSquareStructure3D data)
struct Numbering {
  int64_t ncells = 0;
  int nb = 8;
  int64_t N = 0, active = 0;
  std::vector<int> begin_index, global_numbers;  // reference numbering
  std::vector<int> lat2dof, dof2lat;
  std::vector<int> rowptr, kcol;                 // built on demand
  bool have_csr = false;
  // per lattice cell (cx + n (cy + n cz)): the signs of e1 + e2 + e3 of the
  // cell's frame in the refinement tree (bit d set: +1 along axis d), i.e. the
  // direction from the cell centre to its reference point (a, a, a)
  std::vector<uint8_t> corner;
};

void build_layout(Layout &L, int nref, int order, int block);
void build_numbering(Numbering &Nb, const Layout &L);
void build_csr(Numbering &Nb, const Layout &L);

// legacy-VTK file of nvar scalar functions (reference numbering) on the
// level's hexahedra, in the layout of TOutput3D::WriteVtk
void write_vtk(const Layout &L, const Numbering &Nb, int nvar, const char *const *names,
               const double *const *values, const char *path);

// reference-numbering colour of a DOF (the multicolour smoother class)
int node_colour(const Layout &L, int x, int y, int z);

}  // namespace pmx
