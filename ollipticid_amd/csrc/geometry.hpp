// Box hierarchy geometry (host side).
//
// Restates the reference's grid construction so that every box has exactly
// the reference's point count, spacing and origin:
//   Valid_N                 interface.cc:883-949
//   domain Make/Set_N       domain.cc:27-99, 293-445, 451-525
//   Get_i binary search     domain.cc:1493-1675
//   Find_cell_xyz           domain.cc:1682-1706
//   operator+ / operator&&  domain.cc:1811-1871, 1878-1922, 1972-2005
//   Set_Grids               elliptic.cc:300-921
// and builds the per-axis transfer tables the device kernels use for
// restriction, prolongation and ghost-shell interpolation
// (ffunction.cc:2436-2475, interpolation.cc:125-375).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace oll {

size_t valid_n(size_t n);

struct Box {
  int level = 0;
  int order = 2;        // stencil order of the box
  int R = 1;            // ghost width = max(1, order/2)
  size_t n[3] = {0, 0, 0};
  double d[3] = {0, 0, 0};
  double c[3] = {0, 0, 0};    // centre
  double g0[3] = {0, 0, 0};   // extents before the boundary fix-up
  double g1[3] = {0, 0, 0};
  double lo[3] = {0, 0, 0};   // coordinate of index 0
  double hi[3] = {0, 0, 0};   // coordinate of index n-1
  double len[3] = {0, 0, 0};

  size_t npts() const { return n[0] * n[1] * n[2]; }
  double x(int a, long i) const { return lo[a] + double(i) * d[a]; }
  // first index with x_i >= x (within 3 eps), clamped to 1..n-1
  size_t cell(int a, double x) const;
  bool in_nb(int a, double x) const;   // strictly inside the non-ghost part
  bool in_box(int a, double x) const;  // inside, 0.1 dx tolerance
};

// A box built from (points per side, spacing, centre, order).
Box make_box(const size_t N[3], const double d[3], const double c[3], int order);

struct Hierarchy {
  int levels = 0;   // total levels incl. multigrid coarsenings
  int mgrid = 0;    // number of coarsenings; the base box is level mgrid
  std::vector<Box> boxes;          // coarsest level first
  std::vector<int> level_first, level_count;
};

struct PunctureSet {
  std::vector<double> x, y, z;
};

// Builds the hierarchy for N points per side on a cube of side L with
// `levels` refinement levels around the given punctures.  Returns an empty
// string on success, otherwise the error message.
// coarsen = false builds the grids of -m AppPuncture: no multigrid levels
// below the base box (elliptic.cc:510, 678).
std::string build_hierarchy(size_t N, double L, int levels, int order,
                            const PunctureSet &p, Hierarchy &h, bool coarsen = true);

// ---------------------------------------------------------------------
// Per-axis transfer tables.  Every transfer of the reference is a tensor
// product of one-dimensional interpolations whose node choice and weights
// depend only on the target coordinate along each axis, and whose
// "point is inside the source" tests are conjunctions of per-axis tests.
// One table per (target box, source box, axis) therefore describes the
// whole transfer.
// ---------------------------------------------------------------------
enum AxisFlag : uint8_t {
  AX_NB = 1,       // target coordinate inside the source's non-ghost part
  AX_IN = 2,       // inside the source box (0.1 dx tolerance)
  AX_COINC = 4,    // coincides (3 eps) with source node `coinc`
  AX_FW = 8,       // coincides (1e-6 dx) with a source node whose +-1
                   // neighbours are in the source interior
};

struct AxisTable {
  int k = 0;                    // nodes per axis (ord + 1)
  std::vector<int> start;       // first source node per target index
  std::vector<int> coinc;       // coincident source node (AX_COINC)
  std::vector<uint8_t> flags;   // AxisFlag bits
  std::vector<double> w;        // k weights per target index
};

// method: 0 = linear (cell i-1, i), 1 = Lagrange of order `ord`.
AxisTable build_axis_table(const Box &target, const Box &source, int axis,
                           int method, int ord);

// Whether a linear table is a 2:1 nesting along its axis: for an offset m
// and every target index t inside the source (AX_NB or AX_IN), q = t + m
// >= 0, an even q sits on source node q/2 (coincident, or weight 1 on that
// node) and an odd q midway between nodes (q-1)/2 and (q+1)/2 with weights
// 1/2 (weights to 1e-12).  Sets m.
// The prolongation then forms its nodes and weights from q (no tables).
bool nested_axis(const AxisTable &t, int &m);

}  // namespace oll
