// Device-side descriptors shared by the kernels and the host driver.
//
// HBM layout: one slab per hierarchy.  Each field (u, b, c, d, e, rho, r,
// v, w, ww) is one contiguous float64 array holding every box of every
// level back to back; a box is stored in plain lexicographic order
// (i fastest, then j, then k) including its ghost shell of width R.
#pragma once

#include <cstdint>

namespace oll {

enum Field { F_U = 0, F_B, F_C, F_D, F_E, F_RHO, F_R, F_V, F_W, F_WW, NFIELDS };

struct DevBox {
  int n[3];          // points per axis incl. ghosts
  int R;             // ghost width (stencil radius)
  int order;         // 2,4,6,8
  int level;
  long long off;     // offset of point (0,0,0) in every field array
  double lo[3];      // coordinate of index 0
  double d[3];
  double idx[3];     // 1/d
  double idx2[3];    // (1/d)^2
  double dulap;      // diagonal of the Laplacian stencil
};

// One-axis interpolation table (see geometry.hpp AxisTable), flattened into
// a device buffer.
struct DevAxis {
  int n;             // target points along the axis
  int k;             // nodes per target point
  int start_off;     // offsets into the int / byte / double pools
  int coinc_off;
  int flags_off;
  int w_off;
};

// A transfer between one target box and one source box: three axis tables.
struct DevXfer {
  int target, source;    // box ids
  DevAxis ax[3];
};

struct Puncture {
  double m, x, y, z, px, py, pz, sx, sy, sz;
};

constexpr int kMaxBH = 16;

struct PunctureArgs {
  int n;
  int eq;
  Puncture p[kMaxBH];
  // scalar-field constants (elliptic.cc:71-81)
  int ps_case;
  double ps_r0, ps_isigma, ps_phi0, ps_d1, ps_e1, ps_e2, ps_de1;
};

}  // namespace oll
