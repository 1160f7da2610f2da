// Device descriptors and host-side launchers of the CDNA4 kernels.
//
// HBM layout: one slab per hierarchy (request).  Each field (u, b, c, d, e,
// rho, r, v, w, ww) of a box is a contiguous float64 array in plain
// lexicographic order (i fastest) including the ghost shell of width R; a
// DevBox carries the base pointers of its fields, so one launch can work on
// boxes of many hierarchies.  Launches read device-resident item lists.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace oll {

enum Field { F_U = 0, F_B, F_C, F_D, F_E, F_RHO, F_R, F_V, F_W, F_WW, NFIELDS };

struct DevBox {
  int n[3];          // points per axis incl. ghosts
  int R;             // ghost width (stencil radius)
  int order;         // 2,4,6,8
  int level;
  double lo[3];      // coordinate of index 0
  double d[3];
  double idx[3];     // 1/d
  double idx2[3];    // (1/d)^2
  double dulap;      // diagonal of the Laplacian stencil
  double ipn;        // 1 / Robin fall-off power
  double *f[NFIELDS];
};

// one axis of a transfer table (geometry.hpp AxisTable)
struct DevAxis {
  const int *start;
  const int *coinc;
  const unsigned char *flags;
  const double *w;
  int k;
  int n;
};

// a transfer into one target box from one source box
struct DevXfer {
  const DevBox *src;
  DevAxis ax[3];
};

// row-parallel launches (smoother, residual): item owns rows row0..row0+rows
struct RowItem {
  const DevBox *box;
  unsigned long long *norm;
  int row0;
  int rows;
};

// point-parallel launches (transfers): item owns target points p0..p0+np
// (all points of the box, or its ghost-shell points for the ghost fill)
struct PtItem {
  const DevBox *box;
  const DevXfer *src;  // source tables in source order
  int nsrc;
  int mg;              // multigrid level (Robin shell, "all" coverage)
  long long p0;
  long long np;
};

// Robin stages: face columns col0..col0+ncol of a box
struct RobinItem {
  const DevBox *box;
  int col0;
  int ncol;
};

struct CoarseItem {
  const DevBox *box;
  unsigned long long *norm;
  int *iters;
  double omega;
};

struct Puncture {
  double m, x, y, z, px, py, pz, sx, sy, sz;
};

constexpr int kMaxBH = 16;

struct PunctureArgs {
  int n;
  int eq;
  Puncture p[kMaxBH];
  int ps_case;
  double ps_r0, ps_isigma, ps_phi0, ps_d1, ps_e1, ps_e2, ps_de1;
};

// which points a transfer visits (ffunction.cc Set_Iterator)
enum Region { REG_ALL = 0, REG_INSIDE = 1, REG_BOUNDARY = 2 };
enum RobinMode { ROBIN_UPDATE = 0, ROBIN_OPERATOR_R = 1, ROBIN_DEFECT_W = 2 };

constexpr int kNormSubSlots = 8;  // u64 words per logical norm slot
constexpr int kCoarseOwn = 8;     // interior points per thread in the coarse solve

void launch_coefficients_box(hipStream_t s, const DevBox &B, const PunctureArgs &pa,
                             double robin_a);
void launch_smooth(hipStream_t s, const RowItem *items, int n, int total_rows, int max_row_len,
                   int order, int eq, int colour);
void launch_residual(hipStream_t s, const RowItem *items, int n, int total_rows, int max_row_len,
                     int order, int eq, int write_defect);
void launch_robin(hipStream_t s, const RobinItem *items, int n, int total_cols, int mode);
void launch_restrict_u(hipStream_t s, const PtItem *items, int n, long long total);
void launch_restrict_defect(hipStream_t s, const PtItem *items, int n, long long total, int order,
                            int eq);
void launch_prolong(hipStream_t s, const PtItem *items, int n, long long total);
void launch_ghost_fill(hipStream_t s, const PtItem *items, int n, long long total);
// false when the largest box does not fit the LDS-resident kernel
bool launch_coarse_solve(hipStream_t s, const CoarseItem *items, int n, int max_pts, int max_int,
                         int eq, double eps, int max_iter);
void launch_zero_u64(hipStream_t s, unsigned long long *p, long long n);
void launch_zero_lists(hipStream_t s, unsigned long long *const *ptrs, int n, int count);
// one-time kernel attributes (dynamic LDS of the coarse solve); call before
// any graph capture
void kernels_init();
// dynamic LDS the coarse solve can use
constexpr int kCoarseLdsLimit = 160 * 1024;
void launch_adm_sphere(hipStream_t s, const DevBox *boxes, int nbox, const double *centres,
                       double *sph, int nt);
void launch_interp_points(hipStream_t s, const DevBox *boxes, int field, const int *pbox,
                          const double *pxyz, int npts, int ord, double *out);

}  // namespace oll
