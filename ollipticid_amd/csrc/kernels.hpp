// Host-side launchers of the CDNA4 kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace oll {

constexpr int kMaxLevelBoxes = 16;

struct Fields {
  double *f[NFIELDS];
};

// Boxes of one level, with the prefix of interior rows (ny-2R)(nz-2R) and
// of all points, used to map a flat block index to (box, row).
struct LevelView {
  int nbox;
  int box[kMaxLevelBoxes];
  int row0[kMaxLevelBoxes + 1];     // interior rows prefix
  long long pt0[kMaxLevelBoxes + 1];  // all points prefix
  long long shell0[kMaxLevelBoxes + 1];  // ghost-shell points prefix
  int max_row_len;                  // longest interior row (points)
};

// Transfers into the boxes of one target level: for target box t the
// source tables are xfer[x0[t] .. x0[t+1]) in source order.
struct XferView {
  LevelView tgt;
  int x0[kMaxLevelBoxes + 1];
  const DevXfer *xfer;
  const int *ipool;
  const unsigned char *bpool;
  const double *dpool;
};

// Which points of a target box a transfer visits (ffunction.cc Set_Iterator)
enum Region { REG_ALL = 0, REG_INSIDE = 1, REG_BOUNDARY = 2 };
// Robin kernel modes
enum RobinMode { ROBIN_UPDATE = 0, ROBIN_OPERATOR_R = 1, ROBIN_DEFECT_W = 2 };

void launch_coefficients_box(hipStream_t s, const DevBox &B, Fields F, const PunctureArgs &pa,
                             double robin_a);
void launch_smooth(hipStream_t s, const LevelView &lv, const DevBox *boxes, Fields F,
                   int order, int eq, int colour, unsigned long long *norm);
void launch_residual(hipStream_t s, const LevelView &lv, const DevBox *boxes, Fields F,
                     int order, int eq, int write_defect, unsigned long long *norm);
void launch_robin(hipStream_t s, const DevBox *boxes, int box, const DevBox &hb, Fields F,
                  int mode, double robin_n);
void launch_restrict_u(hipStream_t s, const XferView &xv, const DevBox *boxes, Fields F);
void launch_restrict_defect(hipStream_t s, const XferView &xv, const DevBox *boxes, Fields F,
                            int order, int eq, int mg, double robin_n);
void launch_prolong(hipStream_t s, const XferView &xv, const DevBox *boxes, Fields F, int region);
void launch_ghost_fill(hipStream_t s, const XferView &xv, const DevBox *boxes, Fields F);
// false when the box does not fit the LDS-resident kernel
bool launch_coarse_solve(hipStream_t s, const DevBox *boxes, int box, const DevBox &hb,
                         Fields F, int eq, double ipn, double eps, double omega, int max_iter,
                         unsigned long long *norm, int *iters);
constexpr int kNormSubSlots = 8;  // u64 words per logical norm slot
void launch_zero_u64(hipStream_t s, unsigned long long *p, int n);
void launch_copy(hipStream_t s, double *dst, const double *src, long long n);
void launch_adm_sphere(hipStream_t s, const DevBox *boxes, int nbox, Fields F,
                       const double *centres, double *sph, int nt);
void launch_interp_points(hipStream_t s, const DevBox *boxes, Fields F, int field,
                          const int *pbox, const double *pxyz, int npts, int ord,
                          double *out);

}  // namespace oll
