// Device descriptors and host-side launchers of the CDNA4 kernels.
//
// HBM layout: one slab per hierarchy (request).  Each field (u, b, c, d, e,
// rho, r, v, w, ww) of a box is a contiguous float64 array covering the box
// including its ghost shell of width R, stored by colour: the points of
// colour c = (i+j+k) mod nc (nc = R+1, the smoother's colours) form block c
// of the array, and inside a block point (i, j, k) sits in row (j, k) at
// slot i / nc (rl = ceil(n0/nc) slots per row, rs = rl, plane stride ps,
// block stride cs).  A colour pass of the smoother then reads its own
// points' u, b, c, rho as one unit-stride stream per field and the
// neighbours from the other blocks, rows and planes as further unit-stride
// streams; every neighbour of a point (one step of m < nc along any axis)
// lies in block (c+m) mod nc.  A DevBox carries the base pointers of its
// fields, so one launch can work on boxes of many hierarchies.  Launches
// read device-resident item lists.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace oll {

// F_X0 .. F_X5 exist for PUNCTURES_EM only: grad eta and grad phi, with
// eta in F_B and phi in F_D (F_C, F_E stay 0)
enum Field { F_U = 0, F_B, F_C, F_D, F_E, F_RHO, F_R, F_V, F_W, F_WW,
             F_X0, F_X1, F_X2, F_X3, F_X4, F_X5, NFIELDS };
constexpr int kBaseFields = F_X0;
constexpr int kEqPuncturesEm = 4;
inline int fields_of(int eq) { return eq == kEqPuncturesEm ? int(NFIELDS) : kBaseFields; }
// float32 fields of the mixed-precision smoother (same element layout):
// correction e, linearised diagonal q = dN/du - dulap, defect r = rho - N(u)
enum Field32 { G_E = 0, G_Q, G_R, NFIELDS32 };

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
  int nc;            // colours (R+1)
  int rl;            // slots per row of a colour block, ceil(n0 / nc)
  int rs;            // row stride inside a colour block (= rl)
  unsigned ncm;      // ceil(65536 / nc): x / nc == (x * ncm) >> 16 for x < 16384
  long long ps;      // plane stride inside a colour block, rs * n1
  long long cs;      // colour block stride, ps * n2 rounded up to 32 elements
  long long size;    // elements per field, cs * nc
  float inyi;        // 1 / (n1 - 2R): interior row -> (j, k) by float division
  float in0, in01;   // 1 / n0, 1 / (n0 n1): point index -> (i, j, k)
  double *f[NFIELDS];
  float *g[NFIELDS32];  // null unless the hierarchy runs mixed precision
  // two-colour boxes whose points coinciding with the coarser level's are
  // of colour 1 sweep their colours in descending order (colour pass q
  // updates colour 1 - q): the coincident colour is swept first
  int flip;
  // index box [rlo, rhi) outside which rho is +0 at every interior point:
  // the coefficients write 0 there and afterwards only the FAS source of a
  // restriction writes rho, at the points some finer box covers.  The
  // smoother and the defect read rho inside it only (x - (+0) and (+0) - x
  // are the same bits as x - rho and rho - x there)
  int rlo[3], rhi[3];
};

constexpr int kMaxAxisPoints = 4096;

// x / nc for 0 <= x < 16384
__host__ __device__ inline int split_div(const DevBox &B, int x) {
  return int((unsigned(x) * B.ncm) >> 16);
}
// colour of point (i, j, k)
__host__ __device__ inline int colour_of(const DevBox &B, int i, int j, int k) {
  const int s = i + j + k;
  return s - split_div(B, s) * B.nc;
}
// block shift of a step of m (|m| < nc) from colour c: colour(c+m) - c
__host__ __device__ inline int colour_shift(int nc, int c, int m) {
  return c + m >= nc ? m - nc : (c + m < 0 ? m + nc : m);
}
__host__ __device__ inline long long pidx(const DevBox &B, int i, int j, int k) {
  return (long long)colour_of(B, i, j, k) * B.cs + (long long)k * B.ps + (long long)j * B.rs +
         split_div(B, i);
}
// colours of the multi-colour sweeps: R + 1 for the axis-star Laplacian of
// reach R; 5 for Punctures_EM at orders 4-8, whose one-sided first
// derivatives reach 4 points to one side
constexpr int colours_of(int order, int eq) {
  return (eq == kEqPuncturesEm && order >= 4) ? 5 : order / 2 + 1;
}
// fills the layout members from n, R and the colour count nc
inline void set_layout(DevBox &B, int nc) {
  B.nc = nc;
  B.rl = (B.n[0] + B.nc - 1) / B.nc;
  B.rs = B.rl;
  B.ncm = (65536u + unsigned(B.nc) - 1u) / unsigned(B.nc);
  B.ps = (long long)B.rs * B.n[1];
  B.cs = (B.ps * B.n[2] + 31) & ~31LL;
  B.size = B.cs * B.nc;
  B.inyi = 1.0f / float(B.n[1] - 2 * B.R);
  B.in0 = 1.0f / float(B.n[0]);
  B.in01 = 1.0f / float(B.n[0] * B.n[1]);
}
// Offsets, in units of ES bytes and modulo 2^32, from a point of colour c
// whose i is s mod NC to its neighbours at +-m (m = 1..NC-1) along x, y
// and z: the neighbour lies in colour block (c+m) mod NC, in the same slot
// one row (rs) or plane (ps) away, or along x in the next or previous slot.
template <int NC>
struct Nbr {
  unsigned xp[NC - 1], xm[NC - 1], yp[NC - 1], ym[NC - 1], zp[NC - 1], zm[NC - 1];
};
template <int NC, int ES>
__host__ __device__ inline Nbr<NC> stencil_offs(int c, int s, unsigned cs, unsigned rs, unsigned ps) {
  Nbr<NC> n;
#pragma unroll
  for (int m = 1; m < NC; ++m) {
    const unsigned up = unsigned(colour_shift(NC, c, m)) * cs;
    const unsigned dn = unsigned(colour_shift(NC, c, -m)) * cs;
    n.xp[m - 1] = (up + (s + m >= NC ? 1u : 0u)) * unsigned(ES);
    n.xm[m - 1] = (dn - (s - m < 0 ? 1u : 0u)) * unsigned(ES);
    n.yp[m - 1] = (up + unsigned(m) * rs) * unsigned(ES);
    n.ym[m - 1] = (dn - unsigned(m) * rs) * unsigned(ES);
    n.zp[m - 1] = (up + unsigned(m) * ps) * unsigned(ES);
    n.zm[m - 1] = (dn - unsigned(m) * ps) * unsigned(ES);
  }
  return n;
}

// element offset from a point of colour c with i mod nc = s to its
// neighbour m steps (any sign, |m| <= 4) along axis a
__host__ __device__ inline long long axis_offset(const DevBox &B, int c, int s, int a, int m) {
  // with |m| <= 4 and nc >= 2 two wraps either way suffice
  const int nc = B.nc;
  int c2 = c + m;  // the neighbour's colour, reduced into [0, nc)
  c2 += c2 < 0 ? nc : 0;
  c2 += c2 < 0 ? nc : 0;
  c2 -= c2 >= nc ? nc : 0;
  c2 -= c2 >= nc ? nc : 0;
  long long off = (long long)(c2 - c) * B.cs;
  if (a == 0) {
    // slot change floor((s + m) / nc), s + m in [-4, nc + 3]
    const int q = s + m;
    off += (q < 0 ? -1 : 0) + (q < -nc ? -1 : 0) + (q >= nc ? 1 : 0) + (q >= 2 * nc ? 1 : 0);
  } else {
    off += (long long)m * (a == 1 ? (long long)B.rs : B.ps);
  }
  return off;
}
inline long long layout_size(const int n[3], int nc) {
  const long long block = (long long)((n[0] + nc - 1) / nc) * n[1] * n[2];
  return nc * ((block + 31) & ~31LL);
}

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
  // linear tables that are 2:1 nestings on all axes (affine = 1): target
  // index t maps to q = t + aff[axis], source node q/2 (even q) or the
  // midpoint of nodes (q-1)/2, (q+1)/2 (odd q)
  int aff[3];
  int affine;
  // defect restrictions from a source off the target's lattice: the
  // Lagrange nodes read the filtered defect (F_V, k_filter_defect)
  int filt;
};

// row-parallel launches: one workgroup per RowBlock, which owns `rows`
// consecutive rows of one box starting at row0.  For the colour passes
// (k_smooth, k_smooth32) a row is an interior row (j, k) and a thread owns
// one colour point of it; for the whole-row kernels (residual, defect,
// update) a row is the points of one colour of an interior row (one row of
// that colour's block) and a thread owns one slot of it.
struct RowBlock {
  const DevBox *box;
  unsigned long long *norm;  // null: no norm
  int row0;
  int rows;
  // segments per row: a row longer than the launch's threads per row is
  // split into nseg segments of that many slots, each a row of its own in
  // the block numbering (row index = row * nseg + segment)
  int nseg;
};
constexpr int kRowThreads = 256;  // threads per row workgroup
// the colour passes (k_smooth, k_smooth32) take 64..kMaxRowThreads threads
// per workgroup, a multiple of 64 picked per (precision, order) by the
// launch-shape table (sweep_threads); their results do not depend on it
constexpr int kMaxRowThreads = 512;
int sweep_threads(int mixed, int order);
bool set_sweep_threads(int mixed, int order, int threads);

// fused colour passes (k_smooth_fused): one workgroup per tile of ty x tz
// interior rows (full rows along x) starting at row (j0, k0) of a box
struct TileBlock {
  const DevBox *box;
  unsigned long long *norm;
  int j0, k0;
  int ty, tz;
};
constexpr int kFusedThreads = 512;  // at most; the tile-shape table picks 64..512

// Fused passes (k_smooth_fused): a neighbour m steps away (m < 0: below)
// of a point of pass index pi has pass index (pi + m) mod nc (the colours
// along every axis follow the pass order; two-colour boxes whose order is
// flipped alike); it was updated earlier in the sweep, and lives in the
// written field Y, when that index precedes pi.
__host__ __device__ constexpr bool nbr_in_y(int nc, int pi, int m) {
  return (((pi + m) % nc) + nc) % nc < pi;
}
// i of the first interior point of colour c in row (j, k)
__host__ __device__ inline int colour_first(int c, int j, int k, int R, int nc) {
  return R + int(unsigned(c + nc * 3 * kMaxAxisPoints - j - k - R) % unsigned(nc));
}
// interior row r of a tile's row list: the tile's rows (y fastest), then
// the rows its second pass reads outside it: yb rows below (0 or 1) and ya
// above (up to R) along y for the tile's planes, zb planes below and the
// rest above along z for the tile's rows, those inside the interior
__host__ __device__ inline void tile_row(const TileBlock &T, int r, int yb, int ya, int zb, int &j,
                                         int &k) {
  const int nt = T.ty * T.tz;
  if (r < nt) {
    const int kk = r / T.ty;
    j = T.j0 + (r - kk * T.ty);
    k = T.k0 + kk;
    return;
  }
  r -= nt;
  if (r < yb * T.tz) {
    j = T.j0 - 1;
    k = T.k0 + r;
    return;
  }
  r -= yb * T.tz;
  if (r < ya * T.tz) {
    const int l = r / T.tz;
    j = T.j0 + T.ty + l;
    k = T.k0 + (r - l * T.tz);
    return;
  }
  r -= ya * T.tz;
  if (r < zb * T.ty) {
    j = T.j0 + r;
    k = T.k0 - 1;
    return;
  }
  r -= zb * T.ty;
  const int l = r / T.ty;
  k = T.k0 + T.tz + l;
  j = T.j0 + (r - l * T.ty);
}
// halo counts of a tile of box B with ghost width R (rows inside the interior)
__host__ __device__ inline void tile_halo(const TileBlock &T, const int n[3], int R, int &yb, int &ya,
                                          int &zb, int &za) {
  yb = T.j0 - 1 >= R ? 1 : 0;
  ya = n[1] - R - (T.j0 + T.ty);
  ya = ya < R ? ya : R;
  zb = T.k0 - 1 >= R ? 1 : 0;
  za = n[2] - R - (T.k0 + T.tz);
  za = za < R ? za : R;
}
// tile shape and threads per workgroup of the fused passes per order (the
// autotuner's knobs); results do not depend on them
struct FusedShape {
  int ty, tz, threads;
};
FusedShape fused_shape(int order);
bool set_fused_shape(int order, int ty, int tz, int threads);

// point-parallel launches (transfers): one workgroup per PtBlock, which
// owns target points q0 .. q0+nq-1 (nq <= kPtThreads) of one box, either
// its lexicographic points (i fastest) or its ghost-shell points in shell
// order (the ghost fill)
struct PtBlock {
  const DevBox *box;
  const DevXfer *src;  // source tables in source order
  int nsrc;
  int mg;              // multigrid level (Robin shell, "all" coverage)
  unsigned q0;
  unsigned nq;
  int i0, j0, k0;      // lexicographic blocks: coordinates of point q0
  // w > 0: a patch of whole rows of w points in plane k0, from (i0, j0):
  // offset off is point (i0 + off % w, j0 + off / w, k0); inw = 1 / w
  int w;
  float inw;
};
constexpr int kPtThreads = 256;

// Robin stages: face columns col0..col0+ncol of a box
struct RobinItem {
  const DevBox *box;
  int col0;   // first face column of the box in the launch
  int ncol;
  int ecol0;  // first edge cross-section of the box in the launch
};

struct CoarseItem {
  const DevBox *box;
  unsigned long long *norm;
  int *iters;
  double omega;
};

struct Puncture {
  double m, x, y, z, px, py, pz, sx, sy, sz, q;
};

constexpr int kMaxBH = 16;

// elliptic.cc:2037-2213 (the same terms at :2286-2419): puncture b's term of
// the analytic approximation at (x, y, z), the spin (J), momentum (P) and
// cross (P x J) parts of the perturbative single-puncture solution.  Host
// and device evaluate it operation for operation alike.
__host__ __device__ inline double approx_puncture_u(const Puncture &b, double x, double y, double z) {
#pragma clang fp contract(off)
  const double sbl = sqrt(b.sx * b.sx + b.sy * b.sy + b.sz * b.sz);
  const double imass = 1.0 / b.m;
  const double imass2 = imass * imass;
  const double Jx = 4.0 * b.sx * imass2, Jy = 4.0 * b.sy * imass2, Jz = 4.0 * b.sz * imass2;
  const double J2 = Jx * Jx + Jy * Jy + Jz * Jz;
  const double pbl = sqrt(b.px * b.px + b.py * b.py + b.pz * b.pz);
  const double Px = 2.0 * b.px * imass, Py = 2.0 * b.py * imass, Pz = 2.0 * b.pz * imass;
  const double P2 = Px * Px + Py * Py + Pz * Pz;
  const double xi = x - b.x, yi = y - b.y, zi = z - b.z;
  double r = sqrt(xi * xi + yi * yi + zi * zi);
  if (fabs(r) <= 0.00000000001) r = 0.0000000001;
  const double Rx = 2.0 * xi * imass, Ry = 2.0 * yi * imass, Rz = 2.0 * zi * imass;
  const double R = 2.0 * r * imass;
  const double l1 = 1.0 / (1.0 + R);
  const double l2 = l1 * l1, l3 = l2 * l1, l4 = l3 * l1, l5 = l4 * l1;
  const double u0j = 0.025 * (l1 + l2 + l3 - 4.0 * l4 + 2.0 * l5);
  const double u2j = -0.05 * l5;
  const double muj = (sbl != 0.0 && r != 0.0) ? (b.sx * xi + b.sy * yi + b.sz * zi) / (r * sbl) : 0.0;
  const double PL2j = 0.5 * (3.0 * muj * muj - 1.0);
  const double uj = J2 * (u0j + u2j * R * R * PL2j);
  const double u0p = 0.15625 * (l1 - 2.0 * l2 + 2.0 * l3 - l4 + 0.2 * l5);
  const double u2p = 0.0125 *
                     (15.0 * l1 + 132.0 * l2 + 53.0 * l3 + 96.0 * l4 + 82 * l5 + 84.0 * (l5 + log(l1) / R) / R) /
                     R;
  const double mup = (pbl != 0.0 && r != 0.0) ? (b.px * xi + b.py * yi + b.pz * zi) / (r * pbl) : 0.0;
  const double PL2p = 0.5 * (3.0 * mup * mup - 1.0);
  const double up = P2 * (u0p + u2p * PL2p);
  const double c0 = Py * Jz - Pz * Jy, c1 = Pz * Jx - Px * Jz, c2 = Px * Jy - Py * Jx;
  const double uc = 0.0125 * (c0 * Rx + c1 * Ry + c2 * Rz) * (1.0 + 5.0 * R + 10.0 * R * R) * l5;
  return uj + up + uc;
}

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

// A logical norm slot holds kNormSubSlots partial maxima, each in its own
// 64-byte line (kNormStride u64 words apart); blocks spread their atomics
// over them and the host takes the max.
constexpr int kNormSubSlots = 16;
constexpr int kNormStride = 8;
constexpr int kNormWords = kNormSubSlots * kNormStride;  // u64 words per logical slot
// colour slot list of the coarse solve: ceil(ni / 2) * nj * nk points
inline long long coarse_colour_slots(const int n[3]) {
  return (long long)((n[0] - 1) / 2) * (n[1] - 2) * (n[2] - 2);
}
// largest coarsest box the LDS-resident solve (512 threads) takes: <= 8
// points per colour per thread, edge points <= 512, u and rho within the LDS
inline bool coarse_box_fits(const int n[3], long long size) {
  const long long ni = n[0] - 2, nj = n[1] - 2, nk = n[2] - 2;
  return coarse_colour_slots(n) <= 8 * 512 && 4 * (ni + nj + nk) <= 512 &&
         2 * 8 * size + 2048 <= 160 * 1024;
}

void launch_coefficients_box(hipStream_t s, const DevBox &B, const PunctureArgs &pa,
                             double robin_a);
// per_row: threads per row (colour points per interior row for the colour
// passes, slots per colour row for the whole-row kernels); the host sizes
// RowBlock::rows to at most kRowThreads / per_row
// Row kernels address fields with 32-bit byte offsets: a box field must stay
// below 4 GiB (checked when the hierarchy is built).
constexpr long long kMaxFieldBytes = 1LL << 32;
// e0, e1 (optional): events the runtime stamps at the kernel's start and end
// box1: the device descriptor of the one box every block lists, or null
void launch_smooth(hipStream_t s, const RowBlock *blocks, int nblocks, int per_row, int order,
                   int eq, int colour, int threads, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr,
                   const DevBox *box1 = nullptr);
// what k_residual writes on the interior: r = N(u), w = rho - N(u), or
// r = dN/du_p (the Newton diagonal, elliptic.h:337-345 duOPERATOR)
enum ResidualMode { RES_OPERATOR = 0, RES_DEFECT = 1, RES_DERIVATIVE = 2 };
void launch_residual(hipStream_t s, const RowBlock *blocks, int nblocks, int per_row, int order,
                     int eq, int mode);
// passes q and q+1 (pair) or pass q alone of one sweep, double-buffered:
// parity 0 reads u and writes r, parity 1 the reverse (eq 0-3)
void launch_smooth_fused(hipStream_t s, const TileBlock *blocks, int nblocks, int order, int eq,
                         int q, int pair, int parity, int threads, hipEvent_t e0 = nullptr,
                         hipEvent_t e1 = nullptr);
// r = u on the ghost shell (PtBlocks over shell points)
void launch_copy_shell(hipStream_t s, const PtBlock *blocks, int nblocks);
// mixed precision: FP64 defect -> FP32 (r, q, e = 0), FP32 colour pass of
// the linearised correction equation, FP64 update u += e
void launch_defect32(hipStream_t s, const RowBlock *blocks, int nblocks, int per_row, int order,
                     int eq);
void launch_smooth32(hipStream_t s, const RowBlock *blocks, int nblocks, int per_row, int order,
                     int colour, int threads, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_update32(hipStream_t s, const RowBlock *blocks, int nblocks, int per_row, int order);
// faces (total_cols columns), then edges (total_edges cross-sections),
// then corners of the listed boxes: three launches
void launch_robin(hipStream_t s, const RobinItem *items, int n, int total_cols, int total_edges,
                  int mode);
// Ghost-shell columns by faces (z, then y, then x slabs, low side first),
// each face's columns in tiles of kColSuper x kColSuper, one per workgroup
// of 256 threads, each made of four kColTile x kColTile tiles, one per wave
// (so a wave's and a workgroup's stencils cover compact patches of the
// coarser box); the tiles at a face's far edges are padded to full size and
// their extra slots idle.  Slots of one face side of F0 x F1 columns:
constexpr int kColTile = 8;
constexpr int kColSuper = 16;
__host__ __device__ inline unsigned col_tiles(unsigned f) { return (f + kColSuper - 1) / kColSuper; }
__host__ __device__ inline unsigned face_slots(unsigned f0, unsigned f1) {
  return col_tiles(f0) * col_tiles(f1) * unsigned(kColSuper * kColSuper);
}
// every column slot of a box's ghost shell
inline long long ghost_column_slots(const DevBox &b) {
  const unsigned R = unsigned(b.R), n0 = unsigned(b.n[0]), n1 = unsigned(b.n[1]),
                 n2 = unsigned(b.n[2]);
  return 2LL * (face_slots(n0, n1) + face_slots(n0, n2 - 2 * R) + face_slots(n1 - 2 * R, n2 - 2 * R));
}
// edge cross-sections of a box: 4 (n0 + n1 + n2 - 6R)
inline int robin_edge_sections(const DevBox &B) { return 4 * (B.n[0] + B.n[1] + B.n[2] - 6 * B.R); }
void launch_restrict_u(hipStream_t s, const PtBlock *blocks, int nblocks);
void launch_restrict_defect(hipStream_t s, const PtBlock *blocks, int nblocks, int order, int eq);
// generic transfer of field ff of the listed blocks' sources into field
// tf of their targets (multigrid::Transferir): PtBlock::mg packs the
// region (bits 0-1), `all` (bit 2) and `to_zero` (bit 3); nodes = 2 for
// linear interpolation, else the Lagrange nodes per axis (3..10); false
// for other node counts
bool launch_transfer_fields(hipStream_t s, const PtBlock *blocks, int nblocks, int ff, int tf,
                            int nodes);
// v = full-weighting filter of the defect w (all points of the listed
// boxes); runs = 1: blocks over runs of filter_run_length() points along x
// (ceil(n0 / length) runs per row), 0: blocks over points
void launch_filter_defect(hipStream_t s, const PtBlock *blocks, int nblocks, int runs);
int filter_run_length();
// v = u - ww on the listed coarse points (before a prolongation)
void launch_corr_v(hipStream_t s, const PtBlock *blocks, int nblocks);
void launch_prolong(hipStream_t s, const PtBlock *blocks, int nblocks);
// order: stencil order of the target (refinement) boxes; Lagrange order+1
// columns = 1: blocks over shell columns (R points across a face slab, one
// thread each); 0: blocks over shell points
void launch_ghost_fill(hipStream_t s, const PtBlock *blocks, int nblocks, int order, int columns);
// false when the largest box does not fit the LDS-resident kernel
bool launch_coarse_solve(hipStream_t s, const CoarseItem *items, int n, int max_pts, int max_int,
                         int eq, double eps, int max_iter);
void launch_zero_u64(hipStream_t s, unsigned long long *p, long long n);
// words to clear at one address (each member's own count: the arrays differ)
struct ZeroItem {
  unsigned long long *p;
  long long count;
};
void launch_zero_lists(hipStream_t s, const ZeroItem *items, int n);
// one-time kernel attributes (dynamic LDS of the coarse solve); call before
// any graph capture
void kernels_init();
// dynamic LDS the coarse solve can use
constexpr int kCoarseLdsLimit = 160 * 1024;
void launch_adm_sphere(hipStream_t s, const DevBox *boxes, int nbox, const double *centres,
                       double *sph, int nt);
// u on every point of box B set to the analytic approximation
// (elliptic.cc:1977-2222)
void launch_approx_box(hipStream_t s, const DevBox &B, const PunctureArgs &pa);
void launch_interp_points(hipStream_t s, const DevBox *boxes, int field, const int *pbox,
                          const double *pxyz, int npts, int ord, double *out);

}  // namespace oll
