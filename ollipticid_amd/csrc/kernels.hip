// CDNA4 (gfx950) kernels of the FAS puncture solver.
//
// All stencils are axis stars of radius R = order/2, so the points split
// into R+1 colours (i+j+k) mod (R+1) with no same-colour neighbours: one
// launch updates one colour of every listed box.  A launch works on a
// device-resident item list whose boxes may belong to many hierarchies
// (requests); arithmetic is float64 throughout.
//
// Reference counterparts (WuShichao/OllipticID):
//   k_smooth            elliptic.cc:1126-1242 Smooth, elliptic.h:405-445 ops
//   k_residual          elliptic.cc:944-991 App_Ell_Operator
//   k_robin_*           boundary.cc:447-1101 App_Bound, :1118-1710 operator
//   k_restrict_u        elliptic.cc:1543-1546 (Transferir INSIDE LAGRANGE 3)
//   k_restrict_defect   elliptic.cc:1625-1747 (W = S - R, restrict, S = W + R)
//   k_prolong           elliptic.cc:1827-1902 (V = u - ww, LINEAR, u += V)
//   k_ghost_fill        elliptic.cc:1905-1908 (BOUNDARY LAGRANGE order+1)
//   k_coarse_solve      elliptic.cc:1782-1816 (coarsest-level loop)
//   k_coefficients      operator.cc:752-868, 2193-2380 (b, c, d, e)
//   k_adm_sphere        elliptic.cc:2447-2611 Compute_ADM_Mass
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdio>

#include "kernels.hpp"

// No implicit contraction: every kernel's rounding is fixed by its source,
// whatever the launch shape or unrolling (a request solves to the same bits
// alone and in a batch).  Hot expressions use explicit fma.
#pragma clang fp contract(off)

namespace oll {

namespace {

constexpr double kEps = DBL_EPSILON;

// ------------------------------------------------------------------
// point operators
// ------------------------------------------------------------------

// axis-star second differences (explicit fma; shared by all stencil kernels)
template <int ORD, typename T>
__device__ __forceinline__ T star(T u0, T s1, T s2, T s3, T s4) {
  if constexpr (ORD == 2) {
    return fma(T(-2), u0, s1);
  } else if constexpr (ORD == 4) {
    return fma(T(-30), u0, fma(T(16), s1, -s2));
  } else if constexpr (ORD == 6) {
    return fma(T(-490), u0, fma(T(270), s1, fma(T(-27), s2, T(2) * s3)));
  } else {
    return fma(T(-14350), u0, fma(T(8064), s1, fma(T(-1008), s2, fma(T(128), s3, T(-9) * s4))));
  }
}
template <int ORD, typename T>
__device__ __forceinline__ T star_sum(T ax, T ay, T az, T ix2, T iy2, T iz2) {
  const T s = fma(ax, ix2, fma(ay, iy2, az * iz2));
  if constexpr (ORD == 2) return s;
  else if constexpr (ORD == 4) return s * (T(1) / T(12));
  else if constexpr (ORD == 6) return s * (T(1) / T(180));
  else return s * (T(1) / T(5040));
}

// axis-star Laplacian at element index p of an array (LDS); neighbour
// element offsets from stencil_offs<NC, 1>
template <int ORD, typename T = double, int NC = ORD / 2 + 1>
__device__ __forceinline__ T lap(const T *__restrict__ u, int p, const Nbr<NC> &n, T ix2,
                                 T iy2, T iz2) {
  constexpr int R = ORD / 2;
  T sx[4] = {0, 0, 0, 0}, syv[4] = {0, 0, 0, 0}, szv[4] = {0, 0, 0, 0};
#pragma unroll
  for (int m = 1; m <= R; ++m) {
    sx[m - 1] = u[p + int(n.xp[m - 1])] + u[p + int(n.xm[m - 1])];
    syv[m - 1] = u[p + int(n.yp[m - 1])] + u[p + int(n.ym[m - 1])];
    szv[m - 1] = u[p + int(n.zp[m - 1])] + u[p + int(n.zm[m - 1])];
  }
  const T u0 = u[p];
  return star_sum<ORD, T>(star<ORD, T>(u0, sx[0], sx[1], sx[2], sx[3]),
                          star<ORD, T>(u0, syv[0], syv[1], syv[2], syv[3]),
                          star<ORD, T>(u0, szv[0], szv[1], szv[2], szv[3]), ix2, iy2, iz2);
}

// N(u) and dN/du_p: EQ 0 Lap u + b (c+u)^-7, EQ 1 + d (c+u) + e (c+u)^5,
// EQ 2 Lap u + u^2, EQ 3 Lap u + b u (c = 0 for EQ 2, 3, so cu = u)
template <int EQ>
__device__ __forceinline__ void nonlinear(double lapu, double dulap, double b, double cu,
                                          double dd, double ee, double &N, double &dN) {
  if constexpr (EQ == 2) {
    N = lapu + cu * cu;
    dN = dulap + 2.0 * cu;
    return;
  } else if constexpr (EQ == 3) {
    N = lapu + b * cu;
    dN = dulap + b;
    return;
  }
  const double inv = 1.0 / cu;
  const double inv2 = inv * inv;
  const double inv4 = inv2 * inv2;
  const double inv7 = inv4 * inv2 * inv;
  const double inv8 = inv4 * inv4;
  N = fma(b, inv7, lapu);
  dN = fma(-7.0 * b, inv8, dulap);
  if constexpr (EQ == 1) {
    const double c2 = cu * cu;
    const double c4 = c2 * c2;
    N = fma(ee, c4 * cu, fma(dd, cu, N));
    dN = fma(5.0 * ee, c4, dN + dd);
  }
}

// Punctures_EM (elliptic.h:510-626) at an interior point: u, its gradient g
// (du: each derivative's coefficient of u_p) and, from the mass and charge
// potentials eta, phi and their gradients ge, gf, the thirteen coefficients
// (elliptic.h:1458-1530)
//   N  = Lap u + num / den,
//   num = B |g|^2 + (B1 u + B2).g + C u^2 + D u + E,
//   den = u^3 + A1 u^2 + A2 u + A3,
// and dN/du_p
__device__ __forceinline__ void em_nonlinear(double lapu, double dulap, double u, const double g[3],
                                             const double du[3], double eta, double phi,
                                             const double ge[3], const double gf[3], double &N,
                                             double &dN) {
  const double f2 = gf[0] * gf[0] + gf[1] * gf[1] + gf[2] * gf[2];
  const double fe = gf[0] * ge[0] + gf[1] * ge[1] + gf[2] * ge[2];
  const double e2 = ge[0] * ge[0] + ge[1] * ge[1] + ge[2] * ge[2];
  const double a1 = 3 * eta;
  const double a2 = 3 * (eta * eta) - 0.25 * (phi * phi);
  const double a3 = eta * eta * eta - 0.25 * eta * phi * phi;
  const double bb = -0.25 * (phi * phi);
  const double cc = -0.25 * f2;
  const double dd = 0.5 * phi * fe - 0.5 * eta * f2;
  const double ee = -0.25 * phi * phi * e2 + 0.5 * phi * eta * fe + 0.25 * (1 - eta * eta) * f2;
  double b1[3], b2[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    b1[a] = 0.5 * phi * gf[a];
    b2[a] = 0.5 * phi * (eta * gf[a] - phi * ge[a]);
  }
  const double u2 = u * u;
  const double num = bb * (g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) + (b1[0] * u + b2[0]) * g[0] +
                     (b1[1] * u + b2[1]) * g[1] + (b1[2] * u + b2[2]) * g[2] + cc * u2 + dd * u + ee;
  const double den = u2 * u + a1 * u2 + a2 * u + a3;
  const double dnum = 2 * bb * (g[0] * du[0] + g[1] * du[1] + g[2] * du[2]) + (b1[0] * u + b2[0]) * du[0] +
                      b1[0] * g[0] + (b1[1] * u + b2[1]) * du[1] + b1[1] * g[1] +
                      (b1[2] * u + b2[2]) * du[2] + b1[2] * g[2] + 2 * cc * u + dd;
  N = lapu + num / den;
  dN = dulap + (dnum * den - num * (3 * u2 + 2 * a1 * u + a2)) / (den * den);
}


// Norm slots: non-negative doubles order like their bit patterns, so a max
// is an unsigned atomic max.  The atomic returns nothing, so the issuing
// wave does not wait for it.
__device__ __forceinline__ void norm_update(unsigned long long *slot, double v) {
  unsigned long long *s = slot + (blockIdx.x & (kNormSubSlots - 1)) * kNormStride;
  atomicMax(s, (unsigned long long)__double_as_longlong(v));
}

// exact g / d for g < 2^24 via a float reciprocal and one correction step
__device__ __forceinline__ unsigned fast_div(unsigned g, unsigned d, float inv_d) {
  unsigned q = unsigned(float(g) * inv_d);
  if (q * d > g) --q;
  else if ((q + 1) * d <= g) ++q;
  return q;
}

// coordinate of index i, lo + i d rounded twice as on the host (the file's
// contract(off) keeps the two operations apart; the __dmul_rn/__dadd_rn
// helpers are compiled under the headers' own contraction setting and fuse)
__device__ __forceinline__ double coord(const DevBox &B, int a, int i) {
  return B.lo[a] + (double)i * B.d[a];
}

__device__ __forceinline__ bool interior(const DevBox &B, int i, int j, int k) {
  const int R = B.R;
  return i >= R && i < B.n[0] - R && j >= R && j < B.n[1] - R && k >= R && k < B.n[2] - R;
}

__device__ __forceinline__ void point_ijk(const DevBox &B, long long q, int &i, int &j, int &k) {
  const unsigned n0 = unsigned(B.n[0]), n1 = unsigned(B.n[1]);
  const unsigned qq = unsigned(q);
  const unsigned plane = n0 * n1;
  k = int(qq / plane);
  const unsigned r = qq - unsigned(k) * plane;
  j = int(r / n0);
  i = int(r - unsigned(j) * n0);
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Remapping so that XCD x runs one contiguous run of logical blocks
// keeps the rows and planes a stencil shares inside one L2.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_block() {
  const unsigned b = blockIdx.x, nb = gridDim.x;
  const unsigned q = nb / kXcds, rem = nb % kXcds;
  const unsigned x = b % kXcds, idx = b / kXcds;
  return x * q + (x < rem ? x : rem) + idx;
}

// Row kernels read their box descriptor through the constant address space
// (scalar loads: kernels never write descriptors) and address fields with
// 32-bit byte offsets from a uniform global base.
typedef const __attribute__((address_space(4))) DevBox ConstBox;
typedef __attribute__((address_space(1))) char gchar;

__device__ __forceinline__ ConstBox &cbox(const DevBox *b) { return *(ConstBox *)b; }
// brings every cache line of a box descriptor into the scalar cache with
// loads issued together (one dword per 64 bytes, and the last), and waits
// once for them
__device__ __forceinline__ void fetch_box(const DevBox *b) {
  typedef const __attribute__((address_space(4))) unsigned CWord;
  CWord *w = (CWord *)b;
  constexpr int kWords = int(sizeof(DevBox) / 4);
  unsigned acc = w[kWords - 1];
#pragma unroll
  for (int q = 0; q < kWords; q += 16) acc ^= w[q];
  asm volatile("" ::"s"(acc));
}
// the generic helpers (robin_op_at) take a generic reference
__device__ __forceinline__ const DevBox &cbox_plain(ConstBox &b) { return *(const DevBox *)&b; }
// x / nc and the colour of (i, j, k) from a constant-space box (kernels.hpp
// split_div, colour_of)
__device__ __forceinline__ unsigned divc(ConstBox &B, unsigned x) { return (x * B.ncm) >> 16; }
__device__ __forceinline__ unsigned colourc(ConstBox &B, int i, int j, int k) {
  const unsigned s = unsigned(i + j + k);
  return s - divc(B, s) * unsigned(B.nc);
}

template <typename T>
__device__ __forceinline__ T ld(const void *base, unsigned off) {
  return *(const __attribute__((address_space(1))) T *)((const gchar *)base + off);
}
template <typename T>
__device__ __forceinline__ void st(void *base, unsigned off, T v) {
  *(__attribute__((address_space(1))) T *)((gchar *)base + off) = v;
}
// rho at byte offset o of point (i, j, k): +0 outside B's rho box without
// a load.  Callers take it before the stencil so that the load, when there
// is one, is in flight with the stencil's
__device__ __forceinline__ double rho_at(ConstBox &B, unsigned o, int i, int j, int k) {
  const bool in = (unsigned(i - B.rlo[0]) < unsigned(B.rhi[0] - B.rlo[0])) &
                  (unsigned(j - B.rlo[1]) < unsigned(B.rhi[1] - B.rlo[1])) &
                  (unsigned(k - B.rlo[2]) < unsigned(B.rhi[2] - B.rlo[2]));
  double r = 0.0;
  if (in) r = ld<double>(B.f[F_RHO], o);
  return r;
}

// axis-star Laplacian at byte offset o of a field; neighbour byte offsets
// from stencil_offs<NC, sizeof(T)>
template <int ORD, typename T, int NC = ORD / 2 + 1>
__device__ __forceinline__ T lapb(const void *u, unsigned o, const Nbr<NC> &n, T ix2, T iy2,
                                  T iz2) {
  constexpr int R = ORD / 2;
  T sx[4] = {0, 0, 0, 0}, syv[4] = {0, 0, 0, 0}, szv[4] = {0, 0, 0, 0};
#pragma unroll
  for (int m = 1; m <= R; ++m) {
    sx[m - 1] = ld<T>(u, o + n.xp[m - 1]) + ld<T>(u, o + n.xm[m - 1]);
    syv[m - 1] = ld<T>(u, o + n.yp[m - 1]) + ld<T>(u, o + n.ym[m - 1]);
    szv[m - 1] = ld<T>(u, o + n.zp[m - 1]) + ld<T>(u, o + n.zm[m - 1]);
  }
  const T u0 = ld<T>(u, o);
  return star_sum<ORD, T>(star<ORD, T>(u0, sx[0], sx[1], sx[2], sx[3]),
                          star<ORD, T>(u0, syv[0], syv[1], syv[2], syv[3]),
                          star<ORD, T>(u0, szv[0], szv[1], szv[2], szv[3]), ix2, iy2, iz2);
}

// Punctures_EM's first derivative along one axis at an interior point, as
// the reference's solver evaluates it: dx2 (centred, diff_2nd.cc:238-300) at
// order 2; at orders 4-8 the one-sided five-point stencil, forward on the
// lower half of the axis (index <= n/2, rounded down) and backward above,
// with du its coefficient of the point itself.  p, m: the +1..+4 and
// -1..-4 neighbour byte offsets; ia, na: index and points along the axis
template <int ORD>
__device__ __forceinline__ double em_deriv(const void *u, unsigned o, double u0, const unsigned *pp,
                                           const unsigned *mm, int ia, int na, double id, double &du) {
  if constexpr (ORD == 2) {
    du = 0.0;
    return (ld<double>(u, o + pp[0]) - ld<double>(u, o + mm[0])) * 0.5 * id;
  } else {
    const int nm = na % 2 == 0 ? na / 2 : (na - 1) / 2;
    if (ia <= nm) {
      du = -25 * id / 12.0;
      return -(25 * u0 - 48 * ld<double>(u, o + pp[0]) + 36 * ld<double>(u, o + pp[1]) -
               16 * ld<double>(u, o + pp[2]) + 3 * ld<double>(u, o + pp[3])) *
             id / 12.0;
    }
    du = 25 * id / 12.0;
    return (25 * u0 - 48 * ld<double>(u, o + mm[0]) + 36 * ld<double>(u, o + mm[1]) -
            16 * ld<double>(u, o + mm[2]) + 3 * ld<double>(u, o + mm[3])) *
           id / 12.0;
  }
}

// N(u) and dN/du at byte offset o of box B (u0 = u there, (i, j, k) its
// indices): Lap u from the neighbour offsets nb, the other terms from the
// coefficient fields
template <int ORD, int EQ, class BoxT>
__device__ __forceinline__ void point_op(BoxT &B, const void *u, unsigned o,
                                         const Nbr<colours_of(ORD, EQ)> &nb, double u0, int i, int j,
                                         int k, double &N, double &dN) {
  constexpr int NC = colours_of(ORD, EQ);
  const double lu = lapb<ORD, double, NC>(u, o, nb, B.idx2[0], B.idx2[1], B.idx2[2]);
  if constexpr (EQ == kEqPuncturesEm) {
    double g[3], du[3];
    g[0] = em_deriv<ORD>(u, o, u0, nb.xp, nb.xm, i, B.n[0], B.idx[0], du[0]);
    g[1] = em_deriv<ORD>(u, o, u0, nb.yp, nb.ym, j, B.n[1], B.idx[1], du[1]);
    g[2] = em_deriv<ORD>(u, o, u0, nb.zp, nb.zm, k, B.n[2], B.idx[2], du[2]);
    const double ge[3] = {ld<double>(B.f[F_X0], o), ld<double>(B.f[F_X1], o), ld<double>(B.f[F_X2], o)};
    const double gf[3] = {ld<double>(B.f[F_X3], o), ld<double>(B.f[F_X4], o), ld<double>(B.f[F_X5], o)};
    em_nonlinear(lu, B.dulap, u0, g, du, ld<double>(B.f[F_B], o), ld<double>(B.f[F_D], o), ge, gf, N, dN);
  } else {
    (void)i;
    (void)j;
    (void)k;
    nonlinear<EQ>(lu, B.dulap, ld<double>(B.f[F_B], o), ld<double>(B.f[F_C], o) + u0,
                  EQ == 1 ? ld<double>(B.f[F_D], o) : 0.0, EQ == 1 ? ld<double>(B.f[F_E], o) : 0.0, N, dN);
  }
}

// one DPP step on a double: lanes of rows in ROW_MASK read lane CTRL's value
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = int(b), hi = int(b >> 32);
  const int l2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xF, false);
  const int h2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)h2 << 32) | (unsigned)l2));
}

// max over the 64 lanes of a wave (all lanes active); uniform result
__device__ __forceinline__ double wave_max_dpp(double v) {
  v = fmax(v, dpp_f64<0xB1, 0xF>(v));   // quad_perm [1,0,3,2]
  v = fmax(v, dpp_f64<0x4E, 0xF>(v));   // quad_perm [2,3,0,1]
  v = fmax(v, dpp_f64<0x141, 0xF>(v));  // row_half_mirror
  v = fmax(v, dpp_f64<0x140, 0xF>(v));  // row_mirror
  v = fmax(v, dpp_f64<0x142, 0xA>(v));  // row_bcast15
  v = fmax(v, dpp_f64<0x143, 0xC>(v));  // row_bcast31: lane 63 holds the max
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(b), 63);
  const int hi = __builtin_amdgcn_readlane(int(b >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// block max of v (every thread must call; blockDim.x a multiple of 64, at
// most kMaxRowThreads), one norm update per block
__device__ __forceinline__ void wave_norm(unsigned long long *slot, double v) {
  __shared__ double sh[kMaxRowThreads / 64];
  const double m = wave_max_dpp(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0 && slot) {
    double b = sh[0];
    const int nw = int(blockDim.x >> 6);
    for (int w = 1; w < nw; ++w) b = fmax(b, sh[w]);
    if (b > 0.0) norm_update(slot, b);
  }
}

// the block's interior row r -> (j, k)
template <int R>
__device__ __forceinline__ void row_jk(ConstBox &B, unsigned r, int &j, int &k) {
  const unsigned nyi = unsigned(B.n[1] - 2 * R);
  const unsigned kq = fast_div(r, nyi, B.inyi);
  k = R + int(kq);
  j = R + int(r - kq * nyi);
}

// colour passes (NC colours): the thread's slot t of local row lr: element
// index p of its point (i, j, k) of colour `colour` and that point's i mod
// NC; false when the row is past the block or the slot past the row's
// interior colour points
template <int R, int NC = R + 1>
__device__ __forceinline__ bool colour_slot(const RowBlock &Bk, ConstBox &B, unsigned lr, int t,
                                            unsigned per_row, int colour, unsigned &p, int &s,
                                            int &i, int &j, int &k) {
  if (lr >= unsigned(Bk.rows)) return false;
  unsigned row = unsigned(Bk.row0) + lr;
  if (Bk.nseg > 1) {  // segment `seg` of the row: colour points seg * per_row on
    const unsigned seg = row % unsigned(Bk.nseg);
    row /= unsigned(Bk.nseg);
    t += int(seg * per_row);
  }
  row_jk<R>(B, row, j, k);
  // i of the row's first interior point of this colour
  const int first = R + int(unsigned(colour + NC * 3 * kMaxAxisPoints - j - k - R) % NC);
  i = first + t * NC;
  if (i >= B.n[0] - R) return false;
  s = first % NC;
  p = unsigned(colour) * unsigned(B.cs) + unsigned(k) * unsigned(B.ps) + unsigned(j) * unsigned(B.rs) +
      unsigned(first / NC + t);
  return true;
}

// whole-row kernels: sub-row sr = NC * row + c is colour c of interior row
// `row` (with segments, sr * nseg + segment); the thread owns slot t of it
// (of the segment).  Element index p, colour c and i mod NC; false for
// ghost and padding slots
template <int R, int NC = R + 1>
__device__ __forceinline__ bool row_slot(const RowBlock &Bk, ConstBox &B, unsigned per_row,
                                         float inv_per_row, unsigned rpb, unsigned lr0, unsigned &p,
                                         int &c, int &s, int &i, int &j, int &k) {
  const unsigned lrl = fast_div(threadIdx.x, per_row, inv_per_row);
  int t = int(threadIdx.x - lrl * per_row);
  // threads past the chunk's whole rows idle (the next chunk has them)
  const unsigned lr = lr0 + lrl;
  if (lrl >= rpb || lr >= unsigned(Bk.rows)) return false;
  unsigned sr = unsigned(Bk.row0) + lr;
  if (Bk.nseg > 1) {  // segment `seg` of the colour row: slots seg * per_row on
    const unsigned seg = sr % unsigned(Bk.nseg);
    sr /= unsigned(Bk.nseg);
    t += int(seg * per_row);
  }
  const unsigned row = sr / NC;
  c = int(sr - row * NC);
  row_jk<R>(B, row, j, k);
  s = int(unsigned(c + NC * 3 * kMaxAxisPoints - j - k) % NC);
  i = s + t * NC;
  if (i < R || i >= B.n[0] - R) return false;
  p = unsigned(c) * unsigned(B.cs) + unsigned(k) * unsigned(B.ps) + unsigned(j) * unsigned(B.rs) +
      unsigned(t);
  return true;
}

// the larger of acc and a, a NaN in either kept (one step from acc = 0
// gives a itself, NaN included)
__device__ __forceinline__ double nan_max(double acc, double a) {
  return (a > acc || !(a == a)) ? a : acc;
}

// a whole-row block's rows, rpb = blockDim / per_row at a time: f(rpb,
// lr0) for the first local row lr0 of each chunk (blocks of one launch
// over many requests hold several chunks)
template <class F>
__device__ __forceinline__ void for_rows(const RowBlock &Bk, unsigned per_row, float inv_per_row,
                                         F &&f) {
  const unsigned rpb = fast_div(blockDim.x, per_row, inv_per_row);
#pragma unroll 1
  for (unsigned lr0 = 0; lr0 < unsigned(Bk.rows); lr0 += rpb) f(rpb, lr0);
}

// neighbour byte offsets (ES bytes per element) of a point of colour c
// with i mod NC = s in box B
template <int NC, int ES>
__device__ __forceinline__ Nbr<NC> box_offs(ConstBox &B, int c, int s) {
  return stencil_offs<NC, ES>(c, s, unsigned(B.cs), unsigned(B.rs), unsigned(B.ps));
}

// ------------------------------------------------------------------
// smoother: one colour of one Newton-Gauss-Seidel sweep on the rows of
// every listed block
// ------------------------------------------------------------------
// smoother: one colour of one Newton-Gauss-Seidel sweep on the rows of
// every listed block; thread (lr, t) owns colour point t of local row lr
template <int ORD, int EQ>
__device__ __forceinline__ void smooth_rows(const RowBlock &Bk, ConstBox &B, unsigned per_row,
                                            float inv_per_row, int colour) {
  constexpr int R = ORD / 2;
  constexpr int NC = colours_of(ORD, EQ);
  const unsigned lr = fast_div(threadIdx.x, per_row, inv_per_row);
  const int t = int(threadIdx.x - lr * per_row);
  double rmax = 0.0;
  unsigned p;
  int s, i, j, k;
  const int col = B.flip ? NC - 1 - colour : colour;
  if (colour_slot<R, NC>(Bk, B, lr, t, per_row, col, p, s, i, j, k)) {
    const Nbr<NC> nb = box_offs<NC, 8>(B, col, s);
    const unsigned o = p * 8u;
    double *u = B.f[F_U];
    const double rho = rho_at(B, o, i, j, k);
    const double u0 = ld<double>(u, o);
    double N, dN;
    point_op<ORD, EQ>(B, u, o, nb, u0, i, j, k, N, dN);
    const double res = N - rho;
    const double un = u0 - res / dN;
    st<double>(u, o, un);
    rmax = fabs(res);
    if (!(un == un)) rmax = __longlong_as_double(0x7ff0000000000000LL);  // inf flags NaN
  }
  wave_norm(Bk.norm, rmax);
}

// box1: the one box of every listed block, when the launch covers a single
// box (else null); its descriptor is then fetched without waiting for the
// block's.  Either way the descriptor's cache lines are fetched in one batch
// first: the kernel branches on several of its fields, and loads issued
// branch by branch would each wait for the one before.
template <int ORD, int EQ>
__global__ __launch_bounds__(kMaxRowThreads) void k_smooth(const RowBlock *__restrict__ blocks,
                                                        unsigned per_row, float inv_per_row,
                                                        int colour, const DevBox *box1) {
  if (box1) {
    const RowBlock Bk = blocks[xcd_block()];
    fetch_box(box1);
    smooth_rows<ORD, EQ>(Bk, cbox(box1), per_row, inv_per_row, colour);
  } else {
    const RowBlock Bk = blocks[xcd_block()];
    fetch_box(Bk.box);
    smooth_rows<ORD, EQ>(Bk, cbox(Bk.box), per_row, inv_per_row, colour);
  }
}

// ------------------------------------------------------------------
// fused colour passes.  A sweep reads u from field X and writes every
// updated point to field Y (X and Y are u and r, swapped from sweep to
// sweep; the ghost shell is copied into r first, k_copy_shell): within a
// sweep, a point's neighbours of colours already updated are read from Y,
// the others and the point itself from X, which no launch of the sweep
// writes.  One launch does two consecutive colour passes (pass indices Q
// and Q + 1) on a tile of interior rows: the first pass also on the rows
// the second pass reads (one row and plane below the tile, R above), the
// second on the tile alone.  Every launch that computes a first-pass point
// computes it from the same X and Y values, so tiles that overlap write
// identical values, and each workgroup reads back only what it wrote itself
// (after a barrier).  Each point's arithmetic is k_smooth's: the sweeps
// give the same bits as the colour passes one launch at a time.
// ------------------------------------------------------------------

template <int ORD, int NC, int PI, int M>
__device__ __forceinline__ void star_terms(const void *X, const void *Y, unsigned o,
                                           const Nbr<NC> &n, double *sx, double *sy, double *sz) {
  if constexpr (M <= ORD / 2) {
    const void *bp = nbr_in_y(NC, PI, M) ? Y : X;
    const void *bm = nbr_in_y(NC, PI, -M) ? Y : X;
    sx[M - 1] = ld<double>(bp, o + n.xp[M - 1]) + ld<double>(bm, o + n.xm[M - 1]);
    sy[M - 1] = ld<double>(bp, o + n.yp[M - 1]) + ld<double>(bm, o + n.ym[M - 1]);
    sz[M - 1] = ld<double>(bp, o + n.zp[M - 1]) + ld<double>(bm, o + n.zm[M - 1]);
    star_terms<ORD, NC, PI, M + 1>(X, Y, o, n, sx, sy, sz);
  }
}

// one point of pass PI: reads X and Y as above, writes the new u to Y;
// the arithmetic of k_smooth / point_op / lapb, operation for operation
template <int ORD, int EQ, int PI>
__device__ __forceinline__ double fused_point(ConstBox &B, const void *X, void *Y, int c, int i,
                                              int j, int k) {
  constexpr int R = ORD / 2;
  constexpr int NC = R + 1;
  const int s = i - (i / NC) * NC;
  const unsigned p = unsigned(c) * unsigned(B.cs) + unsigned(k) * unsigned(B.ps) +
                     unsigned(j) * unsigned(B.rs) + unsigned(i / NC);
  const unsigned o = p * 8u;
  const Nbr<NC> nb = box_offs<NC, 8>(B, c, s);
  const double u0 = ld<double>(X, o);
  double sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0}, sz[4] = {0, 0, 0, 0};
  star_terms<ORD, NC, PI, 1>(X, Y, o, nb, sx, sy, sz);
  const double lu = star_sum<ORD, double>(star<ORD, double>(u0, sx[0], sx[1], sx[2], sx[3]),
                                          star<ORD, double>(u0, sy[0], sy[1], sy[2], sy[3]),
                                          star<ORD, double>(u0, sz[0], sz[1], sz[2], sz[3]),
                                          B.idx2[0], B.idx2[1], B.idx2[2]);
  double N, dN;
  nonlinear<EQ>(lu, B.dulap, ld<double>(B.f[F_B], o), ld<double>(B.f[F_C], o) + u0,
                EQ == 1 ? ld<double>(B.f[F_D], o) : 0.0, EQ == 1 ? ld<double>(B.f[F_E], o) : 0.0,
                N, dN);
  const double res = N - ld<double>(B.f[F_RHO], o);
  const double un = u0 - res / dN;
  st<double>(Y, o, un);
  double r = fabs(res);
  if (!(un == un)) r = __longlong_as_double(0x7ff0000000000000LL);  // inf flags NaN
  return r;
}

// pass PI over the first nrows rows of the tile's row list, the flattened
// (row, colour point) index strided over the workgroup; returns the max
// |residual| over the tile's own rows
template <int ORD, int EQ, int PI>
__device__ __forceinline__ double fused_pass(const TileBlock &T, ConstBox &B, const void *X, void *Y,
                                             int nrows, int yb, int ya, int zb) {
  constexpr int R = ORD / 2;
  constexpr int NC = R + 1;
  const int P = (B.n[0] - 2 * R + NC - 1) / NC;  // colour points per interior row, at most
  const int nt = T.ty * T.tz;
  const int c = B.flip ? NC - 1 - PI : PI;
  double rmax = 0.0;
  for (int f = int(threadIdx.x); f < nrows * P; f += int(blockDim.x)) {
    const int r = f / P, t = f - r * P;
    int j, k;
    tile_row(T, r, yb, ya, zb, j, k);
    const int i = colour_first(c, j, k, R, NC) + t * NC;
    if (i >= B.n[0] - R) continue;
    const double res = fused_point<ORD, EQ, PI>(B, X, Y, c, i, j, k);
    if (r < nt) rmax = fmax(rmax, res);
  }
  return rmax;
}

// passes Q and Q + 1 (PAIR) or pass Q alone of one sweep on every listed
// tile; parity 0 reads X = u and writes Y = r, parity 1 the reverse
template <int ORD, int EQ, int Q, int PAIR>
__global__ __launch_bounds__(kFusedThreads) void k_smooth_fused(const TileBlock *__restrict__ blocks,
                                                              int parity) {
  constexpr int R = ORD / 2;
  const TileBlock T = blocks[xcd_block()];
  ConstBox &B = cbox(T.box);
  const void *X = parity ? B.f[F_R] : B.f[F_U];
  void *Y = parity ? B.f[F_U] : B.f[F_R];
  const int nt = T.ty * T.tz;
  double rmax;
  if constexpr (PAIR) {
    int yb, ya, zb, za;
    const int nn[3] = {B.n[0], B.n[1], B.n[2]};
    tile_halo(T, nn, R, yb, ya, zb, za);
    const int nrows = nt + (yb + ya) * T.tz + (zb + za) * T.ty;
    rmax = fused_pass<ORD, EQ, Q>(T, B, X, Y, nrows, yb, ya, zb);
    // the second pass reads the first pass's values this workgroup wrote
    __syncthreads();
    rmax = fmax(rmax, fused_pass<ORD, EQ, Q + 1>(T, B, X, Y, nt, 0, 0, 0));
  } else {
    rmax = fused_pass<ORD, EQ, Q>(T, B, X, Y, nt, 0, 0, 0);
  }
  wave_norm(T.norm, rmax);
}

// operator on the interior: mode RES_OPERATOR r = N(u), RES_DEFECT
// w = rho - N(u), RES_DERIVATIVE r = dN/du_p; the norm is max |N(u)|
template <int ORD, int EQ>
__global__ __launch_bounds__(kRowThreads) void k_residual(const RowBlock *__restrict__ blocks,
                                                          unsigned per_row, float inv_per_row,
                                                          int mode) {
  constexpr int R = ORD / 2;
  constexpr int NC = colours_of(ORD, EQ);
  const RowBlock Bk = blocks[xcd_block()];
  ConstBox &B = cbox(Bk.box);
  double rmax = 0.0;
  for_rows(Bk, per_row, inv_per_row, [&](unsigned rpb, unsigned lr0) {
    unsigned p;
    int c, s, i, j, k;
    if (!row_slot<R, NC>(Bk, B, per_row, inv_per_row, rpb, lr0, p, c, s, i, j, k)) return;
    const Nbr<NC> nb = box_offs<NC, 8>(B, c, s);
    const unsigned o = p * 8u;
    const double *u = B.f[F_U];
    const double rho = mode == RES_DEFECT ? rho_at(B, o, i, j, k) : 0.0;
    double N, dN;
    point_op<ORD, EQ>(B, u, o, nb, ld<double>(u, o), i, j, k, N, dN);
    if (mode == RES_DEFECT)
      st<double>(B.f[F_W], o, rho - N);
    else
      st<double>(B.f[F_R], o, mode == RES_DERIVATIVE ? dN : N);
    rmax = nan_max(rmax, fabs(N));
  });
  wave_norm(Bk.norm, rmax);
}

// ------------------------------------------------------------------
// mixed precision.  One smoothing block of nu sweeps becomes: the FP64
// defect r = rho - N(u) and the linearised diagonal q = dN/du - dulap of
// the current u (k_defect32, also the block's norm), nu multi-colour
// Gauss-Seidel sweeps in FP32 on the correction equation
//   Lap e + q e = r,   e = 0 on the ghost shell          (k_smooth32)
// and the FP64 update u += e (k_update32).  At a converged u the defect is
// zero, e stays zero and u is unchanged: the fixed point is the FP64 one.
// ------------------------------------------------------------------
template <int ORD, int EQ>
__global__ __launch_bounds__(kRowThreads) void k_defect32(const RowBlock *__restrict__ blocks,
                                                          unsigned per_row, float inv_per_row) {
  constexpr int R = ORD / 2;
  constexpr int NC = colours_of(ORD, EQ);
  const RowBlock Bk = blocks[xcd_block()];
  ConstBox &B = cbox(Bk.box);
  double rmax = 0.0;
  for_rows(Bk, per_row, inv_per_row, [&](unsigned rpb, unsigned lr0) {
    unsigned p;
    int c, s, i, j, k;
    if (!row_slot<R, NC>(Bk, B, per_row, inv_per_row, rpb, lr0, p, c, s, i, j, k)) return;
    const Nbr<NC> nb = box_offs<NC, 8>(B, c, s);
    const unsigned o = p * 8u, o4 = p * 4u;
    const double *u = B.f[F_U];
    const double rho = rho_at(B, o, i, j, k);
    double N, dN;
    point_op<ORD, EQ>(B, u, o, nb, ld<double>(u, o), i, j, k, N, dN);
    const double res = rho - N;
    st<float>(B.g[G_R], o4, float(res));
    st<float>(B.g[G_Q], o4, float(dN - B.dulap));
    st<float>(B.g[G_E], o4, 0.0f);
    rmax = nan_max(rmax, fabs(res));
    if (!(res == res)) rmax = __longlong_as_double(0x7ff0000000000000LL);
  });
  wave_norm(Bk.norm, rmax);
}

template <int ORD>
__global__ __launch_bounds__(kMaxRowThreads) void k_smooth32(const RowBlock *__restrict__ blocks,
                                                          unsigned per_row, float inv_per_row,
                                                          int colour) {
  constexpr int R = ORD / 2;
  constexpr int NC = R + 1;
  const RowBlock Bk = blocks[xcd_block()];
  ConstBox &B = cbox(Bk.box);
  const unsigned lr = fast_div(threadIdx.x, per_row, inv_per_row);
  const int t = int(threadIdx.x - lr * per_row);
  unsigned p;
  int s, i, j, k;
  const int col = B.flip ? NC - 1 - colour : colour;
  if (!colour_slot<R>(Bk, B, lr, t, per_row, col, p, s, i, j, k)) return;
  const Nbr<NC> nb = box_offs<NC, 4>(B, col, s);
  const unsigned o = p * 4u;
  float *e = B.g[G_E];
  const float le = lapb<ORD, float>(e, o, nb, float(B.idx2[0]), float(B.idx2[1]), float(B.idx2[2]));
  const float q = ld<float>(B.g[G_Q], o);
  const float ep = ld<float>(e, o);
  const float res = fmaf(q, ep, le) - ld<float>(B.g[G_R], o);
  // float32 relaxation of a correction: the hardware reciprocal suffices
  st<float>(e, o, fmaf(-res, __builtin_amdgcn_rcpf(float(B.dulap) + q), ep));
}

template <int ORD>
__global__ __launch_bounds__(kRowThreads) void k_update32(const RowBlock *__restrict__ blocks,
                                                          unsigned per_row, float inv_per_row) {
  constexpr int R = ORD / 2;
  const RowBlock Bk = blocks[xcd_block()];
  ConstBox &B = cbox(Bk.box);
  for_rows(Bk, per_row, inv_per_row, [&](unsigned rpb, unsigned lr0) {
    unsigned p;
    int c, s, i, j, k;
    if (!row_slot<R>(Bk, B, per_row, inv_per_row, rpb, lr0, p, c, s, i, j, k)) return;
    double *u = B.f[F_U];
    st<double>(u, p * 8u, ld<double>(u, p * 8u) + double(ld<float>(B.g[G_E], p * 4u)));
  });
}

// ------------------------------------------------------------------
// Robin boundary r^2/n sum_d (1/x_d) d_d u + u = S on the ghost shell
// ------------------------------------------------------------------

// value of u at offset m (any sign, |m| <= 4) along axis a from the point
// at storage index p with colour c and i mod nc = s
__device__ __forceinline__ double axis_at(const DevBox &B, const double *__restrict__ u, long long p,
                                          int c, int s, int a, int m) {
  return u[p + axis_offset(B, c, s, a, m)];
}

// first derivative along axis a at (i, j, k), storage index p:
// dx2 (diff_2nd.cc:237-335) for order 2, dx4 (diff_4th.cc:6-89) otherwise.
// coef = coefficient of u at the point itself.
__device__ __forceinline__ double deriv(const DevBox &B, const double *__restrict__ u, long long p,
                                        int a, int i, int j, int k, double &coef) {
  const int n = B.n[a];
  const int ia = a == 0 ? i : (a == 1 ? j : k);
  const double id = B.idx[a];
  const int c = colour_of(B, i, j, k), s = i - split_div(B, i) * B.nc;
#define UA(m) axis_at(B, u, p, c, s, a, (m))
  if (B.order == 2) {
    if (ia == 0) {
      coef = -3.0 * 0.5 * id;
      return (-3.0 * u[p] + 4.0 * UA(1) - UA(2)) * 0.5 * id;
    } else if (ia == n - 1) {
      coef = 3.0 * 0.5 * id;
      return (3.0 * u[p] - 4.0 * UA(-1) + UA(-2)) * 0.5 * id;
    }
    coef = 0.0;
    return (UA(1) - UA(-1)) * 0.5 * id;
  }
  const int nm = (n % 2 == 0) ? n / 2 : (n - 1) / 2;
  if (ia <= nm) {
    coef = -25.0 * id / 12.0;
    return -(25.0 * u[p] - 48.0 * UA(1) + 36.0 * UA(2) - 16.0 * UA(3) + 3.0 * UA(4)) * id / 12.0;
  }
  coef = 25.0 * id / 12.0;
  return (25.0 * u[p] - 48.0 * UA(-1) + 36.0 * UA(-2) - 16.0 * UA(-3) + 3.0 * UA(-4)) * id / 12.0;
#undef UA
}

// box-local views of the fields a Robin stage touches; u may be in LDS
struct BoxData {
  double *u;
  const double *rho;
  double *r;
  double *w;
};

__device__ __forceinline__ BoxData box_data(const DevBox &B) {
  BoxData D;
  D.u = B.f[F_U];
  D.rho = B.f[F_RHO];
  D.r = B.f[F_R];
  D.w = B.f[F_W];
  return D;
}

// Robin update / operator at one shell point; `mask` has bit a set for
// each axis a whose index lies in the ghost shell.
__device__ __forceinline__ void robin_point(const DevBox &B, BoxData D, int i, int j, int k,
                                            int mask, int mode) {
  const long long p = pidx(B, i, j, k);
  double *u = D.u;
  const int ijk[3] = {i, j, k};
  double X[3];
  for (int a = 0; a < 3; ++a) X[a] = coord(B, a, ijk[a]);
  const double r2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
  double sF = 0.0, sdF = 0.0;
  for (int a = 0; a < 3; ++a) {
    if (!(mask & (1 << a))) continue;
    double cf;
    const double dv = deriv(B, u, p, a, i, j, k, cf);
    sF += dv / X[a];
    sdF += cf / X[a];
  }
  const double op = r2 * B.ipn * sF + u[p];
  if (mode == ROBIN_UPDATE) {
    const double Fv = op - D.rho[p];
    const double dF = r2 * B.ipn * sdF + 1.0;
    u[p] = u[p] - Fv / dF;
  } else if (mode == ROBIN_OPERATOR_R) {
    D.r[p] = op;
  } else {
    D.w[p] = D.rho[p] - op;
  }
}

// face column t of a box: all R layers, inner -> outer
__device__ __forceinline__ void robin_face_column(const DevBox &B, BoxData D, int t, int mode) {
  const int R = B.R;
  const int ni = B.n[0] - 2 * R, nj = B.n[1] - 2 * R, nk = B.n[2] - 2 * R;
  const int fx = nj * nk, fy = ni * nk, fz = ni * nj;
  int face, a0, a1;
  if (t < 2 * fx) {
    face = t / fx; t %= fx; a0 = t % nj; a1 = t / nj;  // j, k
  } else if ((t -= 2 * fx) < 2 * fy) {
    face = 2 + t / fy; t %= fy; a0 = t % ni; a1 = t / ni;  // i, k
  } else if ((t -= 2 * fy) < 2 * fz) {
    face = 4 + t / fz; t %= fz; a0 = t % ni; a1 = t / ni;  // i, j
  } else {
    return;
  }
  const int axis = face >> 1, hi = face & 1;
  for (int l = 0; l < R; ++l) {
    const int g = hi ? B.n[axis] - R + l : R - 1 - l;
    int i, j, k;
    if (axis == 0) {
      i = g; j = R + a0; k = R + a1;
    } else if (axis == 1) {
      i = R + a0; j = g; k = R + a1;
    } else {
      i = R + a0; j = R + a1; k = g;
    }
    robin_point(B, D, i, j, k, 1 << axis, mode);
  }
}

// ghost index of position q (0..R-1) on side `hi` of an axis, in the
// reference's iteration order (domain.cc:575-1388): low sides run R-1 -> 0,
// high sides n-R -> n-1
__device__ __forceinline__ int shell_index(int q, int hi, int n, int R) {
  return hi ? n - R + q : R - 1 - q;
}

// The corners' Robin updates form one chain per corner (each point reads
// the ones updated before it), so they run from LDS: the whole block stages
// each corner's cube of kCornerW^3 points (R <= 4 shell layers plus the 4
// inward points the one-sided derivatives read) and rho of its R^3 points,
// eight threads walk the chains, and the updated points are written back.
constexpr int kCornerR = 4;
constexpr int kCornerW = kCornerR + 4;
constexpr int kCornerCube = kCornerW * kCornerW * kCornerW;
constexpr int kCornerR3 = kCornerR * kCornerR * kCornerR;

// box index of local coordinate l (counted inwards from side h) along an axis
__device__ __forceinline__ int corner_index(int l, int h, int n) { return h ? n - 1 - l : l; }

// one Robin update of corner point (li, lj, lk) (local, counted inwards)
// of the cube cu of corner (hx, hy, hz): robin_point's arithmetic, with the
// derivatives' inward neighbours read from the cube
__device__ __forceinline__ void robin_corner_update(const DevBox &B, double *cu, double rho, int hx,
                                                    int hy, int hz, int li, int lj, int lk) {
  const int l[3] = {li, lj, lk}, h[3] = {hx, hy, hz};
  const int st[3] = {1, kCornerW, kCornerW * kCornerW};
  const int q = (lk * kCornerW + lj) * kCornerW + li;
  double X[3];
  for (int a = 0; a < 3; ++a) X[a] = coord(B, a, corner_index(l[a], h[a], B.n[a]));
  const double r2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
  double sF = 0.0, sdF = 0.0;
  const double u0 = cu[q];
  for (int a = 0; a < 3; ++a) {
    const double id = B.idx[a];
    const double *w = cu + q;
    const int m = st[a];
    double cf, dv;
    // deriv(): dx2 at the box faces for order 2, the one-sided dx4 otherwise
    if (B.order == 2) {
      if (!h[a]) {
        cf = -3.0 * 0.5 * id;
        dv = (-3.0 * u0 + 4.0 * w[m] - w[2 * m]) * 0.5 * id;
      } else {
        cf = 3.0 * 0.5 * id;
        dv = (3.0 * u0 - 4.0 * w[m] + w[2 * m]) * 0.5 * id;
      }
    } else if (!h[a]) {
      cf = -25.0 * id / 12.0;
      dv = -(25.0 * u0 - 48.0 * w[m] + 36.0 * w[2 * m] - 16.0 * w[3 * m] + 3.0 * w[4 * m]) * id / 12.0;
    } else {
      cf = 25.0 * id / 12.0;
      dv = (25.0 * u0 - 48.0 * w[m] + 36.0 * w[2 * m] - 16.0 * w[3 * m] + 3.0 * w[4 * m]) * id / 12.0;
    }
    sF += dv / X[a];
    sdF += cf / X[a];
  }
  const double op = r2 * B.ipn * sF + u0;
  const double Fv = op - rho;
  const double dF = r2 * B.ipn * sdF + 1.0;
  cu[q] = u0 - Fv / dF;
}

// One edge cross-section's Robin updates in the reference's order (outer
// axis A, inner axis B, both counted inwards from the shell's sides), from
// registers: the L-shaped set of points its one-sided derivatives read (the
// R x R shell points and the 4 inward points along A and along B of each)
// is loaded first, the R^2 updates chain through the registers with
// robin_point's arithmetic, and the updated points are written back.
// axa, axb: axes A and B; ha, hb: their sides; e_idx: the index along the
// edge's own axis (the third).
template <int R>
__device__ void robin_edge_chain(const DevBox &B, BoxData D, int axa, int axb, int ha, int hb,
                                 int e_idx) {
  constexpr int W = R + 4;
  double va[R][W];  // local A < R: every local B
  double vb[4][R];  // local A >= R: local B < R
  double rho[R][R];
  // box indices of local point (la, lb)
  auto index = [&](int la, int lb, int &i, int &j, int &k) {
    const int pa = ha ? B.n[axa] - 1 - la : la, pb = hb ? B.n[axb] - 1 - lb : lb;
    i = axa == 0 ? pa : (axb == 0 ? pb : e_idx);
    j = axa == 1 ? pa : (axb == 1 ? pb : e_idx);
    k = axa == 2 ? pa : (axb == 2 ? pb : e_idx);
  };
#pragma unroll
  for (int la = 0; la < W; ++la)
#pragma unroll
    for (int lb = 0; lb < W; ++lb) {
      if (la >= R && lb >= R) continue;
      int i, j, k;
      index(la, lb, i, j, k);
      const long long p = pidx(B, i, j, k);
      if (la < R) va[la][lb] = D.u[p]; else vb[la - R][lb] = D.u[p];
      if (la < R && lb < R) rho[la][lb] = D.rho[p];
    }
  // robin_point takes the masked axes in ascending order
  const bool a_first = axa < axb;
#pragma unroll
  for (int qa = 0; qa < R; ++qa)
#pragma unroll
    for (int qb = 0; qb < R; ++qb) {
      const int la = R - 1 - qa, lb = R - 1 - qb;
      int i, j, k;
      index(la, lb, i, j, k);
      const double X[3] = {coord(B, 0, i), coord(B, 1, j), coord(B, 2, k)};
      const double r2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
      double sF = 0.0, sdF = 0.0;
      const double u0 = va[la][lb];
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        const bool along_a = a_first == (o == 0);
        // the 4 inward points along the axis (la + m < W, lb + m < W)
        double w[5];
        w[0] = u0;
#pragma unroll
        for (int m = 1; m <= 4; ++m)
          w[m] = along_a ? (la + m < R ? va[la + m][lb] : vb[la + m - R][lb]) : va[la][lb + m];
        const int a = along_a ? axa : axb;
        const int hs = along_a ? ha : hb;
        const double id = B.idx[a];
        const double xa = a == 0 ? X[0] : (a == 1 ? X[1] : X[2]);
        double cf, dv;
        if (B.order == 2) {
          if (!hs) {
            cf = -3.0 * 0.5 * id;
            dv = (-3.0 * u0 + 4.0 * w[1] - w[2]) * 0.5 * id;
          } else {
            cf = 3.0 * 0.5 * id;
            dv = (3.0 * u0 - 4.0 * w[1] + w[2]) * 0.5 * id;
          }
        } else if (!hs) {
          cf = -25.0 * id / 12.0;
          dv = -(25.0 * u0 - 48.0 * w[1] + 36.0 * w[2] - 16.0 * w[3] + 3.0 * w[4]) * id / 12.0;
        } else {
          cf = 25.0 * id / 12.0;
          dv = (25.0 * u0 - 48.0 * w[1] + 36.0 * w[2] - 16.0 * w[3] + 3.0 * w[4]) * id / 12.0;
        }
        sF += dv / xa;
        sdF += cf / xa;
      }
      const double op = r2 * B.ipn * sF + u0;
      const double Fv = op - rho[la][lb];
      const double dF = r2 * B.ipn * sdF + 1.0;
      va[la][lb] = u0 - Fv / dF;
    }
#pragma unroll
  for (int la = 0; la < R; ++la)
#pragma unroll
    for (int lb = 0; lb < R; ++lb) {
      int i, j, k;
      index(la, lb, i, j, k);
      D.u[pidx(B, i, j, k)] = va[la][lb];
    }
}

// Edge cross-section t of a box (faces must be done): its R x R shell
// points in the reference's order.  Edges XY(4) run along z, XZ(4) along
// y, YZ(4) along x; within an edge the cross-section is visited outer axis
// first (x before y for XY, x before z for XZ, z before y for YZ), as
// Fix_Points orders them.
__device__ void robin_edge_section(const DevBox &B, BoxData D, int mode, int t) {
  const int R = B.R;
  const int nint[3] = {B.n[0] - 2 * R, B.n[1] - 2 * R, B.n[2] - 2 * R};
  int grp, e, s;
  if (t < 4 * nint[2]) {
    grp = 0; e = t / nint[2]; s = t % nint[2];
  } else if (t < 4 * (nint[2] + nint[1])) {
    grp = 1; e = (t - 4 * nint[2]) / nint[1]; s = (t - 4 * nint[2]) % nint[1];
  } else {
    grp = 2;
    e = (t - 4 * (nint[2] + nint[1])) / nint[0];
    s = (t - 4 * (nint[2] + nint[1])) % nint[0];
  }
  const int hA = (e >> 1) & 1, hB = e & 1;
  if (mode == ROBIN_UPDATE && R <= 3) {
    // outer axis A and inner axis B of the cross-section and their sides
    // (the orders below); 8th-order boxes (R = 4) would not fit the
    // registers and take the point-by-point path
    const int axa = grp == 2 ? 2 : 0, axb = grp == 1 ? 2 : 1;
    const int ha = grp == 2 ? hB : hA, hb = grp == 2 ? hA : hB;
    if (R == 1)
      robin_edge_chain<1>(B, D, axa, axb, ha, hb, R + s);
    else if (R == 2)
      robin_edge_chain<2>(B, D, axa, axb, ha, hb, R + s);
    else
      robin_edge_chain<3>(B, D, axa, axb, ha, hb, R + s);
    return;
  }
  for (int qa = 0; qa < R; ++qa)
    for (int qb = 0; qb < R; ++qb) {
      int i, j, k, mask;
      if (grp == 0) {  // XiYi, XiYf, XfYi, XfYf: x outer, y inner; along z
        i = shell_index(qa, hA, B.n[0], R); j = shell_index(qb, hB, B.n[1], R); k = R + s; mask = 3;
      } else if (grp == 1) {  // XiZi, XiZf, XfZi, XfZf: x outer, z inner; along y
        i = shell_index(qa, hA, B.n[0], R); k = shell_index(qb, hB, B.n[2], R); j = R + s; mask = 5;
      } else {  // YiZi, YiZf, YfZi, YfZf: z outer, y inner; along x
        j = shell_index(qb, hA, B.n[1], R); k = shell_index(qa, hB, B.n[2], R); i = R + s; mask = 6;
      }
      robin_point(B, D, i, j, k, mask, mode);
    }
}

// corners of one box, by one block (edges must be done): k outer, j, i
// inner (domain.cc:592-695); corner c has hz = bit 0, hx = bit 1, hy = bit
// 2.  cube and crho: 8 kCornerCube and 8 kCornerR3 doubles of LDS, for the
// corner chains (R <= kCornerR)
__device__ void robin_corners(const DevBox &B, BoxData D, int mode, double *cube, double *crho) {
  const int R = B.R;
  if (mode != ROBIN_UPDATE) {
    // the operator modes leave u alone: every corner point at once
    const int R3 = R * R * R;
    for (int t = threadIdx.x; t < 8 * R3; t += blockDim.x) {
      const int c = t / R3, q = t - c * R3;
      const int hz = c & 1, hx = (c >> 1) & 1, hy = (c >> 2) & 1;
      const int qk = q / (R * R), qj = (q / R) % R, qi = q % R;
      robin_point(B, D, shell_index(qi, hx, B.n[0], R), shell_index(qj, hy, B.n[1], R),
                  shell_index(qk, hz, B.n[2], R), 7, mode);
    }
    return;
  }
  // stage the eight cubes (points beyond a small box are never read) and
  // the corner points' rho
  for (int t = threadIdx.x; t < 8 * kCornerCube; t += blockDim.x) {
    const int c = t / kCornerCube, q = t - c * kCornerCube;
    const int hz = c & 1, hx = (c >> 1) & 1, hy = (c >> 2) & 1;
    const int li = q % kCornerW, lj = (q / kCornerW) % kCornerW, lk = q / (kCornerW * kCornerW);
    double v = 0.0;
    if (li < B.n[0] && lj < B.n[1] && lk < B.n[2]) {
      const long long p = pidx(B, corner_index(li, hx, B.n[0]), corner_index(lj, hy, B.n[1]),
                               corner_index(lk, hz, B.n[2]));
      v = D.u[p];
      if (li < R && lj < R && lk < R) crho[c * kCornerR3 + (lk * kCornerR + lj) * kCornerR + li] = D.rho[p];
    }
    cube[t] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int c = threadIdx.x;
    const int hz = c & 1, hx = (c >> 1) & 1, hy = (c >> 2) & 1;
    double *cu = cube + c * kCornerCube;
    const double *cr = crho + c * kCornerR3;
    // shell_index(q, h, n, R) is local coordinate R - 1 - q from either side
    for (int qk = 0; qk < R; ++qk)
      for (int qj = 0; qj < R; ++qj)
        for (int qi = 0; qi < R; ++qi) {
          const int li = R - 1 - qi, lj = R - 1 - qj, lk = R - 1 - qk;
          robin_corner_update(B, cu, cr[(lk * kCornerR + lj) * kCornerR + li], hx, hy, hz, li, lj, lk);
        }
  }
  __syncthreads();
  const int R3 = R * R * R;
  for (int t = threadIdx.x; t < 8 * R3; t += blockDim.x) {
    const int c = t / R3, q = t - c * R3;
    const int hz = c & 1, hx = (c >> 1) & 1, hy = (c >> 2) & 1;
    const int lk = q / (R * R), lj = (q / R) % R, li = q % R;
    D.u[pidx(B, corner_index(li, hx, B.n[0]), corner_index(lj, hy, B.n[1]), corner_index(lk, hz, B.n[2]))] =
        cube[c * kCornerCube + (lk * kCornerW + lj) * kCornerW + li];
  }
}

// faces: one thread per (box, face column)
__global__ __launch_bounds__(256) void k_robin_faces(const RobinItem *__restrict__ items, int n,
                                                      int total, int mode) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].col0 <= g) lo = mid; else hi = mid - 1;
  }
  const DevBox &B = *items[lo].box;
  robin_face_column(B, box_data(B), g - items[lo].col0, mode);
}

// edges: one thread per (box, edge cross-section), spread over many
// workgroups (a box's edges are a chain of R^2 dependent updates each)
constexpr int kRobinEdgeThreads = 128;
__global__ __launch_bounds__(kRobinEdgeThreads) void k_robin_edges(const RobinItem *__restrict__ items,
                                                                   int n, int total, int mode) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].ecol0 <= g) lo = mid; else hi = mid - 1;
  }
  const DevBox &B = *items[lo].box;
  robin_edge_section(B, box_data(B), mode, g - items[lo].ecol0);
}

// corners: one block per box
__global__ __launch_bounds__(512) void k_robin_corners(const RobinItem *__restrict__ items,
                                                       int mode) {
  __shared__ double cube[8 * kCornerCube], crho[8 * kCornerR3];
  const DevBox &B = *items[blockIdx.x].box;
  robin_corners(B, box_data(B), mode, cube, crho);
}

// Robin operator value at a shell point (no update)
__device__ __forceinline__ double robin_op_at(const DevBox &B, int i, int j, int k) {
  const int R = B.R;
  int mask = 0;
  if (i < R || i >= B.n[0] - R) mask |= 1;
  if (j < R || j >= B.n[1] - R) mask |= 2;
  if (k < R || k >= B.n[2] - R) mask |= 4;
  const long long p = pidx(B, i, j, k);
  const double *__restrict__ u = B.f[F_U];
  const int ijk[3] = {i, j, k};
  double X[3];
  for (int a = 0; a < 3; ++a) X[a] = coord(B, a, ijk[a]);
  const double r2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
  double sF = 0.0;
  for (int a = 0; a < 3; ++a) {
    if (!(mask & (1 << a))) continue;
    double cf;
    sF += deriv(B, u, p, a, i, j, k, cf) / X[a];
  }
  return r2 * B.ipn * sF + u[p];
}

// ------------------------------------------------------------------
// transfers.  Descriptors (target box, transfer, source box) are read
// through the constant address space; fields through 32-bit byte offsets.
// ------------------------------------------------------------------
typedef const __attribute__((address_space(4))) DevXfer ConstXfer;
typedef const __attribute__((address_space(4))) PtBlock ConstPtBlock;

// byte offset of point (i, j, k) of a float64 field
__device__ __forceinline__ unsigned boff(ConstBox &S, int i, int j, int k) {
  return (colourc(S, i, j, k) * unsigned(S.cs) + unsigned(k) * unsigned(S.ps) +
          unsigned(j) * unsigned(S.rs) + divc(S, unsigned(i))) *
         8u;
}

// Walks a float64 field of box S one step at a time from a boff: every step
// of +1 along an axis moves to the next colour block (wrapping after the
// last), plus one row (y), one plane (z), or, along x, one slot when the
// point's x leaves the last residue mod nc.
struct Walk {
  unsigned nc, cs8, wrap8, rs8, ps8;
  __device__ __forceinline__ Walk(ConstBox &S)
      : nc(unsigned(S.nc)), cs8(unsigned(S.cs) * 8u), wrap8((unsigned(S.nc) - 1u) * unsigned(S.cs) * 8u),
        rs8(unsigned(S.rs) * 8u), ps8(unsigned(S.ps) * 8u) {}
  // byte delta of the colour step from colour c; advances c
  __device__ __forceinline__ unsigned colour(unsigned &c) const {
    const bool last = c + 1u == nc;
    c = last ? 0u : c + 1u;
    return last ? 0u - wrap8 : cs8;
  }
  // x of residue r (mod nc) -> x + 1; advances c and r
  __device__ __forceinline__ unsigned x(unsigned &c, unsigned &r) const {
    const bool slot = r + 1u == nc;
    r = slot ? 0u : r + 1u;
    return colour(c) + (slot ? 8u : 0u);
  }
  __device__ __forceinline__ unsigned y(unsigned &c) const { return colour(c) + rs8; }
  __device__ __forceinline__ unsigned z(unsigned &c) const { return colour(c) + ps8; }
};

// point q0 + off of a block in lexicographic order, stepped from the
// block's first point (i0, j0, k0) so that only offsets below 2^24 are
// divided; or point off of a block's patch of rows
__device__ __forceinline__ void pt_lex(const PtBlock &Bk, ConstBox &T, unsigned off, int &i, int &j,
                                       int &k) {
  if (Bk.w > 0) {  // a patch of rows
    const unsigned r = fast_div(off, unsigned(Bk.w), Bk.inw);
    i = Bk.i0 + int(off - r * unsigned(Bk.w));
    j = Bk.j0 + int(r);
    k = Bk.k0;
    return;
  }
  const unsigned n0 = unsigned(T.n[0]), n1 = unsigned(T.n[1]);
  const unsigned x = unsigned(Bk.i0) + off;
  const unsigned dj = fast_div(x, n0, T.in0);
  unsigned jj = unsigned(Bk.j0) + dj, kk = unsigned(Bk.k0);
  if (jj >= n1) {
    const unsigned dk = fast_div(jj, n1, 1.0f / float(n1));
    jj -= dk * n1;
    kk += dk;
  }
  i = int(x - dj * n0);
  j = int(jj);
  k = int(kk);
}

// f(off) for the offsets off < nq of a block's points that fall to this
// thread (kPtThreads threads step through blocks of any length)
template <class F>
__device__ __forceinline__ void for_points(const PtBlock &Bk, F &&f) {
#pragma unroll 1
  for (unsigned off = threadIdx.x; off < Bk.nq; off += kPtThreads) f(off);
}

__device__ __forceinline__ bool is_interior(ConstBox &B, int i, int j, int k) {
  const int R = B.R;
  return i >= R && i < B.n[0] - R && j >= R && j < B.n[1] - R && k >= R && k < B.n[2] - R;
}

__device__ __forceinline__ unsigned axis_flags(ConstXfer &X, const int t[3], unsigned fl[3]) {
  fl[0] = X.ax[0].flags[t[0]];
  fl[1] = X.ax[1].flags[t[1]];
  fl[2] = X.ax[2].flags[t[2]];
  return fl[0] & fl[1] & fl[2];
}

// tensor Lagrange interpolation of field f of box S with KP nodes per axis
// (reference accumulation order, interpolation.cc:257-296); the nodes of
// coincident axes are skipped, their weights being exactly one unit vector
template <int KP>
__device__ __forceinline__ double tinterp(ConstBox &S, const void *f, ConstXfer &X, const int t[3],
                                          const unsigned fl[3]) {
  int s0[3], kk[3];
  const double *w[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (fl[a] & 4) {
      s0[a] = X.ax[a].coinc[t[a]];
      kk[a] = 1;
      w[a] = nullptr;
    } else {
      s0[a] = X.ax[a].start[t[a]];
      kk[a] = KP;
      w[a] = X.ax[a].w + t[a] * KP;
    }
  }
  double wz[KP], wy[KP];
#pragma unroll
  for (int K = 0; K < KP; ++K) {
    wz[K] = w[2] ? w[2][K] : 1.0;
    wy[K] = w[1] ? w[1][K] : 1.0;
  }
  const Walk wk(S);
  unsigned oI = boff(S, s0[0], s0[1], s0[2]);
  unsigned cI = colourc(S, s0[0], s0[1], s0[2]);
  unsigned rI = unsigned(s0[0]) - divc(S, unsigned(s0[0])) * wk.nc;
  double acc = 0.0;
  for (int I = 0; I < kk[0]; ++I) {
    double uj = 0.0;
    unsigned oJ = oI, cJ = cI;
    if constexpr (KP <= 6) {
      // the (J, K) plane of nodes: every load of it issued before the sums
      double v[KP][KP];
#pragma unroll
      for (int J = 0; J < KP; ++J) {
        unsigned off = oJ, c = cJ;
#pragma unroll
        for (int K = 0; K < KP; ++K) {
          v[J][K] = (J < kk[1] && K < kk[2]) ? ld<double>(f, off) : 0.0;
          off += wk.z(c);
        }
        oJ += wk.y(cJ);
      }
#pragma unroll
      for (int J = 0; J < KP; ++J)
        if (J < kk[1]) {
          double uk = 0.0;
#pragma unroll
          for (int K = 0; K < KP; ++K)
            if (K < kk[2]) uk += v[J][K] * wz[K];
          uj += uk * wy[J];
        }
    } else {
      // wider stencils: CJ rows of nodes at a time, every load of them
      // issued before their sums (register budget); the sums keep the
      // order of the row-by-row form
      constexpr int CJ = KP / 2;
      for (int J0 = 0; J0 < kk[1]; J0 += CJ) {
        double v[CJ][KP];
#pragma unroll
        for (int J = 0; J < CJ; ++J) {
          unsigned off = oJ, c = cJ;
#pragma unroll
          for (int K = 0; K < KP; ++K) {
            v[J][K] = (J0 + J < kk[1] && K < kk[2]) ? ld<double>(f, off) : 0.0;
            off += wk.z(c);
          }
          if (J0 + J < kk[1]) oJ += wk.y(cJ);
        }
#pragma unroll
        for (int J = 0; J < CJ; ++J)
          if (J0 + J < kk[1]) {
            double uk = 0.0;
#pragma unroll
            for (int K = 0; K < KP; ++K)
              if (K < kk[2]) uk += v[J][K] * wz[K];
            uj += uk * wy[J0 + J];
          }
      }
    }
    acc += uj * (w[0] ? w[0][I] : 1.0);
    oI += wk.x(cI, rI);
  }
  return acc;
}

// trilinear interpolation of the coarse correction v = u - ww of box S
// (interpolation.cc:125-182; k_corr_v forms v)
__device__ __forceinline__ double trilinear_corr(ConstBox &S, ConstXfer &X, const int t[3]) {
  const void *v = S.f[F_V];
  const int x0 = X.ax[0].start[t[0]], y0 = X.ax[1].start[t[1]], z0 = X.ax[2].start[t[2]];
  const double *wx = X.ax[0].w + 2 * t[0];
  const double *wy = X.ax[1].w + 2 * t[1];
  const double *wzp = X.ax[2].w + 2 * t[2];
  const double dxi = wx[0], dxf = wx[1], dyi = wy[0], dyf = wy[1], dzi = wzp[0], dzf = wzp[1];
  // the eight corners, walked from (x0, y0, z0)
  const Walk wk(S);
  const unsigned c0 = colourc(S, x0, y0, z0);
  unsigned r = unsigned(x0) - divc(S, unsigned(x0)) * wk.nc;
  const unsigned o000 = boff(S, x0, y0, z0);
  unsigned c = c0;
  const unsigned o100 = o000 + wk.x(c, r);  // c = c0 + 1
  unsigned c1 = c;
  const unsigned o110 = o100 + wk.y(c1);     // c1 = c0 + 2
  unsigned c2 = c1;
  const unsigned o111 = o110 + wk.z(c2);
  unsigned cb = c;
  const unsigned o101 = o100 + wk.z(cb);
  unsigned cy = c0;
  const unsigned o010 = o000 + wk.y(cy);    // cy = c0 + 1
  unsigned cz = cy;
  const unsigned o011 = o010 + wk.z(cz);
  unsigned cw = c0;
  const unsigned o001 = o000 + wk.z(cw);
#define V(o) ld<double>(v, o)
  return ((V(o000) * dxi + V(o100) * dxf) * dyi + (V(o010) * dxi + V(o110) * dxf) * dyf) * dzi +
         ((V(o001) * dxi + V(o101) * dxf) * dyi + (V(o011) * dxi + V(o111) * dxf) * dyf) * dzf;
#undef V
}

// the same for a 2:1 nested source (DevXfer::affine): corner nodes and
// weights (1, 0 at a coincident axis; 1/2, 1/2 at a midpoint) from the index
// map instead of the tables, the eight corners walked as above
__device__ __forceinline__ double trilinear_nested(ConstBox &S, ConstXfer &X, const int t[3]) {
  const void *v = S.f[F_V];
  int c[3];
  double wi[3], wf[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int q = t[a] + X.aff[a];
    c[a] = q >> 1;
    const bool mid = (q & 1) != 0;
    wi[a] = mid ? 0.5 : 1.0;
    wf[a] = mid ? 0.5 : 0.0;
    // a coincident axis reads node c twice (weight 0 on the second): keep
    // the second node in range at the source's last node
    if (!mid && c[a] + 1 >= S.n[a]) c[a] = S.n[a] - 2, wi[a] = 0.0, wf[a] = 1.0;
  }
  const Walk wk(S);
  const unsigned c0 = colourc(S, c[0], c[1], c[2]);
  unsigned r = unsigned(c[0]) - divc(S, unsigned(c[0])) * wk.nc;
  const unsigned o000 = boff(S, c[0], c[1], c[2]);
  unsigned cc = c0;
  const unsigned o100 = o000 + wk.x(cc, r);
  unsigned c1 = cc;
  const unsigned o110 = o100 + wk.y(c1);
  unsigned c2 = c1;
  const unsigned o111 = o110 + wk.z(c2);
  unsigned cb = cc;
  const unsigned o101 = o100 + wk.z(cb);
  unsigned cy = c0;
  const unsigned o010 = o000 + wk.y(cy);
  unsigned cz = cy;
  const unsigned o011 = o010 + wk.z(cz);
  unsigned cw = c0;
  const unsigned o001 = o000 + wk.z(cw);
#define V(o) ld<double>(v, o)
  return ((V(o000) * wi[0] + V(o100) * wf[0]) * wi[1] + (V(o010) * wi[0] + V(o110) * wf[0]) * wf[1]) * wi[2] +
         ((V(o001) * wi[0] + V(o101) * wf[0]) * wi[1] + (V(o011) * wi[0] + V(o111) * wf[0]) * wf[1]) * wf[2];
#undef V
}

// trilinear interpolation of field f of box S (interpolation.cc:125-182),
// the corners walked as in trilinear_corr
__device__ __forceinline__ double trilinear_field(ConstBox &S, const void *f, ConstXfer &X,
                                                  const int t[3]) {
  const int x0 = X.ax[0].start[t[0]], y0 = X.ax[1].start[t[1]], z0 = X.ax[2].start[t[2]];
  const double *wx = X.ax[0].w + 2 * t[0];
  const double *wy = X.ax[1].w + 2 * t[1];
  const double *wzp = X.ax[2].w + 2 * t[2];
  const double dxi = wx[0], dxf = wx[1], dyi = wy[0], dyf = wy[1], dzi = wzp[0], dzf = wzp[1];
  const Walk wk(S);
  const unsigned c0 = colourc(S, x0, y0, z0);
  unsigned r = unsigned(x0) - divc(S, unsigned(x0)) * wk.nc;
  const unsigned o000 = boff(S, x0, y0, z0);
  unsigned c = c0;
  const unsigned o100 = o000 + wk.x(c, r);
  unsigned c1 = c;
  const unsigned o110 = o100 + wk.y(c1);
  unsigned c2 = c1;
  const unsigned o111 = o110 + wk.z(c2);
  unsigned cb = c;
  const unsigned o101 = o100 + wk.z(cb);
  unsigned cy = c0;
  const unsigned o010 = o000 + wk.y(cy);
  unsigned cz = cy;
  const unsigned o011 = o010 + wk.z(cz);
  unsigned cw = c0;
  const unsigned o001 = o000 + wk.z(cw);
#define V(o) ld<double>(f, o)
  return ((V(o000) * dxi + V(o100) * dxf) * dyi + (V(o010) * dxi + V(o110) * dxf) * dyf) * dzi +
         ((V(o001) * dxi + V(o101) * dxf) * dyi + (V(o011) * dxi + V(o111) * dxf) * dyf) * dzf;
#undef V
}

// multigrid::Transferir (multigrid.cc:122-141, ffunction.cc:2436-2475):
// every point of the region (PtBlock::mg bits 0-1: 0 all, 1 interior, 2
// ghost shell) of the listed target boxes takes field ff of the source
// boxes interpolated at its coordinates, linear (KP = 2) or Lagrange with
// KP nodes per axis: the last source whose non-ghost part holds the point
// (or, with `all` (bit 2), whose extent does); with `to_zero` (bit 3) the
// last source alone decides and a point it does not hold is set to 0 (the
// reference's source loop overwrites every point on each pass)
template <int KP>
__global__ __launch_bounds__(kPtThreads) void k_transfer_fields(const PtBlock *__restrict__ blocks,
                                                                int ff, int tf) {
  const PtBlock Bk = blocks[xcd_block()];
  ConstBox &T = cbox(Bk.box);
  for_points(Bk, [&](unsigned off) {
    int i, j, k;
    pt_lex(Bk, T, off, i, j, k);
    const int region = Bk.mg & 3, all = (Bk.mg >> 2) & 1, to_zero = (Bk.mg >> 3) & 1;
    const bool inside = is_interior(T, i, j, k);
    if ((region == 1 && !inside) || (region == 2 && inside)) return;
    const int t[3] = {i, j, k};
    const unsigned o = boff(T, i, j, k);
    ConstXfer *xs = (ConstXfer *)Bk.src;
    for (int x = Bk.nsrc - 1; x >= 0; --x) {
      ConstXfer &X = xs[x];
      unsigned fl[3];
      const unsigned f = axis_flags(X, t, fl);
      if ((f & 1) || (all && (f & 2))) {
        ConstBox &S = cbox(X.src);
        double v;
        if constexpr (KP == 2) {
          if (f & 4)  // a source node: its value (interpolation.cc:140-147)
            v = ld<double>(S.f[ff], boff(S, X.ax[0].coinc[i], X.ax[1].coinc[j], X.ax[2].coinc[k]));
          else
            v = trilinear_field(S, S.f[ff], X, t);
        } else {
          v = tinterp<KP>(S, S.f[ff], X, t, fl);
        }
        st<double>(T.f[tf], o, v);
        return;
      }
      if (to_zero) {
        st<double>(T.f[tf], o, 0.0);
        return;
      }
    }
  });
}

// u restriction: coarse INSIDE points inside a fine box's non-ghost part
// (all = false), injection at coincident points (Lagrange 3 elsewhere)
__global__ __launch_bounds__(kPtThreads) void k_restrict_u(const PtBlock *__restrict__ blocks) {
  const PtBlock Bk = blocks[xcd_block()];
  ConstBox &T = cbox(Bk.box);
  for_points(Bk, [&](unsigned off) {
    int i, j, k;
    pt_lex(Bk, T, off, i, j, k);
    if (!is_interior(T, i, j, k)) return;
    const int t[3] = {i, j, k};
    ConstXfer *xs = (ConstXfer *)Bk.src;
    for (int x = Bk.nsrc - 1; x >= 0; --x) {
      ConstXfer &X = xs[x];
      unsigned fl[3];
      if (!(axis_flags(X, t, fl) & 1)) continue;
      ConstBox &S = cbox(X.src);
      st<double>(T.f[F_U], boff(T, i, j, k), tinterp<4>(S, S.f[F_U], X, t, fl));
      return;
    }
  });
}

// full weighting (1/4, 1/2, 1/4)^3 of field w of box S around node
// (ci, cj, ck), whose 3x3x3 neighbourhood must lie in the box
__device__ __forceinline__ double fw27(ConstBox &S, const void *w, int ci, int cj, int ck) {
  // walk the 27 points from (ci-1, cj-1, ck-1)
  const Walk wk(S);
  unsigned oz = boff(S, ci - 1, cj - 1, ck - 1);
  unsigned cz = colourc(S, ci - 1, cj - 1, ck - 1);
  const unsigned r0 = unsigned(ci - 1) - divc(S, unsigned(ci - 1)) * wk.nc;
  double acc = 0.0;
  for (int dz = -1; dz <= 1; ++dz) {
    const double wz = dz == 0 ? 0.5 : 0.25;
    unsigned oy = oz, cy = cz;
    for (int dy = -1; dy <= 1; ++dy) {
      const double wy = dy == 0 ? 0.5 : 0.25;
      unsigned cx = cy, r = r0;
      const unsigned om = oy;
      const unsigned o0 = om + wk.x(cx, r);
      const unsigned op = o0 + wk.x(cx, r);
      const double wm = ld<double>(w, om);
      const double w0 = ld<double>(w, o0);
      const double wp = ld<double>(w, op);
      acc += wz * wy * (0.25 * wm + 0.5 * w0 + 0.25 * wp);
      oy += wk.y(cy);
    }
    oz += wk.z(cz);
  }
  return acc;
}

// Filtered defect of a fine box whose defect is restricted to points of a
// coarser level that do not coincide with its nodes (a merged box off the
// coarser lattice): v = the full weighting of w at every node whose 3x3x3
// neighbourhood is interior, v = w elsewhere.  The restriction then
// interpolates v instead of w: the filter removes the (i+j+k)-parity
// pattern that the multi-colour sweeps leave in the defect, which the
// Lagrange interpolation would otherwise alias into the coarse correction.
// At a converged u the interior defect is zero, so the fixed point is the
// one of the unfiltered restriction.
__global__ __launch_bounds__(kPtThreads) void k_filter_defect(const PtBlock *__restrict__ blocks) {
  const PtBlock Bk = blocks[xcd_block()];
  ConstBox &S = cbox(Bk.box);
  for_points(Bk, [&](unsigned off) {
    int i, j, k;
    pt_lex(Bk, S, off, i, j, k);
    const int R = S.R;
    const unsigned o = boff(S, i, j, k);
    const bool in = i - 1 >= R && i + 1 < S.n[0] - R && j - 1 >= R && j + 1 < S.n[1] - R &&
                    k - 1 >= R && k + 1 < S.n[2] - R;
    st<double>(S.f[F_V], o, in ? fw27(S, S.f[F_W], i, j, k) : ld<double>(S.f[F_W], o));
  });
}

// The same filter, one thread per run of kFilterCx points along x of one
// row: the 9 neighbour rows' kFilterCx + 2 values are loaded once for the
// run (the lanes of a wave then read consecutive slots of one colour block
// per load), and each point sums its 27 terms in fw27's order (same bits).
constexpr int kFilterCx = 4;
__global__ __launch_bounds__(kPtThreads) void k_filter_cols(const PtBlock *__restrict__ blocks) {
  constexpr int NX = kFilterCx + 2;
  const PtBlock Bk = blocks[xcd_block()];
  ConstBox &S = cbox(Bk.box);
  if (threadIdx.x >= Bk.nq) return;
  const int n0 = S.n[0], n1 = S.n[1], n2 = S.n[2], R = S.R;
  const int nxc = (n0 + kFilterCx - 1) / kFilterCx;
  const int q = int(Bk.q0 + threadIdx.x);
  const int row = q / nxc;
  const int i0 = (q - row * nxc) * kFilterCx;
  const int k = row / n1, j = row - k * n1;
  const void *w = S.f[F_W];
  const bool in_jk = j - 1 >= R && j + 1 < n1 - R && k - 1 >= R && k + 1 < n2 - R;
  bool in[kFilterCx];
  bool any = false;
#pragma unroll
  for (int p = 0; p < kFilterCx; ++p) {
    const int i = i0 + p;
    in[p] = in_jk && i - 1 >= R && i + 1 < n0 - R;
    any = any || in[p];
  }
  double acc[kFilterCx];
#pragma unroll
  for (int p = 0; p < kFilterCx; ++p) acc[p] = 0.0;
  if (any) {
    // x nodes i0 - 1 .. i0 + kFilterCx of rows j - 1 .. j + 1, planes
    // k - 1 .. k + 1 (all inside the box: in_jk); nodes past the box ends
    // along x are not loaded and feed no filtered point
    const Walk wk(S);
    unsigned ox[NX], cx[NX];
    bool okx[NX];
#pragma unroll
    for (int x = 0; x < NX; ++x) {
      const int xi = i0 - 1 + x;
      okx[x] = xi >= 0 && xi < n0;
      ox[x] = okx[x] ? boff(S, xi, j - 1, k - 1) : 0u;
      cx[x] = okx[x] ? colourc(S, xi, j - 1, k - 1) : 0u;
    }
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz) {
      const double wz = dz == 0 ? 0.5 : 0.25;
      unsigned oy[NX], cy[NX];
#pragma unroll
      for (int x = 0; x < NX; ++x) {
        oy[x] = ox[x];
        cy[x] = cx[x];
      }
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy) {
        const double wy = dy == 0 ? 0.5 : 0.25;
        double val[NX];
#pragma unroll
        for (int x = 0; x < NX; ++x) val[x] = okx[x] ? ld<double>(w, oy[x]) : 0.0;
#pragma unroll
        for (int p = 0; p < kFilterCx; ++p)
          if (in[p]) acc[p] += wz * wy * (0.25 * val[p] + 0.5 * val[p + 1] + 0.25 * val[p + 2]);
#pragma unroll
        for (int x = 0; x < NX; ++x) oy[x] += wk.y(cy[x]);
      }
#pragma unroll
      for (int x = 0; x < NX; ++x) ox[x] += wk.z(cx[x]);
    }
  }
#pragma unroll
  for (int p = 0; p < kFilterCx; ++p) {
    const int i = i0 + p;
    if (i >= n0) continue;
    const unsigned o = boff(S, i, j, k);
    st<double>(S.f[F_V], o, in[p] ? acc[p] : ld<double>(w, o));
  }
}

// Coarse FAS source: W_c = restriction of the fine defect (full weighting
// where the coarse point coincides with a fine point whose 3x3x3
// neighbourhood is fine interior, injection / Lagrange-2 elsewhere, of the
// filtered defect k_filter_defect leaves in v when the source box is off
// the coarse lattice), then
// rho_c = W_c + R_c with R_c = N(u_c) inside, the Robin operator on the
// shell (multigrid levels).  Refinement levels overwrite rho_c only where
// the point is inside both boxes' non-ghost parts (ffunction.cc:2808-2849).
// Also saves ww_c = u_c on every coarse point.
template <int ORD, int EQ>
__global__ __launch_bounds__(kPtThreads) void k_restrict_defect(const PtBlock *__restrict__ blocks) {
  const PtBlock Bk = blocks[xcd_block()];
  ConstBox &T = cbox(Bk.box);
  for_points(Bk, [&](unsigned off) {
    int i, j, k;
    pt_lex(Bk, T, off, i, j, k);
    const unsigned o = boff(T, i, j, k);
    const double uc = ld<double>(T.f[F_U], o);
    st<double>(T.f[F_WW], o, uc);
    const bool inside = is_interior(T, i, j, k);
    if (!Bk.mg && !inside) return;
    const int t[3] = {i, j, k};
    bool covered = false;
    double wc = 0.0;
    ConstXfer *xs = (ConstXfer *)Bk.src;
    for (int x = Bk.nsrc - 1; x >= 0; --x) {
      ConstXfer &X = xs[x];
      unsigned fl[3];
      const unsigned f = axis_flags(X, t, fl);
      if (!((f & 1) || (Bk.mg && (f & 2)))) continue;
      ConstBox &S = cbox(X.src);
      if (f & 8) {
        wc = fw27(S, S.f[F_W], X.ax[0].coinc[i], X.ax[1].coinc[j], X.ax[2].coinc[k]);
      } else {
        wc = tinterp<3>(S, X.filt ? S.f[F_V] : S.f[F_W], X, t, fl);
      }
      covered = true;
      break;
    }
    if (!covered && !Bk.mg) return;
    if (covered) st<double>(T.f[F_W], o, wc); else wc = ld<double>(T.f[F_W], o);
    double Rc;
    if (inside) {
      constexpr int NC = colours_of(ORD, EQ);
      const Nbr<NC> nb = box_offs<NC, 8>(T, int(colourc(T, i, j, k)), i % NC);
      double dN;
      point_op<ORD, EQ>(T, T.f[F_U], o, nb, uc, i, j, k, Rc, dN);
    } else {
      Rc = robin_op_at(cbox_plain(T), i, j, k);
    }
    st<double>(T.f[F_RHO], o, wc + Rc);
  });
}

// v = u - ww on the listed points of a coarse box (the nodes the next
// prolongation reads): the coarse correction, formed once per node instead
// of once per fine point that interpolates from it
__global__ __launch_bounds__(kPtThreads) void k_corr_v(const PtBlock *__restrict__ blocks) {
  const PtBlock Bk = blocks[xcd_block()];
  ConstBox &S = cbox(Bk.box);
  for_points(Bk, [&](unsigned off) {
    int i, j, k;
    pt_lex(Bk, S, off, i, j, k);
    const unsigned o = boff(S, i, j, k);
    st<double>(S.f[F_V], o, ld<double>(S.f[F_U], o) - ld<double>(S.f[F_WW], o));
  });
}

// prolongation of the coarse correction v (k_corr_v) and u += I v (all
// points on multigrid levels, interior points on refinement levels)
__global__ __launch_bounds__(kPtThreads) void k_prolong(const PtBlock *__restrict__ blocks) {
  const PtBlock Bk = blocks[xcd_block()];
  ConstBox &T = cbox(Bk.box);
  for_points(Bk, [&](unsigned off) {
    int i, j, k;
    pt_lex(Bk, T, off, i, j, k);
    if (!Bk.mg && !is_interior(T, i, j, k)) return;
    const int t[3] = {i, j, k};
    // the target value is loaded first; its latency overlaps the source's
    const unsigned o = boff(T, i, j, k);
    const double u0 = ld<double>(T.f[F_U], o);
    ConstXfer *xs = (ConstXfer *)Bk.src;
    for (int x = Bk.nsrc - 1; x >= 0; --x) {
      ConstXfer &X = xs[x];
      unsigned fl[3];
      const unsigned f = axis_flags(X, t, fl);
      if (!(f & 3)) continue;  // all = true
      ConstBox &S = cbox(X.src);
      double v;
      if (f & 4) {
        v = ld<double>(S.f[F_V], boff(S, X.ax[0].coinc[i], X.ax[1].coinc[j], X.ax[2].coinc[k]));
      } else if (X.affine) {
        v = trilinear_nested(S, X, t);
      } else {
        v = trilinear_corr(S, X, t);
      }
      st<double>(T.f[F_U], o, u0 + v);
      return;
    }
  });
}

// shell point q of a box: the R-thick z slabs first, then y slabs, then x slabs
__device__ __forceinline__ void shell_point(ConstBox &B, unsigned q, int &i, int &j, int &k) {
  const unsigned R = unsigned(B.R);
  const unsigned n0 = unsigned(B.n[0]), n1 = unsigned(B.n[1]), n2 = unsigned(B.n[2]);
  const unsigned zs = n0 * n1 * R;  // one z slab
  if (q < 2 * zs) {
    const unsigned s = q >= zs, r = q - s * zs;
    const unsigned plane = n0 * n1;
    const unsigned l = fast_div(r, plane, B.in01), rr = r - l * plane;
    j = int(fast_div(rr, n0, B.in0));
    i = int(rr - unsigned(j) * n0);
    k = s ? int(n2 - R + l) : int(l);
    return;
  }
  q -= 2 * zs;
  const unsigned ys = n0 * R * (n2 - 2 * R);
  if (q < 2 * ys) {
    const unsigned s = q >= ys, r = q - s * ys;
    const unsigned row = fast_div(r, n0, B.in0);
    i = int(r - row * n0);
    const unsigned kk = row / R, l = row - kk * R;
    k = int(R + kk);
    j = s ? int(n1 - R + l) : int(l);
    return;
  }
  q -= 2 * ys;
  const unsigned xs = R * (n1 - 2 * R) * (n2 - 2 * R);
  const unsigned s = q >= xs, r = q - s * xs;
  const unsigned col = r / R, l = r - col * R;
  const unsigned kk = fast_div(col, n1 - 2 * R, B.inyi);
  j = int(R + (col - kk * (n1 - 2 * R)));
  k = int(R + kk);
  i = s ? int(n0 - R + l) : int(l);
}

// column slot q of the ghost shell (ghost_column_slots): the R shell
// points that differ along the slab normal `axis` only (z slabs: columns
// (i, j); y slabs: (i, k) with k interior; x slabs: (j, k) with j, k
// interior; low side first, each face in tiles of kColTile^2 columns);
// returns the column's first point, or false for a padding slot
__device__ __forceinline__ bool face_tile_pos(unsigned r, unsigned f0, unsigned f1, unsigned &a0,
                                              unsigned &a1) {
  const unsigned tx = col_tiles(f0);
  const unsigned st = r / unsigned(kColSuper * kColSuper), w = r % unsigned(kColSuper * kColSuper);
  const unsigned sy = fast_div(st, tx, 1.0f / float(tx));
  const unsigned wave = w / unsigned(kColTile * kColTile), lane = w % unsigned(kColTile * kColTile);
  a0 = (st - sy * tx) * kColSuper + (wave & 1u) * kColTile + lane % kColTile;
  a1 = sy * kColSuper + (wave >> 1) * kColTile + lane / kColTile;
  return a0 < f0 && a1 < f1;
}
__device__ __forceinline__ bool shell_column(ConstBox &B, unsigned q, int &i, int &j, int &k,
                                             int &axis) {
  const unsigned R = unsigned(B.R);
  const unsigned n0 = unsigned(B.n[0]), n1 = unsigned(B.n[1]), n2 = unsigned(B.n[2]);
  unsigned a0, a1;
  const unsigned zc = face_slots(n0, n1);
  if (q < 2 * zc) {
    const unsigned s = q >= zc;
    if (!face_tile_pos(q - s * zc, n0, n1, a0, a1)) return false;
    i = int(a0);
    j = int(a1);
    k = s ? int(n2 - R) : 0;
    axis = 2;
    return true;
  }
  q -= 2 * zc;
  const unsigned yc = face_slots(n0, n2 - 2 * R);
  if (q < 2 * yc) {
    const unsigned s = q >= yc;
    if (!face_tile_pos(q - s * yc, n0, n2 - 2 * R, a0, a1)) return false;
    i = int(a0);
    k = int(R + a1);
    j = s ? int(n1 - R) : 0;
    axis = 1;
    return true;
  }
  q -= 2 * yc;
  const unsigned xc = face_slots(n1 - 2 * R, n2 - 2 * R);
  const unsigned s = q >= xc;
  if (!face_tile_pos(q - s * xc, n1 - 2 * R, n2 - 2 * R, a0, a1)) return false;
  j = int(R + a0);
  k = int(R + a1);
  i = s ? int(n0 - R) : 0;
  axis = 0;
  return true;
}

// ghost shell of refinement boxes from the coarser level (Lagrange
// order+1, KP = order+2 nodes per axis); one thread per shell point,
// point q being layer q mod R of column slot q / R (shell_column), so that
// a wave's points cover a compact patch of a face
template <int KP>
__global__ __launch_bounds__(kPtThreads) void k_ghost_fill(const PtBlock *__restrict__ blocks) {
  const PtBlock Bk = blocks[xcd_block()];
  ConstBox &T = cbox(Bk.box);
  const unsigned R = unsigned(T.R);
  for_points(Bk, [&](unsigned off) {
    const unsigned q = Bk.q0 + off, c = q / R;
    int i0, j0, k0, ax;
    if (!shell_column(T, c, i0, j0, k0, ax)) return;
    const int l = int(q - c * R);
    const int i = i0 + (ax == 0 ? l : 0), j = j0 + (ax == 1 ? l : 0), k = k0 + (ax == 2 ? l : 0);
    const int t[3] = {i, j, k};
    ConstXfer *xs = (ConstXfer *)Bk.src;
    for (int x = Bk.nsrc - 1; x >= 0; --x) {
      ConstXfer &X = xs[x];
      unsigned fl[3];
      const unsigned f = axis_flags(X, t, fl);
      if (!(f & 3)) continue;
      ConstBox &S = cbox(X.src);
      st<double>(T.f[F_U], boff(T, i, j, k), tinterp<KP>(S, S.f[F_U], X, t, fl));
      return;
    }
  });
}

// Ghost shell of refinement boxes, one thread per shell column (above).
// The column's points share the node ranges and weights of the two axes
// across the slab, so the node values are loaded once for all of them over
// the union of their node ranges along the normal (at most KP + 3 nodes);
// each point's sums run over its own nodes in tinterp's order (z innermost,
// then y, then x, ascending), so every value is tinterp's to the bit.
// Columns whose points take different sources fall back to tinterp point
// by point.
// The nodes a column's fast path reads (the per-axis ranges below: the
// two axes across the slab and the union along the normal) and its source;
// false when the column takes the point-by-point fallback
template <int KP, int ax>
__device__ __forceinline__ bool column_region(const PtBlock &Bk, const int p0[3], int &src,
                                              int lo3[3], int hi3[3]) {
  constexpr int KU = KP + 3;
  constexpr int RG = (KP - 2) / 2;
  ConstXfer *xs = (ConstXfer *)Bk.src;
  unsigned fn[RG], fl0[3] = {0, 0, 0};
  src = -1;
#pragma unroll
  for (int l = 0; l < RG; ++l) {
    int sl = -1;
    fn[l] = 0;
    int t[3] = {p0[0], p0[1], p0[2]};
    t[ax] += l;
    for (int x = Bk.nsrc - 1; x >= 0; --x) {
      unsigned fl[3];
      const unsigned f = axis_flags(xs[x], t, fl);
      if (!(f & 3)) continue;
      sl = x;
      fn[l] = fl[ax];
      if (l == 0) {
        fl0[0] = fl[0];
        fl0[1] = fl[1];
        fl0[2] = fl[2];
      }
      break;
    }
    if (l == 0) src = sl;
    if (sl < 0 || sl != src) return false;
  }
  ConstXfer &X = xs[src];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a == ax) continue;
    lo3[a] = (fl0[a] & 4) ? X.ax[a].coinc[p0[a]] : X.ax[a].start[p0[a]];
    hi3[a] = lo3[a] + ((fl0[a] & 4) ? 1 : KP);
  }
  int lo = 1 << 30, hi = -1;
#pragma unroll
  for (int l = 0; l < RG; ++l) {
    const int t = p0[ax] + l;
    const int first = (fn[l] & 4) ? X.ax[ax].coinc[t] : X.ax[ax].start[t];
    lo = min(lo, first);
    hi = max(hi, first + ((fn[l] & 4) ? 1 : KP) - 1);
  }
  if (hi - lo + 1 > KU) return false;
  lo3[ax] = lo;
  hi3[ax] = hi + 1;
  return true;
}

// STAGED: the column's source nodes are read from lds, the workgroup's copy
// of the source box's nodes [lb, lb + (nx, ny, ...)) (k_ghost_cols)
template <int KP, int ax, bool STAGED>
__device__ __forceinline__ void ghost_column(const PtBlock &Bk, ConstBox &T, const int p0[3],
                                             const double *lds, const int lb[3], int nx, int ny) {
  constexpr int KU = KP + 3;       // union of the column's node ranges, at most
  constexpr int RG = (KP - 2) / 2;  // ghost width of the boxes of this stencil
  ConstXfer *xs = (ConstXfer *)Bk.src;
  int src[RG];
  unsigned fn[RG];  // flags of the normal axis per point
  unsigned fl0[3] = {0, 0, 0};
  bool same = true;
#pragma unroll
  for (int l = 0; l < RG; ++l) {
    src[l] = -1;
    fn[l] = 0;
    int t[3] = {p0[0], p0[1], p0[2]};
    t[ax] += l;
    for (int x = Bk.nsrc - 1; x >= 0; --x) {
      unsigned fl[3];
      const unsigned f = axis_flags(xs[x], t, fl);
      if (!(f & 3)) continue;
      src[l] = x;
      fn[l] = fl[ax];
      if (l == 0) {
        fl0[0] = fl[0];
        fl0[1] = fl[1];
        fl0[2] = fl[2];
      }
      break;
    }
    same = same && src[l] == src[0];
  }
  if (src[0] < 0 || !same) {
#pragma unroll
    for (int l = 0; l < RG; ++l) {
      if (src[l] < 0) continue;
      int t[3] = {p0[0], p0[1], p0[2]};
      t[ax] += l;
      ConstXfer &X = xs[src[l]];
      unsigned fl[3];
      axis_flags(X, t, fl);
      ConstBox &S = cbox(X.src);
      st<double>(T.f[F_U], boff(T, t[0], t[1], t[2]), tinterp<KP>(S, S.f[F_U], X, t, fl));
    }
    return;
  }
  ConstXfer &X = xs[src[0]];
  ConstBox &S = cbox(X.src);
  const void *f = S.f[F_U];
  // the two axes across the slab: shared nodes and weights (tinterp's)
  int s0[3], kk[3];
  const double *w[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (fl0[a] & 4) {
      s0[a] = X.ax[a].coinc[p0[a]];
      kk[a] = 1;
      w[a] = nullptr;
    } else {
      s0[a] = X.ax[a].start[p0[a]];
      kk[a] = KP;
      w[a] = X.ax[a].w + p0[a] * KP;
    }
  }
  // the normal: each point's first node and count, their union from lo
  int first[RG], cnt[RG];
  int lo = 1 << 30, hi = -1;
#pragma unroll
  for (int l = 0; l < RG; ++l) {
    const int t = p0[ax] + l;
    if (fn[l] & 4) {
      first[l] = X.ax[ax].coinc[t];
      cnt[l] = 1;
    } else {
      first[l] = X.ax[ax].start[t];
      cnt[l] = KP;
    }
    lo = min(lo, first[l]);
    hi = max(hi, first[l] + cnt[l] - 1);
  }
  const int U = hi - lo + 1;
  if (U > KU) {  // not a refinement by two along the normal: point by point
#pragma unroll
    for (int l = 0; l < RG; ++l) {
      int t[3] = {p0[0], p0[1], p0[2]};
      t[ax] += l;
      unsigned fl[3];
      axis_flags(X, t, fl);
      st<double>(T.f[F_U], boff(T, t[0], t[1], t[2]), tinterp<KP>(S, f, X, t, fl));
    }
    return;
  }
  // each point's weights along the normal at union positions (zero outside
  // its range, where no term is added)
  double wn[RG][KU];
#pragma unroll
  for (int l = 0; l < RG; ++l) {
    const int d = first[l] - lo;
    const double *wl = (fn[l] & 4) ? nullptr : X.ax[ax].w + (p0[ax] + l) * KP;
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      double v = 0.0;
      if (u >= d && u < d + cnt[l]) v = wl ? wl[u - d] : 1.0;
      wn[l][u] = v;
    }
  }
  s0[ax] = lo;
  kk[ax] = U;
  const Walk wk(S);
  unsigned oI = boff(S, s0[0], s0[1], s0[2]);
  unsigned cI = colourc(S, s0[0], s0[1], s0[2]);
  unsigned rI = unsigned(s0[0]) - divc(S, unsigned(s0[0])) * wk.nc;
  double acc[RG];
#pragma unroll
  for (int l = 0; l < RG; ++l) acc[l] = 0.0;
  // a node of the union lies in point l's range
#define OLL_IN(l, u) ((u) >= first[l] - lo && (u) < first[l] - lo + cnt[l])
  for (int I = 0; I < kk[0]; ++I) {
    double uj[RG];
#pragma unroll
    for (int l = 0; l < RG; ++l) uj[l] = 0.0;
    double ujx = 0.0;  // normal along x: the shared (y, z) sum
    unsigned oJ = oI, cJ = cI;
    if (ax == 1) {
      // normal along y: z sums of every union row, then per point over y
      double uk[KU];
#pragma unroll
      for (int J = 0; J < KU; ++J) {
        unsigned off = oJ, c = cJ;
        double v[KP];
#pragma unroll
        for (int K = 0; K < KP; ++K) {
          if constexpr (STAGED)
            v[K] = (J < U && K < kk[2]) ? lds[((s0[2] + K - lb[2]) * ny + (s0[1] + J - lb[1])) * nx +
                                              (s0[0] + I - lb[0])]
                                        : 0.0;
          else
            v[K] = (J < U && K < kk[2]) ? ld<double>(f, off) : 0.0;
          off += wk.z(c);
        }
        double s = 0.0;
#pragma unroll
        for (int K = 0; K < KP; ++K)
          if (K < kk[2]) s += v[K] * (w[2] ? w[2][K] : 1.0);
        uk[J] = s;
        if (J < U) oJ += wk.y(cJ);
      }
#pragma unroll
      for (int l = 0; l < RG; ++l)
#pragma unroll
        for (int J = 0; J < KU; ++J)
          if (OLL_IN(l, J)) uj[l] += uk[J] * wn[l][J];
    } else {
      for (int J = 0; J < kk[1]; ++J) {
        unsigned off = oJ, c = cJ;
        const double wy = w[1] ? w[1][J] : 1.0;
        if (ax == 2) {
          // normal along z: one union row, then per point over z
          double v[KU];
#pragma unroll
          for (int K = 0; K < KU; ++K) {
            if constexpr (STAGED)
              v[K] = K < U ? lds[((s0[2] + K - lb[2]) * ny + (s0[1] + J - lb[1])) * nx +
                                 (s0[0] + I - lb[0])]
                           : 0.0;
            else
              v[K] = K < U ? ld<double>(f, off) : 0.0;
            off += wk.z(c);
          }
#pragma unroll
          for (int l = 0; l < RG; ++l) {
            double s = 0.0;
#pragma unroll
            for (int K = 0; K < KU; ++K)
              if (OLL_IN(l, K)) s += v[K] * wn[l][K];
            uj[l] += s * wy;
          }
        } else {
          // normal along x: the shared row sum
          double v[KP];
#pragma unroll
          for (int K = 0; K < KP; ++K) {
            if constexpr (STAGED)
              v[K] = K < kk[2] ? lds[((s0[2] + K - lb[2]) * ny + (s0[1] + J - lb[1])) * nx +
                                     (s0[0] + I - lb[0])]
                               : 0.0;
            else
              v[K] = K < kk[2] ? ld<double>(f, off) : 0.0;
            off += wk.z(c);
          }
          double s = 0.0;
#pragma unroll
          for (int K = 0; K < KP; ++K)
            if (K < kk[2]) s += v[K] * (w[2] ? w[2][K] : 1.0);
          ujx += s * wy;
        }
        oJ += wk.y(cJ);
      }
    }
    if (ax == 0) {
#pragma unroll
      for (int l = 0; l < RG; ++l)
        if (OLL_IN(l, I)) {
          double wI = 0.0;  // wn[l][I] without a dynamic register index
#pragma unroll
          for (int u = 0; u < KU; ++u)
            if (u == I) wI = wn[l][u];
          acc[l] += ujx * wI;
        }
    } else {
      const double wx = w[0] ? w[0][I] : 1.0;
#pragma unroll
      for (int l = 0; l < RG; ++l) acc[l] += uj[l] * wx;
    }
    oI += wk.x(cI, rI);
  }
#undef OLL_IN
#pragma unroll
  for (int l = 0; l < RG; ++l) {
    int t[3] = {p0[0], p0[1], p0[2]};
    t[ax] += l;
    st<double>(T.f[F_U], boff(T, t[0], t[1], t[2]), acc[l]);
  }
}

// A workgroup takes one kColSuper^2 tile of a face's columns; when the
// columns on the fast path share their source, the nodes they read (their
// bounding box, at most kGhostLds) are staged in LDS once and every column
// reads its stencil from there.  The same values in the same order either
// way.
constexpr int kGhostLds = 4608;
template <int KP>
__global__ __launch_bounds__(kPtThreads) void k_ghost_cols(const PtBlock *__restrict__ blocks) {
  __shared__ double lds[kGhostLds];
  __shared__ int red[9];  // lo[3], hi[3], min and max source, a column off the fast path
  const PtBlock Bk = blocks[xcd_block()];
  ConstBox &T = cbox(Bk.box);
  ConstXfer *xs = (ConstXfer *)Bk.src;
  // every thread walks the same chunks: the barriers below are uniform
  for (unsigned base = 0; base < Bk.nq; base += kPtThreads) {
    int p0[3] = {0, 0, 0}, ax = 0;
    const bool valid =
        base + threadIdx.x < Bk.nq && shell_column(T, Bk.q0 + base + threadIdx.x, p0[0], p0[1], p0[2], ax);
    int src = -1, lo3[3] = {0, 0, 0}, hi3[3] = {0, 0, 0};
    bool fast = false;
    if (valid) {
      if (ax == 0)
        fast = column_region<KP, 0>(Bk, p0, src, lo3, hi3);
      else if (ax == 1)
        fast = column_region<KP, 1>(Bk, p0, src, lo3, hi3);
      else
        fast = column_region<KP, 2>(Bk, p0, src, lo3, hi3);
    }
    if (threadIdx.x < 9) red[threadIdx.x] = threadIdx.x < 3 ? (1 << 30) : (threadIdx.x == 6 ? (1 << 30) : -1);
    __syncthreads();
    if (fast) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        atomicMin(&red[a], lo3[a]);
        atomicMax(&red[3 + a], hi3[a]);
      }
      atomicMin(&red[6], src);
      atomicMax(&red[7], src);
    }
    __syncthreads();
    const int lb[3] = {red[0], red[1], red[2]};
    const int nx = red[3] - red[0], ny = red[4] - red[1], nz = red[5] - red[2];
    const bool staged = red[6] == red[7] && nx > 0 && ny > 0 && nz > 0 && nx * ny * nz <= kGhostLds;
    if (staged) {
      ConstBox &S = cbox(xs[red[6]].src);
      const void *f = S.f[F_U];
      const int nxy = nx * ny;
      for (int q = threadIdx.x; q < nxy * nz; q += kPtThreads) {
        const int z = q / nxy, r = q - z * nxy, y = r / nx, x = r - y * nx;
        lds[q] = ld<double>(f, boff(S, lb[0] + x, lb[1] + y, lb[2] + z));
      }
    }
    __syncthreads();
    if (valid) {
      if (staged && fast) {
        if (ax == 0)
          ghost_column<KP, 0, true>(Bk, T, p0, lds, lb, nx, ny);
        else if (ax == 1)
          ghost_column<KP, 1, true>(Bk, T, p0, lds, lb, nx, ny);
        else
          ghost_column<KP, 2, true>(Bk, T, p0, lds, lb, nx, ny);
      } else {
        if (ax == 0)
          ghost_column<KP, 0, false>(Bk, T, p0, nullptr, lb, 0, 0);
        else if (ax == 1)
          ghost_column<KP, 1, false>(Bk, T, p0, nullptr, lb, 0, 0);
        else
          ghost_column<KP, 2, false>(Bk, T, p0, nullptr, lb, 0, 0);
      }
    }
    __syncthreads();  // lds and red are reused by the next chunk
  }
}

// r = u on the ghost shell of the listed boxes (the fused sweeps read
// ghost values from either field)
__global__ __launch_bounds__(kPtThreads) void k_copy_shell(const PtBlock *__restrict__ blocks) {
  const PtBlock Bk = blocks[xcd_block()];
  ConstBox &B = cbox(Bk.box);
  for_points(Bk, [&](unsigned off) {
    int i, j, k;
    shell_point(B, Bk.q0 + off, i, j, k);
    const unsigned o = boff(B, i, j, k);
    st<double>(B.f[F_R], o, ld<double>(B.f[F_U], o));
  });
}

// ------------------------------------------------------------------
// coarsest level: {Robin, one multi-colour sweep} until the sweep norm is
// <= eps or stops changing (elliptic.cc:1782-1816); one block per
// hierarchy, u and rho in LDS (the box's colour-block layout) for the whole loop, b,
// c (d, e) of the owned interior points in registers.  Interior updates are
// over-relaxed by omega (omega = 1 is the reference's Gauss-Seidel-Newton);
// the loop's fixed point, the coarse solution, does not depend on omega.
// The coarsest level always has order 2 (R = 1): its Robin shell is one
// layer of faces, then edges, then corners (boundary.cc:447-1101), each
// phase reading only the previous ones; every thread decodes its shell
// points once.
// ------------------------------------------------------------------
struct CoarseGeo {
  int n[3];
  int rl, rs, ps, cs, nc;
  unsigned ncm;
  double lo[3], d[3], idx[3], idx2[3];
  double ipn, dulap;
};

// element index of (i, j, k) in the LDS copy (the box's colour-block layout)
__device__ __forceinline__ int coarse_p(const CoarseGeo &G, int i, int j, int k) {
  const unsigned s = unsigned(i + j + k);
  const int c = int(s - ((s * G.ncm) >> 16) * unsigned(G.nc));
  return c * G.cs + k * G.ps + j * G.rs + int((unsigned(i) * G.ncm) >> 16);
}

// Robin update at shell point (i, j, k) of the LDS field; mask has bit a
// for each axis whose index is 0 or n-1 (dx2 one-sided derivative,
// diff_2nd.cc:237-335).  The coarsest level is 2nd order, so its layout has
// two colours: point (i, j, k) is slot i / 2 of row (j, k) in block
// (i + j + k) & 1, and a step along an axis moves to the other block.
__device__ __forceinline__ void coarse_robin(const CoarseGeo &G, double *su, const double *srho,
                                             int i, int j, int k, int mask) {
  const int c = (i + j + k) & 1;
  const int p = c * G.cs + k * G.ps + j * G.rs + (i >> 1);
  const int flip = (1 - 2 * c) * G.cs;
  const int ijk[3] = {i, j, k};
  double X[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) X[a] = G.lo[a] + (double)ijk[a] * G.d[a];
  const double r2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
  double sF = 0.0, sdF = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!(mask & (1 << a))) continue;
    const bool lo = ijk[a] == 0;
    const int dir = lo ? 1 : -1;
    // offsets of the points 1 and 2 steps inwards along axis a
    int o1, o2;
    if (a == 0) {
      o1 = flip + ((i + dir) >> 1) - (i >> 1);
      o2 = dir;
    } else {
      const int st = a == 1 ? G.rs : G.ps;
      o1 = flip + dir * st;
      o2 = 2 * dir * st;
    }
    const double id = G.idx[a];
    double cf, dv;
    if (lo) {
      cf = -3.0 * 0.5 * id;
      dv = (-3.0 * su[p] + 4.0 * su[p + o1] - su[p + o2]) * 0.5 * id;
    } else {
      cf = 3.0 * 0.5 * id;
      dv = (3.0 * su[p] - 4.0 * su[p + o1] + su[p + o2]) * 0.5 * id;
    }
    sF += dv / X[a];
    sdF += cf / X[a];
  }
  const double op = r2 * G.ipn * sF + su[p];
  const double Fv = op - srho[p];
  const double dF = r2 * G.ipn * sdF + 1.0;
  su[p] = su[p] - Fv / dF;
}

// OWN: interior points per colour per thread (colour c's points are
// numbered row by row, ceil(ni / 2) slots per interior row).  With at most
// 2048 points per colour over the workgroup (OWN * NT) the coefficients of
// a thread's points stay in registers for the whole loop; beyond that they
// would not fit the register file (the kernel spilled), so each point's
// coefficients are loaded from the (cache-resident) fields as the point is
// updated.
template <int EQ, int OWN, int NT>
__global__ __launch_bounds__(NT) void k_coarse_solve(const CoarseItem *__restrict__ items,
                                                                 double eps, int max_iter) {
  extern __shared__ double lds[];
  __shared__ double sh[NT / 64];
  const CoarseItem It = items[blockIdx.x];
  ConstBox &B = cbox(It.box);
  CoarseGeo G;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    G.n[a] = B.n[a];
    G.lo[a] = B.lo[a];
    G.d[a] = B.d[a];
    G.idx[a] = B.idx[a];
    G.idx2[a] = B.idx2[a];
  }
  G.rl = B.rl;
  G.rs = B.rs;
  G.ps = int(B.ps);
  G.cs = int(B.cs);
  G.nc = B.nc;
  G.ncm = B.ncm;
  G.ipn = B.ipn;
  G.dulap = B.dulap;
  const double omega = It.omega;
  const int n0 = G.n[0], n1 = G.n[1], n2 = G.n[2];
  const int np = int(B.size);
  double *su = lds;
  double *srho = lds + np;
  double *ug = B.f[F_U];
  const double *rg = B.f[F_RHO];
  for (int q = threadIdx.x; q < np; q += NT) {
    su[q] = ld<double>(ug, unsigned(q) * 8u);
    srho[q] = ld<double>(rg, unsigned(q) * 8u);
  }
  // owned interior points, colour by colour
  const int ni = n0 - 2, nj = n1 - 2, nk = n2 - 2;
  const int per_row = (ni + 1) / 2;
  const int slots = nj * nk * per_row;
  constexpr bool kCached = OWN * NT <= 2048;
  constexpr int NCACHE = kCached ? 2 * OWN : 1;
  int own_p[2 * OWN];
  double own_b[NCACHE], own_cc[NCACHE], own_d[NCACHE], own_e[NCACHE];
  // the coefficients of owned point q (colour col, slot m)
  auto coeffs = [&](int q, double &b, double &cc, double &d, double &e) {
    const unsigned o = unsigned(own_p[q] & 0x3fffffff) * 8u;
    b = ld<double>(B.f[F_B], o);
    cc = ld<double>(B.f[F_C], o);
    d = e = 0.0;
    if (EQ == 1 || EQ == kEqPuncturesEm) {
      d = ld<double>(B.f[F_D], o);
      e = ld<double>(B.f[F_E], o);
    }
  };
#pragma unroll
  for (int col = 0; col < 2; ++col)
#pragma unroll
    for (int m = 0; m < OWN; ++m) {
      const int q = col * OWN + m;
      const int t = threadIdx.x + m * NT;
      own_p[q] = -1;
      if constexpr (kCached) own_b[q] = own_cc[q] = own_d[q] = own_e[q] = 0.0;
      if (t >= slots) continue;
      const int row = t / per_row, tt = t - row * per_row;
      const int j = 1 + row % nj, k = 1 + row / nj;
      const int i = 1 + ((col - j - k - 1) & 1) + 2 * tt;
      if (i >= n0 - 1) continue;
      own_p[q] = coarse_p(G, i, j, k) | ((i & 1) << 30);
      if constexpr (kCached) coeffs(q, own_b[q], own_cc[q], own_d[q], own_e[q]);
    }
  // shell points: faces (x, y, z pairs), edges (XY, XZ, YZ), corners;
  // decoded each sweep (one Robin body keeps the register budget)
  const int fx = nj * nk, fy = ni * nk, fz = ni * nj;
  const int nface = 2 * (fx + fy + fz), nedge = 4 * (ni + nj + nk);
  const float ini = 1.0f / float(ni), inj = 1.0f / float(nj), ink = 1.0f / float(nk);
  __syncthreads();
  double prev = 1.0, cur = 1.0;
  int it = 0;
  do {
    prev = cur;
    for (int t = threadIdx.x; t < nface; t += NT) {
      int i, j, k, mask, r = t;
      if (r < 2 * fx) {
        const int hi = r >= fx;
        r -= hi * fx;
        const int q = int(fast_div(unsigned(r), unsigned(nj), inj));
        i = hi ? n0 - 1 : 0; j = 1 + r - q * nj; k = 1 + q; mask = 1;
      } else if ((r -= 2 * fx) < 2 * fy) {
        const int hi = r >= fy;
        r -= hi * fy;
        const int q = int(fast_div(unsigned(r), unsigned(ni), ini));
        j = hi ? n1 - 1 : 0; i = 1 + r - q * ni; k = 1 + q; mask = 2;
      } else {
        r -= 2 * fy;
        const int hi = r >= fz;
        r -= hi * fz;
        const int q = int(fast_div(unsigned(r), unsigned(ni), ini));
        k = hi ? n2 - 1 : 0; i = 1 + r - q * ni; j = 1 + q; mask = 4;
      }
      coarse_robin(G, su, srho, i, j, k, mask);
    }
    __syncthreads();
    if (threadIdx.x < nedge) {
      int t = threadIdx.x, i, j, k, mask;
      if (t < 4 * nk) {
        const int e = int(fast_div(unsigned(t), unsigned(nk), ink));
        i = (e >> 1) ? n0 - 1 : 0; j = (e & 1) ? n1 - 1 : 0; k = 1 + t - e * nk; mask = 3;
      } else if ((t -= 4 * nk) < 4 * nj) {
        const int e = int(fast_div(unsigned(t), unsigned(nj), inj));
        i = (e >> 1) ? n0 - 1 : 0; k = (e & 1) ? n2 - 1 : 0; j = 1 + t - e * nj; mask = 5;
      } else {
        t -= 4 * nj;
        const int e = int(fast_div(unsigned(t), unsigned(ni), ini));
        k = (e >> 1) ? n2 - 1 : 0; j = (e & 1) ? n1 - 1 : 0; i = 1 + t - e * ni; mask = 6;
      }
      coarse_robin(G, su, srho, i, j, k, mask);
    }
    __syncthreads();
    if (threadIdx.x < 8)
      coarse_robin(G, su, srho, (threadIdx.x >> 1) & 1 ? n0 - 1 : 0, (threadIdx.x >> 2) & 1 ? n1 - 1 : 0,
                   threadIdx.x & 1 ? n2 - 1 : 0, 7);
    __syncthreads();
    double mx = 0.0;
#pragma unroll
    for (int col = 0; col < 2; ++col) {
#pragma unroll
      for (int m = 0; m < OWN; ++m) {
        const int q = col * OWN + m;
        if (own_p[q] < 0) continue;
        // uncached: the point's coefficients loaded here (cache-resident)
        double qb, qc, qd, qe;
        if constexpr (kCached) {
          qb = own_b[q];
          qc = own_cc[q];
          qd = own_d[q];
          qe = own_e[q];
        } else {
          coeffs(q, qb, qc, qd, qe);
        }
        const int p = own_p[q] & 0x3fffffff;
        const Nbr<2> nb = stencil_offs<2, 1>(col, own_p[q] >> 30, unsigned(G.cs), unsigned(G.rs),
                                             unsigned(G.ps));
        const double lu = lap<2>(su, p, nb, G.idx2[0], G.idx2[1], G.idx2[2]);
        double N, dN;
        if constexpr (EQ == kEqPuncturesEm) {
          const unsigned o = unsigned(p) * 8u;
          const double g[3] = {(su[p + int(nb.xp[0])] - su[p + int(nb.xm[0])]) * 0.5 * G.idx[0],
                               (su[p + int(nb.yp[0])] - su[p + int(nb.ym[0])]) * 0.5 * G.idx[1],
                               (su[p + int(nb.zp[0])] - su[p + int(nb.zm[0])]) * 0.5 * G.idx[2]};
          const double ge[3] = {ld<double>(B.f[F_X0], o), ld<double>(B.f[F_X1], o), ld<double>(B.f[F_X2], o)};
          const double gf[3] = {ld<double>(B.f[F_X3], o), ld<double>(B.f[F_X4], o), ld<double>(B.f[F_X5], o)};
          const double du[3] = {0.0, 0.0, 0.0};  // order 2: centred derivatives
          em_nonlinear(lu, G.dulap, su[p], g, du, qb, qd, ge, gf, N, dN);
        } else {
          nonlinear<EQ>(lu, G.dulap, qb, qc + su[p], qd, qe, N, dN);
        }
        const double res = N - srho[p];
        const double un = su[p] - omega * (res / dN);
        su[p] = un;
        mx = fmax(mx, fabs(res));
        if (!(un == un)) mx = __longlong_as_double(0x7ff0000000000000LL);
        // keep each point's live range local (the unrolled bodies would
        // otherwise be interleaved past the register budget)
        __builtin_amdgcn_sched_barrier(0);
      }
      if (col == 0) __syncthreads();
    }
    const double wm = wave_max_dpp(mx);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = wm;
    __syncthreads();
    cur = sh[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) cur = fmax(cur, sh[w]);
    ++it;
  } while (cur > eps && prev != cur && it < max_iter && !isinf(cur));
  for (int q = threadIdx.x; q < np; q += NT) st<double>(ug, unsigned(q) * 8u, su[q]);
  if (threadIdx.x == 0) {
    *It.norm = (unsigned long long)__double_as_longlong(cur);
    *It.iters = it;
  }
}

// ------------------------------------------------------------------
// setup: coefficients b, c [,d, e]; rho = Robin value on the shell, 0
// inside; u = 0
// ------------------------------------------------------------------
__device__ double punctures_c(const PunctureArgs &pa, double X, double Y, double Z) {
  double cval = 1.0;
  for (int i = 0; i < pa.n; ++i) {
    const double x = X - pa.p[i].x, y = Y - pa.p[i].y, z = Z - pa.p[i].z;
    double r = sqrt(x * x + y * y + z * z);
    if (r < 2 * kEps) r = 2 * kEps;
    cval += 0.5 * pa.p[i].m / r;
  }
  return cval;
}

__device__ double punctures_b(const PunctureArgs &pa, double X, double Y, double Z) {
  double A00 = 0, A11 = 0, A22 = 0, A01 = 0, A02 = 0, A12 = 0;
  for (int i = 0; i < pa.n; ++i) {
    const Puncture &q = pa.p[i];
    const double x = X - q.x, y = Y - q.y, z = Z - q.z;
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    double r2 = x2 + y2 + z2;
    if (r2 < 2 * kEps) r2 = 2 * kEps;
    const double r = sqrt(r2);
    const double f = 1.5 / (r2 * r2 * r);
    A00 += f * (q.px * x * (x2 + r2) + q.py * y * (x2 - r2) + q.pz * z * (x2 - r2) +
                4.0 * x * (q.sy * z - q.sz * y));
    A11 += f * (q.px * x * (y2 - r2) + q.py * y * (y2 + r2) + q.pz * z * (y2 - r2) +
                4.0 * y * (q.sz * x - q.sx * z));
    A22 += f * (q.px * x * (z2 - r2) + q.py * y * (z2 - r2) + q.pz * z * (z2 + r2) +
                4.0 * z * (q.sx * y - q.sy * x));
    A01 += f * (q.px * y * (r2 + x2) + q.py * x * (r2 + y2) + q.pz * x * y * z -
                2.0 * q.sx * x * z + 2.0 * q.sy * y * z + 2.0 * q.sz * (x2 - y2));
    A02 += f * (q.px * z * (r2 + x2) + q.py * x * y * z + q.pz * x * (r2 + z2) +
                2.0 * q.sx * x * y + 2.0 * q.sy * (z2 - x2) - 2.0 * q.sz * y * z);
    A12 += f * (q.px * x * y * z + q.py * z * (r2 + y2) + q.pz * y * (r2 + z2) +
                2.0 * q.sx * (y2 - z2) - 2.0 * q.sy * x * y + 2.0 * q.sz * x * z);
  }
  return 0.125 * (A00 * A00 + A11 * A11 + A22 * A22 + 2.0 * (A01 * A01 + A02 * A02 + A12 * A12));
}

__device__ double sf_phi(const PunctureArgs &pa, double r) {
  switch (pa.ps_case) {
    case 1: return pa.ps_phi0 * tanh((r - pa.ps_r0) * pa.ps_isigma);
    case 2: return pa.ps_phi0 * exp(-(r - pa.ps_r0) * (r - pa.ps_r0) * pa.ps_isigma);
    case 3: {
      const double x = r, l2 = 0.6;
      return (0.0481646 * exp(-0.0581545 * (x - 1.8039e-08) * (x - 1.8039e-08) / l2) +
              0.298408 * exp(-0.111412 * (x + 9.6741e-09) * (x + 9.6741e-09) / l2) +
              0.42755 * exp(-0.207156 * (x - 1.09822e-08) * (x - 1.09822e-08) / l2) +
              0.204229 * exp(-0.37742 * (x + 2.13778e-08) * (x + 2.13778e-08) / l2) +
              0.021649 * exp(-0.68406 * (x - 8.78608e-08) * (x - 8.78608e-08) / l2)) * l2;
    }
    default:
      return pa.ps_phi0 * 0.5 * (tanh((r + pa.ps_r0) * pa.ps_isigma) - tanh((r - pa.ps_r0) * pa.ps_isigma));
  }
}

__device__ double sf_d(const PunctureArgs &pa, double r) {
  switch (pa.ps_case) {
    case 1: {
      const double ch = cosh((r - pa.ps_r0) * pa.ps_isigma);
      const double ch2 = ch * ch;
      return pa.ps_d1 / (ch2 * ch2);
    }
    case 2:
      return pa.ps_de1 * (r - pa.ps_r0) * (r - pa.ps_r0) *
             exp(-2 * pa.ps_isigma * (r - pa.ps_r0) * (r - pa.ps_r0));
    case 3: {
      const double x = r, l2 = 10000;
      const double g = -2 * (0.0581545 * (x - 1.8039e-08) * 0.0481646 * exp(-0.0581545 * (x - 1.8039e-08) * (x - 1.8039e-08) / l2) +
                             0.111412 * (x + 9.6741e-09) * 0.298408 * exp(-0.111412 * (x + 9.6741e-09) * (x + 9.6741e-09) / l2) +
                             0.207156 * (x - 1.09822e-08) * 0.42755 * exp(-0.207156 * (x - 1.09822e-08) * (x - 1.09822e-08) / l2) +
                             0.37742 * (x + 2.13778e-08) * 0.204229 * exp(-0.37742 * (x + 2.13778e-08) * (x + 2.13778e-08) / l2) +
                             0.68406 * (x - 8.78608e-08) * 0.021649 * exp(-0.68406 * (x - 8.78608e-08) * (x - 8.78608e-08) / l2)) /
                       (l2 * l2);
      return M_PI * g * g;
    }
    default: {
      const double a = 1. / cosh((r + pa.ps_r0) * pa.ps_isigma);
      const double b = 1. / cosh((r - pa.ps_r0) * pa.ps_isigma);
      const double t = a * a - b * b;
      return pa.ps_de1 * 0.25 * t * t;
    }
  }
}

__device__ double sf_e(const PunctureArgs &pa, double r) {
  if (pa.ps_case == 2 || pa.ps_case == 3 || pa.ps_case == 4) {
    const double ff = pa.ps_e2 * sf_phi(pa, r);
    const double g = 1 - exp(ff);
    return pa.ps_e1 * g * g * exp(-2 * ff);
  }
  return 0.0;
}

__global__ __launch_bounds__(256) void k_coefficients(DevBox B, PunctureArgs pa, double robin_a) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long np = (long long)B.n[0] * B.n[1] * B.n[2];
  if (q >= np) return;
  int i, j, k;
  point_ijk(B, q, i, j, k);
  const long long p = pidx(B, i, j, k);
  const double x = coord(B, 0, i), y = coord(B, 1, j), z = coord(B, 2, k);
  const bool inner = interior(B, i, j, k);
  if (pa.eq == 2) {
    // operator.cc:640-662: rho = Gaussian_source, u = Gaussian (elliptic.h)
    const double r2 = x * x + y * y + z * z, a = 0.004;
    B.f[F_B][p] = 0.0;
    B.f[F_C][p] = 0.0;
    B.f[F_RHO][p] = inner ? a * exp(-r2) * (a + exp(0.5 * r2) * (r2 - 3.0)) : robin_a;
    B.f[F_U][p] = a * exp(-0.5 * r2);
    return;
  }
  if (pa.eq == kEqPuncturesEm) {
    // operator.cc:1010-1100 and 3043-3310: eta = sum m / (2r) and phi =
    // sum q / r with their gradients (r and r^3 clamped at 2 eps); rho = 0,
    // u = 1
    double e[4] = {0, 0, 0, 0}, g[4] = {0, 0, 0, 0};
    for (int b = 0; b < pa.n; ++b) {
      const Puncture &P = pa.p[b];
      const double X = x - P.x, Y = y - P.y, Z = z - P.z;
      double r = sqrt(X * X + Y * Y + Z * Z);
      if (r < 2 * kEps) r = 2 * kEps;
      double r3 = pow(X * X + Y * Y + Z * Z, 1.5);
      if (r3 < 2 * kEps) r3 = 2 * kEps;
      e[0] += P.m / r;
      e[1] -= P.m * X / r3;
      e[2] -= P.m * Y / r3;
      e[3] -= P.m * Z / r3;
      g[0] += P.q / r;
      g[1] -= P.q * X / r3;
      g[2] -= P.q * Y / r3;
      g[3] -= P.q * Z / r3;
    }
    B.f[F_B][p] = 0.5 * e[0];
    B.f[F_C][p] = 0.0;
    B.f[F_D][p] = g[0];
    B.f[F_E][p] = 0.0;
    B.f[F_X0][p] = 0.5 * e[1];
    B.f[F_X1][p] = 0.5 * e[2];
    B.f[F_X2][p] = 0.5 * e[3];
    B.f[F_X3][p] = g[1];
    B.f[F_X4][p] = g[2];
    B.f[F_X5][p] = g[3];
    B.f[F_RHO][p] = inner ? 0.0 : robin_a;
    B.f[F_U][p] = 1.0;
    return;
  }
  if (pa.eq == 3) {
    // operator.cc:707-725 and brill (operator.cc:1688-1698): b = Lq, rho = 0, u = 1
    const double amp = 1.9, z2 = z * z, rho2 = x * x + y * y;
    const double qrr = 2.0 * amp * exp(-rho2 - z2) * (1.0 - 5.0 * rho2 + 2.0 * rho2 * rho2);
    const double qzz = 2.0 * amp * exp(-rho2 - z2) * rho2 * (2.0 * z2 - 1.0);
    B.f[F_B][p] = 0.25 * (qrr + qzz);
    B.f[F_C][p] = 0.0;
    B.f[F_RHO][p] = inner ? 0.0 : robin_a;
    B.f[F_U][p] = 1.0;
    return;
  }
  B.f[F_B][p] = punctures_b(pa, x, y, z);
  B.f[F_C][p] = punctures_c(pa, x, y, z);
  if (pa.eq == 1) {
    const double r = sqrt(x * x + y * y + z * z);
    B.f[F_D][p] = sf_d(pa, r);
    B.f[F_E][p] = sf_e(pa, r);
  }
  B.f[F_RHO][p] = inner ? 0.0 : robin_a;
  B.f[F_U][p] = 0.0;
}

// elliptic.cc:1977-2222 ApproximateID: the terms of the punctures up to the
// first with zero mass, added in puncture order
__global__ __launch_bounds__(256) void k_approx_u(DevBox B, PunctureArgs pa) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long np = (long long)B.n[0] * B.n[1] * B.n[2];
  if (q >= np) return;
  int i, j, k;
  point_ijk(B, q, i, j, k);
  const double x = coord(B, 0, i), y = coord(B, 1, j), z = coord(B, 2, k);
  double u = 0.0;
  for (int b = 0; b < pa.n && pa.p[b].m != 0.0; ++b) u += approx_puncture_u(pa.p[b], x, y, z);
  B.f[F_U][pidx(B, i, j, k)] = u;
}

// ------------------------------------------------------------------
// ADM mass surface integrand: dU = x_hat . grad(c + u) interpolated
// (trilinear) onto a 101 x 101 sphere of radius 0.375 min L around the box
// centre, times sin(theta)
// ------------------------------------------------------------------
__device__ int cell_of(const DevBox &B, int a, double x) {
  const int np = B.n[a];
  const double lo = B.lo[a], hi = coord(B, a, np - 1);
  if (!(lo < x && x < hi)) return (lo >= x) ? 1 : np - 1;
  int lo_i = 0, hi_i = np;
  int i = (np % 2 == 0) ? np / 2 : (np - 1) / 2;
  do {
    const double xi = coord(B, a, i);
    if (xi > x || fabs(xi - x) <= 3 * kEps)
      hi_i = i;
    else
      lo_i = i;
    const int s = hi_i + lo_i;
    i = (s % 2 == 0) ? s / 2 : (s - 1) / 2;
  } while (lo_i + 1 < hi_i);
  return i + 1;
}

__device__ double dU_at(const DevBox &B, const double ctr[3], int i, int j, int k) {
  const long long p = pidx(B, i, j, k);
  const int ijk[3] = {i, j, k};
  double X[3], g[3];
  for (int a = 0; a < 3; ++a) {
    X[a] = coord(B, a, ijk[a]) - ctr[a];
    double cf;
    g[a] = deriv(B, B.f[F_C], p, a, i, j, k, cf) + deriv(B, B.f[F_U], p, a, i, j, k, cf);
  }
  const double ir = 1.0 / sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
  return ir * (X[0] * g[0] + X[1] * g[1] + X[2] * g[2]);
}

__global__ __launch_bounds__(256) void k_adm_sphere(const DevBox *__restrict__ boxes,
                                                     const double *__restrict__ centres,
                                                     double *__restrict__ sph, int nt) {
  const int box = blockIdx.y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nt * nt) return;
  const DevBox &B = boxes[box];
  const double ctr[3] = {centres[3 * box], centres[3 * box + 1], centres[3 * box + 2]};
  const double L0 = coord(B, 0, B.n[0] - 1) - B.lo[0];
  const double L1 = coord(B, 1, B.n[1] - 1) - B.lo[1];
  const double L2 = coord(B, 2, B.n[2] - 1) - B.lo[2];
  const double r0 = 0.375 * fmin(L0, fmin(L1, L2));
  const int it = q / nt, ip = q % nt;
  const double dtheta = M_PI / (nt - 1), dphi = 2 * M_PI / (nt - 1);
  const double th = it * dtheta, ph = ip * dphi;
  const double X[3] = {ctr[0] + r0 * cos(ph) * sin(th), ctr[1] + r0 * sin(ph) * sin(th),
                       ctr[2] + r0 * cos(th)};
  int c[3];
  bool coinc = true;
  for (int a = 0; a < 3; ++a) {
    c[a] = cell_of(B, a, X[a]);
    if (!(fabs(X[a] - coord(B, a, c[a])) <= 3 * kEps)) coinc = false;
  }
  double v;
  if (coinc) {
    v = dU_at(B, ctr, c[0], c[1], c[2]);
  } else {
    for (int a = 0; a < 3; ++a) c[a] = c[a] == 0 ? 1 : c[a];
    const double fx = (X[0] - coord(B, 0, c[0] - 1)) * B.idx[0];
    const double fy = (X[1] - coord(B, 1, c[1] - 1)) * B.idx[1];
    const double fz = (X[2] - coord(B, 2, c[2] - 1)) * B.idx[2];
    const double gx = 1 - fx, gy = 1 - fy, gz = 1 - fz;
    const int I = c[0], J = c[1], K = c[2];
    v = ((dU_at(B, ctr, I - 1, J - 1, K - 1) * gx + dU_at(B, ctr, I, J - 1, K - 1) * fx) * gy +
         (dU_at(B, ctr, I - 1, J, K - 1) * gx + dU_at(B, ctr, I, J, K - 1) * fx) * fy) * gz +
        ((dU_at(B, ctr, I - 1, J - 1, K) * gx + dU_at(B, ctr, I, J - 1, K) * fx) * gy +
         (dU_at(B, ctr, I - 1, J, K) * gx + dU_at(B, ctr, I, J, K) * fx) * fy) * fz;
  }
  sph[(long long)box * nt * nt + q] = v * sin(th);
}

// Lagrange interpolation of one field at arbitrary points (one box per
// point), interpolation.cc:189-303
__global__ __launch_bounds__(64) void k_interp_points(const DevBox *__restrict__ boxes, int field,
                                                       const int *__restrict__ pbox,
                                                       const double *__restrict__ pxyz, int npts,
                                                       int ord, double *__restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  const DevBox &B = boxes[pbox[q]];
  const double *__restrict__ f = B.f[field];
  const double X[3] = {pxyz[3 * q], pxyz[3 * q + 1], pxyz[3 * q + 2]};
  int c[3], s[3];
  bool coinc = true;
  for (int a = 0; a < 3; ++a) {
    c[a] = cell_of(B, a, X[a]);
    if (!(fabs(X[a] - coord(B, a, c[a])) <= 3 * kEps)) coinc = false;
  }
  if (coinc) {
    out[q] = f[pidx(B, c[0], c[1], c[2])];
    return;
  }
  // a box with fewer than ord + 1 points per axis takes the widest stencil
  // it holds (interpolation.cc:221-250 would read outside the box there)
  ord = min(ord, min(B.n[0], min(B.n[1], B.n[2])) - 1);
  double w[3][10];
  for (int a = 0; a < 3; ++a) {
    int s0 = (ord % 2 == 0) ? c[a] - ord / 2 : c[a] - (ord + 1) / 2;
    if (s0 < 0) s0 = 0;
    else if (s0 + ord >= B.n[a]) s0 = B.n[a] - ord - 1;
    s[a] = s0;
    for (int m = 0; m <= ord; ++m) {
      const double xn = coord(B, a, s0 + m);
      double lp = 1.0;
      for (int l = s0; l < s0 + m; ++l) lp *= (X[a] - coord(B, a, l)) / (xn - coord(B, a, l));
      for (int l = s0 + m + 1; l <= s0 + ord; ++l) lp *= (X[a] - coord(B, a, l)) / (xn - coord(B, a, l));
      w[a][m] = lp;
    }
  }
  double acc = 0.0;
  for (int I = 0; I <= ord; ++I) {
    double uj = 0.0;
    for (int J = 0; J <= ord; ++J) {
      double uk = 0.0;
      for (int K = 0; K <= ord; ++K) uk += f[pidx(B, s[0] + I, s[1] + J, s[2] + K)] * w[2][K];
      uj += uk * w[1][J];
    }
    acc += uj * w[0][I];
  }
  out[q] = acc;
}

__global__ void k_zero_u64(unsigned long long *p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// clear item blockIdx.x: its own word count, never another member's
__global__ void k_zero_lists(const ZeroItem *items) {
  const ZeroItem it = items[blockIdx.x];
  for (long long i = threadIdx.x; i < it.count; i += blockDim.x) it.p[i] = 0ull;
}

inline unsigned ceil_div(long long a, long long b) { return unsigned((a + b - 1) / b); }

template <int ORD, int EQ>
void smooth_impl(hipStream_t s, const RowBlock *blocks, int nb, int per_row, int colour,
                 int threads, hipEvent_t e0, hipEvent_t e1, const DevBox *box1) {
  if (nb == 0) return;
  if (e0)  // timed (never under graph capture): the runtime stamps e0, e1
    hipExtLaunchKernelGGL((k_smooth<ORD, EQ>), dim3(nb), dim3(threads), 0, s, e0, e1, 0, blocks,
                          unsigned(per_row), 1.0f / float(per_row), colour, box1);
  else
    hipLaunchKernelGGL((k_smooth<ORD, EQ>), dim3(nb), dim3(threads), 0, s, blocks,
                       unsigned(per_row), 1.0f / float(per_row), colour, box1);
}

template <int ORD, int EQ>
void residual_impl(hipStream_t s, const RowBlock *blocks, int nb, int per_row, int mode) {
  if (nb == 0) return;
  hipLaunchKernelGGL((k_residual<ORD, EQ>), dim3(nb), dim3(kRowThreads), 0, s, blocks,
                     unsigned(per_row), 1.0f / float(per_row), mode);
}

template <int ORD, int EQ>
void defect32_impl(hipStream_t s, const RowBlock *blocks, int nb, int per_row) {
  if (nb == 0) return;
  hipLaunchKernelGGL((k_defect32<ORD, EQ>), dim3(nb), dim3(kRowThreads), 0, s, blocks,
                     unsigned(per_row), 1.0f / float(per_row));
}

template <int ORD, int EQ>
void smooth32_impl(hipStream_t s, const RowBlock *blocks, int nb, int per_row, int colour,
                   int threads, hipEvent_t e0, hipEvent_t e1) {
  if (nb == 0) return;
  if (e0)
    hipExtLaunchKernelGGL((k_smooth32<ORD>), dim3(nb), dim3(threads), 0, s, e0, e1, 0, blocks,
                          unsigned(per_row), 1.0f / float(per_row), colour);
  else
    hipLaunchKernelGGL((k_smooth32<ORD>), dim3(nb), dim3(threads), 0, s, blocks,
                       unsigned(per_row), 1.0f / float(per_row), colour);
}

template <int ORD, int EQ>
void update32_impl(hipStream_t s, const RowBlock *blocks, int nb, int per_row) {
  if (nb == 0) return;
  hipLaunchKernelGGL((k_update32<ORD>), dim3(nb), dim3(kRowThreads), 0, s, blocks,
                     unsigned(per_row), 1.0f / float(per_row));
}

template <int ORD, int EQ>
void restrict_defect_impl(hipStream_t s, const PtBlock *blocks, int nb) {
  if (nb == 0) return;
  hipLaunchKernelGGL((k_restrict_defect<ORD, EQ>), dim3(nb), dim3(kPtThreads), 0, s, blocks);
}

constexpr int kLdsLimit = 160 * 1024;

}  // namespace

// ------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------

#define OLL_DISPATCH_ORDER_EQ(ORDER, E, KERNEL, ...)                  \
  switch (ORDER) {                                                    \
    case 2: KERNEL<2, E>(__VA_ARGS__); break;                         \
    case 4: KERNEL<4, E>(__VA_ARGS__); break;                         \
    case 6: KERNEL<6, E>(__VA_ARGS__); break;                         \
    default: KERNEL<8, E>(__VA_ARGS__); break;                        \
  }
#define OLL_DISPATCH_ORDER(ORDER, EQ, KERNEL, ...)                    \
  do {                                                                \
    switch (EQ) {                                                     \
      case 1: OLL_DISPATCH_ORDER_EQ(ORDER, 1, KERNEL, __VA_ARGS__) break; \
      case 2: OLL_DISPATCH_ORDER_EQ(ORDER, 2, KERNEL, __VA_ARGS__) break; \
      case 3: OLL_DISPATCH_ORDER_EQ(ORDER, 3, KERNEL, __VA_ARGS__) break; \
      case kEqPuncturesEm: OLL_DISPATCH_ORDER_EQ(ORDER, kEqPuncturesEm, KERNEL, __VA_ARGS__) break; \
      default: OLL_DISPATCH_ORDER_EQ(ORDER, 0, KERNEL, __VA_ARGS__) break; \
    }                                                                 \
  } while (0)

void launch_coefficients_box(hipStream_t s, const DevBox &B, const PunctureArgs &pa,
                             double robin_a) {
  const long long np = (long long)B.n[0] * B.n[1] * B.n[2];
  hipLaunchKernelGGL(k_coefficients, dim3(ceil_div(np, 256)), dim3(256), 0, s, B, pa, robin_a);
}

void launch_approx_box(hipStream_t s, const DevBox &B, const PunctureArgs &pa) {
  const long long np = (long long)B.n[0] * B.n[1] * B.n[2];
  hipLaunchKernelGGL(k_approx_u, dim3(ceil_div(np, 256)), dim3(256), 0, s, B, pa);
}

// threads per colour-pass workgroup, [mixed][order / 2 - 1]
static int g_sweep_threads[2][4] = {{256, 256, 256, 256}, {256, 256, 256, 256}};

int sweep_threads(int mixed, int order) {
  return g_sweep_threads[mixed ? 1 : 0][std::min(3, std::max(0, order / 2 - 1))];
}

bool set_sweep_threads(int mixed, int order, int threads) {
  if (!(order == 2 || order == 4 || order == 6 || order == 8)) return false;
  if (threads < 64 || threads > kMaxRowThreads || threads % 64 != 0) return false;
  g_sweep_threads[mixed ? 1 : 0][order / 2 - 1] = threads;
  return true;
}

void launch_smooth(hipStream_t s, const RowBlock *blocks, int nblocks, int per_row, int order,
                   int eq, int colour, int threads, hipEvent_t e0, hipEvent_t e1,
                   const DevBox *box1) {
  OLL_DISPATCH_ORDER(order, eq, smooth_impl, s, blocks, nblocks, per_row, colour, threads, e0, e1,
                     box1);
}

// tile shape of the fused passes, [order / 2 - 1]
static FusedShape g_fused_shape[4] = {{8, 8, 256}, {8, 8, 256}, {8, 8, 256}, {8, 8, 256}};

FusedShape fused_shape(int order) { return g_fused_shape[std::min(3, std::max(0, order / 2 - 1))]; }

bool set_fused_shape(int order, int ty, int tz, int threads) {
  if (!(order == 2 || order == 4 || order == 6 || order == 8)) return false;
  if (ty < 1 || tz < 1 || ty > 64 || tz > 64) return false;
  if (threads < 64 || threads > kFusedThreads || threads % 64 != 0) return false;
  g_fused_shape[order / 2 - 1] = FusedShape{ty, tz, threads};
  return true;
}

template <int ORD, int EQ, int Q, int PAIR>
void fused_impl(hipStream_t s, const TileBlock *blocks, int nb, int parity, int threads,
                hipEvent_t e0, hipEvent_t e1) {
  if (nb == 0) return;
  if (e0)  // timed (never under graph capture): the runtime stamps e0, e1
    hipExtLaunchKernelGGL((k_smooth_fused<ORD, EQ, Q, PAIR>), dim3(nb), dim3(threads), 0, s, e0,
                          e1, 0, blocks, parity);
  else
    hipLaunchKernelGGL((k_smooth_fused<ORD, EQ, Q, PAIR>), dim3(nb), dim3(threads), 0, s, blocks,
                       parity);
}

template <int ORD, int EQ>
void fused_order(hipStream_t s, const TileBlock *blocks, int nb, int q, int pair, int parity,
                 int threads, hipEvent_t e0, hipEvent_t e1) {
  constexpr int NC = ORD / 2 + 1;
  // pairs (0, 1), (2, 3); a single pass for the last colour when NC is odd
  if (pair && q == 0) return fused_impl<ORD, EQ, 0, 1>(s, blocks, nb, parity, threads, e0, e1);
  if constexpr (NC >= 4)
    if (pair && q == 2) return fused_impl<ORD, EQ, 2, 1>(s, blocks, nb, parity, threads, e0, e1);
  if constexpr (NC % 2 == 1)
    if (!pair && q == NC - 1)
      return fused_impl<ORD, EQ, NC - 1, 0>(s, blocks, nb, parity, threads, e0, e1);
}

void launch_smooth_fused(hipStream_t s, const TileBlock *blocks, int nblocks, int order, int eq,
                         int q, int pair, int parity, int threads, hipEvent_t e0, hipEvent_t e1) {
  switch (eq) {
    case 1: OLL_DISPATCH_ORDER_EQ(order, 1, fused_order, s, blocks, nblocks, q, pair, parity, threads, e0, e1) break;
    case 2: OLL_DISPATCH_ORDER_EQ(order, 2, fused_order, s, blocks, nblocks, q, pair, parity, threads, e0, e1) break;
    case 3: OLL_DISPATCH_ORDER_EQ(order, 3, fused_order, s, blocks, nblocks, q, pair, parity, threads, e0, e1) break;
    default: OLL_DISPATCH_ORDER_EQ(order, 0, fused_order, s, blocks, nblocks, q, pair, parity, threads, e0, e1) break;
  }
}

void launch_copy_shell(hipStream_t s, const PtBlock *blocks, int nblocks) {
  if (nblocks == 0) return;
  hipLaunchKernelGGL(k_copy_shell, dim3(nblocks), dim3(kPtThreads), 0, s, blocks);
}

void launch_residual(hipStream_t s, const RowBlock *blocks, int nblocks, int per_row, int order,
                     int eq, int mode) {
  OLL_DISPATCH_ORDER(order, eq, residual_impl, s, blocks, nblocks, per_row, mode);
}

void launch_defect32(hipStream_t s, const RowBlock *blocks, int nblocks, int per_row, int order,
                     int eq) {
  OLL_DISPATCH_ORDER(order, eq, defect32_impl, s, blocks, nblocks, per_row);
}

void launch_smooth32(hipStream_t s, const RowBlock *blocks, int nblocks, int per_row, int order,
                     int colour, int threads, hipEvent_t e0, hipEvent_t e1) {
  OLL_DISPATCH_ORDER(order, 0, smooth32_impl, s, blocks, nblocks, per_row, colour, threads, e0, e1);
}

void launch_update32(hipStream_t s, const RowBlock *blocks, int nblocks, int per_row, int order) {
  OLL_DISPATCH_ORDER(order, 0, update32_impl, s, blocks, nblocks, per_row);
}

void launch_robin(hipStream_t s, const RobinItem *items, int n, int total_cols, int total_edges,
                  int mode) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_robin_faces, dim3(ceil_div(total_cols, 256)), dim3(256), 0, s, items, n,
                     total_cols, mode);
  hipLaunchKernelGGL(k_robin_edges, dim3(ceil_div(total_edges, kRobinEdgeThreads)),
                     dim3(kRobinEdgeThreads), 0, s, items, n, total_edges, mode);
  hipLaunchKernelGGL(k_robin_corners, dim3(n), dim3(512), 0, s, items, mode);
}

void launch_restrict_u(hipStream_t s, const PtBlock *blocks, int nblocks) {
  if (nblocks == 0) return;
  hipLaunchKernelGGL(k_restrict_u, dim3(nblocks), dim3(kPtThreads), 0, s, blocks);
}

void launch_restrict_defect(hipStream_t s, const PtBlock *blocks, int nblocks, int order, int eq) {
  OLL_DISPATCH_ORDER(order, eq, restrict_defect_impl, s, blocks, nblocks);
}

void launch_filter_defect(hipStream_t s, const PtBlock *blocks, int nblocks, int runs) {
  if (nblocks == 0) return;
  if (runs)
    hipLaunchKernelGGL(k_filter_cols, dim3(nblocks), dim3(kPtThreads), 0, s, blocks);
  else
    hipLaunchKernelGGL(k_filter_defect, dim3(nblocks), dim3(kPtThreads), 0, s, blocks);
}

int filter_run_length() { return kFilterCx; }

void launch_corr_v(hipStream_t s, const PtBlock *blocks, int nblocks) {
  if (nblocks == 0) return;
  hipLaunchKernelGGL(k_corr_v, dim3(nblocks), dim3(kPtThreads), 0, s, blocks);
}

void launch_prolong(hipStream_t s, const PtBlock *blocks, int nblocks) {
  if (nblocks == 0) return;
  hipLaunchKernelGGL(k_prolong, dim3(nblocks), dim3(kPtThreads), 0, s, blocks);
}

template <int KP>
void ghost_impl(hipStream_t s, const PtBlock *blocks, int nblocks, int columns) {
  if (columns)
    hipLaunchKernelGGL(k_ghost_cols<KP>, dim3(nblocks), dim3(kPtThreads), 0, s, blocks);
  else
    hipLaunchKernelGGL(k_ghost_fill<KP>, dim3(nblocks), dim3(kPtThreads), 0, s, blocks);
}

void launch_ghost_fill(hipStream_t s, const PtBlock *blocks, int nblocks, int order, int columns) {
  if (nblocks == 0) return;
  switch (order) {
    case 2: ghost_impl<4>(s, blocks, nblocks, columns); break;
    case 4: ghost_impl<6>(s, blocks, nblocks, columns); break;
    case 6: ghost_impl<8>(s, blocks, nblocks, columns); break;
    default: ghost_impl<10>(s, blocks, nblocks, columns); break;
  }
}

// coarse-solve shape by box size: up to 2048 points of a colour, 1024
// threads with 2 points each and their coefficients in registers (512
// threads with 4 each for the scalar-field and EM equations, whose longer
// update would spill at 1024 threads); up to 4096 (the 21^3 coarsest box
// of N=161), 1024 threads with 4 each
bool launch_coarse_solve(hipStream_t s, const CoarseItem *items, int n, int max_pts, int max_int,
                         int eq, double eps, int max_iter) {
  const size_t lds = 2 * sizeof(double) * size_t(max_pts);  // max_pts: largest B.size
  // max_int: largest number of points of one colour slot list, ceil(ni/2) nj nk
  if (lds + 2048 > size_t(kLdsLimit) || max_int > 4096) return false;
  if (n == 0) return true;
  const bool small = max_int <= 2048;
  const bool wide = eq != 1 && eq != kEqPuncturesEm;
#define OLL_COARSE(E)                                                                              \
  if (small && wide)                                                                               \
    hipLaunchKernelGGL((k_coarse_solve<E, 2, 1024>), dim3(n), dim3(1024), lds, s, items, eps, max_iter); \
  else if (small)                                                                                  \
    hipLaunchKernelGGL((k_coarse_solve<E, 4, 512>), dim3(n), dim3(512), lds, s, items, eps, max_iter); \
  else                                                                                             \
    hipLaunchKernelGGL((k_coarse_solve<E, 4, 1024>), dim3(n), dim3(1024), lds, s, items, eps, max_iter);
  switch (eq) {
    case 1: OLL_COARSE(1) break;
    case 2: OLL_COARSE(2) break;
    case 3: OLL_COARSE(3) break;
    case kEqPuncturesEm: OLL_COARSE(kEqPuncturesEm) break;
    default: OLL_COARSE(0) break;
  }
#undef OLL_COARSE
  return true;
}

bool launch_transfer_fields(hipStream_t s, const PtBlock *blocks, int nblocks, int ff, int tf,
                            int nodes) {
  if (nblocks == 0) return true;
#define OLL_TF(KP) hipLaunchKernelGGL(k_transfer_fields<KP>, dim3(nblocks), dim3(kPtThreads), 0, s, blocks, ff, tf)
  switch (nodes) {
    case 2: OLL_TF(2); break;
    case 3: OLL_TF(3); break;
    case 4: OLL_TF(4); break;
    case 5: OLL_TF(5); break;
    case 6: OLL_TF(6); break;
    case 7: OLL_TF(7); break;
    case 8: OLL_TF(8); break;
    case 9: OLL_TF(9); break;
    case 10: OLL_TF(10); break;
    default: return false;
  }
#undef OLL_TF
  return true;
}

void launch_zero_lists(hipStream_t s, const ZeroItem *items, int n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_zero_lists, dim3(n), dim3(256), 0, s, items);
}

void kernels_init() {
  static bool done = false;
  if (done) return;
#define OLL_COARSE_FNS(E)                                          \
  reinterpret_cast<const void *>(&k_coarse_solve<E, 4, 512>),      \
      reinterpret_cast<const void *>(&k_coarse_solve<E, 2, 1024>), \
      reinterpret_cast<const void *>(&k_coarse_solve<E, 4, 1024>)
  const void *fns[] = {OLL_COARSE_FNS(0), OLL_COARSE_FNS(1), OLL_COARSE_FNS(2), OLL_COARSE_FNS(3),
                       OLL_COARSE_FNS(kEqPuncturesEm)};
#undef OLL_COARSE_FNS
  for (const void *f : fns)
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsLimit - 1024);
  done = true;
}

void launch_zero_u64(hipStream_t s, unsigned long long *p, long long n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_zero_u64, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n);
}

void launch_adm_sphere(hipStream_t s, const DevBox *boxes, int nbox, const double *centres,
                       double *sph, int nt) {
  hipLaunchKernelGGL(k_adm_sphere, dim3(ceil_div(nt * nt, 256), nbox), dim3(256), 0, s, boxes,
                     centres, sph, nt);
}

void launch_interp_points(hipStream_t s, const DevBox *boxes, int field, const int *pbox,
                          const double *pxyz, int npts, int ord, double *out) {
  hipLaunchKernelGGL(k_interp_points, dim3(ceil_div(npts, 64)), dim3(64), 0, s, boxes, field, pbox,
                     pxyz, npts, ord, out);
}

}  // namespace oll
