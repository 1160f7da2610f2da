// CDNA4 (gfx950) kernels of the FAS puncture solver.
//
// All stencils here are axis stars of radius R = order/2, so the points are
// split into R+1 colours (i+j+k) mod (R+1): no point reads another point of
// its own colour, and one launch updates one colour of every box of a level
// in parallel.  Arithmetic is float64 throughout.
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
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdio>

#include "kernels.hpp"

namespace oll {

namespace {

constexpr double kEps = DBL_EPSILON;

// ------------------------------------------------------------------
// point operators
// ------------------------------------------------------------------

template <int ORD>
__device__ __forceinline__ double lap(const double *__restrict__ u, long long p, long long sy,
                                      long long sz, double ix2, double iy2, double iz2) {
  const double u0 = u[p];
  if constexpr (ORD == 2) {
    return (u[p + 1] - 2.0 * u0 + u[p - 1]) * ix2 + (u[p + sy] - 2.0 * u0 + u[p - sy]) * iy2 +
           (u[p + sz] - 2.0 * u0 + u[p - sz]) * iz2;
  } else if constexpr (ORD == 4) {
    const double ax = -(u[p + 2] + u[p - 2]) + 16.0 * (u[p + 1] + u[p - 1]) - 30.0 * u0;
    const double ay = -(u[p + 2 * sy] + u[p - 2 * sy]) + 16.0 * (u[p + sy] + u[p - sy]) - 30.0 * u0;
    const double az = -(u[p + 2 * sz] + u[p - 2 * sz]) + 16.0 * (u[p + sz] + u[p - sz]) - 30.0 * u0;
    return (ax * ix2 + ay * iy2 + az * iz2) * (1.0 / 12.0);
  } else if constexpr (ORD == 6) {
    const double ax = 2.0 * (u[p + 3] + u[p - 3]) - 27.0 * (u[p + 2] + u[p - 2]) +
                      270.0 * (u[p + 1] + u[p - 1]) - 490.0 * u0;
    const double ay = 2.0 * (u[p + 3 * sy] + u[p - 3 * sy]) - 27.0 * (u[p + 2 * sy] + u[p - 2 * sy]) +
                      270.0 * (u[p + sy] + u[p - sy]) - 490.0 * u0;
    const double az = 2.0 * (u[p + 3 * sz] + u[p - 3 * sz]) - 27.0 * (u[p + 2 * sz] + u[p - 2 * sz]) +
                      270.0 * (u[p + sz] + u[p - sz]) - 490.0 * u0;
    return (ax * ix2 + ay * iy2 + az * iz2) * (1.0 / 180.0);
  } else {
    const double ax = -9.0 * (u[p + 4] + u[p - 4]) + 128.0 * (u[p + 3] + u[p - 3]) -
                      1008.0 * (u[p + 2] + u[p - 2]) + 8064.0 * (u[p + 1] + u[p - 1]) - 14350.0 * u0;
    const double ay = -9.0 * (u[p + 4 * sy] + u[p - 4 * sy]) + 128.0 * (u[p + 3 * sy] + u[p - 3 * sy]) -
                      1008.0 * (u[p + 2 * sy] + u[p - 2 * sy]) + 8064.0 * (u[p + sy] + u[p - sy]) -
                      14350.0 * u0;
    const double az = -9.0 * (u[p + 4 * sz] + u[p - 4 * sz]) + 128.0 * (u[p + 3 * sz] + u[p - 3 * sz]) -
                      1008.0 * (u[p + 2 * sz] + u[p - 2 * sz]) + 8064.0 * (u[p + sz] + u[p - sz]) -
                      14350.0 * u0;
    return (ax * ix2 + ay * iy2 + az * iz2) * (1.0 / 5040.0);
  }
}

// N(u) = Lap u + b (c+u)^-7 [+ d (c+u) + e (c+u)^5] and dN/du_p
template <int EQ>
__device__ __forceinline__ void nonlinear(double lapu, double dulap, double b, double cu,
                                          double dd, double ee, double &N, double &dN) {
  const double inv = 1.0 / cu;
  const double inv2 = inv * inv;
  const double inv4 = inv2 * inv2;
  const double inv7 = inv4 * inv2 * inv;
  const double inv8 = inv4 * inv4;
  N = lapu + b * inv7;
  dN = dulap - 7.0 * b * inv8;
  if constexpr (EQ == 1) {
    const double c2 = cu * cu;
    const double c4 = c2 * c2;
    N += dd * cu + ee * (c4 * cu);
    dN += dd + 5.0 * ee * c4;
  }
}

__device__ __forceinline__ double warp_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// block max of v; result valid in thread 0; blockDim.x multiple of 64
__device__ __forceinline__ double block_max(double v, double *sh) {
  v = warp_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) r = fmax(r, sh[w]);
  }
  return r;
}

// locate the box of a flat interior row index within a level
__device__ __forceinline__ int find_box_row(const LevelView &lv, int row) {
  int b = 0;
  while (b + 1 < lv.nbox && row >= lv.row0[b + 1]) ++b;
  return b;
}

__device__ __forceinline__ int find_box_pt(const LevelView &lv, long long pt) {
  int b = 0;
  while (b + 1 < lv.nbox && pt >= lv.pt0[b + 1]) ++b;
  return b;
}

// Norm slots: every logical norm is kNormSub consecutive u64 words; blocks
// spread their atomics over them and the host takes the max.  A block only
// issues its atomic when it raises the value it reads.
constexpr int kNormSub = 8;

__device__ __forceinline__ void norm_update(unsigned long long *slot, double v) {
  unsigned long long *s = slot + (blockIdx.x & (kNormSub - 1));
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  if (bits > __hip_atomic_load(s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(s, bits);
}

// ------------------------------------------------------------------
// smoother: one colour of one Newton-Gauss-Seidel sweep on every box of a
// level.  Thread g owns colour point t = g mod per_row of interior row
// g / per_row: i = i_first(row) + t (R+1).
// ------------------------------------------------------------------
// exact g / d for g < 2^24 via a float reciprocal and one correction step
__device__ __forceinline__ unsigned fast_div(unsigned g, unsigned d, float inv_d) {
  unsigned q = unsigned(float(g) * inv_d);
  if (q * d > g) --q;
  else if ((q + 1) * d <= g) ++q;
  return q;
}

// wave-level max, then one (read-checked) atomic per wave
__device__ __forceinline__ void wave_norm_update(unsigned long long *slot, double v) {
  v = warp_max(v);
  if ((threadIdx.x & 63) == 0 && v > 0.0) norm_update(slot, v);
}

template <int ORD, int EQ>
__global__ __launch_bounds__(256) void k_smooth(LevelView lv, const DevBox *__restrict__ boxes,
                                                 Fields F, int colour, unsigned per_row,
                                                 float inv_per_row,
                                                 unsigned long long *__restrict__ norm) {
  constexpr int R = ORD / 2;
  constexpr int NC = R + 1;
  const unsigned g = blockIdx.x * 256u + threadIdx.x;
  const unsigned row = fast_div(g, per_row, inv_per_row);
  const int t = int(g - row * per_row);
  double rmax = 0.0;
  if (row < unsigned(lv.row0[lv.nbox])) {
    const int bi = find_box_row(lv, int(row));
    const DevBox &B = boxes[lv.box[bi]];
    const unsigned r = row - unsigned(lv.row0[bi]);
    const unsigned nyi = unsigned(B.n[1] - 2 * R);
    const unsigned kq = fast_div(r, nyi, 1.0f / float(nyi));
    const int k = R + int(kq);
    const int j = R + int(r - kq * nyi);
    const int first = R + ((colour - j - k - R) % NC + 2 * NC) % NC;
    const int i = first + t * NC;
    if (i < B.n[0] - R) {
      const long long sy = B.n[0], sz = (long long)B.n[0] * B.n[1];
      const long long p = B.off + i + sy * j + sz * k;
      double *__restrict__ u = F.f[F_U];
      const double lu = lap<ORD>(u, p, sy, sz, B.idx2[0], B.idx2[1], B.idx2[2]);
      const double cu = F.f[F_C][p] + u[p];
      double N, dN;
      nonlinear<EQ>(lu, B.dulap, F.f[F_B][p], cu, EQ ? F.f[F_D][p] : 0.0,
                    EQ ? F.f[F_E][p] : 0.0, N, dN);
      const double res = N - F.f[F_RHO][p];
      const double un = u[p] - res / dN;
      u[p] = un;
      rmax = fabs(res);
      if (!(un == un)) rmax = __longlong_as_double(0x7ff0000000000000LL);  // inf flags NaN
    }
  }
  wave_norm_update(norm, rmax);
}

// residual on the interior: r = N(u) (write_defect = 0) or w = rho - N(u)
template <int ORD, int EQ>
__global__ __launch_bounds__(256) void k_residual(LevelView lv, const DevBox *__restrict__ boxes,
                                                   Fields F, int write_defect,
                                                   unsigned long long *__restrict__ norm) {
  constexpr int R = ORD / 2;
  __shared__ double sh[4];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int row = blockIdx.x * 4 + ty;
  double rmax = 0.0;
  if (row < lv.row0[lv.nbox]) {
    const int bi = find_box_row(lv, row);
    const DevBox &B = boxes[lv.box[bi]];
    const int r = row - lv.row0[bi];
    const int nyi = B.n[1] - 2 * R;
    const int j = R + r % nyi, k = R + r / nyi;
    const int i = R + blockIdx.y * 64 + tx;
    if (i < B.n[0] - R) {
      const long long sy = B.n[0], sz = (long long)B.n[0] * B.n[1];
      const long long p = B.off + i + sy * j + sz * k;
      const double *__restrict__ u = F.f[F_U];
      const double lu = lap<ORD>(u, p, sy, sz, B.idx2[0], B.idx2[1], B.idx2[2]);
      double N, dN;
      nonlinear<EQ>(lu, B.dulap, F.f[F_B][p], F.f[F_C][p] + u[p], EQ ? F.f[F_D][p] : 0.0,
                    EQ ? F.f[F_E][p] : 0.0, N, dN);
      if (write_defect)
        F.f[F_W][p] = F.f[F_RHO][p] - N;
      else
        F.f[F_R][p] = N;
      rmax = fabs(N);
    }
  }
  const double m = block_max(rmax, sh);
  if (norm && threadIdx.x == 0 && m > 0.0) norm_update(norm, m);
}

// ------------------------------------------------------------------
// Robin boundary r^2/n sum_d (1/x_d) d_d u + u = S on the ghost shell
// ------------------------------------------------------------------

// first derivative along an axis at index i (n points, stride s):
// dx2 (diff_2nd.cc:237-335) for order 2, dx4 (diff_4th.cc:6-89) otherwise.
// coef = coefficient of u at the point itself.
__device__ __forceinline__ double deriv(const double *__restrict__ u, long long p, long long s,
                                        int i, int n, int order, double id, double &coef) {
  if (order == 2) {
    if (i == 0) {
      coef = -3.0 * 0.5 * id;
      return (-3.0 * u[p] + 4.0 * u[p + s] - u[p + 2 * s]) * 0.5 * id;
    } else if (i == n - 1) {
      coef = 3.0 * 0.5 * id;
      return (3.0 * u[p] - 4.0 * u[p - s] + u[p - 2 * s]) * 0.5 * id;
    }
    coef = 0.0;
    return (u[p + s] - u[p - s]) * 0.5 * id;
  }
  const int nm = (n % 2 == 0) ? n / 2 : (n - 1) / 2;
  if (i <= nm) {
    coef = -25.0 * id / 12.0;
    return -(25.0 * u[p] - 48.0 * u[p + s] + 36.0 * u[p + 2 * s] - 16.0 * u[p + 3 * s] +
             3.0 * u[p + 4 * s]) * id / 12.0;
  }
  coef = 25.0 * id / 12.0;
  return (25.0 * u[p] - 48.0 * u[p - s] + 36.0 * u[p - 2 * s] - 16.0 * u[p - 3 * s] +
          3.0 * u[p - 4 * s]) * id / 12.0;
}

// coordinate of index i, without contraction so it equals the host value
__device__ __forceinline__ double coord(const DevBox &B, int a, int i) {
  return __dadd_rn(B.lo[a], __dmul_rn((double)i, B.d[a]));
}

// Box-local views of the fields a Robin stage touches (point (0,0,0) of the
// box at index 0).  u may live in global memory or in LDS.
struct BoxData {
  double *u;
  const double *rho;
  double *r;
  double *w;
};

__device__ __forceinline__ BoxData box_data(const DevBox &B, Fields F) {
  BoxData D;
  D.u = F.f[F_U] + B.off;
  D.rho = F.f[F_RHO] + B.off;
  D.r = F.f[F_R] + B.off;
  D.w = F.f[F_W] + B.off;
  return D;
}

// Robin update / operator at one shell point; `mask` has bit a set for
// each axis a whose index lies in the ghost shell.
__device__ __forceinline__ void robin_point(const DevBox &B, BoxData D, int i, int j, int k,
                                            int mask, int mode, double ipn) {
  const long long sy = B.n[0], sz = (long long)B.n[0] * B.n[1];
  const long long p = i + sy * j + sz * k;
  double *u = D.u;
  const int ijk[3] = {i, j, k};
  const long long st[3] = {1, sy, sz};
  double X[3];
  for (int a = 0; a < 3; ++a) X[a] = coord(B, a, ijk[a]);
  const double r2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
  double sF = 0.0, sdF = 0.0;
  for (int a = 0; a < 3; ++a) {
    if (!(mask & (1 << a))) continue;
    double cf;
    const double dv = deriv(u, p, st[a], ijk[a], B.n[a], B.order, B.idx[a], cf);
    sF += dv / X[a];
    sdF += cf / X[a];
  }
  const double op = r2 * ipn * sF + u[p];
  if (mode == ROBIN_UPDATE) {
    const double Fv = op - D.rho[p];
    const double dF = r2 * ipn * sdF + 1.0;
    u[p] = u[p] - Fv / dF;
  } else if (mode == ROBIN_OPERATOR_R) {
    D.r[p] = op;
  } else {
    D.w[p] = D.rho[p] - op;
  }
}

// face column t (0 .. 2(nj nk + ni nk + ni nj) - 1) of a box: all R layers,
// inner -> outer
__device__ __forceinline__ void robin_face_column(const DevBox &B, BoxData D, int t, int mode,
                                                  double ipn) {
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
    robin_point(B, D, i, j, k, 1 << axis, mode, ipn);
  }
}

// faces: one thread per (face, column); layers inner -> outer
__global__ __launch_bounds__(256) void k_robin_faces(const DevBox *__restrict__ boxes, int box,
                                                      Fields F, int mode, double ipn) {
  const DevBox &B = boxes[box];
  robin_face_column(B, box_data(B, F), blockIdx.x * blockDim.x + threadIdx.x, mode, ipn);
}

// ghost index of position q (0..R-1) on side `hi` of an axis, in the
// reference's iteration order (domain.cc:575-1388): low sides run R-1 -> 0,
// high sides n-R -> n-1
__device__ __forceinline__ int shell_index(int q, int hi, int n, int R) {
  return hi ? n - R + q : R - 1 - q;
}

// edges then corners, one block per box (faces must be done)
__device__ void robin_edges_corners(const DevBox &B, BoxData D, int mode, double ipn) {
  const int R = B.R;
  const int nint[3] = {B.n[0] - 2 * R, B.n[1] - 2 * R, B.n[2] - 2 * R};
  // edges XY(4) run along z, XZ(4) along y, YZ(4) along x; within an edge
  // the cross-section is visited outer axis first (x before y for XY, x
  // before z for XZ, z before y for YZ), as Fix_Points orders them.
  const int tot = 4 * (nint[2] + nint[1] + nint[0]);
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
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
        robin_point(B, D, i, j, k, mask, mode, ipn);
      }
  }
  __syncthreads();
  // corners: k outer, j, i inner (domain.cc:592-695)
  if (threadIdx.x < 8) {
    const int c = threadIdx.x;
    // reference corner order XiYiZi, XiYiZf, XfYiZi, XfYiZf, XiYfZi, ...
    const int hz = c & 1, hx = (c >> 1) & 1, hy = (c >> 2) & 1;
    for (int qk = 0; qk < R; ++qk)
      for (int qj = 0; qj < R; ++qj)
        for (int qi = 0; qi < R; ++qi)
          robin_point(B, D, shell_index(qi, hx, B.n[0], R), shell_index(qj, hy, B.n[1], R),
                      shell_index(qk, hz, B.n[2], R), 7, mode, ipn);
  }
}

__global__ __launch_bounds__(1024) void k_robin_edges_corners(const DevBox *__restrict__ boxes,
                                                               int box, Fields F, int mode,
                                                               double ipn) {
  robin_edges_corners(boxes[box], box_data(boxes[box], F), mode, ipn);
}

// ------------------------------------------------------------------
// transfers
// ------------------------------------------------------------------

struct AxisRef {
  const int *start, *coinc;
  const unsigned char *flags;
  const double *w;
  int k;
};

__device__ __forceinline__ AxisRef axis_ref(const XferView &xv, const DevAxis &a) {
  AxisRef r;
  r.start = xv.ipool + a.start_off;
  r.coinc = xv.ipool + a.coinc_off;
  r.flags = xv.bpool + a.flags_off;
  r.w = xv.dpool + a.w_off;
  r.k = a.k;
  return r;
}

__device__ __forceinline__ bool in_region(const DevBox &B, int i, int j, int k, int region) {
  if (region == REG_ALL) return true;
  const int R = B.R;
  const bool inside = i >= R && i < B.n[0] - R && j >= R && j < B.n[1] - R && k >= R &&
                      k < B.n[2] - R;
  return region == REG_INSIDE ? inside : !inside;
}

// tensor interpolation of field f of box S (reference accumulation order,
// interpolation.cc:257-296)
__device__ double tensor_interp(const DevBox &S, const double *__restrict__ f, const AxisRef ax[3],
                                const int t[3]) {
  const int kx = ax[0].k;
  const int s0 = ax[0].start[t[0]], s1 = ax[1].start[t[1]], s2 = ax[2].start[t[2]];
  const double *wx = ax[0].w + (long long)t[0] * kx;
  const double *wy = ax[1].w + (long long)t[1] * ax[1].k;
  const double *wz = ax[2].w + (long long)t[2] * ax[2].k;
  const long long sy = S.n[0], sz = (long long)S.n[0] * S.n[1];
  double acc = 0.0;
  for (int I = 0; I < kx; ++I) {
    double uj = 0.0;
    for (int J = 0; J < ax[1].k; ++J) {
      double uk = 0.0;
      const long long base = S.off + (s0 + I) + sy * (s1 + J) + sz * (long long)s2;
      for (int K = 0; K < ax[2].k; ++K) uk += f[base + sz * K] * wz[K];
      uj += uk * wy[J];
    }
    acc += uj * wx[I];
  }
  return acc;
}

// trilinear interpolation of (u - ww) of box S (interpolation.cc:125-182)
__device__ __forceinline__ double trilinear_corr(const DevBox &S, const double *__restrict__ u,
                                                 const double *__restrict__ ww, const AxisRef ax[3],
                                                 const int t[3]) {
  const long long sy = S.n[0], sz = (long long)S.n[0] * S.n[1];
  const long long b = S.off + ax[0].start[t[0]] + sy * ax[1].start[t[1]] + sz * ax[2].start[t[2]];
  const double dxi = ax[0].w[2 * t[0]], dxf = ax[0].w[2 * t[0] + 1];
  const double dyi = ax[1].w[2 * t[1]], dyf = ax[1].w[2 * t[1] + 1];
  const double dzi = ax[2].w[2 * t[2]], dzf = ax[2].w[2 * t[2] + 1];
#define V(o) (u[b + (o)] - ww[b + (o)])
  return ((V(0) * dxi + V(1) * dxf) * dyi + (V(sy) * dxi + V(sy + 1) * dxf) * dyf) * dzi +
         ((V(sz) * dxi + V(sz + 1) * dxf) * dyi + (V(sz + sy) * dxi + V(sz + sy + 1) * dxf) * dyf) *
             dzf;
#undef V
}

// u restriction: coarse INSIDE points inside a fine box's non-ghost part
__global__ __launch_bounds__(256) void k_restrict_u(XferView xv, const DevBox *__restrict__ boxes,
                                                     Fields F) {
  const long long pt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= xv.tgt.pt0[xv.tgt.nbox]) return;
  const int bi = find_box_pt(xv.tgt, pt);
  const DevBox &T = boxes[xv.tgt.box[bi]];
  const long long q = pt - xv.tgt.pt0[bi];
  const int i = int(q % T.n[0]), j = int((q / T.n[0]) % T.n[1]), k = int(q / ((long long)T.n[0] * T.n[1]));
  if (!in_region(T, i, j, k, REG_INSIDE)) return;
  const int t[3] = {i, j, k};
  for (int x = xv.x0[bi + 1] - 1; x >= xv.x0[bi]; --x) {
    const DevXfer &X = xv.xfer[x];
    AxisRef ax[3] = {axis_ref(xv, X.ax[0]), axis_ref(xv, X.ax[1]), axis_ref(xv, X.ax[2])};
    if (!((ax[0].flags[i] & ax[1].flags[j] & ax[2].flags[k]) & 1)) continue;  // all=false: nb only
    const DevBox &S = boxes[X.source];
    double v;
    if ((ax[0].flags[i] & ax[1].flags[j] & ax[2].flags[k]) & 4) {
      v = F.f[F_U][S.off + ax[0].coinc[i] + (long long)S.n[0] * ax[1].coinc[j] +
                   (long long)S.n[0] * S.n[1] * ax[2].coinc[k]];
    } else {
      v = tensor_interp(S, F.f[F_U], ax, t);
    }
    F.f[F_U][T.off + q] = v;
    return;
  }
}

// Robin operator value at a shell point (no update)
__device__ __forceinline__ double robin_op_at(const DevBox &B, Fields F, int i, int j, int k,
                                              double ipn) {
  const int R = B.R;
  int mask = 0;
  if (i < R || i >= B.n[0] - R) mask |= 1;
  if (j < R || j >= B.n[1] - R) mask |= 2;
  if (k < R || k >= B.n[2] - R) mask |= 4;
  const long long sy = B.n[0], sz = (long long)B.n[0] * B.n[1];
  const long long p = B.off + i + sy * j + sz * k;
  const double *__restrict__ u = F.f[F_U];
  const int ijk[3] = {i, j, k};
  const long long st[3] = {1, sy, sz};
  double X[3];
  for (int a = 0; a < 3; ++a) X[a] = coord(B, a, ijk[a]);
  const double r2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
  double sF = 0.0;
  for (int a = 0; a < 3; ++a) {
    if (!(mask & (1 << a))) continue;
    double cf;
    sF += deriv(u, p, st[a], ijk[a], B.n[a], B.order, B.idx[a], cf) / X[a];
  }
  return r2 * ipn * sF + u[p];
}

// Coarse FAS source: W_c = restriction of the fine defect (full weighting
// where the coarse point coincides with a fine point whose 3x3x3
// neighbourhood is fine interior, injection / Lagrange-2 elsewhere), then
// rho_c = W_c + R_c with R_c = N(u_c) inside, Robin operator on the shell
// (multigrid levels).  Refinement levels overwrite rho_c only where the
// point is inside both boxes' non-ghost parts (ffunction.cc:2808-2849).
// Also saves ww_c = u_c on every coarse point.
template <int ORD, int EQ>
__global__ __launch_bounds__(256) void k_restrict_defect(XferView xv, const DevBox *__restrict__ boxes,
                                                          Fields F, int mg, double ipn) {
  const long long pt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= xv.tgt.pt0[xv.tgt.nbox]) return;
  const int bi = find_box_pt(xv.tgt, pt);
  const DevBox &T = boxes[xv.tgt.box[bi]];
  const long long q = pt - xv.tgt.pt0[bi];
  const int i = int(q % T.n[0]), j = int((q / T.n[0]) % T.n[1]), k = int(q / ((long long)T.n[0] * T.n[1]));
  const long long p = T.off + q;
  F.f[F_WW][p] = F.f[F_U][p];
  const bool inside = in_region(T, i, j, k, REG_INSIDE);
  if (!mg && !inside) return;
  const int t[3] = {i, j, k};
  bool covered = false;
  double wc = 0.0;
  for (int x = xv.x0[bi + 1] - 1; x >= xv.x0[bi]; --x) {
    const DevXfer &X = xv.xfer[x];
    AxisRef ax[3] = {axis_ref(xv, X.ax[0]), axis_ref(xv, X.ax[1]), axis_ref(xv, X.ax[2])};
    const unsigned f = ax[0].flags[i] & ax[1].flags[j] & ax[2].flags[k];
    if (!((f & 1) || (mg && (f & 2)))) continue;
    const DevBox &S = boxes[X.source];
    const double *__restrict__ w = F.f[F_W];
    if (f & 8) {
      const long long sy = S.n[0], sz = (long long)S.n[0] * S.n[1];
      const long long c = S.off + ax[0].coinc[i] + sy * ax[1].coinc[j] + sz * ax[2].coinc[k];
      double acc = 0.0;
      for (int dz = -1; dz <= 1; ++dz) {
        const double wz = dz == 0 ? 0.5 : 0.25;
        for (int dy = -1; dy <= 1; ++dy) {
          const double wy = dy == 0 ? 0.5 : 0.25;
          const long long r = c + sy * dy + sz * dz;
          acc += wz * wy * (0.25 * w[r - 1] + 0.5 * w[r] + 0.25 * w[r + 1]);
        }
      }
      wc = acc;
    } else if (f & 4) {
      wc = w[S.off + ax[0].coinc[i] + (long long)S.n[0] * ax[1].coinc[j] +
             (long long)S.n[0] * S.n[1] * ax[2].coinc[k]];
    } else {
      wc = tensor_interp(S, w, ax, t);
    }
    covered = true;
    break;
  }
  if (!covered && !mg) return;
  if (covered) F.f[F_W][p] = wc; else wc = F.f[F_W][p];
  double Rc;
  if (inside) {
    const long long sy = T.n[0], sz = (long long)T.n[0] * T.n[1];
    const double *__restrict__ u = F.f[F_U];
    double dN;
    nonlinear<EQ>(lap<ORD>(u, p, sy, sz, T.idx2[0], T.idx2[1], T.idx2[2]), T.dulap, F.f[F_B][p],
                  F.f[F_C][p] + u[p], EQ ? F.f[F_D][p] : 0.0, EQ ? F.f[F_E][p] : 0.0, Rc, dN);
  } else {
    Rc = robin_op_at(T, F, i, j, k, ipn);
  }
  F.f[F_RHO][p] = wc + Rc;
}

// prolongation of the coarse correction and u += v
__global__ __launch_bounds__(256) void k_prolong(XferView xv, const DevBox *__restrict__ boxes,
                                                  Fields F, int region) {
  const long long pt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= xv.tgt.pt0[xv.tgt.nbox]) return;
  const int bi = find_box_pt(xv.tgt, pt);
  const DevBox &T = boxes[xv.tgt.box[bi]];
  const long long q = pt - xv.tgt.pt0[bi];
  const int i = int(q % T.n[0]), j = int((q / T.n[0]) % T.n[1]), k = int(q / ((long long)T.n[0] * T.n[1]));
  if (!in_region(T, i, j, k, region)) return;
  const int t[3] = {i, j, k};
  for (int x = xv.x0[bi + 1] - 1; x >= xv.x0[bi]; --x) {
    const DevXfer &X = xv.xfer[x];
    AxisRef ax[3] = {axis_ref(xv, X.ax[0]), axis_ref(xv, X.ax[1]), axis_ref(xv, X.ax[2])};
    const unsigned f = ax[0].flags[i] & ax[1].flags[j] & ax[2].flags[k];
    if (!(f & 3)) continue;  // all = true
    const DevBox &S = boxes[X.source];
    double v;
    if (f & 4) {
      const long long c = S.off + ax[0].coinc[i] + (long long)S.n[0] * ax[1].coinc[j] +
                          (long long)S.n[0] * S.n[1] * ax[2].coinc[k];
      v = F.f[F_U][c] - F.f[F_WW][c];
    } else {
      v = trilinear_corr(S, F.f[F_U], F.f[F_WW], ax, t);
    }
    F.f[F_U][T.off + q] += v;
    return;
  }
}

// shell point q (0 .. n0 n1 n2 - (n0-2R)(n1-2R)(n2-2R) - 1) of a box:
// the R-thick z slabs first, then y slabs, then x slabs
__device__ __forceinline__ void shell_point(const DevBox &B, unsigned q, int &i, int &j, int &k) {
  const unsigned R = unsigned(B.R);
  const unsigned n0 = unsigned(B.n[0]), n1 = unsigned(B.n[1]), n2 = unsigned(B.n[2]);
  const unsigned zs = n0 * n1 * R;  // one z slab
  if (q < 2 * zs) {
    const unsigned s = q >= zs, r = q - s * zs;
    const unsigned plane = n0 * n1;
    const unsigned l = r / plane, rr = r - l * plane;
    j = int(rr / n0);
    i = int(rr - unsigned(j) * n0);
    k = s ? int(n2 - R + l) : int(l);
    return;
  }
  q -= 2 * zs;
  const unsigned ys = n0 * R * (n2 - 2 * R);
  if (q < 2 * ys) {
    const unsigned s = q >= ys, r = q - s * ys;
    const unsigned row = r / n0;
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
  const unsigned kk = col / (n1 - 2 * R);
  j = int(R + (col - kk * (n1 - 2 * R)));
  k = int(R + kk);
  i = s ? int(n0 - R + l) : int(l);
}

// Lagrange tensor interpolation that skips the nodes of coincident axes
// (their weights are exactly one unit vector)
__device__ double tensor_interp_sparse(const DevBox &S, const double *__restrict__ f,
                                       const AxisRef ax[3], const int t[3], unsigned fl0,
                                       unsigned fl1, unsigned fl2) {
  int s0[3], kk[3];
  const double *w[3];
  const unsigned fl[3] = {fl0, fl1, fl2};
  for (int a = 0; a < 3; ++a) {
    if (fl[a] & 4) {
      s0[a] = ax[a].coinc[t[a]];
      kk[a] = 1;
      w[a] = nullptr;
    } else {
      s0[a] = ax[a].start[t[a]];
      kk[a] = ax[a].k;
      w[a] = ax[a].w + (long long)t[a] * ax[a].k;
    }
  }
  const long long sy = S.n[0], sz = (long long)S.n[0] * S.n[1];
  double acc = 0.0;
  for (int I = 0; I < kk[0]; ++I) {
    double uj = 0.0;
    for (int J = 0; J < kk[1]; ++J) {
      double uk = 0.0;
      const long long base = S.off + (s0[0] + I) + sy * (s0[1] + J) + sz * (long long)s0[2];
      for (int K = 0; K < kk[2]; ++K) uk += f[base + sz * K] * (w[2] ? w[2][K] : 1.0);
      uj += uk * (w[1] ? w[1][J] : 1.0);
    }
    acc += uj * (w[0] ? w[0][I] : 1.0);
  }
  return acc;
}

// ghost shell of refinement boxes from the coarser level (Lagrange order+1);
// one thread per shell point
__global__ __launch_bounds__(256) void k_ghost_fill(XferView xv, const DevBox *__restrict__ boxes,
                                                     Fields F) {
  const long long pt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nb = xv.tgt.nbox;
  if (pt >= xv.tgt.shell0[nb]) return;
  int bi = 0;
  while (bi + 1 < nb && pt >= xv.tgt.shell0[bi + 1]) ++bi;
  const DevBox &T = boxes[xv.tgt.box[bi]];
  int i, j, k;
  shell_point(T, unsigned(pt - xv.tgt.shell0[bi]), i, j, k);
  const int t[3] = {i, j, k};
  for (int x = xv.x0[bi + 1] - 1; x >= xv.x0[bi]; --x) {
    const DevXfer &X = xv.xfer[x];
    AxisRef ax[3] = {axis_ref(xv, X.ax[0]), axis_ref(xv, X.ax[1]), axis_ref(xv, X.ax[2])};
    const unsigned f0 = ax[0].flags[i], f1 = ax[1].flags[j], f2 = ax[2].flags[k];
    const unsigned f = f0 & f1 & f2;
    if (!(f & 3)) continue;
    const DevBox &S = boxes[X.source];
    const double v = tensor_interp_sparse(S, F.f[F_U], ax, t, f0, f1, f2);
    F.f[F_U][T.off + i + (long long)T.n[0] * j + (long long)T.n[0] * T.n[1] * k] = v;
    return;
  }
}

// ------------------------------------------------------------------
// coarsest level: {Robin, one multi-colour sweep} until the sweep norm is
// <= eps or stops changing (elliptic.cc:1782-1816); one block, one launch.
// u lives in LDS for the whole loop.  Interior updates are over-relaxed by
// omega (omega = 1 is the reference's Gauss-Seidel-Newton); the loop's fixed
// point, the coarse solution, does not depend on omega.
// ------------------------------------------------------------------
constexpr int kCoarseOwn = 8;  // interior points per thread (1024 threads)

template <int EQ>
__global__ __launch_bounds__(1024) void k_coarse_solve(const DevBox *__restrict__ boxes, int box,
                                                        Fields F, double ipn, double eps,
                                                        double omega, int max_iter,
                                                        unsigned long long *__restrict__ norm,
                                                        int *__restrict__ iters) {
  extern __shared__ double lds[];
  __shared__ double sh[16];
  __shared__ double s_norm;
  const DevBox &B = boxes[box];
  const int R = B.R;  // 1: the coarsest level always uses order 2
  const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
  const int np = n0 * n1 * n2;
  double *su = lds;        // u
  double *srho = lds + np; // rho
  const int ni = n0 - 2 * R, nj = n1 - 2 * R, nk = n2 - 2 * R;
  const long long sy = n0, sz = (long long)n0 * n1;
  double *__restrict__ ug = F.f[F_U] + B.off;
  for (int q = threadIdx.x; q < np; q += blockDim.x) {
    su[q] = ug[q];
    srho[q] = F.f[F_RHO][B.off + q];
  }
  // owned interior points: t = threadIdx.x + m * blockDim.x
  const int nint = ni * nj * nk;
  int own_p[kCoarseOwn], own_c[kCoarseOwn];
  double own_b[kCoarseOwn], own_cc[kCoarseOwn], own_d[kCoarseOwn], own_e[kCoarseOwn];
#pragma unroll
  for (int m = 0; m < kCoarseOwn; ++m) {
    const int t = threadIdx.x + m * blockDim.x;
    own_p[m] = -1;
    own_c[m] = -1;
    own_b[m] = own_cc[m] = own_d[m] = own_e[m] = 0.0;
    if (t < nint) {
      const int i = R + t % ni, j = R + (t / ni) % nj, k = R + t / (ni * nj);
      const int p = i + n0 * (j + n1 * k);
      own_p[m] = p;
      own_c[m] = (i + j + k) % (R + 1);
      own_b[m] = F.f[F_B][B.off + p];
      own_cc[m] = F.f[F_C][B.off + p];
      if (EQ) {
        own_d[m] = F.f[F_D][B.off + p];
        own_e[m] = F.f[F_E][B.off + p];
      }
    }
  }
  __syncthreads();
  BoxData D;
  D.u = su;
  D.rho = srho;
  D.r = nullptr;
  D.w = nullptr;
  const int nface = 2 * (nj * nk + ni * nk + ni * nj);
  double prev = 1.0, cur = 1.0;
  int it = 0;
  do {
    prev = cur;
    for (int t = threadIdx.x; t < nface; t += blockDim.x) robin_face_column(B, D, t, ROBIN_UPDATE, ipn);
    __syncthreads();
    robin_edges_corners(B, D, ROBIN_UPDATE, ipn);
    __syncthreads();
    double m = 0.0;
    for (int col = 0; col <= R; ++col) {
#pragma unroll
      for (int q = 0; q < kCoarseOwn; ++q) {
        if (own_c[q] != col) continue;
        const long long p = own_p[q];
        const double lu = lap<2>(su, p, sy, sz, B.idx2[0], B.idx2[1], B.idx2[2]);
        double N, dN;
        nonlinear<EQ>(lu, B.dulap, own_b[q], own_cc[q] + su[p], own_d[q], own_e[q], N, dN);
        const double res = N - srho[p];
        const double un = su[p] - omega * (res / dN);
        su[p] = un;
        m = fmax(m, fabs(res));
        if (!(un == un)) m = __longlong_as_double(0x7ff0000000000000LL);
      }
      __syncthreads();
    }
    const double bm = block_max(m, sh);
    if (threadIdx.x == 0) s_norm = bm;
    __syncthreads();
    cur = s_norm;
    ++it;
    __syncthreads();
  } while (cur > eps && prev != cur && it < max_iter && !isinf(cur));
  for (int q = threadIdx.x; q < np; q += blockDim.x) ug[q] = su[q];
  if (threadIdx.x == 0) {
    *norm = (unsigned long long)__double_as_longlong(cur);
    *iters = it;
  }
}

// ------------------------------------------------------------------
// setup: coefficients b, c [,d, e]; rho = robin value on the shell, 0
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

__global__ __launch_bounds__(256) void k_coefficients(DevBox B, Fields F, PunctureArgs pa,
                                                       double robin_a) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long np = (long long)B.n[0] * B.n[1] * B.n[2];
  if (q >= np) return;
  const int i = int(q % B.n[0]), j = int((q / B.n[0]) % B.n[1]), k = int(q / ((long long)B.n[0] * B.n[1]));
  const double x = coord(B, 0, i), y = coord(B, 1, j), z = coord(B, 2, k);
  const long long p = B.off + q;
  F.f[F_B][p] = punctures_b(pa, x, y, z);
  F.f[F_C][p] = punctures_c(pa, x, y, z);
  if (pa.eq == 1) {
    const double r = sqrt(x * x + y * y + z * z);
    F.f[F_D][p] = sf_d(pa, r);
    F.f[F_E][p] = sf_e(pa, r);
  }
  const int R = B.R;
  const bool inside = i >= R && i < B.n[0] - R && j >= R && j < B.n[1] - R && k >= R && k < B.n[2] - R;
  F.f[F_RHO][p] = inside ? 0.0 : robin_a;
  F.f[F_U][p] = 0.0;
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

__device__ double dU_at(const DevBox &B, Fields F, const double ctr[3], int i, int j, int k) {
  const long long sy = B.n[0], sz = (long long)B.n[0] * B.n[1];
  const long long p = B.off + i + sy * j + sz * k;
  const int ijk[3] = {i, j, k};
  const long long st[3] = {1, sy, sz};
  double X[3], g[3];
  for (int a = 0; a < 3; ++a) {
    X[a] = coord(B, a, ijk[a]) - ctr[a];
    double cf;
    g[a] = deriv(F.f[F_C], p, st[a], ijk[a], B.n[a], B.order, B.idx[a], cf) +
           deriv(F.f[F_U], p, st[a], ijk[a], B.n[a], B.order, B.idx[a], cf);
  }
  const double ir = 1.0 / sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
  return ir * (X[0] * g[0] + X[1] * g[1] + X[2] * g[2]);
}

__global__ __launch_bounds__(256) void k_adm_sphere(const DevBox *__restrict__ boxes, Fields F,
                                                     const double *__restrict__ centres,
                                                     double *__restrict__ sph, int nt) {
  const int box = blockIdx.y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nt * nt) return;
  const DevBox &B = boxes[box];
  const double ctr[3] = {centres[3 * box], centres[3 * box + 1], centres[3 * box + 2]};
  double L0 = coord(B, 0, B.n[0] - 1) - B.lo[0];
  double L1 = coord(B, 1, B.n[1] - 1) - B.lo[1];
  double L2 = coord(B, 2, B.n[2] - 1) - B.lo[2];
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
    v = dU_at(B, F, ctr, c[0], c[1], c[2]);
  } else {
    for (int a = 0; a < 3; ++a) c[a] = c[a] == 0 ? 1 : c[a];
    const double fx = (X[0] - coord(B, 0, c[0] - 1)) * B.idx[0];
    const double fy = (X[1] - coord(B, 1, c[1] - 1)) * B.idx[1];
    const double fz = (X[2] - coord(B, 2, c[2] - 1)) * B.idx[2];
    const double gx = 1 - fx, gy = 1 - fy, gz = 1 - fz;
    const int I = c[0], J = c[1], K = c[2];
    v = ((dU_at(B, F, ctr, I - 1, J - 1, K - 1) * gx + dU_at(B, F, ctr, I, J - 1, K - 1) * fx) * gy +
         (dU_at(B, F, ctr, I - 1, J, K - 1) * gx + dU_at(B, F, ctr, I, J, K - 1) * fx) * fy) * gz +
        ((dU_at(B, F, ctr, I - 1, J - 1, K) * gx + dU_at(B, F, ctr, I, J - 1, K) * fx) * gy +
         (dU_at(B, F, ctr, I - 1, J, K) * gx + dU_at(B, F, ctr, I, J, K) * fx) * fy) * fz;
  }
  sph[(long long)box * nt * nt + q] = v * sin(th);
}

// Lagrange interpolation of one field at arbitrary points (one box per
// point), interpolation.cc:189-303
__global__ __launch_bounds__(64) void k_interp_points(const DevBox *__restrict__ boxes, Fields F,
                                                       int field, const int *__restrict__ pbox,
                                                       const double *__restrict__ pxyz, int npts,
                                                       int ord, double *__restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  const DevBox &B = boxes[pbox[q]];
  const double *__restrict__ f = F.f[field];
  const double X[3] = {pxyz[3 * q], pxyz[3 * q + 1], pxyz[3 * q + 2]};
  int c[3], s[3];
  bool coinc = true;
  for (int a = 0; a < 3; ++a) {
    c[a] = cell_of(B, a, X[a]);
    if (!(fabs(X[a] - coord(B, a, c[a])) <= 3 * kEps)) coinc = false;
  }
  const long long sy = B.n[0], sz = (long long)B.n[0] * B.n[1];
  if (coinc) {
    out[q] = f[B.off + c[0] + sy * c[1] + sz * c[2]];
    return;
  }
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
      for (int K = 0; K <= ord; ++K) uk += f[B.off + (s[0] + I) + sy * (s[1] + J) + sz * (s[2] + K)] * w[2][K];
      uj += uk * w[1][J];
    }
    acc += uj * w[0][I];
  }
  out[q] = acc;
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

inline unsigned ceil_div(long long a, long long b) { return unsigned((a + b - 1) / b); }

inline bool empty_grid(dim3 g) { return g.x == 0 || g.y == 0 || g.z == 0; }

}  // namespace

// ------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------

void launch_coefficients_box(hipStream_t s, const DevBox &B, Fields F, const PunctureArgs &pa,
                             double robin_a) {
  const long long np = (long long)B.n[0] * B.n[1] * B.n[2];
  hipLaunchKernelGGL(k_coefficients, dim3(ceil_div(np, 256)), dim3(256), 0, s, B, F, pa, robin_a);
}

#define OLL_DISPATCH_ORDER(ORDER, EQ, KERNEL, ...)                    \
  do {                                                                \
    if (EQ) {                                                         \
      switch (ORDER) {                                                \
        case 2: KERNEL<2, 1>(__VA_ARGS__); break;                     \
        case 4: KERNEL<4, 1>(__VA_ARGS__); break;                     \
        case 6: KERNEL<6, 1>(__VA_ARGS__); break;                     \
        default: KERNEL<8, 1>(__VA_ARGS__); break;                    \
      }                                                               \
    } else {                                                          \
      switch (ORDER) {                                                \
        case 2: KERNEL<2, 0>(__VA_ARGS__); break;                     \
        case 4: KERNEL<4, 0>(__VA_ARGS__); break;                     \
        case 6: KERNEL<6, 0>(__VA_ARGS__); break;                     \
        default: KERNEL<8, 0>(__VA_ARGS__); break;                    \
      }                                                               \
    }                                                                 \
  } while (0)

namespace {
template <int ORD, int EQ>
void smooth_impl(hipStream_t s, const LevelView &lv, const DevBox *boxes, Fields F, int colour,
                 unsigned long long *norm) {
  constexpr int NC = ORD / 2 + 1;
  const long long rows = lv.row0[lv.nbox];
  const unsigned per_row = unsigned((lv.max_row_len + NC - 1) / NC);
  const dim3 grid(ceil_div(rows * per_row, 256));
  if (empty_grid(grid)) return;
  hipLaunchKernelGGL((k_smooth<ORD, EQ>), grid, dim3(256), 0, s, lv, boxes, F, colour, per_row,
                     1.0f / float(per_row), norm);
}

template <int ORD, int EQ>
void residual_impl(hipStream_t s, const LevelView &lv, const DevBox *boxes, Fields F,
                   int write_defect, unsigned long long *norm) {
  const int rows = lv.row0[lv.nbox];
  dim3 grid(ceil_div(rows, 4), ceil_div(lv.max_row_len, 64));
  if (empty_grid(grid)) return;
  hipLaunchKernelGGL((k_residual<ORD, EQ>), grid, dim3(256), 0, s, lv, boxes, F, write_defect,
                     norm);
}

template <int ORD, int EQ>
void restrict_defect_impl(hipStream_t s, const XferView &xv, const DevBox *boxes, Fields F, int mg,
                          double ipn) {
  const long long np = xv.tgt.pt0[xv.tgt.nbox];
  hipLaunchKernelGGL((k_restrict_defect<ORD, EQ>), dim3(ceil_div(np, 256)), dim3(256), 0, s, xv,
                     boxes, F, mg, ipn);
}
}  // namespace

void launch_smooth(hipStream_t s, const LevelView &lv, const DevBox *boxes, Fields F, int order,
                   int eq, int colour, unsigned long long *norm) {
  OLL_DISPATCH_ORDER(order, eq, smooth_impl, s, lv, boxes, F, colour, norm);
}

void launch_residual(hipStream_t s, const LevelView &lv, const DevBox *boxes, Fields F, int order,
                     int eq, int write_defect, unsigned long long *norm) {
  OLL_DISPATCH_ORDER(order, eq, residual_impl, s, lv, boxes, F, write_defect, norm);
}

void launch_robin(hipStream_t s, const DevBox *boxes, int box, const DevBox &hb, Fields F, int mode,
                  double ipn) {
  const int R = hb.R;
  const int ni = hb.n[0] - 2 * R, nj = hb.n[1] - 2 * R, nk = hb.n[2] - 2 * R;
  const long long cols = 2LL * (nj * nk + ni * nk + ni * nj);
  hipLaunchKernelGGL(k_robin_faces, dim3(ceil_div(cols, 256)), dim3(256), 0, s, boxes, box, F, mode,
                     ipn);
  hipLaunchKernelGGL(k_robin_edges_corners, dim3(1), dim3(1024), 0, s, boxes, box, F, mode, ipn);
}

void launch_restrict_u(hipStream_t s, const XferView &xv, const DevBox *boxes, Fields F) {
  const long long np = xv.tgt.pt0[xv.tgt.nbox];
  hipLaunchKernelGGL(k_restrict_u, dim3(ceil_div(np, 256)), dim3(256), 0, s, xv, boxes, F);
}

void launch_restrict_defect(hipStream_t s, const XferView &xv, const DevBox *boxes, Fields F,
                            int order, int eq, int mg, double ipn) {
  OLL_DISPATCH_ORDER(order, eq, restrict_defect_impl, s, xv, boxes, F, mg, ipn);
}

void launch_prolong(hipStream_t s, const XferView &xv, const DevBox *boxes, Fields F, int region) {
  const long long np = xv.tgt.pt0[xv.tgt.nbox];
  hipLaunchKernelGGL(k_prolong, dim3(ceil_div(np, 256)), dim3(256), 0, s, xv, boxes, F, region);
}

void launch_ghost_fill(hipStream_t s, const XferView &xv, const DevBox *boxes, Fields F) {
  const long long np = xv.tgt.shell0[xv.tgt.nbox];
  if (np == 0) return;
  hipLaunchKernelGGL(k_ghost_fill, dim3(ceil_div(np, 256)), dim3(256), 0, s, xv, boxes, F);
}

int coarse_solve_lds_limit() { return 160 * 1024; }

bool launch_coarse_solve(hipStream_t s, const DevBox *boxes, int box, const DevBox &hb, Fields F,
                         int eq, double ipn, double eps, double omega, int max_iter,
                         unsigned long long *norm, int *iters) {
  const size_t lds = 2 * sizeof(double) * size_t(hb.n[0]) * hb.n[1] * hb.n[2];
  const long long nint = (long long)(hb.n[0] - 2) * (hb.n[1] - 2) * (hb.n[2] - 2);
  if (lds + 2048 > size_t(coarse_solve_lds_limit()) || nint > 1024LL * kCoarseOwn) return false;
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_coarse_solve<0>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, coarse_solve_lds_limit() - 1024);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_coarse_solve<1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, coarse_solve_lds_limit() - 1024);
    attr_set = true;
  }
  if (eq)
    hipLaunchKernelGGL(k_coarse_solve<1>, dim3(1), dim3(1024), lds, s, boxes, box, F, ipn, eps,
                       omega, max_iter, norm, iters);
  else
    hipLaunchKernelGGL(k_coarse_solve<0>, dim3(1), dim3(1024), lds, s, boxes, box, F, ipn, eps,
                       omega, max_iter, norm, iters);
  return true;
}

void launch_zero_u64(hipStream_t s, unsigned long long *p, int n) {
  hipLaunchKernelGGL(k_zero_u64, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n);
}

void launch_copy(hipStream_t s, double *dst, const double *src, long long n) {
  hipLaunchKernelGGL(k_copy, dim3(ceil_div(n, 256)), dim3(256), 0, s, dst, src, n);
}

void launch_adm_sphere(hipStream_t s, const DevBox *boxes, int nbox, Fields F,
                       const double *centres, double *sph, int nt) {
  hipLaunchKernelGGL(k_adm_sphere, dim3(ceil_div(nt * nt, 256), nbox), dim3(256), 0, s, boxes, F,
                     centres, sph, nt);
}

void launch_interp_points(hipStream_t s, const DevBox *boxes, Fields F, int field, const int *pbox,
                          const double *pxyz, int npts, int ord, double *out) {
  hipLaunchKernelGGL(k_interp_points, dim3(ceil_div(npts, 64)), dim3(64), 0, s, boxes, F, field,
                     pbox, pxyz, npts, ord, out);
}

}  // namespace oll
