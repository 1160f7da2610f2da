// Host check of the fused double-buffered colour passes (k_smooth_fused in
// ollipticid_amd/csrc/kernels.hip, with the index helpers of kernels.hpp):
// the kernel's enumeration of a tile's rows (tile_row, tile_halo), of the
// colour points of a row (colour_first) and its choice of field per
// neighbour (nbr_in_y) are replayed on the CPU, tile by tile in a shuffled
// order, for boxes of 2 (both sweep orders), 3, 4 and 5 colours, many box
// shapes and tile shapes, with the launch sequence the engine issues (a
// ghost-shell copy, then per sweep the pass pairs (0, 1), (2, 3) and a
// single last pass for an odd colour count, the fields swapped from sweep
// to sweep).  Checked:
//  * every element touched lies in [0, size) and every written point is an
//    interior point of the pass's colour;
//  * every read from the written field Y of a point updated in the current
//    sweep was written by an earlier launch, or, for the first pass's
//    points read by the second pass, by the same tile in its first pass
//    (no dependence between workgroups of one launch); every read from the
//    field X holds the previous sweep's value; tiles that write one point
//    in the same launch write the same bits;
//  * after the sweeps the field holding u equals, bit for bit, in-place
//    colour-by-colour Gauss-Seidel sweeps of the same point update, and
//    the norm over the tiles' own rows equals the max over all updates.
// Prints "ok" and exits 0, or the first failure and exits 1.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../ollipticid_amd/csrc/kernels.hpp"

using namespace oll;

namespace {

struct Fail {};

void check(bool ok, const char *what, int NC, const DevBox &B, int ty, int tz, int i, int j, int k) {
  if (ok) return;
  std::printf("fused check failed (%s): nc=%d flip=%d n=(%d,%d,%d) tile %dx%d point (%d,%d,%d)\n",
              what, NC, B.flip, B.n[0], B.n[1], B.n[2], ty, tz, i, j, k);
  throw Fail{};
}

bool interior(const DevBox &B, int i, int j, int k) {
  const int R = B.R;
  return i >= R && i < B.n[0] - R && j >= R && j < B.n[1] - R && k >= R && k < B.n[2] - R;
}

// pass index of colour c (two-colour boxes may sweep in descending order)
int pass_of(const DevBox &B, int c) { return B.flip ? B.nc - 1 - c : c; }

// the point update both sides use: any fixed expression of the point's
// value, its 6R star neighbours (summed in pairs, as the kernel does) and
// a per-point constant
double update(double u0, const double *sx, const double *sy, const double *sz, int R, double rho,
              double &res) {
  double acc = 0.0;
  for (int m = 0; m < R; ++m) acc += (0.31 + 0.07 * m) * sx[m] + (0.23 - 0.05 * m) * sy[m] +
                                     (0.17 + 0.03 * m) * sz[m];
  const double un = 0.2 * u0 + 0.5 * acc / (6.0 * R) + rho;
  res = un - u0;
  return un;
}

// in-place colour-by-colour sweeps (the unfused launches' order)
double reference_sweeps(const DevBox &B, std::vector<double> &U, const std::vector<double> &rho,
                        int nsweeps) {
  const int R = B.R, NC = B.nc;
  double rmax = 0.0;
  for (int sw = 0; sw < nsweeps; ++sw)
    for (int q = 0; q < NC; ++q)
      for (int k = R; k < B.n[2] - R; ++k)
        for (int j = R; j < B.n[1] - R; ++j)
          for (int i = R; i < B.n[0] - R; ++i) {
            if (pass_of(B, colour_of(B, i, j, k)) != q) continue;
            const long long p = pidx(B, i, j, k);
            double sx[4], sy[4], sz[4];
            for (int m = 1; m <= R; ++m) {
              sx[m - 1] = U[size_t(pidx(B, i + m, j, k))] + U[size_t(pidx(B, i - m, j, k))];
              sy[m - 1] = U[size_t(pidx(B, i, j + m, k))] + U[size_t(pidx(B, i, j - m, k))];
              sz[m - 1] = U[size_t(pidx(B, i, j, k + m))] + U[size_t(pidx(B, i, j, k - m))];
            }
            double res;
            U[size_t(p)] = update(U[size_t(p)], sx, sy, sz, R, rho[size_t(p)], res);
            rmax = std::fmax(rmax, std::fabs(res));
          }
  return rmax;
}

template <int NC>
struct Emu {
  const DevBox &B;
  int ty, tz;
  std::vector<double> F[2];
  std::vector<long long> stamp[2];  // launch that last wrote each element (-1: never)
  std::vector<long long> owner;     // (launch, tile) of the current launch's first-pass writes
  std::vector<char> interior_el;
  std::vector<int> ci, cj, ck;      // element -> point
  long long launch = 0;
  long long sweep_first_launch = 0;

  Emu(const DevBox &b, int ty_, int tz_) : B(b), ty(ty_), tz(tz_) {}

  void fail(bool ok, const char *what, int i, int j, int k) { check(ok, what, NC, B, ty, tz, i, j, k); }

  // reads element e of field f for a point of pass pi in pass phase `phase`
  // (1: the pass computed on the halo too, 2: the second pass) of the tile
  // `tile`; `want_y` is the kernel's choice, nb_pass the neighbour's pass
  double read(int f_x, bool want_y, long long e, int pi, int phase, long long tile, int i, int j,
              int k) {
    fail(e >= 0 && e < B.size, "element out of range", i, j, k);
    const int fy = 1 - f_x;
    const int f = want_y ? fy : f_x;
    if (!interior_el[size_t(e)]) {  // ghost shell: the same value in both fields
      fail(std::memcmp(&F[0][size_t(e)], &F[1][size_t(e)], 8) == 0, "ghost differs between fields",
           i, j, k);
      return F[f][size_t(e)];
    }
    const int nb_pass = pass_of(B, colour_of(B, ci[size_t(e)], cj[size_t(e)], ck[size_t(e)]));
    if (want_y) {
      fail(nb_pass < pi, "Y read of a point not yet updated", i, j, k);
      fail(stamp[fy][size_t(e)] >= sweep_first_launch, "Y read of a stale value", i, j, k);
      if (stamp[fy][size_t(e)] == launch) {
        fail(phase == 2 && owner[size_t(e)] == tile, "Y read of another workgroup's write", i, j, k);
      }
    } else {
      fail(nb_pass >= pi, "X read of a point already updated", i, j, k);
      fail(stamp[f_x][size_t(e)] < sweep_first_launch, "X written during its sweep", i, j, k);
    }
    return F[f][size_t(e)];
  }

  // one pass of index PI over the first nrows rows of a tile (fused_pass)
  template <int PI>
  double pass(const TileBlock &T, int f_x, int nrows, int yb, int ya, int zb, int phase,
              long long tile, const std::vector<double> &rho) {
    constexpr int R = NC - 1;
    const int P = (B.n[0] - 2 * R + NC - 1) / NC;
    const int nt = T.ty * T.tz;
    const int c = B.flip ? NC - 1 - PI : PI;
    double rmax = 0.0;
    std::vector<char> seen(size_t(B.n[1]) * size_t(B.n[2]), 0);
    for (int r = 0; r < nrows; ++r) {
      int j, k;
      tile_row(T, r, yb, ya, zb, j, k);
      fail(j >= R && j < B.n[1] - R && k >= R && k < B.n[2] - R, "row outside the interior", -1, j, k);
      fail(!seen[size_t(k) * size_t(B.n[1]) + size_t(j)], "row listed twice", -1, j, k);
      seen[size_t(k) * size_t(B.n[1]) + size_t(j)] = 1;
      for (int t = 0; t < P; ++t) {
        const int i = colour_first(c, j, k, R, NC) + t * NC;
        if (i >= B.n[0] - R) continue;
        fail(i >= R && colour_of(B, i, j, k) == c, "slot not of the pass's colour", i, j, k);
        const int s = i - (i / NC) * NC;
        const long long p = (long long)c * B.cs + (long long)k * B.ps + (long long)j * B.rs + i / NC;
        fail(p == pidx(B, i, j, k), "element index", i, j, k);
        const Nbr<NC> nb = stencil_offs<NC, 1>(c, s, unsigned(B.cs), unsigned(B.rs), unsigned(B.ps));
        const double u0 = read(f_x, false, p, PI, phase, tile, i, j, k);
        double sx[4], sy[4], sz[4];
        for (int m = 1; m <= R; ++m) {
          const bool yp = nbr_in_y(NC, PI, m), ym = nbr_in_y(NC, PI, -m);
          auto at = [&](unsigned off) { return (long long)unsigned(unsigned(p) + off); };
          fail(at(nb.xp[m - 1]) == pidx(B, i + m, j, k) && at(nb.xm[m - 1]) == pidx(B, i - m, j, k) &&
                   at(nb.yp[m - 1]) == pidx(B, i, j + m, k) && at(nb.ym[m - 1]) == pidx(B, i, j - m, k) &&
                   at(nb.zp[m - 1]) == pidx(B, i, j, k + m) && at(nb.zm[m - 1]) == pidx(B, i, j, k - m),
               "neighbour offset", i, j, k);
          sx[m - 1] = read(f_x, yp, at(nb.xp[m - 1]), PI, phase, tile, i, j, k) +
                      read(f_x, ym, at(nb.xm[m - 1]), PI, phase, tile, i, j, k);
          sy[m - 1] = read(f_x, yp, at(nb.yp[m - 1]), PI, phase, tile, i, j, k) +
                      read(f_x, ym, at(nb.ym[m - 1]), PI, phase, tile, i, j, k);
          sz[m - 1] = read(f_x, yp, at(nb.zp[m - 1]), PI, phase, tile, i, j, k) +
                      read(f_x, ym, at(nb.zm[m - 1]), PI, phase, tile, i, j, k);
        }
        double res;
        const double un = update(u0, sx, sy, sz, R, rho[size_t(p)], res);
        const int fy = 1 - f_x;
        if (stamp[fy][size_t(p)] == launch)
          fail(std::memcmp(&F[fy][size_t(p)], &un, 8) == 0, "tiles write different values", i, j, k);
        F[fy][size_t(p)] = un;
        stamp[fy][size_t(p)] = launch;
        if (phase == 1) owner[size_t(p)] = tile;
        if (r < nt) rmax = std::fmax(rmax, std::fabs(res));
      }
    }
    return rmax;
  }

  template <int Q>
  double launch_pair(const std::vector<TileBlock> &tiles, const std::vector<int> &order, int f_x,
                     bool pair, const std::vector<double> &rho) {
    constexpr int R = NC - 1;
    double rmax = 0.0;
    std::fill(owner.begin(), owner.end(), -1);
    for (int t : order) {
      const TileBlock &T = tiles[size_t(t)];
      const int nt = T.ty * T.tz;
      if (pair) {
        int yb, ya, zb, za;
        tile_halo(T, B.n, R, yb, ya, zb, za);
        const int nrows = nt + (yb + ya) * T.tz + (zb + za) * T.ty;
        rmax = std::fmax(rmax, pass<Q>(T, f_x, nrows, yb, ya, zb, 1, t, rho));
        if constexpr (Q + 1 < NC) rmax = std::fmax(rmax, pass<Q + 1>(T, f_x, nt, 0, 0, 0, 2, t, rho));
      } else {
        rmax = std::fmax(rmax, pass<Q>(T, f_x, nt, 0, 0, 0, 1, t, rho));
      }
    }
    ++launch;
    return rmax;
  }

  double run(const std::vector<double> &u0, const std::vector<double> &rho, int nsweeps,
             std::mt19937 &rng, int &final_field) {
    const int R = NC - 1;
    F[0] = u0;
    F[1].assign(size_t(B.size), std::nan(""));
    stamp[0].assign(size_t(B.size), -1);
    stamp[1].assign(size_t(B.size), -1);
    owner.assign(size_t(B.size), -1);
    interior_el.assign(size_t(B.size), 0);
    ci.assign(size_t(B.size), -1);
    cj = ci;
    ck = ci;
    for (int k = 0; k < B.n[2]; ++k)
      for (int j = 0; j < B.n[1]; ++j)
        for (int i = 0; i < B.n[0]; ++i) {
          const size_t e = size_t(pidx(B, i, j, k));
          interior_el[e] = interior(B, i, j, k);
          ci[e] = i;
          cj[e] = j;
          ck[e] = k;
          if (!interior_el[e]) F[1][e] = F[0][e];  // k_copy_shell
        }
    // the engine's tiles (tile_blocks)
    std::vector<TileBlock> tiles;
    for (int k0 = R; k0 < B.n[2] - R; k0 += tz)
      for (int j0 = R; j0 < B.n[1] - R; j0 += ty)
        tiles.push_back(TileBlock{&B, nullptr, j0, k0, std::min(ty, B.n[1] - R - j0),
                                  std::min(tz, B.n[2] - R - k0)});
    std::vector<int> order(tiles.size());
    for (size_t t = 0; t < order.size(); ++t) order[t] = int(t);
    double rmax = 0.0;
    for (int sw = 0; sw < nsweeps; ++sw) {
      const int f_x = sw & 1;
      sweep_first_launch = launch;
      for (int q = 0; q < NC; q += 2) {
        std::shuffle(order.begin(), order.end(), rng);
        const bool pair = q + 1 < NC;
        switch (q) {
          case 0: rmax = std::fmax(rmax, launch_pair<0>(tiles, order, f_x, pair, rho)); break;
          case 2: if constexpr (NC > 2) rmax = std::fmax(rmax, launch_pair<2>(tiles, order, f_x, pair, rho)); break;
          case 4: if constexpr (NC > 4) rmax = std::fmax(rmax, launch_pair<4>(tiles, order, f_x, pair, rho)); break;
        }
      }
      // every interior point was written during the sweep
      for (long long e = 0; e < B.size; ++e)
        if (interior_el[size_t(e)])
          fail(stamp[1 - f_x][size_t(e)] >= sweep_first_launch, "point not updated in its sweep",
               ci[size_t(e)], cj[size_t(e)], ck[size_t(e)]);
    }
    final_field = nsweeps & 1;
    return rmax;
  }
};

template <int NC>
void check_case(int n0, int n1, int n2, int flip, int ty, int tz, int nsweeps, std::mt19937 &rng) {
  DevBox B{};
  B.n[0] = n0;
  B.n[1] = n1;
  B.n[2] = n2;
  B.R = NC - 1;
  B.flip = flip;
  set_layout(B, NC);
  std::uniform_real_distribution<double> dist(-1.0, 1.0);
  std::vector<double> u0(size_t(B.size), std::nan("")), rho(size_t(B.size), 0.0);
  for (int k = 0; k < n2; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n0; ++i) {
        const size_t e = size_t(pidx(B, i, j, k));
        u0[e] = dist(rng);
        rho[e] = 0.01 * dist(rng);
      }
  std::vector<double> U = u0;
  const double rref = reference_sweeps(B, U, rho, nsweeps);
  Emu<NC> emu(B, ty, tz);
  int ff;
  const double remu = emu.run(u0, rho, nsweeps, rng, ff);
  for (int k = 0; k < n2; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n0; ++i) {
        const size_t e = size_t(pidx(B, i, j, k));
        check(std::memcmp(&U[e], &emu.F[ff][e], 8) == 0, "result differs from colour passes", NC, B,
              ty, tz, i, j, k);
      }
  check(std::memcmp(&rref, &remu, 8) == 0, "norm differs from colour passes", NC, B, ty, tz, -1, -1, -1);
}

template <int NC>
void check_nc(std::mt19937 &rng) {
  const int R = NC - 1;
  const int interiors[][3] = {{1, 1, 1}, {2, 3, 1}, {7, 5, 6}, {9, 13, 4}, {12, 2, 11}, {5, 17, 9}};
  const int tiles[][2] = {{1, 1}, {2, 3}, {3, 2}, {4, 1}, {1, 5}, {8, 8}, {16, 4}};
  for (int flip = 0; flip <= (NC == 2 ? 1 : 0); ++flip)
    for (const auto &in : interiors)
      for (const auto &t : tiles)
        for (int ns = 1; ns <= 3; ++ns)
          check_case<NC>(in[0] + 2 * R, in[1] + 2 * R, in[2] + 2 * R, flip, t[0], t[1], ns, rng);
}

}  // namespace

int main() {
  std::mt19937 rng(12345);
  try {
    check_nc<2>(rng);
    check_nc<3>(rng);
    check_nc<4>(rng);
    check_nc<5>(rng);
  } catch (Fail &) {
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
