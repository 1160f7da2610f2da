// Box hierarchy geometry (host side); see geometry.hpp for the reference
// functions each part follows.  Floating-point expressions keep the
// reference's operation order: box extents must match it to the last bit
// because later "is this point inside / coincident" tests depend on them.
#include "geometry.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <string>

namespace oll {

namespace {
constexpr double kEps = DBL_EPSILON;

inline bool near3eps(double a, double b) { return std::fabs(a - b) <= 3 * kEps; }

// interface.cc:883-949: sizes the multigrid hierarchy accepts
constexpr size_t kValid[63] = {
    3,   4,   5,   6,   7,   8,   9,   10,  11,  12,  13,  14,  15,  17,  19,  21,
    23,  25,  27,  29,  33,  37,  41,  45,  49,  53,  57,  65,  73,  81,  89,  97,
    105, 113, 129, 145, 161, 177, 193, 209, 225, 257, 289, 321, 353, 385, 417, 449,
    513, 577, 641, 705, 769, 833, 897, 1153, 1281, 1409, 1665, 1793, 2305, 2817, 3329};

// domain.cc:451-525 for a single partition without symmetry: trims or
// extends the point count so the last point does not pass the far end,
// then re-derives the near end from the far one.
void fix_axis(double &lo, double &hi, size_t &n, double glo, double ghi, double h) {
  const double near_end = glo + 0.0;
  const double far_end = ghi - 0.0;
  while (lo + double(n - 1) * h > far_end) --n;
  if (lo + double(n - 1) * h < far_end - 0.5 * h) ++n;
  hi = lo + double(n - 1) * h;
  while (hi - double(n - 1) * h < near_end) --n;
  if (hi - double(n - 1) * h > near_end + 0.5 * h) ++n;
  lo = hi - double(n - 1) * h;
}
}  // namespace

size_t valid_n(size_t n) {
  for (size_t v : kValid)
    if (v == n) return n;
  size_t below = kValid[61], above = kValid[62];
  for (size_t i = 0; i < 62 && kValid[i] < n; ++i) {
    below = kValid[i];
    above = kValid[i + 1];
  }
  return (n - below < above - n) ? below : above;
}

Box make_box(const size_t N[3], const double d[3], const double c[3], int order) {
  Box b;
  b.order = order;
  b.R = order / 2 >= 1 ? order / 2 : 1;
  for (int a = 0; a < 3; ++a) {
    const size_t gw = size_t(b.R);
    size_t n = valid_n(N[a]);
    b.d[a] = d[a];
    b.c[a] = c[a];
    // domain.cc:326-351
    const double glo = c[a] - 0.5 * double(n - 1) * d[a] - double(gw - 1) * d[a];
    n += gw - 1;
    const double ghi = glo + double(n - 1) * d[a] + double(gw - 1) * d[a];
    n += gw - 1;
    b.g0[a] = glo;
    b.g1[a] = ghi;
    b.lo[a] = glo;
    b.hi[a] = ghi;
    b.n[a] = n;
    fix_axis(b.lo[a], b.hi[a], b.n[a], glo, ghi, d[a]);
    b.len[a] = b.hi[a] - b.lo[a];
  }
  return b;
}

// domain.cc:1493-1675: binary search for the first node at or beyond x
size_t Box::cell(int a, double x) const {
  const int np = int(n[a]);
  if (!(lo[a] < x && x < hi[a])) return (lo[a] >= x) ? 1 : size_t(np - 1);
  int lo_i = 0, hi_i = np;
  int i = (np % 2 == 0) ? np / 2 : (np - 1) / 2;
  do {
    const double xi = lo[a] + double(i) * d[a];
    if (xi > x || near3eps(xi, x))
      hi_i = i;
    else
      lo_i = i;
    const int s = hi_i + lo_i;
    i = (s % 2 == 0) ? s / 2 : (s - 1) / 2;
  } while (lo_i + 1 < hi_i);
  return size_t(i + 1);
}

// domain.h:839-850
bool Box::in_nb(int a, double x) const {
  return lo[a] + d[a] * (double(R) - 0.1) < x && x < hi[a] - d[a] * (double(R) - 0.1);
}

// domain.h:767-775
bool Box::in_box(int a, double x) const {
  return lo[a] - d[a] * 0.1 <= x && x <= hi[a] + d[a] * 0.1;
}

namespace {

// domain.cc:1811-1871: union of two boxes (padded extents), re-snapped
Box merge(const Box &A, const Box &B) {
  size_t n[3];
  double c[3], d[3];
  for (int a = 0; a < 3; ++a) {
    const double lo = std::min(A.g0[a], B.g0[a]);
    const double hi = std::max(A.g1[a], B.g1[a]);
    c[a] = 0.5 * (lo + hi);
    n[a] = size_t((hi - lo) / A.d[a] + 1);
    d[a] = A.d[a];
  }
  return make_box(n, d, c, A.order);
}

// domain.cc:1878-1922, 1972-1980, 1998-2005: A && B is B >= (B - A),
// i.e. the boxes overlap or touch on every axis.
bool touches(const Box &A, const Box &B) {
  for (int a = 0; a < 3; ++a) {
    const double ci = std::max(B.lo[a], A.lo[a]);
    const double cf = std::min(B.hi[a], A.hi[a]);
    const double lo = std::min(ci, cf), hi = std::max(ci, cf);
    if (!(B.lo[a] <= lo && hi <= B.hi[a])) return false;
  }
  return true;
}

}  // namespace

std::string build_hierarchy(size_t Nin, double L, int levels, int order,
                            const PunctureSet &p, Hierarchy &h, bool coarsen) {
  if (!(order == 2 || order == 4 || order == 6 || order == 8))
    return "order must be [2,4,6,8].";
  if (levels < 1) return "levels must be >= 1";
  const size_t N = valid_n(Nin);
  const double Dx = std::fabs(L) / double(N - 1.0);  // interface.cc:1061

  // multigrid coarsening of the base box, elliptic.cc:502-651
  std::vector<size_t> ns{N};
  std::vector<double> ds{Dx};
  while (coarsen) {
    const size_t half = (ns.back() + 1) / 2;
    if (half < 12) break;
    ns.push_back(half);
    ds.push_back(ds.back() * 2.0);
  }
  if (coarsen && ns.size() <= 1) return "There are not enough grid-points for multigrid";
  h.mgrid = int(ns.size()) - 1;
  h.levels = levels + h.mgrid;

  std::vector<std::vector<Box>> per_level(size_t(h.levels));
  for (int l = h.mgrid; l >= 0; --l) {
    const size_t n3[3] = {ns[l], ns[l], ns[l]};
    const double d3[3] = {ds[l], ds[l], ds[l]};
    const double c3[3] = {0, 0, 0};
    per_level[size_t(h.mgrid - l)].push_back(make_box(n3, d3, c3, l == 0 ? order : 2));
  }
  const Box &base = per_level[size_t(h.mgrid)][0];

  // refinement boxes centred on the punctures, elliptic.cc:712-828
  const size_t nbh = p.x.size();
  for (int l = h.mgrid + 1; l < h.levels; ++l) {
    const double hs = Dx / std::pow(2.0, double(l - h.mgrid));
    for (size_t i = 0; i < nbh; ++i) {
      const double pos[3] = {p.x[i], p.y[i], p.z[i]};
      double ctr[3];
      for (int a = 0; a < 3; ++a) {
        const size_t ic = base.cell(a, pos[a]);
        double lo = base.x(a, long(ic) - 1), hi = base.x(a, long(ic));
        if (near3eps(lo, pos[a]) || near3eps(hi, pos[a])) {
          ctr[a] = pos[a];
        } else {
          while (lo + hs < pos[a]) lo += hs;
          while (hi - hs > pos[a]) hi -= hs;
          ctr[a] = std::fabs(lo - pos[a]) < std::fabs(hi - pos[a]) ? lo : hi;
        }
      }
      const size_t n3[3] = {N, N, N};
      const double d3[3] = {hs, hs, hs};
      per_level[size_t(l)].push_back(make_box(n3, d3, ctr, order));
    }
  }

  // Merge boxes of a level that overlap or touch, rescanning the level
  // after a merge (elliptic.cc:839-896).  The reference keeps every box of
  // every level in one list D and addresses box i of level l as the level's
  // first index plus i, with i running up to the level's count as it stood
  // when the pair was reached: after a merge of (i, j) with j < i erases box
  // j, D[first + i] is the next box in the list, possibly the first box of
  // level l + 1, and that box is what receives the union.  The same list
  // arithmetic is kept here, so the hierarchy is the reference's one also in
  // those cases; an index past the end of the list (where the reference
  // reads past its vector) is refused.
  std::vector<Box> D;
  std::vector<int> ndl(size_t(h.levels), 0);
  for (int l = 0; l < h.levels; ++l) {
    ndl[size_t(l)] = int(per_level[size_t(l)].size());
    D.insert(D.end(), per_level[size_t(l)].begin(), per_level[size_t(l)].end());
  }
  for (int l = h.mgrid + 1; l < h.levels; ++l) {
    bool restore = false;
    for (int i = 0; i < ndl[size_t(l)]; ++i)
      for (int j = 0; j < ndl[size_t(l)]; ++j) {
        if (i == j) continue;
        size_t first = 0;
        for (int ll = 0; ll < l; ++ll) first += size_t(ndl[size_t(ll)]);
        const size_t ib = first + size_t(i), jb = first + size_t(j);
        if (ib >= D.size() || jb >= D.size())
          return "box merging at level " + std::to_string(l) +
                 " reaches past the box list (the reference's grid set-up fails here)";
        if (touches(D[ib], D[jb])) {
          D[ib] = merge(D[ib], D[jb]);
          D.erase(D.begin() + long(jb));
          --ndl[size_t(l)];
          restore = true;
        }
      }
    if (restore) --l;
  }
  size_t q = 0;
  for (int l = 0; l < h.levels; ++l) {
    per_level[size_t(l)].assign(D.begin() + long(q), D.begin() + long(q) + ndl[size_t(l)]);
    q += size_t(ndl[size_t(l)]);
  }

  h.boxes.clear();
  h.level_first.assign(size_t(h.levels), 0);
  h.level_count.assign(size_t(h.levels), 0);
  for (int l = 0; l < h.levels; ++l) {
    h.level_first[size_t(l)] = int(h.boxes.size());
    h.level_count[size_t(l)] = int(per_level[size_t(l)].size());
    for (Box b : per_level[size_t(l)]) {
      b.level = l;
      h.boxes.push_back(b);
    }
  }
  return "";
}

// interpolation.cc:125-375 restated per axis
AxisTable build_axis_table(const Box &T, const Box &S, int a, int method, int ord) {
  AxisTable t;
  const size_t n = T.n[a];
  t.k = method == 0 ? 2 : ord + 1;
  t.start.assign(n, 0);
  t.coinc.assign(n, 0);
  t.flags.assign(n, 0);
  t.w.assign(n * size_t(t.k), 0.0);
  const double ih = 1.0 / S.d[a];
  for (size_t ti = 0; ti < n; ++ti) {
    const double x = T.x(a, long(ti));
    uint8_t f = 0;
    if (S.in_nb(a, x)) f |= AX_NB;
    if (S.in_box(a, x)) f |= AX_IN;
    const size_t i = S.cell(a, x);
    t.coinc[ti] = int(i);
    if (near3eps(x, S.x(a, long(i)))) f |= AX_COINC;
    // full-weighting eligibility: coincident within 1e-6 dx with a node
    // whose neighbours i-1, i+1 lie in the source interior
    for (long cnd = long(i) - 1; cnd <= long(i); ++cnd) {
      if (cnd < 0 || cnd >= long(S.n[a])) continue;
      if (std::fabs(S.x(a, cnd) - x) <= 1e-6 * S.d[a]) {
        if (cnd - 1 >= S.R && cnd + 1 < long(S.n[a]) - S.R) {
          f |= AX_FW;
          t.coinc[ti] = int(cnd);
        }
        break;
      }
    }
    double *w = &t.w[ti * size_t(t.k)];
    if (method == 0) {
      const size_t ii = i == 0 ? 1 : i;
      const double frac = (x - S.x(a, long(ii) - 1)) * ih;
      t.start[ti] = int(ii) - 1;
      w[0] = 1 - frac;
      w[1] = frac;
    } else {
      long s0 = (ord % 2 == 0) ? long(i) - ord / 2 : long(i) - (ord + 1) / 2;
      if (s0 < 0)
        s0 = 0;
      else if (s0 + ord >= long(S.n[a]))
        s0 = long(S.n[a]) - ord - 1;
      t.start[ti] = int(s0);
      for (int q = 0; q <= ord; ++q) {
        const long node = s0 + q;
        const double xn = S.x(a, node);
        double lp = 1.0;
        for (long m = s0; m < node; ++m) lp *= (x - S.x(a, m)) / (xn - S.x(a, m));
        for (long m = node + 1; m <= s0 + ord; ++m) lp *= (x - S.x(a, m)) / (xn - S.x(a, m));
        w[q] = lp;
      }
    }
    t.flags[ti] = f;
  }
  return t;
}

bool nested_axis(const AxisTable &t, int &m) {
  bool have = false;
  auto near = [](double w, double v) { return std::fabs(w - v) <= 1e-12; };
  for (size_t i = 0; i < t.flags.size(); ++i) {
    if (!(t.flags[i] & (AX_NB | AX_IN))) continue;
    const double w0 = t.w[2 * i], w1 = t.w[2 * i + 1];
    int q;  // position in half source spacings
    if (t.flags[i] & AX_COINC)
      q = 2 * t.coinc[i];
    else if (near(w0, 0.5) && near(w1, 0.5))
      q = 2 * t.start[i] + 1;
    else if (near(w0, 1.0) && near(w1, 0.0))  // on node start (the cell search clamps)
      q = 2 * t.start[i];
    else if (near(w0, 0.0) && near(w1, 1.0))
      q = 2 * t.start[i] + 2;
    else
      return false;
    if (!have) m = q - int(i), have = true;
    if (q - int(i) != m || q < 0) return false;
  }
  return have;
}

}  // namespace oll
