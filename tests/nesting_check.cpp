// Host check of the prolongation's nested fast path
// (ollipticid_amd/csrc/geometry.cpp nested_axis, kernels.hip
// trilinear_nested): for the hierarchies of several problems, every
// (target box, coarser source box, axis) linear table that nested_axis
// accepts gives, at every target index inside the source, the same
// one-dimensional interpolation from the index map (node q/2 with weight 1,
// or nodes (q-1)/2, (q+1)/2 with weights 1/2; the second node clamped to the
// last one) as the table's own nodes and weights, to 1e-13 relative on a
// test function.  Also requires the equal-mass binary workload's tables all
// to be nested.  Prints "ok <nested> <tables>" and exits 0, or the first
// mismatch and exits 1.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../ollipticid_amd/csrc/geometry.hpp"

using namespace oll;

namespace {

double f(long node) { return std::sin(0.37 * double(node)) + 0.01 * double(node); }

struct Problem {
  size_t N;
  double L;
  int levels, order;
  std::vector<double> x, y, z;
  bool all_nested;
};

int check(const Problem &p, int &nested, int &tables) {
  PunctureSet ps;
  ps.x = p.x;
  ps.y = p.y;
  ps.z = p.z;
  Hierarchy g;
  const std::string err = build_hierarchy(p.N, p.L, p.levels, p.order, ps, g);
  if (!err.empty()) {
    std::printf("hierarchy: %s\n", err.c_str());
    return 1;
  }
  for (int k = 1; k < g.levels; ++k)
    for (int t = 0; t < g.level_count[size_t(k)]; ++t)
      for (int s = 0; s < g.level_count[size_t(k - 1)]; ++s) {
        const Box &T = g.boxes[size_t(g.level_first[size_t(k)] + t)];
        const Box &S = g.boxes[size_t(g.level_first[size_t(k - 1)] + s)];
        for (int a = 0; a < 3; ++a) {
          const AxisTable tab = build_axis_table(T, S, a, 0, 1);
          ++tables;
          int m = 0;
          if (!nested_axis(tab, m)) {
            if (p.all_nested && g.level_count[size_t(k - 1)] == 1 && g.level_count[size_t(k)] == 1) {
              std::printf("binary table not nested: level %d box %d axis %d\n", k, t, a);
              return 1;
            }
            continue;
          }
          ++nested;
          for (size_t i = 0; i < tab.flags.size(); ++i) {
            if (!(tab.flags[i] & (AX_NB | AX_IN))) continue;
            const bool co = (tab.flags[i] & AX_COINC) != 0;  // the kernel's coincident path
            // the table's interpolation (the coincident path reads its node)
            const double want = co ? f(tab.coinc[i])
                                   : tab.w[2 * i] * f(tab.start[i]) + tab.w[2 * i + 1] * f(tab.start[i] + 1);
            const long q = long(i) + m;
            long c = q >> 1;
            double wi = (q & 1) ? 0.5 : 1.0, wf = (q & 1) ? 0.5 : 0.0;
            if (!(q & 1) && c + 1 >= long(S.n[a])) c = long(S.n[a]) - 2, wi = 0.0, wf = 1.0;
            if (c < 0 || c + 1 >= long(S.n[a])) {
              std::printf("node out of range: level %d axis %d index %zu\n", k, a, i);
              return 1;
            }
            const double got = wi * f(c) + wf * f(c + 1);
            if (std::fabs(got - want) > 1e-13 * (1.0 + std::fabs(want))) {
              std::printf("mismatch: level %d box %d axis %d index %zu: %.17g vs %.17g\n", k, t, a, i,
                          got, want);
              return 1;
            }
          }
        }
      }
  return 0;
}

}  // namespace

int main() {
  const Problem probs[] = {
      // equal-mass binary (the benchmark): every source level holds one box
      {65, 1000, 9, 4, {0, 0}, {5.5, -5.5}, {0, 0}, true},
      // two punctures whose finer boxes merge, 4th order
      {25, 48, 3, 4, {1.1, -1.3}, {0.3, -0.4}, {-0.2, 0.5}, false},
      // four punctures, 2nd order, merged boxes
      {33, 64, 4, 2, {2.1, -2.3, 0.4, -0.7}, {1.3, -1.1, 3.0, -2.9}, {0.2, -0.4, 1.1, -1.2}, false},
      // off-lattice single puncture, 6th order
      {33, 40, 3, 6, {0.77}, {-0.31}, {0.13}, false},
  };
  int nested = 0, tables = 0;
  for (const Problem &p : probs)
    if (check(p, nested, tables)) return 1;
  std::printf("ok %d %d\n", nested, tables);
  return 0;
}
