// Host check of the colour-block field layout (ollipticid_amd/csrc/kernels.hpp):
// for boxes of every order and a range of shapes, with R + 1 colours and
// with the 5 colours of Punctures_EM at orders 4-8,
//  * pidx maps the points of a box one to one into [0, size);
//  * stencil_offs gives, for every point and step m = 1..NC-1 along each axis,
//    exactly pidx(neighbour) - pidx(point) (modulo 2^32, as the kernels add
//    it to 32-bit byte offsets), for element size 1, 4 and 8;
//  * the colour passes' and whole-row kernels' slot arithmetic (first
//    interior point of a colour in a row, element index of slot t) and the
//    transfers' steps along i, j and k (one colour block on, wrapping after
//    the last block, plus one slot, row or plane) and the Robin derivatives'
//    axis offsets (up to 4 points either way) agree with pidx.
// Prints "ok" and exits 0, or the first mismatch and exits 1.
#include <cstdio>
#include <vector>

#include "../ollipticid_amd/csrc/kernels.hpp"

using namespace oll;

namespace {

int fail(const char *what, int R, int n0, int n1, int n2, int i, int j, int k, int a, int m) {
  std::printf("mismatch (%s): R=%d n=(%d,%d,%d) point (%d,%d,%d) axis %d step %d\n", what, R, n0, n1,
              n2, i, j, k, a, m);
  return 1;
}

template <int NC, int ES>
int check_offs(const DevBox &B) {
  const int R = B.R;
  for (int k = 0; k < B.n[2]; ++k)
    for (int j = 0; j < B.n[1]; ++j)
      for (int i = 0; i < B.n[0]; ++i) {
        const long long p = pidx(B, i, j, k);
        const Nbr<NC> nb = stencil_offs<NC, ES>(colour_of(B, i, j, k), i % NC, unsigned(B.cs),
                                                unsigned(B.rs), unsigned(B.ps));
        for (int m = 1; m < NC; ++m) {  // every offset stencil_offs gives (>= R)
          const int ijk[3] = {i, j, k};
          const unsigned *pos[3] = {nb.xp, nb.yp, nb.zp}, *neg[3] = {nb.xm, nb.ym, nb.zm};
          for (int a = 0; a < 3; ++a) {
            for (int sgn = -1; sgn <= 1; sgn += 2) {
              int q[3] = {ijk[0], ijk[1], ijk[2]};
              q[a] += sgn * m;
              if (q[a] < 0 || q[a] >= B.n[a]) continue;
              const unsigned want = unsigned((pidx(B, q[0], q[1], q[2]) - p) * ES);
              const unsigned got = sgn > 0 ? pos[a][m - 1] : neg[a][m - 1];
              if (want != got) return fail("stencil_offs", R, B.n[0], B.n[1], B.n[2], i, j, k, a, sgn * m);
            }
          }
        }
      }
  return 0;
}

int check_box(int R, int NC, int n0, int n1, int n2) {
  DevBox B{};
  B.n[0] = n0;
  B.n[1] = n1;
  B.n[2] = n2;
  B.R = R;
  set_layout(B, NC);
  if (B.size != layout_size(B.n, NC)) return fail("layout_size", R, n0, n1, n2, 0, 0, 0, 0, 0);
  std::vector<char> seen(size_t(B.size), 0);
  for (int k = 0; k < n2; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n0; ++i) {
        const long long p = pidx(B, i, j, k);
        if (p < 0 || p >= B.size || seen[size_t(p)]) return fail("pidx", R, n0, n1, n2, i, j, k, 0, 0);
        seen[size_t(p)] = 1;
        // transfers walk stencils one step at a time: +1 along any axis
        // moves one colour block on (wrapping after the last) plus one
        // plane (k), one row (j) or, when i leaves residue NC-1, one slot
        const unsigned c = unsigned(colour_of(B, i, j, k));
        const unsigned cs8 = unsigned(B.cs) * 8u;
        const unsigned dc = c + 1u == unsigned(NC) ? 0u - unsigned(NC - 1) * cs8 : cs8;
        if (k + 1 < n2 && unsigned(p * 8) + dc + unsigned(B.ps) * 8u != unsigned(pidx(B, i, j, k + 1) * 8))
          return fail("k step", R, n0, n1, n2, i, j, k, 2, 1);
        if (j + 1 < n1 && unsigned(p * 8) + dc + unsigned(B.rs) * 8u != unsigned(pidx(B, i, j + 1, k) * 8))
          return fail("j step", R, n0, n1, n2, i, j, k, 1, 1);
        if (i + 1 < n0 &&
            unsigned(p * 8) + dc + (i % NC == NC - 1 ? 8u : 0u) != unsigned(pidx(B, i + 1, j, k) * 8))
          return fail("i step", R, n0, n1, n2, i, j, k, 0, 1);
        // one-sided derivatives reach up to 4 points along an axis
        for (int a = 0; a < 3; ++a)
          for (int m = -4; m <= 4; ++m) {
            int q[3] = {i, j, k};
            q[a] += m;
            if (q[a] < 0 || q[a] >= B.n[a]) continue;
            if (p + axis_offset(B, int(c), i % NC, a, m) != pidx(B, q[0], q[1], q[2]))
              return fail("axis_offset", R, n0, n1, n2, i, j, k, a, m);
          }
      }
  // colour passes: slot t of interior row (j, k) for colour c
  for (int c = 0; c < NC; ++c)
    for (int k = R; k < n2 - R; ++k)
      for (int j = R; j < n1 - R; ++j) {
        const int first = R + int(unsigned(c + NC * 3 * kMaxAxisPoints - j - k - R) % NC);
        for (int t = 0; first + t * NC < n0 - R; ++t) {
          const int i = first + t * NC;
          const long long p = (long long)c * B.cs + (long long)k * B.ps + (long long)j * B.rs + first / NC + t;
          if (colour_of(B, i, j, k) != c || p != pidx(B, i, j, k) || first % NC != i % NC)
            return fail("colour slot", R, n0, n1, n2, i, j, k, 0, c);
        }
        // whole-row kernels: colour c of row (j, k), slot t
        const int s = int(unsigned(c + NC * 3 * kMaxAxisPoints - j - k) % NC);
        for (int t = 0; t < B.rl; ++t) {
          const int i = s + t * NC;
          if (i < R || i >= n0 - R) continue;
          const long long p = (long long)c * B.cs + (long long)k * B.ps + (long long)j * B.rs + t;
          if (colour_of(B, i, j, k) != c || p != pidx(B, i, j, k))
            return fail("row slot", R, n0, n1, n2, i, j, k, 0, c);
        }
      }
  int rc = 0;
  switch (NC) {
    case 2: rc = check_offs<2, 1>(B) || check_offs<2, 4>(B) || check_offs<2, 8>(B); break;
    case 3: rc = check_offs<3, 1>(B) || check_offs<3, 4>(B) || check_offs<3, 8>(B); break;
    case 4: rc = check_offs<4, 1>(B) || check_offs<4, 4>(B) || check_offs<4, 8>(B); break;
    case 5: rc = check_offs<5, 1>(B) || check_offs<5, 4>(B) || check_offs<5, 8>(B); break;
  }
  return rc;
}

}  // namespace

int main() {
  const int shapes[][3] = {{5, 5, 5}, {7, 9, 11}, {13, 13, 13}, {17, 17, 17}, {21, 19, 23},
                           {25, 26, 27}, {33, 33, 33}, {34, 35, 36}, {37, 41, 29}, {67, 75, 67}};
  for (int R = 1; R <= 4; ++R)
    for (const auto &s : shapes) {
      if (s[0] <= 2 * R || s[1] <= 2 * R || s[2] <= 2 * R) continue;
      if (check_box(R, R + 1, s[0], s[1], s[2])) return 1;
      if (R >= 2 && R < 4 && check_box(R, colours_of(2 * R, kEqPuncturesEm), s[0], s[1], s[2])) return 1;
    }
  // the magic division covers i + j + k up to three times the largest axis
  for (int nc = 2; nc <= 5; ++nc) {
    DevBox B{};
    B.n[0] = B.n[1] = B.n[2] = 8;
    B.R = nc - 1;
    set_layout(B, nc);
    for (int x = 0; x < 3 * kMaxAxisPoints; ++x)
      if (split_div(B, x) != x / nc) {
        std::printf("split_div(%d) wrong for nc=%d\n", x, nc);
        return 1;
      }
  }
  std::printf("ok\n");
  return 0;
}
