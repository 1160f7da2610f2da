// A host program written against the reference's class API (include/
// interface.h, include/elliptic.h) with the reference's names and
// signatures: first main.cc's call sequence (src/main.cc:101-141), then the
// rest of the elliptic class's hot-path methods (include/elliptic.h:288-345):
// the operator and its Newton derivative at the solution through the
// ffunction iterators, Set_Boundary, App_Bound and Vcycle again from u = 0,
// then multigrid::Transferir on polynomials its interpolation reproduces.
// Prints one "dropin:" line per check for the test-suite to read.
#include <math.h>

#include <iomanip>
#include <iostream>
#include <string>

#include "elliptic.h"
#include "interface.h"

using namespace std;

int main(int argc, char *argv[]) {
  interface Oll(argc, argv);
  elliptic E(&Oll);
  E.Print_Info();
  E.Solve();
  E.Print_Solution(Oll.Get_double_par("Dprint"), Oll.Get_double_par("Lx"),
                   Oll.Get_double_par("Ly"), Oll.Get_double_par("Lz"));
  if (Oll.Get_logic_par("io_bam")) E.IO_bam(Oll.Get_string_par("file_bam"));
  if (Oll.Get_logic_par("io_Zcode")) E.IO_Zcode(Oll.Get_string_par("file_bam"));

  const int cycles = E.Get_cycles();
  const double res = E.Get_residual();
  // the solver's own printing changes the stream's format: set it per line
  cout << scientific << setprecision(17) << "dropin: solve " << cycles << " " << res << endl;

  // N(u) - rho and dN/du on the finest level's interior
  multigrid &mg = E.Get_multigrid();
  const size_t top = mg.Get_levels() - 1;
  double worst = 0, dn_max = -1e300, dn_min = 1e300;
  size_t points = 0;
  for (size_t b = 0; b < mg.Get_num_dom(top); ++b) {
    ffunction *U = mg.Get_ffunction(top, b, "u");
    ffunction *rho = mg.Get_ffunction(top, b, "rho0");
    U->Set_Iterator(INSIDE);
    rho->Set_Iterator(INSIDE);
    do {
      worst = fmax(worst, fabs(E.OPERATOR(0, top, b) - rho->Get_val()));
      const double dn = E.duOPERATOR(0, top, b);
      dn_max = fmax(dn_max, dn);
      dn_min = fmin(dn_min, dn);
      ++points;
      rho->End();
    } while (U->End());
    E.App_Symmetry(U);
  }
  cout << scientific << setprecision(17) << "dropin: operator " << points << " " << worst << " " << dn_min << " " << dn_max << endl;

  // back to the initial state, the Robin shells of the multigrid levels,
  // and the V-cycles again: the same solve
  E.Set_Boundary();
  for (size_t l = 0; l < mg.Get_levels(); ++l)
    for (size_t b = 0; b < mg.Get_num_dom(l); ++b) E.App_Bound(l, b);
  E.Vcycle();
  E.FullMG();
  cout << scientific << setprecision(17) << "dropin: again " << E.Get_cycles() << " " << E.Get_residual() << endl;

  // multigrid::Transferir with the V-cycle's own patterns: the restriction
  // of u (INSIDE, LAGRANGE 3, all = false; elliptic.cc:1545) carries a
  // polynomial of degree 3 per axis over exactly, and the prolongation
  // (ALL, LINEAR, to_zero; elliptic.cc:1868) a multilinear one
  auto cubic = [](double x, double y, double z) {
    return 0.3 + 0.1 * x - 0.2 * y + 0.05 * z + 0.01 * x * y * z + 0.002 * x * x * x -
           0.001 * y * y * z + 1e-4 * z * z * z;
  };
  auto bilinear = [](double x, double y, double z) {
    return -0.7 + 0.03 * x + 0.02 * y - 0.01 * z + 0.004 * x * y - 0.002 * y * z + 1e-3 * x * y * z;
  };
  auto fill = [&](size_t l, const char *name, auto f, double shift) {
    for (size_t b = 0; b < mg.Get_num_dom(l); ++b) {
      ffunction *F = mg.Get_ffunction(l, b, name);
      F->Set_Iterator(ALL);
      do F->Set(f(F->Get_x(), F->Get_y(), F->Get_z()) + shift); while (F->End());
      F->Store();
    }
  };
  fill(top, "u", cubic, 0.0);
  fill(top - 1, "v0", cubic, 1.0);
  mg.Transferir(top, "u", top - 1, "v0", INSIDE, LAGRANGE, 3, false);
  size_t moved = 0;
  double err3 = 0;
  for (size_t b = 0; b < mg.Get_num_dom(top - 1); ++b) {
    ffunction *V = mg.Get_ffunction(top - 1, b, "v0");
    V->Set_Iterator(ALL);
    do {
      const double p = cubic(V->Get_x(), V->Get_y(), V->Get_z());
      const double d = fabs(V->Get_val() - p);
      if (d < 0.5) {
        ++moved;
        err3 = fmax(err3, d);
      }
    } while (V->End());
  }
  fill(top - 1, "u", bilinear, 0.0);
  fill(top, "v0", bilinear, 1.0);
  mg.Transferir(top - 1, "u", top, "v0", ALL, LINEAR, 1, true, true);
  double err1 = 0;
  size_t pts1 = 0;
  for (size_t b = 0; b < mg.Get_num_dom(top); ++b) {
    ffunction *V = mg.Get_ffunction(top, b, "v0");
    V->Set_Iterator(ALL);
    do {
      err1 = fmax(err1, fabs(V->Get_val() - bilinear(V->Get_x(), V->Get_y(), V->Get_z())));
      ++pts1;
    } while (V->End());
  }
  cout << scientific << setprecision(17) << "dropin: transfer " << moved << " " << err3 << " "
       << pts1 << " " << err1 << endl;
  return 0;
}
