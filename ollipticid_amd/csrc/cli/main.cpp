// Olliptic.x: drop-in command line of the reference (src/main.cc:83-169)
// on the MI355X solver.
#include <iostream>

#include "elliptic.h"
#include "interface.h"

int main(int argc, char *argv[]) {
  olliptic::interface Oll(argc, argv);
  if (!Oll.Valid()) {
    // the reference prints its help and exits 0 (interface.cc:341-351)
    std::cout << "Olliptic.x (MI355X): " << Oll.Error() << "\n"
              << "usage: Olliptic.x -eq Punctures|Punctures_scalar_field -B Robin [-BA a -Bn n]\n"
              << "       -NumBH k -bhmp<i> m -bhx<i> .. -bhsz<i> .. -L L -N N -O 2|4|6|8 -l levels\n"
              << "       -s tol -nu1 n -nu2 n -o dir [-psc c -phi0 p -r0 r -sigma s -a2 a]\n"
              << "       [-p1d yes|no] [-pint yes|no] [-int yes|no] [-Dp d] [-p3d yes|no]\n"
              << "       [-iob yes -ioZ yes -iof file] [-so yes|partial|no] [-prec fp64|mixed]\n";
    return 0;
  }
  olliptic::elliptic E(&Oll);
  if (!E.Ok()) {
    std::cerr << "Error: " << E.Error() << "\n";
    return 1;
  }
  E.Print_Info();
  E.Solve();
  if (!E.Ok()) return 1;
  E.Print_Solution(Oll.Get_double_par("Dprint"), Oll.Get_double_par("Lx"),
                   Oll.Get_double_par("Ly"), Oll.Get_double_par("Lz"));
  // the reference exits with status 1 when an export point lies outside
  // the grid (multigrid.cc:1031-1038, 1166-1177)
  if (Oll.Get_logic_par("io_bam") && !E.IO_bam(Oll.Get_string_par("file_bam"))) return 1;
  if (Oll.Get_logic_par("io_Zcode") && !E.IO_Zcode(Oll.Get_string_par("file_bam"))) return 1;
  return 0;
}
