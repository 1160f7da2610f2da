// Olliptic.x: drop-in command line of the reference (src/main.cc:83-169)
// on the MI355X solver.  The call sequence is the reference's: the class
// constructors and methods report command-line errors (usage, status 0) and
// solver errors (message, status 1) themselves, as the reference's do.
#include "elliptic.h"
#include "interface.h"

int main(int argc, char *argv[]) {
  interface Oll(argc, argv);
  elliptic E(&Oll);
  E.Print_Info();
  E.Solve();
  E.Print_Solution(Oll.Get_double_par("Dprint"), Oll.Get_double_par("Lx"),
                   Oll.Get_double_par("Ly"), Oll.Get_double_par("Lz"));
  if (Oll.Get_logic_par("io_bam")) E.IO_bam(Oll.Get_string_par("file_bam"));
  if (Oll.Get_logic_par("io_Zcode")) E.IO_Zcode(Oll.Get_string_par("file_bam"));
  return 0;
}
