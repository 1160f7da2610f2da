// Host engine: one Hier per request (its box hierarchy resident in HBM) and
// Schedules, the launch sequences of one V-cycle for a set of hierarchies.
//
// A schedule aligns its members by depth below their top level: at depth d
// every member works on its own level top - d, so one launch per (stage,
// stencil order, equation, colour) covers the boxes of all members.  The
// coarsest-level solves of all members run together, one block each.  The
// control flow of one member is the reference's FASN recursion
// (elliptic.cc:1469-1968); the outer loop and stopping rule are Vcycle's
// (elliptic.cc:1251-1342).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/olliptic_amd.h"
#include "geometry.hpp"
#include "kernels.hpp"

namespace oll {

struct Hier {
  oll_problem prob{};
  Hierarchy geo;
  std::vector<DevBox> hbox;      // host copies, device field pointers
  DevBox *dbox = nullptr;        // device copies
  double *slab = nullptr;
  float *slab32 = nullptr;       // mixed precision only
  int *ipool = nullptr;
  unsigned char *bpool = nullptr;
  double *dpool = nullptr;
  DevXfer *dxfer = nullptr;
  size_t bytes = 0;              // HBM held
  struct XRef {
    int off = 0, cnt = 0;        // into dxfer
    // index bounding box [lo, hi) of the target points inside some
    // source's non-ghost part (empty when lo >= hi on an axis)
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  };
  // [fine level k][target box within its level]
  std::vector<std::vector<XRef>> ru, rw, pv, gf;
  // per box: bounding box {lo[3], hi[3]} of the nodes prolongations read
  // from it (u and ww; one node of margin), empty when lo >= hi
  std::vector<std::array<int, 6>> prolong_nodes;
  // per box: source of a defect restriction off the coarser lattice, whose
  // defect is filtered before the restriction (k_filter_defect)
  std::vector<char> filt_src;
  unsigned long long *norms = nullptr;  // nslots * kNormWords
  int nslots = 0;
  int *iters = nullptr;
  PunctureArgs pa{};
  double eps_solv = 0, eps_coarse = 0, coarse_omega = 1;
  std::vector<std::vector<double>> colour_pts;  // [level][colour]
  std::vector<std::vector<double>> colour_rho;  // of them inside the boxes' rho boxes
  // V-cycle state
  int cycle = 0;
  double res = 1e12, res_prev = 1e12;
  double best = 1e300;           // lowest top-level residual so far
  int best_cycle = 0;
  bool done = false;
  int status = 0;                // 0 ok, <0 error code
  std::vector<double> hist;      // (cycle, pre-smoothing norm) pairs
  std::vector<double> last_slots;
  long long coarse_sweeps = 0;
  std::vector<unsigned long long> raw;  // staging of the norm slots
  int raw_iters = 0;

  int top() const { return geo.levels - 1; }
  int level_order(int l) const { return hbox[size_t(geo.level_first[size_t(l)])].order; }
  bool mixed() const { return prob.precision == OLL_PREC_MIXED; }
  unsigned long long *slot(int s) const { return norms + size_t(s) * kNormWords; }
  static int slot_pre(int k) { return 2 * k; }
  static int slot_post(int k) { return 2 * k + 1; }
  int slot_scratch() const { return nslots - 1; }
};

// builds geometry, allocates, uploads transfer tables, evaluates the
// coefficients; returns an OLL_* code and sets err
int hier_build(Hier &H, const oll_problem &p, hipStream_t s, std::string &err);
void hier_release(Hier &H);
// per colour pass, the points of level l inside their boxes' rho boxes
std::vector<double> level_rho_points(const Hier &H, int l);
// boxes first .. first + count - 1 read rho everywhere from now on (after a
// caller writes the field); waits for the descriptor copy on s
int hier_rho_whole(Hier &H, int first, int count, hipStream_t s, std::string &err);
int hier_reset(Hier &H, hipStream_t s, std::string &err);
// host-only footprint of a problem's hierarchy (bytes of HBM)
int problem_footprint(const oll_problem &p, size_t &bytes, std::string &err);

struct Launch {
  enum Kind {
    ZERO, ROBIN, SMOOTH, RESIDUAL, RESTRICT_U, RESTRICT_DEFECT, COARSE, PROLONG, GHOST,
    DEFECT32, SMOOTH32, UPDATE32, FILTER, SMOOTH_FUSED, COPY_SHELL, CORR_V
  };
  Kind kind;
  int order = 2, eq = 0, colour = 0, mode = 0;
  size_t off = 0;     // items in the schedule blob
  int n = 0;          // items
  long long total = 0;
  int max_row = 0;    // row kernels: threads per row
  int threads = kRowThreads;  // colour passes: threads per workgroup
  int max_pts = 0, max_int = 0;  // COARSE: largest box, colour slot list; ROBIN: edge cross-sections
  int pair = 0, parity = 0;  // SMOOTH_FUSED: two passes or one; 0 reads u, 1 reads r
  double bytes = 0;   // algorithmic bytes (smoother launches, FP64 or FP32)
  const DevBox *box1 = nullptr;  // row launches over a single box: its descriptor
};

struct Schedule {
  std::vector<Launch> launches;
  std::vector<unsigned char> blob;
  unsigned char *dblob = nullptr;
  size_t dcap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool uploaded = false;
  int runs = 0;  // times run since built: graphs are captured on the second run

  template <typename T>
  size_t add(const std::vector<T> &items) {
    size_t off = (blob.size() + 255) & ~size_t(255);
    blob.resize(off + items.size() * sizeof(T));
    if (!items.empty()) std::memcpy(blob.data() + off, items.data(), items.size() * sizeof(T));
    return off;
  }
  void clear();
};

// events around every smoother launch (SMOOTH or SMOOTH32) when on
struct Timing {
  bool on = false;
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  double bytes = 0;
  std::vector<double> launch_bytes;  // per timed launch, in issue order
  // 2 * order + (0: FP64 smoother colour pass, 1: FP32 correction colour pass)
  std::vector<int> launch_kind;
};

// the full V-cycle of every member; all members must share nu1, nu2
int build_vcycle(const std::vector<Hier *> &members, Schedule &S, std::string &err);
// single stages of one hierarchy on one level (tests, tools and the class
// API); box >= 0 restricts the Robin stages to that box of the level
int build_stage(Hier &H, int level, int stage, int nu, Schedule &S, std::string &err, int box = -1);
enum Stage { STAGE_SMOOTH = 0, STAGE_RESIDUAL, STAGE_ROBIN, STAGE_BOUNDARY_OP, STAGE_DERIVATIVE };

// multigrid::Transferir (multigrid.cc:122-141): field ffield of the boxes
// of level `from` into field tfield of the boxes of level `to` over region
// (0 all, 1 interior, 2 ghost shell), linear (method 0) or Lagrange of
// order ord (method 1, ord + 1 nodes per axis, ord 2..9); transfer tables
// are built for the call and freed after it
int transfer_fields(Hier &H, hipStream_t s, int from, int ffield, int to, int tfield, int region,
                    int method, int ord, int all, int to_zero, std::string &err);
// colour order of box b (0: ascending, 1: descending) for nc colours
int colour_flip(const Hierarchy &g, const Box &b, int nc);
// FP64 sweeps of the puncture, scalar-field, Poisson and Brill equations
// as fused, double-buffered colour-pass pairs (default off); off issues
// one launch per colour.  Process-wide; schedules built afterwards use it.
void set_fused_sweeps(bool on);
// ghost fills one thread per shell column and defect filters one thread
// per run of points along x (default), or both one thread per point; the
// same bits either way.  Process-wide; schedules built afterwards use it.
void set_shared_loads(bool on);
bool shared_loads();
// whether ghost fills of boxes of this stencil order go by columns
bool ghost_columns_for(int order);
bool fused_sweeps();

int upload_schedule(Schedule &S, hipStream_t s, std::string &err);
// runs the schedule; with graphs = true it is captured as a graph on its
// second run (a schedule run once is cheaper to issue directly) and replayed
// from then on
int run_schedule(Schedule &S, hipStream_t s, Timing *t, bool graphs, std::string &err);
void release_schedule(Schedule &S);

// enqueue the copy of a member's norm slots (read_norms), then after a
// stream synchronization decode them into H.last_slots (decode_norms);
// apply_cycle also records the V-cycle history and applies the stopping
// rule; finish_cycle does all three for one member
int read_norms(Hier &H, hipStream_t s, std::string &err);
void decode_norms(Hier &H);
int apply_cycle(Hier &H, std::string &err);
int finish_cycle(Hier &H, hipStream_t s, std::string &err);
// V-cycle schedule whose member h starts at level start[h] (FASN(start))
int build_vcycle_from(const std::vector<Hier *> &members, const std::vector<int> &start,
                      Schedule &S, std::string &err);
// ADM masses (elliptic.cc:2447-2611); surface may be null.  adm_mass
// allocates its device staging per call; adm_mass_with works in the
// caller's buffers: pbox (num_bh ints), pxyz (3 num_bh), up (num_bh), and,
// with surface, ctr (3 nbox) and sph (nbox kAdmNt^2 doubles)
constexpr int kAdmNt = 4 * 25 + 1;  // sphere samples per angle (elliptic.cc:2471-2480)
struct AdmBuffers {
  int *pbox = nullptr;
  double *pxyz = nullptr, *up = nullptr, *ctr = nullptr, *sph = nullptr;
};
int adm_mass(Hier &H, hipStream_t s, double *puncture, double *surface, std::string &err);
int adm_mass_with(Hier &H, hipStream_t s, const AdmBuffers &d, double *puncture, double *surface,
                  std::string &err);
// elliptic.cc:2232-2444 Get_Puncture_App values (host arithmetic)
void puncture_approx(const Hier &H, double *u_at_p);
// Lagrange interpolation of `field` at npts host points (box[q], xyz[3q..3q+2])
int interp_points(Hier &H, hipStream_t s, int field, int npts, const int *box, const double *xyz,
                  int ord, double *out, std::string &err);

}  // namespace oll
