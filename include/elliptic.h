// Solver object of the drop-in Olliptic.x (reference class `elliptic`,
// include/elliptic.h:270-345, src/elliptic.cc), plus the box/field view
// classes `multigrid`, `ffunction` and `domain` (include/multigrid.h,
// include/ffunction.h, include/domain.h) over the device-resident hierarchy
// of include/olliptic_amd.h.  Like the reference's, the classes live in the
// global namespace, so the reference's src/main.cc compiles against this
// header unchanged (without its MPI blocks).
//
// The public methods of the reference's elliptic class on the Punctures
// path keep their names, signatures and meaning: Make, Print_Info, Solve,
// Print_Solution, IO_bam, IO_Zcode, Set_Boundary, App_Bound, App_Symmetry,
// Vcycle, FullMG, ApproximateID, Get_Puncture_App, OPERATOR, duOPERATOR.
// As in the reference, a failure inside them prints its message and ends
// the program with status 1 (elliptic.cc:1194-1207, multigrid.cc:1031-1038);
// Ok()/Error() report the state for callers that construct the object and
// check before calling on.  Field values are read and written through host
// mirrors (ffunction::Get_val/Set/Store), fetched from the device on demand
// and refreshed after every call that changes the device state; they are
// not on the solve path.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "interface.h"
#include "olliptic_amd.h"

// ffunction.h:50 t_iterator: the regions an ffunction iterator walks (the
// reference also names each corner, edge and face block of its permuted
// storage; the solve path uses these three)
enum t_iterator { ALL, INSIDE, BOUNDARY };
// interpolation methods (ffunction.h:82-83); transfers take LINEAR and
// LAGRANGE
enum m_interpol { LINEAR, LAGRANGE, NEWTON, HERMITE, AVERAGE };

class domain {
 public:
  domain() = default;
  explicit domain(const oll_box_desc &d) : b(d) {}
  size_t Get_Nx() const { return size_t(b.n[0]); }
  size_t Get_Ny() const { return size_t(b.n[1]); }
  size_t Get_Nz() const { return size_t(b.n[2]); }
  size_t Get_N() const { return Get_Nx() * Get_Ny() * Get_Nz(); }
  double Get_dx() const { return b.d[0]; }
  double Get_dy() const { return b.d[1]; }
  double Get_dz() const { return b.d[2]; }
  double Get_Xi() const { return b.origin[0]; }
  double Get_Yi() const { return b.origin[1]; }
  double Get_Zi() const { return b.origin[2]; }
  double Get_Xf() const { return b.end[0]; }
  double Get_Yf() const { return b.end[1]; }
  double Get_Zf() const { return b.end[2]; }
  double Get_x(int i) const { return b.origin[0] + double(i) * b.d[0]; }
  double Get_y(int j) const { return b.origin[1] + double(j) * b.d[1]; }
  double Get_z(int k) const { return b.origin[2] + double(k) * b.d[2]; }
  double Get_Center_x() const { return b.centre[0]; }
  double Get_Center_y() const { return b.centre[1]; }
  double Get_Center_z() const { return b.centre[2]; }
  size_t Get_order() const { return size_t(b.order); }
  size_t Get_level() const { return size_t(b.level); }
  size_t Get_buffer() const { return size_t(b.ghost); }
  // domain.cc:1493-1675 Get_i/Get_j/Get_k: the first grid index whose
  // coordinate is at or above x (within 3 DBL_EPSILON); 1 at or below the
  // box, n - 1 at or above it
  size_t Get_i(double x) const { return index_of(0, x); }
  size_t Get_j(double y) const { return index_of(1, y); }
  size_t Get_k(double z) const { return index_of(2, z); }
  // domain.h:767-851
  bool x_is_in(double x) const { return Get_Xi() <= x && x <= Get_Xf(); }
  bool y_is_in(double y) const { return Get_Yi() <= y && y <= Get_Yf(); }
  bool z_is_in(double z) const { return Get_Zi() <= z && z <= Get_Zf(); }
  bool is_in(double x, double y, double z) const;
  // strictly inside the box less its buffer of order/2 points
  bool is_in_nb(double x, double y, double z) const;
  // domain.cc:1947-1954: this box lies strictly inside D on every axis
  bool operator<(const domain &D) const;

 private:
  size_t index_of(int a, double x) const;
  double lo(int a) const { return b.origin[a]; }
  double hi(int a) const { return b.end[a]; }
  oll_box_desc b{};
};

class ffunction {
 public:
  // gen: the owner's device-state counter; host mirrors older than it are
  // fetched again
  ffunction(oll_hier *h, int box, int field, const domain &d,
            std::shared_ptr<unsigned long> gen = nullptr)
      : h_(h), box_(box), field_(field), d_(d), gen_(std::move(gen)) {}
  const domain *Get_domain() const { return &d_; }
  int Field() const { return field_; }
  int Box() const { return box_; }
  size_t Get_nx() const { return d_.Get_Nx(); }
  size_t Get_ny() const { return d_.Get_Ny(); }
  size_t Get_nz() const { return d_.Get_Nz(); }
  size_t Get_N() const { return d_.Get_N(); }
  size_t Get_order() const { return d_.Get_order(); }
  // value at grid point (i, j, k) (ffunction.h:409)
  double Get_val(size_t i, size_t j, size_t k);
  // ffunction.h:391-394 Set_Iterator, :537-545 End, :480 Get_val(),
  // :526 Set(y), :432-478 Get_x/y/z: a walk over the points of a region,
  // used as  U->Set_Iterator(INSIDE); do { ... } while (U->End());
  // The walk is lexicographic (i fastest); the reference's permuted storage
  // visits the same points in another order, which pointwise loops do not
  // see.
  void Set_Iterator(t_iterator it = INSIDE);
  bool End();
  double Get_val();
  void Set(double y);
  void Set(size_t i, size_t j, size_t k, double y);
  double Get_x() const { return d_.Get_x(int(ci_)); }
  double Get_y() const { return d_.Get_y(int(cj_)); }
  double Get_z() const { return d_.Get_z(int(ck_)); }
  // index of the iterator's point in lexicographic order
  size_t Get_Index() const { return ci_ + Get_nx() * (cj_ + Get_ny() * ck_); }
  // device -> host refresh and host -> device write-back of the whole box
  void Fetch();
  void Store();
  std::vector<double> &Host() { return host_; }

 private:
  bool in_region(size_t i, size_t j, size_t k) const;
  void ensure();
  oll_hier *h_;
  int box_, field_;
  domain d_;
  std::shared_ptr<unsigned long> gen_;
  unsigned long seen_ = 0;
  std::vector<double> host_;
  bool fetched_ = false;
  t_iterator it_ = INSIDE;
  size_t ci_ = 0, cj_ = 0, ck_ = 0;
  bool done_ = true;
};

class multigrid {
 public:
  explicit multigrid(oll_hier *h = nullptr);
  size_t Get_levels() const { return size_t(levels_); }
  size_t Get_num_dom(size_t level) const;
  // multigrid.cc:97-116: a box index beyond the level's count falls back
  // to the level's first box
  size_t Get_dom_index(size_t level, size_t box) const;
  const domain &Get_domain(size_t level, size_t box) const;
  // multigrid.h:98-101: the named field of box `box` of `level` (u, b, c,
  // d, e, rho0, r0, v0, w0, ww0); the multigrid owns it
  ffunction *Get_ffunction(size_t level, size_t box, const std::string &name) const;
  // multigrid.cc:122-141: field from_var of the boxes of level `from` into
  // to_var of the boxes of level `to` over `iter`'s points (oll_transfer);
  // NEWTON, HERMITE and AVERAGE end the program with a message
  void Transferir(size_t from, std::string from_var, size_t to, std::string to_var,
                  t_iterator iter = ALL, m_interpol method = LINEAR, size_t ord = 1,
                  bool all = true, bool to_zero = false);
  // the device state changed: host mirrors are fetched again when read
  void Touch() { ++*gen_; }
  unsigned long Generation() const { return *gen_; }
  // multigrid.cc:35-48: the box that holds every other one
  size_t Get_l0() const { return l0_; }
  // multigrid.cc:223-241: the last box, over levels and their boxes in
  // order, whose interior less its buffer holds the point; box l0 otherwise
  size_t Box_Of(double x, double y, double z) const;
  // Lagrange interpolation (order + 1 points per axis) of field `name` at
  // xyz[3q..3q+2] in box[q], one device launch for all points
  std::vector<double> Interpol(const std::string &name, const std::vector<int> &box,
                               const std::vector<double> &xyz, size_t order) const;
  // multigrid.h:312-337 + ffunction.cc:819-1160: per level box a 1-D cut
  // through its middle (or through x1, x2 where the box holds them) to
  // <dir><name>-l<l>b<b>-0.x/.y/.z
  void Print_1D(const std::string &dir, const std::string &name, size_t l, int cut, double x1,
                double x2, bool interpolar, size_t order) const;
  // multigrid.cc:151-548: a cut over [-LD/2, LD/2] with step _dx (LD/1000
  // when _dx <= 0) through the composite grid to <dir><name>.X/.Y/.Z
  void Print_Interpol_1D(const std::string &dir, const std::string &name, double LD, double _dx,
                         int cut, double x1, double x2, size_t order) const;
  // multigrid.cc:554-958: a plane over [-LD1/4, LD1/4] x [-LD2/4, LD2/4]
  // (steps LD/200 when _dx <= 0) to <dir><name>.YZ/.XZ/.XY
  void Print_Interpol_2D(const std::string &dir, const std::string &name, double LD1,
                         double LD2, double _dx1, double _dx2, int cut, double x1,
                         size_t order) const;
  // multigrid.cc:991-1104: each "x y z value" record of a BAM file gets u
  // at (x, y, z) written over it in the widths 16/17/17/24; false when a
  // point lies outside the outermost box
  bool Print_IO_BAM(const std::string &name, const std::string &var, size_t order) const;
  // multigrid.cc:1123-1222: "x y z u(x,y,z)" per record of name into name_m
  bool Print_IO_Zcode(const std::string &name, const std::string &var, size_t order) const;

 private:
  oll_hier *h_;
  int levels_ = 0;
  size_t l0_ = 0;
  std::vector<int> first_, count_;
  std::vector<domain> dom_;
  std::shared_ptr<unsigned long> gen_ = std::make_shared<unsigned long>(1);
  mutable std::map<std::pair<int, int>, std::unique_ptr<ffunction>> fields_;
};

// cut directions (ffunction.h `cut`)
enum { CUT_X = 0, CUT_Y = 1, CUT_Z = 2 };

class elliptic {
 public:
  explicit elliptic(interface *Oll);
  ~elliptic();
  elliptic(const elliptic &) = delete;
  elliptic &operator=(const elliptic &) = delete;

  // elliptic.cc:31-221: parameters, Set_Grids, Set_Equation,
  // Set_Coefficients (the hierarchy built and filled on the device)
  void Make(interface *Oll);
  void Print_Info();
  // elliptic.h:331-334: the method of -m (V-cycles or the analytic
  // approximation); writes dir/Loo.gp and dir/stdout0.out as the reference
  void Solve();
  // elliptic.cc:1251-1342: V-cycles until res <= tol or res stops falling
  void Vcycle();
  // elliptic.h:323-325: empty in the reference too
  void FullMG() {}
  // elliptic.cc:1977-2222 (-m AppPuncture): u set to the analytic
  // approximation of each puncture, no iteration
  void ApproximateID();
  // elliptic.cc:2232-2444: prints the approximation at each puncture's node
  void Get_Puncture_App();
  // boundary.cc:8-84: u = 0 everywhere, rho = the Robin value on the shell
  void Set_Boundary();
  // boundary.cc:447-907: the Robin update of the ghost shell of box ibox of
  // level l
  void App_Bound(size_t l, size_t ibox);
  // boundary.cc:93-441: mirror fills of symmetry domains; a full domain
  // (the only one this solver builds) has none, as in the reference when no
  // bound is -2/-3
  void App_Symmetry(ffunction *U) { (void)U; }
  // elliptic.h:337-345: the equation's operator N(u) and its derivative
  // dN/du at the point itself, at the current iterator point of box ibox of
  // level l (the point of its "u" ffunction's walk, an INSIDE point); one
  // variable (lvar 0) for the puncture equations.  Evaluated on the device
  // for the whole level, then read from a host mirror.
  double OPERATOR(size_t lvar, size_t l, size_t ibox);
  double duOPERATOR(size_t lvar, size_t l, size_t ibox);
  void Print_Solution(double pd = 0.1, double Lx = 1.0, double Ly = 1.0, double Lz = 1.0);
  // output.cc:377-451: rewrites the BAM initial-data files <file><l>,
  // <file><l>a .. <file><l>e (levels l = 0, 1, ... while a level has one)
  // in place with u at their points; a point outside the outermost box ends
  // the program with status 1 after the reference's message
  void IO_bam(std::string file);
  // output.cc:455-617: writes <name>_m beside each Zcode point file
  // <file><ll>-0<b>.mgid and <file>SH-{x,y,z}{m,p}.mgid
  void IO_Zcode(std::string file);

  multigrid &Get_multigrid() { return mg_; }
  const std::vector<double> &Get_ADM_mass() const { return adm_; }
  int Get_cycles() const { return cycles_; }
  double Get_residual() const { return res_; }
  bool Ok() const { return ok_; }
  const std::string &Error() const { return err_; }

 private:
  void print_pre(int k, int top, double norm, int cycle);
  void print_post(int k, int top, double norm, int cycle);
  void compute_adm(bool print);
  bool export_bam(const std::string &file);
  bool export_zcode(const std::string &file);
  double op_at(int which, size_t lvar, size_t l, size_t ibox);
  // records the error and ends the program with status 1
  void failed(const std::string &msg, bool print = true);

  oll_problem prob_{};
  oll_hier *h_ = nullptr;
  multigrid mg_;
  std::string dir_;
  bool print_1d_ = true, print_3d_ = false, print_interpol_ = true, interpol_output_ = true;
  bool quiet_ = false, verbose_ = true, fit_approx_ = false;
  std::vector<double> adm_;
  int cycles_ = 0;
  double res_ = 0, tol_ = 0;
  bool ok_ = true;
  std::string err_;
  // OPERATOR / duOPERATOR: [which] the level evaluated, the state counter
  // it was evaluated at, and the value per box
  int op_level_[2] = {-1, -1};
  unsigned long op_gen_[2] = {0, 0};
  std::map<int, std::vector<double>> op_val_[2];
};
