// Solver object of the drop-in Olliptic.x (reference class `elliptic`,
// include/elliptic.h:270-345, src/elliptic.cc), plus the box/field view
// classes `multigrid`, `ffunction` and `domain` (include/multigrid.h,
// include/ffunction.h, include/domain.h) over the device-resident hierarchy
// of include/olliptic_amd.h.
//
// The public methods used on the reference's path (main.cc:101-141) keep
// their names and meaning: Make, Print_Info, Solve, Print_Solution, IO_bam,
// IO_Zcode.  Field values are read through host mirrors (Get_val), which are
// fetched from the device on demand and are not on the solve path.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "interface.h"
#include "olliptic_amd.h"

namespace olliptic {

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
  double Get_Xf() const { return Get_x(b.n[0] - 1); }
  double Get_Yf() const { return Get_y(b.n[1] - 1); }
  double Get_Zf() const { return Get_z(b.n[2] - 1); }
  double Get_x(int i) const { return b.origin[0] + double(i) * b.d[0]; }
  double Get_y(int j) const { return b.origin[1] + double(j) * b.d[1]; }
  double Get_z(int k) const { return b.origin[2] + double(k) * b.d[2]; }
  double Get_Center_x() const { return b.centre[0]; }
  double Get_Center_y() const { return b.centre[1]; }
  double Get_Center_z() const { return b.centre[2]; }
  size_t Get_order() const { return size_t(b.order); }
  size_t Get_level() const { return size_t(b.level); }
  size_t Get_buffer() const { return size_t(b.ghost); }

 private:
  oll_box_desc b{};
};

class ffunction {
 public:
  ffunction(oll_hier *h, int box, int field, const domain &d) : h_(h), box_(box), field_(field), d_(d) {}
  const domain *Get_domain() const { return &d_; }
  size_t Get_nx() const { return d_.Get_Nx(); }
  size_t Get_ny() const { return d_.Get_Ny(); }
  size_t Get_nz() const { return d_.Get_Nz(); }
  // value at grid point (i, j, k); the first call fetches the box
  double Get_val(size_t i, size_t j, size_t k);
  // device -> host refresh and host -> device write-back of the whole box
  void Fetch();
  void Store();
  std::vector<double> &Host() { return host_; }

 private:
  oll_hier *h_;
  int box_, field_;
  domain d_;
  std::vector<double> host_;
  bool fetched_ = false;
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
  ffunction Get_ffunction(size_t level, size_t box, const std::string &name) const;

 private:
  oll_hier *h_;
  int levels_ = 0;
  std::vector<int> first_, count_;
  std::vector<domain> dom_;
};

class elliptic {
 public:
  explicit elliptic(interface *Oll);
  ~elliptic();
  elliptic(const elliptic &) = delete;
  elliptic &operator=(const elliptic &) = delete;

  void Make(interface *Oll);
  void Print_Info();
  // elliptic.cc:1251-1342: V-cycles until res <= tol or res stops falling;
  // writes dir/Loo.gp and dir/stdout0.out as the reference does
  void Solve();
  void Print_Solution(double Dxyz, double Lx, double Ly, double Lz);
  void IO_bam(const std::string &file);
  void IO_Zcode(const std::string &file);

  const multigrid &Get_multigrid() const { return mg_; }
  const std::vector<double> &Get_ADM_mass() const { return adm_; }
  int Get_cycles() const { return cycles_; }
  double Get_residual() const { return res_; }
  bool Ok() const { return ok_; }
  const std::string &Error() const { return err_; }

 private:
  void print_pre(int k, int top, double norm, int cycle);
  void print_post(int k, int top, double norm, int cycle);
  void compute_adm(bool print);

  oll_problem prob_{};
  oll_hier *h_ = nullptr;
  multigrid mg_;
  std::string dir_;
  bool print_1d_ = true, print_3d_ = false, quiet_ = false;
  std::vector<double> adm_;
  int cycles_ = 0;
  double res_ = 0, tol_ = 0;
  bool ok_ = true;
  std::string err_;
};

}  // namespace olliptic
