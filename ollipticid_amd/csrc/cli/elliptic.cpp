// Solver object and field views of Olliptic.x; see include/elliptic.h.
#include "elliptic.h"

#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

namespace olliptic {

namespace fs = std::filesystem;

// ---------------------------------------------------------------- views
double ffunction::Get_val(size_t i, size_t j, size_t k) {
  if (!fetched_) Fetch();
  return host_[i + Get_nx() * (j + Get_ny() * k)];
}

void ffunction::Fetch() {
  host_.resize(d_.Get_N());
  if (oll_get_field(h_, box_, field_, host_.data()) != OLL_OK)
    throw std::runtime_error(oll_last_error());
  fetched_ = true;
}

void ffunction::Store() {
  if (oll_set_field(h_, box_, field_, host_.data()) != OLL_OK)
    throw std::runtime_error(oll_last_error());
}

multigrid::multigrid(oll_hier *h) : h_(h) {
  if (!h) return;
  levels_ = oll_num_levels(h);
  for (int l = 0; l < levels_; ++l) {
    int f = 0, n = 0;
    oll_level_boxes(h, l, &f, &n);
    first_.push_back(f);
    count_.push_back(n);
  }
  for (int b = 0; b < oll_num_boxes(h); ++b) {
    oll_box_desc d;
    oll_box_info(h, b, &d);
    dom_.emplace_back(d);
  }
}

size_t multigrid::Get_num_dom(size_t level) const {
  return level < count_.size() ? size_t(count_[level]) : 0;
}

size_t multigrid::Get_dom_index(size_t level, size_t box) const {
  if (level >= first_.size()) return 0;
  return size_t(first_[level]) + (box < size_t(count_[level]) ? box : 0);
}

const domain &multigrid::Get_domain(size_t level, size_t box) const {
  return dom_[Get_dom_index(level, box)];
}

ffunction multigrid::Get_ffunction(size_t level, size_t box, const std::string &name) const {
  static const char *names[OLL_NFIELDS] = {"u", "b", "c", "d", "e", "rho0", "r0", "v0", "w0", "ww0"};
  int f = -1;
  for (int i = 0; i < OLL_NFIELDS; ++i)
    if (name == names[i]) f = i;
  if (name == "rho") f = OLL_F_RHO;
  if (f < 0) throw std::runtime_error("unknown field " + name);
  const size_t ib = Get_dom_index(level, box);
  return ffunction(h_, int(ib), f, dom_[ib]);
}

// ---------------------------------------------------------------- elliptic
elliptic::elliptic(interface *Oll) { Make(Oll); }

elliptic::~elliptic() {
  if (h_) oll_hier_destroy(h_);
}

// elliptic.cc:31-221 Make; interface.cc:1497-1527 output directory
void elliptic::Make(interface *Oll) {
  oll_problem_default(&prob_);
  prob_.eq = Oll->Get_Equation() == PUNCTURES_SCALAR ? OLL_EQ_PUNCTURES_SCALAR : OLL_EQ_PUNCTURES;
  const std::vector<double> m = Oll->Get_bhmp(), x = Oll->Get_bhx(), y = Oll->Get_bhy(),
                            z = Oll->Get_bhz(), px = Oll->Get_bhpx(), py = Oll->Get_bhpy(),
                            pz = Oll->Get_bhpz(), sx = Oll->Get_bhsx(), sy = Oll->Get_bhsy(),
                            sz = Oll->Get_bhsz();
  prob_.num_bh = int(m.size());
  for (size_t i = 0; i < m.size() && i < size_t(OLL_MAX_BH); ++i)
    prob_.bh[i] = oll_puncture{m[i], x[i], y[i], z[i], px[i], py[i], pz[i], sx[i], sy[i], sz[i]};
  prob_.L = Oll->Get_double_par("Lx");
  prob_.N = int(Oll->Get_integer_par("nx"));
  prob_.levels = int(Oll->Get_integer_par("levels"));
  prob_.order = int(Oll->Get_integer_par("order"));
  prob_.nu1 = int(Oll->Get_integer_par("nu1"));
  prob_.nu2 = int(Oll->Get_integer_par("nu2"));
  prob_.tol = Oll->Get_double_par("tolerance");
  prob_.robin_a = Oll->Get_double_par("robin_a");
  prob_.robin_n = int(Oll->Get_integer_par("robin_n"));
  prob_.ps_case = int(Oll->Get_integer_par("ps_case"));
  prob_.ps_r0 = Oll->Get_double_par("ps_r0");
  prob_.ps_sigma = Oll->Get_double_par("ps_sigma");
  prob_.ps_phi0 = Oll->Get_double_par("ps_phi0");
  prob_.ps_a2 = Oll->Get_double_par("ps_a2");
  prob_.precision = Oll->Get_string_par("precision") == "mixed" ? OLL_PREC_MIXED : OLL_PREC_FP64;
  tol_ = prob_.tol < 2 * 2.220446049250313e-16 ? 2 * 2.220446049250313e-16 : prob_.tol;
  print_1d_ = Oll->Get_logic_par("print_1d");
  print_3d_ = Oll->Get_logic_par("print_3d");
  quiet_ = Oll->Get_TOutput() == NO;

  const std::string name = Oll->Get_string_par("dir_name");
  dir_ = name + "/";
  std::error_code ec;
  fs::remove_all(name + "_prev", ec);
  if (fs::exists(name)) fs::rename(name, name + "_prev", ec);
  fs::create_directories(name, ec);
  std::ofstream(dir_ + name + ".run") << Oll->Command_Line() << "\n";

  if (oll_hier_create(&prob_, &h_) != OLL_OK) {
    ok_ = false;
    err_ = oll_last_error();
    return;
  }
  mg_ = multigrid(h_);
}

// output.cc:1030-1345 (the puncture and box tables)
void elliptic::Print_Info() {
  if (!ok_ || quiet_) return;
  std::cout << std::scientific << std::setprecision(7);
  std::cout << "Calculating initial data for " << prob_.num_bh << " puncture(s)\n";
  for (int i = 0; i < prob_.num_bh; ++i) {
    const oll_puncture &p = prob_.bh[i];
    std::cout << "Puncture " << i + 1 << "  mass " << p.m << "  position " << p.x << " " << p.y
              << " " << p.z << "  momentum " << p.px << " " << p.py << " " << p.pz << "  spin "
              << p.sx << " " << p.sy << " " << p.sz << "\n";
  }
  for (size_t l = 0; l < mg_.Get_levels(); ++l)
    for (size_t b = 0; b < mg_.Get_num_dom(l); ++b) {
      const domain &d = mg_.Get_domain(l, b);
      std::cout << "  level: " << l << " | box: " << b << "\n"
                << std::scientific << std::setprecision(3) << "  nx = " << d.Get_Nx()
                << ", \t\tny = " << d.Get_Ny() << ", \t\tnz = " << d.Get_Nz() << "\n"
                << "  dx = " << d.Get_dx() << ", \tdy = " << d.Get_dy() << ", \tdz = " << d.Get_dz()
                << "\n"
                << std::fixed << std::setprecision(2) << "  x-->[" << d.Get_Xi() << ","
                << d.Get_Xf() << "]\ty-->[" << d.Get_Yi() << "," << d.Get_Yf() << "]\tz-->["
                << d.Get_Zi() << "," << d.Get_Zf() << "]\n";
    }
  std::cout << std::defaultfloat;
}

// output.cc:1355-1473 Print_Pre_Cycle
void elliptic::print_pre(int k, int top, double norm, int cycle) {
  std::ostringstream o;
  o << std::scientific << std::setprecision(8);
  if (k == top) {
    o << "\n\nCycle: " << cycle << "\n";
    for (int j = 0; j <= top; ++j) o << "========================";
    o << "\n";
  }
  for (int j = 0; j < k; ++j) o << "|\t\t\t";
  o << "|Loo[" << k << "] : " << norm;
  for (int j = k + 1; j <= top + 1; ++j) o << "|\t\t\t";
  o << "\n";
  if (!quiet_) std::cout << o.str();
  std::ofstream(dir_ + "stdout0.out", std::ios::app) << o.str();
  std::ofstream nf(dir_ + "Loo.gp", std::ios::app);
  nf << std::scientific << std::setprecision(18) << cycle << "\t" << norm << "\n";
}

// output.cc:1484-1594 Print_Post_Cycle
void elliptic::print_post(int k, int top, double norm, int cycle) {
  (void)cycle;
  std::ostringstream o;
  o << std::scientific << std::setprecision(8);
  for (int j = 0; j < k; ++j) o << "|\t\t\t";
  o << "|Loo[" << k << "] : " << norm;
  for (int j = k + 1; j <= top + 1; ++j) o << "|\t\t\t";
  o << "\n";
  if (k == top) {
    for (int j = 0; j <= top; ++j) o << "========================";
    o << "\n" << norm << "  -->  " << tol_ << "\n";
  }
  if (!quiet_) std::cout << o.str();
  std::ofstream(dir_ + "stdout0.out", std::ios::app) << o.str();
}

// elliptic.cc:2447-2611 (values) and its stdout lines
void elliptic::compute_adm(bool print) {
  adm_.assign(size_t(prob_.num_bh), 0.0);
  std::vector<double> surf(size_t(oll_num_boxes(h_)));
  if (oll_adm_mass(h_, adm_.data(), surf.data()) != OLL_OK) {
    ok_ = false;
    err_ = oll_last_error();
    return;
  }
  if (!print || quiet_) return;
  std::cout << std::scientific << std::setprecision(8) << "Puncure method: \n";
  for (size_t i = 0; i < adm_.size(); ++i) std::cout << "ADM mass[" << i << "] = " << adm_[i] << "\n";
  for (size_t l = 0; l < mg_.Get_levels(); ++l)
    for (size_t b = 0; b < mg_.Get_num_dom(l); ++b) {
      const domain &d = mg_.Get_domain(l, b);
      const double L = std::min(d.Get_Xf() - d.Get_Xi(),
                                std::min(d.Get_Yf() - d.Get_Yi(), d.Get_Zf() - d.Get_Zi()));
      std::cout << "Level " << l << " box " << b << " r0= " << 0.375 * L
                << " AMDmass= " << surf[mg_.Get_dom_index(l, b)] << "\n";
    }
}

// elliptic.cc:1251-1342 Vcycle with the reference's per-cycle logging
void elliptic::Solve() {
  if (!ok_) return;
  const int top = int(mg_.Get_levels()) - 1;
  std::vector<double> pre(size_t(top) + 1), post(size_t(top) + 1);
  double res = 1e12, res_prev;
  int cycle = 0;
  do {
    res_prev = res;
    if (oll_fas(h_, top, &res) != OLL_OK) {
      ok_ = false;
      err_ = oll_last_error();
      std::cerr << "Error: " << err_ << "\n";
      return;
    }
    oll_last_norms(h_, pre.data(), post.data(), top + 1);
    for (int k = top; k >= 0; --k) print_pre(k, top, pre[size_t(k)], cycle);
    for (int k = 1; k <= top; ++k) print_post(k, top, post[size_t(k)], cycle);
    compute_adm(true);
    ++cycle;
  } while (res > tol_ && res < res_prev);
  cycles_ = cycle;
  res_ = res;
  if (!quiet_) {
    std::cout << std::scientific << std::setprecision(8);
    for (int j = 0; j < prob_.num_bh; ++j) std::cout << "mp [" << j << "] = " << prob_.bh[j].m << "\n";
  }
}

// output.cc:624-729: -p3d writes every box (ffunction.cc:411-467 format);
// the interpolated 1-D/2-D cuts are not produced by this build
void elliptic::Print_Solution(double, double, double, double) {
  if (!ok_) return;
  if (!quiet_) std::cout << "Printing..." << std::endl;
  if (print_3d_) {
    for (size_t l = 0; l < mg_.Get_levels(); ++l)
      for (size_t b = 0; b < mg_.Get_num_dom(l); ++b) {
        ffunction u = mg_.Get_ffunction(l, b, "u");
        const domain &d = *u.Get_domain();
        std::ostringstream fn;
        fn << dir_ << "u-l" << l << "b" << b << "-0.xyz";
        std::ofstream f(fn.str(), std::ios::app);
        f << std::scientific << std::setprecision(12);
        for (size_t k = 0; k < d.Get_Nz(); ++k)
          for (size_t j = 0; j < d.Get_Ny(); ++j)
            for (size_t i = 0; i < d.Get_Nx(); ++i)
              f << d.Get_x(int(i)) << "\t" << d.Get_y(int(j)) << "\t" << d.Get_z(int(k)) << "\t"
                << u.Get_val(i, j, k) << "\t\n";
      }
  }
  if (!quiet_) std::cout << "done" << std::endl;
}

void elliptic::IO_bam(const std::string &) {
  std::cerr << "IO_bam export is not available in this build\n";
}

void elliptic::IO_Zcode(const std::string &) {
  std::cerr << "IO_Zcode export is not available in this build\n";
}

}  // namespace olliptic
