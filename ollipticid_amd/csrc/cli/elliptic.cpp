// Solver object and field views of Olliptic.x; see include/elliptic.h.
#include "elliptic.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>

namespace fs = std::filesystem;

// ---------------------------------------------------------------- views
void ffunction::ensure() {
  if (!fetched_ || (gen_ && seen_ != *gen_)) Fetch();
}

double ffunction::Get_val(size_t i, size_t j, size_t k) {
  ensure();
  return host_[i + Get_nx() * (j + Get_ny() * k)];
}

bool ffunction::in_region(size_t i, size_t j, size_t k) const {
  if (it_ == ALL) return true;
  const size_t R = d_.Get_buffer();
  const bool inner = i >= R && i < Get_nx() - R && j >= R && j < Get_ny() - R && k >= R &&
                     k < Get_nz() - R;
  return it_ == INSIDE ? inner : !inner;
}

void ffunction::Set_Iterator(t_iterator it) {
  it_ = it;
  ci_ = cj_ = ck_ = 0;
  done_ = false;
  if (!in_region(0, 0, 0)) End();
}

bool ffunction::End() {
  if (done_) return false;
  do {
    if (++ci_ == Get_nx()) {
      ci_ = 0;
      if (++cj_ == Get_ny()) {
        cj_ = 0;
        if (++ck_ == Get_nz()) {
          done_ = true;
          return false;
        }
      }
    }
  } while (!in_region(ci_, cj_, ck_));
  return true;
}

double ffunction::Get_val() { return Get_val(ci_, cj_, ck_); }

void ffunction::Set(double y) { Set(ci_, cj_, ck_, y); }

void ffunction::Set(size_t i, size_t j, size_t k, double y) {
  ensure();
  host_[i + Get_nx() * (j + Get_ny() * k)] = y;
}

void ffunction::Fetch() {
  host_.resize(d_.Get_N());
  if (oll_get_field(h_, box_, field_, host_.data()) != OLL_OK)
    throw std::runtime_error(oll_last_error());
  fetched_ = true;
  if (gen_) seen_ = *gen_;
}

void ffunction::Store() {
  if (oll_set_field(h_, box_, field_, host_.data()) != OLL_OK)
    throw std::runtime_error(oll_last_error());
  // every other mirror of the hierarchy may now be stale; this one is not
  if (gen_) seen_ = ++*gen_;
}

size_t domain::index_of(int a, double x) const {
  const int np = b.n[a];
  if (!(lo(a) < x && x < hi(a))) return lo(a) >= x ? 1 : size_t(np - 1);
  int inf = 0, sup = np;
  int i = (np % 2 == 0) ? np / 2 : (np - 1) / 2;
  do {
    const double xi = lo(a) + double(i) * b.d[a];
    if (xi > x || std::fabs(xi - x) <= 3 * std::numeric_limits<double>::epsilon())
      sup = i;
    else
      inf = i;
    i = ((sup + inf) % 2 == 0) ? (sup + inf) / 2 : (sup + inf - 1) / 2;
  } while (inf + 1 < sup);
  return size_t(i + 1);
}

bool domain::is_in(double x, double y, double z) const {
  const double p[3] = {x, y, z};
  for (int a = 0; a < 3; ++a)
    if (!(lo(a) - b.d[a] * 0.1 <= p[a] && p[a] <= hi(a) + b.d[a] * 0.1)) return false;
  return true;
}

bool domain::is_in_nb(double x, double y, double z) const {
  const double p[3] = {x, y, z};
  const double w = double(b.ghost) - 0.1;
  for (int a = 0; a < 3; ++a)
    if (!(lo(a) + b.d[a] * w < p[a] && p[a] < hi(a) - b.d[a] * w)) return false;
  return true;
}

bool domain::operator<(const domain &D) const {
  for (int a = 0; a < 3; ++a)
    if (!(D.lo(a) < lo(a) && hi(a) < D.hi(a))) return false;
  return true;
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
  for (size_t d = 1; d < dom_.size(); ++d)
    if (dom_[l0_] < dom_[d]) l0_ = d;
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

// field id of a variable name (operator.cc:103-124 names)
static int field_id(const std::string &name) {
  static const char *names[OLL_NFIELDS] = {"u", "b", "c", "d", "e", "rho0", "r0", "v0", "w0", "ww0"};
  int f = -1;
  for (int i = 0; i < OLL_NFIELDS; ++i)
    if (name == names[i]) f = i;
  if (name == "rho") f = OLL_F_RHO;
  if (f < 0) throw std::runtime_error("unknown field " + name);
  return f;
}

void multigrid::Transferir(size_t from, std::string from_var, size_t to, std::string to_var,
                           t_iterator iter, m_interpol method, size_t ord, bool all,
                           bool to_zero) {
  if (method != LINEAR && method != LAGRANGE) {
    std::cerr << "Transferir: only LINEAR and LAGRANGE interpolation are available" << std::endl;
    std::exit(1);
  }
  const int region = iter == ALL ? OLL_REGION_ALL : iter == INSIDE ? OLL_REGION_INSIDE
                                                                   : OLL_REGION_BOUNDARY;
  if (oll_transfer(h_, int(from), field_id(from_var), int(to), field_id(to_var), region,
                   method == LINEAR ? OLL_INTERP_LINEAR : OLL_INTERP_LAGRANGE, int(ord),
                   all ? 1 : 0, to_zero ? 1 : 0) != OLL_OK) {
    std::cerr << "Transferir: " << oll_last_error() << std::endl;
    std::exit(1);
  }
  Touch();
}

ffunction *multigrid::Get_ffunction(size_t level, size_t box, const std::string &name) const {
  const int f = field_id(name);
  const size_t ib = Get_dom_index(level, box);
  std::unique_ptr<ffunction> &p = fields_[{int(ib), f}];
  if (!p) p.reset(new ffunction(h_, int(ib), f, dom_[ib], gen_));
  return p.get();
}

size_t multigrid::Box_Of(double x, double y, double z) const {
  size_t in = Get_dom_index(l0_, 0);
  for (size_t l = 0; l < Get_levels(); ++l)
    for (size_t i = 0; i < Get_num_dom(l); ++i) {
      const size_t ib = Get_dom_index(l, i);
      if (dom_[ib].is_in_nb(x, y, z)) in = ib;
    }
  return in;
}

std::vector<double> multigrid::Interpol(const std::string &name, const std::vector<int> &box,
                                        const std::vector<double> &xyz, size_t order) const {
  const ffunction *f = Get_ffunction(0, 0, name);
  std::vector<double> out(box.size());
  if (box.empty()) return out;
  if (oll_interp(h_, f->Field(), int(box.size()), box.data(), xyz.data(), int(order),
                 out.data()) != OLL_OK)
    throw std::runtime_error(oll_last_error());
  return out;
}

namespace {

// point list with the box each point is read from
struct Points {
  std::vector<int> box;
  std::vector<double> xyz;
  void add(size_t b, double x, double y, double z) {
    box.push_back(int(b));
    xyz.insert(xyz.end(), {x, y, z});
  }
};

void open_sci(std::ofstream &f, const std::string &path, int digits) {
  f.open(path, std::ios::app);
  f.setf(std::ios_base::scientific, std::ios_base::floatfield);
  f.precision(digits);
}

}  // namespace

void multigrid::Print_1D(const std::string &dir, const std::string &name, size_t l, int c,
                         double x1, double x2, bool interpolar, size_t order) const {
  static const char *ext[3] = {".x", ".y", ".z"};
  l = l < Get_levels() ? l : Get_levels() - 1;
  for (size_t ib = 0; ib < Get_num_dom(l); ++ib) {
    const size_t b = Get_dom_index(l, ib);
    const domain &D = dom_[b];
    const size_t n[3] = {D.Get_Nx(), D.Get_Ny(), D.Get_Nz()};
    // ffunction.cc:838-844: the middle point, then the two fixed axes
    size_t idx[3];
    for (int a = 0; a < 3; ++a) idx[a] = (n[a] % 2 == 0) ? n[a] / 2 : (n[a] - 1) / 2;
    const int a1 = c == CUT_X ? 1 : 0, a2 = c == CUT_Z ? 1 : 2;
    double p[3] = {D.Get_x(int(idx[0])), D.Get_y(int(idx[1])), D.Get_z(int(idx[2]))};
    const double want[2] = {x1, x2};
    const int fixed[2] = {a1, a2};
    for (int t = 0; t < 2; ++t) {
      const int a = fixed[t];
      const bool in = a == 0 ? D.x_is_in(want[t]) : a == 1 ? D.y_is_in(want[t]) : D.z_is_in(want[t]);
      if (in && interpolar) {
        p[a] = want[t];
      } else if (in) {
        idx[a] = a == 0 ? D.Get_i(want[t]) : a == 1 ? D.Get_j(want[t]) : D.Get_k(want[t]);
        p[a] = a == 0 ? D.Get_x(int(idx[a])) : a == 1 ? D.Get_y(int(idx[a])) : D.Get_z(int(idx[a]));
      }
    }
    // values: Interpol_In at (node, p[a1], p[a2]); without interpolation
    // every point is a grid point and the same call returns its value
    Points pts;
    for (size_t i = 0; i < n[c]; ++i) {
      double q[3] = {p[0], p[1], p[2]};
      q[c] = c == CUT_X ? D.Get_x(int(i)) : c == CUT_Y ? D.Get_y(int(i)) : D.Get_z(int(i));
      pts.add(b, q[0], q[1], q[2]);
    }
    const std::vector<double> v = Interpol(name, pts.box, pts.xyz, order);
    std::ostringstream fn;
    fn << dir << name << "-l" << l << "b" << ib << "-" << 0 << ext[c];
    std::ofstream f;
    open_sci(f, fn.str(), 12);
    for (size_t i = 0; i < v.size(); ++i) f << pts.xyz[3 * i + size_t(c)] << "\t" << v[i] << "\t\n";
    f << "\n";
  }
}

void multigrid::Print_Interpol_1D(const std::string &dir, const std::string &name, double LD,
                                  double _dx, int c, double x1, double x2, size_t order) const {
  static const char *ext[3] = {".X", ".Y", ".Z"};
  const domain &D0 = dom_[l0_];
  double p[3] = {0, 0, 0};
  const double xi = -0.5 * LD, xf = 0.5 * LD;
  const double Dx = _dx > 0.0 ? _dx : LD / 1000;
  // multigrid.cc:204-211, 317-324, 430-437: the fixed coordinates
  if (c == CUT_X) {
    if (D0.y_is_in(x1)) p[1] = x1;
    if (D0.z_is_in(x2)) p[2] = x2;
  } else if (c == CUT_Y) {
    if (D0.x_is_in(x1)) p[0] = x1;
    if (D0.z_is_in(x2)) p[2] = x2;
  } else {
    if (D0.x_is_in(x1)) p[0] = x1;
    if (D0.y_is_in(x2)) p[1] = x2;
  }
  Points pts;
  std::vector<double> t;
  for (double s = xi; s <= xf; s += Dx) {
    p[c] = s;
    // only the x cut is limited to the outer box (multigrid.cc:225)
    if (c == CUT_X && !D0.is_in(p[0], p[1], p[2])) continue;
    pts.add(Box_Of(p[0], p[1], p[2]), p[0], p[1], p[2]);
    t.push_back(s);
  }
  const std::vector<double> v = Interpol(name, pts.box, pts.xyz, order);
  std::ofstream f;
  open_sci(f, dir + name + ext[c], 18);
  for (size_t i = 0; i < v.size(); ++i) f << t[i] << "\t" << v[i] << "\n";
  f << "\n";
}

void multigrid::Print_Interpol_2D(const std::string &dir, const std::string &name, double LD1,
                                  double LD2, double _dx1, double _dx2, int c, double x1,
                                  size_t order) const {
  static const char *ext[3] = {".YZ", ".XZ", ".XY"};
  std::cout << "Print 2D interpol dir: " << dir << " name: " << name
            << " dir+name: " << dir + name + ".XY" << std::endl;
  const domain &D0 = dom_[l0_];
  double p[3] = {0, 0, 0};
  const double s1i = -0.25 * LD1, s2i = -0.25 * LD2, s1f = 0.25 * LD1, s2f = 0.25 * LD2;
  const double D1 = _dx1 > 0.0 ? _dx1 : LD1 / 200;
  const double D2 = _dx2 > 0.0 ? _dx2 : LD2 / 200;
  // the plane's two axes and the fixed one
  const int a1 = c == CUT_X ? 1 : 0, a2 = c == CUT_Z ? 1 : 2;
  if (c == CUT_X && D0.x_is_in(x1)) p[0] = x1;
  if (c == CUT_Y && D0.y_is_in(x1)) p[1] = x1;
  if (c == CUT_Z && D0.z_is_in(x1)) p[2] = x1;
  Points pts;
  std::vector<size_t> row_end;
  for (double s = s1i; s <= s1f; s += D1) {
    for (double r = s2i; r <= s2f; r += D2) {
      p[a1] = s;
      p[a2] = r;
      pts.add(Box_Of(p[0], p[1], p[2]), p[0], p[1], p[2]);
    }
    row_end.push_back(pts.box.size());
  }
  const std::vector<double> v = Interpol(name, pts.box, pts.xyz, order);
  const std::string out = dir + name + ext[c];
  std::ofstream f;
  open_sci(f, out, 18);
  // multigrid.cc:880-900: the XY plane ends each row with a blank line
  size_t row = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    f << pts.xyz[3 * i + size_t(a1)] << "\t" << pts.xyz[3 * i + size_t(a2)] << "\t" << v[i] << "\n";
    if (c == CUT_Z && i + 1 == row_end[row]) {
      f << "\n";
      ++row;
    }
  }
  f << "\n";
  std::cout << "Out file: " << out << std::endl;
}

// ---------------------------------------------------------------- elliptic
elliptic::elliptic(interface *Oll) { Make(Oll); }

elliptic::~elliptic() {
  if (h_) oll_hier_destroy(h_);
}

// elliptic.cc:31-221 Make; interface.cc:1497-1527 output directory
void elliptic::Make(interface *Oll) {
  oll_problem_default(&prob_);
  switch (Oll->Get_Equation()) {
    case PUNCTURES_SCALAR: prob_.eq = OLL_EQ_PUNCTURES_SCALAR; break;
    case POISSON: prob_.eq = OLL_EQ_POISSON; break;
    case BRILL_WAVES: prob_.eq = OLL_EQ_BRILL; break;
    case PUNCTURES_EM: prob_.eq = OLL_EQ_PUNCTURES_EM; break;
    default: prob_.eq = OLL_EQ_PUNCTURES; break;
  }
  const std::vector<double> m = Oll->Get_bhmp(), x = Oll->Get_bhx(), y = Oll->Get_bhy(),
                            z = Oll->Get_bhz(), px = Oll->Get_bhpx(), py = Oll->Get_bhpy(),
                            pz = Oll->Get_bhpz(), sx = Oll->Get_bhsx(), sy = Oll->Get_bhsy(),
                            sz = Oll->Get_bhsz(), qp = Oll->Get_bhqp();
  prob_.num_bh = int(m.size());
  for (size_t i = 0; i < m.size() && i < size_t(OLL_MAX_BH); ++i)
    prob_.bh[i] = oll_puncture{m[i], x[i], y[i], z[i], px[i], py[i], pz[i], sx[i], sy[i], sz[i], qp[i]};
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
  prob_.method = Oll->Get_Method() == APP ? OLL_METHOD_APP : OLL_METHOD_VCYCLE;
  prob_.initial_guess =
      Oll->Get_string_par("initial_guess") == "approx" ? OLL_GUESS_APPROX : OLL_GUESS_ZERO;
  fit_approx_ = Oll->Get_logic_par("fit_approx");
  tol_ = prob_.tol < 2 * 2.220446049250313e-16 ? 2 * 2.220446049250313e-16 : prob_.tol;
  print_1d_ = Oll->Get_logic_par("print_1d");
  print_3d_ = Oll->Get_logic_par("print_3d");
  print_interpol_ = Oll->Get_logic_par("print_interpol");
  interpol_output_ = Oll->Get_logic_par("interpol_output");
  quiet_ = Oll->Get_TOutput() == NO;
  verbose_ = Oll->Get_TOutput() == YES;

  const std::string name = Oll->Get_string_par("dir_name");
  dir_ = name + "/";
  std::error_code ec;
  fs::remove_all(name + "_prev", ec);
  if (fs::exists(name)) fs::rename(name, name + "_prev", ec);
  fs::create_directories(name, ec);
  std::ofstream(dir_ + name + ".run") << Oll->Command_Line() << "\n";

  if (oll_hier_create(&prob_, &h_) != OLL_OK) return failed(oll_last_error());
  mg_ = multigrid(h_);
}

// the reference ends the program from inside on errors (output.cc:20-111
// error_exit, elliptic.cc:1194-1207): so do the class's void methods
void elliptic::failed(const std::string &msg, bool print) {
  ok_ = false;
  err_ = msg;
  if (print) std::cerr << "Error: " << err_ << std::endl;
  std::exit(1);
}

// boundary.cc:8-84
void elliptic::Set_Boundary() {
  if (!ok_) return;
  if (oll_reset(h_) != OLL_OK) return failed(oll_last_error());
  mg_.Touch();
}

// boundary.cc:447-907
void elliptic::App_Bound(size_t l, size_t ibox) {
  if (!ok_) return;
  if (oll_robin_box(h_, int(mg_.Get_dom_index(l, ibox))) != OLL_OK) return failed(oll_last_error());
  mg_.Touch();
}

// elliptic.h:337-345 at the iterator point of box (l, ibox)'s u
double elliptic::op_at(int which, size_t lvar, size_t l, size_t ibox) {
  if (!ok_) return 0.0;
  if (lvar != 0) {
    failed("the puncture equations have one variable (lvar 0)");
    return 0.0;
  }
  ffunction *u = mg_.Get_ffunction(l, ibox, "u");
  const int ib = u->Box();
  if (op_level_[which] != int(l) || op_gen_[which] != mg_.Generation()) {
    double linf = 0;
    if (oll_operator(h_, int(l), which, &linf) != OLL_OK) {
      failed(oll_last_error());
      return 0.0;
    }
    op_level_[which] = int(l);
    op_val_[which].clear();
    op_gen_[which] = mg_.Generation();
  }
  std::vector<double> &v = op_val_[which][ib];
  if (v.empty()) {
    v.resize(u->Get_N());
    if (oll_get_field(h_, ib, OLL_F_R, v.data()) != OLL_OK) {
      failed(oll_last_error());
      return 0.0;
    }
  }
  return v[u->Get_Index()];
}

double elliptic::OPERATOR(size_t lvar, size_t l, size_t ibox) { return op_at(0, lvar, l, ibox); }
double elliptic::duOPERATOR(size_t lvar, size_t l, size_t ibox) { return op_at(1, lvar, l, ibox); }

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
  if (oll_adm_mass(h_, adm_.data(), surf.data()) != OLL_OK) return failed(oll_last_error());
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

// elliptic.cc:2232-2444 Get_Puncture_App: the approximation at each
// puncture's grid node, printed
void elliptic::Get_Puncture_App() {
  if (!ok_) return;
  std::vector<double> up(size_t(prob_.num_bh), 0.0);
  if (oll_puncture_approx(h_, up.data()) != OLL_OK) return;
  std::cout << std::scientific << std::setprecision(7);
  for (int i = 0; i < prob_.num_bh; ++i)
    std::cout << " == Approximate ID: puncture " << i << "\n"
              << " x : " << prob_.bh[i].x << "\t" << " y : " << prob_.bh[i].y << "\t"
              << " z : " << prob_.bh[i].z << "\t" << " u : " << up[size_t(i)] << std::endl;
}

// elliptic.cc:1977-2222 ApproximateID (-m AppPuncture): u is the
// approximation, no iteration
void elliptic::ApproximateID() {
  if (!ok_) return;
  oll_result r{};
  if (oll_solve(h_, &r, nullptr, 0) != OLL_OK) return failed(oll_last_error());
  mg_.Touch();
  for (size_t l = 0; l < mg_.Get_levels(); ++l)
    for (size_t b = 0; b < mg_.Get_num_dom(l); ++b)
      std::cout << " == Approximate ID: level -" << l << "- box -" << b << "-\n" << std::endl;
  Get_Puncture_App();
  cycles_ = 0;
  res_ = 0;
}

// elliptic.cc:1251-1342 Vcycle with the reference's per-cycle logging
void elliptic::Solve() {
  if (!ok_) return;
  if (prob_.method == OLL_METHOD_APP) return ApproximateID();
  Vcycle();
}

void elliptic::Vcycle() {
  if (!ok_) return;
  if (fit_approx_) Get_Puncture_App();
  const int top = int(mg_.Get_levels()) - 1;
  std::vector<double> pre(size_t(top) + 1), post(size_t(top) + 1);
  double res = 1e12, res_prev;
  int cycle = 0;
  do {
    res_prev = res;
    const int rc = oll_fas(h_, top, &res);
    mg_.Touch();
    if (rc != OLL_OK) return failed(oll_last_error());
    oll_last_norms(h_, pre.data(), post.data(), top + 1);
    for (int k = top; k >= 0; --k) print_pre(k, top, pre[size_t(k)], cycle);
    for (int k = 1; k <= top; ++k) print_post(k, top, post[size_t(k)], cycle);
    // elliptic.cc:1939-1941: the ADM mass only for the puncture equations
    if (prob_.eq == OLL_EQ_PUNCTURES || prob_.eq == OLL_EQ_PUNCTURES_SCALAR) compute_adm(true);
    ++cycle;
  } while (res > tol_ && res < res_prev);
  cycles_ = cycle;
  res_ = res;
  if (!quiet_) {
    std::cout << std::scientific << std::setprecision(8);
    // elliptic.cc:1333-1335: up to the first zero mass
    for (int j = 0; j < prob_.num_bh && prob_.bh[j].m != 0.0; ++j)
      std::cout << "mp [" << j << "] = " << prob_.bh[j].m << "\n";
    // elliptic.cc:1337-1339: the charges up to the first zero one
    for (int j = 0; j < prob_.num_bh && prob_.bh[j].q != 0.0; ++j)
      std::cout << "qp [" << j << "] = " << prob_.bh[j].q << "\n";
  }
}

// output.cc:624-729: 1-D cuts of every box (-p1d), every box in 3-D
// (-p3d, ffunction.cc:411-467 format), composite cuts and the z = 0 plane
// (-pint), all interpolated with Lagrange order + 1
void elliptic::Print_Solution(double Dxyz, double Lx, double Ly, double Lz) {
  if (!ok_) return;
  if (!quiet_) std::cout << "Printing..." << std::endl;
  const size_t ord = size_t(prob_.order) + 1;
  if (print_1d_)
    for (size_t l = 0; l < mg_.Get_levels(); ++l)
      for (int c : {CUT_X, CUT_Y, CUT_Z}) mg_.Print_1D(dir_, "u", l, c, 0, 0, interpol_output_, ord);
  if (print_3d_) {
    for (size_t l = 0; l < mg_.Get_levels(); ++l)
      for (size_t b = 0; b < mg_.Get_num_dom(l); ++b) {
        ffunction &u = *mg_.Get_ffunction(l, b, "u");
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
  if (print_interpol_) {
    mg_.Print_Interpol_1D(dir_, "u", Lx, Dxyz, CUT_X, 0, 0, ord);
    mg_.Print_Interpol_1D(dir_, "u", Ly, Dxyz, CUT_Y, 0, 0, ord);
    mg_.Print_Interpol_1D(dir_, "u", Lz, Dxyz, CUT_Z, 0, 0, ord);
    mg_.Print_Interpol_2D(dir_, "u", Lx, Ly, -0.1, -0.1, CUT_Z, 0, ord);
  }
  if (!quiet_) std::cout << "done" << std::endl;
}

namespace {

bool read_file(const std::string &path, std::string &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

bool file_exists(const std::string &path) { return std::ifstream(path).is_open(); }

// Numbers read one at a time as `std::istream >> double` reads them:
// leading whitespace skipped, `eof` once a read reaches the end of the text,
// `fail` (sticky) once a read finds no number; a failed read leaves the
// target unchanged.  Only plain decimal notation is accepted.
struct Reader {
  const std::string &s;
  size_t pos = 0;
  bool eof = false, fail = false, malformed = false;
  bool number(double &v) {
    if (fail) return false;
    while (pos < s.size() && std::isspace(static_cast<unsigned char>(s[pos]))) ++pos;
    if (pos >= s.size()) {
      eof = fail = true;
      return false;
    }
    const char c = s[pos], c1 = pos + 1 < s.size() ? s[pos + 1] : '\0';
    const bool digit = std::isdigit(static_cast<unsigned char>(c)) || c == '.' ||
                       ((c == '+' || c == '-') &&
                        (std::isdigit(static_cast<unsigned char>(c1)) || c1 == '.'));
    char *e = nullptr;
    const double t = digit ? std::strtod(s.c_str() + pos, &e) : 0.0;
    if (!digit || e == s.c_str() + pos) {
      fail = malformed = true;
      return false;
    }
    v = t;
    pos = size_t(e - s.c_str());
    if (pos >= s.size()) eof = true;
    return true;
  }
};

}  // namespace

bool multigrid::Print_IO_BAM(const std::string &name, const std::string &var, size_t order) const {
  // the reference leaves cout in this format (multigrid.cc:1005-1008)
  std::cout.setf(std::ios_base::scientific, std::ios_base::floatfield);
  std::cout.precision(16);
  std::cout << "reading from: " << name << std::endl;
  std::string buf;
  if (!read_file(name, buf)) return true;
  const domain &D0 = dom_[Get_dom_index(l0_, 0)];
  // A record is read from where the previous one's trailing character
  // ended and rewritten (74 bytes) from where its read started.  Values are
  // interpolated in batches; a batch is written out before a read starts
  // inside the bytes one of its records will overwrite (short lines), so
  // every read sees what the reference's sequential loop would see.
  std::vector<size_t> at;
  Points pts;
  size_t write_end = 0;
  auto land = [&]() {
    const std::vector<double> v = Interpol(var, pts.box, pts.xyz, order);
    for (size_t q = 0; q < v.size(); ++q) {
      std::ostringstream o;
      o.setf(std::ios_base::scientific, std::ios_base::floatfield);
      o.precision(9);
      o << std::setw(16) << pts.xyz[3 * q] << std::setw(17) << pts.xyz[3 * q + 1] << std::setw(17)
        << pts.xyz[3 * q + 2];
      o.precision(15);
      o << std::setw(24) << v[q];
      const std::string rec = o.str();
      if (at[q] + rec.size() > buf.size()) buf.resize(at[q] + rec.size());
      buf.replace(at[q], rec.size(), rec);
    }
    at.clear();
    pts = Points();
  };
  bool ok = true;
  size_t cur = 0;
  for (;;) {
    if (write_end > cur) land();
    Reader r{buf, cur};
    double x = 0, y = 0, z = 0, val = 0;
    if (!(r.number(x) && r.number(y) && r.number(z) && r.number(val))) {
      // the reference's loop never leaves a failed read (it seeks back to
      // the start and spins); the records so far are what it wrote
      if (r.malformed) std::cerr << "warning: unreadable record in " << name << " at byte " << cur << "\n";
      break;
    }
    if (!D0.is_in(x, y, z)) {
      std::cout << "Error: bam grid bigger than Olliptic grid" << std::endl;
      ok = false;
      break;
    }
    // a record whose value ends the file is read with the stream at its
    // end, and the reference's write of it does not happen
    if (r.eof) break;
    at.push_back(cur);
    pts.add(Box_Of(x, y, z), x, y, z);
    write_end = cur + 74;
    cur = r.pos + 1;
  }
  land();
  std::ofstream out(name, std::ios::binary | std::ios::trunc);
  out << buf;
  return ok;
}

bool multigrid::Print_IO_Zcode(const std::string &name, const std::string &var, size_t order) const {
  std::cout << "reading from: " << name << std::endl;
  std::string buf;
  if (!read_file(name, buf)) return true;
  const domain &D0 = dom_[Get_dom_index(l0_, 0)];
  Points pts;
  bool ok = true, any = false;
  Reader r{buf};
  double x = 0, y = 0, z = 0, val = 0;
  // multigrid.cc:1153-1209: a record per pass until the stream is at its
  // end; a file ending in a newline therefore repeats its last record (the
  // pass that meets the end keeps the previous values)
  do {
    r.number(x);
    r.number(y);
    r.number(z);
    r.number(val);
    if (r.malformed) {
      std::cerr << "warning: unreadable record in " << name << " at byte " << r.pos << "\n";
      break;
    }
    if (r.fail && !any) break;  // nothing read: the reference would use unset values
    any = true;
    if (!D0.is_in(x, y, z)) {
      std::cout << "X: " << x << "\t" << "Y: " << y << "\t" << "Z: " << z << std::endl;
      std::cout << "Error: bam grid bigger than Olliptic grid" << std::endl;
      ok = false;
      break;
    }
    pts.add(Box_Of(x, y, z), x, y, z);
  } while (!r.eof);
  const std::vector<double> v = Interpol(var, pts.box, pts.xyz, order);
  std::ofstream out(name + "_m", std::ios::trunc);
  out.setf(std::ios_base::scientific, std::ios_base::floatfield);
  out.precision(16);
  for (size_t q = 0; q < v.size(); ++q)
    out << pts.xyz[3 * q] << " " << pts.xyz[3 * q + 1] << " " << pts.xyz[3 * q + 2] << " " << v[q]
        << "\n";
  return ok;
}

void elliptic::IO_bam(std::string file) {
  // the reference's message is printed where the point is found
  if (ok_ && !export_bam(file)) failed("a BAM point lies outside the grid", false);
}

void elliptic::IO_Zcode(std::string file) {
  if (ok_ && !export_zcode(file)) failed("a Zcode point lies outside the grid", false);
}

bool elliptic::export_bam(const std::string &file) {
  static const char *ending[] = {"", "a", "b", "c", "d", "e"};
  std::vector<std::string> names;
  size_t found = 0;
  for (size_t level = 0; level == 0 || found > 0; ++level) {
    found = 0;
    for (const char *e : ending) {
      const std::string f = file + std::to_string(level) + e;
      if (!file_exists(f)) continue;
      ++found;
      if (verbose_) std::cout << "File found: " << f << std::endl;
      names.push_back(f);
    }
  }
  for (const std::string &f : names)
    if (!mg_.Print_IO_BAM(f, "u", size_t(prob_.order) + 1)) return false;
  return true;
}

bool elliptic::export_zcode(const std::string &file) {
  std::vector<std::string> names;
  size_t found = 0;
  for (size_t level = 0; level == 0 || found > 0; ++level) {
    found = 0;
    for (int b = 0; b < 6; ++b) {
      const std::string f = file + (level < 10 ? "0" : "") + std::to_string(level) + "-0" +
                            std::to_string(b) + ".mgid";
      if (!file_exists(f)) continue;
      ++found;
      if (verbose_) std::cout << "File found: " << f << std::endl;
      names.push_back(f);
    }
  }
  for (const char *sh : {"xm", "xp", "ym", "yp", "zm", "zp"}) {
    const std::string f = file + "SH-" + sh + ".mgid";
    if (!file_exists(f)) continue;
    if (verbose_) std::cout << "File found: " << f << std::endl;
    names.push_back(f);
  }
  for (const std::string &f : names)
    if (!mg_.Print_IO_Zcode(f, "u", size_t(prob_.order) + 1)) return false;
  return true;
}

