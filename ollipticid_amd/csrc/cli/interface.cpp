// Command-line registry of Olliptic.x; see include/interface.h.
#include "interface.h"

#include "olliptic_amd.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>

// interface.cc:883-949
void Valid_N(size_t &N) {
  static const size_t v[63] = {
      3,   4,   5,   6,   7,   8,   9,   10,  11,  12,  13,  14,  15,  17,  19,  21,
      23,  25,  27,  29,  33,  37,  41,  45,  49,  53,  57,  65,  73,  81,  89,  97,
      105, 113, 129, 145, 161, 177, 193, 209, 225, 257, 289, 321, 353, 385, 417, 449,
      513, 577, 641, 705, 769, 833, 897, 1153, 1281, 1409, 1665, 1793, 2305, 2817, 3329};
  for (size_t x : v)
    if (x == N) return;
  size_t lo = v[61], hi = v[62];
  for (size_t i = 0; i < 62 && v[i] < N; ++i) {
    lo = v[i];
    hi = v[i + 1];
  }
  N = (N - lo < hi - N) ? lo : hi;
}

void interface::add_double(const std::string &n, double val, const std::string &k1,
                           const std::string &k2, const std::string &help) {
  d_names.push_back(n);
  d_par[n] = val;
  keys[n] = {k1, k2, help};
}
void interface::add_int(const std::string &n, size_t val, const std::string &k1,
                        const std::string &k2, const std::string &help) {
  i_names.push_back(n);
  i_par[n] = val;
  keys[n] = {k1, k2, help};
}
void interface::add_logic(const std::string &n, bool val, const std::string &k1,
                          const std::string &k2, const std::string &help) {
  l_names.push_back(n);
  l_par[n] = val;
  keys[n] = {k1, k2, help};
}
void interface::add_string(const std::string &n, const std::string &val, const std::string &k1,
                           const std::string &k2, const std::string &help) {
  s_names.push_back(n);
  s_par[n] = val;
  keys[n] = {k1, k2, help};
}

void interface::fail(const std::string &msg) {
  exec = false;
  if (!error.empty()) error += "; ";
  error += msg;
}

// registry and defaults: interface.cc:35-355
interface::interface(int argc, char *argv[]) {
  std::ostringstream cl;
  for (int i = 0; i < argc; ++i) cl << argv[i] << " ";
  param = cl.str();

  add_double("dxyz", 0.5, "-d", "--grid-size", "Grid size [dx=dy=dz]");
  add_double("dx", 0.5, "-dx", "--grid-size-x", "Grid size in x");
  add_double("dy", 0.5, "-dy", "--grid-size-y", "Grid size in y");
  add_double("dz", 0.5, "-dz", "--grid-size-z", "Grid size in z");
  add_double("L", 2.0, "-L", "--length", "Domain length [lx=ly=lz]");
  add_double("Lx", 2.0, "-Lx", "--length-x", "Domain length in x");
  add_double("Ly", 2.0, "-Ly", "--length-y", "Domain length in y");
  add_double("Lz", 2.0, "-Lz", "--length-z", "Domain length in z");
  add_double("Dprint", -1.0, "-Dp", "--d-print", "Grid size for interpolated data");
  add_double("tolerance", 1e-8, "-s", "--tolerance", "Residual tolerance");
  add_double("ps_r0", 120, "-r0", "--radius_sf_r0", "Radius of the scalar field perturbation");
  add_double("ps_sigma", 8, "-sigma", "--sigma_sf", "Sigma of the scalar field perturbation");
  add_double("ps_phi0", 0.025, "-phi0", "--phi0_sf", "Amplitude of the scalar field perturbation");
  add_double("ps_a2", 1, "-a2", "--a2_sf", "Scalar field perturbation parameter a2");
  add_double("robin_a", 0.0, "-BA", "--boundary-value",
             "Robin value at infinity (default 1 for Brill-waves, Punctures_EM, else 0)");
  add_logic("io_bam", false, "-iob", "--io-bam", "Produce ID for bam");
  add_logic("io_Zcode", false, "-ioZ", "--io-Zcode", "Produce ID for Zcode");
  add_logic("print_coeff", false, "-pc", "--print-coeff", "Print coefficients");
  add_logic("print_1d", true, "-p1d", "--print-1d", "Print 1d output");
  add_logic("print_2d", false, "-p2d", "--print-2d", "Print 2d output");
  add_logic("print_3d", false, "-p3d", "--print-3d", "Print 3d output");
  add_logic("print_1d_it", false, "-p1d_it", "--print-1d-iteration", "Print 1d output per iteration");
  add_logic("print_interpol", true, "-pint", "--print-interpol", "Print 1d interpolated output");
  add_logic("interpol_output", true, "-int", "--interpol-output", "Interpolate 1d and 2d output");
  add_logic("fit_approx", false, "-fapp", "--fit-appox", "Fit approximate initial data");
  add_int("nxyz", 30, "-n", "-N", "Number of grid points [nx=ny=nz]");
  add_int("nx", 30, "-nx", "--grid-points-x", "Grid points in x");
  add_int("ny", 30, "-ny", "--grid-points-y", "Grid points in y");
  add_int("nz", 30, "-nz", "--grid-points-z", "Grid points in z");
  add_int("levels", 1, "-l", "--levels", "Number of refinement levels");
  add_int("order", 4, "-O", "--order", "Order of stencil [2,4,6,8]");
  add_int("nu1", 3, "-nu1", "--pre-smooth", "Pre-smooth iterations");
  add_int("nu2", 3, "-nu2", "--post-smooth", "Post-smooth iterations");
  add_int("NumBH", 1, "-NumBH", "--num-bh", "Number of punctures");
  add_int("ps_case", 1, "-psc", "-ps_case", "Puncture + scalar field case [1,2,3,4]");
  add_int("robin_n", 1, "-Bn", "--boundary-falloff", "Robin fall-off power");
  add_string("dir_name", "Output", "-o", "--out-dir", "Output directory");
  add_string("file_bam", "none", "-iof", "--io-bam-file", "Bam initial data file");
  add_string("file_TOV", "none", "-tovf", "--tov-file", "TOV ID file");
  add_string("help", "h", "-h", "--help", "Displays this usage screen");
  add_string("equation", "none", "-eq", "--equation",
             "Punctures, Punctures_scalar_field, Poisson, Brill-waves or Punctures_EM");
  add_string("method", "Vcycle", "-m", "--method", "Vcycle");
  add_string("output", "yes", "-so", "--std-output", "Standard output: yes, partial or no");
  add_string("domain_sym", "full", "-ds", "--domain-sym", "full");
  add_string("boundary", "Dirichlet", "-B", "--boundary", "Robin");
  add_string("precision", "fp64", "-prec", "--precision", "fp64 or mixed");
  // no reference counterpart: start the V-cycles from the -m AppPuncture field
  add_string("initial_guess", "zero", "-ig", "--initial-guess", "zero or approx");

  bhmp.assign(1, 0);
  read_console(argc, argv);
  proc_parameters();
  // interface.cc:341-351: a command line the solver cannot run prints the
  // usage and ends the program with status 0
  if (!exec) {
    std::cout << "Olliptic.x (MI355X): " << error << "\n"
              << "usage: Olliptic.x -eq Punctures|Punctures_scalar_field -B Robin [-BA a -Bn n]\n"
              << "       -NumBH k -bhmp<i> m -bhx<i> .. -bhsz<i> .. -L L -N N -O 2|4|6|8 -l levels\n"
              << "       -s tol -nu1 n -nu2 n -o dir [-psc c -phi0 p -r0 r -sigma s -a2 a]\n"
              << "       [-p1d yes|no] [-pint yes|no] [-int yes|no] [-Dp d] [-p3d yes|no]\n"
              << "       [-iob yes -ioZ yes -iof file] [-so yes|partial|no] [-prec fp64|mixed]\n";
    std::exit(0);
  }
}

// interface.cc:366-874
void interface::read_console(int argc, char *argv[]) {
  if (argc > 2) {
    for (int i = 1; i < argc; ++i) {
      const char *a = argv[i];
      const bool has = i + 1 < argc;
      auto match = [&](const std::string &n) {
        const Keys &k = keys[n];
        return std::strcmp(a, k.k1.c_str()) == 0 || std::strcmp(a, k.k2.c_str()) == 0;
      };
      for (const std::string &n : d_names)
        if (match(n)) {
          if (!has) {
            fail("Missing float value for: " + keys[n].k1);
            continue;
          }
          const double v = std::strtod(argv[i + 1], nullptr);
          d_par[n] = v;
          if (n == "robin_a") robin_value_given = true;
          if (n == "L" || n == "Lx" || n == "Ly" || n == "Lz") {
            if (n == "L") d_par["Lx"] = d_par["Ly"] = d_par["Lz"] = v;
            by_domain = true;
          }
          if (n == "dxyz" || n == "dx" || n == "dy" || n == "dz") {
            if (n == "dxyz") d_par["dx"] = d_par["dy"] = d_par["dz"] = v;
            by_grid_size = true;
          }
        }
      for (const std::string &n : i_names)
        if (match(n)) {
          if (!has) {
            fail("Missing integer value for: " + keys[n].k1);
            continue;
          }
          const long v = long(std::strtoul(argv[i + 1], nullptr, 10));
          if (v < 0) {
            fail("Options for integer parameter " + keys[n].k1 + " must be a positive integer");
            continue;
          }
          i_par[n] = size_t(v);
          if (n == "nxyz" || n == "nx" || n == "ny" || n == "nz") {
            if (n == "nxyz") i_par["nx"] = i_par["ny"] = i_par["nz"] = size_t(v);
            by_points = true;
          }
        }
      for (const std::string &n : l_names)
        if (match(n)) {
          if (!has) {
            fail("Missing logic value for: " + keys[n].k1);
            continue;
          }
          const std::string v = argv[i + 1];
          if (v == "yes" || v == "YES" || v == "true")
            l_par[n] = true;
          else if (v == "no" || v == "NO" || v == "false")
            l_par[n] = false;
          else
            fail("Options for logic parameter " + keys[n].k1 + " must be yes/no");
        }
      for (const std::string &n : s_names)
        if (match(n) && n != "help") {
          if (!has) {
            fail("Missing text value for: " + keys[n].k1);
            continue;
          }
          s_par[n] = argv[i + 1];
        }
      if (std::strcmp(a, "--help") == 0 || std::strcmp(a, "-h") == 0) fail("help requested");
    }
  }
  const size_t nbh = i_par["NumBH"];
  for (auto *v : {&bhmp, &bhqp, &bhx, &bhy, &bhz, &bhpx, &bhpy, &bhpz, &bhsx, &bhsy, &bhsz})
    v->assign(nbh, 0.0);
  struct Bh {
    const char *key;
    std::vector<double> *vec;
  } bh[] = {{"-bhmp", &bhmp}, {"-bhqp", &bhqp}, {"-bhx", &bhx},   {"-bhy", &bhy},
            {"-bhz", &bhz},   {"-bhpx", &bhpx}, {"-bhpy", &bhpy}, {"-bhpz", &bhpz},
            {"-bhsx", &bhsx}, {"-bhsy", &bhsy}, {"-bhsz", &bhsz}};
  for (int i = 1; i + 1 < argc; ++i)
    for (size_t j = 0; j < nbh; ++j)
      for (const Bh &b : bh)
        if (std::string(argv[i]) == std::string(b.key) + std::to_string(j + 1))
          (*b.vec)[j] = std::strtod(argv[i + 1], nullptr);
}

// interface.cc:958-1125 (the parts that define the puncture problem)
void interface::proc_parameters() {
  if (i_par["levels"] < 1) fail("levels must be >= 1");
  const size_t order = i_par["order"];
  if (!(order == 2 || order == 4 || order == 6 || order == 8)) fail("order must be [2,4,6,8].");
  if (by_domain) {
    if (by_points && by_grid_size) {
      fail("Domain definition");
    } else if (by_grid_size) {
      i_par["nx"] = size_t(std::fabs(d_par["Lx"]) / d_par["dx"] + 1);
      i_par["ny"] = size_t(std::fabs(d_par["Ly"]) / d_par["dy"] + 1);
      i_par["nz"] = size_t(std::fabs(d_par["Lz"]) / d_par["dz"] + 1);
    } else {
      size_t nx = i_par["nx"], ny = i_par["ny"], nz = i_par["nz"];
      Valid_N(nx);
      Valid_N(ny);
      Valid_N(nz);
      d_par["dx"] = std::fabs(d_par["Lx"]) / double(nx - 1.0);
      d_par["dy"] = std::fabs(d_par["Ly"]) / double(ny - 1.0);
      d_par["dz"] = std::fabs(d_par["Lz"]) / double(nz - 1.0);
      i_par["nx"] = nx;
      i_par["ny"] = ny;
      i_par["nz"] = nz;
    }
  } else {
    size_t nx = i_par["nx"], ny = i_par["ny"], nz = i_par["nz"];
    Valid_N(nx);
    Valid_N(ny);
    Valid_N(nz);
    d_par["Lx"] = d_par["dx"] * (nx - 1.0);
    d_par["Ly"] = d_par["dy"] * (ny - 1.0);
    d_par["Lz"] = d_par["dz"] * (nz - 1.0);
    i_par["nx"] = nx;
    i_par["ny"] = ny;
    i_par["nz"] = nz;
  }
  const std::string eq = s_par["equation"];
  if (eq == "Punctures")
    equation = PUNCTURES;
  else if (eq == "Punctures_scalar_field")
    equation = PUNCTURES_SCALAR;
  else if (eq == "Poisson")
    equation = POISSON;
  else if (eq == "Brill-waves")
    equation = BRILL_WAVES;
  else if (eq == "Punctures_EM")
    equation = PUNCTURES_EM;
  else
    fail("equation '" + eq +
         "' is not solved here (Punctures, Punctures_scalar_field, Poisson, Brill-waves, "
         "Punctures_EM)");
  if ((equation == BRILL_WAVES || equation == PUNCTURES_EM) && !robin_value_given)
    d_par["robin_a"] = 1.0;
  if (equation == PUNCTURES)
    for (double m : bhmp)
      if (m <= 0) fail("Wrong punture mass parameters");
  const std::string b = s_par["boundary"];
  boundary = b == "Robin" ? ROBIN : (b == "Asymptotic" ? ASYMPTOTIC : DIRICHLET);
  if (boundary != ROBIN) fail("boundary '" + b + "' is not solved here (Robin)");
  const std::string m = s_par["method"];
  if (m == "AppPuncture")
    method = APP;
  else if (m == "Vcycle")
    method = VCYCLE;
  else
    fail("method '" + m + "' is not available here (Vcycle, AppPuncture)");
  // interface.cc:1154-1160
  if (method == APP && equation != PUNCTURES)
    fail("*** Error: App puncture only works with puncture equation");
  const std::string g = s_par["initial_guess"];
  if (g != "zero" && g != "approx") fail("initial guess must be zero or approx");
  if (s_par["domain_sym"] != "full") fail("only the full domain is available");
  const std::string o = s_par["output"];
  output = o == "no" ? NO : (o == "partial" ? PARTIAL : YES);
  if (i_par["nx"] != i_par["ny"] || i_par["nx"] != i_par["nz"] || d_par["dx"] != d_par["dy"] ||
      d_par["dx"] != d_par["dz"])
    fail("only cubic domains (equal N and L per axis) are available");
  // interface.cc:1168-1175
  if (l_par["io_bam"] && s_par["file_bam"] == "none") fail("*** Error: Missing file.");
  if (i_par["NumBH"] < 1 || i_par["NumBH"] > OLL_MAX_BH) fail("NumBH must be in 1..16");
}

bool interface::Get_logic_par(const std::string &n) const {
  auto it = l_par.find(n);
  return it == l_par.end() ? false : it->second;
}
size_t interface::Get_integer_par(const std::string &n) const {
  auto it = i_par.find(n);
  return it == i_par.end() ? 0 : it->second;
}
double interface::Get_double_par(const std::string &n) const {
  auto it = d_par.find(n);
  return it == d_par.end() ? 0.0 : it->second;
}
std::string interface::Get_string_par(const std::string &n) const {
  auto it = s_par.find(n);
  return it == s_par.end() ? std::string() : it->second;
}
