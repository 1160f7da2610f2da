// Command-line registry of the drop-in Olliptic.x (reference class
// `interface`, include/interface.h:121-438, src/interface.cc:35-1543).
//
// Same flags, two keys per flag, same defaults and parsing rules: every argv
// entry is matched against every registry, a repeated flag keeps its last
// value, unknown flags are ignored, nothing is parsed when argc <= 2, N is
// snapped by Valid_N and dx = L/(N-1).  -BA/-Bn (the Robin value at infinity
// and fall-off power) are flags of this solver; without -BA the Robin value
// is the one the reference hard-codes per equation (boundary_coeff,
// operator.cc:665, 744, 1093): 1 for Brill-waves and Punctures_EM, 0
// otherwise.
#pragma once

#include <map>
#include <string>
#include <vector>

enum t_equation {
  PUNCTURES = 0, PUNCTURES_SCALAR = 1, POISSON = 2, BRILL_WAVES = 3, PUNCTURES_EM = 4,
  OTHER_EQUATION = 99
};
enum t_boundary { DIRICHLET, ROBIN, ASYMPTOTIC };
enum t_method { GAUSS_SEIDEL, FMG, VCYCLE, WCYCLE, APP };
enum t_output { YES, PARTIAL, NO };

class interface {
 public:
  interface(int argc, char *argv[]);

  bool Get_logic_par(const std::string &name) const;
  size_t Get_integer_par(const std::string &name) const;
  double Get_double_par(const std::string &name) const;
  std::string Get_string_par(const std::string &name) const;

  std::vector<double> Get_bhmp() const { return bhmp; }
  std::vector<double> Get_bhqp() const { return bhqp; }
  std::vector<double> Get_bhx() const { return bhx; }
  std::vector<double> Get_bhy() const { return bhy; }
  std::vector<double> Get_bhz() const { return bhz; }
  std::vector<double> Get_bhpx() const { return bhpx; }
  std::vector<double> Get_bhpy() const { return bhpy; }
  std::vector<double> Get_bhpz() const { return bhpz; }
  std::vector<double> Get_bhsx() const { return bhsx; }
  std::vector<double> Get_bhsy() const { return bhsy; }
  std::vector<double> Get_bhsz() const { return bhsz; }

  t_equation Get_Equation() const { return equation; }
  t_boundary Get_Boundary() const { return boundary; }
  t_method Get_Method() const { return method; }
  t_output Get_TOutput() const { return output; }
  // false when the command line asks for something outside this solver or
  // is malformed; the message says what (the reference prints help, exit 0)
  bool Valid() const { return exec; }
  const std::string &Error() const { return error; }
  const std::string &Command_Line() const { return param; }

 private:
  struct Keys {
    std::string k1, k2, help;
  };
  void add_double(const std::string &n, double v, const std::string &k1, const std::string &k2,
                  const std::string &help);
  void add_int(const std::string &n, size_t v, const std::string &k1, const std::string &k2,
               const std::string &help);
  void add_logic(const std::string &n, bool v, const std::string &k1, const std::string &k2,
                 const std::string &help);
  void add_string(const std::string &n, const std::string &v, const std::string &k1,
                  const std::string &k2, const std::string &help);
  void read_console(int argc, char *argv[]);
  void proc_parameters();
  void fail(const std::string &msg);

  std::string param;
  std::vector<std::string> d_names, i_names, l_names, s_names;
  std::map<std::string, double> d_par;
  std::map<std::string, size_t> i_par;
  std::map<std::string, bool> l_par;
  std::map<std::string, std::string> s_par;
  std::map<std::string, Keys> keys;
  bool by_domain = false, by_points = false, by_grid_size = false, robin_value_given = false;
  bool exec = true;
  std::string error;
  std::vector<double> bhmp, bhqp, bhx, bhy, bhz, bhpx, bhpy, bhpz, bhsx, bhsy, bhsz;
  t_equation equation = OTHER_EQUATION;
  t_boundary boundary = DIRICHLET;
  t_method method = VCYCLE;
  t_output output = YES;
};

// interface.cc:883-949
void Valid_N(size_t &N);
