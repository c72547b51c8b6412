// sdot_prm.cpp -- reader for the `subsection rsot_solver` block of a deal.II
// parameter file (keys and defaults: ParameterManager.cc:64-95,
// ParameterManager.h:56-72).  Grammar: `subsection <name>` ... `end`,
// `set <key> = <value>`, `#` comments, trailing `\` joins lines.  The block is
// recognised at any nesting depth (run/parameters.prm nests it inside
// `subsection SotParameterManager`).  String-valued keys follow the
// reference's reading: solver_control_type is componentwise only for
// "componentwise" (SotSolver.cc:156), distance_threshold_type falls back to
// pointwise for anything but "integral"/"geometric" (SotSolver.cc:392-406).
// `precision` (fp64 | fp32_lse), `line_search` (reference | noise_aware) and
// `preconditioner` (none | diagonal) are this library's own keys.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/sdot.h"

namespace {

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

std::string lower(std::string s) {
  for (auto &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

bool parse_bool(const std::string &v, int32_t &out) {
  const std::string l = lower(v);
  if (l == "true" || l == "yes" || l == "1") { out = 1; return true; }
  if (l == "false" || l == "no" || l == "0") { out = 0; return true; }
  return false;
}

bool parse_double(const std::string &v, double &out) {
  char *end = nullptr;
  out = std::strtod(v.c_str(), &end);
  return end && end != v.c_str() && trim(end).empty();
}

bool parse_uint(const std::string &v, uint32_t &out) {
  char *end = nullptr;
  const long long x = std::strtoll(v.c_str(), &end, 10);
  if (!end || end == v.c_str() || !trim(end).empty() || x < 0 || x > 0xFFFFFFFFll) return false;
  out = (uint32_t)x;
  return true;
}

}  // namespace

extern "C" int sdot_params_default(sdot_params *p) {
  if (!p) return SDOT_ERR_INVALID;
  p->max_iterations = 1000;
  p->tolerance = 1.0;
  p->solver_control_type = SDOT_CTRL_L1NORM;
  p->epsilon = 1e-3;
  p->tau = 1e-8;
  p->distance_threshold_type = SDOT_THRESH_POINTWISE;
  p->verbose_output = 1;
  p->solver_type = 0;
  p->quadrature_order = 3;
  p->nb_points = 1000;
  p->number_of_threads = 0;
  p->use_epsilon_scaling = 0;
  p->epsilon_scaling_factor = 2.0;
  p->epsilon_scaling_steps = 5;
  p->use_log_sum_exp_trick = 0;
  p->precision = SDOT_PREC_FP64;
  p->line_search = SDOT_LS_REFERENCE;
  p->preconditioner = SDOT_PRECOND_NONE;
  return SDOT_OK;
}

extern "C" int sdot_params_parse_prm(const char *text, sdot_params *p) {
  if (!text || !p) return SDOT_ERR_INVALID;
  std::istringstream in(text);
  std::string raw, line;
  std::vector<std::string> stack;
  int lineno = 0;
  while (std::getline(in, raw)) {
    ++lineno;
    // join continuation lines
    while (!raw.empty() && raw.back() == '\\') {
      raw.pop_back();
      std::string more;
      if (!std::getline(in, more)) break;
      ++lineno;
      raw += more;
    }
    const size_t hash = raw.find('#');
    line = trim(hash == std::string::npos ? raw : raw.substr(0, hash));
    if (line.empty()) continue;
    if (line.compare(0, 11, "subsection ") == 0) {
      stack.push_back(trim(line.substr(11)));
      continue;
    }
    if (line == "end") {
      if (stack.empty()) return SDOT_ERR_PARSE;
      stack.pop_back();
      continue;
    }
    if (line.compare(0, 4, "set ") != 0) return SDOT_ERR_PARSE;
    const size_t eq = line.find('=');
    if (eq == std::string::npos) return SDOT_ERR_PARSE;
    if (stack.empty() || stack.back() != "rsot_solver") continue;
    const std::string key = trim(line.substr(4, eq - 4));
    const std::string val = trim(line.substr(eq + 1));
    bool ok = true;
    if (key == "max_iterations") ok = parse_uint(val, p->max_iterations);
    else if (key == "tolerance") ok = parse_double(val, p->tolerance);
    else if (key == "solver_control_type")
      p->solver_control_type = val == "componentwise" ? SDOT_CTRL_COMPONENTWISE : SDOT_CTRL_L1NORM;
    else if (key == "epsilon") ok = parse_double(val, p->epsilon);
    else if (key == "tau") ok = parse_double(val, p->tau);
    else if (key == "distance_threshold_type")
      p->distance_threshold_type = val == "integral"    ? SDOT_THRESH_INTEGRAL
                                   : val == "geometric" ? SDOT_THRESH_GEOMETRIC
                                                        : SDOT_THRESH_POINTWISE;
    else if (key == "verbose_output") ok = parse_bool(val, p->verbose_output);
    else if (key == "solver_type") p->solver_type = 0;
    else if (key == "quadrature_order") ok = parse_uint(val, p->quadrature_order);
    else if (key == "number_of_threads") ok = parse_uint(val, p->number_of_threads);
    else if (key == "use_epsilon_scaling") ok = parse_bool(val, p->use_epsilon_scaling);
    else if (key == "epsilon_scaling_factor") ok = parse_double(val, p->epsilon_scaling_factor);
    else if (key == "epsilon_scaling_steps") ok = parse_uint(val, p->epsilon_scaling_steps);
    else if (key == "use_log_sum_exp_trick") ok = parse_bool(val, p->use_log_sum_exp_trick);
    else if (key == "precision") {
      const std::string l = lower(val);
      if (l == "fp64") p->precision = SDOT_PREC_FP64;
      else if (l == "fp32_lse") p->precision = SDOT_PREC_FP32_LSE;
      else ok = false;
    } else if (key == "line_search") {
      const std::string l = lower(val);
      if (l == "reference") p->line_search = SDOT_LS_REFERENCE;
      else if (l == "noise_aware") p->line_search = SDOT_LS_NOISE_AWARE;
      else ok = false;
    } else if (key == "preconditioner") {
      const std::string l = lower(val);
      if (l == "none") p->preconditioner = SDOT_PRECOND_NONE;
      else if (l == "diagonal") p->preconditioner = SDOT_PRECOND_DIAGONAL;
      else ok = false;
    } else
      ok = false;  // deal.II rejects undeclared entries
    if (!ok) return SDOT_ERR_PARSE;
  }
  if (!stack.empty()) return SDOT_ERR_PARSE;
  return SDOT_OK;
}

extern "C" int sdot_params_load_prm(const char *path, sdot_params *p) {
  if (!path || !p) return SDOT_ERR_INVALID;
  std::ifstream f(path);
  if (!f) return SDOT_ERR_INVALID;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  return sdot_params_parse_prm(s.c_str(), p);
}
