// sdot_prm.cpp -- reader of deal.II parameter files in the grammar of the
// reference's SotParameterManager (ParameterManager.cc:19-146, 412-422;
// fields and defaults ParameterManager.h:42-121, 202-205): the
// `subsection rsot_solver` block (sdot_params_parse_prm) and the whole file
// (sdot_prm_file_parse).  Grammar: `subsection <name>` ... `end`,
// `set <key> = <value>`, `#` comments, trailing `\` joins lines.  The block is
// recognised at any nesting depth (run/parameters.prm nests it inside
// `subsection SotParameterManager`).  String-valued keys follow the
// reference's reading: solver_control_type is componentwise only for
// "componentwise" (SotSolver.cc:156), distance_threshold_type falls back to
// pointwise for anything but "integral"/"geometric" (SotSolver.cc:392-406).
// `precision` (fp64 | fp32_lse), `line_search` (reference | noise_aware) and
// `preconditioner` (none | diagonal) and `prune_bits` (a number) are this
// library's own keys.
#include <cctype>
#include <climits>
#include <cmath>
#include <cstring>
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
  p->prune_bits = -1.0;
  return SDOT_OK;
}

namespace {

// One `set key = value` line with the subsection path it sits in.
struct Entry {
  std::vector<std::string> path;
  std::string key, value;
};

// Splits a parameter file into its entries; false on a grammar error.
bool tokenize(const char *text, std::vector<Entry> &out) {
  std::istringstream in(text);
  std::string raw, line;
  std::vector<std::string> stack;
  while (std::getline(in, raw)) {
    // join continuation lines
    while (!raw.empty() && raw.back() == '\\') {
      raw.pop_back();
      std::string more;
      if (!std::getline(in, more)) break;
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
      if (stack.empty()) return false;
      stack.pop_back();
      continue;
    }
    if (line.compare(0, 4, "set ") != 0) return false;
    const size_t eq = line.find('=');
    if (eq == std::string::npos) return false;
    out.push_back({stack, trim(line.substr(4, eq - 4)), trim(line.substr(eq + 1))});
  }
  return stack.empty();
}

// One rsot_solver key into p; false for an undeclared key or a bad value.
bool solver_key(const std::string &key, const std::string &val, sdot_params *p) {
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
  } else if (key == "prune_bits") {
    char *end = nullptr;
    const double v = std::strtod(val.c_str(), &end);
    ok = end && *end == '\0' && !val.empty() && std::isfinite(v);
    if (ok) p->prune_bits = v;
  } else
    ok = false;  // deal.II rejects undeclared entries
  return ok;
}

template <size_t K>
bool put_str(const std::string &v, char (&dst)[K]) {
  if (v.size() >= K) return false;
  std::memcpy(dst, v.c_str(), v.size() + 1);
  return true;
}

bool parse_int(const std::string &v, int32_t &out) {
  char *end = nullptr;
  const long long x = std::strtoll(v.c_str(), &end, 10);
  if (!end || end == v.c_str() || !trim(end).empty() || x < INT32_MIN || x > INT32_MAX) return false;
  out = (int32_t)x;
  return true;
}

// A comma-separated list of unsigned integers (deal.II Patterns::List of
// Patterns::Integer(0)); an empty value is the empty list.
bool parse_uint_list(const std::string &v, uint32_t *out, int32_t cap, int32_t &n) {
  n = 0;
  if (trim(v).empty()) return true;
  std::stringstream ss(v);
  std::string item;
  while (std::getline(ss, item, ',')) {
    uint32_t x;
    if (n >= cap || !parse_uint(trim(item), x)) return false;
    out[n++] = x;
  }
  return true;
}

bool mesh_key(const std::string &key, const std::string &val, sdot_mesh_params &m) {
  if (key == "number of refinements") return parse_uint(val, m.n_refinements);
  if (key == "grid generator function") return put_str(val, m.grid_generator_function);
  if (key == "grid generator arguments") return put_str(val, m.grid_generator_arguments);
  if (key == "use tetrahedral mesh") return parse_bool(val, m.use_tetrahedral_mesh);
  if (key == "use custom density") return parse_bool(val, m.use_custom_density);
  if (key == "density file path") return put_str(val, m.density_file_path);
  if (key == "density file format") return put_str(val, m.density_file_format);
  if (key == "density field name") return put_str(val, m.density_field_name);
  return false;
}

bool multilevel_key(const std::string &key, const std::string &val, sdot_multilevel_params &m) {
  if (key == "source_enabled") return parse_bool(val, m.source_enabled);
  if (key == "source_min_vertices") return parse_int(val, m.source_min_vertices);
  if (key == "source_max_vertices") return parse_int(val, m.source_max_vertices);
  if (key == "source_hierarchy_dir") return put_str(val, m.source_hierarchy_dir);
  if (key == "target_enabled") return parse_bool(val, m.target_enabled);
  if (key == "target_min_points") return parse_int(val, m.target_min_points);
  if (key == "target_max_points") return parse_int(val, m.target_max_points);
  if (key == "target_hierarchy_dir") return put_str(val, m.target_hierarchy_dir);
  if (key == "use_softmax_potential_transfer")
    return parse_bool(val, m.use_softmax_potential_transfer);
  if (key == "use_python_clustering") return parse_bool(val, m.use_python_clustering);
  if (key == "python_script_name") return put_str(val, m.python_script_name);
  if (key == "output_prefix") return put_str(val, m.output_prefix);
  return false;
}

// The entry's path relative to the manager's own subsection: the entries of
// `subsection SotParameterManager` at any depth (the tutorials nest it in
// their own subsection) and top-level ones (a file may leave the manager's
// name out); nullptr for entries of other objects.
const std::vector<std::string> *manager_relative(const Entry &e, std::vector<std::string> &tmp) {
  for (size_t i = 0; i < e.path.size(); ++i)
    if (e.path[i] == "SotParameterManager") {
      tmp.assign(e.path.begin() + i + 1, e.path.end());
      return &tmp;
    }
  static const char *kOwn[] = {"mesh_generation", "rsot_solver", "multilevel_parameters",
                               "power_diagram_parameters", "transport_map_parameters",
                               "conditional_density_parameters", "lloyd parameters"};
  if (e.path.empty()) return &e.path;
  for (const char *k : kOwn)
    if (e.path[0] == k) return &e.path;
  return nullptr;
}

}  // namespace

extern "C" int sdot_params_parse_prm(const char *text, sdot_params *p) {
  if (!text || !p) return SDOT_ERR_INVALID;
  std::vector<Entry> entries;
  if (!tokenize(text, entries)) return SDOT_ERR_PARSE;
  for (const Entry &e : entries) {
    // the block is recognised at any nesting depth
    if (e.path.empty() || e.path.back() != "rsot_solver") continue;
    if (!solver_key(e.key, e.value, p)) return SDOT_ERR_PARSE;
  }
  return SDOT_OK;
}

extern "C" int sdot_prm_file_default(sdot_prm_file *f) {
  if (!f) return SDOT_ERR_INVALID;
  std::memset(f, 0, sizeof(*f));
  sdot_params_default(&f->solver);
  f->io_coding = 0;
  for (sdot_mesh_params *m : {&f->source, &f->target}) {
    put_str("vtk", m->density_file_format);
    put_str("normalized_density", m->density_field_name);
  }
  sdot_multilevel_params &ml = f->multilevel;
  ml.source_min_vertices = 1000;
  ml.source_max_vertices = 10000;
  put_str("output/data_multilevel/source_multilevel", ml.source_hierarchy_dir);
  ml.target_min_points = 100;
  ml.target_max_points = 1000;
  put_str("output/data_multilevel/target_multilevel", ml.target_hierarchy_dir);
  ml.use_softmax_potential_transfer = 1;
  put_str("multilevel_clustering_scipy.py", ml.python_script_name);
  put_str("output/multilevel/sot", ml.output_prefix);
  put_str("dealii", f->power_diagram_implementation);
  f->transport_map_truncation_radius = -1.0;
  f->density_truncation_radius = -1.0;
  f->lloyd_max_iterations = 1000;
  f->lloyd_relative_tolerance = 1e-8;
  return SDOT_OK;
}

extern "C" int sdot_prm_file_parse(const char *text, sdot_prm_file *f) {
  if (!text || !f) return SDOT_ERR_INVALID;
  std::vector<Entry> entries;
  if (!tokenize(text, entries)) return SDOT_ERR_PARSE;
  std::vector<std::string> tmp;
  for (const Entry &e : entries) {
    const std::vector<std::string> *rel = manager_relative(e, tmp);
    if (!rel) continue;
    const std::vector<std::string> &r = *rel;
    const std::string &k = e.key, &v = e.value;
    bool ok;
    if (r.empty()) {
      if (k == "selected_task") ok = put_str(v, f->selected_task);
      else if (k == "io_coding") {
        ok = true;
        if (v == "txt") f->io_coding = 0;
        else if (v == "bin") f->io_coding = 1;
        else ok = false;
      } else ok = false;
    } else if (r.size() == 1 && r[0] == "rsot_solver") {
      ok = solver_key(k, v, &f->solver);
    } else if (r.size() == 2 && r[0] == "mesh_generation" && (r[1] == "source" || r[1] == "target")) {
      ok = mesh_key(k, v, r[1] == "source" ? f->source : f->target);
    } else if (r.size() == 1 && r[0] == "multilevel_parameters") {
      ok = multilevel_key(k, v, f->multilevel);
    } else if (r.size() == 1 && r[0] == "power_diagram_parameters") {
      ok = k == "implementation" && put_str(v, f->power_diagram_implementation);
    } else if (r.size() == 1 && r[0] == "transport_map_parameters") {
      ok = k == "truncation_radius" && parse_double(v, f->transport_map_truncation_radius);
    } else if (r.size() == 1 && r[0] == "conditional_density_parameters") {
      if (k == "indices")
        ok = parse_uint_list(v, f->density_indices, SDOT_MAX_DENSITY_INDICES, f->n_density_indices);
      else if (k == "potential_folder") ok = put_str(v, f->density_potential_folder);
      else if (k == "truncation_radius") ok = parse_double(v, f->density_truncation_radius);
      else ok = false;
    } else if (r.size() == 1 && r[0] == "lloyd parameters") {
      if (k == "max iterations") ok = parse_uint(v, f->lloyd_max_iterations);
      else if (k == "relative tolerance") ok = parse_double(v, f->lloyd_relative_tolerance);
      else ok = false;
    } else {
      ok = false;  // a subsection the manager does not declare
    }
    if (!ok) return SDOT_ERR_PARSE;
  }
  return SDOT_OK;
}

extern "C" int sdot_prm_file_load(const char *path, sdot_prm_file *f) {
  if (!path || !f) return SDOT_ERR_INVALID;
  std::ifstream in(path);
  if (!in) return SDOT_ERR_INVALID;
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string s = ss.str();
  return sdot_prm_file_parse(s.c_str(), f);
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
