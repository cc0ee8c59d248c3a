// Readers of the reference's input files (see inp.hpp).
#include "inp.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "ou_common.hpp"

namespace ou {
namespace {

std::string trim(const std::string &s) {
  const size_t b = s.find_first_not_of(" \t\r\n");
  if (b == std::string::npos) return "";
  const size_t e = s.find_last_not_of(" \t\r\n");
  return s.substr(b, e - b + 1);
}

// Runs of blanks inside a name count as one blank.
std::string squeeze(const std::string &s) {
  std::string out;
  bool blank = false;
  for (char c : trim(s)) {
    if (c == ' ' || c == '\t') {
      blank = true;
      continue;
    }
    if (blank && !out.empty()) out += ' ';
    blank = false;
    out += c;
  }
  return out;
}

bool parse_long(const std::string &t, long &v) {
  const std::string s = trim(t);
  if (s.empty()) return false;
  char *end = nullptr;
  errno = 0;
  v = std::strtol(s.c_str(), &end, 10);
  return errno == 0 && end && *end == '\0';
}

bool parse_double(const std::string &t, double &v) {
  const std::string s = trim(t);
  if (s.empty()) return false;
  char *end = nullptr;
  errno = 0;
  v = std::strtod(s.c_str(), &end);
  return errno == 0 && end && *end == '\0' && std::isfinite(v);
}

std::vector<std::string> split(const std::string &s, char sep) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : s) {
    if (c == sep) {
      out.push_back(cur);
      cur.clear();
    } else {
      cur += c;
    }
  }
  out.push_back(cur);
  return out;
}

// Does `value` match the entry's pattern (deal.II Patterns::Integer, Double,
// Selection, Bool, FileName, List)?
bool matches(const InpEntry &e, const std::string &value) {
  switch (e.kind) {
    case InpEntry::INTEGER: {
      long v;
      return parse_long(value, v) && v >= e.lo && v <= e.hi;
    }
    case InpEntry::DOUBLE: {
      double v;
      return parse_double(value, v) && v >= e.lo && v <= e.hi;
    }
    case InpEntry::SELECTION: {
      for (const std::string &o : split(e.options, '|'))
        if (o == value) return true;
      return false;
    }
    case InpEntry::BOOL:
      return value == "true" || value == "false" || value == "yes" || value == "no";
    case InpEntry::FILENAME:
      return true;
    case InpEntry::LIST_DOUBLE:
    case InpEntry::LIST_INTEGER: {
      if (trim(value).empty()) return true;
      for (const std::string &item : split(value, ',')) {
        if (e.kind == InpEntry::LIST_DOUBLE) {
          double v;
          if (!parse_double(item, v) || v < e.lo || v > e.hi) return false;
        } else {
          long v;
          if (!parse_long(item, v) || v < e.lo || v > e.hi) return false;
        }
      }
      return true;
    }
  }
  return false;
}

InpEntry integer(const std::string &name, const std::string &def, double lo,
                 double hi = 2147483647.0) {
  InpEntry e{name, def, InpEntry::INTEGER};
  e.lo = lo;
  e.hi = hi;
  return e;
}
InpEntry real(const std::string &name, const std::string &def, double lo = -1e300,
              double hi = 1e300) {
  InpEntry e{name, def, InpEntry::DOUBLE};
  e.lo = lo;
  e.hi = hi;
  return e;
}
InpEntry selection(const std::string &name, const std::string &def, const std::string &opts) {
  InpEntry e{name, def, InpEntry::SELECTION};
  e.options = opts;
  return e;
}
InpEntry flag(const std::string &name, const std::string &def) {
  return InpEntry{name, def, InpEntry::BOOL};
}
InpEntry filename(const std::string &name, const std::string &def) {
  return InpEntry{name, def, InpEntry::FILENAME};
}
InpEntry list(const std::string &name, const std::string &def, InpEntry::Kind kind,
              double lo = -1e300) {
  InpEntry e{name, def, kind};
  e.lo = lo;
  return e;
}

int bc_of(const std::string &v) {
  if (v == "vacuum") return INP_VACUUM;
  if (v == "reflective") return INP_REFLECTIVE;
  if (v == "isotropic") return INP_ISOTROPIC;
  return INP_MOST_NORMAL;
}

}  // namespace

InpFile::InpFile(const std::string &path, const std::vector<InpEntry> &declared,
                 const std::string &what) {
  std::ifstream f(path);
  if (!f) fail(OU_ERR_ARG, what + " file not found: " + path);
  std::map<std::string, const InpEntry *> table;
  for (const InpEntry &e : declared) {
    table[e.name] = &e;
    values_[e.name] = e.def;
  }
  auto bad = [&](int line, const std::string &why) {
    fail(OU_ERR_ARG, "invalid " + what + " file " + path + ", line " + std::to_string(line) +
                         ": " + why);
  };
  std::vector<std::string> sub;
  std::string raw, acc;
  int lineno = 0, start = 0;
  while (std::getline(f, raw)) {
    ++lineno;
    const size_t hash = raw.find('#');
    if (hash != std::string::npos) raw.erase(hash);
    std::string t = trim(raw);
    if (acc.empty()) start = lineno;
    if (!t.empty() && t.back() == '\\') {
      acc += t.substr(0, t.size() - 1) + " ";
      continue;
    }
    t = trim(acc + t);
    acc.clear();
    if (t.empty()) continue;
    if (t.compare(0, 11, "subsection ") == 0) {
      sub.push_back(squeeze(t.substr(11)));
    } else if (t == "end") {
      if (sub.empty()) bad(start, "'end' without a subsection");
      sub.pop_back();
    } else if (t.compare(0, 4, "set ") == 0) {
      const size_t eq = t.find('=');
      if (eq == std::string::npos) bad(start, "'set' without '='");
      std::string key;
      for (const std::string &s : sub) key += s + "/";
      key += squeeze(t.substr(4, eq - 4));
      const std::string value = trim(t.substr(eq + 1));
      const auto it = table.find(key);
      if (it == table.end()) bad(start, "no entry '" + key + "' is declared");
      if (!matches(*it->second, value))
        bad(start, "value '" + value + "' of '" + key + "' does not match its pattern");
      values_[key] = value;
    } else {
      bad(start, "cannot read '" + t + "'");
    }
  }
  if (!acc.empty()) bad(start, "the file ends inside a continued line");
  if (!sub.empty()) bad(lineno, "subsection '" + sub.back() + "' is not closed");
}

const std::string &InpFile::get(const std::string &name) const {
  const auto it = values_.find(name);
  if (it == values_.end()) fail(OU_ERR_ARG, "no entry '" + name + "' is declared");
  return it->second;
}

long InpFile::get_integer(const std::string &name) const {
  long v = 0;
  if (!parse_long(get(name), v)) fail(OU_ERR_ARG, "'" + name + "' is not an integer");
  return v;
}

double InpFile::get_double(const std::string &name) const {
  double v = 0;
  if (!parse_double(get(name), v)) fail(OU_ERR_ARG, "'" + name + "' is not a number");
  return v;
}

bool InpFile::get_bool(const std::string &name) const {
  const std::string &v = get(name);
  return v == "true" || v == "yes";
}

std::vector<double> list_double(const std::string &input, int n) {
  std::string s = input;
  std::replace(s.begin(), s.end(), ',', ' ');
  std::vector<double> values(size_t(std::max(n, 0)), 0.0);
  const char *p = s.c_str();
  for (int i = 0; i < n; ++i) {
    char *end = nullptr;
    values[size_t(i)] = std::strtod(p, &end);
    p = end;
  }
  return values;
}

std::vector<long> list_integer(const std::string &input, int n) {
  std::string s = input;
  std::replace(s.begin(), s.end(), ',', ' ');
  std::vector<long> values(size_t(std::max(n, 0)), 0);
  const char *p = s.c_str();
  for (int i = 0; i < n; ++i) {
    char *end = nullptr;
    values[size_t(i)] = std::strtol(p, &end, 0);  // base 0, as Geometry.cc:140-150
    p = end;
  }
  return values;
}

Parameters::Parameters(const std::string &path) {
  const std::string rt = "Radiative Transfer parameters/";
  const std::string bcs = "vacuum|isotropic|most normal|reflective";
  std::vector<InpEntry> d = {
      integer("Dimension", "2", 2, 3),
      integer("FE order", "1", 1, 5),
      filename("Geometry file", "geometry.inp"),
      integer("Number of refinements", "0", 0),
      real("Refinement factor", "0.", 0.0, 1.0),
      real("Coarsening factor", "0.", 0.0, 1.0),
      filename("Output file", "output"),
      filename(rt + "Cross sections file", "xs.inp"),
      selection(rt + "Solver type", "SI", "SI|GMRES|BICGSTAB"),
      integer(rt + "Max outer iterations", "1", 1),
      integer(rt + "Max inner iterations", "1", 1),
      real(rt + "Tolerance outer solver", "1e-4", 0.0),
      real(rt + "Tolerance inner solver", "1e-6", 0.0),
      integer(rt + "Sn order", "2", 2),
      selection(rt + "Quadrature type", "GLC", "LS|GLC"),
      selection(rt + "Sum weight", "4PI", "1|2PI|4PI"),
      flag(rt + "Galerkin", "false"),
  };
  const char *faces[6] = {"Xmin", "Xmax", "Ymin", "Ymax", "Zmin", "Zmax"};
  for (const char *fc : faces) d.push_back(selection(rt + fc + " BC", "vacuum", bcs));
  for (const char *fc : faces)
    d.push_back(list(rt + fc + " BC values", "0", InpEntry::LIST_DOUBLE, 0.0));
  d.push_back(integer(rt + "Number of sources", "0", 0));
  d.push_back(integer(rt + "Number of groups", "1", 1));
  d.push_back(list(rt + "Source intensity", "1.", InpEntry::LIST_DOUBLE, 0.0));
  const InpFile prm(path, d, "parameters");

  dim = int(prm.get_integer("Dimension"));
  fe_order = int(prm.get_integer("FE order"));
  geometry_filename = prm.get("Geometry file");
  n_refinements = int(prm.get_integer("Number of refinements"));
  refinement_factor = prm.get_double("Refinement factor");
  coarsening_factor = prm.get_double("Coarsening factor");
  output_filename = prm.get("Output file");
  xs_filename = prm.get(rt + "Cross sections file");
  const std::string solver = prm.get(rt + "Solver type");
  solver_type = solver == "SI" ? INP_SI : solver == "BICGSTAB" ? INP_BICGSTAB : INP_GMRES;
  max_out_it = int(prm.get_integer(rt + "Max outer iterations"));
  max_in_it = int(prm.get_integer(rt + "Max inner iterations"));
  outer_tol = prm.get_double(rt + "Tolerance outer solver");
  inner_tol = prm.get_double(rt + "Tolerance inner solver");
  sn_order = int(prm.get_integer(rt + "Sn order"));
  quad_type = prm.get(rt + "Quadrature type") == "LS" ? INP_LS : INP_GLC;
  const std::string ws = prm.get(rt + "Sum weight");
  weight_sum = ws == "1" ? 1.0 : ws == "2PI" ? 2.0 * M_PI : 4.0 * M_PI;
  galerkin = prm.get_bool(rt + "Galerkin");
  bc_type.assign(size_t(2 * dim), INP_VACUUM);
  for (int f = 0; f < 2 * dim; ++f) bc_type[size_t(f)] = bc_of(prm.get(rt + faces[f] + " BC"));
  n_groups = int(prm.get_integer(rt + "Number of groups"));
  inc_flux.assign(size_t(2 * dim), std::vector<double>(size_t(n_groups), 0.0));
  for (int f = 0; f < 2 * dim; ++f)
    if (bc_type[size_t(f)] == INP_ISOTROPIC || bc_type[size_t(f)] == INP_MOST_NORMAL)
      inc_flux[size_t(f)] = list_double(prm.get(rt + faces[f] + " BC values"), n_groups);
  n_src = int(prm.get_integer(rt + "Number of sources"));
  src.assign(size_t(n_src), std::vector<double>(size_t(n_groups), 0.0));
  const std::vector<double> v = list_double(prm.get(rt + "Source intensity"), n_src * n_groups);
  for (int i = 0; i < n_src; ++i)
    for (int g = 0; g < n_groups; ++g)
      src[size_t(i)][size_t(g)] = v[size_t(i) * n_groups + g];
}

Geometry::Geometry(const std::string &path, int dim_) : dim(dim_) {
  if (dim != 2 && dim != 3) fail(OU_ERR_ARG, "geometry dimension must be 2 or 3");
  const std::vector<InpEntry> d = {
      list("Bottom left point", "0.,0.", InpEntry::LIST_DOUBLE),
      list("Upper right point", "0.,0.", InpEntry::LIST_DOUBLE),
      list("Number of subdivisions", "1,1", InpEntry::LIST_INTEGER, 1),
      integer("Number of global refinements", "0", 0),
      list("Material IDs", "0", InpEntry::LIST_INTEGER, 0),
      list("Source IDs", "0", InpEntry::LIST_INTEGER, 0),
  };
  const InpFile prm(path, d, "geometry");
  bottom_left = list_double(prm.get("Bottom left point"), dim);
  upper_right = list_double(prm.get("Upper right point"), dim);
  const std::vector<long> ns = list_integer(prm.get("Number of subdivisions"), dim);
  n_subdivisions.assign(ns.begin(), ns.end());
  n_global_refinements = int(prm.get_integer("Number of global refinements"));
  long n_sub = 1;
  for (int a = 0; a < dim; ++a) {
    if (n_subdivisions[size_t(a)] < 1) fail(OU_ERR_ARG, "geometry: subdivisions must be >= 1");
    n_sub *= n_subdivisions[size_t(a)];
  }
  // cells per axis after the refinements must fit an int
  for (int a = 0; a < dim; ++a)
    if (n_global_refinements > 30 ||
        (long(n_subdivisions[size_t(a)]) << n_global_refinements) > (1L << 30))
      fail(OU_ERR_ARG, "geometry: too many cells per axis after the global refinements");
  const std::vector<long> mids = list_integer(prm.get("Material IDs"), int(n_sub));
  const std::vector<long> sids = list_integer(prm.get("Source IDs"), int(n_sub));
  material_ids.assign(mids.begin(), mids.end());
  source_ids.assign(sids.begin(), sids.end());
  n_materials = *std::max_element(material_ids.begin(), material_ids.end()) + 1;
}

std::vector<int> Geometry::n_cells() const {
  std::vector<int> n(3, 1);
  for (int a = 0; a < dim; ++a)
    n[size_t(a)] = n_subdivisions[size_t(a)] << n_global_refinements;
  return n;
}

std::vector<double> Geometry::cell_size() const {
  const std::vector<int> n = n_cells();
  std::vector<double> h(3, 1.0);
  for (int a = 0; a < dim; ++a)
    h[size_t(a)] = (upper_right[size_t(a)] - bottom_left[size_t(a)]) / n[size_t(a)];
  return h;
}

void Geometry::fine_ids(std::vector<int> &material, std::vector<int> &source) const {
  const std::vector<int> n = n_cells();
  const int r = n_global_refinements;
  const size_t total = size_t(n[0]) * n[1] * n[2];
  material.resize(total);
  source.resize(total);
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        const size_t c = size_t(i) + size_t(n[0]) * (size_t(j) + size_t(n[1]) * k);
        const int ci = i >> r, cj = j >> r, ck = dim == 3 ? k >> r : 0;
        const size_t pc = size_t(ci) + size_t(n_subdivisions[0]) *
                                           (size_t(cj) + size_t(n_subdivisions[1]) * ck);
        material[c] = material_ids[pc];
        source[c] = source_ids[pc];
      }
}

MaterialProperties::MaterialProperties(const std::string &path, int n_materials_,
                                       int n_groups_)
    : n_materials(n_materials_), n_groups(n_groups_) {
  if (n_materials < 1 || n_groups < 1)
    fail(OU_ERR_ARG, "material properties need at least one material and one group");
  const std::vector<InpEntry> d = {
      integer("L_max", "0", 0),
      list("Sigma_t", "1.", InpEntry::LIST_DOUBLE, 0.0),
      list("Sigma_s", "0.", InpEntry::LIST_DOUBLE, 0.0),
  };
  const InpFile prm(path, d, "cross sections");
  L_max = int(prm.get_integer("L_max"));
  sigma_t = list_double(prm.get("Sigma_t"), n_materials * n_groups);
  sigma_s = list_double(prm.get("Sigma_s"), n_materials * n_groups * n_groups * (L_max + 1));
}

}  // namespace ou
