// Readers of the reference's input files: a parameter file, a geometry file
// and a cross-section file, each a subset of deal.II's ParameterHandler
// format (`set <name> = <value>`, `subsection <name>` ... `end`, `#`
// comments, a trailing backslash continues a line).
//
// Parameters            src/Parameters.cc:10-302   (declarations :28-100, parsing :102-302)
// Geometry (the reader) src/Geometry.cc:10-128     (declarations :85-99, parsing :101-128)
// RTMaterialProperties  src/RTMaterialProperties.cc:10-81
#pragma once

#include <map>
#include <string>
#include <vector>

namespace ou {

// One declared entry: name (with its subsection path), default value and the
// pattern the value must match.
struct InpEntry {
  std::string name;
  std::string def;
  enum Kind { INTEGER, DOUBLE, SELECTION, BOOL, FILENAME, LIST_DOUBLE, LIST_INTEGER } kind;
  double lo = -1e300, hi = 1e300;    // INTEGER / DOUBLE / list element bounds
  std::string options;                // SELECTION: "a|b|c"
};

// The values of one file after reading: every declared entry, set or default.
class InpFile {
 public:
  InpFile(const std::string &path, const std::vector<InpEntry> &declared,
          const std::string &what);
  const std::string &get(const std::string &name) const;
  long get_integer(const std::string &name) const;
  double get_double(const std::string &name) const;
  bool get_bool(const std::string &name) const;

 private:
  std::map<std::string, std::string> values_;
};

// The reference's list reader (Parameters.cc:304-328): commas become blanks
// and n values are read with strtod; values past the end of the list are 0.
std::vector<double> list_double(const std::string &input, int n);
std::vector<long> list_integer(const std::string &input, int n);

enum InpSolver { INP_SI = 0, INP_GMRES = 1, INP_BICGSTAB = 2 };
enum InpQuad { INP_LS = 0, INP_GLC = 1 };
enum InpBc { INP_VACUUM = 0, INP_ISOTROPIC = 1, INP_MOST_NORMAL = 2, INP_REFLECTIVE = 3 };

struct Parameters {
  explicit Parameters(const std::string &path);
  int dim, fe_order, n_refinements;
  double refinement_factor, coarsening_factor;
  std::string geometry_filename, output_filename, xs_filename;
  int solver_type, max_out_it, max_in_it;
  double outer_tol, inner_tol;
  int sn_order, quad_type;
  double weight_sum;
  bool galerkin;
  std::vector<int> bc_type;                  // [2*dim]
  int n_groups, n_src;
  std::vector<std::vector<double>> inc_flux;  // [2*dim][n_groups]
  std::vector<std::vector<double>> src;       // [n_src][n_groups]
};

struct Geometry {
  Geometry(const std::string &path, int dim);
  int dim;
  std::vector<double> bottom_left, upper_right;  // [dim]
  std::vector<int> n_subdivisions;              // [dim]
  int n_global_refinements;
  std::vector<int> material_ids, source_ids;    // coarse cells, x fastest
  int n_materials;
  // The mesh after the global refinements: n_subdivisions * 2^refinements
  // cells per axis, each with its coarse cell's ids (a child's centre lies in
  // its parent).  Cells are numbered x fastest.
  std::vector<int> n_cells() const;
  std::vector<double> cell_size() const;
  void fine_ids(std::vector<int> &material, std::vector<int> &source) const;
};

struct MaterialProperties {
  MaterialProperties(const std::string &path, int n_materials, int n_groups);
  int n_materials, n_groups, L_max;
  std::vector<double> sigma_t;  // [mat][g]
  std::vector<double> sigma_s;  // [mat][g][g'][l]  (RTMaterialProperties.hh:92-101)
  double get_sigma_t(int mat, int g) const { return sigma_t[size_t(mat) * n_groups + g]; }
  double get_sigma_s(int mat, int g, int gp, int l) const {
    return sigma_s[size_t(l) + size_t(gp) * (L_max + 1) + size_t(g) * n_groups * (L_max + 1) +
                   size_t(mat) * n_groups * n_groups * (L_max + 1)];
  }
};

}  // namespace ou
