// Legacy VTK output of the flux moments (the reference's output_results,
// src/ouranos.cc:211-253, which writes deal.II DataOut patches).
//
// One file `<filename>0000.vtk` (subdomain 0, four digits as the reference
// names them) and a VisIt record `<filename>.vtk` listing it.  Every mesh
// cell is one VTK cell with its own 2^dim corner points (the DG fields are
// discontinuous across faces, so corners are not shared); the point fields
// flux_<ggg>g_<mmm>m hold the corner values of each group and moment, which
// are dof values because the equidistant element nodes include the corners.
// The cell field `partitioning` is the owning subdomain (0).
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "ou_common.hpp"

namespace ou {

void write_vtk(const std::string &filename, int dim, int p, const int *n_cells,
               const double *cell_size, const double *origin, int n_groups, int n_mom,
               const double *flux) {
  if (dim != 2 && dim != 3) fail(OU_ERR_ARG, "dim must be 2 or 3");
  if (p < 1 || p > 5) fail(OU_ERR_ARG, "FE order must be in 1..5");
  if (n_groups < 1 || n_mom < 1) fail(OU_ERR_ARG, "need at least one group and one moment");
  const int nx = n_cells[0], ny = n_cells[1], nz = dim == 3 ? n_cells[2] : 1;
  if (nx < 1 || ny < 1 || nz < 1) fail(OU_ERR_ARG, "cell counts must be positive");
  const int P1 = p + 1, N = dim == 2 ? P1 * P1 : P1 * P1 * P1, NV = dim == 2 ? 4 : 8;
  const long cells = long(nx) * ny * nz, n_dofs = cells * N;
  // VTK_QUAD / VTK_HEXAHEDRON corner order
  static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                   {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const std::string piece = filename + "0000.vtk";
  std::ofstream out(piece);
  if (!out) fail(OU_ERR_ARG, "cannot write " + piece);
  out.precision(17);
  out << "# vtk DataFile Version 3.0\n"
      << "flux moments, " << n_groups << " group(s) x " << n_mom << " moment(s)\n"
      << "ASCII\nDATASET UNSTRUCTURED_GRID\n\n";
  out << "POINTS " << cells * NV << " double\n";
  for (long c = 0; c < cells; ++c) {
    const long i = c % nx, j = (c / nx) % ny, k = c / (long(nx) * ny);
    for (int v = 0; v < NV; ++v) {
      const double x = origin[0] + double(i + corner[v][0]) * cell_size[0];
      const double y = origin[1] + double(j + corner[v][1]) * cell_size[1];
      const double z = dim == 3 ? origin[2] + double(k + corner[v][2]) * cell_size[2] : 0.0;
      out << x << ' ' << y << ' ' << z << '\n';
    }
  }
  out << "\nCELLS " << cells << ' ' << cells * (NV + 1) << '\n';
  for (long c = 0; c < cells; ++c) {
    out << NV;
    for (int v = 0; v < NV; ++v) out << ' ' << c * NV + v;
    out << '\n';
  }
  out << "\nCELL_TYPES " << cells << '\n';
  for (long c = 0; c < cells; ++c) out << (dim == 2 ? 9 : 12) << '\n';
  out << "\nCELL_DATA " << cells << "\nSCALARS partitioning double 1\nLOOKUP_TABLE default\n";
  for (long c = 0; c < cells; ++c) out << "0\n";
  out << "\nPOINT_DATA " << cells * NV << '\n';
  for (int g = 0; g < n_groups; ++g)
    for (int m = 0; m < n_mom; ++m) {
      char name[32];
      std::snprintf(name, sizeof name, "flux_%03dg_%03dm", g, m);
      out << "SCALARS " << name << " double 1\nLOOKUP_TABLE default\n";
      const double *f = flux + (long(g) * n_mom + m) * n_dofs;
      for (long c = 0; c < cells; ++c)
        for (int v = 0; v < NV; ++v) {
          const int local = corner[v][0] * p + P1 * corner[v][1] * p +
                            (dim == 3 ? P1 * P1 * corner[v][2] * p : 0);
          out << f[c * N + local] << '\n';
        }
    }
  if (!out) fail(OU_ERR_ARG, "writing " + piece + " failed");
  out.close();
  std::ofstream record(filename + ".vtk");
  if (!record) fail(OU_ERR_ARG, "cannot write " + filename + ".vtk");
  const size_t slash = piece.find_last_of('/');
  record << "!NBLOCKS 1\n" << (slash == std::string::npos ? piece : piece.substr(slash + 1))
         << '\n';
}

}  // namespace ou
