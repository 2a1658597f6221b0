// Field output: cell-centre evaluation of ND_p fields and the VTK writer.
#include "hephaestus/outputs.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>

namespace hephaestus
{

namespace
{

// Lagrange basis on `nodes` evaluated at t
std::vector<double>
lagrange_at(const std::vector<double> & nodes, double t)
{
  std::vector<double> v(nodes.size(), 1.0);
  for (size_t m = 0; m < nodes.size(); m++)
    for (size_t k = 0; k < nodes.size(); k++)
      if (k != m)
        v[m] *= (t - nodes[k]) / (nodes[m] - nodes[k]);
  return v;
}

// Gauss-Legendre points (n = 1..3) and Gauss-Lobatto points (n = 2..4) on
// [0, 1], closed forms of the roots of P_n and of (1 - x^2) P'_{n-1}
std::vector<double>
gauss_points(int n)
{
  switch (n)
  {
  case 1:
    return {0.5};
  case 2:
    return {0.5 - 0.5 / std::sqrt(3.0), 0.5 + 0.5 / std::sqrt(3.0)};
  case 3:
    return {0.5 - 0.5 * std::sqrt(0.6), 0.5, 0.5 + 0.5 * std::sqrt(0.6)};
  default:
    throw Error("gauss_points: orders 1..3");
  }
}

std::vector<double>
lobatto_points(int n)
{
  switch (n)
  {
  case 2:
    return {0.0, 1.0};
  case 3:
    return {0.0, 0.5, 1.0};
  case 4:
    return {0.0, 0.5 - 0.5 / std::sqrt(5.0), 0.5 + 0.5 / std::sqrt(5.0), 1.0};
  default:
    throw Error("lobatto_points: orders 1..3");
  }
}

void
vertex(const StructuredHexMesh & m, int i, int j, int k, double x[3])
{
  if (!m.vert.empty())
  {
    const size_t v = ((size_t)k * (m.ny + 1) + j) * (m.nx + 1) + i;
    for (int a = 0; a < 3; a++)
      x[a] = m.vert[3 * v + a];
    return;
  }
  const int idx[3] = {i, j, k}, n[3] = {m.nx, m.ny, m.nz};
  for (int a = 0; a < 3; a++)
    x[a] = m.x0[a] + m.len[a] * idx[a] / n[a];
}

} // namespace

std::vector<double>
EvaluateCellCenters(const StructuredHexMesh & mesh, int p, const std::vector<double> & dofs)
{
  const int n[3] = {mesh.nx, mesh.ny, mesh.nz};
  int64_t cd[3][3], off[4] = {0, 0, 0, 0};
  for (int c = 0; c < 3; c++)
  {
    for (int a = 0; a < 3; a++)
      cd[c][a] = (a == c) ? (int64_t)n[a] * p : (int64_t)n[a] * p + 1;
    off[c + 1] = off[c] + cd[c][0] * cd[c][1] * cd[c][2];
  }
  if ((int64_t)dofs.size() != off[3])
    throw Error("EvaluateCellCenters: the DOF vector does not match the mesh and order");
  const std::vector<double> bo = lagrange_at(gauss_points(p), 0.5), bc = lagrange_at(lobatto_points(p + 1), 0.5);
  std::vector<double> out(3 * (size_t)mesh.NumElements());
  for (int ez = 0; ez < n[2]; ez++)
    for (int ey = 0; ey < n[1]; ey++)
      for (int ex = 0; ex < n[0]; ex++)
      {
        const int e0[3] = {ex, ey, ez};
        double uh[3];
        for (int c = 0; c < 3; c++)
        {
          const int nd[3] = {c == 0 ? p : p + 1, c == 1 ? p : p + 1, c == 2 ? p : p + 1};
          double s = 0.0;
          for (int dz = 0; dz < nd[2]; dz++)
            for (int dy = 0; dy < nd[1]; dy++)
              for (int dx = 0; dx < nd[0]; dx++)
              {
                const int d[3] = {dx, dy, dz};
                double w = 1.0;
                int64_t g[3];
                for (int a = 0; a < 3; a++)
                {
                  w *= (a == c) ? bo[d[a]] : bc[d[a]];
                  g[a] = (int64_t)e0[a] * p + d[a];
                }
                s += w * dofs[off[c] + (g[2] * cd[c][1] + g[1]) * cd[c][0] + g[0]];
              }
          uh[c] = s;
        }
        // J at the centre: column a = mean of the four element edges along a
        double J[3][3] = {{0}};
        for (int a = 0; a < 3; a++)
          for (int t = 0; t < 4; t++)
          {
            int lo[3] = {ex, ey, ez};
            const int b = (a + 1) % 3, cc = (a + 2) % 3;
            lo[b] += t & 1;
            lo[cc] += (t >> 1) & 1;
            int hi[3] = {lo[0], lo[1], lo[2]};
            hi[a] += 1;
            double xl[3], xh[3];
            vertex(mesh, lo[0], lo[1], lo[2], xl);
            vertex(mesh, hi[0], hi[1], hi[2], xh);
            for (int r = 0; r < 3; r++)
              J[r][a] += 0.25 * (xh[r] - xl[r]);
          }
        // u = J^-T u_hat: solve J^T u = u_hat (Cramer's rule)
        double A[3][3];
        for (int r = 0; r < 3; r++)
          for (int q = 0; q < 3; q++)
            A[r][q] = J[q][r];
        const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                           A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                           A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
        const size_t e = ((size_t)ez * n[1] + ey) * n[0] + ex;
        if (!(std::fabs(det) > 0.0) || !std::isfinite(det))
          throw Error("EvaluateCellCenters: degenerate element " + std::to_string(e));
        for (int k = 0; k < 3; k++)
        {
          double M[3][3];
          for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++)
              M[r][q] = (q == k) ? uh[r] : A[r][q];
          const double dk = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                            M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                            M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
          out[3 * e + k] = dk / det;
        }
      }
  return out;
}

void
VTKDataCollection::Save()
{
  if (_fields.empty())
    throw Error("VTKDataCollection::Save: no registered fields");
  const FESpace * fes = _fields[0].second->ParFESpace();
  const StructuredHexMesh & m = *fes->Mesh();
  char buf[64];
  std::snprintf(buf, sizeof(buf), "_%06d.vtk", _cycle);
  _last = _prefix + buf;
  std::ofstream f(_last);
  if (!f)
    throw Error("VTKDataCollection::Save: cannot open " + _last);
  f.precision(17);
  const int64_t E = m.NumElements();
  f << "# vtk DataFile Version 3.0\nhephaestus cycle " << _cycle << " time " << _time << "\nASCII\n"
    << "DATASET STRUCTURED_GRID\nDIMENSIONS " << m.nx + 1 << " " << m.ny + 1 << " " << m.nz + 1 << "\n"
    << "POINTS " << (int64_t)(m.nx + 1) * (m.ny + 1) * (m.nz + 1) << " double\n";
  for (int k = 0; k <= m.nz; k++)
    for (int j = 0; j <= m.ny; j++)
      for (int i = 0; i <= m.nx; i++)
      {
        double x[3];
        vertex(m, i, j, k, x);
        f << x[0] << " " << x[1] << " " << x[2] << "\n";
      }
  f << "CELL_DATA " << E << "\nSCALARS attribute int 1\nLOOKUP_TABLE default\n";
  for (int64_t e = 0; e < E; e++)
    f << (m.attributes.empty() ? 1 : m.attributes[e]) << "\n";
  for (auto & nf : _fields)
  {
    const std::vector<double> u = EvaluateCellCenters(m, nf.second->ParFESpace()->Order(), nf.second->Download());
    f << "VECTORS " << nf.first << " double\n";
    for (int64_t e = 0; e < E; e++)
      f << u[3 * e] << " " << u[3 * e + 1] << " " << u[3 * e + 2] << "\n";
  }
  if (!f)
    throw Error("VTKDataCollection::Save: write failed for " + _last);
}

void
Outputs::Reset()
{
  if (!_gridfunctions)
    throw Error("Outputs::Reset: no grid functions (construct with them or call Init)");
  _cycle = 0;
  const std::vector<std::string> names =
      _output_field_names.empty() ? _gridfunctions->Names() : _output_field_names;
  for (auto & dcn : Names())
    for (auto & n : names)
      Get(dcn)->RegisterField(n, _gridfunctions->Get(n));
  WriteOutputFields(0.0);
}

void
Outputs::Write(double t)
{
  _cycle++;
  WriteOutputFields(t);
}

void
Outputs::WriteOutputFields(double t)
{
  for (auto & dcn : Names())
  {
    DataCollection * dc = Get(dcn);
    dc->SetCycle(_cycle);
    dc->SetTime(t);
    dc->Save();
  }
}

} // namespace hephaestus
