// MFEM mesh v1.0 reader / writer for structured hexahedral meshes.
#include "hephaestus/mesh_io.hpp"

#include <array>
#include <fstream>
#include <sstream>
#include <vector>

namespace hephaestus
{

namespace
{

// next non-empty, non-comment line
bool
next_line(std::istream & in, std::string & line)
{
  while (std::getline(in, line))
  {
    const size_t a = line.find_first_not_of(" \t\r");
    if (a == std::string::npos || line[a] == '#')
      continue;
    line = line.substr(a);
    const size_t b = line.find_last_not_of(" \t\r");
    line = line.substr(0, b + 1);
    return true;
  }
  return false;
}

// integer of a header line, or an Error naming the file
int64_t
to_int(const std::string & s, const std::string & path)
{
  try
  {
    size_t pos = 0;
    const long long v = std::stoll(s, &pos);
    if (pos != s.size())
      throw std::invalid_argument(s);
    return v;
  }
  catch (const std::exception &)
  {
    throw Error("ReadMFEMMesh: expected an integer in " + path + ", got '" + s + "'");
  }
}

// A section count of the header: an integer in [0, 2^31).
int64_t
to_count(const std::string & s, const std::string & path, const char * what)
{
  const int64_t v = to_int(s, path);
  if (v < 0 || v >= (int64_t(1) << 31))
    throw Error(std::string("ReadMFEMMesh: ") + what + " count out of range in " + path + ": " + s);
  return v;
}

constexpr int CUBE = 5;

// corner offsets (dx, dy, dz) of the local vertices of an MFEM hexahedron
constexpr int CORNER[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

} // namespace

StructuredHexMesh
ReadMFEMMesh(const std::string & path)
{
  std::ifstream in(path);
  if (!in)
    throw Error("ReadMFEMMesh: cannot open " + path);
  std::string line;
  if (!next_line(in, line) || line != "MFEM mesh v1.0")
    throw Error("ReadMFEMMesh: " + path + " is not an 'MFEM mesh v1.0' file");
  int dim = 0;
  std::vector<int> attr;
  std::vector<std::array<int64_t, 8>> elems;
  std::vector<double> coords;
  int64_t nverts = -1;
  while (next_line(in, line))
  {
    if (line == "dimension")
    {
      next_line(in, line);
      dim = (int)to_int(line, path);
      if (dim != 3)
        throw Error("ReadMFEMMesh: only 3D meshes are supported");
    }
    else if (line == "elements")
    {
      next_line(in, line);
      const int64_t ne = to_count(line, path, "element");
      for (int64_t e = 0; e < ne; e++)
      {
        if (!next_line(in, line))
          throw Error("ReadMFEMMesh: truncated elements section");
        std::istringstream ss(line);
        int a, geom;
        ss >> a >> geom;
        if (geom != CUBE)
          throw Error("ReadMFEMMesh: only hexahedral elements are supported (found geometry " +
                      std::to_string(geom) + "; tetrahedral and mixed meshes need the unstructured path)");
        std::array<int64_t, 8> v;
        for (auto & x : v)
          ss >> x;
        if (!ss)
          throw Error("ReadMFEMMesh: bad element line '" + line + "'");
        attr.push_back(a);
        elems.push_back(v);
      }
    }
    else if (line == "boundary")
    {
      next_line(in, line);
      const int64_t nb = to_count(line, path, "boundary");
      for (int64_t b = 0; b < nb; b++)
        if (!next_line(in, line))
          throw Error("ReadMFEMMesh: truncated boundary section");
    }
    else if (line == "vertices")
    {
      next_line(in, line);
      nverts = to_count(line, path, "vertex");
      next_line(in, line);
      if (to_int(line, path) != 3)
        throw Error("ReadMFEMMesh: vertices must have 3 coordinates");
      coords.resize(3 * nverts);
      for (int64_t v = 0; v < nverts; v++)
      {
        if (!next_line(in, line))
          throw Error("ReadMFEMMesh: truncated vertices section");
        std::istringstream ss(line);
        ss >> coords[3 * v] >> coords[3 * v + 1] >> coords[3 * v + 2];
        if (!ss)
          throw Error("ReadMFEMMesh: bad vertex line '" + line + "'");
      }
    }
    else if (line == "mfem_mesh_end")
      break;
    else if (line.rfind("nodes", 0) == 0 || line.rfind("FiniteElementSpace", 0) == 0)
      throw Error("ReadMFEMMesh: curved (nodal) meshes are not supported");
  }
  if (dim != 3 || elems.empty() || nverts <= 0)
    throw Error("ReadMFEMMesh: missing dimension, elements or vertices in " + path);
  // nx: elements along the first row (each shares its +x face with the next)
  const int64_t E = (int64_t)elems.size();
  auto shares_x = [&](int64_t a, int64_t b) {
    return elems[a][1] == elems[b][0] && elems[a][2] == elems[b][3] && elems[a][5] == elems[b][4] &&
           elems[a][6] == elems[b][7];
  };
  auto shares_y = [&](int64_t a, int64_t b) {
    return elems[a][3] == elems[b][0] && elems[a][2] == elems[b][1] && elems[a][7] == elems[b][4] &&
           elems[a][6] == elems[b][5];
  };
  int64_t nx = 1;
  while (nx < E && shares_x(nx - 1, nx))
    nx++;
  int64_t ny = 1;
  while (ny * nx < E && shares_y((ny - 1) * nx, ny * nx))
    ny++;
  if (E % (nx * ny) != 0)
    throw Error("ReadMFEMMesh: the elements do not form a structured grid");
  const int64_t nz = E / (nx * ny);
  if ((nx + 1) * (ny + 1) * (nz + 1) != nverts)
    throw Error("ReadMFEMMesh: vertex count does not match a structured " + std::to_string(nx) + "x" +
                std::to_string(ny) + "x" + std::to_string(nz) + " grid");
  // structured vertex (i, j, k) -> file vertex id, consistent over every element
  std::vector<int64_t> map((size_t)nverts, -1);
  std::vector<char> used((size_t)nverts, 0);
  for (int64_t e = 0; e < E; e++)
  {
    const int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    for (int c = 0; c < 8; c++)
    {
      const int64_t s = ((ez + CORNER[c][2]) * (ny + 1) + (ey + CORNER[c][1])) * (nx + 1) + ex + CORNER[c][0];
      const int64_t v = elems[e][c];
      if (v < 0 || v >= nverts)
        throw Error("ReadMFEMMesh: vertex index out of range");
      if (map[s] < 0)
      {
        if (used[v])
          throw Error("ReadMFEMMesh: the elements do not form a structured grid (vertex reused)");
        map[s] = v;
        used[v] = 1;
      }
      else if (map[s] != v)
        throw Error("ReadMFEMMesh: the elements do not form a structured grid (element " + std::to_string(e) + ")");
    }
  }
  StructuredHexMesh m;
  m.nx = (int)nx;
  m.ny = (int)ny;
  m.nz = (int)nz;
  m.attributes = attr;
  m.vert.resize(3 * nverts);
  for (int64_t s = 0; s < nverts; s++)
    for (int r = 0; r < 3; r++)
      m.vert[3 * s + r] = coords[3 * map[s] + r];
  return m;
}

void
WriteMFEMMesh(const StructuredHexMesh & m, const std::string & path)
{
  std::ofstream f(path);
  if (!f)
    throw Error("WriteMFEMMesh: cannot open " + path);
  f.precision(17);
  const int64_t nx = m.nx, ny = m.ny, nz = m.nz, E = m.NumElements();
  auto vid = [&](int64_t i, int64_t j, int64_t k) { return (k * (ny + 1) + j) * (nx + 1) + i; };
  f << "MFEM mesh v1.0\n\ndimension\n3\n\nelements\n" << E << "\n";
  for (int64_t e = 0; e < E; e++)
  {
    const int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    f << (m.attributes.empty() ? 1 : m.attributes[e]) << " " << CUBE;
    for (int c = 0; c < 8; c++)
      f << " " << vid(ex + CORNER[c][0], ey + CORNER[c][1], ez + CORNER[c][2]);
    f << "\n";
  }
  // boundary quads, outward-facing, attribute = face index + 1
  std::ostringstream b;
  int64_t nb = 0;
  for (int64_t k = 0; k < nz; k++)
    for (int64_t j = 0; j < ny; j++)
    {
      b << "1 3 " << vid(0, j, k) << " " << vid(0, j, k + 1) << " " << vid(0, j + 1, k + 1) << " " << vid(0, j + 1, k)
        << "\n";
      b << "2 3 " << vid(nx, j, k) << " " << vid(nx, j + 1, k) << " " << vid(nx, j + 1, k + 1) << " "
        << vid(nx, j, k + 1) << "\n";
      nb += 2;
    }
  for (int64_t k = 0; k < nz; k++)
    for (int64_t i = 0; i < nx; i++)
    {
      b << "3 3 " << vid(i, 0, k) << " " << vid(i + 1, 0, k) << " " << vid(i + 1, 0, k + 1) << " " << vid(i, 0, k + 1)
        << "\n";
      b << "4 3 " << vid(i, ny, k) << " " << vid(i, ny, k + 1) << " " << vid(i + 1, ny, k + 1) << " "
        << vid(i + 1, ny, k) << "\n";
      nb += 2;
    }
  for (int64_t j = 0; j < ny; j++)
    for (int64_t i = 0; i < nx; i++)
    {
      b << "5 3 " << vid(i, j, 0) << " " << vid(i, j + 1, 0) << " " << vid(i + 1, j + 1, 0) << " " << vid(i + 1, j, 0)
        << "\n";
      b << "6 3 " << vid(i, j, nz) << " " << vid(i + 1, j, nz) << " " << vid(i + 1, j + 1, nz) << " "
        << vid(i, j + 1, nz) << "\n";
      nb += 2;
    }
  f << "\nboundary\n" << nb << "\n" << b.str();
  f << "\nvertices\n" << (nx + 1) * (ny + 1) * (nz + 1) << "\n3\n";
  for (int64_t k = 0; k <= nz; k++)
    for (int64_t j = 0; j <= ny; j++)
      for (int64_t i = 0; i <= nx; i++)
      {
        double x[3];
        if (!m.vert.empty())
          for (int r = 0; r < 3; r++)
            x[r] = m.vert[3 * vid(i, j, k) + r];
        else
        {
          const int64_t idx[3] = {i, j, k}, n[3] = {nx, ny, nz};
          for (int r = 0; r < 3; r++)
            x[r] = m.x0[r] + m.len[r] * (double)idx[r] / (double)n[r];
        }
        f << x[0] << " " << x[1] << " " << x[2] << "\n";
      }
  if (!f)
    throw Error("WriteMFEMMesh: write failed for " + path);
}

} // namespace hephaestus
