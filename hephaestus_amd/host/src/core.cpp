// Host framework pieces: coefficients, spaces, device vectors, boundary
// conditions, sources.
#include "hephaestus/core.hpp"

#include "hephaestus/forms.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdlib>

namespace hephaestus
{

void
CheckHc(int rc, const std::string & what)
{
  if (rc != HC_OK)
    throw Error(what + ": " + hc_last_error(), rc);
}

static void
check_hip(hipError_t e, const std::string & what)
{
  if (e != hipSuccess)
    throw Error(what + ": " + hipGetErrorString(e), HC_EHIP);
}

// ---------------------------------------------------------------------------
void
StructuredHexMesh::Centroid(int64_t e, double c[3]) const
{
  const int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / ((int64_t)nx * ny);
  if (vert.empty())
  {
    c[0] = x0[0] + (ex + 0.5) * len[0] / nx;
    c[1] = x0[1] + (ey + 0.5) * len[1] / ny;
    c[2] = x0[2] + (ez + 0.5) * len[2] / nz;
    return;
  }
  c[0] = c[1] = c[2] = 0.0;
  const int64_t sy = nx + 1, sz = (int64_t)(nx + 1) * (ny + 1);
  for (int k = 0; k < 8; k++)
  {
    const int64_t v = (ex + (k & 1)) + (ey + ((k >> 1) & 1)) * sy + (ez + (k >> 2)) * sz;
    for (int r = 0; r < 3; r++)
      c[r] += 0.125 * vert[3 * v + r];
  }
}

StructuredHexMesh
StructuredHexMesh::TwoRegionCube(int n, double lo, double hi)
{
  StructuredHexMesh m;
  m.nx = m.ny = m.nz = n;
  m.attributes.assign(m.NumElements(), 1);
  for (int64_t e = 0; e < m.NumElements(); e++)
  {
    double c[3];
    m.Centroid(e, c);
    bool in = true;
    for (int r = 0; r < 3; r++)
      in = in && c[r] >= lo && c[r] <= hi;
    if (in)
      m.attributes[e] = 2;
  }
  return m;
}

std::vector<double>
Coefficient::ElementValues(const StructuredHexMesh & mesh) const
{
  std::vector<double> v(mesh.NumElements());
  for (int64_t e = 0; e < mesh.NumElements(); e++)
    v[e] = Eval(mesh, e);
  return v;
}

double
PWConstCoefficient::Eval(const StructuredHexMesh & mesh, int64_t e) const
{
  const int a = mesh.attributes.empty() ? 1 : mesh.attributes[e];
  if (a < 1 || a > (int)_values.size())
    throw Error("PWConstCoefficient: no value for attribute " + std::to_string(a));
  return _values[a - 1];
}

double
prodFunc(double a, double b)
{
  return a * b;
}

double
fracFunc(double a, double b)
{
  return a / b;
}

Coefficients::Coefficients()
{
  // the reference registers these defaults (coefficients.cpp)
  _scalars.Register("_one", std::make_shared<ConstantCoefficient>(1.0));
}

// ---------------------------------------------------------------------------
FESpace::FESpace(const StructuredHexMesh * mesh, const std::string & fec_name) : _mesh(mesh)
{
  // "ND_3D_P<p>" or "H1_3D_P<p>"
  const std::string nd = "ND_3D_P", h1 = "H1_3D_P";
  if (fec_name.compare(0, nd.size(), nd) == 0)
    _h1 = false;
  else if (fec_name.compare(0, h1.size(), h1) == 0)
    _h1 = true;
  else
    throw Error("FESpace: only ND_3D_P<p> and H1_3D_P<p> spaces are supported, got '" + fec_name + "'");
  _order = std::atoi(fec_name.c_str() + nd.size());
  if (_order < 1 || _order > 3)
    throw Error("FESpace: order must be 1..3, got '" + fec_name + "'");
}

int64_t
FESpace::GetTrueVSize() const
{
  const int64_t p = _order, nx = _mesh->nx, ny = _mesh->ny, nz = _mesh->nz;
  if (_h1)
    return (nx * p + 1) * (ny * p + 1) * (nz * p + 1);
  return nx * p * (ny * p + 1) * (nz * p + 1) + (nx * p + 1) * ny * p * (nz * p + 1) +
         (nx * p + 1) * (ny * p + 1) * nz * p;
}

// ---------------------------------------------------------------------------
DeviceVector::~DeviceVector()
{
  if (_d)
    (void)hipFree(_d);
}

void
DeviceVector::SetSize(int64_t n)
{
  if (n == _n && _d)
    return;
  if (_d)
    (void)hipFree(_d);
  _d = nullptr;
  _n = 0;
  if (n > 0)
    check_hip(hipMalloc((void **)&_d, sizeof(double) * n), "DeviceVector allocation");
  _n = n;
}

void
DeviceVector::Fill(double v)
{
  if (_n == 0)
    return;
  if (v == 0.0)
  {
    CheckHc(hc_axpby(_n, 0.0, _d, 0.0, _d, nullptr), "DeviceVector::Fill");
    return;
  }
  std::vector<double> h(_n, v);
  Upload(h);
}

void
DeviceVector::CopyFrom(const DeviceVector & o)
{
  SetSize(o._n);
  if (_n)
    check_hip(hipMemcpy(_d, o._d, sizeof(double) * _n, hipMemcpyDeviceToDevice), "DeviceVector copy");
}

void
DeviceVector::Upload(const std::vector<double> & h)
{
  SetSize((int64_t)h.size());
  if (_n)
    check_hip(hipMemcpy(_d, h.data(), sizeof(double) * _n, hipMemcpyHostToDevice), "DeviceVector upload");
}

std::vector<double>
DeviceVector::Download() const
{
  std::vector<double> h(_n);
  if (_n)
    check_hip(hipMemcpy(h.data(), _d, sizeof(double) * _n, hipMemcpyDeviceToHost), "DeviceVector download");
  return h;
}

void
DeviceVector::Axpby(double a, const DeviceVector & x, double b)
{
  if (x._n != _n)
    throw Error("DeviceVector::Axpby: size mismatch");
  CheckHc(hc_axpby(_n, a, x._d, b, _d, nullptr), "DeviceVector::Axpby");
}

int
BCMap::EssentialFaces() const
{
  int f = 0;
  for (auto & n : Names())
    if (dynamic_cast<VectorDirichletBC *>(Get(n)))
      f |= Get(n)->Faces();
  return f;
}

void
CoilSource::Apply(LinearForm * lf)
{
  if (!lf->Operator())
    throw Error("CoilSource: the linear form has no operator (build the bilinear forms first)");
  DeviceVector b(lf->Size());
  CheckHc(hc_op_rhs_coil(lf->Operator(), _seed, b.Data(), nullptr), "CoilSource");
  lf->Axpby(_scale, b, 1.0);
}

DivFreeSource::DivFreeSource(uint64_t seed, bool perform_helmholtz_projection, double tol, int maxit)
  : _seed(seed), _project(perform_helmholtz_projection), _tol(tol), _maxit(maxit)
{
}

DivFreeSource::~DivFreeSource()
{
  delete _g;
  delete _mg;
}

void
DivFreeSource::Apply(LinearForm * lf)
{
  if (!lf->Operator())
    throw Error("DivFreeSource: the linear form has no operator (build the bilinear forms first)");
  const int64_t n = lf->Size();
  DeviceVector j(n);
  CheckHc(hc_op_rhs_coil(lf->Operator(), _seed, j.Data(), nullptr), "DivFreeSource");
  if (!_g)
    _g = new DeviceVector(n);
  if (!_mg)
    _mg = new DeviceVector(n);
  _g->SetSize(n);
  _mg->SetSize(n);
  _mg->Fill(0.0);
  hc_solve_stats st{};
  CheckHc(hc_op_div_free_source(lf->Operator(), j.Data(), _g->Data(), _mg->Data(), nullptr, _project ? 1 : 0, _tol,
                                _maxit, &st, nullptr),
          "DivFreeSource");
  lf->Axpby(1.0, *_mg, 1.0);
}

void
DivFreeSource::SubtractSource(DeviceVector * gf) const
{
  if (!_mg)
    throw Error("DivFreeSource::SubtractSource before Apply");
  gf->Axpby(-1.0, *_mg, 1.0);
}

const DeviceVector &
DivFreeSource::SourceField() const
{
  if (!_g)
    throw Error("DivFreeSource::SourceField before Apply");
  return *_g;
}

} // namespace hephaestus
