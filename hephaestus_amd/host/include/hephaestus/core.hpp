// Host-side framework of the H(curl) path: the names and call structure of
// hephaestus' InputParameters (src/io/inputs.hpp), NamedFieldsMap
// (src/named_fields_map/named_fields_map.hpp), coefficients
// (src/coefficients/coefficients.{hpp,cpp}), finite-element spaces, grid
// functions and boundary conditions, over device memory and libhcurl.so.
//
// Errors: the reference aborts (MFEM_ABORT); this layer throws
// hephaestus::Error carrying the message (and the libhcurl status when one
// caused it).
#pragma once

#include "../../../../include/hc_hcurl.h"

#include <any>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace hephaestus
{

class Error : public std::runtime_error
{
public:
  explicit Error(const std::string & msg, int code = 0) : std::runtime_error(msg), _code(code) {}
  int Code() const { return _code; }

private:
  int _code;
};

// Throws Error with hc_last_error() when rc != HC_OK.
void CheckHc(int rc, const std::string & what);

// ---------------------------------------------------------------------------
// InputParameters (src/io/inputs.hpp)
// ---------------------------------------------------------------------------
class InputParameters
{
public:
  InputParameters() = default;
  explicit InputParameters(std::map<std::string, std::any> params) : _params(std::move(params)) {}
  void SetParam(const std::string & name, std::any value) { _params[name] = std::move(value); }
  bool HasParam(const std::string & name) const { return _params.count(name) != 0; }
  template <typename T>
  T GetParam(const std::string & name) const
  {
    auto it = _params.find(name);
    if (it == _params.end())
      throw Error("required parameter '" + name + "' is missing");
    try
    {
      return std::any_cast<T>(it->second);
    }
    catch (const std::bad_any_cast & e)
    {
      throw Error("Exception raised when trying to cast required parameter '" + name + "': " + e.what());
    }
  }
  template <typename T>
  T GetOptionalParam(const std::string & name, T value) const
  {
    auto it = _params.find(name);
    if (it == _params.end())
      return value;
    try
    {
      return std::any_cast<T>(it->second);
    }
    catch (...)
    {
      return value;
    }
  }

private:
  std::map<std::string, std::any> _params;
};

// ---------------------------------------------------------------------------
// NamedFieldsMap (src/named_fields_map/named_fields_map.hpp): owns its
// entries through shared_ptr; Get returns a non-owning pointer.
// ---------------------------------------------------------------------------
template <typename T>
class NamedFieldsMap
{
public:
  void Register(const std::string & name, std::shared_ptr<T> field)
  {
    if (!field)
      throw Error("NamedFieldsMap: cannot register a null field '" + name + "'");
    _map[name] = std::move(field);
  }
  bool Has(const std::string & name) const { return _map.count(name) != 0; }
  T * Get(const std::string & name) const
  {
    auto it = _map.find(name);
    if (it == _map.end())
      throw Error("NamedFieldsMap: no field named '" + name + "'");
    return it->second.get();
  }
  std::shared_ptr<T> GetShared(const std::string & name) const
  {
    Get(name);
    return _map.at(name);
  }
  void Deregister(const std::string & name) { _map.erase(name); }
  size_t NumFields() const { return _map.size(); }
  std::vector<std::string> Names() const
  {
    std::vector<std::string> out;
    for (auto & kv : _map)
      out.push_back(kv.first);
    return out;
  }

private:
  std::map<std::string, std::shared_ptr<T>> _map;
};

// ---------------------------------------------------------------------------
// Mesh: structured hexahedral mesh (box or vertex array) with one attribute
// per element (the reference's Exodus/MFEM meshes carry element attributes
// that the piecewise coefficients key on).
// ---------------------------------------------------------------------------
struct StructuredHexMesh
{
  int nx = 1, ny = 1, nz = 1;
  double x0[3] = {0, 0, 0};
  double len[3] = {1, 1, 1};
  std::vector<double> vert;     // optional (nx+1)(ny+1)(nz+1)*3, x fastest
  std::vector<int> attributes;  // one per element (default 1)
  int64_t NumElements() const { return (int64_t)nx * ny * nz; }
  // element centroid (box or vertex average)
  void Centroid(int64_t e, double c[3]) const;
  // unit cube with attribute 2 on elements whose centroid lies in [lo,hi]^3, else 1
  static StructuredHexMesh TwoRegionCube(int n, double lo = 0.25, double hi = 0.75);
};

// ---------------------------------------------------------------------------
// Coefficients (src/coefficients/coefficients.{hpp,cpp})
// Scalar coefficients are piecewise constant per element here; Eval returns
// the value on one element.
// ---------------------------------------------------------------------------
class Coefficient
{
public:
  virtual ~Coefficient() = default;
  virtual double Eval(const StructuredHexMesh & mesh, int64_t e) const = 0;
  std::vector<double> ElementValues(const StructuredHexMesh & mesh) const;
};

class ConstantCoefficient : public Coefficient
{
public:
  explicit ConstantCoefficient(double c = 1.0) : constant(c) {}
  double Eval(const StructuredHexMesh &, int64_t) const override { return constant; }
  double constant;
};

// value by element attribute (mfem::PWConstCoefficient semantics: attribute a
// takes values[a-1])
class PWConstCoefficient : public Coefficient
{
public:
  explicit PWConstCoefficient(std::vector<double> values) : _values(std::move(values)) {}
  double Eval(const StructuredHexMesh & mesh, int64_t e) const override;

private:
  std::vector<double> _values;
};

// f(a(x), b(x)) of two coefficients (mfem::TransformedCoefficient); a, b not owned
class TransformedCoefficient : public Coefficient
{
public:
  TransformedCoefficient(const Coefficient * a, const Coefficient * b, std::function<double(double, double)> f)
    : _a(a), _b(b), _f(std::move(f))
  {
  }
  double Eval(const StructuredHexMesh & mesh, int64_t e) const override
  {
    return _f(_a->Eval(mesh, e), _b->Eval(mesh, e));
  }

private:
  const Coefficient * _a;
  const Coefficient * _b;
  std::function<double(double, double)> _f;
};

// prodFunc / fracFunc of src/coefficients/coefficients.cpp:8-17
double prodFunc(double a, double b);
double fracFunc(double a, double b);

class Coefficients
{
public:
  Coefficients();
  NamedFieldsMap<Coefficient> _scalars;
  double _t = 0.0; // time
};

// ---------------------------------------------------------------------------
// Finite-element spaces and grid functions
// ---------------------------------------------------------------------------
// Nedelec (H(curl)) space of order p on a structured hex mesh; name
// "ND_3D_P<p>" as in the reference's FE space registry.
class FESpace
{
public:
  FESpace(const StructuredHexMesh * mesh, const std::string & fec_name);
  FESpace(const StructuredHexMesh * mesh, int order, bool h1 = false) : _mesh(mesh), _order(order), _h1(h1) {}
  const StructuredHexMesh * Mesh() const { return _mesh; }
  int Order() const { return _order; }
  // H1_3D_P<p> (continuous Q_p on the Gauss-Lobatto points) rather than ND_3D_P<p>
  bool IsH1() const { return _h1; }
  int64_t GetTrueVSize() const;
  int64_t GetVSize() const { return GetTrueVSize(); }

private:
  const StructuredHexMesh * _mesh;
  int _order;
  bool _h1 = false;
};

// Device vector (hipMalloc'd) of doubles.
class DeviceVector
{
public:
  DeviceVector() = default;
  explicit DeviceVector(int64_t n) { SetSize(n); }
  ~DeviceVector();
  DeviceVector(const DeviceVector &) = delete;
  DeviceVector & operator=(const DeviceVector &) = delete;
  void SetSize(int64_t n);
  int64_t Size() const { return _n; }
  double * Data() { return _d; }
  const double * Data() const { return _d; }
  void Fill(double v);
  void CopyFrom(const DeviceVector & o);
  void Upload(const std::vector<double> & h);
  std::vector<double> Download() const;
  // this = a * x + b * this
  void Axpby(double a, const DeviceVector & x, double b);

private:
  double * _d = nullptr;
  int64_t _n = 0;
};

class GridFunction : public DeviceVector
{
public:
  explicit GridFunction(FESpace * fes) : DeviceVector(fes->GetTrueVSize()), _fes(fes) { Fill(0.0); }
  FESpace * ParFESpace() const { return _fes; }

private:
  FESpace * _fes;
};

using FESpaces = NamedFieldsMap<FESpace>;
using GridFunctions = NamedFieldsMap<GridFunction>;

// ---------------------------------------------------------------------------
// Boundary conditions (src/boundary_conditions): PEC = zero tangential
// Dirichlet on the listed boundary faces (VectorDirichletBC with TANGENTIAL
// application, vector_dirichlet_bc.cpp:23-45, value zero).
// ---------------------------------------------------------------------------
class BoundaryCondition
{
public:
  BoundaryCondition(std::string name, int faces) : _name(std::move(name)), _faces(faces) {}
  virtual ~BoundaryCondition() = default;
  int Faces() const { return _faces; }
  const std::string & Name() const { return _name; }

private:
  std::string _name;
  int _faces; // HC_FACE_* bits
};

class VectorDirichletBC : public BoundaryCondition
{
public:
  using BoundaryCondition::BoundaryCondition;
};

class BCMap : public NamedFieldsMap<BoundaryCondition>
{
public:
  // union of the essential faces of every registered condition
  int EssentialFaces() const;
};

// ---------------------------------------------------------------------------
// Sources: a linear-form contribution (src/sources).  CoilSource adds the
// coil-like current of hc_op_rhs_coil (a divergence-free J = curl psi; the
// reference's DivFreeSource projects its J onto divergence-free fields, which
// this J already is).
// ---------------------------------------------------------------------------
class LinearForm;
class Source
{
public:
  virtual ~Source() = default;
  virtual void Apply(LinearForm * lf) = 0;
};

class CoilSource : public Source
{
public:
  explicit CoilSource(uint64_t seed, double scale = 1.0) : _seed(seed), _scale(scale) {}
  void Apply(LinearForm * lf) override;

private:
  uint64_t _seed;
  double _scale;
};

// DivFreeSource (src/sources/div_free_source.cpp:71-120) for the coil current
// J of hc_op_rhs_coil: g = M^-1 (J, phi_i) with the unit-coefficient H(curl)
// mass matrix, optionally the Helmholtz projection g - grad Q (no boundary
// conditions), then lf += M g; SubtractSource removes M g again.
class DeviceVector;
class DivFreeSource : public Source
{
public:
  explicit DivFreeSource(uint64_t seed, bool perform_helmholtz_projection = true, double tol = 1e-12,
                         int maxit = 20000);
  ~DivFreeSource() override;
  void Apply(LinearForm * lf) override;
  void SubtractSource(DeviceVector * gf) const;
  // g of the last Apply (the divergence-free source field)
  const DeviceVector & SourceField() const;

private:
  uint64_t _seed;
  bool _project;
  double _tol;
  int _maxit;
  DeviceVector * _g = nullptr;  // g
  DeviceVector * _mg = nullptr; // M g
};

using Sources = NamedFieldsMap<Source>;

} // namespace hephaestus
