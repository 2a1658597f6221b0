// Partial-assembly forms on libhcurl.so and the kernels that fill them.
#include "hephaestus/forms.hpp"

namespace hephaestus
{

BilinearForm::~BilinearForm()
{
  if (_op)
    hc_op_destroy(_op);
}

void
BilinearForm::Assemble()
{
  if (_fes->IsH1())
    return; // recorded only (see AddDomainIntegrator(DiffusionIntegrator))
  const StructuredHexMesh & mesh = *_fes->Mesh();
  const int64_t E = mesh.NumElements();
  std::vector<double> nu(E, 0.0), sig(E, 0.0);
  for (const Coefficient * c : _curl)
  {
    const auto v = c->ElementValues(mesh);
    for (int64_t e = 0; e < E; e++)
      nu[e] += v[e];
  }
  for (const Coefficient * c : _mass)
  {
    const auto v = c->ElementValues(mesh);
    for (int64_t e = 0; e < E; e++)
      sig[e] += v[e];
  }
  _last_work = 0;
  if (!_op)
  {
    hc_mesh_desc d{};
    d.nx = mesh.nx;
    d.ny = mesh.ny;
    d.nz = mesh.nz;
    d.vert = mesh.vert.empty() ? nullptr : mesh.vert.data();
    for (int r = 0; r < 3; r++)
    {
      d.x0[r] = mesh.x0[r];
      d.len[r] = mesh.len[r];
    }
    d.pec_faces = _ess_faces;
    CheckHc(hc_op_create(&d, _fes->Order(), nu.data(), sig.data(), _prec, &_op), "BilinearForm::Assemble");
    _last_work = 1;
  }
  else if (nu != _nu || sig != _sig)
  {
    CheckHc(hc_op_set_coefficients(_op, nu.data(), sig.data()), "BilinearForm::Assemble");
    _last_work = 1;
  }
  // the coefficients carry any dt factor: setup at unit time-step scale
  CheckHc(hc_op_setup(_op, 1.0, nullptr), "BilinearForm::Assemble");
  _nu.swap(nu);
  _sig.swap(sig);
}

void
BilinearForm::Mult(const DeviceVector & x, DeviceVector & y) const
{
  if (!_op)
    throw Error("BilinearForm::Mult: form not assembled");
  if (x.Size() != _fes->GetTrueVSize() || y.Size() != x.Size())
    throw Error("BilinearForm::Mult: vector size mismatch");
  CheckHc(hc_op_apply(_op, x.Data(), y.Data(), 1, nullptr), "BilinearForm::Mult");
}

void
BilinearForm::AddMultTranspose(const DeviceVector & u, DeviceVector & lf, double a) const
{
  _tmp.SetSize(u.Size());
  Mult(u, _tmp);
  lf.Axpby(a, _tmp, 1.0);
}

// ---------------------------------------------------------------------------
CurlCurlKernel::CurlCurlKernel(const InputParameters & params)
  : Kernel(params), _coef_name(params.GetParam<std::string>("CoefficientName"))
{
}

void
CurlCurlKernel::Init(GridFunctions &, const FESpaces &, BCMap &, Coefficients & coefficients)
{
  _coef = coefficients._scalars.Get(_coef_name);
}

void
CurlCurlKernel::Apply(BilinearForm * blf)
{
  blf->AddDomainIntegrator(CurlCurlIntegrator{_coef});
}

DiffusionKernel::DiffusionKernel(const InputParameters & params)
  : Kernel(params), _coef_name(params.GetParam<std::string>("CoefficientName"))
{
}

void
DiffusionKernel::Init(GridFunctions &, const FESpaces &, BCMap &, Coefficients & coefficients)
{
  _coef = coefficients._scalars.Get(_coef_name);
}

void
DiffusionKernel::Apply(BilinearForm * blf)
{
  blf->AddDomainIntegrator(DiffusionIntegrator{_coef});
}

MixedVectorGradientKernel::MixedVectorGradientKernel(const InputParameters & params)
  : Kernel(params), _coef_name(params.GetParam<std::string>("CoefficientName"))
{
}

void
MixedVectorGradientKernel::Init(GridFunctions &, const FESpaces &, BCMap &, Coefficients & coefficients)
{
  _coef = coefficients._scalars.Get(_coef_name);
}

void
MixedVectorGradientKernel::Apply(MixedBilinearForm * mblf)
{
  mblf->AddDomainIntegrator(MixedVectorGradientIntegrator{_coef});
}

VectorFEWeakDivergenceKernel::VectorFEWeakDivergenceKernel(const InputParameters & params)
  : Kernel(params), _coef_name(params.GetParam<std::string>("CoefficientName"))
{
}

void
VectorFEWeakDivergenceKernel::Init(GridFunctions &, const FESpaces &, BCMap &, Coefficients & coefficients)
{
  _coef = coefficients._scalars.Get(_coef_name);
}

void
VectorFEWeakDivergenceKernel::Apply(MixedBilinearForm * mblf)
{
  mblf->AddDomainIntegrator(VectorFEWeakDivergenceIntegrator{_coef});
}

VectorFEMassKernel::VectorFEMassKernel(const InputParameters & params)
  : Kernel(params), _coef_name(params.GetParam<std::string>("CoefficientName"))
{
}

void
VectorFEMassKernel::Init(GridFunctions &, const FESpaces &, BCMap &, Coefficients & coefficients)
{
  _coef = coefficients._scalars.Get(_coef_name);
}

void
VectorFEMassKernel::Apply(BilinearForm * blf)
{
  blf->AddDomainIntegrator(VectorFEMassIntegrator{_coef});
}

WeakCurlCurlKernel::WeakCurlCurlKernel(const InputParameters & params)
  : Kernel(params),
    _coupled_gf_name(params.GetParam<std::string>("CoupledVariableName")),
    _coef_name(params.GetParam<std::string>("CoefficientName"))
{
}

void
WeakCurlCurlKernel::Init(GridFunctions & gridfunctions, const FESpaces &, BCMap & bc_map,
                         Coefficients & coefficients)
{
  _u = gridfunctions.Get(_coupled_gf_name);
  _coef = coefficients._scalars.Get(_coef_name);
  _curl_curl = std::make_unique<BilinearForm>(_u->ParFESpace());
  _curl_curl->SetEssentialFaces(bc_map.EssentialFaces());
  _curl_curl->SetPrecision(HC_PREC_FP64);
  _curl_curl->AddDomainIntegrator(CurlCurlIntegrator{_coef});
  _curl_curl->Assemble();
}

void
WeakCurlCurlKernel::Apply(LinearForm * lf)
{
  _curl_curl->AddMultTranspose(*_u, *lf, -1.0);
}

} // namespace hephaestus
