// Equation systems, Jacobi-PCG solver and the backward-Euler step operator.
#include "hephaestus/equation_system.hpp"

#include <algorithm>
#include <cmath>
#include <map>

namespace hephaestus
{

bool
EquationSystem::VectorContainsName(const std::vector<std::string> & v, const std::string & name) const
{
  return std::find(v.begin(), v.end(), name) != v.end();
}

void
EquationSystem::AddTestVariableNameIfMissing(const std::string & name)
{
  if (!VectorContainsName(_test_var_names, name))
    _test_var_names.push_back(name);
}

void
EquationSystem::AddTrialVariableNameIfMissing(const std::string & name)
{
  if (!VectorContainsName(_trial_var_names, name))
    _trial_var_names.push_back(name);
}

void
EquationSystem::AddKernel(const std::string & test_var_name, std::shared_ptr<ParBilinearFormKernel> blf_kernel)
{
  AddTestVariableNameIfMissing(test_var_name);
  if (!_blf_kernels_map.Has(test_var_name))
    _blf_kernels_map.Register(test_var_name, std::make_shared<std::vector<std::shared_ptr<ParBilinearFormKernel>>>());
  _blf_kernels_map.Get(test_var_name)->push_back(std::move(blf_kernel));
}

void
EquationSystem::AddKernel(const std::string & test_var_name, std::shared_ptr<ParLinearFormKernel> lf_kernel)
{
  AddTestVariableNameIfMissing(test_var_name);
  if (!_lf_kernels_map.Has(test_var_name))
    _lf_kernels_map.Register(test_var_name, std::make_shared<std::vector<std::shared_ptr<ParLinearFormKernel>>>());
  _lf_kernels_map.Get(test_var_name)->push_back(std::move(lf_kernel));
}

void
EquationSystem::ApplyBoundaryConditions(BCMap & bc_map)
{
  // PEC conditions enter the operator (eliminated rows/columns) and
  // FormLinearSystem (B_ess = 0)
  _ess_faces = bc_map.EssentialFaces();
}

void
EquationSystem::Init(GridFunctions & gridfunctions, const FESpaces & fespaces, BCMap & bc_map,
                     Coefficients & coefficients)
{
  // trial variables must be registered gridfunctions; test spaces follow them
  for (auto & test_var_name : _test_var_names)
  {
    if (!gridfunctions.Has(test_var_name))
      throw Error("EquationSystem::Init: test variable '" + test_var_name + "' has no gridfunction");
    _test_pfespaces.push_back(gridfunctions.Get(test_var_name)->ParFESpace());
  }
  _ess_faces = bc_map.EssentialFaces();
  for (auto & test_var_name : _test_var_names)
  {
    if (_blf_kernels_map.Has(test_var_name))
      for (auto & k : *_blf_kernels_map.Get(test_var_name))
        k->Init(gridfunctions, fespaces, bc_map, coefficients);
    if (_lf_kernels_map.Has(test_var_name))
      for (auto & k : *_lf_kernels_map.Get(test_var_name))
        k->Init(gridfunctions, fespaces, bc_map, coefficients);
  }
}

void
EquationSystem::BuildBilinearForms()
{
  for (size_t i = 0; i < _test_var_names.size(); i++)
  {
    const auto & name = _test_var_names[i];
    // the form (and its device factors) persists across rebuilds; kernels
    // re-add their integrators and Assemble recomputes only on changes
    if (!_blfs.Has(name))
      _blfs.Register(name, std::make_shared<BilinearForm>(_test_pfespaces[i]));
    BilinearForm * blf = _blfs.Get(name);
    blf->ClearIntegrators();
    blf->SetEssentialFaces(_ess_faces);
    if (_blf_kernels_map.Has(name))
      for (auto & k : *_blf_kernels_map.Get(name))
        k->Apply(blf);
    blf->Assemble();
  }
}

void
EquationSystem::BuildLinearForms(BCMap & bc_map, Sources & sources)
{
  ApplyBoundaryConditions(bc_map);
  for (size_t i = 0; i < _test_var_names.size(); i++)
  {
    const auto & name = _test_var_names[i];
    if (!_lfs.Has(name))
      _lfs.Register(name, std::make_shared<LinearForm>(_test_pfespaces[i]));
    LinearForm * lf = _lfs.Get(name);
    lf->Fill(0.0);
    if (_blfs.Has(name))
      lf->SetOperator(_blfs.Get(name)->Handle());
    if (_lf_kernels_map.Has(name))
      for (auto & k : *_lf_kernels_map.Get(name))
        k->Apply(lf);
    if (i == 0)
      for (auto & s : sources.Names())
        sources.Get(s)->Apply(lf);
  }
}

void
EquationSystem::BuildEquationSystem(BCMap & bc_map, Sources & sources)
{
  // bilinear forms first: device sources evaluate on the form's operator
  ApplyBoundaryConditions(bc_map);
  BuildBilinearForms();
  BuildLinearForms(bc_map, sources);
}

void
EquationSystem::FormLinearSystem(DeviceVector & X, DeviceVector & B)
{
  if (_test_var_names.size() != 1)
    throw Error("FormLinearSystem: one test variable (the H(curl) system) is supported");
  const auto & name = _test_var_names[0];
  BilinearForm * blf = _blfs.Get(name);
  LinearForm * lf = _lfs.Get(name);
  B.CopyFrom(*lf);
  X.SetSize(B.Size());
  X.Fill(0.0);
  // zero tangential boundary values: B_ess = X_ess = 0, B_free unchanged
  CheckHc(hc_op_zero_ess(blf->Handle(), B.Data(), nullptr), "FormLinearSystem");
}

void
EquationSystem::Mult(const DeviceVector & x, DeviceVector & y) const
{
  GetGradient().Mult(x, y);
}

BilinearForm &
EquationSystem::GetGradient() const
{
  if (_test_var_names.empty())
    throw Error("EquationSystem: no test variable");
  return *_blfs.Get(_test_var_names[0]);
}

void
EquationSystem::RecoverFEMSolution(const DeviceVector & X, GridFunctions & gridfunctions)
{
  gridfunctions.Get(_test_var_names[0])->CopyFrom(X);
}

// ---------------------------------------------------------------------------
TimeDependentEquationSystem::TimeDependentEquationSystem() : _dt_coef(1.0) {}

void
TimeDependentEquationSystem::AddTrialVariableNameIfMissing(const std::string & name)
{
  EquationSystem::AddTrialVariableNameIfMissing(name);
  const std::string d = GetTimeDerivativeName(name);
  if (!VectorContainsName(_trial_var_time_derivative_names, d))
    _trial_var_time_derivative_names.push_back(d);
}

void
TimeDependentEquationSystem::SetTimeStep(double dt)
{
  if (std::fabs(dt - _dt_coef.constant) > 1.0e-12 * dt)
  {
    _dt_coef.constant = dt;
    for (auto & name : _test_var_names)
      if (_blfs.Has(name))
        _blfs.Get(name)->Assemble();
  }
}

void
TimeDependentEquationSystem::UpdateEquationSystem(BCMap & bc_map, Sources & sources)
{
  BuildEquationSystem(bc_map, sources);
}

// ---------------------------------------------------------------------------
CurlCurlEquationSystem::CurlCurlEquationSystem(const InputParameters & params)
  : _h_curl_var_name(params.GetParam<std::string>("HCurlVarName")),
    _alpha_coef_name(params.GetParam<std::string>("AlphaCoefName")),
    _beta_coef_name(params.GetParam<std::string>("BetaCoefName")),
    _dtalpha_coef_name("dt_" + _alpha_coef_name)
{
}

void
CurlCurlEquationSystem::Init(GridFunctions & gridfunctions, const FESpaces & fespaces, BCMap & bc_map,
                             Coefficients & coefficients)
{
  coefficients._scalars.Register(
      _dtalpha_coef_name,
      std::make_shared<TransformedCoefficient>(&_dt_coef, coefficients._scalars.Get(_alpha_coef_name), prodFunc));
  TimeDependentEquationSystem::Init(gridfunctions, fespaces, bc_map, coefficients);
}

void
CurlCurlEquationSystem::AddKernels()
{
  AddTrialVariableNameIfMissing(_h_curl_var_name);
  const std::string du_dt = GetTimeDerivativeName(_h_curl_var_name);
  // (alpha curl u_n, curl u')
  InputParameters weak;
  weak.SetParam("CoupledVariableName", _h_curl_var_name);
  weak.SetParam("CoefficientName", _alpha_coef_name);
  AddKernel(du_dt, std::shared_ptr<ParLinearFormKernel>(std::make_shared<WeakCurlCurlKernel>(weak)));
  // (alpha dt curl du/dt, curl u')
  InputParameters cc;
  cc.SetParam("CoefficientName", _dtalpha_coef_name);
  AddKernel(du_dt, std::shared_ptr<ParBilinearFormKernel>(std::make_shared<CurlCurlKernel>(cc)));
  // (beta du/dt, u')
  InputParameters mass;
  mass.SetParam("CoefficientName", _beta_coef_name);
  AddKernel(du_dt, std::shared_ptr<ParBilinearFormKernel>(std::make_shared<VectorFEMassKernel>(mass)));
}

// ---------------------------------------------------------------------------
AVEquationSystem::AVEquationSystem(const InputParameters & params)
  : _a_name(params.GetParam<std::string>("VectorPotentialName")),
    _v_name(params.GetParam<std::string>("ScalarPotentialName")),
    _alpha_coef_name(params.GetParam<std::string>("AlphaCoefName")),
    _beta_coef_name(params.GetParam<std::string>("BetaCoefName")),
    _dtalpha_coef_name("dt_" + _alpha_coef_name),
    _neg_beta_coef_name("negative_" + _beta_coef_name)
{
}

void
AVEquationSystem::Init(GridFunctions & gridfunctions, const FESpaces & fespaces, BCMap & bc_map,
                       Coefficients & coefficients)
{
  coefficients._scalars.Register(
      _dtalpha_coef_name,
      std::make_shared<TransformedCoefficient>(&_dt_coef, coefficients._scalars.Get(_alpha_coef_name), prodFunc));
  coefficients._scalars.Register(
      _neg_beta_coef_name,
      std::make_shared<TransformedCoefficient>(&_neg_coef, coefficients._scalars.Get(_beta_coef_name), prodFunc));
  TimeDependentEquationSystem::Init(gridfunctions, fespaces, bc_map, coefficients);
  for (auto & kv : _mblf_kernels)
    for (auto & k : kv.second)
      k->Init(gridfunctions, fespaces, bc_map, coefficients);
  _gfs = &gridfunctions;
  _fes = &fespaces;
  _bcs = &bc_map;
  _coefs = &coefficients;
}

void
AVEquationSystem::AddKernel(const std::string & trial_var_name, const std::string & test_var_name,
                            std::shared_ptr<MixedBilinearFormKernel> kernel)
{
  AddTestVariableNameIfMissing(test_var_name);
  _mblf_kernels[{trial_var_name, test_var_name}].push_back(std::move(kernel));
}

void
AVEquationSystem::AddKernels()
{
  AddTrialVariableNameIfMissing(_a_name);
  const std::string da_dt = GetTimeDerivativeName(_a_name);
  AddTrialVariableNameIfMissing(_v_name);
  // (alpha curl A_n, curl dA'/dt)
  InputParameters weak;
  weak.SetParam("CoupledVariableName", _a_name);
  weak.SetParam("CoefficientName", _alpha_coef_name);
  EquationSystem::AddKernel(da_dt, std::shared_ptr<ParLinearFormKernel>(std::make_shared<WeakCurlCurlKernel>(weak)));
  // (alpha dt curl dA/dt, curl dA'/dt)
  InputParameters cc;
  cc.SetParam("CoefficientName", _dtalpha_coef_name);
  EquationSystem::AddKernel(da_dt, std::shared_ptr<ParBilinearFormKernel>(std::make_shared<CurlCurlKernel>(cc)));
  // (beta dA/dt, dA'/dt)
  InputParameters mass;
  mass.SetParam("CoefficientName", _beta_coef_name);
  EquationSystem::AddKernel(da_dt, std::shared_ptr<ParBilinearFormKernel>(std::make_shared<VectorFEMassKernel>(mass)));
  // (sigma grad V, dA'/dt)
  InputParameters grad;
  grad.SetParam("CoefficientName", _beta_coef_name);
  AddKernel(_v_name, da_dt, std::make_shared<MixedVectorGradientKernel>(grad));
  // (sigma grad V, grad V')
  InputParameters diff;
  diff.SetParam("CoefficientName", _beta_coef_name);
  EquationSystem::AddKernel(_v_name, std::shared_ptr<ParBilinearFormKernel>(std::make_shared<DiffusionKernel>(diff)));
  // (sigma dA/dt, grad V')
  InputParameters wdiv;
  wdiv.SetParam("CoefficientName", _beta_coef_name);
  AddKernel(da_dt, _v_name, std::make_shared<VectorFEWeakDivergenceKernel>(wdiv));
}

void
AVEquationSystem::BuildBilinearForms()
{
  EquationSystem::BuildBilinearForms();
  for (auto & kv : _mblf_kernels)
  {
    auto & f = _mblfs[kv.first];
    if (!f)
      f = std::make_unique<MixedBilinearForm>(_gfs->Get(kv.first.first)->ParFESpace(),
                                              _gfs->Get(kv.first.second)->ParFESpace());
    f->ClearIntegrators();
    for (auto & k : kv.second)
      k->Apply(f.get());
  }
  // the block operator applies every block through the A-A operator's mass
  // factors: the mixed and diffusion coefficients must equal its sigma
  const std::string da_dt = GetTimeDerivativeName(_a_name);
  const StructuredHexMesh & mesh = *_gfs->Get(da_dt)->ParFESpace()->Mesh();
  auto sum = [&](const std::vector<const Coefficient *> & cs) {
    std::vector<double> v(mesh.NumElements(), 0.0);
    for (const Coefficient * c : cs)
    {
      const auto e = c->ElementValues(mesh);
      for (size_t i = 0; i < v.size(); i++)
        v[i] += e[i];
    }
    return v;
  };
  const std::vector<double> sig = sum(_blfs.Get(da_dt)->MassCoefficients());
  const std::vector<double> sdiff = sum(_blfs.Get(_v_name)->DiffusionCoefficients());
  const std::vector<double> sgrad = sum(_mblfs.at({_v_name, da_dt})->GradientCoefficients());
  const std::vector<double> swdiv = sum(_mblfs.at({da_dt, _v_name})->WeakDivergenceCoefficients());
  if (sdiff != sig || sgrad != sig || swdiv != sig)
    throw Error("AVEquationSystem: the mass, mixed-gradient, weak-divergence and diffusion coefficients must be "
                "the same sigma");
}

hc_solve_stats
AVEquationSystem::Solve(DeviceVector & dA_dt, DeviceVector & V, double tol, int maxit)
{
  const std::string da_dt = GetTimeDerivativeName(_a_name);
  hc_op_t op = _blfs.Get(da_dt)->Handle();
  LinearForm * la = _lfs.Get(da_dt);
  LinearForm * lv = _lfs.Get(_v_name);
  // B: the linear forms with the PEC rows of the A part set to the (zero)
  // tangential values, as FormLinearSystem does
  DeviceVector Ba(la->Size()), Bv(lv->Size());
  Ba.CopyFrom(*la);
  Bv.CopyFrom(*lv);
  CheckHc(hc_op_zero_ess(op, Ba.Data(), nullptr), "AVEquationSystem::Solve");
  hc_solve_stats st{};
  CheckHc(hc_op_av_solve(op, Ba.Data(), Bv.Data(), dA_dt.Data(), V.Data(), tol, maxit, &st, nullptr),
          "AVEquationSystem::Solve");
  return st;
}

hc_solve_stats
AVEquationSystem::Step(double dt, GridFunctions & gridfunctions, BCMap & bc_map, Sources & sources, double tol,
                       int maxit)
{
  const std::string da_dt = GetTimeDerivativeName(_a_name);
  SetTimeStep(dt);
  UpdateEquationSystem(bc_map, sources);
  GridFunction * a = gridfunctions.Get(_a_name);
  GridFunction * k = gridfunctions.Get(da_dt);
  GridFunction * v = gridfunctions.Get(_v_name);
  k->Fill(0.0);
  v->Fill(0.0);
  const hc_solve_stats st = Solve(*k, *v, tol, maxit);
  a->Axpby(dt, *k, 1.0);
  return st;
}

// ---------------------------------------------------------------------------
DefaultJacobiPCGSolver::DefaultJacobiPCGSolver(const InputParameters & params, BilinearForm & op)
  : _op(op),
    _tol(params.GetOptionalParam<double>("Tolerance", 1e-9)),
    _maxit(params.GetOptionalParam<int>("MaxIter", 1000)),
    _mode(params.GetOptionalParam<std::string>("Precision", "fp64") == "fp32_ir" ? HC_SOLVE_FP32_IR
                                                                                 : HC_SOLVE_FP64)
{
  // "Preconditioner": "Jacobi" (HypreDiagScale, hephaestus_solvers.hpp:52) or
  // "AMS" (the gradient-space correction, the one-level form of the HypreAMS
  // that HCurlFormulation::ConstructJacobianPreconditioner sets)
  const std::string pc = params.GetOptionalParam<std::string>("Preconditioner", "Jacobi");
  if (pc == "AMS")
    _mode |= HC_SOLVE_AUX;
  else if (pc != "Jacobi")
    throw Error("DefaultJacobiPCGSolver: Preconditioner must be \"Jacobi\" or \"AMS\", got \"" + pc + "\"");
}

void
DefaultJacobiPCGSolver::Mult(const DeviceVector & B, DeviceVector & X)
{
  _stats = hc_solve_stats{};
  const int rc = hc_pcg_solve(_op.Handle(), B.Data(), X.Data(), 1, _tol, _maxit, _mode, &_stats, nullptr);
  CheckHc(rc, "DefaultJacobiPCGSolver::Mult");
}

// ---------------------------------------------------------------------------
TimeDomainEquationSystemProblemOperator::TimeDomainEquationSystemProblemOperator(
    TimeDependentEquationSystem & es, GridFunctions & gridfunctions, BCMap & bc_map, Coefficients & coefficients,
    Sources & sources, InputParameters solver_options)
  : _es(es),
    _gfs(gridfunctions),
    _bc(bc_map),
    _coefs(coefficients),
    _sources(sources),
    _solver_options(std::move(solver_options))
{
}

void
TimeDomainEquationSystemProblemOperator::ImplicitSolve(double dt, const DeviceVector & X, DeviceVector & dX_dt)
{
  // bind the trial variable to the state X
  GridFunction * u = _gfs.Get(_es._trial_var_names.at(0));
  if (u->Data() != X.Data())
    u->CopyFrom(X);
  _es.SetTimeStep(dt);
  _es.UpdateEquationSystem(_bc, _sources);
  _es.FormLinearSystem(_X, _B);
  DefaultJacobiPCGSolver solver(_solver_options, _es.GetGradient());
  solver.Mult(_B, _X);
  _stats = solver.Stats();
  _es.RecoverFEMSolution(_X, _gfs);
  dX_dt.CopyFrom(_X);
}

const hc_solve_stats &
TimeDomainEquationSystemProblemOperator::Step(double dt, const std::string & var_name)
{
  GridFunction * u = _gfs.Get(var_name);
  DeviceVector k;
  ImplicitSolve(dt, *u, k);
  u->Axpby(dt, k, 1.0);
  return _stats;
}

} // namespace hephaestus
