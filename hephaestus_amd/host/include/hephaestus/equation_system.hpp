// Equation systems, solver and time-step operator of the H(curl) path:
// EquationSystem / TimeDependentEquationSystem
// (src/equation_systems/equation_system.{hpp,cpp}), CurlCurlEquationSystem
// (src/formulations/HCurl/hcurl_formulation.cpp:109-157),
// DefaultJacobiPCGSolver (src/solvers/hephaestus_solvers.hpp:34-57) and
// TimeDomainEquationSystemProblemOperator::ImplicitSolve
// (src/problem_operators/time_domain_equation_system_problem_operator.cpp:30-59).
#pragma once

#include "forms.hpp"

namespace hephaestus
{

class EquationSystem
{
public:
  using ParBilinearFormKernel = Kernel<BilinearForm>;
  using ParLinearFormKernel = Kernel<LinearForm>;

  EquationSystem() = default;
  virtual ~EquationSystem() = default;

  std::vector<std::string> _trial_var_names;
  std::vector<std::string> _test_var_names;
  std::vector<FESpace *> _test_pfespaces;
  NamedFieldsMap<BilinearForm> _blfs;
  NamedFieldsMap<LinearForm> _lfs;

  virtual void AddTestVariableNameIfMissing(const std::string & name);
  virtual void AddTrialVariableNameIfMissing(const std::string & name);
  void AddKernel(const std::string & test_var_name, std::shared_ptr<ParBilinearFormKernel> blf_kernel);
  void AddKernel(const std::string & test_var_name, std::shared_ptr<ParLinearFormKernel> lf_kernel);
  virtual void ApplyBoundaryConditions(BCMap & bc_map);
  virtual void AddKernels() {}
  virtual void Init(GridFunctions & gridfunctions, const FESpaces & fespaces, BCMap & bc_map,
                    Coefficients & coefficients);
  virtual void BuildLinearForms(BCMap & bc_map, Sources & sources);
  virtual void BuildBilinearForms();
  virtual void BuildEquationSystem(BCMap & bc_map, Sources & sources);
  // X, B: true-DOF solution (initial guess) and right-hand side of the single
  // test variable, with the PEC constraint applied (B_ess = X_ess = 0)
  virtual void FormLinearSystem(DeviceVector & X, DeviceVector & B);
  // y = A x (the operator the solver sees)
  void Mult(const DeviceVector & x, DeviceVector & y) const;
  BilinearForm & GetGradient() const;
  virtual void RecoverFEMSolution(const DeviceVector & X, GridFunctions & gridfunctions);
  int EssentialFaces() const { return _ess_faces; }

protected:
  bool VectorContainsName(const std::vector<std::string> & v, const std::string & name) const;
  NamedFieldsMap<std::vector<std::shared_ptr<ParBilinearFormKernel>>> _blf_kernels_map;
  NamedFieldsMap<std::vector<std::shared_ptr<ParLinearFormKernel>>> _lf_kernels_map;
  int _ess_faces = 0;
};

class TimeDependentEquationSystem : public EquationSystem
{
public:
  TimeDependentEquationSystem();
  static std::string GetTimeDerivativeName(const std::string & name) { return "d" + name + "_dt"; }
  void AddTrialVariableNameIfMissing(const std::string & name) override;
  // re-assemble when |dt - dt_old| > 1e-12 dt (equation_system.cpp:413-426)
  virtual void SetTimeStep(double dt);
  virtual void UpdateEquationSystem(BCMap & bc_map, Sources & sources);
  ConstantCoefficient _dt_coef;
  std::vector<std::string> _trial_var_time_derivative_names;
};

// (dt alpha curl du/dt, curl u') + (beta du/dt, u') = -(alpha curl u_n, curl u') + sources
class CurlCurlEquationSystem : public TimeDependentEquationSystem
{
public:
  explicit CurlCurlEquationSystem(const InputParameters & params);
  void Init(GridFunctions & gridfunctions, const FESpaces & fespaces, BCMap & bc_map,
            Coefficients & coefficients) override;
  void AddKernels() override;
  std::string _h_curl_var_name, _alpha_coef_name, _beta_coef_name, _dtalpha_coef_name;
};

// AV formulation (src/formulations/AV/av_formulation.cpp:119-196): unknowns
// dA/dt (ND) and V (H1) of
//   (dt alpha curl dA/dt, curl A') + (beta dA/dt, A') + (sigma grad V, A')
//       = -(alpha curl A_n, curl A') + sources
//   (sigma dA/dt, grad V') + (sigma grad V, grad V') = 0
// with sigma = beta.  The bilinear forms are those of the reference's kernels;
// BuildBilinearForms assembles the A-A block as one device operator and the
// other blocks are applied through it (hc_op_av_apply).  Solve uses the block
// solver hc_op_av_solve; Step advances A by dt.
class AVEquationSystem : public TimeDependentEquationSystem
{
public:
  using MixedBilinearFormKernel = Kernel<MixedBilinearForm>;
  explicit AVEquationSystem(const InputParameters & params);
  void Init(GridFunctions & gridfunctions, const FESpaces & fespaces, BCMap & bc_map,
            Coefficients & coefficients) override;
  void AddKernels() override;
  using EquationSystem::AddKernel;
  // mixed kernel with trial variable trial_var_name in the rows of test_var_name
  void AddKernel(const std::string & trial_var_name, const std::string & test_var_name,
                 std::shared_ptr<MixedBilinearFormKernel> kernel);
  void BuildBilinearForms() override;
  // the block system with the current forms; (dA_dt, V) hold the initial guess
  hc_solve_stats Solve(DeviceVector & dA_dt, DeviceVector & V, double tol, int maxit);
  // one backward-Euler step: build, solve, A <- A + dt dA/dt (V and dA/dt kept
  // in their grid functions)
  hc_solve_stats Step(double dt, GridFunctions & gridfunctions, BCMap & bc_map, Sources & sources, double tol,
                      int maxit);
  std::string _a_name, _v_name, _alpha_coef_name, _beta_coef_name, _dtalpha_coef_name, _neg_beta_coef_name;
  ConstantCoefficient _neg_coef{-1.0};

private:
  std::map<std::pair<std::string, std::string>, std::vector<std::shared_ptr<MixedBilinearFormKernel>>>
      _mblf_kernels;
  std::map<std::pair<std::string, std::string>, std::unique_ptr<MixedBilinearForm>> _mblfs;
  GridFunctions * _gfs = nullptr;
  const FESpaces * _fes = nullptr;
  BCMap * _bcs = nullptr;
  Coefficients * _coefs = nullptr;
};

// Jacobi-preconditioned CG on the GPU operator; "Tolerance" / "MaxIter" /
// "Precision" ("fp64" or "fp32_ir") / "Preconditioner" ("Jacobi" or "AMS", the
// latter the one-level gradient-space correction) parameters.  Stopping rule
// ||b - A x||_2 <= tol ||b||_2 confirmed in FP64.
class DefaultJacobiPCGSolver
{
public:
  DefaultJacobiPCGSolver(const InputParameters & params, BilinearForm & op);
  void Mult(const DeviceVector & B, DeviceVector & X);
  const hc_solve_stats & Stats() const { return _stats; }

private:
  BilinearForm & _op;
  double _tol;
  int _maxit;
  int _mode;
  hc_solve_stats _stats{};
};

// One backward-Euler step of a time-dependent equation system: ImplicitSolve
// computes dX_dt from X (u_n); Step advances X by dt.
class TimeDomainEquationSystemProblemOperator
{
public:
  TimeDomainEquationSystemProblemOperator(TimeDependentEquationSystem & es, GridFunctions & gridfunctions,
                                          BCMap & bc_map, Coefficients & coefficients, Sources & sources,
                                          InputParameters solver_options);
  void ImplicitSolve(double dt, const DeviceVector & X, DeviceVector & dX_dt);
  // X <- X + dt * dX_dt (BackwardEulerSolver step); returns the solve statistics
  const hc_solve_stats & Step(double dt, const std::string & var_name);
  const hc_solve_stats & LastStats() const { return _stats; }

private:
  TimeDependentEquationSystem & _es;
  GridFunctions & _gfs;
  BCMap & _bc;
  Coefficients & _coefs;
  Sources & _sources;
  InputParameters _solver_options;
  DeviceVector _X, _B;
  hc_solve_stats _stats{};
};

} // namespace hephaestus
