// Bilinear / linear forms of the H(curl) path in partial assembly on the GPU,
// and the kernels that fill them: Kernel<T> (src/kernels/kernel_base.hpp:10-26),
// CurlCurlKernel (src/kernels/curl_curl_kernel.{hpp,cpp}), VectorFEMassKernel
// (src/kernels/vector_fe_mass_kernel.{hpp,cpp}), WeakCurlCurlKernel
// (src/kernels/weak_curl_curl_kernel.{hpp,cpp}).
#pragma once

#include "core.hpp"

namespace hephaestus
{

// Domain integrators a BilinearForm accepts (the MFEM integrator classes the
// reference kernels add): each carries a non-owned coefficient.
struct CurlCurlIntegrator
{
  const Coefficient * coef;
};
struct VectorFEMassIntegrator
{
  const Coefficient * coef;
};
// integrators of the AV formulation's other blocks
struct DiffusionIntegrator
{
  const Coefficient * coef;
};
struct MixedVectorGradientIntegrator
{
  const Coefficient * coef;
};
struct VectorFEWeakDivergenceIntegrator
{
  const Coefficient * coef;
};

// ParBilinearForm counterpart: a(u,w) = sum (a_i curl u, curl w) + sum (b_j u, w)
// assembled as one libhcurl operator (partial assembly: quadrature-point
// factors on the device).  Assemble() re-runs the setup only when the
// coefficient values changed (the dt rule of
// TimeDependentEquationSystem::SetTimeStep, equation_system.cpp:413-426,
// applies through the dt-scaled coefficient).
class BilinearForm
{
public:
  explicit BilinearForm(FESpace * fes) : _fes(fes) {}
  ~BilinearForm();
  BilinearForm(const BilinearForm &) = delete;
  BilinearForm & operator=(const BilinearForm &) = delete;

  void AddDomainIntegrator(CurlCurlIntegrator i) { _curl.push_back(i.coef); }
  void AddDomainIntegrator(VectorFEMassIntegrator i) { _mass.push_back(i.coef); }
  // H1 forms (the AV system's (sigma grad V, grad V')) are recorded, not
  // assembled on their own: AVEquationSystem builds them into its block operator
  void AddDomainIntegrator(DiffusionIntegrator i) { _diff.push_back(i.coef); }
  void ClearIntegrators()
  {
    _curl.clear();
    _mass.clear();
    _diff.clear();
  }
  const std::vector<const Coefficient *> & MassCoefficients() const { return _mass; }
  const std::vector<const Coefficient *> & DiffusionCoefficients() const { return _diff; }
  // HC_FACE_* bits of the PEC faces (fixed once the form is assembled)
  void SetEssentialFaces(int faces) { _ess_faces = faces; }
  void Assemble();
  // y = A x with essential (PEC) rows/columns eliminated, A_ess,ess = I
  void Mult(const DeviceVector & x, DeviceVector & y) const;
  // lf += a * K^T u, with K this form without elimination restricted to the
  // free DOFs (u's PEC entries are zero for PEC problems); K is symmetric
  void AddMultTranspose(const DeviceVector & u, DeviceVector & lf, double a) const;
  hc_op_t Handle() const { return _op; }
  FESpace * ParFESpace() const { return _fes; }
  void SetPrecision(int prec_bits) { _prec = prec_bits; }
  // 0: setup was a no-op, 1: factors recomputed
  int LastAssembleWork() const { return _last_work; }

private:
  FESpace * _fes;
  std::vector<const Coefficient *> _curl, _mass, _diff;
  hc_op_t _op = nullptr;
  std::vector<double> _nu, _sig;
  int _prec = HC_PREC_FP64 | HC_PREC_FP32;
  int _ess_faces = HC_FACE_ALL;
  int _last_work = 0;
  mutable DeviceVector _tmp;
};

// ParMixedBilinearForm counterpart (trial space -> test space): records the
// coefficients of its integrators; AVEquationSystem builds its blocks.
class MixedBilinearForm
{
public:
  MixedBilinearForm(FESpace * trial, FESpace * test) : _trial(trial), _test(test) {}
  void AddDomainIntegrator(MixedVectorGradientIntegrator i) { _grad.push_back(i.coef); }
  void AddDomainIntegrator(VectorFEWeakDivergenceIntegrator i) { _wdiv.push_back(i.coef); }
  void ClearIntegrators()
  {
    _grad.clear();
    _wdiv.clear();
  }
  FESpace * TrialFESpace() const { return _trial; }
  FESpace * TestFESpace() const { return _test; }
  const std::vector<const Coefficient *> & GradientCoefficients() const { return _grad; }
  const std::vector<const Coefficient *> & WeakDivergenceCoefficients() const { return _wdiv; }

private:
  FESpace *_trial, *_test;
  std::vector<const Coefficient *> _grad, _wdiv;
};

// ParLinearForm counterpart: a device vector on the space.
class LinearForm : public DeviceVector
{
public:
  explicit LinearForm(FESpace * fes) : DeviceVector(fes->GetTrueVSize()), _fes(fes) { Fill(0.0); }
  FESpace * ParFESpace() const { return _fes; }
  // operator used by device-side sources (coil right-hand side)
  void SetOperator(hc_op_t op) { _op = op; }
  hc_op_t Operator() const { return _op; }

private:
  FESpace * _fes;
  hc_op_t _op = nullptr;
};

template <typename T>
class Kernel
{
public:
  Kernel() = default;
  virtual ~Kernel() = default;
  explicit Kernel(const InputParameters &) {}
  virtual void Init(GridFunctions & gridfunctions, const FESpaces & fespaces, BCMap & bc_map,
                    Coefficients & coefficients)
  {
  }
  virtual void Apply(T * form) = 0;
};

// (alpha curl u, curl u')
class CurlCurlKernel : public Kernel<BilinearForm>
{
public:
  explicit CurlCurlKernel(const InputParameters & params);
  void Init(GridFunctions & gridfunctions, const FESpaces & fespaces, BCMap & bc_map,
            Coefficients & coefficients) override;
  void Apply(BilinearForm * blf) override;
  std::string _coef_name;
  Coefficient * _coef{nullptr};
};

// (beta u, u')
class VectorFEMassKernel : public Kernel<BilinearForm>
{
public:
  explicit VectorFEMassKernel(const InputParameters & params);
  void Init(GridFunctions & gridfunctions, const FESpaces & fespaces, BCMap & bc_map,
            Coefficients & coefficients) override;
  void Apply(BilinearForm * blf) override;
  std::string _coef_name;
  Coefficient * _coef{nullptr};
};

// (sigma grad V, grad V') (src/kernels/diffusion_kernel.{hpp,cpp})
class DiffusionKernel : public Kernel<BilinearForm>
{
public:
  explicit DiffusionKernel(const InputParameters & params);
  void Init(GridFunctions & gridfunctions, const FESpaces & fespaces, BCMap & bc_map,
            Coefficients & coefficients) override;
  void Apply(BilinearForm * blf) override;
  std::string _coef_name;
  Coefficient * _coef{nullptr};
};

// (sigma grad V, u') (src/kernels/mixed_vector_gradient_kernel.{hpp,cpp})
class MixedVectorGradientKernel : public Kernel<MixedBilinearForm>
{
public:
  explicit MixedVectorGradientKernel(const InputParameters & params);
  void Init(GridFunctions & gridfunctions, const FESpaces & fespaces, BCMap & bc_map,
            Coefficients & coefficients) override;
  void Apply(MixedBilinearForm * mblf) override;
  std::string _coef_name;
  Coefficient * _coef{nullptr};
};

// (sigma u, grad V') (src/kernels/vector_fe_weak_divergence_kernel.{hpp,cpp})
class VectorFEWeakDivergenceKernel : public Kernel<MixedBilinearForm>
{
public:
  explicit VectorFEWeakDivergenceKernel(const InputParameters & params);
  void Init(GridFunctions & gridfunctions, const FESpaces & fespaces, BCMap & bc_map,
            Coefficients & coefficients) override;
  void Apply(MixedBilinearForm * mblf) override;
  std::string _coef_name;
  Coefficient * _coef{nullptr};
};

// -(alpha curl u_n, curl u'): the explicit part of a backward-Euler step
class WeakCurlCurlKernel : public Kernel<LinearForm>
{
public:
  explicit WeakCurlCurlKernel(const InputParameters & params);
  void Init(GridFunctions & gridfunctions, const FESpaces & fespaces, BCMap & bc_map,
            Coefficients & coefficients) override;
  void Apply(LinearForm * lf) override;
  std::string _coupled_gf_name;
  std::string _coef_name;
  GridFunction * _u{nullptr};
  Coefficient * _coef{nullptr};
  std::unique_ptr<BilinearForm> _curl_curl;
};

} // namespace hephaestus
