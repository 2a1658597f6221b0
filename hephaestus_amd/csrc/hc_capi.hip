// C ABI of libhcurl.so (declared in include/hc_hcurl.h): operator handles.
// See that header for the reference interface each entry point replaces.
#include "hc_op.hpp"
#include "hc_aux.hpp"
#include "hc_basis.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace hc
{

namespace
{
thread_local std::string g_last_error;
}

int
fail(int code, const std::string & msg)
{
  g_last_error = msg;
  return code;
}

const char *
last_error()
{
  return g_last_error.c_str();
}

int
fail_exception()
{
  try
  {
    throw;
  }
  catch (const std::bad_alloc &)
  {
    return fail(HC_ENOMEM, "host memory allocation failed");
  }
  catch (const std::exception & ex)
  {
    return fail(HC_EHIP, std::string("unexpected error: ") + ex.what());
  }
  catch (...)
  {
    return fail(HC_EHIP, "unexpected error");
  }
}

} // namespace hc

using hc::fail;

namespace
{

// 1D bases (hc_basis.hpp): the same constexpr functions the tile kernels
// fold at compile time.
using hc::basis::gauss_legendre01;
using hc::basis::gauss_lobatto01;
using hc::basis::lagrange;

// the tables of orders 1..PMAX, evaluated at compile time (the values the
// tile kernels use)
constexpr hc::basis::Tab<double> kTabs[hc::PMAX + 1] = {hc::basis::Tab<double>{}, hc::basis::make_tab(1),
                                                         hc::basis::make_tab(2), hc::basis::make_tab(3)};
static_assert(hc::PMAX == 3, "kTabs lists orders 1..3");

void
make_tables(int p, hc::Tables<double> & t)
{
  std::memset(&t, 0, sizeof(t));
  const hc::basis::Tab<double> & k = kTabs[p];
  for (int i = 0; i < hc::QMAXC * hc::PMAX; i++)
    t.Bo[i] = k.Bo[i];
  for (int i = 0; i < hc::QMAXC * hc::QMAXC; i++)
  {
    t.Bc[i] = k.Bc[i];
    t.Gc[i] = k.Gc[i];
  }
  for (int i = 0; i < hc::QMAXC; i++)
  {
    t.qw[i] = k.qw[i];
    t.qp[i] = k.qp[i];
  }
}

// Dg[i][m] = L_m'(xi_i): derivative of the closed-point Lagrange basis at the
// open points (the 1D factor of the discrete gradient, hc_aux.hpp).
void
make_grad_1d(int p, double * dg)
{
  double on[hc::PMAX], ow[hc::PMAX], cn[hc::PMAX + 1];
  gauss_legendre01(p, on, ow);
  gauss_lobatto01(p + 1, cn);
  for (int i = 0; i < p; i++)
  {
    double v[hc::PMAX + 1], d[hc::PMAX + 1];
    lagrange(p + 1, cn, on[i], v, d);
    for (int m = 0; m <= p; m++)
      dg[i * (p + 1) + m] = d[m];
  }
}

template <typename A>
void
cast_tables(const hc::Tables<double> & a, hc::Tables<A> & b)
{
  for (int i = 0; i < hc::QMAXC * hc::PMAX; i++)
    b.Bo[i] = (A)a.Bo[i];
  for (int i = 0; i < hc::QMAXC * hc::QMAXC; i++)
  {
    b.Bc[i] = (A)a.Bc[i];
    b.Gc[i] = (A)a.Gc[i];
  }
  for (int i = 0; i < hc::QMAXC; i++)
  {
    b.qw[i] = (A)a.qw[i];
    b.qp[i] = (A)a.qp[i];
  }
}

} // namespace

namespace
{

void
set_plan(hc_op_t op, hc::ApplyPlan & pl, const hc::TileCandidate * c)
{
  pl.g = op->g;
  pl.cand = c;
  pl.g.T[0] = c->shape.tx;
  pl.g.T[1] = c->shape.ty;
  pl.g.T[2] = c->shape.tz;
  for (int a = 0; a < 3; a++)
    pl.g.nt[a] = (pl.g.n[a] + pl.g.T[a] - 1) / pl.g.T[a];
  pl.segs = hc::make_plane_segs(pl.g);
}

template <typename T>
hc::ApplyPlan &
plan_of(hc_op_t op)
{
  return sizeof(T) == 8 ? op->plan64 : op->plan32;
}

// y = A x for nrhs right-hand sides (rhs r at r*N) in the precision of T;
// with resid_b, y = resid_b - A x (the kernels form the difference as they
// store).
template <typename T>
int
apply_t(hc_op_t op, const T * x, T * y, int nrhs, const int * done, hipStream_t st, const T * resid_b = nullptr)
{
  const bool f32 = sizeof(T) == 4;
  int rc = op->ensure_ws(nrhs, f32);
  if (rc)
    return rc;
  hc::ApplyPlan & pl = plan_of<T>(op);
  const hc::Tables<T> & tb = *(const hc::Tables<T> *)(f32 ? (const void *)&op->tb32 : (const void *)&op->tb64);
  const T * qd = (const T *)(f32 ? (const void *)op->qd32 : (const void *)op->qd64);
  T * part = (T *)(f32 ? (void *)op->ws.part32 : (void *)op->ws.part64);
  const int ntiles = pl.g.nt[0] * pl.g.nt[1] * pl.g.nt[2];
  const int rb = nrhs == 1 ? 1 : (op->rhs_block > 0 ? std::min(op->rhs_block, nrhs) : hc::default_rhs_block(ntiles, nrhs));
  HC_TRY_HIP(hc::apply_op<T>(pl.g, pl.segs, tb, qd, op->ef(f32), x, y, part, done, st, pl.cand, nrhs, hc::STATE_STRIDE_INT, rb,
                             f32 ? op->desc_dev + 1 : op->desc_dev, resid_b));
  return HC_OK;
}

// Per-element coefficients must be finite and non-negative (a negative value
// makes the operator indefinite; zero switches a term off, as in the forms of
// WeakCurlCurlKernel or of a unit mass matrix).
int
check_coefficients(int64_t E, const double * nu_e, const double * sigma_e, const char * who)
{
  for (int64_t e = 0; e < E; e++)
    if (!(nu_e[e] >= 0) || !std::isfinite(nu_e[e]) || !(sigma_e[e] >= 0) || !std::isfinite(sigma_e[e]))
      return fail(HC_EINVAL, std::string(who) + ": coefficients must be finite and non-negative (element " +
                                 std::to_string(e) + ")");
  return HC_OK;
}

} // namespace

extern "C" {

const char *
hc_last_error(void)
{
  return hc::last_error();
}

int
hc_op_create(const hc_mesh_desc * mesh, int order, const double * nu_e, const double * sigma_e, int precision,
             hc_op_t * out)
try
{
  if (!mesh || !out || !nu_e || !sigma_e)
    return fail(HC_EINVAL, "hc_op_create: null argument");
  *out = nullptr;
  if (order < 1 || order > hc::PMAX)
    return fail(HC_EINVAL, "hc_op_create: order must be 1.." + std::to_string(hc::PMAX));
  if (mesh->nx < 1 || mesh->ny < 1 || mesh->nz < 1)
    return fail(HC_EINVAL, "hc_op_create: mesh needs at least one element per direction");
  {
    // the kernels index DOFs within one component with 32-bit integers
    const int64_t m[3] = {(int64_t)mesh->nx * order + 1, (int64_t)mesh->ny * order + 1, (int64_t)mesh->nz * order + 1};
    if (m[0] * m[1] * m[2] >= (int64_t(1) << 31))
      return fail(HC_EINVAL, "hc_op_create: mesh too large (each vector component must hold fewer than 2^31 DOFs)");
  }
  if (!mesh->vert)
    for (int a = 0; a < 3; a++)
      if (!(mesh->len[a] > 0) || !std::isfinite(mesh->len[a]) || !std::isfinite(mesh->x0[a]))
        return fail(HC_EINVAL, "hc_op_create: box lengths must be positive and finite, the origin finite");
  {
    const int64_t E = (int64_t)mesh->nx * mesh->ny * mesh->nz;
    int rc = check_coefficients(E, nu_e, sigma_e, "hc_op_create");
    if (rc)
      return rc;
  }
  if (!(precision & (HC_PREC_FP64 | HC_PREC_FP32)))
    precision = HC_PREC_FP64;
  hc_op_t op = new hc_op_s();
  op->p = order;
  op->precision = precision | HC_PREC_FP64; // the FP64 copy is always kept (setup, refinement)
  hc::Geom & g = op->g;
  g.n[0] = mesh->nx;
  g.n[1] = mesh->ny;
  g.n[2] = mesh->nz;
  g.p = order;
  g.pec = mesh->pec_faces & HC_FACE_ALL;
  const hc::TileShape ts = hc::default_tile(order);
  g.T[0] = ts.tx;
  g.T[1] = ts.ty;
  g.T[2] = ts.tz;
  for (int a = 0; a < 3; a++)
    g.nt[a] = (g.n[a] + g.T[a] - 1) / g.T[a];
  g.off[0] = 0;
  for (int c = 0; c < 3; c++)
  {
    for (int a = 0; a < 3; a++)
      g.cd[c][a] = (a == c) ? (int64_t)g.n[a] * order : (int64_t)g.n[a] * order + 1;
    g.off[c + 1] = g.off[c] + g.cd[c][0] * g.cd[c][1] * g.cd[c][2];
  }
  op->E = (int64_t)g.n[0] * g.n[1] * g.n[2];
  op->N = g.off[3];
  {
    const hc::TileCandidate * cands;
    hc::tile_candidates(order, &cands);
    set_plan(op, op->plan64, &cands[0]);
    set_plan(op, op->plan32, &cands[0]);
  }
  make_tables(order, op->tb64);
  cast_tables(op->tb64, op->tb32);
  op->nu.assign(nu_e, nu_e + op->E);
  op->sig.assign(sigma_e, sigma_e + op->E);
  op->box = (mesh->vert == nullptr);
  for (int a = 0; a < 3; a++)
  {
    op->h[a] = mesh->len[a] / g.n[a];
    op->x0[a] = mesh->x0[a];
  }
  // PEC count (host, structured)
  int64_t ness = 0;
  for (int c = 0; c < 3; c++)
  {
    int64_t cnt[3];
    for (int a = 0; a < 3; a++)
      cnt[a] = g.cd[c][a];
    // free count = product over closed axes of (dim - #pec faces on that axis) x open dim
    int64_t freec = 1;
    for (int a = 0; a < 3; a++)
    {
      int64_t f = cnt[a];
      if (a != c)
        f -= ((g.pec >> (2 * a)) & 1) + ((g.pec >> (2 * a + 1)) & 1);
      freec *= f;
    }
    ness += cnt[0] * cnt[1] * cnt[2] - freec;
  }
  op->n_ess = ness;
  const int Q = order + 1;
  const int64_t nq = op->E * 12 * Q * Q * Q;
  int rc;
  if ((rc = op->alloc(op->d_nu, op->E)) || (rc = op->alloc(op->d_sig, op->E)) || (rc = op->alloc(op->qd64, nq)) ||
      (rc = op->alloc(op->diag, op->N)) || (rc = op->alloc(op->dinv64, op->N)) || (rc = op->alloc(op->desc_dev, 2)) ||
      (rc = op->alloc(op->d_ratio, hc::TERM_RATIO_BLOCKS)))
  {
    delete op;
    return rc;
  }
  {
    double dg1d[hc::PMAX * (hc::PMAX + 1)];
    make_grad_1d(order, dg1d);
    op->ag = hc::make_aux_geom(g, dg1d);
    if ((rc = op->alloc(op->ginv64, op->ag.NN)))
    {
      delete op;
      return rc;
    }
  }
  if (op->precision & HC_PREC_FP32)
  {
    if ((rc = op->alloc(op->qd32, nq)) || (rc = op->alloc(op->dinv32, op->N)) ||
        (rc = op->alloc(op->ginv32, op->ag.NN)))
    {
      delete op;
      return rc;
    }
  }
  if (op->box)
  {
    // element factors: J = diag(h), |det J| = h0 h1 h2
    if ((rc = op->alloc(op->qc64, 2 * op->E)) || (op->qd32 && (rc = op->alloc(op->qc32, 2 * op->E))))
    {
      delete op;
      return rc;
    }
    const double det = op->h[0] * (op->h[1] * op->h[2]);
    for (int a = 0; a < 3; a++)
    {
      op->ck[a] = op->h[a] * op->h[a] / det;
      op->mk[a] = det / (op->h[a] * op->h[a]);
    }
  }
  else
  {
    const int64_t nv = (int64_t)(g.n[0] + 1) * (g.n[1] + 1) * (g.n[2] + 1) * 3;
    if ((rc = op->alloc(op->d_vert, nv)))
    {
      delete op;
      return rc;
    }
    if (hipMemcpy(op->d_vert, mesh->vert, nv * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    {
      delete op;
      return fail(HC_EHIP, "hc_op_create: vertex upload failed");
    }
  }
  // the scratch address of hc_op_apply_noalloc on this device
  if ((rc = hc::noalloc_setup()))
  {
    delete op;
    return rc;
  }
  *out = op;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_set_coefficients(hc_op_t op, const double * nu_e, const double * sigma_e)
try
{
  if (!op || !nu_e || !sigma_e)
    return fail(HC_EINVAL, "hc_op_set_coefficients: null argument");
  if (int rc = check_coefficients(op->E, nu_e, sigma_e, "hc_op_set_coefficients"))
    return rc;
  op->nu.assign(nu_e, nu_e + op->E);
  op->sig.assign(sigma_e, sigma_e + op->E);
  op->coef_dirty = true;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_setup(hc_op_t op, double dt, void * stream)
try
{
  if (!op)
    return fail(HC_EINVAL, "hc_op_setup: null handle");
  if (!(dt > 0) || !std::isfinite(dt))
    return fail(HC_EINVAL, "hc_op_setup: dt must be positive and finite");
  hipStream_t st = (hipStream_t)stream;
  // same rule as TimeDependentEquationSystem::SetTimeStep: re-assemble only when
  // |dt - dt_old| > 1e-12 dt (or the coefficients changed)
  if (op->setup_done && !op->coef_dirty && std::fabs(dt - op->last_dt) <= 1.0e-12 * dt)
    return HC_OK;
  if (op->coef_dirty)
  {
    HC_TRY_HIP(hipMemcpyAsync(op->d_nu, op->nu.data(), op->E * sizeof(double), hipMemcpyHostToDevice, st));
    HC_TRY_HIP(hipMemcpyAsync(op->d_sig, op->sig.data(), op->E * sizeof(double), hipMemcpyHostToDevice, st));
  }
  HC_TRY_HIP(hc::pa_setup(op->g, op->tb64, op->d_vert, op->h, op->d_nu, op->d_sig, dt, op->qd64, st));
  const int Q = op->p + 1;
  if (op->qd32)
    HC_TRY_HIP(hc::cast_f32(op->qd64, op->qd32, op->E * 12 * Q * Q * Q, st));
  if (op->qc64)
    HC_TRY_HIP(hc::elem_factors(op->E, op->d_nu, op->d_sig, dt, op->qc64, op->qc32, st));
  HC_TRY_HIP(hc::diagonal(op->g, op->tb64, op->qd64, op->diag, st));
  HC_TRY_HIP(hc::invert_diag(op->diag, op->dinv64, op->dinv32, op->N, st));
  // node diagonal of the gradient-space correction (computed in place)
  HC_TRY_HIP(hc::aux_diag_inv(op->g, op->ag, op->tb64, op->qd64, op->ginv64, op->ginv64, op->ginv32, st));
  // sigma h^2 / nu per element (block minima)
  HC_TRY_HIP(hc::term_ratio_blocks(op->qd64, op->E, op->p, op->d_ratio, st));
  double rmin[hc::TERM_RATIO_BLOCKS];
  HC_TRY_HIP(hipMemcpyAsync(rmin, op->d_ratio, sizeof(rmin), hipMemcpyDeviceToHost, st));
  HC_TRY_HIP(hipStreamSynchronize(st));
  {
    double m = 1.0e300;
    for (int i = 0; i < hc::TERM_RATIO_BLOCKS; i++)
      m = std::min(m, rmin[i]);
    op->term_ratio = m / ((double)op->p * op->p);
  }
  op->setup_done = true;
  op->coef_dirty = false;
  op->last_dt = dt;
  op->setup_version++; // a multigrid hierarchy of the old setup is rebuilt on its next use
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int64_t
hc_op_ndofs(hc_op_t op)
{
  return op ? op->N : -1;
}

int
hc_op_ess_count(hc_op_t op, int64_t * n_ess)
try
{
  if (!op || !n_ess)
    return fail(HC_EINVAL, "hc_op_ess_count: null argument");
  *n_ess = op->n_ess;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int64_t
hc_op_device_bytes(hc_op_t op)
{
  return op ? op->bytes : -1;
}

int
hc_op_apply(hc_op_t op, const double * x, double * y, int nrhs, void * stream)
try
{
  if (!op || !x || !y)
    return fail(HC_EINVAL, "hc_op_apply: null argument");
  if (!op->setup_done)
    return fail(HC_EINVAL, "hc_op_apply: hc_op_setup has not been called");
  if (nrhs < 1)
    return fail(HC_EINVAL, "hc_op_apply: nrhs must be >= 1");
  return apply_t<double>(op, x, y, nrhs, nullptr, (hipStream_t)stream);
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_apply_f32(hc_op_t op, const float * x, float * y, int nrhs, void * stream)
try
{
  if (!op || !x || !y)
    return fail(HC_EINVAL, "hc_op_apply_f32: null argument");
  if (!op->setup_done)
    return fail(HC_EINVAL, "hc_op_apply_f32: hc_op_setup has not been called");
  if (!op->qd32)
    return fail(HC_EINVAL, "hc_op_apply_f32: operator created without HC_PREC_FP32");
  if (nrhs < 1)
    return fail(HC_EINVAL, "hc_op_apply_f32: nrhs must be >= 1");
  return apply_t<float>(op, x, y, nrhs, nullptr, (hipStream_t)stream);
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_apply_curl(hc_op_t op, const double * x, double * y, void * stream)
try
{
  if (!op || !x || !y)
    return fail(HC_EINVAL, "hc_op_apply_curl: null argument");
  if (!op->setup_done)
    return fail(HC_EINVAL, "hc_op_apply_curl: hc_op_setup has not been called");
  int rc = op->ensure_ws(1, false);
  if (rc)
    return rc;
  // the quadrature data hold dt * nu: scale the curl term back by 1/dt
  HC_TRY_HIP(hc::apply_scaled<double>(op->plan64.g, op->plan64.segs, op->tb64, op->qd64, op->ef(false), x, y, op->ws.part64, 0.0,
                                      1.0 / op->last_dt, op->plan64.cand, op->desc_dev, (hipStream_t)stream));
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int64_t
hc_op_nnodes(hc_op_t op)
{
  if (!op)
  {
    (void)fail(HC_EINVAL, "hc_op_nnodes: null handle");
    return -1;
  }
  return op->ag.NN;
}

int
hc_op_av_apply(hc_op_t op, const double * a, const double * v, double * ya, double * yv, void * stream)
try
{
  if (!op || !a || !v || !ya || !yv)
    return fail(HC_EINVAL, "hc_op_av_apply: null argument");
  if (!op->setup_done)
    return fail(HC_EINVAL, "hc_op_av_apply: hc_op_setup has not been called");
  if ((const double *)ya == a || (const double *)ya == v || (const double *)yv == a || (const double *)yv == v ||
      ya == yv)
    return fail(HC_EINVAL, "hc_op_av_apply: outputs must not alias the inputs or each other");
  int rc = op->ensure_ws(1, false);
  if (rc)
    return rc;
  hipStream_t st = (hipStream_t)stream;
  const hc::ApplyPlan & pl = op->plan64;
  hc::Geom g0 = pl.g;
  g0.pec = 0; // the mass term of s acts on every DOF (v has no essential nodes)
  double * s = op->ws.r64;
  double * k = op->ws.p64;
  // s = a0 + G v
  HC_TRY_HIP(hc::av_gather(pl.g, op->ag, a, v, s, st));
  // ya = M s (mass term alone, no elimination), yv = G^T M s
  HC_TRY_HIP(hc::apply_scaled<double>(g0, pl.segs, op->tb64, op->qd64, op->ef(false), s, ya, op->ws.part64, 1.0, 0.0, pl.cand,
                                      op->desc_dev, st));
  HC_TRY_HIP(hc::grad_transpose(op->ag, ya, yv, st));
  // k = dt K a0 on the free DOFs (the quadrature data hold dt nu); ya += k there, ya = a on PEC DOFs
  HC_TRY_HIP(hc::apply_scaled<double>(pl.g, pl.segs, op->tb64, op->qd64, op->ef(false), a, k, op->ws.part64, 0.0, 1.0, pl.cand,
                                      op->desc_dev, st));
  HC_TRY_HIP(hc::av_finish(pl.g, op->ag, a, k, ya, st));
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_set_rhs_block(hc_op_t op, int rhs_block)
try
{
  if (!op || rhs_block < 0)
    return fail(HC_EINVAL, "hc_op_set_rhs_block: bad argument");
  op->rhs_block = rhs_block;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_set_factor_mode(hc_op_t op, int mode)
try
{
  if (!op || (mode != 0 && mode != 1))
    return fail(HC_EINVAL, "hc_op_set_factor_mode: null handle or mode other than 0 and 1");
  op->ef_mode = mode;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_factor_mode(hc_op_t op)
try
{
  if (!op)
    return fail(HC_EINVAL, "hc_op_factor_mode: null handle");
  return op->ef(false).qc ? 1 : 0;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_term_ratio(hc_op_t op, double * ratio)
try
{
  if (!op || !ratio)
    return fail(HC_EINVAL, "hc_op_term_ratio: null argument");
  if (!op->setup_done)
    return fail(HC_EINVAL, "hc_op_term_ratio: hc_op_setup has not been called");
  *ratio = op->term_ratio;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_precond_apply(hc_op_t op, int mode, const double * r, double * z, void * stream)
try
{
  if (!op || !r || !z)
    return fail(HC_EINVAL, "hc_op_precond_apply: null argument");
  if (!op->setup_done)
    return fail(HC_EINVAL, "hc_op_precond_apply: hc_op_setup has not been called");
  if ((mode & ~(HC_SOLVE_AUX | 0xF)) || (hc::prec_mode(mode) != HC_SOLVE_FP64 && hc::prec_mode(mode) != HC_SOLVE_FP32_IR))
    return fail(HC_EINVAL, "hc_op_precond_apply: unknown mode");
  hipStream_t st = (hipStream_t)stream;
  if (!(mode & HC_SOLVE_AUX))
  {
    HC_TRY_HIP(hc::mul(op->N, op->dinv64, r, z, st));
    return HC_OK;
  }
  int rc = op->ensure_ws(1, false);
  if (rc)
    return rc;
  HC_TRY_HIP(hc::aux_precond_apply(op->ag, op->dinv64, op->ginv64, r, op->ws.t64, z, st));
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_diagonal(hc_op_t op, double * d, void * stream)
try
{
  if (!op || !d)
    return fail(HC_EINVAL, "hc_op_diagonal: null argument");
  if (!op->setup_done)
    return fail(HC_EINVAL, "hc_op_diagonal: hc_op_setup has not been called");
  HC_TRY_HIP(hipMemcpyAsync(d, op->diag, op->N * sizeof(double), hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

} // extern "C"

namespace hc
{
int
op_residual(hc_op_t op, const double * b, const double * x, double * r, hipStream_t st, int nrhs)
{
  return apply_t<double>(op, x, r, nrhs, nullptr, st, b);
}
} // namespace hc

namespace
{
const hc::TileCandidate *
find_candidate(int p, int tx, int ty, int tz, int prefetch)
{
  const hc::TileCandidate * c;
  const int n = hc::tile_candidates(p, &c);
  for (int i = 0; i < n; i++)
    if (c[i].shape.tx == tx && c[i].shape.ty == ty && c[i].shape.tz == tz &&
        (prefetch < 0 || c[i].mode == prefetch))
      return &c[i];
  return nullptr;
}
} // namespace

extern "C" {

int
hc_tile_list(int order, int * out, int max)
try
{
  const hc::TileCandidate * c;
  const int n = hc::tile_candidates(order, &c);
  for (int i = 0; i < n && i < max && out; i++)
  {
    out[4 * i + 0] = c[i].shape.tx;
    out[4 * i + 1] = c[i].shape.ty;
    out[4 * i + 2] = c[i].shape.tz;
    out[4 * i + 3] = c[i].mode;
  }
  return n;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_set_tile(hc_op_t op, int prec_bits, int tx, int ty, int tz, int prefetch)
try
{
  if (!op || (prec_bits != 64 && prec_bits != 32))
    return fail(HC_EINVAL, "hc_op_set_tile: bad handle or precision (64 or 32)");
  const hc::TileCandidate * c = find_candidate(op->p, tx, ty, tz, prefetch);
  if (!c)
    return fail(HC_EINVAL, "hc_op_set_tile: no compiled tile variant with this shape for order " +
                               std::to_string(op->p));
  set_plan(op, prec_bits == 64 ? op->plan64 : op->plan32, c);
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_get_tile(hc_op_t op, int prec_bits, int * tx, int * ty, int * tz, int * prefetch)
try
{
  if (!op || (prec_bits != 64 && prec_bits != 32) || !tx || !ty || !tz || !prefetch)
    return fail(HC_EINVAL, "hc_op_get_tile: bad argument");
  const hc::ApplyPlan & pl = prec_bits == 64 ? op->plan64 : op->plan32;
  *tx = pl.cand->shape.tx;
  *ty = pl.cand->shape.ty;
  *tz = pl.cand->shape.tz;
  *prefetch = pl.cand->mode;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_autotune(hc_op_t op, int reps, double * best_us, void * stream)
try
{
  if (!op)
    return fail(HC_EINVAL, "hc_op_autotune: null handle");
  if (!op->setup_done)
    return fail(HC_EINVAL, "hc_op_autotune: hc_op_setup has not been called");
  if (reps < 1)
    reps = 5;
  hipStream_t st = (hipStream_t)stream;
  int rc = op->ensure_ws(1, op->qd32 != nullptr);
  if (rc)
    return rc;
  const hc::TileCandidate * c;
  const int n = hc::tile_candidates(op->p, &c);
  hipEvent_t e0, e1;
  HC_TRY_HIP(hipEventCreate(&e0));
  HC_TRY_HIP(hipEventCreate(&e1));
  // input: a PEC-consistent random vector (r64 / r32 of the workspace)
  HC_TRY_HIP(hc::fill_uniform(&op->g, 17, op->ws.r64, op->N, st));
  for (int prec = 64; prec >= 32; prec -= 32)
  {
    if (prec == 32 && !op->qd32)
      break;
    if (prec == 32)
      HC_TRY_HIP(hc::cast_f32(op->ws.r64, op->ws.r32, op->N, st));
    hc::ApplyPlan & pl = prec == 64 ? op->plan64 : op->plan32;
    float best = 1e30f;
    const hc::TileCandidate * bestc = pl.cand;
    for (int i = 0; i < n; i++)
    {
      set_plan(op, pl, &c[i]);
      auto run = [&]() -> int {
        return prec == 64 ? apply_t<double>(op, op->ws.r64, op->ws.q64, 1, nullptr, st)
                          : apply_t<float>(op, op->ws.r32, op->ws.q32, 1, nullptr, st);
      };
      for (int w = 0; w < 2; w++)
        if ((rc = run()))
          return rc;
      HC_TRY_HIP(hipEventRecord(e0, st));
      for (int r = 0; r < reps; r++)
        if ((rc = run()))
          return rc;
      HC_TRY_HIP(hipEventRecord(e1, st));
      HC_TRY_HIP(hipEventSynchronize(e1));
      float ms = 0;
      HC_TRY_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best)
      {
        best = ms;
        bestc = &c[i];
      }
    }
    set_plan(op, pl, bestc);
    if (best_us)
      best_us[prec == 64 ? 0 : 1] = 1000.0 * best / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_destroy(hc_op_t op)
try
{
  if (!op)
    return fail(HC_EINVAL, "hc_op_destroy: null handle");
  (void)hipDeviceSynchronize();
  delete op;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_profile_apply(hc_op_t op, int prec_bits, const void * x, void * y, int reps, double * out_us, void * stream)
try
{
  if (!op || !x || !y || !out_us || reps < 1 || (prec_bits != 64 && prec_bits != 32))
    return fail(HC_EINVAL, "hc_op_profile_apply: bad argument");
  if (!op->setup_done)
    return fail(HC_EINVAL, "hc_op_profile_apply: hc_op_setup has not been called");
  if (prec_bits == 32 && !op->qd32)
    return fail(HC_EINVAL, "hc_op_profile_apply: operator created without HC_PREC_FP32");
  const bool f32 = prec_bits == 32;
  int rc = op->ensure_ws(1, f32);
  if (rc)
    return rc;
  hipStream_t st = (hipStream_t)stream;
  const hc::ApplyPlan & pl = f32 ? op->plan32 : op->plan64;
  hc::ApplyDesc d{};
  d.g = pl.g;
  d.segs = pl.segs;
  d.prows = (int)pl.segs.start[pl.segs.nseg];
  d.qd = f32 ? (const void *)op->qd32 : (const void *)op->qd64;
  d.ef = op->ef(f32);
  d.x = x;
  d.y = y;
  d.part = f32 ? (void *)op->ws.part32 : (void *)op->ws.part64;
  d.nrhs = 1;
  d.rb = 1;
  d.ntiles = pl.g.nt[0] * pl.g.nt[1] * pl.g.nt[2];
  const void * tb = f32 ? (const void *)&op->tb32 : (const void *)&op->tb64;
  std::vector<hipEvent_t> ev(3 * reps);
  for (auto & e : ev)
    HC_TRY_HIP(hipEventCreate(&e));
  for (int r = 0; r < reps; r++)
  {
    HC_TRY_HIP(hipEventRecord(ev[3 * r], st));
    HC_TRY_HIP(pl.cand->launch(d, nullptr, 1, d.ntiles, 0, tb, prec_bits, st));
    HC_TRY_HIP(hipEventRecord(ev[3 * r + 1], st));
    HC_TRY_HIP(f32 ? hc::assemble_planes<float>(d, nullptr, 1, d.prows, st)
                   : hc::assemble_planes<double>(d, nullptr, 1, d.prows, st));
    HC_TRY_HIP(hipEventRecord(ev[3 * r + 2], st));
  }
  HC_TRY_HIP(hipEventSynchronize(ev[3 * reps - 1]));
  double t_tile = 0, t_plane = 0;
  for (int r = 0; r < reps; r++)
  {
    float a = 0, b = 0;
    HC_TRY_HIP(hipEventElapsedTime(&a, ev[3 * r], ev[3 * r + 1]));
    HC_TRY_HIP(hipEventElapsedTime(&b, ev[3 * r + 1], ev[3 * r + 2]));
    t_tile += a;
    t_plane += b;
  }
  for (auto & e : ev)
    (void)hipEventDestroy(e);
  out_us[0] = 1000.0 * t_tile / reps;
  out_us[1] = 1000.0 * t_plane / reps;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_rhs_coil(hc_op_t op, uint64_t seed, double * b, void * stream)
try
{
  if (!op || !b)
    return fail(HC_EINVAL, "hc_op_rhs_coil: null argument");
  hc::CoilParams cp{};
  uint64_t k = 0;
  auto u01 = [&]() {
    uint64_t z = seed + (++k) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int j = 0; j < 4; j++)
  {
    for (int r = 0; r < 3; r++)
      cp.a[j][r] = 2.0 * u01() - 1.0;
    for (int r = 0; r < 3; r++)
      cp.c[j][r] = 0.3 + 0.4 * u01();
    cp.s[j] = 0.05 + 0.1 * u01();
  }
  const int p = op->p, Qr = p + 2;
  double on[hc::PMAX], ow[hc::PMAX], cn[hc::PMAX + 1];
  gauss_legendre01(p, on, ow);
  gauss_lobatto01(p + 1, cn);
  gauss_legendre01(Qr, cp.qp, cp.qw);
  for (int q = 0; q < Qr; q++)
  {
    double v[hc::PMAX + 1];
    lagrange(p, on, cp.qp[q], v, nullptr);
    for (int j = 0; j < p; j++)
      cp.Bo[q * p + j] = v[j];
    lagrange(p + 1, cn, cp.qp[q], v, nullptr);
    for (int j = 0; j <= p; j++)
      cp.Bc[q * (p + 1) + j] = v[j];
  }
  HC_TRY_HIP(hc::rhs_coil(op->g, cp, op->d_vert, op->h, op->x0, b, (hipStream_t)stream));
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_axpby(int64_t n, double a, const double * x, double b, double * y, void * stream)
try
{
  if (n < 0 || (n > 0 && (!x || !y)))
    return fail(HC_EINVAL, "hc_axpby: bad argument");
  HC_TRY_HIP(hc::axpby(n, a, x, b, y, (hipStream_t)stream));
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_op_zero_ess(hc_op_t op, double * v, void * stream)
try
{
  if (!op || !v)
    return fail(HC_EINVAL, "hc_op_zero_ess: null argument");
  HC_TRY_HIP(hc::zero_ess(op->g, v, (hipStream_t)stream));
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_fill_uniform(hc_op_t op, uint64_t seed, double * v, int64_t n, void * stream)
try
{
  if (!v || n < 0)
    return fail(HC_EINVAL, "hc_fill_uniform: bad argument");
  if (op && n != op->N)
    return fail(HC_EINVAL, "hc_fill_uniform: length differs from the operator's DOF count");
  HC_TRY_HIP(hc::fill_uniform(op ? &op->g : nullptr, seed, v, n, (hipStream_t)stream));
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

} // extern "C"
