// C ABI of libhcurl.so (declared in include/hc_hcurl.h).  See that header for
// the reference interface each entry point replaces.
#include "../../include/hc_hcurl.h"
#include "hc_kernels.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace
{

thread_local std::string g_last_error;

int
fail(int code, const std::string & msg)
{
  g_last_error = msg;
  return code;
}

#define HC_TRY_HIP(expr)                                                                                        \
  do                                                                                                             \
  {                                                                                                              \
    hipError_t _e = (expr);                                                                                      \
    if (_e != hipSuccess)                                                                                        \
      return fail(HC_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));                                   \
  } while (0)

// ---------------------------------------------------------------------------
// 1D bases: Gauss-Legendre points/weights and Gauss-Lobatto points on [0,1],
// Lagrange values/derivatives.
// ---------------------------------------------------------------------------
void
legendre_pd(int n, double x, double & P, double & dP)
{
  double p0 = 1.0, p1 = x;
  if (n == 0)
  {
    P = 1.0;
    dP = 0.0;
    return;
  }
  for (int k = 2; k <= n; k++)
  {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  P = p1;
  dP = n * (x * p1 - p0) / (x * x - 1.0);
}

void
gauss_legendre01(int n, double * x, double * w)
{
  for (int i = 0; i < n; i++)
  {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), P, dP;
    for (int it = 0; it < 100; it++)
    {
      legendre_pd(n, z, P, dP);
      const double dz = P / dP;
      z -= dz;
      if (std::fabs(dz) < 1e-16)
        break;
    }
    legendre_pd(n, z, P, dP);
    x[n - 1 - i] = 0.5 * (z + 1.0);
    w[n - 1 - i] = 1.0 / ((1.0 - z * z) * dP * dP);
  }
}

void
gauss_lobatto01(int n, double * x)
{
  const int p = n - 1;
  x[0] = 0.0;
  x[p] = 1.0;
  for (int i = 1; i < p; i++)
  {
    double z = -std::cos(M_PI * i / p), P, dP;
    for (int it = 0; it < 100; it++)
    {
      legendre_pd(p, z, P, dP);
      const double d2P = (2.0 * z * dP - p * (p + 1.0) * P) / (1.0 - z * z);
      const double dz = dP / d2P;
      z -= dz;
      if (std::fabs(dz) < 1e-16)
        break;
    }
    x[i] = 0.5 * (z + 1.0);
  }
}

void
lagrange(int n, const double * nodes, double t, double * val, double * der)
{
  for (int j = 0; j < n; j++)
  {
    double v = 1.0, d = 0.0;
    for (int m = 0; m < n; m++)
      if (m != j)
        v *= (t - nodes[m]) / (nodes[j] - nodes[m]);
    for (int k = 0; k < n; k++)
    {
      if (k == j)
        continue;
      double prod = 1.0 / (nodes[j] - nodes[k]);
      for (int m = 0; m < n; m++)
        if (m != j && m != k)
          prod *= (t - nodes[m]) / (nodes[j] - nodes[m]);
      d += prod;
    }
    val[j] = v;
    if (der)
      der[j] = d;
  }
}

void
make_tables(int p, hc::Tables<double> & t)
{
  std::memset(&t, 0, sizeof(t));
  const int Q = p + 1, D = p + 1;
  double on[hc::PMAX], ow[hc::PMAX], cn[hc::PMAX + 1];
  gauss_legendre01(p, on, ow);
  gauss_lobatto01(p + 1, cn);
  gauss_legendre01(Q, t.qp, t.qw);
  for (int q = 0; q < Q; q++)
  {
    double v[hc::PMAX + 1], d[hc::PMAX + 1];
    lagrange(p, on, t.qp[q], v, nullptr);
    for (int j = 0; j < p; j++)
      t.Bo[q * p + j] = v[j];
    lagrange(D, cn, t.qp[q], v, d);
    for (int j = 0; j < D; j++)
    {
      t.Bc[q * D + j] = v[j];
      t.Gc[q * D + j] = d[j];
    }
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

__global__ void
k_zero_f(float * v, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = 0.0f;
}

} // namespace

struct hc_op_s
{
  hc::Geom g{};
  hc::PlaneSegs segs{};
  hc::Tables<double> tb64{};
  hc::Tables<float> tb32{};
  int p = 0;
  int precision = 0;
  int64_t E = 0, N = 0, n_ess = 0;
  bool box = true;
  double h[3] = {0, 0, 0};
  std::vector<double> nu, sig;
  bool coef_dirty = true;
  bool setup_done = false;
  double last_dt = 0.0;
  int64_t bytes = 0;
  // device
  double * d_vert = nullptr;
  double * d_nu = nullptr;
  double * d_sig = nullptr;
  double * qd64 = nullptr;
  float * qd32 = nullptr;
  double * diag = nullptr;
  double * dinv64 = nullptr;
  float * dinv32 = nullptr;
  double * part64 = nullptr;
  float * part32 = nullptr;
  // solver work (allocated on first solve)
  double *r64 = nullptr, *p64 = nullptr, *q64 = nullptr;
  float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *x32 = nullptr, *b32 = nullptr;
  double * red = nullptr;
  hc::PcgState * st_dev = nullptr; // [0] inner/plain, [1] outer
  hc::PcgState * st_host = nullptr;
  std::vector<void *> allocs;

  template <typename T>
  int alloc(T *& ptr, int64_t count)
  {
    void * v = nullptr;
    const size_t sz = sizeof(T) * (size_t)(count > 0 ? count : 1);
    if (hipMalloc(&v, sz) != hipSuccess)
      return fail(HC_ENOMEM, "hipMalloc of " + std::to_string(sz) + " bytes failed");
    ptr = (T *)v;
    allocs.push_back(v);
    bytes += (int64_t)sz;
    return HC_OK;
  }

  ~hc_op_s()
  {
    for (void * v : allocs)
      (void)hipFree(v);
    if (st_host)
      (void)hipHostFree(st_host);
  }
};

namespace
{

int
ensure_solver_work(hc_op_t op, bool need32)
{
  if (!op->r64)
  {
    int rc;
    if ((rc = op->alloc(op->r64, op->N)) || (rc = op->alloc(op->p64, op->N)) || (rc = op->alloc(op->q64, op->N)) ||
        (rc = op->alloc(op->red, (int64_t)hc::RED_BLOCKS * 4)) || (rc = op->alloc(op->st_dev, 2)))
      return rc;
    if (hipHostMalloc((void **)&op->st_host, 2 * sizeof(hc::PcgState), hipHostMallocDefault) != hipSuccess)
      return fail(HC_ENOMEM, "hipHostMalloc failed");
    HC_TRY_HIP(hipMemset(op->st_dev, 0, 2 * sizeof(hc::PcgState)));
  }
  if (need32 && !op->r32)
  {
    int rc;
    if ((rc = op->alloc(op->r32, op->N)) || (rc = op->alloc(op->p32, op->N)) || (rc = op->alloc(op->q32, op->N)) ||
        (rc = op->alloc(op->x32, op->N)) || (rc = op->alloc(op->b32, op->N)))
      return rc;
  }
  return HC_OK;
}

// Read the device state of solve `k` (0 inner/plain, 1 outer); synchronises.
int
read_state(hc_op_t op, int k, hipStream_t st, hc::PcgState & out)
{
  HC_TRY_HIP(hipMemcpyAsync(op->st_host + k, op->st_dev + k, sizeof(hc::PcgState), hipMemcpyDeviceToHost, st));
  HC_TRY_HIP(hipStreamSynchronize(st));
  out = op->st_host[k];
  return HC_OK;
}

// Jacobi-PCG on A_T (precision T) with RHS b (T), solution x (T, initial guess
// or zero when zero_guess).  Returns iterations; *state gets the final state.
template <typename T>
int
run_pcg(hc_op_t op,
        const T * qd,
        const hc::Tables<T> & tb,
        T * part,
        const T * dinv,
        const T * b,
        T * x,
        T * r,
        T * p,
        T * q,
        bool zero_guess,
        double rtol,
        int maxit,
        hipStream_t st,
        hc::PcgState & fin)
{
  hc::PcgWork w{op->red, op->st_dev};
  if (!zero_guess)
    HC_TRY_HIP(hc::apply_op<T>(op->g, op->segs, tb, qd, x, q, part, nullptr, st));
  HC_TRY_HIP(hc::pcg_init<T>(b, zero_guess ? nullptr : q, dinv, r, p, op->N, rtol, w, st));
  int it = 0, chunk = 8;
  int rc = read_state(op, 0, st, fin);
  if (rc)
    return rc;
  const int * done = &op->st_dev[0].done;
  while (!fin.done && it < maxit)
  {
    const int n = std::min(chunk, maxit - it);
    for (int k = 0; k < n; k++, it++)
    {
      HC_TRY_HIP(hc::apply_op<T>(op->g, op->segs, tb, qd, p, q, part, done, st));
      HC_TRY_HIP(hc::pcg_dot_pq<T>(p, q, op->N, w, st));
      HC_TRY_HIP(hc::pcg_update_xr<T>(it, x, p, r, q, dinv, op->N, w, st));
      HC_TRY_HIP(hc::pcg_update_p<T>(it, p, r, dinv, op->N, w, st));
    }
    if ((rc = read_state(op, 0, st, fin)))
      return rc;
    chunk = std::min(chunk * 2, 64);
  }
  return HC_OK;
}

} // namespace

extern "C" {

const char *
hc_last_error(void)
{
  return g_last_error.c_str();
}

int
hc_op_create(const hc_mesh_desc * mesh, int order, const double * nu_e, const double * sigma_e, int precision,
             hc_op_t * out)
{
  if (!mesh || !out || !nu_e || !sigma_e)
    return fail(HC_EINVAL, "hc_op_create: null argument");
  *out = nullptr;
  if (order < 1 || order > hc::PMAX)
    return fail(HC_EINVAL, "hc_op_create: order must be 1.." + std::to_string(hc::PMAX));
  if (mesh->nx < 1 || mesh->ny < 1 || mesh->nz < 1)
    return fail(HC_EINVAL, "hc_op_create: mesh needs at least one element per direction");
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
  op->segs = hc::make_plane_segs(g);
  make_tables(order, op->tb64);
  cast_tables(op->tb64, op->tb32);
  op->nu.assign(nu_e, nu_e + op->E);
  op->sig.assign(sigma_e, sigma_e + op->E);
  op->box = (mesh->vert == nullptr);
  for (int a = 0; a < 3; a++)
    op->h[a] = mesh->len[a] / g.n[a];
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
      (rc = op->alloc(op->diag, op->N)) || (rc = op->alloc(op->dinv64, op->N)) ||
      (rc = op->alloc(op->part64, 4 * op->N)))
  {
    delete op;
    return rc;
  }
  if (op->precision & HC_PREC_FP32)
  {
    if ((rc = op->alloc(op->qd32, nq)) || (rc = op->alloc(op->dinv32, op->N)) ||
        (rc = op->alloc(op->part32, 4 * op->N)))
    {
      delete op;
      return rc;
    }
  }
  if (!op->box)
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
  *out = op;
  return HC_OK;
}

int
hc_op_set_coefficients(hc_op_t op, const double * nu_e, const double * sigma_e)
{
  if (!op || !nu_e || !sigma_e)
    return fail(HC_EINVAL, "hc_op_set_coefficients: null argument");
  op->nu.assign(nu_e, nu_e + op->E);
  op->sig.assign(sigma_e, sigma_e + op->E);
  op->coef_dirty = true;
  return HC_OK;
}

int
hc_op_setup(hc_op_t op, double dt, void * stream)
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
  HC_TRY_HIP(hc::diagonal(op->g, op->tb64, op->qd64, op->diag, st));
  HC_TRY_HIP(hc::invert_diag(op->diag, op->dinv64, op->dinv32, op->N, st));
  HC_TRY_HIP(hipStreamSynchronize(st));
  op->setup_done = true;
  op->coef_dirty = false;
  op->last_dt = dt;
  return HC_OK;
}

int64_t
hc_op_ndofs(hc_op_t op)
{
  return op ? op->N : -1;
}

int
hc_op_ess_count(hc_op_t op, int64_t * n_ess)
{
  if (!op || !n_ess)
    return fail(HC_EINVAL, "hc_op_ess_count: null argument");
  *n_ess = op->n_ess;
  return HC_OK;
}

int64_t
hc_op_device_bytes(hc_op_t op)
{
  return op ? op->bytes : -1;
}

int
hc_op_apply(hc_op_t op, const double * x, double * y, int nrhs, void * stream)
{
  if (!op || !x || !y)
    return fail(HC_EINVAL, "hc_op_apply: null argument");
  if (!op->setup_done)
    return fail(HC_EINVAL, "hc_op_apply: hc_op_setup has not been called");
  if (nrhs != 1)
    return fail(HC_EINVAL, "hc_op_apply: nrhs > 1 is not supported by this entry point yet");
  HC_TRY_HIP(hc::apply_op<double>(op->g, op->segs, op->tb64, op->qd64, x, y, op->part64, nullptr,
                                  (hipStream_t)stream));
  return HC_OK;
}

int
hc_op_apply_f32(hc_op_t op, const float * x, float * y, int nrhs, void * stream)
{
  if (!op || !x || !y)
    return fail(HC_EINVAL, "hc_op_apply_f32: null argument");
  if (!op->setup_done)
    return fail(HC_EINVAL, "hc_op_apply_f32: hc_op_setup has not been called");
  if (!op->qd32)
    return fail(HC_EINVAL, "hc_op_apply_f32: operator created without HC_PREC_FP32");
  if (nrhs != 1)
    return fail(HC_EINVAL, "hc_op_apply_f32: nrhs > 1 is not supported by this entry point yet");
  HC_TRY_HIP(hc::apply_op<float>(op->g, op->segs, op->tb32, op->qd32, x, y, op->part32, nullptr,
                                 (hipStream_t)stream));
  return HC_OK;
}

int
hc_op_diagonal(hc_op_t op, double * d, void * stream)
{
  if (!op || !d)
    return fail(HC_EINVAL, "hc_op_diagonal: null argument");
  if (!op->setup_done)
    return fail(HC_EINVAL, "hc_op_diagonal: hc_op_setup has not been called");
  HC_TRY_HIP(hipMemcpyAsync(d, op->diag, op->N * sizeof(double), hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return HC_OK;
}

int
hc_pcg_solve(hc_op_t op, const double * b, double * x, int nrhs, double rtol, int maxit, int mode,
             hc_solve_stats * stats, void * stream)
{
  if (!op || !b || !x)
    return fail(HC_EINVAL, "hc_pcg_solve: null argument");
  if (!op->setup_done)
    return fail(HC_EINVAL, "hc_pcg_solve: hc_op_setup has not been called");
  if (nrhs != 1)
    return fail(HC_EINVAL, "hc_pcg_solve: nrhs > 1 goes through the batch interface");
  if (!(rtol > 0) || maxit < 1)
    return fail(HC_EINVAL, "hc_pcg_solve: rtol must be > 0 and maxit >= 1");
  if (mode == HC_SOLVE_FP32_IR && !op->qd32)
    return fail(HC_EINVAL, "hc_pcg_solve: FP32 refinement needs an operator created with HC_PREC_FP32");
  hipStream_t st = (hipStream_t)stream;
  int rc = ensure_solver_work(op, mode == HC_SOLVE_FP32_IR);
  if (rc)
    return rc;
  hipEvent_t e0, e1;
  HC_TRY_HIP(hipEventCreate(&e0));
  HC_TRY_HIP(hipEventCreate(&e1));
  HC_TRY_HIP(hipEventRecord(e0, st));
  hc_solve_stats s{};
  hc::PcgWork w{op->red, op->st_dev};
  hc::PcgWork wo{op->red, op->st_dev + 1};
  hc::PcgState fin{};
  if (mode == HC_SOLVE_FP64)
  {
    // PCG; when the recursive residual meets rtol but the true FP64 residual
    // does not (round-off drift near attainable accuracy), restart from the
    // current x with the true residual (at most 4 restarts).
    int total = 0;
    for (int restart = 0; restart < 5 && total < maxit; restart++)
    {
      rc = run_pcg<double>(op, op->qd64, op->tb64, op->part64, op->dinv64, b, x, op->r64, op->p64, op->q64, false,
                           rtol, maxit - total, st, fin);
      if (rc)
        return rc;
      total += fin.iters;
      s.status = fin.status;
      if (fin.status != 0 || !fin.done)
        break;
      HC_TRY_HIP(hc::apply_op<double>(op->g, op->segs, op->tb64, op->qd64, x, op->q64, op->part64, nullptr, st));
      HC_TRY_HIP(hc::residual64(b, op->q64, op->r64, op->N, wo, true, st));
      hc::PcgState chk{};
      if ((rc = read_state(op, 1, st, chk)))
        return rc;
      if (chk.rr <= rtol * rtol * chk.bb)
        break;
      s.outer_iterations++;
    }
    s.iterations = total;
  }
  else if (mode == HC_SOLVE_FP32_IR)
  {
    // outer residual in FP64
    HC_TRY_HIP(hc::apply_op<double>(op->g, op->segs, op->tb64, op->qd64, x, op->q64, op->part64, nullptr, st));
    HC_TRY_HIP(hc::residual64(b, op->q64, op->r64, op->N, wo, true, st));
    hc::PcgState outer{};
    if ((rc = read_state(op, 1, st, outer)))
      return rc;
    const double bnorm = std::sqrt(outer.bb);
    const int max_outer = 30;
    // inner tolerance: FP32 CG reliably reduces its residual by ~1e-4 to 1e-5
    const double inner_rtol = 3e-5;
    int total = 0, k = 0;
    while (std::sqrt(outer.rr) > rtol * bnorm && k < max_outer && total < maxit && outer.status == 0)
    {
      HC_TRY_HIP(hc::to_f32_scaled(op->r64, op->b32, op->N, op->st_dev + 1, st));
      hipLaunchKernelGGL(k_zero_f, dim3(1024), dim3(256), 0, st, op->x32, op->N);
      const double want = std::max(inner_rtol, 0.5 * rtol * bnorm / std::sqrt(outer.rr));
      rc = run_pcg<float>(op, op->qd32, op->tb32, op->part32, op->dinv32, op->b32, op->x32, op->r32, op->p32,
                          op->q32, true, want, maxit - total, st, fin);
      if (rc)
        return rc;
      total += fin.iters;
      HC_TRY_HIP(hc::ir_update(x, op->x32, op->N, op->st_dev + 1, st));
      HC_TRY_HIP(hc::apply_op<double>(op->g, op->segs, op->tb64, op->qd64, x, op->q64, op->part64, nullptr, st));
      HC_TRY_HIP(hc::residual64(b, op->q64, op->r64, op->N, wo, false, st));
      if ((rc = read_state(op, 1, st, outer)))
        return rc;
      k++;
      if (fin.status != 0)
        break;
    }
    s.iterations = total;
    s.outer_iterations = k;
    s.status = outer.status ? outer.status : fin.status;
  }
  else
    return fail(HC_EINVAL, "hc_pcg_solve: unknown mode");
  // final true residual in FP64
  HC_TRY_HIP(hc::apply_op<double>(op->g, op->segs, op->tb64, op->qd64, x, op->q64, op->part64, nullptr, st));
  HC_TRY_HIP(hc::residual64(b, op->q64, op->r64, op->N, wo, true, st));
  HC_TRY_HIP(hipEventRecord(e1, st));
  hc::PcgState outer{};
  if ((rc = read_state(op, 1, st, outer)))
    return rc;
  float ms = 0;
  HC_TRY_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  s.rel_res = outer.bb > 0 ? std::sqrt(outer.rr / outer.bb) : 0.0;
  s.converged = (s.status == 0 && s.rel_res <= rtol * (1.0 + 1e-6)) ? 1 : 0;
  s.solve_ms = ms;
  if (s.status == 0 && !s.converged)
    s.status = HC_ENOTCONV;
  if (stats)
    *stats = s;
  if (s.status == HC_ENAN)
    return fail(HC_ENAN, "hc_pcg_solve: NaN/Inf met in a reduction");
  if (!s.converged)
  {
    char buf[160];
    std::snprintf(buf, sizeof(buf), "hc_pcg_solve: not converged after %d iterations, true relative residual %.3e",
                  s.iterations, s.rel_res);
    return fail(HC_ENOTCONV, buf);
  }
  return HC_OK;
}

int
hc_op_destroy(hc_op_t op)
{
  if (!op)
    return fail(HC_EINVAL, "hc_op_destroy: null handle");
  (void)hipDeviceSynchronize();
  delete op;
  return HC_OK;
}

int
hc_fill_uniform(hc_op_t op, uint64_t seed, double * v, int64_t n, void * stream)
{
  if (!v || n < 0)
    return fail(HC_EINVAL, "hc_fill_uniform: bad argument");
  if (op && n != op->N)
    return fail(HC_EINVAL, "hc_fill_uniform: length differs from the operator's DOF count");
  HC_TRY_HIP(hc::fill_uniform(op ? &op->g : nullptr, seed, v, n, (hipStream_t)stream));
  return HC_OK;
}

} // extern "C"
