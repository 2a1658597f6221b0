// Solve of the AV formulation's block system (AVEquationSystem,
// src/formulations/AV/av_formulation.cpp:150-196) on the device.
//
// The block operator (hc_op_av_apply) with the weak-divergence block in the
// form the reference states, (sigma u, grad V')
// (src/kernels/vector_fe_weak_divergence_kernel.hpp:8), is symmetric positive
// semi-definite: its energy is dt (curl a, nu curl a) + (a0 + G v, sigma
// (a0 + G v)), zero exactly on the gauge fields (-G phi, phi).  The reference
// solves it with HypreGMRES + BoomerAMG (problem_builder_base.cpp:206-218);
// here it is preconditioned CG with the block-Jacobi preconditioner
// diag(A_AA)^-1 (1 on PEC DOFs) and diag(G^T M G)^-1 on the nodes, which
// converges on consistent right-hand sides (the gauge component of the
// iterate is left as the initial guess has it).  Reductions: per-block
// partial sums on the device, summed on the host in a fixed order, so a solve
// is bitwise reproducible.
#include "hc_aux.hpp"
#include "hc_kernels.hpp"
#include "hc_op.hpp"

#include <cmath>
#include <vector>

namespace hc
{

namespace
{

constexpr int AV_THREADS = 256;
constexpr int AV_BLOCKS = 1024;

__device__ inline double
av_block_sum(double v, double * sh)
{
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = AV_THREADS / 2; s > 0; s >>= 1)
  {
    if (t < s)
      sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads(); // sh is reused by the caller's next reduction
  return r;
}

// out[b] = block b's part of x.y
__global__ void __launch_bounds__(AV_THREADS)
k_av_dot(int64_t n, const double * __restrict__ x, const double * __restrict__ y, double * __restrict__ out)
{
  __shared__ double sh[AV_THREADS];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * AV_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * AV_THREADS)
    s += x[i] * y[i];
  s = av_block_sum(s, sh);
  if (threadIdx.x == 0)
    out[blockIdx.x] = s;
}

// x += alpha p, r -= alpha q, z = d r; out[b] = block parts of (r, z), out[AV_BLOCKS + b] of (r, r)
__global__ void __launch_bounds__(AV_THREADS)
k_av_update(int64_t n, double alpha, const double * __restrict__ p, const double * __restrict__ q,
            const double * __restrict__ d, double * __restrict__ x, double * __restrict__ r, double * __restrict__ z,
            double * __restrict__ out)
{
  __shared__ double sh[AV_THREADS];
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * AV_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * AV_THREADS)
  {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    const double zi = d[i] * ri;
    r[i] = ri;
    z[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
  rz = av_block_sum(rz, sh);
  rr = av_block_sum(rr, sh);
  if (threadIdx.x == 0)
  {
    out[blockIdx.x] = rz;
    out[AV_BLOCKS + blockIdx.x] = rr;
  }
}

// r = b - q, z = d r, p = z; out as k_av_update
__global__ void __launch_bounds__(AV_THREADS)
k_av_start(int64_t n, const double * __restrict__ b, const double * __restrict__ q, const double * __restrict__ d,
           double * __restrict__ r, double * __restrict__ z, double * __restrict__ p, double * __restrict__ out)
{
  __shared__ double sh[AV_THREADS];
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * AV_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * AV_THREADS)
  {
    const double ri = b[i] - q[i];
    const double zi = d[i] * ri;
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
  rz = av_block_sum(rz, sh);
  rr = av_block_sum(rr, sh);
  if (threadIdx.x == 0)
  {
    out[blockIdx.x] = rz;
    out[AV_BLOCKS + blockIdx.x] = rr;
  }
}

// p = z + beta p
__global__ void
k_av_xpby(int64_t n, const double * __restrict__ z, double beta, double * __restrict__ p)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

// r = b - q; out[b] = block parts of (r, r)
__global__ void __launch_bounds__(AV_THREADS)
k_av_resid(int64_t n, const double * __restrict__ b, const double * __restrict__ q, double * __restrict__ out)
{
  __shared__ double sh[AV_THREADS];
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * AV_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * AV_THREADS)
  {
    const double ri = b[i] - q[i];
    rr += ri * ri;
  }
  rr = av_block_sum(rr, sh);
  if (threadIdx.x == 0)
    out[blockIdx.x] = rr;
}

// Device buffers of one solve, released on every path out.
struct AvBuffers
{
  std::vector<void *> p;
  ~AvBuffers()
  {
    for (void * q : p)
      (void)hipFree(q);
  }
  double * get(size_t n)
  {
    void * q = nullptr;
    if (hipMalloc(&q, n * sizeof(double)) != hipSuccess)
      return nullptr;
    p.push_back(q);
    return (double *)q;
  }
};

double
host_sum(const double * v, int n)
{
  double s = 0.0;
  for (int i = 0; i < n; i++)
    s += v[i];
  return s;
}

} // namespace

// Copy nsums arrays of AV_BLOCKS block partials to the host (stream-ordered).
static int
read_partials(double * dev, double * host, int nsums, hipStream_t st)
{
  if (hipMemcpyAsync(host, dev, sizeof(double) * AV_BLOCKS * nsums, hipMemcpyDeviceToHost, st) != hipSuccess)
    return HC_EHIP;
  if (hipStreamSynchronize(st) != hipSuccess)
    return HC_EHIP;
  return HC_OK;
}

__global__ void
k_fill_value(int64_t n, double v, double * __restrict__ y)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = v;
}

// Preconditioned CG on one vector of n entries: apply(p, q) computes q = S p,
// dinv the diagonal preconditioner; x holds the initial guess.  Stops when
// ||r|| <= rtol ||b|| (recursive residual).  r, z, p, q: n-entry work vectors;
// part: 2 * AV_BLOCKS device doubles; h: as many host doubles.
template <typename Apply>
static int
pcg_one(int64_t n, Apply && apply, const double * dinv, const double * b, double * x, double rtol, int maxit,
        double * r, double * z, double * p, double * q, double * part, double * h, hipStream_t st, int & iters,
        bool & conv, int & status)
{
  const dim3 G(AV_BLOCKS), T(AV_THREADS);
  iters = 0;
  status = HC_OK;
  int rc = apply(x, q);
  if (rc)
    return rc;
  hipLaunchKernelGGL(k_av_dot, G, T, 0, st, n, b, b, part);
  if ((rc = read_partials(part, h, 1, st)))
    return fail(rc, "PCG: reduction failed");
  const double bb = host_sum(h, AV_BLOCKS);
  hipLaunchKernelGGL(k_av_start, G, T, 0, st, n, b, (const double *)q, dinv, r, z, p, part);
  if ((rc = read_partials(part, h, 2, st)))
    return fail(rc, "PCG: reduction failed");
  double gam = host_sum(h, AV_BLOCKS), rr = host_sum(h + AV_BLOCKS, AV_BLOCKS);
  const double tol2 = rtol * rtol * bb;
  conv = bb == 0.0 || rr <= tol2;
  while (!conv && iters < maxit)
  {
    if ((rc = apply(p, q)))
      return rc;
    hipLaunchKernelGGL(k_av_dot, G, T, 0, st, n, (const double *)p, (const double *)q, part);
    if ((rc = read_partials(part, h, 1, st)))
      return fail(rc, "PCG: reduction failed");
    const double pq = host_sum(h, AV_BLOCKS);
    if (!std::isfinite(pq) || !(pq > 0.0))
    {
      status = std::isfinite(pq) ? HC_ENOTCONV : HC_ENAN;
      break;
    }
    hipLaunchKernelGGL(k_av_update, G, T, 0, st, n, gam / pq, (const double *)p, (const double *)q, dinv, x, r, z,
                       part);
    if ((rc = read_partials(part, h, 2, st)))
      return fail(rc, "PCG: reduction failed");
    const double gnew = host_sum(h, AV_BLOCKS);
    rr = host_sum(h + AV_BLOCKS, AV_BLOCKS);
    iters++;
    if (!std::isfinite(gnew) || !std::isfinite(rr))
    {
      status = HC_ENAN;
      break;
    }
    conv = rr <= tol2;
    if (conv)
      break;
    hipLaunchKernelGGL(k_av_xpby, dim3(2048), dim3(256), 0, st, n, (const double *)z, gnew / gam, p);
    gam = gnew;
  }
  if (status == HC_OK && !conv)
    status = HC_ENOTCONV;
  return HC_OK;
}

} // namespace hc

using hc::AV_BLOCKS;
using hc::AV_THREADS;

// DivFreeSource::Apply (src/sources/div_free_source.cpp:71-114) for the
// source linear form j = (J, phi_i):
//   g = M^-1 j                 (H(curl) mass with unit coefficient, no
//                               essential DOFs: div_free_source.cpp:76-89)
//   g <- g - G Q,  (G^T M G) Q = G^T M g   (the Helmholtz projection
//                               P(g) = g - grad Q, (grad.P(g), q) = 0 for all q
//                               in H1, div_free_source.cpp:7-17,93-106, no
//                               boundary conditions)
//   lf += M g                  (div_free_source.cpp:108-113)
// Both solves are Jacobi-preconditioned CG to rtol (the reference uses its
// default GMRES, hephaestus_solvers.hpp); the projection's singular system
// (constants) has a consistent right-hand side (G 1 = 0).  src receives the
// projected source field g, Q (optional) the potential.
extern "C" int
hc_op_div_free_source(hc_op_t op, const double * j, double * src, double * lf, double * Q, int project,
                      double rtol, int maxit, hc_solve_stats * stats, void * stream)
{
  if (!op || !j || !src || !lf)
    return hc::fail(HC_EINVAL, "hc_op_div_free_source: null argument");
  if (!(rtol > 0.0) || maxit < 1)
    return hc::fail(HC_EINVAL, "hc_op_div_free_source: rtol must be > 0 and maxit >= 1");
  int rc = op->ensure_ws(1, false);
  if (rc)
    return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = op->N, NN = op->ag.NN, E = op->E;
  const int Qp = op->p + 1;
  hc::AvBuffers buf;
  double * qdu = buf.get((size_t)12 * E * Qp * Qp * Qp);
  double *c0 = buf.get(E), *c1 = buf.get(E), *dM = buf.get(N);
  double *r = buf.get(N), *z = buf.get(N), *p = buf.get(N), *q = buf.get(N), *t = buf.get(N);
  double *rn = buf.get(NN), *zn = buf.get(NN), *pn = buf.get(NN), *qn = buf.get(NN), *dS = buf.get(NN),
         *Qv = buf.get(NN), *bn = buf.get(NN);
  double * part = buf.get(2 * AV_BLOCKS);
  if (!qdu || !c0 || !c1 || !dM || !r || !z || !p || !q || !t || !rn || !zn || !pn || !qn || !dS || !Qv || !bn ||
      !part)
    return hc::fail(HC_ENOMEM, "hc_op_div_free_source: device allocation failed");
  std::vector<double> h(2 * AV_BLOCKS);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return hc::fail(HC_EHIP, "hc_op_div_free_source: event creation failed");
  struct Ev
  {
    hipEvent_t a, b;
    ~Ev()
    {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } ev{e0, e1};
  (void)hipEventRecord(e0, st);
  const hc::ApplyPlan & pl = op->plan64;
  hc::Geom g0 = pl.g;
  g0.pec = 0;
  // unit-coefficient mass factors on the operator's geometry
  hipLaunchKernelGGL(hc::k_fill_value, dim3(256), dim3(256), 0, st, E, 0.0, c0);
  hipLaunchKernelGGL(hc::k_fill_value, dim3(256), dim3(256), 0, st, E, 1.0, c1);
  HC_TRY_HIP(hc::pa_setup(op->g, op->tb64, op->d_vert, op->h, c0, c1, 1.0, qdu, st));
  HC_TRY_HIP(hc::diagonal(g0, op->tb64, qdu, dM, st));
  HC_TRY_HIP(hc::invert_diag(dM, dM, nullptr, N, st));
  auto mass = [&](const double * x, double * y) -> int {
    HC_TRY_HIP(hc::apply_scaled<double>(g0, pl.segs, op->tb64, qdu, x, y, op->ws.part64, 1.0, 0.0, pl.cand,
                                        op->desc_dev, st));
    return HC_OK;
  };
  int it1 = 0, it2 = 0, s1 = HC_OK, s2 = HC_OK;
  bool c1ok = false, c2ok = true;
  // g = M^-1 j
  hipLaunchKernelGGL(hc::k_fill_value, dim3(2048), dim3(256), 0, st, N, 0.0, src);
  if ((rc = hc::pcg_one(N, mass, dM, j, src, rtol, maxit, r, z, p, q, part, h.data(), st, it1, c1ok, s1)))
    return rc;
  hipLaunchKernelGGL(hc::k_fill_value, dim3(2048), dim3(256), 0, st, NN, 0.0, Qv);
  if (project && s1 == HC_OK)
  {
    // b_n = G^T M g; S = G^T M G on the nodes, diag(S)^-1 from the unit mass factors
    if ((rc = mass(src, t)))
      return rc;
    HC_TRY_HIP(hc::grad_transpose(op->ag, t, bn, st));
    HC_TRY_HIP(hc::aux_diag_inv(g0, op->ag, op->tb64, qdu, dS, dS, nullptr, st));
    auto stiff = [&](const double * x, double * y) -> int {
      HC_TRY_HIP(hc::grad_apply(op->ag, x, t, st));
      int e = mass(t, q);
      if (e)
        return e;
      HC_TRY_HIP(hc::grad_transpose(op->ag, q, y, st));
      return HC_OK;
    };
    // the stiffness apply uses t and q: the node PCG's own vectors are rn..qn
    if ((rc = hc::pcg_one(NN, stiff, dS, bn, Qv, rtol, maxit, rn, zn, pn, qn, part, h.data(), st, it2, c2ok, s2)))
      return rc;
    // g <- g - G Q
    HC_TRY_HIP(hc::grad_apply(op->ag, Qv, t, st));
    HC_TRY_HIP(hc::axpby(N, -1.0, t, 1.0, src, st));
  }
  // lf += M g (left untouched when a solve failed)
  if (s1 == HC_OK && s2 == HC_OK)
  {
    if ((rc = mass(src, t)))
      return rc;
    HC_TRY_HIP(hc::axpby(N, 1.0, t, 1.0, lf, st));
  }
  if (Q)
    HC_TRY_HIP(hipMemcpyAsync(Q, Qv, sizeof(double) * NN, hipMemcpyDeviceToDevice, st));
  (void)hipEventRecord(e1, st);
  HC_TRY_HIP(hipStreamSynchronize(st));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  const int status = s1 != HC_OK ? s1 : s2;
  if (stats)
  {
    *stats = hc_solve_stats{};
    stats->iterations = it1 + it2;
    stats->outer_iterations = project ? 2 : 1;
    stats->converged = (c1ok && c2ok && status == HC_OK) ? 1 : 0;
    stats->status = status;
    stats->solve_ms = ms;
  }
  if (status != HC_OK)
    return hc::fail(status, status == HC_ENAN ? "hc_op_div_free_source: NaN/Inf in a reduction"
                                              : "hc_op_div_free_source: tolerance not reached in maxit");
  return HC_OK;
}

extern "C" int
hc_op_av_solve(hc_op_t op, const double * ba, const double * bv, double * a, double * v, double rtol, int maxit,
               hc_solve_stats * stats, void * stream)
{
  if (!op || !ba || !bv || !a || !v)
    return hc::fail(HC_EINVAL, "hc_op_av_solve: null argument");
  if (!op->setup_done)
    return hc::fail(HC_EINVAL, "hc_op_av_solve: hc_op_setup has not been called");
  if (!(rtol > 0.0) || maxit < 1)
    return hc::fail(HC_EINVAL, "hc_op_av_solve: rtol must be > 0 and maxit >= 1");
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = op->N, NN = op->ag.NN;
  hc::AvBuffers buf;
  double *ra = buf.get(N), *za = buf.get(N), *pa = buf.get(N), *qa = buf.get(N);
  double *rv = buf.get(NN), *zv = buf.get(NN), *pv = buf.get(NN), *qv = buf.get(NN), *vinv = buf.get(NN);
  double * part = buf.get(4 * AV_BLOCKS);
  if (!ra || !za || !pa || !qa || !rv || !zv || !pv || !qv || !vinv || !part)
    return hc::fail(HC_ENOMEM, "hc_op_av_solve: device allocation failed");
  std::vector<double> h(4 * AV_BLOCKS);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return hc::fail(HC_EHIP, "hc_op_av_solve: event creation failed");
  struct Ev
  {
    hipEvent_t a, b;
    ~Ev()
    {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } ev{e0, e1};
  (void)hipEventRecord(e0, st);
  // node preconditioner: diag(G^T M G)^-1 on every node (V has no essential nodes)
  hc::Geom g0 = op->g;
  g0.pec = 0;
  if (hc::aux_diag_inv(g0, op->ag, op->tb64, op->qd64, vinv, vinv, nullptr, st) != hipSuccess)
    return hc::fail(HC_EHIP, "hc_op_av_solve: node diagonal failed");
  const dim3 G(AV_BLOCKS), T(AV_THREADS);
  // r = b - AV x, z = D^-1 r, p = z
  int rc = hc_op_av_apply(op, a, v, qa, qv, stream);
  if (rc)
    return rc;
  hipLaunchKernelGGL(hc::k_av_dot, G, T, 0, st, N, ba, ba, part);
  hipLaunchKernelGGL(hc::k_av_dot, G, T, 0, st, NN, bv, bv, part + AV_BLOCKS);
  if ((rc = hc::read_partials(part, h.data(), 2, st)))
    return hc::fail(rc, "hc_op_av_solve: reduction failed");
  const double bb = hc::host_sum(h.data(), AV_BLOCKS) + hc::host_sum(h.data() + AV_BLOCKS, AV_BLOCKS);
  hipLaunchKernelGGL(hc::k_av_start, G, T, 0, st, N, ba, qa, op->dinv64, ra, za, pa, part);
  hipLaunchKernelGGL(hc::k_av_start, G, T, 0, st, NN, bv, qv, vinv, rv, zv, pv, part + 2 * AV_BLOCKS);
  if ((rc = hc::read_partials(part, h.data(), 4, st)))
    return hc::fail(rc, "hc_op_av_solve: reduction failed");
  double gam = hc::host_sum(h.data(), AV_BLOCKS) + hc::host_sum(h.data() + 2 * AV_BLOCKS, AV_BLOCKS);
  double rr = hc::host_sum(h.data() + AV_BLOCKS, AV_BLOCKS) + hc::host_sum(h.data() + 3 * AV_BLOCKS, AV_BLOCKS);
  const double tol2 = rtol * rtol * bb;
  int it = 0, status = HC_OK;
  bool conv = bb == 0.0 || rr <= tol2;
  while (!conv && it < maxit)
  {
    if ((rc = hc_op_av_apply(op, pa, pv, qa, qv, stream)))
      return rc;
    hipLaunchKernelGGL(hc::k_av_dot, G, T, 0, st, N, pa, qa, part);
    hipLaunchKernelGGL(hc::k_av_dot, G, T, 0, st, NN, pv, qv, part + AV_BLOCKS);
    if ((rc = hc::read_partials(part, h.data(), 2, st)))
      return hc::fail(rc, "hc_op_av_solve: reduction failed");
    const double pq = hc::host_sum(h.data(), AV_BLOCKS) + hc::host_sum(h.data() + AV_BLOCKS, AV_BLOCKS);
    if (!std::isfinite(pq) || !(pq > 0.0))
    {
      status = std::isfinite(pq) ? HC_ENOTCONV : HC_ENAN;
      break;
    }
    const double alpha = gam / pq;
    hipLaunchKernelGGL(hc::k_av_update, G, T, 0, st, N, alpha, (const double *)pa, (const double *)qa,
                       (const double *)op->dinv64, a, ra, za, part);
    hipLaunchKernelGGL(hc::k_av_update, G, T, 0, st, NN, alpha, (const double *)pv, (const double *)qv,
                       (const double *)vinv, v, rv, zv, part + 2 * AV_BLOCKS);
    if ((rc = hc::read_partials(part, h.data(), 4, st)))
      return hc::fail(rc, "hc_op_av_solve: reduction failed");
    const double gnew = hc::host_sum(h.data(), AV_BLOCKS) + hc::host_sum(h.data() + 2 * AV_BLOCKS, AV_BLOCKS);
    rr = hc::host_sum(h.data() + AV_BLOCKS, AV_BLOCKS) + hc::host_sum(h.data() + 3 * AV_BLOCKS, AV_BLOCKS);
    it++;
    if (!std::isfinite(gnew) || !std::isfinite(rr))
    {
      status = HC_ENAN;
      break;
    }
    conv = rr <= tol2;
    if (conv)
      break;
    hipLaunchKernelGGL(hc::k_av_xpby, dim3(2048), dim3(256), 0, st, N, (const double *)za, gnew / gam, pa);
    hipLaunchKernelGGL(hc::k_av_xpby, dim3(2048), dim3(256), 0, st, NN, (const double *)zv, gnew / gam, pv);
    gam = gnew;
  }
  // true residual of the final iterate
  if ((rc = hc_op_av_apply(op, a, v, qa, qv, stream)))
    return rc;
  hipLaunchKernelGGL(hc::k_av_resid, G, T, 0, st, N, ba, qa, part);
  hipLaunchKernelGGL(hc::k_av_resid, G, T, 0, st, NN, bv, qv, part + AV_BLOCKS);
  (void)hipEventRecord(e1, st);
  if ((rc = hc::read_partials(part, h.data(), 2, st)))
    return hc::fail(rc, "hc_op_av_solve: reduction failed");
  const double tr = hc::host_sum(h.data(), AV_BLOCKS) + hc::host_sum(h.data() + AV_BLOCKS, AV_BLOCKS);
  const double rel = bb > 0.0 ? std::sqrt(tr / bb) : 0.0;
  conv = conv && status == HC_OK && rel <= rtol * (1.0 + 1e-6);
  if (status == HC_OK && !conv)
    status = HC_ENOTCONV;
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (stats)
  {
    *stats = hc_solve_stats{};
    stats->iterations = it;
    stats->outer_iterations = 1;
    stats->converged = conv ? 1 : 0;
    stats->status = status;
    stats->rel_res = rel;
    stats->solve_ms = ms;
  }
  if (status != HC_OK)
    return hc::fail(status, status == HC_ENAN ? "hc_op_av_solve: NaN/Inf in a reduction"
                                                     : "hc_op_av_solve: tolerance not reached in maxit");
  return HC_OK;
}
