// Solve of the AV formulation's block system (AVEquationSystem,
// src/formulations/AV/av_formulation.cpp:150-196) on the device.
//
// The block operator (hc_op_av_apply) with the weak-divergence block in the
// form the reference states, (sigma u, grad V')
// (src/kernels/vector_fe_weak_divergence_kernel.hpp:8), is symmetric positive
// semi-definite: its energy is dt (curl a, nu curl a) + (a0 + G v, sigma
// (a0 + G v)), zero exactly on the gauge fields (-G phi, phi).  The reference
// solves it with HypreGMRES + BoomerAMG (problem_builder_base.cpp:206-218);
// here it is preconditioned CG with a block-diagonal preconditioner: the
// multigrid V-cycles of hc_mg.hip on the ND block and on the node block
// G^T M G where the mesh has the hierarchies (with exact blocks the
// preconditioned operator has the eigenvalues 0, 1 and 2 only on the free
// DOFs, since K G = 0 makes A G = M G there), else block Jacobi diag(A_AA)^-1 (1 on PEC DOFs) and
// diag(G^T M G)^-1; it converges on consistent right-hand sides.  Reductions:
// per-block partial sums on the device, summed in a fixed order, so a solve
// is bitwise reproducible.
#include "hc_aux.hpp"
#include "hc_kernels.hpp"
#include "hc_op.hpp"

#include <cmath>
#include <initializer_list>
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

// r = b - q, z = d r, p = z; out[b] = block parts of (r, z), out[AV_BLOCKS + b] of (r, r)
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

// Sum of n partials by one workgroup of AV_THREADS in a fixed order.
__device__ inline double
fixed_sum(const double * __restrict__ a, int n, double * sh)
{
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += AV_THREADS)
    s += a[i];
  return av_block_sum(s, sh);
}

} // namespace

// State of a device-driven PCG (in device memory; the host reads it every
// PCG_CHUNK iterations).  Partials: [0, 2 AV_BLOCKS) one dot per vector block,
// [2 AV_BLOCKS, 6 AV_BLOCKS) the (r, z), (r, r) pairs of up to two blocks.
struct PcgDev
{
  double bb, tol2, gam, alpha, beta, rr;
  int it, maxit, done, status;
};
constexpr int PCG_CHUNK = 8;
constexpr int PCG_PARTS = 6 * AV_BLOCKS;

namespace
{

__global__ void __launch_bounds__(AV_THREADS)
k_pcgd_start(const double * __restrict__ part, int nb, double rtol, int maxit, PcgDev * s)
{
  __shared__ double sh[AV_THREADS];
  double bb = 0.0, gam = 0.0, rr = 0.0;
  for (int k = 0; k < nb; k++)
  {
    bb += fixed_sum(part + k * AV_BLOCKS, AV_BLOCKS, sh);
    gam += fixed_sum(part + (2 + 2 * k) * AV_BLOCKS, AV_BLOCKS, sh);
    rr += fixed_sum(part + (3 + 2 * k) * AV_BLOCKS, AV_BLOCKS, sh);
  }
  if (threadIdx.x == 0)
  {
    s->bb = bb;
    s->tol2 = rtol * rtol * bb;
    s->gam = gam;
    s->rr = rr;
    s->alpha = s->beta = 0.0;
    s->it = 0;
    s->maxit = maxit;
    s->status = HC_OK;
    s->done = (bb == 0.0 || rr <= s->tol2) ? 1 : 0;
  }
}

// alpha = gam / (p, q), or stop on a non-positive or non-finite (p, q)
__global__ void __launch_bounds__(AV_THREADS)
k_pcgd_alpha(const double * __restrict__ part, int nb, PcgDev * s)
{
  __shared__ double sh[AV_THREADS];
  if (s->done)
    return;
  double pq = 0.0;
  for (int k = 0; k < nb; k++)
    pq += fixed_sum(part + k * AV_BLOCKS, AV_BLOCKS, sh);
  if (threadIdx.x == 0)
  {
    if (!isfinite(pq) || !(pq > 0.0))
    {
      s->status = isfinite(pq) ? HC_ENOTCONV : HC_ENAN;
      s->done = 1;
    }
    else
      s->alpha = s->gam / pq;
  }
}

// after the update: convergence test on (r, r), beta = (r, z)_new / (r, z)_old
__global__ void __launch_bounds__(AV_THREADS)
k_pcgd_beta(const double * __restrict__ part, int nb, PcgDev * s)
{
  __shared__ double sh[AV_THREADS];
  if (s->done)
    return;
  double gnew = 0.0, rr = 0.0;
  for (int k = 0; k < nb; k++)
  {
    gnew += fixed_sum(part + (2 + 2 * k) * AV_BLOCKS, AV_BLOCKS, sh);
    rr += fixed_sum(part + (3 + 2 * k) * AV_BLOCKS, AV_BLOCKS, sh);
  }
  if (threadIdx.x == 0)
  {
    s->it++;
    s->rr = rr;
    if (!isfinite(gnew) || !isfinite(rr))
    {
      s->status = HC_ENAN;
      s->done = 1;
    }
    else if (rr <= s->tol2)
      s->done = 1;
    else if (s->it >= s->maxit)
    {
      s->status = HC_ENOTCONV;
      s->done = 1;
    }
    else
    {
      s->beta = gnew / s->gam;
      s->gam = gnew;
    }
  }
}

// x += alpha p, r -= alpha q, z = d r with alpha from the state (nothing once done)
__global__ void __launch_bounds__(AV_THREADS)
k_pcgd_update(int64_t n, const PcgDev * __restrict__ s, const double * __restrict__ p, const double * __restrict__ q,
              const double * __restrict__ d, double * __restrict__ x, double * __restrict__ r,
              double * __restrict__ z, double * __restrict__ out)
{
  if (s->done)
    return;
  __shared__ double sh[AV_THREADS];
  const double alpha = s->alpha;
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * AV_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * AV_THREADS)
  {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    if (d)
    {
      const double zi = d[i] * ri;
      z[i] = zi;
      rz += ri * zi;
    }
    rr += ri * ri;
  }
  rz = av_block_sum(rz, sh);
  rr = av_block_sum(rr, sh);
  if (threadIdx.x == 0)
  {
    out[blockIdx.x] = rz; // (without d the caller's (r, z) dot overwrites it)
    out[AV_BLOCKS + blockIdx.x] = rr;
  }
}

// r = b - q; out[AV_BLOCKS + b] = block parts of (r, r)
__global__ void __launch_bounds__(AV_THREADS)
k_pcgd_rstart(int64_t n, const double * __restrict__ b, const double * __restrict__ q, double * __restrict__ r,
              double * __restrict__ out)
{
  __shared__ double sh[AV_THREADS];
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * AV_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * AV_THREADS)
  {
    const double ri = b[i] - q[i];
    r[i] = ri;
    rr += ri * ri;
  }
  rr = av_block_sum(rr, sh);
  if (threadIdx.x == 0)
    out[AV_BLOCKS + blockIdx.x] = rr;
}

// p = z + beta p (nothing once done)
__global__ void
k_pcgd_xpby(int64_t n, const PcgDev * __restrict__ s, const double * __restrict__ z, double * __restrict__ p)
{
  if (s->done)
    return;
  const double beta = s->beta;
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

__global__ void
k_fill_value(int64_t n, double v, double * __restrict__ y)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = v;
}

// Pieces carved out of one device block, 256-byte aligned.
struct Carver
{
  char * p;
  double * take(int64_t n)
  {
    double * v = (double *)p;
    p += align256(sizeof(double) * (size_t)(n > 0 ? n : 1));
    return v;
  }
};

size_t
bytes_of(std::initializer_list<int64_t> counts)
{
  size_t b = 0;
  for (int64_t n : counts)
    b += align256(sizeof(double) * (size_t)(n > 0 ? n : 1));
  return b;
}

} // namespace

// One vector block of a device-driven PCG (nb = 1 or 2 blocks: the AV system
// has an ND block and a node block).
struct PcgBlock
{
  int64_t n;
  const double * dinv;
  const double * b;
  double *x, *r, *z, *p, *q;
};

// Preconditioned CG over nb vector blocks: apply(in0, in1, out0, out1)
// computes out = S in; the preconditioner is diagonal (blk.dinv) or, with
// use_prec, prec(r0, r1, z0, z1) (a fixed symmetric positive definite
// operator, launched without host synchronisation).  Scalars stay on the
// device (fixed-order partial sums, bitwise reproducible); the host reads the
// state every PCG_CHUNK iterations, the kernels of the iterations after the
// stopping point do nothing.  Stops when ||r|| <= rtol ||b|| on the recursive
// residual.  part: PCG_PARTS device doubles; s: device state; hs: pinned
// host copy.
struct NoPrec
{
  int operator()(const double *, const double *, double *, double *) const { return HC_OK; }
};

template <typename Apply, typename Prec = NoPrec>
static int
pcg_dev(int nb, const PcgBlock * bl, Apply && apply, double rtol, int maxit, double * part, PcgDev * s,
        PcgDev * hs, hipStream_t st, int & iters, bool & conv, int & status, bool use_prec = false,
        Prec && prec = Prec{})
{
  const dim3 G(AV_BLOCKS), T(AV_THREADS);
  auto precond = [&]() -> int {
    if (int rc = prec(bl[0].r, nb > 1 ? bl[1].r : nullptr, bl[0].z, nb > 1 ? bl[1].z : nullptr))
      return rc;
    for (int k = 0; k < nb; k++)
      hipLaunchKernelGGL(k_av_dot, G, T, 0, st, bl[k].n, (const double *)bl[k].r, (const double *)bl[k].z,
                         part + (2 + 2 * k) * AV_BLOCKS);
    return HC_OK;
  };
  if (int rc = apply(bl[0].x, nb > 1 ? bl[1].x : nullptr, bl[0].q, nb > 1 ? bl[1].q : nullptr))
    return rc;
  for (int k = 0; k < nb; k++)
  {
    hipLaunchKernelGGL(k_av_dot, G, T, 0, st, bl[k].n, bl[k].b, bl[k].b, part + k * AV_BLOCKS);
    if (use_prec)
      hipLaunchKernelGGL(k_pcgd_rstart, G, T, 0, st, bl[k].n, bl[k].b, (const double *)bl[k].q, bl[k].r,
                         part + (2 + 2 * k) * AV_BLOCKS);
    else
      hipLaunchKernelGGL(k_av_start, G, T, 0, st, bl[k].n, bl[k].b, (const double *)bl[k].q, bl[k].dinv, bl[k].r,
                         bl[k].z, bl[k].p, part + (2 + 2 * k) * AV_BLOCKS);
  }
  if (use_prec)
  {
    if (int rc = precond())
      return rc;
    for (int k = 0; k < nb; k++)
      HC_TRY_HIP(hipMemcpyAsync(bl[k].p, bl[k].z, sizeof(double) * bl[k].n, hipMemcpyDeviceToDevice, st));
  }
  hipLaunchKernelGGL(k_pcgd_start, dim3(1), T, 0, st, (const double *)part, nb, rtol, maxit, s);
  HC_TRY_HIP(hipGetLastError());
  for (;;)
  {
    HC_TRY_HIP(hipMemcpyAsync(hs, s, sizeof(PcgDev), hipMemcpyDeviceToHost, st));
    HC_TRY_HIP(hipStreamSynchronize(st));
    if (hs->done)
      break;
    for (int c = 0; c < PCG_CHUNK; c++)
    {
      if (int rc = apply(bl[0].p, nb > 1 ? bl[1].p : nullptr, bl[0].q, nb > 1 ? bl[1].q : nullptr))
        return rc;
      for (int k = 0; k < nb; k++)
        hipLaunchKernelGGL(k_av_dot, G, T, 0, st, bl[k].n, (const double *)bl[k].p, (const double *)bl[k].q,
                           part + k * AV_BLOCKS);
      hipLaunchKernelGGL(k_pcgd_alpha, dim3(1), T, 0, st, (const double *)part, nb, s);
      for (int k = 0; k < nb; k++)
        hipLaunchKernelGGL(k_pcgd_update, G, T, 0, st, bl[k].n, (const PcgDev *)s, (const double *)bl[k].p,
                           (const double *)bl[k].q, use_prec ? nullptr : bl[k].dinv, bl[k].x, bl[k].r, bl[k].z,
                           part + (2 + 2 * k) * AV_BLOCKS);
      if (use_prec)
        if (int rc = precond())
          return rc;
      hipLaunchKernelGGL(k_pcgd_beta, dim3(1), T, 0, st, (const double *)part, nb, s);
      for (int k = 0; k < nb; k++)
        hipLaunchKernelGGL(k_pcgd_xpby, dim3(2048), dim3(256), 0, st, bl[k].n, (const PcgDev *)s,
                           (const double *)bl[k].z, bl[k].p);
      HC_TRY_HIP(hipGetLastError());
    }
  }
  iters = hs->it;
  status = hs->status;
  conv = status == HC_OK;
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
  const int64_t NQ = (int64_t)12 * E * Qp * Qp * Qp;
  // work vectors, kept with the operator: unit-mass factors and the two
  // diagonals (mesh only, computed once), the ND and node PCG vectors
  if (!op->dfs_mem.base)
  {
    const size_t need = hc::bytes_of({NQ, E, E, N, N, N, N, N, N, NN, NN, NN, NN, NN, NN, NN, hc::PCG_PARTS}) +
                        hc::align256(sizeof(hc::PcgDev));
    if ((rc = op->dfs_mem.alloc(need)))
      return rc;
    op->dfs_ready = false;
  }
  if (!op->av_host && hipHostMalloc((void **)&op->av_host, sizeof(hc::PcgDev)) != hipSuccess)
  {
    op->av_host = nullptr;
    return hc::fail(HC_ENOMEM, "hc_op_div_free_source: pinned host allocation failed");
  }
  hc::Carver cv{(char *)op->dfs_mem.base};
  double *qdu = cv.take(NQ), *c0 = cv.take(E), *c1 = cv.take(E), *dM = cv.take(N);
  double *r = cv.take(N), *z = cv.take(N), *p = cv.take(N), *q = cv.take(N), *t = cv.take(N);
  double *rn = cv.take(NN), *zn = cv.take(NN), *pn = cv.take(NN), *qn = cv.take(NN), *dS = cv.take(NN),
         *Qv = cv.take(NN), *bn = cv.take(NN);
  double * part = cv.take(hc::PCG_PARTS);
  hc::PcgDev * sdev = (hc::PcgDev *)cv.p;
  hc::PcgDev * shost = (hc::PcgDev *)op->av_host;
  // the projection's node Laplacian G^T M G: preconditioned by the node-space
  // V-cycle of a unit-mass operator on the same box mesh (built once, outside the timed region), else Jacobi
  if (project && op->box && op->dfs_mg == 0)
  {
    op->dfs_mg = -1;
    hc_mesh_desc md{};
    md.nx = op->g.n[0];
    md.ny = op->g.n[1];
    md.nz = op->g.n[2];
    md.vert = nullptr;
    for (int a = 0; a < 3; a++)
    {
      md.x0[a] = op->x0[a];
      md.len[a] = op->h[a] * op->g.n[a];
    }
    md.pec_faces = 0;
    std::vector<double> zero(E, 0.0), one(E, 1.0);
    if (!op->dfs_op && hc_op_create(&md, op->p, zero.data(), one.data(), HC_PREC_FP64, &op->dfs_op) != HC_OK)
      op->dfs_op = nullptr;
    if (op->dfs_op && hc_op_setup(op->dfs_op, 1.0, stream) == HC_OK && hc::mg_av_ready(op->dfs_op, st))
      op->dfs_mg = 1;
  }
  const bool node_mg = project && op->dfs_mg == 1;
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
  if (!op->dfs_ready)
  {
    // unit-coefficient mass factors on the operator's geometry, their
    // diagonal and the node diagonal of G^T M G
    hipLaunchKernelGGL(hc::k_fill_value, dim3(256), dim3(256), 0, st, E, 0.0, c0);
    hipLaunchKernelGGL(hc::k_fill_value, dim3(256), dim3(256), 0, st, E, 1.0, c1);
    HC_TRY_HIP(hc::pa_setup(op->g, op->tb64, op->d_vert, op->h, c0, c1, 1.0, qdu, st));
    HC_TRY_HIP(hc::diagonal(g0, op->tb64, qdu, dM, st));
    HC_TRY_HIP(hc::invert_diag(dM, dM, nullptr, N, st));
    HC_TRY_HIP(hc::aux_diag_inv(g0, op->ag, op->tb64, qdu, dS, dS, nullptr, st));
    op->dfs_ready = true;
  }
  auto mass = [&](const double * x, double * y) -> int {
    HC_TRY_HIP(hc::apply_scaled<double>(g0, pl.segs, op->tb64, qdu, x, y, op->ws.part64, 1.0, 0.0, pl.cand,
                                        op->desc_dev, st));
    return HC_OK;
  };
  int it1 = 0, it2 = 0, s1 = HC_OK, s2 = HC_OK;
  bool c1ok = false, c2ok = true;
  // g = M^-1 j
  hipLaunchKernelGGL(hc::k_fill_value, dim3(2048), dim3(256), 0, st, N, 0.0, src);
  {
    const hc::PcgBlock blk{N, dM, j, src, r, z, p, q};
    auto ap = [&](const double * x0, const double *, double * y0, double *) -> int { return mass(x0, y0); };
    if ((rc = hc::pcg_dev(1, &blk, ap, rtol, maxit, part, sdev, shost, st, it1, c1ok, s1)))
      return rc;
  }
  hipLaunchKernelGGL(hc::k_fill_value, dim3(2048), dim3(256), 0, st, NN, 0.0, Qv);
  if (project && s1 == HC_OK)
  {
    // b_n = G^T M g; S = G^T M G on the nodes, diag(S)^-1 from the unit mass factors
    if ((rc = mass(src, t)))
      return rc;
    HC_TRY_HIP(hc::grad_transpose(op->ag, t, bn, st));
    auto stiff = [&](const double * x0, const double *, double * y0, double *) -> int {
      HC_TRY_HIP(hc::grad_apply(op->ag, x0, t, st));
      int e = mass(t, q);
      if (e)
        return e;
      HC_TRY_HIP(hc::grad_transpose(op->ag, q, y0, st));
      return HC_OK;
    };
    // the stiffness apply uses t and q: the node PCG's own vectors are rn..qn
    const hc::PcgBlock blk{NN, dS, bn, Qv, rn, zn, pn, qn};
    auto npr = [&](const double * r0, const double *, double * z0, double *) -> int {
      return hc::mg_node_precond(op->dfs_op, r0, z0, st);
    };
    if ((rc = hc::pcg_dev(1, &blk, stiff, rtol, maxit, part, sdev, shost, st, it2, c2ok, s2, node_mg, npr)))
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
  int rc;
  if (!op->av_mem.base &&
      (rc = op->av_mem.alloc(hc::bytes_of({N, N, N, N, NN, NN, NN, NN, NN, hc::PCG_PARTS}) +
                             hc::align256(sizeof(hc::PcgDev)))))
    return rc;
  if (!op->av_host && hipHostMalloc((void **)&op->av_host, sizeof(hc::PcgDev)) != hipSuccess)
  {
    op->av_host = nullptr;
    return hc::fail(HC_ENOMEM, "hc_op_av_solve: pinned host allocation failed");
  }
  hc::Carver cv{(char *)op->av_mem.base};
  double *ra = cv.take(N), *za = cv.take(N), *pa = cv.take(N), *qa = cv.take(N);
  double *rv = cv.take(NN), *zv = cv.take(NN), *pv = cv.take(NN), *qv = cv.take(NN), *vinv = cv.take(NN);
  double * part = cv.take(hc::PCG_PARTS);
  hc::PcgDev * sdev = (hc::PcgDev *)cv.p;
  hc::PcgDev * shost = (hc::PcgDev *)op->av_host;
  // block preconditioner: multigrid on both blocks where the mesh has the
  // hierarchies (built on first use, outside the timed region), else block Jacobi
  const bool mg = hc::mg_av_ready(op, st);
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
  const hc::PcgBlock blk[2] = {{N, op->dinv64, ba, a, ra, za, pa, qa}, {NN, vinv, bv, v, rv, zv, pv, qv}};
  auto ap = [&](const double * x0, const double * x1, double * y0, double * y1) -> int {
    return hc_op_av_apply(op, x0, x1, y0, y1, stream);
  };
  auto pr = [&](const double * r0, const double * r1, double * z0, double * z1) -> int {
    return hc::mg_av_precond(op, r0, r1, z0, z1, st);
  };
  int it = 0, status = HC_OK;
  bool conv = false;
  if ((rc = hc::pcg_dev(2, blk, ap, rtol, maxit, part, sdev, shost, st, it, conv, status, mg, pr)))
    return rc;
  // true residual of the final iterate
  if ((rc = hc_op_av_apply(op, a, v, qa, qv, stream)))
    return rc;
  const dim3 G(AV_BLOCKS), T(AV_THREADS);
  hipLaunchKernelGGL(hc::k_av_resid, G, T, 0, st, N, ba, qa, part);
  hipLaunchKernelGGL(hc::k_av_resid, G, T, 0, st, NN, bv, qv, part + AV_BLOCKS);
  (void)hipEventRecord(e1, st);
  std::vector<double> h(2 * AV_BLOCKS);
  HC_TRY_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * 2 * AV_BLOCKS, hipMemcpyDeviceToHost, st));
  HC_TRY_HIP(hipStreamSynchronize(st));
  double tr = 0.0;
  for (double x : h)
    tr += x;
  const double bb = shost->bb;
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
    stats->outer_iterations = mg ? 2 : 1; // 2: multigrid block preconditioner, 1: block Jacobi
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
