// Device-driven preconditioned CG over one or two vector blocks: scalars
// stay in device memory (fixed-order partial sums, bitwise reproducible),
// the host reads the solver state after a number of iterations predicted
// from the residual's rate (at most PCG_CHUNK).  Used by the AV block solve
// and DivFreeSource (hc_av.hip).  The kernels have internal linkage: each
// including file gets its own copies.
#pragma once

#include "hc_op.hpp"

#include <algorithm>
#include <cmath>

namespace hc
{

namespace
{

constexpr int AV_THREADS = 256;
constexpr int AV_BLOCKS = 1024;

__device__ inline double
av_block_sum(double v, double * sh)
{
  return group_sum<AV_THREADS>(v, sh); // sh is free again on return
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
// few iterations, at most PCG_CHUNK).  Partials: [0, 2 AV_BLOCKS) one dot per vector block,
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
// state every few iterations (at most PCG_CHUNK), the vector kernels of the
// iterations after the stopping point do nothing.  Stops when ||r|| <= rtol ||b|| on the recursive
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
  // iterations per host check: predicted from the observed rate of the
  // residual so that few iterations run past the stopping point (their
  // kernels do nothing, but the operator and preconditioner launches still run)
  int chunk = 1;
  double rr_prev = -1.0;
  int it_prev = 0;
  for (;;)
  {
    HC_TRY_HIP(hipMemcpyAsync(hs, s, sizeof(PcgDev), hipMemcpyDeviceToHost, st));
    HC_TRY_HIP(hipStreamSynchronize(st));
    if (hs->done)
      break;
    if (rr_prev > 0.0 && hs->it > it_prev && hs->rr > 0.0 && hs->rr < rr_prev)
    {
      const double per_it = std::log(hs->rr / rr_prev) / (hs->it - it_prev); // < 0
      const double need = std::log(hs->tol2 / hs->rr) / per_it;             // iterations left
      chunk = (int)std::min<double>(PCG_CHUNK, std::max(1.0, std::floor(need)));
    }
    else if (rr_prev > 0.0)
      chunk = 2; // no decrease seen (the residual norm of CG need not fall monotonically)
    rr_prev = hs->rr;
    it_prev = hs->it;
    for (int c = 0; c < chunk; c++)
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
