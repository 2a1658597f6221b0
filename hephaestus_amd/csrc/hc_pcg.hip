// Jacobi-preconditioned CG kernels: the device side of DefaultJacobiPCGSolver
// (src/solvers/hephaestus_solvers.hpp:34-57: HyprePCG + HypreDiagScale [ext]).
//
// One iteration is: q = A p (operator kernels), then
//   pcg_dot_pq     partial (p, q) per block
//   pcg_update_xr  alpha = gamma / (p,q);  x += alpha p;  r -= alpha q;
//                  partials (r, D^-1 r), (r, r)          -- one pass over x,p,r,q,d
//   pcg_update_p   beta = gamma_new / gamma;  p = D^-1 r + beta p
// Every scalar stays on the device (PcgState); reductions are two-stage with a
// fixed block grid and a fixed summation order, so results are bitwise
// reproducible and independent of what else runs on the GPU.  Dot products are
// accumulated in FP64 for FP32 vectors too.  Convergence sets state->done; all
// later kernels of the solve return at once, so the host checks the flag only
// every few iterations.
#include "hc_kernels.hpp"

namespace hc
{

// Block-wide sum of K values per thread (fixed tree order).
template <int K>
__device__ inline void
block_sum(double (&v)[K], double * sh)
{
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < K; k++)
    sh[k * RED_THREADS + t] = v[k];
  __syncthreads();
  for (int s = RED_THREADS / 2; s > 0; s >>= 1)
  {
    if (t < s)
    {
#pragma unroll
      for (int k = 0; k < K; k++)
        sh[k * RED_THREADS + t] += sh[k * RED_THREADS + t + s];
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < K; k++)
    v[k] = sh[k * RED_THREADS];
  __syncthreads();
}

// Reduce the RED_BLOCKS partials of slot k (stride 4) in a fixed order; every
// block computes the same value.
template <int K>
__device__ inline void
reduce_partials(const double * __restrict__ part, int k0, double (&out)[K], double * sh)
{
  double v[K];
#pragma unroll
  for (int k = 0; k < K; k++)
    v[k] = 0.0;
  for (int i = threadIdx.x; i < RED_BLOCKS; i += RED_THREADS)
  {
#pragma unroll
    for (int k = 0; k < K; k++)
      v[k] += part[i * 4 + k0 + k];
  }
  block_sum<K>(v, sh);
#pragma unroll
  for (int k = 0; k < K; k++)
    out[k] = v[k];
}

__device__ inline bool
bad(double v)
{
  return !(fabs(v) <= 1.7976931348623157e308);
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_init(const T * __restrict__ b, const T * __restrict__ q, const T * __restrict__ dinv, T * __restrict__ r,
           T * __restrict__ p, int64_t n, double * __restrict__ part)
{
  __shared__ double sh[3 * RED_THREADS];
  double v[3] = {0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * RED_THREADS)
  {
    const double bi = (double)b[i];
    const double ri = q ? bi - (double)q[i] : bi;
    const double zi = (double)dinv[i] * ri;
    r[i] = (T)ri;
    p[i] = (T)zi;
    v[0] += bi * bi;
    v[1] += ri * zi;
    v[2] += ri * ri;
  }
  block_sum<3>(v, sh);
  if (threadIdx.x == 0)
  {
    part[blockIdx.x * 4 + 0] = v[0];
    part[blockIdx.x * 4 + 1] = v[1];
    part[blockIdx.x * 4 + 2] = v[2];
  }
}

__global__ void __launch_bounds__(RED_THREADS)
k_pcg_init_fin(const double * __restrict__ part, PcgState * s, double tol2)
{
  __shared__ double sh[3 * RED_THREADS];
  double o[3];
  reduce_partials<3>(part, 0, o, sh);
  if (threadIdx.x == 0)
  {
    s->bb = o[0];
    s->gam[0] = o[1];
    s->gam[1] = 0.0;
    s->rr = o[2];
    s->tol2 = tol2;
    s->iters = 0;
    s->status = 0;
    const bool nan = bad(o[0]) || bad(o[1]) || bad(o[2]);
    if (nan)
      s->status = -5;
    s->done = (nan || o[2] <= tol2 * o[0]) ? 1 : 0;
  }
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_dot_pq(const T * __restrict__ p, const T * __restrict__ q, int64_t n, double * __restrict__ part,
             const PcgState * __restrict__ s)
{
  if (s->done)
    return;
  __shared__ double sh[RED_THREADS];
  double v[1] = {0};
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * RED_THREADS)
    v[0] += (double)p[i] * (double)q[i];
  block_sum<1>(v, sh);
  if (threadIdx.x == 0)
    part[blockIdx.x * 4 + 3] = v[0];
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_update_xr(int it, T * __restrict__ x, const T * __restrict__ p, T * __restrict__ r, const T * __restrict__ q,
                const T * __restrict__ dinv, int64_t n, double * __restrict__ part, const PcgState * __restrict__ s)
{
  if (s->done)
    return;
  __shared__ double sh[2 * RED_THREADS];
  double pq[1];
  reduce_partials<1>(part, 3, pq, sh);
  const double alpha = s->gam[it & 1] / pq[0];
  const T a = (T)alpha;
  double v[2] = {0, 0};
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * RED_THREADS)
  {
    x[i] = x[i] + a * p[i];
    const T ri = r[i] - a * q[i];
    r[i] = ri;
    const double rd = (double)ri;
    v[0] += rd * ((double)dinv[i] * rd);
    v[1] += rd * rd;
  }
  block_sum<2>(v, sh);
  if (threadIdx.x == 0)
  {
    part[blockIdx.x * 4 + 0] = v[0];
    part[blockIdx.x * 4 + 1] = v[1];
  }
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_update_p(int it, T * __restrict__ p, const T * __restrict__ r, const T * __restrict__ dinv, int64_t n,
               const double * __restrict__ part, PcgState * s)
{
  if (s->done)
    return;
  __shared__ double sh[2 * RED_THREADS];
  double o[2];
  reduce_partials<2>(part, 0, o, sh);
  const double gnew = o[0], rr = o[1];
  const double gold = s->gam[it & 1];
  const bool nan = bad(gnew) || bad(rr);
  const bool conv = nan || rr <= s->tol2 * s->bb;
  if (blockIdx.x == 0 && threadIdx.x == 0)
  {
    s->gam[(it + 1) & 1] = gnew;
    s->rr = rr;
    s->iters = it + 1;
    if (nan)
      s->status = -5;
    if (conv)
      s->done = 1;
  }
  if (conv)
    return;
  const T beta = (T)(gnew / gold);
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * RED_THREADS)
    p[i] = dinv[i] * r[i] + beta * p[i];
}

// ----------------------------------------------------------------------------
// FP64 refinement helpers
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(RED_THREADS)
k_residual64(const double * __restrict__ b, const double * __restrict__ Ax, double * __restrict__ r, int64_t n,
             double * __restrict__ part)
{
  __shared__ double sh[2 * RED_THREADS];
  double v[2] = {0, 0};
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * RED_THREADS)
  {
    const double ri = b[i] - Ax[i];
    r[i] = ri;
    v[0] += ri * ri;
    v[1] += b[i] * b[i];
  }
  block_sum<2>(v, sh);
  if (threadIdx.x == 0)
  {
    part[blockIdx.x * 4 + 0] = v[0];
    part[blockIdx.x * 4 + 1] = v[1];
  }
}

__global__ void __launch_bounds__(RED_THREADS)
k_residual_fin(const double * __restrict__ part, PcgState * s, int set_bb)
{
  __shared__ double sh[2 * RED_THREADS];
  double o[2];
  reduce_partials<2>(part, 0, o, sh);
  if (threadIdx.x == 0)
  {
    s->rr = o[0];
    if (set_bb)
    {
      // start of a new outer sequence
      s->bb = o[1];
      s->status = 0;
      s->done = 0;
      s->iters = 0;
    }
    if (bad(o[0]))
      s->status = -5;
  }
}

__global__ void
k_to_f32_scaled(const double * __restrict__ r64, float * __restrict__ r32, int64_t n, const PcgState * __restrict__ s)
{
  const double inv = s->rr > 0 ? 1.0 / sqrt(s->rr) : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    r32[i] = (float)(r64[i] * inv);
}

__global__ void
k_ir_update(double * __restrict__ x, const float * __restrict__ d32, int64_t n, const PcgState * __restrict__ s)
{
  const double sc = sqrt(s->rr);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] += sc * (double)d32[i];
}

__global__ void
k_invert_diag(const double * __restrict__ d, double * __restrict__ d64, float * __restrict__ d32, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
  {
    const double v = 1.0 / d[i];
    if (d64)
      d64[i] = v;
    if (d32)
      d32[i] = (float)v;
  }
}

__device__ inline uint64_t
splitmix64_at(uint64_t seed, uint64_t i)
{
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ void
k_fill_uniform(Geom g, int use_g, uint64_t seed, double * __restrict__ v, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
  {
    double val = 2.0 * ((double)(splitmix64_at(seed, (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    if (use_g)
    {
      int c;
      int64_t a, b, cc;
      dof_coords(g, i, c, a, b, cc);
      if (is_ess(g, c, a, b, cc))
        val = 0.0;
    }
    v[i] = val;
  }
}

// ----------------------------------------------------------------------------
// launchers
// ----------------------------------------------------------------------------
#define HC_LAUNCH_RED(kern, ...)                                                                                  \
  hipLaunchKernelGGL(kern, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, __VA_ARGS__)

template <typename T>
hipError_t
pcg_init(const T * b, const T * q, const T * dinv, T * r, T * p, int64_t n, double rtol, PcgWork w, hipStream_t st)
{
  HC_LAUNCH_RED(k_pcg_init<T>, b, q, dinv, r, p, n, w.part);
  hipLaunchKernelGGL(k_pcg_init_fin, dim3(1), dim3(RED_THREADS), 0, st, w.part, w.state, rtol * rtol);
  return hipGetLastError();
}

template <typename T>
hipError_t
pcg_dot_pq(const T * p, const T * q, int64_t n, PcgWork w, hipStream_t st)
{
  HC_LAUNCH_RED(k_pcg_dot_pq<T>, p, q, n, w.part, w.state);
  return hipGetLastError();
}

template <typename T>
hipError_t
pcg_update_xr(int it, T * x, const T * p, T * r, const T * q, const T * dinv, int64_t n, PcgWork w, hipStream_t st)
{
  HC_LAUNCH_RED(k_pcg_update_xr<T>, it, x, p, r, q, dinv, n, w.part, w.state);
  return hipGetLastError();
}

template <typename T>
hipError_t
pcg_update_p(int it, T * p, const T * r, const T * dinv, int64_t n, PcgWork w, hipStream_t st)
{
  HC_LAUNCH_RED(k_pcg_update_p<T>, it, p, r, dinv, n, w.part, w.state);
  return hipGetLastError();
}

#define HC_INST(T)                                                                                                 \
  template hipError_t pcg_init<T>(const T *, const T *, const T *, T *, T *, int64_t, double, PcgWork,            \
                                  hipStream_t);                                                                    \
  template hipError_t pcg_dot_pq<T>(const T *, const T *, int64_t, PcgWork, hipStream_t);                         \
  template hipError_t pcg_update_xr<T>(int, T *, const T *, T *, const T *, const T *, int64_t, PcgWork,          \
                                       hipStream_t);                                                               \
  template hipError_t pcg_update_p<T>(int, T *, const T *, const T *, int64_t, PcgWork, hipStream_t);
HC_INST(double)
HC_INST(float)
#undef HC_INST

hipError_t
residual64(const double * b, const double * Ax, double * r, int64_t n, PcgWork w, bool set_bb, hipStream_t st)
{
  HC_LAUNCH_RED(k_residual64, b, Ax, r, n, w.part);
  hipLaunchKernelGGL(k_residual_fin, dim3(1), dim3(RED_THREADS), 0, st, w.part, w.state, set_bb ? 1 : 0);
  return hipGetLastError();
}

hipError_t
to_f32_scaled(const double * r64, float * r32, int64_t n, const PcgState * s, hipStream_t st)
{
  hipLaunchKernelGGL(k_to_f32_scaled, dim3(1024), dim3(256), 0, st, r64, r32, n, s);
  return hipGetLastError();
}

hipError_t
ir_update(double * x, const float * d32, int64_t n, const PcgState * s, hipStream_t st)
{
  hipLaunchKernelGGL(k_ir_update, dim3(1024), dim3(256), 0, st, x, d32, n, s);
  return hipGetLastError();
}

hipError_t
invert_diag(const double * d, double * d64, float * d32, int64_t n, hipStream_t st)
{
  hipLaunchKernelGGL(k_invert_diag, dim3(1024), dim3(256), 0, st, d, d64, d32, n);
  return hipGetLastError();
}

hipError_t
fill_uniform(const Geom * g, uint64_t seed, double * v, int64_t n, hipStream_t st)
{
  Geom gg{};
  if (g)
    gg = *g;
  hipLaunchKernelGGL(k_fill_uniform, dim3(1024), dim3(256), 0, st, gg, g ? 1 : 0, seed, v, n);
  return hipGetLastError();
}

} // namespace hc
