// Jacobi-preconditioned CG kernels: the device side of DefaultJacobiPCGSolver
// (src/solvers/hephaestus_solvers.hpp:34-57: HyprePCG + HypreDiagScale [ext]),
// and the FP64 outer residual / refinement steps, batched over solve slots.
//
// A slot is one right-hand side of one operator instance (its own length n,
// vectors, Jacobi diagonal, states and reduction partials), read from a device
// table; the grid is (blocks per slot, slots).  One PCG iteration is q = A p
// (the operator kernels, gated by each slot's inner done flag), then
//   pcg_dot_pq     partial (p, q) per block
//   pcg_update_xr  alpha = gamma / (p,q);  x += alpha p;  r -= alpha q;
//                  partials (r, D^-1 r), (r, r)          -- one pass over x,p,r,q,d
//   pcg_update_p   beta = gamma_new / gamma;  p = D^-1 r + beta p
// Every scalar stays on the device (PcgState).  Reductions are two-stage with
// a block count fixed by n and a fixed summation order, so the result of a
// slot is bitwise independent of the other slots of the launch.  Dot products
// accumulate in FP64 for FP32 vectors too.  Each slot carries a phase (outer
// state) that gates the outer kernels, and the inner done flag gates the
// iteration kernels, so slots at different stages -- and slots that join a
// running batch -- share every launch; the iteration index of a slot is
// (device counter) - (its start), the counter being advanced by pcg_dot_pq.
#include "hc_kernels.hpp"

namespace hc
{

__device__ inline int
slot_blocks(int64_t n)
{
  return red_blocks(n);
}

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

// Reduce the nb block partials of slots k0.. (stride 4) in a fixed order;
// every block computes the same value.
template <int K>
__device__ inline void
reduce_partials(const double * __restrict__ part, int nb, int k0, double (&out)[K], double * sh)
{
  double v[K];
#pragma unroll
  for (int k = 0; k < K; k++)
    v[k] = 0.0;
  for (int i = threadIdx.x; i < nb; i += RED_THREADS)
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

// grid-stride loop over the entries of one slot handled by this block
#define HC_SLOT_LOOP(i, n, nb)                                                                                    \
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < (n); i += (int64_t)(nb)*RED_THREADS)

// ---------------------------------------------------------------------------
// PCG
// ---------------------------------------------------------------------------
// r = b - q (or b when zero_guess, then x = 0);  p = D^-1 r;
// partials (b,b), (r, D^-1 r), (r,r); slots in PH_INIT.
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_init(const SolveSlot * __restrict__ slots, int zero_guess)
{
  const SolveSlot & sl = slots[blockIdx.y];
  const int nb = slot_blocks(sl.n);
  if ((int)blockIdx.x >= nb || sl.out->phase != PH_INIT)
    return;
  const T * __restrict__ b = (const T *)sl.ib;
  const T * __restrict__ q = (const T *)sl.iq;
  const T * __restrict__ dinv = (const T *)sl.idinv;
  T * __restrict__ x = (T *)sl.ix;
  T * __restrict__ r = (T *)sl.ir;
  T * __restrict__ p = (T *)sl.ip;
  __shared__ double sh[3 * RED_THREADS];
  double v[3] = {0, 0, 0};
  HC_SLOT_LOOP(i, sl.n, nb)
  {
    const double bi = (double)b[i];
    double ri = bi;
    if (zero_guess)
      x[i] = T(0);
    else
      ri = bi - (double)q[i];
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
    sl.part[blockIdx.x * 4 + 0] = v[0];
    sl.part[blockIdx.x * 4 + 1] = v[1];
    sl.part[blockIdx.x * 4 + 2] = v[2];
  }
}

// Inner state of each slot in PH_INIT, which moves to PH_INNER.  Tolerance:
// rtol of the slot, or, for the inner solve of a refinement (inner_floor > 0),
// max(inner_floor, 0.5 rtol ||b|| / ||r||) of the outer state.  Iteration
// budget: the slot's maxit less the inner iterations it has used.
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_init_fin(const SolveSlot * __restrict__ slots, const int * __restrict__ git, double inner_floor)
{
  const SolveSlot & sl = slots[blockIdx.x];
  PcgState * s = sl.in;
  PcgState * o = sl.out;
  if (o->phase != PH_INIT)
    return;
  __shared__ double sh[3 * RED_THREADS];
  double v[3];
  reduce_partials<3>(sl.part, slot_blocks(sl.n), 0, v, sh);
  if (threadIdx.x == 0)
  {
    double tol = sl.rtol;
    if (inner_floor > 0)
      tol = fmax(inner_floor, o->rr > 0 ? 0.5 * sl.rtol * sqrt(o->bb / o->rr) : 1.0);
    s->bb = v[0];
    s->gam[0] = v[1];
    s->gam[1] = 0.0;
    s->rr = v[2];
    s->tol2 = tol * tol;
    s->iters = 0;
    s->it0 = *git;
    s->maxit = sl.maxit - o->inner;
    s->status = 0;
    const bool nan = bad(v[0]) || bad(v[1]) || bad(v[2]);
    if (nan)
      s->status = -5;
    s->done = (nan || v[2] <= s->tol2 * v[0] || s->maxit <= 0) ? 1 : 0;
    o->phase = PH_INNER;
  }
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_dot_pq(const SolveSlot * __restrict__ slots, int * git)
{
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0)
    *git = *git + 1; // read by the next kernels only
  const SolveSlot & sl = slots[blockIdx.y];
  const int nb = slot_blocks(sl.n);
  if ((int)blockIdx.x >= nb || sl.in->done)
    return;
  const T * __restrict__ p = (const T *)sl.ip;
  const T * __restrict__ q = (const T *)sl.iq;
  __shared__ double sh[RED_THREADS];
  double v[1] = {0};
  HC_SLOT_LOOP(i, sl.n, nb)
  v[0] += (double)p[i] * (double)q[i];
  block_sum<1>(v, sh);
  if (threadIdx.x == 0)
    sl.part[blockIdx.x * 4 + 3] = v[0];
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_update_xr(const SolveSlot * __restrict__ slots, const int * __restrict__ git)
{
  const SolveSlot & sl = slots[blockIdx.y];
  const PcgState * s = sl.in;
  const int nb = slot_blocks(sl.n);
  if ((int)blockIdx.x >= nb || s->done)
    return;
  const int it = *git - 1 - s->it0;
  T * __restrict__ x = (T *)sl.ix;
  T * __restrict__ r = (T *)sl.ir;
  const T * __restrict__ p = (const T *)sl.ip;
  const T * __restrict__ q = (const T *)sl.iq;
  const T * __restrict__ dinv = (const T *)sl.idinv;
  __shared__ double sh[2 * RED_THREADS];
  double pq[1];
  reduce_partials<1>(sl.part, nb, 3, pq, sh);
  const double alpha = s->gam[it & 1] / pq[0];
  const T a = (T)alpha;
  double v[2] = {0, 0};
  HC_SLOT_LOOP(i, sl.n, nb)
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
    sl.part[blockIdx.x * 4 + 0] = v[0];
    sl.part[blockIdx.x * 4 + 1] = v[1];
  }
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_update_p(const SolveSlot * __restrict__ slots, const int * __restrict__ git)
{
  const SolveSlot & sl = slots[blockIdx.y];
  PcgState * s = sl.in;
  const int nb = slot_blocks(sl.n);
  if ((int)blockIdx.x >= nb || s->done)
    return;
  const int it = *git - 1 - s->it0;
  __shared__ double sh[2 * RED_THREADS];
  double o[2];
  reduce_partials<2>(sl.part, nb, 0, o, sh);
  const double gnew = o[0], rr = o[1];
  const double gold = s->gam[it & 1];
  const bool nan = bad(gnew) || bad(rr);
  const bool conv = nan || rr <= s->tol2 * s->bb;
  const bool stop = conv || it + 1 >= s->maxit;
  if (blockIdx.x == 0 && threadIdx.x == 0)
  {
    s->gam[(it + 1) & 1] = gnew;
    s->rr = rr;
    s->iters = it + 1;
    if (nan)
      s->status = -5;
    if (stop)
      s->done = 1;
  }
  if (stop)
    return;
  T * __restrict__ p = (T *)sl.ip;
  const T * __restrict__ r = (const T *)sl.ir;
  const T * __restrict__ dinv = (const T *)sl.idinv;
  const T beta = (T)(gnew / gold);
  HC_SLOT_LOOP(i, sl.n, nb)
  p[i] = dinv[i] * r[i] + beta * p[i];
}

// ---------------------------------------------------------------------------
// phases, FP64 outer residual and refinement
// ---------------------------------------------------------------------------
// PH_FIRST -> PH_RESID (residual of the initial guess); PH_INNER with a
// finished inner solve -> PH_RESID.  One thread per slot.
__global__ void
k_phase_prep(const SolveSlot * __restrict__ slots, int nslots)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots)
    return;
  PcgState * o = slots[i].out;
  if (o->phase == PH_FIRST)
  {
    o->phase = PH_RESID;
    o->first = 1;
  }
  else if (o->phase == PH_INNER && slots[i].in->done)
  {
    o->phase = PH_RESID;
    o->first = 0;
  }
}

// x += ||r64|| * ix (float): the refinement update after an FP32 inner solve,
// the scale of the k_to_f32_scaled that made its right-hand side
__global__ void
k_ir_update(const SolveSlot * __restrict__ slots)
{
  const SolveSlot & sl = slots[blockIdx.y];
  const PcgState * s = sl.out;
  if (s->phase != PH_RESID || s->first)
    return;
  const double sc = sqrt(s->rr);
  double * __restrict__ x = sl.x;
  const float * __restrict__ d32 = (const float *)sl.ix;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < sl.n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] += sc * (double)d32[i];
}

// r64 = b - q64 (q64 = A x), partials (r,r), (b,b); slots in PH_RESID
__global__ void __launch_bounds__(RED_THREADS)
k_residual64(const SolveSlot * __restrict__ slots)
{
  const SolveSlot & sl = slots[blockIdx.y];
  const int nb = slot_blocks(sl.n);
  if ((int)blockIdx.x >= nb || sl.out->phase != PH_RESID)
    return;
  const double * __restrict__ b = sl.b;
  const double * __restrict__ Ax = sl.q64;
  double * __restrict__ r = sl.r64;
  __shared__ double sh[2 * RED_THREADS];
  double v[2] = {0, 0};
  HC_SLOT_LOOP(i, sl.n, nb)
  {
    const double ri = b[i] - Ax[i];
    r[i] = ri;
    v[0] += ri * ri;
    v[1] += b[i] * b[i];
  }
  block_sum<2>(v, sh);
  if (threadIdx.x == 0)
  {
    sl.part[blockIdx.x * 4 + 0] = v[0];
    sl.part[blockIdx.x * 4 + 1] = v[1];
  }
}

// Outer state of slots in PH_RESID: rr (and bb on the first residual);
// otherwise the inner iterations of the finished inner solve are added and the
// step counted.  Then PH_DONE once rr <= rtol^2 bb, or on a NaN, after
// max_outer steps or when the iteration budget is spent; PH_INIT otherwise.
__global__ void __launch_bounds__(RED_THREADS)
k_residual_fin(const SolveSlot * __restrict__ slots, int max_outer)
{
  const SolveSlot & sl = slots[blockIdx.x];
  PcgState * s = sl.out;
  if (s->phase != PH_RESID)
    return;
  __shared__ double sh[2 * RED_THREADS];
  double o[2];
  reduce_partials<2>(sl.part, slot_blocks(sl.n), 0, o, sh);
  if (threadIdx.x == 0)
  {
    s->rr = o[0];
    if (s->first)
    {
      s->bb = o[1];
      s->tol2 = sl.rtol * sl.rtol;
      s->status = 0;
      s->iters = 0;
      s->inner = 0;
      s->first = 0;
    }
    else
    {
      s->inner += sl.in->iters;
      s->iters += 1;
      if (sl.in->status)
        s->status = sl.in->status;
    }
    if (bad(o[0]))
      s->status = -5;
    const bool stop =
        s->status != 0 || o[0] <= s->tol2 * s->bb || s->iters >= max_outer || s->inner >= sl.maxit;
    s->done = stop ? 1 : 0;
    s->phase = stop ? PH_DONE : PH_INIT;
  }
}

// ib (float) = r64 / ||r64|| for slots in PH_INIT
__global__ void
k_to_f32_scaled(const SolveSlot * __restrict__ slots)
{
  const SolveSlot & sl = slots[blockIdx.y];
  const PcgState * s = sl.out;
  if (s->phase != PH_INIT)
    return;
  const double inv = s->rr > 0 ? 1.0 / sqrt(s->rr) : 0.0;
  const double * __restrict__ r64 = sl.r64;
  float * __restrict__ r32 = (float *)sl.ib;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < sl.n; i += (int64_t)gridDim.x * blockDim.x)
    r32[i] = (float)(r64[i] * inv);
}

__global__ void
k_gather_states(const SolveSlot * __restrict__ slots, int nslots, PcgState * __restrict__ dst)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nslots)
    dst[i] = *slots[i].out;
}

// ---------------------------------------------------------------------------
// single-vector helpers
// ---------------------------------------------------------------------------
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

__global__ void
k_zero_int(int * v, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    v[i] = 0;
}

// ----------------------------------------------------------------------------
// launchers
// ----------------------------------------------------------------------------
static inline dim3
slot_grid(const SlotTable & t)
{
  return dim3(t.nb_max, t.nslots);
}

template <typename T>
hipError_t
pcg_init(const SlotTable & t, const int * git, bool zero_guess, double inner_floor, hipStream_t st)
{
  hipLaunchKernelGGL(k_pcg_init<T>, slot_grid(t), dim3(RED_THREADS), 0, st, t.slots, zero_guess ? 1 : 0);
  hipLaunchKernelGGL(k_pcg_init_fin, dim3(t.nslots), dim3(RED_THREADS), 0, st, t.slots, git, inner_floor);
  return hipGetLastError();
}

template <typename T>
hipError_t
pcg_iterate(const SlotTable & t, int * git, hipStream_t st)
{
  hipLaunchKernelGGL(k_pcg_dot_pq<T>, slot_grid(t), dim3(RED_THREADS), 0, st, t.slots, git);
  hipLaunchKernelGGL(k_pcg_update_xr<T>, slot_grid(t), dim3(RED_THREADS), 0, st, t.slots, (const int *)git);
  hipLaunchKernelGGL(k_pcg_update_p<T>, slot_grid(t), dim3(RED_THREADS), 0, st, t.slots, (const int *)git);
  return hipGetLastError();
}

template hipError_t pcg_init<double>(const SlotTable &, const int *, bool, double, hipStream_t);
template hipError_t pcg_init<float>(const SlotTable &, const int *, bool, double, hipStream_t);
template hipError_t pcg_iterate<double>(const SlotTable &, int *, hipStream_t);
template hipError_t pcg_iterate<float>(const SlotTable &, int *, hipStream_t);

hipError_t
phase_prep(const SlotTable & t, hipStream_t st)
{
  hipLaunchKernelGGL(k_phase_prep, dim3((t.nslots + 255) / 256), dim3(256), 0, st, t.slots, t.nslots);
  return hipGetLastError();
}

hipError_t
residual64(const SlotTable & t, int max_outer, hipStream_t st)
{
  hipLaunchKernelGGL(k_residual64, slot_grid(t), dim3(RED_THREADS), 0, st, t.slots);
  hipLaunchKernelGGL(k_residual_fin, dim3(t.nslots), dim3(RED_THREADS), 0, st, t.slots, max_outer);
  return hipGetLastError();
}

hipError_t
to_f32_scaled(const SlotTable & t, hipStream_t st)
{
  hipLaunchKernelGGL(k_to_f32_scaled, slot_grid(t), dim3(RED_THREADS), 0, st, t.slots);
  return hipGetLastError();
}

hipError_t
ir_update(const SlotTable & t, hipStream_t st)
{
  hipLaunchKernelGGL(k_ir_update, slot_grid(t), dim3(RED_THREADS), 0, st, t.slots);
  return hipGetLastError();
}

hipError_t
gather_states(const SlotTable & t, PcgState * dst, hipStream_t st)
{
  hipLaunchKernelGGL(k_gather_states, dim3((t.nslots + 255) / 256), dim3(256), 0, st, t.slots, t.nslots, dst);
  return hipGetLastError();
}

hipError_t
zero_int(int * v, int n, hipStream_t st)
{
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_zero_int, dim3((n + 255) / 256), dim3(256), 0, st, v, n);
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
