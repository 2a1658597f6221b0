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
//   pcg_update_p   beta = gamma_new / gamma;  p = z + beta p
// with z = D^-1 r (Jacobi), or z = D^-1 r + G Dg^-1 G^T r for slots with the
// gradient-space correction (hc_aux.hpp): then aux_nodes runs between the two
// updates, t = Dg^-1 G^T r at the nodes and the partial (G^T r, t) that
// completes (r, z), and pcg_update_p adds (G t) while it forms p.
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

// Block-wide sum of K values per thread (wavefront butterflies, then the
// wave sums in order: a fixed order).
template <int K>
__device__ inline void
block_sum(double (&v)[K], double * sh)
{
  group_sum<RED_THREADS, K>(v, sh);
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

// FP32 PCG: FP64 residual replacement at least every this many iterations (a
// stagnating FP32 recursion is found within that many)
constexpr int RELIABLE_MAX_GAP = 256;

__device__ inline bool
bad(double v)
{
  return !(fabs(v) <= 1.7976931348623157e308);
}

// grid-stride loop over the entries of one slot handled by this block
// (bk: the block's index within its slot)
#define HC_SLOT_LOOP(i, n, nb)                                                                                    \
  for (int64_t i = (int64_t)bk * RED_THREADS + threadIdx.x; i < (n); i += (int64_t)(nb)*RED_THREADS)

// Block of the flattened slot grid: slot s and block b within the slot (the
// blocks of slot s are [blk0[s], blk0[s + 1]), blk0[s + 1] - blk0[s] =
// slot_blocks(n of slot s)).
struct SlotBlock
{
  int s, b;
};
__device__ inline SlotBlock
slot_block(const SlotTable & t)
{
  const int g = blockIdx.x;
  int lo = 0, hi = t.nslots - 1;
  while (lo < hi)
  {
    const int mid = (lo + hi + 1) >> 1;
    if (t.blk0[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return {lo, g - t.blk0[lo]};
}

// Vectors of a slot in the precision of T.
template <typename T>
struct SlotVecs;
template <>
struct SlotVecs<double>
{
  static constexpr int RUN = 2;
  const SolveSlot & s;
  __device__ double * x() const { return s.x; }
  __device__ double * r() const { return s.r64; }
  __device__ double * p() const { return s.p64; }
  __device__ double * q() const { return s.q64; }
  __device__ const double * dinv() const { return s.dinv64; }
  __device__ double * t() const { return s.t64; }
  __device__ const double * ginv() const { return s.ginv64; }
};
template <>
struct SlotVecs<float>
{
  static constexpr int RUN = 1;
  const SolveSlot & s;
  __device__ float * x() const { return s.x32; }
  __device__ float * r() const { return s.r32; }
  __device__ float * p() const { return s.p32; }
  __device__ float * q() const { return s.q32; }
  __device__ const float * dinv() const { return s.dinv32; }
  __device__ float * t() const { return s.t32; }
  __device__ const float * ginv() const { return s.ginv32; }
};

// ---------------------------------------------------------------------------
// PCG
// ---------------------------------------------------------------------------
// FP64 start: r = b - q64 (q64 = A x of this round), p = D^-1 r; partials
// (b,b), (r, D^-1 r), (r,r).  FP64 slots in PH_INIT.
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_init64(SlotTable t)
{
  const SlotBlock sb = slot_block(t);
  const int bk = sb.b;
  const SolveSlot & sl = t.slots[sb.s];
  const int nb = slot_blocks(sl.n);
  if (sl.out->phase != PH_INIT || !sl.out->prec64 || !sl.g64)
    return;
  const double * __restrict__ b = sl.b;
  const double * __restrict__ q = sl.q64;
  const double * __restrict__ dinv = sl.dinv64;
  double * __restrict__ r = sl.r64;
  double * __restrict__ p = sl.p64;
  __shared__ double sh[3 * RED_THREADS];
  double v[3] = {0, 0, 0};
  HC_SLOT_LOOP(i, sl.n, nb)
  {
    const double bi = b[i];
    const double ri = bi - q[i];
    const double zi = dinv[i] * ri;
    r[i] = ri;
    if (!sl.aux)
      p[i] = zi; // else set by aux_pset once t is known
    v[0] += bi * bi;
    v[1] += ri * zi;
    v[2] += ri * ri;
  }
  block_sum<3>(v, sh);
  if (threadIdx.x == 0)
  {
    sl.part[bk * 4 + 0] = v[0];
    sl.part[bk * 4 + 1] = v[1];
    sl.part[bk * 4 + 2] = v[2];
  }
}

// Inner state of FP64 slots in PH_INIT, which move to PH_INNER.  Tolerance:
// the slot's rtol; iteration budget: maxit less the inner iterations used.
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_init64_fin(const SolveSlot * __restrict__ slots, const int * __restrict__ git)
{
  const SolveSlot & sl = slots[blockIdx.x];
  PcgState * s = sl.in;
  PcgState * o = sl.out;
  if (o->phase != PH_INIT || !o->prec64 || !sl.g64)
    return;
  __shared__ double sh[4 * RED_THREADS];
  double v[4];
  reduce_partials<4>(sl.part, slot_blocks(sl.n), 0, v, sh);
  if (sl.aux)
    v[1] += v[3]; // (r, G Dg^-1 G^T r) from aux_nodes
  if (threadIdx.x == 0)
  {
    s->bb = v[0];
    s->gam[0] = v[1];
    s->gam[1] = 0.0;
    s->rr = v[2];
    s->rrmax = v[2];
    s->tol2 = sl.rtol * sl.rtol;
    s->iters = 0;
    s->counted = 0;
    s->it0 = *git;
    s->maxit = sl.maxit - o->inner;
    s->status = 0;
    s->first = 0;
    const bool nan = bad(v[0]) || bad(v[1]) || bad(v[2]);
    if (nan)
      s->status = -5;
    s->done = (nan || v[2] <= s->tol2 * v[0] || s->maxit <= 0) ? 1 : 0;
    s->run = s->done ? 0 : 2;
    o->phase = PH_INNER;
  }
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_dot_pq(SlotTable t, int * git)
{
  if (blockIdx.x == 0 && threadIdx.x == 0)
    *git = *git + 1; // read by the next kernels only
  const SlotBlock sb = slot_block(t);
  const int bk = sb.b;
  const SolveSlot & sl = t.slots[sb.s];
  const int nb = slot_blocks(sl.n);
  const SlotVecs<T> V{sl};
  if (sl.in->run != SlotVecs<T>::RUN)
    return;
  const T * __restrict__ p = V.p();
  const T * __restrict__ q = V.q();
  __shared__ double sh[RED_THREADS];
  double v[1] = {0};
  HC_SLOT_LOOP(i, sl.n, nb)
  v[0] += (double)p[i] * (double)q[i];
  block_sum<1>(v, sh);
  if (threadIdx.x == 0)
    sl.part[bk * 4 + 3] = v[0];
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_update_xr(SlotTable t, const int * __restrict__ git)
{
  const SlotBlock sb = slot_block(t);
  const int bk = sb.b;
  const SolveSlot & sl = t.slots[sb.s];
  const PcgState * s = sl.in;
  const int nb = slot_blocks(sl.n);
  const SlotVecs<T> V{sl};
  if (s->run != SlotVecs<T>::RUN)
    return;
  const int it = *git - 1 - s->it0;
  T * __restrict__ x = V.x();
  T * __restrict__ r = V.r();
  const T * __restrict__ p = V.p();
  const T * __restrict__ q = V.q();
  const T * __restrict__ dinv = V.dinv();
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
    sl.part[bk * 4 + 0] = v[0];
    sl.part[bk * 4 + 1] = v[1];
  }
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_update_p(SlotTable t, const int * __restrict__ git)
{
  const SlotBlock sb = slot_block(t);
  const int bk = sb.b;
  const SolveSlot & sl = t.slots[sb.s];
  PcgState * s = sl.in;
  const int nb = slot_blocks(sl.n);
  const SlotVecs<T> V{sl};
  if (s->run != SlotVecs<T>::RUN)
    return;
  const int it = *git - 1 - s->it0;
  __shared__ double sh[3 * RED_THREADS];
  double o[3];
  reduce_partials<3>(sl.part, nb, 0, o, sh);
  const double gnew = sl.aux ? o[0] + o[2] : o[0], rr = o[1];
  const double gold = s->gam[it & 1];
  const bool nan = bad(gnew) || bad(rr);
  const bool conv = nan || rr <= s->tol2 * s->bb;
  const bool stop = conv || it + 1 >= s->maxit;
  if (bk == 0 && threadIdx.x == 0)
  {
    s->gam[(it + 1) & 1] = gnew;
    s->rr = rr;
    s->iters = it + 1;
    if (rr > s->rrmax)
      s->rrmax = rr; // FP32: k_phase_prep compares against it
    if (nan)
      s->status = -5;
    if (stop)
    {
      s->done = 1;
      s->run = 0; // later blocks of this launch may see it: they stop too
    }
  }
  if (stop)
    return;
  T * __restrict__ p = V.p();
  const T * __restrict__ r = V.r();
  const T * __restrict__ dinv = V.dinv();
  const T beta = (T)(gnew / gold);
  if (sl.aux)
  {
    const T * __restrict__ tv = V.t();
    with_order(sl.ag, [&](auto pc) {
      constexpr int P = decltype(pc)::value;
      const DgTab<P> tab = load_dg<P>(sl.ag);
      HC_SLOT_LOOP(i, sl.n, nb)
      p[i] = (dinv[i] * r[i] + (T)aux_g<P>(sl.ag, tab, tv, i)) + beta * p[i];
    });
  }
  else
  {
    HC_SLOT_LOOP(i, sl.n, nb)
    p[i] = dinv[i] * r[i] + beta * p[i];
  }
}

// Gradient-space part of z = B r for the slots with aux set: t = Dg^-1 G^T r
// at the nodes and the block partials of (G^T r, t) = (r, G t) into part[kp].
// which = 0: iterating slots of precision T (kp = 2, beside the (r, D^-1 r)
// and (r, r) partials of pcg_update_xr); 1: FP64 slots starting (PH_INIT);
// 2: FP32 slots starting (kp = 3 for both).
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_aux_nodes(SlotTable t, int which)
{
  const SlotBlock sb = slot_block(t);
  const int bk = sb.b;
  const SolveSlot & sl = t.slots[sb.s];
  const int nb = slot_blocks(sl.n);
  if (!sl.aux)
    return;
  if (which == 0 && sl.in->run != SlotVecs<T>::RUN)
    return;
  if (which == 1 && (sl.out->phase != PH_INIT || !sl.out->prec64 || !sl.g64))
    return;
  if (which == 2 && (sl.out->phase != PH_INIT || sl.out->prec64))
    return;
  const SlotVecs<T> V{sl};
  const T * __restrict__ r = V.r();
  const T * __restrict__ ginv = V.ginv();
  T * __restrict__ tv = V.t();
  __shared__ double sh[RED_THREADS];
  double v[1] = {0};
  with_order(sl.ag, [&](auto pc) {
    constexpr int P = decltype(pc)::value;
    const DgTab<P> tab = load_dg<P>(sl.ag);
    const int S = nb * RED_THREADS;
    int k = bk * RED_THREADS + threadIdx.x;
    NodeWalk nw(sl.ag, k, S);
    for (; k < sl.ag.NN; k += S, nw.step(sl.ag))
    {
      const double g = aux_gt<P>(sl.ag, tab, r, nw.I, nw.J, nw.K);
      const T tk = (T)((double)ginv[k] * g);
      tv[k] = tk;
      v[0] += g * (double)tk;
    }
  });
  block_sum<1>(v, sh);
  if (threadIdx.x == 0)
    sl.part[bk * 4 + (which == 0 ? 2 : 3)] = v[0];
}

// First search direction p = D^-1 r + G t of starting aux slots (which as in
// k_aux_nodes; FP32 slots only on a fresh start, a paused PCG keeps p).
template <typename T>
__global__ void __launch_bounds__(RED_THREADS)
k_aux_pset(SlotTable t, int which)
{
  const SlotBlock sb = slot_block(t);
  const int bk = sb.b;
  const SolveSlot & sl = t.slots[sb.s];
  const int nb = slot_blocks(sl.n);
  const PcgState * o = sl.out;
  if (!sl.aux || o->phase != PH_INIT)
    return;
  if ((which == 1) != (o->prec64 != 0) || (which == 1 && !sl.g64))
    return;
  if (which == 2 && o->iters > 0 && sl.in->first)
    return;
  const SlotVecs<T> V{sl};
  const T * __restrict__ r = V.r();
  const T * __restrict__ dinv = V.dinv();
  const T * __restrict__ tv = V.t();
  T * __restrict__ p = V.p();
  with_order(sl.ag, [&](auto pc) {
    constexpr int P = decltype(pc)::value;
    const DgTab<P> tab = load_dg<P>(sl.ag);
    HC_SLOT_LOOP(i, sl.n, nb)
    p[i] = dinv[i] * r[i] + (T)aux_g<P>(sl.ag, tab, tv, i);
  });
}

// ---------------------------------------------------------------------------
// phases, FP64 outer residual and refinement
// ---------------------------------------------------------------------------
// PH_FIRST -> PH_RESID (residual of the initial guess); PH_INNER with a
// finished inner solve -> PH_RESID; an FP32 PCG whose recursive residual fell
// by delta since the last replacement, or that ran RELIABLE_MAX_GAP iterations
// since, -> PH_RESID too (its iterate is folded into x and the residual
// recomputed in FP64; the search direction is kept).  One thread per slot.
__global__ void
k_phase_prep(const SolveSlot * __restrict__ slots, int nslots)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots)
    return;
  const SolveSlot & sl = slots[i];
  PcgState * o = sl.out;
  PcgState * in = sl.in;
  if (o->phase == PH_FIRST)
  {
    o->phase = PH_RESID;
    o->first = 1;
  }
  else if (o->phase == PH_RESID && o->revert)
    o->revert = 0; // x was restored last round: its residual is recomputed now
  else if (o->phase == PH_INNER &&
           (in->run == 0 ||
            (!o->prec64 && (in->rr < sl.delta2 * in->rrmax || in->iters - in->counted >= RELIABLE_MAX_GAP))))
  {
    in->first = in->run != 0 ? 1 : 0; // 1: the search direction stays valid
    in->run = 0;
    in->done = 1;
    o->phase = PH_RESID;
    o->first = 0;
  }
}

// x += x32: the FP32 iterate since the last replacement folded into the FP64
// solution (FP32 slots entering PH_RESID after iterations)
__global__ void
k_ir_update(SlotTable t)
{
  const SlotBlock sb = slot_block(t);
  const int bk = sb.b;
  const SolveSlot & sl = t.slots[sb.s];
  const PcgState * s = sl.out;
  if (s->phase != PH_RESID || s->first || s->prec64)
    return;
  double * __restrict__ x = sl.x;
  const float * __restrict__ d32 = sl.x32;
  for (int64_t i = (int64_t)bk * blockDim.x + threadIdx.x; i < sl.n; i += (int64_t)slot_blocks(sl.n) * blockDim.x)
    x[i] += (double)d32[i];
}

// r64 = b - q64 (q64 = A x), partials (r,r), (b,b); slots in PH_RESID
__global__ void __launch_bounds__(RED_THREADS)
k_residual64(SlotTable t)
{
  const SlotBlock sb = slot_block(t);
  const int bk = sb.b;
  const SolveSlot & sl = t.slots[sb.s];
  const int nb = slot_blocks(sl.n);
  if (sl.out->phase != PH_RESID)
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
    sl.part[bk * 4 + 0] = v[0];
    sl.part[bk * 4 + 1] = v[1];
  }
}

// Outer state of slots in PH_RESID: rr (and bb on the first residual);
// otherwise the iterations of the paused or finished PCG are counted and the
// step too.  An FP32 slot continues in FP64 when its true residual norm exceeds
// the recursive one by sqrt(10) at a replacement, when two replacements in a
// row did not halve it, or -- after undoing the last FP32 increment -- when
// the true residual grew or a NaN appeared.  Then PH_DONE once
// rr <= rtol^2 bb, on a NaN, after max_outer steps or when the iteration
// budget is spent; PH_INIT otherwise.
__global__ void __launch_bounds__(RED_THREADS)
k_residual_fin(const SolveSlot * __restrict__ slots, int max_outer, int max_outer64)
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
    const double rr = o[0];
    s->rr = rr;
    if (s->first)
    {
      s->bb = o[1];
      s->tol2 = sl.rtol * sl.rtol;
      s->status = 0;
      s->iters = 0;
      s->inner = 0;
      s->first = 0;
      s->stall = 0;
    }
    else
    {
      PcgState * in = sl.in;
      s->inner += in->iters - in->counted;
      in->counted = in->iters;
      s->iters += 1;
      if (s->prec64)
        s->outer64 += 1;
      if (!s->prec64)
      {
        if (in->status || !(rr <= s->rrmax))
        {
          // the FP32 increment blew the residual up (or overflowed): undo it
          // and continue in FP64 from the previous x
          s->prec64 = 1;
          s->revert = 1;
          in->status = 0;
          in->first = 0;
          s->phase = PH_RESID; // its residual is recomputed next round
          s->done = 0;
          return;
        }
        // continue in FP64 from this x when the FP32 recursion no longer
        // tracks the true residual (the FP32 operator cannot resolve what is
        // left of it), or after two replacements in a row that did not halve
        // the residual norm
        s->stall = (rr > 0.25 * s->rrmax) ? s->stall + 1 : 0;
        if (rr > 10.0 * in->rr || s->stall >= 2)
        {
          s->prec64 = 1;
          in->first = 0;
        }
      }
      else if (in->status)
        s->status = in->status;
    }
    s->rrmax = rr;
    if (bad(rr))
      s->status = -5;
    const bool stop = s->status != 0 || rr <= s->tol2 * s->bb || s->iters >= max_outer ||
                      s->outer64 >= max_outer64 || s->inner >= sl.maxit;
    s->done = stop ? 1 : 0;
    s->phase = stop ? PH_DONE : PH_INIT;
  }
}

// x -= x32 for slots that reject their last FP32 increment
__global__ void
k_ir_revert(SlotTable t)
{
  const SlotBlock sb = slot_block(t);
  const int bk = sb.b;
  const SolveSlot & sl = t.slots[sb.s];
  if (!sl.out->revert || sl.out->phase != PH_RESID)
    return;
  double * __restrict__ x = sl.x;
  const float * __restrict__ d32 = sl.x32;
  for (int64_t i = (int64_t)bk * blockDim.x + threadIdx.x; i < sl.n; i += (int64_t)slot_blocks(sl.n) * blockDim.x)
    x[i] -= (double)d32[i];
}

// FP32 PCG (re)start of FP32 slots in PH_INIT: r32 = (float) r64 (the FP64
// true residual), x32 = 0; p = D^-1 r32 unless the PCG paused for a
// replacement (then the search direction is kept).  Partials (r32, D^-1 r32),
// (r32, r32).
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_init_rel(SlotTable t)
{
  const SlotBlock sb = slot_block(t);
  const int bk = sb.b;
  const SolveSlot & sl = t.slots[sb.s];
  const int nb = slot_blocks(sl.n);
  const PcgState * o = sl.out;
  if (o->phase != PH_INIT || o->prec64)
    return;
  const bool fresh = !(o->iters > 0 && sl.in->first);
  const double * __restrict__ r64 = sl.r64;
  const float * __restrict__ dinv = sl.dinv32;
  float * __restrict__ x = sl.x32;
  float * __restrict__ r = sl.r32;
  float * __restrict__ p = sl.p32;
  __shared__ double sh[2 * RED_THREADS];
  double v[2] = {0, 0};
  HC_SLOT_LOOP(i, sl.n, nb)
  {
    const float ri = (float)r64[i];
    r[i] = ri;
    x[i] = 0.0f;
    const float zi = dinv[i] * ri;
    if (fresh && !sl.aux)
      p[i] = zi; // aux: set by aux_pset once t is known
    v[0] += (double)ri * (double)zi;
    v[1] += (double)ri * (double)ri;
  }
  block_sum<2>(v, sh);
  if (threadIdx.x == 0)
  {
    sl.part[bk * 4 + 0] = v[0];
    sl.part[bk * 4 + 1] = v[1];
  }
}

// Inner state of FP32 slots in PH_INIT, which move to PH_INNER.  The
// iteration index continues across replacements, so (r, z) goes to the parity
// slot the next iteration reads.
__global__ void __launch_bounds__(RED_THREADS)
k_pcg_init_rel_fin(const SolveSlot * __restrict__ slots, const int * __restrict__ git)
{
  const SolveSlot & sl = slots[blockIdx.x];
  PcgState * s = sl.in;
  PcgState * o = sl.out;
  if (o->phase != PH_INIT || o->prec64)
    return;
  __shared__ double sh[4 * RED_THREADS];
  double v[4];
  reduce_partials<4>(sl.part, slot_blocks(sl.n), 0, v, sh);
  if (sl.aux)
    v[0] += v[3]; // (r, G Dg^-1 G^T r) from aux_nodes
  if (threadIdx.x == 0)
  {
    const bool fresh = !(o->iters > 0 && s->first);
    const int g = *git;
    if (fresh)
    {
      s->it0 = g;
      s->iters = 0;
      s->counted = 0;
      s->gam[0] = s->gam[1] = 0.0;
      s->status = 0;
      s->maxit = sl.maxit - o->inner;
    }
    s->first = 0;
    s->gam[(g - s->it0) & 1] = v[0];
    s->rr = v[1];
    s->rrmax = v[1];
    s->bb = o->bb;
    s->tol2 = sl.rtol * sl.rtol;
    const bool nan = bad(v[0]) || bad(v[1]);
    if (nan)
      s->status = -5;
    s->done = (nan || v[1] <= s->tol2 * s->bb || s->iters >= s->maxit) ? 1 : 0;
    s->run = s->done ? 0 : 1;
    o->phase = PH_INNER;
  }
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
  return dim3(t.nblk);
}

hipError_t
pcg_init64(const SlotTable & t, const int * git_pair, hipStream_t st)
{
  hipLaunchKernelGGL(k_pcg_init64, slot_grid(t), dim3(RED_THREADS), 0, st, t);
  if (t.any_aux)
  {
    hipLaunchKernelGGL(k_aux_nodes<double>, slot_grid(t), dim3(RED_THREADS), 0, st, t, 1);
    hipLaunchKernelGGL(k_aux_pset<double>, slot_grid(t), dim3(RED_THREADS), 0, st, t, 1);
  }
  hipLaunchKernelGGL(k_pcg_init64_fin, dim3(t.nslots), dim3(RED_THREADS), 0, st, t.slots, git_pair + 1);
  return hipGetLastError();
}

// The iteration counter of precision T: git[0] for FP32, git[1] for FP64
// (each advances once per iteration of its precision group).
template <typename T>
hipError_t
pcg_iterate(const SlotTable & t, int * git_pair, hipStream_t st)
{
  int * git = git_pair + (sizeof(T) == 4 ? 0 : 1);
  hipLaunchKernelGGL(k_pcg_dot_pq<T>, slot_grid(t), dim3(RED_THREADS), 0, st, t, git);
  hipLaunchKernelGGL(k_pcg_update_xr<T>, slot_grid(t), dim3(RED_THREADS), 0, st, t, (const int *)git);
  if (t.any_aux)
    hipLaunchKernelGGL(k_aux_nodes<T>, slot_grid(t), dim3(RED_THREADS), 0, st, t, 0);
  hipLaunchKernelGGL(k_pcg_update_p<T>, slot_grid(t), dim3(RED_THREADS), 0, st, t, (const int *)git);
  return hipGetLastError();
}

template hipError_t pcg_iterate<double>(const SlotTable &, int *, hipStream_t);
template hipError_t pcg_iterate<float>(const SlotTable &, int *, hipStream_t);

hipError_t
phase_prep(const SlotTable & t, hipStream_t st)
{
  hipLaunchKernelGGL(k_phase_prep, dim3((t.nslots + 255) / 256), dim3(256), 0, st, t.slots, t.nslots);
  return hipGetLastError();
}

hipError_t
residual64(const SlotTable & t, int max_outer, int max_outer64, hipStream_t st)
{
  hipLaunchKernelGGL(k_residual64, slot_grid(t), dim3(RED_THREADS), 0, st, t);
  hipLaunchKernelGGL(k_residual_fin, dim3(t.nslots), dim3(RED_THREADS), 0, st, t.slots, max_outer, max_outer64);
  hipLaunchKernelGGL(k_ir_revert, slot_grid(t), dim3(RED_THREADS), 0, st, t);
  return hipGetLastError();
}

hipError_t
pcg_init_reliable(const SlotTable & t, const int * git_pair, hipStream_t st)
{
  hipLaunchKernelGGL(k_pcg_init_rel, slot_grid(t), dim3(RED_THREADS), 0, st, t);
  if (t.any_aux)
  {
    hipLaunchKernelGGL(k_aux_nodes<float>, slot_grid(t), dim3(RED_THREADS), 0, st, t, 2);
    hipLaunchKernelGGL(k_aux_pset<float>, slot_grid(t), dim3(RED_THREADS), 0, st, t, 2);
  }
  hipLaunchKernelGGL(k_pcg_init_rel_fin, dim3(t.nslots), dim3(RED_THREADS), 0, st, t.slots, git_pair);
  return hipGetLastError();
}

hipError_t
ir_update(const SlotTable & t, hipStream_t st)
{
  hipLaunchKernelGGL(k_ir_update, slot_grid(t), dim3(RED_THREADS), 0, st, t);
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
