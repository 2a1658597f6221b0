// Discrete gradient of the structured ND_p space and the auxiliary-space
// (gradient) correction of the Jacobi preconditioner.
//
// H1 space: continuous Q_p on the closed (Gauss-Lobatto) points, one node per
// point of the (n0 p + 1) x (n1 p + 1) x (n2 p + 1) grid, x fastest.  The
// discrete gradient G maps nodal values to ND DOFs exactly (the ND space holds
// the gradients of Q_p): the DOF of component c at (open point i of element e
// along c, closed indices j, k along the other axes) is
//     sum_m Dg[i][m] phi(e p + m along c, j, k),   Dg[i][m] = L_m'(xi_i),
// L_m the closed-point Lagrange basis, xi_i the open (Gauss) points, both on
// the reference interval (0, 1) of the ND basis.  So K G = 0 and
// G^T A G = G^T M_sigma G is the sigma-weighted H1 stiffness matrix.
//
// The preconditioner is the additive two-space form
//     B r = D^-1 r + G Dg^-1 G^T r,   Dg = diag(G^T A G),
// (Hiptmair-Xu with point smoothers in both spaces): SPD, and the gradient
// fields -- the near-null space of the curl-curl term, whose eigenvalues are
// set by sigma alone -- are handled in the space where they are smooth.
// Nodes on PEC faces are essential (Dg^-1 = 0 there).  The reference's
// counterpart for this operator is HypreAMS (hcurl_formulation.cpp:69-79),
// which adds multigrid on both auxiliary spaces; this is its one-level form.
#pragma once

#include "hc_common.hpp"

#include <type_traits>

namespace hc
{

struct AuxGeom
{
  int n[3];                       // elements per direction
  int p;                          // order
  int M[3];                       // nodes per direction (n p + 1)
  int cd[3][3];                   // ND component grid dims [c][axis]
  int64_t off[4];                 // ND component offsets, off[3] = N
  int NN;                         // nodes
  double dg[PMAX * (PMAX + 1)];   // Dg[i][m] at i (p + 1) + m
  double rcd[3][2];               // 1 / cd[c][0], 1 / cd[c][1]
};

// q = n / d, r = n % d for 0 <= n < 2^31, 1 <= d, with rd = 1 / d: the FP64
// product is within one of the quotient, one correction step makes it exact.
__device__ inline void
fast_divmod(int n, int d, double rd, int & q, int & r)
{
  q = (int)((double)n * rd);
  r = n - q * d;
  if (r < 0)
  {
    q -= 1;
    r += d;
  }
  else if (r >= d)
  {
    q += 1;
    r -= d;
  }
}

// The 1D derivative table Dg of order P in registers: its entries are
// wave-uniform, and a row or column chosen by a per-lane index is a short
// select chain instead of an indexed memory load.
template <int P>
struct DgTab
{
  double d[P][P + 1];
};

template <int P>
__device__ inline DgTab<P>
load_dg(const AuxGeom & a)
{
  DgTab<P> t;
#pragma unroll
  for (int i = 0; i < P; i++)
#pragma unroll
    for (int m = 0; m <= P; m++)
      t.d[i][m] = a.dg[i * (P + 1) + m];
  return t;
}

// (G^T r) at node (I, J, K), order P.
template <int P, typename T>
__device__ inline double
aux_gt(const AuxGeom & a, const DgTab<P> & tab, const T * __restrict__ r, int I, int J, int K)
{
  const int idx[3] = {I, J, K};
  double g = 0.0;
#pragma unroll
  for (int c = 0; c < 3; c++)
  {
    const int ic = idx[c];
    // DOF line of component c through the node: stride along c and its base
    int64_t base;
    int stride;
    if (c == 0)
    {
      base = a.off[0] + ((int64_t)K * a.cd[0][1] + J) * a.cd[0][0];
      stride = 1;
    }
    else if (c == 1)
    {
      base = a.off[1] + (int64_t)K * a.cd[1][1] * a.cd[1][0] + I;
      stride = a.cd[1][0];
    }
    else
    {
      base = a.off[2] + (int64_t)J * a.cd[2][0] + I;
      stride = a.cd[2][0] * a.cd[2][1];
    }
    // the (at most two) elements along c holding the node: local index m of
    // the node in the lower-numbered one's upper end (m = P) and in the other
    const int e = ic / P, m = ic % P;
    if (e < a.n[c])
    {
      double col[P];
#pragma unroll
      for (int i = 0; i < P; i++)
      {
        col[i] = tab.d[i][0];
#pragma unroll
        for (int k = 1; k < P; k++)
          col[i] = (m == k) ? tab.d[i][k] : col[i];
      }
      const T * rl = r + base + (int64_t)(e * P) * stride;
#pragma unroll
      for (int i = 0; i < P; i++)
        g += col[i] * (double)rl[(int64_t)i * stride];
    }
    if (m == 0 && ic > 0)
    {
      const T * rl = r + base + (int64_t)((e - 1) * P) * stride;
#pragma unroll
      for (int i = 0; i < P; i++)
        g += tab.d[i][P] * (double)rl[(int64_t)i * stride];
    }
  }
  return g;
}

// Node index walker of a grid-stride loop (start k0, step S): keeps (I, J, K)
// of k up to date with additions instead of divisions.
struct NodeWalk
{
  int I, J, K, sI, sJ, sK;
  __device__ NodeWalk(const AuxGeom & a, int k0, int S)
  {
    const int m01 = a.M[0] * a.M[1];
    I = k0 % a.M[0];
    J = (k0 / a.M[0]) % a.M[1];
    K = k0 / m01;
    sI = S % a.M[0];
    sJ = (S / a.M[0]) % a.M[1];
    sK = S / m01;
  }
  __device__ void step(const AuxGeom & a)
  {
    I += sI;
    int cy = 0;
    if (I >= a.M[0])
    {
      I -= a.M[0];
      cy = 1;
    }
    J += sJ + cy;
    cy = 0;
    if (J >= a.M[1])
    {
      J -= a.M[1];
      cy = 1;
    }
    K += sK + cy;
  }
};

// Component and grid indices of ND DOF i (divisions by the component grid
// sizes through fast_divmod).
__device__ inline void
aux_dof_coords(const AuxGeom & a, int64_t i, int & c, int & i0, int & i1, int & i2)
{
  c = (i < a.off[1]) ? 0 : ((i < a.off[2]) ? 1 : 2);
  const int li = (int)(i - a.off[c]);
  int rest;
  fast_divmod(li, a.cd[c][0], a.rcd[c][0], rest, i0);
  fast_divmod(rest, a.cd[c][1], a.rcd[c][1], i2, i1);
}

// (G t) at the ND DOF of component c at (i0, i1, i2), order P.
template <int P, typename T>
__device__ inline double
aux_g_at(const AuxGeom & a, const DgTab<P> & tab, const T * __restrict__ t, int c, int i0, int i1, int i2)
{
  const int ic = (c == 0) ? i0 : ((c == 1) ? i1 : i2);
  const int e = ic / P, l = ic % P;
  int node[3] = {i0, i1, i2};
  node[c] = e * P;
  const int stride = (c == 0) ? 1 : ((c == 1) ? a.M[0] : a.M[0] * a.M[1]);
  const T * tl = t + ((int64_t)node[2] * a.M[1] + node[1]) * a.M[0] + node[0];
  double row[P + 1];
#pragma unroll
  for (int m = 0; m <= P; m++)
  {
    row[m] = tab.d[0][m];
#pragma unroll
    for (int k = 1; k < P; k++)
      row[m] = (l == k) ? tab.d[k][m] : row[m];
  }
  double v = 0.0;
#pragma unroll
  for (int m = 0; m <= P; m++)
    v += row[m] * (double)tl[(int64_t)m * stride];
  return v;
}

// (G t) at ND DOF i, order P.
template <int P, typename T>
__device__ inline double
aux_g(const AuxGeom & a, const DgTab<P> & tab, const T * __restrict__ t, int64_t i)
{
  int c, i0, i1, i2;
  aux_dof_coords(a, i, c, i0, i1, i2);
  return aux_g_at<P>(a, tab, t, c, i0, i1, i2);
}

// Runs f with the order of a as a compile-time constant (orders 1..PMAX).
static_assert(PMAX == 3, "with_order dispatches orders 1..3");
template <typename F>
__device__ inline void
with_order(const AuxGeom & a, F && f)
{
  switch (a.p)
  {
  case 1:
    f(std::integral_constant<int, 1>{});
    break;
  case 2:
    f(std::integral_constant<int, 2>{});
    break;
  default:
    f(std::integral_constant<int, 3>{});
    break;
  }
}

// Host: geometry of the auxiliary space of g, with the 1D derivative matrix dg1d.
AuxGeom make_aux_geom(const Geom & g, const double * dg1d);
// Node diagonal inverse Dg^-1 (0 on PEC nodes and where Dg = 0) from the
// mass factors of the quadrature data qd (FP64) of the operator; dg_scratch
// (NN doubles) may be ginv64 itself.
hipError_t aux_diag_inv(const Geom & g, const AuxGeom & a, const Tables<double> & tb, const double * qd,
                        double * dg_scratch, double * ginv64, float * ginv32, hipStream_t st);
// Per-block minima (TERM_RATIO_BLOCKS doubles) over the elements of
// trace(mass factors) / trace(curl factors), i.e. sigma h^2 / nu per element.
constexpr int TERM_RATIO_BLOCKS = 256;
hipError_t term_ratio_blocks(const double * qd, int64_t E, int p, double * out, hipStream_t st);
// FP32 inner solves of FP32 + FP64 refinement are used for an operator only
// when min_e sigma h_eff^2 / nu (h_eff = h / p) is at least this: FP32
// rounding of a vector puts curl-curl noise of relative size eps32 = 6e-8 on
// fields whose true operator action (the gradients, eigenvalues ~ sigma) is
// sigma h_eff^2 / nu times smaller; the bound keeps that noise below 1 %.
constexpr double FP32_MIN_TERM_RATIO = 6.0e-6;
// AV block pieces (the blocks of src/formulations/AV/av_formulation.cpp:150-196
// built from G and the ND operator):
// s = a0 + G v with a0 = a zeroed on the PEC DOFs of g (all ND DOFs).
hipError_t av_gather(const Geom & g, const AuxGeom & a, const double * av, const double * v, double * s,
                     hipStream_t st);
// ya = ya + k on the free DOFs, ya = a on the PEC DOFs of g.
hipError_t av_finish(const Geom & g, const AuxGeom & a, const double * av, const double * k, double * ya,
                     hipStream_t st);
// y = G v on every ND DOF.
hipError_t grad_apply(const AuxGeom & a, const double * v, double * y, hipStream_t st);
// yv = G^T w at every node (no essential nodes).
hipError_t grad_transpose(const AuxGeom & a, const double * w, double * yv, hipStream_t st);
// z = D^-1 r + G Dg^-1 G^T r for one vector (t: node scratch, NN entries).
// One step of the Chebyshev iteration with the hybrid smoother B = D^-1 + G Dg^-1 G^T:
// t = Dg^-1 G^T r, then d = ca B r + cb d and x += d in one pass (flags 1:
// fresh d; 2: fresh x, x = d).
hipError_t aux_smooth_step(const AuxGeom & a, const double * dinv, const double * ginv, const double * r, double * t,
                           double ca, double cb, double * d, double * x, int flags, hipStream_t st, int nrhs = 1,
                           int64_t sn = 0, int64_t snn = 0);
hipError_t aux_precond_apply(const AuxGeom & a, const double * dinv, const double * ginv, const double * r, double * t,
                             double * z, hipStream_t st, int nrhs = 1, int64_t sn = 0, int64_t snn = 0);

} // namespace hc
