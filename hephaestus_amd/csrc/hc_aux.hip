// Auxiliary-space (gradient) correction: node diagonal of G^T A G from the
// quadrature data, and the standalone preconditioner apply (the batched PCG
// kernels use the same device functions, hc_aux.hpp).
#include "hc_aux.hpp"

#include <algorithm>
#include "hc_kernels.hpp"

namespace hc
{

AuxGeom
make_aux_geom(const Geom & g, const double * dg1d)
{
  AuxGeom a{};
  a.p = g.p;
  int64_t nn = 1;
  for (int d = 0; d < 3; d++)
  {
    a.n[d] = g.n[d];
    a.M[d] = g.n[d] * g.p + 1;
    nn *= a.M[d];
  }
  a.NN = (int)nn;
  for (int c = 0; c < 3; c++)
    for (int d = 0; d < 3; d++)
      a.cd[c][d] = (int)g.cd[c][d];
  for (int c = 0; c < 4; c++)
    a.off[c] = g.off[c];
  for (int i = 0; i < g.p * (g.p + 1); i++)
    a.dg[i] = dg1d[i];
  for (int c = 0; c < 3; c++)
    for (int d = 0; d < 2; d++)
      a.rcd[c][d] = 1.0 / (double)a.cd[c][d];
  return a;
}

// Dg at the nodes of the elements of one colour (ex, ey, ez parities = bits
// of col): elements of one colour share no node, so each adds its element
// term to its nodes without conflicts and the colours run in a fixed order.
// One workgroup per element, one thread per element node (a fastest):
//   sum_q grad phi_n(q)^T Qm(q) grad phi_n(q)
// with Qm the symmetric mass factors 0..5 of the point (the operator's sigma
// term; the curl term vanishes on gradients) and grad phi_n in reference
// coordinates (the ND basis' own).
template <int P>
__global__ void __launch_bounds__(64)
k_aux_diag(Geom g, Tables<double> tb, const double * __restrict__ qd, int col, double * __restrict__ dg)
{
  constexpr int D = P + 1, Q = P + 1, QQ = Q * Q;
  const int hx = (g.n[0] - (col & 1) + 1) / 2, hy = (g.n[1] - ((col >> 1) & 1) + 1) / 2;
  const int b = blockIdx.x;
  const int ex = 2 * (b % hx) + (col & 1);
  const int ey = 2 * ((b / hx) % hy) + ((col >> 1) & 1);
  const int ez = 2 * (b / (hx * hy)) + ((col >> 2) & 1);
  if (ez >= g.n[2] || threadIdx.x >= D * D * D)
    return;
  const int a = threadIdx.x % D, bb = (threadIdx.x / D) % D, c = threadIdx.x / (D * D);
  const int64_t e = ((int64_t)ez * g.n[1] + ey) * g.n[0] + ex;
  const double * qe = qd + e * (12 * Q * QQ);
  double s = 0.0;
  for (int qz = 0; qz < Q; qz++)
    for (int qy = 0; qy < Q; qy++)
      for (int qx = 0; qx < Q; qx++)
      {
        const double gx = tb.Gc[qx * D + a] * tb.Bc[qy * D + bb] * tb.Bc[qz * D + c];
        const double gy = tb.Bc[qx * D + a] * tb.Gc[qy * D + bb] * tb.Bc[qz * D + c];
        const double gz = tb.Bc[qx * D + a] * tb.Bc[qy * D + bb] * tb.Gc[qz * D + c];
        const double * f = qe + qz * 12 * QQ + qy * Q + qx;
        const double m0 = f[0], m1 = f[QQ], m2 = f[2 * QQ], m3 = f[3 * QQ], m4 = f[4 * QQ], m5 = f[5 * QQ];
        s += gx * (m0 * gx + m1 * gy + m2 * gz) + gy * (m1 * gx + m3 * gy + m4 * gz) +
             gz * (m2 * gx + m4 * gy + m5 * gz);
      }
  const int M0 = g.n[0] * P + 1, M1 = g.n[1] * P + 1;
  const int64_t node = ((int64_t)(ez * P + c) * M1 + (ey * P + bb)) * M0 + (ex * P + a);
  dg[node] += s;
}

__global__ void
k_aux_fill_zero(double * __restrict__ v, int n)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    v[i] = 0.0;
}

// Dg^-1 at the nodes (dg and g64 may be the same array: each entry is read,
// then written, by one thread).
__global__ void
k_aux_invert(AuxGeom a, int pec, const double * dg, double * g64, float * __restrict__ g32)
{
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < a.NN; k += gridDim.x * blockDim.x)
  {
    const int idx[3] = {k % a.M[0], (k / a.M[0]) % a.M[1], k / (a.M[0] * a.M[1])};
    bool ess = false;
#pragma unroll
    for (int d = 0; d < 3; d++)
      ess = ess || (idx[d] == 0 && ((pec >> (2 * d)) & 1)) || (idx[d] == a.M[d] - 1 && ((pec >> (2 * d + 1)) & 1));
    const double v = dg[k];
    const double inv = (ess || !(v > 0.0)) ? 0.0 : 1.0 / v;
    g64[k] = inv;
    if (g32)
      g32[k] = (float)inv;
  }
}

hipError_t
aux_diag_inv(const Geom & g, const AuxGeom & a, const Tables<double> & tb, const double * qd, double * dg_scratch,
             double * ginv64, float * ginv32, hipStream_t st)
{
  hipLaunchKernelGGL(k_aux_fill_zero, dim3(1024), dim3(256), 0, st, dg_scratch, a.NN);
  for (int col = 0; col < 8; col++)
  {
    const int hx = (g.n[0] - (col & 1) + 1) / 2, hy = (g.n[1] - ((col >> 1) & 1) + 1) / 2;
    const int hz = (g.n[2] - ((col >> 2) & 1) + 1) / 2;
    const int nb = hx * hy * hz;
    if (nb <= 0)
      continue;
    switch (g.p)
    {
    case 1:
      hipLaunchKernelGGL(k_aux_diag<1>, dim3(nb), dim3(64), 0, st, g, tb, qd, col, dg_scratch);
      break;
    case 2:
      hipLaunchKernelGGL(k_aux_diag<2>, dim3(nb), dim3(64), 0, st, g, tb, qd, col, dg_scratch);
      break;
    case 3:
      hipLaunchKernelGGL(k_aux_diag<3>, dim3(nb), dim3(64), 0, st, g, tb, qd, col, dg_scratch);
      break;
    default:
      return hipErrorInvalidValue;
    }
  }
  hipLaunchKernelGGL(k_aux_invert, dim3(1024), dim3(256), 0, st, a, g.pec, (const double *)dg_scratch, ginv64,
                     ginv32);
  return hipGetLastError();
}

// Per-block minimum over elements of trace(mass factors) / trace(curl
// factors) at the first quadrature point: sigma h^2 / nu of the element on an
// affine hex (the ratio of the two terms on a field of unit reference size).
__global__ void __launch_bounds__(256)
k_term_ratio(const double * __restrict__ qd, int64_t E, int Q, double * __restrict__ out)
{
  __shared__ double sh[256];
  double v = 1.0e300;
  const int QQ = Q * Q;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < E; e += (int64_t)gridDim.x * 256)
  {
    const double * f = qd + e * (12 * (int64_t)Q * QQ);
    const double m = f[0] + f[3 * QQ] + f[5 * QQ];
    const double c = f[6 * QQ] + f[9 * QQ] + f[11 * QQ];
    const double r = c > 0.0 ? m / c : 1.0e300;
    v = r < v ? r : v;
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1)
  {
    if ((int)threadIdx.x < s2 && sh[threadIdx.x + s2] < sh[threadIdx.x])
      sh[threadIdx.x] = sh[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    out[blockIdx.x] = sh[0];
}

hipError_t
term_ratio_blocks(const double * qd, int64_t E, int p, double * out, hipStream_t st)
{
  hipLaunchKernelGGL(k_term_ratio, dim3(TERM_RATIO_BLOCKS), dim3(256), 0, st, qd, E, p + 1, out);
  return hipGetLastError();
}

// Standalone z = B r: t = Dg^-1 G^T r at the nodes, then z = D^-1 r + G t.
__global__ void
k_aux_nodes_single(AuxGeom a, const double * __restrict__ ginv, const double * __restrict__ r,
                   double * __restrict__ t, int64_t sr, int64_t st_)
{
  r += blockIdx.y * sr; // right-hand side blockIdx.y (strides 0 for one)
  t += blockIdx.y * st_;
  with_order(a, [&](auto pc) {
    constexpr int P = decltype(pc)::value;
    const DgTab<P> tab = load_dg<P>(a);
    const int S = gridDim.x * blockDim.x;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    NodeWalk nw(a, k, S);
    for (; k < a.NN; k += S, nw.step(a))
      t[k] = ginv[k] * aux_gt<P>(a, tab, r, nw.I, nw.J, nw.K);
  });
}

__global__ void
k_aux_z_single(AuxGeom a, const double * __restrict__ dinv, const double * __restrict__ r,
               const double * __restrict__ t, double * __restrict__ z, int64_t sr, int64_t st_)
{
  r += blockIdx.y * sr;
  z += blockIdx.y * sr;
  t += blockIdx.y * st_;
  with_order(a, [&](auto pc) {
    constexpr int P = decltype(pc)::value;
    const DgTab<P> tab = load_dg<P>(a);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.off[3];
         i += (int64_t)gridDim.x * blockDim.x)
      z[i] = dinv[i] * r[i] + aux_g<P>(a, tab, t, i);
  });
}

// One smoothing step after t = Dg^-1 G^T r: z = D^-1 r + G t, then the
// Chebyshev (or damped) update d = ca z + cb d, x += d, in one pass
// (flags 1: fresh d, d = ca z; 2: fresh x, x = d).  The same arithmetic as z
// followed by the separate update, with z kept in a register.
__global__ void
k_aux_z_cheb(AuxGeom a, const double * __restrict__ dinv, const double * __restrict__ r,
             const double * __restrict__ t, double ca, double cb, double * __restrict__ d, double * __restrict__ x,
             int flags, int64_t sr, int64_t st_)
{
  r += blockIdx.y * sr; // right-hand side blockIdx.y (strides 0 for one)
  d += blockIdx.y * sr;
  x += blockIdx.y * sr;
  t += blockIdx.y * st_;
  with_order(a, [&](auto pc) {
    constexpr int P = decltype(pc)::value;
    const DgTab<P> tab = load_dg<P>(a);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.off[3];
         i += (int64_t)gridDim.x * blockDim.x)
    {
      const double z = dinv[i] * r[i] + aux_g<P>(a, tab, t, i);
      const double di = (flags & 1) ? ca * z : ca * z + cb * d[i];
      d[i] = di;
      x[i] = (flags & 2) ? di : x[i] + di;
    }
  });
}

hipError_t
aux_smooth_step(const AuxGeom & a, const double * dinv, const double * ginv, const double * r, double * t,
                double ca, double cb, double * d, double * x, int flags, hipStream_t st, int nrhs, int64_t sn,
                int64_t snn)
{
  // the grids of aux_precond_apply (the per-entry arithmetic does not depend on them)
  const unsigned gn = nrhs > 1 ? std::max(64u, 2048u / (unsigned)nrhs) : 2048u;
  const unsigned gz = nrhs > 1 ? std::max(64u, 4096u / (unsigned)nrhs) : 4096u;
  hipLaunchKernelGGL(k_aux_nodes_single, dim3(gn, nrhs), dim3(256), 0, st, a, ginv, r, t, sn, snn);
  hipLaunchKernelGGL(k_aux_z_cheb, dim3(gz, nrhs), dim3(256), 0, st, a, dinv, r, (const double *)t, ca, cb, d, x,
                     flags, sn, snn);
  return hipGetLastError();
}

hipError_t
aux_precond_apply(const AuxGeom & a, const double * dinv, const double * ginv, const double * r, double * t,
                  double * z, hipStream_t st, int nrhs, int64_t sn, int64_t snn)
{
  // one right-hand side: the full grids; several: fewer blocks per right-hand side
  const unsigned gn = nrhs > 1 ? std::max(64u, 2048u / (unsigned)nrhs) : 2048u;
  const unsigned gz = nrhs > 1 ? std::max(64u, 4096u / (unsigned)nrhs) : 4096u;
  hipLaunchKernelGGL(k_aux_nodes_single, dim3(gn, nrhs), dim3(256), 0, st, a, ginv, r, t, sn, snn);
  hipLaunchKernelGGL(k_aux_z_single, dim3(gz, nrhs), dim3(256), 0, st, a, dinv, r, (const double *)t, z, sn, snn);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// AV block operator pieces.  Unknowns: a = dA/dt on the ND DOFs, v = V on the
// H1 nodes.  With s = a0 + G v (a0: a with its PEC DOFs zeroed),
//   ya = dt K a0 + M s   on the free DOFs, ya = a on the PEC DOFs,
//   yv = G^T M s         at every node,
// i.e. CurlCurl + VectorFEMass + MixedVectorGradient (sigma grad V, u') in the
// A rows and VectorFEWeakDivergence (sigma u, grad V') + Diffusion
// (sigma grad V, grad V') in the V rows, the forms stated at
// src/kernels/{mixed_vector_gradient,vector_fe_weak_divergence,diffusion}_kernel.hpp
// with the kernels registered at src/formulations/AV/av_formulation.cpp:150-196.
// M G is exact (G maps into the ND space) and K G = 0, so the mixed and
// diffusion blocks are M G and G^T M G.
// ---------------------------------------------------------------------------
// Row segments of the ND DOFs: one work item per element edge along a
// component's open axis (the p DOFs that share the p + 1 nodes of that edge),
// items of a component in the memory order of its DOFs with the open axis
// counted in elements.
struct RowItems
{
  int64_t ioff[4]; // first item of each component, ioff[3] = N / p
  int dim[3][2];   // items along the first two axes of each component
  double rd[3][2]; // their reciprocals (fast_divmod)
};

static RowItems
row_items(const AuxGeom & a)
{
  RowItems ri{};
  for (int c = 0; c < 4; c++)
    ri.ioff[c] = a.off[c] / a.p;
  for (int c = 0; c < 3; c++)
    for (int d = 0; d < 2; d++)
    {
      ri.dim[c][d] = (d == c) ? a.n[d] : a.cd[c][d];
      ri.rd[c][d] = 1.0 / (double)ri.dim[c][d];
    }
  return ri;
}

// s = a0 + G v (ADD_A, a0 = a with its PEC DOFs zeroed) or s = G v, one row
// segment per item: the p + 1 node values of the edge are loaded once for
// its p DOFs; each DOF sums its p + 1 terms in node order, as aux_g_at does.
template <bool ADD_A>
__global__ void
k_g_rows(Geom g, AuxGeom a, RowItems ri, const double * __restrict__ av, const double * __restrict__ v,
         double * __restrict__ s)
{
  with_order(a, [&](auto pc) {
    constexpr int P = decltype(pc)::value;
    const DgTab<P> tab = load_dg<P>(a);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ri.ioff[3];
         t += (int64_t)gridDim.x * blockDim.x)
    {
      const int c = (t < ri.ioff[1]) ? 0 : ((t < ri.ioff[2]) ? 1 : 2);
      int i[3], rest;
      fast_divmod((int)(t - ri.ioff[c]), ri.dim[c][0], ri.rd[c][0], rest, i[0]);
      fast_divmod(rest, ri.dim[c][1], ri.rd[c][1], i[2], i[1]);
      i[c] *= P; // first DOF of the edge, and its first node along c
      const int64_t d0 = a.off[c] + ((int64_t)i[2] * a.cd[c][1] + i[1]) * a.cd[c][0] + i[0];
      const int64_t dstride = (c == 0) ? 1 : ((c == 1) ? (int64_t)a.cd[1][0] : (int64_t)a.cd[2][0] * a.cd[2][1]);
      const int64_t nstride = (c == 0) ? 1 : ((c == 1) ? (int64_t)a.M[0] : (int64_t)a.M[0] * a.M[1]);
      const double * tl = v + ((int64_t)i[2] * a.M[1] + i[1]) * a.M[0] + i[0];
      double nv[P + 1];
#pragma unroll
      for (int m = 0; m <= P; m++)
        nv[m] = tl[m * nstride];
      const bool ess = ADD_A && is_ess(g, c, i[0], i[1], i[2]); // the same for the edge's DOFs
#pragma unroll
      for (int l = 0; l < P; l++)
      {
        double gv = 0.0;
#pragma unroll
        for (int m = 0; m <= P; m++)
          gv += tab.d[l][m] * nv[m];
        if constexpr (ADD_A)
        {
          const double a0 = ess ? 0.0 : av[d0 + l * dstride];
          s[d0 + l * dstride] = a0 + gv;
        }
        else
          s[d0 + l * dstride] = gv;
      }
    }
  });
}

__global__ void
k_av_finish(Geom g, AuxGeom a, const double * __restrict__ av, const double * __restrict__ k,
            double * __restrict__ ya)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g.off[3]; i += (int64_t)gridDim.x * blockDim.x)
  {
    int c, i0, i1, i2;
    aux_dof_coords(a, i, c, i0, i1, i2);
    ya[i] = is_ess(g, c, i0, i1, i2) ? av[i] : ya[i] + k[i];
  }
}

__global__ void
k_grad_transpose(AuxGeom a, const double * __restrict__ w, double * __restrict__ yv)
{
  with_order(a, [&](auto pc) {
    constexpr int P = decltype(pc)::value;
    const DgTab<P> tab = load_dg<P>(a);
    const int S = gridDim.x * blockDim.x;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    NodeWalk nw(a, k, S);
    for (; k < a.NN; k += S, nw.step(a))
      yv[k] = aux_gt<P>(a, tab, w, nw.I, nw.J, nw.K);
  });
}

hipError_t
grad_apply(const AuxGeom & a, const double * v, double * y, hipStream_t st)
{
  Geom g{}; // unused without ADD_A
  hipLaunchKernelGGL(k_g_rows<false>, dim3(4096), dim3(256), 0, st, g, a, row_items(a), nullptr, v, y);
  return hipGetLastError();
}

hipError_t
av_gather(const Geom & g, const AuxGeom & a, const double * av, const double * v, double * s, hipStream_t st)
{
  hipLaunchKernelGGL(k_g_rows<true>, dim3(4096), dim3(256), 0, st, g, a, row_items(a), av, v, s);
  return hipGetLastError();
}

hipError_t
av_finish(const Geom & g, const AuxGeom & a, const double * av, const double * k, double * ya, hipStream_t st)
{
  hipLaunchKernelGGL(k_av_finish, dim3(4096), dim3(256), 0, st, g, a, av, k, ya);
  return hipGetLastError();
}

hipError_t
grad_transpose(const AuxGeom & a, const double * w, double * yv, hipStream_t st)
{
  hipLaunchKernelGGL(k_grad_transpose, dim3(2048), dim3(256), 0, st, a, w, yv);
  return hipGetLastError();
}

} // namespace hc
