// Shared definitions of the MI355X H(curl) operator library: structured-mesh
// geometry, 1D basis tables, launch helpers.
//
// The operator is the ND_p hex curl-curl + mass bilinear form that the
// reference registers through CurlCurlKernel (src/kernels/curl_curl_kernel.cpp:20-24)
// and VectorFEMassKernel (src/kernels/vector_fe_mass_kernel.cpp:20-24); the
// element arithmetic follows the published MFEM ND_HexahedronElement
// mathematics (open Gauss-Legendre x closed Gauss-Lobatto tensor bases,
// covariant Piola map).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hc
{

constexpr int PMAX = 3;   // orders 1..3 are compiled
constexpr int QMAXC = PMAX + 1;

// 1D tables of order p (Q = p + 1 Gauss points on (0,1)).  Passed by value as
// a kernel argument so that, after the contraction loops are unrolled, every
// table access is a uniform (scalar) load.
template <typename T>
struct Tables
{
  T Bo[QMAXC * PMAX];    // open basis at quadrature points   [q][d], d < p
  T Bc[QMAXC * QMAXC];   // closed basis                        [q][d], d <= p
  T Gc[QMAXC * QMAXC];   // derivative of closed basis          [q][d]
  T qw[QMAXC];           // Gauss weights (sum to 1)
  T qp[QMAXC];           // Gauss points on (0,1)
};

// Structured hex mesh + ND_p DOF layout.
struct Geom
{
  int n[3];        // elements per direction
  int p;           // order
  int pec;         // PEC face bits (bit 2d: min face of axis d, bit 2d+1: max face)
  int T[3];        // apply tile extent (elements)
  int nt[3];       // number of tiles per direction
  int64_t off[4];  // component offsets, off[3] = N
  int64_t cd[3][3];// component grid dims [c][axis]
};

__host__ __device__ inline bool
is_ess(const Geom & g, int c, int64_t i0, int64_t i1, int64_t i2)
{
  const int64_t idx[3] = {i0, i1, i2};
  bool e = false;
#pragma unroll
  for (int a = 0; a < 3; a++)
  {
    if (a == c)
      continue;
    if (idx[a] == 0 && ((g.pec >> (2 * a)) & 1))
      e = true;
    if (idx[a] == g.cd[c][a] - 1 && ((g.pec >> (2 * a + 1)) & 1))
      e = true;
  }
  return e;
}

// Global DOF -> (component, i0, i1, i2).
__host__ __device__ inline void
dof_coords(const Geom & g, int64_t gid, int & c, int64_t & i0, int64_t & i1, int64_t & i2)
{
  c = (gid >= g.off[1]) + (gid >= g.off[2]);
  int64_t r = gid - g.off[c];
  i0 = r % g.cd[c][0];
  int64_t t = r / g.cd[c][0];
  i1 = t % g.cd[c][1];
  i2 = t / g.cd[c][1];
}

// Tile-boundary plane test for closed direction a of a DOF at index ia.
__host__ __device__ inline bool
on_tile_plane(const Geom & g, int a, int64_t ia)
{
  const int64_t step = (int64_t)g.T[a] * g.p;
  return ia > 0 && ia < (int64_t)g.n[a] * g.p && (ia % step) == 0;
}

inline int
div_up(int64_t a, int64_t b)
{
  return (int)((a + b - 1) / b);
}

// Sum over the 64 lanes of a wavefront by a butterfly of cross-lane
// exchanges.  In every step the two lanes of a pair add the same two operands,
// so all lanes end with the same value and the summation order is fixed.
__device__ inline double
wave_sum(double v)
{
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
    v += __shfl_xor(v, m, 64);
  return v;
}

// Sum over a workgroup of NT threads (a multiple of 64) of K values per
// thread: wavefront sums, then the NT / 64 wave sums in wave order.  Every
// thread receives the results; sh (K * NT / 64 doubles) is free again on
// return.
template <int NT, int K>
__device__ inline void
group_sum(double (&v)[K], double * sh)
{
  constexpr int NW = NT / 64;
  static_assert(NT % 64 == 0, "whole wavefronts");
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; k++)
  {
    v[k] = wave_sum(v[k]);
    if ((threadIdx.x & 63) == 0)
      sh[k * NW + w] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++)
  {
    double r = sh[k * NW];
#pragma unroll
    for (int i = 1; i < NW; i++)
      r += sh[k * NW + i];
    v[k] = r;
  }
  __syncthreads();
}
template <int NT>
__device__ inline double
group_sum(double v, double * sh)
{
  double a[1] = {v};
  group_sum<NT, 1>(a, sh);
  return a[0];
}

} // namespace hc
