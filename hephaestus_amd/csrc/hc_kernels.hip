// CDNA4 (gfx950) kernels of the H(curl) curl-curl + sigma-mass operator:
//   * k_pa_setup      quadrature-point geometric factors (FP64), the partial-
//                     assembly replacement of ParBilinearForm::Assemble as
//                     driven from EquationSystem::BuildBilinearForms
//                     (src/equation_systems/equation_system.cpp:324-347);
//   * (the tiled element apply k_apply_tile lives in hc_apply.hip)
//   * k_assemble_planes  sums the (2 or 4) tile partials of every tile-face DOF
//                     in a fixed order (bitwise-reproducible results);
//   * k_diag_elem     Jacobi diagonal, element-parallel in 8 parity colours (no atomics).
// The apply is the operator EquationSystem::Mult (equation_system.cpp:200-204)
// on the PEC-eliminated system of FormLinearSystem (equation_system.cpp:132-190):
// y_free = (A x_free)_free, y_ess = x_ess.
#include "hc_kernels.hpp"

#include <algorithm>

namespace hc
{

// ---------------------------------------------------------------------------
// PA setup
// ---------------------------------------------------------------------------

// Mass factor ms*|det| (J^T J)^{-1}, curl factor cs * J^T J / |det|.
__device__ inline void
point_factors(const double J[9], double ms, double cs, double out[12])
{
  double G[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int s = 0; s < 3; s++)
      G[r * 3 + s] = J[0 * 3 + r] * J[0 * 3 + s] + J[1 * 3 + r] * J[1 * 3 + s] + J[2 * 3 + r] * J[2 * 3 + s];
  const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                     J[2] * (J[3] * J[7] - J[4] * J[6]);
  const double adet = fabs(det);
  const double A00 = G[4] * G[8] - G[5] * G[7];
  const double A01 = -(G[1] * G[8] - G[2] * G[7]);
  const double A02 = G[1] * G[5] - G[2] * G[4];
  const double A11 = G[0] * G[8] - G[2] * G[6];
  const double A12 = -(G[0] * G[5] - G[2] * G[3]);
  const double A22 = G[0] * G[4] - G[1] * G[3];
  const double sm = ms / adet;
  out[0] = sm * A00;
  out[1] = sm * A01;
  out[2] = sm * A02;
  out[3] = sm * A11;
  out[4] = sm * A12;
  out[5] = sm * A22;
  const double sc = cs / adet;
  out[6] = sc * G[0];
  out[7] = sc * G[1];
  out[8] = sc * G[2];
  out[9] = sc * G[4];
  out[10] = sc * G[5];
  out[11] = sc * G[8];
}

// One thread per (element, quadrature point).  qdata layout [e][qz][k][qy][qx].
__global__ void
k_pa_setup(Geom g,
           Tables<double> tb,
           const double * __restrict__ vert, // may be null (box mesh)
           double hx,
           double hy,
           double hz,
           const double * __restrict__ curl_coef,
           const double * __restrict__ mass_coef,
           double curl_scale,
           double * __restrict__ qd)
{
  const int Q = g.p + 1, Q3 = Q * Q * Q;
  const int64_t E = (int64_t)g.n[0] * g.n[1] * g.n[2];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * Q3)
    return;
  const int64_t e = t / Q3;
  const int q = (int)(t % Q3);
  const int qx = q % Q, qy = (q / Q) % Q, qz = q / (Q * Q);
  const double a = tb.qp[qx], b = tb.qp[qy], c = tb.qp[qz];
  double J[9];
  if (vert == nullptr)
  {
    J[0] = hx;
    J[1] = 0;
    J[2] = 0;
    J[3] = 0;
    J[4] = hy;
    J[5] = 0;
    J[6] = 0;
    J[7] = 0;
    J[8] = hz;
  }
  else
  {
    const int64_t ex = e % g.n[0], ey = (e / g.n[0]) % g.n[1], ez = e / ((int64_t)g.n[0] * g.n[1]);
    const int64_t sy = g.n[0] + 1, sz = (int64_t)(g.n[0] + 1) * (g.n[1] + 1);
    const double L[2][3] = {{1 - a, 1 - b, 1 - c}, {a, b, c}};
    const double dL[2] = {-1.0, 1.0};
#pragma unroll
    for (int i = 0; i < 9; i++)
      J[i] = 0;
#pragma unroll
    for (int kc = 0; kc < 2; kc++)
#pragma unroll
      for (int kb = 0; kb < 2; kb++)
#pragma unroll
        for (int ka = 0; ka < 2; ka++)
        {
          const double * V = vert + 3 * ((ex + ka) + (ey + kb) * sy + (ez + kc) * sz);
          const double g0 = dL[ka] * L[kb][1] * L[kc][2];
          const double g1 = L[ka][0] * dL[kb] * L[kc][2];
          const double g2 = L[ka][0] * L[kb][1] * dL[kc];
#pragma unroll
          for (int r = 0; r < 3; r++)
          {
            J[r * 3 + 0] += V[r] * g0;
            J[r * 3 + 1] += V[r] * g1;
            J[r * 3 + 2] += V[r] * g2;
          }
        }
  }
  const double w = tb.qw[qx] * tb.qw[qy] * tb.qw[qz];
  double f[12];
  point_factors(J, w * mass_coef[e], w * curl_scale * curl_coef[e], f);
#pragma unroll
  for (int k = 0; k < 12; k++)
    qd[((e * Q + qz) * 12 + k) * Q * Q + qy * Q + qx] = f[k];
}

__global__ void
k_cast_f32(const double * __restrict__ a, float * __restrict__ b, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// Sums the tile partials of DOFs on interior tile faces.  One block per
// (instance, rhs, segment, row); a segment is (component c, plane axis a), a
// row is (plane k, position along the second in-plane axis); the threads
// stride the first in-plane axis.  DOFs on the intersection of two plane
// families are summed by the lower-numbered axis; fixed summation order.
template <typename T, bool MULTI>
__global__ void __launch_bounds__(64)
k_assemble_planes(ApplyDesc d0, const ApplyDesc * __restrict__ descs, int ndesc)
{
  const int w = blockIdx.x;
  int di = 0;
  if constexpr (MULTI)
    di = __builtin_amdgcn_readfirstlane(find_instance(w, ndesc, [&](int i) { return descs[i].pw0; }));
  const ApplyDesc & d = MULTI ? descs[di] : d0;
  const int local = w - d.pw0;
  const int rhs = local / d.prows;
  const int row0 = local - rhs * d.prows;
  if (d.done && d.done[rhs * d.done_stride] != d.gate)
    return;
  const Geom & g = d.g;
  int s = 0;
  while (s + 1 < d.segs.nseg && d.segs.start[s + 1] <= row0)
    s++;
  const int row = row0 - (int)d.segs.start[s];
  const int c = d.segs.comp[s], a = d.segs.axis[s];
  const int b0 = (a == 0) ? 1 : 0;
  const int b1 = (a == 2) ? 1 : 2;
  const int n0 = (int)g.cd[c][b0], n1 = (int)g.cd[c][b1];
  const int k = row / n1 + 1; // plane index 1 .. nt[a]-1
  const int j1 = row % n1;
  const int64_t N = g.off[3];
  const int64_t PS = part_stride(d);
  const T * part = (const T *)d.part + (int64_t)rhs * 4 * PS;
  T * y = (T *)d.y + (int64_t)rhs * N;
  const T * rb = d.resid ? (const T *)d.resid + (int64_t)rhs * N : nullptr;
  const int a2 = 3 - c - a; // the other closed axis of component c
  const int step2 = g.T[a2] * g.p;
  const int dlo = (c == 0) ? 1 : 0;
  const int64_t ebase = d.segs.eoff[c][a] + (int64_t)row * n0;
  for (int j0 = threadIdx.x; j0 < n0; j0 += blockDim.x)
  {
    int gi[3];
    gi[a] = k * g.T[a] * g.p;
    gi[b0] = j0;
    gi[b1] = j1;
    const bool on2 = gi[a2] > 0 && gi[a2] < g.n[a2] * g.p && (gi[a2] % step2) == 0;
    if (on2 && a2 < a)
      continue; // summed by the lower-numbered axis
    bool ess = false;
#pragma unroll
    for (int q = 0; q < 3; q++)
    {
      if (q == c)
        continue;
      const int dim = (int)g.cd[c][q];
      ess = ess || (gi[q] == 0 && ((g.pec >> (2 * q)) & 1)) || (gi[q] == dim - 1 && ((g.pec >> (2 * q + 1)) & 1));
    }
    if (ess)
      continue; // written by the owning tile
    const int64_t gid = g.off[c] + ((int64_t)gi[2] * g.cd[c][1] + gi[1]) * g.cd[c][0] + gi[0];
    // slot bits: bit 0 = lower-numbered closed axis, bit 1 = the other
    const bool on_lo = (a == dlo) ? true : on2;
    const bool on_hi = (a == dlo) ? on2 : true;
    const int64_t pid = ebase + j0; // compact entry (contiguous in j0)
    T v = part[pid];
    if (on_lo)
      v += part[PS + pid];
    if (on_hi)
      v += part[2 * PS + pid];
    if (on_lo && on_hi)
      v += part[3 * PS + pid];
    y[gid] = rb ? rb[gid] - v : v;
  }
}

// ---------------------------------------------------------------------------
// Diagonal: element-parallel in 8 parity colours (elements of one colour share
// no DOF, so each adds its element diagonal without conflicts; the colours run
// in a fixed order).  One workgroup per element with its quadrature factors and
// the 1-D tables in LDS, one thread per local DOF (component-major, x fastest):
//   d_i += sum_q  v^2 M_cc + c1^2 K_11 + c2^2 K_22 + 2 c1 c2 K_12
// with v the DOF's basis function (reference component c) and (c1, c2) the
// two nonzero components of its curl at the point.  PEC DOFs are 1.
// ---------------------------------------------------------------------------
template <int P>
__global__ void __launch_bounds__(256)
k_diag_elem(Geom g, Tables<double> tb, const double * __restrict__ qd, int col, double * __restrict__ diag)
{
  constexpr int O = P, D = P + 1, Q = P + 1, QQ = Q * Q, Q3 = Q * QQ;
  constexpr int ND = 3 * O * D * D; // local DOFs
  __shared__ double sq[12 * Q3];
  __shared__ double sBo[Q * O], sBc[Q * D], sGc[Q * D];
  const int hx = (g.n[0] - (col & 1) + 1) / 2, hy = (g.n[1] - ((col >> 1) & 1) + 1) / 2;
  const int b = blockIdx.x;
  const int ex = 2 * (b % hx) + (col & 1);
  const int ey = 2 * ((b / hx) % hy) + ((col >> 1) & 1);
  const int ez = 2 * (b / (hx * hy)) + ((col >> 2) & 1);
  const int64_t e = ((int64_t)ez * g.n[1] + ey) * g.n[0] + ex;
  const double * qe = qd + e * (12 * Q3);
  for (int i = threadIdx.x; i < 12 * Q3; i += blockDim.x)
    sq[i] = qe[i];
  for (int i = threadIdx.x; i < Q * D; i += blockDim.x)
  {
    sBc[i] = tb.Bc[i];
    sGc[i] = tb.Gc[i];
    if (i < Q * O)
      sBo[i] = tb.Bo[i];
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t >= ND)
    return;
  // local DOF: component c, indices (dx, dy, dz) with O along c, D elsewhere
  const int c = t / (O * D * D);
  const int r = t - c * (O * D * D);
  const int n0 = (c == 0) ? O : D, n1 = (c == 1) ? O : D;
  const int dx = r % n0, dy = (r / n0) % n1, dz = r / (n0 * n1);
  // 1-D factors per axis: value (open or closed basis) and derivative (closed)
  const double * fx = (c == 0) ? sBo + dx : sBc + dx;
  const double * fy = (c == 1) ? sBo + dy : sBc + dy;
  const double * fz = (c == 2) ? sBo + dz : sBc + dz;
  const int sx = (c == 0) ? O : D, sy = (c == 1) ? O : D, sz = (c == 2) ? O : D;
  const int km = (c == 0) ? 0 : ((c == 1) ? 3 : 5);
  const int k1 = (c == 0) ? 9 : 6, k2 = (c == 2) ? 9 : 11, k12 = (c == 0) ? 10 : ((c == 1) ? 8 : 7);
  double s = 0.0;
  for (int qz = 0; qz < Q; qz++)
    for (int qy = 0; qy < Q; qy++)
      for (int qx = 0; qx < Q; qx++)
      {
        const double a = fx[qx * sx], bb = fy[qy * sy], cc = fz[qz * sz];
        const double v = a * bb * cc;
        double c1, c2;
        if (c == 0)
        {
          c1 = a * bb * sGc[qz * D + dz];  // curl_y
          c2 = -a * sGc[qy * D + dy] * cc; // curl_z
        }
        else if (c == 1)
        {
          c1 = -a * bb * sGc[qz * D + dz]; // curl_x
          c2 = sGc[qx * D + dx] * bb * cc; // curl_z
        }
        else
        {
          c1 = a * sGc[qy * D + dy] * cc;  // curl_x
          c2 = -sGc[qx * D + dx] * bb * cc; // curl_y
        }
        const double * f = sq + qz * 12 * QQ + qy * Q + qx;
        s += v * v * f[km * QQ] + c1 * c1 * f[k1 * QQ] + c2 * c2 * f[k2 * QQ] + 2.0 * c1 * c2 * f[k12 * QQ];
      }
  const int64_t gi[3] = {(int64_t)ex * P + dx, (int64_t)ey * P + dy, (int64_t)ez * P + dz};
  const int64_t gid = g.off[c] + (gi[2] * g.cd[c][1] + gi[1]) * g.cd[c][0] + gi[0];
  if (is_ess(g, c, gi[0], gi[1], gi[2]))
    diag[gid] = 1.0;
  else
    diag[gid] += s;
}

__global__ void
k_fill0(double * __restrict__ v, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

PlaneSegs
make_plane_segs(const Geom & g)
{
  // segments (component, plane axis) with start[] = prefix of plane rows
  PlaneSegs s{};
  s.nseg = 0;
  s.start[0] = 0;
  int64_t e = 0;
  for (int c = 0; c < 3; c++)
    for (int a = 0; a < 3; a++)
    {
      if (a == c || g.nt[a] <= 1)
        continue;
      const int b0 = (a == 0) ? 1 : 0;
      const int b1 = (a == 2) ? 1 : 2;
      const int64_t rows = (int64_t)(g.nt[a] - 1) * g.cd[c][b1];
      s.comp[s.nseg] = c;
      s.axis[s.nseg] = a;
      s.eoff[c][a] = e;
      e += rows * g.cd[c][b0];
      s.start[s.nseg + 1] = s.start[s.nseg] + rows;
      s.nseg++;
    }
  return s;
}

void
plane_rows(ApplyDesc & d)
{
  d.segs = make_plane_segs(d.g);
  d.prows = (int)d.segs.start[d.segs.nseg];
}

template <typename T>
hipError_t
assemble_planes(const ApplyDesc & d0, const ApplyDesc * descs, int ndesc, int npwork, hipStream_t st)
{
  if (npwork <= 0)
    return hipSuccess;
  if (descs)
    hipLaunchKernelGGL((k_assemble_planes<T, true>), dim3(npwork), dim3(64), 0, st, d0, descs, ndesc);
  else
    hipLaunchKernelGGL((k_assemble_planes<T, false>), dim3(npwork), dim3(64), 0, st, d0, descs, ndesc);
  return hipGetLastError();
}

template hipError_t assemble_planes<double>(const ApplyDesc &, const ApplyDesc *, int, int, hipStream_t);
template hipError_t assemble_planes<float>(const ApplyDesc &, const ApplyDesc *, int, int, hipStream_t);

hipError_t
pa_setup(const Geom & g, const Tables<double> & tb, const double * vert, const double h[3],
         const double * curl_coef, const double * mass_coef, double curl_scale, double * qd, hipStream_t st)
{
  const int Q = g.p + 1;
  const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2] * Q * Q * Q;
  hipLaunchKernelGGL(k_pa_setup, dim3(div_up(n, 256)), dim3(256), 0, st, g, tb, vert, h[0], h[1], h[2], curl_coef,
                     mass_coef, curl_scale, qd);
  return hipGetLastError();
}

// Element factors of a box mesh: qc[2e] = dt nu_e, qc[2e+1] = sigma_e (and
// their FP32 copies when qc32 is set).
__global__ void
k_elem_factors(int64_t E, const double * __restrict__ nu, const double * __restrict__ sig, double dt,
               double * __restrict__ qc, float * __restrict__ qc32)
{
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x)
  {
    const double a = dt * nu[e], b = sig[e];
    qc[2 * e] = a;
    qc[2 * e + 1] = b;
    if (qc32)
    {
      qc32[2 * e] = (float)a;
      qc32[2 * e + 1] = (float)b;
    }
  }
}

hipError_t
elem_factors(int64_t E, const double * nu, const double * sig, double dt, double * qc, float * qc32, hipStream_t st)
{
  hipLaunchKernelGGL(k_elem_factors, dim3(1024), dim3(256), 0, st, E, nu, sig, dt, qc, qc32);
  return hipGetLastError();
}

hipError_t
cast_f32(const double * a, float * b, int64_t n, hipStream_t st)
{
  hipLaunchKernelGGL(k_cast_f32, dim3(1024), dim3(256), 0, st, a, b, n);
  return hipGetLastError();
}

hipError_t
diagonal(const Geom & g, const Tables<double> & tb, const double * qd, double * d, hipStream_t st)
{
  hipLaunchKernelGGL(k_fill0, dim3(4096), dim3(256), 0, st, d, g.off[3]);
  const int nd = 3 * g.p * (g.p + 1) * (g.p + 1);
  const int nt = (nd + 63) / 64 * 64;
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
      hipLaunchKernelGGL(k_diag_elem<1>, dim3(nb), dim3(nt), 0, st, g, tb, qd, col, d);
      break;
    case 2:
      hipLaunchKernelGGL(k_diag_elem<2>, dim3(nb), dim3(nt), 0, st, g, tb, qd, col, d);
      break;
    case 3:
      hipLaunchKernelGGL(k_diag_elem<3>, dim3(nb), dim3(nt), 0, st, g, tb, qd, col, d);
      break;
    default:
      return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Coil-like right-hand side b_i = int J . phi_i, J = curl psi, psi a sum of
// four Gaussian bumps a_k exp(-|x - c_k|^2 / (2 s_k^2)); Gauss rule with
// p + 2 points per direction.  Elements of one parity colour share no DOF, so
// each colour's launch adds its element vectors without conflicts; colours in
// a fixed order make the result deterministic.  PEC DOFs are zeroed after.
// ---------------------------------------------------------------------------
__global__ void
k_rhs_coil(Geom g, CoilParams cp, const double * __restrict__ vert, double hx, double hy, double hz, double ox,
           double oy, double oz, int col, double * __restrict__ b)
{
  const int p = g.p, O = p, D = p + 1, Qr = p + 2;
  const int cx = col & 1, cy = (col >> 1) & 1, cz = (col >> 2) & 1;
  const int64_t hxn = (g.n[0] - cx + 1) / 2, hyn = (g.n[1] - cy + 1) / 2, hzn = (g.n[2] - cz + 1) / 2;
  const int64_t si = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (si >= hxn * hyn * hzn)
    return;
  const int64_t ex = cx + 2 * (si % hxn), ey = cy + 2 * ((si / hxn) % hyn), ez = cz + 2 * (si / (hxn * hyn));
  double be[3 * PMAX * (PMAX + 1) * (PMAX + 1)];
  const int nd = 3 * p * (p + 1) * (p + 1);
  for (int i = 0; i < nd; i++)
    be[i] = 0.0;
  const int64_t sy = g.n[0] + 1, sz = (int64_t)(g.n[0] + 1) * (g.n[1] + 1);
  for (int qz = 0; qz < Qr; qz++)
    for (int qy = 0; qy < Qr; qy++)
      for (int qx = 0; qx < Qr; qx++)
      {
        const double ta = cp.qp[qx], tb = cp.qp[qy], tc = cp.qp[qz];
        double X[3], J[9];
        if (!vert)
        {
          X[0] = ox + (ex + ta) * hx;
          X[1] = oy + (ey + tb) * hy;
          X[2] = oz + (ez + tc) * hz;
          J[0] = hx;
          J[1] = J[2] = J[3] = 0;
          J[4] = hy;
          J[5] = J[6] = J[7] = 0;
          J[8] = hz;
        }
        else
        {
          const double L[2][3] = {{1 - ta, 1 - tb, 1 - tc}, {ta, tb, tc}};
          const double dL[2] = {-1.0, 1.0};
          for (int i = 0; i < 9; i++)
            J[i] = 0;
          X[0] = X[1] = X[2] = 0;
          for (int kc = 0; kc < 2; kc++)
            for (int kb = 0; kb < 2; kb++)
              for (int ka = 0; ka < 2; ka++)
              {
                const double * V = vert + 3 * ((ex + ka) + (ey + kb) * sy + (ez + kc) * sz);
                const double w0 = L[ka][0] * L[kb][1] * L[kc][2];
                const double g0 = dL[ka] * L[kb][1] * L[kc][2];
                const double g1 = L[ka][0] * dL[kb] * L[kc][2];
                const double g2 = L[ka][0] * L[kb][1] * dL[kc];
                for (int r = 0; r < 3; r++)
                {
                  X[r] += V[r] * w0;
                  J[r * 3 + 0] += V[r] * g0;
                  J[r * 3 + 1] += V[r] * g1;
                  J[r * 3 + 2] += V[r] * g2;
                }
              }
        }
        const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                           J[2] * (J[3] * J[7] - J[4] * J[6]);
        double Jc[3] = {0, 0, 0};
        for (int k = 0; k < 4; k++)
        {
          const double d0 = X[0] - cp.c[k][0], d1 = X[1] - cp.c[k][1], d2 = X[2] - cp.c[k][2];
          const double s2 = cp.s[k] * cp.s[k];
          const double e = exp(-(d0 * d0 + d1 * d1 + d2 * d2) / (2 * s2));
          const double gr0 = -d0 / s2 * e, gr1 = -d1 / s2 * e, gr2 = -d2 / s2 * e;
          // curl(e a) = grad e x a
          Jc[0] += gr1 * cp.a[k][2] - gr2 * cp.a[k][1];
          Jc[1] += gr2 * cp.a[k][0] - gr0 * cp.a[k][2];
          Jc[2] += gr0 * cp.a[k][1] - gr1 * cp.a[k][0];
        }
        // pulled back: J^{-1} Jc |det| w = adj(J) Jc sign(det) w
        double adj[9];
        adj[0] = J[4] * J[8] - J[5] * J[7];
        adj[1] = -(J[1] * J[8] - J[2] * J[7]);
        adj[2] = J[1] * J[5] - J[2] * J[4];
        adj[3] = -(J[3] * J[8] - J[5] * J[6]);
        adj[4] = J[0] * J[8] - J[2] * J[6];
        adj[5] = -(J[0] * J[5] - J[2] * J[3]);
        adj[6] = J[3] * J[7] - J[4] * J[6];
        adj[7] = -(J[0] * J[7] - J[1] * J[6]);
        adj[8] = J[0] * J[4] - J[1] * J[3];
        const double w = cp.qw[qx] * cp.qw[qy] * cp.qw[qz] * (det >= 0 ? 1.0 : -1.0);
        double f[3];
        for (int r = 0; r < 3; r++)
          f[r] = w * (adj[r * 3 + 0] * Jc[0] + adj[r * 3 + 1] * Jc[1] + adj[r * 3 + 2] * Jc[2]);
        int l = 0;
        for (int c = 0; c < 3; c++)
        {
          const int n0 = (c == 0) ? O : D, n1 = (c == 1) ? O : D, n2 = (c == 2) ? O : D;
          for (int dz = 0; dz < n2; dz++)
          {
            const double vz = (c == 2) ? cp.Bo[qz * O + dz] : cp.Bc[qz * D + dz];
            for (int dy = 0; dy < n1; dy++)
            {
              const double vy = (c == 1) ? cp.Bo[qy * O + dy] : cp.Bc[qy * D + dy];
              for (int dx = 0; dx < n0; dx++, l++)
              {
                const double vx = (c == 0) ? cp.Bo[qx * O + dx] : cp.Bc[qx * D + dx];
                be[l] += f[c] * vx * vy * vz;
              }
            }
          }
        }
      }
  int l = 0;
  for (int c = 0; c < 3; c++)
  {
    const int n0 = (c == 0) ? O : D, n1 = (c == 1) ? O : D, n2 = (c == 2) ? O : D;
    for (int dz = 0; dz < n2; dz++)
      for (int dy = 0; dy < n1; dy++)
        for (int dx = 0; dx < n0; dx++, l++)
        {
          const int64_t gid = g.off[c] + ((ez * p + dz) * g.cd[c][1] + (ey * p + dy)) * g.cd[c][0] + ex * p + dx;
          b[gid] += be[l];
        }
  }
}

__global__ void
k_fill_zero(double * __restrict__ v, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = 0.0;
}

__global__ void
k_zero_ess(Geom g, double * __restrict__ v, int64_t sv)
{
  v += blockIdx.y * sv; // right-hand side blockIdx.y
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g.off[3]; i += (int64_t)gridDim.x * blockDim.x)
  {
    int c;
    int64_t a0, a1, a2;
    dof_coords(g, i, c, a0, a1, a2);
    if (is_ess(g, c, a0, a1, a2))
      v[i] = 0.0;
  }
}

hipError_t
rhs_coil(const Geom & g, const CoilParams & cp, const double * vert, const double h[3], const double x0[3],
         double * b, hipStream_t st)
{
  hipLaunchKernelGGL(k_fill_zero, dim3(1024), dim3(256), 0, st, b, g.off[3]);
  for (int col = 0; col < 8; col++)
  {
    const int cx = col & 1, cy = (col >> 1) & 1, cz = (col >> 2) & 1;
    const int64_t nsub = (int64_t)((g.n[0] - cx + 1) / 2) * ((g.n[1] - cy + 1) / 2) * ((g.n[2] - cz + 1) / 2);
    if (nsub == 0)
      continue;
    hipLaunchKernelGGL(k_rhs_coil, dim3(div_up(nsub, 64)), dim3(64), 0, st, g, cp, vert, h[0], h[1], h[2], x0[0],
                       x0[1], x0[2], col, b);
  }
  hipLaunchKernelGGL(k_zero_ess, dim3(1024), dim3(256), 0, st, g, b, (int64_t)0);
  return hipGetLastError();
}

__global__ void
k_axpby(int64_t n, double a, const double * __restrict__ x, double b, double * __restrict__ y)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + (b == 0.0 ? 0.0 : b * y[i]);
}

hipError_t
axpby(int64_t n, double a, const double * x, double b, double * y, hipStream_t st)
{
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_axpby, dim3((unsigned)std::min<int64_t>(div_up(n, 256), 4096)), dim3(256), 0, st, n, a, x,
                     b, y);
  return hipGetLastError();
}

__global__ void
k_mul(int64_t n, const double * __restrict__ d, const double * __restrict__ x, double * __restrict__ y)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = d[i] * x[i];
}

hipError_t
mul(int64_t n, const double * d, const double * x, double * y, hipStream_t st)
{
  if (n <= 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_mul, dim3((unsigned)std::min<int64_t>(div_up(n, 256), 4096)), dim3(256), 0, st, n, d, x, y);
  return hipGetLastError();
}

hipError_t
zero_ess(const Geom & g, double * v, hipStream_t st, int nrhs)
{
  hipLaunchKernelGGL(k_zero_ess, dim3(nrhs > 1 ? std::max(64, 1024 / nrhs) : 1024, nrhs), dim3(256), 0, st, g, v,
                     (int64_t)g.off[3]);
  return hipGetLastError();
}

} // namespace hc
