// CDNA4 (gfx950) kernels of the H(curl) curl-curl + sigma-mass operator:
//   * k_pa_setup      quadrature-point geometric factors (FP64), the partial-
//                     assembly replacement of ParBilinearForm::Assemble as
//                     driven from EquationSystem::BuildBilinearForms
//                     (src/equation_systems/equation_system.cpp:324-347);
//   * k_apply_tile    sum-factorised element apply of dt*nu K + sigma M over a
//                     tile of elements, LDS-staged, deterministic in-tile
//                     assembly; the tile-interior DOFs are final, the DOFs on
//                     tile faces go to a partials buffer;
//   * k_assemble_planes  sums the (2 or 4) tile partials of every tile-face DOF
//                     in a fixed order (bitwise-reproducible results);
//   * k_diagonal      Jacobi diagonal, DOF-parallel gather (no atomics).
// The apply is the operator EquationSystem::Mult (equation_system.cpp:200-204)
// on the PEC-eliminated system of FormLinearSystem (equation_system.cpp:132-190):
// y_free = (A x_free)_free, y_ess = x_ess.
//
// The operator is HBM-bound (about 2-3 flop/byte at p = 3): no MFMA; the tile
// kernel streams the 12 per-quadrature-point factors once, reads each x DOF of
// the tile box once and writes each y DOF once (plus the face partials).
#include "hc_kernels.hpp"

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

// ---------------------------------------------------------------------------
// Tiled sum-factorised apply
// ---------------------------------------------------------------------------

template <int P, int TX, int TY, int TZ>
struct TileCfg
{
  static constexpr int O = P, D = P + 1, Q = P + 1, QQ = Q * Q;
  static constexpr int TE = TX * TY * TZ;
  static constexpr int NP = TE * QQ;
  static constexpr int NTHR = (NP + 63) / 64 * 64;
  static constexpr int NC = O * D * D;
  static constexpr int ND = 3 * NC;
  static constexpr int B00 = TX * P, B01 = TY * P + 1, B02 = TZ * P + 1;
  static constexpr int B10 = TX * P + 1, B11 = TY * P, B12 = TZ * P + 1;
  static constexpr int B20 = TX * P + 1, B21 = TY * P + 1, B22 = TZ * P;
  static constexpr int BOX0 = B00 * B01 * B02, BOX1 = B10 * B11 * B12, BOX2 = B20 * B21 * B22;
  static constexpr int BOX = BOX0 + BOX1 + BOX2;
  static constexpr int cmax(int a, int b) { return a > b ? a : b; }
  static constexpr int SA = cmax(D * D * Q, 2 * D * O * Q);
  static constexpr int SB = cmax(2 * D * QQ, 3 * O * QQ);
  static constexpr int UNI = cmax(BOX, TE * ND);
};

// Decode a box index into (component, l0, l1, l2).
template <class C>
__device__ inline void
box_decode(int i, int & c, int & l0, int & l1, int & l2)
{
  int r, b0, b1;
  if (i < C::BOX0)
  {
    c = 0;
    r = i;
    b0 = C::B00;
    b1 = C::B01;
  }
  else if (i < C::BOX0 + C::BOX1)
  {
    c = 1;
    r = i - C::BOX0;
    b0 = C::B10;
    b1 = C::B11;
  }
  else
  {
    c = 2;
    r = i - C::BOX0 - C::BOX1;
    b0 = C::B20;
    b1 = C::B21;
  }
  l0 = r % b0;
  r /= b0;
  l1 = r % b1;
  l2 = r / b1;
}

template <int P, typename T, int TX, int TY, int TZ>
__global__ void __launch_bounds__((TileCfg<P, TX, TY, TZ>::NTHR))
k_apply_tile(Geom g,
             Tables<T> tb,
             const T * __restrict__ qd,
             const T * __restrict__ x,
             T * __restrict__ y,
             T * __restrict__ part,
             const int * __restrict__ done)
{
  using C = TileCfg<P, TX, TY, TZ>;
  constexpr int O = C::O, D = C::D, Q = C::Q, QQ = C::QQ, SA = C::SA, SB = C::SB;
  constexpr int NT = C::NTHR;
  if (done && *done)
    return;
  __shared__ T sU[C::UNI];
  __shared__ T sA[C::TE * SA];
  __shared__ T sB[C::TE * SB];
  const int tid = threadIdx.x;
  int bid = blockIdx.x;
  const int tx = bid % g.nt[0];
  bid /= g.nt[0];
  const int ty = bid % g.nt[1];
  const int tz = bid / g.nt[1];
  const int e0[3] = {tx * TX, ty * TY, tz * TZ};
  const int64_t N = g.off[3];

  // ---- 1. gather the tile's DOF box (PEC DOFs read as zero) ----
  for (int i = tid; i < C::BOX; i += NT)
  {
    int c, l0, l1, l2;
    box_decode<C>(i, c, l0, l1, l2);
    const int64_t i0 = (int64_t)e0[0] * P + l0, i1 = (int64_t)e0[1] * P + l1, i2 = (int64_t)e0[2] * P + l2;
    T v = T(0);
    if (i0 < g.cd[c][0] && i1 < g.cd[c][1] && i2 < g.cd[c][2] && !is_ess(g, c, i0, i1, i2))
      v = x[g.off[c] + (i2 * g.cd[c][1] + i1) * g.cd[c][0] + i0];
    sU[i] = v;
  }
  __syncthreads();

  // pencil owned by this thread in the quadrature-point stages
  const int pel = tid / QQ;
  const int pqx = (tid % QQ) % Q, pqy = (tid % QQ) / Q;
  const bool pact = tid < C::NP;
  T U[3][Q], Cu[3][Q];
#pragma unroll
  for (int q = 0; q < Q; q++)
    Cu[0][q] = Cu[1][q] = Cu[2][q] = T(0);

  // ---------------- forward: component x, X[dz][dy][dx] (D,D,O) ----------------
  for (int it = tid; it < C::TE * D * D; it += NT)
  {
    const int el = it / (D * D), r = it % (D * D), dy = r % D, dz = r / D;
    const int l0 = el % TX, l1 = (el / TX) % TY, l2 = el / (TX * TY);
    const T * X = sU + ((l2 * P + dz) * C::B01 + (l1 * P + dy)) * C::B00 + l0 * P;
    T xv[O];
#pragma unroll
    for (int dx = 0; dx < O; dx++)
      xv[dx] = X[dx];
#pragma unroll
    for (int qx = 0; qx < Q; qx++)
    {
      T s = 0;
#pragma unroll
      for (int dx = 0; dx < O; dx++)
        s += tb.Bo[qx * O + dx] * xv[dx];
      sA[el * SA + (dz * D + dy) * Q + qx] = s;
    }
  }
  __syncthreads();
  for (int it = tid; it < C::TE * D * Q; it += NT)
  {
    const int el = it / (D * Q), r = it % (D * Q), qx = r % Q, dz = r / Q;
    T a[D];
#pragma unroll
    for (int dy = 0; dy < D; dy++)
      a[dy] = sA[el * SA + (dz * D + dy) * Q + qx];
#pragma unroll
    for (int qy = 0; qy < Q; qy++)
    {
      T s = 0, gg = 0;
#pragma unroll
      for (int dy = 0; dy < D; dy++)
      {
        s += tb.Bc[qy * D + dy] * a[dy];
        gg += tb.Gc[qy * D + dy] * a[dy];
      }
      sB[el * SB + (dz * Q + qy) * Q + qx] = s;
      sB[el * SB + D * QQ + (dz * Q + qy) * Q + qx] = gg;
    }
  }
  __syncthreads();
  if (pact)
  {
    T b1[D], b2[D];
#pragma unroll
    for (int dz = 0; dz < D; dz++)
    {
      b1[dz] = sB[pel * SB + (dz * Q + pqy) * Q + pqx];
      b2[dz] = sB[pel * SB + D * QQ + (dz * Q + pqy) * Q + pqx];
    }
#pragma unroll
    for (int qz = 0; qz < Q; qz++)
    {
      T u = 0, uz = 0, uy = 0;
#pragma unroll
      for (int dz = 0; dz < D; dz++)
      {
        u += tb.Bc[qz * D + dz] * b1[dz];
        uz += tb.Gc[qz * D + dz] * b1[dz];
        uy += tb.Bc[qz * D + dz] * b2[dz];
      }
      U[0][qz] = u;
      Cu[1][qz] += uz;
      Cu[2][qz] -= uy;
    }
  }
  __syncthreads();

  // ---------------- forward: component y, Y[dz][dy][dx] (D,O,D) ----------------
  for (int it = tid; it < C::TE * D * O; it += NT)
  {
    const int el = it / (D * O), r = it % (D * O), dy = r % O, dz = r / O;
    const int l0 = el % TX, l1 = (el / TX) % TY, l2 = el / (TX * TY);
    const T * Y = sU + C::BOX0 + ((l2 * P + dz) * C::B11 + (l1 * P + dy)) * C::B10 + l0 * P;
    T yv[D];
#pragma unroll
    for (int dx = 0; dx < D; dx++)
      yv[dx] = Y[dx];
#pragma unroll
    for (int qx = 0; qx < Q; qx++)
    {
      T s = 0, gg = 0;
#pragma unroll
      for (int dx = 0; dx < D; dx++)
      {
        s += tb.Bc[qx * D + dx] * yv[dx];
        gg += tb.Gc[qx * D + dx] * yv[dx];
      }
      sA[el * SA + (dz * O + dy) * Q + qx] = s;
      sA[el * SA + D * O * Q + (dz * O + dy) * Q + qx] = gg;
    }
  }
  __syncthreads();
  for (int it = tid; it < C::TE * D * Q; it += NT)
  {
    const int el = it / (D * Q), r = it % (D * Q), qx = r % Q, dz = r / Q;
    T a1[O], a2[O];
#pragma unroll
    for (int dy = 0; dy < O; dy++)
    {
      a1[dy] = sA[el * SA + (dz * O + dy) * Q + qx];
      a2[dy] = sA[el * SA + D * O * Q + (dz * O + dy) * Q + qx];
    }
#pragma unroll
    for (int qy = 0; qy < Q; qy++)
    {
      T s = 0, gg = 0;
#pragma unroll
      for (int dy = 0; dy < O; dy++)
      {
        s += tb.Bo[qy * O + dy] * a1[dy];
        gg += tb.Bo[qy * O + dy] * a2[dy];
      }
      sB[el * SB + (dz * Q + qy) * Q + qx] = s;
      sB[el * SB + D * QQ + (dz * Q + qy) * Q + qx] = gg;
    }
  }
  __syncthreads();
  if (pact)
  {
    T b1[D], b2[D];
#pragma unroll
    for (int dz = 0; dz < D; dz++)
    {
      b1[dz] = sB[pel * SB + (dz * Q + pqy) * Q + pqx];
      b2[dz] = sB[pel * SB + D * QQ + (dz * Q + pqy) * Q + pqx];
    }
#pragma unroll
    for (int qz = 0; qz < Q; qz++)
    {
      T u = 0, uz = 0, ux = 0;
#pragma unroll
      for (int dz = 0; dz < D; dz++)
      {
        u += tb.Bc[qz * D + dz] * b1[dz];
        uz += tb.Gc[qz * D + dz] * b1[dz];
        ux += tb.Bc[qz * D + dz] * b2[dz];
      }
      U[1][qz] = u;
      Cu[0][qz] -= uz;
      Cu[2][qz] += ux;
    }
  }
  __syncthreads();

  // ---------------- forward: component z, Z[dz][dy][dx] (O,D,D) ----------------
  for (int it = tid; it < C::TE * O * D; it += NT)
  {
    const int el = it / (O * D), r = it % (O * D), dy = r % D, dz = r / D;
    const int l0 = el % TX, l1 = (el / TX) % TY, l2 = el / (TX * TY);
    const T * Z = sU + C::BOX0 + C::BOX1 + ((l2 * P + dz) * C::B21 + (l1 * P + dy)) * C::B20 + l0 * P;
    T zv[D];
#pragma unroll
    for (int dx = 0; dx < D; dx++)
      zv[dx] = Z[dx];
#pragma unroll
    for (int qx = 0; qx < Q; qx++)
    {
      T s = 0, gg = 0;
#pragma unroll
      for (int dx = 0; dx < D; dx++)
      {
        s += tb.Bc[qx * D + dx] * zv[dx];
        gg += tb.Gc[qx * D + dx] * zv[dx];
      }
      sA[el * SA + (dz * D + dy) * Q + qx] = s;
      sA[el * SA + O * D * Q + (dz * D + dy) * Q + qx] = gg;
    }
  }
  __syncthreads();
  for (int it = tid; it < C::TE * O * Q; it += NT)
  {
    const int el = it / (O * Q), r = it % (O * Q), qx = r % Q, dz = r / Q;
    T a1[D], a2[D];
#pragma unroll
    for (int dy = 0; dy < D; dy++)
    {
      a1[dy] = sA[el * SA + (dz * D + dy) * Q + qx];
      a2[dy] = sA[el * SA + O * D * Q + (dz * D + dy) * Q + qx];
    }
#pragma unroll
    for (int qy = 0; qy < Q; qy++)
    {
      T s = 0, gy = 0, gx = 0;
#pragma unroll
      for (int dy = 0; dy < D; dy++)
      {
        s += tb.Bc[qy * D + dy] * a1[dy];
        gy += tb.Gc[qy * D + dy] * a1[dy];
        gx += tb.Bc[qy * D + dy] * a2[dy];
      }
      sB[el * SB + (dz * Q + qy) * Q + qx] = s;
      sB[el * SB + O * QQ + (dz * Q + qy) * Q + qx] = gy;
      sB[el * SB + 2 * O * QQ + (dz * Q + qy) * Q + qx] = gx;
    }
  }
  __syncthreads();
  if (pact)
  {
    T b1[O], b2[O], b3[O];
#pragma unroll
    for (int dz = 0; dz < O; dz++)
    {
      b1[dz] = sB[pel * SB + (dz * Q + pqy) * Q + pqx];
      b2[dz] = sB[pel * SB + O * QQ + (dz * Q + pqy) * Q + pqx];
      b3[dz] = sB[pel * SB + 2 * O * QQ + (dz * Q + pqy) * Q + pqx];
    }
#pragma unroll
    for (int qz = 0; qz < Q; qz++)
    {
      T u = 0, uy = 0, ux = 0;
#pragma unroll
      for (int dz = 0; dz < O; dz++)
      {
        u += tb.Bo[qz * O + dz] * b1[dz];
        uy += tb.Bo[qz * O + dz] * b2[dz];
        ux += tb.Bo[qz * O + dz] * b3[dz];
      }
      U[2][qz] = u;
      Cu[0][qz] += uy;
      Cu[1][qz] -= ux;
    }

    // ---------------- quadrature-point factors ----------------
    const int l0 = pel % TX, l1 = (pel / TX) % TY, l2 = pel / (TX * TY);
    const int64_t ge0 = e0[0] + l0, ge1 = e0[1] + l1, ge2 = e0[2] + l2;
    if (ge0 < g.n[0] && ge1 < g.n[1] && ge2 < g.n[2])
    {
      const int64_t e = (ge2 * g.n[1] + ge1) * g.n[0] + ge0;
      const T * qe = qd + e * (12 * Q * QQ) + pqy * Q + pqx;
#pragma unroll
      for (int qz = 0; qz < Q; qz++)
      {
        const T * qq = qe + qz * 12 * QQ;
        const T m0 = qq[0 * QQ], m1 = qq[1 * QQ], m2 = qq[2 * QQ], m3 = qq[3 * QQ], m4 = qq[4 * QQ],
                m5 = qq[5 * QQ];
        const T k0 = qq[6 * QQ], k1 = qq[7 * QQ], k2 = qq[8 * QQ], k3 = qq[9 * QQ], k4 = qq[10 * QQ],
                k5 = qq[11 * QQ];
        const T u0 = U[0][qz], u1 = U[1][qz], u2 = U[2][qz];
        const T c0 = Cu[0][qz], c1 = Cu[1][qz], c2 = Cu[2][qz];
        U[0][qz] = m0 * u0 + m1 * u1 + m2 * u2;
        U[1][qz] = m1 * u0 + m3 * u1 + m4 * u2;
        U[2][qz] = m2 * u0 + m4 * u1 + m5 * u2;
        Cu[0][qz] = k0 * c0 + k1 * c1 + k2 * c2;
        Cu[1][qz] = k1 * c0 + k3 * c1 + k4 * c2;
        Cu[2][qz] = k2 * c0 + k4 * c1 + k5 * c2;
      }
    }
    else
    {
#pragma unroll
      for (int qz = 0; qz < Q; qz++)
        U[0][qz] = U[1][qz] = U[2][qz] = Cu[0][qz] = Cu[1][qz] = Cu[2][qz] = T(0);
    }

    // ---------------- backward z^T, component x ----------------
#pragma unroll
    for (int dz = 0; dz < D; dz++)
    {
      T s1 = 0, s2 = 0;
#pragma unroll
      for (int qz = 0; qz < Q; qz++)
      {
        s1 += tb.Bc[qz * D + dz] * U[0][qz] + tb.Gc[qz * D + dz] * Cu[1][qz];
        s2 -= tb.Bc[qz * D + dz] * Cu[2][qz];
      }
      sB[pel * SB + (dz * Q + pqy) * Q + pqx] = s1;
      sB[pel * SB + D * QQ + (dz * Q + pqy) * Q + pqx] = s2;
    }
  }
  __syncthreads();
  // y^T, component x
  for (int it = tid; it < C::TE * D * Q; it += NT)
  {
    const int el = it / (D * Q), r = it % (D * Q), qx = r % Q, dz = r / Q;
    T s1[Q], s2[Q];
#pragma unroll
    for (int qy = 0; qy < Q; qy++)
    {
      s1[qy] = sB[el * SB + (dz * Q + qy) * Q + qx];
      s2[qy] = sB[el * SB + D * QQ + (dz * Q + qy) * Q + qx];
    }
#pragma unroll
    for (int dy = 0; dy < D; dy++)
    {
      T t = 0;
#pragma unroll
      for (int qy = 0; qy < Q; qy++)
        t += tb.Bc[qy * D + dy] * s1[qy] + tb.Gc[qy * D + dy] * s2[qy];
      sA[el * SA + (dz * D + dy) * Q + qx] = t;
    }
  }
  __syncthreads();
  // x^T, component x -> element vectors (sU reused)
  for (int it = tid; it < C::TE * D * D; it += NT)
  {
    const int el = it / (D * D), r = it % (D * D), dy = r % D, dz = r / D;
    T t[Q];
#pragma unroll
    for (int qx = 0; qx < Q; qx++)
      t[qx] = sA[el * SA + (dz * D + dy) * Q + qx];
#pragma unroll
    for (int dx = 0; dx < O; dx++)
    {
      T s = 0;
#pragma unroll
      for (int qx = 0; qx < Q; qx++)
        s += tb.Bo[qx * O + dx] * t[qx];
      sU[el * C::ND + (dz * D + dy) * O + dx] = s;
    }
  }
  // z^T, component y (writes sB only; the x^T stage above reads sA only)
  if (pact)
  {
#pragma unroll
    for (int dz = 0; dz < D; dz++)
    {
      T s1 = 0, s2 = 0;
#pragma unroll
      for (int qz = 0; qz < Q; qz++)
      {
        s1 += tb.Bc[qz * D + dz] * U[1][qz] - tb.Gc[qz * D + dz] * Cu[0][qz];
        s2 += tb.Bc[qz * D + dz] * Cu[2][qz];
      }
      sB[pel * SB + (dz * Q + pqy) * Q + pqx] = s1;
      sB[pel * SB + D * QQ + (dz * Q + pqy) * Q + pqx] = s2;
    }
  }
  __syncthreads();
  // y^T, component y
  for (int it = tid; it < C::TE * D * Q; it += NT)
  {
    const int el = it / (D * Q), r = it % (D * Q), qx = r % Q, dz = r / Q;
    T s1[Q], s2[Q];
#pragma unroll
    for (int qy = 0; qy < Q; qy++)
    {
      s1[qy] = sB[el * SB + (dz * Q + qy) * Q + qx];
      s2[qy] = sB[el * SB + D * QQ + (dz * Q + qy) * Q + qx];
    }
#pragma unroll
    for (int dy = 0; dy < O; dy++)
    {
      T ta = 0, tb2 = 0;
#pragma unroll
      for (int qy = 0; qy < Q; qy++)
      {
        ta += tb.Bo[qy * O + dy] * s1[qy];
        tb2 += tb.Bo[qy * O + dy] * s2[qy];
      }
      sA[el * SA + (dz * O + dy) * Q + qx] = ta;
      sA[el * SA + D * O * Q + (dz * O + dy) * Q + qx] = tb2;
    }
  }
  __syncthreads();
  // x^T, component y
  for (int it = tid; it < C::TE * D * O; it += NT)
  {
    const int el = it / (D * O), r = it % (D * O), dy = r % O, dz = r / O;
    T ta[Q], tb2[Q];
#pragma unroll
    for (int qx = 0; qx < Q; qx++)
    {
      ta[qx] = sA[el * SA + (dz * O + dy) * Q + qx];
      tb2[qx] = sA[el * SA + D * O * Q + (dz * O + dy) * Q + qx];
    }
#pragma unroll
    for (int dx = 0; dx < D; dx++)
    {
      T s = 0;
#pragma unroll
      for (int qx = 0; qx < Q; qx++)
        s += tb.Bc[qx * D + dx] * ta[qx] + tb.Gc[qx * D + dx] * tb2[qx];
      sU[el * C::ND + C::NC + (dz * O + dy) * D + dx] = s;
    }
  }
  // z^T, component z
  if (pact)
  {
#pragma unroll
    for (int dz = 0; dz < O; dz++)
    {
      T sa = 0, sb = 0, sc = 0;
#pragma unroll
      for (int qz = 0; qz < Q; qz++)
      {
        const T bo = tb.Bo[qz * O + dz];
        sa += bo * U[2][qz];
        sb += bo * Cu[0][qz];
        sc -= bo * Cu[1][qz];
      }
      sB[pel * SB + (dz * Q + pqy) * Q + pqx] = sa;
      sB[pel * SB + O * QQ + (dz * Q + pqy) * Q + pqx] = sb;
      sB[pel * SB + 2 * O * QQ + (dz * Q + pqy) * Q + pqx] = sc;
    }
  }
  __syncthreads();
  // y^T, component z
  for (int it = tid; it < C::TE * O * Q; it += NT)
  {
    const int el = it / (O * Q), r = it % (O * Q), qx = r % Q, dz = r / Q;
    T sa[Q], sb[Q], sc[Q];
#pragma unroll
    for (int qy = 0; qy < Q; qy++)
    {
      sa[qy] = sB[el * SB + (dz * Q + qy) * Q + qx];
      sb[qy] = sB[el * SB + O * QQ + (dz * Q + qy) * Q + qx];
      sc[qy] = sB[el * SB + 2 * O * QQ + (dz * Q + qy) * Q + qx];
    }
#pragma unroll
    for (int dy = 0; dy < D; dy++)
    {
      T ta = 0, tb2 = 0;
#pragma unroll
      for (int qy = 0; qy < Q; qy++)
      {
        ta += tb.Bc[qy * D + dy] * sa[qy] + tb.Gc[qy * D + dy] * sb[qy];
        tb2 += tb.Bc[qy * D + dy] * sc[qy];
      }
      sA[el * SA + (dz * D + dy) * Q + qx] = ta;
      sA[el * SA + O * D * Q + (dz * D + dy) * Q + qx] = tb2;
    }
  }
  __syncthreads();
  // x^T, component z
  for (int it = tid; it < C::TE * O * D; it += NT)
  {
    const int el = it / (O * D), r = it % (O * D), dy = r % D, dz = r / D;
    T ta[Q], tb2[Q];
#pragma unroll
    for (int qx = 0; qx < Q; qx++)
    {
      ta[qx] = sA[el * SA + (dz * D + dy) * Q + qx];
      tb2[qx] = sA[el * SA + O * D * Q + (dz * D + dy) * Q + qx];
    }
#pragma unroll
    for (int dx = 0; dx < D; dx++)
    {
      T s = 0;
#pragma unroll
      for (int qx = 0; qx < Q; qx++)
        s += tb.Bc[qx * D + dx] * ta[qx] + tb.Gc[qx * D + dx] * tb2[qx];
      sU[el * C::ND + 2 * C::NC + (dz * D + dy) * D + dx] = s;
    }
  }
  __syncthreads();

  // ---- 5. in-tile assembly (fixed order) and write-out ----
  for (int i = tid; i < C::BOX; i += NT)
  {
    int c, l[3];
    box_decode<C>(i, c, l[0], l[1], l[2]);
    const int64_t gi[3] = {(int64_t)e0[0] * P + l[0], (int64_t)e0[1] * P + l[1], (int64_t)e0[2] * P + l[2]};
    if (gi[0] >= g.cd[c][0] || gi[1] >= g.cd[c][1] || gi[2] >= g.cd[c][2])
      continue;
    const int64_t gid = g.off[c] + (gi[2] * g.cd[c][1] + gi[1]) * g.cd[c][0] + gi[0];
    // tile-face membership in the two closed directions
    bool onp[3] = {false, false, false};
    bool owner = true;
    int slot = 0, bit = 1;
#pragma unroll
    for (int a = 0; a < 3; a++)
    {
      if (a == c)
        continue;
      onp[a] = on_tile_plane(g, a, gi[a]);
      if (onp[a] && l[a] != 0)
      {
        owner = false;
        slot |= bit;
      }
      bit <<= 1;
    }
    if (is_ess(g, c, gi[0], gi[1], gi[2]))
    {
      if (owner)
        y[gid] = x[gid];
      continue;
    }
    // elements of the tile containing this DOF, per axis (at most 2)
    int ne[3], el_[3][2], lc[3][2];
    const int Tt[3] = {TX, TY, TZ};
#pragma unroll
    for (int a = 0; a < 3; a++)
    {
      ne[a] = 0;
      if (a == c)
      {
        el_[a][0] = l[a] / P;
        lc[a][0] = l[a] % P;
        ne[a] = 1;
      }
      else
      {
        const int eh = l[a] / P, lh = l[a] % P;
        if (lh == 0 && eh > 0)
        {
          el_[a][ne[a]] = eh - 1;
          lc[a][ne[a]] = P;
          ne[a]++;
        }
        if (eh < Tt[a] && e0[a] + eh < g.n[a])
        {
          el_[a][ne[a]] = eh;
          lc[a][ne[a]] = lh;
          ne[a]++;
        }
      }
    }
    T s = 0;
    for (int kz = 0; kz < ne[2]; kz++)
      for (int ky = 0; ky < ne[1]; ky++)
        for (int kx = 0; kx < ne[0]; kx++)
        {
          const int el = (el_[2][kz] * TY + el_[1][ky]) * TX + el_[0][kx];
          const int d0 = lc[0][kx], d1 = lc[1][ky], d2 = lc[2][kz];
          int li;
          if (c == 0)
            li = (d2 * D + d1) * O + d0;
          else if (c == 1)
            li = C::NC + (d2 * O + d1) * D + d0;
          else
            li = 2 * C::NC + (d2 * D + d1) * D + d0;
          s += sU[el * C::ND + li];
        }
    if (!(onp[0] || onp[1] || onp[2]))
      y[gid] = s;
    else
      part[slot * N + gid] = s;
  }
}

// Sums the tile partials of DOFs on interior tile faces.  Thread space:
// segments (component c, closed axis a) x (nt[a]-1 planes) x in-plane DOFs.
template <typename T>
__global__ void
k_assemble_planes(Geom g, PlaneSegs segs, const T * __restrict__ part, T * __restrict__ y, const int * __restrict__ done)
{
  if (done && *done)
    return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= segs.start[segs.nseg])
    return;
  int s = 0;
  while (t >= segs.start[s + 1])
    s++;
  const int c = segs.comp[s], a = segs.axis[s];
  const int b0 = (a == 0) ? 1 : 0;
  const int b1 = (a == 2) ? 1 : 2;
  int64_t r = t - segs.start[s];
  const int64_t n0 = g.cd[c][b0], n1 = g.cd[c][b1];
  const int64_t j0 = r % n0;
  r /= n0;
  const int64_t j1 = r % n1;
  const int64_t k = r / n1 + 1;
  int64_t gi[3];
  gi[a] = k * g.T[a] * g.p;
  gi[b0] = j0;
  gi[b1] = j1;
  // the other closed axis
  const int a2 = 3 - c - a;
  const bool on2 = on_tile_plane(g, a2, gi[a2]);
  if (on2 && a2 < a)
    return; // enumerated by the lower-numbered axis
  if (is_ess(g, c, gi[0], gi[1], gi[2]))
    return;
  const int64_t gid = g.off[c] + (gi[2] * g.cd[c][1] + gi[1]) * g.cd[c][0] + gi[0];
  const int64_t N = g.off[3];
  // slot bits: bit 0 = lower-numbered closed axis, bit 1 = the other
  const int dlo = (c == 0) ? 1 : 0;
  const bool on_lo = (a == dlo) ? true : on2;
  const bool on_hi = (a == dlo) ? on2 : true;
  T v = part[gid];
  if (on_lo)
    v += part[N + gid];
  if (on_hi)
    v += part[2 * N + gid];
  if (on_lo && on_hi)
    v += part[3 * N + gid];
  y[gid] = v;
}

// ---------------------------------------------------------------------------
// Diagonal (DOF-parallel gather over the <= 4 elements of each DOF)
// ---------------------------------------------------------------------------
__global__ void
k_diagonal(Geom g, Tables<double> tb, const double * __restrict__ qd, double * __restrict__ diag)
{
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= g.off[3])
    return;
  int c;
  int64_t gi[3];
  dof_coords(g, gid, c, gi[0], gi[1], gi[2]);
  if (is_ess(g, c, gi[0], gi[1], gi[2]))
  {
    diag[gid] = 1.0;
    return;
  }
  const int p = g.p, O = p, D = p + 1, Q = p + 1, QQ = Q * Q;
  int ne[3];
  int64_t ec[3][2];
  int lc[3][2];
  for (int a = 0; a < 3; a++)
  {
    ne[a] = 0;
    if (a == c)
    {
      ec[a][0] = gi[a] / p;
      lc[a][0] = (int)(gi[a] % p);
      ne[a] = 1;
    }
    else
    {
      const int64_t eh = gi[a] / p;
      const int lh = (int)(gi[a] % p);
      if (lh == 0 && eh > 0)
      {
        ec[a][ne[a]] = eh - 1;
        lc[a][ne[a]] = p;
        ne[a]++;
      }
      if (eh < g.n[a])
      {
        ec[a][ne[a]] = eh;
        lc[a][ne[a]] = lh;
        ne[a]++;
      }
    }
  }
  double s = 0.0;
  for (int kz = 0; kz < ne[2]; kz++)
    for (int ky = 0; ky < ne[1]; ky++)
      for (int kx = 0; kx < ne[0]; kx++)
      {
        const int64_t e = (ec[2][kz] * g.n[1] + ec[1][ky]) * g.n[0] + ec[0][kx];
        const int dx = lc[0][kx], dy = lc[1][ky], dz = lc[2][kz];
        const double * qe = qd + e * 12 * Q * QQ;
        for (int qz = 0; qz < Q; qz++)
          for (int qy = 0; qy < Q; qy++)
            for (int qx = 0; qx < Q; qx++)
            {
              const double * qq = qe + qz * 12 * QQ + qy * Q + qx;
              double v, c1, c2;
              int k1, k2, k12, km;
              if (c == 0)
              {
                const double a = tb.Bo[qx * O + dx], b = tb.Bc[qy * D + dy], bp = tb.Gc[qy * D + dy];
                const double cc = tb.Bc[qz * D + dz], cp = tb.Gc[qz * D + dz];
                v = a * b * cc;
                c1 = a * b * cp;   // curl_y
                c2 = -a * bp * cc; // curl_z
                km = 0;
                k1 = 9;
                k2 = 11;
                k12 = 10;
              }
              else if (c == 1)
              {
                const double a = tb.Bc[qx * D + dx], ap = tb.Gc[qx * D + dx], b = tb.Bo[qy * O + dy];
                const double cc = tb.Bc[qz * D + dz], cp = tb.Gc[qz * D + dz];
                v = a * b * cc;
                c1 = -a * b * cp; // curl_x
                c2 = ap * b * cc; // curl_z
                km = 3;
                k1 = 6;
                k2 = 11;
                k12 = 8;
              }
              else
              {
                const double a = tb.Bc[qx * D + dx], ap = tb.Gc[qx * D + dx];
                const double b = tb.Bc[qy * D + dy], bp = tb.Gc[qy * D + dy], cc = tb.Bo[qz * O + dz];
                v = a * b * cc;
                c1 = a * bp * cc;  // curl_x
                c2 = -ap * b * cc; // curl_y
                km = 5;
                k1 = 6;
                k2 = 9;
                k12 = 7;
              }
              s += v * v * qq[km * QQ] + c1 * c1 * qq[k1 * QQ] + c2 * c2 * qq[k2 * QQ] +
                   2.0 * c1 * c2 * qq[k12 * QQ];
            }
      }
  diag[gid] = s;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

template <int P, int TX, int TY, int TZ>
static TileShape
shape_of()
{
  return TileShape{TX, TY, TZ, TileCfg<P, TX, TY, TZ>::NTHR};
}

TileShape
default_tile(int p)
{
  switch (p)
  {
    case 1:
      return shape_of<1, 4, 4, 4>();
    case 2:
      return shape_of<2, 4, 4, 2>();
    default:
      return shape_of<3, 4, 2, 2>();
  }
}

template <typename T>
static hipError_t
launch_tile(const Geom & g, const Tables<T> & tb, const T * qd, const T * x, T * y, T * part,
            const int * done, hipStream_t st)
{
  const int nblk = g.nt[0] * g.nt[1] * g.nt[2];
  switch (g.p)
  {
    case 1:
    {
      using C = TileCfg<1, 4, 4, 4>;
      hipLaunchKernelGGL((k_apply_tile<1, T, 4, 4, 4>), dim3(nblk), dim3(C::NTHR), 0, st, g, tb, qd, x, y, part, done);
      break;
    }
    case 2:
    {
      using C = TileCfg<2, 4, 4, 2>;
      hipLaunchKernelGGL((k_apply_tile<2, T, 4, 4, 2>), dim3(nblk), dim3(C::NTHR), 0, st, g, tb, qd, x, y, part, done);
      break;
    }
    case 3:
    {
      using C = TileCfg<3, 4, 2, 2>;
      hipLaunchKernelGGL((k_apply_tile<3, T, 4, 2, 2>), dim3(nblk), dim3(C::NTHR), 0, st, g, tb, qd, x, y, part, done);
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

PlaneSegs
make_plane_segs(const Geom & g)
{
  PlaneSegs s{};
  s.nseg = 0;
  s.start[0] = 0;
  for (int c = 0; c < 3; c++)
    for (int a = 0; a < 3; a++)
    {
      if (a == c || g.nt[a] <= 1)
        continue;
      const int b0 = (a == 0) ? 1 : 0;
      const int b1 = (a == 2) ? 1 : 2;
      const int64_t cnt = (int64_t)(g.nt[a] - 1) * g.cd[c][b0] * g.cd[c][b1];
      s.comp[s.nseg] = c;
      s.axis[s.nseg] = a;
      s.start[s.nseg + 1] = s.start[s.nseg] + cnt;
      s.nseg++;
    }
  return s;
}

template <typename T>
hipError_t
apply_op(const Geom & g, const PlaneSegs & segs, const Tables<T> & tb, const T * qd, const T * x, T * y,
         T * part, const int * done, hipStream_t st)
{
  hipError_t e = launch_tile<T>(g, tb, qd, x, y, part, done, st);
  if (e != hipSuccess)
    return e;
  const int64_t n = segs.start[segs.nseg];
  if (n > 0)
  {
    hipLaunchKernelGGL(k_assemble_planes<T>, dim3(div_up(n, 256)), dim3(256), 0, st, g, segs, part, y, done);
    e = hipGetLastError();
  }
  return e;
}

template hipError_t apply_op<double>(const Geom &, const PlaneSegs &, const Tables<double> &, const double *,
                                     const double *, double *, double *, const int *, hipStream_t);
template hipError_t apply_op<float>(const Geom &, const PlaneSegs &, const Tables<float> &, const float *,
                                    const float *, float *, float *, const int *, hipStream_t);

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

hipError_t
cast_f32(const double * a, float * b, int64_t n, hipStream_t st)
{
  hipLaunchKernelGGL(k_cast_f32, dim3(1024), dim3(256), 0, st, a, b, n);
  return hipGetLastError();
}

hipError_t
diagonal(const Geom & g, const Tables<double> & tb, const double * qd, double * d, hipStream_t st)
{
  hipLaunchKernelGGL(k_diagonal, dim3(div_up(g.off[3], 256)), dim3(256), 0, st, g, tb, qd, d);
  return hipGetLastError();
}

} // namespace hc
