// Tiled sum-factorised apply of dt*nu K + sigma M on ND_p hexes (gfx950).
//
// This is the operator EquationSystem::Mult (src/equation_systems/equation_system.cpp:200-204)
// on the system formed by FormLinearSystem (equation_system.cpp:132-190): PEC
// rows/cols eliminated, y_free = (A x_free)_free, y_ess = x_ess.  The element
// arithmetic is that of MFEM's CurlCurlIntegrator + VectorFEMassIntegrator
// (src/kernels/curl_curl_kernel.cpp:23, src/kernels/vector_fe_mass_kernel.cpp:23)
// in partial-assembly form: 6 + 6 symmetric factors per quadrature point.
//
// One workgroup owns a tile of TX x TY x TZ elements, one thread per
// (element, qx, qy) pencil.  Per vector component the forward contraction runs
// x (HBM rows -> LDS A) -> y (LDS A -> LDS B) -> z (LDS B -> registers); the
// quadrature-point factors (prefetched into registers at entry) are applied in
// registers; the backward contraction runs z^T (registers -> LDS S) -> y^T
// (LDS -> LDS T) -> a row stage that sums the T tensors of the (at most four)
// elements sharing a DOF row in a fixed order, contracts along x and writes the
// result.  DOFs interior to the tile are final; DOFs on a face shared with a
// neighbouring tile go to a partials buffer that k_assemble_planes sums in a
// fixed order -- results are bitwise reproducible.  LDS strides are padded so
// that the 16 lanes of each element hit distinct banks.
#include "hc_kernels.hpp"

#include <type_traits>

namespace hc
{

template <int P, int TX, int TY, int TZ>
struct Tile
{
  static constexpr int O = P, D = P + 1, Q = P + 1, QQ = Q * Q;
  static constexpr int TE = TX * TY * TZ;
  static constexpr int NP = TE * QQ;
  static constexpr int NT = (NP + 63) / 64 * 64;
  // A-type arrays [qx][dz][dy], qx stride padded
  static constexpr int SQX = D * D + 1;
  static constexpr int LA = Q * SQX;
  // B-type arrays [dz][qy][qx], dz stride padded
  static constexpr int SDZ = QQ + 4;
  static constexpr int LB = D * SDZ;
  static constexpr int pad_to(int v, int m, int r) { return ((v - r + m - 1) / m) * m + r; }
  static constexpr int SA = pad_to(2 * LA, 16, 4);
  static constexpr int SB = pad_to(3 * LB, 16, 4);
};

template <int P, typename T, int TX, int TY, int TZ>
__global__ void __launch_bounds__((Tile<P, TX, TY, TZ>::NT))
k_apply_tile(Geom g,
             Tables<T> tb,
             const T * __restrict__ qd,
             const T * __restrict__ x,
             T * __restrict__ y,
             T * __restrict__ part,
             const int * __restrict__ done)
{
  using C = Tile<P, TX, TY, TZ>;
  constexpr int O = C::O, D = C::D, Q = C::Q, QQ = C::QQ;
  constexpr int SQX = C::SQX, LA = C::LA, SDZ = C::SDZ, LB = C::LB, SA = C::SA, SB = C::SB;
  constexpr int NT = C::NT, TE = C::TE;
  if (done && *done)
    return;
  __shared__ T sA[TE * SA];
  __shared__ T sB[TE * SB];

  const int tid = threadIdx.x;
  int bid = blockIdx.x;
  const int tpos[3] = {bid % g.nt[0], (bid / g.nt[0]) % g.nt[1], bid / (g.nt[0] * g.nt[1])};
  const int e0[3] = {tpos[0] * TX, tpos[1] * TY, tpos[2] * TZ};
  // active extent of the tile (partial tiles at the upper domain faces)
  const int ne[3] = {min(TX, g.n[0] - e0[0]), min(TY, g.n[1] - e0[1]), min(TZ, g.n[2] - e0[2])};
  const int64_t N = g.off[3];
  const int cd[3][3] = {{(int)g.cd[0][0], (int)g.cd[0][1], (int)g.cd[0][2]},
                        {(int)g.cd[1][0], (int)g.cd[1][1], (int)g.cd[1][2]},
                        {(int)g.cd[2][0], (int)g.cd[2][1], (int)g.cd[2][2]}};
  const int pec = g.pec;

  // ---- pencil of this thread; prefetch its quadrature data ----
  const int pel = tid / QQ;
  const int pqx = (tid % QQ) % Q, pqy = (tid % QQ) / Q;
  const int pl0 = pel % TX, pl1 = (pel / TX) % TY, pl2 = pel / (TX * TY);
  const bool pin = tid < C::NP && pl0 < ne[0] && pl1 < ne[1] && pl2 < ne[2];
  T qf[Q][12];
  if (pin)
  {
    const int64_t e = ((int64_t)(e0[2] + pl2) * g.n[1] + (e0[1] + pl1)) * g.n[0] + (e0[0] + pl0);
    const T * qe = qd + e * (12 * Q * QQ) + pqy * Q + pqx;
#pragma unroll
    for (int qz = 0; qz < Q; qz++)
#pragma unroll
      for (int k = 0; k < 12; k++)
        qf[qz][k] = __builtin_nontemporal_load(qe + (qz * 12 + k) * QQ);
  }
  else
  {
#pragma unroll
    for (int qz = 0; qz < Q; qz++)
#pragma unroll
      for (int k = 0; k < 12; k++)
        qf[qz][k] = T(0);
  }
  T U[3][Q], Cu[3][Q];
#pragma unroll
  for (int q = 0; q < Q; q++)
    Cu[0][q] = Cu[1][q] = Cu[2][q] = T(0);

  T * const pA = sA + pel * SA; // pencil's element blocks (stage z / z^T)
  T * const pB = sB + pel * SB;

  // ======================= forward =======================
  // (component index as a compile-time constant: one instantiation per component)
  auto forward = [&](auto cc) {
    constexpr int c = decltype(cc)::value;
    // ---- stage x: global rows -> A (contract dx -> qx) ----
    // items: ex_l fastest (coalesced row loads), then dy, dz, ey_l, ez_l
    constexpr int NI = TE * D * D;
    for (int it = tid; it < NI; it += NT)
    {
      const int lx = it % TX;
      int r = it / TX;
      const int dy = r % D;
      r /= D;
      const int dz = r % D;
      r /= D;
      const int ly = r % TY, lz = r / TY;
      const int ny_c = (c == 1) ? O : D, nz_c = (c == 2) ? O : D;
      if (dy >= ny_c || dz >= nz_c)
        continue;
      const int el = (lz * TY + ly) * TX + lx;
      const bool act = lx < ne[0] && ly < ne[1] && lz < ne[2];
      const int iy = (e0[1] + ly) * P + dy, iz = (e0[2] + lz) * P + dz;
      const int ix0 = (e0[0] + lx) * P;
      bool ress = false; // whole row essential
      if (c != 1)
        ress = ress || (iy == 0 && (pec & 4)) || (iy == cd[c][1] - 1 && (pec & 8));
      if (c != 2)
        ress = ress || (iz == 0 && (pec & 16)) || (iz == cd[c][2] - 1 && (pec & 32));
      T * A1 = sA + el * SA + dz * D + dy;
      if (c == 0)
      {
        T xv[O];
#pragma unroll
        for (int d = 0; d < O; d++)
          xv[d] = T(0);
        if (act && !ress)
        {
          const T * row = x + g.off[0] + ((int64_t)iz * cd[0][1] + iy) * cd[0][0] + ix0;
#pragma unroll
          for (int d = 0; d < O; d++)
            xv[d] = row[d];
        }
#pragma unroll
        for (int qx = 0; qx < Q; qx++)
        {
          T s = 0;
#pragma unroll
          for (int d = 0; d < O; d++)
            s += tb.Bo[qx * O + d] * xv[d];
          A1[qx * SQX] = s;
        }
      }
      else
      {
        T xv[D];
#pragma unroll
        for (int d = 0; d < D; d++)
          xv[d] = T(0);
        if (act && !ress)
        {
          const T * row = x + g.off[c] + ((int64_t)iz * cd[c][1] + iy) * cd[c][0] + ix0;
#pragma unroll
          for (int d = 0; d < D; d++)
            xv[d] = row[d];
          if (ix0 == 0 && (pec & 1))
            xv[0] = T(0);
          if (ix0 + P == cd[c][0] - 1 && (pec & 2))
            xv[P] = T(0);
        }
#pragma unroll
        for (int qx = 0; qx < Q; qx++)
        {
          T s = 0, gg = 0;
#pragma unroll
          for (int d = 0; d < D; d++)
          {
            s += tb.Bc[qx * D + d] * xv[d];
            gg += tb.Gc[qx * D + d] * xv[d];
          }
          A1[qx * SQX] = s;
          A1[LA + qx * SQX] = gg;
        }
      }
    }
    __syncthreads();
    // ---- stage y: A -> B (contract dy -> qy) ----
    {
      constexpr int NJ = TE * Q * D;
      for (int it = tid; it < NJ; it += NT)
      {
        const int qx = it % Q;
        const int dz = (it / Q) % D;
        const int el = it / (Q * D);
        const int nz_c = (c == 2) ? O : D;
        if (dz >= nz_c)
          continue;
        const T * A1 = sA + el * SA + qx * SQX + dz * D;
        T * B1 = sB + el * SB + dz * SDZ + qx;
        if (c == 0)
        {
          T a[D];
#pragma unroll
          for (int d = 0; d < D; d++)
            a[d] = A1[d];
#pragma unroll
          for (int qy = 0; qy < Q; qy++)
          {
            T s = 0, gg = 0;
#pragma unroll
            for (int d = 0; d < D; d++)
            {
              s += tb.Bc[qy * D + d] * a[d];
              gg += tb.Gc[qy * D + d] * a[d];
            }
            B1[qy * Q] = s;
            B1[LB + qy * Q] = gg;
          }
        }
        else if (c == 1)
        {
          T a1[O], a2[O];
#pragma unroll
          for (int d = 0; d < O; d++)
          {
            a1[d] = A1[d];
            a2[d] = A1[LA + d];
          }
#pragma unroll
          for (int qy = 0; qy < Q; qy++)
          {
            T s = 0, gg = 0;
#pragma unroll
            for (int d = 0; d < O; d++)
            {
              s += tb.Bo[qy * O + d] * a1[d];
              gg += tb.Bo[qy * O + d] * a2[d];
            }
            B1[qy * Q] = s;
            B1[LB + qy * Q] = gg;
          }
        }
        else
        {
          T a1[D], a2[D];
#pragma unroll
          for (int d = 0; d < D; d++)
          {
            a1[d] = A1[d];
            a2[d] = A1[LA + d];
          }
#pragma unroll
          for (int qy = 0; qy < Q; qy++)
          {
            T s = 0, gy = 0, gx = 0;
#pragma unroll
            for (int d = 0; d < D; d++)
            {
              s += tb.Bc[qy * D + d] * a1[d];
              gy += tb.Gc[qy * D + d] * a1[d];
              gx += tb.Bc[qy * D + d] * a2[d];
            }
            B1[qy * Q] = s;
            B1[LB + qy * Q] = gy;
            B1[2 * LB + qy * Q] = gx;
          }
        }
      }
    }
    __syncthreads();
    // ---- stage z: B -> registers (contract dz -> qz) ----
    if (tid < C::NP)
    {
      const T * B1 = pB + pqy * Q + pqx;
      if (c == 0)
      {
        T b1[D], b2[D];
#pragma unroll
        for (int d = 0; d < D; d++)
        {
          b1[d] = B1[d * SDZ];
          b2[d] = B1[LB + d * SDZ];
        }
#pragma unroll
        for (int qz = 0; qz < Q; qz++)
        {
          T u = 0, uz = 0, uy = 0;
#pragma unroll
          for (int d = 0; d < D; d++)
          {
            u += tb.Bc[qz * D + d] * b1[d];
            uz += tb.Gc[qz * D + d] * b1[d];
            uy += tb.Bc[qz * D + d] * b2[d];
          }
          U[0][qz] = u;
          Cu[1][qz] += uz;
          Cu[2][qz] -= uy;
        }
      }
      else if (c == 1)
      {
        T b1[D], b2[D];
#pragma unroll
        for (int d = 0; d < D; d++)
        {
          b1[d] = B1[d * SDZ];
          b2[d] = B1[LB + d * SDZ];
        }
#pragma unroll
        for (int qz = 0; qz < Q; qz++)
        {
          T u = 0, uz = 0, ux = 0;
#pragma unroll
          for (int d = 0; d < D; d++)
          {
            u += tb.Bc[qz * D + d] * b1[d];
            uz += tb.Gc[qz * D + d] * b1[d];
            ux += tb.Bc[qz * D + d] * b2[d];
          }
          U[1][qz] = u;
          Cu[0][qz] -= uz;
          Cu[2][qz] += ux;
        }
      }
      else
      {
        T b1[O], b2[O], b3[O];
#pragma unroll
        for (int d = 0; d < O; d++)
        {
          b1[d] = B1[d * SDZ];
          b2[d] = B1[LB + d * SDZ];
          b3[d] = B1[2 * LB + d * SDZ];
        }
#pragma unroll
        for (int qz = 0; qz < Q; qz++)
        {
          T u = 0, uy = 0, ux = 0;
#pragma unroll
          for (int d = 0; d < O; d++)
          {
            u += tb.Bo[qz * O + d] * b1[d];
            uy += tb.Bo[qz * O + d] * b2[d];
            ux += tb.Bo[qz * O + d] * b3[d];
          }
          U[2][qz] = u;
          Cu[0][qz] += uy;
          Cu[1][qz] -= ux;
        }
      }
    }
    // (no barrier: the next stage x writes A only; stage z above read B)
  };
  forward(std::integral_constant<int, 0>{});
  forward(std::integral_constant<int, 1>{});
  forward(std::integral_constant<int, 2>{});

  // ======================= quadrature points =======================
#pragma unroll
  for (int qz = 0; qz < Q; qz++)
  {
    const T u0 = U[0][qz], u1 = U[1][qz], u2 = U[2][qz];
    const T c0 = Cu[0][qz], c1 = Cu[1][qz], c2 = Cu[2][qz];
    U[0][qz] = qf[qz][0] * u0 + qf[qz][1] * u1 + qf[qz][2] * u2;
    U[1][qz] = qf[qz][1] * u0 + qf[qz][3] * u1 + qf[qz][4] * u2;
    U[2][qz] = qf[qz][2] * u0 + qf[qz][4] * u1 + qf[qz][5] * u2;
    Cu[0][qz] = qf[qz][6] * c0 + qf[qz][7] * c1 + qf[qz][8] * c2;
    Cu[1][qz] = qf[qz][7] * c0 + qf[qz][9] * c1 + qf[qz][10] * c2;
    Cu[2][qz] = qf[qz][8] * c0 + qf[qz][10] * c1 + qf[qz][11] * c2;
  }

  // ======================= backward =======================
  auto backward = [&](auto cc) {
    constexpr int c = decltype(cc)::value;
    // ---- z^T: registers -> S (B region), own pencil entries only ----
    if (tid < C::NP)
    {
      T * S1 = pB + pqy * Q + pqx;
      if (c == 0)
      {
#pragma unroll
        for (int d = 0; d < D; d++)
        {
          T s1 = 0, s2 = 0;
#pragma unroll
          for (int qz = 0; qz < Q; qz++)
          {
            s1 += tb.Bc[qz * D + d] * U[0][qz] + tb.Gc[qz * D + d] * Cu[1][qz];
            s2 -= tb.Bc[qz * D + d] * Cu[2][qz];
          }
          S1[d * SDZ] = s1;
          S1[LB + d * SDZ] = s2;
        }
      }
      else if (c == 1)
      {
#pragma unroll
        for (int d = 0; d < D; d++)
        {
          T s1 = 0, s2 = 0;
#pragma unroll
          for (int qz = 0; qz < Q; qz++)
          {
            s1 += tb.Bc[qz * D + d] * U[1][qz] - tb.Gc[qz * D + d] * Cu[0][qz];
            s2 += tb.Bc[qz * D + d] * Cu[2][qz];
          }
          S1[d * SDZ] = s1;
          S1[LB + d * SDZ] = s2;
        }
      }
      else
      {
#pragma unroll
        for (int d = 0; d < O; d++)
        {
          T sa = 0, sb = 0, sc = 0;
#pragma unroll
          for (int qz = 0; qz < Q; qz++)
          {
            const T bo = tb.Bo[qz * O + d];
            sa += bo * U[2][qz];
            sb += bo * Cu[0][qz];
            sc -= bo * Cu[1][qz];
          }
          S1[d * SDZ] = sa;
          S1[LB + d * SDZ] = sb;
          S1[2 * LB + d * SDZ] = sc;
        }
      }
    }
    __syncthreads();
    // ---- y^T: S -> T (A region), contract qy -> dy ----
    {
      constexpr int NJ = TE * Q * D;
      for (int it = tid; it < NJ; it += NT)
      {
        const int qx = it % Q;
        const int dz = (it / Q) % D;
        const int el = it / (Q * D);
        const int nz_c = (c == 2) ? O : D;
        if (dz >= nz_c)
          continue;
        const T * S1 = sB + el * SB + dz * SDZ + qx;
        T * T1 = sA + el * SA + qx * SQX + dz * D;
        if (c == 0)
        {
          T s1[Q], s2[Q];
#pragma unroll
          for (int q = 0; q < Q; q++)
          {
            s1[q] = S1[q * Q];
            s2[q] = S1[LB + q * Q];
          }
#pragma unroll
          for (int d = 0; d < D; d++)
          {
            T t = 0;
#pragma unroll
            for (int q = 0; q < Q; q++)
              t += tb.Bc[q * D + d] * s1[q] + tb.Gc[q * D + d] * s2[q];
            T1[d] = t;
          }
        }
        else if (c == 1)
        {
          T s1[Q], s2[Q];
#pragma unroll
          for (int q = 0; q < Q; q++)
          {
            s1[q] = S1[q * Q];
            s2[q] = S1[LB + q * Q];
          }
#pragma unroll
          for (int d = 0; d < O; d++)
          {
            T ta = 0, tb2 = 0;
#pragma unroll
            for (int q = 0; q < Q; q++)
            {
              ta += tb.Bo[q * O + d] * s1[q];
              tb2 += tb.Bo[q * O + d] * s2[q];
            }
            T1[d] = ta;
            T1[LA + d] = tb2;
          }
        }
        else
        {
          T sa[Q], sb[Q], sc[Q];
#pragma unroll
          for (int q = 0; q < Q; q++)
          {
            sa[q] = S1[q * Q];
            sb[q] = S1[LB + q * Q];
            sc[q] = S1[2 * LB + q * Q];
          }
#pragma unroll
          for (int d = 0; d < D; d++)
          {
            T ta = 0, tb2 = 0;
#pragma unroll
            for (int q = 0; q < Q; q++)
            {
              ta += tb.Bc[q * D + d] * sa[q] + tb.Gc[q * D + d] * sb[q];
              tb2 += tb.Bc[q * D + d] * sc[q];
            }
            T1[d] = ta;
            T1[LA + d] = tb2;
          }
        }
      }
    }
    __syncthreads();
    // ---- row stage: sum T over the elements sharing a row, contract x, write ----
    {
      // rows of component c: (lx, iy, iz); open axis has T*P points, closed T*P+1
      constexpr int RY0 = TY * P + 1, RZ0 = TZ * P + 1;
      const int RY = (c == 1) ? TY * P : RY0;
      const int RZ = (c == 2) ? TZ * P : RZ0;
      const int NR = TX * RY * RZ;
      for (int it = tid; it < NR; it += NT)
      {
        const int lx = it % TX;
        const int iy = (it / TX) % RY;
        const int iz = it / (TX * RY);
        // element candidates along y and z: slot 0 = lower element (shared face),
        // slot 1 = the element the index falls in; summed slot 0 first
        const int ehy = iy / P, lhy = iy % P, ehz = iz / P, lhz = iz % P;
        const int ey[2] = {ehy - 1, ehy}, dyc[2] = {P, lhy};
        const int ez[2] = {ehz - 1, ehz}, dzc[2] = {P, lhz};
        const bool vy[2] = {c != 1 && lhy == 0 && ehy > 0, c == 1 || ehy < ne[1]};
        const bool vz[2] = {c != 2 && lhz == 0 && ehz > 0, c == 2 || ehz < ne[2]};
        const int nyc = (vy[0] || vy[1]) ? 1 : 0, nzc = (vz[0] || vz[1]) ? 1 : 0;
        const int giy = e0[1] * P + iy, giz = e0[2] * P + iz;
        if (lx >= ne[0] || giy >= cd[c][1] || giz >= cd[c][2] || nyc == 0 || nzc == 0)
          continue;
        // row-level essential test and tile-face classification (y and z)
        bool ress = false;
        if (c != 1)
          ress = ress || (giy == 0 && (pec & 4)) || (giy == cd[c][1] - 1 && (pec & 8));
        if (c != 2)
          ress = ress || (giz == 0 && (pec & 16)) || (giz == cd[c][2] - 1 && (pec & 32));
        const bool ylo = (c != 1) && iy == 0 && tpos[1] > 0;
        const bool yhi = (c != 1) && iy == TY * P && tpos[1] < g.nt[1] - 1;
        const bool zlo = (c != 2) && iz == 0 && tpos[2] > 0;
        const bool zhi = (c != 2) && iz == TZ * P && tpos[2] < g.nt[2] - 1;
        const int64_t rowbase = g.off[c] + ((int64_t)giz * cd[c][1] + giy) * cd[c][0];
        if (c == 0)
        {
          // open x: each thread owns dx = 0..O-1 of element column lx
          T ts[Q];
#pragma unroll
          for (int q = 0; q < Q; q++)
            ts[q] = T(0);
#pragma unroll
          for (int kz = 0; kz < 2; kz++)
#pragma unroll
            for (int ky = 0; ky < 2; ky++)
            {
              if (!vz[kz] || !vy[ky])
                continue;
              const int el = (ez[kz] * TY + ey[ky]) * TX + lx;
              const T * T1 = sA + el * SA + dzc[kz] * D + dyc[ky];
#pragma unroll
              for (int q = 0; q < Q; q++)
                ts[q] += T1[q * SQX];
            }
          // slot bits: bit0 = y (upper face of a non-last tile), bit1 = z
          const int slot = (yhi ? 1 : 0) | (zhi ? 2 : 0);
          const bool shared = ylo || yhi || zlo || zhi;
          const bool owner = !yhi && !zhi;
          const int64_t base = rowbase + (e0[0] + lx) * P;
#pragma unroll
          for (int d = 0; d < O; d++)
          {
            if (ress)
            {
              if (owner)
                y[base + d] = x[base + d];
              continue;
            }
            T s = 0;
#pragma unroll
            for (int q = 0; q < Q; q++)
              s += tb.Bo[q * O + d] * ts[q];
            if (shared)
              part[slot * N + base + d] = s;
            else
              y[base + d] = s;
          }
        }
        else
        {
          // closed x: element column lx contributes dx = 0..P; dx = 0 also
          // receives dx = P of column lx-1 (same tile), dx = P is written by
          // the last active column only.
          T ta[Q], tb2[Q], pa[Q], pb[Q];
#pragma unroll
          for (int q = 0; q < Q; q++)
            ta[q] = tb2[q] = pa[q] = pb[q] = T(0);
#pragma unroll
          for (int kz = 0; kz < 2; kz++)
#pragma unroll
            for (int ky = 0; ky < 2; ky++)
            {
              if (!vz[kz] || !vy[ky])
                continue;
              const int elb = (ez[kz] * TY + ey[ky]) * TX;
              const int off = dzc[kz] * D + dyc[ky];
              const T * T1 = sA + (elb + lx) * SA + off;
#pragma unroll
              for (int q = 0; q < Q; q++)
              {
                ta[q] += T1[q * SQX];
                tb2[q] += T1[LA + q * SQX];
              }
              if (lx > 0)
              {
                const T * T0 = sA + (elb + lx - 1) * SA + off;
#pragma unroll
                for (int q = 0; q < Q; q++)
                {
                  pa[q] += T0[q * SQX];
                  pb[q] += T0[LA + q * SQX];
                }
              }
            }
          const int gx0 = (e0[0] + lx) * P;
          const bool last = (lx == ne[0] - 1);
          // component-c slot bits: bit0 = x (closed), bit1 = the other closed axis
          const bool olo = (c == 1) ? zlo : ylo;
          const bool ohi = (c == 1) ? zhi : yhi;
#pragma unroll
          for (int d = 0; d <= P; d++)
          {
            if (d == P && !last)
              continue;
            const int gx = gx0 + d;
            const bool xlo = (d == 0 && lx == 0 && tpos[0] > 0);
            const bool xhi = (d == P && lx == TX - 1 && tpos[0] < g.nt[0] - 1);
            const bool ess = ress || (gx == 0 && (pec & 1)) || (gx == cd[c][0] - 1 && (pec & 2));
            const int64_t gid = rowbase + gx;
            if (ess)
            {
              if (!xhi && !ohi)
                y[gid] = x[gid];
              continue;
            }
            T s = 0;
#pragma unroll
            for (int q = 0; q < Q; q++)
              s += tb.Bc[q * D + d] * ta[q] + tb.Gc[q * D + d] * tb2[q];
            if (d == 0 && lx > 0)
            {
              T sp = 0;
#pragma unroll
              for (int q = 0; q < Q; q++)
                sp += tb.Bc[q * D + P] * pa[q] + tb.Gc[q * D + P] * pb[q];
              s = sp + s;
            }
            const int slot = (xhi ? 1 : 0) | (ohi ? 2 : 0);
            if (xlo || xhi || olo || ohi)
              part[slot * N + gid] = s;
            else
              y[gid] = s;
          }
        }
      }
    }
    // (no barrier: the next z^T writes B only; this stage read A)
  };
  backward(std::integral_constant<int, 0>{});
  backward(std::integral_constant<int, 1>{});
  backward(std::integral_constant<int, 2>{});
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int P, int TX, int TY, int TZ>
static TileShape
shape_of()
{
  return TileShape{TX, TY, TZ, Tile<P, TX, TY, TZ>::NT};
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
      hipLaunchKernelGGL((k_apply_tile<1, T, 4, 4, 4>), dim3(nblk), dim3(Tile<1, 4, 4, 4>::NT), 0, st, g, tb, qd, x,
                         y, part, done);
      break;
    case 2:
      hipLaunchKernelGGL((k_apply_tile<2, T, 4, 4, 2>), dim3(nblk), dim3(Tile<2, 4, 4, 2>::NT), 0, st, g, tb, qd, x,
                         y, part, done);
      break;
    case 3:
      hipLaunchKernelGGL((k_apply_tile<3, T, 4, 2, 2>), dim3(nblk), dim3(Tile<3, 4, 2, 2>::NT), 0, st, g, tb, qd, x,
                         y, part, done);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t
apply_op(const Geom & g, const PlaneSegs & segs, const Tables<T> & tb, const T * qd, const T * x, T * y,
         T * part, const int * done, hipStream_t st)
{
  hipError_t e = launch_tile<T>(g, tb, qd, x, y, part, done, st);
  if (e != hipSuccess)
    return e;
  return assemble_planes<T>(g, segs, part, y, done, st);
}

template hipError_t apply_op<double>(const Geom &, const PlaneSegs &, const Tables<double> &, const double *,
                                     const double *, double *, double *, const int *, hipStream_t);
template hipError_t apply_op<float>(const Geom &, const PlaneSegs &, const Tables<float> &, const float *,
                                    const float *, float *, float *, const int *, hipStream_t);

} // namespace hc
