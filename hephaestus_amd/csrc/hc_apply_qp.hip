// Element apply with one lane per quadrature point (the "qp" tile kernel).
//
// Same operator, tile decomposition, descriptors and write-out as
// k_apply_tile (hc_apply.hip), organised for occupancy instead: an element's
// (p+1)^3 quadrature points map to (p+1)^3 lanes of one wave (p = 3: one
// element per wave; p = 2: two; p = 1: eight), so every sum-factorisation
// stage exchanges data through LDS inside the wave and needs no workgroup
// barrier.  Per lane only the point's 12 quadrature factors and 6 field values
// live in registers.  Per element and stage:
//   forward  (component c):  x DOFs --x--> S1 --y--> S2 --z--> u_c, curl parts
//   point                    12 factors: U' = M U, C' = K C
//   backward (component c):  U', C' --z^T--> S2 --y^T--> S1 --x^T--> element vector
// and one barrier before the tile gather: each tile DOF sums the element
// vectors of the (up to four) elements that share it in a fixed order, then
// goes to y, to the tile-face partials, or (PEC) copies x.
#include "hc_kernels.hpp"

#include <type_traits>

namespace hc
{

template <int P, int TX, int TY, int TZ>
struct QpTile
{
  static constexpr int O = P, D = P + 1, Q = P + 1, Q2 = Q * Q, Q3 = Q * Q * Q;
  static constexpr int EW = 64 / Q3; // elements per wave
  static constexpr int TE = TX * TY * TZ;
  static constexpr int NW = (TE + EW - 1) / EW;
  static constexpr int NT = NW * 64;
  static constexpr int AL = (Q3 + 3) / 4 * 4; // LDS array length (>= (p+1)^3)
  static constexpr int EL = 9 * AL;           // per element: 3 DOF arrays, 3 + 3 stage arrays
};

__device__ inline void
qp_wave_sync()
{
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
}

template <int P, typename T, int TX, int TY, int TZ, bool BATCH>
__global__ void __launch_bounds__((QpTile<P, TX, TY, TZ>::NT))
k_apply_qp(ApplyDesc d0, const ApplyDesc * __restrict__ descs, int ndesc, Tables<T> tb)
{
  using C = QpTile<P, TX, TY, TZ>;
  constexpr int O = C::O, D = C::D, Q = C::Q, Q2 = C::Q2, Q3 = C::Q3, EW = C::EW, TE = C::TE, NT = C::NT;
  constexpr int AL = C::AL, EL = C::EL;
  // ---- work item: (instance, tile, rhs group) as in k_apply_tile ----
  const int w = blockIdx.x;
  int di = 0;
  if constexpr (BATCH)
    di = __builtin_amdgcn_readfirstlane(find_instance(w, ndesc, [&](int i) { return descs[i].w0; }));
  const ApplyDesc & ad = BATCH ? descs[di] : d0;
  const int wl = w - ad.w0;
  const int rb = BATCH ? ad.rb : 1;
  const int ngrp = BATCH ? (ad.nrhs + rb - 1) / rb : ad.nrhs;
  const int r_begin = (wl % ngrp) * rb;
  const int r_end = BATCH ? min(ad.nrhs, r_begin + rb) : r_begin + 1;
  const int tcur = wl / ngrp;
  if (!BATCH && ad.done && gmem(ad.done)[r_begin * ad.done_stride] != ad.gate)
    return;
  const Geom & g = ad.g;
  const gptr<const T> __restrict__ qd = gmem((const T *)ad.qd);
  gptr<const T> x = gmem((const T *)ad.x);
  gptr<T> y = gmem((T *)ad.y);
  gptr<T> part = gmem((T *)ad.part);
  __shared__ T sE[TE * EL];
  // 1D tables in LDS: the lanes index them with their own point / DOF index
  __shared__ T sBo[QMAXC * PMAX], sBc[QMAXC * QMAXC], sGc[QMAXC * QMAXC];

  const int tid = threadIdx.x;
  if (tid == 0)
  {
#pragma unroll
    for (int i = 0; i < QMAXC * PMAX; i++)
      sBo[i] = tb.Bo[i];
#pragma unroll
    for (int i = 0; i < QMAXC * QMAXC; i++)
    {
      sBc[i] = tb.Bc[i];
      sGc[i] = tb.Gc[i];
    }
  }
  __syncthreads();
  int tpos[3], e0[3], ne[3];
  tpos[0] = tcur % g.nt[0];
  tpos[1] = (tcur / g.nt[0]) % g.nt[1];
  tpos[2] = tcur / (g.nt[0] * g.nt[1]);
  e0[0] = tpos[0] * TX;
  e0[1] = tpos[1] * TY;
  e0[2] = tpos[2] * TZ;
  ne[0] = min(TX, g.n[0] - e0[0]);
  ne[1] = min(TY, g.n[1] - e0[1]);
  ne[2] = min(TZ, g.n[2] - e0[2]);
  const int64_t N = g.off[3];
  const int pec = g.pec;

  // ---- lane roles: element of the wave, quadrature point / stage item ----
  const int lane = tid & 63;
  const int le = lane / Q3;
  const int qid = lane - le * Q3;
  const int el = (tid >> 6) * EW + le;
  const int elx = el % TX, ely = (el / TX) % TY, elz = el / (TX * TY);
  const bool eact = lane < EW * Q3 && el < TE && elx < ne[0] && ely < ne[1] && elz < ne[2];
  const int gex = e0[0] + elx, gey = e0[1] + ely, gez = e0[2] + elz;
  T * const R0 = sE + (eact ? el : 0) * EL; // DOF arrays: x of the element, then its output
  T * const R1 = R0 + 3 * AL;               // S1 (forward), V / S1 (backward)
  T * const R2 = R0 + 6 * AL;               // S2
  // quadrature point of the lane
  const int qx = qid % Q, qy = (qid / Q) % Q, qz = qid / Q2;

  // ---- the point's 12 quadrature factors (held across right-hand sides) ----
  T qf[12];
  {
    const int64_t e = eact ? ((int64_t)gez * g.n[1] + gey) * g.n[0] + gex : 0;
    const gptr<const T> qe = qd + e * (12 * Q3) + (eact ? qz * 12 * Q2 + qy * Q + qx : 0);
#pragma unroll
    for (int k = 0; k < 12; k++)
    {
      const T v = __builtin_nontemporal_load(qe + k * Q2);
      qf[k] = eact ? v : T(0);
    }
  }

  // component dims: n0 along x, n1 along y, n2 along z (open axis c has p)
  auto nd = [](int c, int a) { return (a == c) ? O : D; };

  for (int rhs = r_begin; rhs < r_end; rhs++)
  {
    if (BATCH && ad.done && gmem(ad.done)[rhs * ad.done_stride] != ad.gate)
      continue;
    x = gmem((const T *)ad.x) + (int64_t)rhs * N;
    y = gmem((T *)ad.y) + (int64_t)rhs * N;
    part = gmem((T *)ad.part) + (int64_t)rhs * 4 * part_stride(ad);

    // ---- element DOFs of x into R0 (PEC DOFs read as zero) ----
    if (eact)
    {
#pragma unroll
      for (int c = 0; c < 3; c++)
      {
        const int n0 = nd(c, 0), n1 = nd(c, 1), n2 = nd(c, 2);
        const int dim1 = g.n[1] * P + (c != 1), dim0 = g.n[0] * P + (c != 0), dim2 = g.n[2] * P + (c != 2);
        for (int j = qid; j < D * D * D; j += Q3)
        {
          const int i0 = j % D, i1 = (j / D) % D, i2 = j / (D * D);
          T v = T(0);
          if (i0 < n0 && i1 < n1 && i2 < n2)
          {
            const int gx = gex * P + i0, gy = gey * P + i1, gz = gez * P + i2;
            bool ess = false;
            if (c != 0)
              ess = ess || (gx == 0 && (pec & 1)) || (gx == dim0 - 1 && (pec & 2));
            if (c != 1)
              ess = ess || (gy == 0 && (pec & 4)) || (gy == dim1 - 1 && (pec & 8));
            if (c != 2)
              ess = ess || (gz == 0 && (pec & 16)) || (gz == dim2 - 1 && (pec & 32));
            if (!ess)
              v = x[g.off[c] + ((int64_t)gz * dim1 + gy) * dim0 + gx];
          }
          R0[c * AL + j] = v;
        }
      }
    }
    qp_wave_sync();

    T U[3] = {T(0), T(0), T(0)}, Cu[3] = {T(0), T(0), T(0)};
    // ======================= forward =======================
    auto forward = [&](auto cc) {
      constexpr int c = decltype(cc)::value;
      constexpr int n0 = (c == 0) ? O : D, n1 = (c == 1) ? O : D, n2 = (c == 2) ? O : D;
      const T * X = R0 + c * AL;
      // stage x: item (qx, d1, d2) -> S1 [d2][d1][qx]
      {
        const int iq = qid % Q, d1 = (qid / Q) % D, d2 = qid / (Q * D);
        if (eact && d1 < n1 && d2 < n2)
        {
          const T * xr = X + (d2 * D + d1) * D;
          T a = 0, gg = 0;
#pragma unroll
          for (int d = 0; d < n0; d++)
          {
            const T v = xr[d];
            if (c == 0)
              a += sBo[iq * O + d] * v;
            else
            {
              a += sBc[iq * D + d] * v;
              gg += sGc[iq * D + d] * v;
            }
          }
          R1[(d2 * D + d1) * Q + iq] = a;
          if (c != 0)
            R1[AL + (d2 * D + d1) * Q + iq] = gg;
        }
      }
      qp_wave_sync();
      // stage y: item (qx, qy, d2) -> S2 [d2][qy][qx]
      {
        const int iq = qid % Q, jq = (qid / Q) % Q, d2 = qid / Q2;
        if (eact && d2 < n2)
        {
          T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
          for (int d = 0; d < n1; d++)
          {
            const T s0 = R1[(d2 * D + d) * Q + iq];
            if (c == 0)
            {
              a0 += sBc[jq * D + d] * s0;
              a1 += sGc[jq * D + d] * s0;
            }
            else if (c == 1)
            {
              const T s1 = R1[AL + (d2 * D + d) * Q + iq];
              a0 += sBo[jq * O + d] * s0;
              a1 += sBo[jq * O + d] * s1;
            }
            else
            {
              const T s1 = R1[AL + (d2 * D + d) * Q + iq];
              a0 += sBc[jq * D + d] * s0;
              a1 += sGc[jq * D + d] * s0;
              a2 += sBc[jq * D + d] * s1;
            }
          }
          R2[(d2 * Q + jq) * Q + iq] = a0;
          R2[AL + (d2 * Q + jq) * Q + iq] = a1;
          if (c == 2)
            R2[2 * AL + (d2 * Q + jq) * Q + iq] = a2;
        }
      }
      qp_wave_sync();
      // stage z: the lane's point
      if (eact)
      {
        T u = 0, v1 = 0, v2 = 0;
#pragma unroll
        for (int d = 0; d < n2; d++)
        {
          const int o = (d * Q + qy) * Q + qx;
          const T b0 = R2[o], b1 = R2[AL + o];
          if (c == 2)
          {
            const T bo = sBo[qz * O + d];
            u += bo * b0;
            v1 += bo * b1;
            v2 += bo * R2[2 * AL + o];
          }
          else
          {
            u += sBc[qz * D + d] * b0;
            v1 += sGc[qz * D + d] * b0;
            v2 += sBc[qz * D + d] * b1;
          }
        }
        U[c] = u;
        if (c == 0)
        {
          Cu[1] += v1; // d/dz u_x
          Cu[2] -= v2; // d/dy u_x
        }
        else if (c == 1)
        {
          Cu[0] -= v1; // d/dz u_y
          Cu[2] += v2; // d/dx u_y
        }
        else
        {
          Cu[0] += v1; // d/dy u_z
          Cu[1] -= v2; // d/dx u_z
        }
      }
      // the next stage x writes S1 only after this wave's reads of S2 and S1
      qp_wave_sync();
    };
    forward(std::integral_constant<int, 0>{});
    forward(std::integral_constant<int, 1>{});
    forward(std::integral_constant<int, 2>{});

    // ======================= quadrature point =======================
    {
      const T u0 = U[0], u1 = U[1], u2 = U[2];
      const T c0 = Cu[0], c1 = Cu[1], c2 = Cu[2];
      U[0] = qf[0] * u0 + qf[1] * u1 + qf[2] * u2;
      U[1] = qf[1] * u0 + qf[3] * u1 + qf[4] * u2;
      U[2] = qf[2] * u0 + qf[4] * u1 + qf[5] * u2;
      Cu[0] = qf[6] * c0 + qf[7] * c1 + qf[8] * c2;
      Cu[1] = qf[7] * c0 + qf[9] * c1 + qf[10] * c2;
      Cu[2] = qf[8] * c0 + qf[10] * c1 + qf[11] * c2;
    }

    // ======================= backward =======================
    auto backward = [&](auto cc) {
      constexpr int c = decltype(cc)::value;
      constexpr int n0 = (c == 0) ? O : D, n1 = (c == 1) ? O : D, n2 = (c == 2) ? O : D;
      // the point values this component's test functions need -> V (R1)
      if (eact)
      {
        if (c == 0)
        {
          R1[qid] = U[0];
          R1[AL + qid] = Cu[1];
          R1[2 * AL + qid] = Cu[2];
        }
        else if (c == 1)
        {
          R1[qid] = U[1];
          R1[AL + qid] = Cu[0];
          R1[2 * AL + qid] = Cu[2];
        }
        else
        {
          R1[qid] = U[2];
          R1[AL + qid] = Cu[0];
          R1[2 * AL + qid] = Cu[1];
        }
      }
      qp_wave_sync();
      // z^T: item (qx, qy, d2) -> S2 [d2][qy][qx]
      {
        const int iq = qid % Q, jq = (qid / Q) % Q, d2 = qid / Q2;
        if (eact && d2 < n2)
        {
          T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
          for (int q = 0; q < Q; q++)
          {
            const int o = (q * Q + jq) * Q + iq;
            const T va = R1[o], vb = R1[AL + o], vc = R1[2 * AL + o];
            if (c == 0)
            {
              s0 += sBc[q * D + d2] * va + sGc[q * D + d2] * vb;
              s1 -= sBc[q * D + d2] * vc;
            }
            else if (c == 1)
            {
              s0 += sBc[q * D + d2] * va - sGc[q * D + d2] * vb;
              s1 += sBc[q * D + d2] * vc;
            }
            else
            {
              const T bo = sBo[q * O + d2];
              s0 += bo * va;
              s1 += bo * vb;
              s2 -= bo * vc;
            }
          }
          R2[(d2 * Q + jq) * Q + iq] = s0;
          R2[AL + (d2 * Q + jq) * Q + iq] = s1;
          if (c == 2)
            R2[2 * AL + (d2 * Q + jq) * Q + iq] = s2;
        }
      }
      qp_wave_sync();
      // y^T: item (qx, d1, d2) -> S1 [d2][d1][qx] (R1, the V arrays are dead)
      {
        const int iq = qid % Q, d1 = (qid / Q) % D, d2 = qid / (Q * D);
        if (eact && d1 < n1 && d2 < n2)
        {
          T ta = 0, tb2 = 0;
#pragma unroll
          for (int q = 0; q < Q; q++)
          {
            const int o = (d2 * Q + q) * Q + iq;
            const T sa = R2[o], sb = R2[AL + o];
            if (c == 0)
              ta += sBc[q * D + d1] * sa + sGc[q * D + d1] * sb;
            else if (c == 1)
            {
              ta += sBo[q * O + d1] * sa;
              tb2 += sBo[q * O + d1] * sb;
            }
            else
            {
              ta += sBc[q * D + d1] * sa + sGc[q * D + d1] * sb;
              tb2 += sBc[q * D + d1] * R2[2 * AL + o];
            }
          }
          R1[(d2 * D + d1) * Q + iq] = ta;
          if (c != 0)
            R1[AL + (d2 * D + d1) * Q + iq] = tb2;
        }
      }
      qp_wave_sync();
      // x^T: item (d0, d1, d2) -> element vector of component c (R0)
      {
        const int d0 = qid % D, d1 = (qid / D) % D, d2 = qid / (D * D);
        if (eact && d0 < n0 && d1 < n1 && d2 < n2)
        {
          const T * t = R1 + (d2 * D + d1) * Q;
          T v = 0;
#pragma unroll
          for (int q = 0; q < Q; q++)
          {
            if (c == 0)
              v += sBo[q * O + d0] * t[q];
            else
              v += sBc[q * D + d0] * t[q] + sGc[q * D + d0] * t[AL + q];
          }
          R0[c * AL + (d2 * D + d1) * D + d0] = v;
        }
      }
      qp_wave_sync();
    };
    backward(std::integral_constant<int, 0>{});
    backward(std::integral_constant<int, 1>{});
    backward(std::integral_constant<int, 2>{});
    __syncthreads(); // the gather reads the element vectors of other waves

    // ======================= tile gather and write-out =======================
    constexpr int BX0 = TX * P, BX1 = TY * P + 1, BX2 = TZ * P + 1;
    constexpr int BY0 = TX * P + 1, BY1 = TY * P, BY2 = TZ * P + 1;
    constexpr int BZ0 = TX * P + 1, BZ1 = TY * P + 1, BZ2 = TZ * P;
    constexpr int OB1 = BX0 * BX1 * BX2, OB2 = OB1 + BY0 * BY1 * BY2, OBN = OB2 + BZ0 * BZ1 * BZ2;
    // descriptor-table instances: segment offsets held in scalar registers;
    // the by-value instance reads them where used (fewer live registers)
    SegOffsets eos{};
    if constexpr (BATCH)
      eos = seg_offsets(ad.segs);
    for (int i = tid; i < OBN; i += NT)
    {
      int c, r, b0, b1;
      if (i < OB1)
      {
        c = 0;
        r = i;
        b0 = BX0;
        b1 = BX1;
      }
      else if (i < OB2)
      {
        c = 1;
        r = i - OB1;
        b0 = BY0;
        b1 = BY1;
      }
      else
      {
        c = 2;
        r = i - OB2;
        b0 = BZ0;
        b1 = BZ1;
      }
      int l[3];
      l[0] = r % b0;
      l[1] = (r / b0) % b1;
      l[2] = r / (b0 * b1);
      const int Tl[3] = {TX * P, TY * P, TZ * P};
      int gi[3], dims[3];
      bool ok = true, ess = false, shared = false, owner = true;
      int slot = 0, bit = 1, alo = 3;
#pragma unroll
      for (int a = 0; a < 3; a++)
      {
        gi[a] = e0[a] * P + l[a];
        const int dim = g.n[a] * P + (a != c ? 1 : 0);
        dims[a] = dim;
        ok = ok && gi[a] < dim;
        if (a == c)
          continue;
        ess = ess || (gi[a] == 0 && ((pec >> (2 * a)) & 1)) || (gi[a] == dim - 1 && ((pec >> (2 * a + 1)) & 1));
        const bool lo = (l[a] == 0 && tpos[a] > 0);
        const bool hi = (l[a] == Tl[a] && tpos[a] < g.nt[a] - 1);
        shared = shared || lo || hi;
        if ((lo || hi) && alo == 3)
          alo = a;
        if (hi)
        {
          owner = false;
          slot |= bit;
        }
        bit <<= 1;
      }
      if (!ok)
        continue;
      const int64_t offc = (c == 0) ? g.off[0] : ((c == 1) ? g.off[1] : g.off[2]);
      const int64_t gid = offc + ((int64_t)gi[2] * dims[1] + gi[1]) * dims[0] + gi[0];
      if (ess)
      {
        if (owner)
          y[gid] = x[gid];
        continue;
      }
      // sum of the element vectors of the (up to 4) tile elements sharing the
      // DOF: along a closed axis the DOF at a multiple of p > 0 also belongs
      // to the lower element (at d = p); lower element first
      int ec[3][2], dc[3][2], nc[3];
#pragma unroll
      for (int a = 0; a < 3; a++)
      {
        const int ea = l[a] / P, da = l[a] % P;
        const int lim = (a == 0) ? ne[0] : ((a == 1) ? ne[1] : ne[2]);
        const bool sec = (a != c) && da == 0 && ea > 0 && ea - 1 < lim; // lower element, at d = p
        const bool pri = ea < lim;
        ec[a][0] = sec ? ea - 1 : ea;
        dc[a][0] = sec ? P : da;
        ec[a][1] = ea;
        dc[a][1] = da;
        nc[a] = (sec ? 1 : 0) + (pri ? 1 : 0);
      }
      T v = T(0);
#pragma unroll
      for (int kz = 0; kz < 2; kz++)
#pragma unroll
        for (int ky = 0; ky < 2; ky++)
#pragma unroll
          for (int kx = 0; kx < 2; kx++)
          {
            if (kz >= nc[2] || ky >= nc[1] || kx >= nc[0])
              continue;
            const int ee = (ec[2][kz] * TY + ec[1][ky]) * TX + ec[0][kx];
            v += sE[ee * EL + c * AL + (dc[2][kz] * D + dc[1][ky]) * D + dc[0][kx]];
          }
      if (shared)
      {
        const int ga = (alo == 0) ? gi[0] : ((alo == 1) ? gi[1] : gi[2]);
        const int ta = (alo == 0) ? TX * P : ((alo == 1) ? TY * P : TZ * P);
        const int j0 = (alo == 0) ? gi[1] : gi[0];
        const int j1 = (alo == 2) ? gi[1] : gi[2];
        const int n0 = (alo == 0) ? dims[1] : dims[0];
        const int n1 = (alo == 2) ? dims[1] : dims[2];
        const int64_t eo = BATCH ? seg_offset(eos, c, alo) : ad.segs.eoff[c][alo];
        const int64_t pid = eo + ((int64_t)(ga / ta - 1) * n1 + j1) * n0 + j0;
        part[slot * part_stride(ad) + pid] = v;
      }
      else
        y[gid] = v;
    }
    if (BATCH && rhs + 1 < r_end)
      __syncthreads(); // the next rhs overwrites the element arrays
  }
}

template <int P, int TX, int TY, int TZ>
hipError_t
launch_qp(const ApplyDesc & d0, const ApplyDesc * descs, int ndesc, int nwork, int grouped, const void * tb,
          int prec, hipStream_t st)
{
  using C = QpTile<P, TX, TY, TZ>;
  if (nwork <= 0)
    return hipSuccess;
  (void)grouped; // this kernel reads the quadrature-point factors (always present)
  if (prec == 64 && descs)
    hipLaunchKernelGGL((k_apply_qp<P, double, TX, TY, TZ, true>), dim3(nwork), dim3(C::NT), 0, st, d0, descs, ndesc,
                       *(const Tables<double> *)tb);
  else if (prec == 64)
    hipLaunchKernelGGL((k_apply_qp<P, double, TX, TY, TZ, false>), dim3(nwork), dim3(C::NT), 0, st, d0, descs,
                       ndesc, *(const Tables<double> *)tb);
  else if (descs)
    hipLaunchKernelGGL((k_apply_qp<P, float, TX, TY, TZ, true>), dim3(nwork), dim3(C::NT), 0, st, d0, descs, ndesc,
                       *(const Tables<float> *)tb);
  else
    hipLaunchKernelGGL((k_apply_qp<P, float, TX, TY, TZ, false>), dim3(nwork), dim3(C::NT), 0, st, d0, descs,
                       ndesc, *(const Tables<float> *)tb);
  return hipGetLastError();
}

#define HC_QP_INST(P, TX, TY, TZ)                                                                                \
  template hipError_t launch_qp<P, TX, TY, TZ>(const ApplyDesc &, const ApplyDesc *, int, int, int, const void *, \
                                               int, hipStream_t);
HC_QP_INST(1, 4, 4, 4)
HC_QP_INST(1, 8, 4, 2)
HC_QP_INST(2, 4, 2, 2)
HC_QP_INST(2, 4, 4, 2)
#undef HC_QP_INST

} // namespace hc
