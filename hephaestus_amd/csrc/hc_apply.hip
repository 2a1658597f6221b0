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
#include "hc_basis.hpp"
#include "hc_kernels.hpp"

#include <type_traits>
#include <utility>

namespace hc
{

template <int P, int TX, int TY, int TZ, int SP = 1>
struct Tile
{
  static constexpr int O = P, D = P + 1, Q = P + 1, QQ = Q * Q;
  static constexpr int TE = TX * TY * TZ;
  static constexpr int NP = TE * QQ;                // pencils
  static constexpr int NPP = (NP + 63) / 64 * 64;   // pencils padded to whole waves
  static constexpr int NT = NPP * SP;               // SP wave-uniform groups share each pencil's qz range
  static constexpr int QH = (Q + SP - 1) / SP;      // quadrature points per thread along z
  // A-type arrays [qx][dz][dy], qx stride padded
  static constexpr int SQX = D * D + 1;
  static constexpr int LA = Q * SQX;
  // B-type arrays [dz][qy][qx], dz stride padded
  static constexpr int SDZ = QQ + 4;
  static constexpr int LB = D * SDZ;
  static constexpr int pad_to(int v, int m, int r) { return ((v - r + m - 1) / m) * m + r; }
  static constexpr int SA = pad_to(5 * LA, 16, 4);
  static constexpr int SB = pad_to(3 * LB, 16, 4);
  // wave-synchronous stages: when the (p+1)^2 pencils, stage items and
  // transposed-stage items of an element always fall in one wave (64 divisible
  // by (p+1)^2, whole waves of pencils, no split of the qz range, and the
  // stage-x items of a row of TX elements -- D*D*TX, element fastest -- in
  // whole fractions of a wave), every LDS exchange between the forward and
  // backward stages stays inside a wave and needs no workgroup barrier --
  // only the x^T rows and the write-out do.
  static constexpr bool WS = (SP == 1) && (64 % QQ == 0) && (NP % 64 == 0) && (64 % (D * D * TX) == 0);
};

// Move the listed entries of a (wave-uniform) array into vector registers:
// an empty asm statement that takes and returns each value in a VGPR, so the
// compiler keeps them there instead of re-creating constants in scalar
// registers at every use.  Static indices only (the array stays in registers).
template <typename V>
__device__ inline void
pin_one(V & v)
{
  asm volatile("" : "+v"(v));
}
template <typename A, size_t... I>
__device__ inline void
pin_vgpr(A & a, std::index_sequence<I...>)
{
  (pin_one(a[I]), ...);
}

// Order the LDS traffic of one wave (all lanes of a wavefront run in lockstep,
// so completing its own LDS operations is enough).
__device__ inline void
wave_lds_sync()
{
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
}

// BATCH = 0: one instance, described by the by-value argument d0 (its
// fields stay in scalar registers), one rhs per work item.  BATCH = 2: the
// instances of the device table descs[], one rhs per work item.  BATCH = 1: the
// instances of the device table descs[]; a work item loops over a group of rb
// right-hand sides with the tile's quadrature data held in registers (read
// from HBM once per group).
// EF = 1: element factors (box meshes, ApplyDesc::ef) instead of the
// quadrature-point factors.
template <int P, typename T, int TX, int TY, int TZ, int MODE, int SP, int BATCH, int EF, int VT = 1>
__global__ void __launch_bounds__((Tile<P, TX, TY, TZ, SP>::NT))
k_apply_tile(ApplyDesc d0, const ApplyDesc * __restrict__ descs, int ndesc, Tables<T> tb_arg)
{
  using C = Tile<P, TX, TY, TZ, SP>;
  // the 1D tables of order P as compile-time constants (the same values as
  // tb_arg): every use has a static index once the loops are unrolled, so no
  // table occupies scalar registers through the kernel
  (void)tb_arg;
  constexpr int QH = C::QH, NPP = C::NPP;
  constexpr int O = C::O, D = C::D, Q = C::Q, QQ = C::QQ;
  constexpr int SQX = C::SQX, LA = C::LA, SDZ = C::SDZ, LB = C::LB, SA = C::SA, SB = C::SB;
  constexpr int NT = C::NT, TE = C::TE;
  constexpr basis::Tab<T> tb_c = basis::tab<P, T>();
  // FP64: gfx950 has no 64-bit instruction literals, so each use of a double
  // constant costs two scalar moves into a register pair (rematerialised
  // whenever scalar registers run short).  With VREG the order-P tables are
  // moved into vector registers once per wave and every contraction reads
  // them there: 88 more VGPRs at order 3, which fit where a SIMD holds two
  // waves of these kernels (one qz group, the LDS of two workgroups per CU)
  // and no quadrature-point factors are held (element factors); every batch
  // form of such a tile does the same, so their results agree.  Measured at
  // 48^3 order 3: 177 -> 156 us on the (4, 2, 2) tile; with two qz groups, or
  // with the quadrature-point factors, the extra registers cost occupancy
  // and the kernels slow down.  FP32 constants are literals of the VOP2
  // encoding.
  // With the quadrature-point factors the order-3 tiles hold 170 VGPRs
  // already: only the closed-basis tables (64 VGPRs) are moved there.
  constexpr bool VREG = VT != 0 && sizeof(T) == 8 && SP == 1 && (EF || (P == 3 && MODE == 1 && BATCH != 1));
  basis::Tab<T> tb = tb_c;
  if constexpr (VREG)
  {
    if constexpr (EF)
      pin_vgpr(tb.Bo, std::make_index_sequence<Q * O>{});
    pin_vgpr(tb.Bc, std::make_index_sequence<Q * D>{});
    pin_vgpr(tb.Gc, std::make_index_sequence<Q * D>{});
  }
  // barrier between stages whose LDS exchanges may cross waves
  auto stage_sync = [&]() {
    if constexpr (C::WS)
      wave_lds_sync();
    else
      __syncthreads();
  };
  // ---- work item: (instance, tile, rhs group), group fastest; a group is
  // rb consecutive right-hand sides (rb = 1 unless BATCH) ----
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  int di = 0;
  constexpr bool TABLE = BATCH != 0, GROUP = BATCH == 1;
  if constexpr (TABLE)
    di = __builtin_amdgcn_readfirstlane(find_instance(w, ndesc, [&](int i) { return descs[i].w0; }));
  const ApplyDesc & ad = TABLE ? descs[di] : d0;
  const int rb = GROUP ? ad.rb : 1;
  const int ngrp = GROUP ? (ad.nrhs + rb - 1) / rb : ad.nrhs;
  const int wl = w - ad.w0;
  const int r_begin = (wl % ngrp) * rb;
  const int r_end = GROUP ? min(ad.nrhs, r_begin + rb) : r_begin + 1;
  const int tcur = wl / ngrp;
  if (!GROUP && ad.done && gmem(ad.done)[r_begin * ad.done_stride] != ad.gate)
    return;
  const Geom & g = ad.g;
  const gptr<const T> __restrict__ qd = gmem((const T *)ad.qd);
  // vectors of the current right-hand side (reset per rhs below)
  gptr<const T> x = gmem((const T *)ad.x);
  gptr<T> y = gmem((T *)ad.y);
  gptr<T> part = gmem((T *)ad.part);
  gptr<const T> rbv = gmem((const T *)ad.resid); // residual mode when set (wave-uniform)
  const bool resid = ad.resid != nullptr;
  __shared__ T sA[TE * SA];
  __shared__ T sB[TE * SB];

  const int tid = threadIdx.x;
  // tile position, element origin and active extent (partial tiles at the
  // upper domain faces)
  auto tile_pos = [&](int t, int * tp, int * e0_, int * ne_) {
    tp[0] = t % g.nt[0];
    tp[1] = (t / g.nt[0]) % g.nt[1];
    tp[2] = t / (g.nt[0] * g.nt[1]);
    e0_[0] = tp[0] * TX;
    e0_[1] = tp[1] * TY;
    e0_[2] = tp[2] * TZ;
    ne_[0] = min(TX, g.n[0] - e0_[0]);
    ne_[1] = min(TY, g.n[1] - e0_[1]);
    ne_[2] = min(TZ, g.n[2] - e0_[2]);
  };
  int tpos[3], e0[3], ne[3];
  tile_pos(tcur, tpos, e0, ne);
  const int64_t N = g.off[3];
  const int cd[3][3] = {{(int)g.cd[0][0], (int)g.cd[0][1], (int)g.cd[0][2]},
                        {(int)g.cd[1][0], (int)g.cd[1][1], (int)g.cd[1][2]},
                        {(int)g.cd[2][0], (int)g.cd[2][1], (int)g.cd[2][2]}};
  const int pec = g.pec;

  // pencil of this thread (stages z, z^T and the quadrature points)
  // group ph (wave-uniform) owns quadrature points qz = ph*QH + j of the pencil
  const int ph = __builtin_amdgcn_readfirstlane(tid / NPP);
  const int ptid = tid - ph * NPP;
  const bool pact = ptid < C::NP;
  const int pel = ptid / QQ;
  const int pqx = (ptid % QQ) % Q, pqy = (ptid % QQ) / Q;
  // run f with the group index as a compile-time constant (keeps every 1D
  // table index static, so the table argument stays in scalar registers)
  auto with_ph = [&](auto f) {
    if (SP == 1 || ph == 0)
      f(std::integral_constant<int, 0>{});
    else
      f(std::integral_constant<int, SP == 1 ? 0 : 1>{});
  };
  const int pl0 = pel % TX, pl1 = (pel / TX) % TY, pl2 = pel / (TX * TY);
  // stage-x item of this thread: (lx fastest, dy, dz, ly, lz); TE*D*D <= NT
  const int ilx = tid % TX, idy = (tid / TX) % D, idz = (tid / (TX * D)) % D;
  const int ily = (tid / (TX * D * D)) % TY, ilz = tid / (TX * D * D * TY);

  // ---- loads of one tile: the pencil's 12 Q quadrature factors (or its
  // element's two factors), the item's x rows of the three components (PEC
  // DOFs read as zero) ----
  T qf[EF ? 1 : QH][EF ? 1 : 12];
  T fe[2]; // element factors dt nu_e, sigma_e
  T xr0[O], xr1[D], xr2[D];
  auto load_qf = [&](const int * e0_, const int * ne_) {
    const bool pin = pact && pl0 < ne_[0] && pl1 < ne_[1] && pl2 < ne_[2];
    // Inactive lanes read element 0 (in bounds, finite values) and keep them:
    // their U, Cu are zero or never reach an output (the x^T rows sum valid
    // elements only; quadrature points past Q meet zero table entries).
    const int64_t e = pin ? ((int64_t)(e0_[2] + pl2) * g.n[1] + (e0_[1] + pl1)) * g.n[0] + (e0_[0] + pl0) : 0;
    if constexpr (EF)
    {
      const gptr<const T> ce = gmem((const T *)ad.ef.qc) + 2 * e;
      fe[0] = ce[0];
      fe[1] = ce[1];
      return;
    }
    const gptr<const T> qe = qd + e * (12 * Q * QQ) + (pin ? pqy * Q + pqx : 0);
#pragma unroll
    for (int j = 0; j < QH; j++)
    {
      const int qz = ph * QH + j;
      const bool ld = pin && qz < Q;
      const gptr<const T> qq = qe + (ld ? qz : 0) * 12 * QQ;
#pragma unroll
      for (int k = 0; k < 12; k++)
        qf[j][k] = __builtin_nontemporal_load(qq + k * QQ);
    }
  };
  auto load_x = [&](const int * e0_, const int * ne_) {
    const bool iin = tid < TE * D * D && ilx < ne_[0] && ily < ne_[1] && ilz < ne_[2];
    const int iiy = (e0_[1] + ily) * P + idy, iiz = (e0_[2] + ilz) * P + idz, iix = (e0_[0] + ilx) * P;
#pragma unroll
    for (int d = 0; d < O; d++)
      xr0[d] = T(0);
#pragma unroll
    for (int d = 0; d < D; d++)
      xr1[d] = xr2[d] = T(0);
    // component x: closed y, z (row PEC), open x
    const bool ress0 = (iiy == 0 && (pec & 4)) || (iiy == cd[0][1] - 1 && (pec & 8)) ||
                       (iiz == 0 && (pec & 16)) || (iiz == cd[0][2] - 1 && (pec & 32));
    if (iin && !ress0)
    {
      const gptr<const T> row = x + g.off[0] + ((int64_t)iiz * cd[0][1] + iiy) * cd[0][0] + iix;
#pragma unroll
      for (int d = 0; d < O; d++)
        xr0[d] = row[d];
    }
    // component y: open y (dy < O), closed z (row PEC), closed x (end PEC)
    const bool ress1 = (iiz == 0 && (pec & 16)) || (iiz == cd[1][2] - 1 && (pec & 32));
    if (iin && idy < O && !ress1)
    {
      const gptr<const T> row = x + g.off[1] + ((int64_t)iiz * cd[1][1] + iiy) * cd[1][0] + iix;
#pragma unroll
      for (int d = 0; d < D; d++)
        xr1[d] = row[d];
      if (iix == 0 && (pec & 1))
        xr1[0] = T(0);
      if (iix + P == cd[1][0] - 1 && (pec & 2))
        xr1[P] = T(0);
    }
    // component z: open z (dz < O), closed y (row PEC), closed x (end PEC)
    const bool ress2 = (iiy == 0 && (pec & 4)) || (iiy == cd[2][1] - 1 && (pec & 8));
    if (iin && idz < O && !ress2)
    {
      const gptr<const T> row = x + g.off[2] + ((int64_t)iiz * cd[2][1] + iiy) * cd[2][0] + iix;
#pragma unroll
      for (int d = 0; d < D; d++)
        xr2[d] = row[d];
      if (iix == 0 && (pec & 1))
        xr2[0] = T(0);
      if (iix + P == cd[2][0] - 1 && (pec & 2))
        xr2[P] = T(0);
    }
  };
  // MODE 0: quadrature data loaded at the quadrature-point stage (fewest
  // registers); MODE 1: at entry, hidden behind the forward contractions.
  // Vector-memory loads complete in issue order (one counter): for a single
  // rhs the x rows go first, so stage x waits for them alone and the
  // quadrature data stays in flight through the forward contractions.
  bool x_ready = false;
  if (MODE >= 1)
  {
    if constexpr (!GROUP)
    {
      x = gmem((const T *)ad.x) + (int64_t)r_begin * g.off[3];
      load_x(e0, ne);
      x_ready = true;
    }
    load_qf(e0, ne);
  }
  else if (EF)
    load_qf(e0, ne);

  for (int rhs = r_begin; rhs < r_end; rhs++)
  {
  if (GROUP && ad.done && gmem(ad.done)[rhs * ad.done_stride] != ad.gate)
    continue;
  x = gmem((const T *)ad.x) + (int64_t)rhs * g.off[3];
  y = gmem((T *)ad.y) + (int64_t)rhs * g.off[3];
  if (resid)
    rbv = gmem((const T *)ad.resid) + (int64_t)rhs * g.off[3];
  part = gmem((T *)ad.part) + (int64_t)rhs * 4 * part_stride(ad);
  if (!x_ready)
    load_x(e0, ne);
  x_ready = false;
  T U[3][QH], Cu[3][QH];
#pragma unroll
  for (int j = 0; j < QH; j++)
    Cu[0][j] = Cu[1][j] = Cu[2][j] = T(0);

  T * const pA = sA + pel * SA; // pencil's element blocks (stage z / z^T)
  T * const pB = sB + pel * SB;

  // ======================= forward =======================
  // (component index as a compile-time constant: one instantiation per component)
  auto forward = [&](auto cc) {
    constexpr int c = decltype(cc)::value;
    // ---- stage x: prefetched rows -> A (contract dx -> qx) ----
    {
      const int ny_c = (c == 1) ? O : D, nz_c = (c == 2) ? O : D;
      if (tid < TE * D * D && idy < ny_c && idz < nz_c)
      {
        const int el = (ilz * TY + ily) * TX + ilx;
        T * A1 = sA + el * SA + idz * D + idy;
        if (c == 0)
        {
#pragma unroll
          for (int qx = 0; qx < Q; qx++)
          {
            T s = 0;
#pragma unroll
            for (int d = 0; d < O; d++)
              s += tb.Bo[qx * O + d] * xr0[d];
            A1[qx * SQX] = s;
          }
        }
        else
        {
#pragma unroll
          for (int qx = 0; qx < Q; qx++)
          {
            T s = 0, gg = 0;
#pragma unroll
            for (int d = 0; d < D; d++)
            {
              const T v = (c == 1) ? xr1[d] : xr2[d];
              s += tb.Bc[qx * D + d] * v;
              gg += tb.Gc[qx * D + d] * v;
            }
            A1[qx * SQX] = s;
            A1[LA + qx * SQX] = gg;
          }
        }
      }
    }
    stage_sync();
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
    stage_sync();
    // ---- stage z: B -> registers (contract dz -> qz), own qz range ----
    if (pact)
    {
      const T * B1 = pB + pqy * Q + pqx;
      constexpr int ND = (c == 2) ? O : D;
      constexpr int NA = (c == 2) ? 3 : 2;
      T b[NA][ND];
#pragma unroll
      for (int d = 0; d < ND; d++)
#pragma unroll
        for (int k = 0; k < NA; k++)
          b[k][d] = B1[k * LB + d * SDZ];
      with_ph([&](auto phc) {
      constexpr int PH = decltype(phc)::value;
#pragma unroll
      for (int j = 0; j < QH; j++)
      {
        const int qz = PH * QH + j; // past Q (odd Q split in two): tables are zero there
        T u = 0, v1 = 0, v2 = 0;
#pragma unroll
        for (int d = 0; d < ND; d++)
        {
          if (c == 2)
          {
            const T bo = tb.Bo[qz * O + d];
            u += bo * b[0][d];
            v1 += bo * b[1][d];
            v2 += bo * b[2][d];
          }
          else
          {
            u += tb.Bc[qz * D + d] * b[0][d];
            v1 += tb.Gc[qz * D + d] * b[0][d];
            v2 += tb.Bc[qz * D + d] * b[1][d];
          }
        }
        U[c][j] = u;
        if (c == 0)
        {
          Cu[1][j] += v1; // d/dz u_x
          Cu[2][j] -= v2; // d/dy u_x
        }
        else if (c == 1)
        {
          Cu[0][j] -= v1; // d/dz u_y
          Cu[2][j] += v2; // d/dx u_y
        }
        else
        {
          Cu[0][j] += v1; // d/dy u_z
          Cu[1][j] -= v2; // d/dx u_z
        }
      }
      });
    }
    // (no barrier: the next stage x writes A only; stage z above read B)
  };
  forward(std::integral_constant<int, 0>{});
  forward(std::integral_constant<int, 1>{});
  forward(std::integral_constant<int, 2>{});
  // With SP > 1 the groups sharing a pencil sit in different waves: all of
  // them must have read the pencil's B entries (stage z of the last
  // component) before z^T below overwrites them.
  if constexpr (SP > 1)
    __syncthreads();

  // ======================= quadrature points =======================
  if (MODE == 0 && !EF)
    load_qf(e0, ne);
  // term weights (batched kernel only; the single-instance kernel applies A)
  const T msc = TABLE ? (T)ad.mscale : T(1), csc = TABLE ? (T)ad.cscale : T(1);
  if constexpr (EF)
  {
    // factor of component a at (qx, qy, qz): sigma_e w mk[a] (mass) and
    // dt nu_e w ck[a] (curl), w = w_qx w_qy w_qz
    auto qw_sel = [&](int q) {
      T v = tb_c.qw[0];
      if constexpr (Q > 1)
        v = q == 1 ? tb_c.qw[1] : v;
      if constexpr (Q > 2)
        v = q == 2 ? tb_c.qw[2] : v;
      if constexpr (Q > 3)
        v = q == 3 ? tb_c.qw[3] : v;
      static_assert(Q <= 4, "weights of at most four points");
      return v;
    };
    const T wxy = qw_sel(pqx) * qw_sel(pqy);
    const T sgm = TABLE ? fe[1] * msc : fe[1], cnu = TABLE ? fe[0] * csc : fe[0];
    const T mk0 = (T)ad.ef.mk[0], mk1 = (T)ad.ef.mk[1], mk2 = (T)ad.ef.mk[2];
    const T ck0 = (T)ad.ef.ck[0], ck1 = (T)ad.ef.ck[1], ck2 = (T)ad.ef.ck[2];
    with_ph([&](auto phc) {
      constexpr int PH = decltype(phc)::value;
#pragma unroll
      for (int j = 0; j < QH; j++)
      {
        const int qz = PH * QH + j; // past Q: zero weight, and U, Cu are zero there
        const T w = wxy * tb_c.qw[qz];
        const T sm = w * sgm, sc = w * cnu;
        U[0][j] = (sm * mk0) * U[0][j];
        U[1][j] = (sm * mk1) * U[1][j];
        U[2][j] = (sm * mk2) * U[2][j];
        Cu[0][j] = (sc * ck0) * Cu[0][j];
        Cu[1][j] = (sc * ck1) * Cu[1][j];
        Cu[2][j] = (sc * ck2) * Cu[2][j];
      }
    });
  }
  else
  {
#pragma unroll
  for (int j = 0; j < QH; j++)
  {
    T u0 = U[0][j], u1 = U[1][j], u2 = U[2][j];
    T c0 = Cu[0][j], c1 = Cu[1][j], c2 = Cu[2][j];
    if constexpr (TABLE)
    {
      u0 *= msc;
      u1 *= msc;
      u2 *= msc;
      c0 *= csc;
      c1 *= csc;
      c2 *= csc;
    }
    U[0][j] = qf[j][0] * u0 + qf[j][1] * u1 + qf[j][2] * u2;
    U[1][j] = qf[j][1] * u0 + qf[j][3] * u1 + qf[j][4] * u2;
    U[2][j] = qf[j][2] * u0 + qf[j][4] * u1 + qf[j][5] * u2;
    Cu[0][j] = qf[j][6] * c0 + qf[j][7] * c1 + qf[j][8] * c2;
    Cu[1][j] = qf[j][7] * c0 + qf[j][9] * c1 + qf[j][10] * c2;
    Cu[2][j] = qf[j][8] * c0 + qf[j][10] * c1 + qf[j][11] * c2;
  }
  }

  // ======================= backward =======================
  auto backward = [&](auto cc) {
    constexpr int c = decltype(cc)::value;
    // ---- z^T: registers -> S (B region), own pencil entries only ----
    // Each group sums over its own qz range; with SP > 1 group 0 writes S and
    // the other groups add theirs after a barrier (fixed order).
    auto ztr = [&](bool add) {
      with_ph([&](auto phc) {
      constexpr int PH = decltype(phc)::value;
      T * S1 = pB + pqy * Q + pqx;
      constexpr int ND = (c == 2) ? O : D;
      constexpr int NA = (c == 2) ? 3 : 2;
#pragma unroll
      for (int d = 0; d < ND; d++)
      {
        T sv[NA];
#pragma unroll
        for (int k = 0; k < NA; k++)
          sv[k] = T(0);
#pragma unroll
        for (int j = 0; j < QH; j++)
        {
          const int qz = PH * QH + j; // past Q: zero table entries and zero V, W
          if (c == 0)
          {
            sv[0] += tb.Bc[qz * D + d] * U[0][j] + tb.Gc[qz * D + d] * Cu[1][j];
            sv[1] -= tb.Bc[qz * D + d] * Cu[2][j];
          }
          else if (c == 1)
          {
            sv[0] += tb.Bc[qz * D + d] * U[1][j] - tb.Gc[qz * D + d] * Cu[0][j];
            sv[1] += tb.Bc[qz * D + d] * Cu[2][j];
          }
          else
          {
            const T bo = tb.Bo[qz * O + d];
            sv[0] += bo * U[2][j];
            sv[1] += bo * Cu[0][j];
            sv[2] -= bo * Cu[1][j];
          }
        }
#pragma unroll
        for (int k = 0; k < NA; k++)
        {
          if (add)
            S1[k * LB + d * SDZ] += sv[k];
          else
            S1[k * LB + d * SDZ] = sv[k];
        }
      }
      });
    };
    if (pact && ph == 0)
      ztr(false);
#pragma unroll
    for (int h = 1; h < SP; h++)
    {
      __syncthreads();
      if (pact && ph == h)
        ztr(true);
    }
    stage_sync();
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
        T * T1 = sA + el * SA + (c == 0 ? 0 : (c == 1 ? LA : 3 * LA)) + qx * SQX + dz * D;
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
    stage_sync();
    // (the barrier above also orders this y^T before the next z^T overwrites B)
  };
  backward(std::integral_constant<int, 0>{});
  backward(std::integral_constant<int, 1>{});
  backward(std::integral_constant<int, 2>{});
  if constexpr (C::WS)
    __syncthreads(); // the x^T rows read the T arrays of other waves' elements

  // ======================= x^T rows =======================
  // Item (component, DOF row (iy, iz) of the tile box, element column lx):
  // sums the T tensors of the (at most four) elements sharing the row in a
  // fixed order (lower element first), contracts along x and stores the
  // column's DOFs in an LDS box (B region, free now).  For the components
  // with a closed x axis the DOF on an element x-face inside the tile is
  // (column lx-1 at dx = P) + (column lx at dx = 0): column lx-1 stores its
  // dx = P value in the box, column lx keeps its dx = 0 value in sC, and a
  // second pass adds the two (a sum of two terms, so the order is immaterial).
  constexpr int BX0 = TX * P, BX1 = TY * P + 1, BX2 = TZ * P + 1; // comp x box
  constexpr int BY0 = TX * P + 1, BY1 = TY * P, BY2 = TZ * P + 1; // comp y box
  constexpr int BZ0 = TX * P + 1, BZ1 = TY * P + 1, BZ2 = TZ * P; // comp z box
  constexpr int OB1 = BX0 * BX1 * BX2, OB2 = OB1 + BY0 * BY1 * BY2, OBN = OB2 + BZ0 * BZ1 * BZ2;
  constexpr int NR0 = BX1 * BX2 * TX, NR1 = NR0 + BY1 * BY2 * TX, NRN = NR1 + BZ1 * BZ2 * TX;
  static_assert(OBN + NRN <= TE * SB, "output box and x-face values must fit in the B region");
  T * const sOut = sB;
  T * const sC = sB + OBN;
  auto xrow = [&](auto cc, int it, int j) {
    constexpr int c = decltype(cc)::value;
    constexpr int RY = (c == 1) ? TY * P : TY * P + 1;
    constexpr int RX = (c == 0) ? TX * P : TX * P + 1;
    const int lx = j % TX, r = j / TX;
    if (lx >= ne[0])
      return;
    const int iy = r % RY, iz = r / RY;
    const int ehy = iy / P, lhy = iy % P, ehz = iz / P, lhz = iz % P;
    const int ey[2] = {ehy - 1, ehy}, dyc[2] = {P, lhy};
    const int ez[2] = {ehz - 1, ehz}, dzc[2] = {P, lhz};
    const bool vy[2] = {c != 1 && lhy == 0 && ehy > 0, c == 1 || ehy < ne[1]};
    const bool vz[2] = {c != 2 && lhz == 0 && ehz > 0, c == 2 || ehz < ne[2]};
    if (!(vy[0] || vy[1]) || !(vz[0] || vz[1]))
      return;
    const T * TA = sA + (c == 0 ? 0 : (c == 1 ? LA : 3 * LA));
    T * const orow = sOut + (c == 0 ? 0 : (c == 1 ? OB1 : OB2)) + r * RX + lx * P;
    if (c == 0)
    {
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
          const T * T1 = TA + ((ez[kz] * TY + ey[ky]) * TX + lx) * SA + dzc[kz] * D + dyc[ky];
#pragma unroll
          for (int q = 0; q < Q; q++)
            ts[q] += T1[q * SQX];
        }
#pragma unroll
      for (int d = 0; d < O; d++)
      {
        T v = 0;
#pragma unroll
        for (int q = 0; q < Q; q++)
          v += tb.Bo[q * O + d] * ts[q];
        orow[d] = v;
      }
    }
    else
    {
      T ta[Q], tb2[Q];
#pragma unroll
      for (int q = 0; q < Q; q++)
        ta[q] = tb2[q] = T(0);
#pragma unroll
      for (int kz = 0; kz < 2; kz++)
#pragma unroll
        for (int ky = 0; ky < 2; ky++)
        {
          if (!vz[kz] || !vy[ky])
            continue;
          const T * T1 = TA + ((ez[kz] * TY + ey[ky]) * TX + lx) * SA + dzc[kz] * D + dyc[ky];
#pragma unroll
          for (int q = 0; q < Q; q++)
          {
            ta[q] += T1[q * SQX];
            tb2[q] += T1[LA + q * SQX];
          }
        }
#pragma unroll
      for (int d = 0; d <= P; d++)
      {
        T v = 0;
#pragma unroll
        for (int q = 0; q < Q; q++)
          v += tb.Bc[q * D + d] * ta[q] + tb.Gc[q * D + d] * tb2[q];
        if (d == 0 && lx > 0)
          sC[it] = v;
        else
          orow[d] = v;
      }
    }
  };
  // items padded so that each component starts on a wave boundary (a wave
  // runs one component's code)
  constexpr int XP1 = (NR0 + 63) / 64 * 64, XP2 = XP1 + (NR1 - NR0 + 63) / 64 * 64;
  constexpr int XPN = XP2 + (NRN - NR1 + 63) / 64 * 64;
  for (int i = tid; i < XPN; i += NT)
  {
    const int iw = __builtin_amdgcn_readfirstlane(i - (i & 63)); // the wave's first item
    if (iw < XP1)
    {
      if (i < NR0)
        xrow(std::integral_constant<int, 0>{}, i, i);
    }
    else if (iw < XP2)
    {
      if (i - XP1 < NR1 - NR0)
        xrow(std::integral_constant<int, 1>{}, NR0 + i - XP1, i - XP1);
    }
    else if (i - XP2 < NRN - NR1)
      xrow(std::integral_constant<int, 2>{}, NR1 + i - XP2, i - XP2);
  }
  __syncthreads();
  // second pass: x-faces between the tile's element columns (closed-x components);
  // rows outside the active extent hold unused values and are never written out
  for (int it = NR0 + tid; it < NRN; it += NT)
  {
    const bool c1 = it < NR1;
    const int j = c1 ? it - NR0 : it - NR1;
    const int lx = j % TX, r = j / TX;
    if (lx == 0 || lx >= ne[0])
      continue;
    T * const o = sOut + (c1 ? OB1 : OB2) + r * (TX * P + 1) + lx * P;
    *o = *o + sC[it];
  }
  __syncthreads();
  // ---- write-out: interior DOFs to y, DOFs on a face shared with another
  // tile to the partials (slot bit0 = lower-numbered closed axis, bit1 = the
  // other; set when this tile holds the face as its upper face), PEC DOFs copy
  // x (written by the tile that holds them as a lower face) ----
  // Each component's items start on a wave boundary, so a wave runs the code
  // of one component; the tile's bases, strides and face flags are
  // wave-uniform (scalar) and a DOF costs a few compares, two 24-bit
  // multiply-adds and one select of its destination.
  constexpr int NB0 = OB1, NB1 = OB2 - OB1, NB2 = OBN - OB2;
  constexpr int WP1 = (NB0 + 63) / 64 * 64, WP2 = WP1 + (NB1 + 63) / 64 * 64, WPN = WP2 + (NB2 + 63) / 64 * 64;
  // items [wbeg, wbeg + nb) of the padded item space hold component c; the
  // thread takes those congruent to tid modulo NT (one loop per component, so
  // that only one component's tile values are live at a time)
  auto wloop = [&](auto cc, int wbeg, int nb) {
    constexpr int c = decltype(cc)::value;
    constexpr int a1 = (c == 0) ? 1 : 0, a2 = (c == 2) ? 1 : 2; // closed axes, a1 < a2
    constexpr int BB0 = (c == 0) ? TX * P : TX * P + 1;
    constexpr int BB1 = (c == 1) ? TY * P : TY * P + 1;
    constexpr int Tl[3] = {TX * P, TY * P, TZ * P};
    constexpr int OBc = (c == 0) ? 0 : ((c == 1) ? OB1 : OB2);
    // in-plane axes (first, second) of the plane normal to a1 and to a2
    constexpr int p10 = (a1 == 0) ? 1 : 0, p11 = 2;
    constexpr int p20 = 0, p21 = (a2 == 2) ? 1 : 2;
    int dims[3], o[3];
#pragma unroll
    for (int a = 0; a < 3; a++)
    {
      dims[a] = g.n[a] * P + (a != c ? 1 : 0);
      o[a] = e0[a] * P;
    }
    const int64_t ybase = g.off[c] + ((int64_t)o[2] * dims[1] + o[1]) * dims[0] + o[0];
    for (int i = wbeg + ((tid - wbeg) % NT + NT) % NT; i < wbeg + nb; i += NT)
    {
      const int j = i - wbeg;
      const int l[3] = {j % BB0, (j / BB0) % BB1, j / (BB0 * BB1)};
      if (l[0] >= dims[0] - o[0] || l[1] >= dims[1] - o[1] || l[2] >= dims[2] - o[2])
        continue;
      const uint32_t off = __umul24(__umul24(l[2], dims[1]) + l[1], dims[0]) + l[0];
      const bool lo1 = l[a1] == 0 && tpos[a1] > 0, hi1 = l[a1] == Tl[a1] && tpos[a1] < g.nt[a1] - 1;
      const bool lo2 = l[a2] == 0 && tpos[a2] > 0, hi2 = l[a2] == Tl[a2] && tpos[a2] < g.nt[a2] - 1;
      const bool ess = (l[a1] == 0 && tpos[a1] == 0 && ((pec >> (2 * a1)) & 1)) ||
                       (l[a1] == dims[a1] - 1 - o[a1] && ((pec >> (2 * a1 + 1)) & 1)) ||
                       (l[a2] == 0 && tpos[a2] == 0 && ((pec >> (2 * a2)) & 1)) ||
                       (l[a2] == dims[a2] - 1 - o[a2] && ((pec >> (2 * a2 + 1)) & 1));
      if (ess)
      {
        if (!hi1 && !hi2)
          y[ybase + off] = resid ? rbv[ybase + off] - x[ybase + off] : x[ybase + off];
        continue;
      }
      const T v = sOut[OBc + j];
      if (lo1 || hi1 || lo2 || hi2)
      {
        // compact entry of the segment (c, alo), alo the lowest axis with an
        // interior tile plane through the DOF: row (plane k, position along
        // the second in-plane axis), entry along the first; k = the tile's
        // plane index along alo minus one, plus one for its upper face
        const int64_t ps = part_stride(ad);
        int64_t pid;
        if (lo1 || hi1)
          pid = ad.segs.eoff[c][a1] +
                ((int64_t)(tpos[a1] - 1 + (hi1 ? 1 : 0)) * dims[p11] + o[p11] + l[p11]) * dims[p10] + o[p10] +
                l[p10] + (hi1 ? ps : 0) + (hi2 ? 2 * ps : 0);
        else
          pid = ad.segs.eoff[c][a2] +
                ((int64_t)(tpos[a2] - 1 + (hi2 ? 1 : 0)) * dims[p21] + o[p21] + l[p21]) * dims[p20] + o[p20] +
                l[p20] + (hi2 ? 2 * ps : 0);
        part[pid] = v;
      }
      else
        y[ybase + off] = resid ? rbv[ybase + off] - v : v;
    }
  };
  wloop(std::integral_constant<int, 0>{}, 0, NB0);
  wloop(std::integral_constant<int, 1>{}, WP1, NB1);
  wloop(std::integral_constant<int, 2>{}, WP2, NB2);
  (void)WPN;
  if (GROUP && rhs + 1 < r_end)
    __syncthreads(); // the next rhs overwrites the LDS stages
  }
}

// ---------------------------------------------------------------------------
// launchers: candidate tile shapes per order (chosen by the autotuner)
// ---------------------------------------------------------------------------
template <int P, int TX, int TY, int TZ, int MODE, int SP>
static hipError_t
launch_one(const ApplyDesc & d0, const ApplyDesc * descs, int ndesc, int nwork, int grouped, const void * tb,
           int prec, hipStream_t st)
{
  using C = Tile<P, TX, TY, TZ, SP>;
  if (nwork <= 0)
    return hipSuccess;
  // BATCH variant: 0 single instance, 1 table with rhs groups, 2 table with one rhs per item
  const int bv = !descs ? 0 : ((grouped & 1) ? 1 : 2);
  const bool ef = descs ? (grouped & LAUNCH_EF) != 0 : d0.ef.qc != nullptr;
  auto go = [&](auto kern, const auto & t) {
    hipLaunchKernelGGL(kern, dim3(nwork), dim3(C::NT), 0, st, d0, descs, ndesc, t);
  };
  if (prec == 64)
  {
    const Tables<double> & t = *(const Tables<double> *)tb;
    if (bv == 0)
      ef ? go(k_apply_tile<P, double, TX, TY, TZ, MODE, SP, 0, 1>, t) : go(k_apply_tile<P, double, TX, TY, TZ, MODE, SP, 0, 0>, t);
    else if (bv == 1)
      ef ? go(k_apply_tile<P, double, TX, TY, TZ, MODE, SP, 1, 1>, t) : go(k_apply_tile<P, double, TX, TY, TZ, MODE, SP, 1, 0>, t);
    else
      ef ? go(k_apply_tile<P, double, TX, TY, TZ, MODE, SP, 2, 1>, t) : go(k_apply_tile<P, double, TX, TY, TZ, MODE, SP, 2, 0>, t);
  }
  else
  {
    const Tables<float> & t = *(const Tables<float> *)tb;
    if (bv == 0)
      ef ? go(k_apply_tile<P, float, TX, TY, TZ, MODE, SP, 0, 1>, t) : go(k_apply_tile<P, float, TX, TY, TZ, MODE, SP, 0, 0>, t);
    else if (bv == 1)
      ef ? go(k_apply_tile<P, float, TX, TY, TZ, MODE, SP, 1, 1>, t) : go(k_apply_tile<P, float, TX, TY, TZ, MODE, SP, 1, 0>, t);
    else
      ef ? go(k_apply_tile<P, float, TX, TY, TZ, MODE, SP, 2, 1>, t) : go(k_apply_tile<P, float, TX, TY, TZ, MODE, SP, 2, 0>, t);
  }
  return hipGetLastError();
}

// mode field: load mode (0, 1) + 4 * (number of qz groups - 1)
#define HC_CAND(P, TX, TY, TZ, M, SP)                                                                            \
  {P, {TX, TY, TZ, Tile<P, TX, TY, TZ, SP>::NT}, M + 4 * (SP - 1), &launch_one<P, TX, TY, TZ, M, SP>}

#define HC_QP(P, TX, TY, TZ) {P, {TX, TY, TZ, 0}, MODE_QP, &launch_qp<P, TX, TY, TZ>}

static const TileCandidate g_cands[] = {
    // first entry of each order = default before tuning (order 3: one qz group,
    // whose FP64 kernels hold the order tables in vector registers -- the
    // fastest element-factor tile on box meshes, within a few per cent of the
    // best with quadrature-point factors)
    HC_CAND(1, 4, 4, 4, 1, 2), HC_CAND(1, 4, 4, 4, 1, 1), HC_CAND(1, 8, 4, 2, 1, 1), HC_CAND(1, 8, 4, 2, 0, 2),
    HC_CAND(1, 8, 8, 1, 0, 1), HC_CAND(1, 16, 2, 2, 1, 1), HC_QP(1, 4, 4, 4), HC_QP(1, 8, 4, 2),
    HC_CAND(2, 4, 2, 2, 1, 1), HC_CAND(2, 4, 2, 2, 1, 2), HC_CAND(2, 4, 2, 2, 0, 2), HC_CAND(2, 4, 4, 2, 1, 1),
    HC_CAND(2, 4, 4, 2, 0, 2), HC_CAND(2, 8, 2, 2, 0, 1), HC_CAND(2, 8, 2, 2, 1, 1), HC_CAND(2, 8, 4, 1, 1, 1),
    HC_QP(2, 4, 2, 2), HC_QP(2, 4, 4, 2),
    HC_CAND(3, 4, 2, 2, 1, 1), HC_CAND(3, 4, 2, 2, 1, 2), HC_CAND(3, 4, 2, 2, 0, 2), HC_CAND(3, 4, 4, 1, 1, 2),
    HC_CAND(3, 2, 2, 4, 1, 2), HC_CAND(3, 2, 2, 2, 1, 1), HC_CAND(3, 4, 2, 1, 1, 1), HC_CAND(3, 8, 2, 1, 1, 2),
    HC_CAND(3, 8, 1, 2, 1, 2), HC_CAND(3, 8, 2, 1, 1, 1), HC_CAND(3, 4, 2, 1, 1, 2), HC_CAND(3, 2, 2, 2, 1, 2),
};
#undef HC_CAND
#undef HC_QP

int
tile_candidates(int p, const TileCandidate ** out)
{
  int first = -1, n = 0;
  for (int i = 0; i < (int)(sizeof(g_cands) / sizeof(g_cands[0])); i++)
    if (g_cands[i].p == p)
    {
      if (first < 0)
        first = i;
      n++;
    }
  *out = first >= 0 ? &g_cands[first] : nullptr;
  return n;
}

TileShape
default_tile(int p)
{
  const TileCandidate * c;
  tile_candidates(p, &c);
  return c[0].shape;
}

template <typename T>
hipError_t
apply_op(const Geom & g, const PlaneSegs & segs, const Tables<T> & tb, const T * qd, const ElemFactors & ef,
         const T * x, T * y, T * part, const int * done, hipStream_t st, const TileCandidate * cand, int nrhs,
         int done_stride, int rb, ApplyDesc * desc_dev, const T * resid_b)
{
  if (!cand)
  {
    const TileCandidate * c0;
    const int n = tile_candidates(g.p, &c0);
    for (int i = 0; i < n && !cand; i++)
      if (c0[i].shape.tx == g.T[0] && c0[i].shape.ty == g.T[1] && c0[i].shape.tz == g.T[2])
        cand = &c0[i];
    if (!cand)
      return hipErrorInvalidValue;
  }
  ApplyDesc d{};
  d.g = g;
  d.segs = segs;
  d.prows = (int)segs.start[segs.nseg];
  d.qd = qd;
  d.ef = ef;
  d.resid = resid_b;
  d.x = x;
  d.y = y;
  d.part = part;
  d.done = done;
  d.done_stride = done_stride;
  d.nrhs = nrhs;
  d.rb = 1;
  d.w0 = 0;
  d.ntiles = g.nt[0] * g.nt[1] * g.nt[2];
  d.pw0 = 0;
  hipError_t e;
  if (rb > 1 && nrhs > 1)
  {
    if (!desc_dev)
      return hipErrorInvalidValue;
    d.rb = rb < nrhs ? rb : nrhs;
    if ((e = hipMemcpyAsync(desc_dev, &d, sizeof(d), hipMemcpyHostToDevice, st)) != hipSuccess)
      return e;
    e = cand->launch(d, desc_dev, 1, d.ntiles * ((nrhs + d.rb - 1) / d.rb), 1 | (ef.qc ? LAUNCH_EF : 0), &tb,
                     sizeof(T) == 8 ? 64 : 32, st);
  }
  else
    e = cand->launch(d, nullptr, 1, d.ntiles * nrhs, 0, &tb, sizeof(T) == 8 ? 64 : 32, st);
  if (e != hipSuccess)
    return e;
  return assemble_planes<T>(d, nullptr, 1, d.prows * nrhs, st);
}

template <typename T>
hipError_t
apply_scaled(const Geom & g, const PlaneSegs & segs, const Tables<T> & tb, const T * qd, const ElemFactors & ef,
             const T * x, T * y, T * part, double mscale, double cscale, const TileCandidate * cand,
             ApplyDesc * desc_dev, hipStream_t st)
{
  ApplyDesc d{};
  d.g = g;
  d.segs = segs;
  d.prows = (int)segs.start[segs.nseg];
  d.qd = qd;
  d.ef = ef;
  d.x = x;
  d.y = y;
  d.part = part;
  d.nrhs = 1;
  d.rb = 1;
  d.ntiles = g.nt[0] * g.nt[1] * g.nt[2];
  d.mscale = mscale;
  d.cscale = cscale;
  // term weights are honoured by the batched kernel (descriptor in device memory)
  hipError_t e = hipMemcpyAsync(desc_dev, &d, sizeof(d), hipMemcpyHostToDevice, st);
  if (e != hipSuccess)
    return e;
  e = cand->launch(d, desc_dev, 1, d.ntiles, ef.qc ? LAUNCH_EF : 0, &tb, sizeof(T) == 8 ? 64 : 32, st);
  if (e != hipSuccess)
    return e;
  return assemble_planes<T>(d, nullptr, 1, d.prows, st);
}
template hipError_t apply_scaled<double>(const Geom &, const PlaneSegs &, const Tables<double> &, const double *,
                                         const ElemFactors &, const double *, double *, double *, double, double,
                                         const TileCandidate *, ApplyDesc *, hipStream_t);

int
default_rhs_block(int ntiles, int nrhs)
{
  // One right-hand side per work item.  The tile kernel is bound by
  // instruction issue and latency, not by its quadrature-data reads: holding a
  // tile's factors in registers across a group of right-hand sides raises the
  // register count and cuts the waves per SIMD, while the work items of one
  // tile (consecutive, right-hand side fastest, on one XCD) re-read the
  // factors from L2.  Measured on 32^3 order 2 with 32 right-hand sides:
  // 31 us per rhs at rb = 1 vs 42 us at rb = 32 (FP64); on 24^3 order 3 with
  // 16: 28 vs 50 us.  Groups stay available through the explicit setting.
  (void)ntiles;
  (void)nrhs;
  return 1;
}

template hipError_t apply_op<double>(const Geom &, const PlaneSegs &, const Tables<double> &, const double *,
                                     const ElemFactors &, const double *, double *, double *, const int *,
                                     hipStream_t, const TileCandidate *, int, int, int, ApplyDesc *,
                                     const double *);
template hipError_t apply_op<float>(const Geom &, const PlaneSegs &, const Tables<float> &, const float *,
                                    const ElemFactors &, const float *, float *, float *, const int *, hipStream_t,
                                    const TileCandidate *, int, int, int, ApplyDesc *, const float *);

template <typename T>
hipError_t
apply_multi(const TileCandidate * cand, const Tables<T> & tb, const ApplyDesc * descs_dev, int ndesc, int nwork,
            int npwork, int grouped, hipStream_t st)
{
  ApplyDesc dummy{};
  hipError_t e = cand->launch(dummy, descs_dev, ndesc, nwork, grouped, &tb, sizeof(T) == 8 ? 64 : 32, st);
  if (e != hipSuccess)
    return e;
  return assemble_planes<T>(dummy, descs_dev, ndesc, npwork, st);
}

template hipError_t apply_multi<double>(const TileCandidate *, const Tables<double> &, const ApplyDesc *, int, int, int,
                                        int, hipStream_t);
template hipError_t apply_multi<float>(const TileCandidate *, const Tables<float> &, const ApplyDesc *, int, int, int, int,
                                       hipStream_t);

} // namespace hc
