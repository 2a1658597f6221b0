/*
 * hc_oracle.c -- CPU FP64 oracle for the H(curl) curl-curl + sigma-mass path.
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py as the checker / CPU baseline.  The product
 * library never links or calls this code.
 *
 * Reference behaviour restated (paths relative to the hephaestus source root):
 *   hcurl_formulation.cpp:130-157   operator = dt*alpha K + beta M (kernels)
 *   curl_curl_kernel.cpp:20-24      CurlCurlIntegrator(coef)   -> K part
 *   vector_fe_mass_kernel.cpp:20-24 VectorFEMassIntegrator(coef) -> M part
 *   equation_system.cpp:132-190     FormLinearSystem: PEC rows/cols eliminated
 *   equation_system.cpp:200-204     Mult: y = A x
 *   boundary_conditions.cpp:32-51   essential (tangential) true DOFs
 *   hephaestus_solvers.hpp:34-57    DefaultJacobiPCGSolver (PCG + DiagScale)
 * The MFEM/hypre arithmetic behind those calls is restated from its published
 * mathematics (ND_p hex = open Gauss-Legendre x closed Gauss-Lobatto tensor
 * basis, covariant Piola transform, PCG with diagonal scaling).
 */
#include "hc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define DMAX (HCO_PMAX + 1)
#define QMAX (HCO_PMAX + 1)
#define Q3MAX (QMAX * QMAX * QMAX)
#define NDMAX (3 * HCO_PMAX * DMAX * DMAX)

int
hco_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* 1D points, weights, Lagrange tables                                 */
/* ------------------------------------------------------------------ */

/* Legendre P_n(x) and P_n'(x) on [-1,1] by the three-term recurrence. */
static void
legendre(int n, double x, double * P, double * dP)
{
  double p0 = 1.0, p1 = x;
  if (n == 0)
  {
    *P = 1.0;
    *dP = 0.0;
    return;
  }
  for (int k = 2; k <= n; k++)
  {
    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *P = p1;
  *dP = n * (x * p1 - p0) / (x * x - 1.0);
}

void
hco_gauss_legendre(int n, double * x, double * w)
{
  for (int i = 0; i < n; i++)
  {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; it++)
    {
      double P, dP;
      legendre(n, z, &P, &dP);
      double dz = P / dP;
      z -= dz;
      if (fabs(dz) < 1e-16)
        break;
    }
    double P, dP;
    legendre(n, z, &P, &dP);
    /* ascending order on [0,1] */
    x[n - 1 - i] = 0.5 * (z + 1.0);
    w[n - 1 - i] = 1.0 / ((1.0 - z * z) * dP * dP); /* (2/((1-z^2)P'^2))/2 */
  }
}

void
hco_gauss_lobatto(int n, double * x)
{
  /* n points: endpoints plus roots of P'_{n-1}. */
  int p = n - 1;
  x[0] = 0.0;
  x[n - 1] = 1.0;
  for (int i = 1; i < p; i++)
  {
    double z = -cos(M_PI * i / p);
    for (int it = 0; it < 100; it++)
    {
      double P, dP;
      legendre(p, z, &P, &dP);
      double d2P = (2.0 * z * dP - p * (p + 1.0) * P) / (1.0 - z * z);
      double dz = dP / d2P;
      z -= dz;
      if (fabs(dz) < 1e-16)
        break;
    }
    x[i] = 0.5 * (z + 1.0);
  }
}

static void
lagrange_eval(int n, const double * nodes, double t, double * val, double * der)
{
  for (int j = 0; j < n; j++)
  {
    double v = 1.0;
    for (int m = 0; m < n; m++)
      if (m != j)
        v *= (t - nodes[m]) / (nodes[j] - nodes[m]);
    val[j] = v;
    if (der)
    {
      double d = 0.0;
      for (int k = 0; k < n; k++)
      {
        if (k == j)
          continue;
        double prod = 1.0 / (nodes[j] - nodes[k]);
        for (int m = 0; m < n; m++)
          if (m != j && m != k)
            prod *= (t - nodes[m]) / (nodes[j] - nodes[m]);
        d += prod;
      }
      der[j] = d;
    }
  }
}

/* Tables at an arbitrary set of nq points. */
static void
tables_at(int p, int nq, const double * pts, double * Bo, double * Bc, double * Gc)
{
  double on[HCO_PMAX], cn[DMAX], ow[HCO_PMAX];
  hco_gauss_legendre(p, on, ow);
  hco_gauss_lobatto(p + 1, cn);
  for (int q = 0; q < nq; q++)
  {
    lagrange_eval(p, on, pts[q], Bo + q * p, NULL);
    lagrange_eval(p + 1, cn, pts[q], Bc + q * (p + 1), Gc + q * (p + 1));
  }
}

int
hco_basis(int p, double * Bo, double * Bc, double * Gc, double * qpts, double * qw)
{
  if (p < 1 || p > HCO_PMAX)
    return -1;
  int Q = p + 1;
  hco_gauss_legendre(Q, qpts, qw);
  tables_at(p, Q, qpts, Bo, Bc, Gc);
  return 0;
}

typedef struct
{
  int p, D, O, Q;
  double Bo[QMAX * HCO_PMAX], Bc[QMAX * DMAX], Gc[QMAX * DMAX], qp[QMAX], qw[QMAX];
} tabs_t;

static void
make_tabs(int p, tabs_t * t)
{
  t->p = p;
  t->O = p;
  t->D = p + 1;
  t->Q = p + 1;
  hco_basis(p, t->Bo, t->Bc, t->Gc, t->qp, t->qw);
}

/* ------------------------------------------------------------------ */
/* Mesh / DOF numbering                                                 */
/* ------------------------------------------------------------------ */

void
hco_comp_dims(const hco_mesh * m, int p, int c, int64_t dims[3])
{
  int64_t ne[3] = {m->nx, m->ny, m->nz};
  for (int d = 0; d < 3; d++)
    dims[d] = (d == c) ? ne[d] * p : ne[d] * p + 1;
}

static void
comp_offsets(const hco_mesh * m, int p, int64_t off[4])
{
  off[0] = 0;
  for (int c = 0; c < 3; c++)
  {
    int64_t d[3];
    hco_comp_dims(m, p, c, d);
    off[c + 1] = off[c] + d[0] * d[1] * d[2];
  }
}

int64_t
hco_ndofs(const hco_mesh * m, int p)
{
  int64_t off[4];
  comp_offsets(m, p, off);
  return off[3];
}

void
hco_ess_mask(const hco_mesh * m, int p, uint8_t * ess)
{
  int64_t off[4];
  comp_offsets(m, p, off);
  for (int c = 0; c < 3; c++)
  {
    int64_t d[3];
    hco_comp_dims(m, p, c, d);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < d[2]; k++)
      for (int64_t j = 0; j < d[1]; j++)
        for (int64_t i = 0; i < d[0]; i++)
        {
          int64_t idx[3] = {i, j, k};
          int e = 0;
          for (int a = 0; a < 3; a++)
          {
            if (a == c)
              continue; /* tangential: closed directions only */
            if (idx[a] == 0 && (m->pec_faces >> (2 * a) & 1))
              e = 1;
            if (idx[a] == d[a] - 1 && (m->pec_faces >> (2 * a + 1) & 1))
              e = 1;
          }
          ess[off[c] + (k * d[1] + j) * d[0] + i] = (uint8_t)e;
        }
  }
}

void
hco_elem_dofs(const hco_mesh * m, int p, int64_t e, int64_t * gid)
{
  int64_t off[4];
  comp_offsets(m, p, off);
  int64_t ex = e % m->nx, ey = (e / m->nx) % m->ny, ez = e / ((int64_t)m->nx * m->ny);
  int D = p + 1, O = p;
  int l = 0;
  for (int c = 0; c < 3; c++)
  {
    int64_t d[3];
    hco_comp_dims(m, p, c, d);
    int n0 = (c == 0) ? O : D, n1 = (c == 1) ? O : D, n2 = (c == 2) ? O : D;
    for (int dz = 0; dz < n2; dz++)
      for (int dy = 0; dy < n1; dy++)
        for (int dx = 0; dx < n0; dx++)
        {
          int64_t ix = ex * p + dx, iy = ey * p + dy, iz = ez * p + dz;
          gid[l++] = off[c] + (iz * d[1] + iy) * d[0] + ix;
        }
  }
}

/* Jacobian of the trilinear element map at reference point (a,b,c). */
static void
elem_jacobian(const hco_mesh * m, int64_t e, double a, double b, double c, double J[9])
{
  int64_t ex = e % m->nx, ey = (e / m->nx) % m->ny, ez = e / ((int64_t)m->nx * m->ny);
  if (!m->vert)
  {
    memset(J, 0, 9 * sizeof(double));
    J[0] = m->len[0] / m->nx;
    J[4] = m->len[1] / m->ny;
    J[8] = m->len[2] / m->nz;
    return;
  }
  double L[2][3] = {{1 - a, 1 - b, 1 - c}, {a, b, c}};
  double dL[2] = {-1.0, 1.0};
  memset(J, 0, 9 * sizeof(double));
  int64_t sx = 1, sy = m->nx + 1, sz = (int64_t)(m->nx + 1) * (m->ny + 1);
  for (int kc = 0; kc < 2; kc++)
    for (int kb = 0; kb < 2; kb++)
      for (int ka = 0; ka < 2; ka++)
      {
        const double * V = m->vert + 3 * ((ex + ka) * sx + (ey + kb) * sy + (ez + kc) * sz);
        double g0 = dL[ka] * L[kb][1] * L[kc][2];
        double g1 = L[ka][0] * dL[kb] * L[kc][2];
        double g2 = L[ka][0] * L[kb][1] * dL[kc];
        for (int r = 0; r < 3; r++)
        {
          J[r * 3 + 0] += V[r] * g0;
          J[r * 3 + 1] += V[r] * g1;
          J[r * 3 + 2] += V[r] * g2;
        }
      }
}

static void
elem_point(const hco_mesh * m, int64_t e, double a, double b, double c, double X[3])
{
  int64_t ex = e % m->nx, ey = (e / m->nx) % m->ny, ez = e / ((int64_t)m->nx * m->ny);
  if (!m->vert)
  {
    X[0] = m->x0[0] + (ex + a) * m->len[0] / m->nx;
    X[1] = m->x0[1] + (ey + b) * m->len[1] / m->ny;
    X[2] = m->x0[2] + (ez + c) * m->len[2] / m->nz;
    return;
  }
  double L[2][3] = {{1 - a, 1 - b, 1 - c}, {a, b, c}};
  X[0] = X[1] = X[2] = 0.0;
  int64_t sx = 1, sy = m->nx + 1, sz = (int64_t)(m->nx + 1) * (m->ny + 1);
  for (int kc = 0; kc < 2; kc++)
    for (int kb = 0; kb < 2; kb++)
      for (int ka = 0; ka < 2; ka++)
      {
        const double * V = m->vert + 3 * ((ex + ka) * sx + (ey + kb) * sy + (ez + kc) * sz);
        double g = L[ka][0] * L[kb][1] * L[kc][2];
        for (int r = 0; r < 3; r++)
          X[r] += V[r] * g;
      }
}

/* ------------------------------------------------------------------ */
/* PA setup                                                             */
/* ------------------------------------------------------------------ */

/* Symmetric factors at one point: mass = ms * |det| (J^T J)^{-1},
 * curl = cs * J^T J / |det|. */
static void
point_factors(const double J[9], double ms, double cs, double out[12])
{
  double G[9];
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 3; s++)
      G[r * 3 + s] = J[0 * 3 + r] * J[0 * 3 + s] + J[1 * 3 + r] * J[1 * 3 + s] + J[2 * 3 + r] * J[2 * 3 + s];
  double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
               J[2] * (J[3] * J[7] - J[4] * J[6]);
  double adet = fabs(det);
  /* adj(G) (G symmetric); det(G) = det^2 */
  double A00 = G[4] * G[8] - G[5] * G[7];
  double A01 = -(G[1] * G[8] - G[2] * G[7]);
  double A02 = G[1] * G[5] - G[2] * G[4];
  double A11 = G[0] * G[8] - G[2] * G[6];
  double A12 = -(G[0] * G[5] - G[2] * G[3]);
  double A22 = G[0] * G[4] - G[1] * G[3];
  double sm = ms / adet; /* ms*|det| * adj(G)/det^2 */
  out[0] = sm * A00;
  out[1] = sm * A01;
  out[2] = sm * A02;
  out[3] = sm * A11;
  out[4] = sm * A12;
  out[5] = sm * A22;
  double sc = cs / adet;
  out[6] = sc * G[0];
  out[7] = sc * G[1];
  out[8] = sc * G[2];
  out[9] = sc * G[4];
  out[10] = sc * G[5];
  out[11] = sc * G[8];
}

void
hco_pa_setup(const hco_mesh * m,
             int p,
             const double * curl_coef,
             const double * mass_coef,
             double * qdata)
{
  tabs_t t;
  make_tabs(p, &t);
  int Q = t.Q;
  int64_t E = (int64_t)m->nx * m->ny * m->nz;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; e++)
  {
    for (int qz = 0; qz < Q; qz++)
      for (int qy = 0; qy < Q; qy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double J[9], f[12];
          elem_jacobian(m, e, t.qp[qx], t.qp[qy], t.qp[qz], J);
          double w = t.qw[qx] * t.qw[qy] * t.qw[qz];
          point_factors(J, w * mass_coef[e], w * curl_coef[e], f);
          for (int k = 0; k < 12; k++)
            qdata[((e * Q + qz) * 12 + k) * Q * Q + qy * Q + qx] = f[k];
        }
  }
}

/* ------------------------------------------------------------------ */
/* Sum-factorised element apply                                         */
/* ------------------------------------------------------------------ */

#define QD(k, q) qd[((q) / (Q * Q) * 12 + (k)) * Q * Q + (q) % (Q * Q)]

static void
elem_apply(const tabs_t * t, const double * qd, const double * xe, double * ye)
{
  const int O = t->O, D = t->D, Q = t->Q;
  const double *Bo = t->Bo, *Bc = t->Bc, *Gc = t->Gc;
  const int nc = O * D * D;
  double U[3][Q3MAX], C[3][Q3MAX];
  double A1[DMAX * DMAX * QMAX], A2[DMAX * DMAX * QMAX];
  double B1[DMAX * QMAX * QMAX], B2[DMAX * QMAX * QMAX], B3[DMAX * QMAX * QMAX];
  const int Q3 = Q * Q * Q;
  for (int q = 0; q < Q3; q++)
    C[0][q] = C[1][q] = C[2][q] = 0.0;

  /* ---- component x: X[dz][dy][dx], (D, D, O) ---- */
  {
    const double * X = xe;
    for (int dz = 0; dz < D; dz++)
      for (int dy = 0; dy < D; dy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double s = 0;
          for (int dx = 0; dx < O; dx++)
            s += Bo[qx * O + dx] * X[(dz * D + dy) * O + dx];
          A1[(dz * D + dy) * Q + qx] = s;
        }
    for (int dz = 0; dz < D; dz++)
      for (int qy = 0; qy < Q; qy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double s = 0, g = 0;
          for (int dy = 0; dy < D; dy++)
          {
            double a = A1[(dz * D + dy) * Q + qx];
            s += Bc[qy * D + dy] * a;
            g += Gc[qy * D + dy] * a;
          }
          B1[(dz * Q + qy) * Q + qx] = s;
          B2[(dz * Q + qy) * Q + qx] = g;
        }
    for (int qz = 0; qz < Q; qz++)
      for (int qy = 0; qy < Q; qy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double u = 0, uz = 0, uy = 0;
          for (int dz = 0; dz < D; dz++)
          {
            double b1 = B1[(dz * Q + qy) * Q + qx], b2 = B2[(dz * Q + qy) * Q + qx];
            u += Bc[qz * D + dz] * b1;
            uz += Gc[qz * D + dz] * b1;
            uy += Bc[qz * D + dz] * b2;
          }
          int q = (qz * Q + qy) * Q + qx;
          U[0][q] = u;
          C[1][q] += uz;
          C[2][q] -= uy;
        }
  }
  /* ---- component y: Y[dz][dy][dx], (D, O, D) ---- */
  {
    const double * Y = xe + nc;
    for (int dz = 0; dz < D; dz++)
      for (int dy = 0; dy < O; dy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double s = 0, g = 0;
          for (int dx = 0; dx < D; dx++)
          {
            double v = Y[(dz * O + dy) * D + dx];
            s += Bc[qx * D + dx] * v;
            g += Gc[qx * D + dx] * v;
          }
          A1[(dz * O + dy) * Q + qx] = s;
          A2[(dz * O + dy) * Q + qx] = g;
        }
    for (int dz = 0; dz < D; dz++)
      for (int qy = 0; qy < Q; qy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double s = 0, g = 0;
          for (int dy = 0; dy < O; dy++)
          {
            s += Bo[qy * O + dy] * A1[(dz * O + dy) * Q + qx];
            g += Bo[qy * O + dy] * A2[(dz * O + dy) * Q + qx];
          }
          B1[(dz * Q + qy) * Q + qx] = s;
          B2[(dz * Q + qy) * Q + qx] = g;
        }
    for (int qz = 0; qz < Q; qz++)
      for (int qy = 0; qy < Q; qy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double u = 0, uz = 0, ux = 0;
          for (int dz = 0; dz < D; dz++)
          {
            double b1 = B1[(dz * Q + qy) * Q + qx], b2 = B2[(dz * Q + qy) * Q + qx];
            u += Bc[qz * D + dz] * b1;
            uz += Gc[qz * D + dz] * b1;
            ux += Bc[qz * D + dz] * b2;
          }
          int q = (qz * Q + qy) * Q + qx;
          U[1][q] = u;
          C[0][q] -= uz;
          C[2][q] += ux;
        }
  }
  /* ---- component z: Z[dz][dy][dx], (O, D, D) ---- */
  {
    const double * Z = xe + 2 * nc;
    for (int dz = 0; dz < O; dz++)
      for (int dy = 0; dy < D; dy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double s = 0, g = 0;
          for (int dx = 0; dx < D; dx++)
          {
            double v = Z[(dz * D + dy) * D + dx];
            s += Bc[qx * D + dx] * v;
            g += Gc[qx * D + dx] * v;
          }
          A1[(dz * D + dy) * Q + qx] = s;
          A2[(dz * D + dy) * Q + qx] = g;
        }
    for (int dz = 0; dz < O; dz++)
      for (int qy = 0; qy < Q; qy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double s = 0, gy = 0, gx = 0;
          for (int dy = 0; dy < D; dy++)
          {
            double a1 = A1[(dz * D + dy) * Q + qx], a2 = A2[(dz * D + dy) * Q + qx];
            s += Bc[qy * D + dy] * a1;
            gy += Gc[qy * D + dy] * a1;
            gx += Bc[qy * D + dy] * a2;
          }
          B1[(dz * Q + qy) * Q + qx] = s;
          B2[(dz * Q + qy) * Q + qx] = gy;
          B3[(dz * Q + qy) * Q + qx] = gx;
        }
    for (int qz = 0; qz < Q; qz++)
      for (int qy = 0; qy < Q; qy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double u = 0, uy = 0, ux = 0;
          for (int dz = 0; dz < O; dz++)
          {
            int i = (dz * Q + qy) * Q + qx;
            u += Bo[qz * O + dz] * B1[i];
            uy += Bo[qz * O + dz] * B2[i];
            ux += Bo[qz * O + dz] * B3[i];
          }
          int q = (qz * Q + qy) * Q + qx;
          U[2][q] = u;
          C[0][q] += uy;
          C[1][q] -= ux;
        }
  }
  /* ---- quadrature-point factors ---- */
  double V[3][Q3MAX], W[3][Q3MAX];
  for (int q = 0; q < Q3; q++)
  {
    double m0 = QD(0, q), m1 = QD(1, q), m2 = QD(2, q), m3 = QD(3, q), m4 = QD(4, q), m5 = QD(5, q);
    double k0 = QD(6, q), k1 = QD(7, q), k2 = QD(8, q), k3 = QD(9, q), k4 = QD(10, q), k5 = QD(11, q);
    double u0 = U[0][q], u1 = U[1][q], u2 = U[2][q];
    double c0 = C[0][q], c1 = C[1][q], c2 = C[2][q];
    V[0][q] = m0 * u0 + m1 * u1 + m2 * u2;
    V[1][q] = m1 * u0 + m3 * u1 + m4 * u2;
    V[2][q] = m2 * u0 + m4 * u1 + m5 * u2;
    W[0][q] = k0 * c0 + k1 * c1 + k2 * c2;
    W[1][q] = k1 * c0 + k3 * c1 + k4 * c2;
    W[2][q] = k2 * c0 + k4 * c1 + k5 * c2;
  }
  /* ---- backward, component x ---- */
  {
    double * Yx = ye;
    for (int dz = 0; dz < D; dz++)
      for (int qy = 0; qy < Q; qy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double sa = 0, sb = 0;
          for (int qz = 0; qz < Q; qz++)
          {
            int q = (qz * Q + qy) * Q + qx;
            sa += Bc[qz * D + dz] * V[0][q] + Gc[qz * D + dz] * W[1][q];
            sb -= Bc[qz * D + dz] * W[2][q];
          }
          B1[(dz * Q + qy) * Q + qx] = sa;
          B2[(dz * Q + qy) * Q + qx] = sb;
        }
    for (int dz = 0; dz < D; dz++)
      for (int dy = 0; dy < D; dy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double s = 0;
          for (int qy = 0; qy < Q; qy++)
            s += Bc[qy * D + dy] * B1[(dz * Q + qy) * Q + qx] + Gc[qy * D + dy] * B2[(dz * Q + qy) * Q + qx];
          A1[(dz * D + dy) * Q + qx] = s;
        }
    for (int dz = 0; dz < D; dz++)
      for (int dy = 0; dy < D; dy++)
        for (int dx = 0; dx < O; dx++)
        {
          double s = 0;
          for (int qx = 0; qx < Q; qx++)
            s += Bo[qx * O + dx] * A1[(dz * D + dy) * Q + qx];
          Yx[(dz * D + dy) * O + dx] += s;
        }
  }
  /* ---- backward, component y ---- */
  {
    double * Yy = ye + nc;
    for (int dz = 0; dz < D; dz++)
      for (int qy = 0; qy < Q; qy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double sa = 0, sb = 0;
          for (int qz = 0; qz < Q; qz++)
          {
            int q = (qz * Q + qy) * Q + qx;
            sa += Bc[qz * D + dz] * V[1][q] - Gc[qz * D + dz] * W[0][q];
            sb += Bc[qz * D + dz] * W[2][q];
          }
          B1[(dz * Q + qy) * Q + qx] = sa;
          B2[(dz * Q + qy) * Q + qx] = sb;
        }
    for (int dz = 0; dz < D; dz++)
      for (int dy = 0; dy < O; dy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double ta = 0, tb = 0;
          for (int qy = 0; qy < Q; qy++)
          {
            ta += Bo[qy * O + dy] * B1[(dz * Q + qy) * Q + qx];
            tb += Bo[qy * O + dy] * B2[(dz * Q + qy) * Q + qx];
          }
          A1[(dz * O + dy) * Q + qx] = ta;
          A2[(dz * O + dy) * Q + qx] = tb;
        }
    for (int dz = 0; dz < D; dz++)
      for (int dy = 0; dy < O; dy++)
        for (int dx = 0; dx < D; dx++)
        {
          double s = 0;
          for (int qx = 0; qx < Q; qx++)
            s += Bc[qx * D + dx] * A1[(dz * O + dy) * Q + qx] + Gc[qx * D + dx] * A2[(dz * O + dy) * Q + qx];
          Yy[(dz * O + dy) * D + dx] += s;
        }
  }
  /* ---- backward, component z ---- */
  {
    double * Yz = ye + 2 * nc;
    for (int dz = 0; dz < O; dz++)
      for (int qy = 0; qy < Q; qy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double sa = 0, sb = 0, sc = 0;
          for (int qz = 0; qz < Q; qz++)
          {
            int q = (qz * Q + qy) * Q + qx;
            double bo = Bo[qz * O + dz];
            sa += bo * V[2][q];
            sb += bo * W[0][q];
            sc -= bo * W[1][q];
          }
          B1[(dz * Q + qy) * Q + qx] = sa;
          B2[(dz * Q + qy) * Q + qx] = sb;
          B3[(dz * Q + qy) * Q + qx] = sc;
        }
    for (int dz = 0; dz < O; dz++)
      for (int dy = 0; dy < D; dy++)
        for (int qx = 0; qx < Q; qx++)
        {
          double ta = 0, tb = 0;
          for (int qy = 0; qy < Q; qy++)
          {
            int i = (dz * Q + qy) * Q + qx;
            ta += Bc[qy * D + dy] * B1[i] + Gc[qy * D + dy] * B2[i];
            tb += Bc[qy * D + dy] * B3[i];
          }
          A1[(dz * D + dy) * Q + qx] = ta;
          A2[(dz * D + dy) * Q + qx] = tb;
        }
    for (int dz = 0; dz < O; dz++)
      for (int dy = 0; dy < D; dy++)
        for (int dx = 0; dx < D; dx++)
        {
          double s = 0;
          for (int qx = 0; qx < Q; qx++)
            s += Bc[qx * D + dx] * A1[(dz * D + dy) * Q + qx] + Gc[qx * D + dx] * A2[(dz * D + dy) * Q + qx];
          Yz[(dz * D + dy) * D + dx] += s;
        }
  }
}

void
hco_pa_apply_raw(const hco_mesh * m,
                 int p,
                 const double * qdata,
                 int64_t e0,
                 int64_t e1,
                 const double * x,
                 double * y)
{
  tabs_t t;
  make_tabs(p, &t);
  int Q = t.Q, nd = 3 * p * (p + 1) * (p + 1);
  for (int64_t e = e0; e < e1; e++)
  {
    int64_t gid[NDMAX];
    double xe[NDMAX], ye[NDMAX];
    hco_elem_dofs(m, p, e, gid);
    for (int i = 0; i < nd; i++)
    {
      xe[i] = x[gid[i]];
      ye[i] = 0.0;
    }
    elem_apply(&t, qdata + e * 12 * Q * Q * Q, xe, ye);
    for (int i = 0; i < nd; i++)
      y[gid[i]] += ye[i];
  }
}

void
hco_pa_apply(const hco_mesh * m,
             int p,
             const double * qdata,
             const uint8_t * ess,
             const double * x,
             double * y)
{
  tabs_t t;
  make_tabs(p, &t);
  const int Q = t.Q, nd = 3 * p * (p + 1) * (p + 1);
  const int64_t N = hco_ndofs(m, p);
  const int nx = m->nx, ny = m->ny, nz = m->nz;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++)
    y[i] = 0.0;
  /* 8-colour element sweep: same-colour elements share no DOF. */
  for (int col = 0; col < 8; col++)
  {
    int cx = col & 1, cy = (col >> 1) & 1, cz = (col >> 2) & 1;
    int64_t hx = (nx - cx + 1) / 2, hy = (ny - cy + 1) / 2, hz = (nz - cz + 1) / 2;
    int64_t nsub = hx * hy * hz;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < nsub; s++)
    {
      int64_t ex = cx + 2 * (s % hx), ey = cy + 2 * ((s / hx) % hy), ez = cz + 2 * (s / (hx * hy));
      int64_t e = (ez * ny + ey) * nx + ex;
      int64_t gid[NDMAX];
      double xe[NDMAX], ye[NDMAX];
      hco_elem_dofs(m, p, e, gid);
      for (int i = 0; i < nd; i++)
      {
        xe[i] = ess[gid[i]] ? 0.0 : x[gid[i]];
        ye[i] = 0.0;
      }
      elem_apply(&t, qdata + e * 12 * Q * Q * Q, xe, ye);
      for (int i = 0; i < nd; i++)
        y[gid[i]] += ye[i];
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++)
    if (ess[i])
      y[i] = x[i];
}

/* ------------------------------------------------------------------ */
/* Diagonal                                                             */
/* ------------------------------------------------------------------ */

static void
elem_diag(const tabs_t * t, const double * qd, double * de)
{
  const int O = t->O, D = t->D, Q = t->Q;
  const double *Bo = t->Bo, *Bc = t->Bc, *Gc = t->Gc;
  int l = 0;
  /* component x: phi = (a b c, 0, 0), curl = (0, a b c', -a b' c) */
  for (int dz = 0; dz < D; dz++)
    for (int dy = 0; dy < D; dy++)
      for (int dx = 0; dx < O; dx++)
      {
        double s = 0;
        for (int qz = 0; qz < Q; qz++)
          for (int qy = 0; qy < Q; qy++)
            for (int qx = 0; qx < Q; qx++)
            {
              int q = (qz * Q + qy) * Q + qx;
              double a = Bo[qx * O + dx], b = Bc[qy * D + dy], bp = Gc[qy * D + dy];
              double c = Bc[qz * D + dz], cp = Gc[qz * D + dz];
              double v = a * b * c, cy = a * b * cp, cz = -a * bp * c;
              s += v * v * QD(0, q) + cy * cy * QD(9, q) + cz * cz * QD(11, q) + 2 * cy * cz * QD(10, q);
            }
        de[l++] = s;
      }
  /* component y: phi = (0, a b c, 0), curl = (-a b c', 0, a' b c) */
  for (int dz = 0; dz < D; dz++)
    for (int dy = 0; dy < O; dy++)
      for (int dx = 0; dx < D; dx++)
      {
        double s = 0;
        for (int qz = 0; qz < Q; qz++)
          for (int qy = 0; qy < Q; qy++)
            for (int qx = 0; qx < Q; qx++)
            {
              int q = (qz * Q + qy) * Q + qx;
              double a = Bc[qx * D + dx], ap = Gc[qx * D + dx], b = Bo[qy * O + dy];
              double c = Bc[qz * D + dz], cp = Gc[qz * D + dz];
              double v = a * b * c, cx = -a * b * cp, cz = ap * b * c;
              s += v * v * QD(3, q) + cx * cx * QD(6, q) + cz * cz * QD(11, q) + 2 * cx * cz * QD(8, q);
            }
        de[l++] = s;
      }
  /* component z: phi = (0, 0, a b c), curl = (a b' c, -a' b c, 0) */
  for (int dz = 0; dz < O; dz++)
    for (int dy = 0; dy < D; dy++)
      for (int dx = 0; dx < D; dx++)
      {
        double s = 0;
        for (int qz = 0; qz < Q; qz++)
          for (int qy = 0; qy < Q; qy++)
            for (int qx = 0; qx < Q; qx++)
            {
              int q = (qz * Q + qy) * Q + qx;
              double a = Bc[qx * D + dx], ap = Gc[qx * D + dx], b = Bc[qy * D + dy], bp = Gc[qy * D + dy];
              double c = Bo[qz * O + dz];
              double v = a * b * c, cx = a * bp * c, cy = -ap * b * c;
              s += v * v * QD(5, q) + cx * cx * QD(6, q) + cy * cy * QD(9, q) + 2 * cx * cy * QD(7, q);
            }
        de[l++] = s;
      }
}

void
hco_pa_diagonal(const hco_mesh * m, int p, const double * qdata, const uint8_t * ess, double * diag)
{
  tabs_t t;
  make_tabs(p, &t);
  const int Q = t.Q, nd = 3 * p * (p + 1) * (p + 1);
  const int64_t N = hco_ndofs(m, p);
  const int nx = m->nx, ny = m->ny, nz = m->nz;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++)
    diag[i] = 0.0;
  for (int col = 0; col < 8; col++)
  {
    int cx = col & 1, cy = (col >> 1) & 1, cz = (col >> 2) & 1;
    int64_t hx = (nx - cx + 1) / 2, hy = (ny - cy + 1) / 2, hz = (nz - cz + 1) / 2;
    int64_t nsub = hx * hy * hz;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < nsub; s++)
    {
      int64_t ex = cx + 2 * (s % hx), ey = cy + 2 * ((s / hx) % hy), ez = cz + 2 * (s / (hx * hy));
      int64_t e = (ez * ny + ey) * nx + ex;
      int64_t gid[NDMAX];
      double de[NDMAX];
      hco_elem_dofs(m, p, e, gid);
      elem_diag(&t, qdata + e * 12 * Q * Q * Q, de);
      for (int i = 0; i < nd; i++)
        diag[gid[i]] += de[i];
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++)
    if (ess[i])
      diag[i] = 1.0;
}

/* ------------------------------------------------------------------ */
/* Dense element matrices and CSR full assembly                         */
/* ------------------------------------------------------------------ */

/* Reference basis values and curls at all Q^3 points: phi[i][q][3], cphi[i][q][3]. */
static void
ref_basis(const tabs_t * t, double * phi, double * cphi)
{
  const int O = t->O, D = t->D, Q = t->Q, Q3 = Q * Q * Q;
  const double *Bo = t->Bo, *Bc = t->Bc, *Gc = t->Gc;
  int l = 0;
  for (int c = 0; c < 3; c++)
  {
    int n0 = (c == 0) ? O : D, n1 = (c == 1) ? O : D, n2 = (c == 2) ? O : D;
    for (int dz = 0; dz < n2; dz++)
      for (int dy = 0; dy < n1; dy++)
        for (int dx = 0; dx < n0; dx++, l++)
          for (int qz = 0; qz < Q; qz++)
            for (int qy = 0; qy < Q; qy++)
              for (int qx = 0; qx < Q; qx++)
              {
                int q = (qz * Q + qy) * Q + qx;
                double fx, fy, fz, gx = 0, gy = 0, gz = 0;
                /* 1D factors along each axis with their derivatives */
                if (c == 0)
                {
                  fx = Bo[qx * O + dx];
                  fy = Bc[qy * D + dy];
                  gy = Gc[qy * D + dy];
                  fz = Bc[qz * D + dz];
                  gz = Gc[qz * D + dz];
                }
                else if (c == 1)
                {
                  fx = Bc[qx * D + dx];
                  gx = Gc[qx * D + dx];
                  fy = Bo[qy * O + dy];
                  fz = Bc[qz * D + dz];
                  gz = Gc[qz * D + dz];
                }
                else
                {
                  fx = Bc[qx * D + dx];
                  gx = Gc[qx * D + dx];
                  fy = Bc[qy * D + dy];
                  gy = Gc[qy * D + dy];
                  fz = Bo[qz * O + dz];
                }
                double * ph = phi + ((size_t)l * Q3 + q) * 3;
                double * cp = cphi + ((size_t)l * Q3 + q) * 3;
                ph[0] = ph[1] = ph[2] = 0.0;
                ph[c] = fx * fy * fz;
                /* curl of f e_c */
                double dfx = gx * fy * fz, dfy = fx * gy * fz, dfz = fx * fy * gz;
                if (c == 0)
                {
                  cp[0] = 0.0;
                  cp[1] = dfz;
                  cp[2] = -dfy;
                }
                else if (c == 1)
                {
                  cp[0] = -dfz;
                  cp[1] = 0.0;
                  cp[2] = dfx;
                }
                else
                {
                  cp[0] = dfy;
                  cp[1] = -dfx;
                  cp[2] = 0.0;
                }
              }
  }
}

static void
elem_matrix_from_basis(int nd, int Q3, int Q, const double * phi, const double * cphi, const double * qd, double * Ae)
{
  for (int i = 0; i < nd; i++)
    for (int j = 0; j < nd; j++)
    {
      double s = 0.0;
      for (int q = 0; q < Q3; q++)
      {
        const double * a = phi + ((size_t)i * Q3 + q) * 3;
        const double * b = phi + ((size_t)j * Q3 + q) * 3;
        const double * ca = cphi + ((size_t)i * Q3 + q) * 3;
        const double * cb = cphi + ((size_t)j * Q3 + q) * 3;
        double M[9] = {QD(0, q), QD(1, q), QD(2, q), QD(1, q), QD(3, q), QD(4, q), QD(2, q), QD(4, q), QD(5, q)};
        double K[9] = {QD(6, q), QD(7, q), QD(8, q), QD(7, q), QD(9, q), QD(10, q), QD(8, q), QD(10, q), QD(11, q)};
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++)
            s += a[r] * M[r * 3 + c] * b[c] + ca[r] * K[r * 3 + c] * cb[c];
      }
      Ae[(size_t)i * nd + j] = s;
    }
}

void
hco_elem_matrix(const hco_mesh * m, int p, const double * qdata, int64_t e, double * Ae)
{
  tabs_t t;
  make_tabs(p, &t);
  const int Q = t.Q, Q3 = Q * Q * Q, nd = 3 * p * (p + 1) * (p + 1);
  double * phi = (double *)malloc(sizeof(double) * nd * Q3 * 3);
  double * cphi = (double *)malloc(sizeof(double) * nd * Q3 * 3);
  ref_basis(&t, phi, cphi);
  elem_matrix_from_basis(nd, Q3, Q, phi, cphi, qdata + e * 12 * Q3, Ae);
  free(phi);
  free(cphi);
  (void)m;
}

static int
cmp_i64(const void * a, const void * b)
{
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

/* Elements containing global DOF g (up to 4) and the local index in each. */
static int
dof_elements(const hco_mesh * m, int p, int64_t g, int64_t * elems, int * locs)
{
  int64_t off[4];
  comp_offsets(m, p, off);
  int c = (g >= off[1]) + (g >= off[2]);
  int64_t d[3];
  hco_comp_dims(m, p, c, d);
  int64_t r = g - off[c];
  int64_t idx[3] = {r % d[0], (r / d[0]) % d[1], r / (d[0] * d[1])};
  int64_t ne[3] = {m->nx, m->ny, m->nz};
  /* candidate element coordinate(s) and local index per axis */
  int64_t ec[3][2];
  int lc[3][2];
  int nc[3];
  for (int a = 0; a < 3; a++)
  {
    if (a == c)
    {
      ec[a][0] = idx[a] / p;
      lc[a][0] = (int)(idx[a] % p);
      nc[a] = 1;
    }
    else
    {
      nc[a] = 0;
      int64_t e_hi = idx[a] / p;
      int l_hi = (int)(idx[a] % p);
      if (e_hi < ne[a])
      {
        ec[a][nc[a]] = e_hi;
        lc[a][nc[a]] = l_hi;
        nc[a]++;
      }
      if (l_hi == 0 && e_hi > 0)
      {
        ec[a][nc[a]] = e_hi - 1;
        lc[a][nc[a]] = p;
        nc[a]++;
      }
    }
  }
  int O = p, D = p + 1;
  int n0 = (c == 0) ? O : D, n1 = (c == 1) ? O : D;
  int ncomp = O * D * D;
  int k = 0;
  for (int iz = 0; iz < nc[2]; iz++)
    for (int iy = 0; iy < nc[1]; iy++)
      for (int ix = 0; ix < nc[0]; ix++)
      {
        elems[k] = (ec[2][iz] * ne[1] + ec[1][iy]) * ne[0] + ec[0][ix];
        locs[k] = c * ncomp + (lc[2][iz] * n1 + lc[1][iy]) * n0 + lc[0][ix];
        k++;
      }
  return k;
}

int
hco_csr_assemble(const hco_mesh * m, int p, const double * qdata, const uint8_t * ess, hco_csr * A)
{
  tabs_t t;
  make_tabs(p, &t);
  const int Q = t.Q, Q3 = Q * Q * Q, nd = 3 * p * (p + 1) * (p + 1);
  const int64_t N = hco_ndofs(m, p);
  const int nx = m->nx, ny = m->ny, nz = m->nz;
  A->n = N;
  A->rowptr = (int64_t *)malloc(sizeof(int64_t) * (N + 1));
  int64_t * cnt = (int64_t *)calloc(N + 1, sizeof(int64_t));
  /* pass 1: row lengths */
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t g = 0; g < N; g++)
  {
    if (ess[g])
    {
      cnt[g] = 1;
      continue;
    }
    int64_t el[4], cols[4 * NDMAX], gid[NDMAX];
    int lo[4];
    int ne = dof_elements(m, p, g, el, lo);
    int nc = 0;
    for (int k = 0; k < ne; k++)
    {
      hco_elem_dofs(m, p, el[k], gid);
      for (int j = 0; j < nd; j++)
        if (!ess[gid[j]])
          cols[nc++] = gid[j];
    }
    qsort(cols, nc, sizeof(int64_t), cmp_i64);
    int u = 0;
    for (int j = 0; j < nc; j++)
      if (j == 0 || cols[j] != cols[j - 1])
        u++;
    cnt[g] = u;
  }
  A->rowptr[0] = 0;
  for (int64_t g = 0; g < N; g++)
    A->rowptr[g + 1] = A->rowptr[g] + cnt[g];
  free(cnt);
  A->nnz = A->rowptr[N];
  A->col = (int64_t *)malloc(sizeof(int64_t) * A->nnz);
  A->val = (double *)calloc(A->nnz, sizeof(double));
  if (!A->col || !A->val)
    return -1;
  /* pass 2: columns */
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t g = 0; g < N; g++)
  {
    int64_t * out = A->col + A->rowptr[g];
    if (ess[g])
    {
      out[0] = g;
      A->val[A->rowptr[g]] = 1.0;
      continue;
    }
    int64_t el[4], cols[4 * NDMAX], gid[NDMAX];
    int lo[4];
    int ne = dof_elements(m, p, g, el, lo);
    int nc = 0;
    for (int k = 0; k < ne; k++)
    {
      hco_elem_dofs(m, p, el[k], gid);
      for (int j = 0; j < nd; j++)
        if (!ess[gid[j]])
          cols[nc++] = gid[j];
    }
    qsort(cols, nc, sizeof(int64_t), cmp_i64);
    int u = 0;
    for (int j = 0; j < nc; j++)
      if (j == 0 || cols[j] != cols[j - 1])
        out[u++] = cols[j];
  }
  /* pass 3: element matrices scattered by colour */
  double * phi = (double *)malloc(sizeof(double) * nd * Q3 * 3);
  double * cphi = (double *)malloc(sizeof(double) * nd * Q3 * 3);
  ref_basis(&t, phi, cphi);
  for (int col = 0; col < 8; col++)
  {
    int cx = col & 1, cy = (col >> 1) & 1, cz = (col >> 2) & 1;
    int64_t hx = (nx - cx + 1) / 2, hy = (ny - cy + 1) / 2, hz = (nz - cz + 1) / 2;
    int64_t nsub = hx * hy * hz;
#pragma omp parallel
    {
      double * Ae = (double *)malloc(sizeof(double) * nd * nd);
#pragma omp for schedule(static)
      for (int64_t s = 0; s < nsub; s++)
      {
        int64_t ex = cx + 2 * (s % hx), ey = cy + 2 * ((s / hx) % hy), ez = cz + 2 * (s / (hx * hy));
        int64_t e = (ez * ny + ey) * nx + ex;
        int64_t gid[NDMAX];
        hco_elem_dofs(m, p, e, gid);
        elem_matrix_from_basis(nd, Q3, Q, phi, cphi, qdata + e * 12 * Q3, Ae);
        for (int i = 0; i < nd; i++)
        {
          int64_t r = gid[i];
          if (ess[r])
            continue;
          const int64_t * rc = A->col + A->rowptr[r];
          int64_t len = A->rowptr[r + 1] - A->rowptr[r];
          for (int j = 0; j < nd; j++)
          {
            int64_t cidx = gid[j];
            if (ess[cidx])
              continue;
            int64_t lo = 0, hi = len - 1;
            while (lo < hi)
            {
              int64_t mid = (lo + hi) / 2;
              if (rc[mid] < cidx)
                lo = mid + 1;
              else
                hi = mid;
            }
            A->val[A->rowptr[r] + lo] += Ae[(size_t)i * nd + j];
          }
        }
      }
      free(Ae);
    }
  }
  free(phi);
  free(cphi);
  return 0;
}

void
hco_csr_free(hco_csr * A)
{
  free(A->rowptr);
  free(A->col);
  free(A->val);
  A->rowptr = NULL;
  A->col = NULL;
  A->val = NULL;
}

void
hco_csr_apply(const hco_csr * A, const double * x, double * y)
{
#pragma omp parallel for schedule(static, 256)
  for (int64_t r = 0; r < A->n; r++)
  {
    double s = 0.0;
    for (int64_t k = A->rowptr[r]; k < A->rowptr[r + 1]; k++)
      s += A->val[k] * x[A->col[k]];
    y[r] = s;
  }
}

/* ------------------------------------------------------------------ */
/* Jacobi-PCG                                                           */
/* ------------------------------------------------------------------ */

typedef void (*apply_fn)(const void * ctx, const double * x, double * y);

/* Deterministic for a fixed thread count: static blocks summed in order. */
static double
pdot(int64_t n, const double * a, const double * b)
{
  int nt = hco_num_threads();
  double part[256];
  if (nt > 256)
    nt = 256;
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int tid = omp_get_thread_num(), T = omp_get_num_threads();
#else
    int tid = 0, T = 1;
#endif
    int64_t lo = n * tid / T, hi = n * (tid + 1) / T;
    double s = 0.0;
    for (int64_t i = lo; i < hi; i++)
      s += a[i] * b[i];
    part[tid] = s;
#pragma omp barrier
#pragma omp single
    {
      for (int k = T; k < nt; k++)
        part[k] = 0.0;
    }
  }
  double s = 0.0;
  for (int k = 0; k < nt; k++)
    s += part[k];
  return s;
}

static int
pcg_core(apply_fn A,
         const void * ctx,
         int64_t n,
         const double * dinv_src,
         const double * b,
         double * x,
         double rtol,
         int maxit,
         hco_pcg_stats * st,
         double * hist)
{
  double * r = (double *)malloc(sizeof(double) * n);
  double * z = (double *)malloc(sizeof(double) * n);
  double * pv = (double *)malloc(sizeof(double) * n);
  double * q = (double *)malloc(sizeof(double) * n);
  double * dinv = (double *)malloc(sizeof(double) * n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    dinv[i] = 1.0 / dinv_src[i];
  A(ctx, x, q);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
  {
    r[i] = b[i] - q[i];
    z[i] = dinv[i] * r[i];
    pv[i] = z[i];
  }
  double bb = pdot(n, b, b);
  double gamma = pdot(n, r, z);
  double rr = pdot(n, r, r);
  double bnorm = sqrt(bb);
  if (hist)
    hist[0] = bnorm > 0 ? sqrt(rr) / bnorm : 0.0;
  int it = 0, conv = (rr <= rtol * rtol * bb);
  while (!conv && it < maxit)
  {
    A(ctx, pv, q);
    double pq = pdot(n, pv, q);
    double alpha = gamma / pq;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
    {
      x[i] += alpha * pv[i];
      r[i] -= alpha * q[i];
      z[i] = dinv[i] * r[i];
    }
    double gnew = pdot(n, r, z);
    rr = pdot(n, r, r);
    it++;
    if (hist)
      hist[it] = sqrt(rr) / bnorm;
    if (rr <= rtol * rtol * bb)
    {
      conv = 1;
      break;
    }
    double beta = gnew / gamma;
    gamma = gnew;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
      pv[i] = z[i] + beta * pv[i];
  }
  /* true residual */
  A(ctx, x, q);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    r[i] = b[i] - q[i];
  double tr = pdot(n, r, r);
  if (st)
  {
    st->iterations = it;
    st->converged = conv;
    st->rel_res = bnorm > 0 ? sqrt(rr) / bnorm : 0.0;
    st->true_rel_res = bnorm > 0 ? sqrt(tr) / bnorm : 0.0;
  }
  free(r);
  free(z);
  free(pv);
  free(q);
  free(dinv);
  return conv ? it : -1;
}

typedef struct
{
  const hco_mesh * m;
  int p;
  const double * qdata;
  const uint8_t * ess;
} pa_ctx;

static void
pa_apply_cb(const void * c, const double * x, double * y)
{
  const pa_ctx * k = (const pa_ctx *)c;
  hco_pa_apply(k->m, k->p, k->qdata, k->ess, x, y);
}

static void
csr_apply_cb(const void * c, const double * x, double * y)
{
  hco_csr_apply((const hco_csr *)c, x, y);
}

int
hco_pcg_pa(const hco_mesh * m,
           int p,
           const double * qdata,
           const uint8_t * ess,
           const double * diag,
           const double * b,
           double * x,
           double rtol,
           int maxit,
           hco_pcg_stats * st,
           double * res_hist)
{
  pa_ctx c = {m, p, qdata, ess};
  return pcg_core(pa_apply_cb, &c, hco_ndofs(m, p), diag, b, x, rtol, maxit, st, res_hist);
}

int
hco_pcg_csr(const hco_csr * A,
            const double * diag,
            const double * b,
            double * x,
            double rtol,
            int maxit,
            hco_pcg_stats * st,
            double * res_hist)
{
  return pcg_core(csr_apply_cb, A, A->n, diag, b, x, rtol, maxit, st, res_hist);
}

/* ------------------------------------------------------------------ */
/* Inputs                                                               */
/* ------------------------------------------------------------------ */

static inline uint64_t
splitmix64_at(uint64_t seed, uint64_t i)
{
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static inline double
u01(uint64_t u)
{
  return (double)(u >> 11) * (1.0 / 9007199254740992.0);
}

void
hco_rand_uniform(uint64_t seed, int64_t n, double * out)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    out[i] = 2.0 * u01(splitmix64_at(seed, (uint64_t)i)) - 1.0;
}

void
hco_rhs_uniform(const hco_mesh * m, int p, uint64_t seed, const uint8_t * ess, double * b)
{
  int64_t N = hco_ndofs(m, p);
  hco_rand_uniform(seed, N, b);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++)
    if (ess[i])
      b[i] = 0.0;
}

void
hco_rhs_coil(const hco_mesh * m, int p, uint64_t seed, const uint8_t * ess, double * b)
{
  /* four Gaussian bumps of a vector potential psi; J = curl psi */
  double a[4][3], c[4][3], s[4];
  uint64_t k = 0;
  for (int j = 0; j < 4; j++)
  {
    for (int r = 0; r < 3; r++)
      a[j][r] = 2.0 * u01(splitmix64_at(seed, k++)) - 1.0;
    for (int r = 0; r < 3; r++)
      c[j][r] = 0.3 + 0.4 * u01(splitmix64_at(seed, k++));
    s[j] = 0.05 + 0.1 * u01(splitmix64_at(seed, k++));
  }
  const int Qr = p + 2, O = p, D = p + 1, nd = 3 * p * (p + 1) * (p + 1);
  double qp[QMAX + 1], qw[QMAX + 1];
  double Bo[(QMAX + 1) * HCO_PMAX], Bc[(QMAX + 1) * DMAX], Gc[(QMAX + 1) * DMAX];
  hco_gauss_legendre(Qr, qp, qw);
  tables_at(p, Qr, qp, Bo, Bc, Gc);
  int64_t N = hco_ndofs(m, p);
  const int nx = m->nx, ny = m->ny, nz = m->nz;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++)
    b[i] = 0.0;
  for (int col = 0; col < 8; col++)
  {
    int cx = col & 1, cy = (col >> 1) & 1, cz = (col >> 2) & 1;
    int64_t hx = (nx - cx + 1) / 2, hy = (ny - cy + 1) / 2, hz = (nz - cz + 1) / 2;
    int64_t nsub = hx * hy * hz;
#pragma omp parallel for schedule(static)
    for (int64_t si = 0; si < nsub; si++)
    {
      int64_t ex = cx + 2 * (si % hx), ey = cy + 2 * ((si / hx) % hy), ez = cz + 2 * (si / (hx * hy));
      int64_t e = (ez * ny + ey) * nx + ex;
      int64_t gid[NDMAX];
      double be[NDMAX];
      hco_elem_dofs(m, p, e, gid);
      for (int i = 0; i < nd; i++)
        be[i] = 0.0;
      for (int qz = 0; qz < Qr; qz++)
        for (int qy = 0; qy < Qr; qy++)
          for (int qx = 0; qx < Qr; qx++)
          {
            double X[3], J[9];
            elem_point(m, e, qp[qx], qp[qy], qp[qz], X);
            elem_jacobian(m, e, qp[qx], qp[qy], qp[qz], J);
            double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                         J[2] * (J[3] * J[7] - J[4] * J[6]);
            /* current density at X */
            double Jc[3] = {0, 0, 0};
            for (int j = 0; j < 4; j++)
            {
              double d0 = X[0] - c[j][0], d1 = X[1] - c[j][1], d2 = X[2] - c[j][2];
              double g = exp(-(d0 * d0 + d1 * d1 + d2 * d2) / (2 * s[j] * s[j]));
              double gr[3] = {-d0 / (s[j] * s[j]) * g, -d1 / (s[j] * s[j]) * g, -d2 / (s[j] * s[j]) * g};
              /* curl(g a) = grad g x a */
              Jc[0] += gr[1] * a[j][2] - gr[2] * a[j][1];
              Jc[1] += gr[2] * a[j][0] - gr[0] * a[j][2];
              Jc[2] += gr[0] * a[j][1] - gr[1] * a[j][0];
            }
            /* pulled back: J^{-1} Jc * |det| * w  ==  adj(J) Jc * sign(det) * w */
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
            double sg = det >= 0 ? 1.0 : -1.0;
            double w = qw[qx] * qw[qy] * qw[qz] * sg;
            double f[3];
            for (int r = 0; r < 3; r++)
              f[r] = w * (adj[r * 3 + 0] * Jc[0] + adj[r * 3 + 1] * Jc[1] + adj[r * 3 + 2] * Jc[2]);
            int l = 0;
            for (int cc = 0; cc < 3; cc++)
            {
              int n0 = (cc == 0) ? O : D, n1 = (cc == 1) ? O : D, n2 = (cc == 2) ? O : D;
              for (int dz = 0; dz < n2; dz++)
                for (int dy = 0; dy < n1; dy++)
                  for (int dx = 0; dx < n0; dx++, l++)
                  {
                    double vx = (cc == 0) ? Bo[qx * O + dx] : Bc[qx * D + dx];
                    double vy = (cc == 1) ? Bo[qy * O + dy] : Bc[qy * D + dy];
                    double vz = (cc == 2) ? Bo[qz * O + dz] : Bc[qz * D + dz];
                    be[l] += f[cc] * vx * vy * vz;
                  }
            }
          }
      for (int i = 0; i < nd; i++)
        b[gid[i]] += be[i];
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++)
    if (ess[i])
      b[i] = 0.0;
}
