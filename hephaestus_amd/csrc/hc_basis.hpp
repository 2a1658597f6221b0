// 1D bases of the ND_p hex element on [0,1]: Gauss-Legendre points and
// weights, Gauss-Lobatto points, Lagrange values and derivatives, and the
// kernel tables built from them (open basis at the quadrature points, closed
// basis and its derivative).  Everything here is constexpr: the tile kernels
// fold the tables of their order into instruction constants (no table in
// scalar registers for the whole kernel), and the host evaluates the same
// functions at run time for every other consumer, so all of them see the
// same numbers.
#pragma once

#include "hc_common.hpp"

namespace hc
{
namespace basis
{

constexpr double PI = 3.14159265358979323846;

constexpr double
abs_(double x)
{
  return x < 0 ? -x : x;
}

// cos on [0, pi] from its Taylor series (Newton starting values only).
constexpr double
cos_start(double x)
{
  double s = 1.0;
  if (x > PI / 2)
  {
    x = PI - x;
    s = -1.0;
  }
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 12; k++)
  {
    term *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return s * sum;
}

// Legendre P_n(x) and P_n'(x) by the three-term recurrence.
constexpr void
legendre_pd(int n, double x, double & P, double & dP)
{
  if (n == 0)
  {
    P = 1.0;
    dP = 0.0;
    return;
  }
  double p0 = 1.0, p1 = x;
  for (int k = 2; k <= n; k++)
  {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  P = p1;
  dP = n * (x * p1 - p0) / (x * x - 1.0);
}

// n Gauss-Legendre points (ascending) and weights (summing to 1) on (0,1):
// Newton on P_n from cos(pi (i + 3/4) / (n + 1/2)).
constexpr void
gauss_legendre01(int n, double * x, double * w)
{
  for (int i = 0; i < n; i++)
  {
    double z = cos_start(PI * (i + 0.75) / (n + 0.5)), P = 0, dP = 0;
    for (int it = 0; it < 100; it++)
    {
      legendre_pd(n, z, P, dP);
      const double dz = P / dP;
      z -= dz;
      if (abs_(dz) < 1e-16)
        break;
    }
    legendre_pd(n, z, P, dP);
    x[n - 1 - i] = 0.5 * (z + 1.0);
    w[n - 1 - i] = 1.0 / ((1.0 - z * z) * dP * dP);
  }
}

// n Gauss-Lobatto points on [0,1] (ends included): Newton on P'_{n-1}.
constexpr void
gauss_lobatto01(int n, double * x)
{
  const int p = n - 1;
  x[0] = 0.0;
  x[p] = 1.0;
  for (int i = 1; i < p; i++)
  {
    double z = -cos_start(PI * i / p), P = 0, dP = 0;
    for (int it = 0; it < 100; it++)
    {
      legendre_pd(p, z, P, dP);
      const double d2P = (2.0 * z * dP - p * (p + 1.0) * P) / (1.0 - z * z);
      const double dz = dP / d2P;
      z -= dz;
      if (abs_(dz) < 1e-16)
        break;
    }
    x[i] = 0.5 * (z + 1.0);
  }
}

// Lagrange basis of the n nodes at t: values (and derivatives when der).
constexpr void
lagrange(int n, const double * nodes, double t, double * val, double * der)
{
  for (int j = 0; j < n; j++)
  {
    double v = 1.0, d = 0.0;
    for (int m = 0; m < n; m++)
      if (m != j)
        v *= (t - nodes[m]) / (nodes[j] - nodes[m]);
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
    val[j] = v;
    if (der)
      der[j] = d;
  }
}

// The kernel tables of order p (entries past Q = p + 1 and past the basis
// size are zero).
template <typename T>
struct Tab
{
  T Bo[QMAXC * PMAX] = {};
  T Bc[QMAXC * QMAXC] = {};
  T Gc[QMAXC * QMAXC] = {};
  T qw[QMAXC] = {};
  T qp[QMAXC] = {};
};

constexpr Tab<double>
make_tab(int p)
{
  Tab<double> t{};
  const int Q = p + 1, D = p + 1;
  double on[PMAX] = {}, ow[PMAX] = {}, cn[PMAX + 1] = {};
  gauss_legendre01(p, on, ow);
  gauss_lobatto01(p + 1, cn);
  gauss_legendre01(Q, t.qp, t.qw);
  for (int q = 0; q < Q; q++)
  {
    double v[PMAX + 1] = {}, d[PMAX + 1] = {};
    lagrange(p, on, t.qp[q], v, nullptr);
    for (int j = 0; j < p; j++)
      t.Bo[q * p + j] = v[j];
    lagrange(D, cn, t.qp[q], v, d);
    for (int j = 0; j < D; j++)
    {
      t.Bc[q * D + j] = v[j];
      t.Gc[q * D + j] = d[j];
    }
  }
  return t;
}

template <typename T>
constexpr Tab<T>
cast_tab(const Tab<double> & a)
{
  Tab<T> b{};
  for (int i = 0; i < QMAXC * PMAX; i++)
    b.Bo[i] = (T)a.Bo[i];
  for (int i = 0; i < QMAXC * QMAXC; i++)
  {
    b.Bc[i] = (T)a.Bc[i];
    b.Gc[i] = (T)a.Gc[i];
  }
  for (int i = 0; i < QMAXC; i++)
  {
    b.qw[i] = (T)a.qw[i];
    b.qp[i] = (T)a.qp[i];
  }
  return b;
}

// Tables of order P in element type T, evaluated at compile time.
template <int P, typename T>
__host__ __device__ constexpr Tab<T>
tab()
{
  return cast_tab<T>(make_tab(P));
}

} // namespace basis
} // namespace hc
